"""GPU parity of the provers that draw their challenges on the device (the
Keccak duplex run by the kernels, bitz_amd/csrc/fs.hpp): bit for bit against
the oracle over its own Keccak duplex (oracle/transcript.c, pinned to
hashlib's SHA3/SHAKE in test_keccak), and against the same prover with every
challenge drawn by the host transcript.

GKR (gpgkr_prove, crates/gkr/src/lib.rs:24-159): the leaf point, the claim,
the narg string (every prover message in order) and the transcript state
after the proof (the next squeeze), over layers that end in the LDS tail
only and layers with grid-wide rounds first, top points with a zero
coordinate (the at-one sums), and a transcript that has absorbed a message
before the proof (the duplex does not start at position 0)."""
import numpy as np
import pytest

import bitz_amd as G
from common import elements, elements_fast
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, scope="module")
def _fast_oracle():
    O.set_kernel("pclmul")
    yield
    O.set_kernel("portable")
    G.set_device_fiat_shamir(True)


def gkr_case(log_leaves, seed):
    leaves = elements_fast(seed, 1 << log_leaves)
    zero = (leaves[:, 0] == 0) & (leaves[:, 1] == 0)
    leaves[zero, 0] = 1
    return leaves


def prove_both(leaves, log_groups, point, domain, prefix, device):
    """(oracle, library) proofs over Keccak transcripts of `domain` that
    first absorbed `prefix` as a public message."""
    log_leaves = leaves.shape[0].bit_length() - 1
    store, _ = O.gkr_fold_layers(leaves, log_groups)
    ot = O.KeccakTranscript(domain)
    gt = G.Transcript.keccak(domain)
    if prefix:
        O.absorb(ot, 0, prefix)
        gt.absorb(0, prefix)
    ref = O.gkr_prove(ot, point, store, log_leaves)
    _, wit = G.grand_product(G.DeviceTable(leaves), log_groups)
    G.set_device_fiat_shamir(device)
    try:
        got = G.gpgkr_prove(gt, point, wit)
    finally:
        G.set_device_fiat_shamir(True)
    return (ot, ref), (gt, got)


@pytest.mark.parametrize("log_leaves,log_groups,zero,prefix",
                         [(3, 1, None, b""), (6, 2, None, b""), (10, 3, 0, b"x" * 50), (12, 1, None, b""),
                          (14, 7, 6, b"y" * 130), (16, 1, 0, b""), (18, 2, 1, b"z")])
def test_gkr_device_fs_matches_oracle(log_leaves, log_groups, zero, prefix):
    leaves = gkr_case(log_leaves, 300 + log_leaves)
    point = elements(310 + log_leaves, log_groups)
    if zero is not None:
        point[zero] = 0
    domain = b"gkr/device-fs/%d" % log_leaves
    (ot, (ref_point, ref_claim)), (gt, (gp, gc)) = prove_both(leaves, log_groups, point, domain, prefix, True)
    assert np.array_equal(gp, ref_point)
    assert gc == ref_claim
    assert gt.narg() == ot.narg()
    assert gt.squeeze() == O.squeeze(ot)  # the transcript continues from the device's state
    assert gc == O.mle_evaluate(leaves, gp)


@pytest.mark.parametrize("log_leaves,log_groups", [(11, 2), (15, 1)])
def test_gkr_device_fs_equals_host_fs(log_leaves, log_groups):
    leaves = gkr_case(log_leaves, 320 + log_leaves)
    point = elements(330 + log_leaves, log_groups)
    _, (gd, (pd, cd)) = prove_both(leaves, log_groups, point, b"gkr/both", b"p", True)
    _, (gh, (ph, ch)) = prove_both(leaves, log_groups, point, b"gkr/both", b"p", False)
    assert np.array_equal(pd, ph) and cd == ch
    assert gd.narg() == gh.narg()
    assert gd.squeeze() == gh.squeeze()


def test_gkr_keccak_bits128_grinds_on_the_host():
    """Grinding needs a proof-of-work search: a Bits128 proof over a Keccak
    transcript runs the host path and matches the oracle's grinding."""
    leaves = gkr_case(9, 340)
    point = elements(341, 2)
    store, _ = O.gkr_fold_layers(leaves, 2)
    ot = O.KeccakTranscript(b"gkr/bits128")
    ref_point, ref_claim = O.gkr_prove(ot, point, store, 9, security_bits=128)
    _, wit = G.grand_product(G.DeviceTable(leaves), 2)
    gt = G.Transcript.keccak(b"gkr/bits128")
    gp, gc = G.gpgkr_prove(gt, point, wit, G.BITS128)
    assert np.array_equal(gp, ref_point) and gc == ref_claim
    assert gt.narg() == ot.narg()
