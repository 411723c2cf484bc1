"""GPU parity of the post-GKR inner-product sumcheck (crates/post_gkr) with
the oracle, bit for bit: the narg string (every round message, the
nonces at Bits128, the closing evaluation), the point and the evaluation,
on shapes from the smallest (t = 7) to the end-to-end one (2^7 x 2^15 bits),
host and device inputs; the call order through a grind hook; the errors."""
import numpy as np
import pytest

import bitz_amd as G
from common import elements, elements_fast, ints
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, scope="module")
def _fast_oracle():
    O.set_kernel("pclmul")
    yield
    O.set_kernel("portable")


def instance(t, s, seed, density=0.5):
    rng = np.random.default_rng(seed)
    rows, cols = elements_fast(seed, 1 << t), elements_fast(seed + 1, 1 << s)
    bits = (rng.random(1 << (t + s)) < density).astype(np.uint8)
    packed = np.frombuffer(np.packbits(bits, bitorder="little").tobytes(), dtype=np.uint64).reshape(-1, 2).copy()
    # the target from the oracle's own weighted sum: a prover run with a wrong
    # target fails, so search nothing and read it off a one-element check
    return rows, cols, packed, bits


def target_of(rows, cols, bits):
    t = len(rows).bit_length() - 1
    idx = np.nonzero(bits)[0]
    w = O.mul_batch(rows[idx & ((1 << t) - 1)], cols[idx >> t])
    return int(np.bitwise_xor.reduce(w[:, 0]) if len(w) else 0) | (
        int(np.bitwise_xor.reduce(w[:, 1]) if len(w) else 0) << 64)


@pytest.mark.parametrize("t,s,security", [(7, 0, 100), (7, 3, 128), (8, 6, 100), (11, 4, 128), (7, 15, 100),
                                          (7, 15, 128), (9, 12, 100)])
def test_post_gkr_matches_the_oracle(t, s, security):
    rows, cols, packed, bits = instance(t, s, 1000 + 31 * t + s)
    target = target_of(rows, cols, bits)
    ot = O.DuplexTranscript(b"post_gkr/gpu")
    rc, ref_point, ref_ev = O.post_gkr_prove(ot, rows, cols, target, packed, security)
    assert rc == 0
    gt = G.Transcript.duplex(b"post_gkr/gpu")
    point, ev = G.post_gkr_prove(gt, rows, cols, target, packed, security)
    assert np.array_equal(point, ref_point) and ev == ref_ev
    assert gt.narg() == ot.narg()
    assert O.squeeze(ot) == gt.squeeze()
    if t + s <= 14:  # the closing evaluation is MLE[f](point)
        f = np.zeros((1 << (t + s), 2), np.uint64)
        f[:, 0] = bits
        assert ev == O.mle_evaluate(f, point)
    dt = G.Transcript.duplex(b"post_gkr/gpu")
    dev = [G.DeviceTable(x) for x in (rows, cols, packed)]
    point2, ev2 = G.post_gkr_prove(dt, *dev[:2], target, dev[2], security)
    assert np.array_equal(point2, ref_point) and ev2 == ref_ev and dt.narg() == ot.narg()


def test_post_gkr_hands_grinding_to_the_callers_transcript():
    rows, cols, packed, bits = instance(7, 4, 77)
    target = target_of(rows, cols, bits)
    ch = ints(elements(78, 11))
    ot = O.ListTranscript(ch, grind_hook=True)
    rc, ref_point, ref_ev = O.post_gkr_prove(ot, rows, cols, target, packed, 128)
    assert rc == 0
    events, it = [], iter(ch)
    gt = G.Transcript.callbacks(lambda kind, data: events.append(("absorb", kind, data)),
                                lambda: (events.append(("squeeze",)), next(it).to_bytes(16, "little"))[1],
                                grind=lambda label, b: events.append(("grind", label, b)))
    point, ev = G.post_gkr_prove(gt, rows, cols, target, packed, G.BITS128)
    assert events == ot.events
    assert sum(e[0] == "grind" for e in events) == 11


def test_post_gkr_errors_come_before_the_transcript():
    rows, cols, packed, bits = instance(7, 2, 5)
    target = target_of(rows, cols, bits)
    t = G.Transcript.duplex(b"e")
    with pytest.raises(G.BitzError) as e:
        G.post_gkr_prove(t, rows, cols, target ^ 1, packed)
    assert e.value.variant == "ClaimDoesNotHold"
    with pytest.raises(G.BitzError) as e:
        G.post_gkr_prove(t, rows, cols, target, packed[:2])
    assert e.value.variant == "WitnessLengthMismatch"
    with pytest.raises(G.BitzError) as e:
        G.post_gkr_prove(t, rows[:64], cols, target, packed)
    assert e.value.variant == "InvalidArgument"
    assert t.narg() == b"" and t.squeeze() == G.Transcript.duplex(b"e").squeeze()
