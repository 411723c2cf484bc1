"""The oracle's post-GKR inner-product sumcheck (crates/post_gkr) pinned
without its own structure: every round message against the dense
definition on the unpacked bits (the big-int model in tests/models.py),
the closing evaluation against MLE[f](rho) by definition, the reference's
verifier replay (accepting, and rejecting tampering), the two prover errors,
and the grinding boundary before each challenge at Bits128."""
import numpy as np

from common import elements, ints, mle_eval_def
from models import gf_mul
from oracle import oracle as O


def instance(t, s, seed, density=0.5):
    rng = np.random.default_rng(seed)
    rows, cols = elements(seed, 1 << t), elements(seed + 1, 1 << s)
    bits = (rng.random(1 << (t + s)) < density).astype(np.uint8)
    packed = np.frombuffer(np.packbits(bits, bitorder="little").tobytes(), dtype=np.uint64).reshape(-1, 2).copy()
    ri, ci = ints(rows), ints(cols)
    w = [gf_mul(ri[i & ((1 << t) - 1)], ci[i >> t]) for i in range(1 << (t + s))]
    target = 0
    for i, b in enumerate(bits):
        if b:
            target ^= w[i]
    return rows, cols, packed, bits, w, target


def dense_rounds(w, f, challenges):
    """Each round's (a0, a2) and the running tables, by definition."""
    w, f, out = list(w), list(f), []
    for r in challenges:
        a0 = a2 = 0
        for j in range(0, len(w), 2):
            a0 ^= gf_mul(w[j], f[j])
            a2 ^= gf_mul(w[j] ^ w[j + 1], f[j] ^ f[j + 1])
        out.append((a0, a2))
        w = [w[j] ^ gf_mul(r, w[j] ^ w[j + 1]) for j in range(0, len(w), 2)]
        f = [f[j] ^ gf_mul(r, f[j] ^ f[j + 1]) for j in range(0, len(f), 2)]
    return out, f[0]


def messages_of(tr):
    sent = [m for kind, m in tr.messages if kind == 1]
    pairs = [np.frombuffer(m, dtype=np.uint64).reshape(2, 2) for m in sent[:-1]]
    return pairs, np.frombuffer(sent[-1], dtype=np.uint64).reshape(1, 2)


def test_rounds_match_the_dense_definition_and_the_evaluation_is_the_mle():
    for t, s, seed in ((7, 0, 1), (7, 1, 2), (8, 1, 3)):
        rows, cols, packed, bits, w, target = instance(t, s, seed)
        ch = ints(elements(100 + seed, t + s))
        tr = O.ListTranscript(ch)
        rc, point, ev = O.post_gkr_prove(tr, rows, cols, target, packed)
        assert rc == 0
        assert ints(point) == ch
        expect, v = dense_rounds(w, [int(b) for b in bits], ch)
        pairs, last = messages_of(tr)
        assert [tuple(ints(p)) for p in pairs] == expect
        assert ev == v == mle_eval_def([int(b) for b in bits], ch) == ints(last)[0]


def test_the_verifier_accepts_and_rejects_tampering():
    rows, cols, packed, bits, w, target = instance(9, 2, 7)
    for security in (100, 128):
        dt = O.DuplexTranscript(b"post_gkr")
        rc, point, ev = O.post_gkr_prove(dt, rows, cols, target, packed, security)
        assert rc == 0
        lt = O.DuplexTranscript(b"post_gkr")
        # the messages, read back from a recording run on the same challenges
        rec = O.ListTranscript(ints(point), grind_hook=True)
        O.post_gkr_prove(rec, rows, cols, target, packed, security)
        pairs, _ = messages_of(rec)
        msgs = np.concatenate(pairs)
        rc, vpoint = O.post_gkr_verify(lt, rows, cols, target, msgs, ev, security)
        assert rc == 0 and np.array_equal(vpoint, point)
        assert dt.narg() == lt.narg()
        bad = msgs.copy()
        bad[3, 0] ^= 1
        assert O.post_gkr_verify(O.DuplexTranscript(b"post_gkr"), rows, cols, target, bad, ev, security)[0] != 0
        assert O.post_gkr_verify(O.DuplexTranscript(b"post_gkr"), rows, cols, target, msgs, ev ^ 1, security)[0] != 0


def test_prover_errors():
    rows, cols, packed, bits, w, target = instance(7, 1, 5)
    assert O.post_gkr_prove(O.SeededTranscript(1), rows, cols, target ^ 1, packed)[0] == -14  # ClaimDoesNotHold
    assert O.post_gkr_prove(O.SeededTranscript(1), rows, cols, target, packed[:1])[0] == -13  # WitnessLengthMismatch


def test_bits128_grinds_one_bit_before_every_challenge():
    """lib.rs:142, sumcheck.rs:131: grinding_bits(2) = 1 at Bits128, 0 at Bits100."""
    rows, cols, packed, bits, w, target = instance(7, 2, 9)
    for security, bits_expected in ((128, 1), (100, 0)):
        tr = O.ListTranscript(ints(elements(3, 9)), grind_hook=True)
        assert O.post_gkr_prove(tr, rows, cols, target, packed, security)[0] == 0
        kinds = [e[0] for e in tr.events]
        grinds = [e for e in tr.events if e[0] == "grind"]
        if bits_expected:
            assert len(grinds) == 9 and all(g == ("grind", b"bitz/pcs/sumcheck/v1", 1) for g in grinds)
            assert all(kinds[i + 1] == "squeeze" for i, k in enumerate(kinds) if k == "grind")
        else:
            assert not grinds
