"""The batched PIOP engine (bitz_piop_engine_*): proofs of different sizes
stepped together give, each, exactly what prove_spartan_piop gives for the
same transcript and inputs (which tests/test_gpu_piop.py and
tests/test_gpu_sha256.py check against the oracle), whether submitted
together or joining while others are in flight, two proofs sharing one
circuit included."""
import numpy as np
import pytest

import bitz_amd as G
import bitz_amd.sha256 as S

pytestmark = pytest.mark.gpu


def same(a, b):
    return (np.array_equal(a.outer_rounds, b.outer_rounds) and a.outer_claims == b.outer_claims and
            np.array_equal(a.inner_rounds, b.inner_rounds) and np.array_equal(a.inner_point, b.inner_point) and
            a.matrix_evaluation == b.matrix_evaluation and a.final_claim == b.final_claim)


_circuits = {}


def request(blocks, seed):
    if blocks not in _circuits:
        c = S.Sha256Circuit(S.COMPRESSION_CHAIN, blocks)
        _circuits[blocks] = (c, *c.prepare())
    c, matrices, digest = _circuits[blocks]
    rng = np.random.default_rng(seed)
    msg = rng.integers(0, 256, 64 * blocks, dtype=np.uint8).tobytes()
    iv = [int(x) for x in rng.integers(0, 2**32, 8, dtype=np.uint64)]
    w = c.witness(msg, iv=iv)
    h = G.assignment_mle_words(w.h, c.columns, c.columns, device=True)
    return matrices, digest, matrices.r1cs_products(h), h


def alone(tag, matrices, digest, products, h):
    return G.prove_spartan_piop(G.Transcript.duplex(b"engine/%d" % tag), matrices, digest, products, h)


def test_engine_matches_one_proof_at_a_time():
    reqs = {tag: request(b, 100 + tag) for tag, b in enumerate([1, 2, 4, 8, 2, 16, 1])}
    ref = {tag: alone(tag, *r) for tag, r in reqs.items()}
    eng = G.PiopEngine()
    transcripts = {tag: G.Transcript.duplex(b"engine/%d" % tag) for tag in reqs}
    for tag, (m, dg, pr, h) in reqs.items():
        eng.submit(tag, transcripts[tag], m, dg, pr, h)
    steps = 0
    while eng.step():
        steps += 1
        assert steps < 200
    assert sorted(eng.done()) == sorted(reqs)
    for tag in reqs:
        out = eng.take(tag)
        assert same(out, ref[tag]), tag
    assert eng.done() == []


def test_engine_admits_proofs_while_others_run():
    reqs = {tag: request(b, 200 + tag) for tag, b in enumerate([8, 1, 4, 2, 32])}
    ref = {tag: alone(tag, *r) for tag, r in reqs.items()}
    eng = G.PiopEngine()
    pending = list(reqs)
    results, steps = {}, 0
    while pending or eng.step():
        if pending and steps % 3 == 0:  # a new proof joins every third step
            tag = pending.pop(0)
            eng.submit(tag, G.Transcript.duplex(b"engine/%d" % tag), *reqs[tag])
        for tag in eng.done():
            results[tag] = eng.take(tag)
        steps += 1
        assert steps < 400
    for tag in eng.done():
        results[tag] = eng.take(tag)
    assert sorted(results) == sorted(reqs)
    for tag in reqs:
        assert same(results[tag], ref[tag]), tag


def test_background_engine_with_submitters_on_threads():
    from concurrent.futures import ThreadPoolExecutor
    reqs = {tag: request(b, 300 + tag) for tag, b in enumerate([4, 1, 16, 2, 8, 1, 2])}
    ref = {tag: alone(tag, *r) for tag, r in reqs.items()}
    eng = G.PiopEngine(background=True)
    with ThreadPoolExecutor(3) as pool:
        list(pool.map(lambda t: eng.submit(t, G.Transcript.duplex(b"engine/%d" % t), *reqs[t]), reqs))
    results = {}
    for _ in range(2000):
        if eng.wait(0.01):
            for tag in eng.done():
                results[tag] = eng.take(tag)
        if len(results) == len(reqs):
            break
    assert sorted(results) == sorted(reqs)
    for tag in reqs:
        assert same(results[tag], ref[tag]), tag


def test_engine_with_callback_transcripts():
    """The caller's own transcript (callbacks, as the Rust side's spongefish
    would be) called from the engine's thread: the same messages and
    challenges as prove_spartan_piop under the same callbacks."""
    import hashlib

    class Sponge:
        def __init__(self, label):
            self.state = hashlib.sha256(label).digest()
            self.log = []

        def absorb(self, kind, data):
            self.log.append((kind, data))
            self.state = hashlib.sha256(self.state + bytes([kind]) + data).digest()

        def squeeze(self):
            self.state = hashlib.sha256(self.state + b"sq").digest()
            return self.state[:16]

    reqs = {tag: request(b, 400 + tag) for tag, b in enumerate([2, 8, 1])}
    ref, ref_logs = {}, {}
    for tag, r in reqs.items():
        sp = Sponge(b"cb/%d" % tag)
        ref[tag] = G.prove_spartan_piop(G.Transcript.callbacks(sp.absorb, sp.squeeze), *r)
        ref_logs[tag] = sp.log
    eng = G.PiopEngine(background=True)
    sponges = {tag: Sponge(b"cb/%d" % tag) for tag in reqs}
    for tag, r in reqs.items():
        eng.submit(tag, G.Transcript.callbacks(sponges[tag].absorb, sponges[tag].squeeze), *r)
    results = {}
    for _ in range(2000):
        if eng.wait(0.01):
            for tag in eng.done():
                results[tag] = eng.take(tag)
        if len(results) == len(reqs):
            break
    for tag in reqs:
        assert same(results[tag], ref[tag]), tag
        assert sponges[tag].log == ref_logs[tag]


def test_engine_errors():
    m, dg, pr, h = request(1, 7)
    eng = G.PiopEngine()
    eng.submit(1, G.Transcript.duplex(b"e"), m, dg, pr, h)
    with pytest.raises(G.BitzError):
        eng.submit(1, G.Transcript.duplex(b"e"), m, dg, pr, h)  # tag in use
    with pytest.raises(G.BitzError):
        eng.take(1)  # still in flight
    while eng.step():
        pass
    eng.take(1)
