"""The inner sumcheck, Az/Bz/Cz, the PIOP and the engine over a Boolean
assignment kept as packed bits (the *_packed entry points) against the
oracle's inner sumcheck over the table build_assignment_mle gives
(crates/spartan/src/matrix.rs:392-420, sumcheck.rs:442-552), and against the
same entry points over that table.  Tables with zero tails (D longer than h,
h longer than D), sparse D, all-zero and all-one assignments, sizes from one
entry to 2^20 (every bit round, the switch to elements, the grid-wide and
single-workgroup rounds after it)."""
import numpy as np
import pytest

import bitz_amd as G
import bitz_amd.sha256 as S
from common import elements_fast, ints
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def assignment(rng, n, density=0.5, h_used=1.0):
    """Random bits over 2^n entries with h[0] = 1, zero past the first h_used
    fraction."""
    N = 1 << n
    bits = (rng.random(N) < density).astype(np.uint8)
    if h_used < 1:
        bits[max(1, int(N * h_used)) | 1:] = 0
    bits[0] = 1
    return bits


def matrix(seed, n, d_used=1.0, sparsity=0.0):
    """D: uniform elements, zero past the first d_used fraction and at a
    `sparsity` fraction of places."""
    N = 1 << n
    d = elements_fast(seed, N)
    rng = np.random.default_rng(seed + 1)
    if sparsity:
        d[rng.random(N) < sparsity] = 0
    if d_used < 1:
        d[max(1, int(N * d_used)):] = 0
    return d


def table(bits):
    h = np.zeros((len(bits), 2), np.uint64)
    h[:, 0] = bits
    return h


CASES = [
    # n, density, h_used, d_used, sparsity
    (0, 0.5, 1.0, 1.0, 0.0),
    (1, 0.5, 1.0, 1.0, 0.0),
    (2, 0.5, 1.0, 1.0, 0.0),
    (3, 0.5, 1.0, 1.0, 0.0),
    (4, 0.5, 1.0, 1.0, 0.0),
    (5, 0.3, 1.0, 1.0, 0.0),
    (6, 0.5, 1.0, 0.7, 0.0),
    (7, 1.0, 1.0, 1.0, 0.0),
    (9, 0.0, 1.0, 1.0, 0.0),
    (10, 0.5, 0.6, 1.0, 0.0),
    (12, 0.5, 1.0, 0.3, 0.97),
    (13, 0.1, 0.62, 0.8, 0.9),
    (16, 0.5, 1.0, 1.0, 0.0),
    (20, 0.4, 0.62, 1.0, 0.97),
]


@pytest.mark.parametrize("n,density,h_used,d_used,sparsity", CASES)
def test_packed_inner_sumcheck_matches_oracle(n, density, h_used, d_used, sparsity):
    rng = np.random.default_rng(1000 + n)
    bits = assignment(rng, n, density, h_used)
    d, h = matrix(2000 + n, n, d_used, sparsity), table(bits)
    claim = O.inner_product(d, h)
    ot = O.DuplexTranscript(b"inner/packed")
    rc, ref = O.inner_sumcheck(ot, claim, d, h)
    assert rc == 0
    for dev in (False, True):
        packed = G.PackedAssignment.from_bits(bits, 1 << n, device=dev)
        assert np.array_equal(packed.to_table(), h)
        gt = G.Transcript.duplex(b"inner/packed")
        out = G.prove_inner_sumcheck_packed(gt, claim, G.DeviceTable(d) if dev else d, packed)
        assert np.array_equal(out.round_polynomials, ref["rounds"]), dev
        assert np.array_equal(out.eval_points, ref["points"])
        assert out.final_claim == ref["final_claim"]
        assert [out.batched_matrix_evaluation, out.witness_evaluation] == ints(ref["evals"])


def test_packed_inner_sumcheck_rejects_mismatched_sizes():
    bits = np.ones(8, np.uint8)
    with pytest.raises(G.SumcheckError) as e:
        G.prove_inner_sumcheck_packed(G.Transcript.seeded(1), 0, elements_fast(1, 4),
                                      G.PackedAssignment.from_bits(bits, 8))
    assert e.value.variant == "InvalidProductDimensions"


def test_assignment_pack_checks_and_clears_the_tail():
    """build_assignment_mle's errors; bits past the assignment cleared."""
    with pytest.raises(G.BitzError) as e:
        G.PackedAssignment(np.array([2], np.uint64), 5, 5)
    assert e.value.variant == "InvalidAssignmentConstant"
    with pytest.raises(G.BitzError) as e:
        G.PackedAssignment(np.array([1], np.uint64), 5, 6)
    assert e.value.variant == "InvalidAssignmentLength"
    words = np.array([~np.uint64(0), ~np.uint64(0)], np.uint64)
    for dev in (False, True):
        p = G.PackedAssignment(words, 70, 70, device=dev)
        t = p.to_table()
        assert len(t) == 128 and t[:70, 0].all() and not t[70:].any()
        assert np.array_equal(t, G.assignment_mle_words(words, 70, 70))


def test_packed_r1cs_products_match_the_table_form():
    rng = np.random.default_rng(5)
    rows, row_vars, col_vars = 700, 10, 12
    csr = []
    for m in range(3):
        starts, cols = [0], []
        for _ in range(rows):
            cs = np.unique(rng.integers(0, 1 << col_vars, size=7))
            cols += cs.tolist()
            starts.append(len(cols))
        csr.append((np.array(starts, np.uint64), np.array(cols, np.uint32), elements_fast(90 + m, len(cols))))
    prepared = G.PreparedConstraintMatrices(rows, row_vars, col_vars, csr)
    bits = assignment(rng, col_vars, 0.5, 0.8)
    h = table(bits)
    ref = prepared.r1cs_products(h)
    for dev in (False, True):
        out = prepared.r1cs_products(G.PackedAssignment.from_bits(bits, 1 << col_vars, device=dev))
        for a, b in ((out.az, ref.az), (out.bz, ref.bz), (out.cz, ref.cz)):
            assert np.array_equal(a.to_host() if dev else a, b)


_circuits = {}


def request(blocks, seed):
    """A compression-chain request: the table form and the packed form of
    its assignment on the device, and its products."""
    if blocks not in _circuits:
        c = S.Sha256Circuit(S.COMPRESSION_CHAIN, blocks)
        _circuits[blocks] = (c, *c.prepare())
    c, matrices, digest = _circuits[blocks]
    rng = np.random.default_rng(seed)
    msg = rng.integers(0, 256, 64 * blocks, dtype=np.uint8).tobytes()
    iv = [int(x) for x in rng.integers(0, 2**32, 8, dtype=np.uint64)]
    w = c.witness(msg, iv=iv)
    h = G.assignment_mle_words(w.h, c.columns, c.columns, device=True)
    packed = G.PackedAssignment(w.h, c.columns, c.columns, device=True)
    return matrices, digest, h, packed


def same(a, b):
    return (np.array_equal(a.outer_rounds, b.outer_rounds) and a.outer_claims == b.outer_claims and
            np.array_equal(a.inner_rounds, b.inner_rounds) and np.array_equal(a.inner_point, b.inner_point) and
            a.matrix_evaluation == b.matrix_evaluation and a.final_claim == b.final_claim)


@pytest.mark.parametrize("blocks", [1, 3, 16])
def test_packed_piop_equals_the_table_piop(blocks):
    matrices, digest, h, packed = request(blocks, 40 + blocks)
    products = matrices.r1cs_products(h)
    pp = matrices.r1cs_products(packed)
    for a, b in ((pp.az, products.az), (pp.bz, products.bz), (pp.cz, products.cz)):
        assert np.array_equal(a.to_host(), b.to_host())
    ref = G.prove_spartan_piop(G.Transcript.duplex(b"packed/%d" % blocks), matrices, digest, products, h)
    out = G.prove_spartan_piop(G.Transcript.duplex(b"packed/%d" % blocks), matrices, digest, pp, packed)
    assert same(out, ref)


def test_packed_engine_matches_one_proof_at_a_time():
    """Packed and table proofs in one engine, joining at different steps;
    each equals its single proof."""
    sizes = [1, 2, 8, 1, 4, 32, 2]
    reqs = {tag: request(b, 300 + tag) for tag, b in enumerate(sizes)}
    refs = {}
    for tag, (m, dg, h, packed) in reqs.items():
        refs[tag] = G.prove_spartan_piop(G.Transcript.duplex(b"engine/%d" % tag), m, dg, m.r1cs_products(h), h)
    eng = G.PiopEngine()
    keep = {}
    order = list(reqs)
    steps = 0
    while order or eng.in_flight:
        for _ in range(2):  # two proofs join per step
            if order:
                tag = order.pop(0)
                m, dg, h, packed = reqs[tag]
                use = packed if tag % 3 else h
                keep[tag] = (G.Transcript.duplex(b"engine/%d" % tag), m.r1cs_products(packed))
                eng.submit(tag, keep[tag][0], m, dg, keep[tag][1], use)
        eng.step()
        steps += 1
        assert steps < 400
    for tag in reqs:
        assert same(eng.take(tag), refs[tag]), tag


def test_engine_device_transcript_matches_the_host_transcript():
    """Proofs with a Keccak transcript in the engine (challenges drawn on the
    device, several rounds per step, no host round trip per round) equal the
    single prover's with the same transcript drawn on the host, and the
    caller's transcript continues from the same state; proofs with a host
    transcript share the engine's steps."""
    sizes = [1, 2, 8, 1, 4, 32, 2, 64, 1]
    reqs = {tag: request(b, 500 + tag) for tag, b in enumerate(sizes)}
    host_fs = {tag for tag in reqs if tag % 4 == 3}

    def transcript(tag):
        return G.Transcript.duplex(b"fs/%d" % tag) if tag in host_fs else G.Transcript.keccak(b"fs/%d" % tag)

    refs, after = {}, {}
    for tag, (m, dg, h, packed) in reqs.items():
        t = transcript(tag)
        refs[tag] = G.prove_spartan_piop(t, m, dg, m.r1cs_products(packed), packed)
        after[tag] = t.squeeze()
    for background in (False, True):
        eng = G.PiopEngine(background=background)
        keep = {}
        for tag, (m, dg, h, packed) in reqs.items():
            keep[tag] = (transcript(tag), m.r1cs_products(packed))
            eng.submit(tag, keep[tag][0], m, dg, keep[tag][1], packed)
        if background:
            while len(eng.done()) < len(reqs):
                eng.wait(0.01)
        else:
            steps = 0
            while eng.step():
                steps += 1
                assert steps < 200
        for tag in reqs:
            assert same(eng.take(tag), refs[tag]), (background, tag)
            assert keep[tag][0].squeeze() == after[tag], (background, tag)


def test_engine_from_the_assignment_words():
    """bitz_piop_engine_submit_assignment (host words in; the engine uploads
    them and computes Az, Bz, Cz) gives the single prover's proof, with a
    Keccak transcript (device challenges) and a duplex one (host challenges);
    a wrong length or constant bit is rejected before the transcript."""
    sizes = [1, 4, 2, 16]
    circuits = {b: (lambda c: (c, *c.prepare()))(S.Sha256Circuit(S.COMPRESSION_CHAIN, b)) for b in set(sizes)}
    eng = G.PiopEngine(background=True)
    refs, keep = {}, {}
    for tag, b in enumerate(sizes):
        c, m, dg = circuits[b]
        rng = np.random.default_rng(900 + tag)
        msg = rng.integers(0, 256, 64 * b, dtype=np.uint8).tobytes()
        iv = [int(x) for x in rng.integers(0, 2**32, 8, dtype=np.uint64)]
        w = c.witness(msg, iv=iv)
        mk = (lambda t: G.Transcript.keccak(b"words/%d" % t)) if tag % 2 == 0 else \
            (lambda t: G.Transcript.duplex(b"words/%d" % t))
        packed = G.PackedAssignment(w.h, c.columns, c.columns, device=True)
        refs[tag] = G.prove_spartan_piop(mk(tag), m, dg, m.r1cs_products(packed), packed)
        keep[tag] = mk(tag)
        eng.submit_assignment(tag, keep[tag], m, dg, w.h, c.columns, c.columns)
    while len(eng.done()) < len(sizes):
        eng.wait(0.01)
    for tag in range(len(sizes)):
        assert same(eng.take(tag), refs[tag]), tag
    c, m, dg = circuits[1]
    t = G.Transcript.keccak(b"words/err")
    before = G.Transcript.keccak(b"words/err").squeeze()
    with pytest.raises(G.SumcheckError) as e:
        eng.submit_assignment(99, t, m, dg, np.zeros((c.columns + 63) // 64, np.uint64), c.columns, c.columns)
    assert e.value.variant == "InvalidAssignmentConstant"
    with pytest.raises(G.SumcheckError) as e:
        eng.submit_assignment(99, t, m, dg, np.ones(1000, np.uint64), c.columns - 1, c.columns)
    assert e.value.variant == "InvalidAssignmentLength"
    assert t.squeeze() == before  # nothing reached the transcript
