"""Benchmark: SHA-256 blocks proved per second (sumcheck stage) with p50/p99
proof latency on one MI355X -- the serving workload, by default.

Default workload (serve): proving requests for SHA-256 compression chains of
2^U blocks (U uniform in 0..8; message bytes and chaining value uniform from
the seed).  Each request is its witness (host), the assignment (packed
bits), Az/Bz/Cz and prove_spartan_piop (outer sumcheck, bind_and_batch, inner
sumcheck) with every table on the device, served with continuous batching
(bitz_amd/serving.py).  A step is one set of --serve-requests requests, all
queued at once, served to completion; value = blocks proved per second over
the timed steps.  p50/p99 are per-request latencies (arrival to final claim)
under Poisson arrivals at --serve-rate requests/s; the latency at other
rates, the concurrent-slots mode and the plain one-request-at-a-time path
are reported beside it, as are the isolated outer sumcheck at 2^24 and
the roofline of the serving run's dominant kernel (the batched inner fold,
timed live with HIP events on the engines' streams).

Other workloads (--workload): outer (one 2^log_n outer sumcheck per step),
batch (256 outer sumchecks of 2^12..2^22), gkr (grand product over 2^22
leaves), inner, bind, chain (the sha256-chain statement), products (each
GF(2^128) product primitive over 2^log_n elements, the constant-table and
int8-MFMA products by a fixed element among them), mt (M^T apply and
transpose_inner over the SHA-256 circuits), eval (a multilinear's evaluation
over 2^26 device-resident entries).

Prints one JSON line.  The CPU baseline is the oracle (a C restatement of
the reference with the bit-serial multiply the reference selects on x86),
timed on a bounded sample before the GPU is used.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("SHA-256 blocks proved/sec (sumcheck stage) + p50/p99 proof latency, 1 MI355X")
HBM_PEAK_GBS = 8000.0


def measured_traffic(n: int):
    """HBM bytes per fold launch from the calibrated rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs recorded under profiles/ (2^24 workload only)."""
    path = os.path.join(ROOT, "profiles", "outer_fold_n24_traffic.json")
    if n != 24 or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["traffic_bytes_per_launch"]


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def make_inputs(n: int, seed: int):
    rng = np.random.default_rng(seed)
    a, b, c = (rng.integers(0, 2**64, size=(1 << n, 2), dtype=np.uint64) for _ in range(3))
    tau = rng.integers(0, 2**64, size=(n, 2), dtype=np.uint64)
    return a, b, c, tau


def cpu_baseline(n_sample: int, n_target: int) -> dict:
    """The oracle (bit-serial multiply, the kernel the reference selects on
    x86) proving one outer sumcheck at 2^n_sample on one host core, scaled
    linearly in the table length to 2^n_target."""
    from oracle import oracle as O
    O.set_kernel("portable")
    a, b, c, tau = make_inputs(n_sample, 7)
    lo, hi = O.eq_table(tau[: n_sample // 2]), O.eq_table(tau[n_sample // 2:])
    t0 = time.perf_counter()
    rc, _ = O.outer_sumcheck(O.SeededTranscript(1), 0, lo, hi, a, b, c)
    dt = time.perf_counter() - t0
    assert rc == 0
    per_proof = dt * (1 << (n_target - n_sample))
    return {
        "value": 1.0 / per_proof,
        "unit": "outer-sumcheck proofs/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle outer sumcheck, bit-serial portable multiply, 3 tables of 2^{n_sample} "
                   f"elements, 1 thread: {dt:.2f} s, scaled x{1 << (n_target - n_sample)} to 2^{n_target}"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log-n", type=int, default=24)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-sample-log-n", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["serve", "outer", "batch", "gkr", "inner", "bind", "chain", "products", "mt", "eval"],
                    default="serve",
                    help="serve: SHA-256 proving requests served with continuous batching (default); "
                         "outer: one 2^log_n outer sumcheck per step; batch: 256 outer sumchecks "
                         "of 2^12..2^22 elements per step, batched; gkr: product tree and layered sumcheck "
                         "over 2^22 leaves per step; inner: one inner sumcheck over 2 tables of 2^inner_log_n")
    ap.add_argument("--serve-rate", type=float, default=1000.0,
                    help="serve: Poisson arrivals per second for the reported p50/p99")
    ap.add_argument("--serve-rates", default="300,1000,2000,3000",
                    help="serve: the Poisson rates of the latency curve")
    ap.add_argument("--serve-requests", type=int, default=200, help="serve: requests per step")
    ap.add_argument("--serve-latency-requests", type=int, default=400,
                    help="serve: requests per Poisson run of the latency curve")
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="serve: worker processes of the full-core CPU baseline (0: min(16, cpu count))")
    ap.add_argument("--no-outer", action="store_true", help="serve: skip the outer 2^24 extra figures")
    ap.add_argument("--capacity-only", action="store_true",
                    help="serve: only the timed capacity steps (no latency curve, slots or plain runs; "
                         "for kernel profiles of the batched mode)")
    ap.add_argument("--serve-positions", type=int, default=0,
                    help="serve: rounds an engine step advances each proof by (0: the library's default)")
    ap.add_argument("--engine-stats", action="store_true",
                    help="serve: also time every engine step (device span by HIP events, host launch and "
                         "finish, wait) and report per-step averages")
    ap.add_argument("--chain-blocks", type=int, default=64)
    ap.add_argument("--serve-slots", type=int, default=192)
    ap.add_argument("--serve-engines", type=int, default=4,
                    help="serve: PIOP engines, one stream each (0: one per 8 slots)")
    ap.add_argument("--serve-preparers", type=int, default=12,
                    help="serve: host threads preparing requests (witness, assignment, Az/Bz/Cz)")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
    ap.add_argument("--inner-log-n", type=int, default=23,
                    help="inner: log2 table length (23: the column count of a 256-block SHA-256 chain)")
    ap.add_argument("--eval-log-n", type=int, default=26, help="eval: the multilinear's variables")
    ap.add_argument("--gkr-log-leaves", type=int, default=22)
    ap.add_argument("--gkr-log-groups", type=int, default=1,
                    help="gkr: log2 of the number of trees (1: full depth, two outputs; 15: 7 layers)")
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--batch-eager", action="store_true",
                    help="batch: launch the kernels directly instead of replaying the captured graph "
                         "(for kernel tracers that do not follow graph launches)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device twice (the known-size reference for "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path (a proof would shard "
              "its tables by high index bits, one exchange of 48 bytes per GPU and round). Use --gpus 1.",
              file=sys.stderr)
        return 2

    import bitz_amd as G

    if args.copy_calibration:
        G.init(0)
        src, dst = G.DeviceTable(length=1 << 24), G.DeviceTable(length=1 << 24)
        for _ in range(2):
            dst.copy_from(src)
        G.synchronize()
        log("bench: copied 2 x 256 MiB device to device")
        return 0

    if args.workload == "batch":
        return bench_batch(G, args)
    if args.workload == "gkr":
        return bench_gkr(G, args)
    if args.workload == "inner":
        return bench_inner(G, args)
    if args.workload == "serve":
        return bench_serve(G, args)
    if args.workload == "chain":
        return bench_chain(G, args)
    if args.workload == "bind":
        return bench_bind(G, args)
    if args.workload == "products":
        return bench_products(G, args)
    if args.workload == "mt":
        return bench_mt(G, args)
    if args.workload == "eval":
        return bench_eval(G, args)

    return bench_outer(G, args)


def run_outer(G, n: int, steps: int, warmup: int, seed: int) -> dict:
    """The isolated outer sumcheck: 2^n tables, rounds queued with the
    seeded challenges; returns the JSON line without its CPU baseline."""
    G.init(0)
    log(f"bench: outer sumcheck n={n}, generating inputs")
    a, b, c, tau = make_inputs(n, seed)
    lo, hi = G.make_equality_factors(tau)
    eq = (G.DeviceTable(lo), G.DeviceTable(hi))
    prods = G.R1csProductMles(G.DeviceTable(a), G.DeviceTable(b), G.DeviceTable(c))
    del a, b, c
    claim = G.outer_claim(eq, prods)
    log("bench: inputs resident in HBM; initial claim computed on device")

    def seeded_challenges(i: int) -> np.ndarray:
        t = G.Transcript.seeded(seed + i)
        return np.array([[(x := t.squeeze()) & (2**64 - 1), x >> 64] for _ in range(n)], dtype=np.uint64)

    # Challenges are inputs: drawn before any timing.
    warm_ch = [seeded_challenges(i) for i in range(warmup)]
    timed_ch = [seeded_challenges(1000 + i) for i in range(steps)]

    def step(ch: np.ndarray) -> float:
        """One proof with the seeded challenges: the rounds are queued back to
        back (the challenges do not depend on the messages)."""
        t0 = time.perf_counter()
        out = G.prove_outer_sumcheck_with_challenges(ch, claim, eq, prods)
        dt = time.perf_counter() - t0
        assert out.round_polynomials.shape[0] == n
        return dt

    def step_transcript(i: int) -> float:
        """One proof through a message-bound transcript: a host round trip
        per round for the challenge."""
        t = G.Transcript.duplex(b"bench/%d" % i)
        t0 = time.perf_counter()
        G.prove_outer_sumcheck(t, claim, eq, prods)
        return time.perf_counter() - t0

    for i in range(warmup):
        log(f"bench: warmup {i + 1}/{warmup}: {step(warm_ch[i]) * 1e3:.2f} ms")
    G.synchronize()
    # HIP events bracket only the grid-wide fold launches: one event pair per
    # proof spans its back-to-back fold launches.
    G.profile_enable(True)
    G.profile_only("outer_fold_next")
    times = []
    t_start = time.perf_counter()
    for i in range(steps):
        times.append(step(timed_ch[i]))
    G.synchronize()
    wall = time.perf_counter() - t_start
    for i, dt in enumerate(times):
        log(f"bench: step {i + 1}/{steps}: {dt * 1e3:.2f} ms")
    fold = G.profile_read("outer_fold_next")
    G.profile_only(None)
    G.profile_enable(False)
    tt = sorted(step_transcript(i) for i in range(max(3, steps // 2)))
    log(f"bench: transcript-bound proofs: {[round(x * 1e3, 2) for x in tt]} ms")

    N = 1 << n
    ms_per_step = wall / steps * 1e3
    achieved = fold["bytes"] / (fold["ms"] * 1e-3) / 1e9 if fold["ms"] else 0.0
    total_algo = 192.0 * N  # bytes the whole proof must move (read 3N, then fold rounds)
    line = {
        "metric": f"outer sumchecks proved/sec (eq*(A*B+C) over 2^{n} GF(2^128) elements per table), 1 MI355X",
        "value": steps / wall,
        "unit": "outer-sumcheck proofs/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms_per_step,
        "p50_ms": float(np.percentile(np.array(times) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(times) * 1e3, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "synthetic (uniform random tables and point from the seed)",
        "config": {
            "workload": f"outer sumcheck eq*(A*B+C), 3 tables of 2^{n} GF(2^128) elements, "
                        f"{n} rounds, challenges from the seed (rounds queued back to back)",
            "log_n": n,
            "elements_per_table": N,
            "parallelism": "single GPU",
        },
        "transcript_mode_p50_ms": float(np.percentile(np.array(tt) * 1e3, 50)),
        "elements_folded_per_s": 3 * (N - 1) * steps / wall,
        "proof_hbm_gbs": total_algo * steps / wall / 1e9,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_outer_fold (grid-wide fold rounds 1-5)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(n),
            "launches": fold["launches"],
            "kernel_ms_per_step": fold["ms"] / steps,
        },
    }
    del eq, prods
    return line


def bench_outer(G, args) -> int:
    line = run_outer(G, args.log_n, args.steps, args.warmup, args.seed)
    if not args.no_cpu:
        log(f"bench: CPU oracle sample at 2^{args.cpu_sample_log_n}")
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample_log_n, args.log_n)
    print(json.dumps(line), flush=True)
    return 0


def batch_workload(seed: int, count: int):
    """The batched workload's instances: num_vars uniform in [12, 22] from the
    seed; A, B, C of instance i are windows (at a seeded offset) of three
    uniform pools of 2^(max num_vars + 1) elements; tau and the challenges
    uniform.  Returns (pools, [(num_vars, offset, tau, challenges)])."""
    rng = np.random.default_rng(seed)
    sizes = [int(x) for x in rng.integers(12, 23, size=count)]
    pool_log = max(sizes) + 1
    pools = [rng.integers(0, 2**64, size=(1 << pool_log, 2), dtype=np.uint64) for _ in range(3)]
    insts = []
    for n in sizes:
        off = int(rng.integers(0, (1 << pool_log) - (1 << n) + 1))
        tau = rng.integers(0, 2**64, size=(n, 2), dtype=np.uint64)
        ch = rng.integers(0, 2**64, size=(n, 2), dtype=np.uint64)
        insts.append((n, off, tau, ch))
    return pools, insts


def bench_batch(G, args) -> int:
    """256 independent outer sumchecks, num_vars uniform in [12, 22] from the
    seed, each with A, B, C, tau and challenges as the single workload.  A
    step proves the whole batch in one bitz_prove_outer_batch call (a
    replay of the captured launch graph, tile shapes from the autotuner,
    which runs during the first warmup call); the same instances proved one
    call each give the plain single-instance rate."""
    G.init(0)
    host_pools, spec = batch_workload(args.seed, args.batch)
    sizes = [n for n, *_ in spec]
    log(f"bench: batch of {args.batch} outer sumchecks, n in [12, 22], total {sum(1 << n for n in sizes)} "
        f"elements per table")
    pools = [G.DeviceTable(p) for p in host_pools]
    del host_pools

    class View(G.DeviceTable):  # a window of a device pool (owns nothing)
        def __init__(self, pool, off, n):
            self.ptr, self.len, self.pool = pool.ptr + 16 * off, n, pool

        def free(self):
            pass

    insts = []
    for n, off, tau, ch in spec:
        prods = G.R1csProductMles(*(View(p, off, 1 << n) for p in pools))
        lo, hi = G.make_equality_factors(tau)
        eq = (G.DeviceTable(lo), G.DeviceTable(hi))
        claim = G.outer_claim(eq, prods)
        insts.append((ch, claim, eq, prods))
    folded = sum(3 * ((1 << n) - 1) for n in sizes)
    graph = not args.batch_eager
    # The instance records are marshalled once (a caller's own structures
    # over the C ABI); every step is one full bitz_prove_outer_batch call.
    batch = G.OuterBatch(insts, graph=graph)
    for i in range(args.warmup):
        t0 = time.perf_counter()
        batch.prove()
        log(f"bench: warmup {i + 1}: {(time.perf_counter() - t0) * 1e3:.2f} ms")
    G.synchronize()
    times = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        batch.prove()
        times.append(time.perf_counter() - t0)
    G.synchronize()
    wall = time.perf_counter() - t_start
    tile = G.outer_batch_config()
    # The same batch with the kernels launched directly (no graph).
    eager_batch = G.OuterBatch(insts, graph=False)
    t0 = time.perf_counter()
    for _ in range(2):
        eager_batch.prove()
    G.synchronize()
    eager = (time.perf_counter() - t0) / 2
    # The plain path: the same instances, one pipelined proof per call.
    t0 = time.perf_counter()
    for ch, claim, eq, prods in insts:
        G.prove_outer_sumcheck_with_challenges(ch, claim, eq, prods)
    G.synchronize()
    plain = time.perf_counter() - t0
    line = {
        "metric": "GF(2^128) elements folded per second, batched outer sumcheck",
        "value": folded * args.steps / wall,
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "p50_ms": float(np.percentile(np.array(times) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(times) * 1e3, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "synthetic (uniform random tables, points and challenges from the seed)",
        "config": {"workload": f"{args.batch} outer sumchecks, num_vars uniform in [12, 22] (seed {args.seed}), "
                               "challenges fixed in advance, one batched call per step",
                   "instances": args.batch, "elements_per_table": sum(1 << n for n in sizes)},
        "tile_shapes": {"thread_bits": tile[0], "tail_log": tile[1], "chosen_by": "autotuner"},
        "launch": "eager" if args.batch_eager else "captured graph",
        "eager_launch_ms_per_step": eager * 1e3,
        "plain_single_instance": {"ms_for_all_instances": plain * 1e3,
                                  "elements_folded_per_s": folded / plain},
        "speedup_vs_plain": (folded * args.steps / wall) / (folded / plain),
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_gkr(G, args) -> int:
    """2^L nonzero uniform leaves from the seed (default L = 22), 2^s trees
    (default s = 1: full depth, two outputs).  A step builds the product
    tree on the device (grand_product) and proves every layer
    (gpgkr_prove) with the seeded challenges known in advance, all rounds
    queued back to back; the transcript-driven form (one host round trip per
    round) is timed beside it."""
    L, s = args.gkr_log_leaves, args.gkr_log_groups
    G.init(0)
    rng = np.random.default_rng(args.seed)
    leaves = rng.integers(0, 2**64, size=(1 << L, 2), dtype=np.uint64)
    leaves[(leaves[:, 0] == 0) & (leaves[:, 1] == 0), 0] = 1  # nonzero leaves
    dev = G.DeviceTable(leaves)
    point = rng.integers(0, 2**64, size=(s, 2), dtype=np.uint64)
    count = sum(n + 1 for n in range(s, L))

    def seeded(i: int) -> np.ndarray:
        t = G.Transcript.seeded(args.seed + i)
        return np.array([[(x := t.squeeze()) & (2**64 - 1), x >> 64] for _ in range(count)], dtype=np.uint64)

    warm_ch = [seeded(i) for i in range(args.warmup)]
    timed_ch = [seeded(1000 + i) for i in range(args.steps)]
    log(f"bench: GKR over 2^{L} leaves, 2^{s} trees ({L - s} layers, {count - (L - s)} rounds)")

    def step(ch) -> float:
        t0 = time.perf_counter()
        out, wit = G.grand_product(dev, s)
        _, _, lp, claim = G.gpgkr_prove_with_challenges(point, ch, wit)
        dt = time.perf_counter() - t0
        assert len(lp) == L
        return dt

    for i in range(args.warmup):
        log(f"bench: warmup {i + 1}/{args.warmup}: {step(warm_ch[i]) * 1e3:.2f} ms")
    G.synchronize()
    G.profile_enable(True)
    G.profile_only("gkr_round_batch")
    times = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        times.append(step(timed_ch[i]))
    G.synchronize()
    wall = time.perf_counter() - t_start
    bind = G.profile_read("gkr_round_batch")
    G.profile_only(None)
    G.profile_enable(False)
    tt = []
    for i in range(3):
        t0 = time.perf_counter()
        _, wit = G.grand_product(dev, s)
        G.gpgkr_prove(G.Transcript.seeded(args.seed + 2000 + i), point, wit)
        tt.append(time.perf_counter() - t0)

    def keccak_p50(device: bool) -> float:
        """A Keccak transcript: challenges drawn on the device, or by the
        host transcript after each round's read-back."""
        G.set_device_fiat_shamir(device)
        ts = []
        for i in range(4):
            t0 = time.perf_counter()
            _, wit = G.grand_product(dev, s)
            G.gpgkr_prove(G.Transcript.keccak(b"bench/gkr/%d" % i), point, wit)
            ts.append(time.perf_counter() - t0)
        G.set_device_fiat_shamir(True)
        return float(np.percentile(np.array(ts[1:]) * 1e3, 50))

    keccak_device, keccak_host = keccak_p50(True), keccak_p50(False)
    bittable = bittable_shape(G, args)
    log(f"bench: Keccak transcript: {keccak_device:.2f} ms with device Fiat-Shamir, {keccak_host:.2f} ms through the host")
    achieved = bind["bytes"] / (bind["ms"] * 1e-3) / 1e9 if bind["ms"] else 0.0
    children = sum(1 << (n + 1) for n in range(s, L))
    line = {
        "metric": "GKR grand products proved/sec (product tree + layered sumcheck), 1 MI355X",
        "value": args.steps / wall,
        "unit": "grand-product proofs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "p50_ms": float(np.percentile(np.array(times) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(times) * 1e3, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "synthetic (uniform nonzero leaves, point and challenges from the seed)",
        "config": {"workload": f"GKR grand product over 2^{L} leaves, 2^{s} trees, {L - s} layers, "
                               "seeded challenges queued", "log_leaves": L, "log_groups": s},
        "leaves_per_s": (1 << L) * args.steps / wall,
        "transcript_mode_p50_ms": float(np.percentile(np.array(tt) * 1e3, 50)),
        "keccak_transcript_p50_ms": {"device_fiat_shamir": keccak_device, "host_round_trips": keccak_host},
        "bittable_t7_s15": bittable,
        "algorithmic_gb_per_s": 92.0 * children * args.steps / wall / 1e9,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_gkr_round_batch (grid-wide rounds of every large layer, one launch per round index)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "launches": bind["launches"],
            "kernel_ms_per_step": bind["ms"] / args.steps,
        },
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_gkr(20, L, s)
    print(json.dumps(line), flush=True)
    return 0


def bittable_shape(G, args, t: int = 7, s: int = 15) -> dict:
    """The grand product over a BitTable at the end-to-end prover's shape
    (2^t rows of images, 2^s trees: 2^(t+s) leaves, a uniform random table):
    what expanding the table into dense leaves costs (the k_bittable_leaves
    kernel and the HBM it takes) beside the tree build and a queued proof,
    against the memory the reference's lookup-table form would hold."""
    rng = np.random.default_rng(args.seed + 77)
    images = rng.integers(0, 2**64, size=(1 << t, 2), dtype=np.uint64)
    images[(images[:, 0] == 0) & (images[:, 1] == 0), 0] = 1
    packed = rng.integers(0, 2**64, size=(1 << (t + s)) // 64, dtype=np.uint64)
    point = rng.integers(0, 2**64, size=(s, 2), dtype=np.uint64)
    count = sum(n + 1 for n in range(s, t + s))
    tr = G.Transcript.seeded(args.seed + 3000)
    ch = np.array([[(x := tr.squeeze()) & (2**64 - 1), x >> 64] for _ in range(count)], dtype=np.uint64)
    for _ in range(2):
        _, wit = G.grand_product_bittable(images, packed, t, s)
        G.gpgkr_prove_with_challenges(point, ch, wit)
    G.synchronize()
    G.profile_enable(True)
    G.profile_only(None)
    builds, proves, reps = [], [], 5
    for _ in range(reps):
        t0 = time.perf_counter()
        _, wit = G.grand_product_bittable(images, packed, t, s)
        t1 = time.perf_counter()
        G.gpgkr_prove_with_challenges(point, ch, wit)
        builds.append(t1 - t0)
        proves.append(time.perf_counter() - t1)
    ex, ly = G.profile_read("bittable_leaves"), G.profile_read("gkr_layers")
    G.profile_enable(False)
    leaves = 1 << (t + s)
    # The lookup-table form: the table's bits, one 16-entry table per four
    # images, layer 3 and up in full, and round 1's bound children (as many
    # as layer 1 has nodes), lib.rs:44-78, 813-894.
    lut_bytes = leaves // 8 + 16 * 16 * ((1 << t) // 4) + 16 * (leaves // 4) + 16 * (leaves // 2)
    out = {
        "log_rows": t, "log_trees": s, "leaves": leaves,
        "expand_kernel_ms": ex["ms"] / max(ex["launches"], 1),
        "expand_gbs": ex["bytes"] / max(ex["ms"], 1e-9) / 1e6,
        "tree_layers_ms": ly["ms"] / max(ly["launches"], 1),
        "build_call_ms_p50": float(np.percentile(np.array(builds) * 1e3, 50)),
        "queued_proof_ms_p50": float(np.percentile(np.array(proves) * 1e3, 50)),
        "dense_tree_bytes": 2 * 16 * leaves,
        "lut_form_bytes": lut_bytes,
    }
    log(f"bench: BitTable t={t} s={s}: expansion {out['expand_kernel_ms'] * 1e3:.1f} us, tree "
        f"{out['tree_layers_ms'] * 1e3:.1f} us, proof {out['queued_proof_ms_p50']:.2f} ms")
    return out


def bench_inner(G, args) -> int:
    """The inner sumcheck sum_y D(y) h(y) over two uniform tables of 2^n
    elements (default n = 23), challenges from the seed known in advance (all
    rounds queued); the transcript-driven form is timed beside it."""
    n = args.inner_log_n
    G.init(0)
    rng = np.random.default_rng(args.seed)
    d = G.DeviceTable(rng.integers(0, 2**64, size=(1 << n, 2), dtype=np.uint64))
    w = G.DeviceTable(rng.integers(0, 2**64, size=(1 << n, 2), dtype=np.uint64))
    claim = 0  # the claim only shifts c1; the work is the same

    def seeded(i: int) -> np.ndarray:
        t = G.Transcript.seeded(args.seed + i)
        return np.array([[(x := t.squeeze()) & (2**64 - 1), x >> 64] for _ in range(n)], dtype=np.uint64)

    warm_ch = [seeded(i) for i in range(args.warmup)]
    timed_ch = [seeded(1000 + i) for i in range(args.steps)]
    log(f"bench: inner sumcheck over 2 tables of 2^{n}")
    for i in range(args.warmup):
        t0 = time.perf_counter()
        G.prove_inner_sumcheck_with_challenges(warm_ch[i], claim, d, w)
        log(f"bench: warmup {i + 1}/{args.warmup}: {(time.perf_counter() - t0) * 1e3:.2f} ms")
    G.synchronize()
    G.profile_enable(True)
    G.profile_only("inner_fold_lane")
    times = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        G.prove_inner_sumcheck_with_challenges(timed_ch[i], claim, d, w)
        times.append(time.perf_counter() - t0)
    G.synchronize()
    wall = time.perf_counter() - t_start
    fold = G.profile_read("inner_fold_lane")
    G.profile_only(None)
    G.profile_enable(False)
    tt = []
    for i in range(3):
        t0 = time.perf_counter()
        G.prove_inner_sumcheck(G.Transcript.seeded(args.seed + 2000 + i), claim, d, w)
        tt.append(time.perf_counter() - t0)
    achieved = fold["bytes"] / (fold["ms"] * 1e-3) / 1e9 if fold["ms"] else 0.0
    N = 1 << n
    line = {
        "metric": "inner sumchecks proved/sec, 1 MI355X",
        "value": args.steps / wall,
        "unit": "inner-sumcheck proofs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "p50_ms": float(np.percentile(np.array(times) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(times) * 1e3, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "synthetic (uniform tables and challenges from the seed)",
        "config": {"workload": f"inner sumcheck D*h, 2 tables of 2^{n} GF(2^128) elements, seeded challenges queued",
                   "log_n": n},
        "algorithmic_gb_per_s": 128.0 * N * args.steps / wall / 1e9,
        "transcript_mode_p50_ms": float(np.percentile(np.array(tt) * 1e3, 50)),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_inner_fold_lane (grid-wide fold rounds)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "launches": fold["launches"],
            "kernel_ms_per_step": fold["ms"] / args.steps,
        },
    }
    print(json.dumps(line), flush=True)
    return 0


def serve_mix(seed: int, count: int, rate: float):
    """count requests of the serving workload (blocks 2^U, U uniform in 0..8;
    message and chaining value uniform), Poisson arrivals at `rate`/s."""
    from bitz_amd import serving as S
    return S.poisson_requests(rate, count, seed)


def bench_serve(G, args) -> int:
    """The serving workload (see the module notes).  The CPU baseline runs
    first, before anything touches the GPU; then the circuits are prepared
    and the kernels warmed, --warmup untimed steps run, and --steps timed
    steps each serve a fresh set of --serve-requests requests, all queued at
    once, with continuous batching."""
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_serve(args)
    from bitz_amd import serving as S
    if args.serve_positions:
        G.set_engine_positions(args.serve_positions)
    srv = S.ProvingServer(hbm_budget=int(args.serve_budget_gib * (1 << 30)), slots=args.serve_slots)
    log("bench: preparing the nine SHA-256 circuits (1 .. 256 blocks)")
    rng = np.random.default_rng(args.seed + 1)
    for u in range(9):  # circuits and kernels warm, like a deployed service
        b = 1 << u
        srv.prove(10**6 + u, b, *S.random_input(rng, b))
    G.synchronize()

    engines = args.serve_engines or max(1, args.serve_slots // 8)

    def queued(seed: int):
        rs = serve_mix(seed, args.serve_requests, 1.0)
        for r in rs:
            r.arrival = 0.0
        return rs

    for i in range(args.warmup):
        t0 = time.perf_counter()
        srv.serve_batched(queued(args.seed + 500 + i), engines=engines, preparers=args.serve_preparers)
        log(f"bench: warmup {i + 1}/{args.warmup}: {(time.perf_counter() - t0) * 1e3:.1f} ms")
    # Steps: inputs drawn before the clock starts; HIP events on the engines'
    # streams time every batched inner fold launch.
    step_reqs = [queued(args.seed + i) for i in range(args.steps)]
    blocks = sum(r.blocks for rs in step_reqs for r in rs)
    G.synchronize()
    # Under rocprofv3 (which exports ROCPROFILER_LIBRARY_CTOR to the program)
    # the kernel tracer times the launches and the events stay off: a run
    # with both ended in a host fault inside the tracer's hipStreamSynchronize
    # on an engine thread.
    events = "ROCPROFILER_LIBRARY_CTOR" not in os.environ
    G.profile_enable(events)
    G.profile_only(None if args.engine_stats else "inner_fold_batch")
    step_ms = []
    t_start = time.perf_counter()
    for rs in step_reqs:
        t0 = time.perf_counter()
        srv.serve_batched(rs, engines=engines, preparers=args.serve_preparers)
        step_ms.append((time.perf_counter() - t0) * 1e3)
    G.synchronize()
    wall = time.perf_counter() - t_start
    fold = G.profile_read("inner_fold_batch")
    engine = None
    if args.engine_stats:
        st = {k: G.profile_read("piop_engine_" + k) for k in ("step", "host_launch", "wait", "host_finish")}
        n = max(st["step"]["launches"], 1)
        engine = {"steps": st["step"]["launches"], "proofs_per_step": st["step"]["bytes"] / n,
                  **{k + "_us_per_step": v["ms"] * 1e3 / n for k, v in st.items()},
                  "engine_streams": engines, "wall_us_per_step_per_engine": wall * 1e6 * engines / n}
    G.profile_only(None)
    G.profile_enable(False)
    queued_lat = np.array([r.latency for rs in step_reqs for r in rs]) * 1e3
    value = blocks / wall
    log(f"bench: {args.steps} steps, {blocks} blocks in {wall:.3f} s: {value:.0f} blocks/s")

    def poisson(mode: str, rate: float, seed: int) -> dict:
        rs = serve_mix(seed, args.serve_latency_requests, rate)
        t0 = time.perf_counter()
        if mode == "batched":
            srv.serve_batched(rs, engines=engines, preparers=args.serve_preparers)
        else:
            srv.serve(rs)
        span = max(r.done for r in rs) - min(r.arrival for r in rs)
        lat = np.array([r.latency for r in rs]) * 1e3
        b = sum(r.blocks for r in rs)
        return {"rate_per_s": rate, "offered_blocks_per_s": b / max(r.arrival for r in rs),
                "blocks_per_s": b / span, "p50_ms": float(np.percentile(lat, 50)),
                "p90_ms": float(np.percentile(lat, 90)), "p99_ms": float(np.percentile(lat, 99)),
                "peak_in_flight": srv.peak_in_flight, "peak_hbm_gib": srv.peak_bytes / (1 << 30),
                "wall_s": time.perf_counter() - t0}

    if args.capacity_only:
        print(json.dumps({"metric": METRIC, "value": value, "unit": "SHA-256 blocks/s", "steps": args.steps,
                          "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
                          "inner_fold_batch": fold, "engine": engine}), flush=True)
        return 0
    rates = sorted({float(x) for x in args.serve_rates.split(",") if x} | {args.serve_rate})
    curve = [poisson("batched", rate, args.seed + 100 + k) for k, rate in enumerate(rates)]
    for c in curve:
        log(f"bench: poisson {c['rate_per_s']:.0f}/s: p50 {c['p50_ms']:.2f} ms, p99 {c['p99_ms']:.2f} ms")
    head = next(c for c in curve if c["rate_per_s"] == args.serve_rate)
    # The concurrent-slots mode and the plain path on the first step's requests.
    rs = queued(args.seed)
    t0 = time.perf_counter()
    srv.serve(rs)
    slots_sat = sum(r.blocks for r in rs) / (time.perf_counter() - t0)
    slots_poisson = poisson("slots", args.serve_rate, args.seed + 100 + rates.index(args.serve_rate))
    rs = queued(args.seed)
    t0 = time.perf_counter()
    for r in rs:
        srv.prove(r.rid, r.blocks, r.message, r.iv)
    G.synchronize()
    plain = sum(r.blocks for r in rs) / (time.perf_counter() - t0)
    c256 = srv.circuit(256)
    achieved = fold["bytes"] / (fold["ms"] * 1e-3) / 1e9 if fold["ms"] else 0.0
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "SHA-256 blocks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "p50_ms": head["p50_ms"],
        "p99_ms": head["p99_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "synthetic (message bytes and chaining value per request uniform from the seed)",
        "config": {"workload": f"serving SHA-256 compression chains (config C5): a step serves {args.serve_requests} "
                               "requests of 2^U blocks (U uniform in 0..8) queued at once, with continuous batching "
                               f"({engines} PIOP engines, {args.serve_slots} slots, HBM budget "
                               f"{args.serve_budget_gib} GiB); each request: witness (host), packed assignment, "
                               "Az/Bz/Cz, prove_spartan_piop (duplex transcript), every table on the device. "
                               f"p50/p99: per-request latency under Poisson arrivals at {args.serve_rate:g}/s",
                   "requests_per_step": args.serve_requests, "blocks_timed": blocks,
                   "poisson_rate_per_s": args.serve_rate,
                   "circuit_256_blocks": {"rows": c256.rows, "columns": c256.columns, "nnz": c256.nnz}},
        "step_ms": {"p50": float(np.percentile(step_ms, 50)), "max": float(max(step_ms))},
        "saturated_latency_ms": {"p50": float(np.percentile(queued_lat, 50)),
                                 "p99": float(np.percentile(queued_lat, 99))},
        "latency_curve": curve,
        "slots_mode": {"blocks_per_s": slots_sat, "poisson": slots_poisson},
        "plain_one_at_a_time": {"blocks_per_s": plain},
        "speedup_vs_plain": value / plain,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_inner_fold_batch (the engines' batched inner-sumcheck fold rounds)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "launches": fold["launches"],
            "avg_launch_us": fold["ms"] * 1e3 / max(fold["launches"], 1),
            "algorithmic_bytes_per_launch": fold["bytes"] / max(fold["launches"], 1),
        },
    }
    if not args.no_outer:
        del srv
        o = run_outer(G, 24, 10, 3, args.seed)
        line["outer_2e24"] = {k: o[k] for k in ("value", "unit", "p50_ms", "p99_ms", "transcript_mode_p50_ms",
                                                "elements_folded_per_s", "proof_hbm_gbs", "roofline")}
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)
    return 0


def bench_chain(G, args) -> int:
    """The sha256-chain statement of the CLI (tooling/cli/src/circuits.rs:
    167-217) at --chain-blocks blocks (default 64): a padded message from the
    seed, its rows built and digested once (setup, not timed, as the
    reference's benchmark prepares its instance); a step is the proof's
    sumcheck stage: witness (host), the packed assignment, Az/Bz/Cz and
    prove_spartan_piop under the duplex transcript (one host round trip per
    round)."""
    import struct
    from bitz_amd import sha256 as SH
    G.init(0)
    B = args.chain_blocks
    rng = np.random.default_rng(args.seed)
    raw = rng.integers(0, 256, 64 * B - 9, dtype=np.uint8).tobytes()
    msg = raw + b"\x80" + struct.pack(">Q", 8 * len(raw))
    c = SH.Sha256Circuit(SH.CHAIN, B, statement=msg)
    t0 = time.perf_counter()
    matrices, digest = c.prepare()
    setup = time.perf_counter() - t0

    def step(i: int) -> float:
        t0 = time.perf_counter()
        w = c.witness(msg)
        h = G.PackedAssignment(w.h, c.columns, c.columns, device=True)
        products = matrices.r1cs_products(h)
        G.prove_spartan_piop(G.Transcript.duplex(b"chain/%d" % i), matrices, digest, products, h)
        return time.perf_counter() - t0

    for i in range(args.warmup):
        step(i)
    G.synchronize()
    times = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        times.append(step(1000 + i))
    G.synchronize()
    wall = time.perf_counter() - t_start
    lat = np.array(times) * 1e3
    line = {
        "metric": METRIC,
        "value": B * args.steps / wall,
        "unit": "SHA-256 blocks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "synthetic (message bytes from the seed, padded as SHA-256 pads)",
        "config": {"workload": f"sha256-chain statement, {B} blocks, one proof at a time: witness, packed assignment, "
                               "Az/Bz/Cz, prove_spartan_piop (duplex transcript)",
                   "blocks": B, "rows": c.rows, "columns": c.columns, "nnz": sum(c.nnz),
                   "row_vars": c.num_row_vars, "col_vars": c.num_column_vars, "setup_s": setup},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_sha256(SH.CHAIN, B, msg, args.seed)
    print(json.dumps(line), flush=True)
    return 0


class OracleCircuit:
    """A SHA-256 circuit for the oracle: the producer's matrices in the
    parity projection as oracle CSR matrices, and its statement digest."""

    def __init__(self, kind: int, blocks: int, statement: bytes | None = None):
        from bitz_amd import sha256 as SH
        from oracle import oracle as O
        self.c = SH.Sha256Circuit(kind, blocks, statement=statement)
        _, self.digest = self.c.prepare(matrices=False)
        ex = self.c.export()
        self.mats = [O.CsrMatrix(ex[k][0], ex[k][1], ex[k][3]) for k in "abc"]

    def prove(self, message: bytes, iv=None, label: bytes = b"cpu/sha256") -> None:
        """One request on the oracle: witness (the producer, host), the
        assignment, Az/Bz/Cz (the reference tests' multiply), and
        prove_spartan_piop's sequence (crates/spartan/src/piop.rs:67-112):
        digest, tau, outer sumcheck, rho, bind_and_batch, inner sumcheck."""
        from oracle import oracle as O
        c = self.c
        rv, cv = c.num_row_vars, c.num_column_vars
        w = c.witness(message, iv=iv)
        bits = np.unpackbits(w.h.view(np.uint8), bitorder="little")[: c.columns]
        h = np.zeros((1 << cv, 2), np.uint64)
        h[: c.columns, 0] = bits
        prods = [O.csr_multiply(m, c.rows, h, 1 << rv) for m in self.mats]
        t = O.DuplexTranscript(label)
        O.absorb(t, 0, self.digest)
        tau = np.array([[(x := O.squeeze(t)) & (2**64 - 1), x >> 64] for _ in range(rv)], np.uint64)
        lo, hi = O.eq_table(tau[: rv // 2]), O.eq_table(tau[rv // 2:])
        rc, outer = O.outer_sumcheck(t, 0, lo, hi, *prods)
        rho = O.squeeze(t)
        a_c, b_c, c_c = (O.to_int(x) for x in outer["claims"])
        claim = a_c ^ O.to_int(O.mul(rho, b_c)) ^ O.to_int(O.mul(O.to_int(O.mul(rho, rho)), c_c))
        d = O.bind_and_batch(self.mats, c.rows, rv, cv, outer["points"], rho)
        rc2, _ = O.inner_sumcheck(t, claim, d, h)
        assert rc == 0 and rc2 == 0


def _cpu_time_requests(blocks_list, seed: int, kernel: str) -> list:
    """Seconds per request for each block count in blocks_list (compression
    chains, message and chaining value from the seed), on this process."""
    from bitz_amd import sha256 as SH
    from oracle import oracle as O
    O.set_kernel(kernel)
    rng = np.random.default_rng(seed)
    circuits, out = {}, []
    for b in blocks_list:
        if b not in circuits:
            circuits[b] = OracleCircuit(SH.COMPRESSION_CHAIN, b)
        msg = rng.integers(0, 256, 64 * b, dtype=np.uint8).tobytes()
        iv = [int(x) for x in rng.integers(0, 2**32, 8, dtype=np.uint64)]
        t0 = time.perf_counter()
        circuits[b].prove(msg, iv)
        out.append(time.perf_counter() - t0)
    return out


def cpu_baseline_serve(args) -> dict:
    """The oracle proving the serving workload's requests on the host, with
    the bit-serial multiply the reference selects on x86
    (crates/field/src/gf128.rs:40-49), before the GPU is touched.

    1 thread: one request of each block count 1, 2, ..., 256 (the mix draws
    them with equal probability), timed one after another; requests of 1
    block, the configs' CPU case, 11 times (median).  The mix rate is the
    blocks of the first timed step's requests over the sum of their sizes'
    times.  Full cores: P worker processes each prove the same list of
    requests (16 blocks, 4 times) at once; their aggregate rate over the
    1-thread rate of the same list scales the mix rate."""
    import multiprocessing as mp
    kernel = "portable"
    log("bench: CPU baseline (oracle, bit-serial multiply), 1 thread")
    one = _cpu_time_requests([1] * 11, args.seed + 7, kernel)
    sizes = [1 << u for u in range(9)]
    per = dict(zip(sizes, _cpu_time_requests(sizes, args.seed + 8, kernel)))
    per[1] = float(np.median(one))
    mix = serve_mix(args.seed, args.serve_requests, 1.0)
    mix_blocks = sum(r.blocks for r in mix)
    mix_s = sum(per[r.blocks] for r in mix)
    rate1 = mix_blocks / mix_s
    workers = args.cpu_workers or min(16, os.cpu_count() or 1)
    sample = [16] * 4
    single = sum(_cpu_time_requests(sample, args.seed + 9, kernel))
    log(f"bench: CPU baseline, {workers} worker processes")
    ctx = mp.get_context("spawn")
    t0 = time.perf_counter()
    with ctx.Pool(workers) as pool:
        pool.starmap(_cpu_time_requests, [(sample, args.seed + 10 + w, kernel) for w in range(workers)])
    par = time.perf_counter() - t0
    agg = workers * sum(sample) / par
    scale = agg / (sum(sample) / single)
    return {"value": rate1, "unit": "SHA-256 blocks/s", "cores": 1, "kind": "port",
            "sample": (f"oracle (C restatement of the reference, bit-serial multiply), 1 thread: one compression "
                       f"chain request of each size 1..256 blocks ({sum(per[b] for b in sizes):.1f} s), "
                       f"weighted by the {args.serve_requests} requests ({mix_blocks} blocks) of the first timed "
                       "step; each request: witness, assignment, Az/Bz/Cz, outer sumcheck, bind_and_batch, "
                       "inner sumcheck under the duplex transcript"),
            "seconds_per_request": {str(b): per[b] for b in sizes},
            "c1_one_block_median_ms": 1e3 * per[1], "c1_runs": len(one),
            "full_cores": {"value": rate1 * scale, "cores": workers,
                           "sample": (f"{workers} processes each proving 4 requests of 16 blocks at once: "
                                      f"{agg:.1f} blocks/s aggregate, {scale:.2f}x the 1-thread rate of the same "
                                      "requests, applied to the mix rate (process start-up included)")}}


def cpu_baseline_sha256(kind: int, blocks: int, statement: bytes, seed: int, runs: int = 3) -> dict:
    """The oracle proving the given SHA-256 statement on one host thread with
    the bit-serial multiply, median of `runs`."""
    from oracle import oracle as O
    O.set_kernel("portable")
    oc = OracleCircuit(kind, blocks, statement=statement)
    times = []
    for i in range(runs):
        t0 = time.perf_counter()
        oc.prove(statement, label=b"cpu/%d" % i)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return {"value": blocks / med, "unit": "SHA-256 blocks/s", "cores": 1, "kind": "port",
            "sample": f"oracle, bit-serial multiply, 1 thread: median of {runs} proofs of the statement, {med:.2f} s"}


def bench_bind(G, args) -> int:
    """bind_and_batch over the positions of the 256-block SHA-256 compression
    chain (47104 rows, 2^23 columns, 13.9 M positions from the producer) with
    every value a uniform nonzero element (as a projection onto a large field
    leaves them), the output D left on the device.  The same call on the
    parity projection itself (where even coefficients map to zero and the
    device keeps only the odd ones) is reported beside it."""
    from bitz_amd import serving as S
    G.init(0)
    c = S.make_circuit(256)
    ex = c.circuit.export()
    rng = np.random.default_rng(args.seed)
    csr = []
    for k in "abc":
        st, co = ex[k][0], ex[k][1]
        vals = rng.integers(0, 2**64, size=(len(co), 2), dtype=np.uint64)
        vals[(vals[:, 0] == 0) & (vals[:, 1] == 0), 0] = 1
        csr.append((st, co, vals))
    generic = G.PreparedConstraintMatrices(c.rows, c.row_vars, c.col_vars, csr)
    nnz = c.nnz
    rx = G.DeviceTable(rng.integers(0, 2**64, size=(c.row_vars, 2), dtype=np.uint64))
    rho = 0x1234567890ABCDEF1234567890ABCDEF
    out = G.DeviceTable(length=1 << c.col_vars)

    def timed(m):
        for _ in range(args.warmup):
            m.bind_and_batch(rx, rho, device=True, out=out)
        G.synchronize()
        times = []
        t_start = time.perf_counter()
        for _ in range(args.steps):
            t0 = time.perf_counter()
            m.bind_and_batch(rx, rho, device=True, out=out)
            times.append(time.perf_counter() - t0)
        G.synchronize()
        return time.perf_counter() - t_start, np.array(times) * 1e3

    wall, times = timed(generic)
    pwall, ptimes = timed(c.matrices)
    live = int(sum(int((ex[k][2] & 1).sum()) for k in "abc"))
    algo = nnz * (4 + 1 + 16) + 16 * (1 << c.col_vars) + 16 * 4 * (1 << c.row_vars)
    line = {
        "metric": "bind_and_batch calls/sec, 1 MI355X",
        "value": args.steps / wall,
        "unit": "bind_and_batch/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "p50_ms": float(np.percentile(times, 50)),
        "p99_ms": float(np.percentile(times, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "positions of the 256-block SHA-256 compression chain (producer), uniform nonzero values, "
                "random row point",
        "config": {"workload": f"bind_and_batch: {c.rows} rows (2^{c.row_vars}), 2^{c.col_vars} columns, "
                               f"{nnz} nonzeros", "nnz": nnz},
        "algorithmic_gb_per_s": algo * args.steps / wall / 1e9,
        "algorithmic_bytes": algo,
        "parity_projection": {"ms_per_step": pwall / args.steps * 1e3, "p50_ms": float(np.percentile(ptimes, 50)),
                              "nonzero_values": live},
    }
    print(json.dumps(line), flush=True)
    return 0


# Published MI355X peaks: dense int8 matrix ops, and the vector ALU (256 CUs
# x 64 lanes at 2.4 GHz) that the table products run on.
INT8_PEAK_TOPS = 5000.0
VALU_LANE_OPS_PER_S = 256 * 64 * 2.4e9


def bench_products(G, args) -> int:
    """GF(2^128) products per second of each product primitive over 2^log_n
    device-resident elements (one launch each, timed with HIP events on the
    launch's stream): the comb, the lane-table product and its two-operand
    form, the four-per-wave row product, and the two ways of multiplying by
    a fixed element, its 4-bit constant table on the vector ALU and its F2
    matrix on the int8 matrix cores (the fold-by-challenge product).  The
    constant-table product is bound by HBM at 32 B per product; the MFMA
    one by the matrix cores (128 x 128 x 2 = 32768 int8 ops per product)
    and by the VALU work that lays the bits out as bytes and gathers the
    parities back."""
    G.init(0)
    n = 1 << args.log_n
    rng = np.random.default_rng(args.seed)
    a = G.DeviceTable(rng.integers(0, 2**64, size=(n, 2), dtype=np.uint64))
    b = G.DeviceTable(rng.integers(0, 2**64, size=(n, 2), dtype=np.uint64))
    out = G.DeviceTable(length=n)
    modes = [("mul_comb", 0), ("mul_lane", G.MUL_LANE), ("mul_lane2", G.MUL_LANE2), ("mul_rows", G.MUL_ROWS),
             ("mul_const", G.MUL_CONST), ("mul_mfma", G.MUL_MFMA)]
    res = {}
    for name, mode in modes:
        G.profile_enable(False)
        for _ in range(args.warmup):
            G.mul_batch_device(a, b, out, mode)
        G.profile_enable(True)
        G.profile_only(name)
        for _ in range(args.steps):
            G.mul_batch_device(a, b, out, mode)
        st = G.profile_read(name)
        ms = st["ms"] / max(st["launches"], 1)
        rate = n / (ms * 1e-3)
        res[name] = {"ms_per_launch": ms, "products_per_s": rate,
                     "hbm_gbs": st["bytes"] / max(st["launches"], 1) / (ms * 1e-3) / 1e9,
                     "valu_lane_ops_per_product_at_peak": VALU_LANE_OPS_PER_S / rate}
        log(f"bench: {name}: {ms:.3f} ms per 2^{args.log_n} products, {rate / 1e9:.1f} G/s")
    G.profile_enable(False)
    const, mfma = res["mul_const"], res["mul_mfma"]
    mfma_tops = mfma["products_per_s"] * 32768 / 1e12
    line = {
        "metric": "GF(2^128) multiplications by a fixed element per second, 1 MI355X",
        "value": const["products_per_s"],
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": const["ms_per_launch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "synthetic (uniform elements from the seed)",
        "config": {"workload": f"a[i] * c over 2^{args.log_n} device-resident elements, one launch per step"},
        "roofline": {"bound": "hbm", "kernel": "k_mul_const", "achieved": const["hbm_gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": const["hbm_gbs"] / HBM_PEAK_GBS,
                     "traffic": None},
        "mfma_roofline": {"bound": "mfma", "kernel": "k_mul_mfma", "achieved": mfma_tops,
                          "peak": INT8_PEAK_TOPS, "unit": "TOP/s (int8)", "frac": mfma_tops / INT8_PEAK_TOPS},
        "mfma_vs_const_table": mfma["products_per_s"] / const["products_per_s"],
        "primitives": res,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_eval(G, args) -> int:
    """DenseMultilinearExtension::evaluate (crates/poly/src/mle.rs:158-172) of
    a 2^eval_log_n table resident on the device at a uniform point: the
    table folded two variables a pass through the challenges' constant
    tables.  The roofline is the first pass (every input read once, a quarter
    as many outputs written: 20 B per input), timed with HIP events on its
    stream; the oracle (bit-serial multiply, one host core) evaluates a 2^20
    table as the CPU baseline, scaled to the same size."""
    from oracle import oracle as O
    n = args.eval_log_n
    cpu = None
    if not args.no_cpu:
        O.set_kernel("portable")
        rng = np.random.default_rng(3)
        t = rng.integers(0, 2**64, size=(1 << 20, 2), dtype=np.uint64)
        p = rng.integers(0, 2**64, size=(20, 2), dtype=np.uint64)
        t0 = time.perf_counter()
        O.mle_evaluate(t, p)
        dt = time.perf_counter() - t0
        cpu = {"value": 1.0 / (dt * (1 << (n - 20))), "unit": "evaluations/s", "cores": 1, "kind": "port",
               "sample": f"oracle mle_evaluate, bit-serial portable multiply, 2^20 entries, 1 thread: {dt:.2f} s, "
                         f"scaled x{1 << (n - 20)} to 2^{n}"}
    G.init(0)
    rng = np.random.default_rng(args.seed)
    table = G.DeviceTable(length=1 << n)
    chunk = 1 << 22
    for o in range(0, 1 << n, chunk):  # uniform entries, uploaded in pieces
        table.copy_from(G.DeviceTable(rng.integers(0, 2**64, size=(min(chunk, 1 << n), 2), dtype=np.uint64)), o)
    points = [rng.integers(0, 2**64, size=(n, 2), dtype=np.uint64) for _ in range(args.warmup + args.steps)]
    for i in range(args.warmup):
        G.mle_evaluate_device(table, points[i])
    G.profile_enable(True)
    G.profile_only("mle_fold_pass")
    times = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        G.mle_evaluate_device(table, points[args.warmup + i])
        times.append(time.perf_counter() - t0)
    wall = time.perf_counter() - t_start
    st = G.profile_read("mle_fold_pass")
    G.profile_enable(False)
    ms = st["ms"] / max(st["launches"], 1)
    achieved = st["bytes"] / max(st["launches"], 1) / (ms * 1e-3) / 1e9
    line = {
        "metric": "multilinear evaluations/sec (2^%d entries, device-resident), 1 MI355X" % n,
        "value": args.steps / wall,
        "unit": "evaluations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "p50_ms": float(np.percentile(np.array(times) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(times) * 1e3, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "synthetic (uniform table and points from the seed)",
        "config": {"workload": f"DenseMultilinearExtension::evaluate over 2^{n} device-resident entries", "log_n": n},
        "roofline": {"bound": "hbm", "kernel": "k_fold2_t (the first fold pass: two variables)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "kernel_ms": ms},
        "cpu_baseline": cpu,
    }
    log(f"bench: evaluate 2^{n}: {line['p50_ms']:.3f} ms, first pass {ms * 1e3:.1f} us at {achieved:.0f} GB/s")
    print(json.dumps(line), flush=True)
    return 0


def bench_mt(G, args) -> int:
    """M^T over the SHA-256 circuits: apply (r M, one challenge per row of M)
    and transpose_inner (the row weight split as a low and a high eq table),
    with device-resident inputs and outputs, each kernel timed with HIP events
    on its stream.  Circuits: the 2 KiB block-aligned message that the
    reference's benchmark uses (crates/circuit/benches/sha256_matrix_products.rs)
    and the 256-block compression chain of the serving workload.  The oracle
    runs apply on the 2 KiB circuit on one host core as the CPU baseline."""
    from bitz_amd import sha256 as S
    from oracle import oracle as O
    G.init(0)
    rng = np.random.default_rng(args.seed)
    cases = {}
    cpu = None
    for name, kind, blocks in (("sha256_2kib", S.BLOCK_ALIGNED, 32), ("chain_256", S.COMPRESSION_CHAIN, 256)):
        c = S.Sha256Circuit(kind, blocks)
        offsets, rows = c.m_transpose()
        mt = G.MaterializedMTranspose(c.columns, offsets, rows)
        nnz, cols = len(rows), len(offsets) - 1
        ch = rng.integers(0, 2**64, size=(c.columns, 2), dtype=np.uint64)
        dch, dout = G.DeviceTable(ch), G.DeviceTable(length=cols)
        vars_ = max(c.columns - 1, 0).bit_length()
        lo = (vars_ + 1) // 2
        drw = G.DeviceTable(rng.integers(0, 2**64, size=(1 << lo, 2), dtype=np.uint64))
        dcw = G.DeviceTable(rng.integers(0, 2**64, size=(1 << (vars_ - lo), 2), dtype=np.uint64))
        res = {"m_rows": c.columns, "m_columns": cols, "nonzeros": nnz}
        for kern, call in (("mt_apply", lambda: mt.apply_device(dch, dout)),
                           ("mt_transpose_inner", lambda: mt.transpose_inner_device(drw, dcw, dout))):
            G.profile_enable(False)
            for _ in range(args.warmup):
                call()
            G.profile_enable(True)
            G.profile_only(kern)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                call()
            wall = time.perf_counter() - t0
            st = G.profile_read(kern)
            ms = st["ms"] / max(st["launches"], 1)
            gbs = st["bytes"] / max(st["launches"], 1) / (ms * 1e-3) / 1e9
            res[kern] = {"kernel_ms": ms, "call_ms": wall / args.steps * 1e3, "algorithmic_gbs": gbs,
                         "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": st["bytes"] / max(st["launches"], 1)}
            log(f"bench: {name} {kern}: {ms * 1e3:.1f} us per launch, {gbs:.0f} GB/s algorithmic")
        G.profile_enable(False)
        if name == "sha256_2kib" and not args.no_cpu:
            runs = []
            for _ in range(5):
                t0 = time.perf_counter()
                O.mt_apply(offsets, rows, ch)
                runs.append(time.perf_counter() - t0)
            cpu = {"value": 1.0 / float(np.median(runs)), "unit": "apply/s", "cores": 1, "kind": "port",
                   "sample": f"oracle mt_apply on the 2 KiB circuit ({cols} columns, {nnz} nonzeros), "
                             f"1 thread, median of 5: {float(np.median(runs)) * 1e3:.2f} ms"}
        cases[name] = res
    head = cases["sha256_2kib"]["mt_apply"]
    line = {
        "metric": "M^T apply calls/sec (SHA-256 2 KiB circuit), 1 MI355X",
        "value": 1e3 / head["kernel_ms"],
        "unit": "apply/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["kernel_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128",
        "data": "SHA-256 circuits from the producer, uniform challenges and weights from the seed",
        "config": {"workload": "MaterializedMTranspose apply and transpose_inner on the SHA-256 2 KiB "
                               "block-aligned circuit and the 256-block compression chain"},
        "roofline": {"bound": "hbm", "kernel": "k_mt_apply", "achieved": head["algorithmic_gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": head["frac"], "traffic": None},
        "cases": cases,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline_gkr(l_sample: int, l_target: int, s: int) -> dict:
    """The oracle (bit-serial multiply) building and proving a 2^l_sample-leaf
    product tree on one host core, scaled linearly to 2^l_target leaves."""
    from oracle import oracle as O
    O.set_kernel("portable")
    rng = np.random.default_rng(7)
    leaves = rng.integers(1, 2**64, size=(1 << l_sample, 2), dtype=np.uint64)
    s = min(s, l_sample - 1)
    t0 = time.perf_counter()
    store, _ = O.gkr_fold_layers(leaves, s)
    O.gkr_prove(O.SeededTranscript(1), rng.integers(0, 2**64, size=(s, 2), dtype=np.uint64), store, l_sample)
    dt = time.perf_counter() - t0
    per = dt * (1 << (l_target - l_sample))
    return {
        "value": 1.0 / per,
        "unit": "grand-product proofs/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle product tree + GKR prove, bit-serial portable multiply, 2^{l_sample} leaves, 2^{s} "
                   f"trees, 1 thread: {dt:.2f} s, scaled x{1 << (l_target - l_sample)} to 2^{l_target}"),
    }


if __name__ == "__main__":
    sys.exit(main())
