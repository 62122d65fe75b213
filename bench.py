"""Benchmark of the MI355X WHIR prover hot path.

Workload (one GPU): WHIR commit + prove of one multilinear polynomial with
2^24 KoalaBear coefficients — the reference tests' seed-1 polynomial
Poly::rand(&mut rng(1), 24), generated by the library (whir_poly_rand) before
timing — rate 1/2
(log_inv_rate 1), folding factor 4, initial reduction 3, JohnsonBound
soundness at 123 bits, one evaluation claim at a point drawn from the
transcript (the reference's run_whir_poseidon flow,
crates/whir/src/test.rs:174-254). A step is one full proof: RS encode +
Poseidon2 Merkle commit, claim evaluation, SVO sumcheck, 4 WHIR rounds, final
polynomial and STIR openings, proof bytes copied back to the host. The input
polynomial is resident in HBM before the timed region starts.

The JSON line also carries
  * roofline: the dominant kernel, the base-tree Poseidon2-24 leaf hash
    (one permutation per lane, VALU-bound). Its launches are timed live with HIP
    events on the prover's stream, in a second pass of the timed proofs (the
    events stay out of the timed loop); achieved =
    algorithmic 32-bit multiply operations per second: per W24 permutation 982
    modular multiplications (8 x 24 full-round cubes and 23 partial-round cubes
    at 2 each, 23 x 24 internal-diagonal products) of 4 multiplies each,
    against the MI355X VALU lane rate (256 CUs x 4 SIMD x 16 lanes x 2.4 GHz).
    Beside it, the issued-instruction utilisation: the kernel's VALU
    instructions per lane (from its own gfx950 code object,
    whir_amd/kernel_stats.json) per second against the same peak. `traffic`
    comes from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary.
  * roofline_dft: the base RS encode (three NTT passes), timed the same way.
    It issues 4.9 G VALU lane instructions (rocprofv3 SQ_INSTS_VALU, read from
    the same summary) against 0.74 GB of HBM traffic, so its bound is the VALU
    issue rate (125 us at peak) rather than HBM (92 us at 8 TB/s); both
    views are in the object.
  * cpu_baseline: the CPU oracle (a multithreaded C++ restatement of the
    reference prover, "port") proving the same polynomial on the host cores
    this process may use (its CPU affinity, capped by OMP_NUM_THREADS where set:
    the box's CPU share), median / min / max over 5 runs; its proof bytes are
    compared with the GPU's (bit_exact).
  * autotune: the tuner's choices for the headline shape's transforms and
    trees (whir_ctx_set_autotune: candidates timed at workspace creation,
    before the timed region; the batched and served legs reuse them).
  * batched: 64 independent jobs, variable counts uniform in 16..22 (sizes
    log-uniform), rates 1/2 and 1/4 mixed, folding 4, seeded; grouped by shape
    into lockstep batches (every kernel launched once per batch, one grid
    slice per job) running side by side on 12 streams, the whole set captured
    as one hipGraph (whir_multi); against the same 64 jobs one by one on the
    plain path (eager launches, one stream, workspaces cached per shape), with
    the proof server (one captured graph per job over several streams) beside
    them. Proof bytes of every job are compared between all three.
  * serving: Poisson arrivals of jobs with 16..24 variables, rate 1/2 or 1/4,
    1..4 evaluation claims, seeded; offered at three loads (fractions of the
    plain path's measured capacity on this job mix); proofs/s and p50/p99
    latency (arrival to proof on the host) for the server and for the plain
    per-job FIFO path; every job's bytes compared between the two.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FOLDING, RATE, IR, SOUNDNESS, SECURITY, N_POINTS = 4, 1, 3, 1, 123, 1
METRIC = "commit+prove ms @2^24 vars; batched proofs/s & p99 vs plain GPU, bit-exact"
VALU_PEAK_T = 256 * 4 * 16 * 2.4e9 / 1e12  # T lane-instructions / s
HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "whir_prove_k24_pmc.txt")


MODMULS_PER_P24 = 8 * 24 * 2 + 23 * 2 + 23 * 24  # 982: cubes of full and partial rounds, internal diagonal
MULOPS_PER_MODMUL = 4


def synthetic_poly(seed, k):
    """Seeded coefficients, uniform in [0, p), as Montgomery words (the batched
    and served legs' many jobs)."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, 0x7F000001, size=1 << k, dtype=np.uint64).astype(np.uint32)


def host_cpu():
    """(CPU model, logical CPUs of the host, CPUs this process may use)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit():
        usable = min(usable, int(omp))
    return model, os.cpu_count(), usable


def read_traffic(key="traffic_bytes_per_launch"):
    """HBM bytes per launch from the committed rocprofv3 PMC summary."""
    try:
        with open(PMC_SUMMARY) as fh:
            for line in fh:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def cpu_baseline(poly, k, gpu_proof, threads, runs=5):
    from oracle import pyoracle as O

    O.set_threads(threads)
    times, same = [], True
    for _ in range(runs):
        t0 = time.perf_counter()
        proof, cp, root = O.whir_prove(k, FOLDING, RATE, IR, SOUNDNESS, SECURITY, N_POINTS, poly)
        times.append(time.perf_counter() - t0)
        same &= proof == gpu_proof
    rc, _ = O.whir_verify(k, FOLDING, RATE, IR, SOUNDNESS, SECURITY, N_POINTS, gpu_proof)
    model, n_host, _ = host_cpu()
    return {
        "value": round(float(np.median(times)) * 1e3, 1),
        "unit": "ms",
        "cores": threads,
        "kind": "port",
        "sample": (f"{runs} full commit+prove runs at k={k} of the same polynomial by the CPU oracle "
                   "(C++/OpenMP restatement of the reference prover, not Plonky3); value = median"),
        "min_ms": round(min(times) * 1e3, 1),
        "max_ms": round(max(times) * 1e3, 1),
        "runs": runs,
        "host_logical_cpus": n_host,
        "cpu_model": model,
        "bit_exact": bool(same),
        "oracle_verifier_accepts_gpu_proof": rc == 0,
    }


class PlainPath:
    """The plain GPU path run job by job: one stream, eager launches, a cached
    workspace per proof shape, the host waits for each proof."""

    def __init__(self, ctx):
        self.ctx = ctx
        self.provers = {}

    def prove(self, prm, d_poly):
        import whir_amd

        key = tuple(getattr(prm, f) for f, _ in prm._fields_)
        p = self.provers.get(key)
        if p is None:
            p = self.provers[key] = whir_amd.Prover(prm, self.ctx)
        p.enqueue(d_poly)
        return p.result()[0]

    def close(self):
        for p in self.provers.values():
            p.close()
        self.provers = {}


def _percentile(xs, q):
    return float(np.percentile(np.asarray(xs), q)) if len(xs) else None


def _upload_jobs(specs, poly_of):
    import whir_amd

    bufs = []
    for i, s in enumerate(specs):
        poly = poly_of(i, s)
        d = whir_amd.api.DeviceBuffer(poly.nbytes)
        d.upload(poly)
        bufs.append(d)
    return bufs


def _drain(srv, specs, bufs):
    """Submit every job at once and wait for all; returns the proofs in job order."""
    ids = [srv.submit(p, b.ptr) for p, b in zip(specs, bufs)]
    pending = set(ids)
    while pending:
        pending.difference_update(srv.poll())
    return [srv.take(j, 1 << 22)[0] for j in ids]


def batched_leg(ctx, slots, reps=3, n_jobs=64, seed=4, streams=12, queue_groups=(1, 2)):
    """64 independent proofs of mixed sizes: grouped by shape into lockstep
    batches side by side, the whole set one captured hipGraph (whir_multi);
    beside it the proof server (one graph per job over `slots` streams) and the
    plain path job by job."""
    import whir_amd

    rng = np.random.default_rng(seed)
    ks = rng.integers(16, 23, n_jobs)
    rates = rng.choice([1, 2], n_jobs)
    specs = [whir_amd.params(int(k), FOLDING, int(r), IR, SOUNDNESS, SECURITY, 1) for k, r in zip(ks, rates)]
    bufs = _upload_jobs(specs, lambda i, s: synthetic_poly(1000 + i, s.k))
    plain = PlainPath(ctx)
    srv = whir_amd.Server(slots, 64 << 30, ctx=ctx)
    multi = whir_amd.MultiBatch(specs, streams=streams, ctx=ctx, graph=1)
    # the same set with the commit work queue: every job's base LDE, leaves and
    # wide tree layers from persistent kernels over one device job table
    multi_q = {g: whir_amd.MultiBatch(specs, streams=streams, ctx=ctx, graph=1, commit_queue=g)
               for g in queue_groups}
    ref = [plain.prove(p, b.ptr) for p, b in zip(specs, bufs)]  # warm-up: workspaces
    _drain(srv, specs, bufs)  # warm-up: workspaces + graph capture

    def run_multi(m):
        m.enqueue([b.ptr for b in bufs])
        return [m.result(i)[0] for i in range(n_jobs)]

    run_multi(multi)  # warm-up: graph capture
    for m in multi_q.values():
        run_multi(m)
    t_plain, t_srv, t_multi, same = [], [], [], True
    t_queue = {g: [] for g in multi_q}
    for _ in range(reps):
        whir_amd.lib().whir_device_synchronize()
        t0 = time.perf_counter()
        got = [plain.prove(p, b.ptr) for p, b in zip(specs, bufs)]
        t_plain.append(time.perf_counter() - t0)
        same &= got == ref
        t0 = time.perf_counter()
        got = _drain(srv, specs, bufs)
        t_srv.append(time.perf_counter() - t0)
        same &= got == ref
        t0 = time.perf_counter()
        got = run_multi(multi)
        t_multi.append(time.perf_counter() - t0)
        same &= got == ref
        for g, m in multi_q.items():
            t0 = time.perf_counter()
            got = run_multi(m)
            t_queue[g].append(time.perf_counter() - t0)
            same &= got == ref
    queue_jobs = {g: m.queue_jobs for g, m in multi_q.items()}
    for m in multi_q.values():
        m.close()
    multi.close()
    srv.close()
    plain.close()
    tp, ts, tm = min(t_plain), min(t_srv), min(t_multi)
    tq = {g: min(v) for g, v in t_queue.items()}
    gbest = min(tq, key=tq.get)
    return {
        "workload": f"{n_jobs} jobs, k uniform in [16,22] (seed {seed}), rates 1/2+1/4 mixed, folding 4, 1 claim",
        "jobs_by_k": {int(k): int((ks == k).sum()) for k in sorted(set(ks.tolist()))},
        "method": (f"lockstep batches per shape ({len(set((int(k), int(r)) for k, r in zip(ks, rates)))} shapes), "
                   f"side by side on {streams} streams, the whole set one captured hipGraph"),
        "plain_proofs_per_s": round(n_jobs / tp, 1),
        "batched_proofs_per_s": round(n_jobs / tm, 1),
        "speedup": round(tp / tm, 3),
        "plain_ms": round(tp * 1e3, 2),
        "batched_ms": round(tm * 1e3, 2),
        "commit_queue": {
            "method": ("the same set with the commit work queue: every job's base LDE, leaf hashes and wide "
                       "tree layers from persistent kernels draining a device job table (one launch per "
                       "stage), the batches in G groups whose queues run one after another, each group's "
                       "openings overlapping the next group's commit; one captured hipGraph"),
            "jobs_in_queue": queue_jobs[gbest],
            "groups": gbest,
            "proofs_per_s": round(n_jobs / tq[gbest], 1),
            "ms": round(tq[gbest] * 1e3, 2),
            "proofs_per_s_by_groups": {g: round(n_jobs / t, 1) for g, t in sorted(tq.items())},
        },
        "server_slots": slots,
        "server_proofs_per_s": round(n_jobs / ts, 1),
        "server_ms": round(ts * 1e3, 2),
        "reps": reps,
        "bit_identical": bool(same),
        "tuned_plans_for_batches": sum(1 for t in ctx.tuning() if t.get("jobs", 1) > 1),
    }


def _serving_specs(n, seed):
    import whir_amd

    rng = np.random.default_rng(seed)
    ks = rng.integers(16, 25, n)
    rates = rng.choice([1, 2], n)
    npts = rng.integers(1, 5, n)
    return [whir_amd.params(int(k), FOLDING, int(r), IR, SOUNDNESS, SECURITY, int(m))
            for k, r, m in zip(ks, rates, npts)]


def _run_plain_arrivals(plain, specs, bufs, arrivals):
    done = []
    proofs = []
    t0 = time.perf_counter()
    for p, b, a in zip(specs, bufs, arrivals):
        while time.perf_counter() - t0 < a:
            pass
        proofs.append(plain.prove(p, b.ptr))
        done.append(time.perf_counter() - t0)
    return np.asarray(done), proofs


def _run_server_arrivals(srv, specs, bufs, arrivals):
    n = len(specs)
    done = np.zeros(n)
    idx = {}
    i, finished = 0, 0
    t0 = time.perf_counter()
    while finished < n:
        now = time.perf_counter() - t0
        while i < n and arrivals[i] <= now:
            idx[srv.submit(specs[i], bufs[i].ptr)] = i
            i += 1
        for jid in srv.poll():
            done[idx[jid]] = time.perf_counter() - t0
            finished += 1
    proofs = [None] * n
    for jid, j in idx.items():
        proofs[j] = srv.take(jid, 1 << 22)[0]
    return done, proofs


def serving_leg(ctx, slots, n_jobs=240, seed=5, loads=(0.5, 1.0, 1.5)):
    """Poisson arrivals at three offered loads; server versus plain FIFO."""
    import whir_amd

    specs = _serving_specs(n_jobs, seed)
    # three seeded polynomials per variable count, shared by the jobs of that size
    pool = {}

    def poly_of(i, s):
        key = (s.k, i % 3)
        if key not in pool:
            pool[key] = synthetic_poly(2000 + 10 * s.k + i % 3, s.k)
        return pool[key]

    bufs = _upload_jobs(specs, poly_of)
    pool.clear()
    plain = PlainPath(ctx)
    srv = whir_amd.Server(slots, 64 << 30, ctx=ctx)
    ref = [plain.prove(p, b.ptr) for p, b in zip(specs, bufs)]  # warm-up
    _drain(srv, specs, bufs)
    t0 = time.perf_counter()
    for p, b in zip(specs, bufs):
        plain.prove(p, b.ptr)
    mean_service = (time.perf_counter() - t0) / n_jobs
    plain_capacity = 1.0 / mean_service
    rng = np.random.default_rng(seed + 1)
    out = {
        "workload": f"{n_jobs} jobs, k uniform in [16,24], rate 1/2 or 1/4, 1..4 claims (seed {seed}), Poisson arrivals",
        "server_slots": slots,
        "plain_capacity_proofs_per_s": round(plain_capacity, 1),
        "loads": [],
    }
    same = True
    # cold start: a fresh server at load 1.0, so every shape's first job pays
    # its workspace allocation and graph capture inside the served latency
    cold = whir_amd.Server(slots, 64 << 30, ctx=ctx)
    arr = np.cumsum(np.random.default_rng(seed + 2).exponential(1.0 / plain_capacity, n_jobs))
    arr -= arr[0]
    done, proofs = _run_server_arrivals(cold, specs, bufs, arr)
    cold.close()
    lat = (done - arr) * 1e3
    out["server_cold_start_load_1.0"] = {
        "proofs_per_s": round(n_jobs / done.max(), 1),
        "p50_ms": round(_percentile(lat, 50), 3),
        "p99_ms": round(_percentile(lat, 99), 3),
        "shapes": len({tuple(getattr(p, f) for f, _ in p._fields_) for p in specs}),
    }
    same &= proofs == ref
    for load in loads:
        lam = load * plain_capacity
        arrivals = np.cumsum(rng.exponential(1.0 / lam, n_jobs))
        arrivals -= arrivals[0]
        row = {"offered_load": load, "offered_proofs_per_s": round(lam, 1)}
        for name, fn, obj in (("plain", _run_plain_arrivals, plain), ("server", _run_server_arrivals, srv)):
            done, proofs = fn(obj, specs, bufs, arrivals)
            lat = (done - arrivals) * 1e3
            row[name] = {
                "proofs_per_s": round(n_jobs / done.max(), 1),
                "p50_ms": round(_percentile(lat, 50), 3),
                "p99_ms": round(_percentile(lat, 99), 3),
            }
            same &= proofs == ref
        out["loads"].append(row)
    out["bit_identical"] = bool(same)
    srv.close()
    plain.close()
    for b in bufs:
        b.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--k", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default NTT / Merkle shapes instead of the tuner's per-size choices")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--slots", type=int, default=6,
                    help="proof-server streams (batched and serving legs); 6 measured best on MI355X")
    ap.add_argument("--serve-jobs", type=int, default=240)
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a device-to-device copy of this size (for HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration_mib:
        import whir_amd

        n = args.copy_calibration_mib << 20
        a, b = whir_amd.api.DeviceBuffer(n), whir_amd.api.DeviceBuffer(n)
        for _ in range(3):
            whir_amd.lib().whir_memcpy_d2d(b.ptr, a.ptr, n, None)
        whir_amd.lib().whir_device_synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only; --gpus {args.gpus} is not supported "
              "(a proof runs on a single GPU; batched proving scales by running one "
              "process per GPU).", file=sys.stderr)
        sys.exit(2)

    import whir_amd
    from whir_amd._lib import whir_params

    ctx = whir_amd.Context(0, 1000, autotune=not args.no_autotune)
    params = whir_params(args.k, FOLDING, RATE, IR, SOUNDNESS, SECURITY, N_POINTS)
    prover = whir_amd.Prover(params, ctx)
    poly = whir_amd.poly_rand(1, args.k)  # Poly::rand(&mut rng(1), k), test.rs:213-219
    prover.upload(poly)  # resident in HBM before timing

    for _ in range(args.warmup):
        prover.enqueue()
        proof, cp, root = prover.result()

    whir_amd.lib().whir_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prover.enqueue()
        proof, cp, root = prover.result()
    whir_amd.lib().whir_device_synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps

    # the leaf-hash and base-DFT kernel times for the roofline: HIP events
    # around those launches, in a second run of the same proofs (the events
    # stay out of the timed run above)
    prover.set_timers(True)
    for _ in range(args.steps):
        prover.enqueue()
        tproof = prover.result()[0]
    leaf_ms, leaf_n, perms = prover.timer(prover.TIMER_LEAVES_BASE)
    dft_ms, dft_n, dft_words = prover.timer(prover.TIMER_DFT_BASE)
    prover.set_timers(False)
    assert tproof == proof, "the timed proofs and the event-timed proofs differ"

    # the same proof as one captured hipGraph per launch (no per-kernel timers)
    prover.set_graph(True)
    for _ in range(2):
        prover.enqueue()
        gproof = prover.result()[0]
    whir_amd.lib().whir_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        prover.enqueue()
        gproof = prover.result()[0]
    whir_amd.lib().whir_device_synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    prover.set_graph(False)
    try:  # written by the build from the library's own code object
        with open(os.path.join(ROOT, "whir_amd", "kernel_stats.json")) as fh:
            kstats = json.load(fh)
    except OSError:
        kstats = {}
    instr = kstats.get("leaves16_valu_per_lane")
    leaf_s = leaf_ms / leaf_n / 1e3
    achieved = perms * MODMULS_PER_P24 * MULOPS_PER_MODMUL / leaf_s / 1e12
    issued = instr * perms / leaf_s / 1e12 if instr else None
    roofline = {
        "kernel": "merkle_leaves_kernel<16> (Poseidon2-24 leaf hash of the base codeword)",
        "bound": "valu",
        "achieved": round(achieved, 2),
        "peak": round(VALU_PEAK_T, 2),
        "unit": "T int32-mul-op/s (982 modmul x 4 per W24 permutation, algorithmic)",
        "frac": round(achieved / VALU_PEAK_T, 3),
        "traffic": read_traffic(),
        "algorithmic_bytes": perms * (64 + 32),
        "avg_launch_us": round(leaf_s * 1e6, 1),
        "perms_per_launch": perms,
        "perms_per_s": round(perms / leaf_s, 0),
        "issued_valu_instr_per_perm": instr,
        "issued_valu_T_per_s": round(issued, 2) if issued else None,
        "issued_valu_frac": round(issued / VALU_PEAK_T, 3) if issued else None,
    }
    dft_s = dft_ms / dft_n / 1e3
    # the PMC summary's counts are for the default plan of the base transform
    # (three passes, a radix-16 register pass first, two words per line);
    # with another tuned plan only the timing-based fields apply
    base_plan = next((t for t in (ctx.tuning() if not args.no_autotune else [])
                      if t["kind"] == "ntt" and t["log_rows"] == args.k - FOLDING + RATE
                      and t["log_cols"] == FOLDING and t.get("jobs", 1) == 1), None)
    default_plan = base_plan is None or (base_plan["two_pass"] == 0 and base_plan["first_pass_log"] == 4
                                         and base_plan["log_vec"] == 1)
    dft_traffic = read_traffic("ntt_base_traffic_bytes") if default_plan else None
    dft_valu = read_traffic("ntt_base_valu_lane_instr") if default_plan else None
    dft_valu_t = dft_valu / dft_s / 1e12 if dft_valu else None
    roofline_dft = {
        "kernel": "base RS encode (coset-form LDE 2^20 -> 2^21 rows x 16 columns; "
                  + ("3 NTT passes)" if default_plan else f"tuned plan {base_plan})"),
        "bound": "valu",
        "achieved": round(dft_valu_t, 2) if dft_valu_t else None,
        "peak": round(VALU_PEAK_T, 2),
        "unit": "T VALU lane-instr/s",
        "frac": round(dft_valu_t / VALU_PEAK_T, 3) if dft_valu_t else None,
        "valu_lane_instr": dft_valu,
        "traffic": dft_traffic,
        "traffic_GBps": round(dft_traffic / dft_s / 1e9, 1) if dft_traffic else None,
        "traffic_frac_of_hbm_peak": round(dft_traffic / dft_s / 1e9 / HBM_PEAK_GBS, 3) if dft_traffic else None,
        "algorithmic_bytes": dft_words * 4,
        "algorithmic_GBps": round(dft_words * 4 / dft_s / 1e9, 1),
        "avg_launch_us": round(dft_s * 1e6, 1),
    }

    out = {
        "metric": METRIC,
        "value": round(ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"whir commit+prove, 2^{args.k} KoalaBear coeffs, rate 1/2, folding 4, ir 3, JB-123, 1 ext claim",
            "k": args.k,
            "proof_bytes": len(proof),
            "proofs_per_s": round(1e3 / ms, 2),
            "graph_ms_per_step": round(graph_ms, 4),
            "graph_proof_identical": gproof == proof,
        },
        "roofline": roofline,
        "roofline_dft": roofline_dft,
    }
    if not args.no_autotune:
        out["autotune"] = [t for t in ctx.tuning() if t["kind"] == "ntt" and t["n_in"] >= (1 << 12)
                           or t["kind"] == "merkle" and t["log_rows"] >= 15][:12]
    prover.close()
    if not args.no_batched:
        out["batched"] = batched_leg(ctx, args.slots)
    if not args.no_serving:
        out["serving"] = serving_leg(ctx, args.slots, n_jobs=args.serve_jobs)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_cpu()[2]
        out["cpu_baseline"] = cpu_baseline(poly, args.k, proof, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
