"""The measured configurations and the reference's own test grid through the GPU
path, each proof compared byte for byte with the CPU oracle:

* the headline job: 2^24 coefficients (the seed-1 polynomial of
  crates/whir/src/test.rs:213), rate 1/2, folding 4, ir 3, JohnsonBound-123,
  one claim; root, bytes and checkpoint equal, the oracle verifier accepts;
* the reference's test_whir_poseidon grid (crates/whir/src/test.rs:256-282:
  JB/CB/UD x k 4..12 x folding 1..3 x initial_reduction 1..folding x 1..4
  points, 32-bit security) and the committed golden fixtures
  (tests/golden/oracle_proofs.json);
* a batch shaped like the 64-job workload (k 16..22, rates 1/2 and 1/4) through
  the proof server, and a Poisson-timed stream shaped like the serving workload
  (k 16..24, 1..4 claims): every job's bytes equal the oracle's and the job's
  proof alone.
"""
import hashlib
import json
import os
import time

import numpy as np
import pytest

from tests.stack_specs import rand_polys, stacked_k

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _first_diff(a, b):
    n = min(len(a), len(b))
    for i in range(n):
        if a[i] != b[i]:
            return i
    return n if len(a) != len(b) else -1


def test_headline_k24(gpu_ctx, oracle):
    import whir_amd

    cfg = (24, 4, 1, 3, 1, 123, 1)
    k, f, rate, ir, s, sec, n = cfg
    poly = oracle.rand_poly(1, k)
    print("prove k=24 rate 1/2 folding 4 ir 3 JB-123, 1 claim", flush=True)
    gproof, gcp, groot = whir_amd.Whir(k, f, rate, ir, s, sec, ctx=gpu_ctx).prove(poly, n)
    oproof, ocp, oroot = oracle.whir_prove(*cfg, poly)
    assert np.array_equal(groot, oroot)
    d = _first_diff(gproof, oproof)
    assert d == -1, f"proof differs at byte {d}"
    assert gcp == ocp
    rc, vcp = oracle.whir_verify(*cfg, gproof)
    assert rc == 0 and vcp == gcp


GRID = [(s, k, f, ir, n) for s in (1, 2, 0) for k in range(4, 13) for f in range(1, 4) for ir in range(1, f + 1)
        for n in range(1, 5)]


@pytest.mark.parametrize("s", [1, 2, 0], ids=["JB", "CB", "UD"])
def test_reference_grid(gpu_ctx, oracle, s):
    import whir_amd

    cases = [c for c in GRID if c[0] == s]
    print(f"grid soundness={s}: {len(cases)} proofs", flush=True)
    for (_, k, f, ir, n) in cases:
        poly = oracle.rand_poly(1, k)
        w = whir_amd.Whir(k, f, 1, ir, s, 32, ctx=gpu_ctx)
        gproof, gcp, _ = w.prove(poly, n)
        oproof, ocp, _ = oracle.whir_prove(k, f, 1, ir, s, 32, n, poly)
        assert gproof == oproof and gcp == ocp, (s, k, f, ir, n)


def test_golden_fixtures(gpu_ctx, oracle):
    import whir_amd

    with open(os.path.join(HERE, "golden", "oracle_proofs.json")) as fh:
        data = json.load(fh)
    for g in data["proofs"]:
        poly = oracle.rand_poly(g["poly_seed"], g["k"])
        print(f"golden k={g['k']} f={g['folding']} rate={g['rate']} points={g['n_points']}", flush=True)
        w = whir_amd.Whir(g["k"], g["folding"], g["rate"], g["initial_reduction"], g["soundness"],
                          g["security_level"], ctx=gpu_ctx)
        proof, cp, root = w.prove(poly, g["n_points"])
        assert len(proof) == g["proof_len"]
        assert hashlib.sha256(proof).hexdigest() == g["proof_sha256"]
        assert cp == g["checkpoint"] and [int(x) for x in root] == g["root"]
    for g in data["stacked"]:
        k_polys = g["k_polys"]
        claims = [(int(a), bool(b)) for a, b in g["claims"]]
        assert stacked_k(k_polys) == g["k"]
        polys = rand_polys(oracle, k_polys, g["poly_seed"])
        print(f"golden stacked k={g['k']} polys={len(k_polys)} claims={len(claims)}", flush=True)
        pr = whir_amd.StackedProver(whir_amd.params(g["k"], g["folding"], g["rate"], g["initial_reduction"],
                                                    g["soundness"], g["security_level"], len(claims)),
                                    k_polys, claims, ctx=gpu_ctx)
        try:
            pr.upload(polys)
            pr.enqueue()
            proof, cp, root = pr.result()
        finally:
            pr.close()
        assert hashlib.sha256(proof).hexdigest() == g["proof_sha256"]
        assert cp == g["checkpoint"] and [int(x) for x in root] == g["root"]


def _server_run(gpu_ctx, specs, bufs, arrivals=None, slots=6):
    """Submits every job (at its arrival time when given) and polls until all
    have finished; returns the proofs in job order."""
    import whir_amd

    srv = whir_amd.Server(slots, 64 << 30, ctx=gpu_ctx)
    try:
        ids = [None] * len(specs)
        done = set()
        i = 0
        t0 = time.perf_counter()
        deadline = t0 + 120.0
        while len(done) < len(specs):
            now = time.perf_counter() - t0
            while i < len(specs) and (arrivals is None or arrivals[i] <= now):
                ids[i] = srv.submit(specs[i], bufs[i].ptr)
                i += 1
            done.update(srv.poll())
            assert time.perf_counter() < deadline, "server did not finish"
        return [srv.take(j, 1 << 23)[0] for j in ids]
    finally:
        srv.close()


def _jobs(oracle, cfgs, seed0):
    import whir_amd

    specs, bufs, polys = [], [], []
    for i, (k, rate, n) in enumerate(cfgs):
        poly = oracle.rand_poly(seed0 + i, k)
        d = whir_amd.api.DeviceBuffer(poly.nbytes)
        d.upload(poly)
        specs.append(whir_amd.params(k, 4, rate, 3, 1, 123, n))
        bufs.append(d)
        polys.append(poly)
    return specs, bufs, polys


def test_batched_mix_against_oracle(gpu_ctx, oracle):
    """A batch shaped like the 64-job workload: k over 16..22, rates 1/2 and 1/4."""
    cfgs = [(16, 1, 1), (22, 2, 1), (17, 2, 1), (21, 1, 1), (18, 1, 1), (20, 2, 1), (19, 1, 1), (16, 2, 1),
            (22, 1, 1), (18, 2, 1)]
    specs, bufs, polys = _jobs(oracle, cfgs, 300)
    print(f"batched mix: {len(cfgs)} jobs", flush=True)
    got = _server_run(gpu_ctx, specs, bufs)
    for (k, rate, n), poly, proof in zip(cfgs, polys, got):
        want = oracle.whir_prove(k, 4, rate, 3, 1, 123, n, poly)[0]
        assert proof == want, (k, rate)


def test_served_stream_against_oracle_and_alone(gpu_ctx, oracle):
    """A Poisson-timed stream shaped like the serving workload: k up to 24, 1..4
    claims; each job equals the oracle's proof and its proof alone."""
    import whir_amd

    cfgs = [(16, 1, 3), (24, 1, 1), (19, 2, 4), (17, 1, 2), (23, 2, 1), (20, 1, 3), (16, 2, 2), (21, 1, 4)]
    specs, bufs, polys = _jobs(oracle, cfgs, 500)
    rng = np.random.default_rng(17)
    arrivals = np.cumsum(rng.exponential(0.002, len(cfgs)))
    arrivals -= arrivals[0]
    print(f"served stream: {len(cfgs)} jobs over {arrivals[-1] * 1e3:.1f} ms", flush=True)
    got = _server_run(gpu_ctx, specs, bufs, arrivals)
    for (k, rate, n), spec, buf, poly, proof in zip(cfgs, specs, bufs, polys, got):
        alone = whir_amd.Prover(spec, gpu_ctx)
        try:
            alone.enqueue(buf.ptr)
            assert proof == alone.result()[0], (k, rate, n)
        finally:
            alone.close()
        assert proof == oracle.whir_prove(k, 4, rate, 3, 1, 123, n, poly)[0], (k, rate, n)


def test_batched_mix_commit_queue_against_oracle(gpu_ctx, oracle):
    """The 64-job workload's shape mix (k 16..22, rates 1/2 and 1/4, shapes
    repeated) through a mixed-shape set with the commit work queue, one
    captured graph: every job equals the oracle's proof."""
    import whir_amd

    cfgs = [(16, 1, 1), (22, 2, 1), (17, 2, 1), (21, 1, 1), (18, 1, 1), (20, 2, 1), (19, 1, 1), (16, 2, 1),
            (22, 1, 1), (18, 2, 1), (16, 1, 1), (19, 1, 1)]
    specs, bufs, polys = _jobs(oracle, cfgs, 800)
    print(f"batched mix, commit queue: {len(cfgs)} jobs", flush=True)
    m = whir_amd.MultiBatch(specs, streams=6, ctx=gpu_ctx, graph=1, commit_queue=2)
    try:
        assert m.queue_jobs == len(cfgs)
        m.enqueue([d.ptr for d in bufs])
        got = [m.result(i)[0] for i in range(len(cfgs))]
    finally:
        m.close()
    for (k, rate, n), poly, proof in zip(cfgs, polys, got):
        assert proof == oracle.whir_prove(k, 4, rate, 3, 1, 123, n, poly)[0], (k, rate)
