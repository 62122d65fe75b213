"""Graph replay and the proof server.

* A hipGraph-captured proof, replayed several times with eager work and
  synchronisations in between, gives the eager proof's bytes every time (and a
  new input gives that input's bytes).
* The server (continuous batching over streams, device-resident workspaces,
  HBM-gated admission) returns for every job exactly the bytes the job gets
  when proven alone, also when the HBM budget forces jobs to wait and
  workspaces to be evicted; invalid shapes and impossible budgets fail cleanly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _alone(oracle, cfg, seed):
    import whir_amd

    k, f, rate, ir, s, sec, n = cfg
    poly = oracle.rand_poly(seed, k)
    return poly, whir_amd.Whir(k, f, rate, ir, s, sec).prove(poly, n)[0]


def test_graph_replays_match_eager(gpu_ctx, oracle):
    import whir_amd

    cfg = (12, 3, 1, 2, 1, 32, 2)
    k, f, rate, ir, s, sec, n = cfg
    poly = oracle.rand_poly(11, k)
    p = whir_amd.Prover(whir_amd.params(k, f, rate, ir, s, sec, n), gpu_ctx)
    p.upload(poly)
    print("graph: eager reference", flush=True)
    p.enqueue()
    eager = p.result()[0]
    assert eager == oracle.whir_prove(k, f, rate, ir, s, sec, n, poly)[0]
    p.set_graph(True)
    rng = np.random.default_rng(0)
    for rep in range(4):
        print(f"graph replay {rep}", flush=True)
        p.enqueue()
        assert p.result()[0] == eager
        # eager work and synchronisation between replays
        whir_amd.dft_batch(rng.integers(0, 1000, (64, 4)).astype(np.uint32), 8, ctx=gpu_ctx)
    poly2 = oracle.rand_poly(12, k)
    d = whir_amd.api.DeviceBuffer(poly2.nbytes)
    d.upload(poly2)
    p.enqueue(d.ptr)
    assert p.result()[0] == oracle.whir_prove(k, f, rate, ir, s, sec, n, poly2)[0]


MIX = [
    (8, 2, 1, 1, 1, 32, 1),
    (10, 3, 1, 2, 1, 32, 2),
    (12, 4, 1, 3, 1, 123, 1),
    (12, 4, 2, 3, 1, 123, 3),
    (13, 4, 1, 3, 1, 123, 4),
    (10, 3, 1, 2, 1, 32, 2),
    (14, 4, 2, 3, 1, 123, 1),
    (12, 4, 1, 3, 1, 123, 1),
    (9, 2, 2, 2, 2, 32, 2),
    (11, 3, 1, 3, 0, 32, 1),
]


@pytest.mark.parametrize("slots,budget", [(4, 8 << 30), (3, 0)])
def test_server_mix_equals_alone(gpu_ctx, oracle, slots, budget):
    import whir_amd

    expected, bufs, prms = [], [], []
    for i, cfg in enumerate(MIX):
        poly, proof = _alone(oracle, cfg, 100 + i)
        expected.append(proof)
        d = whir_amd.api.DeviceBuffer(poly.nbytes)
        d.upload(poly)
        bufs.append(d)
        prms.append(whir_amd.params(*cfg))
    if budget == 0:  # room for about two of the largest workspaces: forces waits and evictions
        budget = 2 * max(whir_amd.prover_bytes_for(p) for p in prms) + 1
    print(f"server slots={slots} budget={budget}", flush=True)
    srv = whir_amd.Server(slots, budget, ctx=gpu_ctx)
    ids = [srv.submit(p, d.ptr) for p, d in zip(prms, bufs)]
    done = set()
    for _ in range(200000):
        done.update(srv.poll())
        if len(done) == len(ids):
            break
    assert len(done) == len(ids)
    for jid, want in zip(ids, expected):
        got, _ = srv.take(jid, 1 << 22)
        assert got == want
    running, queued, held = srv.stats()
    assert running == 0 and queued == 0 and held <= budget
    srv.close()


def test_server_errors(gpu_ctx, oracle):
    import whir_amd

    srv = whir_amd.Server(2, 1 << 20, ctx=gpu_ctx)  # 1 MiB: no workspace fits
    poly = oracle.rand_poly(1, 12)
    d = whir_amd.api.DeviceBuffer(poly.nbytes)
    d.upload(poly)
    with pytest.raises(whir_amd.WhirError):
        srv.submit(whir_amd.params(12, 13, 1, 3), d.ptr)  # folding > k
    jid = srv.submit(whir_amd.params(12, 4, 1, 3), d.ptr)
    done = []
    for _ in range(1000):
        done += srv.poll()
        if done:
            break
    assert done == [jid]
    with pytest.raises(whir_amd.WhirError) as e:
        srv.take(jid, 1 << 20)
    assert e.value.code == 5  # WHIR_EOOM
    srv.close()


def test_server_poll_reports_every_job_with_max_one(gpu_ctx, oracle):
    """poll(max=1): every finished job is reported exactly once, by later polls
    if need be, and its proof equals the job's proof alone."""
    import ctypes

    import whir_amd
    from whir_amd._lib import lib

    cfgs = [(10, 3, 1, 2, 1, 32, 1), (11, 3, 1, 2, 1, 32, 2), (12, 4, 1, 3, 1, 123, 1), (10, 3, 1, 2, 1, 32, 1),
            (9, 2, 2, 2, 2, 32, 2)]
    expected, bufs, prms = [], [], []
    for i, cfg in enumerate(cfgs):
        poly, proof = _alone(oracle, cfg, 400 + i)
        expected.append(proof)
        d = whir_amd.api.DeviceBuffer(poly.nbytes)
        d.upload(poly)
        bufs.append(d)
        prms.append(whir_amd.params(*cfg))
    srv = whir_amd.Server(4, 8 << 30, ctx=gpu_ctx)
    try:
        ids = [srv.submit(p, d.ptr) for p, d in zip(prms, bufs)]
        one = (ctypes.c_int64 * 1)()
        n = ctypes.c_int()
        reported = []
        for _ in range(400000):
            assert lib().whir_server_poll(srv.ptr, one, 1, ctypes.byref(n)) == 0
            assert n.value in (0, 1)
            if n.value:
                reported.append(one[0])
            if len(reported) == len(ids):
                break
        assert sorted(reported) == sorted(ids)
        # nothing is reported twice
        assert lib().whir_server_poll(srv.ptr, one, 1, ctypes.byref(n)) == 0 and n.value == 0
        for jid, want in zip(ids, expected):
            assert srv.take(jid, 1 << 22)[0] == want
    finally:
        srv.close()
