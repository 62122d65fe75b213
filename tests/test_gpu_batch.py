"""Lockstep batches (whir_batch_*): `jobs` proofs of one shape as one kernel
sequence, each kernel launched once over all jobs. Every job's proof must be
the oracle's proof of that job's polynomial (run_whir_poseidon's flow,
crates/whir/src/test.rs:174-254) and the job's proof alone; the shapes cover
every kernel family of the prover (round 0's fold with code, later rounds on
the grid, in the sumcheck tail and whole in one workgroup, several claims,
unique decoding without OOD claims); graph replays give the same bytes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [
    # (k, folding, rate, ir, soundness, security, n_points), jobs
    ((10, 2, 1, 1, 1, 32, 1), 3),
    ((12, 3, 1, 2, 1, 32, 2), 4),
    ((11, 3, 1, 3, 0, 32, 1), 2),     # unique decoding: no OOD claims
    ((14, 4, 2, 3, 1, 123, 3), 3),
    ((18, 3, 1, 3, 1, 123, 2), 2),    # a later round of 2^12 on the grid, sumcheck tail
    ((17, 4, 1, 3, 1, 123, 1), 5),
]


@pytest.mark.parametrize("cfg,jobs", SHAPES, ids=lambda v: str(v))
def test_batch_jobs_match_oracle(gpu_ctx, oracle, cfg, jobs):
    import whir_amd

    k, f, rate, ir, s, sec, n = cfg
    polys = [oracle.rand_poly(700 + 13 * j + k, k) for j in range(jobs)]
    print(f"batch k={k} f={f} rate={rate} points={n} jobs={jobs}", flush=True)
    b = whir_amd.Batch(whir_amd.params(*cfg), jobs, ctx=gpu_ctx)
    try:
        for j, poly in enumerate(polys):
            b.upload(j, poly)
        b.enqueue()
        got = [b.result(j) for j in range(jobs)]
    finally:
        b.close()
    for j, poly in enumerate(polys):
        want = oracle.whir_prove(*cfg, poly)
        assert got[j][0] == want[0], f"job {j} differs from the oracle"
        assert got[j][1] == want[1] and np.array_equal(got[j][2], want[2])
    alone = whir_amd.Whir(k, f, rate, ir, s, sec, ctx=gpu_ctx).prove(polys[-1], n)
    assert alone[0] == got[-1][0]


def test_batch_graph_replays_and_new_inputs(gpu_ctx, oracle):
    import whir_amd

    cfg = (12, 3, 1, 2, 1, 32, 2)
    jobs = 3
    polys = [oracle.rand_poly(40 + j, 12) for j in range(jobs)]
    want = [oracle.whir_prove(*cfg, p)[0] for p in polys]
    b = whir_amd.Batch(whir_amd.params(*cfg), jobs, ctx=gpu_ctx, graph=True)
    bufs = []
    try:
        for rep in range(3):
            print(f"batch graph replay {rep}", flush=True)
            for j, p in enumerate(polys):
                b.upload(j, p)
            b.enqueue()
            assert [b.result(j)[0] for j in range(jobs)] == want
            whir_amd.dft_batch(np.arange(256, dtype=np.uint32).reshape(64, 4), 8, ctx=gpu_ctx)  # eager work between
        # inputs copied from other device buffers into the captured graph's inputs
        new = [oracle.rand_poly(90 + j, 12) for j in range(jobs)]
        for p in new:
            d = whir_amd.api.DeviceBuffer(p.nbytes)
            d.upload(p)
            bufs.append(d)
        b.enqueue([d.ptr for d in bufs])
        assert [b.result(j)[0] for j in range(jobs)] == [oracle.whir_prove(*cfg, p)[0] for p in new]
    finally:
        b.close()


def test_batch_of_one_is_the_plain_prover(gpu_ctx, oracle):
    import whir_amd

    cfg = (13, 4, 1, 3, 1, 123, 1)
    poly = oracle.rand_poly(5, 13)
    b = whir_amd.Batch(whir_amd.params(*cfg), 1, ctx=gpu_ctx)
    try:
        b.upload(0, poly)
        b.enqueue()
        got = b.result(0)
    finally:
        b.close()
    assert got[0] == whir_amd.Whir(13, 4, 1, 3, 1, 123, ctx=gpu_ctx).prove(poly, 1)[0]


@pytest.mark.parametrize("graph", [0, 1, 2])
def test_multi_shape_set_matches_oracle(gpu_ctx, oracle, graph):
    """Jobs of mixed shapes grouped into lockstep batches side by side (one
    captured graph for the whole set in graph mode), replayed twice."""
    import whir_amd

    cfgs = [(12, 4, 1, 3, 1, 123, 1), (14, 4, 2, 3, 1, 123, 1), (12, 4, 1, 3, 1, 123, 1), (13, 4, 1, 3, 1, 123, 2),
            (14, 4, 2, 3, 1, 123, 1), (12, 4, 1, 3, 1, 123, 1), (16, 4, 1, 3, 1, 123, 1), (10, 2, 1, 1, 1, 32, 3),
            (13, 4, 1, 3, 1, 123, 2)]
    polys = [oracle.rand_poly(900 + i, c[0]) for i, c in enumerate(cfgs)]
    want = [oracle.whir_prove(*c, p)[0] for c, p in zip(cfgs, polys)]
    bufs = []
    for p in polys:
        d = whir_amd.api.DeviceBuffer(p.nbytes)
        d.upload(p)
        bufs.append(d)
    print(f"multi-shape set: {len(cfgs)} jobs, graph={graph}", flush=True)
    m = whir_amd.MultiBatch([whir_amd.params(*c) for c in cfgs], streams=3, ctx=gpu_ctx, graph=graph)
    try:
        for rep in range(2):
            m.enqueue([d.ptr for d in bufs])
            got = [m.result(i)[0] for i in range(len(cfgs))]
            for i in range(len(cfgs)):
                assert got[i] == want[i], (rep, i, cfgs[i])
    finally:
        m.close()


def test_copying_results_equal_the_views(gpu_ctx, oracle):
    """whir_*_result (copy into the caller's buffer) and whir_*_result_view
    (a pointer into the library's pinned copy) give the same proof, checkpoint
    and root, for a prover, a batch and a mixed-shape set; a buffer one byte
    short is refused."""
    import ctypes

    import whir_amd
    from whir_amd._lib import check

    L = gpu_ctx.L
    cfg = (12, 4, 1, 3, 1, 123, 1)
    polys = [oracle.rand_poly(60 + j, 12) for j in range(2)]

    def copied(fn, *args):
        ln = ctypes.c_size_t()
        cp = ctypes.c_uint32()
        root = np.zeros(8, np.uint32)
        check(fn(*args, None, 0, ctypes.byref(ln), ctypes.byref(cp), root.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
              "len")
        buf = (ctypes.c_uint8 * ln.value)()
        assert fn(*args, buf, ln.value - 1, None, None, None) == 1  # WHIR_ETRANSCRIPT: too small
        check(fn(*args, buf, ln.value, ctypes.byref(ln), None, None), "copy")
        return bytes(buf), int(cp.value), root

    pr = whir_amd.Prover(whir_amd.params(*cfg), gpu_ctx)
    b = whir_amd.Batch(whir_amd.params(*cfg), 2, ctx=gpu_ctx)
    m = whir_amd.MultiBatch([whir_amd.params(*cfg), whir_amd.params(13, 4, 1, 3, 1, 123, 1)], streams=2, ctx=gpu_ctx)
    bufs = []
    try:
        pr.upload(polys[0])
        pr.enqueue()
        v = pr.result()
        c = copied(L.whir_prover_result, pr.ptr)
        assert c[0] == v[0] and c[1] == v[1] and np.array_equal(c[2], v[2])
        for j, p in enumerate(polys):
            b.upload(j, p)
        b.enqueue()
        for j in range(2):
            v = b.result(j)
            c = copied(L.whir_batch_result, b.ptr, j)
            assert c[0] == v[0] and c[1] == v[1] and np.array_equal(c[2], v[2])
        for p in (polys[0], oracle.rand_poly(70, 13)):
            d = whir_amd.api.DeviceBuffer(p.nbytes)
            d.upload(p)
            bufs.append(d)
        m.enqueue([d.ptr for d in bufs])
        for j in range(2):
            v = m.result(j)
            c = copied(L.whir_multi_result, m.ptr, j)
            assert c[0] == v[0] and c[1] == v[1] and np.array_equal(c[2], v[2])
        assert m.result(0)[0] == oracle.whir_prove(*cfg, polys[0])[0]
    finally:
        pr.close()
        b.close()
        m.close()


@pytest.mark.parametrize("graph,groups", [(0, 1), (1, 2), (2, 3), (1, 8)])
def test_multi_commit_queue_matches_oracle(gpu_ctx, oracle, graph, groups):
    """The commit work queue (persistent NTT / leaf / layer kernels draining a
    device job table of every folding-4 job, uneven sizes and rates; the
    batches in `groups` pipelined groups) ahead of the batches: every job's
    bytes equal the oracle's, replayed twice; the job of folding 2 keeps its
    own commit; switching the queue off and on again gives the same bytes."""
    import whir_amd

    cfgs = [(12, 4, 1, 3, 1, 123, 1), (16, 4, 2, 3, 1, 123, 1), (13, 4, 1, 3, 1, 123, 2), (12, 4, 1, 3, 1, 123, 1),
            (17, 4, 1, 3, 1, 123, 1), (10, 2, 1, 1, 1, 32, 3), (15, 4, 2, 3, 1, 123, 1), (14, 4, 1, 3, 1, 123, 1)]
    polys = [oracle.rand_poly(1300 + i, c[0]) for i, c in enumerate(cfgs)]
    want = [oracle.whir_prove(*c, p)[0] for c, p in zip(cfgs, polys)]
    bufs = []
    for p in polys:
        d = whir_amd.api.DeviceBuffer(p.nbytes)
        d.upload(p)
        bufs.append(d)
    print(f"commit queue: {len(cfgs)} jobs, graph={graph}, groups={groups}", flush=True)
    m = whir_amd.MultiBatch([whir_amd.params(*c) for c in cfgs], streams=3, ctx=gpu_ctx, graph=graph,
                            commit_queue=groups)
    try:
        assert m.queue_jobs == len(cfgs) - 1  # every job but the folding-2 one
        for rep in range(2):
            m.enqueue([d.ptr for d in bufs])
            got = [m.result(i)[0] for i in range(len(cfgs))]
            for i in range(len(cfgs)):
                assert got[i] == want[i], (rep, i, cfgs[i])
        assert m.set_commit_queue(0) == 0
        m.enqueue([d.ptr for d in bufs])
        assert [m.result(i)[0] for i in range(len(cfgs))] == want
        assert m.set_commit_queue(groups) == len(cfgs) - 1
        m.enqueue([d.ptr for d in bufs])
        assert [m.result(i)[0] for i in range(len(cfgs))] == want
    finally:
        m.close()
