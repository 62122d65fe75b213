"""GPU parity of the BabyBear build (libwhir_amd_bb.so: the prover's kernels
compiled over p = 15 * 2^27 + 1, Ext4 with W = 11, Poseidon2 with the x^7
S-box and R_P = 13 / 21; the field of crates/transcript/src/lib.rs:389-396)
against the oracle's BabyBear build, bit-exact: Poseidon2 lanes and the
transcript's cooperative W16, the column DFT, the Merkle tree, the eq / eval /
sumcheck kernels, and whole proofs (run_whir_poseidon's flow,
crates/whir/src/test.rs:174-254) accepted by the oracle's verifier, alone and
in a lockstep batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 15 * (1 << 27) + 1


@pytest.fixture(scope="module")
def bb(oracle):
    o = oracle.for_field("babybear")
    o.lib()
    return o


@pytest.fixture(scope="module")
def bb_ctx():
    import whir_amd

    ctx = whir_amd.Context(0, 1000, field="babybear")
    yield ctx
    ctx.close()


def _rand(rng, shape):
    return rng.integers(0, P, size=shape, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("width,coop", [(16, False), (24, False), (16, True)])
def test_poseidon2(bb_ctx, bb, width, coop):
    import whir_amd

    rng = np.random.default_rng(width + coop)
    states = _rand(rng, (640, width))
    states[0] = 0
    states[1] = P - 1
    print(f"babybear poseidon2 width={width} cooperative={coop}", flush=True)
    got = whir_amd.poseidon2_permute(states, width, cooperative=coop, ctx=bb_ctx)
    want = np.stack([bb.permute(width, s) for s in states])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n_in,log_h,width", [(8, 4, 2), (1 << 12, 13, 16), (1 << 16, 17, 4), (1 << 20, 21, 16)])
def test_dft(bb_ctx, bb, n_in, log_h, width):
    import whir_amd

    rng = np.random.default_rng(log_h)
    mat = _rand(rng, (n_in, width))
    print(f"babybear dft n_in={n_in} log_h={log_h} width={width}", flush=True)
    got = whir_amd.dft_batch(mat, log_h, ctx=bb_ctx)
    pad = np.zeros((1 << log_h, width), np.uint32)
    pad[:n_in] = mat
    assert np.array_equal(got, bb.dft_batch(pad))


@pytest.mark.parametrize("log_h,row_words", [(3, 4), (10, 16), (14, 64), (17, 16)])
def test_merkle(bb_ctx, bb, log_h, row_words):
    import whir_amd

    rng = np.random.default_rng(log_h * 7 + row_words)
    rows = _rand(rng, (1 << log_h, row_words))
    print(f"babybear merkle log_h={log_h} row_words={row_words}", flush=True)
    root, layers = whir_amd.merkle_commit(rows, ctx=bb_ctx)
    want_root, want_layers = bb.merkle_commit(rows, want_layers=True)
    assert np.array_equal(root, want_root)
    assert np.array_equal(layers, want_layers)


def test_eq_eval_and_sumcheck(bb_ctx, bb):
    import whir_amd

    api = whir_amd.api
    k = 12
    poly = bb.rand_poly(3, k)
    pe = bb.rand_ext(7, 1 << k)
    z = bb.rand_ext(4, k)
    print("babybear eval / compress / sumcheck coeffs", flush=True)
    assert np.array_equal(api.eval_ext(pe, z, ctx=bb_ctx), bb.eval_ext(pe, z))
    assert np.array_equal(api.compress_hi(poly, k, z[:8], ctx=bb_ctx), bb.compress_hi(poly, z[:8]))
    assert np.array_equal(api.compress_lo(poly, k, z[:3], ctx=bb_ctx), bb.compress_lo(poly, z[:3]))
    f = bb.rand_ext(5, 1 << 10)
    w = bb.rand_ext(6, 1 << 10)
    assert np.array_equal(api.sumcheck_coeffs(f, w, ctx=bb_ctx), bb.coeffs(f, w))


PROOFS = [
    (10, 2, 1, 1, 1, 32, 2),
    (12, 3, 1, 2, 2, 32, 1),
    (11, 3, 1, 3, 0, 32, 1),    # unique decoding: no OOD claims
    (14, 4, 2, 3, 1, 123, 3),
    (18, 4, 1, 3, 1, 123, 1),   # the reference's CPU-runnable shape
]


@pytest.mark.parametrize("cfg", PROOFS, ids=lambda c: "k{}f{}r{}ir{}s{}sec{}n{}".format(*c))
def test_proof_bytes_match_oracle(bb_ctx, bb, cfg):
    import whir_amd

    k, f, rate, ir, s, sec, n = cfg
    poly = bb.rand_poly(1, k)
    print(f"babybear prove {cfg}", flush=True)
    gproof, gcp, groot = whir_amd.Whir(k, f, rate, ir, s, sec, ctx=bb_ctx).prove(poly, n)
    oproof, ocp, oroot = bb.whir_prove(*cfg, poly)
    assert np.array_equal(groot, oroot)
    assert gproof == oproof and gcp == ocp
    rc, vcp = bb.whir_verify(*cfg, gproof)
    assert rc == 0 and vcp == gcp


def test_batch_matches_oracle(bb_ctx, bb):
    import whir_amd

    cfg = (13, 4, 1, 3, 1, 123, 2)
    polys = [bb.rand_poly(50 + j, 13) for j in range(3)]
    print("babybear lockstep batch of 3", flush=True)
    b = whir_amd.Batch(whir_amd.params(*cfg), 3, ctx=bb_ctx)
    try:
        for j, p in enumerate(polys):
            b.upload(j, p)
        b.enqueue()
        got = [b.result(j)[0] for j in range(3)]
    finally:
        b.close()
    assert got == [bb.whir_prove(*cfg, p)[0] for p in polys]


def test_multi_commit_queue_matches_oracle(bb_ctx, bb):
    """A mixed-shape set with the commit work queue over the BabyBear build."""
    import whir_amd

    cfgs = [(13, 4, 1, 3, 1, 123, 1), (15, 4, 2, 3, 1, 123, 1), (13, 4, 1, 3, 1, 123, 1), (14, 4, 1, 3, 1, 123, 2)]
    polys = [bb.rand_poly(80 + i, c[0]) for i, c in enumerate(cfgs)]
    bufs = []
    for p in polys:
        d = whir_amd.api.DeviceBuffer(p.nbytes)
        d.upload(p)
        bufs.append(d)
    print("babybear mixed-shape set with the commit queue", flush=True)
    m = whir_amd.MultiBatch([whir_amd.params(*c) for c in cfgs], streams=2, ctx=bb_ctx, graph=1, commit_queue=1)
    try:
        assert m.queue_jobs == len(cfgs)
        m.enqueue([d.ptr for d in bufs])
        got = [m.result(i)[0] for i in range(len(cfgs))]
    finally:
        m.close()
    assert got == [bb.whir_prove(*c, p)[0] for c, p in zip(cfgs, polys)]


def test_session_with_prefix_and_given_point(bb_ctx, bb):
    """Whir::commit / ProverStack::eval / Whir::open inside a caller's
    BabyBear transcript (crates/whir/src/whir.rs:127-180): the GPU's bytes
    follow the caller's prefix as in the oracle's session proof, the
    handed-back challenger is the oracle's, and its verifier accepts."""
    import whir_amd
    from kbpy import DuplexPy

    k, f, rate, ir, s, sec = 12, 3, 1, 2, 1, 32
    rng = np.random.default_rng(3)
    poly = bb.rand_poly(9, k)
    prefix = _rand(rng, 5)
    given = [_rand(rng, (k, 4))]
    claims = [(0, 1), (0, 2)]
    oproof, ocp, oroot, ost = bb.whir_prove_session(k, f, rate, ir, s, sec, [poly], claims, prefix, given)
    d = DuplexPy(bb)
    for x in prefix:
        d.observe(int(x))
    print("babybear session k=12", flush=True)
    pr = whir_amd.Prover(whir_amd.params(k, f, rate, ir, s, sec, 2), bb_ctx)
    try:
        pr.upload(poly)
        pr.commit(None, d.as_state())
        pr.eval(None)
        pr.eval(given[0])
        pr.open()
        gproof, gcp, groot = pr.result()
        gst = pr.challenger()
    finally:
        pr.close()
    assert gproof == oproof[4 * len(prefix):] and gcp == ocp and np.array_equal(groot, oroot)
    assert np.array_equal(np.asarray(gst["sponge_state"]), np.asarray(ost["sponge_state"]))
    rc, _ = bb.whir_verify_session(k, f, rate, ir, s, sec, [k], claims, oproof[:4 * len(prefix)] + gproof, prefix,
                                   given)
    assert rc == 0


def test_stacked_proof(bb_ctx, bb):
    """Several BabyBear polys committed and opened as one (ProverStack,
    crates/whir/src/sumcheck/prover.rs:31-124) against the oracle's stack."""
    import whir_amd
    from tests.stack_specs import MIXED, rand_polys, stacked_k

    k_polys, claims = MIXED[1]
    f, rate, ir, s, sec = 3, 1, 2, 1, 32
    k = stacked_k(k_polys)
    polys = rand_polys(bb, k_polys, 31)
    print(f"babybear stacked k={k} polys={k_polys}", flush=True)
    pr = whir_amd.StackedProver(whir_amd.params(k, f, rate, ir, s, sec, len(claims)), k_polys, claims, ctx=bb_ctx)
    try:
        pr.upload(polys)
        pr.enqueue()
        got = pr.result()
    finally:
        pr.close()
    want = bb.whir_prove_stacked(k, f, rate, ir, s, sec, polys, claims)
    assert got[0] == want[0] and got[1] == want[1]


def test_checked_proof(bb_ctx, bb):
    """The device self-checks on a BabyBear proof with a grid round after
    round 0: no check fails and the bytes are the oracle's."""
    import whir_amd

    cfg = (20, 3, 1, 3, 1, 123, 1)
    poly = bb.rand_poly(2, 20)
    print("babybear checked k=20", flush=True)
    pr = whir_amd.Prover(whir_amd.params(*cfg), bb_ctx)
    try:
        pr.set_checks(True)
        pr.upload(poly)
        pr.enqueue()
        got = pr.result()
    finally:
        pr.close()
    assert got[0] == bb.whir_prove(*cfg, poly)[0]
