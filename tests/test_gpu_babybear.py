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
