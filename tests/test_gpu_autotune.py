"""The autotuner (whir_ctx_set_autotune): every candidate it may choose gives the
same words, and proofs on an autotuned context are the oracle's bytes.

* NTT candidates: two passes (three or two register stages each) or three
  passes x first-pass radix x workgroup size x
  vector width (the space ntt_candidates enumerates), forced through the
  context knobs, each DFT equal to the oracle's (Radix2DFTSmallBatch
  semantics, crates/whir/src/whir.rs:34,59);
* Merkle candidates: two-lane threshold x cooperative-subtree depth, each tree
  equal to the oracle's layers (crates/merkle/src/poseidon_packed.rs);
* whole proofs on an autotuned context (its choices reported by
  whir_ctx_tuning) equal to the oracle's and accepted by its verifier."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 0x7F000001


def _rand(rng, shape):
    return rng.integers(0, P, size=shape, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("n_in,log_h,width", [(1 << 18, 19, 16), (1 << 14, 16, 64), ((1 << 17) + 3, 19, 4)])
def test_ntt_candidates_identical(oracle, n_in, log_h, width):
    import whir_amd
    from whir_amd._lib import lib

    rng = np.random.default_rng(n_in ^ log_h)
    mat = _rand(rng, (n_in, width))
    padded = np.zeros((1 << log_h, width), np.uint32)
    padded[:n_in] = mat
    want = oracle.dft_batch(padded)
    ctx = whir_amd.Context(0, 1000)
    try:
        for two in (0, 1, 2):
            for first in (4, 3, 0):
                for lt in (8, 7):
                    for lv in (2, 1):
                        assert lib().whir_ctx_set_ntt_passes(ctx.ptr, two) == 0
                        assert lib().whir_ctx_set_ntt_shape(ctx.ptr, first, lv) == 0
                        assert lib().whir_ctx_set_ntt_tuning(ctx.ptr, 8, lt) == 0
                        print(f"ntt candidate two_pass={two} first={first} log_threads={lt} log_vec={lv} "
                              f"n_in={n_in} log_h={log_h} width={width}", flush=True)
                        got = whir_amd.dft_batch(mat, log_h, ctx=ctx)
                        assert np.array_equal(got, want), (two, first, lt, lv)
    finally:
        ctx.close()


@pytest.mark.parametrize("log_h,row_words", [(17, 16), (14, 64), (12, 32)])
def test_merkle_candidates_identical(oracle, log_h, row_words):
    import whir_amd
    from whir_amd._lib import lib

    rng = np.random.default_rng(log_h * 100 + row_words)
    rows = _rand(rng, (1 << log_h, row_words))
    want_root, want_layers = oracle.merkle_commit(rows, want_layers=True)
    ctx = whir_amd.Context(0, 1000)
    try:
        for pm in (13, 14, 15, 16):
            for cf in (13, 14, 15, 16, 17):
                assert lib().whir_ctx_set_merkle_tuning(ctx.ptr, pm, cf) == 0
                print(f"merkle candidate pair_max_log={pm} coop_from_log={cf} log_h={log_h} words={row_words}",
                      flush=True)
                root, layers = whir_amd.merkle_commit(rows, ctx=ctx)
                assert np.array_equal(layers, want_layers) and np.array_equal(root, want_root), (pm, cf)
    finally:
        ctx.close()


def test_autotuned_proofs_match_oracle(oracle):
    import whir_amd

    ctx = whir_amd.Context(0, 1000, autotune=True)
    try:
        for cfg, seed in (((18, 4, 1, 3, 1, 123, 1), 1), ((20, 4, 2, 3, 1, 123, 2), 2), ((19, 3, 1, 2, 1, 32, 1), 3)):
            k, f, rate, ir, s, sec, n = cfg
            poly = oracle.rand_poly(seed, k)
            print(f"autotuned prove k={k} f={f} rate={rate}", flush=True)
            proof, cp, root = whir_amd.Whir(k, f, rate, ir, s, sec, ctx=ctx).prove(poly, n)
            want = oracle.whir_prove(*cfg, poly)
            assert proof == want[0] and cp == want[1]
            rc, _ = oracle.whir_verify(*cfg, proof)
            assert rc == 0
        tuned = ctx.tuning()
        kinds = {t["kind"] for t in tuned}
        assert kinds == {"ntt", "merkle"}
        for t in tuned:
            assert t["candidates"] >= 2 and t["best_us"] > 0
            assert t["default_us"] <= 0 or t["best_us"] <= t["default_us"] + 1e-6
    finally:
        ctx.close()


def test_batches_tuned_with_their_job_count(oracle):
    """A lockstep batch on an autotuned context is timed with its own job
    count (its plans kept for that job-count class, beside the single-proof
    ones; wider Merkle levels are candidates) and every job's proof is still
    the oracle's."""
    import whir_amd

    cfg = (18, 4, 1, 3, 1, 123, 1)
    ctx = whir_amd.Context(0, 1000, autotune=True)
    try:
        polys = [oracle.rand_poly(80 + j, 18) for j in range(4)]
        print("autotuned batch of 4, k=18", flush=True)
        b = whir_amd.Batch(whir_amd.params(*cfg), 4, ctx=ctx)
        try:
            for j, p in enumerate(polys):
                b.upload(j, p)
            b.enqueue()
            got = [b.result(j)[0] for j in range(4)]
        finally:
            b.close()
        tuned = ctx.tuning()
        assert {t["jobs"] for t in tuned} == {4}
        assert {t["kind"] for t in tuned} == {"ntt", "merkle"}
        # the single-proof path on the same context tunes its own plans
        single = whir_amd.Whir(*cfg[:6], ctx=ctx).prove(polys[0], 1)[0]
        assert {t["jobs"] for t in ctx.tuning()} == {1, 4}
    finally:
        ctx.close()
    assert got == [oracle.whir_prove(*cfg, p)[0] for p in polys]
    assert single == got[0]
