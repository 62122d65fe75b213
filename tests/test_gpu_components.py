"""Component C-ABI entry points (eq / SplitEq / sumcheck / combine / Merkle
openings) against the oracle, bit for bit, on seeded inputs; Merkle openings
are checked by recomputing the oracle root from each row and its path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 0x7F000001


def _ext(rng, n):
    return rng.integers(0, P, (n, 4), dtype=np.uint64).astype(np.uint32)


def _base(rng, n):
    return rng.integers(0, P, n, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("k", [1, 2, 5, 9, 14, 20])
def test_eq_table(gpu_ctx, oracle, k):
    import whir_amd

    rng = np.random.default_rng(k)
    z, scale = _ext(rng, k), _ext(rng, 1)[0]
    print(f"eq_table k={k}", flush=True)
    assert np.array_equal(whir_amd.api.eq_table(z, scale, ctx=gpu_ctx), oracle.eq_table(z, scale))


@pytest.mark.parametrize("k", [1, 3, 8, 13, 18])
def test_eval_ext(gpu_ctx, oracle, k):
    import whir_amd

    rng = np.random.default_rng(100 + k)
    poly, z = _ext(rng, 1 << k), _ext(rng, k)
    print(f"eval_ext k={k}", flush=True)
    assert np.array_equal(whir_amd.api.eval_ext(poly, z, ctx=gpu_ctx), oracle.eval_ext(poly, z))


@pytest.mark.parametrize("k,m", [(4, 4), (8, 4), (10, 6), (12, 12), (16, 12), (20, 15), (9, 8)])
def test_compress_hi(gpu_ctx, oracle, k, m):
    import whir_amd

    rng = np.random.default_rng(200 + k * 7 + m)
    poly, z = _base(rng, 1 << k), _ext(rng, m)
    print(f"compress_hi k={k} m={m}", flush=True)
    got = whir_amd.api.compress_hi(poly, k, z, ctx=gpu_ctx)
    assert np.array_equal(got, oracle.compress_hi(poly, z))
    if m == k:  # eval_base
        assert np.array_equal(got[0], oracle.eval_base(poly, z))


@pytest.mark.parametrize("k,m", [(4, 4), (8, 1), (10, 4), (16, 5), (20, 4)])
def test_compress_lo(gpu_ctx, oracle, k, m):
    import whir_amd

    rng = np.random.default_rng(300 + k * 7 + m)
    poly, z = _base(rng, 1 << k), _ext(rng, m)
    print(f"compress_lo k={k} m={m}", flush=True)
    assert np.array_equal(whir_amd.api.compress_lo(poly, k, z, ctx=gpu_ctx), oracle.compress_lo(poly, z))


@pytest.mark.parametrize("k", [1, 2, 7, 12, 19])
def test_sumcheck_coeffs_and_fold(gpu_ctx, oracle, k):
    import whir_amd

    rng = np.random.default_rng(400 + k)
    f, w, r = _ext(rng, 1 << k), _ext(rng, 1 << k), _ext(rng, 1)[0]
    print(f"coeffs/fix_lo_var k={k}", flush=True)
    assert np.array_equal(whir_amd.api.sumcheck_coeffs(f, w, ctx=gpu_ctx), oracle.coeffs(f, w))
    assert np.array_equal(whir_amd.api.fix_lo_var(f, r, ctx=gpu_ctx), oracle.fix_lo_var(f, r))


@pytest.mark.parametrize("k,n", [(1, 1), (6, 3), (11, 2), (16, 5), (13, 9), (12, 1)])
def test_combine_eqs(gpu_ctx, oracle, k, n):
    import whir_amd

    rng = np.random.default_rng(500 + k * 11 + n)
    w, pts, alpha = _ext(rng, 1 << k), _ext(rng, n * k).reshape(n, k, 4), _ext(rng, 1)[0]
    print(f"combine_eqs k={k} n={n}", flush=True)
    assert np.array_equal(whir_amd.api.combine_eqs(w, pts, alpha, ctx=gpu_ctx), oracle.combine_eqs(w, pts, alpha))


@pytest.mark.parametrize("k,n,shift", [(1, 1, 0), (5, 7, 2), (12, 40, 3), (16, 133, 2), (9, 1, 0)])
def test_combine_pows(gpu_ctx, oracle, k, n, shift):
    import whir_amd

    rng = np.random.default_rng(600 + k * 13 + n)
    w, vs, alpha = _ext(rng, 1 << k), _base(rng, n), _ext(rng, 1)[0]
    print(f"combine_pows k={k} n={n} shift={shift}", flush=True)
    got = whir_amd.api.combine_pows(w, vs, alpha, shift, ctx=gpu_ctx)
    assert np.array_equal(got, oracle.combine_pows(w, vs, alpha, shift))


@pytest.mark.parametrize("k,n", [(1, 1), (6, 5), (13, 40), (16, 133)])
def test_combine_pows_scaled(gpu_ctx, oracle, k, n):
    import whir_amd

    rng = np.random.default_rng(650 + k * 13 + n)
    w, vs, scales = _ext(rng, 1 << k), _base(rng, n), _ext(rng, n)
    print(f"combine_pows_scaled k={k} n={n}", flush=True)
    got = whir_amd.api.combine_pows_scaled(w, vs, scales, ctx=gpu_ctx)
    assert np.array_equal(got, oracle.combine_pows_scaled(w, vs, scales))


@pytest.mark.parametrize("log_h,row_words", [(1, 16), (5, 64), (10, 16), (13, 8)])
def test_merkle_open_paths_verify(gpu_ctx, oracle, log_h, row_words):
    import whir_amd

    rng = np.random.default_rng(700 + log_h)
    rows = _base(rng, (1 << log_h) * row_words).reshape(1 << log_h, row_words)
    root, layers = oracle.merkle_commit(rows, want_layers=True)
    idx = rng.integers(0, 1 << log_h, 9).astype(np.uint32)
    print(f"merkle_open log_h={log_h} row_words={row_words}", flush=True)
    got = whir_amd.api.merkle_open(layers, rows, log_h, idx, ctx=gpu_ctx)
    for q, i in enumerate(idx):
        row = got[q, :row_words]
        assert np.array_equal(row, rows[i])
        node = oracle.merkle_commit(row.reshape(1, row_words))  # leaf digest (a 1-row tree's root)
        cur = int(i)
        for lvl in range(log_h):
            sib = got[q, row_words + 8 * lvl: row_words + 8 * (lvl + 1)]
            mid = 1 << (log_h - lvl - 1)
            pos = cur & ((mid << 1) - 1)
            node = oracle.compress16(node, sib) if pos < mid else oracle.compress16(sib, node)
        assert np.array_equal(node, root)


SORT_NS = [0, 1, 2, 3, 63, 64, 65, 127, 128, 129, 255, 256, 257, 287, 511, 512, 513, 1000, 1024]


@pytest.mark.parametrize("n", SORT_NS)
def test_sort_dedup(gpu_ctx, n):
    """sorted().dedup() of stir_indexes (whir.rs:63-77) is np.unique of the
    values: random indices with duplicates, all equal, ascending, descending
    and values next to the padding word."""
    import whir_amd

    rng = np.random.default_rng(7000 + n)
    cases = [
        rng.integers(0, max(1, n // 2), n, dtype=np.uint64).astype(np.uint32),
        rng.integers(0, 1 << 31, n, dtype=np.uint64).astype(np.uint32),
        np.full(n, 12345, np.uint32),
        np.arange(n, dtype=np.uint32),
        np.arange(n, dtype=np.uint32)[::-1].copy(),
        (0xFFFFFFFE - rng.integers(0, 3, n, dtype=np.uint64)).astype(np.uint32),
    ]
    print(f"sort_dedup n={n}", flush=True)
    for v in cases:
        assert np.array_equal(whir_amd.api.sort_dedup(v, ctx=gpu_ctx), np.unique(v))
