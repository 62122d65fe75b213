"""Pins the oracle's DFT, eq-polynomial and sumcheck kernels by mathematics and
by the reference's own cross-implementation tests, restated:

* dft_batch == the O(n^2) DFT definition with omega = two_adic_generator(log h),
  and the layout property of test_simple_whir_no_transpose
  (crates/whir/src/test.rs:106-172): codeword row i, evaluated at the folding
  point, equals the folded polynomial at omega^i;
* eq / eval / compress_lo / compress_hi identities of test_poly and test_eq
  (crates/poly/src/test.rs:24-124) against a naive Python evaluation;
* coeffs + fix_lo_var + extrapolate_012 satisfy the sumcheck round identity
  (the reference's debug invariant expr.prod() == sum, expr.rs:68);
* combine_eqs / combine_pows == the naive weighted sums (combine.rs tests)."""
import random

import numpy as np
import pytest

from kbpy import ONE, P, ZERO, e_from_m, e_to_m, eadd, emul, eq_naive, escale, esub, from_m, to_m


def rand_ext_list(rnd, n):
    return [[rnd.randrange(P) for _ in range(4)] for _ in range(n)]


def ext_arr(lst):
    return np.array([e_to_m(e) for e in lst], np.uint32).reshape(-1, 4)


def arr_ext(a):
    return [e_from_m(r) for r in np.asarray(a).reshape(-1, 4)]


@pytest.mark.parametrize("h,w", [(2, 1), (4, 3), (8, 2), (16, 4), (32, 1), (64, 2)])
def test_dft_matches_definition(oracle, h, w):
    rnd = np.random.default_rng(h * 10 + w)
    mat = rnd.integers(0, P, (h, w), dtype=np.uint64).astype(np.uint32)
    got = oracle.dft_batch(mat)
    g = from_m(oracle.two_adic_generator(h.bit_length() - 1))
    for c in range(w):
        col = [from_m(int(x)) for x in mat[:, c]]
        for i in range(h):
            want = sum(col[j] * pow(g, i * j, P) for j in range(h)) % P
            assert from_m(int(got[i, c])) == want
    assert np.array_equal(got, oracle.dft_batch_naive(mat))


def test_codeword_layout_property(oracle):
    """test_simple_whir_no_transpose restated on KoalaBear with extension points:
    encode(f, rate, folding) = DFT of f as 2^folding-wide rows padded to
    2^(k - folding + rate) rows; row i evaluated at the folding point alpha equals
    f1(omega^i) where f1 = f with its low `folding` variables bound to alpha."""
    rnd = random.Random(3)
    k = 6
    f0 = oracle.rand_poly(1, k)
    for folding in range(0, 4):
        for rate in range(0, 3):
            alpha = rand_ext_list(rnd, folding)
            kd = k - folding + rate
            mat = np.zeros((1 << kd, 1 << folding), np.uint32)
            mat[: 1 << (k - folding)] = f0.reshape(-1, 1 << folding)
            cw = oracle.dft_batch(mat)
            f1 = oracle.compress_lo(f0, ext_arr(alpha)) if folding else ext_arr(
                [[from_m(int(x)), 0, 0, 0] for x in f0])
            om = oracle.two_adic_generator(kd)
            for i in range(0, 1 << kd, max(1, (1 << kd) // 16)):
                w = from_m(om)
                x = np.array(e_to_m([pow(w, i, P), 0, 0, 0]), np.uint32)
                y0 = oracle.eval_univariate(f1, x)
                row = cw[i]
                y1 = oracle.eval_base(row, ext_arr(alpha)) if folding else np.array(
                    e_to_m([from_m(int(row[0])), 0, 0, 0]), np.uint32)
                assert np.array_equal(y0, y1)


def naive_eval(poly_ext, z):
    acc = ZERO
    for i, c in enumerate(poly_ext):
        acc = eadd(acc, emul(c, eq_naive(z, i)))
    return acc


@pytest.mark.parametrize("k", [1, 2, 3, 5, 6])
def test_eval_and_compress_identities(oracle, k):
    rnd = random.Random(k)
    poly = oracle.rand_poly(k + 10, k)
    pe = [[from_m(int(x)), 0, 0, 0] for x in poly]
    z = rand_ext_list(rnd, k)
    e0 = naive_eval(pe, z)
    assert arr_ext(oracle.eval_base(poly, ext_arr(z)))[0] == e0
    assert arr_ext(oracle.compress_lo(poly, ext_arr(z)))[0] == e0
    assert arr_ext(oracle.compress_hi(poly, ext_arr(z)))[0] == e0
    for d in range(1, k):
        z0, z1 = z[:d], z[d:]
        lo = oracle.compress_lo(poly, ext_arr(z0))
        assert arr_ext(oracle.eval_ext(lo, ext_arr(z1)))[0] == e0
        hi = oracle.compress_hi(poly, ext_arr(z1))
        assert arr_ext(oracle.eval_ext(hi, ext_arr(z0)))[0] == e0
    # eq table entries
    table = arr_ext(oracle.eq_table(ext_arr(z)))
    assert table == [eq_naive(z, i) for i in range(1 << k)]


@pytest.mark.parametrize("k", [1, 3, 6])
def test_sumcheck_round_identity(oracle, k):
    rnd = random.Random(20 + k)
    f = rand_ext_list(rnd, 1 << k)
    w = rand_ext_list(rnd, 1 << k)
    total = ZERO
    for a, b in zip(f, w):
        total = eadd(total, emul(a, b))
    c0c2 = arr_ext(oracle.coeffs(ext_arr(f), ext_arr(w)))
    c0, c2 = c0c2
    # h(X) = sum_i (f[2i] + X (f[2i+1]-f[2i])) (w[2i] + X (w[2i+1]-w[2i])): h(0) = c0, h(2) = c2
    h0, h2 = ZERO, ZERO
    for i in range(len(f) // 2):
        h0 = eadd(h0, emul(f[2 * i], w[2 * i]))
        F2 = esub(escale(f[2 * i + 1], 2), f[2 * i])
        W2 = esub(escale(w[2 * i + 1], 2), w[2 * i])
        h2 = eadd(h2, emul(F2, W2))
    assert (c0, c2) == (h0, h2)
    r = rand_ext_list(rnd, 1)[0]
    ff = arr_ext(oracle.fix_lo_var(ext_arr(f), ext_arr([r])[0]))
    ww = arr_ext(oracle.fix_lo_var(ext_arr(w), ext_arr([r])[0]))
    folded = ZERO
    for a, b in zip(ff, ww):
        folded = eadd(folded, emul(a, b))
    # extrapolate_012(c0, sum - c0, c2, r) == <fold f, fold w>
    inv2 = pow(2, -1, P)
    one, two = ONE, [2, 0, 0, 0]
    l0 = escale(emul(esub(r, one), esub(r, two)), inv2)
    l1 = emul(r, esub(two, r))
    l2 = escale(emul(r, esub(r, one)), inv2)
    h1 = esub(total, c0)
    assert eadd(eadd(emul(c0, l0), emul(h1, l1)), emul(c2, l2)) == folded


def test_combine_eqs_and_pows(oracle):
    rnd = random.Random(9)
    for k in (1, 4, 7):
        for n in (1, 2, 5):
            alpha = rand_ext_list(rnd, 1)[0]
            pts = [rand_ext_list(rnd, k) for _ in range(n)]
            base = rand_ext_list(rnd, 1 << k)
            got = arr_ext(oracle.combine_eqs(ext_arr(base), np.stack([ext_arr(p) for p in pts]), ext_arr([alpha])[0]))
            want = []
            for x in range(1 << k):
                acc = base[x]
                a = ONE
                for p in pts:
                    acc = eadd(acc, emul(a, eq_naive(p, x)))
                    a = emul(a, alpha)
                want.append(acc)
            assert got == want
            vars_ = [rnd.randrange(P) for _ in range(n)]
            shift = 3
            got = arr_ext(oracle.combine_pows(ext_arr(base), np.array([to_m(v) for v in vars_], np.uint32),
                                              ext_arr([alpha])[0], shift))
            want = []
            for x in range(1 << k):
                acc = base[x]
                for j, v in enumerate(vars_):
                    a = ONE
                    for _ in range(shift + j):
                        a = emul(a, alpha)
                    acc = eadd(acc, escale(a, pow(v, x, P)))
                want.append(acc)
            assert got == want
            # combine_pows_scaled (combine.rs:440-468) with arbitrary scales
            scales = rand_ext_list(rnd, n)
            got = arr_ext(oracle.combine_pows_scaled(ext_arr(base), np.array([to_m(v) for v in vars_], np.uint32),
                                                     ext_arr(scales)))
            want = []
            for x in range(1 << k):
                acc = base[x]
                for v, sc in zip(vars_, scales):
                    acc = eadd(acc, escale(sc, pow(v, x, P)))
                want.append(acc)
            assert got == want
