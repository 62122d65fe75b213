"""Pins the oracle's field layer by closed form: KoalaBear Montgomery arithmetic
against Python big integers, the two-adic generators against 3^((p-1)/2^b),
irreducibility of x^4 - 3, and extension multiplication against polynomial
multiplication mod x^4 - 3."""
import random

from kbpy import P, R, emul, from_m, to_m


def test_modulus_and_montgomery_constants(oracle):
    assert P == 2**31 - 2**24 + 1
    assert oracle.to_monty(1) == R
    for c in (0, 1, 2, P - 1, 12345):
        assert oracle.from_monty(oracle.to_monty(c)) == c


def test_base_arithmetic_matches_bigints(oracle):
    rnd = random.Random(1)
    for _ in range(2000):
        a, b = rnd.randrange(P), rnd.randrange(P)
        ma, mb = to_m(a), to_m(b)
        assert from_m(oracle.fmul(ma, mb)) == a * b % P
        assert from_m(oracle.fadd(ma, mb)) == (a + b) % P
        assert from_m(oracle.lib().oracle_fsub(ma, mb)) == (a - b) % P
        if a:
            assert from_m(oracle.finv(ma)) == pow(a, -1, P)


def test_two_adic_generators_closed_form(oracle):
    for b in range(25):
        g = from_m(oracle.two_adic_generator(b))
        assert g == pow(3, (P - 1) >> b, P)
        assert pow(g, 1 << b, P) == 1
        if b:
            assert pow(g, 1 << (b - 1), P) == P - 1  # exact order 2^b
    # the last two entries of Plonky3's KoalaBear TWO_ADIC_GENERATORS table
    assert pow(3, (P - 1) >> 23, P) == 0x68E7DD49
    assert pow(3, (P - 1) >> 24, P) == 0x6AC49F88


def test_extension_is_a_field():
    # x^4 - W is irreducible over F_p (p = 1 mod 4) iff W is a non-square.
    assert pow(3, (P - 1) // 2, P) == P - 1
    assert pow(2, (P - 1) // 2, P) == 1  # 2 is a square: 3 is the smallest non-residue
    assert P % 4 == 1


def test_ext_mul_matches_polynomial_product(oracle):
    import numpy as np

    rnd = random.Random(2)
    for _ in range(500):
        a = [rnd.randrange(P) for _ in range(4)]
        b = [rnd.randrange(P) for _ in range(4)]
        out = np.zeros(4, np.uint32)
        am = np.array([to_m(x) for x in a], np.uint32)
        bm = np.array([to_m(x) for x in b], np.uint32)
        u32p = oracle.u32p
        oracle.lib().oracle_emul(am.ctypes.data_as(u32p), bm.ctypes.data_as(u32p), out.ctypes.data_as(u32p))
        assert [from_m(int(x)) for x in out] == emul(a, b)
