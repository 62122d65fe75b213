"""ORACLE — test infrastructure only.

ctypes binding of oracle/liboracle.so (the CPU restatement of the reference
WHIR prover/verifier). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, as the checker; the product path never
does. Field elements are numpy uint32 arrays of Montgomery words; extension
elements are 4 consecutive words (coefficient 0 first).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

FIELD = "koalabear"
P = 0x7F000001
_FIELDS = {"koalabear": ("liboracle.so", 0x7F000001), "babybear": ("liboracle_bb.so", 0x78000001)}
_by_field = {}


def for_field(name):
    """This module's functions over the oracle built for `name` ("koalabear",
    the default, or "babybear": the same restatement over BabyBear)."""
    import importlib.util
    import sys

    if name == FIELD:
        return sys.modules[__name__]
    if name not in _by_field:
        libname, p = _FIELDS[name]
        spec = importlib.util.spec_from_file_location(f"oracle.pyoracle_{name}", __file__)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        mod.FIELD, mod.P, mod._LIB_PATH, mod._lib = name, p, os.path.join(_HERE, libname), None
        _by_field[name] = mod
    return _by_field[name]
UD, JB, CB = 0, 1, 2

u32p = ctypes.POINTER(ctypes.c_uint32)
u8p = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_fmul.restype = ctypes.c_uint32
        L.oracle_fadd.restype = ctypes.c_uint32
        L.oracle_fsub.restype = ctypes.c_uint32
        L.oracle_finv.restype = ctypes.c_uint32
        L.oracle_to_monty.restype = ctypes.c_uint32
        L.oracle_from_monty.restype = ctypes.c_uint32
        L.oracle_two_adic_generator.restype = ctypes.c_uint32
        for name in ("oracle_fmul", "oracle_fadd", "oracle_fsub"):
            getattr(L, name).argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.oracle_finv.argtypes = [ctypes.c_uint32]
        L.oracle_emul.argtypes = [u32p, u32p, u32p]
        L.oracle_to_monty.argtypes = [ctypes.c_uint32]
        L.oracle_from_monty.argtypes = [ctypes.c_uint32]
        L.oracle_two_adic_generator.argtypes = [ctypes.c_uint]
        L.oracle_chacha_block.argtypes = [u32p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, u32p]
        L.oracle_rng_u32.argtypes = [ctypes.c_uint64, ctypes.c_size_t, u32p]
        L.oracle_stdrng_key.argtypes = [ctypes.c_uint64, u32p]
        L.oracle_rand_poly.argtypes = [ctypes.c_uint64, ctypes.c_int, u32p]
        L.oracle_rand_ext.argtypes = [ctypes.c_uint64, ctypes.c_size_t, u32p]
        L.oracle_poseidon2_constants.argtypes = [ctypes.c_int, ctypes.c_uint64, u32p, u32p, u32p, u32p]
        L.oracle_poseidon2_permute.argtypes = [ctypes.c_int, u32p]
        L.oracle_mds_light.argtypes = [ctypes.c_int, u32p]
        L.oracle_internal_layer.argtypes = [ctypes.c_int, u32p]
        L.oracle_sponge24.argtypes = [u32p, ctypes.c_size_t, u32p]
        L.oracle_compress16.argtypes = [u32p, u32p, u32p]
        L.oracle_dft_batch.argtypes = [u32p, ctypes.c_size_t, ctypes.c_size_t]
        L.oracle_dft_batch_naive.argtypes = [u32p, ctypes.c_size_t, ctypes.c_size_t]
        L.oracle_merkle_commit.argtypes = [u32p, ctypes.c_size_t, ctypes.c_size_t, u32p, u32p]
        L.oracle_challenger_probe.argtypes = [u32p, ctypes.c_size_t, ctypes.c_size_t, u32p]
        L.oracle_eq_table.argtypes = [u32p, ctypes.c_int, u32p, u32p]
        L.oracle_eval_base.argtypes = [u32p, ctypes.c_int, u32p, u32p]
        L.oracle_eval_ext.argtypes = [u32p, ctypes.c_int, u32p, u32p]
        L.oracle_compress_hi.argtypes = [u32p, ctypes.c_int, u32p, ctypes.c_int, u32p]
        L.oracle_compress_lo.argtypes = [u32p, ctypes.c_int, u32p, ctypes.c_int, u32p]
        L.oracle_coeffs.argtypes = [u32p, u32p, ctypes.c_int, u32p]
        L.oracle_fix_lo_var.argtypes = [u32p, ctypes.c_int, u32p]
        L.oracle_combine_eqs.argtypes = [u32p, ctypes.c_int, u32p, ctypes.c_int, u32p]
        L.oracle_combine_pows.argtypes = [u32p, ctypes.c_int, u32p, ctypes.c_int, u32p, ctypes.c_int]
        L.oracle_combine_pows_scaled.argtypes = [u32p, ctypes.c_int, u32p, ctypes.c_int, u32p]
        L.oracle_eval_univariate.argtypes = [u32p, ctypes.c_size_t, u32p, u32p]
        L.oracle_n_ood_queries.argtypes = [ctypes.c_int] * 4
        L.oracle_n_stir_queries.argtypes = [ctypes.c_int] * 3
        L.oracle_number_of_rounds.argtypes = [ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.oracle_whir_prove.argtypes = [ctypes.c_int] * 7 + [u32p, u8p, ctypes.c_size_t,
                                                             ctypes.POINTER(ctypes.c_size_t), u32p, u32p]
        L.oracle_whir_verify.argtypes = [ctypes.c_int] * 7 + [u8p, ctypes.c_size_t, u32p]
        i32p = ctypes.POINTER(ctypes.c_int)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.oracle_stack_layout.argtypes = [ctypes.c_int, i32p, ctypes.c_int, i32p, u64p, i32p]
        L.oracle_stack_claim_ranks.argtypes = [ctypes.c_int, i32p, ctypes.c_int, ctypes.c_int, i32p, u8p, u64p]
        L.oracle_stack_virtual_check.argtypes = [ctypes.c_int, i32p, u32p, ctypes.c_int, u32p]
        L.oracle_whir_prove_stacked.argtypes = ([ctypes.c_int] * 7 + [i32p, u32p, ctypes.c_int, i32p, u8p, u8p,
                                                ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), u32p, u32p])
        L.oracle_whir_verify_stacked.argtypes = ([ctypes.c_int] * 7 + [i32p, ctypes.c_int, i32p, u8p, u8p,
                                                 ctypes.c_size_t, u32p])
        L.oracle_whir_prove_session.argtypes = ([ctypes.c_int] * 7 + [i32p, u32p, ctypes.c_int, i32p, u8p,
                                                ctypes.c_int, u32p, u32p, u8p, ctypes.c_size_t,
                                                ctypes.POINTER(ctypes.c_size_t), u32p, u32p, u32p])
        L.oracle_whir_verify_session.argtypes = ([ctypes.c_int] * 7 + [i32p, ctypes.c_int, i32p, u8p, ctypes.c_int,
                                                 u32p, u32p, u8p, ctypes.c_size_t, u32p])
        L.oracle_set_threads.argtypes = [ctypes.c_int]
        L.oracle_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t=u32p):
    return a.ctypes.data_as(t)


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


# ---------------- field ----------------
def fmul(a, b):
    return lib().oracle_fmul(a, b)


def fadd(a, b):
    return lib().oracle_fadd(a, b)


def finv(a):
    return lib().oracle_finv(a)


def to_monty(c):
    return lib().oracle_to_monty(c)


def from_monty(m):
    return lib().oracle_from_monty(m)


def two_adic_generator(bits):
    return lib().oracle_two_adic_generator(bits)


def set_threads(n):
    lib().oracle_set_threads(int(n))


def get_threads():
    return lib().oracle_get_threads()


# ---------------- rng ----------------
def chacha_block(key, counter, stream, rounds):
    out = np.zeros(16, np.uint32)
    k = _u32(key)
    lib().oracle_chacha_block(_p(k), counter, stream, rounds, _p(out))
    return out


def rng_u32(seed, n):
    out = np.zeros(n, np.uint32)
    lib().oracle_rng_u32(seed, n, _p(out))
    return out


def stdrng_key(seed):
    """The 8-word ChaCha key StdRng::seed_from_u64(seed) derives (PCG32 expansion)."""
    out = np.zeros(8, np.uint32)
    lib().oracle_stdrng_key(seed, _p(out))
    return out


def rand_poly(seed, k):
    out = np.zeros(1 << k, np.uint32)
    lib().oracle_rand_poly(seed, k, _p(out))
    return out


def rand_ext(seed, n):
    out = np.zeros(4 * n, np.uint32)
    lib().oracle_rand_ext(seed, n, _p(out))
    return out.reshape(n, 4)


# ---------------- poseidon2 ----------------
ROUNDS_P = {"koalabear": {16: 20, 24: 23}, "babybear": {16: 13, 24: 21}}


def poseidon2_constants(width, seed=1000):
    rp = ROUNDS_P[FIELD][width]
    a = np.zeros(4 * width, np.uint32)
    b = np.zeros(4 * width, np.uint32)
    c = np.zeros(rp, np.uint32)
    d = np.zeros(width, np.uint32)
    lib().oracle_poseidon2_constants(width, seed, _p(a), _p(b), _p(c), _p(d))
    return a.reshape(4, width), b.reshape(4, width), c, d


def permute(width, state):
    s = _u32(state).copy()
    lib().oracle_poseidon2_permute(width, _p(s))
    return s


def mds_light(width, state):
    s = _u32(state).copy()
    lib().oracle_mds_light(width, _p(s))
    return s


def internal_layer(width, state):
    s = _u32(state).copy()
    lib().oracle_internal_layer(width, _p(s))
    return s


def sponge24(values):
    v = _u32(values)
    out = np.zeros(8, np.uint32)
    lib().oracle_sponge24(_p(v), v.size, _p(out))
    return out


def compress16(l, r):
    a, b = _u32(l), _u32(r)
    out = np.zeros(8, np.uint32)
    lib().oracle_compress16(_p(a), _p(b), _p(out))
    return out


def challenger_probe(obs, n_samples):
    o = _u32(obs)
    out = np.zeros(n_samples, np.uint32)
    lib().oracle_challenger_probe(_p(o), o.size, n_samples, _p(out))
    return out


# ---------------- dft / merkle ----------------
def dft_batch(mat):
    m = _u32(mat).copy()
    h, w = m.shape
    lib().oracle_dft_batch(_p(m), h, w)
    return m


def dft_batch_naive(mat):
    m = _u32(mat).copy()
    h, w = m.shape
    lib().oracle_dft_batch_naive(_p(m), h, w)
    return m


def merkle_commit(rows, want_layers=False):
    r = _u32(rows)
    h, w = r.shape
    root = np.zeros(8, np.uint32)
    layers = np.zeros((2 * h - 1, 8), np.uint32) if want_layers else None
    lib().oracle_merkle_commit(_p(r), h, w, _p(root), _p(layers) if want_layers else None)
    return (root, layers) if want_layers else root


# ---------------- poly ----------------
def eq_table(z, scale=None):
    z = _u32(z).reshape(-1, 4)
    k = z.shape[0]
    sc = _u32(scale if scale is not None else [to_monty(1), 0, 0, 0])
    out = np.zeros((1 << k, 4), np.uint32)
    lib().oracle_eq_table(_p(z), k, _p(sc), _p(out))
    return out


def eval_base(poly, z):
    p = _u32(poly)
    z = _u32(z).reshape(-1, 4)
    out = np.zeros(4, np.uint32)
    lib().oracle_eval_base(_p(p), z.shape[0], _p(z), _p(out))
    return out


def eval_ext(poly, z):
    p = _u32(poly).reshape(-1, 4)
    z = _u32(z).reshape(-1, 4)
    out = np.zeros(4, np.uint32)
    lib().oracle_eval_ext(_p(p), z.shape[0], _p(z), _p(out))
    return out


def compress_hi(poly, z):
    p = _u32(poly)
    z = _u32(z).reshape(-1, 4)
    k = int(np.log2(p.size))
    out = np.zeros((1 << (k - z.shape[0]), 4), np.uint32)
    lib().oracle_compress_hi(_p(p), k, _p(z), z.shape[0], _p(out))
    return out


def compress_lo(poly, z):
    p = _u32(poly)
    z = _u32(z).reshape(-1, 4)
    k = int(np.log2(p.size))
    out = np.zeros((1 << (k - z.shape[0]), 4), np.uint32)
    lib().oracle_compress_lo(_p(p), k, _p(z), z.shape[0], _p(out))
    return out


def coeffs(poly, w):
    p = _u32(poly).reshape(-1, 4)
    ww = _u32(w).reshape(-1, 4)
    out = np.zeros(8, np.uint32)
    lib().oracle_coeffs(_p(p), _p(ww), int(np.log2(p.shape[0])), _p(out))
    return out.reshape(2, 4)


def fix_lo_var(poly, r):
    p = _u32(poly).reshape(-1, 4).copy()
    rr = _u32(r)
    k = int(np.log2(p.shape[0]))
    lib().oracle_fix_lo_var(_p(p), k, _p(rr))
    return p[: p.shape[0] // 2].copy()


def combine_eqs(out, points, alpha):
    o = _u32(out).reshape(-1, 4).copy()
    k = int(np.log2(o.shape[0]))
    pts = _u32(points).reshape(-1, k, 4)
    a = _u32(alpha)
    lib().oracle_combine_eqs(_p(o), k, _p(pts), pts.shape[0], _p(a))
    return o


def combine_pows(out, vars_, alpha, shift):
    o = _u32(out).reshape(-1, 4).copy()
    k = int(np.log2(o.shape[0]))
    v = _u32(vars_)
    a = _u32(alpha)
    lib().oracle_combine_pows(_p(o), k, _p(v), v.size, _p(a), shift)
    return o


def combine_pows_scaled(out, vars_, scales):
    o = _u32(out).reshape(-1, 4).copy()
    k = int(np.log2(o.shape[0]))
    v = _u32(vars_)
    sc = _u32(scales).reshape(-1, 4)
    assert sc.shape[0] == v.size
    lib().oracle_combine_pows_scaled(_p(o), k, _p(v), v.size, _p(sc))
    return o


def eval_univariate(poly, x):
    p = _u32(poly).reshape(-1, 4)
    xx = _u32(x)
    out = np.zeros(4, np.uint32)
    lib().oracle_eval_univariate(_p(p), p.shape[0], _p(xx), _p(out))
    return out


# ---------------- params ----------------
def n_ood_queries(soundness, sec, log_degree, log_inv_rate):
    return lib().oracle_n_ood_queries(soundness, sec, log_degree, log_inv_rate)


def n_stir_queries(soundness, sec, log_inv_rate):
    return lib().oracle_n_stir_queries(soundness, sec, log_inv_rate)


def number_of_rounds(f, k):
    a, b = ctypes.c_int(), ctypes.c_int()
    lib().oracle_number_of_rounds(f, k, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


# ---------------- WHIR ----------------
def whir_prove(k, folding, rate, ir, soundness, security, n_points, poly):
    p = _u32(poly)
    assert p.size == 1 << k
    cap = 1 << 16
    while True:
        buf = np.zeros(cap, np.uint8)
        ln = ctypes.c_size_t()
        cp = np.zeros(1, np.uint32)
        root = np.zeros(8, np.uint32)
        rc = lib().oracle_whir_prove(k, folding, rate, ir, soundness, security, n_points, _p(p),
                                     _p(buf, u8p), cap, ctypes.byref(ln), _p(cp), _p(root))
        if rc == 0:
            return bytes(buf[: ln.value]), int(cp[0]), root
        cap = ln.value


def whir_verify(k, folding, rate, ir, soundness, security, n_points, proof):
    b = np.frombuffer(bytes(proof), dtype=np.uint8).copy()
    cp = np.zeros(1, np.uint32)
    rc = lib().oracle_whir_verify(k, folding, rate, ir, soundness, security, n_points,
                                  _p(b, u8p), b.size, _p(cp))
    return rc, int(cp[0])


# ---------------- multi-polynomial stacking ----------------
def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _claims(claims):
    """claims: list of (poly id, fresh point flag) in transcript order."""
    cp = _i32([c[0] for c in claims])
    cn = np.ascontiguousarray(np.asarray([1 if c[1] else 0 for c in claims], dtype=np.uint8))
    return cp, cn


def stack_layout(k_polys, k_svo):
    """(k, offsets per poly id, poly ids in Selector order)."""
    kp = _i32(k_polys)
    kout = np.zeros(1, np.int32)
    offs = np.zeros(len(kp), np.uint64)
    order = np.zeros(len(kp), np.int32)
    i32p = ctypes.POINTER(ctypes.c_int)
    rc = lib().oracle_stack_layout(len(kp), _p(kp, i32p), k_svo, _p(kout, i32p),
                                   _p(offs, ctypes.POINTER(ctypes.c_uint64)), _p(order, i32p))
    if rc:
        raise ValueError("bad stack")
    return int(kout[0]), [int(o) for o in offs], [int(o) for o in order]


def stack_claim_ranks(k_polys, k_svo, claims):
    kp = _i32(k_polys)
    cp, cn = _claims(claims)
    out = np.zeros(len(claims), np.uint64)
    i32p = ctypes.POINTER(ctypes.c_int)
    rc = lib().oracle_stack_claim_ranks(len(kp), _p(kp, i32p), k_svo, len(claims), _p(cp, i32p), _p(cn, u8p),
                                        _p(out, ctypes.POINTER(ctypes.c_uint64)))
    if rc:
        raise ValueError("bad stack")
    return [int(r) for r in out]


def stack_virtual_check(polys, k_svo, x):
    kp = _i32([int(np.log2(len(p))) for p in polys])
    data = _u32(np.concatenate([np.asarray(p, np.uint32) for p in polys]))
    xe = _u32(x)
    i32p = ctypes.POINTER(ctypes.c_int)
    return lib().oracle_stack_virtual_check(len(kp), _p(kp, i32p), _p(data), k_svo, _p(xe))


def whir_prove_stacked(k, folding, rate, ir, soundness, security, polys, claims):
    kp = _i32([int(np.log2(len(p))) for p in polys])
    data = _u32(np.concatenate([np.asarray(p, np.uint32) for p in polys]))
    cp, cn = _claims(claims)
    i32p = ctypes.POINTER(ctypes.c_int)
    cap = 1 << 16
    while True:
        buf = np.zeros(cap, np.uint8)
        ln = ctypes.c_size_t()
        chk = np.zeros(1, np.uint32)
        root = np.zeros(8, np.uint32)
        rc = lib().oracle_whir_prove_stacked(k, folding, rate, ir, soundness, security, len(kp), _p(kp, i32p),
                                             _p(data), len(claims), _p(cp, i32p), _p(cn, u8p), _p(buf, u8p), cap,
                                             ctypes.byref(ln), _p(chk), _p(root))
        if rc == -2:
            raise ValueError("bad stack")
        if rc == 0:
            return bytes(buf[: ln.value]), int(chk[0]), root
        cap = ln.value


def whir_verify_stacked(k, folding, rate, ir, soundness, security, k_polys, claims, proof):
    kp = _i32(k_polys)
    cp, cn = _claims(claims)
    b = np.frombuffer(bytes(proof), dtype=np.uint8).copy()
    chk = np.zeros(1, np.uint32)
    i32p = ctypes.POINTER(ctypes.c_int)
    rc = lib().oracle_whir_verify_stacked(k, folding, rate, ir, soundness, security, len(kp), _p(kp, i32p),
                                          len(claims), _p(cp, i32p), _p(cn, u8p), _p(b, u8p), b.size, _p(chk))
    return rc, int(chk[0])


# ---------------- a proof inside a caller's transcript session ----------------
GIVEN = 2  # claim kind: a point given by the caller (0 shared, 1 drawn)


def _session_claims(claims):
    """claims: (poly id, kind) with kind 0 shared / 1 drawn / 2 given, or a bool
    for drawn/shared."""
    cp = _i32([c[0] for c in claims])
    cn = np.ascontiguousarray(np.asarray([int(c[1]) for c in claims], dtype=np.uint8))
    return cp, cn


def whir_prove_session(k, folding, rate, ir, soundness, security, polys, claims, prefix=(), given=()):
    """The prover inside a caller's transcript: `prefix` words written before the
    commitment, claims of kind 2 take the next point of `given` (k_poly x 4
    words each). Returns (proof bytes including the prefix's, checkpoint, root,
    challenger state after open as a dict of the whir_challenger fields)."""
    kp = _i32([int(np.log2(len(p))) for p in polys])
    data = _u32(np.concatenate([np.asarray(p, np.uint32) for p in polys]))
    cp, cn = _session_claims(claims)
    pre = _u32(np.asarray(list(prefix), np.uint32)) if len(prefix) else np.zeros(1, np.uint32)
    gv = _u32(np.concatenate([np.asarray(g, np.uint32).reshape(-1) for g in given])) if len(given) else \
        np.zeros(1, np.uint32)
    i32p = ctypes.POINTER(ctypes.c_int)
    cap = 1 << 16
    while True:
        buf = np.zeros(cap, np.uint8)
        ln = ctypes.c_size_t()
        chk = np.zeros(1, np.uint32)
        root = np.zeros(8, np.uint32)
        st = np.zeros(40, np.uint32)
        rc = lib().oracle_whir_prove_session(k, folding, rate, ir, soundness, security, len(kp), _p(kp, i32p),
                                             _p(data), len(claims), _p(cp, i32p), _p(cn, u8p), len(prefix), _p(pre),
                                             _p(gv), _p(buf, u8p), cap, ctypes.byref(ln), _p(chk), _p(root), _p(st))
        if rc == -2:
            raise ValueError("bad session")
        if rc == 0:
            state = {"sponge_state": st[:16].copy(), "input_buffer": st[16:24].copy(),
                     "output_buffer": st[24:32].copy(), "input_len": int(st[32]), "output_len": int(st[33])}
            return bytes(buf[: ln.value]), int(chk[0]), root, state
        cap = ln.value


def whir_verify_session(k, folding, rate, ir, soundness, security, k_polys, claims, proof, prefix=(), given=()):
    kp = _i32(k_polys)
    cp, cn = _session_claims(claims)
    pre = _u32(np.asarray(list(prefix), np.uint32)) if len(prefix) else np.zeros(1, np.uint32)
    gv = _u32(np.concatenate([np.asarray(g, np.uint32).reshape(-1) for g in given])) if len(given) else \
        np.zeros(1, np.uint32)
    b = np.frombuffer(bytes(proof), dtype=np.uint8).copy()
    chk = np.zeros(1, np.uint32)
    i32p = ctypes.POINTER(ctypes.c_int)
    rc = lib().oracle_whir_verify_session(k, folding, rate, ir, soundness, security, len(kp), _p(kp, i32p),
                                          len(claims), _p(cp, i32p), _p(cn, u8p), len(prefix), _p(pre), _p(gv),
                                          _p(b, u8p), b.size, _p(chk))
    return rc, int(chk[0])
