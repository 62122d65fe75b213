"""whir_compress_hi_noalloc (no counterpart in the reference: the allocation-free
form of SplitEq::compress_hi, crates/poly/src/eq.rs:290-398) through the C ABI:
two calls in a row, the second on a smaller input, each equal to
whir_compress_hi on the same input and to the oracle; a call whose scratch
would exceed the static array is rejected before anything is launched."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 0x7F000001


def _dev(words):
    import whir_amd

    a = np.ascontiguousarray(np.asarray(words, np.uint32).reshape(-1))
    d = whir_amd.api.DeviceBuffer(a.nbytes)
    d.upload(a)
    return d


def _run(fn, ctx, d_poly, k, d_z, m):
    import whir_amd
    from whir_amd._lib import lib

    out = whir_amd.api.DeviceBuffer((4 << (k - m)) * 4)
    rc = fn(ctx.ptr, d_poly.ptr, k, d_z.ptr, m, out.ptr, None)
    assert rc == 0
    assert lib().whir_device_synchronize() == 0
    return out.download(np.uint32, 4 << (k - m)).reshape(-1, 4)


def test_noalloc_matches_allocating_entry_point(gpu_ctx, oracle):
    from whir_amd._lib import lib

    rng = np.random.default_rng(77)
    for k, m in ((20, 16), (12, 10)):  # the second call is smaller
        poly = rng.integers(0, P, 1 << k, dtype=np.uint64).astype(np.uint32)
        z = rng.integers(0, P, 4 * m, dtype=np.uint64).astype(np.uint32).reshape(m, 4)
        print(f"compress_hi_noalloc k={k} m={m}", flush=True)
        d_poly, d_z = _dev(poly), _dev(z)
        got = _run(lib().whir_compress_hi_noalloc, gpu_ctx, d_poly, k, d_z, m)
        want = _run(lib().whir_compress_hi, gpu_ctx, d_poly, k, d_z, m)
        assert np.array_equal(got, want)
        assert np.array_equal(got, oracle.compress_hi(poly, z))


def test_noalloc_rejects_oversized_scratch(gpu_ctx):
    from whir_amd._lib import lib

    d = _dev(np.zeros(64, np.uint32))
    # k - m = 5 and m = 29: 2^15 blocks x 32 partials x 16 bytes alone fill the
    # 16 MiB array; rejected on the host, before any launch
    rc = lib().whir_compress_hi_noalloc(gpu_ctx.ptr, d.ptr, 34, d.ptr, 29, d.ptr, None)
    assert rc == 4  # WHIR_EINVAL
