"""GPU parity of the individual HIP kernels against the CPU oracle, bit-exact.

Poseidon2 (per-lane W16/W24 and the transcript's cooperative W16), the
column-batched DFT (Radix2DFTSmallBatch semantics, whir.rs:34,59) with rate
padding, and the split-half Poseidon2 Merkle tree (poseidon_packed.rs)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 0x7F000001


def _rand_field(rng, shape):
    return rng.integers(0, P, size=shape, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("width", [16, 24])
def test_poseidon2_lanes(gpu_ctx, oracle, width):
    import whir_amd

    rng = np.random.default_rng(width)
    states = _rand_field(rng, (1000, width))
    states[0] = 0
    print(f"poseidon2 width={width}", flush=True)
    got = whir_amd.poseidon2_permute(states, width, ctx=gpu_ctx)
    want = np.stack([oracle.permute(width, s) for s in states])
    assert np.array_equal(got, want)


def test_poseidon2_cooperative16(gpu_ctx, oracle):
    import whir_amd

    rng = np.random.default_rng(7)
    states = _rand_field(rng, (64, 16))
    print("poseidon2 cooperative16", flush=True)
    got = whir_amd.poseidon2_permute(states, 16, cooperative=True, ctx=gpu_ctx)
    want = np.stack([oracle.permute(16, s) for s in states])
    assert np.array_equal(got, want)


def test_poseidon2_wave16(gpu_ctx, oracle):
    """The transcript's one-wavefront P16 (linearised internal rounds), on
    random states and on states with words near p and 0 (the accumulators'
    reduction bounds)."""
    import whir_amd

    rng = np.random.default_rng(11)
    states = _rand_field(rng, (96, 16))
    states[64:80] = P - 1 - rng.integers(0, 4, (16, 16), dtype=np.uint32)
    states[80:96] = rng.integers(0, 4, (16, 16), dtype=np.uint32)
    print("poseidon2 wave16", flush=True)
    got = whir_amd.poseidon2_permute(states, 16, wave=True, ctx=gpu_ctx)
    want = np.stack([oracle.permute(16, s) for s in states])
    assert np.array_equal(got, want)


DFT_CASES = [
    # (n_in rows, log_h, width)
    (1, 1, 1),
    (2, 2, 2),
    (16, 5, 16),
    (64, 7, 16),
    (1 << 10, 11, 16),
    (1 << 12, 13, 16),
    (1 << 14, 15, 64),
    (16, 15, 64),     # tiny polynomial, large rate (late WHIR rounds)
    (1 << 15, 17, 8),
    (1 << 16, 17, 16),
    (1 << 13, 16, 32),
    (1 << 20, 21, 16),  # the base encode of 2^24 coefficients (default plan: register pass + two LDS passes)
    # ragged inputs: first-pass lines whose rows run past n_in (clamped loads, zeroed rows)
    (3, 5, 8),
    (1000, 12, 16),
    ((1 << 16) - 7, 18, 16),
    ((1 << 20) - 1, 21, 16),
]


@pytest.mark.parametrize("n_in,log_h,width", DFT_CASES)
def test_dft_batch(gpu_ctx, oracle, n_in, log_h, width):
    import whir_amd

    rng = np.random.default_rng(n_in * 31 + log_h)
    mat = _rand_field(rng, (n_in, width))
    print(f"dft n_in={n_in} log_h={log_h} width={width}", flush=True)
    got = whir_amd.dft_batch(mat, log_h, ctx=gpu_ctx)
    padded = np.zeros((1 << log_h, width), np.uint32)
    padded[:n_in] = mat
    want = oracle.dft_batch(padded)
    assert np.array_equal(got, want)


# pass layouts of whir_ctx_set_ntt_tuning / whir_ctx_set_ntt_shape on one fresh
# context: (max_log_pass, log_threads, first_pass_log, log_vec); a small
# max_log_pass makes these shapes take three passes
NTT_LAYOUTS = [(5, 8, 4, 2), (5, 8, 3, 2), (5, 7, 2, 1), (4, 6, 4, 0), (8, 8, 4, 2), (6, 8, 0, 1)]
LAYOUT_CASES = [(1 << 12, 13, 16), (1 << 11, 13, 64), (1 << 13, 14, 8), (1 << 10, 12, 4), (1 << 17, 18, 16)]

# two passes of up to 1024 rows (1: three register stages per pass; 2: two
# register stages of up to 32 points, 256-thread workgroups, for >= 16 columns)
# against three passes, for the transform sizes where the choice exists
# (2^17..2^20 rows)
TWO_PASS_CASES = [(1 << 17, 18, 16), (1 << 18, 19, 4), ((1 << 18) - 5, 19, 16), (1 << 19, 21, 2), (1 << 20, 21, 16),
                  (1 << 20, 20, 1), (1 << 17, 19, 64), (1 << 19, 20, 16), (1 << 18, 20, 16), ((1 << 20) - 3, 21, 16),
                  (1 << 19, 21, 64)]


@pytest.mark.parametrize("two_pass", [0, 1, 2])
def test_dft_two_pass_plan(oracle, two_pass):
    import whir_amd
    from whir_amd._lib import lib

    ctx = whir_amd.Context(0, 1000)
    try:
        assert lib().whir_ctx_set_ntt_passes(ctx.ptr, two_pass) == 0
        for layout in ((8, 8), (8, 7), (8, 6)):
            assert lib().whir_ctx_set_ntt_tuning(ctx.ptr, *layout) == 0
            for n_in, log_h, width in TWO_PASS_CASES:
                if layout != (8, 8) and n_in < (1 << 19):
                    continue
                rng = np.random.default_rng(n_in + log_h + width)
                mat = _rand_field(rng, (n_in, width))
                print(f"dft two_pass={two_pass} tuning={layout} n_in={n_in} log_h={log_h} width={width}", flush=True)
                got = whir_amd.dft_batch(mat, log_h, ctx=ctx)
                padded = np.zeros((1 << log_h, width), np.uint32)
                padded[:n_in] = mat
                assert np.array_equal(got, oracle.dft_batch(padded)), (two_pass, layout, n_in, log_h, width)
    finally:
        ctx.close()


@pytest.mark.parametrize("layout", NTT_LAYOUTS, ids=lambda t: "-".join(map(str, t)))
def test_dft_pass_layouts(oracle, layout):
    import whir_amd
    from whir_amd._lib import lib

    ctx = whir_amd.Context(0, 1000)
    try:
        assert lib().whir_ctx_set_ntt_tuning(ctx.ptr, layout[0], layout[1]) == 0
        assert lib().whir_ctx_set_ntt_shape(ctx.ptr, layout[2], layout[3]) == 0
        for n_in, log_h, width in LAYOUT_CASES:
            rng = np.random.default_rng(n_in + log_h + width)
            mat = _rand_field(rng, (n_in, width))
            print(f"dft layout={layout} n_in={n_in} log_h={log_h} width={width}", flush=True)
            got = whir_amd.dft_batch(mat, log_h, ctx=ctx)
            padded = np.zeros((1 << log_h, width), np.uint32)
            padded[:n_in] = mat
            assert np.array_equal(got, oracle.dft_batch(padded)), (layout, n_in, log_h, width)
    finally:
        ctx.close()


@pytest.mark.parametrize("log_h,row_words", [(0, 16), (1, 2), (3, 16), (4, 8), (6, 64), (10, 16), (11, 64), (13, 16),
                                             (15, 64), (16, 16), (17, 32)])
def test_merkle_commit(gpu_ctx, oracle, log_h, row_words):
    import whir_amd

    rng = np.random.default_rng(log_h * 7 + row_words)
    rows = _rand_field(rng, (1 << log_h, row_words))
    print(f"merkle log_h={log_h} row_words={row_words}", flush=True)
    root, layers = whir_amd.merkle_commit(rows, ctx=gpu_ctx)
    want_root, want_layers = oracle.merkle_commit(rows, want_layers=True)
    assert np.array_equal(layers, want_layers)
    assert np.array_equal(root, want_root)
