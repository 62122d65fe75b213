"""Host model of the one-wavefront Poseidon2-16 (permute16_wave in
whir_amd/csrc/tx.h, coefficient columns from wave16_consts_build in
poseidon2.h), checked against the oracle permutation.

The model runs the internal rounds the way the kernel does: 64 lanes, each
with one 64-bit accumulator (lane 16 + r: round r's S-box input, lane i < 16:
output word i), seeded from input words 1..15, one multiply-add per round with
the cube of lane 16 + r, a min-subtract on the high word after each term and a
Montgomery REDC. Every REDC input is asserted below p * 2^32 and every high
word below 2p before its reduction, on random states and on states whose
words sit at 0 and at the top of the near-canonical range [0, p + 2^7) that
the external rounds hand over. The external rounds are the word-space
composition of test_oracle_poseidon2 (MDS-light, x^3 with R^-2 in Montgomery
words), itself pinned to the oracle; the device form is pinned bit for bit by
tests/test_gpu_kernels.py::test_poseidon2_wave16."""
import numpy as np
import pytest

from kbpy import P, R, RINV
from test_oracle_poseidon2 import vdiag

NPINV = (-pow(P, -1, 1 << 32)) % (1 << 32)
M32 = (1 << 32) - 1


def redc(t):
    assert 0 <= t < P << 32, "REDC input bound"
    m = (t * NPINV) & M32
    u = t + m * P
    assert u < 1 << 64 and u & M32 == 0
    return u >> 32


def canon2(t):
    assert t < 2 * P
    return t - P if t >= P else t


def mulr(a, b):
    return redc(a * b)


def hi_reduce(t):
    h, lo = t >> 32, t & M32
    assert h < 2 * P, "high word bound"
    return ((h - P if h >= P else h) << 32) | lo


def cube_w(w):  # the S-box on a Montgomery word
    return w ** 3 * RINV * RINV % P


def m4(x):
    s = sum(x)
    return [(s + x[a] + 2 * x[(a + 1) % 4]) % P for a in range(4)]


def mds(x):
    y = []
    for c in range(4):
        y += m4(x[4 * c:4 * c + 4])
    col = [sum(y[4 * c + a] for c in range(4)) for a in range(4)]
    return [(y[i] + col[i % 4]) % P for i in range(16)]


class Model:
    def __init__(self, oracle):
        ei, et, ic, _ = oracle.poseidon2_constants(16, 1000)
        self.ei = [[int(v) for v in r] for r in ei]
        self.et = [[int(v) for v in r] for r in et]
        self.ic = [int(v) for v in ic]
        V = vdiag(16)
        N = 16
        A = [[(1 + (V[i] if i == k else 0)) % P if k else 0 for k in range(N)] for i in range(N)]
        b = [(1 + (V[i] if i == 0 else 0)) % P for i in range(N)]
        pw = [[[int(i == k) for k in range(N)] for i in range(N)]]
        for _ in range(20):
            prev = pw[-1]
            pw.append([[sum(A[i][j] * prev[j][k] for j in range(N)) % P for k in range(N)] for i in range(N)])
        apply_b = lambda n, i: sum(pw[n][i][j] * b[j] for j in range(N)) % P
        # the lazy scales of the external rounds (lazy_consts_build)
        a = RINV
        for _ in range(4):
            a = a ** 3 * RINV % P
        self.sigma = sigma = a
        for _ in range(20):
            a = a * RINV % P
        self.a_t = a_t = a
        si = pow(sigma, -1, P)
        col = [[0] * 36 for _ in range(64)]
        for i in range(16):
            ra = R * a_t % P
            for j in range(20):
                col[i][j] = ra * apply_b(19 - j, i) * si ** 3 % P
            for k in range(1, 16):
                col[i][19 + k] = ra * pw[20][i][k] * si % P
            col[i][35] = ra * self.et[0][i] % P
        for r in range(1, 20):
            c = col[16 + r]
            for j in range(r):
                c[j] = R * apply_b(r - 1 - j, 0) * si * si % P
            for k in range(1, 16):
                c[19 + k] = R * pw[r][0][k] % P
            c[35] = R * sigma * self.ic[r] % P
        self.col = col

    def reference(self, st):
        x = mds(st)
        for r in range(4):
            x = mds([cube_w((x[i] + self.ei[r][i]) % P) for i in range(16)])
        u0 = x
        for r in range(20):
            x = list(x)
            x[0] = cube_w((x[0] + self.ic[r]) % P)
            s = sum(x) % P
            x = [(vdiag(16)[i] * x[i] + s) % P for i in range(16)]
        for r in range(4):
            x = mds([cube_w((x[i] + self.et[r][i]) % P) for i in range(16)])
        return u0, x

    def internal(self, xt):
        """xt: the 16 scaled words entering the internal rounds (< p + 2^7);
        returns the 16 scaled words entering the terminal rounds."""
        acc = [self.col[L][35] for L in range(64)]
        for k in range(1, 16):
            acc = [acc[L] + self.col[L][19 + k] * xt[k] for L in range(64)]
            assert all(t < 1 << 64 for t in acc)
            if k >= 4 and (k % 2 == 0 or k == 15):
                acc = [hi_reduce(t) for t in acc]
        xr = [xt[0]] * 64  # only lane 0's word matters in round 0
        for r in range(20):
            y = mulr(mulr(xr[0 if r == 0 else 16 + r], xr[0 if r == 0 else 16 + r]), xr[0 if r == 0 else 16 + r])
            assert y < 2 * P
            acc = [hi_reduce(acc[L] + self.col[L][r] * y) for L in range(64)]
            xr = [canon2(redc(t)) for t in acc]
        return xr[:16]


@pytest.fixture(scope="module")
def model(oracle):
    return Model(oracle)


def test_word_space_model_matches_oracle(model, oracle):
    rng = np.random.default_rng(21)
    for st in [[int(v) for v in rng.integers(0, P, 16)] for _ in range(4)] + [[0] * 16, [P - 1] * 16]:
        _, want = model.reference(st)
        assert want == [int(v) for v in oracle.permute(16, np.array(st, np.uint32))]


def _inputs():
    """Scaled words entering the internal rounds, as the external rounds hand
    them over: near-canonical, in [0, p + 2^7)."""
    rng = np.random.default_rng(22)
    top = P + (1 << 7)
    out = [[int(v) for v in rng.integers(0, top, 16)] for _ in range(24)]
    out.append([top - 1] * 16)
    out.append([0] * 16)
    out.append([P + int(v) for v in rng.integers(0, 1 << 7, 16)])
    out.append([int(v) if i % 2 else top - 1 for i, v in enumerate(rng.integers(0, 3, 16))])
    return out


def test_linearised_internal_rounds(model):
    si = pow(model.sigma, -1, P)
    for n, xt in enumerate(_inputs()):
        x = [v * si % P for v in xt]  # u0, with round 0's constant on word 0
        x[0] = (x[0] - model.ic[0]) % P
        for r in range(20):
            x[0] = cube_w((x[0] + model.ic[r]) % P)
            s = sum(x) % P
            x = [(vdiag(16)[i] * x[i] + s) % P for i in range(16)]
        want = [model.a_t * (x[i] + model.et[0][i]) % P for i in range(16)]
        assert model.internal(xt) == want, f"input {n}"
