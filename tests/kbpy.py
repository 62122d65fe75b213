"""Pure-Python KoalaBear / quartic-extension arithmetic on canonical integers,
used by the CPU tests as an independent check of the oracle. Values are
canonical residues; the oracle's words are Montgomery representatives
(x * 2^32 mod p), converted with to_m / from_m."""
P = 0x7F000001
R = (1 << 32) % P
RINV = pow(R, -1, P)
W = 3


def to_m(x):
    return (x % P) * R % P


def from_m(m):
    return m * RINV % P


def e_from_m(words):
    return [from_m(int(w)) for w in words]


def e_to_m(e):
    return [to_m(c) for c in e]


def eadd(a, b):
    return [(x + y) % P for x, y in zip(a, b)]


def esub(a, b):
    return [(x - y) % P for x, y in zip(a, b)]


def emul(a, b):
    t = [0] * 7
    for i in range(4):
        for j in range(4):
            t[i + j] += a[i] * b[j]
    return [(t[0] + W * t[4]) % P, (t[1] + W * t[5]) % P, (t[2] + W * t[6]) % P, t[3] % P]


def escale(a, c):
    return [x * c % P for x in a]


ONE = [1, 0, 0, 0]
ZERO = [0, 0, 0, 0]


def epow(a, e):
    r = ONE
    while e:
        if e & 1:
            r = emul(r, a)
        a = emul(a, a)
        e >>= 1
    return r


def eq_naive(z, x):
    """eq(z, x) = prod_j (x_j ? z_j : 1 - z_j), bit j of x bound to z_j."""
    acc = ONE
    for j, zj in enumerate(z):
        acc = emul(acc, zj if (x >> j) & 1 else esub(ONE, zj))
    return acc


def ext_list_from_m(arr):
    return [e_from_m(row) for row in arr.reshape(-1, 4)]


class DuplexPy:
    """DuplexChallenger<F, P16, 16, 8> restated over the oracle's permutation
    (observe clears the output and absorbs at 8 inputs; sample duplexes when
    input is pending or the output is empty, then pops from the end)."""

    def __init__(self, oracle, state=None):
        import numpy as np

        self.o = oracle
        self.state = np.zeros(16, np.uint32)
        self.inp, self.out = [], []
        if state is not None:  # a whir_challenger dict
            self.state = np.array(state["sponge_state"], np.uint32)
            self.inp = [int(x) for x in state["input_buffer"][: state["input_len"]]]
            self.out = [int(x) for x in state["output_buffer"][: state["output_len"]]]

    def duplex(self):
        for i, v in enumerate(self.inp):
            self.state[i] = v
        self.inp = []
        self.state = self.o.permute(16, self.state)
        self.out = list(self.state[:8])

    def observe(self, v):
        self.out = []
        self.inp.append(v)
        if len(self.inp) == 8:
            self.duplex()

    def sample(self):
        if self.inp or not self.out:
            self.duplex()
        return self.out.pop()

    def as_state(self):
        """The whir_challenger fields of this state."""
        import numpy as np

        ib = np.zeros(8, np.uint32)
        ob = np.zeros(8, np.uint32)
        ib[: len(self.inp)] = self.inp
        ob[: len(self.out)] = self.out
        return {"sponge_state": np.array(self.state, np.uint32), "input_buffer": ib, "output_buffer": ob,
                "input_len": len(self.inp), "output_len": len(self.out)}
