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
