"""Stacked-polynomial job shapes shared by the oracle and GPU stacking tests.

A stack is (k_polys, claims): the size of each registered polynomial (ids in
registration order) and the evaluation claims in transcript order as
(poly id, fresh point) pairs, where a claim without a fresh point reuses the
previous claim's point.

`zkvm_shape` follows the claim pattern of the reference's test_zkvm
(crates/whir/src/sumcheck/test.rs:94-147): table A of 20 polys, tables B (29)
and C (96) of small polys, and eight shared points, each evaluated on a subset
of one table's polys. The reference uses k = 20 and k = 8 (stacked k = 25);
`ka` and `kb` scale the two sizes."""


def shared_point(ids, selected):
    return [(ids[i], j == 0) for j, i in enumerate(selected)]


def zkvm_shape(ka=20, kb=8):
    a = list(range(0, 20))
    b = list(range(20, 49))
    c = list(range(49, 145))
    k_polys = [ka] * 20 + [kb] * 125
    claims = []
    claims += shared_point(a, [0] + list(range(2, 20)))  # p0
    claims += shared_point(a, [0, 1])  # p1
    claims += shared_point(a, list(range(20)))  # p2
    claims += shared_point(b, list(range(6, 24)))  # p3
    claims += shared_point(b, list(range(0, 8)) + list(range(24, 29)))  # p4
    claims += shared_point(b, list(range(29)))  # p5
    claims += shared_point(c, list(range(1, 20)) + list(range(88, 96)))  # p6
    claims += shared_point(c, list(range(96)))  # p7
    return k_polys, claims


def stacked_k(k_polys):
    total = sum(1 << k for k in k_polys)
    k = 0
    while (1 << k) < total:
        k += 1
    return k


# Small mixed stacks: (k_polys, claims)
MIXED = [
    # two equal polys, one claim each at separate points
    ([6, 6], [(0, True), (1, True)]),
    # three sizes, a shared point over two equal polys, a poly with no claim
    ([8, 5, 8, 4], [(2, True), (0, False), (1, True), (1, True)]),
    # total not a power of two (padding), the later registered large poly first
    ([9, 7, 7, 7, 9, 5], [(4, True), (0, True), (5, True), (1, True), (2, False), (3, False)]),
    # one big and many tiny polys
    ([10] + [4] * 9, [(0, True)] + [(i, i == 1) for i in range(1, 10)] + [(0, True)]),
]


def rand_polys(oracle, k_polys, seed=1000):
    return [oracle.rand_poly(seed + i, k) for i, k in enumerate(k_polys)]
