"""Stacked multi-polynomial proofs on the GPU (whir_prover_create_stacked /
whir_prover_enqueue_stacked) against the oracle: proof bytes, checkpoint and
root identical, the oracle's stacked verifier accepts, and the layout offsets
agree. Shapes from tests/stack_specs.py, including the reference's test_zkvm
claim pattern at its own sizes (20 polys of 2^20, 125 of 2^8, stacked k = 25)."""
import numpy as np
import pytest

from tests.stack_specs import MIXED, rand_polys, stacked_k, zkvm_shape

pytestmark = pytest.mark.gpu

CASES = [
    # (stack index or "zkvm"), folding, rate, ir, soundness, security, poly seed
    (0, 2, 1, 1, 1, 32, 3),
    (1, 3, 1, 2, 1, 32, 7),
    (2, 4, 1, 3, 1, 123, 8),
    (3, 2, 2, 1, 2, 32, 9),
    (3, 4, 1, 3, 0, 32, 10),
    ("zkvm-small", 4, 1, 3, 1, 123, 1000),
]


def _shape(which):
    if which == "zkvm-small":
        return zkvm_shape(14, 6)
    if which == "zkvm":
        return zkvm_shape()
    return MIXED[which]


def _prove_gpu(gpu_ctx, k, f, rate, ir, s, sec, k_polys, claims, polys, repeat=1):
    import whir_amd

    p = whir_amd.params(k, f, rate, ir, s, sec, len(claims))
    pr = whir_amd.StackedProver(p, k_polys, claims, ctx=gpu_ctx)
    try:
        pr.upload(polys)
        outs = []
        for _ in range(repeat):
            pr.enqueue()
            outs.append(pr.result())
        offs = [pr.offset(i) for i in range(len(k_polys))]
    finally:
        pr.close()
    return outs, offs


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}-f{c[1]}-r{c[2]}-s{c[4]}")
def test_stacked_matches_oracle(gpu_ctx, oracle, case):
    which, f, rate, ir, s, sec, seed = case
    k_polys, claims = _shape(which)
    k = stacked_k(k_polys)
    polys = rand_polys(oracle, k_polys, seed)
    print(f"stacked {which}: k={k} polys={len(k_polys)} claims={len(claims)} f={f}", flush=True)
    want, wcp, wroot = oracle.whir_prove_stacked(k, f, rate, ir, s, sec, polys, claims)
    (got,), offs = _prove_gpu(gpu_ctx, k, f, rate, ir, s, sec, k_polys, claims, polys)
    proof, cp, root = got
    assert offs == oracle.stack_layout(k_polys, f)[1]
    assert len(proof) == len(want)
    assert proof == want
    assert cp == wcp and np.array_equal(root, wroot)
    rc, vcp = oracle.whir_verify_stacked(k, f, rate, ir, s, sec, k_polys, claims, proof)
    assert rc == 0 and vcp == cp


def test_one_poly_with_shared_points(gpu_ctx, oracle):
    # the same poly evaluated twice at one point, then at a fresh one; the
    # prover is reused, so the second proof checks the workspace is rebuilt
    claims = [(0, True), (0, False), (0, True)]
    poly = oracle.rand_poly(4, 12)
    want, wcp, _ = oracle.whir_prove_stacked(12, 3, 1, 2, 1, 32, [poly], claims)
    outs, _ = _prove_gpu(gpu_ctx, 12, 3, 1, 2, 1, 32, [12], claims, [poly], repeat=2)
    for proof, cp, _ in outs:
        assert proof == want and cp == wcp


def test_zkvm_shape_full_size(gpu_ctx, oracle):
    # test_zkvm's stack (sumcheck/test.rs:94-147) inside a WHIR proof: k = 25,
    # folding 4, rate 1/2, JB-123
    k_polys, claims = zkvm_shape()
    k = stacked_k(k_polys)
    polys = rand_polys(oracle, k_polys, 2000)
    print(f"stacked zkvm: k={k} claims={len(claims)}", flush=True)
    want, wcp, wroot = oracle.whir_prove_stacked(k, 4, 1, 3, 1, 123, polys, claims)
    (got,), _ = _prove_gpu(gpu_ctx, k, 4, 1, 3, 1, 123, k_polys, claims, polys)
    assert got[0] == want and got[1] == wcp and np.array_equal(got[2], wroot)


def test_bad_stacks_rejected(gpu_ctx):
    import whir_amd

    bad = [
        ([6, 5], [(0, True), (1, False)], 7),   # shared point across sizes
        ([6, 5], [(0, False)], 7),              # first claim reuses a point
        ([6, 1], [(0, True)], 7),               # poly smaller than the folding factor
        ([6, 5], [(0, True)], 8),               # k is not the stacked k
        ([6, 5], [(2, True)], 7),               # no such poly
    ]
    for k_polys, claims, k in bad:
        with pytest.raises(whir_amd.WhirError):
            whir_amd.StackedProver(whir_amd.params(k, 2, 1, 1, 1, 32, 1), k_polys, claims, ctx=gpu_ctx)
