"""Device self-checks (whir_prover_set_checks / whir_batch_set_checks): the
reference's debug assertions on the prover path, run on the GPU beside the
proof (the Merkle opening re-verification of MerkleTree::query,
crates/merkle/src/poseidon_packed.rs:158-230, and sum == prod() after each
sumcheck fold, crates/whir/src/sumcheck/expr.rs:68). With the checks on, the
proof bytes are unchanged (still the oracle's) and no check fails; a codeword
overwritten from the host between commit and open makes the opening checks
fail (WHIR_ECHECK) while the unchecked prover returns a proof the oracle's
verifier rejects."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WHIR_ECHECK = 8

CHECKED = [
    (12, 3, 1, 2, 1, 32, 2),     # small rounds only: opening checks of round 0
    (20, 3, 1, 3, 1, 123, 2),    # round 1 on the grid (2^14 coefficients): sumcheck checks
    (22, 4, 1, 3, 1, 123, 1),    # round 1 on the grid (2^14)
    (24, 4, 1, 3, 1, 123, 1),    # the headline shape (round 1: 2^16)
]


@pytest.mark.parametrize("cfg", CHECKED, ids=lambda c: "k{}f{}r{}n{}".format(c[0], c[1], c[2], c[6]))
def test_checked_proof_is_the_oracles(gpu_ctx, oracle, cfg):
    import whir_amd

    k, f, rate, ir, s, sec, n = cfg
    poly = oracle.rand_poly(3, k)
    print(f"checked prove {cfg}", flush=True)
    pr = whir_amd.Prover(whir_amd.params(*cfg), gpu_ctx)
    try:
        pr.set_checks(True)
        pr.upload(poly)
        pr.enqueue()
        got = pr.result()  # raises WhirError(WHIR_ECHECK) if a check failed
        pr.set_graph(True)  # the checks inside a captured graph, replayed
        for _ in range(2):
            pr.enqueue()
            assert pr.result()[0] == got[0]
    finally:
        pr.close()
    want = oracle.whir_prove(*cfg, poly)
    assert got[0] == want[0] and got[1] == want[1]


def test_checked_batch(gpu_ctx, oracle):
    import whir_amd

    cfg = (20, 3, 1, 3, 1, 123, 1)
    polys = [oracle.rand_poly(70 + j, 20) for j in range(2)]
    print("checked batch of 2", flush=True)
    b = whir_amd.Batch(whir_amd.params(*cfg), 2, ctx=gpu_ctx)
    try:
        b.set_checks(True)
        for j, p in enumerate(polys):
            b.upload(j, p)
        b.enqueue()
        got = [b.result(j)[0] for j in range(2)]
    finally:
        b.close()
    assert got == [oracle.whir_prove(*cfg, p)[0] for p in polys]


def _prove_with_cleared_codeword(gpu_ctx, cfg, poly, checks):
    """Session commit, then the base codeword overwritten with zeros from the
    host, then open. The workspace starts with the transcript scalars, the log
    and the 2^k-word input; the base codeword (2^(k + rate) words) follows the
    input (256-byte aligned)."""
    import whir_amd
    from whir_amd._lib import check

    k, f, rate = cfg[0], cfg[1], cfg[2]
    pr = whir_amd.Prover(whir_amd.params(*cfg), gpu_ctx)
    try:
        pr.set_checks(checks)
        pr.upload(poly)
        pr.commit()
        pr.eval(None)
        check(gpu_ctx.L.whir_device_synchronize(), "sync")
        cw0 = pr.input_ptr + (((1 << k) * 4 + 255) & ~255)
        zeros = np.zeros(1 << (k + rate), np.uint32)
        check(gpu_ctx.L.whir_memcpy_h2d(cw0, zeros.ctypes.data, zeros.nbytes), "whir_memcpy_h2d")
        pr.open()
        try:
            return pr.result(), None
        except whir_amd.WhirError as e:
            return None, e.code
    finally:
        pr.close()


def test_cleared_codeword_fails_the_opening_checks(gpu_ctx, oracle):
    cfg = (14, 4, 1, 3, 1, 123, 1)
    poly = oracle.rand_poly(9, 14)
    print("cleared codeword, checks on", flush=True)
    res, code = _prove_with_cleared_codeword(gpu_ctx, cfg, poly, True)
    assert res is None and code == WHIR_ECHECK
    print("cleared codeword, checks off", flush=True)
    res, code = _prove_with_cleared_codeword(gpu_ctx, cfg, poly, False)
    assert code is None
    rc, _ = oracle.whir_verify(*cfg, res[0])
    assert rc != 0  # the verifier rejects what the checks flagged
