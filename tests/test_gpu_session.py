"""The drop-in boundary Whir::commit / ProverStack::eval / Whir::open
(crates/whir/src/whir.rs:127-180, crates/whir/src/sumcheck/prover.rs:64-90) on
the GPU: whir_prover_commit continues a caller's transcript (a DuplexChallenger
state written on the CPU), whir_prover_eval takes drawn or caller-given claim
points, whir_prover_open hands the transcript back. Every case is compared with
the oracle's session prover (oracle.whir_prove_session): the GPU's bytes are
the oracle proof after the caller's prefix, the roots, checkpoints and
handed-back challenger states are equal, and the oracle's session verifier
accepts prefix + GPU bytes."""
import numpy as np
import pytest

from kbpy import DuplexPy
from tests.stack_specs import MIXED, rand_polys, stacked_k

pytestmark = pytest.mark.gpu

P = 0x7F000001


def _state_after_prefix(oracle, prefix):
    d = DuplexPy(oracle)
    for x in prefix:
        d.observe(int(x))
    return d.as_state()


def _same_state(a, b):
    return (np.array_equal(np.asarray(a["sponge_state"]), np.asarray(b["sponge_state"]))
            and a["input_len"] == b["input_len"] and a["output_len"] == b["output_len"]
            and list(a["input_buffer"][: a["input_len"]]) == list(b["input_buffer"][: b["input_len"]])
            and list(a["output_buffer"][: a["output_len"]]) == list(b["output_buffer"][: b["output_len"]]))


def _gpu_session(gpu_ctx, prover, d_poly, state, kinds, given):
    prover.commit(d_poly, state)
    gi = 0
    for kind in kinds:
        if kind == 2:
            prover.eval(given[gi])
            gi += 1
        else:
            prover.eval(None)
    prover.open()
    proof, cp, root = prover.result()
    return proof, cp, root, prover.challenger()


SESSIONS = [
    # k, folding, rate, ir, soundness, security, claim kinds (1 drawn, 2 given), prefix words
    (12, 3, 1, 2, 1, 32, [1, 2, 1], 5),
    (14, 4, 1, 3, 1, 123, [2], 0),
    (14, 4, 2, 3, 1, 123, [2, 1, 2, 2], 8),
    (10, 2, 1, 1, 0, 32, [1, 1], 3),       # unique decoding: no virtual claims
    (18, 4, 1, 3, 1, 123, [2, 1], 11),
]


@pytest.mark.parametrize("cfg", SESSIONS, ids=lambda c: "k{}f{}r{}-{}-p{}".format(c[0], c[1], c[2],
                                                                               "".join(map(str, c[6])), c[7]))
def test_session_matches_oracle(gpu_ctx, oracle, cfg):
    import whir_amd

    k, f, rate, ir, s, sec, kinds, n_prefix = cfg
    rng = np.random.default_rng(k * 100 + n_prefix)
    poly = oracle.rand_poly(9, k)
    prefix = rng.integers(0, P, n_prefix, dtype=np.uint64).astype(np.uint32)
    given = [rng.integers(0, P, 4 * k, dtype=np.uint64).astype(np.uint32).reshape(k, 4)
             for kk in kinds if kk == 2]
    claims = [(0, kk) for kk in kinds]
    oproof, ocp, oroot, ost = oracle.whir_prove_session(k, f, rate, ir, s, sec, [poly], claims, prefix, given)
    print(f"session k={k} f={f} rate={rate} claims={kinds} prefix={n_prefix}", flush=True)
    pr = whir_amd.Prover(whir_amd.params(k, f, rate, ir, s, sec, len(kinds)), gpu_ctx)
    try:
        pr.upload(poly)
        state = _state_after_prefix(oracle, prefix) if n_prefix else None
        gproof, gcp, groot, gst = _gpu_session(gpu_ctx, pr, None, state, kinds, given)
        assert gproof == oproof[4 * n_prefix:]
        assert np.array_equal(groot, oroot) and gcp == ocp
        assert _same_state(gst, ost)
        rc, vcp = oracle.whir_verify_session(k, f, rate, ir, s, sec, [k], claims, oproof[:4 * n_prefix] + gproof,
                                             prefix, given)
        assert rc == 0 and vcp == gcp
        # the same prover then runs the fused test-driver proof unchanged
        pr.enqueue()
        assert pr.result()[0] == oracle.whir_prove(k, f, rate, ir, s, sec, len(kinds), poly)[0]
    finally:
        pr.close()


def test_plain_session_equals_fused_proof(gpu_ctx, oracle):
    """commit (fresh transcript) + drawn evals + open = whir_prover_enqueue."""
    import whir_amd

    k, f, rate, ir, s, sec, n = 13, 4, 1, 3, 1, 123, 3
    poly = oracle.rand_poly(4, k)
    pr = whir_amd.Prover(whir_amd.params(k, f, rate, ir, s, sec, n), gpu_ctx)
    try:
        pr.upload(poly)
        pr.enqueue()
        fused = pr.result()
        sess = _gpu_session(gpu_ctx, pr, None, None, [1] * n, [])
        assert sess[0] == fused[0] and sess[1] == fused[1]
        want = oracle.whir_prove(k, f, rate, ir, s, sec, n, poly)
        assert fused[0] == want[0]
        # the handed-back state draws the checkpoint next
        assert int(DuplexPy(oracle, sess[3]).sample()) == fused[1]
    finally:
        pr.close()


def test_stacked_session(gpu_ctx, oracle):
    """A stack with shared-point claim groups; each group's point given."""
    import whir_amd

    k_polys, claims = MIXED[1]
    f, rate, ir, s, sec = 3, 1, 2, 1, 32
    k = stacked_k(k_polys)
    polys = rand_polys(oracle, k_polys, 21)
    rng = np.random.default_rng(5)
    kinds = [2 if fresh else 0 for _, fresh in claims]
    given = [rng.integers(0, P, 4 * k_polys[pid], dtype=np.uint64).astype(np.uint32).reshape(-1, 4)
             for (pid, fresh) in claims if fresh]
    sclaims = [(pid, kk) for (pid, _), kk in zip(claims, kinds)]
    prefix = np.array([1, 2, 3, 4, 5, 6, 7], np.uint32)
    oproof, ocp, oroot, ost = oracle.whir_prove_session(k, f, rate, ir, s, sec, polys, sclaims, prefix, given)
    print(f"stacked session k={k} polys={k_polys}", flush=True)
    pr = whir_amd.StackedProver(whir_amd.params(k, f, rate, ir, s, sec, len(claims)), k_polys, claims, ctx=gpu_ctx)
    try:
        stacked = np.zeros(1 << k, np.uint32)
        for pid, poly in enumerate(polys):
            off = pr.offset(pid)
            stacked[off: off + len(poly)] = poly
        d = whir_amd.api.DeviceBuffer(stacked.nbytes)
        d.upload(stacked)
        gproof, gcp, groot, gst = _gpu_session(gpu_ctx, pr, d.ptr, _state_after_prefix(oracle, prefix), kinds,
                                               given)
        assert gproof == oproof[4 * len(prefix):]
        assert gcp == ocp and _same_state(gst, ost)
        rc, _ = oracle.whir_verify_session(k, f, rate, ir, s, sec, k_polys, sclaims, oproof[:28] + gproof, prefix,
                                           given)
        assert rc == 0
    finally:
        pr.close()


def test_session_order_and_state_errors(gpu_ctx, oracle):
    import whir_amd
    from whir_amd._lib import lib

    k, f, rate, ir, s, sec = 12, 3, 1, 2, 1, 32
    pr = whir_amd.Prover(whir_amd.params(k, f, rate, ir, s, sec, 2), gpu_ctx)
    try:
        pr.upload(oracle.rand_poly(1, k))
        with pytest.raises(whir_amd.WhirError):
            pr.eval(None)  # before commit
        bad = DuplexPy(oracle).as_state()
        bad["output_len"] = 3  # an output buffer that is not the state's prefix
        bad["output_buffer"][:3] = [1, 2, 3]
        with pytest.raises(whir_amd.WhirError):
            pr.commit(None, bad)
        bad2 = DuplexPy(oracle).as_state()
        bad2["sponge_state"][0] = P  # not canonical
        with pytest.raises(whir_amd.WhirError):
            pr.commit(None, bad2)
        pr.commit(None, None)
        pr.eval(None)
        with pytest.raises(whir_amd.WhirError):
            pr.open()  # one claim still to evaluate
        pr.eval(None)
        with pytest.raises(whir_amd.WhirError):
            pr.eval(None)  # no third claim
        pr.open()
        proof, cp, _ = pr.result()
        assert proof == oracle.whir_prove(k, f, rate, ir, s, sec, 2, oracle.rand_poly(1, k))[0]
        assert lib().whir_prover_open(pr.ptr, None) != 0  # the session has ended
    finally:
        pr.close()
