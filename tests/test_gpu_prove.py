"""End-to-end parity: the GPU prover's proof bytes equal the CPU oracle's for the
same seeded polynomial (the reference's `run_whir_poseidon` flow,
crates/whir/src/test.rs:174-254), the checkpoint draws agree, and the oracle's
restatement of the reference verifier (whir.rs:361-522) accepts the GPU proof.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (k, folding, rate, initial_reduction, soundness, security, n_points)
SMALL = [
    (4, 1, 1, 1, 1, 32, 1),
    (7, 2, 1, 2, 1, 32, 2),
    (8, 3, 1, 1, 2, 32, 1),
    (10, 2, 1, 1, 1, 32, 3),
    (10, 3, 1, 3, 0, 32, 4),
    (12, 3, 1, 2, 2, 32, 2),
    (12, 2, 1, 2, 1, 32, 4),
    (13, 4, 1, 3, 1, 123, 1),
    (14, 4, 2, 3, 1, 123, 2),
    (16, 4, 1, 1, 1, 123, 1),
    (18, 4, 1, 3, 1, 123, 1),   # the reference's CPU-runnable configuration
    (17, 5, 1, 3, 1, 123, 1),   # folding 5 (whir_bench shape, smaller k)
    (18, 3, 1, 3, 1, 123, 2),   # a later round of 2^12 coefficients (above the one-workgroup rounds)
    (17, 3, 2, 2, 1, 123, 1),   # a later round of 2^11 coefficients, rate 1/4
]


def _prove_both(oracle, cfg, seed=1):
    import whir_amd

    k, f, rate, ir, s, sec, npts = cfg
    poly = oracle.rand_poly(seed, k)
    print(f"prove k={k} f={f} rate={rate} ir={ir} soundness={s} sec={sec} points={npts}", flush=True)
    w = whir_amd.Whir(k, f, rate, ir, s, sec)
    gproof, gcp, groot = w.prove(poly, npts)
    oproof, ocp, oroot = oracle.whir_prove(k, f, rate, ir, s, sec, npts, poly)
    return gproof, gcp, groot, oproof, ocp, oroot


def _first_diff(a, b):
    n = min(len(a), len(b))
    for i in range(n):
        if a[i] != b[i]:
            return i
    return n if len(a) != len(b) else -1


@pytest.mark.parametrize("cfg", SMALL, ids=lambda c: "k{}f{}r{}ir{}s{}sec{}n{}".format(*c))
def test_proof_bytes_match_oracle(gpu_ctx, oracle, cfg):
    gproof, gcp, groot, oproof, ocp, oroot = _prove_both(oracle, cfg)
    assert np.array_equal(groot, oroot), "commitment root differs"
    d = _first_diff(gproof, oproof)
    assert d == -1, f"proof differs at byte {d} (len gpu {len(gproof)} oracle {len(oproof)})"
    assert gcp == ocp
    k, f, rate, ir, s, sec, npts = cfg
    rc, vcp = oracle.whir_verify(k, f, rate, ir, s, sec, npts, gproof)
    assert rc == 0 and vcp == gcp


def test_same_result_alone_and_repeated(gpu_ctx, oracle):
    """A prover object reused for several polynomials gives each one the bytes
    it gets alone."""
    import whir_amd

    cfg = (12, 3, 1, 2, 1, 32, 2)
    k, f, rate, ir, s, sec, npts = cfg
    w = whir_amd.Whir(k, f, rate, ir, s, sec)
    outs = []
    for seed in (3, 4, 3):
        poly = oracle.rand_poly(seed, k)
        outs.append(w.prove(poly, npts)[0])
    assert outs[0] == outs[2] and outs[0] != outs[1]
    want = oracle.whir_prove(k, f, rate, ir, s, sec, npts, oracle.rand_poly(4, k))[0]
    assert outs[1] == want


def test_large_job_eight_claims(gpu_ctx, oracle):
    """The large single-job configuration: 2^26 coefficients, rate 1/4 (a 1 GiB
    base codeword), folding 4, eight evaluation claims — bytes equal to the
    oracle's and accepted by the oracle verifier."""
    cfg = (26, 4, 2, 3, 1, 123, 8)
    gproof, gcp, groot, oproof, ocp, oroot = _prove_both(oracle, cfg, seed=3)
    assert np.array_equal(groot, oroot)
    d = _first_diff(gproof, oproof)
    assert d == -1, f"proof differs at byte {d}"
    rc, vcp = oracle.whir_verify(*cfg, gproof)
    assert rc == 0 and vcp == gcp
