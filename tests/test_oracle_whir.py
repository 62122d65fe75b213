"""Protocol-level checks of the oracle:

* the proof schedule (OOD / STIR counts, rounds) from the restated params.rs
  against an independent Python restatement and the published plan of the
  measured configurations;
* the reference's end-to-end round trip test_whir_poseidon
  (crates/whir/src/test.rs:256-282): JB/CB/UD x k 4..12 x folding 1..3 x
  initial_reduction 1..folding x 1..4 points, 32-bit security: the restated
  verifier accepts and the checkpoint draws agree;
* failure handling: a flipped byte is rejected, a truncated proof and a
  non-canonical word give the Transcript error (transcript/src/lib.rs:352-353);
* golden fixtures tests/golden/oracle_proofs.json (oracle regressions)."""
import hashlib
import json
import math
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LOG2_10 = math.log2(10)


def py_n_stir(s, sec, r):
    le = {0: 0.0, 1: -(0.5 * r + LOG2_10 + 1.0), 2: -(r + LOG2_10 + 1.0)}[s]
    rate = 1.0 / (1 << r)
    delta = {0: 0.5 * (1 - rate), 1: 1 - math.sqrt(rate) - 2.0 ** le, 2: 1 - rate - 2.0 ** le}[s]
    return math.ceil(-sec / math.log2(1 - delta))


def py_n_ood(s, sec, d, r, fb=124):
    if s == 0:
        return 0
    le = {1: -(0.5 * r + LOG2_10 + 1.0), 2: -(r + LOG2_10 + 1.0)}[s]
    lsb = (r / 2.0 - (1.0 + le)) if s == 1 else ((d + r) - le)
    for n in range(1, 64):
        if (n * fb) + 1.0 - (2.0 * lsb + d * n) >= sec:
            return n
    raise AssertionError


def test_params_match_restatement(oracle):
    for s in (0, 1, 2):
        for sec in (32, 80, 100, 123, 128):
            for r in range(1, 14):
                assert oracle.n_stir_queries(s, sec, r) == py_n_stir(s, sec, r)
                for d in (4, 10, 20, 26):
                    assert oracle.n_ood_queries(s, sec, d, r) == py_n_ood(s, sec, d, r)


def test_schedule_of_measured_configs(oracle):
    # k=24, rate 1/2, folding 4, ir 3, JB-123: 2 OOD everywhere, STIR 287/133/51/32/23
    assert oracle.number_of_rounds(4, 24) == (4, 4)
    assert oracle.n_ood_queries(1, 123, 24, 1) == 2
    prev_rates = [1, 2, 5, 8, 11]  # rate + r*f - prev_reduction
    assert [oracle.n_stir_queries(1, 123, r) for r in prev_rates] == [287, 133, 51, 32, 23]
    # k=18 (the CPU-runnable configuration): 2 rounds? (18-6)/4 = 3 -> 2 rounds, final 6
    assert oracle.number_of_rounds(4, 18) == (2, 6)
    # k=26, rate 1/4
    assert oracle.number_of_rounds(4, 26) == (4, 6)
    assert [oracle.n_stir_queries(1, 123, r) for r in (2, 3, 6, 9, 12)] == [133, 87, 42, 28, 21]


GRID = [(s, k, f, ir, n) for s in (1, 2, 0) for k in range(4, 13) for f in range(1, 4) for ir in range(1, f + 1)
        for n in range(1, 5)]


def test_reference_round_trip_grid(oracle):
    for (s, k, f, ir, n) in GRID:
        poly = oracle.rand_poly(1, k)
        proof, cp, _ = oracle.whir_prove(k, f, 1, ir, s, 32, n, poly)
        rc, vcp = oracle.whir_verify(k, f, 1, ir, s, 32, n, proof)
        assert rc == 0 and vcp == cp, (s, k, f, ir, n)


@pytest.mark.parametrize("cfg", [(10, 2, 1, 1, 1, 32, 2), (14, 4, 1, 3, 1, 123, 1)])
def test_tampered_and_truncated_proofs(oracle, cfg):
    k, f, rate, ir, s, sec, n = cfg
    poly = oracle.rand_poly(1, k)
    proof, _, _ = oracle.whir_prove(*cfg, poly)
    for pos in (0, 40, len(proof) // 3, len(proof) // 2, len(proof) - 20):
        b = bytearray(proof)
        b[pos] ^= 0x01
        rc, _ = oracle.whir_verify(*cfg, bytes(b))
        assert rc != 0
    rc, _ = oracle.whir_verify(*cfg, proof[:-4])
    assert rc == 1  # Transcript error: read past the end
    b = bytearray(proof)
    b[0:4] = (0x7F000001).to_bytes(4, "little")  # not a canonical Montgomery word
    rc, _ = oracle.whir_verify(*cfg, bytes(b))
    assert rc == 1


def test_golden_fixtures(oracle):
    with open(os.path.join(HERE, "golden", "oracle_proofs.json")) as fh:
        data = json.load(fh)
    for g in data["proofs"]:
        poly = oracle.rand_poly(g["poly_seed"], g["k"])
        proof, cp, root = oracle.whir_prove(g["k"], g["folding"], g["rate"], g["initial_reduction"],
                                            g["soundness"], g["security_level"], g["n_points"], poly)
        assert len(proof) == g["proof_len"]
        assert hashlib.sha256(proof).hexdigest() == g["proof_sha256"]
        assert cp == g["checkpoint"]
        assert [int(x) for x in root] == g["root"]
