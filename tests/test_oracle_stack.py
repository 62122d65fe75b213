"""Multi-polynomial stacking in the oracle (ProverStack / ProverLayout /
VerifierLayout, crates/whir/src/sumcheck/{prover,verifier,mod}.rs) inside a
WHIR proof (whir.rs:168-359 with a stacked layout).

Pins:
* layout and alpha order: known answers worked out by hand from
  ProverStack::layout (prover.rs:92-124) and the derived Ord of Selector
  (mod.rs:30-37), chosen so that numeric selector order, registration order
  and size order all give different answers; the test_zkvm shape stacks to
  k = 25 (sumcheck/test.rs:150);
* write_virtual_eval's per-poly value and SVO accumulators equal one claim on
  the stacked vector (the reference's own debug check, prover.rs:258-266), and
  equal eval_base of the stacked vector at the point;
* round trips: the restated VerifierLayout / SumcheckVerifier accepts, with the
  same checkpoint draw, and rejects flipped bytes; a bad claim pattern raises;
* a stack of one poly gives the plain proof (golden-pinned path)."""
import hashlib
import json
import os

import numpy as np
import pytest

from tests.stack_specs import MIXED, rand_polys, stacked_k, zkvm_shape

HERE = os.path.dirname(os.path.abspath(__file__))
P = 0x7F000001


def test_layout_known_answer(oracle):
    # ids 0,1,2: k=20; ids 3,4: k=8; total 3*2^20 + 512 -> k = 22.
    # Stable sort by k then walk backwards: 2 @0, 1 @2^20, 0 @2^21, 4 @3*2^20, 3 @3*2^20+256.
    # Selectors (bit i of index, LSB first): id2 [0,0]  id1 [1,0]  id0 [0,1]
    #   id4 = 12288 over 14 bits: [0]*12 + [1,1]   id3 = 12289: [1] + [0]*11 + [1,1]
    # Vec<bool> order: [0,0] < [0,0,0..] (prefix) < [0,1] < [1,0] < [1,0,0..]
    k, offs, order = oracle.stack_layout([20, 20, 20, 8, 8], 4)
    assert k == 22
    assert offs == [2 << 20, 1 << 20, 0, (3 << 20) + 256, 3 << 20]
    assert order == [2, 4, 0, 1, 3]


def test_claim_ranks_known_answer(oracle):
    # layout order [2, 4, 0, 1, 3]; claims (poly, fresh):
    claims = [(0, True), (1, False), (2, False), (3, True), (4, False), (2, True)]
    # poly 2: claims 2, 5 -> ranks 0, 1; poly 4: claim 4 -> 2; poly 0: claim 0 -> 3;
    # poly 1: claim 1 -> 4; poly 3: claim 3 -> 5
    assert oracle.stack_claim_ranks([20, 20, 20, 8, 8], 4, claims) == [3, 4, 0, 5, 2, 1]


def test_zkvm_layout(oracle):
    k_polys, claims = zkvm_shape()
    k, offs, order = oracle.stack_layout(k_polys, 7)
    assert k == 25 == stacked_k(k_polys)  # sumcheck/test.rs:150
    # blocks are aligned, disjoint and inside 2^k
    spans = sorted((o, o + (1 << kp)) for o, kp in zip(offs, k_polys))
    assert all(o % (1 << kp) == 0 for o, kp in zip(offs, k_polys))
    assert all(a[1] <= b[0] for a, b in zip(spans, spans[1:])) and spans[-1][1] <= 1 << k
    assert sorted(order) == list(range(len(k_polys)))
    ranks = oracle.stack_claim_ranks(k_polys, 7, claims)
    assert sorted(ranks) == list(range(len(claims)))


@pytest.mark.parametrize("idx", range(len(MIXED)))
def test_virtual_claim_equals_stacked_claim(oracle, idx):
    k_polys, _ = MIXED[idx]
    polys = rand_polys(oracle, k_polys, seed=50 + idx)
    for s in range(3):
        x = oracle.rand_ext(900 + 10 * idx + s, 1)[0]
        for k_svo in (1, 3, 4):
            if k_svo <= min(k_polys):
                assert oracle.stack_virtual_check(polys, k_svo, x) == 1


def _prove_verify(oracle, k_polys, claims, f, rate, ir, s, sec, seed=1000):
    polys = rand_polys(oracle, k_polys, seed)
    k = stacked_k(k_polys)
    proof, cp, root = oracle.whir_prove_stacked(k, f, rate, ir, s, sec, polys, claims)
    rc, vcp = oracle.whir_verify_stacked(k, f, rate, ir, s, sec, k_polys, claims, proof)
    assert rc == 0 and vcp == cp, (k_polys, f, s)
    return proof, cp, root


@pytest.mark.parametrize("idx", range(len(MIXED)))
@pytest.mark.parametrize("f,s", [(1, 1), (2, 2), (3, 0), (4, 1)])
def test_stacked_round_trip(oracle, idx, f, s):
    k_polys, claims = MIXED[idx]
    if f > min(k_polys):
        pytest.skip("a poly smaller than the folding factor")
    _prove_verify(oracle, k_polys, claims, f, 1, min(f, 2), s, 32)


def test_stacked_tampered(oracle):
    k_polys, claims = MIXED[2]
    k = stacked_k(k_polys)
    proof, _, _ = _prove_verify(oracle, k_polys, claims, 3, 1, 2, 1, 32)
    for pos in (0, 33, 40 + 16 * len(claims), len(proof) // 2, len(proof) - 20):
        b = bytearray(proof)
        b[pos] ^= 0x01
        rc, _ = oracle.whir_verify_stacked(k, 3, 1, 2, 1, 32, k_polys, claims, bytes(b))
        assert rc != 0, pos
    # the same bytes under another claim pattern or layout do not verify
    rc, _ = oracle.whir_verify_stacked(k, 3, 1, 2, 1, 32, k_polys, claims[::-1], proof)
    assert rc != 0
    rc, _ = oracle.whir_verify_stacked(k, 3, 1, 2, 1, 32, k_polys[::-1], claims, proof)
    assert rc != 0


def test_bad_claim_patterns(oracle):
    polys = rand_polys(oracle, [6, 5])
    with pytest.raises(ValueError):  # shared point across different sizes
        oracle.whir_prove_stacked(7, 2, 1, 1, 1, 32, polys, [(0, True), (1, False)])
    with pytest.raises(ValueError):  # first claim must draw its point
        oracle.whir_prove_stacked(7, 2, 1, 1, 1, 32, polys, [(0, False)])
    with pytest.raises(ValueError):  # poly smaller than the folding factor
        oracle.whir_prove_stacked(7, 6, 1, 1, 1, 32, polys, [(1, True)])
    with pytest.raises(ValueError):  # stacked k must be the WHIR k
        oracle.whir_prove_stacked(8, 2, 1, 1, 1, 32, polys, [(0, True)])


def test_stack_of_one_is_plain_proof(oracle):
    poly = oracle.rand_poly(1, 12)
    plain = oracle.whir_prove(12, 3, 1, 3, 0, 32, 3, poly)
    stacked = oracle.whir_prove_stacked(12, 3, 1, 3, 0, 32, [poly], [(0, True)] * 3)
    assert plain[0] == stacked[0] and plain[1] == stacked[1]


def test_zkvm_shape_round_trip(oracle):
    # the test_zkvm claim pattern at reduced sizes (A: k=14, B and C: k=6; stacked k=19),
    # WHIR folding 4, rate 1/2, JB-123
    k_polys, claims = zkvm_shape(14, 6)
    _prove_verify(oracle, k_polys, claims, 4, 1, 3, 1, 123)


def test_golden_stacked_fixtures(oracle):
    with open(os.path.join(HERE, "golden", "oracle_proofs.json")) as fh:
        data = json.load(fh)
    assert data["stacked"]
    for g in data["stacked"]:
        k_polys, claims = g["k_polys"], [tuple(c) for c in g["claims"]]
        polys = rand_polys(oracle, k_polys, g["poly_seed"])
        proof, cp, root = oracle.whir_prove_stacked(g["k"], g["folding"], g["rate"], g["initial_reduction"],
                                                    g["soundness"], g["security_level"], polys, claims)
        assert len(proof) == g["proof_len"]
        assert hashlib.sha256(proof).hexdigest() == g["proof_sha256"]
        assert cp == g["checkpoint"]
        assert [int(x) for x in root] == g["root"]
