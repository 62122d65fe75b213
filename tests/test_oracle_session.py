"""The oracle inside a caller's transcript session (Whir::commit /
ProverStack::eval / Whir::open driven by a caller, crates/whir/src/whir.rs:127-180,
crates/whir/src/sumcheck/prover.rs:64-90): words written before the
commitment, claims at points the caller gives, and the challenger state open
hands back. These are the checks the GPU session entry points are compared
against.

* no prefix and only drawn points is the test driver's proof (test.rs:218-235);
* the prefix's bytes start the proof (bincode: 4 LE bytes per word,
  transcript/src/field.rs:6-25);
* a given claim's value in the proof is the polynomial's value at that point
  (naive multilinear evaluation, which the mathematics fixes);
* the session verifier accepts, and rejects another prefix or another point;
* sampling from the handed-back state (DuplexPy, a restatement of the
  challenger) returns the checkpoint draw.
"""
import numpy as np
import pytest

from kbpy import DuplexPy, e_from_m, e_to_m, emul

P = 0x7F000001


def _rand_words(rng, n):
    return rng.integers(0, P, n, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("cfg", [(10, 2, 1, 1, 1, 32, 2), (12, 4, 1, 3, 1, 123, 1)])
def test_plain_session_is_the_test_driver(oracle, cfg):
    k, f, rate, ir, s, sec, n = cfg
    poly = oracle.rand_poly(1, k)
    want = oracle.whir_prove(k, f, rate, ir, s, sec, n, poly)
    got = oracle.whir_prove_session(k, f, rate, ir, s, sec, [poly], [(0, 1)] * n)
    assert got[0] == want[0] and got[1] == want[1] and np.array_equal(got[2], want[2])


@pytest.mark.parametrize("n_prefix", [0, 3, 8, 11])
def test_prefix_and_given_points(oracle, n_prefix):
    k, f, rate, ir, s, sec = 12, 3, 1, 2, 1, 32
    rng = np.random.default_rng(n_prefix)
    poly = oracle.rand_poly(5, k)
    prefix = _rand_words(rng, n_prefix)
    z = _rand_words(rng, 4 * k).reshape(k, 4)
    claims = [(0, 1), (0, oracle.GIVEN), (0, 1)]
    proof, cp, root, st = oracle.whir_prove_session(k, f, rate, ir, s, sec, [poly], claims, prefix, [z])
    # the prefix's bytes, then the root
    assert proof[: 4 * n_prefix] == b"".join(int(x).to_bytes(4, "little") for x in prefix)
    assert np.frombuffer(proof[4 * n_prefix: 4 * n_prefix + 32], np.uint32).tolist() == [int(x) for x in root]
    # claim 1's value (after the root and claim 0's 16 bytes) is poly(z)
    off = 4 * n_prefix + 32 + 16
    assert np.frombuffer(proof[off: off + 16], np.uint32).tolist() == [int(x) for x in oracle.eval_base(poly, z)]
    rc, vcp = oracle.whir_verify_session(k, f, rate, ir, s, sec, [k], claims, proof, prefix, [z])
    assert rc == 0 and vcp == cp
    z2 = z.copy()
    z2[0, 0] = (int(z2[0, 0]) + 1) % P
    assert oracle.whir_verify_session(k, f, rate, ir, s, sec, [k], claims, proof, prefix, [z2])[0] != 0
    if n_prefix:
        p2 = prefix.copy()
        p2[-1] = (int(p2[-1]) + 1) % P
        assert oracle.whir_verify_session(k, f, rate, ir, s, sec, [k], claims, proof, p2, [z])[0] != 0
    # the handed-back challenger continues the transcript: its next sample is the checkpoint
    assert int(DuplexPy(oracle, st).sample()) == cp


def test_given_point_equal_to_the_drawn_one(oracle):
    """A claim given the very point the transcript would draw gives the drawn
    proof's bytes (only the draw itself differs: the given claim does not
    sample, so the transcript before it is replayed to find the point)."""
    k, f, rate, ir, s, sec = 10, 2, 1, 1, 1, 32
    poly = oracle.rand_poly(2, k)
    prefix = np.array([7, 8, 9], np.uint32)
    drawn = oracle.whir_prove_session(k, f, rate, ir, s, sec, [poly], [(0, 1)], prefix)
    # replay: prefix observed, root observed, then draw 4 words -> z, point = expand(k, z)
    d = DuplexPy(oracle)
    for x in prefix:
        d.observe(int(x))
    for x in drawn[2]:
        d.observe(int(x))
    z = [int(d.sample()) for _ in range(4)]
    pt = [e_from_m(z)]
    for _ in range(k - 1):
        pt.append(emul(pt[-1], pt[-1]))
    pt = np.array([e_to_m(e) for e in pt], np.uint32)
    given = oracle.whir_prove_session(k, f, rate, ir, s, sec, [poly], [(0, oracle.GIVEN)], prefix, [pt])
    # identical claim value; the later transcript differs because no draw happened
    off = 12 + 32
    assert drawn[0][:off + 16] == given[0][:off + 16]
    rc, _ = oracle.whir_verify_session(k, f, rate, ir, s, sec, [k], [(0, oracle.GIVEN)], given[0], prefix, [pt])
    assert rc == 0
