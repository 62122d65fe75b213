"""Writes tests/golden/oracle_proofs.json: SHA-256 digests, lengths and
checkpoints of CPU-oracle proofs for small seeded configurations, single-poly
("proofs") and stacked ("stacked", shapes from tests/stack_specs.py).

These fixtures come from this repository's oracle (the Rust reference and its
Plonky3 dependency cannot be built or imported here). They pin the oracle and
the GPU path against regressions and against each other; they do not pin the
oracle to the reference crate."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import pyoracle as O  # noqa: E402
from tests.stack_specs import MIXED, rand_polys, stacked_k, zkvm_shape  # noqa: E402

CONFIGS = [
    # k, folding, rate, ir, soundness, security, n_points, poly seed
    (8, 2, 1, 1, 1, 32, 1, 1),
    (10, 3, 1, 2, 2, 32, 2, 1),
    (12, 3, 1, 3, 0, 32, 3, 1),
    (14, 4, 1, 3, 1, 123, 1, 1),
    (16, 4, 2, 3, 1, 123, 4, 2),
    (18, 4, 1, 3, 1, 123, 1, 1),
]


def make():
    out = []
    for (k, f, rate, ir, s, sec, n, seed) in CONFIGS:
        poly = O.rand_poly(seed, k)
        proof, cp, root = O.whir_prove(k, f, rate, ir, s, sec, n, poly)
        out.append({
            "k": k, "folding": f, "rate": rate, "initial_reduction": ir, "soundness": s,
            "security_level": sec, "n_points": n, "poly_seed": seed,
            "proof_len": len(proof), "proof_sha256": hashlib.sha256(proof).hexdigest(),
            "checkpoint": cp, "root": [int(x) for x in root],
            "proof_head_hex": proof[:64].hex(),
        })
    return out


STACKED = [
    # (k_polys, claims), folding, rate, ir, soundness, security, poly seed
    (MIXED[1], 3, 1, 2, 1, 32, 7),
    (MIXED[2], 4, 1, 3, 1, 123, 8),
    (MIXED[3], 2, 2, 1, 2, 32, 9),
    (zkvm_shape(14, 6), 4, 1, 3, 1, 123, 1000),
]


def make_stacked():
    out = []
    for ((k_polys, claims), f, rate, ir, s, sec, seed) in STACKED:
        k = stacked_k(k_polys)
        polys = rand_polys(O, k_polys, seed)
        proof, cp, root = O.whir_prove_stacked(k, f, rate, ir, s, sec, polys, claims)
        out.append({
            "k": k, "k_polys": list(k_polys), "claims": [[int(a), bool(b)] for a, b in claims],
            "folding": f, "rate": rate, "initial_reduction": ir, "soundness": s, "security_level": sec,
            "poly_seed": seed, "proof_len": len(proof), "proof_sha256": hashlib.sha256(proof).hexdigest(),
            "checkpoint": cp, "root": [int(x) for x in root],
        })
    return out


if __name__ == "__main__":
    path = os.path.join(HERE, "oracle_proofs.json")
    with open(path, "w") as fh:
        json.dump({"source": "CPU oracle (oracle/), not the Rust reference", "proofs": make(),
                   "stacked": make_stacked()}, fh, indent=1)
    print("wrote", path)
