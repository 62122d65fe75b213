"""Writes tests/golden/oracle_proofs.json: SHA-256 digests, lengths and
checkpoints of CPU-oracle proofs for small seeded configurations.

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


if __name__ == "__main__":
    path = os.path.join(HERE, "oracle_proofs.json")
    with open(path, "w") as fh:
        json.dump({"source": "CPU oracle (oracle/), not the Rust reference", "proofs": make()}, fh, indent=1)
    print("wrote", path)
