"""CPU checks of the drop-in boundary: the HIP library loads, exports every
entry point include/whir_gpu.h declares, and the Python mirror binds each one
(no device calls are made here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "whir_gpu.h")) as fh:
        text = fh.read()
    names = re.findall(r"^\s*(?:int|size_t|uint32_t\s*\*)\s+(whir_[a-z0-9_]+)\s*\(", text, re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("whir_ctx_create", "whir_dft_batch", "whir_merkle_commit", "whir_prover_enqueue",
                 "whir_prover_result", "whir_prove_host"):
        assert must in names


@pytest.mark.parametrize("field", ["koalabear", "babybear"])
def test_library_exports_every_declared_symbol(field):
    from whir_amd._lib import library_paths

    lib = ctypes.CDLL(library_paths[field])
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_python_binding_covers_header():
    from whir_amd._lib import SIGNATURES

    bound = {n for n, _, _ in SIGNATURES}
    assert set(declared_symbols()) <= bound


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    import whir_amd._lib as L

    monkeypatch.setattr(L, "library_path", str(tmp_path / "absent.so"))
    monkeypatch.setattr(L, "library_paths", {"babybear": str(tmp_path / "absent_bb.so")})
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "_libs", {})
    for field in ("koalabear", "babybear"):
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            L.lib(field)


def test_poly_rand_is_the_oracles_seeded_polynomial(oracle):
    """whir_poly_rand is host code (no device call): Poly::rand(&mut rng(seed), k)
    of the library against the oracle's restatement, for several seeds."""
    import numpy as np

    import whir_amd

    bb = oracle.for_field("babybear")
    for seed, k in ((1, 10), (2, 0), (1000, 13), (7, 16)):
        assert np.array_equal(whir_amd.poly_rand(seed, k), oracle.rand_poly(seed, k)), (seed, k)
        assert np.array_equal(whir_amd.poly_rand(seed, k, field="babybear"), bb.rand_poly(seed, k)), (seed, k)
    # the two fields' libraries are distinct builds: rejection sampling and
    # Montgomery words differ with the modulus
    assert not np.array_equal(whir_amd.poly_rand(1, 10), whir_amd.poly_rand(1, 10, field="babybear"))
