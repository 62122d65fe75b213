"""whir_amd — MI355X (gfx950) WHIR prover hot path.

Host-side mirror of the reference's prover interface (kilic/whir):
  * ``Whir(k, folding, rate, initial_reduction, soundness, security_level)``
    mirrors ``Whir::new`` (crates/whir/src/whir.rs:103-125); ``Whir.prove``
    runs ``Whir::commit`` + ``ProverStack::eval`` + ``Whir::open`` as driven by
    ``run_whir_poseidon`` (crates/whir/src/test.rs:218-235) and returns the
    serialized PoseidonWriter proof.
  * ``dft_batch`` mirrors ``TwoAdicSubgroupDft::dft_batch`` (whir.rs:34,59).
  * ``merkle_commit`` mirrors ``MerkleTree::commit`` (crates/merkle/src/lib.rs:22-29).
Everything runs in HIP kernels through the C ABI of include/whir_gpu.h; there
is no CPU fallback: if libwhir_amd.so is missing or no GPU is present the calls
raise.
"""
from ._lib import WhirError, lib, library_path  # noqa: F401
from .api import (  # noqa: F401
    CAPACITY_BOUND,
    Batch,
    JOHNSON_BOUND,
    UNIQUE_DECODING,
    Context,
    MultiBatch,
    Prover,
    Server,
    StackedProver,
    Whir,
    params,
    poly_rand,
    prover_bytes_for,
    dft_batch,
    merkle_commit,
    poseidon2_permute,
)
