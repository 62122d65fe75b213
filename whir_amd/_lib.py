"""Loads libwhir_amd.so (built in-tree by whir_amd/build.py) and declares the
C ABI of include/whir_gpu.h for ctypes."""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
library_path = os.path.join(_HERE, "libwhir_amd.so")
# the same C ABI built for each field the reference instantiates
library_paths = {"koalabear": library_path, "babybear": os.path.join(_HERE, "libwhir_amd_bb.so")}

STATUS = {
    0: "WHIR_OK",
    1: "WHIR_ETRANSCRIPT",
    2: "WHIR_EVERIFY",
    3: "WHIR_EIO",
    4: "WHIR_EINVAL",
    5: "WHIR_EOOM",
    6: "WHIR_EDEVICE",
    7: "WHIR_EBUSY",
    8: "WHIR_ECHECK",
}


class WhirError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where}: {STATUS.get(code, code)}")
        self.code = code


class whir_params(ctypes.Structure):
    _fields_ = [
        ("k", ctypes.c_int32),
        ("folding", ctypes.c_int32),
        ("rate", ctypes.c_int32),
        ("initial_reduction", ctypes.c_int32),
        ("soundness", ctypes.c_int32),
        ("security_level", ctypes.c_int32),
        ("n_points", ctypes.c_int32),
    ]


class whir_stack(ctypes.Structure):
    _fields_ = [
        ("n_polys", ctypes.c_int32),
        ("k_polys", ctypes.POINTER(ctypes.c_int32)),
        ("n_claims", ctypes.c_int32),
        ("claim_poly", ctypes.POINTER(ctypes.c_int32)),
        ("claim_fresh", ctypes.POINTER(ctypes.c_uint8)),
    ]


class whir_challenger(ctypes.Structure):
    _fields_ = [
        ("sponge_state", ctypes.c_uint32 * 16),
        ("input_buffer", ctypes.c_uint32 * 8),
        ("output_buffer", ctypes.c_uint32 * 8),
        ("input_len", ctypes.c_uint32),
        ("output_len", ctypes.c_uint32),
    ]


# (name, restype, argtypes) for every entry point declared in include/whir_gpu.h
_vp = ctypes.c_void_p
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u8pp = ctypes.POINTER(_u8p)
SIGNATURES = [
    ("whir_ctx_create", ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("whir_ctx_destroy", ctypes.c_int, [_vp]),
    ("whir_ctx_set_ntt_tuning", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    ("whir_ctx_set_ntt_shape", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    ("whir_ctx_set_ntt_passes", ctypes.c_int, [_vp, ctypes.c_int]),
    ("whir_ctx_set_merkle_tuning", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    ("whir_ctx_set_autotune", ctypes.c_int, [_vp, ctypes.c_int]),
    ("whir_ctx_tuning", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    ("whir_dft_batch", ctypes.c_int, [_vp, _vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp]),
    ("whir_merkle_commit", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint32, _vp, _u32p, _vp]),
    ("whir_poseidon2_permute", ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_uint64, _vp]),
    ("whir_merkle_open", ctypes.c_int,
     [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, _vp]),
    ("whir_eq_table", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, _vp, _vp]),
    ("whir_eval_ext", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp, _vp]),
    ("whir_sort_dedup", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp, _vp]),
    ("whir_compress_hi", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, _vp]),
    ("whir_compress_lo", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, _vp]),
    ("whir_compress_hi_noalloc", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, _vp]),
    ("whir_sumcheck_coeffs", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint32, _vp, _vp]),
    ("whir_fix_lo_var", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, _vp, _vp]),
    ("whir_combine_eqs", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _u32p, _vp]),
    ("whir_combine_pows", ctypes.c_int,
     [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _u32p, ctypes.c_uint64, _vp]),
    ("whir_combine_pows_scaled", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, _vp]),
    ("whir_prover_create", ctypes.c_int, [_vp, ctypes.POINTER(whir_params), ctypes.POINTER(_vp)]),
    ("whir_prover_destroy", ctypes.c_int, [_vp]),
    ("whir_prover_proof_capacity", ctypes.c_size_t, [_vp]),
    ("whir_prover_input", _vp, [_vp]),
    ("whir_prover_enqueue", ctypes.c_int, [_vp, _vp, _vp]),
    ("whir_prover_result", ctypes.c_int,
     [_vp, _u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), _u32p, _u32p]),
    ("whir_prover_result_view", ctypes.c_int, [_vp, _u8pp, ctypes.POINTER(ctypes.c_size_t), _u32p, _u32p]),
    ("whir_prover_commit", ctypes.c_int, [_vp, _vp, ctypes.POINTER(whir_challenger), _vp]),
    ("whir_prover_eval", ctypes.c_int, [_vp, _u32p, _vp]),
    ("whir_prover_open", ctypes.c_int, [_vp, _vp]),
    ("whir_prover_challenger", ctypes.c_int, [_vp, ctypes.POINTER(whir_challenger)]),
    ("whir_prover_create_stacked", ctypes.c_int,
     [_vp, ctypes.POINTER(whir_params), ctypes.POINTER(whir_stack), ctypes.POINTER(_vp)]),
    ("whir_prover_stack_offset", ctypes.c_int64, [_vp, ctypes.c_int32]),
    ("whir_prover_enqueue_stacked", ctypes.c_int, [_vp, ctypes.POINTER(_vp), _vp]),
    ("whir_prover_set_graph", ctypes.c_int, [_vp, ctypes.c_int]),
    ("whir_prover_set_checks", ctypes.c_int, [_vp, ctypes.c_int]),
    ("whir_prover_query", ctypes.c_int, [_vp]),
    ("whir_prover_device_bytes", ctypes.c_size_t, [_vp]),
    ("whir_prover_bytes_for", ctypes.c_size_t, [ctypes.POINTER(whir_params)]),
    ("whir_server_create", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("whir_server_destroy", ctypes.c_int, [_vp]),
    ("whir_server_submit", ctypes.c_int, [_vp, ctypes.POINTER(whir_params), _vp, ctypes.POINTER(ctypes.c_int64)]),
    ("whir_server_poll", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int)]),
    ("whir_server_take", ctypes.c_int, [_vp, ctypes.c_int64, _u8p, ctypes.c_size_t,
                                         ctypes.POINTER(ctypes.c_size_t), _u32p]),
    ("whir_server_stats", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_uint64)]),
    ("whir_prover_set_timers", ctypes.c_int, [_vp, ctypes.c_int]),
    ("whir_prover_timer", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    ("whir_prove_host", ctypes.c_int,
     [_vp, ctypes.POINTER(whir_params), _u32p, _u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), _u32p, _u32p]),
    ("whir_batch_create", ctypes.c_int, [_vp, ctypes.POINTER(whir_params), ctypes.c_int, ctypes.POINTER(_vp)]),
    ("whir_batch_destroy", ctypes.c_int, [_vp]),
    ("whir_batch_input", _vp, [_vp, ctypes.c_int]),
    ("whir_batch_set_graph", ctypes.c_int, [_vp, ctypes.c_int]),
    ("whir_batch_set_checks", ctypes.c_int, [_vp, ctypes.c_int]),
    ("whir_batch_enqueue", ctypes.c_int, [_vp, ctypes.POINTER(_vp), _vp]),
    ("whir_batch_query", ctypes.c_int, [_vp]),
    ("whir_batch_result", ctypes.c_int,
     [_vp, ctypes.c_int, _u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), _u32p, _u32p]),
    ("whir_batch_result_view", ctypes.c_int, [_vp, ctypes.c_int, _u8pp, ctypes.POINTER(ctypes.c_size_t), _u32p, _u32p]),
    ("whir_batch_proof_capacity", ctypes.c_size_t, [_vp]),
    ("whir_batch_device_bytes", ctypes.c_size_t, [_vp]),
    ("whir_multi_create", ctypes.c_int,
     [_vp, ctypes.POINTER(whir_params), ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    ("whir_multi_destroy", ctypes.c_int, [_vp]),
    ("whir_multi_input", _vp, [_vp, ctypes.c_int]),
    ("whir_multi_set_graph", ctypes.c_int, [_vp, ctypes.c_int]),
    ("whir_multi_set_commit_queue", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    ("whir_multi_enqueue", ctypes.c_int, [_vp, ctypes.POINTER(_vp), _vp]),
    ("whir_multi_result", ctypes.c_int,
     [_vp, ctypes.c_int, _u8p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t), _u32p, _u32p]),
    ("whir_multi_result_view", ctypes.c_int, [_vp, ctypes.c_int, _u8pp, ctypes.POINTER(ctypes.c_size_t), _u32p, _u32p]),
    ("whir_multi_proof_capacity", ctypes.c_size_t, [_vp, ctypes.c_int]),
    ("whir_poly_rand", ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, _u32p]),
    ("whir_device_alloc", ctypes.c_int, [ctypes.c_size_t, ctypes.POINTER(_vp)]),
    ("whir_device_free", ctypes.c_int, [_vp]),
    ("whir_memcpy_h2d", ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    ("whir_memcpy_d2h", ctypes.c_int, [_vp, _vp, ctypes.c_size_t]),
    ("whir_memcpy_d2d", ctypes.c_int, [_vp, _vp, ctypes.c_size_t, _vp]),
    ("whir_device_synchronize", ctypes.c_int, []),
]

_lib = None
_libs = {}


def lib(field="koalabear"):
    """The loaded library of `field`. Raises if it has not been built: there is
    no fallback."""
    global _lib
    if field == "koalabear" and _lib is not None:
        return _lib
    if field not in _libs:
        path = library_path if field == "koalabear" else library_paths[field]
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing; build it with `python whir_amd/build.py` "
                "(or __graft_entry__.build()). The WHIR path has no CPU fallback.")
        L = ctypes.CDLL(path)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _libs[field] = L
    if field == "koalabear":
        _lib = _libs[field]
    return _libs[field]


def check(rc, where):
    if rc != 0:
        raise WhirError(rc, where)
