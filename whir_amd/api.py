"""Python mirror of the reference's prover API over the HIP C ABI."""
import ctypes

import numpy as np

from ._lib import check, lib, whir_challenger, whir_params, whir_stack

UNIQUE_DECODING, JOHNSON_BOUND, CAPACITY_BOUND = 0, 1, 2
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


def _view_result(fn, name, *args):
    """(proof bytes, checkpoint, root) through a *_result_view entry point:
    one copy from the library's pinned host buffer into the bytes object."""
    ptr = _u8p()
    ln = ctypes.c_size_t()
    cp = ctypes.c_uint32()
    root = np.zeros(8, np.uint32)
    check(fn(*args, ctypes.byref(ptr), ctypes.byref(ln), ctypes.byref(cp), root.ctypes.data_as(_u32p)), name)
    return ctypes.string_at(ptr, ln.value), int(cp.value), root


class DeviceBuffer:
    """A plain hipMalloc allocation owned by Python."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().whir_device_alloc(max(int(nbytes), 4), ctypes.byref(p)), "whir_device_alloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(lib().whir_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes), "whir_memcpy_h2d")

    def download(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        check(lib().whir_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "whir_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().whir_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """whir_ctx: device, Poseidon2 constants (seed 1000 by default), twiddles.
    With autotune, workspaces created on it tune their NTT and Merkle shapes."""

    def __init__(self, device=0, poseidon_seed=1000, autotune=False, field="koalabear"):
        self.field = field
        self.L = lib(field)
        p = ctypes.c_void_p()
        check(self.L.whir_ctx_create(device, poseidon_seed, ctypes.byref(p)), "whir_ctx_create")
        self.ptr = p.value
        if autotune:
            check(self.L.whir_ctx_set_autotune(self.ptr, 1), "whir_ctx_set_autotune")

    def tuning(self):
        """The autotuner's choices so far (list of dicts)."""
        import json

        n = ctypes.c_size_t()
        check(self.L.whir_ctx_tuning(self.ptr, None, 0, ctypes.byref(n)), "whir_ctx_tuning")
        buf = ctypes.create_string_buffer(n.value + 1)
        check(self.L.whir_ctx_tuning(self.ptr, buf, n.value + 1, ctypes.byref(n)), "whir_ctx_tuning")
        return json.loads(buf.value.decode())

    def close(self):
        if self.ptr:
            self.L.whir_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context()
    return _default_ctx


def dft_batch(mat, log_h, ctx=None):
    """Column DFT of `mat` (n_in rows x width, Montgomery words) zero-padded to
    2^log_h rows (TwoAdicSubgroupDft::dft_batch, whir.rs:34)."""
    ctx = ctx or default_context()
    m = _u32(mat)
    n_in, width = m.shape
    din = DeviceBuffer(m.nbytes)
    din.upload(m)
    dout = DeviceBuffer((width << log_h) * 4)
    check(ctx.L.whir_dft_batch(ctx.ptr, din.ptr, n_in, log_h, width, dout.ptr, None), "whir_dft_batch")
    check(ctx.L.whir_device_synchronize(), "sync")
    return dout.download(np.uint32, width << log_h).reshape(1 << log_h, width)


def merkle_commit(rows, ctx=None):
    """Poseidon2 Merkle tree of `rows` (height x row_words); returns (root, layers)."""
    ctx = ctx or default_context()
    r = _u32(rows)
    h, w = r.shape
    log_h = h.bit_length() - 1
    assert 1 << log_h == h
    dr = DeviceBuffer(r.nbytes)
    dr.upload(r)
    dl = DeviceBuffer((2 * h - 1) * 32)
    root = np.zeros(8, np.uint32)
    check(ctx.L.whir_merkle_commit(ctx.ptr, dr.ptr, log_h, w, dl.ptr, root.ctypes.data_as(_u32p), None),
          "whir_merkle_commit")
    layers = dl.download(np.uint32, (2 * h - 1) * 8).reshape(2 * h - 1, 8)
    return root, layers


def poseidon2_permute(states, width, cooperative=False, wave=False, ctx=None):
    """Poseidon2-{16,24} of each row of `states` (n x width). Width 16 only:
    `cooperative` runs the 16-lane row form of the Merkle tree tops, `wave`
    the device transcript's one-wavefront form."""
    ctx = ctx or default_context()
    s = _u32(states).reshape(-1, width)
    d = DeviceBuffer(s.nbytes)
    d.upload(s)
    w = -64 if wave else -16 if cooperative else width
    check(ctx.L.whir_poseidon2_permute(ctx.ptr, w, d.ptr, s.shape[0], None), "whir_poseidon2_permute")
    check(ctx.L.whir_device_synchronize(), "sync")
    return d.download(np.uint32, s.size).reshape(s.shape)


def _dev(arr):
    a = _u32(arr)
    d = DeviceBuffer(a.nbytes)
    d.upload(a)
    return d, a


def _ext4(v):
    a = _u32(v).reshape(4)
    return a, a.ctypes.data_as(_u32p)


def _run(rc, name, out, count):
    check(rc, name)
    check(lib().whir_device_synchronize(), "sync")
    return out.download(np.uint32, count)


def merkle_open(layers, rows, log_h, indices, ctx=None):
    """MerkleTree::query for each index: row words then log_h sibling digests
    (rows: 2^log_h x row_words; layers: output of merkle_commit)."""
    ctx = ctx or default_context()
    r = _u32(rows)
    row_words = r.shape[1]
    dl, _ = _dev(layers)
    dr, _ = _dev(r)
    di, idx = _dev(np.asarray(indices, np.uint32))
    per = row_words + 8 * log_h
    do = DeviceBuffer(max(4, len(idx) * per * 4))
    out = _run(ctx.L.whir_merkle_open(ctx.ptr, dl.ptr, log_h, dr.ptr, row_words, di.ptr, len(idx), do.ptr, None),
               "whir_merkle_open", do, len(idx) * per)
    return out.reshape(len(idx), per)


def sort_dedup(values, ctx=None):
    """The sorted().dedup() of stir_indexes on the device (len(values) <= 1024)."""
    ctx = ctx or default_context()
    v = np.asarray(values, np.uint32).ravel()
    dv = DeviceBuffer(max(4, v.size * 4))
    if v.size:
        dv.upload(v)
    do = DeviceBuffer(max(4, v.size * 4))
    dc = DeviceBuffer(4)
    check(ctx.L.whir_sort_dedup(ctx.ptr, dv.ptr, v.size, do.ptr, dc.ptr, None), "whir_sort_dedup")
    check(lib().whir_device_synchronize(), "sync")
    count = int(dc.download(np.uint32, 1)[0])
    return do.download(np.uint32, max(count, 1))[:count]


def eq_table(z, scale, ctx=None):
    """Point::eq: 2^k x 4 words for a point z (k x 4 words)."""
    ctx = ctx or default_context()
    dz, zz = _dev(np.asarray(z).reshape(-1, 4))
    k = zz.shape[0]
    sc, scp = _ext4(scale)
    do = DeviceBuffer((4 << k) * 4)
    return _run(ctx.L.whir_eq_table(ctx.ptr, dz.ptr, k, scp, do.ptr, None), "whir_eq_table", do,
                4 << k).reshape(-1, 4)


def eval_ext(poly, z, ctx=None):
    """SplitEq::eval_ext of a 2^k extension polynomial (2^k x 4 words)."""
    ctx = ctx or default_context()
    dp, p = _dev(np.asarray(poly).reshape(-1, 4))
    dz, zz = _dev(np.asarray(z).reshape(-1, 4))
    do = DeviceBuffer(16)
    return _run(ctx.L.whir_eval_ext(ctx.ptr, dp.ptr, zz.shape[0], dz.ptr, do.ptr, None), "whir_eval_ext", do, 4)


def compress_hi(poly, k, z, ctx=None):
    """SplitEq::compress_hi: bind the top len(z) variables of a base polynomial."""
    ctx = ctx or default_context()
    dp, _ = _dev(poly)
    dz, zz = _dev(np.asarray(z).reshape(-1, 4))
    m = zz.shape[0]
    do = DeviceBuffer((4 << (k - m)) * 4)
    return _run(ctx.L.whir_compress_hi(ctx.ptr, dp.ptr, k, dz.ptr, m, do.ptr, None), "whir_compress_hi", do,
                4 << (k - m)).reshape(-1, 4)


def compress_lo(poly, k, z, ctx=None):
    """SplitEq::compress_lo: bind the low len(z) variables of a base polynomial."""
    ctx = ctx or default_context()
    dp, _ = _dev(poly)
    dz, zz = _dev(np.asarray(z).reshape(-1, 4))
    m = zz.shape[0]
    do = DeviceBuffer((4 << (k - m)) * 4)
    return _run(ctx.L.whir_compress_lo(ctx.ptr, dp.ptr, k, dz.ptr, m, do.ptr, None), "whir_compress_lo", do,
                4 << (k - m)).reshape(-1, 4)


def sumcheck_coeffs(f, w, ctx=None):
    """Sumcheck round coefficients (c0, c2) of two 2^k extension arrays."""
    ctx = ctx or default_context()
    df, ff = _dev(np.asarray(f).reshape(-1, 4))
    dw, _ = _dev(np.asarray(w).reshape(-1, 4))
    k = int(ff.shape[0]).bit_length() - 1
    do = DeviceBuffer(32)
    return _run(ctx.L.whir_sumcheck_coeffs(ctx.ptr, df.ptr, dw.ptr, k, do.ptr, None), "whir_sumcheck_coeffs", do,
                8).reshape(2, 4)


def fix_lo_var(poly, r, ctx=None):
    """Poly::fix_lo_var_mut of a 2^k extension array."""
    ctx = ctx or default_context()
    dp, p = _dev(np.asarray(poly).reshape(-1, 4))
    k = int(p.shape[0]).bit_length() - 1
    rr, rp = _ext4(r)
    do = DeviceBuffer((4 << (k - 1)) * 4)
    return _run(ctx.L.whir_fix_lo_var(ctx.ptr, dp.ptr, k, rp, do.ptr, None), "whir_fix_lo_var", do,
                4 << (k - 1)).reshape(-1, 4)


def combine_eqs(w, points, alpha, ctx=None):
    """w += sum_j alpha^j eq(points_j); w is 2^k x 4 words, points n x k x 4."""
    ctx = ctx or default_context()
    dw, ww = _dev(np.asarray(w).reshape(-1, 4))
    k = int(ww.shape[0]).bit_length() - 1
    pts = np.asarray(points, np.uint32).reshape(-1, k, 4)
    dpt, _ = _dev(pts)
    aa, ap = _ext4(alpha)
    return _run(ctx.L.whir_combine_eqs(ctx.ptr, dw.ptr, k, dpt.ptr, pts.shape[0], ap, None), "whir_combine_eqs",
                dw, ww.size).reshape(-1, 4)


def combine_pows(w, variables, alpha, shift, ctx=None):
    """w[i] += sum_j alpha^(shift+j) vars_j^i; w is 2^k x 4 words."""
    ctx = ctx or default_context()
    dw, ww = _dev(np.asarray(w).reshape(-1, 4))
    k = int(ww.shape[0]).bit_length() - 1
    dv, vv = _dev(np.asarray(variables, np.uint32))
    aa, ap = _ext4(alpha)
    return _run(ctx.L.whir_combine_pows(ctx.ptr, dw.ptr, k, dv.ptr, vv.size, ap, shift, None), "whir_combine_pows",
                dw, ww.size).reshape(-1, 4)


def combine_pows_scaled(w, variables, scales, ctx=None):
    """w[i] += sum_j scales_j vars_j^i; w is 2^k x 4 words, scales n x 4."""
    ctx = ctx or default_context()
    dw, ww = _dev(np.asarray(w).reshape(-1, 4))
    k = int(ww.shape[0]).bit_length() - 1
    dv, vv = _dev(np.asarray(variables, np.uint32))
    ds, _ = _dev(np.asarray(scales).reshape(-1, 4))
    return _run(ctx.L.whir_combine_pows_scaled(ctx.ptr, dw.ptr, k, dv.ptr, vv.size, ds.ptr, None),
                "whir_combine_pows_scaled", dw, ww.size).reshape(-1, 4)


class Prover:
    """Device workspace for one proof shape (whir_prover)."""

    def __init__(self, params, ctx=None):
        self.ctx = ctx or default_context()
        self.params = params
        p = ctypes.c_void_p()
        check(self.ctx.L.whir_prover_create(self.ctx.ptr, ctypes.byref(params), ctypes.byref(p)), "whir_prover_create")
        self.ptr = p.value
        self.capacity = self.ctx.L.whir_prover_proof_capacity(self.ptr)
        self.input_ptr = self.ctx.L.whir_prover_input(self.ptr)

    def upload(self, poly):
        a = _u32(poly)
        assert a.size == 1 << self.params.k
        check(self.ctx.L.whir_memcpy_h2d(self.input_ptr, a.ctypes.data, a.nbytes), "whir_memcpy_h2d")

    def enqueue(self, d_poly=None, stream=None):
        check(self.ctx.L.whir_prover_enqueue(self.ptr, d_poly or self.input_ptr, stream), "whir_prover_enqueue")

    def result(self):
        return _view_result(self.ctx.L.whir_prover_result_view, "whir_prover_result_view", self.ptr)

    # Whir::commit / ProverStack::eval / Whir::open with the caller's transcript
    # (whir.rs:127-180, prover.rs:64-90)
    def commit(self, d_poly=None, challenger=None, stream=None):
        """Commits d_poly (default: the uploaded input); `challenger` is a dict of
        whir_challenger fields (None: a fresh transcript)."""
        st = None
        if challenger is not None:
            st = whir_challenger()
            for i, v in enumerate(challenger["sponge_state"]):
                st.sponge_state[i] = int(v)
            for i in range(int(challenger["input_len"])):
                st.input_buffer[i] = int(challenger["input_buffer"][i])
            for i in range(int(challenger["output_len"])):
                st.output_buffer[i] = int(challenger["output_buffer"][i])
            st.input_len = int(challenger["input_len"])
            st.output_len = int(challenger["output_len"])
        check(self.ctx.L.whir_prover_commit(self.ptr, d_poly or self.input_ptr, ctypes.byref(st) if st else None, stream),
              "whir_prover_commit")

    def eval(self, point=None, stream=None):
        """The next evaluation claim: at `point` (k_poly x 4 Montgomery words) or,
        with None, at expand(k_poly, draw())."""
        if point is None:
            check(self.ctx.L.whir_prover_eval(self.ptr, None, stream), "whir_prover_eval")
        else:
            pt = _u32(point).reshape(-1)
            check(self.ctx.L.whir_prover_eval(self.ptr, pt.ctypes.data_as(_u32p), stream), "whir_prover_eval")

    def open(self, stream=None):
        check(self.ctx.L.whir_prover_open(self.ptr, stream), "whir_prover_open")

    def challenger(self):
        """The transcript state open handed back, as whir_challenger fields."""
        st = whir_challenger()
        check(self.ctx.L.whir_prover_challenger(self.ptr, ctypes.byref(st)), "whir_prover_challenger")
        return {"sponge_state": np.array(st.sponge_state[:], np.uint32),
                "input_buffer": np.array(st.input_buffer[:], np.uint32),
                "output_buffer": np.array(st.output_buffer[:], np.uint32),
                "input_len": int(st.input_len), "output_len": int(st.output_len)}

    TIMER_LEAVES_BASE, TIMER_DFT_BASE = 0, 1

    def set_graph(self, enable=True):
        check(self.ctx.L.whir_prover_set_graph(self.ptr, int(enable)), "whir_prover_set_graph")

    def set_checks(self, enable=True):
        """Device self-checks of each proof; result() then raises WhirError
        (WHIR_ECHECK) when one failed."""
        check(self.ctx.L.whir_prover_set_checks(self.ptr, int(enable)), "whir_prover_set_checks")

    def done(self):
        rc = self.ctx.L.whir_prover_query(self.ptr)
        if rc not in (0, 7):
            check(rc, "whir_prover_query")
        return rc == 0

    def set_timers(self, enable=True):
        check(self.ctx.L.whir_prover_set_timers(self.ptr, int(enable)), "whir_prover_set_timers")

    def timer(self, which):
        ms, n, work = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
        check(self.ctx.L.whir_prover_timer(self.ptr, which, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(work)),
              "whir_prover_timer")
        return ms.value, n.value, work.value

    def close(self):
        if self.ptr:
            self.ctx.L.whir_prover_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class StackedProver(Prover):
    """Device workspace for one stacked proof shape: `k_polys` registered polys
    (ProverStack::register, crates/whir/src/sumcheck/prover.rs:59-62) and
    evaluation claims `claims` = [(poly id, fresh point), ...] in transcript
    order (ProverStack::eval, prover.rs:64-90). ``params.k`` must be the
    stacked k; the proof is Whir::open over the stack's layout (whir.rs:168-359)
    after the commitment of the stacked vector."""

    def __init__(self, params, k_polys, claims, ctx=None):
        self.ctx = ctx or default_context()
        self.params = params
        self.k_polys = [int(k) for k in k_polys]
        kp = np.asarray(self.k_polys, np.int32)
        cp = np.asarray([int(c[0]) for c in claims], np.int32)
        cf = np.asarray([1 if c[1] else 0 for c in claims], np.uint8)
        st = whir_stack(len(kp), kp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(cp),
                        cp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), cf.ctypes.data_as(_u8p))
        p = ctypes.c_void_p()
        check(self.ctx.L.whir_prover_create_stacked(self.ctx.ptr, ctypes.byref(params), ctypes.byref(st),
                                               ctypes.byref(p)), "whir_prover_create_stacked")
        self.ptr = p.value
        self.capacity = self.ctx.L.whir_prover_proof_capacity(self.ptr)
        self.input_ptr = self.ctx.L.whir_prover_input(self.ptr)

    def offset(self, poly_id):
        return self.ctx.L.whir_prover_stack_offset(self.ptr, poly_id)

    def upload(self, polys):
        """Copies the polys to the device and lays them out with the zero tail."""
        bufs = []
        for kp, poly in zip(self.k_polys, polys):
            a = _u32(poly)
            assert a.size == 1 << kp
            b = DeviceBuffer(a.nbytes)
            b.upload(a)
            bufs.append(b)
        ptrs = (ctypes.c_void_p * len(bufs))(*[b.ptr for b in bufs])
        self._staged = (bufs, ptrs)

    def enqueue(self, d_polys=None, stream=None):
        """Lays out device polys (default: the staged upload) and enqueues the proof."""
        if d_polys is None:
            ptrs = self._staged[1]
        else:
            ptrs = (ctypes.c_void_p * len(d_polys))(*d_polys)
        check(self.ctx.L.whir_prover_enqueue_stacked(self.ptr, ptrs, stream), "whir_prover_enqueue_stacked")


class Batch:
    """whir_batch: `jobs` proofs of one shape as one lockstep kernel sequence
    (every launch covers all jobs; each job's bytes are its bytes alone)."""

    def __init__(self, params, jobs, ctx=None, graph=False):
        self.ctx = ctx or default_context()
        self.params = params
        self.jobs = int(jobs)
        p = ctypes.c_void_p()
        check(self.ctx.L.whir_batch_create(self.ctx.ptr, ctypes.byref(params), self.jobs, ctypes.byref(p)),
              "whir_batch_create")
        self.ptr = p.value
        self.capacity = self.ctx.L.whir_batch_proof_capacity(self.ptr)
        if graph:
            check(self.ctx.L.whir_batch_set_graph(self.ptr, 1), "whir_batch_set_graph")

    def set_checks(self, enable=True):
        check(self.ctx.L.whir_batch_set_checks(self.ptr, int(enable)), "whir_batch_set_checks")

    def input_ptr(self, job):
        return self.ctx.L.whir_batch_input(self.ptr, job)

    def upload(self, job, poly):
        a = _u32(poly)
        assert a.size == 1 << self.params.k
        check(self.ctx.L.whir_memcpy_h2d(self.input_ptr(job), a.ctypes.data, a.nbytes), "whir_memcpy_h2d")

    def enqueue(self, d_polys=None, stream=None):
        ptrs = None if d_polys is None else (ctypes.c_void_p * self.jobs)(*d_polys)
        check(self.ctx.L.whir_batch_enqueue(self.ptr, ptrs, stream), "whir_batch_enqueue")

    def done(self):
        rc = self.ctx.L.whir_batch_query(self.ptr)
        if rc not in (0, 7):
            check(rc, "whir_batch_query")
        return rc == 0

    def result(self, job):
        return _view_result(self.ctx.L.whir_batch_result_view, "whir_batch_result_view", self.ptr, job)

    def close(self):
        if self.ptr:
            self.ctx.L.whir_batch_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MultiBatch:
    """whir_multi: independent jobs of mixed shapes, grouped by shape into
    lockstep batches that run side by side on `streams` streams (graph: the
    whole set as one captured hipGraph)."""

    def __init__(self, params_list, streams=8, ctx=None, graph=1, commit_queue=0):
        """graph: 0 eager, 1 one captured graph for the whole set, 2 one per batch.
        commit_queue: number of job groups whose base LDEs, leaf hashes and
        wide tree layers come from persistent kernels over a device job table
        (whir_multi_set_commit_queue); 0 off."""
        self.ctx = ctx or default_context()
        self.n = len(params_list)
        arr = (whir_params * self.n)(*params_list)
        p = ctypes.c_void_p()
        check(self.ctx.L.whir_multi_create(self.ctx.ptr, arr, self.n, streams, ctypes.byref(p)), "whir_multi_create")
        self.ptr = p.value
        check(self.ctx.L.whir_multi_set_graph(self.ptr, int(graph)), "whir_multi_set_graph")
        self.queue_jobs = 0
        if commit_queue:
            self.set_commit_queue(int(commit_queue))

    def set_commit_queue(self, groups):
        """groups: 0 off, else the number of pipelined job groups. Returns the
        number of jobs the commit work queue covers."""
        n = ctypes.c_int(0)
        check(self.ctx.L.whir_multi_set_commit_queue(self.ptr, int(groups), ctypes.byref(n)),
              "whir_multi_set_commit_queue")
        self.queue_jobs = n.value
        return n.value

    def input_ptr(self, job):
        return self.ctx.L.whir_multi_input(self.ptr, job)

    def enqueue(self, d_polys=None, stream=None):
        ptrs = None if d_polys is None else (ctypes.c_void_p * self.n)(*d_polys)
        check(self.ctx.L.whir_multi_enqueue(self.ptr, ptrs, stream), "whir_multi_enqueue")

    def result(self, job):
        return _view_result(self.ctx.L.whir_multi_result_view, "whir_multi_result_view", self.ptr, job)

    def close(self):
        if self.ptr:
            self.ctx.L.whir_multi_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Whir:
    """Mirror of ``Whir::new`` (crates/whir/src/whir.rs:103-125) with the
    reference's Poseidon2 trees and transcript built in."""

    def __init__(self, k, folding, rate, initial_reduction, soundness=JOHNSON_BOUND, security_level=123,
                 ctx=None):
        self.k, self.folding, self.rate = k, folding, rate
        self.initial_reduction = initial_reduction
        self.soundness, self.security_level = soundness, security_level
        self.ctx = ctx or default_context()
        self._provers = {}

    def params(self, n_points):
        return whir_params(self.k, self.folding, self.rate, self.initial_reduction, self.soundness,
                           self.security_level, n_points)

    def prover(self, n_points):
        if n_points not in self._provers:
            self._provers[n_points] = Prover(self.params(n_points), self.ctx)
        return self._provers[n_points]

    def prove(self, poly, n_points=1):
        """commit + n_points evaluation claims + open. Returns (proof bytes,
        checkpoint draw, commitment root)."""
        pr = self.prover(n_points)
        pr.upload(poly)
        pr.enqueue()
        return pr.result()


def poly_rand(seed, k, field="koalabear"):
    """Poly::<F>::rand(&mut rng(seed), k) (crates/poly/src/poly.rs:47-52): the
    polynomial the reference's tests prove for `seed`, as Montgomery words of
    `field`."""
    out = np.empty(1 << k, np.uint32)
    check(lib(field).whir_poly_rand(seed, k, out.ctypes.data_as(_u32p)), "whir_poly_rand")
    return out


def params(k, folding=4, rate=1, initial_reduction=3, soundness=JOHNSON_BOUND, security_level=123, n_points=1):
    return whir_params(k, folding, rate, initial_reduction, soundness, security_level, n_points)


def prover_bytes_for(p):
    return lib().whir_prover_bytes_for(ctypes.byref(p))


class Server:
    """whir_server: continuous batching of independent proofs (see include/whir_gpu.h)."""

    def __init__(self, max_slots=6, hbm_budget_bytes=64 << 30, ctx=None):
        self.ctx = ctx or default_context()
        p = ctypes.c_void_p()
        check(self.ctx.L.whir_server_create(self.ctx.ptr, max_slots, hbm_budget_bytes, ctypes.byref(p)),
              "whir_server_create")
        self.ptr = p.value
        self._ids = (ctypes.c_int64 * 256)()

    def submit(self, p, d_poly):
        jid = ctypes.c_int64()
        check(self.ctx.L.whir_server_submit(self.ptr, ctypes.byref(p), d_poly, ctypes.byref(jid)), "whir_server_submit")
        return jid.value

    def poll(self):
        n = ctypes.c_int()
        check(self.ctx.L.whir_server_poll(self.ptr, self._ids, 256, ctypes.byref(n)), "whir_server_poll")
        return [self._ids[i] for i in range(n.value)]

    def take(self, job_id, capacity):
        buf = np.zeros(capacity, np.uint8)
        ln = ctypes.c_size_t()
        cp = np.zeros(1, np.uint32)
        check(self.ctx.L.whir_server_take(self.ptr, job_id, buf.ctypes.data_as(_u8p), buf.size, ctypes.byref(ln),
                                     cp.ctypes.data_as(_u32p)), "whir_server_take")
        return bytes(buf[: ln.value]), int(cp[0])

    def stats(self):
        r, q, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()
        check(self.ctx.L.whir_server_stats(self.ptr, ctypes.byref(r), ctypes.byref(q), ctypes.byref(b)),
              "whir_server_stats")
        return r.value, q.value, b.value

    def close(self):
        if self.ptr:
            self.ctx.L.whir_server_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
