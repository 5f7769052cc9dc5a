"""Python mirror of the vkhel C API (include/vkhel/vkhel.h).

Names and argument meaning follow the reference's public functions
(include/vkhel/vkhel.h:8-53): ``Context`` is ``vkhel_ctx``, ``NttTables``
is ``vkhel_ntt_tables``, ``Vector`` is ``vkhel_vector`` and its methods are
the ``vkhel_vector_*`` calls.  Errors raise ``VkhelError`` instead of
asserting.
"""
import ctypes

import numpy as np

from . import lib as _lib

NTT_AUTO, NTT_RADIX2, NTT_FUSED = 0, 1, 2
JOB_NTT_FWD, JOB_NTT_INV, JOB_POLYMUL, JOB_MUL, JOB_ADD = range(5)
JOB_CENTER, JOB_GTADD, JOB_GTSUB = 5, 6, 7  # RNS elementwise (add_rns)
JOB_INTT_CENTER = 8  # fused INTT + centered lift (add_rns, params [p])
JOB_RESCALE = 9  # drop the last limb, rounding (add_rns, no params)
JOB_RESCALE_NTT = 10  # the same on NTT-form limbs, result in NTT form
JOB_ADD_REF = 11  # elemfma(a, b, 1, q) per limb: the reference's vector add


def _L():
    return _lib.load()


def _u64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


class Context:
    """vkhel_ctx_create / vkhel_ctx_destroy (src/vkhel.c:4-13)."""

    def __init__(self):
        L = _L()
        self._h = L.vkhel_ctx_create()
        if not self._h:
            _lib.check()
            raise _lib.VkhelError(2, "vkhel_ctx_create failed")

    @property
    def handle(self):
        return self._h

    @property
    def device_name(self):
        return _L().vkhel_ext_ctx_device_name(self._h).decode()

    @property
    def stream(self):
        return _L().vkhel_ext_ctx_stream(self._h)

    def set_staging_pool(self, enabled):
        """vkhel_ext_set_staging_pool: reuse map staging buffers (default)."""
        _lib.check(_L().vkhel_ext_set_staging_pool(self._h, int(bool(enabled))))

    @property
    def staging_bytes(self):
        return int(_L().vkhel_ext_staging_bytes(self._h))

    def synchronize(self):
        _lib.check(_L().vkhel_ext_ctx_synchronize(self._h))

    def set_ntt_algo(self, algo):
        _lib.check(_L().vkhel_ext_set_ntt_algo(self._h, algo))

    def autotune(self, logn_lo=12, logn_hi=17, batch=0):
        """Times each compiled tile size per transform size; keeps the best."""
        _lib.check(_L().vkhel_ext_autotune(self._h, logn_lo, logn_hi, batch))

    def tile_bits(self, logn):
        tb = _L().vkhel_ext_tile_bits(self._h, logn)
        _lib.check()
        return tb

    def set_tile_bits(self, logn, tile_bits):
        _lib.check(_L().vkhel_ext_set_tile_bits(self._h, logn, tile_bits))

    def pass_mode(self, logn):
        m = _L().vkhel_ext_pass_mode(self._h, logn)
        _lib.check()
        return m

    def set_pass_mode(self, logn, mode):
        _lib.check(_L().vkhel_ext_set_pass_mode(self._h, logn, mode))

    def split(self, logn):
        """vkhel_ext_split: stage split choice of log2 n (0 default, 1 alt)."""
        s = _L().vkhel_ext_split(self._h, logn)
        _lib.check()
        return s

    def set_split(self, logn, split):
        _lib.check(_L().vkhel_ext_set_split(self._h, logn, split))

    def tile_bits2(self, logn):
        """vkhel_ext_tile_bits2: the row pass's tile of log2 n (0: same)."""
        tb = _L().vkhel_ext_tile_bits2(self._h, logn)
        _lib.check()
        return tb

    def set_tile_bits2(self, logn, tile_bits):
        _lib.check(_L().vkhel_ext_set_tile_bits2(self._h, logn, tile_bits))

    def set_handoff_ahead(self, polys):
        """vkhel_ext_set_handoff_ahead: look-ahead of pass mode 3's schedule
        (0: from the occupancy)."""
        _lib.check(_L().vkhel_ext_set_handoff_ahead(self._h, polys))

    def autotune_save(self, path):
        _lib.check(_L().vkhel_ext_autotune_save(self._h, str(path).encode()))

    def autotune_load(self, path):
        _lib.check(_L().vkhel_ext_autotune_load(self._h, str(path).encode()))

    def vector(self, length, zero=True):
        return Vector(self, length, zero)

    def polymul_noalloc(self, a, b, res, tables, stream=None):
        """vkhel_ext_polymul_noalloc: res = a*b mod (X^n + 1, q_i) per limb,
        b's spectra in the library's static per-GPU scratch (no device
        allocation); enqueued on `stream` (a hipstream.Stream; None: the
        context's stream) without waiting.  Tables must be resident
        (NttTables.upload)."""
        tabs = list(tables)
        arr = (ctypes.c_void_p * len(tabs))(*[t.handle for t in tabs])
        s = stream.handle if stream is not None else None
        _lib.check(_L().vkhel_ext_polymul_noalloc(self._h, a.handle, b.handle,
                                                  res.handle, arr, len(tabs), s))

    def from_array(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint64)
        v = Vector(self, a.size, zero=False)
        v.copy_from_host(a)
        return v

    def close(self):
        if self._h:
            _L().vkhel_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class NttTables:
    """vkhel_ntt_tables_create(n, q, w) (src/ntt_tables.c:65-80)."""

    def __init__(self, n, q, w, _handle=None):
        self.n, self.q, self.w = int(n), int(q), int(w)
        self._h = _handle if _handle is not None else \
            _L().vkhel_ntt_tables_create(self.n, self.q, self.w)
        if not self._h:
            _lib.check()
            raise _lib.VkhelError(1, "vkhel_ntt_tables_create failed")

    @classmethod
    def on_device(cls, ctx, n, q, w):
        """vkhel_ext_tables_create_device: tables generated on the GPU of
        ctx; w must be a primitive 2n-th root of unity mod q."""
        h = _L().vkhel_ext_tables_create_device(ctx.handle, int(n), int(q), int(w))
        if not h:
            _lib.check()
            raise _lib.VkhelError(1, "vkhel_ext_tables_create_device failed")
        return cls(n, q, w, _handle=h)

    @property
    def handle(self):
        return self._h

    @property
    def root_ok(self):
        """True if w is a primitive 2n-th root of unity mod q."""
        r = _L().vkhel_ext_tables_root_ok(self._h)
        if r < 0:
            _lib.check()
        return r == 1

    def read_device(self, ctx):
        """(roots, roots_shoup, inv_roots, inv_roots_shoup) of the device
        copy on ctx in reference index order; the companions are the packed
        63-bit ones the kernels read."""
        out = [np.zeros(self.n, dtype=np.uint64) for _ in range(4)]
        _lib.check(_L().vkhel_ext_tables_read_device(
            ctx.handle, self._h, *[_u64p(a) for a in out]))
        return tuple(out)

    def upload(self, ctx):
        """vkhel_ext_tables_upload: make the device copy on ctx now."""
        _lib.check(_L().vkhel_ext_tables_upload(ctx.handle, self._h))

    def arrays(self):
        """(roots, inv_roots, roots_bf, inv_roots_bf), reference layout."""
        out = [np.zeros(self.n, dtype=np.uint64) for _ in range(4)]
        _lib.check(_L().vkhel_ext_tables_copy(self._h, *[_u64p(a) for a in out]))
        return tuple(out)

    def close(self):
        if self._h:
            _L().vkhel_ntt_tables_destroy(self._h)
            self._h = None


class Vector:
    """vkhel_vector and its operations (src/vector.c)."""

    def __init__(self, ctx, length, zero=True, _handle=None):
        self.ctx = ctx
        if _handle is not None:
            self._h = _handle
        else:
            self._h = _L().vkhel_vector_create2(ctx.handle, int(length), bool(zero))
            if not self._h:
                _lib.check()
                raise _lib.VkhelError(3, "vkhel_vector_create2 failed")
        self.length = int(_L().vkhel_ext_vector_length(self._h))

    @property
    def handle(self):
        return self._h

    def copy_from_host(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint64)
        if a.size < self.length:
            raise ValueError("host array shorter than the vector")
        _L().vkhel_vector_copy_from_host(self._h, _u64p(a))
        _lib.check()

    def map_read(self):
        """vkhel_vector_map + copy out + vkhel_vector_unmap."""
        L = _L()
        p = ctypes.c_void_p()
        L.vkhel_vector_map(self._h, ctypes.byref(p), self.length * 8)
        _lib.check()
        out = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint64)),
                                    shape=(max(self.length, 1),))[:self.length].copy()
        L.vkhel_vector_unmap(self._h)
        _lib.check()
        return out

    def read(self, offset=0, count=None):
        count = self.length - offset if count is None else count
        out = np.empty(count, dtype=np.uint64)
        _lib.check(_L().vkhel_ext_vector_read(self._h, offset, count, _u64p(out)))
        return out

    def write(self, arr, offset=0):
        a = np.ascontiguousarray(arr, dtype=np.uint64)
        _lib.check(_L().vkhel_ext_vector_write(self._h, offset, a.size, _u64p(a)))

    def write_async(self, arr, offset=0, stream=None):
        """Enqueue a host->device copy of `arr` (ideally a HostBuffer view)
        on `stream` (a hipstream.Stream; None: the context's stream)."""
        if not (arr.dtype == np.uint64 and arr.flags.c_contiguous):
            raise ValueError("write_async needs a contiguous uint64 array")
        s = stream.handle if stream is not None else None
        _lib.check(_L().vkhel_ext_vector_write_async(self._h, offset, arr.size,
                                                     _u64p(arr), s))

    def read_async(self, out, offset=0, stream=None):
        """Enqueue a device->host copy into `out` (ideally a HostBuffer
        view) on `stream`; `out` holds the words once the stream passes."""
        if not (out.dtype == np.uint64 and out.flags.c_contiguous):
            raise ValueError("read_async needs a contiguous uint64 array")
        s = stream.handle if stream is not None else None
        _lib.check(_L().vkhel_ext_vector_read_async(self._h, offset, out.size,
                                                    _u64p(out), s))

    def fill_uniform(self, q, seed, offset=0, count=None):
        """Synthetic residues in [0, q), generated on the device."""
        count = self.length - offset if count is None else count
        _lib.check(_L().vkhel_ext_vector_fill_uniform(self._h, offset, count, q, seed))

    def view(self, offset, length):
        """vkhel_ext_vector_view: a vector aliasing [offset, offset+length)."""
        h = _L().vkhel_ext_vector_view(self._h, offset, length)
        if not h:
            _lib.check()
        v = Vector(self.ctx, 0, _handle=h)
        v._base = self  # keep the aliased vector alive
        return v

    def dup(self):
        h = _L().vkhel_vector_dup(self._h)
        if not h:
            _lib.check()
        return Vector(self.ctx, 0, _handle=h)

    def device_ptr(self):
        return _L().vkhel_ext_vector_device_ptr(self._h)

    # ---- ops: self is the first operand, `out` the result vector ----
    def elemfma(self, b, out, multiplier, mod):
        _L().vkhel_vector_elemfma(self._h, b.handle, out.handle, multiplier, mod)
        _lib.check()

    def elemmod(self, out, mod, q):
        _L().vkhel_vector_elemmod(self._h, out.handle, mod, q)
        _lib.check()

    def elemmul(self, b, out, mod):
        _L().vkhel_vector_elemmul(self._h, b.handle, out.handle, mod)
        _lib.check()

    def elemgtadd(self, out, bound, diff):
        _L().vkhel_vector_elemgtadd(self._h, out.handle, bound, diff)
        _lib.check()

    def elemgtsub(self, out, bound, diff, mod):
        _L().vkhel_vector_elemgtsub(self._h, out.handle, bound, diff, mod)
        _lib.check()

    def forward_transform(self, out, tables):
        _L().vkhel_vector_forward_transform(self._h, out.handle, tables.handle)
        _lib.check()

    def inverse_transform(self, out, tables):
        _L().vkhel_vector_inverse_transform(self._h, out.handle, tables.handle)
        _lib.check()

    def close(self):
        if self._h:
            _L().vkhel_vector_destroy(self._h)
            self._h = None


class HostBuffer:
    """Page-locked host memory (vkhel_ext_host_alloc) of `words` uint64
    words; `array` is a numpy view of it."""

    def __init__(self, words):
        L = _L()
        self._p = L.vkhel_ext_host_alloc(max(words, 1) * 8)
        if not self._p:
            _lib.check()
            raise _lib.VkhelError(-1, "host_alloc failed")
        self.array = np.ctypeslib.as_array(
            ctypes.cast(self._p, ctypes.POINTER(ctypes.c_uint64)),
            shape=(max(words, 1),))[:words]

    def close(self):
        if self._p:
            self.array = None
            _L().vkhel_ext_host_free(self._p)
            self._p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """vkhel_batch_* (include/vkhel/vkhel_ext.h): many RNS jobs, one graph."""

    def __init__(self, ctx):
        self.ctx = ctx
        self._h = _L().vkhel_batch_create(ctx.handle)
        if not self._h:
            _lib.check()

    def add(self, op, a, b, res, tables):
        tabs = list(tables)
        arr = (ctypes.c_void_p * len(tabs))(*[t.handle for t in tabs])
        _lib.check(_L().vkhel_batch_add(self._h, op, a.handle,
                                        b.handle if b is not None else None,
                                        res.handle, arr, len(tabs)))

    def add_rns(self, op, a, res, tables, params=None):
        """vkhel_batch_add_rns: JOB_CENTER or JOB_INTT_CENTER (params [p]),
        JOB_GTADD ([bound_i, diff_i] per limb), JOB_GTSUB
        ([bound_i, diff_i, mod_i]) or JOB_RESCALE / JOB_RESCALE_NTT (no
        params)."""
        tabs = list(tables)
        arr = (ctypes.c_void_p * len(tabs))(*[t.handle for t in tabs])
        prm = None if params is None else np.ascontiguousarray(params, dtype=np.uint64)
        _lib.check(_L().vkhel_batch_add_rns(self._h, op, a.handle, res.handle,
                                            arr, len(tabs),
                                            None if prm is None else _u64p(prm)))

    def clear(self):
        _L().vkhel_batch_clear(self._h)

    def launch(self):
        _lib.check(_L().vkhel_batch_launch(self._h))

    def wait(self):
        """vkhel_batch_wait: blocks until the last launch finished."""
        _lib.check(_L().vkhel_batch_wait(self._h))

    def done(self):
        """vkhel_batch_query: True once the last launch finished."""
        r = _L().vkhel_batch_query(self._h)
        if r < 0:
            _lib.check()
        return r == 1

    def run(self):
        _lib.check(_L().vkhel_batch_run(self._h))

    def __len__(self):
        return int(_L().vkhel_batch_num_jobs(self._h))

    def close(self):
        if self._h:
            _L().vkhel_batch_destroy(self._h)
            self._h = None


class ServerConfig(ctypes.Structure):
    """struct vkhel_server_config (include/vkhel/vkhel_ext.h)."""
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64),
                ("max_inflight", ctypes.c_uint32),
                ("max_batch", ctypes.c_uint32),
                ("streams", ctypes.c_uint32),
                ("queue_capacity", ctypes.c_uint32),
                ("scratch_bytes", ctypes.c_uint64),
                ("task_limbs", ctypes.c_uint32),
                ("per_call", ctypes.c_uint32)]


class RequestInfo(ctypes.Structure):
    """struct vkhel_request_info."""
    _fields_ = [("status", ctypes.c_int),
                ("submit_ns", ctypes.c_uint64),
                ("admit_ns", ctypes.c_uint64),
                ("complete_ns", ctypes.c_uint64),
                ("batch_size", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    """struct vkhel_server_stats."""
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64),
                ("cancelled", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("transform_coeffs", ctypes.c_uint64),
                ("inflight_bytes", ctypes.c_uint64),
                ("peak_inflight_bytes", ctypes.c_uint64),
                ("inflight", ctypes.c_uint32), ("waiting", ctypes.c_uint32),
                ("reserved_bytes", ctypes.c_uint64)]


OK, ERR_INVALID, ERR_DEVICE, ERR_NOMEM, ERR_BUDGET, ERR_CANCELLED, ERR_TIMEOUT, \
    ERR_BUSY = range(8)


class JobDesc(ctypes.Structure):
    """struct vkhel_job_desc."""
    _fields_ = [("op", ctypes.c_int), ("a", ctypes.c_void_p), ("b", ctypes.c_void_p),
                ("res", ctypes.c_void_p), ("tables", ctypes.c_void_p),
                ("nlimbs", ctypes.c_uint32)]


class Server:
    """vkhel_server_*: continuous batching over device-resident operands."""

    def __init__(self, ctx, hbm_budget_bytes=0, max_inflight=0, max_batch=0,
                 streams=0, queue_capacity=0, scratch_bytes=0, task_limbs=0,
                 per_call=False):
        self.ctx = ctx
        self.config = ServerConfig(hbm_budget_bytes, max_inflight, max_batch,
                                   streams, queue_capacity, scratch_bytes,
                                   task_limbs, int(bool(per_call)))
        self._h = _L().vkhel_server_create(ctx.handle, ctypes.byref(self.config))
        if not self._h:
            _lib.check()

    def submit(self, op, a, b, res, tables):
        """Returns the request id; raises VkhelError (BUDGET, BUSY, ...)."""
        tabs = list(tables)
        arr = (ctypes.c_void_p * len(tabs))(*[t.handle for t in tabs])
        rid = _L().vkhel_server_submit(self._h, op, a.handle,
                                       b.handle if b is not None else None,
                                       res.handle, arr, len(tabs))
        if not rid:
            _lib.check()
            raise _lib.VkhelError(ERR_INVALID, "submit failed")
        return int(rid)

    def submit_rns(self, op, a, res, tables, params=None):
        """vkhel_server_submit_rns: the RNS jobs of Batch.add_rns."""
        tabs = list(tables)
        arr = (ctypes.c_void_p * len(tabs))(*[t.handle for t in tabs])
        prm = None if params is None else np.ascontiguousarray(params, dtype=np.uint64)
        rid = _L().vkhel_server_submit_rns(self._h, op, a.handle, res.handle, arr,
                                           len(tabs),
                                           None if prm is None else _u64p(prm))
        if not rid:
            _lib.check()
            raise _lib.VkhelError(ERR_INVALID, "submit failed")
        return int(rid)

    def wait(self, rid, timeout_ns=0):
        """(status, info); status ERR_TIMEOUT leaves the request pending."""
        info = RequestInfo()
        st = _L().vkhel_server_wait(self._h, rid, timeout_ns, ctypes.byref(info))
        if st == ERR_INVALID:
            _lib.check()
        return st, info

    def cancel(self, rid):
        _lib.check(_L().vkhel_server_cancel(self._h, rid))

    def drain(self):
        _lib.check(_L().vkhel_server_drain(self._h))

    def replay(self, jobs, rate_per_s, seed=1):
        """Submits (op, a, b, res, tables) jobs at Poisson arrivals, waits for
        all; returns (status, [RequestInfo])."""
        n = len(jobs)
        descs = (JobDesc * n)()
        keep = []
        for d, (op, a, b, res, tables) in zip(descs, jobs):
            arr = (ctypes.c_void_p * len(tables))(*[t.handle for t in tables])
            keep.append(arr)
            d.op, d.a = op, a.handle
            d.b = b.handle if b is not None else None
            d.res, d.tables = res.handle, ctypes.cast(arr, ctypes.c_void_p)
            d.nlimbs = len(tables)
        infos = (RequestInfo * n)()
        st = _L().vkhel_server_replay(self._h, descs, n, float(rate_per_s), seed, infos)
        if st == ERR_INVALID:
            _lib.check()
        _L().vkhel_ext_clear_error()
        return st, list(infos)

    def stats(self):
        s = ServerStats()
        _lib.check(_L().vkhel_server_get_stats(self._h, ctypes.byref(s)))
        return s

    def close(self):
        if self._h:
            _L().vkhel_server_destroy(self._h)
            self._h = None
