"""MI355X (gfx950) closed-shell CCSD amplitude engine.

Python access to libaqamd.so (the C ABI declared in include/aqamd.h) for
tests, benchmarks and scripting.  There is no CPU fallback: importing the
compute entry points without the built library raises, and every call that
fails in the library raises AqError with the library's message.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaqamd.so")

AQ_PREC_FP64, AQ_PREC_FP32, AQ_PREC_BF16X3 = 0, 1, 2
PRECISIONS = {"fp64": AQ_PREC_FP64, "fp32": AQ_PREC_FP32, "bf16x3": AQ_PREC_BF16X3}


class AqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[aqamd error {code}] {msg}")
        self.code = code


class TimingRecord(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("flops", C.c_double), ("bytes", C.c_double)]


class RHFCCSDConfig(C.Structure):
    _fields_ = [("convergence", C.c_double), ("max_iterations", C.c_int), ("diis_order", C.c_int),
                ("diis_start", C.c_int), ("diis_damping", C.c_double), ("precision", C.c_int),
                ("refine_below", C.c_double), ("min_reduced_flops", C.c_double), ("fused_update", C.c_int),
                ("diis_reset_on_refine", C.c_int), ("fp32_stage_below", C.c_double), ("graph", C.c_int),
                ("packed_ladder", C.c_int)]


class TuneEntry(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("trans_a", C.c_int), ("trans_b", C.c_int),
                ("m_bucket", C.c_int), ("n_bucket", C.c_int), ("k_bucket", C.c_int), ("batch_bucket", C.c_int),
                ("variant", C.c_int), ("ms", C.c_float * 5)]


ITER_CB = C.CFUNCTYPE(None, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_void_p)

_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)

# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    "aq_last_error": [],
    "aq_version": [],
    "aq_ctx_create": [C.c_int, C.c_size_t, C.POINTER(_P)],
    "aq_ctx_destroy": [_P],
    "aq_ctx_sync": [_P],
    "aq_ctx_set_precision": [_P, C.c_int, C.c_double],
    "aq_ctx_get_precision": [_P, C.POINTER(C.c_int)],
    "aq_ctx_set_gemm_variant": [_P, C.c_int, C.c_int],
    "aq_ctx_set_autotune": [_P, C.c_int],
    "aq_autotune_table": [C.POINTER(TuneEntry), C.c_int, C.POINTER(C.c_int)],
    "aq_ctx_memory": [_P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "aq_ctx_trim": [_P],
    "aq_ctx_enable_timing": [_P, C.c_int],
    "aq_ctx_timing": [_P, C.POINTER(TimingRecord), C.c_int, C.POINTER(C.c_int)],
    "aq_ctx_timing_reset": [_P],
    "aq_tensor_create": [_P, C.c_int, _I64P, C.c_int, C.POINTER(_P)],
    "aq_tensor_copy": [_P, _P, C.POINTER(_P)],
    "aq_tensor_destroy": [_P],
    "aq_tensor_ndim": [_P, C.POINTER(C.c_int)],
    "aq_tensor_lengths": [_P, _I64P],
    "aq_tensor_size": [_P, _I64P],
    "aq_tensor_data": [_P, C.POINTER(_P)],
    "aq_tensor_touch": [_P],
    "aq_tensor_resize": [_P, C.c_int, C.POINTER(C.c_int64), C.c_int],
    "aq_contract": [_P, C.c_double, _P, C.c_char_p, _P, C.c_char_p, C.c_double, _P, C.c_char_p],
    "aq_sum": [_P, C.c_double, _P, C.c_char_p, C.c_double, _P, C.c_char_p],
    "aq_scale": [_P, _P, C.c_double, C.c_char_p],
    "aq_fill": [_P, _P, C.c_double, C.c_double],
    "aq_dot": [_P, _P, C.c_char_p, _P, C.c_char_p, _DP],
    "aq_weight": [_P, _P, C.c_int, C.POINTER(_DP), C.c_double],
    "aq_weight_noalloc": [_P, _P, C.c_int, C.POINTER(_DP), C.c_double],
    "aq_norm": [_P, _P, C.c_int, _DP],
    "aq_read": [_P, _P, C.c_int64, _I64P, _DP],
    "aq_write": [_P, _P, C.c_int64, _I64P, _DP, C.c_double, C.c_double],
    "aq_read_all": [_P, _P, _DP],
    "aq_write_all": [_P, _P, _DP],
    "aq_slice": [_P, C.c_double, _P, _I64P, C.c_double, _P, _I64P, _I64P],
    "aq_div": [_P, C.c_double, _P, _P, C.c_double, _P],
    "aq_invert": [_P, C.c_double, _P, C.c_double, _P],
    "aq_rhfccsd_default_config": [C.POINTER(RHFCCSDConfig)],
    "aq_rhfccsd_create": [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(RHFCCSDConfig), C.POINTER(_P)],
    "aq_rhfccsd_destroy": [_P],
    "aq_rhfccsd_mp2": [_P, _DP],
    "aq_rhfccsd_iterate": [_P, _DP, _DP],
    "aq_rhfccsd_run": [_P, ITER_CB, C.c_void_p, _DP, _DP, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "aq_rhfccsd_tensor": [_P, C.c_int, C.POINTER(_P)],
    "aq_rhfccsd_hbar": [_P, C.c_int],
    "aq_rhfccsd_current_precision": [_P, C.POINTER(C.c_int)],
    "aq_rhfccsd_graph_stats": [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "aq_rhfccsd_packed_ladder": [_P, C.POINTER(C.c_int)],
    "aq_synth_host": [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_double, _DP, _DP],
    "aq_synth_default_sigma": [C.c_int, C.c_int],
    "aq_synth_build": [_P, C.c_int, C.c_int, C.c_int, _DP, _DP, C.POINTER(_P)],
}
_RESTYPE = {"aq_last_error": C.c_char_p, "aq_version": C.c_int, "aq_synth_default_sigma": C.c_double}

_lib = None


def lib():
    """Load libaqamd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m aquarius_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, C.c_int)
        _lib = L
    return _lib


def _ck(rc):
    if rc != 0:
        raise AqError(rc, lib().aq_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(_DP)


class Context:
    """aq_ctx: one device, one stream, one HBM budget."""

    def __init__(self, device=0, hbm_budget=0):
        self.h = _P()
        _ck(lib().aq_ctx_create(device, hbm_budget, C.byref(self.h)))

    def close(self):
        if self.h:
            _ck(lib().aq_ctx_destroy(self.h))
            self.h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        _ck(lib().aq_ctx_sync(self.h))

    def set_precision(self, prec, min_reduced_flops=0.0):
        prec = PRECISIONS.get(prec, prec)
        _ck(lib().aq_ctx_set_precision(self.h, int(prec), float(min_reduced_flops)))

    def set_gemm_variant(self, prec, variant):
        _ck(lib().aq_ctx_set_gemm_variant(self.h, int(PRECISIONS.get(prec, prec)), int(variant)))

    def set_autotune(self, enable=True):
        _ck(lib().aq_ctx_set_autotune(self.h, 1 if enable else 0))

    def memory(self):
        a, b = C.c_size_t(), C.c_size_t()
        _ck(lib().aq_ctx_memory(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def trim(self):
        _ck(lib().aq_ctx_trim(self.h))

    def enable_timing(self, on=True):
        _ck(lib().aq_ctx_enable_timing(self.h, 1 if on else 0))

    def timing(self):
        n = C.c_int()
        _ck(lib().aq_ctx_timing(self.h, None, 0, C.byref(n)))
        recs = (TimingRecord * max(n.value, 1))()
        _ck(lib().aq_ctx_timing(self.h, recs, n.value, C.byref(n)))
        return {r.name.decode(): {"launches": r.launches, "total_ms": r.total_ms, "flops": r.flops,
                                  "bytes": r.bytes} for r in recs[:n.value]}

    def timing_reset(self):
        _ck(lib().aq_ctx_timing_reset(self.h))

    # tensor factory
    def tensor(self, shape, zero=True):
        return Tensor(self, shape, zero)

    def from_numpy(self, arr):
        arr = np.asarray(arr, dtype=np.float64)
        t = Tensor(self, arr.shape, zero=False)
        t.write_all(arr)
        return t

    # operators (CTFTensor surface)
    def contract(self, alpha, A, ia, B, ib, beta, Cc, ic):
        _ck(lib().aq_contract(self.h, alpha, A.h, ia.encode(), B.h, ib.encode(), beta, Cc.h, ic.encode()))

    def sum(self, alpha, A, ia, beta, B, ib):
        _ck(lib().aq_sum(self.h, alpha, A.h, ia.encode(), beta, B.h, ib.encode()))

    def scale(self, X, alpha, idx):
        _ck(lib().aq_scale(self.h, X.h, alpha, idx.encode()))

    def fill(self, X, alpha, beta):
        _ck(lib().aq_fill(self.h, X.h, alpha, beta))

    def dot(self, A, ia, B, ib):
        r = C.c_double()
        _ck(lib().aq_dot(self.h, A.h, ia.encode(), B.h, ib.encode(), C.byref(r)))
        return r.value

    def weight(self, X, ds, shift=0.0):
        arrs = [np.ascontiguousarray(d, dtype=np.float64) for d in ds]
        ptrs = (_DP * len(arrs))(*[_dp(a) for a in arrs])
        _ck(lib().aq_weight(self.h, X.h, len(arrs), ptrs, shift))

    def norm(self, X, p):
        r = C.c_double()
        _ck(lib().aq_norm(self.h, X.h, p, C.byref(r)))
        return r.value

    def div(self, alpha, A, B, beta, Cc):
        _ck(lib().aq_div(self.h, alpha, A.h, B.h, beta, Cc.h))

    def invert(self, alpha, A, beta, Cc):
        _ck(lib().aq_invert(self.h, alpha, A.h, beta, Cc.h))

    def slice(self, alpha, A, startA, beta, B, startB, extent):
        a = np.asarray(startA, dtype=np.int64)
        b = np.asarray(startB, dtype=np.int64)
        e = np.asarray(extent, dtype=np.int64)
        _ck(lib().aq_slice(self.h, alpha, A.h, a.ctypes.data_as(_I64P), beta, B.h, b.ctypes.data_as(_I64P),
                           e.ctypes.data_as(_I64P)))


class Tensor:
    """Dense column-major FP64 device tensor (aq_tensor)."""

    def __init__(self, ctx, shape, zero=True, handle=None, owned=True):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.owned = owned
        if handle is not None:
            self.h = handle
            return
        self.h = _P()
        lens = np.asarray(self.shape, dtype=np.int64)
        _ck(lib().aq_tensor_create(ctx.h, len(self.shape), lens.ctypes.data_as(_I64P), 1 if zero else 0,
                                   C.byref(self.h)))

    @property
    def size(self):
        n = 1
        for s in self.shape:
            n *= s
        return n

    def copy(self):
        """Deep copy on the device (aq_tensor_copy)."""
        h = _P()
        _ck(lib().aq_tensor_copy(self.ctx.h, self.h, C.byref(h)))
        return Tensor(self.ctx, self.shape, handle=h)

    def free(self):
        if self.h and self.owned:
            _ck(lib().aq_tensor_destroy(self.h))
        self.h = _P()

    def data_ptr(self):
        """Device address of the column-major data (aq_tensor_data)."""
        p = _P()
        _ck(lib().aq_tensor_data(self.h, C.byref(p)))
        return p.value

    def touch(self):
        """Declare an outside write through data_ptr() (aq_tensor_touch)."""
        _ck(lib().aq_tensor_touch(self.h))

    def resize(self, shape, zero=True):
        """New lengths and fresh storage (aq_tensor_resize; CTFTensor::resize)."""
        shape = tuple(int(x) for x in shape)
        ln = (C.c_int64 * max(len(shape), 1))(*shape)
        _ck(lib().aq_tensor_resize(self.h, len(shape), ln, 1 if zero else 0))
        self.shape = shape

    def write_all(self, arr):
        arr = np.asarray(arr, dtype=np.float64)
        if arr.shape != self.shape:
            raise ValueError(f"shape {arr.shape} != tensor shape {self.shape}")
        flat = np.ascontiguousarray(arr.ravel(order="F"))
        _ck(lib().aq_write_all(self.ctx.h, self.h, _dp(flat)))

    def numpy(self):
        flat = np.empty(self.size, dtype=np.float64)
        _ck(lib().aq_read_all(self.ctx.h, self.h, _dp(flat)))
        return flat.reshape(self.shape, order="F")

    def read(self, keys):
        keys = np.ascontiguousarray(keys, dtype=np.int64)
        vals = np.empty(len(keys))
        _ck(lib().aq_read(self.ctx.h, self.h, len(keys), keys.ctypes.data_as(_I64P), _dp(vals)))
        return vals

    def write(self, keys, vals, alpha=1.0, beta=0.0):
        keys = np.ascontiguousarray(keys, dtype=np.int64)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        _ck(lib().aq_write(self.ctx.h, self.h, len(keys), keys.ctypes.data_as(_I64P), _dp(vals), alpha, beta))


def autotune_table():
    """The GEMM tuning table: one dict per shape bucket."""
    n = C.c_int()
    _ck(lib().aq_autotune_table(None, 0, C.byref(n)))
    arr = (TuneEntry * max(n.value, 1))()
    _ck(lib().aq_autotune_table(arr, n.value, C.byref(n)))
    return [{f: (list(getattr(e, f)) if f == "ms" else getattr(e, f)) for f, _ in TuneEntry._fields_}
            for e in arr[:n.value]]


def default_config(**over):
    cfg = RHFCCSDConfig()
    _ck(lib().aq_rhfccsd_default_config(C.byref(cfg)))
    for k, v in over.items():
        if k == "precision":
            v = PRECISIONS.get(v, v)
        setattr(cfg, k, v)
    return cfg


FOCK = ("fAI", "fAB", "fIJ")
BLOCKS = ("VABCD", "VABCI", "VABIJ", "VAIBJ", "VAIJB", "VAIJK", "VIJKL")
STATE = {"T1": 0, "T2": 1, "Z1": 2, "Z2": 3, "Tau": 4, "FAE": 10, "FMI": 11, "FME": 12, "WMNIJ": 13,
         "WMNEJ": 14, "WAMIJ": 15, "WAMEI": 16, "WAMIE": 17, "WABEJ": 18}


class Integrals:
    """Fock blocks + Dirac blocks resident on the device."""

    def __init__(self, ctx, tensors):
        self.ctx = ctx
        self.t = tensors   # dict name -> Tensor

    @classmethod
    def from_numpy(cls, ctx, f, V):
        t = {k: ctx.from_numpy(f[k]) for k in FOCK}
        t.update({k: ctx.from_numpy(V[k]) for k in BLOCKS})
        return cls(ctx, t)

    @classmethod
    def synthetic(cls, ctx, seed, o, v, nQ=None, host=None, sigma=0.0):
        """Seeded synthetic construction built on the device.  Returns the
        integrals; `host` (eps, B) may be passed to reuse a host draw."""
        nQ = 3 * (o + v) if nQ is None else nQ
        eps, B = synth_host(seed, o, v, nQ, sigma) if host is None else host
        out = (_P * 10)()
        _ck(lib().aq_synth_build(ctx.h, o, v, nQ, _dp(eps), _dp(B), out))
        shapes = [(v, o), (v, v), (o, o), (v, v, v, v), (v, v, v, o), (v, v, o, o), (v, o, v, o), (v, o, o, v),
                  (v, o, o, o), (o, o, o, o)]
        names = FOCK + BLOCKS
        return cls(ctx, {n: Tensor(ctx, s, handle=_P(out[k])) for k, (n, s) in enumerate(zip(names, shapes))})

    def free(self):
        for t in self.t.values():
            t.free()


def synth_host(seed, o, v, nQ, sigma=0.0):
    """(eps[o+v], B[nQ, o+v, o+v]) of the seeded synthetic construction
    (sigma <= 0: the size-normalised default scale)."""
    n = o + v
    eps = np.empty(n)
    B = np.empty((nQ, n, n))
    _ck(lib().aq_synth_host(seed, o, v, nQ, float(sigma), _dp(eps), _dp(B)))
    return eps, B


def synth_default_sigma(v, nQ):
    return lib().aq_synth_default_sigma(v, nQ)


class RHFCCSD:
    """aq_rhfccsd: the closed-shell CCSD solver (RHFCCSD, rhfccsd.cxx)."""

    def __init__(self, ctx, ints, config=None, **over):
        self.ctx = ctx
        self.ints = ints
        cfg = config if config is not None else default_config(**over)
        f3 = (_P * 3)(*[ints.t[k].h for k in FOCK])
        v7 = (_P * 7)(*[ints.t[k].h for k in BLOCKS])
        self.h = _P()
        _ck(lib().aq_rhfccsd_create(ctx.h, f3, v7, C.byref(cfg), C.byref(self.h)))
        self.cfg = cfg
        self.v, self.o = ints.t["fAI"].shape
        self.history = []

    def close(self):
        if self.h:
            _ck(lib().aq_rhfccsd_destroy(self.h))
            self.h = _P()

    @property
    def mp2(self):
        r = C.c_double()
        _ck(lib().aq_rhfccsd_mp2(self.h, C.byref(r)))
        return r.value

    def iterate(self):
        e, c = C.c_double(), C.c_double()
        _ck(lib().aq_rhfccsd_iterate(self.h, C.byref(e), C.byref(c)))
        return e.value, c.value

    def run(self, callback=None):
        hist = self.history

        def _cb(it, e, c, dt, prec, user):
            hist.append({"iter": it, "energy": e, "conv": c, "seconds": dt, "precision": prec})
            if callback is not None:
                callback(hist[-1])

        cb = ITER_CB(_cb)
        e, c, n, ok = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        _ck(lib().aq_rhfccsd_run(self.h, cb, None, C.byref(e), C.byref(c), C.byref(n), C.byref(ok)))
        return {"energy": e.value, "conv": c.value, "iterations": n.value, "converged": bool(ok.value)}

    def hbar(self, with_wabej=False):
        """Form the Hbar products of the current amplitudes (rhfccsd.cxx:108-188)."""
        _ck(lib().aq_rhfccsd_hbar(self.h, 1 if with_wabej else 0))

    def tensor(self, name):
        p = _P()
        _ck(lib().aq_rhfccsd_tensor(self.h, STATE[name], C.byref(p)))
        n = C.c_int()
        _ck(lib().aq_tensor_ndim(p, C.byref(n)))
        lens = np.zeros(n.value, dtype=np.int64)
        _ck(lib().aq_tensor_lengths(p, lens.ctypes.data_as(_I64P)))
        return Tensor(self.ctx, tuple(lens), handle=p, owned=False)

    def current_precision(self):
        r = C.c_int()
        _ck(lib().aq_rhfccsd_current_precision(self.h, C.byref(r)))
        return r.value

    @property
    def packed_ladder(self):
        """True when the ladder runs on packed (anti)symmetric pairs."""
        r = C.c_int()
        _ck(lib().aq_rhfccsd_packed_ladder(self.h, C.byref(r)))
        return bool(r.value)

    def graph_stats(self):
        """(captures, replays, abandoned captures) of the contraction-stage HIP graph."""
        c, r, f = C.c_int(), C.c_int(), C.c_int()
        _ck(lib().aq_rhfccsd_graph_stats(self.h, C.byref(c), C.byref(r), C.byref(f)))
        return c.value, r.value, f.value


# ---------------------------------------------------------------- DIIS

_SIGS.update({
    "aq_diis_create": [_P, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.POINTER(_P)],
    "aq_diis_destroy": [_P],
    "aq_diis_extrapolate": [_P, C.POINTER(_P), C.POINTER(_P)],
    "aq_diis_coefficients": [_P, _DP, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)],
})


class DIIS:
    """Device DIIS extrapolator (aq_diis_*; DIIS<T>, diis.hpp:36-227)."""

    def __init__(self, ctx, order, start=1, damping=0.0, nx=1, ndx=1):
        self.ctx, self.order = ctx, order
        h = _P()
        _ck(lib().aq_diis_create(ctx.h, order, start, damping, nx, ndx, C.byref(h)))
        self.h = h

    def extrapolate(self, x, dx):
        px = (_P * len(x))(*[t.h for t in x])
        pdx = (_P * len(dx))(*[t.h for t in dx])
        _ck(lib().aq_diis_extrapolate(self.h, px, pdx))

    def coefficients(self):
        buf = (C.c_double * (self.order + 1))()
        n, retry = C.c_int(), C.c_int()
        _ck(lib().aq_diis_coefficients(self.h, buf, self.order + 1, C.byref(n), C.byref(retry)))
        return np.array(buf[:n.value]), bool(retry.value)

    def close(self):
        if self.h:
            _ck(lib().aq_diis_destroy(self.h))
        self.h = _P()

# ---------------------------------------------------------------- serving

class JobSpec(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("o", C.c_int), ("v", C.c_int), ("nQ", C.c_int), ("sigma", C.c_double),
                ("cfg", RHFCCSDConfig), ("inputs", C.POINTER(_P)), ("fcidump", C.c_char_p)]


class JobResult(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int), ("slot", C.c_int), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_end", C.c_double), ("energy", C.c_double), ("conv", C.c_double),
                ("mp2", C.c_double), ("iterations", C.c_int), ("converged", C.c_int),
                ("reduced_iterations", C.c_int), ("bytes_reserved", C.c_size_t), ("bytes_peak", C.c_size_t),
                ("error_code", C.c_int), ("error", C.c_char * 160)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "error"}
        d["error"] = self.error.decode(errors="replace")
        return d


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("failed", C.c_int64), ("queued", C.c_int64),
                ("running", C.c_int64), ("budget", C.c_size_t), ("reserved", C.c_size_t), ("elapsed_s", C.c_double)]


JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3

_SIGS.update({
    "aq_ctx_peak_memory": [_P, C.POINTER(C.c_size_t), C.c_int],
    "aq_job_bytes_estimate": [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "aq_server_create": [C.c_int, C.c_size_t, C.c_int, C.c_double, C.POINTER(_P)],
    "aq_server_submit": [_P, C.POINTER(JobSpec), C.POINTER(C.c_int64)],
    "aq_server_poll": [_P, C.c_int64, C.POINTER(JobResult)],
    "aq_server_wait": [_P, C.c_int64, C.POINTER(JobResult)],
    "aq_server_drain": [_P],
    "aq_server_stats_get": [_P, C.POINTER(ServerStats)],
    "aq_server_destroy": [_P],
    "aq_server_set_grouping": [_P, C.c_int],
    "aq_server_grouping_stats": [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                 C.POINTER(C.c_int64)],
})
_RESTYPE["aq_job_bytes_estimate"] = C.c_size_t


def job_bytes_estimate(o, v, nQ=0, diis_order=8, precision="fp64"):
    return lib().aq_job_bytes_estimate(o, v, nQ, diis_order, int(PRECISIONS.get(precision, precision)))


class Server:
    """aq_server: continuous batching of synthetic CCSD jobs on one device."""

    def __init__(self, device=0, hbm_budget=0, slots=4, max_bypass_s=30.0):
        self.h = _P()
        _ck(lib().aq_server_create(device, int(hbm_budget), slots, float(max_bypass_s), C.byref(self.h)))

    def submit(self, seed, o, v, nQ=0, sigma=0.0, config=None, inputs=None, fcidump=None, **over):
        """Queue a job: the synthetic system (seed, o, v, nQ, sigma), or the
        caller's integrals (`inputs`: an Integrals on this device, borrowed
        until the job is done), or an FCIDUMP file (`fcidump`: a path)."""
        spec = JobSpec(seed=seed, o=o, v=v, nQ=nQ, sigma=sigma,
                       cfg=config if config is not None else default_config(**over))
        keep = None
        if inputs is not None:
            keep = (_P * 10)(*[inputs.t[k].h for k in FOCK + BLOCKS])
            spec.inputs = C.cast(keep, C.POINTER(_P))
        if fcidump is not None:
            spec.fcidump = os.fsencode(str(fcidump))
        jid = C.c_int64()
        _ck(lib().aq_server_submit(self.h, C.byref(spec), C.byref(jid)))
        del keep
        return jid.value

    def poll(self, jid):
        r = JobResult()
        _ck(lib().aq_server_poll(self.h, jid, C.byref(r)))
        return r.as_dict()

    def wait(self, jid):
        r = JobResult()
        _ck(lib().aq_server_wait(self.h, jid, C.byref(r)))
        return r.as_dict()

    def drain(self):
        _ck(lib().aq_server_drain(self.h))

    def stats(self):
        s = ServerStats()
        _ck(lib().aq_server_stats_get(self.h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def set_grouping(self, window_us):
        """Cross-job grouped GEMM launches with a host wait of at most
        window_us per GEMM (0: off); only while no job is queued or running."""
        _ck(lib().aq_server_set_grouping(self.h, int(window_us)))

    def grouping_stats(self):
        n, g, mx, solo = C.c_int64(), C.c_int64(), C.c_int(), C.c_int64()
        _ck(lib().aq_server_grouping_stats(self.h, C.byref(n), C.byref(g), C.byref(mx), C.byref(solo)))
        return {"launches": n.value, "gemms": g.value, "max_group": mx.value, "solo": solo.value}

    def close(self):
        if self.h:
            _ck(lib().aq_server_destroy(self.h))
            self.h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ------------------------------------------------------ integral front end

ANGSTROM_TO_BOHR = 1.88972612456506198632428439        # molecule.cxx:1335-1336
BASIS_DIR = os.path.join(HERE, "basis")
SYMBOLS = ("H He Li Be B C N O F Ne Na Mg Al Si P S Cl Ar "
           "K Ca Sc Ti V Cr Mn Fe Co Ni Cu Zn Ga Ge As Se Br Kr").split()


class SCFConfig(C.Structure):
    _fields_ = [("convergence", C.c_double), ("max_iterations", C.c_int), ("diis_order", C.c_int)]


class SCFResult(C.Structure):
    _fields_ = [("energy", C.c_double), ("conv", C.c_double), ("iterations", C.c_int), ("converged", C.c_int),
                ("nocc", C.c_int), ("nocc_beta", C.c_int)]


_SIGS.update({
    "aq_molecule_create": [C.c_int, C.POINTER(C.c_int), _DP, C.c_char_p, C.c_int, C.c_int, C.POINTER(_P)],
    "aq_molecule_destroy": [_P],
    "aq_molecule_info": [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), _DP],
    "aq_ints_1e": [_P, _DP, _DP, _DP],
    "aq_ints_2e": [_P, C.c_int, _DP],
    "aq_ints_2e_device": [_P, _P, C.c_int, C.POINTER(_P)],
    "aq_scf_default_config": [C.POINTER(SCFConfig)],
    "aq_scf_rhf": [_P, _P, _P, C.POINTER(SCFConfig), C.POINTER(SCFResult), _DP, _DP, _DP],
    "aq_aomo_rhf": [_P, _P, C.c_int, _DP, _DP, C.POINTER(_P)],
})


def basis_text(name):
    """Text of a GENBAS file shipped under aquarius_amd/basis (cc-pVDZ, DZ)."""
    with open(os.path.join(BASIS_DIR, name + ".gbs")) as f:
        return f.read()


class Molecule:
    """aq_molecule: atoms + GENBAS basis (Molecule / BasisSet / Shell of the
    reference).  atoms: [(symbol, x, y, z)] in `units` ('bohr' or 'angstrom')."""

    def __init__(self, atoms, basis="cc-pVDZ", units="angstrom", spherical=True, charge=0, genbas=None):
        scale = ANGSTROM_TO_BOHR if units == "angstrom" else 1.0
        Z = np.array([[s.upper() for s in SYMBOLS].index(a[0].upper()) + 1 for a in atoms], dtype=np.int32)
        xyz = np.array([[a[1] * scale, a[2] * scale, a[3] * scale] for a in atoms], dtype=np.float64)
        text = genbas if genbas is not None else basis_text(basis)
        self.h = _P()
        _ck(lib().aq_molecule_create(len(atoms), Z.ctypes.data_as(C.POINTER(C.c_int)), _dp(xyz), text.encode(),
                                     1 if spherical else 0, charge, C.byref(self.h)))
        nao, ne, ns, en = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        _ck(lib().aq_molecule_info(self.h, C.byref(nao), C.byref(ne), C.byref(ns), C.byref(en)))
        self.nao, self.nelectrons, self.nshells, self.nuclear_repulsion = nao.value, ne.value, ns.value, en.value
        self.atoms = [(a[0], *xyz[k]) for k, a in enumerate(atoms)]

    def ints_1e(self):
        n = self.nao
        S, T, V = np.empty((n, n)), np.empty((n, n)), np.empty((n, n))
        _ck(lib().aq_ints_1e(self.h, _dp(S), _dp(T), _dp(V)))
        return S, T, V

    def ints_2e(self, nthreads=0):
        """Host AO tensor eri[m, n, l, s] = (mn|ls)."""
        n = self.nao
        flat = np.empty(n ** 4)
        _ck(lib().aq_ints_2e(self.h, nthreads, _dp(flat)))
        return flat.reshape((n, n, n, n), order="F")

    def ints_2e_device(self, ctx, nthreads=0):
        h = _P()
        _ck(lib().aq_ints_2e_device(ctx.h, self.h, nthreads, C.byref(h)))
        return Tensor(ctx, (self.nao,) * 4, handle=h)

    def close(self):
        if self.h:
            _ck(lib().aq_molecule_destroy(self.h))
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SCF:
    """Closed-shell SCF on the device (localaoscf) and the AO->MO step
    (aomoints) that turns it into CCSD input."""

    def __init__(self, ctx, mol, eri=None, convergence=None, max_iterations=None, diis_order=None, nthreads=0):
        self.ctx, self.mol = ctx, mol
        self.eri = eri if eri is not None else mol.ints_2e_device(ctx, nthreads)
        cfg = SCFConfig()
        _ck(lib().aq_scf_default_config(C.byref(cfg)))
        if convergence is not None:
            cfg.convergence = convergence
        if max_iterations is not None:
            cfg.max_iterations = max_iterations
        if diis_order is not None:
            cfg.diis_order = diis_order
        n = mol.nao
        self.eps = np.empty(n)
        self.C = np.empty((n, n), order="F")
        self.F = np.empty((n, n))
        r = SCFResult()
        _ck(lib().aq_scf_rhf(ctx.h, mol.h, self.eri.h, C.byref(cfg), C.byref(r), _dp(self.eps),
                             self.C.ctypes.data_as(_DP), _dp(self.F)))
        self.energy, self.conv, self.iterations = r.energy, r.conv, r.iterations
        self.converged, self.nocc = bool(r.converged), r.nocc

    def integrals(self):
        """Integrals (fAI, fAB, fIJ + seven Dirac blocks) on the device."""
        out = (_P * 10)()
        _ck(lib().aq_aomo_rhf(self.ctx.h, self.eri.h, self.nocc, self.C.ctypes.data_as(_DP), _dp(self.F), out))
        o, v = self.nocc, self.mol.nao - self.nocc
        shapes = [(v, o), (v, v), (o, o), (v, v, v, v), (v, v, v, o), (v, v, o, o), (v, o, v, o), (v, o, o, v),
                  (v, o, o, o), (o, o, o, o)]
        return Integrals(self.ctx, {nm: Tensor(self.ctx, s, handle=_P(out[k]))
                                    for k, (nm, s) in enumerate(zip(FOCK + BLOCKS, shapes))})


# ------------------------------------------------------------------ FCIDUMP

_SIGS.update({
    "aq_mo_rhf_blocks": [_P, _P, C.c_int, _DP, _DP, _DP, C.POINTER(_P)],
    "aq_fcidump_header": [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "aq_fcidump_read": [C.c_char_p, C.c_int, _DP, _DP, _DP],
    "aq_fcidump_write": [C.c_char_p, C.c_int, C.c_int, C.c_int, _DP, _DP, C.c_double, C.c_double],
})


def fcidump_read(path, full_1e=False):
    """(h[norb, norb], eri[norb]*4 Mulliken, ecore, norb, nelec, ms2) of an FCIDUMP file."""
    n, ne, ms2 = C.c_int(), C.c_int(), C.c_int()
    _ck(lib().aq_fcidump_header(str(path).encode(), C.byref(n), C.byref(ne), C.byref(ms2)))
    N = n.value
    h = np.empty(N * N)
    eri = np.empty(N ** 4)
    ec = C.c_double()
    _ck(lib().aq_fcidump_read(str(path).encode(), 1 if full_1e else 0, _dp(h), _dp(eri), C.byref(ec)))
    return h.reshape((N, N), order="F"), eri.reshape((N,) * 4, order="F"), ec.value, N, ne.value, ms2.value


def fcidump_write(path, h, eri, ecore, nelec, ms2=0, threshold=1e-14):
    h = np.asfortranarray(h, dtype=np.float64)
    eri = np.asfortranarray(eri, dtype=np.float64)
    n = h.shape[0]
    _ck(lib().aq_fcidump_write(str(path).encode(), n, nelec, ms2, h.ctypes.data_as(_DP), eri.ctypes.data_as(_DP),
                               float(ecore), float(threshold)))


def integrals_from_mo(ctx, h, mo_eri, nocc):
    """CCSD inputs from MO-basis integrals (aq_mo_rhf_blocks): returns
    (Integrals, electronic SCF energy, f)."""
    n = h.shape[0]
    t = ctx.from_numpy(mo_eri)
    out = (_P * 10)()
    e = C.c_double()
    f = np.empty((n, n), order="F")
    hh = np.asfortranarray(h, dtype=np.float64)
    try:
        _ck(lib().aq_mo_rhf_blocks(ctx.h, t.h, nocc, hh.ctypes.data_as(_DP), C.byref(e), f.ctypes.data_as(_DP), out))
    finally:
        t.free()
    o, v = nocc, n - nocc
    shapes = [(v, o), (v, v), (o, o), (v, v, v, v), (v, v, v, o), (v, v, o, o), (v, o, v, o), (v, o, o, v),
              (v, o, o, o), (o, o, o, o)]
    ints = Integrals(ctx, {nm: Tensor(ctx, s, handle=_P(out[k])) for k, (nm, s) in enumerate(zip(FOCK + BLOCKS, shapes))})
    return ints, e.value, f


def integrals_from_fcidump(ctx, path, full_1e=False):
    """Read an FCIDUMP file and build the CCSD inputs: (Integrals, E_SCF total, norb, nelec)."""
    h, eri, ecore, n, nelec, ms2 = fcidump_read(path, full_1e)
    if nelec % 2 or ms2 != 0:
        raise AqError(4, "the closed-shell engine needs an even electron count and MS2 = 0")
    ints, e, _ = integrals_from_mo(ctx, h, eri, nelec // 2)
    return ints, ecore + e, n, nelec


# ------------------------------------------------- open shells (spin-orbital)

SO_INPUTS = ("fAI", "fME", "fAE", "fMI", "VABIJ", "VMNEF", "VAMEF", "VABEJ", "VABEF", "VMNIJ", "VMNEJ", "VAMIJ",
             "VAMEI")

_SIGS.update({
    "aq_scf_uhf": [_P, _P, _P, C.c_int, C.POINTER(SCFConfig), C.POINTER(SCFResult), _DP, _DP, _DP],
    "aq_aomo_uhf": [_P, _P, C.c_int, C.c_int, _DP, _DP, C.POINTER(_P)],
    "aq_soccsd_create": [_P, C.POINTER(_P), C.POINTER(RHFCCSDConfig), C.c_int, C.POINTER(_P)],
    "aq_soccsd_destroy": [_P],
    "aq_soccsd_mp2": [_P, _DP],
    "aq_soccsd_iterate": [_P, _DP, _DP],
    "aq_soccsd_run": [_P, ITER_CB, C.c_void_p, _DP, _DP, C.POINTER(C.c_int), C.POINTER(C.c_int)],
    "aq_soccsd_tensor": [_P, C.c_int, C.POINTER(_P)],
})


class UHF:
    """Unrestricted SCF on the device (localaoscf, uhf.cxx) and the
    spin-orbital integrals of the spin-orbital CCSD path."""

    def __init__(self, ctx, mol, multiplicity=1, eri=None, convergence=None, max_iterations=None, diis_order=None,
                 nthreads=0):
        self.ctx, self.mol = ctx, mol
        self.eri = eri if eri is not None else mol.ints_2e_device(ctx, nthreads)
        cfg = SCFConfig()
        _ck(lib().aq_scf_default_config(C.byref(cfg)))
        for k, val in (("convergence", convergence), ("max_iterations", max_iterations), ("diis_order", diis_order)):
            if val is not None:
                setattr(cfg, k, val)
        n = mol.nao
        self.eps = np.empty(2 * n)
        self.C = np.empty(2 * n * n)
        self.F = np.empty(2 * n * n)
        r = SCFResult()
        _ck(lib().aq_scf_uhf(ctx.h, mol.h, self.eri.h, multiplicity, C.byref(cfg), C.byref(r), _dp(self.eps),
                             _dp(self.C), _dp(self.F)))
        self.energy, self.conv, self.iterations = r.energy, r.conv, r.iterations
        self.converged, self.nalpha, self.nbeta = bool(r.converged), r.nocc, r.nocc_beta

    def spin_orbital_integrals(self):
        out = (_P * 13)()
        _ck(lib().aq_aomo_uhf(self.ctx.h, self.eri.h, self.nalpha, self.nbeta, _dp(self.C), _dp(self.F), out))
        n = self.mol.nao
        o = self.nalpha + self.nbeta
        v = 2 * n - o
        shapes = [(v, o), (o, v), (v, v), (o, o), (v, v, o, o), (o, o, v, v), (v, o, v, v), (v, v, v, o),
                  (v, v, v, v), (o, o, o, o), (o, o, v, o), (v, o, o, o), (v, o, v, o)]
        return Integrals(self.ctx, {nm: Tensor(self.ctx, s, handle=_P(out[k]))
                                    for k, (nm, s) in enumerate(zip(SO_INPUTS, shapes))})


class SpinOrbitalCCSD:
    """aq_soccsd: spin-orbital CCSD (CCSD<U>, ccsd.cxx) on UHF integrals."""

    def __init__(self, ctx, ints, config=None, ccd=False, **over):
        self.ctx = ctx
        cfg = config if config is not None else default_config(**over)
        ins = (_P * 13)(*[ints.t[k].h for k in SO_INPUTS])
        self.h = _P()
        _ck(lib().aq_soccsd_create(ctx.h, ins, C.byref(cfg), 1 if ccd else 0, C.byref(self.h)))
        self.history = []

    @property
    def mp2(self):
        r = C.c_double()
        _ck(lib().aq_soccsd_mp2(self.h, C.byref(r)))
        return r.value

    def iterate(self):
        e, c = C.c_double(), C.c_double()
        _ck(lib().aq_soccsd_iterate(self.h, C.byref(e), C.byref(c)))
        return e.value, c.value

    def run(self):
        hist = self.history

        def _cb(it, e, c, dt, prec, user):
            hist.append({"iter": it, "energy": e, "conv": c, "seconds": dt, "precision": prec})

        cb = ITER_CB(_cb)
        e, c, n, ok = C.c_double(), C.c_double(), C.c_int(), C.c_int()
        _ck(lib().aq_soccsd_run(self.h, cb, None, C.byref(e), C.byref(c), C.byref(n), C.byref(ok)))
        return {"energy": e.value, "conv": c.value, "iterations": n.value, "converged": bool(ok.value)}

    def tensor(self, name):
        which = {"T1": 0, "T2": 1, "Z1": 2, "Z2": 3, "Tau": 4}[name]
        p = _P()
        _ck(lib().aq_soccsd_tensor(self.h, which, C.byref(p)))
        n = C.c_int()
        _ck(lib().aq_tensor_ndim(p, C.byref(n)))
        lens = np.zeros(n.value, dtype=np.int64)
        _ck(lib().aq_tensor_lengths(p, lens.ctypes.data_as(_I64P)))
        return Tensor(self.ctx, tuple(lens), handle=p, owned=False)

    def close(self):
        if self.h:
            _ck(lib().aq_soccsd_destroy(self.h))
            self.h = _P()
