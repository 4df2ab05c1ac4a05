"""ctypes binding of libitg (include/itg.h).

Loads the in-tree libitg.so next to this file and raises if it is missing:
there is no CPU fallback. Device buffers are torch tensors on the GPU (torch
is the allocator and the stream owner); libitg is switched onto torch's
current stream so library kernels and torch copies are ordered.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libitg.so")

ITG_OK, ITG_EDIM, ITG_EDTYPE, ITG_EOOM, ITG_EHIP, ITG_EUNSUPPORTED, ITG_EARG, ITG_ECONV = \
    0, -1, -2, -3, -4, -5, -6, -7
F64, F32, C64 = 0, 1, 2
MODE_NATIVE, MODE_F32, MODE_BF16X3 = 0, 1, 2
TRUNC_NDTENSORS, TRUNC_REFGPU = 0, 1
MODES = {"fp64": MODE_NATIVE, "native": MODE_NATIVE, "fp32": MODE_F32, "bf16x3": MODE_BF16X3}
# kernel classes of itg_timing
K_GEMM_SPLIT, K_GETT, K_WPASS, K_TRD_SYMV, K_TRD_POST = 0, 1, 2, 3, 4
KERNEL_NAMES = {K_GEMM_SPLIT: "gemm_split_wide_kernel / gemm_split_kernel (split-BF16 GEMM)",
                K_GETT: "contraction GEMM (GETT kernel or rocBLAS dgemm)",
                K_WPASS: "heff_wpass_kernel (W1.W2 pass)",
                K_TRD_SYMV: "trd_symv_kernel (tridiagonalisation symv)",
                K_TRD_POST: "trd_post_kernel (tridiagonalisation update)"}


class ITGError(RuntimeError):
    pass


class DimensionMismatch(ITGError, ValueError):
    """Raised where the reference throws DimensionMismatch."""


class ArgumentError(ITGError, ValueError):
    """Raised where the reference throws ArgumentError."""


_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p
_ctxp = C.c_void_p

# name: (restype, argtypes)
_SIGS = {
    "itg_init": (C.c_int, [C.c_int, C.POINTER(_ctxp)]),
    "itg_destroy": (C.c_int, [_ctxp]),
    "itg_sync": (C.c_int, [_ctxp]),
    "itg_last_error": (C.c_char_p, [_ctxp]),
    "itg_stream": (_vp, [_ctxp]),
    "itg_set_stream": (C.c_int, [_ctxp, _vp]),
    "itg_device_info": (C.c_int, [_ctxp, C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]),
    "itg_timing": (C.c_int, [_ctxp, C.c_int]),
    "itg_timing_set": (C.c_int, [_ctxp, C.c_int]),
    "itg_timing_report": (C.c_int, [_ctxp, C.c_int, _i64p, _dp, _dp, _dp]),
    "itg_malloc": (C.c_int, [_ctxp, C.c_size_t, C.POINTER(_vp)]),
    "itg_free": (C.c_int, [_ctxp, _vp]),
    "itg_memcpy_h2d": (C.c_int, [_ctxp, _vp, _vp, C.c_size_t]),
    "itg_memcpy_d2h": (C.c_int, [_ctxp, _vp, _vp, C.c_size_t]),
    "itg_memcpy_d2d": (C.c_int, [_ctxp, _vp, _vp, C.c_size_t]),
    "itg_fill": (C.c_int, [_ctxp, C.c_int, C.c_int64, C.c_double, _vp]),
    "itg_convert": (C.c_int, [_ctxp, C.c_int64, C.c_int, _vp, C.c_int, _vp]),
    "itg_randn": (C.c_int, [_ctxp, C.c_int, C.c_int64, C.c_uint64, C.c_uint64, _vp]),
    "itg_permute": (C.c_int, [_ctxp, C.c_int, C.c_int, _i64p, _i32p, C.c_double, _vp, _vp]),
    "itg_add_permuted": (C.c_int, [_ctxp, C.c_int, C.c_int, _i64p, _i32p, C.c_double, _vp,
                                   C.c_double, _vp]),
    "itg_scal": (C.c_int, [_ctxp, C.c_int, C.c_int64, C.c_double, _vp]),
    "itg_axpby": (C.c_int, [_ctxp, C.c_int, C.c_int64, C.c_double, _vp, C.c_double, _vp]),
    "itg_precond_diag": (C.c_int, [_ctxp, C.c_int, C.c_int64, _vp, _vp, C.c_double, C.c_double,
                                   _vp]),
    "itg_zscal": (C.c_int, [_ctxp, C.c_int64, C.c_double, C.c_double, _vp]),
    "itg_zaxpby": (C.c_int, [_ctxp, C.c_int64, C.c_double, C.c_double, _vp, C.c_double,
                             C.c_double, _vp]),
    "itg_conj": (C.c_int, [_ctxp, C.c_int, C.c_int64, _vp]),
    "itg_dot": (C.c_int, [_ctxp, C.c_int, C.c_int64, _vp, _vp, _dp]),
    "itg_nrm2": (C.c_int, [_ctxp, C.c_int, C.c_int64, _vp, _dp]),
    "itg_contract": (C.c_int, [_ctxp, C.c_int, C.c_int,
                               C.c_int, _i64p, _i32p, _vp,
                               C.c_int, _i64p, _i32p, _vp,
                               C.c_int, _i64p, _i32p, _vp,
                               C.c_double, C.c_double]),
    "itg_diag_contract": (C.c_int, [_ctxp, C.c_int, C.c_int, _i64p, _i32p, _vp, C.c_int64, _dp,
                                    C.c_int, _i64p, _i32p, _vp, C.c_int, _i64p, _i32p, _vp,
                                    C.c_double, C.c_double]),
    "itg_hadamard": (C.c_int, [_ctxp, C.c_int, C.c_int64, _vp, _vp, _vp]),
    "itg_embed": (C.c_int, [_ctxp, C.c_int, C.c_int, _i64p, _vp, _i64p, _i64p, _vp]),
    "itg_gemm": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_char, C.c_char, C.c_int64, C.c_int64,
                           C.c_int64, C.c_double, _vp, C.c_int64, _vp, C.c_int64, C.c_double,
                           _vp, C.c_int64]),
    "itg_truncate": (C.c_int, [_ctxp, _dp, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_int,
                               C.c_int, C.c_int, _i64p, _dp, _dp]),
    "itg_truncate_device": (C.c_int, [_ctxp, _vp, C.c_int64, C.c_int64, C.c_int64, C.c_double,
                                      C.c_int, C.c_int, C.c_int, _i64p, _dp, _dp]),
    "itg_eigh_trunc": (C.c_int, [_ctxp, C.c_int, C.c_int64, _vp, _vp, _vp, C.c_int64, C.c_double,
                                 C.c_int, C.c_int, _i64p, _dp, _dp]),
    "itg_svd_trunc": (C.c_int, [_ctxp, C.c_int, C.c_int64, C.c_int64, _vp, _vp, _vp, _vp,
                                C.c_int64, C.c_int64, C.c_double, C.c_int, C.c_int, _i64p, _dp,
                                _dp]),
    "itg_qr": (C.c_int, [_ctxp, C.c_int, C.c_int64, C.c_int64, _vp, _vp, _vp]),
    "itg_heff2_work_elems": (C.c_int64, [C.c_int64] * 6),
    "itg_heff2_apply": (C.c_int, [_ctxp, C.c_int, C.c_int] + [C.c_int64] * 6 + [_vp] * 7),
    "itg_heff2_noalloc_max_bytes": (C.c_int64, []),
    "itg_heff2_apply_noalloc": (C.c_int, [_ctxp, C.c_int, C.c_int] + [C.c_int64] * 6 +
                                [_vp] * 6),
    "itg_heff2_diag": (C.c_int, [_ctxp, C.c_int] + [C.c_int64] * 6 + [_vp] * 5),
    "itg_heff2_prep_elems": (C.c_int64, [C.c_int64] * 6),
    "itg_heff2_prepare": (C.c_int, [_ctxp, C.c_int, C.c_int] + [C.c_int64] * 6 + [_vp] * 5),
    "itg_heff2_apply_prepared": (C.c_int, [_ctxp, C.c_int, C.c_int] + [C.c_int64] * 6 +
                                 [_vp] * 6),
    "itg_davidson2_work_elems": (C.c_int64, [C.c_int64, C.c_int64, C.c_int64, C.c_int]),
    "itg_davidson2": (C.c_int, [_ctxp, C.c_int, C.c_int] + [C.c_int64] * 6 +
                      [_vp, _vp, _vp, _vp, C.c_int, C.c_double, _vp, _vp, _vp, _dp,
                       C.POINTER(C.c_int)]),
    "itg_davidson2_mlp_work_elems": (C.c_int64, [C.c_int64, C.c_int64, C.c_int64, C.c_int]),
    "itg_davidson2_mlp": (C.c_int, [_ctxp, C.c_int, C.c_int] + [C.c_int64] * 6 +
                          [_vp, _vp, _vp, _vp, C.c_int, C.c_double, _vp, _vp, C.c_int, C.c_int,
                           C.c_int, C.c_double, _vp, _vp, _vp, _dp, C.POINTER(C.c_int),
                           C.POINTER(C.c_int)]),
    "itg_mlp_param_count": (C.c_int64, [C.c_int, C.c_int, C.c_int]),
    "itg_mlp_forward": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                  _vp]),
    "itg_mlp_infer": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                  _vp]),
    "itg_mlp_backward": (C.c_int, [_ctxp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp,
                                   _vp, _vp, _vp, _dp]),
    "itg_mlp_adam": (C.c_int, [_ctxp, C.c_int64, _vp, _vp, _vp, _vp, C.c_int, C.c_float,
                               C.c_float, C.c_float, C.c_float]),
    "itg_mlp_features": (C.c_int, [_ctxp] + [C.c_int64] * 3 + [_vp, _vp, C.c_double, C.c_double]
                         + [C.c_int64] * 3 + [_vp]),
    "itg_mlp_gather": (C.c_int, [_ctxp] + [C.c_int64] * 3 + [_vp, C.c_double] + [C.c_int64] * 3
                       + [_vp]),
    "itg_mlp_scatter": (C.c_int, [_ctxp] + [C.c_int64] * 3 + [_vp, C.c_int64, C.c_int64, _vp]),
}

EXPORTED = tuple(_SIGS)

_lib = None


def lib():
    """The loaded libitg; raises if the library was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ITGError("libitg.so not found at %s: build it first (make -C "
                           "itensorsgpu.jl_amd or __graft_entry__.build())" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def load_only():
    """Load the shared object and resolve every exported symbol (no GPU use)."""
    L = C.CDLL(LIB_PATH)
    missing = [n for n in _SIGS if not hasattr(L, n)]
    return L, missing


class Context:
    """One libitg context on one device, bound to torch's current stream."""

    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise ITGError("no GPU visible: libitg needs an MI355X (gfx950)")
        self.device = device
        self.ptr = _ctxp()
        rc = lib().itg_init(device, C.byref(self.ptr))
        if rc != ITG_OK:
            raise ITGError("itg_init failed with status %d" % rc)
        torch.cuda.set_device(device)
        # a dedicated stream (default priority) shared by torch's allocator
        # and libitg, so events recorded on it time the library's kernels
        self.stream = torch.cuda.Stream(device=device)
        torch.cuda.set_stream(self.stream)
        self.bind_stream(self.stream.cuda_stream)

    def bind_stream(self, stream_handle):
        self.check(lib().itg_set_stream(self.ptr, _vp(stream_handle)))

    def check(self, rc):
        if rc == ITG_OK:
            return
        msg = lib().itg_last_error(self.ptr).decode(errors="replace")
        if rc == ITG_EDIM:
            raise DimensionMismatch(msg)
        if rc in (ITG_EARG, ITG_EDTYPE):
            raise ArgumentError(msg)
        raise ITGError("libitg status %d: %s" % (rc, msg))

    def info(self):
        cu, gfx, hbm = C.c_int(), C.c_int(), C.c_double()
        self.check(lib().itg_device_info(self.ptr, C.byref(cu), C.byref(gfx), C.byref(hbm)))
        return {"num_cu": cu.value, "gfx": hex(gfx.value), "hbm_gb": hbm.value}

    def sync(self):
        self.check(lib().itg_sync(self.ptr))

    def timing(self, kinds=()):
        """Start (clearing) event timing of the kernel classes `kinds`
        (K_* below); an empty set stops it (itg_timing)."""
        mask = 0
        for k in kinds:
            mask |= 1 << int(k)
        self.check(lib().itg_timing(self.ptr, mask))

    def timing_set(self, kinds=()):
        """Change the timed kernel classes, keeping what was recorded."""
        mask = 0
        for k in kinds:
            mask |= 1 << int(k)
        self.check(lib().itg_timing_set(self.ptr, mask))

    def timing_report(self, kind):
        """{launches, ms, flops, bytes} of one kernel class (K_* below)."""
        n, ms, fl, by = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        self.check(lib().itg_timing_report(self.ptr, int(kind), C.byref(n), C.byref(ms),
                                           C.byref(fl), C.byref(by)))
        return {"launches": n.value, "ms": ms.value, "flops": fl.value, "bytes": by.value}


_ctx = None


def ctx():
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


def i64arr(xs):
    a = (C.c_int64 * max(len(xs), 1))(*[int(x) for x in xs])
    return a


def i32arr(xs):
    return (C.c_int32 * max(len(xs), 1))(*[int(x) for x in xs])


def dtype_code(t):
    import torch
    if t.dtype == torch.float64:
        return F64
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.complex128:
        return C64
    raise ArgumentError("unsupported dtype %s" % (t.dtype,))


def ptr(t):
    return _vp(t.data_ptr())


def host_f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)
