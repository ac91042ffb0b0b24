"""Python binding of the gfx950 sparse-MTTKRP library (ctypes over include/mmcsf.h).

The library ``libmmcsf.so`` sits next to this package's parent directory
(``mm-csf_amd/libmmcsf.so``) and is built by ``mm-csf_amd/Makefile``.  Loading
fails loudly when it is missing: there is no CPU fallback for any GPU entry
point.  PyTorch is used only for device memory and streams.

Mirrors the reference driver's objects (src/util.h): ``Tensor`` (class Tensor,
util.h:80-108), ``Format`` (TiledTensor / HYBTensor, util.h:110-173),
``Options`` (util.h:189-236) and the MTTKRP launchers of src/mttkrp_gpu.h.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = [
    "lib", "MMCSFError", "Options", "Tensor", "Format", "mttkrp", "mttkrp_allmode",
    "TYPE_COO_CPU", "TYPE_COO", "TYPE_BCSF", "TYPE_BCSF_ANYMODE", "TYPE_HBCSF", "TYPE_MMCSF", "TYPE_SINGLECSF",
    "TYPE_ONECSF",
    "F32", "BF16",
    "parse_options", "to_bf16", "cpd_als", "gram",
]

TYPE_COO_CPU, TYPE_COO, TYPE_BCSF, TYPE_BCSF_ANYMODE, TYPE_HBCSF, TYPE_MMCSF = 1, 3, 8, 9, 10, 12
TYPE_SINGLECSF, TYPE_ONECSF = 13, 14
F32, BF16 = 0, 1

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libmmcsf.so")


class MMCSFError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mmcsf error {code}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [
        ("R", C.c_uint64), ("mode", C.c_uint64), ("impType", C.c_uint64),
        ("warpPerSlice", C.c_uint64), ("nTile", C.c_uint64), ("gridSize", C.c_uint64),
        ("TBsize", C.c_uint64), ("MIfbTh", C.c_uint64), ("fiberPerWarp", C.c_uint64),
        ("verbose", C.c_int), ("correctness", C.c_int), ("nBin", C.c_uint64),
        ("cpdIters", C.c_uint64), ("natOrdering", C.c_int), ("fbrThreashold", C.c_uint64),
        ("factorDtype", C.c_int), ("seed", C.c_uint64), ("modeTrees", C.c_int),
        ("mixedPart", C.c_int), ("inFileName", C.c_char * 4096), ("outFileName", C.c_char * 4096),
        ("cacheFileName", C.c_char * 4096),
        ("m0", C.c_char * 16), ("m1", C.c_char * 16), ("m2", C.c_char * 16),
    ]


class FormatInfo(C.Structure):
    _fields_ = [
        ("type", C.c_int), ("ndims", C.c_int), ("dims", C.c_uint64 * 4), ("totNnz", C.c_uint64),
        ("modeOrder", C.c_int * 4), ("nLevel", C.c_uint64 * 4), ("nFibers", C.c_uint64),
        ("droppedSlices", C.c_uint64), ("COOnnz", C.c_uint64), ("CSLnnz", C.c_uint64),
        ("HCSRnnz", C.c_uint64), ("redunMode", C.c_int), ("modeTrees", C.c_int),
        ("modeLevel", (C.c_uint64 * 4) * 4), ("deviceBytes", C.c_uint64), ("partNnz", C.c_uint64 * 4),
        ("mixedPart", C.c_int), ("modeTreeOrder", (C.c_int * 4) * 4),
    ]


class CpdStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("fit", C.c_double), ("fit_history", C.c_double * 64),
        ("ms_total", C.c_double), ("ms_mttkrp", C.c_double), ("lu_fallbacks", C.c_int),
    ]


_LIB = None
VP = C.c_void_p
U64P = C.POINTER(C.c_uint64)


def lib():
    """The loaded libmmcsf.so (raises when it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libmmcsf.so not built: {LIB_PATH} (run make -C mm-csf_amd)")
    L = C.CDLL(LIB_PATH)
    L.mmcsf_last_error.restype = C.c_char_p
    L.mmcsf_version.restype = C.c_char_p
    L.mmcsf_options_default.argtypes = [C.POINTER(Options)]
    L.mmcsf_options_parse.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.POINTER(Options)]
    L.mmcsf_tensor_load_tns.argtypes = [C.c_char_p, C.POINTER(VP)]
    L.mmcsf_tensor_load_bin.argtypes = [C.c_char_p, C.POINTER(VP)]
    L.mmcsf_tensor_save_bin.argtypes = [VP, C.c_char_p]
    L.mmcsf_tensor_save_tns.argtypes = [VP, C.c_char_p]
    L.mmcsf_tensor_from_host.argtypes = [C.c_int, U64P, C.c_uint64, C.POINTER(U64P),
                                         C.POINTER(C.c_float), C.POINTER(VP)]
    L.mmcsf_tensor_from_device.argtypes = [C.c_int, U64P, C.c_uint64, C.POINTER(VP), VP, VP,
                                           C.POINTER(VP)]
    L.mmcsf_tensor_generate.argtypes = [C.c_int, U64P, C.c_uint64, C.POINTER(C.c_int),
                                        C.POINTER(C.c_double), C.c_uint64, VP, C.POINTER(VP)]
    L.mmcsf_tensor_info.argtypes = [VP, C.POINTER(C.c_int), U64P, U64P]
    L.mmcsf_tensor_device_arrays.argtypes = [VP, C.POINTER(VP), C.POINTER(VP)]
    L.mmcsf_tensor_copy_to_host.argtypes = [VP, C.POINTER(U64P), C.POINTER(C.c_float)]
    L.mmcsf_tensor_copy_to_device.argtypes = [VP, C.POINTER(VP), VP, VP]
    L.mmcsf_tensor_free.argtypes = [VP]
    L.mmcsf_tensor_free.restype = None
    L.mmcsf_format_build.argtypes = [VP, C.c_int, C.POINTER(Options), VP, C.POINTER(VP)]
    L.mmcsf_format_info_get.argtypes = [VP, C.POINTER(FormatInfo)]
    L.mmcsf_format_export.argtypes = [VP, C.c_char_p, C.c_int, VP, C.c_uint64, U64P]
    L.mmcsf_format_save.argtypes = [VP, C.c_char_p]
    L.mmcsf_format_load.argtypes = [C.c_char_p, C.POINTER(VP)]
    L.mmcsf_format_is_cache.argtypes = [C.c_char_p]
    L.mmcsf_format_to_tensor.argtypes = [VP, C.POINTER(VP)]
    L.mmcsf_format_free.argtypes = [VP]
    L.mmcsf_format_free.restype = None
    L.mmcsf_mttkrp.argtypes = [VP, C.c_int, C.c_int, C.c_int, C.POINTER(VP), VP, VP]
    L.mmcsf_mttkrp_allmode.argtypes = [VP, C.c_int, C.c_int, C.POINTER(VP), C.POINTER(VP), VP,
                                       C.POINTER(C.c_float)]
    L.mmcsf_cpd_als.argtypes = [VP, C.c_int, C.c_int, C.c_double, C.c_int, C.POINTER(VP),
                                C.POINTER(C.c_double), C.POINTER(CpdStats), VP]
    L.mmcsf_cpd_als_noalloc.argtypes = L.mmcsf_cpd_als.argtypes
    L.mmcsf_gram.argtypes = [VP, C.c_uint64, C.c_int, VP, VP]
    L.mmcsf_mttkrp_coo_host.argtypes = [VP, C.c_int, C.c_int, C.POINTER(C.POINTER(C.c_float)),
                                        C.POINTER(C.c_float)]
    L.mmcsf_ttm.argtypes = [VP, C.c_int, C.c_int, VP, VP, VP]
    L.mmcsf_ttm_host.argtypes = [VP, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.mmcsf_cp_forward.argtypes = [VP, C.c_int, C.c_int, C.POINTER(VP), VP, VP, VP]
    L.mmcsf_cp_adjoint.argtypes = [VP, C.c_int, C.c_int, C.c_int, C.POINTER(VP), VP, C.c_float, VP, VP]
    L.mmcsf_cp_loss_grad.argtypes = [VP, C.c_int, C.c_int, C.POINTER(VP), C.c_float, C.POINTER(VP), VP, VP]
    _LIB = L
    return L


def _check(rc):
    if rc != 0:
        raise MMCSFError(rc, lib().mmcsf_last_error().decode(errors="replace"))


def parse_options(argv):
    """parse_cmd_options (util.h:3341-3442) -> Options."""
    o = Options()
    args = [b"mttkrp"] + [a.encode() if isinstance(a, str) else a for a in argv]
    arr = (C.c_char_p * len(args))(*args)
    _check(lib().mmcsf_options_parse(len(args), arr, C.byref(o)))
    return o


def default_options():
    o = Options()
    _check(lib().mmcsf_options_default(C.byref(o)))
    return o


def _stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return VP(s.cuda_stream)


class Tensor:
    """Device COO tensor (class Tensor, util.h:80-108)."""

    def __init__(self, handle):
        self._h = VP(handle) if not isinstance(handle, VP) else handle

    @classmethod
    def from_numpy(cls, dims, inds, vals):
        dims = np.ascontiguousarray(dims, dtype=np.uint64)
        inds = [np.ascontiguousarray(a, dtype=np.uint64) for a in inds]
        vals = np.ascontiguousarray(vals, dtype=np.float32)
        nnz = len(vals)
        ptrs = (U64P * len(inds))(*[a.ctypes.data_as(U64P) for a in inds])
        h = VP()
        _check(lib().mmcsf_tensor_from_host(len(dims), dims.ctypes.data_as(U64P), nnz, ptrs,
                                            vals.ctypes.data_as(C.POINTER(C.c_float)), C.byref(h)))
        return cls(h)

    @classmethod
    def from_torch(cls, dims, inds, vals, stream=None):
        """Device COO from device tensors (int32 coordinates per mode, float32
        values); the data is copied into the handle."""
        import torch
        dims = np.ascontiguousarray(dims, dtype=np.uint64)
        if len(inds) != len(dims):
            raise ValueError("one coordinate tensor per mode")
        for a in inds:
            if not a.is_cuda or a.dtype != torch.int32 or not a.is_contiguous() or a.numel() != vals.numel():
                raise ValueError("coordinates must be contiguous int32 device tensors, one entry per nonzero")
        if not vals.is_cuda or vals.dtype != torch.float32 or not vals.is_contiguous():
            raise ValueError("values must be a contiguous float32 device tensor")
        ptrs = (VP * len(inds))(*[VP(a.data_ptr()) for a in inds])
        h = VP()
        _check(lib().mmcsf_tensor_from_device(len(dims), dims.ctypes.data_as(U64P), vals.numel(), ptrs,
                                              VP(vals.data_ptr()), _stream_ptr(stream), C.byref(h)))
        return cls(h)

    @classmethod
    def load_bin(cls, path):
        """Binary COO file written by save_bin."""
        h = VP()
        _check(lib().mmcsf_tensor_load_bin(str(path).encode(), C.byref(h)))
        return cls(h)

    def save_bin(self, path):
        _check(lib().mmcsf_tensor_save_bin(self._h, str(path).encode()))

    def save_tns(self, path):
        _check(lib().mmcsf_tensor_save_tns(self._h, str(path).encode()))

    @classmethod
    def load_tns(cls, path):
        h = VP()
        _check(lib().mmcsf_tensor_load_tns(str(path).encode(), C.byref(h)))
        return cls(h)

    @classmethod
    def generate(cls, dims, nnz, law, alpha, seed, stream=None):
        dims = np.ascontiguousarray(dims, dtype=np.uint64)
        n = len(dims)
        lw = (C.c_int * n)(*[int(x) for x in law])
        al = (C.c_double * n)(*[float(x) for x in alpha])
        h = VP()
        _check(lib().mmcsf_tensor_generate(n, dims.ctypes.data_as(U64P), int(nnz), lw, al, int(seed),
                                           _stream_ptr(stream), C.byref(h)))
        return cls(h)

    def info(self):
        nd = C.c_int()
        dims = (C.c_uint64 * 4)()
        nnz = C.c_uint64()
        _check(lib().mmcsf_tensor_info(self._h, C.byref(nd), dims, C.byref(nnz)))
        return nd.value, [int(dims[i]) for i in range(nd.value)], nnz.value

    def to_numpy(self):
        nd, dims, nnz = self.info()
        inds = [np.zeros(nnz, dtype=np.uint64) for _ in range(nd)]
        vals = np.zeros(nnz, dtype=np.float32)
        ptrs = (U64P * nd)(*[a.ctypes.data_as(U64P) for a in inds])
        _check(lib().mmcsf_tensor_copy_to_host(self._h, ptrs, vals.ctypes.data_as(C.POINTER(C.c_float))))
        return dims, [a.astype(np.int64) for a in inds], vals

    def to_torch(self, device=None, stream=None):
        """(inds, vals) as device torch tensors: int32 coordinates per mode, float32 values."""
        import torch
        nd, dims, nnz = self.info()
        dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        inds = [torch.empty(nnz, dtype=torch.int32, device=dev) for _ in range(nd)]
        vals = torch.empty(nnz, dtype=torch.float32, device=dev)
        ip = (VP * nd)(*[VP(a.data_ptr()) for a in inds])
        _check(lib().mmcsf_tensor_copy_to_device(self._h, ip, VP(vals.data_ptr()), _stream_ptr(stream)))
        return inds, vals

    def device_arrays(self):
        nd, _, _ = self.info()
        ip = (VP * 4)()
        vp = VP()
        _check(lib().mmcsf_tensor_device_arrays(self._h, ip, C.byref(vp)))
        return [ip[i] for i in range(nd)], vp.value

    def free(self):
        if self._h is not None and self._h.value:
            lib().mmcsf_tensor_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Format:
    """Device format of one driver type (-t 1/3/8/10/12)."""

    def __init__(self, tensor, fmt_type, mode=0, fbr_threshold=99999999, stream=None, opts=None, mode_trees=True,
                 mixed=False, fb_th=1):
        o = opts if opts is not None else default_options()
        o.mode = int(mode)
        o.modeTrees = 1 if mode_trees else 0
        if mixed:
            o.mixedPart = 1
            o.MIfbTh = int(fb_th)
        o.fbrThreashold = int(fbr_threshold) & 0xFFFFFFFFFFFFFFFF
        h = VP()
        _check(lib().mmcsf_format_build(tensor._h, int(fmt_type), C.byref(o), _stream_ptr(stream),
                                        C.byref(h)))
        self._h = h
        self.type = fmt_type

    def info(self):
        i = FormatInfo()
        _check(lib().mmcsf_format_info_get(self._h, C.byref(i)))
        return {
            "type": i.type, "ndims": i.ndims, "dims": list(i.dims[: i.ndims]), "totNnz": i.totNnz,
            "modeOrder": list(i.modeOrder[: i.ndims]), "nLevel": list(i.nLevel[: i.ndims]),
            "nFibers": i.nFibers, "droppedSlices": i.droppedSlices, "COOnnz": i.COOnnz,
            "CSLnnz": i.CSLnnz, "HCSRnnz": i.HCSRnnz, "redunMode": i.redunMode,
            "modeTrees": i.modeTrees,
            "modeLevel": [list(i.modeLevel[m][: i.ndims]) for m in range(i.ndims)],
            "deviceBytes": i.deviceBytes, "partNnz": list(i.partNnz[: i.ndims]), "mixedPart": i.mixedPart,
            "modeTreeOrder": [list(i.modeTreeOrder[m][: i.ndims]) for m in range(i.ndims)],
        }

    def export(self, name, k=0):
        """A reference member as a numpy array (uint64, or float32 for values)."""
        n = C.c_uint64()
        _check(lib().mmcsf_format_export(self._h, name.encode(), int(k), None, 0, C.byref(n)))
        is_float = name in ("vals", "COOvals", "CSLvals", "P.vals")
        out = np.zeros(max(n.value, 1), dtype=np.float32 if is_float else np.uint64)
        _check(lib().mmcsf_format_export(self._h, name.encode(), int(k), out.ctypes.data_as(VP),
                                         n.value, C.byref(n)))
        return out[: n.value]

    def save(self, path):
        """Binary cache of the built format (mmcsf_format_save)."""
        _check(lib().mmcsf_format_save(self._h, str(path).encode()))

    @classmethod
    def load(cls, path):
        """A format from its binary cache (mmcsf_format_load), without building."""
        obj = cls.__new__(cls)
        h = VP()
        _check(lib().mmcsf_format_load(str(path).encode(), C.byref(h)))
        obj._h = h
        obj.type = obj.info()["type"]
        return obj

    def to_tensor(self):
        """The nonzeros the format stores, as a Tensor (mmcsf_format_to_tensor)."""
        h = VP()
        _check(lib().mmcsf_format_to_tensor(self._h, C.byref(h)))
        return Tensor(h)

    def free(self):
        if self._h is not None and self._h.value:
            lib().mmcsf_format_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_bf16(t):
    """float32 torch tensor -> bfloat16 bit patterns (round to nearest even)."""
    import torch
    return t.to(torch.bfloat16).contiguous()


def _factor_ptrs(factors):
    return (VP * len(factors))(*[VP(f.data_ptr()) if f is not None else VP(0) for f in factors])


def _dtype_of(factors):
    import torch
    dts = {f.dtype for f in factors if f is not None}
    if dts == {torch.float32}:
        return F32
    if dts == {torch.bfloat16}:
        return BF16
    raise ValueError("factors must all be float32 or all bfloat16")


def _check_factors(dims, factors, R, skip=None):
    if len(factors) != len(dims):
        raise ValueError(f"expected {len(dims)} factor matrices, got {len(factors)}")
    for m, f in enumerate(factors):
        if m == skip and f is None:
            continue
        if f is None or not f.is_cuda or not f.is_contiguous() or tuple(f.shape) != (dims[m], R):
            raise ValueError(f"factor {m} must be a contiguous device tensor of shape ({dims[m]}, {R})")


def mttkrp(fmt, mode, factors, out=None, stream=None):
    """MTTKRP of ``mode``; factors are device torch tensors (float32 or bfloat16)."""
    import torch
    R = next(f for f in factors if f is not None).shape[1]
    dims = fmt.info()["dims"]
    if not 0 <= int(mode) < len(dims):
        raise MMCSFError(-1, "Mode cannot be larger than tensor dimension.")
    rows = dims[mode]
    _check_factors(dims, factors, R, skip=mode)
    if out is None:
        dev = next(f for f in factors if f is not None).device
        out = torch.empty((rows, R), dtype=torch.float32, device=dev)
    if tuple(out.shape) != (rows, R) or out.dtype != torch.float32 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float32 tensor of shape ({rows}, {R})")
    _check(lib().mmcsf_mttkrp(fmt._h, int(mode), int(R), _dtype_of(factors), _factor_ptrs(factors),
                              VP(out.data_ptr()), _stream_ptr(stream)))
    return out


def mttkrp_allmode(fmt, factors, outs=None, stream=None, timed=False):
    """All-mode MTTKRP of an MM-CSF or one-tree (types 9 / 13 / 14) format;
    returns (outs, ms_per_mode or None)."""
    import torch
    R = factors[0].shape[1]
    info = fmt.info()
    _check_factors(info["dims"], factors, R)
    if outs is None:
        outs = [torch.empty((d, R), dtype=torch.float32, device=factors[0].device) for d in info["dims"]]
    for o, d in zip(outs, info["dims"]):
        if tuple(o.shape) != (d, R) or o.dtype != torch.float32 or not o.is_contiguous():
            raise ValueError(f"outputs must be contiguous float32 tensors of shape (dims[m], {R})")
    op = (VP * len(outs))(*[VP(o.data_ptr()) for o in outs])
    ms = (C.c_float * len(outs))() if timed else None
    _check(lib().mmcsf_mttkrp_allmode(fmt._h, int(R), _dtype_of(factors), _factor_ptrs(factors), op,
                                      _stream_ptr(stream), ms))
    return outs, ([float(v) for v in ms] if timed else None)


def ttm(fmt, U_leaf, out=None, stream=None):
    """Tensor-times-matrix on the leaf mode (TTM_GPU, ttm_gpu.cu:140-259) of a
    B-CSF-family format: one float32 row of R per leaf fiber."""
    import torch
    info = fmt.info()
    R = U_leaf.shape[1]
    leaf_mode = info["modeOrder"][-1]
    if (not U_leaf.is_cuda or not U_leaf.is_contiguous()
            or tuple(U_leaf.shape) != (info["dims"][leaf_mode], R)):
        raise ValueError(f"the leaf factor must be a contiguous device tensor of shape "
                         f"({info['dims'][leaf_mode]}, {R})")
    rows = info["nLevel"][info["ndims"] - 2]
    if out is None:
        out = torch.empty((rows, R), dtype=torch.float32, device=U_leaf.device)
    _check(lib().mmcsf_ttm(fmt._h, int(R), _dtype_of([U_leaf]), VP(U_leaf.data_ptr()), VP(out.data_ptr()),
                           _stream_ptr(stream)))
    return out


def ttm_host(fmt, U_leaf):
    """TTM_CPU (ttm_cpu.cpp:60-86) on the host: numpy float32 in and out."""
    info = fmt.info()
    U = np.ascontiguousarray(U_leaf, dtype=np.float32)
    R = U.shape[1]
    Y = np.zeros((info["nLevel"][info["ndims"] - 2], R), dtype=np.float32)
    _check(lib().mmcsf_ttm_host(fmt._h, int(R), U.ctypes.data_as(C.POINTER(C.c_float)),
                                Y.ctypes.data_as(C.POINTER(C.c_float))))
    return Y


def gram(U, stream=None):
    """U^T U (float64) of a device float32 matrix, via MFMA."""
    import torch
    assert U.is_cuda and U.dtype == torch.float32 and U.is_contiguous() and U.dim() == 2
    R = U.shape[1]
    G = torch.empty((R, R), dtype=torch.float64, device=U.device)
    _check(lib().mmcsf_gram(VP(U.data_ptr()), U.shape[0], R, VP(G.data_ptr()), _stream_ptr(stream)))
    return G


def cpd_als(fmt, factors, iters=10, tol=1e-5, mttkrp_bf16=False, stream=None, no_alloc=False):
    """CP-ALS (cpd(), cpd_cpu.cu:484-691) on an MM-CSF format.

    ``factors``: device float32 tensors (dims[m], R), updated in place to the
    final normalised factors.  Returns (lambda, stats dict).  ``no_alloc``
    runs mmcsf_cpd_als_noalloc, whose work space is the library's fixed
    device scratch (MMCSF_CPD_NOALLOC_SCRATCH_BYTES per GPU)."""
    info = fmt.info()
    R = factors[0].shape[1]
    _check_factors(info["dims"], factors, R)
    fp = (VP * len(factors))(*[VP(f.data_ptr()) for f in factors])
    lam = (C.c_double * R)()
    st = CpdStats()
    fn = lib().mmcsf_cpd_als_noalloc if no_alloc else lib().mmcsf_cpd_als
    _check(fn(fmt._h, int(R), int(iters), float(tol), BF16 if mttkrp_bf16 else F32, fp, lam, C.byref(st),
              _stream_ptr(stream)))
    n = st.iterations
    return np.array(lam[:], dtype=np.float64), {
        "iterations": n, "fit": st.fit, "fit_history": list(st.fit_history[: min(n, 64)]),
        "ms_total": st.ms_total, "ms_mttkrp": st.ms_mttkrp, "lu_fallbacks": st.lu_fallbacks,
    }


# ---- CP completion -------------------------------------------------------------

def _cp_factors(fmt, factors):
    info = fmt.info()
    R = factors[0].shape[1]
    _check_factors(info["dims"], factors, R)
    return info, R, _dtype_of(factors)


def cp_forward(fmt, factors, residual=False, stream=None):
    """Loss 1/2 sum_x e_x^2 of the CP model on the observed entries, with
    e_x = sum_r prod_n U_n[i_n(x)][r] - v_x.  Returns (loss, residual): loss is
    a device float64 scalar tensor, residual (float32, tensor order) or None."""
    import torch
    info, R, dt = _cp_factors(fmt, factors)
    dev = factors[0].device
    loss = torch.empty((), dtype=torch.float64, device=dev)
    res = torch.empty(info["totNnz"], dtype=torch.float32, device=dev) if residual else None
    _check(lib().mmcsf_cp_forward(fmt._h, int(R), dt, _factor_ptrs(factors),
                                  VP(res.data_ptr()) if res is not None else VP(0), VP(loss.data_ptr()),
                                  _stream_ptr(stream)))
    return loss, res


def cp_adjoint(fmt, mode, factors, weights=None, alpha=1.0, out=None, stream=None):
    """alpha * sum_x w_x * d(model_x)/dU_mode as a (dims[mode], R) float32 tensor.
    ``weights`` (float32, tensor order) or None for w = the residual, which
    makes the result alpha * dL/dU_mode."""
    import torch
    info, R, dt = _cp_factors(fmt, factors)
    rows = info["dims"][mode]
    if out is None:
        out = torch.empty((rows, R), dtype=torch.float32, device=factors[0].device)
    if weights is not None:
        if (not weights.is_cuda or weights.dtype != torch.float32 or not weights.is_contiguous()
                or weights.numel() != info["totNnz"]):
            raise ValueError("weights must be a contiguous float32 device tensor with one entry per nonzero")
    _check(lib().mmcsf_cp_adjoint(fmt._h, int(mode), int(R), dt, _factor_ptrs(factors),
                                  VP(weights.data_ptr()) if weights is not None else VP(0), float(alpha),
                                  VP(out.data_ptr()), _stream_ptr(stream)))
    return out


def cp_loss_grad(fmt, factors, alpha=1.0, need=None, stream=None):
    """(loss, grads): loss as in cp_forward, grads[m] = alpha * dL/dU_m (float32)
    for every m with need[m] (default: all), else None; one pass per mode."""
    import torch
    info, R, dt = _cp_factors(fmt, factors)
    dev = factors[0].device
    need = need if need is not None else [True] * len(factors)
    grads = [torch.empty((d, R), dtype=torch.float32, device=dev) if n else None
             for d, n in zip(info["dims"], need)]
    loss = torch.empty((), dtype=torch.float64, device=dev)
    gp = (VP * len(grads))(*[VP(g.data_ptr()) if g is not None else VP(0) for g in grads])
    _check(lib().mmcsf_cp_loss_grad(fmt._h, int(R), dt, _factor_ptrs(factors), float(alpha), gp,
                                    VP(loss.data_ptr()), _stream_ptr(stream)))
    return loss, grads


def _autograd_function():
    import torch

    class CPCompletionLoss(torch.autograd.Function):
        """L(U_0..U_{N-1}) = 1/2 sum_x (sum_r prod_n U_n[i_n(x)][r] - v_x)^2 on the
        observed entries of an MM-CSF format.  The forward pass computes the loss
        and, when a gradient is needed, every factor's gradient in the same
        per-mode traversals; backward scales them by the incoming gradient."""

        @staticmethod
        def forward(ctx, fmt, *factors):
            need = list(ctx.needs_input_grad[1:])
            if any(need):
                loss, grads = cp_loss_grad(fmt, list(factors), need=need)
                ctx.save_for_backward(*[g if g is not None else torch.empty(0) for g in grads])
                ctx.need = need
                ctx.dtypes = [f.dtype for f in factors]
            else:
                loss, _ = cp_forward(fmt, list(factors))
            return loss

        @staticmethod
        def backward(ctx, gout):
            grads = ctx.saved_tensors
            out = [None]
            for u_need, g, dt in zip(ctx.need, grads, ctx.dtypes):
                out.append((g * gout.to(g.dtype)).to(dt) if u_need else None)
            return tuple(out)

    return CPCompletionLoss


_CP_FN = None


def cp_completion_loss(fmt, factors):
    """Differentiable CP-completion loss (float64 device scalar).  Factors are
    device tensors (all float32 or all bfloat16); their gradients come back
    in float32 for float32 factors and in bfloat16 for bfloat16 factors."""
    global _CP_FN
    if _CP_FN is None:
        _CP_FN = _autograd_function()
    return _CP_FN.apply(fmt, *factors)
