"""CPU oracle for the MM-CSF MTTKRP path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package; it is the checker, never the thing measured or
shipped.  The product (``mm-csf_amd/``) does not import it.

The arithmetic lives in ``mmcsf_oracle.c`` (a plain-C restatement of the
reference's builders and COO CPU MTTKRP, built into ``liboracle.so`` by
``oracle/Makefile``).  This module wraps it with ctypes and restates the
driver flow of the reference's ``mttkrp`` binary for each format type
(src/mttkrp.cu:71-472):

* type 1 / 3  sort_COOtensor (mttkrp.cu:79) -> COO
* type 8      create_HCSR -> make_KTiling -> create_TiledHCSR -> make_TiledBin
              -> create_fbrLikeSlcInds (mttkrp.cu:153-219)
* type 10     create_HCSR -> create_HYB[_4D] -> make_HybBin (mttkrp.cu:138-150)
* type 12     mm_partition_reuseBased (every nonzero -> partition 2, util.h:2517)
              -> populate_paritions -> sort_MI_CSF -> create_TiledHCSR
              (fbrS split at 128 for 4-D) -> create_fbrLikeSlcInds
              (mttkrp.cu:289-395)

Float64 CP-ALS and CP-completion references live in ``oracle/dense.py``.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

U64P = C.POINTER(C.c_uint64)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_order_tensormode.argtypes = [C.c_int, U64P, C.c_int, C.c_int, C.POINTER(C.c_int)]
        L.orc_sort_perm.argtypes = [C.c_int, C.c_uint64, C.POINTER(U64P), C.POINTER(C.c_int), C.c_int, U64P]
        L.orc_build_csf.argtypes = [C.c_int, C.c_uint64, C.POINTER(U64P), C.c_uint64, C.c_uint64,
                                    C.POINTER(U64P), C.POINTER(U64P), U64P, U64P]
        L.orc_build_csf.restype = C.c_int64
        L.orc_fbr_like_slc_inds.argtypes = [C.c_uint64, U64P, U64P, U64P]
        L.orc_tiled_bin.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.POINTER(U64P),
                                    C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.orc_tiled_bin.restype = C.c_int64
        L.orc_hyb_bin.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.POINTER(U64P), C.POINTER(C.c_int32)]
        L.orc_hyb_bin.restype = C.c_int64
        L.orc_csl_bin.argtypes = [C.c_uint64, C.c_uint64, U64P, C.POINTER(C.c_int32)]
        L.orc_csl_bin.restype = C.c_int64
        F32P = C.POINTER(C.c_float)
        L.orc_create_hyb.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.POINTER(U64P), C.POINTER(U64P),
                                     U64P, F32P, C.POINTER(U64P), F32P,
                                     U64P, U64P, U64P, U64P, F32P,
                                     C.POINTER(U64P), C.POINTER(U64P), U64P, U64P, U64P, F32P, U64P]
        L.orc_mttkrp_coo_f32.argtypes = [C.c_int, C.c_uint64, C.POINTER(U64P), F32P, C.c_int,
                                         C.c_uint64, C.POINTER(F32P), F32P]
        L.orc_mttkrp_coo_f64.argtypes = [C.c_int, C.c_uint64, C.POINTER(U64P), F32P, C.c_int,
                                         C.c_uint64, C.POINTER(F32P), C.POINTER(C.c_double)]
        U32PP = C.POINTER(C.POINTER(C.c_uint32))
        D64P = C.POINTER(C.c_double)
        L.orc_mttkrp_f64_mt.argtypes = [C.c_int, C.c_uint64, U32PP, F32P, D64P, C.c_int, C.c_uint64,
                                        C.c_uint64, C.POINTER(D64P), D64P, C.c_int]
        L.orc_mttkrp_f64_mt.restype = C.c_int
        L.orc_cp_resid_f64_mt.argtypes = [C.c_int, C.c_uint64, U32PP, F32P, C.c_uint64, C.POINTER(D64P),
                                          D64P, D64P, C.c_int]
        L.orc_cp_resid_f64_mt.restype = C.c_int
        L.orc_mm_partition.argtypes = [C.c_int, C.c_uint64, C.POINTER(U64P), U64P, C.c_int, C.c_int, C.c_int,
                                       C.POINTER(C.POINTER(C.c_int)), C.POINTER(U64P), C.POINTER(U64P),
                                       C.POINTER(U64P), U64P]
        L.orc_mm_partition.restype = C.c_int
        _LIB = L
    return _LIB


def _u64(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def _p64(a):
    return a.ctypes.data_as(U64P)


def _parr(arrs, ptype):
    t = (ptype * len(arrs))(*[a.ctypes.data_as(ptype) for a in arrs])
    return t


# --------------------------------------------------------------------------
# primitives
# --------------------------------------------------------------------------
def order_tensormode(dims, mode, nat=False):
    """util.h:255-311."""
    d = _u64(dims)
    out = (C.c_int * len(dims))()
    lib().orc_order_tensormode(len(dims), _p64(d), int(mode), int(bool(nat)), out)
    return [int(v) for v in out]


def sort_perm(inds, order, nkeys=None):
    """Full-key lexicographic sort permutation (stable on input position)."""
    inds = [_u64(a) for a in inds]
    n = len(inds)
    nnz = len(inds[0])
    nkeys = n if nkeys is None else nkeys
    perm = np.zeros(nnz, dtype=np.uint64)
    o = (C.c_int * n)(*order)
    lib().orc_sort_perm(n, nnz, _parr(inds, U64P), o, nkeys, _p64(perm))
    return perm.astype(np.int64)


def build_csf(keys, fbr_threshold, fbrs_threshold):
    """create_TiledHCSR (util.h:1320-1436) on sorted per-level keys."""
    keys = [_u64(k) for k in keys]
    ndims = len(keys)
    nnz = len(keys[0])
    L = ndims - 1
    ptr = [np.zeros(nnz + 2, dtype=np.uint64) for _ in range(L)]
    idx = [np.zeros(nnz + 2, dtype=np.uint64) for _ in range(L)]
    plen = np.zeros(L, dtype=np.uint64)
    ilen = np.zeros(L, dtype=np.uint64)
    nf = lib().orc_build_csf(ndims, nnz, _parr(keys, U64P), C.c_uint64(fbr_threshold),
                             C.c_uint64(fbrs_threshold), _parr(ptr, U64P), _parr(idx, U64P),
                             _p64(plen), _p64(ilen))
    fbrPtr = [ptr[l][: int(plen[l])].copy() for l in range(L)]
    fbrIdx = [idx[l][: int(ilen[l])].copy() for l in range(L)]
    return fbrPtr, fbrIdx, int(nf)


def fbr_like_slc_inds(fbrPtr, fbrIdx, nFibers):
    out = np.zeros(max(nFibers, 1), dtype=np.uint64)
    nslc = len(fbrIdx[0])
    p0 = _u64(fbrPtr[0])
    i0 = _u64(fbrIdx[0])
    lib().orc_fbr_like_slc_inds(nslc, _p64(p0), _p64(i0), _p64(out))
    return out[:nFibers]


def tiled_bin(ndims, totnnz, fbrPtr):
    nslc = len(fbrPtr[0]) - 1
    ptr = [_u64(p) for p in fbrPtr]
    bins = np.zeros(max(nslc, 1), dtype=np.int32)
    w = np.zeros(max(nslc, 1), dtype=np.int64)
    dropped = lib().orc_tiled_bin(ndims, totnnz, nslc, _parr(ptr, U64P),
                                  bins.ctypes.data_as(C.POINTER(C.c_int32)),
                                  w.ctypes.data_as(C.POINTER(C.c_int64)))
    return bins[:nslc], w[:nslc], int(dropped)


def hyb_bin(ndims, totnnz, fbrPtr):
    nslc = len(fbrPtr[0]) - 1
    ptr = [_u64(p) for p in fbrPtr]
    bins = np.zeros(max(nslc, 1), dtype=np.int32)
    dropped = lib().orc_hyb_bin(ndims, totnnz, nslc, _parr(ptr, U64P),
                                bins.ctypes.data_as(C.POINTER(C.c_int32)))
    return bins[:nslc], int(dropped)


def csl_bin(totnnz, cslPtr):
    n = len(cslPtr) - 1
    p = _u64(cslPtr)
    bins = np.zeros(max(n, 1), dtype=np.int32)
    dropped = lib().orc_csl_bin(totnnz, n, _p64(p), bins.ctypes.data_as(C.POINTER(C.c_int32)))
    return bins[:n], int(dropped)


def mapper_bins(bin_of, nbin=10):
    """slcMapperBin[b] = ascending slice ordinals with bin b (push order)."""
    return [np.nonzero(bin_of == b)[0].astype(np.uint64) for b in range(nbin)]


def create_hyb(ndims, fbrPtr, fbrIdx, leaf, vals):
    nnz = len(vals)
    L = ndims - 1
    nslc = len(fbrIdx[0])
    n2 = nnz + 2
    ptr = [_u64(p) for p in fbrPtr]
    idx = [_u64(p) for p in fbrIdx]
    leaf = _u64(leaf)
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    F32P = C.POINTER(C.c_float)
    coo_ind = [np.zeros(n2, np.uint64) for _ in range(ndims)]
    coo_val = np.zeros(n2, np.float32)
    csl_ptr = np.zeros(n2, np.uint64)
    csl_idx = np.zeros(n2, np.uint64)
    csl_i1 = np.zeros(n2, np.uint64)
    csl_i2 = np.zeros(n2, np.uint64)
    csl_val = np.zeros(n2, np.float32)
    hptr = [np.zeros(n2, np.uint64) for _ in range(L)]
    hidx = [np.zeros(n2, np.uint64) for _ in range(L)]
    hplen = np.zeros(L, np.uint64)
    hilen = np.zeros(L, np.uint64)
    hleaf = np.zeros(n2, np.uint64)
    hval = np.zeros(n2, np.float32)
    counts = np.zeros(5, np.uint64)
    lib().orc_create_hyb(ndims, nnz, nslc, _parr(ptr, U64P), _parr(idx, U64P), _p64(leaf),
                         vals.ctypes.data_as(F32P), _parr(coo_ind, U64P), coo_val.ctypes.data_as(F32P),
                         _p64(csl_ptr), _p64(csl_idx), _p64(csl_i1), _p64(csl_i2),
                         csl_val.ctypes.data_as(F32P), _parr(hptr, U64P), _parr(hidx, U64P),
                         _p64(hplen), _p64(hilen), _p64(hleaf), hval.ctypes.data_as(F32P), _p64(counts))
    coo_n, csl_n, h_n, ncsl, nfib = [int(c) for c in counts]
    return {
        "COOnnz": coo_n, "CSLnnz": csl_n, "HCSRnnz": h_n, "nFibers": nfib,
        "COOinds": [c[:coo_n].copy() for c in coo_ind],          # by level
        "COOvals": coo_val[:coo_n].copy(),
        "CSLslicePtr": csl_ptr[: ncsl + 1].copy(),
        "CSLsliceIdx": csl_idx[:ncsl].copy(),
        "CSLinds1": csl_i1[:csl_n].copy(), "CSLinds2": csl_i2[:csl_n].copy(),
        "CSLvals": csl_val[:csl_n].copy(),
        "fbrPtr": [hptr[l][: int(hplen[l])].copy() for l in range(L)],
        "fbrIdx": [hidx[l][: int(hilen[l])].copy() for l in range(L)],
        "leaf": hleaf[:h_n].copy(), "vals": hval[:h_n].copy(),
    }


def mttkrp_coo_f32(inds, vals, mode, U, out):
    """MTTKRP_COO_CPU[_4D] (mttkrp_cpu.cpp:60-113): accumulates into ``out``."""
    inds = [_u64(a) for a in inds]
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    U = [np.ascontiguousarray(u, dtype=np.float32) for u in U]
    R = U[0].shape[1]
    assert out.dtype == np.float32 and out.flags.c_contiguous
    F32P = C.POINTER(C.c_float)
    lib().orc_mttkrp_coo_f32(len(inds), len(vals), _parr(inds, U64P), vals.ctypes.data_as(F32P),
                             int(mode), R, _parr(U, F32P), out.ctypes.data_as(F32P))
    return out


def mttkrp_coo_f64(inds, vals, mode, U, dims=None):
    """Float64 MTTKRP oracle from float64 copies of the float32 inputs."""
    inds = [_u64(a) for a in inds]
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    U = [np.ascontiguousarray(u, dtype=np.float32) for u in U]
    R = U[0].shape[1]
    out = np.zeros((U[mode].shape[0], R), dtype=np.float64)
    F32P = C.POINTER(C.c_float)
    lib().orc_mttkrp_coo_f64(len(inds), len(vals), _parr(inds, U64P), vals.ctypes.data_as(F32P),
                             int(mode), R, _parr(U, F32P), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def default_threads():
    """Host threads for the split oracle: at most 16 (the CPU share of one GPU)."""
    return max(1, min(16, os.cpu_count() or 1))


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def mttkrp_f64_mt(inds, vals, mode, U, rows=None, threads=None):
    """Float64 MTTKRP split over host threads by output-row ranges; equal bit
    for bit to ``mttkrp_coo_f64`` (each row summed in stored order).  ``vals``
    may be float32 (tensor values) or float64 (e.g. residuals); ``U`` are
    converted to float64 copies."""
    inds = [_u32(a) for a in inds]
    nnz = len(inds[0])
    U = [np.ascontiguousarray(u, dtype=np.float64) for u in U]
    R = U[0].shape[1]
    rows = U[mode].shape[0] if rows is None else rows
    out = np.empty((rows, R), dtype=np.float64)
    D64P = C.POINTER(C.c_double)
    F32P = C.POINTER(C.c_float)
    vals = np.ascontiguousarray(vals)
    vf = vals.ctypes.data_as(F32P) if vals.dtype == np.float32 else None
    if vals.dtype != np.float32:
        vals = np.ascontiguousarray(vals, dtype=np.float64)
    vd = vals.ctypes.data_as(D64P) if vals.dtype == np.float64 else None
    rc = lib().orc_mttkrp_f64_mt(len(inds), nnz, _parr(inds, C.POINTER(C.c_uint32)), vf, vd, int(mode), rows, R,
                                 _parr(U, D64P), out.ctypes.data_as(D64P), int(threads or default_threads()))
    if rc != 0:
        raise RuntimeError("orc_mttkrp_f64_mt failed")
    return out


def cp_resid_f64_mt(inds, vals, U, threads=None):
    """(1/2 sum e^2, e) with e_x = sum_r prod_n U_n[i_n(x)][r] - v_x in float64."""
    inds = [_u32(a) for a in inds]
    nnz = len(inds[0])
    U = [np.ascontiguousarray(u, dtype=np.float64) for u in U]
    vals = np.ascontiguousarray(vals, dtype=np.float32)
    e = np.empty(nnz, dtype=np.float64)
    loss = C.c_double(0.0)
    D64P = C.POINTER(C.c_double)
    rc = lib().orc_cp_resid_f64_mt(len(inds), nnz, _parr(inds, C.POINTER(C.c_uint32)),
                                   vals.ctypes.data_as(C.POINTER(C.c_float)), U[0].shape[1], _parr(U, D64P),
                                   e.ctypes.data_as(D64P), C.byref(loss), int(threads or default_threads()))
    if rc != 0:
        raise RuntimeError("orc_cp_resid_f64_mt failed")
    return float(loss.value), e


# --------------------------------------------------------------------------
# driver flows (mttkrp.cu)
# --------------------------------------------------------------------------
FBRS_THRESHOLD_MMCSF = 128          # util.h:1326
NO_SPLIT = 999999999                # util.h:1323
MMCSF_CREATE_HCSR_THRESHOLD = 99999999   # util.h:763-764


def _sorted_keys(inds, order):
    perm = sort_perm(inds, order)
    keys = [np.asarray(inds[m], dtype=np.uint64)[perm] for m in order]
    return perm, keys


def build_format(fmt_type, dims, inds, vals, mode=0, fbr_threshold=99999999):
    """Reference format for one driver type; returns a dict of arrays.

    Level arrays are in modeOrder order; ``inds`` are 0-based per tensor mode.
    """
    ndims = len(dims)
    nnz = len(vals)
    vals = np.asarray(vals, dtype=np.float32)
    if fmt_type in (1, 3, 8, 10, 13):
        order = order_tensormode(dims, mode)
        perm, keys = _sorted_keys(inds, order)
        out = {"type": fmt_type, "modeOrder": order, "perm": perm,
               "keys": keys, "vals": vals[perm]}
        if fmt_type in (1, 3):
            return out
        if fmt_type == 8:
            # make_KTiling copies the sorted COO; create_TiledHCSR with no fbrS split
            ptr, idx, nf = build_csf(keys, fbr_threshold, NO_SPLIT)
        else:
            # create_HCSR (util.h:760-850): no fbrS rule at all
            ptr, idx, nf = build_csf(keys, fbr_threshold, NO_SPLIT)
        out.update(fbrPtr=ptr, fbrIdx=idx, nFibers=nf, leaf=keys[-1])
        if fmt_type == 13:
            # sort_COOtensor + create_HCSR + create_fbrLikeSlcInds (mttkrp.cu:244-253,
            # util.h:1755-1766); no bins
            out["fbrLikeSlcInds"] = fbr_like_slc_inds(ptr, idx, nf)
            return out
        if fmt_type == 8:
            out["fbrLikeSlcInds"] = fbr_like_slc_inds(ptr, idx, nf)
            bins, w, dropped = tiled_bin(ndims, nnz, ptr)
            out.update(binOf=bins, weight=w, dropped=dropped, slcMapperBin=mapper_bins(bins))
        else:
            hyb = create_hyb(ndims, ptr, idx, keys[-1], vals[perm])
            hb, hd = hyb_bin(ndims, nnz, hyb["fbrPtr"])
            hyb.update(binOf=hb, dropped=hd, slcMapperBin=mapper_bins(hb))
            if ndims == 3:
                cb, cd = csl_bin(nnz, hyb["CSLslicePtr"])
                hyb.update(CSLbinOf=cb, CSLslcMapperBin=mapper_bins(cb))
            out["hyb"] = hyb
        return out
    if fmt_type == 12:
        # all nonzeros land in partition 2 (util.h:2517); its order is
        # order_tensormode(X, 2) (util.h:2248 via init_tensor)
        order = order_tensormode(dims, 2)
        perm, keys = _sorted_keys(inds, order)
        fs = FBRS_THRESHOLD_MMCSF
        ptr, idx, nf = build_csf(keys, fbr_threshold, fs)
        return {"type": 12, "modeOrder": order, "perm": perm, "keys": keys,
                "vals": vals[perm], "fbrPtr": ptr, "fbrIdx": idx, "nFibers": nf,
                "leaf": keys[-1], "fbrLikeSlcInds": fbr_like_slc_inds(ptr, idx, nf),
                "partPerNnz": np.full(nnz, 2, dtype=np.uint64),
                "redunMode": redundant_mode(dims)}
    raise ValueError(f"unsupported format type {fmt_type}")


def dense_slice_ptr(ndims, dims_slice, fbrPtr, fbrIdx):
    """denseSlcPtr of util.h:2203-2209: the first level-1 entry of each slice
    coordinate; the entry after the last slice holds that slice's START (the
    reference writes fbrPtr[0][slc] there, not fbrPtr[0][slc + 1])."""
    sp = np.zeros(dims_slice + 1, dtype=np.uint64)
    p0 = np.asarray(fbrPtr[0], dtype=np.uint64)
    i0 = np.asarray(fbrIdx[0], dtype=np.int64)
    ns = len(i0)
    if ns:
        sp[i0] = p0[:ns]
        ends = p0[1:ns + 1].copy()
        ends[-1] = p0[ns - 1]
        sp[i0 + 1] = ends
    return sp


def nnz_per_fiber(ndims, fbrPtr):
    """nnzPerFiber of util.h:2211-2224: nonzeros under each level-1 entry
    (fbrPtr[1].size() entries, the last one 0)."""
    p1 = np.asarray(fbrPtr[1], dtype=np.int64)
    if ndims == 3:
        c = np.diff(p1)
    else:
        p2 = np.asarray(fbrPtr[2], dtype=np.int64)
        c = p2[p1[1:]] - p2[p1[:-1]]
    return np.concatenate([c, [0]]).astype(np.uint64)


def mm_partition(dims, inds, vals, fb_th=1, load_mode=0, fbr_threshold=99999999):
    """The MM-CSF mixed-mode partition (mm_partition_reuseBased,
    util.h:2167-2512, without the override of 2517) and its partitions
    (populate_paritions, sort_MI_CSF, create_TiledHCSR, create_fbrLikeSlcInds;
    mttkrp.cu:306-395).  The per-nonzero loop runs sequentially in the order
    the reference walks the tensor: sorted by every non-redundant mode's order
    in turn (the arrays are shared, util.h:380-381), i.e. by the last one's,
    with the canonical full-key order inside equal keys.

    Returns dict: order (tensor positions in loop order), part (per loop
    position), partitions[p] = None or a dict like build_format(12)."""
    N = len(dims)
    nnz = len(vals)
    vals = np.asarray(vals, dtype=np.float32)
    orders = [order_tensormode(dims, m) for m in range(N)]
    red = redundant_mode(dims)
    red = -1 if red is None else red
    sp, f1, cnt = [None] * N, [None] * N, [None] * N
    last = None
    for m in range(N):
        if m == red:
            continue
        perm, keys = _sorted_keys(inds, orders[m])
        ptr, idx, nf = build_csf(keys, MMCSF_CREATE_HCSR_THRESHOLD, NO_SPLIT)
        sp[m] = dense_slice_ptr(N, dims[orders[m][0]], ptr, idx)
        f1[m] = np.ascontiguousarray(idx[1], dtype=np.uint64)
        cnt[m] = nnz_per_fiber(N, ptr)
        last = perm
    xi = [np.ascontiguousarray(np.asarray(a)[last], dtype=np.uint64) for a in inds]
    dummy = np.zeros(1, dtype=np.uint64)
    sp_p = (U64P * N)(*[(a if a is not None else dummy).ctypes.data_as(U64P) for a in sp])
    f1_p = (U64P * N)(*[(a if a is not None else dummy).ctypes.data_as(U64P) for a in f1])
    cnt_p = (U64P * N)(*[(a if a is not None else dummy).ctypes.data_as(U64P) for a in cnt])
    oarr = [(C.c_int * N)(*o) for o in orders]
    o_p = (C.POINTER(C.c_int) * N)(*[C.cast(o, C.POINTER(C.c_int)) for o in oarr])
    part = np.zeros(nnz, dtype=np.uint64)
    d = _u64(dims)
    lib().orc_mm_partition(N, nnz, _parr(xi, U64P), _p64(d), red, int(fb_th), int(load_mode), o_p, sp_p,
                           f1_p, cnt_p, _p64(part))
    parts = []
    for p in range(N):
        sel = np.nonzero(part == p)[0]
        if len(sel) == 0:
            parts.append(None)
            continue
        pin = [np.asarray(a)[sel] for a in xi]
        perm, keys = _sorted_keys(pin, orders[p])
        fs = FBRS_THRESHOLD_MMCSF if N == 4 else NO_SPLIT
        ptr, idx, nf = build_csf(keys, fbr_threshold, fs)
        parts.append({"modeOrder": orders[p], "keys": keys, "vals": vals[last][sel][perm], "fbrPtr": ptr,
                      "fbrIdx": idx, "nFibers": nf, "fbrLikeSlcInds": fbr_like_slc_inds(ptr, idx, nf),
                      "nnz": len(sel)})
    return {"order": np.asarray(last, dtype=np.int64), "part": part, "partitions": parts, "redunMode": red}


def redundant_mode(dims):
    """mttkrp.cu:315-347: last m whose (order[0], order[1]) swaps an earlier one's."""
    orders = [order_tensormode(dims, m) for m in range(len(dims))]
    red = None
    for m in range(1, len(dims)):
        for q in range(m):
            if orders[m][1] == orders[q][0] and orders[m][0] == orders[q][1]:
                red = m
    return red


def load_tns(path):
    """load_tensor (util.h:313-370): 1-based text -> 0-based COO."""
    with open(path) as f:
        lines = f.read().split("\n")
    toks = " ".join(lines).split()
    ndims = int(toks[0])
    dims = [int(t) for t in toks[1:1 + ndims]]
    body = toks[1 + ndims:]
    n = len(body) // (ndims + 1)
    arr = np.array(body[: n * (ndims + 1)], dtype=np.float64).reshape(n, ndims + 1)
    inds = [arr[:, m].astype(np.int64) - 1 for m in range(ndims)]
    vals = arr[:, ndims].astype(np.float32)
    return dims, inds, vals
