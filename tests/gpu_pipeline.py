"""Shared helpers of the GPU parity tests: run the setup path (ILU(0), pair lists,
Dinv, ISAI(L/U), BCSC->BCSR) through the C ABI on device buffers."""
from __future__ import annotations

import ctypes

import numpy as np
import scipy.sparse as sp

from isai_kernel_amd import DeviceArray, call, problems


def dev(a):
    return DeviceArray.from_numpy(a)


def gpu_ilu(rp, ci, val):
    nb, nnzb, bs = len(rp) - 1, len(ci), val.shape[1]
    d_rp, d_ci, d_v = dev(rp), dev(ci), dev(val)
    zp = ctypes.c_int()
    call("isai_bsrilu0_inplace", nb, nnzb, bs, d_rp.ptr, d_ci.ptr, d_v.ptr, ctypes.byref(zp), None)
    return d_rp, d_ci, d_v


def gpu_pairs(d_rp, d_ci, mb, mc, upper):
    nb = len(mb) - 1
    mnnz = int(mb[-1])
    d_mb, d_mc = dev(mb), dev(mc)
    d_pp = DeviceArray((mnnz + 1,), np.int32)
    total = ctypes.c_int()
    call("isai_pairs_count", nb, mnnz, d_rp.ptr, d_ci.ptr, d_mb.ptr, d_mc.ptr, int(upper), d_pp.ptr,
         ctypes.byref(total), None)
    d_pl = DeviceArray((max(total.value, 1),), np.int32)
    d_ps = DeviceArray((max(total.value, 1),), np.int32)
    call("isai_pairs_fill", nb, d_rp.ptr, d_ci.ptr, d_mb.ptr, d_mc.ptr, int(upper), d_pp.ptr, d_pl.ptr,
         d_ps.ptr, None)
    return d_mb, d_mc, d_pp, d_pl, d_ps, total.value


def gpu_isai(rp, ci, d_rp, d_ci, d_lu, bs=4, one_launch=True):
    """Returns host copies: dict with L/U BCSC patterns, values, pairs, dinv."""
    nb = len(rp) - 1
    out = {}
    fac = {}
    for lower in (True, False):
        mb, mc = problems.bcsc_pattern(rp, ci, lower)
        d_mb, d_mc, d_pp, d_pl, d_ps, tot = gpu_pairs(d_rp, d_ci, mb, mc, upper=not lower)
        d_dinv = DeviceArray((nb, bs, bs), np.float64)
        call("isai_build_dinv", nb, bs, d_rp.ptr, d_ci.ptr, d_lu.ptr, int(lower), d_dinv.ptr, None)
        d_mv = DeviceArray((len(mc), bs, bs), np.float64)
        fac[lower] = (mb, mc, d_mb, d_mc, d_pp, d_pl, d_ps, tot, d_dinv, d_mv)
    nmax = max(int(np.diff(fac[True][0]).max()), int(np.diff(fac[False][0]).max()))
    L, U = fac[True], fac[False]
    if one_launch:
        call("isai_build_lu", nb, bs, d_lu.ptr, L[2].ptr, L[3].ptr, L[4].ptr, L[5].ptr, L[6].ptr, L[8].ptr,
             L[9].ptr, U[2].ptr, U[3].ptr, U[4].ptr, U[5].ptr, U[6].ptr, U[8].ptr, U[9].ptr, nmax, None)
    else:
        call("isai_build_lower", nb, bs, d_lu.ptr, L[2].ptr, L[3].ptr, L[4].ptr, L[5].ptr, L[6].ptr, L[8].ptr,
             L[9].ptr, nmax, None)
        call("isai_build_upper", nb, bs, d_lu.ptr, U[2].ptr, U[3].ptr, U[4].ptr, U[5].ptr, U[6].ptr, U[8].ptr,
             U[9].ptr, nmax, None)
    call("isai_device_synchronize")
    for lower, key in ((True, "L"), (False, "U")):
        mb, mc, d_mb, d_mc, d_pp, d_pl, d_ps, tot, d_dinv, d_mv = fac[lower]
        out[key] = dict(mb=mb, mc=mc, pp=d_pp.to_numpy(), pl=d_pl.to_numpy()[:tot], ps=d_ps.to_numpy()[:tot],
                        dinv=d_dinv.to_numpy(), mval=d_mv.to_numpy(), d_mval=d_mv, d_mb=d_mb, d_mc=d_mc)
    out["nmax"] = nmax
    return out


def to_bcsr(rp, ci, f, lower):
    """BCSC ISAI values -> BCSR through the host map and the device permutation."""
    trp, tci = problems.tri_bcsr_pattern(rp, ci, lower)
    mp = problems.map_bcsc2bcsr(f["mb"], f["mc"], trp, tci)
    d_map = dev(mp)
    d_dst = DeviceArray(f["mval"].shape, np.float64)
    call("isai_reorder_bcsc_to_bcsr", len(mp), f["mval"].shape[1], d_map.ptr, f["d_mval"].ptr, d_dst.ptr, None)
    call("isai_device_synchronize")
    return trp, tci, d_dst


def scipy_bsr(rp, ci, val):
    nb, bs = len(rp) - 1, val.shape[1]
    return sp.bsr_matrix((val, ci, rp), shape=(nb * bs, nb * bs))


def split_lu(rp, ci, lu):
    """Unit-lower L and upper U (scipy CSR) from a packed block factor."""
    A = scipy_bsr(rp, ci, lu).tocoo()
    low = A.row > A.col
    n = A.shape[0]
    L = sp.csr_matrix((A.data[low], (A.row[low], A.col[low])), shape=A.shape) + sp.identity(n, format="csr")
    up = A.row <= A.col
    U = sp.csr_matrix((A.data[up], (A.row[up], A.col[up])), shape=A.shape)
    return L, U


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def round_values(val, bf16):
    """FP64 block values rounded as the library stores them: FP32 (RNE), or BF16
    (RNE on the top 16 bits of the FP32 value)."""
    f = np.asarray(val, np.float64).astype(np.float32)
    if not bf16:
        return f.astype(np.float64)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def reduced_apply_bound(L, U, v, y_ref, bf16):
    """Row-wise bound on |z_gpu - z_oracle| for z = M_U (M_L v) with FP32 / BF16
    factor values and FP32 accumulation, when both sides use the same rounded
    operands and differ only in the order of the FP32 sums (k terms per row:
    |sum error| <= k 2^-24 sum |terms| for either order).  The first product's
    difference dy passes through |M_U|; the second adds its own order term."""
    from _oracle import spmv
    u = 2.0**-24
    vf = np.abs(np.asarray(v, np.float32).astype(np.float64))
    kL = 4 * int(np.diff(L[0]).max())
    kU = 4 * int(np.diff(U[0]).max())
    aL = np.abs(round_values(L[2], bf16))
    aU = np.abs(round_values(U[2], bf16))
    dy = 2 * kL * u * spmv(L[0], L[1], aL, vf)
    return (spmv(U[0], U[1], aU, dy) + 2 * kU * u * spmv(U[0], U[1], aU, np.abs(y_ref))) * 1.01
