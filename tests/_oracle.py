"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- the checker.

Only tests, the smoke check and bench.py's cpu_baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_SO = ROOT / "oracle" / "liboracle.so"

if not ORACLE_SO.exists():
    raise ImportError(f"{ORACLE_SO} is missing: run `make -C oracle`")

ora = C.CDLL(str(ORACLE_SO))

P, I, D = C.c_void_p, C.c_int, C.c_double

_SIGS = {
    "ora_csr2bsr_count": ([I, I, P, P, P, P], I),
    "ora_csr2bsr_fill": ([I, I, P, P, P, P, P, P], I),
    "ora_bsr2csr": ([I, I, P, P, P, P, P, P], I),
    "ora_bcsc_pattern": ([I, P, P, I, P, P, I, P], I),
    "ora_tri_bcsr_pattern": ([I, P, P, I, P, P, I, P], I),
    "ora_map_bcsc2bcsr": ([I, P, P, P, P, P], I),
    "ora_reorder_bcsc_to_bcsr": ([I, I, P, P, P], None),
    "ora_pairs_lower": ([I, P, P, P, P, P, P, P, I, P], I),
    "ora_pairs_upper": ([I, P, P, P, P, P, P, P, I, P], I),
    "ora_invert_block": ([I, P, P, D], I),
    "ora_build_dinv": ([I, I, P, P, P, I, P], None),
    "ora_isai_lower": ([I, I, P, P, P, P, P, P, P, P], None),
    "ora_isai_upper": ([I, I, P, P, P, P, P, P, P], None),
    "ora_bsrilu0": ([I, I, P, P, P, P], I),
    "ora_levels": ([I, P, P, I, P], I),
    "ora_bsr_spmv": ([I, I, P, P, P, P, P], None),
    "ora_bsr_spmv_f32acc": ([I, I, P, P, P, I, P, P], None),
    "ora_round_f32": ([D], C.c_float),
    "ora_round_bf16": ([D], C.c_float),
    "ora_bsr_sptrsv_lu": ([I, I, P, P, P, P, P], None),
    "ora_csr_sptrsv_lu": ([I, P, P, P, P, P], I),
    "ora_rel_residual_bsr": ([I, I, P, P, P, P, P], D),
    "ora_bicgstab": ([I, I, P, P, P, P, P, P, D, I, P, P], I),
    "ora_fgmres": ([I, I, P, P, P, P, P, P, I, D, I, P, P], I),
    "ora_fgmres_mixed": ([I, I, P, P, P, P, P, P, I, D, I, D, I, P, P, P], I),
    "ora_fgmres_mixed_gmres": ([I, I, P, P, P, P, P, P, I, D, I, D, I, P, P, P], I),
    "ora_precond_apply": ([P, P, P], None),
    "ora_num_threads": ([], I),
    "ora_bsr_transpose": ([I, I, P, P, P, P, P, P], None),
    "ora_transpose_packed_lu": ([I, I, P, P, P, P, P, P], I),
    "ora_bsr_outer_grad": ([I, I, P, P, P, P, D, P], None),
    "ora_pair_stats": ([I, P, P, P, P, P, P, P, P], I),
}
for _n, (_a, _r) in _SIGS.items():
    _f = getattr(ora, _n)
    _f.argtypes = _a
    _f.restype = _r


def _p(a):
    return None if a is None else a.ctypes.data


# Converted copies stay referenced here until well after the call that uses their
# address: `_p(_i32(x))` would otherwise hand C the address of a freed temporary
# whenever x is not already contiguous int32 / float64.
_KEEP: list = []


def _keep(a, r):
    if r is not a:
        _KEEP.append(r)
        if len(_KEEP) > 256:
            del _KEEP[:128]
    return r


def _i32(a):
    return _keep(a, np.ascontiguousarray(a, np.int32))


def _f64(a):
    return _keep(a, np.ascontiguousarray(a, np.float64))


class PrecondT(C.Structure):
    _fields_ = [
        ("precond", I), ("nb", I), ("bs", I),
        ("Lrp", P), ("Lci", P), ("Lval", P),
        ("Urp", P), ("Uci", P), ("Uval", P),
        ("LUrp", P), ("LUci", P), ("LUval", P),
        ("isai_bf16", I), ("isai_f32", I),
    ]


def csr2bsr(n, bs, rp, ci, v):
    rp, ci, v = _i32(rp), _i32(ci), _f64(v)
    nb, nnzb = C.c_int(), C.c_int()
    st = ora.ora_csr2bsr_count(n, bs, _p(rp), _p(ci), C.byref(nb), C.byref(nnzb))
    if st:
        raise ValueError(f"ora_csr2bsr_count failed {st}")
    brp = np.empty(nb.value + 1, np.int32)
    bci = np.empty(nnzb.value, np.int32)
    bv = np.empty((nnzb.value, bs, bs))
    ora.ora_csr2bsr_fill(n, bs, _p(rp), _p(ci), _p(v), _p(brp), _p(bci), _p(bv))
    return brp, bci, bv


def bsr2csr(rp, ci, val):
    nb, bs = len(rp) - 1, val.shape[1]
    nnz = len(ci) * bs * bs
    crp = np.empty(nb * bs + 1, np.int32)
    cci = np.empty(nnz, np.int32)
    cv = np.empty(nnz)
    ora.ora_bsr2csr(nb, bs, _p(_i32(rp)), _p(_i32(ci)), _p(_f64(val)), _p(crp), _p(cci), _p(cv))
    return crp, cci, cv


def bcsc_pattern(rp, ci, lower):
    nb = len(rp) - 1
    cap = len(ci) + nb
    mb = np.empty(nb + 1, np.int32)
    mc = np.empty(max(cap, 1), np.int32)
    nnz = C.c_int()
    st = ora.ora_bcsc_pattern(nb, _p(rp), _p(ci), int(lower), _p(mb), _p(mc), cap, C.byref(nnz))
    assert st == 0
    return mb, mc[: nnz.value].copy()


def tri_bcsr_pattern(rp, ci, lower):
    nb = len(rp) - 1
    cap = len(ci)
    trp = np.empty(nb + 1, np.int32)
    tci = np.empty(max(cap, 1), np.int32)
    nnz = C.c_int()
    st = ora.ora_tri_bcsr_pattern(nb, _p(rp), _p(ci), int(lower), _p(trp), _p(tci), cap, C.byref(nnz))
    assert st == 0
    return trp, tci[: nnz.value].copy()


def map_bcsc2bcsr(colptr, rowind, rowptr, colind):
    m = np.empty(len(rowind), np.int32)
    st = ora.ora_map_bcsc2bcsr(len(colptr) - 1, _p(colptr), _p(rowind), _p(rowptr), _p(colind), _p(m))
    return st, m


def reorder(mp, src, bs=4):
    dst = np.zeros_like(src)
    ora.ora_reorder_bcsc_to_bcsr(len(mp), bs, _p(_i32(mp)), _p(_f64(src)), _p(dst))
    return dst


def pairs(rp, ci, mb, mc, upper=False):
    nb = len(rp) - 1
    mnnz = int(mb[-1])
    pp = np.empty(mnnz + 1, np.int32)
    total = C.c_int()
    fn = ora.ora_pairs_upper if upper else ora.ora_pairs_lower
    st = fn(nb, _p(rp), _p(ci), _p(mb), _p(mc), _p(pp), None, None, 0, C.byref(total))
    cap = total.value
    pl = np.empty(max(cap, 1), np.int32)
    ps = np.empty(max(cap, 1), np.int32)
    st = fn(nb, _p(rp), _p(ci), _p(mb), _p(mc), _p(pp), _p(pl), _p(ps), cap, C.byref(total))
    assert st == 0
    return pp, pl[:cap].copy(), ps[:cap].copy()


def invert_block(A, eps=1e-14):
    A = _f64(A)
    out = np.empty_like(A)
    ok = ora.ora_invert_block(A.shape[0], _p(A), _p(out), eps)
    return bool(ok), out


def build_dinv(rp, ci, val, lower):
    nb, bs = len(rp) - 1, val.shape[1]
    out = np.empty((nb, bs, bs))
    ora.ora_build_dinv(nb, bs, _p(rp), _p(ci), _p(_f64(val)), int(lower), _p(out))
    return out


def isai_lower(val, mb, mc, pp, pl, ps, dinv):
    nb, bs = len(mb) - 1, val.shape[1]
    out = np.zeros((int(mb[-1]), bs, bs))
    ora.ora_isai_lower(nb, bs, _p(_f64(val)), _p(mb), _p(mc), _p(pp), _p(pl), _p(ps), _p(_f64(dinv)), _p(out))
    return out


def isai_upper(rp, ci, val, mb, mc, dinv):
    nb, bs = len(mb) - 1, val.shape[1]
    out = np.zeros((int(mb[-1]), bs, bs))
    ora.ora_isai_upper(nb, bs, _p(rp), _p(ci), _p(_f64(val)), _p(mb), _p(mc), _p(_f64(dinv)), _p(out))
    return out


def bsrilu0(rp, ci, val):
    """Returns (packed LU, status, zero_pivot)."""
    nb, bs = len(rp) - 1, val.shape[1]
    lu = np.array(val, np.float64, copy=True)
    zp = C.c_int()
    st = ora.ora_bsrilu0(nb, bs, _p(rp), _p(ci), _p(lu), C.byref(zp))
    return lu, st, zp.value


def levels(rp, ci, lower=True):
    nb = len(rp) - 1
    lev = np.empty(nb, np.int32)
    n = ora.ora_levels(nb, _p(rp), _p(ci), int(lower), _p(lev))
    return lev, n


def spmv(rp, ci, val, x):
    nb, bs = len(rp) - 1, val.shape[1]
    y = np.empty(nb * bs)
    ora.ora_bsr_spmv(nb, bs, _p(rp), _p(ci), _p(_f64(val)), _p(_f64(x)), _p(y))
    return y


def spmv_f32acc(rp, ci, val, x, bf16=False):
    nb, bs = len(rp) - 1, val.shape[1]
    y = np.empty(nb * bs)
    ora.ora_bsr_spmv_f32acc(nb, bs, _p(rp), _p(ci), _p(_f64(val)), int(bf16), _p(_f64(x)), _p(y))
    return y


def sptrsv_lu(rp, ci, lu, b):
    nb, bs = len(rp) - 1, lu.shape[1]
    x = np.empty(nb * bs)
    ora.ora_bsr_sptrsv_lu(nb, bs, _p(rp), _p(ci), _p(_f64(lu)), _p(_f64(b)), _p(x))
    return x


def csr_sptrsv_lu(rp, ci, v, b):
    m = len(rp) - 1
    x = np.empty(m)
    st = ora.ora_csr_sptrsv_lu(m, _p(_i32(rp)), _p(_i32(ci)), _p(_f64(v)), _p(_f64(b)), _p(x))
    return st, x


def rel_residual(rp, ci, val, b, x):
    nb, bs = len(rp) - 1, val.shape[1]
    return ora.ora_rel_residual_bsr(nb, bs, _p(rp), _p(ci), _p(_f64(val)), _p(_f64(b)), _p(_f64(x)))


def make_precond(kind, nb, bs, L=None, U=None, LU=None, f32=False, bf16=False):
    """kind: 0 none, 1 ISAI (L, U = (rp, ci, val) BCSR), 2 ILU solves (LU = (rp, ci, val))."""
    P_ = PrecondT()
    P_.precond, P_.nb, P_.bs = kind, nb, bs
    keep = []
    if L is not None:
        keep += list(L) + list(U)
        P_.Lrp, P_.Lci, P_.Lval = _p(L[0]), _p(L[1]), _p(L[2])
        P_.Urp, P_.Uci, P_.Uval = _p(U[0]), _p(U[1]), _p(U[2])
    if LU is not None:
        keep += list(LU)
        P_.LUrp, P_.LUci, P_.LUval = _p(LU[0]), _p(LU[1]), _p(LU[2])
    P_.isai_f32, P_.isai_bf16 = int(f32), int(bf16)
    P_._keep = keep
    return P_


def bicgstab(rp, ci, val, P_, b, x0=None, tol=1e-10, maxit=1000):
    nb, bs = len(rp) - 1, val.shape[1]
    x = np.zeros(nb * bs) if x0 is None else _f64(x0).copy()
    it, rr = C.c_int(), C.c_double()
    st = ora.ora_bicgstab(nb, bs, _p(rp), _p(ci), _p(_f64(val)), C.byref(P_), _p(_f64(b)), _p(x),
                          tol, maxit, C.byref(it), C.byref(rr))
    return x, st, it.value, rr.value


def fgmres(rp, ci, val, P_, b, m=30, x0=None, tol=1e-10, maxit=1000):
    nb, bs = len(rp) - 1, val.shape[1]
    x = np.zeros(nb * bs) if x0 is None else _f64(x0).copy()
    it, rr = C.c_int(), C.c_double()
    st = ora.ora_fgmres(nb, bs, _p(rp), _p(ci), _p(_f64(val)), C.byref(P_), _p(_f64(b)), _p(x), m,
                        tol, maxit, C.byref(it), C.byref(rr))
    return x, st, it.value, rr.value


def fgmres_mixed(rp, ci, val, P_, b, m=10, x0=None, tol=1e-10, maxit=400, inner_tol=1e-5, inner_maxit=40,
                 inner="bicgstab"):
    """FP64 FGMRES(m) with an inner FP32-storage solve as flexible preconditioner:
    BiCGSTAB (inner="bicgstab") or one GMRES cycle of at most inner_maxit steps
    (inner="gmres").  Returns (x, status, inner iterations, outer steps, true relres)."""
    nb, bs = len(rp) - 1, val.shape[1]
    x = np.zeros(nb * bs) if x0 is None else _f64(x0).copy()
    it, out, rr = C.c_int(), C.c_int(), C.c_double()
    fn = ora.ora_fgmres_mixed_gmres if inner == "gmres" else ora.ora_fgmres_mixed
    st = fn(nb, bs, _p(rp), _p(ci), _p(_f64(val)), C.byref(P_), _p(_f64(b)), _p(x), m, tol,
            maxit, inner_tol, inner_maxit, C.byref(it), C.byref(out), C.byref(rr))
    return x, st, it.value, out.value, rr.value


def isai_factors(rp, ci, val):
    """The oracle's whole setup path: packed ILU(0), then ISAI(L) and ISAI(U) in BCSC
    reordered to BCSR.  Returns (lu, (rpL, ciL, valL), (rpU, ciU, valU))."""
    lu, _, _ = bsrilu0(rp, ci, val)
    mbL, mcL = bcsc_pattern(rp, ci, True)
    mbU, mcU = bcsc_pattern(rp, ci, False)
    pp, pl, ps = pairs(rp, ci, mbL, mcL)
    ML = isai_lower(lu, mbL, mcL, pp, pl, ps, build_dinv(rp, ci, lu, True))
    MU = isai_upper(rp, ci, lu, mbU, mcU, build_dinv(rp, ci, lu, False))
    trpL, tciL = tri_bcsr_pattern(rp, ci, True)
    trpU, tciU = tri_bcsr_pattern(rp, ci, False)
    _, mapL = map_bcsc2bcsr(mbL, mcL, trpL, tciL)
    _, mapU = map_bcsc2bcsr(mbU, mcU, trpU, tciU)
    return lu, (trpL, tciL, reorder(mapL, ML)), (trpU, tciU, reorder(mapU, MU))


def precond_apply(P_, v):
    z = np.empty(P_.nb * P_.bs)
    ora.ora_precond_apply(C.byref(P_), _p(_f64(v)), _p(z))
    return z


def bsr_transpose(rp, ci, val):
    nb, bs = len(rp) - 1, val.shape[1]
    rp, ci = _i32(rp), _i32(ci)
    trp = np.empty(nb + 1, np.int32)
    tci = np.empty(len(ci), np.int32)
    tval = np.empty((len(ci), bs, bs))
    ora.ora_bsr_transpose(nb, bs, _p(rp), _p(ci), _p(_f64(val)), _p(trp), _p(tci), _p(tval))
    return trp, tci, tval


def transpose_packed_lu(rp, ci, lu):
    """(trp, tci, packed ILU(0) factor of A^T) from the packed factor of A."""
    nb, bs = len(rp) - 1, lu.shape[1]
    trp, tci, _ = bsr_transpose(rp, ci, lu)
    lut = np.empty_like(lu)
    st = ora.ora_transpose_packed_lu(nb, bs, _p(_i32(rp)), _p(_i32(ci)), _p(_f64(lu)), _p(trp), _p(tci), _p(lut))
    if st:
        raise ValueError(f"ora_transpose_packed_lu failed {st}")
    return trp, tci, lut


def outer_grad(rp, ci, lam, x, alpha, bs=4):
    nb = len(rp) - 1
    g = np.empty((len(ci), bs, bs))
    ora.ora_bsr_outer_grad(nb, bs, _p(_i32(rp)), _p(_i32(ci)), _p(_f64(lam)), _p(_f64(x)), alpha, _p(g))
    return g


PCTS = (50, 90, 95, 99)


def stats_dict(ints, i64, dbl):
    """Pair statistics as a dict, from the 39 ints / 21 int64 / 3 doubles layout
    shared by ora_pair_stats and isai_pair_stats."""
    d = {"nb": int(ints[0]), "mnnz": int(ints[1]), "n_max": int(ints[2]), "q_max": int(ints[3]),
         "qlen_max": int(ints[4]), "total_pairs": int(i64[0]),
         "n_avg": float(dbl[0]), "q_avg": float(dbl[1]), "qlen_avg": float(dbl[2]),
         "qlen_hist": [int(v) for v in i64[1:11]], "bin_sum_q": [int(v) for v in i64[11:21]],
         "bin_cols": [int(v) for v in ints[17:27]], "bin_max_q": [int(v) for v in ints[27:37]]}
    for k, pc in enumerate(PCTS):
        d[f"n_p{pc}"], d[f"q_p{pc}"], d[f"qlen_p{pc}"] = int(ints[5 + k]), int(ints[9 + k]), int(ints[13 + k])
    return d


def pair_stats(mb, pair_ptr):
    """analyze_pair_lengths / analyze_column_q_bins figures and the per-column
    (N, Q, max qlen) arrays."""
    nb = len(mb) - 1
    ints, i64, dbl = np.zeros(39, np.int32), np.zeros(21, np.int64), np.zeros(3)
    cn, cq, cm = np.zeros(nb, np.int32), np.zeros(nb, np.int32), np.zeros(nb, np.int32)
    st = ora.ora_pair_stats(nb, _p(_i32(mb)), _p(_i32(pair_ptr)), _p(ints), _p(i64), _p(dbl), _p(cn), _p(cq), _p(cm))
    if st:
        raise ValueError(f"ora_pair_stats failed {st}")
    return stats_dict(ints, i64, dbl), cn, cq, cm


def num_threads():
    return ora.ora_num_threads()


# ------------------------------------------------------------ dense helpers
def bsr_to_dense(rp, ci, val):
    nb, bs = len(rp) - 1, val.shape[1]
    A = np.zeros((nb * bs, nb * bs))
    for i in range(nb):
        for k in range(rp[i], rp[i + 1]):
            j = ci[k]
            A[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] += val[k]
    return A


def bcsc_to_dense(mb, mc, mval, nb):
    bs = mval.shape[1]
    M = np.zeros((nb * bs, nb * bs))
    for j in range(nb):
        for k in range(mb[j], mb[j + 1]):
            i = mc[k]
            M[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = mval[k]
    return M
