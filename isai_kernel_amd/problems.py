"""Host-side wrappers (numpy in, numpy out) over the library's host entry points:
synthetic problem generation and the pattern / format conversions."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._ffi import call, ptr

BS = 4

# BiCGSTAB / FGMRES workloads, by problem id (generator in include/isai_gen.h).
PROBLEM_NAMES = {
    0: "2-D 5-pt 64^2, bs=4",
    1: "3-D 7-pt 96^3, bs=4",
    2: "3-D 27-pt convection-diffusion 64^3, bs=4",
    3: "2-D 9-pt LxL, bs=4",
    4: "2-D 9-pt 256^2, bs=4",
}


def generate(problem: int, L: int = 0, seed: int = 0):
    """BSR arrays (rowptr int32, colind int32, val float64 [nnzb,4,4]) of a problem."""
    nb, nnzb = C.c_int(), C.c_int()
    call("isai_gen_problem_count", problem, L, C.byref(nb), C.byref(nnzb))
    rp = np.empty(nb.value + 1, np.int32)
    ci = np.empty(nnzb.value, np.int32)
    val = np.empty((nnzb.value, BS, BS), np.float64)
    call("isai_gen_problem_fill", problem, L, seed, ptr(rp), ptr(ci), ptr(val))
    return rp, ci, val


def batch(nsys: int, lmin: int, lmax: int, seed: int = 0, rhs_max: int = 16, problem: int = 3):
    """Block-diagonal BSR of nsys problem-`problem` systems (2-D 9-pt by default)
    with grid side L_s uniform in [lmin, lmax] and 1..rhs_max right-hand sides each.

    L_s and the right-hand-side counts come from numpy's default_rng(seed); system s
    is generated with seed + s.  Returns (rp, ci, val, sys_ptr, Ls, nrhs)."""
    rng = np.random.default_rng(seed)
    Ls = rng.integers(lmin, lmax + 1, nsys)
    nrhs = rng.integers(1, rhs_max + 1, nsys).astype(np.int32)
    counts = []
    for L in Ls:
        nb, nnzb = C.c_int(), C.c_int()
        call("isai_gen_problem_count", problem, int(L), C.byref(nb), C.byref(nnzb))
        counts.append((nb.value, nnzb.value))
    sys_ptr = np.zeros(nsys + 1, np.int32)
    sys_ptr[1:] = np.cumsum([c[0] for c in counts])
    blk_ptr = np.zeros(nsys + 1, np.int64)
    blk_ptr[1:] = np.cumsum([c[1] for c in counts])
    rp = np.empty(sys_ptr[-1] + 1, np.int32)
    ci = np.empty(blk_ptr[-1], np.int32)
    val = np.empty((blk_ptr[-1], BS, BS), np.float64)
    for s, L in enumerate(Ls):
        r0, r1, k0, k1 = sys_ptr[s], sys_ptr[s + 1], blk_ptr[s], blk_ptr[s + 1]
        srp, sci, sval = generate(problem, int(L), seed + s)
        rp[r0:r1] = srp[:-1] + k0
        ci[k0:k1] = sci + r0
        val[k0:k1] = sval
    rp[-1] = blk_ptr[-1]
    return rp, ci, val, sys_ptr, Ls, nrhs


def normal(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    call("isai_gen_normal", seed, n, ptr(out))
    return out


def csr_to_bsr(n: int, bs: int, rp: np.ndarray, ci: np.ndarray, v: np.ndarray):
    nb, nnzb = C.c_int(), C.c_int()
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    call("isai_csr_to_bsr_count", n, bs, ptr(rp), ptr(ci), C.byref(nb), C.byref(nnzb))
    brp = np.empty(nb.value + 1, np.int32)
    bci = np.empty(nnzb.value, np.int32)
    bv = np.empty((nnzb.value, bs, bs), np.float64)
    call("isai_csr_to_bsr_fill", n, bs, ptr(rp), ptr(ci), ptr(v), ptr(brp), ptr(bci), ptr(bv))
    return brp, bci, bv


def tri_bcsr_pattern(rp: np.ndarray, ci: np.ndarray, lower: bool):
    nb = len(rp) - 1
    nnz = C.c_int()
    call("isai_tri_bcsr_pattern_count", nb, ptr(rp), ptr(ci), int(lower), C.byref(nnz))
    trp = np.empty(nb + 1, np.int32)
    tci = np.empty(nnz.value, np.int32)
    call("isai_tri_bcsr_pattern", nb, ptr(rp), ptr(ci), int(lower), ptr(trp), ptr(tci))
    return trp, tci


def bcsc_pattern(rp: np.ndarray, ci: np.ndarray, lower: bool):
    nb = len(rp) - 1
    nnz = C.c_int()
    call("isai_bcsc_pattern_count", nb, ptr(rp), ptr(ci), int(lower), C.byref(nnz))
    mb = np.empty(nb + 1, np.int32)
    mc = np.empty(nnz.value, np.int32)
    call("isai_bcsc_pattern", nb, ptr(rp), ptr(ci), int(lower), ptr(mb), ptr(mc))
    return mb, mc


def map_bcsc2bcsr(colptr, rowind, rowptr, colind) -> np.ndarray:
    nb = len(colptr) - 1
    out = np.empty(len(rowind), np.int32)
    call("isai_map_bcsc2bcsr", nb, ptr(colptr), ptr(rowind), ptr(rowptr), ptr(colind), ptr(out))
    return out
