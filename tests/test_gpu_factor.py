"""GPU parity of ILU(0), pair lists, Dinv and the ISAI build against the oracle,
plus full-size invariants on the 96^3 benchmark matrix."""
import ctypes

import numpy as np
import pytest

import _oracle as O
from gpu_pipeline import dev, gpu_ilu, gpu_isai, rel_err, scipy_bsr, split_lu, to_bcsr
import scipy.sparse as sp

from isai_kernel_amd import DeviceArray, IsaiError, call, lib, problems

pytestmark = pytest.mark.gpu
TOL = 1e-12

CASES = [(0, 16), (1, 8), (2, 5), (3, 12), (4, 10)]


def _random_bsr(nb, bs, density, seed):
    rng = np.random.default_rng(seed)
    mask = rng.random((nb, nb)) < density
    np.fill_diagonal(mask, True)
    rp = np.concatenate([[0], np.cumsum(mask.sum(1))]).astype(np.int32)
    ci = np.concatenate([np.nonzero(r)[0] for r in mask]).astype(np.int32)
    val = rng.uniform(-1, 1, (len(ci), bs, bs))
    for i in range(nb):
        k = rp[i] + int(np.nonzero(ci[rp[i]:rp[i + 1]] == i)[0][0])
        val[k] += np.eye(bs) * 4 * bs * (rp[i + 1] - rp[i])
    return rp, ci, val


@pytest.mark.parametrize("problem,L", CASES)
def test_ilu0_matches_oracle(gpu, problem, L):
    print(f"ilu0 problem={problem} L={L}", flush=True)
    rp, ci, val = problems.generate(problem, L, seed=5)
    _, _, d_lu = gpu_ilu(rp, ci, val)
    lu = d_lu.to_numpy()
    ref, st, _ = O.bsrilu0(rp, ci, val)
    assert st == 0
    # same scalar operations in the same order as the oracle: bit-identical
    assert np.array_equal(lu, ref), rel_err(lu, ref)
    # determinism: a second factorization is bit-identical
    assert np.array_equal(gpu_ilu(rp, ci, val)[2].to_numpy(), lu)


@pytest.mark.parametrize("bs", [2, 3, 5])
def test_ilu0_other_block_sizes(gpu, bs):
    print(f"ilu0 random bs={bs}", flush=True)
    rp, ci, val = _random_bsr(30, bs, 0.15, seed=bs)
    lu = gpu_ilu(rp, ci, val)[2].to_numpy()
    ref, st, _ = O.bsrilu0(rp, ci, val)
    assert st == 0 and np.array_equal(lu, ref), rel_err(lu, ref)


def test_ilu0_zero_pivots(gpu):
    print("ilu0 zero pivots", flush=True)
    # structural: row 1 has no diagonal block
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 0], np.int32)
    val = np.stack([np.eye(4), np.eye(4)])
    d = [dev(a) for a in (rp, ci, val)]
    zp = ctypes.c_int()
    with pytest.raises(IsaiError) as e:
        call("isai_bsrilu0_inplace", 2, 2, 4, d[0].ptr, d[1].ptr, d[2].ptr, ctypes.byref(zp), None)
    assert e.value.status == -5
    # numerical: U(0,0) gets an exact zero on its diagonal
    v = np.eye(4)
    v[2, 2] = 0.0
    d = [dev(a) for a in (np.array([0, 1], np.int32), np.array([0], np.int32), v[None])]
    zp = ctypes.c_int(-7)
    st = lib.isai_bsrilu0_inplace(1, 1, 4, d[0].ptr, d[1].ptr, d[2].ptr, ctypes.byref(zp), None)
    assert st == -5 and zp.value == 0


@pytest.mark.parametrize("problem,L", CASES)
@pytest.mark.parametrize("one_launch", [True, False])
def test_isai_build_matches_oracle(gpu, problem, L, one_launch):
    print(f"isai build problem={problem} L={L} one_launch={one_launch}", flush=True)
    rp, ci, val = problems.generate(problem, L, seed=6)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    lu = d_lu.to_numpy()
    f = gpu_isai(rp, ci, d_rp, d_ci, d_lu, one_launch=one_launch)
    for lower, key in ((True, "L"), (False, "U")):
        F = f[key]
        pp, pl, ps = O.pairs(rp, ci, F["mb"], F["mc"], upper=not lower)
        assert np.array_equal(F["pp"], pp) and np.array_equal(F["pl"], pl) and np.array_equal(F["ps"], ps)
        dref = O.build_dinv(rp, ci, lu, lower)
        assert np.array_equal(F["dinv"], dref), rel_err(F["dinv"], dref)
        if lower:
            mref = O.isai_lower(lu, F["mb"], F["mc"], pp, pl, ps, dref)
        else:
            mref = O.isai_upper(rp, ci, lu, F["mb"], F["mc"], dref)
        assert np.array_equal(F["mval"], mref), (key, rel_err(F["mval"], mref))


@pytest.mark.parametrize("bs", [3, 5, 7])
def test_isai_build_other_block_sizes(gpu, bs):
    print(f"isai build bs={bs}", flush=True)
    rp, ci, val = _random_bsr(40, bs, 0.12, seed=10 + bs)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    lu = d_lu.to_numpy()
    f = gpu_isai(rp, ci, d_rp, d_ci, d_lu, bs=bs)
    for lower, key in ((True, "L"), (False, "U")):
        F = f[key]
        pp, pl, ps = O.pairs(rp, ci, F["mb"], F["mc"], upper=not lower)
        dref = O.build_dinv(rp, ci, lu, lower)
        mref = (O.isai_lower(lu, F["mb"], F["mc"], pp, pl, ps, dref) if lower
                else O.isai_upper(rp, ci, lu, F["mb"], F["mc"], dref))
        assert np.array_equal(F["mval"], mref), rel_err(F["mval"], mref)


def test_isai_upper_lookup_form_and_small_nmax(gpu):
    """The reference-shaped ISAI(U) entry, and an Nmax smaller than the longest
    column (falls back to reading Y from Mval) give the oracle's values."""
    print("isai upper lookup form", flush=True)
    rp, ci, val = problems.generate(2, 5, seed=8)
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    lu = d_lu.to_numpy()
    nb = len(rp) - 1
    mb, mc = problems.bcsc_pattern(rp, ci, False)
    d_mb, d_mc = dev(mb), dev(mc)
    d_dinv = DeviceArray((nb, 4, 4), np.float64)
    call("isai_build_dinv", nb, 4, d_rp.ptr, d_ci.ptr, d_lu.ptr, 0, d_dinv.ptr, None)
    d_mv = DeviceArray((len(mc), 4, 4), np.float64)
    call("isai_build_upper_lookup", nb, 4, d_rp.ptr, d_ci.ptr, d_lu.ptr, d_mb.ptr, d_mc.ptr, len(mc),
         d_dinv.ptr, d_mv.ptr, 2, None)
    mref = O.isai_upper(rp, ci, lu, mb, mc, O.build_dinv(rp, ci, lu, False))
    assert np.array_equal(d_mv.to_numpy(), mref)


def test_full_size_ilu_and_isai_invariants(gpu):
    """96^3 7-point: (L U)_IJ = A_IJ on sampled stored blocks, and
    (L M_L)(i,j) = delta_ij I, (U M_U)(i,j) = delta_ij I on sampled columns."""
    print("full-size ILU + ISAI 96^3", flush=True)
    rp, ci, val = problems.generate(1, 96, seed=0)
    nb = len(rp) - 1
    d_rp, d_ci, d_lu = gpu_ilu(rp, ci, val)
    lu = d_lu.to_numpy()
    Ls, Us = split_lu(rp, ci, lu)
    rng = np.random.default_rng(0)
    rows = np.unique(np.concatenate([rng.integers(0, nb, 300), [0, nb - 1, nb // 2]]))
    sel = np.concatenate([np.arange(r * 4, r * 4 + 4) for r in rows])
    LU_rows = (Ls[sel, :] @ Us).toarray()
    A = scipy_bsr(rp, ci, val).tocsr()
    A_rows = A[sel, :].toarray()
    mask = A_rows != 0
    for idx, r in enumerate(rows):  # all 16 entries of every stored block
        for k in range(rp[r], rp[r + 1]):
            c = ci[k]
            blk_lu = LU_rows[idx * 4:idx * 4 + 4, c * 4:c * 4 + 4]
            assert np.allclose(blk_lu, val[k], rtol=1e-13, atol=1e-13 * np.abs(val[k]).max())
    del LU_rows, A_rows, mask
    f = gpu_isai(rp, ci, d_rp, d_ci, d_lu)
    cols = rows[:120]
    n = nb * 4
    for key, F, T in (("L", f["L"], Ls), ("U", f["U"], Us)):
        mb, mc, mv = F["mb"], F["mc"], F["mval"]
        for j in cols:
            S = mc[mb[j]:mb[j + 1]]
            srows = (S[:, None] * 4 + np.arange(4)[None, :]).reshape(-1)
            Msp = sp.csr_matrix((mv[mb[j]:mb[j + 1]].reshape(-1, 4).reshape(-1),
                                 (np.repeat(srows, 4), np.tile(np.arange(4), len(srows)))), shape=(n, 4))
            P = (T[srows, :] @ Msp).toarray()
            for q, t in enumerate(S):
                exp = np.eye(4) if t == j else np.zeros((4, 4))
                assert np.allclose(P[q * 4:q * 4 + 4], exp, atol=1e-12), (key, j, t)
    # BCSC -> BCSR permutation of the factor keeps every block
    trp, tci, d_bcsr = to_bcsr(rp, ci, f["L"], True)
    bcsr = d_bcsr.to_numpy()
    assert np.array_equal(np.sort(bcsr.reshape(-1)), np.sort(f["L"]["mval"].reshape(-1)))
