"""Pins the CPU oracle against mathematics, brute force and hand-worked fixtures.

The reference ships no tests or golden vectors, so each oracle function is
checked here against something that does not reuse its own formula: dense
numpy algebra, textbook dense elimination with pattern dropping, closed forms,
and a hand-worked conversion fixture.  CPU only.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import _oracle as O
from isai_kernel_amd import problems

GOLDEN = Path(__file__).resolve().parent / "golden"
RNG = np.random.default_rng(1234)


def _f(v):
    return float(v) if not isinstance(v, str) else float(v)


def load_csr2bsr_fixture():
    d = json.loads((GOLDEN / "csr2bsr_small.json").read_text())
    v = np.array([_f(x) for x in d["csr_val"]])
    bv = np.array([[[_f(x) for x in row] for row in blk] for blk in d["bsr_val"]])
    return d, v, bv


def random_bsr(nb, bs, density, rng, diag=True, dominant=True):
    """Random block pattern (diagonal always present when diag), random values."""
    mask = rng.random((nb, nb)) < density
    if diag:
        np.fill_diagonal(mask, True)
    rp = np.zeros(nb + 1, np.int32)
    cols, vals = [], []
    for i in range(nb):
        js = np.nonzero(mask[i])[0]
        rp[i + 1] = rp[i] + len(js)
        for j in js:
            cols.append(j)
            b = rng.uniform(-1, 1, (bs, bs))
            if i == j and dominant:
                b += np.eye(bs) * (2.0 * bs * (1 + mask[i].sum()))
            vals.append(b)
    return rp, np.array(cols, np.int32), np.array(vals).reshape(-1, bs, bs)


def dense_mask(rp, ci, nb):
    m = np.zeros((nb, nb), bool)
    for i in range(nb):
        m[i, ci[rp[i]:rp[i + 1]]] = True
    return m


def dense_ilu0(A, mask):
    """Textbook IKJ ILU(0) on a scalar pattern (Saad, Alg. 10.4), packed L\\U."""
    W = A.copy()
    n = A.shape[0]
    for i in range(1, n):
        for k in range(i):
            if not mask[i, k]:
                continue
            W[i, k] /= W[k, k]
            sel = mask[i, k + 1:]
            W[i, k + 1:] -= W[i, k] * W[k, k + 1:] * sel
    return W


# ------------------------------------------------------------------ csr2bsr
def test_csr2bsr_hand_fixture():
    d, v, bv = load_csr2bsr_fixture()
    brp, bci, out = O.csr2bsr(d["n"], d["bs"], d["csr_rowptr"], d["csr_colind"], v)
    assert brp.tolist() == d["bsr_rowptr"] and bci.tolist() == d["bsr_colind"]
    # bitwise, so the -0.0 of the block-opening entry is checked too
    assert out.tobytes() == bv.tobytes()


def test_csr2bsr_random_vs_dense_and_roundtrip():
    for trial in range(5):
        nb, bs = 7, [1, 2, 3, 4, 5][trial]
        rp, ci, val = random_bsr(nb, bs, 0.4, RNG)
        crp, cci, cv = O.bsr2csr(rp, ci, val)
        A = O.bsr_to_dense(rp, ci, val)
        # bsr2csr closed form: dense equality and nnz = nnzb*bs^2
        D = np.zeros_like(A)
        for r in range(nb * bs):
            D[r, cci[crp[r]:crp[r + 1]]] += cv[crp[r]:crp[r + 1]]
        assert np.array_equal(A, D) and len(cci) == len(ci) * bs * bs
        brp, bci, bv = O.csr2bsr(nb * bs, bs, crp, cci, cv)
        assert np.array_equal(brp, rp) and np.array_equal(bci, ci) and np.array_equal(bv, val)


def test_csr2bsr_duplicates_summed_and_bad_n():
    rp = np.array([0, 3, 3, 3, 3], np.int32)
    ci = np.array([1, 1, 1], np.int32)
    v = np.array([0.25, 0.5, 1.0])
    _, _, bv = O.csr2bsr(4, 2, rp, ci, v)
    assert bv[0, 0, 1] == 1.75
    with pytest.raises(ValueError):
        O.csr2bsr(5, 2, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0))


# ----------------------------------------------------------------- patterns
@pytest.mark.parametrize("lower", [True, False])
def test_patterns_vs_dense_mask(lower):
    rp, ci, _ = random_bsr(23, 1, 0.2, RNG, diag=False)
    nb = 23
    m = dense_mask(rp, ci, nb)
    tri = np.tril(m) if lower else np.triu(m)
    mb, mc = O.bcsc_pattern(rp, ci, lower)
    for j in range(nb):
        expect = set(np.nonzero(tri[:, j])[0]) | {j}
        got = mc[mb[j]:mb[j + 1]]
        assert list(got) == sorted(expect)
    trp, tci = O.tri_bcsr_pattern(rp, ci, lower)
    for i in range(nb):
        assert list(tci[trp[i]:trp[i + 1]]) == list(np.nonzero(tri[i])[0])


def test_map_is_a_permutation_and_detects_mismatch():
    rp, ci, _ = random_bsr(30, 1, 0.15, RNG)
    mb, mc = O.bcsc_pattern(rp, ci, True)
    trp, tci = O.tri_bcsr_pattern(rp, ci, True)
    st, m = O.map_bcsc2bcsr(mb, mc, trp, tci)
    assert st == 0
    assert sorted(m.tolist()) == list(range(len(tci)))
    for j in range(30):
        for k in range(mb[j], mb[j + 1]):
            p = m[k]
            assert tci[p] == j and trp[mc[k]] <= p < trp[mc[k] + 1]
    # drop one row's diagonal from the BCSR pattern -> "Pattern mismatch"
    trp2, tci2 = trp.copy(), tci.copy()
    r = 5
    keep = np.ones(len(tci2), bool)
    keep[trp2[r]:trp2[r + 1]][tci2[trp2[r]:trp2[r + 1]] == r] = False
    tci2 = tci2[keep]
    trp2[r + 1:] -= 1
    st, _ = O.map_bcsc2bcsr(mb, mc, trp2, tci2)
    assert st == -4


def test_reorder_is_map_scatter():
    src = RNG.standard_normal((9, 4, 4))
    mp = RNG.permutation(9).astype(np.int32)
    dst = O.reorder(mp, src)
    for k in range(9):
        assert np.array_equal(dst[mp[k]], src[k])


# -------------------------------------------------------------------- pairs
@pytest.mark.parametrize("upper", [False, True])
def test_pairs_brute_force(upper):
    nb = 25
    rp, ci, _ = random_bsr(nb, 1, 0.2, RNG)
    mask = dense_mask(rp, ci, nb)
    mb, mc = O.bcsc_pattern(rp, ci, not upper)
    pp, pl, ps = O.pairs(rp, ci, mb, mc, upper=upper)
    for j in range(nb):
        S = mc[mb[j]:mb[j + 1]]
        for idx, t in enumerate(S):
            got = list(zip(pl[pp[mb[j] + idx]:pp[mb[j] + idx + 1]], ps[pp[mb[j] + idx]:pp[mb[j] + idx + 1]]))
            exp = []
            rng_p = range(idx) if not upper else range(len(S) - 1, idx, -1)
            for p in rng_p:
                if mask[t, S[p]]:
                    pos = rp[t] + int(np.nonzero(ci[rp[t]:rp[t + 1]] == S[p])[0][0])
                    exp.append((p, pos))
            assert got == exp


def _stats_brute(mb, pp):
    """analyze_pair_lengths / analyze_column_q_bins by numpy: sorted arrays indexed
    at int((n-1)*p), np.digitize for the qlen bins, explicit column bins."""
    nb = len(mb) - 1
    qlen = np.diff(pp)[: mb[-1]]
    N = np.diff(mb)
    Q = np.array([qlen[mb[j]:mb[j + 1]].sum() for j in range(nb)], np.int64)
    pct = lambda v, p: int(np.sort(v)[int((len(v) - 1) * p)]) if len(v) else 0
    edges = [1, 2, 4, 8, 16, 32, 64, 128, 256]
    hist = np.bincount(np.digitize(qlen, edges), minlength=10)
    b = np.minimum(np.arange(nb, dtype=np.int64) * 10 // nb, 9)
    d = {"nb": nb, "mnnz": int(mb[-1]), "n_max": int(N.max()), "q_max": int(Q.max()),
         "qlen_max": int(qlen.max()) if len(qlen) else 0, "total_pairs": int(Q.sum()),
         "n_avg": N.sum() / nb, "q_avg": Q.sum() / nb, "qlen_avg": Q.sum() / len(qlen) if len(qlen) else 0.0,
         "qlen_hist": [int(v) for v in hist], "bin_sum_q": [int(Q[b == k].sum()) for k in range(10)],
         "bin_cols": [int((b == k).sum()) for k in range(10)],
         "bin_max_q": [int(Q[b == k].max()) if (b == k).any() else 0 for k in range(10)]}
    for pc in O.PCTS:
        d[f"n_p{pc}"], d[f"q_p{pc}"], d[f"qlen_p{pc}"] = pct(N, pc / 100), pct(Q, pc / 100), pct(qlen, pc / 100)
    return d, N, Q


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pair_stats_brute_force(seed):
    """analyze_pair.h statistics against numpy on random column lengths and pair
    counts, with qlen values spread over every histogram bin (0 .. 300)."""
    rng = np.random.default_rng(seed)
    nb = int(rng.integers(1, 400))
    N = rng.integers(0, 12, nb)
    mb = np.concatenate([[0], np.cumsum(N)]).astype(np.int32)
    qlen = rng.choice([0, 1, 2, 3, 5, 9, 17, 40, 90, 200, 300], size=mb[-1])
    pp = np.concatenate([[0], np.cumsum(qlen)]).astype(np.int32)
    got, cn, cq, cm = O.pair_stats(mb, pp)
    exp, EN, EQ = _stats_brute(mb, pp)
    assert got == exp
    assert np.array_equal(cn, EN) and np.array_equal(cq, EQ)
    assert np.array_equal(cm, [qlen[mb[j]:mb[j + 1]].max() if N[j] else 0 for j in range(nb)])


def test_pair_stats_hand_case():
    """Three columns with N = 1, 2, 0 and pair counts (3), (0, 1): counted by hand."""
    got, cn, cq, cm = O.pair_stats(np.array([0, 1, 3, 3]), np.array([0, 3, 3, 4]))
    assert (got["mnnz"], got["total_pairs"], got["n_max"], got["q_max"], got["qlen_max"]) == (3, 4, 2, 3, 3)
    assert list(cn) == [1, 2, 0] and list(cq) == [3, 1, 0] and list(cm) == [3, 1, 0]
    # sorted N = 0 1 2, Q = 0 1 3, qlen = 0 1 3: index int(2 p) is 1 for p = .5, .9, .95, .99
    assert (got["n_p50"], got["n_p99"], got["q_p50"], got["q_p90"], got["qlen_p95"]) == (1, 1, 1, 1, 1)
    assert got["qlen_hist"][:3] == [1, 1, 1] and sum(got["qlen_hist"]) == 3
    # column bins j * 10 // 3 = 0, 3, 6
    assert got["bin_cols"][0] == got["bin_cols"][3] == got["bin_cols"][6] == 1
    assert got["bin_sum_q"][0] == 3 and got["bin_sum_q"][3] == 1 and got["bin_max_q"][6] == 0
    assert got["n_avg"] == 1.0 and got["q_avg"] == 4 / 3 and got["qlen_avg"] == 4 / 3


# ------------------------------------------------------------ block inverse
def test_invert_block_and_guard():
    for bs in (1, 3, 4, 5, 7):
        A = RNG.uniform(-1, 1, (bs, bs)) + 3 * bs * np.eye(bs)
        ok, Ai = O.invert_block(A)
        assert ok and np.allclose(A @ Ai, np.eye(bs), atol=1e-13)
    ok, _ = O.invert_block(np.array([[0.0, 1.0], [1.0, 0.0]]))
    assert not ok  # zero pivot without pivoting (isai_bsr.h:69)
    # triangular closed form: inverse of unit lower [[1,0],[a,1]] is [[1,0],[-a,1]]
    ok, Ai = O.invert_block(np.array([[1.0, 0.0], [2.5, 1.0]]))
    assert ok and np.array_equal(Ai, np.array([[1.0, 0.0], [-2.5, 1.0]]))


def test_dinv_parts():
    rp, ci, val = problems.generate(1, 4, seed=3)
    lu, st, _ = O.bsrilu0(rp, ci, val)
    assert st == 0
    for lower in (True, False):
        dinv = O.build_dinv(rp, ci, lu, lower)
        for k in range(len(rp) - 1):
            pos = rp[k] + int(np.nonzero(ci[rp[k]:rp[k + 1]] == k)[0][0])
            D = lu[pos]
            part = np.tril(D, -1) + np.eye(4) if lower else np.triu(D)
            assert np.allclose(part @ dinv[k], np.eye(4), atol=1e-13)


# ------------------------------------------------------------------ ILU(0)
@pytest.mark.parametrize("density", [0.25, 1.0])
def test_ilu0_vs_dense_masked_elimination(density):
    nb, bs = 12, 4
    rp, ci, val = random_bsr(nb, bs, density, RNG)
    lu, st, zp = O.bsrilu0(rp, ci, val)
    assert st == 0 and zp == -1
    A = O.bsr_to_dense(rp, ci, val)
    mask = np.kron(dense_mask(rp, ci, nb), np.ones((bs, bs), bool))
    W = dense_ilu0(A, mask)
    LUd = O.bsr_to_dense(rp, ci, lu)
    assert np.allclose(LUd[mask], W[mask], rtol=1e-12, atol=1e-12)
    # defining invariant: (L U)_IJ = A_IJ on every stored block
    L = np.tril(LUd, -1) + np.eye(nb * bs)
    U = np.triu(LUd)
    assert np.allclose((L @ U)[mask], A[mask], rtol=1e-12, atol=1e-12)
    if density == 1.0:  # full pattern: exact LU
        assert np.allclose(L @ U, A, rtol=1e-12, atol=1e-12)


def test_ilu0_block_tridiagonal_closed_form():
    nb, bs = 6, 4
    rp = np.array([0] + [min(i + 2, nb) - max(i - 1, 0) for i in range(nb)], np.int32).cumsum().astype(np.int32)
    ci = np.array([j for i in range(nb) for j in range(max(i - 1, 0), min(i + 2, nb))], np.int32)
    val = RNG.uniform(-1, 1, (len(ci), bs, bs))
    for i in range(nb):
        val[rp[i] + (1 if i > 0 else 0)] += 10 * np.eye(bs)
    lu, st, _ = O.bsrilu0(rp, ci, val)
    assert st == 0
    A = O.bsr_to_dense(rp, ci, val)
    # D_i = A_ii - A_{i,i-1} D_{i-1}^{-1} A_{i-1,i}; L_{i,i-1} = A_{i,i-1} U_{i-1}^{-1}
    D_prev = None
    for i in range(nb):
        Aii = A[i*bs:(i+1)*bs, i*bs:(i+1)*bs]
        D = Aii if i == 0 else Aii - A[i*bs:(i+1)*bs, (i-1)*bs:i*bs] @ np.linalg.solve(
            D_prev, A[(i-1)*bs:i*bs, i*bs:(i+1)*bs])
        packed = lu[rp[i] + (1 if i > 0 else 0)]
        Lii, Uii = np.tril(packed, -1) + np.eye(bs), np.triu(packed)
        assert np.allclose(Lii @ Uii, D, rtol=1e-12, atol=1e-12)
        D_prev = D


def test_ilu0_zero_pivot_reporting():
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 0], np.int32)  # row 1 has no diagonal block: structural zero
    val = np.stack([np.eye(4), np.eye(4)])
    _, st, zp = O.bsrilu0(rp, ci, val)
    assert st == -5 and zp == 1
    rp = np.array([0, 1], np.int32)
    ci = np.array([0], np.int32)
    v = np.eye(4)
    v[2, 2] = 0.0  # numerical zero pivot in U
    _, st, zp = O.bsrilu0(rp, ci, v[None])
    assert st == -5 and zp == 0


# ------------------------------------------------------------------- ISAI
def _isai_factors(rp, ci, lu):
    nb = len(rp) - 1
    mbL, mcL = O.bcsc_pattern(rp, ci, True)
    mbU, mcU = O.bcsc_pattern(rp, ci, False)
    pp, pl, ps = O.pairs(rp, ci, mbL, mcL)
    ML = O.isai_lower(lu, mbL, mcL, pp, pl, ps, O.build_dinv(rp, ci, lu, True))
    MU = O.isai_upper(rp, ci, lu, mbU, mcU, O.build_dinv(rp, ci, lu, False))
    return (mbL, mcL, ML), (mbU, mcU, MU)


@pytest.mark.parametrize("density", [0.3, 1.0])
def test_isai_defining_invariant(density):
    nb, bs = 14, 4
    rp, ci, val = random_bsr(nb, bs, density, RNG)
    lu, st, _ = O.bsrilu0(rp, ci, val)
    (mbL, mcL, ML), (mbU, mcU, MU) = _isai_factors(rp, ci, lu)
    LUd = O.bsr_to_dense(rp, ci, lu)
    Ld = np.tril(LUd, -1) + np.eye(nb * bs)
    Ud = np.triu(LUd)
    for (mb, mc, M, F) in ((mbL, mcL, ML, Ld), (mbU, mcU, MU, Ud)):
        Md = O.bcsc_to_dense(mb, mc, M, nb)
        P = F @ Md
        for j in range(nb):
            for k in range(mb[j], mb[j + 1]):
                i = mc[k]
                blk = P[i*bs:(i+1)*bs, j*bs:(j+1)*bs]
                assert np.allclose(blk, np.eye(bs) if i == j else 0, atol=1e-12)
        if density == 1.0:  # full pattern: ISAI is the exact inverse
            assert np.allclose(Md, np.linalg.inv(F), atol=1e-12)


def test_isai_lower_bidiagonal_closed_form():
    nb, bs = 5, 4
    rp = np.array([0, 1, 3, 5, 7, 9], np.int32)
    ci = np.array([0, 0, 1, 1, 2, 2, 3, 3, 4], np.int32)
    val = RNG.uniform(-1, 1, (9, bs, bs))
    lu = val.copy()  # use values directly as a packed factor (L unit-lower part)
    (mbL, mcL, ML), _ = _isai_factors(rp, ci, lu)
    def Ldiag(i):
        return np.tril(lu[rp[i + 1] - 1], -1) + np.eye(bs)
    for j in range(nb):
        k0 = mbL[j]
        assert np.allclose(ML[k0], np.linalg.inv(Ldiag(j)), atol=1e-13)
        if j + 1 < nb:
            Lsub = lu[rp[j + 1]]  # block (j+1, j)
            exp = -np.linalg.inv(Ldiag(j + 1)) @ Lsub @ np.linalg.inv(Ldiag(j))
            assert np.allclose(ML[k0 + 1], exp, atol=1e-12)


# ------------------------------------------------------------------ levels
@pytest.mark.parametrize("problem,L,formula", [
    (0, 9, lambda x, y, z: x + y),
    (1, 6, lambda x, y, z: x + y + z),
    (3, 8, lambda x, y, z: x + 2 * y),
    (2, 5, lambda x, y, z: x + 2 * y + 4 * z),
])
def test_levels_closed_forms(problem, L, formula):
    rp, ci, _ = problems.generate(problem, L, seed=0)
    lev, n = O.levels(rp, ci, True)
    Lz = L if problem in (1, 2) else 1
    exp = np.array([formula(x, y, z) for z in range(Lz) for y in range(L) for x in range(L)])
    assert np.array_equal(lev, exp) and n == exp.max() + 1
    levU, nU = O.levels(rp, ci, False)
    assert nU == n and np.array_equal(levU, exp.max() - exp)


# --------------------------------------------------------------- SpMV/solve
def test_spmv_dense_and_reduced_precision():
    rp, ci, val = problems.generate(1, 5, seed=1)
    x = RNG.standard_normal((len(rp) - 1) * 4)
    A = O.bsr_to_dense(rp, ci, val)
    assert np.allclose(O.spmv(rp, ci, val, x), A @ x, rtol=1e-14, atol=1e-13)
    A32 = O.bsr_to_dense(rp, ci, val.astype(np.float32).astype(np.float64))
    y32 = O.spmv_f32acc(rp, ci, val, x)
    ref = A32 @ x.astype(np.float32).astype(np.float64)
    assert np.max(np.abs(y32 - ref)) <= 8 * 2**-24 * np.max(np.abs(A32) @ np.abs(x))
    # BF16 rounding: round-to-nearest-even on the top 16 bits of the float
    for v, exp in ((1.0 + 2**-8, 1.0), (1.0 + 3 * 2**-8, 1.0 + 2**-6), (-2.0 - 2**-7, -2.0), (1.00390625, 1.0)):
        assert O.ora.ora_round_bf16(v) == exp


def test_sptrsv_block_and_csr():
    rp, ci, val = problems.generate(1, 4, seed=2)
    lu, st, _ = O.bsrilu0(rp, ci, val)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    x = O.sptrsv_lu(rp, ci, lu, b)
    LUd = O.bsr_to_dense(rp, ci, lu)
    L = np.tril(LUd, -1) + np.eye(n)
    U = np.triu(LUd)
    assert np.allclose(L @ (U @ x), b, rtol=1e-12, atol=1e-12)
    crp, cci, cv = O.bsr2csr(rp, ci, lu)
    st, xc = O.csr_sptrsv_lu(crp, cci, cv, b)
    assert st == 0 and np.allclose(xc, x, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("solver", ["bicgstab", "fgmres"])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_krylov_solvers_converge_to_direct_solution(solver, kind):
    rp, ci, val = problems.generate(0, 8, seed=0)
    nb, n = len(rp) - 1, (len(rp) - 1) * 4
    b = O.spmv(rp, ci, val, np.ones(n))
    lu, _, _ = O.bsrilu0(rp, ci, val)
    (mbL, mcL, ML), (mbU, mcU, MU) = _isai_factors(rp, ci, lu)
    trpL, tciL = O.tri_bcsr_pattern(rp, ci, True)
    trpU, tciU = O.tri_bcsr_pattern(rp, ci, False)
    _, mapL = O.map_bcsc2bcsr(mbL, mcL, trpL, tciL)
    _, mapU = O.map_bcsc2bcsr(mbU, mcU, trpU, tciU)
    L = (trpL, tciL, O.reorder(mapL, ML))
    U = (trpU, tciU, O.reorder(mapU, MU))
    P = O.make_precond(kind, nb, 4, L=L, U=U, LU=(rp, ci, lu))
    fn = O.bicgstab if solver == "bicgstab" else O.fgmres
    x, st, it, rr = fn(rp, ci, val, P, b, tol=1e-10, maxit=500)
    assert st == 0 and rr <= 1.5e-10
    xd = np.linalg.solve(O.bsr_to_dense(rp, ci, val), b)
    assert np.allclose(x, xd, rtol=1e-8, atol=1e-8)
    assert np.allclose(xd, 1.0, atol=1e-10)


def _isai_precond(rp, ci, val, f32=False, bf16=False):
    nb = len(rp) - 1
    lu, _, _ = O.bsrilu0(rp, ci, val)
    (mbL, mcL, ML), (mbU, mcU, MU) = _isai_factors(rp, ci, lu)
    trpL, tciL = O.tri_bcsr_pattern(rp, ci, True)
    trpU, tciU = O.tri_bcsr_pattern(rp, ci, False)
    _, mapL = O.map_bcsc2bcsr(mbL, mcL, trpL, tciL)
    _, mapU = O.map_bcsc2bcsr(mbU, mcU, trpU, tciU)
    L = (trpL, tciL, O.reorder(mapL, ML))
    U = (trpU, tciU, O.reorder(mapU, MU))
    return O.make_precond(1, nb, 4, L=L, U=U, f32=f32, bf16=bf16)


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("problem,L", [(0, 8), (1, 4), (2, 4)])
def test_fgmres_mixed_converges_to_direct_solution(prec, problem, L):
    """The inner-outer mixed scheme reaches the FP64 answer: the outer FGMRES wins
    back what the FP32 inner solves (FP32 A, FP32 / BF16 ISAI, FP32 vectors) lose."""
    rp, ci, val = problems.generate(problem, L, seed=3)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    P = _isai_precond(rp, ci, val, f32=prec == "fp32", bf16=prec == "bf16")
    x, st, it, outer, rr = O.fgmres_mixed(rp, ci, val, P, b, m=10, tol=1e-10, maxit=400, inner_tol=1e-5)
    assert st == 0 and rr <= 1e-10
    assert 2 <= outer <= 3, outer  # a 1e-5 inner tolerance: two or three outer steps for 1e-10
    xd = np.linalg.solve(O.bsr_to_dense(rp, ci, val), b)
    assert np.max(np.abs(x - xd)) <= 1e-8 * np.max(np.abs(xd))


def test_fgmres_mixed_inner_solve_has_the_fp32_floor():
    """Twelve inner iterations in FP32 storage leave a true FP64 residual at the FP32
    floor (between 1e-8 and 1e-5); twelve FP64 BiCGSTAB iterations with the same
    preconditioner pass 1e-10.  An inner solve that forgot to round A or its vectors
    to FP32 would pass 1e-8."""
    rp, ci, val = problems.generate(1, 6, seed=4)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    P = _isai_precond(rp, ci, val, f32=True)
    x, st, it, outer, rr = O.fgmres_mixed(rp, ci, val, P, b, m=1, tol=1e-13, maxit=12, inner_tol=1e-14,
                                          inner_maxit=12)
    assert st != 0 and outer == 1 and it == 12
    assert 1e-8 < rr < 1e-5, rr
    _, _, it32, rr32 = O.bicgstab(rp, ci, val, P, b, tol=1e-16, maxit=12)
    assert it32 == 12 and rr32 <= 1e-10, rr32


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
@pytest.mark.parametrize("problem,L", [(0, 8), (1, 4), (2, 4)])
def test_fgmres_mixed_gmres_converges_to_direct_solution(prec, problem, L):
    """The inner-GMRES mixed scheme (FP32-storage GMRES cycles as the flexible
    preconditioner of FP64 FGMRES) reaches the FP64 answer of a direct solve."""
    rp, ci, val = problems.generate(problem, L, seed=3)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    P = _isai_precond(rp, ci, val, f32=prec == "fp32", bf16=prec == "bf16")
    x, st, it, outer, rr = O.fgmres_mixed(rp, ci, val, P, b, m=10, tol=1e-10, maxit=400, inner_tol=1e-5,
                                          inner_maxit=30, inner="gmres")
    assert st == 0 and rr <= 1e-10
    assert 2 <= outer <= 4, outer
    xd = np.linalg.solve(O.bsr_to_dense(rp, ci, val), b)
    assert np.max(np.abs(x - xd)) <= 1e-8 * np.max(np.abs(xd))


def _np_gmres_cycle(Ad, apply, v, k):
    """Right-preconditioned GMRES(k) from x = 0 on A z = v in FP64 numpy (Arnoldi
    with modified Gram-Schmidt and a dense least-squares solve): z = M Q_k y."""
    Q = np.zeros((k + 1, len(v)))
    H = np.zeros((k + 1, k))
    beta = np.linalg.norm(v)
    Q[0] = v / beta
    for j in range(k):
        w = Ad @ apply(Q[j])
        for i in range(j + 1):
            H[i, j] = w @ Q[i]
            w = w - H[i, j] * Q[i]
        H[j + 1, j] = np.linalg.norm(w)
        Q[j + 1] = w / H[j + 1, j]
    e1 = np.zeros(k + 1)
    e1[0] = beta
    y = np.linalg.lstsq(H, e1, rcond=None)[0]
    return apply(Q[:k].T @ y)


@pytest.mark.parametrize("k", [3, 5])
def test_fgmres_mixed_gmres_inner_cycle_matches_numpy_gmres(k):
    """One outer step with one inner cycle of k steps: the outer residual is
    min_a ||b - a A z|| for the inner solution z, so it equals that of FP64 GMRES(k)
    written independently in numpy with the same preconditioner, up to the FP32
    storage of the inner cycle (1e-3 of the residual; a dropped Givens term, a
    wrong basis index or a missing Gram-Schmidt pass changes it by far more)."""
    rp, ci, val = problems.generate(2, 8, seed=5)
    n = (len(rp) - 1) * 4
    b = RNG.standard_normal(n)
    P = _isai_precond(rp, ci, val, f32=True)
    _, st, it, outer, rr = O.fgmres_mixed(rp, ci, val, P, b, m=1, tol=1e-14, maxit=k, inner_tol=1e-14,
                                          inner_maxit=k, inner="gmres")
    assert outer == 1 and it == k
    Ad = O.bsr_to_dense(rp, ci, val)
    z = _np_gmres_cycle(Ad, lambda v: O.precond_apply(P, v), b / np.linalg.norm(b), k)
    Az = Ad @ z
    a = (Az @ b) / (Az @ Az)
    rr_np = np.linalg.norm(b - a * Az) / np.linalg.norm(b)
    # the cycle made progress and stays far above the FP32 floor (about 1e-7),
    # so the comparison has content
    assert 1e-4 < rr_np < 0.5
    assert abs(rr - rr_np) <= 1e-3 * rr_np, (rr, rr_np)
    # one step fewer is distinguishably worse, so k is pinned too
    z1 = _np_gmres_cycle(Ad, lambda v: O.precond_apply(P, v), b / np.linalg.norm(b), k - 1)
    Az1 = Ad @ z1
    rr1 = np.linalg.norm(b - ((Az1 @ b) / (Az1 @ Az1)) * Az1) / np.linalg.norm(b)
    assert rr1 > rr * 1.01


def test_differentiation_config_matrix_as_specified_is_indefinite():
    """The differentiation form's matrix as its text specifies it (2-D 9-point
    stencil with the coefficients of the 5-point plain case: diagonal 4.5 I + S
    against eight neighbours -I U(0.9, 1.1)) is indefinite: at L = 24 the real
    parts of its eigenvalues span about -3.41 .. 8.47.  The oracle's
    ISAI-preconditioned BiCGSTAB and FGMRES(30) do not converge on it, so x = A^-1 b
    and its gradient are not computable by the specified method; the form runs on
    the same stencil with the batching form's coefficients (diagonal 8.5 I), whose
    eigenvalues all have positive real parts and where the oracle converges."""
    rp, ci, val = problems.generate(4, 24, seed=0)
    ev = np.linalg.eigvals(O.bsr_to_dense(rp, ci, val)).real
    assert -3.5 < ev.min() < -3.3 and 8.4 < ev.max() < 8.55 and (ev < 0).sum() > 100
    _, L, U = O.isai_factors(rp, ci, val)
    P = O.make_precond(1, len(rp) - 1, 4, L=L, U=U)
    b = problems.normal(1, len(ev))
    _, st, it, rr = O.bicgstab(rp, ci, val, P, b, tol=1e-12, maxit=500)
    assert st != 0 and rr > 1.0
    _, st, it, rr = O.fgmres(rp, ci, val, P, b, m=30, tol=1e-12, maxit=500)
    assert st != 0 and rr > 0.5
    rp3, ci3, val3 = problems.generate(3, 24, seed=0)
    ev3 = np.linalg.eigvals(O.bsr_to_dense(rp3, ci3, val3)).real
    assert ev3.min() > 0.4
    _, L3, U3 = O.isai_factors(rp3, ci3, val3)
    P3 = O.make_precond(1, len(rp3) - 1, 4, L=L3, U=U3)
    _, st, it, rr = O.bicgstab(rp3, ci3, val3, P3, b, tol=1e-12, maxit=500)
    assert st == 0 and rr <= 1e-12


# ---------------------------------------------------------- differentiation form
def test_bsr_transpose_dense():
    nb, bs = 14, 4
    rp, ci, val = random_bsr(nb, bs, 0.3, RNG)
    trp, tci, tval = O.bsr_transpose(rp, ci, val)
    assert np.array_equal(O.bsr_to_dense(trp, tci, tval), O.bsr_to_dense(rp, ci, val).T)
    for i in range(nb):
        assert np.all(np.diff(tci[trp[i]:trp[i + 1]]) > 0)


@pytest.mark.parametrize("density", [0.3, 1.0])
def test_transpose_packed_lu_identity(density):
    """L'U' = (LU)^T with L' unit lower, U' upper (dense algebra)."""
    nb, bs = 10, 4
    rp, ci, val = random_bsr(nb, bs, density, RNG)
    lu, st, _ = O.bsrilu0(rp, ci, val)
    assert st == 0
    trp, tci, lut = O.transpose_packed_lu(rp, ci, lu)
    LUd, LUt = O.bsr_to_dense(rp, ci, lu), O.bsr_to_dense(trp, tci, lut)
    L, U = np.tril(LUd, -1) + np.eye(nb * bs), np.triu(LUd)
    L2, U2 = np.tril(LUt, -1) + np.eye(nb * bs), np.triu(LUt)
    assert np.allclose(L2 @ U2, (L @ U).T, rtol=1e-12, atol=1e-12)


def test_transpose_packed_lu_equals_ilu0_of_transpose():
    """On a structurally symmetric pattern ILU(0) is unique, so refactoring A^T
    gives the derived factor."""
    rp, ci, val = problems.generate(3, 7, seed=5)
    lu, st, _ = O.bsrilu0(rp, ci, val)
    trp, tci, lut = O.transpose_packed_lu(rp, ci, lu)
    _, _, tval = O.bsr_transpose(rp, ci, val)
    lu_t, st2, _ = O.bsrilu0(trp, tci, tval)
    assert st == 0 and st2 == 0
    assert np.allclose(lut, lu_t, rtol=1e-11, atol=1e-12)


def test_outer_grad_dense():
    nb, bs = 9, 4
    rp, ci, val = random_bsr(nb, bs, 0.4, RNG)
    lam, x = RNG.standard_normal(nb * bs), RNG.standard_normal(nb * bs)
    g = O.outer_grad(rp, ci, lam, x, -1.0)
    mask = np.kron(dense_mask(rp, ci, nb), np.ones((bs, bs), bool))
    G = O.bsr_to_dense(rp, ci, g)
    assert np.array_equal(G[mask], (-np.outer(lam, x))[mask]) and not G[~mask].any()
