"""CPU tests of the C-ABI library: it loads, exports every declared entry point,
and its host-side format code matches the oracle bit for bit."""
import ctypes
import json
from pathlib import Path

import numpy as np
import pytest

import _oracle as O
from isai_kernel_amd import IsaiError, lib, problems
from isai_kernel_amd._ffi import SIGNATURES, declared_symbols

ROOT = Path(__file__).resolve().parent.parent
RNG = np.random.default_rng(99)


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_symbols(ROOT / "include")
    assert len(names) > 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    unbound = [n for n in names if n not in SIGNATURES]
    assert not unbound, f"declared but without a ctypes signature: {unbound}"
    stale = [n for n in SIGNATURES if n not in names]
    assert not stale, f"bound but not declared in include/: {stale}"


def test_error_path_reports_message():
    nb, nnzb = ctypes.c_int(), ctypes.c_int()
    rp = np.zeros(6, np.int32)
    st = lib.isai_csr_to_bsr_count(5, 4, rp.ctypes.data, rp.ctypes.data, ctypes.byref(nb), ctypes.byref(nnzb))
    assert st == -1
    assert b"not a multiple" in lib.isai_last_error()
    with pytest.raises(IsaiError):
        problems.csr_to_bsr(5, 4, rp, np.zeros(0, np.int32), np.zeros(0))


def test_csr2bsr_hand_fixture_bitexact():
    d = json.loads((ROOT / "tests/golden/csr2bsr_small.json").read_text())
    v = np.array([float(x) for x in d["csr_val"]])
    exp = np.array([[[float(x) for x in r] for r in b] for b in d["bsr_val"]])
    brp, bci, bv = problems.csr_to_bsr(d["n"], d["bs"], np.array(d["csr_rowptr"]), np.array(d["csr_colind"]), v)
    assert brp.tolist() == d["bsr_rowptr"] and bci.tolist() == d["bsr_colind"]
    assert bv.tobytes() == exp.tobytes()


@pytest.mark.parametrize("problem,L", [(0, 16), (1, 7), (2, 5), (3, 11)])
def test_csr2bsr_matches_oracle_on_generated(problem, L):
    rp, ci, val = problems.generate(problem, L, seed=7)
    crp, cci, cv = O.bsr2csr(rp, ci, val)
    # shuffle entries within each CSR row and duplicate a few (summed by both)
    n = len(crp) - 1
    rows = np.repeat(np.arange(n), np.diff(crp))
    dup = RNG.random(len(cci)) < 0.05
    rows2 = np.concatenate([rows, rows[dup]])
    cols2 = np.concatenate([cci, cci[dup]])
    vals2 = np.concatenate([cv, 0.125 * cv[dup]])
    order = np.lexsort((RNG.random(len(rows2)), rows2))
    rows2, cols2, vals2 = rows2[order], cols2[order], vals2[order]
    rp2 = np.zeros(n + 1, np.int32)
    np.add.at(rp2, rows2 + 1, 1)
    rp2 = np.cumsum(rp2).astype(np.int32)
    a = problems.csr_to_bsr(n, 4, rp2, cols2.astype(np.int32), vals2)
    b = O.csr2bsr(n, 4, rp2, cols2, vals2)
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()


@pytest.mark.parametrize("lower", [True, False])
def test_patterns_and_map_match_oracle(lower):
    rp, ci, _ = problems.generate(2, 6, seed=1)
    mb, mc = problems.bcsc_pattern(rp, ci, lower)
    omb, omc = O.bcsc_pattern(rp, ci, lower)
    assert np.array_equal(mb, omb) and np.array_equal(mc, omc)
    trp, tci = problems.tri_bcsr_pattern(rp, ci, lower)
    otrp, otci = O.tri_bcsr_pattern(rp, ci, lower)
    assert np.array_equal(trp, otrp) and np.array_equal(tci, otci)
    assert len(tci) == len(mc)  # main.cu:537-546 count check
    m = problems.map_bcsc2bcsr(mb, mc, trp, tci)
    st, om = O.map_bcsc2bcsr(mb, mc, trp, tci)
    assert st == 0 and np.array_equal(m, om)


def test_map_mismatch_is_an_error():
    rp, ci, _ = problems.generate(0, 5, seed=1)
    mb, mc = problems.bcsc_pattern(rp, ci, True)
    trp, tci = problems.tri_bcsr_pattern(rp, ci, True)
    tci = tci.copy()
    tci[trp[3]:trp[4]] = 0  # corrupt one row
    with pytest.raises(IsaiError) as e:
        problems.map_bcsc2bcsr(mb, mc, trp, tci)
    assert e.value.status == -4 and "Pattern mismatch" in str(e.value)


def test_generator_sizes_match_closed_forms():
    # cfg1: 96^3 7-pt -> nnzb = n + 6 * 96^2 * 95
    nb, nnzb = ctypes.c_int(), ctypes.c_int()
    for prob, L, exp_nb, exp_nnzb in [
        (0, 0, 4096, 4096 + 4 * 64 * 63),
        (1, 0, 96**3, 96**3 + 6 * 96 * 96 * 95),
        (2, 0, 64**3, (3 * 64 - 2) ** 3),
        (4, 0, 65536, (3 * 256 - 2) ** 2),
        (3, 24, 576, (3 * 24 - 2) ** 2),
    ]:
        assert lib.isai_gen_problem_count(prob, L, ctypes.byref(nb), ctypes.byref(nnzb)) == 0
        assert (nb.value, nnzb.value) == (exp_nb, exp_nnzb)


def test_generator_coefficients():
    rp, ci, val = problems.generate(1, 5, seed=0)
    nb = len(rp) - 1
    for i in range(nb):
        for k in range(rp[i], rp[i + 1]):
            d = val[k] - (6.5 if ci[k] == i else -1.0) * np.eye(4)
            assert np.all(np.abs(d) <= (0.1 if ci[k] == i else 0.05))
    rp, ci, val = problems.generate(0, 6, seed=0)
    for i in range(len(rp) - 1):
        for k in range(rp[i], rp[i + 1]):
            if ci[k] == i:
                assert np.array_equal(val[k], val[k].T) and np.all(np.abs(val[k] - 4.5 * np.eye(4)) < 0.1)
            else:
                u = -val[k][0, 0]
                assert 0.9 <= u <= 1.1 and np.array_equal(val[k], -u * np.eye(4))
    # same seed -> same matrix; other seed -> different values
    a = problems.generate(2, 4, seed=5)[2]
    assert np.array_equal(a, problems.generate(2, 4, seed=5)[2])
    assert not np.array_equal(a, problems.generate(2, 4, seed=6)[2])


@pytest.mark.parametrize("lower", [True, False])
def test_levels_match_oracle(lower):
    rp, ci, _ = problems.generate(2, 7, seed=0)
    lev = np.empty(len(rp) - 1, np.int32)
    n = ctypes.c_int()
    assert lib.isai_bsr_levels(len(rp) - 1, rp.ctypes.data, ci.ctypes.data, int(lower), lev.ctypes.data,
                               ctypes.byref(n)) == 0
    olev, on = O.levels(rp, ci, lower)
    assert np.array_equal(lev, olev) and n.value == on


def test_matrix_market_reader_and_residual():
    """Reader semantics (1-based -> 0-based, sorted by (row, col), duplicates kept
    in file order) on a hand-written file; csr2bsr of it sums the duplicate."""
    path = str(ROOT / "tests/golden/small.mtx").encode()
    nr, nc, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    prp, pci, pv = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    assert lib.isai_read_matrix_market(path, ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(nnz),
                                       ctypes.byref(prp), ctypes.byref(pci), ctypes.byref(pv)) == 0
    assert (nr.value, nc.value, nnz.value) == (8, 8, 13)
    rp = np.ctypeslib.as_array(ctypes.cast(prp, ctypes.POINTER(ctypes.c_int)), (9,)).copy()
    ci = np.ctypeslib.as_array(ctypes.cast(pci, ctypes.POINTER(ctypes.c_int)), (13,)).copy()
    v = np.ctypeslib.as_array(ctypes.cast(pv, ctypes.POINTER(ctypes.c_double)), (13,)).copy()
    for p_ in (prp, pci, pv):
        lib.isai_host_free(p_)
    assert rp.tolist() == [0, 2, 4, 5, 6, 8, 9, 11, 13]
    assert ci.tolist() == [0, 5, 1, 1, 6, 0, 3, 4, 5, 2, 6, 0, 7]
    assert v[2:4].tolist() == [3.0, 0.5] and np.signbit(v[6]) and v[6] == 0.0
    brp, bci, bv = problems.csr_to_bsr(8, 4, rp, ci, v)
    assert bv[0, 1, 1] == 3.5 and np.signbit(bv[2, 0, 3])
    # residual check: exact solution of a diagonal system
    drp = np.arange(5, dtype=np.int32)
    dci = np.arange(4, dtype=np.int32)
    dv = np.array([2.0, 4.0, 8.0, 16.0])
    b = np.ones(4)
    x = b / dv
    ex, rr = ctypes.c_double(), ctypes.c_double()
    assert lib.isai_rel_errors_csr(4, drp.ctypes.data, dci.ctypes.data, dv.ctypes.data, b.ctypes.data,
                                   x.ctypes.data, x.ctypes.data, ctypes.byref(ex), ctypes.byref(rr)) == 0
    assert ex.value == 0.0 and rr.value == 0.0
    assert lib.isai_read_matrix_market(b"/nonexistent.mtx", ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(nnz),
                                       ctypes.byref(prp), ctypes.byref(pci), ctypes.byref(pv)) == -10


def test_driver_usage_without_arguments():
    """The C driver (built with the library) parses its arguments before touching
    the GPU: without a matrix it prints its usage and exits."""
    import subprocess
    drv = ROOT / "isai_kernel_amd" / "bin" / "isai_driver"
    if not drv.exists():
        pytest.skip("driver not built")
    p = subprocess.run([str(drv)], capture_output=True, text=True, timeout=30)
    assert p.returncode == 0 and "Usage" in p.stderr
    p = subprocess.run([str(drv), "--bogus"], capture_output=True, text=True, timeout=30)
    assert p.returncode == 2 and "unknown option" in p.stderr
