"""isai_pairs_count_noalloc (the pair count whose scan scratch is a fixed device
array of the library, for callers that cannot allocate) against isai_pairs_count
through the C ABI: PairPtr and the total must be bit-identical on two calls in a
row, the second on a smaller pattern, so that scratch left by the first call
cannot leak into the second."""
import ctypes

import numpy as np
import pytest

from isai_kernel_amd import DeviceArray, call, problems
from isai_kernel_amd._ffi import IsaiError

pytestmark = pytest.mark.gpu

ISAI_ERR_CAPACITY = -8


def _stencil7_pattern(L):
    """Block CSR pattern of the 7-point stencil on an L^3 grid, natural order."""
    nb = L ** 3
    r = np.arange(nb, dtype=np.int64)
    i, j, k = r % L, r // L % L, r // (L * L)
    offs = np.array([-L * L, -L, -1, 0, 1, L, L * L], np.int64)
    ok = np.stack([k > 0, j > 0, i > 0, np.ones(nb, bool), i < L - 1, j < L - 1, k < L - 1], axis=1)
    cols = r[:, None] + offs[None, :]
    rp = np.concatenate([[0], np.cumsum(ok.sum(axis=1))]).astype(np.int32)
    ci = cols[ok].astype(np.int32)
    return rp, ci


def _count(name, d_rp, d_ci, nb, mb, mc, upper):
    d_mb, d_mc = DeviceArray.from_numpy(mb), DeviceArray.from_numpy(mc)
    d_pp = DeviceArray((len(mc) + 1,), np.int32)
    total = ctypes.c_int(-1)
    call(name, nb, len(mc), d_rp.ptr, d_ci.ptr, d_mb.ptr, d_mc.ptr, int(upper), d_pp.ptr, ctypes.byref(total),
         None)
    return d_pp.to_numpy(), total.value


def _compare(L, upper):
    rp, ci = _stencil7_pattern(L)
    mb, mc = problems.bcsc_pattern(rp, ci, not upper)
    nb = len(rp) - 1
    d_rp, d_ci = DeviceArray.from_numpy(rp), DeviceArray.from_numpy(ci)
    print(f"pairs_count_noalloc L={L} upper={upper} nb={nb} Mnnz={len(mc)}", flush=True)
    ref_pp, ref_total = _count("isai_pairs_count", d_rp, d_ci, nb, mb, mc, upper)
    pp, total = _count("isai_pairs_count_noalloc", d_rp, d_ci, nb, mb, mc, upper)
    assert total == ref_total
    assert np.array_equal(pp, ref_pp)
    assert pp[-1] == total and pp[0] == 0 and np.all(np.diff(pp) >= 0)
    return total


def test_noalloc_pair_count_matches_allocating_count():
    # 104^3 lower: Mnnz + 1 = 4.4M entries, so the scan recurses twice (2178 tiles,
    # then 2); then a small upper pattern whose scan is a single tile.
    big = _compare(104, 0)
    assert big > 0
    small = _compare(6, 1)
    assert 0 < small < big


def test_noalloc_pair_count_rejects_oversize():
    rp, ci = _stencil7_pattern(2)
    d_rp, d_ci = DeviceArray.from_numpy(rp), DeviceArray.from_numpy(ci)
    d_pp = DeviceArray((16,), np.int32)
    total = ctypes.c_int(-1)
    with pytest.raises(IsaiError) as err:
        call("isai_pairs_count_noalloc", len(rp) - 1, (1 << 28), d_rp.ptr, d_ci.ptr, d_rp.ptr, d_ci.ptr, 0,
             d_pp.ptr, ctypes.byref(total), None)
    assert err.value.status == ISAI_ERR_CAPACITY
    assert "scratch limit" in str(err.value)
