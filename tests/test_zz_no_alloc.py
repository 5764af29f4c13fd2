"""tlbm_step_noalloc (the entry point that allocates nothing) against tlbm_step, through the
C ABI: the same populations, report fields and FP64 sums, bit for bit, on two calls in a row
(the second on a smaller lattice, so the shared partials array holds the first call's data),
and a clear refusal of a report step larger than its partials array."""

import ctypes

import numpy as np
import pytest

from cases import Case, random_fields, random_populations

pytestmark = pytest.mark.gpu


def _both(case, storage, seed, block_x=0):
    """One report step through tlbm_step and one through tlbm_step_noalloc from the same
    inputs; returns ((populations, fields, sums) of each)."""
    import tlbm
    from tlbm import _capi
    L = _capi.lib()
    lat, phys, geo, flow = case.device()
    raw = tlbm.encode_populations(random_populations(case.n, seed).astype(np.float32), storage)
    old = random_fields(case.n, seed + 1)
    h = tlbm.kernel_handle(phys.code, geo.code, _capi.REPORT, flow.side.code,
                           tlbm.STORAGES[storage])
    p = tlbm.make_params(lat, phys, geo, flow)
    p.block_x = block_x
    out = []
    for fn in (L.tlbm_step, L.tlbm_step_noalloc):
        f_in = tlbm.array(raw, storage)
        f_out = tlbm.empty((19 * case.n,), storage)
        fields = {k: tlbm.array(old[k]) for k in tlbm.FIELDS}
        fs = _capi.Fields(*[fields[k].ptr for k in _capi.FIELD_NAMES])
        sums = tlbm.HostBuffer(_capi.TLBM_NSUMS)
        sums.numpy[:] = np.nan
        _capi.check(fn(h, f_in.ptr, f_out.ptr, ctypes.byref(p), ctypes.byref(fs), sums.ptr, None),
                    "step")
        tlbm.sync()
        out.append((f_out.numpy(), {k: fields[k].numpy() for k in tlbm.FIELDS},
                    sums.numpy.copy()))
    return out


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_noalloc_step_matches_the_allocating_step_twice(storage):
    big = Case((160, 48, 40), "les", "sphere", "lid", geometry=(40.3, 22.2, 19.4, 6.1))
    small = Case((48, 20, 12), "trt", "cylinder", "periodic")
    for case, seed in ((big, 5), (small, 9)):
        (f_a, fl_a, s_a), (f_b, fl_b, s_b) = _both(case, storage, seed)
        assert np.array_equal(f_a, f_b)
        for k in fl_a:
            assert np.array_equal(fl_a[k], fl_b[k]), k
        assert np.array_equal(s_a, s_b) and np.isfinite(s_b).all()


def test_noalloc_refuses_a_report_step_beyond_its_partials():
    """A lattice of 33,554,433 cells at the default 256 threads per block needs one block more
    than the array holds: refused before any launch (the buffers are never touched)."""
    import tlbm
    from tlbm import _capi
    L = _capi.lib()
    lat = tlbm.D3Q19(33554433, 1, 1)
    case = Case((24, 12, 9))
    _, phys, geo, flow = case.device()
    h = tlbm.kernel_handle(phys.code, geo.code, _capi.REPORT, flow.side.code, _capi.F32)
    p = tlbm.make_params(lat, phys, geo, flow)
    dummy = tlbm.empty((64,))
    other = tlbm.empty((64,))
    sums = tlbm.HostBuffer(_capi.TLBM_NSUMS)
    rc = L.tlbm_step_noalloc(h, dummy.ptr, other.ptr, ctypes.byref(p), None, sums.ptr, None)
    assert rc != 0 and "131072 blocks" in _capi.last_error()
