"""The tiled population layout on the host: the index formula the header states, the
tile/untile converters, buffer lengths (Python against the C entry point) and the argument
checks of the entry points that take q-major planes only.  No GPU."""

import ctypes

import numpy as np
import pytest


def _lat(nx, ny, nz):
    import tlbm
    return tlbm.D3Q19(nx, ny, nz)


def _params(dims, qstride, tile_rows):
    from tlbm import _capi
    p = _capi.Params()
    p.nx, p.ny, p.nz = dims
    p.qstride, p.tile_rows = qstride, tile_rows
    p.obs[0], p.obs[1], p.obs[2] = 2.0, 2.0, 1.0
    return p


@pytest.mark.parametrize("nq", [19, 7])
@pytest.mark.parametrize("dims,k", [((8, 3, 4), 1), ((8, 3, 4), 4), ((8, 3, 4), 8),
                                    ((5, 7, 3), 2), ((64, 6, 5), 16)])
def test_tile_index_formula_and_round_trip(dims, k, nq):
    """Population q of cell (x, r = z ny + y) sits at (r // k) Q k nx + q k nx + (r % k) nx + x
    for a buffer of Q planes (19 populations, or the 7 temperature planes), checked element by
    element against an independent loop; untile inverts tile."""
    import tlbm
    lat = _lat(*dims)
    nx, ny, nz = dims
    rng = np.random.default_rng(7)
    planes = rng.standard_normal((nq, lat.n)).astype(np.float32)
    flat = tlbm.tile_populations(planes, lat, k)
    assert flat.size == tlbm.population_elems(lat, None, k, nq)
    for q in range(nq):
        for z in range(nz):
            for y in range(ny):
                r = z * ny + y
                for x in range(nx):
                    addr = (r // k) * nq * k * nx + q * k * nx + (r % k) * nx + x
                    assert flat[addr] == planes[q, r * nx + x]
    assert np.array_equal(tlbm.untile_populations(flat, lat, k, nq), planes)


def test_population_elems_matches_the_library():
    import tlbm
    from tlbm import _capi
    lib = _capi.lib()
    for dims, qstride, k in [((8, 3, 4), 96, 0), ((8, 3, 4), 200, 0), ((8, 3, 4), 96, 4),
                             ((512, 256, 256), None, 4), ((96, 7, 5), 3360, 8),
                             ((384, 128, 64), None, 32)]:
        lat = _lat(*dims)
        p = _params(dims, lat.n if qstride is None else qstride, k)
        got = ctypes.c_uint64()
        assert lib.tlbm_population_elems(ctypes.byref(p), ctypes.byref(got)) == 0, \
            _capi.last_error()
        assert got.value == tlbm.population_elems(lat, qstride, k)
    for bad in (3, -1, 1 << 21):
        p.tile_rows = bad
        assert lib.tlbm_population_elems(ctypes.byref(p), ctypes.byref(got)) != 0
        assert "tile_rows" in _capi.last_error()
    p.tile_rows = 4
    for bad in (3, -2, 128):
        p.xcd_group = bad
        assert lib.tlbm_population_elems(ctypes.byref(p), ctypes.byref(got)) != 0
        assert "xcd_group" in _capi.last_error()


def test_default_tile_rows():
    """Tiles whose plane segment holds at most 32 KiB at FP32 and 8 KiB at FP16/BF16, only for
    lattices of 2^20 cells or more with rows of whole waves."""
    from tlbm import default_tile_rows
    assert default_tile_rows(512, 256, 256) == 16
    assert default_tile_rows(512, 256, 256, "f16") == 8
    assert default_tile_rows(512, 256, 256, "bf16") == 8
    assert default_tile_rows(384, 256, 256) == 16
    assert default_tile_rows(256, 256, 256) == 32
    assert default_tile_rows(256, 64, 64) == 32
    assert default_tile_rows(1024, 256, 128) == 8
    assert default_tile_rows(1024, 256, 128, "f16") == 4
    assert default_tile_rows(2048, 32, 32) == 4
    assert default_tile_rows(8192, 16, 16) == 1
    assert default_tile_rows(16384, 8, 8) == 0           # a row beyond 32 KiB at FP32
    assert default_tile_rows(4096, 32, 32, "f16") == 1
    assert default_tile_rows(8192, 16, 16, "f16") == 0   # a row beyond 8 KiB at FP16
    assert default_tile_rows(480, 256, 256) == 0         # rows not whole waves
    assert default_tile_rows(128, 64, 64) == 0           # below 2^20 cells


def test_adjoint_rejects_tiles():
    """The adjoint takes q-major planes: a tiled params is refused before any launch."""
    from tlbm import _capi
    lib = _capi.lib()
    p = _params((8, 6, 4), 192, 2)
    p.tau0 = 0.56
    a, b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    org = (ctypes.c_int32 * 3)(0, 0, 0)
    size = (ctypes.c_int32 * 3)(1, 1, 1)
    p.phi = 0x3000
    assert lib.tlbm_adjoint_step(ctypes.byref(p), a, b, a, b, 1.0, 1.0, org, size,
                                 ctypes.c_void_p(0x4000), 0, None) != 0
    assert "q-major" in _capi.last_error()


def test_batch_tile_rows():
    """Batched domains: tiles of 16 KiB per plane at FP32 (the largest power of two k with
    k nx <= 4096), q-major planes for 16-bit storage and for rows longer than a tile."""
    from tlbm.batch import batch_tile_rows
    assert [batch_tile_rows(nx) for nx in (64, 100, 128, 130, 192, 256, 4096, 5000)] == \
        [64, 32, 32, 16, 16, 16, 1, 0]
    for nx in (64, 100, 4096):
        k = batch_tile_rows(nx)
        assert k * nx <= 4096 < 2 * k * nx
    assert batch_tile_rows(128, "f16") == batch_tile_rows(128, "bf16") == 0
