"""The C ABI library loads and exports what include/tlbm.h declares (no GPU calls)."""

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tlbm.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(tlbm_\w+)\s*\(", text, re.M)))


def test_header_declares_the_surface():
    names = _declared()
    for must in ("tlbm_alloc_managed", "tlbm_kernel_lookup", "tlbm_init", "tlbm_step",
                 "tlbm_coarsen", "tlbm_last_error", "tlbm_stream_sync"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from tlbm import _capi
    lib = _capi.lib()
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_declared()) == set(_capi.SIGNATURES), "ctypes table and header disagree"


def test_struct_layout_matches_the_c_compiler(tmp_path):
    """ctypes structs against offsetof/sizeof from gcc on include/tlbm.h."""
    from tlbm import _capi
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "tlbm.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(tlbm_params), offsetof(tlbm_params, qstride),
         offsetof(tlbm_params, obs), offsetof(tlbm_params, lid), offsetof(tlbm_params, omega_rot),
         sizeof(tlbm_spec), sizeof(tlbm_fields), offsetof(tlbm_params, tile_rows));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    P = _capi.Params
    expect = [ctypes.sizeof(P), P.qstride.offset, P.obs.offset, P.lid.offset, P.omega_rot.offset,
              ctypes.sizeof(_capi.Spec), ctypes.sizeof(_capi.Fields), P.tile_rows.offset]
    assert got == expect


def test_kernel_lookup_covers_every_combination():
    """3 storages x 3 collisions x 5 obstacles x 3 sides x 2 diagnostics, plus the temperature
    variant of every FP32 combination: distinct handles."""
    from tlbm import _capi
    lib = _capi.lib()
    seen = set()
    for heat in range(2):
        for st in range(1 if heat else 3):
            for c in range(3):
                for o in range(5):
                    for s in range(3):
                        for d in range(2):
                            h = ctypes.c_int(-1)
                            spec = _capi.Spec(c, o, d, s, st, heat)
                            assert lib.tlbm_kernel_lookup(ctypes.byref(spec),
                                                          ctypes.byref(h)) == 0
                            seen.add(h.value)
    assert len(seen) == 270 + 90


@pytest.mark.parametrize("bad", [(3, 0, 0, 0, 0, 0), (0, -1, 0, 0, 0, 0), (0, 5, 0, 0, 0, 0),
                                 (0, 0, 2, 0, 0, 0), (0, 0, 0, 3, 0, 0), (0, 0, 0, 0, 3, 0),
                                 (0, 0, 0, 0, 0, 2), (0, 0, 0, 0, 1, 1)])
def test_kernel_lookup_rejects_unknown_specs(bad):
    from tlbm import _capi
    lib = _capi.lib()
    h = ctypes.c_int(-1)
    assert lib.tlbm_kernel_lookup(ctypes.byref(_capi.Spec(*bad)), ctypes.byref(h)) != 0
    assert _capi.last_error()


def test_heat_and_flow_entry_points_reject_each_others_handles():
    from tlbm import _capi
    lib = _capi.lib()
    flow, heat = ctypes.c_int(-1), ctypes.c_int(-1)
    lib.tlbm_kernel_lookup(ctypes.byref(_capi.Spec(0, 4, 0, 0, 0, 0)), ctypes.byref(flow))
    lib.tlbm_kernel_lookup(ctypes.byref(_capi.Spec(0, 4, 0, 0, 0, 1)), ctypes.byref(heat))
    p = _capi.Params()
    p.nx, p.ny, p.nz, p.qstride = 8, 6, 4, 192
    p.obs[6], p.obs[7] = 1.0, 3.0
    a, b, c, d = (ctypes.c_void_p(v) for v in (64, 128, 256, 512))
    assert lib.tlbm_step(heat, a, b, ctypes.byref(p), None, None, None) != 0
    assert "tlbm_step_heat" in _capi.last_error()
    ht = _capi.Heat(None, None, None, 0.01, 0.01, 0.85)
    assert lib.tlbm_step_heat(flow, a, b, c, d, ctypes.byref(p), ctypes.byref(ht), None, None,
                              None) != 0
    assert lib.tlbm_step_heat(heat, a, b, c, d, ctypes.byref(p), ctypes.byref(ht), None, None,
                              None) != 0                                     # no t_solid
    assert "t_solid" in _capi.last_error()
    ht.t_solid = 1024
    assert lib.tlbm_step_heat(heat, a, b, c, c, ctypes.byref(p), ctypes.byref(ht), None, None,
                              None) != 0                                     # g_in == g_out
    p.obs[7] = 0.0
    assert lib.tlbm_step_heat(heat, a, b, c, d, ctypes.byref(p), ctypes.byref(ht), None, None,
                              None) != 0                                     # fin pitch 0


def test_step_argument_errors_are_reported_before_any_launch():
    from tlbm import _capi
    lib = _capi.lib()
    h = ctypes.c_int(-1)
    lib.tlbm_kernel_lookup(ctypes.byref(_capi.Spec(0, 0, 0, 0, 0)), ctypes.byref(h))
    p = _capi.Params()
    p.nx, p.ny, p.nz, p.qstride = 4, 4, 4, 64
    fake_a, fake_b = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    assert lib.tlbm_step(h, fake_a, fake_a, ctypes.byref(p), None, None, None) != 0
    assert "distinct" in _capi.last_error()
    p.nx = 0
    assert lib.tlbm_step(h, fake_a, fake_b, ctypes.byref(p), None, None, None) != 0
    assert "positive" in _capi.last_error()
    p.nx, p.qstride = 4, 63
    assert lib.tlbm_step(h, fake_a, fake_b, ctypes.byref(p), None, None, None) != 0
    assert "qstride" in _capi.last_error()
    p.qstride = 64
    assert lib.tlbm_step(12345, fake_a, fake_b, ctypes.byref(p), None, None, None) != 0
    rep = ctypes.c_int(-1)
    lib.tlbm_kernel_lookup(ctypes.byref(_capi.Spec(0, 0, 1, 0, 0)), ctypes.byref(rep))
    assert lib.tlbm_step(rep, fake_a, fake_b, ctypes.byref(p), None, None, None) != 0
    assert "report" in _capi.last_error()
    assert lib.tlbm_set_launch(h, 96) != 0
    assert lib.tlbm_set_launch(12345, 64) != 0
    assert lib.tlbm_set_launch(h, 320) != 0 and lib.tlbm_set_launch(h, -64) != 0
    for bx in (64, 192, 0):
        assert lib.tlbm_set_launch(h, bx) == 0


def test_batch_create_checks_domains_before_any_allocation():
    """A batched domain may hold any number of cells below 2^32 (the cell index is 32-bit)."""
    from tlbm import _capi
    lib = _capi.lib()
    h = ctypes.c_int(-1)
    lib.tlbm_kernel_lookup(ctypes.byref(_capi.Spec(1, 0, 0, 0, 0)), ctypes.byref(h))
    p = (_capi.Params * 2)()
    for q in p:
        q.nx, q.ny, q.nz = 8, 6, 4
    p[1].nx, p[1].ny, p[1].nz = 2 ** 20, 2 ** 10, 4
    out = ctypes.c_void_p()
    assert lib.tlbm_batch_create(h, p, 2, ctypes.byref(out)) != 0
    assert "32-bit" in _capi.last_error() and not out.value
    p[1].nx = 0
    assert lib.tlbm_batch_create(h, p, 2, ctypes.byref(out)) != 0
    assert "positive" in _capi.last_error()
    assert lib.tlbm_batch_create(h, p, 0, ctypes.byref(out)) != 0


def test_capture_entry_points_reject_the_null_stream():
    from tlbm import _capi
    lib = _capi.lib()
    out = ctypes.c_void_p()
    assert lib.tlbm_capture_begin(None) != 0
    assert "stream" in _capi.last_error()
    assert lib.tlbm_capture_end(None, ctypes.byref(out)) != 0
    assert out.value is None
    assert lib.tlbm_graph_destroy(None) == 0


def test_shifted_population_codecs():
    """encode/decode of f - w_q: f16 within half an ulp of the deviation, bf16 equal to an
    independent round-to-nearest-even (torch's bfloat16 conversion)."""
    import torch
    from tlbm.lbm import W32, decode_populations, encode_populations
    rng = np.random.default_rng(0)
    f = (W32[:, None] + 0.02 * rng.standard_normal((19, 1000))).astype(np.float32)
    dev = (f - W32[:, None]).astype(np.float32)
    back = decode_populations(encode_populations(f, "f16"), "f16")
    assert np.all(np.abs(back - f) <= np.abs(dev) * 2.0 ** -11 + 1e-7)
    raw = encode_populations(f, "bf16")
    ref = torch.from_numpy(dev).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(raw, ref)
    assert np.array_equal(encode_populations(f, "f32"), f)
    assert not encode_populations(np.tile(W32[:, None], (1, 5)), "bf16").any()
