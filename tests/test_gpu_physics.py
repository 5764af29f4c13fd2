"""Whole-run physics of the HIP step against the FP64 C oracle.

Each case runs the same configuration on the GPU (FP32, FP16 and BF16 population storage, FP32
arithmetic) and in the FP64 oracle, sampling the monitors every `every` steps the way
collision_check.py:14-22 does, and compares the drag/lift records and the final velocity field.
The errors are printed (pytest -s) with the bound each is held to.
"""

import os

import numpy as np
import pytest

from configs import C1, C1_STEPS, C2, C2_SAMPLE_EVERY, C2_STEPS, ST, ST_EVERY, ST_STEPS, ST_WINDOW

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _oracle():
    from oracle import c_oracle
    if not c_oracle.available():
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       capture_output=True)
    return c_oracle


def _run_gpu(cfg, steps, every, storage):
    import tlbm
    kw = dict(cfg)
    s = tlbm.Solver(**kw, storage=storage)
    drag, lift = [], []
    while s.step_count < steps:
        for _ in range(every - 1):
            s.step()
        s.step(report=True)
        m = s.monitors(every)
        drag.append(m["drag"])
        lift.append(m["lift"])
    tlbm.sync()
    u = np.stack([s.fields[k].numpy for k in ("ux", "uy", "uz")]).astype(np.float64)
    return np.array(drag), np.array(lift), u


def _run_oracle(cfg, steps, every):
    C = _oracle()
    s = C.CSolver(**cfg)
    drag, lift = [], []
    while s.step_count < steps:
        for _ in range(every - 1):
            s.step()
        s.step(report=True)
        m = s.monitors(every)
        drag.append(m["drag"])
        lift.append(m["lift"])
    u = np.stack([s.fields[k] for k in ("ux", "uy", "uz")])
    return np.array(drag), np.array(lift), u


_C1_REF = {}


def _c1_reference():
    if not _C1_REF:
        _C1_REF["r"] = _run_oracle(C1, C1_STEPS, 50)
    return _C1_REF["r"]


# bounds per storage: (max |du| / u0, rms |du| / u0, max |d drag| / mean drag, max |d lift| /
# mean drag), about 3x the errors measured on MI355X (f32 3.4e-5, 1.2e-5, 1.7e-5, 2.4e-6;
# f16 1.3e-3, 3.0e-4, 5.0e-4, 3.2e-4; bf16 1.1e-2, 2.8e-3, 8.3e-3, 2.1e-3)
C1_BOUNDS = {"f32": (1e-4, 4e-5, 6e-5, 1e-5), "f16": (4e-3, 1e-3, 2e-3, 1e-3),
             "bf16": (4e-2, 1e-2, 2.5e-2, 7e-3)}


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_c1_bgk_cylinder_against_fp64_oracle(storage):
    """BGK 96x48x48, cylinder r = 6 at (24, 24), Re 100, seeded noise, 1000 steps."""
    d_ref, l_ref, u_ref = _c1_reference()
    d, l, u = _run_gpu(C1, C1_STEPS, 50, storage)
    u0 = C1["u0"]
    du = np.sqrt(((u - u_ref) ** 2).sum(0))
    err = (du.max() / u0, np.sqrt((du ** 2).mean()) / u0,
           np.abs(d - d_ref).max() / abs(d_ref.mean()), np.abs(l - l_ref).max() / abs(d_ref.mean()))
    print(f"\nC1 {storage}: max|du|/u0 {err[0]:.2e}  rms|du|/u0 {err[1]:.2e}  "
          f"drag {d[-1]:.5f} (oracle {d_ref[-1]:.5f}) max rel {err[2]:.2e}  lift max {err[3]:.2e}")
    for e, b in zip(err, C1_BOUNDS[storage]):
        assert e <= b, (err, C1_BOUNDS[storage])


@pytest.mark.parametrize("collision", ["bgk", "trt", "les"])
def test_re300_strouhal_with_the_collision_check_estimator(collision):
    """collision_check.py:12-31 on the default Solver at Re 300, run for 12000 steps: its FFT
    arg-max estimator (bin width D / (u0 * 50 * 120) = 0.05 in St) gives the published 0.200
    (README.md:66-67) for all three collision models.  The bin-free estimate of the same record
    (shedding saturated, steps 8000-16000) is 0.22: the published value is that bin."""
    from oracle import lbm_oracle as O
    cfg = dict(ST, collision=collision)
    d, lift, _ = _run_gpu(cfg, ST_STEPS, ST_EVERY, "f32")
    st_bin = O.strouhal_fft(lift[:12000 // ST_EVERY], ST_EVERY, cfg["diameter"], cfg["u0"])
    a, b = (w // ST_EVERY for w in ST_WINDOW)
    st = O.strouhal_refined(lift[a:b], ST_EVERY, cfg["diameter"], cfg["u0"])
    print(f"\nRe 300 {collision}: St (collision_check estimator, 12000 steps) {st_bin:.3f}; "
          f"bin-free St {st:.4f}; mean drag {d[a:b].mean():.4f}")
    assert abs(st_bin - 0.200) < 1e-9
    assert 0.2 < st < 0.25


# St and mean drag against the FP64 record: measured on MI355X f32 -2.2e-4 / +1.2e-3, f16
# +7.2e-3 / -4.5e-3, bf16 -2.2e-2 / -2.8e-2.  FP32 is held to 0.5 %, FP16 to 1 %; BF16 keeps 8
# significant bits of each deviation f - w (FP16: 11) and misses 1 %: it is held to 4 %.
RE300_BOUNDS = {"f32": 0.005, "f16": 0.01, "bf16": 0.04}


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_re300_strouhal_and_drag_against_fp64_record(storage):
    """The same case (TRT) against the FP64 C-oracle record tests/golden/st_trt_re300_lift.npz:
    bin-free St and mean drag over steps 8000-16000 within RE300_BOUNDS[storage]."""
    from oracle import lbm_oracle as O
    path = os.path.join(ROOT, "tests", "golden", "st_trt_re300_lift.npz")
    if not os.path.exists(path):
        pytest.skip("tests/golden/st_trt_re300_lift.npz not generated")
    ref = np.load(path)
    d, lift, _ = _run_gpu(ST, ST_STEPS, ST_EVERY, storage)
    a, b = (w // ST_EVERY for w in ST_WINDOW)
    assert np.array_equal(ref["step"][a:b], np.arange(a + 1, b + 1) * ST_EVERY)
    st = O.strouhal_refined(lift[a:b], ST_EVERY, ST["diameter"], ST["u0"])
    st_ref = O.strouhal_refined(ref["lift"][a:b], ST_EVERY, ST["diameter"], ST["u0"])
    dm, dm_ref = d[a:b].mean(), ref["drag"][a:b].mean()
    bound = RE300_BOUNDS[storage]
    print(f"\nRe 300 TRT {storage}: St {st:.4f} (FP64 {st_ref:.4f}, {st / st_ref - 1:+.2e}); "
          f"mean drag {dm:.5f} (FP64 {dm_ref:.5f}, {dm / dm_ref - 1:+.2e})")
    assert abs(st / st_ref - 1) <= bound
    assert abs(dm / dm_ref - 1) <= bound


# C2 follows its FP64 record closely through the start-up transient; from about step 600 the
# flow amplifies any difference (FP32 against FP64: 2.2e-6 of the peak drag at step 500, 1.0e-3
# at step 1000, 1.8e-2 at step 1500), so only the first 500 steps are a trajectory test.  Bounds:
# about 4x the errors measured on MI355X (f32 2.2e-6, f16 2.1e-4, bf16 3.1e-3 of max |drag|).
C2_AGREE_STEPS = 500
C2_BOUNDS = {"f32": 1e-5, "f16": 1e-3, "bf16": 1.2e-2}


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_c2_trt_smagorinsky_against_fp64_record(storage):
    """C2 (TRT + Smagorinsky, 256x128x128, cylinder D = 24, Re 300) against the FP64 C-oracle
    record tests/golden/c2_trt_re300_lift.npz (monitors every 10 steps): drag and lift over the
    first 500 steps within the storage format's bound, relative to the peak drag."""
    path = os.path.join(ROOT, "tests", "golden", "c2_trt_re300_lift.npz")
    if not os.path.exists(path):
        pytest.skip("tests/golden/c2_trt_re300_lift.npz not generated")
    ref = np.load(path)
    k = C2_AGREE_STEPS // C2_SAMPLE_EVERY
    assert len(ref["step"]) == C2_STEPS // C2_SAMPLE_EVERY
    assert np.array_equal(ref["step"][:k], np.arange(1, k + 1) * C2_SAMPLE_EVERY)
    d, lift, _ = _run_gpu(C2, C2_AGREE_STEPS, C2_SAMPLE_EVERY, storage)
    scale = np.abs(ref["drag"][:k]).max()
    ed = np.abs(d - ref["drag"][:k]).max() / scale
    el = np.abs(lift - ref["lift"][:k]).max() / scale
    print(f"\nC2 {storage}, steps <= {C2_AGREE_STEPS}: max |d drag| / max |drag| {ed:.2e}, "
          f"max |d lift| / max |drag| {el:.2e} (bound {C2_BOUNDS[storage]:.0e})")
    assert ed <= C2_BOUNDS[storage] and el <= C2_BOUNDS[storage]


# C2 over its whole 20000 steps.  Past the start-up transient the flow is irregular and every
# run leaves the FP64 trajectory (any perturbation grows), so the second half is compared as
# statistics: mean drag, lift RMS and the dominant lift frequency over steps 10001-20000, each
# with a moving-block bootstrap standard error (blocks of 1000 steps, about the drag's
# decorrelation time).  A storage format passes when each difference is within three combined
# standard errors, and the Strouhal numbers within one spectral bin of the 10000-step window.
C2_HALF = (10000, 20000)
C2_BLOCK_STEPS = 1000


def _block_bootstrap_se(x, block, stat, reps=2000, seed=0):
    """Standard error of stat(x) from a moving-block bootstrap (blocks of `block` samples)."""
    rng = np.random.default_rng(seed)
    n = len(x)
    nb = n // block
    starts = rng.integers(0, n - block + 1, size=(reps, nb))
    idx = (starts[:, :, None] + np.arange(block)[None, None, :]).reshape(reps, -1)
    return float(np.std([stat(x[i]) for i in idx]))


def _c2_statistics(drag, lift):
    every = C2_SAMPLE_EVERY
    a, b = (w // every for w in C2_HALF)
    d, l = np.asarray(drag[a:b]), np.asarray(lift[a:b])
    blk = C2_BLOCK_STEPS // every
    from oracle import lbm_oracle as O
    st = O.strouhal_refined(l, every, C2["diameter"], C2["u0"])
    return {"drag": (float(d.mean()), _block_bootstrap_se(d, blk, np.mean)),
            "lift_rms": (float(l.std()), _block_bootstrap_se(l, blk, np.std)),
            "st": st}


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_c2_whole_run_statistics_against_fp64_record(storage):
    """C2 (TRT + Smagorinsky 256x128x128, cylinder D = 24, Re 300) for all 20000 steps, monitors
    every 10 as in tests/golden/c2_trt_re300_lift.npz: the second half's mean drag, lift RMS and
    dominant lift frequency against the FP64 record's, within their bootstrap uncertainty."""
    path = os.path.join(ROOT, "tests", "golden", "c2_trt_re300_lift.npz")
    if not os.path.exists(path):
        pytest.skip("tests/golden/c2_trt_re300_lift.npz not generated")
    ref = np.load(path)
    assert np.array_equal(ref["step"], np.arange(1, len(ref["step"]) + 1) * C2_SAMPLE_EVERY)
    d, lift, _ = _run_gpu(C2, C2_STEPS, C2_SAMPLE_EVERY, storage)
    assert np.all(np.isfinite(d)) and np.all(np.isfinite(lift))
    got, want = _c2_statistics(d, lift), _c2_statistics(ref["drag"], ref["lift"])
    bin_st = C2["diameter"] / (C2["u0"] * (C2_HALF[1] - C2_HALF[0]))
    lines = [f"\nC2 {storage}, steps {C2_HALF[0] + 1}-{C2_HALF[1]}:"]
    for key in ("drag", "lift_rms"):
        (g, se_g), (w, se_w) = got[key], want[key]
        z = abs(g - w) / np.hypot(se_g, se_w)
        lines.append(f"{key} {g:.4f} +- {se_g:.4f} (FP64 {w:.4f} +- {se_w:.4f}, "
                     f"{g / w - 1:+.2%}, {z:.2f} SE)")
        assert z <= 3.0, (key, got[key], want[key])
    lines.append(f"St {got['st']:.4f} (FP64 {want['st']:.4f}, {got['st'] / want['st'] - 1:+.2%}, "
                 f"bin {bin_st:.3f})")
    print("  ".join(lines))
    assert abs(got["st"] - want["st"]) <= bin_st


@pytest.mark.parametrize("reynolds", [300.0, 1000.0, 3000.0])
@pytest.mark.parametrize("collision", ["bgk", "trt", "les"])
def test_collision_check_stability_matrix(collision, reynolds):
    """collision_check.py:9-26 at its default 8000 steps: the default Solver at Re 300, 1000 and
    3000, monitors every 50 steps, stopping at the first non-finite drag.  README.md:65-66: BGK
    diverges at Re 3000; TRT and LES (Smagorinsky subgrid viscosity) stay stable."""
    import tlbm
    s = tlbm.Solver(collision=collision, reynolds=reynolds)
    blew_up = None
    while s.step_count < 8000:
        for _ in range(49):
            s.step()
        s.step(report=True)
        if not np.isfinite(s.monitors(50)["drag"]):
            blew_up = s.step_count
            break
    print(f"\nRe {reynolds:5.0f} {collision}: "
          + (f"diverged at step {blew_up}" if blew_up else "stable for 8000 steps"))
    if collision == "bgk" and reynolds == 3000.0:
        assert blew_up is not None
    elif collision != "bgk":
        assert blew_up is None


_C3_F32 = {}


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_headline_workload_long_run_in_16_bit_storage(storage):
    """The headline workload (LES 512x256x256, sphere, moving lid, Re 3000) over 6000 steps,
    monitors every 100: finite in every storage format, and the 16-bit formats' mean drag over
    steps 3001-6000 within 2 % of FP32's (measured on MI355X: FP16 +0.1 %, BF16 -0.04 %)."""
    import tlbm
    from configs import C3

    def run(st):
        s = tlbm.Solver(**C3, storage=st)
        drag = []
        while s.step_count < 6000:
            for _ in range(99):
                s.step()
            s.step(report=True)
            drag.append(s.monitors(100)["drag"])
        return np.array(drag)
    if "d" not in _C3_F32:
        _C3_F32["d"] = run("f32")
    ref, d = _C3_F32["d"], run(storage)
    assert np.all(np.isfinite(ref)) and np.all(np.isfinite(d))
    m, m_ref = d[30:].mean(), ref[30:].mean()
    print(f"\nC3 {storage}: mean drag (steps 3001-6000) {m:.5f}, FP32 {m_ref:.5f} "
          f"({m / m_ref - 1:+.2e})")
    assert abs(m / m_ref - 1) < 0.02


# ---- velocity fields of C2, C3 and C4 against FP64 records (tests/golden/config_fields.npz,
# written by tests/golden/make_config_fields.py from the C oracle) ----------------------------
CONFIG_FIELDS = os.path.join(ROOT, "tests", "golden", "config_fields.npz")
C4_DOMAINS = (1, 4, 7)


def _config_record():
    if not os.path.exists(CONFIG_FIELDS):
        pytest.skip("tests/golden/config_fields.npz not generated")
    return np.load(CONFIG_FIELDS)


def _velocity_errors(u, u_ref, u0):
    du = np.sqrt(((np.asarray(u, np.float64) - u_ref) ** 2).sum(0))
    return du.max() / u0, np.sqrt((du ** 2).mean()) / u0


# (max |du| / u0, rms |du| / u0, max |d drag| / |drag|) per case and storage, about 3x the errors
# measured on MI355X.  C2 at step 500 and C4 domain 7 (Re 370, shedding) have left the FP64
# trajectory locally in the 16-bit formats (the flow amplifies any difference: C2's FP16 max
# error is 2.9 u0 where a start-up shear-layer feature has moved, while its drag is within
# 2e-3), so their velocity bounds are loose there.  Measured:
#   c2 f32 1.4e-2 3.4e-4 3.1e-5 | f16 2.9 9.4e-2 2.0e-3 | bf16 4.6 2.2e-1 1.7e-2
#   c3 f32 2.5e-5 8.3e-6 2.4e-5 | f16 1.7e-3 5.1e-4 2.8e-4 | bf16 1.4e-2 4.0e-3 2.5e-3
#   c4 domains 1, 4 (Re 115, 103): f32 6.5e-5 2.0e-5 3.1e-6 | f16 2.1e-3 4.7e-4 6.9e-4
#                                  | bf16 7.4e-2 1.9e-2 3.3e-2
#   c4 domain 7 (Re 370): f32 2.9e-3 1.7e-4 4.4e-5 | f16 9.2e-1 5.9e-2 5.5e-2
#                         | bf16 1.4 1.3e-1 1.7e-1
FIELD_BOUNDS = {
    ("c2", "f32"): (4e-2, 1e-3, 1e-4), ("c2", "f16"): (9.0, 0.3, 6e-3),
    ("c2", "bf16"): (14.0, 0.7, 5e-2),
    ("c3", "f32"): (8e-5, 2.5e-5, 8e-5), ("c3", "f16"): (5e-3, 1.5e-3, 1e-3),
    ("c3", "bf16"): (4.5e-2, 1.2e-2, 8e-3),
    ("c4", "f32"): (2e-4, 6e-5, 1e-5), ("c4", "f16"): (6e-3, 1.5e-3, 2e-3),
    ("c4", "bf16"): (0.2, 6e-2, 0.1),
    ("c4_7", "f32"): (1e-2, 5e-4, 1.5e-4), ("c4_7", "f16"): (3.0, 0.18, 0.15),
    ("c4_7", "bf16"): (4.5, 0.4, 0.5),
}


def _check(errors, bounds):
    for e, b in zip(errors, bounds):
        assert b is None or e <= b, (errors, bounds)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("case", ["c2", "c3"])
def test_config_velocity_against_fp64_record(case, storage):
    """C2 (500 steps, plane z = 64) and the headline C3 (200 steps, plane z = 128 at every other
    x and y): max and RMS velocity error over the recorded plane, and the drag at each report
    step (every 100), against the FP64 C oracle."""
    from configs import C3
    rec = _config_record()
    cfg, steps = (C2, 500) if case == "c2" else (C3, 200)
    d, _, u = _run_gpu(cfg, steps, 100, storage)
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    u = u.reshape(3, nz, ny, nx)
    u = u[:, 64] if case == "c2" else u[:, 128, ::2, ::2]
    emax, erms = _velocity_errors(u, rec[f"{case}_u"], cfg["u0"])
    ed = float(np.abs(d - rec[f"{case}_drag"]).max() / np.abs(rec[f"{case}_drag"]).max())
    print(f"\n{case} {storage}: max|du|/u0 {emax:.2e}  rms|du|/u0 {erms:.2e}  "
          f"drag {d[-1]:.6f} (FP64 {rec[f'{case}_drag'][-1]:.6f}) max rel {ed:.2e}")
    _check((emax, erms, ed), FIELD_BOUNDS[(case, storage)])


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_c4_sweep_domains_against_fp64_record(storage):
    """C4: the 64-channel sweep stepped 1000 times by the batched kernel (report at the last);
    domains 1, 4 (laminar) and 7 (Re 370, shedding) against their FP64 C-oracle runs: velocity
    on the plane z = nz / 2 and the drag at step 1000."""
    import tlbm
    from tlbm.batch import BatchSolver, sweep_domains
    rec = _config_record()
    domains = sweep_domains(64, seed=0)
    b = BatchSolver(domains, collision="trt", storage=storage, noise_amp=1e-3, fields=True)
    for _ in range(999):
        b.step()
    b.step(report=True)
    mons = b.monitors(1000)
    tlbm.sync()
    fields = {k: b.fields[k].numpy() for k in ("ux", "uy", "uz")}
    for i in C4_DOMAINS:
        dom = domains[i]
        _, c = b.offsets(i)
        u = np.stack([fields[k][c:c + dom.n].reshape(dom.nz, dom.ny, dom.nx)[dom.nz // 2]
                      for k in ("ux", "uy", "uz")])
        emax, erms = _velocity_errors(u, rec[f"c4_{i}_u"], 0.08)
        ref_drag = float(rec[f"c4_{i}_drag"][-1])
        ed = abs(mons[i]["drag"] - ref_drag) / abs(ref_drag)
        print(f"\nc4 domain {i} {storage}: max|du|/u0 {emax:.2e}  rms|du|/u0 {erms:.2e}  "
              f"drag {mons[i]['drag']:.6f} (FP64 {ref_drag:.6f}) rel {ed:.2e}")
        _check((emax, erms, ed), FIELD_BOUNDS[("c4_7" if i == 7 else "c4", storage)])
    b.close()
