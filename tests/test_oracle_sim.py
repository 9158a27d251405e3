"""What pins the oracle's walks independently of their own code, on the CPU: orc_sim,
the restatement of src/sim/kernels.cu:72-233 in the reference's float/double mix,
and its FP64 twin orc_sim_f64 (the same walk with field, phase, rotations and
relaxation in double), against which every precision mode's error is stated.
Each closed-form test runs on both twins (``prec`` "f32" / "f64"); the tolerance
table TOL gives the bound per twin.  The FP64 twin is held to 1e-9 relative (or
1e-12 absolute on unit magnetisation), which a dropped term, a wrong sign, a
float-rounded operand or a transposed index would exceed by orders of magnitude.

Every expectation below comes from the mathematics of the sequence, not from the
oracle's formulas:
  - free precession and relaxation (Bloch closed forms) at one or several echoes;
  - static dephasing: with no diffusion a spin keeps its voxel, so its phase is
    (TE index + 1) times that voxel's field, found here from the position alone;
  - a spin echo refocuses any static field: |Mxy(TE)| = exp(-TE/T2) for every spin;
  - a gradient sample adds gamma G x dt (the reference's config/gradient.ini case);
  - ideal dephasing gives spin i the phase i * D / n_spins (kernels.cu:175-178);
  - free diffusion: per-axis steps are N(0, 2 D dt) (simulation_parameters.cuh:236-237);
  - PGSE: the echo equals exp(-Var(phase)/2) with Var computed exactly for the
    discrete waveform (Stejskal-Tanner in discrete time);
  - bSSFP with alternating RF phase reaches the Freeman-Hill steady state;
  - permeability 0 keeps every spin in its start label, permeability 1 leaves the
    trajectories exactly those of a single-label phantom;
  - reflecting and wrapping boundaries keep positions inside the FoV;
  - a spin starting outside the volume is lost and its outputs are left as they were.
  - an excitation at a general phase (phase cycling) and a refocusing pulse about a
    general axis in the transverse plane are rotations by the flip angle about
    (cos ph, sin ph, 0) (Rodrigues' formula), with relaxation between events and to
    the end of every TR;
  - the two twins walk identical trajectories and labels: positions are double in
    both (kernels.cu:93, 131).
Rotation directions are the ones the reference's own known answers fix
(tests/test_kernel.cpp:36-89, re-checked in test_oracle_kat.py).  B0 and gradient
samples are float in the reference (simulation_parameters.cuh:181,
config_reader.cpp:16-23), so the expectations use their float32 values.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from tests import cases

GAMMA = 267515315.0             # rad/s/T (src/definitions.h:20)
RAD2DEG = 57.2957795130823      # src/sim/simulation_parameters.cuh:25


def rx(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def ry(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def rz(deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def relax(m, t_s, T1_ms, T2_ms):
    e1, e2 = math.exp(-t_s / (T1_ms * 1e-3)), math.exp(-t_s / (T2_ms * 1e-3))
    return np.array([m[0] * e2, m[1] * e2, 1 + e1 * (m[2] - 1)])


B0 = float(np.float32(9.4))     # parameters.B0 is float (simulation_parameters.cuh:181)

# Tolerances per twin: the float walk rounds phases and rotations in float, the
# FP64 twin only in double.
TOL = {
    "f32": dict(m_atol=2e-6, deph_atol=2e-4, se_rtol=2e-5, se_atol=1e-4, grad_atol=2e-5, ens_atol=1e-5,
                ideal_atol=2e-5, fh_rtol=2e-4, general_atol=1e-5),
    "f64": dict(m_atol=1e-12, deph_atol=1e-12, se_rtol=1e-9, se_atol=1e-12, grad_atol=1e-12, ens_atol=1e-12,
                ideal_atol=1e-12, fh_rtol=1e-9, general_atol=1e-12),
}
PRECS = ["f32", "f64"]


def run(cfg, mask, field_T, fov_um, n, xyz0=None, prec="f32", **kw):
    cfg = dict(cfg, NUMBER_OF_SPINS=n)
    c = orc.ConfigCase(cfg, mask, field_T, fov_um * 1e-6, XYZ0=xyz0)
    d = c.oracle()
    d.update(kw)
    return c, d, orc.sim(d, n_threads=8, f64=(prec == "f64"))


def deg_per_step(field_T, dt):
    """Field in degrees per step, in double: B0 (float) * dt * gamma * 180/pi."""
    return field_T.astype(np.float64) * (B0 * dt * 1e-6 * GAMMA * RAD2DEG)


def rot_axis(deg, ph_deg):
    """Rotation by `deg` about (cos ph, sin ph, 0): Rodrigues' formula."""
    u = np.array([math.cos(math.radians(ph_deg)), math.sin(math.radians(ph_deg)), 0.0])
    a = math.radians(deg)
    K = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
    return math.cos(a) * np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * np.outer(u, u)


def voxel_of(xyz0, shape, fov_m):
    g = np.floor(xyz0.astype(np.float64) * (np.asarray(shape) / np.float64(np.float32(fov_m))))
    return g.astype(int)


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("prec", PRECS)
def test_free_precession_and_relaxation_at_several_echoes(prec):
    cfg = cases.base_cfg(1, TR=10050, TE=[2000, 5000, 8000], TIME_STEP=50, DIFFUSIVITY=[0.0],
                         T1=[2200.0], T2=[41.0])
    _, _, r = run(cfg, np.zeros((4, 4, 4), np.uint8), None, 10.0, 8, prec=prec)
    m_exc = rx(90) @ np.array([0.0, 0.0, 1.0])
    for k, te_us in enumerate((2000, 5000, 8000)):
        expect = relax(m_exc, te_us * 1e-6, 2200.0, 41.0)
        np.testing.assert_allclose(r["M1"][:, k, :], np.tile(expect, (8, 1)), atol=TOL[prec]["m_atol"], rtol=0)
    assert r["n_lost"] == 0


@pytest.mark.parametrize("prec", PRECS)
def test_static_dephasing_brute_force(prec):
    rng = np.random.default_rng(5)
    shape = (6, 5, 7)                                       # non-cubic: a swapped axis fails
    field = (rng.standard_normal(shape) * 1e-7).astype(np.float32)
    cfg = cases.gre(1, TE=5000, dt=50, DIFFUSIVITY=[0.0], T1=[1500.0], T2=[60.0])
    c, d, r = run(cfg, np.zeros(shape, np.uint8), field, 30.0, 300, prec=prec)
    te_idx = 5000 // 50
    f_deg = deg_per_step(field, 50)
    v = voxel_of(c.XYZ0, shape, 30e-6)
    m_exc = rx(90) @ np.array([0.0, 0.0, 1.0])
    for i in range(300):
        phi = (te_idx + 1) * f_deg[tuple(v[i])]
        expect = relax(rz(phi) @ m_exc, te_idx * 50e-6, 1500.0, 60.0)
        np.testing.assert_allclose(r["M1"][i, 0], expect, atol=TOL[prec]["deph_atol"], rtol=0)


@pytest.mark.parametrize("prec", PRECS)
def test_spin_echo_refocuses_any_static_field(prec):
    m, f = cases.cylinder_phantom(n=16, fov_um=32.0, radius_um=4.0, n_cyl=2, seed=8)
    cfg = cases.se(2, TE=10000, dt=50, DIFFUSIVITY=[0.0, 0.0], T2=[40.0, 25.0], T1=[2200.0, 1500.0])
    c, _, r = run(cfg, m, f * 50, 32.0, 400, prec=prec)
    v = voxel_of(c.XYZ0, m.shape, 32e-6)
    lab = m[v[:, 0], v[:, 1], v[:, 2]]
    T2 = np.where(lab == 0, 40.0, 25.0)
    mag = np.hypot(r["M1"][:, 0, 0].astype(np.float64), r["M1"][:, 0, 1].astype(np.float64))
    np.testing.assert_allclose(mag, np.exp(-10.0 / T2), rtol=TOL[prec]["se_rtol"])
    np.testing.assert_array_equal(r["T"][:, 0], lab)
    # Within a step the field is added before the RF is applied (kernels.cu:171-199),
    # so the first half holds 101 field steps (0..100) and the second 100: the echo
    # keeps one step of phase.  Every spin, rotations and relaxation in float64:
    f_deg = deg_per_step(f * 50, 50)
    T1 = np.where(lab == 0, 2200.0, 1500.0)
    m_exc = rx(90) @ np.array([0.0, 0.0, 1.0])
    for i in range(400):
        fi = f_deg[tuple(v[i])]
        m = relax(rz(101 * fi) @ m_exc, 100 * 50e-6, T1[i], T2[i])
        m = relax(rz(100 * fi) @ (ry(180) @ m), 100 * 50e-6, T1[i], T2[i])
        np.testing.assert_allclose(r["M1"][i, 0], m, atol=TOL[prec]["se_atol"], rtol=0)


@pytest.mark.parametrize("prec", PRECS)
def test_gradient_sample_phase_known_answer(prec):
    # config/gradient.ini: D = 0, excitation about y (RF_PH 90), one gradient
    # sample of 521.9377 mT/m on x for one dt at t = 10 ms, echo at 20 ms.
    G, dt = 521.9377, 50
    cfg = cases.base_cfg(1, TR=40000, TE=[20000], RF_FA=[90.0], RF_PH=[90.0], RF_T=[0], TIME_STEP=dt,
                         DIFFUSIVITY=[0.0], T1=[1000e3], T2=[1000e3],
                         GRADIENT_X=[G], GRADIENT_Y=[0.0], GRADIENT_Z=[0.0], GRADIENT_T=[10000])
    n = 11
    g = (np.arange(n) + 0.5) / n
    xyz0 = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3).astype(np.float32) * np.float32(100e-6)
    _, _, r = run(cfg, np.zeros((4, 4, 4), np.uint8), None, 100.0, n ** 3, xyz0=xyz0, prec=prec)
    m_exc = ry(90) @ np.array([0.0, 0.0, 1.0])
    Gf = float(np.float32(G))                      # gradient samples are float (config_reader.cpp:16-23)
    phi = Gf * xyz0[:, 0].astype(np.float64) * 1e-3 * dt * 1e-6 * GAMMA * RAD2DEG
    for i in range(0, n ** 3, 7):
        expect = relax(rz(phi[i]) @ m_exc, 20000e-6, 1000e3, 1000e3)
        np.testing.assert_allclose(r["M1"][i, 0], expect, atol=TOL[prec]["grad_atol"], rtol=0)
    # the ensemble signal is the sinc of the phase ramp over the sampled grid
    s = np.abs(np.mean(np.exp(1j * np.radians(phi))))
    got = np.hypot(r["M1"][:, 0, 0].astype(np.float64).mean(), r["M1"][:, 0, 1].astype(np.float64).mean())
    assert abs(got - s * math.exp(-0.02 / 1000.0)) < TOL[prec]["ens_atol"]


@pytest.mark.parametrize("prec", PRECS)
def test_ideal_dephasing_is_linear_in_the_spin_index(prec):
    # 67 spins and 357.3 degrees: i * D / n is inexact in float, so a twin that
    # formed it in float would miss the FP64 bound
    n, deph = 67, float(np.float32(357.3))
    cfg = cases.base_cfg(1, TR=5050, TE=[5000], TIME_STEP=50, DIFFUSIVITY=[0.0], T1=[1e6], T2=[1e6],
                         DEPHASING=[357.3], DEPHASING_T=[1000])
    _, _, r = run(cfg, np.zeros((4, 4, 4), np.uint8), None, 10.0, n, prec=prec)
    m_exc = rx(90) @ np.array([0.0, 0.0, 1.0])
    for i in range(n):
        np.testing.assert_allclose(r["M1"][i, 0], relax(rz(deph * i / n) @ m_exc, 5e-3, 1e6, 1e6),
                                   atol=TOL[prec]["ideal_atol"], rtol=0)


@pytest.mark.parametrize("prec", PRECS)
def test_general_phase_excitation_and_refocusing_with_relaxation(prec):
    # Phase-cycled excitations (linear 37, quadratic 11 degrees: general phases,
    # kernels.cu:110-120), a refocusing pulse about a general transverse axis
    # (RF_PH 33, kernels.cuh:183-195) and relaxation between events and to the end
    # of every TR (kernels.cu:190-231), over six TRs; D = 0 and no field.
    nd, fa0, fa1, ph1, dt, TR, t_rf, TE = 5, 50.0, 120.0, 33.0, 50, 10000, 3000, 7000
    T1, T2 = 700.0, 90.0
    cfg = cases.base_cfg(1, TR=TR, TE=[TE], RF_FA=[fa0, fa1], RF_PH=[0.0, ph1], RF_T=[0, t_rf], TIME_STEP=dt,
                         DUMMY_SCAN=nd, LINEAR_PHASE_CYCLING=37.0, QUADRATIC_PHASE_CYCLING=11.0,
                         DIFFUSIVITY=[0.0], T1=[T1], T2=[T2])
    _, _, r = run(cfg, np.zeros((4, 4, 4), np.uint8), None, 10.0, 4, prec=prec)
    m = np.array([0.0, 0.0, 1.0])
    n_tp = TR // dt
    for d in range(nd + 1):
        ph = 0.0 + d * float(np.float32(37.0)) + d * (d + 1) / 2.0 * float(np.float32(11.0))
        m = rot_axis(fa0, ph) @ m
        m = rot_axis(fa1, ph1) @ relax(m, (t_rf // dt) * dt * 1e-6, T1, T2)
        if d == nd:
            m = relax(m, (TE - t_rf) // dt * dt * 1e-6, T1, T2)
            echo = m.copy()
            m = relax(m, (n_tp - TE // dt) * dt * 1e-6, T1, T2)
        else:
            m = relax(m, (n_tp - t_rf // dt) * dt * 1e-6, T1, T2)
    np.testing.assert_allclose(r["M1"][:, 0, :], np.tile(echo, (4, 1)), atol=TOL[prec]["general_atol"], rtol=0)


def test_free_diffusion_steps_are_gaussian_with_variance_2Ddt():
    D, dt, n = 2e-9, 50, 1000
    cfg = cases.gre(1, TE=5000, dt=dt, DIFFUSIVITY=[D], RECORD_TRAJECTORY=1, T1=[1e9], T2=[1e9])
    rng = np.random.default_rng(1)
    xyz0 = rng.uniform(0.4, 0.6, size=(n, 3)).astype(np.float32) * np.float32(1000e-6)
    _, _, r = run(cfg, np.zeros((8, 8, 8), np.uint8), None, 1000.0, n, xyz0=xyz0)
    inc = np.diff(r["XYZ1"].astype(np.float64), axis=1).reshape(-1, 3)
    sigma2 = 2 * np.float32(D) * dt * 1e-6
    m = inc.shape[0]
    for k in range(3):
        assert abs(inc[:, k].mean()) < 5 * math.sqrt(sigma2 / m)
        assert abs(inc[:, k].var() / sigma2 - 1) < 5 * math.sqrt(2 / m) + 1e-3
        z = inc[:, k] / math.sqrt(sigma2)
        assert abs((z ** 4).mean() - 3.0) < 5 * math.sqrt(96 / m)         # Gaussian kurtosis
    assert np.abs(np.corrcoef(inc.T) - np.eye(3)).max() < 5 / math.sqrt(m)
    # the FP64 twin walks the same trajectory
    _, _, r64 = run(cfg, np.zeros((8, 8, 8), np.uint8), None, 1000.0, n, xyz0=xyz0, prec="f64")
    np.testing.assert_array_equal(r64["XYZ1"], r["XYZ1"])


@pytest.mark.parametrize("prec", PRECS)
def test_pgse_echo_matches_discrete_stejskal_tanner(prec):
    D, b, delta, DELTA, dt = 2e-9, 500.0, 5.0, 10.0, 50
    cfg, G = cases.pgse(1, b=b, delta_ms=delta, DELTA_ms=DELTA, dt=dt, axis=(1, 0, 0), start_ms=1.0)
    cfg.update(DIFFUSIVITY=[D], T1=[1e9], T2=[1e9])
    n = 12000
    rng = np.random.default_rng(2)
    xyz0 = rng.uniform(0.45, 0.55, size=(n, 3)).astype(np.float32) * np.float32(1000e-6)
    _, d, r = run(cfg, np.zeros((8, 8, 8), np.uint8), None, 1000.0, n, xyz0=xyz0, prec=prec)
    t_g, Gx = np.asarray(d["gradient_t"]), np.asarray(d["gx"], np.float64)
    s = np.where(t_g <= int(d["RF_t"][1]), -1.0, 1.0)        # sign flips at the refocusing pulse
    gp = 1e-3 * dt * 1e-6 * GAMMA
    F = np.array([(s * Gx)[t_g >= j].sum() for j in range(d["n_timepoints"])])
    expect = math.exp(-0.5 * gp * gp * d["diffusivity"][0] ** 2 * float((F * F).sum()))
    v = r["M1"][:, 0, :2].astype(np.float64)
    mean = v.mean(axis=0)
    got = float(np.linalg.norm(mean))
    u = mean / got
    se = math.sqrt(float(u @ np.cov(v.T) @ u) / n)
    assert 0.2 < expect < 0.8
    assert abs(got - expect) < 4.5 * se + 1e-4, (got, expect, se)


@pytest.mark.parametrize("prec", PRECS)
def test_bssfp_reaches_freeman_hill_steady_state(prec):
    # The transient decays about as E1^n: the FP64 twin needs it below 1e-9, so
    # its case relaxes faster (T1 100 ms) and runs 700 dummy scans.
    T1, T2, dummies = (200.0, 100.0, 400) if prec == "f32" else (100.0, 50.0, 700)
    TR, TE, fa = 5000, 2500, 30.0
    cfg = cases.bssfp(1, TE=TE, dt=50, dummies=dummies, lin=180.0, fa=fa, DIFFUSIVITY=[0.0],
                      T1=[T1], T2=[T2])
    _, _, r = run(cfg, np.zeros((4, 4, 4), np.uint8), None, 10.0, 4, prec=prec)
    E1, E2 = math.exp(-TR * 1e-3 / T1), math.exp(-TR * 1e-3 / T2)
    a = math.radians(fa)
    mxy = (1 - E1) * math.sin(a) / (1 - (E1 - E2) * math.cos(a) - E1 * E2)
    expect = mxy * math.exp(-TE * 1e-3 / T2)
    got = np.hypot(r["M1"][:, 0, 0].astype(np.float64), r["M1"][:, 0, 1].astype(np.float64))
    np.testing.assert_allclose(got, expect, rtol=TOL[prec]["fh_rtol"])


@pytest.mark.parametrize("case", ["se_permeable_3labels", "pgse_wrap", "gre_reflect_lost"])
def test_twins_walk_identical_trajectories_and_labels(case):
    # Positions, draws, lookups and permeability decisions are the same code in
    # both twins (double positions, kernels.cu:93, 131), so trajectories, labels
    # and lost spins agree bit for bit; magnetisations differ by float rounding.
    if case == "se_permeable_3labels":
        m, f = cases.three_substrate_phantom(n=16, fov_um=32.0, seed=37)
        cfg = cases.se(3, TE=8000, dt=50, DIFFUSIVITY=[2e-9, 1e-9, 1.5e-9], P_XY=[[1, .5, .3], [.4, 1, .6], [.2, .7, 1]],
                       T1=[2000.0, 1600.0, 1200.0], T2=[40.0, 20.0, 30.0], RECORD_TRAJECTORY=1)
        fov, n = 32.0, 300
    elif case == "pgse_wrap":
        m, f = cases.cylinder_phantom(n=16, fov_um=32.0, radius_um=4.0, n_cyl=2, seed=9)
        cfg, _ = cases.pgse(2, b=1000.0, delta_ms=3.0, DELTA_ms=6.0, dt=50, axis=(1, 1, 0), start_ms=1.0)
        cfg.update(DIFFUSIVITY=[2e-9, 2e-9], P_XY=[[1, .5], [.5, 1]], CROSS_FOV=1, RECORD_TRAJECTORY=1)
        fov, n = 32.0, 300
    else:
        m, f = cases.two_substrate_phantom(n=12, fov_um=12.0, seed=38)
        cfg = cases.gre(2, TE=6000, dt=50, DIFFUSIVITY=[3e-9, 3e-9], P_XY=[[1, 0.0], [0.0, 1]], MAX_ITERATIONS=3)
        fov, n = 12.0, 400
    _, _, a = run(cfg, m, f, fov, n, prec="f32")
    _, _, b = run(cfg, m, f, fov, n, prec="f64")
    np.testing.assert_array_equal(a["XYZ1"], b["XYZ1"])
    np.testing.assert_array_equal(a["T"], b["T"])
    np.testing.assert_array_equal(a["lost"], b["lost"])
    ok = a["lost"] == 0
    assert ok.sum() > n // 2
    np.testing.assert_allclose(a["M1"][ok].astype(np.float64), b["M1"][ok], atol=2e-4, rtol=0)
    if case == "gre_reflect_lost":
        assert a["n_lost"] > 0          # stuck spins (MAX_ITERATIONS 3) are abandoned in both


def test_impermeable_labels_keep_every_spin_inside_its_start_label():
    m, f = cases.three_substrate_phantom(n=16, fov_um=32.0, seed=35)
    cfg = cases.se(3, TE=6000, dt=50, DIFFUSIVITY=[2e-9] * 3, MAX_ITERATIONS=1e6, RECORD_TRAJECTORY=1)
    c, _, r = run(cfg, m, f, 32.0, 500)
    v0 = voxel_of(c.XYZ0, m.shape, 32e-6)
    lab0 = m[v0[:, 0], v0[:, 1], v0[:, 2]]
    assert r["n_lost"] == 0
    np.testing.assert_array_equal(r["T"][:, 0], lab0)
    traj = r["XYZ1"].reshape(-1, 3)
    vt = voxel_of(traj, m.shape, 32e-6).reshape(500, -1, 3)
    labs = m[vt[..., 0], vt[..., 1], vt[..., 2]]
    assert (labs == lab0[:, None]).all()


def test_permeability_one_leaves_trajectories_of_a_single_label():
    m, f = cases.three_substrate_phantom(n=16, fov_um=32.0, seed=36)
    kw = dict(DIFFUSIVITY=[1.5e-9] * 3, T1=[2000.0] * 3, T2=[50.0] * 3, RECORD_TRAJECTORY=1)
    cfg3 = cases.se(3, TE=4000, dt=50, P_XY=np.ones((3, 3)).tolist(), **kw)
    kw1 = {k: v[:1] if isinstance(v, list) else v for k, v in kw.items()}
    cfg1 = cases.se(1, TE=4000, dt=50, **kw1)
    _, _, a = run(cfg3, m, None, 32.0, 300)
    _, _, b = run(cfg1, np.zeros_like(m), None, 32.0, 300)
    np.testing.assert_array_equal(a["XYZ1"], b["XYZ1"])
    np.testing.assert_array_equal(a["M1"], b["M1"])


def test_one_way_membrane_uses_row_from_column_to():
    # P_XY is indexed [from * n + to] (kernels.cu:154): with P(0->1) = 1 and
    # P(1->0) = 0 spins enter label 1 freely and never leave it.
    m = np.zeros((16, 16, 16), np.uint8)
    m[8:] = 1
    cfg = cases.gre(2, TE=20000, dt=50, DIFFUSIVITY=[2e-9, 2e-9], P_XY=[[1.0, 1.0], [0.0, 1.0]],
                    MAX_ITERATIONS=1e6, RECORD_TRAJECTORY=1)
    c, _, r = run(cfg, m, None, 16.0, 400)
    v0 = voxel_of(c.XYZ0, m.shape, 16e-6)
    lab0 = m[v0[:, 0], v0[:, 1], v0[:, 2]]
    assert r["n_lost"] == 0
    assert (r["T"][lab0 == 1, 0] == 1).all()
    assert (r["T"][lab0 == 0, 0] == 1).sum() > 10
    vt = voxel_of(r["XYZ1"].reshape(-1, 3), m.shape, 16e-6).reshape(400, -1, 3)
    labs = m[vt[..., 0], vt[..., 1], vt[..., 2]]
    assert (np.diff(labs.astype(int), axis=1) >= 0).all()          # never 1 -> 0


@pytest.mark.parametrize("cross", [0, 1])
def test_boundaries_keep_positions_inside_the_fov(cross):
    # Steps (sigma 0.55 um) far below half the FoV: a reflection off the far wall
    # can then not carry a spin past the near one (kernels.cu:133-136).
    cfg = cases.gre(1, TE=5000, dt=50, DIFFUSIVITY=[3e-9], RECORD_TRAJECTORY=1, CROSS_FOV=cross)
    _, d, r = run(cfg, np.zeros((5, 5, 5), np.uint8), None, 20.0, 300)
    assert r["n_lost"] == 0
    x = r["XYZ1"].astype(np.float64)
    fov = np.asarray(d["fov"])
    assert (x >= 0).all() and (x < fov).all()
    # the steps are a good fraction of the FoV, so the boundary was crossed many times
    assert (np.abs(np.diff(x, axis=1)) > 0.5 * fov).any() == bool(cross)


def test_spin_outside_the_volume_is_lost_and_keeps_prior_outputs():
    cfg = cases.gre(1, TE=2000, dt=50, DIFFUSIVITY=[0.0])
    xyz0 = np.float32([[5e-6, 5e-6, 5e-6], [12e-6, 5e-6, 5e-6], [5e-6, 5e-6, 9.99e-6]])
    c = orc.ConfigCase(dict(cfg, NUMBER_OF_SPINS=3), np.zeros((4, 4, 4), np.uint8), None, 10e-6, XYZ0=xyz0)
    d = c.oracle()
    d["M1_init"] = np.full((3, 1, 3), 7.0, np.float32)
    d["T_init"] = np.full((3, 1), 9, np.uint8)
    r = orc.sim(d, n_threads=2)
    # x beyond the FoV on the slowest axis indexes past the volume: the spin is
    # abandoned before its first step (kernels.cu:102-106 reads out of bounds there)
    assert r["lost"].tolist() == [0, 1, 0]
    assert (r["M1"][1] == 7.0).all() and r["T"][1, 0] == 9
    assert r["T"][0, 0] == 0 and r["T"][2, 0] == 0
