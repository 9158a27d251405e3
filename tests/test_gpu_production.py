"""The production walk (Philox4x32-10 draws on the device).

Its draws differ from the reference's, so results are pinned statistically and
by what the mathematics fixes:
  - the generator: Random123 known answers are checked on the host build of the
    same code (test_philox_kat), the device draws for normal/uniform statistics;
  - signals agree with the oracle (the reference CPU path) within 4.5 standard
    errors of the Monte-Carlo difference;
  - free diffusion: increments are N(0, 2 D dt) per axis (simulation_parameters.cuh:236-237);
  - PGSE: the echo equals exp(-Var(phi)/2) with Var(phi) computed exactly for the
    discrete gradient waveform and Gaussian steps (Stejskal-Tanner in discrete time);
  - spin echo: |Mxy(TE)| = exp(-TE/T2) for every spin whatever its path;
  - results do not depend on how spins are ordered or split into launches;
  - precision modes: identical trajectories and labels; M within the stated
    error of the FP64 path; the int16 fixed-point field holds a field that is
    one constant inside the inclusions at its float32 value, where float16
    rounds it.
"""
import math

import numpy as np
import pytest

from tests import cases
from oracle import oracle as orc
from spinwalk_amd import engine as E

pytestmark = pytest.mark.gpu

GAMMA = 267515315.0


def run(engine, case, scale_index=0, precision=E.SW_PREC_F32, field_format=E.SW_FIELD_F32, spatial_order=1,
        spin_offset=0, n=None, formats=(E.SW_FIELD_F32,), step_loop=E.SW_STEP_LOOP_AUTO):
    p, xs, gs = case.scale_params(case.scales[scale_index])
    n = case.n_spins - spin_offset if n is None else n
    sl = slice(spin_offset, spin_offset + n)
    ph = engine.phantom(case.mask, case.field_T, case.fov, formats=formats)
    seq = engine.sequence(p, case.tables)
    dt = np.float64 if precision == E.SW_PREC_F64 else np.float32
    M1 = np.zeros((n, case.n_te, 3), dt)
    T = np.zeros((n, case.n_te), np.uint8)
    XYZ1 = np.zeros((n, case.traj_len, 3), np.float32)
    st = engine.walk_host(ph, seq, p, np.ascontiguousarray(case.M0[sl]), np.ascontiguousarray(case.XYZ0[sl]), M1,
                          XYZ1, T, spin_offset=spin_offset, xyz_scale=xs, gradient_scale=gs, precision=precision,
                          field_format=field_format, spatial_order=spatial_order, step_loop=step_loop)
    return dict(M1=M1, T=T, XYZ1=XYZ1, stats=st)


def signal(M1, T, te, label=None):
    sel = np.ones(T.shape[0], bool) if label is None else (T[:, te] == label)
    v = M1[sel, te, :2].astype(np.float64)
    mean = v.mean(axis=0)
    s = np.linalg.norm(mean)
    u = mean / max(s, 1e-300)
    var = float(u @ np.cov(v.T) @ u) / max(1, len(v))       # delta method
    return s, var, len(v)


# ---------------------------------------------------------------------------
def test_device_draws_equal_the_restatement(engine):
    """The device's draws (sw_philox_draws_device: the bundle each attempted step
    takes) against the numpy restatement of Philox4x32-10 and of the bundle, with
    a 64-bit seed and a spin offset: the permeability uniform bit for bit, the
    normals within the error of the hardware log2, sqrt, sin and cos."""
    import torch
    from tests.test_philox import draw_bundle
    seed, spin0, n_spins, n_att = 0x0123456789ABCDEF, 1000, 512, 64
    out = torch.zeros((n_spins, n_att, 4), dtype=torch.float32, device="cuda")
    engine.philox_draws(seed, spin0, n_spins, n_att, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = draw_bundle(seed, spin0, n_spins, n_att)
    np.testing.assert_array_equal(got[..., 3], want[..., 3].astype(np.float32))
    np.testing.assert_allclose(got[..., :3], want[..., :3], rtol=0, atol=2e-5)


def test_device_draw_statistics(engine):
    import torch
    from scipy import stats
    n_spins, n_att = 4096, 256
    out = torch.zeros((n_spins, n_att, 4), dtype=torch.float32, device="cuda")
    engine.philox_draws(10, 0, n_spins, n_att, out)
    torch.cuda.synchronize()
    d = out.cpu().numpy().astype(np.float64)
    for k in range(3):
        x = d[..., k].ravel()
        assert abs(x.mean()) < 5 / math.sqrt(x.size)
        assert abs(x.var() - 1) < 5 * math.sqrt(2 / x.size)
        assert abs(stats.kurtosis(x, fisher=False) - 3) < 0.02
        assert stats.kstest(x[:200000], "norm").pvalue > 1e-4
    u = d[..., 3].ravel()
    assert u.min() >= 0 and u.max() < 1
    assert stats.kstest(u[:200000], "uniform").pvalue > 1e-4
    lag = np.concatenate([d[:, 1:, 0], np.zeros((n_spins, 1))], axis=1).ravel()     # next attempt's x
    c = np.corrcoef(np.stack([d[..., 0].ravel(), d[..., 1].ravel(), d[..., 2].ravel(), d[..., 3].ravel(), lag]))
    assert np.abs(c - np.eye(5)).max() < 5 / math.sqrt(u.size)
    # a different seed gives different streams
    out2 = torch.zeros_like(out)
    engine.philox_draws(11, 0, n_spins, 4, out2[:, :4].contiguous())
    assert not torch.equal(out2[:, :4], out[:, :4])


@pytest.mark.parametrize("kind", ["se3", "gre2_field", "bssfp"])
def test_signals_agree_with_oracle_within_mc_error(engine, kind):
    if kind == "se3":
        m, f = cases.three_substrate_phantom(n=24, fov_um=48.0, seed=31)
        cfg = cases.se(3, TE=8000, dt=50, P_XY=[[1, 0.5, 0.5], [0.5, 1, 0.5], [0.5, 0.5, 1]],
                       T1=[2000.0, 1600.0, 1600.0], T2=[40.0, 20.0, 20.0])
        fov = 48.0
    elif kind == "gre2_field":
        m, f = cases.cylinder_phantom(n=24, fov_um=24.0, radius_um=3.0, n_cyl=3, seed=32)
        f = f * 10
        cfg = cases.gre(2, TE=6000, dt=50)
        fov = 24.0
    else:
        m, f = cases.cylinder_phantom(n=20, fov_um=40.0, radius_um=4.0, n_cyl=2, seed=33)
        cfg = cases.bssfp(2, TE=1000, dt=50, dummies=5, lin=180.0, T1=[300.0, 250.0], T2=[50.0, 40.0],
                          P_XY=[[1, 0.3], [0.3, 1]])
        fov = 40.0
    case = cases.make_case(cfg, m, f, fov, n_spins=20000)
    g = run(engine, case)
    o = orc.sim(cases.to_oracle(case))
    assert g["stats"]["n_lost"] == 0 and o["n_lost"] == 0
    labels = [None] + list(range(int(m.max()) + 1))
    for te in range(case.n_te):
        for lab in labels:
            sg, vg, ng = signal(g["M1"], g["T"], te, lab)
            so, vo, no = signal(o["M1"], o["T"], te, lab)
            if min(ng, no) < 200:
                continue
            z = abs(sg - so) / math.sqrt(vg + vo + 1e-30)
            assert z < 4.5, (kind, te, lab, sg, so, z)


def test_free_diffusion_increments(engine):
    mask = np.zeros((8, 8, 8), np.uint8)
    cfg = cases.gre(1, TE=5000, dt=50, DIFFUSIVITY=[2e-9], RECORD_TRAJECTORY=1, T1=[1e9], T2=[1e9])
    rng = np.random.default_rng(1)
    xyz0 = rng.uniform(0.4, 0.6, size=(2000, 3)).astype(np.float32) * np.float32(1000e-6)
    case = cases.make_case(cfg, mask, None, 1000.0, n_spins=2000, xyz0=xyz0)
    g = run(engine, case)
    X = g["XYZ1"].astype(np.float64)
    inc = np.diff(X, axis=1).reshape(-1, 3)
    sigma2 = 2 * np.float32(2e-9) * 50e-6
    n = inc.shape[0]
    for k in range(3):
        assert abs(inc[:, k].mean()) < 5 * math.sqrt(sigma2 / n)
        assert abs(inc[:, k].var() / sigma2 - 1) < 5 * math.sqrt(2 / n) + 1e-3
    c = np.corrcoef(inc.T)
    assert np.abs(c - np.eye(3)).max() < 5 / math.sqrt(n)


def test_walk_consumes_the_materialized_draws(engine):
    """Every attempted step k of spin s moves by draw[s, k, 0:3] times the step
    size sqrt(2 D dt) (simulation_parameters.cuh:236-237, kernels.cu:131), with
    draw = sw_philox_draws_device(seed, s, k): each increment of the recorded
    trajectory equals it to the float rounding of the positions (the production
    walk keeps positions in float voxel units).  One substrate, no walls within
    reach: every attempt is accepted."""
    import torch
    mask = np.zeros((8, 8, 8), np.uint8)
    cfg = cases.gre(1, TE=5000, dt=50, DIFFUSIVITY=[2e-9], RECORD_TRAJECTORY=1, T1=[1e9], T2=[1e9])
    rng = np.random.default_rng(3)
    n = 1000
    xyz0 = rng.uniform(0.45, 0.55, size=(n, 3)).astype(np.float32) * np.float32(1000e-6)
    case = cases.make_case(cfg, mask, None, 1000.0, n_spins=n, xyz0=xyz0)
    g = run(engine, case)
    n_t = case.traj_len
    assert g["stats"]["n_lost"] == 0 and n_t > 50
    assert g["stats"]["n_attempts"] == g["stats"]["n_steps"] == n * n_t
    d = torch.zeros((n, n_t, 4), dtype=torch.float32, device="cuda")
    engine.philox_draws(case.params.seed, 0, n, n_t, d)
    torch.cuda.synchronize()
    draws = d.cpu().numpy().astype(np.float64)
    diff = float(case.tables["diffusivity"][0])
    X = g["XYZ1"].astype(np.float64)
    prev = np.concatenate([xyz0.astype(np.float64)[:, None, :], X[:, :-1]], axis=1)
    err = np.abs((X - prev) - draws[..., :3] * diff)
    ulp = float(np.spacing(np.float32(600e-6)))                  # one float step of a position near 0.6 mm
    assert err.max() < 4 * ulp, (err.max(), ulp, diff)
    assert diff > 1e3 * ulp                                      # the check resolves a step to 1e-3 of its size


def test_production_boundaries_match_the_reference(engine):
    """The FoV boundary in the production walk (float voxel units) against the
    oracle's (the reference CPU path, kernels.cu:133-136): spins that start
    uniform over a small FoV and cross its faces many times stay inside it with
    none lost, and their final positions follow the oracle's distribution
    (two-sample KS per axis).  The reference's CROSS_FOV = 0 correction turns a
    crossing step back from the spin's old position (x_old + |step|), which is
    not a mirror image and thins the spins near the walls; wrap-around
    (CROSS_FOV = 1) keeps them uniform."""
    from scipy import stats
    mask = np.zeros((8, 8, 8), np.uint8)
    for cross in (0, 1):
        cfg = cases.gre(1, TE=20000, dt=50, DIFFUSIVITY=[3e-9], CROSS_FOV=cross, T1=[1e9], T2=[1e9])
        case = cases.make_case(cfg, mask, None, 8.0, n_spins=20000)      # steps of 0.55 um in an 8 um box
        g = run(engine, case)
        o = orc.sim(cases.to_oracle(case))
        assert g["stats"]["n_lost"] == 0 and o["n_lost"] == 0
        fov = float(case.fov[0])
        x = g["XYZ1"][:, 0, :].astype(np.float64) / fov
        xo = o["XYZ1"][:, 0, :].astype(np.float64) / fov
        assert (x >= 0).all() and (x <= 1).all()
        for k in range(3):
            assert stats.ks_2samp(x[:, k], xo[:, k]).pvalue > 1e-4, (cross, k)
            if cross:
                assert stats.kstest(x[:, k], "uniform").pvalue > 1e-4, k
        if not cross:                                   # the thinning near the walls is there in both
            edge = lambda v: float(((v < 0.03) | (v > 0.97)).mean())
            assert edge(x[:, 0]) < 0.05 and edge(xo[:, 0]) < 0.05


@pytest.mark.parametrize("dims", [(13, 7, 10), (8, 9, 12), (10, 12, 8)])
def test_production_gathers_every_voxel_exactly(engine, dims):
    """With zero diffusivity nothing moves and no draw is taken, so the
    production walk is deterministic: one spin at the centre of every voxel of a
    phantom whose sides are not multiples of the record bricks (2 x 3 x 4 for the
    production walk's float32 records, 4 x 2 x 2 for the reference layout),
    or whose brick counts along y and z are even, which the production layout
    pads to odd counts (10 x 12 x 8: 4 and 2 bricks laid out as 5 and 3),
    must see that voxel's label and field, as the oracle does:
    labels equal, M within 1e-5 (a GRE of 20 steps, each adding the voxel's
    field as a phase)."""
    m, f = cases.blocky_phantom(dims, n_labels=3, seed=7, block=2)
    f = f * 30
    cfg = cases.gre(3, TE=1000, dt=50, DIFFUSIVITY=[0.0, 0.0, 0.0], T1=[1000.0, 900.0, 800.0], T2=[60.0, 50.0, 40.0])
    fov_um = 2.0 * np.asarray(dims, float)
    idx = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, 3)
    xyz0 = ((idx + 0.5) * 2e-6).astype(np.float32)
    case = cases.make_case(cfg, m, f, fov_um, n_spins=len(xyz0), xyz0=xyz0)
    for prec in (E.SW_PREC_F32, E.SW_PREC_F64):
        g = run(engine, case, precision=prec)
        o = orc.sim(cases.to_oracle(case), f64=prec == E.SW_PREC_F64)
        assert g["stats"]["n_lost"] == 0 and g["stats"]["n_attempts"] == g["stats"]["n_steps"]
        np.testing.assert_array_equal(g["T"][:, 0], m.reshape(-1))
        np.testing.assert_array_equal(g["T"], o["T"])
        mag = np.linalg.norm(o["M1"], axis=-1, keepdims=True)
        assert (np.abs(g["M1"] - o["M1"]) <= 1e-5 * mag + 1e-7).all()


def _dequantized(f, fmt):
    """The float32 field a 16-bit record holds for each voxel, as the engine
    quantizes it (phantom.hip): float16 / bfloat16 of f / s, s = 2^(e - 15)
    with max|f| = m 2^e, m in [0.5, 1); int16 rint(f / q), q = max|f| / 32767."""
    f = np.asarray(f, np.float32)
    fmax = np.float32(np.abs(f).max())
    if fmt == E.SW_FIELD_Q16:
        q = np.float32(np.float64(fmax) / 32767.0)
        inv = np.float32(1.0) / q
        k = np.clip(np.rint(f.astype(np.float64) * np.float64(inv)), -32767, 32767)
        return (k.astype(np.float32) * q).astype(np.float32)
    _, ex = np.frexp(fmax)
    s = np.float32(2.0 ** (int(ex) - 15))
    x = (f * (np.float32(1.0) / s)).astype(np.float32)
    if fmt == E.SW_FIELD_F16:
        h = x.astype(np.float16).astype(np.float32)
    else:
        b = x.view(np.uint32).astype(np.uint64)
        h = (((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16).astype(np.uint32).view(np.float32)
    return (h * s).astype(np.float32)


@pytest.mark.parametrize("dims", [(13, 7, 10), (8, 9, 12), (10, 12, 8)])
@pytest.mark.parametrize("fmt", [E.SW_FIELD_F16, E.SW_FIELD_BF16, E.SW_FIELD_Q16])
def test_production_gathers_16bit_records_exactly(engine, dims, fmt):
    """The 16-bit records (packed 3-byte records in 3 x 3 x 4 bricks) on sides
    that are not multiples of the bricks, or whose brick counts along y and z
    are even and laid out padded to odd ones: with zero diffusivity every spin stays
    at the centre of its voxel, so it must see that voxel's label, and its
    magnetisation must be the oracle's for the field the record holds (the
    float32 field quantized as the engine quantizes it): labels equal, M within
    1e-5, in the Kahan phase the 16-bit fields are meant for."""
    m, f = cases.blocky_phantom(dims, n_labels=3, seed=8, block=2)
    f = (f * 30).astype(np.float32)
    cfg = cases.gre(3, TE=1000, dt=50, DIFFUSIVITY=[0.0, 0.0, 0.0], T1=[1000.0, 900.0, 800.0], T2=[60.0, 50.0, 40.0])
    fov_um = 2.0 * np.asarray(dims, float)
    idx = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, 3)
    xyz0 = ((idx + 0.5) * 2e-6).astype(np.float32)
    case = cases.make_case(cfg, m, f, fov_um, n_spins=len(xyz0), xyz0=xyz0)
    g = run(engine, case, precision=E.SW_PREC_F32_KAHAN, field_format=fmt, formats=(E.SW_FIELD_F32, fmt))
    ref = cases.make_case(cfg, m, _dequantized(f, fmt), fov_um, n_spins=len(xyz0), xyz0=xyz0)
    o = orc.sim(cases.to_oracle(ref))
    assert g["stats"]["n_lost"] == 0 and g["stats"]["n_attempts"] == g["stats"]["n_steps"]
    np.testing.assert_array_equal(g["T"][:, 0], m.reshape(-1))
    np.testing.assert_array_equal(g["T"], o["T"])
    mag = np.linalg.norm(o["M1"], axis=-1, keepdims=True)
    assert (np.abs(g["M1"] - o["M1"]) <= 1e-5 * mag + 1e-7).all()


def test_pgse_discrete_stejskal_tanner(engine):
    D, b, delta, DELTA, dt = 2e-9, 500.0, 5.0, 10.0, 50
    mask = np.zeros((8, 8, 8), np.uint8)
    cfg, G = cases.pgse(1, b=b, delta_ms=delta, DELTA_ms=DELTA, dt=dt, axis=(1, 0, 0), start_ms=1.0)
    cfg.update(DIFFUSIVITY=[D], T1=[1e9], T2=[1e9])
    n = 60000
    rng = np.random.default_rng(2)
    xyz0 = rng.uniform(0.45, 0.55, size=(n, 3)).astype(np.float32) * np.float32(1000e-6)
    case = cases.make_case(cfg, mask, None, 1000.0, n_spins=n, xyz0=xyz0)
    g = run(engine, case)
    # exact attenuation of the discrete sequence: phase = gamma' sum_k s_k G_k x(t_k),
    # x(t) = x0 + sum_{j<=t} xi_j, xi ~ N(0, sigma^2); s = -1 up to the refocusing pulse
    t_g = np.asarray(case.tables["gradient_t"])
    Gx = np.asarray(case.tables["gx"], np.float64)
    t_rf = int(case.tables["RF_t"][1])
    s = np.where(t_g <= t_rf, -1.0, 1.0)
    gp = 1e-3 * dt * 1e-6 * GAMMA
    assert abs((s * Gx).sum()) < 1e-9 * np.abs(Gx).sum()
    n_t = case.params.n_timepoints
    F = np.array([(s * Gx)[t_g >= j].sum() for j in range(n_t)])
    sigma2 = case.tables["diffusivity"][0] ** 2
    expect = math.exp(-0.5 * gp * gp * sigma2 * float((F * F).sum()))
    got, var, _ = signal(g["M1"], g["T"], 0)
    assert 0.2 < expect < 0.8
    assert abs(got - expect) < 4.5 * math.sqrt(var) + 1e-4, (got, expect)
    # and the nominal b-value of the waveform (src/dwi/pgse.cpp:80-83) is close to it
    assert abs(math.exp(-b * 1e6 * D) - expect) < 0.03


def test_spin_echo_magnitude_every_spin(engine):
    m, f = cases.cylinder_phantom(n=24, fov_um=48.0, radius_um=5.0, n_cyl=3, seed=34)
    cfg = cases.se(2, TE=10000, dt=50, T2=[40.0, 40.0], T1=[2200.0, 2200.0], P_XY=[[1, 0.5], [0.5, 1]])
    case = cases.make_case(cfg, m, f * 20, 48.0, n_spins=50000)
    g = run(engine, case)
    mag = np.linalg.norm(g["M1"][:, 0, :2].astype(np.float64), axis=1)
    np.testing.assert_allclose(mag, math.exp(-10.0 / 40.0), rtol=2e-5)


def test_impermeable_labels_never_change(engine):
    m, f = cases.three_substrate_phantom(n=24, fov_um=48.0, seed=35)
    cfg = cases.ste(3, dt=50, DIFFUSIVITY=[2e-9, 2e-9, 2e-9], MAX_ITERATIONS=1e6)
    case = cases.make_case(cfg, m, f, 48.0, n_spins=20000)
    g = run(engine, case)
    idx = np.floor(case.XYZ0.astype(np.float64) * (np.asarray(m.shape) / case.fov.astype(np.float64))).astype(int)
    lab0 = m[idx[:, 0], idx[:, 1], idx[:, 2]]
    assert g["stats"]["n_lost"] == 0
    for te in range(case.n_te):
        np.testing.assert_array_equal(g["T"][:, te], lab0)


def test_still_tissue_beside_mobile_tissue(engine):
    """A tissue of zero diffusivity beside a mobile one, spins spread over both
    and walked in index order, so nearly every wave holds lanes that move and
    lanes that do not (kernels.cu:130: no step without diffusivity).  A spin in
    the still tissue keeps its start position (to the float rounding of the
    voxel-unit conversion) and its label; a mobile spin walks, reflects off the
    impermeable interface and never enters the still tissue."""
    n_vox = 16
    m = np.zeros((n_vox,) * 3, np.uint8)
    m[: n_vox // 2] = 1                                           # x < fov/2: still tissue
    cfg = cases.gre(2, TE=4000, dt=50, DIFFUSIVITY=[2e-9, 0.0], P_XY=[[1, 0], [0, 1]],
                    T1=[1e9, 1e9], T2=[1e9, 1e9], MAX_ITERATIONS=1e6)
    rng = np.random.default_rng(37)
    n = 4096
    fov_um = 32.0
    xyz0 = rng.uniform(0.02, 0.98, size=(n, 3)) * fov_um * 1e-6
    case = cases.make_case(cfg, m, None, fov_um, n_spins=n, xyz0=xyz0.astype(np.float32))
    g = run(engine, case, spatial_order=0)
    still = case.XYZ0[:, 0].astype(np.float64) < 0.5 * fov_um * 1e-6
    assert 0.4 < still.mean() < 0.6 and g["stats"]["n_lost"] == 0
    X0, X1 = case.XYZ0.astype(np.float64), g["XYZ1"][:, 0, :].astype(np.float64)
    ulp = float(np.spacing(np.float32(fov_um * 1e-6)))
    assert np.abs(X1[still] - X0[still]).max() <= 2 * ulp
    np.testing.assert_array_equal(g["T"][still, 0], 1)
    np.testing.assert_array_equal(g["T"][~still, 0], 0)
    assert (X1[~still, 0] >= 0.5 * fov_um * 1e-6).all()
    step = np.sqrt(2 * 2e-9 * 50e-6)                              # metres per axis per step
    msd = ((X1[~still] - X0[~still]) ** 2).sum(axis=1).mean()
    assert msd > 10 * step ** 2                                   # 80 steps: they walked


def test_order_and_launch_split_invariance(engine):
    m, f = cases.three_substrate_phantom(n=24, fov_um=48.0, seed=36)
    cfg = cases.se(3, TE=6000, dt=50, P_XY=[[1, 0.5, 0.5], [0.5, 1, 0.5], [0.5, 0.5, 1]])
    case = cases.make_case(cfg, m, f, 48.0, n_spins=10000)
    a = run(engine, case, spatial_order=0)
    b = run(engine, case, spatial_order=1)
    c1 = run(engine, case, spatial_order=1, spin_offset=0, n=3333)
    c2 = run(engine, case, spatial_order=0, spin_offset=3333, n=10000 - 3333)
    for k in ("M1", "T", "XYZ1"):
        np.testing.assert_array_equal(a[k], b[k])
        np.testing.assert_array_equal(a[k], np.concatenate([c1[k], c2[k]]))
    assert a["stats"]["n_attempts"] > a["stats"]["n_steps"]


def test_precision_modes_against_fp64(engine):
    m, f = cases.cylinder_phantom(n=32, fov_um=64.0, radius_um=6.0, n_cyl=4, seed=37)
    cfg = cases.se(2, TE=40000, dt=20, P_XY=[[1, 0.5], [0.5, 1]], T2=[40.0, 20.0], T1=[2000.0, 1600.0])
    case = cases.make_case(cfg, m, f, 64.0, n_spins=20000)
    fm = (E.SW_FIELD_F32, E.SW_FIELD_F16, E.SW_FIELD_BF16, E.SW_FIELD_Q16)
    ref = run(engine, case, precision=E.SW_PREC_F64, formats=fm)
    res = {}
    for name, prec, ff in (("f32", E.SW_PREC_F32, E.SW_FIELD_F32), ("kahan", E.SW_PREC_F32_KAHAN, E.SW_FIELD_F32),
                           ("f16", E.SW_PREC_F32_KAHAN, E.SW_FIELD_F16),
                           ("bf16", E.SW_PREC_F32_KAHAN, E.SW_FIELD_BF16),
                           ("q16", E.SW_PREC_F32_KAHAN, E.SW_FIELD_Q16)):
        r = run(engine, case, precision=prec, field_format=ff, formats=fm)
        np.testing.assert_array_equal(r["T"], ref["T"])
        np.testing.assert_array_equal(r["XYZ1"], ref["XYZ1"])
        s_ref = signal(ref["M1"], ref["T"], 0)[0]
        s = signal(r["M1"], r["T"], 0)[0]
        rms = float(np.sqrt(np.mean((r["M1"].astype(np.float64) - ref["M1"]) ** 2)))
        res[name] = (abs(s - s_ref) / s_ref, rms)
    assert res["f32"][0] < 1e-5 and res["kahan"][0] < 1e-5
    assert res["kahan"][1] <= res["f32"][1]
    assert res["f16"][0] < 1e-4, res
    assert res["bf16"][0] < 2e-3, res
    assert res["q16"][0] < 1e-4, res


def test_q16_field_holds_a_constant_inside_field(engine):
    """Cylinders along B0: no field outside them and one constant c inside
    (2 pi (1-Y) dchi (cos^2 theta - 1/3) at theta = 0).  dchi is set so that c,
    after the power-of-two scale, falls 7 units of 16 above a float16 value
    (c / half_scale = 16391, relative rounding error 4.3e-4).  The int16 copy
    holds c as its largest step, so its walk matches the float32 field: the mean
    transverse magnetisation, compared as a complex number (a common phase error
    counts), stays within 1e-5 of the FP64 walk, at least ten times closer than
    the float16 copy."""
    _, f0 = cases.cylinder_phantom(n=32, fov_um=64.0, radius_um=8.0, n_cyl=4, seed=42, random_axis=False)
    c0 = float(np.abs(f0).max())
    mant, ex = math.frexp(c0)
    dchi = 0.11e-6 * 0.4 * (16391.0 / 32768.0) / mant
    m, f = cases.cylinder_phantom(n=32, fov_um=64.0, radius_um=8.0, n_cyl=4, seed=42, random_axis=False, dchi=dchi)
    vals = np.unique(f)
    assert len(vals) == 2 and vals[0] == 0.0
    c = float(vals[1])
    hs = math.ldexp(1.0, math.frexp(c)[1] - 15)
    assert abs(float(np.float16(c / hs)) - c / hs) / (c / hs) > 4e-4      # float16 rounds c by 4.3e-4
    cfg = cases.gre(2, TE=20000, dt=50, P_XY=[[1, 0.5], [0.5, 1]], T2=[40.0, 40.0], T1=[2000.0, 2000.0])
    case = cases.make_case(cfg, m, f, 64.0, n_spins=20000)
    fm = (E.SW_FIELD_F32, E.SW_FIELD_F16, E.SW_FIELD_Q16)
    ref = run(engine, case, precision=E.SW_PREC_F64, formats=fm)
    z_ref = ref["M1"][:, 0, 0].astype(np.float64).mean() + 1j * ref["M1"][:, 0, 1].astype(np.float64).mean()
    assert abs(z_ref) > 0.1
    err = {}
    for name, ff in (("f16", E.SW_FIELD_F16), ("q16", E.SW_FIELD_Q16)):
        r = run(engine, case, precision=E.SW_PREC_F32_KAHAN, field_format=ff, formats=fm)
        np.testing.assert_array_equal(r["T"], ref["T"])
        np.testing.assert_array_equal(r["XYZ1"], ref["XYZ1"])
        z = r["M1"][:, 0, 0].astype(np.float64).mean() + 1j * r["M1"][:, 0, 1].astype(np.float64).mean()
        err[name] = abs(z - z_ref) / abs(z_ref)
    assert err["q16"] < 1e-5, err
    assert err["q16"] * 10 < err["f16"], err


def test_echo_reduce_matches_numpy(engine):
    import torch
    rng = np.random.default_rng(3)
    M1 = rng.normal(size=(70000, 3, 3)).astype(np.float32)
    T = rng.integers(0, 4, size=(70000, 3)).astype(np.uint8)
    out = torch.zeros((3, 4, 3), dtype=torch.float64, device="cuda")
    engine.echo_reduce(torch.from_numpy(M1).cuda(), torch.from_numpy(T).cuda(), 4, out)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = orc.echo_reduce(M1, T, 4)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-9)


def test_walk_rejects_mismatched_buffers(engine):
    """The C entry points take bare pointers; the bindings refuse buffers whose
    shape or dtype does not match the walk, and the C side refuses what it can
    check itself (missing buffers, unprepared parameters)."""
    m, f = cases.cylinder_phantom(n=16, fov_um=32.0, radius_um=4.0, n_cyl=2, seed=5)
    case = cases.make_case(cases.gre(2, TE=2000, dt=50), m, f, 32.0, n_spins=64)
    p, xs, gs = case.scale_params(case.scales[0])
    ph = engine.phantom(case.mask, case.field_T, case.fov)
    seq = engine.sequence(p, case.tables)
    n = case.n_spins
    good = dict(M0=case.M0, XYZ0=case.XYZ0, M1=np.zeros((n, case.n_te, 3), np.float32),
                XYZ1=np.zeros((n, case.traj_len, 3), np.float32), T=np.zeros((n, case.n_te), np.uint8))

    def call(precision=E.SW_PREC_F32, params=p, **kw):
        b = dict(good, **kw)
        return engine.walk_host(ph, seq, params, b["M0"], b["XYZ0"], b["M1"], b["XYZ1"], b["T"], precision=precision)

    assert call()["n_lost"] == 0
    with pytest.raises(E.EngineError, match="M1"):
        call(precision=E.SW_PREC_F64)                                   # float32 M1 for the FP64 path
    with pytest.raises(E.EngineError, match="XYZ1"):
        call(XYZ1=np.zeros((n - 1, case.traj_len, 3), np.float32))
    with pytest.raises(E.EngineError, match="T"):
        call(T=np.zeros((n, case.n_te), np.int32))
    with pytest.raises(E.EngineError, match="field format"):
        engine.walk_host(ph, seq, p, good["M0"], good["XYZ0"], good["M1"], good["XYZ1"], good["T"],
                         field_format=E.SW_FIELD_Q16)                    # the phantom holds no int16 copy
    q = E.SwParams.from_buffer_copy(p)
    q.n_dummy_scan = -1
    with pytest.raises(E.EngineError, match="n_dummy_scan"):
        call(params=q)
    q = E.SwParams.from_buffer_copy(p)
    q.n_timepoints = 1 << 31                                            # the walk flags lost spins in bit 31
    with pytest.raises(E.EngineError, match="n_timepoints"):
        call(params=q)
    seq.close()
    ph.close()
