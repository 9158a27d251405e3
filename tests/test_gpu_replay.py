"""Replay parity: the HIP walk fed the reference CPU path's own draw streams
reproduces the oracle (a restatement of src/sim/kernels.cu:72-233) bit for bit.

Bar: trajectories (XYZ1, float32) and tissue labels T are bit-identical, the set
of abandoned spins is identical, and M agrees within 1e-5 relative to the
magnitude of each recorded vector (plus 1e-7 absolute for vectors near zero).
M is not bit-exact only because sinf/cosf/exp of the device math library and of
glibc may differ in the last place.  The FP64 path agrees with the oracle's FP64
twin within 1e-12 relative.
"""
import numpy as np
import pytest

from tests import cases
from oracle import oracle as orc
from spinwalk_amd import engine as E

pytestmark = pytest.mark.gpu

M_RTOL, M_ATOL = 1e-5, 1e-7


def _replay_streams(case, n_att_factor=3.0):
    steps = case.steps_per_spin()
    nn = int(3 * steps * n_att_factor) + 96
    nu = int(steps * n_att_factor) + 64
    return E.replay_fill(case.params.seed, 0, case.n_spins, nn, nu)


def _run_engine(engine, case, scale_index=0, precision=E.SW_PREC_F32, replay=None, spin_offset=0, n=None,
                M1_init=None, T_init=None, spatial_order=0):
    p, xs, gs = case.scale_params(case.scales[scale_index])
    n = case.n_spins - spin_offset if n is None else n
    sl = slice(spin_offset, spin_offset + n)
    ph = engine.phantom(case.mask, case.field_T, case.fov)
    seq = engine.sequence(p, case.tables)
    dt = np.float64 if precision == E.SW_PREC_F64 else np.float32
    M1 = np.zeros((n, case.n_te, 3), dt) if M1_init is None else np.array(M1_init[sl], dt)
    T = np.zeros((n, case.n_te), np.uint8) if T_init is None else np.array(T_init[sl], np.uint8)
    XYZ1 = np.zeros((n, case.traj_len, 3), np.float32)
    rp = None if replay is None else (replay[0][sl], replay[1][sl])
    st = engine.walk_host(ph, seq, p, np.ascontiguousarray(case.M0[sl]), np.ascontiguousarray(case.XYZ0[sl]),
                          M1, XYZ1, T, spin_offset=spin_offset, xyz_scale=xs, gradient_scale=gs,
                          rng=E.SW_RNG_REPLAY if replay is not None else E.SW_RNG_PHILOX, precision=precision,
                          field_format=E.SW_FIELD_F32, replay=rp, spatial_order=spatial_order)
    return dict(M1=M1, T=T, XYZ1=XYZ1, stats=st)


def _assert_parity(g, o, f64=False, sl=slice(None)):
    assert g["stats"]["n_replay_exhausted"] == 0
    assert g["stats"]["n_lost"] == int(o["lost"][sl].sum())
    np.testing.assert_array_equal(g["XYZ1"], o["XYZ1"][sl])
    np.testing.assert_array_equal(g["T"], o["T"][sl])
    gm, om = g["M1"], o["M1"][sl]
    mag = np.linalg.norm(om, axis=-1, keepdims=True)
    tol = (1e-12 * mag + 1e-14) if f64 else (M_RTOL * mag + M_ATOL)
    bad = np.abs(gm - om) > tol
    assert not bad.any(), f"M mismatch at {np.argwhere(bad)[:5]}: gpu={gm[bad][:5]} oracle={om[bad][:5]}"


def _check(engine, cfg, mask, field, fov_um, n_spins=256, scales=(0,), f64=True, **kw):
    case = cases.make_case(cfg, mask, field, fov_um, n_spins=n_spins)
    rp = _replay_streams(case)
    for si in scales:
        d = cases.to_oracle(case, si)
        o = orc.sim(d)
        g = _run_engine(engine, case, si, replay=rp, **kw)
        _assert_parity(g, o)
        if f64:
            o64 = orc.sim(d, f64=True)
            g64 = _run_engine(engine, case, si, precision=E.SW_PREC_F64, replay=rp, **kw)
            _assert_parity(g64, o64, f64=True)
    return case


def test_replay_gre_two_substrates(engine):
    m, f = cases.cylinder_phantom(n=24, fov_um=48.0, radius_um=5.0, n_cyl=3, seed=11)
    _check(engine, cases.gre(2, TE=5000, dt=50), m, f, 48.0)


def test_replay_se_three_substrates_permeable_wrap(engine):
    m, f = cases.three_substrate_phantom()
    pxy = [[1.0, 0.5, 0.3], [0.5, 1.0, 0.2], [0.7, 0.4, 1.0]]
    cfg = cases.se(3, TE=10000, dt=25, P_XY=pxy, CROSS_FOV=1,
                   DIFFUSIVITY=[1e-9, 0.5e-9, 2e-9], T1=[2000.0, 1600.0, 1200.0], T2=[40.0, 20.0, 30.0])
    _check(engine, cfg, m, f, 48.0, n_spins=384)


def test_replay_bssfp_dummies_phase_cycling(engine):
    m, f = cases.cylinder_phantom(n=20, fov_um=40.0, radius_um=4.0, n_cyl=2, seed=5)
    cfg = cases.bssfp(2, TE=1000, dt=50, dummies=4, lin=180.0, QUADRATIC_PHASE_CYCLING=37.5,
                      T1=[400.0, 300.0], T2=[60.0, 50.0], P_XY=[[1.0, 0.2], [0.2, 1.0]])
    _check(engine, cfg, m, f, 40.0)


def test_replay_stimulated_echo_dephasing(engine):
    m, f = cases.cylinder_phantom(n=16, fov_um=32.0, radius_um=3.0, n_cyl=2, seed=7)
    _check(engine, cases.ste(2, dt=100, NUMBER_OF_SPINS=200), m, f, 32.0, n_spins=200)


def test_replay_pgse_gradient_scales(engine):
    m, f = cases.cylinder_phantom(n=16, fov_um=32.0, radius_um=3.0, n_cyl=2, seed=8)
    cfg, _ = cases.pgse(2, b=1000.0, delta_ms=2.0, DELTA_ms=4.0, dt=50, axis=(1, 1, 0))
    cfg.update(WHAT_TO_SCALE=1, SCALE=[0.5, 1.0, 1.7], CROSS_FOV=1)
    _check(engine, cfg, m, f, 32.0, scales=(0, 1, 2))


def test_replay_fov_scales(engine):
    m, f = cases.cylinder_phantom(n=16, fov_um=32.0, radius_um=3.0, n_cyl=3, seed=9)
    cfg = cases.gre(2, TE=4000, dt=50, WHAT_TO_SCALE=0, SCALE=[0.3, 1.0, 2.5], P_XY=[[1, 0.3], [0.6, 1]])
    _check(engine, cfg, m, f, 32.0, scales=(0, 1, 2))


def test_replay_phase_cycling_scales(engine):
    m, f = cases.cylinder_phantom(n=16, fov_um=32.0, radius_um=3.0, n_cyl=2, seed=10)
    cfg = cases.bssfp(2, TE=500, dt=50, dummies=3, lin=90.0, WHAT_TO_SCALE=2, SCALE=[0.0, 1.0, 2.0, 3.0])
    _check(engine, cfg, m, f, 32.0, scales=(0, 1, 2, 3))


def test_replay_trajectory(engine):
    m, f = cases.cylinder_phantom(n=16, fov_um=32.0, radius_um=3.0, n_cyl=2, seed=12)
    cfg = cases.bssfp(2, TE=500, dt=50, dummies=2, RECORD_TRAJECTORY=1, P_XY=[[1, 0.5], [0.5, 1]])
    _check(engine, cfg, m, f, 32.0, n_spins=128)


def test_replay_no_field_negative_T1_zero_diffusivity(engine):
    m, _ = cases.three_substrate_phantom(n=16, fov_um=32.0, seed=13)
    cfg = cases.se(3, TE=4000, dt=50, DIFFUSIVITY=[1e-9, 0.0, 1.5e-9], T1=[-1.0, 1000.0, 800.0],
                   T2=[40.0, -1.0, 30.0], P_XY=[[1, 0.5, 0.5], [0.5, 1, 0.5], [0.5, 0.5, 1]])
    _check(engine, cfg, m, None, 32.0)


def test_replay_lost_spins_out_of_range_and_stuck(engine):
    # huge steps with reflection push spins past the FoV (kernels.cu:141-147);
    # an impermeable pocket with a tiny MAX_ITERATIONS abandons stuck spins (:154-159)
    m, f = cases.cylinder_phantom(n=12, fov_um=24.0, radius_um=2.0, n_cyl=3, seed=14)
    cfg = cases.gre(2, TE=2000, dt=50, DIFFUSIVITY=[3e-7, 3e-7], P_XY=[[1, 0], [0, 1]], MAX_ITERATIONS=3)
    case = _check(engine, cfg, m, f, 24.0, n_spins=256)
    o = orc.sim(cases.to_oracle(case))
    assert o["n_lost"] > 0


def test_replay_spin_offset_and_prior_outputs(engine):
    # a slice of spins [96, 224) of a 256-spin run: RNG streams and dephasing use
    # the global spin index; outputs of spins abandoned before an echo keep their
    # prior contents (the reference's behaviour on reused host buffers)
    m, f = cases.cylinder_phantom(n=12, fov_um=24.0, radius_um=2.0, n_cyl=3, seed=15)
    cfg = cases.ste(2, dt=100, DIFFUSIVITY=[2e-7, 2e-7], P_XY=[[1, 0], [0, 1]], MAX_ITERATIONS=2)
    case = cases.make_case(cfg, m, f, 24.0, n_spins=256)
    rp = _replay_streams(case)
    d = cases.to_oracle(case)
    rng = np.random.default_rng(0)
    d["M1_init"] = rng.normal(size=(256, case.n_te, 3)).astype(np.float32)
    d["T_init"] = rng.integers(0, 200, size=(256, case.n_te)).astype(np.uint8)
    o = orc.sim(d)
    assert o["n_lost"] > 0
    sl = slice(96, 224)
    g = _run_engine(engine, case, replay=rp, spin_offset=96, n=128, M1_init=d["M1_init"], T_init=d["T_init"])
    _assert_parity(g, o, sl=sl)


def test_replay_spatial_order_is_invisible(engine):
    m, f = cases.three_substrate_phantom(n=24, fov_um=48.0, seed=16)
    cfg = cases.se(3, TE=6000, dt=50, P_XY=[[1, 0.5, 0.5], [0.5, 1, 0.5], [0.5, 0.5, 1]])
    case = cases.make_case(cfg, m, f, 48.0, n_spins=2048)
    rp = _replay_streams(case)
    o = orc.sim(cases.to_oracle(case))
    g = _run_engine(engine, case, replay=rp, spatial_order=1)
    _assert_parity(g, o)


def test_replay_odd_dims_wild_wrap_out_of_axis_subscripts(engine):
    # non-cubic, non-multiple-of-brick dims; steps far larger than the FoV with
    # wrap-around leave positions outside [0, fov), where the reference's int64
    # sub2ind lands in neighbouring rows or outside the volume (kernels.cu:133-147)
    m, f = cases.blocky_phantom((13, 11, 7), n_labels=3, seed=3)
    cfg = cases.gre(3, TE=1000, dt=50, CROSS_FOV=1, DIFFUSIVITY=[4e-6, 3e-6, 5e-6],
                    P_XY=[[1, 0.6, 0.4], [0.5, 1, 0.7], [0.3, 0.8, 1]])
    case = cases.make_case(cfg, m, f, (26.0, 22.0, 14.0), n_spins=512)
    rp = _replay_streams(case)
    d = cases.to_oracle(case)
    o = orc.sim(d)
    assert 0 < o["n_lost"] < 512
    g = _run_engine(engine, case, replay=rp)
    _assert_parity(g, o)


def test_replay_long_slab_64bit_index_path(engine):
    # nx+ny+nz > 16384: the per-axis tables would not fit in 64 KiB of LDS, so the
    # walk takes the 64-bit index arithmetic
    m, f = cases.blocky_phantom((16400, 3, 3), n_labels=2, seed=4, block=5)
    cfg = cases.gre(2, TE=500, dt=50, P_XY=[[1, 0.5], [0.5, 1]], DIFFUSIVITY=[1e-7, 1e-7])
    case = cases.make_case(cfg, m, f, (16400.0, 3.0, 3.0), n_spins=256)
    rp = _replay_streams(case)
    o = orc.sim(cases.to_oracle(case))
    g = _run_engine(engine, case, replay=rp)
    _assert_parity(g, o)


def test_replay_records_beyond_4gb(engine):
    """1024^3 voxels with an FP32 field: 8.6 GB of voxel records, beyond 32-bit byte
    offsets, so the walk takes the 64-bit index arithmetic at full size.  The
    spins walk near the far x end, where every record lies beyond 4 GB."""
    n = 1024
    m = np.zeros((n, n, n), np.uint8)
    f = np.zeros((n, n, n), np.float32)
    sub = (slice(960, 1024), slice(480, 544), slice(480, 544))
    bm, bf = cases.blocky_phantom((64, 64, 64), n_labels=2, seed=9, block=4)
    m[sub] = bm
    f[sub] = bf
    cfg = cases.gre(2, TE=2500, dt=50, P_XY=[[1, 0.5], [0.5, 1]], DIFFUSIVITY=[1e-9, 1e-9])
    rng = np.random.default_rng(9)
    lo, hi = np.float32([975, 495, 495]), np.float32([1010, 530, 530])
    xyz0 = (rng.uniform(0, 1, size=(256, 3)).astype(np.float32) * (hi - lo) + lo) * np.float32(1e-6)
    case = cases.make_case(cfg, m, f, float(n), n_spins=256, xyz0=xyz0)
    rp = _replay_streams(case)
    o = orc.sim(cases.to_oracle(case))
    g = _run_engine(engine, case, replay=rp)
    assert g["stats"]["n_lost"] == 0
    _assert_parity(g, o)
