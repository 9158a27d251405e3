"""Pins of the oracle's primitives and host transforms (CPU only).

Every check here comes from outside the oracle's own code: the reference's unit
tests (tests/test_kernel.cpp:17-89), closed forms (axis-angle rotation, Bloch
relaxation), the C++ standard's known answer for mt19937, an independent numpy
restatement of the published libstdc++ algorithms (Marsaglia polar method over
generate_canonical<float,24>), and the arithmetic the reference's host code
performs (integer division, the T->deg/step factor).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc

TOL = 1e-5   # tests/test_kernel.cpp uses 1e-5 absolute


# ---- tests/test_kernel.cpp:17-34 -------------------------------------------
def test_sub2ind_3d_row_major():
    assert orc.sub2ind3(1, 2, 3, 10, 10, 10) == 1 * 10 * 10 + 2 * 10 + 3


def test_sub2ind_4d_row_major():
    assert orc.sub2ind4(1, 2, 3, 4, 10, 10, 10, 10) == 1 * 1000 + 2 * 100 + 3 * 10 + 4


def test_sub2ind_non_cubic_z_fastest():
    # row-major, z fastest: neighbours in z are adjacent in memory
    nx, ny, nz = 5, 7, 11
    a = np.arange(nx * ny * nz).reshape(nx, ny, nz)
    for (x, y, z) in [(0, 0, 0), (4, 6, 10), (2, 3, 7), (1, 0, 9)]:
        assert orc.sub2ind3(x, y, z, nx, ny, nz) == a[x, y, z]


def _sc(deg):
    return np.float32(math.sin(deg * math.pi / 180.0)), np.float32(math.cos(deg * math.pi / 180.0))


# ---- tests/test_kernel.cpp:36-89 -------------------------------------------
def test_xrot_90():
    s, c = _sc(90.0)
    np.testing.assert_allclose(orc.xrot(s, c, [0, 0, 1]), [0, -1, 0], atol=TOL)


def test_yrot_90():
    s, c = _sc(90.0)
    np.testing.assert_allclose(orc.yrot(s, c, [1, 0, 0]), [0, 0, -1], atol=TOL)


def test_zrot_90():
    s, c = _sc(90.0)
    np.testing.assert_allclose(orc.zrot(s, c, [1, 0, 0]), [0, 1, 0], atol=TOL)


def test_relax():
    m0 = np.array([1.0, 0.5, -0.5], np.float32)
    m1 = orc.relax(0.9, 0.8, m0)
    np.testing.assert_allclose(m1, [m0[0] * 0.8, m0[1] * 0.8, 1.0 + 0.9 * (m0[2] - 1.0)], atol=TOL)


# ---- closed forms ------------------------------------------------------------
def _R(axis, a):
    """Right-handed rotation by angle a about a unit axis (Rodrigues)."""
    n = np.asarray(axis, float)
    n = n / np.linalg.norm(n)
    K = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
    return np.eye(3) + math.sin(a) * K + (1 - math.cos(a)) * (K @ K)


@pytest.mark.parametrize("deg", [-170.0, -33.0, 12.5, 90.0, 179.0, 321.0])
def test_axis_rotations_match_rodrigues(deg):
    rng = np.random.default_rng(1)
    m0 = rng.normal(size=3).astype(np.float32)
    a = math.radians(deg)
    for f, axis in ((orc.xrot_deg, (1, 0, 0)), (orc.yrot_deg, (0, 1, 0)), (orc.zrot_deg, (0, 0, 1))):
        np.testing.assert_allclose(f(deg, m0), _R(axis, a) @ m0, atol=2e-6)


@pytest.mark.parametrize("ph", [0.0, 90.0, 180.0, 270.0, -90.0, 33.0, 123.4, 200.0, 359.0])
@pytest.mark.parametrize("fa", [16.0, 90.0, 180.0, 63.0])
def test_xrot_withphase_is_rotation_about_phase_axis(ph, fa):
    # Rz(ph) Rx(fa) Rz(-ph): rotation by fa about (cos ph, sin ph, 0) (kernels.cuh:155-206)
    rng = np.random.default_rng(int(abs(ph) * 10 + fa))
    for _ in range(4):
        m0 = rng.normal(size=3).astype(np.float32)
        want = _R((math.cos(math.radians(ph)), math.sin(math.radians(ph)), 0.0), math.radians(fa)) @ m0
        np.testing.assert_allclose(orc.xrot_withphase_deg(fa, ph, m0), want, atol=3e-6)
        s, c = _sc(fa)
        np.testing.assert_allclose(orc.xrot_withphase(s, c, ph, m0), want, atol=3e-6)


def test_xrot_withphase_special_cases_are_exact():
    m0 = np.array([0.3, -0.7, 0.64], np.float32)
    s, c = _sc(37.0)
    np.testing.assert_array_equal(orc.xrot_withphase(s, c, 0.0, m0), orc.xrot(s, c, m0))
    np.testing.assert_array_equal(orc.xrot_withphase(s, c, 180.0, m0), orc.xrot(-s, c, m0))
    np.testing.assert_array_equal(orc.xrot_withphase(s, c, 90.0, m0), orc.yrot(s, c, m0))
    np.testing.assert_array_equal(orc.xrot_withphase(s, c, 270.0, m0), orc.yrot(-s, c, m0))
    np.testing.assert_array_equal(orc.xrot_withphase(s, c, -90.0, m0), orc.yrot(-s, c, m0))


@pytest.mark.parametrize("phase,T1,T2,t", [(37.0, 1.2, 0.04, 0.01), (-400.0, 2.0, 0.05, 0.02),
                                           (12.0, -1.0, 0.04, 0.01), (12.0, 1.0, -1.0, 0.01)])
def test_dephase_relax_closed_form(phase, T1, T2, t):
    m0 = np.array([0.2, -0.9, 0.3], np.float32)
    m = _R((0, 0, 1), math.radians(phase)) @ m0.astype(float)
    if T1 >= 0 and T2 >= 0:
        e1, e2 = math.exp(-t / T1), math.exp(-t / T2)
        m = np.array([m[0] * e2, m[1] * e2, 1 + e1 * (m[2] - 1)])
    np.testing.assert_allclose(orc.dephase_relax(m0, phase, T1, T2, t), m, atol=3e-6)


# ---- libstdc++ draw streams --------------------------------------------------
def test_mt19937_standard_known_answer():
    # C++ standard [rand.predef]: the 10000th consecutive output of a
    # default-constructed mt19937 (seed 5489) is 4123659995
    assert orc.mt19937_raw(5489, 10000)[-1] == 4123659995


def _canonical_f32(r):
    """generate_canonical<float,24>(mt19937): one 32-bit word, float(r) / 2^32,
    clamped below 1 (libstdc++ random.tcc, published algorithm)."""
    u = (np.float32(r) / np.float32(4294967296.0)).astype(np.float32)
    return np.where(u >= 1, np.nextafter(np.float32(1), np.float32(0)), u).astype(np.float32)


def _polar_normals(raw, n):
    """Marsaglia polar method as normal_distribution<float> applies it: rejects
    r2 > 1 or r2 == 0, returns y*mult first and caches x*mult for the next call."""
    out, i = [], 0
    while len(out) < n:
        u1, u2 = _canonical_f32(raw[i]), _canonical_f32(raw[i + 1])
        i += 2
        x = np.float32(np.float64(np.float32(2.0) * u1) - 1.0)
        y = np.float32(np.float64(np.float32(2.0) * u2) - 1.0)
        r2 = np.float32(np.float32(x * x) + np.float32(y * y))
        if r2 > 1.0 or r2 == 0.0:
            continue
        mult = np.float32(math.sqrt(np.float32(np.float32(-2.0 * math.log(float(r2))) / r2)))
        out.append(np.float32(y * mult))
        out.append(np.float32(x * mult))
    return np.asarray(out[:n], np.float32), i


@pytest.mark.parametrize("seed,spin", [(10, 0), (10, 7), (123, 1000)])
def test_recorded_draws_follow_published_algorithms(seed, spin):
    nn, nu = 300, 50
    normals, uniforms = orc.record_draws(seed, spin, nn, nu)
    skip = seed + spin                      # gen.discard(seed+spin) (kernels.cu:87-88)
    raw = orc.mt19937_raw(seed + spin, skip + 4 * nn + nu)[skip:]
    want_n, _ = _polar_normals(raw, nn)
    # log/sqrt of numpy vs glibc may differ in the last place
    np.testing.assert_allclose(normals, want_n, rtol=4e-7, atol=1e-7)
    np.testing.assert_array_equal(uniforms, _canonical_f32(raw[:nu]))


def test_recorded_normals_are_standard_normal():
    x, _ = orc.record_draws(10, 3, 200000, 1)
    assert abs(x.mean()) < 5 * (1 / math.sqrt(x.size))
    assert abs(x.var() - 1.0) < 5 * math.sqrt(2.0 / x.size)


# ---- host transforms ---------------------------------------------------------
def test_timing_scale_truncates_toward_zero():
    v = np.array([0, 49, 50, 99, 10000, -75, -125, 20000], np.int32)
    got = orc.timing_scale(v, 50)
    np.testing.assert_array_equal(got, np.trunc(v / 50).astype(np.int32))


def test_str2vec_parses_floats_then_truncates():
    np.testing.assert_array_equal(orc.str2vec_i32("5e3 6e3 7e3"), [5000, 6000, 7000])
    np.testing.assert_array_equal(orc.str2vec_i32(" 0 100e3  200e3 "), [0, 100000, 200000])
    np.testing.assert_array_equal(orc.str2vec_i32("1.9 -2.5"), [1, -2])


def test_tesla2deg_closed_form():
    for B0, dt in ((9.4, 50), (3.0, 20), (7.0, 5)):
        want = B0 * dt * 1e-6 * 267515315.0 * 180.0 / math.pi
        got = orc.tesla2deg(B0, dt)
        assert abs(got - want) <= 2 * np.spacing(np.float32(want))


def test_prepare_closed_forms():
    p = orc.prepare(90.0, 20050, 50, 2200.0, [1e-9, 2e-9, 0.0], -1)
    assert p["n_timepoints"] == 20050 // 50
    np.testing.assert_allclose(p["diffusivity"], np.sqrt(2 * np.array([1e-9, 2e-9, 0.0]) * 50e-6), rtol=1e-12)
    assert p["n_dummy_scan"] == int(5 * 2200.0 / (20050 * 1e-3))
    assert abs(p["c"]) < 1e-7 and abs(p["s"] - 1.0) < 1e-7
    q = orc.prepare(16.0, 5000, 50, 1000.0, [1e-9], 3)
    assert q["n_dummy_scan"] == 3
    assert abs(q["c"] - math.cos(math.radians(16))) < 1e-7


def test_init_positions_inside_98_percent_and_published_algorithm():
    fov = np.array([100e-6, 50e-6, 80e-6], np.float32)
    xyz = orc.init_positions(10, fov, 2000)
    lo, hi = (0.01 * fov.astype(float)).astype(np.float32), (0.99 * fov.astype(float)).astype(np.float32)
    assert (xyz >= lo).all() and (xyz <= hi).all()
    # uniform_real_distribution<float>(a,b): canonical*(b-a)+a in float, x,y,z per spin
    raw = orc.mt19937_raw(10, 6)
    u = _canonical_f32(raw)
    want = (u[:3] * (hi - lo) + lo).astype(np.float32)
    np.testing.assert_array_equal(xyz[0], want)


def test_echo_reduce_brute_force():
    rng = np.random.default_rng(0)
    M1 = rng.normal(size=(500, 3, 3)).astype(np.float32)
    T = rng.integers(0, 3, size=(500, 3)).astype(np.uint8)
    got = orc.echo_reduce(M1, T, 3)
    for te in range(3):
        for lab in range(3):
            sel = T[:, te] == lab
            np.testing.assert_allclose(got[te, lab, 0], M1[sel, te, 0].astype(np.float64).sum(), rtol=1e-12)
            np.testing.assert_allclose(got[te, lab, 1], M1[sel, te, 1].astype(np.float64).sum(), rtol=1e-12)
            assert got[te, lab, 2] == sel.sum()
