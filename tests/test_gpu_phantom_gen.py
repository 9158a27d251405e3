"""The device phantom generator (sw_phantom_generate).

What fixes the expected values:
  - volume fraction: the mask's voxel fraction is within 2 percentage points of
    the target, the reference's own acceptance test (tests/test_phantom.cpp:15-30);
  - closed-form fields (src/phantom/phantom_cylinder.cpp:254-265,
    src/phantom/phantom_sphere.cpp:185-187), checked from the mask alone:
      * axons parallel to B0 have no outside field and 2 pi (1-Y) dChi (1 - 1/3)
        inside, so the field map equals that constant times the mask, exactly;
      * a single axon perpendicular to B0: 2 pi (1-Y) dChi a^2/r^2 cos(2 phi)
        outside, with centre and radius measured from its voxels;
      * a single sphere: 4 pi (1-Y) dChi a^3/r^3 (cos^2 theta - 1/3) outside, 0 inside;
  - packed axons with random radii reach a volume fraction beyond what equal
    disks reach by random sequential addition (about 55 %), and no voxel carries
    two shapes' interior fields.
"""
import math

import numpy as np
import pytest

from spinwalk_amd import engine as E
from spinwalk_amd import workloads as W

pytestmark = pytest.mark.gpu
DCHI, Y = W.DCHI, W.Y_OXY


def axons(radius, vf, seed, orientation=0.0, rmin=0.0, trunc=0.0, label=1):
    return dict(kind=E.SW_SHAPE_AXONS, label=label, radius_um=radius, volume_fraction=vf, random_orientation=0,
                orientation_deg=orientation, dchi=DCHI, Y=Y, seed=seed, trunc_radii=trunc, radius_min_um=rmin)


def _gen(engine, n, fov, specs):
    ph, vf = engine.generate_phantom((n, n, n), (fov, fov, fov), specs, formats=(E.SW_FIELD_F32,))
    m, f = ph.download(field=True)
    ph.close()
    return m, f, vf


@pytest.mark.parametrize("spec", [
    W.cylinders(5.0, 4.0, seed=1),
    W.spheres(4.0, 6.0, seed=2),
    axons(3.0, 30.0, seed=3, orientation=90.0),
])
def test_volume_fraction_within_two_points(engine, spec):
    m, f, vf = _gen(engine, 160, 160.0, [spec])
    got = 100.0 * np.count_nonzero(m) / m.size
    assert abs(got - spec["volume_fraction"]) < 2.0, (got, vf)


def test_packed_axons_parallel_to_b0_field_is_the_inside_constant(engine):
    m, f, vf = _gen(engine, 128, 256.0, [axons(-3.0, 62.0, seed=4, rmin=0.5)])
    frac = 100.0 * np.count_nonzero(m) / m.size
    assert frac > 57.0, (frac, vf)            # beyond random sequential addition of equal disks
    k_in = np.float32(2 * math.pi * (1 - np.float32(Y)) * np.float32(DCHI) * (1 - 1 / 3.0))
    want = np.where(m > 0, k_in, np.float32(0)).astype(np.float32)
    np.testing.assert_allclose(f, want, rtol=1e-6, atol=0)
    # axons are parallel to z: every z slice of the mask is the same
    assert (m == m[:, :, :1]).all()


def test_single_axon_perpendicular_to_b0_closed_form(engine):
    n, fov = 96, 96.0
    m, f, _ = _gen(engine, n, fov, [axons(6.0, 0.01, seed=5, orientation=90.0)])
    sl = m[:, :, 0] > 0
    assert 0.9 * math.pi * 36 < sl.sum() < 1.1 * math.pi * 36
    g = (np.arange(n) + 0.5) * fov / n
    X, Yg = np.meshgrid(g, g, indexing="ij")
    x0, y0 = X[sl].mean(), Yg[sl].mean()
    a2 = sl.sum() * (fov / n) ** 2 / math.pi
    dx, dy = X - x0, Yg - y0
    r2 = dx * dx + dy * dy
    k = 2 * math.pi * (1 - Y) * DCHI
    want = k * a2 / r2 * (2 * dx * dx / r2 - 1)              # B0 projection along x
    ring = (r2 > 2.0 * a2) & (r2 < 36.0 * a2)
    err = np.abs(f[:, :, 0][ring] - want[ring]).max()
    assert err < 0.05 * k, err
    np.testing.assert_allclose(f[:, :, 0][sl], np.float32(k * (0 - 1 / 3.0)), rtol=1e-5)


def test_single_sphere_closed_form(engine):
    n, fov = 96, 96.0
    m, f, _ = _gen(engine, n, fov, [W.spheres(8.0, 0.01, seed=6)])
    inside = m > 0
    g = (np.arange(n) + 0.5) * fov / n
    X, Yg, Z = np.meshgrid(g, g, g, indexing="ij")
    c = [X[inside].mean(), Yg[inside].mean(), Z[inside].mean()]
    a3 = inside.sum() * (fov / n) ** 3 * 3 / (4 * math.pi)
    dx, dy, dz = X - c[0], Yg - c[1], Z - c[2]
    r2 = dx * dx + dy * dy + dz * dz
    k = 4 * math.pi * (1 - Y) * DCHI
    want = k * a3 / r2 ** 1.5 * (dz * dz / r2 - 1 / 3.0)
    a = a3 ** (1 / 3.0)
    shell = (r2 > (1.5 * a) ** 2) & (r2 < (6 * a) ** 2)
    assert np.abs(f[shell] - want[shell]).max() < 0.05 * k
    np.testing.assert_array_equal(f[inside], 0)


def test_rejects_bad_specs(engine):
    with pytest.raises(E.EngineError, match="radius"):
        engine.generate_phantom((8, 8, 8), (8.0,) * 3, [W.cylinders(-1.0, 4.0, seed=1)])
    with pytest.raises(E.EngineError, match="radius_min"):
        engine.generate_phantom((8, 8, 8), (8.0,) * 3, [dict(W.spheres(1.0, 4.0, seed=1), radius_min_um=0.5)])


def test_int16_field_copy_needs_a_finite_fieldmap(engine):
    """The int16 copy scales by max|field|, which a NaN or an infinity leaves
    undefined: the phantom is refused (the float32 copy of the same field is
    accepted, as the reference accepts any float fieldmap)."""
    mask = np.zeros((8, 8, 8), np.uint8)
    field = np.zeros((8, 8, 8), np.float32)
    field[3, 4, 5] = np.inf
    with pytest.raises(E.EngineError, match="finite"):
        engine.phantom(mask, field, [8e-6] * 3, formats=(E.SW_FIELD_F32, E.SW_FIELD_Q16))
    ph = engine.phantom(mask, field, [8e-6] * 3, formats=(E.SW_FIELD_F32,))
    ph.close()
