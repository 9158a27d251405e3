"""Synthetic simulation cases for the tests.

Phantoms are small (<= 32^3) and built in numpy from the closed-form vessel
fields the reference's generator uses: infinite cylinder
2*pi*(1-Y)*dChi * (a^2/r^2) * cos(2 phi) * sin^2(theta) outside and
2*pi*(1-Y)*dChi * (cos^2(theta) - 1/3) inside (src/phantom/phantom_cylinder.cpp:254-265);
sphere 4*pi*(1-Y)*dChi * a^3/r^3 * (cos^2(phi) - 1/3) outside, 0 inside
(src/phantom/phantom_sphere.cpp:185-187).  Fields are in tesla per 1 T of B0.

``to_oracle`` turns an engine ``Case`` plus a scale into the oracle's input dict,
recomputing every host transform with the oracle's own restatement and checking
that the engine's values agree bit for bit.
"""
from __future__ import annotations

import numpy as np

import sys, os
sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle import oracle as orc  # noqa: E402


def grid_centres(n, fov):
    v = fov / n
    return (np.arange(n) + 0.5) * v


def cylinder_phantom(n=24, fov_um=48.0, radius_um=4.0, n_cyl=3, Y=0.7, dchi=0.11e-6 * 0.4, seed=1,
                     random_axis=True, label=1, mask=None, field=None):
    """Mask and field_T of n_cyl infinite cylinders (labels `label`)."""
    rng = np.random.default_rng(seed)
    c = grid_centres(n, fov_um)
    X, Yg, Z = np.meshgrid(c, c, c, indexing="ij")
    P = np.stack([X, Yg, Z], -1)
    mask = np.zeros((n, n, n), np.uint8) if mask is None else mask
    field = np.zeros((n, n, n), np.float64) if field is None else field.astype(np.float64)
    b0 = np.array([0.0, 0.0, 1.0])
    k = 2 * np.pi * (1 - Y) * dchi
    for _ in range(n_cyl):
        p0 = rng.uniform(0, fov_um, 3)
        if random_axis:
            u = rng.normal(size=3)
            u /= np.linalg.norm(u)
        else:
            u = np.array([0.0, 0.0, 1.0])
        ct = float(u @ b0)
        st2 = 1 - ct * ct
        d = P - p0
        d = d - (d @ u)[..., None] * u
        r2 = np.einsum("...i,...i->...", d, d)
        inside = r2 <= radius_um ** 2
        mask[inside] = label
        bp = b0 - ct * u
        nb = np.linalg.norm(bp)
        if nb > 1e-12:
            bp /= nb
            cphi = (d @ bp) / np.sqrt(np.maximum(r2, 1e-30))
            out = k * (radius_um ** 2 / np.maximum(r2, 1e-30)) * (2 * cphi * cphi - 1) * st2
        else:
            out = np.zeros_like(r2)
        field += np.where(inside, k * (ct * ct - 1.0 / 3.0), out)
    return mask, field.astype(np.float32)


def sphere_phantom(n=24, fov_um=48.0, radius_um=3.0, n_sph=4, Y=0.7, dchi=0.11e-6 * 0.4, seed=2, label=2,
                   mask=None, field=None):
    rng = np.random.default_rng(seed)
    c = grid_centres(n, fov_um)
    X, Yg, Z = np.meshgrid(c, c, c, indexing="ij")
    P = np.stack([X, Yg, Z], -1)
    mask = np.zeros((n, n, n), np.uint8) if mask is None else mask
    field = np.zeros((n, n, n), np.float64) if field is None else field.astype(np.float64)
    k = 4 * np.pi * (1 - Y) * dchi
    for _ in range(n_sph):
        c0 = rng.uniform(0, fov_um, 3)
        d = P - c0
        r2 = np.einsum("...i,...i->...", d, d)
        inside = r2 <= radius_um ** 2
        mask[inside] = label
        cphi2 = d[..., 2] ** 2 / np.maximum(r2, 1e-30)
        field += np.where(inside, 0.0, k * radius_um ** 3 / np.maximum(r2, 1e-30) ** 1.5 * (cphi2 - 1.0 / 3.0))
    return mask, field.astype(np.float32)


def two_substrate_phantom(n=24, fov_um=48.0, seed=3):
    m, f = cylinder_phantom(n, fov_um, radius_um=5.0, n_cyl=3, seed=seed, label=1)
    return m, f


def three_substrate_phantom(n=24, fov_um=48.0, seed=4):
    m, f = cylinder_phantom(n, fov_um, radius_um=5.0, n_cyl=2, seed=seed, label=1)
    m, f = sphere_phantom(n, fov_um, radius_um=4.0, n_sph=4, seed=seed + 1, label=2, mask=m, field=f)
    return m, f


# ---------------------------------------------------------------------------
# sequences (config dicts in the reference's INI key names and units)
# ---------------------------------------------------------------------------
def base_cfg(n_sub, **kw):
    cfg = dict(
        TR=10000, TE=[5000], RF_FA=[90.0], RF_PH=[0.0], RF_T=[0], TIME_STEP=50,
        DUMMY_SCAN=0, LINEAR_PHASE_CYCLING=0.0, QUADRATIC_PHASE_CYCLING=0.0,
        B0=9.4, SEED=10, NUMBER_OF_SPINS=256, CROSS_FOV=0, RECORD_TRAJECTORY=0, MAX_ITERATIONS=1e4,
        DIFFUSIVITY=[1e-9] * n_sub, T1=[2200.0] * n_sub, T2=[41.0] * n_sub,
        P_XY=np.eye(n_sub).tolist(), WHAT_TO_SCALE=0, SCALE=[1.0],
    )
    cfg.update(kw)
    return cfg


def gre(n_sub, TE=5000, dt=50, **kw):
    # config_generator.cpp:92-117: TR = TE + dt, 90 degree excitation
    return base_cfg(n_sub, TR=TE + dt, TE=[TE], RF_FA=[90.0], RF_PH=[0.0], RF_T=[0], TIME_STEP=dt, **kw)


def se(n_sub, TE=10000, dt=50, **kw):
    # config_generator.cpp:120-144: 90/180 with phases 0/90, refocus at TE/2
    return base_cfg(n_sub, TR=TE + dt, TE=[TE], RF_FA=[90.0, 180.0], RF_PH=[0.0, 90.0],
                    RF_T=[0, TE // 2], TIME_STEP=dt, **kw)


def bssfp(n_sub, TE=2500, dt=50, dummies=3, lin=180.0, fa=16.0, **kw):
    # config_generator.cpp:147-174: TR = 2 TE, FA 16, phase cycling
    return base_cfg(n_sub, TR=2 * TE, TE=[TE], RF_FA=[fa], RF_PH=[0.0], RF_T=[0], TIME_STEP=dt,
                    DUMMY_SCAN=dummies, LINEAR_PHASE_CYCLING=lin, **kw)


def ste(n_sub, dt=50, **kw):
    # config/stimulated_echo.ini: three 90s, two echoes, dephasing events
    return base_cfg(n_sub, TR=12000, TE=[4000, 8000], RF_FA=[90.0, 90.0, 90.0], RF_PH=[0.0, 90.0, 90.0],
                    RF_T=[0, 2000, 6000], DEPHASING=[360.0, 360.0, 720.0, 360.0],
                    DEPHASING_T=[1000, 3000, 5000, 7000], TIME_STEP=dt, **kw)


def pgse(n_sub, b=1000.0, delta_ms=2.0, DELTA_ms=4.0, dt=50, axis=(1.0, 0.0, 0.0), start_ms=0.5, **kw):
    """PGSE waveform as src/dwi/pgse.cpp:67-153 writes it (one direction)."""
    gamma = 267515315.0
    d, D = delta_ms * 1e-3, DELTA_ms * 1e-3
    G = np.sqrt(b * 1e6 / (gamma * gamma * d * d * (D - d / 3.0))) * 1000.0
    u = np.asarray(axis, float)
    u = u / np.linalg.norm(u)
    start_us = int(start_ms * 1000)
    npts = int(delta_ms * 1000 / dt)
    amp = [0.0] + [G] * npts + [0.0] + [0.0] + [G] * npts + [0.0]
    t = [start_us - dt] + [start_us + i * dt for i in range(npts)] + [start_us + npts * dt + dt]
    t += [int(start_us + (DELTA_ms - delta_ms) * 1000 + npts * dt - dt)]
    t += [int(start_us + (DELTA_ms - delta_ms) * 1000 + i * dt) for i in range(npts, 2 * npts)]
    t += [int(start_us + (DELTA_ms - delta_ms) * 1000 + 2 * npts * dt + dt)]
    TE = int(2 * (start_us + delta_ms * 1000 + (DELTA_ms - delta_ms) * 1000 / 2))
    amp = np.asarray(amp)
    return base_cfg(n_sub, TR=TE + dt, TE=[TE], RF_FA=[90.0, 180.0], RF_PH=[0.0, 90.0],
                    RF_T=[0, int(start_us + delta_ms * 1000 + (DELTA_ms - delta_ms) * 1000 / 2)],
                    GRADIENT_X=(amp * u[0]).tolist(), GRADIENT_Y=(amp * u[1]).tolist(),
                    GRADIENT_Z=(amp * u[2]).tolist(), GRADIENT_T=t, TIME_STEP=dt, **kw), G


# ---------------------------------------------------------------------------
# engine Case -> oracle inputs
# ---------------------------------------------------------------------------
def to_oracle(case, scale_index=0, check_transforms=True):
    """Oracle input dict for one scale of an engine Case, checking that the
    engine's host transforms equal the oracle's restatement bit for bit."""
    d = orc.from_engine_case(case, scale_index)
    if check_transforms:
        p, prep = case.params, d["prep"]
        for name in ("TE_t", "RF_t", "dephasing_t", "gradient_t"):
            assert np.array_equal(d[name], case.tables[name]), name
        assert prep["c"] == p.c and prep["s"] == p.s
        assert prep["n_timepoints"] == p.n_timepoints
        assert prep["n_dummy_scan"] == p.n_dummy_scan
        assert np.array_equal(prep["diffusivity"], case.tables["diffusivity"])
        if case.field_T is not None:
            from spinwalk_amd import engine as E
            assert orc.tesla2deg(d["B0"], d["timestep_us"]) == E.tesla2deg(d["B0"], d["timestep_us"])
    return d


def blocky_phantom(dims, n_labels=2, seed=0, block=3):
    """Non-cubic phantom: labels constant on blocks of `block` voxels, smooth field."""
    rng = np.random.default_rng(seed)
    coarse = rng.integers(0, n_labels, size=[(d + block - 1) // block for d in dims])
    mask = np.kron(coarse, np.ones((block,) * 3, np.int64))[:dims[0], :dims[1], :dims[2]].astype(np.uint8)
    g = np.meshgrid(*[np.linspace(0, 1, d) for d in dims], indexing="ij")
    field = (3e-8 * np.sin(7 * g[0] + 2) * np.cos(5 * g[1]) + 2e-8 * g[2] + 1e-8 * (mask == 1)).astype(np.float32)
    return mask, field


def make_case(cfg, mask, field_T, fov_um, n_spins=None, xyz0=None):
    from spinwalk_amd import engine as E
    if n_spins is not None:
        cfg = dict(cfg, NUMBER_OF_SPINS=n_spins)
    fov = (np.asarray(fov_um, np.float64) * 1e-6).astype(np.float32) * np.ones(3, np.float32)
    return E.prepare_case(cfg, mask, field_T, fov, XYZ0=xyz0)


def write_config(path, cfg, phantoms, output_dir="out", seq_name="test", xyz0=None):
    """Write `cfg` (reference key names) as an INI file the host reader accepts."""
    from spinwalk_amd import host as H
    files = {"OUTPUT_DIR": output_dir}
    for i, p in enumerate(phantoms):
        files[f"PHANTOM[{i}]"] = p
    for i, p in enumerate(xyz0 or []):
        files[f"XYZ0[{i}]"] = p
    tissue = {}
    for i, v in enumerate(cfg["DIFFUSIVITY"]):
        tissue[f"DIFFUSIVITY[{i}]"] = repr(float(v))
    for i, row in enumerate(np.asarray(cfg["P_XY"], float)):
        tissue[f"P_XY[{i}]"] = " ".join(repr(float(x)) for x in row)
    for i, v in enumerate(cfg["T1"]):
        tissue[f"T1[{i}]"] = repr(float(v))
    for i, v in enumerate(cfg["T2"]):
        tissue[f"T2[{i}]"] = repr(float(v))
    scan = {k: cfg[k] for k in ("TR", "TE", "RF_FA", "RF_PH", "RF_T", "TIME_STEP", "DUMMY_SCAN",
                                "LINEAR_PHASE_CYCLING", "QUADRATIC_PHASE_CYCLING") if k in cfg}
    for k in ("DEPHASING", "DEPHASING_T", "GRADIENT_X", "GRADIENT_Y", "GRADIENT_Z", "GRADIENT_T"):
        if cfg.get(k) is not None and len(cfg.get(k)):
            scan[k] = cfg[k]
    sim = {k: cfg[k] for k in ("B0", "SEED", "NUMBER_OF_SPINS", "CROSS_FOV", "RECORD_TRAJECTORY", "MAX_ITERATIONS",
                               "WHAT_TO_SCALE") if k in cfg}
    for i, s in enumerate(cfg.get("SCALE", [1.0])):
        sim[f"SCALE[{i}]"] = repr(float(s))
    H.write_ini(path, {"GENERAL": {"SEQ_NAME": seq_name}, "FILES": files, "TISSUE_PARAMETERS": tissue,
                       "SCAN_PARAMETERS": scan, "SIMULATION_PARAMETERS": sim})
