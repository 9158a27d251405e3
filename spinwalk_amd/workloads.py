"""The benchmark workloads: synthetic phantoms and sequences of the five
configurations the headline metric is quoted on, in the reference's INI key
names and units (config/config_default.ini:1-127).

Sequences follow the reference's config generator defaults
(src/config/config_generator.cpp:92-174): GRE TR = TE + dt with a 90 degree
excitation; SE 90/180 with phases 0/90 and the refocusing pulse at TE/2; bSSFP
TR = 2 TE with FA 16.
"""
from __future__ import annotations

import math

import numpy as np

from . import engine as E

DCHI = 0.11e-6 * 0.4          # susceptibility difference x hematocrit
Y_OXY = 0.7


def cylinders(radius_um, vf, seed, label=1, random_orientation=1, orientation_deg=90.0):
    return dict(kind=E.SW_SHAPE_CYLINDER, label=label, radius_um=radius_um, volume_fraction=vf,
                random_orientation=random_orientation, orientation_deg=orientation_deg, dchi=DCHI, Y=Y_OXY,
                seed=seed, trunc_radii=0.0)


def spheres(radius_um, vf, seed, label=2):
    return dict(kind=E.SW_SHAPE_SPHERE, label=label, radius_um=radius_um, volume_fraction=vf,
                random_orientation=0, orientation_deg=0.0, dchi=DCHI, Y=Y_OXY, seed=seed, trunc_radii=10.0)


def _base(n_sub, **kw):
    cfg = dict(B0=9.4, SEED=10, CROSS_FOV=0, RECORD_TRAJECTORY=0, MAX_ITERATIONS=1e4, DUMMY_SCAN=0,
               LINEAR_PHASE_CYCLING=0.0, QUADRATIC_PHASE_CYCLING=0.0, WHAT_TO_SCALE=0, SCALE=[1.0],
               DIFFUSIVITY=[1e-9] * n_sub, P_XY=np.eye(n_sub).tolist())
    cfg.update(kw)
    return cfg


def c1():
    """100^3 over 100 um, random cylinders r 5 um at 4 %, GRE TE 20 ms dt 50 us, 1e5 spins."""
    phantom = dict(dims=(100, 100, 100), fov_um=(100.0, 100.0, 100.0), specs=[cylinders(5.0, 4.0, seed=1)])
    cfg = _base(2, TR=20050, TE=[20000], RF_FA=[90.0], RF_PH=[0.0], RF_T=[0], TIME_STEP=50,
                NUMBER_OF_SPINS=100000, T1=[2200.0, 2200.0], T2=[40.0, 40.0])
    return dict(name="c1_gre_100cube_1e5", phantom=phantom, cfg=cfg)


def c2():
    """500^3 over 500 um, random cylinders r 5 um at 4 % (label 1) + spheres r 3 um
    at 2 % (label 2), P_XY 0.5 between substrates, T1/T2 2000/40 (tissue) and
    1600/20 (inclusions), SE 90/180 TE 40 ms dt 20 us, 1e6 spins."""
    phantom = dict(dims=(500, 500, 500), fov_um=(500.0, 500.0, 500.0),
                   specs=[cylinders(5.0, 4.0, seed=1), spheres(3.0, 2.0, seed=2)])
    pxy = [[1.0, 0.5, 0.5], [0.5, 1.0, 0.5], [0.5, 0.5, 1.0]]
    cfg = _base(3, TR=40020, TE=[40000], RF_FA=[90.0, 180.0], RF_PH=[0.0, 90.0], RF_T=[0, 20000], TIME_STEP=20,
                NUMBER_OF_SPINS=1000000, T1=[2000.0, 1600.0, 1600.0], T2=[40.0, 20.0, 20.0], P_XY=pxy)
    return dict(name="c2_se_500cube_1e6", phantom=phantom, cfg=cfg)


def _divisor_at_most(n, limit, lowest):
    """The largest divisor of n in [lowest, limit] (lowest if none is)."""
    best = lowest
    for d in range(lowest, max(lowest, limit) + 1):
        if n % d == 0:
            best = d
    return best


def c4():
    """FoV-scaling sweep: one 256^3 phantom of random cylinders (r 8 um in a
    256 um FoV, 4 %) walked at 32 FoV scales s, log-spaced so that the scaled
    radius 8 s um spans 1-100 um (WHAT_TO_SCALE = 0), each with GRE TE 20 ms and SE
    TE 40 ms and 1e5 spins: 64 tiles.  Each tile has its own time step: the largest
    divisor of TE (so echo and refocusing fall on the grid) with
    sqrt(2 D dt) <= 0.3 voxel, kept within 1k-20k steps per TR."""
    phantom = dict(dims=(256, 256, 256), fov_um=(256.0, 256.0, 256.0), specs=[cylinders(8.0, 4.0, seed=4)])
    scales = np.logspace(math.log10(1 / 8), math.log10(100 / 8), 32)
    D = 1e-9
    tiles = []
    for seq, TE in (("gre", 20000), ("se", 40000)):
        for s in scales:
            ideal = (0.3 * s * 1e-6) ** 2 / (2 * D) * 1e6                 # us
            lo, hi = -(-TE // 20000), TE // 1000
            dt = _divisor_at_most(TE, int(min(max(ideal, lo), hi)), lo)
            kw = dict(TR=TE + dt, TE=[TE], TIME_STEP=dt, NUMBER_OF_SPINS=100000, T1=[2200.0, 2200.0],
                      T2=[41.0, 41.0], DIFFUSIVITY=[D, D], SCALE=[float(np.float32(s))], WHAT_TO_SCALE=0)
            if seq == "gre":
                cfg = _base(2, RF_FA=[90.0], RF_PH=[0.0], RF_T=[0], **kw)
            else:
                cfg = _base(2, RF_FA=[90.0, 180.0], RF_PH=[0.0, 90.0], RF_T=[0, TE // 2], **kw)
            tiles.append(dict(seq=seq, scale=float(s), dt=dt, cfg=cfg))
    return dict(name="c4_fov_sweep_256cube_32scales_gre_se_1e5", phantom=phantom, tiles=tiles)


def c3():
    """Precision form: 1000^3 voxels over 1000 um (1 GB mask) of random cylinders
    (r 5 um, 4 %), 1e7 spins, bSSFP (TR 5 ms, TE 2.5 ms, FA 16, dt 50 us) with 100
    dummy scans and 4-step linear phase cycling (LINEAR_PHASE_CYCLING 90 scaled by
    0, 1, 2, 3; WHAT_TO_SCALE = 2): 4 tiles of 1e7 spins x 10 100 steps."""
    phantom = dict(dims=(1000, 1000, 1000), fov_um=(1000.0, 1000.0, 1000.0), specs=[cylinders(5.0, 4.0, seed=3)])
    tiles = []
    for k in range(4):
        cfg = _base(2, TR=5000, TE=[2500], RF_FA=[16.0], RF_PH=[0.0], RF_T=[0], TIME_STEP=50, DUMMY_SCAN=100,
                    LINEAR_PHASE_CYCLING=90.0, WHAT_TO_SCALE=2, SCALE=[float(k)], NUMBER_OF_SPINS=10000000,
                    T1=[2200.0, 2200.0], T2=[41.0, 41.0])
        tiles.append(dict(seq="bssfp", scale=float(k), dt=50, cfg=cfg))
    return dict(name="c3_bssfp_1000cube_1e7_4cycles", phantom=phantom, tiles=tiles)


def axons(radius_max_um, radius_min_um, vf, seed, label=1, orientation_deg=0.0):
    """Packed parallel axons: random radii in [min, max] shrunk to fit, along z."""
    return dict(kind=E.SW_SHAPE_AXONS, label=label, radius_um=radius_max_um, volume_fraction=vf,
                random_orientation=0, orientation_deg=orientation_deg, dchi=DCHI, Y=Y_OXY, seed=seed,
                trunc_radii=10.0, radius_min_um=radius_min_um)


GAMMA = 267515315.0


def pgse_tables(b_first, axis, start_ms, delta_ms, DELTA_ms, dt):
    """The PGSE RF pair and gradient waveform written by the `dwi` subcommand
    (src/dwi/pgse.cpp:78-135), in the reference's INI keys."""
    d, D = delta_ms * 1e-3, DELTA_ms * 1e-3
    G = math.sqrt(b_first * 1e6 / (GAMMA * GAMMA * d * d * (D - d / 3.0))) * 1000.0
    u = np.asarray(axis, float) / np.linalg.norm(axis)
    start_us = start_ms * 1000
    n = delta_ms * 1000 // dt
    amp = [0.0] + [G] * n + [0.0] + [0.0] + [G] * n + [0.0]
    gap = (DELTA_ms - delta_ms) * 1000
    t = [start_us - dt] + [start_us + i * dt for i in range(n)] + [start_us + n * dt + dt]
    t += [start_us + gap + n * dt - dt] + [start_us + gap + i * dt for i in range(n, 2 * n)]
    t += [start_us + gap + 2 * n * dt + dt]
    amp = np.asarray(amp)
    return dict(RF_FA=[90.0, 180.0], RF_PH=[0.0, 90.0],
                RF_T=[0, start_us + delta_ms * 1000 + (DELTA_ms - delta_ms) * 1000 // 2],
                GRADIENT_X=(amp * u[0]).tolist(), GRADIENT_Y=(amp * u[1]).tolist(),
                GRADIENT_Z=(amp * u[2]).tolist(), GRADIENT_T=t)


def c5():
    """Heterogeneous batch: 64 phantoms of 128^3-512^3 voxels (1 um voxels), each
    of one kind, in turn:
      - randomly oriented cylinders, r 2-6 um, 2-6 % volume fraction;
      - spheres, r 2-5 um, 2-6 %;
      - packed parallel axons, r 0.5-3 um, 30-70 %.
    The radius 0.5-3 um and the 30-70 % fraction belong to the packed axons:
    shapes placed at random without overlap (the reference's cylinder and sphere
    generators, src/phantom/phantom_cylinder.cpp:85-130,
    src/phantom/phantom_sphere.cpp:79-113) stop fitting far below 30 % (random
    sequential placement of equal spheres jams near 38 %, of randomly oriented
    infinite cylinders at a few percent), so only packing reaches that range.
    Each phantom is walked with one of GRE (TE 20 ms), SE (TE 40 ms) or PGSE along
    x, y or z (delta 10 ms, Delta 30 ms, b = 1000, 0, 2000, 3000 s/mm^2 as gradient
    scales of the first), with 1e4-1e6 spins (log-uniform), dt 50 us.  Every tile
    of every phantom goes into one launch."""
    rng = np.random.default_rng(2024)
    phantoms, tiles = [], []
    for i in range(64):
        n = int(rng.choice([128, 192, 256, 320, 384, 448, 512]))
        kind = i % 3
        if kind == 0:
            spec = cylinders(float(rng.uniform(2.0, 6.0)), float(rng.uniform(2.0, 6.0)), seed=100 + i)
        elif kind == 1:
            spec = spheres(float(rng.uniform(2.0, 5.0)), float(rng.uniform(2.0, 6.0)), seed=100 + i, label=1)
        else:
            spec = axons(3.0, 0.5, float(rng.uniform(30.0, 70.0)), seed=100 + i)
        phantoms.append(dict(dims=(n, n, n), fov_um=(float(n),) * 3, specs=[spec]))
        spins = int(round(10 ** rng.uniform(4.0, 6.0)))
        seq = ["gre", "se", "pgse_x", "pgse_y", "pgse_z"][int(rng.integers(0, 5))]
        common = dict(NUMBER_OF_SPINS=spins, TIME_STEP=50, T1=[2200.0, 2200.0], T2=[41.0, 41.0],
                      DIFFUSIVITY=[1e-9, 1e-9], P_XY=[[1.0, 0.5], [0.5, 1.0]], SEED=1000 + i)
        if seq == "gre":
            cfgs = [_base(2, TR=20050, TE=[20000], RF_FA=[90.0], RF_PH=[0.0], RF_T=[0], **common)]
        elif seq == "se":
            cfgs = [_base(2, TR=40050, TE=[40000], RF_FA=[90.0, 180.0], RF_PH=[0.0, 90.0], RF_T=[0, 20000],
                          **common)]
        else:
            axis = {"pgse_x": (1, 0, 0), "pgse_y": (0, 1, 0), "pgse_z": (0, 0, 1)}[seq]
            tab = pgse_tables(1000.0, axis, 5, 10, 30, 50)
            TE = 2 * (5000 + 10000 + 10000)
            cfgs = [_base(2, TR=TE + 50, TE=[TE], WHAT_TO_SCALE=1, SCALE=[float(np.float32(math.sqrt(b / 1000.0)))],
                          **tab, **common) for b in (1000.0, 0.0, 2000.0, 3000.0)]
        for cfg in cfgs:
            tiles.append(dict(seq=seq, scale=float(cfg["SCALE"][0]), dt=50, cfg=cfg, phantom=i))
    return dict(name="c5_hetero_64_phantoms", phantoms=phantoms, tiles=tiles)


WORKLOADS = {"c1": c1, "c2": c2, "c3": c3, "c4": c4, "c5": c5}


def build(engine, wl, formats=(E.SW_FIELD_F32,), n_spins=None):
    """Generate the phantom on the device and prepare the case (host transforms
    through the engine).  Returns (phantom, case, vf)."""
    ph, vf = engine.generate_phantom(wl["phantom"]["dims"], wl["phantom"]["fov_um"], wl["phantom"]["specs"],
                                     formats=formats)
    cfg = dict(wl["cfg"])
    if n_spins is not None:
        cfg["NUMBER_OF_SPINS"] = n_spins
    n_labels, _ = ph.info()
    dims = wl["phantom"]["dims"]
    # the case needs the mask only for its shape on this path; the phantom stays on the device
    case = E.prepare_case(cfg, np.zeros((1, 1, 1), np.uint8), None, ph.fov_m, XYZ0=None)
    p = case.params
    for i in range(3):
        p.phantom_size[i] = dims[i]
    p.matrix_length = int(np.prod(dims))
    p.fieldmap_exist = 1
    case.mask = None
    return ph, case, vf


def build_tiles(engine, wl, formats=(E.SW_FIELD_F32,)):
    """For a workload with tiles: the device phantoms and a case per tile.
    Returns ([phantom], [case], [phantom index per tile], [vf])."""
    specs = wl["phantoms"] if "phantoms" in wl else [wl["phantom"]]
    phs, vfs = [], []
    for sp in specs:
        ph, vf = engine.generate_phantom(sp["dims"], sp["fov_um"], sp["specs"], formats=formats)
        phs.append(ph)
        vfs.append(vf)
    out, idx = [], []
    for t in wl["tiles"]:
        k = t.get("phantom", 0)
        dims = specs[k]["dims"]
        case = E.prepare_case(dict(t["cfg"]), np.zeros((1, 1, 1), np.uint8), None, phs[k].fov_m, XYZ0=None)
        p = case.params
        for i in range(3):
            p.phantom_size[i] = dims[i]
        p.matrix_length = int(np.prod(dims))
        p.fieldmap_exist = 1
        case.mask = None
        out.append(case)
        idx.append(k)
    return phs, out, idx, vfs
