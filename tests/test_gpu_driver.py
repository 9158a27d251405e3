"""The drop-in path end to end: INI config + HDF5 phantoms -> the host driver
(monte_carlo::run, src/sim/monte_carlo.cu:199-355) on the GPU -> HDF5 results in
the reference's layout (:168-197), checked against the oracle.

With the reference CPU path's own draws replayed, the written trajectories and
tissue labels equal the oracle's bit for bit and M agrees within 1e-5 relative;
across phantoms the output buffers carry the previous phantom's values for
abandoned spins, as the reference's reused host vectors do."""
import os
import subprocess

import numpy as np
import pytest

from tests import cases
from oracle import oracle as orc
from spinwalk_amd import engine as E
from spinwalk_amd import host as H

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _check(got_M, got_T, got_X, o):
    np.testing.assert_array_equal(got_X, o["XYZ1"])
    np.testing.assert_array_equal(got_T, o["T"])
    mag = np.linalg.norm(o["M1"], axis=-1, keepdims=True)
    assert (np.abs(got_M - o["M1"]) <= 1e-5 * mag + 1e-7).all()


@pytest.mark.parametrize("record,chunk", [(0, 0), (1, 0), (1, 37)])
def test_run_config_replay_matches_oracle(engine, tmp_path, record, chunk):
    fov_um = 24.0
    m1, f1 = cases.cylinder_phantom(n=12, fov_um=fov_um, radius_um=2.0, n_cyl=3, seed=21)
    m2, f2 = cases.three_substrate_phantom(n=12, fov_um=fov_um, seed=22)
    fov = np.full(3, np.float32(fov_um * 1e-6))
    H.write_phantom(tmp_path / "ph1.h5", m1, f1, fov)
    H.write_phantom(tmp_path / "ph2.h5", m2, f2, fov)
    # stuck spins (MAX_ITERATIONS=2 with impermeable pockets) exercise the stale outputs
    cfg = cases.se(3, TE=3000, dt=50, NUMBER_OF_SPINS=200, MAX_ITERATIONS=2, DIFFUSIVITY=[2e-7, 2e-7, 2e-7],
                   P_XY=[[1, 0, 0.5], [0, 1, 0.5], [0.5, 0.5, 1]], WHAT_TO_SCALE=0, SCALE=[0.5, 1.0],
                   RECORD_TRAJECTORY=record)
    cfg["TE"] = [1500, 3000]
    traj = 3050 // 50 if record else 1
    ini = tmp_path / "se.ini"
    cases.write_config(ini, cfg, ["ph1.h5", "ph2.h5"], output_dir="out", seq_name="se")
    # with trajectories the spins of a scale stream through the device in chunks
    # (a chunk of 37 spins leaves a ragged last chunk)
    H.run_config(engine, ini, rng=E.SW_RNG_REPLAY, trajectory_chunk_spins=chunk)

    prev = None
    for k, (m, f) in enumerate(((m1, f1), (m2, f2))):
        out = tmp_path / "out" / f"se_ph{k + 1}.h5"
        M = H.h5_read(out, "M")
        T = H.h5_read(out, "T", np.uint8)
        X = H.h5_read(out, "XYZ")
        assert M.shape == (2, 200, 2, 3) and T.shape == (2, 200, 2, 1) and X.shape == (2, 200, traj, 3)
        np.testing.assert_array_equal(H.h5_read(out, "scales").ravel(), np.float32([0.5, 1.0]))
        np.testing.assert_array_equal(H.h5_read(out, "TE").ravel(), np.float32([1500e-6, 3000e-6]))
        case = cases.make_case(cfg, m, f, fov_um)
        lost = 0
        for si in range(2):
            d = cases.to_oracle(case, si)
            if prev is not None:
                d["M1_init"], d["T_init"], d["XYZ1_init"] = prev[0][si], prev[1][si], prev[2][si]
            o = orc.sim(d)
            lost += o["n_lost"]
            _check(M[si], T[si][..., 0], X[si], o)
        assert lost > 0
        prev = (M, T[..., 0], X)


def test_cli_runs_a_config(tmp_path):
    m, f = cases.cylinder_phantom(n=10, fov_um=20.0, radius_um=2.0, n_cyl=2, seed=5)
    H.write_phantom(tmp_path / "p.h5", m, f, np.full(3, np.float32(20e-6)))
    cfg = cases.gre(2, TE=2000, dt=50, NUMBER_OF_SPINS=500)
    cases.write_config(tmp_path / "g.ini", cfg, ["p.h5"], output_dir="res", seq_name="gre")
    exe = os.path.join(ROOT, "spinwalk_amd", "bin", "spinwalk")
    r = subprocess.run([exe, "sim", "-c", str(tmp_path / "g.ini")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    M = H.h5_read(tmp_path / "res" / "gre_p.h5", "M")
    assert M.shape == (1, 500, 1, 3) and np.isfinite(M).all()
    # the mixed-precision options: same trajectories and labels, M within 1e-4 of the default run
    r = subprocess.run([exe, "sim", "-c", str(tmp_path / "g.ini"), "--precision", "kahan", "--field", "q16"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    M2 = H.h5_read(tmp_path / "res" / "gre_p.h5", "M")
    np.testing.assert_allclose(M2, M, rtol=0, atol=1e-4)
    # -p (the CPU path) and unknown option values are refused
    r = subprocess.run([exe, "sim", "-c", str(tmp_path / "g.ini"), "-p"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    r = subprocess.run([exe, "sim", "-c", str(tmp_path / "g.ini"), "--field", "f8"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "unknown field format" in r.stderr


def test_generated_phantoms_and_config_run_end_to_end(engine, tmp_path):
    """The reference's own simulation test (tests/test_sim.cu:21-43) on this engine:
    two phantoms of parallel cylinders with random radii up to 20 um (the
    reference's negative radius), 5 % volume fraction, FoV 600 um on 300^3 voxels,
    dChi 0.11e-6, Y 0.85 and 0.78, seed 45, along B0; a GRE config from the
    `config` generator (TE 20 ms, dt 25 us) with NUMBER_OF_SPINS = 1000; run
    through sim::handler::execute.

    Beyond running, the closed form of this geometry: cylinders along B0 have no
    field outside and a constant field inside (src/phantom/phantom_cylinder.cpp:
    262-265), and the default substrates are impermeable.  So every spin keeps
    |Mxy| = exp(-TE/T2), outside spins keep the excitation's angle, and inside
    spins are all turned by the same angle: the inside field's phase per step
    summed over the 801 steps up to the echo (field added before the echo is
    recorded, src/sim/kernels.cu:171-215), which grows with 1 - Y."""
    import ctypes as C
    names, f_in = [], []
    for k, Y in enumerate((0.85, 0.78)):
        spec = dict(kind=E.SW_SHAPE_AXONS, label=1, radius_um=-20.0, volume_fraction=5.0, random_orientation=0,
                    orientation_deg=0.0, dchi=0.11e-6, Y=Y, seed=45, trunc_radii=0.0, radius_min_um=0.0)
        ph, _ = engine.generate_phantom((300, 300, 300), (600.0,) * 3, [spec])
        m, f = ph.download(field=True)
        ph.close()
        assert abs(100.0 * np.count_nonzero(m) / m.size - 5.0) < 2.0     # tests/test_phantom.cpp:15-22
        inside = np.unique(f[m == 1])
        assert inside.size == 1 and (f[m == 0] == 0).all()
        f_in.append(inside[0])
        names.append(f"phantom{k + 1}.h5")
        H.write_phantom(tmp_path / names[-1], m, f, np.full(3, np.float32(600e-6)))
    H.generate_config("GRE", 20000, 25, names, tmp_path / "gre.ini")
    H.write_ini(tmp_path / "gre_sim.ini", {"GENERAL": {"SEQ_NAME": "gre", "PARENT_CONFIG": "gre.ini"},
                                           "SIMULATION_PARAMETERS": {"NUMBER_OF_SPINS": 1000}})
    files = (C.c_char_p * 1)(str(tmp_path / "gre_sim.ini").encode())
    L = H.load()
    assert L.sw_sim_execute(0, 0, files, 1) == E.SW_OK, L.sw_host_last_error().decode()
    c = H.load_config(tmp_path / "gre_sim.ini")
    assert c["n_timepoints"] == 801 and list(c["TE_t"]) == [800] and list(c["pXY"]) == [1, 0, 0, 1]
    e2 = np.exp(-20.0 / float(c["T2_ms"][0]))
    step_deg = [float(np.float32(x) * np.float32(E.tesla2deg(c["B0"], 25))) for x in f_in]
    for out, sd in zip(c["outputs"], step_deg):
        M = H.h5_read(out, "M")
        T = H.h5_read(out, "T", np.uint8)
        assert M.shape == (1, 1000, 1, 3) and T.shape == (1, 1000, 1, 1)
        np.testing.assert_array_equal(H.h5_read(out, "TE").ravel(), np.float32([20000e-6]))
        mxy = M[0, :, 0, :2].astype(np.float64)
        np.testing.assert_allclose(np.linalg.norm(mxy, axis=-1), e2, rtol=2e-5)
        z = mxy[:, 0] + 1j * mxy[:, 1]
        lab = T[0, :, 0, 0]
        assert (lab == 0).sum() > 800 and (lab == 1).sum() > 10
        z0, z1 = z[lab == 0], z[lab == 1]
        for zz in (z0, z1):                                 # one angle per label, to float rounding
            assert np.degrees(np.abs(np.angle(zz / zz.mean())).max()) < 1e-3
        d = np.degrees(np.angle(z1.mean() / z0.mean())) % 360.0
        want = (801 * sd) % 360.0
        assert min(abs(d - want), abs((360.0 - want) - d)) < 0.05, (d, want)      # either sense of rotation
    assert step_deg[1] / step_deg[0] == pytest.approx((1 - 0.78) / (1 - 0.85), rel=1e-5)


def test_run_config_reads_xyz0_file_and_keeps_the_fov_edge(engine, tmp_path):
    """Initial positions from an HDF5 `XYZ` dataset (monte_carlo.cu:127-139).  The
    range check lets a coordinate equal the FoV: on the first axis such a spin's
    start index lies beyond the volume and the spin is abandoned (its outputs keep
    the buffer's zeros); on the last axis the index runs into the next row, and the
    spin walks from that voxel, as the reference's sub2ind makes it."""
    fov_um, n = 24.0, 64
    m, f = cases.cylinder_phantom(n=12, fov_um=fov_um, radius_um=3.0, n_cyl=3, seed=31)
    fov = np.full(3, np.float32(fov_um * 1e-6))
    H.write_phantom(tmp_path / "ph.h5", m, f, fov)
    rng = np.random.default_rng(7)
    xyz0 = (rng.uniform(0.0, 1.0, (n, 3)) * fov).astype(np.float32)
    xyz0[0] = (fov[0], fov[1] / 2, fov[2] / 2)          # x == fov: abandoned
    xyz0[1] = (fov[0] / 3, fov[1] / 3, fov[2])          # z == fov: walks from the next row's voxel
    xyz0[2] = (0.0, 0.0, 0.0)
    H.h5_write(tmp_path / "xyz.h5", "XYZ", xyz0)
    cfg = cases.gre(2, TE=2000, dt=50, NUMBER_OF_SPINS=n, DIFFUSIVITY=[1e-9, 1e-9], P_XY=[[1, 0.5], [0.5, 1]])
    cases.write_config(tmp_path / "g.ini", cfg, ["ph.h5"], output_dir="out", seq_name="gre", xyz0=["xyz.h5"])
    H.run_config(engine, tmp_path / "g.ini", rng=E.SW_RNG_REPLAY)
    M = H.h5_read(tmp_path / "out" / "gre_ph.h5", "M")
    T = H.h5_read(tmp_path / "out" / "gre_ph.h5", "T", np.uint8)
    X = H.h5_read(tmp_path / "out" / "gre_ph.h5", "XYZ")
    case = cases.make_case(cfg, m, f, fov_um, xyz0=xyz0)
    o = orc.sim(cases.to_oracle(case, 0))
    _check(M[0], T[0][..., 0], X[0], o)
    assert o["lost"][0] == 1 and o["lost"][1:].sum() == 0
    assert (M[0, 0] == 0).all() and (T[0, 0] == 0).all()
    np.testing.assert_array_equal(X[0, 0, 0], xyz0[0])
    assert np.abs(M[0, 1:, 0, :2]).max() > 0.5


@pytest.mark.parametrize("where", ["above", "below"])
def test_run_config_rejects_xyz0_outside_the_fov(engine, tmp_path, where):
    fov_um, n = 24.0, 16
    m, f = cases.cylinder_phantom(n=12, fov_um=fov_um, radius_um=3.0, n_cyl=2, seed=32)
    fov = np.full(3, np.float32(fov_um * 1e-6))
    H.write_phantom(tmp_path / "ph.h5", m, f, fov)
    xyz0 = np.full((n, 3), fov[0] / 2, np.float32)
    xyz0[5, 1] = np.nextafter(fov[1], np.float32(1)) if where == "above" else np.float32(-1e-12)
    H.h5_write(tmp_path / "xyz.h5", "XYZ", xyz0)
    cfg = cases.gre(2, TE=2000, dt=50, NUMBER_OF_SPINS=n)
    cases.write_config(tmp_path / "g.ini", cfg, ["ph.h5"], output_dir="out", seq_name="gre", xyz0=["xyz.h5"])
    with pytest.raises(E.EngineError, match="Initial positions are outside the FoV"):
        H.run_config(engine, tmp_path / "g.ini")
    assert not (tmp_path / "out" / "gre_ph.h5").exists()


def test_run_config_rejects_a_mask_with_more_labels_than_substrates(engine, tmp_path):
    """read_phantom's check (monte_carlo.cu:113-118): max(mask) + 1 substrates
    must be described by the config."""
    fov_um = 24.0
    m, f = cases.three_substrate_phantom(n=12, fov_um=fov_um, seed=33)
    assert m.max() == 2
    H.write_phantom(tmp_path / "ph.h5", m, f, np.full(3, np.float32(fov_um * 1e-6)))
    cfg = cases.gre(2, TE=2000, dt=50, NUMBER_OF_SPINS=16)
    cases.write_config(tmp_path / "g.ini", cfg, ["ph.h5"], output_dir="out", seq_name="gre")
    with pytest.raises(E.EngineError, match="The number of substrate types in the mask does not match the number "
                                            "of substrate types in the config file: 3 vs 2"):
        H.run_config(engine, tmp_path / "g.ini")
    # the same phantom with a third substrate in the config runs
    cfg3 = cases.gre(3, TE=2000, dt=50, NUMBER_OF_SPINS=16)
    cases.write_config(tmp_path / "g3.ini", cfg3, ["ph.h5"], output_dir="out", seq_name="gre3")
    H.run_config(engine, tmp_path / "g3.ini")
    assert H.h5_read(tmp_path / "out" / "gre3_ph.h5", "M").shape == (1, 16, 1, 3)
