"""The host config reader against the reference's own config files.

The INI files shipped in the reference's config/ directory are read as data
(never run) wherever the reference tree is present; the test is skipped
elsewhere.  Each file goes through the engine's reader (`host.load_config`,
C++), and through an independent Python restatement of the reference's reader
written here from src/sim/config_reader.cpp:16-193 (read), :195-324 (check),
:39-46 (timing_scale) and src/sim/simulation_parameters.cuh:227-245 (prepare).
The two must agree on every parameter, table, file path and output name, and
they must reject the same files with the same check.

Two of the reference's configs exercise its quirks:
  - grase.ini writes TE[i] and RF_FA[i] keys, which the reader never looks up,
    so the echo and RF tables stay the parent's;
  - gradient.ini replaces T1/T2/DIFFUSIVITY with one substrate but inherits
    the parent's 2x2 P_XY, so the reference's own check rejects it.
"""
import math
import os
import shutil

import numpy as np
import pytest

from spinwalk_amd import host as H

REF_CONFIG = "/root/reference/config"

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_CONFIG), reason="the reference tree is not present")


def _ini(path):
    """section -> {key: value}; ';' and '#' start comment lines; values trimmed."""
    out, sec = {}, None
    for raw in open(path, encoding="utf-8").read().splitlines():
        line = raw.strip()
        if not line or line[0] in ";#":
            continue
        if line.startswith("[") and line.endswith("]"):
            sec = line[1:-1].strip()
            out.setdefault(sec, {})
            continue
        if "=" in line and sec is not None:
            k, v = line.split("=", 1)
            out[sec][k.strip()] = v.strip()
    return out


def _f32(x):
    return float(np.float32(x))


def _vec(s, int_out):
    """str2vec: whitespace-separated values read as float, then cast."""
    vals = [np.float32(float(t)) for t in s.split()]
    return [int(v) for v in vals] if int_out else [float(v) for v in vals]


class _Reader:
    """config_reader::read, one file at a time, parent first."""

    def __init__(self, top):
        self.top = os.path.abspath(top)
        self.p = {}
        self.files = {"PHANTOM": [], "XYZ0": [], "M0": []}
        self.scales = []
        self.output_dir = ""
        self.seq_name = ""

    def read(self, path):
        ini = _ini(path)
        here = os.path.dirname(os.path.abspath(path))
        g = lambda s, k: ini.get(s, {}).get(k, "")
        parent = g("GENERAL", "PARENT_CONFIG")
        if parent:
            self.read(parent if os.path.isabs(parent) else os.path.join(here, parent))
        self.seq_name = g("GENERAL", "SEQ_NAME") or self.seq_name
        for key in self.files:
            if f"{key}[0]" in ini.get("FILES", {}):
                self.files[key] = []
            i = 0
            while g("FILES", f"{key}[{i}]"):
                v = g("FILES", f"{key}[{i}]")
                self.files[key].append(v if os.path.isabs(v) else os.path.normpath(os.path.join(here, v)))
                i += 1
        od = g("FILES", "OUTPUT_DIR")
        if od:
            self.output_dir = od if os.path.isabs(od) else os.path.normpath(
                os.path.join(os.path.dirname(self.top), od))
        for key, conv in (("TR", "f"), ("TIME_STEP", "f"), ("DUMMY_SCAN", "i"), ("LINEAR_PHASE_CYCLING", "f"),
                          ("QUADRATIC_PHASE_CYCLING", "f")):
            if g("SCAN_PARAMETERS", key):
                v = g("SCAN_PARAMETERS", key)
                self.p[key] = int(float(v)) if conv == "i" else _f32(v)
        for key, is_int in (("TE", True), ("RF_T", True), ("RF_FA", False), ("RF_PH", False),
                            ("DEPHASING_T", True), ("DEPHASING", False), ("GRADIENT_T", True),
                            ("GRADIENT_X", False), ("GRADIENT_Y", False), ("GRADIENT_Z", False)):
            if g("SCAN_PARAMETERS", key):
                self.p[key] = _vec(g("SCAN_PARAMETERS", key), is_int)
        for key, conv in (("B0", "f"), ("SEED", "i"), ("NUMBER_OF_SPINS", "d"), ("CROSS_FOV", "i"),
                          ("RECORD_TRAJECTORY", "i"), ("MAX_ITERATIONS", "d"), ("WHAT_TO_SCALE", "i")):
            if g("SIMULATION_PARAMETERS", key):
                v = g("SIMULATION_PARAMETERS", key)
                self.p[key] = {"f": _f32, "i": lambda x: int(float(x)), "d": float}[conv](v)
        if "SCALE[0]" in ini.get("SIMULATION_PARAMETERS", {}):
            self.scales = []
        i = 0
        while g("SIMULATION_PARAMETERS", f"SCALE[{i}]"):
            self.scales.append(_f32(float(g("SIMULATION_PARAMETERS", f"SCALE[{i}]"))))
            i += 1
        for key in ("DIFFUSIVITY", "T1", "T2"):
            vals, i = [], 0
            while g("TISSUE_PARAMETERS", f"{key}[{i}]"):
                vals.append(_f32(g("TISSUE_PARAMETERS", f"{key}[{i}]")))
                i += 1
            if vals:
                self.p[key] = vals
        rows, i = [], 0
        while g("TISSUE_PARAMETERS", f"P_XY[{i}]"):
            rows += [_f32(float(t)) for t in g("TISSUE_PARAMETERS", f"P_XY[{i}]").split()]
            i += 1
        if rows:
            self.p["P_XY"] = rows

    def check(self):
        """config_reader::check; returns the error text or None."""
        q = self.p
        n1 = len(q.get("T1", []))
        if not (len(q.get("RF_FA", [])) == len(q.get("RF_T", [])) == len(q.get("RF_PH", []))):
            return "RF_FA, RF_PH and RF_us must have the same number of elements"
        if len(q.get("DEPHASING_T", [])) != len(q.get("DEPHASING", [])):
            return "DEPHASING and DEPHASING_T must have the same number of elements"
        if n1 != len(q.get("T2", [])):
            return "T1 and T2 must have the same number of elements"
        if n1 != len(q.get("DIFFUSIVITY", [])):
            return "T1 and diffusivity must have the same number of elements"
        if n1 * n1 != len(q.get("P_XY", [])):
            return "T1 and P_XY must have the same number of elements"
        if not self.scales:
            self.scales = [1.0]
        return None


def _expected(top):
    r = _Reader(top)
    r.read(top)
    err = r.check()
    if err:
        return None, err
    q = r.p
    dt = int(q["TIME_STEP"])
    tr = int(q["TR"])
    e = dict(seq_name=r.seq_name, TR_us=tr, timestep_us=dt, n_timepoints=tr // dt, B0=q["B0"],
             seed=q["SEED"], n_spins=int(q["NUMBER_OF_SPINS"]), max_iterations=int(q["MAX_ITERATIONS"]),
             cross_fov=q["CROSS_FOV"], record_trajectory=q["RECORD_TRAJECTORY"], what_to_scale=q["WHAT_TO_SCALE"],
             linear_phase_cycling=q["LINEAR_PHASE_CYCLING"], quadratic_phase_cycling=q["QUADRATIC_PHASE_CYCLING"],
             scales=r.scales)
    for key, name in (("TE", "TE_t"), ("RF_T", "RF_t"), ("DEPHASING_T", "dephasing_t"), ("GRADIENT_T", "gradient_t")):
        e[name] = [int(math.trunc(v / dt)) for v in q.get(key, [])]          # C++ integer division
    e["RF_FA_deg"], e["RF_PH_deg"] = q["RF_FA"], q["RF_PH"]
    e["dephasing_deg"] = q.get("DEPHASING", [])
    e["gx"], e["gy"], e["gz"] = q.get("GRADIENT_X", []), q.get("GRADIENT_Y", []), q.get("GRADIENT_Z", [])
    e["T1_ms"], e["T2_ms"], e["pXY"] = q["T1"], q["T2"], q["P_XY"]
    e["diffusivity"] = [1e-3 * math.sqrt(2.0 * d * dt) for d in q["DIFFUSIVITY"]]
    nd = q["DUMMY_SCAN"]
    if nd < 0:                                   # simulation_parameters.cuh:239-241
        nd = int(5.0 * float(np.float32(q["T1"][0])) / float(np.float32(tr * 1e-3)))
    e["n_dummy_scan"] = nd
    # cosf of the double angle narrowed to float (simulation_parameters.cuh:229)
    e["c"] = float(np.cos(np.float32(float(np.float32(q["RF_FA"][0])) * math.pi / 180.0)))
    files = dict(r.files)
    files["XYZ0"] = (files["XYZ0"] + [""] * len(files["PHANTOM"]))[:len(files["PHANTOM"])]
    files["M0"] = (files["M0"] + [""] * len(files["PHANTOM"]))[:len(files["PHANTOM"])]
    e["files"] = files
    e["outputs"] = [os.path.join(r.output_dir, r.seq_name + "_" + os.path.splitext(os.path.basename(p))[0] + ".h5")
                    for p in files["PHANTOM"]]
    return e, None


@pytest.fixture(scope="module")
def tree(tmp_path_factory):
    """The reference's config directory at <root>/r/config, with every file its
    configs name (../../phantoms/...) created as a small phantom."""
    root = tmp_path_factory.mktemp("refcfg")
    cfg = root / "r" / "config"
    cfg.mkdir(parents=True)
    for f in sorted(os.listdir(REF_CONFIG)):
        if f.endswith(".ini"):
            shutil.copy(os.path.join(REF_CONFIG, f), cfg / f)
    for f in os.listdir(cfg):
        for sec in _ini(cfg / f).values():
            for k, v in sec.items():
                if k.split("[")[0] in ("PHANTOM", "XYZ0", "M0") and v:
                    p = os.path.normpath(os.path.join(cfg, v))
                    os.makedirs(os.path.dirname(p), exist_ok=True)
                    if not os.path.exists(p):
                        H.write_phantom(p, np.zeros((4, 4, 4), np.uint8), np.zeros((4, 4, 4), np.float32),
                                        [1e-5] * 3)
    return cfg


CONFIGS = sorted(f for f in os.listdir(REF_CONFIG) if f.endswith(".ini")) if os.path.isdir(REF_CONFIG) else []


@pytest.mark.parametrize("name", CONFIGS)
def test_reference_config_reads_as_the_reference_reads_it(tree, name):
    want, err = _expected(str(tree / name))
    if err is not None:
        with pytest.raises(H.ConfigError, match=err.split(" must")[0]):
            H.load_config(tree / name)
        return
    got = H.load_config(tree / name)
    for k, v in want.items():
        g = got[k]
        if k == "seed" and v == 0:               # SEED = 0 draws a random seed
            assert g != 0
        elif k in ("files", "outputs", "seq_name", "scales"):
            assert g == v, (k, g, v)
        elif isinstance(v, list):
            np.testing.assert_array_equal(np.asarray(g, np.float64), np.asarray(v, np.float64), err_msg=k)
        elif k == "c":                           # one float ulp: numpy's cos against the C library's cosf
            assert abs(g - v) <= np.spacing(np.float32(abs(v))), (k, g, v)
        elif k in ("B0", "linear_phase_cycling", "quadratic_phase_cycling"):
            assert float(np.float32(g)) == float(np.float32(v)), (k, g, v)
        else:
            assert g == v, (k, g, v)


def test_reference_quirks_in_its_own_configs(tree):
    # grase.ini: TE[i] / RF_FA[i] are never looked up; the parent's tables stay
    g = H.load_config(tree / "grase.ini")
    d = H.load_config(tree / "config_default.ini")
    np.testing.assert_array_equal(g["TE_t"] * g["timestep_us"], d["TE_t"] * d["timestep_us"])
    np.testing.assert_array_equal(g["RF_FA_deg"], d["RF_FA_deg"])
    assert g["TR_us"] == 200000
    # gradient.ini: one substrate against the parent's 2x2 P_XY fails the check
    with pytest.raises(H.ConfigError, match="P_XY"):
        H.load_config(tree / "gradient.ini")
    # ssfp.ini: DUMMY_SCAN = -1 gives 5 T1 / TR dummy scans; PHASE_CYCLING is ignored
    s = H.load_config(tree / "ssfp.ini")
    assert s["n_dummy_scan"] == 1100 and s["linear_phase_cycling"] == 0.0
