"""The C++ config reader and HDF5 layer of the host library (CPU only).

Expected values are written out from the reference's documented semantics
(config/config_default.ini:1-127; src/sim/config_reader.cpp:61-324): parent
inheritance with child overrides, KEY[0] in a child replacing an inherited list,
times parsed as floats then truncated and divided by TIME_STEP, P_XY rows
concatenated, paths resolved against the file that names them except OUTPUT_DIR
(against the top-level config), and the consistency checks."""
import math
import os

import numpy as np
import pytest

from spinwalk_amd import host as H


def _phantom(path):
    H.write_phantom(path, np.zeros((4, 4, 4), np.uint8), np.zeros((4, 4, 4), np.float32), [1e-5] * 3)


PARENT = """
; parent config (comment line)
[GENERAL]
PARENT_CONFIG =
SEQ_NAME      = noname

[FILES]
OUTPUT_DIR   = out
PHANTOM[0]  = ph/a.h5
PHANTOM[1]  = ph/b.h5
XYZ0[0]   =
M0[0]   =

[TISSUE_PARAMETERS]
DIFFUSIVITY[0] = 1.0e-9
DIFFUSIVITY[1] = 2e-9
P_XY[0] = 1.0 0.25
P_XY[1] = 0.5 1.0
T1[0] = 2200
T1[1] = 1500
T2[0] = 41
T2[1] = 30.5

[SCAN_PARAMETERS]
TR = 10e3
TE = 5e3 6e3 7e3
RF_FA = 15.0 0.0 0.0
RF_PH = 0.0 0.0 0.0
RF_T = 0 100e3 200e3
DEPHASING =
DEPHASING_T =
GRADIENT_X =
GRADIENT_Y =
GRADIENT_Z =
GRADIENT_T =
TIME_STEP  = 50
DUMMY_SCAN  = 0
LINEAR_PHASE_CYCLING = 0
QUADRATIC_PHASE_CYCLING = 0

[SIMULATION_PARAMETERS]
B0   = 9.4
SEED = 0
NUMBER_OF_SPINS = 1e5
CROSS_FOV = 0
RECORD_TRAJECTORY = 0
MAX_ITERATIONS = 1e4
WHAT_TO_SCALE = 0
SCALE[0] = 0.0125
SCALE[1] = 0.0147
"""

CHILD = """[GENERAL]
PARENT_CONFIG = base/parent.ini
SEQ_NAME     = se

[FILES]
PHANTOM[0]  = c.h5

[SCAN_PARAMETERS]
TR = 40000
TE = 20000

RF_FA = 90.0 180.0
RF_PH = 0.0 90
RF_T = 0 10000

[SIMULATION_PARAMETERS]
; use 0 for random seed generation
SEED = 10
SCALE[0] = 1.0
"""


@pytest.fixture
def tree(tmp_path):
    (tmp_path / "base" / "ph").mkdir(parents=True)
    _phantom(tmp_path / "base" / "ph" / "a.h5")
    _phantom(tmp_path / "base" / "ph" / "b.h5")
    _phantom(tmp_path / "c.h5")
    (tmp_path / "base" / "parent.ini").write_text(PARENT)
    (tmp_path / "child.ini").write_text(CHILD)
    return tmp_path


def test_parent_config_alone(tree):
    c = H.load_config(tree / "base" / "parent.ini")
    assert c["TR_us"] == 10000 and c["timestep_us"] == 50
    assert c["n_timepoints"] == 10000 // 50
    np.testing.assert_array_equal(c["TE_t"], [100, 120, 140])
    np.testing.assert_array_equal(c["RF_t"], [0, 2000, 4000])
    assert c["n_spins"] == 100000 and c["max_iterations"] == 10000
    assert c["cross_fov"] == 0 and c["what_to_scale"] == 0
    np.testing.assert_allclose(c["scales"], [0.0125, 0.0147], rtol=1e-7)
    np.testing.assert_array_equal(c["pXY"], np.float32([1.0, 0.25, 0.5, 1.0]))
    np.testing.assert_array_equal(c["T2_ms"], np.float32([41, 30.5]))
    # diffusivity becomes the per-axis step in metres: sqrt(2 D dt)
    np.testing.assert_allclose(c["diffusivity"], np.sqrt(2 * np.float32([1e-9, 2e-9]).astype(float) * 50e-6),
                               rtol=1e-12)
    assert abs(c["c"] - math.cos(math.radians(15))) < 1e-7
    assert c["seed"] != 0                    # SEED = 0 draws a random seed
    assert c["files"]["PHANTOM"] == [str(tree / "base" / "ph" / "a.h5"), str(tree / "base" / "ph" / "b.h5")]
    assert c["files"]["XYZ0"] == ["", ""] and c["files"]["M0"] == ["", ""]
    assert c["outputs"] == [str(tree / "base" / "out" / "noname_a.h5"), str(tree / "base" / "out" / "noname_b.h5")]


def test_child_overrides_and_inherits(tree):
    c = H.load_config(tree / "child.ini")
    assert c["seq_name"] == "se" and c["seed"] == 10
    np.testing.assert_array_equal(c["TE_t"], [400])
    np.testing.assert_array_equal(c["RF_t"], [0, 200])
    np.testing.assert_array_equal(c["RF_FA_deg"], np.float32([90, 180]))
    np.testing.assert_array_equal(c["RF_PH_deg"], np.float32([0, 90]))
    assert c["n_timepoints"] == 800
    # inherited tissue, replaced lists
    np.testing.assert_array_equal(c["T1_ms"], np.float32([2200, 1500]))
    assert c["scales"] == [1.0]
    assert c["files"]["PHANTOM"] == [str(tree / "c.h5")]
    # OUTPUT_DIR resolves against the top-level config, not the parent that sets it
    assert c["outputs"] == [str(tree / "out" / "se_c.h5")]
    assert os.path.isdir(tree / "out")


def test_value_comments_and_crlf(tmp_path):
    _phantom(tmp_path / "p.h5")
    txt = CHILD.replace("PARENT_CONFIG = base/parent.ini", "PARENT_CONFIG = parent.ini")
    txt = txt.replace("TE = 20000", "TE = 20e3 ; an inline note is part of the value").replace("\n", "\r\n")
    (tmp_path / "parent.ini").write_text(PARENT.replace("ph/a.h5", "p.h5").replace("PHANTOM[1]  = ph/b.h5", ""))
    (tmp_path / "c.h5").write_bytes((tmp_path / "p.h5").read_bytes())
    (tmp_path / "child.ini").write_bytes(b"\xef\xbb\xbf" + txt.encode())
    c = H.load_config(tmp_path / "child.ini")
    np.testing.assert_array_equal(c["TE_t"], [400])


@pytest.mark.parametrize("patch,msg", [
    (("RF_PH = 0.0 90", "RF_PH = 0.0"), "same number"),
    (("TE = 20000", "TE = 20000 10000"), "TE must"),
    (("RF_T = 0 10000", "RF_T = 50 10000"), "RF times"),
    (("SEED = 10", "SEED = 10\nWHAT_TO_SCALE = 3"), "WHAT_TO_SCALE"),
    (("PHANTOM[0]  = c.h5", "PHANTOM[0]  = missing.h5"), "does not exist"),
    (("SEED = 10", "SEED = ten"), "cannot parse"),
])
def test_check_rejects(tree, patch, msg):
    (tree / "bad.ini").write_text(CHILD.replace(*patch))
    with pytest.raises(H.ConfigError, match=msg):
        H.load_config(tree / "bad.ini")


def test_pxy_size_must_match(tree):
    (tree / "bad.ini").write_text(CHILD + "\n[TISSUE_PARAMETERS]\nP_XY[0] = 1 0 0\n")
    with pytest.raises(H.ConfigError, match="P_XY"):
        H.load_config(tree / "bad.ini")


def test_dummy_scans_auto(tree):
    (tree / "b.ini").write_text(CHILD.replace("RF_T = 0 10000", "RF_T = 0 10000\nDUMMY_SCAN = -1"))
    c = H.load_config(tree / "b.ini")
    assert c["n_dummy_scan"] == int(5.0 * 2200 / (40000 * 1e-3))


def test_hdf5_roundtrip(tmp_path):
    f = tmp_path / "x.h5"
    a = np.arange(24, dtype=np.float32).reshape(2, 3, 4)
    m = (np.arange(60) % 7).astype(np.uint8).reshape(3, 4, 5)
    H.h5_write(f, "fieldmap", a)
    H.h5_write(f, "mask", m)
    H.h5_write(f, "fieldmap", a * 2)            # replaces the dataset
    np.testing.assert_array_equal(H.h5_read(f, "fieldmap"), a * 2)
    np.testing.assert_array_equal(H.h5_read(f, "mask", np.uint8), m)
    # float data read back converted to double
    np.testing.assert_array_equal(H.h5_read(f, "fieldmap", np.float64), (a * 2).astype(np.float64))
    with pytest.raises(OSError):
        H.h5_read(f, "nope")
