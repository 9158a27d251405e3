"""ctypes loader for the CPU oracle (oracle/build/libsw_oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()`` (as
the checker) and by the ``cpu_baseline`` leg of ``bench.py`` (as the timed CPU
baseline).  The product package ``spinwalk_amd`` never imports this module.

The oracle restates the reference CPU path (src/sim/kernels.cu:72-233 and the
host transforms of src/sim/monte_carlo.cu, src/sim/config_reader.cpp,
src/sim/simulation_parameters.cuh); see oracle/sw_oracle.h for the per-function
citations and for what pins each function.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SW_ORACLE_LIB points the tests at another build of sw_oracle.cpp (a copy with a
# planted mistake, when checking that the tests catch it).
LIB_PATH = os.environ.get("SW_ORACLE_LIB") or os.path.join(_HERE, "build", "libsw_oracle.so")


class OrcParams(C.Structure):
    _fields_ = [
        ("fov", C.c_double * 3),
        ("phantom_size", C.c_uint64 * 3),
        ("matrix_length", C.c_int64),
        ("c", C.c_float), ("s", C.c_float),
        ("linear_phase_cycling", C.c_float), ("quadratic_phase_cycling", C.c_float),
        ("timestep_us", C.c_int32), ("TR_us", C.c_int32), ("n_dummy_scan", C.c_int32),
        ("n_spins", C.c_uint32), ("n_timepoints", C.c_uint32), ("n_substrate", C.c_uint32),
        ("seed", C.c_uint64), ("max_iterations", C.c_uint64),
        ("enCrossFOV", C.c_int32), ("enRecordTrajectory", C.c_int32),
    ]


_P = C.c_void_p


class OrcTables(C.Structure):
    _fields_ = [
        ("diffusivity", _P), ("T1_ms", _P), ("T2_ms", _P), ("pXY", _P),
        ("RF_FA_deg", _P), ("RF_PH_deg", _P), ("RF_t", _P), ("n_rf", C.c_uint32),
        ("TE_t", _P), ("n_te", C.c_uint32),
        ("dephasing_deg", _P), ("dephasing_t", _P), ("n_dephasing", C.c_uint32),
        ("gradientX_mTm", _P), ("gradientY_mTm", _P), ("gradientZ_mTm", _P),
        ("gradient_t", _P), ("n_gradient", C.c_uint32),
    ]


_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, no FMA contraction)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        f = C.POINTER(C.c_float)
        L.orc_sub2ind3.restype = C.c_int64
        L.orc_sub2ind3.argtypes = [C.c_int64] * 6
        L.orc_sub2ind4.restype = C.c_int64
        L.orc_sub2ind4.argtypes = [C.c_int64] * 8
        for n in ("orc_xrot_sc", "orc_yrot_sc", "orc_zrot_sc"):
            getattr(L, n).argtypes = [C.c_float, C.c_float, _P, _P]
        for n in ("orc_xrot_deg", "orc_yrot_deg", "orc_zrot_deg"):
            getattr(L, n).argtypes = [C.c_float, _P, _P]
        L.orc_xrot_withphase_sc.argtypes = [C.c_float, C.c_float, C.c_float, _P, _P]
        L.orc_xrot_withphase_deg.argtypes = [C.c_float, C.c_float, _P, _P]
        L.orc_relax.argtypes = [C.c_float, C.c_float, _P, _P]
        L.orc_dephase_relax.argtypes = [_P, _P, C.c_float, C.c_float, C.c_float, C.c_float]
        L.orc_prepare.restype = C.c_int
        L.orc_prepare.argtypes = [C.c_float, C.c_int32, C.c_int32, C.c_float, _P, C.c_uint32,
                                  _P, _P, _P, _P]
        L.orc_timing_scale.argtypes = [_P, C.c_uint32, C.c_int32]
        L.orc_tesla2deg.restype = C.c_float
        L.orc_tesla2deg.argtypes = [C.c_float, C.c_int32]
        L.orc_scale_field.argtypes = [_P, C.c_uint64, C.c_float]
        L.orc_init_positions.argtypes = [C.c_uint64, _P, C.c_uint32, _P]
        L.orc_str2vec_i32.restype = C.c_int
        L.orc_str2vec_i32.argtypes = [C.c_char_p, _P, C.c_int]
        L.orc_sim.restype = C.c_uint32
        L.orc_sim.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcTables), _P, _P, _P, _P, _P, _P,
                              _P, C.c_uint32, C.c_uint32, C.c_int, _P, _P]
        L.orc_sim_f64.restype = C.c_uint32
        L.orc_sim_f64.argtypes = [C.POINTER(OrcParams), C.POINTER(OrcTables), C.c_float, _P, _P,
                                  _P, _P, _P, _P, _P, C.c_uint32, C.c_uint32, C.c_int, _P]
        L.orc_record_draws.argtypes = [C.c_uint64, C.c_uint32, _P, C.c_uint64, _P, C.c_uint64]
        L.orc_mt19937_raw.argtypes = [C.c_uint32, _P, C.c_uint64]
        L.orc_echo_reduce.argtypes = [_P, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P]
        _lib = L
        del f
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# primitives
# ---------------------------------------------------------------------------
def _rot(name, *scalars, m0):
    m0 = np.ascontiguousarray(m0, dtype=np.float32)
    m1 = np.zeros(3, np.float32)
    getattr(lib(), name)(*[C.c_float(s) for s in scalars], _ptr(m0), _ptr(m1))
    return m1


def xrot(s, c, m0): return _rot("orc_xrot_sc", s, c, m0=m0)
def yrot(s, c, m0): return _rot("orc_yrot_sc", s, c, m0=m0)
def zrot(s, c, m0): return _rot("orc_zrot_sc", s, c, m0=m0)
def xrot_deg(t, m0): return _rot("orc_xrot_deg", t, m0=m0)
def yrot_deg(t, m0): return _rot("orc_yrot_deg", t, m0=m0)
def zrot_deg(t, m0): return _rot("orc_zrot_deg", t, m0=m0)
def xrot_withphase(s, c, ph, m0): return _rot("orc_xrot_withphase_sc", s, c, ph, m0=m0)
def xrot_withphase_deg(t, ph, m0): return _rot("orc_xrot_withphase_deg", t, ph, m0=m0)
def relax(e1, e2, m0): return _rot("orc_relax", e1, e2, m0=m0)


def dephase_relax(m0, phase, T1, T2, t):
    m0 = np.ascontiguousarray(m0, dtype=np.float32)
    m1 = np.zeros(3, np.float32)
    lib().orc_dephase_relax(_ptr(m0), _ptr(m1), C.c_float(phase), C.c_float(T1), C.c_float(T2),
                            C.c_float(t))
    return m1


def sub2ind3(x, y, z, nx, ny, nz): return lib().orc_sub2ind3(x, y, z, nx, ny, nz)
def sub2ind4(x, y, z, o, nx, ny, nz, no): return lib().orc_sub2ind4(x, y, z, o, nx, ny, nz, no)


def tesla2deg(B0, dt_us): return lib().orc_tesla2deg(C.c_float(B0), dt_us)


def scale_field(field_T, factor):
    f = np.array(field_T, dtype=np.float32, copy=True).ravel()
    lib().orc_scale_field(_ptr(f), f.size, C.c_float(factor))
    return f.reshape(np.shape(field_T))


def init_positions(seed, fov, n_spins):
    fov = np.ascontiguousarray(fov, dtype=np.float32)
    out = np.zeros((n_spins, 3), np.float32)
    lib().orc_init_positions(seed, _ptr(fov), n_spins, _ptr(out))
    return out


def str2vec_i32(s: str):
    buf = np.zeros(4096, np.int32)
    n = lib().orc_str2vec_i32(s.encode(), _ptr(buf), buf.size)
    return buf[:n].copy()


def timing_scale(v, dt):
    a = np.ascontiguousarray(v, dtype=np.int32).copy()
    lib().orc_timing_scale(_ptr(a), a.size, dt)
    return a


def prepare(RF_FA0, TR_us, dt_us, T1_0, diffusivity, n_dummy):
    d = np.ascontiguousarray(diffusivity, dtype=np.float64).copy()
    nd = C.c_int32(n_dummy)
    c, s, nt = C.c_float(), C.c_float(), C.c_uint32()
    rc = lib().orc_prepare(C.c_float(RF_FA0), TR_us, dt_us, C.c_float(T1_0), _ptr(d), d.size,
                           C.byref(nd), C.byref(c), C.byref(s), C.byref(nt))
    if rc:
        raise ValueError("prepare failed")
    return dict(c=c.value, s=s.value, n_timepoints=nt.value, diffusivity=d, n_dummy_scan=nd.value)


def record_draws(seed, spin, n_normals, n_uniforms):
    n = np.zeros(n_normals, np.float32)
    u = np.zeros(n_uniforms, np.float32)
    lib().orc_record_draws(seed, spin, _ptr(n), n_normals, _ptr(u), n_uniforms)
    return n, u


def mt19937_raw(seed, n):
    out = np.zeros(n, np.uint32)
    lib().orc_mt19937_raw(seed, _ptr(out), n)
    return out


def echo_reduce(M1, T, n_sub):
    M1 = np.ascontiguousarray(M1, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.uint8)
    n_spins, n_te = T.shape[0], T.shape[1]
    out = np.zeros((n_te, n_sub, 3), np.float64)
    lib().orc_echo_reduce(_ptr(M1), _ptr(T), n_spins, n_te, n_sub, _ptr(out))
    return out


# ---------------------------------------------------------------------------
# the walk
# ---------------------------------------------------------------------------
class _Keep:
    """Holds numpy arrays alive while ctypes structs point into them."""

    def __init__(self):
        self.items = []

    def __call__(self, a, dtype):
        if a is None or len(a) == 0:
            arr = np.zeros(1, dtype)
            self.items.append(arr)
            return arr, 0
        arr = np.ascontiguousarray(a, dtype=dtype)
        self.items.append(arr)
        return arr, arr.size


def make_structs(sim):
    """Build (OrcParams, OrcTables, keepalive) from a prepared simulation dict.

    ``sim`` is the dict produced by ``tests.simcase.prepare_case`` (times already
    converted to step indices, diffusivity already converted by prepare())."""
    keep = _Keep()
    p = OrcParams()
    for i in range(3):
        p.fov[i] = float(sim["fov"][i])
        p.phantom_size[i] = int(sim["phantom_size"][i])
    p.matrix_length = int(np.prod(sim["phantom_size"]))
    p.c, p.s = sim["c"], sim["s"]
    p.linear_phase_cycling = sim.get("linear_phase_cycling", 0.0)
    p.quadratic_phase_cycling = sim.get("quadratic_phase_cycling", 0.0)
    p.timestep_us, p.TR_us, p.n_dummy_scan = sim["timestep_us"], sim["TR_us"], sim["n_dummy_scan"]
    p.n_spins, p.n_timepoints = sim["n_spins"], sim["n_timepoints"]
    p.n_substrate = len(sim["diffusivity"])
    p.seed, p.max_iterations = sim["seed"], sim["max_iterations"]
    p.enCrossFOV, p.enRecordTrajectory = int(sim["cross_fov"]), int(sim["record_trajectory"])
    q = OrcTables()
    a, _ = keep(sim["diffusivity"], np.float64); q.diffusivity = _ptr(a)
    a, _ = keep(sim["T1_ms"], np.float32); q.T1_ms = _ptr(a)
    a, _ = keep(sim["T2_ms"], np.float32); q.T2_ms = _ptr(a)
    a, _ = keep(np.ravel(sim["pXY"]), np.float32); q.pXY = _ptr(a)
    a, _ = keep(sim["RF_FA_deg"], np.float32); q.RF_FA_deg = _ptr(a)
    a, _ = keep(sim["RF_PH_deg"], np.float32); q.RF_PH_deg = _ptr(a)
    a, n = keep(sim["RF_t"], np.int32); q.RF_t = _ptr(a); q.n_rf = n
    a, n = keep(sim["TE_t"], np.int32); q.TE_t = _ptr(a); q.n_te = n
    a, _ = keep(sim.get("dephasing_deg"), np.float32); q.dephasing_deg = _ptr(a)
    a, n = keep(sim.get("dephasing_t"), np.int32); q.dephasing_t = _ptr(a); q.n_dephasing = n
    gs = float(sim.get("gradient_scale", 1.0))
    for k, key in (("gradientX_mTm", "gx"), ("gradientY_mTm", "gy"), ("gradientZ_mTm", "gz")):
        g = sim.get(key)
        if g is not None and len(g):
            # host-side gradient scaling in float (monte_carlo.cu:288-290)
            g = (np.asarray(g, np.float32) * np.float32(gs)).astype(np.float32)
        a, _ = keep(g, np.float32); setattr(q, k, _ptr(a))
    a, n = keep(sim.get("gradient_t"), np.int32); q.gradient_t = _ptr(a); q.n_gradient = n
    return p, q, keep


def sim(case, spin_begin=0, spin_end=None, n_threads=8, f64=False, timing=False):
    """Run the oracle walk for one scale of ``case``.

    Returns dict(M1, XYZ1, T, lost, timing)."""
    p, q, keep = make_structs(case)
    n = case["n_spins"]
    spin_end = n if spin_end is None else spin_end
    n_te = q.n_te
    traj = (case["n_dummy_scan"] + 1) * case["n_timepoints"] if case["record_trajectory"] else 1
    M1 = np.array(case["M1_init"], dtype=np.float64 if f64 else np.float32, copy=True) \
        if "M1_init" in case else np.zeros((n, n_te, 3), np.float64 if f64 else np.float32)
    T = np.array(case["T_init"], np.uint8, copy=True) if "T_init" in case else \
        np.zeros((n, n_te), np.uint8)
    XYZ1 = np.array(case["XYZ1_init"], np.float32, copy=True).reshape(n, traj, 3) if "XYZ1_init" in case else \
        np.zeros((n, traj, 3), np.float32)
    lost = np.zeros(n, np.uint8)
    mask = np.ascontiguousarray(case["mask"], dtype=np.uint8)
    M0 = np.ascontiguousarray(case["M0"], dtype=np.float32)
    XYZ0 = np.ascontiguousarray(case["XYZ0_scaled"], dtype=np.float32)
    tm = np.zeros(2, np.float64)
    if f64:
        fT = None if case.get("field_T") is None else np.ascontiguousarray(case["field_T"], np.float32)
        nlost = lib().orc_sim_f64(C.byref(p), C.byref(q), C.c_float(case["B0"]), _ptr(fT),
                                  _ptr(mask), _ptr(M0), _ptr(XYZ0), _ptr(M1), _ptr(XYZ1), _ptr(T),
                                  spin_begin, spin_end, n_threads, _ptr(lost))
    else:
        f = None if case.get("field_deg") is None else np.ascontiguousarray(case["field_deg"], np.float32)
        nlost = lib().orc_sim(C.byref(p), C.byref(q), _ptr(f), _ptr(mask), _ptr(M0), _ptr(XYZ0),
                              _ptr(M1), _ptr(XYZ1), _ptr(T), spin_begin, spin_end, n_threads,
                              _ptr(lost), _ptr(tm) if timing else None)
    del keep
    return dict(M1=M1, XYZ1=XYZ1, T=T, lost=lost, n_lost=int(nlost),
                timing=dict(discard_s=float(tm[0]), total_s=float(tm[1])))


def from_engine_case(case, scale_index=0, mask=None, field_T=None):
    """Oracle inputs for one scale of an engine ``Case`` (monte_carlo.cu:241-320),
    every host transform recomputed by the oracle's own restatement.  ``mask`` and
    ``field_T`` override the case's arrays (a device-generated phantom)."""
    p = case.params
    raw = case.raw
    dt = int(np.float32(raw["TIME_STEP"]))
    TR = int(np.float32(raw["TR"]))
    tabs = {}
    for key, name in (("TE", "TE_t"), ("RF_T", "RF_t"), ("DEPHASING_T", "dephasing_t"), ("GRADIENT_T", "gradient_t")):
        v = raw.get(key, [])
        tabs[name] = timing_scale([int(np.float32(x)) for x in v], dt) if len(v) else np.zeros(0, np.int32)
    diff = np.asarray([float(np.float32(x)) for x in raw["DIFFUSIVITY"]], np.float64)
    prep = prepare(np.float32(raw["RF_FA"][0]), TR, dt, np.float32(raw["T1"][0]), diff, int(raw.get("DUMMY_SCAN", 0)))
    scale = np.float32(case.scales[scale_index])
    st = case.scale_type
    fov = np.asarray(case.fov, np.float32)
    if st == 0:
        fov_s = (scale * fov).astype(np.float32)                       # monte_carlo.cu:280
        xyz0 = (case.XYZ0 * scale).astype(np.float32)                  # monte_carlo.cu:278
    else:
        fov_s, xyz0 = fov, case.XYZ0
    mask = case.mask if mask is None else mask
    field_T = case.field_T if field_T is None else field_T
    lin = np.float32(raw.get("LINEAR_PHASE_CYCLING", 0.0))
    d = dict(
        fov=[float(x) for x in fov_s], phantom_size=list(mask.shape),
        c=prep["c"], s=prep["s"],
        linear_phase_cycling=float(lin * scale) if st == 2 else float(lin),
        quadratic_phase_cycling=float(np.float32(raw.get("QUADRATIC_PHASE_CYCLING", 0.0))),
        timestep_us=dt, TR_us=TR, n_dummy_scan=prep["n_dummy_scan"], n_spins=int(p.n_spins),
        n_timepoints=prep["n_timepoints"], seed=int(p.seed), max_iterations=int(p.max_iterations),
        cross_fov=int(p.enCrossFOV), record_trajectory=int(p.enRecordTrajectory),
        diffusivity=prep["diffusivity"], T1_ms=np.asarray(raw["T1"], np.float32),
        T2_ms=np.asarray(raw["T2"], np.float32), pXY=np.asarray(raw["P_XY"], np.float32).ravel(),
        RF_FA_deg=np.asarray(raw["RF_FA"], np.float32), RF_PH_deg=np.asarray(raw["RF_PH"], np.float32),
        RF_t=tabs["RF_t"], TE_t=tabs["TE_t"],
        dephasing_deg=np.asarray(raw.get("DEPHASING", []), np.float32), dephasing_t=tabs["dephasing_t"],
        gx=np.asarray(raw.get("GRADIENT_X", []), np.float32), gy=np.asarray(raw.get("GRADIENT_Y", []), np.float32),
        gz=np.asarray(raw.get("GRADIENT_Z", []), np.float32), gradient_t=tabs["gradient_t"],
        gradient_scale=float(scale) if st == 1 else 1.0,
        mask=mask, M0=case.M0, XYZ0_scaled=xyz0, B0=float(np.float32(raw.get("B0", 9.4))),
        field_T=field_T, prep=prep,
    )
    d["field_deg"] = None if field_T is None else scale_field(field_T, tesla2deg(d["B0"], dt))
    return d


class ConfigCase:
    """A simulation set up from a config dict with the oracle's own transforms only
    (no engine library): the object ``from_engine_case`` reads.  Key names and
    units are the reference's INI keys (config/config_default.ini)."""

    def __init__(self, cfg, mask, field_T, fov_m, XYZ0=None, M0=None):
        self.raw = dict(cfg)
        self.mask = np.ascontiguousarray(mask, np.uint8)
        self.field_T = None if field_T is None else np.ascontiguousarray(field_T, np.float32)
        self.fov = np.asarray(fov_m, np.float32) * np.ones(3, np.float32)
        n = int(float(cfg.get("NUMBER_OF_SPINS", 1000)))               # std::stod (config_reader.cpp:145)
        seed = int(cfg.get("SEED", 10))
        self.XYZ0 = init_positions(seed, self.fov, n) if XYZ0 is None else \
            np.ascontiguousarray(XYZ0, np.float32).reshape(-1, 3)
        self.M0 = np.tile(np.float32([0, 0, 1]), (n, 1)) if M0 is None else np.ascontiguousarray(M0, np.float32)
        self.scales = [float(np.float32(s)) for s in cfg.get("SCALE", [1.0])]
        self.scale_type = int(cfg.get("WHAT_TO_SCALE", 0))

        class _P:
            pass
        self.params = _P()
        self.params.n_spins = n
        self.params.seed = seed
        self.params.max_iterations = int(float(cfg.get("MAX_ITERATIONS", 1e4)))
        self.params.enCrossFOV = int(cfg.get("CROSS_FOV", 0))
        self.params.enRecordTrajectory = int(cfg.get("RECORD_TRAJECTORY", 0))

    def oracle(self, scale_index=0):
        return from_engine_case(self, scale_index)
