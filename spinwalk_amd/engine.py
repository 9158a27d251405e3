"""ctypes binding of the engine's C ABI (include/spinwalk_engine.h).

This is the Python-side mirror of the reference's kernel seam and host
transforms: ``Engine.walk_device`` is ``cu_sim`` (src/sim/kernels.cu:55-63) as
launched per scale by ``monte_carlo::run`` (src/sim/monte_carlo.cu:324-333);
``prepare_case`` restates the parameter path of ``config_reader::read`` /
``timing_scale`` and ``parameters::prepare`` (src/sim/config_reader.cpp:61-193,
39-46; src/sim/simulation_parameters.cuh:227-245) through the engine's own C
entry points.  PyTorch supplies device memory and streams only.

The shared library is required: there is no CPU fallback, and loading fails
loudly when it has not been built (``python -c "import __graft_entry__ as g;
g.build()"``).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libspinwalk_engine.so")

SW_OK = 0
SW_RNG_PHILOX, SW_RNG_REPLAY = 0, 1
SW_PREC_F32, SW_PREC_F32_KAHAN, SW_PREC_F64 = 0, 1, 2
SW_FIELD_NONE, SW_FIELD_F32, SW_FIELD_F16, SW_FIELD_BF16, SW_FIELD_Q16 = 0, 1, 2, 3, 4
SW_DTYPE_F32, SW_DTYPE_F64 = 0, 1
SW_STEP_LOOP_AUTO, SW_STEP_LOOP_ONE, SW_STEP_LOOP_TWO = 0, 1, 2

_P = C.c_void_p


class SwParams(C.Structure):
    """Mirror of ``parameters`` (src/sim/simulation_parameters.cuh:176-246)."""
    _fields_ = [
        ("fov", C.c_double * 3),
        ("phantom_size", C.c_uint64 * 3),
        ("matrix_length", C.c_int64),
        ("B0", C.c_float), ("c", C.c_float), ("s", C.c_float),
        ("linear_phase_cycling", C.c_float), ("quadratic_phase_cycling", C.c_float),
        ("timestep_us", C.c_int32), ("TR_us", C.c_int32), ("n_dummy_scan", C.c_int32),
        ("n_spins", C.c_uint32), ("n_timepoints", C.c_uint32), ("n_substrate", C.c_uint32),
        ("n_scales", C.c_uint32),
        ("seed", C.c_uint64), ("max_iterations", C.c_uint64),
        ("enCrossFOV", C.c_int32), ("enRecordTrajectory", C.c_int32), ("fieldmap_exist", C.c_int32),
    ]


class SwSeq(C.Structure):
    """Mirror of ``parameters_hvec`` (src/sim/simulation_parameters.cuh:39-65)."""
    _fields_ = [
        ("diffusivity", _P), ("T1_ms", _P), ("T2_ms", _P), ("pXY", _P), ("n_substrate", C.c_uint32),
        ("RF_FA_deg", _P), ("RF_PH_deg", _P), ("RF_t", _P), ("n_rf", C.c_uint32),
        ("TE_t", _P), ("n_te", C.c_uint32),
        ("dephasing_deg", _P), ("dephasing_t", _P), ("n_dephasing", C.c_uint32),
        ("gradientX_mTm", _P), ("gradientY_mTm", _P), ("gradientZ_mTm", _P),
        ("gradient_t", _P), ("n_gradient", C.c_uint32),
    ]


class SwWalkOpts(C.Structure):
    _fields_ = [
        ("spin_offset", C.c_uint32), ("n_local", C.c_uint32),
        ("xyz_scale", C.c_float), ("gradient_scale", C.c_float),
        ("rng", C.c_int32), ("precision", C.c_int32), ("field_format", C.c_int32),
        ("spatial_order", C.c_int32),
        ("replay_normals", _P), ("replay_normal_stride", C.c_uint64),
        ("replay_uniforms", _P), ("replay_uniform_stride", C.c_uint64),
        ("step_loop", C.c_int32),
    ]


class SwBatchSchedule(C.Structure):
    _fields_ = [("chunk_spins", C.c_uint32), ("bands_shared", C.c_uint32), ("bands_mixed", C.c_uint32),
                ("spins_per_lane", C.c_uint32), ("waves_per_simd", C.c_uint32), ("order_brick", C.c_uint32)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class SwBatchItem(C.Structure):
    _fields_ = [
        ("phantom", _P), ("sequence", _P), ("params", SwParams), ("opts", SwWalkOpts),
        ("d_M0", _P), ("d_XYZ0", _P), ("d_M1", _P), ("d_XYZ1", _P), ("d_T", _P), ("d_echo_sums", _P),
    ]


class SwShapeSpec(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("label", C.c_uint32), ("radius_um", C.c_float), ("volume_fraction", C.c_float),
        ("random_orientation", C.c_int32), ("orientation_deg", C.c_float), ("dchi", C.c_float), ("Y", C.c_float),
        ("seed", C.c_uint32), ("trunc_radii", C.c_float), ("radius_min_um", C.c_float),
    ]


SW_SHAPE_CYLINDER, SW_SHAPE_SPHERE, SW_SHAPE_AXONS = 0, 1, 2


class SwWalkStats(C.Structure):
    _fields_ = [
        ("n_lost", C.c_uint64), ("n_replay_exhausted", C.c_uint64),
        ("n_steps", C.c_uint64), ("n_attempts", C.c_uint64), ("kernel_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# Every entry point include/spinwalk_engine.h declares (checked by the ABI test).
EXPORTS = (
    "sw_engine_create", "sw_engine_destroy", "sw_last_error", "sw_device_info", "sw_device_count",
    "sw_version", "sw_params_prepare", "sw_timing_scale", "sw_tesla2deg", "sw_init_positions",
    "sw_replay_fill", "sw_phantom_create", "sw_phantom_create_device", "sw_phantom_destroy",
    "sw_phantom_info", "sw_phantom_download", "sw_phantom_generate", "sw_sequence_create", "sw_sequence_destroy", "sw_walk_device", "sw_walk_device_noalloc", "sw_walk_host",
    "sw_batch_create", "sw_batch_run", "sw_batch_destroy", "sw_batch_get_schedule", "sw_batch_set_schedule",
    "sw_batch_tune", "sw_batch_info", "sw_batch_graph_info",
    "sw_echo_reduce_device", "sw_philox_draws_device", "sw_probe_gather_device", "sw_stream_sync",
)

_lib = None


class EngineError(RuntimeError):
    pass


def load():
    """Load libspinwalk_engine.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineError(f"{LIB_PATH} is not built: run __graft_entry__.build() (make -C spinwalk_amd/csrc)")
    # PyTorch-ROCm bundles its own libamdhip64.so.7 and asks for it by a name that
    # does not match an already-loaded ROCm copy, so loading the engine first puts
    # two HIP runtimes in the process and torch then finds no GPU.  Loading torch
    # first makes the engine bind to torch's copy (same soname): one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    L.sw_engine_create.argtypes = [C.c_int, C.POINTER(_P)]
    L.sw_engine_destroy.argtypes = [_P]
    L.sw_engine_destroy.restype = None
    L.sw_last_error.argtypes = [_P]
    L.sw_last_error.restype = C.c_char_p
    L.sw_device_info.argtypes = [C.c_int, C.c_char_p, C.c_size_t]
    L.sw_device_count.argtypes = []
    L.sw_version.restype = C.c_char_p
    L.sw_params_prepare.argtypes = [C.POINTER(SwParams), _P, C.c_float, C.c_float]
    L.sw_timing_scale.argtypes = [_P, C.c_uint32, C.c_int32]
    L.sw_timing_scale.restype = None
    L.sw_tesla2deg.argtypes = [C.c_float, C.c_int32]
    L.sw_tesla2deg.restype = C.c_float
    L.sw_init_positions.argtypes = [C.c_uint64, _P, C.c_uint32, _P]
    L.sw_init_positions.restype = None
    L.sw_replay_fill.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, _P, C.c_uint64, _P, C.c_uint64, C.c_int]
    L.sw_phantom_create.argtypes = [_P, _P, _P, _P, _P, C.c_uint32, C.POINTER(_P)]
    L.sw_phantom_create_device.argtypes = [_P, _P, _P, _P, _P, C.c_uint32, C.POINTER(_P)]
    L.sw_phantom_destroy.argtypes = [_P]
    L.sw_phantom_destroy.restype = None
    L.sw_phantom_info.argtypes = [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_float)]
    L.sw_phantom_download.argtypes = [_P, _P, _P]
    L.sw_phantom_generate.argtypes = [_P, _P, _P, C.POINTER(SwShapeSpec), C.c_uint32, C.c_uint32, C.POINTER(_P), _P]
    L.sw_sequence_create.argtypes = [_P, C.POINTER(SwParams), C.POINTER(SwSeq), C.POINTER(_P)]
    L.sw_sequence_destroy.argtypes = [_P]
    L.sw_sequence_destroy.restype = None
    L.sw_walk_device.argtypes = [_P, _P, _P, C.POINTER(SwParams), C.POINTER(SwWalkOpts), _P, _P, _P, _P, _P,
                                 _P, C.POINTER(SwWalkStats)]
    L.sw_walk_device_noalloc.argtypes = L.sw_walk_device.argtypes
    L.sw_walk_host.argtypes = [_P, _P, _P, C.POINTER(SwParams), C.POINTER(SwWalkOpts), _P, _P, _P, _P, _P,
                               _P, _P, C.POINTER(SwWalkStats)]
    L.sw_echo_reduce_device.argtypes = [_P, _P, C.c_int, _P, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P]
    L.sw_philox_draws_device.argtypes = [_P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P]
    L.sw_probe_gather_device.argtypes = [_P, _P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, _P, _P,
                                         C.POINTER(C.c_float)]
    L.sw_stream_sync.argtypes = [_P, _P]
    L.sw_batch_create.argtypes = [_P, C.POINTER(SwBatchItem), C.c_uint32, C.POINTER(_P)]
    L.sw_batch_run.argtypes = [_P, _P, C.POINTER(SwWalkStats), C.POINTER(SwWalkStats)]
    L.sw_batch_destroy.argtypes = [_P]
    L.sw_batch_destroy.restype = None
    L.sw_batch_info.argtypes = [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.sw_batch_graph_info.argtypes = [_P, C.POINTER(C.c_int32), C.POINTER(C.c_uint32)]
    L.sw_batch_get_schedule.argtypes = [_P, C.POINTER(SwBatchSchedule)]
    L.sw_batch_set_schedule.argtypes = [_P, C.POINTER(SwBatchSchedule)]
    L.sw_batch_tune.argtypes = [_P, _P, C.c_uint32, C.POINTER(SwBatchSchedule), C.POINTER(C.c_float),
                                C.POINTER(C.c_float)]
    _lib = L
    return L


def _stream_arg(stream):
    """(hipStream_t handle or None, sync_after) for a torch stream argument.

    None selects the engine's own stream (the caller synchronises, e.g. with
    torch.cuda.synchronize()).  A NULL handle at the C ABI also means the engine's
    stream, so torch's legacy default stream (handle 0) cannot be named there: work
    given that stream runs on the engine's stream and the call waits for it."""
    if stream is None:
        return None, False
    h = int(stream.cuda_stream)
    return (None, True) if h == 0 else (_P(h), False)


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# host transforms
# ---------------------------------------------------------------------------
def _i32_via_float(v):
    """str2vec<int32_t>: istream_iterator<float> then static_cast (config_reader.cpp:16-23)."""
    return np.asarray([int(np.float32(x)) for x in v], dtype=np.int32)


def tesla2deg(B0, dt_us):
    return float(load().sw_tesla2deg(C.c_float(B0), int(dt_us)))


def init_positions(seed, fov, n_spins):
    fov = np.ascontiguousarray(fov, np.float32)
    out = np.zeros((n_spins, 3), np.float32)
    load().sw_init_positions(int(seed), _np_ptr(fov), int(n_spins), _np_ptr(out))
    return out


def replay_fill(seed, spin_begin, n, n_normals, n_uniforms, n_threads=8):
    """Record the reference CPU path's draw streams for spins [spin_begin, spin_begin+n)."""
    nr = np.zeros((n, n_normals), np.float32)
    ur = np.zeros((n, n_uniforms), np.float32)
    rc = load().sw_replay_fill(int(seed), int(spin_begin), int(n), _np_ptr(nr), int(n_normals), _np_ptr(ur),
                               int(n_uniforms), int(n_threads))
    if rc != SW_OK:
        raise EngineError(f"sw_replay_fill failed ({rc})")
    return nr, ur


@dataclass
class Case:
    """A prepared simulation of one phantom: what monte_carlo::run holds when it
    launches the kernel (parameters, host tables, phantom, initial spins)."""
    params: SwParams
    tables: dict
    mask: np.ndarray
    field_T: np.ndarray | None
    fov: np.ndarray                    # float32 metres, as read from the phantom file
    XYZ0: np.ndarray                   # float32 [n,3] metres (unscaled)
    M0: np.ndarray                     # float32 [n,3]
    scales: list = field(default_factory=lambda: [1.0])
    scale_type: int = 0                # WHAT_TO_SCALE: 0 FoV, 1 gradient, 2 phase cycling
    raw: dict = field(default_factory=dict)

    @property
    def n_spins(self):
        return int(self.params.n_spins)

    @property
    def n_te(self):
        return int(len(self.tables["TE_t"]))

    @property
    def traj_len(self):
        p = self.params
        return (p.n_dummy_scan + 1) * p.n_timepoints if p.enRecordTrajectory else 1

    def steps_per_spin(self):
        p = self.params
        return (p.n_dummy_scan + 1) * p.n_timepoints

    def scale_params(self, scale):
        """Per-scale parameters and launch scales (monte_carlo.cu:273-305)."""
        p = SwParams()
        C.pointer(p)[0] = self.params
        xyz_scale, grad_scale = 1.0, 1.0
        s = np.float32(scale)
        if self.scale_type == 0:
            for i in range(3):
                p.fov[i] = float(np.float32(s * self.fov[i]))      # float*float (monte_carlo.cu:280)
            xyz_scale = float(s)
        else:
            for i in range(3):
                p.fov[i] = float(self.fov[i])
            if self.scale_type == 1:
                grad_scale = float(s)
            elif self.scale_type == 2:
                p.linear_phase_cycling = float(np.float32(self.raw["LINEAR_PHASE_CYCLING"]) * s)
        return p, xyz_scale, grad_scale


def prepare_case(cfg, mask, field_T, fov, XYZ0=None, M0=None):
    """Turn a config dict (reference units and key names) into a Case.

    Numeric conversions follow config_reader::read (src/sim/config_reader.cpp:107-189):
    TR/TIME_STEP through std::stof, time lists through str2vec (float then int),
    NUMBER_OF_SPINS and MAX_ITERATIONS through std::stod, SEED through std::stoi;
    then timing_scale (:39-46), the Tesla->deg factor and parameters::prepare via
    the engine's own entry points."""
    L = load()
    g = lambda k, d=None: cfg.get(k, d)
    p = SwParams()
    mask = np.ascontiguousarray(mask, dtype=np.uint8)
    dims = mask.shape
    fov = np.ascontiguousarray(fov, dtype=np.float32)
    for i in range(3):
        p.phantom_size[i] = int(dims[i])
        p.fov[i] = float(fov[i])
    p.matrix_length = int(mask.size)
    p.B0 = float(np.float32(g("B0", 9.4)))
    p.TR_us = int(np.float32(g("TR")))
    p.timestep_us = int(np.float32(g("TIME_STEP")))
    p.n_dummy_scan = int(g("DUMMY_SCAN", 0))
    p.linear_phase_cycling = float(np.float32(g("LINEAR_PHASE_CYCLING", 0.0)))
    p.quadratic_phase_cycling = float(np.float32(g("QUADRATIC_PHASE_CYCLING", 0.0)))
    p.seed = int(g("SEED", 0))
    p.n_spins = int(float(g("NUMBER_OF_SPINS", 1000)))
    p.enCrossFOV = int(g("CROSS_FOV", 0))
    p.enRecordTrajectory = int(g("RECORD_TRAJECTORY", 0))
    p.max_iterations = int(float(g("MAX_ITERATIONS", 1e4)))
    diff = np.asarray([float(np.float32(x)) for x in g("DIFFUSIVITY")], np.float64)
    T1 = np.asarray(g("T1"), np.float32)
    T2 = np.asarray(g("T2"), np.float32)
    pXY = np.asarray(g("P_XY"), np.float32).ravel()
    p.n_substrate = len(diff)
    scales = [float(np.float32(s)) for s in g("SCALE", [1.0])]
    p.n_scales = len(scales)
    p.fieldmap_exist = int(field_T is not None)
    tabs = {}
    for key, name in (("TE", "TE_t"), ("RF_T", "RF_t"), ("DEPHASING_T", "dephasing_t"), ("GRADIENT_T", "gradient_t")):
        v = _i32_via_float(g(key, []))
        if len(v):
            L.sw_timing_scale(_np_ptr(v), len(v), p.timestep_us)
        tabs[name] = v
    tabs["RF_FA_deg"] = np.asarray(g("RF_FA"), np.float32)
    tabs["RF_PH_deg"] = np.asarray(g("RF_PH"), np.float32)
    tabs["dephasing_deg"] = np.asarray(g("DEPHASING", []), np.float32)
    for k, n in (("GRADIENT_X", "gx"), ("GRADIENT_Y", "gy"), ("GRADIENT_Z", "gz")):
        tabs[n] = np.asarray(g(k, []), np.float32)
    rc = L.sw_params_prepare(C.byref(p), _np_ptr(diff), C.c_float(tabs["RF_FA_deg"][0]), C.c_float(T1[0]))
    if rc != SW_OK:
        raise EngineError("sw_params_prepare failed")
    tabs.update(diffusivity=diff, T1_ms=T1, T2_ms=T2, pXY=pXY)
    if XYZ0 is None:
        XYZ0 = init_positions(p.seed, fov, p.n_spins)
    XYZ0 = np.ascontiguousarray(XYZ0, np.float32).reshape(-1, 3)
    if M0 is None:
        M0 = np.tile(np.asarray([0, 0, 1], np.float32), (p.n_spins, 1))   # monte_carlo.cu:162-164
    ft = None if field_T is None else np.ascontiguousarray(field_T, np.float32)
    return Case(params=p, tables=tabs, mask=mask, field_T=ft, fov=fov, XYZ0=XYZ0,
                M0=np.ascontiguousarray(M0, np.float32), scales=scales,
                scale_type=int(g("WHAT_TO_SCALE", 0)), raw=dict(cfg))


def make_seq(tables):
    """SwSeq view of a tables dict (arrays kept alive by the returned tuple)."""
    keep = []

    def a(key, dt):
        v = tables.get(key)
        if v is None or len(v) == 0:
            return None, 0
        arr = np.ascontiguousarray(v, dtype=dt)
        keep.append(arr)
        return _np_ptr(arr), arr.size

    q = SwSeq()
    q.diffusivity, q.n_substrate = a("diffusivity", np.float64)
    q.T1_ms, _ = a("T1_ms", np.float32)
    q.T2_ms, _ = a("T2_ms", np.float32)
    q.pXY, _ = a("pXY", np.float32)
    q.RF_FA_deg, _ = a("RF_FA_deg", np.float32)
    q.RF_PH_deg, _ = a("RF_PH_deg", np.float32)
    q.RF_t, q.n_rf = a("RF_t", np.int32)
    q.TE_t, q.n_te = a("TE_t", np.int32)
    q.dephasing_deg, _ = a("dephasing_deg", np.float32)
    q.dephasing_t, q.n_dephasing = a("dephasing_t", np.int32)
    q.gradientX_mTm, _ = a("gx", np.float32)
    q.gradientY_mTm, _ = a("gy", np.float32)
    q.gradientZ_mTm, _ = a("gz", np.float32)
    q.gradient_t, q.n_gradient = a("gradient_t", np.int32)
    return q, keep


# ---------------------------------------------------------------------------
# engine objects
# ---------------------------------------------------------------------------
class Engine:
    """One engine per host thread, bound to one gfx950 device."""

    def __init__(self, device=0):
        self.L = load()
        h = _P()
        rc = self.L.sw_engine_create(int(device), C.byref(h))
        if rc != SW_OK:
            raise EngineError(f"sw_engine_create: {self.L.sw_last_error(None).decode()} ({rc})")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.L.sw_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != SW_OK:
            raise EngineError(f"{what}: {self.L.sw_last_error(self.h).decode()} ({rc})")

    def last_error(self):
        return self.L.sw_last_error(self.h).decode()

    # ---- phantom / sequence
    def phantom(self, mask, field_T, fov, formats=(SW_FIELD_F32,)):
        mask = np.ascontiguousarray(mask, np.uint8)
        dims = np.asarray(mask.shape, np.uint64)
        fov = np.ascontiguousarray(fov, np.float32)
        ft = None if field_T is None else np.ascontiguousarray(field_T, np.float32)
        bits = 0
        for f in formats:
            bits |= 1 << f
        h = _P()
        self._check(self.L.sw_phantom_create(self.h, _np_ptr(mask), _np_ptr(ft), _np_ptr(dims), _np_ptr(fov),
                                             bits, C.byref(h)), "sw_phantom_create")
        ph = Phantom(self, h)
        ph.dims = tuple(int(d) for d in dims)
        ph.fov_m = fov
        return ph

    def generate_phantom(self, dims, fov_um, specs, formats=(SW_FIELD_F32,)):
        """Device phantom generator.  specs: list of dicts with SwShapeSpec fields.
        Returns (Phantom, analytic volume fraction placed per spec, percent)."""
        arr = (SwShapeSpec * len(specs))()
        for i, sp in enumerate(specs):
            for k, v in sp.items():
                setattr(arr[i], k, v)
        dims = np.asarray(dims, np.uint64)
        fov = np.ascontiguousarray(fov_um, np.float32)
        vf = np.zeros(len(specs), np.float64)
        bits = 0
        for f in formats:
            bits |= 1 << f
        h = _P()
        self._check(self.L.sw_phantom_generate(self.h, _np_ptr(dims), _np_ptr(fov), arr, len(specs), bits,
                                               C.byref(h), _np_ptr(vf)), "sw_phantom_generate")
        ph = Phantom(self, h)
        ph.dims = tuple(int(d) for d in dims)
        ph.fov_m = (fov * np.float32(1e-6)).astype(np.float32)
        return ph, vf

    def phantom_device(self, d_mask, d_field, dims, fov, formats=(SW_FIELD_F32,)):
        dims = np.asarray(dims, np.uint64)
        fov = np.ascontiguousarray(fov, np.float32)
        bits = 0
        for f in formats:
            bits |= 1 << f
        h = _P()
        self._check(self.L.sw_phantom_create_device(self.h, _P(d_mask.data_ptr()),
                                                    None if d_field is None else _P(d_field.data_ptr()),
                                                    _np_ptr(dims), _np_ptr(fov), bits, C.byref(h)),
                    "sw_phantom_create_device")
        return Phantom(self, h)

    def sequence(self, params, tables):
        q, keep = make_seq(tables)
        h = _P()
        self._check(self.L.sw_sequence_create(self.h, C.byref(params), C.byref(q), C.byref(h)),
                    "sw_sequence_create")
        del keep
        sq = Sequence(self, h)
        sq.n_te, sq.n_sub = int(q.n_te), int(q.n_substrate)
        return sq

    @staticmethod
    def _check_spin_buffers(seq, params, n, precision, M0, XYZ0, M1, XYZ1, T, replay, rng, exact=True):
        """Shapes and dtypes of a walk's buffers (numpy arrays or torch tensors):
        the C entry points take bare pointers, so a buffer smaller than the walk
        writes would be overrun.  exact=False accepts buffers with more than n
        rows (a walk of their first n spins)."""
        def shape_dtype(a):
            sh, dt = tuple(a.shape), str(a.dtype).replace("torch.", "")
            if not exact and len(sh) >= 1 and sh[0] >= n:
                sh = (n,) + sh[1:]
            return sh, dt

        traj = (params.n_dummy_scan + 1) * params.n_timepoints if params.enRecordTrajectory else 1
        m1 = "float64" if precision == SW_PREC_F64 else "float32"
        want = [("M0", M0, (n, 3), "float32"), ("XYZ0", XYZ0, (n, 3), "float32"),
                ("M1", M1, (n, seq.n_te, 3), m1), ("XYZ1", XYZ1, (n, traj, 3), "float32"),
                ("T", T, (n, seq.n_te), "uint8")]
        for name, a, shape, dt in want:
            got = shape_dtype(a)
            if got != (shape, dt):
                raise EngineError(f"{name}: expected shape {shape} {dt}, got {got[0]} {got[1]}")
        if rng == SW_RNG_REPLAY:
            if replay is None:
                raise EngineError("replay mode needs normal and uniform streams")
            for name, a in zip(("replay normals", "replay uniforms"), replay):
                sh, dt = shape_dtype(a)
                if len(sh) != 2 or sh[0] != n or dt != "float32":
                    raise EngineError(f"{name}: expected shape ({n}, stride) float32, got {sh} {dt}")

    # ---- walks
    def walk_host(self, ph, seq, params, M0, XYZ0, M1, XYZ1, T, *, spin_offset=0, xyz_scale=1.0,
                  gradient_scale=1.0, rng=SW_RNG_PHILOX, precision=SW_PREC_F32, field_format=SW_FIELD_F32,
                  spatial_order=0, replay=None, step_loop=SW_STEP_LOOP_AUTO):
        """Host arrays in, host arrays out (in place), reference semantics."""
        o = SwWalkOpts()
        o.spin_offset, o.n_local = int(spin_offset), int(M0.shape[0])
        o.xyz_scale, o.gradient_scale = float(xyz_scale), float(gradient_scale)
        o.rng, o.precision, o.field_format, o.spatial_order = rng, precision, field_format, spatial_order
        o.step_loop = int(step_loop)
        rn = ru = None
        if replay is not None:
            rn = np.ascontiguousarray(replay[0], np.float32)
            ru = np.ascontiguousarray(replay[1], np.float32)
            o.replay_normal_stride, o.replay_uniform_stride = rn.shape[1], ru.shape[1]
        self._check_spin_buffers(seq, params, o.n_local, precision, M0, XYZ0, M1, XYZ1, T,
                                 None if replay is None else (rn, ru), rng)
        st = SwWalkStats()
        for arr in (M0, XYZ0, M1, XYZ1, T):
            if not arr.flags["C_CONTIGUOUS"]:
                raise EngineError("walk_host needs C-contiguous arrays")
        rc = self.L.sw_walk_host(self.h, ph.h, seq.h, C.byref(params), C.byref(o), _np_ptr(M0), _np_ptr(XYZ0),
                                 _np_ptr(M1), _np_ptr(XYZ1), _np_ptr(T), _np_ptr(rn), _np_ptr(ru), C.byref(st))
        self._check(rc, "sw_walk_host")
        return st.as_dict()

    def walk_device(self, ph, seq, params, M0, XYZ0, M1, XYZ1, T, *, spin_offset=0, n_local=None, xyz_scale=1.0,
                    gradient_scale=1.0, rng=SW_RNG_PHILOX, precision=SW_PREC_F32, field_format=SW_FIELD_F32,
                    spatial_order=0, replay=None, stream=None, stats=True, noalloc=False,
                    step_loop=SW_STEP_LOOP_AUTO):
        """torch device tensors; launches on `stream` (torch stream or None).
        noalloc: walk through sw_walk_device_noalloc (fixed device scratch, no
        allocation) instead of sw_walk_device.  step_loop: SW_STEP_LOOP_* (the
        production walk's one- or two-attempt step loop; same results)."""
        o = SwWalkOpts()
        o.spin_offset = int(spin_offset)
        o.n_local = int(M0.shape[0] if n_local is None else n_local)
        o.xyz_scale, o.gradient_scale = float(xyz_scale), float(gradient_scale)
        o.rng, o.precision, o.field_format, o.spatial_order = rng, precision, field_format, spatial_order
        o.step_loop = int(step_loop)
        self._check_spin_buffers(seq, params, o.n_local, precision, M0, XYZ0, M1, XYZ1, T, replay, rng,
                                 exact=n_local is None)
        for t in (M0, XYZ0, M1, XYZ1, T) + (tuple(replay) if replay is not None else ()):
            if not t.is_contiguous():
                raise EngineError("walk_device needs contiguous tensors")
        if replay is not None:
            o.replay_normals, o.replay_uniforms = replay[0].data_ptr(), replay[1].data_ptr()
            o.replay_normal_stride, o.replay_uniform_stride = replay[0].shape[1], replay[1].shape[1]
        st = SwWalkStats()
        sp, wait = _stream_arg(stream)
        ptr = lambda t: None if t is None else _P(t.data_ptr())
        fn = self.L.sw_walk_device_noalloc if noalloc else self.L.sw_walk_device
        rc = fn(self.h, ph.h, seq.h, C.byref(params), C.byref(o), ptr(M0), ptr(XYZ0), ptr(M1), ptr(XYZ1), ptr(T), sp,
                C.byref(st) if stats else None)
        self._check(rc, "sw_walk_device_noalloc" if noalloc else "sw_walk_device")
        if wait:
            self.sync()
        return st.as_dict() if stats else None

    def batch(self, items):
        """A Batch of tiles; each item is a dict with phantom, sequence, params and
        device tensors M0, XYZ0, M1, XYZ1, T (optional: spin_offset, n_local,
        xyz_scale, gradient_scale, rng, precision, field_format, spatial_order,
        replay=(normals, uniforms), echo_sums)."""
        return Batch(self, items)

    def echo_reduce(self, d_M1, d_T, n_sub, d_out, stream=None):
        import torch
        n_spins, n_te = d_T.shape[0], d_T.shape[1]
        if d_T.dtype != torch.uint8 or d_T.dim() != 2:
            raise EngineError("T: expected a (n_spins, n_te) uint8 tensor")
        if tuple(d_M1.shape) != (n_spins, n_te, 3) or d_M1.dtype not in (torch.float32, torch.float64):
            raise EngineError(f"M1: expected shape {(n_spins, n_te, 3)} float32 or float64, got "
                              f"{tuple(d_M1.shape)} {d_M1.dtype}")
        if tuple(d_out.shape) != (n_te, n_sub, 3) or d_out.dtype != torch.float64:
            raise EngineError(f"echo sums: expected shape {(n_te, n_sub, 3)} float64")
        if not (d_M1.is_contiguous() and d_T.is_contiguous() and d_out.is_contiguous()):
            raise EngineError("echo_reduce needs contiguous tensors")
        dt = SW_DTYPE_F64 if d_M1.dtype == torch.float64 else SW_DTYPE_F32
        sp, wait = _stream_arg(stream)
        self._check(self.L.sw_echo_reduce_device(self.h, _P(d_M1.data_ptr()), dt, _P(d_T.data_ptr()), n_spins, n_te,
                                                 n_sub, _P(d_out.data_ptr()), sp), "sw_echo_reduce_device")
        if wait:
            self.sync()

    def probe_gather(self, d_buf, n_lanes, n_reads, d_sink, seed=1, stream=None):
        """Random 128-byte line reads over the tensor d_buf; returns the kernel time
        in ms.  d_sink: uint64/int64 device tensor of n_lanes entries."""
        if d_sink.numel() < n_lanes or d_sink.element_size() != 8:
            raise EngineError("d_sink needs n_lanes 8-byte entries")
        ms = C.c_float()
        sp, _ = _stream_arg(stream)
        self._check(self.L.sw_probe_gather_device(self.h, _P(d_buf.data_ptr()), d_buf.numel() * d_buf.element_size(),
                                                  int(n_lanes), int(n_reads), int(seed), _P(d_sink.data_ptr()), sp,
                                                  C.byref(ms)), "sw_probe_gather_device")
        return ms.value

    def philox_draws(self, seed, spin0, n_spins, n_att, d_out, stream=None):
        sp, wait = _stream_arg(stream)
        self._check(self.L.sw_philox_draws_device(self.h, int(seed), int(spin0), int(n_spins), int(n_att),
                                                  _P(d_out.data_ptr()), sp), "sw_philox_draws_device")
        if wait:
            self.sync()

    def sync(self, stream=None):
        sp, _ = _stream_arg(stream)
        self._check(self.L.sw_stream_sync(self.h, sp), "sw_stream_sync")


class Batch:
    """Many tiles walked in one launch (sw_batch_create / sw_batch_run).  Keeps the
    tiles' tensors, phantoms and sequences alive while it lives."""

    def __init__(self, eng, items):
        self.eng = eng
        self.keep = list(items)
        arr = (SwBatchItem * len(items))()
        ptr = lambda t: None if t is None else _P(t.data_ptr())
        for i, it in enumerate(items):
            b = arr[i]
            b.phantom, b.sequence = it["phantom"].h, it["sequence"].h
            C.pointer(b.params)[0] = it["params"]
            o = b.opts
            o.spin_offset = int(it.get("spin_offset", 0))
            o.n_local = int(it.get("n_local", it["M0"].shape[0]))
            o.xyz_scale, o.gradient_scale = float(it.get("xyz_scale", 1.0)), float(it.get("gradient_scale", 1.0))
            o.rng, o.precision = int(it.get("rng", SW_RNG_PHILOX)), int(it.get("precision", SW_PREC_F32))
            o.field_format, o.spatial_order = int(it.get("field_format", SW_FIELD_F32)), int(it.get("spatial_order", 1))
            rp = it.get("replay")
            eng._check_spin_buffers(it["sequence"], it["params"], o.n_local, o.precision, it["M0"], it["XYZ0"],
                                    it["M1"], it["XYZ1"], it["T"], rp, o.rng, exact=False)
            if rp is not None:
                o.replay_normals, o.replay_uniforms = rp[0].data_ptr(), rp[1].data_ptr()
                o.replay_normal_stride, o.replay_uniform_stride = rp[0].shape[1], rp[1].shape[1]
            b.d_M0, b.d_XYZ0, b.d_M1 = ptr(it["M0"]), ptr(it["XYZ0"]), ptr(it["M1"])
            b.d_XYZ1, b.d_T, b.d_echo_sums = ptr(it["XYZ1"]), ptr(it["T"]), ptr(it.get("echo_sums"))
        h = _P()
        eng._check(eng.L.sw_batch_create(eng.h, arr, len(items), C.byref(h)), "sw_batch_create")
        self.h = h
        self.n = len(items)

    def run(self, stream=None, stats=True):
        sp, wait = _stream_arg(stream)
        if not stats:
            self.eng._check(self.eng.L.sw_batch_run(self.h, sp, None, None), "sw_batch_run")
            if wait:
                self.eng.sync()
            return None
        per = (SwWalkStats * self.n)()
        tot = SwWalkStats()
        self.eng._check(self.eng.L.sw_batch_run(self.h, sp, per, C.byref(tot)), "sw_batch_run")
        return dict(total=tot.as_dict(), items=[per[i].as_dict() for i in range(self.n)])

    def info(self):
        """Walks of the batch and tiles walked inside shared (fused) walks."""
        w, f = C.c_uint32(), C.c_uint32()
        self.eng._check(self.eng.L.sw_batch_info(self.h, C.byref(w), C.byref(f)), "sw_batch_info")
        return dict(walks=w.value, fused_tiles=f.value)

    def graph_info(self):
        """Whether runs launch the batch's captured HIP graph, and its node count."""
        g, n = C.c_int32(), C.c_uint32()
        self.eng._check(self.eng.L.sw_batch_graph_info(self.h, C.byref(g), C.byref(n)), "sw_batch_graph_info")
        return dict(uses_graph=bool(g.value), nodes=n.value)

    @property
    def schedule(self):
        sc = SwBatchSchedule()
        self.eng._check(self.eng.L.sw_batch_get_schedule(self.h, C.byref(sc)), "sw_batch_get_schedule")
        return sc.as_dict()

    def set_schedule(self, chunk_spins, bands_shared, bands_mixed, spins_per_lane=1, waves_per_simd=0, order_brick=8):
        sc = SwBatchSchedule(int(chunk_spins), int(bands_shared), int(bands_mixed), int(spins_per_lane),
                             int(waves_per_simd), int(order_brick))
        self.eng._check(self.eng.L.sw_batch_set_schedule(self.h, C.byref(sc)), "sw_batch_set_schedule")

    def tune(self, max_trials=0, stream=None):
        """The autotuner (sw_batch_tune): times full runs of the batch under
        candidate schedules and keeps the fastest.  Returns the schedule kept and
        the run times (ms) of it and of the starting schedule."""
        sp, _ = _stream_arg(stream)
        sc, best, first = SwBatchSchedule(), C.c_float(), C.c_float()
        self.eng._check(self.eng.L.sw_batch_tune(self.h, sp, int(max_trials), C.byref(sc), C.byref(best),
                                                 C.byref(first)), "sw_batch_tune")
        return dict(schedule=sc.as_dict(), best_ms=best.value, first_ms=first.value)

    def close(self):
        if getattr(self, "h", None):
            self.eng.L.sw_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Phantom:
    def __init__(self, eng, h):
        self.eng, self.h = eng, h

    def download(self, field=True):
        dims = self.dims
        mask = np.zeros(dims, np.uint8)
        f = np.zeros(dims, np.float32) if field else None
        self.eng._check(self.eng.L.sw_phantom_download(self.h, _np_ptr(mask), _np_ptr(f)), "sw_phantom_download")
        return mask, f

    def info(self):
        n, s = C.c_uint32(), C.c_float()
        self.eng._check(self.eng.L.sw_phantom_info(self.h, C.byref(n), C.byref(s)), "sw_phantom_info")
        return int(n.value), float(s.value)

    def close(self):
        if self.h:
            self.eng.L.sw_phantom_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Sequence:
    def __init__(self, eng, h):
        self.eng, self.h = eng, h

    def close(self):
        if self.h:
            self.eng.L.sw_sequence_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
