"""ctypes binding of the host driver library (include/spinwalk_host.h).

``load_config`` mirrors ``config_reader::prepare`` + ``parameters::prepare``
(src/sim/config_reader.cpp:25-37, src/sim/simulation_parameters.cuh:227-245);
``run_config`` mirrors ``monte_carlo::run`` (src/sim/monte_carlo.cu:199-355);
``h5_read``/``h5_write`` give the HDF5 layout of ``h5_helper``
(src/sim/h5_helper.h:48-129) to callers without an HDF5 binding.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import engine as E

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libspinwalk_host.so")
_P = C.c_void_p
_lib = None

EXPORTS = ("sw_config_load", "sw_config_destroy", "sw_config_free_string", "sw_config_params", "sw_config_scales",
           "sw_config_file", "sw_config_n_files", "sw_config_output", "sw_config_seq_name", "sw_run_config",
           "sw_run_config_ex", "sw_run_config_opts", "sw_sim_execute", "sw_h5_read", "sw_h5_write", "sw_generate_config",
           "sw_generate_pgse", "sw_host_last_error")


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise E.EngineError(f"{LIB_PATH} is not built (make -C spinwalk_amd/csrc)")
    E.load()
    L = C.CDLL(LIB_PATH)
    L.sw_config_load.argtypes = [C.c_char_p, C.POINTER(_P), C.POINTER(C.c_char_p)]
    L.sw_config_destroy.argtypes = [_P]
    L.sw_config_destroy.restype = None
    L.sw_config_free_string.argtypes = [C.c_char_p]
    L.sw_config_free_string.restype = None
    L.sw_config_params.argtypes = [_P, C.POINTER(E.SwParams), C.POINTER(E.SwSeq)]
    L.sw_config_scales.argtypes = [_P, C.POINTER(C.POINTER(C.c_float)), C.POINTER(C.c_uint32), C.POINTER(C.c_int32)]
    L.sw_config_file.argtypes = [_P, C.c_char_p, C.c_uint32]
    L.sw_config_file.restype = C.c_char_p
    L.sw_config_n_files.argtypes = [_P, C.c_char_p]
    L.sw_config_n_files.restype = C.c_uint32
    L.sw_config_output.argtypes = [_P, C.c_uint32]
    L.sw_config_output.restype = C.c_char_p
    L.sw_config_seq_name.argtypes = [_P]
    L.sw_config_seq_name.restype = C.c_char_p
    L.sw_run_config.argtypes = [_P, C.c_char_p]
    L.sw_run_config_ex.argtypes = [_P, C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32]
    L.sw_run_config_opts.argtypes = [_P, C.c_char_p, C.POINTER(SwRunOptions)]
    L.sw_sim_execute.argtypes = [C.c_int, C.c_uint32, C.POINTER(C.c_char_p), C.c_uint32]
    L.sw_h5_read.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, _P, C.c_uint64, C.POINTER(C.c_uint64),
                             C.POINTER(C.c_uint32)]
    L.sw_h5_write.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, _P, C.POINTER(C.c_uint64), C.c_uint32]
    L.sw_generate_config.argtypes = [C.c_char_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_char_p), C.c_uint32,
                                     C.c_char_p]
    L.sw_generate_pgse.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.c_uint32, C.POINTER(C.c_float),
                                   C.c_uint32, C.c_uint32, C.c_uint32]
    L.sw_host_last_error.restype = C.c_char_p
    _lib = L
    return L


class ConfigError(ValueError):
    pass


class SwRunOptions(C.Structure):
    _fields_ = [("rng", C.c_int32), ("precision", C.c_int32), ("field_format", C.c_int32),
                ("spatial_order", C.c_int32), ("trajectory_chunk_spins", C.c_uint32)]


def _arr(ptr, n, dt):
    if n == 0 or not ptr:
        return np.zeros(0, dt)
    ct = {np.float32: C.c_float, np.float64: C.c_double, np.int32: C.c_int32}[dt]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()


def load_config(path):
    """Parse, check and prepare a config; returns a dict (raises ConfigError)."""
    L = load()
    h = _P()
    err = C.c_char_p()
    rc = L.sw_config_load(str(path).encode(), C.byref(h), C.byref(err))
    if rc != E.SW_OK:
        msg = err.value.decode() if err.value else "config rejected"
        raise ConfigError(msg)
    try:
        p, q = E.SwParams(), E.SwSeq()
        L.sw_config_params(h, C.byref(p), C.byref(q))
        sc = C.POINTER(C.c_float)()
        n, what = C.c_uint32(), C.c_int32()
        L.sw_config_scales(h, C.byref(sc), C.byref(n), C.byref(what))
        ns = q.n_substrate
        out = dict(
            params=p,
            TR_us=p.TR_us, timestep_us=p.timestep_us, n_timepoints=p.n_timepoints, n_dummy_scan=p.n_dummy_scan,
            c=p.c, s=p.s, B0=p.B0, seed=p.seed, n_spins=p.n_spins, max_iterations=p.max_iterations,
            cross_fov=p.enCrossFOV, record_trajectory=p.enRecordTrajectory,
            linear_phase_cycling=p.linear_phase_cycling, quadratic_phase_cycling=p.quadratic_phase_cycling,
            diffusivity=_arr(q.diffusivity, ns, np.float64), T1_ms=_arr(q.T1_ms, ns, np.float32),
            T2_ms=_arr(q.T2_ms, ns, np.float32), pXY=_arr(q.pXY, ns * ns, np.float32),
            RF_FA_deg=_arr(q.RF_FA_deg, q.n_rf, np.float32), RF_PH_deg=_arr(q.RF_PH_deg, q.n_rf, np.float32),
            RF_t=_arr(q.RF_t, q.n_rf, np.int32), TE_t=_arr(q.TE_t, q.n_te, np.int32),
            dephasing_deg=_arr(q.dephasing_deg, q.n_dephasing, np.float32),
            dephasing_t=_arr(q.dephasing_t, q.n_dephasing, np.int32),
            gx=_arr(q.gradientX_mTm, q.n_gradient, np.float32), gy=_arr(q.gradientY_mTm, q.n_gradient, np.float32),
            gz=_arr(q.gradientZ_mTm, q.n_gradient, np.float32), gradient_t=_arr(q.gradient_t, q.n_gradient, np.int32),
            scales=[sc[i] for i in range(n.value)], what_to_scale=what.value,
            seq_name=L.sw_config_seq_name(h).decode(),
            files={k: [L.sw_config_file(h, k.encode(), i).decode() for i in range(L.sw_config_n_files(h, k.encode()))]
                   for k in ("PHANTOM", "XYZ0", "M0")},
        )
        out["outputs"] = [L.sw_config_output(h, i).decode() for i in range(len(out["files"]["PHANTOM"]))]
        return out
    finally:
        L.sw_config_destroy(h)


def h5_write(path, name, arr):
    arr = np.ascontiguousarray(arr)
    dt = {np.dtype(np.float32): 0, np.dtype(np.uint8): 1, np.dtype(np.float64): 2}[arr.dtype]
    dims = (C.c_uint64 * max(1, arr.ndim))(*arr.shape)
    rc = load().sw_h5_write(str(path).encode(), name.encode(), dt, arr.ctypes.data_as(_P), dims, arr.ndim)
    if rc != E.SW_OK:
        raise OSError(load().sw_host_last_error().decode())


def h5_read(path, name, dtype=np.float32):
    L = load()
    dt = {np.dtype(np.float32): 0, np.dtype(np.uint8): 1, np.dtype(np.float64): 2}[np.dtype(dtype)]
    dims = (C.c_uint64 * 8)()
    nd = C.c_uint32()
    if L.sw_h5_read(str(path).encode(), name.encode(), dt, None, 0, dims, C.byref(nd)) != E.SW_OK:
        raise OSError(L.sw_host_last_error().decode())
    shape = tuple(int(dims[i]) for i in range(nd.value))
    out = np.zeros(shape, dtype)
    if L.sw_h5_read(str(path).encode(), name.encode(), dt, out.ctypes.data_as(_P), out.size, dims,
                    C.byref(nd)) != E.SW_OK:
        raise OSError(L.sw_host_last_error().decode())
    return out


def write_phantom(path, mask, field_T, fov_m):
    """A phantom file in the reference's layout (src/phantom/phantom_base.cpp:65-105)."""
    if field_T is not None:
        h5_write(path, "fieldmap", np.asarray(field_T, np.float32))
    h5_write(path, "mask", np.asarray(mask, np.uint8))
    h5_write(path, "fov", np.asarray(fov_m, np.float32).reshape(3))


def run_config(engine, path, rng=E.SW_RNG_PHILOX, precision=E.SW_PREC_F32, field_format=E.SW_FIELD_F32,
               spatial_order=1, trajectory_chunk_spins=0):
    L = load()
    o = SwRunOptions(rng, precision, field_format, spatial_order, trajectory_chunk_spins)
    rc = L.sw_run_config_opts(engine.h, str(path).encode(), C.byref(o))
    if rc != E.SW_OK:
        raise E.EngineError(f"sw_run_config: {L.sw_host_last_error().decode()} ({rc})")


def write_ini(path, sections):
    """Write an INI file from {section: {key: value}} (lists joined by spaces)."""
    lines = []
    for sec, kv in sections.items():
        lines.append(f"[{sec}]")
        for k, v in kv.items():
            if isinstance(v, (list, tuple, np.ndarray)):
                v = " ".join(repr(float(x)) if isinstance(x, float) else str(x) for x in v)
            lines.append(f"{k} = {v}")
        lines.append("")
    with open(path, "w") as fh:
        fh.write("\n".join(lines))


def generate_config(seq_name, TE_us, timestep_us, phantoms, output):
    """The `config` subcommand (src/config/config_generator.cpp:14-207)."""
    L = load()
    arr = (C.c_char_p * max(1, len(phantoms)))(*[str(x).encode() for x in phantoms])
    if L.sw_generate_config(seq_name.encode(), int(TE_us), int(timestep_us), arr, len(phantoms),
                            str(output).encode()) != E.SW_OK:
        raise ConfigError(L.sw_host_last_error().decode())


def generate_pgse(config_file, b_values, direction, start_ms, delta_ms, DELTA_ms):
    """The `dwi` subcommand (src/dwi/pgse.cpp:67-153): rewrites config_file in place."""
    L = load()
    b = (C.c_double * len(b_values))(*[float(x) for x in b_values])
    d = (C.c_float * 3)(*[float(x) for x in direction])
    if L.sw_generate_pgse(str(config_file).encode(), b, len(b_values), d, int(start_ms), int(delta_ms),
                          int(DELTA_ms)) != E.SW_OK:
        raise ConfigError(L.sw_host_last_error().decode())
