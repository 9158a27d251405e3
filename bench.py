#!/usr/bin/env python3
"""Throughput of the SpinWalk `sim` hot path on one MI355X, in spin-steps/s.

Default workload (``c4``), the FoV-scaling sweep: one 256^3 voxel, 256 um phantom
of randomly oriented cylinders (r 8 um, 4 %) generated on the device, walked at
32 FoV scales (scaled radii 1-100 um), each with GRE (TE 20 ms) and SE (TE 40 ms)
and 1e5 spins: 64 tiles whose per-tile time steps give 1 001 to 20 001 steps per
spin, 3.16e10 spin-steps in all.  One "step" of this benchmark is one complete
simulation of every tile followed by the FP64 echo reduction of each; inputs
are resident in HBM when the timed region starts.  `value` is the batched
engine at the reference's precision (FP32), all tiles in one launch of one
captured HIP graph.  Beside it, the same tiles walked as the reference launches
them (one plain FP32 walk per tile in turn) and the batched engine in mixed
precision (Kahan-compensated FP32 phase, int16 fixed-point field), with every
signal's error against the FP64 walk of the same draws.  Before the timed
region each batch runs the engine's autotuner (sw_batch_tune: chunk size,
bands, spins per lane, resident waves per SIMD, visiting-order brick) and
keeps the schedule it picks; the line reports the schedule and the autotuner's
times (`--no-tune` keeps the default schedule, `--schedule` sets one).

``--workload c1|c2|c3|c5`` selects another workload (c1: 100^3 GRE, 1e5 spins;
c2: 500^3 spin echo, 1e6 spins, one walk; c3: 1000^3 bSSFP, four phase-cycling
tiles of 1e7 spins; c5: 64 phantoms).

Prints one JSON line.  ``--gpus N`` with N > 1 is refused: this engine runs on one
GPU (spins shard trivially across GPUs; that path is not built here).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
PRECISIONS = {"f32": 0, "kahan": 1, "f64": 2}
FIELDS = {"f32": 1, "f16": 2, "bf16": 3, "q16": 4}
STEP_LOOPS = {"auto": 0, "one": 1, "two": 2}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(case, ph, budget_s, threads):
    """The oracle (a C++ restatement of the reference CPU path, with its
    per-spin mt19937 discard) on the first spins of the same workload, spread
    over `threads` host threads, until `budget_s` seconds of wall time are used."""
    from oracle import oracle as orc
    mask, field = ph.download(field=True)
    d = orc.from_engine_case(case, 0, mask=mask, field_T=field)
    n = case.n_spins
    chunk = max(1000, min(n, 2000 * threads))
    done, steps, wall, disc = 0, 0, 0.0, 0.0
    while done < n and wall < budget_s:
        b, e = done, min(n, done + chunk)
        t0 = time.perf_counter()
        r = orc.sim(d, spin_begin=b, spin_end=e, n_threads=threads, timing=True)
        wall += time.perf_counter() - t0
        disc += r["timing"]["discard_s"]
        lost = int(r["lost"][b:e].sum())
        steps += (e - b - lost) * case.steps_per_spin()
        done = e
    rate = steps / wall
    walk_only = steps / max(wall - disc, 1e-9)
    tail = ("which grows with the spin index, so the full run would be slower" if done < n else
            "over every spin of the workload")
    return dict(value=rate, unit="spin-steps/s", cores=threads, kind="port",
                sample=f"spins [0, {done}) of {n} (all {case.steps_per_spin()} steps each, "
                       f"{steps:.3g} spin-steps, {wall:.1f} s wall); the rate includes the reference's per-spin "
                       f"mt19937 discard(seed+spin), {tail}; walk-only rate {walk_only:.4g} spin-steps/s",
                walk_only_value=walk_only)


def _signals(sums):
    """|<Mxy>| per (echo, label) from FP64 echo sums, None where no spin counts."""
    out = {}
    for te in range(sums.shape[0]):
        for lab in range(sums.shape[1]):
            if sums[te, lab, 2] > 0:
                out[(te, lab)] = float(np.hypot(sums[te, lab, 0], sums[te, lab, 1]) / sums[te, lab, 2])
    return out


def _max_rel(a, ref):
    errs = [abs(a[k] - ref[k]) / ref[k] for k in ref if k in a and ref[k] > 0]
    return max(errs) if errs else None


def _brick_morton_order(XYZ0, xyz_scale, s2g, dims):
    """A spatial visiting order of spins: sorted by the Morton index of the 8^3
    voxel brick of the start position (the order the engine builds on the
    device for spatial_order = 1), computed once on the host."""
    v = np.floor(XYZ0.astype(np.float64) * float(np.float32(xyz_scale)) * np.asarray(s2g)[None, :]).astype(np.int64)
    v = np.clip(v, 0, np.asarray(dims, np.int64)[None, :] - 1) >> 3
    key = np.zeros(len(v), np.int64)
    for bit in range(10):
        for ax in range(3):
            key |= ((v[:, ax] >> bit) & 1) << (3 * bit + 2 - ax)
    return np.argsort(key, kind="stable")


def cpu_info():
    """The host CPU model and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return model, allowed


def cpu_threads():
    """Threads of the CPU baseline: every CPU this process may use, capped by
    OMP_NUM_THREADS when the machine sets it (the host's share of a shared GPU box)."""
    _, allowed = cpu_info()
    cap = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(allowed, int(cap))) if cap and cap.isdigit() else allowed


def main_tiles(args, torch, E, W, eng, stream, dev):
    """A workload of many tiles (default: the FoV-scaling sweep, 32 scales x GRE/SE
    of 1e5 spins on one 256^3 phantom).  Three ways of walking the same tiles:
      - the reference's launch structure: one plain FP32 walk per tile, in turn
        (src/sim/monte_carlo.cu:271-342 launches one cu_sim per scale);
      - the batched engine in plain FP32, every tile in one launch of one
        captured HIP graph: this is `value`;
      - the batched engine in the mixed precision of --precision/--field
        (default Kahan-compensated FP32 phase, int16 fixed-point field).
    The signal errors are against the FP64 batch of the same draws."""
    wl = W.WORKLOADS[args.workload]()
    prec = PRECISIONS[args.precision or "kahan"]
    fmt = FIELDS[args.field or "q16"]
    log(f"[bench] workload {wl['name']}: generating phantoms on the device")
    phs, cases_, pidx, vf = W.build_tiles(eng, wl, formats=tuple(sorted({E.SW_FIELD_F32, fmt})))
    tiles, seq_tiles = [], []
    inputs = {}        # one device copy of each distinct start array, shared by the tiles that use it

    def resident(a):
        key = (a.shape, a.dtype.str, hash(a.tobytes()))
        if key not in inputs or not np.array_equal(inputs[key][0], a):
            inputs[key] = (a, torch.from_numpy(a).to(dev))
        return inputs[key][1]

    z = lambda *sh, dt=torch.float32: torch.zeros(sh, dtype=dt, device=dev)
    for case, k in zip(cases_, pidx):
        p, xs, gs = case.scale_params(case.scales[0])
        n = case.n_spins
        tiles.append(dict(phantom=phs[k], sequence=eng.sequence(p, case.tables), params=p, xyz_scale=xs,
                          gradient_scale=gs, spin_offset=0, n_local=n, spatial_order=args.order,
                          M0=resident(case.M0), XYZ0=resident(case.XYZ0),
                          M1=z(n, case.n_te, 3), T=z(n, case.n_te, dt=torch.uint8), XYZ1=z(n, case.traj_len, 3),
                          echo_sums=z(case.n_te, p.n_substrate, 3, dt=torch.float64), steps=case.steps_per_spin()))
        # The per-tile baseline visits the spins in the same spatial order as the
        # batch, with the order built here once (the batch builds its orders once,
        # at creation), so the timed region holds the walks alone.
        X0 = case.XYZ0
        if args.order:
            s2g = [p.phantom_size[i] / p.fov[i] for i in range(3)]
            X0 = np.ascontiguousarray(X0[_brick_morton_order(X0, xs, s2g, list(p.phantom_size))])
        seq_tiles.append(dict(tiles[-1], XYZ0=torch.from_numpy(X0).to(dev)))
    total_steps = sum(t["n_local"] * t["steps"] for t in tiles)

    def sequential():
        for t in seq_tiles:
            eng.walk_device(t["phantom"], t["sequence"], t["params"], t["M0"], t["XYZ0"], t["M1"], t["XYZ1"], t["T"],
                            xyz_scale=t["xyz_scale"], gradient_scale=t["gradient_scale"], precision=0,
                            field_format=E.SW_FIELD_F32, spatial_order=0, stream=stream, stats=False,
                            step_loop=STEP_LOOPS[args.step_loop])
            eng.echo_reduce(t["M1"], t["T"], t["params"].n_substrate, t["echo_sums"], stream=stream)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / args.steps

    def make_batch(precision, field):
        dt = torch.float64 if precision == 2 else torch.float32
        items = []
        for t in tiles:
            it = dict(t)
            it.update(precision=precision, field_format=field, M1=torch.zeros(tuple(t["M1"].shape), dtype=dt, device=dev),
                      T=torch.zeros_like(t["T"]), XYZ1=torch.zeros_like(t["XYZ1"]),
                      echo_sums=torch.zeros_like(t["echo_sums"]))
            items.append(it)
        return eng.batch(items), items

    def run_batch(precision, field, label, schedule):
        b, items = make_batch(precision, field)
        tuning = None
        if schedule:                       # a given schedule (a profiled run repeats a tuned one)
            b.set_schedule(*[int(x) for x in schedule.split(",")])
            log(f"[bench] schedule ({label}): {b.schedule}")
        elif not args.no_tune:             # the autotuner, before the timed region
            tuning = b.tune(stream=stream)
            log(f"[bench] autotuner ({label}): {tuning['schedule']}, {tuning['best_ms']:.2f} ms per run "
                f"(default schedule {tuning['first_ms']:.2f} ms)")
        st = {}

        def batched():
            st.update(b.run(stream=stream, stats=True)["total"])

        secs = timed(batched)
        g = b.graph_info()
        sig = [_signals(it["echo_sums"].cpu().numpy()) for it in items]
        log(f"[bench] batched {label}: {secs * 1e3:.2f} ms, {total_steps / secs:.4g} spin-steps/s, "
            f"kernel {st['kernel_ms']:.2f} ms, lost {st['n_lost']}, HIP graph {g}")
        return dict(secs=secs, stats=dict(st), graph=g, signals=sig, schedule=b.schedule, tuning=tuning,
                    info=b.info(), batch=b, items=items)

    seq_s = timed(sequential)
    log(f"[bench] one plain FP32 walk per tile: {seq_s * 1e3:.2f} ms, {total_steps / seq_s:.4g} spin-steps/s")
    f32 = run_batch(0, E.SW_FIELD_F32, "FP32", args.schedule)
    mixed = run_batch(prec, fmt, f"{args.precision or 'kahan'} + {args.field or 'q16'} field", args.schedule_mixed)
    for r in (f32, mixed):
        r["batch"].close()
        del r["items"]
    b64, items64 = make_batch(2, E.SW_FIELD_F32)
    b64.run(stream=stream, stats=True)
    ref_sig = [_signals(it["echo_sums"].cpu().numpy()) for it in items64]
    b64.close()
    del items64

    def errors(sig):
        rows = []
        for i, (a, r) in enumerate(zip(sig, ref_sig)):
            for key in r:
                if key in a and r[key] > 0:
                    rows.append(dict(tile=i, seq=wl["tiles"][i]["seq"], scale=float(wl["tiles"][i]["scale"]),
                                     echo=key[0], label=key[1], signal=a[key], f64=r[key],
                                     abs_err=abs(a[key] - r[key]), rel_err=abs(a[key] - r[key]) / r[key]))
        worst = sorted(rows, key=lambda x: -x["rel_err"])[:5]
        return dict(max_rel=worst[0]["rel_err"] if worst else None,
                    max_abs=max((w["abs_err"] for w in rows), default=None), worst=worst)

    if args.copy_mib > 0:              # profiler calibration: one device copy of known size
        a = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=dev)
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()
        del a
    err32, errm = errors(f32["signals"]), errors(mixed["signals"])
    log(f"[bench] max relative signal error against FP64: FP32 batch {err32['max_rel']:.2e}, "
        f"mixed batch {errm['max_rel']:.2e}")
    for w in errm["worst"]:
        log(f"[bench]   mixed: tile {w['tile']} {w['seq']} scale {w['scale']:.3g} label {w['label']}: signal "
            f"{w['signal']:.6g} vs FP64 {w['f64']:.6g}, abs {w['abs_err']:.2e}, rel {w['rel_err']:.2e}")

    def roofline(res, field, tag):
        bpe = 5 if field == E.SW_FIELD_F32 else 3           # u8 mask + field bytes per spin-step
        kernel_s = res["stats"]["kernel_ms"] * 1e-3
        achieved = total_steps * bpe / kernel_s / 1e9
        traffic, traffic_note = None, "no rocprof measurement for this workload and form in profiles/"
        tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{tag}.json")
        if os.path.exists(tpath):
            with open(tpath) as fh:
                t = json.load(fh)
            # The line traffic depends on the schedule: the file holds one measurement per schedule,
            # and one is reported only for the schedule it was measured under.
            runs = t.get("measurements", [t])
            hit = [m for m in runs if m.get("schedule") == res["schedule"]]
            if hit:
                traffic, traffic_note = hit[0].get("hbm_bytes_per_launch"), f"measured, {os.path.relpath(tpath, ROOT)}"
            else:
                traffic_note = (f"{os.path.relpath(tpath, ROOT)} was measured under schedules "
                                f"{[m.get('schedule') for m in runs]}")
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_note": traffic_note,
                "traffic_rate": None if traffic is None else traffic / kernel_s / 1e9,
                "bytes_per_spin_step": bpe, "kernel_ms": res["stats"]["kernel_ms"], "kernel": "walk_batch_kernel"}

    mixed_tag = f"{args.precision or 'kahan'}_{args.field or 'q16'}"
    out = {
        "metric": "spin-steps/s", "value": total_steps / f32["secs"], "unit": "spin-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": f32["secs"] * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": wl["name"],
                   "phantom": ("x".join(str(d) for d in wl["phantom"]["dims"]) +
                               f" voxels, {wl['phantom']['fov_um'][0]:g} um, random cylinders")
                   if "phantom" in wl else f"{len(phs)} phantoms of cylinders, spheres and packed axons",
                   "tiles": len(tiles), "spins_per_tile": tiles[0]["n_local"], "spin_steps": total_steps,
                   "precision": "f32", "field": "f32",
                   "mode": "batched" + (", one HIP graph" if f32["graph"]["uses_graph"] else ", stream launches"),
                   "hip_graph": f32["graph"], "walks": f32["info"], "schedule": f32["schedule"],
                   "autotuned": f32["tuning"] is not None, "rng": "philox4x32-10"},
        "roofline": roofline(f32, E.SW_FIELD_F32, "f32_f32"),
        "signal_error_vs_f64": err32,
        "sequential_f32": {"value": total_steps / seq_s, "ms_per_step": seq_s * 1e3,
                           "mode": "one plain FP32 walk per tile in turn (the reference's launch structure)"},
        "mixed_batched": {"value": total_steps / mixed["secs"], "ms_per_step": mixed["secs"] * 1e3,
                          "precision": args.precision or "kahan", "field": args.field or "q16",
                          "dtype": {0: "f32", 1: "f32-kahan", 2: "f64"}[prec] + f"+{args.field or 'q16'}-field",
                          "hip_graph": mixed["graph"], "schedule": mixed["schedule"],
                          "roofline": roofline(mixed, fmt, mixed_tag), "signal_error_vs_f64": errm,
                          "speedup_vs_sequential_f32": seq_s / mixed["secs"],
                          "speedup_vs_batched_f32": f32["secs"] / mixed["secs"]},
    }
    for key, r in (("autotuner", f32), ("autotuner_mixed", mixed)):
        if r["tuning"] is not None:
            out[key] = r["tuning"]
    if not args.no_cpu:
        threads = cpu_threads()
        model, allowed = cpu_info()
        log(f"[bench] CPU baseline: oracle on {threads} threads ({model}, {allowed} CPUs allowed), "
            f"{args.cpu_seconds} s budget, first tile")
        out["cpu_baseline"] = cpu_baseline(cases_[0], phs[pidx[0]], args.cpu_seconds, threads)
        out["cpu_baseline"].update(cpu_model=model, cpus_allowed=allowed, os_cpu_count=os.cpu_count())
    print(json.dumps(out), flush=True)
    return 0


# Working sets of the probe: the L2 of one XCD, under the 256 MB Infinity Cache,
# and the FP32 voxel records of the 500^3 and 1000^3 workloads (bricks of 16
# eight-byte records: 1.0e9 and 8.0e9 bytes).
PROBE_SIZES = ((4 << 20, "4 MiB"), (192 << 20, "192 MiB"), (1_000_000_000, "1.0 GB (500^3 FP32 records)"),
               (8_000_000_000, "8.0 GB (1000^3 FP32 records)"))


def main_probe(args, torch, eng, stream, dev):
    """Random-line read rates and the streaming copy rate on this GPU: the
    ceilings that the walk's gather traffic is compared with."""
    n_lanes, n_reads = 1 << 20, 128
    sink = torch.empty(n_lanes, dtype=torch.int64, device=dev)
    rows = []
    for nbytes, name in PROBE_SIZES:
        buf = torch.empty(nbytes // 8, dtype=torch.int64, device=dev)
        torch.cuda.synchronize()
        for _ in range(2):
            eng.probe_gather(buf, n_lanes, n_reads, sink, stream=stream)
        ms = sorted(eng.probe_gather(buf, n_lanes, n_reads, sink, seed=k + 2, stream=stream) for k in range(5))[2]
        lines = n_lanes * n_reads
        rows.append(dict(working_set=name, bytes=nbytes, ms=ms, lines_per_s=lines / (ms * 1e-3),
                         line_GBps=lines * 128 / (ms * 1e-3) / 1e9))
        log(f"[probe] random 128-byte lines over {name}: {rows[-1]['line_GBps']:.0f} GB/s of lines "
            f"({rows[-1]['lines_per_s']:.3g} lines/s)")
        del buf
    a = torch.empty(1 << 28, dtype=torch.float32, device=dev)          # 1 GiB
    b = torch.empty_like(a)
    b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(10):
        b.copy_(a)
    e1.record(stream)
    torch.cuda.synchronize()
    copy_GBps = 10 * 2 * a.numel() * 4 / (e0.elapsed_time(e1) * 1e-3) / 1e9
    log(f"[probe] 1 GiB device copy: {copy_GBps:.0f} GB/s (read + write)")
    print(json.dumps({"probe": "memory ceilings", "random_line_reads": rows, "copy_GBps": copy_GBps,
                      "hbm_peak_GBps": HBM_PEAK_GBS, "n_lanes": n_lanes, "reads_per_lane": n_reads}), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c4", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--precision", default=None, choices=sorted(PRECISIONS),
                    help="single-walk workloads: the walk's precision (default f32); tile workloads: the "
                         "mixed batch's (default kahan)")
    ap.add_argument("--field", default=None, choices=sorted(FIELDS),
                    help="single-walk workloads: the field format (default f32); tile workloads: the mixed "
                         "batch's (default q16)")
    ap.add_argument("--order", type=int, default=1, help="spatial visiting order (1) or spin order (0)")
    ap.add_argument("--step-loop", default="auto", choices=sorted(STEP_LOOPS),
                    help="step loop of the single walks: one or two attempts per trip, or auto (by record size)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="wall-time budget of the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--host-io", action="store_true",
                    help="after the timed region, time one host-buffer call (sw_walk_host: uploads, walk, downloads)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the timed region, run one device copy of this size (profiler calibration)")
    ap.add_argument("--no-tune", action="store_true",
                    help="tile workloads: keep the default schedule instead of running the batch autotuner "
                         "(sw_batch_tune) before the timed region")
    ap.add_argument("--schedule", default=None,
                    help="tile workloads: the FP32 batch's schedule, 'chunk,bands_shared,bands_mixed,"
                         "spins_per_lane,waves_per_simd,order_brick' (no autotuning)")
    ap.add_argument("--schedule-mixed", default=None, help="the same for the mixed-precision batch")
    ap.add_argument("--probe", action="store_true",
                    help="instead of the walk: the memory ceilings of the gather (random 128-byte line reads at "
                         "several working-set sizes) and of a streaming copy")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; --gpus must be 1", file=sys.stderr)
        return 2

    log("[bench] importing torch")
    import torch
    from spinwalk_amd import engine as E
    from spinwalk_amd import workloads as W

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    # A stream of our own (not the legacy default stream): torch allocates and fills
    # on it, the engine launches on it, its HIP events time the walk on it, and
    # torch's readbacks are ordered after.
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    eng = E.Engine(0)
    if args.probe:
        return main_probe(args, torch, eng, stream, dev)
    if "tiles" in W.WORKLOADS[args.workload]():
        return main_tiles(args, torch, E, W, eng, stream, dev)
    wl = W.WORKLOADS[args.workload]()
    args.precision = args.precision or "f32"
    args.field = args.field or "f32"
    fmt = FIELDS[args.field]
    prec = PRECISIONS[args.precision]
    log(f"[bench] workload {wl['name']}: generating phantom on the device")
    t0 = time.perf_counter()
    ph, case, vf = W.build(eng, wl, formats=(E.SW_FIELD_F32, fmt))
    log(f"[bench] phantom ready in {time.perf_counter() - t0:.1f} s, analytic volume fractions {vf}")
    p = case.params
    seq = eng.sequence(p, case.tables)
    n, n_te = case.n_spins, case.n_te
    M0 = torch.from_numpy(case.M0).to(dev)
    XYZ0 = torch.from_numpy(case.XYZ0).to(dev)
    M1 = torch.zeros((n, n_te, 3), dtype=torch.float64 if prec == 2 else torch.float32, device=dev)
    T = torch.zeros((n, n_te), dtype=torch.uint8, device=dev)
    XYZ1 = torch.zeros((n, case.traj_len, 3), dtype=torch.float32, device=dev)
    sums = torch.zeros((n_te, p.n_substrate, 3), dtype=torch.float64, device=dev)

    def one_step():
        st = eng.walk_device(ph, seq, p, M0, XYZ0, M1, XYZ1, T, rng=E.SW_RNG_PHILOX, precision=prec,
                             field_format=fmt, spatial_order=args.order, stream=stream, stats=True,
                             step_loop=STEP_LOOPS[args.step_loop])
        eng.echo_reduce(M1, T, p.n_substrate, sums, stream=stream)
        return st

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    kms, steps, lost = [], 0, 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = one_step()
        kms.append(st["kernel_ms"])
        steps += st["n_steps"]
        lost += st["n_lost"]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    if args.host_io:
        hM1 = np.zeros(tuple(M1.shape), np.float64 if prec == 2 else np.float32)
        hT = np.zeros(tuple(T.shape), np.uint8)
        hX = np.zeros(tuple(XYZ1.shape), np.float32)
        t1 = time.perf_counter()
        hst = eng.walk_host(ph, seq, p, case.M0, case.XYZ0, hM1, hX, hT, rng=E.SW_RNG_PHILOX, precision=prec,
                            field_format=fmt, spatial_order=args.order)
        ht = time.perf_counter() - t1
        log(f"[bench] host-buffer call (sw_walk_host, PCIe transfers included): {ht * 1e3:.3f} ms, "
            f"{hst['n_steps'] / ht:.4g} spin-steps/s (kernel {hst['kernel_ms']:.3f} ms)")
    if args.copy_mib > 0:
        a = torch.empty(args.copy_mib << 18, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
    value = steps / wall
    bpe = 5 if fmt == E.SW_FIELD_F32 else 3           # u8 mask + field bytes per spin-step
    if prec == 2:
        bpe = 5
    steps_per_launch = steps / max(1, args.steps)
    kernel_s = float(np.mean(kms)) * 1e-3
    achieved = steps_per_launch * bpe / kernel_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{args.precision}_{args.field}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    s = sums.cpu().numpy()
    # The same walk in FP64 (same draws, so the same trajectories): the signal error
    # of the measured precision against it, per echo and label.
    M1d = torch.zeros((n, n_te, 3), dtype=torch.float64, device=dev)
    sums_d = torch.zeros_like(sums)
    eng.walk_device(ph, seq, p, M0, XYZ0, M1d, XYZ1, T, rng=E.SW_RNG_PHILOX, precision=PRECISIONS["f64"],
                    field_format=E.SW_FIELD_F32, spatial_order=args.order, stream=stream, stats=True)
    eng.echo_reduce(M1d, T, p.n_substrate, sums_d, stream=stream)
    sd = sums_d.cpu().numpy()
    del M1d
    errs = []
    for te in range(n_te):
        for lab in range(p.n_substrate):
            if s[te, lab, 2] > 0 and sd[te, lab, 2] > 0:
                mag = float(np.hypot(s[te, lab, 0], s[te, lab, 1]) / s[te, lab, 2])
                ref = float(np.hypot(sd[te, lab, 0], sd[te, lab, 1]) / sd[te, lab, 2])
                errs.append(dict(echo=te, label=lab, signal=mag, signal_f64=ref, rel_err=abs(mag - ref) / ref))
                log(f"[bench] echo {te} label {lab}: |<Mxy>| = {mag:.6f} over {int(s[te, lab, 2])} spins, "
                    f"FP64 {ref:.6f}, rel. error {abs(mag - ref) / ref:.2e}")
    log(f"[bench] walk kernel {np.mean(kms):.3f} ms (min {np.min(kms):.3f}), {steps_per_launch:.4g} accepted "
        f"spin-steps/launch, lost {lost}")
    out = {
        "metric": "spin-steps/s",
        "value": value,
        "unit": "spin-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {0: "f32", 1: "f32-kahan", 2: "f64"}[prec] + ("" if fmt == 1 else f"+{args.field}-field"),
        "data": "synthetic",
        "config": {"workload": wl["name"], "phantom": "500^3 voxels, 500 um, random cylinders r5 4% + spheres r3 2%"
                   if args.workload == "c2" else str(wl["phantom"]["dims"]), "spins": n,
                   "steps_per_spin": case.steps_per_spin(), "sequence": "SE 90/180 TE 40 ms dt 20 us"
                   if args.workload == "c2" else "GRE", "precision": args.precision, "field": args.field,
                   "spatial_order": args.order, "rng": "philox4x32-10", "step_loop": args.step_loop},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     # the measured line bytes of one launch (rocprof FETCH_SIZE + WRITE_SIZE, calibrated)
                     # over this run's kernel time: what the memory system moved, in GB/s
                     "traffic_rate": None if traffic is None else traffic / kernel_s / 1e9,
                     "bytes_per_spin_step": bpe, "kernel_ms": float(np.mean(kms))},
    }
    out["signal_error_vs_f64"] = {"max_rel": max(e["rel_err"] for e in errs) if errs else None, "per_echo_label": errs}
    if not args.no_cpu:
        threads = cpu_threads()
        model, allowed = cpu_info()
        log(f"[bench] CPU baseline: oracle on {threads} threads ({model}, {allowed} CPUs allowed), "
            f"{args.cpu_seconds} s budget")
        out["cpu_baseline"] = cpu_baseline(case, ph, args.cpu_seconds, threads)
        out["cpu_baseline"].update(cpu_model=model, cpus_allowed=allowed, os_cpu_count=os.cpu_count())
    print(json.dumps(out), flush=True)
    ph.close()
    seq.close()
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
