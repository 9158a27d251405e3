"""Parity and invariants at the benchmark workloads' full sizes.

100^3 GRE (100^3 voxels, 1e5 spins, 401 steps, impermeable cylinders): the one
workload the oracle walks whole in a test's time, so every spin is checked:
  - all 1e5 spins replayed with the reference CPU path's own draw streams:
    trajectories, labels and lost spins bit-identical to the oracle, M within
    1e-5 relative, the FP64 path within 1e-12 of the oracle's FP64 twin;
  - all 1e5 spins with production draws: each label's signal |<Mxy>| agrees
    with the oracle's (the reference's own draws) within five standard errors
    of the difference of two independent Monte-Carlo estimates.

For the larger workloads the oracle cannot walk every spin in a test's time, so
the walk is held to what does not depend on size, and to the oracle on a sample:

500^3 spin echo (500^3 voxels generated on the device, 1e6 spins, 2001 steps,
three substrates, permeability 0.5):
  - spins [0, 2048) of the full case, replayed with the reference CPU path's own
    draw streams on the full phantom: trajectories, labels and lost spins are
    bit-identical to the oracle's (a restatement of src/sim/kernels.cu:72-233), M
    within 1e-5 relative, the FP64 path within 1e-12;
  - production draws, all 1e6 spins: every spin takes every step (accepted steps
    = n_spins x steps) and none is lost; labels are the phantom's; positions stay
    in [0, fov) (kernels.cu:133-136); the FP32, Kahan and Kahan + 16-bit field
    walks take the FP64 walk's trajectories and labels exactly, and their signals
    agree with it within 1e-4 relative per echo and label; the device's FP64 echo
    sums equal numpy's float64 sums of the downloaded M and T.

FoV sweep (64 tiles of 1e5 spins, 1 001-20 001 steps) and 64 phantoms (175
tiles):
  - a sample of spins of single tiles, replayed on the full phantom, bit-exact
    against the oracle as above: GRE at FoV scale 2.44 and SE at FoV scale 0.41
    (8 001 steps) of the sweep; PGSE along z at b = 3000 s/mm^2 (a gradient scale)
    on the 512^3 packed-axon and the 384^3 sphere phantoms;
  - the mixed-precision batch (Kahan phase, int16 fixed-point field) keeps every
    tile's signals within 1e-4 relative of the FP64 batch of the same draws.

1000^3 bSSFP (four phase-cycling tiles of 1e7 spins, 10 100 steps each):
  - spins [0, 1024) of the phase-cycling tile at scale 1, replayed on the full
    phantom (its 8 GB of FP32 records take the 64-bit index path), through 100
    dummy scans: bit-exact against the oracle as above;
  - the four tiles fuse into one walk, and each tile's M, T and positions are
    bit-identical to its own separate walk.
"""
import numpy as np
import pytest

from tests.test_gpu_replay import _assert_parity
from oracle import oracle as orc
from spinwalk_amd import engine as E
from spinwalk_amd import workloads as W

pytestmark = pytest.mark.gpu

SIGNAL_RTOL = 1e-4        # mixed precision against FP64, per echo and label


@pytest.fixture(scope="module")
def c2(engine):
    ph, case, _ = W.build(engine, W.c2(), formats=(E.SW_FIELD_F32, E.SW_FIELD_Q16))
    yield ph, case
    ph.close()


def _signals(sums):
    out = {}
    for te in range(sums.shape[0]):
        for lab in range(sums.shape[1]):
            if sums[te, lab, 2] > 0:
                out[(te, lab)] = float(np.hypot(sums[te, lab, 0], sums[te, lab, 1]) / sums[te, lab, 2])
    return out


def _max_rel(a, ref):
    assert a.keys() == ref.keys()
    return max(abs(a[k] - ref[k]) / ref[k] for k in ref if ref[k] > 0)


@pytest.fixture(scope="module")
def c1(engine):
    ph, case, _ = W.build(engine, W.c1(), formats=(E.SW_FIELD_F32,))
    mask, field = ph.download(field=True)
    d = orc.from_engine_case(case, 0, mask=mask, field_T=field)
    yield ph, case, d
    ph.close()


def test_c1_every_spin_replayed_bit_exact(engine, c1):
    ph, case, d = c1
    n, steps = case.n_spins, case.steps_per_spin()
    assert n == 100000 and steps == 401 and case.n_te == 1
    # impermeable walls: a rejected attempt draws again, so room for three attempts a step
    normals, uniforms = E.replay_fill(case.params.seed, 0, n, 9 * steps + 96, 3 * steps + 64, n_threads=16)
    p, xs, gs = case.scale_params(case.scales[0])
    seq = engine.sequence(p, case.tables)
    for prec, f64 in ((E.SW_PREC_F32, False), (E.SW_PREC_F64, True)):
        print(f"[c1] replay all {n} spins, precision {prec}", flush=True)
        o = orc.sim(d, n_threads=16, f64=f64)
        M1 = np.zeros((n, case.n_te, 3), np.float64 if f64 else np.float32)
        T = np.zeros((n, case.n_te), np.uint8)
        XYZ1 = np.zeros((n, case.traj_len, 3), np.float32)
        st = engine.walk_host(ph, seq, p, case.M0, case.XYZ0, M1, XYZ1, T, xyz_scale=xs, gradient_scale=gs,
                              rng=E.SW_RNG_REPLAY, precision=prec, field_format=E.SW_FIELD_F32,
                              replay=(normals, uniforms), spatial_order=1)
        _assert_parity(dict(M1=M1, T=T, XYZ1=XYZ1, stats=st), o, f64=f64)
        if o["n_lost"] == 0:
            assert st["n_steps"] == n * steps
        assert st["n_attempts"] > st["n_steps"]           # the walls reject attempts
    seq.close()


def _label_signals(M1, T):
    """Per label: |<Mxy>| and the standard error of it (the spread of the spins'
    Mxy along the mean direction over the square root of their number)."""
    out = {}
    z = M1[:, 0, 0].astype(np.float64) + 1j * M1[:, 0, 1].astype(np.float64)
    for lab in np.unique(T[:, 0]):
        zz = z[T[:, 0] == lab]
        m = zz.mean()
        proj = (zz * np.conj(m) / abs(m)).real
        out[int(lab)] = (abs(m), proj.std(ddof=1) / np.sqrt(zz.size), zz.size)
    return out


def test_c1_production_signals_within_monte_carlo_error_of_oracle(engine, c1):
    ph, case, d = c1
    n = case.n_spins
    o = orc.sim(d, n_threads=16)
    p, xs, gs = case.scale_params(case.scales[0])
    seq = engine.sequence(p, case.tables)
    M1 = np.zeros((n, case.n_te, 3), np.float32)
    T = np.zeros((n, case.n_te), np.uint8)
    XYZ1 = np.zeros((n, case.traj_len, 3), np.float32)
    st = engine.walk_host(ph, seq, p, case.M0, case.XYZ0, M1, XYZ1, T, xyz_scale=xs, gradient_scale=gs,
                          rng=E.SW_RNG_PHILOX, precision=E.SW_PREC_F32, field_format=E.SW_FIELD_F32,
                          spatial_order=1)
    seq.close()
    assert st["n_lost"] == 0 and o["n_lost"] == 0
    assert st["n_steps"] == n * case.steps_per_spin()
    g, r = _label_signals(M1, T), _label_signals(o["M1"], o["T"])
    assert g.keys() == r.keys() == {0, 1}
    for lab in r:
        # impermeable walls keep every spin in its start label
        assert g[lab][2] == r[lab][2] == int((T[:, 0] == lab).sum())
        se = np.hypot(g[lab][1], r[lab][1])
        assert abs(g[lab][0] - r[lab][0]) < 5 * se, (lab, g[lab], r[lab])
        assert se < 0.05 * r[lab][0]                      # the test resolves each signal to 5 % or better


def test_c2_replay_sample_on_the_full_phantom(engine, c2):
    ph, case = c2
    n = 2048
    steps = case.steps_per_spin()
    assert steps == 2001 and case.n_spins == 1000000
    mask, field = ph.download(field=True)
    d = orc.from_engine_case(case, 0, mask=mask, field_T=field)
    normals, uniforms = E.replay_fill(case.params.seed, 0, n, 6 * steps + 96, 2 * steps + 64)
    p, xs, gs = case.scale_params(case.scales[0])
    seq = engine.sequence(p, case.tables)
    for prec, f64 in ((E.SW_PREC_F32, False), (E.SW_PREC_F64, True)):
        o = orc.sim(d, spin_begin=0, spin_end=n, n_threads=8, f64=f64)
        dt = np.float64 if f64 else np.float32
        M1 = np.zeros((n, case.n_te, 3), dt)
        T = np.zeros((n, case.n_te), np.uint8)
        XYZ1 = np.zeros((n, case.traj_len, 3), np.float32)
        st = engine.walk_host(ph, seq, p, np.ascontiguousarray(case.M0[:n]), np.ascontiguousarray(case.XYZ0[:n]),
                              M1, XYZ1, T, xyz_scale=xs, gradient_scale=gs, rng=E.SW_RNG_REPLAY, precision=prec,
                              field_format=E.SW_FIELD_F32, replay=(normals, uniforms), spatial_order=1)
        g = dict(M1=M1, T=T, XYZ1=XYZ1, stats=st)
        _assert_parity(g, o, f64=f64, sl=slice(0, n))
        assert st["n_steps"] == n * steps
    seq.close()


def test_c2_full_size_production_invariants(engine, c2):
    import torch
    ph, case = c2
    dev = torch.device("cuda:0")
    p, xs, gs = case.scale_params(case.scales[0])
    seq = engine.sequence(p, case.tables)
    n, n_te, steps = case.n_spins, case.n_te, case.steps_per_spin()
    M0 = torch.from_numpy(case.M0).to(dev)
    XYZ0 = torch.from_numpy(case.XYZ0).to(dev)
    out = {}
    for name, prec, ff in (("f64", E.SW_PREC_F64, E.SW_FIELD_F32), ("f32", E.SW_PREC_F32, E.SW_FIELD_F32),
                           ("kahan", E.SW_PREC_F32_KAHAN, E.SW_FIELD_F32),
                           ("kahan_q16", E.SW_PREC_F32_KAHAN, E.SW_FIELD_Q16)):
        M1 = torch.zeros((n, n_te, 3), dtype=torch.float64 if prec == E.SW_PREC_F64 else torch.float32, device=dev)
        T = torch.zeros((n, n_te), dtype=torch.uint8, device=dev)
        XYZ1 = torch.zeros((n, case.traj_len, 3), dtype=torch.float32, device=dev)
        sums = torch.zeros((n_te, p.n_substrate, 3), dtype=torch.float64, device=dev)
        st = engine.walk_device(ph, seq, p, M0, XYZ0, M1, XYZ1, T, rng=E.SW_RNG_PHILOX, precision=prec,
                                field_format=ff, spatial_order=1, stats=True)
        engine.echo_reduce(M1, T, p.n_substrate, sums)
        torch.cuda.synchronize()
        assert st["n_lost"] == 0 and st["n_steps"] == n * steps, st
        out[name] = dict(M1=M1.cpu().numpy(), T=T.cpu().numpy(), XYZ1=XYZ1.cpu().numpy(), sums=sums.cpu().numpy())
    ref = out["f64"]
    n_labels, _ = ph.info()
    assert ref["T"].max() < n_labels
    x = ref["XYZ1"].astype(np.float64)
    fov = np.asarray([p.fov[i] for i in range(3)], np.float64)
    assert (x >= 0).all() and (x < fov).all()
    # the FP64 echo sums: numpy float64 over the downloaded M and T
    M = ref["M1"]
    for te in range(n_te):
        for lab in range(p.n_substrate):
            sel = ref["T"][:, te] == lab
            want = [M[sel, te, 0].sum(), M[sel, te, 1].sum(), float(sel.sum())]
            np.testing.assert_allclose(ref["sums"][te, lab], want, rtol=1e-12, atol=1e-9)
    s_ref = _signals(ref["sums"])
    assert len(s_ref) == 3
    for name in ("f32", "kahan", "kahan_q16"):
        np.testing.assert_array_equal(out[name]["T"], ref["T"])
        np.testing.assert_array_equal(out[name]["XYZ1"], ref["XYZ1"])
        err = _max_rel(_signals(out[name]["sums"]), s_ref)
        assert err < SIGNAL_RTOL, (name, err)
    seq.close()


def _tile_items(engine, torch, wl, precision, field_format, formats):
    dev = torch.device("cuda:0")
    phs, cases_, pidx, _ = W.build_tiles(engine, wl, formats=formats)
    inputs, items = {}, []

    def resident(a):                   # tiles that share start arrays share one device copy
        key = (a.shape, hash(a.tobytes()))
        if key not in inputs:
            inputs[key] = torch.from_numpy(a).to(dev)
        return inputs[key]

    for case, k in zip(cases_, pidx):
        p, xs, gs = case.scale_params(case.scales[0])
        n = case.n_spins
        items.append(dict(phantom=phs[k], sequence=engine.sequence(p, case.tables), params=p, xyz_scale=xs,
                          gradient_scale=gs, n_local=n, spatial_order=1, precision=precision,
                          field_format=field_format, M0=resident(case.M0), XYZ0=resident(case.XYZ0),
                          M1=torch.zeros((n, case.n_te, 3), device=dev,
                                         dtype=torch.float64 if precision == E.SW_PREC_F64 else torch.float32),
                          T=torch.zeros((n, case.n_te), dtype=torch.uint8, device=dev),
                          XYZ1=torch.zeros((n, case.traj_len, 3), dtype=torch.float32, device=dev),
                          echo_sums=torch.zeros((case.n_te, p.n_substrate, 3), dtype=torch.float64, device=dev),
                          steps=case.steps_per_spin()))
    return phs, items


def _close(phs, *item_lists):
    for items in item_lists:
        for it in items:
            it["sequence"].close()
    for ph in phs:
        ph.close()


@pytest.mark.parametrize("workload", ["c4", "c5"])
def test_mixed_precision_batch_within_1e4_of_fp64(engine, workload):
    import torch
    wl = W.WORKLOADS[workload]()
    fm = (E.SW_FIELD_F32, E.SW_FIELD_Q16)
    phs, mixed = _tile_items(engine, torch, wl, E.SW_PREC_F32_KAHAN, E.SW_FIELD_Q16, fm)
    ref = []
    for it in mixed:                   # the FP64 batch walks the same phantoms and draws
        r = dict(it)
        r.update(precision=E.SW_PREC_F64, field_format=E.SW_FIELD_F32,
                 M1=torch.zeros(tuple(it["M1"].shape), dtype=torch.float64, device="cuda"),
                 T=torch.zeros_like(it["T"]), XYZ1=torch.zeros_like(it["XYZ1"]),
                 echo_sums=torch.zeros_like(it["echo_sums"]))
        ref.append(r)
    bm, br = engine.batch(mixed), engine.batch(ref)
    sm, sr = bm.run(stats=True), br.run(stats=True)
    torch.cuda.synchronize()
    assert sm["total"]["n_lost"] == 0 and sr["total"]["n_lost"] == 0
    assert sm["total"]["n_steps"] == sr["total"]["n_steps"] == sum(it["n_local"] * it["steps"] for it in mixed)
    worst = 0.0
    for i, (a, b) in enumerate(zip(mixed, ref)):
        assert torch.equal(a["T"], b["T"]) and torch.equal(a["XYZ1"], b["XYZ1"]), i
        err = _max_rel(_signals(a["echo_sums"].cpu().numpy()), _signals(b["echo_sums"].cpu().numpy()))
        worst = max(worst, err)
    assert worst < SIGNAL_RTOL, worst
    bm.close()
    br.close()
    _close(phs, mixed)


@pytest.mark.parametrize("workload,tile,n", [("c4", 20, 1024), ("c4", 40, 512), ("c5", 47, 1024),
                                             ("c5", 100, 1024)])
def test_tile_replay_sample_on_the_full_phantom(engine, workload, tile, n):
    wl = W.WORKLOADS[workload]()
    t = dict(wl["tiles"][tile])
    if "phantoms" in wl:                       # only the phantom this tile walks
        wl = dict(wl, phantoms=[wl["phantoms"][t["phantom"]]])
        t["phantom"] = 0
    phs, cases_, _, _ = W.build_tiles(engine, dict(wl, tiles=[t]), formats=(E.SW_FIELD_F32,))
    ph, case = phs[0], cases_[0]
    steps = case.steps_per_spin()
    mask, field = ph.download(field=True)
    d = orc.from_engine_case(case, 0, mask=mask, field_T=field)
    del mask, field
    normals, uniforms = E.replay_fill(case.params.seed, 0, n, 9 * steps + 96, 3 * steps + 64)
    p, xs, gs = case.scale_params(case.scales[0])
    seq = engine.sequence(p, case.tables)
    o = orc.sim(d, spin_begin=0, spin_end=n, n_threads=8)
    M1 = np.zeros((n, case.n_te, 3), np.float32)
    T = np.zeros((n, case.n_te), np.uint8)
    XYZ1 = np.zeros((n, case.traj_len, 3), np.float32)
    st = engine.walk_host(ph, seq, p, np.ascontiguousarray(case.M0[:n]), np.ascontiguousarray(case.XYZ0[:n]),
                          M1, XYZ1, T, xyz_scale=xs, gradient_scale=gs, rng=E.SW_RNG_REPLAY,
                          field_format=E.SW_FIELD_F32, replay=(normals, uniforms), spatial_order=1)
    _assert_parity(dict(M1=M1, T=T, XYZ1=XYZ1, stats=st), o, sl=slice(0, n))
    if st["n_lost"] == 0:
        assert st["n_steps"] == n * steps
    seq.close()
    ph.close()


def test_c3_replay_sample_on_the_full_phantom(engine):
    wl = W.c3()
    phs, cases_, _, _ = W.build_tiles(engine, dict(wl, tiles=[wl["tiles"][1]]), formats=(E.SW_FIELD_F32,))
    ph, case = phs[0], cases_[0]
    n = 1024
    steps = case.steps_per_spin()
    assert steps == 101 * 100 and case.n_spins == 10000000 and case.scale_type == 2
    mask, field = ph.download(field=True)
    d = orc.from_engine_case(case, 0, mask=mask, field_T=field)
    del mask, field
    # impermeable walls: rejected attempts draw again, so room for three attempts a step
    normals, uniforms = E.replay_fill(case.params.seed, 0, n, 9 * steps + 96, 3 * steps + 64)
    p, xs, gs = case.scale_params(case.scales[0])
    assert p.linear_phase_cycling == 90.0
    seq = engine.sequence(p, case.tables)
    o = orc.sim(d, spin_begin=0, spin_end=n, n_threads=8)
    M1 = np.zeros((n, case.n_te, 3), np.float32)
    T = np.zeros((n, case.n_te), np.uint8)
    XYZ1 = np.zeros((n, case.traj_len, 3), np.float32)
    st = engine.walk_host(ph, seq, p, np.ascontiguousarray(case.M0[:n]), np.ascontiguousarray(case.XYZ0[:n]),
                          M1, XYZ1, T, xyz_scale=xs, gradient_scale=gs, rng=E.SW_RNG_REPLAY,
                          field_format=E.SW_FIELD_F32, replay=(normals, uniforms), spatial_order=1)
    _assert_parity(dict(M1=M1, T=T, XYZ1=XYZ1, stats=st), o, sl=slice(0, n))
    assert st["n_steps"] == n * steps
    seq.close()
    ph.close()


def test_c3_fused_batch_equals_separate_walks(engine):
    import torch
    wl = W.c3()
    phs, items = _tile_items(engine, torch, wl, E.SW_PREC_F32_KAHAN, E.SW_FIELD_Q16, (E.SW_FIELD_Q16,))
    assert len(items) == 4 and items[0]["n_local"] == 10000000 and items[0]["steps"] == 10100
    b = engine.batch(items)
    assert b.info() == dict(walks=1, fused_tiles=4)
    st = b.run(stats=True)
    torch.cuda.synchronize()
    assert st["total"]["n_lost"] == 0
    for it in items:
        M1 = torch.zeros_like(it["M1"])
        T = torch.zeros_like(it["T"])
        XYZ1 = torch.zeros_like(it["XYZ1"])
        s = engine.walk_device(it["phantom"], it["sequence"], it["params"], it["M0"], it["XYZ0"], M1, XYZ1, T,
                               xyz_scale=it["xyz_scale"], gradient_scale=it["gradient_scale"],
                               precision=E.SW_PREC_F32_KAHAN, field_format=E.SW_FIELD_Q16, spatial_order=1,
                               stats=True)
        torch.cuda.synchronize()
        assert s["n_lost"] == 0
        assert torch.equal(M1, it["M1"]) and torch.equal(T, it["T"]) and torch.equal(XYZ1, it["XYZ1"])
        del M1, T, XYZ1
    b.close()
    _close(phs, items)
