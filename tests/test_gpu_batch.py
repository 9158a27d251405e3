"""The batched walk (sw_batch_*): many tiles in one launch, captured as a HIP graph.

A tile of a batch is exactly what one sw_walk_device call walks with the same
arguments, and its draws depend only on (seed, global spin, attempt) or on its own
replay streams, so every tile must reproduce the separate call bit for bit:
trajectories, labels, magnetisation, counters and FP64 echo sums.  The graph is
replayed several times with eager work of the program in between, and every
replay is checked against the separate calls.
"""
import numpy as np
import pytest

from tests import cases
from spinwalk_amd import engine as E

pytestmark = pytest.mark.gpu


def _tile(engine, torch, case, scale_index=0, *, spin_offset=0, n=None, precision=E.SW_PREC_F32,
          field_format=E.SW_FIELD_F32, rng=E.SW_RNG_PHILOX, formats=(E.SW_FIELD_F32,), ph=None, shared=None):
    p, xs, gs = case.scale_params(case.scales[scale_index])
    n = case.n_spins - spin_offset if n is None else n
    sl = slice(spin_offset, spin_offset + n)
    if ph is None:
        ph = engine.phantom(case.mask, case.field_T, case.fov, formats=formats)
    seq = engine.sequence(p, case.tables)
    dt = torch.float64 if precision == E.SW_PREC_F64 else torch.float32
    t = dict(phantom=ph, sequence=seq, params=p, spin_offset=spin_offset, n_local=n, xyz_scale=xs,
             gradient_scale=gs, precision=precision, field_format=field_format, rng=rng, spatial_order=1,
             M0=torch.from_numpy(np.ascontiguousarray(case.M0[sl])).cuda(),
             XYZ0=torch.from_numpy(np.ascontiguousarray(case.XYZ0[sl])).cuda(),
             M1=torch.zeros((n, case.n_te, 3), dtype=dt, device="cuda"),
             T=torch.zeros((n, case.n_te), dtype=torch.uint8, device="cuda"),
             XYZ1=torch.zeros((n, case.traj_len, 3), dtype=torch.float32, device="cuda"),
             echo_sums=torch.zeros((case.n_te, p.n_substrate, 3), dtype=torch.float64, device="cuda"))
    if rng == E.SW_RNG_REPLAY:
        steps = case.steps_per_spin()
        nn, uu = E.replay_fill(int(p.seed), spin_offset, n, 9 * steps + 96, 3 * steps + 64)
        t["replay"] = (torch.from_numpy(nn).cuda(), torch.from_numpy(uu).cuda())
    if shared is not None:                     # another tile's start and draw buffers (same contents)
        t.update(shared)
    return t


def _separate(engine, torch, t):
    """The same tile through sw_walk_device, into fresh buffers."""
    M1, T, X = torch.zeros_like(t["M1"]), torch.zeros_like(t["T"]), torch.zeros_like(t["XYZ1"])
    st = engine.walk_device(t["phantom"], t["sequence"], t["params"], t["M0"], t["XYZ0"], M1, X, T,
                            spin_offset=t["spin_offset"], n_local=t["n_local"], xyz_scale=t["xyz_scale"],
                            gradient_scale=t["gradient_scale"], rng=t["rng"], precision=t["precision"],
                            field_format=t["field_format"], spatial_order=0, replay=t.get("replay"))
    sums = torch.zeros_like(t["echo_sums"])
    engine.echo_reduce(M1, T, t["params"].n_substrate, sums)
    torch.cuda.synchronize()
    return dict(M1=M1.cpu().numpy(), T=T.cpu().numpy(), XYZ1=X.cpu().numpy(), sums=sums.cpu().numpy(), stats=st)


def _check(t, ref, st=None):
    np.testing.assert_array_equal(t["XYZ1"].cpu().numpy(), ref["XYZ1"])
    np.testing.assert_array_equal(t["T"].cpu().numpy(), ref["T"])
    np.testing.assert_array_equal(t["M1"].cpu().numpy(), ref["M1"])
    np.testing.assert_array_equal(t["echo_sums"].cpu().numpy(), ref["sums"])
    if st is not None:
        for k in ("n_lost", "n_steps", "n_attempts"):
            assert st[k] == ref["stats"][k], k


def _clear(tiles):
    for t in tiles:
        t["M1"].zero_(); t["T"].zero_(); t["XYZ1"].zero_(); t["echo_sums"].zero_()


def _cases():
    m1, f1 = cases.cylinder_phantom(n=24, fov_um=48.0, radius_um=5.0, n_cyl=3, seed=41)
    m2, f2 = cases.three_substrate_phantom(n=24, fov_um=48.0, seed=42)
    m3, f3 = cases.blocky_phantom((17, 9, 31), n_labels=2, seed=43)
    gre = cases.gre(2, TE=4000, dt=50, P_XY=[[1, 0.5], [0.5, 1]], SCALE=[1.0, 1.7])
    se = cases.se(3, TE=6000, dt=50, CROSS_FOV=1, P_XY=[[1, 0.3, 0.3], [0.3, 1, 0.3], [0.3, 0.3, 1]])
    ssfp = cases.bssfp(2, TE=1000, dt=50, dummies=4, lin=90.0, WHAT_TO_SCALE=2, SCALE=[1.0, 0.5],
                       P_XY=[[1, 0.5], [0.5, 1]])
    pg, _ = cases.pgse(2, b=800.0, delta_ms=1.0, DELTA_ms=2.0, dt=50, axis=(0.3, 1.0, 0.2), start_ms=0.5,
                       P_XY=[[1, 0.5], [0.5, 1]], WHAT_TO_SCALE=1, SCALE=[1.0, 0.7])
    return [
        (cases.make_case(gre, m1, f1 * 30, 48.0, n_spins=3000), 0, {}),
        (cases.make_case(gre, m1, f1 * 30, 48.0, n_spins=3000), 1, {}),               # FoV scale
        (cases.make_case(se, m2, f2, 48.0, n_spins=2000), 0, dict(spin_offset=300, n=1300)),
        (cases.make_case(ssfp, m3, f3, (17.0, 9.0, 31.0), n_spins=1500), 1, {}),      # phase-cycling scale
        (cases.make_case(pg, m1, f1, 48.0, n_spins=2500), 1, {}),                     # gradient scale
        (cases.make_case(gre, m1, None, 48.0, n_spins=700), 0, {}),                   # no fieldmap
    ]


def test_batch_equals_separate_walks_over_graph_replays(engine):
    import torch
    cs = _cases()
    # the no-field tile gathers label-only records: a batch shares one record
    # format, so it gets a batch of its own
    tiles = [_tile(engine, torch, c, si, **kw) for c, si, kw in cs[:-1]]
    refs = [_separate(engine, torch, t) for t in tiles]
    b = engine.batch(tiles)
    gi = b.graph_info()
    assert gi["uses_graph"] and gi["nodes"] >= 2, gi      # counter reset + walks (+ echo sums), one graph
    other = _tile(engine, torch, cs[0][0])
    for rep in range(3):
        _clear(tiles)
        st = b.run()
        torch.cuda.synchronize()
        for t, r, s in zip(tiles, refs, st["items"]):
            _check(t, r, s)
        assert st["total"]["n_steps"] == sum(r["stats"]["n_steps"] for r in refs)
        # eager work of the program between replays
        _separate(engine, torch, other)
        torch.cuda.synchronize()
    b.close()
    nf = [_tile(engine, torch, cs[-1][0])]
    ref = _separate(engine, torch, nf[0])
    b2 = engine.batch(nf)
    b2.run()
    torch.cuda.synchronize()
    _check(nf[0], ref)


def test_batch_replay_mode_and_precisions(engine):
    import torch
    cs = _cases()
    for prec, ff, rng in ((E.SW_PREC_F32, E.SW_FIELD_F32, E.SW_RNG_REPLAY),
                          (E.SW_PREC_F64, E.SW_FIELD_F32, E.SW_RNG_PHILOX),
                          (E.SW_PREC_F32_KAHAN, E.SW_FIELD_F16, E.SW_RNG_PHILOX),
                          (E.SW_PREC_F32, E.SW_FIELD_Q16, E.SW_RNG_REPLAY),
                          (E.SW_PREC_F32_KAHAN, E.SW_FIELD_Q16, E.SW_RNG_PHILOX)):
        fm = (E.SW_FIELD_F32, E.SW_FIELD_F16, E.SW_FIELD_Q16)
        tiles = [_tile(engine, torch, c, si, precision=prec, field_format=ff, rng=rng, formats=fm, **kw)
                 for c, si, kw in cs[:3]]
        refs = [_separate(engine, torch, t) for t in tiles]
        b = engine.batch(tiles)
        for _ in range(2):
            _clear(tiles)
            st = b.run()
            torch.cuda.synchronize()
            for t, r, s in zip(tiles, refs, st["items"]):
                _check(t, r, s)
        b.close()


def test_batch_rejects_mixed_modes(engine):
    import torch
    cs = _cases()
    a = _tile(engine, torch, cs[0][0])
    c = _tile(engine, torch, cs[0][0], precision=E.SW_PREC_F32_KAHAN)
    with pytest.raises(E.EngineError, match="shares rng and precision"):
        engine.batch([a, c])
    d = _tile(engine, torch, cs[-1][0])
    with pytest.raises(E.EngineError, match="same field format"):
        engine.batch([a, d])


def test_batch_schedules_and_autotuner_keep_results(engine):
    """The batch autotuner (sw_batch_tune) and explicit schedules change only the
    order in which chunks are walked: under every schedule, and after tuning,
    every tile still equals its separate walk bit for bit.  Tiles 0-2 share one
    phantom (a shared cost class), the others have their own."""
    import torch
    cs = _cases()
    shared = engine.phantom(cs[0][0].mask, cs[0][0].field_T, cs[0][0].fov)
    tiles = [_tile(engine, torch, cs[0][0], 0, ph=shared), _tile(engine, torch, cs[1][0], 1, ph=shared),
             _tile(engine, torch, cs[4][0], 1, ph=shared)]
    tiles += [_tile(engine, torch, c, si, **kw) for c, si, kw in (cs[2], cs[3])]
    refs = [_separate(engine, torch, t) for t in tiles]
    b = engine.batch(tiles)
    keys = ("chunk_spins", "bands_shared", "bands_mixed", "spins_per_lane", "waves_per_simd", "order_brick")
    assert b.schedule == dict(zip(keys, (256, 1 << 20, 4, 1, 0, 8)))

    def check_run():
        _clear(tiles)
        st = b.run()
        torch.cuda.synchronize()
        for t, r, x in zip(tiles, refs, st["items"]):
            _check(t, r, x)

    # chunk sizes and bands; two spins per lane; caps of 2 to 4 waves per SIMD;
    # visiting orders of 4^3, 16^3 and 32^3 voxel bricks (rebuilt in place)
    for sc in ((128, 64, 4), (64, 1, 1), (256, 16, 64), (128, 1 << 20, 16), (256, 1 << 20, 4, 2),
               (64, 4, 4, 2, 3), (256, 1 << 20, 4, 1, 2, 4), (128, 16, 16, 2, 0, 16), (256, 1 << 20, 4, 1, 4, 32),
               (256, 1 << 20, 4, 1, 0, 8)):
        b.set_schedule(*sc)
        want = dict(zip(keys, (256, 1 << 20, 4, 1, 0, 8)))
        want.update(zip(keys, sc))
        assert b.schedule == want
        assert b.graph_info()["uses_graph"]
        check_run()
    res = b.tune(max_trials=16)
    assert res["schedule"] == b.schedule and res["best_ms"] > 0 and res["first_ms"] > 0
    assert res["schedule"]["chunk_spins"] in (64, 128, 256)
    check_run()
    for bad in ((100, 1, 1), (256, 0, 1), (256, 1, 1, 3), (256, 1, 1, 1, 1), (256, 1, 1, 1, 9),
                (256, 1, 1, 1, 0, 6)):
        with pytest.raises(E.EngineError, match="chunk_spins"):
            b.set_schedule(*bad)
    check_run()                          # a rejected schedule leaves the batch as it was
    b.close()


def test_batch_fuses_same_trajectory_tiles(engine):
    """Gradient scales (PGSE, WHAT_TO_SCALE = 1) and phase-cycling scales (bSSFP,
    WHAT_TO_SCALE = 2) of one config walk the same trajectories when they share
    the start and draw buffers: the batch walks them together, four at most per
    walk (five gradient scales: a walk of four and a single walk; three
    phase-cycling scales: a walk of three).  Every tile still equals its separate
    walk bit for bit, in the production, Kahan + FP16 and replay modes."""
    import torch
    m1, f1 = cases.cylinder_phantom(n=24, fov_um=48.0, radius_um=5.0, n_cyl=3, seed=41)
    m3, f3 = cases.blocky_phantom((17, 9, 31), n_labels=2, seed=43)
    pg, _ = cases.pgse(2, b=800.0, delta_ms=1.0, DELTA_ms=2.0, dt=50, axis=(0.3, 1.0, 0.2), start_ms=0.5,
                       P_XY=[[1, 0.5], [0.5, 1]], WHAT_TO_SCALE=1, SCALE=[1.0, 0.7, 0.4, 0.1, 0.55])
    ssfp = cases.bssfp(2, TE=1000, dt=50, dummies=4, lin=90.0, WHAT_TO_SCALE=2, SCALE=[0.0, 1.0, 2.0],
                       P_XY=[[1, 0.5], [0.5, 1]])
    fm = (E.SW_FIELD_F32, E.SW_FIELD_F16)
    for prec, ff, rng in ((E.SW_PREC_F32, E.SW_FIELD_F32, E.SW_RNG_PHILOX),
                          (E.SW_PREC_F32_KAHAN, E.SW_FIELD_F16, E.SW_RNG_PHILOX),
                          (E.SW_PREC_F32, E.SW_FIELD_F32, E.SW_RNG_REPLAY)):
        tiles = []
        for cfg, m, f, fov, n_sc in ((pg, m1, f1, 48.0, 5), (ssfp, m3, f3, (17.0, 9.0, 31.0), 3)):
            case = cases.make_case(cfg, m, f, fov, n_spins=1500)
            ph = engine.phantom(case.mask, case.field_T, case.fov, formats=fm)
            shared = None
            for si in range(n_sc):
                t = _tile(engine, torch, case, si, precision=prec, field_format=ff, rng=rng, ph=ph, shared=shared)
                if shared is None:
                    shared = dict(M0=t["M0"], XYZ0=t["XYZ0"])
                    if "replay" in t:
                        shared["replay"] = t["replay"]
                tiles.append(t)
        refs = [_separate(engine, torch, t) for t in tiles]
        b = engine.batch(tiles)
        assert b.info() == dict(walks=3, fused_tiles=7)
        for _ in range(2):
            _clear(tiles)
            st = b.run()
            torch.cuda.synchronize()
            for t, r, x in zip(tiles, refs, st["items"]):
                _check(t, r, x)
        b.close()


def test_batch_mixes_record_layouts(engine):
    """The records of a phantom whose per-axis tables overflow LDS (16400 x 4 x 4
    voxels) stay in single bricks and are addressed in the 64-bit mode; a small
    phantom's records are grouped in super-blocks of bricks and addressed from
    the LDS tables. In one batch both tiles run in the 64-bit mode, which then
    reads each tile's layout: every tile equals its separate walk bit for bit,
    in both precisions the fused and single batches use."""
    import torch
    m_big, f_big = cases.blocky_phantom((16400, 4, 4), n_labels=2, seed=51, block=3)
    m_small, f_small = cases.cylinder_phantom(n=24, fov_um=48.0, radius_um=5.0, n_cyl=3, seed=52)
    gre = cases.gre(2, TE=2000, dt=50, P_XY=[[1, 0.5], [0.5, 1]])
    for prec in (E.SW_PREC_F32, E.SW_PREC_F32_KAHAN):
        tiles = [_tile(engine, torch, cases.make_case(gre, m_small, f_small * 30, 48.0, n_spins=1500),
                       precision=prec),
                 _tile(engine, torch, cases.make_case(gre, m_big, f_big * 30, (16400.0, 4.0, 4.0), n_spins=1500),
                       precision=prec)]
        refs = [_separate(engine, torch, t) for t in tiles]
        b = engine.batch(tiles)
        st = b.run()
        torch.cuda.synchronize()
        for t, r, s in zip(tiles, refs, st["items"]):
            _check(t, r, s)
        b.close()


def test_empty_inputs(engine):
    """Zero spins everywhere an entry point takes spins: a walk of none is a
    no-op that reports zero steps, echo sums over no spins are zero (and are
    written), and an empty tile in a batch leaves the other tiles' results
    bit-identical to their separate walks."""
    import torch
    cs = _cases()
    case = cs[0][0]
    p, xs, gs = case.scale_params(case.scales[0])
    ph = engine.phantom(case.mask, case.field_T, case.fov)
    seq = engine.sequence(p, case.tables)
    z = lambda *sh, dt=np.float32: np.zeros(sh, dt)
    st = engine.walk_host(ph, seq, p, z(0, 3), z(0, 3), z(0, case.n_te, 3), z(0, case.traj_len, 3),
                          z(0, case.n_te, dt=np.uint8))
    assert st["n_steps"] == 0 and st["n_lost"] == 0 and st["n_attempts"] == 0
    zt = lambda *sh, dt=torch.float32: torch.zeros(sh, dtype=dt, device="cuda")
    st = engine.walk_device(ph, seq, p, zt(0, 3), zt(0, 3), zt(0, case.n_te, 3), zt(0, case.traj_len, 3),
                            zt(0, case.n_te, dt=torch.uint8))
    assert st["n_steps"] == 0
    sums = torch.full((case.n_te, p.n_substrate, 3), 7.0, dtype=torch.float64, device="cuda")
    engine.echo_reduce(zt(0, case.n_te, 3), zt(0, case.n_te, dt=torch.uint8), p.n_substrate, sums)
    torch.cuda.synchronize()
    assert (sums.cpu().numpy() == 0).all()
    seq.close()
    empty = _tile(engine, torch, case, n=0)
    full = _tile(engine, torch, case, ph=empty["phantom"])
    ref = _separate(engine, torch, full)
    empty["echo_sums"].fill_(7.0)
    b = engine.batch([empty, full])
    st = b.run()
    torch.cuda.synchronize()
    assert st["items"][0]["n_steps"] == 0 and st["items"][0]["n_attempts"] == 0
    assert (empty["echo_sums"].cpu().numpy() == 0).all()
    _check(full, ref, st["items"][1])
    b.close()


def test_walks_on_two_streams_share_the_engine_scratch_safely(engine):
    """sw_walk_device keeps its counters, visiting order and excitation tables in
    engine-owned buffers.  Two walks enqueued back to back on two streams, the
    second with more spins (so its visiting order rewrites the buffer the first
    walk reads), must each equal their walk on the engine's own stream: every
    spin walked exactly once, counters per call."""
    import torch
    m, f = cases.three_substrate_phantom(n=24, fov_um=48.0, seed=51)
    cfg = cases.se(3, TE=8000, dt=50, P_XY=[[1, 0.5, 0.5], [0.5, 1, 0.5], [0.5, 0.5, 1]])
    ca = cases.make_case(cfg, m, f, 48.0, n_spins=6000)
    cb = cases.make_case(dict(cfg, SEED=77), m, f, 48.0, n_spins=24000)
    ph = engine.phantom(m, f, ca.fov)
    ta = _tile(engine, torch, ca, ph=ph)
    tb = _tile(engine, torch, cb, ph=ph)
    ra, rb = _separate(engine, torch, ta), _separate(engine, torch, tb)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(2):
        _clear([ta, tb])
        torch.cuda.synchronize()
        outs = []
        for t, s in ((ta, sa), (tb, sb)):
            with torch.cuda.stream(s):
                engine.walk_device(t["phantom"], t["sequence"], t["params"], t["M0"], t["XYZ0"], t["M1"], t["XYZ1"],
                                   t["T"], spatial_order=1, stream=s, stats=False)
                engine.echo_reduce(t["M1"], t["T"], t["params"].n_substrate, t["echo_sums"], stream=s)
        torch.cuda.synchronize()
        for t, r in ((ta, ra), (tb, rb)):
            _check(t, r)
