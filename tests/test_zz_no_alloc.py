"""sw_walk_device_noalloc: the walk of sw_walk_device with its counters, visiting
order and sort buffers in the library's fixed device scratch instead of
engine-owned allocations.  Its results are sw_walk_device's bit for bit: two
calls in a row, the second with fewer spins (so the scratch holds the first
call's order and counters when the second starts), with and without the
spatial visiting order, in production and replay mode.  Walks that need more
scratch than the fixed array holds are refused before anything is launched."""
import numpy as np
import pytest

from tests import cases
from spinwalk_amd import engine as E

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(engine):
    import torch
    m, f = cases.three_substrate_phantom(n=24, fov_um=48.0, seed=41)
    cfg = cases.se(3, TE=4000, dt=50, NUMBER_OF_SPINS=20000, DIFFUSIVITY=[1e-9, 1e-9, 1e-9],
                   P_XY=[[1, 0.5, 0.5], [0.5, 1, 0.5], [0.5, 0.5, 1]])
    case = cases.make_case(cfg, m, f, 48.0)
    ph = engine.phantom(m, f, case.fov, formats=(E.SW_FIELD_F32, E.SW_FIELD_Q16))
    p, xs, gs = case.scale_params(case.scales[0])
    seq = engine.sequence(p, case.tables)
    yield torch, case, ph, seq, p
    seq.close()
    ph.close()


def _walk(engine, torch, case, ph, seq, p, n, noalloc, **kw):
    dev = torch.device("cuda:0")
    M0 = torch.from_numpy(np.ascontiguousarray(case.M0[:n])).to(dev)
    XYZ0 = torch.from_numpy(np.ascontiguousarray(case.XYZ0[:n])).to(dev)
    dt = torch.float64 if kw.get("precision") == E.SW_PREC_F64 else torch.float32
    M1 = torch.zeros((n, case.n_te, 3), dtype=dt, device=dev)
    T = torch.zeros((n, case.n_te), dtype=torch.uint8, device=dev)
    XYZ1 = torch.zeros((n, case.traj_len, 3), dtype=torch.float32, device=dev)
    st = engine.walk_device(ph, seq, p, M0, XYZ0, M1, XYZ1, T, n_local=n, noalloc=noalloc, stats=True, **kw)
    torch.cuda.synchronize()
    return dict(M1=M1.cpu().numpy(), T=T.cpu().numpy(), XYZ1=XYZ1.cpu().numpy(), stats=st)


@pytest.mark.parametrize("order", [1, 0])
@pytest.mark.parametrize("precision,field", [(E.SW_PREC_F32, E.SW_FIELD_F32),
                                             (E.SW_PREC_F32_KAHAN, E.SW_FIELD_Q16)])
def test_noalloc_equals_walk_device_on_two_calls(engine, setup, order, precision, field):
    torch, case, ph, seq, p = setup
    for n in (20000, 7000):
        kw = dict(spatial_order=order, precision=precision, field_format=field)
        ref = _walk(engine, torch, case, ph, seq, p, n, False, **kw)
        got = _walk(engine, torch, case, ph, seq, p, n, True, **kw)
        for k in ("M1", "T", "XYZ1"):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{k}, n={n}")
        for k in ("n_lost", "n_replay_exhausted", "n_steps", "n_attempts"):
            assert got["stats"][k] == ref["stats"][k], (k, n)
        assert ref["stats"]["n_steps"] == n * case.steps_per_spin()


def test_noalloc_replay_equals_walk_device(engine, setup):
    torch, case, ph, seq, p = setup
    dev = torch.device("cuda:0")
    steps = case.steps_per_spin()
    for n in (3000, 1000):
        rn, ru = E.replay_fill(case.params.seed, 0, n, 9 * steps + 96, 3 * steps + 64)
        rp = (torch.from_numpy(rn).to(dev), torch.from_numpy(ru).to(dev))
        kw = dict(spatial_order=1, rng=E.SW_RNG_REPLAY, replay=rp)
        ref = _walk(engine, torch, case, ph, seq, p, n, False, **kw)
        got = _walk(engine, torch, case, ph, seq, p, n, True, **kw)
        for k in ("M1", "T", "XYZ1"):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=f"{k}, n={n}")
        assert got["stats"] == dict(ref["stats"], kernel_ms=got["stats"]["kernel_ms"])


def test_noalloc_refuses_a_walk_beyond_its_scratch(engine):
    import torch
    m, f = cases.two_substrate_phantom(n=24, fov_um=48.0, seed=42)
    n = 3000000                        # order + sort buffers: about 36 MB > 32 MiB
    case = cases.make_case(cases.gre(2, TE=1000, dt=50, NUMBER_OF_SPINS=n), m, f, 48.0)
    ph = engine.phantom(m, f, case.fov)
    p, xs, gs = case.scale_params(case.scales[0])
    seq = engine.sequence(p, case.tables)
    dev = torch.device("cuda:0")
    M0 = torch.from_numpy(case.M0).to(dev)
    XYZ0 = torch.from_numpy(case.XYZ0).to(dev)
    M1 = torch.zeros((n, 1, 3), device=dev)
    T = torch.zeros((n, 1), dtype=torch.uint8, device=dev)
    XYZ1 = torch.zeros((n, 1, 3), device=dev)
    with pytest.raises(E.EngineError, match="no-allocation scratch holds"):
        engine.walk_device(ph, seq, p, M0, XYZ0, M1, XYZ1, T, spatial_order=1, noalloc=True)
    torch.cuda.synchronize()
    assert not M1.any()                # nothing was launched
    # without the spatial order the same walk needs only the counters
    st = engine.walk_device(ph, seq, p, M0, XYZ0, M1, XYZ1, T, spatial_order=0, noalloc=True)
    assert st["n_steps"] == n * case.steps_per_spin() and st["n_lost"] == 0
    seq.close()
    ph.close()


def test_noalloc_refuses_a_sequence_built_for_other_phase_cycling(engine, setup):
    torch, case, ph, seq, p = setup
    q = type(p).from_buffer_copy(p)
    q.linear_phase_cycling = 45.0
    n = 16
    dev = torch.device("cuda:0")
    M0 = torch.from_numpy(np.ascontiguousarray(case.M0[:n])).to(dev)
    XYZ0 = torch.from_numpy(np.ascontiguousarray(case.XYZ0[:n])).to(dev)
    M1 = torch.zeros((n, case.n_te, 3), device=dev)
    T = torch.zeros((n, case.n_te), dtype=torch.uint8, device=dev)
    XYZ1 = torch.zeros((n, case.traj_len, 3), device=dev)
    with pytest.raises(E.EngineError, match="excitation table"):
        engine.walk_device(ph, seq, q, M0, XYZ0, M1, XYZ1, T, n_local=n, noalloc=True)
