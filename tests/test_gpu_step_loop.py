"""The production walk's two step loops give the same walk.

A single production walk runs its step loop with one attempt per trip or with
two, the second formed from the first's position before the first's gather
returns and dropped when the first is rejected, changes tissue or ends the TR
(sw_walk_opts.step_loop).  Every attempt takes the draws of its index either
way, so the two loops must agree bit for bit: magnetisation, labels, positions
(every recorded one) and the step and attempt counts.  The cases are chosen so
that the second attempt is dropped for each of its reasons: rejected crossings,
accepted crossings into tissues with other step sizes, lanes of still tissue
beside moving ones, spins that get stuck and are abandoned, the end of every TR
(dummy scans), and events of every kind on either attempt of a trip.
"""
import numpy as np
import pytest

from tests import cases
from tests.test_gpu_production import run
from spinwalk_amd import engine as E

pytestmark = pytest.mark.gpu


def both(engine, case, **kw):
    one = run(engine, case, step_loop=E.SW_STEP_LOOP_ONE, **kw)
    two = run(engine, case, step_loop=E.SW_STEP_LOOP_TWO, **kw)
    for k in ("M1", "T", "XYZ1"):
        np.testing.assert_array_equal(one[k], two[k], err_msg=k)
    for k in ("n_steps", "n_attempts", "n_lost", "n_replay_exhausted"):
        assert one["stats"][k] == two["stats"][k], k
    return one


@pytest.mark.parametrize("precision,field_format", [(E.SW_PREC_F32, E.SW_FIELD_F32),
                                                    (E.SW_PREC_F32_KAHAN, E.SW_FIELD_Q16),
                                                    (E.SW_PREC_F32_KAHAN, E.SW_FIELD_F16),
                                                    (E.SW_PREC_F64, E.SW_FIELD_F32)])
def test_crossings_with_other_step_sizes(engine, precision, field_format):
    """Three tissues of different diffusivity, crossings accepted half the time."""
    m, f = cases.three_substrate_phantom(n=24, fov_um=48.0, seed=41)
    cfg = cases.se(3, TE=6000, dt=50, DIFFUSIVITY=[1e-9, 2.5e-9, 0.4e-9],
                   P_XY=[[1, 0.5, 0.5], [0.5, 1, 0.5], [0.5, 0.5, 1]])
    case = cases.make_case(cfg, m, f, 48.0, n_spins=8192)
    g = both(engine, case, precision=precision, field_format=field_format,
             formats=(E.SW_FIELD_F32, E.SW_FIELD_F16, E.SW_FIELD_Q16))
    st = g["stats"]
    assert st["n_attempts"] > st["n_steps"] > 0 and st["n_lost"] == 0


def test_trajectories_dummy_scans_and_phase_cycling(engine):
    """bSSFP with dummy scans (the TR ends with trips of either parity) and every
    accepted position recorded."""
    m, f = cases.two_substrate_phantom(n=20, fov_um=40.0, seed=42)
    cfg = cases.bssfp(2, TE=1000, dt=50, dummies=3, RECORD_TRAJECTORY=1, P_XY=[[1, 0.3], [0.3, 1]])
    case = cases.make_case(cfg, m, f, 40.0, n_spins=2048)
    g = both(engine, case)
    assert g["XYZ1"].shape[1] == case.traj_len and case.traj_len > 100


def test_gradients_and_dephasing(engine):
    """PGSE gradients and the stimulated-echo dephasing events."""
    m, f = cases.two_substrate_phantom(n=24, fov_um=48.0, seed=43)
    for cfg in (cases.pgse(2, b=1000.0, delta_ms=2.0, DELTA_ms=4.0, dt=50, P_XY=[[1, 0.5], [0.5, 1]])[0],
                cases.ste(2, dt=50, P_XY=[[1, 0.5], [0.5, 1]])):
        case = cases.make_case(cfg, m, f, 48.0, n_spins=4096)
        both(engine, case)


def test_still_tissue_stuck_spins_and_cross_fov(engine):
    """Still tissue beside mobile tissue (lanes that never move), a crossing limit
    of two rejections in a row so that spins get stuck and are abandoned, and a
    walk that crosses the FoV."""
    n_vox = 16
    m = np.zeros((n_vox,) * 3, np.uint8)
    m[: n_vox // 2] = 1
    m[n_vox // 2:, :4] = 2
    rng = np.random.default_rng(44)
    xyz0 = (rng.uniform(0.02, 0.98, size=(4096, 3)) * 32e-6).astype(np.float32)
    cfg = cases.gre(3, TE=4000, dt=50, DIFFUSIVITY=[2e-9, 0.0, 2e-9], P_XY=[[1, 0, 0], [0, 1, 0], [0, 0, 1]],
                    MAX_ITERATIONS=2)
    case = cases.make_case(cfg, m, None, 32.0, n_spins=4096, xyz0=xyz0)
    g = both(engine, case, spatial_order=0)
    assert g["stats"]["n_lost"] > 0                                # some spins were abandoned
    cfg = cases.gre(3, TE=4000, dt=50, DIFFUSIVITY=[2e-9, 1e-9, 2e-9], P_XY=[[1, 0.5, 0.5], [0.5, 1, 0.5],
                                                                              [0.5, 0.5, 1]], CROSS_FOV=1)
    case = cases.make_case(cfg, m, None, 32.0, n_spins=4096, xyz0=xyz0)
    both(engine, case)


def test_auto_takes_one_attempt_for_a_small_phantom(engine):
    """The default (auto) loop on a phantom far below 256 MiB of records is the
    one-attempt loop: same results as either explicit loop."""
    m, f = cases.two_substrate_phantom(n=16, fov_um=32.0, seed=45)
    case = cases.make_case(cases.gre(2, TE=2000, dt=50), m, f, 32.0, n_spins=1024)
    a = run(engine, case)
    b = both(engine, case)
    for k in ("M1", "T", "XYZ1"):
        np.testing.assert_array_equal(a[k], b[k])


def test_rejects_unknown_step_loop(engine):
    m, f = cases.two_substrate_phantom(n=16, fov_um=32.0, seed=46)
    case = cases.make_case(cases.gre(2, TE=2000, dt=50), m, f, 32.0, n_spins=64)
    with pytest.raises(E.EngineError, match="step_loop"):
        run(engine, case, step_loop=3)
