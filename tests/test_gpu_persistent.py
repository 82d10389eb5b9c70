"""The persistent step (tis_options.persistent): each replica advances on its
own workgroups inside one cooperative launch, its phases separated by barriers
among those workgroups.  It runs the step order and device bodies of the
per-kernel step, so from one state the two agree to the rounding of the
atomically added forces (whose order differs), with the same step counters,
rebuild counts and HB outcomes; an odd number of steps ends on the other state
buffers."""
import numpy as np
import pytest

import tis_ctypes as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eight(lib, cfg1):
    """8 replicas of the 5 mM system (one per XCD), set-up draws of seeds 1..8."""
    hs = [T.SystemHandle(cfg1[0])]
    for k in range(2, 9):
        d, _ = T.build_16scd(lib, T=0.596, c_mg=0.005, c_k=0.050, seed=k)
        hs.append(T.SystemHandle(d))
    return hs


@pytest.mark.parametrize("prec,tol", [(T.PREC_FP64, 1e-9), (T.PREC_MIXED, 1e-5)])
def test_persistent_step_tracks_the_kernel_step(lib, eight, prec, tol):
    steps = 151   # odd: the persistent launch ends on the other buffer set
    a = T.Engine(lib, eight, precision=prec, rng=T.RNG_PHILOX, seed=77, persistent=1, hb_groups=4)
    b = T.Engine(lib, eight, precision=prec, rng=T.RNG_PHILOX, seed=77, persistent=0, hb_groups=4)
    pc = a.persistent_counts()
    assert pc["active"], pc["why"]
    assert not b.persistent_counts()["active"]
    a.step(steps)
    b.step(steps)
    pc = a.persistent_counts()
    assert pc["steps"] == steps and pc["launches"] == 1, pc
    assert b.launch_counts()["graph_steps"] + b.launch_counts()["eager_steps"] == steps
    worst = 0.0
    for r in range(len(eight)):
        xa, oa, va = a.get_state(r)
        xb, ob, vb = b.get_state(r)
        assert np.isfinite(xa).all()
        worst = max(worst, np.abs(xa - xb).max(), np.abs(oa - ob).max())
        ea, eb = a.energies(r), b.energies(r)
        assert ea["klok"] == eb["klok"] == steps + 1
        assert ea["n_rebuilds"] == eb["n_rebuilds"]
        assert np.array_equal(a.hb_outcome(r), b.hb_outcome(r))
        assert ea["e3_stack"] == pytest.approx(eb["e3_stack"], rel=1e-6, abs=1e-6)
    print(f"largest position difference after {steps} steps: {worst:.3e} A")
    assert worst < tol
    a.close()
    b.close()


def test_persistent_step_switches_off_and_on(lib, eight):
    """tis_set_persistent moves a context between the two step forms mid-run;
    the trajectory continues (counters and buffers carried over)."""
    a = T.Engine(lib, eight, precision=T.PREC_MIXED, rng=T.RNG_PHILOX, seed=5, persistent=1)
    b = T.Engine(lib, eight, precision=T.PREC_MIXED, rng=T.RNG_PHILOX, seed=5, persistent=0)
    a.step(25)
    a.set_persistent(False)
    a.step(30)
    a.set_persistent(True)
    a.step(25)
    b.step(80)
    assert a.persistent_counts()["steps"] == 50
    for r in (0, 7):
        xa, _, _ = a.get_state(r)
        xb, _, _ = b.get_state(r)
        assert a.energies(r)["klok"] == b.energies(r)["klok"] == 81
        assert np.abs(xa - xb).max() < 1e-5
    a.close()
    b.close()


def test_persistent_step_is_refused_where_it_cannot_run(lib, cfg1):
    """One replica (fewer than the 8 XCDs) keeps the per-kernel step; switching
    the persistent step on is refused with the reason."""
    e = T.Engine(lib, [T.SystemHandle(cfg1[0])], precision=T.PREC_MIXED, persistent=1)
    pc = e.persistent_counts()
    assert not pc["active"] and "XCD" in pc["why"]
    with pytest.raises(T.TisError):
        e.set_persistent(True)
    e.step(5)
    e.close()
