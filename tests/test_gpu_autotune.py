"""Launch shapes and the autotuner: every pair-gather shape gives the default
shape's forces to rounding (only the summation order of a particle's pair
forces changes), and tis_autotune returns a valid, no slower setting."""
import numpy as np
import pytest

import tis_ctypes as T

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sys1(cfg1):
    return T.SystemHandle(cfg1[0])


def _forces(lib, h, **kw):
    eng = T.Engine(lib, [h], precision=T.PREC_FP64, **kw)
    f, e = eng.eval(T.TERM_SR | T.TERM_LR)
    pairs = eng.pairs(0)
    eng.close()
    return f, e, pairs


@pytest.mark.parametrize("lanes,block,waves", [(1, 256, 4096), (2, 64, 1024), (8, 128, 2048), (16, 256, 8192),
                                               (64, 64, 16384)])
def test_pair_gather_shapes_agree(lib, sys1, lanes, block, waves):
    f0, e0, p0 = _forces(lib, sys1)
    f, e, p = _forces(lib, sys1, pair_lanes=lanes, pair_block=block, list_waves=waves)
    assert np.array_equal(p, p0)
    assert np.linalg.norm(f - f0) / np.linalg.norm(f0) < 1e-13
    for k in ("e_wca", "e_coul_sr", "e_coul_lr"):
        assert e[k] == pytest.approx(e0[k], rel=1e-12), k


def test_autotune_returns_a_valid_setting(lib, sys1, cfg1):
    d2, _ = T.build_16scd(lib, T=0.596, c_mg=0.010, c_k=0.050, seed=2, box=400.0)
    hs = [sys1, T.SystemHandle(d2)]
    best, rep = T.autotune(lib, hs, trial_steps=100, precision=T.PREC_MIXED)
    print(best, rep)
    assert rep["trials"] >= 10
    assert 0 < rep["ms_per_step_best"] <= rep["ms_per_step_start"]
    assert best["pair_lanes"] in (2, 4, 8, 16) and best["pair_block"] in (64, 128, 256)
    assert best["graph_steps"] in (25, 50, 100, 200) and best["list_waves"] in (1024, 2048, 4096, 8192, 16384)
    assert best["hb_groups"] in (1, 2, 4, 8)
    eng = T.Engine(lib, hs, precision=T.PREC_MIXED, **best)
    eng.step(200)
    for r in range(2):
        pos, _, _ = eng.get_state(r)
        assert np.isfinite(pos).all()
    eng.close()


@pytest.mark.parametrize("prec", [T.PREC_FP64, T.PREC_MIXED])
def test_hb_outcome_does_not_depend_on_the_resolution_groups(lib, sys1, prec):
    """Components of the Philox resolution are shared out over 1..16 workgroups
    per replica; each component draws from its own stream, so at one state the
    outcome, the energies and the forces are those of one workgroup (forces to
    the rounding of the per-bead atomics), and 30-step trajectories agree to the
    rounding those atomics leave (amplified by FP32 force arithmetic in mixed
    precision)."""
    d2, _ = T.build_16scd(lib, T=0.596, c_mg=0.020, c_k=0.050, seed=4, box=450.0)
    hs = [sys1, T.SystemHandle(d2)]
    ref, states = None, None
    for g in (1, 2, 4, 8, 16):
        eng = T.Engine(lib, hs, precision=prec, rng=T.RNG_PHILOX, seed=3, hb_groups=g)
        eng.step(30)
        x30 = [eng.get_state(r)[0] for r in range(2)]
        if states is None:
            states = [eng.get_state(r) for r in range(2)]
        res = []
        for r in range(2):
            eng.set_state(*states[r], klok=31, rep=r)
            f, e = eng.eval(T.TERM_HB, rep=r)
            res.append((eng.hb_outcome(r), f, e["e2_hb"] + e["e3_hb"], x30[r]))
        eng.close()
        if ref is None:
            ref = res
            assert all((o == 2).any() for o, _, _, _ in res)   # the resolution switched candidates off
            continue
        for (o, f, e, x), (o0, f0, e0, x0) in zip(res, ref):
            assert np.array_equal(o, o0), g
            assert np.linalg.norm(f - f0) <= 1e-12 * np.linalg.norm(f0), g
            assert e == pytest.approx(e0, rel=1e-12), g
            assert np.abs(x - x0).max() < (1e-9 if prec == T.PREC_FP64 else 1e-6), g


@pytest.mark.parametrize("prec,rng", [(T.PREC_FP64, T.RNG_PHILOX), (T.PREC_MIXED, T.RNG_PHILOX),
                                      (T.PREC_FP64, T.RNG_STREAM)])
def test_hb_tables_built_beside_the_resolution_change_nothing(lib, sys1, prec, rng):
    """hb_table_lag = 1 builds the HB tables of a step's list rebuild beside that
    step's valence resolution, which reads the previous rebuild's tables; every
    candidate within its gate is listed in both, so the rebuild counts, the HB
    outcomes and the trajectory match hb_table_lag = 0 (the trajectory to the
    rounding the per-bead force atomics leave, amplified over 150 steps), with
    Philox draws and with the reference's stream (whose consumption must not
    change either)."""
    d2, _ = T.build_16scd(lib, T=0.596, c_mg=0.020, c_k=0.050, seed=4, box=450.0)
    hs = [sys1, T.SystemHandle(d2)]
    runs = []
    for lag in (0, 1):
        eng = T.Engine(lib, hs, precision=prec, rng=rng, seed=3, hb_table_lag=lag, graph_steps=25)
        if rng == T.RNG_STREAM:   # the reference's draws: the same stream in both runs
            for r in range(2):
                eng.set_stream(T.glibc_stream(lib, 11 + r, 1_500_000), rep=r)
        out = []
        for _ in range(6):
            eng.step(25)
            out.append([(eng.hb_outcome(r), eng.get_state(r)[0], eng.energies(r)["n_rebuilds"]) for r in range(2)])
        eng.close()
        runs.append(out)
    assert all(rb > 1 for _, _, rb in runs[0][-1])          # several rebuilds inside the steps
    tol = 1e-8 if prec == T.PREC_FP64 else 1e-4
    for b0, b1 in zip(*runs):
        for (o0, x0, n0), (o1, x1, n1) in zip(b0, b1):
            assert n0 == n1
            assert np.array_equal(o0, o1)
            assert np.abs(x0 - x1).max() < tol
