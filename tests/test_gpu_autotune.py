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
    eng = T.Engine(lib, hs, precision=T.PREC_MIXED, **best)
    eng.step(200)
    for r in range(2):
        pos, _, _ = eng.get_state(r)
        assert np.isfinite(pos).all()
    eng.close()
