"""The HB resolution's phase clock (tis_trace_hb): switching it on changes no
result (trajectories agree to the rounding that the per-bead force atomics
leave in any two runs), and the ring holds, for each traced step, replica and
resolution group, stamps in phase order and the sizes the step reports."""
import numpy as np
import pytest

import tis_ctypes as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [T.PREC_FP64, T.PREC_MIXED])
def test_trace_leaves_the_trajectory_unchanged(lib, cfg1, prec):
    d2, _ = T.build_16scd(lib, T=0.596, c_mg=0.010, c_k=0.050, seed=2, box=400.0)
    hs = [T.SystemHandle(cfg1[0]), T.SystemHandle(d2)]
    runs = []
    for on in (False, True):
        eng = T.Engine(lib, hs, precision=prec, seed=7, hb_groups=4, graph_steps=32)
        eng.trace_hb(on)
        eng.step(128)                    # four graph replays
        runs.append(([eng.get_state(r)[0] for r in range(2)], [eng.energies(r) for r in range(2)],
                     [eng.hb_outcome(r) for r in range(2)], eng.hb_trace() if on else None))
        eng.close()
    (p0, e0, o0, _), (p1, e1, o1, tr) = runs
    tol = 1e-8 if prec == T.PREC_FP64 else 1e-4
    for r in range(2):
        assert np.abs(p0[r] - p1[r]).max() < tol, r
        assert np.array_equal(o0[r], o1[r]), r
        for k in ("klok", "n_rebuilds", "n_hb_active"):
            assert e0[r][k] == e1[r][k], (r, k)
    # steps 65..128 fill the 64 slots; groups 0..3 hold stamps, the others stay 0
    assert tr.shape == (64, 2, 16, 10)
    assert not tr[:, :, 4:, :].any()
    st = tr[:, :, :4, :8].astype(np.int64)
    assert (st > 0).all()
    assert (np.diff(st, axis=-1) >= 0).all()
    # the last step's sizes: live candidates of the table it resolved over
    live = (tr[:, :, :4, 9] >> np.uint64(32)).astype(np.int64)
    assert (live > 0).all() and (live <= 3072).all()
    assert (live == live[:, :, :1]).all()            # every group of a replica sees the same table
    n_over = (tr[:, :, :4, 9] & np.uint64(0xffff)).astype(np.int64)
    n_comp = ((tr[:, :, :4, 9] >> np.uint64(16)) & np.uint64(0xffff)).astype(np.int64)
    assert (n_comp <= n_over).all() and ((n_over == 0) == (n_comp == 0)).all()
    # the serial depth the step's close reports is the largest group's
    depth = tr[:, :, :4, 8].astype(np.int64).max(axis=2)
    slot = (128 % 64)                                 # klok of the last step: 128
    for r in range(2):
        assert depth[slot, r] == e1[r]["hb_rounds"], r


def test_trace_entry_points_reject_bad_arguments(lib, cfg1):
    eng = T.Engine(lib, [T.SystemHandle(cfg1[0])], precision=T.PREC_FP64)
    with pytest.raises(RuntimeError):
        eng.hb_trace()                   # never switched on
    assert lib.tis_trace_hb(eng.ctx, 2) == -1       # TIS_EINVAL
    eng.close()
