"""Degenerate geometry and non-finite states.

A straight HB angle: the cosine of pi, formed in FP32 from the bead
differences, can round just past -1, where acos is NaN; the HB energy, the
resolution and the forces must stay finite (the cosine is clamped to [-1, 1],
which leaves every cosine inside that range unchanged).  A 16-replica batch in
mixed precision reached such a state near step 700 of its seed-3 trajectory.
Coincident particles give a NaN pair force; the run must stop with
TIS_ENONFINITE rather than carry NaN coordinates on.  Non-finite input is
refused by tis_set_state."""
import numpy as np
import pytest

import tis_ctypes as T

pytestmark = pytest.mark.gpu

ENONFINITE = -7


@pytest.fixture(scope="module")
def sys1(cfg1):
    return T.SystemHandle(cfg1[0])


def _cos_f32(p0, p1, p4):
    """Cosine of the angle (p0, p1, p4) as the FP32 kernels form it: differences
    in FP64 rounded to FP32, then dot products summed x, y, z without fusion."""
    f = np.float32
    ba = [f(a - b) for a, b in zip(p0, p1)]
    bc = [f(a - b) for a, b in zip(p4, p1)]
    dot = lambda u, v: f(f(f(u[0] * v[0]) + f(u[1] * v[1])) + f(u[2] * v[2]))
    ia, ib = f(f(1) / dot(ba, ba)), f(f(1) / dot(bc, bc))
    return f(dot(ba, bc) * np.sqrt(f(ia * ib)))


def _straight_state(d, outcome):
    """The initial state with bead b0 of a kept HB candidate moved onto the line
    b4 -> b1, beyond b1, at a distance whose FP32 cosine rounds below -1."""
    pos = d["pos"].reshape(-1, 3).copy()
    beads = np.asarray(d["hb_beads"]).reshape(-1, 6)
    for g in np.where(outcome == 1)[0]:
        b0, b1, b4 = beads[g, 0], beads[g, 1], beads[g, 4]
        if len({b0, b1, b4}) < 3:
            continue
        u = pos[b1] - pos[b4]
        r0 = np.linalg.norm(pos[b0] - pos[b1])
        for t in np.linspace(0.9, 1.1, 201):
            p0 = pos[b1] + u * (t * r0 / np.linalg.norm(u))
            if _cos_f32(p0, pos[b1], pos[b4]) < np.float32(-1):
                pos[b0] = p0
                return pos, int(g)
    raise AssertionError("no candidate placement rounds past -1")


def test_straight_hb_angle_keeps_the_run_finite(lib, sys1):
    d = sys1.d
    eng = T.Engine(lib, [sys1], precision=T.PREC_FP64, rng=T.RNG_PHILOX, seed=2)
    eng.eval(T.TERM_HB)
    pos, g = _straight_state(d, eng.hb_outcome())
    eng.close()
    zero = np.zeros_like(pos)
    for prec in (T.PREC_MIXED, T.PREC_FP64):
        eng = T.Engine(lib, [sys1], precision=prec, rng=T.RNG_PHILOX, seed=2)
        eng.set_state(pos, pos, zero, klok=5)
        f, e = eng.eval(T.TERM_HB)
        assert eng.hb_outcome()[g] != -1, prec       # the candidate takes part
        assert np.isfinite(f).all(), prec
        assert np.isfinite(e["e2_hb"]) and np.isfinite(e["e3_hb"]), prec
        eng.set_state(pos, pos, zero, klok=5)
        eng.step(100)
        x, xo, v = eng.get_state()
        assert np.isfinite(x).all() and np.isfinite(v).all(), prec
        eng.close()


def test_batched_mixed_trajectory_passes_the_straight_angle(lib):
    """The 16-replica batch (c_Mg = 0.5*40^(k/15) mM, box 350+100k/15 A) in
    mixed precision with Philox seed 3: replica 8 formed a straight HB angle at
    step 702 and became NaN before the clamp."""
    hs = []
    for k in range(16):
        dk, _ = T.build_16scd(lib, T=0.596, c_mg=0.5e-3 * 40.0 ** (k / 15.0), c_k=0.050, seed=k + 1,
                              box=350.0 + 100.0 * k / 15.0)
        hs.append(T.SystemHandle(dk))
    eng = T.Engine(lib, hs, precision=T.PREC_MIXED, rng=T.RNG_PHILOX, seed=3, hb_groups=4, graph_steps=100,
                   pair_lanes=8)
    eng.step(1500)
    for r in range(16):
        x, _, v = eng.get_state(r)
        assert np.isfinite(x).all() and np.isfinite(v).all(), r
    eng.close()


def test_coincident_particles_stop_the_run(lib, sys1):
    d = sys1.d
    pos = d["pos"].reshape(-1, 3).copy()
    n = d["n_particles"]
    pos[n - 1] = pos[n - 2]                      # two ions on one point: a NaN pair force
    eng = T.Engine(lib, [sys1], precision=T.PREC_FP64, rng=T.RNG_PHILOX, seed=2)
    eng.set_state(pos, pos, np.zeros_like(pos), klok=1)
    with pytest.raises(T.TisError) as ei:
        eng.step(5)
    assert ei.value.code == ENONFINITE and "non-finite" in str(ei.value)
    eng.close()


def test_non_finite_input_is_refused(lib, sys1):
    pos = sys1.d["pos"].reshape(-1, 3).copy()
    pos[7, 1] = np.nan
    eng = T.Engine(lib, [sys1], precision=T.PREC_FP64)
    with pytest.raises(T.TisError) as ei:
        eng.set_state(pos, pos, np.zeros_like(pos))
    assert ei.value.code == -1 and "non-finite" in str(ei.value)
    eng.close()


EOVERFLOW = -5


def test_exhausted_reference_stream_stops_the_run(lib, sys1):
    """A reference-stream run fed 6,000 glibc values, fewer than two steps draw
    (the HB resolution's draws, then 3,000 for the step's noise table): draws
    at or past the end read 0.5 inside the kernels (every read stays in the
    stream's region) and the call stops with TIS_EOVERFLOW, naming the
    exhausted stream, instead of running on."""
    eng = T.Engine(lib, [sys1], precision=T.PREC_FP64, rng=T.RNG_STREAM)
    eng.set_stream(T.glibc_stream(lib, 31, 6000))
    with pytest.raises(T.TisError) as ei:
        eng.step(10)
    assert ei.value.code == EOVERFLOW and "stream exhausted" in str(ei.value)
    eng.close()
