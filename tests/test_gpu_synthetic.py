"""GPU parity on the large synthetic chain: 1,500 nt in a 600 A box with
10 mM MgCl2 + 100 mM KCl (32,914 particles, 400 HB candidates, 868 secondary
stacks), against the FP64 CPU oracle on the same inputs.

Bars: pair sets and HB statuses bit-exact; mixed-precision (FP32 force
arithmetic) per-term forces within 1e-5 relative of the oracle (norm-wise) and,
per particle, |dF_i| below 1e-4 of max(|F_i|, RMS(F)); the HB valence
resolution fed the same glibc stream consumes the same draws; a Philox run
stays finite.  At 110 mM salt a few ion pairs sit at WCA contact with
|F_i| ~ 30 kcal/mol/A, 60 times the RMS force, so the per-particle bar is
scaled by each particle's own force (FP32 gives a few 1e-6 of it) rather than
by the RMS alone.
"""
import numpy as np
import pytest

import oracle_ctypes as O
import tis_ctypes as T

pytestmark = pytest.mark.gpu

TERMS = [("stack", T.TERM_STACK), ("backbone", T.TERM_BACKBONE), ("sr", T.TERM_SR), ("lr", T.TERM_LR)]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def chain(lib):
    """The built chain sits with every bond and angle at its minimum, where the
    backbone force is zero and a relative error means nothing; the beads are
    displaced by seeded N(0, 0.05 A) noise first (the walk keeps non-bonded
    beads 3.5 A apart, so no contact comes near the WCA core)."""
    d = T.build_synthetic(lib)
    assert d["n_particles"] == 32914
    rng = np.random.default_rng(11)
    pos = d["pos"].reshape(-1, 3).copy()
    pos[:d["n_beads"]] += rng.normal(0.0, 0.05, (d["n_beads"], 3))
    vel = d["vel"].reshape(-1, 3)
    state = (pos, pos - 0.05 * vel, vel)
    h = T.SystemHandle(d)
    o = O.Oracle(h)
    o.set_state(*state)
    return h, o, state


def engine_at(lib, h, state, **kw):
    eng = T.Engine(lib, [h], **kw)
    eng.set_state(*state)
    return eng


def test_chain_lists_and_terms_mixed(lib, chain):
    h, o, state = chain
    eng = engine_at(lib, h, state, precision=T.PREC_MIXED)
    assert np.array_equal(eng.pairs(0), o.pairs(0))
    assert np.array_equal(eng.pairs(1), o.pairs(1))
    assert np.array_equal(eng.hb_status(), o.hb_status())
    for name, mask in TERMS:
        f, _ = eng.eval(mask)
        _, fo, _ = o.eval(mask)
        mag = np.linalg.norm(fo, axis=1)
        rms = np.sqrt(np.mean(mag ** 2))
        worst = (np.linalg.norm(f - fo, axis=1) / np.maximum(mag, rms)).max()
        print(name, "rel", rel(f, fo), "per-particle", worst, flush=True)
        assert rel(f, fo) < 1e-5, name
        assert worst < 1e-4, name
    eng.close()


@pytest.mark.parametrize("precision,tol", [(T.PREC_FP64, 1e-10), (T.PREC_MIXED, 1e-5)])
def test_chain_hb_with_reference_stream(lib, chain, precision, tol):
    h, o, state = chain
    raw = O.glibc_stream(4242, 8000)
    eng = engine_at(lib, h, state, precision=precision, rng=T.RNG_STREAM)
    eng.set_stream(raw)
    f, e = eng.eval(T.TERM_HB)
    o.set_stream(raw)
    _, fo, eo = o.eval(T.TERM_HB)
    print("hb rel", rel(f, fo), "cursor", eng.cursor(), o.cursor(), flush=True)
    assert eng.cursor() == o.cursor() > 0
    assert rel(f, fo) < tol
    assert e["e2_hb"] == pytest.approx(eo["e2_hb"], rel=tol)
    assert e["e3_hb"] == pytest.approx(eo["e3_hb"], rel=tol)
    eng.close()


def test_chain_philox_run_is_stable(lib, chain):
    h, _, _ = chain
    eng = T.Engine(lib, [h], precision=T.PREC_MIXED)
    eng.step(1000)
    p, _, v = eng.get_state()
    e = eng.energies()
    print("rebuilds", e["n_rebuilds"], "hb active", e["n_hb_active"], "rg", e["rg"], flush=True)
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
    assert e["klok"] == 1001
    assert e["n_rebuilds"] > 0
    eng.close()


def test_chain_without_hb_candidates_or_stacks(lib):
    """A 60-nt chain with no HB candidates and no stacks (586 particles, 150 A
    box), its beads displaced by seeded N(0, 0.05 A) noise as in the chain
    fixture (the built chain has zero backbone force): the empty HB tables,
    resolution and stack slices must leave every term equal to the oracle's
    (FP64, 1e-10 relative; the HB and stack terms exactly zero), the pair sets
    bit-exact, and a Philox run of 500 steps in either precision finite with
    zero HB and stack energies."""
    d = T.build_synthetic(lib, n_nt=60, stack_frac=0.0, n_hb=0, radius=20.0, box=150.0)
    assert d["n_hb"] == 0 and d["n_st3"] == 0 and d["n_st2"] == 0
    rng = np.random.default_rng(12)
    pos = d["pos"].reshape(-1, 3).copy()
    pos[:d["n_beads"]] += rng.normal(0.0, 0.05, (d["n_beads"], 3))
    vel = d["vel"].reshape(-1, 3)
    state = (pos, pos - 0.05 * vel, vel)
    h = T.SystemHandle(d)
    o = O.Oracle(h)
    o.set_state(*state)
    print("case: no-HB chain, FP64 terms", flush=True)
    eng = engine_at(lib, h, state, precision=T.PREC_FP64)
    assert np.array_equal(eng.pairs(0), o.pairs(0))
    assert np.array_equal(eng.pairs(1), o.pairs(1))
    for name, mask in TERMS + [("hb", T.TERM_HB)]:
        f, e = eng.eval(mask)
        _, fo, eo = o.eval(mask)
        if name in ("hb", "stack"):
            assert not np.any(f) and not np.any(fo), name
        else:
            print(name, "rel", rel(f, fo))
            assert rel(f, fo) < 1e-10, name
    eng.close()
    for prec in (T.PREC_FP64, T.PREC_MIXED):
        print("case: no-HB chain, Philox run, precision", prec, flush=True)
        eng = T.Engine(lib, [h], precision=prec)
        eng.step(500)
        p, _, v = eng.get_state()
        e = eng.energies()
        assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
        assert e["klok"] == 501
        for k in ("e2_hb", "e3_hb", "e2_stack", "e3_stack"):
            assert e[k] == 0.0, k
        eng.close()
