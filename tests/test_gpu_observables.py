"""Device-accumulated observables (tis_observe / tis_observe_now /
tis_get_observables) against the observables oracle on the same states.

  - one sample: per-phosphate Mg / Cl / K contact counts and bound-ion totals
    bit-exact against the brute-force minimum-image count; Rg, D, S within
    1e-10 of the eigenvalue form of the gyration tensor; energies, HB outcomes
    and tertiary-stack energies equal to a separate evaluation of the same state;
  - a run sampled every 20 steps: the cadence, the means equal to the mean of
    per-state oracle values, and the same trajectory as an unsampled run (to the
    rounding of the force atomics).
"""
import numpy as np
import pytest

import observables as OB
import tis_ctypes as T

pytestmark = pytest.mark.gpu


def _engine(lib, d, **kw):
    return T.Engine(lib, [T.SystemHandle(d)], precision=kw.pop("precision", T.PREC_FP64), rng=T.RNG_PHILOX,
                    seed=7, **kw)


def _oracle_sample(d, pos):
    box = np.asarray(d["box"], dtype=float)
    counts, ps = OB.p_ion_counts(pos, d["cls"], box)
    bound = OB.bound_totals(pos, d["cls"], box)
    rg, D, S = OB.shape(pos[:d["n_beads"]])
    return counts, ps, bound, (rg, D, S)


@pytest.fixture(scope="module")
def system(cfg1):
    return cfg1[0]


def test_one_sample_matches_the_oracle(lib, system):
    d = system
    eng = _engine(lib, d)
    eng.step(300)
    eng.observe(0)
    before = (eng.energies(), eng.stack_energies(), eng.hb_outcome())
    eng.observe_now()
    # the sample leaves what the last step's force call reported (Energies and
    # Contacts file columns, HB outcomes) as it was
    after = (eng.energies(), eng.stack_energies(), eng.hb_outcome())
    for k in ("e2_hb", "e3_hb", "e2_stack", "e3_stack", "hb_rounds"):
        assert after[0][k] == before[0][k], k
    assert np.array_equal(after[1], before[1]) and np.array_equal(after[2], before[2])
    o = eng.observables()
    pos, _, _ = eng.get_state()
    counts, ps, bound, (rg, D, S) = _oracle_sample(d, pos)
    assert o["n_samples"] == 1
    pb = o["per_bead"]
    assert np.array_equal(pb[ps, :3], counts.astype(float))
    assert np.array_equal(pb[ps, 3], (counts[:, 0] > 0).astype(float))
    others = np.setdiff1d(np.arange(d["n_beads"]), ps)
    assert not pb[others].any()
    assert [o["bound_mg"], o["bound_cl"], o["bound_k"]] == bound.tolist()
    assert counts.sum() > 0 and bound.sum() > 0
    assert o["rg"] == pytest.approx(rg, rel=1e-10)
    assert o["rg2"] == pytest.approx(rg * rg, rel=1e-10)
    assert o["asphericity"] == pytest.approx(D, rel=1e-9, abs=1e-12)
    assert o["shape"] == pytest.approx(S, rel=1e-9, abs=1e-12)
    # the sample's energies and outcomes are those of an evaluation of this state
    _, e = eng.eval(T.TERM_ALL)
    kept = (eng.hb_outcome() == 1).astype(float)
    assert np.array_equal(o["hb_kept"], kept)
    assert np.array_equal(o["st3_mean"], eng.stack_energies())
    for k in ("e2_hb", "e3_hb", "e2_stack", "e3_stack", "e_wca", "e_coul_sr", "e_coul_lr", "e_bond", "e_angle"):
        assert o[k] == pytest.approx(e[k], rel=1e-12, abs=1e-12), k
    eng.close()


def test_cadence_means_and_unchanged_trajectory(lib, system):
    d = system
    a = _engine(lib, d, graph_steps=10)
    a.observe(20)
    a.step(200)
    oa = a.observables()
    assert oa["n_samples"] == 10
    # same run sampled by hand: identical accumulators, oracle means
    b = _engine(lib, d, graph_steps=10)
    b.observe(0)
    per, shp, bnd = [], [], []
    for _ in range(10):
        b.step(20)
        b.observe_now()
        pos, _, _ = b.get_state()
        counts, ps, bound, sh = _oracle_sample(d, pos)
        per.append(counts)
        bnd.append(bound)
        shp.append(sh)
    ob = b.observables()
    # multi-body forces and energy sums are reduced with FP64 atomics, whose
    # summation order varies, so two runs agree to rounding, not bit for bit
    for k in ("bound_mg", "bound_cl", "bound_k"):
        assert oa[k] == ob[k], k
    for k in ("rg", "rg2", "asphericity", "shape", "e2_hb", "e3_hb", "e2_stack", "e3_stack", "e_wca",
              "e_coul_sr", "e_coul_lr", "e_bond", "e_angle"):
        assert oa[k] == pytest.approx(ob[k], rel=1e-12, abs=1e-12), k
    assert np.array_equal(oa["per_bead"], ob["per_bead"])
    assert np.array_equal(oa["hb_kept"], ob["hb_kept"])
    assert np.allclose(oa["st3_mean"], ob["st3_mean"], rtol=1e-12, atol=1e-14)
    mean_counts = np.mean(per, axis=0)
    assert np.allclose(oa["per_bead"][ps, :3], mean_counts, rtol=0, atol=1e-12)
    assert np.allclose([oa["bound_mg"], oa["bound_cl"], oa["bound_k"]], np.mean(bnd, axis=0), rtol=0, atol=1e-12)
    shp = np.array(shp)
    assert oa["rg"] == pytest.approx(shp[:, 0].mean(), rel=1e-10)
    assert oa["asphericity"] == pytest.approx(shp[:, 1].mean(), rel=1e-9)
    # sampling leaves the trajectory alone (to the rounding of the atomics)
    c = _engine(lib, d, graph_steps=10)
    c.step(200)
    for x, y in zip(a.get_state(), c.get_state()):
        assert np.abs(x - y).max() < 1e-8
    for e in (a, b, c):
        e.close()


def test_observe_rejects_a_cut_beyond_the_lists(lib, system):
    eng = _engine(lib, system)
    with pytest.raises(T.TisError):
        eng.observe(10, cut=14.0)
    eng.observe(10, cut=T.R_KT)
    eng.close()


def test_batched_replicas_sample_independently(lib, system):
    """Two replicas of different salt in one batch: each replica's sample is its
    own state's (the second replica's indices are offset in the batch)."""
    d2, _ = T.build_16scd(lib, T=0.596, c_mg=0.010, c_k=0.050, seed=3, box=380.0)
    eng = T.Engine(lib, [T.SystemHandle(system), T.SystemHandle(d2)], precision=T.PREC_MIXED, rng=T.RNG_PHILOX,
                   seed=5)
    eng.step(100)
    eng.observe(0)
    eng.observe_now()
    for rep, d in enumerate((system, d2)):
        o = eng.observables(rep)
        pos, _, _ = eng.get_state(rep)
        counts, ps, bound, (rg, _, _) = _oracle_sample(d, pos)
        assert np.array_equal(o["per_bead"][ps, :3], counts.astype(float)), rep
        assert [o["bound_mg"], o["bound_cl"], o["bound_k"]] == bound.tolist(), rep
        assert o["rg"] == pytest.approx(rg, rel=1e-10)
    eng.close()
