"""GPU parity of the HIP engine against the FP64 CPU oracle (same inputs).

Bars: FP64 mode, per-term forces within 1e-10 relative (norm-wise) of the
oracle; mixed mode (FP32 force arithmetic) within 1e-5 relative (norm-wise,
the target stated for this path) and max |dF| / RMS(F) below 1e-4; pair sets
and HB candidate statuses bit-exact; HB valence resolution fed the same glibc
stream consumes the same number of draws.
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
def sys1(cfg1):
    d, _ = cfg1
    return T.SystemHandle(d)


@pytest.fixture(scope="module")
def oracle1(sys1):
    o = O.Oracle(sys1)
    d = sys1.d
    o.set_state(d["pos"], d["pos_old"], d["vel"])
    return o


@pytest.mark.parametrize("name,mask", TERMS)
def test_terms_fp64(lib, sys1, oracle1, name, mask):
    eng = T.Engine(lib, [sys1], precision=T.PREC_FP64)
    f, e = eng.eval(mask)
    _, fo, eo = oracle1.eval(mask)
    print(name, "rel", rel(f, fo))
    assert rel(f, fo) < 1e-10
    for k in ("e2_stack", "e3_stack", "e_wca", "e_coul_sr", "e_coul_lr", "e_bond", "e_angle"):
        assert e[k] == pytest.approx(eo[k], rel=1e-10, abs=1e-9), k
    eng.close()


@pytest.mark.parametrize("name,mask", TERMS)
def test_terms_mixed(lib, sys1, oracle1, name, mask):
    eng = T.Engine(lib, [sys1], precision=T.PREC_MIXED)
    f, _ = eng.eval(mask)
    _, fo, _ = oracle1.eval(mask)
    rms = np.sqrt(np.mean(np.sum(fo ** 2, axis=1)))
    print(name, "rel", rel(f, fo), "max/rms", np.abs(f - fo).max() / rms)
    assert rel(f, fo) < 1e-5
    assert np.abs(f - fo).max() / rms < 1e-4
    eng.close()


def test_pair_sets_and_hb_status(lib, sys1, oracle1):
    eng = T.Engine(lib, [sys1])
    assert np.array_equal(eng.pairs(0), oracle1.pairs(0))
    assert np.array_equal(eng.pairs(1), oracle1.pairs(1))
    assert np.array_equal(eng.hb_status(), oracle1.hb_status())
    eng.close()


def test_pair_sets_with_particles_one_wrap_leaves_outside_the_box(lib, sys1):
    """The cell sort wraps each coordinate once (separate_in_cells,
    fast.cpp:1510-1530), so a particle two box lengths away stays outside the
    box and its cell is clamped to the grid: it does not lie inside its cell,
    and the list search cut to the cells within the list radius of each
    particle must fall back to whole cells for the replica.  Two ions placed so
    (+2 L in x, -1.5 L in y): SR and LR pair sets bit-exact and pair forces
    within 1e-10 (FP64) of the oracle at the same state."""
    d = sys1.d
    L = np.asarray(d["box"], dtype=np.float64)
    pos = np.array(d["pos"], dtype=np.float64).reshape(-1, 3)
    old = np.array(d["pos_old"], dtype=np.float64).reshape(-1, 3)
    vel = np.array(d["vel"], dtype=np.float64).reshape(-1, 3)
    k1, k2 = d["n_beads"], d["n_particles"] - 1
    for k, ax, s in ((k1, 0, 2.0), (k2, 1, -1.5)):
        pos[k, ax] += s * L[ax]
        old[k, ax] += s * L[ax]
    o = O.Oracle(sys1)
    o.set_state(pos, old, vel)
    eng = T.Engine(lib, [sys1], precision=T.PREC_FP64)
    eng.set_state(pos, old, vel)
    assert np.array_equal(eng.pairs(0), o.pairs(0))
    assert np.array_equal(eng.pairs(1), o.pairs(1))
    for name, mask in (("sr", T.TERM_SR), ("lr", T.TERM_LR)):
        f, _ = eng.eval(mask)
        _, fo, _ = o.eval(mask)
        print(name, "rel", rel(f, fo))
        assert rel(f, fo) < 1e-10
    eng.close()


@pytest.mark.parametrize("precision,tol", [(T.PREC_FP64, 1e-10), (T.PREC_MIXED, 1e-5)])
def test_hb_with_reference_stream(lib, sys1, precision, tol):
    raw = O.glibc_stream(12345, 4000)
    eng = T.Engine(lib, [sys1], precision=precision, rng=T.RNG_STREAM)
    eng.set_stream(raw)
    f, e = eng.eval(T.TERM_HB)
    o = O.Oracle(sys1)
    d = sys1.d
    o.set_state(d["pos"], d["pos_old"], d["vel"])
    o.set_stream(raw)
    _, fo, eo = o.eval(T.TERM_HB)
    print("hb rel", rel(f, fo), "cursor", eng.cursor(), o.cursor())
    assert eng.cursor() == o.cursor()
    assert rel(f, fo) < tol
    assert e["e2_hb"] == pytest.approx(eo["e2_hb"], rel=tol)
    assert e["e3_hb"] == pytest.approx(eo["e3_hb"], rel=tol)
    eng.close()


def test_trajectory_reference_stream_fp64(lib, sys1):
    """50 full steps (noise on) with the glibc stream: positions track the oracle."""
    raw = O.glibc_stream(777, 50 * 4200)
    eng = T.Engine(lib, [sys1], rng=T.RNG_STREAM, use_graph=0)
    eng.set_stream(raw)
    o = O.Oracle(sys1)
    d = sys1.d
    o.set_state(d["pos"], d["pos_old"], d["vel"])
    o.set_stream(raw)
    for chunk in range(5):
        eng.step(10)
        assert o.step(10, 1.0) == 0
        p, _, v = eng.get_state()
        po, _, vo = o.get_state()
        err = np.abs(p - po).max()
        print("step", 10 * (chunk + 1), "max |dx|", err, "cursor", eng.cursor(), o.cursor())
        assert eng.cursor() == o.cursor()
        assert err < 1e-8
    e = eng.energies()
    assert e["n_rebuilds"] == o.rebuilds()
    eng.close()


def test_philox_run_is_stable(lib, sys1):
    eng = T.Engine(lib, [sys1], precision=T.PREC_MIXED)
    eng.step(500)
    p, _, v = eng.get_state()
    e = eng.energies()
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
    assert e["klok"] == 501
    assert 30.0 < e["rg"] < 50.0
    eng.close()


def test_rng_epoch_selects_fresh_draws(lib, sys1):
    """Philox draws are keyed by (seed, replica, epoch) and counted by klok: a run
    restarted at the same klok repeats its draws in the same epoch and draws
    afresh in another (the _DD second deck restarts klok at 1, DD.cpp:6428)."""
    d = sys1.d
    state = [d[k].reshape(-1, 3) for k in ("pos", "pos_old", "vel")]
    eng = T.Engine(lib, [sys1], precision=T.PREC_FP64, use_graph=0)
    runs = []
    for epoch in (0, 1, 0):
        eng.set_rng_epoch(epoch)
        eng.set_state(*state, klok=1)
        eng.step(5)
        runs.append(eng.get_state()[2])
    same, other = np.abs(runs[2] - runs[0]).max(), np.abs(runs[1] - runs[0]).max()
    print("same epoch", same, "other epoch", other)
    assert same < 1e-9
    assert other > 1e-3
    eng.close()


@pytest.mark.parametrize("prec", [T.PREC_FP64, T.PREC_MIXED])
def test_one_three_pair_at_the_wca_singularity_stays_finite(lib, sys1, prec):
    """P 833 and B 835 (a 1-3 pair through S 834) placed 1e-4 A beyond the point
    where the shifted WCA distance r + 1.5852 - D_LJ (D_LJ = 3.2) vanishes.  The
    reference adds that pair's WCA in the pair sum and removes it in the bonded
    terms; at this distance both are ~1e58 (inf in FP32).  The step path leaves the
    pair out of the sum, so the trajectory stays finite.  A weak P-S-B angle
    (k = 5 kcal/mol/rad^2) lets a batched FP32 run reach this point."""
    d = sys1.d
    pos, old, vel = (d[k].reshape(-1, 3).copy() for k in ("pos", "pos_old", "vel"))
    assert tuple(d["type"][833:836]) == (0, 1, 4)
    u = pos[833] - pos[835]
    u /= np.linalg.norm(u)
    shift = pos[835] + u * (3.2 - 1.5852 + 1e-4) - pos[833]
    pos[833] += shift
    old[833] += shift
    eng = T.Engine(lib, [sys1], precision=prec, use_graph=0)
    eng.set_state(pos, old, vel)
    f, _ = eng.eval(T.TERM_SR)          # the reference's split: the pair's own WCA
    assert not np.isfinite(np.linalg.norm(f[833])) or np.linalg.norm(f[833]) > 1e40
    eng.step(20)
    p, _, v = eng.get_state()
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
    assert np.linalg.norm(p[833] - p[835]) < 6.0
    eng.close()


def test_batched_replicas_equal_single_runs(lib):
    """Three uneven replicas in one context give what each gives alone
    (reference-stream mode, FP64: same arithmetic, same draws)."""
    systems = []
    for k, (c, box) in enumerate([(0.0005, 350.0), (0.004, 400.0), (0.02, 450.0)]):
        d, _ = T.build_16scd(lib, T=0.596, c_mg=c, c_k=0.050, seed=k + 1, box=box)
        systems.append(T.SystemHandle(d))
    raws = [O.glibc_stream(100 + k, 30 * 4200) for k in range(3)]
    batch = T.Engine(lib, systems, rng=T.RNG_STREAM, use_graph=0)
    for k in range(3):
        batch.set_stream(raws[k], rep=k)
    batch.step(30)
    for k in range(3):
        one = T.Engine(lib, [systems[k]], rng=T.RNG_STREAM, use_graph=0)
        one.set_stream(raws[k])
        one.step(30)
        pb, _, vb = batch.get_state(k)
        po, _, vo = one.get_state()
        print(k, systems[k].d["n_particles"], np.abs(pb - po).max())
        assert batch.cursor(k) == one.cursor()
        assert np.abs(pb - po).max() < 1e-9
        one.close()
    batch.close()


@pytest.fixture(scope="module")
def high_salt(lib):
    """The 450 A, 20 mM MgCl2 + 100 mM KCl system (14,957 particles) and the FP64
    oracle at its set-up state."""
    d, _ = T.build_16scd(lib, T=0.596, c_mg=0.020, c_k=0.100, seed=1, box=450.0)
    assert d["n_particles"] == 14957
    h = T.SystemHandle(d)
    o = O.Oracle(h)
    o.set_state(d["pos"], d["pos_old"], d["vel"])
    return h, o


def test_high_salt_450A_lists_and_status(lib, high_salt):
    """Pair sets and HB candidate statuses bit-exact."""
    h, o = high_salt
    eng = T.Engine(lib, [h], precision=T.PREC_MIXED)
    assert np.array_equal(eng.pairs(0), o.pairs(0))
    assert np.array_equal(eng.pairs(1), o.pairs(1))
    assert np.array_equal(eng.hb_status(), o.hb_status())
    eng.close()


@pytest.mark.parametrize("name,mask", TERMS + [("hb", T.TERM_HB)])
def test_high_salt_450A_per_term_mixed(lib, high_salt, name, mask):
    """Per-term FP32 forces of the 450 A, 20 mM system against the FP64 oracle
    (SR_/LR_interactions_lists fast.cpp:1783-1907, stacking_interactions
    :1911-2141, polymer_constraints :2354-2520, hydrogen_bonds_interactions
    :2145-2350): norm-wise within 1e-5 and max |dF| / RMS(F) within 1e-4.  The HB
    term runs with the reference's glibc stream on both sides (same draw count,
    so the valence resolution takes the same branch)."""
    h, o = high_salt
    d = h.d
    stream = name == "hb"
    eng = T.Engine(lib, [h], precision=T.PREC_MIXED, rng=T.RNG_STREAM if stream else T.RNG_PHILOX)
    if stream:
        raw = O.glibc_stream(4242, 20000)
        eng.set_stream(raw)
        o.set_state(d["pos"], d["pos_old"], d["vel"])
        o.set_stream(raw)
    f, _ = eng.eval(mask)
    _, fo, _ = o.eval(mask)
    rms = np.sqrt(np.mean(np.sum(fo ** 2, axis=1)))
    r_norm, r_max = rel(f, fo), np.abs(f - fo).max() / rms
    print(f"high salt {name}: |dF|/|F| = {r_norm:.3e}, max|dF|/RMS(F) = {r_max:.3e}")
    if stream:
        assert eng.cursor() == o.cursor() and eng.cursor() > 0
    assert np.linalg.norm(fo) > 0
    assert r_norm < 1e-5
    assert r_max < 1e-4
    eng.close()


@pytest.mark.parametrize("gsteps", [50, 25, 7])
def test_graph_replays_match_eager_steps(lib, sys1, gsteps):
    """hipGraph replays give the same trajectory as eager launches.  Each step
    writes the other of two state buffers, so a graph of an odd number of steps
    (25, 7) ends on the other buffer and the next replay uses the graph captured
    for that one; 50-step calls with a 7-step graph also run an eager remainder
    of one step between replays."""
    raw = O.glibc_stream(31, 100 * 4200)
    a = T.Engine(lib, [sys1], rng=T.RNG_STREAM, use_graph=1, graph_steps=gsteps)
    b = T.Engine(lib, [sys1], rng=T.RNG_STREAM, use_graph=0)
    a.set_stream(raw)
    b.set_stream(raw)
    for _ in range(2):
        a.step(50)
        b.step(50)
        pa, _, _ = a.get_state()
        pb, _, _ = b.get_state()
        assert a.cursor() == b.cursor()
        assert np.abs(pa - pb).max() < 1e-9
    cnt = a.launch_counts()
    assert cnt["graph_steps"] == 2 * (50 // gsteps) * gsteps and cnt["eager_steps"] == 2 * (50 % gsteps)
    a.close()
    b.close()


@pytest.fixture(scope="module")
def sequential_hb(lib, sys1):
    """Kept counts and resolved HB energies of 1500 resolutions of the reference's
    sequential process (oracle, glibc draws) at the configuration of sys1."""
    n = 1500
    o = O.Oracle(sys1)
    d = sys1.d
    o.set_state(d["pos"], d["pos_old"], d["vel"])
    kept_o, e_o = 0, []
    for s in range(n):
        o.set_stream(T.glibc_stream(lib, 5000 + s, 5000))
        _, _, e = o.eval(T.TERM_HB)
        kept_o = kept_o + (o.hb_outcome() == 1)
        e_o.append(e["e2_hb"] + e["e3_hb"])
    return kept_o, np.array(e_o), o.hb_iterations()


def test_component_resolution_has_the_sequential_distribution(lib, sys1, sequential_hb):
    """Production HB valence resolution (independent components on separate
    wavefronts, Philox draws) against the reference's sequential process (oracle,
    glibc draws) on the same configuration: for every candidate, the frequency
    with which it is kept over 1500 independent resolutions must agree
    (two-proportion z-score below 5 for every candidate; about 1,000 candidates
    are tested), and so must the mean of the resolved HB energy (4 standard
    errors).  The serial depth (iterations of the busiest wavefront) must drop
    well below the sequential iteration count."""
    n = 1500
    eng = T.Engine(lib, [sys1], precision=T.PREC_FP64, rng=T.RNG_PHILOX, seed=11)
    kept_g, e_g, rounds = 0, [], []
    for s in range(n):
        eng.set_step_counter(1000 + s)
        _, e = eng.eval(T.TERM_HB)
        kept_g = kept_g + (eng.hb_outcome() == 1)
        e_g.append(e["e2_hb"] + e["e3_hb"])
        if s < 50:
            rounds.append(eng.energies()["hb_rounds"])
    eng.close()
    kept_o, e_o, iters = sequential_hb
    pg, po = kept_g / n, kept_o / n
    pp = (kept_g + kept_o) / (2 * n)
    var = pp * (1 - pp) * 2 / n
    z = np.where(var > 0, (pg - po) / np.sqrt(np.where(var > 0, var, 1)), 0.0)
    e_g = np.array(e_g)
    se = np.sqrt(e_g.var() / n + e_o.var() / n)
    contested = int(np.sum((pp > 0) & (pp < 1)))
    print("contested candidates", contested, "max |z|", np.abs(z).max(), "mean E gpu", e_g.mean(),
          "ref", e_o.mean(), "se", se, "rounds", np.mean(rounds), max(rounds), "sequential iterations", iters)
    assert contested > 50
    assert np.abs(z).max() < 5.0
    assert abs(e_g.mean() - e_o.mean()) < 4 * se
    assert max(rounds) < iters / 2


def test_philox_resolution_is_deterministic(lib, sys1):
    a = T.Engine(lib, [sys1], rng=T.RNG_PHILOX, seed=5)
    b = T.Engine(lib, [sys1], rng=T.RNG_PHILOX, seed=5)
    fa, ea = a.eval(T.TERM_HB)
    fb, eb = b.eval(T.TERM_HB)
    assert ea["e2_hb"] == eb["e2_hb"] and ea["e3_hb"] == eb["e3_hb"]
    assert np.abs(fa - fb).max() < 1e-12
    a.close()
    b.close()


def _replica(lib, T_kcal=0.596, c_mg=0.005, seed=1):
    d, _ = T.build_16scd(lib, T=T_kcal, c_mg=c_mg, c_k=0.050, seed=seed, box=350.0)
    return T.SystemHandle(d)


def _energy_parts(eng, rep, mask):
    _, e = eng.eval(mask, rep=rep)
    nonhb = sum(e[k] for k in ("e2_stack", "e3_stack", "e_wca", "e_coul_sr", "e_coul_lr", "e_bond", "e_angle"))
    return nonhb, e["e2_hb"] + e["e3_hb"]


def test_exchange_exponent_matches_energies_at_two_temperatures(lib):
    """Two replicas of the 5 mM system at T = 0.596 and 0.640 kcal/mol (lB and the
    secondary-stack energies depend on T).  The on-device attempt's exponent
    D = (U_0(x_1) - U_0(x_0))/T_0 + (U_1(x_0) - U_1(x_1))/T_1 + (H(x_1) - H(x_0))(1/T_0 - 1/T_1)
    must equal the one composed from per-term evaluations of the same states
    (each per-term energy is itself checked against the FP64 oracle above); an
    accepted attempt swaps the configurations with velocities scaled by
    sqrt(T_to / T_from), a rejected one leaves both replicas bit-identical."""
    temps = (0.596, 0.640)
    hs = [_replica(lib, t) for t in temps]
    eng = T.Engine(lib, hs, precision=T.PREC_FP64, seed=3, use_graph=0)
    eng.step(30)
    st = [eng.get_state(r) for r in range(2)]
    kl = [eng.energies(r)["klok"] for r in range(2)]
    host = T.Engine(lib, hs, precision=T.PREC_FP64, seed=3, use_graph=0)
    self_e = []
    for r in range(2):
        host.set_state(*st[r], klok=kl[r], rep=r)
        self_e.append(_energy_parts(host, r, T.TERM_ALL))
    cross = []
    for r in range(2):
        host.set_state(*st[1 - r], klok=kl[r], rep=r)
        cross.append(_energy_parts(host, r, T.TERM_ALL & ~T.TERM_HB)[0])
    b0, b1 = 1.0 / temps[0], 1.0 / temps[1]
    d_host = (b0 * (cross[0] - self_e[0][0]) + b1 * (cross[1] - self_e[1][0])
              + (self_e[1][1] - self_e[0][1]) * (b0 - b1))
    eng.exchange_now()
    att, acc, d = eng.exchange_stats()
    print("D host", d_host, "device", d[0], "accepted", acc[0])
    assert att[0] == 1
    assert d[0] == pytest.approx(d_host, rel=1e-9, abs=1e-7)
    after = [eng.get_state(r) for r in range(2)]
    if acc[0] == 1:
        s01 = np.sqrt(temps[0] / temps[1])
        assert np.abs(after[0][0] - st[1][0]).max() == 0.0
        assert np.abs(after[0][2] - s01 * st[1][2]).max() < 1e-12
        assert list(eng.walkers()) == [1, 0]
    else:
        for r in range(2):
            for a, b in zip(after[r], st[r]):
                assert np.array_equal(a, b)
        assert list(eng.walkers()) == [0, 1]
    eng.close()
    host.close()


def test_exchange_between_equal_temperatures_always_swaps(lib):
    """Same parameters in both slots: the non-HB terms cancel and the HB term has
    a zero factor, so D is rounding noise and the attempt is accepted; the slots
    then hold each other's configurations (velocity scale 1)."""
    h = _replica(lib)
    eng = T.Engine(lib, [h, h], precision=T.PREC_FP64, seed=9, use_graph=0)
    eng.step(5)
    a = eng.get_state(0)
    other = T.Engine(lib, [h], precision=T.PREC_FP64, seed=21, use_graph=0)
    other.step(40)
    b = other.get_state(0)
    eng.set_state(*b, klok=6, rep=1)
    eng.exchange_now()
    att, acc, d = eng.exchange_stats()
    assert att[0] == 1 and acc[0] == 1 and abs(d[0]) < 1e-6
    p0, o0, v0 = eng.get_state(0)
    p1, o1, v1 = eng.get_state(1)
    assert np.array_equal(p0, b[0]) and np.array_equal(p1, a[0])
    assert np.array_equal(v0, b[2]) and np.array_equal(v1, a[2])
    assert np.abs(o0 - b[1]).max() < 1e-12 and np.abs(o1 - a[1]).max() < 1e-12
    assert list(eng.walkers()) == [1, 0]
    eng.close()
    other.close()


def test_exchange_cadence_in_steps_and_incompatible_neighbours(lib):
    """exchange_every = 50: 200 steps make 4 attempts, alternating the pairs
    (0,1),(2,3) and (1,2); a neighbour with another [Mg] (other particle count)
    is never paired.  Graph-captured chunks run between the attempts."""
    hs = [_replica(lib, t) for t in (0.596, 0.605, 0.615, 0.625)] + [_replica(lib, 0.596, c_mg=0.010)]
    eng = T.Engine(lib, hs, precision=T.PREC_MIXED, seed=4, use_graph=1, graph_steps=25, exchange_every=50)
    eng.step(200)
    att, acc, _ = eng.exchange_stats()
    print("attempts", att, "accepts", acc, "walkers", eng.walkers())
    assert list(att) == [2, 2, 2, 0]
    assert sorted(eng.walkers()[:4]) == [0, 1, 2, 3] and eng.walkers()[4] == 4
    for r in range(5):
        p, _, v = eng.get_state(r)
        assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
    eng.close()


@pytest.mark.parametrize("prec,tol", [(T.PREC_FP64, 1e-12), (T.PREC_MIXED, 1e-6)])
def test_step_stack_energies_equal_an_evaluation(lib, prec, tol):
    """In a step the stacks of a replica are spread over several workgroups and
    their energy sums added up after the streams join, and the HB energy sums of
    the resolution groups are added up when the step closes; the E2 / E3 stack
    and HB energies a step reports are those of an evaluation of the state it
    started from at the same step counter (two replicas of different systems in
    one batch; Philox draws depend on the step counter only)."""
    hs = []
    for c, box, seed in ((0.005, 350.0, 1), (0.02, 450.0, 2)):
        d, _ = T.build_16scd(lib, T=0.596, c_mg=c, c_k=0.050, seed=seed, box=box)
        hs.append(T.SystemHandle(d))
    eng = T.Engine(lib, hs, precision=prec, rng=T.RNG_PHILOX, seed=9, use_graph=0)
    eng.step(20)
    want = []
    for r in range(2):
        _, e = eng.eval(T.TERM_STACK | T.TERM_HB, rep=r)
        want.append((e["e2_stack"], e["e3_stack"], e["e2_hb"], e["e3_hb"]))
    eng.step(1)
    for r in range(2):
        e = eng.energies(r)
        assert e["e2_stack"] == pytest.approx(want[r][0], rel=tol), r
        assert e["e3_stack"] == pytest.approx(want[r][1], rel=tol), r
        assert e["e2_hb"] == pytest.approx(want[r][2], rel=1e-12), r
        assert e["e3_hb"] == pytest.approx(want[r][3], rel=1e-12), r
        assert all(w != 0.0 for w in want[r])
    eng.close()


@pytest.mark.parametrize("prec,bar", [(T.PREC_FP64, 1e-10), (T.PREC_MIXED, 1e-4)])
def test_noise_free_trajectory_tracks_the_reference(lib, sys1, prec, bar):
    """200 noise-free steps (friction and forces only; the HB valence resolution
    still draws from the same glibc stream) against the FP64 oracle: FP64 tracks
    to rounding; FP32 force arithmetic drifts by its rounding of each step's
    forces, amplified over the 200 steps (bar in A, largest coordinate
    difference, printed per 20 steps; measured 7e-13 for FP64, bit-identical
    for the first 140 steps, and 8e-6 for FP32 forces)."""
    raw = O.glibc_stream(4242, 200 * 4200)
    eng = T.Engine(lib, [sys1], precision=prec, rng=T.RNG_STREAM, noise_scale=0.0, use_graph=0)
    eng.set_stream(raw)
    o = O.Oracle(sys1)
    d = sys1.d
    o.set_state(d["pos"], d["pos_old"], d["vel"])
    o.set_stream(raw)
    worst = 0.0
    for chunk in range(10):
        eng.step(20)
        assert o.step(20, 0.0) == 0
        p, _, _ = eng.get_state()
        po, _, _ = o.get_state()
        err = np.abs(p - po).max()
        worst = max(worst, err)
        print("step", 20 * (chunk + 1), "max |dx|", err, flush=True)
        assert eng.cursor() == o.cursor()
    assert worst < bar
    eng.close()
