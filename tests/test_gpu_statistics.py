"""Statistical equality of the production dynamics and the reference's process.

From one start state (the 5 mM 16SCD set-up), independent trajectories of 400
steps are run two ways:
  - device, production settings: FP32 forces / FP64 state, Philox noise per
    particle, component-parallel HB resolution; 64 replicas of the system in one
    batch (each replica has its own Philox key);
  - oracle: the reference's algorithm in FP64 with glibc rand() streams of
    different seeds (sequential HB resolution, the reference's noise-slot sharing).
At step 400 each trajectory gives observables of its state: the mean-square
displacement of Mg, Cl and K ions and of the beads (the integrator's friction,
time step and noise amplitude), the kinetic energy per particle of ions and
beads, Rg, the number of Mg ions bound to phosphates, and the energy terms of
the state.  The two ensembles (64 and 16 trajectories) must agree in mean: Welch
|z| < 4.5 for every one of the 13 observables (the per-trajectory values are
independent; with about 15 degrees of freedom on the oracle side the chance of a
false alarm over all 13 is about 0.5 %).
"""
import os

import numpy as np
import pytest

import observables as OB
import oracle_ctypes as O
import tis_ctypes as T

pytestmark = pytest.mark.gpu

STEPS = 400
N_GPU = 64
N_CPU = 16


def _per_state(d, x0, pos, vel, energies):
    box = np.asarray(d["box"], dtype=float)
    cls = np.asarray(d["cls"])
    masses = np.asarray(d["type_mass"])[np.asarray(d["type"])]
    disp = OB.min_image(pos - x0, box)
    msd = np.einsum("ij,ij->i", disp, disp)
    ke = 0.5 * masses * np.einsum("ij,ij->i", vel, vel)
    beads = np.arange(d["n_beads"])
    out = {
        "msd_beads": msd[beads].mean(),
        "msd_mg": msd[cls == OB.CLS_MG].mean(),
        "msd_cl": msd[cls == OB.CLS_CL].mean(),
        "msd_k": msd[cls == OB.CLS_K].mean(),
        "ke_beads": ke[beads].mean(),
        "ke_ions": ke[cls >= OB.CLS_MG].mean(),
        "rg": OB.radius_of_gyration(pos[beads]),
        "bound_mg": OB.bound_totals(pos, cls, box)[0],
        "e_hb": energies["e2_hb"] + energies["e3_hb"],
        "e_stack": energies["e2_stack"] + energies["e3_stack"],
        "e_wca": energies["e_wca"],
        "e_coul": energies["e_coul_sr"] + energies["e_coul_lr"],
        "e_backbone": energies["e_bond"] + energies["e_angle"],
    }
    return out


def test_ensemble_at_fixed_time_matches_the_reference_process(lib, cfg1):
    d = cfg1[0]
    h = T.SystemHandle(d)
    x0 = d["pos"].reshape(-1, 3)
    eng = T.Engine(lib, [h] * N_GPU, precision=T.PREC_MIXED, rng=T.RNG_PHILOX, seed=2024)
    eng.step(STEPS)
    gpu = []
    for r in range(N_GPU):
        pos, _, vel = eng.get_state(r)
        _, e = eng.eval(T.TERM_ALL, rep=r)
        gpu.append(_per_state(d, x0, pos, vel, e))
    eng.close()
    cpu = []
    o = O.Oracle(h)
    for s in range(N_CPU):
        o.set_state(d["pos"], d["pos_old"], d["vel"])
        o.set_stream_glibc(101 + s, 4_000 * (STEPS + 20))
        o.step(STEPS)
        pos, _, vel = o.get_state()
        _, _, e = o.eval(T.TERM_ALL)
        cpu.append(_per_state(d, x0, pos.reshape(-1, 3), vel.reshape(-1, 3), e))
    worst = 0.0
    for k in gpu[0]:
        g = np.array([v[k] for v in gpu], dtype=float)
        c = np.array([v[k] for v in cpu], dtype=float)
        se = np.sqrt(g.var(ddof=1) / len(g) + c.var(ddof=1) / len(c))
        z = (g.mean() - c.mean()) / se if se > 0 else 0.0
        worst = max(worst, abs(z))
        print(f"{k:12s} device {g.mean():14.6f} +- {g.std(ddof=1):10.6f}   reference {c.mean():14.6f} "
              f"+- {c.std(ddof=1):10.6f}   z {z:+.2f}", flush=True)
        assert abs(z) < 4.5, k
    print("max |z|", worst)


# ---------------------------------------------------------------- long window
# tests/golden/stats_5mM_1e5.npz: 32 trajectories of the reference's process (the
# FP64 oracle with glibc streams, sequential HB resolution and the reference's
# noise-slot sharing) from the same start state, 10^5 steps each, sampled every
# 10^4 steps (tests/golden/make_stats_fixture.py).
STATS = os.path.join(os.path.dirname(__file__), "golden", "stats_5mM_1e5.npz")
N_LONG = 64
COLUMNS = ("rg", "bound_mg", "e_hb", "e_stack", "n_kept", "p_mg_1", "p_mg_2", "p_mg_3", "p_mg_4", "p_mg_5",
           "st3_formed")


def _rows(rg, bound_mg, e_hb, e_stack, n_kept, p_mg, st3, st3_E):
    """Observables of one state: the scalars, Mg contacts summed over five
    blocks of consecutive phosphates, and the tertiary stacks below half their
    stacking energy (Q < 1: formed)."""
    blocks = np.array_split(np.arange(p_mg.shape[-1]), 5)
    return [rg, bound_mg, e_hb, e_stack, n_kept] + [float(p_mg[b].sum()) for b in blocks] + \
        [float(np.count_nonzero(st3 < 0.5 * st3_E))]


@pytest.mark.skipif(not os.path.exists(STATS), reason="statistics fixture not generated")
def test_long_window_observables_match_the_reference_process(lib, cfg1):
    """Rg, bound Mg, Mg contacts along the chain, HB and stacking energies, kept
    HB candidates and formed tertiary stacks at 10 fixed times up to 10^5 steps
    (244 ps): 64 device trajectories (mixed precision, Philox noise per particle,
    component-parallel HB resolution) against the 32 reference-process
    trajectories of the fixture.  Every (time, observable) pair must agree in
    mean, Welch |z| < 5 (110 comparisons, each between independent
    trajectories: about 0.2 % chance of a false alarm in all with ~31 degrees
    of freedom on the reference side).  The 10^9-step equilibrium means the
    reference reports (analysis/Analysis_storedvalues_ave.txt) lie far beyond a
    test and stay unpinned."""
    ref = np.load(STATS)
    times = ref["times"]
    d = cfg1[0]
    st3_E = np.asarray(d["st3_E"])
    p_beads = np.nonzero(np.asarray(d["cls"])[: d["n_beads"]] == OB.CLS_P)[0]
    eng = T.Engine(lib, [T.SystemHandle(d)] * N_LONG, precision=T.PREC_MIXED, rng=T.RNG_PHILOX, seed=2024)
    done, worst = 0, 0.0
    for k, t in enumerate(times):
        eng.step(int(t) - done)
        done = int(t)
        eng.observe(0)
        eng.observe_now()
        g = []
        for r in range(N_LONG):
            o = eng.observables(r)
            g.append(_rows(o["rg"], o["bound_mg"], o["e2_hb"] + o["e3_hb"], o["e2_stack"] + o["e3_stack"],
                           o["hb_kept"].sum(), o["per_bead"][p_beads, 0], o["st3_mean"], st3_E))
        c = [_rows(ref["rg"][j, k], ref["bound_mg"][j, k], ref["e_hb"][j, k], ref["e_stack"][j, k],
                   ref["n_kept"][j, k], ref["p_mg"][j, k].astype(float), ref["st3"][j, k], st3_E)
             for j in range(ref["rg"].shape[0])]
        g, c = np.asarray(g, dtype=float), np.asarray(c, dtype=float)
        line = []
        for q, name in enumerate(COLUMNS):
            se = np.sqrt(g[:, q].var(ddof=1) / len(g) + c[:, q].var(ddof=1) / len(c))
            z = 0.0 if se == 0 else (g[:, q].mean() - c[:, q].mean()) / se
            worst = max(worst, abs(z))
            line.append(f"{name} {g[:, q].mean():.3f}/{c[:, q].mean():.3f} ({z:+.1f})")
            assert abs(z) < 5.0, (int(t), name, g[:, q].mean(), c[:, q].mean(), z)
        print(f"step {int(t)}: " + ", ".join(line), flush=True)
    print("max |z|", worst)
    eng.close()
