"""The FP64 CPU oracle, pinned two ways.

1. Against the compiled reference (golden ref_cfg1.npz, made by
   tests/golden/make_fixtures.py from oracle/_ref): per-term forces and energies at
   the 5 mM / 350 A / T = 0.596 start state, the SR/LR pair sets, HB statuses and the
   HB valence resolution fed the same glibc stream, and a 200-step noisy trajectory.
2. Against what the mathematics fixes, independently of the reference: forces are
   minus the finite-difference gradient of the term energies, every term is
   translation invariant, the non-periodic terms are rotation invariant, and the
   cell-list pair sets equal a brute-force O(N^2) enumeration.
"""
import numpy as np
import pytest

import oracle_ctypes as O
import tis_ctypes as T

TERMS = {"stack": T.TERM_STACK, "hb": T.TERM_HB, "poly": T.TERM_BACKBONE, "SR": T.TERM_SR, "LR": T.TERM_LR}


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def handle(cfg1):
    return T.SystemHandle(cfg1[0])


@pytest.fixture()
def orc(handle):
    o = O.Oracle(handle)
    d = handle.d
    o.set_state(d["pos"], d["pos_old"], d["vel"])
    return o


@pytest.mark.parametrize("name,tol", [("stack", 1e-12), ("poly", 1e-13), ("SR", 1e-15), ("LR", 1e-15)])
def test_terms_match_reference(orc, golden_cfg1, name, tol):
    rc, f, _ = orc.eval(TERMS[name])
    assert rc == 0
    assert rel(f, golden_cfg1["F_" + name]) < tol


def test_hb_matches_reference(orc, golden_cfg1):
    g = golden_cfg1
    orc.set_stream(O.glibc_stream(int(g["hb_rng_seed"][0]), 5000))
    rc, f, e = orc.eval(T.TERM_HB)
    assert rc == 0
    assert orc.cursor() == int(g["hb_rand_calls"][0])
    assert rel(f, g["F_hb"]) < 1e-14
    assert e["e2_hb"] == pytest.approx(g["energies"][0], rel=1e-14)
    assert e["e3_hb"] == pytest.approx(g["energies"][1], rel=1e-14)


def test_energies_match_reference(orc, golden_cfg1):
    e_ref = golden_cfg1["energies"]
    _, _, es = orc.eval(T.TERM_STACK)
    assert es["e2_stack"] == pytest.approx(e_ref[2], rel=1e-13)
    assert es["e3_stack"] == pytest.approx(e_ref[3], rel=1e-13)
    _, _, e = orc.eval(T.TERM_SR)
    assert e["e_wca"] == pytest.approx(e_ref[4], rel=1e-12)
    assert e["e_coul_sr"] == pytest.approx(e_ref[5], rel=1e-13)
    _, _, e = orc.eval(T.TERM_LR)
    assert e["e_coul_lr"] == pytest.approx(e_ref[6], rel=1e-13)
    assert es["rg"] == pytest.approx(e_ref[7], rel=1e-14)
    assert np.allclose(orc.stack_energies(), golden_cfg1["st_energy"][:49], rtol=1e-13)


def test_lists_match_reference(orc, golden_cfg1):
    import hashlib
    assert np.array_equal(orc.pairs(0), golden_cfg1["sr_pairs"])
    lr = orc.pairs(1)
    assert len(lr) == int(golden_cfg1["lr_count"][0])
    assert hashlib.sha256(np.ascontiguousarray(lr.astype(np.int32)).tobytes()).hexdigest() == \
        golden_cfg1["lr_sha"][0]
    assert np.array_equal(orc.hb_status(), golden_cfg1["hb_status0"].astype(np.int32))


def test_trajectory_tracks_reference(handle, golden_cfg1):
    g = golden_cfg1
    d = handle.d
    o = O.Oracle(handle)
    o.set_state(d["pos"], d["pos_old"], d["vel"])
    o.set_stream(O.glibc_stream(int(g["traj_rng_seed"][0]), int(g["traj_draws"][-1]) + 10))
    for s in range(1, 201):
        assert o.step(1, 1.0) == 0
        assert o.cursor() == g["traj_draws"][s - 1], s
        if s in (50, 100, 200):
            p, _, v = o.get_state()
            assert np.abs(p - g[f"t{s}_pos"]).max() < 1e-9, s
            assert np.abs(v - g[f"t{s}_vel"]).max() < 1e-9, s
    assert o.rebuilds() == int(g["traj_rebuilt"].sum())


def _energy(o, mask, pos):
    o.set_state(pos, None, None)
    _, _, e = o.eval(mask)
    if mask == T.TERM_STACK:
        return e["e2_stack"] + e["e3_stack"]
    if mask == T.TERM_HB:
        return e["e2_hb"] + e["e3_hb"]
    if mask == T.TERM_BACKBONE:
        return e["e_bond"] + e["e_angle"] + e["e_wca"]
    if mask == T.TERM_SR:
        return e["e_wca"] + e["e_coul_sr"]
    return e["e_coul_lr"]


@pytest.mark.parametrize("name", ["stack", "hb", "poly", "SR", "LR"])
def test_forces_are_minus_energy_gradient(handle, name):
    """Central differences (h = 1e-5 A) of the term energy on 12 coordinates."""
    d = handle.d
    mask = TERMS[name]
    o = O.Oracle(handle)
    pos = d["pos"].reshape(-1, 3).copy()
    raw = O.glibc_stream(4242, 4000)
    o.set_state(pos, d["pos_old"], d["vel"])
    o.set_stream(raw)
    _, f, _ = o.eval(mask)
    mag = np.linalg.norm(f, axis=1)
    rng = np.random.default_rng(7)
    cand = np.nonzero(mag > 1e-3 * mag.max())[0]
    picks = rng.choice(cand, size=min(12, len(cand)), replace=False)
    h = 1e-5
    for i in picks:
        c = int(rng.integers(3))
        vals = []
        for sgn in (+1, -1):
            p = pos.copy()
            p[i, c] += sgn * h
            o.set_stream(raw)   # same draws: the HB resolution takes the same branch
            vals.append(_energy(o, mask, p))
        fd = -(vals[0] - vals[1]) / (2 * h)
        assert fd == pytest.approx(f[i, c], rel=2e-5, abs=2e-5 * mag.max()), (name, i, c)


@pytest.mark.parametrize("name", ["stack", "hb", "poly", "SR", "LR"])
def test_translation_invariance(orc, name):
    orc.set_stream(O.glibc_stream(99, 4000))
    _, f, _ = orc.eval(TERMS[name])
    scale = np.abs(f).sum()
    assert np.abs(f.sum(axis=0)).max() < 1e-10 * scale


@pytest.mark.parametrize("name", ["stack", "hb", "poly"])
def test_rotation_invariance(orc, handle, name):
    orc.set_stream(O.glibc_stream(99, 4000))
    _, f, _ = orc.eval(TERMS[name])
    x = handle.d["pos"].reshape(-1, 3)
    torque = np.cross(x - x.mean(axis=0), f).sum(axis=0)
    scale = np.abs(np.cross(x - x.mean(axis=0), f)).sum()
    assert np.abs(torque).max() < 1e-10 * scale


def test_cell_lists_equal_brute_force(handle):
    d = handle.d
    a = O.Oracle(handle)
    a.set_state(d["pos"], d["pos_old"], d["vel"])
    b = O.Oracle(handle)
    b.set_state(d["pos"], d["pos_old"], d["vel"], brute=True)
    assert np.array_equal(a.pairs(0), b.pairs(0))
    assert np.array_equal(a.pairs(1), b.pairs(1))
    assert np.array_equal(a.hb_status(), b.hb_status())
