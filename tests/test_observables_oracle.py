"""The observables oracle (oracle/observables.py) against closed forms and
hand-counted cases, before it is used to check the device accumulators.

  - shape parameters of point sets whose gyration eigenvalues are known:
    collinear (w, 0, 0): D = 1, S = 2; planar ring (w, w, 0): D = 1/4, S = -1/4;
    cube vertices (w, w, w): D = 0, S = 0;
  - tr T of the pair formula = Rg^2 of the reference's radius_of_gyration;
  - contacts across the periodic boundary counted by hand;
  - bound totals count an ion near two phosphates once;
  - 1 mol/L is 6.02214e-4 ions per A^3.
"""
import itertools

import numpy as np
import pytest

import observables as OB


def test_collinear_points_are_rods():
    rng = np.random.default_rng(0)
    t = rng.normal(size=40)
    x = np.outer(t, [0.3, -0.5, 0.8]) + [10.0, 20.0, 30.0]
    rg, D, S = OB.shape(x)
    assert D == pytest.approx(1.0, abs=1e-10)
    assert S == pytest.approx(2.0, abs=1e-9)


def test_planar_ring_is_oblate():
    a = 2 * np.pi * np.arange(12) / 12
    x = np.stack([5 * np.cos(a), 5 * np.sin(a), np.zeros_like(a)], axis=1)
    rg, D, S = OB.shape(x)
    assert rg == pytest.approx(5.0, rel=1e-12)
    assert D == pytest.approx(0.25, abs=1e-12)
    assert S == pytest.approx(-0.25, abs=1e-12)


def test_cube_vertices_are_isotropic():
    x = np.array(list(itertools.product([-1.0, 1.0], repeat=3))) * 3.0
    rg, D, S = OB.shape(x)
    assert rg == pytest.approx(3.0 * np.sqrt(3.0), rel=1e-12)
    assert abs(D) < 1e-14 and abs(S) < 1e-14


def test_trace_is_the_reference_rg_squared():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(300, 3)) * [20.0, 5.0, 9.0] + 175.0
    rg, _, _ = OB.shape(x)
    assert rg == pytest.approx(OB.radius_of_gyration(x), rel=1e-12)


def _layout():
    # box 100 A; P beads at (1, 50, 50) and (50, 50, 50)
    pos = [(1, 50, 50), (50, 50, 50),          # P, P
           (99, 50, 50),                         # Mg: 2 A from P0 through the boundary
           (9, 50, 50),                          # Mg: 8 A from P0 (outside 7.2684)
           (50, 50, 56),                         # Mg: 6 A from P1
           (93.0, 50, 50),                       # Cl: 8 A from P0 through the boundary
           (1, 50, 57.2),                        # K: 7.2 A from P0
           (50, 57.5, 50)]                       # K: 7.5 A from P1 (outside)
    cls = [0, 0, 3, 3, 3, 4, 5, 5]
    return np.array(pos, dtype=float), np.array(cls), np.array([100.0, 100.0, 100.0])


def test_contacts_counted_by_hand():
    pos, cls, box = _layout()
    counts, ps = OB.p_ion_counts(pos, cls, box)
    assert list(ps) == [0, 1]
    assert counts.tolist() == [[1, 0, 1], [1, 0, 0]]
    counts, _ = OB.p_ion_counts(pos, cls, box, cut=8.0)
    assert counts.tolist() == [[2, 1, 1], [1, 0, 1]]


def test_bound_totals_count_each_ion_once():
    pos, cls, box = _layout()
    pos = np.vstack([pos, [(25.5, 50, 50), (25.5, 50, 53)]])   # Mg and Cl 24.5 A from both Ps: unbound
    cls = np.concatenate([cls, [3, 4]])
    assert OB.bound_totals(pos, cls, box).tolist() == [2, 0, 1]
    pos[-2] = (4.0, 50, 50)                                     # now 3 A from P0 and 3 A from Mg(99)
    assert OB.bound_totals(pos, cls, box).tolist() == [3, 0, 1]
    # a K within the cut of both P beads is one bound K
    pos2 = np.array([(0, 0, 0), (4, 0, 0), (2, 1, 0)], dtype=float)
    assert OB.bound_totals(pos2, np.array([0, 0, 5]), box).tolist() == [0, 0, 1]


def test_one_molar_is_avogadro_per_litre():
    v = 4.0 / 3.0 * np.pi * OB.R_KT ** 3
    assert OB.concentration(1.0 * 6.02214e-4 * v) == pytest.approx(1.0, rel=1e-12)
