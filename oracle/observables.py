"""Oracle (test infrastructure only): numpy restatement of the reference's
offline observables, the definitions the device accumulators follow.

Only tests and bench.py's CPU leg use this module, as the checker; the product
computes the same quantities on the device (csrc/observe.h).

  p_ion_counts   per phosphate bead, ions of each kind within `cut`, minimum image
                 with the box side (analysis/mg_correlation/mg_bound.F90:97-126:
                 v > L/2 -> v - L, v < -L/2 -> v + L, then |v|^2 <= cut^2)
  bound_totals   distinct ions of each kind within `cut` of any phosphate
                 (count_bound_total, called at analysis/salt_nt/dcd_traj_SaltCount_lB.py:52-62
                 with cut r_kT = 7.2684 A; the routine lives in the external cafysis
                 package, absent here: restated from its call site and output use)
  shape          gyration tensor T = sum_ij (x_i - x_j)(x_i - x_j)^T / (2 N^2), its
                 eigenvalues w, Rg = sqrt(tr T), D = 1.5 sum (w - wbar)^2 / (tr T)^2,
                 S = 27 prod (w - wbar) / (tr T)^3 (analysis/shape/dcd_traj_shape.F90:193-226)
  radius_of_gyration  fast.cpp:2844-2876 (unweighted, about the centroid)
  concentration  mean count / (4/3 pi cut^3 * 6.02214e-4) mol/L
                 (analysis/salt_nt/integrate_num_ions.py:51-55)

Pinned by tests/test_observables_oracle.py: closed forms of D and S for
collinear, planar-ring and isotropic point sets; tr T = Rg^2 against the
reference's radius_of_gyration formula; hand-counted contacts across the
periodic boundary; bound totals against a per-ion brute force.
"""
import numpy as np

R_KT = 7.2684
CLS_P, CLS_MG, CLS_CL, CLS_K = 0, 3, 4, 5


def min_image(v, box):
    """mg_bound.F90:104-112, per component."""
    v = np.array(v, dtype=np.float64, copy=True)
    L = np.asarray(box, dtype=np.float64)
    hi = v > 0.5 * L
    v[hi] -= np.broadcast_to(L, v.shape)[hi]
    lo = v < -0.5 * L
    v[lo] += np.broadcast_to(L, v.shape)[lo]
    return v


def p_ion_counts(pos, cls, box, cut=R_KT):
    """(n_P, 3) counts of Mg, Cl, K within cut of every P bead, in bead order,
    and the P bead indices."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    cls = np.asarray(cls)
    ps = np.nonzero(cls == CLS_P)[0]
    out = np.zeros((len(ps), 3), dtype=np.int64)
    for c, kind in enumerate((CLS_MG, CLS_CL, CLS_K)):
        ions = np.nonzero(cls == kind)[0]
        if len(ions) == 0:
            continue
        for a, p in enumerate(ps):
            v = min_image(pos[ions] - pos[p], box)
            out[a, c] = int(np.count_nonzero(np.einsum("ij,ij->i", v, v) <= cut * cut))
    return out, ps


def bound_totals(pos, cls, box, cut=R_KT):
    """Distinct Mg, Cl, K ions within cut of at least one P bead."""
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    cls = np.asarray(cls)
    ps = np.nonzero(cls == CLS_P)[0]
    out = []
    for kind in (CLS_MG, CLS_CL, CLS_K):
        n = 0
        for i in np.nonzero(cls == kind)[0]:
            v = min_image(pos[ps] - pos[i], box)
            n += bool(np.any(np.einsum("ij,ij->i", v, v) <= cut * cut))
        out.append(n)
    return np.array(out)


def gyration_tensor(x):
    """dcd_traj_shape.F90:200-212: the double sum over pairs, divided by 2 N^2."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    n = len(x)
    d = x[:, None, :] - x[None, :, :]
    return np.einsum("ija,ijb->ab", d, d) / (2.0 * n * n)


def shape(x):
    """(Rg, D, S) of dcd_traj_shape.F90:214-222."""
    w = np.linalg.eigvalsh(gyration_tensor(x))
    tr = w.sum()
    wa = tr / 3.0
    rg = np.sqrt(tr)
    D = 1.5 * np.sum((w - wa) ** 2) / tr ** 2
    S = 27.0 * np.prod(w - wa) / tr ** 3
    return rg, D, S


def radius_of_gyration(x):
    """fast.cpp:2844-2876."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    c = x.mean(axis=0)
    return float(np.sqrt(np.sum((x - c) ** 2) / len(x)))


def concentration(mean_count, cut=R_KT):
    """integrate_num_ions.py:51-55 (mol/L)."""
    return np.asarray(mean_count) / (4.0 / 3.0 * np.pi * cut ** 3 * 6.02214 / 1e4)
