"""Force-matching backward pass (tis_force_match) on the GPU.

Loss L = (1/M) sum |F(x; theta) - F_target|^2 over snapshots of the 5 mM
system (native beads + N(0, 0.5^2 A^2) per coordinate, ions as placed), with
targets from the same snapshots under parameters perturbed by +-10 %.
Checked against:
  * central finite differences of the FP64 device loss (h = 1e-4) for
    positions, HB strengths, tertiary / secondary stack energies and the ion LJ
    radii (1e-4 relative, the stated bar for parameter gradients), for the FP64
    and the mixed-precision (production) gradients;
  * an FP64 reference built from the CPU oracle: its forces, its loss, its
    J v = dF/dx . v by central differences of its forces along v = F - target,
    and the finite differences of its loss in every parameter set, against the
    FP64 and the mixed-precision device (reference-stream mode, so that the HB
    valence resolution takes the same branch);
  * the torch.autograd.Function wrapper (gradients through loss.backward(),
    torch.autograd.gradcheck on the ion radii).
"""
import numpy as np
import pytest

import force_match as FM
import oracle_ctypes as O
import tis_ctypes as T

pytestmark = pytest.mark.gpu


def _snapshots(d, n_snap, seed=0):
    return FM.snapshots(d, n_snap, sigma=0.5, seed=seed)


_perturbed = FM.perturbed_params
_base_params = FM.base_params
Problem = FM.Problem


@pytest.fixture(scope="module")
def sys1(cfg1):
    return T.SystemHandle(cfg1[0])


def _targets(lib, sys1, snaps, precision=T.PREC_FP64, **kw):
    d = sys1.d
    tp = Problem(lib, sys1, snaps, precision, **kw)
    n = len(snaps) * d["n_particles"]
    out = tp.run(snaps, _perturbed(d), np.zeros((n, 3)), grads=False)
    tp.close()
    return out["forces"]


def _fd_check(fd, g, scale, what, floor=0.0, tol=1e-4):
    """tol relative (1e-4 unless stated), with floor = the rounding error of a
    central difference of the loss (about 10 eps |L| / h)."""
    err = abs(fd - g)
    print(what, "fd", fd, "grad", g, "err", err)
    assert err <= tol * max(abs(fd), 1e-3 * scale) + floor, what


def _without_contacts(d, x, lim=2.2):
    """No bonded / 1-3 pair and no other pair of particles closer than lim (A)."""
    from scipy.spatial import cKDTree
    b = d["bond_ij"].reshape(-1, 2)
    a = d["angle_ijk"].reshape(-1, 3)
    pairs = np.concatenate([b, a[:, [0, 2]]])
    if np.linalg.norm(x[pairs[:, 0]] - x[pairs[:, 1]], axis=1).min() <= lim:
        return False
    bonded = {(min(i, k), max(i, k)) for i, k in pairs}
    close = cKDTree(x).query_pairs(lim, output_type="ndarray")
    return all((min(i, j), max(i, j)) in bonded for i, j in close)


@pytest.mark.parametrize("prec,seed", [(T.PREC_FP64, 0), (T.PREC_FP64, 227), (T.PREC_MIXED, 227)],
                         ids=["fp64-contacts", "fp64", "mixed"])
def test_gradients_match_finite_differences(lib, sys1, prec, seed):
    """Device gradients in either precision against central differences (h =
    1e-4) of the FP64 device loss, the function both precisions compute.  The
    mixed gradients (FP32 forces in the residual, FP64 tangents) are checked
    against the same differences: a difference of the mixed loss itself at h =
    1e-4 would measure the FP32 rounding of the forces it touches (~1e-7 of
    each force, against a change of ~1e-4 of the term), not the derivative.

    Snapshots: seed 0 pushes beads of snapshot 0 into near-contact (a non-bonded
    pair at ~1.7 A carries a WCA force of ~1e15 kcal/mol/A), which only FP64
    resolves: FP32 rounds that force by ~1e8, and such a residual swamps every
    other term of the mixed loss.  Seed 227 gives two snapshots with no pair of
    particles within 2.2 A except bonded / 1-3 pairs (force-matching data from a
    trajectory has no such contacts); both precisions are checked on it.

    Bars: parameters 1e-4 relative in both precisions.  Positions 1e-4 in FP64
    and 1e-3 in mixed precision: dL/dx = (2/M) J v carries the FP32 rounding of
    the forces in the residual v = F - target (~1e-7 of each force, a few 1e-6
    kcal/mol/A) through the stiff bond and contact terms of J (measured 3e-4
    relative on the median coordinate); the tangent arithmetic itself is FP64
    and matches the oracle's J along the same v to 1e-6
    (test_gradients_match_the_fp64_oracle)."""
    d = sys1.d
    snaps = _snapshots(d, 2, seed=seed)
    if seed == 227:
        assert all(_without_contacts(d, x) for x in snaps)
    target = _targets(lib, sys1, snaps)
    pb = Problem(lib, sys1, snaps, prec)
    base = _base_params(d)
    out = pb.run(snaps, base, target)
    print("loss", out["loss"])
    assert out["loss"] > 0
    n = d["n_particles"]
    h = 1e-4
    pf = pb if prec == T.PREC_FP64 else Problem(lib, sys1, snaps, T.PREC_FP64)

    def loss_at(xs=snaps, **over):
        p = dict(base)
        p.update(over)
        return pf.run(xs, p, target, grads=False)["loss"]

    # positions: five bead coordinates of snapshot 0 with gradients between the
    # 50th and 95th percentiles (the largest belong to beads the noise pushed
    # into contact, where the loss varies as r^-26 and a finite difference of
    # step 1e-4 does not resolve it) and three ion coordinates of snapshot 1
    gp = out["pos"]
    gb = np.abs(gp[:d["n_beads"]]).ravel()
    lo, hi = np.percentile(gb, 50), np.percentile(gb, 95)
    gscale = hi
    cand = np.nonzero((gb >= lo) & (gb <= hi))[0]
    gi = np.abs(gp[n + d["n_beads"]:2 * n]).ravel()
    ions = [3 * (n + d["n_beads"]) + int(k) for k in np.argsort(-gi)[:3]]
    picks = list(np.random.default_rng(11).choice(cand, size=5, replace=False)) + ions
    floor = 10 * np.finfo(float).eps * out["loss"] / h
    for flat in picks:
        i, c = divmod(int(flat), 3)
        r, li = divmod(i, n)
        vals = []
        for s in (+1, -1):
            xs = [x.copy() for x in snaps]
            xs[r][li, c] += s * h
            vals.append(loss_at(xs))
        _fd_check((vals[0] - vals[1]) / (2 * h), gp[i, c], gscale, f"x[{i},{c}]", floor,
                  tol=1e-4 if prec == T.PREC_FP64 else 1e-3)
    # parameters
    for key, count in (("hb_E", 4), ("st3_E", 3), ("st2_E", 4), ("ion_R", 3)):
        g = out[key]
        idx = np.argsort(-np.abs(g))[:count]
        # every candidate / stack picked has a nonzero gradient; an ion radius may
        # have none when no pair of that ion type is in WCA contact (the placement
        # keeps ions apart), and then the finite difference must vanish too
        assert np.abs(g[idx]).max() > 0, key
        if key != "ion_R":
            assert np.abs(g[idx]).min() > 0, key
        for k in idx:
            vals = []
            for s in (+1, -1):
                p = base[key].copy()
                p[k] += s * h
                vals.append(loss_at(**{key: p}))
            _fd_check((vals[0] - vals[1]) / (2 * h), g[k], np.abs(g).max(), f"{key}[{k}]", floor)
    pb.close()
    if pf is not pb:
        pf.close()


def test_gradients_match_the_fp64_oracle(lib, sys1):
    """Reference-stream mode: the oracle's forces, loss, J v and parameter
    finite differences against the device (FP64 and mixed precision)."""
    d = sys1.d
    # The reference adds the WCA of bonded and 1-3 pairs in the pair sum and
    # subtracts it in the bonded terms; a noise-pushed 1-3 pair near the shifted
    # singularity (r ~ 1.61 A) makes both ~1e29 and their FP64 difference rounding
    # noise, while the device leaves the pair out (the same force, exactly).  Most
    # sigma = 0.5 A snapshots have such a pair (closest 1-3 distance 1.3-2.3 A)
    # or a non-bonded contact whose r^-13 force makes the loss ~1e24, beyond what
    # a finite difference of the loss resolves; this one (seed 36) has its closest
    # 1-3 pair at 2.42 A and its closest other pair at 2.27 A.
    from scipy.spatial import cKDTree
    snaps = _snapshots(d, 1, seed=36)
    b = d["bond_ij"].reshape(-1, 2)
    a = d["angle_ijk"].reshape(-1, 3)
    pairs = np.concatenate([b, a[:, [0, 2]]])
    assert np.linalg.norm(snaps[0][pairs[:, 0]] - snaps[0][pairs[:, 1]], axis=1).min() > 2.2
    bonded = {(min(i, k), max(i, k)) for i, k in pairs}
    close = cKDTree(snaps[0]).query_pairs(2.2, output_type="ndarray")
    assert all((min(i, j), max(i, j)) in bonded for i, j in close)
    raw = O.glibc_stream(77, 20000)
    target = _targets(lib, sys1, snaps, rng=T.RNG_STREAM, stream=raw)
    n = d["n_particles"]
    M = 3 * n

    def oracle_forces(handle, x):
        o = O.Oracle(handle)
        o.set_state(x, x, np.zeros_like(x), klok=101)
        o.set_stream(raw)
        _, f, _ = o.eval(T.TERM_ALL)
        return f

    f_o = oracle_forces(sys1, snaps[0])
    v = f_o - target
    loss_o = float(np.sum(v * v)) / M
    def oracle_jv(w):
        """J w of the oracle's forces by the fourth-order central stencil
        (largest displacement 2e-4 A)."""
        eps = 1e-4 / np.abs(w).max()
        fo = {k: oracle_forces(sys1, snaps[0] + k * eps * w) for k in (-2, -1, 1, 2)}
        return (-fo[2] + 8 * fo[1] - 8 * fo[-1] + fo[-2]) / (12 * eps)

    for prec, tol_f, tol_g in ((T.PREC_FP64, 1e-10, 1e-6), (T.PREC_MIXED, 1e-5, 1e-6)):
        pb = Problem(lib, sys1, snaps, prec, rng=T.RNG_STREAM, stream=raw)
        out = pb.run(snaps, _base_params(d), target)
        rel_f = np.linalg.norm(out["forces"] - f_o) / np.linalg.norm(f_o)
        # the device's J v in FP64 forward-mode derivatives along its own
        # residual v (FP32 forces in mixed mode: stiff contacts amplify their
        # rounding in J v, so the reference uses the same v)
        g_ref = 2.0 / M * oracle_jv(out["forces"] - target)
        rel_g = np.linalg.norm(out["pos"] - g_ref) / np.linalg.norm(g_ref)
        print(prec, "loss", out["loss"], "oracle", loss_o, "rel F", rel_f, "rel grad", rel_g)
        assert rel_f < tol_f
        assert out["loss"] == pytest.approx(loss_o, rel=10 * tol_f)
        assert rel_g < tol_g
        # parameter derivatives of the oracle's FP64 loss for the largest device
        # gradients, in both precisions (1e-4 relative, the bar for parameters)
        for key, pkey in (("hb_E", "hb_E"), ("st3_E", "st3_E"), ("st2_E", "st2_E"), ("ion_R", "type_R")):
            g = out[key]
            for k in np.argsort(-np.abs(g))[:2]:
                vals = []
                for s in (+1, -1):
                    dd = dict(d)
                    arr = dd[pkey].copy()
                    arr[k + (6 if key == "ion_R" else 0)] += s * 1e-4
                    dd[pkey] = arr
                    f = oracle_forces(T.SystemHandle(dd), snaps[0])
                    vals.append(float(np.sum((f - target) ** 2)) / M)
                fd = (vals[0] - vals[1]) / 2e-4
                print(key, k, "oracle fd", fd, "device", g[k])
                floor = 10 * np.finfo(float).eps * loss_o / 1e-4
                assert abs(fd - g[k]) <= 1e-4 * abs(fd) + floor
        pb.close()


def test_autograd_function_returns_the_device_gradients(lib, sys1):
    torch = pytest.importorskip("torch")
    d = sys1.d
    snaps = _snapshots(d, 2, seed=3)
    target = _targets(lib, sys1, snaps)
    pb = Problem(lib, sys1, snaps, T.PREC_FP64)
    base = _base_params(d)
    direct = pb.run(snaps, base, target)
    pos = torch.tensor(np.concatenate(snaps), dtype=torch.float64, requires_grad=True)
    prm = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in base.items()}
    loss = FM.force_match_loss(pos, prm["hb_E"], prm["st3_E"], prm["st2_E"], prm["ion_R"], pb.eng, target, 101)
    loss.backward()
    assert loss.item() == pytest.approx(direct["loss"], rel=1e-12)
    # equal up to the order of the FP64 force atomics
    same = lambda a, b: np.allclose(a, b, rtol=1e-10, atol=1e-12 * np.abs(b).max())
    assert same(pos.grad.numpy(), direct["pos"])
    for k in base:
        assert same(prm[k].grad.numpy(), direct[k]), k
    pos_c = pos.detach()
    fixed = {k: prm[k].detach() for k in ("hb_E", "st3_E", "st2_E")}
    f = lambda r: FM.force_match_loss(pos_c, fixed["hb_E"], fixed["st3_E"], fixed["st2_E"], r, pb.eng, target, 101)
    assert torch.autograd.gradcheck(f, (prm["ion_R"].detach().clone().requires_grad_(True),), eps=1e-4,
                                    atol=1e-6, rtol=1e-4)
    pb.close()


def test_256_snapshots_linearity(lib, sys1):
    """The full force-matching set (256 snapshots, 936,704 particles in one
    context): the loss and every gradient equal the combination of four
    64-snapshot evaluations (L = mean of the four losses, parameter gradients
    their mean, position gradients theirs divided by 4).  Every replica reads the
    same reference-stream draws (Philox keys depend on the replica's slot, and a
    snapshot's HB resolution must not change with its position in the batch).
    Finite differences are checked on two snapshots above: with sigma = 0.5 A
    some of the 256 snapshots put non-bonded beads almost in contact, the r^-26
    WCA residual makes the set's loss ~1e60, and a difference of that loss
    cannot resolve any parameter gradient."""
    d = sys1.d
    snaps = _snapshots(d, 256, seed=0)
    n = d["n_particles"]
    stream = T.glibc_stream(lib, 5, 50000)
    target = _targets(lib, sys1, snaps, rng=T.RNG_STREAM, stream=stream)
    base = _base_params(d)
    full = Problem(lib, sys1, snaps, T.PREC_FP64, rng=T.RNG_STREAM, stream=stream)
    g = full.run(snaps, base, target)
    full.close()
    parts = []
    for b in range(4):
        sl = slice(64 * b, 64 * (b + 1))
        pb = Problem(lib, sys1, snaps[sl], T.PREC_FP64, rng=T.RNG_STREAM, stream=stream)
        parts.append(pb.run(snaps[sl], base, target[64 * n * b:64 * n * (b + 1)]))
        pb.close()
    print("loss", g["loss"])
    assert g["loss"] == pytest.approx(np.mean([p["loss"] for p in parts]), rel=1e-12)
    for k in ("hb_E", "st3_E", "st2_E", "ion_R"):
        want = np.mean([p[k] for p in parts], axis=0)
        assert np.linalg.norm(g[k] - want) <= 1e-10 * np.linalg.norm(want), k
    want = np.concatenate([p["pos"] for p in parts]) / 4.0
    assert np.linalg.norm(g["pos"] - want) <= 1e-10 * np.linalg.norm(want)
