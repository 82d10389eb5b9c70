"""Force-matching loss over the C ABI as a torch.autograd.Function.

A set of snapshots of one system is held by an Engine as its replicas; the
loss is tis_force_match's
    L = (1/M) sum |F(x; hb_E, st3_E, st2_E, R_ion) - F_target|^2
and its backward returns the device's gradients (forward-mode derivatives of
the force kernels, see include/tis_gpu.h).  Positions enter as a (N_total, 3)
tensor (replicas concatenated); each call sets every replica's state (which
rebuilds its lists) at the step counter given, so the HB valence resolution
draws the same numbers on every call.
"""
import numpy as np
import torch

import tis_ctypes as T


def snapshots(d, n_snap, sigma=0.5, seed=0):
    """n_snap configurations of system d: its beads displaced by seeded
    N(0, sigma^2) noise per coordinate, its ions as placed."""
    rng = np.random.default_rng(seed)
    base = d["pos"].reshape(-1, 3)
    out = []
    for _ in range(n_snap):
        x = base.copy()
        x[:d["n_beads"]] += rng.normal(0.0, sigma, size=(d["n_beads"], 3))
        out.append(x)
    return out


def perturbed_params(d, frac=0.1, seed=1):
    """Force-field parameters of d (HB strengths, tertiary / secondary stack
    energies, Mg / Cl / K radii), each scaled by 1 +- frac with a seeded sign."""
    rng = np.random.default_rng(seed)
    sgn = lambda n: np.where(rng.random(n) < 0.5, -1.0, 1.0)
    return dict(hb_E=d["hb_E"] * (1 + frac * sgn(d["n_hb"])), st3_E=d["st3_E"] * (1 + frac * sgn(d["n_st3"])),
                st2_E=d["st2_E"] * (1 + frac * sgn(d["n_st2"])), ion_R=d["type_R"][6:9] * (1 + frac * sgn(3)))


def base_params(d):
    return dict(hb_E=d["hb_E"].copy(), st3_E=d["st3_E"].copy(), st2_E=d["st2_E"].copy(),
                ion_R=d["type_R"][6:9].copy())


class Problem:
    """An Engine whose replicas are snapshots of one system; loss and gradients
    of tis_force_match at (snapshots, parameters), every snapshot at a fixed
    step counter so that the HB valence resolution draws the same numbers on
    every call."""

    def __init__(self, lib, handle, snaps, precision, rng=T.RNG_PHILOX, stream=None, klok=101):
        self.h = handle
        self.snaps = snaps
        self.stream = stream
        self.klok = klok
        self.eng = T.Engine(lib, [handle] * len(snaps), precision=precision, rng=rng, seed=7, use_graph=0)

    def set(self, xs, params):
        for r, x in enumerate(xs):
            self.eng.set_state(x, x, np.zeros_like(x), klok=self.klok, rep=r)
            if self.stream is not None:
                self.eng.set_stream(self.stream, rep=r)
        self.eng.set_force_field(**params)

    def run(self, xs, params, target, grads=True):
        self.set(xs, params)
        return self.eng.force_match(target, want_grads=grads)

    def close(self):
        self.eng.close()


class ForceMatchLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, hb_E, st3_E, st2_E, ion_R, engine, target, klok):
        x = pos.detach().cpu().numpy().astype(np.float64)
        off = 0
        for r, h in enumerate(engine.handles):
            n = h.d["n_particles"]
            xr = x[off:off + n]
            engine.set_state(xr, xr, np.zeros_like(xr), klok=klok, rep=r)
            off += n
        engine.set_force_field(hb_E.detach().cpu().numpy(), st3_E.detach().cpu().numpy(),
                               st2_E.detach().cpu().numpy(), ion_R.detach().cpu().numpy())
        out = engine.force_match(target, want_grads=True)
        ctx.grads = [torch.from_numpy(out[k].copy()) for k in ("pos", "hb_E", "st3_E", "st2_E", "ion_R")]
        return pos.new_tensor(out["loss"])

    @staticmethod
    def backward(ctx, g):
        gp, ghb, gs3, gs2, gion = [t.to(g.dtype) * g for t in ctx.grads]
        return gp, ghb, gs3, gs2, gion, None, None, None


def force_match_loss(pos, hb_E, st3_E, st2_E, ion_R, engine, target, klok=1):
    return ForceMatchLoss.apply(pos, hb_E, st3_E, st2_E, ion_R, engine, target, klok)
