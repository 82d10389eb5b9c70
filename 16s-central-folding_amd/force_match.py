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
