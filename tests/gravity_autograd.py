"""Test-harness wrapper of the adjoint C-ABI as a torch.autograd.Function.

TreeGravity.apply(pos, mass, cfg) returns the tree acceleration (particle-cell
and particle-particle lists, precision form) of every particle in the input
order. The forward pass keeps only its inputs: the tree, the interaction
lists and the moments are freed, and backward rebuilds them from the saved
positions (checkpoint and recompute), then calls changaGravityAdjoint with
the incoming gradient as w."""
import numpy as np
import torch

from changa_amd import engine


def _evaluation(pos, mass, cfg):
    w = dict(cfg)
    w["pos"] = np.ascontiguousarray(pos.detach().cpu().numpy(), dtype=np.float64)
    w["mass"] = np.ascontiguousarray(mass.detach().cpu().numpy(), dtype=np.float64)
    return engine.Evaluation(w, precise=True)


class TreeGravity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, mass, cfg):
        ev = _evaluation(pos, mass, cfg)
        ev.zero()
        ev.run_pc()
        ev.run_pp()
        torch.cuda.synchronize()
        perm = torch.as_tensor(ev.tree["perm"], dtype=torch.long)
        acc = torch.zeros_like(pos)
        acc[perm] = ev.d_vars[:, :3].double().to(pos.device)
        ctx.save_for_backward(pos, mass)
        ctx.cfg = cfg
        del ev
        return acc

    @staticmethod
    def backward(ctx, gacc):
        pos, mass = ctx.saved_tensors
        ev = _evaluation(pos, mass, ctx.cfg)  # recompute from the checkpoint
        perm = torch.as_tensor(ev.tree["perm"], dtype=torch.long)
        g = ev.adjoint(gacc.detach()[perm].float())
        torch.cuda.synchronize()
        gpos = torch.zeros_like(pos)
        gmass = torch.zeros_like(mass)
        gpos[perm] = g[:, :3].to(pos.device)
        gmass[perm] = g[:, 3].to(mass.device)
        del ev
        return gpos, gmass, None
