"""torch.autograd binding of the AVBD adjoint (a test and training binding).

``replay(ctx, options, state, velocities, angular_velocities, inv_mass,
inv_moment, friction)`` replays the steps recorded in ``ctx`` (after
``ctx.adjoint_begin`` and ``ctx.step``) from ``state`` with the given
velocities, inverse masses / moments and friction, and returns the final
positions (n x 3) and orientations (n x 4).  Its backward is the hand-written
reverse mode of the library (``dio_adjoint_backward``): the gradient of the
frozen-contact replay, in which each recorded step keeps its contact set and
only the AVBD solve is differentiated.

Tensors cross the C boundary as host float32 arrays; outputs and gradients
come back on the inputs' device.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import torch

from . import records as R


def _host(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float32).contiguous().numpy()


class _Replay(torch.autograd.Function):
    @staticmethod
    def forward(fctx, ctx, options, state, velocities, angular, inv_mass, inv_moment, friction):
        st = np.array(state, R.BODY_STATE, copy=True)
        st["vel"][:, :3] = _host(velocities)
        st["ang"][:, :3] = _host(angular)
        st["inv_mass"] = _host(inv_mass)
        st["inv_moment"][:, :3] = _host(inv_moment)
        opts = dataclasses.replace(options, friction=float(friction))
        pos, rot = ctx.adjoint_replay(opts, st)
        # friction may be a Python number (no gradient) or a one-element
        # tensor of any shape (its gradient takes that shape)
        fctx.saved = (ctx, opts, st, tuple(friction.shape) if torch.is_tensor(friction) else None)
        dev = velocities.device
        return torch.from_numpy(pos).to(dev), torch.from_numpy(rot).to(dev)

    @staticmethod
    def backward(fctx, g_pos, g_rot):
        ctx, opts, st, friction_shape = fctx.saved
        g = ctx.adjoint_backward(opts, d_positions=None if g_pos is None else _host(g_pos),
                                 d_orientations=None if g_rot is None else _host(g_rot), state0=st)
        dev = g_pos.device if g_pos is not None else g_rot.device
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        g_friction = None
        if friction_shape is not None:
            g_friction = torch.tensor(g["friction"], dtype=torch.float32, device=dev).reshape(friction_shape)
        return (None, None, None, t(g["velocities"]), t(g["angular_velocities"]), t(g["inv_mass"]),
                t(g["inv_moment"]), g_friction)


def replay(ctx, options, state, velocities, angular_velocities, inv_mass, inv_moment, friction):
    """Differentiable replay of ctx's recording (see the module docstring)."""
    return _Replay.apply(ctx, options, state, velocities, angular_velocities, inv_mass, inv_moment, friction)
