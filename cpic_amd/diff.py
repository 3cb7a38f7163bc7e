"""Differentiable PIC behind torch.autograd.Function.

`FieldEnergyLoss.apply(r0, u0, run, nsteps, checkpoint_every)` returns the
time-integrated field energy  L = sum_{it < nsteps} W(it)  of the FP64 PIC
loop started from particle positions r0 and velocities u0 (both (n, 2)
float64 tensors); its backward returns dL/dr0 and dL/du0 from the library's
hand-written adjoint kernels (cpic_adj_* in include/cpic_hip.h), with
checkpoints every `checkpoint_every` iterations and recompute.  Parameters
that the initial state is built from (drift v0, thermal spread sigma, ...)
get their gradients from torch's chain rule on top.

The tensors cross the C ABI as host arrays, like every other entry point;
the simulation runs on the GPU inside the library.
"""
import ctypes as C

import numpy as np

import cpic_amd as ca

ca.lib()          # the library (and its HIP runtime) loads before torch
import torch      # noqa: E402


class AdjointRun:
    """One cpic_adj_t: an FP64 run of n particles with fixed species."""

    def __init__(self, nx, ny, L, dt, q, m, species, e0=1.0, B=(0.0, 0.0, 0.0), nchunks=1,
                 device=0):
        self.desc = ca.make_desc(nx, ny, L, dt, q, m, e0=e0, B=B, nchunks=nchunks, precision=64,
                                 device=device)
        sp = np.ascontiguousarray(species, dtype=np.int32)
        self.n = len(sp)
        h = C.c_void_p()
        ca.check(ca.lib().cpic_adj_create(C.byref(self.desc), self.n,
                                          sp.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(h)))
        self._h = h
        self.generation = 0

    def close(self):
        if getattr(self, "_h", None):
            ca.lib().cpic_adj_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def forward(self, x, y, ux, uy, nsteps, checkpoint_every=20):
        x, y, ux, uy = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, ux, uy))
        loss = C.c_double()
        energy = np.empty(int(nsteps))
        ca.check(ca.lib().cpic_adj_forward(self._h, ca._dp(x), ca._dp(y), ca._dp(ux), ca._dp(uy),
                                           int(nsteps), int(checkpoint_every), C.byref(loss),
                                           ca._dp(energy)))
        self.generation += 1
        return loss.value, energy

    def backward(self, dloss=1.0):
        out = [np.empty(self.n) for _ in range(4)]
        ca.check(ca.lib().cpic_adj_backward(self._h, float(dloss), *(ca._dp(a) for a in out)))
        return out

    def state(self):
        out = [np.empty(self.n) for _ in range(4)]
        ca.check(ca.lib().cpic_adj_state(self._h, *(ca._dp(a) for a in out)))
        return out

    def set_field_history(self, on):
        """Whether later forwards keep every iteration's rho and E (when they
        fit) for the backward, instead of its recomputing them."""
        ca.check(ca.lib().cpic_adj_set_field_history(self._h, 1 if on else 0))

    def device_ms(self):
        """Device time of the last forward and backward in ms (-1 if not run)."""
        f, b = C.c_double(), C.c_double()
        ca.check(ca.lib().cpic_adj_device_ms(self._h, C.byref(f), C.byref(b)))
        return f.value, b.value


class FieldEnergyLoss(torch.autograd.Function):
    """loss = sum of field energies over nsteps iterations from (r0, u0)."""

    @staticmethod
    def forward(ctx, r0, u0, run, nsteps, checkpoint_every=20):
        r = r0.detach().to("cpu", torch.float64).numpy()
        u = u0.detach().to("cpu", torch.float64).numpy()
        loss, energy = run.forward(r[:, 0], r[:, 1], u[:, 0], u[:, 1], nsteps, checkpoint_every)
        ctx.run = run
        ctx.generation = run.generation
        ctx.device = r0.device
        ctx.energy = energy
        return torch.tensor(loss, dtype=torch.float64, device=r0.device)

    @staticmethod
    def backward(ctx, g):
        run = ctx.run
        if run.generation != ctx.generation:
            raise RuntimeError("AdjointRun was run forward again before this backward")
        gx, gy, gux, guy = run.backward(float(g))
        gr = torch.from_numpy(np.stack([gx, gy], axis=1)).to(ctx.device)
        gu = torch.from_numpy(np.stack([gux, guy], axis=1)).to(ctx.device)
        return gr, gu, None, None, None


def field_energy_loss(r0, u0, run, nsteps, checkpoint_every=20):
    return FieldEnergyLoss.apply(r0, u0, run, nsteps, checkpoint_every)
