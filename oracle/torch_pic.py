"""FP64 torch restatement of the PIC loop, differentiated by torch autograd.

Test infrastructure only (it checks the library's adjoint kernels):
it follows the same discrete model as the C oracle -- CIC weights of
relative_position_grid (src/interpolate.c:39-66), the MFT Green's function
(src/solver.c:269-280) applied with torch.fft, the central-difference E
(src/field.c:381-406), the Boris push for B = (0,0,Bz) (src/mover.c:23-83),
the iteration-0 half step (src/mover.c:204-214) and the periodic wrap -- so
that torch's reverse mode gives reference gradients of
L = sum_it W(it), W = -(e0/2) sum rho*phi.  Runs on the CPU.
"""
import math

import torch


class TorchPIC:
    def __init__(self, nx, ny, L, dt, q, m, e0=1.0, Bz=0.0):
        self.nx, self.ny = nx, ny
        self.L = (float(L[0]), float(L[1]))
        self.dt, self.e0, self.Bz = float(dt), float(e0), float(Bz)
        self.q = [float(v) for v in q]
        self.m = [float(v) for v in m]
        self.dx, self.dy = self.L[0] / nx, self.L[1] / ny
        hx = nx // 2 + 1
        cx = torch.tensor([math.cos(2.0 * math.pi / nx * k) for k in range(hx)], dtype=torch.float64)
        cy = torch.tensor([math.cos(2.0 * math.pi / ny * k) for k in range(ny)], dtype=torch.float64)
        den = 2.0 * (cx[None, :] + cy[:, None]) - 4.0
        den[0, 0] = 1.0
        G = 1.0 / den
        G[0, 0] = 0.0
        self.G = G

    def solve(self, rho):
        return torch.fft.irfft2(torch.fft.rfft2(rho) * self.G, s=(self.ny, self.nx))

    def efield(self, phi):
        ex = (torch.roll(phi, 1, dims=1) - torch.roll(phi, -1, dims=1)) / (2 * self.dx)
        ey = (torch.roll(phi, 1, dims=0) - torch.roll(phi, -1, dims=0)) / (2 * self.dy)
        return ex, ey

    def cic(self, x, y):
        idx, idy = 1.0 / self.dx, 1.0 / self.dy
        sx = torch.floor(x.detach() * idx)
        sy = torch.floor(y.detach() * idy)
        rx = (x - sx * self.dx) * idx
        ry = (y - sy * self.dx) * idy          # dx in the Y term, as the reference
        ix = sx.long() % self.nx
        iy = sy.long() % self.ny
        i1x = (ix + 1) % self.nx
        i1y = (iy + 1) % self.ny
        nodes = (iy * self.nx + ix, i1y * self.nx + ix, iy * self.nx + i1x, i1y * self.nx + i1x)
        dxw, dyw = 1.0 - rx, 1.0 - ry
        w = (dxw * dyw, dxw * ry, rx * dyw, rx * ry)
        return nodes, w

    def deposit(self, x, y, c):
        nodes, w = self.cic(x, y)
        rho = torch.zeros(self.nx * self.ny, dtype=torch.float64)
        for k in range(4):
            rho = rho.index_add(0, nodes[k], w[k] * c)
        return rho.view(self.ny, self.nx)

    def gather(self, ex, ey, x, y):
        nodes, w = self.cic(x, y)
        fx, fy = ex.reshape(-1), ey.reshape(-1)
        gx = sum(w[k] * fx[nodes[k]] for k in range(4))
        gy = sum(w[k] * fy[nodes[k]] for k in range(4))
        return gx, gy

    def run(self, x, y, ux, uy, species, nsteps):
        """Returns (loss, [W(it)]) from initial state tensors of shape (n,)."""
        sp = torch.as_tensor(species, dtype=torch.long)
        q = torch.tensor(self.q, dtype=torch.float64)[sp]
        m = torch.tensor(self.m, dtype=torch.float64)[sp]
        c = -q / self.e0
        Ws = []
        for it in range(nsteps):
            rho = self.deposit(x, y, c)
            phi = self.solve(rho)
            Ws.append(-0.5 * self.e0 * torch.sum(rho * phi))
            ex, ey = self.efield(phi)
            gx, gy = self.gather(ex, ey, x, y)
            dtp = -self.dt / 2 if it == 0 else self.dt
            h = 0.5 * dtp * q / m
            tz = self.Bz * h
            sz = 2.0 * tz / (tz * tz + 1.0)
            vmx, vmy = ux + h * gx, uy + h * gy
            vpx, vpy = vmx + vmy * tz, vmy - vmx * tz
            vplx, vply = vmx + vpy * sz, vmy - vpx * sz
            ux, uy = vplx + h * gx, vply + h * gy
            if it > 0:
                x = torch.remainder(x + dtp * ux, self.L[0])
                y = torch.remainder(y + dtp * uy, self.L[1])
        return sum(Ws), Ws
