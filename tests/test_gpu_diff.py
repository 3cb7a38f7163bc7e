"""Differentiable PIC on the GPU (cpic_adj_* through cpic_amd.diff).

- the adjoint module's forward is the production FP64 loop: its field-energy
  history and final state match the oracle (1e-12 relative L2, 20 steps);
- its gradients match torch autograd through the FP64 torch restatement
  (oracle/torch_pic.py) on a 64x64 / 10,240-particle two-stream over 200
  steps, and do not depend on the checkpoint interval;
- on 256^2 with 4,194,304 particles over 200 steps, dL/dv0 and dL/dsigma
  from the adjoint match central finite differences (h = 1e-6)."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import pyoracle as po                        # noqa: E402
from common import rel_l2, two_stream        # noqa: E402


def diff():
    from cpic_amd import diff as D
    return D


def c1_state(nx=64, n=5120, v0=0.1, sigma=0.01, seed=31, Bz=0.0):
    L = (float(nx), float(nx))
    m = 2 * n / (L[0] * L[1])
    parts = two_stream(n, L, v0, sigma, seed)
    sp = np.concatenate([np.full(n, 0), np.full(n, 1)]).astype(np.int32)
    st = [np.concatenate([p[k] for p in parts]) for k in range(1, 5)]
    return dict(nx=nx, L=L, dt=0.1, q=[-1.0, -1.0], m=[m, m], Bz=Bz, parts=parts, sp=sp, st=st)


def test_adjoint_forward_matches_oracle():
    D = diff()
    c = c1_state(Bz=0.2)
    run = D.AdjointRun(c["nx"], c["nx"], c["L"], c["dt"], c["q"], c["m"], c["sp"],
                       B=(0.0, 0.0, c["Bz"]))
    loss, W = run.forward(*c["st"], 20, 7)
    o = po.Oracle(c["nx"], c["nx"], c["L"], c["dt"], c["q"], c["m"], B=(0.0, 0.0, c["Bz"]))
    for s, (ids, x, y, ux, uy) in enumerate(c["parts"]):
        o.add(s, ids, x, y, ux, uy)
    o.pre_step()
    Wo = []
    for it in range(20):
        o.stage("field_E")
        Wo.append(o.field_energy())
        o.stage("plasma_E")
        o.stage("plasma_r")
        o.stage("field_rho")
        o.iter = it + 1
    assert rel_l2(W, Wo) < 1e-12
    assert abs(loss - sum(Wo)) < 1e-12 * abs(sum(Wo))
    x, y, ux, uy = run.state()
    n = len(c["parts"][0][0])
    for s in range(2):
        a = o.particles(s)
        sl = slice(s * n, (s + 1) * n)
        for k, v in (("x", x), ("y", y), ("ux", ux), ("uy", uy)):
            assert rel_l2(v[sl], a[k]) < 1e-12, (s, k)
    run.close()


def test_adjoint_gradient_matches_torch_reference_64():
    """64x64, 10,240 particles, 200 steps, checkpoints every 20."""
    D = diff()
    from torch_pic import TorchPIC
    c = c1_state()
    run = D.AdjointRun(c["nx"], c["nx"], c["L"], c["dt"], c["q"], c["m"], c["sp"])
    r0 = torch.tensor(np.stack(c["st"][:2], 1), requires_grad=True)
    u0 = torch.tensor(np.stack(c["st"][2:], 1), requires_grad=True)
    loss = D.field_energy_loss(r0, u0, run, 200, 20)
    loss.backward()
    t0 = time.time()
    ref = TorchPIC(c["nx"], c["nx"], c["L"], c["dt"], c["q"], c["m"])
    st = [torch.tensor(a, requires_grad=True) for a in c["st"]]
    lref, _ = ref.run(*st, c["sp"], 200)
    lref.backward()
    print(f"torch reference: {time.time() - t0:.1f} s", flush=True)
    assert abs(float(loss) - float(lref)) < 1e-10 * abs(float(lref))
    errs = {}
    for k, (g, gref) in enumerate(((r0.grad[:, 0], st[0].grad), (r0.grad[:, 1], st[1].grad),
                                   (u0.grad[:, 0], st[2].grad), (u0.grad[:, 1], st[3].grad))):
        errs[k] = rel_l2(g.numpy(), gref.numpy())
    print("gradient rel L2 vs torch reference (x, y, ux, uy):", errs, flush=True)
    assert max(errs.values()) < 1e-7, errs
    # the checkpoint interval changes only the recompute schedule
    run.forward(*c["st"], 200, 1)
    f_ms, b_ms = run.device_ms()
    assert f_ms > 0 and b_ms == -1.0          # this forward's backward has not run
    g1 = run.backward(1.0)
    f2, b_ms = run.device_ms()
    assert f2 == f_ms and b_ms > 0
    run.forward(*c["st"], 200, 200)
    g200 = run.backward(1.0)
    for a, b in zip(g1, g200):
        assert rel_l2(a, b) < 1e-9
    # a partial last segment (200 = 6 * 30 + 20) and an interval beyond the
    # run (one segment of 200 iterations, sized by the run, not by K)
    grefs = [g.grad.numpy() for g in st]
    for K in (30, 500):
        run.forward(*c["st"], 200, K)
        gK = run.backward(1.0)
        for a, b, gr in zip(g1, gK, grefs):
            assert rel_l2(b, a) < 1e-9, K
            assert rel_l2(b, gr) < 1e-7, K
    # without the forward's field history the reverse sweep recomputes each
    # segment's rho and E with its positions: the same gradients
    run.set_field_history(False)
    for K in (20, 30):
        lossK, _ = run.forward(*c["st"], 200, K)
        gK = run.backward(1.0)
        assert abs(lossK - float(lref)) < 1e-10 * abs(float(lref))
        for a, b, gr in zip(g1, gK, grefs):
            assert rel_l2(b, a) < 1e-9, K
            assert rel_l2(b, gr) < 1e-7, K
    run.set_field_history(True)
    run.close()


def test_adjoint_lds_boxes_across_the_periodic_boundary():
    """64x64 at 64 particles per cell (262,144 particles): each block's 1,024
    particles span a quarter row, so every deposit uses an LDS box, and the
    chunks that end one grid row and begin the next, or hold particles that
    cross x = 0 (v0 = 0.5), use a box that wraps.  Forward state and field
    energies vs the oracle at 1e-12; gradients vs torch autograd at 1e-7
    over 10 steps."""
    D = diff()
    from torch_pic import TorchPIC
    c = c1_state(nx=64, n=131072, v0=0.5, sigma=0.05, seed=37)
    nsteps = 10
    run = D.AdjointRun(c["nx"], c["nx"], c["L"], c["dt"], c["q"], c["m"], c["sp"])
    loss, W = run.forward(*c["st"], nsteps, 5)
    o = po.Oracle(c["nx"], c["nx"], c["L"], c["dt"], c["q"], c["m"])
    for s, (ids, x, y, ux, uy) in enumerate(c["parts"]):
        o.add(s, ids, x, y, ux, uy)
    o.pre_step()
    Wo = []
    for it in range(nsteps):
        o.stage("field_E")
        Wo.append(o.field_energy())
        o.stage("plasma_E")
        o.stage("plasma_r")
        o.stage("field_rho")
        o.iter = it + 1
    assert rel_l2(W, Wo) < 1e-12
    x, y, ux, uy = run.state()
    n = len(c["parts"][0][0])
    crossed = 0
    for s in range(2):
        a = o.particles(s)
        sl = slice(s * n, (s + 1) * n)
        for k, v in (("x", x), ("y", y), ("ux", ux), ("uy", uy)):
            assert rel_l2(v[sl], a[k]) < 1e-12, (s, k)
        crossed += int(np.sum(np.abs(x[sl] - c["st"][0][sl]) > c["L"][0] / 2))
    assert crossed > 1000, crossed
    r0 = torch.tensor(np.stack(c["st"][:2], 1), requires_grad=True)
    u0 = torch.tensor(np.stack(c["st"][2:], 1), requires_grad=True)
    lg = D.field_energy_loss(r0, u0, run, nsteps, 5)
    lg.backward()
    ref = TorchPIC(c["nx"], c["nx"], c["L"], c["dt"], c["q"], c["m"])
    st = [torch.tensor(a, requires_grad=True) for a in c["st"]]
    lref, _ = ref.run(*st, c["sp"], nsteps)
    lref.backward()
    assert abs(float(lg) - float(lref)) < 1e-10 * abs(float(lref))
    errs = [rel_l2(g.numpy(), gref.numpy())
            for g, gref in ((r0.grad[:, 0], st[0].grad), (r0.grad[:, 1], st[1].grad),
                            (u0.grad[:, 0], st[2].grad), (u0.grad[:, 1], st[3].grad))]
    print("gradient rel L2 vs torch reference (x, y, ux, uy):", errs, flush=True)
    assert max(errs) < 1e-7, errs
    run.close()


def test_c5_gradients_v0_sigma_vs_finite_differences():
    """256^2, 4,194,304 particles (64 ppc), two-stream, B = 0, 200 steps,
    checkpoint every 20: d loss / d v0 and d loss / d sigma through the
    adjoint and torch's chain rule vs central differences with h = 1e-6."""
    D = diff()
    nx, ppc = 256, 64
    n = nx * nx * ppc // 2
    L = (float(nx), float(nx))
    m = float(ppc)
    rng = np.random.default_rng(55)
    pos = rng.uniform(0.0, nx, (2 * n, 2))
    z = rng.normal(0.0, 1.0, (2 * n, 2))
    sign = np.concatenate([np.ones(n), -np.ones(n)])
    sp = np.concatenate([np.zeros(n), np.ones(n)]).astype(np.int32)
    run = D.AdjointRun(nx, nx, L, 0.1, [-1.0, -1.0], [m, m], sp)
    r0 = torch.tensor(pos)
    zt = torch.tensor(z)
    st = torch.tensor(sign)

    def velocities(v0, sigma):
        return torch.stack([st * v0 + sigma * zt[:, 0], sigma * zt[:, 1]], 1)

    v0 = torch.tensor(0.1, dtype=torch.float64, requires_grad=True)
    sigma = torch.tensor(0.01, dtype=torch.float64, requires_grad=True)
    t0 = time.time()
    loss = D.field_energy_loss(r0, velocities(v0, sigma), run, 200, 20)
    loss.backward()
    print(f"forward+backward: {time.time() - t0:.1f} s, loss {float(loss):.12e}", flush=True)
    h = 1e-6

    def L_at(a, b):
        with torch.no_grad():
            return float(D.field_energy_loss(r0, velocities(torch.tensor(a, dtype=torch.float64),
                                                            torch.tensor(b, dtype=torch.float64)),
                                             run, 200, 20))

    fd_v0 = (L_at(0.1 + h, 0.01) - L_at(0.1 - h, 0.01)) / (2 * h)
    fd_s = (L_at(0.1, 0.01 + h) - L_at(0.1, 0.01 - h)) / (2 * h)
    print(f"dL/dv0 adjoint {float(v0.grad):.10e} fd {fd_v0:.10e}; "
          f"dL/dsigma adjoint {float(sigma.grad):.10e} fd {fd_s:.10e}", flush=True)
    assert abs(float(v0.grad) - fd_v0) <= 1e-4 * abs(fd_v0)
    assert abs(float(sigma.grad) - fd_s) <= 1e-4 * abs(fd_s)
    run.close()
