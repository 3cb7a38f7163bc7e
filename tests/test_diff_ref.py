"""The torch FP64 restatement used as the reference for gradients
(oracle/torch_pic.py), pinned on the CPU: its forward equals the C oracle's
field-energy history and particle state, and its autograd gradient equals
central finite differences of itself."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import pyoracle as po                        # noqa: E402
from common import rel_l2, two_stream        # noqa: E402
from torch_pic import TorchPIC               # noqa: E402


def case(nx=32, n=1024, Bz=0.0, seed=21):
    L = (float(nx), float(nx))
    m = 2 * n / (L[0] * L[1])
    parts = two_stream(n, L, 0.1, 0.02, seed)
    sp = np.concatenate([np.full(n, 0), np.full(n, 1)]).astype(np.int32)
    cat = [np.concatenate([p[k] for p in parts]) for k in range(1, 5)]
    return dict(nx=nx, L=L, dt=0.1, q=[-1.0, -1.0], m=[m, m], Bz=Bz, parts=parts, sp=sp,
                state=cat)


@pytest.mark.parametrize("Bz", [0.0, 0.3])
def test_torch_restatement_matches_oracle(Bz):
    c = case(Bz=Bz)
    o = po.Oracle(c["nx"], c["nx"], c["L"], c["dt"], c["q"], c["m"], B=(0.0, 0.0, Bz))
    for s, (ids, x, y, ux, uy) in enumerate(c["parts"]):
        o.add(s, ids, x, y, ux, uy)
    o.pre_step()
    W = []
    for it in range(10):
        o.stage("field_E")
        W.append(o.field_energy())
        o.stage("plasma_E")
        o.stage("plasma_r")
        o.stage("field_rho")
        o.iter = it + 1
    t = TorchPIC(c["nx"], c["nx"], c["L"], c["dt"], c["q"], c["m"], Bz=Bz)
    x, y, ux, uy = (torch.tensor(a) for a in c["state"])
    _, Wt = t.run(x, y, ux, uy, c["sp"], 10)
    assert rel_l2([float(w) for w in Wt], W) < 1e-12


def test_torch_gradient_matches_finite_differences():
    c = case(nx=16, n=128, Bz=0.2, seed=22)
    t = TorchPIC(c["nx"], c["nx"], c["L"], c["dt"], c["q"], c["m"], Bz=c["Bz"])
    st = [torch.tensor(a, requires_grad=True) for a in c["state"]]
    loss, _ = t.run(*st, c["sp"], 6)
    loss.backward()
    rng = np.random.default_rng(3)
    for k in range(4):
        for i in rng.choice(len(c["sp"]), 4, replace=False):
            h = 1e-6
            vals = []
            for sgn in (1, -1):
                s2 = [torch.tensor(a) for a in c["state"]]
                s2[k][i] += sgn * h
                vals.append(float(t.run(*s2, c["sp"], 6)[0]))
            fd = (vals[0] - vals[1]) / (2 * h)
            g = float(st[k].grad[i])
            assert abs(fd - g) <= 1e-5 * max(1.0, abs(g)), (k, i, fd, g)
