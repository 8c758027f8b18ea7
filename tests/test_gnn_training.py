"""The training script's torch statement of the learned prolongation against the f64 numpy oracle: the same
weights from the same parameters, and torch autograd's parameter gradient of the two-level loss equal to the
oracle's hand-written backward pass (itself pinned by finite differences)."""
import numpy as np
import pytest
import torch

import oracle_lib as O
from oracle import gnn as G
from drape_amd.learned import train as T


@pytest.fixture(scope="module")
def pair():
    od = O.Drape(O.reference_scene(17, sphere=False, placed=True))
    a = od.problem_arrays()
    rest = od.rest_positions()
    x_hat = rest.copy()
    x_hat[2::3] -= 1e-3
    od.set_time_step(0.1)
    _, vals = od.assemble(rest, x_hat, 0.0, True)
    g = G.Graph(17, a["row_vertex"], a["row_ptr"], a["col_idx"], vals)
    s = T.System(17, a["row_vertex"], a["row_ptr"], a["col_idx"], vals, g.parent, g.w_bil, g.coarse_rows,
                 torch.device("cpu"))
    return g, s


def test_torch_forward_matches_oracle(pair):
    g, s = pair
    rng = np.random.default_rng(3)
    theta = G.init_params(seed=2) + 0.1 * rng.standard_normal(G.param_count())
    w_np = G.forward(theta, g)
    w_t = T.forward(torch.tensor(theta), s).numpy()
    assert np.allclose(w_t, w_np, rtol=1e-12, atol=1e-14)


def test_torch_gradient_matches_oracle_backward(pair):
    g, s = pair
    rng = np.random.default_rng(4)
    theta = G.init_params(seed=5) + 0.05 * rng.standard_normal(G.param_count())
    E = G.smooth_vectors(g, 3, seed=6)
    L_np, grad_np = G.backward(theta, g, E)
    Et = torch.tensor(E)
    AE = s.Acsr @ Et
    den = (Et * AE).sum(dim=0)
    th = torch.tensor(theta, requires_grad=True)
    L_t = T.loss(T.forward(th, s), s, Et, AE, den)
    L_t.backward()
    assert float(L_t) == pytest.approx(L_np, rel=1e-10)
    scale = np.abs(grad_np).max()
    assert np.abs(th.grad.numpy() - grad_np).max() <= 1e-7 * scale


def test_torch_pcg_loss_matches_oracle_and_its_gradient_finite_differences(pair):
    """The training loss (A-norm error after k two-level PCG iterations) in torch equals the numpy oracle's,
    is 1 at k = 0, and its autograd gradient matches central differences of the oracle's loss."""
    g, s = pair
    T.attach(s, 1, seed=0)
    rng = np.random.default_rng(9)
    xstar = rng.standard_normal(3 * g.R)
    b = g.A @ xstar
    bt, xt = torch.tensor(b), torch.tensor(xstar)
    theta = G.init_params(seed=10) + 0.05 * rng.standard_normal(G.param_count())
    w_np = G.forward(theta, g)
    for k in (0, 6):
        L_np = G.pcg_loss(w_np, g, b, xstar, k)
        th = torch.tensor(theta, requires_grad=True)
        L_t = T.pcg_loss(T.forward(th, s), s, bt, xt, k)
        assert float(L_t) == pytest.approx(L_np, rel=1e-9)
        if k == 0:
            assert L_np == pytest.approx(1.0, rel=1e-12)
    L_t.backward()
    grad = th.grad.numpy()
    for j in list(rng.choice(theta.size, 6, replace=False)) + [theta.size - 2, theta.size - 40]:
        h = 1e-6
        tp, tm = theta.copy(), theta.copy()
        tp[j] += h
        tm[j] -= h
        fd = (G.pcg_loss(G.forward(tp, g), g, b, xstar, 6) - G.pcg_loss(G.forward(tm, g), g, b, xstar, 6)) / (2 * h)
        assert fd == pytest.approx(grad[j], rel=2e-4, abs=1e-10), j
