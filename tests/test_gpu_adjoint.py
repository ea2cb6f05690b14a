"""The variable-coefficient solve and its adjoint through torch.autograd
(hpcg_amd.adjoint over the C ABI), against the oracle's fp64 CPU gradient
and central finite differences."""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def setup_problem(n, seed=3):
    g = np.random.default_rng(seed)
    kappa = np.exp(0.5 * g.standard_normal(n ** 3))
    b = g.uniform(-1.0, 1.0, n ** 3)
    xt = g.standard_normal(n ** 3)
    return kappa, b, xt


def oracle_gradient(n, kappa, b, xt):
    P = po.Problem(n, n, n, kappa=kappa)
    x = P.cg(b, max_iter=1000, tol=1e-14, mode=2)["x"]
    lam = P.cg(x - xt, max_iter=1000, tol=1e-14, mode=2)["x"]
    return x, lam, P.kappa_grad(lam, x)


@pytest.mark.parametrize("variant,tol", [("fp64", 1e-13), ("mixed_fp32", 1e-10)])
def test_autograd_matches_fp64_cpu_gradient(variant, tol):
    import torch

    from hpcg_amd.adjoint import KappaSolve, kappa_solve
    n = 16
    kappa, b, xt = setup_problem(n)
    k = torch.tensor(kappa, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda", requires_grad=True)
    x = kappa_solve(k, bt, (n, n, n), tol, variant)
    loss = 0.5 * ((x - torch.tensor(xt, device="cuda")) ** 2).sum()
    loss.backward()
    xo, lamo, go = oracle_gradient(n, kappa, b, xt)
    assert np.linalg.norm(x.detach().cpu().numpy() - xo) <= 1e-6 * np.linalg.norm(xo)
    gk = k.grad.cpu().numpy()
    gb = bt.grad.cpu().numpy()
    assert np.linalg.norm(gk - go) <= 1e-6 * np.linalg.norm(go), np.linalg.norm(gk - go) / np.linalg.norm(go)
    assert np.linalg.norm(gb - lamo) <= 1e-6 * np.linalg.norm(lamo)
    t = KappaSolve.last_timing
    assert t["forward_iters"] > 0 and t["adjoint_iters"] > 0


def test_gradient_finite_differences():
    """Central differences of L = 1/2 |x - xt|^2 (h = 1e-6, fp64 solves to 1e-14)
    on random kappa entries; L(+h) - L(-h) = 1/2 (x+ - x-).(x+ + x- - 2 xt)."""
    import torch

    from hpcg_amd.adjoint import KappaOperator
    n = 16
    kappa, b, xt = setup_problem(n, seed=5)
    bt = torch.tensor(b, device="cuda")
    xtt = torch.tensor(xt, device="cuda")
    op = KappaOperator(kappa, (n, n, n), variant="fp64")
    x = op.solve(bt, tol=1e-14)
    lam = op.solve(x - xtt, tol=1e-14)
    g = op.grad_kappa(lam, x).cpu().numpy()
    op.close()
    rng = np.random.default_rng(6)
    h = 1e-6
    scale = np.max(np.abs(g))
    for j in rng.choice(n ** 3, 8, replace=False):
        xs = []
        for sgn in (1.0, -1.0):
            kp = kappa.copy()
            kp[j] += sgn * h
            o = KappaOperator(kp, (n, n, n), variant="fp64")
            xs.append(o.solve(bt, tol=1e-14))
            o.close()
        dl = 0.5 * torch.dot(xs[0] - xs[1], xs[0] + xs[1] - 2 * xtt).item()
        fd = dl / (2 * h)
        assert abs(fd - g[j]) <= 1e-5 * scale + 1e-4 * abs(g[j]), (j, fd, g[j])
