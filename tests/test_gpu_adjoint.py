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


def test_adjoint_64_smoothed_target():
    """The adjoint case at 64^3 (kappa and b seeded as for the kappa
    operator): x_t a seeded N(0,1) field smoothed by 3 symmetric Gauss-Seidel
    sweeps (r = 0), L = 1/2 |x - x_t|^2, forward and adjoint solves in the
    mixed FP32-MG / FP64-outer PCG to 1e-10.  dL/dkappa and dL/db within
    1e-6 of the oracle's FP64 CPU gradient (FP64 CG to 1e-13, forward and
    adjoint); dL/dkappa against central finite differences on 64 random
    entries (h = 1e-6, FP64 device solves to 1e-13)."""
    import torch

    from hpcg_amd.adjoint import KappaOperator, kappa_solve
    n = 64
    kappa, b, x0 = setup_problem(n, seed=2027)
    P = po.Problem(n, n, n, kappa=kappa)
    xt = x0.copy()
    for _ in range(3):
        xt = P.symgs(np.zeros(P.n), xt, mode=2)
    k = torch.tensor(kappa, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda", requires_grad=True)
    xtt = torch.tensor(xt, device="cuda")
    x = kappa_solve(k, bt, (n, n, n), 1e-10, "mixed_fp32")
    loss = 0.5 * ((x - xtt) ** 2).sum()
    loss.backward()
    gk, gb = k.grad.cpu().numpy(), bt.grad.cpu().numpy()
    xo = P.cg(b, max_iter=2000, tol=1e-13, mode=2)["x"]
    lamo = P.cg(xo - xt, max_iter=2000, tol=1e-13, mode=2)["x"]
    go = P.kappa_grad(lamo, xo)
    assert np.linalg.norm(gk - go) <= 1e-6 * np.linalg.norm(go), np.linalg.norm(gk - go) / np.linalg.norm(go)
    assert np.linalg.norm(gb - lamo) <= 1e-6 * np.linalg.norm(lamo)
    bd = torch.tensor(b, device="cuda")
    h = 1e-6
    scale = np.max(np.abs(go))
    for j in np.random.default_rng(77).choice(n ** 3, 64, replace=False):
        xs = []
        for sgn in (1.0, -1.0):
            kp = kappa.copy()
            kp[j] += sgn * h
            o = KappaOperator(kp, (n, n, n), variant="fp64")
            xs.append(o.solve(bd, tol=1e-13))
            o.close()
        fd = 0.5 * torch.dot(xs[0] - xs[1], xs[0] + xs[1] - 2 * xtt).item() / (2 * h)
        # FD error measured at 128^3: 1.3e-5 of max |g|; a wrong term of the
        # formula (sign, factor 1/2, the boundary term) is O(1)
        assert abs(fd - gk[j]) <= 1e-4 * scale + 1e-4 * abs(gk[j]), (j, fd, gk[j])


def test_gradients_keep_input_shape_and_device():
    """kappa as a 3-D CPU tensor gets its gradient as a 3-D CPU tensor, equal
    to the 1-D device case's; a numpy kappa takes no gradient and b still
    gets one."""
    import torch

    from hpcg_amd.adjoint import kappa_solve
    n = 16
    kappa, b, xt = setup_problem(n, seed=9)
    xtt = torch.tensor(xt, device="cuda")
    k1 = torch.tensor(kappa, device="cuda", requires_grad=True)
    b1 = torch.tensor(b, device="cuda", requires_grad=True)
    (0.5 * ((kappa_solve(k1, b1, (n, n, n), 1e-12, "fp64") - xtt) ** 2).sum()).backward()
    k3 = torch.tensor(kappa.reshape(n, n, n), requires_grad=True)  # CPU, 3-D
    b3 = torch.tensor(b, device="cuda", requires_grad=True)
    (0.5 * ((kappa_solve(k3, b3, (n, n, n), 1e-12, "fp64") - xtt) ** 2).sum()).backward()
    assert k3.grad.shape == (n, n, n) and k3.grad.device.type == "cpu"
    assert torch.equal(k3.grad.reshape(-1), k1.grad.cpu())
    assert torch.equal(b3.grad, b1.grad)
    bn = torch.tensor(b, device="cuda", requires_grad=True)
    (0.5 * ((kappa_solve(kappa, bn, (n, n, n), 1e-12, "fp64") - xtt) ** 2).sum()).backward()
    assert torch.equal(bn.grad, b1.grad)
