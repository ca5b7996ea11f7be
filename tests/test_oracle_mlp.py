"""Pins on the oracle of the learned warm start (no GPU): finite differences
and an independent autograd for the loss gradient, known-answer weight
layouts for the forward pass, and closed forms / brute force for the
feature map."""
import numpy as np
import pytest

import orc
import sphx

NF, H, NO, NP = orc.MLP_NF, orc.MLP_H, orc.MLP_NO, orc.MLP_NP
W1, B1 = 0, NF * H
W2, B2 = B1 + H, B1 + H + H * H
W3, B3 = B2 + H, B2 + H + H * NO


def random_net(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    w = rng.normal(0.0, scale, NP)
    return w


def test_layout_constants():
    assert NP == 6020
    assert (W1, B1, W2, B2, W3, B3) == (0, 1536, 1600, 5696, 5760, 6016)


def test_forward_known_answers():
    """Weights that realise y_o = relu(x_o) + c_o, and single-entry weights
    that expose a transposed or shifted layout."""
    w = np.zeros(NP)
    for k in range(NO):
        w[W1 + k * H + k] = 1.0      # h1_k = relu(x_k)
        w[W2 + k * H + k] = 1.0      # h2_k = h1_k
        w[W3 + k * NO + k] = 1.0     # y_k = h2_k
    w[B3:B3 + NO] = [0.5, -1.0, 2.0, 0.0]
    rng = np.random.default_rng(3)
    X = rng.normal(size=(64, NF))
    Y, H1, H2 = orc.mlp_forward(w, X)
    np.testing.assert_allclose(Y, np.maximum(X[:, :NO], 0) + w[B3:B3 + NO], rtol=0, atol=1e-15)
    # one path: x_5 -> unit 7 (w=2) -> unit 11 (w=-3, bias 1) -> output 2 (w=0.25)
    w = np.zeros(NP)
    w[W1 + 5 * H + 7] = 2.0
    w[W2 + 7 * H + 11] = -3.0
    w[B2 + 11] = 1.0
    w[W3 + 11 * NO + 2] = 0.25
    X = np.zeros((3, NF))
    X[:, 5] = [0.1, -1.0, 0.2]
    Y, H1, H2 = orc.mlp_forward(w, X)
    h2 = np.maximum(1.0 - 3.0 * np.maximum(2.0 * X[:, 5], 0), 0)
    np.testing.assert_allclose(Y[:, 2], 0.25 * h2, atol=1e-15)
    assert np.all(Y[:, [0, 1, 3]] == 0)
    np.testing.assert_allclose(H1[:, 7], np.maximum(2.0 * X[:, 5], 0))


def test_loss_gradient_finite_differences():
    w = random_net(1)
    rng = np.random.default_rng(2)
    X = rng.normal(size=(40, NF))
    T = rng.normal(size=(40, NO))
    lw = np.array([1.0, 0.3, 0.2, 0.1])
    loss, g = orc.mlp_loss_grad(w, X, T, lw)
    Y, _, _ = orc.mlp_forward(w, X)
    assert loss == pytest.approx(np.mean(np.sum(lw * (Y - T) ** 2, axis=1)), rel=1e-12)
    idx = np.concatenate([rng.choice(NP, 200, replace=False), np.arange(B3, NP),
                          np.arange(B1, B1 + 8), np.arange(B2, B2 + 8)])
    eps = 1e-6
    for e in idx:
        wp, wm = w.copy(), w.copy()
        wp[e] += eps
        wm[e] -= eps
        fd = (orc.mlp_loss_grad(wp, X, T, lw)[0] - orc.mlp_loss_grad(wm, X, T, lw)[0]) / (2 * eps)
        assert abs(fd - g[e]) <= 1e-6 * max(1.0, abs(fd)), (e, fd, g[e])


def test_loss_gradient_against_torch_autograd():
    torch = pytest.importorskip("torch")
    w = random_net(4)
    rng = np.random.default_rng(5)
    X = rng.normal(size=(33, NF))
    T = rng.normal(size=(33, NO))
    lw = np.array([1.0, 0.1, 0.5, 2.0])
    loss, g = orc.mlp_loss_grad(w, X, T, lw)
    tw = torch.tensor(w, dtype=torch.float64, requires_grad=True)
    x = torch.tensor(X)
    a = torch.nn.functional.linear(x, tw[W1:B1].view(NF, H).t(), tw[B1:W2]).relu()
    b = torch.nn.functional.linear(a, tw[W2:B2].view(H, H).t(), tw[B2:W3]).relu()
    y = torch.nn.functional.linear(b, tw[W3:B3].view(H, NO).t(), tw[B3:])
    tl = ((y - torch.tensor(T)) ** 2 * torch.tensor(lw)).sum(dim=1).mean()
    tl.backward()
    assert loss == pytest.approx(tl.item(), rel=1e-12)
    np.testing.assert_allclose(g, tw.grad.numpy(), rtol=1e-10, atol=1e-13)


# --------------------------------------------------------------- features --
def lattice_state(n=14, s=0.01, h=0.02):
    p, disp, calc, info = sphx.scene_dambreak(n, n, n, s, (n * s, n * s, n * s), h, 0, 0.0, 1, 1)
    p.dt = 1e-3
    st = orc.State(disp, calc, p, np.float32)
    st.grid_build()
    st.ii_density()
    st.ii_dii()
    st.ii_aii()
    return p, st


def interior(st, p, margin):
    pos = st["pos"].astype(np.float64)
    lo = pos.min(axis=0) + margin
    hi = pos.max(axis=0) - margin
    return np.all((pos > lo) & (pos < hi), axis=1)


def test_features_uniform_and_linear_pressure():
    # 2h = 3.8 s keeps every lattice distance clear of the support edge
    p, st = lattice_state(h=0.019)
    h = p.smoothradius
    P = 5000.0
    inn = interior(st, p, 2 * h + 1e-4)
    assert inn.sum() > 50
    # uniform pressure: neighbour statistics equal p_i, zero gradient
    st["pressure"][:] = 3000.0
    X = orc.mlp_features(st, None, P)
    fl = inn
    np.testing.assert_allclose(X[fl, 0], 0.6, rtol=1e-6)
    for q in (5, 6, 7):
        np.testing.assert_allclose(X[fl, q], 0.6, rtol=1e-6)
    assert np.all(np.abs(X[fl, 8]) < 1e-6)
    assert np.all(X[:, 19:22][fl] == 0)
    assert np.all(X[:, 23] == 1.0)
    # hydrostatic p = rho0 g (H - y): the SPH gradient recovers rho0 g, so
    # the feature along gravity is 1 (lattice sums of the cubic spline at
    # h = 1.9 s are within 2 %) and the horizontal ones vanish
    rho0, g = 1000.0, 9.81
    y = st["pos"][:, 1].astype(np.float64)
    st["pressure"][:] = (rho0 * g * (0.5 - y)).astype(np.float32)
    # the gradient weights read m_j / rho_j: use the exact lattice volume
    st["dens"][:] = (st["mass"] / 1e-6).astype(np.float32)
    X = orc.mlp_features(st, None, P)
    np.testing.assert_allclose(X[fl, 19], 1.0, rtol=0.02)
    assert np.all(np.abs(X[fl, 21]) < 1e-3)
    # a linear field's kernel-weighted average is p_i on a symmetric lattice
    assert np.all(np.abs(X[fl, 8]) < 1e-4)
    # max / min of a linear field over the 2h ball: p_i -+ rho0 g 3s (the
    # farthest lattice rows below / above inside the support)
    s = 0.01
    np.testing.assert_allclose(X[fl, 6] - X[fl, 0], rho0 * g * 3 * s / P, rtol=1e-4)
    np.testing.assert_allclose(X[fl, 0] - X[fl, 7], rho0 * g * 3 * s / P, rtol=1e-4)


def test_features_pointwise_and_counts_against_brute_force():
    p, st = lattice_state(n=10)
    rng = np.random.default_rng(9)
    st["pressure"][:] = rng.uniform(0, 4000, st.n).astype(np.float32)
    alpha4 = np.zeros((st.n, 4), np.float32)
    alpha4[:, 1] = rng.uniform(0, 1, st.n)
    alpha4[:, 0] = 1 - alpha4[:, 1]
    alpha4[:, 3] = alpha4[:, 0] * 1000 + alpha4[:, 1] * 1.2
    P = 2000.0
    dpl = rng.uniform(-300, 300, st.n).astype(np.float32)
    X = orc.mlp_features(st, alpha4, P, dpl)
    np.testing.assert_allclose(X[:, 20], dpl.astype(np.float64) / P * 10, rtol=1e-12)
    rho0 = st["restdens"].astype(np.float64)
    np.testing.assert_allclose(X[:, 0], st["pressure"] / P, rtol=1e-6)
    np.testing.assert_allclose(X[:, 1], (st["rho_adv"] - rho0) / rho0 * 100, rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(X[:, 2], (st["dens"] - rho0) / rho0 * 100, rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(X[:, 3], -st["aii"] * P / rho0, rtol=1e-6)
    np.testing.assert_allclose(X[:, 11], alpha4[:, 1], rtol=1e-7)
    np.testing.assert_allclose(X[:, 12], alpha4[:, 3] / rho0, rtol=1e-6)
    np.testing.assert_allclose(X[:, 13:16], st["dii"] * P / p.smoothradius, rtol=1e-6)
    np.testing.assert_allclose(X[:, 16:19], st["veleval"] * p.dt / p.smoothradius, rtol=1e-6,
                               atol=1e-12)
    # neighbour count within 2h, brute force over all pairs
    pos = st["pos"].astype(np.float64)
    d = np.sqrt(((pos[:, None, :] - pos[None, :, :]) ** 2).sum(-1))
    nn = ((d > 0) & (d < 2 * p.smoothradius)).sum(1)
    np.testing.assert_array_equal(np.rint(X[:, 9] * 64).astype(int), nn)
    assert np.all(X[:, 10] == 0)
    # max / min of neighbour pressures, brute force
    pr = np.maximum(st["pressure"].astype(np.float64), 0)
    nb = (d > 0) & (d < 2 * p.smoothradius)
    pmax = np.array([max(pr[i], pr[nb[i]].max()) for i in range(st.n)])
    pmin = np.array([min(pr[i], pr[nb[i]].min()) for i in range(st.n)])
    np.testing.assert_allclose(X[:, 6], pmax / P, rtol=1e-7)
    np.testing.assert_allclose(X[:, 7], pmin / P, rtol=1e-7)
