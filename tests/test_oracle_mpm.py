"""Pins on the MPM oracle (no GPU): B-spline identities, P2G conservation,
APIC reproduction of affine node velocity fields, and the Drucker-Prager
return mapping's consistency condition."""
import numpy as np
import pytest

import orc
import sphx

UNDEF = 0xFFFFFFFF


@pytest.mark.parametrize("kernel", [0, 1])
def test_spline_partition_of_unity_and_moments(kernel):
    """sum_k N(x - k) = 1, sum_k N(x - k)(k - x) = 0 and sum_k N(x - k)(k - x)^2
    = 1/3 (cubic) or 1/4 (quadratic) for every x; the reference's cubic uses
    truncated constants (1.33333, 0.666666), so it holds to 1e-5."""
    tol = 1e-5 if kernel == 0 else 1e-12
    m2 = 1.0 / 3.0 if kernel == 0 else 0.25
    for x in np.linspace(0.0, 1.0, 37):
        ks = np.arange(-3, 5)
        w = np.array([orc.mpm_weight(kernel, x - k) for k in ks])
        assert abs(w.sum() - 1.0) < tol
        assert abs((w * (ks - x)).sum()) < tol
        assert abs((w * (ks - x) ** 2).sum() - m2) < 4 * tol


@pytest.mark.parametrize("kernel", [0, 1])
def test_spline_derivative_is_derivative(kernel):
    eps = 1e-6
    for x in np.linspace(-2.2, 2.2, 89):
        if any(abs(abs(x) - b) < 1e-3 for b in (0.0, 0.5, 1.0, 1.5, 2.0)):
            continue
        fd = (orc.mpm_weight(kernel, x + eps) - orc.mpm_weight(kernel, x - eps)) / (2 * eps)
        assert abs(fd - orc.mpm_dweight(kernel, x)) < 2e-5, x


def sand_state(kernel, n=1500, seed=3):
    p, disp, calc, info = sphx.scene_sand(n=n, seed=seed)
    st = orc.State(disp, calc, p, np.float64)
    st.grid_build()
    rng = np.random.default_rng(seed)
    st["vel"][:] = rng.normal(0, 0.3, (st.n, 3))
    mp = orc.MPM(st, p, kernel, 0, calc)
    mp["B"][:] = rng.normal(0, 0.01, (st.n, 9))
    return p, st, mp


def node_positions(p):
    res = (p.mpmRes.x, p.mpmRes.y, p.mpmRes.z)
    y, z, x = np.meshgrid(np.arange(res[1]), np.arange(res[2]), np.arange(res[0]), indexing="ij")
    xs = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float64)
    return np.array([p.mpmXmin.x, p.mpmXmin.y, p.mpmXmin.z]) + xs * p.mpmcellsize


@pytest.mark.parametrize("kernel", [0, 1])
def test_p2g_conserves_mass_and_momentum(kernel):
    p, st, mp = sand_state(kernel)
    fl = (st["type"] != 1) & (st["mgcell"] != UNDEF)
    mass = st["mass"][fl]
    tol = 1e-5 if kernel == 0 else 1e-12
    mp.p2g(apic=0)
    assert abs(mp["node_m"].sum() - mass.sum()) < tol * mass.sum()
    mom_n = (mp["node_v"] * mp["node_m"][:, None]).sum(0)
    mom_p = (st["vel"][fl] * mass[:, None]).sum(0)
    np.testing.assert_allclose(mom_n, mom_p, atol=tol * np.abs(mom_p).max() * 10)
    # APIC: the affine term has zero first moment, so momentum is unchanged
    mp.p2g(apic=1)
    mom_a = (mp["node_v"] * mp["node_m"][:, None]).sum(0)
    np.testing.assert_allclose(mom_a, mom_p, atol=tol * np.abs(mom_p).max() * 10)


@pytest.mark.parametrize("kernel", [0, 1])
def test_g2p_reproduces_affine_node_velocity(kernel):
    """Node velocities v_I = A x_I + b: G2P gives v_p = A x_p + b and
    B_p = A D with D = dx^2/3 (cubic) or dx^2/4 (quadratic)."""
    p, st, mp = sand_state(kernel)
    rng = np.random.default_rng(8)
    A = rng.normal(0, 1, (3, 3))
    b = rng.normal(0, 1, 3)
    xn = node_positions(p)
    mp["node_v"][:] = xn @ A.T + b
    mp["node_m"][:] = 1.0
    x0 = st["pos"].copy()
    p.dt = 0.0
    mp.g2p(apic=1)
    tol = 2e-5 if kernel == 0 else 1e-10
    np.testing.assert_allclose(st["vel"], x0 @ A.T + b, atol=tol * 10)
    D = p.mpmcellsize ** 2 / (3.0 if kernel == 0 else 4.0)
    # B = A D + (A x_p + b) (sum_I w (x_I - x_p))^T: the second term is the
    # first-moment defect of the truncated cubic constants (~1e-5 dx)
    vmax = np.abs(x0 @ A.T + b).max()
    atol_B = 3e-5 * p.mpmcellsize * vmax if kernel == 0 else 1e-10 * D
    np.testing.assert_allclose(mp["B"].reshape(-1, 3, 3), np.broadcast_to(A * D, (st.n, 3, 3)),
                               atol=atol_B)
    np.testing.assert_array_equal(st["pos"], x0)   # dt = 0


def dp_params():
    p, _, _, _ = sphx.scene_sand(n=8)
    return p


def test_drucker_prager_constants():
    p = dp_params()
    s = np.sin(np.radians(30.0))
    assert p.a_phi == pytest.approx(2 * s / np.sqrt(3) / (3 - s), rel=1e-6)
    assert p.a_psi == 0.0 and p.k_c == 0.0


def test_drucker_prager_hydrostatic_compression_does_not_yield():
    p = dp_params()
    sig = -1000.0 * np.eye(3)
    eps = np.diag([-1.0, -2.0, -0.5])
    assert np.all(orc.drucker_prager(p, 1e5, 2e5, eps, sig) == 0)


def test_drucker_prager_consistency_condition():
    """On the yield surface (f = 0) the plastic correction keeps the trial
    stress rate tangent to it: f(sigma + dt (C eps - plastic)) = O(dt^2),
    while the elastic rate alone leaves it at O(dt)."""
    p = dp_params()
    G, K = 134615.0, 291667.0
    a = p.a_phi
    # pure shear plus pressure, just outside the surface f = sqrt(J2) + a I1
    s_dev = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    j2 = np.sqrt(0.5 * (s_dev ** 2).sum())
    pr = j2 / (3 * a) * (1 - 1e-9)
    sig = s_dev - pr * np.eye(3)

    def f(sg):
        d = sg - np.trace(sg) / 3 * np.eye(3)
        return np.sqrt(0.5 * (d ** 2).sum()) + a * np.trace(sg)

    assert 0 < f(sig) < 1e-8
    eps = np.array([[0.0, 0.3, 0.0], [0.3, 0.0, 0.0], [0.0, 0.0, 0.0]])   # loading shear
    C_eps = 2 * G * eps + (K - 2 * G / 3) * np.trace(eps) * np.eye(3)
    pl = orc.drucker_prager(p, G, K, eps, sig)
    assert np.abs(pl).max() > 0
    dt = 1e-7
    f_el = f(sig + dt * C_eps)
    f_pl = f(sig + dt * (C_eps - pl))
    assert f_el > 1e-3 * dt * G
    assert abs(f_pl) < 1e-6 * abs(f_el)


# -- rate-form stress (MpmSolver.cu:477-560) ---------------------------------
# For a node velocity field v_I = L x_I + b the B-spline gradient is exactly
# grad v = L (quadratic spline, physical units), and the Jaumann rate form
# gives sigma' = sigma + dt (2G D + (K - 2G/3) tr(D) I + W sigma - sigma W)
# with D = (L + L^T)/2, W = (L - L^T)/2.  Under a pure rotation the stress
# must rotate rigidly: sigma' = R sigma R^T + O(dt^2), R = exp(W dt).

def stress_state(L, b=(0.1, -0.2, 0.05), sigma0=None, typ=2, m0=0.5):
    p, disp, calc, info = sphx.scene_sand(n=600, seed=11)
    p.dt = 5e-5
    disp["type"] = typ
    st = orc.State(disp, calc, p, np.float64)
    st.grid_build()
    mp = orc.MPM(st, p, 1, 1, calc)
    xn = node_positions(p)
    mp["node_v"][:] = xn @ np.asarray(L).T + np.asarray(b)
    mp["node_m"][:] = m0
    s0 = np.array([[-3000.0, 200.0, -50.0], [200.0, -2000.0, 120.0], [-50.0, 120.0, -1000.0]])
    if sigma0 is not None:
        s0 = sigma0
    mp["stress"][:] = s0.reshape(9)
    G, K = 134615.0, 291667.0
    mp["G"][:] = G
    mp["K"][:] = K
    ok = st["mgcell"] != UNDEF
    assert ok.all()
    mp.stress()
    return p, st, mp, s0, G, K


def jaumann_step(s0, L, G, K, dt):
    D = 0.5 * (L + L.T)
    W = 0.5 * (L - L.T)
    return s0 + dt * (2 * G * D + (K - 2 * G / 3) * np.trace(D) * np.eye(3) + W @ s0 - s0 @ W)


def test_stress_rate_pure_rotation_is_objective():
    w = np.array([0.0, 1.5, -2.0])
    L = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])   # v = w x x
    p, st, mp, s0, G, K = stress_state(L)
    dt = p.dt
    s1 = mp["stress"].reshape(-1, 3, 3)
    np.testing.assert_allclose(s1, np.broadcast_to(jaumann_step(s0, L, G, K, dt), s1.shape),
                               rtol=0, atol=1e-9 * np.abs(s0).max())
    # rigid rotation of the stress, to second order in dt
    th = np.linalg.norm(w) * dt
    k = w / np.linalg.norm(w)
    Kx = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    R = np.eye(3) + np.sin(th) * Kx + (1 - np.cos(th)) * Kx @ Kx
    rot = R @ s0 @ R.T
    change = np.abs(s1[0] - s0).max()
    assert change > 0
    assert np.abs(s1 - rot).max() < 1e-3 * change
    # particle density from the node masses: sum_I m_I w_Ip / dx^3 = m0 / dx^3
    rc = p.mpmcellsize * p.simscale
    np.testing.assert_allclose(st["dens"], 0.5 / rc ** 3, rtol=1e-12)


def test_stress_rate_uniaxial_strain_moduli():
    """v = (e x, 0, 0): d sigma_xx = (K + 4G/3) e dt, d sigma_yy = d sigma_zz
    = (K - 2G/3) e dt, no shear change (constrained modulus and lambda)."""
    e = 0.8
    L = np.diag([e, 0.0, 0.0])
    s0 = -1000.0 * np.eye(3)
    p, st, mp, _, G, K = stress_state(L, sigma0=s0)
    ds = (mp["stress"].reshape(-1, 3, 3) - s0) / (e * p.dt)
    want = np.diag([K + 4 * G / 3, K - 2 * G / 3, K - 2 * G / 3])
    np.testing.assert_allclose(ds, np.broadcast_to(want, ds.shape), rtol=1e-9, atol=1e-6 * K)


def test_stress_rate_general_gradient():
    rng = np.random.default_rng(4)
    L = rng.normal(0, 1.0, (3, 3))
    p, st, mp, s0, G, K = stress_state(L)
    s1 = mp["stress"].reshape(-1, 3, 3)
    np.testing.assert_allclose(s1, np.broadcast_to(jaumann_step(s0, L, G, K, p.dt), s1.shape),
                               rtol=0, atol=1e-9 * np.abs(s0).max())


# -- node update and grid collision (MpmSolver.cu:595-704) -------------------
def test_node_update_gravity_and_soft_boundaries():
    """Zero stress: every massive node gains g dt; then the soft bounds act
    as a sticky floor (a node below the floor moving down stops), free-slip
    side walls (the inward normal component of a node outside a wall is
    removed) and an open top."""
    p, disp, calc, info = sphx.scene_sand(n=600, seed=12)
    p.dt = 1e-3
    st = orc.State(disp, calc, p, np.float64)
    st.grid_build()
    mp = orc.MPM(st, p, 1, 1, calc)
    mp["stress"][:] = 0.0
    rng = np.random.default_rng(5)
    N = len(mp["node_m"])
    v0 = rng.normal(0, 1.0, (N, 3))
    mp["node_v"][:] = v0
    mp["node_m"][:] = 1.0
    mp["node_m"][::7] = 0.0                # massless nodes are left alone
    mp.node_update()
    g = np.array(p.gravity.tolist(), np.float64)
    x = node_positions(p)
    lo = np.array(p.softminx.tolist(), np.float64)
    hi = np.array(p.softmaxx.tolist(), np.float64)
    v = v0 + g * p.dt
    floor = (lo[1] - x[:, 1] > 1e-6) & (v[:, 1] < 0)
    v[floor] = 0.0
    for a in (0, 2):
        v[(lo[a] - x[:, a] > 1e-6) & (v[:, a] < 0), a] = 0.0
        v[(x[:, a] - hi[a] > 1e-6) & (v[:, a] > 0), a] = 0.0
    v[::7] = v0[::7]
    assert floor.sum() > 100 and (x[:, 1] - hi[1] > 1e-6).sum() > 100
    np.testing.assert_allclose(mp["node_v"], v, rtol=1e-12, atol=1e-12)


# -- MpmGetMomentum (MpmSolver.cu:107-153) -----------------------------------
def node_cells_f32(p):
    """initMpm's node cells (FP32, truncation) and validity 1 <= g <= res-1."""
    res = (p.mpmRes.x, p.mpmRes.y, p.mpmRes.z)
    y, z, x = np.meshgrid(np.arange(res[1]), np.arange(res[2]), np.arange(res[0]), indexing="ij")
    ids = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32)
    x0 = np.array([p.mpmXmin.x, p.mpmXmin.y, p.mpmXmin.z], np.float32)
    pos = x0 + ids * np.float32(p.mpmcellsize)
    lo = np.array([p.gridmin.x, p.gridmin.y, p.gridmin.z], np.float32)
    fac = np.array([p.gridIdfac.x, p.gridIdfac.y, p.gridIdfac.z], np.float32)
    g = ((pos - lo) * fac).astype(np.int64)
    gr = np.array([p.gridres.x, p.gridres.y, p.gridres.z])
    ok = np.all((g >= 1) & (g <= gr - 1), axis=1)
    return pos, ok


def test_node_momentum_brute_force_and_uniform_field():
    """Every node against an O(N x nodes) trilinear-weighted mean over all
    particles (no cell search), and a uniform particle velocity reproduced
    exactly at every node that has particles within one spacing."""
    p, disp, calc, info = sphx.scene_sand(n=1500, seed=13)
    # nodes off the SPH cell lattice, so that every one of the 27 cells can
    # hold particles within a node spacing
    p.mpmXmin.x += 0.4 * p.mpmcellsize
    p.mpmXmin.y += 0.7 * p.mpmcellsize
    p.mpmXmin.z += 0.15 * p.mpmcellsize
    st = orc.State(disp, calc, p, np.float64)
    st.grid_build()
    rng = np.random.default_rng(6)
    st["vel"][:] = rng.normal(0, 0.5, (st.n, 3))
    mp = orc.MPM(st, p, 1, 1, calc)
    mp.momentum()
    xn, ok = node_cells_f32(p)
    cs = float(np.float32(p.mpmcellsize))
    x = st["pos"]
    v = st["vel"]
    want = np.zeros((len(xn), 3))
    hits = 0
    for i in np.where(ok)[0]:
        d = np.abs(xn[i].astype(np.float64) - x)
        m = np.all(d < cs, axis=1)
        if not m.any():
            continue
        w = np.prod(1.0 - d[m] / cs, axis=1)
        if w.sum() > 1e-6:
            want[i] = (w[:, None] * v[m]).sum(0) / w.sum()
            hits += 1
    assert hits > 500
    np.testing.assert_allclose(mp["node_v"], want, rtol=0, atol=1e-12)
    st["vel"][:] = [0.25, -1.5, 0.75]
    mp.momentum()
    nz = np.abs(want).sum(1) > 0
    np.testing.assert_allclose(mp["node_v"][nz], np.broadcast_to([0.25, -1.5, 0.75], (nz.sum(), 3)),
                               rtol=1e-13)
    assert np.all(mp["node_v"][~nz] == 0)
