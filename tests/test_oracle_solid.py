"""Pins on the oracle's SPH elastic solid and surface tension (no GPU):
closed forms on a perfect lattice and pair antisymmetry, chosen so a
transposed product, a swapped index or a wrong sign fails one of them."""
import numpy as np

import orc
import sphx

ELASTIC = 2


def lattice(n=7, s=0.01, h=0.021, jitter=0.0, seed=1):
    """n^3 elastic particles of spacing s (no boundary), WCSPH kernels
    (support h, cells 2h); mass 1e-3, dens = 1/rho = 1e-3 as after the
    density pass."""
    p, disp, calc, info = sphx.scene_dambreak(n, n, n, s, (n * s, n * s, n * s), h, 0, jitter,
                                              seed, 0)
    disp["type"] = ELASTIC
    calc["mass"] = 1e-3
    calc["dens"] = 1e-3
    p.dt = 2e-2
    p.solidK = 1e5
    st = orc.State(disp, calc, p, np.float64)
    st.grid_build()
    assert np.all(st["mgcell"] != 0xFFFFFFFF)
    return p, st


def centre(st):
    c = st["pos"].mean(axis=0)
    return int(np.argmin(np.sum((st["pos"] - c) ** 2, axis=1)))


def test_tensor_linear_velocity_field_closed_form():
    """v = A x on a perfect cubic lattice: sum_j cm_ij x_ij (x) x_ij = c I by
    the lattice symmetry, so grad v = -c A and the update is F' = F - c dt A F
    (not F - c dt F A, nor with A^T); stress = solidK (F' F'^T - I)."""
    p, st = lattice()
    assert p.kspikydiff < 0
    A = np.array([[0.3, -1.2, 0.5], [0.8, 0.1, -0.4], [-0.6, 0.9, -0.2]])
    F0 = np.array([[1.1, 0.2, -0.1], [0.05, 0.95, 0.3], [-0.2, 0.1, 1.05]])
    st.a["vel"][:] = st["pos"] @ A.T
    F = np.tile(F0.reshape(9), (st.n, 1))
    S = np.zeros((st.n, 9))
    st.solid_tensor(F, S)
    i = centre(st)
    Fi = F[i].reshape(3, 3)
    AF = A @ F0
    c = (F0[0, 0] - Fi[0, 0]) / (p.dt * AF[0, 0])
    print("lattice constant c =", c)
    assert abs(c * p.dt) > 1e-3
    np.testing.assert_allclose(Fi, F0 - c * p.dt * AF, rtol=1e-10, atol=1e-12)
    assert np.max(np.abs(Fi - (F0 - c * p.dt * F0 @ A))) > 1e-3   # the other order differs
    np.testing.assert_allclose(S[i].reshape(3, 3), p.solidK * (Fi @ Fi.T - np.eye(3)),
                               rtol=1e-10, atol=1e-6)


def test_tensor_uniform_velocity_keeps_F():
    p, st = lattice(n=5, jitter=1e-3)
    st.a["vel"][:] = [0.3, -0.2, 0.1]
    F0 = np.array([[1.0, 0.1, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    F = np.tile(F0.reshape(9), (st.n, 1))
    S = np.zeros((st.n, 9))
    st.solid_tensor(F, S)
    np.testing.assert_array_equal(F, np.tile(F0.reshape(9), (st.n, 1)))
    want = p.solidK * (F0 @ F0.T - np.eye(3))
    np.testing.assert_allclose(S.reshape(-1, 3, 3), np.broadcast_to(want, (st.n, 3, 3)),
                               rtol=1e-12, atol=1e-9)


def test_force_uniform_stress_vanishes_inside():
    """Uniform stress: an interior particle's neighbourhood is symmetric, so
    its acceleration vanishes (to the FP32 lattice positions' asymmetry,
    ~1e-7 of the particle spacing); particles at the lattice's faces feel
    the full one-sided sum."""
    p, st = lattice()
    s0 = np.array([[2.0, 0.5, -1.0], [0.5, -3.0, 0.2], [-1.0, 0.2, 1.5]]) * 1e3
    S = np.tile(s0.reshape(9), (st.n, 1))
    st.solid_force(S)
    i = centre(st)
    a = st["accel"]
    print("centre", a[i], "max", np.max(np.abs(a)))
    assert np.max(np.abs(a)) > 0
    assert np.max(np.abs(a[i])) <= 1e-5 * np.max(np.abs(a))


def test_force_conserves_momentum():
    """Equal masses and densities: each pair adds cm (sigma_i + sigma_j) x_ij
    to i and the negative to j, so the accelerations sum to zero (a
    difference sigma_i - sigma_j would not)."""
    p, st = lattice(n=6, jitter=1e-3, seed=5)
    rng = np.random.default_rng(3)
    S = rng.normal(size=(st.n, 9)) * 1e3
    st.solid_force(S)
    a = st["accel"]
    tot = np.abs(a.sum(axis=0)).max()
    print("sum", a.sum(axis=0), "scale", np.abs(a).sum())
    assert np.abs(a).sum() > 0
    assert tot <= 1e-12 * np.abs(a).sum()


def test_surface_tension_pair_antisymmetry():
    p, st = lattice(n=6, jitter=1e-3, seed=9)
    p.surfaceTensionK = 2.0
    p.surfaceTensionFluidC = 0.5
    st.a["accel"][:] = 0.0
    st.surface_tension()
    a = st["accel"]
    assert np.abs(a).sum() > 0
    assert np.abs(a.sum(axis=0)).max() <= 1e-12 * np.abs(a).sum()


def test_surface_tension_pair_and_cutoff():
    """Two particles: the force on 0 is C cos(1.5 pi r / tR) / r x_01, and the
    reference's cutoff compares r^2 with tR = K h (not tR^2)."""
    h = 0.021
    K = 0.01
    tR = np.sqrt(np.float32(K * K) * np.float32(h * h))
    for r, inside in ((0.0140, True), (0.0150, False)):
        p, disp, calc, info = sphx.scene_dambreak(2, 1, 1, r, (0.1, 0.1, 0.1), h, 0, 0.0, 1, 0)
        p.surfaceTensionK = K
        p.surfaceTensionFluidC = 0.7
        st = orc.State(disp, calc, p, np.float64)
        st.grid_build()
        st.a["accel"][:] = 0.0
        st.surface_tension()
        d = st["pos"][0] - st["pos"][1]
        rr = np.linalg.norm(d)
        assert (rr * rr <= tR) == inside
        want = 0.7 * np.cos(1.5 * 3.14159 * rr / tR) / rr * d if inside else np.zeros(3)
        np.testing.assert_allclose(st["accel"][0], want, rtol=1e-6, atol=1e-12)
        np.testing.assert_allclose(st["accel"][1], -want, rtol=1e-6, atol=1e-12)
