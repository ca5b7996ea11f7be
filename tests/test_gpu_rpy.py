"""Stokesian mobility on the GPU: the MFMA polydisperse RPY apply against
the numpy FP64 fixture and the FP64 oracle (all three precision modes,
padding, split counts), and the reference walkthrough / dense product on
device buffers.

Tolerances (normwise, infinity norm, vs FP64): bf16x3 with the
second-order split terms (mode 6) 1e-5; bf16x3 (mode 3) 1e-4; plain bf16
(mode 1) 2e-2, shown for contrast; walkthrough 1e-5; dense product 1e-6."""
from pathlib import Path

import numpy as np
import pytest

import orc
import sphx

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
BOUND = {1: 2e-2, 3: 1e-4, 6: 1e-5}


@pytest.mark.parametrize("mode", [1, 3, 6])
def test_rpy_apply_matches_fixture(gpu, mode):
    d = np.load(GOLDEN / "rpy_n256.npz")
    n, K = len(d["pos"]), d["F"].shape[1]
    with sphx.Mobility(n, K) as mob:
        mob.set_spheres(d["pos"], d["rad"], float(d["mu"]))
        U = mob.apply(d["F"], mode)
    e = orc.relerr(U, d["U"])
    print(f"mode {mode}: gpu vs fp64 {e:.2e}")
    assert e <= BOUND[mode]


@pytest.mark.parametrize("n,K,split", [(1000, 16, 7), (97, 32, 1), (2048, 64, 16)])
def test_rpy_apply_matches_oracle(gpu, n, K, split):
    rng = np.random.default_rng(n)
    side = 2.0 * (n / 4096) ** (1 / 3)
    # a jittered lattice keeps the spheres apart without rejection sampling
    m = int(np.ceil(n ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(m)] * 3, indexing="ij"), -1).reshape(-1, 3)[:n]
    pos = ((g + 0.5 + rng.uniform(-0.2, 0.2, (n, 3))) * side / m).astype(np.float32)
    rad = rng.uniform(0.01, 0.03, n).astype(np.float32)
    F = rng.standard_normal((3 * n, K)).astype(np.float32)
    U64 = orc.rpy_apply(pos, rad, 1.0, F.astype(np.float64))
    with sphx.Mobility(n, K) as mob:
        mob.set_spheres(pos, rad, 1.0)
        mob.set_split(split)
        U = mob.apply(F, 6)
    e = orc.relerr(U, U64)
    print(f"n={n} K={K} split={split}: gpu vs fp64 {e:.2e}")
    assert e <= BOUND[6]


def test_walkthrough_and_dense_product(gpu):
    import ctypes as C
    rng = np.random.default_rng(3)
    n = 300
    pos = rng.uniform(0, 40, (n, 3)).astype(np.float32)
    f0 = rng.normal(0, 1, (n, 3)).astype(np.float32)
    disp = np.zeros(n, sphx.display_dtype)
    disp["pos"] = pos
    d_disp = sphx.DeviceArray(disp)
    d_force = sphx.DeviceArray(f0)
    d_u = sphx.DeviceArray(np.zeros((n, 3), np.float32))
    d_w = sphx.DeviceArray(np.zeros((n, 3), np.float32))
    b = sphx.StokesianBufList(d_u.ptr.value, d_disp.ptr.value, d_w.ptr.value, d_force.ptr.value,
                              0.01, 1.0)
    L = sphx.lib()
    sphx.check(L.sphx_rpy_walkthrough(None, C.byref(b), n, 1), "walkthrough")
    f_gpu = d_force.get()
    # near-field force (Kern_getForce :257-300) on the host, FP64
    d = pos[None, :, :].astype(np.float64) - pos[:, None, :]
    r = np.linalg.norm(d, axis=-1)
    np.fill_diagonal(r, np.inf)
    fac = np.where(r < 1.5, 5 * (np.exp(1.5 / r) - 2.7128 * r / 1.5), 0.0)
    f_ref = f0 + (-(d * fac[:, :, None])).sum(1)
    assert orc.relerr(f_gpu, f_ref) <= 1e-5
    u, w, pn = orc.rpy_walkthrough(pos, f_gpu, np.zeros((n, 3)), 0.01)
    out = d_disp.get()
    eu, ew = orc.relerr(d_u.get(), u), orc.relerr(d_w.get(), w)
    print(f"walkthrough: U {eu:.2e} Omega {ew:.2e}")
    assert eu <= 1e-5 and ew <= 1e-5
    assert orc.relerr(out["pos"], pn) <= 1e-6
    # getMobU: dense (6n x 6n) product, through the legacy C alias
    m = 40
    mat = rng.normal(0, 1, (6 * m, 6 * m)).astype(np.float32)
    fv = rng.normal(0, 1, 6 * m).astype(np.float32)
    d_mat, d_f = sphx.DeviceArray(mat), sphx.DeviceArray(fv)
    d_out = sphx.DeviceArray(np.zeros(6 * m, np.float32))
    L.getMobU.argtypes = [C.c_void_p] * 3 + [C.c_int]
    L.getMobU.restype = None
    L.getMobU(d_mat.ptr, d_f.ptr, d_out.ptr, m)
    ref = mat.astype(np.float64) @ fv.astype(np.float64)
    assert orc.relerr(d_out.get(), ref) <= 1e-6


def test_walkthrough_background_from_node_grid(gpu):
    """Kern_interpolateVelocity (Mobility.cu:322-366): trilinear node
    velocities at the spheres, divided by simscale.  A linear node field
    v_I = A x_I + b is reproduced exactly inside the grid, so the
    background is A x + b in closed form; the walkthrough then follows the
    FP64 oracle from that background."""
    import ctypes as C
    p, disp, calc, info = sphx.scene_sand(n=512, seed=32)
    rng = np.random.default_rng(9)
    A = rng.normal(0, 1.0, (3, 3))
    b = rng.normal(0, 1.0, 3)
    ss = 0.5
    with sphx.Context(p) as nodes:
        nodes.upload(disp, calc)
        nodes.mpm_setup(1, 1)
        res = (p.mpmRes.x, p.mpmRes.y, p.mpmRes.z)
        y, z, x = np.meshgrid(*[np.arange(r) for r in (res[1], res[2], res[0])], indexing="ij")
        xs = np.stack([x.ravel(), y.ravel(), z.ravel()], 1) * np.float64(p.mpmcellsize)
        nodes.mpm_set_nodes((xs @ A.T + b).astype(np.float32), np.ones(len(xs), np.float32))
        n = 200
        hi = (res[0] - 1) * p.mpmcellsize
        pos = rng.uniform(0.05 * hi, 0.95 * hi, (n, 3)).astype(np.float32)
        # forces small enough that the background is a visible part of U
        # (the spheres are ~0.1 apart in units of their radius)
        f0 = rng.normal(0, 1e-5, (n, 3)).astype(np.float32)
        sd = np.zeros(n, sphx.display_dtype)
        sd["pos"] = pos
        d_disp = sphx.DeviceArray(sd)
        d_force = sphx.DeviceArray(f0)
        d_u = sphx.DeviceArray(np.zeros((n, 3), np.float32))
        d_w = sphx.DeviceArray(np.zeros((n, 3), np.float32))
        bl = sphx.StokesianBufList(d_u.ptr.value, d_disp.ptr.value, d_w.ptr.value,
                                   d_force.ptr.value, 1e-5, ss)
        L = sphx.lib()
        sphx.check(L.sphx_rpy_walkthrough(nodes.h, C.byref(bl), n, 0), "walkthrough")
        u_gpu, w_gpu, p_gpu = d_u.get(), d_w.get(), d_disp.get()["pos"]
    ubg = (pos.astype(np.float64) @ A.T + b) / ss
    u, w, pn = orc.rpy_walkthrough(pos, f0, ubg, 1e-5)
    eu, ew = orc.relerr(u_gpu, u), orc.relerr(w_gpu, w)
    print(f"walkthrough with node background: U {eu:.2e} Omega {ew:.2e}")
    assert eu <= 1e-5 and ew <= 1e-5
    assert orc.relerr(p_gpu, pn) <= 1e-6
    # the background is a visible part of U
    assert np.abs(ubg).max() > 0.05 * np.abs(u).max()
