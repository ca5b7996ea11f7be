"""sphx_rpy_walkthrough_noalloc (static device scratch, no allocation)
against sphx_rpy_walkthrough on identical device buffers: bitwise equal
results on two calls in a row, the second with fewer spheres, one with a
live MPM node grid as the background velocity; n above the scratch's
maximum is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

import sphx

pytestmark = pytest.mark.gpu


def buffers(pos, force):
    n = len(pos)
    disp = np.zeros(n, sphx.display_dtype)
    disp["pos"] = pos
    arrs = [sphx.DeviceArray(np.zeros((n, 3), np.float32)), sphx.DeviceArray(disp),
            sphx.DeviceArray(np.zeros((n, 3), np.float32)), sphx.DeviceArray(force)]
    return arrs


def buflist(arrs, dt, ss):
    u, d, w, f = arrs
    return sphx.StokesianBufList(u.ptr.value, d.ptr.value, w.ptr.value, f.ptr.value, dt, ss)


def node_context():
    """A sand context whose node grid holds a smooth velocity field."""
    p, disp, calc, info = sphx.scene_sand(n=512, seed=31)
    ctx = sphx.Context(p)
    ctx.upload(disp, calc)
    ctx.mpm_setup(1, 1)
    res = (p.mpmRes.x, p.mpmRes.y, p.mpmRes.z)
    y, z, x = np.meshgrid(*[np.arange(r) for r in (res[1], res[2], res[0])], indexing="ij")
    xs = np.stack([x.ravel(), y.ravel(), z.ravel()], 1) * p.mpmcellsize
    v = np.stack([np.sin(3 * xs[:, 1]), 0.5 * xs[:, 0] - xs[:, 2], np.cos(2 * xs[:, 0])], 1)
    ctx.mpm_set_nodes(v.astype(np.float32), np.ones(len(v), np.float32))
    return ctx


def run_both(nodes, pos, force, dt, ss, with_force):
    L = sphx.lib()
    a = buffers(pos, force)
    b = buffers(pos, force)
    n = len(pos)
    h = nodes.h if nodes is not None else None
    print(f"walkthrough n={n} nodes={nodes is not None} force={with_force}", flush=True)
    sphx.check(L.sphx_rpy_walkthrough(h, C.byref(buflist(a, dt, ss)), n, with_force),
               "walkthrough")
    sphx.check(L.sphx_rpy_walkthrough_noalloc(h, C.byref(buflist(b, dt, ss)), n, with_force),
               "walkthrough_noalloc")
    for x, y, name in zip(a, b, ("U", "disp", "Omega", "force")):
        np.testing.assert_array_equal(y.get(), x.get(), err_msg=name)
    return a


def test_noalloc_walkthrough_matches_allocating_one(gpu):
    rng = np.random.default_rng(41)
    nodes = node_context()
    try:
        n1 = 700
        pos1 = rng.uniform(0.1, 0.8, (n1, 3)).astype(np.float32)
        f1 = rng.normal(0, 1, (n1, 3)).astype(np.float32)
        a = run_both(nodes, pos1, f1, 1e-4, 1.0, 0)
        assert np.abs(a[0].get()).max() > 0
    finally:
        nodes.close()
    n2 = 260
    pos2 = rng.uniform(0, 40, (n2, 3)).astype(np.float32)
    f2 = rng.normal(0, 1, (n2, 3)).astype(np.float32)
    a = run_both(None, pos2, f2, 0.01, 1.0, 1)
    assert np.abs(a[1].get()["pos"] - pos2).max() > 0


def test_noalloc_walkthrough_refuses_more_than_its_scratch(gpu):
    L = sphx.lib()
    a = buffers(np.zeros((4, 3), np.float32), np.zeros((4, 3), np.float32))
    rc = L.sphx_rpy_walkthrough_noalloc(None, C.byref(buflist(a, 0.01, 1.0)), 65537, 1)
    assert rc != 0
    assert "scratch" in sphx.last_error()
