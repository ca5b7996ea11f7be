"""Grid build on the GPU vs the oracle: hash, stable sort, cell ranges and
reorder must be bit-exact (fluid_system_kern.cu:36-121,
fluid_system_host.cu:147-210)."""
import numpy as np
import pytest

import orc
import sphx
from helpers import shuffled_scene

pytestmark = pytest.mark.gpu

UNDEF = 0xFFFFFFFF


def gpu_build(p, disp, calc):
    with sphx.Context(p) as ctx:
        ctx.upload(disp, calc)
        ctx.call("grid_build")
        g = ctx.download_grid()
        d, c = ctx.download()
    return g, d, c


def oracle_build(p, disp, calc):
    st = orc.State(disp, calc, p)
    st.grid_build()
    return st


@pytest.mark.parametrize("n_out", [0, 5])
def test_grid_build_bitexact(gpu, n_out):
    p, disp, calc, info = shuffled_scene(n_out=n_out)
    print(f"grid build: n={len(disp)} gridtotal={p.gridtotal} n_out={n_out}", flush=True)
    g, d, c = gpu_build(p, disp, calc)
    st = oracle_build(p, disp, calc)
    assert np.array_equal(g["mgcell"], st["mgcell"])
    assert np.array_equal(g["midsort"], st["midsort"])
    assert np.array_equal(g["gridoff"], st["gridoff"])
    assert np.array_equal(g["gridcnt"], st["gridcnt"])
    valid = st["mgcell"] != UNDEF
    assert np.array_equal(g["mgndx"][valid], st["mgndx"][valid])
    assert np.array_equal(g["MFidTable"][calc["bornid"][st["midsort"][valid]]],
                          np.where(valid)[0])
    # reordered particle data: bit-exact copy (out-of-grid slots parked)
    assert np.array_equal(d["pos"], st["pos"].astype(np.float32))
    for k in ["vel", "veleval", "accel"]:
        assert np.array_equal(c[k], st[k])
    for k in ["pressure", "dens", "restdens", "mass", "visc"]:
        assert np.array_equal(c[k], st[k])
    assert np.array_equal(d["type"], st["type"])
    assert np.array_equal(c["bornid"], st["bornid"])
    # cold fields travel with the permutation (valid slots)
    src = st["midsort"][valid]
    assert np.array_equal(d["color"][valid], disp["color"][src])
    assert np.array_equal(c["deformGrad"][valid], calc["deformGrad"][src])
    assert np.array_equal(c["X"][valid], calc["X"][src])
    if n_out:
        assert np.all(d["pos"][~valid] == -1000.0)


def test_grid_build_twice_nearly_sorted(gpu):
    """Second build on already-sorted data (the per-step case)."""
    p, disp, calc, info = shuffled_scene(n_out=3, seed=11)
    with sphx.Context(p) as ctx:
        ctx.upload(disp, calc)
        ctx.call("grid_build")
        d1, c1 = ctx.download()
        ctx.call("grid_build")
        g = ctx.download_grid()
        d, c = ctx.download()
    st = oracle_build(p, d1, c1)
    assert np.array_equal(g["mgcell"], st["mgcell"])
    assert np.array_equal(g["midsort"], st["midsort"])
    assert np.array_equal(g["gridoff"], st["gridoff"])
    assert np.array_equal(g["gridcnt"], st["gridcnt"])
    assert np.array_equal(d["pos"], st["pos"])


@pytest.mark.parametrize("frac", [0.03, 0.5])
def test_grid_rebuild_after_moves(gpu, frac):
    """Rebuild after a fraction of the particles moved to other cells: the
    sort's digit runs see nearly sorted keys with strays (or, at frac 0.5,
    mostly unsorted ones), and the cell ranges start their search from the
    last build's ranges, which are then off."""
    p, disp, calc, info = shuffled_scene(n_out=3, seed=12)
    rng = np.random.default_rng(5)
    with sphx.Context(p) as ctx:
        ctx.upload(disp, calc)
        ctx.call("grid_build")
        d1, c1 = ctx.download()
        cell = 1.0 / p.gridIdfac.x
        moved = rng.random(len(d1)) < frac
        step = rng.integers(-2, 3, size=(len(d1), 3)).astype(np.float32) * np.float32(cell)
        inside = d1["pos"][:, 0] > -999.0
        sel = moved & inside
        d1["pos"][sel] = d1["pos"][sel] + step[sel]
        print(f"rebuild: n={len(d1)} moved={int(sel.sum())}", flush=True)
        ctx.upload(d1, c1)
        ctx.call("grid_build")
        g = ctx.download_grid()
        d, c = ctx.download()
    st = oracle_build(p, d1, c1)
    assert np.array_equal(g["mgcell"], st["mgcell"])
    assert np.array_equal(g["midsort"], st["midsort"])
    assert np.array_equal(g["gridoff"], st["gridoff"])
    assert np.array_equal(g["gridcnt"], st["gridcnt"])
    valid = st["mgcell"] != UNDEF
    assert np.array_equal(g["mgndx"][valid], st["mgndx"][valid])
    assert np.array_equal(d["pos"], st["pos"].astype(np.float32))


def test_grid_empty_and_single(gpu):
    p, disp, calc, info = shuffled_scene(nx=2, ny=2, nz=2, layers=0)
    p.num = 0
    for n in [0, 1]:
        with sphx.Context(p) as ctx:
            if n:
                ctx.upload(disp[:n], calc[:n])
            ctx.call("grid_build")
            g = ctx.download_grid()
            assert ctx.num == n
            assert g["gridcnt"].sum() == n


def test_grid_all_out_of_range(gpu):
    p, disp, calc, info = shuffled_scene(nx=4, ny=4, nz=4, layers=0)
    disp["pos"][:] = [-9.0, -9.0, -9.0]
    g, d, c = gpu_build(p, disp, calc)
    assert np.all(g["mgcell"] == UNDEF)
    assert np.array_equal(g["midsort"], np.arange(len(disp)))
    assert g["gridcnt"].sum() == 0


def test_grid_full_size_properties(gpu):
    """1,048,576 particles scattered over a 52^3 grid: size-independent
    properties of a stable sort and of the cell table."""
    n = 1 << 20
    rng = np.random.default_rng(5)
    p, disp0, calc0, info = sphx.scene_dambreak(2, 2, 2, 0.0091, (1.0, 1.0, 1.0), 0.01, 0, 0.0,
                                                1, 1)
    p.num = p.maxNum = n
    disp = np.zeros(n, sphx.display_dtype)
    calc = np.zeros(n, sphx.calc_dtype)
    lo = np.array([p.gridmin.x, p.gridmin.y, p.gridmin.z], np.float32)
    hi = np.array([p.gridmax.x, p.gridmax.y, p.gridmax.z], np.float32)
    disp["pos"] = (lo + (hi - lo) * rng.random((n, 3))).astype(np.float32)
    calc["bornid"] = np.arange(n)
    calc["mass"] = 1.0
    print(f"full-size grid: n={n} gridtotal={p.gridtotal}", flush=True)
    g, d, c = gpu_build(p, disp, calc)
    key, ids = g["mgcell"].astype(np.int64), g["midsort"].astype(np.int64)
    # permutation
    assert np.array_equal(np.sort(ids), np.arange(n))
    # brute-force hash of every particle (FP32, truncation, margin rule)
    pos = disp["pos"][ids]
    fac = np.array([p.gridIdfac.x, p.gridIdfac.y, p.gridIdfac.z], np.float32)
    gc = ((pos - lo) * fac).astype(np.int64)
    res = np.array([p.gridres.x, p.gridres.y, p.gridres.z])
    ok = np.all((gc >= 1) & (gc < res - 1), axis=1)
    want = np.where(ok, (gc[:, 1] * res[2] + gc[:, 2]) * res[0] + gc[:, 0], UNDEF)
    assert np.array_equal(key, want)
    # sorted, stable (equal keys keep index order)
    assert np.all(np.diff(key) >= 0)
    same = np.diff(key) == 0
    assert np.all(np.diff(ids)[same] > 0)
    # cell table consistent with the keys
    cnt = g["gridcnt"]
    assert cnt.sum() == np.count_nonzero(key != UNDEF)
    nz = np.nonzero(cnt)[0]
    assert np.all(key[g["gridoff"][nz]] == nz)
    assert np.array_equal(np.bincount(key[key != UNDEF], minlength=p.gridtotal), cnt)
    # reorder gathered the right particles
    assert np.array_equal(c["bornid"][key != UNDEF], ids[key != UNDEF])
