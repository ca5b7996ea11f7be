"""State I/O in the reference's formats (host only, no GPU): scene.xml
parameters as FluidSystem::ParseXML reads them, the text frame of
FluidSystem::outputFile, and AoS checkpoints."""
import math

import numpy as np
import pytest

import sphx

SCENE = """<?xml version="1.0"?>
<!-- a scene file in the reference's layout -->
<Fluid>
	<AccelLimit>	1500.0		</AccelLimit>
	<VelLimit>		100		</VelLimit>
	<Gravity>	0,-9.8,0.5		</Gravity>
	<VolMin> -60, -20, -60 </VolMin>
	<VolMax> 60,  200,  60 </VolMax>
</Fluid>

<MultiScene id="1">
	<Num> 1000 </Num>
	<SimScale> 0.01 </SimScale>
</MultiScene>

<MultiScene id="2">
	<Num> 500000 </Num>
	<SimScale>		0.005		</SimScale>
	<DT>			0.00005 </DT>
	<Radius> 0.01 </Radius>
	<SmoothRadius> 0.015 </SmoothRadius>
	<Viscosity>	 1.5		</Viscosity>
	<Mass> 0.00010543 </Mass>
	<RestDensity> 1000 </RestDensity>
	<BoundStiff>	50000.0	</BoundStiff>
	<ScaleP>	1	</ScaleP>
	<FluidCount> 3 </FluidCount>
	<SoftBoundMin>	-40, 0, -40 </SoftBoundMin>
	<SoftBoundMax>	 40, 100, 40 </SoftBoundMax>
	<VolMin0>		 -20,   2, -20	</VolMin0>
	<VolMax0>		  20,  10,  20	</VolMax0>
	<!--VolMin0>		 -7,   2, -7	</VolMin0-->
	<VolMin1>		 -7,   2, -7	</VolMin1>
	<VolMax1>		  7,  16,  7	</VolMax1>
	<MassRatio>    1.,    2.,   3.  </MassRatio>
	<DensityRatio> 1.,    0.5,   0.25  </DensityRatio>
	<ViscRatio>    1,   4,   9   </ViscRatio>
	<SolidK> 1000000000 </SolidK>
	<SolidG> 100000000 </SolidG>
	<DP_Phi>  15  </DP_Phi>
	<DP_Psi>  10  </DP_Psi>
	<DP_cohesion> 10.0 </DP_cohesion>
</MultiScene>

<BoundInfo id="1">
	<Viscosity>	100.0		</Viscosity>
	<RestDensity>	600.0		</RestDensity>
	<Mass>		0.00010543	</Mass>
</BoundInfo>
"""


def test_scene_xml_values():
    p, sx = sphx.scene_xml_parse(SCENE, 2)
    f32 = np.float32
    assert (p.acclimit, p.vlimit) == (f32(1500.0), f32(100.0))
    assert (p.gravity.x, p.gravity.y, p.gravity.z) == (0.0, f32(-9.8), 0.5)
    assert (p.volmin.x, p.volmin.y, p.volmin.z) == (-60.0, -20.0, -60.0)
    assert (p.volmax.x, p.volmax.y, p.volmax.z) == (60.0, 200.0, 60.0)
    assert p.maxNum == 500000 and sx.maxPointNum == 500000
    assert p.simscale == f32(0.005) and p.dt == f32(0.00005)
    assert p.radius == f32(0.01) and p.smoothradius == f32(0.015)
    assert p.viscosity == f32(1.5) and p.mass == f32(0.00010543)
    assert p.restdensity == 1000.0 and p.boundstiff == 50000.0
    assert (p.softminx.x, p.softminx.y, p.softminx.z) == (-40.0, 0.0, -40.0)
    assert (p.softmaxx.x, p.softmaxx.y, p.softmaxx.z) == (40.0, 100.0, 40.0)
    assert p.solidK == f32(1e9) and p.solidG == f32(1e8)
    assert sx.scaleP == 1.0 and sx.fluidnum == 3
    # the commented-out VolMin0 is skipped
    assert (sx.fvmin[0].x, sx.fvmin[0].y, sx.fvmin[0].z) == (-20.0, 2.0, -20.0)
    assert (sx.fvmax[1].x, sx.fvmax[1].y, sx.fvmax[1].z) == (7.0, 16.0, 7.0)
    assert list(sx.massratio) == [1.0, 2.0, 3.0]
    assert list(sx.densratio) == [1.0, 0.5, 0.25]
    assert list(sx.viscratio) == [1.0, 4.0, 9.0]
    # boundary block
    assert p.bvisc == 100.0 and p.bRestdensity == 600.0 and p.bmass == f32(0.00010543)


def test_scene_xml_drucker_prager_constants():
    """a_phi = 2 sin(phi) / (sqrt 3 (3 - sin phi)), k_c = 3 c cos(psi) /
    (sqrt 3 (3 - sin phi)) with the angles converted from degrees."""
    p, _ = sphx.scene_xml_parse(SCENE, 2)
    phi, psi = math.radians(15), math.radians(10)
    assert p.a_phi == pytest.approx(2 * math.sin(phi) / math.sqrt(3) / (3 - math.sin(phi)), rel=1e-6)
    assert p.a_psi == pytest.approx(2 * math.sin(psi) / math.sqrt(3) / (3 - math.sin(psi)), rel=1e-6)
    assert p.k_c == pytest.approx(3 * 10.0 * math.cos(psi) / math.sqrt(3) / (3 - math.sin(phi)),
                                  rel=1e-6)


def test_scene_xml_absent_values_and_errors():
    p, sx = sphx.scene_xml_parse(SCENE, 1)
    assert p.maxNum == 1000 and p.simscale == np.float32(0.01)
    assert p.dt == 0.0 and p.smoothradius == 0.0 and list(sx.massratio) == [0.0, 0.0, 0.0]
    with pytest.raises(sphx.SphxError, match="scene 7 not found"):
        sphx.scene_xml_parse(SCENE, 7)
    with pytest.raises(sphx.SphxError, match="missing Fluid"):
        sphx.scene_xml_parse("<MultiScene id=\"2\"></MultiScene>", 2)


def test_frame_round_trip(tmp_path):
    p, disp, calc, info = sphx.scene_dambreak(6, 5, 4, 0.01, (0.12, 0.1, 0.08), 0.02, 1, 1e-4,
                                              3, 1)
    path = tmp_path / "data_0000.txt"
    sphx.frame_write(path, disp, calc)
    lines = path.read_text().splitlines()
    assert int(lines[0]) == len(disp) and len(lines) == len(disp) + 1
    assert len(lines[1].split()) == 6
    d2, c2 = sphx.frame_read(path)
    assert len(d2) == len(disp)
    np.testing.assert_allclose(d2["pos"], disp["pos"], rtol=0, atol=5e-7)
    np.testing.assert_allclose(c2["mass"], calc["mass"], rtol=0, atol=5e-7)
    np.testing.assert_allclose(c2["restdens"], calc["restdens"], rtol=0, atol=5e-7)
    np.testing.assert_array_equal(d2["type"], disp["type"])


def test_checkpoint_round_trip_and_checks(tmp_path):
    p, disp, calc, info = sphx.scene_dambreak(6, 5, 4, 0.01, (0.12, 0.1, 0.08), 0.02, 1, 1e-4,
                                              3, 1)
    rng = np.random.default_rng(4)
    calc["vel"] = rng.normal(size=calc["vel"].shape)
    calc["stress"] = rng.normal(size=calc["stress"].shape)
    path = tmp_path / "state.ckp"
    sphx.checkpoint_write(path, p, disp, calc)
    p2, d2, c2 = sphx.checkpoint_read(path)
    assert bytes(p2) == bytes(p)
    assert d2.tobytes() == disp.tobytes() and c2.tobytes() == calc.tobytes()
    bad = tmp_path / "bad.ckp"
    bad.write_bytes(b"NOTACKPT" + path.read_bytes()[8:])
    with pytest.raises(sphx.SphxError, match="not a checkpoint"):
        sphx.checkpoint_read(bad)
    short = tmp_path / "short.ckp"
    short.write_bytes(path.read_bytes()[:-100])
    with pytest.raises(sphx.SphxError, match="truncated"):
        sphx.checkpoint_read(short)


def boundary_params(spacing=0.5, simscale=1.0):
    p = sphx.ParamCarrier()
    p.bRestdensity = 1000.0
    p.bmass = 1000.0 * (spacing * simscale) ** 3
    p.simscale = simscale
    p.bvisc = 7.0
    return p


def wall_count(mn, mx, s):
    """Columns ceil(extent / s) in x and z; layers min.y, min.y + s, ... <=
    max.y with the spacing added in float32 as the reference's loop does."""
    cx = int(np.ceil(np.float32(mx[0] - mn[0]) / np.float32(s)))
    cz = int(np.ceil(np.float32(mx[2] - mn[2]) / np.float32(s)))
    ny, y = 0, np.float32(mn[1])
    while y <= np.float32(mx[1]):
        ny += 1
        y = np.float32(y + np.float32(s))
    return cx * cz * ny


def test_boundary_wall_lattice():
    p = boundary_params(0.5)
    s = np.float32(np.cbrt(np.float32(p.bmass / p.bRestdensity)))
    disp, calc = sphx.boundary_parse("WALL\n-2, 0, -1.5\n2, 3, 1.5\n", p, start=5)
    assert len(disp) == wall_count((-2, 0, -1.5), (2, 3, 1.5), s) == 8 * 6 * 7
    pos = disp["pos"].astype(np.float64)
    assert np.all(disp["type"] == 1)
    np.testing.assert_array_equal(calc["bornid"], np.arange(5, 5 + len(disp)))
    assert np.all(calc["mass"] == np.float32(p.bmass)) and np.all(calc["visc"] == 7.0)
    assert np.all(calc["restdens"] == 1000.0)
    np.testing.assert_array_equal(calc["deformGrad"][:, [0, 4, 8]], 1.0)
    # lattice columns from min, layers from min.y
    kx = (pos[:, 0] + 2) / s
    kz = (pos[:, 2] + 1.5) / s
    np.testing.assert_allclose(kx, np.round(kx), atol=1e-5)
    np.testing.assert_allclose(kz, np.round(kz), atol=1e-5)
    assert set(np.round(kx).astype(int)) == set(range(8))
    assert set(np.round(kz).astype(int)) == set(range(6))
    # colour alpha 0 above z = 0, 0.2 at and below
    np.testing.assert_array_equal(disp["color"][:, 3], np.where(pos[:, 2] > 0, 0.0, np.float32(0.2)))


def test_boundary_file_commands_and_open_box():
    p = boundary_params(0.5)
    s = np.float32(np.cbrt(np.float32(p.bmass / p.bRestdensity)))
    walls = [((-5, 0, -5), (-4, 5, 5)), ((4, 0, -5), (5, 5, 5)), ((-5, -1, -5), (5, 0, 5))]
    text = "".join(f"WALL\n{a[0]}, {a[1]}, {a[2]}\n {b[0]},{b[1]},{b[2]}\n" for a, b in walls)
    disp, calc = sphx.boundary_parse(text, p)
    assert len(disp) == sum(wall_count(a, b, s) for a, b in walls)
    # OPEN_BOX = ground, front, back, left, right walls around the box
    mn, mx, t = (-3.0, 0.0, -2.0), (3.0, 4.0, 2.0), 1.0
    box = [((mn[0], mn[1] - t, mn[2]), (mx[0], mn[1], mx[2])),
           ((mn[0] - t, mn[1], mn[2] - t), (mx[0] + t, mx[1], mn[2])),
           ((mn[0] - t, mn[1], mx[2]), (mx[0] + t, mx[1], mx[2] + t)),
           ((mn[0] - t, mn[1], mn[2]), (mn[0], mx[1], mx[2])),
           ((mx[0], mn[1], mn[2]), (mx[0] + t, mx[1], mx[2]))]
    d2, c2 = sphx.boundary_parse("OPEN_BOX\n-3, 0, -2\n3, 4, 2\n1", p)
    assert len(d2) == sum(wall_count(a, b, s) for a, b in box)
    parts = [sphx.boundary_parse(f"WALL {a[0]},{a[1]},{a[2]} {b[0]},{b[1]},{b[2]}", p)[0]["pos"]
             for a, b in box]
    np.testing.assert_array_equal(d2["pos"], np.concatenate(parts))
    # a cap below the count: the count is still reported, nothing written past it
    L = sphx.lib()
    n = sphx.C.c_int()
    buf_d = np.zeros(10, sphx.display_dtype)
    buf_c = np.zeros(10, sphx.calc_dtype)
    assert L.sphx_boundary_parse(b"OPEN_BOX -3,0,-2 3,4,2 1", sphx.C.byref(p), sphx.ptr(buf_d),
                                 sphx.ptr(buf_c), 2, 10, sphx.C.byref(n)) == 0
    assert n.value == len(d2)
    np.testing.assert_array_equal(buf_d["pos"][2:], d2["pos"][:8])
    assert np.all(buf_d["pos"][:2] == 0)
