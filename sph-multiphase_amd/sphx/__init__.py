"""ctypes binding of libsphx.so (the C ABI declared in include/sphx.h).

Particle data crosses the boundary in the reference's AoS formats
(SPH_Solver/common_header.h:28-137); here they are numpy structured dtypes
with the same byte layout.  The library is the HIP implementation; there is
no CPU fallback: without the built library or without a GPU the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent.parent
LIB_PATH = _PKG / "libsphx.so"

f3 = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4")])
f4 = np.dtype([("x", "<f4"), ("y", "<f4"), ("z", "<f4"), ("w", "<f4")])
i3 = np.dtype([("x", "<i4"), ("y", "<i4"), ("z", "<i4")])

display_dtype = np.dtype([("pos", "<f4", (3,)), ("color", "<f4", (4,)), ("type", "<i4")])
calc_dtype = np.dtype([
    ("vel", "<f4", (3,)), ("veleval", "<f4", (3,)), ("pressure", "<f4"), ("dens", "<f4"),
    ("restdens", "<f4"), ("mass", "<f4"), ("visc", "<f4"), ("bornid", "<i4"),
    ("accel", "<f4", (3,)), ("deformGrad", "<f4", (9,)), ("stress", "<f4", (9,)),
    ("B", "<f4", (9,)), ("solidG", "<f4"), ("solidK", "<f4"), ("X", "<f4", (3,)),
    ("invA", "<f4", (9,)),
])
assert display_dtype.itemsize == 32 and calc_dtype.itemsize == 224

TYPE_FLUID, TYPE_BOUNDARY, TYPE_ELASTIC, TYPE_GRANULAR = 0, 1, 2, 3
GRID_UNDEF = 0xFFFFFFFF


class Float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

    def tolist(self):
        return [self.x, self.y, self.z]


class Int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class ParamCarrier(C.Structure):
    """common_header.h:28-109 (464 bytes)."""
    _fields_ = [
        ("acclimit", C.c_float), ("vlimit", C.c_float), ("gravity", Float3),
        ("volmin", Float3), ("volmax", Float3), ("softminx", Float3), ("softmaxx", Float3),
        ("num", C.c_int), ("maxNum", C.c_int), ("dt", C.c_float), ("simscale", C.c_float),
        ("viscosity", C.c_float), ("restdensity", C.c_float), ("mass", C.c_float),
        ("radius", C.c_float), ("smoothradius", C.c_float), ("scalep", C.c_float),
        ("massArr", C.c_float * 3), ("densArr", C.c_float * 3), ("viscArr", C.c_float * 3),
        ("boundstiff", C.c_float), ("kpoly6", C.c_float), ("kspiky", C.c_float),
        ("kspikydiff", C.c_float), ("klaplacian", C.c_float), ("kspline", C.c_float),
        ("bmass", C.c_float), ("bRestdensity", C.c_float), ("bvisc", C.c_float),
        ("cellsize", C.c_float), ("searchnum", C.c_int), ("neighbornum", C.c_int),
        ("neighborid", C.c_int * 27), ("gridmin", Float3), ("gridmax", Float3),
        ("gridsize", Float3), ("gridIdfac", Float3), ("gridres", Int3), ("gridtotal", C.c_int),
        ("mpmcellsize", C.c_float), ("mpmXmin", Float3), ("mpmXmax", Float3),
        ("mpmRes", Int3), ("mpmNodeNum", C.c_int), ("solidK", C.c_float),
        ("solidG", C.c_float), ("a_phi", C.c_float), ("a_psi", C.c_float), ("k_c", C.c_float),
        ("w_deltax", C.c_float), ("fijn", C.c_float), ("SFlipControl", C.c_float),
        ("LFlipControl", C.c_float), ("mpmSplitNum", C.c_int), ("collisionStiff", C.c_float),
        ("surfaceTensionK", C.c_float), ("surfaceTensionFluidC", C.c_float),
        ("fBlending", C.c_float),
    ]


assert C.sizeof(ParamCarrier) == 464


class SolverOptions(C.Structure):
    _fields_ = [
        ("omega", C.c_float), ("warm_start", C.c_float), ("full_support", C.c_int),
        ("boundary_pressure", C.c_int), ("tol", C.c_float), ("min_iters", C.c_int),
        ("max_iters", C.c_int), ("clamp_to_grid", C.c_int), ("wcsph_kernel", C.c_int),
        ("eos_B", C.c_float), ("eos_gamma", C.c_float), ("art_visc_alpha", C.c_float),
        ("sound_speed", C.c_float), ("drift_relax", C.c_float), ("drift_diffusion", C.c_float),
        ("drift_pressure_scale", C.c_float), ("drift_diffusion_scale", C.c_float),
        ("boussinesq", C.c_int), ("rho_ref", C.c_float), ("clamp_pressure", C.c_int),
        ("reserved", C.c_int * 7),
    ]


class SceneInfo(C.Structure):
    _fields_ = [("n_fluid", C.c_int), ("n_boundary", C.c_int), ("n_total", C.c_int),
                ("n_phase", C.c_int), ("spacing", C.c_float)]


class SceneXml(C.Structure):
    _fields_ = [("maxPointNum", C.c_int), ("scaleP", C.c_float), ("fluidnum", C.c_int),
                ("fvmin", Float3 * 2), ("fvmax", Float3 * 2), ("massratio", C.c_float * 3),
                ("densratio", C.c_float * 3), ("viscratio", C.c_float * 3)]


class SphxError(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    """Load libsphx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SphxError(f"{LIB_PATH} not built; run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        P, VP, I, D = C.POINTER, C.c_void_p, C.c_int, C.c_double
        L.sphx_last_error.restype = C.c_char_p
        L.sphx_version.restype = C.c_char_p
        L.sphx_device_count.restype = I
        L.sphx_create.argtypes = [P(VP), P(ParamCarrier)]
        L.sphx_destroy.argtypes = [VP]
        L.sphx_set_params.argtypes = [VP, P(ParamCarrier)]
        L.sphx_get_params.argtypes = [VP, P(ParamCarrier)]
        L.sphx_default_options.argtypes = [P(SolverOptions)]
        L.sphx_default_options.restype = None
        L.sphx_set_options.argtypes = [VP, P(SolverOptions)]
        L.sphx_sync.argtypes = [VP]
        L.sphx_upload.argtypes = [VP, VP, VP, I, I, I]
        L.sphx_download.argtypes = [VP, VP, VP, I, I]
        L.sphx_num_particles.argtypes = [VP]
        L.sphx_download_grid.argtypes = [VP, VP, VP, VP, VP, VP, VP]
        L.sphx_nbr_counts.argtypes = [VP, VP, VP]
        for name in ["sphx_grid_hash", "sphx_grid_sort", "sphx_grid_reorder", "sphx_grid_build",
                     "sphx_boundary_volume", "sphx_wcsph_density_pressure", "sphx_wcsph_force",
                     "sphx_wcsph_advance", "sphx_wcsph_step", "sphx_iisph_density",
                     "sphx_iisph_predict_advection", "sphx_iisph_integrate"]:
            getattr(L, name).argtypes = [VP]
        L.sphx_iisph_pressure_iteration.argtypes = [VP, P(D), P(D)]
        L.sphx_iisph_pressure_solve.argtypes = [VP, P(I), P(D)]
        L.sphx_iisph_step.argtypes = [VP, P(I), P(D)]
        L.sphx_download_iisph.argtypes = [VP] * 8
        L.sphx_upload_iisph.argtypes = [VP] * 6
        L.sphx_timing_enable.argtypes = [VP, I]
        L.sphx_timing_reset.argtypes = [VP]
        L.sphx_timing_only.argtypes = [VP, C.c_char_p]
        L.sphx_timing_get.argtypes = [VP, C.c_char_p, P(D), P(I)]
        L.sphx_setup_grid.argtypes = [P(ParamCarrier)]
        L.sphx_setup_grid.restype = None
        L.sphx_setup_kernels.argtypes = [P(ParamCarrier)]
        L.sphx_setup_kernels.restype = None
        L.sphx_scene_dambreak.argtypes = [I, I, I, C.c_float, C.c_float, C.c_float, C.c_float,
                                          C.c_float, I, C.c_float, C.c_uint64, I,
                                          P(ParamCarrier), VP, VP, P(SceneInfo)]
        for name in ["sphx_iisph_predict_dii", "sphx_iisph_predict_aii", "sphx_iisph_density_dii",
                     "sphx_mf_drift_velocity",
                     "sphx_mf_alpha_advance", "sphx_mf_correction", "sphx_mf_force"]:
            getattr(L, name).argtypes = [VP]
        L.sphx_mf_setup.argtypes = [VP, I, VP, VP]
        L.sphx_mf_upload.argtypes = [VP, VP, I, I]
        L.sphx_mf_download.argtypes = [VP, VP, VP, VP, VP]
        L.sphx_mf_step.argtypes = [VP, P(I), P(D)]
        L.sphx_scene_bubbles.argtypes = [I, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                         I, C.c_float, C.c_float, I, C.c_uint64, P(ParamCarrier),
                                         VP, VP, VP, P(SceneInfo)]
        I64 = C.c_int64
        L.sphx_mlp_nparam.restype = I
        L.sphx_mlp_init.argtypes = [VP, C.c_uint64, C.c_float, C.c_float]
        L.sphx_mlp_set_weights.argtypes = [VP, VP, I]
        L.sphx_mlp_get_weights.argtypes = [VP, VP, I]
        L.sphx_mlp_mode.argtypes = [VP, I, I, I]
        L.sphx_mlp_set_drift_blend.argtypes = [VP, C.c_float]
        L.sphx_mlp_reserve.argtypes = [VP, I64]
        L.sphx_mlp_clear_samples.argtypes = [VP]
        L.sphx_mlp_samples.argtypes = [VP, P(I64), P(I64)]
        L.sphx_mlp_download_samples.argtypes = [VP, I64, I64, VP, VP, VP]
        L.sphx_mlp_features.argtypes = [VP, VP, VP, VP]
        L.sphx_mlp_online.argtypes = [VP, I, I, I, C.c_float, VP]
        L.sphx_mlp_online_async.argtypes = [VP, I]
        L.sphx_mlp_download_features.argtypes = [VP, VP, VP]
        L.sphx_io_last_error.restype = C.c_char_p
        L.sphx_scene_xml_parse.argtypes = [C.c_char_p, I, VP, VP]
        L.sphx_frame_write.argtypes = [C.c_char_p, VP, VP, I]
        L.sphx_frame_read.argtypes = [C.c_char_p, VP, VP, I, VP]
        L.sphx_checkpoint_write.argtypes = [C.c_char_p, VP, VP, VP, I]
        L.sphx_checkpoint_read.argtypes = [C.c_char_p, VP, VP, VP, I, VP]
        L.sphx_checkpoint_save.argtypes = [VP, C.c_char_p]
        L.sphx_boundary_parse.argtypes = [C.c_char_p, VP, VP, VP, I, I, VP]
        L.sphx_checkpoint_restore.argtypes = [VP, C.c_char_p]
        L.sphx_mlp_record_stride.argtypes = [VP, I]
        L.sphx_mlp_eval.argtypes = [VP, VP, I, VP]
        L.sphx_mlp_grad.argtypes = [VP, VP, VP, VP, I, VP, VP, P(D)]
        L.sphx_mlp_train.argtypes = [VP, I, I, C.c_float, VP, C.c_uint64, VP]
        L.sphx_mlp_stats.argtypes = [VP, P(I64), P(I64), P(I)]
        L.sphx_mpm_setup.argtypes = [VP, I, I]
        L.sphx_mpm_p2g.argtypes = [VP, I]
        L.sphx_mpm_g2p.argtypes = [VP, I]
        for name in ["sphx_mpm_stress", "sphx_mpm_node_update", "sphx_mpm_substep"]:
            getattr(L, name).argtypes = [VP]
        L.sphx_mpm_download_nodes.argtypes = [VP, VP, VP]
        L.sphx_mpm_upload_nodes.argtypes = [VP, VP, VP]
        L.sphx_mpm_active_blocks.argtypes = [VP, P(I)]
        L.sphx_setup_mpm_grid.argtypes = [P(ParamCarrier), C.c_float]
        L.sphx_setup_mpm_grid.restype = None
        L.sphx_set_drucker_prager.argtypes = [P(ParamCarrier)]
        L.sphx_set_drucker_prager.restype = None
        L.sphx_scene_sand.argtypes = [I, C.c_float, C.c_float, C.c_float, I, C.c_float, C.c_float,
                                      C.c_float, C.c_float, C.c_float, C.c_uint64,
                                      P(ParamCarrier), VP, VP, P(SceneInfo)]
        L.sphx_rpy_create.argtypes = [P(VP), I, I]
        L.sphx_rpy_destroy.argtypes = [VP]
        L.sphx_rpy_set_spheres.argtypes = [VP, VP, VP, C.c_float]
        L.sphx_rpy_set_split.argtypes = [VP, I]
        L.sphx_rpy_apply.argtypes = [VP, VP, VP, I]
        L.sphx_rpy_apply_device.argtypes = [VP, VP, VP, I]
        L.sphx_rpy_sync.argtypes = [VP]
        L.sphx_rpy_timing.argtypes = [VP, I]
        L.sphx_rpy_timing_get.argtypes = [VP, P(D), P(I)]
        L.sphx_rpy_walkthrough.argtypes = [VP, VP, I, I]
        L.sphx_mob_gemv.argtypes = [VP, VP, VP, I]
        L.sphx_dev_alloc.argtypes = [P(VP), C.c_size_t]
        L.sphx_dev_free.argtypes = [VP]
        L.sphx_dev_copy.argtypes = [VP, VP, C.c_size_t, I]
        L.sphx_scene_ensemble.argtypes = [I, VP, C.c_float, C.c_float, C.c_float, C.c_float,
                                          C.c_float, I, C.c_uint64, P(ParamCarrier), VP, VP, VP,
                                          VP, P(SceneInfo)]
        L.sphx_ensemble_setup.argtypes = [VP, I]
        L.sphx_scene_stats.argtypes = [VP, VP, VP]
        L.sphx_graph_build.argtypes = [VP, I]
        L.sphx_graph_step.argtypes = [VP, P(I), P(D)]
        L.sphx_graph_release.argtypes = [VP]
        L.sphx_tune_set.argtypes = [VP, C.c_char_p, I]
        L.sphx_tune_get.argtypes = [VP, C.c_char_p, P(I)]
        L.sphx_autotune.argtypes = [VP, I]
        _lib = L
    return _lib


def last_error() -> str:
    return lib().sphx_last_error().decode()


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        raise SphxError(f"{what}: {last_error()} (rc={rc})")


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_options() -> SolverOptions:
    o = SolverOptions()
    lib().sphx_default_options(C.byref(o))
    return o


# ------------------------------------------------------------------ scenes --
def scene_dambreak(nx=32, ny=32, nz=32, s=0.01, L=(0.64, 0.32, 0.32), h=0.02, layers=2,
                   jitter=1e-4, seed=20261202, kernel_family=1):
    """Seeded dam-break scene; returns (params, disp, calc, info)."""
    L_ = lib()
    info = SceneInfo()
    p = ParamCarrier()
    L_.sphx_scene_dambreak(nx, ny, nz, s, L[0], L[1], L[2], h, layers, jitter, seed,
                           kernel_family, C.byref(p), None, None, C.byref(info))
    disp = np.zeros(info.n_total, display_dtype)
    calc = np.zeros(info.n_total, calc_dtype)
    L_.sphx_scene_dambreak(nx, ny, nz, s, L[0], L[1], L[2], h, layers, jitter, seed,
                           kernel_family, C.byref(p), ptr(disp), ptr(calc), C.byref(info))
    return p, disp, calc, info


def scene_bubbles(n_fluid=1 << 20, L=(1.0, 1.0, 1.0), fill=0.8, h=0.01, n_bubbles=512,
                  rmin=0.01, rmax=0.04, layers=2, seed=20261203):
    """Two-phase bubble tank; returns (params, disp, calc, alpha[n][2], info)."""
    L_ = lib()
    info = SceneInfo()
    p = ParamCarrier()
    args = (n_fluid, L[0], L[1], L[2], fill, h, n_bubbles, rmin, rmax, layers, seed)
    rc = L_.sphx_scene_bubbles(*args, C.byref(p), None, None, None, C.byref(info))
    if rc < 0:
        raise SphxError("sphx_scene_bubbles: bad arguments")
    disp = np.zeros(info.n_total, display_dtype)
    calc = np.zeros(info.n_total, calc_dtype)
    alpha = np.zeros((info.n_total, 2), np.float32)
    L_.sphx_scene_bubbles(*args, C.byref(p), ptr(disp), ptr(calc), ptr(alpha), C.byref(info))
    return p, disp, calc, alpha, info


# ----------------------------------------------------------------- context --
def scene_ensemble(n_fluid, fill=0.8, h=0.01, bubbles_per_million=512.0, rmin=0.01, rmax=0.04,
                   layers=2, seed=20261207):
    """Batched ensemble of two-phase tanks (sphx_scene_ensemble); returns
    params, disp, calc, alpha [n][2], scene_first [nscenes + 1], info."""
    L = lib()
    nf = np.ascontiguousarray(n_fluid, np.int32)
    S = len(nf)
    p = ParamCarrier()
    info = SceneInfo()
    first = np.zeros(S + 1, np.int32)
    n = L.sphx_scene_ensemble(S, ptr(nf), fill, h, bubbles_per_million, rmin, rmax, layers, seed,
                              C.byref(p), None, None, None, ptr(first), C.byref(info))
    if n < 0:
        raise SphxError("sphx_scene_ensemble: bad arguments")
    disp = np.zeros(n, display_dtype)
    calc = np.zeros(n, calc_dtype)
    alpha = np.zeros((n, 2), np.float32)
    L.sphx_scene_ensemble(S, ptr(nf), fill, h, bubbles_per_million, rmin, rmax, layers, seed,
                          C.byref(p), ptr(disp), ptr(calc), ptr(alpha), ptr(first), C.byref(info))
    return p, disp, calc, alpha, first, info


def scene_sand(n=4096, lo=0.3, hi=0.6, dx=1.0 / 32, res=32, E=3.5e5, nu=0.3, phi=30.0,
               rho=1600.0, f_noise=1e-3, seed=20261205):
    """Granular block over a res^3 MPM node grid (sphx_scene_sand)."""
    L = lib()
    p = ParamCarrier()
    info = SceneInfo()
    disp = np.zeros(n, display_dtype)
    calc = np.zeros(n, calc_dtype)
    rc = L.sphx_scene_sand(n, lo, hi, dx, res, E, nu, phi, rho, f_noise, seed, C.byref(p),
                           ptr(disp), ptr(calc), C.byref(info))
    if rc < 0:
        raise SphxError("sphx_scene_sand: bad arguments")
    return p, disp, calc, info


# ----------------------------------------------------------- state I/O --
def _io_check(rc, what):
    if rc != 0:
        raise SphxError(f"{what}: {lib().sphx_io_last_error().decode()}")


def scene_xml_parse(text: str, caseid: int):
    """(ParamCarrier, SceneXml) from scene.xml text (sphx_scene_xml_parse)."""
    L = lib()
    p = ParamCarrier()
    sx = SceneXml()
    _io_check(L.sphx_scene_xml_parse(text.encode(), caseid, C.byref(p), C.byref(sx)),
              "sphx_scene_xml_parse")
    return p, sx


def frame_write(path, disp, calc):
    _io_check(lib().sphx_frame_write(str(path).encode(), ptr(disp), ptr(calc), len(disp)),
              "sphx_frame_write")


def frame_read(path):
    """(disp, calc) with pos, type, mass and restdens of a text frame."""
    L = lib()
    n = C.c_int()
    _io_check(L.sphx_frame_read(str(path).encode(), None, None, 0, C.byref(n)), "sphx_frame_read")
    disp = np.zeros(n.value, display_dtype)
    calc = np.zeros(n.value, calc_dtype)
    _io_check(L.sphx_frame_read(str(path).encode(), ptr(disp), ptr(calc), n.value, C.byref(n)),
              "sphx_frame_read")
    return disp, calc


def checkpoint_write(path, params, disp, calc):
    _io_check(lib().sphx_checkpoint_write(str(path).encode(), C.byref(params), ptr(disp),
                                          ptr(calc), len(disp)), "sphx_checkpoint_write")


def checkpoint_read(path):
    """(params, disp, calc) of a checkpoint file."""
    L = lib()
    p = ParamCarrier()
    n = C.c_int()
    _io_check(L.sphx_checkpoint_read(str(path).encode(), C.byref(p), None, None, 0, C.byref(n)),
              "sphx_checkpoint_read")
    disp = np.zeros(n.value, display_dtype)
    calc = np.zeros(n.value, calc_dtype)
    _io_check(L.sphx_checkpoint_read(str(path).encode(), C.byref(p), ptr(disp), ptr(calc),
                                     n.value, C.byref(n)), "sphx_checkpoint_read")
    return p, disp, calc


def boundary_parse(text: str, params, start=0):
    """Boundary particles of a shape file's text (WALL / OPEN_BOX commands);
    returns (disp, calc) of the particles, bornid counted from `start`."""
    L = lib()
    n = C.c_int()
    _io_check(L.sphx_boundary_parse(text.encode(), C.byref(params), None, None, start, start,
                                    C.byref(n)), "sphx_boundary_parse")
    disp = np.zeros(start + n.value, display_dtype)
    calc = np.zeros(start + n.value, calc_dtype)
    _io_check(L.sphx_boundary_parse(text.encode(), C.byref(params), ptr(disp), ptr(calc), start,
                                    start + n.value, C.byref(n)), "sphx_boundary_parse")
    return disp[start:], calc[start:]


class Context:
    """Owns one sphx_context (device buffers for params.maxNum particles)."""

    def __init__(self, params: ParamCarrier, options: SolverOptions | None = None):
        self.L = lib()
        self.h = C.c_void_p()
        check(self.L.sphx_create(C.byref(self.h), C.byref(params)), "sphx_create")
        self.params = params
        if options is not None:
            self.set_options(options)

    def close(self):
        if self.h:
            self.L.sphx_destroy(self.h)
            self.h = C.c_void_p()

    def checkpoint_save(self, path):
        _io_check(self.L.sphx_checkpoint_save(self.h, str(path).encode()), "sphx_checkpoint_save")

    def checkpoint_restore(self, path):
        _io_check(self.L.sphx_checkpoint_restore(self.h, str(path).encode()),
                  "sphx_checkpoint_restore")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_options(self, o: SolverOptions):
        check(self.L.sphx_set_options(self.h, C.byref(o)), "sphx_set_options")

    def set_params(self, p: ParamCarrier):
        check(self.L.sphx_set_params(self.h, C.byref(p)), "sphx_set_params")
        self.params = p

    def upload(self, disp, calc, start=0, grow=1):
        disp = np.ascontiguousarray(disp, display_dtype)
        calc = np.ascontiguousarray(calc, calc_dtype)
        check(self.L.sphx_upload(self.h, ptr(disp), ptr(calc), start, len(disp), grow),
              "sphx_upload")

    @property
    def num(self) -> int:
        return self.L.sphx_num_particles(self.h)

    def download(self):
        n = self.num
        disp = np.zeros(n, display_dtype)
        calc = np.zeros(n, calc_dtype)
        check(self.L.sphx_download(self.h, ptr(disp), ptr(calc), 0, n), "sphx_download")
        return disp, calc

    def download_grid(self):
        n, gt = self.num, self.params.gridtotal
        out = dict(mgcell=np.zeros(n, np.uint32), midsort=np.zeros(n, np.uint32),
                   mgndx=np.zeros(n, np.uint32), gridoff=np.zeros(gt, np.int32),
                   gridcnt=np.zeros(gt, np.int32), MFidTable=np.zeros(n, np.int32))
        check(self.L.sphx_download_grid(self.h, *[ptr(out[k]) for k in
                                                  ["mgcell", "midsort", "mgndx", "gridoff",
                                                   "gridcnt", "MFidTable"]]),
              "sphx_download_grid")
        return out

    def nbr_counts(self):
        """Per-particle neighbour counts of the step's list (-1: overflowed to
        the cell scan) and the list capacity."""
        out = np.zeros(self.num, np.int32)
        cap = C.c_int()
        check(self.L.sphx_nbr_counts(self.h, ptr(out), C.byref(cap)), "sphx_nbr_counts")
        return out, cap.value

    def download_iisph(self):
        n = self.num
        out = dict(dii=np.zeros((n, 3), np.float32), dijpj=np.zeros((n, 3), np.float32),
                   aii=np.zeros(n, np.float32), rho_adv=np.zeros(n, np.float32),
                   press_l=np.zeros(n, np.float32), press_l1=np.zeros(n, np.float32),
                   residue=np.zeros(n, np.float32))
        check(self.L.sphx_download_iisph(self.h, *[ptr(out[k]) for k in
                                                   ["dii", "dijpj", "aii", "rho_adv", "press_l",
                                                    "press_l1", "residue"]]),
              "sphx_download_iisph")
        return out

    def call(self, name: str):
        check(getattr(self.L, "sphx_" + name)(self.h), name)

    def sync(self):
        check(self.L.sphx_sync(self.h), "sphx_sync")

    def pressure_iteration(self):
        ms, ma = C.c_double(), C.c_double()
        check(self.L.sphx_iisph_pressure_iteration(self.h, C.byref(ms), C.byref(ma)),
              "pressure_iteration")
        return ms.value, ma.value

    def pressure_solve(self):
        it, err = C.c_int(), C.c_double()
        check(self.L.sphx_iisph_pressure_solve(self.h, C.byref(it), C.byref(err)),
              "pressure_solve")
        return it.value, err.value

    def iisph_step(self):
        it, err = C.c_int(), C.c_double()
        check(self.L.sphx_iisph_step(self.h, C.byref(it), C.byref(err)), "iisph_step")
        return it.value, err.value

    # -- multiphase --
    def mf_setup(self, rho, nu):
        rho = np.ascontiguousarray(rho, np.float32)
        nu = np.ascontiguousarray(nu, np.float32)
        self.nphase = len(rho)
        check(self.L.sphx_mf_setup(self.h, len(rho), ptr(rho), ptr(nu)), "sphx_mf_setup")

    def mf_upload(self, alpha, start=0):
        alpha = np.ascontiguousarray(alpha, np.float32)
        check(self.L.sphx_mf_upload(self.h, ptr(alpha), start, len(alpha)), "sphx_mf_upload")

    def mf_download(self):
        n, k = self.num, self.nphase
        out = dict(alpha=np.zeros((n, k), np.float32), rho_m=np.zeros(n, np.float32),
                   drift=np.zeros((n, k, 3), np.float32), mforce=np.zeros((n, 3), np.float32))
        check(self.L.sphx_mf_download(self.h, ptr(out["alpha"]), ptr(out["rho_m"]),
                                      ptr(out["drift"]), ptr(out["mforce"])), "sphx_mf_download")
        return out

    def mf_step(self):
        it, err = C.c_int(), C.c_double()
        check(self.L.sphx_mf_step(self.h, C.byref(it), C.byref(err)), "mf_step")
        return it.value, err.value

    # -- learned warm start / drift closure --
    MLP_NF, MLP_NO = 24, 4

    def mlp_init(self, seed=1, p_scale=1.0, u_scale=1.0):
        check(self.L.sphx_mlp_init(self.h, seed, p_scale, u_scale), "sphx_mlp_init")

    def mlp_weights(self):
        w = np.zeros(self.L.sphx_mlp_nparam(), np.float32)
        check(self.L.sphx_mlp_get_weights(self.h, ptr(w), len(w)), "sphx_mlp_get_weights")
        return w

    def mlp_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        check(self.L.sphx_mlp_set_weights(self.h, ptr(w), len(w)), "sphx_mlp_set_weights")

    def mlp_mode(self, warm=False, drift=False, record=False):
        check(self.L.sphx_mlp_mode(self.h, int(warm), int(drift), int(record)), "sphx_mlp_mode")

    def mlp_set_drift_blend(self, b):
        check(self.L.sphx_mlp_set_drift_blend(self.h, b), "sphx_mlp_set_drift_blend")

    def mlp_reserve(self, n):
        check(self.L.sphx_mlp_reserve(self.h, int(n)), "sphx_mlp_reserve")

    def mlp_clear_samples(self):
        check(self.L.sphx_mlp_clear_samples(self.h), "sphx_mlp_clear_samples")

    def mlp_samples(self):
        t, v = C.c_int64(), C.c_int64()
        check(self.L.sphx_mlp_samples(self.h, C.byref(t), C.byref(v)), "sphx_mlp_samples")
        return t.value, v.value

    def mlp_download_samples(self, start, count):
        X = np.zeros((count, self.MLP_NF), np.float32)
        T = np.zeros((count, self.MLP_NO), np.float32)
        M = np.zeros(count, np.float32)
        check(self.L.sphx_mlp_download_samples(self.h, start, count, ptr(X), ptr(T), ptr(M)),
              "sphx_mlp_download_samples")
        return X, T, M

    def mlp_download_features(self):
        """(X [n][24], p(t-1) [n]) as the last pass left them."""
        X = np.zeros((self.num, self.MLP_NF), np.float32)
        pp = np.zeros(self.num, np.float32)
        check(self.L.sphx_mlp_download_features(self.h, ptr(X), ptr(pp)),
              "sphx_mlp_download_features")
        return X, pp

    def mlp_features(self):
        X = np.zeros((self.num, self.MLP_NF), np.float32)
        pp = np.zeros(self.num, np.float32)
        dp = np.zeros(self.num, np.float32)
        check(self.L.sphx_mlp_features(self.h, ptr(X), ptr(pp), ptr(dp)), "sphx_mlp_features")
        return X, pp, dp

    def mlp_online(self, every, iters=0, batch=0, lr=0.0, lw=(1.0, 1.0, 1.0, 1.0)):
        lw = np.ascontiguousarray(lw, np.float32)
        check(self.L.sphx_mlp_online(self.h, every, iters, batch, lr, ptr(lw)), "sphx_mlp_online")

    def mlp_online_async(self, on=True):
        check(self.L.sphx_mlp_online_async(self.h, int(bool(on))), "sphx_mlp_online_async")

    def mlp_record_stride(self, stride):
        check(self.L.sphx_mlp_record_stride(self.h, int(stride)), "sphx_mlp_record_stride")

    def mlp_eval(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((len(X), self.MLP_NO), np.float32)
        check(self.L.sphx_mlp_eval(self.h, ptr(X), len(X), ptr(Y)), "sphx_mlp_eval")
        return Y

    def mlp_grad(self, X, T, lw, M=None):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        lw = np.ascontiguousarray(lw, np.float32)
        M = None if M is None else np.ascontiguousarray(M, np.float32)
        g = np.zeros(self.L.sphx_mlp_nparam(), np.float32)
        loss = C.c_double()
        check(self.L.sphx_mlp_grad(self.h, ptr(X), ptr(T), ptr(M), len(X), ptr(lw), ptr(g),
                                   C.byref(loss)), "sphx_mlp_grad")
        return loss.value, g

    def mlp_train(self, epochs, batch, lr, lw=(1.0, 0.1, 0.1, 0.1), seed=1):
        lw = np.ascontiguousarray(lw, np.float32)
        out = np.zeros(max(epochs, 1), np.float64)
        check(self.L.sphx_mlp_train(self.h, epochs, batch, lr, ptr(lw), seed, ptr(out)),
              "sphx_mlp_train")
        return out[:epochs]

    def mlp_guard(self, factor=10.0):
        """Probe the plain start when the learned start's first iteration
        leaves an error above factor * tol (0: guard off)."""
        self.L.sphx_mlp_guard.argtypes = [C.c_void_p, C.c_float]
        check(self.L.sphx_mlp_guard(self.h, float(factor)), "sphx_mlp_guard")

    def mlp_guard_stats(self):
        a, b = C.c_int64(), C.c_int64()
        check(self.L.sphx_mlp_guard_stats(self.h, C.byref(a), C.byref(b)), "sphx_mlp_guard_stats")
        return {"probes": a.value, "plain_picks": b.value}

    def mlp_stats(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int()
        check(self.L.sphx_mlp_stats(self.h, C.byref(a), C.byref(b), C.byref(c)), "sphx_mlp_stats")
        return dict(learned=a.value, fallbacks=b.value, last_fallback=c.value)

    # -- batched ensemble / step graphs --
    def ensemble_setup(self, nscenes):
        check(self.L.sphx_ensemble_setup(self.h, nscenes), "sphx_ensemble_setup")
        self.nscenes = nscenes

    def scene_stats(self):
        it = np.zeros(self.nscenes, np.int32)
        err = np.zeros(self.nscenes, np.float64)
        check(self.L.sphx_scene_stats(self.h, ptr(it), ptr(err)), "sphx_scene_stats")
        return it, err

    def graph_build(self, budget):
        check(self.L.sphx_graph_build(self.h, budget), "sphx_graph_build")

    def tune_cache_save(self, path):
        check(self.L.sphx_tune_cache_save(self.h, str(path).encode()), "sphx_tune_cache_save")

    def tune_cache_load(self, path):
        n = C.c_int()
        check(self.L.sphx_tune_cache_load(self.h, str(path).encode(), C.byref(n)),
              "sphx_tune_cache_load")
        return n.value

    def persistent_solve(self, on=True):
        check(self.L.sphx_persistent_solve(self.h, 1 if on else 0), "sphx_persistent_solve")

    def solve_history(self, max_solves=4096):
        it = np.zeros(max_solves, np.int32)
        err = np.zeros(max_solves, np.float64)
        n = C.c_int()
        check(self.L.sphx_solve_history(self.h, ptr(it), ptr(err), max_solves, C.byref(n)),
              "sphx_solve_history")
        return it[:n.value], err[:n.value]

    def graph_step_async(self):
        """One captured step with the persistent solve, no read-back."""
        check(self.L.sphx_graph_step(self.h, None, None), "sphx_graph_step")

    def graph_step(self):
        it, err = C.c_int(), C.c_double()
        check(self.L.sphx_graph_step(self.h, C.byref(it), C.byref(err)), "sphx_graph_step")
        return it.value, err.value

    TUNE_KERNELS = ("iisph_density", "iisph_dii", "iisph_aii", "iisph_dp", "iisph_iter",
                    "mf_drift", "mf_alpha", "mf_force", "mlp_features", "mlp_infer",
                    "iisph_density_dii")

    def tune_set(self, kernel, block):
        check(self.L.sphx_tune_set(self.h, kernel.encode(), block), "sphx_tune_set")

    def tuning(self):
        out = {}
        for k in self.TUNE_KERNELS:
            v = C.c_int()
            check(self.L.sphx_tune_get(self.h, k.encode(), C.byref(v)), "sphx_tune_get")
            out[k] = v.value
        return out

    def autotune(self, steps=2):
        check(self.L.sphx_autotune(self.h, steps), "sphx_autotune")
        return self.tuning()

    # -- MPM --
    def mpm_setup(self, kernel=0, physical=0):
        check(self.L.sphx_mpm_setup(self.h, kernel, physical), "sphx_mpm_setup")
        pr = self.params
        self.mpm_res = (pr.mpmRes.x, pr.mpmRes.y, pr.mpmRes.z)

    def mpm_p2g(self, apic=1):
        check(self.L.sphx_mpm_p2g(self.h, apic), "sphx_mpm_p2g")

    def mpm_g2p(self, apic=1):
        check(self.L.sphx_mpm_g2p(self.h, apic), "sphx_mpm_g2p")

    def mpm_nodes(self):
        N = int(np.prod(self.mpm_res))
        v = np.zeros((N, 3), np.float32)
        m = np.zeros(N, np.float32)
        check(self.L.sphx_mpm_download_nodes(self.h, ptr(v), ptr(m)), "sphx_mpm_download_nodes")
        return v, m

    def mpm_set_nodes(self, v, m):
        v = np.ascontiguousarray(v, np.float32)
        m = np.ascontiguousarray(m, np.float32)
        check(self.L.sphx_mpm_upload_nodes(self.h, ptr(v), ptr(m)), "sphx_mpm_upload_nodes")

    # -- FP64 conservation totals --
    def conservation_totals(self):
        """[scenes][8]: sum m, sum m v (3), sum m |v|^2 / 2, sum m x (3)."""
        S = max(1, getattr(self, "nscenes", 0))
        out = np.zeros((S, 8), np.float64)
        check(self.L.sphx_conservation_totals(self.h, ptr(out)), "sphx_conservation_totals")
        return out

    def conservation_track(self, on=True):
        check(self.L.sphx_conservation_track(self.h, 1 if on else 0), "sphx_conservation_track")

    def conservation_history(self, max_steps=4096):
        S = max(1, getattr(self, "nscenes", 0))
        out = np.zeros((max_steps, S, 8), np.float64)
        n = C.c_int()
        check(self.L.sphx_conservation_history(self.h, ptr(out), max_steps, C.byref(n)),
              "sphx_conservation_history")
        return out[:n.value]

    def mpm_active_blocks(self):
        n = C.c_int()
        check(self.L.sphx_mpm_active_blocks(self.h, C.byref(n)), "sphx_mpm_active_blocks")
        return n.value

    def timing(self, on=True):
        check(self.L.sphx_timing_enable(self.h, 1 if on else 0), "timing")

    def timing_reset(self):
        check(self.L.sphx_timing_reset(self.h), "timing_reset")

    def timing_only(self, name=None):
        """Time only the named kernel (None: every kernel)."""
        check(self.L.sphx_timing_only(self.h, name.encode() if name else None), "timing_only")

    def timing_get(self, name: str):
        t, n = C.c_double(), C.c_int()
        check(self.L.sphx_timing_get(self.h, name.encode(), C.byref(t), C.byref(n)), "timing_get")
        return t.value, n.value


def device_count() -> int:
    return lib().sphx_device_count()


class Mobility:
    """Polydisperse RPY mobility applied on MFMA (sphx_rpy_*)."""

    def __init__(self, n: int, nrhs: int = 64):
        self.L = lib()
        self.h = C.c_void_p()
        check(self.L.sphx_rpy_create(C.byref(self.h), n, nrhs), "sphx_rpy_create")
        self.n, self.nrhs = n, nrhs

    def set_spheres(self, pos, radius, mu=1.0):
        pos = np.ascontiguousarray(pos, np.float32)
        radius = np.ascontiguousarray(radius, np.float32)
        check(self.L.sphx_rpy_set_spheres(self.h, ptr(pos), ptr(radius), mu), "set_spheres")

    def set_split(self, split):
        check(self.L.sphx_rpy_set_split(self.h, split), "set_split")

    def apply(self, F, mode=6):
        F = np.ascontiguousarray(F, np.float32)
        assert F.shape == (3 * self.n, self.nrhs)
        U = np.zeros_like(F)
        check(self.L.sphx_rpy_apply(self.h, ptr(F), ptr(U), mode), "sphx_rpy_apply")
        return U

    def apply_device(self, dF: int, dU: int, mode=6):
        check(self.L.sphx_rpy_apply_device(self.h, C.c_void_p(dF), C.c_void_p(dU), mode),
              "sphx_rpy_apply_device")

    def sync(self):
        check(self.L.sphx_rpy_sync(self.h), "sphx_rpy_sync")

    def timing(self, on=True):
        check(self.L.sphx_rpy_timing(self.h, int(on)), "sphx_rpy_timing")

    def timing_get(self):
        """(total ms, launches) of the tile kernel since timing(True)."""
        ms, n = C.c_double(), C.c_int()
        check(self.L.sphx_rpy_timing_get(self.h, C.byref(ms), C.byref(n)), "sphx_rpy_timing_get")
        return ms.value, n.value

    def close(self):
        if self.h:
            self.L.sphx_rpy_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class StokesianBufList(C.Structure):
    _fields_ = [("cuUnew", C.c_void_p), ("dispBuffer", C.c_void_p), ("cuOmega", C.c_void_p),
                ("cuForce", C.c_void_p), ("dt", C.c_float), ("simscale", C.c_float)]


class DeviceArray:
    """A device copy of a numpy array (sphx_dev_alloc); .ptr is the device
    address, .get() copies it back."""

    def __init__(self, a):
        self.L = lib()
        self.a = np.ascontiguousarray(a)
        self.ptr = C.c_void_p()
        check(self.L.sphx_dev_alloc(C.byref(self.ptr), max(self.a.nbytes, 1)), "sphx_dev_alloc")
        if self.a.nbytes:
            check(self.L.sphx_dev_copy(self.ptr, ptr(self.a), self.a.nbytes, 1), "sphx_dev_copy")

    def get(self):
        out = np.empty_like(self.a)
        if out.nbytes:
            check(self.L.sphx_dev_copy(ptr(out), self.ptr, out.nbytes, 2), "sphx_dev_copy")
        return out

    def free(self):
        if self.ptr:
            self.L.sphx_dev_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
