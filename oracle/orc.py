"""Python binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, as the checker or as the CPU baseline.  It converts the
reference's AoS particle arrays into the oracle's typed SoA state, runs the
restated reference kernels in FP32 (faithful) or FP64 (twin), and converts
back.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
# SPHX_ORACLE_LIB selects another build of the same sources (the sanitizer
# build of tests/san)
LIB_PATH = Path(os.environ.get("SPHX_ORACLE_LIB", str(_HERE / "liboracle.so")))
sys.path.insert(0, str(_HERE.parent / "sph-multiphase_amd"))
import sphx  # noqa: E402  (types only)

UNDEF = 0xFFFFFFFF


class OrcState(C.Structure):
    _fields_ = [("n", C.c_int), ("gridtotal", C.c_int)] + [
        (name, C.c_void_p) for name in
        ["pos", "type", "vel", "veleval", "accel", "pressure", "dens", "restdens", "mass",
         "visc", "bornid", "mgcell", "midsort", "mgndx", "gridoff", "gridcnt", "MFidTable",
         "dii", "dijpj", "aii", "rho_adv", "press_l", "press_l1", "residue"]]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        _lib = C.CDLL(str(LIB_PATH))
        _lib.orc_f32_ii_solve.restype = C.c_int
        _lib.orc_f64_ii_solve.restype = C.c_int
    return _lib


VEC = ["pos", "vel", "veleval", "accel", "dii", "dijpj"]
SCAL = ["pressure", "dens", "restdens", "mass", "visc", "aii", "rho_adv", "press_l",
        "press_l1", "residue"]


class State:
    """Oracle state for n particles in a grid of gridtotal cells."""

    def __init__(self, disp, calc, params: sphx.ParamCarrier, dtype=np.float32):
        n = len(disp)
        self.dtype = np.dtype(dtype)
        self.params = params
        self.n = n
        self.a = {}
        for k in VEC:
            self.a[k] = np.zeros((n, 3), self.dtype)
        for k in SCAL:
            self.a[k] = np.zeros(n, self.dtype)
        self.a["pos"][:] = disp["pos"]
        self.a["type"] = np.ascontiguousarray(disp["type"], np.int32)
        for k in ["vel", "veleval", "accel"]:
            self.a[k][:] = calc[k]
        for k in ["pressure", "dens", "restdens", "mass", "visc"]:
            self.a[k][:] = calc[k]
        self.a["bornid"] = np.ascontiguousarray(calc["bornid"], np.int32)
        gt = params.gridtotal
        self.a["mgcell"] = np.zeros(n, np.uint32)
        self.a["midsort"] = np.zeros(n, np.uint32)
        self.a["mgndx"] = np.zeros(n, np.uint32)
        self.a["gridoff"] = np.zeros(gt, np.int32)
        self.a["gridcnt"] = np.zeros(gt, np.int32)
        self.a["MFidTable"] = np.full(n, -1, np.int32)
        self.s = OrcState()
        self.s.n = n
        self.s.gridtotal = gt
        for name, _ in OrcState._fields_[2:]:
            setattr(self.s, name, self.a[name].ctypes.data)
        self.opts = sphx.default_options()

    @property
    def pfx(self):
        return "orc_f32_" if self.dtype == np.float32 else "orc_f64_"

    def __getitem__(self, k):
        return self.a[k]

    def copy_grid_from(self, other: "State"):
        for k in ["mgcell", "midsort", "mgndx", "gridoff", "gridcnt", "MFidTable"]:
            self.a[k][:] = other.a[k]

    # -- grid build (FP32 positions, integer outputs) --
    def grid_build(self, mfid_mode=0):
        L = lib()
        if self.dtype == np.float32:
            L.orc_initial_sort(C.byref(self.params), C.byref(self.s))
            L.orc_sort_by_key(C.byref(self.s))
            L.orc_cell_ranges(C.byref(self.s))
            L.orc_rearrange(C.byref(self.s), mfid_mode)
            return
        # FP64 twin: the hash is defined on FP32 positions (kern.cu:52), so
        # the index arrays come from an FP32 build and the same gather is
        # applied to the FP64 fields
        disp = np.zeros(self.n, sphx.display_dtype)
        calc = np.zeros(self.n, sphx.calc_dtype)
        disp["pos"] = self.a["pos"].astype(np.float32)
        disp["type"] = self.a["type"]
        calc["bornid"] = self.a["bornid"]
        f = State(disp, calc, self.params, np.float32)
        f.grid_build(mfid_mode)
        self.copy_grid_from(f)
        valid = f.a["mgcell"] != UNDEF
        src = np.where(valid, f.a["midsort"], np.arange(self.n)).astype(np.int64)
        for k in ["pos", "vel", "veleval", "accel", "pressure", "dens", "restdens", "mass",
                  "visc", "type", "bornid"]:
            self.a[k][:] = self.a[k][src]
        self.a["pos"][~valid] = -1000.0

    def hash_only(self):
        lib().orc_initial_sort(C.byref(self.params), C.byref(self.s))

    def _call(self, name, with_opts=True):
        f = getattr(lib(), self.pfx + name)
        if with_opts:
            f(C.byref(self.params), C.byref(self.opts), C.byref(self.s))
        else:
            f(C.byref(self.params), C.byref(self.s))

    def boundary_volume(self): self._call("boundary_volume")
    def density_pressure(self): self._call("density_pressure")
    def force(self): self._call("force")
    def advance(self): self._call("advance", False)
    def ii_density(self): self._call("ii_density", False)
    def ii_dii(self): self._call("ii_dii")
    def ii_aii(self): self._call("ii_aii")
    def ii_dp(self): self._call("ii_dp")
    def ii_iter(self): self._call("ii_iter")
    def ii_integrate(self): self._call("ii_integrate", False)

    # -- SPH elastic solid and surface tension (oracle_solid.cpp) --
    def solid_tensor(self, F, stress):
        """F [n][9] (row-major deformation gradients, this state's dtype) is
        updated in place; stress [n][9] is written for elastic particles."""
        assert F.dtype == self.dtype and stress.dtype == self.dtype
        assert F.flags.c_contiguous and stress.flags.c_contiguous
        getattr(lib(), self.pfx + "solid_tensor")(C.byref(self.params), C.byref(self.s),
                                                  C.c_void_p(F.ctypes.data),
                                                  C.c_void_p(stress.ctypes.data))

    def solid_force(self, stress):
        """accel of elastic particles from the stresses [n][9]."""
        assert stress.dtype == self.dtype and stress.flags.c_contiguous
        getattr(lib(), self.pfx + "solid_force")(C.byref(self.params), C.byref(self.s),
                                                 C.c_void_p(stress.ctypes.data))

    def surface_tension(self):
        getattr(lib(), self.pfx + "surface_tension")(C.byref(self.params), C.byref(self.s))

    def ii_iteration(self):
        ms, ma = C.c_double(), C.c_double()
        getattr(lib(), self.pfx + "ii_iteration")(C.byref(self.params), C.byref(self.opts),
                                                  C.byref(self.s), C.byref(ms), C.byref(ma))
        return ms.value, ma.value

    def ii_solve(self):
        err = C.c_double()
        it = getattr(lib(), self.pfx + "ii_solve")(C.byref(self.params), C.byref(self.opts),
                                                   C.byref(self.s), C.byref(err))
        return it, err.value


class OrcMF(C.Structure):
    _fields_ = [("nphase", C.c_int), ("rho", C.c_float * 3), ("nu", C.c_float * 3),
                ("relax", C.c_float), ("sigma", C.c_float), ("cpress", C.c_float),
                ("cdiff", C.c_float), ("boussinesq", C.c_int), ("rho_ref", C.c_float),
                ("alpha", C.c_void_p), ("alpha_new", C.c_void_p), ("drift", C.c_void_p),
                ("mforce", C.c_void_p), ("amf", C.c_void_p)]


class MF:
    """Multiphase arrays attached to an oracle State (same dtype)."""

    def __init__(self, st: State, rho, nu, opts: sphx.SolverOptions):
        n, dt = st.n, st.dtype
        self.st = st
        self.nphase = len(rho)
        self.a = dict(alpha=np.zeros((n, 4), dt), alpha_new=np.zeros((n, 4), dt),
                      drift=np.zeros((n, 3, 3), dt), mforce=np.zeros((n, 3), dt),
                      amf=np.zeros((n, 3), dt))
        self.m = OrcMF()
        self.m.nphase = len(rho)
        for k in range(len(rho)):
            self.m.rho[k] = rho[k]
            self.m.nu[k] = nu[k]
        self.m.relax, self.m.sigma = opts.drift_relax, opts.drift_diffusion
        self.m.cpress, self.m.cdiff = opts.drift_pressure_scale, opts.drift_diffusion_scale
        self.m.boussinesq, self.m.rho_ref = opts.boussinesq, opts.rho_ref
        for k in ["alpha", "alpha_new", "drift", "mforce", "amf"]:
            setattr(self.m, k, self.a[k].ctypes.data)

    def __getitem__(self, k):
        return self.a[k]

    def set_alpha(self, alpha, rho_m):
        k = alpha.shape[1]
        self.a["alpha"][:] = 0
        self.a["alpha"][:, :k] = alpha
        self.a["alpha"][:, 3] = rho_m

    def _call(self, name):
        getattr(lib(), self.st.pfx + name)(C.byref(self.st.params), C.byref(self.st.s),
                                           C.byref(self.m))

    def drift(self): self._call("mf_drift")
    def alpha_advance(self): self._call("mf_alpha")
    def correction(self): self._call("mf_correction")
    def force(self): self._call("mf_force")


def relerr(a, b, ref=None):
    """max |a-b| / max |ref| (normwise relative, infinity norm)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    r = np.asarray(b if ref is None else ref, np.float64)
    den = np.max(np.abs(r)) if r.size else 0.0
    num = np.max(np.abs(a - b)) if a.size else 0.0
    return num / den if den > 0 else num


# -- learned warm start (FP64 network; features from an FP32 state) --------
MLP_NF, MLP_H, MLP_NO = 24, 64, 4
MLP_NP = MLP_NF * MLP_H + MLP_H + MLP_H * MLP_H + MLP_H + MLP_H * MLP_NO + MLP_NO


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def mlp_features(st: State, alpha4=None, p_scale=1.0, dp_last=None):
    """24 features per particle of the FP32 state `st` after its a_ii pass;
    alpha4 [n][4] = (alpha_0..2, rho_m) or None for one phase; dp_last [n]
    the pressure increment of each particle's last solve or None."""
    assert st.dtype == np.float32
    X = np.zeros((st.n, MLP_NF), np.float64)
    fp = C.POINTER(C.c_float)
    a = None if alpha4 is None else np.ascontiguousarray(alpha4, np.float32)
    d = None if dp_last is None else np.ascontiguousarray(dp_last, np.float32)
    lib().orc_mlp_features(C.byref(st.params), C.byref(st.s),
                           None if a is None else a.ctypes.data_as(fp),
                           None if d is None else d.ctypes.data_as(fp),
                           C.c_double(p_scale), _dp(X))
    return X


def mlp_forward(w, X):
    """Y [n][4], H1, H2 [n][64] (post-ReLU) in FP64."""
    w = np.ascontiguousarray(w, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    n = len(X)
    Y = np.zeros((n, MLP_NO))
    H1 = np.zeros((n, MLP_H))
    H2 = np.zeros((n, MLP_H))
    lib().orc_mlp_forward(_dp(w), _dp(X), C.c_int(n), _dp(Y), _dp(H1), _dp(H2))
    return Y, H1, H2


def mlp_loss_grad(w, X, T, lw):
    """Loss (1/n) sum lw_o (y_so - T_so)^2 and its gradient (FP64)."""
    w = np.ascontiguousarray(w, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    T = np.ascontiguousarray(T, np.float64)
    lw = np.ascontiguousarray(lw, np.float64)
    g = np.zeros(MLP_NP)
    L = lib()
    L.orc_mlp_loss_grad.restype = C.c_double
    loss = L.orc_mlp_loss_grad(_dp(w), _dp(X), _dp(T), C.c_int(len(X)), _dp(lw), _dp(g))
    return loss, g


# -- MPM transfer -----------------------------------------------------------
class OrcMPM(C.Structure):
    _fields_ = [("kernel", C.c_int), ("physical", C.c_int), ("node_v", C.c_void_p),
                ("node_m", C.c_void_p), ("B", C.c_void_p), ("stress", C.c_void_p),
                ("G", C.c_void_p), ("K", C.c_void_p)]


class MPM:
    """MPM arrays attached to an oracle State (same dtype); node arrays in
    the y-major node order of the library."""

    def __init__(self, st: State, params, kernel=0, physical=0, calc=None):
        dt = st.dtype
        N = params.mpmRes.x * params.mpmRes.y * params.mpmRes.z
        self.st = st
        self.a = dict(node_v=np.zeros((N, 3), dt), node_m=np.zeros(N, dt),
                      B=np.zeros((st.n, 9), dt), stress=np.zeros((st.n, 9), dt),
                      G=np.zeros(st.n, dt), K=np.zeros(st.n, dt))
        if calc is not None:
            self.a["B"][:] = calc["B"].reshape(st.n, 9)
            self.a["stress"][:] = calc["stress"].reshape(st.n, 9)
            self.a["G"][:] = calc["solidG"]
            self.a["K"][:] = calc["solidK"]
        self.m = OrcMPM()
        self.m.kernel, self.m.physical = kernel, physical
        for k in ["node_v", "node_m", "B", "stress", "G", "K"]:
            setattr(self.m, k, self.a[k].ctypes.data)

    def __getitem__(self, k):
        return self.a[k]

    def _call(self, name, *extra):
        getattr(lib(), self.st.pfx + name)(C.byref(self.st.params), C.byref(self.st.s),
                                           C.byref(self.m), *extra)

    def p2g(self, apic=1): self._call("mpm_p2g", C.c_int(apic))
    def stress(self): self._call("mpm_stress")
    def node_update(self): self._call("mpm_node_update")
    def g2p(self, apic=1): self._call("mpm_g2p", C.c_int(apic))
    def momentum(self): self._call("mpm_momentum")


def mpm_weight(kernel, x, dtype=np.float64):
    L = lib()
    f = L.orc_f64_mpm_weight if dtype == np.float64 else L.orc_f32_mpm_weight
    f.restype = C.c_double
    return f(C.c_int(kernel), C.c_double(x))


def mpm_dweight(kernel, x, dtype=np.float64):
    L = lib()
    f = L.orc_f64_mpm_dweight if dtype == np.float64 else L.orc_f32_mpm_dweight
    f.restype = C.c_double
    return f(C.c_int(kernel), C.c_double(x))


def drucker_prager(params, G, K, eps, sig):
    out = np.zeros(9)
    e = np.ascontiguousarray(eps, np.float64).reshape(9)
    s = np.ascontiguousarray(sig, np.float64).reshape(9)
    lib().orc_f64_mpm_drucker_prager(C.byref(params), C.c_double(G), C.c_double(K), _dp(e), _dp(s),
                                     _dp(out))
    return out.reshape(3, 3)


# -- Stokesian mobility -------------------------------------------------------
def rpy_pair_ref(d):
    """PairwiseMobMatrix's A1N and B1N (equal spheres a = 1) for d = x_k - x_i."""
    d = np.ascontiguousarray(d, np.float64)
    A = np.zeros(9)
    B = np.zeros(9)
    lib().orc_rpy_pair_ref(_dp(d), _dp(A), _dp(B))
    return A.reshape(3, 3), B.reshape(3, 3)


def rpy_walkthrough(pos, force, ubg, dt):
    pos = np.ascontiguousarray(pos, np.float64)
    force = np.ascontiguousarray(force, np.float64)
    u = np.ascontiguousarray(ubg, np.float64).copy()
    w = np.zeros_like(u)
    pn = np.zeros_like(pos)
    lib().orc_rpy_walkthrough(C.c_int(len(pos)), _dp(pos), _dp(force), C.c_double(dt), _dp(u),
                              _dp(w), _dp(pn))
    return u, w, pn


def rpy_apply(pos, rad, mu, F):
    """Polydisperse RPY U = M F in FP64 (F [3n][K])."""
    pos = np.ascontiguousarray(pos, np.float64)
    rad = np.ascontiguousarray(rad, np.float64)
    F = np.ascontiguousarray(F, np.float64)
    U = np.zeros_like(F)
    lib().orc_rpy_apply(C.c_int(len(pos)), _dp(pos), _dp(rad), C.c_double(mu), _dp(F),
                        C.c_int(F.shape[1]), _dp(U))
    return U
