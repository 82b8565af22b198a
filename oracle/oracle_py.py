"""ctypes wrapper of the FP64 CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker or the timed CPU baseline.  The product path
(nature-control_amd/qsolver) never imports it.

Arrays follow the reference layouts (Basic/Mesh.cuh:48-113,
FldOp/QFldfuncs.cuh:67,122):
  physical  float64   [Ny, Nx]      phys[j*Nx + i]
  spectral  complex128 [Ny, Nx//2+1] spec[j*Nxh + i]
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

IFE, IFRK2, RK4 = 0, 1, 2
_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "or_mesh_new": (_P, [_I, _I, _D, _D]),
            "or_mesh_free": (None, [_P]),
            "or_tables": (None, [_P, _P, _P, _P, _P]),
            "or_field_new": (_P, [_P]),
            "or_field_free": (None, [_P]),
            "or_field_spec": (_P, [_P]),
            "or_field_phys": (_P, [_P]),
            "or_fft1d": (None, [_I, _I, _P, _P]),
            "or_FwdTrans": (None, [_P, _P, _P]),
            "or_BwdTrans": (None, [_P, _P, _P]),
            "or_single": (None, [_P, _P, _P, _D]),
            "or_cross": (None, [_P, _P, _P, _P]),
            "or_p11func": (None, [_P, _P, _P, _P, _D]),
            "or_p12func": (None, [_P, _P, _P, _P, _P, _D, _P, _P]),
            "or_p21func": (None, [_P, _P, _P, _P, _P, _D, _P, _P]),
            "or_wnonl_func": (None, [_P] * 11 + [_D, _D, _D, _P, _P]),
            "or_r1nonl_func": (None, [_P] * 7 + [_D, _P, _P]),
            "or_r2nonl_func": (None, [_P] * 7 + [_D, _P, _P]),
            "or_curr_func": (None, [_P] * 6),
            "or_solver_new": (_P, [_P, _P]),
            "or_solver_free": (None, [_P]),
            "or_solver_set_activity": (None, [_P, _P]),
            "or_solver_set_state": (None, [_P, _P, _P, _P]),
            "or_solver_get_state": (None, [_P, _P, _P, _P]),
            "or_solver_step": (None, [_P, _I, _I]),
            "or_solver_rhs": (None, [_P] * 7),
            "or_solver_lin": (None, [_P] * 7),
            "or_adj_B": (None, [_P] * 3),
            "or_adj_BT": (None, [_P] * 3),
            "or_adj_F": (None, [_P] * 3),
            "or_adj_FT": (None, [_P] * 3),
            "or_adj_rhs_vjp": (None, [_P] * 11),
            "or_adj_step_vjp": (None, [_P] * 11),
            "or_adj_loss_grad": (_D, [_P, _P, _I, _P]),
            "or_adj_gradient": (_D, [_P, _I, _I, _P, _P, _P]),
            "or_Etot_func": (_D, [_P, _P, _I, _I, _D, _D]),
            "or_Espec_func": (_D, [_P, _D, _P, _P, _D]),
            "or_Espec_shells": (_I, [_P, _P, _P, _D, _D, _I, _P]),
            "or_set_threads": (None, [_I]),
            "or_get_threads": (_I, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def set_threads(n):
    lib().or_set_threads(int(n))


def get_threads():
    return lib().or_get_threads()


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().or_fft1d(x.size, sign, _ptr(x), _ptr(out))
    return out


class Mesh:
    """Mirror of the reference Mesh (Basic/Mesh.cuh:48-113)."""

    def __init__(self, Nx, Ny, Lx=2 * np.pi, Ly=None):
        Ly = Lx if Ly is None else Ly
        self.Nx, self.Ny, self.Nxh = Nx, Ny, Nx // 2 + 1
        self.Lx, self.Ly = Lx, Ly
        self.h = lib().or_mesh_new(Nx, Ny, Lx, Ly)
        kx, ky, k2, cut = (np.empty((Ny, self.Nxh)) for _ in range(4))
        lib().or_tables(self.h, _ptr(kx), _ptr(ky), _ptr(k2), _ptr(cut))
        self.kx, self.ky, self.k_squared, self.cutoff = kx, ky, k2, cut

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_mesh_free(self.h)
            self.h = None

    @property
    def spec_shape(self):
        return (self.Ny, self.Nxh)

    @property
    def phys_shape(self):
        return (self.Ny, self.Nx)

    def grid(self):
        x = np.arange(self.Nx) * (self.Lx / self.Nx)
        y = np.arange(self.Ny) * (self.Ly / self.Ny)
        return np.meshgrid(x, y)  # X[j, i] = i*dx, Y[j, i] = j*dy

    def fwd(self, f):
        """FwdTrans (FldOp/QFldfuncs.cuh:45-48): unnormalised R2C."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        ft = np.empty(self.spec_shape, np.complex128)
        lib().or_FwdTrans(_ptr(ft), _ptr(f), self.h)
        return ft

    def bwd(self, ft):
        """BwdTrans (FldOp/QFldfuncs.cuh:50-55).  ft is symmetrised in place."""
        assert ft.dtype == np.complex128 and ft.flags.c_contiguous
        f = np.empty(self.phys_shape)
        lib().or_BwdTrans(_ptr(f), _ptr(ft), self.h)
        return f

    def vel(self, wt):
        """vel_func (FldOp/QFldfuncs.cuh:204-224): (u, v) spectra of a vorticity spectrum."""
        k2 = self.k_squared.copy()
        k2[0, 0] = 1.0
        psi = wt / k2
        psi[0, 0] = 0.0
        return 1j * self.ky * psi, -1j * self.kx * psi

    def Etot(self, u, v):
        """Etot_func (post_process/postprocess.cu:54-65)."""
        u = np.ascontiguousarray(u, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        return lib().or_Etot_func(_ptr(u), _ptr(v), self.Nx, self.Ny, self.Lx / self.Nx, self.Ly / self.Ny)

    def Espec(self, k, ut, vt, alpha):
        """Espec_func (post_process/EnergySpec.cu:81-96), one shell."""
        ut = np.ascontiguousarray(ut, np.complex128)
        vt = np.ascontiguousarray(vt, np.complex128)
        return lib().or_Espec_func(self.h, k, _ptr(ut), _ptr(vt), alpha)

    def Espec_shells(self, ut, vt, dk, kmax, nmax=4096):
        """The EnergySpec.cu:194-204 driver loop: one value per shell."""
        ut = np.ascontiguousarray(ut, np.complex128)
        vt = np.ascontiguousarray(vt, np.complex128)
        E = np.empty(nmax)
        n = lib().or_Espec_shells(self.h, _ptr(ut), _ptr(vt), dk, kmax, nmax, _ptr(E))
        return E[:n].copy()


class Field:
    """Mirror of the reference Field (Basic/Mesh.cuh:115-129)."""

    def __init__(self, mesh):
        self.mesh = mesh
        self.h = lib().or_field_new(mesh.h)
        ns = mesh.Ny * mesh.Nxh
        npx = mesh.Ny * mesh.Nx
        sp = ctypes.cast(lib().or_field_spec(self.h), ctypes.POINTER(ctypes.c_double))
        ph = ctypes.cast(lib().or_field_phys(self.h), ctypes.POINTER(ctypes.c_double))
        self.spec = np.ctypeslib.as_array(sp, shape=(2 * ns,)).view(np.complex128).reshape(mesh.spec_shape)
        self.phys = np.ctypeslib.as_array(ph, shape=(npx,)).reshape(mesh.phys_shape)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_field_free(self.h)
            self.h = None


def params_array(lam=0.1, Re=0.1, Er=0.1, Rf=7.5e-5, alpha=0.2, DQ=1.0, aQ=1.0, cQ=1.0, dt=1e-3,
                 advect_w=1):
    return np.array([lam, Re, Er, Rf, alpha, DQ, aQ, cQ, dt, float(advect_w)], dtype=np.float64)


class Solver:
    """Whole-step oracle: friction.cu:150-260 with the IFE / IFRK2 / RK4 stages."""

    def __init__(self, mesh, params):
        self.mesh = mesh
        self.p = np.ascontiguousarray(params, dtype=np.float64)
        self.h = lib().or_solver_new(mesh.h, _ptr(self.p))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_solver_free(self.h)
            self.h = None

    def set_activity(self, Ra=None):
        Ra = None if Ra is None else np.ascontiguousarray(Ra, dtype=np.float64)
        lib().or_solver_set_activity(self.h, _ptr(Ra))

    def set_state(self, w, r1, r2):
        w, r1, r2 = (np.ascontiguousarray(a, dtype=np.complex128) for a in (w, r1, r2))
        lib().or_solver_set_state(self.h, _ptr(w), _ptr(r1), _ptr(r2))

    def get_state(self):
        out = [np.empty(self.mesh.spec_shape, np.complex128) for _ in range(3)]
        lib().or_solver_get_state(self.h, *(_ptr(a) for a in out))
        return out

    def step(self, nsteps=1, method=IFRK2):
        lib().or_solver_step(self.h, method, nsteps)

    def rhs(self, w, r1, r2):
        """Returns (N_w, N_r1, N_r2) for stage spectra w, r1, r2 (copies)."""
        ins = [np.array(a, dtype=np.complex128, order="C") for a in (w, r1, r2)]
        outs = [np.empty(self.mesh.spec_shape, np.complex128) for _ in range(3)]
        lib().or_solver_rhs(self.h, *(_ptr(a) for a in ins + outs))
        return outs

    def lin_tables(self):
        out = [np.empty(self.mesh.spec_shape) for _ in range(6)]
        lib().or_solver_lin(self.h, *(_ptr(a) for a in out))
        return out


class Adjoint:
    """Discrete adjoint of the oracle's IFRK2 step (exact transposes in the
    real inner products of the stored arrays).  The control is the solver's
    activity field; loss kind 0 = 1/2 mean(w^2), 1 = 1/2 mean(u^2 + v^2)."""

    ENSTROPHY, KINETIC = 0, 1

    def __init__(self, solver):
        self.s = solver
        self.m = solver.mesh

    def _c(self, a):
        return np.ascontiguousarray(a, dtype=np.complex128)

    def _r(self, a):
        return np.ascontiguousarray(a, dtype=np.float64)

    def B(self, G):
        f = np.empty(self.m.phys_shape)
        lib().or_adj_B(self.s.h, _ptr(self._c(G)), _ptr(f))
        return f

    def BT(self, f):
        G = np.empty(self.m.spec_shape, np.complex128)
        lib().or_adj_BT(self.s.h, _ptr(self._r(f)), _ptr(G))
        return G

    def F(self, f):
        G = np.empty(self.m.spec_shape, np.complex128)
        lib().or_adj_F(self.s.h, _ptr(self._r(f)), _ptr(G))
        return G

    def FT(self, G):
        f = np.empty(self.m.phys_shape)
        lib().or_adj_FT(self.s.h, _ptr(self._c(G)), _ptr(f))
        return f

    def rhs_vjp(self, X, Nbar):
        """Returns (Xbar (3 spectra), abar (physical)) = VJP of the right-hand side."""
        ins = [self._c(a) for a in list(X) + list(Nbar)]
        outs = [np.empty(self.m.spec_shape, np.complex128) for _ in range(3)]
        abar = np.zeros(self.m.phys_shape)
        lib().or_adj_rhs_vjp(self.s.h, *(_ptr(a) for a in ins + outs), _ptr(abar))
        return outs, abar

    def step_vjp(self, u, gbar):
        ins = [self._c(a) for a in list(u) + list(gbar)]
        outs = [np.empty(self.m.spec_shape, np.complex128) for _ in range(3)]
        abar = np.zeros(self.m.phys_shape)
        lib().or_adj_step_vjp(self.s.h, *(_ptr(a) for a in ins + outs), _ptr(abar))
        return outs, abar

    def loss_grad(self, W, kind):
        Wb = np.empty(self.m.spec_shape, np.complex128)
        L = lib().or_adj_loss_grad(self.s.h, _ptr(self._c(W)), int(kind), _ptr(Wb))
        return L, Wb

    def gradient(self, nsteps, kind, alpha_t=None):
        """Loss after nsteps from the solver's state, dL/d(activity field) and,
        with a per-step uniform activity alpha_t, dL/dalpha_t.  Leaves the
        solver at the final state."""
        ab = np.empty(self.m.phys_shape)
        at = None if alpha_t is None else self._r(alpha_t)
        abt = None if alpha_t is None else np.empty(nsteps)
        L = lib().or_adj_gradient(self.s.h, int(nsteps), int(kind), _ptr(at), _ptr(ab), _ptr(abt))
        return L, ab, abt
