"""Python bindings of libqsolver.so (the gfx950 Qsolver C-ABI, include/qsolver.h).

Device memory and streams come from PyTorch-ROCm: every field is a CUDA
tensor (float64 [Ny, Nx] physical, complex128 [Ny, Nx//2+1] spectral, the
reference layouts) passed by data_ptr(); every Mesh runs on torch's current
stream.  There is no CPU fallback: without the built library or a GPU the
calls raise.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqsolver.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "qsolver.h")

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int

QS_OK = 0
(QS_ERR_INVALID_ARG, QS_ERR_UNSUPPORTED_SIZE, QS_ERR_HIP, QS_ERR_NONFINITE, QS_ERR_OOM, QS_ERR_STATE,
 QS_ERR_IO) = range(1, 8)
STATUS = {0: "ok", 1: "invalid argument", 2: "unsupported size", 3: "HIP runtime error",
          4: "non-finite value in state", 5: "out of device memory", 6: "invalid object state",
          7: "file I/O error"}


class QsolverError(RuntimeError):
    def __init__(self, status, fn, msg):
        super().__init__(f"{fn} failed: {STATUS.get(status, status)}: {msg}")
        self.status = status


class Qcomp(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double)]


class MeshStruct(ctypes.Structure):
    """Basic/Mesh.cuh:48-113 mirror (include/qsolver.h)."""
    _fields_ = [("Nx", _I), ("Nxh", _I), ("Ny", _I),
                ("Lx", _D), ("Ly", _D), ("dx", _D), ("dy", _D),
                ("specsize", _I), ("physsize", _I), ("wavesize", _I),
                ("BSZ", _I),
                ("spec", _P), ("phys", _P),
                ("kx", _P), ("ky", _P), ("k_squared", _P), ("cutoff", _P),
                ("priv", _P)]


class FieldStruct(ctypes.Structure):
    _fields_ = [("spec", _P), ("phys", _P), ("mesh", ctypes.POINTER(MeshStruct))]


class QsParams(ctypes.Structure):
    _fields_ = [("lam", _D), ("Re", _D), ("Er", _D), ("Rf", _D), ("alpha", _D),
                ("D_Q", _D), ("a_Q", _D), ("c_Q", _D), ("dt", _D),
                ("advect_w", _I), ("reserved", _I)]


_lib = None

# Objects that own device memory, closed in dependency order at interpreter
# exit (solvers, then fields, then meshes) before the HIP runtime goes away.
import atexit  # noqa: E402
import weakref  # noqa: E402

_live = {"solver": weakref.WeakSet(), "field": weakref.WeakSet(), "mesh": weakref.WeakSet()}


@atexit.register
def _close_all():
    for kind in ("solver", "field", "mesh"):
        for obj in list(_live[kind]):
            try:
                obj.close()
            except Exception:
                pass

# name -> argtypes (all functions return int status unless listed in _RESTYPE)
_SIGS = {
    "qs_mesh_create": [_I, _I, _D, _D, _I, ctypes.POINTER(ctypes.POINTER(MeshStruct))],
    "qs_mesh_destroy": [_P],
    "qs_mesh_set_stream": [_P, _P],
    "qs_mesh_synchronize": [_P],
    "qs_field_create": [_P, ctypes.POINTER(ctypes.POINTER(FieldStruct))],
    "qs_field_destroy": [_P],
    "qs_last_error": [],
    "qs_status_string": [_I],
    "qs_FwdTrans": [_P, _P, _P],
    "qs_BwdTrans": [_P, _P, _P],
    "qs_coeff": [_P, _P],
    "qs_symmetry_func": [_P, _P],
    "qs_FldMul": [_P, _P, _P, _D, _P],
    "qs_FldAdd": [_P, _D, _P, _D, _P, _P],
    "qs_FldAddScalar": [_P, _D, _P, _D, _P],
    "qs_FldSet": [_P, _D, _P],
    "qs_FldCopy": [_P, _P, _P],
    "qs_SpecSet": [_P, _P, _P],
    "qs_SpecSetConst": [_P, Qcomp, _P],
    "qs_SpecAdd": [_P, _D, _P, _D, _P, _P],
    "qs_SpecMul": [_P, _P, _D, _P],
    "qs_xDeriv": [_P, _P, _P],
    "qs_yDeriv": [_P, _P, _P],
    "qs_Laplacian": [_P, _P, _P],
    "qs_vel_func": [_P, _P, _P, _P],
    "qs_wlin_func": [_P, _P, _D, _D, _D, _P],
    "qs_r1lin_func": [_P, _P, _D, _P],
    "qs_r2lin_func": [_P, _P, _D, _P],
    "qs_S_func": [_P, _P, _P, _P],
    "qs_Rar": [_P, _P, _P, _P],
    "qs_single1": [_P, _P, _P, _D, _P],
    "qs_S2r": [_P, _P, _P, _P],
    "qs_PiDeriv": [_P, _P, _P, _P, _D, _D, _P],
    "qs_dealias": [_P, _P],
    "qs_single": [_P, _P, _P, _D],
    "qs_cross": [_P, _P, _P, _P],
    "qs_p11func": [_P, _P, _P, _P, _D],
    "qs_p12func": [_P, _P, _P, _P, _P, _D, _P, _P],
    "qs_p21func": [_P, _P, _P, _P, _P, _D, _P, _P],
    "qs_wnonl_func": [_P] * 11 + [_D, _D, _D, _P, _P],
    "qs_r1nonl_func": [_P] * 7 + [_D, _P, _P],
    "qs_r2nonl_func": [_P] * 7 + [_D, _P, _P],
    "qs_curr_func": [_P] * 6,
    "qs_ife_func1D": [_P, _P, _P, _P, _P, _D, _P],
    "qs_ifrk2_func0D": [_P, _P, _P, _P, _P, _D, _P],
    "qs_ifrk2_func1D": [_P, _P, _P, _P, _P, _P, _D, _P],
    "qs_ifrk2_func2D": [_P, _P, _P, _P, _P, _P, _D, _P],
    "qs_rk4_func0D": [_P, _P, _P, _P, _P, _D, _P],
    "qs_rk4_func1D": [_P, _P, _P, _P, _P, _P, _D, _P],
    "qs_rk4_func2D": [_P, _P, _P, _P, _P, _P, _D, _P],
    "qs_rk4_func3D": [_P, _P, _P, _P, _P, _P, _D, _P],
    "qs_rk4_func4D": [_P, _P, _P, _P, _P, _P, _D, _P],
    "qs_solver_create": [_P, ctypes.POINTER(QsParams), ctypes.POINTER(_P)],
    "qs_solver_destroy": [_P],
    "qs_solver_set_activity": [_P, _P],
    "qs_solver_set_state": [_P, _P, _P, _P],
    "qs_solver_get_state": [_P, _P, _P, _P],
    "qs_step_ifrk2": [_P, _I],
    "qs_step_ifrk2_unfused": [_P, _I],
    "qs_step_rk4_unfused": [_P, _I],
    "qs_solver_nonfinite": [_P, ctypes.POINTER(_I)],
    "qs_solver_profile": [_P, _I],
    "qs_solver_debug_phase": [_P, _I],
    "qs_solver_debug_array": [_P, _I, _P],
    "qs_debug_wave_fft": [_P, _I, _I, _P, _P],
    "qs_solver_profile_read": [_P, ctypes.POINTER(_D), ctypes.POINTER(_I)],
    "qs_synth_uniform": [ctypes.c_uint64, _D, _D, ctypes.c_size_t, _P],
    "qs_step_rk4": [_P, _I],
    "qs_step_ifrk2_noalloc": [_P, ctypes.POINTER(QsParams), _P, _P, _P, _P, _I, _P],
    "qs_step_ife": [_P, _I],
    "qs_adj_forward": [_P, _I, _I, _P],
    "qs_adj_backward": [_P] * 9,
    "qs_adj_loss": [_P, _I, _P, ctypes.POINTER(_D), _P],
    "qs_adj_memory": [_P, ctypes.POINTER(ctypes.c_size_t)],
    "qs_adj_release": [_P],
    "qs_ensemble_create": [_I, _P, _P, _P],
    "qs_ensemble_destroy": [_P],
    "qs_ensemble_set_stream": [_P, _P],
    "qs_ensemble_size": [_P, ctypes.POINTER(_I)],
    "qs_ensemble_solver": [_P, _I, _P],
    "qs_ensemble_set_state": [_P, _I, _P, _P, _P],
    "qs_ensemble_get_state": [_P, _I, _P, _P, _P],
    "qs_ensemble_set_activity": [_P, _I, _P],
    "qs_ensemble_step": [_P, _I],
    "qs_ensemble_get_config": [_P, _P],
    "qs_ensemble_set_config": [_P, _P],
    "qs_ensemble_tune": [_P, _I, _P, ctypes.POINTER(_D)],
    "qs_ensemble_adj_forward": [_P, _I, _I, _P],
    "qs_ensemble_adj_loss": [_P, _I, _P],
    "qs_ensemble_adj_backward": [_P, _P, _P],
    "qs_ensemble_adj_memory": [_P, ctypes.POINTER(ctypes.c_size_t)],
    "qs_Etot_func": [_P, _P, _P, _P],
    "qs_Espec_func": [_P, _D, _P, _P, _D, _P],
    "qs_Espec_shells": [_P, _P, _P, _D, _D, _I, _P, ctypes.POINTER(_I)],
    "qs_field_visual": [_P, ctypes.c_char_p, _I, _I],
    "qs_wave_visual": [_P, ctypes.c_char_p, _I, _I],
    "qs_spec_visual": [_P, ctypes.c_char_p, _I, _I],
    "qs_coord": [_D, _D, _I, _I, ctypes.c_char_p, ctypes.c_char_p],
    "qs_file_init": [ctypes.c_char_p, _P, _P],
    "qs_solver_snapshot_csv": [_P, ctypes.c_char_p, _I],
    "qs_solver_init_csv": [_P, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p],
    "qs_solver_save": [_P, ctypes.c_char_p, ctypes.c_int64],
    "qs_solver_load": [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(QsParams)],
}
_RESTYPE = {"qs_last_error": ctypes.c_char_p, "qs_status_string": ctypes.c_char_p}


def build():
    """Compiles libqsolver.so in-tree (hipcc, gfx950)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", os.path.dirname(_HERE)])


def load():
    """Loads the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, _I)
        _lib = L
    return _lib


def call(name, *args):
    L = load()
    st = getattr(L, name)(*args)
    if st != QS_OK:
        raise QsolverError(st, name, L.qs_last_error().decode(errors="replace"))
    return st


def last_error():
    return load().qs_last_error().decode(errors="replace")


def ptr(t):
    """Device pointer of a CUDA tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a CUDA tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def current_stream_handle():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def synth_uniform(seed, lo, hi, n):
    """splitmix64 uniform stream on the host (numpy float64 array)."""
    import numpy as np
    out = np.empty(n, dtype=np.float64)
    call("qs_synth_uniform", ctypes.c_uint64(seed), lo, hi, n, out.ctypes.data_as(_P))
    return out


class Mesh:
    """Mirror of the reference Mesh (Basic/Mesh.cuh:48-113) on one GPU."""

    def __init__(self, Nx, Ny, Lx, Ly=None, BSZ=16):
        Ly = Lx if Ly is None else Ly
        self._m = ctypes.POINTER(MeshStruct)()
        call("qs_mesh_create", Nx, Ny, float(Lx), float(Ly), BSZ, ctypes.byref(self._m))
        _live["mesh"].add(self)
        self.attach_stream()
        s = self._m.contents
        self.Nx, self.Ny, self.Nxh = s.Nx, s.Ny, s.Nxh
        self.Lx, self.Ly = s.Lx, s.Ly

    @property
    def handle(self):
        return self._m

    def attach_stream(self, stream=None):
        h = ctypes.c_void_p(stream.cuda_stream) if stream is not None else current_stream_handle()
        call("qs_mesh_set_stream", self._m, h)

    def close(self):
        if getattr(self, "_m", None):
            if _lib is not None:
                _lib.qs_mesh_destroy(self._m)
            self._m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def spec_shape(self):
        return (self.Ny, self.Nxh)

    @property
    def phys_shape(self):
        return (self.Ny, self.Nx)

    def spec(self):
        return torch.zeros(self.spec_shape, dtype=torch.complex128, device="cuda")

    def phys(self):
        return torch.zeros(self.phys_shape, dtype=torch.float64, device="cuda")

    def tables(self):
        """(kx, ky, k_squared, cutoff) copied into float64 [Ny, Nxh] tensors."""
        m = self._m.contents
        return tuple(_view(getattr(m, nm), self.spec_shape, torch.float64, "<f8").clone()
                     for nm in ("kx", "ky", "k_squared", "cutoff"))

    # ---- reference-named operations (FldOp/QFldfuncs.cuh:45-55) ----------
    def FwdTrans(self, ft, f):
        call("qs_FwdTrans", ptr(ft), ptr(f), self._m)

    def BwdTrans(self, f, ft):
        call("qs_BwdTrans", ptr(f), ptr(ft), self._m)

    def op(self, name, *args):
        """Calls qs_<name>(*args, mesh) with tensors converted to pointers."""
        conv = [ptr(a) if isinstance(a, torch.Tensor) else a for a in args]
        call("qs_" + name, *conv, self._m)

    # ---- diagnostics (post_process/postprocess.cu, EnergySpec.cu) ---------
    def Etot(self, u, v):
        """Etot_func of physical velocity fields: sum (u^2+v^2) dx dy / 2
        (divide by Lx*Ly for the driver's Etot)."""
        E = torch.zeros(1, dtype=torch.float64, device="cuda")
        call("qs_Etot_func", self._m, ptr(u), ptr(v), ptr(E))
        return float(E.item())

    def Espec(self, k, ut, vt, alpha):
        """Espec_func of one shell k^2 <= |K|^2 < (k+alpha)^2."""
        E = torch.zeros(1, dtype=torch.float64, device="cuda")
        call("qs_Espec_func", self._m, float(k), ptr(ut), ptr(vt), float(alpha), ptr(E))
        return float(E.item())

    def Espec_shells(self, ut, vt, dk, kmax, nmax=1 << 16):
        """Every shell of the EnergySpec driver loop (k = 0, dk, ... < kmax)
        as a cuda float64 tensor."""
        E = torch.zeros(nmax, dtype=torch.float64, device="cuda")
        n = _I(0)
        call("qs_Espec_shells", self._m, ptr(ut), ptr(vt), float(dk), float(kmax), int(nmax), ptr(E),
             ctypes.byref(n))
        return E[: n.value].clone()

    def file_init(self, path):
        """file_init (friction.cu:62-97): a physical CSV as a float64 [Ny, Nx] numpy array."""
        import numpy as np
        out = np.empty(self.phys_shape, dtype=np.float64)
        call("qs_file_init", os.fsencode(path), out.ctypes.data_as(_P), self._m)
        return out


class Field:
    """Mirror of the reference Field (Basic/Mesh.cuh:115-129)."""

    def __init__(self, mesh):
        self.mesh = mesh
        self._f = ctypes.POINTER(FieldStruct)()
        call("qs_field_create", mesh.handle, ctypes.byref(self._f))
        _live["field"].add(self)
        self.spec = _view(self._f.contents.spec, mesh.spec_shape, torch.complex128, "<c16")
        self.phys = _view(self._f.contents.phys, mesh.phys_shape, torch.float64, "<f8")

    @property
    def handle(self):
        return self._f

    def close(self):
        if getattr(self, "_f", None):
            if _lib is not None:
                _lib.qs_field_destroy(self._f)
            self._f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _view(p, shape, dtype, typestr):
    class _Holder:
        __cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr,
                                    "data": (int(p), False), "version": 3}
    return torch.as_tensor(_Holder(), device="cuda")


def params(lam=1.0, Re=1.0, Er=1.0, Rf=0.1, alpha=1.0, D_Q=1.0, a_Q=1.0, c_Q=1.0, dt=1e-3,
           advect_w=1):
    return QsParams(lam, Re, Er, Rf, alpha, D_Q, a_Q, c_Q, dt, int(advect_w), 0)


class Solver:
    """Fused IFRK2 solver (qs_solver_*, qs_step_ifrk2)."""

    def __init__(self, mesh, p):
        self.mesh = mesh
        self.p = p
        self._s = ctypes.c_void_p()
        call("qs_solver_create", mesh.handle, ctypes.byref(p), ctypes.byref(self._s))
        _live["solver"].add(self)

    def close(self):
        if getattr(self, "_s", None) and self._s.value:
            if _lib is not None:
                _lib.qs_solver_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_activity(self, Ra=None):
        call("qs_solver_set_activity", self._s, ptr(Ra))

    def set_state(self, w, r1, r2):
        call("qs_solver_set_state", self._s, ptr(w), ptr(r1), ptr(r2))

    def get_state(self):
        out = [self.mesh.spec() for _ in range(3)]
        call("qs_solver_get_state", self._s, *(ptr(t) for t in out))
        return out

    def step(self, nsteps=1):
        call("qs_step_ifrk2", self._s, int(nsteps))

    def step_rk4(self, nsteps=1):
        call("qs_step_rk4", self._s, int(nsteps))

    def step_ife(self, nsteps=1):
        call("qs_step_ife", self._s, int(nsteps))

    def step_unfused(self, nsteps=1):
        call("qs_step_ifrk2_unfused", self._s, int(nsteps))

    def step_rk4_unfused(self, nsteps=1):
        call("qs_step_rk4_unfused", self._s, int(nsteps))

    KERNEL_CLASSES = ("col_prologue", "row_rhs", "col_stage1", "col_stage2")

    def profile(self, enable=True):
        call("qs_solver_profile", self._s, int(bool(enable)))

    def profile_read(self):
        """{kernel class: (average ms, launches)} since the last read."""
        avg = (_D * 4)()
        cnt = (_I * 4)()
        call("qs_solver_profile_read", self._s, avg, cnt)
        return {k: (avg[i], cnt[i]) for i, k in enumerate(self.KERNEL_CLASSES)}

    def debug_phase(self, phase):
        call("qs_solver_debug_phase", self._s, int(phase))

    def debug_array(self, which):
        out = self.mesh.spec()
        call("qs_solver_debug_array", self._s, int(which), ptr(out))
        return out

    def nonfinite(self):
        flag = _I(0)
        call("qs_solver_nonfinite", self._s, ctypes.byref(flag))
        return bool(flag.value)

    # ---- snapshot and restart (friction.cu:163-185, :262-279) --------------
    def snapshot_csv(self, prefix, step):
        """<prefix><step>{w,r1,r2,S}.csv of the current state."""
        call("qs_solver_snapshot_csv", self._s, os.fsencode(str(prefix)), int(step))

    def init_csv(self, w_csv, r1_csv, r2_csv):
        call("qs_solver_init_csv", self._s, os.fsencode(w_csv), os.fsencode(r1_csv), os.fsencode(r2_csv))

    def save(self, path, step=0):
        """Exact binary checkpoint of state and activity."""
        call("qs_solver_save", self._s, os.fsencode(path), int(step))

    def load(self, path):
        """Restores state and activity; returns (stored step, stored qs_params)."""
        st = ctypes.c_int64(0)
        p = QsParams()
        call("qs_solver_load", self._s, os.fsencode(path), ctypes.byref(st), ctypes.byref(p))
        return st.value, p

    # ---- discrete adjoint (qs_adj_*) ---------------------------------------
    def adj_forward(self, nsteps, ckpt_every=0, alpha_t=None):
        """nsteps steps from the current state with checkpoints every
        ckpt_every steps; alpha_t: cuda float64 [nsteps] per-step activity."""
        call("qs_adj_forward", self._s, int(nsteps), int(ckpt_every), ptr(alpha_t))

    def adj_backward(self, gW, gR1, gR2, want_field=True, want_sched=False, want_state=False, nsteps=None):
        """Reverse sweep: returns (dL/dRa field or None, dL/dalpha_t or None,
        (dL/dW0, dL/dR1_0, dL/dR2_0) or None)."""
        m = self.mesh
        ab = torch.empty(m.phys_shape, dtype=torch.float64, device="cuda") if want_field else None
        at = torch.empty(nsteps, dtype=torch.float64, device="cuda") if want_sched else None
        u = [m.spec() for _ in range(3)] if want_state else [None, None, None]
        call("qs_adj_backward", self._s, ptr(gW), ptr(gR1), ptr(gR2), ptr(ab), ptr(at), *(ptr(t) for t in u))
        return ab, at, (u if want_state else None)

    def adj_memory(self):
        b = ctypes.c_size_t(0)
        call("qs_adj_memory", self._s, ctypes.byref(b))
        return b.value

    def adj_release(self):
        call("qs_adj_release", self._s)


ENSTROPHY, KINETIC = 0, 1


def _host(a, dtype):
    import numpy as np
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(_P)


def field_visual(f, name):
    """field_visual (BasicUtils.cuh:38-50) of a host [Ny, Nx] float64 array."""
    import numpy as np
    a, p = _host(f, np.float64)
    call("qs_field_visual", p, os.fsencode(name), a.shape[1], a.shape[0])


def wave_visual(f, name):
    """wave_visual (EnergySpec.cu:8-20) of a host [Ny, Nxh] float64 array."""
    import numpy as np
    a, p = _host(f, np.float64)
    call("qs_wave_visual", p, os.fsencode(name), a.shape[1], a.shape[0])


def spec_visual(f, name):
    """spec_visual (EnergySpec.cu:22-34) of a host [Ny, Nxh] complex128 array."""
    import numpy as np
    a, p = _host(f, np.complex128)
    call("qs_spec_visual", p, os.fsencode(name), a.shape[1], a.shape[0])


def coord(dx, dy, Nx, Ny, xname="x.csv", yname="y.csv"):
    """coord (BasicUtils.cuh:52-67)."""
    call("qs_coord", float(dx), float(dy), int(Nx), int(Ny), os.fsencode(xname), os.fsencode(yname))


def adj_loss(mesh, W, kind=ENSTROPHY):
    """(loss, dL/dW) of a vorticity spectrum (qs_adj_loss)."""
    L = _D(0.0)
    Wb = mesh.spec()
    call("qs_adj_loss", mesh.handle, int(kind), ptr(W), ctypes.byref(L), ptr(Wb))
    return L.value, Wb


class EnsembleConfig(ctypes.Structure):
    """qs_ensemble_config."""
    _fields_ = [("graph_steps", _I), ("concurrent", _I), ("order", _I),
                ("row_blocks", _I * 8), ("col_blocks", _I * 8), ("lanes", _I * 8), ("merged", _I)]

    def as_dict(self):
        return {"graph_steps": self.graph_steps, "concurrent": self.concurrent, "order": self.order,
                "row_blocks": list(self.row_blocks), "col_blocks": list(self.col_blocks),
                "lanes": list(self.lanes), "merged": self.merged}


class Ensemble:
    """Independent IFRK2 runs of uneven square meshes advanced together
    (qs_ensemble_*).  sizes: mesh size per run; params: qs_params per run."""

    def __init__(self, sizes, params_list):
        n = len(sizes)
        self.sizes = list(sizes)
        Ns = (_I * n)(*sizes)
        ps = (QsParams * n)(*params_list)
        self._e = ctypes.c_void_p()
        call("qs_ensemble_create", n, Ns, ps, ctypes.byref(self._e))
        self._meshes = {}
        _live["solver"].add(self)

    def close(self):
        if getattr(self, "_e", None) and self._e.value:
            if _lib is not None:
                _lib.qs_ensemble_destroy(self._e)
            self._e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def spec(self, r):
        N = self.sizes[r]
        return torch.zeros((N, N // 2 + 1), dtype=torch.complex128, device="cuda")

    def set_state(self, r, w, r1, r2):
        call("qs_ensemble_set_state", self._e, int(r), ptr(w), ptr(r1), ptr(r2))

    def get_state(self, r):
        out = [self.spec(r) for _ in range(3)]
        call("qs_ensemble_get_state", self._e, int(r), *(ptr(t) for t in out))
        return out

    def set_activity(self, r, Ra):
        call("qs_ensemble_set_activity", self._e, int(r), ptr(Ra))

    def step(self, nsteps=1):
        call("qs_ensemble_step", self._e, int(nsteps))

    def config(self):
        c = EnsembleConfig()
        call("qs_ensemble_get_config", self._e, ctypes.byref(c))
        return c

    def set_config(self, c):
        call("qs_ensemble_set_config", self._e, ctypes.byref(c))

    # ---- batched discrete adjoint (runs of one mesh size) -------------------
    def adj_forward(self, nsteps, ckpt_every=0, alpha_t=None):
        """alpha_t: cuda float64 [n, nsteps] per-run per-step activity, or None."""
        call("qs_ensemble_adj_forward", self._e, int(nsteps), int(ckpt_every), ptr(alpha_t))

    def adj_loss(self, kind=0):
        """Per-run losses (numpy [n]) of the final states; sets the reverse seeds."""
        import numpy as np
        out = np.zeros(len(self.sizes))
        call("qs_ensemble_adj_loss", self._e, int(kind), out.ctypes.data_as(_P))
        return out

    def adj_backward(self, nsteps, want_sched=True, want_field=False):
        n = len(self.sizes)
        N = self.sizes[0]
        at = torch.empty((n, nsteps), dtype=torch.float64, device="cuda") if want_sched else None
        ab = torch.empty((n, N, N), dtype=torch.float64, device="cuda") if want_field else None
        call("qs_ensemble_adj_backward", self._e, ptr(at), ptr(ab))
        return at, ab

    def adj_memory(self):
        b = ctypes.c_size_t(0)
        call("qs_ensemble_adj_memory", self._e, ctypes.byref(b))
        return b.value

    def tune(self, trial_steps=5):
        c = EnsembleConfig()
        ms = _D(0.0)
        call("qs_ensemble_tune", self._e, int(trial_steps), ctypes.byref(c), ctypes.byref(ms))
        return c, ms.value
