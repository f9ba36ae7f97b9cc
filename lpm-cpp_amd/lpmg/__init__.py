"""Python binding of liblpmg.so (the C ABI declared in include/lpm_gpu.h).

`System` is one LPM particle system resident on the GPU: the Assembly,
Stiffness and Solver state of LPM-CPP's quasi-static Newton step.  Its methods
keep the reference's names (forward = updateGeometry + updateForceState,
assemble = calcStiffness + updateStiffnessDispBC, newton = NewtonIteration,
solve_step = solveProblemStep, ...).  The library is the only compute path:
there is no CPU fallback, and a missing or unloadable liblpmg.so raises.
"""
import ctypes
import os

import numpy as np

from . import lattice  # noqa: F401

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "liblpmg.so")
CSRC = os.path.join(_PKG, "csrc")
_lib = None

FIELDS = dict(XYZ=0, XYZ_INITIAL=1, XYZ_LAST=2, PIN=3, PEX=4, STRESS=5, DAMAGE=6,
              DAMAGE_VISUAL=7, STATE_VAR=8, RESIDUAL=9, BDAMAGE=10, DLP=11, BFORCE=12,
              DIS=13, CS=14, CONSTRAINT=15, DDOT_LOCAL=16, DDOT_NONLOCAL=17, DLE_TOTAL=18,
              CS_SUM=19, BSTRAIN=20, DLE=21, DAMAGE_LAST=22, TYPE=23, TDLE_TOTAL=24, FLOATING=25)

# every symbol include/lpm_gpu.h declares
EXPORTS = ("lpmg_version", "lpmg_last_error", "lpmg_unit_cell", "lpmg_create", "lpmg_create_with_bonds",
           "lpmg_destroy", "lpmg_sizes", "lpmg_export_bonds", "lpmg_export_conns", "lpmg_export_csr",
           "lpmg_set_types", "lpmg_set_material", "lpmg_forward", "lpmg_update_rr", "lpmg_assemble",
           "lpmg_spmv", "lpmg_spmv_noalloc", "lpmg_solve_linear", "lpmg_newton", "lpmg_update_state_var",
           "lpmg_update_static_damage", "lpmg_update_fatigue_damage", "lpmg_store_state", "lpmg_reset_state", "lpmg_search_nonlocal",
           "lpmg_export_nonlocal",
           "lpmg_define_group", "lpmg_disp_bc", "lpmg_force_bc", "lpmg_solve_step", "lpmg_restart", "lpmg_field_len",
           "lpmg_get_field", "lpmg_set_field", "lpmg_profile", "lpmg_profile_read", "lpmg_time_kernel",
           "lpmg_synchronize", "lpmg_request_bytes", "lpmg_server_create", "lpmg_server_submit",
           "lpmg_server_wait", "lpmg_server_info", "lpmg_server_destroy", "lpmg_multigrid_levels")


class LpmgError(RuntimeError):
    pass


class NotConverged(LpmgError):
    pass


def lib():
    """Load liblpmg.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32p, i64p, dp = (ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                              ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double))
        for name in EXPORTS:
            getattr(L, name).restype = ctypes.c_int
        L.lpmg_last_error.restype = ctypes.c_char_p
        L.lpmg_request_bytes.restype = ctypes.c_int64
        L.lpmg_server_destroy.restype = None
        L.lpmg_destroy.restype = None
        L.lpmg_destroy.argtypes = [vp]
        L.lpmg_field_len.restype = ctypes.c_int64
        L.lpmg_field_len.argtypes = [vp, ctypes.c_int]
        L.lpmg_unit_cell.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.lpmg_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp,
                                  ctypes.POINTER(vp)]
        L.lpmg_create_with_bonds.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp,
                                             ctypes.c_int64, i32p, i32p, i32p, ctypes.POINTER(vp)]
        L.lpmg_sizes.argtypes = [vp, i64p]
        L.lpmg_export_bonds.argtypes = [vp, i32p, i32p, i32p]
        L.lpmg_export_conns.argtypes = [vp, i32p, i32p, i32p]
        L.lpmg_export_nonlocal.argtypes = [vp, i32p, i32p]
        L.lpmg_export_csr.argtypes = [vp, i64p, i64p, i64p, dp]
        L.lpmg_set_types.argtypes = [vp, i32p]
        L.lpmg_set_material.argtypes = [vp, ctypes.c_int, i32p, ctypes.c_int, dp]
        for name in ("lpmg_forward", "lpmg_update_state_var", "lpmg_store_state", "lpmg_synchronize",
                     "lpmg_restart"):
            getattr(L, name).argtypes = [vp]
        L.lpmg_update_rr.argtypes = [vp, dp]
        L.lpmg_assemble.argtypes = [vp, ctypes.c_int]
        L.lpmg_spmv.argtypes = [vp, ctypes.c_int, dp, dp]
        L.lpmg_spmv_noalloc.argtypes = [vp, ctypes.c_int, dp, dp]
        L.lpmg_solve_linear.argtypes = [vp, dp, dp]
        L.lpmg_newton.argtypes = [vp, ctypes.c_int, ctypes.c_double, dp, dp]
        L.lpmg_update_static_damage.argtypes = [vp, ctypes.c_int]
        L.lpmg_update_fatigue_damage.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.lpmg_reset_state.argtypes = [vp, ctypes.c_int]
        L.lpmg_search_nonlocal.argtypes = [vp, ctypes.c_double]
        L.lpmg_define_group.argtypes = [vp, ctypes.c_int, i32p]
        L.lpmg_disp_bc.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.lpmg_force_bc.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double]
        L.lpmg_solve_step.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp]
        L.lpmg_get_field.argtypes = [vp, ctypes.c_int, dp]
        L.lpmg_set_field.argtypes = [vp, ctypes.c_int, dp]
        L.lpmg_profile.argtypes = [vp, ctypes.c_int]
        L.lpmg_profile_read.argtypes = [vp, dp]
        L.lpmg_time_kernel.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.lpmg_multigrid_levels.argtypes = [vp, i32p, ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _err():
    return lib().lpmg_last_error().decode()


def solver_opts(precond=1, rtol=1e-12, atol=1e-12, maxit=0, precision=0, chunk=16, predict=0.0):
    """opts[8] of lpmg_solve_linear / lpmg_newton / lpmg_solve_step.  precond:
    0 plain CG (the reference's MKL call), 1 Jacobi CG, 2 multigrid CG, 3 Jacobi
    BiCGSTAB, 4 smoothed-aggregation multigrid CG; precision 1: FP32 stiffness
    and inner solve in the FP64 Newton."""
    return np.array([precond, rtol, atol, maxit, precision, chunk, predict, 0], dtype=np.float64)


class System:
    """An LPM particle system on the GPU (LPM-CPP's Assembly + Stiffness + Solver)."""

    def __init__(self, lattice_type, radius, ptype, xyz, bonds=None):
        L = lib()
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        h = ctypes.c_void_p()
        if bonds is None:
            rc = L.lpmg_create(int(lattice_type), float(radius), int(ptype), xyz.shape[0],
                               _p(xyz, ctypes.c_double), ctypes.byref(h))
        else:
            p1, p2, layer = (np.ascontiguousarray(b, dtype=np.int32) for b in bonds)
            rc = L.lpmg_create_with_bonds(int(lattice_type), float(radius), int(ptype), xyz.shape[0],
                                          _p(xyz, ctypes.c_double), len(p1), _p(p1, ctypes.c_int32),
                                          _p(p2, ctypes.c_int32), _p(layer, ctypes.c_int32),
                                          ctypes.byref(h))
        if rc != 0:
            raise LpmgError(f"lpmg_create failed ({rc}): {_err()}")
        self.h = h
        s = np.zeros(8, dtype=np.int64)
        self._chk(L.lpmg_sizes(self.h, _p(s, ctypes.c_int64)))
        self.N, self.dim, self.nbond, self.nnz, self.nnzb, self.max_nb, self.max_nconn, self.ptype = (
            int(v) for v in s)

    def close(self):
        if getattr(self, "h", None):
            lib().lpmg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc == -4:
            raise NotConverged(_err())
        if rc < 0:
            raise LpmgError(f"liblpmg error {rc}: {_err()}")
        return rc

    # -- structure exports
    def bonds(self):
        bptr = np.zeros(self.N + 1, np.int32)
        nb0 = np.zeros(self.N, np.int32)
        p2 = np.zeros(self.nbond, np.int32)
        self._chk(lib().lpmg_export_bonds(self.h, _p(bptr, ctypes.c_int32), _p(nb0, ctypes.c_int32),
                                          _p(p2, ctypes.c_int32)))
        return bptr, nb0, p2

    def conns(self):
        cptr = np.zeros(self.N + 1, np.int32)
        conns = np.zeros(self.nnzb, np.int32)
        ncl = np.zeros(self.N, np.int32)
        self._chk(lib().lpmg_export_conns(self.h, _p(cptr, ctypes.c_int32), _p(conns, ctypes.c_int32),
                                          _p(ncl, ctypes.c_int32)))
        return cptr, conns, ncl

    def csr(self):
        Kp = np.zeros(self.N + 1, np.int64)
        IK = np.zeros(self.dim * self.N + 1, np.int64)
        JK = np.zeros(self.nnz, np.int64)
        K = np.zeros(self.nnz)
        self._chk(lib().lpmg_export_csr(self.h, _p(Kp, ctypes.c_int64), _p(IK, ctypes.c_int64),
                                        _p(JK, ctypes.c_int64), _p(K, ctypes.c_double)))
        return Kp, IK, JK, K

    # -- setup
    def set_types(self, t):
        t = np.ascontiguousarray(t, dtype=np.int32)
        self._chk(lib().lpmg_set_types(self.h, _p(t, ctypes.c_int32)))

    def set_material(self, model, params, idx=None):
        p = np.zeros(12)
        p[:len(params)] = params
        if idx is None:
            self._chk(lib().lpmg_set_material(self.h, 0, None, int(model), _p(p, ctypes.c_double)))
        else:
            idx = np.ascontiguousarray(idx, dtype=np.int32)
            self._chk(lib().lpmg_set_material(self.h, len(idx), _p(idx, ctypes.c_int32), int(model),
                                              _p(p, ctypes.c_double)))

    def define_group(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        return self._chk(lib().lpmg_define_group(self.h, len(idx), _p(idx, ctypes.c_int32)))

    def disp_bc(self, gid, mode, flag, step):
        self._chk(lib().lpmg_disp_bc(self.h, gid, mode, ord(flag), float(step)))

    def force_bc(self, gid, mode, fx, fy, fz):
        self._chk(lib().lpmg_force_bc(self.h, gid, mode, float(fx), float(fy), float(fz)))

    # -- hot path
    def forward(self):
        self._chk(lib().lpmg_forward(self.h))

    def update_rr(self):
        out = np.zeros(2)
        self._chk(lib().lpmg_update_rr(self.h, _p(out, ctypes.c_double)))
        return out

    def assemble(self, precision=0):
        self._chk(lib().lpmg_assemble(self.h, int(precision)))

    def spmv(self, x, precision=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self._chk(lib().lpmg_spmv(self.h, int(precision), _p(x, ctypes.c_double), _p(y, ctypes.c_double)))
        return y

    def spmv_noalloc(self, x, precision=0):
        """spmv through the allocation-free entry point (static device staging)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        self._chk(lib().lpmg_spmv_noalloc(self.h, int(precision), _p(x, ctypes.c_double),
                                          _p(y, ctypes.c_double)))
        return y

    def solve_linear(self, **kw):
        o = solver_opts(**kw)
        st = np.zeros(4)
        self._chk(lib().lpmg_solve_linear(self.h, _p(o, ctypes.c_double), _p(st, ctypes.c_double)))
        return st

    def newton(self, max_iter=30, tol=1e-5, **kw):
        o = solver_opts(**kw)
        st = np.zeros(4)
        n = self._chk(lib().lpmg_newton(self.h, int(max_iter), float(tol), _p(o, ctypes.c_double),
                                        _p(st, ctypes.c_double)))
        return n, st

    def solve_step(self, undamaged=-1, max_iter=30, tol=1e-5, **kw):
        o = solver_opts(**kw)
        st = np.zeros(4)
        ok = self._chk(lib().lpmg_solve_step(self.h, int(undamaged), int(max_iter), float(tol),
                                             _p(o, ctypes.c_double), _p(st, ctypes.c_double)))
        return bool(ok), st

    def update_state_var(self):
        self._chk(lib().lpmg_update_state_var(self.h))

    def update_static_damage(self, undamaged=-1):
        return bool(self._chk(lib().lpmg_update_static_damage(self.h, int(undamaged))))

    def update_fatigue_damage(self, undamaged=-1, dNdt=1.0):
        return bool(self._chk(lib().lpmg_update_fatigue_damage(self.h, int(undamaged), float(dNdt))))

    def store_state(self):
        self._chk(lib().lpmg_store_state(self.h))

    def reset_state(self, reset_xyz=True):
        self._chk(lib().lpmg_reset_state(self.h, int(reset_xyz)))

    def search_nonlocal(self, ratio):
        self._chk(lib().lpmg_search_nonlocal(self.h, float(ratio)))

    def nonlocal_neighbors(self):
        """(nlptr, nlidx): Particle::neighbors_nonlocal of every particle, id order."""
        nlptr = np.zeros(self.N + 1, np.int32)
        self._chk(lib().lpmg_export_nonlocal(self.h, _p(nlptr, ctypes.c_int32), None))
        nlidx = np.zeros(max(int(nlptr[-1]), 1), np.int32)
        self._chk(lib().lpmg_export_nonlocal(self.h, None, _p(nlidx, ctypes.c_int32)))
        return nlptr, nlidx[:int(nlptr[-1])]

    def synchronize(self):
        self._chk(lib().lpmg_synchronize(self.h))

    def restart(self):
        """Back to the initial configuration, topology kept (see lpmg_restart)."""
        self._chk(lib().lpmg_restart(self.h))

    # -- instrumentation
    def profile(self, enable=True):
        self._chk(lib().lpmg_profile(self.h, int(enable)))

    def profile_read(self):
        out = np.zeros(4)
        self._chk(lib().lpmg_profile_read(self.h, _p(out, ctypes.c_double)))
        return dict(spmv_ms=out[0], spmv_launches=int(out[1]), spmv_bytes_fp64=out[2],
                    spmv_bytes_fp32=out[3])

    def multigrid_levels(self):
        out = np.zeros(32, dtype=np.int32)
        k = self._chk(lib().lpmg_multigrid_levels(self.h, _p(out, ctypes.c_int32), 32))
        return out[:k].tolist()

    def time_kernel(self, which, reps=20):
        out = np.zeros(2)
        self._chk(lib().lpmg_time_kernel(self.h, int(which), int(reps), _p(out, ctypes.c_double)))
        return out[0], out[1]

    # -- fields
    def get(self, name):
        f = FIELDS[name]
        n = lib().lpmg_field_len(self.h, f)
        out = np.zeros(n)
        self._chk(lib().lpmg_get_field(self.h, f, _p(out, ctypes.c_double)))
        return out

    def set(self, name, values):
        f = FIELDS[name]
        n = lib().lpmg_field_len(self.h, f)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if v.size != n:
            raise ValueError(f"{name}: expected {n} values, got {v.size}")
        self._chk(lib().lpmg_set_field(self.h, f, _p(v, ctypes.c_double)))


def unit_cell(lattice_type, radius):
    out = np.zeros(16)
    if lib().lpmg_unit_cell(int(lattice_type), float(radius), _p(out, ctypes.c_double)) != 0:
        raise LpmgError(_err())
    return out
