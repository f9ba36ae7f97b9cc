"""ctypes binding of the CPU FP64 oracle (oracle/liblpm_oracle.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it as the checker.  Nothing in the product imports it.
The C functions restate LPM-CPP (see the citations in lpm_oracle.cpp).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblpm_oracle.so")
_lib = None

# field ids shared with the product library's get/set_field
FIELDS = dict(XYZ=0, XYZ_INITIAL=1, XYZ_LAST=2, PIN=3, PEX=4, STRESS=5, DAMAGE=6,
              DAMAGE_VISUAL=7, STATE_VAR=8, RESIDUAL=9, BDAMAGE=10, DLP=11, BFORCE=12,
              DIS=13, CS=14, CONSTRAINT=15, DDOT_LOCAL=16, DDOT_NONLOCAL=17, DLE_TOTAL=18,
              CS_SUM=19, BSTRAIN=20, DLE=21, DAMAGE_LAST=22, TYPE=23, TDLE_TOTAL=24)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, i32p, i64p, dp = (ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                              ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double))
        L.lpmo_last_error.restype = ctypes.c_char_p
        L.lpmo_create.restype = vp
        L.lpmo_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int]
        L.lpmo_create_with_bonds.restype = vp
        L.lpmo_create_with_bonds.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp,
                                             ctypes.c_int64, i32p, i32p, i32p]
        L.lpmo_destroy.argtypes = [vp]
        L.lpmo_field_len.restype = ctypes.c_int64
        L.lpmo_field_len.argtypes = [vp, ctypes.c_int]
        for name in ("lpmo_sizes", "lpmo_export_bonds", "lpmo_export_conns", "lpmo_export_csr",
                     "lpmo_set_types", "lpmo_set_material", "lpmo_update_geometry",
                     "lpmo_update_force_state", "lpmo_forward", "lpmo_update_rr", "lpmo_assemble",
                     "lpmo_spmv", "lpmo_solve_linear", "lpmo_newton", "lpmo_update_state_var",
                     "lpmo_update_static_damage", "lpmo_update_fatigue_damage", "lpmo_store_state",
                     "lpmo_reset_state",
                     "lpmo_search_nonlocal", "lpmo_export_nonlocal", "lpmo_define_group", "lpmo_disp_bc", "lpmo_force_bc",
                     "lpmo_solve_step", "lpmo_get_field", "lpmo_set_field", "lpmo_unit_cell"):
            getattr(L, name).restype = ctypes.c_int
        L.lpmo_set_material.argtypes = [vp, ctypes.c_int, i32p, ctypes.c_int, dp]
        L.lpmo_newton.argtypes = [vp, ctypes.c_int, ctypes.c_double, dp, dp]
        L.lpmo_solve_step.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp]
        L.lpmo_disp_bc.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.lpmo_force_bc.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double]
        L.lpmo_search_nonlocal.argtypes = [vp, ctypes.c_double]
        L.lpmo_export_nonlocal.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        L.lpmo_update_fatigue_damage.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.lpmo_unit_cell.argtypes = [ctypes.c_int, ctypes.c_double, dp]
        L.lpmo_set_threads.argtypes = [ctypes.c_int]
        L.lpmo_ct_search_neighbor.restype = ctypes.c_int
        L.lpmo_ct_search_neighbor.argtypes = [ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                              ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int64,
                                              ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(ctypes.c_int32)]
        L.lpmo_ct_valid.restype = ctypes.c_int
        L.lpmo_ct_valid.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, dp,
                                    ctypes.POINTER(ctypes.c_int32)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def unit_cell(lattice, radius):
    out = np.zeros(16)
    if lib().lpmo_unit_cell(int(lattice), float(radius), _p(out, ctypes.c_double)) != 0:
        raise ValueError(lib().lpmo_last_error().decode())
    return out


def set_threads(n):
    lib().lpmo_set_threads(int(n))


def _regions(region):
    circ = np.ascontiguousarray(np.array(region.circles, dtype=np.float64).reshape(-1, 3))
    notch = np.ascontiguousarray(np.array(region.notches, dtype=np.float64).reshape(-1, 3))
    return circ, notch


def ct_search_neighbor(xyz, cut, nneighbors, region):
    """geo1's searchNeighbor by brute force: (layer, p1, p2)."""
    L = lib()
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    n = xyz.shape[0]
    circ, notch = _regions(region)
    cap = n * nneighbors
    out = [np.zeros(cap, dtype=np.int32) for _ in range(3)]
    m = L.lpmo_ct_search_neighbor(n, _p(xyz, ctypes.c_double), float(cut[0]), float(cut[1]), int(nneighbors),
                                  len(circ), _p(circ, ctypes.c_double), len(notch), _p(notch, ctypes.c_double),
                                  cap, *(_p(o, ctypes.c_int32) for o in out))
    if m < 0:
        raise RuntimeError(L.lpmo_last_error().decode())
    return tuple(o[:m] for o in out)


def ct_valid(xyz, region):
    """geo1's isValid per particle."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    circ, notch = _regions(region)
    v = np.zeros(xyz.shape[0], dtype=np.int32)
    lib().lpmo_ct_valid(xyz.shape[0], _p(xyz, ctypes.c_double), len(circ), _p(circ, ctypes.c_double),
                        len(notch), _p(notch, ctypes.c_double), _p(v, ctypes.c_int32))
    return v.astype(bool)


class Oracle:
    """One particle system (an LPM-CPP Assembly + Stiffness + Solver) on the CPU."""

    def __init__(self, lattice, radius, ptype, xyz, bonds=None, brute=False):
        L = lib()
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.n = xyz.shape[0]
        if bonds is None:
            h = L.lpmo_create(int(lattice), float(radius), int(ptype), self.n,
                              _p(xyz, ctypes.c_double), int(brute))
        else:
            p1, p2, layer = (np.ascontiguousarray(b, dtype=np.int32) for b in bonds)
            h = L.lpmo_create_with_bonds(int(lattice), float(radius), int(ptype), self.n,
                                         _p(xyz, ctypes.c_double), len(p1), _p(p1, ctypes.c_int32),
                                         _p(p2, ctypes.c_int32), _p(layer, ctypes.c_int32))
        if not h:
            raise RuntimeError("oracle create failed: " + L.lpmo_last_error().decode())
        self.h = ctypes.c_void_p(h)
        s = np.zeros(8, dtype=np.int64)
        L.lpmo_sizes(self.h, _p(s, ctypes.c_int64))
        self.N, self.dim, self.nbond, self.nnz, self.nnzb, self.max_nb, self.max_nconn, self.ptype = (
            int(v) for v in s)

    def close(self):
        if self.h:
            lib().lpmo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc < 0:
            raise RuntimeError(lib().lpmo_last_error().decode())
        return rc

    # -- structure
    def bonds(self):
        bptr = np.zeros(self.N + 1, np.int32)
        nb0 = np.zeros(self.N, np.int32)
        p2 = np.zeros(self.nbond, np.int32)
        lib().lpmo_export_bonds(self.h, _p(bptr, ctypes.c_int32), _p(nb0, ctypes.c_int32),
                                _p(p2, ctypes.c_int32))
        return bptr, nb0, p2

    def conns(self):
        cptr = np.zeros(self.N + 1, np.int32)
        conns = np.zeros(self.nnzb, np.int32)
        ncl = np.zeros(self.N, np.int32)
        lib().lpmo_export_conns(self.h, _p(cptr, ctypes.c_int32), _p(conns, ctypes.c_int32),
                                _p(ncl, ctypes.c_int32))
        return cptr, conns, ncl

    def csr(self):
        Kp = np.zeros(self.N + 1, np.int64)
        IK = np.zeros(self.dim * self.N + 1, np.int64)
        JK = np.zeros(self.nnz, np.int64)
        K = np.zeros(self.nnz)
        lib().lpmo_export_csr(self.h, _p(Kp, ctypes.c_int64), _p(IK, ctypes.c_int64),
                              _p(JK, ctypes.c_int64), _p(K, ctypes.c_double))
        return Kp, IK, JK, K

    # -- setup
    def set_types(self, t):
        t = np.ascontiguousarray(t, dtype=np.int32)
        lib().lpmo_set_types(self.h, _p(t, ctypes.c_int32))

    def set_material(self, model, params, idx=None):
        p = np.zeros(12)
        p[:len(params)] = params
        if idx is None:
            self._chk(lib().lpmo_set_material(self.h, 0, None, int(model), _p(p, ctypes.c_double)))
        else:
            idx = np.ascontiguousarray(idx, dtype=np.int32)
            self._chk(lib().lpmo_set_material(self.h, len(idx), _p(idx, ctypes.c_int32), int(model),
                                              _p(p, ctypes.c_double)))

    def define_group(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        return self._chk(lib().lpmo_define_group(self.h, len(idx), _p(idx, ctypes.c_int32)))

    def disp_bc(self, gid, mode, flag, step):
        self._chk(lib().lpmo_disp_bc(self.h, gid, mode, ord(flag), float(step)))

    def force_bc(self, gid, mode, fx, fy, fz):
        self._chk(lib().lpmo_force_bc(self.h, gid, mode, float(fx), float(fy), float(fz)))

    # -- hot path
    def forward(self):
        lib().lpmo_forward(self.h)

    def update_geometry(self):
        lib().lpmo_update_geometry(self.h)

    def update_force_state(self):
        lib().lpmo_update_force_state(self.h)

    def update_rr(self):
        out = np.zeros(2)
        lib().lpmo_update_rr(self.h, _p(out, ctypes.c_double))
        return out

    def assemble(self):
        lib().lpmo_assemble(self.h)

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().lpmo_spmv(self.h, _p(x, ctypes.c_double), _p(y, ctypes.c_double))
        return y

    @staticmethod
    def _opts(precond=0, rtol=1e-12, atol=1e-12, maxit=0):
        return np.array([precond, rtol, atol, maxit], dtype=np.float64)

    def solve_linear(self, **kw):
        o = self._opts(**kw)
        st = np.zeros(4)
        lib().lpmo_solve_linear(self.h, _p(o, ctypes.c_double), _p(st, ctypes.c_double))
        return st

    def newton(self, max_iter=30, tol=1e-5, **kw):
        o = self._opts(**kw)
        st = np.zeros(4)
        n = lib().lpmo_newton(self.h, int(max_iter), float(tol), _p(o, ctypes.c_double),
                              _p(st, ctypes.c_double))
        return n, st

    def solve_step(self, undamaged=-1, max_iter=30, tol=1e-5, **kw):
        o = self._opts(**kw)
        st = np.zeros(4)
        ok = lib().lpmo_solve_step(self.h, int(undamaged), int(max_iter), float(tol),
                                   _p(o, ctypes.c_double), _p(st, ctypes.c_double))
        return bool(ok), st

    def update_state_var(self):
        lib().lpmo_update_state_var(self.h)

    def update_static_damage(self, undamaged=-1):
        return bool(lib().lpmo_update_static_damage(self.h, int(undamaged)))

    def update_fatigue_damage(self, undamaged=-1, dNdt=1.0):
        return bool(lib().lpmo_update_fatigue_damage(self.h, int(undamaged), float(dNdt)))

    def store_state(self):
        lib().lpmo_store_state(self.h)

    def reset_state(self, reset_xyz=True):
        lib().lpmo_reset_state(self.h, int(reset_xyz))

    def search_nonlocal(self, ratio):
        lib().lpmo_search_nonlocal(self.h, float(ratio))

    def nonlocal_neighbors(self):
        """(nlptr, nlidx) of the last search: every particle's list in id order."""
        nlptr = np.zeros(self.N + 1, np.int32)
        self._chk(lib().lpmo_export_nonlocal(self.h, _p(nlptr, ctypes.c_int32), None))
        nlidx = np.zeros(max(int(nlptr[-1]), 1), np.int32)
        self._chk(lib().lpmo_export_nonlocal(self.h, None, _p(nlidx, ctypes.c_int32)))
        return nlptr, nlidx[:int(nlptr[-1])]

    # -- fields
    def get(self, name):
        f = FIELDS[name]
        n = lib().lpmo_field_len(self.h, f)
        out = np.zeros(n)
        self._chk(lib().lpmo_get_field(self.h, f, _p(out, ctypes.c_double)))
        return out

    def set(self, name, values):
        f = FIELDS[name]
        n = lib().lpmo_field_len(self.h, f)
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        assert v.size == n, (name, v.size, n)
        self._chk(lib().lpmo_set_field(self.h, f, _p(v, ctypes.c_double)))
