"""Python binding of the serving runtime (lpmg_server_* in include/lpm_gpu.h).

A `Server` runs many load-stepped quasi-static solves at once on the GPU:
FIFO admission by worker slot and device-memory budget, one device-resident
handle per request, and (batching on) the CG iterations of every request in a
linear solve fused into shared launches.  `request_from_problem` turns a
`problems.Problem` into the C request struct; the arrays it points to are kept
alive by the returned `Request`.
"""
import ctypes

import numpy as np

from . import FIELDS, LpmgError, _err, _p, lib

RS_FIELDS = ("status", "steps", "halvings", "newton", "cg", "passes", "queue_ms", "run_ms", "total_ms",
             "bytes_est", "bytes")


class Bc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("group", ctypes.c_int), ("mode", ctypes.c_int), ("flag", ctypes.c_int),
                ("value", ctypes.c_double * 3)]


class OutField(ctypes.Structure):
    _fields_ = [("field", ctypes.c_int), ("capacity", ctypes.c_int64), ("data", ctypes.POINTER(ctypes.c_double))]


_i32p = ctypes.POINTER(ctypes.c_int32)
_dp = ctypes.POINTER(ctypes.c_double)


class RequestStruct(ctypes.Structure):
    _fields_ = [("lattice", ctypes.c_int), ("radius", ctypes.c_double), ("ptype", ctypes.c_int),
                ("n", ctypes.c_int), ("xyz", _dp), ("nbond", ctypes.c_int64), ("p1", _i32p), ("p2", _i32p),
                ("layer", _i32p), ("type", _i32p), ("model", ctypes.c_int), ("n_params", ctypes.c_int),
                ("params", _dp), ("nonlocal_ratio", ctypes.c_double), ("prepare_state", ctypes.c_int),
                ("n_groups", ctypes.c_int), ("group_size", _i32p), ("group_idx", _i32p), ("n_bc", ctypes.c_int),
                ("bc", ctypes.POINTER(Bc)), ("n_steps", ctypes.c_int), ("undamaged", ctypes.c_int),
                ("max_iter", ctypes.c_int), ("tol", ctypes.c_double), ("opts", ctypes.c_double * 8),
                ("max_halvings", ctypes.c_int), ("n_out", ctypes.c_int), ("out", ctypes.POINTER(OutField)),
                ("step_stats", _dp), ("step_stats_cap", ctypes.c_int)]


_bound = False


def _bind():
    global _bound
    if _bound:
        return lib()
    L = lib()
    vp = ctypes.c_void_p
    L.lpmg_request_bytes.restype = ctypes.c_int64
    L.lpmg_request_bytes.argtypes = [ctypes.POINTER(RequestStruct)]
    L.lpmg_server_create.restype = ctypes.c_int
    L.lpmg_server_create.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(vp)]
    L.lpmg_server_submit.restype = ctypes.c_int
    L.lpmg_server_submit.argtypes = [vp, ctypes.POINTER(RequestStruct), ctypes.POINTER(ctypes.c_int64)]
    L.lpmg_server_wait.restype = ctypes.c_int
    L.lpmg_server_wait.argtypes = [vp, ctypes.c_int64, ctypes.c_double, _dp]
    L.lpmg_server_info.restype = ctypes.c_int
    L.lpmg_server_info.argtypes = [vp, _dp]
    L.lpmg_server_destroy.restype = None
    L.lpmg_server_destroy.argtypes = [vp]
    _bound = True
    return L


class Request:
    """One solve: the C struct, the arrays behind it and the output buffers."""

    def __init__(self, prob, opts, out_fields=("XYZ",), nonlocal_ratio=0.0, prepare_state=False,
                 max_halvings=8, sizes=None):
        xyz = np.ascontiguousarray(prob.xyz, dtype=np.float64).reshape(-1, 3)
        n = xyz.shape[0]
        self.prob = prob
        self._keep = [xyz]
        r = RequestStruct()
        r.lattice, r.radius, r.ptype, r.n = int(prob.lattice), float(prob.radius), int(prob.ptype), n
        r.xyz = _p(xyz, ctypes.c_double)
        bonds = prob.extra.get("bonds")
        nbond = 0
        if bonds is not None:
            p1, p2, layer = (np.ascontiguousarray(b, dtype=np.int32) for b in bonds)
            self._keep += [p1, p2, layer]
            nbond = len(p1)
            r.nbond, r.p1, r.p2, r.layer = nbond, _p(p1, ctypes.c_int32), _p(p2, ctypes.c_int32), \
                _p(layer, ctypes.c_int32)
        if prob.types is not None:
            t = np.ascontiguousarray(prob.types, dtype=np.int32)
            self._keep.append(t)
            r.type = _p(t, ctypes.c_int32)
        model, params = prob.material
        params = np.ascontiguousarray(params, dtype=np.float64)
        self._keep.append(params)
        r.model, r.n_params, r.params = int(model), len(params), _p(params, ctypes.c_double)
        r.nonlocal_ratio = float(nonlocal_ratio)
        r.prepare_state = int(bool(prepare_state))
        names = list(prob.groups)
        gsize = np.array([len(prob.groups[g]) for g in names], dtype=np.int32)
        gidx = (np.concatenate([np.asarray(prob.groups[g], dtype=np.int32) for g in names]) if names
                else np.zeros(0, dtype=np.int32))
        gidx = np.ascontiguousarray(gidx, dtype=np.int32)
        self._keep += [gsize, gidx]
        r.n_groups, r.group_size, r.group_idx = len(names), _p(gsize, ctypes.c_int32), _p(gidx, ctypes.c_int32)
        steps = prob.steps
        if any(s != steps[0] for s in steps):
            raise ValueError("the serving runtime applies one boundary-condition template per step")
        tmpl = steps[0] if steps else []
        bcs = (Bc * max(1, len(tmpl)))()
        for k, bc in enumerate(tmpl):
            if bc[0] == "disp":
                _, g, mode, flag, d = bc
                bcs[k].kind, bcs[k].flag = 0, ord(flag) if isinstance(flag, str) else int(flag)
                bcs[k].value[0] = d
            else:
                _, g, mode, fx, fy, fz = bc
                bcs[k].kind = 1
                bcs[k].value[0], bcs[k].value[1], bcs[k].value[2] = fx, fy, fz
            bcs[k].group, bcs[k].mode = names.index(g), int(mode)
        self._keep.append(bcs)
        r.n_bc, r.bc, r.n_steps = len(tmpl), bcs, len(steps)
        r.undamaged, r.max_iter, r.tol = int(prob.undamaged), int(prob.max_iter), float(prob.tol)
        for k, v in enumerate(np.asarray(opts, dtype=np.float64)[:8]):
            r.opts[k] = v
        r.max_halvings = int(max_halvings)
        sizes = sizes or {}
        self.out = {}
        outs = (OutField * max(1, len(out_fields)))()
        for k, name in enumerate(out_fields):
            length = sizes.get(name)
            if length is None:
                length = _field_len(name, n, nbond, 2 if prob.lattice in (0, 1) else 3)
            buf = np.zeros(length, dtype=np.float64)
            self.out[name] = buf
            outs[k].field, outs[k].capacity, outs[k].data = FIELDS[name], length, _p(buf, ctypes.c_double)
        self._keep.append(outs)
        r.n_out, r.out = len(out_fields), outs
        cap = 4 * max(1, len(steps)) + 64
        self.step_stats = np.zeros(3 * cap, dtype=np.float64)
        r.step_stats, r.step_stats_cap = _p(self.step_stats, ctypes.c_double), cap
        self.struct = r

    def bytes(self):
        return int(_bind().lpmg_request_bytes(ctypes.byref(self.struct)))


def _field_len(name, n, nbond, dim):
    per = dict(XYZ=3, XYZ_INITIAL=3, XYZ_LAST=3, PIN=3, PEX=3, CONSTRAINT=3, STRESS=6, CS_SUM=6, DAMAGE=1,
               DAMAGE_VISUAL=1, DDOT_LOCAL=1, DDOT_NONLOCAL=1, DAMAGE_LAST=1, TYPE=1, STATE_VAR=7,
               DLE_TOTAL=2, TDLE_TOTAL=2)
    if name in per:
        return per[name] * n
    if name == "RESIDUAL":
        return dim * n
    if name in ("BDAMAGE", "DLP", "BFORCE", "DIS", "BSTRAIN", "DLE", "CS"):
        if nbond <= 0:
            raise ValueError(f"{name} needs a bond count: pass sizes={{'{name}': ...}}")
        return nbond * (3 if name == "CS" else 1)
    raise KeyError(name)


class Server:
    """lpmg_server: worker slots, a device-memory budget and fused CG launches.

    batching: True fuses every request's CG, False none; an int fuses the
    requests of at most that many particles (larger ones use their own stream)."""

    def __init__(self, max_active=8, mem_budget=0, batching=True):
        L = _bind()
        h = ctypes.c_void_p()
        fuse = -1 if batching is True else 0 if batching is False else int(batching)
        rc = L.lpmg_server_create(int(max_active), int(mem_budget), fuse, ctypes.byref(h))
        if rc != 0:
            raise LpmgError(f"lpmg_server_create failed ({rc}): {_err()}")
        self.h = h
        self.pending = {}

    def submit(self, req):
        t = ctypes.c_int64()
        rc = _bind().lpmg_server_submit(self.h, ctypes.byref(req.struct), ctypes.byref(t))
        if rc != 0:
            raise LpmgError(f"lpmg_server_submit failed ({rc}): {_err()}")
        self.pending[t.value] = req
        return t.value

    def wait(self, ticket, timeout=-1.0):
        """Returns (status, stats dict) once finished, or None on timeout."""
        st = np.zeros(len(RS_FIELDS), dtype=np.float64)
        rc = _bind().lpmg_server_wait(self.h, int(ticket), float(timeout), _p(st, ctypes.c_double))
        if rc == 1:
            return None
        stats = dict(zip(RS_FIELDS, st.tolist()))
        if rc != 0:
            stats["error"] = _err()
        self.pending.pop(ticket, None)
        return rc, stats

    def info(self):
        out = np.zeros(11, dtype=np.float64)
        if _bind().lpmg_server_info(self.h, _p(out, ctypes.c_double)) != 0:
            raise LpmgError(_err())
        keys = ("queued", "running", "reserved", "budget", "peak_running", "peak_reserved", "finished",
                "chunks", "system_chunks", "max_batch", "graphs")
        return dict(zip(keys, out.tolist()))

    def close(self):
        if getattr(self, "h", None):
            _bind().lpmg_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
