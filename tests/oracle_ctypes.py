"""ctypes bindings to the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg import this module.  It builds the oracle on first use when the
shared library is missing and a compiler is present.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# FVM_ORACLE_LIBRARY selects another build of the oracle (a copy with a
# planted mistake, to check that the tests catch it)
LIB_PATH = os.environ.get("FVM_ORACLE_LIBRARY", os.path.join(ORACLE_DIR, "liboracle.so"))
# the same oracle with its per-cell / per-edge loops on OpenMP threads (CPU
# baseline on all cores; the same bits as liboracle.so)
LIB_OMP_PATH = os.environ.get("FVM_ORACLE_OMP_LIBRARY", os.path.join(ORACLE_DIR, "liboracle_omp.so"))

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OrcFlow(C.Structure):
    _fields_ = [("mach", C.c_double), ("alpha_deg", C.c_double), ("gamma", C.c_double),
                ("p_in", C.c_double), ("t_in", C.c_double), ("cfl", C.c_double)]


class OrcOpts(C.Structure):
    _fields_ = [("order", C.c_int), ("smoothing", C.c_int), ("eps", C.c_double),
                ("global_dt", C.c_int)]


_libs = {}


def lib(path=LIB_PATH):
    L = _libs.get(path)
    if L is None:
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        L = C.CDLL(path)
        L.orc_mesh_build.argtypes = [C.c_int, C.c_int, _dp, _dp, _ip, C.c_int, C.POINTER(C.c_void_p)]
        L.orc_mesh_read_dat.argtypes = [C.c_char_p, C.POINTER(C.c_void_p), C.c_int]
        L.orc_mesh_free.argtypes = [C.c_void_p]
        L.orc_mesh_counts.argtypes = [C.c_void_p, _ip]
        L.orc_mesh_export.argtypes = [C.c_void_p] + [_ip] * 10 + [_dp] * 8
        L.orc_create.argtypes = [C.c_void_p, C.POINTER(OrcFlow), C.POINTER(OrcOpts)]
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_run.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp]
        L.orc_step_count.argtypes = [C.c_void_p]
        L.orc_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_forces.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_freestream.argtypes = [C.c_void_p, _dp]
        L.orc_eval_residual.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_smooth.argtypes = [C.c_void_p, _dp]
        L.orc_pta.argtypes = [C.c_double, C.c_double, _dp, _dp]
        L.orc_roe.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_far_ghost.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, _dp]
        L.orc_wall_ghost.argtypes = [_dp, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_limiter.argtypes = [_dp, _dp, _dp, _dp]
        L.orc_spectral_radius.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double]
        L.orc_spectral_radius.restype = C.c_double
        L.orc_time_steps.argtypes = [C.c_void_p, C.c_int]
        L.orc_time_steps.restype = C.c_double
        L.orc_threads.argtypes = []
        L.orc_threads.restype = C.c_int
        _libs[path] = L
    return L


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleMesh:
    """Mesh topology built by the oracle (1-based arrays, index 0 unused)."""

    def __init__(self, x, y, tri, literal=False, library=LIB_PATH):
        self.L = L = lib(library)
        x, y = f64(x), f64(y)
        tri = np.ascontiguousarray(tri, dtype=np.int32).reshape(-1)
        h = C.c_void_p()
        st = L.orc_mesh_build(len(x), len(tri) // 3, dptr(x), dptr(y), iptr(tri),
                                  int(literal), C.byref(h))
        if st != 0:
            raise ValueError(f"oracle mesh build failed with status {st}")
        self.h = h
        cnt = np.zeros(5, np.int32)
        L.orc_mesh_counts(h, iptr(cnt))
        self.nnode, self.ncell, self.nedge, self.nwall, self.nfar = (int(v) for v in cnt)
        C1, E1 = self.ncell + 1, self.nedge + 1
        self.celledge = np.zeros(3 * C1, np.int32)
        self.neicell = np.zeros(3 * C1, np.int32)
        self.clog = np.zeros(C1, np.int32)
        ints = {k: np.zeros(E1, np.int32) for k in
                ("elog", "left", "right", "farid", "wallid", "node1", "node2")}
        self.area, self.cx, self.cy = (np.zeros(C1) for _ in range(3))
        dbl = {k: np.zeros(E1) for k in ("vx", "vy", "vn", "midx", "midy")}
        L.orc_mesh_export(h, iptr(self.celledge), iptr(self.neicell), iptr(self.clog),
                              *(iptr(ints[k]) for k in ("elog", "left", "right", "farid", "wallid",
                                                        "node1", "node2")),
                              dptr(self.area), dptr(self.cx), dptr(self.cy),
                              *(dptr(dbl[k]) for k in ("vx", "vy", "vn", "midx", "midy")))
        for k, v in ints.items():
            setattr(self, k, v)
        for k, v in dbl.items():
            setattr(self, k, v)
        self.celledge = self.celledge.reshape(C1, 3)
        self.neicell = self.neicell.reshape(C1, 3)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_mesh_free(self.h)
            self.h = None


class OracleSolver:
    def __init__(self, mesh, mach, alpha_deg, cfl=1.5, gamma=1.4, p_in=101325.0, t_in=288.15,
                 order=2, smoothing=True, eps=0.8, global_dt=False):
        self.mesh = mesh  # keeps the mesh alive
        self.L = mesh.L  # the library that built the mesh
        self.flow = OrcFlow(mach, alpha_deg, gamma, p_in, t_in, cfl)
        self.opts = OrcOpts(order, int(bool(smoothing)), eps, int(bool(global_dt)))
        self.h = self.L.orc_create(mesh.h, C.byref(self.flow), C.byref(self.opts))
        if not self.h:
            raise MemoryError("oracle solver allocation failed")

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def run(self, maxstep, epsl=-1.0):
        hist = np.zeros(maxstep + 1)
        n = self.L.orc_run(self.h, maxstep, epsl, dptr(hist))
        steps = self.L.orc_step_count(self.h)
        return hist[: steps + 1], n

    def state(self):
        n = self.mesh.ncell
        out = [np.zeros(n) for _ in range(4)]
        self.L.orc_get_state(self.h, *(dptr(a) for a in out))
        return np.stack(out, axis=1)

    def set_state(self, W):
        W = f64(W)
        cols = [f64(W[:, k]) for k in range(4)]
        self.L.orc_set_state(self.h, *(dptr(a) for a in cols))

    def forces(self):
        cl, cd = C.c_double(), C.c_double()
        self.L.orc_forces(self.h, C.byref(cl), C.byref(cd))
        return cl.value, cd.value

    def freestream(self):
        out = np.zeros(7)
        self.L.orc_freestream(self.h, dptr(out))
        return dict(zip(("UI", "VI", "AIN", "VIN", "ROUIN", "PIS", "rhoE"), out))

    def eval_residual(self):
        n = self.mesh.ncell
        res, sr, dt = np.zeros(4 * n), np.zeros(n), np.zeros(n)
        grad, umax, umin = np.zeros(8 * n), np.zeros(4 * n), np.zeros(4 * n)
        self.L.orc_eval_residual(self.h, dptr(res), dptr(sr), dptr(dt), dptr(grad), dptr(umax),
                                dptr(umin))
        return dict(res=res.reshape(n, 4), sr=sr, dt=dt, grad=grad.reshape(n, 2, 4),
                    umax=umax.reshape(n, 4), umin=umin.reshape(n, 4))

    def smooth(self, res):
        r = f64(res).reshape(-1).copy()
        self.L.orc_smooth(self.h, dptr(r))
        return r.reshape(-1, 4)

    def time_steps(self, steps):
        return self.L.orc_time_steps(self.h, steps)


def read_dat(path):
    """Reads a mesh in the reference's .dat format (optionally gzip-compressed)."""
    import gzip
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rt") as f:
        tok = f.read().split()
    nnode, ncell = int(tok[0]), int(tok[1])
    vals = tok[2:]
    x = np.array(vals[:nnode], dtype=np.float64)
    y = np.array(vals[nnode:2 * nnode], dtype=np.float64)
    tri = np.array(vals[2 * nnode:2 * nnode + 3 * ncell], dtype=np.int32).reshape(ncell, 3)
    return x, y, tri
