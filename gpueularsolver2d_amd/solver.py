"""Python binding of libfvm_hip.so through its C ABI (include/fvm_abi.h).

The reference's entry points (gpu_fvm.cu:27-68) map to:
  inputMesh            -> meshgen.read_dat / fvm_mesh_desc
  Initialization       -> done inside fvm_create from the flow parameters
  CalculationMeshGeo   -> mesh_topology (host-only) / done inside fvm_create
  ROE_SOLVER           -> Solver.run (fvm_run)
Batches of flow cases on several meshes go through Batch (fvm_batch_*).

There is no fallback: if libfvm_hip.so is missing or no GPU is present the
calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FVM_LIBRARY selects another build of the library (e.g. a variant under test)
LIB_PATH = os.environ.get("FVM_LIBRARY", os.path.join(HERE, "libfvm_hip.so"))

FVM_FP64, FVM_FP32 = 0, 1
FVM_KERNELS_TILED, FVM_KERNELS_PLAIN, FVM_KERNELS_SPLIT = 0, 1, 2

STATUS = {0: "FVM_OK", 1: "FVM_ERR_ARGUMENT", 2: "FVM_ERR_MESH", 3: "FVM_ERR_AREA",
          4: "FVM_ERR_HIP", 5: "FVM_ERR_NAN", 6: "FVM_ERR_NOT_CONVERGED",
          7: "FVM_ERR_NO_DEVICE", 8: "FVM_ERR_ALLOC"}

# Every function the header declares; tests check the library exports them.
EXPORTS = ("fvm_default_opts", "fvm_status_string", "fvm_last_error", "fvm_mesh_topology",
           "fvm_create", "fvm_run", "fvm_get_state", "fvm_set_state", "fvm_get_forces",
           "fvm_get_timing", "fvm_destroy", "fvm_batch_create", "fvm_batch_run",
           "fvm_batch_get_state", "fvm_batch_set_state", "fvm_batch_get_forces",
           "fvm_batch_get_timing", "fvm_batch_get_config", "fvm_get_config",
           "fvm_batch_kernel_times", "fvm_kernel_times", "fvm_batch_destroy", "fvm_run_noalloc")


class FvmError(RuntimeError):
    def __init__(self, status, msg):
        self.status = status
        self.name = STATUS.get(status, str(status))
        super().__init__(f"{self.name}: {msg}")


class MeshDesc(C.Structure):
    _fields_ = [("nnode", C.c_int32), ("ncell", C.c_int32), ("x", C.POINTER(C.c_double)),
                ("y", C.POINTER(C.c_double)), ("tri", C.POINTER(C.c_int32))]


class FlowParams(C.Structure):
    _fields_ = [("mach", C.c_double), ("alpha_deg", C.c_double), ("gamma", C.c_double),
                ("p_in", C.c_double), ("t_in", C.c_double), ("cfl", C.c_double)]


class SchemeOpts(C.Structure):
    _fields_ = [("order", C.c_int32), ("smoothing", C.c_int32), ("eps_smooth", C.c_double),
                ("global_dt", C.c_int32), ("precision", C.c_int32), ("kernels", C.c_int32),
                ("tile_cells", C.c_int32), ("threads", C.c_int32),
                ("steps_per_graph", C.c_int32), ("device", C.c_int32)]


_lib = None
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


def load_library(path=LIB_PATH):
    """Loads libfvm_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m gpueularsolver2d_amd.build` "
                           "(the solver has no CPU fallback)")
    L = C.CDLL(path)
    L.fvm_default_opts.argtypes = [C.POINTER(SchemeOpts)]
    L.fvm_default_opts.restype = None
    L.fvm_status_string.argtypes = [C.c_int]
    L.fvm_status_string.restype = C.c_char_p
    L.fvm_last_error.argtypes = []
    L.fvm_last_error.restype = C.c_char_p
    L.fvm_mesh_topology.argtypes = [C.POINTER(MeshDesc), _ip] + [_ip] * 10
    L.fvm_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(FlowParams), C.POINTER(SchemeOpts),
                             C.POINTER(C.c_void_p)]
    L.fvm_run.argtypes = [C.c_void_p, C.c_int32, C.c_double, _dp, _ip]
    L.fvm_run_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_double, _dp, _ip]
    L.fvm_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.fvm_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.fvm_get_forces.argtypes = [C.c_void_p, _dp, _dp]
    L.fvm_get_timing.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int64)]
    L.fvm_destroy.argtypes = [C.c_void_p]
    L.fvm_destroy.restype = None
    L.fvm_batch_create.argtypes = [C.c_int32, C.POINTER(MeshDesc), C.c_int32, _ip,
                                   C.POINTER(FlowParams), C.POINTER(SchemeOpts),
                                   C.POINTER(C.c_void_p)]
    L.fvm_batch_run.argtypes = [C.c_void_p, C.c_int32, C.c_double, _dp, _ip, _ip]
    L.fvm_batch_get_state.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp]
    L.fvm_batch_set_state.argtypes = [C.c_void_p, C.c_int32, _dp, _dp, _dp, _dp]
    L.fvm_batch_get_forces.argtypes = [C.c_void_p, _dp, _dp]
    L.fvm_batch_get_timing.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_int64)]
    L.fvm_batch_get_config.argtypes = [C.c_void_p, _ip]
    L.fvm_get_config.argtypes = [C.c_void_p, _ip]
    L.fvm_batch_kernel_times.argtypes = [C.c_void_p, C.c_int32, _dp]
    L.fvm_kernel_times.argtypes = [C.c_void_p, C.c_int32, _dp]
    L.fvm_batch_destroy.argtypes = [C.c_void_p]
    L.fvm_batch_destroy.restype = None
    for name in ("fvm_mesh_topology", "fvm_create", "fvm_run", "fvm_run_noalloc", "fvm_get_state",
                 "fvm_set_state",
                 "fvm_get_forces", "fvm_get_timing", "fvm_batch_create", "fvm_batch_run",
                 "fvm_batch_get_state", "fvm_batch_set_state", "fvm_batch_get_forces",
                 "fvm_batch_get_timing", "fvm_batch_get_config", "fvm_get_config",
                 "fvm_batch_kernel_times", "fvm_kernel_times"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(st):
    if st != 0:
        raise FvmError(st, load_library().fvm_last_error().decode())


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _iptr(a):
    return a.ctypes.data_as(_ip)


class _MeshHold:
    """Keeps the numpy arrays behind an fvm_mesh_desc alive."""

    def __init__(self, mesh):
        self.x = np.ascontiguousarray(mesh.x, dtype=np.float64)
        self.y = np.ascontiguousarray(mesh.y, dtype=np.float64)
        self.tri = np.ascontiguousarray(mesh.tri, dtype=np.int32).reshape(-1)
        self.desc = MeshDesc(len(self.x), len(self.tri) // 3, _dptr(self.x), _dptr(self.y),
                             _iptr(self.tri))


def scheme_opts(order=2, smoothing=True, eps_smooth=0.8, global_dt=False, precision="fp64",
                kernels="tiled", tile_cells=0, threads=0, steps_per_graph=0, device=0):
    o = SchemeOpts()
    load_library().fvm_default_opts(C.byref(o))
    o.order = order
    o.smoothing = int(bool(smoothing))
    o.eps_smooth = eps_smooth
    o.global_dt = int(bool(global_dt))
    o.precision = {"fp64": FVM_FP64, "fp32": FVM_FP32}[precision]
    o.kernels = {"tiled": FVM_KERNELS_TILED, "plain": FVM_KERNELS_PLAIN,
                 "split": FVM_KERNELS_SPLIT}[kernels]
    o.tile_cells = tile_cells
    o.threads = threads
    o.steps_per_graph = steps_per_graph
    o.device = device
    return o


def flow_params(mach, alpha_deg, cfl=1.5, gamma=1.4, p_in=101325.0, t_in=288.15):
    return FlowParams(mach, alpha_deg, gamma, p_in, t_in, cfl)


def mesh_topology(mesh):
    """The reference's topology (CalculationMeshGeo + changeEdge), 1-based ids.
    Host-only; needs no GPU."""
    L = load_library()
    h = _MeshHold(mesh)
    counts = np.zeros(5, np.int32)
    none = C.cast(None, _ip)
    _check(L.fvm_mesh_topology(C.byref(h.desc), _iptr(counts), *([none] * 10)))
    nnode, ncell, nedge, nwall, nfar = (int(v) for v in counts)
    cell = {k: np.zeros(3 * ncell, np.int32) for k in ("celledge", "neicell")}
    clog = np.zeros(ncell, np.int32)
    edge = {k: np.zeros(nedge, np.int32) for k in ("elog", "left", "right", "farid", "wallid",
                                                    "node1", "node2")}
    _check(L.fvm_mesh_topology(C.byref(h.desc), _iptr(counts), _iptr(cell["celledge"]),
                               _iptr(cell["neicell"]), _iptr(clog),
                               *(_iptr(edge[k]) for k in ("elog", "left", "right", "farid",
                                                          "wallid", "node1", "node2"))))
    out = dict(nnode=nnode, ncell=ncell, nedge=nedge, nwall=nwall, nfar=nfar, clog=clog,
               celledge=cell["celledge"].reshape(ncell, 3), neicell=cell["neicell"].reshape(ncell, 3))
    out.update(edge)
    return out


class Solver:
    """One flow case on one mesh (fvm_create / fvm_run / fvm_get_*)."""

    def __init__(self, mesh, mach, alpha_deg, cfl=1.5, gamma=1.4, p_in=101325.0, t_in=288.15,
                 **opts):
        L = load_library()
        self._mesh = _MeshHold(mesh)
        self.ncell = self._mesh.desc.ncell
        self.flow = FlowParams(mach, alpha_deg, gamma, p_in, t_in, cfl)
        self.opts = scheme_opts(**opts)
        h = C.c_void_p()
        _check(L.fvm_create(C.byref(self._mesh.desc), C.byref(self.flow), C.byref(self.opts),
                            C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load_library().fvm_destroy(self.h)
            self.h = None

    __del__ = close

    def run(self, max_steps, epsl=-1.0, check=True):
        """Returns (residual history Resd[0..steps], steps done, status)."""
        hist = np.zeros(max_steps + 1)
        steps = C.c_int32()
        st = load_library().fvm_run(self.h, max_steps, epsl, _dptr(hist), C.byref(steps))
        if check and st not in (0,):
            _check(st)
        return hist[: steps.value + 1], steps.value, st

    def run_noalloc(self, max_steps, epsl=-1.0, check=True):
        """run() through fvm_run_noalloc: no device allocation, scratch held by
        the library (at most FVM_NOALLOC_SCRATCH_BYTES per device)."""
        hist = np.zeros(max_steps + 1)
        steps = C.c_int32()
        st = load_library().fvm_run_noalloc(self.h, max_steps, epsl, _dptr(hist), C.byref(steps))
        if check and st not in (0,):
            _check(st)
        return hist[: steps.value + 1], steps.value, st

    def state(self):
        out = [np.zeros(self.ncell) for _ in range(4)]
        _check(load_library().fvm_get_state(self.h, *(_dptr(a) for a in out)))
        return np.stack(out, axis=1)

    def set_state(self, W):
        cols = [np.ascontiguousarray(W[:, k], dtype=np.float64) for k in range(4)]
        _check(load_library().fvm_set_state(self.h, *(_dptr(a) for a in cols)))

    def forces(self):
        cl, cd = C.c_double(), C.c_double()
        _check(load_library().fvm_get_forces(self.h, C.byref(cl), C.byref(cd)))
        return cl.value, cd.value

    def timing(self):
        ms, csu = C.c_double(), C.c_int64()
        _check(load_library().fvm_get_timing(self.h, C.byref(ms), C.byref(csu)))
        return ms.value, csu.value

    def config(self):
        cfg = np.zeros(6, np.int32)
        _check(load_library().fvm_get_config(self.h, _iptr(cfg)))
        return dict(tile_cells=int(cfg[0]), threads=int(cfg[1]), steps_per_graph=int(cfg[2]),
                    tiles=int(cfg[3]), blocks_per_cu=int(cfg[4]), cus=int(cfg[5]))

    def kernel_times(self, reps=20):
        """Average ms per launch: (stage_tile, 0) with one kernel per stage,
        (flux_tile, update_tile) with two."""
        ms = np.zeros(2)
        _check(load_library().fvm_kernel_times(self.h, reps, _dptr(ms)))
        return float(ms[0]), float(ms[1])


class Batch:
    """Many flow cases on several meshes in every launch (fvm_batch_*)."""

    def __init__(self, meshes, case_mesh, flows, **opts):
        L = load_library()
        self._meshes = [_MeshHold(m) for m in meshes]
        descs = (MeshDesc * len(meshes))(*[m.desc for m in self._meshes])
        self._descs = descs
        self.case_mesh = np.ascontiguousarray(case_mesh, dtype=np.int32)
        self.ncase = len(self.case_mesh)
        fl = (FlowParams * self.ncase)(*[f if isinstance(f, FlowParams) else flow_params(**f)
                                         for f in flows])
        self._flows = fl
        self.opts = scheme_opts(**opts)
        h = C.c_void_p()
        _check(L.fvm_batch_create(len(meshes), descs, self.ncase, _iptr(self.case_mesh), fl,
                                  C.byref(self.opts), C.byref(h)))
        self.h = h
        self.ncells = [self._meshes[m].desc.ncell for m in self.case_mesh]

    def close(self):
        if getattr(self, "h", None):
            load_library().fvm_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def run(self, max_steps, epsl=-1.0, check=True):
        hist = np.zeros((self.ncase, max_steps + 1))
        steps = np.zeros(self.ncase, np.int32)
        status = np.zeros(self.ncase, np.int32)
        st = load_library().fvm_batch_run(self.h, max_steps, epsl, _dptr(hist), _iptr(steps),
                                          _iptr(status))
        if check and st != 0:
            _check(st)
        return hist, steps, status

    def state(self, icase):
        out = [np.zeros(self.ncells[icase]) for _ in range(4)]
        _check(load_library().fvm_batch_get_state(self.h, icase, *(_dptr(a) for a in out)))
        return np.stack(out, axis=1)

    def set_state(self, icase, W):
        cols = [np.ascontiguousarray(W[:, k], dtype=np.float64) for k in range(4)]
        _check(load_library().fvm_batch_set_state(self.h, icase, *(_dptr(a) for a in cols)))

    def forces(self):
        cl, cd = np.zeros(self.ncase), np.zeros(self.ncase)
        _check(load_library().fvm_batch_get_forces(self.h, _dptr(cl), _dptr(cd)))
        return cl, cd

    def timing(self):
        ms, csu = C.c_double(), C.c_int64()
        _check(load_library().fvm_batch_get_timing(self.h, C.byref(ms), C.byref(csu)))
        return ms.value, csu.value

    def kernel_times(self, reps=20):
        ms = np.zeros(2)
        _check(load_library().fvm_batch_kernel_times(self.h, reps, _dptr(ms)))
        return float(ms[0]), float(ms[1])

    def config(self):
        cfg = np.zeros(6, np.int32)
        _check(load_library().fvm_batch_get_config(self.h, _iptr(cfg)))
        return dict(tile_cells=int(cfg[0]), threads=int(cfg[1]), steps_per_graph=int(cfg[2]),
                    tiles=int(cfg[3]), blocks_per_cu=int(cfg[4]), cus=int(cfg[5]))
