"""The C-ABI library on the CPU: it loads, exports every function the header
declares, its host-side topology equals the oracle's literal restatement of
CalculationMeshGeo, and bad inputs come back as status codes (never exit)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from gpueularsolver2d_amd import Solver, mesh_topology, naca0012_mesh, read_dat
from gpueularsolver2d_amd.meshgen import Mesh
from gpueularsolver2d_amd.solver import EXPORTS, FvmError, load_library
from oracle_ctypes import OracleMesh
from small_meshes import annulus, square2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_functions():
    src = open(os.path.join(ROOT, "include", "fvm_abi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fvm_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_function():
    L = load_library()
    declared = header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(EXPORTS) == declared


def test_status_strings():
    L = load_library()
    for s in range(9):
        assert L.fvm_status_string(s)
    assert L.fvm_status_string(0) == b"ok"


@pytest.mark.parametrize("name", ["2822-3k", "2822-7k", "naca3000", "annulus8x1", "annulus24x3",
                                  "square2"])
def test_topology_equals_oracle(name):
    if name.startswith("naca"):
        m = naca0012_mesh(3000, seed=1)
    elif name.startswith("annulus"):
        m = annulus(*map(int, name[7:].split("x")))
    elif name == "square2":
        m = square2()
    else:
        m = read_dat(os.path.join(GOLDEN, name + ".dat.gz"))
    t = mesh_topology(m)
    om = OracleMesh(m.x, m.y, m.tri, literal=(m.ncell < 15000))
    assert (t["nnode"], t["ncell"], t["nedge"], t["nwall"], t["nfar"]) == \
        (om.nnode, om.ncell, om.nedge, om.nwall, om.nfar)
    for k in ("elog", "left", "right", "farid", "wallid", "node1", "node2"):
        assert np.array_equal(t[k], getattr(om, k)[1:]), k
    assert np.array_equal(t["celledge"], om.celledge[1:])
    assert np.array_equal(t["neicell"], om.neicell[1:])
    assert np.array_equal(t["clog"], om.clog[1:])


def square():
    x = np.array([0.0, 1.0, 1.0, 0.0])
    y = np.array([0.0, 0.0, 1.0, 1.0])
    return x, y


def test_bad_node_id_is_mesh_error():
    x, y = square()
    with pytest.raises(FvmError) as e:
        mesh_topology(Mesh(x, y, np.array([[1, 2, 5]])))
    assert e.value.name == "FVM_ERR_MESH"


def test_clockwise_cell_is_area_error():
    x, y = square()
    with pytest.raises(FvmError) as e:
        mesh_topology(Mesh(x, y, np.array([[1, 3, 2]])))
    assert e.value.name == "FVM_ERR_AREA"


def test_nonmanifold_edge_is_mesh_error():
    x = np.array([0.0, 1.0, 0.5, 0.5, 0.5])
    y = np.array([0.0, 0.0, 1.0, 2.0, -1.0])
    # edge (1,2) shared by three counter-clockwise triangles
    tri = np.array([[1, 2, 3], [1, 2, 4], [2, 1, 5]])
    with pytest.raises(FvmError) as e:
        mesh_topology(Mesh(x, y, tri))
    assert e.value.name == "FVM_ERR_MESH"


def test_null_arguments_are_argument_errors():
    L = load_library()
    cnt = np.zeros(5, np.int32)
    none = C.cast(None, C.POINTER(C.c_int32))
    st = L.fvm_mesh_topology(None, cnt.ctypes.data_as(C.POINTER(C.c_int32)), *([none] * 10))
    assert st == 1
    out = C.c_void_p()
    assert L.fvm_create(None, None, None, C.byref(out)) == 1
    assert L.fvm_run(None, 10, 0.0, None, None) == 1
    assert L.fvm_batch_run(None, 10, 0.0, None, None, None) == 1


def test_invalid_options_rejected_before_device():
    m = naca0012_mesh(3000, seed=1)
    for bad in (dict(order=3), dict(precision="fp32", kernels="plain"), dict(threads=100)):
        with pytest.raises(FvmError) as e:
            Solver(m, 0.5, 0.0, **bad)
        assert e.value.name == "FVM_ERR_ARGUMENT"


def test_tiling_validated_on_host_then_device_required():
    # on a machine without a GPU, creation gets through the host-side mesh
    # pre-processing and tile self-checks and stops at the device probe
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    m = naca0012_mesh(3000, seed=1)
    for tile in (8, 64, 256, 2048):
        with pytest.raises(FvmError) as e:
            Solver(m, 0.5, 0.0, tile_cells=tile)
        assert e.value.name == "FVM_ERR_NO_DEVICE"
