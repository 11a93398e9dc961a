"""The C-ABI library loads and exports every function include/cmg_hip.h declares (no
compute calls: these run without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cmg_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(cmg_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_reference_subset():
    names = declared_functions()
    # the stable C ABI voxel/PLC subset of bindings/c/include/cmg_c/cmg_c.h:36-153
    for n in ["cmg_plc_create", "cmg_plc_destroy", "cmg_plc_set_points", "cmg_plc_add_facet",
              "cmg_plc_voxelize", "cmg_voxels_destroy", "cmg_voxels_dims",
              "cmg_voxels_origin", "cmg_voxels_spacing", "cmg_voxels_occupancy",
              "cmg_voxels_distance", "cmg_version",
              # the device layer of include/cmg/backend/cuda_kernels.hpp:14-21
              "cmg_hip_available", "cmg_hip_orient3d_signs"]:
        assert n in names, n


def test_library_exports_every_declared_symbol():
    import cybermeshgenerator_amd as cm
    lib = ctypes.CDLL(cm.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_handles_and_input_validation_without_gpu():
    """Input checks run before any device work, like the reference C ABI
    (cmg_c.cpp:79-106, 242-249)."""
    import cybermeshgenerator_amd as cm
    lib = cm.lib()
    h = lib.cmg_plc_create()
    assert h
    assert lib.cmg_plc_add_facet(h, None, 0, 0) == 1          # count < 3
    out = ctypes.c_void_p(123)
    err = ctypes.create_string_buffer(64)
    st = lib.cmg_plc_voxelize(h, 0, 0, 1, ctypes.byref(out), err, 64)
    assert st == 1 and out.value is None                       # resolution < 1
    assert b"resolution" in err.value
    st = lib.cmg_plc_voxelize(h, 8, 0, 1, ctypes.byref(out), err, 64)
    assert st == 1 and b"no points" in err.value               # empty PLC
    lib.cmg_plc_destroy(h)
    assert cm.version().endswith("hip")


def test_no_cpu_fallback_without_device():
    import cybermeshgenerator_amd as cm
    if cm.hip_available():
        pytest.skip("a GPU is present; the no-device error path is not reachable")
    import numpy as np
    with pytest.raises(RuntimeError, match="no usable HIP device"):
        cm.orient3d_signs(np.zeros((4, 3)), np.array([[0, 1, 2, 3]]))
