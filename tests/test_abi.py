"""C-ABI checks that need no GPU: the library loads, exports every entry point
include/rmh.h declares, its structs match the Python mirror byte for byte, and
host-only entry points behave."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import REPO, load_package


def header_functions():
    text = open(os.path.join(REPO, "include", "rmh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rmh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(rmh):
    L = rmh.load_library()
    declared = header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(rmh.EXPORTED) == declared


def test_struct_layouts_match_header(rmh):
    abi = rmh.abi
    assert abi.TU_DTYPE.itemsize == 16 and abi.PU_DTYPE.itemsize == 16
    assert ctypes.sizeof(abi.Plane) == 24
    assert ctypes.sizeof(abi.McFrame) == 144
    assert ctypes.sizeof(abi.DbkFrame) == 104
    assert ctypes.sizeof(abi.GraphNode) == 24
    assert ctypes.sizeof(abi.Config) == 12
    from workloads import EDGE_CELL_DTYPE, PU_DTYPE, TU_DTYPE
    assert TU_DTYPE == abi.TU_DTYPE and PU_DTYPE == abi.PU_DTYPE and EDGE_CELL_DTYPE == abi.EDGE_CELL_DTYPE
    assert ctypes.sizeof(abi.EdgeFrame) == 40
    assert ctypes.sizeof(abi.LmcsAps) == 36 and ctypes.sizeof(abi.LmcsFrame) == 88
    assert ctypes.sizeof(abi.AlfAps) == 1186 and abi.AlfAps.chroma_coeff.offset == 930
    assert ctypes.sizeof(abi.AlfFrame) == 208 and abi.ALF_CTB_DTYPE.itemsize == 8


def test_kernel_matrices_equal_oracle_matrices(rmh, po):
    for t, sizes in [(0, [2, 4, 8, 16, 32, 64]), (1, [4, 8, 16, 32]), (2, [4, 8, 16, 32])]:
        for n in sizes:
            assert (rmh.transform_matrix(t, n) == po.transform_matrix(t, n)).all(), (t, n)
    for t, n in [(1, 64), (2, 2), (0, 3), (7, 4)]:
        with pytest.raises(rmh.InvalidInput):
            rmh.transform_matrix(t, n)


def test_lfnst_tables_match_oracle(po):
    """The kernel-side LFNST header and the oracle header hold the same kernels."""
    a = open(os.path.join(REPO, "oracle", "lfnst_data.h")).read()
    b = open(os.path.join(REPO, "rustmpeg-next_amd", "csrc", "lfnst_data.h")).read()
    nums = lambda t: re.findall(r"-?\d+", t.split("=", 1)[1])
    assert nums(a.replace("ORC", "X")) == nums(b.replace("RMH", "X"))


def test_create_reports_typed_status_without_usable_device(rmh):
    """Here (no GPU) rmh_create must fail with a typed status, never crash."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); from conftest import load_package; m = load_package();\n"
            "try:\n    m.Accelerator(0); print('ok')\nexcept m.AcceleratorError as e:\n    print(type(e).__name__)\n"
            % os.path.join(REPO, "tests"))
    env = dict(os.environ, HIP_VISIBLE_DEVICES=os.environ.get("HIP_VISIBLE_DEVICES", ""))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, env=env)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() in ("Unavailable", "ok")


def test_null_arguments_are_rejected(rmh):
    L = rmh.load_library()
    assert L.rmh_create(None, None) == rmh.abi.RMH_INVALID_INPUT
    assert L.rmh_residual_batch(None, None, 0, None, 0, None) == rmh.abi.RMH_INVALID_INPUT
    assert L.rmh_info(None, None) == rmh.abi.RMH_INVALID_INPUT
    assert L.rmh_plan_bytes(None) == 0
    assert L.rmh_last_error(None) == b"null context"
