"""The C ABI library: it loads on a machine without a GPU, exports every symbol
declared in include/*.h, and rejects bad arguments before touching a device.
No compute calls are made here."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

INCLUDE = os.path.join(ROOT, "include")


def declared_c_functions():
    names = set()
    for h in ("gardenia_graph.h", "gardenia_hip.h"):
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(gd_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_loads_and_exports_every_declared_symbol():
    from gardenia_ocl_amd import _native

    lib = _native.lib()
    declared = declared_c_functions()
    assert len(declared) >= 20
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # the ctypes signature table covers exactly the declared API
    assert declared == set(_native.SIGNATURES), declared ^ set(_native.SIGNATURES)


def test_reference_plugin_symbols_exported():
    """PRSolver / SpmvSolver with the reference's C++ signatures
    (src/pr/pr.h:23, src/spmv/spmv.h:29) are defined by the library."""
    from gardenia_ocl_amd import LIB_PATH

    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    assert "_Z8PRSolveriiPiS_S_S_S_Pf" in out  # PRSolver(int,int,int*,int*,int*,int*,int*,float*)
    assert "_Z10SpmvSolveriiPiS_PfS0_S0_" in out  # SpmvSolver(int,int,int*,int*,float*,float*,float*)
    # SymGSSolver(int,int,int*,int*,int*,float*,float*,float*,std::vector<int>)
    assert "_Z11SymGSSolveriiPiS_S_PfS0_S0_St6vectorIiSaIiEE" in out


def test_argument_errors_before_device():
    from gardenia_ocl_amd import _native as N

    lib = N.lib()
    h = ctypes.c_void_p()
    assert lib.gd_pr_plan_create(-1, 0, None, None, None, None, ctypes.byref(h)) == N.GD_ERR_INVALID
    assert b"out of range" in lib.gd_last_error()
    assert lib.gd_pr_plan_create(4, 3, None, None, None, None, ctypes.byref(h)) == N.GD_ERR_INVALID
    assert lib.gd_spmv_plan_create(4, 4, 3, None, None, None, None, ctypes.byref(h)) == N.GD_ERR_INVALID
    assert lib.gd_pr_plan_run(None, None, None, None, None) == N.GD_ERR_INVALID
    assert lib.gd_spmv_plan_run(None, None, None) == N.GD_ERR_INVALID
    assert lib.gd_pr_plan_destroy(None) == N.GD_OK
    co = (ctypes.c_int32 * 3)(0, 5, 3)
    assert lib.gd_symgs_plan_create(8, 0, None, None, None, None, co, 2, None,
                                    ctypes.byref(h)) == N.GD_ERR_INVALID
    assert lib.gd_symgs_plan_run(None, None, None, 1, None) == N.GD_ERR_INVALID
    assert lib.gd_symgs_plan_destroy(None) == N.GD_OK


def test_pr_plan_config_rejected_before_device():
    """Out-of-range plan knobs are argument errors, checked before any device
    work (an empty graph, so no device array is needed)."""
    from gardenia_ocl_amd import _native as N

    lib = N.lib()
    h = ctypes.c_void_p()
    for field, value, text in (("hub_threshold", 63, b"hub_threshold"),
                               ("hub_threshold", 4097, b"hub_threshold"),
                               ("block", 300, b"block"), ("items_per_thread", 4, b"items_per_thread"),
                               ("stage_hubs", 2, b"stage_hubs"), ("stage_hubs", -2, b"stage_hubs")):
        cfg = N.PrPlanConfig()
        lib.gd_pr_default_plan_config(ctypes.byref(cfg))
        assert cfg.hub_threshold == 0 and cfg.block == 0 and cfg.stage_hubs == -1
        setattr(cfg, field, value)
        assert lib.gd_pr_plan_create_ex(0, 0, None, None, None, None, ctypes.byref(cfg),
                                        ctypes.byref(h)) == N.GD_ERR_INVALID, field
        assert text in lib.gd_last_error()
    cfg = N.PrPlanConfig()
    lib.gd_pr_default_plan_config(ctypes.byref(cfg))
    cfg.items_per_thread, cfg.hub_threshold = 8, 2049  # above 256*8
    assert lib.gd_pr_plan_create_ex(0, 0, None, None, None, None, ctypes.byref(cfg),
                                    ctypes.byref(h)) == N.GD_ERR_INVALID


def test_default_options_are_reference_constants():
    from gardenia_ocl_amd import _native as N

    o = N.PrOptions()
    N.lib().gd_pr_default_options(ctypes.byref(o))
    assert o.damping == np.float32(0.85)  # kDamp, include/common.h:18
    assert o.tol == 1e-3 and o.max_iter == 30  # EPSILON, MAX_ITER, src/pr/pr.h:4-5
    assert o.seed_vertex == -1


def test_no_device_is_reported_not_faked():
    """Without a GPU the device entry points fail loudly (no CPU fallback)."""
    from gardenia_ocl_amd import _native as N

    lib = N.lib()
    if lib.gd_device_count() > 0:
        pytest.skip("a device is visible")
    off = np.array([0, 1], np.int32)
    col = np.array([0], np.int32)
    deg = np.array([1], np.int32)
    sc = np.array([1.0], np.float32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    st = lib.gd_pr_solve(1, 1, p(off), p(col), None, None, p(deg), p(sc), None, None, None)
    assert st == N.GD_ERR_NO_DEVICE
    from gardenia_ocl_amd import GardeniaError, PRSolver

    with pytest.raises(GardeniaError, match="GD_ERR_NO_DEVICE"):
        PRSolver(1, 1, off, col, None, None, deg, sc)


def test_oracle_is_not_imported_by_the_product():
    """The product never loads the oracle: no module of gardenia_ocl_amd names it."""
    pkg = os.path.join(ROOT, "gardenia_ocl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "liboracle" not in text, f


def test_noalloc_limits_rejected_before_device():
    """gd_pr_run_noalloc rejects sizes beyond its static scratch before any
    device work (so also here, without a GPU)."""
    from gardenia_ocl_amd import _native as N
    from gardenia_ocl_amd.pr import make_options

    lib = N.lib()
    o = make_options(tol=1e-3, max_iter=10)
    big = (1 << 21) + 1
    assert lib.gd_pr_run_noalloc(big, 0, None, None, None, None, ctypes.byref(o), None, None,
                                 None) == N.GD_ERR_INVALID
    assert b"GD_NOALLOC_MAX_VERTICES" in lib.gd_last_error()
    o.max_iter = 4097
    assert lib.gd_pr_run_noalloc(4, 0, None, None, None, None, ctypes.byref(o), None, None,
                                 None) == N.GD_ERR_INVALID
    assert b"max_iter" in lib.gd_last_error()
