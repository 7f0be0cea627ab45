"""The C-ABI library loads and exports every symbol include/*.h declares (no compute calls)."""
import ctypes
import glob
import os
import re

import moris_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**\s*(mgpu_[A-Za-z0-9_]+)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = M.load_library()
    names = declared_symbols()
    assert "mgpu_set_assemble" in names and "mgpu_set_create" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(M.EXPORTED) <= names


def test_error_string_after_bad_call():
    lib = M.load_library()
    out = ctypes.c_void_p()
    rc = lib.mgpu_problem_build(None, ctypes.byref(out))
    assert rc == 1
    assert b"null" in lib.mgpu_last_error_string()


def _create_status(desc):
    lib = M.load_library()
    out = ctypes.c_void_p()
    rc = lib.mgpu_set_create(ctypes.byref(desc), ctypes.byref(out))
    if rc == 0:
        lib.mgpu_set_destroy(out)
    return rc, lib.mgpu_last_error_string().decode()


def test_create_rejects_bad_material_before_touching_a_device():
    """Non-positive or non-finite k, E <= 0 and nu outside (-1, 1/2) are refused by mgpu_set_create
    with MGPU_ERR_INVALID_ARG (the stencil encodes void and cut cells beside these values), before
    any device is needed."""
    for iwg, bad in ((M.IWG_SPATIALDIFF_BULK, [0.0, -0.5, float("nan"), float("inf")]),
                     (M.IWG_STRUC_LINEAR_BULK, [(0.0, 0.3), (-1.0, 0.3), (1.0, 0.5), (1.0, -1.0), (float("nan"), 0.3)])):
        spec, keep = M.make_spec(3, (4, 4, 4), iwg=iwg, spheres=[[0.5, 0.5, 0.5, 0.3]])
        P = M.Problem(spec, keep)
        a = P.arrays()
        for b in bad:
            m = a["material"].copy()
            if iwg == M.IWG_SPATIALDIFF_BULK:
                m[3] = b
            else:
                m[6:8] = b
            d, k = M.desc_from_arrays(P.desc, dict(a, material=m))
            rc, msg = _create_status(d)
            assert rc == 1, (iwg, b, rc, msg)
            assert ("conductivity" in msg) if iwg == M.IWG_SPATIALDIFF_BULK else ("nu" in msg)


def test_gpu_setup_validates_the_spec_before_touching_a_device():
    """mgpu_setup_build refuses THB, p > 2, Lagrange order > 2 and enrichment with MGPU_ERR_UNSUPPORTED and
    a null spec with MGPU_ERR_INVALID_ARG before it looks for a device; a valid spec (HEX8, HEX27 with
    p = 2 or nodal rows) without a device returns MGPU_ERR_NO_DEVICE (no CPU fallback)."""
    from moris_amd import gpu_setup as G
    lib = G._lib()
    out = ctypes.c_void_p()
    assert lib.mgpu_setup_build(None, None, ctypes.byref(out)) == 1
    for kw in (dict(lagrange_order=3), dict(lagrange_order=2, bspline_order=2, refine_levels=1),
               dict(refine_levels=1, bspline_order=2), dict(bspline_order=3), dict(enrich=1)):
        spec, keep = M.make_spec(3, (4, 4, 4), iwg=M.IWG_STRUC_LINEAR_BULK, spheres=[[0.5, 0.5, 0.5, 0.3]], **kw)
        assert lib.mgpu_setup_build(ctypes.byref(spec), None, ctypes.byref(out)) == 3, kw
    if M.device_count() == 0:
        for kw in ({}, dict(lagrange_order=2, bspline_order=2), dict(lagrange_order=2)):
            spec, keep = M.make_spec(3, (4, 4, 4), spheres=[[0.5, 0.5, 0.5, 0.3]], **kw)
            assert lib.mgpu_setup_build(ctypes.byref(spec), None, ctypes.byref(out)) == 4, kw
