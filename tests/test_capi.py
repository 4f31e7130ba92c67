"""The C-ABI library: it is built, loads without a GPU, and exports exactly the
entry points include/adpres_gpu.h declares (no compute call is made here)."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, golden
from adpres_amd import engine
from adpres_amd.configs import generate
from adpres_amd.problem import problem_from_dict, read_ref_dump


def declared_functions():
    text = open(os.path.join(ROOT, "include", "adpres_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(adg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_exports_the_binding_uses():
    assert declared_functions() == sorted(engine.EXPORTS)


def test_library_loads_and_exports_every_declared_symbol():
    assert os.path.exists(engine.LIB_PATH), "build first: python -m adpres_amd._build"
    lib = ctypes.CDLL(engine.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_status_strings_without_gpu():
    lib = engine.load_library()
    assert lib.adg_status_string(0) == b"ok"
    assert b"OUTER" in lib.adg_status_string(1)
    assert b"not stable" in lib.adg_status_string(2)


def test_create_rejects_bad_geometry_without_touching_the_gpu():
    lib = engine.load_library()
    h = ctypes.c_void_p()
    assert lib.adg_create(None, ctypes.byref(h)) == engine.ADG_E_ARG
    g = engine.AdgGeom()
    g.nnod = 0
    assert lib.adg_create(ctypes.byref(g), ctypes.byref(h)) == engine.ADG_E_ARG


def test_differentiation_entry_points_reject_bad_arguments_without_touching_the_gpu():
    lib = engine.load_library()
    w = np.zeros(12)
    it = engine.AdgIter()
    out = np.zeros(8)
    dp = ctypes.POINTER(ctypes.c_double)
    # no context, no cotangent, no output, tol <= 0, maxit < 1
    assert lib.adg_nodal_vjp(None, 2, w.ctypes.data_as(dp), None, None, None, None) == engine.ADG_E_ARG
    assert lib.adg_nodal_vjp(None, 2, None, None, None, None, None) == engine.ADG_E_ARG
    for tol, maxit, grad in ((1e-9, 10, None), (0.0, 10, out), (1e-9, 0, out)):
        g = grad.ctypes.data_as(dp) if grad is not None else None
        assert lib.adg_backward_nodal(None, ctypes.byref(it), 2, 1.0, None, tol, maxit, g, None, None) \
            == engine.ADG_E_ARG


def test_generators_reproduce_the_fixture_cores():
    import hashlib
    d = golden("core3d_2g_s1_solve")
    p = generate("core3d_2g", int(d["seed"])).problem
    h = hashlib.sha256()
    for k in ("ix", "iy", "iz", "mat", "xdel", "ydel", "zdel", "sigtr", "siga", "nuf", "sigf", "sigs", "chi", "dc"):
        h.update(np.ascontiguousarray(getattr(p, k)).tobytes())
    assert h.hexdigest() == str(d["problem_sha"])
    d2 = golden("core2d_small_s1_solve")
    q = generate("core2d_small", 1).problem
    ref = problem_from_dict(d2)
    for k in ("ix", "iy", "iz", "mat", "sigtr", "sigs", "D", "sigr"):
        assert np.array_equal(getattr(q, k), getattr(ref, k)), k


def test_batch_needs_one_group_count_without_touching_the_gpu():
    import pytest
    from adpres_amd.configs import batch_core
    from adpres_amd.engine import Batch
    cores = [batch_core(3, i).problem for i in range(6)]
    assert len({q.ng for q in cores}) > 1
    with pytest.raises(ValueError):
        Batch(cores)
    with pytest.raises(ValueError):
        Batch([])


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes structures of the Python binding have the sizes and field
    offsets of the C structs in include/adpres_gpu.h (compiled here with the
    host compiler), so a binding that lays them out differently fails on the
    CPU rather than on the GPU."""
    import subprocess
    structs = {"adg_geom": engine.AdgGeom, "adg_iter": engine.AdgIter, "adg_trace": engine.AdgTrace,
               "adg_xs_tables": engine.AdgXsTables, "adg_xtab_tables": engine.AdgXtabTables}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "adpres_gpu.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(ln.split()[:2]): int(ln.split()[2]) for ln in out if ln.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


def test_fortran_binding_types_match_the_header(tmp_path):
    """The bind(C) derived types of the Fortran binding
    (adpres_amd/fortran/adpres_gpu_mod.f90) have the sizes of the C structs
    (compiled here with amdflang and gcc; skipped without amdflang)."""
    import shutil
    import subprocess
    fc = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)
    if not fc:
        pytest.skip("no Fortran compiler")
    types = {"adg_geom": "adg_geom_t", "adg_iter": "adg_iter_t", "adg_trace": "adg_trace_t",
             "adg_xs_tables": "adg_xs_tables_t", "adg_xtab_tables": "adg_xtab_tables_t"}
    prog = ["program sizes", "  use iso_c_binding", "  use adpres_gpu"]
    prog += [f"  type({ft}) :: v_{c}" for c, ft in types.items()]
    prog += [f"  print '(A,1X,I0)', '{c}', c_sizeof(v_{c})" for c in types]
    prog += ["end program sizes"]
    (tmp_path / "sizes.f90").write_text("\n".join(prog) + "\n")
    mod = os.path.join(ROOT, "adpres_amd", "fortran", "adpres_gpu_mod.f90")
    subprocess.run([fc, "-J", str(tmp_path), mod, str(tmp_path / "sizes.f90"), "-o", str(tmp_path / "fsizes")],
                   check=True, cwd=tmp_path, capture_output=True)
    fout = subprocess.run([str(tmp_path / "fsizes")], check=True, capture_output=True, text=True).stdout.split()
    fs = dict(zip(fout[0::2], (int(v) for v in fout[1::2])))
    csrc = ['#include <stdio.h>', '#include "adpres_gpu.h"', "int main(void) {"]
    csrc += [f'  printf("{c} %zu\\n", sizeof({c}));' for c in types]
    csrc += ["  return 0;", "}"]
    (tmp_path / "csizes.c").write_text("\n".join(csrc))
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(tmp_path / "csizes.c"), "-o",
                    str(tmp_path / "csizes")], check=True)
    cout = subprocess.run([str(tmp_path / "csizes")], check=True, capture_output=True, text=True).stdout.split()
    cs = dict(zip(cout[0::2], (int(v) for v in cout[1::2])))
    assert fs == cs


def test_persistent_plans_are_checked_on_the_host():
    """adg_persistent_plan (no GPU): every fixture geometry, the measured
    cores and seeded random boxes get a brick plan whose every kernel index
    is in bounds, every node owned once and every stencil neighbour on the
    tile."""
    from adpres_amd.engine import persistent_plan
    shapes = [(68, 68, 96), (34, 34, 26), (18, 18, 2), (17, 17, 19), (136, 136, 2), (9, 9, 2), (2, 2, 2),
              (24, 24, 2), (200, 3, 2), (3, 200, 2), (2, 2, 300)]
    rng = np.random.default_rng(17)
    shapes += [tuple(int(v) for v in rng.integers(2, 90, 3)) for _ in range(12)]
    for sh in shapes:
        for prec in ("fp64", "fp32_refine"):
            for ncu in (256, 64):
                pl = persistent_plan(*sh, precision=prec, ncu=ncu)
                if pl["ok"]:
                    assert pl["checked"], (sh, prec, ncu, pl)
                    assert pl["blocks"] % 8 == 0 and pl["blocks"] <= ncu
    # the 8-group core gets one brick per CU, 8 nodes per thread
    pl = persistent_plan(68, 68, 96)
    assert pl["ok"] and pl["blocks"] == 256 and pl["nodes_per_thread"] == 8
