"""The C-ABI library loads on the CPU and exports every entry point that
include/windhip.h declares; argument checks answer before any device call."""
import ctypes as C
import os
import re
import subprocess

import pytest

from hicar_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "windhip.h")).read()
    return sorted(set(re.findall(r"\b(windhip_[a-z_0-9]+)\s*\(", txt)))


def test_header_and_exports_agree():
    lib = nat.load()
    decl = _declared()
    assert decl, "no declarations found"
    missing = [f for f in decl if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(nat.EXPORTS) == decl


def test_struct_layouts_match_header(tmp_path):
    """Every field offset of the ctypes mirrors equals the C compiler's."""
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "windhip.h"', "int main(void){"]
    pairs = [("windhip_desc", nat.Desc), ("windhip_solve_opts", nat.SolveOpts), ("windhip_solve_info", nat.SolveInfo)]
    for cname, py in pairs:
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = []
    for _, py in pairs:
        want.append(C.sizeof(py))
        want += [getattr(py, f).offset for f, _ in py._fields_]
    assert got == want


def test_default_opts_are_the_reference_defaults():
    o = nat.default_opts()
    assert o.tol_rel == 1e-5 and o.tol_abs == 1e-10 and o.max_iters == 1500  # wind_iterative.F90:57-61
    assert o.start == nat.START_WARM and o.retry == 1 and o.precision == nat.PLAIN_F64


@pytest.mark.parametrize("field,val", [("nx", 2), ("nz", 1), ("dx", 0.0), ("ims", 2), ("field_dtype", 2),
                                       ("field_mem", 7)])
def test_create_rejects_bad_descriptors(field, val):
    lib = nat.load()
    d = nat.Desc(nx=8, ny=8, nz=4, dx=100.0, ims=1, ime=8, kms=1, kme=4, jms=1, jme=8,
                 field_dtype=nat.WINDHIP_F64, field_mem=nat.WINDHIP_HOST)
    setattr(d, field, val)
    h = C.c_void_p()
    assert lib.windhip_create(C.byref(d), C.byref(h)) == nat.WINDHIP_ERR_ARG
    assert not h.value
    assert lib.windhip_last_error()


def test_null_handle_calls_fail_cleanly():
    lib = nat.load()
    assert lib.windhip_solve(None, None, None, None, None) == nat.WINDHIP_ERR_ARG
    assert lib.windhip_destroy(None) == nat.WINDHIP_OK
