"""The Fortran host program (tests/fortran/test_drop_in.F90) drives the solver
through the ISO_C_BINDING interface exactly as HICAR would call it."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "test_drop_in")


def test_fortran_program_is_built():
    assert os.path.exists(EXE), "run __graft_entry__.build() (tests/fortran/Makefile)"


@pytest.mark.gpu
def test_fortran_drop_in_reduces_divergence():
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "PASS" in out.stdout
