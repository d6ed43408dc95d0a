"""The Fortran host programs drive the solver through the ISO_C_BINDING
interface as HICAR would: tests/fortran/test_drop_in.F90 through the
array-level module, tests/fortran/test_domain_adapter.F90 through the
`wind_iterative` adapter with the reference's domain_t / options_t
signatures (against stub HICAR modules), two nests and the nest cache."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "test_drop_in")
ADAPTER = os.path.join(ROOT, "tests", "fortran", "_build", "test_domain_adapter")


def test_fortran_program_is_built():
    assert os.path.exists(EXE), "run __graft_entry__.build() (tests/fortran/Makefile)"
    assert os.path.exists(ADAPTER), "run __graft_entry__.build() (tests/fortran/Makefile)"


@pytest.mark.gpu
def test_fortran_drop_in_reduces_divergence():
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "PASS" in out.stdout


@pytest.mark.gpu
def test_fortran_domain_adapter_two_nests():
    out = subprocess.run([ADAPTER], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    assert "PASS" in out.stdout and "FAIL" not in out.stdout
