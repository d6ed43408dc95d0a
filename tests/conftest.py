"""Shared fixtures. GPU tests are marked `gpu`; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def small_case(name: str, seed: int = 3):
    """Small seeded cases that the CPU oracle solves in well under a second."""
    from hicar_amd import synthetic as S
    if name == "hill":
        return S.gaussian_hill_case(nx=24, ny=20, nz=10, dx=250.0, seed=seed)
    if name == "fractal":
        return S.fractal_case(nx=21, ny=19, nz=8, dx=50.0, relief=400.0, seed=seed, lowest=10.0, top=6000.0)
    if name == "fractal_alpha":
        return S.fractal_case(nx=18, ny=23, nz=9, dx=50.0, relief=400.0, seed=seed, lowest=10.0, top=6000.0,
                              varying_alpha=True)
    if name == "flat":
        return S.flat_plane_case(nx=16, ny=14, nz=6)
    if name == "config0":
        return S.gaussian_hill_case(seed=seed)
    raise KeyError(name)


def oracle_for(case):
    from oracle import Oracle
    o = Oracle(case.nx, case.ny, case.nz, case.dx)
    o.set_metrics(case.metrics)
    return o


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
