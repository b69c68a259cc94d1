"""Shared test setup: import paths, the `gpu` marker and scene helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pbd-simulation-framework-cuda_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs through the C-ABI of libpbd.so")


@pytest.fixture(scope="session")
def F():
    import pbdfluid

    return pbdfluid


@pytest.fixture(scope="session")
def O():
    import pyoracle

    return pyoracle


def small_dam(F, n_side=16, variant="textbook", **overrides):
    """The 32k dam-break parameter set scaled to an n_side^3 lattice."""
    p = F.load_params(F.config_path("dam_break_32k"))
    p.box_width = p.box_height = p.box_depth = n_side
    for k, v in overrides.items():
        setattr(p, k, v)
    return p
