"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def say(*a):
    """Progress line naming the case before a GPU step (flushed)."""
    print("[case]", *a, flush=True)


@pytest.fixture(scope="session")
def oracle():
    import cgs_oracle
    cgs_oracle.lib()
    return cgs_oracle


@pytest.fixture(scope="session")
def ctx():
    import cgspline_amd as cgs
    say("cgs_open(0)")
    c = cgs.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def c1():
    from cgspline_amd import scenes
    return scenes.heightfield_c1()


@pytest.fixture(scope="session")
def c2_small():
    """The closed-surface workload at 320×180 (same scene, fewer rays)."""
    from cgspline_amd import scenes
    return scenes.closed_surfaces(width=320, height=180)


@pytest.fixture(scope="session")
def c3_small():
    """The displaced-grid generator with 96 grids (24,576 patches) at 256²."""
    from cgspline_amd import scenes
    return scenes.displaced_grids(grids=96, width=256, height=256, box=6.0)


@pytest.fixture(scope="session")
def grazing():
    """A 128×128-patch height field at a grazing angle (128² rays): deep
    enough that traversals postpone more than 8 siblings at once."""
    from cgspline_amd import scenes
    return scenes.grazing_heightfield()


def rng(seed=0):
    return np.random.default_rng(seed)


@pytest.fixture(scope="session")
def stacked():
    """180,000 objects stacked along y and 64 rays through all of them
    (scenes.stacked_arcs)."""
    from cgspline_amd import scenes
    return scenes.stacked_arcs()
