"""Shared test setup: import paths, the `gpu` marker, scene helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "mls-mpm_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle

    oracle.lib("f32")
    oracle.lib("f64")
    return oracle


def dam_block(grid, lo, hi, seed=0, jitter=0.1):
    from mlsmpm import scenes

    return scenes.lattice_block(lo, hi, jitter, np.random.default_rng(seed))
