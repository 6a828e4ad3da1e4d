"""Shared test setup: the `gpu` marker, repo on sys.path, oracle helpers."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libfsim_hip)")


def oracle_cfg(cfg: dict, solver: int = 0, nthreads: int = 1, rtol=None) -> dict:
    """The library configuration in the CPU oracle's layout."""
    d = dict(cfg)
    d["rtol"] = cfg.get("lapl_rtol", 1e-10) if rtol is None else rtol
    d["maxit"] = cfg.get("lapl_maxit", 20000)
    d["solver"] = solver
    d["nthreads"] = nthreads
    return d


def rel_l2(a, b):
    import numpy as np

    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope="session")
def O():
    from oracle import oracle

    oracle.build()
    return oracle
