import sys
from functools import lru_cache
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


@lru_cache(maxsize=1)
def maps():
    from mit_jax_amd.exchange import ExchangeMaps
    return ExchangeMaps.load()


@lru_cache(maxsize=1)
def oracle_tables():
    from oracle import tables as OT
    return OT.tables(OT.Topology.from_facets(OT.LLC90_FACETS))


@lru_cache(maxsize=2)
def dst3_case(seed=0):
    """Config-1 inputs on the seed grid: (grid, u, v, w, tracer)."""
    from mit_jax_amd import synthetic as S
    g = S.make_grid(seed)
    u, v, w = S.flow(g, seed + 2)
    tr = S.tracer(g, seed + 3)
    return g, u, v, w, tr


@pytest.fixture(scope="session")
def exmaps():
    return maps()


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def to_dev(a, dev):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), device=dev)
