"""Shared fixtures.  GPU tests are marked `gpu`; everything else runs on the CPU.
The oracle (oracle/) is imported here only as the checker."""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

GOLDEN_DIR = ROOT / "tests" / "golden"
GRAPH_DIR = GOLDEN_DIR / "graphs"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgraphalgo.so)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


@pytest.fixture(scope="session")
def golden():
    return json.loads((GOLDEN_DIR / "golden.json").read_text())


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def G():
    from graphalgo_amd import graphs
    return graphs


def small_graph_names():
    return ["G43.mtx", "iris_dataset_30NN.mtx", "yeast_30NN.mtx.gz", "cage3.mtx"]


@pytest.fixture(scope="session")
def small_graphs(G):
    """name -> (graph as read, graph as the reference's CLI feeds TC/k-truss)."""
    out = {}
    for name in small_graph_names():
        g = G.read_mtx(str(GRAPH_DIR / name))
        out[name] = (g, G.symm_from_lower(g))
    return out


@pytest.fixture(scope="session")
def gpu():
    """Initialises the GPU library (GPU tests only)."""
    from graphalgo_amd import api, _native
    _native.init(0)
    return api
