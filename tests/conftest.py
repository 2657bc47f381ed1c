import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU and the built HIP library")


def load_graph_fixture(path):
    """(SNNGraph, LIFParams, fixture dict) from a fwd_*.npz golden file."""
    from snc_amd.graph import SNNGraph
    from snc_amd.runtime import LIFParams

    z = dict(np.load(path, allow_pickle=False))
    g = SNNGraph(z["row_ptr"], z["post_ids"], z["weights"], z["delays"], z["sign"], z["role"],
                 z["channel"])
    lif = LIFParams(float(z["threshold"]), float(z["decay"]), float(z["reset"]),
                    int(z["refractory"]), 0)
    return g, lif, z


def unpack_raster(raster_words, n):
    """[T][W] uint32 bitmap -> [T][n] uint8."""
    r = np.ascontiguousarray(raster_words, dtype=np.uint32)
    bits = np.unpackbits(r.view(np.uint8), axis=1, bitorder="little")
    return bits[:, :n]


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN
