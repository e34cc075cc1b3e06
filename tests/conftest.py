import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def c1_setup():
    from gemini_unified_amd import ExternalRadialNodes, global_model
    nem = global_model(1, 200)
    nodes = ExternalRadialNodes.from_blocks(nem, [1], [30000.0])
    return nem, nodes, nodes.node_index(nem.rearth)
