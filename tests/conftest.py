import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


@pytest.fixture(autouse=True)
def _name_the_case(request):
    # name the case before any GPU work so a fault can be attributed
    if request.node.get_closest_marker("gpu") is not None:
        print(f"\n[gpu case] {request.node.nodeid}", flush=True)
    yield


@pytest.fixture(scope="session")
def ctx():
    from tsc_tri_t1t2j1j2_amd import capi
    c = capi.Context(0)
    print(f"[gpu] MFMA accumulator layouts (f64, f32, bf16): {c.mfma_layout()}", flush=True)
    yield c
