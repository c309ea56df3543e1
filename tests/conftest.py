import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; run on the GPU box")


def _ensure_built():
    from retrocam_amd import _native as N
    from oracle import rco
    missing = [p for p in (N.HIP_LIB_PATH, N.HOST_LIB_PATH) if not os.path.exists(p)]
    if missing:
        from retrocam_amd import build
        build.build_all()
    if not os.path.exists(rco.LIB_PATH):
        rco.build()


_ensure_built()


@pytest.fixture(scope="session")
def gpu_ctx():
    from retrocam_amd.renderer import RenderContext
    ctx = RenderContext(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def cornell():
    from retrocam_amd import workloads
    return workloads.cornell_small()


@pytest.fixture(scope="session")
def full_frame_scene():
    from retrocam_amd import workloads
    return workloads.full_frame()
