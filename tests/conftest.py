import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

SCENES = os.path.join(ROOT, "tests", "golden", "scenes")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (runs on the MI355X box)")


def scene_text(name):
    with open(os.path.join(SCENES, name)) as f:
        return f.read()


@pytest.fixture(scope="session")
def gpu_ctx_factory():
    import interflect_amd as ia

    def make(scene_name=None, text=None):
        ctx = ia.Context(0)
        ctx.load_scene(text if text is not None else scene_text(scene_name))
        return ctx

    return make
