import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libpthip.so")


@pytest.fixture(scope="session")
def scene_dir(tmp_path_factory):
    d = os.environ.get("PT_SCENE_DIR")
    if d:
        return d
    return str(tmp_path_factory.mktemp("scenes"))


@pytest.fixture(scope="session")
def box_scn(scene_dir):
    from pt_amd.scenes import generate
    return generate.cached_scene("box", root=scene_dir, seed=1)
