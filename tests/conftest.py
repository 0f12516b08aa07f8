import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP solver)")


def _built(path):
    return os.path.exists(os.path.join(ROOT, path))


@pytest.fixture(scope="session")
def built_libs():
    missing = [p for p in ("multiviewdynrecon_amd/libmvdr_solver.so", "multiviewdynrecon_amd/libmvdr_scene.so",
                           "oracle/libmvdr_oracle.so") if not _built(p)]
    if missing:
        from multiviewdynrecon_amd import _build

        _build.build_all()
    return True


@pytest.fixture(scope="session")
def scene0(built_libs):
    from multiviewdynrecon_amd.scene import make_scene

    return make_scene(0)


@pytest.fixture(scope="session")
def scene1(built_libs):
    from multiviewdynrecon_amd.scene import make_scene

    return make_scene(1)
