import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def renderer():
    from path_amd.renderer import Renderer
    return Renderer(0)


@pytest.fixture(scope="session")
def cornell(tmp_path_factory):
    from path_amd import scenes
    from path_amd.renderer import Scene
    d = tmp_path_factory.mktemp("cornell")
    scene, ftw = Scene.load(scenes.cornell_box(d), 1.0)
    return scene, ftw


@pytest.fixture(scope="session")
def cornell_sphere(tmp_path_factory):
    """Cornell box + icosphere level 4 (5,120 triangles) for fast parity runs."""
    from path_amd import scenes
    from path_amd.renderer import Scene
    d = tmp_path_factory.mktemp("cornell_sphere")
    scene, ftw = Scene.load(scenes.cornell_icosphere(d, level=4, seed=1), 1.0)
    return scene, ftw


@pytest.fixture(scope="session")
def benchmark_scene(tmp_path_factory):
    """The throughput benchmark's scene: Cornell box + icosphere level 6
    (81,956 triangles)."""
    from path_amd import scenes
    from path_amd.renderer import Scene
    d = tmp_path_factory.mktemp("benchmark_scene")
    scene, ftw = Scene.load(scenes.cornell_icosphere(d, level=6, seed=0), 1.0)
    return scene, ftw
