import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the C-ABI kernels")
    config.addinivalue_line("markers", "slow: long CPU test")


@pytest.fixture(scope="session")
def oracle():
    import oracle.binding as ob
    ob.load()
    return ob


@pytest.fixture(scope="session")
def krr():
    import kiraray_amd
    kiraray_amd.load_library()
    return kiraray_amd


@pytest.fixture(scope="session")
def gpu():
    """HIP runtime helpers of the tracer library (one runtime per process)."""
    from kiraray_amd import hip
    if hip.device_count() == 0:
        pytest.fail("GPU test collected but no HIP device is visible")
    hip.check(hip.runtime().hipSetDevice(0))
    return hip


CORNELL = {"model": [{"type": "cornell_box"}], "resolution": [256, 256],
           "camera": {"eye": [0, 1, 2.8], "target": [0, 1, 0]}}


def icosphere_scene(count=1024, level=3, res=(1920, 1080), seed=7, media=False, grid_res=256):
    cfg = {"model": [{"type": "icosphere_field", "count": count, "level": level, "seed": seed}],
           "resolution": list(res),
           "camera": {"eye": [0.5, 0.75, 2.3], "target": [0.5, 0.3, 0.45]}}
    if media:
        cfg["media"] = [{"type": "value_noise_grid", "res": grid_res, "octaves": 5, "seed": 11,
                         "sigma_t_max": 20.0, "albedo": 0.8, "g": 0.3}]
        cfg["camera"]["medium"] = 0
    return cfg
