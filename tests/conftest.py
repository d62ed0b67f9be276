import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("grid-rt_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

# Scene seeds of the benchmark configurations (splitmix64 -> xoshiro256**).
SEED_C1 = 0x6772740000000001
SEED_C2 = 0x6772740000000002


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu_ctx():
    import grt
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    ctx = grt.Context(0)
    yield ctx
    ctx.synchronize()


@pytest.fixture(scope="session")
def c1_scene():
    """Config 1 geometry: 64 UV-spheres x 256 triangles, seeded."""
    import grt
    v, n = grt.gen_spheres(SEED_C1, 64, 0.02, 0.08)
    return v, n


def c1_camera(width=512, height=512):
    import grt
    return grt.camera_lookat((0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 60.0, width,
                             height)


def rng(seed=0):
    return np.random.default_rng(seed)
