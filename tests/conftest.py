import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and librtx_hip.so")


@pytest.fixture(scope="session")
def twin():
    import oracle
    oracle.lib()
    return oracle


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ren():
    """The product package; GPU tests require its HIP library to load (no fallback)."""
    import rendertoy_amd
    from rendertoy_amd import _native
    if not gpu_available():
        pytest.skip("no HIP device")
    _native.load()
    return rendertoy_amd


def c1_scene():
    """Workload C1: seeded icosphere f=64 (81,920 tris), positions only."""
    from rendertoy_amd import scenes
    d = scenes.icosphere(64, 1000)
    return [{"pos": d["pos"], "idx": d["idx"], "nrm": None, "uv": None}]


def c2_scene(f=256):
    """Workload C2: icosphere f=256 (1,310,720 tris) with N and UV, plus a 2-triangle ground
    plane; value-noise RGBA8 texture 1024^2."""
    from rendertoy_amd import scenes
    d = scenes.icosphere(f, 2000)
    g = scenes.ground_plane()
    tex = scenes.value_noise_rgba8(1024, 2001)
    return [d, g], tex
