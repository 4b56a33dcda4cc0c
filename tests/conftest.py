import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the native HIP library)")
    config.addinivalue_line("markers", "slow: full-size case, still bounded to well under a minute")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    o._load()
    return o


@pytest.fixture(scope="session")
def capi():
    from psrdada_cpp_amd import capi as c
    c.lib()
    return c


@pytest.fixture(scope="session")
def gpu(capi):
    """Selects device 0; fails loudly if no GPU is visible (no CPU fallback)."""
    import ctypes
    n = ctypes.c_int()
    capi.call("fbf_device_count", ctypes.byref(n))
    assert n.value >= 1, "no GPU visible"
    capi.call("fbf_set_device", 0)
    name = ctypes.create_string_buffer(256)
    capi.call("fbf_device_name", name, 256)
    print("device:", name.value.decode(), flush=True)
    return capi


def rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def gaussian_int8(seed, n, sigma=16.0, clip=True):
    r = rng(seed).normal(0.0, sigma, size=n)
    v = np.rint(r).astype(np.int64)
    if clip:
        v = np.clip(v, -127, 127)
    return (v & 0xFF).astype(np.uint8).view(np.int8)


def unit_weights(seed, n):
    """round(127 e^{i theta}), theta ~ U(0, 2 pi) as char2 bytes (CoherentBeamformerTester.cu:172-179)."""
    th = rng(seed).uniform(0.0, 2 * np.pi, size=n)
    w = np.empty((n, 2), dtype=np.int8)
    w[:, 0] = np.round(127 * np.cos(th)).astype(np.int8)
    w[:, 1] = np.round(127 * np.sin(th)).astype(np.int8)
    return w.reshape(-1)
