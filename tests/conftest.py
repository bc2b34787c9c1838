"""Shared fixtures.  `gpu`-marked tests need an MI355X; everything else runs on CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def golden(name):
    """A known-answer fixture from tests/golden (written by make_kats.py)."""
    import json
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def _ensure_built():
    from mambatracer_amd import _lib
    from oracle import orc_py
    if not os.path.exists(_lib.LIB_PATH):
        from mambatracer_amd import build
        build.build()
    if not os.path.exists(orc_py.LIB):
        orc_py.build()


_ensure_built()


@pytest.fixture(scope="session")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")
    from mambatracer_amd import api
    c = api.Context(0)
    yield c
    c.sync()
    c.close()


def c0_scene_arrays():
    """Config 0: room box [-1,1]^3 (12 tris) + 20 000 random triangles, seed 1."""
    from mambatracer_amd import api
    rp, rf = api.gen_room(0)
    sp, sf = api.gen_soup(1, 20000, -0.8, 0.8, 0.01, 0.05, first_vertex=8, mat=0)
    points = np.concatenate([rp, sp])
    faces = np.concatenate([rf, sf])
    nodes, bounds, depth = api.CNormalBVH.build(points, faces)
    return points, faces, nodes, bounds, depth


def c0_rays(w=512, h=512):
    from mambatracer_amd import api
    return api.gen_camera_rays([0, 0, -0.95], [0, 0, 1], [0, 1, 0], 60.0, w, h)


@pytest.fixture(scope="session")
def c0():
    return c0_scene_arrays()
