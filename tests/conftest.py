import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
for p in (HERE, os.path.join(HERE, "golden"), os.path.join(REPO, "graphical-edmd_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(HERE, "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libedmd_hip.so)")


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def golden_state(g):
    return {"x": g["x"], "y": g["y"], "vx": g["vx"], "vy": g["vy"], "rad": g["rad"],
            "m": g["m"], "Lx": float(g["Lx"]), "Ly": float(g["Ly"])}


TRACED = ["sq2000_phi070", "sq2000_phi070_r1", "n400_phi060", "tri1024_phi080", "poly500_phi050",
          "tiny_ncx3"]
SWEEP_ONLY = ["tiny_ncx2", "tiny_ncx1"]
GROWN = ["grow_mono200", "grow_poly150"]
RESCALED = ["rescale_n400", "rescale_res09_n400", "rescale_grow_poly400"]


def rescale_params(g):
    """Oracle / engine keyword arguments of a thermostat golden vector."""
    return dict(noise=2, dtnoise=float(g["dtnoise"]), temperature=float(g["temperature"]),
                vr=float(g["vr"]), res=float(g["res"]))


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="session")
def oracle():
    import oracle_bind
    if not os.path.exists(oracle_bind.LIB_PATH):
        pytest.skip("oracle not built (make -C oracle oracle)")
    return oracle_bind


@pytest.fixture(scope="session")
def gpu():
    import edmd_hip
    if edmd_hip.device_count() < 1:
        pytest.fail("no GPU visible to libedmd_hip.so")
    return edmd_hip
