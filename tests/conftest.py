import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "lammps-tensormd_amd"))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_EXAMPLES = "/root/reference/examples/tensormd"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def goldens():
    with open(os.path.join(GOLDEN, "reference_goldens.json")) as f:
        return json.load(f)


def model_path(name):
    """Golden model file: committed fixture, else the reference tree (CPU only)."""
    p = os.path.join(GOLDEN, name)
    if os.path.exists(p):
        return p
    for root, _, files in os.walk(REF_EXAMPLES):
        if name in files:
            return os.path.join(root, name)
    pytest.skip(f"model {name} not available")


@pytest.fixture(scope="session")
def synth_models(tmp_path_factory):
    """Seeded synthetic models with the shapes of the benchmark configs."""
    from tensormd_amd import workloads as W
    d = tmp_path_factory.mktemp("models")
    out = {}
    out["c1"] = W.write_model(str(d / "c1.npz"), symbols=["Cu"], masses=[63.546], rmax=6.0,
                              num_filters=16, layer_sizes=[128, 64, 1], actfn="tanh", seed=11,
                              radial_scale=0.2)
    out["c2"] = W.write_model(str(d / "c2.npz"), symbols=["Mo", "Nb", "Ta", "W"],
                              masses=[95.95, 92.906, 180.95, 183.84], rmax=6.5, num_filters=24,
                              layer_sizes=[256, 128, 64, 1], actfn="tanh", seed=12,
                              radial_scale=0.2)
    out["sq2"] = W.write_model(str(d / "sq2.npz"), symbols=["Al", "Ni"], masses=[26.98, 58.69],
                               rmax=5.5, num_filters=32, layer_sizes=[64, 64, 64, 1],
                               actfn="squareplus", use_resnet_dt=True, seed=13, radial_scale=0.2)
    out["sp3"] = W.write_model(str(d / "sp3.npz"), symbols=["A", "B", "C"],
                               masses=[10.0, 20.0, 30.0], rmax=5.0, num_filters=16,
                               layer_sizes=[32, 32, 1], actfn="softplus", use_resnet_dt=True,
                               seed=14, radial_scale=0.2)
    return out


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
