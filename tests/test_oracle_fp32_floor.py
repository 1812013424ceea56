"""The oracle's reduced-precision emulation (oracle/tmd_oracle.cpp,
atom_chain_emul): the same chain with chosen stages in FP32.

It prices the rounding floor of an FP32 evaluation, against which the GPU
FP32 path's errors on the reference's trained FP64 models are measured
(tests/test_gpu_parity.py).  Pinned here by:
  * with every stage in FP64 it equals the FP64 chain (energies, forces,
    virial, descriptor rows and dE/dG) to rounding, so it runs the same
    algorithm;
  * on the benchmark-shaped tanh models its FP32 floor sits below the
    absolute bars (1e-7 eV/atom, 1e-5 eV/A), which the GPU FP32 path meets;
  * on the reference's W models it does not: their FP32 floor is the
    error the GPU shows.
"""
import numpy as np
import pytest

import oracle
from conftest import model_path
from tensormd_amd import workloads as W

ALL_FP32 = (oracle.OracleModel.EMUL_DESC | oracle.OracleModel.EMUL_MLP
            | oracle.OracleModel.EMUL_FORCE)


def _errs(a, ref, n):
    return dict(dF=float(abs(a["f"] - ref["f"]).max()),
                dE=abs(a["energy"] - ref["energy"]) / n,
                dEi=float(abs(a["eatom"] - ref["eatom"]).max()),
                dV=float(abs(a["virial"] - ref["virial"]).max()) / n)


def _run(om, x, t, cell, flags, want_desc=False):
    om.set_emulation(flags)
    try:
        return om.compute_periodic(x, t - 1, cell, want_desc=want_desc)
    finally:
        om.set_emulation(0)


@pytest.mark.parametrize("which", ["c2", "sq2", "sp3", "W_k32_snapshot.npz"])
def test_emulation_in_fp64_is_the_oracle_chain(synth_models, which):
    if which.endswith(".npz"):
        om = oracle.OracleModel(model_path(which), 0.001)
        x, t, cell = W.crystal("bcc", 3.189, (3, 3, 3), sigma=0.05, seed=3)
    else:
        om = oracle.OracleModel(synth_models[which], 0.001)
        x, t, cell = W.crystal("bcc", 3.2, (3, 3, 3), nelt=om.nelt, sigma=0.08, seed=4)
    ref = om.compute_periodic(x, t - 1, cell, want_desc=True)
    emu = _run(om, x, t, cell, oracle.OracleModel.EMUL_FORCE_TEMPLATE, want_desc=True)
    scale_f = max(1.0, float(abs(ref["f"]).max()))
    assert abs(emu["f"] - ref["f"]).max() <= 1e-11 * scale_f
    assert abs(emu["eatom"] - ref["eatom"]).max() <= 1e-11 * max(1.0, abs(ref["eatom"]).max())
    assert abs(emu["virial"] - ref["virial"]).max() <= 1e-10 * max(1.0, abs(ref["virial"]).max())
    assert abs(emu["G"] - ref["G"]).max() <= 1e-11 * max(1.0, abs(ref["G"]).max())
    assert abs(emu["dEdG"] - ref["dEdG"]).max() <= 1e-10 * max(1.0, abs(ref["dEdG"]).max())


def test_emulation_stages_change_results(synth_models):
    """Each FP32 stage moves its own outputs: the descriptor stage moves G,
    the MLP stage moves energies and dE/dG but not G, the force stage moves
    forces only (compared with the same emulated chain in FP64)."""
    om = oracle.OracleModel(synth_models["c2"], 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (3, 3, 3), nelt=4, sigma=0.08, seed=5)
    ref = _run(om, x, t, cell, oracle.OracleModel.EMUL_FORCE_TEMPLATE, want_desc=True)
    d = _run(om, x, t, cell, oracle.OracleModel.EMUL_DESC, want_desc=True)
    m = _run(om, x, t, cell, oracle.OracleModel.EMUL_MLP, want_desc=True)
    f = _run(om, x, t, cell, oracle.OracleModel.EMUL_FORCE, want_desc=True)
    assert 0 < abs(d["G"] - ref["G"]).max() <= 1e-5 * abs(ref["G"]).max()
    assert abs(m["G"] - ref["G"]).max() == 0.0
    assert 0 < abs(m["eatom"] - ref["eatom"]).max() <= 1e-5
    assert abs(f["eatom"] - ref["eatom"]).max() == 0.0
    assert 0 < abs(f["f"] - ref["f"]).max() <= 1e-5


def test_fp32_floor_of_benchmark_models_is_below_the_bars(synth_models):
    """The C1- and C2-shaped tanh models: an FP32 evaluation can hold the
    absolute bars (1e-7 eV/atom, 1e-5 eV/A)."""
    for key, lat, a, cells, nelt in [("c1", "fcc", 3.615, (3, 3, 3), 1),
                                     ("c2", "bcc", 3.2, (4, 4, 4), 4)]:
        om = oracle.OracleModel(synth_models[key], 0.001)
        x, t, cell = W.crystal(lat, a, cells, nelt=nelt, sigma=0.05, seed=6)
        ref = om.compute_periodic(x, t - 1, cell)
        e = _errs(_run(om, x, t, cell, ALL_FP32), ref, len(x))
        assert e["dE"] <= 1e-7 and e["dF"] <= 1e-5, (key, e)


def test_fp32_floor_of_reference_w_models_exceeds_the_energy_bar():
    """The reference's trained W models (|E| ~ 13 eV/atom) on their logged
    perfect BCC lattice: every atom sees the same environment, so the FP32
    rounding of its energy does not average out, and an FP32 evaluation
    misses 1e-7 eV/atom by an order of magnitude.  The rounding enters in
    the FP32 sums (P = sum H M over ~58 neighbors; the hidden GEMMs), not
    in the rounding of stored weights, biases or activations."""
    x, cell = W.lattice("bcc", 3.189, (3, 3, 3))
    t = np.ones(len(x), dtype=np.int32)
    for name in ["W_k32_snapshot.npz", "W_m3_k64_n128_squareplus.npz"]:
        om = oracle.OracleModel(model_path(name), 0.001)
        ref = om.compute_periodic(x, t - 1, cell)
        e = _errs(_run(om, x, t, cell, ALL_FP32), ref, len(x))
        assert 5e-7 <= e["dE"] <= 1e-5, (name, e)
