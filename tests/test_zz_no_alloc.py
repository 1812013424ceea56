"""tmd_compute_list_noalloc (fixed device scratch, no allocation in the
call) against tmd_compute_list on the same inputs, through the C ABI.

Two calls in a row, the second on a smaller system, so that the second
call runs on scratch that still holds the first call's data.  Both entry
points run the same kernels: energies and per-atom energies are equal to
the bit (fixed-order sums); forces and per-atom virials are accumulated
with FP64 atomics whose order varies, so they agree to FP64 rounding.
"""
import sys

import numpy as np
import pytest

import oracle
from tensormd_amd import TensorMD, TmdError, workloads as W

pytestmark = pytest.mark.gpu

F_TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _list_case(om, cells, seed):
    x, t, cell = W.crystal("bcc", 3.2, cells, nelt=4, sigma=0.08, seed=seed)
    xall, owner, ilist, numneigh, flat = W.lammps_ghost_list(x, cell, om.rc + 1.0)
    return len(x), xall, t[owner], ilist, numneigh, flat


def test_noalloc_matches_compute_list(synth_models):
    om = oracle.OracleModel(synth_models["c2"], 0.001)
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "fp32", 0.001)
    for cells, seed in [((4, 4, 4), 81), ((3, 3, 3), 82)]:
        say("case: noalloc vs compute_list, C2 model, bcc", cells)
        n, xall, tall, ilist, numneigh, flat = _list_case(om, cells, seed)
        base = m.compute_list(n, xall, tall, ilist, numneigh, flat)
        got = m.compute_list(n, xall, tall, ilist, numneigh, flat, noalloc=True)
        assert got["energy"] == base["energy"]
        assert np.array_equal(got["eatom"], base["eatom"])
        assert abs(got["f"] - base["f"]).max() <= 1e-10
        assert abs(got["vatom"] - base["vatom"]).max() <= 1e-10
        assert abs(got["virial"] - base["virial"]).max() <= 1e-9 * max(1.0, abs(base["virial"]).max())
        ref = om.compute_list(n, xall, tall - 1, ilist, numneigh, flat)
        df = float(abs(got["f"] - ref["f"]).max())
        de = abs(got["energy"] - ref["energy"]) / n
        say("  vs oracle: max|dF|", df, "|dE|/atom", de)
        assert df <= F_TOL and de <= 1e-7


def test_noalloc_refuses_beyond_its_scratch(synth_models):
    """30,000 isolated atoms of a 4-element k = 24 model need 92 MB of
    descriptor rows: more than the 64 MiB scratch, refused before any
    device work; the context stays usable."""
    say("case: noalloc capacity refusal")
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "fp32", 0.001)
    n = 30000
    x = np.zeros((n, 3))
    x[:, 0] = np.arange(n) * 20.0
    t = (np.arange(n) % 4 + 1).astype(np.int32)
    ilist = np.arange(n, dtype=np.int32)
    numneigh = np.zeros(n, dtype=np.int32)
    flat = np.zeros(0, dtype=np.int32)
    with pytest.raises(TmdError) as ei:
        m.compute_list(n, x, t, ilist, numneigh, flat, noalloc=True)
    assert ei.value.code == -5 and "scratch" in str(ei.value)
    # the same system through the allocating entry point still works
    got = m.compute_list(n, x, t, ilist, numneigh, flat)
    assert np.isfinite(got["energy"])
