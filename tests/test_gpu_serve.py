"""Batched single-point serving path vs the FP64 oracle and vs itself
(each structure alone and in a mixed batch), through the C ABI (needs an
MI355X).

Tolerances: FP32 path vs the FP64 oracle as for the periodic system
(max-abs force 1e-5 eV/A, energy 1e-7 eV/atom, virial 1e-6 eV/atom, every
atomic energy within 1e-6 eV).  The energy bar per atom is 1e-7 for
structures of 100 atoms or more; below that it is 1e-6/sqrt(n): atomic
energies here are ~1.5 eV, so each carries an FP32 rounding error of ~1e-7
that is independent between atoms and only averages out over many atoms.  Alone
vs in the mix: energies and virials are fixed-order FP64 sums of per-atom
values that do not depend on the batch, so they must agree to 1e-12 relative;
forces are FP64 atomic sums (order varies) and must agree to 1e-10 eV/A.
"""
import sys

import numpy as np
import pytest

import oracle
from tensormd_amd import TensorMD, TmdError, workloads as W

pytestmark = pytest.mark.gpu

SYMBOLS = ["Mo", "Nb", "Ta", "W"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _structures(count, seed, nmax=1200):
    return W.serving_requests(count, seed=seed, nmax=nmax)


def test_serve_fp32_vs_oracle(synth_models):
    say("case: serving FP32 vs oracle, 12 mixed structures incl. 16-atom and triclinic cells")
    reqs = _structures(10, seed=41, nmax=600)
    # a tilted cell and the smallest BCC supercell (multi-image neighbors)
    x, t, cell = W.crystal("bcc", 3.2, (3, 4, 3), nelt=4, sigma=0.1, seed=42, equiatomic=False)
    cell = cell.copy()
    cell[1, 0] += 1.1
    cell[2, 1] -= 0.6
    reqs.append((W.wrap(x, cell), t, cell))
    reqs.append(W.crystal("bcc", 3.2, (2, 2, 2), nelt=4, sigma=0.1, seed=43, equiatomic=False))
    m = TensorMD(synth_models["c2"], SYMBOLS, "fp32", 0.001)
    srv = m.serve(max_batch_atoms=8192)
    tickets = [srv.submit(x, t, c, eatom=True) for x, t, c in reqs]
    res = [srv.wait(tk) for tk in tickets]
    st = srv.stats()
    srv.stop()
    say("stats", st)
    om = oracle.OracleModel(synth_models["c2"], 0.001)
    for k, ((x, t, c), r) in enumerate(zip(reqs, res)):
        assert r["status"] == 0
        ref = om.compute_periodic(x, t - 1, c)
        n = len(x)
        dF = float(abs(r["f"] - ref["f"]).max())
        dE = abs(r["energy"] - ref["energy"]) / n
        dV = float(abs(r["virial"] - ref["virial"]).max()) / n
        dEa = float(abs(r["eatom"] - ref["eatom"]).max())
        say(f"struct {k}: n={n} batch={r['batch_structs']} dF={dF:.2e} dE={dE:.2e} dV={dV:.2e}")
        assert dF <= 1e-5
        assert dE <= max(1e-7, 1e-6 / np.sqrt(n))
        assert dV <= 1e-6
        # individual atomic energies (|E_i| <~ 1 eV here, FP32 ulp 1.2e-7)
        # carry the accumulated FP32 rounding of the 384-256-128-64 MLP,
        # ~10 ulp; the structure energy is held to 1e-7 eV/atom above
        assert dEa <= 2e-6


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp32"])
def test_serve_alone_equals_mixed(synth_models, precision):
    say("case: serving alone vs in the mix,", precision)
    probe = _structures(6, seed=51, nmax=2000)
    fill = _structures(60, seed=52, nmax=2000)
    m = TensorMD(synth_models["c2"], SYMBOLS, precision, 0.001)
    srv = m.serve(max_batch_atoms=65536)
    alone = []
    for x, t, c in probe:  # one at a time: each runs in a batch of its own
        r = srv.wait(srv.submit(x, t, c))
        assert r["status"] == 0 and r["batch_structs"] == 1
        alone.append(r)
    # all at once, probes interleaved with other structures
    mix = []
    pi = 0
    for k, item in enumerate(fill):
        mix.append((item, None))
        if k % 10 == 5 and pi < len(probe):
            mix.append((probe[pi], pi))
            pi += 1
    tickets = [srv.submit(x, t, c) for (x, t, c), _ in mix]
    order = [o for _, o in mix]
    mixed = {}
    for tk, o in zip(tickets, order):
        r = srv.wait(tk)
        assert r["status"] == 0
        if o is not None:
            mixed[o] = r
    st = srv.stats()
    srv.stop()
    say("stats", st)
    assert len(mixed) == len(probe)
    assert max(r["batch_structs"] for r in mixed.values()) > 1
    for p, a in enumerate(alone):
        b = mixed[p]
        say(f"probe {p}: n={len(probe[p][0])} mixed batch={b['batch_structs']} structs "
            f"dE={abs(a['energy'] - b['energy']):.1e} dF={abs(a['f'] - b['f']).max():.1e}")
        assert abs(a["energy"] - b["energy"]) <= 1e-12 * max(1.0, abs(a["energy"]))
        assert abs(a["virial"] - b["virial"]).max() <= 1e-12 * max(1.0, abs(a["virial"]).max())
        assert abs(a["f"] - b["f"]).max() <= 1e-10


def test_serve_graph_matches_eager_and_system(synth_models):
    say("case: serving graph vs eager vs tmd_system (bf16)")
    reqs = _structures(8, seed=61, nmax=1500)
    m = TensorMD(synth_models["c2"], SYMBOLS, "bf16", 0.001)
    out = []
    for graph in (True, False):
        srv = m.serve(use_graph=graph, max_batch_atoms=16384)
        tickets = [srv.submit(x, t, c) for x, t, c in reqs]
        out.append([srv.wait(tk) for tk in tickets])
        srv.stop()
    for a, b in zip(*out):
        assert abs(a["energy"] - b["energy"]) <= 1e-12 * max(1.0, abs(a["energy"]))
        assert abs(a["f"] - b["f"]).max() <= 1e-10
    # the same structure as a device-resident system (skin 0), FP32: its bins
    # differ, so the descriptor sums run in another order (FP32 rounding)
    m32 = TensorMD(synth_models["c2"], SYMBOLS, "fp32", 0.001)
    srv = m32.serve()
    x, t, c = reqs[0]
    r = srv.wait(srv.submit(x, t, c))
    srv.stop()
    s = m32.system(x, t, c, skin=0.0)
    ref = s.compute()
    s.close()
    assert abs(r["energy"] - ref["energy"]) / len(x) <= 1e-7
    assert abs(r["f"] - ref["f"]).max() <= 2e-6
    assert abs(r["virial"] - ref["virial"]).max() / len(x) <= 1e-6


def test_serve_rejections(synth_models):
    say("case: serving argument errors")
    m = TensorMD(synth_models["c2"], SYMBOLS, "bf16", 0.001)
    srv = m.serve(max_batch_atoms=2048)
    x, t, c = W.crystal("bcc", 3.2, (1, 2, 2), nelt=4, sigma=0.0, seed=1)
    with pytest.raises(TmdError) as e:
        srv.submit(x, t, c)  # 3.2 A wide: narrower than rc/2
    assert e.value.code == -1
    x, t, c = W.crystal("bcc", 3.2, (10, 10, 11), nelt=4, seed=2)
    with pytest.raises(TmdError) as e:
        srv.submit(x, t, c)  # 2200 atoms > slot capacity
    assert e.value.code == -4
    x, t, c = W.crystal("bcc", 3.2, (2, 2, 2), nelt=4, seed=3)
    t = t.copy()
    t[0] = 7
    with pytest.raises(TmdError) as e:
        srv.submit(x, t, c)  # unmapped type
    assert e.value.code == -1
    # the loop still serves after the rejections
    x, t, c = W.crystal("bcc", 3.2, (2, 2, 2), nelt=4, sigma=0.1, seed=4)
    r = srv.wait(srv.submit(x, t, c))
    assert r["status"] == 0
    srv.stop()
