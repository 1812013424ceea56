"""HIP path vs the FP64 oracle, through the C ABI (needs an MI355X).

Tolerances (FP32 path vs FP64 oracle): max-abs force error <= 1e-5 eV/A,
energy error <= 1e-7 eV/atom.  BF16 MLP path: max-abs force error
<= 5e-3 eV/A.  Integer work (the neighbor set) is compared bit-exactly.
"""
import sys

import numpy as np
import pytest

import oracle
from conftest import model_path
from tensormd_amd import TensorMD, workloads as W

pytestmark = pytest.mark.gpu

F_TOL = 1e-5
E_TOL = 1e-7


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _errs(gpu, ref, n):
    return dict(dF=float(abs(gpu["f"] - ref["f"]).max()),
                dE=abs(gpu["energy"] - ref["energy"]) / n,
                dV=float(abs(gpu["virial"] - ref["virial"]).max()) / n)


def _fp32_floor(om, x, types, cell, ref):
    """Errors of the oracle's FP32 emulation of the same chain (every stage
    in FP32) against its FP64 result: the rounding floor of any FP32
    evaluation of this model on this structure."""
    om.set_emulation(om.EMUL_DESC | om.EMUL_MLP | om.EMUL_FORCE)
    try:
        emu = om.compute_periodic(x, types - 1, cell)
    finally:
        om.set_emulation(0)
    e = _errs(emu, ref, len(x))
    e["dEi_rms"] = float(np.sqrt(np.mean((emu["eatom"] - ref["eatom"]) ** 2)))
    e["n"] = len(x)
    return e


def _floor_bars(fl):
    """Energy and force bars for the FP32 path on a model whose FP32 floor
    exceeds the absolute bars (the reference's trained FP64 models): the
    absolute bar, or the emulated floor where that is larger.  The device
    sums in another order, so its rounding differs from the emulation's:
    forces within twice the emulated max error; the energy per atom within
    twice the emulated one or, where the atomic energy errors are random
    (a jittered cell), three standard errors of their mean,
    3 rms(dE_i) / sqrt(N)."""
    e_bar = max(E_TOL, 2.0 * fl["dE"], 3.0 * fl["dEi_rms"] / np.sqrt(fl["n"]))
    return e_bar, max(F_TOL, 2.0 * fl["dF"])


def test_list_mode_fcc_c1(synth_models):
    say("case: list mode, C1 model, fcc 4x4x4 jittered")
    om = oracle.OracleModel(synth_models["c1"], 0.001)
    x, t, cell = W.crystal("fcc", 3.615, (4, 4, 4), sigma=0.05, seed=21)
    xall, owner, ilist, numneigh, flat = W.lammps_ghost_list(x, cell, om.rc + 1.0)
    tall = t[owner]
    ref = om.compute_list(len(x), xall, tall - 1, ilist, numneigh, flat)
    m = TensorMD(synth_models["c1"], ["Cu"], "fp32", 0.001)
    got = m.compute_list(len(x), xall, tall, ilist, numneigh, flat)
    e = _errs(got, ref, len(x))
    say("errors", e)
    assert e["dF"] <= F_TOL
    assert e["dE"] <= E_TOL
    assert abs(got["eatom"] - ref["eatom"]).max() <= 1e-6
    assert abs(got["vatom"] - ref["vatom"]).max() <= 1e-5
    # fdotr virial of the reference equals the pair-sum virial returned here
    assert e["dV"] <= 1e-6


def _sub_list(ilist, numneigh, flat, sub):
    """CSR neighbor data of the centers `sub` (in that order) from a full list."""
    off = np.concatenate([[0], np.cumsum(numneigh[ilist])])
    start = {int(i): int(off[k]) for k, i in enumerate(ilist)}
    nn = np.zeros_like(numneigh)
    nn[sub] = numneigh[sub]
    sflat = np.concatenate([flat[start[int(i)]:start[int(i)] + numneigh[i]] for i in sub])
    return np.asarray(sub, dtype=np.int32), nn, sflat


def test_list_mode_edge_cases(synth_models):
    """Pair::compute inputs beyond a plain full list: the scale factor, a
    partial shuffled ilist (eatom of the other atoms is zero), an isolated
    atom with no neighbor inside rc, and an empty ilist that must still
    overwrite every output buffer."""
    import ctypes
    from tensormd_amd import lib, _p, _dp, _ip
    say("case: list mode edge cases, C2 model, bcc 3x3x3 4 elements")
    om = oracle.OracleModel(synth_models["c2"], 0.001)
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "fp32", 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (3, 3, 3), nelt=4, sigma=0.08, seed=32)
    xall, owner, ilist, numneigh, flat = W.lammps_ghost_list(x, cell, om.rc + 1.0)
    tall = t[owner]
    n = len(x)
    say("case: scale 0.37")
    ref = om.compute_list(n, xall, tall - 1, ilist, numneigh, flat, scale=0.37)
    got = m.compute_list(n, xall, tall, ilist, numneigh, flat, scale=0.37)
    e = _errs(got, ref, n)
    say("errors", e)
    assert e["dF"] <= F_TOL and e["dE"] <= E_TOL and e["dV"] <= 1e-6
    assert abs(got["eatom"] - ref["eatom"]).max() <= 1e-6
    assert abs(got["vatom"] - ref["vatom"]).max() <= 1e-5
    say("case: partial shuffled ilist")
    sub = np.random.default_rng(6).permutation(n)[: n // 3]
    sl, snn, sflat = _sub_list(ilist, numneigh, flat, sub)
    ref = om.compute_list(n, xall, tall - 1, sl, snn, sflat)
    got = m.compute_list(n, xall, tall, sl, snn, sflat)
    e = _errs(got, ref, len(sub))
    say("errors", e)
    assert e["dF"] <= F_TOL and e["dE"] <= E_TOL and e["dV"] <= 1e-6
    rest = np.setdiff1d(np.arange(n), sub)
    assert np.array_equal(got["eatom"][rest], np.zeros(len(rest)))
    assert abs(got["eatom"][sub] - ref["eatom"][sub]).max() <= 1e-6
    say("case: isolated atom")
    xi = np.array([[0.0, 0.0, 0.0], [2.6, 0.3, -0.2], [30.0, 30.0, 30.0]])
    ti = np.array([1, 4, 2], dtype=np.int32)
    il = np.arange(3, dtype=np.int32)
    nn = np.array([1, 1, 0], dtype=np.int32)
    fl = np.array([1, 0], dtype=np.int32)
    ref = om.compute_list(3, xi, ti - 1, il, nn, fl)
    got = m.compute_list(3, xi, ti, il, nn, fl)
    e = _errs(got, ref, 3)
    say("errors", e)
    assert e["dF"] <= F_TOL and e["dE"] <= E_TOL
    assert np.array_equal(got["f"][2], np.zeros(3))
    assert abs(got["eatom"] - ref["eatom"]).max() <= 1e-6
    say("case: empty ilist, outputs pre-filled")
    xa, ta = _f64_c(xall), np.ascontiguousarray(tall, dtype=np.int32)
    f = np.full((len(xa), 3), 7.0)
    ea = np.full(n, 7.0)
    va = np.full((len(xa), 6), 7.0)
    vir = np.full(6, 7.0)
    en = ctypes.c_double(7.0)
    empty = np.zeros(1, dtype=np.int32)
    rc = lib().tmd_compute_list(m.h, n, len(xa), _p(xa, _dp), _p(ta, _ip), 0, _p(empty, _ip),
                                _p(empty, _ip), None, 1.0, _p(f, _dp), _p(ea, _dp), _p(va, _dp),
                                ctypes.byref(en), _p(vir, _dp))
    assert rc == 0
    assert en.value == 0.0
    for a in (f, ea, va, vir):
        assert np.array_equal(a, np.zeros_like(a))


def _f64_c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def test_model_file_variants(synth_models, tmp_path):
    """The same model stored deflate-compressed gives the same results (to
    the FP64 rounding of the atomically summed f and vatom, whose order of
    summation varies from run to run; eatom exactly); stored with float32
    arrays and `precision` 32 (the reference's float models) it matches the
    oracle reading that same file."""
    say("case: model file variants (deflate, float32 arrays), C1 model, fcc 4x4x4")
    src = np.load(synth_models["c1"])  # allow_pickle stays False
    arrays = {k: src[k] for k in src.files}
    comp = str(tmp_path / "c1_deflate.npz")
    np.savez_compressed(comp, **arrays)
    f32 = str(tmp_path / "c1_f32.npz")
    a32 = {k: (v.astype(np.float32) if v.dtype == np.float64 else v) for k, v in arrays.items()}
    a32["precision"] = np.int32(32)
    np.savez(f32, **a32)
    x, t, cell = W.crystal("fcc", 3.615, (4, 4, 4), sigma=0.05, seed=24)
    xall, owner, ilist, numneigh, flat = W.lammps_ghost_list(x, cell, 7.0)
    tall = t[owner]
    base = TensorMD(synth_models["c1"], ["Cu"], "fp32", 0.001).compute_list(len(x), xall, tall, ilist,
                                                                            numneigh, flat)
    got = TensorMD(comp, ["Cu"], "fp32", 0.001).compute_list(len(x), xall, tall, ilist, numneigh, flat)
    assert np.array_equal(got["eatom"], base["eatom"])
    for k in ("f", "vatom", "virial"):
        assert np.allclose(got[k], base[k], rtol=0, atol=1e-12), k
    assert abs(got["energy"] - base["energy"]) <= 1e-12 * abs(base["energy"])
    om = oracle.OracleModel(f32, 0.001)
    ref = om.compute_list(len(x), xall, tall - 1, ilist, numneigh, flat)
    got = TensorMD(f32, ["Cu"], "fp32", 0.001).compute_list(len(x), xall, tall, ilist, numneigh, flat)
    e = _errs(got, ref, len(x))
    say("float32 file errors", e)
    assert e["dF"] <= F_TOL and e["dE"] <= E_TOL and e["dV"] <= 1e-6


@pytest.mark.parametrize("mm", [3, 1])
def test_descriptors_list(tmp_path, mm):
    """compute tensormd/density over a LAMMPS full list: the descriptor row of
    each listed atom in the model file's layout, zero rows for the others,
    against the oracle's G (FP32 descriptor: 1e-5 of the largest |G|)."""
    say(f"case: list-mode descriptors, max_moment {mm}, 4 elements, bcc 3x3x3")
    path = W.write_model(str(tmp_path / f"d{mm}.npz"), symbols=["Mo", "Nb", "Ta", "W"],
                         masses=[95.95, 92.906, 180.95, 183.84], rmax=6.5, num_filters=24,
                         layer_sizes=[64, 32, 1], actfn="tanh", max_moment=mm, seed=50 + mm,
                         radial_scale=0.2)
    om = oracle.OracleModel(path, 0.001)
    m = TensorMD(path, ["Mo", "Nb", "Ta", "W"], "fp32", 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (3, 3, 3), nelt=4, sigma=0.08, seed=51)
    xall, owner, ilist, numneigh, flat = W.lammps_ghost_list(x, cell, om.rc + 1.0)
    tall = t[owner]
    n = len(x)
    sub = np.random.default_rng(7).permutation(n)[: n // 2]
    for il, nn, fl in ((ilist, numneigh, flat), _sub_list(ilist, numneigh, flat, sub)):
        ref = om.compute_list(n, xall, tall - 1, il, nn, fl, want_desc=True)["G"]
        got = m.descriptors_list(n, xall, tall, il, nn, fl)
        assert got.shape == (n, 4 * 24 * (mm + 1))
        scale = max(1.0, float(abs(ref).max()))
        err = float(abs(got[il] - ref).max())
        say("rows", len(il), "max |dG|", err, "max |G|", scale)
        assert err <= 1e-5 * scale
        rest = np.setdiff1d(np.arange(n), il)
        assert np.array_equal(got[rest], np.zeros((len(rest), got.shape[1])))


def test_list_mode_serial_reuses_the_list(synth_models):
    """tmd_compute_list_serial: while the serial repeats, only x is uploaded.
    Positions moved within the skin with the same list give what
    tmd_compute_list gives; a new serial repacks; a plain call in between
    makes the next serial call repack even with the old serial."""
    say("case: list serial, C2 model, bcc 4x4x4 and 3x3x3")
    om = oracle.OracleModel(synth_models["c2"], 0.001)
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "fp32", 0.001)
    rng = np.random.default_rng(91)

    def case(cells, seed):
        x, t, cell = W.crystal("bcc", 3.2, cells, nelt=4, sigma=0.08, seed=seed)
        xall, owner, ilist, numneigh, flat = W.lammps_ghost_list(x, cell, om.rc + 1.0)
        return len(x), xall, t[owner], ilist, numneigh, flat

    def same(a, b):
        assert a["energy"] == b["energy"] and np.array_equal(a["eatom"], b["eatom"])
        assert abs(a["f"] - b["f"]).max() <= 1e-10 and abs(a["vatom"] - b["vatom"]).max() <= 1e-10

    A = case((4, 4, 4), 92)
    B = case((3, 3, 3), 93)
    n, xall, tall, ilist, numneigh, flat = A
    same(m.compute_list(*A, serial=7), m.compute_list(*A))
    for step in range(3):  # same list, moved positions: list stays on the device
        xm = xall + rng.normal(0.0, 0.05, xall.shape)
        got = m.compute_list(n, xm, tall, ilist, numneigh, flat, serial=8)
        base = m.compute_list(n, xm, tall, ilist, numneigh, flat)
        same(got, base)
        if step == 2:
            ref = om.compute_list(n, xm, tall - 1, ilist, numneigh, flat)
            assert abs(got["f"] - ref["f"]).max() <= F_TOL
            assert abs(got["energy"] - ref["energy"]) / n <= E_TOL
    # a different system under a new serial
    same(m.compute_list(*B, serial=10), m.compute_list(*B))
    # a plain call on another list of the same shape (half of every atom's
    # neighbors) between two calls with one serial: the second call repacks
    starts = np.concatenate([[0], np.cumsum(numneigh)[:-1]])
    half = numneigh // 2
    flat_h = np.concatenate([flat[s:s + h] for s, h in zip(starts, half)]).astype(np.int32)
    H = (n, xall, tall, ilist, half, flat_h)
    full = m.compute_list(*A, serial=12)
    halved = m.compute_list(*H)
    assert abs(halved["energy"] - full["energy"]) > 1e-3
    same(m.compute_list(*A, serial=12), full)


def test_list_mode_rejections(synth_models):
    """Malformed Pair::compute inputs are refused with TMD_ERR_ARG before any
    device work: an unmapped type, a neighbor index past nall, nall < nlocal,
    and an ilist entry past nlocal."""
    import ctypes
    from tensormd_amd import TmdError, lib, _p, _dp, _ip
    say("case: list mode argument errors")
    m = TensorMD(synth_models["c1"], ["Cu"], "fp32", 0.001)
    x, t, cell = W.crystal("fcc", 3.615, (2, 2, 2), sigma=0.05, seed=23)
    xall, owner, ilist, numneigh, flat = W.lammps_ghost_list(x, cell, 7.0)
    tall = t[owner].astype(np.int32)
    n = len(x)
    bad_t = tall.copy()
    bad_t[-1] = 2  # one type past the single mapped type
    bad_flat = flat.copy()
    bad_flat[0] = len(xall)
    for what, args in (("unmapped type", (n, xall, bad_t, ilist, numneigh, flat)),
                       ("neighbor past nall", (n, xall, tall, ilist, numneigh, bad_flat)),
                       ("nall < nlocal", (len(xall) + 1, xall, tall, ilist,
                                          np.concatenate([numneigh, np.zeros(len(xall) + 1 - n,
                                                                             dtype=numneigh.dtype)]),
                                          flat))):
        say("case:", what)
        with pytest.raises(TmdError) as e:
            m.compute_list(*args)
        assert e.value.code == -1, what
    say("case: ilist past nlocal")
    xa, ta = _f64_c(xall), np.ascontiguousarray(tall)
    il = np.array([0, n], dtype=np.int32)
    nn = np.zeros(n + 1, dtype=np.int32)
    ptrs = (_ip * (n + 1))()
    f = np.zeros((len(xa), 3))
    en = ctypes.c_double()
    rc = lib().tmd_compute_list(m.h, n, len(xa), _p(xa, _dp), _p(ta, _ip), 2, _p(il, _ip),
                                _p(nn, _ip), ptrs, 1.0, _p(f, _dp), None, None,
                                ctypes.byref(en), None)
    assert rc == -1
    say("case: descriptor rows without an output buffer")
    rc = lib().tmd_descriptors_list(m.h, n, len(xa), _p(xa, _dp), _p(ta, _ip), 0, _p(il, _ip),
                                    _p(nn, _ip), None, None)
    assert rc == -1


def test_system_c2_shape(synth_models):
    say("case: periodic system, C2 model, bcc 8x8x8 4 elements")
    om = oracle.OracleModel(synth_models["c2"], 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (8, 8, 8), nelt=4, sigma=0.08, seed=22)
    ref = om.compute_periodic(x, t - 1, cell)
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "fp32", 0.001)
    s = m.system(x, t, cell, skin=1.0)
    got = s.compute()
    e = _errs(got, ref, len(x))
    say("errors", e)
    assert e["dF"] <= F_TOL
    assert e["dE"] <= E_TOL
    assert abs(got["eatom"] - ref["eatom"]).max() <= 1e-6
    assert e["dV"] <= 1e-6


@pytest.mark.parametrize("skin,lo,hi", [(2.0, 128, 192), (3.0, 192, 256), (4.5, 256, 400)])
def test_compaction_paths(synth_models, skin, lo, hi):
    """Raw lists of ~160, ~215 and ~333 entries per atom (rc + skin = 8.5, 9.5,
    11 A): the descriptor's 3- and 4-batch compaction and its two-sweep path
    for lists over 256 entries.  The kept set, and so the result, does not
    depend on the skin."""
    say(f"case: compaction path, C2 model, bcc 8x8x8, skin {skin}")
    om = oracle.OracleModel(synth_models["c2"], 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (8, 8, 8), nelt=4, sigma=0.08, seed=28)
    ref = om.compute_periodic(x, t - 1, cell)
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "fp32", 0.001)
    s = m.system(x, t, cell, skin=skin)
    got = s.compute()
    n = np.diff(s.neighbors()[0])
    say("raw list entries per atom", int(n.min()), int(n.max()))
    assert lo < n.min() and n.max() <= hi
    e = _errs(got, ref, len(x))
    say("errors", e)
    assert e["dF"] <= F_TOL
    assert e["dE"] <= E_TOL
    assert e["dV"] <= 1e-6


def test_kernel_timing_samples(synth_models):
    """Sampled kernel timing inside md_run: every 5th of 20 steps launched
    kernel by kernel with events; the trajectory is the one of the graph
    replay, and the sampled kept-neighbor mean is the count within rc."""
    say("case: sampled kernel timing in md_run, C2 model, bcc 6x6x6")
    x, t, cell = W.crystal("bcc", 3.2, (6, 6, 6), nelt=4, sigma=0.08, seed=33)
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "fp32", 0.001)
    v = W.maxwell_boltzmann(np.array([95.95, 92.906, 180.95, 183.84])[t - 1], 300.0, seed=34)
    outs, kt = [], None
    for every in (0, 5):
        s = m.system(x, t, cell, skin=1.0)
        s.set_velocities(v)
        s.set_kernel_timing(every)
        s.md_run(20, 0.001, history=False)
        outs.append(s.get())
        kt = s.kernel_timing()
        s.close()
    say("timing", kt)
    assert kt["samples"] == 4
    assert all(kt[k] > 0 for k in ("descriptor", "mlp", "force"))
    ro, _, _ = oracle.neighbors(outs[1]["x"], cell, m.cutoff)
    assert abs(kt["kept_per_atom"] - np.diff(ro).mean()) < 1.0
    assert abs(outs[0]["x"] - outs[1]["x"]).max() < 1e-9
    assert abs(outs[0]["f"] - outs[1]["f"]).max() < 1e-7


@pytest.mark.parametrize("mm,tsym", [(0, False), (1, False), (2, False), (2, True)])
def test_lower_max_moment(tmp_path, mm, tsym):
    """Models with max_moment 0-2 (the reference takes 0-5; the shipped models
    use 3) run on the max_moment-3 kernels with the missing moment channels
    given zero first-layer weights.  Forces, energies, virial and the
    exported descriptors (in the file's [b][k][max_moment + 1] layout) match
    the oracle, which evaluates the model's own max_moment."""
    say(f"case: max_moment {mm}, T-symmetric {tsym}, 2 elements, bcc 6x6x6")
    path = W.write_model(str(tmp_path / f"mm{mm}{int(tsym)}.npz"), symbols=["Mo", "W"],
                         masses=[95.95, 183.84], rmax=6.0, num_filters=16, layer_sizes=[64, 32, 1],
                         actfn="tanh", max_moment=mm, t_symmetric=tsym, seed=40 + mm,
                         radial_scale=0.2)
    om = oracle.OracleModel(path, 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (6, 6, 6), nelt=2, sigma=0.08, seed=41)
    ref = om.compute_periodic(x, t - 1, cell, want_desc=True)
    m = TensorMD(path, ["Mo", "W"], "fp32", 0.001)
    info = m.info()
    assert info["max_moment"] == mm and info["width"] == 2 * 16 * (mm + 1)
    s = m.system(x, t, cell, skin=1.0)
    got = s.compute()
    e = _errs(got, ref, len(x))
    say("errors", e)
    assert e["dF"] <= F_TOL
    assert e["dE"] <= E_TOL
    assert e["dV"] <= 1e-6
    G, dG = s.descriptors(info["width"])
    scale = max(1.0, float(abs(ref["G"]).max()))
    assert abs(G - ref["G"]).max() <= 1e-5 * scale
    assert abs(dG - ref["dEdG"]).max() <= 1e-5 * max(1.0, float(abs(ref["dEdG"]).max()))


def test_polynomial_cutoff(tmp_path):
    """fctype 1: the polynomial cutoff 1 + g Z^6 - (g + 1) Z^5 with g = 5
    (tensormd.cpp:158-162, cutoff.cpp:47-51), folded into the radial table."""
    say("case: polynomial cutoff (fctype 1), 2 elements, bcc 6x6x6")
    path = W.write_model(str(tmp_path / "poly.npz"), symbols=["Mo", "W"], masses=[95.95, 183.84],
                         rmax=6.0, num_filters=16, layer_sizes=[64, 32, 1], actfn="tanh", fctype=1,
                         seed=47, radial_scale=0.2)
    om = oracle.OracleModel(path, 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (6, 6, 6), nelt=2, sigma=0.08, seed=48)
    ref = om.compute_periodic(x, t - 1, cell)
    m = TensorMD(path, ["Mo", "W"], "fp32", 0.001)
    got = m.system(x, t, cell, skin=1.0).compute()
    e = _errs(got, ref, len(x))
    say("errors", e)
    assert e["dF"] <= F_TOL
    assert e["dE"] <= E_TOL
    assert e["dV"] <= 1e-6


def test_neighbor_set_bit_exact(synth_models):
    say("case: neighbor set, bcc 10x9x8 tilted cell")
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "fp32", 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (10, 9, 8), nelt=4, sigma=0.1, seed=23)
    cell = cell.copy()
    cell[1, 0] += 1.3
    cell[2, 1] -= 0.7
    x = W.wrap(x, cell)
    s = m.system(x, t, cell, skin=1.0)
    s.compute()
    off, j, sh = s.neighbors()
    xd = s.get()["x"]  # positions as wrapped on the device
    ro, rj, rs = oracle.neighbors(xd, cell, m.cutoff + 1.0, method="brute")
    assert np.array_equal(np.diff(off), np.diff(ro))
    for i in range(len(x)):
        a = sorted(zip(j[off[i]:off[i + 1]], map(tuple, sh[off[i]:off[i + 1]])))
        b = sorted(zip(rj[ro[i]:ro[i + 1]], map(tuple, rs[ro[i]:ro[i + 1]])))
        assert a == b, i


@pytest.mark.parametrize("case", ["W_k32", "W_m3_k64", "MoNi_m3", "MoNi_squareplus_500k"])
def test_golden_models(goldens, case):
    say("case: golden model", case)
    c = goldens["cases"][case]
    if case.startswith("W_"):
        x, cell = W.lattice("bcc", c["a"], (3, 3, 3))
        types = np.ones(len(x), dtype=np.int32)
        symbols = ["W"]
    else:
        s = goldens["structures"]["Mo8_Ni8"]
        x1 = np.array(s["x"])
        c1 = np.array(s["cell"])
        # replicate 2x2x2 so that the cell is wider than the list cutoff
        x = np.concatenate([W.wrap(x1, c1) + np.array([a, b, cc]) @ c1
                            for a in range(2) for b in range(2) for cc in range(2)])
        cell = 2.0 * c1
        x = W.wrap(x, cell)
        types = np.tile(np.array(s["types"]), 8)
        symbols = c["symbols"]
    m = TensorMD(model_path(c["model"]), symbols, "fp32", c["interp"])
    got = m.system(x, types, cell, skin=1.0).compute()
    om = oracle.OracleModel(model_path(c["model"]), c["interp"])
    ref = om.compute_periodic(x, types - 1, cell)
    e = _errs(got, ref, len(x))
    fl = _fp32_floor(om, x, types, cell, ref)
    say("errors", e, "FP32 floor", fl, "E/atom", got["energy"] / len(x),
        "golden", c["pe"] / c["natoms"])
    e_tol, f_tol = _floor_bars(fl)
    assert e["dF"] <= f_tol
    assert e["dE"] <= e_tol
    assert abs(got["energy"] / len(x) - c["pe"] / c["natoms"]) <= e_tol + 5e-7
    if c["model"] == goldens["forces_MoNi"]["model"]:  # the force dump's model
        fy = np.array([r[2:] for r in goldens["forces_MoNi"]["rows"]])
        assert abs(got["f"][:16] - fy).max() <= F_TOL


def test_k128_model_vs_oracle():
    """The widest radial basis shipped with the reference (k = 128, four
    128-wide squareplus resnet layers), jittered so that forces are not zero
    by symmetry: FP32 path vs the FP64 oracle."""
    say("case: W_k128_3H_ver_2c, bcc 4x4x4 jittered")
    path = model_path("W_k128_3H_ver_2c.npz")
    x, t, cell = W.crystal("bcc", 3.189, (4, 4, 4), sigma=0.06, seed=46)
    om = oracle.OracleModel(path, 0.001)
    ref = om.compute_periodic(x, t - 1, cell)
    m = TensorMD(path, ["W"], "fp32", 0.001)
    got = m.system(x, t, cell, skin=1.0).compute()
    e = _errs(got, ref, len(x))
    fl = _fp32_floor(om, x, t, cell, ref)
    say("errors", e, "FP32 floor", fl, "E/atom", ref["energy"] / len(x),
        "max |F|", float(abs(ref["f"]).max()))
    e_tol, f_tol = _floor_bars(fl)
    assert e["dF"] <= f_tol
    assert e["dE"] <= e_tol


@pytest.mark.parametrize("method", ["pexp", "morse", "density", "sf"])
def test_analytic_radial_models(tmp_path, method):
    """use_fnn 0 models (analytic radial functions of
    src/ML-TENSORMD/descriptor.cpp:51-132, tabulated like the radial
    network, descriptor.cpp:146-172): FP32 path vs the FP64 oracle."""
    say(f"case: analytic radial {method}, 3 elements, bcc 4x4x4 jittered")
    p = W.write_model(str(tmp_path / f"{method}.npz"), symbols=["Mo", "Nb", "Ta"],
                      masses=[95.95, 92.906, 180.95], rmax=5.5, num_filters=24,
                      layer_sizes=[64, 32, 1], actfn="tanh", seed=70, radial=method)
    om = oracle.OracleModel(p, 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (4, 4, 4), nelt=3, sigma=0.08, seed=71)
    ref = om.compute_periodic(x, t - 1, cell)
    m = TensorMD(p, ["Mo", "Nb", "Ta"], "fp32", 0.001)
    got = m.system(x, t, cell, skin=1.0).compute()
    e = _errs(got, ref, len(x))
    say("errors", e, "max |F|", float(abs(ref["f"]).max()))
    assert float(abs(ref["f"]).max()) > 1e-3  # the model exerts forces
    assert e["dF"] <= F_TOL and e["dE"] <= E_TOL and e["dV"] <= 1e-6


@pytest.mark.parametrize("case", ["W_k32", "MoNi_squareplus_500k", "W_k128", "tanh_poly", "pexp",
                                  "morse", "density", "sf"])
def test_device_table_build(goldens, tmp_path, case):
    """The radial table built on the device at load (table_device = 1, the
    work of forward_fnn_GPU, gpu/cuda/nnp.cu:59-116, and the Kerley set-up,
    ration.hpp:166-225) against the host FP64 build: values, slopes and
    Kerley coefficients, then forces and energy of the same structure."""
    say("case: device table build,", case)
    if case in goldens["cases"]:
        c = goldens["cases"][case]
        path, symbols, dx = model_path(c["model"]), c.get("symbols") or ["W"], c["interp"]
    elif case == "W_k128":
        path, symbols, dx = model_path("W_k128_3H_ver_2c.npz"), ["W"], 0.001
    else:
        kw = dict(fctype=1, seed=47) if case == "tanh_poly" else dict(radial=case, seed=70)
        path = W.write_model(str(tmp_path / f"{case}.npz"), symbols=["Mo", "W"], masses=[95.95, 183.84],
                             rmax=5.5, num_filters=16, layer_sizes=[64, 32, 1], actfn="tanh",
                             radial_scale=0.2, **kw)
        symbols, dx = ["Mo", "W"], 0.001
    host = TensorMD(path, symbols, "fp32", dx)
    dev = TensorMD(path, symbols, "fp32", dx, table_device=True)
    th, dxh = host.radial_table()
    td, dxd = dev.radial_table()
    assert th.shape == td.shape and dxh == dxd
    d = [float(abs(td[..., q] - th[..., q]).max()) for q in range(4)]
    say("max |difference| of y, slope, c1, c2:", d, "max |y|", float(abs(th[..., 0]).max()))
    # y differs by the rounding of the elementary functions (allowed: 16 ulps
    # of max |y|); the slope is a difference of two values over dx, and c1,
    # c2 are differences of at most three slopes over dx, so their bounds
    # grow by 2/dx and a further 4/dx
    u = 16 * np.finfo(np.float64).eps * float(abs(th[..., 0]).max())
    assert d[0] <= u and d[1] <= 2 * u / dxh and max(d[2], d[3]) <= 8 * u / dxh ** 2
    nelt = len(symbols)
    x, t, cell = W.crystal("bcc", 3.2, (4, 4, 4), nelt=nelt, sigma=0.08, seed=91)
    a = host.system(x, t, cell, skin=1.0).compute()
    b = dev.system(x, t, cell, skin=1.0).compute()
    fmax = max(1.0, float(abs(a["f"]).max()))
    say("max |dF|", float(abs(a["f"] - b["f"]).max()), "|dE|", abs(a["energy"] - b["energy"]))
    assert float(abs(a["f"] - b["f"]).max()) <= 1e-6 * fmax
    assert abs(a["energy"] - b["energy"]) <= 1e-9 * max(1.0, abs(a["energy"]))


@pytest.mark.parametrize("mm,k,nelt", [(4, 24, 4), (5, 16, 2), (5, 32, 3), (4, 128, 1)])
def test_high_moment_models(tmp_path, mm, k, nelt):
    """max_moment 4 and 5 models (35 and 56 monomials, tensormd.cpp:234-273;
    the reference GPU path's calculate_dM_g/_h, kernel.cu:553-556): FP32 path
    vs the FP64 oracle, energies, forces, virial and the descriptor rows."""
    sym = ["Mo", "Nb", "Ta", "W"][:nelt]
    say(f"case: max_moment {mm}, k={k}, {nelt} elements, bcc 4x4x4 jittered")
    p = W.write_model(str(tmp_path / "hm.npz"), symbols=sym, masses=[50.0 + 10 * i for i in range(nelt)],
                      rmax=5.5, num_filters=k, layer_sizes=[64, 32, 1], actfn="tanh", max_moment=mm,
                      seed=80 + mm + k, radial_scale=0.2)
    om = oracle.OracleModel(p, 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (4, 4, 4), nelt=nelt, sigma=0.08, seed=81)
    ref = om.compute_periodic(x, t - 1, cell, want_desc=True)
    m = TensorMD(p, sym, "fp32", 0.001)
    assert m.info()["max_moment"] == mm and m.info()["width"] == nelt * k * (mm + 1)
    s = m.system(x, t, cell, skin=1.0)
    got = s.compute()
    e = _errs(got, ref, len(x))
    G, _ = s.descriptors(m.info()["width"])
    dG = float(abs(G - ref["G"]).max())
    say("errors", e, "max |dG|", dG, "max |G|", float(abs(ref["G"]).max()))
    assert dG <= 1e-5 * max(1.0, float(abs(ref["G"]).max()))
    assert e["dF"] <= F_TOL and e["dE"] <= E_TOL and e["dV"] <= 1e-6


@pytest.mark.parametrize("sommerfeld,resnet", [(False, False), (True, True)])
def test_tdnp_models(tmp_path, sommerfeld, resnet):
    """TDNP models (tdnp.cpp:150-218): the atomic energies are the electron
    internal energy u, the forces -dA/dx with A = U - T S, and the entropy
    and free energy (sums and per atom) are returned beside them.  List mode
    and the device-resident system against the FP64 oracle, at two electron
    temperatures set on the same context."""
    say(f"case: TDNP sommerfeld={sommerfeld} resnet={resnet}, 2 elements, bcc 4x4x4")
    p = W.write_model(str(tmp_path / "tdnp.npz"), symbols=["Mo", "Ni"], masses=[95.95, 58.69], rmax=5.5,
                      num_filters=16, layer_sizes=[1], seed=90 + int(sommerfeld),
                      tdnp=dict(H=[64, 32, 16], U=[32, 1], S=[32, 1], sommerfeld=sommerfeld,
                                resnet=resnet, actfn="softplus"))
    om = oracle.OracleModel(p, 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (4, 4, 4), nelt=2, sigma=0.08, seed=91)
    xall, owner, ilist, numneigh, flat = W.lammps_ghost_list(x, cell, om.rc + 1.0)
    tall = t[owner]
    n = len(x)
    m = TensorMD(p, ["Mo", "Ni"], "fp32", 0.001)
    sysm = m.system(x, t, cell, skin=1.0)
    for T in (0.25, 0.6):
        om.set_etemp(T)
        m.set_electron_temperature(T)
        ref = om.compute_list(n, xall, tall - 1, ilist, numneigh, flat, scale=0.9)
        rtd = om.tdnp_last(n)
        got = m.compute_list(n, xall, tall, ilist, numneigh, flat, scale=0.9)
        gtd = m.tdnp_results(n)
        e = _errs(got, ref, n)
        dS = abs(gtd["entropy"] - rtd["entropy"]) / n
        dA = abs(gtd["free_energy"] - rtd["free_energy"]) / n
        say("T", T, "list errors", e, "dS/atom", dS, "dA/atom", dA,
            "A-U per atom", (rtd["free_energy"] - ref["energy"]) / n)
        assert abs(rtd["free_energy"] - ref["energy"]) / n > 1e-4  # the entropy term matters
        assert e["dF"] <= F_TOL and e["dE"] <= E_TOL and e["dV"] <= 1e-6
        assert dS <= E_TOL and dA <= E_TOL
        assert abs(gtd["satom"] - rtd["satom"]).max() <= 1e-6
        assert abs(gtd["aatom"] - rtd["aatom"]).max() <= 1e-6
        assert abs(got["eatom"] - ref["eatom"]).max() <= 1e-6
        # the device-resident system reads the same temperature
        refp = om.compute_periodic(x, t - 1, cell)
        sysm.set_positions(x)
        gp = sysm.compute()
        ep = _errs(gp, refp, n)
        say("T", T, "system errors", ep)
        assert ep["dF"] <= F_TOL and ep["dE"] <= E_TOL


@pytest.mark.parametrize("k,nelt", [(64, 5), (128, 3), (128, 5), (32, 8)])
def test_wide_slot_models(tmp_path, k, nelt):
    """Models whose slots take several register passes in the descriptor and
    force kernels (k = 64 with 5 elements, the reference's element limit,
    tensormd.h:27; k = 128 with 3 and 5 elements), and an 8-element model
    in one pass: FP32 path vs the FP64 oracle, energies, forces, virial and
    the descriptor rows themselves."""
    sym = ["Mo", "Nb", "Ta", "W", "V", "Cr", "Ti", "Zr"][:nelt]
    say(f"case: k={k}, {nelt} elements, bcc 4x4x4 jittered")
    p = W.write_model(str(tmp_path / "wide.npz"), symbols=sym, masses=[50.0 + 10 * i for i in range(nelt)],
                      rmax=5.5, num_filters=k, layer_sizes=[64, 32, 1], actfn="tanh", seed=50 + k + nelt,
                      radial_scale=0.2)
    x, t, cell = W.crystal("bcc", 3.2, (4, 4, 4), nelt=nelt, sigma=0.08, seed=47)
    om = oracle.OracleModel(p, 0.001)
    ref = om.compute_periodic(x, t - 1, cell, want_desc=True)
    m = TensorMD(p, sym, "fp32", 0.001)
    s = m.system(x, t, cell, skin=1.0)
    got = s.compute()
    e = _errs(got, ref, len(x))
    G, dG = s.descriptors(m.info()["width"])
    rG = float(abs(G - ref["G"]).max() / max(1.0, abs(ref["G"]).max()))
    rdG = float(abs(dG - ref["dEdG"]).max() / max(1.0, abs(ref["dEdG"]).max()))
    say("errors", e, "rel dG", rG, "rel d(dE/dG)", rdG)
    assert rG <= 1e-5 and rdG <= 1e-5
    assert e["dF"] <= F_TOL and e["dE"] <= E_TOL and e["dV"] <= 1e-6


@pytest.mark.parametrize("resnet", [False, True])
def test_relu_model(tmp_path, resnet):
    """ReLU main network (actfn 0), with and without resnet skips: FP32 path
    vs the FP64 oracle (whose ReLU network is pinned by torch autograd)."""
    say(f"case: relu model resnet={resnet}, 2 elements, fcc 4x4x4 jittered")
    p = W.write_model(str(tmp_path / "relu.npz"), symbols=["Al", "Ni"], masses=[26.98, 58.69], rmax=5.0,
                      num_filters=16, layer_sizes=[64, 64, 32, 1], actfn="relu", use_resnet_dt=resnet,
                      seed=22, radial_scale=0.2, bias_sigma=0.3)
    x, t, cell = W.crystal("fcc", 3.8, (4, 4, 4), nelt=2, sigma=0.06, seed=48)
    ref = oracle.OracleModel(p, 0.001).compute_periodic(x, t - 1, cell)
    got = TensorMD(p, ["Al", "Ni"], "fp32", 0.001).system(x, t, cell, skin=1.0).compute()
    e = _errs(got, ref, len(x))
    say("errors", e)
    assert e["dF"] <= F_TOL and e["dE"] <= E_TOL and e["dV"] <= 1e-6


def test_bf16_mlp_error(synth_models):
    say("case: bf16 MLP, C2 model, bcc 8x8x8")
    om = oracle.OracleModel(synth_models["c2"], 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (8, 8, 8), nelt=4, sigma=0.08, seed=24)
    ref = om.compute_periodic(x, t - 1, cell)
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "bf16", 0.001)
    got = m.system(x, t, cell, skin=1.0).compute()
    e = _errs(got, ref, len(x))
    rms = float(np.sqrt(((got["f"] - ref["f"]) ** 2).mean()))
    say("bf16 errors", e, "rmsF", rms)
    assert e["dF"] <= 5e-3


def _bf16_rne(a):
    """FP32 -> BF16 (round to nearest even) -> FP32, for finite values."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def test_bf16_descriptor_rows(synth_models):
    """The BF16 path of a benchmark-shaped network (384-256-128-64-1, tanh)
    stores its descriptor rows as BF16: the exported G equals the FP32
    path's G rounded to nearest even, element for element (and differs from
    the unrounded rows, so the BF16 rows were taken); the forces stay
    within the BF16 bar of the oracle."""
    say("case: BF16 descriptor rows, C2 model, bcc 6x6x6")
    x, t, cell = W.crystal("bcc", 3.2, (6, 6, 6), nelt=4, sigma=0.08, seed=31)
    syms = ["Mo", "Nb", "Ta", "W"]
    m32 = TensorMD(synth_models["c2"], syms, "fp32", 0.001)
    m16 = TensorMD(synth_models["c2"], syms, "bf16", 0.001)
    width = m32.info()["width"]
    s32 = m32.system(x, t, cell, skin=1.0)
    s32.compute()
    G32, _ = s32.descriptors(width)
    s16 = m16.system(x, t, cell, skin=1.0)
    got = s16.compute()
    G16, dG16 = s16.descriptors(width)
    rounded = _bf16_rne(G32)
    say("G rows differing from their BF16 rounding:", int((rounded != G32).sum()), "of", G32.size)
    assert not np.array_equal(rounded, G32)
    assert np.array_equal(G16, rounded)
    assert np.isfinite(dG16).all()
    ref = oracle.OracleModel(synth_models["c2"], 0.001).compute_periodic(x, t - 1, cell)
    assert _errs(got, ref, len(x))["dF"] <= 5e-3


def test_bf16x3_mlp_error(synth_models):
    """Split-BF16 MLP (operands as hi + lo BF16 pairs, three BF16 MFMA
    products per product): the C2 model on a jittered bcc 8x8x8 lattice
    against the FP64 oracle, beside plain BF16 on the same input."""
    say("case: split-bf16 MLP, C2 model, bcc 8x8x8")
    om = oracle.OracleModel(synth_models["c2"], 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (8, 8, 8), nelt=4, sigma=0.08, seed=24)
    ref = om.compute_periodic(x, t - 1, cell)
    errs = {}
    for prec in ("bf16", "bf16x3"):
        m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], prec, 0.001)
        errs[prec] = _errs(m.system(x, t, cell, skin=1.0).compute(), ref, len(x))
        m.close()
    # the same structure through list mode (the LAMMPS path)
    xall, owner, ilist, numneigh, flat = W.lammps_ghost_list(x, cell, om.rc + 1.0)
    tall = t[owner]
    ref_l = om.compute_list(len(x), xall, tall - 1, ilist, numneigh, flat)
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "bf16x3", 0.001)
    errs["bf16x3_list"] = _errs(m.compute_list(len(x), xall, tall, ilist, numneigh, flat), ref_l, len(x))
    m.close()
    say("errors", errs)
    for key in ("bf16x3", "bf16x3_list"):
        assert errs[key]["dF"] <= 5e-4 and errs[key]["dE"] <= 1e-5
    assert errs["bf16x3"]["dF"] <= 0.1 * errs["bf16"]["dF"]


def _guard_atoms(ref, G_gpu, x, cell, k, nelt, rc):
    """Atoms whose forces hang on the reference's |G_0| < 1e-8 guard
    (dE/dG_0 is dropped below it: gpu/cuda/kernel.cu:754-759): every atom
    with a G_0 that the FP32 device path and the FP64 oracle put on
    different sides of the threshold (G_0 = P_0 sums radial values that can
    cancel to within FP32 rounding of 1e-8), and its neighbors within rc,
    which receive that atom's pair forces.  There the two force fields
    differ by the whole dropped term; elsewhere the guard acts alike."""
    g64 = np.abs(ref["G"].reshape(len(x), nelt, k, 4)[..., 0]) < 1e-8
    g32 = np.abs(G_gpu.reshape(len(x), nelt, k, 4)[..., 0]) < 1e-8
    near = np.where((g64 != g32).any(axis=(1, 2)))[0]
    mask = np.zeros(len(x), bool)
    inv = np.linalg.inv(cell)
    for a in near:
        d = x - x[a]
        fr = d @ inv
        fr -= np.round(fr)
        d = fr @ cell
        mask |= (d * d).sum(1) < (rc + 1e-6) ** 2
    return near, mask


def test_c2_trajectory_errors_and_drift(tmp_path):
    """The 128,000-atom BCC benchmark system (same model, lattice and velocity
    seeds as bench.py): forces, energies and virials of the BF16 path along
    its 1000-step NVE run, and of the split-BF16 and FP32 paths at the same
    positions, against the FP64 oracle every 250 steps; total-energy drift of
    the BF16 and FP32 paths.

    FP32 meets its bars (1e-5 eV/A, 1e-7 eV/atom) all along, and split BF16
    holds 5e-3 eV/A at every checkpoint.  Plain BF16 meets 5e-3 eV/A on the
    starting lattice (max |F| ~0.5 eV/A); the random-init model then heats
    the system to ~2,200 K (max |F| up to ~2.4 eV/A) and the plain BF16
    max-abs error grows with the forces to 0.5-2 % of max |F| (RMS ~1e-3
    eV/A): the bounds below pin that measured behaviour."""
    say("case: C2 128k-atom trajectory, BF16 and FP32 vs the FP64 oracle, NVE drift")
    syms = ["Mo", "Nb", "Ta", "W"]
    masses = np.array([95.95, 92.906, 180.95, 183.84])
    path = W.write_model(str(tmp_path / "c2.npz"), symbols=syms, masses=list(masses), rmax=6.5,
                         num_filters=24, layer_sizes=[256, 128, 64, 1], actfn="tanh", seed=2024,
                         radial_scale=0.2)
    x, t, cell = W.crystal("bcc", 3.2, (40, 40, 40), nelt=4, sigma=0.08, seed=7)
    n = len(x)
    v = W.maxwell_boltzmann(masses[t - 1], 600.0, seed=8)
    om = oracle.OracleModel(path, 0.001)
    mb = TensorMD(path, syms, "bf16", 0.001)
    mf = TensorMD(path, syms, "fp32", 0.001)
    m3 = TensorMD(path, syms, "bf16x3", 0.001)
    sb = mb.system(x, t, cell, skin=1.0)
    sb.set_velocities(v)
    sb.compute()
    probe = mf.system(x, t, cell, skin=1.0)
    probe3 = m3.system(x, t, cell, skin=1.0)
    ep_b, ek_b = [], []
    for seg in range(5):
        if seg:
            ep, ek = sb.md_run(250, 0.001, history=True)
            ep_b.append(ep)
            ek_b.append(ek)
        got = sb.get()
        ref = om.compute_periodic(got["x"], t - 1, cell, want_desc=True)
        probe.set_positions(got["x"])
        gf = probe.compute()
        probe3.set_positions(got["x"])
        g3 = probe3.compute()
        # forces compared away from the reference's G_0 guard (its atoms are
        # reported; they are a handful per 10^5)
        Gp, _ = probe.descriptors(mf.info()["width"])
        near, gmask = _guard_atoms(ref, Gp, got["x"], cell, 24, 4, om.rc)
        keep = ~gmask
        eb, ef, e3 = _errs(got, ref, n), _errs(gf, ref, n), _errs(g3, ref, n)
        for e, g in ((eb, got), (ef, gf), (e3, g3)):
            e["dF_all"] = e["dF"]
            e["dF"] = float(abs(g["f"][keep] - ref["f"][keep]).max())
        maxf = float(abs(ref["f"]).max())
        rms = float(np.sqrt(((got["f"] - ref["f"]) ** 2).mean()))
        say(f"step {250 * seg}: max|F| {maxf:.3f} guard atoms {list(near)} ({int(gmask.sum())} excluded) "
            f"bf16 {eb} rms {rms:.2e} bf16x3 {e3} fp32 {ef}")
        assert gmask.sum() <= 0.005 * n
        assert ef["dF"] <= F_TOL and ef["dE"] <= E_TOL and ef["dV"] <= 1e-6
        # the 5e-3 eV/A target, and a regression bar 10x above the measured
        # split-BF16 error (max 4.7e-5 eV/A at the hottest checkpoint)
        assert e3["dF"] <= 5e-3 and e3["dF"] <= 5e-4 and e3["dE"] <= 1e-5
        if seg == 0:
            assert eb["dF"] <= 5e-3
        assert eb["dF"] <= 0.03 * maxf and rms <= 2e-3
    sf = mf.system(x, t, cell, skin=1.0)
    sf.set_velocities(v)
    sf.compute()
    ep_f, ek_f = sf.md_run(1000, 0.001, history=True)
    drift = {}
    for tag, ep, ek in (("bf16", np.concatenate(ep_b), np.concatenate(ek_b)), ("fp32", ep_f, ek_f)):
        et = (ep + ek) / n
        drift[tag] = (float(et[-1] - et[0]), float(abs(et - et[0]).max()))
    say("NVE total energy per atom over 1000 steps (end - start, max |deviation|):", drift)
    assert abs(drift["fp32"][0]) <= 5e-6 and drift["fp32"][1] <= 5e-6
    assert abs(drift["bf16"][0]) <= 5e-5 and drift["bf16"][1] <= 1e-4


def test_nve_c1_trajectory(synth_models):
    say("case: NVE 100 steps, C1 2048 atoms")
    om = oracle.OracleModel(synth_models["c1"], 0.001)
    x, t, cell = W.crystal("fcc", 3.615, (8, 8, 8), sigma=0.05, seed=25)
    v = W.maxwell_boltzmann(np.full(len(x), om.masses[0]), 300.0, seed=26)
    nsteps = 100
    ref = om.nve(x, v, t - 1, cell, 0.001, nsteps)
    m = TensorMD(synth_models["c1"], ["Cu"], "fp32", 0.001)
    s = m.system(x, t, cell, skin=1.0)
    s.set_velocities(v)
    s.compute()
    ep, ek = s.md_run(nsteps, 0.001)
    n = len(x)
    de = np.abs(ep - ref["epot"][1:]) / n
    dk = np.abs(ek - ref["ekin"][1:]) / n
    say("max |dEpot|/atom", de.max(), "max |dEkin|/atom", dk.max(), "builds", s.nbuilds)
    assert de.max() <= 1e-6
    assert dk.max() <= 1e-6
    # energy conservation of the GPU trajectory
    etot = ep + ek
    assert abs(etot - etot[0]).max() / n < 1e-5


def test_graph_replay_matches_eager(synth_models):
    """The captured-step graph gives the same trajectory as eager launches
    (FP64 force atomics may reorder: tolerance, not bit equality)."""
    say("case: graph vs eager, C2 model, bcc 6x6x6, 30 steps")
    x, t, cell = W.crystal("bcc", 3.2, (6, 6, 6), nelt=4, sigma=0.08, seed=31)
    v = W.maxwell_boltzmann(np.array([95.95, 92.906, 180.95, 183.84])[t - 1], 600.0, seed=32)
    m = TensorMD(synth_models["c2"], ["Mo", "Nb", "Ta", "W"], "fp32", 0.001)
    out = []
    for graph in (True, False):
        s = m.system(x, t, cell, skin=1.0)
        s.set_graph(graph)
        s.set_velocities(v)
        s.compute()
        ep, ek = [], []
        for rep in range(3):  # several replays with host syncs in between
            a, b = s.md_run(10, 0.001)
            ep.append(a)
            ek.append(b)
        out.append((np.concatenate(ep), np.concatenate(ek), s.get()["x"]))
        s.close()
    n = len(x)
    assert np.abs(out[0][0] - out[1][0]).max() / n < 1e-9
    assert np.abs(out[0][1] - out[1][1]).max() / n < 1e-9
    assert np.abs(out[0][2] - out[1][2]).max() < 1e-9


@pytest.mark.parametrize("stepping", ["md_run", "async"])
def test_md_run_grows_list_and_rolls_back(synth_models, stepping):
    """A converging flow raises the local density past the neighbor-list
    capacity during the run: the run must grow the list, rerun the segment
    from its snapshot and end with the forces of a fresh evaluation.  With
    one tmd_md_step_async call per step the overflowed step is run again
    when the next call takes up its check."""
    say(f"case: list overflow during {stepping}, C1 2048 atoms, converging flow")
    x, t, cell = W.crystal("fcc", 3.615, (8, 8, 8), sigma=0.05, seed=27)
    L = np.diag(cell)
    v = -1.5 * (x - 0.5 * L)  # A/ps toward the box center
    m = TensorMD(synth_models["c1"], ["Cu"], "fp32", 0.001)
    s = m.system(x, t, cell, skin=1.0)
    s.set_velocities(v)
    s.compute()
    off0, _, _ = s.neighbors()
    cap0 = int(np.diff(off0).max())
    if stepping == "md_run":
        s.md_run(150, 0.001, history=False)
    else:
        for q in range(150):
            s.step_async(0.001)
    off, _, _ = s.neighbors()
    out = s.get()
    say("max list length before", cap0, "after", int(np.diff(off).max()), "builds", s.nbuilds)
    # rc + skin = 7 A, 0.0847 atoms/A^3: the initial capacity is 1.5 x 121.7 + 16 -> 200
    assert int(np.diff(off).max()) > 200
    # a fresh system lists the neighbors in another order (FP32 sums differ
    # in rounding only); a truncated list would be off by whole pair terms
    fresh = m.system(out["x"], t, cell, skin=1.0)
    ref = fresh.compute()
    fmax = max(1.0, float(abs(ref["f"]).max()))
    say("max |dF|", float(abs(out["f"] - ref["f"]).max()), "max |F|", fmax)
    assert abs(out["f"] - ref["f"]).max() <= 1e-5 * fmax
    assert abs(out["energy"] - ref["energy"]) / len(x) <= 1e-6 * max(1.0, abs(ref["energy"]) / len(x))


def test_lammps_fix_call_sequence(synth_models):
    """fix nve/tensormd/amd drives the device system as LAMMPS calls it: at
    each run's setup a system is made from the host state (positions wrapped
    into the box, the wraps counted in image flags, as LAMMPS' setup leaves
    them), every step is one tmd_md_step_async call, thermo steps read the energy,
    and the end of a run copies x and v back.  Two runs of 25 steps agree
    with one continuous 50-step run and with the oracle's NVE."""
    say("case: LAMMPS fix call sequence, C1 2048 atoms, 2 runs x 25 steps")
    om = oracle.OracleModel(synth_models["c1"], 0.001)
    x, t, cell = W.crystal("fcc", 3.615, (8, 8, 8), sigma=0.05, seed=25)
    v = W.maxwell_boltzmann(np.full(len(x), om.masses[0]), 300.0, seed=26)
    nsteps, every = 50, 5
    ref = om.nve(x, v, t - 1, cell, 0.001, nsteps)
    m = TensorMD(synth_models["c1"], ["Cu"], "fp32", 0.001)
    L = np.diag(cell)
    hx, hv, img = x.copy(), v.copy(), np.zeros_like(x)
    epot = []
    for run in range(2):
        wraps = np.floor(hx / L)
        hx, img = hx - wraps * L, img + wraps
        s = m.system(hx, t, cell, skin=1.0)
        s.set_velocities(hv)
        for step in range(nsteps // 2):
            s.step_async(0.001)
            if (step + 1) % every == 0:
                epot.append(s.get()["energy"])
        out = s.get()
        hx, hv = out["x"], out["v"]
        s.close()
    cont = m.system(x, t, cell, skin=1.0)
    cont.set_velocities(v)
    ep_c, _ = cont.md_run(nsteps, 0.001)
    end = cont.get()
    n = len(x)
    epot = np.array(epot)
    dx = float(np.abs(hx + img * L - end["x"]).max())
    de_c = float(np.abs(epot - ep_c[every - 1::every]).max()) / n
    de_o = float(np.abs(epot - ref["epot"][every::every]).max()) / n
    say("max |dx| vs one run", dx, "max |dE|/atom vs one run", de_c, "vs oracle", de_o)
    assert dx <= 1e-6
    assert float(np.abs(hv - end["v"]).max()) <= 1e-5
    assert de_c <= 1e-7
    assert de_o <= 1e-6


def test_fp8_mlp_error(tmp_path):
    """FP8 e4m3 MLP (weights with per-output-channel scales, activations and
    gradients with per-row scales, FP32 accumulation) on the 256-256-128-1
    architecture: errors against the FP64 oracle are reported beside the BF16
    ones; the bar is loose (the experiment's result is the error itself)."""
    say("case: fp8 vs bf16 MLP, 256-256-128-1 model, bcc 8x8x8 4 elements")
    model = W.write_model(str(tmp_path / "c4.npz"), symbols=["Mo", "Nb", "Ta", "W"],
                          masses=[95.95, 92.906, 180.95, 183.84], rmax=6.5, num_filters=24,
                          layer_sizes=[256, 256, 128, 1], actfn="tanh", seed=44,
                          radial_scale=0.2)
    om = oracle.OracleModel(model, 0.001)
    x, t, cell = W.crystal("bcc", 3.2, (8, 8, 8), nelt=4, sigma=0.08, seed=45)
    ref = om.compute_periodic(x, t - 1, cell)
    out = {}
    for prec in ("bf16", "fp8"):
        m = TensorMD(model, ["Mo", "Nb", "Ta", "W"], prec, 0.001)
        got = m.system(x, t, cell, skin=1.0).compute()
        e = _errs(got, ref, len(x))
        e["rmsF"] = float(np.sqrt(((got["f"] - ref["f"]) ** 2).mean()))
        out[prec] = e
        say(prec, "errors", e, "max |F|", float(abs(ref["f"]).max()))
    assert out["bf16"]["dF"] <= 5e-3
    assert out["fp8"]["dF"] <= 0.1
    assert out["fp8"]["dE"] <= 0.01
    # the Pair::compute drop-in (LAMMPS-style ghosts) in FP8, against the
    # oracle's list evaluation, to the same bars
    m = TensorMD(model, ["Mo", "Nb", "Ta", "W"], "fp8", 0.001)
    xall, owner, ilist, numneigh, flat = W.lammps_ghost_list(x, cell, om.rc + 1.0)
    tall = t[owner]
    ref_l = om.compute_list(len(x), xall, tall - 1, ilist, numneigh, flat)
    got_l = m.compute_list(len(x), xall, tall, ilist, numneigh, flat)
    e_l = _errs(got_l, ref_l, len(x))
    say("fp8 list-mode errors", e_l)
    assert e_l["dF"] <= 0.1 and e_l["dE"] <= 0.01
