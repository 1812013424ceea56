"""C ABI: the library loads, exports every declared symbol, and rejects bad
inputs with status codes (no GPU needed: model parsing and argument checks
happen before the device is touched)."""
import ctypes
import os
import re
import zipfile

import numpy as np
import pytest

import tensormd_amd as T
from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "tensormd_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tmd_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(T.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_options_default():
    o = T.Options()
    T.lib().tmd_options_default(ctypes.byref(o))
    assert o.mlp_precision == 0 and o.interp_dx == 0.001 and o.device == -1


def _create_rc(path, symbols, precision="fp32", dx=0.001):
    try:
        T.TensorMD(path, symbols, precision, dx)
        return 0, ""
    except T.TmdError as e:
        return e.code, str(e)


def test_missing_file_is_model_error(tmp_path):
    rc, msg = _create_rc(str(tmp_path / "nope.npz"), ["W"])
    assert rc == -2 and "cannot open" in msg


def test_corrupt_file_is_model_error(tmp_path):
    p = tmp_path / "bad.npz"
    p.write_bytes(b"PK\x03\x04 this is not a zip")
    rc, _ = _create_rc(str(p), ["W"])
    assert rc == -2


def test_missing_key_is_model_error(tmp_path, synth_models):
    src = np.load(synth_models["c1"])
    d = {k: src[k] for k in src.files if k != "fnn::num_filters"}
    p = tmp_path / "nokey.npz"
    np.savez(p, **d)
    rc, msg = _create_rc(str(p), ["Cu"])
    assert rc == -2 and "fnn::num_filters" in msg


def test_unknown_element_is_arg_error(synth_models):
    rc, msg = _create_rc(synth_models["c1"], ["Xx"])
    assert rc == -1 and "pair coefficients" in msg


def test_bad_interp_is_model_error(synth_models):
    rc, _ = _create_rc(synth_models["c1"], ["Cu"], dx=-1.0)
    assert rc == -2


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device path")
def test_fp8_precision_is_accepted(synth_models):
    """The FP8 and split-BF16 MLP modes pass option checks (the model's
    widths are checked against the fast kernel's alignment after device
    initialisation); an unknown mode is an argument error."""
    for prec in ("fp8", "bf16x3"):
        rc, msg = _create_rc(synth_models["c1"], ["Cu"], precision=prec)
        assert rc == -3, msg
    L = T.lib()
    opt = T.Options()
    L.tmd_options_default(ctypes.byref(opt))
    opt.mlp_precision = 4  # no such mode
    arr = (ctypes.c_char_p * 1)(b"Cu")
    h = ctypes.c_void_p()
    rc = L.tmd_create(str(synth_models["c1"]).encode(), arr, 1, ctypes.byref(opt), ctypes.byref(h))
    assert rc == -1


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device path")
def test_valid_model_reaches_device_init(synth_models, tmp_path):
    """Valid models (stored and deflated npz) parse; without a GPU the call
    fails at device initialisation, not in the model reader."""
    rc, msg = _create_rc(synth_models["c1"], ["Cu"])
    assert rc == -3, msg
    src = np.load(synth_models["c1"])
    p = tmp_path / "deflated.npz"
    np.savez_compressed(p, **{k: src[k] for k in src.files})
    with zipfile.ZipFile(p) as z:
        assert any(i.compress_type == zipfile.ZIP_DEFLATED for i in z.infolist())
    rc, msg = _create_rc(str(p), ["Cu"])
    assert rc == -3, msg


@pytest.mark.parametrize("lang,std", [("c", "c99"), ("c", "c11"), ("c++", "c++11"), ("c++", "c++17")])
def test_header_is_standalone(tmp_path, lang, std):
    """include/tensormd_amd.h compiles on its own as C and as C++."""
    import shutil
    import subprocess
    cc = shutil.which("gcc" if lang == "c" else "g++")
    if cc is None:
        pytest.skip("no host compiler")
    src = tmp_path / ("use." + ("c" if lang == "c" else "cpp"))
    src.write_text('#include "tensormd_amd.h"\n'
                   "int probe(size_t n) { tmd_options o; tmd_options_default(&o); return (int) n + o.device; }\n")
    r = subprocess.run([cc, "-x", lang, f"-std={std}", "-Wall", "-Werror", "-pedantic", "-fsyntax-only",
                        "-I" + os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("k", [12, 48, 96])
def test_unsupported_filter_count_is_model_error(tmp_path, k):
    """Radial widths the descriptor kernels are not built for are refused at
    load (TMD_ERR_MODEL), before any device work, not at the first compute."""
    from tensormd_amd import workloads as W
    p = W.write_model(str(tmp_path / f"k{k}.npz"), symbols=["Mo", "W"], masses=[95.95, 183.84], rmax=5.0,
                      num_filters=k, layer_sizes=[32, 1], seed=3)
    rc, msg = _create_rc(p, ["Mo", "W"])
    assert rc == -2 and "num_filters" in msg, msg


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-device path")
@pytest.mark.parametrize("k,nelt", [(64, 5), (128, 3), (128, 5), (24, 8)])
def test_wide_slot_models_load(tmp_path, k, nelt):
    """Models whose slots need several register passes (k = 64 with 5
    elements, k = 128 with 3 or more) and 8-element models pass the model
    checks: without a GPU the call fails at device initialisation."""
    from tensormd_amd import workloads as W
    sym = ["Mo", "Nb", "Ta", "W", "V", "Cr", "Ti", "Zr"][:nelt]
    p = W.write_model(str(tmp_path / "m.npz"), symbols=sym, masses=[50.0 + 10 * i for i in range(nelt)],
                      rmax=5.0, num_filters=k, layer_sizes=[32, 1], seed=4)
    rc, msg = _create_rc(p, sym)
    assert rc == -3, msg


def test_list_entry_points_reject_null_context():
    """Argument checks of the list-mode variants come before any device work."""
    L = T.lib()
    e = ctypes.c_double()
    f = np.zeros(3)
    x = np.zeros(3)
    t = np.ones(1, dtype=np.int32)
    args = (1, 1, x.ctypes.data_as(T._dp), t.ctypes.data_as(T._ip), 0, None, None, None, 1.0,
            f.ctypes.data_as(T._dp), None, None, ctypes.byref(e), None)
    assert L.tmd_compute_list(None, *args) == -1
    assert L.tmd_compute_list_serial(None, 0, *args) == -1
    assert L.tmd_compute_list_noalloc(None, *args) == -1


def test_tdnp_entry_points_reject_null_context():
    L = T.lib()
    d = ctypes.c_double()
    assert L.tmd_set_electron_temperature(None, 0.1) == -1
    assert L.tmd_tdnp_results(None, ctypes.byref(d), ctypes.byref(d), None, None) == -1


def test_system_step_and_table_entry_points_reject_null():
    """The per-step entry point and the table export check their arguments
    before any device work; the device-built table is off by default."""
    L = T.lib()
    assert L.tmd_md_step_async(None, 0.001) == -1
    assert L.tmd_md_run(None, 1, 0.001, None, None) == -1
    n, dx = ctypes.c_int(), ctypes.c_double()
    assert L.tmd_radial_table(None, None, 0, ctypes.byref(n), ctypes.byref(dx)) == -1
    opt = T.Options()
    L.tmd_options_default(ctypes.byref(opt))
    assert opt.table_device == 0


def test_device_table_build_needs_a_device(synth_models):
    """table_device = 1 loads the model on the host, then fails at device
    initialisation here (no GPU) instead of falling back to the host table."""
    with pytest.raises(T.TmdError) as ei:
        T.TensorMD(synth_models["c1"], ["Cu"], "fp32", 0.001, table_device=True)
    assert ei.value.code == -3
