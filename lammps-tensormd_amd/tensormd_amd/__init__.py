"""Python binding of libtensormd_amd (the MI355X pair_style tensormd).

Thin ctypes layer over the C ABI in include/tensormd_amd.h.  There is no
fallback: if the HIP library is missing or cannot be loaded, importing the
binding's entry points raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libtensormd_amd.so")
CSRC = os.path.join(os.path.dirname(_HERE), "csrc")

TMD_OK = 0
ERRORS = {-1: "TMD_ERR_ARG", -2: "TMD_ERR_MODEL", -3: "TMD_ERR_DEVICE", -4: "TMD_ERR_NOMEM",
          -5: "TMD_ERR_CAPACITY", -6: "TMD_ERR_STATE"}
PRECISION = {"fp32": 0, "bf16": 1, "fp8": 2}

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)
_vp = ctypes.c_void_p


class TmdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("mlp_precision", ctypes.c_int), ("interp_dx", ctypes.c_double),
                ("cmax_exact", ctypes.c_int), ("device", ctypes.c_int)]


_lib = None


def build():
    """Compile the HIP library in-tree (hipcc, gfx950)."""
    import subprocess
    subprocess.run(["make", "-C", CSRC, "-j8"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {CSRC}` "
                          "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    L.tmd_options_default.argtypes = [ctypes.POINTER(Options)]
    L.tmd_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.c_int,
                             ctypes.POINTER(Options), ctypes.POINTER(_vp)]
    L.tmd_destroy.argtypes = [_vp]
    L.tmd_last_error.argtypes = [_vp]
    L.tmd_last_error.restype = ctypes.c_char_p
    L.tmd_cutoff.argtypes = [_vp, _dp]
    L.tmd_masses.argtypes = [_vp, _dp]
    L.tmd_type_map.argtypes = [_vp, _ip]
    L.tmd_model_info.argtypes = [_vp, _ip]
    L.tmd_compute_list.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _ip, ctypes.c_int, _ip,
                                   _ip, ctypes.POINTER(_ip), ctypes.c_double, _dp, _dp, _dp, _dp,
                                   _dp]
    L.tmd_system_create.argtypes = [_vp, ctypes.c_int, _dp, _ip, _dp, ctypes.c_double,
                                    ctypes.POINTER(_vp)]
    L.tmd_system_destroy.argtypes = [_vp]
    L.tmd_system_set_positions.argtypes = [_vp, _dp]
    L.tmd_system_set_velocities.argtypes = [_vp, _dp]
    L.tmd_system_compute.argtypes = [_vp]
    L.tmd_system_get.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.tmd_md_run.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp, _dp]
    L.tmd_system_nbuilds.argtypes = [_vp]
    L.tmd_md_profile.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp]
    L.tmd_system_set_graph.argtypes = [_vp, ctypes.c_int]
    L.tmd_bench_copy.argtypes = [ctypes.c_size_t, ctypes.c_int, _dp]
    L.tmd_system_neighbors.argtypes = [_vp, ctypes.c_long, _lp, _ip, _ip, _lp]
    L.tmd_system_descriptors.argtypes = [_vp, _fp, _fp]
    _lib = L
    return L


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _check(rc, ctx=None):
    if rc != TMD_OK:
        raise TmdError(rc, lib().tmd_last_error(ctx).decode(errors="replace"))


def bench_copy(nbytes, iters=20):
    """Achieved HBM read+write GB/s of a plain device copy of `nbytes`."""
    g = ctypes.c_double()
    _check(lib().tmd_bench_copy(int(nbytes), int(iters), ctypes.byref(g)), None)
    return g.value


class TensorMD:
    """One model on one device (PairTensorMD::coeff + read_potential_file)."""

    def __init__(self, npz_path, symbols, precision="fp32", interp_dx=0.001, device=-1,
                 cmax_exact=True):
        L = lib()
        opt = Options()
        L.tmd_options_default(ctypes.byref(opt))
        opt.mlp_precision = PRECISION[precision]
        opt.interp_dx = float(interp_dx)
        opt.device = int(device)
        opt.cmax_exact = 1 if cmax_exact else 0
        arr = (ctypes.c_char_p * len(symbols))(*[s.encode() if s else None for s in symbols])
        h = _vp()
        rc = L.tmd_create(str(npz_path).encode(), arr, len(symbols), ctypes.byref(opt),
                          ctypes.byref(h))
        _check(rc, None)
        self.h = h
        self.ntypes = len(symbols)
        self.precision = precision

    def close(self):
        if getattr(self, "h", None):
            lib().tmd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def cutoff(self):
        rc = ctypes.c_double()
        _check(lib().tmd_cutoff(self.h, ctypes.byref(rc)), self.h)
        return rc.value

    def masses(self):
        m = np.zeros(self.ntypes + 1)
        _check(lib().tmd_masses(self.h, _p(m, _dp)), self.h)
        return m

    def type_map(self):
        m = np.zeros(self.ntypes + 1, dtype=np.int32)
        _check(lib().tmd_type_map(self.h, _p(m, _ip)), self.h)
        return m

    def info(self):
        v = np.zeros(5, dtype=np.int32)
        _check(lib().tmd_model_info(self.h, _p(v, _ip)), self.h)
        return dict(nelt=int(v[0]), num_filters=int(v[1]), max_moment=int(v[2]),
                    width=int(v[3]), nlayers=int(v[4]))

    def compute_list(self, nlocal, x, types, ilist, numneigh, neigh_flat, scale=1.0,
                     eatom=True, vatom=True):
        """Pair::compute over a LAMMPS full neighbor list (CSR in ilist order)."""
        x = _f64(x).reshape(-1, 3)
        nall = x.shape[0]
        types = _i32(types)
        ilist = _i32(ilist)
        numneigh = _i32(numneigh)
        neigh_flat = _i32(neigh_flat)
        inum = ilist.size
        # firstneigh[i] points into neigh_flat (entries stored in ilist order)
        ptrs = (_ip * max(nlocal, 1))()
        base = neigh_flat.ctypes.data
        off = 0
        for ii in range(inum):
            i = int(ilist[ii])
            ptrs[i] = ctypes.cast(base + 4 * off, _ip)
            off += int(numneigh[i])
        f = np.zeros((nall, 3))
        ea = np.zeros(nlocal) if eatom else None
        va = np.zeros((nall, 6)) if vatom else None
        e = ctypes.c_double()
        vir = np.zeros(6)
        rc = lib().tmd_compute_list(self.h, int(nlocal), int(nall), _p(x, _dp), _p(types, _ip),
                                    int(inum), _p(ilist, _ip), _p(numneigh, _ip), ptrs,
                                    float(scale), _p(f, _dp), _p(ea, _dp), _p(va, _dp),
                                    ctypes.byref(e), _p(vir, _dp))
        _check(rc, self.h)
        return dict(f=f, eatom=ea, vatom=va, energy=e.value, virial=vir)

    def system(self, x, types, cell, skin=1.0):
        return System(self, x, types, cell, skin)


class System:
    """Device-resident periodic system (positions/velocities stay on the GPU)."""

    def __init__(self, model, x, types, cell, skin):
        self.model = model
        x = _f64(x).reshape(-1, 3)
        self.n = x.shape[0]
        types = _i32(types)
        cell = _f64(cell).reshape(9)
        h = _vp()
        rc = lib().tmd_system_create(model.h, self.n, _p(x, _dp), _p(types, _ip), _p(cell, _dp),
                                     float(skin), ctypes.byref(h))
        _check(rc, model.h)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().tmd_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_positions(self, x):
        x = _f64(x).reshape(-1, 3)
        _check(lib().tmd_system_set_positions(self.h, _p(x, _dp)), self.model.h)

    def set_velocities(self, v):
        v = _f64(v).reshape(-1, 3)
        _check(lib().tmd_system_set_velocities(self.h, _p(v, _dp)), self.model.h)

    def compute(self):
        _check(lib().tmd_system_compute(self.h), self.model.h)
        return self.get()

    def get(self):
        x = np.zeros((self.n, 3))
        v = np.zeros((self.n, 3))
        f = np.zeros((self.n, 3))
        ea = np.zeros(self.n)
        e = ctypes.c_double()
        vir = np.zeros(6)
        _check(lib().tmd_system_get(self.h, _p(x, _dp), _p(v, _dp), _p(f, _dp), _p(ea, _dp),
                                    ctypes.byref(e), _p(vir, _dp)), self.model.h)
        return dict(x=x, v=v, f=f, eatom=ea, energy=e.value, virial=vir)

    def md_run(self, nsteps, dt, history=True):
        ep = np.zeros(nsteps) if history else None
        ek = np.zeros(nsteps) if history else None
        _check(lib().tmd_md_run(self.h, int(nsteps), float(dt), _p(ep, _dp), _p(ek, _dp)),
               self.model.h)
        return ep, ek

    def set_graph(self, enable):
        _check(lib().tmd_system_set_graph(self.h, 1 if enable else 0), self.model.h)

    def profile(self, nsteps, dt):
        """Mean ms per step of each phase, timed with HIP events on the library stream."""
        ms = np.zeros(7)
        _check(lib().tmd_md_profile(self.h, int(nsteps), float(dt), _p(ms, _dp)), self.model.h)
        keys = ["neighbor", "descriptor", "mlp", "force", "reduce", "integrate", "step"]
        return dict(zip(keys, ms.tolist()))

    @property
    def nbuilds(self):
        return int(lib().tmd_system_nbuilds(self.h))

    def neighbors(self):
        off = np.zeros(self.n + 1, dtype=np.int64)
        tot = ctypes.c_long()
        _check(lib().tmd_system_neighbors(self.h, 0, _p(off, _lp), None, None, ctypes.byref(tot)),
               self.model.h)
        cap = tot.value
        j = np.zeros(max(cap, 1), dtype=np.int32)
        s = np.zeros((max(cap, 1), 3), dtype=np.int32)
        _check(lib().tmd_system_neighbors(self.h, cap, _p(off, _lp), _p(j, _ip), _p(s, _ip),
                                          ctypes.byref(tot)), self.model.h)
        return off, j[:cap], s[:cap]

    def descriptors(self, width):
        G = np.zeros((self.n, width), dtype=np.float32)
        dG = np.zeros((self.n, width), dtype=np.float32)
        _check(lib().tmd_system_descriptors(self.h, _p(G, _fp), _p(dG, _fp)), self.model.h)
        return G, dG
