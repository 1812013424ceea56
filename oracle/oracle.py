"""Python front end of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; it is the checker of the HIP product path, never a fallback for
it.  The model file is read with numpy (no pickle) and pushed key by key into
the C++ restatement in tmd_oracle.cpp, so the oracle shares no model-reading
code with the product library.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libtmd_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_long)


def build(force=False):
    """Compile the oracle with make (g++ only)."""
    args = ["make", "-C", _HERE]
    if force:
        args.append("-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.tmdo_new.restype = ctypes.c_void_p
        L.tmdo_free.argtypes = [ctypes.c_void_p]
        L.tmdo_set_ints.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _ip, ctypes.c_int]
        L.tmdo_set_doubles.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp, ctypes.c_long]
        L.tmdo_finalize.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.tmdo_last_error.restype = ctypes.c_char_p
        L.tmdo_info.argtypes = [ctypes.c_void_p, _ip]
        L.tmdo_get_T.argtypes = [ctypes.c_void_p, _dp]
        L.tmdo_get_table.argtypes = [ctypes.c_void_p, _dp]
        L.tmdo_radial.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.tmdo_fnn.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.tmdo_monomials.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.tmdo_mlp.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.tmdo_compute_list.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _ip,
                                        ctypes.c_int, _ip, _lp, _ip, ctypes.c_double, _dp, _dp,
                                        _dp, _dp, _dp, _dp, _dp]
        L.tmdo_neighbors_periodic.restype = ctypes.c_long
        L.tmdo_neighbors_periodic.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double,
                                              ctypes.c_long, _lp, _ip, _ip]
        L.tmdo_neighbors_binned.restype = ctypes.c_long
        L.tmdo_neighbors_binned.argtypes = L.tmdo_neighbors_periodic.argtypes
        L.tmdo_compute_periodic.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _ip, _dp, _lp,
                                            _ip, _ip, ctypes.c_double, _dp, _dp, _dp, _dp, _dp,
                                            _dp, _dp]
        L.tmdo_nve.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _ip, _dp, _dp,
                               ctypes.c_double, ctypes.c_int, _dp, _dp, _dp]
        L.tmdo_num_threads.restype = ctypes.c_int
        L.tmdo_set_threads.argtypes = [ctypes.c_int]
        L.tmdo_set_emulation.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.tmdo_set_etemp.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.tmdo_tdnp_last.argtypes = [ctypes.c_void_p, ctypes.c_long, _dp, _dp, _dp, _dp]
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


class OracleModel:
    """FP64 TensorMD model: pushes every npz key into the C++ oracle."""

    def __init__(self, npz_path, interp_dx):
        L = lib()
        self._path = str(npz_path)
        self.h = L.tmdo_new()
        with np.load(npz_path, allow_pickle=False) as z:
            for key in z.files:
                a = z[key]
                if a.dtype.kind in "iu":
                    v = _i32(a.ravel())
                    L.tmdo_set_ints(self.h, key.encode(), _p(v, _ip), v.size)
                else:
                    v = _f64(a.ravel())
                    L.tmdo_set_doubles(self.h, key.encode(), _p(v, _dp), v.size)
            self.masses = np.array(z["masses"], dtype=np.float64).ravel()
            nums = np.array(z["numbers"]).ravel()
        self.symbols = []
        for e in range(len(self.masses)):
            s = chr(int(nums[2 * e]))
            if 2 * e + 1 < nums.size and nums[2 * e + 1] > 0:
                s += chr(int(nums[2 * e + 1]))
            self.symbols.append(s)
        if L.tmdo_finalize(self.h, float(interp_dx)) != 0:
            raise ValueError(L.tmdo_last_error().decode())
        info = np.zeros(6, dtype=np.int32)
        L.tmdo_info(self.h, _p(info, _ip))
        self.nelt, self.k, self.nm, self.nd, self.tab_n, self.gwidth = [int(v) for v in info]
        with np.load(npz_path, allow_pickle=False) as z:
            self.rc = float(np.array(z["rmax"]).ravel()[0])

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.tmdo_free(self.h)
            self.h = None

    # Stages of the chain evaluated in FP32 by the reduced-precision
    # emulation (tmd_oracle.cpp atom_chain_emul): it prices the rounding
    # floor of an FP32 evaluation; it is not the reference.
    EMUL_DESC, EMUL_MLP, EMUL_FORCE, EMUL_FORCE_TEMPLATE = 1, 2, 4, 8

    def set_emulation(self, flags):
        """0 = the FP64 chain; otherwise a bitmask of EMUL_* stages in FP32."""
        lib().tmdo_set_emulation(self.h, int(flags))
        return self

    def set_etemp(self, etemp_ev):
        """Electron temperature (eV) of a TDNP model."""
        lib().tmdo_set_etemp(self.h, float(etemp_ev))
        return self

    def tdnp_last(self, n):
        """TDNP results of the last compute call on n atoms (nlocal in list
        mode): per-atom electron entropy and free energy (scaled), and their
        sums S and A."""
        sa = np.zeros(n)
        aa = np.zeros(n)
        S = np.zeros(1)
        A = np.zeros(1)
        if lib().tmdo_tdnp_last(self.h, int(n), _p(sa, _dp), _p(aa, _dp), _p(S, _dp), _p(A, _dp)) != 0:
            raise RuntimeError(lib().tmdo_last_error().decode())
        return dict(satom=sa, aatom=aa, entropy=float(S[0]), free_energy=float(A[0]))

    # ---- pieces -------------------------------------------------------
    def T(self):
        t = np.zeros((self.nm, self.nd))
        lib().tmdo_get_T(self.h, _p(t, _dp))
        return t

    def table(self):
        t = np.zeros((self.tab_n, 4, self.k))
        lib().tmdo_get_table(self.h, _p(t, _dp))
        return t

    def radial(self, r):
        r = _f64(r)
        H = np.zeros((r.size, self.k))
        dH = np.zeros((r.size, self.k))
        lib().tmdo_radial(self.h, r.size, _p(r, _dp), _p(H, _dp), _p(dH, _dp))
        return H, dH

    def fnn(self, r):
        r = _f64(r)
        y = np.zeros((r.size, self.k))
        lib().tmdo_fnn(self.h, r.size, _p(r, _dp), _p(y, _dp))
        return y

    def monomials(self, u):
        u = _f64(u).reshape(-1, 3)
        M = np.zeros((u.shape[0], self.nd))
        dM = np.zeros((u.shape[0], 3, self.nd))
        lib().tmdo_monomials(self.h, u.shape[0], _p(u, _dp), _p(M, _dp), _p(dM, _dp))
        return M, dM

    def mlp(self, elem, G):
        G = _f64(G).reshape(-1, self.gwidth)
        E = np.zeros(G.shape[0])
        dG = np.zeros_like(G)
        lib().tmdo_mlp(self.h, int(elem), G.shape[0], _p(G, _dp), _p(E, _dp), _p(dG, _dp))
        return E, dG

    # ---- LAMMPS-shaped compute ----------------------------------------
    def compute_list(self, nlocal, x, elem, ilist, numneigh, neigh_flat, scale=1.0,
                     want_desc=False):
        x = _f64(x).reshape(-1, 3)
        nall = x.shape[0]
        elem = _i32(elem).reshape(-1)
        _same_len(elem, nall, "elem")
        ilist = _i32(ilist)
        numneigh = np.asarray(numneigh, dtype=np.int64)
        off = _i64(np.concatenate([[0], np.cumsum(numneigh[ilist])]))
        neigh_flat = _i32(neigh_flat)
        f = np.zeros((nall, 3))
        eatom = np.zeros(nlocal)
        vatom = np.zeros((nall, 6))
        e = np.zeros(1)
        vir = np.zeros(6)
        G = dG = None
        if want_desc:
            G = np.zeros((ilist.size, self.gwidth))
            dG = np.zeros_like(G)
        rc = lib().tmdo_compute_list(self.h, nlocal, nall, _p(x, _dp), _p(elem, _ip), ilist.size,
                                     _p(ilist, _ip), _p(off, _lp), _p(neigh_flat, _ip),
                                     float(scale), _p(f, _dp), _p(eatom, _dp), _p(vatom, _dp),
                                     _p(e, _dp), _p(vir, _dp), _p(G, _dp), _p(dG, _dp))
        if rc != 0:
            raise RuntimeError(lib().tmdo_last_error().decode())
        out = dict(f=f, eatom=eatom, vatom=vatom, energy=float(e[0]), virial=vir)
        if want_desc:
            out["G"] = G
            out["dEdG"] = dG
        return out

    # ---- periodic structures ------------------------------------------
    def compute_periodic(self, x, elem, cell, nbr=None, scale=1.0, want_desc=False):
        x = _f64(x).reshape(-1, 3)
        n = x.shape[0]
        cell = _f64(cell).reshape(3, 3)
        elem = _i32(elem).reshape(-1)
        _same_len(elem, n, "elem")
        if nbr is None:
            nbr = neighbors(x, cell, self.rc)
        off, nj, ns = nbr
        f = np.zeros((n, 3))
        eatom = np.zeros(n)
        vatom = np.zeros((n, 6))
        e = np.zeros(1)
        vir = np.zeros(6)
        G = dG = None
        if want_desc:
            G = np.zeros((n, self.gwidth))
            dG = np.zeros_like(G)
        rc = lib().tmdo_compute_periodic(self.h, n, _p(x, _dp), _p(elem, _ip), _p(cell, _dp),
                                         _p(off, _lp), _p(nj, _ip), _p(ns, _ip), float(scale),
                                         _p(f, _dp), _p(e, _dp), _p(eatom, _dp), _p(vatom, _dp),
                                         _p(vir, _dp), _p(G, _dp), _p(dG, _dp))
        if rc != 0:
            raise RuntimeError(lib().tmdo_last_error().decode())
        out = dict(f=f, eatom=eatom, vatom=vatom, energy=float(e[0]), virial=vir)
        if want_desc:
            out["G"] = G
            out["dEdG"] = dG
        return out

    def nve(self, x, v, elem, cell, dt, nsteps, keep_forces=False):
        x = _f64(x).reshape(-1, 3).copy()
        v = _f64(v).reshape(-1, 3).copy()
        n = x.shape[0]
        elem = _i32(elem).reshape(-1)
        _same_len(elem, n, "elem")
        _same_len(v, n, "v")
        cell = _f64(cell).reshape(3, 3)
        masses = _f64(self.masses)
        ep = np.zeros(nsteps + 1)
        ek = np.zeros(nsteps + 1)
        fh = np.zeros((nsteps + 1, n, 3)) if keep_forces else None
        rc = lib().tmdo_nve(self.h, n, _p(x, _dp), _p(v, _dp), _p(elem, _ip), _p(cell, _dp),
                            _p(masses, _dp), float(dt), int(nsteps), _p(ep, _dp), _p(ek, _dp),
                            _p(fh, _dp))
        if rc != 0:
            raise RuntimeError(lib().tmdo_last_error().decode())
        return dict(x=x, v=v, epot=ep, ekin=ek, forces=fh)


def _same_len(a, n, name):
    """The C functions read n entries: refuse shorter (or longer) arrays."""
    if a.shape[0] != n:
        raise ValueError(f"oracle: {name} has {a.shape[0]} rows, expected {n}")


def neighbors(x, cell, cut, method="binned", cap_per_atom=400):
    """Full periodic neighbor list (j, image shift) within `cut` (r^2 <= cut^2).

    Positions must be wrapped into the cell.  Returns (off[n+1], j[npairs],
    shift[npairs][3]); within an atom, entries are sorted by (j, shift).
    """
    x = _f64(x).reshape(-1, 3)
    cell = _f64(cell).reshape(3, 3)
    n = x.shape[0]
    cap = n * cap_per_atom
    while True:
        off = np.zeros(n + 1, dtype=np.int64)
        nj = np.zeros(cap, dtype=np.int32)
        ns = np.zeros((cap, 3), dtype=np.int32)
        fn = lib().tmdo_neighbors_binned if method == "binned" else lib().tmdo_neighbors_periodic
        tot = fn(n, _p(x, _dp), _p(cell, _dp), float(cut), cap, _p(off, _lp), _p(nj, _ip),
                 _p(ns, _ip))
        if tot <= cap:
            return off, nj[:tot].copy(), ns[:tot].copy()
        cap = int(tot)


def num_threads():
    return int(lib().tmdo_num_threads())


def set_threads(n):
    lib().tmdo_set_threads(int(n))
