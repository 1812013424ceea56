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
TMD_PENDING = 1
ERRORS = {-1: "TMD_ERR_ARG", -2: "TMD_ERR_MODEL", -3: "TMD_ERR_DEVICE", -4: "TMD_ERR_NOMEM",
          -5: "TMD_ERR_CAPACITY", -6: "TMD_ERR_STATE"}
PRECISION = {"fp32": 0, "bf16": 1, "fp8": 2, "bf16x3": 3}

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
                ("cmax_exact", ctypes.c_int), ("device", ctypes.c_int),
                ("table_device", ctypes.c_int)]


class ServeOptions(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("max_batch_atoms", ctypes.c_int),
                ("max_batch_structs", ctypes.c_int), ("hbm_fraction", ctypes.c_double),
                ("use_graph", ctypes.c_int), ("autotune", ctypes.c_int),
                ("max_queue", ctypes.c_int)]


class Structure(ctypes.Structure):
    _fields_ = [("natoms", ctypes.c_int), ("x", ctypes.POINTER(ctypes.c_double)),
                ("type", ctypes.POINTER(ctypes.c_int)), ("cell", ctypes.c_double * 9),
                ("f", ctypes.POINTER(ctypes.c_double)),
                ("eatom", ctypes.POINTER(ctypes.c_double))]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("energy", ctypes.c_double),
                ("virial", ctypes.c_double * 6), ("t_submit", ctypes.c_double),
                ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
                ("batch_structs", ctypes.c_int), ("batch_atoms", ctypes.c_int)]


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
    L.tmd_radial_table.argtypes = [_vp, _dp, ctypes.c_long, _ip, _dp]
    L.tmd_compute_list.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _ip, ctypes.c_int, _ip,
                                   _ip, ctypes.POINTER(_ip), ctypes.c_double, _dp, _dp, _dp, _dp,
                                   _dp]
    L.tmd_compute_list_serial.argtypes = [_vp, ctypes.c_longlong] + L.tmd_compute_list.argtypes[1:]
    L.tmd_compute_list_noalloc.argtypes = L.tmd_compute_list.argtypes
    L.tmd_descriptors_list.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _dp, _ip, ctypes.c_int, _ip,
                                       _ip, ctypes.POINTER(_ip), _dp]
    L.tmd_set_electron_temperature.argtypes = [_vp, ctypes.c_double]
    L.tmd_tdnp_results.argtypes = [_vp, _dp, _dp, _dp, _dp]
    L.tmd_system_create.argtypes = [_vp, ctypes.c_int, _dp, _ip, _dp, ctypes.c_double,
                                    ctypes.POINTER(_vp)]
    L.tmd_system_destroy.argtypes = [_vp]
    L.tmd_md_step_async.argtypes = [_vp, ctypes.c_double]
    L.tmd_system_set_positions.argtypes = [_vp, _dp]
    L.tmd_system_set_velocities.argtypes = [_vp, _dp]
    L.tmd_system_compute.argtypes = [_vp]
    L.tmd_system_get.argtypes = [_vp, _dp, _dp, _dp, _dp, _dp, _dp]
    L.tmd_md_run.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp, _dp]
    L.tmd_system_nbuilds.argtypes = [_vp]
    L.tmd_md_profile.argtypes = [_vp, ctypes.c_int, ctypes.c_double, _dp]
    L.tmd_system_set_graph.argtypes = [_vp, ctypes.c_int]
    L.tmd_system_set_kernel_timing.argtypes = [_vp, ctypes.c_int]
    L.tmd_system_kernel_timing.argtypes = [_vp, _dp, _dp, ctypes.POINTER(ctypes.c_longlong)]
    L.tmd_bench_copy.argtypes = [ctypes.c_size_t, ctypes.c_int, _dp]
    L.tmd_device_memory.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    L.tmd_system_neighbors.argtypes = [_vp, ctypes.c_long, _lp, _ip, _ip, _lp]
    L.tmd_system_descriptors.argtypes = [_vp, _fp, _fp]
    L.tmd_serve_options_default.argtypes = [ctypes.POINTER(ServeOptions)]
    L.tmd_serve_start.argtypes = [_vp, ctypes.POINTER(ServeOptions)]
    L.tmd_serve_submit.argtypes = [_vp, ctypes.POINTER(Structure), ctypes.POINTER(ctypes.c_longlong)]
    L.tmd_serve_poll.argtypes = [_vp, ctypes.c_longlong, ctypes.POINTER(Result)]
    L.tmd_serve_wait.argtypes = [_vp, ctypes.c_longlong, ctypes.POINTER(Result)]
    L.tmd_serve_stop.argtypes = [_vp]
    L.tmd_serve_stats.argtypes = [_vp, _dp]
    L.tmd_serve_replay.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(Structure), _dp,
                                   ctypes.POINTER(Result)]
    L.tmd_clock.restype = ctypes.c_double
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


def device_memory():
    """(free, total) bytes of device memory on the current device."""
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    _check(lib().tmd_device_memory(ctypes.byref(f), ctypes.byref(t)), None)
    return f.value, t.value


class NeighborList:
    """A LAMMPS-style full neighbor list held for the C ABI: ilist,
    numneigh and the firstneigh pointer array into one CSR buffer (entries
    stored in ilist order)."""

    def __init__(self, nlocal, ilist, numneigh, neigh_flat):
        self.nlocal = int(nlocal)
        self.ilist = _i32(ilist)
        self.numneigh = _i32(numneigh)
        self.flat = _i32(neigh_flat)
        self.ptrs = (_ip * max(self.nlocal, 1))()
        counts = self.numneigh[self.ilist].astype(np.int64)
        starts = np.concatenate([[0], np.cumsum(counts)[:-1]]) if counts.size else counts
        base = self.flat.ctypes.data
        addr = (ctypes.c_void_p * max(self.nlocal, 1)).from_buffer(self.ptrs)
        for i, st in zip(self.ilist.tolist(), starts.tolist()):
            addr[i] = base + 4 * st


class TensorMD:
    """One model on one device (PairTensorMD::coeff + read_potential_file)."""

    def __init__(self, npz_path, symbols, precision="fp32", interp_dx=0.001, device=-1,
                 cmax_exact=True, table_device=False):
        L = lib()
        opt = Options()
        L.tmd_options_default(ctypes.byref(opt))
        opt.mlp_precision = PRECISION[precision]
        opt.interp_dx = float(interp_dx)
        opt.device = int(device)
        opt.cmax_exact = 1 if cmax_exact else 0
        opt.table_device = 1 if table_device else 0
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

    def radial_table(self):
        """(table[n][num_filters][4] = y, slope, c1, c2 in FP64, dx)."""
        n, dx = ctypes.c_int(), ctypes.c_double()
        _check(lib().tmd_radial_table(self.h, None, 0, ctypes.byref(n), ctypes.byref(dx)), self.h)
        tab = np.zeros((n.value, self.info()["num_filters"], 4))
        _check(lib().tmd_radial_table(self.h, _p(tab, _dp), tab.size, ctypes.byref(n),
                                      ctypes.byref(dx)), self.h)
        return tab, dx.value

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
                     eatom=True, vatom=True, serial=None, noalloc=False):
        """Pair::compute over a LAMMPS full neighbor list (CSR in ilist order).

        serial: call tmd_compute_list_serial with this list serial (the list
        stays on the device while the serial repeats); noalloc: call
        tmd_compute_list_noalloc (fixed device scratch, no allocation)."""
        lst = NeighborList(nlocal, ilist, numneigh, neigh_flat)
        return self.compute_list_ptrs(lst, x, types, scale=scale, eatom=eatom, vatom=vatom,
                                      serial=serial, noalloc=noalloc)

    def compute_list_ptrs(self, lst, x, types, scale=1.0, eatom=True, vatom=True, serial=None,
                          noalloc=False, out=None):
        """compute_list over a prepared NeighborList; `out` (a dict from an
        earlier call with the same shapes) is overwritten in place."""
        x = _f64(x).reshape(-1, 3)
        nall = x.shape[0]
        types = _i32(types)
        nlocal = lst.nlocal
        if out is None:
            out = dict(f=np.zeros((nall, 3)), eatom=np.zeros(nlocal) if eatom else None,
                       vatom=np.zeros((nall, 6)) if vatom else None, energy=0.0,
                       virial=np.zeros(6))
        e = ctypes.c_double()
        args = (int(nlocal), int(nall), _p(x, _dp), _p(types, _ip), int(lst.ilist.size),
                _p(lst.ilist, _ip), _p(lst.numneigh, _ip), lst.ptrs, float(scale),
                _p(out["f"], _dp), _p(out["eatom"], _dp), _p(out["vatom"], _dp),
                ctypes.byref(e), _p(out["virial"], _dp))
        if noalloc:
            rc = lib().tmd_compute_list_noalloc(self.h, *args)
        elif serial is not None:
            rc = lib().tmd_compute_list_serial(self.h, int(serial), *args)
        else:
            rc = lib().tmd_compute_list(self.h, *args)
        _check(rc, self.h)
        out["energy"] = e.value
        return out

    def set_electron_temperature(self, etemp_ev):
        """Electron temperature (eV) read by TDNP models."""
        _check(lib().tmd_set_electron_temperature(self.h, float(etemp_ev)), self.h)

    def tdnp_results(self, nlocal):
        """Electron entropy and free energy (scaled sums) and their per-atom
        values of the last compute_list call of a TDNP model."""
        S = ctypes.c_double()
        A = ctypes.c_double()
        sa = np.zeros(nlocal)
        aa = np.zeros(nlocal)
        _check(lib().tmd_tdnp_results(self.h, ctypes.byref(S), ctypes.byref(A), _p(sa, _dp), _p(aa, _dp)),
               self.h)
        return dict(entropy=S.value, free_energy=A.value, satom=sa, aatom=aa)

    def descriptors_list(self, nlocal, x, types, ilist, numneigh, neigh_flat):
        """compute tensormd/density: descriptor rows [nlocal, width] (zero for
        atoms not in ilist) from a LAMMPS-style full neighbor list."""
        x = _f64(x).reshape(-1, 3)
        types = _i32(types)
        ilist = _i32(ilist)
        numneigh = _i32(numneigh)
        neigh_flat = _i32(neigh_flat)
        ptrs = (_ip * max(nlocal, 1))()
        base = neigh_flat.ctypes.data
        off = 0
        for i in ilist:
            ptrs[int(i)] = ctypes.cast(base + 4 * off, _ip)
            off += int(numneigh[i])
        G = np.zeros((nlocal, self.info()["width"]))
        rc = lib().tmd_descriptors_list(self.h, int(nlocal), x.shape[0], _p(x, _dp), _p(types, _ip),
                                        ilist.size, _p(ilist, _ip), _p(numneigh, _ip), ptrs,
                                        _p(G, _dp))
        _check(rc, self.h)
        return G

    def system(self, x, types, cell, skin=1.0):
        return System(self, x, types, cell, skin)

    def serve(self, **opts):
        return Server(self, **opts)


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

    def step_async(self, dt):
        """One MD step enqueued without waiting (tmd_md_step_async)."""
        _check(lib().tmd_md_step_async(self.h, float(dt)), self.model.h)

    def md_run(self, nsteps, dt, history=True):
        ep = np.zeros(nsteps) if history else None
        ek = np.zeros(nsteps) if history else None
        _check(lib().tmd_md_run(self.h, int(nsteps), float(dt), _p(ep, _dp), _p(ek, _dp)),
               self.model.h)
        return ep, ek

    def set_graph(self, enable):
        _check(lib().tmd_system_set_graph(self.h, 1 if enable else 0), self.model.h)

    def set_kernel_timing(self, every):
        """Sample every `every`-th step of later md_run calls kernel by kernel,
        with HIP events around the descriptor, MLP and force kernels (0: off).
        Clears the samples."""
        _check(lib().tmd_system_set_kernel_timing(self.h, int(every)), self.model.h)

    def kernel_timing(self):
        """Mean kernel ms over the sampled steps, the mean kept neighbors per
        atom at those steps, and the number of samples."""
        ms = np.zeros(3)
        kept = np.zeros(1)
        n = ctypes.c_longlong()
        _check(lib().tmd_system_kernel_timing(self.h, _p(ms, _dp), _p(kept, _dp), ctypes.byref(n)),
               self.model.h)
        return dict(descriptor=float(ms[0]), mlp=float(ms[1]), force=float(ms[2]),
                    kept_per_atom=float(kept[0]), samples=int(n.value))

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


class Ticket:
    """One submitted request; holds the output arrays until it completes."""

    def __init__(self, server, tid, natoms, f, eatom, keep):
        self.server = server
        self.id = tid
        self.natoms = natoms
        self.f = f
        self.eatom = eatom
        self._keep = keep
        self.result = None

    def _fill(self, r):
        self.result = dict(status=r.status, energy=r.energy, virial=np.array(r.virial[:]),
                           f=self.f, eatom=self.eatom, t_submit=r.t_submit, t_start=r.t_start,
                           t_done=r.t_done, batch_structs=r.batch_structs,
                           batch_atoms=r.batch_atoms)
        self._keep = None
        return self.result


class Server:
    """Batched single-point requests on one model (continuous batching)."""

    def __init__(self, model, slots=2, max_batch_atoms=65536, max_batch_structs=512,
                 hbm_fraction=0.25, use_graph=True, autotune=True, max_queue=65536):
        self.model = model
        o = ServeOptions()
        lib().tmd_serve_options_default(ctypes.byref(o))
        o.slots, o.max_batch_atoms, o.max_batch_structs = slots, max_batch_atoms, max_batch_structs
        o.hbm_fraction = hbm_fraction
        o.use_graph = 1 if use_graph else 0
        o.autotune = 1 if autotune else 0
        o.max_queue = max_queue
        _check(lib().tmd_serve_start(model.h, ctypes.byref(o)), model.h)
        self.running = True
        self._live = {}  # outstanding tickets keep their output arrays alive

    def submit(self, x, types, cell, forces=True, eatom=False):
        x = _f64(x).reshape(-1, 3)
        types = _i32(types)
        cell = _f64(cell).reshape(9)
        n = x.shape[0]
        f = np.zeros((n, 3)) if forces else None
        ea = np.zeros(n) if eatom else None
        st = Structure()
        st.natoms = n
        st.x = _p(x, _dp)
        st.type = _p(types, _ip)
        for q in range(9):
            st.cell[q] = cell[q]
        st.f = _p(f, _dp)
        st.eatom = _p(ea, _dp)
        tid = ctypes.c_longlong()
        _check(lib().tmd_serve_submit(self.model.h, ctypes.byref(st), ctypes.byref(tid)),
               self.model.h)
        t = Ticket(self, tid.value, n, f, ea, (x, types))
        self._live[t.id] = t
        return t

    def poll(self, ticket):
        """Result dict if done, else None."""
        if ticket.result is not None:
            return ticket.result
        r = Result()
        rc = lib().tmd_serve_poll(self.model.h, ticket.id, ctypes.byref(r))
        if rc == TMD_PENDING:
            return None
        _check(rc, self.model.h)
        self._live.pop(ticket.id, None)
        return ticket._fill(r)

    def wait(self, ticket):
        if ticket.result is not None:
            return ticket.result
        r = Result()
        _check(lib().tmd_serve_wait(self.model.h, ticket.id, ctypes.byref(r)), self.model.h)
        self._live.pop(ticket.id, None)
        return ticket._fill(r)

    def stats(self):
        v = np.zeros(24)
        _check(lib().tmd_serve_stats(self.model.h, _p(v, _dp)), self.model.h)
        nb = int((v[16:24] > 0).sum())
        return dict(batches=int(v[0]), structs=int(v[1]), atoms=int(v[2]), captures=int(v[3]),
                    reruns=int(v[4]), gpu_ms=float(v[5]), slot_atoms=int(v[6]), slots=int(v[7]),
                    row_tiles={int(v[16 + b]): int(v[8 + b]) for b in range(nb)})

    def replay(self, structures, arrival, forces=True):
        """Submit structures[k] = (x, types, cell) at arrival[k] s after the
        call from a library loop (no Python in the timing), wait for all.
        Returns (list of result dicts, force arrays)."""
        n = len(structures)
        arr = (Structure * max(n, 1))()
        keep = []
        fs = []
        for k, (x, types, cell) in enumerate(structures):
            x = _f64(x).reshape(-1, 3)
            types = _i32(types)
            cell = _f64(cell).reshape(9)
            f = np.zeros((x.shape[0], 3)) if forces else None
            keep.append((x, types))
            fs.append(f)
            arr[k].natoms = x.shape[0]
            arr[k].x = _p(x, _dp)
            arr[k].type = _p(types, _ip)
            for q in range(9):
                arr[k].cell[q] = cell[q]
            arr[k].f = _p(f, _dp)
            arr[k].eatom = None
        at = _f64(arrival)
        res = (Result * max(n, 1))()
        _check(lib().tmd_serve_replay(self.model.h, n, arr, _p(at, _dp), res), self.model.h)
        out = [dict(status=r.status, energy=r.energy, virial=np.array(r.virial[:]),
                    t_submit=r.t_submit, t_start=r.t_start, t_done=r.t_done,
                    batch_structs=r.batch_structs, batch_atoms=r.batch_atoms)
               for r in res[:n]]
        return out, fs

    def stop(self):
        """Finish every queued request, then stop the loop."""
        if self.running:
            self.running = False
            _check(lib().tmd_serve_stop(self.model.h), self.model.h)
            self._live.clear()


def clock():
    """The library's monotonic clock (seconds), the clock of request times."""
    return lib().tmd_clock()
