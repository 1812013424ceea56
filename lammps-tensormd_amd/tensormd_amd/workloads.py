"""Synthetic workloads: TensorMD model files and crystals, all seeded.

The model writer produces .npz files in the reference schema (keys read by
TensorMD::read_npz, src/ML-TENSORMD/tensormd.cpp:108-177; NNP
setup_global, src/ML-TENSORMD/nnp.cpp:110-145; FeaturePotential,
src/ML-TENSORMD/fnn.cpp:32-55): int32 scalars, float64 arrays, weights
row-major [in][out], ZIP_STORED like the shipped models.
"""
import numpy as np

# LAMMPS metal units
KB = 8.617343e-5          # eV/K
MVV2E = 1.0364269e-4      # amu (A/ps)^2 -> eV
FTM2V = 1.0 / MVV2E
NKTV2P = 1.6021765e6      # eV/A^3 -> bar

ACT = {"relu": 0, "softplus": 1, "tanh": 2, "squareplus": 3}


def _symbol_codes(symbols):
    codes = []
    for s in symbols:
        c = [ord(ch) for ch in s[:2]]
        codes += c + [0] * (2 - len(c))
    return np.array(codes, dtype=np.int32)


RADIAL = {"pexp": 0, "morse": 1, "density": 2, "sf": 3}


def analytic_radial_params(method, k, rmax, rng):
    """Seeded parameter sets of the analytic radial functions
    (src/ML-TENSORMD/descriptor.cpp:51-132), k filters spread over (0, rmax)
    and scaled so that a dense crystal's descriptors are of order one."""
    centers = np.linspace(0.3 * rmax, 0.9 * rmax, k)
    if method == "pexp":   # exp(-(r / rl)^pl)
        return {"rl": centers, "pl": rng.uniform(1.0, 4.0, k)}
    if method == "morse":  # D (exp(-2 gamma (r - r0)) - 2 exp(-gamma (r - r0)))
        return {"D": rng.uniform(0.005, 0.02, k), "gamma": rng.uniform(0.2, 0.6, k),
                "r0": np.linspace(0.45 * rmax, 0.9 * rmax, k)}
    if method == "density":  # A exp(-beta (r - re) / re)
        return {"A": rng.uniform(0.1, 0.5, k), "beta": rng.uniform(1.0, 4.0, k), "re": centers}
    if method == "sf":  # exp(-eta (r - omega)^2 / rc^2)
        return {"eta": rng.uniform(5.0, 40.0, k), "omega": centers}
    raise ValueError(method)


def write_model(path, *, symbols, masses, rmax, num_filters, layer_sizes, actfn="tanh",
                fnn_hidden=None, fnn_actfn="softplus", max_moment=3, t_symmetric=False,
                use_resnet_dt=False,
                apply_output_bias=True, fctype=0, seed=0, radial_scale=0.1, bias_sigma=0.05,
                radial=None, tdnp=None):
    """Write a random-init TensorMD model (Xavier-normal FP64 weights).

    layer_sizes: main MLP widths, last must be 1 (e.g. [256, 128, 64, 1]).
    fnn_hidden: hidden widths of the radial network (default [num_filters]);
    its output width is num_filters.  radial_scale multiplies the radial
    network's output layer so that the descriptors of a dense crystal are of
    order one (trained models are normalised the same way by training).
    radial: None for the radial network, or one of "pexp", "morse",
    "density", "sf" for analytic radial functions (use_fnn 0, the
    descriptor:: keys of src/ML-TENSORMD/descriptor.h:37-87) with seeded
    parameters.
    tdnp: None, or a dict describing a temperature-dependent model (tdnp 1,
    the H::, U:: and S:: networks of src/ML-TENSORMD/tdnp.cpp:60-148) that
    replaces the main MLP: "H" hidden and output widths of the network from
    the descriptor (its output gains the electron-temperature column when
    read), "U" and "S" widths ending in 1, optional "actfn", "resnet" and
    "sommerfeld".  layer_sizes is then ignored.
    """
    rng = np.random.default_rng(seed)
    nelt = len(symbols)
    k = int(num_filters)
    fnn_hidden = list(fnn_hidden) if fnn_hidden is not None else [k]
    fnn_sizes = fnn_hidden + [k]
    gw = nelt * k * (max_moment + 1)
    z = {}
    i32 = lambda v: np.array(v, dtype=np.int32)
    z["rmax"] = np.array(float(rmax))
    z["nelt"] = i32(nelt)
    z["masses"] = np.array(masses, dtype=np.float64)
    z["numbers"] = _symbol_codes(symbols)
    z["tdnp"] = i32(0)
    z["precision"] = i32(64)

    def xavier(n_in, n_out):
        return rng.normal(0.0, np.sqrt(2.0 / (n_in + n_out)), size=(n_in, n_out))

    if radial is None:
        z["use_fnn"] = i32(1)
        z["fnn::nlayers"] = i32(len(fnn_sizes))
        z["fnn::layer_sizes"] = i32(fnn_sizes)
        z["fnn::num_filters"] = i32(k)
        z["fnn::actfn"] = i32(ACT[fnn_actfn])
        z["fnn::use_resnet_dt"] = i32(0)
        z["fnn::apply_output_bias"] = i32(0)
        prev = 1
        for l, w in enumerate(fnn_sizes):
            W = xavier(prev, w)
            if l == len(fnn_sizes) - 1:
                W *= radial_scale
            z[f"fnn::weights_0_{l}"] = W.reshape(w) if prev == 1 else W
            if l < len(fnn_sizes) - 1:
                z[f"fnn::biases_0_{l}"] = rng.normal(0.0, 0.5, size=w)
            prev = w
    else:
        z["use_fnn"] = i32(0)
        z["descriptor::method"] = i32(RADIAL[radial])
        for key, val in analytic_radial_params(radial, k, rmax, rng).items():
            z[f"descriptor::{key}"] = np.asarray(val, dtype=np.float64)
    z["max_moment"] = i32(max_moment)
    z["fctype"] = i32(fctype)
    z["is_T_symmetric"] = i32(1 if t_symmetric else 0)

    def nets(scope, sizes, n_in, act, resnet, outbias):
        z[f"{scope}nlayers"] = i32(len(sizes))
        z[f"{scope}actfn"] = i32(ACT[act])
        z[f"{scope}layer_sizes"] = i32(sizes)
        z[f"{scope}use_resnet_dt"] = i32(1 if resnet else 0)
        z[f"{scope}apply_output_bias"] = i32(1 if outbias else 0)
        for e in range(nelt):
            prev = n_in
            for l, w in enumerate(sizes):
                W = xavier(prev, w)
                z[f"{scope}weights_{e}_{l}"] = W.reshape(prev) if w == 1 else W
                if l < len(sizes) - 1:
                    z[f"{scope}biases_{e}_{l}"] = rng.normal(0.0, bias_sigma, size=w)
                elif outbias:
                    z[f"{scope}biases_{e}_{l}"] = (np.array(rng.normal(0.0, 0.1)) if w == 1
                                                   else rng.normal(0.0, 0.1, size=w))
                prev = w

    if tdnp is None:
        sizes = list(layer_sizes)
        assert sizes[-1] == 1
        nets("", sizes, gw, actfn, use_resnet_dt, apply_output_bias)
    else:
        z["tdnp"] = i32(1)
        if tdnp.get("sommerfeld"):
            z["tdnp::Sommerfeld"] = i32(1)
        act = tdnp.get("actfn", actfn)
        res = tdnp.get("resnet", use_resnet_dt)
        h = list(tdnp["H"])
        # H's last layer is linear and without bias: the reference appends a
        # zero column to its weights only (tdnp.cpp:91-104)
        nets("H::", h, gw, act, res, False)
        for scope in ("U", "S"):
            sz = list(tdnp[scope])
            assert sz[-1] == 1
            nets(f"{scope}::", sz, h[-1] + 1, act, res, apply_output_bias)
    np.savez(path, **z)
    return path


# ---------------------------------------------------------------------------
def lattice(kind, a, n):
    """Perfect fcc/bcc supercell: (positions [N,3], cell [3,3] rows)."""
    if kind == "fcc":
        base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]], dtype=np.float64)
    elif kind == "bcc":
        base = np.array([[0, 0, 0], [0.5, 0.5, 0.5]], dtype=np.float64)
    else:
        raise ValueError(kind)
    nx, ny, nz = n
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1)
    g = g.reshape(-1, 1, 3).astype(np.float64)
    x = ((g + base[None]) * a).reshape(-1, 3)
    cell = np.diag([nx * a, ny * a, nz * a]).astype(np.float64)
    return x, cell


def wrap(x, cell):
    f = x @ np.linalg.inv(cell)
    f -= np.floor(f)
    f[f >= 1.0] = 0.0
    return f @ cell


def crystal(kind, a, n, *, nelt=1, sigma=0.0, seed=0, equiatomic=True):
    """Jittered crystal with (exact) equiatomic random site occupancy.

    Returns (x [N,3], types [N] 1-based, cell [3,3])."""
    rng = np.random.default_rng(seed)
    x, cell = lattice(kind, a, n)
    N = x.shape[0]
    if nelt == 1:
        types = np.ones(N, dtype=np.int32)
    elif equiatomic:
        types = (np.arange(N) % nelt + 1).astype(np.int32)
        rng.shuffle(types)
    else:
        types = rng.integers(1, nelt + 1, size=N).astype(np.int32)
    if sigma > 0:
        x = x + rng.normal(0.0, sigma, size=x.shape)
    return wrap(x, cell), types, cell


def serving_requests(count, *, seed=0, nmin=16, nmax=4096, nelt=4, sigma=0.1,
                     a_bcc=3.2, a_fcc=4.032):
    """Single-point requests: periodic BCC or FCC supercells whose atom count
    is drawn log-uniform in [nmin, nmax] and rounded to whole cells (at least
    2 BCC cells per axis so that every cell is wider than rc/2), random
    composition over `nelt` elements, Gaussian jitter `sigma`.

    Returns a list of (x [N,3], types [N] 1-based, cell [3,3])."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        target = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        kind = "bcc" if rng.random() < 0.5 else "fcc"
        basis, a, nlo = (2, a_bcc, 2) if kind == "bcc" else (4, a_fcc, 1)
        cells = max(target / basis, nlo ** 3)
        n1 = max(nlo, int(round(cells ** (1.0 / 3.0))))
        n2 = max(nlo, int(round((cells / n1) ** 0.5)))
        n3 = max(nlo, int(round(cells / (n1 * n2))))
        x, cell = lattice(kind, a, (n1, n2, n3))
        types = rng.integers(1, nelt + 1, size=x.shape[0]).astype(np.int32)
        x = x + rng.normal(0.0, sigma, size=x.shape)
        out.append((wrap(x, cell), types, cell))
    return out


def maxwell_boltzmann(masses, temperature, seed=0):
    """Velocities (A/ps) at exactly `temperature` (dof 3N-3), zero momentum."""
    rng = np.random.default_rng(seed)
    m = np.asarray(masses, dtype=np.float64)
    N = m.size
    v = rng.normal(size=(N, 3)) * np.sqrt(KB * temperature / (m * MVV2E))[:, None]
    p = (m[:, None] * v).sum(0) / m.sum()
    v -= p
    ke = 0.5 * MVV2E * (m[:, None] * v * v).sum()
    t_now = 2.0 * ke / ((3 * N - 3) * KB)
    if t_now > 0:
        v *= np.sqrt(temperature / t_now)
    return v


def kinetic_energy(masses, v):
    return 0.5 * MVV2E * (np.asarray(masses)[:, None] * v * v).sum()


# ---------------------------------------------------------------------------
def lammps_ghost_list(x, cell, cut):
    """Build a LAMMPS-style full neighbor list with explicit ghost atoms.

    Returns (xall [nall,3], owner [nall], ilist, numneigh, neigh_flat) where
    atoms 0..N-1 are local and ghosts are periodic images (j, shift) that
    fall within `cut` of some local atom, like LAMMPS' ghost shell after
    Comm::borders.  Brute force over images: for test-sized systems.
    """
    x = np.asarray(x, dtype=np.float64)
    cell = np.asarray(cell, dtype=np.float64)
    N = x.shape[0]
    inv = np.linalg.inv(cell)
    widths = 1.0 / np.linalg.norm(inv, axis=0)
    nimg = [int(np.ceil(cut / w)) + 1 for w in widths]
    shifts = np.array([(a, b, c) for a in range(-nimg[0], nimg[0] + 1)
                       for b in range(-nimg[1], nimg[1] + 1)
                       for c in range(-nimg[2], nimg[2] + 1)], dtype=np.int64)
    ghost_index = {}
    xall = [x]
    owner = list(range(N))
    lists = []
    cut2 = cut * cut
    for i in range(N):
        nb = []
        for s in shifts:
            sh = s @ cell
            d = (x - x[i]) + sh
            r2 = (d * d).sum(1)
            ok = np.nonzero(r2 <= cut2)[0]
            for j in ok:
                if j == i and not s.any():
                    continue
                if not s.any():
                    nb.append(int(j))
                    continue
                key = (int(j), int(s[0]), int(s[1]), int(s[2]))
                g = ghost_index.get(key)
                if g is None:
                    g = N + len(ghost_index)
                    ghost_index[key] = g
                    xall.append((x[j] + sh)[None])
                    owner.append(int(j))
                nb.append(g)
        lists.append(nb)
    xall = np.concatenate(xall, 0)
    numneigh = np.array([len(l) for l in lists], dtype=np.int32)
    neigh_flat = np.array([j for l in lists for j in l], dtype=np.int32)
    ilist = np.arange(N, dtype=np.int32)
    return xall, np.array(owner, dtype=np.int64), ilist, numneigh, neigh_flat
