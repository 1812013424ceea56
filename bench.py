"""Benchmark: device-resident MD with pair_style tensormd on one MI355X.

Workload (the production-scale configuration): 4-element BCC random solid
solution, a = 3.2 A, 40x40x40 cells = 128,000 atoms, Gaussian jitter
sigma = 0.08 A, Maxwell-Boltzmann velocities at 600 K, random-init model
(cutoff 6.5 A, 24 radial functions, max moment 3, per-element MLP
256-128-64-1 tanh), velocity-Verlet at dt = 1 fs with device neighbor
rebuilds (skin 1 A).  A step = velocity Verlet + neighbor check/rebuild +
descriptor + MLP forward/backward + forces/virial + FP64 energy/virial sums.

`value` is the BF16-MLP path's atom-steps/s; the plain FP32 path is timed in
the same run (`fp32_atom_steps_per_s`).  `roofline` prices the dominant
kernel's algorithmic FLOPs at its mean duration over the timed run (every
10th step launched kernel by kernel with HIP events on the library stream)
and at the kept-neighbor count of those steps.  `nve_drift` is each path's
total-energy drift over a separate untimed 1000-step run from the same
initial state.

Serving (the `serving` object): 4096 single-point requests, periodic BCC/FCC
supercells with log-uniform atom counts in [16, 4096] and random
composition over the 4 elements (jitter 0.1 A), same model.  Measured:
one-at-a-time FP32 requests (one slot, one structure per batch), the
saturated BF16 request loop (all submitted at once), and Poisson arrivals at
0.5x and 0.9x of that capacity; latency = submit -> results on the host.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lammps-tensormd_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

PEAK = {"hbm_gbs": 8000.0, "fp32_tflops": 157.3, "bf16_tflops": 2500.0}
KB_EV = 8.617343e-5          # Boltzmann constant, eV/K
MVV2E = 1.0364269e-4         # (g/mol) (A/ps)^2 -> eV


def per_atom_counts(c, *, d, k, b, m, gw, sizes, g_bytes=4):
    """Algorithmic FLOPs and compulsory HBM bytes per atom of each kernel.

    c = mean neighbors within rc, d = 20 monomials, k radial functions,
    b elements, m + 1 moments, gw = b k (m + 1) descriptor width.
    descriptor: P = sum_c M (x) H (2 c d k) and the radial interpolation
    (19 c k); G from P is small.  force: P again and U, V (2 x 2 c d k),
    per-neighbor chain rule c (2 k + 2 d + 6), S from dE/dG 2 b k d (m + 1).
    MLP: forward + input gradient 4 sum n_l n_{l+1}.
    bytes: list entries within rc (4 c), x + type (28), G write or read
    (g_bytes gw: 4 for FP32 rows, 2 where the BF16 path stores them as BF16),
    dE/dG write or read (4 gw), force write (24)."""
    mlp = 4 * sum(a * b_ for a, b_ in zip(sizes[:-1], sizes[1:]))
    return dict(
        descriptor_flop=2 * c * d * k + 19 * c * k,
        descriptor_bytes=4 * c + 28 + g_bytes * gw,
        force_flop=2 * (2 * c * d * k) + c * (2 * k + 2 * d + 6) + 2 * b * k * d * (m + 1),
        force_bytes=4 * c + 28 + 4 * gw + 24,
        mlp_flop=mlp,
        mlp_bytes=(g_bytes + 4) * gw + 8,
    )


def neighbors_within(system, x, cell, rc):
    """Mean number of neighbors within rc (from the device list, rc + skin)."""
    off, j, sh = system.neighbors()
    i = np.repeat(np.arange(len(off) - 1), np.diff(off))
    d = x[j] + sh @ cell - x[i]
    return float((np.einsum("ij,ij->i", d, d) < rc * rc).sum() / (len(off) - 1))


def kinetic_temperature(mass, v):
    """Kinetic temperature (K) of velocities v (A/ps) with masses (g/mol)."""
    mv2 = float((np.asarray(mass)[:, None] * np.asarray(v) ** 2).sum())
    return mv2 * MVV2E / (3.0 * len(mass) * KB_EV)


def _traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), if present."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_128k_bcc4.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None
    return dict(bytes_per_launch=k["bytes_per_launch"], source=os.path.relpath(path, ROOT))


def _span(res):
    """Seconds from the first submit to the last result (library clock)."""
    return max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)


def _lat(res, start="t_submit"):
    lat = np.array([r["t_done"] - r[start] for r in res]) * 1e3
    return dict(p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                max_ms=float(lat.max()))


def serve_bench(model, symbols, nreq, seed=11):
    """Request-loop measurements (see the module docstring)."""
    from tensormd_amd import TensorMD, workloads as W
    reqs = W.serving_requests(nreq, seed=seed)
    natoms = int(sum(len(r[0]) for r in reqs))
    zero = np.zeros(nreq)
    out = dict(requests=nreq, atoms=natoms, mean_atoms=natoms / nreq)

    m32 = TensorMD(model, symbols, "fp32", 0.001, device=0)
    srv = m32.serve(slots=1, max_batch_structs=1)
    srv.replay(reqs[:64], zero[:64])  # warm: captures and tile tuning
    res, _ = srv.replay(reqs, zero)
    dt = _span(res)
    # one at a time: one slot, one structure per batch; all requests queue at
    # once, so the latency reported is each request's own service time
    # (batch launch -> results on the host)
    serial = dict(structs_per_s=nreq / dt, **_lat(res, "t_start"))
    assert all(r["status"] == 0 for r in res)
    srv.stop()
    m32.close()
    out["serial_fp32"] = serial

    mb = TensorMD(model, symbols, "bf16", 0.001, device=0)
    srv = mb.serve()
    srv.replay(reqs[:512], zero[:512])  # warm every bucket
    res, _ = srv.replay(reqs, zero)
    dt = _span(res)
    assert all(r["status"] == 0 for r in res)
    cap = nreq / dt
    out["saturated_bf16"] = dict(structs_per_s=cap, atoms_per_s=natoms / dt,
                                 mean_batch=float(np.mean([r["batch_structs"] for r in res])),
                                 **_lat(res))
    rng = np.random.default_rng(seed + 1)
    for frac in (0.5, 0.9):
        rate = frac * cap
        arr = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
        arr -= arr[0]
        res, _ = srv.replay(reqs, arr)
        dt = _span(res)
        assert all(r["status"] == 0 for r in res)
        out[f"poisson_{frac}x"] = dict(offered_per_s=rate, structs_per_s=nreq / dt,
                                        mean_batch=float(np.mean([r["batch_structs"] for r in res])),
                                        **_lat(res))
    out["stats"] = srv.stats()
    srv.stop()
    mb.close()
    out["batched_over_serial"] = out["saturated_bf16"]["structs_per_s"] / serial["structs_per_s"]
    return out


def ghost_list_from_system(system, x, cell):
    """LAMMPS-style full neighbor list with explicit ghosts, from the device
    system's list (cutoff rc + skin): an entry with a nonzero image shift
    becomes a ghost position x_j + shift cell, one ghost per distinct
    (j, shift), numbered after the local atoms."""
    off, j, sh = system.neighbors()
    n = len(x)
    code = ((sh[:, 0] + 3) * 49 + (sh[:, 1] + 3) * 7 + (sh[:, 2] + 3)).astype(np.int64)
    key = j.astype(np.int64) * 343 + code
    image = code != 3 * 49 + 3 * 7 + 3
    gkey, inv = np.unique(key[image], return_inverse=True)
    flat = j.astype(np.int64)
    flat[image] = n + inv
    gj, gc = gkey // 343, gkey % 343
    gsh = np.stack([gc // 49 - 3, (gc // 7) % 7 - 3, gc % 7 - 3], 1).astype(np.float64)
    xall = np.concatenate([x, x[gj] + gsh @ cell])
    owner = np.concatenate([np.arange(n), gj])
    return (xall, owner, np.arange(n, dtype=np.int32), np.diff(off).astype(np.int32),
            flat.astype(np.int32))


def list_mode_bench(model, symbols, x, t, cell, prec, calls):
    """The Pair::compute drop-in (tmd_compute_list) on the same system with
    LAMMPS-style ghosts: host buffers in and out every call (x and types up,
    f down, PCIe included).  Timed twice: packing and uploading the whole
    list every call (what the reference does every step), and with the list
    kept on the device between neighbor builds (tmd_compute_list_serial with
    an unchanged serial: only x goes up)."""
    from tensormd_amd import NeighborList, TensorMD
    m = TensorMD(model, symbols, prec, 0.001, device=0)
    s = m.system(x, t, cell, skin=1.0)
    s.compute()
    xall, owner, ilist, numneigh, flat = ghost_list_from_system(s, s.get()["x"], cell)
    s.close()
    tall = t[owner]
    n = len(x)
    lst = NeighborList(n, ilist, numneigh, flat)
    out = m.compute_list_ptrs(lst, xall, tall, eatom=False, vatom=False)
    t0 = time.perf_counter()
    for _ in range(calls):
        m.compute_list_ptrs(lst, xall, tall, eatom=False, vatom=False, out=out)
    dt_full = time.perf_counter() - t0
    m.compute_list_ptrs(lst, xall, tall, eatom=False, vatom=False, serial=1, out=out)
    t0 = time.perf_counter()
    for _ in range(calls):
        m.compute_list_ptrs(lst, xall, tall, eatom=False, vatom=False, serial=1, out=out)
    dt_ser = time.perf_counter() - t0
    m.close()
    return dict(natoms=n, nall=int(len(xall)), list_entries=int(len(flat)), calls=calls,
                repack_every_call=dict(atom_steps_per_s=n * calls / dt_full,
                                       ms_per_call=1e3 * dt_full / calls),
                list_kept_on_device=dict(atom_steps_per_s=n * calls / dt_ser,
                                         ms_per_call=1e3 * dt_ser / calls))


def fix_path_bench(model, symbols, x, t, cell, v, steps, host_ms=0.5):
    """The step of fix nve/tensormd/amd (one tmd_md_step_async per LAMMPS
    step) against one tmd_md_run of the same steps, BF16, each from the same
    state; with host_ms of host work per step (LAMMPS' own bookkeeping,
    modelled by a sleep) overlapped by the asynchronous step, or waited for
    after each synchronous one.  Median of three."""
    from tensormd_amd import TensorMD
    m = TensorMD(model, symbols, "bf16", 0.001, device=0)
    s = m.system(x, t, cell, skin=1.0)
    s.set_velocities(v)
    s.compute()
    s.md_run(20, 0.001, history=False)
    st = s.get()

    def one_run():
        s.md_run(steps, 0.001, history=False)

    def async_steps(work):
        for _ in range(steps):
            s.step_async(0.001)
            if work:
                time.sleep(1e-3 * host_ms)

    def sync_steps(work):
        for _ in range(steps):
            s.md_run(1, 0.001, history=False)
            if work:
                time.sleep(1e-3 * host_ms)

    modes = {"one_md_run": one_run, "async_steps": lambda: async_steps(False),
             "async_steps_with_host_work": lambda: async_steps(True),
             "sync_steps_with_host_work": lambda: sync_steps(True)}
    out = {k: [] for k in modes}
    for _ in range(3):
        for k, fn in modes.items():
            s.set_positions(st["x"])
            s.set_velocities(st["v"])
            s.compute()
            t0 = time.perf_counter()
            fn()
            s.get()
            out[k].append(1e3 * (time.perf_counter() - t0) / steps)
    s.close()
    m.close()
    res = {f"{k}_ms_per_step": float(np.median(a)) for k, a in out.items()}
    res.update(steps=steps, host_work_ms=host_ms, precision="bf16")
    return res


def _force_errors(got, ref, n):
    df = got["f"] - ref["f"]
    return dict(max_abs_dF=float(abs(df).max()), rms_dF=float(np.sqrt(np.mean(df * df))),
                dE_per_atom=abs(got["energy"] - ref["energy"]) / n,
                max_abs_F=float(abs(ref["f"]).max()))


def fp8_experiment(steps, warmup, keep=None):
    """1,024,000-atom 4-element BCC alloy, MLP 256-256-128-1: FP32, BF16 and
    FP8 e4m3 MLPs side by side, with the device memory each system holds.
    keep (a dict) receives the model, the start and end positions and each
    path's forces there, for a check against the FP64 oracle."""
    from tensormd_amd import TensorMD, device_memory, workloads as W
    symbols = ["Mo", "Nb", "Ta", "W"]
    masses = [95.95, 92.906, 180.95, 183.84]
    tmp = tempfile.mkdtemp()
    model = W.write_model(os.path.join(tmp, "c4.npz"), symbols=symbols, masses=masses, rmax=6.5,
                          num_filters=24, layer_sizes=[256, 256, 128, 1], actfn="tanh", seed=4096,
                          radial_scale=0.2)
    x, t, cell = W.crystal("bcc", 3.2, (80, 80, 80), nelt=4, sigma=0.08, seed=9)
    v = W.maxwell_boltzmann(np.array(masses)[t - 1], 600.0, seed=10)
    n = len(x)
    out = dict(natoms=n, steps=steps, mlp="256-256-128-1 tanh",
               errors_note="errors_vs_fp32_*: BF16 and FP8 forces and energies against the FP32 "
                           "GPU path at the same positions: the starting state, and the FP32 run's "
                           "state after warmup + steps; errors_vs_fp64_oracle_* (added by the CPU "
                           "leg when it runs): every path against the FP64 oracle at those "
                           "positions")
    ref = {}
    for prec in ("fp32", "bf16", "fp8"):
        m = TensorMD(model, symbols, prec, 0.001, device=0)
        free0, _ = device_memory()
        s = m.system(x, t, cell, skin=1.0)
        s.set_velocities(v)
        r0 = s.compute()
        free1, _ = device_memory()
        s.md_run(warmup, 0.001, history=False)
        s.get()
        t0 = time.perf_counter()
        s.md_run(steps, 0.001, history=False)
        s.get()
        dt = time.perf_counter() - t0
        prof = s.profile(5, 0.001)
        out[prec] = dict(atom_steps_per_s=n * steps / dt, ms_per_step=1e3 * dt / steps,
                         mlp_ms=prof["mlp"], system_hbm_bytes=int(free0 - free1),
                         hbm_bytes_per_atom=(free0 - free1) / n)
        if prec == "fp32":
            x_end = s.get()["x"]
            s.set_positions(x_end)
            ref = dict(start=r0, end=s.compute(), x_end=x_end)
            r1 = ref["end"]
        else:
            s.set_positions(ref["x_end"])
            r1 = s.compute()
            out[prec]["errors_vs_fp32_start"] = _force_errors(r0, ref["start"], n)
            out[prec]["errors_vs_fp32_end"] = _force_errors(r1, ref["end"], n)
        if keep is not None:
            keep.update(model=model, x=x, t=t, cell=cell, x_end=ref["x_end"])
            keep[prec] = dict(start=r0, end=r1)
        print(f"bench: c4 {prec}: {out[prec]}", flush=True)
        s.close()
        m.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cells", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--precisions", default="fp32,bf16,bf16x3")
    ap.add_argument("--serve-requests", type=int, default=4096)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--c4-steps", type=int, default=200)
    # calls of the Pair::compute drop-in (host buffers, LAMMPS-style ghosts)
    # on the same system, per precision and list mode
    ap.add_argument("--list-calls", type=int, default=20)
    # steps of the fix nve/tensormd/amd timing (asynchronous per-step calls)
    ap.add_argument("--fix-steps", type=int, default=100)
    # every N-th timed step of the headline precision is launched kernel by
    # kernel with HIP events around the kernels (0: off; the roofline then
    # uses the kernel times of the profile steps after the run)
    ap.add_argument("--kernel-timing-every", type=int, default=10)
    # NVE total-energy drift of each precision over this many steps, from the
    # same initial state, in a separate untimed run with per-step energies
    ap.add_argument("--drift-steps", type=int, default=1000)
    # launch each MD step's kernels one by one instead of replaying the
    # captured step graph (same kernels and results; for kernel tracers)
    ap.add_argument("--no-graph", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    from tensormd_amd import TensorMD, workloads as W

    symbols = ["Mo", "Nb", "Ta", "W"]
    masses = [95.95, 92.906, 180.95, 183.84]
    sizes = [384, 256, 128, 64, 1]
    tmp = tempfile.mkdtemp()
    model = W.write_model(os.path.join(tmp, "c2.npz"), symbols=symbols, masses=masses, rmax=6.5,
                          num_filters=24, layer_sizes=sizes[1:], actfn="tanh", seed=2024,
                          radial_scale=0.2)
    n = args.cells
    x, t, cell = W.crystal("bcc", 3.2, (n, n, n), nelt=4, sigma=0.08, seed=7)
    natoms = len(x)
    v = W.maxwell_boltzmann(np.array(masses)[t - 1], 600.0, seed=8)
    print(f"bench: {natoms} atoms, steps={args.steps} warmup={args.warmup}", flush=True)

    res = {}
    prof = {}
    state = {}
    ktime = {}
    drift = {}
    kt_every = args.kernel_timing_every
    nbr_within_rc = None
    head = "bf16" if "bf16" in args.precisions.split(",") else args.precisions.split(",")[0]
    for prec in args.precisions.split(","):
        m = TensorMD(model, symbols, prec, 0.001, device=0)
        s = m.system(x, t, cell, skin=1.0)
        if args.no_graph:
            s.set_graph(False)
        s.set_velocities(v)
        s.compute()
        if nbr_within_rc is None:
            nbr_within_rc = neighbors_within(s, s.get()["x"], cell, 6.5)
        s.md_run(args.warmup, 0.001, history=False)
        s.get()  # synchronise
        if kt_every > 0:
            # kernel durations over the timed region: every kt_every-th step
            # is launched kernel by kernel with HIP events on the library
            # stream (the others replay the captured step graph)
            s.set_kernel_timing(kt_every)
        t0 = time.perf_counter()
        s.md_run(args.steps, 0.001, history=False)
        out = s.get()  # synchronises the stream
        dt = time.perf_counter() - t0
        if kt_every > 0:
            ktime[prec] = s.kernel_timing()
            s.set_kernel_timing(0)
        res[prec] = dict(time=dt, rate=natoms * args.steps / dt, energy=out["energy"],
                         builds=s.nbuilds)
        prof[prec] = s.profile(args.profile_steps, 0.001)
        if prec == head:
            # the state the kernel times above were taken in: the random-init
            # model does not hold the lattice, so the system heats and the
            # neighbor count within rc grows during the run
            st = s.get()
            state = dict(neighbors_within_rc=neighbors_within(s, st["x"], cell, 6.5),
                         temperature_K=kinetic_temperature(np.array(masses)[t - 1], st["v"]))
        if args.drift_steps > 0:
            sd = m.system(x, t, cell, skin=1.0)
            sd.set_velocities(v)
            sd.compute()
            ep, ek = sd.md_run(args.drift_steps, 0.001, history=True)
            et = (ep + ek) / natoms
            drift[prec] = dict(steps=args.drift_steps, etot_per_atom_start=float(et[0]),
                               drift_per_atom=float(et[-1] - et[0]),
                               max_dev_per_atom=float(abs(et - et[0]).max()),
                               epot_per_atom_start=float(ep[0] / natoms),
                               epot_per_atom_end=float(ep[-1] / natoms))
            sd.close()
        print(f"bench: {prec}: {res[prec]['rate'] / 1e6:.2f} M atom-steps/s, "
              f"{1e3 * dt / args.steps:.3f} ms/step, builds={s.nbuilds}, phases(ms)={prof[prec]}",
              flush=True)
        s.close()
        m.close()

    # roofline of the dominant kernel of the headline (bf16) path
    from tensormd_amd import bench_copy
    copy_gbs = bench_copy(1 << 30, 20)
    p = prof[head]
    phases = {k: p[k] for k in ["neighbor", "descriptor", "mlp", "force", "reduce", "integrate"]}
    if head in ktime:
        kt = ktime[head]
        timing = (f"HIP events on the library stream around the kernels of every {kt_every}th "
                  f"step of the timed run")
    else:
        kt = dict(descriptor=p["descriptor"], mlp=p["mlp"], force=p["force"],
                  kept_per_atom=state.get("neighbors_within_rc", nbr_within_rc),
                  samples=args.profile_steps)
        timing = "HIP events around the kernels of the profile steps after the timed run"
    # work per launch at the sampled steps: their mean kept-neighbor count
    c = kt["kept_per_atom"]
    per = per_atom_counts(c, d=20, k=24, b=4, m=3, gw=384, sizes=sizes,
                          g_bytes=2 if head == "bf16" else 4)
    dom = max(["descriptor", "mlp", "force"], key=lambda k: kt[k])

    def kernel_row(kname, prec, ms, c_kept):
        """Roofline row of one kernel: algorithmic FLOP/s against its
        arithmetic roof and algorithmic bytes/s against HBM."""
        # the BF16 path of this tanh benchmark network stores G as BF16
        pk = per_atom_counts(c_kept, d=20, k=24, b=4, m=3, gw=384, sizes=sizes,
                             g_bytes=2 if prec == "bf16" else 4)
        t_s = ms * 1e-3
        flops = pk[kname + "_flop"] * natoms / t_s / 1e12
        gbs = pk[kname + "_bytes"] * natoms / t_s / 1e9
        if kname == "mlp":
            peak = PEAK["bf16_tflops"] if prec.startswith("bf16") else PEAK["fp32_tflops"]
            unit = {"bf16": "BF16", "bf16x3": "BF16, three products per product (split operands)"}
            row = dict(kernel=f"k_mlp_fast {prec} (fused per-element MLP fwd+bwd on "
                              f"{unit.get(prec, 'FP32')} MFMA)", bound="mfma")
        else:
            # fused FP32 gather/scatter kernels without MFMA: arithmetic
            # intensity 50-80 FLOP/B, so the binding roof is FP32 vector
            # issue (VALU), not HBM
            peak = PEAK["fp32_tflops"]
            row = dict(kernel=f"k_{kname} (fused FP32 {kname} kernel, VALU)", bound="valu",
                       valu_peak_without_packed_fp32=PEAK["fp32_tflops"] / 2)
        row.update(precision=prec, achieved=flops, peak=peak, unit="TFLOP/s", frac=flops / peak,
                   hbm_achieved_gbs=gbs, hbm_peak_gbs=PEAK["hbm_gbs"], hbm_frac=gbs / PEAK["hbm_gbs"],
                   avg_launch_ms=ms, neighbors_within_rc=c_kept)
        return row

    roof = dict(kernel_row(dom, head, kt[dom], c))
    roof["kernels"] = [kernel_row(k, pr, ktime[pr][k], ktime[pr]["kept_per_atom"])
                       for pr in ktime for k in ("descriptor", "mlp", "force")]
    roof["avg_launch_ms"] = kt[dom]
    roof["avg_launch_samples"] = kt["samples"]
    roof["timing"] = timing
    roof["kernel_ms_timed_run"] = {k: kt[k] for k in ("descriptor", "mlp", "force")}
    # the committed rocprofv3 kernel trace of this system covers a 300-step run;
    # kernel durations grow with the neighbor count as the system heats, so
    # they agree with avg_launch_ms at equal step counts
    roof["kernel_trace"] = dict(summary="profiles/md_step_128k_bcc4_kernel_stats.md", run_steps=300,
                                this_run_steps=args.steps)
    roof["algorithmic_per_atom"] = {k: v for k, v in per.items() if k.startswith(dom)}
    roof["traffic"] = None
    tr = _traffic(dom)
    if tr is not None:
        roof["traffic"] = tr["bytes_per_launch"]
        roof["traffic_source"] = tr["source"]
    roof["measured_copy_gbs"] = copy_gbs
    roof["phases_ms_end_state"] = phases
    roof["neighbors_within_rc"] = c
    roof["neighbors_within_rc_start"] = nbr_within_rc
    roof["neighbors_within_rc_end"] = state.get("neighbors_within_rc")
    roof["temperature_K_end"] = state.get("temperature_K")

    listmode = None
    if args.list_calls > 0:
        listmode = {prec: list_mode_bench(model, symbols, x, t, cell, prec, args.list_calls)
                    for prec in ("fp32", "bf16")}
        print(f"bench: list mode {json.dumps(listmode)}", flush=True)
    fixpath = None
    if args.fix_steps > 0:
        fixpath = fix_path_bench(model, symbols, x, t, cell, v, args.fix_steps)
        print(f"bench: lammps fix path {json.dumps(fixpath)}", flush=True)

    fp8 = None
    c4_keep = {}
    if args.c4_steps > 0:
        fp8 = fp8_experiment(args.c4_steps, 20, c4_keep)
        print("bench: c4 " + json.dumps({k: dict(ms_per_step=v["ms_per_step"], mlp_ms=v["mlp_ms"])
                                         for k, v in fp8.items() if isinstance(v, dict) and "mlp_ms" in v}),
              flush=True)

    serving = None
    if args.serve_requests > 0:
        serving = serve_bench(model, symbols, args.serve_requests)
        print(f"bench: serving {json.dumps(serving)}", flush=True)

    cpu = None
    if not args.no_cpu:
        # threads pinned to consecutive cores (set before the oracle's OpenMP
        # runtime starts)
        os.environ.setdefault("OMP_PROC_BIND", "close")
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu_model = "unknown"
        try:
            with open("/proc/cpuinfo") as fh:
                cpu_model = next(l.split(":", 1)[1].strip() for l in fh if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        om = oracle.OracleModel(model, 0.001)
        oracle.set_threads(threads)
        nsteps = args.cpu_steps
        t0 = time.perf_counter()
        om.nve(x, v, t - 1, cell, 0.001, nsteps)
        dtc = time.perf_counter() - t0
        # nve evaluates forces nsteps + 1 times (initial forces + one per step)
        cpu = dict(value=natoms * (nsteps + 1) / dtc, unit="atom-steps/s",
                   cores=oracle.num_threads(), kind="port",
                   sample=f"FP64 C++/OpenMP restatement of the TensorMD semantics (not the "
                          f"reference binary) running velocity-Verlet on the same {natoms}-atom "
                          f"system: {nsteps} steps = {nsteps + 1} force evaluations with a "
                          f"neighbor build each, {dtc:.1f} s, OMP threads={threads} "
                          f"(OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}), CPU: {cpu_model}")
        # the single-element FCC C1 system (2048 atoms), 1 thread and all threads
        c1 = W.write_model(os.path.join(tmp, "c1.npz"), symbols=["Cu"], masses=[63.546], rmax=6.0,
                           num_filters=16, layer_sizes=[128, 64, 1], actfn="tanh", seed=11,
                           radial_scale=0.2)
        x1, t1, cell1 = W.crystal("fcc", 3.615, (8, 8, 8), sigma=0.05, seed=3)
        v1 = W.maxwell_boltzmann(np.full(len(x1), 63.546), 300.0, seed=4)
        o1 = oracle.OracleModel(c1, 0.001)
        c1_rates = {}
        for nth in (1, threads):
            oracle.set_threads(nth)
            t0 = time.perf_counter()
            o1.nve(x1, v1, t1 - 1, cell1, 0.001, args.cpu_steps)
            c1_rates[f"threads_{nth}"] = len(x1) * (args.cpu_steps + 1) / (time.perf_counter() - t0)
        oracle.set_threads(threads)
        cpu["c1_2048_atoms_atom_steps_per_s"] = c1_rates
        print(f"bench: cpu baseline {cpu}", flush=True)
        # the C4 experiment's paths against the FP64 oracle (the checker) at
        # the start and end positions: one single-point evaluation each
        if fp8 is not None and "model" in c4_keep:
            o4 = oracle.OracleModel(c4_keep["model"], 0.001)
            t0 = time.perf_counter()
            refs = dict(start=o4.compute_periodic(c4_keep["x"], c4_keep["t"] - 1, c4_keep["cell"]),
                        end=o4.compute_periodic(c4_keep["x_end"], c4_keep["t"] - 1, c4_keep["cell"]))
            n4 = len(c4_keep["x"])
            for prec in ("fp32", "bf16", "fp8"):
                if prec in c4_keep and prec in fp8:
                    for at in ("start", "end"):
                        fp8[prec][f"errors_vs_fp64_oracle_{at}"] = _force_errors(c4_keep[prec][at], refs[at], n4)
            fp8["oracle_check_s"] = time.perf_counter() - t0
            print("bench: c4 vs FP64 oracle " + json.dumps(
                {p: {k: v for k, v in fp8[p].items() if k.startswith("errors_vs_fp64")}
                 for p in ("fp32", "bf16", "fp8") if p in fp8}), flush=True)

    if "bf16" not in res or "fp32" not in res:
        print(json.dumps({"partial": True, "rates": {k: v["rate"] for k, v in res.items()},
                          "roofline": roof, "cpu_baseline": cpu, "serving": serving,
                          "fp8_experiment": fp8, "list_mode": listmode, "lammps_fix_path": fixpath,
                          "nve_drift": drift or None}),
              flush=True)
        return
    bf, fp = res["bf16"], res["fp32"]
    line = {
        "metric": "atom-steps/s at 128k atoms BF16 vs FP32 GPU; p99 latency of batched E/F requests",
        "value": bf["rate"],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * bf["time"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded jittered BCC alloy, random-init model)",
        "config": {"workload": f"{natoms}-atom 4-element BCC NVE, 256-128-64-1 tanh MLP, "
                               f"rc 6.5, k 24, BF16 MLP / FP32 descriptors",
                   "natoms": natoms, "dt_fs": 1.0, "skin": 1.0},
        "fp32_atom_steps_per_s": fp["rate"],
        "fp32_ms_per_step": 1e3 * fp["time"] / args.steps,
        "bf16_over_fp32": bf["rate"] / fp["rate"],
        "bf16x3_atom_steps_per_s": res["bf16x3"]["rate"] if "bf16x3" in res else None,
        "bf16x3_ms_per_step": 1e3 * res["bf16x3"]["time"] / args.steps if "bf16x3" in res else None,
        "bf16x3_over_fp32": res["bf16x3"]["rate"] / fp["rate"] if "bf16x3" in res else None,
        "roofline": roof,
        "cpu_baseline": cpu,
        "serving": serving,
        "fp8_experiment": fp8,
        "list_mode": listmode,
        "lammps_fix_path": fixpath,
        "nve_drift": drift or None,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
