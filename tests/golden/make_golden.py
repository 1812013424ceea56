"""Extract the reference's golden numbers into tests/golden/ (run once, here).

Reads only data from the reference tree: LAMMPS logs (step-0 thermo),
structure data files, the force dump force.yaml, and copies two small model
files used as inputs.  Writes reference_goldens.json with every value and
the file:line it came from.  The GPU box never runs this.

    python tests/golden/make_golden.py [/root/reference]
"""
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
EX = os.path.join(REF, "examples", "tensormd")
OUT = os.path.dirname(os.path.abspath(__file__))


def thermo_step0(log):
    """Return (header->value dict, line number) of the step-0 thermo row."""
    lines = open(log).read().splitlines()
    for n, l in enumerate(lines):
        t = l.split()
        if t and t[0] == "Step":
            vals = lines[n + 1].split()
            return dict(zip(t, vals)), n + 2
    raise RuntimeError(log)


def read_data(path):
    lo = {}
    tilt = (0.0, 0.0, 0.0)
    atoms = []
    mode = None
    for l in open(path).read().splitlines():
        t = l.split()
        if len(t) == 4 and t[2].endswith("lo"):
            lo[t[2][0]] = (float(t[0]), float(t[1]))
        elif len(t) == 6 and t[3] == "xy":
            tilt = tuple(map(float, t[:3]))
        elif l.startswith("Atoms"):
            mode = "atoms"
        elif mode == "atoms" and len(t) == 5:
            atoms.append((int(t[0]), int(t[1]), float(t[2]), float(t[3]), float(t[4])))
    atoms.sort()
    lx = lo["x"][1] - lo["x"][0]
    ly = lo["y"][1] - lo["y"][0]
    lz = lo["z"][1] - lo["z"][0]
    xy, xz, yz = tilt
    cell = [[lx, 0, 0], [xy, ly, 0], [xz, yz, lz]]
    return dict(cell=cell, types=[a[1] for a in atoms], x=[list(a[2:]) for a in atoms])


def main():
    g = {"source": "reference examples/tensormd logs, data files and force.yaml", "cases": {}}
    rel = lambda p: os.path.relpath(p, REF)

    # W_k32 perfect bcc, a=3.189, 24x24x20 (in.lammps of W/tensormd/speedup)
    log = os.path.join(EX, "W/tensormd/speedup/a100_log/log_0114.lammps")
    th, ln = thermo_step0(log)
    log2 = os.path.join(EX, "scaling/strong/200/log200.lammps")
    th2, ln2 = thermo_step0(log2)
    g["cases"]["W_k32"] = dict(
        model="W_k32_snapshot.npz", interp=0.001, lattice="bcc", a=3.189, reps=[24, 24, 20],
        natoms=23040, temp=float(th["Temp"]), pe=float(th["PotEng"]), press=float(th["Press"]),
        pe_digits=7, src=f"{rel(log)}:{ln}",
        big=dict(natoms=400 * 500 * 500 * 2, pe=float(th2["PotEng"]), pe_digits=8,
                 src=f"{rel(log2)}:{ln2}"))
    # W_m3_k64, same lattice 24x24x40
    log = os.path.join(EX, "W/tensormd/compare/a100_log/log2.lammps")
    th, ln = thermo_step0(log)
    g["cases"]["W_m3_k64"] = dict(
        model="W_m3_k64_n128_squareplus.npz", interp=0.001, lattice="bcc", a=3.189,
        reps=[24, 24, 40], natoms=46080, temp=float(th["Temp"]), pe=float(th["PotEng"]),
        press=float(th["Press"]), pe_digits=8, src=f"{rel(log)}:{ln}")
    # MoNi: 16-atom triclinic cell x10x10x10
    data = read_data(os.path.join(EX, "binary/Mo8_Ni8_id3210.data"))
    for name, logp in [("MoNi_squareplus_500k", "binary/medium/a100_log/log_release.lammps"),
                       ("MoNi_m3", "binary/medium/orise_log/log_release.lammps")]:
        log = os.path.join(EX, logp)
        th, ln = thermo_step0(log)
        g["cases"][name] = dict(
            model=name + ".npz", interp=0.0001, symbols=["Mo", "Ni"], structure="Mo8_Ni8",
            reps=[10, 10, 10], natoms=16000, temp=float(th["Temp"]), pe=float(th["PotEng"]),
            press=float(th["Press"]), pe_digits=8, src=f"{rel(log)}:{ln}")
    g["structures"] = {"Mo8_Ni8": dict(data, src=rel(os.path.join(EX, "binary/Mo8_Ni8_id3210.data")))}
    # force.yaml: 128 atoms = Mo8_Ni8 x2x2x2, forces at timestep 0
    fy = os.path.join(EX, "binary/medium/force.yaml")
    forces = []
    for l in open(fy):
        m = re.match(r"\s*- \[\s*(\d+)\s*,\s*(\d+)\s*,\s*([-\d.]+)\s*,\s*([-\d.]+)\s*,\s*([-\d.]+)", l)
        if m:
            forces.append([int(m.group(1)), int(m.group(2))] + [float(m.group(i)) for i in (3, 4, 5)])
    g["forces_MoNi"] = dict(src=rel(fy), model="MoNi_m3.npz", decimals=8,
                            note="atoms 1-16 of the 2x2x2 replica; forces repeat per replica",
                            rows=forces[:16], replica_max_dev=max(
                                abs(forces[i][k] - forces[i % 16][k]) for i in range(len(forces))
                                for k in (2, 3, 4)))
    # AlMgSi: 108-atom cell x10x10x3
    data = read_data(os.path.join(EX, "trinary/AlMgSi.data"))
    g["structures"]["AlMgSi"] = dict(data, src=rel(os.path.join(EX, "trinary/AlMgSi.data")))
    log = os.path.join(EX, "trinary/a100_log/log_release.lammps")
    th, ln = thermo_step0(log)
    g["cases"]["AlMgSi"] = dict(
        model="AlMgSi.24aug13.npz", interp=0.0001, symbols=["Al", "Mg", "Si"], structure="AlMgSi",
        reps=[10, 10, 3], natoms=32400, temp=float(th["Temp"]), pe=float(th["PotEng"]),
        press=float(th["Press"]), pe_digits=8, src=f"{rel(log)}:{ln}")
    with open(os.path.join(OUT, "reference_goldens.json"), "w") as f:
        json.dump(g, f, indent=1)
    # small model files used as test inputs (data, not source)
    for p in ["W/tensormd/W_k32_snapshot.npz", "binary/medium/MoNi_m3.npz"]:
        shutil.copy(os.path.join(EX, p), os.path.join(OUT, os.path.basename(p)))
    print("wrote", os.path.join(OUT, "reference_goldens.json"))


if __name__ == "__main__":
    main()
