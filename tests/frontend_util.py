"""Helpers for the drop-in front end tests: the `mft` executable, reference
syntax input writers and LAMMPS-style data files."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MFT_BIN = os.path.join(ROOT, "matilda.ft_amd", "bin", "mft")
INPUTS = os.path.join(ROOT, "tests", "inputs")


def run_mft(args, cwd, timeout=300):
    return subprocess.run([MFT_BIN] + list(args), cwd=cwd, capture_output=True, text=True, timeout=timeout)


def write_data_file(path, cfg):
    """Data file in the layout PS_Box::readDataConfig reads
    (PS_Box_initialization.cu:787-990): atoms id mol type [q] x y z, bonds
    listed once per pair in chain order."""
    topo = cfg["topo"]
    x = np.asarray(cfg["x"], np.float32).reshape(-1, 3)
    n = x.shape[0]
    nb, bt = topo["nBonds"], topo["bondedTo"].reshape(n, -1)
    btype = topo["bondType"].reshape(n, -1)
    bonds = [(int(btype[i, k]) + 1, i + 1, int(bt[i, k]) + 1) for i in range(n) for k in range(nb[i]) if bt[i, k] > i]
    charged = cfg["has_charges"]
    angles = []
    ang = cfg.get("angles")
    if ang is not None:
        na, ag, at = ang["nAngles"], ang["angleGroup"], ang["angleType"]
        angles = [(int(at[i, a]) + 1, *(int(v) + 1 for v in ag[i, a])) for i in range(n) for a in range(na[i])
                  if ag[i, a, 1] == i]
    nat = len(ang["k"]) if ang is not None else 0
    with open(path, "w") as f:
        f.write("data file\n\n%d atoms\n%d bonds\n%d angles\n\n" % (n, len(bonds), len(angles)))
        f.write("%d atom types\n%d bond types\n%d angle types\n\n" % (len(cfg["species"]), len(cfg["bond_k"]), nat))
        for a, l in zip("xyz", cfg["L"]):
            f.write("0 %r %slo %shi\n" % (float(l), a, a))
        f.write("\nMasses\n\n")
        for s in range(len(cfg["species"])):
            f.write("%d %r\n" % (s + 1, float(cfg["mass"][s])))
        f.write("\nAtoms\n\n")
        mol = topo.get("mID", np.zeros(n, np.int32))
        for i in range(n):
            q = (" %r" % float(topo["charge"][i])) if charged else ""
            f.write("%d %d %d%s %.9g %.9g %.9g\n" % (i + 1, int(mol[i]) + 1, int(topo["type"][i]) + 1, q,
                                                     x[i, 0], x[i, 1], x[i, 2]))
        f.write("\nBonds\n\n")
        for k, (t, i, j) in enumerate(bonds):
            f.write("%d %d %d %d\n" % (k + 1, t, i, j))
        if angles:
            f.write("\nAngles\n\n")
            for k, a in enumerate(angles):
                f.write("%d %d %d %d %d\n" % ((k + 1,) + a))


def ps_input(cfg, data_name, steps, log_freq, integrator="GJF"):
    lines = ["box ps", "Dim 3", "boxLengths %r %r %r" % tuple(float(l) for l in cfg["L"]),
             "grid %d %d %d" % tuple(cfg["Nx"]), "pmeorder %d" % cfg["pmeorder"], "randSeed %d" % cfg["seed"]]
    if cfg["has_charges"]:
        lines.append("doCharges")
    lines += ["species %s" % s for s in cfg["species"]]
    for k, (kk, r0, st) in enumerate(zip(cfg["bond_k"], cfg["bond_r0"], cfg["bond_style"])):
        lines.append("bond %d %s %r %r" % (k + 1, ["harmonic", "fene"][st], float(kk), float(r0)))
    ang = cfg.get("angles")
    if ang is not None:
        lines.append("MAXANGLES %d" % ang["maxangles"])
        for t, (k, th, st) in enumerate(zip(ang["k"], ang["theta"], ang["style"])):
            deg = " %r" % round(float(th) * 180.0 / float(np.float32(3.141592654)), 4) if st == 1 else ""
            lines.append("angle %d %s %r%s" % (t + 1, ["wlc", "harmonic"][st], float(k), deg))
    lines.append("readData %s" % data_name)
    names = cfg["species"]
    for p in cfg["potentials"]:
        if p["kind"] == "charges":
            lines.append("potential charges %r %r" % (float(p["A"]), float(p["sigma"])))
            continue
        gi = names[int(np.log2(p["maskI"]))]
        if p["kind"] == "langevin":
            lines.append("potential Langevin %s drag %r delt %r" % (gi, float(p["A"]), float(p["sigma"])))
            continue
        if p["kind"] == "dpd":
            lines.append("neighbor_list %s %r" % (gi, float(p["Rp"])))
            lines.append("potential dpd %s %r %r delt %r" % (gi, float(p["A"]), float(p["sigma"]),
                                                             float(p["wall_pos"])))
            continue
        gj = names[int(np.log2(p["maskJ"]))]
        ramp = (" ramp %r" % float(p["A_final"])) if p.get("ramp") else ""
        lines.append("potential %s %s %s %r %r%s" % (p["kind"], gi, gj, float(p["A"]), float(p["sigma"]), ramp))
    lines += ["integrator all %s delt %r" % (integrator, float(cfg["delt"])), "logFreq %d" % log_freq, "endBox",
              "run nvt %d" % steps]
    return "\n".join(lines) + "\n"


def read_data_positions(path, n):
    with open(path) as f:
        lines = f.read().split("\n")
    i = next(k for k, l in enumerate(lines) if l.strip() == "Atoms") + 2
    x = np.zeros((n, 3), np.float32)
    for l in lines[i:i + n]:
        w = l.split()
        x[int(w[0]) - 1] = [float(v) for v in w[-3:]]
    return x.reshape(-1)


def read_log(path):
    with open(path) as f:
        return [[float(v) for v in l.split()] for l in f if l.strip()]
