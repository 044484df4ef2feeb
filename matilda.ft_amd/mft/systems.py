"""Synthetic TILD and CL-FTS workloads with the shapes of the benchmark configs.

Particle topology follows PS_Box::makeLinear (PS_Box_moleculeMakers.cu:15-311):
beads are numbered molecule-major then block-major; bead i bonds first to i-1
(unless it starts a chain) then to i+1 (unless it ends one), with bond type
blockBondType-1 of its own block.  Coordinates are a synthetic input: chain
heads uniform in the box, then a random walk with N(0, 1/3) steps per axis,
wrapped into [0, L).
"""
from __future__ import annotations

import numpy as np


def make_linear(nmolecs, blocks, block_charge=None, block_bond_type=None, maxbonds=2, start=0):
    """blocks = [(N_b, type_b)]; returns type, charge, nBonds, bondedTo, bondType,
    mID arrays for nmolecs new molecules whose first bead has index `start`."""
    Ns = np.array([n for n, _ in blocks], np.int64)
    Ntot = int(Ns.sum())
    nb = len(blocks)
    if block_bond_type is None:
        block_bond_type = [1] * nb
    block_of_bead = np.repeat(np.arange(nb), Ns)
    pos_in_chain = np.arange(Ntot)
    ns = nmolecs * Ntot
    bead = np.tile(pos_in_chain, nmolecs)
    blk = np.tile(block_of_bead, nmolecs)
    gidx = start + np.arange(ns, dtype=np.int64)
    types = np.array([t for _, t in blocks], np.int32)[blk]
    charge = np.zeros(ns, np.float32)
    if block_charge is not None:
        charge = np.asarray(block_charge, np.float32)[blk]
    btype = np.asarray(block_bond_type, np.int32)[blk] - 1
    has_prev = bead > 0
    has_next = bead < Ntot - 1
    nBonds = has_prev.astype(np.int32) + has_next.astype(np.int32)
    bondedTo = np.zeros((ns, maxbonds), np.int32)
    bondType = np.zeros((ns, maxbonds), np.int32)
    # slot 0 holds the previous bead when there is one, else the next bead
    bondedTo[:, 0] = np.where(has_prev, gidx - 1, np.where(has_next, gidx + 1, 0))
    bondType[:, 0] = np.where(has_prev | has_next, btype, 0)
    both = has_prev & has_next
    bondedTo[:, 1] = np.where(both, gidx + 1, 0)
    bondType[:, 1] = np.where(both, btype, 0)
    mID = np.repeat(np.arange(nmolecs, dtype=np.int32), Ntot)
    return dict(type=types, charge=charge, nBonds=nBonds, bondedTo=bondedTo.reshape(-1),
                bondType=bondType.reshape(-1), mID=mID)


def concat_topologies(parts):
    out = {}
    for k in ("type", "charge", "nBonds", "bondedTo", "bondType"):
        out[k] = np.concatenate([p[k] for p in parts])
    return out


def random_walk(nmolecs, Ntot, L, rng):
    L = np.asarray(L, np.float64)
    heads = rng.random((nmolecs, 1, 3)) * L
    steps = rng.normal(0.0, np.sqrt(1.0 / 3.0), (nmolecs, Ntot, 3))
    steps[:, 0, :] = 0.0
    x = heads + np.cumsum(steps, axis=1)
    x = np.mod(x, L)
    x = x.astype(np.float32).reshape(-1, 3)
    # float rounding of mod can land exactly on L; the reference keeps [0, L]
    return x


def gaussian_sigma(a):
    """sigma of the gaussian potential for per-particle smearing a (sigma^2 = 2a^2)."""
    return float(np.sqrt(2.0) * a)


def tild_diblock(nchains=2000, NA=10, NB=10, box=20.0, grid=32, pmeorder=3, kappa=50.0, chi=1.0, a=0.5,
                 bond_k=1.5, delt=0.005, seed=1234):
    """Plain TILD AB diblock melt (config 1 shape by default)."""
    rng = np.random.default_rng(seed)
    topo = make_linear(nchains, [(NA, 0), (NB, 1)])
    N = topo["type"].size
    L = (box, box, box)
    rho0 = N / box ** 3
    sig = gaussian_sigma(a)
    pots = [
        dict(kind="gaussian", maskI=1, maskJ=1, A=kappa / (2 * rho0), sigma=sig),
        dict(kind="gaussian", maskI=2, maskJ=2, A=kappa / (2 * rho0), sigma=sig),
        dict(kind="gaussian", maskI=1, maskJ=2, A=(kappa + chi) / rho0, sigma=sig),
    ]
    x = random_walk(nchains, NA + NB, L, rng).reshape(-1)
    return dict(name="tild_diblock", Nx=(grid, grid, grid), L=L, pmeorder=pmeorder, species=["A", "B"],
                topo=topo, x=x, potentials=pots, bond_k=[bond_k], bond_r0=[0.0], bond_style=[0],
                delt=delt, mass=[1.0, 1.0], mobility=[1.0, 1.0], has_charges=False, seed=seed)


def tild_coacervate(nchains=40000, Nchain=20, nsolvent=200000, box=60.0, grid=128, pmeorder=3, kappa=50.0,
                    chi_sp=0.5, lB=2.0, charge_smear=0.5, a=0.5, bond_k=1.5, delt=0.005, seed=1234):
    """Large TILD polyelectrolyte coacervate (config 2 shape by default):
    nchains polycations + nchains polyanions (charge +-1 per bead) + solvent."""
    rng = np.random.default_rng(seed)
    t_pos = make_linear(nchains, [(Nchain, 0)], block_charge=[1.0])
    t_neg = make_linear(nchains, [(Nchain, 1)], block_charge=[-1.0], start=nchains * Nchain)
    t_sol = make_linear(nsolvent, [(1, 2)], start=2 * nchains * Nchain)
    topo = concat_topologies([t_pos, t_neg, t_sol])
    N = topo["type"].size
    L = (box, box, box)
    rho0 = N / box ** 3
    sig = gaussian_sigma(a)
    A_self = kappa / (2 * rho0)
    pots = [
        dict(kind="gaussian", maskI=1, maskJ=1, A=A_self, sigma=sig),
        dict(kind="gaussian", maskI=2, maskJ=2, A=A_self, sigma=sig),
        dict(kind="gaussian", maskI=4, maskJ=4, A=A_self, sigma=sig),
        dict(kind="gaussian", maskI=1, maskJ=2, A=kappa / rho0, sigma=sig),
        dict(kind="gaussian", maskI=1, maskJ=4, A=(kappa + chi_sp) / rho0, sigma=sig),
        dict(kind="gaussian", maskI=2, maskJ=4, A=(kappa + chi_sp) / rho0, sigma=sig),
        dict(kind="charges", A=lB, sigma=charge_smear),
    ]
    xp = random_walk(nchains, Nchain, L, rng)
    xn = random_walk(nchains, Nchain, L, rng)
    xs = (rng.random((nsolvent, 3)) * np.asarray(L)).astype(np.float32)
    x = np.concatenate([xp, xn, xs]).reshape(-1)
    return dict(name="tild_coacervate", Nx=(grid, grid, grid), L=L, pmeorder=pmeorder, species=["P", "N", "S"],
                topo=topo, x=x, potentials=pots, bond_k=[bond_k], bond_r0=[0.0], bond_style=[0],
                delt=delt, mass=[1.0] * 3, mobility=[1.0] * 3, has_charges=True, seed=seed)


def add_chain_angles(cfg, chain_len, bead_type, k, theta, style, maxangles=3):
    """Angles on every interior bead of the linear chains of cfg (chains of
    chain_len consecutive ids), the layout of the linear molecule maker
    (PS_Box_moleculeMakers.cu:242-288): each angle (i-1, i, i+1) is entered on
    the lists of its three members.  bead_type(i) gives the 0-based angle type
    of the vertex bead, or -1 for none.  k, theta (radians), style per type."""
    ns = cfg["topo"]["type"].size
    nA = np.zeros(ns, np.int32)
    aT = np.zeros((ns, maxangles), np.int32)
    aG = np.zeros((ns, maxangles, 3), np.int32)
    for i in range(ns):
        pos = i % chain_len
        t = bead_type(i)
        if t < 0 or pos == 0 or pos == chain_len - 1:
            continue
        for p in (i - 1, i, i + 1):
            aG[p, nA[p]] = (i - 1, i, i + 1)
            aT[p, nA[p]] = t
            nA[p] += 1
    cfg["angles"] = dict(nAngles=nA, angleType=aT, angleGroup=aG, maxangles=maxangles,
                         k=list(k), theta=list(theta), style=list(style))
    return cfg


def tild_engine(cfg, integrator="GJF"):
    """Build a TildEngine for a config dict from this module."""
    from . import TildEngine
    topo = cfg["topo"]
    ang = cfg.get("angles")
    akw = {}
    if ang is not None:
        akw = dict(angle_k=ang["k"], angle_theta=ang["theta"], angle_style=ang["style"], maxangles=ang["maxangles"])
    eng = TildEngine(cfg["Nx"], cfg["L"], cfg["pmeorder"], len(cfg["species"]), topo["type"].size,
                     cfg["potentials"], cfg["bond_k"], cfg["bond_r0"], cfg["bond_style"], integrator=integrator,
                     delt=cfg["delt"], mass=cfg["mass"], mobility=cfg["mobility"], seed=cfg["seed"],
                     has_charges=cfg["has_charges"], **akw)
    eng.set_topology(topo["type"], topo["charge"] if cfg["has_charges"] else None, topo["nBonds"],
                     topo["bondedTo"], topo["bondType"])
    if ang is not None:
        eng.set_angles(ang["nAngles"], ang["angleType"], ang["angleGroup"])
    eng.set_state(cfg["x"])
    return eng


def fts_diblock(grid=64, L=8.0, Nr=100.0, NA=50, NB=50, chiN=20.0, kappaN=100.0, C=50.0, dt=0.01, sigma=0.1,
                seed=1234, cl=True, grid_shape=None):
    """Plain CL-FTS AB diblock melt (config 3 shape by default): Helfand then
    Flory potentials, real parts of w+ (Helfand), w+ and w- (Flory) from
    N(0, sigma^2), imaginary parts zero."""
    Nx = tuple(grid_shape) if grid_shape is not None else (grid, grid, grid)
    Ls = tuple(L) if np.ndim(L) else (L, L, L)
    spec = dict(Nx=Nx, L=Ls, Nr=Nr, C=C, nspecies=2, cl=cl, seed=seed,
                potentials=[dict(kind="Helfand", kappaN=kappaN, dt_pl=dt),
                            dict(kind="Flory", I=0, J=1, chiN=chiN, dt_pl=dt, dt_mi=dt)],
                molecules=[dict(blocks=[(NA, 0), (NB, 1)], phi=1.0)])
    rng = np.random.default_rng(seed)
    M = int(np.prod(Nx))
    fields = {
        (0, "pl"): (rng.normal(0.0, sigma, M)).astype(np.complex64),
        (1, "pl"): (rng.normal(0.0, sigma, M)).astype(np.complex64),
        (1, "mi"): (rng.normal(0.0, sigma, M)).astype(np.complex64),
    }
    return spec, fields


def cl_step(spacing, dt_max=0.001, h_ref=0.125):
    """Explicit Euler-Maruyama CL step for a grid spacing h (in Rg):
    dt_max (h / h_ref)^3, at most dt_max.  At C = 50 the explicit step runs
    away (in the FP64 oracle as in the engine) once dt exceeds what the
    finest grid modes tolerate: the 64^3, 8 Rg diblock at dt = 0.01 does so
    within 200 iterations and stays finite over 2000 at dt = 0.001; the
    noise variance per site, 2 dt / gvol, grows as h^-3, so the step shrinks
    with h^3 below h_ref."""
    return min(dt_max, dt_max * (spacing / h_ref) ** 3)


def fts_sweep_replicas(n=256, seed=7, iterations=1000, dt=None):
    """Batched sweep (config 4 shape): grids from {24,32,48,64}^3, box
    4-10 Rg, chiN in [10,20], f_A in [0.3,0.5], Ns in {50,100}, kappaN = 100,
    C = 50; one seed per replica; CL step dt, or cl_step(L / grid) per
    replica when dt is None."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(n):
        g = int(rng.choice([24, 32, 48, 64]))
        Lr = float(rng.uniform(4.0, 10.0))
        chiN = float(rng.uniform(10.0, 20.0))
        fA = float(rng.uniform(0.3, 0.5))
        Ns = int(rng.choice([50, 100]))
        NA = int(round(fA * Ns))
        dtr = cl_step(Lr / g) if dt is None else dt
        spec, fields = fts_diblock(grid=g, L=Lr, Nr=float(Ns), NA=NA, NB=Ns - NA, chiN=chiN, dt=dtr,
                                   seed=seed * 100003 + r)
        out.append((spec, fields))
    return out


def fts_system(spec, fields):
    """Build and initialise an FtsSystem."""
    from . import FtsSystem
    s = FtsSystem(spec)
    load_fields(s.set_fields, spec, fields)
    s.init()
    return s


def load_fields(setter, spec, fields, rep=None):
    for p, pot in enumerate(spec["potentials"]):
        wpl = fields.get((p, "pl"))
        wmi = fields.get((p, "mi"))
        if wpl is None and wmi is None:
            continue
        if rep is None:
            setter(p, wpl, wmi)
        else:
            setter(rep, p, wpl, wmi)
