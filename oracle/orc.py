"""CPU oracle for the MATILDA.FT hot paths (TEST INFRASTRUCTURE ONLY).

Only the test suite, the smoke check and the benchmark's CPU-baseline leg may
import this module; it is the checker, never the thing measured or shipped.

The C primitives in this folder restate the reference kernels; this module
binds them with ctypes and restates the reference's per-step orchestration:

* TILD step: PS_Box::doTimeStep (PS_Box.cu:73-207) and PS_Box::forces
  (PS_Box.cu:240-278): grid weights -> per-group densities -> zero forces ->
  bonds -> every potential's CalcForces into group grid forces -> mapForces per
  group -> integrator.
* CL-FTS iteration: FTS_Box::doTimeStep (FTS_Box.cu:19-123): every potential's
  updateFields (from the previous densities) -> species fields
  (fts_species.cu:53-132) -> molecule densities (fts_molecule_linear.cu:278-363).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_i = C.c_int
_d = C.c_double
_p = C.c_void_p


_OPENMP = {"on": False}


def use_openmp(on: bool, threads: int = 1):
    """Switch lib() to the OpenMP builds (the benchmark's CPU-baseline leg) with
    the given thread count, or back to the single-thread builds the tests use."""
    _OPENMP["on"] = bool(on)
    if on:
        for bits in (64, 32):
            lib(bits).orc_set_threads(int(threads))


def lib(bits: int = 64):
    """Load liboracle64.so (FP64 oracle) or liboracle32.so (FP32 host build),
    or their OpenMP builds after use_openmp(True)."""
    key = (bits, _OPENMP["on"])
    if key in _LIBS:
        return _LIBS[key]
    # ORC_SANITIZE=_asan selects the AddressSanitizer/UBSan build (the CPU
    # sanitizer run loads it with libasan preloaded)
    san = os.environ.get("ORC_SANITIZE", "")
    path = os.path.join(_HERE, f"liboracle{bits}{'_omp' if _OPENMP['on'] else ''}{san}.so")
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: build the oracle (make -C oracle)")
    L = C.CDLL(path)
    sig = {
        "orc_philox4x32_10": (None, [_p, _p, _p]),
        "orc_philox_normals": (None, [_p, C.c_long, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
        "orc_philox_normal4": (None, [_p, _p, _p]),
        "orc_fft1d": (None, [_p, _i, _i, _i]),
        "orc_fft3d": (None, [_p, _i, _i, _i, _i]),
        "orc_box_fft": (None, [_p, _p, _i, _i]),
        "orc_kvec_f": (C.c_float, [_i, _p, _p, _i, _p]),
        "orc_kvec_d": (C.c_double, [_i, _p, _p, _i, _p]),
        "orc_spline_weights": (None, [_d if bits == 64 else C.c_float, _i, _p]),
        "orc_grid_weights": (None, [_p, _i, _p, _p, _i, _i, _p, _p]),
        "orc_fill_density": (None, [_p, _p, _p, _p, _i, _p, _p, _i]),
        "orc_uk_table": (None, [_i, _d, _d, _d, _p, _p, _i, _p]),
        "orc_potential_grid_force": (None, [_p, _p, _p, _p, _i, _d, _p]),
        "orc_potential_energy": (C.c_double, [_p, _p, _p, _p, _i, _d]),
        "orc_map_forces": (None, [_p, _p, _p, _p, _i, _p, _p, _i, _d, _i]),
        "orc_bonds": (None, [_p] * 10 + [_i, _i, _i]),
        "orc_bond_energy": (None, [_p] * 10 + [_i, _i, _i]),
        "orc_bond_virial": (None, [_p] * 10 + [_i, _i, _i]),
        "orc_gjf_coeffs": (None, [_d, _d, _d, _p]),
        "orc_gjf_step": (None, [_p, _p, _p, _p, _p, _p, _p, _p, _i, _i, _p]),
        "orc_em_step": (None, [_p, _p, _d, _p, _p, _i, _i, _p]),
        "orc_fts_bond_fft": (None, [_p, _p, _i, _d, _p]),
        "orc_fts_flory_force": (None, [_p] * 6 + [_d, _d, _d, _i]),
        "orc_fts_helfand_force": (None, [_p, _p, _p, _d, _d, _d, _i]),
        "orc_fts_incompress_force": (None, [_p, _p, _p, _d, _d, _i]),
        "orc_fts_edwards_force": (None, [_p, _p, _p, _d, _d, _i]),
        "orc_fts_update_em": (None, [_p, _p, _p, _d, _i]),
        "orc_fts_update_empc": (None, [_p, _p, _p, _p, _p, _d, _i]),
        "orc_fts_update_1s": (None, [_p, _p, _p, _d, _i]),
    }
    sig["orc_set_threads"] = (None, [_i])
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIBS[key] = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _rdt(bits):
    return np.float64 if bits == 64 else np.float32


def _cdt(bits):
    return np.complex128 if bits == 64 else np.complex64


def _ia(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32))


def _fa(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32))


# ------------------------------------------------------------------ RNG ----
def philox(ctr, key, bits=64):
    out = np.zeros(4, np.uint32)
    lib(bits).orc_philox4x32_10(_ptr(np.asarray(ctr, np.uint32)), _ptr(np.asarray(key, np.uint32)), _ptr(out))
    return out


def normals(n, seed, c1, c2, c3, bits=64):
    out = np.zeros(n, np.float64)
    lib(bits).orc_philox_normals(_ptr(out), n, seed, c1, c2, c3)
    return out


# ------------------------------------------------------------------ FFT ----
def fft3d(a, sign, bits=64):
    """Unnormalised transform of an x-fastest grid array shaped (nz, ny, nx)."""
    a = np.ascontiguousarray(a, dtype=_cdt(bits)).copy()
    nz, ny, nx = a.shape
    lib(bits).orc_fft3d(_ptr(a), nx, ny, nz, sign)
    return a


def box_fft(a, Nx, direction, bits=64):
    a = np.ascontiguousarray(a, dtype=_cdt(bits)).copy()
    lib(bits).orc_box_fft(_ptr(a), _ptr(_ia(Nx)), len(Nx), direction)
    return a


def kvec_f(i, Nx, L):
    k = np.zeros(3, np.float32)
    k2 = lib().orc_kvec_f(i, _ptr(_ia(list(Nx) + [1] * (3 - len(Nx)))), _ptr(_fa(list(L) + [1] * (3 - len(L)))), len(Nx), _ptr(k))
    return k2, k


def kvec_d(i, Nx, L):
    k = np.zeros(3, np.float64)
    k2 = lib().orc_kvec_d(i, _ptr(_ia(list(Nx) + [1] * (3 - len(Nx)))), _ptr(_fa(list(L) + [1] * (3 - len(L)))), len(Nx), _ptr(k))
    return k2, k


# ----------------------------------------------------------------- TILD ----
def spline_weights(s, p, bits=64):
    W = np.zeros(6, _rdt(bits))
    lib(bits).orc_spline_weights(s, p, _ptr(W))
    return W[: p + 1]


def grid_weights(x, Nx, dx, p, bits=64):
    x = _fa(x).reshape(-1)
    Dim = len(Nx)
    ns = x.size // Dim
    G = (p + 1) ** Dim
    inds = np.zeros(ns * G, np.int32)
    W = np.zeros(ns * G, _rdt(bits))
    lib(bits).orc_grid_weights(_ptr(x), ns, _ptr(_ia(list(Nx) + [1])), _ptr(_fa(list(dx) + [1])), p, Dim, _ptr(inds), _ptr(W))
    return inds, W


def fill_density(M, sites, inds, W, G, q=None, bits=64):
    rho = np.zeros(M, _rdt(bits))
    rhoq = np.zeros(M, _rdt(bits)) if q is not None else None
    sites = _ia(sites)
    lib(bits).orc_fill_density(_ptr(rho), _ptr(rhoq), _ptr(_fa(q) if q is not None else None), _ptr(sites), sites.size, _ptr(inds), _ptr(W), G)
    return rho, rhoq


POT_KIND = {"gaussian": 0, "erf2": 1, "erfG": 2, "charges": 3}


def uk_table(kind, A, sigma, Rp, Nx, L, bits=64):
    M = int(np.prod(Nx))
    uk = np.zeros(M, _rdt(bits))
    lib(bits).orc_uk_table(POT_KIND[kind], A, sigma, Rp, _ptr(_ia(list(Nx) + [1])), _ptr(_fa(list(L) + [1])), len(Nx), _ptr(uk))
    return uk


def potential_grid_force(rhoJ, uk, Nx, L, factor, gridF, bits=64):
    lib(bits).orc_potential_grid_force(_ptr(rhoJ), _ptr(uk), _ptr(_ia(list(Nx) + [1])), _ptr(_fa(list(L) + [1])), len(Nx), factor, _ptr(gridF))


def potential_energy(rhoI, rhoJ, uk, Nx, gvol, bits=64):
    return lib(bits).orc_potential_energy(_ptr(rhoI), _ptr(rhoJ), _ptr(uk), _ptr(_ia(list(Nx) + [1])), len(Nx), gvol)


def pair_energy(kind, rhoI, rhoJ, uk, Nx, gvol, bits=64):
    """One potential's CalcEnergy: gvol sum rho_I (u * rho_J)
    (PS_Potential::CalcEnergy, ps_potential.cu:133-178); the charges potential
    takes half of it, 0.5 gvol sum rho_q (u * rho_q) (NBCharge::CalcEnergy,
    ps_potentialCharges.cu:115)."""
    e = potential_energy(rhoI, rhoJ, uk, Nx, gvol, bits)
    return 0.5 * e if kind == "charges" else e


def map_forces(f, gridF, sites, inds, W, G, gvol, Dim, q=None, bits=64):
    sites = _ia(sites)
    lib(bits).orc_map_forces(_ptr(f), _ptr(gridF), _ptr(_fa(q) if q is not None else None), _ptr(sites), sites.size, _ptr(inds), _ptr(W), G, gvol, Dim)


def _angle_terms(ang, x, L):
    """Per (particle, angle) entries of d_angles / d_anglesStressEnergy
    (ps_deviceAngles.cu:8-290) in FP64: the owner particle, the force on it
    and the angle's energy, with the reference's skips (a bond vector shorter
    than 1e-2, |cos| > 1 for the harmonic style)."""
    Dim = len(L)
    Lf = np.asarray(_fa(L), np.float64)
    Lh = np.asarray(_fa(0.5 * _fa(L)), np.float64)
    xr = np.asarray(x).reshape(-1, Dim).astype(np.float64)
    na = np.asarray(ang["nAngles"])
    ma = int(ang["maxangles"])
    at = np.asarray(ang["angleType"]).reshape(-1, ma)
    ag = np.asarray(ang["angleGroup"]).reshape(-1, ma, 3)
    owner, slot = np.nonzero(np.arange(ma)[None, :] < na[:, None])
    t = at[owner, slot]
    ai, aj, ak = ag[owner, slot, 0], ag[owner, slot, 1], ag[owner, slot, 2]

    def mi(d):
        d = np.where(d > Lh, d - Lf, d)
        return np.where(d < -Lh, d + Lf, d)

    rij = mi(xr[ai] - xr[aj])
    rkj = mi(xr[ak] - xr[aj])
    mrij2 = np.sum(rij * rij, axis=1)
    mrkj2 = np.sum(rkj * rkj, axis=1)
    ok = (mrij2 >= 1e-4) & (mrkj2 >= 1e-4)
    mrij = np.sqrt(np.maximum(mrij2, 1e-300))
    mrkj = np.sqrt(np.maximum(mrkj2, 1e-300))
    cos = np.sum(rij * rkj, axis=1) / mrij / mrkj
    k = np.asarray(ang["k"], np.float32).astype(np.float64)[t]
    th0 = np.asarray(ang["theta"], np.float32).astype(np.float64)[t]
    style = np.asarray(ang["style"])[t]
    dci = rkj / (mrij * mrkj)[:, None] - (cos / mrij2)[:, None] * rij
    dck = rij / (mrij * mrkj)[:, None] - (cos / mrkj2)[:, None] * rkj
    harm = style == 1
    ok &= ~(harm & ((cos < -1.0) | (cos > 1.0)))
    cc = np.clip(cos, -1.0, 1.0)
    theta = np.arccos(cc)
    denom = np.sqrt(1.0 - cc * cc)
    isin = np.where(denom > 1e-4, 1.0 / np.maximum(denom, 1e-300), 0.0)
    dth = theta - th0
    pre = np.where(harm, 2.0 * k * dth * isin, -k)
    fi = pre[:, None] * dci
    fk = pre[:, None] * dck
    en = np.where(harm, k * dth * dth, k * (1.0 + cos))
    fo = np.where((owner == ai)[:, None], fi,
                  np.where((owner == aj)[:, None], -(fi + fk), np.where((owner == ak)[:, None], fk, 0.0)))
    fo[~ok] = 0.0
    en = np.where(ok, en, 0.0)
    return owner, fo, en


def angle_forces(f, ang, x, L):
    """Adds the angle forces of every particle to f (d_angles, ps_deviceAngles.cu:8-145)."""
    Dim = len(L)
    owner, fo, _ = _angle_terms(ang, x, L)
    fr = f.reshape(-1, Dim)
    np.add.at(fr, owner, fo.astype(fr.dtype))


def angle_energy(ang, x, L):
    """Uangle = sum over particles of their angles' energies / 3 (PS_Box.cu:296-305)."""
    _, _, en = _angle_terms(ang, x, L)
    return float(np.sum(en)) / 3.0


def bonds(f, topo, x, L, bits=64):
    Lf = _fa(L)
    Lh = _fa(0.5 * Lf)
    ns = topo["nBonds"].size
    lib(bits).orc_bonds(_ptr(f), _ptr(topo["nBonds"]), _ptr(topo["bondedTo"]), _ptr(topo["bondType"]),
                        _ptr(_fa(topo["bondReq"])), _ptr(_fa(topo["bondK"])), _ptr(_ia(topo["bondStyle"])),
                        _ptr(_fa(x)), _ptr(Lf), _ptr(Lh), ns, topo["maxbonds"], len(L))


def bond_energy(topo, x, L, bits=64):
    Lf = _fa(L)
    Lh = _fa(0.5 * Lf)
    ns = topo["nBonds"].size
    e = np.zeros(ns, _rdt(bits))
    lib(bits).orc_bond_energy(_ptr(e), _ptr(topo["nBonds"]), _ptr(topo["bondedTo"]), _ptr(topo["bondType"]),
                              _ptr(_fa(topo["bondReq"])), _ptr(_fa(topo["bondK"])), _ptr(_ia(topo["bondStyle"])),
                              _ptr(_fa(x)), _ptr(Lf), _ptr(Lh), ns, topo["maxbonds"], len(L))
    return e


def bond_virial(topo, x, L, bits=64):
    """Per-particle bond virial terms of d_bondStressEnergy
    (ps_deviceBonds.cu:85-171): [ns, 6] as xx, yy, zz, xy, xz, yz."""
    Lf = _fa(L)
    Lh = _fa(0.5 * Lf)
    ns = topo["nBonds"].size
    v = np.zeros(ns * 6, _rdt(bits))
    lib(bits).orc_bond_virial(_ptr(v), _ptr(topo["nBonds"]), _ptr(topo["bondedTo"]), _ptr(topo["bondType"]),
                              _ptr(_fa(topo["bondReq"])), _ptr(_fa(topo["bondK"])), _ptr(_ia(topo["bondStyle"])),
                              _ptr(_fa(x)), _ptr(Lf), _ptr(Lh), ns, topo["maxbonds"], len(L))
    return v.reshape(ns, 6)


def gjf_coeffs(delt, mass, mobility):
    out = np.zeros(5, np.float64)
    lib().orc_gjf_coeffs(delt, mass, mobility, _ptr(out))
    return out


def make_linear(nmolecs, blocks, species_index, block_bond_type=None, charges=None, maxbonds=2, start=0,
                mid0=0):
    """Topology of PS_Box::makeLinear (PS_Box_moleculeMakers.cu:15-311).

    blocks = [(N_b, species_name)], species_index maps names to integer types.
    Returns intSpecies, nBonds, bondedTo, bondType, charge, mID arrays for the new
    molecules; bead order is molecule-major then block-major.  start is the
    index of the first new bead; mid0 the id of the first new molecule, which
    the reference takes as the largest id already in the box, 0 when empty
    (PS_Box_moleculeMakers.cu:140-145; ids then grow by one per molecule, :306).
    """
    nb = len(blocks)
    if block_bond_type is None:
        block_bond_type = [1] * nb
    Ntot = sum(n for n, _ in blocks)
    ns = nmolecs * Ntot
    intSpecies = np.zeros(ns, np.int32)
    nBonds = np.zeros(ns, np.int32)
    bondedTo = np.zeros(ns * maxbonds, np.int32)
    bondType = np.zeros(ns * maxbonds, np.int32)
    q = np.zeros(ns, np.float32)
    mID = np.zeros(ns, np.int32)
    ind = 0
    for m in range(nmolecs):
        for j, (Nb, sp) in enumerate(blocks):
            for s in range(Nb):
                gi = start + ind
                intSpecies[ind] = species_index[sp]
                if charges is not None:
                    q[ind] = charges[j]
                if j != 0 or s != 0:
                    bondedTo[ind * maxbonds + nBonds[ind]] = gi - 1
                    bondType[ind * maxbonds + nBonds[ind]] = block_bond_type[j] - 1
                    nBonds[ind] += 1
                if j != nb - 1 or s != Nb - 1:
                    bondedTo[ind * maxbonds + nBonds[ind]] = gi + 1
                    bondType[ind * maxbonds + nBonds[ind]] = block_bond_type[j] - 1
                    nBonds[ind] += 1
                mID[ind] = mid0 + m
                ind += 1
    return dict(intSpecies=intSpecies, nBonds=nBonds, bondedTo=bondedTo, bondType=bondType, charge=q, mID=mID)


@dataclass
class TildPotential:
    kind: str            # gaussian | erf2 | erfG | charges | bias
    I: str = "charges"   # group names (species names, "all", or custom groups)
    J: str = "charges"
    A: float = 0.0       # amplitude (l_B for charges, Ao for bias)
    sigma: float = 0.0
    Rp: float = 0.0
    phase: str = "L"     # bias: L | H | G
    dir: int = 0         # bias direction, wall normal axis
    n_periods: int = 1
    wall_pos: float = 0.0
    wall_sign: int = 1


@dataclass
class TildSystem:
    """Reference PS_Box state needed by one time step (FP32 positions)."""
    Nx: tuple
    L: tuple
    pmeorder: int
    species: list        # names, index = integer type
    x: np.ndarray        # [ns*Dim] float32
    topo: dict           # make_linear output merged over molecules
    bond_k: list
    bond_r0: list
    bond_style: list
    potentials: list
    mass: list = None
    mobility: list = None
    delt: float = 0.005
    integrator: str = "GJF"
    groups: dict = field(default_factory=dict)   # extra named groups -> list of species names
    xo: np.ndarray = None
    noise_old: np.ndarray = None
    # nAngles [ns], angleType [ns, maxangles], angleGroup [ns, maxangles, 3],
    # k, theta (radians), style per angle type (None: no angles)
    angles: dict = None
    # velocity Verlet state: velocities and the force array of the last step
    # [ns*Dim] (None: zero, as the reference allocates them)
    v: np.ndarray = None
    f_prev: np.ndarray = None
    seed: int = 0        # Langevin noise key, with the step below
    step: int = 0

    @property
    def Dim(self):
        return len(self.Nx)

    @property
    def M(self):
        return int(np.prod(self.Nx))

    @property
    def ns(self):
        return self.topo["intSpecies"].size

    def dxf(self):
        # PS_Box_initialization.cu:378 (double) then :727-731 (cast to float)
        return np.array([np.float32(float(np.float32(l)) / n) for l, n in zip(self.L, self.Nx)], np.float32)

    def gvol(self):
        # the device kernels use the float product of the float spacings
        g = np.float32(1.0)
        for d in self.dxf():
            g = np.float32(g * d)
        return float(g)

    def group_sites(self, name):
        t = self.topo["intSpecies"]
        if name in ("all", "All"):
            return np.arange(self.ns, dtype=np.int32)
        if name == "charges":
            return np.nonzero(self.topo["charge"] != 0.0)[0].astype(np.int32)
        if name in self.species:
            return np.nonzero(t == self.species.index(name))[0].astype(np.int32)
        if name in self.groups:
            ids = [self.species.index(s) for s in self.groups[name]]
            return np.nonzero(np.isin(t, ids))[0].astype(np.int32)
        raise KeyError(name)


def tild_system_from_config(cfg, integrator="GJF"):
    """TildSystem for a workload dict of mft.systems (species masks -> group names)."""
    topo = cfg["topo"]
    names = cfg["species"]
    pots = []
    for p in cfg["potentials"]:
        if p["kind"] == "charges":
            pots.append(TildPotential("charges", A=p["A"], sigma=p["sigma"]))
            continue
        if p["kind"] == "dpd":
            gI = [n for i, n in enumerate(names) if (p["maskI"] >> i) & 1]
            assert len(gI) == 1
            pots.append(TildPotential("dpd", gI[0], gI[0], A=p["A"], sigma=p["sigma"], Rp=p["Rp"],
                                      wall_pos=p["wall_pos"]))
            continue
        if p["kind"] == "langevin":
            gI = [n for i, n in enumerate(names) if (p["maskI"] >> i) & 1]
            assert len(gI) == 1
            pots.append(TildPotential("langevin", gI[0], gI[0], A=p["A"], sigma=p["sigma"]))
            continue
        if p["kind"] == "wall":
            gI = [n for i, n in enumerate(names) if (p["maskI"] >> i) & 1]
            assert len(gI) == 1
            pots.append(TildPotential("wall", gI[0], gI[0], A=p["A"], dir=p.get("dir", 0), wall_pos=p["wall_pos"],
                                      wall_sign=p.get("wall_sign", 1)))
            continue
        if p["kind"] == "bias":
            gI = [n for i, n in enumerate(names) if (p["maskI"] >> i) & 1]
            assert len(gI) == 1
            pots.append(TildPotential("bias", gI[0], gI[0], A=p["A"], phase=p.get("phase", "L"), dir=p.get("dir", 0),
                                      n_periods=p.get("n_periods", 1)))
            continue
        gI = [n for i, n in enumerate(names) if (p["maskI"] >> i) & 1]
        gJ = [n for i, n in enumerate(names) if (p["maskJ"] >> i) & 1]
        assert len(gI) == 1 and len(gJ) == 1
        pots.append(TildPotential(p["kind"], gI[0], gJ[0], A=p["A"], sigma=p["sigma"], Rp=p.get("Rp", 0.0)))
    t = dict(intSpecies=topo["type"].astype(np.int32), nBonds=topo["nBonds"].astype(np.int32),
             bondedTo=topo["bondedTo"].astype(np.int32), bondType=topo["bondType"].astype(np.int32),
             charge=topo["charge"].astype(np.float32))
    x = cfg["x"].astype(np.float32)
    return TildSystem(Nx=cfg["Nx"], L=cfg["L"], pmeorder=cfg["pmeorder"], species=names, x=x, topo=t,
                      bond_k=cfg["bond_k"], bond_r0=cfg["bond_r0"], bond_style=cfg["bond_style"],
                      potentials=pots, mass=cfg["mass"], mobility=cfg["mobility"], delt=cfg["delt"],
                      integrator=integrator, xo=x.copy(), noise_old=np.zeros_like(x), angles=cfg.get("angles"),
                      seed=cfg["seed"])


def bias_field(pot, Nx, L):
    """A u(r) of Box::make_bias_field (Box.cu:308-458) on the grid, x fastest,
    r = n dx with the float dx of the box, PI = 3.141592654f (constants.h);
    lamellar A sin(2 pi n x_dir / L_dir), cylinders along dir A cos cos,
    gyroid A ((sin x cos y + sin y cos z + sin z cos x)^2 - 1.44)."""
    f32 = np.float32
    PI = f32(3.141592654)
    n = f32(pot.n_periods)
    Lf = [f32(l) for l in L]
    dx = [f32(np.float64(Lf[j]) / Nx[j]) for j in range(3)]
    iz, iy, ix = np.meshgrid(np.arange(Nx[2]), np.arange(Nx[1]), np.arange(Nx[0]), indexing="ij")
    r = [(ix.astype(f32) * dx[0]).ravel(), (iy.astype(f32) * dx[1]).ravel(), (iz.astype(f32) * dx[2]).ravel()]
    arg = [f32(f32(f32(2.0) * PI) * n) / Lf[j] for j in range(3)]
    A = float(pot.A)
    if pot.phase == "L":
        return A * np.sin((arg[pot.dir] * r[pot.dir]).astype(np.float64))
    if pot.phase == "H":
        d1, d2 = 0, 1
        if pot.dir == d1:
            d1 = 2
        elif pot.dir == d2:
            d2 = 2
        return A * np.cos((arg[d1] * r[d1]).astype(np.float64)) * np.cos((arg[d2] * r[d2]).astype(np.float64))
    if pot.phase == "G":
        c = [np.cos((arg[j] * r[j]).astype(np.float64)).astype(f32) for j in range(3)]
        s = [np.sin((arg[j] * r[j]).astype(np.float64)).astype(f32) for j in range(3)]
        e = (s[0] * c[1] + s[1] * c[2] + s[2] * c[0]).astype(f32)
        return A * ((e * e).astype(np.float64) - 1.44)
    raise ValueError("PHASE NOT FOUND in Box.cu")


def bias_force(u, Nx, L):
    """-grad u spectrally (ps_potentialBias.cu:100-185): f_j = IFFT[-i k_j FFT[u]]
    with the float k of Box::get_kD, the Nyquist index kept; AoS [M*3]."""
    M = int(np.prod(Nx))
    uk = box_fft(np.asarray(u, np.complex128), list(Nx), 1)
    k = np.array([kvec_f(i, Nx, L)[1] for i in range(M)], np.float64)
    out = np.zeros((M, 3))
    for j in range(3):
        out[:, j] = box_fft(-1j * k[:, j] * uk, list(Nx), -1).real
    return out.reshape(-1)


def tild_forces(sys: TildSystem, bits=64, want_energy=False):
    """Particle forces at the current positions (PS_Box.cu:96-128, 240-278)."""
    Dim, M = sys.Dim, sys.M
    dx = sys.dxf()
    G = (sys.pmeorder + 1) ** Dim
    inds, W = grid_weights(sys.x, sys.Nx, dx, sys.pmeorder, bits)
    gvol_dev = sys.gvol()
    names = ["all"] + list(sys.species) + list(sys.groups)
    has_q = bool(np.any(sys.topo["charge"] != 0.0))
    if has_q:
        names.append("charges")
    rho, rhoq, gridF = {}, None, {}
    for g in names:
        sites = sys.group_sites(g)
        if g == "charges":
            rho[g], rhoq = fill_density(M, sites, inds, W, G, q=sys.topo["charge"], bits=bits)
        else:
            rho[g], _ = fill_density(M, sites, inds, W, G, bits=bits)
    f = np.zeros(sys.ns * Dim, _rdt(bits))
    bonds(f, dict(sys.topo, bondReq=sys.bond_r0, bondK=sys.bond_k, bondStyle=sys.bond_style, maxbonds=2), sys.x, sys.L, bits)
    if sys.angles is not None:
        angle_forces(f, sys.angles, sys.x, sys.L)
    energies = []
    nlang = 0
    for pot in sys.potentials:
        if pot.kind == "dpd":
            sites = np.asarray(sys.group_sites(pot.I))
            fr = f.reshape(-1, Dim)
            fr[sites] += dpd_pair_forces(sys, pot, sites)
            if want_energy:
                energies.append(0.0)
            continue
        if pot.kind == "langevin":
            # d_add_Langevin_forces (ps_potentialLangevin.cu:60-133): -v drag +
            # sqrtf(2 drag / delt) N(0, 1), normals keyed (particle, step,
            # 'LANG', potential + 1); v is VV's velocity, zero otherwise
            nlang += 1
            sites = np.asarray(sys.group_sites(pot.I))
            drag = float(np.float32(pot.A))
            nmag = float(np.sqrt(np.float32(2.0 * drag / float(np.float32(pot.sigma)))).astype(np.float32))
            z = normal4_np(sites.astype(np.uint64), sys.step, TAG_LANGEVIN, nlang, sys.seed)[:, :Dim]
            v = np.zeros((sites.size, Dim))
            if sys.integrator == "VV" and sys.v is not None:
                v = np.asarray(sys.v, np.float64).reshape(-1, Dim)[sites]
            fr = f.reshape(-1, Dim)
            fr[sites] += -v * drag + nmag * z
            if want_energy:
                energies.append(0.0)
            continue
        if pot.kind == "wall":
            # Wall (ps_potentialWall.cu:92-147): k (pos - x_a) where sign (pos - x_a) > 0
            sites = np.asarray(sys.group_sites(pot.I))
            xa = sys.x.reshape(-1, Dim)[sites, pot.dir].astype(np.float64)
            disp = np.float64(np.float32(pot.wall_pos)) - xa
            on = pot.wall_sign * disp > 0
            fr = f.reshape(-1, Dim)
            fr[sites[on], pot.dir] += float(np.float32(pot.A)) * disp[on]
            if want_energy:
                energies.append(float(np.sum(0.5 * float(np.float32(pot.A)) * disp[on] ** 2)))
            continue
        if pot.kind == "bias":
            u = bias_field(pot, sys.Nx, sys.L)
            gI = gridF.setdefault(pot.I, np.zeros(M * Dim, _rdt(bits)))
            gI += bias_force(u, sys.Nx, sys.L).astype(_rdt(bits))
            if want_energy:
                energies.append(gvol_dev * float(np.sum(rho[pot.I].astype(np.float64) * u)))
            continue
        uk = uk_table(pot.kind, pot.A, pot.sigma, pot.Rp, sys.Nx, sys.L, bits)
        if pot.kind == "charges":
            gf = gridF.setdefault("charges", np.zeros(M * Dim, _rdt(bits)))
            potential_grid_force(rhoq, uk, sys.Nx, sys.L, 1.0, gf, bits)
            if want_energy:
                energies.append(pair_energy("charges", rhoq, rhoq, uk, sys.Nx, gvol_dev, bits))
            continue
        gI = gridF.setdefault(pot.I, np.zeros(M * Dim, _rdt(bits)))
        potential_grid_force(rho[pot.J], uk, sys.Nx, sys.L, 2.0 if pot.I == pot.J else 1.0, gI, bits)
        if pot.I != pot.J:
            gJ = gridF.setdefault(pot.J, np.zeros(M * Dim, _rdt(bits)))
            potential_grid_force(rho[pot.I], uk, sys.Nx, sys.L, 1.0, gJ, bits)
        if want_energy:
            energies.append(pair_energy(pot.kind, rho[pot.I], rho[pot.J], uk, sys.Nx, gvol_dev, bits))
    for g, gf in gridF.items():
        q = sys.topo["charge"] if g == "charges" else None
        map_forces(f, gf, sys.group_sites(g), inds, W, G, gvol_dev, Dim, q=q, bits=bits)
    out = dict(f=f, inds=inds, W=W, rho=rho, rhoq=rhoq, gridF=gridF)
    if want_energy:
        out["energies"] = energies
    return out


def philox_np(c0, c1, c2, c3, seed):
    """Vectorised Philox4x32-10 over counter arrays (same algorithm as
    orc_philox4x32_10; cross-checked in the tests)."""
    M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    k0 = np.uint64(seed & 0xFFFFFFFF)
    k1 = np.uint64((seed >> 32) & 0xFFFFFFFF)
    mask = np.uint64(0xFFFFFFFF)
    n = max(np.size(c0), np.size(c1), np.size(c2), np.size(c3))
    c = [np.broadcast_to(np.asarray(v, np.uint64), (n,)).copy() for v in (c0, c1, c2, c3)]
    for _ in range(10):
        p0 = M0 * c[0]
        p1 = M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c = [(hi1 ^ c[1] ^ k0) & mask, lo1, (hi0 ^ c[3] ^ k1) & mask, lo0]
        k0 = (k0 + np.uint64(0x9E3779B9)) & mask
        k1 = (k1 + np.uint64(0xBB67AE85)) & mask
    return [v.astype(np.uint32) for v in c]


def normal4_np(c0, c1, c2, c3, seed):
    """Box-Muller of philox_np output -> array [n, 4] of unit normals."""
    r = philox_np(c0, c1, c2, c3, seed)
    out = []
    for h in range(2):
        u1 = ((r[2 * h] >> np.uint32(8)).astype(np.float64) + 1.0) * 2.0 ** -24
        u2 = (r[2 * h + 1] >> np.uint32(8)).astype(np.float64) * 2.0 ** -24
        rad = np.sqrt(-2.0 * np.log(u1))
        out += [rad * np.cos(2.0 * np.pi * u2), rad * np.sin(2.0 * np.pi * u2)]
    return np.stack(out, axis=1)


TAG_TILD = 0x54494C44  # 'TILD'
TAG_LANGEVIN = 0x4C414E47  # 'LANG'
TAG_DPD = 0x44504450       # 'DPDP'
TAG_FTS = 0x4654534E   # 'FTSN'


def tild_noise(ns, seed, step, Dim=3):
    """Unit normals for the integrator: counter (particle, step, 'TILD', 0),
    normals 0..Dim-1 of each Philox block."""
    z = normal4_np(np.arange(ns, dtype=np.uint64), step, TAG_TILD, 0, seed)
    return np.ascontiguousarray(z[:, :Dim]).reshape(-1)


def tild_step(sys: TildSystem, xi=None, bits=64):
    """One PS_Box::doTimeStep with the GJF, EM or VV integrator on group 'all'."""
    if sys.integrator == "VV":
        return _vv_step(sys, bits)
    out = tild_forces(sys, bits)
    f = out["f"]
    sites = np.arange(sys.ns, dtype=np.int32)
    Lf = _fa(sys.L)
    x = sys.x.astype(_rdt(bits))
    if sys.integrator == "GJF":
        coef = np.concatenate([gjf_coeffs(sys.delt, m, D) for m, D in zip(sys.mass, sys.mobility)])
        xo = sys.xo.astype(_rdt(bits))
        no = sys.noise_old.astype(_rdt(bits))
        lib(bits).orc_gjf_step(_ptr(x), _ptr(xo), _ptr(no), _ptr(f), _ptr(coef), _ptr(sys.topo["intSpecies"]), _ptr(Lf), _ptr(sites), sys.ns, sys.Dim, _ptr(np.ascontiguousarray(xi, np.float64)))
        out.update(x=x, xo=xo, noise_old=no)
    else:
        lib(bits).orc_em_step(_ptr(x), _ptr(f), sys.delt, _ptr(Lf), _ptr(sites), sys.ns, sys.Dim, _ptr(np.ascontiguousarray(xi, np.float64)))
        out.update(x=x)
    return out


def dpd_pair_forces(sys: TildSystem, pot: TildPotential, sites):
    """d_DPD_forces (ps_potentialDPD.cu:63-137) over every pair of the group
    closer than rcut (PS_NeighborList has no skin and is rebuilt each step,
    ps_neighborList.cu:200-272): F_i += Fmag r^, F_j -= Fmag r^ with r = x_i - x_j,
    Fmag = -gamma w^2 (v_i - v_j).r^ + sqrtf(2 gamma kT) w xi / sqrtf(delt),
    w = 1 - r / rcut.  pot.A = gamma, pot.sigma = kT, pot.Rp = rcut,
    pot.wall_pos = delt; xi is one normal per pair keyed (lower id, higher
    id, 'DPDP', step); v is VV's velocity, zero otherwise."""
    D = sys.Dim
    x = np.asarray(sys.x, np.float64).reshape(-1, D)[sites]
    v = np.zeros_like(x)
    if sys.integrator == "VV" and sys.v is not None:
        v = np.asarray(sys.v, np.float64).reshape(-1, D)[sites]
    L = np.asarray(_fa(sys.L), np.float64)
    Lh = np.asarray(_fa(0.5 * _fa(sys.L)), np.float64)
    g32, kT32 = np.float32(pot.A), np.float32(pot.sigma)
    gamma = float(g32)
    sig = float(np.sqrt(np.float32(np.float32(2.0) * g32 * kT32)))
    rcut = float(np.float32(pot.Rp))
    rc2 = float(np.float32(rcut) * np.float32(rcut))
    isd = float(np.float32(1.0) / np.sqrt(np.float32(pot.wall_pos)))
    n = x.shape[0]
    i, j = np.triu_indices(n, 1)
    dr = x[i] - x[j]
    dr = np.where(dr > Lh, dr - L, np.where(dr < -Lh, dr + L, dr))
    r2 = np.sum(dr * dr, axis=1)
    keep = (r2 < rc2) & (r2 > 0.0)
    i, j, dr, r2 = i[keep], j[keep], dr[keep], r2[keep]
    r = np.sqrt(r2)
    w = 1.0 - r / rcut
    rhat = dr / r[:, None]
    vdot = np.sum((v[i] - v[j]) * rhat, axis=1)
    a, b = sites[i].astype(np.uint64), sites[j].astype(np.uint64)
    xi = normal4_np(np.minimum(a, b), np.maximum(a, b), TAG_DPD, sys.step, sys.seed)[:, 0]
    fmag = -gamma * w * w * vdot + sig * w * xi * isd
    F = np.zeros((n, D))
    np.add.at(F, i, fmag[:, None] * rhat)
    np.add.at(F, j, -fmag[:, None] * rhat)
    return F


def _vv_step(sys: TildSystem, bits=64):
    """PS_Box::doTimeStep (PS_Box.cu:73-207) with VV (ps_integratorVV.cu:15-94):
    Integrate_1 -- v += (dt/2) f_prev / m, x += dt v, wrapped into [0, L) --
    then the forces at the new positions, then Integrate_2 -- v += (dt/2) f / m.
    Positions are rounded to FP32 before the forces, as the device stores them."""
    D = sys.Dim
    t = sys.topo["intSpecies"]
    inv_m = np.array([float(np.float32(1.0) / np.float32(m)) for m in sys.mass], np.float64)[t][:, None]
    hd = float(np.float32(sys.delt) * np.float32(0.5))
    L = np.asarray(_fa(sys.L), np.float64)
    v = np.zeros((sys.ns, D)) if sys.v is None else np.asarray(sys.v, np.float64).reshape(-1, D).copy()
    fp = np.zeros((sys.ns, D)) if sys.f_prev is None else np.asarray(sys.f_prev, np.float64).reshape(-1, D)
    v += hd * inv_m * fp
    x = np.asarray(sys.x, np.float64).reshape(-1, D) + (2.0 * hd) * v
    x = np.where(x >= L, x - L, np.where(x < 0.0, x + L, x))
    x0, v0 = sys.x, sys.v
    sys.x = x.astype(np.float32).reshape(-1)
    sys.v = v.reshape(-1)          # the velocities the forces see (Langevin drag)
    out = tild_forces(sys, bits)
    sys.x, sys.v = x0, v0
    f = out["f"].reshape(-1, D).astype(np.float64)
    v += hd * inv_m * f
    out.update(x=x.reshape(-1), v=v.reshape(-1), f_prev=f.reshape(-1))
    return out


# ------------------------------------------------------------------ FTS ----
def fts_bond_fft(Nx, L, Nr, bits=64):
    g = np.zeros(int(np.prod(Nx)), _cdt(bits))
    lib(bits).orc_fts_bond_fft(_ptr(_ia(list(Nx) + [1])), _ptr(_fa(list(L) + [1])), len(Nx), Nr, _ptr(g))
    return g


@dataclass
class FtsPotential:
    kind: str                 # Helfand | Flory | Incompress | Edwards (kappaN holds B)
    I: int = -1               # species indices (Flory)
    J: int = -1
    chiN: float = 0.0
    kappaN: float = 0.0
    dt_pl: float = 0.0
    dt_mi: float = 0.0
    wpl: np.ndarray = None
    wmi: np.ndarray = None
    scheme: str = "EM"        # EM | EMPC | 1S (updateScheme keyword)
    zero_mean: bool = False   # modify zeroMean (no effect on Flory, fts_potential_flory.cu:94-98)
    saved: dict = None        # EMPC predictor data: fields, forces and noise
    Akpl: np.ndarray = None   # 1S linear coefficients of w+ and w- (real, k space)
    Akmi: np.ndarray = None


@dataclass
class FtsMolecule:
    blocks: list              # [(N_b, species index)]
    phi: float = -1.0
    nmolecs: float = -1.0
    activity: float = -1.0    # > 0: grand canonical (fts_molecule.cu:30-31)


@dataclass
class FtsSystem:
    """Reference FTS_Box state for CL iterations (FP64 fields)."""
    Nx: tuple
    L: tuple
    Nr: float
    C: float
    nspecies: int
    potentials: list
    molecules: list
    rho: np.ndarray = None    # [nspecies, M] species densities (previous iteration)
    W: np.ndarray = None      # [nspecies, M] species fields
    Q: list = None

    @property
    def M(self):
        return int(np.prod(self.Nx))

    def V(self):
        v = 1.0
        for l in self.L:
            v *= float(np.float32(l))
        return v

    def gvol(self):
        g = 1.0
        for l, n in zip(self.L, self.Nx):
            g *= float(np.float32(l)) / float(n)
        return g

    def init_nmolecs(self):
        # LinearMolec::computeLinearTerms (fts_molecule_linear.cu:433-435); a
        # grand canonical molecule keeps phi = -1 (fts_molecule.cu:19), so its
        # count comes out negative and the activity branches apply
        for m in self.molecules:
            Ntot = sum(n for n, _ in m.blocks)
            phi = -1.0 if m.activity > 0.0 else m.phi
            m.nmolecs = self.C * phi * self.V() * self.Nr / float(Ntot)


def fts_build_species_fields(sys: FtsSystem):
    """FTS_Species::buildPotentialField (fts_species.cu:53-132)."""
    M = sys.M
    W = np.zeros((sys.nspecies, M), np.complex128)
    for s in range(sys.nspecies):
        for p in sys.potentials:
            if p.kind in ("Helfand", "Incompress", "Edwards"):
                W[s] += 1j * p.wpl
            elif p.kind == "Flory":
                if s == p.I:
                    W[s] -= p.wmi
                    W[s] += 1j * p.wpl
                if s == p.J:
                    W[s] += p.wmi
                    W[s] += 1j * p.wpl
        W[s] *= 1.0 / sys.Nr
    sys.W = W
    return W


def fts_molecule_densities(sys: FtsSystem, bits=64, keep=False):
    """LinearMolec::calcDensity for every molecule (fts_molecule_linear.cu:278-363)."""
    M = sys.M
    g = fts_bond_fft(sys.Nx, sys.L, sys.Nr, bits)
    rho = np.zeros((sys.nspecies, M), _cdt(bits))
    Ws = np.ascontiguousarray(sys.W, _cdt(bits))
    sys.Q = []
    extra = []
    L = lib(bits)
    for m in sys.molecules:
        nb = len(m.blocks)
        N = _ia([n for n, _ in m.blocks])
        spec = _ia([s for _, s in m.blocks])
        Ntot = int(N.sum())
        q = np.zeros(Ntot * M, _cdt(bits))
        qd = np.zeros(Ntot * M, _cdt(bits))
        rb = np.zeros(nb * M, _cdt(bits))
        Qc = _propagators(L, Ws, nb, N, spec, g, sys.Nx, sys.gvol(), sys.V(), q, qd, bits)
        # fts_molecule_linear.cu:286-291
        factor = m.nmolecs / Qc / sys.V() if m.nmolecs > 0.0 else m.activity
        _density(L, Ws, nb, N, spec, q, qd, factor, M, rb, rho, bits)
        sys.Q.append(Qc)
        if keep:
            extra.append(dict(q=q, qdag=qd, rhoBlock=rb))
    sys.rho = rho.astype(np.complex128)
    return extra


class _c64(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class _c32(C.Structure):
    _fields_ = [("re", C.c_float), ("im", C.c_float)]


def _propagators(L, W, nb, N, spec, g, Nx, gvol, V, q, qd, bits):
    fn = L.orc_fts_propagators
    fn.restype = _c64 if bits == 64 else _c32
    fn.argtypes = [_p, _i, _p, _p, _p, _p, _i, _d, _d, _p, _p]
    r = fn(_ptr(W), nb, _ptr(N), _ptr(spec), _ptr(g), _ptr(_ia(list(Nx) + [1])), len(Nx), gvol, V, _ptr(q), _ptr(qd))
    return complex(r.re, r.im)


def _density(L, W, nb, N, spec, q, qd, factor, M, rb, rho, bits):
    fn = L.orc_fts_density
    ct = _c64 if bits == 64 else _c32
    fn.restype = None
    fn.argtypes = [_p, _i, _p, _p, _p, _p, ct, _i, _p, _p]
    fn(_ptr(W), nb, _ptr(N), _ptr(spec), _ptr(q), _ptr(qd), ct(factor.real, factor.imag), M, _ptr(rb), _ptr(rho))


def fts_noise(M, seed, it, stream=0):
    """CL noise normals: counter (site, iteration, 'FTSN', stream); normal j of
    the block feeds the j-th noisy field in potential declaration order."""
    return normal4_np(np.arange(M, dtype=np.uint64), it, TAG_FTS, stream, seed)


def _forces(sys: FtsSystem, p: FtsPotential):
    """dH/dw of one potential at its current fields and the current species
    densities: d_makeHelfandForce (fts_potential_helfand.cu:223-240),
    d_makeIncompressForce (fts_potential_incompress.cu:202-218),
    d_makeFloryForce (fts_potential_flory.cu:279-304)."""
    M = sys.M
    L = lib()
    rho_tot = sys.rho.sum(axis=0)
    if p.kind == "Helfand":
        dH = np.zeros(M, np.complex128)
        L.orc_fts_helfand_force(_ptr(dH), _ptr(p.wpl), _ptr(rho_tot), sys.C, p.kappaN, sys.Nr, M)
        return dH, None
    if p.kind == "Incompress":
        dH = np.zeros(M, np.complex128)
        L.orc_fts_incompress_force(_ptr(dH), _ptr(p.wpl), _ptr(rho_tot), sys.C, sys.Nr, M)
        return dH, None
    if p.kind == "Edwards":
        dH = np.zeros(M, np.complex128)
        L.orc_fts_edwards_force(_ptr(dH), _ptr(p.wpl), _ptr(rho_tot), p.kappaN, sys.Nr, M)
        return dH, None
    dHp = np.zeros(M, np.complex128)
    dHm = np.zeros(M, np.complex128)
    rI = np.ascontiguousarray(sys.rho[p.I])
    rJ = np.ascontiguousarray(sys.rho[p.J])
    L.orc_fts_flory_force(_ptr(dHp), _ptr(dHm), _ptr(p.wpl), _ptr(p.wmi), _ptr(rI), _ptr(rJ), sys.C, p.chiN, sys.Nr, M)
    return dHp, dHm


def fts_k2(sys: FtsSystem):
    """|k|^2 of every grid index (Box::get_kD double version, Box.cu:181-203)."""
    return np.array([kvec_d(i, sys.Nx, sys.L)[0] for i in range(sys.M)])


def debye_homopolymer(k2, alpha):
    """FTS_Box::computeHomopolyDebye (FTS_Box.cu:537-550): g(0) = 1,
    g = 2 (exp(-k^2 a) + k^2 a - 1) / k^4."""
    g = np.ones_like(k2)
    k = k2[1:]
    g[1:] = 2.0 * (np.exp(-k * alpha) + k * alpha - 1.0) / (k * k)
    return g


def debye_intra_block(k2, f, alpha):
    """FTS_Box::computeIntRABlockDebye (FTS_Box.cu:552-566): the same with
    alpha f, and 0 at k = 0."""
    g = np.zeros_like(k2)
    k = k2[1:]
    g[1:] = 2.0 * (np.exp(-k * alpha * f) + k * alpha * f - 1.0) / (k * k)
    return g


def debye_inter_block(k2, fA, fB, fC, alpha):
    """FTS_Box::computeIntERBlockDebye (FTS_Box.cu:571-587): blocks A and C
    separated by B, with the factor 2; 0 at k = 0."""
    g = np.zeros_like(k2)
    k = k2[1:]
    g[1:] = 2.0 * (1.0 - np.exp(-k * alpha * fA)) * np.exp(-k * alpha * fB) * (1.0 - np.exp(-k * alpha * fC)) / (k * k)
    return g


def fts_linear_coeffs(sys: FtsSystem):
    """initLinearCoeffs of every potential (fts_potential_helfand.cu:277-280,
    fts_potential_flory.cu:323-326, fts_potential_incompress.cu:249-252) then
    LinearMolec::computeLinearTerms of every molecule
    (fts_molecule_linear.cu:433-560): the Debye terms go to A+ of Helfand and
    Flory potentials only (A- of Flory stays 2C/chiN)."""
    k2 = fts_k2(sys)
    M = sys.M
    for p in sys.potentials:
        if p.kind == "Helfand":
            p.Akpl = np.full(M, sys.C / p.kappaN)
        elif p.kind == "Edwards":
            p.Akpl = np.full(M, 1.0 / p.kappaN)
        elif p.kind == "Flory":
            p.Akpl = np.full(M, 2.0 * sys.C / p.chiN)
            p.Akmi = p.Akpl.copy()
        else:
            p.Akpl = np.zeros(M)
    for m in sys.molecules:
        N = [n for n, _ in m.blocks]
        spec = [s for _, s in m.blocks]
        Ntot = sum(N)
        alpha = Ntot / sys.Nr
        pref = m.nmolecs * alpha * Ntot / sys.V()
        h = pref * debye_homopolymer(k2, alpha)
        for p in sys.potentials:
            if p.kind in ("Helfand", "Edwards"):   # fts_molecule_linear.cu:459
                p.Akpl = p.Akpl + h
        if len(N) == 1:
            for p in sys.potentials:
                if p.kind == "Flory":
                    p.Akpl = p.Akpl + h
            continue
        for p in sys.potentials:
            if p.kind != "Flory":
                continue
            acts = (p.I, p.J)
            for j in range(len(N) - 1):
                if spec[j] not in acts:
                    continue
                fA = N[j] / Ntot
                gMT = debye_intra_block(k2, fA, alpha)
                nbetween = 0
                for k in range(j + 1, len(N)):
                    if spec[k] not in acts:
                        continue
                    fC = N[k] / Ntot
                    fB = nbetween / Ntot
                    gMT = gMT + debye_intra_block(k2, fC, alpha) + debye_inter_block(k2, fA, fB, fC, alpha)
                    p.Akpl = p.Akpl + pref * gMT
                    nbetween += N[k]


def _noises(sys: FtsSystem, p: FtsPotential, z, col):
    """The potential's CL noise fields (real, sqrt(2 dt / gvol) N(0,1)); the
    incompressible update has none (fts_potential_incompress.cu:105)."""
    gvol = sys.gvol()
    if z is None or p.kind == "Incompress":
        return None, None, col
    npl = (np.sqrt(2.0 * p.dt_pl / gvol) * z[:, col]).astype(np.complex128)
    if p.kind in ("Helfand", "Edwards"):
        return npl, None, col + 1
    nmi = (np.sqrt(2.0 * p.dt_mi / gvol) * z[:, col + 1]).astype(np.complex128)
    return npl, nmi, col + 2


def fts_update_fields(sys: FtsSystem, z=None):
    """Every potential's updateFields with EM (+ real CL noise when z is given):
    fts_potential_helfand.cu:77-152, fts_potential_flory.cu:123-167,
    fts_potential_incompress.cu:84-113.  An EMPC potential first keeps its
    fields, forces and noise for the corrector (storePredictorData,
    fts_potential_flory.cu:243-249)."""
    M = sys.M
    L = lib()
    col = 0
    for p in sys.potentials:
        dHp, dHm = _forces(sys, p)
        npl, nmi, col = _noises(sys, p, z, col)
        if p.scheme == "1S":
            # k-space semi-implicit step without noise (fts_potential_flory.cu:172-193)
            if p.Akpl is None:
                fts_linear_coeffs(sys)
            pairs = [(p.wpl, dHp, p.Akpl, p.dt_pl)]
            if p.kind == "Flory":
                pairs.append((p.wmi, dHm, p.Akmi, p.dt_mi))
            for w, dH, A, dt in pairs:
                wk = box_fft(w, list(sys.Nx), 1)
                fk = box_fft(dH, list(sys.Nx), 1)
                L.orc_fts_update_1s(_ptr(wk), _ptr(fk), _ptr(np.ascontiguousarray(A, np.float64)), dt, M)
                w[:] = box_fft(wk, list(sys.Nx), -1)
            continue
        if p.scheme == "EMPC":
            p.saved = dict(wpl=p.wpl.copy(), wmi=None if p.wmi is None else p.wmi.copy(), dHp=dHp, dHm=dHm,
                           npl=npl, nmi=nmi)
        L.orc_fts_update_em(_ptr(p.wpl), _ptr(dHp), _ptr(npl), p.dt_pl, M)
        if p.kind == "Flory":
            L.orc_fts_update_em(_ptr(p.wmi), _ptr(dHm), _ptr(nmi), p.dt_mi, M)
    _zero_means(sys)


def _zero_means(sys: FtsSystem, only_empc=False):
    """modify zeroMean: w+ minus its grid mean after an update or an EMPC
    correction (fts_potential_helfand.cu:140-146,198-204,
    fts_potential_incompress.cu:125-131,177-183, fts_potential_edwards.cu:
    146-152,202-208); Flory only parses the keyword."""
    for p in sys.potentials:
        if only_empc and p.scheme != "EMPC":
            continue
        if p.zero_mean and p.kind != "Flory":
            p.wpl -= p.wpl.sum() / p.wpl.size


def fts_correct_fields(sys: FtsSystem):
    """Every EMPC potential's correctFields (fts_potential_flory.cu:198-236,
    and likewise for Helfand and Incompress): forces at the predicted fields
    and densities, then w = w_o - dt/2 (F + F_o) + the predictor's noise."""
    M = sys.M
    L = lib()
    for p in sys.potentials:
        if p.scheme != "EMPC":
            continue
        dHp, dHm = _forces(sys, p)
        sv = p.saved
        L.orc_fts_update_empc(_ptr(p.wpl), _ptr(sv["wpl"]), _ptr(dHp), _ptr(sv["dHp"]), _ptr(sv["npl"]), p.dt_pl, M)
        if p.kind == "Flory":
            L.orc_fts_update_empc(_ptr(p.wmi), _ptr(sv["wmi"]), _ptr(dHm), _ptr(sv["dHm"]), _ptr(sv["nmi"]),
                                  p.dt_mi, M)
    _zero_means(sys, only_empc=True)


def fts_hamiltonian(sys: FtsSystem):
    """FTS_Box::computeHamiltonian (FTS_Box.cu:590-602) with the potential terms
    of fts_potential_flory.cu:252-274, fts_potential_helfand.cu:254-275,
    fts_potential_incompress.cu:232-247 and LinearMolec::calcHTerm
    (fts_molecule_linear.cu:368-382); also returns the parts."""
    gvol = sys.gvol()
    H = 0j
    parts = []
    for p in sys.potentials:
        if p.kind == "Flory":
            h = (np.sum(p.wpl * p.wpl) * gvol) * sys.C / p.chiN + (np.sum(p.wmi * p.wmi) * gvol) * sys.C / p.chiN
        elif p.kind == "Helfand":
            h = (np.sum(p.wpl * p.wpl) * gvol) * sys.C / 2.0 / p.kappaN - 1j * sys.C * (np.sum(p.wpl) * gvol)
        elif p.kind == "Edwards":
            # PotentialEdwards::calcHamiltonian (fts_potential_edwards.cu:256-270)
            h = (np.sum(p.wpl * p.wpl) * gvol) / 2.0 / p.kappaN
        else:
            h = -1j * sys.C * (np.sum(p.wpl) * gvol)
        parts.append(h)
        H += h
    for m, Q in zip(sys.molecules, sys.Q):
        # LinearMolec::calcHTerm: -n log Q, or -z V Q when grand canonical
        h = -m.nmolecs * np.log(Q) if m.nmolecs > 0.0 else -m.activity * sys.V() * Q
        parts.append(h)
        H += h
    return H, parts


def fts_chemical_potentials(sys: FtsSystem):
    """Chemical potential of every molecule: ln(n / (V Q)) when canonical, the
    ideal-chain form of the density factor n/(Q V) (fts_molecule_linear.cu:
    286-291), ln z when grand canonical (factor z).  The reference has no
    chemical-potential output; this is the definition the engine follows."""
    out = []
    for m, Q in zip(sys.molecules, sys.Q):
        if m.nmolecs > 0.0:
            out.append(np.log(complex(m.nmolecs / sys.V())) - np.log(complex(Q)))
        else:
            out.append(np.log(complex(m.activity)))
    return out


def fts_init(sys: FtsSystem, bits=64):
    """FTS_Box::finishInitialization density pass (FTS_Box.cu:193-256)."""
    sys.init_nmolecs()
    fts_build_species_fields(sys)
    fts_molecule_densities(sys, bits)


def fts_step(sys: FtsSystem, z=None, bits=64):
    """One CL iteration (FTS_Box.cu:19-81): update fields, rebuild species
    fields, recompute densities; with a predictor-corrector potential
    (PCflag), correct the fields and rebuild species fields and densities
    again."""
    fts_update_fields(sys, z)
    fts_build_species_fields(sys)
    fts_molecule_densities(sys, bits)
    if any(p.scheme == "EMPC" for p in sys.potentials):
        fts_correct_fields(sys)
        fts_build_species_fields(sys)
        fts_molecule_densities(sys, bits)


def fts_system_from_spec(spec, fields):
    """FtsSystem for a CL-FTS spec of mft.systems (fields as complex128 copies)."""
    M = int(np.prod(spec["Nx"]))
    pots = []
    for k, p in enumerate(spec["potentials"]):
        fp = FtsPotential(kind=p["kind"], I=p.get("I", -1), J=p.get("J", -1), chiN=p.get("chiN", 0.0),
                          kappaN=p.get("kappaN", p.get("B", 0.0)), dt_pl=p.get("dt_pl", 0.0), dt_mi=p.get("dt_mi", 0.0),
                          scheme=p.get("scheme", "EM"), zero_mean=bool(p.get("zero_mean", False)))
        fp.wpl = np.array(fields.get((k, "pl"), np.zeros(M)), np.complex128)
        if p["kind"] == "Flory":
            fp.wmi = np.array(fields.get((k, "mi"), np.zeros(M)), np.complex128)
        pots.append(fp)
    mols = [FtsMolecule(blocks=list(m["blocks"]), phi=m.get("phi", 1.0), activity=m.get("activity", -1.0) or -1.0)
            for m in spec["molecules"]]
    return FtsSystem(Nx=tuple(spec["Nx"]), L=tuple(spec["L"]), Nr=spec["Nr"], C=spec["C"], nspecies=spec["nspecies"],
                     potentials=pots, molecules=mols)
