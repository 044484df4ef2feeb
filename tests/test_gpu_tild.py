"""GPU parity of the TILD path: the HIP engine (through the C ABI) against the
FP64 CPU oracle on the same seeded inputs.

Tolerances (TILD, FP32 on the device): grid indices and topology bit-exact;
densities, grid forces and particle forces within 1e-5 normwise relative
error, rel(a, b) = max|a - b| / max|b|; positions after a step within 1e-6 of
the box length (periodic distance); energies within 1e-5 relative.
"""
import os
import sys

import numpy as np
import pytest

import orc
from mft import systems
from test_oracle_tild import oracle_system, small_system

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    """max|a - b| / max|b| over real or complex arrays."""
    a = np.asarray(a)
    b = np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = a.astype(dt)
    b = b.astype(dt)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def say(*a):
    print(*a, flush=True)


def test_philox_bit_exact(gpu):
    say("case: philox raw 4096 blocks")
    got = gpu.philox_raw(4096, (7 << 32) | 12345, 3, orc.TAG_TILD, 9)
    want = orc.philox_np(np.arange(4096, dtype=np.uint64), 3, orc.TAG_TILD, 9, (7 << 32) | 12345)
    for j in range(4):
        assert np.array_equal(got[:, j], want[j])


def test_normals_match_oracle(gpu):
    say("case: philox normals")
    got = gpu.philox_normal(20000, 99, 5, orc.TAG_FTS, 0)
    want = orc.normal4_np(np.arange(20000, dtype=np.uint64), 5, orc.TAG_FTS, 0, 99)
    assert np.max(np.abs(got - want)) < 2e-5


@pytest.mark.parametrize("shp", ["4x6x8", "8x12x16", "16x8x24"])
def test_fft_golden(gpu, shp):
    say("case: fft golden", shp)
    G = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))
    x = G[f"x3_{shp}"]
    M = x.size
    fwd = gpu.fft3d(x, 1)  # forward scaled by 1/M
    assert rel(fwd * M, G[f"fwd3_{shp}"]) < 2e-6
    inv = gpu.fft3d(x, -1)
    assert rel(inv, G[f"inv3_{shp}"]) < 2e-6


@pytest.mark.parametrize("n", [12, 20, 24, 32, 40, 48, 64, 80, 96, 128, 160, 192, 256])
def test_fft_sizes_against_oracle(gpu, n):
    say("case: fft size", n)
    rng = np.random.default_rng(n)
    shp = (8, n, 8) if n > 64 else (n, 8, n)   # (nz, ny, nx)
    x = rng.normal(size=shp) + 1j * rng.normal(size=shp)
    want = orc.fft3d(x, -1) / x.size
    got = gpu.fft3d(x, 1)
    assert rel(got, want) < 2e-6
    back = gpu.fft3d(got, -1)
    assert rel(back, x) < 2e-6


def test_fft_batched(gpu):
    say("case: fft batch of 3 grids 24x32x16")
    rng = np.random.default_rng(2)
    x = rng.normal(size=(3, 16, 32, 24)) + 1j * rng.normal(size=(3, 16, 32, 24))
    got = gpu.fft3d(x, 1)
    for b in range(3):
        assert rel(got[b], orc.fft3d(x[b], -1) / x[b].size) < 2e-6


def edge_positions(L, Nx, rng, n):
    L = np.asarray(L, np.float32)
    dx = (L / np.asarray(Nx, np.float32)).astype(np.float32)
    x = (rng.random((n, 3)) * L).astype(np.float32)
    x[0] = 0.0
    x[1] = L                                            # x = L exactly
    x[2] = dx * np.array([3, 5, 7], np.float32)         # grid nodes
    x[3] = (dx * np.array([3.5, 5.5, 7.5])).astype(np.float32)   # cell midpoints
    x[4] = np.nextafter(L, np.float32(0))               # just below L
    x[5] = np.nextafter(dx * 2, np.float32(0))           # just below a node
    x[6] = np.nextafter(dx * 2, np.float32(10))          # just above a node
    x[7] = np.float32(1e-7)
    return x.reshape(-1), dx


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4])
def test_grid_weights_indices_bit_exact(gpu, p):
    say("case: grid weights p =", p)
    rng = np.random.default_rng(10 + p)
    Nx, L = (24, 16, 20), (7.3, 5.1, 6.7)
    x, dx = edge_positions(L, Nx, rng, 20000)
    inds_g, W_g = gpu.grid_weights(x, Nx, dx, p)
    inds_o, W_o = orc.grid_weights(x, Nx, dx, p)
    assert np.array_equal(inds_g, inds_o)
    assert rel(W_g, W_o) < TOL


def test_table_kernels_match_oracle(gpu):
    say("case: fill_density / map_forces tables")
    rng = np.random.default_rng(4)
    Nx, L = (16, 24, 12), (5.0, 7.0, 4.0)
    x, dx = edge_positions(L, Nx, rng, 3000)
    inds, W = orc.grid_weights(x, Nx, dx, 3)
    M = int(np.prod(Nx))
    sites = np.arange(0, 3000, 3, dtype=np.int32)
    q = rng.choice([-1.0, 1.0], 3000).astype(np.float32)
    rho_o, rhoq_o = orc.fill_density(M, sites, inds, W, 64, q=q)
    rho_g, rhoq_g = gpu.fill_density(M, sites, inds, W.astype(np.float32), 64, q=q)
    assert rel(rho_g, rho_o) < TOL
    assert rel(rhoq_g, rhoq_o) < TOL
    F = rng.normal(size=3 * M)
    gvol = float(np.float32(np.float32(dx[0] * dx[1]) * dx[2]))
    f_o = np.zeros(9000)
    orc.map_forces(f_o, F, sites, inds, W, 64, gvol, 3, q=q)
    f_g = gpu.map_forces(3000, F.astype(np.float32), sites, inds, W.astype(np.float32), 64, gvol, q=q)
    assert rel(f_g, f_o) < TOL


def test_bonds_match_oracle(gpu):
    say("case: bonds")
    cfg = small_system(False)
    topo = cfg["topo"]
    f_g = gpu.bonds(cfg["x"], topo, [1.5, 2.0], [0.0, 0.3], [0, 1], cfg["L"])
    f_o = np.zeros(cfg["x"].size)
    orc.bonds(f_o, dict(nBonds=topo["nBonds"], bondedTo=topo["bondedTo"], bondType=topo["bondType"],
                        bondReq=[0.0, 0.3], bondK=[1.5, 2.0], bondStyle=[0, 1], maxbonds=2), cfg["x"], cfg["L"])
    assert rel(f_g, f_o) < TOL


def _engine_vs_oracle(gpu, cfg, integrator="GJF"):
    eng = systems.tild_engine(cfg, integrator)
    sys_o = oracle_system(cfg, integrator)
    return eng, sys_o


def mixed_charge_system():
    """Charged system whose species do not carry one charge each (every other
    bead of the polycation/polyanion chains is neutral): the engine keeps a
    separate charge-density channel."""
    cfg = small_system(True, seed=6)
    q = cfg["topo"]["charge"].copy()
    q[(np.arange(q.size) % 2 == 1) & (q != 0)] = 0.0
    cfg["topo"]["charge"] = q
    return cfg


@pytest.mark.parametrize("case", ["neutral", "charged", "mixed_charges"])
def test_engine_forces_and_densities(gpu, case):
    cfg = mixed_charge_system() if case == "mixed_charges" else small_system(case == "charged")
    charged = case != "neutral"
    say("case: engine forces", case, "N =", cfg["topo"]["type"].size)
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    f_g = eng.compute_forces()
    out = orc.tild_forces(sys_o)
    rho_g, rhoq_g = eng.densities(charge=charged)
    for s, name in enumerate(cfg["species"]):
        assert rel(rho_g[s], out["rho"][name]) < TOL, name
    if charged:
        assert rel(rhoq_g, out["rhoq"]) < TOL
    # grid force of each species group (ps_groups.cu:164-178); when every
    # species has one charge q_s the engine folds q_s times the charges group's
    # field into the species field
    M = int(np.prod(cfg["Nx"]))
    q = cfg["topo"]["charge"]
    t = cfg["topo"]["type"]
    folded = charged and all(np.unique(q[t == s]).size <= 1 for s in range(len(cfg["species"])))
    for s, name in enumerate(cfg["species"]):
        want = np.array(out["gridF"].get(name, np.zeros(3 * M)))
        if folded:
            qs = float(q[t == s][0]) if np.any(t == s) else 0.0
            want = want + qs * out["gridF"]["charges"]
        if np.any(want != 0):
            assert rel(eng.grid_force(s), want) < TOL, name
    if charged and not folded:
        assert rel(eng.grid_force(len(cfg["species"])), out["gridF"]["charges"]) < TOL
    assert rel(f_g, out["f"]) < TOL
    eng.close()


def periodic_dist(a, b, L):
    d = (np.asarray(a, np.float64) - np.asarray(b, np.float64)).reshape(-1, 3)
    L = np.asarray(L, np.float64)
    return np.abs(d - L * np.round(d / L))


@pytest.mark.parametrize("integrator", ["GJF", "EM"])
@pytest.mark.parametrize("charged", [False, True])
def test_engine_steps_match_oracle(gpu, integrator, charged):
    cfg = small_system(charged, seed=11)
    say("case: engine steps", cfg["name"], integrator)
    eng, sys_o = _engine_vs_oracle(gpu, cfg, integrator)
    ns = cfg["topo"]["type"].size
    for step in range(3):
        eng.step(1)
        xi = orc.tild_noise(ns, cfg["seed"], step)
        out = orc.tild_step(sys_o, xi)
        sys_o.x = out["x"].astype(np.float32)
        if integrator == "GJF":
            sys_o.xo = out["xo"].astype(np.float32)
            sys_o.noise_old = out["noise_old"].astype(np.float32)
        x_g, xo_g, no_g = eng.get_state()
        assert np.max(periodic_dist(x_g, out["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
        if integrator == "GJF":
            assert rel(no_g, out["noise_old"]) < 1e-5
    eng.close()


@pytest.mark.parametrize("case", ["charged", "mixed_charges"])
def test_energies_match_oracle(gpu, case):
    cfg = small_system(True, seed=3) if case == "charged" else mixed_charge_system()
    say("case: energies", case)
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    pe, be = eng.energies()
    out = orc.tild_forces(sys_o, want_energy=True)
    for g, o in zip(pe, out["energies"]):
        assert abs(g - o) <= TOL * max(abs(o), 1e-12), (g, o)
    topo = sys_o.topo
    eb = orc.bond_energy(dict(nBonds=topo["nBonds"], bondedTo=topo["bondedTo"], bondType=topo["bondType"],
                              bondReq=cfg["bond_r0"], bondK=cfg["bond_k"], bondStyle=cfg["bond_style"], maxbonds=2),
                         sys_o.x, cfg["L"])
    assert be == pytest.approx(0.5 * eb.sum(), rel=TOL)
    # the same energies through a device buffer, without a host round trip
    pd, bd = eng.energies_device()
    assert np.array_equal(pd, np.asarray(pe)) and bd == be
    eng.close()


@pytest.mark.parametrize("case", ["erf_kinds", "many_shapes"])
def test_engine_potential_shapes(gpu, case):
    """erf2/erfG tables next to Gaussians (potentials grouped by k-space shape),
    and more distinct shapes than the kernel groups (each potential evaluated
    on its own)."""
    cfg = small_system(False, seed=8)
    A = cfg["potentials"][0]["A"]
    if case == "erf_kinds":
        cfg["potentials"] = [dict(kind="erf2", maskI=1, maskJ=1, A=A, sigma=0.4, Rp=0.6),
                             dict(kind="erfG", maskI=2, maskJ=2, A=A, sigma=0.5, Rp=0.5),
                             dict(kind="gaussian", maskI=1, maskJ=2, A=2 * A, sigma=0.6),
                             dict(kind="gaussian", maskI=1, maskJ=1, A=0.5 * A, sigma=0.6)]
    else:
        cfg["potentials"] = [dict(kind="gaussian", maskI=1 + (k % 2), maskJ=1 + (k // 2) % 2, A=A * (1 + 0.1 * k),
                                  sigma=0.4 + 0.05 * k) for k in range(10)]
    say("case: potential shapes", case)
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    f_g = eng.compute_forces()
    out = orc.tild_forces(sys_o, want_energy=True)
    assert rel(f_g, out["f"]) < TOL
    pe, _ = eng.energies()
    for g, o in zip(pe, out["energies"]):
        assert abs(g - o) <= TOL * max(abs(o), 1e-12), (g, o)
    eng.close()


def test_ramp_schedule(gpu):
    # PS_Potential::update_prefactor: step i uses Ao_i + (Ao_f - Ao_i) i / nsteps
    cfg = small_system(False, seed=2)
    say("case: ramp")
    pots = [dict(p) for p in cfg["potentials"]]
    pots[2].update(ramp=True, A_final=3 * pots[2]["A"])
    cfg_r = dict(cfg, potentials=pots)
    eng = systems.tild_engine(cfg_r, "GJF")
    eng.step(4)   # one NVT run of 4 steps
    x_g, _, _ = eng.get_state()
    sys_o = oracle_system(cfg)
    ns = cfg["topo"]["type"].size
    A0 = pots[2]["A"]
    for step in range(4):
        sys_o.potentials[2].A = A0 + (3 * A0 - A0) * step / 4.0
        out = orc.tild_step(sys_o, orc.tild_noise(ns, cfg["seed"], step))
        sys_o.x, sys_o.xo, sys_o.noise_old = (out[k].astype(np.float32) for k in ("x", "xo", "noise_old"))
    assert np.max(periodic_dist(x_g, sys_o.x, cfg["L"])) < 1e-5 * max(cfg["L"])
    eng.close()


def test_large_coacervate_invariants(gpu):
    """Config-2 shape (1.8 M particles, 128^3): size-independent properties --
    species mass conservation, zero net force, zero net charge density, and
    bitwise determinism of two identical runs."""
    cfg = systems.tild_coacervate()
    say("case: coacervate 1.8M invariants")
    eng = systems.tild_engine(cfg)
    f = eng.compute_forces().reshape(-1, 3).astype(np.float64)
    rho, rhoq = eng.densities(charge=True)
    gvol = np.prod(np.float32(np.asarray(cfg["L"], np.float32) / np.asarray(cfg["Nx"], np.float32)))
    counts = np.bincount(cfg["topo"]["type"], minlength=3)
    for s in range(3):
        assert rho[s].astype(np.float64).sum() * gvol == pytest.approx(counts[s], rel=1e-6)
    assert abs(rhoq.astype(np.float64).sum() * gvol) < 1e-3 * counts[0]
    fmax = np.max(np.abs(f))
    assert np.max(np.abs(f.sum(axis=0))) < 1e-3 * fmax * np.sqrt(f.shape[0])
    eng.step(3)
    x1, _, _ = eng.get_state()
    eng.close()
    eng2 = systems.tild_engine(cfg)
    eng2.step(3)
    x2, _, _ = eng2.get_state()
    eng2.close()
    assert np.array_equal(x1, x2), "two identical runs differ"


def test_config1_diblock_matches_oracle(gpu):
    """The plain diblock melt at its full size (2000 chains of 10 A + 10 B,
    40 000 particles, 20^3 box, 32^3 grid, cubic splines, GJF dt = 0.005):
    densities, grid forces, particle forces and energies, then three GJF
    steps, against the FP64 oracle on the same positions and noise."""
    cfg = systems.tild_diblock()
    say("case: diblock melt 40k particles, 32^3")
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    f_g = eng.compute_forces()
    out = orc.tild_forces(sys_o, want_energy=True)
    rho_g, _ = eng.densities()
    errs = {name: rel(rho_g[s], out["rho"][name]) for s, name in enumerate(cfg["species"])}
    for s, name in enumerate(cfg["species"]):
        errs[f"grid force {name}"] = rel(eng.grid_force(s), out["gridF"][name])
    errs["forces"] = rel(f_g, out["f"])
    pe, be = eng.energies()
    for k, (g, o) in enumerate(zip(pe, out["energies"])):
        errs[f"energy {k}"] = abs(g - o) / max(abs(o), 1e-12)
    say("  " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()))
    assert all(v < TOL for v in errs.values()), errs
    ns = cfg["topo"]["type"].size
    for step in range(3):
        eng.step(1)
        o = orc.tild_step(sys_o, orc.tild_noise(ns, cfg["seed"], step))
        sys_o.x, sys_o.xo, sys_o.noise_old = (o[k].astype(np.float32) for k in ("x", "xo", "noise_old"))
        x_g, _, no_g = eng.get_state()
        dx = np.max(periodic_dist(x_g, o["x"], cfg["L"]))
        say(f"  step {step}: max |dx| {dx:.2e}")
        assert dx < 1e-6 * max(cfg["L"]), step
        assert rel(no_g, o["noise_old"]) < TOL
    eng.close()


def edge_case_system(case):
    """Small diblock melts at the edges of the engine's layouts."""
    if case in ("p1", "p2", "p4"):
        return systems.tild_diblock(nchains=60, NA=5, NB=5, box=6.0, grid=16, pmeorder=int(case[1]), seed=31)
    cfg = systems.tild_diblock(nchains=60, NA=5, NB=5, box=6.0, grid=16, seed=32)
    if case == "ragged_grid":
        # grid lengths that are not multiples of the brick edge (short last
        # bricks) on a box that is not a cube
        cfg["Nx"], cfg["L"] = (20, 12, 24), (7.0, 4.5, 9.0)
        x = cfg["x"].reshape(-1, 3) % 6.0
        x = x * (np.asarray(cfg["L"]) / 6.0)
        cfg["x"] = x.astype(np.float32).reshape(-1)
    elif case == "empty_species":
        # a third species with potentials acting on it but no particles
        cfg["species"] = ["A", "B", "C"]
        cfg["mass"], cfg["mobility"] = [1.0] * 3, [1.0] * 3
        cfg["potentials"] = cfg["potentials"] + [dict(kind="gaussian", maskI=4, maskJ=1, A=1.5, sigma=0.5)]
    elif case == "edge_positions":
        # particles exactly on the lower faces and just below the upper faces
        x = cfg["x"].reshape(-1, 3).copy()
        L = np.float32(6.0)
        x[0::7, 0] = 0.0
        x[1::7, 1] = np.nextafter(L, np.float32(0.0))
        x[2::7, 2] = 0.0
        x[3::7, :] = np.nextafter(L, np.float32(0.0))
        cfg["x"] = x.astype(np.float32).reshape(-1)
    return cfg


@pytest.mark.parametrize("case", ["p1", "p2", "p4", "ragged_grid", "empty_species", "edge_positions"])
def test_engine_edge_cases(gpu, case):
    """Forces, densities and two GJF steps against the oracle for spline
    orders 1, 2, 4 (4 with the reference's defect), a grid that the bricks do
    not divide, a species with no particles, and particles on the box faces."""
    cfg = edge_case_system(case)
    say("case: engine edge case", case, "Nx", cfg["Nx"], "p", cfg["pmeorder"])
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    f_g = eng.compute_forces()
    out = orc.tild_forces(sys_o)
    rho_g, _ = eng.densities()
    for s, name in enumerate(cfg["species"]):
        want = out["rho"].get(name)
        if want is None or not np.any(want):
            assert not np.any(rho_g[s]), name
        else:
            assert rel(rho_g[s], want) < TOL, name
    assert rel(f_g, out["f"]) < TOL
    ns = cfg["topo"]["type"].size
    for step in range(2):
        eng.step(1)
        o = orc.tild_step(sys_o, orc.tild_noise(ns, cfg["seed"], step))
        sys_o.x = o["x"].astype(np.float32)
        sys_o.xo = o["xo"].astype(np.float32)
        sys_o.noise_old = o["noise_old"].astype(np.float32)
        x_g, _, _ = eng.get_state()
        assert np.max(periodic_dist(x_g, o["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
    eng.close()


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("case", ["neutral", "charged", "mixed_charges", "ragged_grid"])
def test_gather_variants_match_oracle(gpu, case, variant):
    """The gather variants (direct loads, LDS windows per brick and species,
    LDS windows of every species per brick, pipelined windows per brick and
    species) against the oracle's forces and a short trajectory."""
    if case == "mixed_charges":
        cfg = mixed_charge_system()
    elif case == "ragged_grid":
        cfg = edge_case_system("ragged_grid")
    else:
        cfg = small_system(case == "charged", seed=19)
    say("case: gather variant", variant, case)
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    eng.set_gather(variant)
    assert rel(eng.compute_forces(), orc.tild_forces(sys_o)["f"]) < TOL
    ns = cfg["topo"]["type"].size
    for step in range(2):
        eng.step(1)
        o = orc.tild_step(sys_o, orc.tild_noise(ns, cfg["seed"], step))
        sys_o.x, sys_o.xo = o["x"].astype(np.float32), o["xo"].astype(np.float32)
        sys_o.noise_old = o["noise_old"].astype(np.float32)
        assert np.max(periodic_dist(eng.get_state()[0], o["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
    eng.close()


@pytest.mark.parametrize("mask", [1, 2, 3])
def test_structure_factor_matches_oracle_fft(gpu, mask):
    # PS_ComputeSK::do_compute (ps_compute_sk.cu:18-45): |FFT[rho_group]|^2
    # with the 1/M forward scaling, against the oracle's FFT of the engine's
    # own densities; S(0) = (N_group / V)^2
    say("case: structure factor, mask", mask)
    cfg = small_system(False)
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    eng.compute_forces()
    rho, _ = eng.densities()
    grp = sum(rho[s].astype(np.float64) for s in range(len(cfg["species"])) if (mask >> s) & 1)
    want = np.abs(orc.box_fft(grp.astype(np.complex128), list(cfg["Nx"]), 1)) ** 2
    got = eng.structure_factor(mask)
    assert rel(got, want) < 1e-5
    t = np.asarray(cfg["topo"]["type"])
    n = sum(int((t == s).sum()) for s in range(len(cfg["species"])) if (mask >> s) & 1)
    V = float(np.prod(np.asarray(cfg["L"], np.float64)))
    assert abs(got[0] - (n / V) ** 2) < 1e-5 * (n / V) ** 2
    eng.close()


@pytest.mark.parametrize("phase, direction, with_pairs", [("L", 1, True), ("H", 2, True), ("G", 0, False)])
def test_bias_field_matches_oracle(gpu, phase, direction, with_pairs):
    # BiasField (ps_potentialBias.cu): static A u(r) on one species' particles,
    # beside the pair potentials or alone; forces, energies and a step
    cfg = small_system(False)
    bias = dict(kind="bias", maskI=1, maskJ=0, A=2.5, phase=phase, dir=direction, n_periods=1)
    cfg["potentials"] = (list(cfg["potentials"]) if with_pairs else []) + [bias]
    say("case: bias", phase, direction, "with pairs" if with_pairs else "alone")
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    f_g = eng.compute_forces()
    out = orc.tild_forces(sys_o, want_energy=True)
    assert rel(f_g, out["f"]) < TOL
    pe, _ = eng.energies()
    for g, o in zip(pe, out["energies"]):
        assert abs(g - o) <= TOL * max(abs(o), 1e-12), (g, o)
    eng.close()


def test_walls_match_oracle(gpu):
    # Wall (ps_potentialWall.cu): harmonic walls on one species along z at the
    # mid-plane (pushing up) and on the other along x (pushing down), beside
    # the pair potentials; forces, energies and three GJF steps
    cfg = small_system(False, seed=12)
    L = cfg["L"]
    cfg["potentials"] = list(cfg["potentials"]) + [
        dict(kind="wall", maskI=1, maskJ=0, A=3.0, dir=2, wall_pos=0.5 * L[2], wall_sign=1),
        dict(kind="wall", maskI=2, maskJ=0, A=1.5, dir=0, wall_pos=0.3 * L[0], wall_sign=-1)]
    say("case: walls")
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    f_g = eng.compute_forces()
    out = orc.tild_forces(sys_o, want_energy=True)
    assert rel(f_g, out["f"]) < TOL
    pe, _ = eng.energies()
    for g, o in zip(pe, out["energies"]):
        assert abs(g - o) <= TOL * max(abs(o), 1e-12), (g, o)
    # closed form of the first wall's energy: sum over A particles below the plane
    x = cfg["x"].reshape(-1, 3).astype(np.float64)
    a = np.asarray(cfg["topo"]["type"]) == 0
    d = np.float32(0.5 * L[2]) - x[a, 2]
    assert pe[-2] == pytest.approx(np.sum(0.5 * 3.0 * d[d > 0] ** 2), rel=1e-5)
    ns = cfg["topo"]["type"].size
    for step in range(3):
        eng.step(1)
        xi = orc.tild_noise(ns, cfg["seed"], step)
        o = orc.tild_step(sys_o, xi)
        sys_o.x = o["x"].astype(np.float32)
        sys_o.xo = o["xo"].astype(np.float32)
        sys_o.noise_old = o["noise_old"].astype(np.float32)
        x_g, _, _ = eng.get_state()
        assert np.max(periodic_dist(x_g, o["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
    eng.close()


def test_angles_match_oracle(gpu):
    # d_angles / d_anglesStressEnergy (ps_deviceAngles.cu): WLC and harmonic
    # angles on the diblock chains beside the pair potentials; forces, the
    # angle energy and five GJF steps (each re-sorts the slots, so the angle
    # lists are permuted and remapped every step)
    from test_oracle_tild import angle_system
    cfg = angle_system()
    say("case: angles")
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    f_g = eng.compute_forces()
    out = orc.tild_forces(sys_o)
    assert rel(f_g, out["f"]) < TOL
    ea = eng.angle_energy()
    eo = orc.angle_energy(cfg["angles"], sys_o.x, cfg["L"])
    assert ea == pytest.approx(eo, rel=1e-5)
    ns = cfg["topo"]["type"].size
    for step in range(5):
        eng.step(1)
        xi = orc.tild_noise(ns, cfg["seed"], step)
        o = orc.tild_step(sys_o, xi)
        sys_o.x = o["x"].astype(np.float32)
        sys_o.xo = o["xo"].astype(np.float32)
        sys_o.noise_old = o["noise_old"].astype(np.float32)
        x_g, _, _ = eng.get_state()
        assert np.max(periodic_dist(x_g, o["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
    assert eng.angle_energy() == pytest.approx(orc.angle_energy(cfg["angles"], sys_o.x, cfg["L"]), rel=1e-5)
    eng.close()


def test_velocity_verlet_matches_oracle(gpu):
    # VV (ps_integratorVV.cu) from rest: the engine fuses each step's second
    # half kick into the next step, so its velocities between steps are the
    # oracle's less (dt/2) f / m; positions agree step by step
    cfg = small_system(False, seed=14)
    say("case: velocity Verlet")
    eng = systems.tild_engine(cfg, "VV")
    sys_o = oracle_system(cfg, "VV")
    t = cfg["topo"]["type"]
    inv_m = np.array([float(np.float32(1.0) / np.float32(m)) for m in cfg["mass"]])[t][:, None]
    hd = float(np.float32(cfg["delt"]) * np.float32(0.5))
    x_first = None
    for step in range(6):
        eng.step(1)
        o = orc.tild_step(sys_o)
        sys_o.x = o["x"].astype(np.float32)
        sys_o.v, sys_o.f_prev = o["v"], o["f_prev"]
        x_g, _, v_g = eng.get_state()
        assert np.max(periodic_dist(x_g, o["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
        v_o = o["v"].reshape(-1, 3) - hd * inv_m * o["f_prev"].reshape(-1, 3)
        if step == 0:
            assert np.all(v_g == 0.0)
        else:
            assert rel(v_g.reshape(-1, 3), v_o) < 1e-4, step
        if step == 0:
            x_first = x_g
    # the first step from rest does not move the particles
    assert np.array_equal(x_first, cfg["x"].astype(np.float32).reshape(x_first.shape))
    eng.close()


@pytest.mark.parametrize("integrator", ["VV", "GJF"])
def test_langevin_matches_oracle(gpu, integrator):
    # Langevin (ps_potentialLangevin.cu) on species A beside the pair
    # potentials: under VV the drag acts on the velocities, under GJF the
    # reference's velocities stay zero and only the noise acts
    cfg = small_system(False, seed=15)
    cfg["potentials"] = list(cfg["potentials"]) + [dict(kind="langevin", maskI=1, maskJ=0, A=0.8, sigma=0.005)]
    say("case: langevin", integrator)
    eng = systems.tild_engine(cfg, integrator)
    sys_o = oracle_system(cfg, integrator)
    f_g = eng.compute_forces()
    assert rel(f_g, orc.tild_forces(sys_o)["f"]) < TOL
    ns = cfg["topo"]["type"].size
    for step in range(4):
        eng.step(1)
        sys_o.step = step
        if integrator == "VV":
            o = orc.tild_step(sys_o)
            sys_o.v, sys_o.f_prev = o["v"], o["f_prev"]
        else:
            o = orc.tild_step(sys_o, orc.tild_noise(ns, cfg["seed"], step))
            sys_o.xo = o["xo"].astype(np.float32)
            sys_o.noise_old = o["noise_old"].astype(np.float32)
        sys_o.x = o["x"].astype(np.float32)
        x_g, _, _ = eng.get_state()
        assert np.max(periodic_dist(x_g, o["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
    pe, _ = eng.energies()
    assert pe[-1] == 0.0
    eng.close()


@pytest.mark.parametrize("integrator", ["VV", "GJF"])
def test_dpd_matches_oracle(gpu, integrator):
    # DPD pair forces (ps_potentialDPD.cu) on species A through the cell
    # search (rcut 1.0 in a 6^3 box: 6 cells per axis) beside the pair
    # potentials; forces and four steps against the oracle's all-pairs sum
    cfg = small_system(False, seed=16)
    cfg["potentials"] = list(cfg["potentials"]) + [
        dict(kind="dpd", maskI=1, maskJ=0, A=4.5, sigma=1.0, Rp=1.0, wall_pos=cfg["delt"])]
    say("case: dpd", integrator)
    eng = systems.tild_engine(cfg, integrator)
    sys_o = oracle_system(cfg, integrator)
    f_g = eng.compute_forces()
    assert rel(f_g, orc.tild_forces(sys_o)["f"]) < TOL
    ns = cfg["topo"]["type"].size
    for step in range(4):
        eng.step(1)
        sys_o.step = step
        if integrator == "VV":
            o = orc.tild_step(sys_o)
            sys_o.v, sys_o.f_prev = o["v"], o["f_prev"]
        else:
            o = orc.tild_step(sys_o, orc.tild_noise(ns, cfg["seed"], step))
            sys_o.xo = o["xo"].astype(np.float32)
            sys_o.noise_old = o["noise_old"].astype(np.float32)
        sys_o.x = o["x"].astype(np.float32)
        x_g, _, _ = eng.get_state()
        assert np.max(periodic_dist(x_g, o["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
    eng.close()


@pytest.mark.parametrize("grid, nchains, nsolvent, box", [(64, 300, 3000, 14.0), (96, 400, 4000, 16.0),
                                                         (128, 500, 5000, 18.0)])
def test_plane_passes_equal_line_passes(gpu, grid, nchains, nsolvent, box):
    """The x-y transforms of each z-plane in one workgroup run the arithmetic
    of one pass per axis (same plans, twiddles and tile sums): the densities
    are the same bits, the forces agree to float rounding (the compiler may
    contract different multiply-adds in the two kernels; at 64 and 128 points
    they are the same bits) and both match the oracle's forces; so do three
    GJF steps (whose previous noise is regenerated from Philox after the
    first)."""
    say("case: plane passes, grid", grid)
    cfg = systems.tild_coacervate(nchains=nchains, Nchain=10, nsolvent=nsolvent, box=box, grid=grid, seed=23)
    engs = []
    for variant in (0, 1):
        eng = systems.tild_engine(cfg, "GJF")
        eng.set_grid_passes(variant)
        engs.append(eng)
    f0, f1 = engs[0].compute_forces(), engs[1].compute_forces()
    r0, r1 = engs[0].densities()[0], engs[1].densities()[0]
    assert np.array_equal(r0, r1)
    assert rel(f1, f0) < 1e-6
    if grid in (64, 128):
        assert np.array_equal(f0, f1)
    sys_o = oracle_system(cfg, "GJF")
    assert rel(f1, orc.tild_forces(sys_o)["f"]) < TOL
    for eng in engs:
        eng.step(3)
    s0, s1 = engs[0].get_state(), engs[1].get_state()
    assert np.max(periodic_dist(s0[0], s1[0], cfg["L"])) < 1e-6 * max(cfg["L"])
    assert rel(s1[2], s0[2]) == 0.0          # regenerated noise: the same normals
    for eng in engs:
        eng.close()


def test_gjf_previous_noise_is_regenerated_bit_for_bit(gpu):
    """After GJF steps the engine regenerates the previous noise from
    (particle, step - 1) instead of carrying it: get_state's noise_old and
    every position equal the oracle's trajectory that carries it."""
    say("case: gjf regenerated noise")
    cfg = small_system(True, seed=31)
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    ns = cfg["topo"]["type"].size
    for step in range(4):
        eng.step(1)
        o = orc.tild_step(sys_o, orc.tild_noise(ns, cfg["seed"], step))
        sys_o.x, sys_o.xo = o["x"].astype(np.float32), o["xo"].astype(np.float32)
        sys_o.noise_old = o["noise_old"].astype(np.float32)
        x, xo, no = eng.get_state()
        assert np.max(periodic_dist(x, o["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
        assert rel(no, o["noise_old"]) < 1e-6, step
    # a new state restarts from the given previous noise (the oracle from the
    # same state, so one step's error is compared)
    x, xo, no = eng.get_state()
    eng.set_state(x, xo, no)
    sys_o.x, sys_o.xo, sys_o.noise_old = x.copy(), xo.copy(), no.copy()
    eng.step(1)
    o = orc.tild_step(sys_o, orc.tild_noise(ns, cfg["seed"], 4))
    assert np.max(periodic_dist(eng.get_state()[0], o["x"], cfg["L"])) < 1e-6 * max(cfg["L"])
    eng.close()


def test_bond_virial_matches_oracle(gpu):
    """Per-particle bond virial terms of d_bondStressEnergy
    (ps_deviceBonds.cu:85-171) after two GJF steps (particles re-keyed), in
    particle-id order, against the oracle at the engine's positions; bonds
    longer than L/4 wrap against 0.5*Lh as in the reference."""
    cfg = small_system(True, seed=41)
    say("case: bond virial")
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    eng.step(2)
    x, _, _ = eng.get_state()
    v = eng.bond_virial()
    topo = sys_o.topo
    vo = orc.bond_virial(dict(nBonds=topo["nBonds"], bondedTo=topo["bondedTo"], bondType=topo["bondType"],
                              bondReq=cfg["bond_r0"], bondK=cfg["bond_k"], bondStyle=cfg["bond_style"], maxbonds=2),
                         x, cfg["L"])
    assert v.shape == vo.shape
    assert rel(v, vo) < TOL
    # summed over particles: twice the pair virial, trace = -sum mf r^2 < 0 for stretched harmonic bonds
    assert np.sum(v[:, :3], dtype=np.float64) == pytest.approx(np.sum(vo[:, :3]), rel=TOL)
    eng.close()


def test_two_bond_types_match_oracle(gpu):
    """Bond types carried through the key order: a diblock whose A and B
    blocks have their own harmonic bond (k, r0), so the permutation must move
    every slot's types with it.  Three GJF steps, the bond energy and the bond
    virial after the re-keying, against the oracle.  (With one bond type the
    slots' types are not carried: every bond is type 0.)"""
    cfg = systems.tild_diblock(nchains=60, NA=5, NB=5, box=6.0, grid=16, seed=21)
    cfg["topo"] = systems.make_linear(60, [(5, 0), (5, 1)], block_bond_type=[1, 2])
    cfg["bond_k"], cfg["bond_r0"], cfg["bond_style"] = [1.5, 2.5], [0.0, 0.4], [0, 0]
    assert set(np.unique(cfg["topo"]["bondType"][cfg["topo"]["nBonds"].repeat(2) > 0])) == {0, 1}
    say("case: two bond types")
    eng, sys_o = _engine_vs_oracle(gpu, cfg)
    ns = cfg["topo"]["type"].size
    for step in range(3):
        eng.step(1)
        xi = orc.tild_noise(ns, cfg["seed"], step)
        out = orc.tild_step(sys_o, xi)
        sys_o.x = out["x"].astype(np.float32)
        sys_o.xo = out["xo"].astype(np.float32)
        sys_o.noise_old = out["noise_old"].astype(np.float32)
        x_g, _, _ = eng.get_state()
        assert np.max(periodic_dist(x_g, out["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
    topo = sys_o.topo
    bonds = dict(nBonds=topo["nBonds"], bondedTo=topo["bondedTo"], bondType=topo["bondType"],
                 bondReq=cfg["bond_r0"], bondK=cfg["bond_k"], bondStyle=cfg["bond_style"], maxbonds=2)
    x, _, _ = eng.get_state()
    _, be = eng.energies()
    eb = orc.bond_energy(bonds, x, cfg["L"])
    assert be == pytest.approx(0.5 * eb.sum(), rel=TOL)
    v = eng.bond_virial()
    assert rel(v, orc.bond_virial(bonds, x, cfg["L"])) < TOL
    eng.close()


def test_three_bond_slots_match_oracle(gpu):
    """MAXBONDS = 3 (the general bond path: each slot's partner list ends at
    its first -1 entry; the two-bond path is specialised): the diblock's bonds
    with an empty third slot per particle, against the oracle's two-slot
    layout of the same bonds.  Three GJF steps (re-keyed every step), then the
    bond energy and virial."""
    from mft import TildEngine
    cfg = small_system(False, seed=23)
    topo = cfg["topo"]
    ns = topo["type"].size
    bto = np.zeros((ns, 3), np.int32)
    bty = np.zeros((ns, 3), np.int32)
    bto[:, :2] = topo["bondedTo"].reshape(ns, 2)
    bty[:, :2] = topo["bondType"].reshape(ns, 2)
    say("case: three bond slots")
    eng = TildEngine(cfg["Nx"], cfg["L"], cfg["pmeorder"], len(cfg["species"]), ns, cfg["potentials"],
                     cfg["bond_k"], cfg["bond_r0"], cfg["bond_style"], integrator="GJF", delt=cfg["delt"],
                     mass=cfg["mass"], mobility=cfg["mobility"], seed=cfg["seed"], has_charges=False, maxbonds=3)
    eng.set_topology(topo["type"], None, topo["nBonds"], bto.reshape(-1), bty.reshape(-1))
    eng.set_state(cfg["x"])
    sys_o = oracle_system(cfg)
    for step in range(3):
        eng.step(1)
        out = orc.tild_step(sys_o, orc.tild_noise(ns, cfg["seed"], step))
        sys_o.x = out["x"].astype(np.float32)
        sys_o.xo = out["xo"].astype(np.float32)
        sys_o.noise_old = out["noise_old"].astype(np.float32)
        x_g, _, _ = eng.get_state()
        assert np.max(periodic_dist(x_g, out["x"], cfg["L"])) < 1e-6 * max(cfg["L"]), step
    bonds = dict(nBonds=topo["nBonds"], bondedTo=topo["bondedTo"], bondType=topo["bondType"],
                 bondReq=cfg["bond_r0"], bondK=cfg["bond_k"], bondStyle=cfg["bond_style"], maxbonds=2)
    x, _, _ = eng.get_state()
    _, be = eng.energies()
    assert be == pytest.approx(0.5 * orc.bond_energy(bonds, x, cfg["L"]).sum(), rel=TOL)
    assert rel(eng.bond_virial(), orc.bond_virial(bonds, x, cfg["L"])) < TOL
    eng.close()


@pytest.mark.parametrize("integrator", ["GJF", "EM"])
def test_replica_group_graph_equals_engines_alone(gpu, integrator):
    """Replicas advanced by the group's HIP graph (eight steps of every engine
    per replay) give the same bits as the same engines stepped alone: three
    uneven systems, 1 + 17 + 3 + 16 steps (ordinary steps where a replay does
    not fit, several replays with ordinary steps and host reads between
    them), every replica checked after each call."""
    say("case: replica group", integrator)
    cfgs = [small_system(True, seed=51), small_system(False, seed=52), mixed_charge_system()]
    alone = [systems.tild_engine(c, integrator) for c in cfgs]
    grouped = [systems.tild_engine(c, integrator) for c in cfgs]
    grp = gpu.TildGroup(grouped)
    for n in (1, 17, 3, 16):
        for e in alone:
            e.step(n)
        grp.step(n)
        grp.synchronize()
        for a, b in zip(alone, grouped):
            sa, sb = a.get_state(), b.get_state()
            for u, v in zip(sa, sb):
                assert np.array_equal(u, v), n
    # the oracle agrees with the grouped trajectory's next step
    cfg = cfgs[0]
    sys_o = oracle_system(cfg, integrator)
    x, xo, no = grouped[0].get_state()
    sys_o.x, sys_o.xo, sys_o.noise_old = x.copy(), xo.copy(), no.copy()
    grp.step(2)
    grp.synchronize()
    ns = cfg["topo"]["type"].size
    for st in (37, 38):
        o = orc.tild_step(sys_o, orc.tild_noise(ns, cfg["seed"], st))
        sys_o.x = o["x"].astype(np.float32)
        if "xo" in o:
            sys_o.xo, sys_o.noise_old = o["xo"].astype(np.float32), o["noise_old"].astype(np.float32)
    assert np.max(periodic_dist(grouped[0].get_state()[0], o["x"], cfg["L"])) < 2e-6 * max(cfg["L"])
    grp.close()
    for e in alone + grouped:
        e.close()


@pytest.mark.parametrize("grid, nchains, nsolvent, box", [(32, 200, 2000, 10.0), (64, 300, 3000, 14.0),
                                                         (128, 500, 5000, 18.0)])
def test_staged_x_pass_equals_direct(gpu, grid, nchains, nsolvent, box):
    """The forward x pass with its tile rows staged in LDS sums the same
    entries in the same order as reading them directly: densities, forces and
    three GJF steps are the same bits, and the forces match the oracle."""
    say("case: staged x pass, grid", grid)
    cfg = systems.tild_coacervate(nchains=nchains, Nchain=10, nsolvent=nsolvent, box=box, grid=grid, seed=29)
    engs = []
    for variant in (0, 2):
        eng = systems.tild_engine(cfg, "GJF")
        eng.set_grid_passes(variant)
        engs.append(eng)
    f0, f1 = engs[0].compute_forces(), engs[1].compute_forces()
    assert np.array_equal(engs[0].densities()[0], engs[1].densities()[0])
    assert np.array_equal(f0, f1)
    sys_o = oracle_system(cfg, "GJF")
    assert rel(f1, orc.tild_forces(sys_o)["f"]) < TOL
    for eng in engs:
        eng.step(3)
    for a, b in zip(engs[0].get_state(), engs[1].get_state()):
        assert np.array_equal(a, b)
    for eng in engs:
        eng.close()
