"""Pins the TILD oracle against closed forms, invariants and brute force.
CPU only."""
import numpy as np
import pytest

import orc
from mft import systems
from test_oracle_core import bspline


def brute_force_stencil(x, L, N, p):
    """Grid points (index, weight) of one coordinate by direct enumeration: the
    weight of grid point k (at k*dx) is the cardinal B-spline of order p+1
    evaluated at its distance from the particle in units of dx."""
    dx = L / N
    out = {}
    for k in range(-N, 2 * N):
        w = bspline((k * dx - x) / dx, p + 1)
        if abs(w) > 0:
            out[k % N] = out.get(k % N, 0.0) + w
    return out


@pytest.mark.parametrize("p", [1, 2, 3])
def test_grid_weights_brute_force(p):
    rng = np.random.default_rng(p)
    Nx = (8, 6, 10)
    L = (4.0, 3.3, 5.1)
    dx = [np.float32(np.float32(l) / n) for l, n in zip(L, Nx)]
    x = (rng.random((40, 3)) * np.asarray(L)).astype(np.float32)
    x[0] = [0.0, 0.0, 0.0]                       # particle on a corner
    x[1] = [np.float32(L[0]), 1.0, 2.0]          # x = L exactly (GJF wraps to [0, L])
    x[2] = [dx[0] * 3, dx[1] * 2, dx[2] * 5]     # on grid nodes
    inds, W = orc.grid_weights(x.reshape(-1), Nx, dx, p)
    G = (p + 1) ** 3
    gvol = float(np.float32(np.float32(dx[0] * dx[1]) * dx[2]))
    for i in range(x.shape[0]):
        got = {}
        for j in range(G):
            got[inds[i * G + j]] = got.get(inds[i * G + j], 0.0) + W[i * G + j] * gvol
        per_axis = [brute_force_stencil(float(x[i, a]), float(np.float32(L[a])), Nx[a], p) for a in range(3)]
        want = {}
        for kx, wx in per_axis[0].items():
            for ky, wy in per_axis[1].items():
                for kz, wz in per_axis[2].items():
                    idx = kx + Nx[0] * (ky + Nx[1] * kz)
                    want[idx] = want.get(idx, 0.0) + wx * wy * wz
        keys = set(k for k, v in want.items() if v > 1e-12) | set(k for k, v in got.items() if v > 1e-12)
        for k in keys:
            assert abs(got.get(k, 0.0) - want.get(k, 0.0)) < 2e-6, (p, i, k)


def test_density_conserves_mass():
    rng = np.random.default_rng(3)
    Nx, L = (12, 8, 16), (6.0, 4.0, 8.0)
    dx = [np.float32(np.float32(l) / n) for l, n in zip(L, Nx)]
    x = (rng.random((500, 3)) * np.asarray(L)).astype(np.float32).reshape(-1)
    inds, W = orc.grid_weights(x, Nx, dx, 3)
    M = int(np.prod(Nx))
    sites = np.arange(0, 500, 2, dtype=np.int32)
    q = np.where(np.arange(500) % 4 == 0, 1.0, -1.0).astype(np.float32)
    rho, rhoq = orc.fill_density(M, sites, inds, W, 64, q=q)
    gvol = float(np.float32(np.float32(dx[0] * dx[1]) * dx[2]))
    assert abs(rho.sum() * gvol - sites.size) < 1e-5
    assert abs(rhoq.sum() * gvol - q[sites].sum()) < 1e-5


def test_uk_tables_known_values():
    Nx, L = (16, 16, 16), (8.0, 8.0, 8.0)
    A, sig, Rp = 2.5, 0.7, 1.1
    g = orc.uk_table("gaussian", A, sig, 0.0, Nx, L)
    assert g[0] == pytest.approx(A)
    k2, k = orc.kvec_f(5, Nx, L)
    assert g[5] == pytest.approx(A * np.exp(-float(k2) * sig * sig / 2), rel=1e-12)
    e2 = orc.uk_table("erf2", A, sig, Rp, Nx, L)
    vol = 4 * np.pi * Rp ** 3 / 3
    assert e2[0] == pytest.approx(A * vol * vol, rel=1e-12)
    eg = orc.uk_table("erfG", A, sig, Rp, Nx, L)
    assert eg[0] == pytest.approx(A * vol, rel=1e-12)
    # 4 pi (sin x - x cos x)/k^3 is the transform of a sphere: check by quadrature
    kk = float(np.sqrt(k2))
    r = np.linspace(0, Rp, 20001)
    quad = np.trapezoid(4 * np.pi * r * np.sin(kk * r) / kk, r)
    assert eg[5] == pytest.approx(A * np.exp(-k2 * sig * sig / 2) * quad, rel=1e-7)
    c = orc.uk_table("charges", 2.0, 0.5, 0.0, Nx, L)
    assert c[0] == 0.0
    assert c[5] == pytest.approx(4 * np.pi * 2.0 / float(k2) * np.exp(-float(k2) * 0.25), rel=1e-12)


def small_system(charged=False, seed=5):
    if charged:
        cfg = systems.tild_coacervate(nchains=30, Nchain=8, nsolvent=200, box=6.0, grid=16, seed=seed)
    else:
        cfg = systems.tild_diblock(nchains=60, NA=5, NB=5, box=6.0, grid=16, seed=seed)
    return cfg


def oracle_system(cfg, integrator="GJF"):
    return orc.tild_system_from_config(cfg, integrator)


@pytest.mark.parametrize("charged", [False, True])
def test_net_force_vanishes(charged):
    # the reference's own logNetForce invariant (PS_Box.cu:323-335)
    sys = oracle_system(small_system(charged))
    out = orc.tild_forces(sys)
    f = out["f"].reshape(-1, 3)
    assert np.max(np.abs(f.sum(axis=0))) < 1e-9 * np.max(np.abs(f)) * f.shape[0]


def test_gather_is_transpose_of_scatter():
    rng = np.random.default_rng(9)
    Nx, L = (8, 12, 10), (4.0, 6.0, 5.0)
    dx = [np.float32(np.float32(l) / n) for l, n in zip(L, Nx)]
    x = (rng.random((300, 3)) * np.asarray(L)).astype(np.float32).reshape(-1)
    inds, W = orc.grid_weights(x, Nx, dx, 3)
    M = int(np.prod(Nx))
    gvol = float(np.float32(np.float32(dx[0] * dx[1]) * dx[2]))
    sites = np.arange(300, dtype=np.int32)
    rho, _ = orc.fill_density(M, sites, inds, W, 64)
    F = rng.normal(size=3 * M)
    f = np.zeros(900)
    orc.map_forces(f, F, sites, inds, W, 64, gvol, 3)
    lhs = f.reshape(-1, 3).sum(axis=0)
    rhs = gvol * (F.reshape(-1, 3) * rho[:, None]).sum(axis=0)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-10)


def test_gaussian_pair_force_matches_analytic():
    # two particles, large fine grid: the grid force approaches -grad of the
    # Gaussian pair potential smeared twice by the assignment function; with
    # p = 3 on a fine grid the smearing is small relative to sigma = 1.5
    Nx, L = (64, 64, 64), (16.0, 16.0, 16.0)
    A, sig = 3.0, 1.5
    x = np.array([[7.0, 8.0, 8.0], [9.0, 8.0, 8.0]], np.float32)
    topo = dict(intSpecies=np.array([0, 1], np.int32), nBonds=np.zeros(2, np.int32),
                bondedTo=np.zeros(4, np.int32), bondType=np.zeros(4, np.int32), charge=np.zeros(2, np.float32))
    sys = orc.TildSystem(Nx=Nx, L=L, pmeorder=3, species=["A", "B"], x=x.reshape(-1), topo=topo, bond_k=[0.0],
                         bond_r0=[0.0], bond_style=[0], potentials=[orc.TildPotential("gaussian", "A", "B", A=A, sigma=sig)],
                         mass=[1, 1], mobility=[1, 1])
    f = orc.tild_forces(sys)["f"].reshape(2, 3)
    r = 2.0
    # u(r) = A (2 pi s^2)^{-3/2} exp(-r^2 / 2 s^2) with s^2 = sigma^2 plus the
    # variance dx^2/3 of each particle's cubic assignment spline
    dx = 16.0 / 64
    s2 = sig * sig + 2 * dx * dx / 3
    du = A * (2 * np.pi * s2) ** -1.5 * np.exp(-r * r / (2 * s2)) * (-r / s2)
    assert f[0, 0] == pytest.approx(du, rel=3e-3)   # pushed away (negative x)
    assert f[1, 0] == pytest.approx(-du, rel=3e-3)
    assert abs(f[0, 1]) < 1e-9 and abs(f[0, 2]) < 1e-9


def test_bond_force_and_energy_analytic():
    L = (10.0, 10.0, 10.0)
    x = np.array([[1.0, 1.0, 1.0], [9.5, 1.0, 1.0]], np.float32)  # min-image distance 1.5
    topo = dict(nBonds=np.array([1, 1], np.int32), bondedTo=np.array([1, 0, 0, 0], np.int32),
                bondType=np.zeros(4, np.int32), bondReq=[0.5], bondK=[1.5], bondStyle=[0], maxbonds=2)
    f = np.zeros(6)
    orc.bonds(f, topo, x.reshape(-1), L)
    # u = k (r - r0)^2 -> |f| = 2 k (r - r0) toward the partner
    mag = 2 * 1.5 * (1.5 - 0.5)
    # the partner's image sits at x = -0.5, so particle 0 is pulled toward -x
    np.testing.assert_allclose(f.reshape(2, 3)[0], [-mag, 0, 0], atol=1e-6)
    np.testing.assert_allclose(f.reshape(2, 3)[1], [mag, 0, 0], atol=1e-6)
    e = orc.bond_energy(topo, x.reshape(-1), L)
    assert e[0] == pytest.approx(1.5 * (1.5 - 0.5) ** 2, rel=1e-6)
    # separation 3 (between L/4 and L/2): forces use the true minimum image,
    # the logged energy wraps against 0.5*Lh = 2.5 (ps_deviceBonds.cu:132-133)
    x2 = np.array([[1.0, 1.0, 1.0], [4.0, 1.0, 1.0]], np.float32)
    f2 = np.zeros(6)
    orc.bonds(f2, topo, x2.reshape(-1), L)
    np.testing.assert_allclose(f2.reshape(2, 3)[0], [2 * 1.5 * (3.0 - 0.5), 0, 0], atol=1e-6)
    e2 = orc.bond_energy(topo, x2.reshape(-1), L)
    assert e2[0] == pytest.approx(1.5 * (7.0 - 0.5) ** 2, rel=1e-6)


def test_bond_virial_closed_form():
    # d_bondStressEnergy's virial terms (ps_deviceBonds.cu:85-171):
    # -mf dr_a dr_b with mf = 2k(r - r0)/r, i.e. for a harmonic bond the tensor
    # -2k(r - r0) r n n^T (n the unit bond vector), the same for both members;
    # its trace is -2k(r - r0) r, the pair's contribution to -sum r.f
    L = (10.0, 10.0, 10.0)
    d = np.array([0.6, -0.8, 1.2])
    x = np.array([[2.0, 3.0, 4.0], [2.0, 3.0, 4.0] + d], np.float32)
    topo = dict(nBonds=np.array([1, 1], np.int32), bondedTo=np.array([1, 0, 0, 0], np.int32),
                bondType=np.zeros(4, np.int32), bondReq=[0.5], bondK=[1.5], bondStyle=[0], maxbonds=2)
    v = orc.bond_virial(topo, x.reshape(-1), L)
    dr = (x[0] - x[1]).astype(np.float64)
    r = np.linalg.norm(dr)
    mf = 2 * 1.5 * (r - 0.5) / r
    want = -mf * np.array([dr[0] ** 2, dr[1] ** 2, dr[2] ** 2, dr[0] * dr[1], dr[0] * dr[2], dr[1] * dr[2]])
    np.testing.assert_allclose(v[0], want, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(v[1], want, rtol=1e-12, atol=1e-12)
    assert v[0, :3].sum() == pytest.approx(-2 * 1.5 * (r - 0.5) * r, rel=1e-12)
    # the trace is r_ij . f_i of the bond force the oracle applies to i
    f = np.zeros(6)
    orc.bonds(f, topo, x.reshape(-1), L)
    assert v[0, :3].sum() == pytest.approx(float(dr @ f.reshape(2, 3)[0]), rel=1e-6)
    # FENE below r0: mf = 2k / (1 - (r/r0)^2); at r >= r0 no term
    topo_f = dict(topo, bondReq=[2.5], bondStyle=[1])
    v = orc.bond_virial(topo_f, x.reshape(-1), L)
    u = r / 2.5
    np.testing.assert_allclose(v[0, 0], -2 * 1.5 / (1 - u * u) * dr[0] ** 2, rtol=1e-12)
    topo_f = dict(topo, bondReq=[1.0], bondStyle=[1])
    assert np.all(orc.bond_virial(topo_f, x.reshape(-1), L) == 0.0)


def test_gjf_harmonic_configurational_temperature():
    # GJF samples <x^2> = kT/kappa exactly for a harmonic well at any stable dt
    # (Gronbech-Jensen & Farago 2013), here with dt = 0.2, kappa = 4
    n, kappa, dt = 4000, 4.0, 0.2
    L = np.array([1e6], np.float32)
    coef = orc.gjf_coeffs(dt, 1.0, 1.0)
    x = np.zeros(n)
    xo = np.zeros(n)
    no = np.zeros(n)
    types = np.zeros(n, np.int32)
    sites = np.arange(n, dtype=np.int32)
    lib = orc.lib()
    samples = []
    for step in range(3000):
        xs = x + 5e5
        f = -kappa * x
        xi = orc.normal4_np(np.arange(n, dtype=np.uint64), step, orc.TAG_TILD, 0, 77)[:, 0]
        xo_s = xo + 5e5
        lib.orc_gjf_step(orc._ptr(xs), orc._ptr(xo_s), orc._ptr(no), orc._ptr(f), orc._ptr(coef), orc._ptr(types),
                         orc._ptr(L), orc._ptr(sites), n, 1, orc._ptr(np.ascontiguousarray(xi)))
        x = xs - 5e5
        xo = xo_s - 5e5
        if step > 500:
            samples.append(np.mean(x * x))
    assert np.mean(samples) == pytest.approx(1.0 / kappa, rel=0.03)


def test_make_linear_topology_known_answer():
    # two diblocks (2 A + 1 B): beads 0..5; block bond types (1, 2)
    t = orc.make_linear(2, [(2, "A"), (1, "B")], {"A": 0, "B": 1}, block_bond_type=[1, 2])
    assert list(t["intSpecies"]) == [0, 0, 1, 0, 0, 1]
    assert list(t["nBonds"]) == [1, 2, 1, 1, 2, 1]
    # slot 0 = previous bead when present (PS_Box_moleculeMakers.cu:250-256)
    assert list(t["bondedTo"]) == [1, 0, 0, 2, 1, 0, 4, 0, 3, 5, 4, 0]
    assert list(t["bondType"]) == [0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 1, 0]
    # the product-side builder produces the same topology bit-for-bit
    p = systems.make_linear(2, [(2, 0), (1, 1)], block_bond_type=[1, 2])
    assert list(p["type"]) == list(t["intSpecies"])
    assert list(p["nBonds"]) == list(t["nBonds"])
    mask = (np.arange(12) % 2) < np.repeat(t["nBonds"], 2)
    assert list(p["bondedTo"][mask]) == list(t["bondedTo"][mask])
    assert list(p["bondType"][mask]) == list(t["bondType"][mask])


@pytest.mark.parametrize("phase, direction", [("L", 0), ("L", 2), ("H", 1), ("H", 2), ("G", 0)])
def test_bias_force_is_the_analytic_gradient(phase, direction):
    # the bias fields of Box::make_bias_field are trigonometric polynomials
    # below the Nyquist frequency, so the spectral force of
    # ps_potentialBias.cu:100-185 is their exact -grad (up to the float
    # rounding of the sample positions and, for the gyroid, of its terms)
    Nx, L = (16, 12, 20), (4.0, 3.0, 5.0)
    pot = orc.TildPotential("bias", "A", "A", A=1.7, phase=phase, dir=direction, n_periods=1)
    u = orc.bias_field(pot, Nx, L)
    f = orc.bias_force(u, Nx, L).reshape(-1, 3)
    iz, iy, ix = np.meshgrid(np.arange(Nx[2]), np.arange(Nx[1]), np.arange(Nx[0]), indexing="ij")
    r = [ix.ravel() * L[0] / Nx[0], iy.ravel() * L[1] / Nx[1], iz.ravel() * L[2] / Nx[2]]
    a = [2 * np.pi / l for l in L]
    g = np.zeros((u.size, 3))
    if phase == "L":
        g[:, direction] = 1.7 * a[direction] * np.cos(a[direction] * r[direction])
        assert np.allclose(u, 1.7 * np.sin(a[direction] * r[direction]), atol=1e-5)
    elif phase == "H":
        # Box.cu:405-411: dir 0 -> axes (2, 1), dir 1 -> (0, 2), dir 2 -> (0, 1)
        d1, d2 = (2, 1) if direction == 0 else ((0, 2) if direction == 1 else (0, 1))
        g[:, d1] = -1.7 * a[d1] * np.sin(a[d1] * r[d1]) * np.cos(a[d2] * r[d2])
        g[:, d2] = -1.7 * a[d2] * np.cos(a[d1] * r[d1]) * np.sin(a[d2] * r[d2])
    else:
        s = [np.sin(a[j] * r[j]) for j in range(3)]
        c = [np.cos(a[j] * r[j]) for j in range(3)]
        e = s[0] * c[1] + s[1] * c[2] + s[2] * c[0]
        g[:, 0] = 1.7 * 2 * e * a[0] * (c[0] * c[1] - s[2] * s[0])
        g[:, 1] = 1.7 * 2 * e * a[1] * (c[1] * c[2] - s[0] * s[1])
        g[:, 2] = 1.7 * 2 * e * a[2] * (c[2] * c[0] - s[1] * s[2])
    # force = -grad u
    assert np.max(np.abs(f + g)) < 2e-5 * max(1.0, np.abs(g).max())


def angle_system(seed=7):
    """Diblock chains of 10 with an angle at every interior bead: WLC
    (k = 2) at A vertices, harmonic (k = 3, theta_eq = 100 degrees, converted
    as the input reader does: float degrees times PI / 180 in double) at B."""
    cfg = small_system(False, seed=seed)
    t = cfg["topo"]["type"]
    th = float(np.float32(100.0 * (float(np.float32(3.141592654)) / 180.0)))
    return systems.add_chain_angles(cfg, 10, lambda i: int(t[i]), k=[2.0, 3.0], theta=[0.0, th], style=[0, 1])


def test_angle_forces_are_minus_energy_gradient():
    # d_angles (ps_deviceAngles.cu:8-145) against central differences of the
    # energy d_anglesStressEnergy / 3 (PS_Box.cu:296-305) in FP64: a wrong
    # sign, factor or member role fails here
    cfg = angle_system()
    ang, L = cfg["angles"], cfg["L"]
    x = cfg["x"].astype(np.float64).reshape(-1)
    f = np.zeros_like(x)
    orc.angle_forces(f, ang, x, L)
    rng = np.random.default_rng(3)
    h = 1e-6
    for c in rng.choice(x.size, 60, replace=False):
        xp, xm = x.copy(), x.copy()
        xp[c] += h
        xm[c] -= h
        g = (orc.angle_energy(ang, xp, L) - orc.angle_energy(ang, xm, L)) / (2 * h)
        assert f[c] == pytest.approx(-g, rel=1e-5, abs=1e-6), c
    # each angle's forces sum to zero
    assert np.max(np.abs(f.reshape(-1, 3).sum(axis=0))) < 1e-9


def test_angle_closed_forms():
    # a right angle: WLC k (1 + cos) = k; harmonic at its theta_eq: no energy
    # and no force; WLC on a straight line: no energy and no force
    L = (10.0, 10.0, 10.0)

    def one(x, style, k, th):
        ang = dict(nAngles=np.array([1, 1, 1], np.int32), angleType=np.zeros((3, 1), np.int32),
                   angleGroup=np.tile(np.array([0, 1, 2], np.int32), (3, 1, 1)), maxangles=1,
                   k=[k], theta=[th], style=[style])
        f = np.zeros(9)
        orc.angle_forces(f, ang, np.asarray(x, np.float64).reshape(-1), L)
        return orc.angle_energy(ang, np.asarray(x, np.float64).reshape(-1), L), f

    right = [[6.0, 5.0, 5.0], [5.0, 5.0, 5.0], [5.0, 6.5, 5.0]]
    e, f = one(right, 0, 2.5, 0.0)
    assert e == pytest.approx(2.5)
    # WLC force on i: -k (rkj / |rij||rkj| - cos rij / |rij|^2) = -k (0, 1, 0)
    np.testing.assert_allclose(f[0:3], [0.0, -2.5, 0.0], atol=1e-12)
    e, f = one(right, 1, 4.0, np.pi / 2)
    assert abs(e) < 1e-12 and np.max(np.abs(f)) < 1e-6
    e, f = one([[4.0, 5.0, 5.0], [5.0, 5.0, 5.0], [6.2, 5.0, 5.0]], 0, 2.0, 0.0)
    assert abs(e) < 1e-12 and np.max(np.abs(f)) < 1e-12
    # across the periodic boundary: the minimum image is used
    e2, _ = one([[9.5, 5.0, 5.0], [0.5, 5.0, 5.0], [0.5, 6.0, 5.0]], 0, 2.5, 0.0)
    assert e2 == pytest.approx(2.5)


def _dimer(v0=None, k=1.5):
    """Two particles joined by a harmonic bond (k (r - 0)^2) and no grid
    potentials, in a 10^3 box."""
    x = np.array([4.0, 5.0, 5.0, 5.2, 5.3, 4.9], np.float32)
    topo = dict(intSpecies=np.array([0, 0], np.int32), nBonds=np.array([1, 1], np.int32),
                bondedTo=np.array([1, 0, 0, 0], np.int32), bondType=np.zeros(4, np.int32),
                charge=np.zeros(2, np.float32))
    return orc.TildSystem(Nx=(8, 8, 8), L=(10.0, 10.0, 10.0), pmeorder=1, species=["A"], x=x, topo=topo,
                          bond_k=[k], bond_r0=[0.0], bond_style=[0], potentials=[], mass=[2.0], mobility=[1.0],
                          delt=0.01, integrator="VV", v=v0)


def test_vv_free_flight_and_first_step():
    # no forces: x_n = x_0 + n dt v_0, wrapped into [0, L); from rest with no
    # stored force the first step does not move the particles
    sys = _dimer(v0=np.array([30.0, -1.0, 0.5, -2.0, 40.0, 0.0]), k=0.0)
    x0 = sys.x.astype(np.float64).copy()
    for n in range(1, 4):
        o = orc.tild_step(sys)
        sys.x, sys.v, sys.f_prev = o["x"].astype(np.float32), o["v"], o["f_prev"]
        want = (x0 + n * 0.01 * np.array([30.0, -1.0, 0.5, -2.0, 40.0, 0.0])) % 10.0
        np.testing.assert_allclose(sys.x, want, atol=2e-6)
    rest = _dimer()
    o = orc.tild_step(rest)
    np.testing.assert_array_equal(o["x"], rest.x.astype(np.float64))
    assert np.all(o["v"] != 0.0)


def test_vv_is_time_reversible():
    # velocity Verlet is reversible: N steps, flip the velocities, N steps
    # back to the start (a kick in the wrong place breaks this); the stored
    # force starts consistent with the positions
    sys = _dimer(v0=np.array([0.3, 0.0, -0.2, -0.3, 0.1, 0.2]))
    sys.f_prev = orc.tild_forces(sys)["f"]
    x0 = sys.x.astype(np.float64).copy()
    xs = sys.x.astype(np.float64)
    for sign in (1, -1):
        if sign < 0:
            sys.v = -sys.v
        for _ in range(25):
            o = orc.tild_step(sys)
            xs, sys.v, sys.f_prev = o["x"], o["v"], o["f_prev"]
            sys.x = xs        # FP64 positions for this property
    np.testing.assert_allclose(xs, x0, atol=1e-9)


def test_langevin_with_vv_thermalizes_to_unit_temperature():
    # Langevin drag + noise (ps_potentialLangevin.cu) with velocity Verlet on
    # free particles of mass 2: equipartition m <v^2> = kT = 1 per degree of
    # freedom (a wrong drag sign, noise magnitude or mass factor fails)
    rng = np.random.default_rng(4)
    n = 1500
    x = (rng.random(3 * n) * 10.0).astype(np.float32)
    topo = dict(intSpecies=np.zeros(n, np.int32), nBonds=np.zeros(n, np.int32), bondedTo=np.zeros(2 * n, np.int32),
                bondType=np.zeros(2 * n, np.int32), charge=np.zeros(n, np.float32))
    sys = orc.TildSystem(Nx=(8, 8, 8), L=(10.0, 10.0, 10.0), pmeorder=1, species=["A"], x=x, topo=topo,
                         bond_k=[0.0], bond_r0=[0.0], bond_style=[0], mass=[2.0], mobility=[1.0], delt=0.01,
                         integrator="VV", seed=11,
                         potentials=[orc.TildPotential("langevin", "A", "A", A=1.0, sigma=0.01)])
    acc = []
    for step in range(500):
        sys.step = step
        o = orc.tild_step(sys)
        sys.x, sys.v, sys.f_prev = o["x"].astype(np.float32), o["v"], o["f_prev"]
        if step >= 250:
            acc.append(np.mean(2.0 * sys.v ** 2))
    assert np.mean(acc) == pytest.approx(1.0, rel=0.06)


def _dpd_fluid(n=800, box=8.0, kT=1.5, gamma=4.5, seed=21):
    rng = np.random.default_rng(seed)
    x = (rng.random(3 * n) * box).astype(np.float32)
    topo = dict(intSpecies=np.zeros(n, np.int32), nBonds=np.zeros(n, np.int32), bondedTo=np.zeros(2 * n, np.int32),
                bondType=np.zeros(2 * n, np.int32), charge=np.zeros(n, np.float32))
    return orc.TildSystem(Nx=(8, 8, 8), L=(box, box, box), pmeorder=1, species=["A"], x=x, topo=topo,
                          bond_k=[0.0], bond_r0=[0.0], bond_style=[0], mass=[1.0], mobility=[1.0], delt=0.02,
                          integrator="VV", seed=seed,
                          potentials=[orc.TildPotential("dpd", "A", "A", A=gamma, sigma=kT, Rp=1.0, wall_pos=0.02)])


def test_dpd_conserves_momentum_and_thermalizes():
    # DPD pair forces (ps_potentialDPD.cu) obey Newton's third law, so the
    # total force vanishes; with velocity Verlet the fluid reaches the set
    # temperature, <v^2> = kT (fluctuation-dissipation: sigma^2 = 2 gamma kT)
    sys = _dpd_fluid()
    sys.v = np.random.default_rng(1).normal(size=3 * 800) * 0.2
    f = orc.tild_forces(sys)["f"].reshape(-1, 3)
    assert np.max(np.abs(f.sum(axis=0))) < 1e-9 * np.max(np.abs(f)) * f.shape[0]
    sys.v = None
    acc = []
    for step in range(300):
        sys.step = step
        o = orc.tild_step(sys)
        sys.x, sys.v, sys.f_prev = o["x"].astype(np.float32), o["v"], o["f_prev"]
        if step >= 150:
            acc.append(np.mean(sys.v.reshape(-1, 3) ** 2))
    assert np.mean(acc) == pytest.approx(1.5, rel=0.08)


def _em(x, f, dt, L, xi):
    """orc_em_step (d_EM_integrator, ps_integratorEM.cu:41-82) on [n, 3] arrays."""
    n = x.shape[0]
    xs = np.ascontiguousarray(x, np.float64).reshape(-1).copy()
    fs = np.ascontiguousarray(f, np.float64).reshape(-1)
    Lf = np.asarray(L, np.float32)
    orc.lib().orc_em_step(orc._ptr(xs), orc._ptr(fs), dt, orc._ptr(Lf), orc._ptr(np.arange(n, dtype=np.int32)), n, 3,
                          orc._ptr(np.ascontiguousarray(xi, np.float64).reshape(-1)))
    return xs.reshape(n, 3)


def test_em_step_without_noise_is_identity_or_pure_drift():
    # f = 0 and xi = 0 leave x unchanged; xi = 0 gives x + dt f exactly, with
    # the single wrap into [0, L] of the reference (one particle leaves through
    # each face)
    rng = np.random.default_rng(4)
    L = (10.0, 12.0, 8.0)
    x = 1.0 + rng.random((200, 3)) * (np.asarray(L) - 2.0)   # one unit from every face
    zero = np.zeros_like(x)
    assert np.array_equal(_em(x, zero, 0.01, L, zero), x)
    f = rng.normal(0.0, 3.0, x.shape)
    x[0] = (9.99, 6.0, 4.0)
    f[0] = (2.0, 0.0, 0.0)
    x[1] = (5.0, 0.005, 4.0)
    f[1] = (0.0, -1.0, 0.0)
    dt = 0.01
    got = _em(x, f, dt, L, zero)
    want = x + dt * f
    want[0, 0] -= L[0]
    want[1, 1] += L[1]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


def test_em_free_diffusion_msd_is_2dt_per_axis():
    # force-free EM: x(t) - x(0) is a sum of sqrt(2 dt) N(0,1) steps, so
    # <dx^2> = 2 dt n per axis (unit mobility); sqrt(dt) noise would give half
    n, steps, dt = 20000, 40, 0.005
    L = (1e4, 1e4, 1e4)
    x0 = np.full((n, 3), 5e3)
    x = x0.copy()
    zero = np.zeros_like(x)
    for step in range(steps):
        x = _em(x, zero, dt, L, orc.tild_noise(n, 31, step).reshape(n, 3))
    d = x - x0
    msd = np.mean(d * d, axis=0)
    np.testing.assert_allclose(msd, 2 * dt * steps, rtol=0.025)
    assert np.all(np.abs(np.mean(d, axis=0)) < 4 * np.sqrt(2 * dt * steps / n))


@pytest.mark.parametrize("kind", ["charges", "gaussian"])
def test_pair_energy_of_a_single_fourier_mode(kind):
    # rho_I = rho_J = a cos(k x) on the grid: u * rho = a u(k) cos(k x) and
    # sum cos^2 = M/2, so E = gvol a^2 u(k) M / 2 = V a^2 u(k) / 2, halved for
    # the charges potential (ps_potentialCharges.cu:115) whose
    # u(k) = 4 pi l_B / k^2 exp(-k^2 sigma^2); the Gaussian's is
    # A exp(-k^2 sigma^2 / 2) (ps_potentialGaussian.cu:43-45)
    Nx, L = (16, 12, 10), (8.0, 6.0, 5.0)
    M = int(np.prod(Nx))
    a, n = 0.7, 3
    kx = 2 * np.pi * n / L[0]
    ix = np.arange(M) % Nx[0]
    rho = a * np.cos(kx * ix * L[0] / Nx[0])
    gvol = float(np.prod(np.asarray(L) / np.asarray(Nx)))
    V = float(np.prod(L))
    if kind == "charges":
        lB, sig = 2.0, 0.5
        uk = orc.uk_table("charges", lB, sig, 0.0, Nx, L)
        want = 0.5 * V * a * a / 2 * 4 * np.pi * lB / kx ** 2 * np.exp(-kx ** 2 * sig ** 2)
    else:
        A, sig = 3.0, 0.8
        uk = orc.uk_table("gaussian", A, sig, 0.0, Nx, L)
        want = V * a * a / 2 * A * np.exp(-kx ** 2 * sig ** 2 / 2)
    got = orc.pair_energy(kind, rho, rho, uk, Nx, gvol)
    # the tables use the reference's float k (Box::get_kD), hence 1e-6
    assert got == pytest.approx(want, rel=1e-6)
