"""Pins the CL-FTS oracle with closed forms and invariants.  CPU only."""
import numpy as np
import pytest

import orc
from mft import systems


def tiny_spec(Nx=(8, 6, 10), L=(3.0, 2.5, 4.0), blocks=((4, 0), (3, 1)), Nr=8.0, C=3.0, pots=None):
    if pots is None:
        pots = [dict(kind="Helfand", kappaN=30.0, dt_pl=0.02),
                dict(kind="Flory", I=0, J=1, chiN=12.0, dt_pl=0.02, dt_mi=0.03)]
    return dict(Nx=Nx, L=L, Nr=Nr, C=C, nspecies=2, cl=True, seed=5, potentials=pots,
                molecules=[dict(blocks=list(blocks), phi=1.0)])


def random_fields(spec, scale, seed=1, imag=True):
    rng = np.random.default_rng(seed)
    M = int(np.prod(spec["Nx"]))
    f = {}
    for k, p in enumerate(spec["potentials"]):
        f[(k, "pl")] = rng.normal(0, scale, M) + (1j * rng.normal(0, scale, M) if imag else 0)
        if p["kind"] == "Flory":
            f[(k, "mi")] = rng.normal(0, scale, M) + (1j * rng.normal(0, scale, M) if imag else 0)
    return f


def test_zero_field_gives_unit_Q_and_uniform_density():
    spec = tiny_spec()
    s = orc.fts_system_from_spec(spec, {})
    orc.fts_init(s)
    assert abs(s.Q[0] - 1.0) < 1e-12
    Ntot = 7
    n = s.molecules[0].nmolecs
    total = s.rho.sum(axis=0)
    np.testing.assert_allclose(total, n * Ntot / s.V(), rtol=1e-12)


def test_constant_field_closed_form():
    # Incompress w+ = -i c Nr makes W = c on every species: q_s = e^{-c(s+1)}
    c = 0.37
    spec = tiny_spec(pots=[dict(kind="Incompress", dt_pl=0.01)])
    M = int(np.prod(spec["Nx"]))
    s = orc.fts_system_from_spec(spec, {(0, "pl"): np.full(M, -1j * c * spec["Nr"])})
    orc.fts_init(s)
    assert np.log(s.Q[0]) == pytest.approx(-c * 7, rel=1e-12)


def debye_S(k2, Nr, Ntot):
    g = np.exp(-k2 / (Nr - 1.0))
    i = np.arange(Ntot)
    return np.sum(g ** np.abs(i[:, None] - i[None, :]))


def test_linear_response_is_discrete_chain_debye_function():
    # W = eps cos(k x) on every bead: delta rho(r) = -(n/V) S(k) eps cos(k x),
    # S(k) = sum_ij g^|i-j|, g = exp(-k^2/(Nr-1)), to first order in eps
    eps = 1e-5
    Nx, L = (16, 4, 4), (6.0, 2.0, 2.0)
    spec = tiny_spec(Nx=Nx, L=L, blocks=((9, 0),), Nr=10.0, pots=[dict(kind="Incompress", dt_pl=0.01)])
    x = np.arange(Nx[0]) * L[0] / Nx[0]
    k = float(np.float32(np.float32(6.2831853071) * np.float32(1.0)) / np.float32(L[0]))
    W = eps * np.cos(k * x)
    Wfull = np.broadcast_to(W[None, None, :], (Nx[2], Nx[1], Nx[0])).reshape(-1)
    s = orc.fts_system_from_spec(spec, {(0, "pl"): -1j * Wfull * spec["Nr"]})
    orc.fts_init(s)
    n = s.molecules[0].nmolecs
    rho = s.rho.sum(axis=0).reshape(Nx[2], Nx[1], Nx[0])[0, 0, :]
    drho = rho.real - n * 9 / s.V()
    want = -(n / s.V()) * debye_S(k * k, spec["Nr"], 9) * eps * np.cos(k * x)
    np.testing.assert_allclose(drho, want, rtol=0, atol=2e-4 * np.max(np.abs(want)))


def test_density_integrates_to_total_sites():
    spec = tiny_spec()
    s = orc.fts_system_from_spec(spec, random_fields(spec, 0.3))
    orc.fts_init(s)
    total = s.rho.sum() * s.gvol()
    assert total == pytest.approx(s.molecules[0].nmolecs * 7, rel=1e-10)


def test_symmetric_molecule_complementary_equals_forward():
    spec = tiny_spec(blocks=((2, 0), (3, 1), (2, 0)))
    s = orc.fts_system_from_spec(spec, random_fields(spec, 0.3))
    orc.fts_build_species_fields(s)
    s.init_nmolecs()
    out = orc.fts_molecule_densities(s, keep=True)[0]
    np.testing.assert_array_equal(out["q"], out["qdag"])
    # reversing the block order of an asymmetric chain swaps q and qdag
    spec2 = tiny_spec(blocks=((4, 0), (3, 1)))
    spec3 = tiny_spec(blocks=((3, 1), (4, 0)))
    f = random_fields(spec2, 0.3, seed=4)
    a = orc.fts_system_from_spec(spec2, f)
    b = orc.fts_system_from_spec(spec3, f)
    for t in (a, b):
        orc.fts_build_species_fields(t)
        t.init_nmolecs()
    oa = orc.fts_molecule_densities(a, keep=True)[0]
    ob = orc.fts_molecule_densities(b, keep=True)[0]
    np.testing.assert_allclose(oa["q"], ob["qdag"], rtol=1e-12)
    assert a.Q[0] == pytest.approx(b.Q[0], rel=1e-12)


@pytest.mark.parametrize("which", [(0, "pl"), (1, "pl"), (1, "mi")])
def test_forces_are_field_derivatives_of_H(which):
    # the CL drift dH/dw(r) (fts_potential_*.cu force kernels) equals
    # (1/gvol) dH/dw(r) of the Hamiltonian (FTS_Box::computeHamiltonian); H
    # is holomorphic in w, so a real finite difference gives the derivative
    spec = tiny_spec()
    f = random_fields(spec, 0.2, seed=7)
    s = orc.fts_system_from_spec(spec, f)
    orc.fts_init(s)
    M = s.M
    gvol = s.gvol()
    rho_tot = s.rho.sum(axis=0)
    p = s.potentials[which[0]]
    if p.kind == "Helfand":
        dH = np.zeros(M, complex)
        orc.lib().orc_fts_helfand_force(orc._ptr(dH), orc._ptr(p.wpl), orc._ptr(rho_tot), s.C, p.kappaN, s.Nr, M)
    else:
        dHp = np.zeros(M, complex)
        dHm = np.zeros(M, complex)
        rI = np.ascontiguousarray(s.rho[p.I])
        rJ = np.ascontiguousarray(s.rho[p.J])
        orc.lib().orc_fts_flory_force(orc._ptr(dHp), orc._ptr(dHm), orc._ptr(p.wpl), orc._ptr(p.wmi), orc._ptr(rI),
                                      orc._ptr(rJ), s.C, p.chiN, s.Nr, M)
        dH = dHp if which[1] == "pl" else dHm
    h = 1e-5
    for site in (0, 17, M - 1):
        vals = []
        for sgn in (1, -1):
            f2 = {k: v.copy() for k, v in f.items()}
            f2[which][site] += sgn * h
            t = orc.fts_system_from_spec(spec, f2)
            orc.fts_init(t)
            vals.append(orc.fts_hamiltonian(t)[0])
        deriv = (vals[0] - vals[1]) / (2 * h) / gvol
        assert abs(deriv - dH[site]) < 1e-5 * max(1.0, abs(dH[site])), (site, deriv, dH[site])


def test_em_with_zero_step_is_identity():
    spec = tiny_spec()
    for p in spec["potentials"]:
        p["dt_pl"] = 0.0
        p["dt_mi"] = 0.0
    f = random_fields(spec, 0.2)
    s = orc.fts_system_from_spec(spec, f)
    orc.fts_init(s)
    orc.fts_update_fields(s, z=None)
    np.testing.assert_array_equal(s.potentials[0].wpl, f[(0, "pl")])
    np.testing.assert_array_equal(s.potentials[1].wmi, f[(1, "mi")])


def test_cl_noise_is_real_with_variance_2dt_over_gvol():
    spec = tiny_spec()
    spec["potentials"][1]["chiN"] = 1e30   # switch off the Flory drift
    spec["potentials"] = spec["potentials"][1:]
    M = int(np.prod(spec["Nx"]))
    s = orc.fts_system_from_spec(spec, {})
    orc.fts_init(s)
    s.rho[:] = s.rho.mean()          # uniform densities: zero Flory w- drift
    w0 = s.potentials[0].wmi.copy()
    incs = []
    for it in range(40):
        z = orc.fts_noise(M, 11, it)
        orc.fts_update_fields(s, z)
        incs.append(s.potentials[0].wmi - w0)
        w0 = s.potentials[0].wmi.copy()
    incs = np.array(incs)
    assert np.all(incs.imag == 0)
    assert np.var(incs.real) == pytest.approx(2 * 0.03 / s.gvol(), rel=0.05)


def _deterministic_run(scheme, dt, T, spec0, fields):
    """Field relaxation without noise over time T with step dt."""
    spec = dict(spec0)
    spec["potentials"] = [dict(p, dt_pl=dt, dt_mi=dt, scheme=scheme) for p in spec0["potentials"]]
    s = orc.fts_system_from_spec(spec, fields)
    orc.fts_init(s)
    for _ in range(int(round(T / dt))):
        orc.fts_step(s, None)
    return np.concatenate([np.concatenate([p.wpl, p.wmi if p.wmi is not None else []]) for p in s.potentials])


def test_empc_is_second_order_and_em_first_order():
    """Without noise the field update is an ODE; the predictor-corrector step
    (Heun's method) converges as dt^2 and Euler-Maruyama as dt.  Errors
    against a fine EMPC run at T = 0.16: the ratio for halving dt is ~4 for
    EMPC and ~2 for EM (a dropped or mis-signed corrector term breaks it)."""
    spec = tiny_spec(Nx=(6, 6, 6), L=(2.0, 2.0, 2.0), blocks=((3, 0), (3, 1)), Nr=6.0, C=2.0)
    fields = random_fields(spec, 0.3, seed=4, imag=False)
    T = 0.16
    ref = _deterministic_run("EMPC", 0.0025, T, spec, fields)
    err = {}
    for scheme in ("EM", "EMPC"):
        e = [np.max(np.abs(_deterministic_run(scheme, dt, T, spec, fields) - ref)) for dt in (0.04, 0.02)]
        err[scheme] = e[0] / e[1]
    assert 1.7 < err["EM"] < 2.4, err
    assert 3.3 < err["EMPC"] < 4.7, err


def test_empc_with_zero_step_is_identity():
    spec = tiny_spec()
    spec["potentials"] = [dict(p, scheme="EMPC", dt_pl=0.0, dt_mi=0.0) for p in spec["potentials"]]
    fields = random_fields(spec, 0.2, seed=2)
    s = orc.fts_system_from_spec(spec, fields)
    orc.fts_init(s)
    before = [p.wpl.copy() for p in s.potentials]
    orc.fts_step(s, orc.fts_noise(s.M, 5, 0))
    for p, b in zip(s.potentials, before):
        assert np.array_equal(p.wpl, b)


def test_block_debye_terms_sum_to_chain_debye():
    """Intra-block A + intra-block C + the A-C cross term of a diblock equal
    the Debye function of the whole chain at every k != 0 (fA + fC = 1)."""
    k2 = np.linspace(0.0, 40.0, 201)
    for fA in (0.25, 0.5, 0.7):
        for alpha in (0.5, 1.0, 2.5):
            whole = orc.debye_homopolymer(k2, alpha)
            parts = (orc.debye_intra_block(k2, fA, alpha) + orc.debye_intra_block(k2, 1.0 - fA, alpha)
                     + orc.debye_inter_block(k2, fA, 0.0, 1.0 - fA, alpha))
            np.testing.assert_allclose(parts[1:], whole[1:], rtol=1e-12, atol=1e-14)
            assert parts[0] == 0.0 and whole[0] == 1.0   # the reference's k = 0 values


def test_debye_function_limits():
    x = np.array([0.0, 1e-3, 2e-3, 500.0, 1000.0])
    g = orc.debye_homopolymer(x, 1.0)
    np.testing.assert_allclose(g[1:3], 1 - x[1:3] / 3 + x[1:3] ** 2 / 12, rtol=1e-8)
    np.testing.assert_allclose(g[3:], 2.0 / x[3:] - 2.0 / x[3:] ** 2, rtol=1e-9)


def test_linear_coefficients_of_a_diblock():
    """A+ of Flory (block terms) and of Helfand (whole-chain term) carry the
    same Debye part at k != 0; at k = 0 Helfand keeps the prefactor and Flory
    none; A- of Flory is the constant 2C/chiN."""
    spec = tiny_spec(blocks=((4, 0), (3, 1)))
    s = orc.fts_system_from_spec(spec, {})
    s.init_nmolecs()
    orc.fts_linear_coeffs(s)
    h, f = s.potentials
    dh = h.Akpl - s.C / h.kappaN
    df = f.Akpl - 2.0 * s.C / f.chiN
    np.testing.assert_allclose(df[1:], dh[1:], rtol=1e-12)
    m = s.molecules[0]
    assert df[0] == 0.0 and dh[0] == pytest.approx(m.nmolecs * (7 / s.Nr) * 7 / s.V(), rel=1e-12)
    assert np.all(f.Akmi == 2.0 * s.C / f.chiN)


def test_one_step_scheme_is_consistent_with_em():
    """Without noise, one 1S step and one EM step from the same state differ
    by O(dt^2) (the semi-implicit terms cancel to first order): the
    difference shrinks ~4x when dt halves."""
    spec = tiny_spec(Nx=(6, 6, 6), L=(2.0, 2.0, 2.0), blocks=((3, 0), (3, 1)), Nr=6.0, C=2.0)
    fields = random_fields(spec, 0.3, seed=4, imag=False)
    d = []
    for dt in (0.02, 0.01):
        out = []
        for scheme in ("EM", "1S"):
            sp = dict(spec)
            sp["potentials"] = [dict(p, dt_pl=dt, dt_mi=dt, scheme=scheme) for p in spec["potentials"]]
            s = orc.fts_system_from_spec(sp, fields)
            orc.fts_init(s)
            w0 = np.concatenate([s.potentials[0].wpl, s.potentials[1].wpl, s.potentials[1].wmi])
            orc.fts_step(s, None)
            out.append(np.concatenate([s.potentials[0].wpl, s.potentials[1].wpl, s.potentials[1].wmi]))
        d.append(np.max(np.abs(out[0] - out[1])) / np.max(np.abs(out[0] - w0)))
    assert 1.7 < d[0] / d[1] < 2.3, d   # relative to the O(dt) change: O(dt)


def _fd_check_forces(spec, f):
    """Every potential's drift (orc._forces) against (1/gvol) dH/dw by central
    differences at three sites."""
    s = orc.fts_system_from_spec(spec, f)
    orc.fts_init(s)
    M, gvol = s.M, s.gvol()
    h = 1e-5
    for k, p in enumerate(s.potentials):
        dHp, dHm = orc._forces(s, p)
        for which, dH in (("pl", dHp), ("mi", dHm)):
            if dH is None:
                continue
            for site in (0, 17, M - 1):
                vals = []
                for sgn in (1, -1):
                    f2 = {key: v.copy() for key, v in f.items()}
                    f2[(k, which)][site] += sgn * h
                    t = orc.fts_system_from_spec(spec, f2)
                    orc.fts_init(t)
                    vals.append(orc.fts_hamiltonian(t)[0])
                deriv = (vals[0] - vals[1]) / (2 * h) / gvol
                assert abs(deriv - dH[site]) < 1e-5 * max(1.0, abs(dH[site])), (k, which, site, deriv, dH[site])


def test_edwards_force_is_field_derivative_of_H():
    # d_makeEdwardsForce (fts_potential_edwards.cu:227-244) against
    # PotentialEdwards::calcHamiltonian and -n log Q
    spec = tiny_spec(pots=[dict(kind="Edwards", B=0.7, dt_pl=0.02),
                           dict(kind="Flory", I=0, J=1, chiN=12.0, dt_pl=0.02, dt_mi=0.03)])
    _fd_check_forces(spec, random_fields(spec, 0.2, seed=7))


def test_grand_canonical_forces_are_field_derivatives_of_H():
    # with H term -z V Q (fts_molecule_linear.cu:374-376) the drift keeps the
    # form i rho / Nr, rho built with the factor z
    spec = tiny_spec()
    spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], activity=0.9)]
    _fd_check_forces(spec, random_fields(spec, 0.2, seed=8))


def test_edwards_is_helfand_without_the_constant():
    # Edwards(B) = Helfand(kappa N = C B) + i C, the same excluded-volume force
    spec = tiny_spec(pots=[dict(kind="Edwards", B=0.7, dt_pl=0.02)])
    f = random_fields(spec, 0.2, seed=9)
    s = orc.fts_system_from_spec(spec, f)
    orc.fts_init(s)
    e, _ = orc._forces(s, s.potentials[0])
    s.potentials[0].kind, s.potentials[0].kappaN = "Helfand", s.C * 0.7
    hf, _ = orc._forces(s, s.potentials[0])
    np.testing.assert_allclose(e, hf + 1j * s.C, rtol=1e-13, atol=1e-13)


def test_grand_canonical_zero_field_closed_form():
    # W = 0: q = 1, Q = 1, every block density z N_b (factor z,
    # fts_molecule_linear.cu:289-291), nSites = z V Q Ntot, H term -z V
    z = 0.35
    spec = tiny_spec()
    spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], activity=z)]
    s = orc.fts_system_from_spec(spec, {})
    orc.fts_init(s)
    assert abs(s.Q[0] - 1.0) < 1e-12
    np.testing.assert_allclose(s.rho[0], z * 4, rtol=1e-12)
    np.testing.assert_allclose(s.rho[1], z * 3, rtol=1e-12)
    H, parts = orc.fts_hamiltonian(s)
    assert parts[-1] == pytest.approx(-z * s.V(), rel=1e-12)


def test_grand_canonical_density_is_canonical_density_rescaled():
    # for the same fields, factor z versus n / (Q V): the densities differ by
    # z Q V / n exactly
    spec = tiny_spec()
    f = random_fields(spec, 0.2, seed=10)
    c = orc.fts_system_from_spec(spec, f)
    orc.fts_init(c)
    spec_g = dict(spec, molecules=[dict(blocks=[(4, 0), (3, 1)], activity=0.6)])
    g = orc.fts_system_from_spec(spec_g, f)
    orc.fts_init(g)
    scale = 0.6 * c.Q[0] * c.V() / c.molecules[0].nmolecs
    np.testing.assert_allclose(g.rho, c.rho * scale, rtol=1e-12, atol=1e-14)


def test_zero_mean_recentres_w_plus_after_the_update():
    # modify zeroMean (fts_potential_helfand.cu:140-146): w+ has zero grid
    # mean after the update; Flory only parses it (fts_potential_flory.cu:94-98)
    pots = [dict(kind="Helfand", kappaN=30.0, dt_pl=0.02, zero_mean=True),
            dict(kind="Flory", I=0, J=1, chiN=12.0, dt_pl=0.02, dt_mi=0.03, zero_mean=True)]
    spec = tiny_spec(pots=pots)
    f = random_fields(spec, 0.2, seed=14)
    s = orc.fts_system_from_spec(spec, f)
    orc.fts_init(s)
    orc.fts_step(s, orc.fts_noise(s.M, 5, 0))
    assert abs(s.potentials[0].wpl.mean()) < 1e-15
    assert abs(s.potentials[1].wpl.mean()) > 1e-3


def test_zero_mean_incompress_leaves_densities_unchanged():
    # the incompressible drift i rho/Nr - iC has no w term and a constant shift
    # of w+ only changes Q by a phase, so with zeroMean the densities and the
    # Flory fields are those of the run without it, and w+ differs by a constant
    pots = [dict(kind="Incompress", dt_pl=0.01), dict(kind="Flory", I=0, J=1, chiN=10.0, dt_pl=0.01, dt_mi=0.01)]
    spec = tiny_spec(pots=pots)
    f = random_fields(spec, 0.1, seed=15, imag=False)
    a = orc.fts_system_from_spec(spec, f)
    spec_z = dict(spec, potentials=[dict(pots[0], zero_mean=True), pots[1]])
    b = orc.fts_system_from_spec(spec_z, f)
    orc.fts_init(a)
    orc.fts_init(b)
    for it in range(3):
        z = orc.fts_noise(a.M, 5, it)
        orc.fts_step(a, z)
        orc.fts_step(b, z)
    np.testing.assert_allclose(b.rho, a.rho, rtol=1e-9, atol=1e-12 * np.abs(a.rho).max())
    np.testing.assert_allclose(b.potentials[1].wmi, a.potentials[1].wmi, rtol=1e-9, atol=1e-12)
    d = b.potentials[0].wpl - a.potentials[0].wpl
    np.testing.assert_allclose(d, d.mean(), atol=1e-12)
    assert abs(d.mean() + a.potentials[0].wpl.mean()) < 1e-12


def test_incompress_force_is_field_derivative_of_H():
    # d_makeIncompressForce i rho_tot / Nr - i C (fts_potential_incompress.cu:
    # 202-218) against its H term -i C gvol sum w (:232-247) and -n log Q
    spec = tiny_spec(pots=[dict(kind="Incompress", dt_pl=0.02),
                           dict(kind="Flory", I=0, J=1, chiN=12.0, dt_pl=0.02, dt_mi=0.03)])
    spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], phi=0.6), dict(blocks=[(5, 1)], phi=0.4)]
    _fd_check_forces(spec, random_fields(spec, 0.2, seed=11))


def test_chemical_potential_closed_forms():
    # W = 0: Q = 1, mu = ln(n/V); W = c on every bead: Q = exp(-c Ntot),
    # mu = ln(n/V) + c Ntot; grand canonical: mu = ln z whatever the fields
    c = 0.23
    spec = tiny_spec(pots=[dict(kind="Incompress", dt_pl=0.01)])
    M = int(np.prod(spec["Nx"]))
    s = orc.fts_system_from_spec(spec, {})
    orc.fts_init(s)
    n = s.molecules[0].nmolecs
    assert orc.fts_chemical_potentials(s)[0] == pytest.approx(np.log(n / s.V()), rel=1e-13)
    s = orc.fts_system_from_spec(spec, {(0, "pl"): np.full(M, -1j * c * spec["Nr"])})
    orc.fts_init(s)
    assert orc.fts_chemical_potentials(s)[0] == pytest.approx(np.log(n / s.V()) + c * 7, rel=1e-12)
    spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], activity=0.35)]
    s = orc.fts_system_from_spec(spec, {(0, "pl"): np.full(M, -1j * c * spec["Nr"])})
    orc.fts_init(s)
    assert orc.fts_chemical_potentials(s)[0] == pytest.approx(np.log(0.35), rel=1e-14)
