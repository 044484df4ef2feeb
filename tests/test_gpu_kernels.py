"""GPU parity of the per-kernel entry points (caller-owned arrays in the
reference's layouts) against the FP64 oracle on the same inputs.

Tolerances: k-space tables within 1e-6 normwise relative error (one float
rounding of the same formula); grid forces and energies, which go through an
FP32 FFT, within 1e-5; positions within 1e-6 of the box length (periodic
distance) and the saved positions bit-exact; CL fields, propagators (per bead)
and densities within 1e-5; log Q within 1e-6 * max(1, |log Q|); FP64 sums of
the same float inputs (Q from a given q, Hamiltonian terms) within 1e-9.
rel(a, b) = max|a - b| / max|b|.
"""
import numpy as np
import pytest

import orc
from test_gpu_fts import FIELD_TOL, LOGQ_TOL, logq_err
from test_oracle_fts import random_fields, tiny_spec

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a).astype(np.complex128)
    b = np.asarray(b).astype(np.complex128)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def say(*a):
    print(*a, flush=True)


NX, LX = (12, 20, 16), (5.0, 7.5, 4.0)


# ------------------------------------------------------------------ TILD ----
@pytest.mark.parametrize("kind, A, sigma, Rp", [("gaussian", 3.5, 0.7, 0.0), ("erf2", 2.0, 0.3, 0.8),
                                                ("erfG", 1.5, 0.4, 0.6), ("charges", 0.7, 0.5, 0.0)])
def test_kspace_table_matches_oracle(gpu, kind, A, sigma, Rp):
    say("case: kspace table", kind)
    uk, fk = gpu.kspace_table(kind, A, sigma, Rp, NX, LX)
    uo = orc.uk_table(kind, A, sigma, Rp, NX, LX)
    assert np.all(uk.imag == 0)
    assert rel(uk.real, uo) < 1e-6
    # f_j(k) = -i k_j u(k) (ps_potentialGaussian.cu:42-45), with the float k of Box::get_kD
    kf = np.array([orc.kvec_f(i, NX, LX)[1] for i in range(uo.size)], np.float64)
    assert np.all(fk.real == 0)
    for j in range(3):
        assert rel(fk[:, j].imag, -kf[:, j] * uo) < 1e-6


@pytest.mark.parametrize("factor", [1.0, 2.0])
def test_pair_force_matches_oracle(gpu, factor):
    say("case: pair force, factor", factor)
    rng = np.random.default_rng(11)
    M = int(np.prod(NX))
    rho = (1.0 + 0.3 * rng.random(M)).astype(np.float32)
    F0 = rng.normal(0.0, 0.2, 3 * M).astype(np.float32)
    uk, fk = gpu.kspace_table("gaussian", 4.0, 0.7, 0.0, NX, LX)
    Fg = gpu.pair_force(F0, rho, fk, NX, factor)
    Fo = F0.astype(np.float64)
    orc.potential_grid_force(rho.astype(np.float64), orc.uk_table("gaussian", 4.0, 0.7, 0.0, NX, LX), NX, LX,
                             factor, Fo)
    assert rel(Fg, Fo) < 1e-5
    # the reference's net-force invariant: the convolution force sums to zero
    # (only the random start is left)
    assert np.allclose(Fg.reshape(-1, 3).sum(axis=0), F0.reshape(-1, 3).astype(np.float64).sum(axis=0),
                       atol=1e-3 * np.abs(Fo).max())


def test_potential_energy_matches_oracle(gpu):
    say("case: potential energy")
    rng = np.random.default_rng(12)
    M = int(np.prod(NX))
    rI = (1.0 + 0.3 * rng.random(M)).astype(np.float32)
    rJ = (0.5 + 0.5 * rng.random(M)).astype(np.float32)
    gvol = float(np.prod(np.asarray(LX, np.float64) / np.asarray(NX)))
    for kind, A, s, Rp in (("gaussian", 2.0, 0.6, 0.0), ("erfG", 1.0, 0.3, 0.5)):
        uk, _ = gpu.kspace_table(kind, A, s, Rp, NX, LX)
        eg = gpu.potential_energy(rI, rJ, uk, NX, gvol)
        eo = orc.potential_energy(rI.astype(np.float64), rJ.astype(np.float64), orc.uk_table(kind, A, s, Rp, NX, LX),
                                  NX, gvol)
        assert abs(eg - eo) < 1e-5 * abs(eo), (kind, eg, eo)
    # a uniform density with the Gaussian: gvol * M * rho_I rho_J u(k=0)
    one = np.ones(M, np.float32)
    uk, _ = gpu.kspace_table("gaussian", 2.0, 0.6, 0.0, NX, LX)
    eg = gpu.potential_energy(one, one, uk, NX, gvol)
    assert abs(eg - gvol * M * 2.0) < 1e-5 * gvol * M * 2.0


def _integrator_inputs(n, L, rng):
    L = np.asarray(L, np.float32)
    x = (rng.random((n, 3)) * L).astype(np.float32)
    xo = (x + rng.normal(0, 0.05, (n, 3))).astype(np.float32)
    xo = np.where(xo < 0, xo + L, np.where(xo >= L, xo - L, xo)).astype(np.float32)   # wrapped old positions
    x[:4] = np.float32(1e-4)                                 # pushed below 0 by the force
    x[4:8] = (L - np.float32(1e-4)).astype(np.float32)       # pushed past L
    f = rng.normal(0, 20.0, (n, 3)).astype(np.float32)
    f[:4] = -500.0
    f[4:8] = 500.0
    no = rng.normal(0, 0.1, (n, 3)).astype(np.float32)
    return x.reshape(-1), xo.reshape(-1), no.reshape(-1), f.reshape(-1)


def _pdist(a, b, L):
    d = np.abs(a.reshape(-1, 3).astype(np.float64) - b.reshape(-1, 3))
    L = np.asarray(L, np.float64)
    return np.minimum(d, L - d).max()


@pytest.mark.parametrize("step", [0, 17])
def test_gjf_step_matches_oracle(gpu, step):
    say("case: gjf step", step)
    rng = np.random.default_rng(13 + step)
    n, L, delt, seed = 5000, (5.0, 7.0, 4.0), 0.005, (3 << 32) | 77
    types = rng.integers(0, 2, n).astype(np.int32)
    mass, mob = [1.0, 2.0], [1.0, 0.5]
    x, xo, no, f = _integrator_inputs(n, L, rng)
    xg, xog, nog = gpu.gjf_step(x, xo, no, f, types, mass, mob, delt, L, seed, step)
    coef = np.concatenate([orc.gjf_coeffs(float(np.float32(delt)), m, D) for m, D in zip(mass, mob)])
    xi = orc.tild_noise(n, seed, step)
    X, XO, NO = x.astype(np.float64), xo.astype(np.float64), no.astype(np.float64)
    orc.lib().orc_gjf_step(orc._ptr(X), orc._ptr(XO), orc._ptr(NO), orc._ptr(f.astype(np.float64)), orc._ptr(coef),
                           orc._ptr(types), orc._ptr(np.asarray(L, np.float32)), orc._ptr(np.arange(n, dtype=np.int32)),
                           n, 3, orc._ptr(np.ascontiguousarray(xi, np.float64)))
    assert _pdist(xg, X, L) < 1e-6 * max(L)
    assert np.array_equal(xog, x)
    assert rel(nog, NO) < 1e-5
    assert np.all((xg >= 0) & (xg <= np.tile(np.asarray(L, np.float32), n)))


def test_em_step_matches_oracle(gpu):
    say("case: em step")
    rng = np.random.default_rng(19)
    n, L, delt, seed, step = 5000, (5.0, 7.0, 4.0), 0.005, 99, 4
    x, _, _, f = _integrator_inputs(n, L, rng)
    xg = gpu.em_step(x, f, delt, L, seed, step)
    X = x.astype(np.float64)
    xi = orc.tild_noise(n, seed, step)
    orc.lib().orc_em_step(orc._ptr(X), orc._ptr(f.astype(np.float64)), float(np.float32(delt)),
                          orc._ptr(np.asarray(L, np.float32)), orc._ptr(np.arange(n, dtype=np.int32)), n, 3,
                          orc._ptr(np.ascontiguousarray(xi, np.float64)))
    assert _pdist(xg, X, L) < 1e-6 * max(L)


# ------------------------------------------------------------------- FTS ----
def _fts_case():
    pots = [dict(kind="Helfand", kappaN=30.0, dt_pl=0.02), dict(kind="Flory", I=0, J=1, chiN=12.0, dt_pl=0.02,
                                                                 dt_mi=0.03),
            dict(kind="Flory", I=1, J=2, chiN=8.0, dt_pl=0.01, dt_mi=0.02), dict(kind="Incompress", dt_pl=0.01),
            dict(kind="Edwards", B=1.5, dt_pl=0.01)]
    spec = tiny_spec(Nx=(12, 10, 16), L=(3.0, 2.5, 4.0), pots=pots)
    spec["nspecies"] = 3
    spec["molecules"] = [dict(blocks=[(4, 0), (3, 1), (2, 2)], phi=0.5), dict(blocks=[(2, 1), (3, 2), (2, 1)], phi=0.3),
                         dict(blocks=[(5, 2)], phi=0.2)]
    fields = random_fields(spec, 0.1, seed=31)
    o = orc.fts_system_from_spec(spec, {k: np.asarray(v, np.complex64).astype(np.complex128)
                                        for k, v in fields.items()})
    orc.fts_init(o)
    return spec, o


def _fields(o):
    wpl = [p.wpl for p in o.potentials]
    wmi = [p.wmi for p in o.potentials]
    return wpl, wmi


def test_species_fields_match_oracle(gpu):
    say("case: fts species fields")
    spec, o = _fts_case()
    wpl, wmi = _fields(o)
    Wg = gpu.fts_species_fields(spec["potentials"], wpl, wmi, spec["nspecies"], spec["Nr"])
    for s in range(spec["nspecies"]):
        assert rel(Wg[s], o.W[s]) < 1e-6, s


@pytest.mark.parametrize("cl", [0, 1])
def test_update_fields_match_oracle(gpu, cl):
    say("case: fts update fields, cl", cl)
    spec, o = _fts_case()
    M = o.M
    it = 6
    rho = [r.astype(np.complex64) for r in o.rho]
    o.rho = np.array(rho).astype(np.complex128)
    before = [(p.wpl.copy(), None if p.wmi is None else p.wmi.copy()) for p in o.potentials]
    got = []
    col = 0
    for p, (wp, wm) in zip(spec["potentials"], before):
        got.append(gpu.fts_update_fields(p, wp, wm, rho, o.C, o.Nr, o.Nx, o.L, cl, spec["seed"], it, col))
        col += {"Helfand": 1, "Flory": 2, "Incompress": 0, "Edwards": 1}[p["kind"]]
    z = np.concatenate([orc.fts_noise(M, spec["seed"], it), orc.normal4_np(np.arange(M, dtype=np.uint64), it,
                                                                            orc.TAG_FTS, 1, spec["seed"])], axis=1)
    orc.fts_update_fields(o, z if cl else None)
    for k, (p, (wp, wm)) in enumerate(zip(o.potentials, got)):
        say(f"  potential {k}: w+ error {rel(wp, p.wpl):.1e}"
            + ("" if p.wmi is None else f", w- error {rel(wm, p.wmi):.1e}"))
        assert rel(wp, p.wpl) < FIELD_TOL, k
        if p.wmi is not None:
            assert rel(wm, p.wmi) < FIELD_TOL, k


def test_propagators_density_and_q_match_oracle(gpu):
    say("case: fts propagate / density / Q")
    spec, o = _fts_case()
    extra = orc.fts_molecule_densities(o, keep=True)
    g = gpu.fts_bond_factor(o.Nx, o.L, o.Nr)
    assert rel(g, orc.fts_bond_fft(o.Nx, o.L, o.Nr)) < 1e-6
    W = o.W.astype(np.complex64)
    M = o.M
    rho_g = np.zeros((o.nspecies, M), np.complex64)
    for m, (mol, ex) in enumerate(zip(o.molecules, extra)):
        Ntot = sum(n for n, _ in mol.blocks)
        q, qd = gpu.fts_propagate(W, g, mol.blocks, o.Nx)
        qo = ex["q"].reshape(Ntot, M)
        qdo = ex["qdag"].reshape(Ntot, M)
        eq = max(max(rel(q[s], qo[s]), rel(qd[s], qdo[s])) for s in range(Ntot))
        Qg = gpu.fts_partition_function(q[-1], o.gvol(), o.V())
        e = logq_err(Qg, o.Q[m])
        bud = LOGQ_TOL * max(1.0, abs(np.log(o.Q[m])))
        say(f"  molecule {m}: propagator error {eq:.1e}, log Q error {e:.2e}, bar {bud:.2e}")
        assert eq < FIELD_TOL, m
        assert e < bud
        # the FP64 sum itself, on the oracle's own last propagator
        Qs = gpu.fts_partition_function(qo[-1].astype(np.complex64), o.gvol(), o.V())
        Qs_o = np.sum(qo[-1].astype(np.complex64).astype(np.complex128)) * o.gvol() / o.V()
        assert abs(Qs - Qs_o) < 1e-9 * abs(Qs_o)
        # densities from the oracle's propagators, factor n / (Q V)
        factor = mol.nmolecs / o.Q[m] / o.V()
        rho_g = gpu.fts_density(rho_g, qo.astype(np.complex64), qdo.astype(np.complex64), W, mol.blocks, factor)
    say(f"  densities: error {rel(rho_g, o.rho):.1e}")
    assert rel(rho_g, o.rho) < FIELD_TOL


def test_hamiltonian_terms_match_oracle(gpu):
    say("case: fts hamiltonian terms")
    spec, o = _fts_case()
    _, parts = orc.fts_hamiltonian(o)
    for k, p in enumerate(spec["potentials"]):
        po = o.potentials[k]
        Hg = gpu.fts_hamiltonian_term(p, po.wpl, po.wmi, o.C, o.gvol())
        assert abs(Hg - parts[k]) < 1e-9 * max(abs(parts[k]), 1e-30), (k, Hg, parts[k])


def test_kernel_entry_points_reject_bad_arguments(gpu):
    say("case: argument checks")
    with pytest.raises(gpu.MftError):
        gpu.kspace_table("gaussian", 1.0, 0.5, 0.0, (12, 14, 16), LX)   # 14 is not a supported FFT length
    with pytest.raises(gpu.MftError):
        gpu.fts_update_fields(dict(kind="Flory", I=0, J=1, chiN=10.0, dt_pl=0.01, dt_mi=0.01),
                              np.zeros(8 * 8 * 8), np.zeros(8 * 8 * 8), [np.zeros(8 * 8 * 8)] * 2, 1.0, 10.0,
                              (8, 8, 8), (2.0, 2.0, 2.0), 1, 1, 0, 7)                  # w- column 8 does not exist
