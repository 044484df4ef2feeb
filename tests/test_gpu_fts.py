"""GPU parity of the CL-FTS path: the HIP engine through the C ABI against
the FP64 oracle on the same seeded fields and noise.

Tolerances (complex FP32 fields, FP64 reductions), fixed in advance and never
widened by the data: fields and densities within 1e-5 normwise relative
error, rel(a, b) = max|a - b| / max|b|; log Q within 1e-6 * max(1, |log Q|)
of the FP64 oracle.  H = sum_i H_i carries -n log Q, so its error budget is
1e-6 * sum_i |H_i| plus n times the log Q budget.  Each comparison prints the
device error next to that of the FP32 host build of the same oracle.
Trajectories are compared step by step from identical state and noise (CL
dynamics is chaotic, so long runs are compared by invariants).
"""
import numpy as np
import pytest

import orc
from mft import systems
from test_oracle_fts import random_fields, tiny_spec

pytestmark = pytest.mark.gpu

FIELD_TOL = 1e-5
LOGQ_TOL = 1e-6


def rel(a, b):
    a = np.asarray(a).astype(np.complex128)
    b = np.asarray(b).astype(np.complex128)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def say(*a):
    print(*a, flush=True)


def logq_err(Qg, Qo):
    lg, lo = np.log(Qg), np.log(Qo)
    return abs(lg - lo) / max(1.0, abs(lo))


def build(gpu, spec, fields):
    s = gpu.FtsSystem(spec)
    systems.load_fields(s.set_fields, spec, {k: np.asarray(v, np.complex64) for k, v in fields.items()})
    s.init()
    o = orc.fts_system_from_spec(spec, {k: np.asarray(v, np.complex64).astype(np.complex128) for k, v in fields.items()})
    orc.fts_init(o)
    return s, o


def compare(s, o, spec, what="", o32=None):
    """Device state s (FtsSystem or one replica of a batch) against the oracle o:
    log Q, chemical potentials, H, densities, species fields and potential
    fields; every error is printed before the asserts."""
    Qg, Hg, parts_g = s.observables()
    mu_g = s.chemical_potentials()
    mu_o = orc.fts_chemical_potentials(o)
    qbud, errs = [], []
    for m in range(len(spec["molecules"])):
        e = logq_err(Qg[m], o.Q[m])
        e32 = logq_err(o32.Q[m], o.Q[m]) if o32 is not None else float("nan")
        bud = LOGQ_TOL * max(1.0, abs(np.log(o.Q[m])))
        qbud.append(bud)
        # mu = ln(n/V) - ln Q carries the log Q error itself, so it shares the bar
        emu = abs(mu_g[m] - mu_o[m])
        mbud = LOGQ_TOL * max(1.0, abs(np.log(o.Q[m])), abs(mu_o[m]))
        say(f"  {what}: log Q error device {e:.2e}, FP32 host build {e32:.2e}, bar {bud:.2e}; "
            f"mu error {emu:.2e} (bar {mbud:.2e})")
        errs += [(e < bud, f"log Q of molecule {m}"), (emu <= mbud, f"mu of molecule {m}")]
    Ho, parts_o = orc.fts_hamiltonian(o)
    budget = 1e-6 * sum(abs(h) for h in parts_o)
    # -n log Q moves by n dlogQ; the grand canonical -z V Q by z V |Q| dlogQ
    budget += sum((m.nmolecs if m.nmolecs > 0 else m.activity * o.V() * abs(Q)) * b
                  for m, Q, b in zip(o.molecules, o.Q, qbud))
    errs.append((abs(Hg - Ho) <= budget, f"H {Hg} vs {Ho}, budget {budget:.2e}"))
    fe = {"densities": rel(s.densities(), o.rho)}
    if hasattr(s, "species_fields"):   # a served result carries the potential fields only
        fe["species fields"] = rel(s.species_fields(), o.W)
    for k, p in enumerate(spec["potentials"]):
        if p["kind"] == "Flory":
            wp, wm = s.get_fields(k, minus=True)
            fe[f"w- of potential {k}"] = rel(wm, o.potentials[k].wmi)
        else:
            wp = s.get_fields(k)
        fe[f"w+ of potential {k}"] = rel(wp, o.potentials[k].wpl)
    say(f"  {what}: |dH| {abs(Hg - Ho):.2e} (budget {budget:.2e}); "
        + ", ".join(f"{n} {v:.1e}" for n, v in fe.items()))
    errs += [(v < FIELD_TOL, n) for n, v in fe.items()]
    bad = [n for ok, n in errs if not ok]
    assert not bad, (what, bad)


@pytest.mark.parametrize("blocks", [((4, 0), (3, 1)), ((2, 0), (3, 1), (2, 0)), ((6, 1),), ((1, 0),), ((1, 0), (1, 1)),
                                    ((1, 0), (5, 1), (1, 0), (2, 1))])
def test_init_matches_oracle(gpu, blocks):
    spec = tiny_spec(Nx=(16, 12, 20), L=(4.0, 3.0, 5.0), blocks=blocks)
    say("case: fts init", blocks)
    f = random_fields(spec, 0.2, seed=3)
    s, o = build(gpu, spec, f)
    o32 = orc.fts_system_from_spec(spec, {k: np.asarray(v, np.complex64).astype(np.complex128) for k, v in f.items()})
    orc.fts_init(o32, bits=32)
    compare(s, o, spec, "init", o32)
    # propagators in the reference's q-form
    q, qd = s.propagators(0)
    ref = orc.fts_molecule_densities(o, keep=True)[0]
    assert rel(q, ref["q"]) < FIELD_TOL
    assert rel(qd, ref["qdag"]) < FIELD_TOL
    s.close()


def test_closed_forms_on_device(gpu):
    say("case: fts closed forms")
    spec = tiny_spec(Nx=(16, 16, 16), L=(4.0, 4.0, 4.0), pots=[dict(kind="Incompress", dt_pl=0.01)])
    M = 16 ** 3
    s = gpu.FtsSystem(spec)
    s.init()
    Q, H, _ = s.observables()
    assert abs(Q[0] - 1.0) < 1e-6
    c = 0.21
    s.set_fields(0, np.full(M, -1j * c * spec["Nr"], np.complex64))
    s.init()
    Q, H, _ = s.observables()
    assert abs(np.log(Q[0]) - (-c * 7)) < 1e-6 * 7 * c
    s.close()


@pytest.mark.parametrize("case", ["diblock", "diblock_square", "incompress_blend", "diblock_empc",
                                  "blend_empc_mixed", "diblock_1s", "blend_schemes_mixed", "edwards_diblock",
                                  "edwards_1s", "grand_canonical_blend", "grand_canonical_1s", "zero_mean_em",
                                  "zero_mean_empc_mixed", "zero_mean_1s"])
def test_cl_steps_match_oracle(gpu, case):
    say("case: fts CL steps", case)
    if case == "diblock":
        spec, fields = systems.fts_diblock(grid_shape=(16, 24, 12), L=(3.0, 4.0, 2.5), Nr=20.0, NA=8, NB=12,
                                           chiN=15.0, kappaN=50.0, C=5.0, dt=0.01, sigma=0.1, seed=21)
    elif case == "diblock_square":
        # square x-y planes take the fused plane kernel
        spec, fields = systems.fts_diblock(grid_shape=(20, 20, 12), L=(3.5, 3.5, 2.5), Nr=20.0, NA=8, NB=12,
                                           chiN=15.0, kappaN=50.0, C=5.0, dt=0.01, sigma=0.1, seed=23)
    elif case == "diblock_empc":
        # predictor-corrector on both potentials (fts_updateSchemes.cu:56-81)
        spec, fields = systems.fts_diblock(grid_shape=(16, 24, 12), L=(3.0, 4.0, 2.5), Nr=20.0, NA=8, NB=12,
                                           chiN=15.0, kappaN=50.0, C=5.0, dt=0.01, sigma=0.1, seed=27)
        spec["potentials"] = [dict(p, scheme="EMPC") for p in spec["potentials"]]
    elif case == "diblock_1s":
        # one-step k-space scheme with the Debye coefficients (no noise, as the
        # reference's d_fts_update1S)
        spec, fields = systems.fts_diblock(grid_shape=(16, 24, 12), L=(3.0, 4.0, 2.5), Nr=20.0, NA=8, NB=12,
                                           chiN=15.0, kappaN=50.0, C=5.0, dt=0.01, sigma=0.1, seed=29)
        spec["potentials"] = [dict(p, scheme="1S") for p in spec["potentials"]]
    elif case == "blend_schemes_mixed":
        # 1S incompressibility, EMPC Flory and a homopolymer (the Flory A+
        # takes the homopolymer Debye term)
        spec = tiny_spec(Nx=(12, 16, 20), L=(3.0, 4.0, 5.0), pots=[dict(kind="Incompress", dt_pl=0.01, scheme="1S"),
                                                                dict(kind="Flory", I=0, J=1, chiN=10.0, dt_pl=0.01,
                                                                     dt_mi=0.01, scheme="EMPC")])
        spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], phi=0.6), dict(blocks=[(5, 1)], phi=0.4)]
        fields = random_fields(spec, 0.1, seed=10, imag=False)
    elif case in ("edwards_diblock", "edwards_1s"):
        # Edwards excluded volume (fts_potential_edwards.cu) beside Flory
        spec, fields = systems.fts_diblock(grid_shape=(16, 24, 12), L=(3.0, 4.0, 2.5), Nr=20.0, NA=8, NB=12,
                                           chiN=15.0, kappaN=50.0, C=5.0, dt=0.01, sigma=0.1, seed=31)
        spec["potentials"][0] = dict(kind="Edwards", B=2.0, dt_pl=0.01)
        if case == "edwards_1s":
            spec["potentials"] = [dict(p, scheme="1S") for p in spec["potentials"]]
    elif case in ("zero_mean_em", "zero_mean_1s"):
        # modify zeroMean on Helfand / Edwards (Flory parses it without effect)
        spec, fields = systems.fts_diblock(grid_shape=(16, 24, 12), L=(3.0, 4.0, 2.5), Nr=20.0, NA=8, NB=12,
                                           chiN=15.0, kappaN=50.0, C=5.0, dt=0.01, sigma=0.1, seed=33)
        spec["potentials"] = [dict(p, zero_mean=True) for p in spec["potentials"]]
        if case == "zero_mean_1s":
            spec["potentials"][0] = dict(kind="Edwards", B=2.0, dt_pl=0.01, zero_mean=True)
            spec["potentials"] = [dict(p, scheme="1S") for p in spec["potentials"]]
    elif case == "zero_mean_empc_mixed":
        # a zero-mean EMPC incompressibility is re-centred after the corrector too
        spec = tiny_spec(Nx=(12, 16, 20), L=(3.0, 4.0, 5.0), pots=[
            dict(kind="Incompress", dt_pl=0.01, scheme="EMPC", zero_mean=True),
            dict(kind="Helfand", kappaN=40.0, dt_pl=0.01, zero_mean=True),
            dict(kind="Flory", I=0, J=1, chiN=10.0, dt_pl=0.01, dt_mi=0.01)])
        spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], phi=0.6), dict(blocks=[(5, 1)], phi=0.4)]
        fields = random_fields(spec, 0.1, seed=16, imag=False)
    elif case in ("grand_canonical_blend", "grand_canonical_1s"):
        # a grand canonical homopolymer beside a canonical diblock
        # (fts_molecule.cu:30-31, fts_molecule_linear.cu:286-291,374-376)
        pots = [dict(kind="Helfand", kappaN=30.0, dt_pl=0.01),
                dict(kind="Flory", I=0, J=1, chiN=10.0, dt_pl=0.01, dt_mi=0.01)]
        if case == "grand_canonical_1s":
            pots = [dict(p, scheme="1S") for p in pots]
        spec = tiny_spec(Nx=(12, 16, 20), L=(3.0, 4.0, 5.0), pots=pots)
        spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], phi=0.6), dict(blocks=[(5, 1)], activity=0.05)]
        fields = random_fields(spec, 0.1, seed=12, imag=False)
    elif case == "blend_empc_mixed":
        # EMPC Flory beside an EM incompressibility: the EM field keeps its
        # predictor value through the corrector pass
        spec = tiny_spec(Nx=(12, 16, 20), L=(3.0, 4.0, 5.0), pots=[dict(kind="Incompress", dt_pl=0.01),
                                                                dict(kind="Flory", I=0, J=1, chiN=10.0, dt_pl=0.01,
                                                                     dt_mi=0.01, scheme="EMPC")])
        spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], phi=0.6), dict(blocks=[(5, 1)], phi=0.4)]
        fields = random_fields(spec, 0.1, seed=9, imag=False)
    else:
        spec = tiny_spec(Nx=(12, 16, 20), L=(3.0, 4.0, 5.0), pots=[dict(kind="Incompress", dt_pl=0.01),
                                                                dict(kind="Flory", I=0, J=1, chiN=10.0, dt_pl=0.01,
                                                                     dt_mi=0.01)])
        spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], phi=0.6), dict(blocks=[(5, 1)], phi=0.4)]
        fields = random_fields(spec, 0.1, seed=8, imag=False)
    s, o = build(gpu, spec, fields)
    compare(s, o, spec, "init")
    M = int(np.prod(spec["Nx"]))
    for it in range(3):
        s.step(1)
        orc.fts_step(o, orc.fts_noise(M, spec["seed"], it))
        compare(s, o, spec, f"step {it}")
    s.close()


def test_batch_equals_alone(gpu):
    """Replicas of uneven grids, chain lengths and update schemes advanced
    together give the same bits as each replica alone."""
    say("case: fts batch vs alone")
    reps = []
    for k, (g, L, NA, NB) in enumerate([((24, 24, 24), 5.0, 10, 10), ((16, 16, 16), 4.0, 6, 4),
                                         ((24, 24, 24), 6.0, 12, 8), ((32, 32, 32), 7.0, 5, 15),
                                         ((24, 24, 24), 5.5, 9, 11)]):
        reps.append(systems.fts_diblock(grid_shape=g, L=L, Nr=float(NA + NB), NA=NA, NB=NB, chiN=12.0, kappaN=60.0,
                                        C=4.0, dt=0.01, sigma=0.1, seed=100 + k))
    # the last replica uses the predictor-corrector: a group of its own
    reps[-1][0]["potentials"] = [dict(p, scheme="EMPC") for p in reps[-1][0]["potentials"]]
    b = gpu.FtsBatch([r[0] for r in reps])
    for k, (spec, fields) in enumerate(reps):
        systems.load_fields(b.set_fields, spec, fields, rep=k)
    b.init()
    b.step(4)
    for k, (spec, fields) in enumerate(reps):
        s = systems.fts_system(spec, fields)
        s.step(4)
        Qa, Ha, _ = s.observables()
        Qb, Hb, _ = b.observables(k)
        assert Qa[0] == Qb[0] and Ha == Hb, k
        assert np.array_equal(s.densities(), b.densities(k)), k
        s.close()
    b.close()


def test_chunked_propagation_and_autotune_keep_the_bits(gpu):
    """Propagating a group's replicas chunk by chunk (here one replica per
    chunk, an EMPC group included) and the autotuner's timing passes leave
    every replica's trajectory bit for bit the same as one chunk per group."""
    say("case: fts chunked propagation")
    reps = []
    for k, (g, L, NA, NB) in enumerate([((24, 24, 24), 5.0, 10, 10), ((24, 24, 24), 6.0, 12, 8),
                                         ((24, 24, 24), 5.5, 9, 11), ((32, 32, 32), 7.0, 5, 15),
                                         ((32, 32, 32), 6.5, 8, 8)]):
        reps.append(systems.fts_diblock(grid_shape=g, L=L, Nr=float(NA + NB), NA=NA, NB=NB, chiN=12.0, kappaN=60.0,
                                        C=4.0, dt=0.01, sigma=0.1, seed=200 + k))
    reps[-1][0]["potentials"] = [dict(p, scheme="EMPC") for p in reps[-1][0]["potentials"]]
    batches = []
    for mode in ("one", "chunked", "tuned"):
        b = gpu.FtsBatch([r[0] for r in reps])
        for k, (spec, fields) in enumerate(reps):
            systems.load_fields(b.set_fields, spec, fields, rep=k)
        b.init()
        if mode == "chunked":
            b.set_chunk_bytes(1 << 20)    # below one replica: a chunk each
        elif mode == "tuned":
            b.step(1)
            assert b.autotune(2) > 0.0
            b.step(1)
        if mode != "tuned":
            b.step(2)
        b.step(2)
        batches.append(b)
    for k in range(len(reps)):
        Q0, H0, _ = batches[0].observables(k)
        for b in batches[1:]:
            Q1, H1, _ = b.observables(k)
            assert Q0[0] == Q1[0] and H0 == H1, k
            assert np.array_equal(batches[0].densities(k), b.densities(k)), k
    for b in batches:
        b.close()


def test_config3_size_invariants(gpu):
    """64^3, 100-bead diblock: mass conservation and finite H over a short run."""
    say("case: fts 64^3 Ns=100 invariants")
    spec, fields = systems.fts_diblock()
    s = systems.fts_system(spec, fields)
    s.step(5)
    Q, H, parts = s.observables()
    assert np.isfinite(H)
    rho = s.densities()
    L = spec["L"]
    gvol = np.prod([float(np.float32(l)) / n for l, n in zip(L, spec["Nx"])])
    V = np.prod([float(np.float32(l)) for l in L])
    n = spec["C"] * 1.0 * V * spec["Nr"] / 100
    assert rho.astype(np.complex128).sum() * gvol == pytest.approx(n * 100, rel=1e-5)
    s.close()


def test_config3_log_q_against_oracle(gpu):
    """64^3, 100-bead diblock at the benchmark shape: log Q and densities of
    the initial state against the FP64 oracle (and the FP32 host build)."""
    say("case: fts 64^3 Ns=100 log Q vs oracle")
    spec, fields = systems.fts_diblock()
    s = systems.fts_system(spec, fields)
    f64 = {k: np.asarray(v, np.complex64).astype(np.complex128) for k, v in fields.items()}
    o = orc.fts_system_from_spec(spec, f64)
    orc.fts_init(o)
    o32 = orc.fts_system_from_spec(spec, f64)
    orc.fts_init(o32, bits=32)
    compare(s, o, spec, "C3 init", o32)
    s.close()


@pytest.mark.parametrize("grid", [(16, 16, 12), (20, 20, 10), (24, 24, 16), (32, 32, 12), (40, 40, 8),
                                  (48, 48, 8), (64, 64, 10), (96, 96, 6)])
def test_plane_kernel_matches_line_kernels(gpu, grid, monkeypatch):
    """The fused x-y plane kernel against the separate x and y line kernels
    (same arithmetic) over a few CL steps."""
    say("case: fts plane kernel", grid)
    spec, fields = systems.fts_diblock(grid_shape=grid, L=(0.12 * grid[0], 0.12 * grid[1], 0.12 * grid[2]),
                                       Nr=20.0, NA=8, NB=12, chiN=15.0, kappaN=50.0, C=5.0, dt=0.01, sigma=0.1,
                                       seed=31)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("MFT_FTS_PLANE_XY", flag)
        s = systems.fts_system(spec, fields)
        s.step(3)
        Q, H, _ = s.observables()
        out.append((Q[0], H, s.densities()))
        s.close()
    (Q0, H0, r0), (Q1, H1, r1) = out
    say(f"  bitwise equal: {np.array_equal(r0, r1)}, rel {rel(r1, r0):.2e}")
    assert rel(r1, r0) < 1e-6
    assert abs(np.log(Q1) - np.log(Q0)) < 1e-6 * max(1.0, abs(np.log(Q0)))


def test_graph_replay_matches_direct_launches(gpu, monkeypatch):
    """CL iterations replayed from a captured HIP graph against the same
    iterations launched directly: every replay, with host reads and
    synchronisations between replays, gives the same bits."""
    say("case: fts graph replay vs direct launches")
    reps = [systems.fts_diblock(grid_shape=g, L=0.12 * g[0], Nr=20.0, NA=8, NB=12, chiN=12.0, kappaN=50.0, C=5.0,
                                dt=0.01, sigma=0.1, seed=41 + k)
            for k, g in enumerate([(24, 24, 24), (32, 32, 16)])]
    runs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("MFT_FTS_GRAPH", flag)
        b = gpu.FtsBatch([r[0] for r in reps])
        for k, (spec, fields) in enumerate(reps):
            systems.load_fields(b.set_fields, spec, fields, rep=k)
        b.init()
        seq = []
        for n in (1, 1, 2, 1, 3):
            b.step(n)
            seq.append([(b.observables(k)[0][0], b.observables(k)[1], b.densities(k)) for k in range(len(reps))])
        runs.append(seq)
        b.close()
    for i, (d, g) in enumerate(zip(*runs)):
        for k in range(len(reps)):
            assert d[k][0] == g[k][0] and d[k][1] == g[k][1], (i, k)
            assert np.array_equal(d[k][2], g[k][2]), (i, k)


def test_activity_change_takes_effect_after_one_density_pass(gpu):
    # modify molecule 1 activity scale f (fts_molecule_linear.cu:386-412): the
    # next update still uses the densities of the old activity
    say("case: fts activity change")
    spec = tiny_spec(Nx=(12, 16, 20), L=(3.0, 4.0, 5.0))
    spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], activity=0.04)]
    fields = random_fields(spec, 0.1, seed=13, imag=False)
    s, o = build(gpu, spec, fields)
    M = int(np.prod(spec["Nx"]))
    it = 0
    for _ in range(2):
        s.step(1)
        orc.fts_step(o, orc.fts_noise(M, spec["seed"], it))
        it += 1
    s.set_activity(0, 0.04 * 0.8)
    o.molecules[0].activity = 0.04 * 0.8
    compare(s, o, spec, "after the change")
    n_g = s.nsites()[0]
    for k in range(3):
        s.step(1)
        orc.fts_step(o, orc.fts_noise(M, spec["seed"], it))
        it += 1
        compare(s, o, spec, f"step {k} after the change")
    # nSites = z V Re Q Ntot (fts_molecule_linear.cu:344-346)
    n_o = o.molecules[0].activity * o.V() * o.Q[0].real * 7
    assert abs(s.nsites()[0] - n_o) < 1e-5 * abs(n_o)
    assert n_g != s.nsites()[0]
    s.close()


def c4_replicas():
    """Replicas drawn like the batched sweep (box 4-10 Rg, chi N in [10, 20],
    f_A in [0.3, 0.5], kappa N = 100, C = 50) at the sweep's large shapes:
    48^3 and 64^3 grids with 50- and 100-bead chains, beside two small ones."""
    rng = np.random.default_rng(404)
    out = []
    for k, (g, Ns) in enumerate([(64, 100), (48, 50), (64, 50), (48, 100), (24, 100), (32, 50)]):
        Lr, chiN, fA = rng.uniform(4.0, 10.0), rng.uniform(10.0, 20.0), rng.uniform(0.3, 0.5)
        NA = int(round(fA * Ns))
        out.append(systems.fts_diblock(grid=g, L=float(Lr), Nr=float(Ns), NA=NA, NB=Ns - NA, chiN=float(chiN),
                                       seed=9000 + k))
    return out


def test_config4_batch_matches_oracle(gpu):
    """The batched path at the sweep's shapes against the FP64 oracle: every
    replica of one batch at the initial density pass and after two CL
    iterations (log Q, mu, H, densities, species and potential fields)."""
    say("case: fts batch at sweep shapes vs oracle")
    reps = c4_replicas()
    b = gpu.FtsBatch([r[0] for r in reps])
    for k, (spec, fields) in enumerate(reps):
        systems.load_fields(b.set_fields, spec, fields, rep=k)
    b.init()
    oracles = []
    for k, (spec, fields) in enumerate(reps):
        o = orc.fts_system_from_spec(spec, {key: np.asarray(v, np.complex64).astype(np.complex128)
                                            for key, v in fields.items()})
        orc.fts_init(o)
        oracles.append(o)
        compare(b.replica(k), o, spec, f"replica {k} {spec['Nx'][0]}^3 Ns={spec['Nr']:g} init")
    for it in range(2):
        b.step(1)
        for k, (spec, _) in enumerate(reps):
            o = oracles[k]
            orc.fts_step(o, orc.fts_noise(o.M, spec["seed"], it))
            compare(b.replica(k), o, spec, f"replica {k} {spec['Nx'][0]}^3 Ns={spec['Nr']:g} step {it}")
    b.close()


def test_snapshot_restore_returns_to_the_saved_state(gpu):
    """Device snapshots of a sweep: fields and activities saved in a slot,
    changed by iterations and an activity change, then restored; the density
    pass of the restored fields gives the saved observables bit for bit."""
    say("case: fts snapshot / restore")
    spec = tiny_spec(Nx=(12, 16, 20), L=(3.0, 4.0, 5.0))
    spec["molecules"] = [dict(blocks=[(4, 0), (3, 1)], phi=0.6), dict(blocks=[(5, 1)], activity=0.05)]
    s = systems.fts_system(spec, random_fields(spec, 0.1, seed=17, imag=False))
    s.step(3)
    Q0, H0, _ = s.observables()
    rho0 = s.densities()
    w0 = s.get_fields(1, minus=True)
    s.snapshot(2)
    s.step(2)
    s.set_activity(1, 0.03)
    s.step(1)
    assert s.observables()[1] != H0
    s.restore(2)
    Q1, H1, _ = s.observables()
    assert np.array_equal(Q0, Q1) and H0 == H1
    assert np.array_equal(rho0, s.densities())
    w1 = s.get_fields(1, minus=True)
    assert np.array_equal(w0[0], w1[0]) and np.array_equal(w0[1], w1[1])
    s.close()
