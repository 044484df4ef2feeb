"""Continuous batching in the request loop: CL-FTS solves and TILD runs
admitted at different times, while others are on the device, give the same
results as each run alone (bit for bit: a replica's arithmetic and noise do
not depend on its batch), and admission respects the device-memory budget."""
import numpy as np
import pytest

import mft
from mft import systems
from test_oracle_tild import small_system

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def fts_case(grid, Ns, chiN, seed):
    NA = Ns // 2
    spec, fields = systems.fts_diblock(grid_shape=grid, L=tuple(0.25 * g for g in grid), Nr=float(Ns), NA=NA,
                                       NB=Ns - NA, chiN=chiN, kappaN=60.0, C=4.0, dt=0.01, sigma=0.1, seed=seed)
    return spec, fields


def alone_fts(spec, fields, niter):
    s = systems.fts_system(spec, fields)
    s.step(niter)
    Q, H, parts = s.observables()
    rho = s.densities()
    s.close()
    return Q, H, parts, rho


def test_results_same_alone_and_in_the_mix(gpu):
    cases = [(fts_case((16, 16, 16), 10, 12.0, 1), 20), (fts_case((24, 24, 24), 12, 15.0, 2), 35),
             (fts_case((16, 16, 16), 8, 18.0, 3), 15), (fts_case((32, 32, 32), 10, 11.0, 4), 30),
             (fts_case((24, 24, 24), 12, 9.0, 5), 12)]
    tcfg = small_system(True, seed=3)
    srv = mft.Server(hbm_budget_bytes=4e9, max_fts_replicas=8)
    say("case: serve mix, 5 CL-FTS + 1 TILD")
    ids = [srv.submit_fts(cases[0][0][0], cases[0][0][1], cases[0][1]),
           srv.submit_fts(cases[1][0][0], cases[1][0][1], cases[1][1])]
    srv.pump(10)
    ids.append(srv.submit_fts(cases[2][0][0], cases[2][0][1], cases[2][1]))
    tid = srv.submit_tild(tcfg, 9)
    srv.pump(7)
    ids += [srv.submit_fts(c[0][0], c[0][1], c[1]) for c in cases[3:]]
    for _ in range(200):
        nr, nq = srv.pump(8)
        if nr == 0 and nq == 0:
            break
    assert nr == 0 and nq == 0
    for rid, ((spec, fields), niter) in zip(ids, cases):
        assert srv.status(rid) == "done"
        Q, H, parts, rho = srv.fts_result(rid)
        Qa, Ha, pa, rhoa = alone_fts(spec, fields, niter)
        say("  fts", spec["Nx"], niter, "H", H, Ha)
        assert np.array_equal(Q, Qa) and H == Ha and np.array_equal(parts, pa)
        assert np.array_equal(rho, rhoa)
    x, pe, be = srv.tild_result(tid)
    eng = systems.tild_engine(tcfg, "GJF")
    eng.step(9)
    xa, _, _ = eng.get_state()
    pea, bea = eng.energies()
    eng.close()
    assert np.array_equal(x, xa) and np.array_equal(pe, pea) and be == bea
    srv.close()


def test_admission_respects_budget(gpu):
    (spec, fields) = fts_case((24, 24, 24), 12, 14.0, 7)
    one = 1.2 * 8 * 24 ** 3 * (3 + 4 + 24 + 2 + 2)   # about one replica's buffers
    srv = mft.Server(hbm_budget_bytes=one, max_fts_replicas=4)
    say("case: serve budget")
    a = srv.submit_fts(spec, fields, 10)
    b = srv.submit_fts(spec, fields, 10)
    srv.pump(5)
    assert srv.status(a) == "running" and srv.status(b) == "queued"
    for _ in range(20):
        nr, nq = srv.pump(5)
        if nr == 0 and nq == 0:
            break
    Qa = srv.fts_result(a)[0]
    Qb = srv.fts_result(b)[0]
    assert np.array_equal(Qa, Qb)   # same request, same result, whenever it ran
    srv.close()


def test_runaway_solve_fails_early_without_touching_the_others(gpu):
    # a solve whose propagators overflow ends at the next pump with the
    # reference's NaN message (FTS_Box.cu:148-152) instead of holding its slot
    # for all its iterations; the solve beside it is unaffected
    say("case: serve runaway solve")
    spec, fields = fts_case((16, 16, 16), 10, 12.0, 7)
    bad_fields = dict(fields)
    M = 16 ** 3
    bad_fields[(1, "mi")] = np.full(M, 1.0e4, np.complex64)   # W = -/+200: exp(-W) overflows at once
    good, bad = fts_case((16, 16, 16), 10, 12.0, 8), (spec, bad_fields)
    srv = mft.Server(hbm_budget_bytes=4e9, max_fts_replicas=8)
    gid = srv.submit_fts(good[0], good[1], 60)
    bid = srv.submit_fts(bad[0], bad[1], 3000)
    pumps = 0
    while srv.status(bid) not in ("done", "failed"):
        srv.pump(10)
        pumps += 1
        assert pumps < 20
    assert srv.status(bid) == "failed"
    err = srv.error(bid)
    assert "Found Heff = NaN" in err and "not finite after" in err
    assert int(err.split("after ")[1].split()[0]) < 3000
    for _ in range(50):
        nr, nq = srv.pump(10)
        if nr == 0 and nq == 0:
            break
    assert srv.status(gid) == "done"
    Q, H, parts, rho = srv.fts_result(gid)
    Qa, Ha, pa, rhoa = alone_fts(good[0], good[1], 60)
    assert np.array_equal(Q, Qa) and H == Ha and np.array_equal(rho, rhoa)
    srv.close()


def test_served_results_match_oracle(gpu):
    """Every result of a served mix against the FP64 oracle: each CL-FTS
    request is replayed by the oracle from its initial fields and its own
    seed's noise (log Q, mu, H, densities and the final fields), its final
    fields are re-evaluated by the oracle's density pass (the check used for
    long served solves), and the TILD request's positions are replayed."""
    import orc
    from test_gpu_fts import compare, rel

    say("case: served results vs oracle")
    cases = [(fts_case((32, 32, 32), 20, 14.0, 11), 3), (fts_case((48, 48, 48), 16, 19.0, 12), 2),
             (fts_case((32, 32, 32), 24, 9.0, 13), 4)]
    tcfg = small_system(True, seed=9)
    srv = mft.Server(hbm_budget_bytes=8e9, max_fts_replicas=8)
    ids = [srv.submit_fts(cases[0][0][0], cases[0][0][1], cases[0][1])]
    tid = srv.submit_tild(tcfg, 3)
    srv.pump(1)
    ids += [srv.submit_fts(c[0][0], c[0][1], c[1]) for c in cases[1:]]
    for _ in range(50):
        nr, nq = srv.pump(2)
        if nr == 0 and nq == 0:
            break
    assert nr == 0 and nq == 0

    class Served:
        def __init__(self, rid):
            self.r = srv.fts_result(rid)
            self.f, self.mu = srv.fts_final_state(rid)

        def observables(self):
            return self.r[:3]

        def densities(self):
            return self.r[3]

        def chemical_potentials(self):
            return self.mu

        def get_fields(self, k, minus=False):
            return (self.f[(k, "pl")], self.f[(k, "mi")]) if minus else self.f[(k, "pl")]

    for rid, ((spec, fields), niter) in zip(ids, cases):
        assert srv.status(rid) == "done"
        say("  request", spec["Nx"], niter)
        o = orc.fts_system_from_spec(spec, {k: np.asarray(v, np.complex64).astype(np.complex128)
                                            for k, v in fields.items()})
        orc.fts_init(o)
        for it in range(niter):
            orc.fts_step(o, orc.fts_noise(o.M, spec["seed"], it))
        sv = Served(rid)
        compare(sv, o, spec, f"served {spec['Nx']} replay")
        # the final-state check: the oracle's density pass on the served fields
        f = orc.fts_system_from_spec(spec, {k: np.asarray(v).astype(np.complex128) for k, v in sv.f.items()})
        orc.fts_init(f)
        compare(sv, f, spec, f"served {spec['Nx']} final state")
    x, pe, be = srv.tild_result(tid)
    from test_oracle_tild import oracle_system
    sys_o = oracle_system(tcfg, "GJF")
    ns = tcfg["topo"]["type"].size
    for step in range(3):
        out = orc.tild_step(sys_o, orc.tild_noise(ns, tcfg["seed"], step))
        sys_o.x, sys_o.xo, sys_o.noise_old = (out[k].astype(np.float32) for k in ("x", "xo", "noise_old"))
    d = (x.astype(np.float64) - sys_o.x.astype(np.float64)).reshape(-1, 3)
    L = np.asarray(tcfg["L"], np.float64)
    d -= L * np.round(d / L)
    say(f"  tild request: max |dx| {np.max(np.abs(d)):.2e}")
    assert np.max(np.abs(d)) < 1e-6 * max(tcfg["L"])
    out = orc.tild_forces(sys_o, want_energy=True)
    assert rel(pe, np.asarray(out["energies"])) < 1e-5
    srv.close()
