"""The drop-in front end drives the same engine as the C ABI: a reference-
syntax input run through the `mft` executable gives the trajectory and the
logged energies / Hamiltonian of the engine called directly on the same
system (positions within 1e-5 of the box length, logged values within 1e-5
relative; the log prints 6 significant digits)."""
import numpy as np
import pytest

from frontend_util import ps_input, read_data_positions, read_log, run_mft, write_data_file
from mft import systems
from test_oracle_tild import small_system

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def close(a, b, rel=1e-5, atol=1e-6):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= rel * np.abs(b) + atol * max(1.0, np.max(np.abs(b))))


@pytest.mark.parametrize("charged", [False, True])
def test_ps_front_end_matches_engine(gpu, tmp_path, charged):
    cfg = small_system(charged, seed=7)
    say("case: front end ps, charged =", charged)
    write_data_file(tmp_path / "init.data", cfg)
    (tmp_path / "input").write_text(ps_input(cfg, "init.data", steps=12, log_freq=5))
    r = run_mft(["-in", "input"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    n = cfg["topo"]["type"].size
    eng = systems.tild_engine(cfg, "GJF")
    eng.set_run(12)
    # log rows after steps: -1 initial, 0, 5, 10, then the final row (12)
    done = 0
    for row in read_log(tmp_path / "ps_data.dat"):
        k = {-1: 0, 0: 1, 5: 6, 10: 11, 12: 12}[int(row[0])]
        eng.step(k - done)
        done = k
        pe, be = eng.energies()
        assert close(row[1:], [be] + list(pe)), (row, be, pe)
    assert done == 12
    x_fe = read_data_positions(tmp_path / "final.input.data", n)
    x_api, _, _ = eng.get_state()
    d = (x_fe.astype(np.float64) - x_api).reshape(-1, 3)
    L = np.asarray(cfg["L"])
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-5 * max(cfg["L"])
    eng.close()


FTS_INPUT = """box fts cl
Dim 3
grid 16 12 20
boxLengths 3.0 2.5 4.0
Nr 12
C 4
randSeed 9
species A
species B
molecule linear phi 1.0 2 5 A 7 B
potential Helfand 40 0.01 initialize value 0.3 0.0
potential Flory A B 12 0.01 0.01 initialize value 0.1 0.0
logFreq 2
endBox
run vt 5
"""


def test_fts_front_end_matches_engine(gpu, tmp_path):
    say("case: front end fts")
    (tmp_path / "input").write_text(FTS_INPUT)
    r = run_mft(["-in", "input"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    spec = dict(Nx=(16, 12, 20), L=(3.0, 2.5, 4.0), Nr=12.0, C=4.0, nspecies=2, cl=True, seed=9,
                potentials=[dict(kind="Helfand", kappaN=40.0, dt_pl=0.01),
                            dict(kind="Flory", I=0, J=1, chiN=12.0, dt_pl=0.01, dt_mi=0.01)],
                molecules=[dict(blocks=[(5, 0), (7, 1)], phi=1.0)])
    M = 16 * 12 * 20
    # Helfand's initialize never reaches the device; Flory's sets w- and w+ = -1.5i
    fields = {(1, "pl"): np.full(M, -1.5j, np.complex64), (1, "mi"): np.full(M, 0.1, np.complex64)}
    s = systems.fts_system(spec, fields)
    done = 0
    rows = read_log(tmp_path / "fts_data.dat")
    assert [int(r[0]) for r in rows] == [0, 2, 4]
    for row in rows:
        k = int(row[0]) + 1
        s.step(k - done)
        done = k
        Q, H, parts = s.observables()
        want = [H.real, H.imag] + [p.real for p in parts]
        assert close(row[1:], want), (row, want)
    s.close()
