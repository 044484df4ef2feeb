"""The drop-in front end drives the same engine as the C ABI: a reference-
syntax input run through the `mft` executable gives the trajectory and the
logged energies / Hamiltonian of the engine called directly on the same
system (positions within 1e-5 of the box length, logged values within 1e-5
relative; the log prints 6 significant digits)."""
import ctypes as C
import os

import numpy as np
import pytest

import orc

from frontend_util import ps_input, read_data_positions, read_log, run_mft, write_data_file
from mft import systems
from test_oracle_tild import oracle_system, small_system

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
    # log rows after steps: -1 initial, 0, 5, 10, then the final row (12);
    # a row after a step holds the pair energies of that step's density pass
    # and the bond energy at the positions the step ends with
    rows = read_log(tmp_path / "ps_data.dat")
    assert [int(r[0]) for r in rows] == [-1, 0, 5, 10, 12]
    done = 0
    for row in rows:
        k = {-1: 0, 0: 1, 5: 6, 10: 11, 12: 12}[int(row[0])]
        pe, be = eng.step_energies(k - done) if k > done else eng.energies()
        done = k
        assert close(row[1:], [be] + list(pe)), (row, be, pe)
    assert done == 12
    # the same rows from the FP64 oracle in the reference's order
    # (PS_Box.cu:73-207: densities, forces, Integrate_2, update_prefactor,
    # writeData -> CalcEnergy on the step's d_rho, bonds at the new positions)
    sys_o = oracle_system(cfg, "GJF")
    topo = sys_o.topo

    def bond_e():
        eb = orc.bond_energy(dict(nBonds=topo["nBonds"], bondedTo=topo["bondedTo"], bondType=topo["bondType"],
                                  bondReq=cfg["bond_r0"], bondK=cfg["bond_k"], bondStyle=cfg["bond_style"],
                                  maxbonds=2), sys_o.x, cfg["L"])
        return 0.5 * eb.sum()

    want = {-1: [bond_e()] + orc.tild_forces(sys_o, want_energy=True)["energies"]}
    for step in range(12):
        pair = orc.tild_forces(sys_o, want_energy=True)["energies"]
        out = orc.tild_step(sys_o, orc.tild_noise(n, cfg["seed"], step))
        sys_o.x, sys_o.xo, sys_o.noise_old = (out[key].astype(np.float32) for key in ("x", "xo", "noise_old"))
        if step in (0, 5, 10, 11):
            want[12 if step == 11 else step] = [bond_e()] + pair
    for row in rows:
        w = want[int(row[0])]
        say(f"  row {int(row[0])}: front end {row[1:]}, oracle {np.round(w, 6)}")
        assert close(row[1:], w), (row, w)
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
    mus = read_log(tmp_path / "fts_mu.dat")
    assert [int(r[0]) for r in rows] == [0, 2, 4] and [int(r[0]) for r in mus] == [0, 2, 4]
    for row, mrow in zip(rows, mus):
        k = int(row[0]) + 1
        s.step(k - done)
        done = k
        Q, H, parts = s.observables()
        want = [H.real, H.imag] + [p.real for p in parts]
        assert close(row[1:], want), (row, want)
        mu = s.chemical_potentials()
        assert np.allclose(mrow[1:], [mu[0].real, mu[0].imag], rtol=1e-9, atol=1e-9), (mrow, mu)
    s.close()


SPINODAL_INPUT = """box fts cl
Dim 3
grid 12 12 16
boxLengths 3.0 3.0 4.0
Nr 12
C 4
randSeed 5
species A
species B
molecule linear phi 0.7 2 5 A 7 B
molecule linear activity 0.02 1 6 B
potential Edwards 3.0 0.01
potential Flory A B 12 0.01 0.01 initialize value 0.1 0.0
logFreq 4
endBox
run vt 4
modify molecule 2 activity scale 0.5
run vt 4
run findspinodal molecule 2 nSteps 3 initial_f 0.9 maxTries 4 datFile fs.dat
"""


def test_modify_and_find_spinodal_match_engine(gpu, tmp_path):
    # modifyBox (FTS_Box.cu:674-708) and findSpinodal (FTS_BoxFindSpinodal.cu)
    # through the front end against the same sequence of engine calls
    say("case: front end modify + findSpinodal")
    (tmp_path / "input").write_text(SPINODAL_INPUT)
    r = run_mft(["-in", "input"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    spec = dict(Nx=(12, 12, 16), L=(3.0, 3.0, 4.0), Nr=12.0, C=4.0, nspecies=2, cl=True, seed=5,
                potentials=[dict(kind="Edwards", B=3.0, dt_pl=0.01),
                            dict(kind="Flory", I=0, J=1, chiN=12.0, dt_pl=0.01, dt_mi=0.01)],
                molecules=[dict(blocks=[(5, 0), (7, 1)], phi=0.7), dict(blocks=[(6, 1)], activity=0.02)])
    M = 12 * 12 * 16
    fields = {(1, "pl"): np.full(M, -1.5j, np.complex64), (1, "mi"): np.full(M, 0.1, np.complex64)}
    s = systems.fts_system(spec, fields)
    rows = read_log(tmp_path / "fts_data.dat")
    # run vt 4 logs step 0; the grand canonical molecule adds -zVQ and nSites
    s.step(1)
    Q, H, parts = s.observables()
    want = [H.real, H.imag] + [p.real for p in parts] + [s.nsites()[1]]
    assert close(rows[0][1:], want), (rows[0], want)
    s.step(3)
    z = 0.02 * 0.5
    s.set_activity(1, z)
    s.step(4 + 3)
    tries = np.loadtxt(tmp_path / "fs.dat", ndmin=2)
    assert 1 <= len(tries) <= 4
    for t in tries:
        z *= 0.9
        s.set_activity(1, z)
        s.step(3)
        n = s.nsites()
        assert close(t, [-1.0, n[0], z, n[1]]), (t, z, n)
    # sweep_data.dat: Re H, nSites, activity before every change
    sweep = np.loadtxt(tmp_path / "sweep_data.dat", ndmin=2)
    assert len(sweep) == 1 + len(tries)
    assert close(sweep[:, 2], [0.02] + [0.01 * 0.9 ** k for k in range(len(tries))])
    s.close()


def read_lammpstrj(path):
    """Frames of a LAMMPS dump (PS_Box::writeLammpsTraj): [(step, ids, types, mols, x)]."""
    frames = []
    lines = open(path).read().split("\n")
    k = 0
    while k < len(lines) and lines[k].startswith("ITEM: TIMESTEP"):
        step = int(lines[k + 1])
        n = int(lines[k + 3])
        assert lines[k + 4] == "ITEM: BOX BOUNDS pp pp pp" and lines[k + 8] == "ITEM: ATOMS id type mol x y z"
        rows = np.array([l.split() for l in lines[k + 9:k + 9 + n]], np.float64)
        frames.append((step, rows[:, 0].astype(int), rows[:, 1].astype(int), rows[:, 2].astype(int), rows[:, 3:6]))
        k += 9 + n
    return frames


def test_ps_trajectory_frames_match_engine(gpu, tmp_path):
    # trajFreq (PS_Box.cu:190-194, 756-801): frames after steps 0, 5, 10
    cfg = small_system(False, seed=8)
    say("case: front end lammpstrj")
    write_data_file(tmp_path / "init.data", cfg)
    text = ps_input(cfg, "init.data", steps=12, log_freq=100).replace("endBox", "trajFreq 5\nendBox")
    (tmp_path / "input").write_text(text)
    r = run_mft(["-in", "input"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    frames = read_lammpstrj(tmp_path / "traj.lammpstrj")
    assert [f[0] for f in frames] == [0, 5, 10]
    n = cfg["topo"]["type"].size
    L = np.asarray(cfg["L"], np.float64)
    eng = systems.tild_engine(cfg, "GJF")
    eng.set_run(12)
    done = 0
    for step, ids, types, mols, x in frames:
        assert np.array_equal(ids, np.arange(1, n + 1))
        assert np.array_equal(types, np.asarray(cfg["topo"]["type"]) + 1)
        eng.step(step + 1 - done)
        done = step + 1
        xe, _, _ = eng.get_state()
        d = (x - xe.reshape(-1, 3)) / L
        d -= np.round(d)
        assert np.max(np.abs(d)) < 1e-5, step
        assert np.all((x >= 0) & (x <= L + 1e-6))
    eng.close()


def read_kfield(path):
    return np.loadtxt(path)


def test_ps_compute_avg_sk_matches_engine(gpu, tmp_path):
    # compute avg_sk (ps_compute.cu, ps_compute_sk.cu): |rho_B(k)|^2 after steps
    # 2, 4, 6, 8, 10 (step > wait 1, freq 2), averaged and written at the log
    # steps 5 and 10 in PS_Box::writeKFieldFloat's format
    cfg = small_system(False, seed=9)
    say("case: front end compute avg_sk")
    write_data_file(tmp_path / "init.data", cfg)
    text = ps_input(cfg, "init.data", steps=12, log_freq=5).replace(
        "endBox", "compute avg_sk %s freq 2 wait 1 file_name sk.dat\nendBox" % cfg["species"][1])
    (tmp_path / "input").write_text(text)
    r = run_mft(["-in", "input"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    got = read_kfield(tmp_path / "sk.dat")
    eng = systems.tild_engine(cfg, "GJF")
    eng.set_run(12)
    acc, num = 0.0, 0
    for s in range(11):
        eng.step(1)
        if s > 1 and s % 2 == 0:
            acc = acc + eng.structure_factor(2).astype(np.float64)
            num += 1
    want = acc / num
    M = int(np.prod(cfg["Nx"]))
    assert got.shape == (M, 6)
    assert close(got[:, 5], want, rel=1e-5, atol=1e-7)
    assert close(got[:, 3], np.abs(want), rel=1e-5, atol=1e-7)
    # k columns: Box::get_kD with the Nyquist index negative
    nx = cfg["Nx"][0]
    kx = np.where(2 * np.arange(nx) < nx, np.arange(nx), np.arange(nx) - nx) * 2 * np.pi / cfg["L"][0]
    assert close(got[:nx, 0], kx, rel=1e-5, atol=1e-5)
    eng.close()


GSD_REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                       "libgsd_ref.so")


class _GsdEntry(C.Structure):
    _fields_ = [("frame", C.c_uint64), ("N", C.c_uint64), ("location", C.c_int64), ("M", C.c_uint32),
                ("id", C.c_uint16), ("type", C.c_uint8), ("flags", C.c_uint8)]


def read_gsd(path):
    """Every chunk of every frame, read with the reference's own GSD library
    (gsd.c compiled from its sources into oracle/_ref)."""
    lib = C.CDLL(GSD_REF)
    lib.gsd_find_chunk.restype = C.POINTER(_GsdEntry)
    lib.gsd_get_nframes.restype = C.c_uint64
    h = C.create_string_buffer(1 << 16)      # room for struct gsd_handle
    assert lib.gsd_open(h, str(path).encode(), 2) == 0   # GSD_OPEN_READONLY
    dtypes = {1: np.uint8, 3: np.uint32, 4: np.uint64, 5: np.int8, 9: np.float32}
    names = ["configuration/step", "configuration/dimensions", "configuration/box", "particles/N",
             "particles/mass", "particles/typeid", "log/particles/moleculeid", "particles/types", "bonds/N",
             "bonds/typeid", "bonds/group", "angles/N", "angles/typeid", "angles/group", "particles/position",
             "particles/charge"]
    frames = []
    for f in range(lib.gsd_get_nframes(h)):
        fr = {}
        for name in names:
            e = lib.gsd_find_chunk(h, C.c_uint64(f), name.encode())
            if not e:
                continue
            ent = e.contents
            out = np.zeros(max(1, ent.N * ent.M), dtypes[ent.type])
            # gsd_read_chunk refuses an empty chunk (GSD_ERROR_FILE_CORRUPT)
            if ent.N * ent.M > 0:
                assert lib.gsd_read_chunk(h, out.ctypes.data_as(C.c_void_p), e) == 0
            fr[name] = out[: ent.N * ent.M].reshape(ent.N, ent.M) if ent.M > 1 else out[: ent.N]
        frames.append(fr)
    lib.gsd_close(h)
    return frames


def test_ps_gsd_frames_read_by_the_reference_library(gpu, tmp_path):
    # gsdFreq (PS_Box_gsdUtils.cu:14-160): frames after steps 0, 5, 10 that the
    # reference's gsd.c opens and reads back
    if not os.path.exists(GSD_REF):
        pytest.skip("oracle/_ref/libgsd_ref.so not built (reference sources absent where the oracle was built)")
    cfg = small_system(True, seed=10)
    say("case: front end gsd")
    write_data_file(tmp_path / "init.data", cfg)
    text = ps_input(cfg, "init.data", steps=12, log_freq=100).replace("endBox", "gsdFreq 5\nendBox")
    (tmp_path / "input").write_text(text)
    r = run_mft(["-in", "input"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    frames = read_gsd(tmp_path / "traj.gsd")
    assert len(frames) == 3
    topo = cfg["topo"]
    n = topo["type"].size
    L = np.asarray(cfg["L"], np.float32)
    f0 = frames[0]
    assert f0["configuration/dimensions"][0] == 3 and f0["particles/N"][0] == n
    assert np.array_equal(f0["configuration/box"], np.r_[L, 0, 0, 0].astype(np.float32))
    assert np.array_equal(f0["particles/typeid"], topo["type"] + 1)
    nb, mb = topo["nBonds"], topo["bondedTo"].reshape(n, -1)
    pairs = sorted((i, int(mb[i, j])) for i in range(n) for j in range(nb[i]) if mb[i, j] > i)
    assert f0["bonds/N"][0] == len(pairs)
    assert [tuple(p) for p in f0["bonds/group"].tolist()] == pairs
    assert f0["angles/N"][0] == 0
    assert "bonds/N" not in frames[1]
    eng = systems.tild_engine(cfg, "GJF")
    eng.set_run(12)
    done = 0
    for k, fr in enumerate(frames):
        assert fr["configuration/step"][0] == 5 * k
        eng.step(5 * k + 1 - done)
        done = 5 * k + 1
        xe, _, _ = eng.get_state()
        assert np.array_equal(fr["particles/position"], xe.reshape(-1, 3).astype(np.float32) - 0.5 * L)
        assert np.array_equal(fr["particles/charge"], topo["charge"].astype(np.float32))
    eng.close()


def test_ps_front_end_angles_match_engine(gpu, tmp_path):
    # angle / MAXANGLES keywords and the data file's Angles section: the log
    # carries Ubond, Uangle, then the potentials (PS_Box.cu:339-372), and the
    # final positions agree with the engine given the same angle lists
    from test_oracle_tild import angle_system
    cfg = angle_system(seed=8)
    say("case: front end ps with angles")
    write_data_file(tmp_path / "init.data", cfg)
    text = ps_input(cfg, "init.data", steps=12, log_freq=5)
    if os.path.exists(GSD_REF):
        text = text.replace("endBox", "gsdFreq 100\nendBox")
    (tmp_path / "input").write_text(text)
    r = run_mft(["-in", "input"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    n = cfg["topo"]["type"].size
    eng = systems.tild_engine(cfg, "GJF")
    eng.set_run(12)
    done = 0
    for row in read_log(tmp_path / "ps_data.dat"):
        k = {-1: 0, 0: 1, 5: 6, 10: 11, 12: 12}[int(row[0])]
        pe, be = eng.step_energies(k - done) if k > done else eng.energies()
        done = k
        assert close(row[1:], [be, eng.angle_energy()] + list(pe)), (row, be, pe)
    assert done == 12
    x_fe = read_data_positions(tmp_path / "final.input.data", n)
    x_api, _, _ = eng.get_state()
    d = (x_fe.astype(np.float64) - x_api).reshape(-1, 3)
    L = np.asarray(cfg["L"])
    d -= L * np.round(d / L)
    assert np.max(np.abs(d)) < 1e-5 * max(cfg["L"])
    eng.close()
    if os.path.exists(GSD_REF):
        f0 = read_gsd(tmp_path / "traj.gsd")[0]
        ag = cfg["angles"]["angleGroup"]
        verts = [tuple(int(v) for v in ag[i, a]) for i in range(n) for a in range(cfg["angles"]["nAngles"][i])
                 if ag[i, a, 1] == i]
        assert f0["angles/N"][0] == len(verts) == 60 * 8
        assert [tuple(g) for g in f0["angles/group"].tolist()] == verts


def test_ps_front_end_vv_dpd_langevin_match_engine(gpu, tmp_path):
    # integrator VV with a dpd potential on A (through its neighbor_list) and
    # a Langevin potential on B: the front end's log rows and final positions
    # agree with the engine driven directly
    cfg = small_system(False, seed=11)
    cfg["potentials"] = list(cfg["potentials"]) + [
        dict(kind="dpd", maskI=1, maskJ=0, A=4.5, sigma=1.0, Rp=1.0, wall_pos=cfg["delt"]),
        dict(kind="langevin", maskI=2, maskJ=0, A=0.7, sigma=cfg["delt"])]
    say("case: front end VV + dpd + Langevin")
    write_data_file(tmp_path / "init.data", cfg)
    (tmp_path / "input").write_text(ps_input(cfg, "init.data", steps=12, log_freq=5, integrator="VV"))
    r = run_mft(["-in", "input"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    n = cfg["topo"]["type"].size
    eng = systems.tild_engine(cfg, "VV")
    eng.set_run(12)
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


def read_bin_frames(path, M):
    """Binary density file of Box::initBinaryDataFile / writeBinaryData
    (Box.cu:553-582): int Dim, int Nx[Dim], float L[Dim], then M floats a frame."""
    raw = open(path, "rb").read()
    dim = np.frombuffer(raw[:4], np.int32)[0]
    nx = np.frombuffer(raw[4:4 + 4 * dim], np.int32)
    L = np.frombuffer(raw[4 + 4 * dim:4 + 8 * dim], np.float32)
    body = np.frombuffer(raw[4 + 8 * dim:], np.float32)
    assert body.size % M == 0
    return dim, nx, L, body.reshape(-1, M)


@pytest.mark.parametrize("charged", [False, True])
def test_ps_binary_density_frames_match_engine(gpu, tmp_path, charged):
    # fieldFreq 5 over 12 steps: frames after steps 0, 5, 10 and the final one
    # (PS_Box.cu:198-203, 228), each group's density from that step's density pass
    cfg = small_system(charged, seed=12)
    say("case: front end binary density frames, charged =", charged)
    write_data_file(tmp_path / "init.data", cfg)
    text = ps_input(cfg, "init.data", steps=12, log_freq=100).replace("endBox", "fieldFreq 5\nendBox")
    (tmp_path / "input").write_text(text)
    r = run_mft(["-in", "input"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    M = int(np.prod(cfg["Nx"]))
    names = ["all"] + list(cfg["species"]) + (["charges"] if charged else [])
    frames = {}
    for nm in names:
        dim, nx, L, fr = read_bin_frames(tmp_path / f"density-{nm}.bin", M)
        assert dim == 3 and list(nx) == list(cfg["Nx"]) and np.allclose(L, cfg["L"])
        assert fr.shape[0] == 4, (nm, fr.shape)
        frames[nm] = fr
    eng = systems.tild_engine(cfg, "GJF")
    eng.set_run(12)
    q = cfg["topo"]["charge"]
    t = cfg["topo"]["type"]
    charged_species = [s for s in range(len(cfg["species"])) if np.any(q[t == s] != 0)]
    done = 0
    for k, step in enumerate([1, 6, 11, 12]):
        eng.step(step - done)
        done = step
        rho, _ = eng.densities()
        want = {"all": rho.sum(axis=0)}
        for s, nm in enumerate(cfg["species"]):
            want[nm] = rho[s]
        if charged:
            want["charges"] = rho[charged_species].sum(axis=0)
        for nm in names:
            assert np.allclose(frames[nm][k], want[nm], rtol=1e-6, atol=1e-6 * np.abs(want[nm]).max()), (nm, k)
    eng.close()


FTS_FIELDS_INPUT = """box fts cl
Dim 3
grid 16 12 20
boxLengths 3.0 2.5 4.0
Nr 12
C 4
randSeed 9
species A
species B
molecule linear phi 1.0 2 5 A 7 B
potential Incompress 0.01 initialize sin 2 0.3 0.1 2
potential Flory A B 12 0.01 0.01 initialize L 0.2 dir 1 n_periods 2
chemFieldFreq 2
densityFieldFreq 2
clSampleFreq 1
logFreq 100
endBox
run vt 5
"""


def read_grid_text(path):
    a = np.loadtxt(path)
    return a[:, :3], a[:, 3] + 1j * a[:, 4]


def test_fts_field_files_and_initialize_options(gpu, tmp_path):
    # initialize sin / ordered phase (fts_potential.cu:88-138), then the field
    # files of FTS_Box::doTimeStep (FTS_Box.cu:84-113) after iterations 0, 2, 4
    say("case: front end fts field files")
    (tmp_path / "input").write_text(FTS_FIELDS_INPUT)
    r = run_mft(["-in", "input"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    Nx, L = (16, 12, 20), (3.0, 2.5, 4.0)
    M = int(np.prod(Nx))
    n = np.stack([np.arange(M) % Nx[0], (np.arange(M) // Nx[0]) % Nx[1], np.arange(M) // (Nx[0] * Nx[1])], axis=1)
    dx = np.float32(np.asarray(L, np.float32) / np.asarray(Nx, np.float32)).astype(np.float64)
    # initial fields: sin along z with real and imaginary amplitudes; a
    # lamellar w- along y (the Flory initialize also sets w+ = -1.5i)
    PI = float(np.float32(3.141592654))
    wpl0 = (0.3 + 0.1j) * np.sin(2 * PI * n[:, 2] * dx[2] * 2 / float(np.float32(L[2])))
    arg = np.float32(2.0 * np.float32(PI) * np.float32(2) / np.float32(L[1]))
    wmi0 = 0.2 * np.sin((arg * (n[:, 1].astype(np.float32) * np.float32(dx[1]))).astype(np.float32).astype(np.float64))
    spec = dict(Nx=Nx, L=L, Nr=12.0, C=4.0, nspecies=2, cl=True, seed=9,
                potentials=[dict(kind="Incompress", dt_pl=0.01),
                            dict(kind="Flory", I=0, J=1, chiN=12.0, dt_pl=0.01, dt_mi=0.01)],
                molecules=[dict(blocks=[(5, 0), (7, 1)], phi=1.0)])
    fields = {(0, "pl"): wpl0.astype(np.complex64), (1, "pl"): np.full(M, -1.5j, np.complex64),
              (1, "mi"): wmi0.astype(np.complex64)}
    s = systems.fts_system(spec, fields)
    avg = np.zeros((2, M), np.complex128)
    for it in range(5):
        s.step(1)
        avg += s.densities()
        if it % 2:
            continue
    # after the run: the files hold iteration 4's state
    coords, w = read_grid_text(tmp_path / "wpl_Incompress0.dat")
    assert np.allclose(coords, n * dx, atol=1e-6)
    tol = dict(rtol=2e-4, atol=2e-4)   # %1.4e: five significant digits
    assert np.allclose(w, s.get_fields(0), **tol)
    wp, wm = s.get_fields(1, minus=True)
    assert np.allclose(read_grid_text(tmp_path / "wpl_Flory1.dat")[1], wp, **tol)
    assert np.allclose(read_grid_text(tmp_path / "wmi_Flory1.dat")[1], wm, **tol)
    rho = s.densities()
    W = s.species_fields()
    for sp in range(2):
        got = read_grid_text(tmp_path / f"rhoSpecies{sp}.dat")[1]
        assert np.allclose(got, rho[sp], rtol=2e-4, atol=2e-4 * np.abs(rho[sp]).max())
        got = read_grid_text(tmp_path / f"rhoSpeciesAvg{sp}.dat")[1]
        assert np.allclose(got, avg[sp] / 5, rtol=2e-4, atol=2e-4 * np.abs(avg[sp]).max())
        got = read_grid_text(tmp_path / f"fieldSpecies{sp}.dat")[1]
        assert np.allclose(got, W[sp], rtol=2e-4, atol=2e-4 * np.abs(W[sp]).max())
    s.close()
    # readDatFile reads the written w+ back as the initial field of a new run
    text = FTS_FIELDS_INPUT.replace("initialize sin 2 0.3 0.1 2", "initialize readDatFile wpl_Incompress0.dat")
    text = text.replace("run vt 5", "run vt 1").replace("chemFieldFreq 2", "chemFieldFreq 0")
    (tmp_path / "input2").write_text(text)
    w_read = read_grid_text(tmp_path / "wpl_Incompress0.dat")[1]
    r = run_mft(["-in", "input2"], cwd=tmp_path)
    assert r.returncode == 0, r.stdout[-3000:]
    fields2 = dict(fields)
    fields2[(0, "pl")] = w_read.astype(np.complex64)
    s = systems.fts_system(spec, fields2)
    s.step(1)
    rows = read_log(tmp_path / "fts_data.dat")
    Q, H, parts = s.observables()
    assert close(rows[-1][1:3], [H.real, H.imag]), (rows[-1], H)
    s.close()
