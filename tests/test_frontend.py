"""The drop-in front end parses the reference's input syntax (the hot-path
configurations in its v2 grammar) into the engine description; checked with
--dry-run, which needs no GPU."""
import json
import os

import pytest

from frontend_util import INPUTS, MFT_BIN, run_mft

pytestmark = pytest.mark.skipif(not os.path.exists(MFT_BIN), reason="front end not built")


def dry(tmp_path, text=None, name=None):
    if text is not None:
        p = tmp_path / "input"
        p.write_text(text)
        name = str(p)
    r = run_mft(["-in", name, "--dry-run"], cwd=tmp_path, timeout=120)
    return r


def desc(tmp_path, name):
    r = dry(tmp_path, name=os.path.join(INPUTS, name))
    assert r.returncode == 0, r.stdout[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_c1_diblock(tmp_path):
    d = desc(tmp_path, "c1_diblock.in")
    assert d["grid"] == [32, 32, 32] and d["L"] == [20, 20, 20] and d["pmeorder"] == 3
    assert d["nparticles"] == 40000 and d["nspecies"] == 2 and d["nbonds"] == 2000 * 19
    assert d["integrator"] == 0 and d["delt"] == pytest.approx(0.005) and d["seed"] == 1234
    pots = d["potentials"]
    assert [(p["maskI"], p["maskJ"]) for p in pots] == [(1, 1), (2, 2), (1, 2)]
    assert [p["A"] for p in pots] == pytest.approx([5.0, 5.0, 10.2])
    assert all(p["sigma"] == pytest.approx(0.70711) and p["kind"] == 0 for p in pots)
    assert d["bonds"] == [[0, 1.5, 0.0]]


def test_c2_coacervate(tmp_path):
    d = desc(tmp_path, "c2_coacervate.in")
    assert d["nparticles"] == 1800000 and d["nspecies"] == 3 and d["has_charges"] == 1
    assert d["nbonds"] == 2 * 40000 * 19
    assert d["potentials"][-1]["kind"] == 3 and d["potentials"][-1]["A"] == pytest.approx(2.0)
    assert d["potentials"][-1]["sigma"] == pytest.approx(0.5)


def test_c3_fts(tmp_path):
    d = desc(tmp_path, "c3_fts_diblock.in")
    assert d["style"] == "cl" and d["grid"] == [64, 64, 64] and d["Nr"] == 100 and d["C"] == 50
    h, f = d["potentials"]
    assert h["kind"] == 0 and h["kappaN"] == 100 and h["delt_pl"] == pytest.approx(0.01)
    assert f["kind"] == 1 and (f["I"], f["J"]) == (0, 1) and f["chiN"] == 20
    assert d["molecules"] == [{"phi": 1.0, "activity": 0.0, "blocks": [[50, 0], [50, 1]]}]


def test_update_scheme_keyword(tmp_path):
    # fts_potential_flory.cu:90-109: updateScheme EMPC selects the predictor-corrector
    text = open(os.path.join(INPUTS, "c3_fts_diblock.in")).read().replace(
        "potential Flory A B 20 0.01 0.01", "potential Flory A B 20 0.01 0.01 updateScheme EMPC")
    r = dry(tmp_path, text)
    assert r.returncode == 0, r.stdout[-1000:]
    h, f = json.loads(r.stdout.strip().splitlines()[-1])["potentials"]
    assert h["scheme"] == 0 and f["scheme"] == 1
    r = dry(tmp_path, text.replace("updateScheme EMPC", "updateScheme 1S"))
    assert json.loads(r.stdout.strip().splitlines()[-1])["potentials"][1]["scheme"] == 2


def test_flory_initialize_sets_wplus(tmp_path):
    # fts_potential_flory.cu:83-88: initialize also sets w+ = -1.5i
    text = open(os.path.join(INPUTS, "c3_fts_diblock.in")).read().replace(
        "potential Flory A B 20 0.01 0.01", "potential Flory A B 20 0.01 0.01 initialize value 0.2 0.0")
    r = dry(tmp_path, text)
    assert r.returncode == 0, r.stdout[-1000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["potentials"][1]["wpl0"] == [0.0, -1.5]


def test_ramp_and_groups(tmp_path):
    text = open(os.path.join(INPUTS, "c1_diblock.in")).read().replace(
        "potential gaussian A B 10.2 0.70711", "group AB types A B\npotential gaussian AB all 10.2 0.70711 ramp 20.4")
    r = dry(tmp_path, text)
    assert r.returncode == 0, r.stdout[-1000:]
    p = json.loads(r.stdout.strip().splitlines()[-1])["potentials"][2]
    assert (p["maskI"], p["maskJ"], p["ramp"]) == (3, 3, 1) and p["A_final"] == pytest.approx(20.4)


@pytest.mark.parametrize("bad, msg", [("Dim 3\nbogus 1\n", "Invalid keyword bogus"),
                                       ("Dim 3\npotential wobble A A 1 1\n", "is not a supported potential"),
                                       ("Dim 3\nintegrator all XYZ\n", "XYZ is not a valid Integrator")])
def test_errors_die_like_the_reference(tmp_path, bad, msg):
    r = dry(tmp_path, "box ps\n" + bad + "endBox\n")
    assert r.returncode == 1 and msg in r.stdout


EDWARDS_GC = """box fts cl
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
"""


def test_edwards_and_activity_keywords(tmp_path):
    # fts_potential.cu:30-31 (Edwards B delt), fts_molecule.cu:30-31 (activity)
    r = dry(tmp_path, EDWARDS_GC)
    assert r.returncode == 0, r.stdout[-1000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    e, f = d["potentials"]
    assert e["kind"] == 3 and e["kappaN"] == pytest.approx(3.0) and e["delt_pl"] == pytest.approx(0.01)
    assert e["wpl0"] == [0.0, 0.0]
    m1, m2 = d["molecules"]
    assert m1["activity"] == 0.0 and m1["phi"] == pytest.approx(0.7)
    assert m2["activity"] == pytest.approx(0.02) and m2["blocks"] == [[6, 1]]


def test_find_spinodal_arguments(tmp_path):
    # FTS_BoxFindSpinodal.cu:27-67: defaults and keywords, molecule 1-indexed
    r = dry(tmp_path, EDWARDS_GC + "run findspinodal molecule 2 nSteps 30 initial_f 0.9 fScale 0.25 datFile a.dat\n")
    assert r.returncode == 0, r.stdout[-1000:]
    fs = json.loads(r.stdout.strip().splitlines()[-1])["findSpinodal"]
    assert fs["molecule"] == 1 and fs["nSteps"] == 30 and fs["f"] == pytest.approx(0.9)
    assert fs["fScale"] == pytest.approx(0.25) and fs["datFile"] == "a.dat" and fs["phiTol"] == pytest.approx(1e-4)
    r = dry(tmp_path, EDWARDS_GC + "run findspinodal\n")
    assert r.returncode == 1 and "grand canonical" in r.stdout


@pytest.mark.parametrize("cmd, msg", [("modify molecule 3 activity scale 0.5", "Invalid molecule index - out of range!"),
                                      ("modify molecule 1 phi scale 0.5", "changes in FTS:molec:phi not yet supported"),
                                      ("modify potential 1 chiN 3", "potential modifications not yet supported")])
def test_modify_errors_die_like_the_reference(tmp_path, cmd, msg):
    r = dry(tmp_path, EDWARDS_GC + cmd + "\n")
    assert r.returncode == 1 and msg in r.stdout


def test_modify_zero_mean_keyword(tmp_path):
    # fts_potential_helfand.cu:55-59 / fts_potential_flory.cu:94-98
    text = EDWARDS_GC.replace("potential Edwards 3.0 0.01", "potential Edwards 3.0 0.01 modify zeroMean")
    r = dry(tmp_path, text)
    assert r.returncode == 0, r.stdout[-1000:]
    e, f = json.loads(r.stdout.strip().splitlines()[-1])["potentials"]
    assert e["zero_mean"] == 1 and f["zero_mean"] == 0
    r = dry(tmp_path, text.replace("modify zeroMean", "modify sideways"))
    assert r.returncode == 1 and "Invalid modify option on Edwards potential" in r.stdout


def test_compute_avg_sk_keyword(tmp_path):
    # PS_Compute (ps_compute.cu:8-43): group, then wait / freq / file_name
    text = open(os.path.join(INPUTS, "c1_diblock.in")).read().replace(
        "logFreq", "compute avg_sk all freq 5 wait 10\ncompute avg_sk B file_name skB.dat\nlogFreq")
    r = dry(tmp_path, text)
    assert r.returncode == 0, r.stdout[-1000:]
    c = json.loads(r.stdout.strip().splitlines()[-1])["computes"]
    assert c[0] == {"style": "avg_sk", "mask": 3, "wait": 10, "freq": 5, "file": "compute-0.dat"}
    assert c[1] == {"style": "avg_sk", "mask": 2, "wait": 0, "freq": 100, "file": "skB.dat"}
    r = dry(tmp_path, text.replace("wait 10", "wiat 10"))
    assert r.returncode == 1 and "PS_Compute: wiat invalid option for compute!" in r.stdout


def test_bias_potential_keyword(tmp_path):
    # ps_potentialBias.cu:17-66 and Box::known_phase (Box.cu:283-303)
    base = open(os.path.join(INPUTS, "c1_diblock.in")).read()
    text = base.replace("logFreq", "potential bias A 2.5 CYL dir 2 n_periods 2\nlogFreq")
    r = dry(tmp_path, text)
    assert r.returncode == 0, r.stdout[-1000:]
    p = json.loads(r.stdout.strip().splitlines()[-1])["potentials"][-1]
    assert (p["kind"], p["maskI"], p["phase"], p["dir"], p["n_periods"]) == (4, 1, 1, 2, 2)
    assert p["A"] == pytest.approx(2.5)
    for bad, msg in (("potential bias A 2.5 CUBE", "PHASE NOT FOUND in Box.cu"),
                     ("potential bias A 2.5 L wobble 3", "wobble is not a valid initialize option"),
                     ("potential bias A 2.5 BCC", "uninitialised")):
        r = dry(tmp_path, base.replace("logFreq", bad + "\nlogFreq"))
        assert r.returncode == 1 and msg in r.stdout, (bad, r.stdout[-300:])


def test_wall_potential_keyword(tmp_path):
    # ps_potentialWall.cu:10-48: group axis wallPos dirSign k, with its die texts
    base = open(os.path.join(INPUTS, "c1_diblock.in")).read()
    r = dry(tmp_path, base.replace("logFreq", "potential wall B y 7.5 -1 4.0\nlogFreq"))
    assert r.returncode == 0, r.stdout[-1000:]
    p = json.loads(r.stdout.strip().splitlines()[-1])["potentials"][-1]
    assert (p["kind"], p["maskI"], p["dir"], p["wall_sign"]) == (5, 2, 1, -1)
    assert p["wall_pos"] == pytest.approx(7.5) and p["A"] == pytest.approx(4.0)
    for bad, msg in (("potential wall B w 7.5 1 4.0", "axis must be x, y, or z (got 'w')"),
                     ("potential wall B y 7.5 2 4.0", "dirSign must be +1 or -1 (got 2)"),
                     ("potential wall B y 7.5 1 4.0 ramp 5", "ramping applied on an unsupported potential!")):
        r = dry(tmp_path, base.replace("logFreq", bad + "\nlogFreq"))
        assert r.returncode == 1 and msg in r.stdout, (bad, r.stdout[-300:])


def test_angle_keywords(tmp_path):
    # angle type style k [theta] (PS_Box_initialization.cu:83-106), MAXANGLES,
    # and angleType on linear molecules (PS_Box_moleculeMakers.cu:242-288):
    # every interior bead of the chains is the vertex of one angle
    base = open(os.path.join(INPUTS, "c1_diblock.in")).read()
    text = base.replace("bond 1 harmonic 1.5 0.0", "bond 1 harmonic 1.5 0.0\nangle 1 wlc 2.0\nangle 2 harmonic 3.0 120")
    text = text.replace("molecule linear nmolecs 2000 2 10 A 10 B",
                        "molecule linear nmolecs 2000 2 10 A 10 B angleType 1 2")
    r = dry(tmp_path, text)
    assert r.returncode == 0, r.stdout[-1000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["nangles"] == 2000 * 18 and d["maxangles"] == 3
    assert d["angles"][0] == [0, 2.0, 0.0]
    assert d["angles"][1][0] == 1 and d["angles"][1][2] == pytest.approx(120 * 3.141592654 / 180, rel=1e-6)
    for bad, msg in (("angle 1 cosine 2.0", "Angle style not supported!"),):
        r = dry(tmp_path, base.replace("logFreq", bad + "\nlogFreq"))
        assert r.returncode == 1 and msg in r.stdout, (bad, r.stdout[-300:])
    late = base.replace("logFreq", "MAXANGLES 4\nlogFreq")
    r = dry(tmp_path, late)
    assert r.returncode == 1 and "MAXANGLES must be set before defining any molecules" in r.stdout
    undefined = base.replace("molecule linear nmolecs 2000 2 10 A 10 B",
                             "molecule linear nmolecs 2000 2 10 A 10 B angleType 1 1")
    r = dry(tmp_path, undefined)
    assert r.returncode == 1 and "angle type 1 is not defined" in r.stdout


def test_vv_integrator_keyword(tmp_path):
    # ps_integrator.cu:52-68: VV is one of the three integrators
    base = open(os.path.join(INPUTS, "c1_diblock.in")).read()
    r = dry(tmp_path, base.replace("integrator all GJF delt 0.005", "integrator all VV delt 0.002"))
    assert r.returncode == 0, r.stdout[-1000:]
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["integrator"] == 3 and d["delt"] == pytest.approx(0.002)
    r = dry(tmp_path, base.replace("integrator all GJF delt 0.005", "integrator all Langevin"))
    assert r.returncode == 1 and "Langevin is not a valid Integrator" in r.stdout


def test_langevin_potential_keyword(tmp_path):
    # ps_potentialLangevin.cu:16-57: group [drag d] [delt dt], delt defaulting
    # to the integrator's, with its die texts
    base = open(os.path.join(INPUTS, "c1_diblock.in")).read()
    text = base.replace("integrator all GJF delt 0.005", "integrator all VV delt 0.004")
    r = dry(tmp_path, text.replace("logFreq", "potential Langevin A drag 0.5\nlogFreq"))
    assert r.returncode == 0, r.stdout[-1000:]
    p = json.loads(r.stdout.strip().splitlines()[-1])["potentials"][-1]
    assert (p["kind"], p["maskI"]) == (6, 1)
    assert p["A"] == pytest.approx(0.5) and p["sigma"] == pytest.approx(0.004)
    r = dry(tmp_path, text.replace("logFreq", "potential langevin B delt 0.01\nlogFreq"))
    p = json.loads(r.stdout.strip().splitlines()[-1])["potentials"][-1]
    assert p["A"] == pytest.approx(1.0) and p["sigma"] == pytest.approx(0.01) and p["maskI"] == 2
    r = dry(tmp_path, text.replace("logFreq", "potential Langevin A friction 2\nlogFreq"))
    assert r.returncode == 1 and "friction is not a valid initialize option in fts_potential" in r.stdout
    # before any integrator: no delt to default to
    early = base.replace("potential gaussian A A", "potential Langevin A\npotential gaussian A A")
    r = dry(tmp_path, early)
    assert r.returncode == 1 and "did not find appropriate delt for the Langevin noise" in r.stdout


def test_dpd_potential_keyword(tmp_path):
    # ps_potentialDPD.cu:6-43 and ps_neighborList.cu:132-141: dpd needs a
    # neighbor_list on its group; delt defaults to the integrator's
    base = open(os.path.join(INPUTS, "c1_diblock.in")).read()
    text = base.replace("logFreq", "neighbor_list A 1.2\npotential dpd A 4.5 1.0\nlogFreq")
    r = dry(tmp_path, text)
    assert r.returncode == 0, r.stdout[-1000:]
    p = json.loads(r.stdout.strip().splitlines()[-1])["potentials"][-1]
    assert (p["kind"], p["maskI"]) == (7, 1)
    assert (p["A"], p["sigma"], p["Rp"]) == pytest.approx((4.5, 1.0, 1.2)) and p["wall_pos"] == pytest.approx(0.005)
    r = dry(tmp_path, base.replace("logFreq", "potential dpd A 4.5 1.0\nlogFreq"))
    assert r.returncode == 1 and "dpd: no neighbor_list found for group A" in r.stdout
    r = dry(tmp_path, base.replace("logFreq", "neighbor_list A 1.2\npotential dpd A 4.5 1.0 rc 2\nlogFreq"))
    assert r.returncode == 1 and "Invalid keyword in dpd potential: rc" in r.stdout


@pytest.mark.parametrize("bad", ["partner", "type"])
def test_read_data_rejects_bad_bond_lines(tmp_path, bad):
    # a bond whose partner id or type is outside the data file's ranges stops
    # the front end with a message before anything indexes with it
    from frontend_util import ps_input, write_data_file
    from test_oracle_tild import small_system
    cfg = small_system(False, seed=3)
    write_data_file(tmp_path / "init.data", cfg)
    lines = (tmp_path / "init.data").read_text().split("\n")
    k = lines.index("Bonds") + 2
    cnt, t, i, j = lines[k].split()
    n = cfg["topo"]["type"].size
    lines[k] = f"{cnt} {t} {i} {n + 5}" if bad == "partner" else f"{cnt} 7 {i} {j}"
    (tmp_path / "init.data").write_text("\n".join(lines))
    r = dry(tmp_path, ps_input(cfg, "init.data", steps=2, log_freq=1))
    assert r.returncode == 1, r.stdout[-1000:]
    assert ("bond partner out of range" if bad == "partner" else "bond type out of range") in r.stdout


MAKELINEAR_INPUT = """box ps
Dim 3
boxLengths 12 10 8
grid 16 16 16
pmeorder 3
randSeed 77
rho0 3.0
species A
species B
species C
bond 1 harmonic 1.5 0.0
bond 2 harmonic 2.0 0.5
molecule linear nmolecs 7 2 4 A 3 B bondType 1 2 charge 1.0 -1.0
molecule linear nmolecs 5 3 1 C 2 A 1 B bondType 2 1 1
molecule linear nmolecs 9 1 1 C
molecule linear phi 0.05 1 6 B
potential gaussian A B 1.0 0.5
integrator all GJF delt 0.005
logFreq 100
endBox
run nvt 1
"""


def test_make_linear_topology_of_the_front_end_is_the_oracles(tmp_path):
    # the C++ PS_Box::makeLinear (the drop-in's topology source) against the
    # oracle's make_linear, bit for bit, over four molecule commands: block
    # bond types, charges, a one-bead molecule, phi-derived counts
    # (PS_Box_moleculeMakers.cu:15-311) and molecule ids continuing from the
    # largest id already present (:140-145)
    import numpy as np
    import orc
    (tmp_path / "input").write_text(MAKELINEAR_INPUT)
    r = run_mft(["-in", "input", "--dry-run", "--dump-topology", "topo.txt"], cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:]
    got = np.loadtxt(tmp_path / "topo.txt")
    sp = {"A": 0, "B": 1, "C": 2}
    nphi = int(np.float32(3.0 * 12 * 10 * 8 * 0.05) / np.float32(6))   # int(rho0 V phi / Ntot)
    cmds = [(7, [(4, "A"), (3, "B")], [1, 2], [1.0, -1.0]), (5, [(1, "C"), (2, "A"), (1, "B")], [2, 1, 1], None),
            (9, [(1, "C")], None, None), (nphi, [(6, "B")], None, None)]
    start, mids, rows = 0, [], []
    for nm, blocks, bt, q in cmds:
        mid0 = max([0] + mids)
        t = orc.make_linear(nm, blocks, sp, block_bond_type=bt, charges=q, start=start, mid0=mid0)
        n = len(t["intSpecies"])
        for i in range(n):
            nb = int(t["nBonds"][i])
            bto = [int(t["bondedTo"][2 * i + b]) if b < nb else -1 for b in range(2)]
            bty = [int(t["bondType"][2 * i + b]) if b < nb else -1 for b in range(2)]
            rows.append([t["intSpecies"][i], t["charge"][i], t["mID"][i], nb] + bto + bty)
        mids += list(t["mID"])
        start += n
    want = np.array(rows, np.float64)
    assert got.shape == want.shape, (got.shape, want.shape)
    assert np.array_equal(got, want)
    # the second command's first molecule shares the id of the first command's last one
    assert got[7 * 7, 2] == got[7 * 7 - 1, 2] == 6
