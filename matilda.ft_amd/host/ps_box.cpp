// Particle (TILD) box of the drop-in front end: the reference's PS_Box input
// keywords, molecule makers and log output over the mft_tild engine.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <iostream>

#include "matilda.h"

// ----------------------------------------------------------- species ---
// PS_Species::PS_Species (ps_species.cu:13-38)
PS_Species::PS_Species(std::istringstream& iss, PS_Box* box)
{
    (void)box;
    if (!(iss >> speciesLabel)) die("Missing label for species");
    std::string word;
    while (iss >> word) {
        if (word == "mass" || word == "Mass") iss >> mass;
        else if (word == "mobility" || word == "Mobility" || word == "Diffusivity" || word == "diffusivity")
            iss >> mobility;
        else die("Invalid species option " + word);
    }
}

// ------------------------------------------------------------- groups ---
// PS_Group::PS_Group (ps_groups.cu:12-60): "group name types A B ..."
PS_Group::PS_Group(std::istringstream& iss, PS_Box* box)
{
    iss >> name;
    std::string s1;
    iss >> s1;
    if (s1 != "types") die("group style " + s1 + " is not supported");
    while (iss >> s1) speciesMask |= 1u << box->findSpeciesInteger(s1);
}

PS_Group::PS_Group(const std::string& n, uint32_t mask) : name(n), speciesMask(mask) {}

// --------------------------------------------------------- potentials ---
PS_Potential::PS_Potential(std::istringstream& iss, PS_Box* box) : input_command(iss.str()), mybox(box) {}

int PS_Potential::ramp_check_input(std::istringstream& iss, float A)
{
    if (iss.fail()) die("Error during input script; failed to properly read:\n" + iss.str());
    std::string convert;
    if (!(iss >> convert)) return 0;
    if (convert != "ramp") die("Invalid keyword: " + convert);
    ramp = true;
    Ao_initial = A;
    if (!(iss >> Ao_final)) die("no final prefactor specified");
    std::cout << "Ramping potential prefactor from " << Ao_initial << " to " << Ao_final << std::endl;
    return 1;
}

void PS_Potential::update_prefactor(int step, int maxSteps)
{
    // the engine applies the same schedule on the device; the host copy is
    // kept for the log
    if (ramp && maxSteps > 0) Ao = Ao_initial + (Ao_final - Ao_initial) / maxSteps * (step + 1);
}

static mft_tild_potential pair_desc(const PS_Potential& p, PS_Box* box, int kind)
{
    mft_tild_potential d{};
    d.kind = kind;
    d.maskI = box->groupMask(p.grpI);
    d.maskJ = box->groupMask(p.grpJ);
    d.A = p.ramp ? p.Ao_initial : p.Ao;
    d.sigma = std::sqrt(p.sig2);
    d.Rp = p.Rp;
    d.ramp = p.ramp ? 1 : 0;
    d.A_final = p.Ao_final;
    return d;
}

NBGauss::NBGauss(std::istringstream& iss, PS_Box* box) : PS_Potential(iss, box)
{
    float sigma;
    iss >> grpI >> grpJ >> Ao >> sigma;
    sig2 = sigma * sigma;
    ramp_check_input(iss, Ao);
}
mft_tild_potential NBGauss::descriptor() const { return pair_desc(*this, mybox, MFT_POT_GAUSSIAN); }

NBErf2::NBErf2(std::istringstream& iss, PS_Box* box) : PS_Potential(iss, box)
{
    float sigma;
    iss >> grpI >> grpJ >> Ao >> Rp >> sigma;
    sig2 = sigma * sigma;
    ramp_check_input(iss, Ao);
}
mft_tild_potential NBErf2::descriptor() const { return pair_desc(*this, mybox, MFT_POT_ERF2); }

NBErfG::NBErfG(std::istringstream& iss, PS_Box* box) : PS_Potential(iss, box)
{
    float sigma;
    iss >> grpI >> grpJ >> Ao >> Rp >> sigma;
    sig2 = sigma * sigma;
    ramp_check_input(iss, Ao);
}
mft_tild_potential NBErfG::descriptor() const { return pair_desc(*this, mybox, MFT_POT_ERFG); }

NBCharge::NBCharge(std::istringstream& iss, PS_Box* box) : PS_Potential(iss, box)
{
    grpI = grpJ = "charges";
    float sigma;
    iss >> LB >> sigma;
    sig2 = sigma * sigma;
    Ao = LB;
}
mft_tild_potential NBCharge::descriptor() const
{
    mft_tild_potential d{};
    d.kind = MFT_POT_CHARGES;
    d.A = LB;
    d.sigma = std::sqrt(sig2);
    return d;
}

// BiasField (ps_potentialBias.cu:17-66): group Ao phase [dir d] [n_periods n];
// Box::known_phase (Box.cu:283-303) names the phase, an unknown one stops at
// set-up ("PHASE NOT FOUND in Box.cu")
BiasField::BiasField(std::istringstream& iss, PS_Box* box) : PS_Potential(iss, box)
{
    std::string s1;
    iss >> grpI >> Ao >> s1;
    grpJ = grpI;
    if (s1 == "L" || s1 == "LAM") phase = "L";
    else if (s1 == "G" || s1 == "GYR") phase = "G";
    else if (s1 == "BCC" || s1 == "SPH" || s1 == "S") phase = "S";
    else if (s1 == "C" || s1 == "CYL" || s1 == "H" || s1 == "HEX") phase = "H";
    std::string w;
    while (iss >> w) {
        if (w == "dir") iss >> dir;
        else if (w == "n_periods") iss >> n_periods;
        else die(w + " is not a valid initialize option in fts_potential");
    }
    if (phase.empty()) die("PHASE NOT FOUND in Box.cu");
    // Box.cu:338-395 reads pbc_dr2's output argument before setting it
    if (phase == "S") die("the BCC bias field of the reference reads an uninitialised vector (Box.cu:359)");
}
mft_tild_potential BiasField::descriptor() const
{
    mft_tild_potential d{};
    d.kind = MFT_POT_BIAS;
    d.maskI = mybox->groupMask(grpI);
    d.A = Ao;
    d.phase = phase == "L" ? MFT_BIAS_LAM : (phase == "H" ? MFT_BIAS_HEX : MFT_BIAS_GYR);
    d.dir = dir;
    d.n_periods = n_periods;
    return d;
}

// Wall (ps_potentialWall.cu:10-48): group axis wallPos dirSign k; no ramp
Wall::Wall(std::istringstream& iss, PS_Box* box) : PS_Potential(iss, box)
{
    std::string axis;
    iss >> grpI >> axis;
    grpJ = grpI;
    if (axis == "x" || axis == "X") normalDim = 0;
    else if (axis == "y" || axis == "Y") normalDim = 1;
    else if (axis == "z" || axis == "Z") normalDim = 2;
    else die("ps_potentialWall.cu: axis must be x, y, or z (got '" + axis + "')");
    iss >> wallPos >> dirSign;
    if (dirSign != 1 && dirSign != -1)
        die("ps_potentialWall.cu: dirSign must be +1 or -1 (got " + std::to_string(dirSign) + ")");
    iss >> k;
    std::string s1;
    if (iss >> s1) {
        if (s1 == "ramp") die("ramping applied on an unsupported potential!");
        die("Invalid option " + s1 + " on wall potential");
    }
}
mft_tild_potential Wall::descriptor() const
{
    mft_tild_potential d{};
    d.kind = MFT_POT_WALL;
    d.maskI = mybox->groupMask(grpI);
    d.A = k;
    d.dir = normalDim;
    d.wall_pos = wallPos;
    d.wall_sign = dirSign;
    return d;
}

DPD::DPD(std::istringstream& iss, PS_Box* box) : PS_Potential(iss, box)
{
    iss >> grpI >> gamma >> kT;
    grpJ = grpI;
    if (!mybox->integrators.empty()) delt = mybox->integrators[0]->delt;
    std::string word;
    while (iss >> word) {
        if (word == "delt") iss >> delt;
        else die("Invalid keyword in dpd potential: " + word);
    }
    if (delt <= 0.0f) die("ps_potentialDPD: delt not set or invalid");
}
mft_tild_potential DPD::descriptor() const
{
    // DPD::initializePotential (ps_potentialDPD.cu:33-43): the group's neighbour list
    const PS_Box::NeighborList* nl = nullptr;
    for (const auto& n : mybox->neighborLists)
        if (n.group == grpI) nl = &n;
    if (!nl)
        die("dpd: no neighbor_list found for group " + grpI + " \xe2\x80\x94 add 'neighbor_list " + grpI +
            " <rcut>' before this potential");
    mft_tild_potential d{};
    d.kind = MFT_POT_DPD;
    d.maskI = mybox->groupMask(grpI);
    d.A = gamma;
    d.sigma = kT;
    d.Rp = nl->rcut;
    d.wall_pos = delt;
    return d;
}

Langevin::Langevin(std::istringstream& iss, PS_Box* box) : PS_Potential(iss, box)
{
    iss >> grpI;
    grpJ = grpI;
    // the noise's time step defaults to the first integrator's
    if (!mybox->integrators.empty()) delt = mybox->integrators[0]->delt;
    std::string s1;
    while (iss >> s1) {
        if (s1 == "drag") iss >> drag;
        else if (s1 == "delt") iss >> delt;
        else die(s1 + " is not a valid initialize option in fts_potential");
    }
    if (delt < 0.0f) die("ps-potentialLangevin.cu: did not find appropriate delt for the Langevin noise");
}
mft_tild_potential Langevin::descriptor() const
{
    mft_tild_potential d{};
    d.kind = MFT_POT_LANGEVIN;
    d.maskI = mybox->groupMask(grpI);
    d.A = drag;
    d.sigma = delt;
    return d;
}

PS_Potential* PSPotentialFactory(std::istringstream& iss, PS_Box* box)
{
    std::string s1;
    iss >> s1;
    if (s1 == "charges") return new NBCharge(iss, box);
    if (s1 == "erf2") return new NBErf2(iss, box);
    if (s1 == "erfG") return new NBErfG(iss, box);
    if (s1 == "gaussian") return new NBGauss(iss, box);
    if (s1 == "bias") return new BiasField(iss, box);
    if (s1 == "wall") return new Wall(iss, box);
    if (s1 == "Langevin" || s1 == "langevin") return new Langevin(iss, box);
    if (s1 == "dpd") return new DPD(iss, box);
    if (s1 == "bias_lc" ||
        s1 == "maiersaupe")
        die("ps_potential: " + s1 + " is not part of this build's particle-mesh path");
    die("ps_potential.cu: " + s1 + " is not a supported potential");
}

// -------------------------------------------------------- integrators ---
Integrator::Integrator(std::istringstream& iss, PS_Box* box) : mybox(box)
{
    if (!(iss >> groupName >> name)) die("Integrator needs a group and a style");
    std::string word;
    if (iss >> word) {
        if (word == "delt") iss >> delt;
        else die("Invalid keyword in Integrator command");
    }
}

Integrator* IntegratorFactory(std::istringstream& iss, PS_Box* box)
{
    std::streampos pos = iss.tellg();
    std::string s1;
    iss >> s1 >> s1;
    iss.clear();
    iss.seekg(pos);
    if (s1 == "EM") return new EM(iss, box);
    if (s1 == "GJF") return new GJF(iss, box);
    if (s1 == "VV") return new VV(iss, box);
    die(s1 + " is not a valid Integrator");
}

// --------------------------------------------------------------- box ----
PS_Box::PS_Box(std::istringstream& iss) : Box(iss) { boxStyle = "ps"; }

PS_Box::~PS_Box()
{
    for (auto* p : potentials) delete p;
    for (auto* i : integrators) delete i;
    for (auto* c : computes) delete c;
    if (eng) mft_tild_destroy(eng);
}

int PS_Box::findSpeciesInteger(const std::string& s) const
{
    for (size_t i = 0; i < species.size(); ++i)
        if (species[i].speciesLabel == s) return (int)i;
    die("Species " + s + " not found");
}

int PS_Box::findGroupInteger(const std::string& s) const
{
    for (size_t i = 0; i < psGroup.size(); ++i)
        if (psGroup[i].name == s) return (int)i;
    die("Group " + s + " not found");
}

// groups of the reference: "all", one per species label, "charges", and
// user groups (PS_Box_initialization.cu:571-585)
uint32_t PS_Box::groupMask(const std::string& s) const
{
    if (s == "all" || s == "All") return species.size() >= 32 ? 0xffffffffu : (1u << species.size()) - 1u;
    for (size_t i = 0; i < species.size(); ++i)
        if (species[i].speciesLabel == s) return 1u << i;
    for (const auto& g : psGroup)
        if (g.name == s) return g.speciesMask;
    die("Group " + s + " not found");
}

void PS_Box::grow(int n)
{
    x.resize(3 * (size_t)n, 0.f);
    charges.resize(n, 0.f);
    intSpecies.resize(n, 0);
    mID.resize(n, 0);
    nBonds.resize(n, 0);
    bondedTo.resize((size_t)n * MAXBONDS, 0);
    bondType.resize((size_t)n * MAXBONDS, 0);
    nAngles.resize(n, 0);
    angleType.resize((size_t)n * MAXANGLES, 0);
    angleGroup.resize((size_t)n * MAXANGLES * 3, 0);
}

// one angle (i1, i2 vertex, i3) entered on the lists of its three members
// (PS_Box_moleculeMakers.cu:242-288, PS_Box_initialization.cu:962-985)
void PS_Box::addAngle(int i1, int i2, int i3, int type)
{
    for (int p : {i1, i2, i3}) {
        const int na = nAngles[p];
        if (na >= MAXANGLES) die("MAXANGLES too small for the angles of particle " + std::to_string(p + 1));
        int* g = &angleGroup[((size_t)p * MAXANGLES + na) * 3];
        g[0] = i1;
        g[1] = i2;
        g[2] = i3;
        angleType[(size_t)p * MAXANGLES + na] = type;
        nAngles[p] = na + 1;
    }
}

void PS_Box::enableCharges() { doCharges = 1; }

void PS_Box::readInput(std::ifstream& inp)
{
    std::string line, firstWord;
    bool ended = false;
    while (!ended && std::getline(inp, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream iss(line);
        while (iss >> firstWord) {
            if (firstWord == "endBox") {
                ended = true;
                break;
            }
            if (firstWord == "angle") {
                int atype;
                std::string style;
                iss >> atype >> style;
                if (atype < 1) die("angle type must be >= 1");
                if (atype > (int)angleK.size()) {
                    angleK.resize(atype, 0.f);
                    angleTheq.resize(atype, 0.f);
                    angleStyle.resize(atype, 0);
                }
                iss >> angleK[atype - 1];
                if (style == "wlc") {
                    angleStyle[atype - 1] = 0;
                } else if (style == "harmonic") {
                    angleStyle[atype - 1] = 1;
                    float th;
                    iss >> th;
                    angleTheq[atype - 1] = (float)(th * (3.141592654f / 180.0));
                } else {
                    die("Angle style not supported!");
                }
            }
            else if (firstWord == "blocksize" || firstWord == "blockSize") { int b; iss >> b; }
            else if (firstWord == "bond") {
                int btype;
                std::string style;
                iss >> btype >> style;
                if (btype < 1) die("bond type must be >= 1");
                if (btype > (int)bondK.size()) {
                    bondK.resize(btype);
                    bondReq.resize(btype);
                    bondStyle.resize(btype);
                }
                iss >> bondK[btype - 1] >> bondReq[btype - 1];
                if (style == "harmonic") bondStyle[btype - 1] = 0;
                else if (style == "fene" || style == "FENE") bondStyle[btype - 1] = 1;
                else die("Bond type not supported!");
            } else if (firstWord == "boxLengths") {
                read_lengths(iss);
            } else if (firstWord == "compute") {
                computes.push_back(PSComputeFactory(iss, this));
            } else if (firstWord == "datFileName" || firstWord == "dat_file_name") {
                iss >> datFileName;
            } else if (firstWord == "Dim") {
                iss >> Dim;
                if (Dim != 3) die("this build runs 3D boxes (Dim 3)");
                readDimension = true;
            } else if (firstWord == "doCharges") {
                enableCharges();
            } else if (firstWord == "trajFreq" || firstWord == "traj_freq") {
                // PS_Box_initialization.cu:314-315
                iss >> trajFreq;
            } else if (firstWord == "trajFileName") {
                iss >> trajFileName;
            } else if (firstWord == "gsdFreq" || firstWord == "gsd_freq") {
                // PS_Box_initialization.cu:199-211
                iss >> gsdFreq;
                std::string tps;
                if (iss >> tps) {
                    if (tps != "logspace" && tps != "log_space" && tps != "logSpace")
                        die("PS_Box_init::readInput: invalid option to gsdFreq");
                    logSpaceGSDflag = 1;
                    iss >> logScaleGSD;
                }
            } else if (firstWord == "gsdName" || firstWord == "gsd_name") {
                iss >> gsd_name;
            } else if (firstWord == "fieldFreq" || firstWord == "field_freq") {
                iss >> fieldFreq;
                std::string tps;
                if (iss >> tps) {
                    if (tps != "logspace" && tps != "log_space" && tps != "logSpace")
                        die("PS_Box_init::readInput: invalid option to gsdFreq");
                    logSpaceFieldflag = 1;
                    iss >> logScaleFields;
                }
            } else if (firstWord == "initDataFileName") {
                std::string s;
                iss >> s;
            } else if (firstWord == "grid") {
                read_grid(iss);
            } else if (firstWord == "group") {
                psGroup.push_back(PS_Group(iss, this));
            } else if (firstWord == "integrator") {
                integrators.push_back(IntegratorFactory(iss, this));
            } else if (firstWord == "logFreq" || firstWord == "logfreq") {
                iss >> logFreq;
            } else if (firstWord == "MAXANGLES") {
                if (nstot > 0) die("MAXANGLES must be set before defining any molecules");
                iss >> MAXANGLES;
                if (MAXANGLES < 1) die("MAXANGLES must be >= 1");
            } else if (firstWord == "MAXBONDS") {
                if (nstot > 0) die("MAXBONDS must be set before defining any molecules");
                iss >> MAXBONDS;
            } else if (firstWord == "molecule") {
                std::string next;
                iss >> next;
                if (next == "linear") makeLinear(iss);
                else if (next == "grafted" || next == "sclc")
                    die("molecule " + next + " is not part of this build's particle-mesh path");
                else die("PS_Box_init:readInput: invalid molecule type read");
            } else if (firstWord == "Nr") {
                double nr;
                iss >> nr;
            } else if (firstWord == "NSEXTRA" || firstWord == "nsextra") {
                int n;
                iss >> n;
            } else if (firstWord == "Nx" || firstWord == "Ny" || firstWord == "Nz") {
                int a = firstWord == "Nx" ? 0 : (firstWord == "Ny" ? 1 : 2);
                iss >> Nx[a];
                M *= Nx[a];
            } else if (firstWord == "pmeorder") {
                iss >> pmeorder;
            } else if (firstWord == "potential") {
                potentials.push_back(PSPotentialFactory(iss, this));
            } else if (firstWord == "neighbor_list") {
                // PS_NeighborList (ps_neighborList.cu:132-141): group rcut [maxNeighbors];
                // the engine searches cells of width >= rcut every step, with no cap
                NeighborList nl;
                int maxn;
                iss >> nl.group >> nl.rcut;
                if (iss >> maxn) nl.maxNeighbors = maxn;
                neighborLists.push_back(nl);
            } else if (firstWord == "randSeed" || firstWord == "RAND_SEED" || firstWord == "RANDSEED") {
                long long sd;
                iss >> sd;
                RANDSEED = (uint64_t)sd;
                rng.reseed(RANDSEED);
            } else if (firstWord == "readData") {
                std::string name;
                iss >> name;
                readDataConfig(name);
            } else if (firstWord == "rho0") {
                iss >> rho0;
            } else if (firstWord == "species") {
                species.push_back(PS_Species(iss, this));
            } else if (firstWord == "verbose") {
                int v;
                iss >> v;
            } else {
                die("Invalid keyword " + firstWord + " in FTS_Box::readInput()");
            }
        }
    }
    finishInitialization();
}

// PS_Box::makeLinear (PS_Box_moleculeMakers.cu:15-311): chain ends uniform in
// the (x/y/z)range, each further bead one unit vector from the previous,
// wrapped once; bonds to the previous and next beads.
void PS_Box::makeLinear(std::istringstream& iss)
{
    if (rho0 < 0.0) die("rho0 must be defined before molecules created!");
    double phi = -1.0;
    int nmolecs = -1, numBlocks = 0, Ntot = 0;
    std::string s1;
    iss >> s1;
    if (s1 == "phi") iss >> phi;
    else if (s1 == "nmolecs") iss >> nmolecs;
    else die("invalid keyword after molecule linear. must be 'phi' or 'nmolecs'");
    iss >> numBlocks;
    if (numBlocks < 1) die("molecule linear needs at least one block");
    std::vector<int> Nb(numBlocks), bType(numBlocks, 1), aType(numBlocks, 0);
    bool angleFlag = false;
    std::vector<std::string> sp(numBlocks);
    std::vector<float> bq(numBlocks, 0.f);
    bool molCharges = false;
    float Rmin[3] = {0.f, 0.f, 0.f}, Rmax[3] = {L[0], L[1], L[2]};
    for (int j = 0; j < numBlocks; ++j) {
        iss >> Nb[j] >> sp[j];
        Ntot += Nb[j];
    }
    while (iss >> s1) {
        if (s1 == "charge") {
            enableCharges();
            molCharges = true;
            for (int j = 0; j < numBlocks; ++j) iss >> bq[j];
        } else if (s1 == "bondType" || s1 == "bondtype") {
            for (int j = 0; j < numBlocks; ++j) iss >> bType[j];
        } else if (s1 == "angleType" || s1 == "angletype") {
            for (int j = 0; j < numBlocks; ++j) iss >> aType[j];
            angleFlag = true;
        } else if (s1 == "xrange") {
            iss >> Rmin[0] >> Rmax[0];
        } else if (s1 == "yrange") {
            iss >> Rmin[1] >> Rmax[1];
        } else if (s1 == "zrange") {
            iss >> Rmin[2] >> Rmax[2];
        } else if (s1 == "drude") {
            die("drude oscillators not implemented!");
        } else {
            die("invalid molecule option " + s1);
        }
    }
    if (phi > 0.0) nmolecs = int(rho0 * V * phi / float(Ntot));
    if (nmolecs < 0) die("molecule linear: negative molecule count");
    std::cout << "Generating " << nmolecs << " molecules each with " << Ntot << " sites" << std::endl;
    int ind = nstot;
    // the first new molecule takes the largest molecule id already in the box
    // (0 in an empty box), so a second molecule command starts at the id of
    // the previous command's last molecule, as the reference's search over
    // mID does (PS_Box_moleculeMakers.cu:140-145, 296, 306)
    const int molec0 = nstot > 0 ? std::max(0, *std::max_element(mID.begin(), mID.begin() + nstot)) : 0;
    nstot += nmolecs * Ntot;
    grow(nstot);
    std::vector<int> spi(numBlocks);
    for (int j = 0; j < numBlocks; ++j) spi[j] = findSpeciesInteger(sp[j]);
    for (int i = nstot - nmolecs * Ntot; i < nstot; ++i) nAngles[i] = 0;
    for (int i = 0; i < nmolecs; ++i) {
        // with angles the chain is laid out straight: one bond direction per
        // molecule, drawn at its second bead (PS_Box_moleculeMakers.cu:186-190)
        double ru[3] = {0.0, 0.0, 0.0};
        bool have_ru = false;
        for (int j = 0; j < numBlocks; ++j) {
            for (int s = 0; s < Nb[j]; ++s) {
                intSpecies[ind] = spi[j];
                if (j == 0 && s == 0) {
                    for (int a = 0; a < 3; ++a) x[ind * 3 + a] = (float)(rng.uniform() * (Rmax[a] - Rmin[a]) + Rmin[a]);
                } else {
                    if ((s == 1 && j == 0) || !angleFlag || !have_ru) rng.unit_vector(ru);
                    have_ru = true;
                    for (int a = 0; a < 3; ++a) {
                        float v = x[(ind - 1) * 3 + a] + (float)ru[a];
                        if (v > L[a]) v -= L[a];
                        else if (v < 0.f) v += L[a];
                        x[ind * 3 + a] = v;
                    }
                }
                if (molCharges) charges[ind] = bq[j];
                nBonds[ind] = 0;
                if (j != 0 || s != 0) {
                    bondedTo[ind * MAXBONDS + nBonds[ind]] = ind - 1;
                    bondType[ind * MAXBONDS + nBonds[ind]] = bType[j] - 1;
                    nBonds[ind]++;
                }
                if (j != numBlocks - 1 || s != Nb[j] - 1) {
                    if (nBonds[ind] >= MAXBONDS) die("MAXBONDS too small for linear molecules");
                    bondedTo[ind * MAXBONDS + nBonds[ind]] = ind + 1;
                    bondType[ind * MAXBONDS + nBonds[ind]] = bType[j] - 1;
                    nBonds[ind]++;
                }
                // an angle at every interior bead of a block with an angle type
                if (aType[j] != 0 && (j > 0 || s > 0) && (j < numBlocks - 1 || s < Nb[j] - 1))
                    addAngle(ind - 1, ind, ind + 1, aType[j] - 1);
                mID[ind] = molec0 + i;
                ++ind;
            }
        }
    }
}

// PS_Box::readDataConfig (PS_Box_initialization.cu:787-990): LAMMPS-style
// data file with Masses, Atoms (id mol type [q] x y z), Bonds, Angles.
void PS_Box::readDataConfig(const std::string& name)
{
    if (nstot != 0) die("Must read data file before making additional molecules");
    std::ifstream in(name);
    if (!in) die("Failed to open " + name);
    std::string line;
    auto next_nonblank = [&](std::string& out) {
        while (std::getline(in, out))
            if (out.find_first_not_of(" \t\r") != std::string::npos) return true;
        return false;
    };
    std::getline(in, line); // title
    int nat = 0, nb = 0, nang = 0, ntypes = 0, nbt = 0, nat_t = 0;
    double lo, hi;
    while (next_nonblank(line)) {
        std::istringstream ls(line);
        if (line.find("atoms") != std::string::npos && line.find("types") == std::string::npos) ls >> nat;
        else if (line.find("bonds") != std::string::npos && line.find("types") == std::string::npos) ls >> nb;
        else if (line.find("angles") != std::string::npos && line.find("types") == std::string::npos) ls >> nang;
        else if (line.find("atom types") != std::string::npos) ls >> ntypes;
        else if (line.find("bond types") != std::string::npos) ls >> nbt;
        else if (line.find("angle types") != std::string::npos) ls >> nat_t;
        else if (line.find("xlo") != std::string::npos) { ls >> lo >> hi; L[0] = (float)(hi - lo); }
        else if (line.find("ylo") != std::string::npos) { ls >> lo >> hi; L[1] = (float)(hi - lo); }
        else if (line.find("zlo") != std::string::npos) { ls >> lo >> hi; L[2] = (float)(hi - lo); }
        else if (line.find("Masses") != std::string::npos) break;
    }
    (void)nat_t;
    (void)nbt;
    V = (double)L[0] * L[1] * L[2];
    if ((int)species.size() < ntypes) die("readData: define every species before readData");
    for (int i = 0; i < ntypes; ++i) {
        next_nonblank(line);
        std::istringstream ls(line);
        int t;
        float m;
        ls >> t >> m;
        if (t >= 1 && t <= (int)species.size()) species[t - 1].mass = m;
    }
    next_nonblank(line); // Atoms
    nstot = nat;
    grow(nstot);
    for (int i = 0; i < nat; ++i) {
        if (!next_nonblank(line)) die("Premature end of " + name);
        std::istringstream ls(line);
        int id, mol, typ;
        ls >> id >> mol >> typ;
        if (id < 1 || id > nat) die("readData: atom id out of range");
        int k = id - 1;
        mID[k] = mol - 1;
        intSpecies[k] = typ - 1;
        if (doCharges) ls >> charges[k];
        for (int a = 0; a < 3; ++a) ls >> x[k * 3 + a];
    }
    if (nb > 0) {
        next_nonblank(line); // Bonds
        for (int i = 0; i < nb; ++i) {
            if (!next_nonblank(line)) die("Premature end of " + name);
            std::istringstream ls(line);
            int cnt = 0, bt = 0, i1 = 0, i2 = 0;
            ls >> cnt >> bt >> i1 >> i2;
            if (!ls) die("readData: malformed bond line '" + line + "'");
            if (std::min(i1, i2) < 1 || std::max(i1, i2) > nat) die("readData: bond partner out of range");
            if (bt < 1 || bt > std::max(nbt, (int)bondK.size())) die("readData: bond type out of range");
            --bt;
            --i1;
            --i2;
            if (i2 < i1) std::swap(i1, i2);
            if (nBonds[i1] >= MAXBONDS || nBonds[i2] >= MAXBONDS) die("readData: MAXBONDS exceeded");
            bondedTo[i1 * MAXBONDS + nBonds[i1]] = i2;
            bondType[i1 * MAXBONDS + nBonds[i1]++] = bt;
            bondedTo[i2 * MAXBONDS + nBonds[i2]] = i1;
            bondType[i2 * MAXBONDS + nBonds[i2]++] = bt;
        }
    }
    if (nang > 0) {
        next_nonblank(line); // Angles
        for (int i = 0; i < nang; ++i) {
            if (!next_nonblank(line)) die("Premature end of " + name);
            std::istringstream ls(line);
            int cnt, at, i1, i2, i3;
            ls >> cnt >> at >> i1 >> i2 >> i3;
            if (i3 < i1) std::swap(i1, i3);
            if (std::min(i1, std::min(i2, i3)) < 1 || std::max(i1, std::max(i2, i3)) > nat)
                die("readData: angle member out of range");
            addAngle(i1 - 1, i2 - 1, i3 - 1, at - 1);
        }
    }
    nBondsTot = nb;
}

// finishInitialization (PS_Box_initialization.cu:352-426, 551-587): groups,
// the engine description, topology and state on the device.
void PS_Box::finishInitialization()
{
    if (M <= 1 || Nx[0] <= 0) die("grid must be defined");
    if (L[0] <= 0.f) die("boxLengths must be defined");
    if (species.empty()) die("no species defined");
    if (nstot == 0) die("Box created with no particles");
    if (!topology_dump.empty()) {
        FILE* fp = std::fopen(topology_dump.c_str(), "w");
        if (!fp) die("Failed to open " + topology_dump);
        for (int i = 0; i < nstot; ++i) {
            std::fprintf(fp, "%d %.9g %d %d", intSpecies[i], doCharges ? charges[i] : 0.f, mID[i], nBonds[i]);
            for (int b = 0; b < MAXBONDS; ++b) std::fprintf(fp, " %d", b < nBonds[i] ? bondedTo[i * MAXBONDS + b] : -1);
            for (int b = 0; b < MAXBONDS; ++b) std::fprintf(fp, " %d", b < nBonds[i] ? bondType[i * MAXBONDS + b] : -1);
            std::fprintf(fp, "\n");
        }
        std::fclose(fp);
    }
    if (integrators.size() != 1) die("this build runs exactly one integrator on group all");
    if (integrators[0]->groupName != "all") die("integrator group must be 'all' in this build");
    nBondsTot = 0;
    for (int i = 0; i < nstot; ++i) nBondsTot += nBonds[i];
    nBondsTot /= 2;
    // angles counted once, at their vertex (PS_Box_initialization.cu:429-444)
    nAnglesTot = 0;
    for (int i = 0; i < nstot; ++i)
        for (int j = 0; j < nAngles[i]; ++j)
            if (angleGroup[((size_t)i * MAXANGLES + j) * 3 + 1] == i) ++nAnglesTot;
    for (int i = 0; i < nstot; ++i)
        for (int j = 0; j < nAngles[i]; ++j) {
            const int t = angleType[(size_t)i * MAXANGLES + j];
            if (t < 0 || t >= (int)angleK.size()) die("angle type " + std::to_string(t + 1) + " is not defined");
        }
    // the reference checks neutrality when charges are on (PS_Box_initialization.cu:398-412)
    std::vector<mft_tild_potential> pd;
    for (auto* p : potentials) pd.push_back(p->descriptor());
    std::vector<float> mass, mob;
    for (const auto& s : species) {
        mass.push_back(s.mass);
        mob.push_back(s.mobility);
    }
    mft_tild_desc d{};
    d.nx = Nx[0];
    d.ny = Nx[1];
    d.nz = Nx[2];
    d.lx = L[0];
    d.ly = L[1];
    d.lz = L[2];
    d.pmeorder = pmeorder;
    d.nparticles = nstot;
    d.nspecies = (int)species.size();
    d.maxbonds = MAXBONDS;
    d.nbondtypes = (int)bondK.size();
    d.bond_k = bondK.data();
    d.bond_r0 = bondReq.data();
    d.bond_style = bondStyle.data();
    d.npotentials = (int)pd.size();
    d.potentials = pd.data();
    d.integrator = integrators[0]->engine_kind();
    d.delt = integrators[0]->delt;
    d.mass = mass.data();
    d.mobility = mob.data();
    d.seed = RANDSEED;
    d.has_charges = doCharges;
    if (nAnglesTot > 0) {
        d.maxangles = MAXANGLES;
        d.nangletypes = (int)angleK.size();
        d.angle_k = angleK.data();
        d.angle_theta = angleTheq.data();
        d.angle_style = angleStyle.data();
    }
    if (dry_run) {
        std::printf("{\"box\": \"ps\", \"grid\": [%d, %d, %d], \"L\": [%.9g, %.9g, %.9g], \"pmeorder\": %d, "
                    "\"nparticles\": %d, \"nspecies\": %d, \"maxbonds\": %d, \"nbonds\": %d, \"integrator\": %d, "
                    "\"delt\": %.9g, \"seed\": %llu, \"has_charges\": %d, \"potentials\": [",
                    d.nx, d.ny, d.nz, d.lx, d.ly, d.lz, d.pmeorder, d.nparticles, d.nspecies, d.maxbonds, nBondsTot,
                    d.integrator, d.delt, (unsigned long long)d.seed, d.has_charges);
        for (size_t i = 0; i < pd.size(); ++i)
            std::printf("%s{\"kind\": %d, \"maskI\": %u, \"maskJ\": %u, \"A\": %.9g, \"sigma\": %.9g, \"Rp\": %.9g, "
                        "\"ramp\": %d, \"A_final\": %.9g, \"phase\": %d, \"dir\": %d, \"n_periods\": %d, "
                        "\"wall_pos\": %.9g, \"wall_sign\": %d}",
                        i ? ", " : "", pd[i].kind, pd[i].maskI, pd[i].maskJ, pd[i].A, pd[i].sigma, pd[i].Rp,
                        pd[i].ramp, pd[i].A_final, pd[i].phase, pd[i].dir, pd[i].n_periods, pd[i].wall_pos,
                        pd[i].wall_sign);
        std::printf("], \"bonds\": [");
        for (size_t i = 0; i < bondK.size(); ++i)
            std::printf("%s[%d, %.9g, %.9g]", i ? ", " : "", bondStyle[i], bondK[i], bondReq[i]);
        std::printf("], \"nangles\": %d, \"maxangles\": %d, \"angles\": [", nAnglesTot, MAXANGLES);
        for (size_t i = 0; i < angleK.size(); ++i)
            std::printf("%s[%d, %.9g, %.9g]", i ? ", " : "", angleStyle[i], angleK[i], angleTheq[i]);
        std::printf("], \"computes\": [");
        for (size_t i = 0; i < computes.size(); ++i)
            std::printf("%s{\"style\": \"avg_sk\", \"mask\": %u, \"wait\": %d, \"freq\": %d, \"file\": \"%s\"}",
                        i ? ", " : "", groupMask(computes[i]->group), computes[i]->compute_wait,
                        computes[i]->compute_freq, computes[i]->output_name.c_str());
        std::printf("]}\n");
        std::fflush(stdout);
        initialized = true;
        return;
    }
    check_mft(mft_tild_create(&eng, &d), "mft_tild_create");
    check_mft(mft_tild_set_topology(eng, intSpecies.data(), doCharges ? charges.data() : nullptr, nBonds.data(),
                                    bondedTo.data(), bondType.data()),
              "mft_tild_set_topology");
    if (nAnglesTot > 0)
        check_mft(mft_tild_set_angles(eng, nAngles.data(), angleType.data(), angleGroup.data()), "mft_tild_set_angles");
    check_mft(mft_tild_set_state(eng, x.data(), nullptr, nullptr), "mft_tild_set_state");
    int choice = 0;
    check_mft(mft_tild_autotune(eng, 3, &choice), "mft_tild_autotune");
    for (auto* p : potentials) p->initializePotential();
    for (auto* i : integrators) i->finishInitialization();
    for (auto* c : computes) c->initialize_compute();
    initFieldFiles();
    initialized = true;
}

// writeData (PS_Box.cu:339-372): step, Ubond (with bonds), potential energies
void PS_Box::computeThermoProps()
{
    std::vector<double> pe(std::max<size_t>(1, potentials.size()), 0.0);
    double be = 0.0;
    if (stepEnergiesValid) {
        // the energies of the step just taken (PS_Box::writeData after doTimeStep)
        pe = stepPe;
        be = stepBe;
    } else {
        check_mft(mft_tild_energies(eng, pe.data(), &be), "mft_tild_energies");
    }
    Ubond = (float)be;
    double ae = 0.0;
    if (nAnglesTot > 0) check_mft(mft_tild_angle_energy(eng, &ae), "mft_tild_angle_energy");
    Uangle = (float)ae;
    Upe = Ubond + Uangle;
    for (size_t i = 0; i < potentials.size(); ++i) {
        potentials[i]->energy = (float)pe[i];
        Upe += (float)pe[i];
    }
}

void PS_Box::writeData(int step)
{
    computeThermoProps();
    std::ofstream otp(datFileName, std::ios_base::app);
    otp << step;
    std::cout << "step: " << step << " Upe: " << Upe;
    if (nBondsTot > 0) {
        otp << " " << Ubond;
        std::cout << " Ubond: " << Ubond;
    }
    if (nAnglesTot > 0) {
        otp << " " << Uangle;
        std::cout << " Uangle: " << Uangle;
    }
    for (size_t i = 0; i < potentials.size(); ++i) {
        otp << " " << potentials[i]->energy;
        std::cout << " pot[" << i << "]: " << potentials[i]->energy;
    }
    otp << std::endl;
    std::cout << std::endl;
}

// doTimeStep (PS_Box.cu:73-207) as one engine step; potentials' host copies
// follow update_prefactor; the log at step % logFreq == 0.
void PS_Box::doTimeStep(int step)
{
    for (auto* i : integrators) i->Integrate_1();
    if (logFreq > 0 && step % logFreq == 0) {
        takeSteps(1, true);
    } else {
        takeSteps(1, false);
    }
    for (auto* p : potentials) p->update_prefactor(step, maxSteps);
    for (auto* c : computes) c->do_compute(step);
    if (logFreq > 0 && step % logFreq == 0) {
        writeData(step);
        for (auto* c : computes) c->write_output();
    }
    if (gsdFreq > 0 && step % gsdFreq == 0) {
        writeGSDtraj();
        // PS_Box.cu:181-188: log-spaced frames
        if (step > 1 && logSpaceGSDflag == 1) gsdFreq = int(float(gsdFreq) * logScaleGSD);
    }
    if (trajFreq > 0 && step % trajFreq == 0) writeLammpsTraj(step);
    if (fieldFreq > 0 && step % fieldFreq == 0) {
        writeFields();
        // PS_Box.cu:198-203: log-spaced frames
        if (step > 1 && logSpaceFieldflag == 1) fieldFreq = int(float(fieldFreq) * logScaleFields);
    }
}

// PS_Box::writeGSDtraj (PS_Box_gsdUtils.cu:14-160): the first frame (at
// totSteps 0) creates the file and carries the static data -- box, N, masses,
// type ids (species + 1), molecule ids (+ 1), type names, bonds (i < j pairs,
// GSDinit, PS_Box_gsdUtils.cu:586-605) and zero angles; every frame carries the
// step and positions shifted by -L/2 (and charges when they are on)
void PS_Box::writeGSDtraj()
{
    const int n = nstot;
    if (totSteps == 0 || !gsd) {
        gsd.reset(new GsdWriter(gsd_name, "gpu-tild", "hoomd", (1u << 16) | 4u));
        const uint8_t dims = 3;
        const uint64_t step = (uint64_t)totSteps;
        gsd->chunk("configuration/step", GsdWriter::UINT64, 1, 1, &step);
        gsd->chunk("configuration/dimensions", GsdWriter::UINT8, 1, 1, &dims);
        const float box[6] = {L[0], L[1], L[2], 0.f, 0.f, 0.f};
        gsd->chunk("configuration/box", GsdWriter::FLOAT, 6, 1, box);
        const uint32_t N = (uint32_t)n;
        gsd->chunk("particles/N", GsdWriter::UINT32, 1, 1, &N);
        std::vector<float> masses(n);
        std::vector<uint32_t> types(n), mols(n);
        for (int i = 0; i < n; ++i) {
            masses[i] = species[intSpecies[i]].mass;
            types[i] = (uint32_t)intSpecies[i] + 1;
            mols[i] = (uint32_t)mID[i] + 1;
        }
        gsd->chunk("particles/mass", GsdWriter::FLOAT, n, 1, masses.data());
        gsd->chunk("particles/typeid", GsdWriter::UINT32, n, 1, types.data());
        gsd->chunk("log/particles/moleculeid", GsdWriter::UINT32, n, 1, mols.data());
        const int ntypes = (int)species.size() + 1, max_len = 10;
        std::vector<char> names((size_t)ntypes * max_len, 0);
        for (int i = 0; i < ntypes; ++i) {
            const std::string t = "type" + std::to_string(i);
            std::memcpy(&names[(size_t)i * max_len], t.c_str(), std::min<size_t>(t.size() + 1, max_len));
        }
        gsd->chunk("particles/types", GsdWriter::INT8, ntypes, max_len, names.data());
        std::vector<uint32_t> btype, bpair;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < nBonds[i]; ++j) {
                const int k = bondedTo[(size_t)i * MAXBONDS + j];
                if (k > i) {
                    bpair.push_back((uint32_t)i);
                    bpair.push_back((uint32_t)k);
                    btype.push_back((uint32_t)bondType[(size_t)i * MAXBONDS + j]);
                }
            }
        std::vector<uint32_t> atype, agroup;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < nAngles[i]; ++j) {
                const int* g = &angleGroup[((size_t)i * MAXANGLES + j) * 3];
                if (g[1] != i) continue; // each angle once, at its vertex
                atype.push_back((uint32_t)angleType[(size_t)i * MAXANGLES + j]);
                for (int c = 0; c < 3; ++c) agroup.push_back((uint32_t)g[c]);
            }
        const uint32_t nb = (uint32_t)btype.size(), na = (uint32_t)atype.size();
        gsd->chunk("bonds/N", GsdWriter::UINT32, 1, 1, &nb);
        gsd->chunk("bonds/typeid", GsdWriter::UINT32, nb, 1, btype.data());
        gsd->chunk("bonds/group", GsdWriter::UINT32, nb, 2, bpair.data());
        gsd->chunk("angles/N", GsdWriter::UINT32, 1, 1, &na);
        gsd->chunk("angles/typeid", GsdWriter::UINT32, na, 1, atype.data());
        gsd->chunk("angles/group", GsdWriter::UINT32, na, 3, agroup.data());
    } else {
        const uint64_t step = (uint64_t)totSteps;
        gsd->chunk("configuration/step", GsdWriter::UINT64, 1, 1, &step);
    }
    std::vector<float> xc(3 * (size_t)n);
    check_mft(mft_tild_get_state(eng, xc.data(), nullptr, nullptr), "mft_tild_get_state");
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < 3; ++j) xc[(size_t)i * 3 + j] -= 0.5f * L[j];
    gsd->chunk("particles/position", GsdWriter::FLOAT, n, 3, xc.data());
    if (doCharges) gsd->chunk("particles/charge", GsdWriter::FLOAT, n, 1, charges.data());
    gsd->end_frame();
}

// PS_Box::writeKFieldFloat (PS_Box.cu:531-560): kx ky kz |v| |k| v per grid
// point, k from Box::get_kD (float, PI2 = 6.2831853071, Nyquist negative)
void PS_Box::writeKFieldFloat(const std::string& name, const float* dat)
{
    FILE* otp = std::fopen(name.c_str(), "w");
    if (!otp) die("Failed to open output file in writeFieldFloat!");
    const float PI2 = 6.2831853071f;
    for (long i = 0; i < M; ++i) {
        const int n[3] = {(int)(i % Nx[0]), (int)((i / Nx[0]) % Nx[1]), (int)(i / ((long)Nx[0] * Nx[1]))};
        float kv[3], k2 = 0.0f;
        for (int j = 0; j < 3; ++j) {
            kv[j] = (2 * n[j] < Nx[j] ? PI2 * (float)n[j] : PI2 * (float)(n[j] - Nx[j])) / L[j];
            k2 += kv[j] * kv[j];
        }
        for (int j = 0; j < 3; ++j) std::fprintf(otp, "%f ", kv[j]);
        std::fprintf(otp, "%1.8e %1.8e %1.8e\n", std::fabs(dat[i]), std::sqrt(k2), dat[i]);
    }
    std::fclose(otp);
}

int PS_Compute::total_computes = 0;

PS_Compute::PS_Compute(std::istringstream& iss, PS_Box* box) : mybox(box)
{
    compute_id = total_computes++;
    output_name = "compute-" + std::to_string(compute_id) + ".dat";
    iss >> group;
    std::string s1;
    while (iss >> s1) {
        if (s1 == "wait") iss >> compute_wait;
        else if (s1 == "freq") iss >> compute_freq;
        else if (s1 == "file_name") iss >> output_name;
        else die("PS_Compute: " + s1 + " invalid option for compute!");
    }
}

PS_Compute* PSComputeFactory(std::istringstream& iss, PS_Box* box)
{
    std::string s1;
    iss >> s1;
    if (s1 == "avg_sk") return new PS_ComputeSK(iss, box);
    die("Invalid compute option!");
}

PS_ComputeSK::PS_ComputeSK(std::istringstream& iss, PS_Box* box) : PS_Compute(iss, box)
{
    std::cout << "  Computing s(k) for group " << group << std::endl;
}

PS_ComputeSK::~PS_ComputeSK()
{
    if (d_sk) mft_free(d_sk);
    if (d_work) mft_free(d_work);
}

void PS_ComputeSK::initialize_compute()
{
    mask = mybox->groupMask(group);
    const size_t M = mybox->gridSize();
    check_mft(mft_malloc((void**)&d_sk, sizeof(float) * M), "mft_malloc");
    check_mft(mft_malloc(&d_work, 8 * M), "mft_malloc");
    check_mft(mft_memset(d_sk, 0, sizeof(float) * M), "mft_memset");
}

void PS_ComputeSK::do_compute(int step)
{
    if (!active(step)) return;
    check_mft(mft_tild_structure_factor(mybox->engine(), mask, d_sk, 1, d_work), "mft_tild_structure_factor");
    num_data_pts += 1;
}

// the average over the accumulated frames (ps_compute_sk.cu:49-59)
void PS_ComputeSK::write_output()
{
    const size_t M = mybox->gridSize();
    std::vector<float> sk(M);
    check_mft(mft_memcpy_d2h(sk.data(), d_sk, sizeof(float) * M), "mft_memcpy_d2h");
    for (auto& v : sk) v *= 1.0 / float(num_data_pts);
    mybox->writeKFieldFloat(output_name, sk.data());
}

// PS_Box::writeLammpsTraj (PS_Box.cu:756-801): positions wrapped once into
// the box, "id type mol x y z" per particle; step 0 starts the file
void PS_Box::writeLammpsTraj(int step)
{
    std::vector<float> x(3 * (size_t)nstot);
    check_mft(mft_tild_get_state(eng, x.data(), nullptr, nullptr), "mft_tild_get_state");
    for (int i = 0; i < nstot; ++i)
        for (int j = 0; j < 3; ++j) {
            float& v = x[(size_t)i * 3 + j];
            if (v > L[j]) v -= L[j];
            else if (v < 0.0f) v += L[j];
        }
    FILE* otp = std::fopen(trajFileName.c_str(), step == 0 ? "w" : "a");
    if (!otp) die("Failed to open " + trajFileName);
    std::fprintf(otp, "ITEM: TIMESTEP\n%d\nITEM: NUMBER OF ATOMS\n%d\n", step, nstot);
    std::fprintf(otp, "ITEM: BOX BOUNDS pp pp pp\n");
    std::fprintf(otp, "%f %f\n%f %f\n%f %f\n", 0.f, L[0], 0.f, L[1], 0.f, L[2]);
    std::fprintf(otp, "ITEM: ATOMS id type mol x y z\n");
    for (int i = 0; i < nstot; ++i) {
        std::fprintf(otp, "%d %d %d  ", i + 1, intSpecies[i] + 1, mID[i] + 1);
        for (int j = 0; j < 3; ++j) std::fprintf(otp, "%f ", x[(size_t)i * 3 + j]);
        std::fprintf(otp, "\n");
    }
    std::fclose(otp);
}

void PS_Box::NVT(int nSteps)
{
    if (dry_run) return;
    writeData(-1);
    maxSteps = nSteps;
    check_mft(mft_tild_set_run(eng, nSteps), "mft_tild_set_run");
    // steps between log points run as one engine call
    int i = 0;
    while (i < maxSteps) {
        // next log, trajectory or compute step
        int next = maxSteps;
        if (logFreq > 0) next = std::min(next, (i / logFreq + (i % logFreq ? 1 : 0)) * logFreq);
        if (trajFreq > 0) next = std::min(next, (i / trajFreq + (i % trajFreq ? 1 : 0)) * trajFreq);
        if (gsdFreq > 0) next = std::min(next, (i / gsdFreq + (i % gsdFreq ? 1 : 0)) * gsdFreq);
        if (fieldFreq > 0) next = std::min(next, (i / fieldFreq + (i % fieldFreq ? 1 : 0)) * fieldFreq);
        for (auto* c : computes)
            for (int s = i; s < next; ++s)
                if (c->active(s)) {
                    next = s;
                    break;
                }
        if (next > i) {
            // the last chunk keeps its energies for the final log row
            takeSteps(next - i, next == maxSteps);
            for (int s = i; s < next; ++s)
                for (auto* p : potentials) p->update_prefactor(s, maxSteps);
            totSteps += next - i;
            i = next;
            continue;
        }
        doTimeStep(i);
        ++totSteps;
        ++i;
    }
    check_mft(mft_tild_synchronize(eng), "mft_tild_synchronize");
    std::cout << "NVT Finished, writing final output" << std::endl;
    writeData(maxSteps);
    writeFields();
    writeDataConfig("final.input.data");
}

// Binary density files of the reference (Box::initBinaryDataFile and
// writeBinaryData, Box.cu:553-582; PS_Group::writeDensityFieldBinary,
// ps_groups.cu:257-262): "density-<group>.bin" for "all", every species, the
// "charges" group and the user groups, starting with int Dim, int Nx[Dim],
// float L[Dim], then one frame of M floats (the group's number density, x
// fastest) per write.
namespace {
struct DensityGroup {
    std::string name;
    uint32_t mask;     // species whose densities sum to the group's
};
}

static std::vector<DensityGroup> density_groups(const PS_Box& b, const std::vector<PS_Species>& species,
                                                const std::vector<PS_Group>& groups, bool charges,
                                                uint32_t charged_mask)
{
    (void)b;
    std::vector<DensityGroup> out;
    const uint32_t all = species.size() >= 32 ? 0xffffffffu : (1u << species.size()) - 1u;
    out.push_back({"all", all});
    for (size_t i = 0; i < species.size(); ++i) out.push_back({species[i].speciesLabel, 1u << i});
    if (charges) out.push_back({"charges", charged_mask});
    for (const auto& g : groups) out.push_back({g.name, g.speciesMask});
    return out;
}

// species whose every particle carries a charge (their union is the charges
// group); a species with charged and neutral particles makes it unavailable
static bool charged_species(const std::vector<int>& type, const std::vector<float>& q, int ns, int nspecies,
                            uint32_t* mask)
{
    std::vector<int> nq(nspecies, 0), n(nspecies, 0);
    for (int i = 0; i < ns; ++i) {
        n[type[i]]++;
        if (q[i] != 0.f) nq[type[i]]++;
    }
    *mask = 0;
    for (int s = 0; s < nspecies; ++s) {
        if (nq[s] != 0 && nq[s] != n[s]) return false;
        if (nq[s] > 0) *mask |= 1u << s;
    }
    return true;
}

void PS_Box::initFieldFiles()
{
    uint32_t qmask = 0;
    const bool uniform = charged_species(intSpecies, charges, nstot, (int)species.size(), &qmask);
    for (const auto& g : density_groups(*this, species, psGroup, doCharges, qmask)) {
        if (g.name == "charges" && !uniform) {
            std::cout << "note: density-charges.bin is not written (a species holds charged and neutral "
                         "particles)" << std::endl;
            continue;
        }
        std::ofstream otp("density-" + g.name + ".bin", std::ios::binary);
        if (!otp.is_open()) die("Failed to initialize output binary file!");
        const int dim = 3;
        const int nx[3] = {Nx[0], Nx[1], Nx[2]};
        otp.write(reinterpret_cast<const char*>(&dim), sizeof(int));
        otp.write(reinterpret_cast<const char*>(nx), 3 * sizeof(int));
        otp.write(reinterpret_cast<const char*>(L), 3 * sizeof(float));
    }
}

void PS_Box::writeFields()
{
    if (dry_run) return;
    // the densities of the last density pass (the step's, before the
    // integrator moved the particles); the velocity-Verlet step takes its
    // densities after the drift, at the positions it ends with
    if (integrators.size() == 1 && integrators[0]->name == "VV") {
        std::vector<double> pe(std::max<size_t>(1, potentials.size()));
        check_mft(mft_tild_energies(eng, pe.data(), nullptr), "mft_tild_energies");
    }
    const size_t ns = species.size();
    std::vector<float> rho(ns * (size_t)M);
    check_mft(mft_tild_get_densities(eng, rho.data(), nullptr), "mft_tild_get_densities");
    uint32_t qmask = 0;
    const bool uniform = charged_species(intSpecies, charges, nstot, (int)ns, &qmask);
    std::vector<float> frame((size_t)M);
    for (const auto& g : density_groups(*this, species, psGroup, doCharges, qmask)) {
        if (g.name == "charges" && !uniform) continue;
        for (long i = 0; i < M; ++i) {
            float v = 0.f;
            for (size_t sp = 0; sp < ns; ++sp)
                if ((g.mask >> sp) & 1u) v += rho[sp * (size_t)M + i];
            frame[i] = v;
        }
        std::ofstream otp("density-" + g.name + ".bin", std::ios::app | std::ios::binary);
        if (!otp.is_open()) die("Failed to open output binary file!");
        otp.write(reinterpret_cast<const char*>(frame.data()), sizeof(float) * (size_t)M);
    }
}

// n engine steps; with energies, the log energies of the last one are kept
// for writeData (pair energies from that step's densities, PS_Box.cu:152-173)
void PS_Box::takeSteps(int n, bool energies)
{
    if (energies) {
        stepPe.assign(std::max<size_t>(1, potentials.size()), 0.0);
        check_mft(mft_tild_step_energies(eng, n, stepPe.data(), &stepBe), "mft_tild_step_energies");
        stepEnergiesValid = true;
    } else {
        check_mft(mft_tild_step(eng, n), "mft_tild_step");
        stepEnergiesValid = false;
    }
}

// LAMMPS-style data file of the current state (PS_Box.cu:629-720)
void PS_Box::writeDataConfig(const std::string& name)
{
    std::vector<float> xc(3 * (size_t)nstot);
    check_mft(mft_tild_get_state(eng, xc.data(), nullptr, nullptr), "mft_tild_get_state");
    FILE* otp = std::fopen(name.c_str(), "w");
    if (!otp) die("Failed to open " + name);
    std::fprintf(otp, "Generated by the MI355X front end\n\n%d atoms\n%d bonds\n%d angles\n\n", nstot, nBondsTot,
                 nAnglesTot);
    std::fprintf(otp, "%d atom types\n%d bond types\n%d angle types\n\n", (int)species.size(), (int)bondK.size(),
                 (int)angleK.size());
    std::fprintf(otp, "0 %f xlo xhi\n0 %f ylo yhi\n0 %f zlo zhi\n\nMasses\n\n", L[0], L[1], L[2]);
    for (size_t i = 0; i < species.size(); ++i) std::fprintf(otp, "%zu %f\n", i + 1, species[i].mass);
    std::fprintf(otp, "\nAtoms\n\n");
    for (int i = 0; i < nstot; ++i) {
        std::fprintf(otp, "%d %d %d ", i + 1, mID[i] + 1, intSpecies[i] + 1);
        if (doCharges) std::fprintf(otp, "%f ", charges[i]);
        std::fprintf(otp, "%.9g %.9g %.9g\n", xc[3 * i], xc[3 * i + 1], xc[3 * i + 2]);
    }
    std::fprintf(otp, "\nBonds\n\n");
    int b = 0;
    for (int i = 0; i < nstot; ++i)
        for (int k = 0; k < nBonds[i]; ++k) {
            int j = bondedTo[i * MAXBONDS + k];
            if (j > i) std::fprintf(otp, "%d %d %d %d\n", ++b, bondType[i * MAXBONDS + k] + 1, i + 1, j + 1);
        }
    if (nAnglesTot > 0) {
        // PS_Box.cu:728-745: each angle once, from its first member
        std::fprintf(otp, "\nAngles\n\n");
        int a = 0;
        for (int i = 0; i < nstot; ++i)
            for (int k = 0; k < nAngles[i]; ++k) {
                const int* g = &angleGroup[((size_t)i * MAXANGLES + k) * 3];
                if (g[0] == i)
                    std::fprintf(otp, "%d %d %d %d %d\n", ++a, angleType[(size_t)i * MAXANGLES + k] + 1, g[0] + 1,
                                 g[1] + 1, g[2] + 1);
            }
    }
    std::fclose(otp);
}
