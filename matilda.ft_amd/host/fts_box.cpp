// Field-theoretic (CL-FTS / SCFT) box of the drop-in front end: the
// reference's FTS_Box keywords, potential and molecule classes and log output
// over the mft_fts engine.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <deque>
#include <sstream>
#include <fstream>
#include <iostream>

#include "matilda.h"

// FTS_Species (fts_species.cu:10-30): "species A"
FTS_Species::FTS_Species(std::istringstream& iss, FTS_Box* box)
{
    (void)box;
    if (!(iss >> fts_species)) die("Missing label for species");
}

// --------------------------------------------------------- potentials ---
FTS_Potential::FTS_Potential(std::istringstream& iss, FTS_Box* box) : input_command(iss.str()), mybox(box)
{
    wpl.assign(box->gridSize(), {0.0, 0.0});
}

// FTS_Potential::initializeField (fts_potential.cu:62-150): value, random, sine
void FTS_Potential::initializeField(std::istringstream& iss, std::vector<std::complex<double>>& w)
{
    std::string s1;
    iss >> s1;
    const long M = mybox->gridSize();
    if (s1 == "value") {
        double rv, iv;
        iss >> rv >> iv;
        for (auto& v : w) v = {rv, iv};
    } else if (s1 == "random") {
        double ra, ia;
        iss >> ra >> ia;
        for (long i = 0; i < M; ++i) {
            double r = ra * mybox->rng.uniform();
            double im = ia * mybox->rng.uniform();
            w[i] = {r, im};
        }
    } else if (s1 == "sin" || s1 == "sine") {
        // (real_amp + i imag_amp) sin(2 PI r_dir period / L_dir), r = n dx
        // (fts_potential.cu:88-100, Box::get_r Box.cu:219-232, PI of constants.h)
        int dir = 0;
        double ra = 0.0, ia = 0.0, period = 1.0;
        iss >> dir >> ra >> ia >> period;
        if (dir < 0 || dir > 2) die("initialize sin: direction must be 0, 1 or 2");
        const double PIf = (double)3.141592654f;
        const int* Nx = mybox->gridNx();
        const float* L = mybox->boxL();
        const float dxd = L[dir] / (float)Nx[dir];
        for (long i = 0; i < M; ++i) {
            const long n[3] = {i % Nx[0], (i / Nx[0]) % Nx[1], i / ((long)Nx[0] * Nx[1])};
            const double r = (double)n[dir] * (double)dxd;
            const double v = std::sin(2.0 * PIf * r * period / (double)L[dir]);
            w[i] = {ra * v, ia * v};
        }
    } else if (s1 == "L" || s1 == "LAM" || s1 == "C" || s1 == "CYL" || s1 == "H" || s1 == "HEX" || s1 == "G" ||
               s1 == "GYR") {
        // an ordered phase: Ao [dir d] [n_periods n] (fts_potential.cu:102-138)
        const int phase = (s1 == "L" || s1 == "LAM") ? MFT_BIAS_LAM : (s1 == "G" || s1 == "GYR") ? MFT_BIAS_GYR
                                                                                                  : MFT_BIAS_HEX;
        double Ao = 0.0;
        int dir = 0, nper = 1;
        iss >> Ao;
        std::string opt;
        while (iss >> opt) {
            if (opt == "dir") iss >> dir;
            else if (opt == "n_periods") iss >> nper;
            else die(opt + " is not a valid initialize option in fts_potential");
        }
        std::cout << "phase read as " << s1 << " Ao: " << Ao << std::endl;
        std::vector<double> u((size_t)M);
        check_mft(mft_bias_field(u.data(), mybox->gridNx(), mybox->boxL(), phase, Ao, dir, nper), "mft_bias_field");
        for (long i = 0; i < M; ++i) w[i] = {u[i], 0.0};
    } else if (s1 == "BCC" || s1 == "SPH" || s1 == "S") {
        die("initialize " + s1 + ": the BCC phase is not offered (the reference builds it from an uninitialised "
            "vector, Box.cu:359)");
    } else if (s1 == "readDatFile") {
        // Box::readDatFile (Box.cu:622-656): per line the coordinates, then
        // Re w and Im w.  Like the reference (fts_potential.cu:142-146), the
        // file always fills this potential's w+, whichever field the
        // initialize names (so a Flory potential's w+ is then set to -1.5i
        // anyway, and Helfand's and Edwards' never reach the device)
        std::vector<std::complex<double>>& dst = (&w == &wmi) ? wpl : w;
        std::string name;
        iss >> name;
        std::ifstream inp(name);
        if (!inp.is_open()) die("Failed to open " + name);
        std::string line;
        for (long i = 0; i < M; ++i) {
            if (!std::getline(inp, line)) die("readDatFile: " + name + " ends before the grid does");
            std::istringstream ls(line);
            double xd, wr = 0.0, wi = 0.0;
            for (int j = 0; j < 3; ++j) ls >> xd;
            ls >> wr >> wi;
            dst[i] = {wr, wi};
        }
    } else {
        die("Invalid initialize option on potential " + potentialStyle);
    }
}

void FTS_Potential::parse_options(std::istringstream& iss, std::vector<std::complex<double>>* init_target, bool flory)
{
    std::string s1;
    while (iss >> s1) {
        if (s1 == "initialize") {
            std::vector<std::complex<double>> scratch(mybox->gridSize());
            initializeField(iss, init_target ? *init_target : scratch);
            if (flory) {
                // fts_potential_flory.cu:83-88: initialize also sets w+ = -1.5i
                for (auto& v : wpl) v = {0.0, -1.5};
            }
        } else if (s1 == "updateScheme") {
            iss >> updateScheme;
            if (updateScheme != "EM" && updateScheme != "EMPC" && updateScheme != "1S")
                die("updateScheme " + updateScheme + " is not part of this build (EM, EMPC, 1S)");
        } else if (s1 == "modify") {
            iss >> s1;
            // the incompressible potential reuses Helfand's message
            // (fts_potential_incompress.cu:50-53)
            if (s1 != "zeromean" && s1 != "zeroMean")
                die("Invalid modify option on " + std::string(potentialStyle == "Incompress" ? "Helfand" : potentialStyle) +
                    " potential");
            zeroMean = true;
        } else {
            die("Invalid option " + s1 + " on " + potentialStyle + " potential");
        }
    }
}

// writeFields of each potential (fts_potential_flory.cu:307-321,
// fts_potential_helfand.cu:243-250, fts_potential_incompress.cu:221-228,
// fts_potential_edwards.cu:245-252): the current fields as text grids
void FTS_Potential::writeFields(int potInd, mft_fts* sys)
{
    const long M = mybox->gridSize();
    std::vector<float> a(2 * (size_t)M), b(has_wmi ? 2 * (size_t)M : 0);
    check_mft(mft_fts_get_fields(sys, potInd, a.data(), has_wmi ? b.data() : nullptr), "mft_fts_get_fields");
    std::vector<std::complex<double>> w((size_t)M);
    for (long i = 0; i < M; ++i) w[i] = {a[2 * i], a[2 * i + 1]};
    mybox->writeTComplexGridData("wpl_" + potentialStyle + std::to_string(potInd) + ".dat", w.data());
    if (has_wmi) {
        for (long i = 0; i < M; ++i) w[i] = {b[2 * i], b[2 * i + 1]};
        mybox->writeTComplexGridData("wmi_" + potentialStyle + std::to_string(potInd) + ".dat", w.data());
    }
}

// PotentialHelfand (fts_potential_helfand.cu:15-70): kappaN delt [options];
// its initialize never reaches the device (the field stays zero)
PotentialHelfand::PotentialHelfand(std::istringstream& iss, FTS_Box* box) : FTS_Potential(iss, box)
{
    potentialStyle = "Helfand";
    iss >> kappaN >> delt;
    parse_options(iss, nullptr, false);
}
mft_fts_potential PotentialHelfand::descriptor() const
{
    mft_fts_potential d{};
    d.kind = MFT_FTS_HELFAND;
    d.scheme = updateScheme == "EMPC" ? MFT_FTS_EMPC : (updateScheme == "1S" ? MFT_FTS_1S : MFT_FTS_EM);
    d.zero_mean = zeroMean ? 1 : 0;
    d.kappaN = kappaN;
    d.delt_pl = delt;
    return d;
}

// PotentialFlory (fts_potential_flory.cu:21-120): I J chiN deltPlus deltMinus
PotentialFlory::PotentialFlory(std::istringstream& iss, FTS_Box* box) : FTS_Potential(iss, box)
{
    potentialStyle = "Flory";
    std::string a, b;
    iss >> a >> b;
    actsOn = {a, b};
    I = box->findSpeciesInteger(a);
    J = box->findSpeciesInteger(b);
    iss >> chiN >> deltPlus >> deltMinus;
    wmi.assign(box->gridSize(), {0.0, 0.0});
    has_wmi = true;
    parse_options(iss, &wmi, true);
}
mft_fts_potential PotentialFlory::descriptor() const
{
    mft_fts_potential d{};
    d.kind = MFT_FTS_FLORY;
    d.scheme = updateScheme == "EMPC" ? MFT_FTS_EMPC : (updateScheme == "1S" ? MFT_FTS_1S : MFT_FTS_EM);
    d.zero_mean = zeroMean ? 1 : 0;
    d.I = I;
    d.J = J;
    d.chiN = chiN;
    d.delt_pl = deltPlus;
    d.delt_mi = deltMinus;
    return d;
}

// PotentialIncompress (fts_potential_incompress.cu:15-60): delt [options]
PotentialIncompress::PotentialIncompress(std::istringstream& iss, FTS_Box* box) : FTS_Potential(iss, box)
{
    potentialStyle = "Incompress";
    iss >> delt;
    parse_options(iss, &wpl, false);
}
mft_fts_potential PotentialIncompress::descriptor() const
{
    mft_fts_potential d{};
    d.kind = MFT_FTS_INCOMPRESS;
    d.scheme = updateScheme == "EMPC" ? MFT_FTS_EMPC : (updateScheme == "1S" ? MFT_FTS_1S : MFT_FTS_EM);
    d.zero_mean = zeroMean ? 1 : 0;
    d.delt_pl = delt;
    return d;
}

// PotentialEdwards (fts_potential_edwards.cu:12-75): B delt [options]; like
// Helfand's, its initialize never reaches the device
PotentialEdwards::PotentialEdwards(std::istringstream& iss, FTS_Box* box) : FTS_Potential(iss, box)
{
    potentialStyle = "Edwards";
    iss >> B >> delt;
    parse_options(iss, nullptr, false);
}
mft_fts_potential PotentialEdwards::descriptor() const
{
    mft_fts_potential d{};
    d.kind = MFT_FTS_EDWARDS;
    d.scheme = updateScheme == "EMPC" ? MFT_FTS_EMPC : (updateScheme == "1S" ? MFT_FTS_1S : MFT_FTS_EM);
    d.zero_mean = zeroMean ? 1 : 0;
    d.kappaN = B;
    d.delt_pl = delt;
    return d;
}

FTS_Potential* FTS_PotentialFactory(std::istringstream& iss, FTS_Box* box)
{
    std::string s1;
    iss >> s1;
    if (s1 == "Helfand" || s1 == "helfand") return new PotentialHelfand(iss, box);
    if (s1 == "incompress" || s1 == "Incompress" || s1 == "incompressible") return new PotentialIncompress(iss, box);
    if (s1 == "Flory" || s1 == "flory") return new PotentialFlory(iss, box);
    if (s1 == "Edwards" || s1 == "edwards") return new PotentialEdwards(iss, box);
    if (s1 == "particle" || s1 == "Particle") die(s1 + " is not part of this build's field-theory path");
    die(s1 + " is not a valid FTS_Potential");
}

// ---------------------------------------------------------- molecules ---
// FTS_Molec::FTS_Molec (fts_molecule.cu:13-46): type, then phi | nmolecs | activity
FTS_Molec::FTS_Molec(std::istringstream& iss, FTS_Box* box) : input_command(iss.str()), mybox(box)
{
    iss >> molec_type;
    std::string s1;
    iss >> s1;
    if (s1 == "phi") iss >> phi;
    else if (s1 == "nmolecs") iss >> nmolecs;
    else if (s1 == "activity") iss >> activity;
    else die(s1 + " is not a valid way to define the number of particles/molecules");
    // the reference recomputes nmolecs from phi (fts_molecule_linear.cu:433-435),
    // which leaves a molecule given by 'nmolecs' without a density factor
    if (phi <= 0.0 && activity <= 0.0) die("this build's linear molecules take 'phi' or 'activity'");
}

// modifyMolecParam (fts_molecule.cu:77-88)
static void modifyMolecParam(double& param, const std::string& how2Mod, double change)
{
    if (how2Mod == "scale") param *= change;
    else if (how2Mod == "delta") param += change;
    else if (how2Mod == "value") param = change;
    else die(how2Mod + " is not a valid way to modify a molecule parameter");
}

// LinearMolec::modifyMolecule (fts_molecule_linear.cu:386-425)
void FTS_Molec::modifyMolecule(std::istringstream& iss)
{
    std::string property, how2Mod;
    iss >> property;
    if (property == "activity") {
        double change = 0.0;
        iss >> how2Mod >> change;
        std::cout << "activity changed from " << activity;
        modifyMolecParam(activity, how2Mod, change);
        std::cout << " to " << activity << std::endl;
        mybox->setActivity(this, activity);
    } else if (property == "phi") {
        die("changes in FTS:molec:phi not yet supported");
    } else if (property == "nmolecs") {
        die("changes in FTS:molec:nmolecs not yet supported");
    } else {
        die(property + " is not a valid property to change in FTS:molecule");
    }
}

// LinearMolec (fts_molecule_linear.cu:17-140): numBlocks, (N, species) per block
LinearMolec::LinearMolec(std::istringstream& iss, FTS_Box* box) : FTS_Molec(iss, box)
{
    int numBlocks = 0;
    iss >> numBlocks;
    if (numBlocks < 1 || numBlocks > 16) die("linear molecule needs 1..16 blocks");
    N.resize(numBlocks);
    blockSpecies.resize(numBlocks);
    for (int j = 0; j < numBlocks; ++j) {
        iss >> N[j] >> blockSpecies[j];
        Ntot += N[j];
    }
    std::string s1;
    if (iss >> s1) {
        if (s1 == "smear" || s1 == "Smear") die("smeared molecules are not part of this build");
        die("Invalid option " + s1 + " on molecule linear");
    }
}

mft_fts_molecule LinearMolec::descriptor() const
{
    mft_fts_molecule d{};
    d.nblocks = (int)N.size();
    for (size_t j = 0; j < N.size(); ++j) {
        d.N[j] = N[j];
        d.species[j] = mybox->findSpeciesInteger(blockSpecies[j]);
    }
    d.phi = phi;
    d.activity = activity > 0.0 ? activity : 0.0;
    return d;
}

FTS_Molec* FTS_MolecFactory(std::istringstream& iss, FTS_Box* box)
{
    std::streampos pos = iss.tellg();
    std::string s1;
    iss >> s1;
    iss.clear();
    iss.seekg(pos);
    if (s1 == "linear") return new LinearMolec(iss, box);
    if (s1 == "HParticle") die("HParticle molecules are not part of this build");
    die(s1 + " is not a valid FTS_Molec");
}

// --------------------------------------------------------------- box ----
FTS_Box::FTS_Box(std::istringstream& iss) : Box(iss)
{
    boxStyle = "fts";
    iss >> ftsStyle;
    if (ftsStyle != "scft" && ftsStyle != "cl") die("Invalid FTS simulation box style! Must be scft or cl.");
}

FTS_Box::~FTS_Box()
{
    for (auto* p : Potentials) delete p;
    for (auto* m : Molecs) delete m;
    if (sys) mft_fts_destroy(sys);
}

int FTS_Box::findSpeciesInteger(const std::string& s) const
{
    for (size_t i = 0; i < Species.size(); ++i)
        if (Species[i].fts_species == s) return (int)i;
    die(s + " species not found");
}

void FTS_Box::readInput(std::ifstream& inp)
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
            if (firstWord == "boxLengths") read_lengths(iss);
            else if (firstWord == "C") iss >> C;
            else if (firstWord == "chemFieldFreq" || firstWord == "chemfieldfreq") iss >> chemFieldFreq;
            else if (firstWord == "clSampleFreq" || firstWord == "clsamplefreq") iss >> clSampleFreq;
            else if (firstWord == "densityFieldFreq" || firstWord == "densityfieldfreq") iss >> densityFieldFreq; else if (firstWord == "Dim") {
                iss >> Dim;
                if (Dim != 3) die("this build runs 3D boxes (Dim 3)");
                readDimension = true;
            } else if (firstWord == "grid") {
                read_grid(iss);
            } else if (firstWord == "logFreq" || firstWord == "logfreq") {
                iss >> logFreq;
            } else if (firstWord == "molecule") {
                Molecs.push_back(FTS_MolecFactory(iss, this));
            } else if (firstWord == "Nr") {
                iss >> Nr;
            } else if (firstWord == "Nx" || firstWord == "Ny" || firstWord == "Nz") {
                int a = firstWord == "Nx" ? 0 : (firstWord == "Ny" ? 1 : 2);
                iss >> Nx[a];
                M *= Nx[a];
            } else if (firstWord == "potential") {
                Potentials.push_back(FTS_PotentialFactory(iss, this));
            } else if (firstWord == "randSeed" || firstWord == "RAND_SEED") {
                long long sd;
                iss >> sd;
                RANDSEED = (uint64_t)sd;
                rng.reseed(RANDSEED);
            } else if (firstWord == "rho0") {
                iss >> rho0;
            } else if (firstWord == "species") {
                Species.push_back(FTS_Species(iss, this));
            } else if (firstWord == "threads") {
                int t;
                iss >> t;
            } else if (firstWord == "tolerance") {
                iss >> tolMetric >> tolerance;
                if (tolMetric != "Heff") die("tolerance metric " + tolMetric + " is not part of this build (Heff)");
            } else {
                die("Invalid keyword " + firstWord + " in FTS_Box::readInput()");
            }
        }
    }
    finishInitialization();
}

void FTS_Box::finishInitialization()
{
    if (Nr <= 0.0) die("Nr must be defined for FTS boxes");
    if (C <= 0.0) die("C must be defined for FTS boxes");
    if (M <= 1) die("grid must be defined");
    if (Species.empty() || Molecs.empty() || Potentials.empty()) die("FTS box needs species, molecules and potentials");
    if (Species.size() > 8 || Potentials.size() > 8 || Molecs.size() > 4)
        die("at most 8 species, 8 potentials and 4 molecule types");
    mft_fts_desc d{};
    d.nx = Nx[0];
    d.ny = Nx[1];
    d.nz = Nx[2];
    d.lx = L[0];
    d.ly = L[1];
    d.lz = L[2];
    d.Nr = Nr;
    d.C = C;
    d.nspecies = (int)Species.size();
    d.npotentials = (int)Potentials.size();
    for (size_t i = 0; i < Potentials.size(); ++i) d.potentials[i] = Potentials[i]->descriptor();
    d.nmolecules = (int)Molecs.size();
    for (size_t i = 0; i < Molecs.size(); ++i) d.molecules[i] = Molecs[i]->descriptor();
    d.cl = ftsStyle == "cl" ? 1 : 0;
    d.seed = RANDSEED;
    if (dry_run) {
        std::printf("{\"box\": \"fts\", \"style\": \"%s\", \"grid\": [%d, %d, %d], \"L\": [%.9g, %.9g, %.9g], "
                    "\"Nr\": %.9g, \"C\": %.9g, \"nspecies\": %d, \"seed\": %llu, \"potentials\": [",
                    ftsStyle.c_str(), d.nx, d.ny, d.nz, d.lx, d.ly, d.lz, d.Nr, d.C, d.nspecies,
                    (unsigned long long)d.seed);
        for (int i = 0; i < d.npotentials; ++i) {
            const auto& p = d.potentials[i];
            std::printf("%s{\"kind\": %d, \"I\": %d, \"J\": %d, \"chiN\": %.9g, \"kappaN\": %.9g, \"delt_pl\": %.9g, "
                        "\"delt_mi\": %.9g, \"scheme\": %d, \"zero_mean\": %d, \"wpl0\": [%.9g, %.9g]}",
                        i ? ", " : "", p.kind, p.I, p.J, p.chiN, p.kappaN, p.delt_pl, p.delt_mi, p.scheme, p.zero_mean,
                        Potentials[i]->wpl[0].real(), Potentials[i]->wpl[0].imag());
        }
        std::printf("], \"molecules\": [");
        for (int m = 0; m < d.nmolecules; ++m) {
            std::printf("%s{\"phi\": %.9g, \"activity\": %.9g, \"blocks\": [", m ? ", " : "", d.molecules[m].phi,
                        d.molecules[m].activity);
            for (int j = 0; j < d.molecules[m].nblocks; ++j)
                std::printf("%s[%d, %d]", j ? ", " : "", d.molecules[m].N[j], d.molecules[m].species[j]);
            std::printf("]}");
        }
        std::printf("]}\n");
        std::fflush(stdout);
        initialized = true;
        return;
    }
    check_mft(mft_fts_create(&sys, &d), "mft_fts_create");
    std::vector<float> a(2 * (size_t)M), b(2 * (size_t)M);
    for (size_t p = 0; p < Potentials.size(); ++p) {
        // Helfand's and Edwards' initialize never reach the device
        // (fts_potential_helfand.cu, fts_potential_edwards.cu)
        const bool helfand = Potentials[p]->potentialStyle == "Helfand" || Potentials[p]->potentialStyle == "Edwards";
        for (long i = 0; i < M; ++i) {
            a[2 * i] = helfand ? 0.f : (float)Potentials[p]->wpl[i].real();
            a[2 * i + 1] = helfand ? 0.f : (float)Potentials[p]->wpl[i].imag();
            if (Potentials[p]->has_wmi) {
                b[2 * i] = (float)Potentials[p]->wmi[i].real();
                b[2 * i + 1] = (float)Potentials[p]->wmi[i].imag();
            }
        }
        check_mft(mft_fts_set_fields(sys, (int)p, a.data(), Potentials[p]->has_wmi ? b.data() : nullptr),
                  "mft_fts_set_fields");
    }
    check_mft(mft_fts_init(sys), "mft_fts_init");
    initialized = true;
}

// computeHamiltonian (FTS_Box.cu:590-602): FP64 reductions on the device
void FTS_Box::computeHamiltonian()
{
    std::vector<double> Q(2 * Molecs.size()), parts(2 * (Potentials.size() + Molecs.size()));
    double H[2];
    check_mft(mft_fts_observables(sys, Q.data(), H, parts.data()), "mft_fts_observables");
    Heff = {H[0], H[1]};
    for (size_t p = 0; p < Potentials.size(); ++p) Potentials[p]->Hterm = {parts[2 * p], parts[2 * p + 1]};
    for (size_t m = 0; m < Molecs.size(); ++m) {
        Molecs[m]->Q = {Q[2 * m], Q[2 * m + 1]};
        size_t k = Potentials.size() + m;
        Molecs[m]->Hterm = {parts[2 * k], parts[2 * k + 1]};
    }
    updateNsites();
}

// nSites of every molecule at the last density pass: activity V Re Q Ntot
// (fts_molecule_linear.cu:344-346); canonical molecules report nmolecs Ntot
void FTS_Box::updateNsites()
{
    std::vector<double> n(Molecs.size());
    check_mft(mft_fts_nsites(sys, n.data()), "mft_fts_nsites");
    for (size_t m = 0; m < Molecs.size(); ++m) Molecs[m]->nSites = n[m];
}

void FTS_Box::setActivity(const FTS_Molec* mol, double activity)
{
    if (dry_run) return;
    for (size_t m = 0; m < Molecs.size(); ++m)
        if (Molecs[m] == mol) check_mft(mft_fts_set_activity(sys, (int)m, activity), "mft_fts_set_activity");
}

// writeData (FTS_Box.cu:145-190): step Re(H) [Im(H) | error] Hterms -n logQ
void FTS_Box::writeData(int step)
{
    computeHamiltonian();
    if (std::isnan(Heff.real())) die("Found Heff = NaN, quitting!");
    std::ofstream otp("fts_data.dat", std::ios_base::app);
    std::ostringstream line;
    line << step << " " << Heff.real() << " ";
    if (ftsStyle == "cl") line << Heff.imag() << " ";
    else line << error << " ";
    for (auto* p : Potentials) line << p->Hterm.real() << " ";
    // -n log Q, or -z V Q followed by nSites for a grand canonical molecule
    for (auto* m : Molecs) {
        line << m->calcHTerm().real() << " ";
        if (m->activity > 0.0) line << m->nSites << "  ";
    }
    otp << line.str() << std::endl;
    std::cout << line.str() << std::endl;
    // chemical potentials mu = ln(n / (V Q)) of every molecule (ln z for a
    // grand canonical one) in a file of their own, so fts_data.dat keeps the
    // reference's columns: step Re(mu) Im(mu) per molecule
    std::vector<double> mu(2 * Molecs.size());
    check_mft(mft_fts_chemical_potentials(sys, mu.data()), "mft_fts_chemical_potentials");
    std::ofstream omu("fts_mu.dat", std::ios_base::app);
    omu << step;
    omu.precision(12);
    for (size_t m = 0; m < Molecs.size(); ++m) omu << " " << mu[2 * m] << " " << mu[2 * m + 1];
    omu << std::endl;
}

// converged (FTS_Box.cu:715-760): SCFT only, Heff change at log steps
int FTS_Box::converged(int step)
{
    if (ftsStyle != "scft") return 0;
    if (logFreq <= 0 || step % logFreq != 0 || step < logFreq) return 0;
    error = std::fabs(Heff.real() - Hold);
    Hold = Heff.real();
    if (error < tolerance) {
        std::cout << "SCFT converged on step " << step << " using " << tolMetric << " < " << tolerance << std::endl;
        return 1;
    }
    return 0;
}

// the I/O of FTS_Box::doTimeStep (FTS_Box.cu:84-119) after iteration step
bool FTS_Box::ioStep(int step) const
{
    return (logFreq > 0 && step % logFreq == 0) || (densityFieldFreq > 0 && step % densityFieldFreq == 0) ||
           (chemFieldFreq > 0 && step % chemFieldFreq == 0) ||
           (ftsStyle == "cl" && clSampleFreq > 0 && densityFieldFreq > 0 && step % clSampleFreq == 0);
}

// FTS_Box::writeTComplexGridData (FTS_Box.cu:284-306): per grid point the
// coordinates r = n dx ("%lf "), then "%1.4e %1.4e" of Re and Im
void FTS_Box::writeTComplexGridData(const std::string& name, const std::complex<double>* field) const
{
    FILE* otp = std::fopen(name.c_str(), "w");
    if (!otp) die("Failed to open " + name);
    const float dx[3] = {L[0] / (float)Nx[0], L[1] / (float)Nx[1], L[2] / (float)Nx[2]};
    for (long i = 0; i < M; ++i) {
        const long n[3] = {i % Nx[0], (i / Nx[0]) % Nx[1], i / ((long)Nx[0] * Nx[1])};
        for (int j = 0; j < 3; ++j) std::fprintf(otp, "%lf ", (double)n[j] * (double)dx[j]);
        std::fprintf(otp, "%1.4e %1.4e\n", field[i].real(), field[i].imag());
    }
    std::fclose(otp);
}

// CL density samples every clSampleFreq steps (their running average is
// written with the densities), rhoSpecies<i>.dat [rhoSpeciesAvg<i>.dat] every
// densityFieldFreq steps, the potentials' fields and fieldSpecies<i>.dat every
// chemFieldFreq steps (fts_species.cu:136-172)
void FTS_Box::writeFieldFiles(int step)
{
    const size_t ns = Species.size(), m = (size_t)M;
    const bool sample = ftsStyle == "cl" && clSampleFreq > 0 && densityFieldFreq > 0 && step % clSampleFreq == 0;
    const bool dens = densityFieldFreq > 0 && step % densityFieldFreq == 0;
    const bool chem = chemFieldFreq > 0 && step % chemFieldFreq == 0;
    if (!sample && !dens && !chem) return;
    std::vector<float> buf(2 * ns * m);
    std::vector<std::complex<double>> f(m);
    if (sample || dens) {
        check_mft(mft_fts_get_densities(sys, buf.data()), "mft_fts_get_densities");
        if (sample) {
            if (densityAvg.empty()) densityAvg.assign(ns * m, {0.0, 0.0});
            for (size_t i = 0; i < ns * m; ++i) densityAvg[i] += std::complex<double>(buf[2 * i], buf[2 * i + 1]);
            ++nCLSamples;
        }
        if (dens) {
            for (size_t s = 0; s < ns; ++s) {
                for (size_t i = 0; i < m; ++i) f[i] = {buf[2 * (s * m + i)], buf[2 * (s * m + i) + 1]};
                writeTComplexGridData("rhoSpecies" + std::to_string(s) + ".dat", f.data());
                if (ftsStyle == "cl" && nCLSamples > 0) {
                    for (size_t i = 0; i < m; ++i) f[i] = densityAvg[s * m + i] / (double)nCLSamples;
                    writeTComplexGridData("rhoSpeciesAvg" + std::to_string(s) + ".dat", f.data());
                }
            }
        }
    }
    if (chem) {
        for (size_t p = 0; p < Potentials.size(); ++p) Potentials[p]->writeFields((int)p, sys);
        check_mft(mft_fts_get_species_fields(sys, buf.data()), "mft_fts_get_species_fields");
        for (size_t s = 0; s < ns; ++s) {
            for (size_t i = 0; i < m; ++i) f[i] = {buf[2 * (s * m + i)], buf[2 * (s * m + i) + 1]};
            writeTComplexGridData("fieldSpecies" + std::to_string(s) + ".dat", f.data());
        }
    }
}

void FTS_Box::doTimeStep(int step)
{
    for (auto* p : Potentials) p->updateFields();
    for (auto* m : Molecs) m->calcDensity();
    check_mft(mft_fts_step(sys, 1), "mft_fts_step");
    writeFieldFiles(step);
    if (logFreq > 0 && step % logFreq == 0) writeData(step);
}

// NVT (FTS_Box.cu:127-142): steps between log points run as one engine call
void FTS_Box::NVT(int nsteps)
{
    if (dry_run) return;
    int i = 0;
    while (i < nsteps) {
        // the next iteration with I/O (log, field files, CL samples)
        int next = i;
        while (next < nsteps && !ioStep(next)) ++next;
        if (next > i) {
            check_mft(mft_fts_step(sys, next - i), "mft_fts_step");
            totSteps += next - i;
            i = next;
            continue;
        }
        doTimeStep(i);
        if (i >= 50 && converged(i)) break;
        ++totSteps;
        ++i;
    }
    check_mft(mft_fts_synchronize(sys), "mft_fts_synchronize");
}

// FTS_Box::modifyBox (FTS_Box.cu:674-708): appends Re H and (nSites,
// activity) of every grand canonical molecule to sweep_data.dat, then applies
// "molecule i <property> ..." to molecule i (1-indexed)
void FTS_Box::modifyBox(std::istringstream& iss)
{
    std::ofstream otp("sweep_data.dat", std::ios_base::app);
    std::string word;
    iss >> word;
    if (word == "potential") die("potential modifications not yet supported");
    if (word == "molecule") {
        if (!dry_run && sys) updateNsites();
        otp << Heff.real() << " ";
        for (auto* m : Molecs)
            if (m->activity > 0.0) otp << m->nSites << " " << m->activity << " ";
        otp << std::endl;
        int molecInd = 0;
        iss >> molecInd;
        if (molecInd > (int)Molecs.size() || molecInd < 1) die("Invalid molecule index - out of range!");
        Molecs[molecInd - 1]->modifyMolecule(iss);
    }
}

// Spinodal search by activity sweep (the algorithm of
// FTS_BoxFindSpinodal.cu:25-240).  Between sweep points the activity of one
// grand canonical molecule is multiplied by f and the system re-equilibrated
// for nSteps; each point records every molecule's activity and site count,
// and the fields are kept in device snapshot slots (a ring of three, on the
// GPU; the reference copies them to the host).  The composition of the swept
// molecule is phi = n_k / sum_m n_m.  With the three latest points the search
// stops when the slope |d phi / d z| over the newest pair has fallen below a
// third of the slope over the previous pair.  Like the reference, slopes are
// first compared at the third scaled point, after the starting point has
// left the window, and the search ends at that first estimate (its
// refinement loop sits behind an unconditional break,
// FTS_BoxFindSpinodal.cu:214).  find_spinodal.dat gets one line per scaled
// point: "activity nSites" of every molecule.  maxTries (absent from the
// reference, which has no bound) ends a search that never meets the test.
namespace {
struct SweepPoint {
    std::vector<double> z, n;   // activity and site count of every molecule
    int slot = 0;               // device snapshot slot of the fields
};
}

void FTS_Box::findSpinodal(std::istringstream& iss)
{
    double phiTol = 1.0E-4, f = 0.8, fScaleFactor = 0.5;
    int molecIndex = 0, nSteps = 10000;
    long maxTries = -1;
    std::string fileName = "find_spinodal.dat", word;
    while (iss >> word) {
        if (word == "phiTolerance") iss >> phiTol;
        else if (word == "fScale") iss >> fScaleFactor;
        else if (word == "datFile") iss >> fileName;
        else if (word == "molecule") {
            iss >> molecIndex;
            molecIndex -= 1;
        } else if (word == "nSteps") iss >> nSteps;
        else if (word == "initial_f") iss >> f;
        else if (word == "maxTries") iss >> maxTries;
    }
    if (molecIndex < 0 || molecIndex >= (int)Molecs.size()) die("Invalid molecule index - out of range!");
    if (Molecs[molecIndex]->activity <= 0.0) die("findSpinodal scales the activity of a grand canonical molecule");
    if (dry_run) {
        std::printf("{\"findSpinodal\": {\"phiTol\": %.9g, \"f\": %.9g, \"fScale\": %.9g, \"molecule\": %d, "
                    "\"nSteps\": %d, \"datFile\": \"%s\", \"maxTries\": %ld}}\n",
                    phiTol, f, fScaleFactor, molecIndex, nSteps, fileName.c_str(), maxTries);
        std::fflush(stdout);
        return;
    }
    std::cout << "findSpinodal: molecule " << molecIndex + 1 << ", activity factor " << f << " per point, "
              << nSteps << " steps per point (phiTolerance " << phiTol << ", fScale " << fScaleFactor
              << " are read for the refinement stage, which the reference never reaches)" << std::endl;

    const size_t nMol = Molecs.size();
    std::ofstream dat(fileName);
    std::deque<SweepPoint> window;
    int nextSlot = 0;
    auto record = [&]() {
        updateNsites();
        SweepPoint p;
        for (auto* m : Molecs) {
            p.z.push_back(m->activity);
            p.n.push_back(m->nSites);
        }
        p.slot = nextSlot;
        nextSlot = (nextSlot + 1) % 3;
        check_mft(mft_fts_snapshot(sys, p.slot, 0), "mft_fts_snapshot");
        window.push_back(p);
        if (window.size() > 3) window.pop_front();
    };
    auto fraction = [&](const SweepPoint& p) {
        double tot = 0.0;
        for (size_t m = 0; m < nMol; ++m) tot += p.n[m];
        return p.n[molecIndex] / tot;
    };
    auto slope = [&](const SweepPoint& a, const SweepPoint& b) {
        return std::fabs((fraction(b) - fraction(a)) / (b.z[molecIndex] - a.z[molecIndex]));
    };

    NVT(nSteps);
    record();
    const std::string scale = "molecule " + std::to_string(molecIndex + 1) + " activity scale " + std::to_string(f);
    for (long tries = 0;; ++tries) {
        std::istringstream cmd(scale);
        modifyBox(cmd);
        NVT(nSteps);
        record();
        for (auto* m : Molecs) dat << m->activity << " " << m->nSites << " ";
        dat << std::endl;
        const SweepPoint& p = window.back();
        std::cout << "  point " << tries + 1 << ": z = " << p.z[molecIndex] << ", phi = " << fraction(p) << std::endl;
        if (tries >= 2) {
            const double older = slope(window[0], window[1]), newer = slope(window[1], window[2]);
            std::cout << "    |dphi/dz| " << older << " -> " << newer << std::endl;
            if (newer < older / 3.0) {
                std::cout << "findSpinodal: spinodal estimate between z = " << window[0].z[molecIndex] << " and "
                          << window[2].z[molecIndex] << " (phi " << fraction(window[0]) << " .. "
                          << fraction(window[2]) << ")" << std::endl;
                return;
            }
        }
        if (maxTries > 0 && tries + 1 >= maxTries) {
            std::cout << "findSpinodal: no spinodal point within " << maxTries << " tries" << std::endl;
            return;
        }
    }
}
