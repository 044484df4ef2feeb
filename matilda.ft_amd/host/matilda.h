// Drop-in host front end: the reference's input-file interpreter and its
// Box / Potential / Integrator / Group / Molecule class interfaces, kept with
// the same names, keywords and argument order, with bodies that drive the
// MI355X engine through the C ABI of mft.h.
//
// Interfaces mirrored (reference file:line):
//   Box           readInput, NVT, doTimeStep, writeData, findSpinodal, modifyBox (Box.h:80-101)
//   PS_Potential  initializePotential, CalcForces, CalcEnergy, update_prefactor(step, max),
//                 initBinaryOutput, writeBinaryOutput; grpI/J, Iind/Jind, energy (ps_potential.h:13-53)
//   Integrator    Integrate_1, Integrate_2, finishInitialization, delt (ps_integrator.h:21-39)
//   PS_Group      zeroFields, makeDensityField, mapForces, accumulateGridForceComp,
//                 enableForce (ps_groups.h:27-63)
//   FTS_Potential updateFields, correctFields, calcHamiltonian, initLinearCoeffs,
//                 storePredictorData, writeFields; wpl, wmi, Hterm, actsOn, updateScheme,
//                 zeroMean (fts_potential.h:27-64)
//   FTS_Molec     calcDensity, calcHTerm, computeLinearTerms, modifyMolecule,
//                 recomputeNmolecs; Q, nmolecs, activity, nSites (fts_molecule.h:25-76)
// Factories keep the reference's keywords (ps_potential.cu:294-333,
// ps_integrator.cu:51-69, fts_potential.cu:25-55, fts_molecule.cu:53-70,
// Box.cu:491-507).  The engine runs a whole step as one fused sequence, so
// per-stage methods whose work the engine does inside that sequence are kept
// for interface compatibility and do nothing on their own.
#pragma once
#include <complex>
#include <cstdint>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "mft.h"

// die() prints the message and exits with status 1 (die.cu:7-11); engine
// status codes other than MFT_OK are reported through it with mft_last_error().
[[noreturn]] void die(const std::string& msg);
void check_mft(int rc, const std::string& what);

// GSD 2.0 file writer for the trajectory frames of PS_Box::writeGSDtraj
// (gsd_writer.cpp): chunks of the current frame, then end_frame().
class GsdWriter {
public:
    enum Type { UINT8 = 1, UINT32 = 3, UINT64 = 4, INT8 = 5, INT32 = 7, FLOAT = 9 };
    GsdWriter(const std::string& path, const std::string& application, const std::string& schema,
              uint32_t schema_version);
    ~GsdWriter();
    void chunk(const std::string& name, Type type, uint64_t N, uint32_t M, const void* data);
    void end_frame();

private:
    struct Entry {
        uint64_t frame, N;
        int64_t location;
        uint32_t M;
        uint16_t id;
        uint8_t type;
    };
    void flush();
    std::string path_, application_, schema_;
    uint32_t schema_version_;
    int fd_ = -1;
    uint64_t data_end_ = 0, frame_ = 0;
    std::vector<std::string> names_;
    std::map<std::string, uint16_t> ids_;
    std::vector<Entry> entries_, pending_;
};

// Uniform deviates for the molecule makers and random field initialisation.
// A 64-bit xorshift* stream seeded by randSeed (the reference uses its own
// CPU generator; positions are statistically equivalent, not identical).
class HostRng {
public:
    explicit HostRng(uint64_t seed = 1) { reseed(seed); }
    void reseed(uint64_t seed) { s_ = seed * 0x9E3779B97F4A7C15ull + 0x2545F4914F6CDD1Dull; if (!s_) s_ = 1; }
    double uniform();            // (0, 1)
    void unit_vector(double* r); // uniform direction in 3D
private:
    uint64_t s_;
};

class Box {
public:
    explicit Box(std::istringstream& iss) { (void)iss; }
    virtual ~Box() {}
    virtual void readInput(std::ifstream& inp) = 0;
    virtual void NVT(int nsteps) = 0;
    virtual void doTimeStep(int step) = 0;
    virtual void writeData(int step) = 0;
    virtual void findSpinodal(std::istringstream& iss);
    virtual void modifyBox(std::istringstream& iss);
    virtual int converged(int step) { (void)step; return 0; }
    std::string returnBoxStyle() const { return boxStyle; }
    int returnDimension() const { return Dim; }
    const int* gridNx() const { return Nx; }
    const float* boxL() const { return L; }
    // print the engine descriptor instead of creating the engine
    static bool dry_run;
    // when not empty: the particle box writes its topology here once its
    // molecules and data files are read (one text line per particle)
    static std::string topology_dump;

protected:
    std::string boxStyle;
    int Dim = 3;
    int Nx[3] = {0, 0, 0};
    float L[3] = {0.f, 0.f, 0.f};
    long M = 1;
    double V = 1.0;
    int logFreq = 100;
    uint64_t RANDSEED = 0;
    bool readDimension = false;
    void read_grid(std::istringstream& iss);
    void read_lengths(std::istringstream& iss);
};

Box* BoxFactory(std::istringstream& iss);

// ------------------------------------------------------------------- PS ---
class PS_Box;

class PS_Species {
public:
    PS_Species(std::istringstream& iss, PS_Box* box);
    std::string speciesLabel;
    float mass = 1.0f, mobility = 1.0f;
};

class PS_Group {
public:
    PS_Group(std::istringstream& iss, PS_Box* box);     // "group name types A B ..."
    PS_Group(const std::string& name, uint32_t mask);
    std::string name;
    uint32_t speciesMask = 0; // species the group holds
    int forceFlag = 0;
    // density and force mapping run inside the engine's fused step
    void zeroFields() {}
    void makeDensityField() {}
    void mapForces() {}
    void accumulateGridForceComp() {}
    void enableForce() { forceFlag = 1; }
};

class PS_Potential {
public:
    PS_Potential(std::istringstream& iss, PS_Box* box);
    virtual ~PS_Potential() {}
    virtual void initializePotential() {}
    virtual void CalcForces() {}              // in the engine's fused step
    virtual float CalcEnergy() { return energy; }
    virtual void update_prefactor(int step, int maxSteps); // ps_potential.cu:227-248
    virtual void initBinaryOutput() {}
    virtual void writeBinaryOutput() {}
    virtual mft_tild_potential descriptor() const = 0;
    std::string input_command, grpI, grpJ;
    int Iind = -1, Jind = -1;
    float Ao = 0.f, Ao_initial = 0.f, Ao_final = 0.f, sig2 = 0.f, Rp = 0.f;
    bool ramp = false;
    float energy = 0.f;

protected:
    PS_Box* mybox;
    int ramp_check_input(std::istringstream& iss, float A); // ps_potential.cu:253-281
};

class NBGauss : public PS_Potential {
public:
    NBGauss(std::istringstream& iss, PS_Box* box); // ps_potentialGaussian.cu:14-26
    mft_tild_potential descriptor() const override;
};
class NBErf2 : public PS_Potential {
public:
    NBErf2(std::istringstream& iss, PS_Box* box); // ps_potentialErf2.cu:15-26
    mft_tild_potential descriptor() const override;
};
class NBErfG : public PS_Potential {
public:
    NBErfG(std::istringstream& iss, PS_Box* box); // ps_potentialErfG.cu
    mft_tild_potential descriptor() const override;
};
class Wall : public PS_Potential {
public:
    Wall(std::istringstream& iss, PS_Box* box); // ps_potentialWall.cu:10-48
    mft_tild_potential descriptor() const override;
    int normalDim = 0, dirSign = 1;
    float wallPos = 0.f, k = 0.f;
};
class DPD : public PS_Potential {
public:
    DPD(std::istringstream& iss, PS_Box* box); // ps_potentialDPD.cu:6-29
    mft_tild_potential descriptor() const override;
    float gamma = 0.f, kT = 0.f, delt = -1.0f;
};
class Langevin : public PS_Potential {
public:
    Langevin(std::istringstream& iss, PS_Box* box); // ps_potentialLangevin.cu:16-57
    mft_tild_potential descriptor() const override;
    float drag = 1.0f, delt = -5.0f;
};
class BiasField : public PS_Potential {
public:
    BiasField(std::istringstream& iss, PS_Box* box); // ps_potentialBias.cu:17-66
    mft_tild_potential descriptor() const override;
    std::string phase;
    int dir = 0, n_periods = 1;
};
class NBCharge : public PS_Potential {
public:
    NBCharge(std::istringstream& iss, PS_Box* box); // ps_potentialCharges.cu:14-23
    mft_tild_potential descriptor() const override;
    float LB = 0.f;
};

PS_Potential* PSPotentialFactory(std::istringstream& iss, PS_Box* box);

class Integrator {
public:
    Integrator(std::istringstream& iss, PS_Box* box); // ps_integrator.cu:18-39
    virtual ~Integrator() {}
    virtual void Integrate_1() {}
    virtual void Integrate_2() {}       // in the engine's fused step
    virtual void finishInitialization() {}
    virtual int engine_kind() const = 0;
    std::string groupName, name;
    float delt = 0.002f;

protected:
    PS_Box* mybox;
};
class GJF : public Integrator {
public:
    GJF(std::istringstream& iss, PS_Box* box) : Integrator(iss, box) {}
    int engine_kind() const override { return MFT_INT_GJF; }
};
class EM : public Integrator {
public:
    EM(std::istringstream& iss, PS_Box* box) : Integrator(iss, box) {}
    int engine_kind() const override { return MFT_INT_EM; }
};
// velocity Verlet (ps_integratorVV.cu): both half steps in the engine's step
class VV : public Integrator {
public:
    VV(std::istringstream& iss, PS_Box* box) : Integrator(iss, box) {}
    int engine_kind() const override { return MFT_INT_VV; }
};
Integrator* IntegratorFactory(std::istringstream& iss, PS_Box* box);

// PS_Compute (ps_compute.cu:8-43): "compute avg_sk group [wait W] [freq F]
// [file_name name]"; PS_ComputeSK (ps_compute_sk.cu) accumulates
// |rho_group(k)|^2 on the device every freq steps after wait and writes the
// average at log steps.
class PS_Compute {
public:
    PS_Compute(std::istringstream& iss, PS_Box* box);
    virtual ~PS_Compute() {}
    virtual void do_compute(int step) = 0;
    virtual void write_output() = 0;
    virtual void initialize_compute() = 0;
    bool active(int step) const { return step > compute_wait && compute_freq > 0 && step % compute_freq == 0; }
    int compute_id = 0, compute_wait = 0, compute_freq = 100, num_data_pts = 0;
    std::string group, output_name;
    static int total_computes;

protected:
    PS_Box* mybox;
};
class PS_ComputeSK : public PS_Compute {
public:
    PS_ComputeSK(std::istringstream& iss, PS_Box* box);
    ~PS_ComputeSK() override;
    void do_compute(int step) override;
    void write_output() override;
    void initialize_compute() override;

private:
    uint32_t mask = 0;
    float* d_sk = nullptr;
    void* d_work = nullptr;
};
PS_Compute* PSComputeFactory(std::istringstream& iss, PS_Box* box);

class PS_Box : public Box {
public:
    explicit PS_Box(std::istringstream& iss);
    ~PS_Box() override;
    void readInput(std::ifstream& inp) override;       // PS_Box_initialization.cu:32-345
    void NVT(int nsteps) override;                      // PS_Box.cu:210-231
    void doTimeStep(int step) override;                 // PS_Box.cu:73-207
    void writeData(int step) override;                  // PS_Box.cu:339-372
    void computeThermoProps();                          // PS_Box.cu:282-317
    void takeSteps(int n, bool energies);               // engine steps; keeps the log energies of the last
    std::vector<double> stepPe;                         // energies of the last step taken with takeSteps(.., true)
    double stepBe = 0.0;
    bool stepEnergiesValid = false;
    void writeDataConfig(const std::string& name);      // LAMMPS data, PS_Box.cu:629-720
    void writeLammpsTraj(int step);                     // PS_Box.cu:756-801
    void writeKFieldFloat(const std::string& name, const float* dat); // PS_Box.cu:531-560
    mft_tild* engine() const { return eng; }
    long gridSize() const { return M; }
    std::vector<PS_Compute*> computes;
    int trajFreq = 0;
    std::string trajFileName = "traj.lammpstrj";
    void writeGSDtraj();                                // PS_Box_gsdUtils.cu:14-160
    int gsdFreq = 0, logSpaceGSDflag = 0;
    float logScaleGSD = 1.f;
    int fieldFreq = 0, logSpaceFieldflag = 0;            // PS_Box_initialization.cu:171-182
    float logScaleFields = 1.f;
    void initFieldFiles();                              // binary density headers, PS_Box_initialization.cu:528-535
    void writeFields();                                 // one binary density frame per group, PS_Box.cu:376-388
    std::string gsd_name = "traj.gsd";
    std::unique_ptr<GsdWriter> gsd;
    int findSpeciesInteger(const std::string& s) const;
    int findGroupInteger(const std::string& s) const;
    uint32_t groupMask(const std::string& s) const;

    std::vector<PS_Species> species;
    std::vector<PS_Group> psGroup;
    std::vector<PS_Potential*> potentials;
    std::vector<Integrator*> integrators;
    int nstot = 0, MAXBONDS = 2, pmeorder = 1, doCharges = 0, nBondsTot = 0;
    double rho0 = -1.0;
    std::vector<float> bondK, bondReq;
    std::vector<int> bondStyle;
    // particles (reference layouts: x AoS [nstot*3], bonds [nstot*MAXBONDS])
    std::vector<float> x, charges;
    std::vector<int> intSpecies, mID, nBonds, bondedTo, bondType;
    // angles (PS_Box_initialization.cu:83-106, 229-231): per type k, theta_eq
    // (radians), style 0 WLC / 1 harmonic; per particle [nstot*MAXANGLES(*3)]
    int MAXANGLES = 3, nAnglesTot = 0;
    struct NeighborList {
        std::string group;
        float rcut = 0.f;
        int maxNeighbors = 64;
    };
    std::vector<NeighborList> neighborLists;
    std::vector<float> angleK, angleTheq;
    std::vector<int> angleStyle, nAngles, angleType, angleGroup;
    std::string datFileName = "ps_data.dat";
    int maxSteps = 0, totSteps = 0;
    float Upe = 0.f, Ubond = 0.f, Uangle = 0.f;

private:
    HostRng rng;
    mft_tild* eng = nullptr;
    bool initialized = false;
    void makeLinear(std::istringstream& iss);           // PS_Box_moleculeMakers.cu:15-311
    void readDataConfig(const std::string& name);       // PS_Box_initialization.cu:787-990
    void enableCharges();
    void grow(int n);
    void addAngle(int i1, int i2, int i3, int type);
    void finishInitialization();                        // PS_Box_initialization.cu:352-426
};

// ------------------------------------------------------------------ FTS ---
class FTS_Box;

class FTS_Species {
public:
    FTS_Species(std::istringstream& iss, FTS_Box* box);
    std::string fts_species;
};

class FTS_Potential {
public:
    FTS_Potential(std::istringstream& iss, FTS_Box* box);
    virtual ~FTS_Potential() {}
    // the engine updates every field in one fused pass per iteration
    virtual void updateFields() {}
    virtual void correctFields() {}
    virtual std::complex<double> calcHamiltonian() { return Hterm; }
    virtual void initLinearCoeffs() {}
    virtual void storePredictorData() {}
    virtual void writeFields(int potInd, mft_fts* sys); // wpl_<style><i>.dat [, wmi_Flory<i>.dat]
    virtual mft_fts_potential descriptor() const = 0;
    std::string input_command, potentialStyle, updateScheme = "EM";
    std::vector<std::string> actsOn;
    bool zeroMean = false;
    std::complex<double> Hterm{0.0, 0.0};
    std::vector<std::complex<double>> wpl, wmi; // initial fields (host)
    bool has_wmi = false;

protected:
    FTS_Box* mybox;
    void initializeField(std::istringstream& iss, std::vector<std::complex<double>>& w); // fts_potential.cu:62-150
    void parse_options(std::istringstream& iss, std::vector<std::complex<double>>* init_target, bool flory);
};
class PotentialHelfand : public FTS_Potential {
public:
    PotentialHelfand(std::istringstream& iss, FTS_Box* box); // fts_potential_helfand.cu:15-70
    mft_fts_potential descriptor() const override;
    double kappaN = 0.0, delt = 0.0;
};
class PotentialFlory : public FTS_Potential {
public:
    PotentialFlory(std::istringstream& iss, FTS_Box* box); // fts_potential_flory.cu:21-120
    mft_fts_potential descriptor() const override;
    double chiN = 0.0, deltPlus = 0.0, deltMinus = 0.0;
    int I = -1, J = -1;
};
class PotentialIncompress : public FTS_Potential {
public:
    PotentialIncompress(std::istringstream& iss, FTS_Box* box); // fts_potential_incompress.cu:15-60
    mft_fts_potential descriptor() const override;
    double delt = 0.0;
};
class PotentialEdwards : public FTS_Potential {
public:
    PotentialEdwards(std::istringstream& iss, FTS_Box* box); // fts_potential_edwards.cu:12-75
    mft_fts_potential descriptor() const override;
    double B = 0.0, delt = 0.0;
};
FTS_Potential* FTS_PotentialFactory(std::istringstream& iss, FTS_Box* box);

class FTS_Molec {
public:
    FTS_Molec(std::istringstream& iss, FTS_Box* box); // fts_molecule.cu:13-46
    virtual ~FTS_Molec() {}
    virtual void calcDensity() {}      // in the engine's fused iteration
    virtual std::complex<double> calcHTerm() { return Hterm; }
    virtual void computeLinearTerms() {}
    virtual void modifyMolecule(std::istringstream& iss);
    virtual void recomputeNmolecs() {}
    virtual mft_fts_molecule descriptor() const = 0;
    std::string input_command, molec_type;
    double phi = -1.0, nmolecs = -1.0, activity = -1.0, nSites = 0.0;
    std::complex<double> Q{1.0, 0.0}, Hterm{0.0, 0.0};

protected:
    FTS_Box* mybox;
};
class LinearMolec : public FTS_Molec {
public:
    LinearMolec(std::istringstream& iss, FTS_Box* box); // fts_molecule_linear.cu:17-140
    mft_fts_molecule descriptor() const override;
    std::vector<int> N;
    std::vector<std::string> blockSpecies;
    int Ntot = 0;
};
FTS_Molec* FTS_MolecFactory(std::istringstream& iss, FTS_Box* box);

class FTS_Box : public Box {
public:
    explicit FTS_Box(std::istringstream& iss);          // FTS_Box.cu:637-644
    ~FTS_Box() override;
    void readInput(std::ifstream& inp) override;        // FTS_Box.cu:308-470
    void NVT(int nsteps) override;                      // FTS_Box.cu:127-142
    void doTimeStep(int step) override;                 // FTS_Box.cu:19-124
    void writeData(int step) override;                  // FTS_Box.cu:145-190
    void computeHamiltonian();                          // FTS_Box.cu:590-602
    int converged(int step) override;                   // FTS_Box.cu:715-760
    int findSpeciesInteger(const std::string& s) const;
    void findSpinodal(std::istringstream& iss) override;  // algorithm of FTS_BoxFindSpinodal.cu:25-240
    void modifyBox(std::istringstream& iss) override;     // FTS_Box.cu:674-708
    void setActivity(const FTS_Molec* m, double activity);
    void updateNsites();

    std::string ftsStyle;
    double Nr = -1.0, C = -1.0, rho0 = -1.0;
    std::vector<FTS_Species> Species;
    std::vector<FTS_Potential*> Potentials;
    std::vector<FTS_Molec*> Molecs;
    std::complex<double> Heff{0.0, 0.0};
    double Hold = 1.0e8, tolerance = 1.0e-5, error = 0.0;
    // field files (FTS_Box.cu:84-113, 318-320, 373-377)
    int chemFieldFreq = 0, densityFieldFreq = 0, clSampleFreq = 100, nCLSamples = 0;
    std::vector<std::complex<double>> densityAvg;       // running sum of CL density samples [nspecies*M]
    void writeTComplexGridData(const std::string& name, const std::complex<double>* field) const; // FTS_Box.cu:284-306
    void writeFieldFiles(int step);
    bool ioStep(int step) const;
    std::string tolMetric = "Heff";
    int totSteps = 0;
    HostRng rng;
    long gridSize() const { return M; }

private:
    mft_fts* sys = nullptr;
    bool initialized = false;
    bool stepped = false;
    void finishInitialization();
};
