/*
 * mft.h -- C ABI of the MI355X engine for the MATILDA.FT hot paths.
 *
 * Plain C: device and host pointers, sizes and an opaque stream handle
 * (hipStream_t passed as void*, NULL = the null stream).  Every call returns an
 * mft_status; mft_last_error() describes the last failure on the calling
 * thread.  Nothing here exits the process (the reference's die(), die.cu:7-11,
 * becomes a status code that the drop-in layer maps back to the same message).
 *
 * Reference paths below are relative to src/ of MATILDA.FT.
 */
#ifndef MFT_H
#define MFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MFT_OK = 0,
    MFT_ERR_INVALID = 1,     /* bad argument or unsupported configuration */
    MFT_ERR_OOM = 2,         /* device allocation failed or over budget */
    MFT_ERR_NONFINITE = 3,   /* NaN/Inf detected (FTS_Box.cu:150-152 analogue) */
    MFT_ERR_DEVICE = 4,      /* HIP runtime error */
    MFT_ERR_UNSUPPORTED = 5  /* valid reference input this engine does not run */
} mft_status;

const char* mft_last_error(void);
int mft_version(void);

/* ------------------------------------------------------------ device ---- */
/* Thin helpers so that hosts without a GPU framework can move data. */
int mft_device_count(int* n);
/* Free and total device memory of the current device (hipMemGetInfo). */
int mft_mem_info(size_t* free_bytes, size_t* total_bytes);
int mft_malloc(void** dptr, size_t bytes);
int mft_free(void* dptr);
int mft_memcpy_h2d(void* dst, const void* src, size_t bytes);
int mft_memcpy_d2h(void* dst, const void* src, size_t bytes);
int mft_memcpy_d2d(void* dst, const void* src, size_t bytes);
int mft_memset(void* dst, int value, size_t bytes);
int mft_synchronize(void);
int mft_stream_create(void** stream);
int mft_stream_destroy(void* stream);
int mft_stream_synchronize(void* stream);

/* --------------------------------------------------------------- RNG ---- */
/* Replaces curand_init/curand_normal (device_utils.cu:465-472,
 * Box.cu:465-488, fts_make_CLnoise.cu:7-24): Philox4x32-10 blocks for
 * counters (i, c1, c2, c3), i = 0..nblocks-1, keyed by seed.  raw: 4 uint32 per
 * block; normal: 4 float normals per block (Box-Muller on 24-bit uniforms). */
int mft_philox_raw(uint32_t* out, int64_t nblocks, uint64_t seed, uint32_t c1, uint32_t c2,
                   uint32_t c3, void* stream);
int mft_philox_normal(float* out, int64_t nblocks, uint64_t seed, uint32_t c1, uint32_t c2,
                      uint32_t c3, void* stream);

/* --------------------------------------------------------------- FFT ---- */
/* Replaces cufftPlan3d/cufftExecC2C (PS_Box_initialization.cu:369-372,
 * Box.cu:107-127) and cufftExecZ2Z (FTS_Box.cu:466-469, Box.cu:76-104).
 * Complex interleaved float, x fastest.  dir = 1: forward scaled by 1/M (the
 * reference wrappers' convention); dir = -1: inverse, unscaled.  nz = 1 for 2D.
 * batch grids are consecutive.  Every axis length must be one of
 * 2,3,4,5,6,8,10,12,16,20,24,32,40,48,64,80,96,128,160,192,256. */
typedef struct mft_fft_plan mft_fft_plan;
int mft_fft_plan_create(mft_fft_plan** plan, int nx, int ny, int nz, int batch);
int mft_fft_exec(mft_fft_plan* plan, const void* in, void* out, int dir, void* stream);
int mft_fft_plan_destroy(mft_fft_plan* plan);
int mft_fft_supported(int n);

/* ============================================================== TILD ==== */
/* Per-kernel entry points (device pointers, reference layouts: positions
 * x[i*3+j], grids x-fastest, vector grid fields AoS F[g*3+j]).  Dim = 3. */

/* d_calcGridWeights (ps_grid_utils.cu:74-204): (p+1)^3 grid indices and
 * weights/gvol per particle; indices identical to the reference's. */
int mft_tild_grid_weights(float* gridW, int* gridInds, const float* x, int ns, const int Nx[3],
                          const float dx[3], int pmeorder, void* stream);
/* d_fillDensityGrid / d_fillChargeDensityGrid (device_utils.cu:282-352) from
 * the tables above for the listed sites (ns particles in total, M grid
 * points); rho (and rhoq when q != NULL) are accumulated into, like the
 * reference (zero them first).  Deterministic (integer accumulation). */
int mft_tild_fill_density(float* rho, float* rhoq, const float* q, const int* sites, int nsites,
                          int ns, const int* gridInds, const float* gridW, int G, int M,
                          void* stream);
/* d_mapGridForcesToPartics / d_mapGridChargeForcesToPartics
 * (device_utils.cu:358-461): f[p] += sum_g F[g*3+j] W gvol (x q_p if q). */
int mft_tild_map_forces(float* f, const float* gridF, const float* q, const int* sites,
                        int nsites, int ns, const int* gridInds, const float* gridW, int G,
                        float gvol, int M, void* stream);
/* d_bonds (ps_deviceBonds.cu:7-79).  The per-type tables bondReq, bondK,
 * bondStyle [nbondtypes] are host arrays; the rest are device arrays. */
int mft_tild_bonds(float* f, const int* nBonds, const int* bondedTo, const int* bondType,
                   const float* bondReq, const float* bondK, const int* bondStyle, int nbondtypes,
                   const float* x, const float L[3], int ns, int maxbonds, void* stream);

/* Potential kinds (ps_potential.cu:294-333). */
/* MFT_POT_LANGEVIN (Langevin, ps_potentialLangevin.cu): drag and random
 * forces on group maskI, f += -v A + sqrtf(2 A / sigma) N(0, 1) with A the
 * drag and sigma the time step of the noise; v is the velocity under
 * MFT_INT_VV and zero otherwise (the reference's never-updated velocities);
 * its energy is zero */
enum { MFT_POT_GAUSSIAN = 0, MFT_POT_ERF2 = 1, MFT_POT_ERFG = 2, MFT_POT_CHARGES = 3, MFT_POT_BIAS = 4,
       MFT_POT_WALL = 5, MFT_POT_LANGEVIN = 6, MFT_POT_DPD = 7 };
/* MFT_POT_DPD (DPD, ps_potentialDPD.cu, with the group's neighbor_list,
 * ps_neighborList.cu): pairwise dissipative and random forces between the
 * group's particles closer than Rp (the list's rcut), F = (-A w^2 (v_ij . r^)
 * + sqrtf(2 A sigma) w xi / sqrtf(wall_pos)) r^ with w = 1 - r/Rp, A = gamma,
 * sigma = kT, wall_pos = the noise's time step; xi is one normal per pair and
 * step; v as for MFT_POT_LANGEVIN; at most one; Rp at most a third of every box
 * length; energy zero */
/* Phases of the static bias field (Box::make_bias_field, Box.cu:308-458):
 * lamellar A sin(2 pi n x_dir / L_dir), hexagonal cylinders along dir
 * A cos(.) cos(.), gyroid A ((sin x cos y + sin y cos z + sin z cos x)^2 - 1.44).
 * The reference's BCC phase reads an uninitialised vector (Box.cu:359 passes
 * the position as pbc_dr2's output), so it is rejected. */
enum { MFT_BIAS_LAM = 0, MFT_BIAS_HEX = 1, MFT_BIAS_GYR = 2 };
/* Box::make_bias_field (Box.cu:308-458) on the host: u[M] = A times the
 * lamellar (sin along dir), cylinder (cos cos across dir) or gyroid shape
 * with n_periods periods, at r = n dx in float like the reference; the
 * ordered-phase initialize of an FTS field (fts_potential.cu:102-138) and
 * the TILD bias potential use it.  The BCC phase is not offered (the
 * reference evaluates it on an uninitialised vector, Box.cu:359). */
int mft_bias_field(double* u, const int Nx[3], const float L[3], int phase, double A, int dir, int n_periods);

/* initializePotential of one potential (ps_potentialGaussian.cu:28-71,
 * ps_potentialErf2.cu:31-67, ps_potentialErfG.cu:30-66,
 * ps_potentialCharges.cu:30-80): uk[M] = u(k) and fk[M*3] = -i k_j u(k),
 * complex interleaved, device arrays (either may be NULL).  The erf shapes
 * are evaluated in FP64 and rounded once.  Synchronises the stream. */
int mft_tild_kspace_table(float* uk, float* fk, int kind, float A, float sigma, float Rp, const int Nx[3],
                          const float L[3], void* stream);
/* One direction of PS_Potential::CalcForces (ps_potential.cu:41-120):
 * gridF[g*3+j] += factor * Re IFFT[fk_j FFT[rho]] with the reference's
 * forward scaling by 1/M (Box.cu:58-127); call it again with the roles of
 * the groups swapped, or with factor 2 when a group acts on itself.  work:
 * device scratch of 2*M complex floats (16*M bytes). */
int mft_tild_pair_force(float* gridF, const float* rho, const float* fk, const int Nx[3], float factor,
                        void* work, void* stream);
/* PS_Potential::CalcEnergy (ps_potential.cu:127-178): gvol * sum rhoI *
 * Re IFFT[uk FFT[rhoJ]], summed in FP64 (the reference sums in float).
 * work: device scratch of M complex floats + one double per 256 grid points.
 * Synchronises the stream; *energy is a host pointer. */
int mft_tild_potential_energy(double* energy, const float* rhoI, const float* rhoJ, const float* uk,
                              const int Nx[3], double gvol, void* work, void* stream);
/* GJF Integrate_2 on group "all" (ps_integratorGJF.cu:102-163) with the
 * coefficients of GJF::finishInitialization (ps_integratorGJF.cu:56-82, in
 * float): x, xo, noise_old [n*3] updated in place, f [n*3], type [n]; mass
 * and mobility are host arrays [nspecies] (NULL = 1).  Noise: Philox
 * normals keyed (particle, step, 'TILD', 0) by seed, the same stream the
 * engine draws. */
int mft_tild_gjf_step(float* x, float* xo, float* noise_old, const float* f, const int* type, int nparticles,
                      int nspecies, const float* mass, const float* mobility, float delt, const float L[3],
                      uint64_t seed, int64_t step, void* stream);
/* EM Integrate_2 (ps_integratorEM.cu:39-82): x += delt f + sqrt(2 delt) N(0,1),
 * wrapped into the box; noise as above. */
int mft_tild_em_step(float* x, const float* f, int nparticles, float delt, const float L[3], uint64_t seed,
                     int64_t step, void* stream);
/* Integrators (ps_integrator.cu:51-69). */
/* MFT_INT_VV: velocity Verlet (ps_integratorVV.cu); the velocities travel in
 * the noise_old slot of mft_tild_set_state / mft_tild_get_state, start at zero
 * and, between steps, hold the half kick of the last step */
enum { MFT_INT_GJF = 0, MFT_INT_EM = 1, MFT_INT_NONE = 2, MFT_INT_VV = 3 };

typedef struct {
    int kind;          /* MFT_POT_* */
    uint32_t maskI;    /* species bitmask of group I (ignored for charges) */
    uint32_t maskJ;    /* species bitmask of group J */
    float A;           /* amplitude; l_B for charges */
    float sigma;       /* Gaussian width (ps_potentialGaussian.cu:21-23) */
    float Rp;          /* erf radius (ps_potentialErf2.cu:91-92) */
    int ramp;          /* PS_Potential::ramp_check_input (ps_potential.cu:253-283) */
    float A_final;
    /* MFT_POT_BIAS (BiasField, ps_potentialBias.cu:17-200): a static external
     * field A u(r) on the species of maskI (maskJ unused); its force -grad u
     * is taken spectrally once, like the reference, and its energy is
     * gvol sum rho_I A u.  No ramp (the reference's ramp scaling never reaches
     * the precomputed bias forces). */
    int phase;         /* MFT_BIAS_* */
    int dir;           /* 0..2 */
    int n_periods;     /* >= 1 */
    /* MFT_POT_WALL (Wall, ps_potentialWall.cu:10-147): a harmonic wall on
     * axis dir for the species of maskI: f_dir += A (wall_pos - x_dir) and
     * energy A/2 (wall_pos - x_dir)^2 where wall_sign (wall_pos - x_dir) > 0. */
    float wall_pos;
    int wall_sign;     /* +1 or -1 */
} mft_tild_potential;

/* Everything PS_Box::readInput/finishInitialization (PS_Box_initialization.cu:
 * 32-569) derives for the hot loop.  Host pointers; copied at create. */
typedef struct {
    int nx, ny, nz;
    float lx, ly, lz;
    int pmeorder;            /* 0..4 (p = 4 reproduces ps_grid_utils.cu:57) */
    int nparticles;
    int nspecies;            /* <= 16 */
    int maxbonds;            /* MAXBONDS (PS_Box_initialization.cu:40) */
    int nbondtypes;
    const float* bond_k;     /* [nbondtypes] */
    const float* bond_r0;
    const int* bond_style;   /* 0 harmonic, 1 FENE */
    int npotentials;         /* <= 32 */
    const mft_tild_potential* potentials;
    int integrator;          /* MFT_INT_* on group "all" */
    float delt;
    const float* mass;       /* [nspecies] (ps_species.cu:13-38) */
    const float* mobility;
    uint64_t seed;
    int has_charges;
    /* angles (ps_deviceAngles.cu, PS_Box_initialization.cu:83-106): up to
     * maxangles per particle (0 = none), per type k, theta_eq (radians) and
     * style 0 = WLC k (1 + cos theta), 1 = harmonic k (theta - theta_eq)^2 */
    int maxangles;
    int nangletypes;
    const float* angle_k;
    const float* angle_theta;
    const int* angle_style;
} mft_tild_desc;

typedef struct mft_tild mft_tild;

/* Engine with device-resident state.  create copies the description; the
 * state arrays below are host pointers copied in (set_*) or out (get_*).
 * An engine holds at most 2^27 - 1 particles (MFT_ERR_UNSUPPORTED beyond). */
int mft_tild_create(mft_tild** eng, const mft_tild_desc* desc);
int mft_tild_destroy(mft_tild* eng);
void* mft_tild_stream(mft_tild* eng);
/* Topology (PS_Box::makeLinear layout, PS_Box_moleculeMakers.cu:15-311):
 * type[N], charge[N] (may be NULL), nBonds[N], bondedTo/bondType[N*maxbonds]. */
int mft_tild_set_topology(mft_tild* eng, const int* type, const float* charge,
                          const int* nBonds, const int* bondedTo, const int* bondType);
/* Angles (PS_Box_moleculeMakers.cu:242-281 layout): nAngles[N], angleType[N*maxangles],
 * angleGroup[N*maxangles*3] (particle ids i, j, k of each angle the particle
 * is in); host arrays, before mft_tild_set_state. */
int mft_tild_set_angles(mft_tild* eng, const int* nAngles, const int* angleType, const int* angleGroup);
/* Angle energy at the current state (d_anglesStressEnergy / 3,
 * PS_Box.cu:296-305), FP64. */
int mft_tild_angle_energy(mft_tild* eng, double* energy);
/* Positions, previous positions and previous noise (GJF state,
 * ps_integratorGJF.cu:43-51); xo/noise_old may be NULL (x_old = x, 0). */
int mft_tild_set_state(mft_tild* eng, const float* x, const float* xo, const float* noise_old);
int mft_tild_get_state(mft_tild* eng, float* x, float* xo, float* noise_old);
/* Forces at the current positions (PS_Box.cu:96-128 + forces() :240-278)
 * into f[3N] (host). */
int mft_tild_compute_forces(mft_tild* eng, float* f);
/* Species densities rho[nspecies*M] and charge density rhoq[M] (may be NULL)
 * from the last force evaluation (group densities are sums over species). */
int mft_tild_get_densities(mft_tild* eng, float* rho, float* rhoq);
/* Total grid force field of species s, AoS [M*3] (the group's d_gridForce,
 * ps_groups.cu:164-178); s = nspecies gives the charge field before x q. */
int mft_tild_get_grid_force(mft_tild* eng, int s, float* F);
/* nsteps of PS_Box::doTimeStep (PS_Box.cu:73-207) without I/O; step counters
 * advance the RNG and potential ramps. Asynchronous on the engine stream. */
int mft_tild_step(mft_tild* eng, int nsteps);
/* Declares a run of max_steps steps (PS_Box::NVT's maxSteps, PS_Box.cu:210-231):
 * potential ramps (PS_Potential::update_prefactor, ps_potential.cu:227-248)
 * then follow the step index within the run across any number of
 * mft_tild_step calls.  Without it each mft_tild_step call is a run. */
int mft_tild_set_run(mft_tild* eng, int64_t max_steps);
int mft_tild_synchronize(mft_tild* eng);
/* Energies of computeThermoProps (PS_Box.cu:282-317): per-potential
 * energies [npotentials] (charges include the 1/2, ps_potentialCharges.cu:115)
 * and bond energy (x 1/2, PS_Box.cu:292, with the reference's 0.5*Lh minimum
 * image of ps_deviceBonds.cu:132-133). */
int mft_tild_energies(mft_tild* eng, double* pot_energy, double* bond_energy);
/* mft_tild_energies into device memory, without synchronising: d_out
 * [npotentials + 1] receives, on the engine stream, the pair-potential
 * energies (FP64; zero for bias, wall, DPD and Langevin entries) and, last,
 * the bond energy (PS_Box::computeThermoProps, PS_Box.cu:282-292).  The
 * caller reads it back when it needs it (mft_tild_stream orders the copy). */
int mft_tild_energies_device(mft_tild* eng, double* d_out);
/* Virial part of d_bondStressEnergy (ps_deviceBonds.cu:85-171) at the current
 * state: vir[6N] (host, particle-id order, the layout of the reference's
 * d_bondVirScratch), per particle the terms -mf dr_a dr_b over its bonds as
 * xx, yy, zz, xy, xz, yz, with the reference's 0.5*Lh minimum image. */
int mft_tild_bond_virial(mft_tild* eng, float* vir);
/* nsteps steps (nsteps >= 1) that also return the energies PS_Box::writeData
 * logs after the last of them (PS_Box.cu:152-173, 282-317): pair potentials
 * from that step's own density pass (the positions before the integrator
 * moved them) with the amplitudes after its ramp update (update_prefactor
 * runs before writeData); bias and wall energies, and the bond energy (half
 * the per-particle sum, 0.5 Lh minimum image), at the positions after the
 * step.  Either output may be NULL. */
int mft_tild_step_energies(mft_tild* eng, int nsteps, double* pot_energy, double* bond_energy);
/* Per-kernel timing: while enabled, every kernel group of step() is
 * bracketed by HIP events on the engine stream (no host synchronisation);
 * each call resolves the recorded pairs and returns the average ms per launch
 * and the launch counts since timing was last switched on. */
enum {
    MFT_TK_BIN = 0, MFT_TK_SCAN, MFT_TK_SCATTER, MFT_TK_DENSITY, MFT_TK_FFT_X_FWD,
    MFT_TK_FFT_Y_FWD, MFT_TK_CONV_Z, MFT_TK_FFT_Y_INV, MFT_TK_FFT_X_INV, MFT_TK_GATHER,
    MFT_TK_INTEGRATE, MFT_TK_COUNT
};
int mft_tild_timing(mft_tild* eng, int enable, double* avg_ms, int64_t* launches);
/* Which kernel groups timing brackets: bit k of mask = MFT_TK k (default all).
 * Each bracket costs a few microseconds of stream time, so a run that times
 * one kernel inside a throughput measurement selects only that kernel. */
int mft_tild_timing_select(mft_tild* eng, uint32_t mask);
/* Algorithmic HBM bytes per launch of each kernel id at this configuration. */
int mft_tild_kernel_bytes(mft_tild* eng, double* bytes);
int mft_tild_num_outputs(mft_tild* eng, int* nout);
/* Times the kernel variants of this configuration on the device (reps
 * launches each, state unchanged) and keeps the fastest; choice receives the
 * selected variants: bits 0-1 the gather (0 direct loads, 1 LDS windows per
 * brick and species, 2 LDS windows of every species per brick, 3 as 1 with
 * the next window's loads overlapping the current sums; the direct and LDS
 * variants can differ in the last bit of a force), bit 2 the x-y
 * transform passes (0 one pass per axis, 1 one pass per z-plane; equal to
 * rounding), bit 3 the forward x pass (0 tile entries read directly, 1 tile
 * rows staged in LDS first; the same bits). */
int mft_tild_autotune(mft_tild* eng, int reps, int* choice);
/* The last autotune's timings, ms[n] (up to 8; -1 where not timed): gather
 * variants 0-3 (ms per launch), then x-y passes per axis / per plane and the
 * forward x pass direct / staged (ms per grid-force evaluation). */
int mft_tild_autotune_timings(mft_tild* eng, double* ms, int n);
/* Selects a gather variant directly (as above; the default is 1). */
int mft_tild_set_gather(mft_tild* eng, int variant);
/* Selects the x-y transform passes directly: 0 one pass per axis (default),
 * 1 both axes of a z-plane in one workgroup (square 64, 96 or 128 point
 * planes), 2 one pass per axis with the forward x pass staging its tile rows
 * in LDS; MFT_ERR_UNSUPPORTED where the grid does not allow the variant. */
int mft_tild_set_grid_passes(mft_tild* eng, int variant);
/* Replica groups (no counterpart in the reference, which steps one box):
 * independent TILD engines advanced together by one HIP graph holding eight
 * steps of every engine (each engine's kernels on its own stream, so the
 * replicas overlap on the device; remainders take ordinary steps).  Each engine gives the same bits as when
 * stepped alone.  Engines with potential ramps, DPD, Langevin potentials,
 * velocity Verlet or timing enabled, and a GJF engine's first step after its
 * state is set, take ordinary steps.  The engines stay owned by the caller;
 * destroy the group before them. */
typedef struct mft_tild_group mft_tild_group;
int mft_tild_group_create(mft_tild_group** grp, mft_tild* const* engines, int n);
int mft_tild_group_destroy(mft_tild_group* grp);
int mft_tild_group_step(mft_tild_group* grp, int nsteps);
int mft_tild_group_synchronize(mft_tild_group* grp);
/* PS_ComputeSK::do_compute (ps_compute_sk.cu:18-45): sk[M] (device) = or +=
 * |FFT[rho_group]|^2 with the forward transform scaled by 1/M, rho_group the
 * sum of the densities of the species in species_mask at the last density
 * pass (the start of the last step).  work: device scratch of 8*M bytes. */
int mft_tild_structure_factor(mft_tild* eng, uint32_t species_mask, float* sk, int accumulate, void* work);

/* =============================================================== FTS ==== */
/* Per-kernel entry points (complex interleaved float fields [M]). */

/* LinearMolec bond factor exp(-k^2/(Nr-1)) (fts_molecule_linear.cu:109-122),
 * as complex values like the reference's d_bond_fft. */
int mft_fts_bond_factor(float* g2, const int Nx[3], const float L[3], double Nr, void* stream);

/* Potential styles of FTS_PotentialFactory (fts_potential.cu:25-55).  Edwards
 * (fts_potential_edwards.cu) is Helfand's excluded volume with force
 * w/B + i rho/Nr (no C factor, no -iC term) and H = gvol sum w^2 / (2B). */
enum { MFT_FTS_HELFAND = 0, MFT_FTS_FLORY = 1, MFT_FTS_INCOMPRESS = 2, MFT_FTS_EDWARDS = 3 };
/* updateScheme (fts_potential_flory.cu:90-109): Euler-Maruyama; the
 * predictor-corrector of fts_updateSchemes.cu:56-81 (a box with any EMPC
 * potential recomputes its densities twice per iteration, FTS_Box.cu:53-81);
 * or the semi-implicit k-space step of fts_updateSchemes.cu:34-52 with the
 * Debye linear coefficients of fts_molecule_linear.cu:433-560 (like the
 * reference, it adds no noise). */
enum { MFT_FTS_EM = 0, MFT_FTS_EMPC = 1, MFT_FTS_1S = 2 };

typedef struct {
    int kind;             /* MFT_FTS_* (fts_potential.cu:25-55) */
    int I, J;             /* species indices (Flory) */
    double chiN;          /* Flory chi N (fts_potential_flory.cu:49) */
    double kappaN;        /* Helfand kappa N (fts_potential_helfand.cu:26); Edwards B
                             (fts_potential_edwards.cu:18) */
    double delt_pl;       /* time step of w+ (delt for Helfand/Incompress) */
    double delt_mi;       /* time step of w- (Flory) */
    int scheme;           /* MFT_FTS_EM (0, the reference's default), _EMPC or _1S */
    int zero_mean;        /* modify zeroMean: subtract the grid mean of w+ after every
                             update and correction (fts_potential_helfand.cu:140-146,
                             fts_potential_incompress.cu:125-131, fts_potential_edwards.cu:
                             146-152); Flory parses it without effect (fts_potential_flory.cu:94-98) */
} mft_fts_potential;

typedef struct {
    int nblocks;          /* <= 16 (fts_molecule_linear.cu:26-47) */
    int N[16];            /* beads per block */
    int species[16];      /* species index per block */
    double phi;           /* volume fraction (canonical); nmolecs derived as in
                             fts_molecule_linear.cu:433-435 */
    double activity;      /* > 0: grand canonical (fts_molecule.cu:30-31): density
                             factor = activity, H term -activity V Q, nSites =
                             activity V Re Q Ntot (fts_molecule_linear.cu:278-382);
                             phi is then ignored */
} mft_fts_molecule;

typedef struct {
    int nx, ny, nz;
    float lx, ly, lz;     /* box lengths in Rg */
    double Nr;            /* reference chain length (no default in the reference) */
    double C;             /* dimensionless concentration (FTS_Box.cu:486-497) */
    int nspecies;         /* <= 8 */
    int npotentials;      /* <= 8 */
    mft_fts_potential potentials[8];
    int nmolecules;       /* <= 4 */
    mft_fts_molecule molecules[4];
    int cl;               /* 1: complex Langevin (noise), 0: SCFT (no noise) */
    uint64_t seed;        /* noise key; give every replica/request its own seed */
} mft_fts_desc;

typedef struct mft_fts mft_fts;

/* One CL-FTS system (a batch of one).  Fields are complex float host arrays
 * of M values: wpl/wmi per potential (wmi only for Flory). */
int mft_fts_create(mft_fts** sys, const mft_fts_desc* desc);
int mft_fts_destroy(mft_fts* sys);
void* mft_fts_stream(mft_fts* sys);
int mft_fts_set_fields(mft_fts* sys, int pot, const float* wpl, const float* wmi);
/* LinearMolec::modifyMolecule "activity" (fts_molecule_linear.cu:386-412): the
 * new activity of molecule m (grand canonical only) enters the observables at
 * once and the densities from the next density pass on, so the next field
 * update still uses the densities of the old activity, as in the reference. */
int mft_fts_set_activity(mft_fts* sys, int m, double activity);
/* Sites of every molecule: activity V Re Q Ntot (grand canonical,
 * fts_molecule_linear.cu:344-346) or nmolecs Ntot (canonical). */
int mft_fts_nsites(mft_fts* sys, double* nsites);
int mft_fts_get_fields(mft_fts* sys, int pot, float* wpl, float* wmi);
/* Snapshot slot 0..7 of the state a sweep may return to, kept in device
 * memory: restore = 0 copies every potential field and the activities into
 * the slot, restore = 1 copies them back and recomputes the densities of the
 * restored fields (the reference keeps such copies on the host,
 * FTS_Box::storePotentials / restorePotentials, FTS_BoxFindSpinodal.cu:289-371). */
int mft_fts_snapshot(mft_fts* sys, int slot, int restore);
/* FTS_Box::finishInitialization density pass (FTS_Box.cu:193-256). */
int mft_fts_init(mft_fts* sys);
/* niter CL iterations of FTS_Box::doTimeStep (FTS_Box.cu:19-47). */
int mft_fts_step(mft_fts* sys, int niter);
int mft_fts_synchronize(mft_fts* sys);
/* Species densities rho[nspecies*M] (complex) after the last iteration. */
int mft_fts_get_densities(mft_fts* sys, float* rho);
/* Species fields W[nspecies*M] (complex, fts_species.cu:53-132). */
int mft_fts_get_species_fields(mft_fts* sys, float* W);
/* Propagators of molecule m: q[Ntot*M], qdag[Ntot*M] in the reference's
 * q-form (fts_molecule_linear.cu:144-269), the form the engine stores;
 * either may be NULL. */
int mft_fts_get_propagators(mft_fts* sys, int m, float* q, float* qdag);
/* Q[m] (complex), H (complex) and its parts: potentials then molecules
 * (FTS_Box::computeHamiltonian, FTS_Box.cu:590-602), FP64 reductions. */
int mft_fts_observables(mft_fts* sys, double* Q, double* H, double* parts);
/* Chemical potential of every molecule, mu[2*nmolecules] (complex): ln(n/(V Q))
 * for a canonical molecule, ln(activity) for a grand canonical one.  The
 * reference prints none; this is the ideal-chain form its density factor
 * n/(Q V) implies (fts_molecule_linear.cu:286-291, H term :368-382). */
int mft_fts_chemical_potentials(mft_fts* sys, double* mu);

/* Per-kernel FTS entry points on caller-owned device arrays in the
 * reference's q-form layouts (complex interleaved float [M] per field,
 * x-fastest).  The engine above fuses the same arithmetic; these follow the
 * reference's data flow one step at a time. */

/* FTS_Species::buildPotentialField (fts_species.cu:53-132) for every
 * species: W[s*M+i] = (sum over potentials of i w+ and -/+ w- of Flory
 * pairs on I/J) / Nr, accumulated in FP64.  wpl[p], wmi[p]: device fields
 * of potential p (wmi may be NULL when no potential is Flory). */
int mft_fts_species_fields(float* W, int nspecies, int npotentials, const mft_fts_potential* pots,
                           const float* const* wpl, const float* const* wmi, double Nr, long M, void* stream);
/* One potential's updateFields with Euler-Maruyama (fts_potential_helfand.cu:77-152,
 * fts_potential_flory.cu:123-167, fts_potential_incompress.cu:84-113; force
 * d_make*Force, step d_fts_updateEM fts_updateSchemes.cu:8-29) at the
 * species densities rho[s] (device [M] complex each).  cl != 0 adds the real
 * noise sqrt(2 delt / gvol) N(0,1) (fts_make_CLnoise.cu:7-24), the normal in
 * column noise_col (w+) and noise_col+1 (Flory w-) of the 8 drawn per grid
 * point and iteration: counters (i, iter, 'FTSN', 0) and (i, iter, 'FTSN', 1)
 * by seed, the engine's columns for potentials in declaration order. */
int mft_fts_update_fields(float* wpl, float* wmi, const mft_fts_potential* pot, const float* const* rho,
                          int nspecies, double C, double Nr, const int Nx[3], const float L[3], int cl,
                          uint64_t seed, int64_t iter, int noise_col, void* stream);
/* LinearMolec::calcPropagators (fts_molecule_linear.cu:144-269): q[Ntot*M]
 * and qdag[Ntot*M] (either may be NULL) from the species fields W and the
 * bond factor g (mft_fts_bond_factor), q_s = e^{-W} IFFT[g FFT[q_{s-1}]]. */
int mft_fts_propagate(float* q, float* qdag, const float* W, const float* g, const mft_fts_molecule* mol,
                      const int Nx[3], void* stream);
/* LinearMolec::calcDensity (fts_molecule_linear.cu:278-363): adds
 * factor e^{+W_b} sum_s q_s qdag_{Ntot-1-s} of every block b to rho of its
 * species (rho [nspecies*M]); factor = nmolecs / (Q V), complex. */
int mft_fts_density(float* rho, const float* q, const float* qdag, const float* W, const mft_fts_molecule* mol,
                    double factor_re, double factor_im, long M, void* stream);
/* Q = gvol / V * sum q_{Ntot-1} (fts_molecule_linear.cu:265-266), FP64 sum.
 * Q: host double[2]; work: device scratch of 32 bytes per 256 grid points.
 * Synchronises the stream. */
int mft_fts_partition_function(double* Q, const float* q_last, double gvol, double V, long M, void* work,
                               void* stream);
/* Allocation-free variant of mft_fts_partition_function, with no counterpart
 * in the reference: for callers that cannot allocate device memory (code
 * around HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result, bit for bit, without the work argument: the FP64
 * partial sums live in a library-owned static device array (2 MiB per
 * device, held until the process ends), so M is limited to 2^24 points
 * (256^3); larger grids return MFT_ERR_INVALID.  Calls no hipMalloc/hipFree.
 * Calls that overlap on one device share that array and race.  Synchronises
 * the stream. */
int mft_fts_partition_function_noalloc(double* Q, const float* q_last, double gvol, double V, long M,
                                       void* stream);
/* One potential's Hamiltonian term (fts_potential_flory.cu:252-274,
 * fts_potential_helfand.cu:254-275, fts_potential_incompress.cu:232-247),
 * FP64 sums.  H: host double[2]; work as above.  Synchronises the stream. */
int mft_fts_hamiltonian_term(double* H, const mft_fts_potential* pot, const float* wpl, const float* wmi, double C,
                             double gvol, long M, void* work, void* stream);

/* Batched CL-FTS: many independent replicas of uneven sizes advanced by the
 * same launches (replicas with equal grids share every launch).  Replica r
 * uses descs[r]; its noise depends only on descs[r].seed and its own
 * iteration count, so a replica gives the same result alone or in a batch. */
typedef struct mft_fts_batch mft_fts_batch;
int mft_fts_batch_create(mft_fts_batch** b, const mft_fts_desc* descs, int nrep);
int mft_fts_batch_destroy(mft_fts_batch* b);
int mft_fts_batch_set_fields(mft_fts_batch* b, int rep, int pot, const float* wpl,
                             const float* wmi);
int mft_fts_batch_init(mft_fts_batch* b);
int mft_fts_batch_step(mft_fts_batch* b, int niter);
int mft_fts_batch_synchronize(mft_fts_batch* b);
int mft_fts_batch_observables(mft_fts_batch* b, int rep, double* Q, double* H, double* parts);
int mft_fts_batch_get_densities(mft_fts_batch* b, int rep, float* rho);
/* Fields, species fields and chemical potentials of replica rep, as
 * mft_fts_get_fields, mft_fts_get_species_fields, mft_fts_chemical_potentials. */
int mft_fts_batch_get_fields(mft_fts_batch* b, int rep, int pot, float* wpl, float* wmi);
/* Propagation in chunks of replicas (no counterpart in the reference): each
 * group of equal grids runs its replicas' propagators chunk by chunk, all
 * beads of one chunk before the next, so a chunk's slices are read back while
 * they are still cached.  bytes bounds a chunk's per-bead working set (four
 * slices per chain); 0 keeps one chunk per group.  Results are the same bits
 * whatever the chunking. */
int mft_fts_batch_set_chunk_bytes(mft_fts_batch* b, double bytes);
/* Times each group's density pass (propagators, densities, Q from the current
 * fields, which it leaves unchanged) with one chunk and with chunks of 512,
 * 256, 128 and 64 MB, reps passes each, and keeps the fastest per group;
 * ms_per_pass (may be NULL) receives the chosen passes' summed time.  Call
 * after mft_fts_batch_init. */
int mft_fts_batch_autotune(mft_fts_batch* b, int reps, double* ms_per_pass);
int mft_fts_batch_get_species_fields(mft_fts_batch* b, int rep, float* W);
int mft_fts_batch_chemical_potentials(mft_fts_batch* b, int rep, double* mu);

/* ============================================================ serving ==== */
/* Request loop with continuous batching (no counterpart in the reference,
 * whose sweeps re-solve one system at a time, FTS_BoxFindSpinodal.cu:25-371):
 * CL-FTS solves and TILD runs are submitted at any time, admitted while the
 * device-memory estimate of the admitted work stays within hbm_budget_bytes,
 * and advanced together: every admitted CL-FTS replica joins the one
 * continuously batched pool (replicas of one grid share each launch), every
 * TILD run has its own engine stream, so their work overlaps on the device.
 * A request's result does not depend on what it was batched with. */
typedef struct mft_server mft_server;
enum { MFT_REQ_QUEUED = 0, MFT_REQ_RUNNING = 1, MFT_REQ_DONE = 2, MFT_REQ_FAILED = 3 };
int mft_server_create(mft_server** srv, double hbm_budget_bytes, int max_fts_replicas);
int mft_server_destroy(mft_server* srv);
/* niter iterations from the given initial fields: wpl/wmi are arrays of
 * desc->npotentials host pointers to M complex values (NULL = zero field). */
int mft_server_submit_fts(mft_server* srv, const mft_fts_desc* desc, const float* const* wpl,
                          const float* const* wmi, int niter, int64_t* id);
/* nsteps of a TILD system; topology and positions as for mft_tild_set_topology
 * and mft_tild_set_state (all copied at submission). */
int mft_server_submit_tild(mft_server* srv, const mft_tild_desc* desc, const int* type, const float* charge,
                           const int* nBonds, const int* bondedTo, const int* bondType, const float* x,
                           int nsteps, int64_t* id);
/* One pass of the loop: admit queued requests that fit, run up to max_iters
 * pool iterations (fewer when an admitted solve needs fewer), retire finished
 * work.  Returns the number of running and queued requests. */
int mft_server_pump(mft_server* srv, int max_iters, int* nrunning, int* nqueued);
int mft_server_status(mft_server* srv, int64_t id, int* state);
/* Message of a failed request (empty otherwise). */
const char* mft_server_error(mft_server* srv, int64_t id);
/* Results of a finished CL-FTS request: Q[2*nmolecules], H[2],
 * parts[2*(npotentials+nmolecules)] as mft_fts_observables; rho (may be NULL)
 * receives the species densities [nspecies*M] complex. */
int mft_server_fts_result(mft_server* srv, int64_t id, double* Q, double* H, double* parts, float* rho);
/* Final state of a finished CL-FTS request: the fields of potential pot
 * (wpl, and wmi for Flory; either may be NULL) and the chemical potentials
 * mu[2*nmolecules] (may be NULL), as mft_fts_get_fields and
 * mft_fts_chemical_potentials; with the results above they let a caller
 * re-evaluate the final densities and H from the fields. */
int mft_server_fts_final_state(mft_server* srv, int64_t id, int pot, float* wpl, float* wmi, double* mu);
/* Results of a finished TILD request: positions x[3N] (may be NULL) and the
 * energies of mft_tild_energies at the final state. */
int mft_server_tild_result(mft_server* srv, int64_t id, float* x, double* pot_energy, double* bond_energy);
int mft_server_release(mft_server* srv, int64_t id);

#ifdef __cplusplus
}
#endif
#endif
