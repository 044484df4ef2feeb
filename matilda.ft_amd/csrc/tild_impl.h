// Shared state and launchers of the TILD particle-mesh engine.
//
// One step (PS_Box::doTimeStep, PS_Box.cu:73-207) runs as
//   particles live in key order ((brick, species, sub-cell) counting sort);
//   -> tile density: one workgroup per brick of grid points accumulates the
//      (B+p)^3 stencil footprint of the brick's own particles in LDS (integer
//      fixed point, order independent), writes the tile, and a reduction sums
//      the overlapping tiles per grid point in a fixed order
//   -> FFT x, FFT y (all density channels in one launch each)
//   -> fused z pass: forward z FFT, every potential applied in k-space as one
//      per-species kernel sum_c K_tc(k) rho_c(k) times -i k_j, inverse z FFT
//   -> inverse FFT y, inverse FFT x keeping the real part (float4 fields)
//   -> gather per (brick, species): the species' (B+p)^3 force window staged
//      in LDS, each particle sums its (p+1)^3 stencil from it
//   -> bonds + GJF/EM per particle in key order, counting the next positions'
//      keys (wave-aggregated atomics) -> scan -> permutation of the particle
//      state into the next key order (nearly the identity step to step) and
//      remapping of the bond partner slots.
// The reference instead materialises (p+1)^3 index/weight tables per particle,
// scatters with FP32 atomics per group, and runs 4 or 8 FFTs per potential.
#pragma once
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "fft_core.h"
#include "spline.h"

namespace mft {

const float2* fft_twiddles_device(int n); // fft.hip

constexpr int kMaxSpecies = 16;

// A slot's particle id (bits 0-26) and species (bits 27-31) in one word: the
// permutation and the integrator move and read 4 bytes of identity per slot.
constexpr int kMetaIdBits = 27;
struct PMeta {
    uint32_t v;
    __host__ __device__ int id() const { return (int)(v & ((1u << kMetaIdBits) - 1u)); }
    __host__ __device__ int sp() const { return (int)(v >> kMetaIdBits); }
};
__host__ __device__ inline PMeta make_meta(int id, int species)
{
    return PMeta{(uint32_t)id | ((uint32_t)species << kMetaIdBits)};
}
static_assert(kMaxSpecies <= (1 << (32 - kMetaIdBits)), "species field of PMeta");
constexpr int kMaxPots = 32;
constexpr int kMaxBondTypes = 16;
constexpr int kMaxAngleTypes = 8;
constexpr int kMaxAngles = 8;

struct AngleParams {
    int ntypes, maxangles;
    float k[kMaxAngleTypes], theta[kMaxAngleTypes];
    int style[kMaxAngleTypes];
};
constexpr int kMaxChan = 8;

struct KPot {
    int kind;
    uint32_t maskI, maskJ;
    float A;      // current amplitude (ramped)
    float eA;     // amplitude of the energy sum (A, or the amplitude after this step's ramp update)
    float sig2;
    float Rp;
    const float* table; // u(k)/A for erf kinds, else null
};

// Potentials sharing one k-space shape: Z_f(k) = sum_g shape_g(k) sum_c
// coef_g[f][c] rho_c(k) for force slot f and density channel c.  Shapes:
// Gaussian exp(-k^2 sig2 / 2), folded charges exp(-k^2 sig2) / k^2, erf kinds
// a table of u(k)/A.
constexpr int kMaxGroups = 8;
struct KGroup {
    int kind;
    float sig2;
    const float* table;
    float coef[kMaxChan][kMaxChan];
};

struct KParams {
    int npots;
    int ngroups;        // -1: more shapes than kMaxGroups, evaluate per potential
    KGroup grp[kMaxGroups];
    KPot pot[kMaxPots];
    int charge_pot;     // index of the charges potential or -1
    int nchan;          // species channels (+1 for charge density)
    int qchan;          // channel index of the charge density or -1
    int nforce;         // species receiving grid forces
    int force_species[kMaxSpecies];
    int fold;           // 1: every species has one charge; rho_q = sum_s q_s rho_s
    float qspec[kMaxSpecies];
    int energy;         // accumulate potential energies
};

static_assert(sizeof(KParams) <= 3800, "KParams is passed as a kernel argument");

struct BondParams {
    int ntypes;
    float k[kMaxBondTypes];
    float r0[kMaxBondTypes];
    int style[kMaxBondTypes];
};

// harmonic walls (Wall, ps_potentialWall.cu): per-particle forces added in
// the integrate pass
constexpr int kMaxWalls = 4;
struct WallParams {
    int n;
    uint32_t mask[kMaxWalls];
    int axis[kMaxWalls], sign[kMaxWalls];
    float pos[kMaxWalls], k[kMaxWalls];
};

// Langevin potentials (ps_potentialLangevin.cu:60-133), up to kMaxWalls
struct LangevinParams {
    int n;
    uint32_t mask[kMaxWalls];
    float drag[kMaxWalls], nmag[kMaxWalls];
};

// DPD pair forces through a cell search of width >= rcut (ps_potentialDPD.cu,
// ps_neighborList.cu)
struct DpdParams {
    int on;
    uint32_t mask;
    float gamma, sigma, rcut, rcut2, inv_sqrt_delt;
    int nc[3];
    float cw[3];
};

struct GjfParams {
    float a[kMaxSpecies], b[kMaxSpecies], nmag[kMaxSpecies], cf[kMaxSpecies], cn[kMaxSpecies];
    float inv_m[kMaxSpecies]; // VV::finishInitialization (ps_integratorVV.cu:15-26)
};

// Brick decomposition of the grid.  A particle belongs to the brick holding
// its base grid index g (the index the spline stencil starts from, before the
// -p/2 shift); its stencil then lies in the brick's tile of b+p points per
// axis starting at the brick origin - p/2.  Tiles have the fixed pitch BMAX+p
// so stencil offsets are compile-time constants, and are stored in a
// halo-expanded grid (brick k at k*pitch per axis) so rows of neighbouring
// bricks are adjacent.  Sort keys are ((brick * nspecies + species) << 6) |
// sub-cell (4^3 sub-cells of the brick), so the particles of one brick, and of
// one species in it, are contiguous.
struct TileGeom {
    int b[3];    // brick extent per axis (power of two, <= BMAX)
    int sb[3];   // log2 of b
    int nb[3];   // bricks per axis
    int lbmax;   // log2 BMAX
    int pitch;   // BMAX + p
    int nbricks;
    int nspec;   // species per brick in the key
    int ex[3];   // halo-expanded tile grid: nb * pitch per axis
};

__device__ __host__ __forceinline__ int tile_key(int g0, int g1, int g2, int s, const TileGeom& tg)
{
    int k0 = g0 >> tg.sb[0], k1 = g1 >> tg.sb[1], k2 = g2 >> tg.sb[2];
    int sh = tg.lbmax - 2; // sub-cells of (BMAX/4)^3 cells
    int l0 = (g0 & (tg.b[0] - 1)) >> sh, l1 = (g1 & (tg.b[1] - 1)) >> sh, l2 = (g2 & (tg.b[2] - 1)) >> sh;
    int brick = k0 + tg.nb[0] * (k1 + tg.nb[1] * k2);
    return ((brick * tg.nspec + s) << 6) | l0 | (l1 << 2) | (l2 << 4);
}

// Workgroup -> brick: each of the 8 XCDs (round-robin dispatch) gets a
// contiguous slab of bricks, so tiles that share halos meet in one L2.
// Density at one grid point from the halo tiles that cover it: tz, ty, tx
// are the point's rows of the cover table (count, then up to 4 offsets into
// the halo-expanded tile grid) along z, y and x; the entries are added in
// the fixed order z, then y, then x.  Points covered at most twice along
// every axis (all but tiny grids) issue their loads together.
__device__ __forceinline__ float tile_sum(const float* __restrict__ tc, const int* tz, const int* ty, const int* tx)
{
    const int cz = tz[0], cy = ty[0], cx = tx[0];
    float s = 0.f;
    if (cz <= 2 && cy <= 2 && cx <= 2) {
        float v[8];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    const bool in = a < cz && b < cy && d < cx;
                    v[4 * a + 2 * b + d] = in ? tc[tz[1 + a] + ty[1 + b] + tx[1 + d]] : 0.f;
                }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int d = 0; d < 2; ++d)
                    if (a < cz && b < cy && d < cx) s += v[4 * a + 2 * b + d];
        return s;
    }
    for (int a = 0; a < cz; ++a)
        for (int b = 0; b < cy; ++b) {
            const float* row = tc + tz[1 + a] + ty[1 + b];
            for (int d = 0; d < cx; ++d) s += row[tx[1 + d]];
        }
    return s;
}

__device__ __forceinline__ int brick_of_block(int blk, int nbricks) { return xcd_slab(blk, nbricks); }

// Force fields (nout4 of them, M points each): a point's (fx, fy, fz) as three
// consecutive floats, 12 bytes a point, so a window of points spans a quarter
// fewer cache lines than padded float4 points.  The buffers keep 16 bytes a
// point: field f still starts at F4 + f M and uses its first 3 M floats.
__device__ __forceinline__ float3 ff_load(const float4* __restrict__ F, long i)
{
    const float* p = reinterpret_cast<const float*>(F) + 3 * i;
    return make_float3(p[0], p[1], p[2]);
}
__device__ __forceinline__ void ff_store(float4* __restrict__ F, long i, float x, float y, float z)
{
    float* p = reinterpret_cast<float*>(F) + 3 * i;
    p[0] = x;
    p[1] = y;
    p[2] = z;
}

struct TildEngine {
    mft_tild_desc d;
    std::vector<mft_tild_potential> pots;
    std::vector<float> mass, mob;
    int N = 0, M = 0, nchan = 0, qchan = -1, nforce = 0, nout = 0, nout4 = 0, npk = 0, qslot = -1, charge_pot = -1;
    int configured = 0;
    float dx[3] = {0, 0, 0};
    float gvol = 0;
    hipStream_t stream = nullptr;
    // particle state in key order (brick, species, sub-cell), double-buffered
    // for the per-step permutation into the next step's order
    float4* X[2] = {nullptr, nullptr};   // x, y, z, charge
    float4* XO[2] = {nullptr, nullptr};  // previous positions (GJF), 12 bytes a slot (ff_load / ff_store)
    float4* NZ[2] = {nullptr, nullptr};  // previous noise (GJF)
    PMeta* META[2] = {nullptr, nullptr}; // particle id, species
    int* BS[2] = {nullptr, nullptr};     // bond partners as slots [N][maxbonds], -1 past the last
    int* BT[2] = {nullptr, nullptr};     // bond types [N][maxbonds]
    int* NA[2] = {nullptr, nullptr};     // angle count (angles only)
    int* AT[2] = {nullptr, nullptr};     // angle types [N][maxangles]
    int* AG[2] = {nullptr, nullptr};     // angle members as slots [N][maxangles][3]
    int *nAngles = nullptr, *angleType = nullptr, *angleGroup = nullptr; // particle-id order
    AngleParams ap{};
    int cur = 0;
    int vv_state = 0; // velocity Verlet: 0 before the first step, 1 afterwards
    // GJF: the previous noise of every particle is nmag(type) times the
    // Philox normals keyed by (particle, step - 1), because the last step was
    // a GJF step of this engine; the integrator then regenerates it instead of
    // reading NZ, and NZ is not carried through the permutation
    bool nz_keyed = false;
    // step index on the device for graph-captured steps: the integrator
    // reads it and the key scan advances it, so replays key their noise by
    // the step they run (null kernel argument on the ordinary path)
    uint32_t* d_step = nullptr;
    bool use_dstep = false;
    // step temporaries in slot order
    float4 *XN = nullptr, *NN = nullptr; // next positions (12 bytes a slot: ff_load / ff_store) and noise
    float4* fgrid = nullptr;             // grid force per slot, 12 bytes a slot (ff_load / ff_store)
    float4* ftot = nullptr;              // total force per slot (force evaluations)
    int *key = nullptr, *rank = nullptr, *map = nullptr;
    // topology and state staging in particle-id order
    int *type = nullptr, *nBonds = nullptr, *bondedTo = nullptr, *bondType = nullptr;
    float* charge = nullptr;
    float* pio = nullptr;                // [3N] host-order positions / forces
    float* pio2 = nullptr;               // [3N]
    float* pio3 = nullptr;               // [3N]
    int* slot_of = nullptr;              // particle id -> slot
    // binning
    TileGeom tg;
    int nkeys = 0;
    int *count = nullptr, *start = nullptr, *partial = nullptr;
    int nscan_blocks = 0;
    // grids
    float* rho = nullptr;   // [nchan][M], reduced from the tiles when read
    bool rho_valid = false;
    float* tiles = nullptr; // [nchan][halo-expanded grid] density tiles
    int* tile_tab = nullptr; // per-axis tile offsets contributing to each grid coordinate
    float2* S = nullptr;    // [nchan][M]
    float2* Y = nullptr;    // [npk][M] force spectra, two real components per complex field
    float4* F4 = nullptr;   // [nout4][M] force fields, 12 bytes a point (ff_load / ff_store)
    int* force_slot = nullptr; // [nspecies] -> force field slot or -1
    std::vector<int> force_slot_h;
    std::vector<float*> tables; // erf tables per potential (device)
    // static bias fields (BiasField): potential index, species mask, the
    // spectral force field and A u(r) on the grid (device)
    struct Bias {
        int pot;
        uint32_t mask;
        float4* F;
        float* u;
    };
    std::vector<Bias> bias;
    WallParams walls{};
    LangevinParams lang{};
    bool is_vv = false; // created with MFT_INT_VV (velocities live in the noise slot)
    DpdParams dpd{};
    int *dcount = nullptr, *dstart = nullptr, *dcursor = nullptr, *dcell = nullptr, *dslots = nullptr;
    float4* fdpd = nullptr;
    std::vector<int> wall_pot;  // potential index of each wall
    double* epart = nullptr;
    double* eout = nullptr;   // [kMaxPots + 1]
    double* bpart = nullptr;
    BondParams bp;
    GjfParams gp;
    KParams kp;
    size_t dens_lds = 0;
    uint32_t tk_mask = ~0u;      // kernel groups bracketed by timing events
    int gather_mode = 1;    // 0 direct, 1 LDS windows per (brick, species), 2 per brick (mft_tild_autotune picks)
    int plane_passes = 0;   // 1: x and y transforms fused per z-plane (mft_tild_autotune picks)
    int xfwd_staged = 0;    // 1: forward x pass stages its tile rows in LDS (mft_tild_autotune picks)
    int tile_maxcover = 4;  // most tiles covering one coordinate along any axis
    // the autotuner's timings (ms per launch or per grid-force pass; -1 where
    // not timed): gather variants 0-3, x-y per axis / per plane, x pass
    // direct / staged
    double tune_ms[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    int64_t step = 0;
    int64_t run_max = 0, run_pos = 0; // declared run (mft_tild_set_run)
    int have_state = 0;
    // timing: event pairs recorded around kernel launches on the engine
    // stream without host synchronisation, resolved when timing is read
    bool timing = false;
    std::vector<hipEvent_t> evpool;
    std::vector<int> evkind; // kernel id of pair i
    size_t evused = 0;
    double tk_ms[MFT_TK_COUNT];
    int64_t tk_n[MFT_TK_COUNT];
};

// Bond types of buffer c in key order, or null with at most one bond type:
// every bond is then type 0 (set_topology checks the range), so the slots'
// types are neither carried through the permutation nor read.
inline int* slot_bond_types(const TildEngine* e, int c) { return e->bp.ntypes > 1 ? e->BT[c] : nullptr; }

constexpr int kBlock = 256;
inline unsigned nblk(long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanChunk = kScanThreads * kScanItems;

template <typename T>
static int dalloc(T** p, size_t n)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) {
        set_error(MFT_ERR_OOM, "hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
        return MFT_ERR_OOM;
    }
    return MFT_OK;
}

// ---- grid passes (tild_grid.hip, tild_conv.hip) ----
int tild_axis(int sign, int real_in, int real_out, const void* in, void* out, AxisGeom g, int nf, long fs,
              float scale, hipStream_t st);
int tild_x_fwd_tiles(const float* tiles, const int* tab, float2* S, int nx, int ny, int nz, int nchan, long cstride,
                     hipStream_t st, int ex0 = 0, bool staged = false, bool cover2 = false);
bool tild_x_fwd_staged_ok(int nx, int ex0, int maxcover);
int tild_x_inv_packed(const float2* Y, float4* F4, int nx, int ny, int nz, int nfield, hipStream_t st);
bool tild_plane_supported(int nx, int ny);
int tild_xy_fwd_tiles_plane(const float* tiles, const int* tab, float2* S, int nx, int ny, int nz, int nchan,
                            long cstride, hipStream_t st);
int tild_yx_inv_packed_plane(const float2* Y, float4* F4, int nx, int ny, int nz, int nfield, hipStream_t st);
int tild_conv_blocks(int nx, int ny, int nz);
int tild_conv(TildEngine* e);

// ---- particle passes (tild_particles.hip) ----
int tild_tiles_configure(TildEngine* e); // brick geometry, LDS sizes, tile tables
int tild_sort_state(TildEngine* e);     // particle-id-order staging -> key order
int tild_unsort(TildEngine* e, int what); // key order -> particle-id-order staging
int tild_scan(TildEngine* e, bool advance_step = false);
int tild_permute(TildEngine* e);         // next positions into the next key order
int tild_density(TildEngine* e);          // tiles of the current positions
int tild_density_reduce(TildEngine* e);   // tiles -> rho (when the densities are read)
int tild_gather(TildEngine* e, int mode);
bool tild_gather_brick_fits(const TildEngine* e);   // gather variant 2's windows fit in LDS
int tild_integrate(TildEngine* e, bool want_f);
void erf_table(int kind, double sigma, double Rp, int nx, int ny, int nz, float lx, float ly, float lz,
               std::vector<float>& out);
int tild_bond_energy(TildEngine* e);
int tild_bond_virial(TildEngine* e, float* vir);   // [6N] device, particle-id order
int tild_bias_setup(TildEngine* e, int pot);      // the field and its force (tild_bias.hip)
int tild_bias_add(TildEngine* e);                 // F4 += bias forces of the biased species
int tild_bias_energy(TildEngine* e, int k, double* energy);
int tild_wall_energy(TildEngine* e, int w, double* energy); // tild_particles.hip
int tild_dpd_setup(TildEngine* e, int pot);                  // tild_dpd.hip
int tild_dpd_forces(TildEngine* e);
int tild_angle_energy(TildEngine* e, double* energy);         // tild_particles.hip

} // namespace mft
