// Internal device-side data structures of the persiansph_amd GPU path (gfx950).
//
// Particles live in HBM as structure-of-arrays, sorted by the reference's cell key
// kx + nx*(ky + ny*kz) (the CellInitiate grid, Source/Domain.cpp:549-660), and inside a
// cell by original (reference) index.  A sort is a counting sort on the cell key (a one-digit
// radix sort whose digit is the whole key), cell-start table, per-cell order by original index,
// then a gather of the persistent columns into the new order.  It runs every step, or, with
// neighbour-list reuse, at every list build; steps between builds keep the order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace psph {

constexpr int kMaxClasses = 64;

// Constant Particle fields shared by a group of particles (Source/Particle.h:29-149).
// Built on the host at upload from the distinct tuples present; particles carry an index.
struct ClassParams {
    double h, Cs, P0, RefDensity, FPMassC;
    double Alpha, Beta, MuRef, T0, m, CSmag;
    double TI, TIn, TIInitDist;
    double G, K, Sigmay, c, phi, psi;
    // derived constants (computed once on the host from the fields above)
    double dp_alpha2, dp_kf2, dp_alpha3, dp_kf3;   // Drucker-Prager alpha, kf (plane strain, 3-D)
    double sin_psi;                                // sin(psi) for the non-associated flow rule
    double V, RhoF, k, k2, d, n0;                  // soil-water interaction (Particle.h:84-93)
    int32_t IsFree, NoSlip, Shepard, LES, Material, PresEq, VisM, Fail, ID, ShepardStep, InOut, SeepageType;
};

// Flags word per particle.
constexpr uint32_t kClsMask = 0xffffu;
constexpr uint32_t kFree = 1u << 16;
constexpr uint32_t kNoSlip = 1u << 17;
constexpr int kMatShift = 18;      // 2 bits: Material 1..3
constexpr uint32_t kSat = 1u << 20;       // IsSat: saturated soil (density and RefDensity less RhoF)
constexpr uint32_t kSatCheck = 1u << 21;  // SatCheck of the last force pass
constexpr int kInOutShift = 22;           // 2 bits: InOut (1 inflow, 2 outflow)
__host__ __device__ inline int flag_inout(uint32_t f) { return (int)((f >> kInOutShift) & 3u); }

__host__ __device__ inline int flag_cls(uint32_t f) { return (int)(f & kClsMask); }
__host__ __device__ inline bool flag_free(uint32_t f) { return (f & kFree) != 0; }
__host__ __device__ inline bool flag_noslip(uint32_t f) { return (f & kNoSlip) != 0; }
__host__ __device__ inline int flag_mat(uint32_t f) { return (int)((f >> kMatShift) & 3u); }

// Halt reasons written by the device into Ctl::halt.
enum : int {
    kRun = 0,
    kHaltSteps = 1,    // max_steps reached
    kHaltTend = 2,     // Time >= tf at a step head
    kHaltOutput = 3,   // Time >= tout after the force pass (mid-step)
    kHaltLeave = 4,    // particles outside [BLPF, TRPR] after Move
    kHaltError = 5,    // see Ctl::error
    kHaltHookBefore = 6,   // host hook before the wall smoothing and the pair sums (mid-step)
    kHaltHookAfter = 7,    // host hook after the pair sums, before Move (mid-step)
    kHaltSort = 8,     // list reuse: the next step rebuilds, but was queued without its sort
                       // and list-build kernels (Params::lean); the host queues it in full
};

// Solve-loop scalars, resident in device memory so steps run without host round trips.
struct Ctl {
    double time, deltat, deltatint, deltatmin;
    double tf, tout;
    unsigned long long dtmin_bits;        // atomic min of sqrt_h_a*sqrt(h/|a|) (positive doubles)
    long long steps_done, max_steps;
    unsigned long long pairs_ff, pairs_fw;  // directed free-free and free-fixed in-support pairs
    unsigned long long last_pairs;        // unordered in-support pairs of the last force pass
    // mid_step: 1 = stopped after the force pass (Move pending), 2 = stopped before the wall
    // smoothing (host hook GeneralBefore); resume (set by the run head): 1 = the step continues
    // past the hook stop before the wall pass, 2 = past the stop after the force pass
    int halt, error, mid_step, resume;
    // the pair sums of this step start from acc (accelerations and density rates the host wrote
    // at a stop before the wall pass) instead of Gravity plus the body force
    int acc_init;
    int leave, xmat;
    int nl_over;                        // targets whose neighbours overflowed the lists (cumulative)
    int nl_q;                           // overflow queue length of the current force pass
    // rebuild = this step sorts (always 1 without list reuse, see Params::verlet);
    // need_rebuild = some free particle moved more than half the skin since the last build
    // (set by integrate, turned into rebuild by the tail); list_over = bit 0: a free target's
    // skin list overflowed, bit 1: a wall target's did (every step rebuilds until the host
    // has grown the lists)
    int rebuild, need_rebuild, list_over;
    int side;                           // buffer of the persistent columns at the step's start
    int nwact;                          // wall targets with skin neighbours at the last build
    long long rebuilds;                 // list builds (sorts) since upload
    int ct, shep_counter, first_step, n_out;   // n_out: outflow particles of this step
    long long reorders;                 // completed reorders (buffer parity)
    int ntiles;                         // tiles of the FP64 fast path at the last list build
};

// Run-wide constants.
struct Params {
    double blpf[3], trpr[3], csize[3];
    double hmax, cellfac, sqrt_h_a, xsph;
    double gravity[3];
    double eos_cs, eos_p0, eos_rho0;   // the free particles' EOS (uniform, checked at upload)
    int eos_eq;
    // per material (index 1..3): the EOS of its free particles, which a wall of that material
    // takes for its density (uniform per material, checked at upload)
    int eos_eq_m[4];
    double eos_cs_m[4], eos_p0_m[4], eos_rho0_m[4];
    // pairs of different materials (fluid with solid, or fluid with soil): CalcForce12/13 in
    // k_cross, the FSI smoothing (fsi) and the saturation of soil (swi_type != 2)
    int multi, fsi, swi_type, any_soil_fluid, any_solid_fluid;
    int swi1;                          // SWIType 1 with soil and fluid: S and ZWab go to swz
    // in/out-flow boundaries (Boundary::InOutFlow 1, 2, 3) with the constant default conditions:
    // inflow particles (InOut 1) carry in_v and in_density, outflow particles (x >= out_loc,
    // InOut 2) out_v scaled by the mass-conservation factor n_in/n_out and out_density if
    // positive; an inflow particle past loc1 or a particle outside the box halts the step for
    // the host's recycling
    int inout, mass_cons, n_in;
    double in_v[3], out_v[3], in_density, out_density, loc1, out_loc;
    int nc[3], ncells;
    int n, nfree, nfixed, nclasses;
    int dim, kernel_type, visc_eq, gradient_type, scheme, precision;
    int shepard_step;                  // uniform over free particles
    int any_ti, any_sr, any_xsph, any_shepard, any_solid, uniform_h, any_les_t0;
    double h_uniform;
    // fast fluid path (cubic spline, Morris, squared-density form, uniform h and viscosity)
    int fast11;
    double mu_ff;                      // 2 mu mu / (mu + mu) of two free particles
    double mu_ff_wall;                 // the free particles' mu (a wall pair uses the free side's mu)
    double kw, kg, kgh, invh, h2c;     // kernel constants: C of W, C of (1/r)dW/dr, C/h, 1/h, 0.01 h^2
    double sup2_lo, sup2_hi;           // (Cellfac h)^2 band around the exact rij/h <= Cellfac test
    double pair_alpha[64], pair_beta[64];   // (Alpha_i + Alpha_j)/2 per class pair (<= 8 classes)
    // every free particle of one class whose pair constants with any fixed particle are the
    // same: (alpha, beta) of a free-free pair (ab_free) and of a free-fixed pair (ab_fixed);
    // ab_one: the two pairs are equal
    int ab_two, ab_one;
    double ab_free[2], ab_fixed[2];
    // the same pairs as the gathers use them: -alpha/2 (free-free, free-fixed), FP32 copies
    // (-alpha/2, beta of free-free, then of free-fixed), and whether alpha or beta is nonzero
    double ab_nha[2];
    float ab32[4];
    int ab_on[2];
    // FP32-mixed tier (fast fluid configuration only)
    int mixed32;
    // its constants rounded to FP32 once by the host, so that the gathers take them as scalar
    // operands: kernel constants, 1/h, h, 0.01 h^2, cell size, the skin test's FP32 band
    // (r^2 bounds and the adjacency bound, widened by 1e-4), and 2 mu of both pair kinds
    struct {
        float kw, kg, kgh, invh, h, h2c, cs[3], band_lo, band_hi, band_adj, mu2_ff, mu2_wall;
    } f32;
    // neighbour lists of the fast fluid path: capacity per free target (a multiple of 4), and
    // whether entries carry the neighbour cell code (needs fewer than 2^27 particles)
    int nl_cap, nl_codes;
    // FP32-mixed tier of solid/soil runs: CalcForce2233 pair arithmetic and sums in FP32
    int soil32;
    // neighbour-list reuse of the fast fluid path: the particle order, the cell table and the
    // lists are rebuilt only when some free particle has moved more than half the skin since
    // the last build.  Lists then hold every candidate within the support radius plus the skin
    // (reach); each step keeps exactly the reference's pairs from them.
    int verlet, verlet_force;          // reuse on; rebuild every step anyway (test switch)
    int wl_cap;                        // wall sub-list capacity per (wall target, lane), multiple of 4
    double reach2;                     // (support + skin)^2 with a small margin
    double skin_trig2;                 // (skin/2)^2 less a small margin: the rebuild trigger
    double adj2;                       // squared smallest cell size (less a margin): pairs closer
                                       // than this are in neighbouring cells
    // periodic directions (Boundary::Periodic): nc[] counts the cells that hold particles (the
    // reference's two alias layers are not stored; neighbour cells wrap instead); domsize =
    // DomSize, the period
    int any_periodic, periodic[3];
    double domsize[3];
    // general kernels: the pairs of each free target are listed by a light kernel (k_glist) in
    // the force kernel's walk order, in nl/ncnt with capacity nl_cap; the force kernel walks
    // the cells itself only for the entries beyond the capacity
    int glist;
    // a free particle's acceleration at the start of its force pass: Gravity plus the body force
    // of the body-force hook (GeneralBefore adding a constant to every free particle's a)
    double afree[3];
    // host hooks: stop before the wall smoothing (GeneralBefore) and after the pair sums
    // (GeneralAfter) of every step
    int hook_before, hook_after;
    // FP64 fast path with list reuse: the gather runs per tile with the neighbours' records in
    // LDS (see Dev::tile); tile_max bounds the tile count (grid size), tl_stride is the list
    // slots per tile (kTileTargets times the list capacity rounded up to 4 kTileLanes)
    int tiles, tile_max, tl_stride;
    int tile_items;                    // items a tile may stage (kTileItems; less: test switch)
    // list reuse, set per launch by the host: the step is queued without the kernels that only
    // work on rebuilding steps (sort, list build, tile union), and the tail halts (kHaltSort)
    // when the step after it must rebuild
    int lean;
    // set per launch by the host: the step after this one is queued lean, and this step's
    // k_integrate writes that step's free-particle views (k_reorder is then not queued for it)
    int views_next;
    // runs of one material whose free particles are all soil: the pair records of Dev::srec
    int srec;
};

// One set of the persistent columns.
struct Cols {
    double4 *pos, *vel, *vh;
    double2 *mmu;
    uint32_t *flags;
    int32_t *orig;
    double *sig, *sigh, *tau, *tauh, *eps, *epsh;
};

// Device buffers.  The persistent columns are double-buffered (buf[0], buf[1]).  Kernels see
// them through pos..epsh, the state at the step's start (buf[Ctl::side]), and pos_b..epsh_b,
// the step's working copy (buf[side ^ rebuild]): the reorder writes it, permuted, on steps that
// sort, and it is the same buffer on steps that keep the order.  The kernels bind these views
// from the control block at their start (bind() in the kernels).
struct Dev {
    Ctl *ctl;
    ClassParams *cls;
    Cols buf[2];
    // persistent (current order)
    double4 *pos;      // x, y, z, Density
    double4 *vel;      // vx, vy, vz, Pressure
    double2 *mmu;      // Mass, Mu
    double4 *vh;       // Leapfrog: va, Densitya (free) / MVerlet: vb, Densityb; fixed: vb, Densityb
    uint32_t *flags;
    int32_t *orig;
    double *sig;       // Sigma (6: xx xy xz yy yz zz)
    double *sigh;      // Leapfrog Sigmaa / MVerlet Sigmab
    double *tau;       // ShearStress (6)
    double *tauh;      // Leapfrog ShearStressa / MVerlet ShearStressb
    double *eps;       // Strain (6)
    double *epsh;      // Leapfrog Straina / MVerlet Strainb
    // persistent, other buffer (reorder destination)
    double4 *pos_b, *vel_b, *vh_b;
    double2 *mmu_b;
    uint32_t *flags_b;
    int32_t *orig_b;
    double *sig_b, *sigh_b, *tau_b, *tauh_b, *eps_b, *epsh_b;
    // transient (current order)
    double4 *aux;      // Mass, Mu, sound speed, P/d^2  (fixed: wall density-based values)
    // gather views for the fast fluid path, written per step by reorder (free) and wall (fixed):
    // g0 = x, y, z, d   g1 = v, P/d^2   g2 = c, meta, m (free) or FPMassC (fixed), 1/d
    // gw = v_visc (fixed particles only)
    // (d, P, c: the density, pressure and sound speed the pair formulas use for that particle;
    //  meta = class for free particles, -1 - class for fixed ones; v_visc = 2 v - NSv for no-slip
    //  walls, v otherwise, and a free neighbour's v_visc is its v)
    double4 *g0, *g1, *g2, *gw;
    // FP32-mixed gather views: f0 = position relative to the particle's cell origin, d;
    // f1 = v, P/d^2; f2 = c, meta, m or FPMassC, 1/d; fw = v_visc (fixed particles only)
    float4 *f0, *f1, *f2, *fw;
    double4 *wal;      // fixed: NSv x,y,z, wall density d_w
    double4 *acc;      // a x,y,z, dDensity
    double4 *xsp;      // VXSPH x,y,z, ZWab
    double4 *fsi;      // FSINSv x,y,z, FSIPressure (free solid particles next to fluid; zero otherwise)
    double *fsk;       // FSISumKernel
    double2 *swz;      // SWIType 1: S and ZWab of the last force pass (fixed: 0 and 0/1)
    int32_t *xflag;    // free target has pairs with particles of another material (per target)
    double2 *shep;     // SumDen, SumKernel
    double *sr;        // StrainRate (6)
    double *rr;        // RotationRate (3: xy xz yz)
    double *tir;       // TIR (6)
    // free soil particles' pair records (Params::srec), 10 double2 each: x y | z d | vx vy |
    // vz m | Sigma (6) | TIR (6), one contiguous 160-B read per neighbour of CalcForce2233
    double2 *srec;
    double *shr;       // ShearRate, SBar (2)
    int32_t *key;      // cell key of each particle in the current order
    int32_t *rank;     // arrival rank within its cell (counting sort)
    int32_t *key_b;    // key in the new order
    int32_t *src;      // new position -> old position, arrival order inside a cell
    uint32_t *sorig;   // original index of src[p] (arrival order), bit 31 set for free particles
    int32_t *perm;     // new position -> old position, original-index order inside a cell
    int32_t *cstart;   // cell start table (ncells + 1)
    int32_t *fstart;   // free particles before each cell (ncells + 1): free-target list offsets
    int32_t *ccount;   // per-cell counts (zeroed by the scan)
    int32_t *cfcnt;    // per-cell free-particle counts being accumulated for the next step
    int32_t *cfree;    // per-cell free-particle counts of the current step
    unsigned long long *bsum;   // scan block sums (packed all/free counts)
    int32_t *free_list, *fixed_list;
    // neighbour lists of the free targets (fast fluid path): entry e of target t at
    // nl[((e/4)*nfree + t)*4 + e%4], holding j | code << 27 (code = 0..26, the neighbour cell
    // (dx+1) + 3(dy+1) + 9(dz+1) relative to the target's cell); ncnt[t] = in-support count,
    // which may exceed the capacity (the gathers then rescan for the rest).  With list reuse
    // the lists are skin lists: every candidate within the support radius plus the skin at the
    // last build, j | code << 25 with the code over the 5^DIM cells around the target's cell
    int32_t *nl, *ncnt;
    // overflow queue of the force pass: target t and its partial sums (k_force11_nl_over ends it)
    int32_t *ovq;
    double4 *ovs;
    // list reuse: skin sub-lists of the wall targets, one per (target, lane) gt = 4 t + lane,
    // entry e at wl[((e/4)*4 nfixed + gt)*4 + e%4], and their counts; positions at the last
    // build.  (The free targets' skin lists are nl/ncnt themselves.)
    int32_t *wl, *wcnt;
    int32_t *wact;     // wall targets with skin neighbours at the last build (Ctl::nwact of them)
    double4 *xref;
    // tiles of the FP64 fast path with list reuse (Params::tiles): the free targets of one cell
    // row in groups of at most kTileTargets, tile = (first target, targets, free items, wall
    // items; -1 wall items: the tile's neighbours do not fit, it reads them from HBM).  A tile's
    // items are the union of its targets' skin-list neighbours (free ones first, then walls, each
    // in the order of the cell rows around the tile's row), and its targets' lists hold item
    // slots (16 bits) at tl[tile*tl_stride +
    // tl_index(target, e)]: lane e % kTileLanes of a target reads its slots four at a time
    int4 *tile;
    int32_t *items;
    uint16_t *tl;
};

// tiles of the FP64 fast path: targets per tile, lanes per target, the items and wall items
// whose gather records fit a workgroup's LDS, and the bitmap words of the union build
constexpr int kTileTargets = 64;
constexpr int kTileLanes = 8;
constexpr int kTileItems = 720;
constexpr int kTileWalls = 320;
constexpr int kTileWords = 2048;
constexpr int kSlotWords = 6;          // slot words a lane holds in registers (4 slots each)
constexpr int kUnionBlocks = 1024;     // grid of the tile union (four blocks per CU)
static_assert(kTileItems <= 2*kTileTargets*kTileLanes, "staging takes at most two items a thread");

}  // namespace psph
