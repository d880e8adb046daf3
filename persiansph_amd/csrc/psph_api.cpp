// C ABI of persiansph_amd (include/persiansph_amd.h): host side of the GPU Solve path.
//
// Upload restates Solve's initialisation (Source/Domain.cpp:1735-1754): InitialChecks,
// CellInitiate, TimestepCheck, then builds the per-particle class table and the device
// structure-of-arrays.  psph_run queues steps on the solver's stream in batches; the device
// control block (Ctl) carries Time, dt and the stop conditions, so the host only reads it back
// once per batch.  Deletion of particles leaving the domain (CheckParticleLeave,
// Domain.cpp:517-547) is handled here by an order-preserving compaction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/persiansph_amd.h"
#include "psph_dev.h"

namespace psph {
void launch_head(const Dev &D, const Params &P, hipStream_t s);
void launch_keys(const Dev &D, const Params &P, hipStream_t s);
void launch_scan(const Dev &D, const Params &P, hipStream_t s);
void launch_scatter(const Dev &D, const Params &P, hipStream_t s);
void launch_cellsort(const Dev &D, const Params &P, hipStream_t s);
void launch_reorder(const Dev &D, const Params &P, hipStream_t s);
void launch_wall(const Dev &D, const Params &P, hipStream_t s, int after_hook);
void launch_hook_accel(const Dev &D, const Params &P, const double *a, const double *dd, hipStream_t s);
void launch_fixed_sums(const Dev &D, const Params &P, hipStream_t s);
void launch_force(const Dev &D, const Params &P, hipStream_t s);
void launch_integrate(const Dev &D, const Params &P, hipStream_t s);
void launch_tail(const Dev &D, const Params &P, hipStream_t s);
void launch_pairs(const Dev &D, const Params &P, hipStream_t s, int pass, const long long *off,
                  int *cnt, long long *out);
int scan_blocks(const Params &P);
int output_floats();
void launch_pack_output(const Dev &D, const Params &P, float *out, int acc_ok, int tensors, hipStream_t s);
}  // namespace psph

using namespace psph;

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// error reporting for the other entry points of the library (the surrogate's psph_gnn_*)
int psph_internal_fail(int code, const char *msg) { return fail(code, "%s", msg); }

#define HIPCHECK(expr)                                                                        \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) return fail(PSPH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------------------------------
// Staging of psph_neighbour_pairs_noalloc: a file-scope device array (one copy per GPU, held
// until the process ends) in place of the per-call allocation of psph_neighbour_pairs.  Host
// code never takes the array's name as an address: device_scratch() looks up the current
// device's copy with hipGetSymbolAddress, once per device, and every use goes through that
// pointer.  All calls on one GPU share its copy, so pair queries that overlap on one GPU race.
// ------------------------------------------------------------------------------------------
constexpr size_t kScratchBytes = size_t(32) << 20;                     // 32 MiB
constexpr size_t kScratchPairs = kScratchBytes/(2*sizeof(long long));  // 2 Mi pairs
static __device__ __align__(256) long long scratch[kScratchBytes/sizeof(long long)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
long long *g_scratch_dev[kMaxDevices] = {};
}  // namespace

// The current device's copy of `scratch` (the caller has made `device` current).
static int device_scratch(int device, long long **out)
{
    if (device < 0 || device >= kMaxDevices) return fail(PSPH_ERR_INVALID, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PSPH_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch_dev[device] = static_cast<long long *>(p);
    }
    *out = g_scratch_dev[device];
    return 0;
}

namespace {

enum KernelId { KHead, KKeys, KScan, KScatter, KCellsort, KReorder, KWall, KForce, KIntegrate, KTail, KCount };
const char *kKernelNames[KCount] = {"head", "keys", "scan", "scatter", "cellsort", "reorder",
                                    "wall", "force", "integrate", "tail"};

struct Column { void *p = nullptr; size_t bytes = 0; };

}  // namespace

struct psph_handle {
    psph_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    Params P{};
    Dev D{};                    // device buffers; buf[Ctl::side] holds the state between steps
    std::vector<void *> allocs;
    Ctl *ctl_dev = nullptr;
    Ctl *ctl_host = nullptr;    // pinned
    ClassParams *cls_dev = nullptr;
    std::vector<ClassParams> classes;
    std::vector<psph_particle> host;   // reference-order mirror (fields the GPU does not keep)
    std::vector<int32_t> upload_index; // position of each mirrored particle in the uploaded array
    size_t cap = 0;
    bool uploaded = false;
    bool broken = false;
    psph_grid grid{};
    bool trans_ok = false;             // transient columns match the current order
    bool fixed_ok = false;             // fixed particles' pair sums formed for this stop
    bool no_lean = false;              // PSPH_NO_LEAN: queue every step in full (test switch)
    bool no_fused_views = false;       // PSPH_NO_FUSED_VIEWS: every step's views by k_reorder (test switch)
    double body[3] = {0.0, 0.0, 0.0};  // body-force hook (psph_set_body_force)
    bool hook_b = false, hook_a = false;   // host hook stops (psph_set_hooks)
    int regrowths = 0;                 // device buffers re-created for more particles (in/out-flow)
    double *hook_dev = nullptr;        // staging of psph_hook_accel's values (4 doubles per particle)
    size_t hook_cap = 0;
    // in/out-flow: Boundary::InPart (host indices, the reference's order), the zone limits,
    // inoutcounter; the next upload index for particles the recycling creates
    std::vector<int32_t> inpart;
    double loc2 = 0.0, loc3 = 0.0;
    int inoutcounter = 0;
    int32_t next_index = 0;
    // timing
    bool timing = false;
    bool timing_force_only = false;   // events around the force pass only (PSPH_TIMING_FORCE)
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
    double t_ms[KCount] = {0};
    long long t_n[KCount] = {0};
    // scratch for the pair query; the device's copy of the static staging array of
    // psph_neighbour_pairs_noalloc (looked up by psph_create) or the lookup's error
    int *pair_cnt = nullptr;
    long long *pair_off = nullptr;
    long long *pair_scratch = nullptr;
    std::string pair_scratch_err;
    // neighbour lists of the fast fluid path (capacity in Params::nl_cap; skin lists with list
    // reuse); skin sub-lists of the wall targets (capacity Params::wl_cap)
    int32_t *nl = nullptr;
    int32_t *wl = nullptr;
    // visualisation output: FP32 staging on the device, its pinned host copy (out_cap particles
    // each), the copy's completion event; fields packed on the host instead (reference order)
    float *out_dev = nullptr;
    float *out_host = nullptr;
    size_t out_cap = 0, out_n = 0;
    hipEvent_t out_event = nullptr;
    bool out_pending = false;
    std::vector<float> out_acc, out_sig, out_eps;
};

// ------------------------------------------------------------------------------------------
// helpers
// ------------------------------------------------------------------------------------------
static int dev_alloc(psph_handle *h, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    HIPCHECK(hipMalloc(p, bytes));
    HIPCHECK(hipMemset(*p, 0, bytes));
    h->allocs.push_back(*p);
    return 0;
}

static int dev_free(psph_handle *h, void **p)
{
    if (*p) {
        HIPCHECK(hipFree(*p));
        for (auto &q : h->allocs) if (q == *p) q = nullptr;
        *p = nullptr;
    }
    return 0;
}

// (Re)allocates the neighbour lists for `cap` entries per free target.  The stream must be
// idle; lists reallocated under list reuse are rebuilt by the next step (the caller requests
// it).
static int nl_alloc(psph_handle *h, int cap)
{
    Params &P = h->P;
    // the gathers tally a lane's pairs in 16-bit fields (pair_tally)
    if (cap > 65535) return fail(PSPH_ERR_INVALID, "neighbour lists of %d entries per particle (at most 65535)", cap);
    int rc = dev_free(h, (void **)&h->nl);
    if (rc) return rc;
    size_t nf = P.nfree > 0 ? (size_t)P.nfree : 1;
    rc = dev_alloc(h, (void **)&h->nl, sizeof(int32_t)*(size_t)cap*nf);
    if (rc) return rc;
    P.nl_cap = cap;
    h->D.nl = h->nl;
    if (P.tiles) {
        // tiles of the fast path: at most nfree / kTileTargets + one per cell row
        P.tile_max = (int)(nf/kTileTargets) + P.nc[1]*P.nc[2] + 1;
        // slot lists: per target, groups of 4*kTileLanes entries (one 8-byte word per lane)
        P.tl_stride = kTileTargets*((cap + 4*kTileLanes - 1)/(4*kTileLanes))*4*kTileLanes;
        if ((rc = dev_free(h, (void **)&h->D.tile)) || (rc = dev_free(h, (void **)&h->D.items)) ||
            (rc = dev_free(h, (void **)&h->D.tl)))
            return rc;
        if ((rc = dev_alloc(h, (void **)&h->D.tile, sizeof(int4)*(size_t)P.tile_max)) ||
            (rc = dev_alloc(h, (void **)&h->D.items, sizeof(int32_t)*(size_t)P.tile_max*kTileItems)) ||
            (rc = dev_alloc(h, (void **)&h->D.tl, sizeof(uint16_t)*(size_t)P.tile_max*P.tl_stride)))
            return rc;
    }
    return 0;
}

// (Re)allocates the wall targets' skin sub-lists for `cap` entries per (target, lane).
static int wl_alloc(psph_handle *h, int cap)
{
    Params &P = h->P;
    int rc = dev_free(h, (void **)&h->wl);
    if (rc) return rc;
    size_t nw = P.nfixed > 0 ? 4*(size_t)P.nfixed : 4;
    rc = dev_alloc(h, (void **)&h->wl, sizeof(int32_t)*(size_t)cap*nw);
    if (rc) return rc;
    P.wl_cap = cap;
    h->D.wl = h->wl;
    return 0;
}

// First list capacity: 1.5x the neighbour count a uniform spread of the free particles over
// their bounding box would give inside the support radius, plus a margin; a multiple of 4.
// Targets that still overflow are handled exactly by the gathers, and psph_run grows the lists.
static int nl_initial_cap(const psph_handle *h, double radius)
{
    const Params &P = h->P;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    size_t nf = 0;
    for (const auto &q : h->host) {
        if (!q.IsFree) continue;
        ++nf;
        for (int c = 0; c < 3; ++c) { lo[c] = std::min(lo[c], q.x[c]); hi[c] = std::max(hi[c], q.x[c]); }
    }
    double R = radius;
    double vol = 1.0;
    for (int c = 0; c < P.dim; ++c) vol *= std::max(hi[c] - lo[c], R);
    double est = (double)nf/vol*((P.dim == 3) ? 4.0/3.0*M_PI*R*R*R : M_PI*R*R);
    double capd = std::min(1.5*est + 8.0, 1024.0);
    int cap = std::max(16, (int)std::ceil(capd));
    return (cap + 3) & ~3;
}

// First wall sub-list capacity: a wall next to the fluid sees about half the sphere, and one of
// its four lanes up to ~35 % of that; 1.5x margin, a multiple of 4.  Overflows make every step
// rebuild until psph_run has grown the sub-lists.
static int wl_initial_cap(const psph_handle *h, double radius)
{
    double est = (double)nl_initial_cap(h, radius);       // ~1.5x the full-sphere count + 8
    int cap = std::max(8, (int)std::ceil(0.5*0.35*est + 8.0));
    return (cap + 3) & ~3;
}

static void free_all(psph_handle *h)
{
    if (h->stream) hipStreamSynchronize(h->stream);
    for (void *p : h->allocs) hipFree(p);
    h->allocs.clear();
    if (h->ctl_host) { hipHostFree(h->ctl_host); h->ctl_host = nullptr; }
    if (h->out_host) { hipHostFree(h->out_host); h->out_host = nullptr; }
    if (h->out_event) { hipEventDestroy(h->out_event); h->out_event = nullptr; }
    h->out_dev = nullptr;               // in allocs
    h->out_cap = 0;
    h->out_pending = false;
    for (auto &e : h->pending) { hipEventDestroy(e.second.first); hipEventDestroy(e.second.second); }
    h->pending.clear();
    for (auto e : h->event_pool) hipEventDestroy(e);
    h->event_pool.clear();
}

static hipEvent_t get_event(psph_handle *h)
{
    if (!h->event_pool.empty()) {
        hipEvent_t e = h->event_pool.back();
        h->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

template <class F>
static void timed(psph_handle *h, int id, F &&f)
{
    if (!h->timing || (h->timing_force_only && id != KForce)) { f(); return; }
    hipEvent_t a = get_event(h), b = get_event(h);
    (void)hipEventRecord(a, h->stream);
    f();
    (void)hipEventRecord(b, h->stream);
    h->pending.push_back({id, {a, b}});
}

static void collect_timing(psph_handle *h)
{
    for (auto &e : h->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e.second.first, e.second.second) == hipSuccess) {
            h->t_ms[e.first] += ms;
            h->t_n[e.first] += 1;
        }
        h->event_pool.push_back(e.second.first);
        h->event_pool.push_back(e.second.second);
    }
    h->pending.clear();
}

static inline double norm3(const double *a) { return std::sqrt(a[0]*a[0] + a[1]*a[1] + a[2]*a[2]); }

// EOS, Functions.cpp:188-212 (host copy for InitialChecks)
static double host_eos(int EQ, double Cs0, double P00, double Density, double Density0)
{
    switch (EQ) {
    case 0: return P00+(Cs0*Cs0)*(Density-Density0);
    case 1: return P00+(Density0*Cs0*Cs0/7.0)*(std::pow(Density/Density0,7.0)-1);
    default: return (Cs0*Cs0)*Density;
    }
}

static void fill_derived(ClassParams &c)
{
    double tp = std::tan(c.phi);
    c.dp_alpha2 = tp / std::sqrt(9.0+12.0*std::tan(c.phi)*std::tan(c.phi));
    c.dp_kf2 = 3.0 * c.c / std::sqrt(9.0+12.0*std::tan(c.phi)*std::tan(c.phi));
    c.dp_alpha3 = (2.0* std::sin(c.phi)) / (std::sqrt(3.0)*(3.0-std::sin(c.phi)));
    c.dp_kf3 = (6.0*c.c*std::cos(c.phi)) / (std::sqrt(3.0)*(3.0-std::sin(c.phi)));
    c.sin_psi = std::sin(c.psi);
}

static ClassParams class_of(const psph_particle &p)
{
    ClassParams c;
    std::memset(&c, 0, sizeof c);
    c.h = p.h; c.Cs = p.Cs; c.P0 = p.P0; c.RefDensity = p.RefDensity; c.FPMassC = p.FPMassC;
    c.Alpha = p.Alpha; c.Beta = p.Beta; c.MuRef = p.MuRef; c.T0 = p.T0; c.m = p.m; c.CSmag = p.CSmag;
    c.TI = p.TI; c.TIn = p.TIn; c.TIInitDist = p.TIInitDist;
    c.G = p.G; c.K = p.K; c.Sigmay = p.Sigmay; c.c = p.c; c.phi = p.phi; c.psi = p.psi;
    c.IsFree = p.IsFree ? 1 : 0; c.NoSlip = p.NoSlip ? 1 : 0; c.Shepard = p.Shepard ? 1 : 0;
    c.LES = p.LES ? 1 : 0; c.Material = p.Material; c.PresEq = p.PresEq; c.VisM = p.VisM;
    c.Fail = p.Fail; c.ID = p.ID; c.ShepardStep = p.ShepardStep; c.InOut = 0;   // InOut: flag bits
    c.V = p.V; c.RhoF = p.RhoF; c.k = p.k; c.k2 = p.k2; c.d = p.d; c.n0 = p.n0;
    c.SeepageType = p.SeepageType;
    // a saturated soil particle belongs to the class of its unsaturated RefDensity
    if (p.Material == 3 && p.IsSat) c.RefDensity = p.RefDensity + p.RhoF;
    fill_derived(c);
    return c;
}

struct ClassLess {
    bool operator()(const ClassParams &a, const ClassParams &b) const
    { return std::memcmp(&a, &b, sizeof a) < 0; }
};

static inline void to_sym(const double *M, double *s)
{
    s[0] = M[0]; s[1] = M[1]; s[2] = M[2]; s[3] = M[4]; s[4] = M[5]; s[5] = M[8];
}

static inline void from_sym(const double *s, double *M)
{
    M[0] = s[0]; M[1] = s[1]; M[2] = s[2];
    M[3] = s[1]; M[4] = s[3]; M[5] = s[4];
    M[6] = s[2]; M[7] = s[4]; M[8] = s[5];
}

// CellInitiate, Domain.cpp:549-660.  A periodic direction takes the margin R instead of hmax/2
// and two more cell layers (the reference's aliases of the first two; reported in cell_no, not
// stored on the device).
static int cell_initiate(psph_handle *h, const psph_particle *p, size_t n)
{
    psph_grid &g = h->grid;
    std::memset(&g, 0, sizeof g);
    double TRPR[3], BLPF[3];
    std::memcpy(TRPR, h->cfg.trpr, sizeof TRPR);
    std::memcpy(BLPF, h->cfg.blpf, sizeof BLPF);
    double hmax = 0.0, rhomax = 0.0, MuMax = h->cfg.mu_max, CsMax = h->cfg.cs_max;
    if (!(norm3(TRPR) > 0.0) && !(norm3(BLPF) > 0.0)) {
        std::memcpy(BLPF, p[0].x, sizeof BLPF);
        std::memcpy(TRPR, p[0].x, sizeof TRPR);
        hmax = p[0].h;
        rhomax = p[0].Density;
        for (size_t i = 0; i < n; ++i) {
            for (int c = 0; c < 3; ++c) {
                if (p[i].x[c] > TRPR[c]) TRPR[c] = p[i].x[c];
                if (p[i].x[c] < BLPF[c]) BLPF[c] = p[i].x[c];
            }
            if (p[i].h > hmax) hmax = p[i].h;
            if (p[i].Density > rhomax) rhomax = p[i].Density;
            if (p[i].Mu > MuMax) MuMax = p[i].Mu;
            if (p[i].Cs > CsMax) CsMax = p[i].Cs;
        }
    } else {
        return fail(PSPH_ERR_UNSUPPORTED, "a preset TRPR/BLPF box (second Solve on one Domain) is not supported");
    }
    for (int c = 0; c < 3; ++c) {
        if (h->cfg.dom_max[c] > TRPR[c]) TRPR[c] = h->cfg.dom_max[c];
        if (h->cfg.dom_min[c] < BLPF[c]) BLPF[c] = h->cfg.dom_min[c];
    }
    for (int c = 0; c < 3; ++c) {
        if (!h->cfg.periodic[c]) { TRPR[c] += hmax/2; BLPF[c] -= hmax/2; }
        else { TRPR[c] += h->cfg.R; BLPF[c] -= h->cfg.R; }
    }
    double Cellfac = h->cfg.cellfac;
    int dims = h->cfg.dimension;
    int CellNo[3] = {1, 1, 1};
    double CellSize[3] = {0, 0, 0};
    for (int c = 0; c < dims; ++c) {
        double L = (TRPR[c]-BLPF[c])/(Cellfac*hmax);
        if ((double)(std::ceil(L) - L) < (hmax/10.0)) CellNo[c] = (int)std::ceil(L);
        else CellNo[c] = (int)std::floor(L);
        if (CellNo[c] < 1) return fail(PSPH_ERR_INVALID, "domain smaller than one cell in dimension %d", c);
        CellSize[c] = (TRPR[c]-BLPF[c])/CellNo[c];
        // with fewer than three cells across a period the reference's alias layers pair some
        // cells twice
        if (h->cfg.periodic[c] && CellNo[c] < 3)
            return fail(PSPH_ERR_UNSUPPORTED, "a periodic direction needs at least 3 cells across the period "
                        "(dimension %d has %d)", c, CellNo[c]);
    }
    for (int c = 0; c < 3; ++c)
        if (h->cfg.periodic[c]) CellNo[c] += 2;
    double ncells = (double)CellNo[0]*CellNo[1]*CellNo[2];
    if (ncells > 2.0e9) return fail(PSPH_ERR_CAPACITY, "too many cells (%g)", ncells);
    std::memcpy(g.blpf, BLPF, sizeof BLPF);
    std::memcpy(g.trpr, TRPR, sizeof TRPR);
    std::memcpy(g.cell_size, CellSize, sizeof CellSize);
    std::memcpy(g.cell_no, CellNo, sizeof CellNo);
    g.hmax = hmax; g.rhomax = rhomax; g.mu_max = MuMax; g.cs_max = CsMax;
    return 0;
}

// ------------------------------------------------------------------------------------------
// device allocation and state transfer
// ------------------------------------------------------------------------------------------
static int allocate(psph_handle *h, size_t n)
{
    Params &P = h->P;
    Dev A{};
    auto al = [&](void **p, size_t b) { return dev_alloc(h, p, b); };
    int rc = 0;
#define AL(ptr, bytes) do { rc = al((void **)&(ptr), (bytes)); if (rc) return rc; } while (0)
    size_t N = n ? n : 1;
    Dev &a = A;
    AL(h->ctl_dev, sizeof(Ctl));
    AL(h->cls_dev, sizeof(ClassParams)*kMaxClasses);
    // persistent, two buffers
    double4 *pos[2], *vel[2], *vh[2];
    double2 *mmu[2];
    uint32_t *flags[2];
    int32_t *orig[2];
    double *sig[2] = {nullptr, nullptr}, *sigh[2] = {nullptr, nullptr}, *tau[2] = {nullptr, nullptr},
           *tauh[2] = {nullptr, nullptr}, *eps[2] = {nullptr, nullptr}, *epsh[2] = {nullptr, nullptr};
    for (int b = 0; b < 2; ++b) {
        AL(pos[b], sizeof(double4)*N);
        AL(vel[b], sizeof(double4)*N);
        AL(vh[b], sizeof(double4)*N);
        AL(mmu[b], sizeof(double2)*N);
        AL(flags[b], sizeof(uint32_t)*N);
        AL(orig[b], sizeof(int32_t)*N);
        if (P.any_solid) {
            AL(sig[b], sizeof(double)*6*N);
            AL(sigh[b], sizeof(double)*6*N);
            AL(tau[b], sizeof(double)*6*N);
            AL(tauh[b], sizeof(double)*6*N);
            AL(eps[b], sizeof(double)*6*N);
            AL(epsh[b], sizeof(double)*6*N);
        }
    }
    AL(a.aux, sizeof(double4)*N);
    if (P.mixed32) {
        AL(a.f0, sizeof(float4)*N);
        AL(a.f1, sizeof(float4)*N);
        AL(a.f2, sizeof(float4)*N);
        AL(a.fw, sizeof(float4)*N);
    } else if (P.fast11) {
        AL(a.g0, sizeof(double4)*N);
        AL(a.g1, sizeof(double4)*N);
        AL(a.g2, sizeof(double4)*N);
        AL(a.gw, sizeof(double4)*N);
    }
    AL(a.wal, sizeof(double4)*N);
    AL(a.acc, sizeof(double4)*N);
    AL(a.xsp, sizeof(double4)*N);
    if (P.multi) {
        AL(a.fsi, sizeof(double4)*N);
        AL(a.fsk, sizeof(double)*N);
        AL(a.xflag, sizeof(int32_t)*N);
        AL(a.swz, sizeof(double2)*N);
    }
    AL(a.shep, sizeof(double2)*N);
    if (P.any_solid || P.any_sr) AL(a.sr, sizeof(double)*6*N);
    if (P.any_solid) { AL(a.rr, sizeof(double)*3*N); AL(a.tir, sizeof(double)*6*N); }
    if (P.srec) AL(a.srec, sizeof(double2)*10*N);
    if (P.any_les_t0) AL(a.shr, sizeof(double)*2*N);
    AL(a.key, sizeof(int32_t)*N);
    AL(a.rank, sizeof(int32_t)*N);
    AL(a.key_b, sizeof(int32_t)*N);
    AL(a.src, sizeof(int32_t)*N);
    AL(a.sorig, sizeof(uint32_t)*N);
    AL(a.perm, sizeof(int32_t)*N);
    AL(a.cstart, sizeof(int32_t)*(P.ncells + 1));
    AL(a.ccount, sizeof(int32_t)*(P.ncells + 1));
    AL(a.cfcnt, sizeof(int32_t)*(P.ncells + 1));
    AL(a.cfree, sizeof(int32_t)*(P.ncells + 1));
    AL(a.fstart, sizeof(int32_t)*(P.ncells + 1));
    AL(a.bsum, sizeof(unsigned long long)*(scan_blocks(P) + 1));
    AL(a.free_list, sizeof(int32_t)*N);
    AL(a.fixed_list, sizeof(int32_t)*N);
    AL(a.ncnt, sizeof(int32_t)*N);
    AL(a.ovq, sizeof(int32_t)*N);
    AL(a.ovs, sizeof(double4)*2*N);
    if (P.verlet) {
        AL(a.wcnt, sizeof(int32_t)*4*N);
        AL(a.wact, sizeof(int32_t)*N);
        AL(a.xref, sizeof(double4)*N);
    }
    AL(h->pair_cnt, sizeof(int)*N);
    AL(h->pair_off, sizeof(long long)*N);
#undef AL
    a.ctl = h->ctl_dev;
    a.cls = h->cls_dev;
    for (int b = 0; b < 2; ++b)
        a.buf[b] = Cols{pos[b], vel[b], vh[b], mmu[b], flags[b], orig[b],
                        sig[b], sigh[b], tau[b], tauh[b], eps[b], epsh[b]};
    h->D = a;
    if (!h->ctl_host) HIPCHECK(hipHostMalloc((void **)&h->ctl_host, sizeof(Ctl), hipHostMallocDefault));
    h->cap = N;
    return 0;
}

// Buffer of the persistent columns that holds the current state (control block read back):
// buf[side] between steps; during a step halted after its force pass, the step's working copy.
static int state_buffer(const psph_handle *h)
{
    const Ctl &C = *h->ctl_host;
    return C.mid_step ? (C.side ^ C.rebuild) : C.side;
}

// The device buffers with the persistent columns of buffer b as the current ones (host reads).
static Dev with_state(const psph_handle *h, int b)
{
    Dev D = h->D;
    const Cols &S = h->D.buf[b];
    D.pos = S.pos; D.vel = S.vel; D.vh = S.vh; D.mmu = S.mmu; D.flags = S.flags; D.orig = S.orig;
    D.sig = S.sig; D.sigh = S.sigh; D.tau = S.tau; D.tauh = S.tauh; D.eps = S.eps; D.epsh = S.epsh;
    return D;
}

// Writes particles (reference order) into persistent buffer `buf`; classes must be built.
static int write_state(psph_handle *h, const psph_particle *p, size_t n, int buf)
{
    Params &P = h->P;
    std::vector<double4> pos(n), vel(n), vh(n);
    std::vector<double2> mmu(n);
    std::vector<uint32_t> flags(n);
    std::vector<int32_t> orig(n);
    std::vector<double> sig, sigh, tau, tauh, eps, epsh;
    if (P.any_solid) { sig.resize(6*n); sigh.resize(6*n); tau.resize(6*n); tauh.resize(6*n); eps.resize(6*n); epsh.resize(6*n); }
    std::map<ClassParams, int, ClassLess> idx;
    for (size_t k = 0; k < h->classes.size(); ++k) idx[h->classes[k]] = (int)k;
    for (size_t i = 0; i < n; ++i) {
        const psph_particle &q = p[i];
        ClassParams c = class_of(q);
        auto it = idx.find(c);
        if (it == idx.end()) {
            // a parameter set new since upload (a particle the in/out-flow recycling created)
            if ((int)h->classes.size() >= kMaxClasses)
                return fail(PSPH_ERR_CAPACITY, "more than %d distinct particle parameter sets", kMaxClasses);
            h->classes.push_back(c);
            HIPCHECK(hipMemcpy(h->cls_dev, h->classes.data(), sizeof(ClassParams)*h->classes.size(), hipMemcpyHostToDevice));
            P.nclasses = (int)h->classes.size();
            it = idx.emplace(c, (int)h->classes.size() - 1).first;
        }
        int ci = it->second;
        pos[i] = make_double4(q.x[0], q.x[1], q.x[2], q.Density);
        vel[i] = make_double4(q.v[0], q.v[1], q.v[2], q.Pressure);
        bool lf = (P.scheme == 1) && q.IsFree;
        if (lf) vh[i] = make_double4(q.va[0], q.va[1], q.va[2], q.Densitya);
        else vh[i] = make_double4(q.vb[0], q.vb[1], q.vb[2], q.Densityb);
        mmu[i] = make_double2(q.Mass, q.Mu);
        flags[i] = (uint32_t)ci | (q.IsFree ? kFree : 0u) | (q.NoSlip ? kNoSlip : 0u) |
                   ((uint32_t)(q.Material & 3) << kMatShift);
        flags[i] |= ((uint32_t)(q.InOut & 3)) << kInOutShift;
        if (q.Material == 3 && q.IsSat) flags[i] |= kSat;
        orig[i] = (int32_t)i;
        if (P.any_solid) {
            to_sym(q.Sigma, &sig[6*i]);
            to_sym(P.scheme == 1 ? q.Sigmaa : q.Sigmab, &sigh[6*i]);
            to_sym(q.ShearStress, &tau[6*i]);
            to_sym(P.scheme == 1 ? q.ShearStressa : q.ShearStressb, &tauh[6*i]);
            to_sym(q.Strain, &eps[6*i]);
            to_sym(P.scheme == 1 ? q.Straina : q.Strainb, &epsh[6*i]);
        }
    }
    const Cols &D = h->D.buf[buf];
    // blocking copies on the null stream do not wait for the (non-blocking) solver stream
    HIPCHECK(hipStreamSynchronize(h->stream));
    HIPCHECK(hipMemcpy(D.pos, pos.data(), sizeof(double4)*n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.vel, vel.data(), sizeof(double4)*n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.vh, vh.data(), sizeof(double4)*n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.mmu, mmu.data(), sizeof(double2)*n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.flags, flags.data(), sizeof(uint32_t)*n, hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(D.orig, orig.data(), sizeof(int32_t)*n, hipMemcpyHostToDevice));
    if (P.any_solid) {
        HIPCHECK(hipMemcpy(D.sig, sig.data(), sizeof(double)*6*n, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(D.sigh, sigh.data(), sizeof(double)*6*n, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(D.tau, tau.data(), sizeof(double)*6*n, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(D.tauh, tauh.data(), sizeof(double)*6*n, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(D.eps, eps.data(), sizeof(double)*6*n, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(D.epsh, epsh.data(), sizeof(double)*6*n, hipMemcpyHostToDevice));
    }
    return 0;
}

// Reads the current device state into h->host (reference order by the orig column).
// After a completed step the transient columns share the current order.
// The fixed particles' pair sums of a step stopped after its force pass (outputs and the
// GeneralAfter hook), formed once per stop: they feed nothing in the dynamics, so only a state
// the host reads carries them.  The control block on the host must be current.
static int form_fixed_sums(psph_handle *h)
{
    if (h->fixed_ok || !h->trans_ok || h->ctl_host->mid_step != 1) return 0;
    launch_fixed_sums(h->D, h->P, h->stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(h->stream));
    h->fixed_ok = true;
    return 0;
}

static int read_state(psph_handle *h)
{
    Params &P = h->P;
    size_t n = (size_t)P.n;
    HIPCHECK(hipStreamSynchronize(h->stream));
    HIPCHECK(hipMemcpy(h->ctl_host, h->ctl_dev, sizeof(Ctl), hipMemcpyDeviceToHost));
    {
        int rc0 = form_fixed_sums(h);
        if (rc0) return rc0;
    }
    const Dev D = with_state(h, state_buffer(h));
    std::vector<double4> pos(n), vel(n), vh(n), aux(n), wal(n), acc(n), xsp(n);
    std::vector<double2> mmu(n), shep(n);
    std::vector<uint32_t> flags(n);
    std::vector<int32_t> orig(n);
    HIPCHECK(hipMemcpy(pos.data(), D.pos, sizeof(double4)*n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(vel.data(), D.vel, sizeof(double4)*n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(vh.data(), D.vh, sizeof(double4)*n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(mmu.data(), D.mmu, sizeof(double2)*n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(flags.data(), D.flags, sizeof(uint32_t)*n, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(orig.data(), D.orig, sizeof(int32_t)*n, hipMemcpyDeviceToHost));
    bool have_transients = h->trans_ok;
    std::vector<double4> fsi;
    std::vector<double> fsk;
    if (have_transients && P.multi) {
        fsi.resize(n); fsk.resize(n);
        HIPCHECK(hipMemcpy(fsi.data(), D.fsi, sizeof(double4)*n, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(fsk.data(), D.fsk, sizeof(double)*n, hipMemcpyDeviceToHost));
    }
    if (have_transients) {
        HIPCHECK(hipMemcpy(wal.data(), D.wal, sizeof(double4)*n, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(acc.data(), D.acc, sizeof(double4)*n, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(xsp.data(), D.xsp, sizeof(double4)*n, hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(shep.data(), D.shep, sizeof(double2)*n, hipMemcpyDeviceToHost));
    }
    std::vector<double> sig, sigh, tau, tauh, eps, epsh, sr, rr, tir, shr;
    auto rd = [&](std::vector<double> &v, const double *src, size_t k) -> int {
        v.resize(k*n);
        if (src) HIPCHECK(hipMemcpy(v.data(), src, sizeof(double)*k*n, hipMemcpyDeviceToHost));
        return 0;
    };
    if (P.any_solid) {
        if (rd(sig, D.sig, 6) || rd(sigh, D.sigh, 6) || rd(tau, D.tau, 6) || rd(tauh, D.tauh, 6) ||
            rd(eps, D.eps, 6) || rd(epsh, D.epsh, 6))
            return PSPH_ERR_HIP;
    }
    if (have_transients) {
        if ((P.any_solid || P.any_sr) && rd(sr, D.sr, 6)) return PSPH_ERR_HIP;
        if (P.any_solid && (rd(rr, D.rr, 3) || rd(tir, D.tir, 6))) return PSPH_ERR_HIP;
        if (P.any_les_t0 && rd(shr, D.shr, 2)) return PSPH_ERR_HIP;
    }
    bool shep_on = P.any_shepard;
    // stopped before the wall pass (GeneralBefore hook): the sums are those StartAcceleration
    // leaves (Domain.cpp:947-1017), and a free particle's a is where its pair sums will start
    const bool before_walls = h->ctl_host->mid_step == 2;
    const bool acc_init = h->ctl_host->acc_init != 0;
    if (before_walls && acc_init && !have_transients)
        HIPCHECK(hipMemcpy(acc.data(), D.acc, sizeof(double4)*n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) {
        int o = orig[i];
        if (o < 0 || (size_t)o >= h->host.size()) return fail(PSPH_ERR_STATE, "corrupt particle index %d", o);
        psph_particle &q = h->host[o];
        bool fr = flag_free(flags[i]);
        q.x[0] = pos[i].x; q.x[1] = pos[i].y; q.x[2] = pos[i].z; q.Density = pos[i].w;
        q.v[0] = vel[i].x; q.v[1] = vel[i].y; q.v[2] = vel[i].z; q.Pressure = vel[i].w;
        if (P.scheme == 1 && fr) { q.va[0] = vh[i].x; q.va[1] = vh[i].y; q.va[2] = vh[i].z; q.Densitya = vh[i].w; }
        else { q.vb[0] = vh[i].x; q.vb[1] = vh[i].y; q.vb[2] = vh[i].z; q.Densityb = vh[i].w; }
        q.Mass = mmu[i].x; q.Mu = mmu[i].y;
        q.InOut = flag_inout(flags[i]);
        if (P.multi) {
            const ClassParams &cl = h->classes[flag_cls(flags[i])];
            q.IsSat = (flags[i] & kSat) ? 1 : 0;
            q.SatCheck = (flags[i] & kSatCheck) ? 1 : 0;
            if (q.Material == 3) q.RefDensity = q.IsSat ? cl.RefDensity - cl.RhoF : cl.RefDensity;
            if (have_transients) {
                q.FSIPressure = fsi[i].w; q.FSISumKernel = fsk[i];
                q.FSINSv[0] = fsi[i].x; q.FSINSv[1] = fsi[i].y; q.FSINSv[2] = fsi[i].z;
            }
        }
        if (have_transients) {
            if (fr) {
                q.a[0] = acc[i].x; q.a[1] = acc[i].y; q.a[2] = acc[i].z; q.dDensity = acc[i].w;
                q.SumKernel = 0.0;
                if (P.any_xsph) { q.VXSPH[0] = xsp[i].x; q.VXSPH[1] = xsp[i].y; q.VXSPH[2] = xsp[i].z; }
                else { q.VXSPH[0] = q.VXSPH[1] = q.VXSPH[2] = 0.0; }
                if (shep_on) { q.ZWab = xsp[i].w; q.SumDen = shep[i].x; }
            } else if (h->fixed_ok) {
                // stopped after the force pass: the pair sums of the fixed particle (k_fixed_sums)
                q.a[0] = acc[i].x; q.a[1] = acc[i].y; q.a[2] = acc[i].z; q.dDensity = acc[i].w;
                q.NSv[0] = wal[i].x; q.NSv[1] = wal[i].y; q.NSv[2] = wal[i].z;
                q.SumKernel = shep[i].y;
            } else {
                // between steps the pair sums of fixed particles are not held; a stays the
                // gravity reset
                for (int c = 0; c < 3; ++c) q.a[c] = h->cfg.gravity[c];
                q.dDensity = 0.0;
                q.NSv[0] = wal[i].x; q.NSv[1] = wal[i].y; q.NSv[2] = wal[i].z;
                q.SumKernel = shep[i].y;
            }
            if (!sr.empty()) { from_sym(&sr[6*i], q.StrainRate); }
            if (!rr.empty()) {
                double *R = q.RotationRate;
                R[0] = 0.0; R[1] = rr[3*i]; R[2] = rr[3*i+1];
                R[3] = -rr[3*i]; R[4] = 0.0; R[5] = rr[3*i+2];
                R[6] = -rr[3*i+1]; R[7] = -rr[3*i+2]; R[8] = 0.0;
            }
            if (!tir.empty()) from_sym(&tir[6*i], q.TIR);
            if (!shr.empty() && fr) { q.ShearRate = shr[2*i]; q.SBar = shr[2*i+1]; }
        }
        if (P.any_solid) {
            from_sym(&sig[6*i], q.Sigma);
            from_sym(&tau[6*i], q.ShearStress);
            from_sym(&eps[6*i], q.Strain);
            if (P.scheme == 1) { from_sym(&sigh[6*i], q.Sigmaa); from_sym(&tauh[6*i], q.ShearStressa); from_sym(&epsh[6*i], q.Straina); }
            else { from_sym(&sigh[6*i], q.Sigmab); from_sym(&tauh[6*i], q.ShearStressb); from_sym(&epsh[6*i], q.Strainb); }
        }
        else if (!fr && h->ctl_host->reorders > 0) {
            // StartAcceleration zeroes a fixed particle's Sigma and ShearStress (Domain.cpp:991-992);
            // fluid runs keep no tensors on the device
            for (int k = 0; k < 9; ++k) { q.Sigma[k] = 0.0; q.ShearStress[k] = 0.0; }
        }
        if (fr) {
            q.FirstStep = h->ctl_host->first_step;
            if (P.scheme == 0) { q.ct = h->ctl_host->ct; q.ShepardCounter = h->ctl_host->shep_counter; }
        }
        if (before_walls) {
            for (int c = 0; c < 3; ++c) {
                q.a[c] = fr ? (acc_init ? (&acc[i].x)[c] : P.afree[c]) : h->cfg.gravity[c];
                q.VXSPH[c] = 0.0; q.FSINSv[c] = 0.0;
            }
            q.dDensity = (fr && acc_init) ? acc[i].w : 0.0;
            q.ZWab = 0.0; q.SumDen = 0.0; q.SumKernel = 0.0; q.S = 0.0;
            q.FSISumKernel = 0.0; q.FSIPressure = 0.0; q.SatCheck = 0;
            for (int k = 0; k < 9; ++k) { q.StrainRate[k] = 0.0; q.RotationRate[k] = 0.0; }
            if (!fr) {
                q.Pressure = 0.0;
                for (int c = 0; c < 3; ++c) q.NSv[c] = 0.0;
                for (int k = 0; k < 9; ++k) { q.Sigma[k] = 0.0; q.ShearStress[k] = 0.0; }
            }
        }
    }
    return 0;
}

// ------------------------------------------------------------------------------------------
// step queueing
// ------------------------------------------------------------------------------------------
// Queues one step.  from = 0: a whole step; 1: the rest of a step stopped before the wall pass
// (GeneralBefore hook stop); 2: the rest of a step stopped after the force pass (Move).
// Queues one step from stage `from` (0: whole step; 1: from the wall pass; 2: Move and tail).
// lean: list reuse, and the step is known not to rebuild, so the sort and list-build kernels
// (no-ops on such steps) are not queued.  tail_lean: the step's tail halts (kHaltSort) when the
// step after it must rebuild, because that step will be queued lean.  fresh: the step before
// it (in the same queue) wrote this lean step's views, so its k_reorder is not queued.
// views_next: this step's k_integrate writes the views of the step queued after it.
static void enqueue_step(psph_handle *h, int from, bool lean, bool tail_lean, bool fresh, bool views_next)
{
    Params P = h->P;
    const Dev &D = h->D;
    P.lean = lean ? 1 : 0;
    P.views_next = views_next ? 1 : 0;
    if (from == 0) {
        if (!lean) {
            timed(h, KScan, [&] {
                if (P.verlet) launch_keys(D, P, h->stream);      // only on steps that rebuild
                launch_scan(D, P, h->stream);
            });
            timed(h, KScatter, [&] { launch_scatter(D, P, h->stream); });
            timed(h, KCellsort, [&] { launch_cellsort(D, P, h->stream); });
        }
        if (!(lean && fresh)) timed(h, KReorder, [&] { launch_reorder(D, P, h->stream); });
    }
    if (from <= 1) {
        timed(h, KWall, [&] { launch_wall(D, P, h->stream, from); });
        timed(h, KForce, [&] { launch_force(D, P, h->stream); });
    }
    timed(h, KIntegrate, [&] { launch_integrate(D, P, h->stream); });
    P.lean = tail_lean ? 1 : 0;
    timed(h, KTail, [&] { launch_tail(D, P, h->stream); });
}

static int sync_ctl(psph_handle *h)
{
    HIPCHECK(hipMemcpyAsync(h->ctl_host, h->ctl_dev, sizeof(Ctl), hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    if (h->timing) collect_timing(h);
    return 0;
}

static int push_ctl(psph_handle *h)
{
    HIPCHECK(hipMemcpyAsync(h->ctl_dev, h->ctl_host, sizeof(Ctl), hipMemcpyHostToDevice, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    return 0;
}

// Readies the next step after the host has written or re-sorted the state held by variant d:
// pushes the control block and computes the cell keys and arrival ranks for the next sort; with
// list reuse the next step rebuilds and its k_keys computes them.
// With list reuse: makes the next step sort and rebuild the lists.  A step pending its Move
// ends with the tail, which turns need_rebuild into rebuild for the step after it.
static void request_rebuild(psph_handle *h)
{
    if (!h->P.verlet) return;
    if (h->ctl_host->mid_step) h->ctl_host->need_rebuild = 1;
    else h->ctl_host->rebuild = 1;
}

static int prime_first_step(psph_handle *h)
{
    request_rebuild(h);
    int rc = push_ctl(h);
    if (rc) return rc;
    if (!h->P.verlet) launch_keys(h->D, h->P, h->stream);
    HIPCHECK(hipGetLastError());
    return 0;
}

// Order-preserving deletion of the particles outside [BLPF, TRPR] (CheckParticleLeave with
// Array::DelItems, External/array.h:726-738); the survivors are re-uploaded.

// ------------------------------------------------------------------------------------------
// in/out-flow boundaries (Domain.cpp:1327-1635) on the host mirror: the inflow list and its
// order, and the recycling of particles that left the box
// ------------------------------------------------------------------------------------------

// ListGenerate's cell of a particle (Domain.cpp:662-784), in the reference's CellNo grid
static void ref_cell(const psph_handle *h, const psph_particle &q, int ijk[3])
{
    const psph_grid &g = h->grid;
    ijk[0] = ijk[1] = ijk[2] = 0;
    for (int c = 0; c < h->cfg.dimension; ++c) {
        int i = (int)(std::floor((q.x[c] - g.blpf[c]) / g.cell_size[c]));
        if (i < 0) { if ((g.blpf[c] - q.x[c]) <= g.hmax) i = 0; }
        if (i >= g.cell_no[c]) { if ((q.x[c] - g.trpr[c]) <= g.hmax) i = g.cell_no[c]-1; }
        ijk[c] = std::min(std::max(i, 0), g.cell_no[c]-1);
    }
}

// The particles the reference's inflow scan meets (Domain.cpp:1514-1528, 1549-1563), in its
// order: cell rows y outer, z middle, x from 0 to x1 inner, each cell's chain from the largest
// index down
static std::vector<int32_t> inflow_scan(const psph_handle *h, int x1, bool need_flag)
{
    struct E { int q2, q3, q1, idx; };
    std::vector<E> e;
    const psph_grid &g = h->grid;
    const double loc1 = h->P.loc1;
    for (size_t k = 0; k < h->host.size(); ++k) {
        const psph_particle &q = h->host[k];
        if (!q.IsFree || !(q.x[0] <= loc1) || (need_flag && q.InOut != 1)) continue;
        int ijk[3];
        ref_cell(h, q, ijk);
        if (ijk[0] > x1) continue;
        if (h->cfg.periodic[1] && ijk[1] >= g.cell_no[1]-2) continue;
        if (h->cfg.periodic[2] && ijk[2] >= g.cell_no[2]-2) continue;
        e.push_back({ijk[1], ijk[2], ijk[0], (int)k});
    }
    std::sort(e.begin(), e.end(), [](const E &a, const E &b) {
        if (a.q2 != b.q2) return a.q2 < b.q2;
        if (a.q3 != b.q3) return a.q3 < b.q3;
        if (a.q1 != b.q1) return a.q1 < b.q1;
        return a.idx > b.idx;
    });
    std::vector<int32_t> out;
    out.reserve(e.size());
    for (auto &x : e) out.push_back(x.idx);
    return out;
}

// InFlowBCFresh's first call (Domain.cpp:1505-1542): the inflow zone and list, the outflow line
static int inflow_init(psph_handle *h)
{
    Params &P = h->P;
    const psph_config &c = h->cfg;
    const psph_grid &g = h->grid;
    h->inpart.clear();
    if (P.inout == 1 || P.inout == 3) {
        if (!(std::sqrt(c.in_v[0]*c.in_v[0] + c.in_v[1]*c.in_v[1] + c.in_v[2]*c.in_v[2]) > 0.0) || !(c.in_density > 0.0))
            return fail(PSPH_ERR_INVALID, "BC.inv or BC.inDensity are not defined");
        P.loc1 = g.blpf[0] + c.bc_cellfac*g.hmax;
        int x1 = (int)(std::floor((P.loc1 - g.blpf[0]) / g.cell_size[0]));
        h->inpart = inflow_scan(h, x1, false);
        if (h->inpart.empty()) return fail(PSPH_ERR_INVALID, "no free particle in the inflow zone");
        for (int32_t k : h->inpart) h->host[k].InOut = 1;
        h->loc2 = h->loc3 = h->host[h->inpart[0]].x[0];
        for (int32_t k : h->inpart) {
            double x0 = h->host[k].x[0];
            if (x0 < h->loc2) h->loc2 = x0;
            if (x0 > h->loc3) h->loc3 = x0;
        }
    }
    if (P.inout == 2 || P.inout == 3) P.out_loc = g.trpr[0] - c.bc_cellfac*g.hmax;
    h->inoutcounter = 2;
    P.n_in = (int)h->inpart.size();
    return 0;
}

// The fields an inflow particle passes on (Domain.cpp:1366-1401)
static void inflow_copy(psph_particle &a, const psph_particle &b, const double *x)
{
    for (int k = 0; k < 3; ++k) a.x[k] = x[k];
    a.Material = 1; a.InOut = 1; a.FirstStep = 0;
    a.P0 = b.P0; a.PresEq = b.PresEq; a.Cs = b.Cs;
    a.Alpha = b.Alpha; a.Beta = b.Beta; a.Mu = b.Mu; a.MuRef = b.MuRef; a.T0 = 0.0;
    a.Mass = b.Mass; a.h = b.h; a.ID = b.ID; a.TI = 0.0; a.RefDensity = b.RefDensity;
    a.ct = b.ct; a.Shepard = b.Shepard; a.ShepardStep = b.ShepardStep;
    a.ShepardCounter = b.ShepardCounter; a.LES = b.LES; a.CSmag = b.CSmag;
}

static int write_back(psph_handle *h);

// Re-creates every device buffer for `cap` particles, keeping the run's settings, list
// capacities, class table and control block; the state itself is on the host (read back) and
// the caller writes it (write_back).  The pinned buffers stay: an output's host copy may still
// be in use by the caller.
static int grow_buffers(psph_handle *h, size_t cap)
{
    HIPCHECK(hipStreamSynchronize(h->stream));
    const Ctl saved = *h->ctl_host;
    const int nl_cap = h->P.nl_cap, wl_cap = h->P.wl_cap;
    for (void *p : h->allocs) if (p) HIPCHECK(hipFree(p));
    h->allocs.clear();
    h->nl = h->wl = nullptr;
    h->out_dev = nullptr;
    h->hook_dev = nullptr;
    h->D.tile = nullptr; h->D.items = nullptr; h->D.tl = nullptr;
    h->hook_cap = 0;
    int rc = allocate(h, cap);
    if (rc) return rc;
    h->P.nl_cap = 0;
    h->P.wl_cap = 0;
    if (nl_cap > 0 && (rc = nl_alloc(h, nl_cap))) return rc;
    if (wl_cap > 0 && (rc = wl_alloc(h, wl_cap))) return rc;
    HIPCHECK(hipMemcpy(h->cls_dev, h->classes.data(), sizeof(ClassParams)*h->classes.size(), hipMemcpyHostToDevice));
    *h->ctl_host = saved;
    return 0;
}

// InFlowBCLeave (Domain.cpp:1327-1496) at a halt of the device: particles that left the box are
// recycled into the inflow positions the inflow particles past InFlowLoc1 leave behind
static int inflow_leave(psph_handle *h, long long *ndel)
{
    int rc = read_state(h);
    if (rc) return rc;
    const double *T = h->grid.trpr, *B = h->grid.blpf;
    const Params &P = h->P;
    std::vector<psph_particle> &H = h->host;
    std::vector<int32_t> delp, addb;
    std::vector<double> addx;
    for (size_t k = 0; k < H.size(); ++k) {
        psph_particle &q = H[k];
        if ((q.x[0] > T[0]) || (q.x[1] > T[1]) || (q.x[2] > T[2]) ||
            (q.x[0] < B[0]) || (q.x[1] < B[1]) || (q.x[2] < B[2])) {
            q.InOut = 0;
            delp.push_back((int32_t)k);
        }
    }
    if (P.inout == 1 || P.inout == 3) {
        std::vector<int32_t> keep;
        for (int32_t k : h->inpart) {
            psph_particle &q = H[k];
            if (q.x[0] > P.loc1) {
                double t[3] = {q.x[0], q.x[1], q.x[2]};
                t[0] -= (h->loc3 - h->loc2 + h->cfg.initial_dist);
                q.InOut = 0; q.FirstStep = 0; q.ShepardCounter = 0;
                addx.insert(addx.end(), t, t + 3);
                addb.push_back(k);
            } else {
                keep.push_back(k);
            }
        }
        h->inpart.swap(keep);
    }
    const size_t nd = delp.size(), na = addb.size();
    if (na >= nd) {
        for (size_t i = 0; i < nd; ++i) {
            inflow_copy(H[delp[i]], H[addb[i]], &addx[3*i]);
            h->inpart.push_back(delp[i]);
        }
        for (size_t i = nd; i < na; ++i) {
            // Particle::Particle (Particle.cpp:25-117) with the inflow particle's ID, v, Mass,
            // RefDensity and h
            const psph_particle b = H[addb[i]];
            psph_particle a;
            std::memset(&a, 0, sizeof a);
            for (int k = 0; k < 3; ++k) a.v[k] = b.v[k];
            a.TIn = 4.0; a.Density = b.RefDensity; a.RefDensity = b.RefDensity; a.Mass = b.Mass;
            a.FPMassC = 1.0; a.IsFree = 1; a.h = b.h; a.ID = b.ID; a.m = 300.0;
            a.FirstStep = 1; a.V = a.Mass/a.RefDensity; a.ShepardStep = 40; a.CSmag = 0.17;
            inflow_copy(a, b, &addx[3*i]);
            H.push_back(a);
            h->upload_index.push_back(h->next_index++);
            h->inpart.push_back((int32_t)H.size() - 1);
        }
    } else {
        for (size_t i = 0; i < na; ++i) {
            inflow_copy(H[delp[i]], H[addb[i]], &addx[3*i]);
            h->inpart.push_back(delp[i]);
        }
        // Particles.DelItems of the rest, order kept (External/array.h:726-738)
        std::vector<char> gone(H.size(), 0);
        for (size_t i = na; i < nd; ++i) gone[delp[i]] = 1;
        std::vector<psph_particle> keep;
        std::vector<int32_t> keep_idx;
        for (size_t k = 0; k < H.size(); ++k) {
            if (gone[k]) { ++*ndel; continue; }
            keep.push_back(H[k]);
            keep_idx.push_back(h->upload_index[k]);
        }
        H.swap(keep);
        h->upload_index.swap(keep_idx);
        h->inoutcounter = 1;
    }
    if (H.size() > h->cap) {
        // the inflow outpaces the outflow: room for half as many particles again (the
        // reference's Particles array simply grows)
        rc = grow_buffers(h, H.size() + H.size()/2 + 256);
        if (rc) return rc;
        ++h->regrowths;
    }
    if (h->inoutcounter == 1) {
        // InFlowBCFresh's rebuild of the inflow list (Domain.cpp:1545-1566), from the cells of
        // these positions, which the next step starts from
        int x1 = (int)(std::floor((h->P.loc1 - h->grid.blpf[0]) / h->grid.cell_size[0]));
        h->inpart = inflow_scan(h, x1, true);
        h->inoutcounter = 2;
    }
    h->P.n_in = (int)h->inpart.size();
    return write_back(h);
}

static int delete_leavers(psph_handle *h, long long *ndel)
{
    int rc = read_state(h);
    if (rc) return rc;
    const double *T = h->grid.trpr, *B = h->grid.blpf;
    std::vector<psph_particle> keep;
    std::vector<int32_t> keep_idx;
    keep.reserve(h->host.size());
    long long del = 0;
    for (size_t k = 0; k < h->host.size(); ++k) {
        const psph_particle &q = h->host[k];
        if ((q.x[0] > T[0]) || (q.x[1] > T[1]) || (q.x[2] > T[2]) ||
            (q.x[0] < B[0]) || (q.x[1] < B[1]) || (q.x[2] < B[2])) { ++del; continue; }
        keep.push_back(q);
        keep_idx.push_back(h->upload_index[k]);
    }
    h->host.swap(keep);
    h->upload_index.swap(keep_idx);
    *ndel = del;
    return write_back(h);
}

// The host mirror replaces the device state (after deletions or recycling)
static int write_back(psph_handle *h)
{
    int rc = 0;
    Params &P = h->P;
    P.n = (int)h->host.size();
    int nf = 0;
    for (auto &q : h->host) nf += q.IsFree ? 1 : 0;
    const bool grew = nf > P.nfree;
    P.nfree = nf;
    P.nfixed = P.n - nf;
    if (grew && P.nl_cap > 0) {                 // lists are per free target
        rc = nl_alloc(h, P.nl_cap);
        if (rc) return rc;
    }
    rc = write_state(h, h->host.data(), h->host.size(), h->ctl_host->side);
    if (rc) return rc;
    HIPCHECK(hipMemsetAsync(h->D.ccount, 0, sizeof(int32_t)*(P.ncells + 1), h->stream));
    HIPCHECK(hipMemsetAsync(h->D.cfcnt, 0, sizeof(int32_t)*(P.ncells + 1), h->stream));
    h->ctl_host->leave = 0;
    h->ctl_host->halt = kRun;
    h->trans_ok = false;
    h->fixed_ok = false;
    return prime_first_step(h);
}

// ------------------------------------------------------------------------------------------
// C ABI
// ------------------------------------------------------------------------------------------
extern "C" {

int psph_version(void) { return 102; }

int psph_last_error(char *buf, size_t len)
{
    if (buf && len) {
        std::snprintf(buf, len, "%s", g_err.c_str());
    }
    return (int)g_err.size();
}

int psph_create(const psph_config *cfg, psph_handle **out)
{
    if (!cfg || !out) return fail(PSPH_ERR_INVALID, "null argument");
    if (cfg->dimension != 2 && cfg->dimension != 3)
        return fail(PSPH_ERR_INVALID, "Please correct the dimension (2=>2D or 3=>3D)");
    if (cfg->kernel_type < 0 || cfg->kernel_type > 2) return fail(PSPH_ERR_INVALID, "Kernel Type No is out of range");
    if (cfg->visc_eq < 0 || cfg->visc_eq > 3) return fail(PSPH_ERR_INVALID, "Viscosity Equation No is out of range");
    if (cfg->gradient_type < 0 || cfg->gradient_type > 1) return fail(PSPH_ERR_INVALID, "Gradient type out of range");
    if (cfg->scheme < 0 || cfg->scheme > 1) return fail(PSPH_ERR_INVALID, "Scheme must be 0 (modified Verlet) or 1 (Leapfrog)");
    if (cfg->precision != PSPH_FP64 && cfg->precision != PSPH_FP32_MIXED)
        return fail(PSPH_ERR_INVALID, "precision must be PSPH_FP64 or PSPH_FP32_MIXED");
    if (cfg->dimension == 2 && cfg->periodic[2])
        return fail(PSPH_ERR_UNSUPPORTED, "a periodic z direction in a 2-D run is outside this path");
    if (cfg->in_out_flow < 0 || cfg->in_out_flow > 3)
        return fail(PSPH_ERR_INVALID, "InOutFlow %d out of range", cfg->in_out_flow);
    if (cfg->in_out_flow > 0 && cfg->periodic[0])
        return fail(PSPH_ERR_INVALID, "Periodic BC in the X direction cannot be used with In/Out-Flow BC simultaneously");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(PSPH_ERR_NO_DEVICE, "no HIP device visible");
    psph_handle *h = new psph_handle();
    h->cfg = *cfg;
    hipGetDevice(&h->device);
    if (device_scratch(h->device, &h->pair_scratch)) h->pair_scratch_err = g_err;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return fail(PSPH_ERR_HIP, "hipStreamCreate failed");
    }
    *out = h;
    return 0;
}

int psph_upload(psph_handle *h, const psph_particle *p, size_t n, double time, double dt)
{
    if (!h || !p) return fail(PSPH_ERR_INVALID, "null argument");
    if (n == 0) return fail(PSPH_ERR_INVALID, "no particles");
    if (n > (size_t)0x7fffffff) return fail(PSPH_ERR_CAPACITY, "too many particles");
    free_all(h);
    h->uploaded = false;
    h->broken = false;
    h->host.assign(p, p + n);
    h->upload_index.resize(n);
    for (size_t i = 0; i < n; ++i) h->upload_index[i] = (int32_t)i;
    // InitialChecks, Domain.cpp:1662-1719: EOS pressure of fluid and solid particles
    for (auto &q : h->host) {
        if (q.Material < 3) q.Pressure = host_eos(q.PresEq, q.Cs, q.P0, q.Density, q.RefDensity);
        // permeability of soil particles (Domain.cpp:1683-1716)
        if (q.Material == 3) {
            double n0 = q.n0, dd = q.d;
            switch (q.SeepageType) {
            case 0: break;
            case 1: q.k = n0*n0*n0*dd*dd/(180.0*(1.0-n0)*(1.0-n0)); break;
            case 2: q.k = n0*n0*n0*dd*dd/(150.0*(1.0-n0)*(1.0-n0)); q.k2 = 1.75*(1.0-n0)/(n0*n0*n0*dd); break;
            case 3: q.k = n0*n0*n0*dd*dd/(150.0*(1.0-n0)*(1.0-n0)); q.k2 = 0.4/(n0*n0*dd); break;
            default: return fail(PSPH_ERR_INVALID, "Seepage Type No is out of range");
            }
            q.n = q.n0;
        }
    }
    int rc = cell_initiate(h, h->host.data(), n);
    if (rc) return rc;
    // TimestepCheck, Domain.cpp:1721-1733
    {
        double t1 = 0.25*h->grid.hmax/(h->grid.cs_max), t2;
        if (h->grid.mu_max > 0.0) t2 = 0.125*h->grid.hmax*h->grid.hmax*h->grid.rhomax/h->grid.mu_max; else t2 = 1000000.0;
        if (dt > std::min(t1, t2))
            return fail(PSPH_ERR_CFL, "Please decrease the time step to the allowable range (dt=%g > %g)", dt, std::min(t1, t2));
    }
    // classes and uniformity checks
    std::map<ClassParams, int, ClassLess> idx;
    h->classes.clear();
    int mat0 = h->host[0].Material;
    bool have_free = false;
    psph_particle ref{};
    int nfree = 0;
    bool mats[4] = {false, false, false, false}, free_m[4] = {false, false, false, false};
    psph_particle ref_m[4];
    for (size_t i = 0; i < n; ++i) {
        const psph_particle &q = h->host[i];
        if (q.Material < 1 || q.Material > 3)
            return fail(PSPH_ERR_INVALID, "Material type %d out of range (1 fluid, 2 solid, 3 soil)", q.Material);
        mats[q.Material] = true;
        if (q.PresEq < 0 || q.PresEq > 2) return fail(PSPH_ERR_INVALID, "Pressure Equation No %d out of range", q.PresEq);
        if (q.Material == 3 && q.Fail > 1 && q.Fail != 2 && q.Fail != 3) return fail(PSPH_ERR_INVALID, "Failure Type %d out of range", q.Fail);
        if (q.Material == 2 && q.Fail > 1) return fail(PSPH_ERR_INVALID, "Undefined failure criteria for solids");
        if (q.Material == 1 && q.T0 > 0.0 && (q.VisM < 0 || q.VisM > 1)) return fail(PSPH_ERR_INVALID, "Non-Newtonian Viscosity Type out of range");
        ClassParams c = class_of(q);
        if (!idx.count(c)) {
            if ((int)h->classes.size() >= kMaxClasses)
                return fail(PSPH_ERR_CAPACITY, "more than %d distinct particle parameter sets", kMaxClasses);
            idx[c] = (int)h->classes.size();
            h->classes.push_back(c);
        }
        if (q.IsFree) {
            ++nfree;
            const int m = q.Material;
            const double r0 = (m == 3 && q.IsSat) ? q.RefDensity + q.RhoF : q.RefDensity;
            if (!free_m[m]) { ref_m[m] = q; ref_m[m].RefDensity = r0; free_m[m] = true; }
            else if (q.PresEq != ref_m[m].PresEq || q.Cs != ref_m[m].Cs || q.P0 != ref_m[m].P0 ||
                     r0 != ref_m[m].RefDensity)
                return fail(PSPH_ERR_UNSUPPORTED, "free particles of one material with different equations of state");
            if (!have_free) { ref = q; have_free = true; }
            else {
                if (q.FirstStep != ref.FirstStep || q.ct != ref.ct || q.ShepardCounter != ref.ShepardCounter ||
                    q.ShepardStep != ref.ShepardStep)
                    return fail(PSPH_ERR_UNSUPPORTED, "free particles with different integrator counters");
            }
        }
    }
    Params &P = h->P;
    std::memset(&P, 0, sizeof P);
    for (int c = 0; c < 3; ++c) {
        P.blpf[c] = h->grid.blpf[c]; P.trpr[c] = h->grid.trpr[c]; P.csize[c] = h->grid.cell_size[c];
        P.nc[c] = h->grid.cell_no[c]; P.gravity[c] = h->cfg.gravity[c];
        P.periodic[c] = h->cfg.periodic[c] ? 1 : 0;
        if (P.periodic[c]) {
            P.nc[c] -= 2;                    // the alias layers are not stored
            P.domsize[c] = h->grid.trpr[c] - h->grid.blpf[c];
            P.any_periodic = 1;
        }
        P.afree[c] = P.gravity[c] + h->body[c];
    P.hook_before = h->hook_b ? 1 : 0;
    P.hook_after = h->hook_a ? 1 : 0;
    }
    P.ncells = P.nc[0]*P.nc[1]*P.nc[2];
    P.hmax = h->grid.hmax;
    P.cellfac = h->cfg.cellfac;
    P.sqrt_h_a = h->cfg.sqrt_h_a;
    P.xsph = h->cfg.xsph;
    P.n = (int)n;
    P.nfree = nfree;
    P.nfixed = (int)n - nfree;
    P.nclasses = (int)h->classes.size();
    P.dim = h->cfg.dimension;
    P.kernel_type = h->cfg.kernel_type;
    P.visc_eq = h->cfg.visc_eq;
    P.gradient_type = h->cfg.gradient_type;
    P.scheme = h->cfg.scheme;
    P.precision = h->cfg.precision;
    if (have_free) {
        P.eos_eq = ref.PresEq; P.eos_cs = ref.Cs; P.eos_p0 = ref.P0; P.eos_rho0 = ref.RefDensity;
        P.shepard_step = ref.ShepardStep;
    }
    for (int m = 1; m <= 3; ++m)
        if (free_m[m]) {
            P.eos_eq_m[m] = ref_m[m].PresEq; P.eos_cs_m[m] = ref_m[m].Cs;
            P.eos_p0_m[m] = ref_m[m].P0; P.eos_rho0_m[m] = ref_m[m].RefDensity;
        }
    // pairs of different materials: fluid with solid (CalcForce12, FSI smoothing) or fluid with
    // soil (CalcForce13, saturation); solid with soil has no interaction in the reference
    P.fsi = h->cfg.fsi ? 1 : 0;
    P.swi_type = h->cfg.swi_type;
    P.multi = ((int)mats[1] + (int)mats[2] + (int)mats[3]) > 1;
    if (mats[2] && mats[3])
        return fail(PSPH_ERR_UNSUPPORTED, "solid and soil particles in one run (no interaction type between them)");
    if (P.multi && mats[3] && (P.swi_type < 0 || P.swi_type > 3))
        return fail(PSPH_ERR_INVALID, "Soil-Water Interaction Type No %d is out of range", P.swi_type);
    P.swi1 = (P.multi && mats[3] && P.swi_type == 1) ? 1 : 0;
    P.any_soil_fluid = P.multi && mats[3];
    // in/out-flow boundaries with the default InCon/OutCon
    P.inout = h->cfg.in_out_flow;
    if (P.inout) {
        const psph_config &c = h->cfg;
        for (int k = 0; k < 3; ++k) { P.in_v[k] = c.in_v[k]; P.out_v[k] = c.out_v[k]; }
        P.in_density = c.in_density;
        P.out_density = c.out_density;
        P.mass_cons = c.mass_conservation ? 1 : 0;
        if (P.multi || mat0 != 1)
            return fail(PSPH_ERR_UNSUPPORTED, "in/out-flow boundaries with particles other than fluid");
        for (auto &c : h->classes)
            if (c.IsFree && c.Shepard && P.scheme == 0)
                return fail(PSPH_ERR_UNSUPPORTED, "in/out-flow boundaries with the Shepard filter (per-particle counters)");
    }
    P.any_solid_fluid = P.multi && mats[2];
    P.uniform_h = 1;
    P.h_uniform = h->classes[0].h;
    for (auto &c : h->classes) {
        if (c.h != P.h_uniform) P.uniform_h = 0;
        if (c.TI > 0.0) P.any_ti = 1;
        if (c.Material == 1 && (c.T0 > 0.0 || c.LES)) { P.any_sr = 1; P.any_les_t0 = 1; }
        if (c.Material > 1) P.any_solid = 1;
        if (c.IsFree && c.Shepard && P.scheme == 0) P.any_shepard = 1;
    }
    P.any_xsph = (h->cfg.xsph != 0.0 && P.scheme == 0) ? 1 : 0;
    if (P.uniform_h) {
        // the band of the r^2 pair test (support_test): outside it r^2 decides exactly as
        // sqrt(r^2)/h <= Cellfac does, inside it that test runs
        const double sup = P.cellfac*P.h_uniform;
        P.sup2_lo = sup*sup*(1.0 - 1e-13);
        P.sup2_hi = sup*sup*(1.0 + 1e-13);
    }
    // fast fluid path: cubic spline, Morris, squared-density form, no artificial stress, no
    // strain-rate viscosity models, no XSPH/Shepard terms, uniform h and free-particle mu
    {
        bool mu_uniform = true;
        double mu0 = have_free ? ref.Mu : 0.0;
        for (size_t i = 0; i < n; ++i)
            if (h->host[i].IsFree && h->host[i].Mu != mu0) { mu_uniform = false; break; }
        bool ok = mat0 == 1 && !P.multi && !P.inout && P.kernel_type == 0 && P.visc_eq == 0 && P.gradient_type == 0 &&
                  !P.any_ti && !P.any_sr && !P.any_xsph && !P.any_shepard && P.uniform_h &&
                  mu_uniform && P.nclasses <= 8 && have_free && !P.any_periodic &&
                  !std::getenv("PSPH_GENERAL_KERNELS");
        P.fast11 = ok ? 1 : 0;
        P.ab_two = P.ab_one = 0;
        if (ok) {
            double hh = P.h_uniform;
            P.mu_ff = 2.0*mu0*mu0/(mu0+mu0);
            P.mu_ff_wall = mu0;
            if (P.dim == 2) { P.kw = 10.0/(7.0*hh*hh*M_PI); P.kg = 10.0/(7.0*hh*hh*hh*M_PI); }
            else { P.kw = 1.0/(hh*hh*hh*M_PI); P.kg = 1.0/(hh*hh*hh*hh*M_PI); }
            P.kgh = P.kg/hh;
            P.invh = 1.0/hh;
            P.h2c = 0.01*hh*hh;
            for (int a = 0; a < P.nclasses; ++a)
                for (int b = 0; b < P.nclasses; ++b) {
                    P.pair_alpha[a*8+b] = (h->classes[a].Alpha + h->classes[b].Alpha)/2.0;
                    P.pair_beta[a*8+b] = (h->classes[a].Beta + h->classes[b].Beta)/2.0;
                }
            // one free class, and the same pair constants with every fixed class: the gathers
            // take (alpha, beta) from the neighbour's kind instead of the class-pair table
            int cf = -1, nfc = 0;
            for (int a = 0; a < P.nclasses; ++a)
                if (h->classes[a].IsFree) { cf = a; ++nfc; }
            P.ab_two = nfc == 1 ? 1 : 0;
            P.ab_free[0] = P.ab_free[1] = P.ab_fixed[0] = P.ab_fixed[1] = 0.0;
            if (P.ab_two) {
                P.ab_free[0] = P.pair_alpha[cf*8+cf];
                P.ab_free[1] = P.pair_beta[cf*8+cf];
                bool first = true;
                for (int b = 0; b < P.nclasses; ++b) {
                    if (h->classes[b].IsFree) continue;
                    const double al = P.pair_alpha[cf*8+b], be = P.pair_beta[cf*8+b];
                    if (first) { P.ab_fixed[0] = al; P.ab_fixed[1] = be; first = false; }
                    else if (al != P.ab_fixed[0] || be != P.ab_fixed[1]) P.ab_two = 0;
                }
                if (first) { P.ab_fixed[0] = P.ab_free[0]; P.ab_fixed[1] = P.ab_free[1]; }
                P.ab_one = P.ab_two && P.ab_fixed[0] == P.ab_free[0] && P.ab_fixed[1] == P.ab_free[1];
            }
            const double *ab[2] = {P.ab_free, P.ab_fixed};
            for (int k = 0; k < 2; ++k) {
                P.ab_nha[k] = -ab[k][0]*0.5;
                P.ab32[2*k] = -(float)ab[k][0]*0.5f;
                P.ab32[2*k + 1] = (float)ab[k][1];
                P.ab_on[k] = (ab[k][0] != 0.0 || ab[k][1] != 0.0) ? 1 : 0;
            }
        }
    }
    P.soil32 = 0;
    if (P.precision == PSPH_FP32_MIXED && P.multi)
        return fail(PSPH_ERR_UNSUPPORTED, "the FP32-mixed tiers cover runs of one material");
    if (P.precision == PSPH_FP32_MIXED) {
        if (mat0 > 1) {
            P.soil32 = 1;            // solid/soil: FP32 pair sums, FP64 stress update
        } else {
            if (!P.fast11)
                return fail(PSPH_ERR_UNSUPPORTED, "the FP32-mixed tier covers fluid runs with the cubic spline, "
                            "Morris viscosity, the squared-density form, uniform h and mu, no artificial stress, "
                            "no XSPH/Shepard terms, and solid/soil runs");
            P.mixed32 = 1;
        }
    }
    // neighbour-list reuse on the fast fluid path: skin = PSPH_SKIN x the support radius
    // (default 0.03; 0 turns reuse off), at most what keeps the reach inside the 5^DIM cells
    // around a particle's cell; entries then pack the index in 25 bits
    P.verlet = 0;
    P.verlet_force = 0;
    P.wl_cap = 0;
    const double support = P.cellfac*P.h_uniform;
    // the general kernels reuse their pair lists the same way in runs of one material with one
    // h, without periodic or in/out-flow boundaries (PSPH_NO_GREUSE: per-step lists instead)
    const bool gen_reuse = !P.fast11 && !P.mixed32 && P.nfree > 0 && !P.multi && !P.inout &&
                           !P.any_periodic && P.uniform_h && !std::getenv("PSPH_NO_GLIST") &&
                           !std::getenv("PSPH_NO_GREUSE");
    if (P.fast11 || P.mixed32 || gen_reuse) {
        double skin = 0.03;
        if (const char *e = std::getenv("PSPH_SKIN")) skin = std::atof(e);
        double csmin = std::min(P.csize[0], P.csize[1]);
        if (P.dim == 3) csmin = std::min(csmin, P.csize[2]);
        skin = std::min(skin, 2.0*csmin*(1.0 - 1e-9)/support - 1.0);
        if (skin > 0.0 && n < ((size_t)1 << 25)) {
            P.verlet = 1;
            double reach = support*(1.0 + skin), half = 0.5*skin*support;
            P.reach2 = reach*reach*(1.0 + 1e-9);
            P.skin_trig2 = half*half*(1.0 - 1e-6);
            P.adj2 = csmin*csmin*(1.0 - 1e-12);
            P.verlet_force = std::getenv("PSPH_REBUILD_EVERY_STEP") ? 1 : 0;
        }
    }
    if (P.mixed32) {
        P.f32.kw = (float)P.kw; P.f32.kg = (float)P.kg; P.f32.kgh = (float)P.kgh;
        P.f32.invh = (float)P.invh; P.f32.h = (float)P.h_uniform; P.f32.h2c = (float)P.h2c;
        for (int d = 0; d < 3; ++d) P.f32.cs[d] = (float)P.csize[d];
        P.f32.band_lo = (float)(P.sup2_lo*(1.0 - 1e-4));
        P.f32.band_hi = (float)(P.sup2_hi*(1.0 + 1e-4));
        P.f32.band_adj = (float)(P.adj2*(1.0 - 1e-4));
        P.f32.mu2_ff = 2.0f*(float)P.mu_ff;
        P.f32.mu2_wall = 2.0f*(float)P.mu_ff_wall;
    }
    // in/out-flow runs may grow by recycling: room for half as many particles again (more is
    // made when needed, grow_buffers); PSPH_INOUT_SLACK sets the extra room (test switch)
    size_t slack = n/2 + 256;
    if (const char *e = std::getenv("PSPH_INOUT_SLACK")) slack = (size_t)std::max(0, std::atoi(e));
    // pair records of free soil particles (single-material soil runs)
    P.srec = 0;
    if (P.any_solid && !P.multi && !std::getenv("PSPH_NO_SREC")) {
        P.srec = 1;
        for (const ClassParams &c : h->classes)
            if (c.IsFree && c.Material != 3) P.srec = 0;
    }
    rc = allocate(h, P.inout ? n + slack : n);
    if (rc) return rc;
    h->next_index = (int32_t)n;
    if (P.inout) {
        rc = inflow_init(h);
        if (rc) return rc;
    }
    // the fast path with list reuse gathers per tile with the neighbours in LDS (PSPH_NO_TILES:
    // the per-target list gather instead); the FP32 tiles pack each neighbour's cell in 10 bits
    // an axis
    const bool cells10 = P.nc[0] <= 1024 && P.nc[1] <= 1024 && P.nc[2] <= 1024;
    P.tiles = (P.fast11 && P.verlet && (!P.mixed32 || cells10) && !std::getenv("PSPH_NO_TILES")) ? 1 : 0;
    P.lean = 0;
    h->no_lean = std::getenv("PSPH_NO_LEAN") != nullptr;
    h->no_fused_views = std::getenv("PSPH_NO_FUSED_VIEWS") != nullptr;
    P.tile_items = kTileItems;
    if (const char *e = std::getenv("PSPH_TILE_ITEMS")) P.tile_items = std::max(0, std::min(kTileItems, std::atoi(e)));
    h->D.tile = nullptr; h->D.items = nullptr; h->D.tl = nullptr;
    // neighbour lists for the fast fluid path (entries pack the index in 27 bits, 25 with reuse)
    P.nl_cap = 0;
    P.nl_codes = n < ((size_t)1 << 27) ? 1 : 0;
    h->nl = h->wl = nullptr;
    if (P.mixed32 && !P.nl_codes)
        return fail(PSPH_ERR_UNSUPPORTED, "the FP32-mixed tier holds at most 2^27 particles");
    P.glist = 0;
    if (!P.fast11 && !P.mixed32 && P.nfree > 0 && (!std::getenv("PSPH_NO_GLIST") || P.multi)) {
        // pair lists of the general kernels: radius Cellfac hmax (the largest pair support), or
        // with list reuse the support plus the skin; the walls' skin lists hold four sub-list
        // capacities each
        const double radius = P.verlet ? std::sqrt(P.reach2) : P.cellfac*P.hmax;
        int cap0 = nl_initial_cap(h, radius);
        if (const char *e = std::getenv("PSPH_NL_CAP")) cap0 = std::max(4, (std::atoi(e) + 3) & ~3);
        rc = nl_alloc(h, cap0);
        if (rc) return rc;
        P.glist = 1;
        if (P.verlet) {
            int wcap = wl_initial_cap(h, radius);
            if (const char *e = std::getenv("PSPH_WL_CAP")) wcap = std::max(4, (std::atoi(e) + 3) & ~3);
            rc = wl_alloc(h, wcap);
            if (rc) return rc;
        }
    }
    if (P.fast11 || P.mixed32) {
        double radius = P.verlet ? std::sqrt(P.reach2) : support;
        int cap0 = nl_initial_cap(h, radius);
        if (const char *e = std::getenv("PSPH_NL_CAP")) cap0 = std::max(4, (std::atoi(e) + 3) & ~3);
        rc = nl_alloc(h, cap0);
        if (rc) return rc;
        if (P.verlet) {
            int wcap = wl_initial_cap(h, radius);
            if (const char *e = std::getenv("PSPH_WL_CAP")) wcap = std::max(4, (std::atoi(e) + 3) & ~3);
            rc = wl_alloc(h, wcap);
            if (rc) return rc;
        }
    }
    HIPCHECK(hipMemcpy(h->cls_dev, h->classes.data(), sizeof(ClassParams)*h->classes.size(), hipMemcpyHostToDevice));
    std::memset(h->ctl_host, 0, sizeof(Ctl));
    Ctl &C = *h->ctl_host;
    C.side = 0;
    C.rebuild = 1;                  // the first step sorts (every step does without list reuse)
    C.time = time;
    C.deltat = C.deltatint = C.deltatmin = dt;
    C.tf = INFINITY;
    C.tout = INFINITY;
    double di = dt;
    std::memcpy(&C.dtmin_bits, &di, sizeof di);
    C.halt = kRun;
    if (have_free) { C.ct = ref.ct; C.shep_counter = ref.ShepardCounter; C.first_step = ref.FirstStep ? 1 : 0; }
    else C.first_step = 1;
    rc = write_state(h, h->host.data(), n, 0);
    if (rc) return rc;
    rc = prime_first_step(h);
    if (rc) return rc;
    HIPCHECK(hipStreamSynchronize(h->stream));
    h->uploaded = true;
    return 0;
}

int psph_run(psph_handle *h, int64_t max_steps, double tf, double tout, psph_status *st)
{
    if (!h) return fail(PSPH_ERR_INVALID, "null handle");
    if (!h->uploaded) return fail(PSPH_ERR_STATE, "psph_run before psph_upload");
    if (h->broken) return fail(PSPH_ERR_STATE, "the solver stopped on an earlier error: %s", g_err.c_str());
    h->fixed_ok = false;
    psph_status s;
    std::memset(&s, 0, sizeof s);
    // read the control block, set the run limits
    int rc = sync_ctl(h);
    if (rc) return rc;
    Ctl &C = *h->ctl_host;
    const long long builds0 = C.rebuilds;
    C.tf = tf;
    C.tout = tout;
    C.max_steps = max_steps;
    C.halt = kRun;
    rc = push_ctl(h);
    if (rc) return rc;
    launch_head(h->D, h->P, h->stream);
    long long queued = 0;
    // a step stopped mid-way is finished first: from its wall pass (stop before the wall pass)
    // or by its Move (stop after the force pass)
    int pending_from = (C.mid_step != 0 && max_steps > 0) ? (C.mid_step == 2 ? 1 : 2) : 0;
    bool pending_finish = pending_from != 0;
    const long long batch = 16;
    // list reuse: steps after the first of a batch are queued lean (their tails halt before a
    // step that must rebuild); the first is lean when the control block says it keeps the lists
    const bool lean_ok = h->P.verlet && !h->no_lean;
    // the fast fluid passes' views written by the step before (a lean step keeps the order)
    const bool fuse_views = lean_ok && (h->P.fast11 || h->P.mixed32) && !h->no_fused_views;
    for (;;) {
        // queue a batch; every kernel is a no-op once the device has halted
        long long nq = std::min<long long>(batch, max_steps - queued);
        if (pending_finish && nq < 1) nq = 1;
        for (long long k = 0; k < nq; ++k) {
            const bool lean = lean_ok && (k > 0 || !C.rebuild);
            enqueue_step(h, pending_finish ? pending_from : 0, lean, lean_ok, fuse_views && k > 0,
                         fuse_views && k + 1 < nq);
            pending_finish = false;
            ++queued;
        }
        HIPCHECK(hipGetLastError());
        rc = sync_ctl(h);
        if (rc) return rc;
        if (C.reorders > 0) h->trans_ok = true;
        const bool grow_nl = (C.nl_over > 0 || (C.list_over & 1)) && h->P.nl_cap > 0;
        const bool grow_wl = (C.list_over & 2) && h->P.wl_cap > 0;
        if (grow_nl || grow_wl) {
            // some targets had more neighbours than the lists hold (their pairs were found
            // again by the gathers, or the steps rebuilt every time): grow the lists for the
            // steps to come; reused lists are rebuilt by the next step
            if (grow_nl) rc = nl_alloc(h, ((h->P.nl_cap*3/2) + 3) & ~3);
            if (!rc && grow_wl) rc = wl_alloc(h, ((h->P.wl_cap*3/2) + 3) & ~3);
            if (rc) return rc;
            C.nl_over = 0;
            C.list_over = 0;
            request_rebuild(h);
            rc = push_ctl(h);
            if (rc) return rc;
        }
        if (C.halt == kRun) {
            if (queued >= max_steps) break;   // cannot happen: the tail halts at max_steps
            continue;
        }
        if (C.halt == kHaltSort) {
            // the next step rebuilds: queue it in full (C.rebuild is set)
            C.halt = kRun;
            rc = push_ctl(h);
            if (rc) return rc;
            queued = C.steps_done;
            continue;
        }
        if (C.halt == kHaltLeave) {
            long long nd = 0;
            rc = h->P.inout ? inflow_leave(h, &nd) : delete_leavers(h, &nd);
            if (rc) { h->broken = true; return rc; }
            s.deleted += nd;
            if (C.steps_done >= max_steps || !(C.time < tf)) {
                C.halt = (C.steps_done >= max_steps) ? kHaltSteps : kHaltTend;
                break;
            }
            queued = C.steps_done;
            continue;
        }
        break;
    }
    s.time = C.time;
    s.deltat = C.deltat;
    s.deltatmin = C.deltatmin;
    s.steps = C.steps_done;
    s.pairs = (int64_t)C.last_pairs;
    s.n = (size_t)h->P.n;
    s.mid_step = C.mid_step;
    s.list_builds = C.rebuilds - builds0;
    s.stop_reason = (C.halt == kHaltOutput) ? PSPH_STOP_OUTPUT : (C.halt == kHaltTend) ? PSPH_STOP_TEND
                  : (C.halt == kHaltHookBefore) ? PSPH_STOP_HOOK_BEFORE : (C.halt == kHaltHookAfter) ? PSPH_STOP_HOOK_AFTER
                  : PSPH_STOP_STEPS;
    if (st) *st = s;
    if (C.halt == kHaltError) {
        h->broken = true;
        if (C.error == PSPH_ERR_TIMESTEP)
            return fail(PSPH_ERR_TIMESTEP, "Too small time step, please choose a smaller time step initially to make the simulation more stable");
        return fail(C.error ? C.error : PSPH_ERR_UNSUPPORTED,
                    "pairs between particles of different materials (CalcForce12/13) are outside this path");
    }
    return 0;
}

int psph_download(psph_handle *h, psph_particle *p, size_t cap, size_t *n)
{
    if (!h) return fail(PSPH_ERR_INVALID, "null handle");
    if (!h->uploaded) return fail(PSPH_ERR_STATE, "download before upload");
    if (n) *n = (size_t)h->P.n;
    if (!p) return 0;
    if (cap < (size_t)h->P.n) return fail(PSPH_ERR_CAPACITY, "need room for %d particles", h->P.n);
    int rc = read_state(h);
    if (rc) return rc;
    std::memcpy(p, h->host.data(), sizeof(psph_particle)*h->host.size());
    for (size_t i = 0; i < h->host.size(); ++i) p[i].Index = h->upload_index[i];
    return 0;
}

int psph_set_body_force(psph_handle *h, const double accel[3])
{
    if (!h || !accel) return fail(PSPH_ERR_INVALID, "null argument");
    for (int c = 0; c < 3; ++c) {
        if (!std::isfinite(accel[c])) return fail(PSPH_ERR_INVALID, "body force component %d is not finite", c);
        h->body[c] = accel[c];
        h->P.afree[c] = h->P.gravity[c] + accel[c];
    }
    return 0;
}

int psph_set_hooks(psph_handle *h, int before, int after)
{
    if (!h) return fail(PSPH_ERR_INVALID, "null handle");
    if (h->uploaded) {
        int rc = sync_ctl(h);
        if (rc) return rc;
        if (h->ctl_host->mid_step) return fail(PSPH_ERR_STATE, "hooks changed while a step is pending");
    }
    h->hook_b = before != 0;
    h->hook_a = after != 0;
    h->P.hook_before = h->hook_b ? 1 : 0;
    h->P.hook_after = h->hook_a ? 1 : 0;
    return 0;
}

int psph_hook_accel(psph_handle *h, const double *a, const double *ddensity, size_t n)
{
    if (!h || !a) return fail(PSPH_ERR_INVALID, "null argument");
    if (!h->uploaded) return fail(PSPH_ERR_STATE, "hook values before upload");
    int rc = sync_ctl(h);
    if (rc) return rc;
    const int ms = h->ctl_host->mid_step;
    if (ms != 1 && ms != 2) return fail(PSPH_ERR_STATE, "psph_hook_accel outside a stop inside a step");
    if (n != (size_t)h->P.n) return fail(PSPH_ERR_INVALID, "%zu values for %d particles", n, h->P.n);
    if (h->hook_cap < 4*n) {
        if (h->hook_dev) { rc = dev_free(h, (void **)&h->hook_dev); if (rc) return rc; }
        h->hook_cap = 0;
        rc = dev_alloc(h, (void **)&h->hook_dev, sizeof(double)*4*n);
        if (rc) return rc;
        h->hook_cap = 4*n;
    }
    // blocking copies from the caller's (pageable) arrays; the stream is idle at a stop
    HIPCHECK(hipMemcpy(h->hook_dev, a, sizeof(double)*3*n, hipMemcpyHostToDevice));
    if (ddensity)
        HIPCHECK(hipMemcpy(h->hook_dev + 3*n, ddensity, sizeof(double)*n, hipMemcpyHostToDevice));
    launch_hook_accel(h->D, h->P, h->hook_dev, ddensity ? h->hook_dev + 3*n : nullptr, h->stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(h->stream));
    return 0;
}

int psph_output_begin(psph_handle *h)
{
    if (!h) return fail(PSPH_ERR_INVALID, "null handle");
    if (!h->uploaded) return fail(PSPH_ERR_STATE, "output before upload");
    if (h->out_pending) return fail(PSPH_ERR_STATE, "an output is pending (call psph_output_end first)");
    int rc = sync_ctl(h);
    if (rc) return rc;
    if (h->ctl_host->mid_step == 2) return fail(PSPH_ERR_STATE, "output at a stop before the pair sums");
    rc = form_fixed_sums(h);
    if (rc) return rc;
    const Params &P = h->P;
    const size_t n = (size_t)P.n, nf = (size_t)output_floats();
    if (n > h->out_cap) {
        if (h->out_dev) { rc = dev_free(h, (void **)&h->out_dev); if (rc) return rc; }
        if (h->out_host) { HIPCHECK(hipHostFree(h->out_host)); h->out_host = nullptr; }
        h->out_cap = 0;
        rc = dev_alloc(h, (void **)&h->out_dev, sizeof(float)*nf*n);
        if (rc) return rc;
        HIPCHECK(hipHostMalloc((void **)&h->out_host, sizeof(float)*nf*n, hipHostMallocDefault));
        h->out_cap = n;
    } else if (!h->out_dev) {
        // the device buffers were re-created for more particles; the pinned copy is kept
        rc = dev_alloc(h, (void **)&h->out_dev, sizeof(float)*nf*h->out_cap);
        if (rc) return rc;
    }
    if (!h->out_event) HIPCHECK(hipEventCreateWithFlags(&h->out_event, hipEventDisableTiming));
    const Dev D = with_state(h, state_buffer(h));
    const bool acc_dev = h->trans_ok, tensors_dev = P.any_solid != 0;
    launch_pack_output(D, P, h->out_dev, acc_dev ? (h->fixed_ok ? 2 : 1) : 0, tensors_dev ? 1 : 0, h->stream);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(h->out_host, h->out_dev, sizeof(float)*nf*n, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipEventRecord(h->out_event, h->stream));
    h->out_pending = true;
    h->out_n = n;
    // fields the device does not hold for this state come from the host mirror (reference
    // order), as psph_download returns them; filled while the copy runs
    h->out_acc.clear();
    h->out_sig.clear();
    h->out_eps.clear();
    if (!acc_dev) {
        h->out_acc.resize(3*n);
        for (size_t o = 0; o < n; ++o)
            for (int c = 0; c < 3; ++c) h->out_acc[3*o + c] = (float)h->host[o].a[c];
    }
    if (!tensors_dev) {
        static const int sym[6] = {0, 1, 2, 4, 5, 8};      // xx xy xz yy yz zz of a row-major 3x3
        const bool zero_fixed = h->ctl_host->reorders > 0;
        h->out_sig.resize(6*n);
        h->out_eps.resize(6*n);
        for (size_t o = 0; o < n; ++o) {
            const psph_particle &q = h->host[o];
            const bool z = zero_fixed && !q.IsFree;
            for (int c = 0; c < 6; ++c) {
                h->out_sig[6*o + c] = z ? 0.0f : (float)q.Sigma[sym[c]];
                h->out_eps[6*o + c] = (float)q.Strain[sym[c]];
            }
        }
    }
    return 0;
}

int psph_output_end(psph_handle *h, psph_output_f32 *out)
{
    if (!h || !out) return fail(PSPH_ERR_INVALID, "null argument");
    if (!h->out_pending) return fail(PSPH_ERR_STATE, "no output pending (call psph_output_begin first)");
    h->out_pending = false;
    HIPCHECK(hipEventSynchronize(h->out_event));
    const size_t n = h->out_n;
    const float *b = h->out_host;
    out->n = n;
    out->position = b;
    out->velocity = b + 3*n;
    out->acceleration = h->out_acc.empty() ? b + 6*n : h->out_acc.data();
    out->pressure = b + 9*n;
    out->density = b + 10*n;
    out->mass = b + 11*n;
    out->h = b + 12*n;
    out->tag = reinterpret_cast<const int32_t *>(b + 13*n);
    out->sigma = h->out_sig.empty() ? b + 14*n : h->out_sig.data();
    out->strain = h->out_eps.empty() ? b + 20*n : h->out_eps.data();
    return 0;
}

// The pair query of psph_neighbour_pairs / psph_neighbour_pairs_noalloc.  staging: device
// memory for 2*staging_pairs values to stage the list in, or null to allocate it for the call.
static int pair_query(psph_handle *h, int64_t *pairs, size_t cap, size_t *npairs,
                      long long *staging, size_t staging_pairs)
{
    if (!h || !npairs) return fail(PSPH_ERR_INVALID, "null argument");
    if (!h->uploaded) return fail(PSPH_ERR_STATE, "pairs before upload");
    int rc = sync_ctl(h);
    if (rc) return rc;
    if (h->ctl_host->mid_step) return fail(PSPH_ERR_STATE, "pair query while a step is pending");
    const Params &P = h->P;
    const Dev &D = h->D;
    // a sort of the current positions (the cell structure after Move; the order is canonical, so
    // an extra sort changes no later result)
    Ctl save = *h->ctl_host;
    h->ctl_host->halt = kRun;
    h->ctl_host->rebuild = 1;
    rc = push_ctl(h);
    if (rc) return rc;
    if (P.verlet) launch_keys(D, P, h->stream);          // ranks are kept only for rebuilds
    launch_scan(D, P, h->stream);
    launch_scatter(D, P, h->stream);
    launch_cellsort(D, P, h->stream);
    launch_reorder(D, P, h->stream);
    launch_pairs(D, P, h->stream, 0, nullptr, h->pair_cnt, nullptr);   // on the sorted copy
    std::vector<int> cnt(P.n);
    HIPCHECK(hipMemcpyAsync(cnt.data(), h->pair_cnt, sizeof(int)*P.n, hipMemcpyDeviceToHost, h->stream));
    HIPCHECK(hipStreamSynchronize(h->stream));
    std::vector<long long> off(P.n);
    long long tot = 0;
    for (int i = 0; i < P.n; ++i) { off[i] = tot; tot += cnt[i]; }
    *npairs = (size_t)tot;
    const bool fits_staging = !staging || (size_t)tot <= staging_pairs;
    if (pairs && (size_t)tot <= cap && tot > 0 && fits_staging) {
        long long *out_dev = staging;
        // blocking copy from the pageable offsets; the stream is idle here
        HIPCHECK(hipMemcpy(h->pair_off, off.data(), sizeof(long long)*P.n, hipMemcpyHostToDevice));
        if (!staging) HIPCHECK(hipMalloc((void **)&out_dev, sizeof(long long)*2*tot));
        // k_pairs writes every one of the 2*tot values: nothing left from an earlier call is read
        launch_pairs(D, P, h->stream, 1, h->pair_off, nullptr, out_dev);
        hipError_t e = hipMemcpyAsync(pairs, out_dev, sizeof(long long)*2*tot, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (!staging) hipFree(out_dev);
        if (e != hipSuccess) return fail(PSPH_ERR_HIP, "pair copy: %s", hipGetErrorString(e));
    }
    // restore the control block; the sorted copy is now the state, and the lists are stale
    rc = sync_ctl(h);
    if (rc) return rc;
    long long reorders = h->ctl_host->reorders, rebuilds = h->ctl_host->rebuilds;
    *h->ctl_host = save;
    h->ctl_host->reorders = reorders;
    h->ctl_host->rebuilds = rebuilds;
    h->ctl_host->side = save.side ^ 1;
    rc = prime_first_step(h);
    if (rc) return rc;
    HIPCHECK(hipStreamSynchronize(h->stream));
    h->trans_ok = false;
    h->fixed_ok = false;
    if (pairs && (size_t)tot > cap) return fail(PSPH_ERR_CAPACITY, "need room for %lld pairs", tot);
    if (pairs && !fits_staging)
        return fail(PSPH_ERR_CAPACITY, "%lld pairs exceed the %zu that the no-allocation pair query "
                    "stages (a %zu MiB device array); use psph_neighbour_pairs", tot, staging_pairs,
                    kScratchBytes >> 20);
    return 0;
}

int psph_neighbour_pairs(psph_handle *h, int64_t *pairs, size_t cap, size_t *npairs)
{
    return pair_query(h, pairs, cap, npairs, nullptr, 0);
}

int psph_neighbour_pairs_noalloc(psph_handle *h, int64_t *pairs, size_t cap, size_t *npairs)
{
    if (!h) return fail(PSPH_ERR_INVALID, "null handle");
    if (!h->pair_scratch) return fail(PSPH_ERR_HIP, "%s", h->pair_scratch_err.c_str());
    return pair_query(h, pairs, cap, npairs, h->pair_scratch, kScratchPairs);
}

int psph_get_grid(psph_handle *h, psph_grid *g)
{
    if (!h || !g) return fail(PSPH_ERR_INVALID, "null argument");
    if (!h->uploaded) return fail(PSPH_ERR_STATE, "grid before upload");
    *g = h->grid;
    return 0;
}

int psph_set_timing(psph_handle *h, int enable)
{
    if (!h) return fail(PSPH_ERR_INVALID, "null handle");
    if (enable < 0 || enable > PSPH_TIMING_FORCE) return fail(PSPH_ERR_INVALID, "timing mode %d", enable);
    h->timing = enable != 0;
    h->timing_force_only = enable == PSPH_TIMING_FORCE;
    for (int k = 0; k < KCount; ++k) { h->t_ms[k] = 0.0; h->t_n[k] = 0; }
    return 0;
}

int psph_kernel_stats(psph_handle *h, char *names, size_t names_len, double *mean_ms,
                      int64_t *launches, size_t cap, size_t *nkernels)
{
    if (!h) return fail(PSPH_ERR_INVALID, "null handle");
    if (nkernels) *nkernels = KCount;
    std::string all;
    for (int k = 0; k < KCount; ++k) {
        if (k) all += ",";
        all += kKernelNames[k];
        if ((size_t)k < cap) {
            if (mean_ms) mean_ms[k] = h->t_n[k] ? h->t_ms[k]/h->t_n[k] : 0.0;
            if (launches) launches[k] = h->t_n[k];
        }
    }
    if (names && names_len) std::snprintf(names, names_len, "%s", all.c_str());
    return 0;
}

int psph_stream(psph_handle *h, void **stream)
{
    if (!h || !stream) return fail(PSPH_ERR_INVALID, "null argument");
    *stream = (void *)h->stream;
    return 0;
}

// Compulsory HBM bytes per step, each array read or written once per kernel, FP64 fluid:
// keys 32, sort 16, cell table 8/cell, reorder 2 x 112, wall 64 + 32 f_fix, force 112 + 24 f_fix,
// integrate 192 f_free (per particle).
int psph_traffic_model(psph_handle *h, double *bytes_per_step, double *force_bytes)
{
    if (!h || !h->uploaded) return fail(PSPH_ERR_STATE, "traffic model before upload");
    double N = h->P.n, Nf = h->P.nfree, Nx = h->P.nfixed, C = h->P.ncells;
    double force, wall, views;
    if (h->P.precision == PSPH_FP32_MIXED) {
        // FP32-mixed: the gather reads FP32 x, v (24 B), rho, P, m (12 B) and the flags (4 B) of
        // every particle once and the FP32 NSv of the walls (12 B), and writes a and dDensity
        // of the free particles in FP64 (32 B); the wall pass reads the FP32 x, v, rho, P
        // (32 B); the reorder writes those FP32 copies (40 B) besides the FP64 state
        force = 40.0*N + 12.0*Nx + 32.0*Nf;
        wall = 32.0*N + 32.0*Nx;
        views = 40.0*N;
    } else {
        // FP64: the gather reads x, v, rho, P, m, flags of every particle once (80 B) and NSv
        // of the walls (24 B), and writes a and dDensity of the free particles (32 B)
        force = 80.0*N + 24.0*Nx + 32.0*Nf;
        wall = 64.0*N + 32.0*Nx;
        views = 0.0;
    }
    // keys 32 B, sort 16 B, cell table 8 B per cell, reorder of the persistent state 2 x 112 B,
    // integration of the free particles 192 B
    double total = 32.0*N + 16.0*N + 8.0*C + 224.0*N + views + wall + force + 192.0*Nf;
    if (h->P.any_solid) {
        // soil and solid: the tensors Sigma, StrainRate, RotationRate, Strain, TIR read and
        // written once, ~650 B per free particle; of these the force pass reads Sigma and TIR
        // of every particle (96 B) and writes StrainRate and RotationRate of the free ones (72 B)
        const double tens = 96.0*N + 72.0*Nf;
        force += tens;
        total += std::max(650.0*Nf, tens);
    }
    if (bytes_per_step) *bytes_per_step = total;
    if (force_bytes) *force_bytes = force;
    return 0;
}

int psph_destroy(psph_handle *h)
{
    if (!h) return 0;
    free_all(h);
    if (h->stream) hipStreamDestroy(h->stream);
    delete h;
    return 0;
}

}  // extern "C"
