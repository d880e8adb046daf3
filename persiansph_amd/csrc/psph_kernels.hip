// HIP kernels of the PersianSPH Solve hot path for MI355X (gfx950, wave64).
//
// One time step (the loop body of Domain::Solve, Source/Domain.cpp:1762-1793) is:
//   scan       cell counts -> cell-start table            (replaces ListGenerate's HOC chains)
//   scatter    counting-sort scatter by cell key
//   cellsort   order inside each cell by original index   (deterministic, history-free order)
//   reorder    gather persistent columns into the new order, target lists, TIR of free soil
//   wall       fixed-particle smoothing gather            (PrimaryComputeAcceleration)
//   force      free-particle pair gather                  (LastComputeAcceleration, CalcForce11/2233)
//   integrate  AdaptiveTimeStep + Move + Mat1/Mat2/Mat3 + leave check + next-step cell keys
//   tail       Time += dt, step counters, stop conditions
// Every kernel returns at once when the device control block says the run has halted, so the
// host may queue many steps ahead without round trips.
//
// Neighbour-list reuse (fast fluid path): the sort (keys, scan, scatter, cellsort, permuting
// reorder) and the list builds run only on steps that rebuild.  Lists then carry every candidate
// within the support radius plus a skin; integrate flags a rebuild once some free particle has
// moved half the skin since the build.  The other steps copy the columns in place order, walk
// the skin lists and keep exactly the pairs the reference's cell search and rij/h <= Cellfac
// test would form.
//
// Pair sets: a free particle gathers every candidate in the 27 (3-D) or 9 (2-D) cells of the
// reference grid around its own and keeps those with rij/h <= Cellfac, exactly the pairs
// MainNeighbourSearch + CalcForce* evaluate.  Each pair contribution is computed with the
// gathering particle as P1; the reference's pair formulas are antisymmetric to the bit, so this
// equals its contribution whichever particle the reference's search made P1.  The one
// orientation-dependent term, the hydrostatic wall term of PrimaryComputeAcceleration
// (Domain.cpp:1045,1054), is oriented by the reference rule (smaller cell key, or larger
// original index within a cell, is P1).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "psph_dev.h"
#include "sph_math.h"

namespace psph {

constexpr int kBlock = 256;
constexpr int kForceBlock = 128;
constexpr int kCap = 24;            // per-lane neighbour list slots in LDS before a flush
constexpr int kScanTile = 2048;     // cells per scan block (256 threads x 8)
constexpr int kOverBlocks = 32;     // workgroups of the neighbour-list overflow passes

__device__ inline unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ inline double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

__device__ inline int wave_any(bool p) { return __ballot(p) != 0ull; }

// Workgroups go to the 8 XCDs of the device round-robin (block b to XCD b mod 8), and each XCD
// has its own L2.  The pair gathers remap the block index so that XCD x works on the x-th
// contiguous share of the (cell-sorted) targets: neighbouring targets read the same neighbour
// data, which then stays in one L2 instead of being fetched into all eight.
__device__ inline int xcd_block()
{
    int nb = (int)gridDim.x, b = (int)blockIdx.x;
    int q = nb >> 3, r = nb & 7;
    int x = b & 7, k = b >> 3;
    return (x < r) ? x*(q + 1) + k : r*(q + 1) + (x - r)*q + k;
}

// The persistent-column views of this step (see Dev): the state at the step's start is in
// buf[side], the working copy in buf[side ^ rebuild].  Only the tail changes side and rebuild,
// at the end of a step, so every kernel of a step binds the same buffers.
__device__ inline Dev bind(const Dev &D0)
{
    Dev D = D0;
    const int s = D0.ctl->side, b = s ^ D0.ctl->rebuild;
    const Cols &S = D0.buf[s], &B = D0.buf[b];
    D.pos = S.pos; D.vel = S.vel; D.vh = S.vh; D.mmu = S.mmu; D.flags = S.flags; D.orig = S.orig;
    D.sig = S.sig; D.sigh = S.sigh; D.tau = S.tau; D.tauh = S.tauh; D.eps = S.eps; D.epsh = S.epsh;
    D.pos_b = B.pos; D.vel_b = B.vel; D.vh_b = B.vh; D.mmu_b = B.mmu; D.flags_b = B.flags; D.orig_b = B.orig;
    D.sig_b = B.sig; D.sigh_b = B.sigh; D.tau_b = B.tau; D.tauh_b = B.tauh; D.eps_b = B.eps; D.epsh_b = B.epsh;
    return D;
}

// ------------------------------------------------------------------------------------------
// cell keys (ListGenerate, Domain.cpp:662-784): floor((x-BLPF)/CellSize), clamped to the grid
// for particles within hmax outside it
// ------------------------------------------------------------------------------------------
__device__ inline int cell_key(const Params &P, double x, double y, double z, int &kx, int &ky, int &kz)
{
    double xs[3] = {x, y, z};
    int ijk[3] = {0, 0, 0};
    for (int c = 0; c < P.dim; ++c) {
        int i = (int)(floor((xs[c] - P.blpf[c]) / P.csize[c]));
        if (i < 0) i = 0;
        if (i >= P.nc[c]) i = P.nc[c] - 1;
        ijk[c] = i;
    }
    kx = ijk[0]; ky = ijk[1]; kz = ijk[2];
    return ijk[0] + P.nc[0]*(ijk[1] + P.nc[1]*ijk[2]);
}

// Arrival rank of a particle inside its next-step cell, with one atomic per distinct cell per
// wave instead of one per particle (a wave holds consecutive, mostly same-cell particles).
// Lanes with key < 0 take no rank.  The free-particle count of the cell is raised alongside.
__device__ inline int cell_rank(const Dev &D, int key, bool isfree)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const unsigned long long fmask = __ballot(isfree && key >= 0);
    unsigned long long todo = __ballot(key >= 0);
    int rank = 0;
    while (todo) {
        int leader = __ffsll((long long)todo) - 1;
        int lk = __builtin_amdgcn_readlane(key, leader);
        unsigned long long same = __ballot(key == lk);
        int base = 0;
        if (lane == leader) {
            base = atomicAdd(&D.ccount[lk], __popcll(same));
            int nf = __popcll(same & fmask);
            if (nf) atomicAdd(&D.cfcnt[lk], nf);
        }
        base = __builtin_amdgcn_readlane(base, leader);
        if (key == lk) rank = base + __popcll(same & lt);
        todo &= ~same;
    }
    return rank;
}

// With neighbour-list reuse the sort and the list builds run only on steps that rebuild.
__device__ inline bool sort_skipped(const Dev &D, const Params &P) { return P.verlet && !D.ctl->rebuild; }

// The reference's pair test rij/h <= Cellfac (CalcForce11, Interaction.cpp:32) on a squared
// distance: an r^2 band decides all but a 1e-13-wide shell around the support radius, where
// the exact test runs on sqrt(r^2) as the reference forms rij.
__device__ inline bool in_support(const Params &P, double r2)
{
    if (r2 <= P.sup2_lo) return true;
    if (r2 > P.sup2_hi) return false;
    return (sqrt(r2)/P.h_uniform) <= P.cellfac;
}

// The reference's pair test rij/h <= Cellfac with rij = sqrt(r2) (CalcForce11/2233): with one h
// for every class it is decided on r2 outside a 1e-13 band around the support radius
// (in_support), else as written.
__device__ inline bool support_test(const Params &P, double r2, double h)
{
    if (P.uniform_h) return in_support(P, r2);
    return (sqrt(r2)/h) <= P.cellfac;
}


// Cell keys and arrival ranks of the current positions: at upload (and after host-side changes)
// for the first step, and with list reuse at the head of every step that rebuilds.
__global__ void __launch_bounds__(kBlock) k_keys(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt || sort_skipped(D, P)) return;
    int i = blockIdx.x*kBlock + threadIdx.x;
    if (i >= P.n) return;
    double4 p = D.pos[i];
    int kx, ky, kz;
    int key = cell_key(P, p.x, p.y, p.z, kx, ky, kz);
    D.key[i] = key;
    D.rank[i] = cell_rank(D, key, flag_free(D.flags[i]));
}

// ------------------------------------------------------------------------------------------
// exclusive scan of the cell counts -> cell-start table; counts are cleared for the next step
// ------------------------------------------------------------------------------------------
template <class T>
__device__ inline T block_exclusive_scan(T v, T *sh, T &total)
{
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T s = 0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) { T t = sh[k]; sh[k] = s; s += t; }
        sh[16] = s;
    }
    __syncthreads();
    total = sh[16];
    T r = x - v + sh[w];
    __syncthreads();
    return r;
}

// The scan carries two counts per cell packed in one 64-bit word, all particles in the high
// half and free particles in the low half; neither half can carry into the other (each total
// is below 2^31).
__device__ inline unsigned long long pack_counts(int all, int free_)
{
    return ((unsigned long long)(unsigned)all << 32) | (unsigned)free_;
}

__global__ void __launch_bounds__(256) k_scan_reduce(Dev D, Params P)
{
    if (D.ctl->halt || sort_skipped(D, P)) return;
    __shared__ unsigned long long sh[17];
    int base = blockIdx.x*kScanTile + threadIdx.x*8;
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) { int c = base + k; if (c < P.ncells) s += pack_counts(D.ccount[c], D.cfcnt[c]); }
    unsigned long long total;
    block_exclusive_scan(s, sh, total);
    if (threadIdx.x == 0) D.bsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) k_scan_top(Dev D, Params P, int nblocks)
{
    if (D.ctl->halt || sort_skipped(D, P)) return;
    __shared__ unsigned long long sh[17];
    __shared__ unsigned long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nblocks; b0 += 1024) {
        int b = b0 + threadIdx.x;
        unsigned long long v = b < nblocks ? D.bsum[b] : 0;
        unsigned long long total;
        unsigned long long ex = block_exclusive_scan(v, sh, total);
        if (b < nblocks) D.bsum[b] = ex + carry;
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        D.cstart[P.ncells] = (int)(carry >> 32);
        D.fstart[P.ncells] = (int)(carry & 0xffffffffull);
    }
}

__global__ void __launch_bounds__(256) k_scan_down(Dev D, Params P)
{
    if (D.ctl->halt || sort_skipped(D, P)) return;
    __shared__ unsigned long long sh[17];
    int base = blockIdx.x*kScanTile + threadIdx.x*8;
    unsigned long long v[8];
    int fc[8];
    unsigned long long s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        int c = base + k;
        fc[k] = (c < P.ncells) ? D.cfcnt[c] : 0;
        v[k] = (c < P.ncells) ? pack_counts(D.ccount[c], fc[k]) : 0;
        s += v[k];
    }
    unsigned long long total;
    unsigned long long ex = block_exclusive_scan(s, sh, total) + D.bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        int c = base + k;
        if (c < P.ncells) {
            D.cstart[c] = (int)(ex >> 32);
            D.fstart[c] = (int)(ex & 0xffffffffull);
            D.ccount[c] = 0; D.cfree[c] = fc[k]; D.cfcnt[c] = 0;
        }
        ex += v[k];
    }
}

// counting-sort scatter: new position = cell start + arrival rank
__global__ void __launch_bounds__(kBlock) k_scatter(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt || sort_skipped(D, P)) return;
    int i = blockIdx.x*kBlock + threadIdx.x;
    if (i >= P.n) return;
    int k = D.key[i];
    int pos = D.cstart[k] + D.rank[i];
    D.src[pos] = i;
    D.sorig[pos] = (unsigned)D.orig[i] | (flag_free(D.flags[i]) ? 0x80000000u : 0u);
    D.key_b[pos] = k;
}

// Order inside each cell by original index: a particle's final slot is its cell's start plus
// the number of particles of the cell with a smaller original index (original indices are
// distinct; cells hold tens of particles, read contiguously).  The same count over the free
// particles of the cell places the particle in the free-target list (free particles in sorted
// order, cell c's share starting at fstart[c]) or in the fixed-target list (fixed particles in
// sorted order, cell c's share starting at cstart[c] - fstart[c]).
__global__ void __launch_bounds__(kBlock) k_cellsort(Dev D, Params P)
{
    if (D.ctl->halt || sort_skipped(D, P)) return;
    int p = blockIdx.x*kBlock + threadIdx.x;
    if (p >= P.n) return;
    int k = D.key_b[p];
    int b = D.cstart[k], e = D.cstart[k+1];
    unsigned so = D.sorig[p];
    unsigned o = so & 0x7fffffffu;
    int r = 0, rf = 0;
    for (int q = b; q < e; ++q) {
        unsigned x = D.sorig[q];
        int lt = ((x & 0x7fffffffu) < o) ? 1 : 0;
        r += lt;
        rf += lt & (int)(x >> 31);
    }
    int dst = b + r;
    D.perm[dst] = D.src[p];
    int fs = D.fstart[k];
    if (so >> 31) D.free_list[fs + rf] = dst;
    else D.fixed_list[(b - fs) + (r - rf)] = dst;
}

__device__ inline void decode_key(const Params &P, int key, int &kx, int &ky, int &kz)
{
    kx = key % P.nc[0];
    int t = key / P.nc[0];
    ky = t % P.nc[1];
    kz = t / P.nc[1];
}

// Origin (lower corner) of a cell of the reference grid: BLPF + k CellSize (FP64).
__device__ inline void cell_origin(const Params &P, int key, double &o0, double &o1, double &o2)
{
    int kx, ky, kz;
    decode_key(P, key, kx, ky, kz);
    o0 = P.blpf[0] + kx*P.csize[0];
    o1 = P.blpf[1] + ky*P.csize[1];
    o2 = (P.dim == 3) ? P.blpf[2] + kz*P.csize[2] : 0.0;
}

// ------------------------------------------------------------------------------------------
// reorder: gather persistent columns into the new order; build target lists; per-particle
// derived values used by the pair gathers; TIR of free soil/solid (StartAcceleration
// Domain.cpp:951-988).  With list reuse, a step that does not rebuild keeps the order: the
// columns are copied as they stand (the double buffer flips every step either way).
// ------------------------------------------------------------------------------------------
// One particle's 6-component tensor (a 16-B aligned 48-B record) as three 16-B loads / stores
__device__ __forceinline__ void ld6(const double *a, int i, double *v)
{
    const double2 *p = reinterpret_cast<const double2 *>(a + 6*(size_t)i);
    const double2 t0 = p[0], t1 = p[1], t2 = p[2];
    v[0] = t0.x; v[1] = t0.y; v[2] = t1.x; v[3] = t1.y; v[4] = t2.x; v[5] = t2.y;
}
__device__ __forceinline__ void st6(double *a, int i, const double *v)
{
    double2 *p = reinterpret_cast<double2 *>(a + 6*(size_t)i);
    p[0] = make_double2(v[0], v[1]); p[1] = make_double2(v[2], v[3]); p[2] = make_double2(v[4], v[5]);
}

// One particle's 6-component tensor from slot s of src to slot i of dst
__device__ __forceinline__ void copy_tensor6(const double *src, double *dst, int s, int i)
{
    const double2 *a = reinterpret_cast<const double2 *>(src + 6*(size_t)s);
    double2 *b = reinterpret_cast<double2 *>(dst + 6*(size_t)i);
    const double2 t0 = a[0], t1 = a[1], t2 = a[2];
    b[0] = t0; b[1] = t1; b[2] = t2;
}

// The views of free particle i (state p, v, mm; class c) that the pair passes read: aux (mass,
// mu, sound speed and P/rho^2: CalcForce11 :67-72, :139), and the fast fluid path's records,
// FP64 (g0..g2) or FP32-mixed (f0..f2, positions relative to the origin of the cell the
// particle had at the last sort).
__device__ inline void free_views(const Dev &D, const Params &P, const ClassParams &c, int i, uint32_t f,
                                  double4 p, double4 v, double2 mm)
{
    double C = sound_speed(c.PresEq, c.Cs, p.w, c.RefDensity);
    double pr = v.w/(p.w*p.w);
    D.aux[i] = make_double4(mm.x, mm.y, C, pr);
    if (P.mixed32) {
        double o0, o1, o2;
        cell_origin(P, D.key_b[i], o0, o1, o2);
        D.f0[i] = make_float4((float)(p.x - o0), (float)(p.y - o1), (float)(p.z - o2), (float)p.w);
        D.f1[i] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)pr);
        D.f2[i] = make_float4((float)C, (float)flag_cls(f), (float)mm.x, (float)(1.0/p.w));
    } else if (P.fast11) {
        D.g0[i] = p;
        D.g1[i] = make_double4(v.x, v.y, v.z, pr);
        D.g2[i] = make_double4(C, (double)flag_cls(f), mm.x, 1.0/p.w);
    }
}

__global__ void __launch_bounds__(kBlock) k_reorder(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    __shared__ ClassParams cp[kMaxClasses];
    for (int k = threadIdx.x; k < P.nclasses; k += kBlock) cp[k] = D.cls[k];
    __syncthreads();
    const bool sorted = !sort_skipped(D, P);
    int i = blockIdx.x*kBlock + threadIdx.x;
    if (i < P.n) {
        const int s = sorted ? D.perm[i] : i;
        const uint32_t f = D.flags[s];
        // a step that keeps the order works in place (pos_b is pos): fixed particles of the
        // fluid path then need nothing here (the wall pass writes their views)
        if (sorted || flag_free(f) || P.any_solid) {
            double4 p = D.pos[s], v = D.vel[s];
            double2 mm = D.mmu[s];
            if (P.dim == 2) v.z = 0.0;          // StartAcceleration zeroes v(2) in 2-D (Domain.cpp:1011)
            if (sorted) {
                D.pos_b[i] = p;
                D.vel_b[i] = v;
                D.vh_b[i] = D.vh[s];
                D.mmu_b[i] = mm;
                D.flags_b[i] = f;
                D.orig_b[i] = D.orig[s];
                D.key[i] = D.key_b[i];
                if (P.verlet) D.xref[i] = p;      // positions the lists are built from
            } else if (P.dim == 2) {
                D.vel_b[i] = v;
            }
            // key_b[i] stays the cell of slot i at the last sort, the origin of the FP32 positions
            const ClassParams &c = cp[flag_cls(f)];
            if (flag_free(f)) {
                free_views(D, P, c, i, f, p, v, mm);
            } else {
                D.aux[i] = make_double4(mm.x, mm.y, 0.0, 0.0);
            }
            if (P.any_solid) {
                // 48-B tensor records (16-B aligned) moved as three 16-B loads and stores
                double sg[6];
                const double2 *sg2 = reinterpret_cast<const double2 *>(D.sig + 6*(size_t)s);
                for (int q = 0; q < 3; ++q) { const double2 t = sg2[q]; sg[2*q] = t.x; sg[2*q+1] = t.y; }
                if (sorted) {
                    double2 *d2 = reinterpret_cast<double2 *>(D.sig_b + 6*(size_t)i);
                    for (int q = 0; q < 3; ++q) d2[q] = make_double2(sg[2*q], sg[2*q+1]);
                    copy_tensor6(D.sigh, D.sigh_b, s, i);
                    copy_tensor6(D.tau, D.tau_b, s, i);
                    copy_tensor6(D.tauh, D.tauh_b, s, i);
                    copy_tensor6(D.eps, D.eps_b, s, i);
                    copy_tensor6(D.epsh, D.epsh_b, s, i);
                }
                if (flag_free(f)) {
                    // TIR is rebuilt every step; entries the 2-D formula leaves alone are zero
                    double t[6] = {0, 0, 0, 0, 0, 0};
                    if (c.Material > 1 && c.TI > 0.0) tir_tensor(P.dim, sg, c.TI, p.w, t);
                    double2 *t2 = reinterpret_cast<double2 *>(D.tir + 6*(size_t)i);
                    for (int q = 0; q < 3; ++q) t2[q] = make_double2(t[2*q], t[2*q+1]);
                    if (P.srec) {
                        // the particle's CalcForce2233 record as a neighbour (see Dev::srec)
                        double2 *r = D.srec + 10*(size_t)i;
                        r[0] = make_double2(p.x, p.y); r[1] = make_double2(p.z, p.w);
                        r[2] = make_double2(v.x, v.y); r[3] = make_double2(v.z, mm.x);
                        for (int q = 0; q < 3; ++q) r[4 + q] = make_double2(sg[2*q], sg[2*q+1]);
                        for (int q = 0; q < 3; ++q) r[7 + q] = make_double2(t[2*q], t[2*q+1]);
                    }
                }
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        D.ctl->reorders += 1;
        if (sorted) { D.ctl->rebuilds += 1; D.ctl->nwact = 0; }
    }
}

// ------------------------------------------------------------------------------------------
// neighbour rows of a cell: 9 (3-D) or 3 (2-D) contiguous ranges of the sorted order
// ------------------------------------------------------------------------------------------
__device__ inline void row_range(const Dev &D, const Params &P, int kx, int ky, int kz, int r,
                                 int &beg, int &end)
{
    int dy = (r % 3) - 1, dz = (P.dim == 3) ? (r / 3) - 1 : 0;
    int yy = ky + dy, zz = kz + dz;
    beg = end = 0;
    if (yy < 0 || yy >= P.nc[1] || zz < 0 || zz >= P.nc[2]) return;
    int base = P.nc[0]*(yy + P.nc[1]*zz);
    int x0 = kx > 0 ? kx - 1 : 0;
    int x1 = kx + 1 < P.nc[0] ? kx + 1 : P.nc[0] - 1;
    beg = D.cstart[base + x0];
    end = D.cstart[base + x1 + 1];
}

// The same row with periodic directions (Domain.cpp:757-783: the reference's alias layers make
// the cells across a period neighbours): y and z wrap, and an x row that crosses the period is
// two ranges of the sorted order (the second empty otherwise).  Needs 3+ cells per period.
__device__ inline void row_range_periodic(const Dev &D, const Params &P, int kx, int ky, int kz, int r,
                                          int &beg0, int &end0, int &beg1, int &end1)
{
    int dy = (r % 3) - 1, dz = (P.dim == 3) ? (r / 3) - 1 : 0;
    int yy = ky + dy, zz = kz + dz;
    beg0 = end0 = beg1 = end1 = 0;
    if (P.periodic[1]) yy = (yy + P.nc[1]) % P.nc[1];
    if (P.periodic[2]) zz = (zz + P.nc[2]) % P.nc[2];
    if (yy < 0 || yy >= P.nc[1] || zz < 0 || zz >= P.nc[2]) return;
    int base = P.nc[0]*(yy + P.nc[1]*zz);
    int nx = P.nc[0];
    if (P.periodic[0] && kx == 0) {
        beg0 = D.cstart[base + nx - 1]; end0 = D.cstart[base + nx];
        beg1 = D.cstart[base];          end1 = D.cstart[base + 2];
    } else if (P.periodic[0] && kx == nx - 1) {
        beg0 = D.cstart[base + nx - 2]; end0 = D.cstart[base + nx];
        beg1 = D.cstart[base];          end1 = D.cstart[base + 1];
    } else {
        int x0 = kx > 0 ? kx - 1 : 0;
        int x1 = kx + 1 < nx ? kx + 1 : nx - 1;
        beg0 = D.cstart[base + x0];
        end0 = D.cstart[base + x1 + 1];
    }
}

// Periodic_X_Correction (Domain.cpp:110-115) of xij = x_i - x_j: a component longer than
// 2 Cellfac h crosses the period and is shifted by DomSize, in the direction set by the two
// particles' cell indices (those of this step's cell keys, as the reference's CC[]).
__device__ inline void periodic_xij(const Dev &D, const Params &P, double h, int key_i, int j,
                                    double &x0, double &x1, double &x2)
{
    const double lim = 2*P.cellfac*h;
    const bool c0 = P.periodic[0] && (x0 > lim || x0 < -lim);
    const bool c1 = P.periodic[1] && (x1 > lim || x1 < -lim);
    const bool c2 = P.periodic[2] && (x2 > lim || x2 < -lim);
    if (!(c0 || c1 || c2)) return;
    int ix, iy, iz, jx, jy, jz;
    decode_key(P, key_i, ix, iy, iz);
    decode_key(P, D.key[j], jx, jy, jz);
    if (c0) x0 = (ix > jx) ? x0 - P.domsize[0] : x0 + P.domsize[0];
    if (c1) x1 = (iy > jy) ? x1 - P.domsize[1] : x1 + P.domsize[1];
    if (c2) x2 = (iz > jz) ? x2 - P.domsize[2] : x2 + P.domsize[2];
}

// Whether particle i is P1 of its pair with j in the reference's search: P1 is the particle
// of the home cell; the partner sits in the same cell (then P1 is the one linked later, the
// larger original index) or at a forward offset of the half stencil (dz > 0, or dz = 0 and
// dy > 0, or dz = dy = 0 and dx > 0; Domain.cpp:824-945).  Without periodic directions this is
// the smaller cell key; across a period the offset is taken through the boundary.
__device__ inline int pair_first(const Params &P, int key_i, int key_j, int orig_i, int orig_j)
{
    if (!P.any_periodic) return (key_i < key_j) | ((key_i == key_j) & (orig_i > orig_j));
    int ix, iy, iz, jx, jy, jz;
    decode_key(P, key_i, ix, iy, iz);
    decode_key(P, key_j, jx, jy, jz);
    int d[3] = {jx - ix, jy - iy, jz - iz};
    for (int c = 0; c < 3; ++c)
        if (P.periodic[c]) d[c] = (d[c] > 1) ? d[c] - P.nc[c] : (d[c] < -1) ? d[c] + P.nc[c] : d[c];
    int zf = d[2] > 0, zs = d[2] == 0, yf = d[1] > 0, ys = d[1] == 0, xf = d[0] > 0, xs = d[0] == 0;
    return zf | (zs & yf) | (zs & ys & xf) | (zs & ys & xs & (orig_i > orig_j));
}


__device__ inline double pair_h(const ClassParams *cp, const Params &P, int ci, int cj)
{
    return P.uniform_h ? P.h_uniform : (cp[ci].h + cp[cj].h)/2;
}

// Whether the cells of two keys are the same or neighbours: the reference pairs particles of
// neighbouring cells only (MainNeighbourSearch, Domain.cpp:808-945).  Cells can be slightly
// smaller than the support (the CellNo rounding of CellInitiate), so a pair closer than the
// support may still be two cells apart and outside the reference's pair set.
__device__ __noinline__ bool keys_adjacent(const Params &P, int a, int b)
{
    int ax, ay, az, bx, by, bz;
    decode_key(P, a, ax, ay, az);
    decode_key(P, b, bx, by, bz);
    return abs(ax - bx) <= 1 && abs(ay - by) <= 1 && abs(az - bz) <= 1;
}

// A lane's list in slot groups of four (one 16-byte store per group): entry e of the lane at
// out[(e/4)*gs] component e%4.  Entries past the capacity are counted, not stored.
struct ListWriter {
    int4 *out;
    size_t gs;
    int cap, cnt;
    int w0, w1, w2, w3;
    __device__ ListWriter(int4 *o, size_t g, int c) : out(o), gs(g), cap(c), cnt(0), w0(0), w1(0), w2(0), w3(0) {}
    __device__ void push(int v)
    {
        int sl = cnt & 3;
        w0 = (sl == 0) ? v : w0; w1 = (sl == 1) ? v : w1; w2 = (sl == 2) ? v : w2; w3 = (sl == 3) ? v : w3;
        if (sl == 3 && cnt < cap) out[(size_t)(cnt >> 2)*gs] = make_int4(w0, w1, w2, w3);
        ++cnt;
    }
    __device__ void flush()
    {
        if ((cnt & 3) != 0 && cnt < cap) out[(size_t)(cnt >> 2)*gs] = make_int4(w0, w1, w2, w3);
    }
};

// ------------------------------------------------------------------------------------------
// wall pass: PrimaryComputeAcceleration (Domain.cpp:1020-1059, 1163-1206) for fixed targets
// ------------------------------------------------------------------------------------------
// SOLID = false compiles the fluid-only smoothing (no Material 2/3 particles in the run).
//
// REACH = 2 (list reuse in single-material runs): a step that rebuilds walks the 5^DIM cells
// around the wall's cell, records every free candidate of the wall's material within the
// support radius plus the skin in the wall's skin list, and adds the pairs of the neighbouring
// cells; the steps until the next build walk the list instead and keep the entries that are
// pairs of the reference's set at the current positions.  Between builds only the walls with
// skin neighbours at the build (Ctl::nwact of them) have pairs; the others keep what the build
// step wrote.
template <bool SOLID, int REACH>
__global__ void __launch_bounds__(kBlock) k_wall(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    __shared__ ClassParams cp[kMaxClasses];
    for (int k = threadIdx.x; k < P.nclasses; k += kBlock) cp[k] = D.cls[k];
    __syncthreads();
    const bool reuse = REACH == 2 && !D.ctl->rebuild;
    int t;
    if (reuse) {
        const int g = blockIdx.x*kBlock + threadIdx.x;
        if (g >= D.ctl->nwact) return;
        t = D.wact[g];
    } else {
        t = xcd_block()*kBlock + threadIdx.x;
        if (t >= P.nfixed) return;
    }
    int i = D.fixed_list[t];
    uint32_t fi = D.flags_b[i];
    int ci = flag_cls(fi);
    const ClassParams &C = cp[ci];
    int mat = SOLID ? flag_mat(fi) : 1;
    double4 pi = D.pos_b[i];
    int key_i = D.key[i];
    int orig_i = D.orig_b[i];
    int kx, ky, kz;
    decode_key(P, key_i, kx, ky, kz);
    double SK = 0.0, Pw = 0.0, n0 = 0.0, n1 = 0.0, n2 = 0.0;
    double sg[6] = {0, 0, 0, 0, 0, 0};
    int npair = 0;
    const double *G = P.gravity;
    const double sup = P.cellfac*P.hmax;
    const double sup2m = sup*sup*(1.0 + 1e-9);
    const double ox = P.blpf[0] + kx*P.csize[0], oy = P.blpf[1] + ky*P.csize[1], oz = P.blpf[2] + kz*P.csize[2];
    // a free neighbour j of the wall's material at xij = x_wall - x_j inside the support: the
    // pair's terms with the reference's orientation (P1 = smaller key / larger index, xij is
    // x_P1 - x_P2; the negation is exact).  The sign is formed without short-circuit branches:
    // hipcc (ROCm 7.2, -O3) dropped the same-cell negation when it was written as
    // `if (!(a || (b && c))) x = -x;`.
    auto add = [&](int j, double4 pj, double x0, double x1, double x2, double rij, double h) {
        ++npair;
        int kj = D.key[j];
        int oj = D.orig_b[j];
        int wall_first = pair_first(P, key_i, kj, orig_i, oj);
        double sgn = wall_first ? 1.0 : -1.0;
        x0 *= sgn; x1 *= sgn; x2 *= sgn;
        double K = kernel_w<double>(P.dim, P.kernel_type, rij/h, h);
        SK += K;
        if (mat < 3) {
            double4 vj = D.vel_b[j];
            Pw += vj.w * K + (G[0]*x0 + G[1]*x1 + G[2]*x2)*pj.w*K;
        }
        if (SOLID && mat > 1) {
            double sj[6];
            ld6(D.sig_b, j, sj);
            for (int q = 0; q < 6; ++q) sg[q] = sg[q] + sj[q]*K;
        }
        if (C.NoSlip) {
            double4 vj = D.vel_b[j];
            n0 += vj.x * K; n1 += vj.y * K; n2 += vj.z * K;
        }
    };
    if constexpr (REACH == 2) {
        ListWriter out(reinterpret_cast<int4 *>(D.wl) + t, (size_t)P.nfixed, 4*P.wl_cap);
        if (reuse) {
            // between builds: the wall's skin list, reduced to the reference's pairs
            const int c = D.wcnt[t], m = c < 4*P.wl_cap ? c : 4*P.wl_cap;
            const int *lst = D.wl;
            for (int e = 0; e < m; ++e) {
                const int j = lst[((size_t)(e >> 2)*P.nfixed + t)*4 + (e & 3)];
                const double4 pj = D.pos_b[j];
                const double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
                const double r2 = x0*x0 + x1*x1 + x2*x2;
                const double h = pair_h(cp, P, ci, flag_cls(D.flags_b[j]));
                const double rij = sqrt(r2);
                if (!((rij/h) <= P.cellfac)) continue;
                if (r2 > P.adj2 && !keys_adjacent(P, key_i, D.key[j])) continue;
                add(j, pj, x0, x1, x2, rij, h);
            }
        } else {
            // a build: the 5^DIM cells, culled against the reach, in cell order
            const double cx = P.csize[0], cy = P.csize[1], cz = P.csize[2];
            const double cull2 = P.reach2*(1.0 + 1e-9);
            const int ncell = (P.dim == 3) ? 125 : 25;
            for (int cc = 0; cc < ncell; ++cc) {
                const int dxk = (cc % 5) - 2, dyk = ((cc / 5) % 5) - 2, dzk = (P.dim == 3) ? (cc / 25) - 2 : 0;
                const int xx = kx + dxk, yy = ky + dyk, zz = kz + dzk;
                if (xx < 0 || xx >= P.nc[0] || yy < 0 || yy >= P.nc[1] || zz < 0 || zz >= P.nc[2]) continue;
                const int cell = xx + P.nc[0]*(yy + P.nc[1]*zz);
                if (D.cfree[cell] == 0) continue;
                double ex = (dxk < 0) ? pi.x - (ox + (dxk + 1)*cx) : (dxk > 0) ? (ox + dxk*cx) - pi.x : 0.0;
                double ey = (dyk < 0) ? pi.y - (oy + (dyk + 1)*cy) : (dyk > 0) ? (oy + dyk*cy) - pi.y : 0.0;
                double ez = (dzk < 0) ? pi.z - (oz + (dzk + 1)*cz) : (dzk > 0) ? (oz + dzk*cz) - pi.z : 0.0;
                ex = ex > 0.0 ? ex : 0.0; ey = ey > 0.0 ? ey : 0.0; ez = ez > 0.0 ? ez : 0.0;
                if (ex*ex + ey*ey + ez*ez > cull2) continue;
                const bool adj = dxk >= -1 && dxk <= 1 && dyk >= -1 && dyk <= 1 && dzk >= -1 && dzk <= 1;
                const int beg = D.cstart[cell], end = D.cstart[cell + 1];
                for (int j = beg; j < end; ++j) {
                    const uint32_t fj = D.flags_b[j];
                    if (!flag_free(fj) || flag_mat(fj) != mat) continue;
                    const double4 pj = D.pos_b[j];
                    const double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
                    const double r2 = x0*x0 + x1*x1 + x2*x2;
                    if (r2 <= P.reach2) out.push(j);
                    if (!adj) continue;
                    const double h = pair_h(cp, P, ci, flag_cls(fj));
                    const double rij = sqrt(r2);
                    if ((rij/h) <= P.cellfac) add(j, pj, x0, x1, x2, rij, h);
                }
            }
            out.flush();
            D.wcnt[t] = out.cnt;
            if (out.cnt > 4*P.wl_cap) atomicOr(&D.ctl->list_over, 2);
            // record the wall as active if it has skin neighbours (one atomic per wave)
            const bool rec = out.cnt != 0;
            const unsigned long long msk = __ballot(rec);
            const int lane = threadIdx.x & 63;
            int base = 0;
            if (msk) {
                const int leader = __ffsll((long long)msk) - 1;
                if (lane == leader) base = atomicAdd(&D.ctl->nwact, __popcll(msk));
                base = __builtin_amdgcn_readlane(base, leader);
            }
            if (rec) D.wact[base + __popcll(msk & ((1ull << lane) - 1ull))] = t;
        }
    } else {
    const int ncell = (P.dim == 3) ? 27 : 9;
    for (int cc = 0; cc < ncell; ++cc) {
        int dxk = (cc % 3) - 1, dyk = ((cc / 3) % 3) - 1, dzk = (P.dim == 3) ? (cc / 9) - 1 : 0;
        int xx = kx + dxk, yy = ky + dyk, zz = kz + dzk;
        if (P.periodic[0]) xx = (xx + P.nc[0]) % P.nc[0];
        if (P.periodic[1]) yy = (yy + P.nc[1]) % P.nc[1];
        if (P.periodic[2]) zz = (zz + P.nc[2]) % P.nc[2];
        if (xx < 0 || xx >= P.nc[0] || yy < 0 || yy >= P.nc[1] || zz < 0 || zz >= P.nc[2]) continue;
        int cell = xx + P.nc[0]*(yy + P.nc[1]*zz);
        if (D.cfree[cell] == 0) continue;                  // no free particle to gather from
        double ex = (dxk < 0) ? pi.x - ox : (dxk > 0) ? ox + P.csize[0] - pi.x : 0.0;
        double ey = (dyk < 0) ? pi.y - oy : (dyk > 0) ? oy + P.csize[1] - pi.y : 0.0;
        double ez = (dzk < 0) ? pi.z - oz : (dzk > 0) ? oz + P.csize[2] - pi.z : 0.0;
        ex = ex > 0.0 ? ex : 0.0; ey = ey > 0.0 ? ey : 0.0; ez = ez > 0.0 ? ez : 0.0;
        if (ex*ex + ey*ey + ez*ez > sup2m) continue;       // the cell lies outside the support
        int beg = D.cstart[cell], end = D.cstart[cell + 1];
        // candidates in groups of 4: the group's loads are issued together (slots past the
        // range re-read `beg`, which is in bounds, and are skipped), then processed in order
        for (int j0 = beg; j0 < end; j0 += 4) {
          uint32_t fg[4];
          double4 pg[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int jj = (j0 + u < end) ? j0 + u : beg;
            fg[u] = D.flags_b[jj];
            pg[u] = D.pos_b[jj];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int j = j0 + u;
            if (j >= end) continue;
            uint32_t fj = fg[u];
            if (!flag_free(fj) || flag_mat(fj) != mat) continue;
            double4 pj = pg[u];
            double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
            double h = pair_h(cp, P, ci, flag_cls(fj));
            if (P.any_periodic) periodic_xij(D, P, h, key_i, j, x0, x1, x2);
            double rij = sqrt(x0*x0 + x1*x1 + x2*x2);
            if (!((rij/h) <= P.cellfac)) continue;      // outside the reference's pair set
            add(j, pj, x0, x1, x2, rij, h);
        }
        }
    }
    }
    // StartAcceleration resets fixed P, NSv, Sigma (Domain.cpp:990-996); then divide by SumKernel
    if (SK != 0.0) {
        if (mat < 3) Pw = Pw/SK;
        if (SOLID && mat > 1) for (int q = 0; q < 6; ++q) sg[q] = sg[q]*(1.0/SK);
        if (C.NoSlip) { n0 = n0/SK; n1 = n1/SK; n2 = n2/SK; }
        else { n0 = n1 = n2 = 0.0; }
    } else {
        Pw = 0.0; n0 = n1 = n2 = 0.0;
        for (int q = 0; q < 6; ++q) sg[q] = 0.0;
    }
    double4 vi = D.vel_b[i];
    vi.w = Pw;
    D.vel_b[i] = vi;
    D.shep[i] = make_double2(0.0, SK);
    // a fixed particle in a pair gets ZWab = 1 and keeps S = 0 (Interaction.cpp:158, 350)
    if (P.swi1) D.swz[i] = make_double2(0.0, npair > 0 ? 1.0 : 0.0);
    // wall density and sound speed as the free partner sees them (CalcForce11 :41-61, :69-70)
    double dw = density_solid(P.eos_eq_m[mat], P.eos_cs_m[mat], P.eos_p0_m[mat], Pw, P.eos_rho0_m[mat]);
    double Cw = sound_speed(P.eos_eq_m[mat], P.eos_cs_m[mat], dw, P.eos_rho0_m[mat]);
    double2 mm = D.mmu_b[i];
    D.aux[i] = make_double4(mm.x, mm.y, Cw, Pw/(dw*dw));
    D.wal[i] = make_double4(n0, n1, n2, dw);
    if (SOLID) {
        const double z6[6] = {0, 0, 0, 0, 0, 0};
        st6(D.sig_b, i, mat > 1 ? sg : z6);
        st6(D.tau_b, i, z6);
        double t6[6] = {0, 0, 0, 0, 0, 0};
        if (mat > 1 && C.TI > 0.0 && SK != 0.0) tir_tensor(P.dim, sg, C.TI, pi.w, t6);
        st6(D.tir, i, t6);
    }
}

// ------------------------------------------------------------------------------------------
// pair membership and cell codes of the fast fluid path
// ------------------------------------------------------------------------------------------


// Lists cover the W^DIM cells around the target's cell, W = 2 REACH + 1: REACH = 1 for lists
// of one step (support radius), REACH = 2 for reused lists (support radius plus a skin of at
// most one cell).  A list entry is j | code << shift with the cell code
// (dx+REACH) + W (dy+REACH) + W^2 (dz+REACH) of the neighbour's cell (dz = 0 in 2-D).
template <int REACH>
struct Codes {
    static constexpr int W = 2*REACH + 1;
    static constexpr int shift = (REACH == 1) ? 27 : 25;
    static constexpr unsigned mask = (1u << shift) - 1u;
};
template <int REACH>
__device__ inline int nl_entry(int j, int code) { return (int)((unsigned)j | ((unsigned)code << Codes<REACH>::shift)); }
template <int REACH>
__device__ inline int nl_index(int w) { return (int)((unsigned)w & Codes<REACH>::mask); }
template <int REACH>
__device__ inline int nl_code(int w) { return (int)((unsigned)w >> Codes<REACH>::shift); }
template <int REACH>
__device__ inline void code_offset(int c, int &dx, int &dy, int &dz)
{
    constexpr int W = Codes<REACH>::W;
    dx = c % W - REACH;
    dy = (c / W) % W - REACH;
    dz = c / (W*W) - REACH;
}
// Wall pass of the fast fluid configuration (cubic spline, uniform h, fluid only): the same
// smoothing as k_wall with the pair test of in_support, cells culled against the support
// sphere, and cells without free particles skipped.  Four lanes per fixed target, lane q
// taking neighbour cells q, q+4, ...; their sums meet in a fixed order ((q0 + q1) + (q2 + q3)),
// so results stay deterministic.
//
// REACH = 2 (list reuse): a step that rebuilds walks the 5^DIM cells around the wall's cell and
// records, per lane, every free candidate within the support radius plus the skin; the steps
// until the next build walk these sub-lists instead of the cells.
template <int DIM, int REACH>
__global__ void __launch_bounds__(kBlock) k_wall_fast(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    // Between builds only the walls with skin neighbours at the build (the active walls) can
    // have pairs; the others keep the values the build step wrote.  Active walls are taken in
    // plain block order, so that their blocks spread over all XCDs.
    const bool reuse = REACH == 2 && !D.ctl->rebuild;
    const int gl = (reuse ? (int)blockIdx.x : xcd_block())*kBlock + threadIdx.x;
    const int quad = gl & 3;
    bool act;
    int t;
    if (reuse) {
        act = (gl >> 2) < D.ctl->nwact;
        t = act ? D.wact[gl >> 2] : 0;
    } else {
        t = gl >> 2;
        act = t < P.nfixed;
    }
    const int gt = 4*t + quad;                               // the lane's sub-list
    int i = act ? D.fixed_list[t] : 0;
    uint32_t fi = D.flags_b[i];
    const bool noslip = flag_noslip(fi);
    double4 pi = D.pos_b[i];
    const int key_i = D.key[i];
    const int orig_i = D.orig_b[i];
    int kx, ky, kz;
    decode_key(P, key_i, kx, ky, kz);
    double SK = 0.0, Pw = 0.0, n0 = 0.0, n1 = 0.0, n2 = 0.0;
    const double gx = P.gravity[0], gy = P.gravity[1], gz = P.gravity[2];
    const double h = P.h_uniform, kw = P.kw;
    // contribution of free neighbour j at squared distance r2 (a pair of the reference's set)
    auto add = [&](int j, double4 pj, double x0, double x1, double x2, double r2) {
        double q = sqrt(r2)/h;
        // orientation of the reference's pair (P1 = smaller key / larger index), formed
        // branch-free as in k_wall
        int kj = D.key[j];
        int oj = D.orig_b[j];
        int wall_first = (key_i < kj) | ((key_i == kj) & (orig_i > oj));
        double sgn = wall_first ? 1.0 : -1.0;
        x0 *= sgn; x1 *= sgn; x2 *= sgn;
        double K;
        if (q < 1.0) K = kw*(1.0-(3.0/2.0)*q*q+(3.0/4.0)*q*q*q);
        else if (q < 2.0) K = kw*((1.0/4.0)*(2.0-q)*(2.0-q)*(2.0-q));
        else K = 0.0;
        SK += K;
        double4 vj = D.vel_b[j];
        Pw += vj.w * K + (gx*x0 + gy*x1 + gz*x2)*pj.w*K;
        if (noslip) { n0 += vj.x * K; n1 += vj.y * K; n2 += vj.z * K; }
    };
    const size_t gs = 4*(size_t)P.nfixed;                   // sub-list slot group stride
    if (reuse) {
        // between builds: the lane's sub-list, reduced to the reference's pairs
        const int cap = P.wl_cap;
        int c = act ? D.wcnt[gt] : 0;
        int m = c < cap ? c : cap;
        const int4 *lst = reinterpret_cast<const int4 *>(D.wl) + gt;
        int4 w = make_int4(0, 0, 0, 0);
        for (int e = 0; e < m; ++e) {
            int sl = e & 3;
            if (sl == 0) w = lst[(size_t)(e >> 2)*gs];
            int j = (sl == 0) ? w.x : (sl == 1) ? w.y : (sl == 2) ? w.z : w.w;
            double4 pj = D.pos_b[j];
            double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
            double r2 = x0*x0 + x1*x1 + x2*x2;
            bool in = in_support(P, r2);
            if (in && r2 > P.adj2) in = keys_adjacent(P, key_i, D.key[j]);
            if (in) add(j, pj, x0, x1, x2, r2);
        }
    } else {
        constexpr int W = Codes<REACH>::W;
        const double cx = P.csize[0], cy = P.csize[1], cz = P.csize[2];
        const double cull2 = ((REACH == 2) ? P.reach2 : P.sup2_hi)*(1.0 + 1e-9);
        const double ox = P.blpf[0] + kx*cx, oy = P.blpf[1] + ky*cy;
        const double oz = (DIM == 3) ? P.blpf[2] + kz*cz : 0.0;
        const int ncell = act ? ((DIM == 3) ? W*W*W : W*W) : 0;
        const int cap = P.wl_cap;
        int4 *out = reinterpret_cast<int4 *>(D.wl) + gt;
        int wc = 0, w0 = 0, w1 = 0, w2 = 0, w3 = 0;
        for (int cc = quad; cc < ncell; cc += 4) {
            int dxk = (cc % W) - REACH, dyk = ((cc / W) % W) - REACH, dzk = (DIM == 3) ? (cc / (W*W)) - REACH : 0;
            int xx = kx + dxk, yy = ky + dyk, zz = kz + dzk;
            if (xx < 0 || xx >= P.nc[0] || yy < 0 || yy >= P.nc[1] || zz < 0 || zz >= P.nc[2]) continue;
            int cell = xx + P.nc[0]*(yy + P.nc[1]*zz);
            if (D.cfree[cell] == 0) continue;                  // no free particle to gather from
            // distance from the wall to the cell's box
            double ex = (dxk < 0) ? pi.x - (ox + (dxk + 1)*cx) : (dxk > 0) ? (ox + dxk*cx) - pi.x : 0.0;
            double ey = (dyk < 0) ? pi.y - (oy + (dyk + 1)*cy) : (dyk > 0) ? (oy + dyk*cy) - pi.y : 0.0;
            double ez = (dzk < 0) ? pi.z - (oz + (dzk + 1)*cz) : (dzk > 0) ? (oz + dzk*cz) - pi.z : 0.0;
            ex = ex > 0.0 ? ex : 0.0; ey = ey > 0.0 ? ey : 0.0; ez = ez > 0.0 ? ez : 0.0;
            if (ex*ex + ey*ey + ez*ez > cull2) continue;       // the cell lies outside the reach
            const bool adj = dxk >= -1 && dxk <= 1 && dyk >= -1 && dyk <= 1 && dzk >= -1 && dzk <= 1;
            int beg = D.cstart[cell], end = D.cstart[cell + 1];
            for (int j = beg; j < end; ++j) {
                if (!flag_free(D.flags_b[j])) continue;
                double4 pj = D.pos_b[j];
                double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
                double r2 = x0*x0 + x1*x1 + x2*x2;
                if (REACH == 2 && r2 <= P.reach2) {
                    int sl = wc & 3;
                    w0 = (sl == 0) ? j : w0; w1 = (sl == 1) ? j : w1; w2 = (sl == 2) ? j : w2; w3 = (sl == 3) ? j : w3;
                    if (sl == 3 && wc < cap) out[(size_t)(wc >> 2)*gs] = make_int4(w0, w1, w2, w3);
                    ++wc;
                }
                if (adj && in_support(P, r2)) add(j, pj, x0, x1, x2, r2);
            }
        }
        if (REACH == 2) {
            if (act) {
                if ((wc & 3) != 0 && wc < cap) out[(size_t)(wc >> 2)*gs] = make_int4(w0, w1, w2, w3);
                D.wcnt[gt] = wc;
                if (wc > cap) atomicOr(&D.ctl->list_over, 2);
            }
            // record the wall as active if any of its four lanes has skin neighbours (one
            // atomic per wave; lane 4k stands for wall k of the wave)
            int any = act ? wc : 0;
            any |= __shfl_xor(any, 1, 64);
            any |= __shfl_xor(any, 2, 64);
            const bool rec = act && quad == 0 && any != 0;
            const unsigned long long m = __ballot(rec);
            const int lane = threadIdx.x & 63;
            int base = 0;
            if (m) {
                const int leader = __ffsll((long long)m) - 1;
                if (lane == leader) base = atomicAdd(&D.ctl->nwact, __popcll(m));
                base = __builtin_amdgcn_readlane(base, leader);
            }
            if (rec) D.wact[base + __popcll(m & ((1ull << lane) - 1ull))] = t;
        }
    }
    SK += __shfl_xor(SK, 1, 64); Pw += __shfl_xor(Pw, 1, 64);
    n0 += __shfl_xor(n0, 1, 64); n1 += __shfl_xor(n1, 1, 64); n2 += __shfl_xor(n2, 1, 64);
    SK += __shfl_xor(SK, 2, 64); Pw += __shfl_xor(Pw, 2, 64);
    n0 += __shfl_xor(n0, 2, 64); n1 += __shfl_xor(n1, 2, 64); n2 += __shfl_xor(n2, 2, 64);
    if (!act || quad != 0) return;
    if (SK != 0.0) {
        Pw = Pw/SK;
        if (noslip) { n0 = n0/SK; n1 = n1/SK; n2 = n2/SK; }
        else { n0 = n1 = n2 = 0.0; }
    } else {
        Pw = 0.0; n0 = n1 = n2 = 0.0;
    }
    double4 vi = D.vel_b[i];
    vi.w = Pw;
    D.vel_b[i] = vi;
    D.shep[i] = make_double2(0.0, SK);
    double dw = density_solid(P.eos_eq, P.eos_cs, P.eos_p0, Pw, P.eos_rho0);
    double Cw = sound_speed(P.eos_eq, P.eos_cs, dw, P.eos_rho0);
    double2 mm = D.mmu_b[i];
    D.aux[i] = make_double4(mm.x, mm.y, Cw, Pw/(dw*dw));
    D.wal[i] = make_double4(n0, n1, n2, dw);
    const double fpm = D.cls[flag_cls(fi)].FPMassC;
    const int ci = flag_cls(fi);
    if (P.mixed32) {
        double o0, o1, o2;
        cell_origin(P, D.key_b[i], o0, o1, o2);      // the cell of slot i at the last sort
        D.f0[i] = make_float4((float)(pi.x - o0), (float)(pi.y - o1), (float)(pi.z - o2), (float)dw);
        D.f1[i] = make_float4((float)vi.x, (float)vi.y, (float)vi.z, (float)(Pw/(dw*dw)));
        D.f2[i] = make_float4((float)Cw, (float)(-1 - ci), (float)fpm, (float)(1.0/dw));
        if (noslip) D.fw[i] = make_float4((float)(2.0*vi.x-n0), (float)(2.0*vi.y-n1), (float)(2.0*vi.z-n2), 0.0f);
        else D.fw[i] = make_float4((float)vi.x, (float)vi.y, (float)vi.z, 0.0f);
    } else {
        D.g0[i] = make_double4(pi.x, pi.y, pi.z, dw);
        D.g1[i] = make_double4(vi.x, vi.y, vi.z, Pw/(dw*dw));
        D.g2[i] = make_double4(Cw, (double)(-1 - ci), fpm, 1.0/dw);
        if (noslip) D.gw[i] = make_double4(2.0*vi.x-n0, 2.0*vi.y-n1, 2.0*vi.z-n2, 0.0);
        else D.gw[i] = make_double4(vi.x, vi.y, vi.z, 0.0);
    }
}

// ------------------------------------------------------------------------------------------
// force pass: CalcForce11 / CalcForce2233 gathered per free particle
// ------------------------------------------------------------------------------------------
// Per-block sums of a force pass, added to the control block with one atomic per quantity and
// block: in-support pair counts (free-free directed, free-fixed) and the minimum dt candidate
// (bits of a positive double, so unsigned order is double order).  Integer sums and a minimum
// are exact in any order, so the result is deterministic.  (A last-block fold of per-block
// partials would need a device-scope fence per block, which on this chip writes back and
// invalidates the L2 the gather lives in.)
__device__ inline void force_partials(const Dev &D, unsigned long long dtb, unsigned nff, unsigned nfw)
{
    __shared__ unsigned long long red[3][16];
    unsigned long long sff = nff, sfw = nfw, sdt = dtb;
    for (int o = 32; o > 0; o >>= 1) {
        sff += __shfl_down(sff, o, 64);
        sfw += __shfl_down(sfw, o, 64);
        unsigned long long t = __shfl_down(sdt, o, 64);
        sdt = t < sdt ? t : sdt;
    }
    int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = sff; red[1][w] = sfw; red[2][w] = sdt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < (int)(blockDim.x >> 6); ++k) {
            sff += red[0][k]; sfw += red[1][k];
            sdt = red[2][k] < sdt ? red[2][k] : sdt;
        }
        Ctl *c = D.ctl;
        if (sff) atomicAdd(&c->pairs_ff, sff);
        if (sfw) atomicAdd(&c->pairs_fw, sfw);
        if (sdt != ~0ull) atomicMin(&c->dtmin_bits, sdt);
    }
}

// The acceleration (x, y, z) and density rate (w) a free target's pair sums start from:
// StartAcceleration's Gravity, plus the body force of the body-force hook; after a host hook
// before the pair sums, the values the host wrote into acc for this step (psph_hook_accel).
__device__ inline double4 a_start(const Dev &D, const Params &P, int i)
{
    if (D.ctl->acc_init) return D.acc[i];
    return make_double4(P.afree[0], P.afree[1], P.afree[2], 0.0);
}

// Per-target sums of a force pass.  T = float for the FP32-mixed soil tier, whose continuity
// sum carries a Kahan compensation term ddc.
template <class T>
struct AccT {
    T a0, a1, a2, dd, ddc;
    T vx0, vx1, vx2, zw, sd;
    T sw;                 // S of the surface-erosion soil-water interaction (SWIType 1)
    T sr[6];
    T rr[3];
};
using Acc = AccT<double>;

// CalcForce11 (Interaction.cpp:25-184) contribution of neighbour j to free particle i as P1.
// tw (when not null) receives the pair's temp[3], temp1, mi, di and dj: what the reference adds to
// P2 as P2->a -= mi temp, P2->dDensity += mi (dj/di) temp1 (Interaction.cpp:166-167).
__device__ __forceinline__ void pair11(const Dev &D, const Params &P, const ClassParams *cp, int i, int j,
                              uint32_t fi, double4 pi, double4 vi, double4 ai, Acc &A, int shep_on,
                              double *tw = nullptr)
{
    uint32_t fj = D.flags_b[j];
    int ci = flag_cls(fi), cj = flag_cls(fj);
    const ClassParams &Ci_ = cp[ci];
    const ClassParams &Cj_ = cp[cj];
    double4 pj = D.pos_b[j];
    double4 vj = D.vel_b[j];
    double4 aj = D.aux[j];
    bool freej = flag_free(fj);
    double h = pair_h(cp, P, ci, cj);
    double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
    if (P.any_periodic) periodic_xij(D, P, h, D.key[i], j, x0, x1, x2);
    double rij = sqrt(x0*x0 + x1*x1 + x2*x2);
    double Alpha = (Ci_.Alpha + Cj_.Alpha)/2.0;
    double Beta = (Ci_.Beta + Cj_.Beta)/2.0;
    double u0 = vi.x - vj.x, u1 = vi.y - vj.y, u2 = vi.z - vj.z;
    double di = pi.w, mi = ai.x;
    double dj, mj;
    double4 wj = make_double4(0.0, 0.0, 0.0, 0.0);
    if (!freej) {
        wj = D.wal[j];
        dj = wj.w;
        mj = Cj_.FPMassC * mi;
    } else {
        dj = pj.w;
        mj = aj.x;
    }
    double q = rij/h;
    double GK = grad_kernel<double>(P.dim, P.kernel_type, q, h);
    double K = kernel_w<double>(P.dim, P.kernel_type, q, h);
    double PIij = 0.0;
    if (Alpha != 0.0 || Beta != 0.0) {
        double Cii = ai.z, Cjj = aj.z;
        double dvx = u0*x0 + u1*x1 + u2*x2;
        double MUij = h*dvx/(rij*rij+0.01*h*h);
        if (dvx < 0) PIij = (-Alpha*0.5*(Cii+Cjj)*MUij+Beta*MUij*MUij)/(0.5*(di+dj));
    }
    double TIij = 0.0;
    if (Ci_.TI > 0.0 || Cj_.TI > 0.0) {
        double Ri = 0.0, Rj = 0.0;
        if ((vi.w+vj.w) < 0.0) {
            if (vi.w < 0.0) Ri = -vi.w/(di*di);
            if (vj.w < 0.0) Rj = -vj.w/(dj*dj);
        }
        TIij = (Ci_.TI*Ri + Cj_.TI*Rj)*pow((K/kernel_w<double>(P.dim, P.kernel_type, (Ci_.TIInitDist + Cj_.TIInitDist)/(2.0*h), h)),(Ci_.TIn+Cj_.TIn)/2.0);
    }
    double VI0 = 0.0, VI1 = 0.0, VI2 = 0.0;
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    bool sr_on = false;
    double s00 = 0, s01 = 0, s02 = 0, s11 = 0, s12 = 0, s22 = 0;
    if ((Ci_.NoSlip || Cj_.NoSlip) || freej) {
        double Mu;
        if (freej) {
            b0 = u0; b1 = u1; b2 = u2;
            Mu = 2.0*ai.y*aj.y/(ai.y+aj.y);
        } else {
            b0 = vi.x - (2.0*vj.x-wj.x);
            b1 = vi.y - (2.0*vj.y-wj.y);
            b2 = vi.z - (2.0*vj.z-wj.z);
            Mu = ai.y;
        }
        if (Ci_.T0 > 0.0 || Cj_.T0 > 0.0 || (Ci_.LES && Cj_.LES)) {
            sr_on = true;
            double g = -GK;
            s00 = (2.0*b0*x0)*g;
            s01 = (b0*x1+b1*x0)*g;
            s02 = (b0*x2+b2*x0)*g;
            s11 = (2.0*b1*x1)*g;
            s12 = (b1*x2+b2*x1)*g;
            s22 = (2.0*b2*x2)*g;
        }
        viscous_force<double>(P.visc_eq, VI0, VI1, VI2, Mu, di, dj, GK, b0, b1, b2, P.dim,
                              P.kernel_type, rij, h, x0, x1, x2, u0, u1, u2);
    }
    if (P.xsph != 0.0 && freej) {
        double sx = P.xsph*mj/(0.5*(di+dj))*K;
        A.vx0 += sx*-u0; A.vx1 += sx*-u1; A.vx2 += sx*-u2;
    }
    double s;
    if (P.gradient_type == 0) s = -1.0*( ai.w + aj.w + PIij + TIij ) * GK;
    else s = -1.0*( (vi.w + vj.w)/(di*dj) + PIij + TIij ) * GK;
    double t0 = s*x0 + VI0, t1 = s*x1 + VI1, t2 = s*x2 + VI2;
    if (P.dim == 2) t2 = 0.0;
    double g0 = GK*x0, g1 = GK*x1, g2 = GK*x2;
    double temp1 = u0*g0 + u1*g1 + u2*g2;
    if (tw) { tw[0] = t0; tw[1] = t1; tw[2] = t2; tw[3] = temp1; tw[4] = mi; tw[5] = di; tw[6] = dj; }
    A.a0 += mj * t0; A.a1 += mj * t1; A.a2 += mj * t2;
    A.dd += mj * (di/dj) * temp1;
    if (sr_on && (Ci_.T0 > 0.0 || Ci_.LES)) {
        double fq = mj/dj;
        A.sr[0] = A.sr[0] + s00*fq; A.sr[1] = A.sr[1] + s01*fq; A.sr[2] = A.sr[2] + s02*fq;
        A.sr[3] = A.sr[3] + s11*fq; A.sr[4] = A.sr[4] + s12*fq; A.sr[5] = A.sr[5] + s22*fq;
    }
    if (P.swi1) A.sw = A.sw + mj/dj*b0*x1*-GK;         // Interaction.cpp:154
    if (shep_on) {
        A.zw += mj/dj* K;
        A.sd += mj*K;
    }
}

// CalcForce2233 (Interaction.cpp:186-377) contribution of neighbour j to free particle i as P1,
// with the pair arithmetic and sums in T (double: the parity tier; float: the FP32-mixed tier,
// positions differenced in FP64 and rounded, Kahan-compensated continuity sum).
// tiw (when not null): 1/W(dp) of each class pair (k_force's table, see ti_table), the kernel
// value at the artificial-stress spacing that the reference evaluates per pair.
template <class T>
__device__ __forceinline__ void pair2233(const Dev &D, const Params &P, const ClassParams *cp, int i, int j,
                                uint32_t fi, double4 pi, double4 vi, double4 ai, const double *sgi,
                                const double *tri, AccT<T> &A, int shep_on, double *tw = nullptr,
                                const double *tiw = nullptr)
{
    // divisions by refined reciprocals and fused products (a few ulp per term, inside the bars)
#pragma clang fp contract(fast)
    uint32_t fj = D.flags_b[j];
    // a free neighbour's state from its pair record (runs of free soil particles, Dev::srec):
    // its first 64 B are read with the flags, the tensors where they are used (the same two
    // cache lines); a fixed neighbour's state from the columns
    const double2 *rp = P.srec ? D.srec + 10*(size_t)j : nullptr;
    double2 rec[4];
    if (P.srec) {
#pragma unroll
        for (int q = 0; q < 4; ++q) rec[q] = rp[q];
    }
    int ci = flag_cls(fi), cj = flag_cls(fj);
    const ClassParams &Ci_ = cp[ci];
    const ClassParams &Cj_ = cp[cj];
    bool freej = flag_free(fj);
    const bool use_rec = P.srec && freej;
    double4 pj, vj, aj;
    if (use_rec) {
        pj = make_double4(rec[0].x, rec[0].y, rec[1].x, rec[1].y);
        vj = make_double4(rec[2].x, rec[2].y, rec[3].x, 0.0);
        aj = make_double4(rec[3].y, 0.0, 0.0, 0.0);
    } else {
        pj = D.pos_b[j];
        vj = D.vel_b[j];
        aj = D.aux[j];
    }
    T h = (T)pair_h(cp, P, ci, cj);
    double dx0 = pi.x - pj.x, dx1 = pi.y - pj.y, dx2 = pi.z - pj.z;
    if (P.any_periodic) periodic_xij(D, P, pair_h(cp, P, ci, cj), D.key[i], j, dx0, dx1, dx2);
    T x0 = (T)dx0, x1 = (T)dx1, x2 = (T)dx2;
    T rij = sqrt(x0*x0 + x1*x1 + x2*x2);
    T Alpha = (T)((Ci_.Alpha + Cj_.Alpha)/2.0);
    T Beta = (T)((Ci_.Beta + Cj_.Beta)/2.0);
    T di = (T)pi.w, mi = (T)ai.x, dj, mj;
    bool soil = (Ci_.Material*Cj_.Material == 9);
    double4 wj = make_double4(0.0, 0.0, 0.0, 0.0);
    if (!freej) wj = D.wal[j];
    if (soil) {
        if (!freej) { dj = (T)pi.w; mj = (T)Cj_.FPMassC * mi; } else { dj = (T)pj.w; mj = (T)aj.x; }
    } else {
        if (!freej) { dj = (T)wj.w; mj = (T)Cj_.FPMassC * mi; } else { dj = (T)pj.w; mj = (T)aj.x; }
    }
    T u0 = (T)(vi.x - vj.x), u1 = (T)(vi.y - vj.y), u2 = (T)(vi.z - vj.z);
    T q = rij*rcpx(h);
    T GK = grad_kernel<T>(P.dim, P.kernel_type, q, h);
    T K = kernel_w<T>(P.dim, P.kernel_type, q, h);
    const T rdj = rcpx(dj), rdm = rcpx((T)0.5*(di+dj));
    T PIs = 0.0;
    bool pi_on = false;
    if (Alpha != (T)0.0 || Beta != (T)0.0) {
        T dvx = u0*x0 + u1*x1 + u2*x2;
        T MUij = h*dvx*rcpx(rij*rij+(T)0.01*h*h);
        T Cij;
        if (soil) Cij = (T)(0.5*(Ci_.Cs+Cj_.Cs));
        else Cij = (T)(0.5*(ai.z+aj.z));
        if (dvx < (T)0) { PIs = (Alpha*Cij*MUij+Beta*MUij*MUij)*rdm; pi_on = true; }
    }
    T TI6[6] = {0, 0, 0, 0, 0, 0};
    if (Ci_.TI > 0.0 || Cj_.TI > 0.0) {
        // pow(f, n) of the reference as exp(n log f): f = W(r)/W(dp) >= 0 (f = 0 gives 0 either
        // way), |n log f| stays below ~10, so the result is within ~1e-15 relative of pow's at
        // about half of its instructions
        const T rwdp = tiw ? (T)tiw[ci*8 + cj]
                           : rcpx(kernel_w<T>(P.dim, P.kernel_type,
                                              (T)((Ci_.TIInitDist + Cj_.TIInitDist)/(2.0*(double)h)), h));
        T sT = exp((T)((Ci_.TIn+Cj_.TIn)/2.0)*log(K*rwdp));
        double trj[6];
        if (use_rec) { for (int q = 0; q < 3; ++q) { const double2 t = rp[7 + q]; trj[2*q] = t.x; trj[2*q+1] = t.y; } }
        else ld6(D.tir, j, trj);
        for (int k = 0; k < 6; ++k) TI6[k] = (T)(tri[k]+trj[k])*sT;
    }
    T b0, b1, b2;
    if (freej) { b0 = u0; b1 = u1; b2 = u2; }
    else if (Ci_.NoSlip || Cj_.NoSlip) {
        b0 = (T)(vi.x - (2.0*vj.x-wj.x)); b1 = (T)(vi.y - (2.0*vj.y-wj.y)); b2 = (T)(vi.z - (2.0*vj.z-wj.z));
    } else {
        double4 vbj = D.vh_b[j];    // fixed particles keep vb in the half-step column
        b0 = (T)(vi.x - vbj.x); b1 = (T)(vi.y - vbj.y); b2 = (T)(vi.z - vbj.z);
    }
    T f05 = (T)-0.5 * GK;
    T s00 = ((T)2.0*b0*x0)*f05, s01 = (b0*x1+b1*x0)*f05, s02 = (b0*x2+b2*x0)*f05;
    T s11 = ((T)2.0*b1*x1)*f05, s12 = (b1*x2+b2*x1)*f05, s22 = ((T)2.0*b2*x2)*f05;
    T r01 = (b0*x1-b1*x0)*f05, r02 = (b0*x2-b2*x0)*f05, r12 = (b1*x2-b2*x1)*f05;
    if (P.xsph != 0.0 && freej) {
        T sx = (T)P.xsph*mj*rdm*K;
        A.vx0 += sx*-u0; A.vx1 += sx*-u1; A.vx2 += sx*-u2;
    }
    // M = Sigma_i/di^2 + Sigma_j/dj^2 + PI I + TI, then temp = (GK xij)^T M (matvec.h:806-811)
    T Mf[9];
    T sgiT[6], sgj[6];
    double sgjd[6];
    if (use_rec) { for (int q = 0; q < 3; ++q) { const double2 t = rp[4 + q]; sgjd[2*q] = t.x; sgjd[2*q+1] = t.y; } }
    else ld6(D.sig_b, j, sgjd);
    for (int k = 0; k < 6; ++k) { sgiT[k] = (T)sgi[k]; sgj[k] = (T)sgjd[k]; }
    T Mi[9], Mj[9], T9[9];
    sym_to_full(sgiT, Mi);
    sym_to_full(sgj, Mj);
    sym_to_full(TI6, T9);
    if (P.gradient_type == 0) {
        T fa = rcpx(di*di), fb = rdj*rdj;
        for (int k = 0; k < 9; ++k) {
            T pim = pi_on ? ((k == 0 || k == 4 || k == 8) ? (T)1.0 : (T)0.0)*PIs : (T)0.0;
            Mf[k] = Mi[k]*fa + Mj[k]*fb + pim + T9[k];
        }
    } else {
        T fa = rcpx(di*dj);
        for (int k = 0; k < 9; ++k) {
            T pim = pi_on ? ((k == 0 || k == 4 || k == 8) ? (T)1.0 : (T)0.0)*PIs : (T)0.0;
            Mf[k] = (Mi[k] + Mj[k])*fa + pim + T9[k];
        }
    }
    T g0 = GK*x0, g1 = GK*x1, g2 = GK*x2;
    T t0 = g0*Mf[0] + g1*Mf[3] + g2*Mf[6];
    T t1 = g0*Mf[1] + g1*Mf[4] + g2*Mf[7];
    T t2 = g0*Mf[2] + g1*Mf[5] + g2*Mf[8];
    if (P.dim == 2) t2 = (T)0.0;
    T temp1 = u0*g0 + u1*g1 + u2*g2;
    if (tw) {
        tw[0] = (double)t0; tw[1] = (double)t1; tw[2] = (double)t2; tw[3] = (double)temp1;
        tw[4] = (double)mi; tw[5] = (double)di; tw[6] = (double)dj;
    }
    A.a0 += mj * t0; A.a1 += mj * t1; A.a2 += mj * t2;
    if constexpr (sizeof(T) == 4) {
        T c = mj * (di*rdj) * temp1;
        T y = c - A.ddc;
        T tt = A.dd + y;
        A.ddc = (tt - A.dd) - y;
        A.dd = tt;
    } else {
        A.dd += mj * (di*rdj) * temp1;
    }
    T fq = mj*rdj;
    A.zw += fq* K;
    if (P.swi1) A.sw = A.sw + fq*b0*x1*-GK;            // Interaction.cpp:347
    A.sr[0] = A.sr[0] + s00*fq; A.sr[1] = A.sr[1] + s01*fq; A.sr[2] = A.sr[2] + s02*fq;
    A.sr[3] = A.sr[3] + s11*fq; A.sr[4] = A.sr[4] + s12*fq; A.sr[5] = A.sr[5] + s22*fq;
    A.rr[0] = A.rr[0] + r01*fq; A.rr[1] = A.rr[1] + r02*fq; A.rr[2] = A.rr[2] + r12*fq;
    if (shep_on) A.sd += mj* K;
}


// Pair lists of the general kernels: each free target's pairs of this step, in the order the
// force kernel's row walk meets them (rows of three cells, y then z; a row across a period in
// two ranges), by the reference's test rij/h <= Cellfac on the periodic-corrected xij.  The
// candidates' tests run here at high occupancy instead of inside the force kernel's
// register-heavy loop.  Entries beyond the capacity are counted (ncnt), raise Ctl::nl_over (the
// host grows the lists) and are found again by the force kernel.  A cross-material candidate
// raises Ctl::xmat.
//
// REACH = 2 (list reuse in single-material runs): a step that rebuilds lists every candidate of
// the target's material within the support radius plus the skin, over the 5^DIM cells around
// the target's cell (rows of five cells, y then z); the steps until the next build keep the
// lists, and the force kernel keeps the entries that are pairs of the reference's set at the
// current positions.  A list that overflows makes every step rebuild (Ctl::list_over) until the
// host has grown the lists.
template <int REACH>
__global__ void __launch_bounds__(kBlock) k_glist(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    if (REACH == 2 && !D.ctl->rebuild) return;
    __shared__ double hcls[kMaxClasses];
    for (int k = threadIdx.x; k < P.nclasses; k += kBlock) hcls[k] = D.cls[k].h;
    __syncthreads();
    int t = xcd_block()*kBlock + threadIdx.x;
    if (t >= P.nfree) return;
    int i = D.free_list[t];
    uint32_t fi = D.flags_b[i];
    int ci = flag_cls(fi), mat = flag_mat(fi);
    double4 pi = D.pos_b[i];
    int key_i = D.key[i];
    int kx, ky, kz;
    decode_key(P, key_i, kx, ky, kz);
    ListWriter out(reinterpret_cast<int4 *>(D.nl) + t, (size_t)P.nfree, P.nl_cap);
    int xmat = 0, xc = 0;
    if constexpr (REACH == 2) {
        const int nr = (P.dim == 3) ? 25 : 5;
        const int xa = kx > 1 ? kx - 2 : 0, xb = kx + 2 < P.nc[0] ? kx + 2 : P.nc[0] - 1;
        for (int r = 0; r < nr; ++r) {
            const int yy = ky + (r % 5) - 2, zz = (P.dim == 3) ? kz + (r / 5) - 2 : kz;
            if (yy < 0 || yy >= P.nc[1] || zz < 0 || zz >= P.nc[2]) continue;
            const int base = P.nc[0]*(yy + P.nc[1]*zz);
            const int beg = D.cstart[base + xa], end = D.cstart[base + xb + 1];
            for (int j0 = beg; j0 < end; j0 += 4) {
                uint32_t fg[4];
                double4 pg[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int jj = (j0 + u < end) ? j0 + u : beg;
                    fg[u] = D.flags_b[jj];
                    pg[u] = D.pos_b[jj];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = j0 + u;
                    if (j >= end || j == i) continue;
                    if (flag_mat(fg[u]) != mat) { xmat = 1; continue; }
                    const double x0 = pi.x - pg[u].x, x1 = pi.y - pg[u].y, x2 = pi.z - pg[u].z;
                    if (x0*x0 + x1*x1 + x2*x2 <= P.reach2) out.push(j);
                }
            }
        }
        out.flush();
        D.ncnt[t] = out.cnt;
        if (out.cnt > P.nl_cap) { atomicAdd(&D.ctl->nl_over, 1); atomicOr(&D.ctl->list_over, 1); }
        if (wave_any(xmat != 0) && (threadIdx.x & 63) == 0) atomicOr(&D.ctl->xmat, 1);
        return;
    }
    const int nrows = (P.dim == 3) ? 9 : 3;
    for (int rs = 0; rs < 2*nrows; ++rs) {
        int beg = 0, end = 0;
        if (P.any_periodic) {
            int b0, e0, b1, e1;
            row_range_periodic(D, P, kx, ky, kz, rs >> 1, b0, e0, b1, e1);
            beg = (rs & 1) ? b1 : b0;
            end = (rs & 1) ? e1 : e0;
        } else if (!(rs & 1)) {
            row_range(D, P, kx, ky, kz, rs >> 1, beg, end);
        }
        // candidates in groups of 4: the group's loads are issued together (slots past the
        // range re-read `beg`, which is in bounds, and are not tested), then tested in walk order
        for (int j0 = beg; j0 < end; j0 += 4) {
            uint32_t fg[4];
            double4 pg[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int jj = (j0 + u < end) ? j0 + u : beg;
                fg[u] = D.flags_b[jj];
                pg[u] = D.pos_b[jj];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + u;
                if (j >= end || j == i) continue;
                const uint32_t fj = fg[u];
                if (flag_mat(fj) != mat) {
                    // pairs of different materials go to k_cross; none between solid and soil
                    const int prod = mat*flag_mat(fj);
                    if (P.multi && (prod == 2 || prod == 3)) xc = 1; else xmat = 1;
                    continue;
                }
                const double4 pj = pg[u];
                double h = P.uniform_h ? P.h_uniform : (hcls[ci] + hcls[flag_cls(fj)])/2;
                double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
                if (P.any_periodic) periodic_xij(D, P, h, key_i, j, x0, x1, x2);
                if (support_test(P, x0*x0 + x1*x1 + x2*x2, h)) out.push(j);
            }
        }
    }
    out.flush();
    D.ncnt[t] = out.cnt;
    if (P.multi) D.xflag[t] = xc;
    if (out.cnt > P.nl_cap) atomicAdd(&D.ctl->nl_over, 1);
    int anyx = wave_any(xmat != 0);
    if ((threadIdx.x & 63) == 0 && anyx) atomicOr(&D.ctl->xmat, 1);
}

// Whether list entry j of free target i is a pair of the reference's set at the current
// positions, given their squared distance: inside the support (in_support) and in neighbouring
// cells, which needs the current cell keys only beyond the smallest cell size.
__device__ inline bool skin_member(const Dev &D, const Params &P, int i, int j, double r2)
{
    bool in = in_support(P, r2);
    if (in && r2 > P.adj2) in = keys_adjacent(P, D.key[i], D.key[j]);
    return in;
}

// FLUID (SOLID kernels): runs with free fluid particles besides soil or solid ones; without
// them the fluid pair function is not compiled in, which keeps the FP64 kernel within 256 VGPRs
// (two waves per SIMD).  T = float (SOLID only): the FP32-mixed soil tier; gravity is added in
// FP64 afterwards.
template <bool SOLID, class T, bool FLUID = true>
__global__ void __launch_bounds__(kForceBlock) k_force(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    __shared__ ClassParams cp[kMaxClasses];
    __shared__ int lst[kCap*kForceBlock];
    for (int k = threadIdx.x; k < P.nclasses; k += kForceBlock) cp[k] = D.cls[k];
    __syncthreads();
    // 1/W(dp) of each class pair for the artificial stress (up to 8 classes; else per pair)
    __shared__ double tiw[64];
    const double *tiwp = nullptr;
    if constexpr (SOLID) {
        if (P.any_ti && P.nclasses <= 8) {
            for (int k = threadIdx.x; k < 64; k += kForceBlock) {
                const int a = k >> 3, b = k & 7;
                double v = 0.0;
                if (a < P.nclasses && b < P.nclasses) {
                    const double h = pair_h(cp, P, a, b);
                    v = rcpx(kernel_w<double>(P.dim, P.kernel_type, (cp[a].TIInitDist + cp[b].TIInitDist)/(2.0*h), h));
                }
                tiw[k] = v;
            }
            tiwp = tiw;
            __syncthreads();
        }
    }
    const int tid = threadIdx.x;
    int t = xcd_block()*kForceBlock + tid;
    unsigned long long dtb = ~0ull;      // no dt candidate from an inactive lane
    bool act = t < P.nfree;
    int i = act ? D.free_list[t] : 0;
    uint32_t fi = D.flags_b[i];
    int ci = flag_cls(fi);
    int mat = flag_mat(fi);
    double4 pi = D.pos_b[i];
    double4 vi = D.vel_b[i];
    double4 ai = D.aux[i];
    int kx, ky, kz;
    decode_key(P, D.key[i], kx, ky, kz);
    int shep_on = P.any_shepard && cp[ci].Shepard && (D.ctl->shep_counter == cp[ci].ShepardStep);
    int zw_on = shep_on;
    constexpr bool F32 = sizeof(T) == 4;
    AccT<T> A;
    const double4 a_init = act ? a_start(D, P, i) : make_double4(0.0, 0.0, 0.0, 0.0);
    A.a0 = F32 ? (T)0 : (T)a_init.x; A.a1 = F32 ? (T)0 : (T)a_init.y; A.a2 = F32 ? (T)0 : (T)a_init.z;
    A.dd = F32 ? (T)0 : (T)a_init.w; A.ddc = 0.0;
    A.vx0 = A.vx1 = A.vx2 = 0.0; A.zw = 0.0; A.sd = 0.0; A.sw = 0.0;
    for (int k = 0; k < 6; ++k) A.sr[k] = 0.0;
    A.rr[0] = A.rr[1] = A.rr[2] = 0.0;
    double sgi[6] = {0, 0, 0, 0, 0, 0}, tri[6] = {0, 0, 0, 0, 0, 0};
    if (SOLID) {
        ld6(D.sig_b, i, sgi);
        ld6(D.tir, i, tri);
    }
    // a pair of the target's own material: CalcForce11 for fluid targets, CalcForce2233 for
    // solid and soil ones (LastComputeAcceleration, Domain.cpp:1244-1254); runs with fluid and
    // soil or solid together take the SOLID kernel
    auto pair_same = [&](int jj) __attribute__((always_inline)) {
        if constexpr (SOLID) {
            if constexpr (sizeof(T) == 8 && FLUID) {
                if (mat == 1) { pair11(D, P, cp, i, jj, fi, pi, vi, ai, A, shep_on); return; }
            }
            pair2233<T>(D, P, cp, i, jj, fi, pi, vi, ai, sgi, tri, A, shep_on, nullptr, tiwp);
        } else {
            pair11(D, P, cp, i, jj, fi, pi, vi, ai, A, shep_on);
        }
    };
    unsigned long long nff = 0, nfw = 0;
    int xmat = 0;
    int cnt = 0;
    // listed pairs first (k_glist, same order as the walk below); the walk then runs only for
    // a target whose pairs overflowed the list, and skips the `skip` pairs already taken.
    // With list reuse the lists are skin lists: an entry counts if it is a pair of the
    // reference's set at the current positions (skin_member); a skin list overflows only at a
    // build, where the cells are current, and the target then takes the walk alone.
    int skip = 0;
    bool walk = act;
    if (P.glist && act) {
        const int total = D.ncnt[t], m = total < P.nl_cap ? total : P.nl_cap;
        const int *lst_g = D.nl;
        if (P.verlet && total > P.nl_cap) {
            skip = 0;
        } else {
            for (int e = 0; e < m; ++e) {
                int jj = lst_g[((size_t)(e >> 2)*P.nfree + t)*4 + (e & 3)];
                if (P.verlet) {
                    const double4 pj = D.pos_b[jj];
                    const double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
                    if (!skin_member(D, P, i, jj, x0*x0 + x1*x1 + x2*x2)) continue;
                }
                if (flag_free(D.flags_b[jj])) ++nff; else ++nfw;
                pair_same(jj);
            }
            skip = m;
            walk = total > P.nl_cap;
        }
    }
    int seen = 0;
    const int nrows = (P.dim == 3) ? 9 : 3;
    for (int rs = 0; rs < 2*nrows; ++rs) {
        // row r = rs/2; its second range is non-empty only when the row crosses a period
        int beg = 0, end = 0;
        if (walk) {
            if (P.any_periodic) {
                int b0, e0, b1, e1;
                row_range_periodic(D, P, kx, ky, kz, rs >> 1, b0, e0, b1, e1);
                beg = (rs & 1) ? b1 : b0;
                end = (rs & 1) ? e1 : e0;
            } else if (!(rs & 1)) {
                row_range(D, P, kx, ky, kz, rs >> 1, beg, end);
            }
        }
        for (int j = beg; j < end; ++j) {
            if (j != i) {
                uint32_t fj = D.flags_b[j];
                if (flag_mat(fj) != mat) {
                    if (!P.multi) xmat = 1;           // with several materials: k_cross
                } else {
                    double4 pj = D.pos_b[j];
                    double h = pair_h(cp, P, ci, flag_cls(fj));
                    double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
                    if (P.any_periodic) periodic_xij(D, P, h, D.key[i], j, x0, x1, x2);
                    if (support_test(P, x0*x0 + x1*x1 + x2*x2, h) && seen++ >= skip) {
                        lst[cnt*kForceBlock + tid] = j;
                        ++cnt;
                        if (flag_free(fj)) ++nff; else ++nfw;
                    }
                }
            }
            if (wave_any(cnt == kCap)) {
                for (int e = 0; e < cnt; ++e) {
                    int jj = lst[e*kForceBlock + tid];
                    pair_same(jj);
                }
                cnt = 0;
            }
        }
    }
    for (int e = 0; e < cnt; ++e) {
        int jj = lst[e*kForceBlock + tid];
        pair_same(jj);
    }
    if (act) {
        double a0 = A.a0, a1 = A.a1, a2 = A.a2, dd = A.dd;
        if (F32) {
            a0 = a_init.x + (double)A.a0; a1 = a_init.y + (double)A.a1; a2 = a_init.z + (double)A.a2;
            dd = a_init.w + ((double)A.dd + (double)(-A.ddc));
        }
        D.acc[i] = make_double4(a0, a1, a2, dd);
        if (P.any_xsph || zw_on) D.xsp[i] = make_double4(A.vx0, A.vx1, A.vx2, A.zw);
        if (P.swi1) D.swz[i] = make_double2((double)A.sw, (double)A.zw);
        if (shep_on) D.shep[i] = make_double2(A.sd, 0.0);
        if (SOLID || P.any_sr) { double srv[6]; for (int k = 0; k < 6; ++k) srv[k] = A.sr[k]; st6(D.sr, i, srv); }
        if (SOLID) for (int k = 0; k < 3; ++k) D.rr[3*i+k] = A.rr[k];
        // dt control: min over free particles of sqrt_h_a*sqrt(h/|a|) (Domain.cpp:1281-1296);
        // targets with pairs of another material take theirs after k_cross
        if (!(P.multi && D.xflag[t])) {
            double na = sqrt(a0*a0 + a1*a1 + a2*a2);
            double test = sqrt(cp[ci].h/na);
            double cand = P.sqrt_h_a*test;
            dtb = dbits(cand);
        }
    }
    // pair statistics and the cross-material check
    int anyx = wave_any(xmat != 0);
    if ((tid & 63) == 0 && anyx) atomicOr(&D.ctl->xmat, 1);
    force_partials(D, dtb, nff, nfw);
}


// ------------------------------------------------------------------------------------------
// in/out-flow boundaries: InFlowBCFresh's imposition (Domain.cpp:1588-1635) with the default
// InCon/OutCon (the boundary's constant velocity and density); the inflow set is the particles
// flagged InOut 1 (the host keeps the reference's list order for the recycling)
// ------------------------------------------------------------------------------------------
// Outflow particles of this step: free and at or past OutFlowLoc; flagged InOut 2 and counted.
__global__ void __launch_bounds__(kBlock) k_io_mark(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    const int i = blockIdx.x*kBlock + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t f = D.flags_b[i];
    if (!flag_free(f) || !(P.inout == 2 || P.inout == 3)) return;
    if (D.pos_b[i].x >= P.out_loc) {
        D.flags_b[i] = (f & ~(3u << kInOutShift)) | (2u << kInOutShift);
        atomicAdd(&D.ctl->n_out, 1);
    }
}

// The imposed state: velocity (v, va, vb), density (Density, Densitya, Densityb) and its EOS
// pressure; outflow velocity scaled by InPart/OutPart with mass conservation, set only if the
// boundary has one or the scale differs from 1; outflow density only if positive.
__global__ void __launch_bounds__(kBlock) k_io_impose(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    const int i = blockIdx.x*kBlock + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t f = D.flags_b[i];
    const int io = flag_inout(f);
    if (!flag_free(f) || io == 0) return;
    double4 p = D.pos_b[i], v = D.vel_b[i], vh = D.vh_b[i];
    bool set_v = false, set_d = false;
    double w0 = 0.0, w1 = 0.0, w2 = 0.0, den = 0.0;
    if (io == 1) {
        set_v = set_d = true;
        w0 = P.in_v[0]; w1 = P.in_v[1]; w2 = P.in_v[2]; den = P.in_density;
    } else if (p.x >= P.out_loc) {
        const double t11 = P.mass_cons ? P.n_in*1.0/(D.ctl->n_out*1.0) : 1.0;
        const double nv = sqrt(P.out_v[0]*P.out_v[0] + P.out_v[1]*P.out_v[1] + P.out_v[2]*P.out_v[2]);
        set_v = (nv > 0.0 || t11 != 1.0);
        w0 = t11*P.out_v[0]; w1 = t11*P.out_v[1]; w2 = t11*P.out_v[2];
        set_d = P.out_density > 0.0;
        den = P.out_density;
    }
    if (!set_v && !set_d) return;
    const ClassParams &c = D.cls[flag_cls(f)];
    if (set_v) { v.x = w0; v.y = w1; v.z = w2; vh.x = w0; vh.y = w1; vh.z = w2; }
    if (set_d) { p.w = den; vh.w = den; v.w = eos(c.PresEq, c.Cs, c.P0, den, c.RefDensity); }
    D.pos_b[i] = p;
    D.vel_b[i] = v;
    D.vh_b[i] = vh;
    // the pair formulas' per-particle values (k_reorder's sound speed and P/d^2)
    double4 a = D.aux[i];
    a.z = sound_speed(c.PresEq, c.Cs, p.w, c.RefDensity);
    a.w = v.w/(p.w*p.w);
    D.aux[i] = a;
}

// ------------------------------------------------------------------------------------------
// pairs of different materials (LastComputeAcceleration's NSM loop, Domain.cpp:1256-1271)
// ------------------------------------------------------------------------------------------

// The 3^DIM cells around a particle's cell (periodic directions wrap), as the reference's
// half-stencil search pairs them: f(cell index, its y index) for every cell in the grid.
template <class F>
__device__ inline void for_cells(const Params &P, int kx, int ky, int kz, F &&f)
{
    const int ncell = (P.dim == 3) ? 27 : 9;
    for (int cc = 0; cc < ncell; ++cc) {
        int xx = kx + (cc % 3) - 1, yy = ky + ((cc / 3) % 3) - 1, zz = (P.dim == 3) ? kz + (cc / 9) - 1 : kz;
        if (P.periodic[0]) xx = (xx + P.nc[0]) % P.nc[0];
        if (P.periodic[1]) yy = (yy + P.nc[1]) % P.nc[1];
        if (P.periodic[2]) zz = (zz + P.nc[2]) % P.nc[2];
        if (xx < 0 || xx >= P.nc[0] || yy < 0 || yy >= P.nc[1] || zz < 0 || zz >= P.nc[2]) continue;
        f(xx + P.nc[0]*(yy + P.nc[1]*zz), yy);
    }
}

// FSI smoothing (PrimaryComputeAcceleration, Domain.cpp:1089-1160) onto free solid particles
// from their free fluid neighbours (same or adjacent cells, no distance test: K vanishes beyond
// the support).  As in the reference, the FSI velocity takes only the pairs whose P1 is the
// solid particle (the others go to its NSv, which nothing reads), and the pressure and velocity
// are divided by the kernel sum once per pair.  Every other particle gets zeros.
__global__ void __launch_bounds__(kBlock) k_fsi(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    const int i = blockIdx.x*kBlock + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t fi = D.flags_b[i];
    double n0 = 0.0, n1 = 0.0, n2 = 0.0, fp = 0.0, sk = 0.0;
    if (P.fsi && flag_free(fi) && flag_mat(fi) == 2) {
        const double4 pi = D.pos_b[i];
        const int key_i = D.key[i], orig_i = D.orig_b[i];
        const double hi = D.cls[flag_cls(fi)].h;
        const double *G = P.gravity;
        int kx, ky, kz, m = 0;
        decode_key(P, key_i, kx, ky, kz);
        for_cells(P, kx, ky, kz, [&](int cell, int) {
            for (int j = D.cstart[cell]; j < D.cstart[cell + 1]; ++j) {
                const uint32_t fj = D.flags_b[j];
                if (!flag_free(fj) || flag_mat(fj) != 1) continue;
                const double4 pj = D.pos_b[j];
                const double h = (hi + D.cls[flag_cls(fj)].h)/2.0;
                double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
                if (P.any_periodic) periodic_xij(D, P, h, key_i, j, x0, x1, x2);
                const int first = pair_first(P, key_i, D.key[j], orig_i, D.orig_b[j]);
                const double sgn = first ? 1.0 : -1.0;          // x_P1 - x_P2, exact negation
                x0 *= sgn; x1 *= sgn; x2 *= sgn;
                const double K = kernel_w<double>(P.dim, P.kernel_type, sqrt(x0*x0 + x1*x1 + x2*x2)/h, h);
                const double4 vj = D.vel_b[j];
                ++m;
                sk += K;
                if (first) { n0 += vj.x * K; n1 += vj.y * K; n2 += vj.z * K; }
                fp += vj.w * K + (G[0]*x0 + G[1]*x1 + G[2]*x2)*pj.w*K;
            }
        });
        for (int q = 0; q < m; ++q) { fp = fp/sk; n0 = n0/sk; n1 = n1/sk; n2 = n2/sk; }
    }
    D.fsi[i] = make_double4(n0, n1, n2, fp);
    D.fsk[i] = sk;
}

// Saturation (Domain.cpp:1060-1088, 1209-1233): a soil particle with a fluid particle in its
// own or a neighbouring cell, in a cell row not below its own and not lower than itself, one
// of the two free, is saturated for this step: its density, Densityb and RefDensity drop by
// RhoF and its mass becomes V (RefDensity - RhoF); it returns when no such neighbour remains.
__global__ void __launch_bounds__(kBlock) k_sat(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    const int i = blockIdx.x*kBlock + threadIdx.x;
    if (i >= P.n) return;
    const uint32_t fi = D.flags_b[i];
    if (flag_mat(fi) != 3) return;
    const double4 pi = D.pos_b[i];
    int kx, ky, kz;
    decode_key(P, D.key[i], kx, ky, kz);
    bool check = false;
    for_cells(P, kx, ky, kz, [&](int cell, int yy) {
        if (check || yy < ky) return;
        for (int j = D.cstart[cell]; j < D.cstart[cell + 1]; ++j) {
            const uint32_t fj = D.flags_b[j];
            if (flag_mat(fj) != 1 || !(flag_free(fi) || flag_free(fj))) continue;
            if (D.pos_b[j].y >= pi.y) { check = true; return; }
        }
    });
    const ClassParams &c = D.cls[flag_cls(fi)];
    const bool sat = (fi & kSat) != 0;
    uint32_t f = (fi & ~kSatCheck) | (check ? kSatCheck : 0u);
    if (check != sat) {
        double4 p = D.pos_b[i];
        double2 mm = D.mmu_b[i];
        double4 vh = D.vh_b[i];
        const double r0 = c.RefDensity - c.RhoF;              // the saturated RefDensity
        if (check) { mm.x = c.V*(c.RefDensity - c.RhoF); p.w = p.w - c.RhoF; vh.w = vh.w - c.RhoF; f |= kSat; }
        else { mm.x = c.V*(r0 + c.RhoF); p.w = p.w + c.RhoF; vh.w = vh.w + c.RhoF; f &= ~kSat; }
        D.pos_b[i] = p;
        D.mmu_b[i] = mm;
        if (P.scheme == 0 || !flag_free(fi)) D.vh_b[i] = vh;   // Densityb (dead under Leapfrog)
        if (flag_free(fi)) { double4 a = D.aux[i]; a.x = mm.x; D.aux[i] = a; }
    }
    D.flags_b[i] = f;
}

struct XAcc { double a0, a1, a2, dd; double sr[6]; };

// CalcForce12 (Interaction.cpp:379-506) contribution to free particle i as P1 from j, one of
// the two fluid and the other solid.  The fluid particle's EOS turns the solid particle's FSI
// pressure into its density; both masses are the fluid particle's; no artificial viscosity
// (the reference zeroes Alpha and Beta before testing them).
__device__ __forceinline__ void pair12(const Dev &D, const Params &P, const ClassParams *cp, int i, int j,
                              uint32_t fi, uint32_t fj, double4 pi, double4 pj, double x0, double x1,
                              double x2, double rij, double h, XAcc &A)
{
    const ClassParams &Ci = cp[flag_cls(fi)], &Cj = cp[flag_cls(fj)];
    const double4 vi = D.vel_b[i], vj = D.vel_b[j];
    const bool f1 = flag_mat(fi) == 1;
    const double GK = grad_kernel<double>(P.dim, P.kernel_type, rij/h, h);
    double di, dj, mi, mj, Pi_, Pj_, Mu, b0, b1, b2;
    if (f1) {
        const double4 fs = D.fsi[j];
        di = pi.w; mi = D.mmu_b[i].x;
        dj = density_solid(Ci.PresEq, Ci.Cs, Ci.P0, fs.w, Ci.RefDensity);
        mj = mi;
        Pi_ = vi.w; Pj_ = fs.w; Mu = D.mmu_b[i].y;
        b0 = vi.x - (2.0*vj.x-fs.x); b1 = vi.y - (2.0*vj.y-fs.y); b2 = vi.z - (2.0*vj.z-fs.z);
    } else {
        const double4 fs = D.fsi[i];
        di = density_solid(Cj.PresEq, Cj.Cs, Cj.P0, fs.w, Cj.RefDensity);
        mi = D.mmu_b[j].x;
        dj = pj.w;
        mj = mi;
        Pi_ = fs.w; Pj_ = vj.w; Mu = D.mmu_b[j].y;
        b0 = (2.0*vi.x-fs.x) - vj.x; b1 = (2.0*vi.y-fs.y) - vj.y; b2 = (2.0*vi.z-fs.z) - vj.z;
    }
    const double u0 = vi.x - vj.x, u1 = vi.y - vj.y, u2 = vi.z - vj.z;
    double VI0 = 0.0, VI1 = 0.0, VI2 = 0.0;
    viscous_force<double>(P.visc_eq, VI0, VI1, VI2, Mu, di, dj, GK, b0, b1, b2, P.dim,
                          P.kernel_type, rij, h, x0, x1, x2, u0, u1, u2);
    const double PIij = 0.0;
    double sf;
    if (P.gradient_type == 0) sf = -1.0*( Pi_/(di*di) + Pj_/(dj*dj) + PIij ) * GK;
    else sf = -1.0*( (Pi_ + Pj_)/(di*dj) + PIij ) * GK;
    double t0 = sf*x0 + VI0, t1 = sf*x1 + VI1, t2 = sf*x2 + VI2;
    if (P.dim == 2) t2 = 0.0;
    A.a0 += mj * t0; A.a1 += mj * t1; A.a2 += mj * t2;
    if (f1) {
        const double temp1 = u0*(GK*x0) + u1*(GK*x1) + u2*(GK*x2);
        A.dd += mj * (di/dj) * temp1;
        if ((Ci.T0 > 0.0 || Cj.T0 > 0.0 || Ci.LES || Cj.LES) && (Ci.T0 > 0.0 || Ci.LES)) {
            const double g = -GK, f = mj/dj;
            A.sr[0] += ((2.0*b0*x0)*g)*f; A.sr[1] += ((b0*x1+b1*x0)*g)*f; A.sr[2] += ((b0*x2+b2*x0)*g)*f;
            A.sr[3] += ((2.0*b1*x1)*g)*f; A.sr[4] += ((b1*x2+b2*x1)*g)*f; A.sr[5] += ((2.0*b2*x2)*g)*f;
        }
    }
}

// Seepage (Functions.cpp:266-298): Darcy (0, 1) or Forchheimer (2, 3) coefficients
__device__ __forceinline__ void seepage(int st, double k, double k2, double mu, double rho, double &SF1, double &SF2)
{
    SF1 = mu/k;
    SF2 = (st >= 2) ? k2*rho : 0.0;
}

// CalcForce13 (Interaction.cpp:508-653) contribution to free particle i as P1 from j, one of
// the two soil and the other water: the seepage force (SWIType 0 and 2), with SWIType 2 the
// water pressure on the soil particle, nothing with SWIType 3.  K is over both densities.
__device__ __forceinline__ void pair13(const Dev &D, const Params &P, const ClassParams *cp, int i, int j,
                              uint32_t fi, uint32_t fj, double4 pi, double4 pj, double x0, double x1,
                              double x2, double rij, double h, XAcc &A)
{
    if (P.swi_type == 3) return;
    const bool s1 = flag_mat(fi) == 3;                 // i (P1) is the soil particle
    const int so = s1 ? i : j, w = s1 ? j : i;
    const ClassParams &Cs = cp[flag_cls(s1 ? fi : fj)], &Cw = cp[flag_cls(s1 ? fj : fi)];
    const double K = kernel_w<double>(P.dim, P.kernel_type, rij/h, h)/(pi.w*pj.w);
    const double4 vw = D.vel_b[w], vs = D.vel_b[so];
    const double v0 = vw.x - vs.x, v1 = vw.y - vs.y, v2 = vw.z - vs.z;
    const double nv = sqrt(v0*v0 + v1*v1 + v2*v2);
    double s0, s1v, s2;
    const double2 zs = P.swi1 ? D.swz[so] : make_double2(0.0, 1.0);   // soil's S, ZWab
    if (P.swi1 && zs.y < 0.25) {
        // surface erosion: drag and lift on a soil particle with few soil neighbours
        const double s_w = D.swz[w].x;                 // the water particle's S
        const double Cd = 24.0*(Cw.MuRef/Cw.RefDensity)/(Cs.d*nv+0.01*h*h) + 2.0;
        s0 = (3.0/(4.0*Cs.d)*Cw.RefDensity*(1.0-Cs.n0)*Cd*nv*v0) *K;
        s1v = (3.0/(4.0*Cs.d)*Cw.RefDensity*(1.0-Cs.n0)*Cd*nv*v1) *K;
        s2 = (3.0/(4.0*Cs.d)*Cw.RefDensity*(1.0-Cs.n0)*Cd*nv*v2) *K;
        s1v += (Cw.RefDensity*(1.0-Cs.n0)*nv*fabs(s_w-zs.x)) *K;
    } else {
        double SF1, SF2;
        seepage(Cs.SeepageType, Cs.k, Cs.k2, Cw.MuRef, Cw.RefDensity, SF1, SF2);
        s0 = (SF1*v0 + SF2*nv*v0) *K; s1v = (SF1*v1 + SF2*nv*v1) *K; s2 = (SF1*v2 + SF2*nv*v2) *K;
    }
    if (P.dim == 2) s2 = 0.0;
    if (s1) {
        const double mw = D.mmu_b[w].x;
        double t0 = mw*s0, t1 = mw*s1v, t2 = mw*s2;
        if (P.swi_type == 2) {
            const double GK = grad_kernel<double>(P.dim, P.kernel_type, rij/h, h)/(pi.w*pj.w);
            const double pw = vw.w;
            t0 = t0 - mw*pw*GK*x0; t1 = t1 - mw*pw*GK*x1; t2 = t2 - mw*pw*GK*x2;
        }
        A.a0 += t0; A.a1 += t1; A.a2 += t2;
    } else {
        const double ms = D.mmu_b[so].x;
        A.a0 -= ms*s0; A.a1 -= ms*s1v; A.a2 -= ms*s2;
    }
}

// The pairs of each free target with particles of another material, after every other pair
// (the reference's NSM loop runs after SM and FSM): CalcForce12 (fluid with solid, mean h) and
// CalcForce13 (fluid with soil, the smaller h).  Adds to a, dDensity and the fluid's strain
// rate, then takes the target's dt candidate.  Targets without such neighbours return at once.
__global__ void __launch_bounds__(kForceBlock) k_cross(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    __shared__ ClassParams cp[kMaxClasses];
    for (int k = threadIdx.x; k < P.nclasses; k += kForceBlock) cp[k] = D.cls[k];
    __syncthreads();
    const int t = xcd_block()*kForceBlock + threadIdx.x;
    unsigned long long dtb = ~0ull, nff = 0, nfw = 0;
    if (t < P.nfree && D.xflag[t]) {
        const int i = D.free_list[t];
        const uint32_t fi = D.flags_b[i];
        const int ci = flag_cls(fi), mat = flag_mat(fi);
        const double4 pi = D.pos_b[i];
        const int key_i = D.key[i];
        int kx, ky, kz;
        decode_key(P, key_i, kx, ky, kz);
        const double4 a = D.acc[i];
        XAcc A;
        A.a0 = a.x; A.a1 = a.y; A.a2 = a.z; A.dd = a.w;
        const bool sr_on = P.any_sr && (cp[ci].T0 > 0.0 || cp[ci].LES);
        for (int k = 0; k < 6; ++k) A.sr[k] = sr_on ? D.sr[6*i+k] : 0.0;
        for_cells(P, kx, ky, kz, [&](int cell, int) {
            for (int j = D.cstart[cell]; j < D.cstart[cell + 1]; ++j) {
                const uint32_t fj = D.flags_b[j];
                const int prod = mat*flag_mat(fj);
                if (flag_mat(fj) == mat || (prod != 2 && prod != 3)) continue;
                const double4 pj = D.pos_b[j];
                const double hi = cp[ci].h, hj = cp[flag_cls(fj)].h;
                const double h = (prod == 3) ? fmin(hi, hj) : (hi + hj)/2;
                double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
                if (P.any_periodic) periodic_xij(D, P, h, key_i, j, x0, x1, x2);
                const double r2 = x0*x0 + x1*x1 + x2*x2;
                if (!support_test(P, r2, h)) continue;
                const double rij = sqrt(r2);
                if (flag_free(fj)) ++nff; else ++nfw;
                if (prod == 2) pair12(D, P, cp, i, j, fi, fj, pi, pj, x0, x1, x2, rij, h, A);
                else pair13(D, P, cp, i, j, fi, fj, pi, pj, x0, x1, x2, rij, h, A);
            }
        });
        D.acc[i] = make_double4(A.a0, A.a1, A.a2, A.dd);
        if (sr_on) for (int k = 0; k < 6; ++k) D.sr[6*i+k] = A.sr[k];
        const double na = sqrt(A.a0*A.a0 + A.a1*A.a1 + A.a2*A.a2);
        dtb = dbits(P.sqrt_h_a*sqrt(cp[ci].h/na));
    }
    force_partials(D, dtb, nff, nfw);
}

// ------------------------------------------------------------------------------------------
// fast fluid force pass: CalcForce11 for cubic spline, Morris viscosity, squared-density
// momentum form, uniform h and viscosity (the dam-break configuration).  Neighbour data comes
// from the gather views; divisions become refined reciprocals (each within ~2 ulp), the pair
// test is decided on rij^2 except in a 1e-13 band around the support radius, where the exact
// rij/h <= Cellfac test of the reference runs.
// ------------------------------------------------------------------------------------------


struct Acc11 { double a0, a1, a2, dd; };

// Artificial-viscosity constants of a pair of a target of class ci with a neighbour (meta: its
// class, or -1 - class for a fixed particle): nha = -alpha/2 and beta, and whether the pair has
// an artificial viscosity term at all (on: alpha or beta nonzero).  ABTable reads them from the
// class-pair table in LDS; ABTwo chooses between the free-free and the free-fixed pair by the
// neighbour's kind (every free particle of one class, the same constants with every fixed
// class: Params::ab_two); ABOne holds the one pair of a run whose free-free and free-fixed
// constants are equal (Params::ab_one).  ABTwo and ABOne read the kernel arguments, which the
// host fills per precision (Params::ab_nha, ab32, ab_on), so that they stay scalar operands.
template <class T> __device__ inline T ab_nha(const Params &P, int k);
template <> __device__ inline double ab_nha<double>(const Params &P, int k) { return P.ab_nha[k]; }
template <> __device__ inline float ab_nha<float>(const Params &P, int k) { return P.ab32[2*k]; }
template <class T> __device__ inline T ab_beta(const Params &P, int k);
template <> __device__ inline double ab_beta<double>(const Params &P, int k) { return k ? P.ab_fixed[1] : P.ab_free[1]; }
template <> __device__ inline float ab_beta<float>(const Params &P, int k) { return P.ab32[2*k + 1]; }

template <class T>
struct ABTable {
    const T *pa, *pb;
    __device__ void get(int ci, T meta, bool freej, T &nha, T &beta) const {
        const int cj = freej ? (int)meta : (int)((T)-1 - meta);
        nha = -pa[ci*8 + cj]*(T)0.5;
        beta = pb[ci*8 + cj];
    }
    __device__ bool on(T nha, T beta, bool) const { return nha != (T)0 || beta != (T)0; }
};
template <class T>
struct ABTwo {
    const Params &P;
    __device__ void get(int, T, bool freej, T &nha, T &beta) const {
        nha = freej ? ab_nha<T>(P, 0) : ab_nha<T>(P, 1);
        beta = freej ? ab_beta<T>(P, 0) : ab_beta<T>(P, 1);
    }
    __device__ bool on(T, T, bool freej) const { return freej ? P.ab_on[0] != 0 : P.ab_on[1] != 0; }
};
template <class T>
struct ABOne {
    const Params &P;
    __device__ void get(int, T, bool, T &nha, T &beta) const { nha = ab_nha<T>(P, 0); beta = ab_beta<T>(P, 0); }
    __device__ bool on(T, T, bool) const { return P.ab_on[0] != 0; }
};
// ABM: 0 the class-pair table, 1 two pairs by the neighbour's kind, 2 one pair
template <int ABM, class T> struct ABPick;
template <class T> struct ABPick<2, T> {
    static __device__ ABOne<T> make(const Params &P, const T *, const T *) { return {P}; }
};
template <class T> struct ABPick<1, T> {
    static __device__ ABTwo<T> make(const Params &P, const T *, const T *) { return {P}; }
};
template <class T> struct ABPick<0, T> {
    static __device__ ABTable<T> make(const Params &, const T *pa, const T *pb) { return {pa, pb}; }
};

// A lane's pair tally: free neighbours in the low 16 bits, fixed ones in the high 16 bits, one
// add per pair (a lane evaluates fewer than 65,536 pairs: the lists hold at most 65,535
// entries per target).
__device__ inline unsigned pair_tally(bool freej) { return freej ? 1u : 0x10000u; }
__device__ inline unsigned tally_free(unsigned npc) { return npc & 0xffffu; }
__device__ inline unsigned tally_fixed(unsigned npc) { return npc >> 16; }

// SKIN: j comes from a skin list and counts only if it is a pair of the reference's set at the
// current positions (skin_member); i is the target (for its cell key).  gi0..gi2 are the
// target's views, ci its class.  A wall neighbour's viscous velocity is read only for walls.
// The pair of free target i (views gi0..gi2, class ci) with the neighbour whose records are
// a0..a2 (and vw, its viscous velocity, read only for walls).  jf() gives the neighbour's index
// in the current order; it is called only for the rare key-adjacency part of the skin test.
template <int DIM, bool SKIN, class AB, class JF>
__device__ inline void pair11_rec(const Dev &D, const Params &P, const AB &ab,
                                  double4 gi0, double4 gi1, double4 gi2, int ci, int ki, JF jf,
                                  double4 a0, double4 a1, double4 a2, double4 vw,
                                  double vsf, double vsw, Acc11 &A, unsigned &npc)
{
    const double cj_snd = a2.x, meta = a2.y, mrec = a2.z, invdj = a2.w;
    double x0 = gi0.x - a0.x, x1 = gi0.y - a0.y, x2 = gi0.z - a0.z;
    // the pair test on the reference's r^2 (no contraction: the same set as a list build)
    double r2 = x0*x0 + x1*x1 + x2*x2;
    if (SKIN) {
        bool in = in_support(P, r2);
        if (in && r2 > P.adj2) in = keys_adjacent(P, ki, D.key[jf()]);
        if (!in) return;
    }
    // the contribution: products and sums may fuse (each fused pair rounds once instead of
    // twice, far inside the 1e-12 bar on the force pass)
    {
#pragma clang fp contract(fast)
    double invr = rsq_nr(r2);
    double rij = r2 > 0.0 ? r2*invr : 0.0;
    double q = rij*P.invh;
    double K, GK;
    if (q < 1.0) {
        K = P.kw*(1.0-1.5*q*q+0.75*q*q*q);
        GK = P.kgh*(-3.0+2.25*q);
    } else if (q < 2.0) {
        double t = 2.0 - q;
        K = P.kw*(0.25*t*t*t);
        GK = P.kg*invr*(-0.75*t*t);
    } else {
        K = 0.0; GK = 0.0;
    }
    (void)K;
    const bool freej = meta >= 0.0;
    npc += pair_tally(freej);
    // viscous velocity of the neighbour (the caller passes a free neighbour's v)
    const double w0 = vw.x, w1 = vw.y, w2 = vw.z;
    double nha, beta;
    ab.get(ci, meta, freej, nha, beta);
    double u0 = gi1.x - a1.x, u1 = gi1.y - a1.y, u2 = gi1.z - a1.z;
    double dvx = u0*x0 + u1*x1 + u2*x2;
    double di = gi0.w, dj = a0.w;
    double PIij = 0.0;
    if (ab.on(nha, beta, freej) && dvx < 0.0) {
        double MU = P.h_uniform*dvx*rcp_nr(r2 + P.h2c);
        PIij = (nha*(gi2.x + cj_snd)*MU + beta*MU*MU)*rcp_nr(0.5*(di + dj));
    }
    double mi = gi2.z;
    double mj = freej ? mrec : mrec*mi;
    // vsf, vsw: 2 mu / d_i of the target with a free and with a fixed neighbour
    double vs = (freej ? vsf : vsw)*invdj*GK;
    double s = -(gi1.w + a1.w + PIij)*GK;
    double b0 = gi1.x - w0, b1 = gi1.y - w1, b2 = gi1.z - w2;
    double t0 = s*x0 + vs*b0, t1 = s*x1 + vs*b1, t2 = s*x2 + vs*b2;
    if (DIM == 2) t2 = 0.0;
    double temp1 = (u0*x0 + u1*x1 + u2*x2)*GK;
    A.a0 += mj*t0; A.a1 += mj*t1; A.a2 += mj*t2;
    A.dd += mj*(di*invdj)*temp1;
    }
}

// SKIN: j comes from a skin list and counts only if it is a pair of the reference's set at the
// current positions (skin_member); i is the target (for its cell key).  gi0..gi2 are the
// target's views, ci its class.  A wall neighbour's viscous velocity is read only for walls.
template <int DIM, bool SKIN>
__device__ inline void pair11_fast(const Dev &D, const Params &P, const double *pa, const double *pb,
                                   double4 gi0, double4 gi1, double4 gi2, int ci, int ki, int j,
                                   double vsf, double vsw, Acc11 &A, unsigned &npc)
{
    double4 a0 = D.g0[j], a1 = D.g1[j], a2 = D.g2[j];
    double4 vw = a1;
    if (a2.y < 0.0) vw = D.gw[j];
    pair11_rec<DIM, SKIN>(D, P, ABTable<double>{pa, pb}, gi0, gi1, gi2, ci, ki, [j] { return j; }, a0, a1, a2, vw,
                          vsf, vsw, A, npc);
}

// ------------------------------------------------------------------------------------------
// neighbour lists of the fast fluid path.  k_nlist finds each free target's in-support
// neighbours (the exact FP64 pair test of the reference); the gathers then run over the lists,
// so every lane of a wave evaluates pairs for nearly all of the loop instead of waiting on
// candidate scans of other lanes.
//
// List reuse (REACH = 2): a step that rebuilds records every candidate within the support
// radius plus a skin (the skin list); until some free particle has moved half the skin, no
// pair can form outside those candidates.  The gathers walk the skin list every step and keep
// the entries that are pairs of the reference's set at the current positions, so every step's
// pair set is the reference's; the sums run in the list's order, the walk order of the build.
// ------------------------------------------------------------------------------------------

// Visits the candidates of free target i (at px, py, pz in cell kx, ky, kz) in list order:
// neighbour rows (dy, dz) in row order over the W^DIM cells around the target's cell, each row
// culled to the cells that intersect the sphere of squared radius cull2 around the target,
// candidates in sorted order.  f(j, code, r2) receives every candidate j != i of the kept cells
// with its cell code and its squared distance, computed as the reference forms xij.
template <int DIM, int REACH, class F>
__device__ inline void walk_rows(const Dev &D, const Params &P, int i, double px, double py, double pz,
                                 int kx, int ky, int kz, double cull2, F &&f)
{
    constexpr int W = Codes<REACH>::W;
    const int nrows = (DIM == 3) ? W*W : W;
    const double cx = P.csize[0], cy = P.csize[1], cz = P.csize[2];
    const double ox = P.blpf[0] + kx*cx, oy = P.blpf[1] + ky*cy, oz = P.blpf[2] + kz*cz;
    double el = px - ox, er = ox + cx - px;
    el = el > 0.0 ? el : 0.0;
    er = er > 0.0 ? er : 0.0;
    for (int r = 0; r < nrows; ++r) {
        int dy = (r % W) - REACH, dz = (DIM == 3) ? (r / W) - REACH : 0;
        int yy = ky + dy, zz = kz + dz;
        if (yy < 0 || yy >= P.nc[1] || zz < 0 || zz >= P.nc[2]) continue;
        double ey = (dy < 0) ? py - (oy + (dy + 1)*cy) : (dy > 0) ? (oy + dy*cy) - py : 0.0;
        double ez = (dz < 0) ? pz - (oz + (dz + 1)*cz) : (dz > 0) ? (oz + dz*cz) - pz : 0.0;
        ey = ey > 0.0 ? ey : 0.0;
        ez = ez > 0.0 ? ez : 0.0;
        double dyz = ey*ey + ez*ez;
        if (dyz > cull2) continue;
        int x0 = kx, x1 = kx;
        if (kx > 0 && dyz + el*el <= cull2) {
            x0 = kx - 1;
            if (REACH == 2 && kx > 1 && dyz + (el + cx)*(el + cx) <= cull2) x0 = kx - 2;
        }
        if (kx + 1 < P.nc[0] && dyz + er*er <= cull2) {
            x1 = kx + 1;
            if (REACH == 2 && kx + 2 < P.nc[0] && dyz + (er + cx)*(er + cx) <= cull2) x1 = kx + 2;
        }
        int base = P.nc[0]*(yy + P.nc[1]*zz);
        int beg = D.cstart[base + x0], end = D.cstart[base + x1 + 1];
        if (beg >= end) continue;
        // the x offset of a candidate's cell from the cell starts of the row (indices clamped to
        // the row, so cells outside the grid take no candidates)
        int b0 = D.cstart[base + kx], b1 = D.cstart[base + kx + 1];
        int bm = (REACH == 2) ? D.cstart[base + (kx > 0 ? kx - 1 : 0)] : 0;
        int bp = (REACH == 2) ? D.cstart[base + (kx + 2 < P.nc[0] ? kx + 2 : P.nc[0])] : 0;
        int crow = W*(dy + REACH) + W*W*(dz + REACH);
        double4 nx = D.pos_b[beg];
        for (int j = beg; j < end; ++j) {
            double4 pj = nx;
            if (j + 1 < end) nx = D.pos_b[j + 1];
            if (j == i) continue;
            double x0d = px - pj.x, x1d = py - pj.y, x2d = pz - pj.z;
            double r2 = x0d*x0d + x1d*x1d + x2d*x2d;
            int cxo = (REACH == 1) ? ((j < b0) ? 0 : (j < b1) ? 1 : 2)
                                   : (int)(j >= bm) + (int)(j >= b0) + (int)(j >= b1) + (int)(j >= bp);
            f(j, crow + cxo, r2);
        }
    }
}


// List build.  REACH = 1: the pairs of this step.  REACH = 2 (list reuse, steps that rebuild
// only): the skin list, every candidate within the support radius plus the skin; one that
// overflows makes every step rebuild until the host has grown the lists.
template <int DIM, int REACH>
__global__ void __launch_bounds__(kBlock) k_nlist(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    if (REACH == 2 && !D.ctl->rebuild) return;
    int t = xcd_block()*kBlock + threadIdx.x;
    if (t >= P.nfree) return;
    int i = D.free_list[t];
    double4 pi = D.pos_b[i];
    int kx, ky, kz;
    decode_key(P, D.key[i], kx, ky, kz);
    ListWriter out(reinterpret_cast<int4 *>(D.nl) + t, (size_t)P.nfree, P.nl_cap);
    const bool codes = P.nl_codes != 0;
    const double reach2 = P.reach2;
    const double cull2 = ((REACH == 2) ? reach2 : P.sup2_hi)*(1.0 + 1e-9);
    walk_rows<DIM, REACH>(D, P, i, pi.x, pi.y, pi.z, kx, ky, kz, cull2, [&](int j, int code, double r2) {
        bool keep = (REACH == 2) ? (r2 <= reach2) : in_support(P, r2);
        if (keep) out.push(codes ? nl_entry<REACH>(j, code) : j);
    });
    out.flush();
    D.ncnt[t] = out.cnt;
    if (REACH == 2 && out.cnt > P.nl_cap) atomicOr(&D.ctl->list_over, 1);
}

// FP64 fluid force pass over the neighbour lists.
template <int DIM, int REACH>
__global__ void __launch_bounds__(kForceBlock) k_force11_nl(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    __shared__ double pa[64], pb[64];
    if (threadIdx.x < 64) { pa[threadIdx.x] = P.pair_alpha[threadIdx.x]; pb[threadIdx.x] = P.pair_beta[threadIdx.x]; }
    __syncthreads();
    const int tid = threadIdx.x;
    int t = xcd_block()*kForceBlock + tid;
    unsigned long long dtb = ~0ull;      // no dt candidate from an inactive lane
    bool act = t < P.nfree;
    int i = act ? D.free_list[t] : 0;
    double4 gi0 = D.g0[i], gi1 = D.g1[i], gi2 = D.g2[i];
    int ci = (int)gi2.y;
    const double vsf = 2.0*P.mu_ff*gi2.w, vsw = 2.0*P.mu_ff_wall*gi2.w;   // 2 mu / d_i
    Acc11 A;
    {
        const double4 a_init = act ? a_start(D, P, i) : make_double4(0.0, 0.0, 0.0, 0.0);
        A.a0 = a_init.x; A.a1 = a_init.y; A.a2 = a_init.z; A.dd = a_init.w;
    }
    unsigned npc = 0;
    const int ki = D.key[i];
    const int cap = P.nl_cap;
    const bool codes = P.nl_codes != 0;
    int cnt = act ? D.ncnt[t] : 0;
    int m = cnt < cap ? cnt : cap;
    const int4 *lst = reinterpret_cast<const int4 *>(D.nl) + t;
    const size_t gs = (size_t)P.nfree;
    int4 w = make_int4(0, 0, 0, 0);
    for (int e = 0; e < m; ++e) {
        int sl = e & 3;
        if (sl == 0) w = lst[(size_t)(e >> 2)*gs];
        int v = (sl == 0) ? w.x : (sl == 1) ? w.y : (sl == 2) ? w.z : w.w;
        pair11_fast<DIM, REACH == 2>(D, P, pa, pb, gi0, gi1, gi2, ci, ki, codes ? nl_index<REACH>(v) : v,
                                         vsf, vsw, A, npc);
    }
    if (act && cnt > cap) {
        // the list overflowed: k_force11_nl_over continues this target from these sums
        int q = atomicAdd(&D.ctl->nl_q, 1);
        D.ovq[q] = t;
        D.ovs[2*q] = make_double4(A.a0, A.a1, A.a2, A.dd);
        act = false;
    }
    if (act) {
        D.acc[i] = make_double4(A.a0, A.a1, A.a2, A.dd);
        double na = sqrt(A.a0*A.a0 + A.a1*A.a1 + A.a2*A.a2);
        double cand = P.sqrt_h_a*sqrt(P.h_uniform/na);
        dtb = dbits(cand);
    }
    force_partials(D, dtb, tally_free(npc), tally_fixed(npc));
}

// Ends the targets of the FP64 list pass whose neighbours overflowed the lists: the entries
// past the capacity are found again in list order and added to the queued partial sums, which
// is the same sequence of operations as one pass over a long enough list.  Lists overflow only
// on steps that build them (a reused list that overflowed makes every step rebuild), so the
// cell table is current.
template <int DIM, int REACH>
__global__ void __launch_bounds__(kForceBlock) k_force11_nl_over(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    const int nq = D.ctl->nl_q;
    if (nq == 0) return;
    __shared__ double pa[64], pb[64];
    if (threadIdx.x < 64) { pa[threadIdx.x] = P.pair_alpha[threadIdx.x]; pb[threadIdx.x] = P.pair_beta[threadIdx.x]; }
    __syncthreads();
    const int cap = P.nl_cap;
    const double reach2 = P.reach2;
    const double cull2 = ((REACH == 2) ? reach2 : P.sup2_hi)*(1.0 + 1e-9);
    for (int q = blockIdx.x*kForceBlock + threadIdx.x; q < nq; q += gridDim.x*kForceBlock) {
        int t = D.ovq[q];
        int i = D.free_list[t];
        double4 gi0 = D.g0[i], gi1 = D.g1[i], gi2 = D.g2[i];
        int ci = (int)gi2.y;
        const double vsf = 2.0*P.mu_ff*gi2.w, vsw = 2.0*P.mu_ff_wall*gi2.w;   // 2 mu / d_i
        double4 s0 = D.ovs[2*q];
        Acc11 A;
        A.a0 = s0.x; A.a1 = s0.y; A.a2 = s0.z; A.dd = s0.w;
        unsigned npc = 0;
        const int ki = D.key[i];
        int kx, ky, kz;
        decode_key(P, ki, kx, ky, kz);
        int k = 0;
        walk_rows<DIM, REACH>(D, P, i, gi0.x, gi0.y, gi0.z, kx, ky, kz, cull2, [&](int j, int, double r2) {
            if (!((REACH == 2) ? (r2 <= reach2) : in_support(P, r2))) return;      // a list entry
            if (k >= cap) pair11_fast<DIM, REACH == 2>(D, P, pa, pb, gi0, gi1, gi2, ci, ki, j, vsf, vsw, A, npc);
            ++k;
        });
        D.acc[i] = make_double4(A.a0, A.a1, A.a2, A.dd);
        double na = sqrt(A.a0*A.a0 + A.a1*A.a1 + A.a2*A.a2);
        double cand = P.sqrt_h_a*sqrt(P.h_uniform/na);
        atomicMin(&D.ctl->dtmin_bits, dbits(cand));
        if (tally_free(npc)) atomicAdd(&D.ctl->pairs_ff, (unsigned long long)tally_free(npc));
        if (tally_fixed(npc)) atomicAdd(&D.ctl->pairs_fw, (unsigned long long)tally_fixed(npc));
    }
}

// ------------------------------------------------------------------------------------------
// FP64 fast path with list reuse: the gather per tile, with the neighbours' records in LDS.
//
// The list-only gather re-reads each neighbour's three 32-B records from L2 once per target
// that lists it (~80 times a step), and is bound by the address processing of those scattered
// loads.  Here the free targets of one cell row are grouped in tiles of at most kTileTargets;
// on the steps that build the skin lists, each tile forms the union of its targets' list
// entries (its items) and rewrites the lists as item slots.  Every step a workgroup then reads
// its tile's items once, with row-contiguous loads, into LDS and walks the slot lists there,
// kTileLanes lanes per target (lane l takes entries l, l + kTileLanes, ...), whose partial sums
// meet in a fixed order.  A tile whose items do not fit reads them from HBM instead.
// ------------------------------------------------------------------------------------------

// Block-wide exclusive scan of one int per thread (blockDim.x a multiple of 64, <= 1024);
// returns the thread's offset and the total in *total.
__device__ inline int block_scan(int v, int *sh, int *total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int k = 0; k < nw; ++k) { int t = sh[k]; sh[k] = run; run += t; }
        sh[nw] = run;
    }
    __syncthreads();
    int r = sh[w] + x - v;
    *total = sh[nw];
    __syncthreads();
    return r;
}

// The XCD-aware order of xcd_block over the first nb blocks of a grid.
__device__ inline int xcd_remap(int b, int nb)
{
    int q = nb >> 3, r = nb & 7;
    int x = b & 7, k = b >> 3;
    return (x < r) ? x*(q + 1) + k : r*(q + 1) + (x - r)*q + k;
}

// Tile table (list builds only): the free targets of each cell row (x fastest) split evenly
// into ceil(count / kTileTargets) tiles.
__global__ void __launch_bounds__(1024) k_tile_table(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt || !D.ctl->rebuild) return;
    __shared__ int sh[17];
    const int rows = P.nc[1]*P.nc[2];
    int running = 0;
    for (int r0 = 0; r0 < rows; r0 += 1024) {
        const int r = r0 + (int)threadIdx.x;
        int f0 = 0, cnt = 0, nt = 0;
        if (r < rows) {
            const int c0 = P.nc[0]*r;
            f0 = D.fstart[c0];
            cnt = D.fstart[c0 + P.nc[0]] - f0;
            nt = (cnt + kTileTargets - 1)/kTileTargets;
        }
        int total = 0;
        const int off = block_scan(nt, sh, &total);
        for (int k = 0; k < nt; ++k) {
            const int a = f0 + (int)((long long)k*cnt/nt), b = f0 + (int)((long long)(k + 1)*cnt/nt);
            D.tile[running + off + k] = make_int4(a, b - a, 0, 0);
        }
        running += total;
    }
    if (threadIdx.x == 0) D.ctl->ntiles = running;
}

// Slot of list entry e of tile target tt in the tile's slot list: lane e % kTileLanes of the
// target reads its entries e, e + kTileLanes, ... four at a time as one 8-byte word, so the
// entries of group g (entry e = l + g kTileLanes) sit at word (g / 4, tt, l), half-word g % 4.
__device__ inline int tl_index(int tt, int e)
{
    const int g = e / kTileLanes, l = e % kTileLanes;
    return (((g >> 2)*kTileTargets + tt)*kTileLanes + l)*4 + (g & 3);
}

// Union of a tile's list entries (list builds only).  A tile's targets lie in one cell row
// (cells kx0..kx1), so their skin-list neighbours lie in the W^(DIM-1) rows around it, cells
// kx0 - REACH .. kx1 + REACH: W^(DIM-1) contiguous runs of the sorted order.  An entry's cell
// code names its run, and the runs laid end to end index two bitmaps (free neighbours and
// walls); their prefix popcounts give each neighbour its item slot (free ones first, each kind
// in run order), and the targets' lists are rewritten as slots.  The same neighbour met in
// several lists writes the same item.  A tile whose runs exceed the bitmaps, whose items exceed
// the LDS staging, or with an entry outside its runs is marked to read from HBM.
template <int REACH>
__global__ void __launch_bounds__(512) k_tile_union(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt || !D.ctl->rebuild) return;
    constexpr int W = Codes<REACH>::W, NQ = W*W;
    __shared__ uint32_t bmf[kTileWords], bmw[kTileWords];
    __shared__ uint16_t pf[kTileWords], pw[kTileWords];
    __shared__ int rbeg[NQ], rlen[NQ], roff[NQ + 1];
    __shared__ int scnt[kTileTargets], smax[2];
    __shared__ int sh[17];
    __shared__ int bad;
    const int ntiles = D.ctl->ntiles, cap = P.nl_cap;
    const size_t gs = (size_t)P.nfree;
    // a bounded grid walks the tiles, so steps without a list build cost one wave of blocks
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int4 ti = D.tile[tile];
        const int t0 = ti.x, nt = ti.y;
        __syncthreads();                       // the previous tile's reads of the shared arrays
        if (threadIdx.x < 64) {
            // wave 0, lane q < NQ: run q = code / W, the row (dy, dz) = (q % W, q / W) - REACH
            // (dz = 0 in 2-D); the runs' offsets by a wave scan
            const int q = threadIdx.x;
            int beg = 0, len = 0;
            if (q < NQ) {
                int kx0, kx1, ky, kz;
                decode_key(P, D.key[D.free_list[t0]], kx0, ky, kz);
                decode_key(P, D.key[D.free_list[t0 + nt - 1]], kx1, ky, kz);
                const int yy = ky + q % W - REACH, zz = kz + q / W - REACH;
                if (yy >= 0 && yy < P.nc[1] && zz >= 0 && zz < P.nc[2]) {
                    const int base = P.nc[0]*(yy + P.nc[1]*zz);
                    const int xa = max(kx0 - REACH, 0), xb = min(kx1 + REACH, P.nc[0] - 1);
                    beg = D.cstart[base + xa];
                    len = D.cstart[base + xb + 1] - beg;
                }
            }
            int x = len;
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(x, o, 64);
                if (q >= o) x += y;
            }
            if (q < NQ) { rbeg[q] = beg; rlen[q] = len; roff[q] = x - len; }
            if (q == NQ - 1) roff[NQ] = x;
            if (q == 0) bad = 0;
        } else if (threadIdx.x < 64 + 128) {
            // waves 1 and 2: the targets' entry counts and their maximum
            static_assert(kTileTargets <= 128, "two waves hold a tile's counts");
            const int tt = (int)threadIdx.x - 64;
            int m = 0;
            if (tt < nt) { const int c = D.ncnt[t0 + tt]; m = c < cap ? c : cap; }
            if (tt < kTileTargets) scnt[tt] = m;
            for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
            if ((threadIdx.x & 63) == 0) smax[((int)threadIdx.x >> 6) - 1] = m;
        }
        __syncthreads();
        const int ngrp = (max(smax[0], smax[1]) + 3) >> 2;
        const int nwords = (roff[NQ] + 31) >> 5;
        if (nwords > kTileWords) {
            if (threadIdx.x == 0) D.tile[tile] = make_int4(t0, nt, 0, -1);
            continue;
        }
        for (int w = threadIdx.x; w < nwords; w += blockDim.x) { bmf[w] = 0u; bmw[w] = 0u; }
        __syncthreads();
        auto bit_of = [&](int v) {
            const int j = nl_index<REACH>(v), q = nl_code<REACH>(v)/W;
            const int d = j - rbeg[q];
            return ((unsigned)d < (unsigned)rlen[q]) ? roff[q] + d : -1;
        };
        // the lists by groups of four entries: group g of the tile's targets is one contiguous
        // run of nt 16-B words, so each thread loads whole words and the neighbours' flags of
        // its four entries together
        const int nwd = ngrp*nt;
        const int4 *l4 = reinterpret_cast<const int4 *>(D.nl);
        for (int k = threadIdx.x; k < nwd; k += blockDim.x) {
            const int g = k / nt, tt = k - g*nt, m = scnt[tt];
            if (4*g >= m) continue;
            const int4 w4 = l4[(size_t)g*gs + (size_t)(t0 + tt)];
            const int vs[4] = {w4.x, w4.y, w4.z, w4.w};
            uint32_t fl[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) fl[u] = (4*g + u < m) ? D.flags_b[nl_index<REACH>(vs[u])] : 0u;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (4*g + u >= m) break;
                const int b = bit_of(vs[u]);
                if (b < 0) { bad = 1; continue; }
                atomicOr(flag_free(fl[u]) ? &bmf[b >> 5] : &bmw[b >> 5], 1u << (b & 31));
            }
        }
        __syncthreads();
        if (bad) {
            if (threadIdx.x == 0) D.tile[tile] = make_int4(t0, nt, 0, -1);
            continue;
        }
        // prefix popcounts: each thread a contiguous run of words
        const int per = (nwords + (int)blockDim.x - 1)/(int)blockDim.x;
        const int w0 = (int)threadIdx.x*per, w1 = min(w0 + per, nwords);
        int cf = 0, cw = 0;
        for (int w = w0; w < w1; ++w) { cf += __popc(bmf[w]); cw += __popc(bmw[w]); }
        int uf = 0, uw = 0;
        int of = block_scan(cf, sh, &uf);
        int ow = block_scan(cw, sh, &uw);
        if (uf + uw > P.tile_items || uw > kTileWalls) {
            if (threadIdx.x == 0) D.tile[tile] = make_int4(t0, nt, 0, -1);
            continue;
        }
        for (int w = w0; w < w1; ++w) {
            pf[w] = (uint16_t)of; of += __popc(bmf[w]);
            pw[w] = (uint16_t)ow; ow += __popc(bmw[w]);
        }
        __syncthreads();
        int32_t *items = D.items + (size_t)tile*kTileItems;
        uint16_t *tl = D.tl + (size_t)tile*P.tl_stride;
        for (int k = threadIdx.x; k < nwd; k += blockDim.x) {
            const int g = k / nt, tt = k - g*nt, m = scnt[tt];
            if (4*g >= m) continue;
            const int4 w4 = l4[(size_t)g*gs + (size_t)(t0 + tt)];
            const int vs[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = 4*g + u;
                if (e >= m) break;
                const int b = bit_of(vs[u]), w = b >> 5;
                const uint32_t bit = 1u << (b & 31), below = bit - 1u;
                const int slot = (bmf[w] & bit) ? pf[w] + __popc(bmf[w] & below)
                                                : uf + pw[w] + __popc(bmw[w] & below);
                items[slot] = nl_index<REACH>(vs[u]);
                tl[tl_index(tt, e)] = (uint16_t)slot;
            }
        }
        if (threadIdx.x == 0) D.tile[tile] = make_int4(t0, nt, uf, uw);
    }
}

// The FP64 gather per tile.  Sums: lane l of a target adds entries l, l + kTileLanes, ... in
// list order, the first lane starting from the target's initial acceleration; the lanes' sums
// meet as ((l0 + l1) + (l2 + l3)) + ((l4 + l5) + (l6 + l7)).  A target whose list overflowed
// queues the result for k_force11_nl_over, which adds the entries past the capacity.
template <int DIM, int ABM>
__global__ void __launch_bounds__(kTileTargets*kTileLanes) k_force11_tile(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    const int nb = D.ctl->ntiles;
    if ((int)blockIdx.x >= nb) return;
    // the staged records in 16-B halves (x y | z d | vx vy | vz P/d^2 | c meta | m 1/d; walls'
    // viscous velocity), so that lanes reading different slots spread over all the LDS banks
    __shared__ double2 sxy[kTileItems], szd[kTileItems], svv[kTileItems], svp[kTileItems],
                       scm[kTileItems], smr[kTileItems], swa[kTileWalls], swb[kTileWalls];
    __shared__ double pa[64], pb[64];
    const int tile = xcd_remap((int)blockIdx.x, nb);
    const int4 ti = D.tile[tile];
    const int t0 = ti.x, nt = ti.y, uf = ti.z, uw = ti.w;
    const bool staged = uw >= 0;
    const int tid = threadIdx.x;
    if (tid < 64) { pa[tid] = P.pair_alpha[tid]; pb[tid] = P.pair_beta[tid]; }
    const int32_t *items = D.items + (size_t)tile*kTileItems;
    // the target's own records first, so their loads are in flight during the staging
    const int tt = tid / kTileLanes, l = tid % kTileLanes;
    const bool act = tt < nt;
    const int t = t0 + (act ? tt : 0);
    const int i = D.free_list[t];
    const double4 gi0 = D.g0[i], gi1 = D.g1[i], gi2 = D.g2[i];
    const int cap = P.nl_cap;
    const int cnt = act ? D.ncnt[t] : 0;
    const int m = cnt < cap ? cnt : cap;
    const uint64_t *tq = reinterpret_cast<const uint64_t *>(D.tl + (size_t)tile*P.tl_stride) +
                         tt*kTileLanes + l;
    constexpr size_t kWordStride = (size_t)kTileTargets*kTileLanes;
    // the lane's first kSlotWords slot words (its entries l + 8 g, g < 4 kSlotWords) are loaded
    // here, in flight during the staging; longer lists read the rest in the loop
    uint64_t qw[kSlotWords];
#pragma unroll
    for (int G = 0; G < kSlotWords; ++G)
        qw[G] = (staged && l + 4*kTileLanes*G < m) ? tq[G*kWordStride] : 0ull;
    if (staged) {
        // at most two items per thread (kTileItems <= 2 blockDim.x): both rounds' loads are
        // issued before the LDS stores
        const int U = uf + uw, u0 = tid, u1 = tid + (int)blockDim.x;
        const int j0 = u0 < U ? items[u0] : -1, j1 = u1 < U ? items[u1] : -1;
        double4 r00, r01, r02, r0w, r10, r11, r12, r1w;
        if (j0 >= 0) { r00 = D.g0[j0]; r01 = D.g1[j0]; r02 = D.g2[j0]; if (u0 >= uf) r0w = D.gw[j0]; }
        if (j1 >= 0) { r10 = D.g0[j1]; r11 = D.g1[j1]; r12 = D.g2[j1]; if (u1 >= uf) r1w = D.gw[j1]; }
        auto put = [&](int u, double4 r0, double4 r1, double4 r2) {
            sxy[u] = make_double2(r0.x, r0.y); szd[u] = make_double2(r0.z, r0.w);
            svv[u] = make_double2(r1.x, r1.y); svp[u] = make_double2(r1.z, r1.w);
            scm[u] = make_double2(r2.x, r2.y); smr[u] = make_double2(r2.z, r2.w);
        };
        if (j0 >= 0) {
            put(u0, r00, r01, r02);
            if (u0 >= uf) { swa[u0 - uf] = make_double2(r0w.x, r0w.y); swb[u0 - uf] = make_double2(r0w.z, r0w.w); }
        }
        if (j1 >= 0) {
            put(u1, r10, r11, r12);
            if (u1 >= uf) { swa[u1 - uf] = make_double2(r1w.x, r1w.y); swb[u1 - uf] = make_double2(r1w.z, r1w.w); }
        }
    }
    __syncthreads();
    const int ci = (int)gi2.y;
    const double vsf = 2.0*P.mu_ff*gi2.w, vsw = 2.0*P.mu_ff_wall*gi2.w;   // 2 mu / d_i
    Acc11 A;
    {
        const double4 a_init = (act && l == 0) ? a_start(D, P, i) : make_double4(0.0, 0.0, 0.0, 0.0);
        A.a0 = a_init.x; A.a1 = a_init.y; A.a2 = a_init.z; A.dd = a_init.w;
    }
    unsigned npc = 0;
    const int ki = D.key[i];
    if (staged) {
        // slots from the lane's slot words, records from LDS (a wall's viscous velocity only
        // for walls; a free neighbour's is its v)
        const auto ab = ABPick<ABM, double>::make(P, pa, pb);
        uint64_t q = 0ull;
        for (int e = l, g = 0; e < m; e += kTileLanes, ++g) {
            if ((g & 3) == 0) {
                // this lane's next four slots (g is the same in every lane of the wave)
                const int G = g >> 2;
                q = qw[0];
#pragma unroll
                for (int k = 1; k < kSlotWords; ++k) q = (G == k) ? qw[k] : q;
                if (G >= kSlotWords) q = tq[(size_t)G*kWordStride];
            }
            const int slot = (int)((q >> (16*(g & 3))) & 0xffffu);
            const double2 r0 = sxy[slot], r1 = szd[slot], r2 = svv[slot], r3 = svp[slot],
                          r4 = scm[slot], r5 = smr[slot];
            const double4 a0 = make_double4(r0.x, r0.y, r1.x, r1.y), a1 = make_double4(r2.x, r2.y, r3.x, r3.y),
                          a2 = make_double4(r4.x, r4.y, r5.x, r5.y);
            double4 w = a1;
            if (r4.y < 0.0) {
                const double2 wa = swa[slot - uf], wb = swb[slot - uf];
                w = make_double4(wa.x, wa.y, wb.x, wb.y);
            }
            pair11_rec<DIM, true>(D, P, ab, gi0, gi1, gi2, ci, ki, [&] { return items[slot]; },
                                  a0, a1, a2, w, vsf, vsw, A, npc);
        }
    } else {
        const size_t gs = (size_t)P.nfree;
        for (int e = l; e < m; e += kTileLanes) {
            const int j = nl_index<2>(D.nl[(((size_t)(e >> 2))*gs + (size_t)t)*4 + (e & 3)]);
            pair11_fast<DIM, true>(D, P, pa, pb, gi0, gi1, gi2, ci, ki, j, vsf, vsw, A, npc);
        }
    }
    // the lanes' sums meet in a fixed order
    for (int o = 1; o < kTileLanes; o <<= 1) {
        A.a0 += __shfl_xor(A.a0, o, 64);
        A.a1 += __shfl_xor(A.a1, o, 64);
        A.a2 += __shfl_xor(A.a2, o, 64);
        A.dd += __shfl_xor(A.dd, o, 64);
    }
    unsigned long long dtb = ~0ull;
    if (act && l == 0) {
        if (cnt > cap) {
            int q = atomicAdd(&D.ctl->nl_q, 1);
            D.ovq[q] = t;
            D.ovs[2*q] = make_double4(A.a0, A.a1, A.a2, A.dd);
        } else {
            D.acc[i] = make_double4(A.a0, A.a1, A.a2, A.dd);
            double na = sqrt(A.a0*A.a0 + A.a1*A.a1 + A.a2*A.a2);
            double cand = P.sqrt_h_a*sqrt(P.h_uniform/na);
            dtb = dbits(cand);
        }
    }
    force_partials(D, dtb, tally_free(npc), tally_fixed(npc));
}

// ------------------------------------------------------------------------------------------
// FP32-mixed fluid force pass: the fast-path formulas in FP32 on cell-relative coordinates
// (x_i - x_j = (r_i - r_j) - dk CellSize, r measured from each particle's cell origin), FP32
// accumulation of the momentum sum and a Kahan-compensated FP32 continuity sum; results are
// widened to FP64 for the FP64 integration.
// ------------------------------------------------------------------------------------------
struct Acc32 { float a0, a1, a2, dd, ddc; };

// Skin-list membership on the FP32 squared distance: its error is below 1e-6 of the support
// radius squared (cell-relative positions), so outside a 1e-4 band around the support radius
// and beyond 1e-4 below the smallest cell size it decides as the FP64 test would; inside the
// band the FP64 test runs on the FP64 positions.
struct Band32 {
    float lo, hi, adj;
    __device__ explicit Band32(const Params &P) : lo(P.f32.band_lo), hi(P.f32.band_hi), adj(P.f32.band_adj) {}
};

__device__ inline bool skin_member64(const Dev &D, const Params &P, int i, int j)
{
    double4 pi = D.pos_b[i], pj = D.pos_b[j];
    double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
    return skin_member(D, P, i, j, x0*x0 + x1*x1 + x2*x2);
}

__device__ inline bool skin_member32(const Dev &D, const Params &P, const Band32 &B, int i, int j, float r2)
{
    if (r2 > B.hi) return false;
    if (r2 >= B.lo) return skin_member64(D, P, i, j);
    return r2 <= B.adj || keys_adjacent(P, D.key[i], D.key[j]);
}

// SKIN: the neighbour comes from a skin list and counts only if it is a pair of the
// reference's set at the current positions (skin_member32); i is the target.  The neighbour's
// records are a0..a2 and vw (its viscous velocity, read only for walls); jf() gives its index,
// called only inside the FP32 test's band.  o = (k_j - k_i) CellSize, the offset of the two
// particles' cells at the last sort.
template <int DIM, bool SKIN, class AB, class JF>
__device__ inline void pair11_f32_rec(const Dev &D, const Params &P, const AB &ab,
                                      float4 gi0, float4 gi1, float4 gi2, int ci, int i, JF jf,
                                      float4 a0, float4 a1, float4 a2, float4 vw,
                                      float o0, float o1, float o2, Acc32 &A, unsigned &npc,
                                      float kw, float kg, float kgh, float invh, float h, float h2c,
                                      float vsf, float vsw, const Band32 &band)
{
    const float cj_snd = a2.x, meta = a2.y, mrec = a2.z, invdj = a2.w;
    float x0 = (gi0.x - a0.x) - o0, x1 = (gi0.y - a0.y) - o1, x2 = (gi0.z - a0.z) - o2;
    float r2 = x0*x0 + x1*x1 + x2*x2;
    if (SKIN) {
        if (r2 > band.hi) return;
        if (r2 >= band.lo) { if (!skin_member64(D, P, i, jf())) return; }
        else if (!(r2 <= band.adj || keys_adjacent(P, D.key[i], D.key[jf()]))) return;
    }
    float invr = __builtin_amdgcn_rsqf(r2);
    float rij = r2 > 0.0f ? r2*invr : 0.0f;
    float q = rij*invh;
    float GK;
    if (q < 1.0f) {
        GK = kgh*(-3.0f+2.25f*q);
    } else if (q < 2.0f) {
        float t = 2.0f - q;
        GK = kg*invr*(-0.75f*t*t);
    } else {
        GK = 0.0f;
    }
    (void)kw;
    const bool freej = meta >= 0.0f;
    npc += pair_tally(freej);
    const float w0 = vw.x, w1 = vw.y, w2 = vw.z;         // viscous velocity (a free neighbour's v)
    float nha, beta;
    ab.get(ci, meta, freej, nha, beta);
    float u0 = gi1.x - a1.x, u1 = gi1.y - a1.y, u2 = gi1.z - a1.z;
    float dvx = u0*x0 + u1*x1 + u2*x2;
    float di = gi0.w, dj = a0.w;
    float PIij = 0.0f;
    if (ab.on(nha, beta, freej) && dvx < 0.0f) {
        float MU = h*dvx*__builtin_amdgcn_rcpf(r2 + h2c);
        PIij = (nha*(gi2.x + cj_snd)*MU + beta*MU*MU)*__builtin_amdgcn_rcpf(0.5f*(di + dj));
    }
    float mi = gi2.z;
    float mj = freej ? mrec : mrec*mi;
    float vs = (freej ? vsf : vsw)*invdj*GK;                 // 2 mu / d_i by the neighbour's kind
    float s = -(gi1.w + a1.w + PIij)*GK;
    float b0 = gi1.x - w0, b1 = gi1.y - w1, b2 = gi1.z - w2;
    float t0 = s*x0 + vs*b0, t1 = s*x1 + vs*b1, t2 = s*x2 + vs*b2;
    if (DIM == 2) t2 = 0.0f;
    A.a0 += mj*t0; A.a1 += mj*t1; A.a2 += mj*t2;
    // Kahan-compensated continuity sum
    float c = mj*(di*invdj)*(dvx*GK);
    float y = c - A.ddc;
    float tt = A.dd + y;
    A.ddc = (tt - A.dd) - y;
    A.dd = tt;
}

template <int DIM, bool SKIN>
__device__ inline void pair11_f32(const Dev &D, const Params &P, const float *pa, const float *pb,
                                  float4 gi0, float4 gi1, float4 gi2, int ci, int i, int j,
                                  float o0, float o1, float o2, Acc32 &A, unsigned &npc,
                                  float kw, float kg, float kgh, float invh, float h, float h2c,
                                  float vsf, float vsw, const Band32 &band)
{
    const float4 a0 = D.f0[j], a1 = D.f1[j], a2 = D.f2[j];
    const float4 vw = (a2.y < 0.0f) ? D.fw[j] : a1;
    pair11_f32_rec<DIM, SKIN>(D, P, ABTable<float>{pa, pb}, gi0, gi1, gi2, ci, i, [j] { return j; }, a0, a1, a2, vw,
                              o0, o1, o2, A, npc, kw, kg, kgh, invh, h, h2c, vsf, vsw, band);
}

// FP32-mixed fluid force pass over the neighbour lists: the list's cell code gives the cell
// offset dk CellSize of the cell-relative FP32 positions (each relative to the particle's cell
// at the last sort).
template <int DIM, int REACH>
__global__ void __launch_bounds__(kForceBlock) k_force11_f32_nl(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    __shared__ float pa[64], pb[64];
    if (threadIdx.x < 64) { pa[threadIdx.x] = (float)P.pair_alpha[threadIdx.x]; pb[threadIdx.x] = (float)P.pair_beta[threadIdx.x]; }
    __syncthreads();
    const int tid = threadIdx.x;
    int t = xcd_block()*kForceBlock + tid;
    unsigned long long dtb = ~0ull;      // no dt candidate from an inactive lane
    bool act = t < P.nfree;
    int i = act ? D.free_list[t] : 0;
    float4 gi0 = D.f0[i], gi1 = D.f1[i], gi2 = D.f2[i];
    int ci = (int)gi2.y;
    const float csx = P.f32.cs[0], csy = P.f32.cs[1], csz = P.f32.cs[2];
    const float kw = P.f32.kw, kg = P.f32.kg, kgh = P.f32.kgh, invh = P.f32.invh;
    const float h = P.f32.h, h2c = P.f32.h2c;
    const float vsf = P.f32.mu2_ff*gi2.w, vsw = P.f32.mu2_wall*gi2.w;   // 2 mu / d_i
    const Band32 band(P);
    Acc32 A;
    A.a0 = 0.0f; A.a1 = 0.0f; A.a2 = 0.0f; A.dd = 0.0f; A.ddc = 0.0f;
    unsigned npc = 0;
    const int cap = P.nl_cap;
    int cnt = act ? D.ncnt[t] : 0;
    int m = cnt < cap ? cnt : cap;
    const int4 *lst = reinterpret_cast<const int4 *>(D.nl) + t;
    const size_t gs = (size_t)P.nfree;
    auto pair = [&](int j, int c) {
        int dx, dy, dz;
        code_offset<REACH>(c, dx, dy, dz);
        float p0 = dx*csx, p1 = dy*csy, p2 = (DIM == 3) ? dz*csz : 0.0f;
        pair11_f32<DIM, REACH == 2>(D, P, pa, pb, gi0, gi1, gi2, ci, i, j, p0, p1, p2,
                                    A, npc, kw, kg, kgh, invh, h, h2c, vsf, vsw, band);
    };
    int4 w = make_int4(0, 0, 0, 0);
    for (int e = 0; e < m; ++e) {
        int sl = e & 3;
        if (sl == 0) w = lst[(size_t)(e >> 2)*gs];
        int v = (sl == 0) ? w.x : (sl == 1) ? w.y : (sl == 2) ? w.z : w.w;
        pair(nl_index<REACH>(v), nl_code<REACH>(v));
    }
    if (act && cnt > cap) {
        // the list overflowed: k_force11_f32_nl_over continues this target from these sums
        int q = atomicAdd(&D.ctl->nl_q, 1);
        D.ovq[q] = t;
        D.ovs[2*q] = make_double4((double)A.a0, (double)A.a1, (double)A.a2, (double)A.dd);
        D.ovs[2*q+1] = make_double4((double)A.ddc, 0.0, 0.0, 0.0);
        act = false;
    }
    if (act) {
        const double4 ai = a_start(D, P, i);
        double a0 = ai.x + (double)A.a0, a1 = ai.y + (double)A.a1, a2 = ai.z + (double)A.a2;
        D.acc[i] = make_double4(a0, a1, a2, ai.w + ((double)A.dd + (double)(-A.ddc)));
        double na = sqrt(a0*a0 + a1*a1 + a2*a2);
        double cand = P.sqrt_h_a*sqrt(P.h_uniform/na);
        dtb = dbits(cand);
    }
    force_partials(D, dtb, tally_free(npc), tally_fixed(npc));
}

// Ends the overflowed targets of the FP32 list and tile passes (see k_force11_nl_over).  After
// the list pass the queued FP32 sums (exact in FP64) are restored as FP32 and continued in list
// order; after the tile pass (P.tiles) the entries past the capacity are summed in FP32 from
// zero and added in FP64 to the queued FP64 sums of the first ones.
template <int DIM, int REACH>
__global__ void __launch_bounds__(kForceBlock) k_force11_f32_nl_over(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    const int nq = D.ctl->nl_q;
    if (nq == 0) return;
    __shared__ float pa[64], pb[64];
    if (threadIdx.x < 64) { pa[threadIdx.x] = (float)P.pair_alpha[threadIdx.x]; pb[threadIdx.x] = (float)P.pair_beta[threadIdx.x]; }
    __syncthreads();
    const float csx = P.f32.cs[0], csy = P.f32.cs[1], csz = P.f32.cs[2];
    const float kw = P.f32.kw, kg = P.f32.kg, kgh = P.f32.kgh, invh = P.f32.invh;
    const float h = P.f32.h, h2c = P.f32.h2c;
    const int cap = P.nl_cap;
    const Band32 band(P);
    const double reach2 = P.reach2;
    const double cull2 = ((REACH == 2) ? reach2 : P.sup2_hi)*(1.0 + 1e-9);
    for (int q = blockIdx.x*kForceBlock + threadIdx.x; q < nq; q += gridDim.x*kForceBlock) {
        int t = D.ovq[q];
        int i = D.free_list[t];
        float4 gi0 = D.f0[i], gi1 = D.f1[i], gi2 = D.f2[i];
        int ci = (int)gi2.y;
        const float vsf = P.f32.mu2_ff*gi2.w, vsw = P.f32.mu2_wall*gi2.w;   // 2 mu / d_i
        double4 s0 = D.ovs[2*q];
        Acc32 A;
        A.a0 = 0.0f; A.a1 = 0.0f; A.a2 = 0.0f; A.dd = 0.0f; A.ddc = 0.0f;
        if (!P.tiles) {
            A.a0 = (float)s0.x; A.a1 = (float)s0.y; A.a2 = (float)s0.z; A.dd = (float)s0.w;
            A.ddc = (float)D.ovs[2*q+1].x;
            s0 = make_double4(0.0, 0.0, 0.0, 0.0);
        }
        unsigned npc = 0;
        int kx, ky, kz;
        decode_key(P, D.key[i], kx, ky, kz);
        double4 pi = D.pos_b[i];
        int k = 0;
        walk_rows<DIM, REACH>(D, P, i, pi.x, pi.y, pi.z, kx, ky, kz, cull2, [&](int j, int c, double r2) {
            if (!((REACH == 2) ? (r2 <= reach2) : in_support(P, r2))) return;      // a list entry
            if (k >= cap) {
                int dx, dy, dz;
                code_offset<REACH>(c, dx, dy, dz);
                float p0 = dx*csx, p1 = dy*csy, p2 = (DIM == 3) ? dz*csz : 0.0f;
                pair11_f32<DIM, REACH == 2>(D, P, pa, pb, gi0, gi1, gi2, ci, i, j, p0, p1, p2,
                                            A, npc, kw, kg, kgh, invh, h, h2c, vsf, vsw, band);
            }
            ++k;
        });
        const double4 ai = a_start(D, P, i);
        double a0 = ai.x + (s0.x + (double)A.a0), a1 = ai.y + (s0.y + (double)A.a1), a2 = ai.z + (s0.z + (double)A.a2);
        D.acc[i] = make_double4(a0, a1, a2, ai.w + (s0.w + ((double)A.dd + (double)(-A.ddc))));
        double na = sqrt(a0*a0 + a1*a1 + a2*a2);
        double cand = P.sqrt_h_a*sqrt(P.h_uniform/na);
        atomicMin(&D.ctl->dtmin_bits, dbits(cand));
        if (tally_free(npc)) atomicAdd(&D.ctl->pairs_ff, (unsigned long long)tally_free(npc));
        if (tally_fixed(npc)) atomicAdd(&D.ctl->pairs_fw, (unsigned long long)tally_fixed(npc));
    }
}

// The FP32-mixed gather per tile (the tile table and union of k_force11_tile): each
// neighbour's three FP32 records and its cell at the last sort (kx | ky << 10 | kz << 20) are
// staged in LDS, walls' viscous velocities beside them; the pair is pair11_f32's, with the cell
// offset from the two cells.  Lane l of a target adds entries l, l + kTileLanes, ... in list order
// in FP32 (Kahan-compensated continuity per lane); the lanes' sums are widened and meet in FP64
// in the fixed order of k_force11_tile.  Overflowed targets go to k_force11_f32_nl_over.
// slot words a lane of the FP32 tile gather holds in registers (16 entries per lane, 128 per
// target; longer lists read the rest in the loop): fewer than the FP64 gather's, so that the
// kernel fits 80 VGPRs and three workgroups share a CU
constexpr int kSlotWords32 = 4;
template <int DIM, int ABM>
__global__ void __launch_bounds__(kTileTargets*kTileLanes, 6) k_force32_tile(Dev D0, Params P)
{
    const Dev D = bind(D0);
    if (D.ctl->halt) return;
    const int nb = D.ctl->ntiles;
    if ((int)blockIdx.x >= nb) return;
    __shared__ float4 s0[kTileItems], s1[kTileItems], s2[kTileItems], sw[kTileWalls];
    __shared__ int sk[kTileItems];
    __shared__ float pa[64], pb[64];
    const int tile = xcd_remap((int)blockIdx.x, nb);
    const int4 ti = D.tile[tile];
    const int t0 = ti.x, nt = ti.y, uf = ti.z, uw = ti.w;
    const bool staged = uw >= 0;
    const int tid = threadIdx.x;
    if (tid < 64) { pa[tid] = (float)P.pair_alpha[tid]; pb[tid] = (float)P.pair_beta[tid]; }
    const int32_t *items = D.items + (size_t)tile*kTileItems;
    const int tt = tid / kTileLanes, l = tid % kTileLanes;
    const bool act = tt < nt;
    const int t = t0 + (act ? tt : 0);
    const int i = D.free_list[t];
    const float4 gi0 = D.f0[i], gi1 = D.f1[i], gi2 = D.f2[i];
    const int cap = P.nl_cap;
    const int cnt = act ? D.ncnt[t] : 0;
    const int m = cnt < cap ? cnt : cap;
    const uint64_t *tq = reinterpret_cast<const uint64_t *>(D.tl + (size_t)tile*P.tl_stride) +
                         tt*kTileLanes + l;
    constexpr size_t kWordStride = (size_t)kTileTargets*kTileLanes;
    uint64_t qw[kSlotWords32];
#pragma unroll
    for (int G = 0; G < kSlotWords32; ++G)
        qw[G] = (staged && l + 4*kTileLanes*G < m) ? tq[G*kWordStride] : 0ull;
    auto pack = [&](int key) {
        int kx, ky, kz;
        decode_key(P, key, kx, ky, kz);
        return kx | (ky << 10) | (kz << 20);
    };
    if (staged) {
        const int U = uf + uw, u0 = tid, u1 = tid + (int)blockDim.x;
        const int j0 = u0 < U ? items[u0] : -1, j1 = u1 < U ? items[u1] : -1;
        float4 r00, r01, r02, r0w, r10, r11, r12, r1w;
        int k0 = 0, k1 = 0;
        if (j0 >= 0) { r00 = D.f0[j0]; r01 = D.f1[j0]; r02 = D.f2[j0]; if (u0 >= uf) r0w = D.fw[j0]; k0 = D.key_b[j0]; }
        if (j1 >= 0) { r10 = D.f0[j1]; r11 = D.f1[j1]; r12 = D.f2[j1]; if (u1 >= uf) r1w = D.fw[j1]; k1 = D.key_b[j1]; }
        if (j0 >= 0) { s0[u0] = r00; s1[u0] = r01; s2[u0] = r02; sk[u0] = pack(k0); if (u0 >= uf) sw[u0 - uf] = r0w; }
        if (j1 >= 0) { s0[u1] = r10; s1[u1] = r11; s2[u1] = r12; sk[u1] = pack(k1); if (u1 >= uf) sw[u1 - uf] = r1w; }
    }
    __syncthreads();
    const int ci = (int)gi2.y;
    const float csx = P.f32.cs[0], csy = P.f32.cs[1], csz = P.f32.cs[2];
    const float kw = P.f32.kw, kg = P.f32.kg, kgh = P.f32.kgh, invh = P.f32.invh;
    const float h = P.f32.h, h2c = P.f32.h2c;
    const float vsf = P.f32.mu2_ff*gi2.w, vsw = P.f32.mu2_wall*gi2.w;   // 2 mu / d_i
    const Band32 band(P);
    Acc32 A;
    A.a0 = 0.0f; A.a1 = 0.0f; A.a2 = 0.0f; A.dd = 0.0f; A.ddc = 0.0f;
    unsigned npc = 0;
    if (staged) {
        // a wall's viscous velocity read only for walls; a free neighbour's is its v
        const auto ab = ABPick<ABM, float>::make(P, pa, pb);
        const int ki = pack(D.key_b[i]);
        const int kix = ki & 1023, kiy = (ki >> 10) & 1023, kiz = ki >> 20;
        uint64_t q = 0ull;
        for (int e = l, g = 0; e < m; e += kTileLanes, ++g) {
            if ((g & 3) == 0) {
                const int G = g >> 2;
                q = qw[0];
#pragma unroll
                for (int k = 1; k < kSlotWords32; ++k) q = (G == k) ? qw[k] : q;
                if (G >= kSlotWords32) q = tq[(size_t)G*kWordStride];
            }
            const int slot = (int)((q >> (16*(g & 3))) & 0xffffu);
            const float4 a0 = s0[slot], a1 = s1[slot], a2 = s2[slot];
            float4 w = a1;
            if (a2.y < 0.0f) w = sw[slot - uf];
            const int kj = sk[slot];
            const float o0 = (float)((kj & 1023) - kix)*csx, o1 = (float)(((kj >> 10) & 1023) - kiy)*csy;
            const float o2 = (DIM == 3) ? (float)((kj >> 20) - kiz)*csz : 0.0f;
            pair11_f32_rec<DIM, true>(D, P, ab, gi0, gi1, gi2, ci, i, [&] { return items[slot]; },
                                      a0, a1, a2, w, o0, o1, o2, A, npc, kw, kg, kgh, invh, h, h2c,
                                      vsf, vsw, band);
        }
    } else {
        const size_t gs = (size_t)P.nfree;
        for (int e = l; e < m; e += kTileLanes) {
            const int v = D.nl[(((size_t)(e >> 2))*gs + (size_t)t)*4 + (e & 3)];
            int dx, dy, dz;
            code_offset<2>(nl_code<2>(v), dx, dy, dz);
            const float p0 = dx*csx, p1 = dy*csy, p2 = (DIM == 3) ? dz*csz : 0.0f;
            pair11_f32<DIM, true>(D, P, pa, pb, gi0, gi1, gi2, ci, i, nl_index<2>(v), p0, p1, p2,
                                  A, npc, kw, kg, kgh, invh, h, h2c, vsf, vsw, band);
        }
    }
    double b0 = (double)A.a0, b1 = (double)A.a1, b2 = (double)A.a2, bd = (double)A.dd + (double)(-A.ddc);
    for (int o = 1; o < kTileLanes; o <<= 1) {
        b0 += __shfl_xor(b0, o, 64);
        b1 += __shfl_xor(b1, o, 64);
        b2 += __shfl_xor(b2, o, 64);
        bd += __shfl_xor(bd, o, 64);
    }
    unsigned long long dtb = ~0ull;
    if (act && l == 0) {
        if (cnt > cap) {
            int q = atomicAdd(&D.ctl->nl_q, 1);
            D.ovq[q] = t;
            D.ovs[2*q] = make_double4(b0, b1, b2, bd);
        } else {
            const double4 ai = a_start(D, P, i);
            const double a0 = ai.x + b0, a1 = ai.y + b1, a2 = ai.z + b2;
            D.acc[i] = make_double4(a0, a1, a2, ai.w + bd);
            double na = sqrt(a0*a0 + a1*a1 + a2*a2);
            double cand = P.sqrt_h_a*sqrt(P.h_uniform/na);
            dtb = dbits(cand);
        }
    }
    force_partials(D, dtb, tally_free(npc), tally_fixed(npc));
}

// ------------------------------------------------------------------------------------------
// integrate: AdaptiveTimeStep (Domain.cpp:133-152), Move (Particle.cpp:119-274, 561-596),
// Mat1/Mat2/Mat3 (Particle.cpp:146-186, 276-482, 598-795), CheckParticleLeave test
// (Domain.cpp:517-547) and the next step's cell keys
// ------------------------------------------------------------------------------------------
__device__ inline double adaptive_dt(double deltat, double deltatint, double deltatmin)
{
    if (deltatint > deltatmin) {
        if (deltat < deltatmin) return 2.0*deltat*deltatmin/(deltat+deltatmin);
        return deltatmin;
    }
    if (deltatint != deltat) return 2.0*deltat*deltatint/(deltat+deltatint);
    return deltatint;
}

__device__ inline double contract6(const double *S)
{
    // S00*S00 + 2 S01 S10 + 2 S02 S20 + S11*S11 + 2 S12 S21 + S22*S22 (Particle.cpp:149-151)
    return S[0]*S[0] + 2.0*S[1]*S[1] + 2.0*S[2]*S[2] + S[3]*S[3] + 2.0*S[4]*S[4] + S[5]*S[5];
}

// R = sum of the Jaumann rate terms for Mat2 (deviatoric, on S) or Mat3 (total, on Sigma)
__device__ inline void jaumann_rate(bool total, const double *S6, const double *sr6, const double *rr3,
                                    double G, double Kb, double R6[6])
{
    double S[9], W[9], WT[9], SRT[9], RS[9];
    sym_to_full(S6, S);
    W[0] = 0.0;     W[1] = rr3[0];  W[2] = rr3[1];
    W[3] = -rr3[0]; W[4] = 0.0;     W[5] = rr3[2];
    W[6] = -rr3[1]; W[7] = -rr3[2]; W[8] = 0.0;
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) WT[3*b+a] = W[3*a+b];
    mat_mult(S, WT, SRT);
    mat_mult(W, S, RS);
    double SR[9];
    sym_to_full(sr6, SR);
    const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double R[9];
    if (total) {
        double I1s = SR[0]+SR[4]+SR[8];
        double iso = I1s*Kb, third = 1.0/3.0*I1s;
        for (int k = 0; k < 9; ++k) R[k] = I[k]*iso + (SR[k] - I[k]*third)*(2.0*G) + SRT[k] + RS[k];
    } else {
        double tr3 = 1.0/3.0*(SR[0]+SR[4]+SR[8]);
        for (int k = 0; k < 9; ++k) R[k] = (SR[k] - I[k]*tr3)*(2.0*G) + SRT[k] + RS[k];
    }
    full_to_sym(R, R6);
}

__device__ inline double apex_cut6(double *S, double alpha, double kf)
{
    double I1 = S[0] + S[3] + S[5];
    if ((kf-alpha*I1) < 0.0) {
        double Ratio = (alpha == 0.0) ? 0.0 : kf/alpha;
        S[0] -= 1.0/3.0*(I1-Ratio);
        S[3] -= 1.0/3.0*(I1-Ratio);
        S[5] -= 1.0/3.0*(I1-Ratio);
        I1 = Ratio;
    }
    return I1;
}

// Drucker-Prager return mapping (Particle.cpp:341-430 / 661-752); St trial, Sn previous level.
__device__ inline void dp_return(const ClassParams &c, double i22, double *St, const double *Sn,
                                 const double *sr6, double fac, double *tau6)
{
    double alpha = (i22 == 0.0) ? c.dp_alpha2 : c.dp_alpha3;
    double kf = (i22 == 0.0) ? c.dp_kf2 : c.dp_kf3;
    const double I6[6] = {1, 0, 0, 1, 0, 1};
    double I1 = apex_cut6(St, alpha, kf);
    double S[6];
    for (int k = 0; k < 6; ++k) S[k] = St[k] - I6[k]*(1.0/3.0*I1);
    double J2 = 0.5*contract6(S);
    if ((sqrt(J2)+alpha*I1-kf) > 0.0) {
        double trn = Sn[0]+Sn[3]+Sn[5];
        for (int k = 0; k < 6; ++k) S[k] = Sn[k] - I6[k]*(1.0/3.0*trn);
        J2 = 0.5*contract6(S);
        if (sqrt(J2) > 0.0) {
            double I1s = sr6[0]+sr6[3]+sr6[5];
            // sum of the Hadamard product over all 9 entries (abab, Functions.cpp:350-357)
            double sum = S[0]*sr6[0]+S[1]*sr6[1]+S[2]*sr6[2]+S[1]*sr6[1]+S[3]*sr6[3]+S[4]*sr6[4]+S[2]*sr6[2]+S[4]*sr6[4]+S[5]*sr6[5];
            double dLanda, s1;
            double s2 = c.G/sqrt(J2);
            if (c.Fail == 2) {
                dLanda = 1.0/(9.0*alpha*alpha*c.K+c.G)*((3.0*alpha*c.K*I1s) + (c.G/sqrt(J2))*sum);
                s1 = 3.0*alpha*c.K;
            } else {
                dLanda = 1.0/(9.0*alpha*c.K*3.0*c.sin_psi+c.G)*((3.0*alpha*c.K*I1s) + (c.G/sqrt(J2))*sum);
                s1 = 3.0*3.0*c.sin_psi*c.K;
            }
            const double Iv[6] = {1, 0, 0, 1, 0, i22};
            for (int k = 0; k < 6; ++k) {
                double pl = Iv[k]*s1 + S[k]*s2;
                St[k] = St[k] - (pl*dLanda)*fac;
            }
        }
        I1 = apex_cut6(St, alpha, kf);
        for (int k = 0; k < 6; ++k) S[k] = St[k] - I6[k]*(1.0/3.0*I1);
        J2 = 0.5*contract6(S);
        if ((sqrt(J2)+alpha*I1-kf) > 0.0 && sqrt(J2) > 0.0) {
            double sc = (kf-alpha*I1)/sqrt(J2);
            for (int k = 0; k < 6; ++k) St[k] = I6[k]*(I1/3.0) + S[k]*sc;
        }
    }
    for (int k = 0; k < 6; ++k) tau6[k] = S[k];
}

// SOLID = false compiles the fluid-only update (no Material 2/3 particles in the run).
template <bool SOLID>
__global__ void __launch_bounds__(kBlock) k_integrate(Dev D0, Params P)
{
    const Dev D = bind(D0);
    Ctl *ctl = D.ctl;
    if (ctl->halt) return;
    if (ctl->resume < 2 && (ctl->time >= ctl->tout || P.hook_after)) return;   // stop before Move
    double deltatmin = bitsd(ctl->dtmin_bits);
    double dt = adaptive_dt(ctl->deltat, ctl->deltatint, deltatmin);
    if (dt < (ctl->deltatint/1.0e5)) return;                  // Fatal: reported by k_tail
    __shared__ ClassParams cp[kMaxClasses];
    for (int k = threadIdx.x; k < P.nclasses; k += kBlock) cp[k] = D.cls[k];
    __syncthreads();
    int i = blockIdx.x*kBlock + threadIdx.x;
    if (i >= P.n) return;
    uint32_t f = D.flags_b[i];
    double4 p = D.pos_b[i];
    const ClassParams &c = cp[flag_cls(f)];
    bool first = ctl->first_step != 0;
    int ct = first ? 30 : ctl->ct;
    if (flag_free(f)) {
        double4 v = D.vel_b[i];
        double4 vh = D.vh_b[i];
        double4 a = D.acc[i];
        if (P.inout && flag_inout(f) > 0) {
            // in/out-flow particles move with the imposed velocity (Domain.cpp:1305-1320)
            a.x = 0.0; a.y = 0.0; a.z = 0.0;
            if (flag_inout(f) == 1 || P.out_density > 0.0) a.w = 0.0;
            D.acc[i] = a;
        }
        double rho = p.w;
        double P_ = v.w;
        if (P.scheme == 1) {
            // Move_Leapfrog, Particle.cpp:561-596 (vh = va, Densitya)
            if (first) {
                vh.w = rho - dt/2.0*a.w;
                vh.x = v.x - dt/2.0*a.x; vh.y = v.y - dt/2.0*a.y; vh.z = v.z - dt/2.0*a.z;
            }
            double rb = vh.w;
            vh.w += dt*a.w;
            rho = (vh.w+rb)/2.0;
            double b0 = vh.x, b1 = vh.y, b2 = vh.z;
            vh.x += dt*a.x; vh.y += dt*a.y; vh.z += dt*a.z;
            v.x = (vh.x + b0)/2.0; v.y = (vh.y + b1)/2.0; v.z = (vh.z + b2)/2.0;
            p.x += dt*vh.x; p.y += dt*vh.y; p.z += dt*vh.z;
        } else {
            // Move_MVerlet, Particle.cpp:189-274 (vh = vb, Densityb)
            double4 xs = P.any_xsph ? D.xsp[i] : make_double4(0.0, 0.0, 0.0, 0.0);
            double s = 0.5*dt*dt;
            p.x += dt*(v.x+xs.x) + s*a.x;
            p.y += dt*(v.y+xs.y) + s*a.y;
            p.z += dt*(v.z+xs.z) + s*a.z;
            bool shep = c.Shepard && (ctl->shep_counter == c.ShepardStep);
            double zw = 0.0, sd = 0.0;
            if (shep) { zw = D.xsp[i].w; sd = D.shep[i].x; }
            if (ct == 30) {
                if (shep && zw > 0.6) { vh.w = sd/zw; rho = sd/zw; }
                else { vh.w = rho; rho += dt*a.w; }
                vh.x = v.x; vh.y = v.y; vh.z = v.z;
                v.x += dt*a.x; v.y += dt*a.y; v.z += dt*a.z;
            } else {
                if (shep && zw > 0.6) { vh.w = sd/zw; rho = sd/zw; }
                else { double dens = rho; rho = vh.w + 2.0*dt*a.w; vh.w = dens; }
                double t2 = 2*dt;
                double o0 = v.x, o1 = v.y, o2 = v.z;
                v.x = vh.x + t2*a.x; v.y = vh.y + t2*a.y; v.z = vh.z + t2*a.z;
                vh.x = o0; vh.y = o1; vh.z = o2;
            }
        }
        double2 mm = D.mmu_b[i];
        if (!SOLID || c.Material == 1) {
            // Mat1, Particle.cpp:146-186
            P_ = eos(c.PresEq, c.Cs, c.P0, rho, c.RefDensity);
            if (P.any_les_t0) {
                double sr[6];
                for (int k = 0; k < 6; ++k) sr[k] = D.sr[6*i+k];
                double temp = contract6(sr);
                double ShearRate = sqrt(0.5*temp);
                double SBar = sqrt(2.0*temp);
                D.shr[2*i] = ShearRate;
                D.shr[2*i+1] = SBar;
                if (c.LES) mm.y = c.MuRef + c.RefDensity*pow((c.CSmag*c.h),2.0)*SBar;
                if (c.T0 > 0.0) {
                    if (c.VisM == 0) {
                        if (ShearRate != 0.0) mm.y = c.MuRef + c.T0*(1-exp(-c.m*ShearRate))/ShearRate;
                        else mm.y = c.MuRef + c.T0*c.m;
                    } else {
                        mm.y = (1000.0*c.MuRef + c.MuRef*c.MuRef*1000.0/c.T0*ShearRate)/(1+1000.0*c.MuRef/c.T0*ShearRate);
                    }
                }
                D.mmu_b[i] = mm;
            }
        } else {
            double sr[6], rr[3], R[6];
            ld6(D.sr, i, sr);
            for (int k = 0; k < 3; ++k) rr[k] = D.rr[3*i+k];
            double sg[6], sh[6], tau[6], th[6], ep[6], eh[6];
            ld6(D.sig_b, i, sg); ld6(D.sigh_b, i, sh);
            ld6(D.tau_b, i, tau); ld6(D.tauh_b, i, th);
            ld6(D.eps_b, i, ep); ld6(D.epsh_b, i, eh);
            double i22 = (P.dim == 2) ? 0.0 : 1.0;
            const double I6[6] = {1, 0, 0, 1, 0, 1};
            if (c.Material == 2) {
                P_ = eos(c.PresEq, c.Cs, c.P0, rho, c.RefDensity);
                jaumann_rate(false, tau, sr, rr, c.G, c.K, R);
                if (P.scheme == 1) {
                    // Mat2Leapfrog, Particle.cpp:598-639 (th = ShearStressa)
                    if (first) for (int k = 0; k < 6; ++k) th[k] = R[k]*(-dt/2.0) + tau[k];
                    double tb[6];
                    for (int k = 0; k < 6; ++k) { tb[k] = th[k]; th[k] = R[k]*dt + th[k]; }
                    if (c.Fail == 1) {
                        double J2 = 0.5*contract6(th);
                        double am = c.Sigmay/sqrt(3.0*J2);
                        double sc = (1.0 < am) ? 1.0 : am;
                        for (int k = 0; k < 6; ++k) th[k] = th[k]*sc;
                    }
                    for (int k = 0; k < 6; ++k) tau[k] = (th[k]+tb[k])*(1.0/2.0);
                    for (int k = 0; k < 6; ++k) sg[k] = I6[k]*(-P_) + tau[k];
                    if (first) for (int k = 0; k < 6; ++k) eh[k] = sr[k]*(-dt/2.0) + ep[k];
                    double ebb[6];
                    for (int k = 0; k < 6; ++k) { ebb[k] = eh[k]; eh[k] = sr[k]*dt + eh[k]; }
                    for (int k = 0; k < 6; ++k) ep[k] = (eh[k]+ebb[k])*(1.0/2.0);
                } else {
                    // Mat2MVerlet, Particle.cpp:276-318 (th = ShearStressb)
                    double st[6];
                    for (int k = 0; k < 6; ++k) st[k] = tau[k];
                    if (ct == 30) for (int k = 0; k < 6; ++k) tau[k] = R[k]*dt + tau[k];
                    else for (int k = 0; k < 6; ++k) tau[k] = R[k]*(2.0*dt) + th[k];
                    for (int k = 0; k < 6; ++k) th[k] = st[k];
                    if (c.Fail == 1) {
                        double J2 = 0.5*contract6(tau);
                        double am = c.Sigmay/sqrt(3.0*J2);
                        double sc = (1.0 < am) ? 1.0 : am;
                        for (int k = 0; k < 6; ++k) tau[k] = tau[k]*sc;
                    }
                    for (int k = 0; k < 6; ++k) sg[k] = I6[k]*(-P_) + tau[k];
                    for (int k = 0; k < 6; ++k) st[k] = ep[k];
                    if (ct == 30) for (int k = 0; k < 6; ++k) ep[k] = sr[k]*dt + ep[k];
                    else for (int k = 0; k < 6; ++k) ep[k] = sr[k]*(2.0*dt) + eh[k];
                    for (int k = 0; k < 6; ++k) eh[k] = st[k];
                }
            } else {
                jaumann_rate(true, sg, sr, rr, c.G, c.K, R);
                if (P.scheme == 1) {
                    // Mat3Leapfrog, Particle.cpp:641-795 (sh = Sigmaa)
                    if (first) for (int k = 0; k < 6; ++k) sh[k] = R[k]*(-dt/2.0) + sg[k];
                    double sb[6];
                    for (int k = 0; k < 6; ++k) { sb[k] = sh[k]; sh[k] = R[k]*dt + sh[k]; }
                    if (c.Fail > 1) dp_return(c, i22, sh, sg, sr, dt, tau);
                    for (int k = 0; k < 6; ++k) sg[k] = (sh[k]+sb[k])*(1.0/2.0);
                    if (first) for (int k = 0; k < 6; ++k) eh[k] = sr[k]*(-dt/2.0) + ep[k];
                    double ebb[6];
                    for (int k = 0; k < 6; ++k) { ebb[k] = eh[k]; eh[k] = sr[k]*dt + eh[k]; }
                    for (int k = 0; k < 6; ++k) ep[k] = (eh[k]+ebb[k])*(1.0/2.0);
                } else {
                    // Mat3MVerlet, Particle.cpp:320-482 (sh = Sigmab)
                    double st[6];
                    for (int k = 0; k < 6; ++k) st[k] = sg[k];
                    if (ct == 30) for (int k = 0; k < 6; ++k) sg[k] = R[k]*dt + sg[k];
                    else for (int k = 0; k < 6; ++k) sg[k] = R[k]*(2.0*dt) + sh[k];
                    for (int k = 0; k < 6; ++k) sh[k] = st[k];
                    if (c.Fail > 1) dp_return(c, i22, sg, st, sr, (ct == 30) ? dt : 2.0*dt, tau);
                    for (int k = 0; k < 6; ++k) st[k] = ep[k];
                    if (ct == 30) for (int k = 0; k < 6; ++k) ep[k] = sr[k]*dt + ep[k];
                    else for (int k = 0; k < 6; ++k) ep[k] = sr[k]*(2.0*dt) + eh[k];
                    for (int k = 0; k < 6; ++k) eh[k] = st[k];
                }
            }
            st6(D.sig_b, i, sg); st6(D.sigh_b, i, sh);
            st6(D.tau_b, i, tau); st6(D.tauh_b, i, th);
            st6(D.eps_b, i, ep); st6(D.epsh_b, i, eh);
        }
        // periodic position update (Particle.cpp:127-142)
        if (P.any_periodic) {
            if (P.periodic[0]) { if (p.x > P.trpr[0]) p.x -= P.domsize[0]; if (p.x < P.blpf[0]) p.x += P.domsize[0]; }
            if (P.periodic[1]) { if (p.y > P.trpr[1]) p.y -= P.domsize[1]; if (p.y < P.blpf[1]) p.y += P.domsize[1]; }
            if (P.periodic[2]) { if (p.z > P.trpr[2]) p.z -= P.domsize[2]; if (p.z < P.blpf[2]) p.z += P.domsize[2]; }
        }
        p.w = rho;
        v.w = P_;
        if (!SOLID && P.views_next && P.dim == 2) v.z = 0.0;   // as k_reorder leaves a 2-D step's v
        D.pos_b[i] = p;
        D.vel_b[i] = v;
        D.vh_b[i] = vh;
        // the next step keeps this order (queued lean): its views, in place of its k_reorder
        if (!SOLID && P.views_next) free_views(D, P, c, i, f, p, v, mm);
    }
    if (P.verlet && flag_free(f)) {
        // list reuse: a particle more than half the skin away from where the lists were built
        // makes the next step rebuild
        double4 r = D.xref[i];
        double d0 = p.x - r.x, d1 = p.y - r.y, d2 = p.z - r.z;
        if (d0*d0 + d1*d1 + d2*d2 > P.skin_trig2) ctl->need_rebuild = 1;
    }
    // an inflow particle past InFlowLoc1 halts the step for the host's recycling
    // (InFlowBCLeave, Domain.cpp:1344-1357)
    if ((P.inout == 1 || P.inout == 3) && flag_free(f) && flag_inout(f) == 1 && p.x > P.loc1)
        atomicAdd(&ctl->leave, 1);
    // CheckParticleLeave test against [BLPF, TRPR]
    if ((p.x > P.trpr[0]) || (p.y > P.trpr[1]) || (p.z > P.trpr[2]) ||
        (p.x < P.blpf[0]) || (p.y < P.blpf[1]) || (p.z < P.blpf[2])) {
        atomicAdd(&ctl->leave, 1);
        return;
    }
    // next step's cell key (ListGenerate after Move); the arrival ranks of the next sort come
    // with it, or, with list reuse, from k_keys on the steps that rebuild
    int kx, ky, kz;
    int key = cell_key(P, p.x, p.y, p.z, kx, ky, kz);
    D.key[i] = key;
    if (!P.verlet) D.rank[i] = cell_rank(D, key, flag_free(f));
}

// Time update and stop conditions (Solve loop, Domain.cpp:1762-1793); one thread.
__global__ void k_tail(Dev D, Params P)
{
    Ctl *c = D.ctl;
    if (c->halt) return;
    c->nl_over += c->nl_q;                 // the force pass's overflow queue has been ended
    c->nl_q = 0;
    c->last_pairs = c->pairs_ff/2 + c->pairs_fw;
    double deltatmin = bitsd(c->dtmin_bits);
    c->deltatmin = deltatmin;
    if (c->xmat) { c->halt = kHaltError; c->error = -2; return; }
    if (c->resume < 2 && (c->time >= c->tout || P.hook_after)) {
        // output time (Domain.cpp:1773), or the GeneralAfter hook of every step (:1770): the
        // host takes over with the pair sums formed and Move pending
        c->halt = (c->time >= c->tout) ? kHaltOutput : kHaltHookAfter;
        c->mid_step = 1;
        return;
    }
    double dt = adaptive_dt(c->deltat, c->deltatint, deltatmin);
    if (dt < (c->deltatint/1.0e5)) { c->halt = kHaltError; c->error = -4; c->deltat = dt; return; }
    c->deltat = dt;
    c->time += dt;
    c->steps_done += 1;
    c->resume = 0;
    c->mid_step = 0;
    c->acc_init = 0;
    c->n_out = 0;
    // the step's working copy is the next step's state
    c->side ^= c->rebuild;
    if (P.verlet) {
        c->rebuild = (c->need_rebuild || c->list_over || P.verlet_force) ? 1 : 0;
        c->need_rebuild = 0;
    }
    if (P.scheme == 0) {
        // Move_MVerlet's per-particle counters, uniform over free particles (Particle.cpp:191-195, 272-273)
        if (c->first_step) c->ct = 30;
        c->ct = (c->ct == 30) ? 0 : c->ct + 1;
        c->shep_counter = (c->shep_counter == P.shepard_step) ? 0 : c->shep_counter + 1;
    }
    c->first_step = 0;
    c->dtmin_bits = dbits(c->deltatint);
    c->pairs_ff = 0;
    c->pairs_fw = 0;
    if (c->leave) { c->halt = kHaltLeave; return; }
    if (c->steps_done >= c->max_steps) { c->halt = kHaltSteps; return; }
    if (!(c->time < c->tf)) { c->halt = kHaltTend; return; }
    // the next step was queued lean but must sort and rebuild: the host queues it in full
    if (P.lean && c->rebuild) c->halt = kHaltSort;
}

// Head of a run call: stop at once if the loop condition fails (Domain.cpp:1762).
__global__ void k_head(Dev D, Params P)
{
    Ctl *c = D.ctl;
    c->halt = kRun;
    c->steps_done = 0;
    c->leave = 0;
    c->xmat = 0;
    if (c->max_steps <= 0) { c->halt = kHaltSteps; return; }
    if (c->mid_step) { c->resume = (c->mid_step == 2) ? 1 : 2; return; }
    if (!(c->time < c->tf)) { c->halt = kHaltTend; return; }
}

// The GeneralBefore hook of every step (Domain.cpp:1767): after StartAcceleration, the
// in/out-flow imposition and the neighbour search, the host takes over before the wall
// smoothing and the pair sums.  A step resumed past this stop does not launch it.
__global__ void k_hook_stop(Dev D, Params P)
{
    Ctl *c = D.ctl;
    if (c->halt || c->resume) return;
    c->halt = kHaltHookBefore;
    c->mid_step = 2;
}

// Accelerations and density rates from the host at a hook stop, in the host's particle order
// (the reference order of psph_download): before the wall pass they are the values the pair
// sums of this step start from (acc_init), after the force pass the values Move uses.  Fixed
// particles keep theirs (their pair sums feed nothing).
__global__ void __launch_bounds__(kBlock) k_hook_accel(Dev D0, Params P, const double *a, const double *dd)
{
    const Dev D = bind(D0);
    const int i = blockIdx.x*kBlock + threadIdx.x;
    const bool before = D.ctl->mid_step == 2;
    if (i == 0 && before) D.ctl->acc_init = 1;
    if (i >= P.n) return;
    if (!flag_free(D.flags_b[i])) return;
    const int o = D.orig_b[i];
    double4 v = D.acc[i];
    v.x = a[3*o]; v.y = a[3*o + 1]; v.z = a[3*o + 2];
    v.w = dd ? dd[o] : (before ? 0.0 : v.w);     // StartAcceleration zeroes dDensity
    D.acc[i] = v;
}

// The pair sums of fixed particles (read at a stop after the force pass: outputs and the
// GeneralAfter hook).  The reference adds every pair into both particles, so a fixed particle
// P2 of a pair with free P1 gets a -= mi temp, dDensity += mi (dj/di) temp1 on top of the
// StartAcceleration reset a = Gravity (Interaction.cpp:166-167, 339-340; Domain.cpp:1002).
// They feed nothing in the dynamics; the force pass does not form them.  Pairs of the fixed
// particle's own material, found by the reference's test (neighbouring cells, rij/h <=
// Cellfac), each evaluated by the free particle's pair function; written to acc of the fixed
// particle.  Pairs with free particles of another material (CalcForce12/13) are not added.
template <bool SOLID>
__global__ void __launch_bounds__(kBlock) k_fixed_sums(Dev D0, Params P)
{
    const Dev D = bind(D0);
    __shared__ ClassParams cp[kMaxClasses];
    for (int k = threadIdx.x; k < P.nclasses; k += kBlock) cp[k] = D.cls[k];
    __syncthreads();
    const int w = blockIdx.x*kBlock + threadIdx.x;
    if (w >= P.n) return;
    const uint32_t fw = D.flags_b[w];
    if (flag_free(fw)) return;
    const int mat = flag_mat(fw), cw = flag_cls(fw);
    const double4 pw = D.pos_b[w];
    int kx, ky, kz;
    decode_key(P, D.key[w], kx, ky, kz);
    double a0 = P.gravity[0], a1 = P.gravity[1], a2 = P.gravity[2], dd = 0.0;
    // rows of three cells around the wall's cell; with list reuse the sort (and so the cell
    // table) is that of the last build, so the rows span five cells and a candidate counts if
    // its current cell neighbours the wall's (keys_adjacent beyond the smallest cell size)
    const int R = P.verlet ? 2 : 1, W = 2*R + 1;
    const int nrows = (P.dim == 3) ? W*W : W;
    const int key_w = D.key[w];
    for (int rs = 0; rs < 2*nrows; ++rs) {
        int beg = 0, end = 0;
        if (P.any_periodic) {
            int b0, e0, b1, e1;
            row_range_periodic(D, P, kx, ky, kz, rs >> 1, b0, e0, b1, e1);
            beg = (rs & 1) ? b1 : b0;
            end = (rs & 1) ? e1 : e0;
        } else if (!(rs & 1)) {
            if (P.verlet) {
                const int r = rs >> 1;
                const int yy = ky + (r % W) - R, zz = (P.dim == 3) ? kz + (r / W) - R : kz;
                if (yy >= 0 && yy < P.nc[1] && zz >= 0 && zz < P.nc[2]) {
                    const int base = P.nc[0]*(yy + P.nc[1]*zz);
                    const int xa = kx > 1 ? kx - 2 : 0, xb = kx + 2 < P.nc[0] ? kx + 2 : P.nc[0] - 1;
                    beg = D.cstart[base + xa];
                    end = D.cstart[base + xb + 1];
                }
            } else {
                row_range(D, P, kx, ky, kz, rs >> 1, beg, end);
            }
        }
        for (int j = beg; j < end; ++j) {
            const uint32_t fj = D.flags_b[j];
            if (!flag_free(fj) || flag_mat(fj) != mat) continue;
            const double4 pj = D.pos_b[j];
            const double h = pair_h(cp, P, flag_cls(fj), cw);
            double x0 = pj.x - pw.x, x1 = pj.y - pw.y, x2 = pj.z - pw.z;
            if (P.any_periodic) periodic_xij(D, P, h, D.key[j], w, x0, x1, x2);
            const double r2 = x0*x0 + x1*x1 + x2*x2;
            if (!support_test(P, r2, h)) continue;
            if (P.verlet && r2 > P.adj2 && !keys_adjacent(P, key_w, D.key[j])) continue;
            const double4 vj = D.vel_b[j], aj = D.aux[j];
            Acc A;
            A.a0 = A.a1 = A.a2 = A.dd = A.ddc = 0.0;
            A.vx0 = A.vx1 = A.vx2 = 0.0; A.zw = 0.0; A.sd = 0.0; A.sw = 0.0;
            for (int k = 0; k < 6; ++k) A.sr[k] = 0.0;
            A.rr[0] = A.rr[1] = A.rr[2] = 0.0;
            double tw[7];
            if (SOLID && mat != 1) {
                double sgj[6], trj[6];
                ld6(D.sig_b, j, sgj);
                ld6(D.tir, j, trj);
                pair2233<double>(D, P, cp, j, w, fj, pj, vj, aj, sgj, trj, A, 0, tw);
            } else {
                pair11(D, P, cp, j, w, fj, pj, vj, aj, A, 0, tw);
            }
            a0 -= tw[4]*tw[0]; a1 -= tw[4]*tw[1]; a2 -= tw[4]*tw[2];
            dd += tw[4]*(tw[6]/tw[5])*tw[3];
        }
    }
    D.acc[w] = make_double4(a0, a1, a2, dd);
}

// ------------------------------------------------------------------------------------------
// test hook: the neighbour pair set with the reference's orientation (MainNeighbourSearch
// Domain.cpp:808-945 + the rij/h <= Cellfac test of CalcForce11/2233).  Pass 0 counts pairs
// per particle, pass 1 writes them at offsets from an exclusive scan of the counts.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_pairs(Dev D0, Params P, int pass, const long long *off,
                                                  int *cnt_out, long long *out)
{
    const Dev D = bind(D0);
    __shared__ ClassParams cp[kMaxClasses];
    for (int k = threadIdx.x; k < P.nclasses; k += kBlock) cp[k] = D.cls[k];
    __syncthreads();
    int i = blockIdx.x*kBlock + threadIdx.x;
    if (i >= P.n) return;
    uint32_t fi = D.flags_b[i];
    double4 pi = D.pos_b[i];
    int key_i = D.key[i], orig_i = D.orig_b[i];
    int kx, ky, kz;
    decode_key(P, key_i, kx, ky, kz);
    int nrows = (P.dim == 3) ? 9 : 3;
    long long w = (pass == 1) ? off[i] : 0;
    int c = 0;
    for (int rs = 0; rs < 2*nrows; ++rs) {
        int beg = 0, end = 0;
        if (P.any_periodic) {
            int b0, e0, b1, e1;
            row_range_periodic(D, P, kx, ky, kz, rs >> 1, b0, e0, b1, e1);
            beg = (rs & 1) ? b1 : b0;
            end = (rs & 1) ? e1 : e0;
        } else if (!(rs & 1)) {
            row_range(D, P, kx, ky, kz, rs >> 1, beg, end);
        }
        for (int j = beg; j < end; ++j) {
            if (j == i) continue;
            uint32_t fj = D.flags_b[j];
            if (!(flag_free(fi) || flag_free(fj))) continue;
            if (flag_mat(fi) != flag_mat(fj)) continue;
            int kj = D.key[j], oj = D.orig_b[j];
            int first = pair_first(P, key_i, kj, orig_i, oj);
            if (!first) continue;
            double4 pj = D.pos_b[j];
            double x0 = pi.x - pj.x, x1 = pi.y - pj.y, x2 = pi.z - pj.z;
            double h = pair_h(cp, P, flag_cls(fi), flag_cls(fj));
            if (P.any_periodic) periodic_xij(D, P, h, key_i, j, x0, x1, x2);
            double rij = sqrt(x0*x0 + x1*x1 + x2*x2);
            if ((rij/h) <= P.cellfac) {
                if (pass == 1) { out[2*w] = orig_i; out[2*w+1] = oj; ++w; }
                ++c;
            }
        }
    }
    if (pass == 0) cnt_out[i] = c;
}

// ------------------------------------------------------------------------------------------
// visualisation output (Domain::WriteXDMF, Domain.cpp:1893-1928): the fields the reference
// converts to float, rounded to nearest, in reference particle order.  out holds, for n
// particles: position 3n | velocity 3n | acceleration 3n | pressure n | density n | mass n |
// h n | tag n (int32) | sigma 6n | strain 6n.  acc_ok: the acceleration columns hold this
// state's force pass; tensors: the run keeps Sigma and Strain on the device.  The columns are
// the caller's choice of state buffer (no bind()).
// ------------------------------------------------------------------------------------------
constexpr int kOutFloats = 26;

__global__ void __launch_bounds__(kBlock) k_pack_output(Dev D, Params P, float *out, int acc_ok, int tensors)
{
    const int i = blockIdx.x*blockDim.x + threadIdx.x;
    if (i >= P.n) return;
    const size_t n = (size_t)P.n, o = (size_t)D.orig[i];
    const uint32_t f = D.flags[i];
    const ClassParams &cp = D.cls[flag_cls(f)];
    const double4 x = D.pos[i], v = D.vel[i];
    float *po = out + 3*o, *vo = out + 3*n + 3*o;
    po[0] = (float)x.x; po[1] = (float)x.y; po[2] = (float)x.z;
    vo[0] = (float)v.x; vo[1] = (float)v.y; vo[2] = (float)v.z;
    if (acc_ok) {
        float *ao = out + 6*n + 3*o;
        if (flag_free(f) || acc_ok == 2) {
            // acc_ok == 2: the fixed particles' pair sums are formed (k_fixed_sums)
            const double4 a = D.acc[i];
            ao[0] = (float)a.x; ao[1] = (float)a.y; ao[2] = (float)a.z;
        } else {
            // a fixed particle's acceleration is the StartAcceleration reset (Domain.cpp:1002)
            ao[0] = (float)P.gravity[0]; ao[1] = (float)P.gravity[1]; ao[2] = (float)P.gravity[2];
        }
    }
    out[9*n + o] = (float)v.w;
    out[10*n + o] = (float)x.w;
    out[11*n + o] = (float)D.mmu[i].x;
    out[12*n + o] = (float)cp.h;
    reinterpret_cast<int32_t *>(out)[13*n + o] = cp.ID;
    if (tensors) {
        float *so = out + 14*n + 6*o, *eo = out + 20*n + 6*o;
        for (int c = 0; c < 6; ++c) {
            so[c] = (float)D.sig[6*(size_t)i + c];
            eo[c] = (float)D.eps[6*(size_t)i + c];
        }
    }
}

// ------------------------------------------------------------------------------------------
// host-side launchers (one per kernel so the host can bracket each with timing events)
// ------------------------------------------------------------------------------------------
static inline int nblk(int n, int b) { return (n + b - 1)/b; }

int scan_blocks(const Params &P) { return nblk(P.ncells, kScanTile); }

void launch_head(const Dev &D, const Params &P, hipStream_t s)
{ hipLaunchKernelGGL(k_head, dim3(1), dim3(1), 0, s, D, P); }

void launch_keys(const Dev &D, const Params &P, hipStream_t s)
{ hipLaunchKernelGGL(k_keys, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P); }

void launch_scan(const Dev &D, const Params &P, hipStream_t s)
{
    int nb = scan_blocks(P);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, s, D, P);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, D, P, nb);
    hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(256), 0, s, D, P);
}

void launch_scatter(const Dev &D, const Params &P, hipStream_t s)
{ hipLaunchKernelGGL(k_scatter, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P); }

void launch_cellsort(const Dev &D, const Params &P, hipStream_t s)
{ hipLaunchKernelGGL(k_cellsort, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P); }

void launch_reorder(const Dev &D, const Params &P, hipStream_t s)
{ hipLaunchKernelGGL(k_reorder, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P); }

static void launch_wall_pass(const Dev &D, const Params &P, hipStream_t s);

// The wall pass, then with several materials the FSI smoothing and the saturation of soil,
// in the order of PrimaryComputeAcceleration (Domain.cpp:1020-1234).
void launch_wall(const Dev &D, const Params &P, hipStream_t s, int after_hook)
{
    if (!after_hook) {
        if (P.inout) {
            // InFlowBCFresh runs before the neighbour search (Domain.cpp:1763-1766)
            hipLaunchKernelGGL(k_io_mark, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P);
            hipLaunchKernelGGL(k_io_impose, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P);
        }
        if (P.hook_before) hipLaunchKernelGGL(k_hook_stop, dim3(1), dim3(1), 0, s, D, P);
    }
    launch_wall_pass(D, P, s);
    if (P.multi) {
        hipLaunchKernelGGL(k_fsi, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P);
        if (P.any_soil_fluid && P.swi_type != 2)
            hipLaunchKernelGGL(k_sat, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P);
    }
}

static void launch_wall_pass(const Dev &D, const Params &P, hipStream_t s)
{
    if (P.nfixed <= 0) return;
    if (P.fast11 || P.mixed32) {
        const dim3 g(nblk(4*P.nfixed, kBlock)), b(kBlock);
        if (P.dim == 3) {
            if (P.verlet) hipLaunchKernelGGL((k_wall_fast<3, 2>), g, b, 0, s, D, P);
            else hipLaunchKernelGGL((k_wall_fast<3, 1>), g, b, 0, s, D, P);
        } else {
            if (P.verlet) hipLaunchKernelGGL((k_wall_fast<2, 2>), g, b, 0, s, D, P);
            else hipLaunchKernelGGL((k_wall_fast<2, 1>), g, b, 0, s, D, P);
        }
        return;
    }
    const dim3 g(nblk(P.nfixed, kBlock)), b(kBlock);
    if (P.any_solid) {
        if (P.verlet) hipLaunchKernelGGL((k_wall<true, 2>), g, b, 0, s, D, P);
        else hipLaunchKernelGGL((k_wall<true, 1>), g, b, 0, s, D, P);
    } else {
        if (P.verlet) hipLaunchKernelGGL((k_wall<false, 2>), g, b, 0, s, D, P);
        else hipLaunchKernelGGL((k_wall<false, 1>), g, b, 0, s, D, P);
    }
}

static void launch_force_pass(const Dev &D, const Params &P, hipStream_t s);

void launch_force(const Dev &D, const Params &P, hipStream_t s)
{
    if (P.nfree <= 0) return;
    launch_force_pass(D, P, s);
    if (P.multi)
        hipLaunchKernelGGL(k_cross, dim3(nblk(P.nfree, kForceBlock)), dim3(kForceBlock), 0, s, D, P);
}

// The list passes of the fast fluid path: build (every step, or with list reuse on the steps
// that rebuild), the gather and its overflow end.
template <int DIM, int REACH>
static void launch_list_pass(const Dev &D, const Params &P, hipStream_t s)
{
    const dim3 gl(nblk(P.nfree, kBlock)), gf(nblk(P.nfree, kForceBlock)), go(kOverBlocks);
    // the list builds are no-ops on steps that keep the lists; a lean step does not queue them,
    // nor the overflow passes: a list overflows only at a build, and a build that overflowed
    // makes every following step rebuild (Ctl::list_over) until the host has grown the lists
    if (!P.lean) hipLaunchKernelGGL((k_nlist<DIM, REACH>), gl, dim3(kBlock), 0, s, D, P);
    if (REACH == 2 && P.tiles) {
        const dim3 gt(P.tile_max);
        if (!P.lean) {
            hipLaunchKernelGGL(k_tile_table, dim3(1), dim3(1024), 0, s, D, P);
            hipLaunchKernelGGL((k_tile_union<REACH>), dim3(std::min(P.tile_max, kUnionBlocks)), dim3(512), 0, s, D, P);
        }
        if (P.mixed32) {
            if (P.ab_one) hipLaunchKernelGGL((k_force32_tile<DIM, 2>), gt, dim3(kTileTargets*kTileLanes), 0, s, D, P);
            else if (P.ab_two) hipLaunchKernelGGL((k_force32_tile<DIM, 1>), gt, dim3(kTileTargets*kTileLanes), 0, s, D, P);
            else hipLaunchKernelGGL((k_force32_tile<DIM, 0>), gt, dim3(kTileTargets*kTileLanes), 0, s, D, P);
            if (!P.lean) hipLaunchKernelGGL((k_force11_f32_nl_over<DIM, REACH>), go, dim3(kForceBlock), 0, s, D, P);
        } else {
            if (P.ab_one) hipLaunchKernelGGL((k_force11_tile<DIM, 2>), gt, dim3(kTileTargets*kTileLanes), 0, s, D, P);
            else if (P.ab_two) hipLaunchKernelGGL((k_force11_tile<DIM, 1>), gt, dim3(kTileTargets*kTileLanes), 0, s, D, P);
            else hipLaunchKernelGGL((k_force11_tile<DIM, 0>), gt, dim3(kTileTargets*kTileLanes), 0, s, D, P);
            if (!P.lean) hipLaunchKernelGGL((k_force11_nl_over<DIM, REACH>), go, dim3(kForceBlock), 0, s, D, P);
        }
        return;
    }
    if (P.mixed32) {
        hipLaunchKernelGGL((k_force11_f32_nl<DIM, REACH>), gf, dim3(kForceBlock), 0, s, D, P);
        if (!P.lean) hipLaunchKernelGGL((k_force11_f32_nl_over<DIM, REACH>), go, dim3(kForceBlock), 0, s, D, P);
    } else {
        hipLaunchKernelGGL((k_force11_nl<DIM, REACH>), gf, dim3(kForceBlock), 0, s, D, P);
        if (!P.lean) hipLaunchKernelGGL((k_force11_nl_over<DIM, REACH>), go, dim3(kForceBlock), 0, s, D, P);
    }
}

static void launch_force_pass(const Dev &D, const Params &P, hipStream_t s)
{
    if (P.glist) {
        // with list reuse the lists are built only on steps that rebuild (a lean step does not
        // queue the build)
        if (!P.verlet) hipLaunchKernelGGL(k_glist<1>, dim3(nblk(P.nfree, kBlock)), dim3(kBlock), 0, s, D, P);
        else if (!P.lean) hipLaunchKernelGGL(k_glist<2>, dim3(nblk(P.nfree, kBlock)), dim3(kBlock), 0, s, D, P);
    }
    if (P.nl_cap > 0 && !P.glist) {
        if (P.dim == 3) {
            if (P.verlet) launch_list_pass<3, 2>(D, P, s);
            else launch_list_pass<3, 1>(D, P, s);
        } else {
            if (P.verlet) launch_list_pass<2, 2>(D, P, s);
            else launch_list_pass<2, 1>(D, P, s);
        }
        return;
    }
    if (P.any_solid && P.soil32)
        hipLaunchKernelGGL((k_force<true, float>), dim3(nblk(P.nfree, kForceBlock)), dim3(kForceBlock), 0, s, D, P);
    else if (P.any_solid && P.multi)
        hipLaunchKernelGGL((k_force<true, double, true>), dim3(nblk(P.nfree, kForceBlock)), dim3(kForceBlock), 0, s, D, P);
    else if (P.any_solid)
        hipLaunchKernelGGL((k_force<true, double, false>), dim3(nblk(P.nfree, kForceBlock)), dim3(kForceBlock), 0, s, D, P);
    else
        hipLaunchKernelGGL((k_force<false, double>), dim3(nblk(P.nfree, kForceBlock)), dim3(kForceBlock), 0, s, D, P);
}

void launch_integrate(const Dev &D, const Params &P, hipStream_t s)
{
    if (P.any_solid)
        hipLaunchKernelGGL(k_integrate<true>, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P);
    else
        hipLaunchKernelGGL(k_integrate<false>, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P);
}

void launch_tail(const Dev &D, const Params &P, hipStream_t s)
{ hipLaunchKernelGGL(k_tail, dim3(1), dim3(1), 0, s, D, P); }

void launch_fixed_sums(const Dev &D, const Params &P, hipStream_t s)
{
    if (P.any_solid) hipLaunchKernelGGL(k_fixed_sums<true>, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P);
    else hipLaunchKernelGGL(k_fixed_sums<false>, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P);
}

void launch_hook_accel(const Dev &D, const Params &P, const double *a, const double *dd, hipStream_t s)
{ hipLaunchKernelGGL(k_hook_accel, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P, a, dd); }

void launch_pairs(const Dev &D, const Params &P, hipStream_t s, int pass, const long long *off,
                  int *cnt, long long *out)
{ hipLaunchKernelGGL(k_pairs, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P, pass, off, cnt, out); }

int output_floats() { return kOutFloats; }

void launch_pack_output(const Dev &D, const Params &P, float *out, int acc_ok, int tensors, hipStream_t s)
{ hipLaunchKernelGGL(k_pack_output, dim3(nblk(P.n, kBlock)), dim3(kBlock), 0, s, D, P, out, acc_ok, tensors); }

}  // namespace psph
