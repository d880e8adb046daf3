// Host side of the learned surrogate (include/psph_gnn.h): parameters, buffers, graph
// construction, and the forward / backward / Adam sequences of kernel launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/persiansph_amd.h"
#include "../../include/psph_gnn.h"
#include "gnn_dev.h"

int psph_internal_fail(int code, const char *msg);

using namespace psgnn;

#define HC(expr)                                                                              \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return psph_internal_fail(PSPH_ERR_HIP, (std::string(#expr) + ": " + hipGetErrorString(e_)).c_str()); \
    } while (0)

namespace {

struct Lin {
    int kin;           // input width the layer multiplies (padded to a multiple of 32)
    size_t oW, ob;     // offsets of W [kin][128] and b [128] in the FP32 parameters
    size_t oWb, oWtb;  // offsets of the bf16 copies W [kin][128] and Wt [128][kin]
};

// splitmix64
struct Rng {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30))*0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27))*0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11)*(1.0/9007199254740992.0); }
};

uint16_t host_bf16(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

}  // namespace

struct psph_gnn {
    psph_gnn_config cfg{};
    hipStream_t s = nullptr;
    int dim = 2, L = 0, dout = 5;
    bool train = false;
    std::vector<Lin> lin;
    size_t nparam = 0, nwbf = 0;
    float *P = nullptr, *G = nullptr, *Mo = nullptr, *Vo = nullptr;
    uint16_t *Wbf = nullptr;
    long long step = 0;
    FeatScale fs{};
    // loaded problem
    int64_t n = 0, ne = 0;
    bool have_state = false, have_targets = false;
    double wsum = 0.0;
    int64_t ncap = -1, ecap = -1, cellcap = 0, dzcap = 0;
    std::vector<void *> nbuf, ebuf, cbuf;   // allocations sized by n, by ne, by cells
    double *x = nullptr, *v = nullptr, *rho = nullptr;
    int32_t *isfree = nullptr;
    int32_t *rp = nullptr, *deg = nullptr, *gkey = nullptr, *gsorted = nullptr;
    int32_t *cols = nullptr, *recv = nullptr, *send = nullptr, *rev = nullptr;
    int32_t *gcnt = nullptr, *gstart = nullptr, *gfill = nullptr;
    uint16_t *Xn = nullptr, *Ue = nullptr, *Zn0 = nullptr, *Ze0 = nullptr, *Zd = nullptr;
    std::vector<uint16_t *> H, E, ZE, AG, ZN;
    float *Y = nullptr, *tgt = nullptr, *wgt = nullptr, *loss = nullptr;
    uint16_t *dy = nullptr, *dZ = nullptr;
    float *dh[2] = {nullptr, nullptr}, *de[2] = {nullptr, nullptr};
    uint16_t *dhb[2] = {nullptr, nullptr}, *deb[2] = {nullptr, nullptr};
    float *dInN = nullptr, *dInE = nullptr, *dEt = nullptr;
    uint16_t *dEtb = nullptr;

    const Lin &enc_n1() const { return lin[0]; }
    const Lin &enc_n2() const { return lin[1]; }
    const Lin &enc_e1() const { return lin[2]; }
    const Lin &enc_e2() const { return lin[3]; }
    const Lin &pe1(int l) const { return lin[4 + 4*l]; }
    const Lin &pe2(int l) const { return lin[5 + 4*l]; }
    const Lin &pn1(int l) const { return lin[6 + 4*l]; }
    const Lin &pn2(int l) const { return lin[7 + 4*l]; }
    const Lin &dec1() const { return lin[4 + 4*L]; }
    const Lin &dec2() const { return lin[5 + 4*L]; }
    const uint16_t *Wt(const Lin &l) const { return Wbf + l.oWtb; }
    const uint16_t *W(const Lin &l) const { return Wbf + l.oWb; }
    const float *b(const Lin &l) const { return P + l.ob; }
    float *dW(const Lin &l) const { return G + l.oW; }
    float *db(const Lin &l) const { return G + l.ob; }
    int slot(int l) const { return train ? l : (l & 1); }
};

static void free_list(std::vector<void *> &v)
{
    for (void *p : v) (void)hipFree(p);
    v.clear();
}

template <class T>
static int dalloc(std::vector<void *> &list, T **p, size_t count)
{
    size_t bytes = std::max<size_t>(count*sizeof(T), 16);
    HC(hipMalloc((void **)p, bytes));
    list.push_back(*p);
    return 0;
}

#define AL(list, ptr, count) do { int rc_ = dalloc(list, &(ptr), (count)); if (rc_) return rc_; } while (0)

static int refresh_bf16(psph_gnn *g)
{
    for (const Lin &l : g->lin) weights_bf16(g->s, l.kin, l.kin, g->P + l.oW, g->Wbf + l.oWb, g->Wbf + l.oWtb);
    HC(hipGetLastError());
    return 0;
}

// node-sized buffers (capacity grows, never shrinks)
static int ensure_nodes(psph_gnn *g, int64_t n)
{
    if (n <= g->ncap) return 0;
    free_list(g->nbuf);
    int64_t c = std::max<int64_t>(n + n/8, 64);
    g->ncap = c;
    int L = g->L, nh = g->train ? L + 1 : 2, nl = g->train ? L : 1;
    auto &B = g->nbuf;
    AL(B, g->x, 3*c); AL(B, g->v, 3*c); AL(B, g->rho, c); AL(B, g->isfree, c);
    AL(B, g->rp, c + 1); AL(B, g->deg, c + 1); AL(B, g->gkey, c); AL(B, g->gsorted, c);
    AL(B, g->Xn, 32*c); AL(B, g->Zn0, 128*c); AL(B, g->Zd, 128*c); AL(B, g->Y, (size_t)g->dout*c);
    g->H.assign(nh, nullptr);
    for (auto &p : g->H) AL(B, p, 128*c);
    g->AG.assign(nl, nullptr);
    for (auto &p : g->AG) AL(B, p, 128*c);
    g->ZN.assign(nl, nullptr);
    for (auto &p : g->ZN) AL(B, p, 128*c);
    if (g->train) {
        AL(B, g->tgt, (size_t)g->dout*c); AL(B, g->wgt, c); AL(B, g->loss, 4);
        AL(B, g->dy, 128*c);
        for (int k = 0; k < 2; ++k) { AL(B, g->dh[k], 128*c); AL(B, g->dhb[k], 128*c); }
        AL(B, g->dInN, 256*c);
    }
    return 0;
}

// edge-sized buffers
static int ensure_edges(psph_gnn *g, int64_t ne)
{
    if (ne <= g->ecap) return 0;
    free_list(g->ebuf);
    int64_t c = std::max<int64_t>(ne + ne/8, 64);
    g->ecap = c;
    int L = g->L, nh = g->train ? L + 1 : 2, nl = g->train ? L : 1;
    auto &B = g->ebuf;
    AL(B, g->cols, c); AL(B, g->recv, c); AL(B, g->send, c); AL(B, g->rev, c);
    AL(B, g->Ue, 32*c); AL(B, g->Ze0, 128*c);
    g->E.assign(nh, nullptr);
    for (auto &p : g->E) AL(B, p, 128*c);
    g->ZE.assign(nl, nullptr);
    for (auto &p : g->ZE) AL(B, p, 128*c);
    if (g->train) {
        for (int k = 0; k < 2; ++k) { AL(B, g->de[k], 128*c); AL(B, g->deb[k], 128*c); }
        AL(B, g->dInE, 384*c); AL(B, g->dEt, 128*c); AL(B, g->dEtb, 128*c);
    }
    return 0;
}

int psph_gnn_create(const psph_gnn_config *cfg, psph_gnn **out)
{
    if (!cfg || !out) return psph_internal_fail(PSPH_ERR_INVALID, "null argument");
    if (cfg->dim != 2 && cfg->dim != 3) return psph_internal_fail(PSPH_ERR_INVALID, "dim must be 2 or 3");
    if (cfg->hidden != 128) return psph_internal_fail(PSPH_ERR_INVALID, "hidden width must be 128");
    if (cfg->layers < 1 || cfg->layers > 64) return psph_internal_fail(PSPH_ERR_INVALID, "layers must be 1..64");
    if (!(cfg->radius > 0.0)) return psph_internal_fail(PSPH_ERR_INVALID, "radius must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return psph_internal_fail(PSPH_ERR_NO_DEVICE, "no GPU visible");
    psph_gnn *g = new psph_gnn();
    g->cfg = *cfg;
    g->dim = cfg->dim;
    g->L = cfg->layers;
    g->dout = 2*cfg->dim + 1;
    g->train = cfg->train != 0;
    g->fs.inv_v = 1.0/cfg->vel_scale;
    g->fs.rho0 = cfg->rho0;
    g->fs.inv_rho = 1.0/cfg->rho_scale;
    g->fs.out_dx = cfg->out_dx;
    g->fs.out_dv = cfg->out_dv;
    g->fs.out_drho = cfg->out_drho;
    g->fs.drift_t = cfg->drift_t;
    g->fs.inv_deg = cfg->deg_scale > 0.0 ? 1.0/cfg->deg_scale : 1.0;
    auto add = [&](int kin) {
        Lin l;
        l.kin = kin;
        l.oW = g->nparam; g->nparam += (size_t)kin*128;
        l.ob = g->nparam; g->nparam += 128;
        l.oWb = g->nwbf; g->nwbf += (size_t)kin*128;
        l.oWtb = g->nwbf; g->nwbf += (size_t)kin*128;
        g->lin.push_back(l);
    };
    add(32); add(128); add(32); add(128);
    for (int l = 0; l < g->L; ++l) { add(384); add(128); add(256); add(128); }
    add(128); add(128);
    // Glorot-uniform weights over the layer's real fan-in, zero biases; the output layers of
    // the residual updates start at 0.3 and the decoder's output layer at 0.1 of that scale,
    // which keeps the latents' magnitude flat through the layers; padded inputs get zeros
    std::vector<float> p(g->nparam, 0.0f);
    Rng rng{cfg->seed ? cfg->seed : 1};
    for (size_t k = 0; k < g->lin.size(); ++k) {
        const Lin &l = g->lin[k];
        int fan_in = l.kin;
        if (k == 0) fan_in = cfg->dim + 3;
        if (k == 2) fan_in = cfg->dim + 1;
        double a = std::sqrt(6.0/(fan_in + 128.0));
        if (k >= 4 && k < 4 + 4*(size_t)g->L && (k - 4) % 2 == 1) a *= 0.3;
        if (k + 1 == g->lin.size()) a *= 0.1;
        for (int i = 0; i < fan_in; ++i)
            for (int j = 0; j < 128; ++j) {
                if (k + 1 == g->lin.size() && j >= g->dout) continue;
                p[l.oW + (size_t)i*128 + j] = (float)((2.0*rng.uniform() - 1.0)*a);
            }
    }
    auto bail = [&](int rc) { psph_gnn_destroy(g); return rc; };
    if (hipStreamCreate(&g->s) != hipSuccess) return bail(psph_internal_fail(PSPH_ERR_HIP, "stream"));
    if (hipMalloc((void **)&g->P, g->nparam*4) != hipSuccess || hipMalloc((void **)&g->G, g->nparam*4) != hipSuccess ||
        hipMalloc((void **)&g->Mo, g->nparam*4) != hipSuccess || hipMalloc((void **)&g->Vo, g->nparam*4) != hipSuccess ||
        hipMalloc((void **)&g->Wbf, g->nwbf*2) != hipSuccess)
        return bail(psph_internal_fail(PSPH_ERR_HIP, "parameter allocation"));
    if (hipMemcpy(g->P, p.data(), g->nparam*4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(g->G, 0, g->nparam*4) != hipSuccess || hipMemset(g->Mo, 0, g->nparam*4) != hipSuccess ||
        hipMemset(g->Vo, 0, g->nparam*4) != hipSuccess)
        return bail(psph_internal_fail(PSPH_ERR_HIP, "parameter upload"));
    int rc = refresh_bf16(g);
    if (rc) return bail(rc);
    if (hipStreamSynchronize(g->s) != hipSuccess) return bail(psph_internal_fail(PSPH_ERR_HIP, "sync"));
    *out = g;
    return 0;
}

int psph_gnn_destroy(psph_gnn *g)
{
    if (!g) return 0;
    if (g->s) (void)hipStreamSynchronize(g->s);
    free_list(g->nbuf);
    free_list(g->ebuf);
    free_list(g->cbuf);
    if (g->dZ) (void)hipFree(g->dZ);
    for (void *q : {(void *)g->P, (void *)g->G, (void *)g->Mo, (void *)g->Vo, (void *)g->Wbf})
        if (q) (void)hipFree(q);
    if (g->s) (void)hipStreamDestroy(g->s);
    delete g;
    return 0;
}

int64_t psph_gnn_num_params(const psph_gnn *g) { return g ? (int64_t)g->nparam : 0; }

int psph_gnn_get_params(psph_gnn *g, float *p, int64_t n)
{
    if (!g || !p || n < (int64_t)g->nparam) return psph_internal_fail(PSPH_ERR_CAPACITY, "need room for all parameters");
    HC(hipStreamSynchronize(g->s));
    HC(hipMemcpy(p, g->P, g->nparam*4, hipMemcpyDeviceToHost));
    return 0;
}

int psph_gnn_set_params(psph_gnn *g, const float *p, int64_t n)
{
    if (!g || !p || n != (int64_t)g->nparam) return psph_internal_fail(PSPH_ERR_INVALID, "parameter count mismatch");
    HC(hipStreamSynchronize(g->s));
    HC(hipMemcpy(g->P, p, g->nparam*4, hipMemcpyHostToDevice));
    HC(hipMemset(g->Mo, 0, g->nparam*4));
    HC(hipMemset(g->Vo, 0, g->nparam*4));
    g->step = 0;
    int rc = refresh_bf16(g);
    if (rc) return rc;
    HC(hipStreamSynchronize(g->s));
    return 0;
}

int psph_gnn_get_grads(psph_gnn *g, float *gr, int64_t n)
{
    if (!g || !gr || n < (int64_t)g->nparam) return psph_internal_fail(PSPH_ERR_CAPACITY, "need room for all gradients");
    HC(hipStreamSynchronize(g->s));
    HC(hipMemcpy(gr, g->G, g->nparam*4, hipMemcpyDeviceToHost));
    return 0;
}

// dZ holds a bf16 [rows][128] gradient of node or edge rows
static int ensure_dz_rows(psph_gnn *g)
{
    if (!g->train) return 0;
    int64_t rows = std::max(g->ncap, g->ecap);
    if (rows <= g->dzcap) return 0;
    if (g->dZ) { (void)hipFree(g->dZ); g->dZ = nullptr; }
    HC(hipMalloc((void **)&g->dZ, std::max<size_t>((size_t)rows*128*2, 16)));
    g->dzcap = rows;
    return 0;
}

// builds the radius graph of the loaded positions (CSR by receiver, ascending senders)
static int build_graph(psph_gnn *g, const double *xh)
{
    const int64_t n = g->n;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], xh[3*i + d]); hi[d] = std::max(hi[d], xh[3*i + d]); }
    GraphGrid gg{};
    gg.radius = g->cfg.radius;
    gg.inv_cell = 1.0/g->cfg.radius;
    gg.dim = g->dim;
    int64_t ncell = 1;
    for (int d = 0; d < 3; ++d) {
        gg.lo[d] = lo[d];
        int c = (d < g->dim && n > 0) ? (int)std::floor((hi[d] - lo[d])*gg.inv_cell) + 1 : 1;
        gg.nc[d] = std::max(c, 1);
        ncell *= gg.nc[d];
    }
    if (ncell > ((int64_t)1 << 28)) return psph_internal_fail(PSPH_ERR_CAPACITY, "graph grid too large");
    if (ncell + 1 > g->cellcap) {
        free_list(g->cbuf);
        g->cellcap = ncell + 1;
        AL(g->cbuf, g->gcnt, ncell + 1); AL(g->cbuf, g->gstart, ncell + 1); AL(g->cbuf, g->gfill, ncell + 1);
    }
    HC(hipMemsetAsync(g->gcnt, 0, sizeof(int32_t)*(ncell + 1), g->s));
    HC(hipMemsetAsync(g->gfill, 0, sizeof(int32_t)*(ncell + 1), g->s));
    graph_bin(g->s, (int)n, gg, g->x, g->gkey, g->gcnt);
    scan1(g->s, (int)ncell, g->gcnt, g->gstart);
    graph_scatter(g->s, (int)n, g->gkey, g->gstart, g->gfill, g->gsorted);
    graph_cellorder(g->s, (int)ncell, g->gstart, g->gsorted);
    graph_edges(g->s, 0, (int)n, gg, g->x, g->gstart, g->gsorted, g->deg, g->rp, nullptr);
    scan1(g->s, (int)n, g->deg, g->rp);
    int32_t ne32 = 0;
    HC(hipMemcpyAsync(&ne32, g->rp + n, 4, hipMemcpyDeviceToHost, g->s));
    HC(hipStreamSynchronize(g->s));
    g->ne = ne32;
    int rc = ensure_edges(g, g->ne);
    if (rc) return rc;
    rc = ensure_dz_rows(g);
    if (rc) return rc;
    graph_edges(g->s, 1, (int)n, gg, g->x, g->gstart, g->gsorted, g->deg, g->rp, g->cols);
    graph_finish(g->s, (int)n, g->rp, g->cols, g->recv, g->send, g->rev);
    HC(hipGetLastError());
    return 0;
}

int psph_gnn_load_state(psph_gnn *g, const double *x, const double *v, const double *rho, const int32_t *isfree,
                        int64_t n)
{
    if (!g || !x || !v || !rho || !isfree || n < 1) return psph_internal_fail(PSPH_ERR_INVALID, "bad state");
    if (n >= ((int64_t)1 << 31)) return psph_internal_fail(PSPH_ERR_CAPACITY, "too many particles");
    int rc = ensure_nodes(g, n);
    if (rc) return rc;
    g->n = n;
    // blocking copies from the caller's (pageable) arrays, after the stream's earlier work
    HC(hipStreamSynchronize(g->s));
    HC(hipMemcpy(g->x, x, sizeof(double)*3*n, hipMemcpyHostToDevice));
    HC(hipMemcpy(g->v, v, sizeof(double)*3*n, hipMemcpyHostToDevice));
    HC(hipMemcpy(g->rho, rho, sizeof(double)*n, hipMemcpyHostToDevice));
    HC(hipMemcpy(g->isfree, isfree, sizeof(int32_t)*n, hipMemcpyHostToDevice));
    g->wsum = 0.0;
    for (int64_t i = 0; i < n; ++i) g->wsum += isfree[i] ? 1.0 : 0.0;
    rc = build_graph(g, x);
    if (rc) return rc;
    node_in(g->s, (int)n, g->dim, g->v, g->rho, g->isfree, g->rp, g->fs, g->Xn);
    if (g->ne > 0) edge_in(g->s, (int)g->ne, g->dim, g->recv, g->send, g->x, 1.0/g->cfg.radius, g->Ue);
    HC(hipGetLastError());
    g->have_state = true;
    g->have_targets = false;
    return 0;
}

int psph_gnn_load_inputs(psph_gnn *g, const float *node_in_h, const float *edge_in_h, const int32_t *rowptr,
                         const int32_t *cols_h, int64_t n, int64_t ne)
{
    if (!g || !node_in_h || !rowptr || n < 1 || ne < 0 || (ne > 0 && (!edge_in_h || !cols_h)))
        return psph_internal_fail(PSPH_ERR_INVALID, "bad inputs");
    if (rowptr[0] != 0 || rowptr[n] != ne) return psph_internal_fail(PSPH_ERR_INVALID, "rowptr inconsistent");
    for (int64_t i = 0; i < n; ++i) {
        if (rowptr[i + 1] < rowptr[i]) return psph_internal_fail(PSPH_ERR_INVALID, "rowptr decreasing");
        for (int32_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            int32_t j = cols_h[e];
            if (j < 0 || j >= n || j == i) return psph_internal_fail(PSPH_ERR_INVALID, "column out of range");
            if (e > rowptr[i] && cols_h[e - 1] >= j) return psph_internal_fail(PSPH_ERR_INVALID, "row not ascending");
            // the reverse edge (j <- i) must exist: gradients reach senders through it
            if (!std::binary_search(cols_h + rowptr[j], cols_h + rowptr[j + 1], (int32_t)i))
                return psph_internal_fail(PSPH_ERR_INVALID, "graph not symmetric");
        }
    }
    int rc = ensure_nodes(g, n);
    if (rc) return rc;
    g->n = n;
    g->ne = ne;
    rc = ensure_edges(g, ne);
    if (rc) return rc;
    rc = ensure_dz_rows(g);
    if (rc) return rc;
    std::vector<uint16_t> xb((size_t)n*32), ub((size_t)ne*32);
    for (size_t k = 0; k < xb.size(); ++k) xb[k] = host_bf16(node_in_h[k]);
    for (size_t k = 0; k < ub.size(); ++k) ub[k] = host_bf16(edge_in_h[k]);
    HC(hipMemcpy(g->Xn, xb.data(), xb.size()*2, hipMemcpyHostToDevice));
    if (ne > 0) {
        HC(hipMemcpy(g->Ue, ub.data(), ub.size()*2, hipMemcpyHostToDevice));
        HC(hipMemcpy(g->cols, cols_h, sizeof(int32_t)*ne, hipMemcpyHostToDevice));
    }
    HC(hipMemcpy(g->rp, rowptr, sizeof(int32_t)*(n + 1), hipMemcpyHostToDevice));
    graph_finish(g->s, (int)n, g->rp, g->cols, g->recv, g->send, g->rev);
    HC(hipGetLastError());
    HC(hipStreamSynchronize(g->s));
    g->have_state = false;
    g->have_targets = false;
    return 0;
}

int psph_gnn_graph(psph_gnn *g, int64_t *nedges, int32_t *rowptr, int32_t *cols_h, int64_t cap)
{
    if (!g || !nedges) return psph_internal_fail(PSPH_ERR_INVALID, "null argument");
    *nedges = g->ne;
    if (!rowptr && !cols_h) return 0;
    if (cap < g->ne) return psph_internal_fail(PSPH_ERR_CAPACITY, "need room for all edges");
    HC(hipStreamSynchronize(g->s));
    if (rowptr) HC(hipMemcpy(rowptr, g->rp, sizeof(int32_t)*(g->n + 1), hipMemcpyDeviceToHost));
    if (cols_h && g->ne > 0) HC(hipMemcpy(cols_h, g->cols, sizeof(int32_t)*g->ne, hipMemcpyDeviceToHost));
    return 0;
}

static int forward(psph_gnn *g)
{
    if (g->n < 1) return psph_internal_fail(PSPH_ERR_STATE, "no inputs loaded");
    hipStream_t s = g->s;
    const int n = (int)g->n, ne = (int)g->ne;
    gemm_fwd_rows(s, n, 32, g->Xn, 32, g->Wt(g->enc_n1()), g->b(g->enc_n1()), 1, g->Zn0);
    gemm_fwd_rows(s, n, 128, g->Zn0, 128, g->Wt(g->enc_n2()), g->b(g->enc_n2()), 0, g->H[0]);
    if (ne > 0) {
        gemm_fwd_rows(s, ne, 32, g->Ue, 32, g->Wt(g->enc_e1()), g->b(g->enc_e1()), 1, g->Ze0);
        gemm_fwd_rows(s, ne, 128, g->Ze0, 128, g->Wt(g->enc_e2()), g->b(g->enc_e2()), 0, g->E[0]);
    }
    for (int l = 0; l < g->L; ++l) {
        uint16_t *hin = g->H[g->slot(l)], *hout = g->H[g->train ? l + 1 : ((l + 1) & 1)];
        uint16_t *ein = g->E[g->slot(l)], *eout = g->E[g->train ? l + 1 : ((l + 1) & 1)];
        uint16_t *ze = g->ZE[g->train ? l : 0], *ag = g->AG[g->train ? l : 0], *zn = g->ZN[g->train ? l : 0];
        if (ne > 0 && g->train) {
            // training keeps the first layer's activation for the backward pass
            gemm_fwd_edgecat(s, ne, ein, hin, g->recv, g->send, g->Wt(g->pe1(l)), g->b(g->pe1(l)), ze);
            gemm_fwd_resid(s, ne, ze, g->Wt(g->pe2(l)), g->b(g->pe2(l)), ein, eout);
        } else if (ne > 0) {
            edge_mlp_fused(s, ne, ein, hin, g->recv, g->send, g->Wt(g->pe1(l)), g->b(g->pe1(l)),
                           g->Wt(g->pe2(l)), g->b(g->pe2(l)), eout);
        }
        segsum(s, n, g->rp, eout, ag);
        gemm_fwd_nodecat(s, n, hin, ag, g->Wt(g->pn1(l)), g->b(g->pn1(l)), zn);
        gemm_fwd_resid(s, n, zn, g->Wt(g->pn2(l)), g->b(g->pn2(l)), hin, hout);
    }
    uint16_t *hL = g->H[g->train ? g->L : (g->L & 1)];
    gemm_fwd_rows(s, n, 128, hL, 128, g->Wt(g->dec1()), g->b(g->dec1()), 1, g->Zd);
    gemm_fwd_out(s, n, g->Zd, g->Wt(g->dec2()), g->b(g->dec2()), g->dout, g->Y);
    HC(hipGetLastError());
    return 0;
}

int psph_gnn_predict(psph_gnn *g, float *y)
{
    if (!g) return psph_internal_fail(PSPH_ERR_INVALID, "null handle");
    int rc = forward(g);
    if (rc) return rc;
    if (y) HC(hipMemcpyAsync(y, g->Y, sizeof(float)*g->dout*g->n, hipMemcpyDeviceToHost, g->s));
    HC(hipStreamSynchronize(g->s));
    return 0;
}

int psph_gnn_advance(psph_gnn *g, double *x, double *v, double *rho)
{
    if (!g) return psph_internal_fail(PSPH_ERR_INVALID, "null handle");
    if (!g->have_state) return psph_internal_fail(PSPH_ERR_STATE, "advance needs a loaded particle state");
    int rc = forward(g);
    if (rc) return rc;
    apply_out(g->s, (int)g->n, g->dim, g->Y, g->isfree, g->fs, g->x, g->v, g->rho);
    HC(hipGetLastError());
    if (x) HC(hipMemcpyAsync(x, g->x, sizeof(double)*3*g->n, hipMemcpyDeviceToHost, g->s));
    if (v) HC(hipMemcpyAsync(v, g->v, sizeof(double)*3*g->n, hipMemcpyDeviceToHost, g->s));
    if (rho) HC(hipMemcpyAsync(rho, g->rho, sizeof(double)*g->n, hipMemcpyDeviceToHost, g->s));
    HC(hipStreamSynchronize(g->s));
    return 0;
}

int psph_gnn_set_target_state(psph_gnn *g, const double *x1, const double *v1, const double *rho1)
{
    if (!g || !x1 || !v1 || !rho1) return psph_internal_fail(PSPH_ERR_INVALID, "null argument");
    if (!g->train) return psph_internal_fail(PSPH_ERR_STATE, "created without training buffers");
    if (!g->have_state) return psph_internal_fail(PSPH_ERR_STATE, "targets from states need a loaded state");
    const int64_t n = g->n;
    double *t = nullptr;
    HC(hipMalloc((void **)&t, sizeof(double)*7*n));
    auto cleanup = [&](int rc) { (void)hipFree(t); return rc; };
    if (hipMemcpy(t, x1, sizeof(double)*3*n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t + 3*n, v1, sizeof(double)*3*n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t + 6*n, rho1, sizeof(double)*n, hipMemcpyHostToDevice) != hipSuccess)
        return cleanup(psph_internal_fail(PSPH_ERR_HIP, "target upload"));
    targets(g->s, (int)n, g->dim, g->x, g->v, g->rho, t, t + 3*n, t + 6*n, g->isfree, g->fs, g->tgt, g->wgt);
    if (hipStreamSynchronize(g->s) != hipSuccess) return cleanup(psph_internal_fail(PSPH_ERR_HIP, "targets"));
    g->have_targets = true;
    return cleanup(0);
}

int psph_gnn_set_targets(psph_gnn *g, const float *tgt_h, const float *wgt_h)
{
    if (!g || !tgt_h || !wgt_h) return psph_internal_fail(PSPH_ERR_INVALID, "null argument");
    if (!g->train) return psph_internal_fail(PSPH_ERR_STATE, "created without training buffers");
    HC(hipMemcpy(g->tgt, tgt_h, sizeof(float)*g->dout*g->n, hipMemcpyHostToDevice));
    HC(hipMemcpy(g->wgt, wgt_h, sizeof(float)*g->n, hipMemcpyHostToDevice));
    g->wsum = 0.0;
    for (int64_t i = 0; i < g->n; ++i) g->wsum += wgt_h[i];
    g->have_targets = true;
    return 0;
}

static int backward(psph_gnn *g)
{
    hipStream_t s = g->s;
    const int n = (int)g->n, ne = (int)g->ne, L = g->L;
    HC(hipMemsetAsync(g->G, 0, g->nparam*4, s));
    // decoder
    gemm_wgrad_rows(s, n, 128, g->Zd, 128, g->dy, g->dW(g->dec2()), g->db(g->dec2()));
    gemm_bwd_mask(s, n, g->dy, g->W(g->dec2()), g->Zd, g->dZ);
    uint16_t *hL = g->H[L];
    gemm_wgrad_rows(s, n, 128, hL, 128, g->dZ, g->dW(g->dec1()), g->db(g->dec1()));
    int cur = 0;
    gemm_bwd_input(s, n, 128, g->dZ, g->W(g->dec1()), g->dh[cur]);
    add_cols(s, n, 128, 0, g->dh[cur], nullptr, g->dh[cur], g->dhb[cur]);
    const float *de_next = nullptr;
    int ecur = 0;
    for (int l = L - 1; l >= 0; --l) {
        const int nxt = cur ^ 1, enxt = ecur ^ 1;
        // node update h' = h + MLP([h, a])
        gemm_wgrad_rows(s, n, 128, g->ZN[l], 128, g->dhb[cur], g->dW(g->pn2(l)), g->db(g->pn2(l)));
        gemm_bwd_mask(s, n, g->dhb[cur], g->W(g->pn2(l)), g->ZN[l], g->dZ);
        gemm_wgrad_nodecat(s, n, g->H[l], g->AG[l], g->dZ, g->dW(g->pn1(l)), g->db(g->pn1(l)));
        gemm_bwd_input(s, n, 256, g->dZ, g->W(g->pn1(l)), g->dInN);
        if (ne > 0) {
            // aggregation: the gradient of each e' gains that of its receiver's sum
            edge_grad(s, ne, g->rp, g->recv, de_next, g->dInN + 128, 256, g->dEt, g->dEtb);
            // edge update e' = e + MLP([e, h_recv, h_send])
            gemm_wgrad_rows(s, ne, 128, g->ZE[l], 128, g->dEtb, g->dW(g->pe2(l)), g->db(g->pe2(l)));
            gemm_bwd_mask(s, ne, g->dEtb, g->W(g->pe2(l)), g->ZE[l], g->dZ);
            gemm_wgrad_edgecat(s, ne, g->E[l], g->H[l], g->recv, g->send, g->dZ, g->dW(g->pe1(l)), g->db(g->pe1(l)));
            gemm_bwd_input(s, ne, 384, g->dZ, g->W(g->pe1(l)), g->dInE);
            add_cols(s, ne, 384, 0, g->dInE, g->dEt, g->de[enxt], g->deb[enxt]);
            de_next = g->de[enxt];
            ecur = enxt;
        }
        node_grad(s, n, g->rp, g->rev, g->dh[cur], g->dInN, g->dInE, g->dh[nxt], g->dhb[nxt]);
        cur = nxt;
    }
    // encoders
    gemm_wgrad_rows(s, n, 128, g->Zn0, 128, g->dhb[cur], g->dW(g->enc_n2()), g->db(g->enc_n2()));
    gemm_bwd_mask(s, n, g->dhb[cur], g->W(g->enc_n2()), g->Zn0, g->dZ);
    gemm_wgrad_rows(s, n, 32, g->Xn, 32, g->dZ, g->dW(g->enc_n1()), g->db(g->enc_n1()));
    if (ne > 0) {
        gemm_wgrad_rows(s, ne, 128, g->Ze0, 128, g->deb[ecur], g->dW(g->enc_e2()), g->db(g->enc_e2()));
        gemm_bwd_mask(s, ne, g->deb[ecur], g->W(g->enc_e2()), g->Ze0, g->dZ);
        gemm_wgrad_rows(s, ne, 32, g->Ue, 32, g->dZ, g->dW(g->enc_e1()), g->db(g->enc_e1()));
    }
    HC(hipGetLastError());
    return 0;
}

int psph_gnn_train_step(psph_gnn *g, double lr, int32_t update, double *loss_out)
{
    if (!g) return psph_internal_fail(PSPH_ERR_INVALID, "null handle");
    if (!g->train) return psph_internal_fail(PSPH_ERR_STATE, "created without training buffers");
    if (!g->have_targets) return psph_internal_fail(PSPH_ERR_STATE, "no targets set");
    if (!(g->wsum > 0.0)) return psph_internal_fail(PSPH_ERR_INVALID, "targets have zero total weight");
    int rc = forward(g);
    if (rc) return rc;
    HC(hipMemsetAsync(g->loss, 0, 4, g->s));
    mse(g->s, (int)g->n, g->dout, g->Y, g->tgt, g->wgt, g->dy, g->loss, (float)(1.0/(g->dout*g->wsum)));
    rc = backward(g);
    if (rc) return rc;
    if (update) {
        g->step += 1;
        const double b1 = 0.9, b2 = 0.999;
        float c1 = (float)(1.0/(1.0 - std::pow(b1, (double)g->step)));
        float c2 = (float)(1.0/(1.0 - std::pow(b2, (double)g->step)));
        adam(g->s, g->nparam, g->P, g->G, g->Mo, g->Vo, (float)lr, (float)b1, (float)b2, 1e-8f, c1, c2);
        rc = refresh_bf16(g);
        if (rc) return rc;
    }
    float lh = 0.0f;
    HC(hipMemcpyAsync(&lh, g->loss, 4, hipMemcpyDeviceToHost, g->s));
    HC(hipStreamSynchronize(g->s));
    if (loss_out) *loss_out = lh;
    return 0;
}
