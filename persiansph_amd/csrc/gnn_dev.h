// Device-side structures and kernel launchers of the learned surrogate (gnn_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace psgnn {

// uniform cell grid for the radius graph (cell size >= radius)
struct GraphGrid {
    double lo[3];
    double inv_cell;
    double radius;
    int nc[3];
    int dim;
};

// input normalisation and output scales of the network
struct FeatScale {
    double inv_v, rho0, inv_rho;       // node inputs: v * inv_v, (rho/rho0 - 1) * inv_rho
    double out_dx, out_dv, out_drho;   // outputs: change = y * out_* (position: plus drift_t v)
    double drift_t;
    double inv_deg;                    // node input: neighbour count * inv_deg
};

void gemm_fwd_rows(hipStream_t s, int M, int K, const uint16_t *A, int lda, const uint16_t *Wt, const float *b,
                   int relu, uint16_t *Y);
void gemm_fwd_edgecat(hipStream_t s, int M, const uint16_t *E, const uint16_t *H, const int32_t *recv,
                      const int32_t *send, const uint16_t *Wt, const float *b, uint16_t *Y);
void gemm_fwd_nodecat(hipStream_t s, int M, const uint16_t *H, const uint16_t *A, const uint16_t *Wt,
                      const float *b, uint16_t *Y);
void gemm_fwd_resid(hipStream_t s, int M, const uint16_t *Z, const uint16_t *Wt, const float *b,
                    const uint16_t *R, uint16_t *Y);
void edge_mlp_fused(hipStream_t s, int M, const uint16_t *E, const uint16_t *H, const int32_t *recv,
                    const int32_t *send, const uint16_t *W1t, const float *b1, const uint16_t *W2t,
                    const float *b2, uint16_t *Eout);
void gemm_fwd_out(hipStream_t s, int M, const uint16_t *Z, const uint16_t *Wt, const float *b, int dout, float *O);
void gemm_bwd_mask(hipStream_t s, int M, const uint16_t *dY, const uint16_t *W, const uint16_t *Z, uint16_t *dZ);
void gemm_bwd_input(hipStream_t s, int M, int kin, const uint16_t *dY, const uint16_t *W, float *dX);
void gemm_wgrad_rows(hipStream_t s, int M, int kin, const uint16_t *X, int ldx, const uint16_t *dY, float *dW, float *db);
void gemm_wgrad_edgecat(hipStream_t s, int M, const uint16_t *E, const uint16_t *H, const int32_t *recv,
                        const int32_t *send, const uint16_t *dY, float *dW, float *db);
void gemm_wgrad_nodecat(hipStream_t s, int M, const uint16_t *H, const uint16_t *A, const uint16_t *dY, float *dW,
                        float *db);
void segsum(hipStream_t s, int n, const int32_t *rp, const uint16_t *E, uint16_t *a);
void edge_grad(hipStream_t s, int ne, const int32_t *rp, const int32_t *recv, const float *gnext, const float *da,
               int lda, float *g, uint16_t *gb);
void node_grad(hipStream_t s, int n, const int32_t *rp, const int32_t *rev, const float *dh_next, const float *dNode,
               const float *dEdge, float *dh, uint16_t *dhb);
void add_cols(hipStream_t s, int M, int ld, int c0, const float *src, const float *base, float *g, uint16_t *gb);
void mse(hipStream_t s, int n, int dout, const float *y, const float *tgt, const float *wgt, uint16_t *dy,
         float *loss, float inv_count);
void adam(hipStream_t s, size_t n, float *p, const float *g, float *m, float *v, float lr, float b1, float b2,
          float eps, float c1, float c2);
void weights_bf16(hipStream_t s, int kin, int kpad, const float *W, uint16_t *Wb, uint16_t *Wtb);
void graph_bin(hipStream_t s, int n, const GraphGrid &g, const double *x, int32_t *key, int32_t *cnt);
void graph_scatter(hipStream_t s, int n, const int32_t *key, const int32_t *cstart, int32_t *fill, int32_t *sorted);
void graph_cellorder(hipStream_t s, int ncell, const int32_t *cstart, int32_t *sorted);
void graph_edges(hipStream_t s, int pass, int n, const GraphGrid &g, const double *x, const int32_t *cstart,
                 const int32_t *sorted, int32_t *deg, const int32_t *rp, int32_t *cols);
void graph_finish(hipStream_t s, int n, const int32_t *rp, const int32_t *cols, int32_t *recv, int32_t *send, int32_t *rev);
void scan1(hipStream_t s, int n, const int32_t *in, int32_t *out);
void node_in(hipStream_t s, int n, int dim, const double *v, const double *rho, const int32_t *isfree,
             const int32_t *rp, const FeatScale &fs, uint16_t *X);
void edge_in(hipStream_t s, int ne, int dim, const int32_t *recv, const int32_t *send, const double *x, double inv_r,
             uint16_t *U);
void apply_out(hipStream_t s, int n, int dim, const float *y, const int32_t *isfree, const FeatScale &fs,
               double *x, double *v, double *rho);
void targets(hipStream_t s, int n, int dim, const double *x0, const double *v0, const double *r0, const double *x1,
             const double *v1, const double *r1, const int32_t *isfree, const FeatScale &fs, float *tgt, float *wgt);

}  // namespace psgnn
