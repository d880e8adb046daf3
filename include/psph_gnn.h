/*
 * psph_gnn — C ABI of the learned surrogate of the persiansph_amd particle path (MI355X).
 *
 * The reference has no learned component; these entry points have no counterpart in it.
 * They serve the surrogate-accelerated stepping engine (persiansph_amd/surrogate.py): a
 * message-passing network over the particle graph (edges between particles closer than the
 * kernel support) that advances a particle state by k solver steps at a time, trained on
 * trajectories of the plain solver by hand-written HIP forward, backward and Adam kernels
 * (BF16 weights and activations, FP32 accumulation on the matrix cores, FP32 master weights).
 *
 * Network: encoder MLPs on node inputs (velocity, density deviation, free flag, neighbour
 * count) and edge inputs (relative position and distance over the radius); `layers`
 * message-passing layers (edge update e += MLP([e, h_recv, h_send]), mean of incoming edges,
 * node update h += MLP([h, mean])); decoder MLP to the normalised change of position (net of
 * the drift), velocity and density (2*dim + 1 outputs).  Every MLP is
 * Linear(k, 128) -> ReLU -> Linear(128, 128).
 *
 * Status codes are those of persiansph_amd.h (message via psph_last_error).
 */
#ifndef PSPH_GNN_H
#define PSPH_GNN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psph_gnn_config {
    int32_t dim;            /* 2 or 3 */
    int32_t layers;         /* message-passing layers */
    int32_t hidden;         /* latent width: 128 */
    int32_t train;          /* nonzero: keep every layer's activations for backward passes */
    double radius;          /* graph radius (the kernel support Cellfac*h) */
    double rho0;            /* reference density of the inputs */
    double vel_scale;       /* node input v / vel_scale */
    double rho_scale;       /* node input (rho/rho0 - 1) / rho_scale */
    double out_dx;          /* output scales: dx = drift_t * v + y[0:dim] * out_dx */
    double out_dv;          /*                dv = y[dim:2dim] * out_dv */
    double out_drho;        /*                drho = y[2dim] * out_drho */
    double drift_t;         /* time the network advances (k solver steps); 0: no drift term */
    double deg_scale;       /* node input: neighbour count / deg_scale (0: 1) */
    uint64_t seed;          /* weight initialisation */
} psph_gnn_config;

typedef struct psph_gnn psph_gnn;

int psph_gnn_create(const psph_gnn_config *cfg, psph_gnn **out);
int psph_gnn_destroy(psph_gnn *g);

/* FP32 master parameters, in the fixed order: for every Linear layer (encoder node 1, 2,
   encoder edge 1, 2, then per message-passing layer edge 1, 2, node 1, 2, then decoder 1, 2)
   its weight [k_in_padded][128] row-major followed by its bias [128].  k_in_padded is 32 for
   the encoder input layers, 384 / 256 for the edge / node update input layers, else 128. */
int64_t psph_gnn_num_params(const psph_gnn *g);
int psph_gnn_get_params(psph_gnn *g, float *p, int64_t n);
int psph_gnn_set_params(psph_gnn *g, const float *p, int64_t n);

/* Loads a particle state (x, v: n x 3 FP64; rho: n; isfree: n) and builds the radius graph
   and the network inputs on the device. */
int psph_gnn_load_state(psph_gnn *g, const double *x, const double *v, const double *rho,
                        const int32_t *isfree, int64_t n);

/* Loads a graph (CSR by receiver: rowptr n+1, cols = senders, ascending in each row, the graph
   symmetric) and network inputs directly (node_in: n x 32, edge_in: nedges x 32, FP32,
   rounded to BF16), instead of a particle state. */
int psph_gnn_load_inputs(psph_gnn *g, const float *node_in, const float *edge_in, const int32_t *rowptr,
                         const int32_t *cols, int64_t n, int64_t nedges);

/* Number of edges of the loaded graph; with rowptr/cols not NULL also the CSR (cap entries). */
int psph_gnn_graph(psph_gnn *g, int64_t *nedges, int32_t *rowptr, int32_t *cols, int64_t cap);

/* Forward pass on the loaded inputs; y (n x (2*dim+1), FP32) may be NULL. */
int psph_gnn_predict(psph_gnn *g, float *y);

/* Forward pass and its state change applied to the loaded state (free particles); the new
   state is written to x, v, rho (n x 3, n x 3, n) when they are not NULL. */
int psph_gnn_advance(psph_gnn *g, double *x, double *v, double *rho);

/* Training targets: the normalised change from the loaded state to (x1, v1, rho1), position
   net of the drift drift_t * v, free particles weighted 1 and fixed ones 0 ... */
int psph_gnn_set_target_state(psph_gnn *g, const double *x1, const double *v1, const double *rho1);
/* ... or given directly (tgt: n x (2*dim+1), wgt: n). */
int psph_gnn_set_targets(psph_gnn *g, const float *tgt, const float *wgt);

/* One training step: forward, loss = sum_i w_i |y_i - t_i|^2 / (dout * sum_i w_i), backward
   into the FP32 gradients, then an Adam update (beta 0.9/0.999, eps 1e-8) when update != 0. */
int psph_gnn_train_step(psph_gnn *g, double lr, int32_t update, double *loss);

/* Gradients of the last train step (same order as the parameters). */
int psph_gnn_get_grads(psph_gnn *g, float *grad, int64_t n);

#ifdef __cplusplus
}
#endif

#endif
