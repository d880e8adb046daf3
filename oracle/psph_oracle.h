/*
 * psph_oracle — CPU FP64 restatement of PersianSPH's Solve hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness oracle and the CPU baseline
 * ("port" kind) for the GPU path in persiansph_amd.  Only tests/, the smoke check in
 * __graft_entry__.py and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * It keeps the reference's algorithm and data structures: an array of pointers to AoS
 * particles, a head-of-chain cell linked list, half-stencil candidate pair lists split by
 * thread, per-particle OpenMP locks, and the reference's expression order in every
 * formula.  Each function cites the reference file:line it restates.
 *
 * Parity status of each part is recorded in oracle/README.md.
 */
#ifndef PSPH_ORACLE_H
#define PSPH_ORACLE_H

#include "../include/persiansph_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct psph_odomain psph_odomain;

/* --- math primitives (Source/Functions.cpp) --- */
double psph_o_kernel(int dim, int kt, double q, double h);
double psph_o_grad_kernel(int dim, int kt, double q, double h);
double psph_o_laplace_kernel(int dim, int kt, double q, double h);
double psph_o_secderiv_kernel(int dim, int kt, double q, double h);
double psph_o_eos(int eq, double cs0, double p00, double density, double density0);
double psph_o_sound_speed(int eq, double cs0, double density, double density0);
double psph_o_density_solid(int eq, double cs0, double p00, double pressure, double density0);
void   psph_o_viscous_force(int viseq, double vi[3], double mu, double di, double dj, double gk,
                            const double vab[3], int dim, int kt, double rij, double h,
                            const double xij[3], const double vij[3]);
/* Symmetric 3x3 eigen decomposition (cyclic Jacobi), eigenvalues ascending; stands in for
 * gsl_eigen_symmv + gsl_eigen_symmv_sort(ASC) behind Rotation (Functions.cpp:334-348). */
void   psph_o_eig_sym3(const double m[9], double vals[3], double vecs[9]);
/* Artificial-stress tensor TIR of one particle (Domain.cpp:955-987). */
void   psph_o_tir(int dim, const double sigma[9], double ti, double density, double tir[9]);

/* --- single-particle integrator + constitutive update (Source/Particle.cpp:119-795) ---
 * i22 is Domain::I(2,2): 0 in 2-D, 1 in 3-D (InitialChecks, Domain.cpp:1665). */
int    psph_o_particle_move(psph_particle *p, double dt, int scheme, double i22);

/* --- domain --- */
/* nthreads = Domain::Nproc.  nthreads == 1 reproduces the reference's pair order exactly.
 * track_scales = 1 additionally accumulates |contribution| sums used by the parity metric. */
int  psph_o_create(const psph_config *cfg, const psph_particle *p, size_t n, double time,
                   double dt, int nthreads, int track_scales, psph_odomain **out);
int  psph_o_run(psph_odomain *d, int64_t max_steps, double tf, double tout, psph_status *st);
int  psph_o_get(psph_odomain *d, psph_particle *p, size_t cap, size_t *n);
int  psph_o_grid(psph_odomain *d, psph_grid *g);
/* In-support pairs (P1, P2) of the current positions, in the reference's search order
 * (before any force pass of the next step): SM pairs then FSM pairs of each thread list. */
int  psph_o_pairs(psph_odomain *d, int64_t *pairs, size_t cap, size_t *npairs);
/* body-force hook: accel is added to every free particle's a after StartAcceleration */
void psph_o_set_body_force(psph_odomain *d, const double accel[3]);
/* User hooks GeneralBefore / GeneralAfter (Domain.cpp:1767, 1770): called with the particles
 * (reference order, a contiguous copy that is copied back afterwards, so a hook may change any
 * field), after the neighbour search and after the pair sums of every step.  NULL: none. */
typedef void (*psph_o_hook)(psph_particle *p, size_t n, void *user);
void psph_o_set_hooks(psph_odomain *d, psph_o_hook before, psph_o_hook after, void *user);
/* Per-particle |contribution| sums of the last force pass, 32 doubles per particle:
 * a[3], dDensity, Pressure(wall), NSv[3], SumKernel, StrainRate[6], RotationRate[3], (pad),
 * VXSPH[3], ZWab, SumDen, FSIPressure, FSINSv[3], S, (pad x3) -- the wall and FSI
 * entries already divided by their kernel sums. */
int  psph_o_scales(psph_odomain *d, double *out, size_t cap);
void psph_o_destroy(psph_odomain *d);
const char *psph_o_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
