/*
 * persiansph_amd — C ABI of the MI355X (gfx950) particle-interaction path of PersianSPH.
 *
 * The reference has no FFI: its hot path is the body of SPH::Domain::Solve
 * (Source/Domain.cpp:1735-1797), reached from user code through the C++ Domain API
 * (Source/Domain.h:44-166).  This header is the seam that a drop-in Domain::Solve binds:
 * plain C, POD structs, int status codes, no exceptions across the boundary, no torch or
 * HIP types in any signature.  Each entry point names the reference code it replaces.
 *
 * Particle order at the boundary is the reference's particle order (the index into
 * Domain::Particles).  Inside the library particles live on the GPU in cell-sorted
 * structure-of-arrays form; upload/download convert.
 */
#ifndef PERSIANSPH_AMD_H
#define PERSIANSPH_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (0 = ok, negative = error; message via psph_last_error) ---- */
#define PSPH_OK               0
#define PSPH_ERR_INVALID     -1  /* bad argument or enum value (reference: abort() in Functions.cpp switches) */
#define PSPH_ERR_UNSUPPORTED -2  /* feature outside this path: in/out-flow with non-fluid particles or
                                    the Shepard filter, solid and soil particles in one run, FP32-mixed
                                    runs of several materials, a 2-D periodic z, a second Solve on a
                                    preset box (the message names which) */
#define PSPH_ERR_HIP         -3  /* HIP runtime failure */
#define PSPH_ERR_TIMESTEP    -4  /* adaptive dt collapsed below deltatint/1e5 (Domain.cpp:150-151) */
#define PSPH_ERR_CFL         -5  /* user dt above the CFL/viscous limit (Domain.cpp:1721-1733) */
#define PSPH_ERR_NO_DEVICE   -6  /* no GPU visible */
#define PSPH_ERR_STATE       -7  /* call out of order (e.g. run before upload) */
#define PSPH_ERR_CAPACITY    -8  /* output buffer too small / size limit exceeded */

/* ---- precision tiers ---- */
#define PSPH_FP64        0  /* every quantity in FP64 (parity tier) */
#define PSPH_FP32_MIXED  1  /* FP32 pair arithmetic on cell-relative coordinates, Kahan-compensated
                               sums; FP64 positions, velocities, densities and integration */

/* ---- why psph_run returned ---- */
#define PSPH_STOP_STEPS   0  /* max_steps steps completed */
#define PSPH_STOP_TEND    1  /* Time >= tf at the head of a step (Domain.cpp:1762 loop condition) */
#define PSPH_STOP_OUTPUT  2  /* Time >= tout after the force pass (Domain.cpp:1773): the state is
                                mid-step (forces of this step present, Move pending); the next
                                psph_run call finishes the step first */
#define PSPH_STOP_HOOK_BEFORE 3  /* host hook point of GeneralBefore (Domain.cpp:1767), enabled by
                                psph_set_hooks: after StartAcceleration, the in/out-flow
                                imposition and the neighbour search, before the wall smoothing
                                and the pair sums (mid_step = 2) */
#define PSPH_STOP_HOOK_AFTER  4  /* host hook point of GeneralAfter (Domain.cpp:1770): after the
                                pair sums, before AdaptiveTimeStep and Move (mid_step = 1); at an
                                output time the stop is PSPH_STOP_OUTPUT instead, at the same point */

/* Solver-wide settings: the Domain fields read inside Solve (Domain.h:89-164). */
typedef struct psph_config {
    int32_t dimension;       /* Domain::Dimension (2 or 3) */
    int32_t kernel_type;     /* Kernel_Set: 0 cubic spline, 1 Wendland "Quintic", 2 quintic spline */
    int32_t visc_eq;         /* Viscosity_Eq_Set: 0 Morris, 1 Shao, 2 incompressible, 3 Takeda */
    int32_t gradient_type;   /* Gradient_Approach_Set: 0 squared density, 1 multiplied density */
    int32_t scheme;          /* Domain::Scheme: 0 modified Verlet, 1 Leapfrog */
    int32_t precision;       /* PSPH_FP64 or PSPH_FP32_MIXED */
    int32_t periodic[3];     /* Boundary::Periodic (z only in 3-D; needs 3+ cells across a period) */
    int32_t in_out_flow;     /* Boundary::InOutFlow: 1 inflow, 2 outflow, 3 both (x direction) */
    int32_t swi_type;        /* Domain::SWIType: the soil-water interaction of multi-material runs
                                (0-3, CalcForce13 and the saturation of Domain.cpp) */
    int32_t fsi;             /* Domain::FSI: nonzero enables the fluid-solid interaction sums
                                (CalcForce12 with the FSI pressure and no-slip velocity) */
    double  cellfac;         /* Domain::Cellfac (2 for kernels 0/1, 3 for kernel 2) */
    double  gravity[3];      /* Domain::Gravity */
    double  xsph;            /* Domain::XSPH */
    double  sqrt_h_a;        /* Domain::sqrt_h_a (default 0.0025) */
    double  dom_max[3];      /* Domain::DomMax (default -1e11 per component) */
    double  dom_min[3];      /* Domain::DomMin (default +1e11 per component) */
    double  trpr[3];         /* Domain::TRPR on entry (all zero with BLPF: computed from particles) */
    double  blpf[3];         /* Domain::BLPF on entry */
    double  R;               /* Domain::R (periodic margin only) */
    double  mu_max;          /* Domain::MuMax on entry (the reference never initialises it; pass 0) */
    double  cs_max;          /* Domain::CsMax on entry (likewise; pass 0) */
    /* in/out-flow boundaries (Boundary_Condition.cpp:25-42; Domain.cpp:1327-1635), with the
     * default InCon/OutCon (constant velocity and density) */
    double  in_v[3], out_v[3];            /* Boundary::inv, Boundary::outv */
    double  in_density, out_density;      /* Boundary::inDensity, Boundary::outDensity */
    double  bc_cellfac;                   /* Boundary::cellfac (default 3.9): zone width / hmax */
    double  initial_dist;                 /* Domain::InitialDist: recycling offset */
    int32_t mass_conservation;            /* Boundary::MassConservation */
    int32_t reserved_io;
} psph_config;

/* One particle: the fields of SPH::Particle (Source/Particle.h:29-149) that the path reads
 * or writes, FP64 tensors row-major 3x3.  Defaults are the Particle constructor's
 * (Source/Particle.cpp:25-117). */
typedef struct psph_particle {
    double x[3], v[3], va[3], vb[3], a[3], NSv[3], VXSPH[3];
    double Density, Densitya, Densityb, dDensity, Pressure, Mass;
    double ZWab, SumDen, SumKernel;
    double Mu, ShearRate, SBar;
    double StrainRate[9], RotationRate[9];
    double ShearStress[9], ShearStressa[9], ShearStressb[9];
    double Sigma[9], Sigmaa[9], Sigmab[9];
    double Strain[9], Straina[9], Strainb[9];
    double TIR[9];
    double h, Cs, P0, RefDensity, FPMassC;
    double Alpha, Beta, MuRef, T0, m, CSmag;
    double TI, TIn, TIInitDist;
    double G, K, Sigmay, c, phi, psi;
    double FSIPressure, FSISumKernel, FSINSv[3];   /* fluid-solid interaction (Domain.cpp:1089-1160) */
    double n, n0, k, k2, d, V, RhoF, S;            /* soil-water interaction (Interaction.cpp:510-653) */
    int32_t IsFree, NoSlip, Shepard, LES;
    int32_t Material, PresEq, VisM, Fail;
    int32_t ID, ct, FirstStep, InOut;
    int32_t ShepardCounter, ShepardStep;
    int32_t SeepageType, IsSat, SatCheck;
    int32_t reserved2;
    int32_t Index;      /* out: position of this particle in the array given to psph_upload */
    int32_t reserved;
} psph_particle;

/* Scalars of the Solve loop after a psph_run call. */
typedef struct psph_status {
    double  time;            /* Domain::Time */
    double  deltat;          /* Domain::deltat (the dt of the last completed Move) */
    double  deltatmin;       /* Domain::deltatmin of the last force pass */
    int64_t steps;           /* steps completed by this call */
    int64_t pairs;           /* in-support pairs (rij/h <= Cellfac) of the last force pass */
    int64_t deleted;         /* particles removed by the leave check during this call */
    size_t  n;               /* particle count after the call */
    int32_t stop_reason;     /* PSPH_STOP_* */
    int32_t mid_step;        /* 1 if stopped after the force pass with Move pending, 2 if stopped
                                before the wall smoothing (PSPH_STOP_HOOK_BEFORE) */
    int64_t list_builds;     /* steps of this call that sorted the particles and built the
                                neighbour lists (all of them unless the lists are reused) */
} psph_status;

/* The cell grid built by CellInitiate (Domain.cpp:549-660). */
typedef struct psph_grid {
    double  blpf[3], trpr[3], cell_size[3];
    double  hmax, rhomax, mu_max, cs_max;
    int32_t cell_no[3];
    int32_t reserved;
} psph_grid;

typedef struct psph_handle psph_handle;

/* Creates a solver bound to the current HIP device, with its own stream.
 * Replaces Domain::Domain's solver state (Domain.cpp:54-102) for the Solve path. */
int psph_create(const psph_config *cfg, psph_handle **out);

/* Loads n particles (reference order) and runs Solve's initialisation: deltat = deltatint =
 * deltatmin = dt; InitialChecks (EOS pressure); CellInitiate; TimestepCheck (PSPH_ERR_CFL).
 * Replaces Domain.cpp:1735-1754 (InitialChecks :1662-1719, CellInitiate :549-660,
 * ListGenerate :662-784, TimestepCheck :1721-1733).  time is Domain::Time on entry. */
int psph_upload(psph_handle *h, const psph_particle *p, size_t n, double time, double dt);

/* Runs the Solve loop body (Domain.cpp:1762-1793) on the GPU: StartAcceleration,
 * MainNeighbourSearch, PrimaryComputeAcceleration, LastComputeAcceleration,
 * AdaptiveTimeStep, Move, Time += deltat, CheckParticleLeave, CellReset/ListGenerate.
 * Stops after max_steps steps, when Time >= tf at a step head, or when Time >= tout after
 * a force pass (then mid-step; pass tout = +inf to disable).  Particles leaving
 * [BLPF, TRPR] are deleted in order-preserving fashion as Array::DelItems does. */
int psph_run(psph_handle *h, int64_t max_steps, double tf, double tout, psph_status *st);

/* Copies the particle state back in reference order (cap = capacity of p). */
int psph_download(psph_handle *h, psph_particle *p, size_t cap, size_t *n);

/* Device form of the body-force hook of 1-Poiseuille.cpp:23-29 (GeneralBefore = UserAcc, which
 * adds a constant to the acceleration of every free particle after StartAcceleration and
 * before the pair sums, Domain.cpp:1766-1770): from the next force pass on, every free
 * particle's acceleration starts at Gravity + accel.  Kept across psph_upload; zero by default.
 * Arbitrary GeneralBefore/GeneralAfter hooks take the host-synchronised path (psph_set_hooks). */
int psph_set_body_force(psph_handle *h, const double accel[3]);

/* Host-synchronised user hooks: arbitrary GeneralBefore / GeneralAfter functions of the reference
 * (Domain.h:133-134, called at Domain.cpp:1767 and :1770 of every step), the slow path beside the
 * device form psph_set_body_force.  With before (after) nonzero every step stops at
 * PSPH_STOP_HOOK_BEFORE (PSPH_STOP_HOOK_AFTER); the caller downloads the state (psph_download:
 * at the stop before the wall pass it shows what StartAcceleration leaves: a = Gravity, or
 * Gravity + the body force for free particles, the other sums zero, fixed particles' Pressure,
 * NSv, Sigma and ShearStress zero), runs its hook, writes back what the hook changed with
 * psph_hook_accel and calls psph_run again, which finishes the step.  Kept across psph_upload;
 * off by default.  PSPH_ERR_STATE while a step is pending. */
int psph_set_hooks(psph_handle *h, int before, int after);

/* At a stop inside a step (PSPH_STOP_HOOK_BEFORE, PSPH_STOP_HOOK_AFTER or PSPH_STOP_OUTPUT):
 * accelerations a (n x 3) and, if not NULL, density rates ddensity (n) of every particle, in the
 * order of psph_download.  Before the wall pass they are the values the step's pair sums start
 * from (the reference's a and dDensity after GeneralBefore, Domain.cpp:1767-1769; without
 * ddensity the sums start from zero); after the pair sums they replace the sums Move uses
 * (GeneralAfter, :1770).  Only free particles' values are used: fixed particles' a and dDensity
 * feed nothing.  Other fields a hook changes are not taken back; the drop-in refuses hooks that
 * change them.  n must be the current particle count. */
int psph_hook_accel(psph_handle *h, const double *a, const double *ddensity, size_t n);

/* The fields Domain::WriteXDMF (Domain.cpp:1869-1964) converts to float for its HDF5 file, in
 * reference particle order, laid out as its datasets are: Position, Velocity, Acceleration
 * (3 per particle), Pressure, Density, Mass, h, Tag (= Particle::ID), Sigma and Strain (6 per
 * particle: xx xy xz yy yz zz).  Each value is the FP64 state psph_download would return,
 * rounded to nearest float.  Fixed particles' Acceleration is Gravity (their pair sums are not
 * formed; they feed nothing).  The pointers belong to the library and stay valid until the
 * next psph_output_begin, psph_upload or psph_destroy. */
typedef struct psph_output_f32 {
    size_t n;
    const float   *position, *velocity, *acceleration;
    const float   *pressure, *density, *mass, *h;
    const int32_t *tag;
    const float   *sigma, *strain;
} psph_output_f32;

/* Starts an output of the current state (at a PSPH_STOP_OUTPUT stop, the mid-step state the
 * reference writes at Domain.cpp:1773-1779): packs the fields to FP32 on the GPU and queues
 * their copy to pinned host memory on the solver's stream, then returns without waiting.
 * Later psph_run calls may be made before psph_output_end; they run after the copy.
 * Replaces the conversion loop of WriteXDMF (Domain.cpp:1893-1928) and the full state
 * download an output needs without it.  One output may be pending at a time. */
int psph_output_begin(psph_handle *h);

/* Waits for the pending output's copy and returns the packed fields. */
int psph_output_end(psph_handle *h, psph_output_f32 *out);

/* Test hook: the neighbour pair set of the current positions, as the reference's
 * MainNeighbourSearch + the rij/h <= Cellfac test in CalcForce11/2233 (Interaction.cpp:32,195)
 * would form it: unordered pairs (P1, P2) in reference particle indices with the reference's
 * orientation (P1 in the cell with the smaller key, or the larger index within one cell),
 * at least one free, same material.  pairs holds 2*cap int64 values. */
int psph_neighbour_pairs(psph_handle *h, int64_t *pairs, size_t cap, size_t *npairs);

/* psph_neighbour_pairs without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and state changes as
 * psph_neighbour_pairs, but the pair list is staged in a static 32 MiB device array (one per
 * GPU, held until the process ends) instead of a buffer allocated for the call, so at most
 * 2,097,152 pairs (2^21) can be returned: with more, *npairs is set and the call fails with
 * PSPH_ERR_CAPACITY.  Every call on one GPU shares that array: calls that overlap on one GPU
 * (from several threads or handles) race. */
int psph_neighbour_pairs_noalloc(psph_handle *h, int64_t *pairs, size_t cap, size_t *npairs);

/* The grid computed at upload. */
int psph_get_grid(psph_handle *h, psph_grid *g);

/* Profiling: with PSPH_TIMING_ALL every kernel launch is bracketed by HIP events on the solver's
 * stream, with PSPH_TIMING_FORCE only the force pass (two events per step instead of ~20, for
 * timed runs); psph_kernel_stats returns per-kernel launch counts and mean durations (ms). */
#define PSPH_TIMING_OFF   0
#define PSPH_TIMING_ALL   1
#define PSPH_TIMING_FORCE 2
int psph_set_timing(psph_handle *h, int enable);
int psph_kernel_stats(psph_handle *h, char *names, size_t names_len,
                      double *mean_ms, int64_t *launches, size_t cap, size_t *nkernels);

/* The HIP stream the solver launches on (hipStream_t as an opaque pointer). */
int psph_stream(psph_handle *h, void **stream);

/* Bytes of algorithmic (compulsory) HBM traffic per step for the current particle set and
 * precision tier (FP32-mixed: FP32 neighbour records), and of one force pass, as used for
 * roofline figures: every array a kernel needs is read once and written once.  Soil and solid
 * runs add their stress-rate tensors (~650 B per free particle; the force pass reads Sigma and
 * TIR of every particle and writes StrainRate and RotationRate of the free ones). */
int psph_traffic_model(psph_handle *h, double *bytes_per_step, double *force_bytes);

int psph_destroy(psph_handle *h);

/* Copies the calling thread's last error message (NUL-terminated); returns its length. */
int psph_last_error(char *buf, size_t len);

/* ABI version: major*100 + minor. */
int psph_version(void);

#ifdef __cplusplus
}
#endif

#endif /* PERSIANSPH_AMD_H */
