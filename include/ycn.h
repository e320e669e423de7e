/*
 * ycn.h -- C-ABI of the MI355X (gfx950) Langevin update of the 2D lattice Yukawa
 * model with coloured noise (libycn.so).
 *
 * The reference (Matteo294/YukawaColouredNoise) has no FFI: its hot path is a set
 * of C++ objects that code/src/main.cu calls once per Langevin step.  Each entry
 * point below names the reference interface it replaces (paths relative to the
 * reference root).  Conventions:
 *   - every call returns YCN_OK (0) or a negative YCN_ERR_* code; no exceptions
 *     cross the ABI; ycn_last_error() gives the message;
 *   - one context per host thread; calls on a context are serialised and ordered
 *     on the context's own HIP stream;
 *   - host arrays use the reference layouts: scalar fields are doubles row-major
 *     (t, x) with x fastest (code/src/Lattice.cu:16-30); spinors are AoS,
 *     4 complex<double> per site [f1s1, f1s2, f2s1, f2s2] as interleaved re/im
 *     doubles (code/src/Spinor.cuh:7-16); batched fields are replica-major;
 *   - device buffers are owned by the context; pointers passed in are borrowed
 *     for the duration of the call.
 * There is no CPU fallback: ycn_create fails with YCN_ERR_NO_DEVICE when no HIP
 * device is present.
 */
#ifndef YCN_H
#define YCN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YCN_ABI_VERSION 2

enum {
    YCN_OK = 0,
    YCN_ERR_INVALID = -1,       /* bad argument or parameter                          */
    YCN_ERR_HIP = -2,           /* HIP runtime error                                  */
    YCN_ERR_NO_DEVICE = -3,     /* no HIP device: the library never falls back to CPU */
    YCN_ERR_UNSUPPORTED = -4,   /* valid request this build does not implement        */
    YCN_ERR_NOT_CONVERGED = -5  /* a CG solve hit cgMaxIter (result still written)    */
};

/* matrix types of CGsolver::solve (code/src/params.h:8) */
enum { YCN_NORMAL = 0, YCN_DAGGER = 1 };
/* CG tolerance modes */
enum { YCN_TOL_ABS_RECURSIVE = 0, YCN_TOL_REL_TRUE = 1 };
/* CG precisions */
enum { YCN_CG_FP64 = 0, YCN_CG_MIXED = 1 };
/* preconditioners */
enum { YCN_PRECOND_NONE = 0, YCN_PRECOND_CNN = 1 };

typedef struct ycn_ctx ycn_ctx;

/* Runtime parameters.  Physics keys are those of code/input.toml after the
 * mapping in code/src/main.cu:113-170 (m2, lambda, g are the on-device values). */
typedef struct {
    int Nt, Nx;            /* lattice (code/src/params.h:29, compile-time there)       */
    int nReplicas;         /* independent chains in one launch (>= 1)                   */
    int nNoise;            /* Gaussian noise vectors per drift estimate (>= 1)          */
    double m2, lambda;     /* main.cu:117-129                                           */
    double g;              /* Yukawa coupling as used on device (main.cu:165-169)       */
    double M;              /* fermions.fermion_mass                                     */
    double r;              /* fermions.WilsonParam                                      */
    double cutFraction;    /* physics.cutFraction (coloured-noise cutoff s)             */
    int driftMode;         /* fermions.driftMode: 0 normal, 1 rescaled by s^2           */
    double epsBar;         /* langevin.averageEpsilon                                   */
    int adaptive;          /* 1: eps = epsBar*30/max|drift| (langevin_gpu_v2.cu:112)    */
    uint64_t seed;         /* random.seed (Philox key)                                  */
    double cgTol;          /* default 1e-8 (params.h:14)                                */
    int cgTolMode;         /* YCN_TOL_*; default absolute recursive (CGsolver.cu:107)   */
    int cgMaxIter;         /* default 1000 (params.h:15)                                */
    int cgPrecision;       /* YCN_CG_* (mixed: any lattice; the fused FP32 tiles where  */
                           /* one divides it, an FP32 two-stencil pair elsewhere)       */
    int reliableEvery;     /* mixed precision: FP64 reliable update period (50)         */
    int precond;           /* YCN_PRECOND_*                                             */
    int device;            /* HIP device ordinal                                        */
    int evenOdd;           /* 1: fermion force (and the measurement's point-source
                              propagator) by the even-odd preconditioned solve of
                              D^dag chi = eta (Schur complement on the even sites; the
                              reference's even-odd blocks, Dirac.cu:256-379, are
                              commented out).  Same stopping rule on |eta - D^dag chi|,
                              FP64 or mixed precision per cgPrecision.  Needs a fused
                              tile dividing the lattice: Nt even and Nx % 32 == 0, or
                              Nt % 16 == 0 and Nx % 16 == 0.  A_o^-1 needs the diagonal
                              2r + M + g phi away from zero (Wilson r > 0); a solve that
                              cannot reach tol returns YCN_ERR_NOT_CONVERGED.  Exclusive
                              with the learned preconditioner.  Default 0 (the
                              reference's solver).                                      */
} ycn_params;

typedef struct {
    int iters;             /* iterations of the last solve (max over systems)           */
    int restarts;          /* true-residual restarts / reliable updates                 */
    int converged;         /* 1 if every system met the stopping rule                   */
    int systems;           /* systems solved together                                   */
    double rr;             /* final recursive |r|^2 (max over systems)                  */
    double true_rel_resid; /* max_s |b - A x|_s / |b|_s, recomputed in FP64             */
    double iters_mean;     /* mean iterations over systems                              */
} ycn_cg_stats;

typedef struct {
    double magnetisation;  /* sum_x phi (gpuMagnetisation, reductions.cu:106-140)       */
    double phi2;           /* sum_x phi^2                                               */
    double trace;          /* sum_x drift_x / g of a fresh fermion drift (main.cu:472-481) */
    double eps;            /* last step size                                            */
    double langevin_time;  /* accumulated Langevin time on the device                   */
    int clamp_events;      /* max-drift clamps since creation (langevin_gpu_v2.cu:83-91)*/
} ycn_observables;

/* Defaults that reproduce the reference (code/src/params.h:13-15, main.cu). */
void ycn_default_params(ycn_params *p);

/* Replaces the construction block of code/src/main.cu:89-279 (Laplace, DiracOP,
 * FermionicDrift, CGsolver, KernelLaunchInfo, ColouredNoise, cudaMemcpyToSymbol). */
int ycn_create(const ycn_params *p, ycn_ctx **out);
void ycn_destroy(ycn_ctx *ctx);
const char *ycn_last_error(const ycn_ctx *ctx);
int ycn_abi_version(void);
/* Stencil-tile autotuner result of ycn_create: the measured time per launch of the
 * fused normal operator for each tile shape that divides the lattice, and the choice.
 * No reference counterpart (the reference sizes launches with
 * cudaOccupancyMaxPotentialBlockSize, code/src/device_info.h:18-66). */
const char *ycn_tuning_report(const ycn_ctx *ctx);

/* Per-replica m^2 and seeds for batched chains (nReplicas entries each; either may
 * be NULL to keep the ycn_params value).  No reference counterpart: the reference
 * runs replicas as separate jobs (runs.py:90-101). */
int ycn_set_replicas(ycn_ctx *ctx, const double *m2, const uint64_t *seeds);

/* Step-size scale after the epsBar calibration of main.cu:325-352
 * (myEpsBar *= myEpsBar / <eps>, then cudaMemcpyToSymbol(epsBar)). */
int ycn_set_eps_bar(ycn_ctx *ctx, double epsBar);

/* Field access (replaces the managed-memory phi of main.cu:92-95,413-414):
 * nReplicas * Nt * Nx doubles. */
int ycn_set_field(ycn_ctx *ctx, const double *phi);
int ycn_get_field(ycn_ctx *ctx, double *phi);
/* phi_i = 1 - 2u_i from Philox stream 99 of initSeed (replaces main.cu:92-95) */
int ycn_hot_start(ycn_ctx *ctx, uint64_t initSeed);

/* ColouredNoise::operator() (code/src/colouredNoise.cpp:155-178): Philox N(0,1)
 * field (stream 0, counter = step), R2C FFT, mask |p| <= s*sqrt(2)*pi (times 1/V),
 * C2R FFT.  Result stays in the context; copied to noise (nReplicas*V) if non-NULL. */
int ycn_coloured_noise(ycn_ctx *ctx, uint64_t step, double *noise);
/* The filter alone on a given white field xi (nReplicas*V): for parity tests. */
int ycn_coloured_filter(ycn_ctx *ctx, const double *xi, double *out);

/* No reference counterpart.  ycn_coloured_noise without a context and without any
 * allocation, for callers that cannot allocate device memory (code under HIP stream
 * capture, serving loops, harnesses that forbid hipMalloc).  Writes the same field as
 * ycn_coloured_noise of a context with these parameters (Nt, Nx, nReplicas,
 * cutFraction, seed or seeds[nReplicas], device) into the DEVICE buffer noise
 * (nReplicas*Nt*Nx doubles), enqueued on `stream` (a hipStream_t of that device; NULL
 * = the null stream); the call returns once the work is enqueued.  Its scratch is a
 * fixed device array of 8 479 232 bytes that the library holds on every device it runs
 * on until the process ends; limits: Nt, Nx <= 2048, nReplicas <= 64 and
 * nReplicas*Nt*(Nx/2+1) <= 1024*513 (one 1024^2 lattice); larger requests fail with
 * YCN_ERR_INVALID before anything is enqueued.  Calls on one device share that array:
 * calls that overlap on one device race (serialise them, e.g. on one stream).  The
 * first call on a device without a context looks the array up (no allocation); errors
 * are reported by ycn_last_error(NULL) on the calling thread. */
int ycn_coloured_noise_noalloc(const ycn_params *p, const uint64_t *seeds, uint64_t step, double *noise,
                               void *stream);

/* DiracOP::applyD (code/src/Dirac.cu:23-47): out = D in (YCN_NORMAL) or D^dag in
 * (YCN_DAGGER) with the context's field of replica 0 (or phi if non-NULL). */
int ycn_dirac_apply(ycn_ctx *ctx, const double *phi, const double *in, double *out, int mtype);

/* CGsolver::solve (code/src/CGsolver.cu:42-59): x = (D^dag D)^-1 b for YCN_DAGGER,
 * (D D^dag)^-1 b for YCN_NORMAL, replica-0 field (or phi if non-NULL). */
int ycn_cg_solve(ycn_ctx *ctx, const double *phi, const double *b, double *x, int mtype,
                 ycn_cg_stats *stats);

/* FermionicDrift::getForce (code/src/FermionicDrift.cu:57-80): eta_j (stream 1+j,
 * counter = step), CG on D^dag D, chi = D y, drift = -g Re sum conj(chi) eta (x s^2
 * in rescaled mode), averaged over nNoise.  drift (nReplicas*V) copied out if non-NULL. */
int ycn_fermion_force(ycn_ctx *ctx, uint64_t step, double *drift, ycn_cg_stats *stats);

/* One full Langevin step (main.cu:305-315 = cn(); getForce(); kli.Run(kAll)):
 * eps (nReplicas) copied out if non-NULL. */
int ycn_langevin_step(ycn_ctx *ctx, uint64_t step, double *eps, ycn_cg_stats *stats);
/* nsteps steps starting at counter step0 (the loop of main.cu:300-502 between two
 * exports, which synchronises and copies eps to the host every step: main.cu:305-315,
 * FermionicDrift.cu:61-78).  The step sizes and the Langevin time stay on the device.
 * With even-odd preconditioning and mixed precision every step after the context's
 * first solve is one replay of a captured graph (noise, solve setup, a fixed chunk
 * schedule sized from the previous solve, drift, update): the host waits once per
 * segment of up to 16 steps and never inside a step; a step whose solve outlasts its
 * schedule is finished by the host at the end of the segment.  Other configurations
 * poll CG convergence inside each step.  Results are the same either way.
 * stats (optional) receives the totals: iters = sum over steps of the per-step max,
 * iters_mean = mean per solve. */
int ycn_langevin_run(ycn_ctx *ctx, uint64_t step0, int nsteps, ycn_cg_stats *stats);

/* Host waits on the context's work since its creation (the CG solvers' convergence
 * polls, the reads of the captured steps' control block) and the captured steps the
 * host had to finish; either pointer may be NULL.  No counterpart in the reference:
 * it is how a caller sees how often the host synchronises inside its Langevin runs. */
int ycn_host_syncs(ycn_ctx *ctx, int64_t *syncs, int64_t *halted_steps);

/* Measurement block (main.cu:386-483): magnetisation, sum phi^2, eps, Langevin time
 * for replica `replica`; slices (Nt doubles, gpuTimeSlices) and corr (Nt doubles,
 * point-source propagator, gpuTimeSlices_spinors) if non-NULL; the trace needs a
 * fresh fermion drift with measurement noise (stream 0x4D00+j, counter = step).  With
 * evenOdd the propagator D^-1 src is gamma_5 chi from the even-odd solve of
 * D^dag chi = gamma_5 src (gamma_5 = sigma_2 per flavour, gamma_5 D gamma_5 = D^dag). */
int ycn_measure(ycn_ctx *ctx, int replica, uint64_t step, ycn_observables *obs,
                double *slices, double *corr);

/* Wait for all work on the context stream. */
int ycn_sync(ycn_ctx *ctx);

/* ---- learned preconditioner (no reference counterpart) ----
 * A 3-layer periodic 3x3 CNN, 8 -> 16 -> 16 -> 8 channels (channel 2c / 2c+1 = Re / Im
 * of spinor component c), no bias, BF16 weights and activations with FP32 accumulation
 * on MFMA.  The data path is linear in r and each hidden activation is gated by the
 * field: u1 = (W1*u0) . (1 + tanh(G1*q)), u2 = (W2*u1) . (1 + tanh(G2*q)), y = W3*u2,
 * q = g phi; so the preconditioner is a linear map of r whose stencil follows the local
 * Wilson diagonal.  It is a learned correction to an exact skip path:
 * z = c0 r + s * y(r / s), c0 = 1 / (2r + M)^2, s = RMS of r's 8V real entries.
 * The default network has a zero output layer (z = c0 r, i.e. plain CG).
 * Weights (YCN_CNN_NWEIGHTS floats): W1[9][8][16], W2[9][16][16], W3[9][16][8],
 * G1[9][16], G2[9][16] as [tap][in][out], tap = (dt+1)*3 + (dx+1),
 * out(site) = sum W * in(site + (dt,dx)).
 * With precond = YCN_PRECOND_CNN the fermion-force solves run flexible CG with this
 * network (FP32 inner iterations, FP64 reliable updates, plain CG on stall); the FP64
 * true residual with the plain operator decides convergence.  Needs Nt, Nx % 16 == 0. */
#define YCN_CNN_NWEIGHTS 4896
typedef struct {
    int pairs;             /* (residual, error) pairs trained on                        */
    double loss_first;     /* loss of the first pair (before its update)                */
    double loss_last;      /* loss of the last pair                                     */
    double eps_last;       /* Langevin step size of the last training step              */
} ycn_train_stats;
int ycn_precond_set_weights(ycn_ctx *ctx, const float *w);
int ycn_precond_get_weights(ycn_ctx *ctx, float *w);
/* z = CNN(r) for one AoS spinor (host arrays, 8V doubles each); the gates read the
 * context's replica-0 field */
int ycn_precond_apply(ycn_ctx *ctx, const double *r, double *z);
/* loss L = |A (e - z)|^2 / |r|^2 (A = D^dag D at phi, or the replica-0 field if phi is
 * NULL; for r = A e this is |r - A z|^2 / |r|^2, the residual the correction leaves)
 * for z = CNN(r) with gates from the same field, and dL/dW (YCN_CNN_NWEIGHTS floats);
 * no update */
int ycn_precond_loss_grad(ycn_ctx *ctx, const double *phi, const double *r, const double *e, double *loss,
                          float *grad);
/* Online training: nsteps Langevin steps from counter step0 (they advance the chain);
 * each fermion solve leaves pairsPerSolve (r_k, e_k = x* - x_k) pairs, each one Adam
 * step with learning rate lr.  nReplicas = nNoise = 1. */
int ycn_precond_train(ycn_ctx *ctx, uint64_t step0, int nsteps, int pairsPerSolve, double lr,
                      ycn_train_stats *stats);

/* Kernel timing with HIP events on the context stream (for roofline figures):
 * between begin and end every launch of the named kernel is bracketed by events
 * (iteration chunks are then launched one kernel at a time, not as graph replays);
 * end returns the summed duration and count of the launches that did work (launches
 * shorter than a quarter of the 90th percentile exited at once for converged systems and
 * are left out) and the algorithmic bytes moved per launch (averaged over the launches: the
 * pipelined iteration's launches alternate between two traffic figures).  The pipelined even-odd FP32
 * iteration with the finalize in the next launch's prologue is the exception: its
 * chunks hold nothing but back-to-back launches of that kernel, so one event pair
 * brackets each chunk and the count is the iterations done in it.  Kernel names:
 * "cg_normal_op", "cg_update", "dirac", "noise_cols", "bosonic", "cnn". */
int ycn_profile_begin(ycn_ctx *ctx, const char *kernel);
int ycn_profile_end(ycn_ctx *ctx, double *total_ms, int64_t *launches, double *bytes_per_launch);

#ifdef __cplusplus
}
#endif
#endif
