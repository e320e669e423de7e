/*
 * ycn_oracle.h -- CPU FP64 restatement of the reference's Langevin update of the
 * 2D lattice Yukawa model with coloured noise.
 *
 * THIS IS TEST INFRASTRUCTURE.  It is what the HIP library is tested against, never the
 * thing measured or shipped: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path (libycn.so) never calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * reference repository root).  Where the reference leaves something unspecified
 * (cuRAND XORWOW stream, cuFFT internals) the replacement is named and pinned by
 * the tests listed in each comment.
 *
 * Layouts are the reference's:
 *   scalar fields (phi, drift, noise): V doubles, row-major (t, x), x fastest
 *       (code/src/Lattice.cu:16-30, code/src/Laplace.cu:8-17)
 *   spinors: AoS, 4 complex<double> per site [f1s1, f1s2, f2s1, f2s2]
 *       (code/src/Spinor.cuh:7-16, code/src/Dirac.cu:59-62), interleaved re/im.
 */
#ifndef YCN_ORACLE_H
#define YCN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int Nt, Nx;             /* lattice extents (code/src/params.h:29)                 */
    double m2, lambda;      /* bosonic mass^2 and quartic coupling (main.cu:117-129)  */
    double g;               /* Yukawa coupling as used on device (main.cu:165-169)    */
    double M;               /* bare fermion mass (input.toml fermions.fermion_mass)   */
    double r;               /* Wilson parameter (fermions.WilsonParam)                */
    double cutFraction;     /* coloured-noise cutoff fraction s (physics.cutFraction) */
    int driftMode;          /* 0 normal, 1 rescaled by s^2 (FermionicDrift.cu:87-97)*/
    double epsBar;          /* averageEpsilon (langevin_gpu_v2.cu:65-117)            */
    int adaptive;           /* 1: eps = epsBar*30/max|drift| (reference); 0: eps=epsBar*/
    uint64_t seed;          /* random.seed                                            */
    double cgTol;           /* CG tolerance (params.h:14, default 1e-8)               */
    int cgTolMode;          /* 0: sqrt(r.r) <= tol absolute, recursive (CGsolver.cu:107)
                               1: relative: recursive sqrt(r.r) <= tol*|b|, then the true
                                  residual |b-Ax|/|b| is recomputed and the solve
                                  restarted from x while it exceeds tol               */
    int cgMaxIter;          /* params.h:15 (1000)                                     */
    int nNoise;             /* noise vectors per drift estimate (reference: 1)        */
    int evenOdd;            /* 1: fermion force through the even-odd preconditioned
                               solve (orc_eo_cg); 0: CG on D^dag D as the reference */
} orc_params;

typedef struct {
    int iters;              /* total CG iterations                                     */
    int restarts;           /* true-residual restarts (cgTolMode 1)                    */
    int converged;          /* 1 if the stopping rule was met before cgMaxIter         */
    double rr;              /* final recursive |r|^2                                    */
    double true_rel_resid;  /* |b - A x| / |b| recomputed in FP64                      */
} orc_cg_stats;

/* ---- random numbers (replaces cuRAND XORWOW, colouredNoise.cpp:31-36,53-67 and
 *      FermionicDrift.cu:50-55,102-110) --------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* n standard normals of stream (seed, stream) at step `step`; element i comes from
 * Philox counter (i/2, 0, step_lo, step_hi) via Box-Muller. */
void orc_normals(uint64_t seed, uint32_t stream, uint64_t step, int64_t n, double *out);
/* hot start phi_i = 1 - 2u_i, u from stream 99 of initSeed (replaces main.cu:92-95) */
void orc_hot_start(uint64_t initSeed, int64_t V, double *phi);

/* ---- FFT and coloured noise (colouredNoise.cpp:16-29,40-51,138-178) ---------- */
void orc_fft_c2c(int n, const double *in, double *out, int sign);  /* unnormalised */
void orc_mask(int Nt, int Nx, double cutFraction, double *mask);   /* Nt*(Nx/2+1)  */
void orc_coloured_filter(int Nt, int Nx, double cutFraction, const double *xi, double *out);
void orc_coloured_noise(const orc_params *p, uint64_t step, double *noise);

/* ---- Wilson-Dirac operator (Dirac.cu:10-70,144-190; Lattice.cu:16-30) -------- */
void orc_dirac(const orc_params *p, const double *phi, const double *in, double *out, int dagger);
/* A = D^dagger D for mtype 1 ("Dagger"), D D^dagger for mtype 0 ("Normal")     */
void orc_normal_op(const orc_params *p, const double *phi, const double *in, double *out, int mtype);

/* ---- CG (CGsolver.cu:72-177) ------------------------------------------------ */
/* maxIterOverride > 0 forces a fixed iteration cap (used to compare at fixed k). */
int orc_cg(const orc_params *p, const double *phi, const double *b, double *x, int mtype,
           int maxIterOverride, orc_cg_stats *st);

/* ---- fermion force (FermionicDrift.cu:57-100) --------------------------------- */
int orc_fermion_drift(const orc_params *p, const double *phi, uint64_t step, double *drift,
                      orc_cg_stats *st);
/* same, with the noise vectors passed in (nNoise * 4V complex) */
int orc_fermion_drift_eta(const orc_params *p, const double *phi, const double *eta,
                          double *drift, orc_cg_stats *st);
/* eta_j: 4V complex, real part N(0,1) from stream 1+j, imaginary part 0 */
void orc_fermion_noise(uint64_t seed, int j, uint64_t step, int64_t V, double *eta);

/* ---- bosonic drift and update (langevin_gpu_v2.cu:9-117, Laplace.cu:61-100,
 *      reductions.cu:30-63) ---------------------------------------------------- */
void orc_laplace_csr(int Nt, int Nx, int *I, int *J, double *val);
/* drift: in = fermion drift, out = total drift; phi updated in place; *eps out.
 * returns the (clamped) max |drift|. */
double orc_bosonic_update(const orc_params *p, double *phi, double *drift, const double *noise,
                          double *eps);

/* ---- one Langevin step (main.cu:305-315) ------------------------------------- */
int orc_langevin_step(const orc_params *p, double *phi, uint64_t step, double *eps,
                      orc_cg_stats *st);

/* ---- measurement (main.cu:386-483, reductions.cu:65-233) --------------------- */
double orc_sum(int64_t n, const double *v);
void orc_time_slices(int Nt, int Nx, const double *phi, double *slices);
/* point-source propagator: CG(DD^dag) then D^dag; corr[t] = sum_x sum_k Re psi */
int orc_correlator(const orc_params *p, const double *phi, double *corr, orc_cg_stats *st);

/* ---- even-odd preconditioning (Schur complement on the even sites; the reference's
 *      even-odd blocks are commented out, Dirac.cu:256-379).  Compact half-lattice
 *      vectors: AoS, site index t * (Nx/2) + (x >> 1) within a parity.  Even Nt, Nx. */
void orc_eo_schur(const orc_params *p, const double *phi, const double *in_e, double *out_e, int dagger);
void orc_eo_rhs(const orc_params *p, const double *phi, const double *eta, double *b_e);
void orc_eo_reconstruct(const orc_params *p, const double *phi, const double *eta, const double *chi_e,
                        double *chi);
/* chi = (D^dag)^-1 eta (full AoS) by CG on S^T S; stats as orc_cg */
int orc_eo_cg(const orc_params *p, const double *phi, const double *eta, double *chi, int maxIterOverride,
              orc_cg_stats *st);

/* ---- learned preconditioner (no reference counterpart) ------------------------
 * 3 layers of 3x3 periodic convolutions, 8 -> 16 -> 16 -> 8 channels, no bias.  The
 * data path is linear in r; each hidden activation is gated by the field:
 *   u1 = (W1 * u0) . (1 + tanh(G1 * q)),  u2 = (W2 * u1) . (1 + tanh(G2 * q)),
 *   y = W3 * u2,  q = g phi (the Yukawa part of the Wilson diagonal 2r + M + g phi),
 * so the preconditioner is a linear map of r whose stencil depends on the local field.
 * Channel 2c+0 / 2c+1 = Re / Im of spinor component c.  The network is a learned
 * correction to an exact scaled identity: z = c0 r + s y(r / s), with s = RMS of the
 * 8V real entries of r.
 * Weights: W1[9][8][16], W2[9][16][16], W3[9][16][8], G1[9][16], G2[9][16] ([tap][in]
 * [out], the gates' single input being q; tap = (dt+1)*3 + (dx+1), output(site) =
 * sum W * input(site + (dt, dx))), flattened in that order.
 * bf16 != 0 rounds the convolution weights and the activations u0, u1, u2 to bfloat16
 * (round to nearest even) as the GPU kernels do; accumulation is FP64 here. */
#define ORC_CNN_NW (9 * 8 * 16 + 9 * 16 * 16 + 9 * 16 * 8 + 2 * 9 * 16)
void orc_cnn_forward(int Nt, int Nx, const float *W, int bf16, double c0, double g, const double *phi,
                     const double *r, double *z);
/* residual-space loss of the correction z = CNN(r) for a solver pair (r, e):
 * L = |A (e - z)|^2 / |r|^2  (A = D^dag D; with r = A e this is |r - A z|^2 / |r|^2, the
 * residual left by the preconditioned correction relative to the residual), and dL/dW
 * by exact backprop in FP64 without bf16 rounding. */
double orc_cnn_loss_grad(const orc_params *p, const double *phi, const float *W, double c0, const double *r,
                         const double *e, double *grad);
double orc_bf16_round(double v);

/* misc */
int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
