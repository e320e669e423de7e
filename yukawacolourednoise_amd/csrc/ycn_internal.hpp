// ycn_internal.hpp -- context and internal launch interfaces of libycn (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/ycn.h"
#include "cnn_mfma.hpp"

namespace ycn {

// ---------------------------------------------------------------------------
// FFT plan for one transform length (mixed radix, Stockham in LDS).
// ---------------------------------------------------------------------------
constexpr int kMaxFactors = 24;
struct FftPlan {
    int n;
    int nf;
    int radix[kMaxFactors];
    const double2 *tw;  // device: tw[e] = exp(-2 pi i e / n), e < n
};

// Per-system CG state (device resident; written by the last block of a kernel).
struct CgState {
    double rr;         // recursive |r|^2
    double bnorm2;     // |b|^2
    double stop;       // stop when sqrt(rr) <= stop
    double beta;
    double2 alpha;
    double true_rel;   // |b - A x| / |b| of the last true-residual check
    double tol;
    int iter;
    int done;
    int restarts;
    int converged;
    int maxIter;
    int mode;          // YCN_TOL_*
    int maxRestarts;
    int final_;        // mixed precision: converged or gave up (no further reliable updates)
    double rr_prev;    // |r|^2 before the last update (mixed precision: beta after a reliable update)
    double rz;         // flexible CG: <r, z> with z = CNN(r)
    double rz_prev;    // flexible CG: <r, z> before the last iteration
    double true_prev;  // flexible CG: true relative residual at the previous reliable update
    int precOff;       // flexible CG: 1 = preconditioner switched off (plain CG fallback)
    int restartNext;   // flexible CG: next beta = 0
    double enorm2;     // even-odd solve: |eta|^2 of the unpreconditioned system
    int iter0;         // pipelined even-odd iteration: iterations done when the chunk began
    double stop0;      // even-odd solve: the unscaled starting threshold tol |r_0|
    double need;       // even-odd solve: threshold the true residual asked for at the last check at the stop
    double rrRel;      // FP32 loops: recursive |r|^2 when the recursion was last anchored (start, reliable update)
    int nRu;           // pipelined even-odd solve: reliable updates so far
    int ruIter[7];     // pipelined even-odd solve: iteration count at each of the first seven
    double2 ab[2];     // pipelined even-odd solve (prologue finalize): alpha, beta of launch idx in ab[idx & 1]
};
constexpr int kMaxRounds = 7;

// FP32 CG loops pause for an FP64 reliable update once the recursive |r| has fallen by
// this factor since the recursion was last anchored to the true residual (Clark et al.,
// CPC 181 (2010), sec. 5).  Further FP32 iterations would take the recursive residual
// below the floor that FP32 rounding puts under the true one, and the direction kept
// across the update, scaled by |r_true|^2 / |r_prev|^2, would then swamp the new residual.
constexpr double kReliableDrop = 1e-5;

// Per-replica Langevin state (device resident).
struct StepState {
    double eps;        // last step size
    double tL;         // accumulated Langevin time
    double maxDrift;   // last (clamped) max |drift|
    int clamps;        // clamp events
    int pad;
};

// Control block of graph-captured Langevin steps (device resident).  Every kernel of a
// captured step outside the CG iteration chunks returns at once while `halt` is set; the
// step's check kernel sets it when the step's solve did not reach its stopping rule within
// the captured chunk schedule, so that the later steps of the run leave the state as it
// is (their chunk kernels keep iterating the unfinished solve) and the host finishes that
// step when it reads the block at the end of the segment of steps.
struct StepCtl {
    int halt;             // a step's solve did not finish within its schedule
    int converged;        // every finished step's solves met the stopping rule
    uint64_t haltStep;    // step counter of that step
    uint64_t step;        // step counter of the next step (noise and eta counters)
    double itersMaxSum;   // sum over finished steps of the largest iteration count of a solve
    double itersMeanSum;  // sum over finished steps of the mean iteration count
    double trueRelMax;    // largest true relative residual of the finished steps
    double rrMax;         // largest final recursive |r|^2 of the finished steps
    double stopScale;     // starting threshold of the next solve / tol |r_0| (relative mode)
    int steps;            // finished steps
    int lastIters;        // largest iteration count of the last finished step
    int restarts;         // reliable updates of the finished steps (largest over systems, summed)
    int nRounds;          // reliable updates (at most kMaxRounds) of that step's longest solve
    int rounds[7];        // and the iteration count at each
};

enum ProfKernel { PK_NONE = -1, PK_NORMAL_OP = 0, PK_CG_UPDATE, PK_DIRAC, PK_NOISE_COLS, PK_BOSONIC, PK_CNN, PK_COUNT };

struct Ctx {
    ycn_params p;
    int V = 0, H = 0, nRep = 0, nNoise = 0, nSys = 0;
    hipStream_t stream = nullptr;
    // second stream (default priority) for the coloured noise of a Langevin step, which
    // does not depend on the force; joined into `stream` before the update
    hipStream_t noiseStream = nullptr;
    hipEvent_t evMain = nullptr, evNoise = nullptr;
    bool noiseSide = false;   // the steps of this context put the noise on noiseStream
    std::string err;

    // host copies of per-replica parameters
    std::vector<double> m2;
    std::vector<uint64_t> seeds;
    double epsBar = 0;

    // device buffers
    double *phi[2] = {nullptr, nullptr};  // ping-pong for the fused fixed-eps update
    int cur = 0;
    double *noise = nullptr, *drift = nullptr;
    double *phiStage = nullptr;           // a host field passed to a single-system entry point (V)
    double2 *spec = nullptr;              // half spectrum [rep][t][H]
    int *kmax = nullptr;                  // mask: keep kx <= kmax[kt]
    int kmaxAll = -1;
    double *m2_d = nullptr;
    uint64_t *seed_d = nullptr;
    StepState *step_d = nullptr;
    FftPlan planX{}, planT{};

    // CG (FP64): [nSys][4][V] double2 each
    double2 *eta = nullptr, *x = nullptr, *r = nullptr, *pv[2] = {nullptr, nullptr}, *t = nullptr, *Ap = nullptr;
    CgState *cg_d = nullptr;
    double2 *partials = nullptr;          // [nSys][maxBlocks]
    int *active = nullptr;                // systems not yet converged
    int *active_h = nullptr;              // pinned host mirror
    CgState *cg_h = nullptr;              // pinned host mirror of the CG states
    hipEvent_t pollEv = nullptr;          // spin-waited by the solvers' polls
    int eoPredIters = 0;                  // iterations of the last even-odd mixed solve (chunk predictor)
    std::vector<int> eoRounds;            // iteration counts at that solve's reliable updates (its longest system)
    double eoStopScale = 1.0;             // starting threshold of even-odd solves / tol |r_0| (relative mode)
    int maxBlocks = 0;
    int lastIters = 32;                   // predictor for the CG launch chunk
    int plainIters = 0;                   // iterations of the last converged plain FP64 fermion solve (relative mode)

    // fused normal-operator tiles chosen by the autotuner (-1: three-kernel path)
    int tile64 = -1, tile32 = -1;
    // even-odd fused Schur normal operator tiles (-1: even-odd not available)
    int tileEo64 = -1, tileEo32 = -1;
    std::string tuneReport;

    // mixed precision: [nSys][4][V] float2
    float2 *rs = nullptr, *ps[2] = {nullptr, nullptr}, *Aps = nullptr, *xs = nullptr;
    // pipelined even-odd FP32 iteration ([nSys][4][V/2]): the second Ap buffer, and the third
    // direction buffer (the directions rotate through ps[0], ps[1], pfP2)
    float2 *pfA1 = nullptr, *pfP2 = nullptr;
    double2 *pfSlot[2] = {nullptr, nullptr};   // block partials of prologue-finalized launches
    bool eoFinFused = false;              // pipelined iteration finalizes in the next launch's prologue
    int eoPfBlocks = 0;                   // blocks per system of those launches (same tile throughout)
    float *phis = nullptr;

    // learned preconditioner
    float *cnnW = nullptr;                // master weights, Adam m, Adam v, gradient (4 x 4608)
    bf16x8 *cnnF1 = nullptr, *cnnF2 = nullptr, *cnnF3 = nullptr, *cnnD3 = nullptr, *cnnD2 = nullptr;
    float2 *zs = nullptr;                 // z = CNN(r) [nSys][4][V]
    float *cnnAct = nullptr;              // training activations
    _Float16 *cnnGates = nullptr;         // field gates of both hidden layers, [rep][site][32] (per solve)
    size_t cnnActV = 0;
    int adamStep = 0;
    double2 *pairR = nullptr, *pairX = nullptr;   // training snapshots [pairs][4][V]
    int pairCap = 0;
    // statistics of the last ycn_precond_train call
    double trainLossFirst = 0, trainLossLast = 0;
    int trainPairs = 0;

    // graph-captured Langevin steps (even-odd mixed precision): control block, its pinned
    // mirror, and one executable per (chunk schedule, field parity)
    StepCtl *ctl_d = nullptr, *ctl_h = nullptr;
    double *phiSolve = nullptr;           // the field of the captured step's solve (FP64)
    struct StepGraph {
        std::vector<int> chunks;   // launches of each chunk of the schedule
        int parity;                // field buffer the step reads
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> stepGraphs;
    bool stepGraphsOn = true;             // YCN_STEP_GRAPHS=0: every step launches eagerly
    int stepTail = 1, stepClean = 0;      // tail rounds of the schedule; clean segments since a halt
    int64_t hostSyncs = 0;                // host waits on the context's work since creation
    int64_t haltedSteps = 0;              // captured steps finished by the host

    // captured CG iteration chunks: key -> executable graph
    struct GraphEntry {
        const void *phi;
        int nsys, mtype, n, precision, tile;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    bool useGraphs = true;

    // profiling
    int profTarget = PK_NONE;
    std::vector<hipEvent_t> evStart, evStop;
    std::vector<int> evWork;   // per event pair: -1 = one launch; n >= 0 = a chunk of n working launches
    int evCount = 0;
    int profChunk = 0;         // 1 while a chunk is bracketed as a whole (per-launch records off)
    double profBytes = 0;      // algorithmic bytes of the profiled kernel's launches, summed
    int64_t profLaunches = 0;  // and their count
};

// error helpers
#define YCN_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            c.err = std::string(#call) + ": " + hipGetErrorString(e_);                  \
            return YCN_ERR_HIP;                                                         \
        }                                                                               \
    } while (0)

// profiling brackets around a launch of kernel `k`
void prof_before(Ctx &c, int k, double bytes);
void prof_after(Ctx &c, int k);
int prof_chunk_begin(Ctx &c, int k);
void prof_chunk_end(Ctx &c, int slot);

// autotuner timing memo (ycn_abi.hip): kind 0 = normal-operator tiles, 1 = even-odd
bool tune_lookup(int kind, const Ctx &c, std::vector<double> &t);
void tune_store(int kind, const Ctx &c, const std::vector<double> &t);

// noise.hip
int noise_init(Ctx &c);
void noise_free(Ctx &c);
// xi == nullptr: Philox field of (seeds, stream 0, step); else filter xi (device, nRep*V)
int noise_generate(Ctx &c, uint64_t step, const double *xi_dev, double *out_dev);
int noise_generate_on(Ctx &c, hipStream_t s, uint64_t step, const double *xi_dev, double *out_dev);
// the noise of a graph-captured step: counter and halt flag from the control block
int noise_generate_ctl(Ctx &c, hipStream_t s, const StepCtl *ctl, double *out_dev);
// the allocation-free entry point's scratch: address lookup on device dev (current),
// and the message of the last failed call on this thread (nullptr if none)
int noise_noalloc_setup(int dev);
const char *noise_noalloc_error();

// cg.hip
struct SolveStats {
    int iters_max = 0, restarts = 0, converged = 1, systems = 0;
    double rr_max = 0, true_rel_max = 0, iters_mean = 0;
};
int cg_alloc(Ctx &c);
void cg_free(Ctx &c);
int cg_autotune(Ctx &c);
// mixed precision (cg_mixed.hip)
int mp_alloc(Ctx &c);
void mp_free(Ctx &c);
int mp_solve(Ctx &c, const double *phi, int nsys, int mtype, bool gen, uint32_t streamBase, uint64_t step,
             SolveStats *st);
int dirac_apply(Ctx &c, const double *phi_dev, const double2 *in, double2 *out, int nsys, int dagger);
// b: if gen, eta is generated from (seeds, streamBase + j, step); else eta must hold b.
int cg_solve(Ctx &c, const double *phi_dev, int nsys, int mtype, bool gen, uint32_t streamBase,
             uint64_t step, SolveStats *st);
int fermion_drift(Ctx &c, const double *phi_dev, double *drift_dev);
// correlator / trace helpers
int spinor_slices(Ctx &c, const double2 *psi, double *out_dev);

// cnn.hip
int cnn_alloc(Ctx &c);
void cnn_free(Ctx &c);
int cnn_make_frags(Ctx &c);
double cnn_c0(const Ctx &c);
int cnn_set_weights(Ctx &c, const float *w);
int cnn_get_weights(Ctx &c, float *w);
int cnn_get_grad(Ctx &c, float *g);
// launches the fused CNN on nsys systems; returns the number of partials per system
// the field gates of the first nrep replicas from c.phis (before cnn_apply, once per field)
int cnn_gates(Ctx &c, int nrep);
int cnn_apply(Ctx &c, int nsys, const CgState *st, const float2 *rs, const float2 *Ap, float2 *z, double2 *part);
int cnn_train_step(Ctx &c, const double2 *r, const double2 *e, const float *phis, float lr, double *loss,
                   bool update);
// flexible CG with the CNN and the online training driver (cg.hip)
int fcg_solve(Ctx &c, const double *phi, int nsys, int mtype, bool gen, uint32_t streamBase, uint64_t step,
              SolveStats *st);
int train_preconditioner(Ctx &c, uint64_t step0, int nsteps, int pairsPerSolve, float lr, int *pairsUsed);

// update.hip
int bosonic_update(Ctx &c, int fixedEps);
// the update of a graph-captured step: phi_in -> phi_out (fixed eps; adaptive: in place
// on phi_in), no-op while *halt is set; leaves c.cur alone
int bosonic_update_gated(Ctx &c, int fixedEps, double *phi_in, double *phi_out, const int *halt);
int field_slices(Ctx &c, int rep, double *slices_dev, double *sums_dev);
int hot_start(Ctx &c, uint64_t initSeed);
int set_int(Ctx &c, int *ptr, int v);
int poll_active(Ctx &c, int &active);
int poll_iters(Ctx &c, int nsys, int &active, int &itMax);
int spin_sync(Ctx &c);   // waits for the context stream (counted in c.hostSyncs)
int stats_from_mirror(Ctx &c, int nsys, SolveStats *st);
int read_stats(Ctx &c, int nsys, SolveStats *st);
// CG building blocks shared with the even-odd solver (cg.hip): finalize alpha / beta from
// the partials in c.partials, the BLAS-1 updates x += alpha p, r -= alpha Ap on
// [nsys][4][n] vectors (c.x, c.r, c.Ap; mixed: c.xs, c.rs, c.Aps), returning their block
// count, the reliable-update accumulation x += x_s, and an FP64 -> FP32 copy.
void launch_fin_alpha(Ctx &c, int nsys, int nb);
void launch_fin_beta(Ctx &c, int nsys, int nb, bool mixed, bool even);
int launch_cg_update(Ctx &c, int nsys, int n, const double2 *p);
int launch_mp_update(Ctx &c, int nsys, int n, const float2 *p);
int launch_mp_accum(Ctx &c, int nsys, int n);
// b = (float)a over n entries; halt: return at once while *halt is set; copy: also
// copy a there (FP64)
int to_f32_n(Ctx &c, const double *a, float *b, int64_t n, const int *halt = nullptr, double *copy = nullptr);

// even-odd preconditioned fermion force (evenodd.hip)
bool eo_available(const Ctx &c);
int eo_pf_alloc(Ctx &c);
int eo_autotune(Ctx &c);
int eo_solve(Ctx &c, const double *phi, int nsys, bool gen, uint32_t streamBase, uint64_t step, SolveStats *st);
// ctl (graph-captured steps, else null): the drift kernel returns at once while halted
int eo_drift(Ctx &c, const double *phi, double *drift, const StepCtl *ctl = nullptr);
// Langevin steps as graph-captured steps (even-odd mixed precision): eligible when the
// context runs that solver with graphs on and no kernel profiled.  eo_graph_run does
// nsteps steps from counter step0 with one host wait per segment of captured steps and
// adds each step's figures to *tot the way a run of single steps sums them (iters_max,
// restarts and iters_mean summed, converged and-ed, rr_max and true_rel_max maxed).
bool eo_graph_eligible(const Ctx &c);
int eo_graph_run(Ctx &c, uint64_t step0, int nsteps, SolveStats *tot);
int eo_propagator(Ctx &c, const double *phi, double2 *out, SolveStats *st);

// precision conversion
int spinor_to_f32(Ctx &c, const double2 *a, float2 *b, int nsys);
int spinor_to_f64(Ctx &c, const float2 *a, double2 *b, int nsys);
int field_to_f32(Ctx &c, const double *a, float *b);

// layout conversion (abi)
int aos_to_soa(Ctx &c, const double *aos_dev, double2 *soa, int nsys);
int soa_to_aos(Ctx &c, const double2 *soa, double *aos_dev, int nsys);

}  // namespace ycn
