// ycn_abi.hip -- the C-ABI of libycn (include/ycn.h): context, buffers, stream,
// profiling and the step sequence of code/src/main.cu:305-315.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "ycn_internal.hpp"

using namespace ycn;

struct ycn_ctx : public Ctx {};

namespace ycn {

void prof_before(Ctx &c, int k, double bytes) {
    if (c.profTarget != k) return;
    c.profBytes += bytes;
    ++c.profLaunches;
    if (c.profChunk || c.evCount >= (int)c.evStart.size()) return;
    (void)hipEventRecord(c.evStart[c.evCount], c.stream);
}
void prof_after(Ctx &c, int k) {
    if (c.profTarget != k || c.profChunk || c.evCount >= (int)c.evStart.size()) return;
    (void)hipEventRecord(c.evStop[c.evCount], c.stream);
    c.evWork[c.evCount] = -1;
    ++c.evCount;
}

// Bracket a whole chunk of back-to-back launches of kernel k with one event pair (no
// events between its launches).  Returns the pair's slot, or -1 when k is not profiled
// or the slots are used up; the caller sets evWork[slot] to the chunk's working launches
// once its iteration count is known.
int prof_chunk_begin(Ctx &c, int k) {
    if (c.profTarget != k || c.evCount >= (int)c.evStart.size()) return -1;
    (void)hipEventRecord(c.evStart[c.evCount], c.stream);
    c.profChunk = 1;
    return c.evCount;
}
void prof_chunk_end(Ctx &c, int slot) {
    if (slot < 0) return;
    (void)hipEventRecord(c.evStop[slot], c.stream);
    c.profChunk = 0;
    c.evWork[slot] = 0;
    c.evCount = slot + 1;
}

// Process-wide memo of the autotuners' measured times: a context whose device,
// lattice, system count, Wilson-r kernel variant and buffers match an earlier
// context's reuses its timings (the choice, including any tile forced through the
// environment, is still made per context).
namespace {
struct TuneKey {
    int kind, device, Nt, Nx, nSys, r1, f32, pf;
    bool operator==(const TuneKey &o) const {
        return kind == o.kind && device == o.device && Nt == o.Nt && Nx == o.Nx && nSys == o.nSys && r1 == o.r1 &&
               f32 == o.f32 && pf == o.pf;
    }
};
std::mutex g_tuneMutex;
std::vector<std::pair<TuneKey, std::vector<double>>> g_tune;
TuneKey tune_key(int kind, const Ctx &c) {
    return {kind, c.p.device, c.p.Nt, c.p.Nx, c.nSys, c.p.r == 1.0, c.rs != nullptr,
            c.pfA1 != nullptr && c.pfP2 != nullptr && c.pfSlot[0] != nullptr && c.pfSlot[1] != nullptr};
}
}  // namespace

bool tune_lookup(int kind, const Ctx &c, std::vector<double> &t) {
    if (const char *m = std::getenv("YCN_TUNE_MEMO"))
        if (std::atoi(m) == 0) return false;
    const TuneKey k = tune_key(kind, c);
    std::lock_guard<std::mutex> lock(g_tuneMutex);
    for (const auto &e : g_tune)
        if (e.first == k) {
            t = e.second;
            return true;
        }
    return false;
}

void tune_store(int kind, const Ctx &c, const std::vector<double> &t) {
    const TuneKey k = tune_key(kind, c);
    std::lock_guard<std::mutex> lock(g_tuneMutex);
    for (auto &e : g_tune)
        if (e.first == k) {
            e.second = t;
            return;
        }
    g_tune.push_back({k, t});
}

}  // namespace ycn

#define CHECK_CTX(ctx)                                  \
    do {                                                \
        if (!(ctx)) return YCN_ERR_INVALID;             \
    } while (0)

static int fail(Ctx &c, int code, const std::string &msg) {
    c.err = msg;
    return code;
}

// Copy a host array into a temporary device buffer owned by the caller of this helper.
struct DevTmp {
    void *p = nullptr;
    ~DevTmp() {
        if (p) (void)hipFree(p);
    }
};

extern "C" {

int ycn_abi_version(void) { return YCN_ABI_VERSION; }

void ycn_default_params(ycn_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->Nt = 16;             // code/src/params.h:29
    p->Nx = 16;
    p->nReplicas = 1;
    p->nNoise = 1;
    p->m2 = 0.0;
    p->lambda = 0.0;
    p->g = 0.0;
    p->M = 1.0;             // code/input.toml:26
    p->r = 1.0;             // code/input.toml:28
    p->cutFraction = 1.0;   // code/input.toml:7
    p->driftMode = 0;
    p->epsBar = 0.02;       // code/input.toml:10
    p->adaptive = 1;        // langevin_gpu_v2.cu:112
    p->seed = 1234;         // code/input.toml:22
    p->cgTol = 1e-8;        // params.h:14
    p->cgTolMode = YCN_TOL_ABS_RECURSIVE;
    p->cgMaxIter = 1000;    // params.h:15
    p->cgPrecision = YCN_CG_FP64;
    p->reliableEvery = 50;
    p->precond = YCN_PRECOND_NONE;
    p->device = 0;
}

const char *ycn_last_error(const ycn_ctx *ctx) {
    if (ctx) return ctx->err.c_str();
    const char *e = noise_noalloc_error();   // context-free calls of this thread
    return e ? e : "null context";
}

const char *ycn_tuning_report(const ycn_ctx *ctx) { return ctx ? ctx->tuneReport.c_str() : ""; }

void ycn_destroy(ycn_ctx *ctx) {
    if (!ctx) return;
    Ctx &c = *ctx;
    (void)hipSetDevice(c.p.device);
    if (c.noiseStream) (void)hipStreamSynchronize(c.noiseStream);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    noise_free(c);
    cg_free(c);
    mp_free(c);
    cnn_free(c);
    if (c.pairR) (void)hipFree(c.pairR);
    if (c.pairX) (void)hipFree(c.pairX);
    double **bufs[] = {&c.phi[0], &c.phi[1], &c.noise, &c.drift, &c.m2_d, &c.phiStage};
    for (double **b : bufs)
        if (*b) (void)hipFree(*b);
    if (c.seed_d) (void)hipFree(c.seed_d);
    if (c.step_d) (void)hipFree(c.step_d);
    for (auto e : c.evStart) (void)hipEventDestroy(e);
    for (auto e : c.evStop) (void)hipEventDestroy(e);
    if (c.evMain) (void)hipEventDestroy(c.evMain);
    if (c.evNoise) (void)hipEventDestroy(c.evNoise);
    if (c.noiseStream) (void)hipStreamDestroy(c.noiseStream);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete ctx;
}

int ycn_create(const ycn_params *p, ycn_ctx **out) {
    if (!p || !out) return YCN_ERR_INVALID;
    *out = nullptr;
    if (p->Nt < 2 || p->Nx < 2 || p->nReplicas < 1 || p->nNoise < 1 || p->cgMaxIter < 0 || p->cgTol < 0 ||
        (p->cgTolMode != YCN_TOL_ABS_RECURSIVE && p->cgTolMode != YCN_TOL_REL_TRUE) ||
        (p->driftMode != 0 && p->driftMode != 1) ||
        (p->cgPrecision != YCN_CG_FP64 && p->cgPrecision != YCN_CG_MIXED) || p->reliableEvery < 1 ||
        (p->precond != YCN_PRECOND_NONE && p->precond != YCN_PRECOND_CNN) || (p->evenOdd != 0 && p->evenOdd != 1))
        return YCN_ERR_INVALID;
    if ((int64_t)p->Nt * p->Nx > (int64_t)1 << 28) return YCN_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return YCN_ERR_NO_DEVICE;
    if (p->device < 0 || p->device >= ndev) return YCN_ERR_INVALID;
    ycn_ctx *ctx = new (std::nothrow) ycn_ctx();
    if (!ctx) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    c.p = *p;
    c.V = p->Nt * p->Nx;
    c.H = p->Nx / 2 + 1;
    c.nRep = p->nReplicas;
    c.nNoise = p->nNoise;
    c.nSys = c.nRep * c.nNoise;
    c.epsBar = p->epsBar;
    c.m2.assign(c.nRep, p->m2);
    c.seeds.assign(c.nRep, p->seed);
    int rc = YCN_OK;
    auto bail = [&](int code) {
        *out = nullptr;
        std::string e = c.err;
        ycn_destroy(ctx);
        return code;
    };
    if (hipSetDevice(p->device) != hipSuccess) return bail(YCN_ERR_HIP);
    // the allocation-free entry point's scratch address on this device (looked up at
    // the first context, or at that entry point's first call on a device without one)
    (void)noise_noalloc_setup(p->device);
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) return bail(YCN_ERR_HIP);
    // the noise stream is created at the first step that uses it; YCN_NOISE_STREAM=0
    // keeps the coloured noise on the context stream
    const char *ns = std::getenv("YCN_NOISE_STREAM");
    c.noiseSide = !(ns && std::atoi(ns) == 0) && c.p.precond == YCN_PRECOND_NONE;
    const size_t fb = sizeof(double) * (size_t)c.nRep * c.V;
    double **bufs[] = {&c.phi[0], &c.phi[1], &c.noise, &c.drift};
    for (double **b : bufs)
        if (hipMalloc(b, fb) != hipSuccess) return bail(YCN_ERR_HIP);
    if (hipMemset(c.phi[0], 0, fb) != hipSuccess || hipMemset(c.phi[1], 0, fb) != hipSuccess)
        return bail(YCN_ERR_HIP);
    if (hipMalloc(&c.m2_d, sizeof(double) * c.nRep) != hipSuccess) return bail(YCN_ERR_HIP);
    if (hipMalloc(&c.seed_d, sizeof(uint64_t) * c.nRep) != hipSuccess) return bail(YCN_ERR_HIP);
    if (hipMalloc(&c.step_d, sizeof(StepState) * c.nRep) != hipSuccess) return bail(YCN_ERR_HIP);
    if (hipMemset(c.step_d, 0, sizeof(StepState) * c.nRep) != hipSuccess) return bail(YCN_ERR_HIP);
    if (hipMemcpy(c.m2_d, c.m2.data(), sizeof(double) * c.nRep, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c.seed_d, c.seeds.data(), sizeof(uint64_t) * c.nRep, hipMemcpyHostToDevice) != hipSuccess)
        return bail(YCN_ERR_HIP);
    if ((rc = noise_init(c))) return bail(rc);
    if ((rc = cg_alloc(c))) return bail(rc);
    // FP32 buffers whenever the lattice admits the FP32 tiles or mixed precision is asked
    // for (lattices no tile divides run the FP32 two-stencil pair); CNN buffers whenever
    // the lattice admits the CNN tiles
    const bool f32ok = (p->Nt % 4 == 0) && (p->Nx % 16 == 0);
    const bool cnnok = (p->Nt % 16 == 0) && (p->Nx % 16 == 0);
    const bool eoDims = (p->Nt % 2 == 0) && (p->Nx % 2 == 0);
    if ((f32ok || p->cgPrecision == YCN_CG_MIXED || (p->evenOdd && eoDims)) && (rc = mp_alloc(c))) return bail(rc);
    if (cnnok && (rc = cnn_alloc(c))) return bail(rc);
    if (c.p.precond == YCN_PRECOND_CNN && !cnnok) {
        c.err = "the learned preconditioner needs Nt % 16 == 0 and Nx % 16 == 0";
        return bail(YCN_ERR_UNSUPPORTED);
    }
    // YCN_GRAPHS=0 launches CG chunks kernel by kernel instead of replaying captured
    // graphs; results are identical.  Without YCN_GRAPHS, chunks also launch eagerly when
    // a profiler tool library is loaded (rocprofv3 announces it in ROCP_TOOL_LIBRARIES):
    // on ROCm 7.2 a kernel trace of a process that captures these graphs ends in a
    // host-side segmentation fault inside the runtime below the first capture.
    if (const char *gr = std::getenv("YCN_GRAPHS"))
        c.useGraphs = std::atoi(gr) != 0;
    else if (std::getenv("ROCP_TOOL_LIBRARIES"))
        c.useGraphs = false;
    // YCN_STEP_GRAPHS=0 keeps ycn_langevin_run on single eagerly launched steps (with
    // captured CG chunks) where it would replay whole captured steps; results are identical
    if (const char *sg = std::getenv("YCN_STEP_GRAPHS")) c.stepGraphsOn = std::atoi(sg) != 0;
    if ((rc = cg_autotune(c))) return bail(rc);
    if (c.p.evenOdd) {
        if (c.p.precond == YCN_PRECOND_CNN) {
            c.err = "even-odd preconditioning and the learned preconditioner are exclusive";
            return bail(YCN_ERR_UNSUPPORTED);
        }
        if ((rc = eo_autotune(c))) return bail(rc);
        if (!eo_available(c)) {
            c.err = "even-odd preconditioning needs Nt even and Nx % 32 == 0, or Nt % 16 == 0 and Nx % 16 == 0";
            return bail(YCN_ERR_UNSUPPORTED);
        }
    }
    if (hipDeviceSynchronize() != hipSuccess) return bail(YCN_ERR_HIP);
    *out = ctx;
    return YCN_OK;
}

int ycn_set_replicas(ycn_ctx *ctx, const double *m2, const uint64_t *seeds) {
    CHECK_CTX(ctx);
    Ctx &c = *ctx;
    if (m2) c.m2.assign(m2, m2 + c.nRep);
    if (seeds) c.seeds.assign(seeds, seeds + c.nRep);
    YCN_HIP(hipMemcpyAsync(c.m2_d, c.m2.data(), sizeof(double) * c.nRep, hipMemcpyHostToDevice, c.stream));
    YCN_HIP(hipMemcpyAsync(c.seed_d, c.seeds.data(), sizeof(uint64_t) * c.nRep, hipMemcpyHostToDevice, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

int ycn_set_eps_bar(ycn_ctx *ctx, double epsBar) {
    CHECK_CTX(ctx);
    if (!(epsBar > 0)) return YCN_ERR_INVALID;
    ctx->epsBar = epsBar;
    return YCN_OK;
}

int ycn_set_field(ycn_ctx *ctx, const double *phi) {
    CHECK_CTX(ctx);
    if (!phi) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    YCN_HIP(hipMemcpyAsync(c.phi[c.cur], phi, sizeof(double) * (size_t)c.nRep * c.V, hipMemcpyHostToDevice, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

int ycn_get_field(ycn_ctx *ctx, double *phi) {
    CHECK_CTX(ctx);
    if (!phi) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    YCN_HIP(hipMemcpyAsync(phi, c.phi[c.cur], sizeof(double) * (size_t)c.nRep * c.V, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

int ycn_hot_start(ycn_ctx *ctx, uint64_t initSeed) {
    CHECK_CTX(ctx);
    Ctx &c = *ctx;
    int rc = hot_start(c, initSeed);
    if (rc) return rc;
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

int ycn_coloured_noise(ycn_ctx *ctx, uint64_t step, double *noise) {
    CHECK_CTX(ctx);
    Ctx &c = *ctx;
    int rc = noise_generate(c, step, nullptr, c.noise);
    if (rc) return rc;
    if (noise)
        YCN_HIP(hipMemcpyAsync(noise, c.noise, sizeof(double) * (size_t)c.nRep * c.V, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

int ycn_coloured_filter(ycn_ctx *ctx, const double *xi, double *out) {
    CHECK_CTX(ctx);
    if (!xi || !out) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    const size_t nb = sizeof(double) * (size_t)c.nRep * c.V;
    DevTmp d;
    YCN_HIP(hipMalloc(&d.p, nb));
    YCN_HIP(hipMemcpyAsync(d.p, xi, nb, hipMemcpyHostToDevice, c.stream));
    int rc = noise_generate(c, 0, (const double *)d.p, c.noise);
    if (rc) return rc;
    YCN_HIP(hipMemcpyAsync(out, c.noise, nb, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

// phi for single-system entry points: replica 0 of the context, or a host array copied
// into the context's staging buffer (one fixed address, so the captured CG chunks that
// read it are keyed by a buffer that lives as long as the context)
static int stage_phi(Ctx &c, const double *phi_host, const double **phi_dev) {
    if (!phi_host) {
        *phi_dev = c.phi[c.cur];
        return YCN_OK;
    }
    if (!c.phiStage) YCN_HIP(hipMalloc(&c.phiStage, sizeof(double) * c.V));
    YCN_HIP(hipMemcpyAsync(c.phiStage, phi_host, sizeof(double) * c.V, hipMemcpyHostToDevice, c.stream));
    *phi_dev = c.phiStage;
    return YCN_OK;
}

int ycn_dirac_apply(ycn_ctx *ctx, const double *phi, const double *in, double *out, int mtype) {
    CHECK_CTX(ctx);
    if (!in || !out || (mtype != YCN_NORMAL && mtype != YCN_DAGGER)) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    const size_t nb = sizeof(double2) * 4 * (size_t)c.V;
    DevTmp aos;
    const double *phi_dev = nullptr;
    int rc = stage_phi(c, phi, &phi_dev);
    if (rc) return rc;
    YCN_HIP(hipMalloc(&aos.p, nb));
    YCN_HIP(hipMemcpyAsync(aos.p, in, nb, hipMemcpyHostToDevice, c.stream));
    if ((rc = aos_to_soa(c, (const double *)aos.p, c.t, 1))) return rc;
    const int save = c.nNoise;
    c.nNoise = 1;  // one system, replica-0 field
    rc = dirac_apply(c, phi_dev, c.t, c.Ap, 1, mtype == YCN_DAGGER);
    c.nNoise = save;
    if (rc) return rc;
    if ((rc = soa_to_aos(c, c.Ap, (double *)aos.p, 1))) return rc;
    YCN_HIP(hipMemcpyAsync(out, aos.p, nb, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

static void fill_stats(const SolveStats &s, ycn_cg_stats *o) {
    if (!o) return;
    o->iters = s.iters_max;
    o->restarts = s.restarts;
    o->converged = s.converged;
    o->systems = s.systems;
    o->rr = s.rr_max;
    o->true_rel_resid = s.true_rel_max;
    o->iters_mean = s.iters_mean;
}

int ycn_cg_solve(ycn_ctx *ctx, const double *phi, const double *b, double *x, int mtype, ycn_cg_stats *stats) {
    CHECK_CTX(ctx);
    if (!b || !x || (mtype != YCN_NORMAL && mtype != YCN_DAGGER)) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    const size_t nb = sizeof(double2) * 4 * (size_t)c.V;
    DevTmp aos;
    const double *phi_dev = nullptr;
    int rc = stage_phi(c, phi, &phi_dev);
    if (rc) return rc;
    YCN_HIP(hipMalloc(&aos.p, nb));
    YCN_HIP(hipMemcpyAsync(aos.p, b, nb, hipMemcpyHostToDevice, c.stream));
    if ((rc = aos_to_soa(c, (const double *)aos.p, c.eta, 1))) return rc;
    const int save = c.nNoise;
    c.nNoise = 1;
    SolveStats st;
    rc = cg_solve(c, phi_dev, 1, mtype, false, 0, 0, &st);
    c.nNoise = save;
    if (rc) return rc;
    if ((rc = soa_to_aos(c, c.x, (double *)aos.p, 1))) return rc;
    YCN_HIP(hipMemcpyAsync(x, aos.p, nb, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    fill_stats(st, stats);
    return st.converged ? YCN_OK : YCN_ERR_NOT_CONVERGED;
}

static int force_impl(Ctx &c, uint64_t step, uint32_t streamBase, SolveStats *st) {
    if (c.p.evenOdd) {   // chi = (D^dag)^-1 eta through the Schur complement on the even sites
        int rc = eo_solve(c, c.phi[c.cur], c.nSys, true, streamBase, step, st);
        if (rc) return rc;
        return eo_drift(c, c.phi[c.cur], c.drift);
    }
    int rc = cg_solve(c, c.phi[c.cur], c.nSys, YCN_DAGGER, true, streamBase, step, st);
    if (rc) return rc;
    return fermion_drift(c, c.phi[c.cur], c.drift);
}

int ycn_fermion_force(ycn_ctx *ctx, uint64_t step, double *drift, ycn_cg_stats *stats) {
    CHECK_CTX(ctx);
    Ctx &c = *ctx;
    SolveStats st;
    int rc = force_impl(c, step, 1u, &st);
    if (rc) return rc;
    if (drift)
        YCN_HIP(hipMemcpyAsync(drift, c.drift, sizeof(double) * (size_t)c.nRep * c.V, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    fill_stats(st, stats);
    return st.converged ? YCN_OK : YCN_ERR_NOT_CONVERGED;
}

// cn() does not depend on the force, so unless a kernel is being profiled it runs on
// the noise stream under the fermion solve: it first waits for everything already on
// the context stream (the previous update reads the buffer it overwrites), and the
// context stream waits for it before the update, also when the solve failed.  Not with
// the learned preconditioner: its flexible-CG steps measured 7-10 % slower once the
// context owned a second stream, used or not (34.2 against 30.8-31.8 steps/s at
// 1024^2), for a cause not yet found; such contexts never create it.
static int step_impl(Ctx &c, uint64_t step, SolveStats *st) {
    if (c.noiseSide && !c.noiseStream) {
        if (hipSetDevice(c.p.device) != hipSuccess ||
            hipStreamCreateWithFlags(&c.noiseStream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c.evMain, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c.evNoise, hipEventDisableTiming) != hipSuccess) {
            // the side stream is only an overlap: without it the noise stays on the
            // context stream
            (void)hipGetLastError();
            if (c.evMain) (void)hipEventDestroy(c.evMain);
            if (c.evNoise) (void)hipEventDestroy(c.evNoise);
            if (c.noiseStream) (void)hipStreamDestroy(c.noiseStream);
            c.evMain = c.evNoise = nullptr;
            c.noiseStream = nullptr;
            c.noiseSide = false;
        }
    }
    const bool side = c.noiseSide && c.noiseStream && c.profTarget == PK_NONE;
    int rc;
    if (side) {
        YCN_HIP(hipEventRecord(c.evMain, c.stream));
        YCN_HIP(hipStreamWaitEvent(c.noiseStream, c.evMain, 0));
        rc = noise_generate_on(c, c.noiseStream, step, nullptr, c.noise);   // cn()
        if (rc) return rc;
        YCN_HIP(hipEventRecord(c.evNoise, c.noiseStream));
    } else {
        rc = noise_generate(c, step, nullptr, c.noise);                      // cn()
        if (rc) return rc;
    }
    rc = force_impl(c, step, 1u, st);                                         // fDrift.getForce(drift)
    if (side) YCN_HIP(hipStreamWaitEvent(c.stream, c.evNoise, 0));
    if (rc) return rc;
    return bosonic_update(c, c.p.adaptive ? 0 : 1);                           // kli.Run(kAll)
}

int ycn_langevin_step(ycn_ctx *ctx, uint64_t step, double *eps, ycn_cg_stats *stats) {
    CHECK_CTX(ctx);
    Ctx &c = *ctx;
    SolveStats st;
    int rc = step_impl(c, step, &st);
    if (rc) return rc;
    if (eps) {
        std::vector<StepState> S(c.nRep);
        YCN_HIP(hipMemcpyAsync(S.data(), c.step_d, sizeof(StepState) * c.nRep, hipMemcpyDeviceToHost, c.stream));
        YCN_HIP(hipStreamSynchronize(c.stream));
        for (int i = 0; i < c.nRep; ++i) eps[i] = S[i].eps;
    }
    YCN_HIP(hipStreamSynchronize(c.stream));
    fill_stats(st, stats);
    return st.converged ? YCN_OK : YCN_ERR_NOT_CONVERGED;
}

// With the even-odd mixed solver (and graphs on, no kernel profiled) the steps after the
// context's first solve replay captured whole steps: no host wait inside a step, one per
// segment of steps (eo_graph_run); every other configuration runs single steps.
int ycn_langevin_run(ycn_ctx *ctx, uint64_t step0, int nsteps, ycn_cg_stats *stats) {
    CHECK_CTX(ctx);
    if (nsteps < 0) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    SolveStats tot;
    tot.systems = c.nSys;
    tot.iters_mean = 0;   // summed over the steps, divided at the end
    int i = 0;
    for (; i < nsteps && !eo_graph_eligible(c); ++i) {
        SolveStats st;
        int rc = step_impl(c, step0 + (uint64_t)i, &st);
        if (rc) return rc;
        tot.iters_max += st.iters_max;
        tot.restarts += st.restarts;
        tot.converged &= st.converged;
        tot.rr_max = std::max(tot.rr_max, st.rr_max);
        tot.true_rel_max = std::max(tot.true_rel_max, st.true_rel_max);
        tot.iters_mean += st.iters_mean;
    }
    if (i < nsteps) {
        const int rc = eo_graph_run(c, step0 + (uint64_t)i, nsteps - i, &tot);
        if (rc) return rc;
    } else {
        YCN_HIP(hipStreamSynchronize(c.stream));
    }
    tot.iters_mean = nsteps ? tot.iters_mean / nsteps : 0;
    fill_stats(tot, stats);
    return tot.converged ? YCN_OK : YCN_ERR_NOT_CONVERGED;
}

int ycn_host_syncs(ycn_ctx *ctx, int64_t *syncs, int64_t *halted_steps) {
    CHECK_CTX(ctx);
    if (syncs) *syncs = ctx->hostSyncs;
    if (halted_steps) *halted_steps = ctx->haltedSteps;
    return YCN_OK;
}

int ycn_measure(ycn_ctx *ctx, int replica, uint64_t step, ycn_observables *obs, double *slices, double *corr) {
    CHECK_CTX(ctx);
    Ctx &c = *ctx;
    if (replica < 0 || replica >= c.nRep) return YCN_ERR_INVALID;
    const int Nt = c.p.Nt;
    DevTmp dsl;
    YCN_HIP(hipMalloc(&dsl.p, sizeof(double) * (3 * Nt + 8)));
    double *sl = (double *)dsl.p, *sq = sl + Nt, *cr = sl + 2 * Nt;
    int rc = field_slices(c, replica, sl, sq);
    if (rc) return rc;
    std::vector<double> hs(2 * Nt);
    YCN_HIP(hipMemcpyAsync(hs.data(), sl, sizeof(double) * 2 * Nt, hipMemcpyDeviceToHost, c.stream));
    std::vector<StepState> S(c.nRep);
    YCN_HIP(hipMemcpyAsync(S.data(), c.step_d, sizeof(StepState) * c.nRep, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    ycn_observables o{};
    for (int t = 0; t < Nt; ++t) { o.magnetisation += hs[t]; o.phi2 += hs[Nt + t]; }
    if (slices) std::copy(hs.begin(), hs.begin() + Nt, slices);
    o.eps = S[replica].eps;
    o.langevin_time = S[replica].tL;
    o.clamp_events = S[replica].clamps;
    const double *phi_rep = c.phi[c.cur] + (size_t)replica * c.V;
    if (corr) {
        // point source (main.cu:441-469): delta_{x,0} (1,1,0,0); CG on D D^dag, then D^dag,
        // i.e. D^-1 src (with even-odd preconditioning: the Schur-complement solve)
        std::vector<double> src(8 * (size_t)c.V, 0.0);
        src[0] = 1.0;
        src[2] = 1.0;
        DevTmp aos;
        const size_t nb = sizeof(double) * src.size();
        YCN_HIP(hipMalloc(&aos.p, nb));
        YCN_HIP(hipMemcpyAsync(aos.p, src.data(), nb, hipMemcpyHostToDevice, c.stream));
        if ((rc = aos_to_soa(c, (const double *)aos.p, c.eta, 1))) return rc;
        const int save = c.nNoise;
        c.nNoise = 1;
        SolveStats st;
        if (c.p.evenOdd && eo_available(c)) {
            rc = eo_propagator(c, phi_rep, c.t, &st);
        } else {
            rc = cg_solve(c, phi_rep, 1, YCN_NORMAL, false, 0, 0, &st);
            if (!rc) rc = dirac_apply(c, phi_rep, c.x, c.t, 1, 1);
        }
        c.nNoise = save;
        if (rc) return rc;
        if ((rc = spinor_slices(c, c.t, cr))) return rc;
        YCN_HIP(hipMemcpyAsync(corr, cr, sizeof(double) * Nt, hipMemcpyDeviceToHost, c.stream));
        YCN_HIP(hipStreamSynchronize(c.stream));
    }
    // condensate: fresh drift with measurement noise (main.cu:472-481)
    {
        SolveStats st;
        if ((rc = force_impl(c, step, 0x4D00u, &st))) return rc;
        std::vector<double> d(c.V);
        YCN_HIP(hipMemcpyAsync(d.data(), c.drift + (size_t)replica * c.V, sizeof(double) * c.V, hipMemcpyDeviceToHost,
                               c.stream));
        YCN_HIP(hipStreamSynchronize(c.stream));
        double s = 0;
        for (double v : d) s += v;
        o.trace = (c.p.g != 0.0) ? s / c.p.g : 0.0;
    }
    if (obs) *obs = o;
    return YCN_OK;
}

static int need_cnn(Ctx &c) {
    if (!c.cnnW || !c.rs || c.tile32 < 0)
        return fail(c, YCN_ERR_UNSUPPORTED, "the learned preconditioner needs Nt % 16 == 0 and Nx % 16 == 0");
    return YCN_OK;
}

int ycn_precond_set_weights(ycn_ctx *ctx, const float *w) {
    CHECK_CTX(ctx);
    if (!w) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    int rc = need_cnn(c);
    return rc ? rc : cnn_set_weights(c, w);
}

int ycn_precond_get_weights(ycn_ctx *ctx, float *w) {
    CHECK_CTX(ctx);
    if (!w) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    int rc = need_cnn(c);
    return rc ? rc : cnn_get_weights(c, w);
}

int ycn_precond_apply(ycn_ctx *ctx, const double *r, double *z) {
    CHECK_CTX(ctx);
    if (!r || !z) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    int rc = need_cnn(c);
    if (rc) return rc;
    const size_t n = (size_t)8 * c.V;
    double rr = 0;
    for (size_t i = 0; i < n; ++i) rr += r[i] * r[i];
    DevTmp aos;
    YCN_HIP(hipMalloc(&aos.p, sizeof(double) * n));
    YCN_HIP(hipMemcpyAsync(aos.p, r, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    if ((rc = aos_to_soa(c, (const double *)aos.p, c.t, 1))) return rc;
    if ((rc = spinor_to_f32(c, c.t, c.rs, 1))) return rc;
    CgState S{};
    S.rr = rr;
    YCN_HIP(hipMemcpyAsync(c.cg_d, &S, sizeof(CgState), hipMemcpyHostToDevice, c.stream));
    if ((rc = field_to_f32(c, c.phi[c.cur], c.phis))) return rc;   // the gates read the field
    if ((rc = cnn_gates(c, 1))) return rc;
    rc = cnn_apply(c, 1, c.cg_d, c.rs, c.Aps, c.zs, c.partials);
    if (rc < 0) return rc;
    if ((rc = spinor_to_f64(c, c.zs, c.t, 1))) return rc;
    if ((rc = soa_to_aos(c, c.t, (double *)aos.p, 1))) return rc;
    YCN_HIP(hipMemcpyAsync(z, aos.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipMemsetAsync(c.cg_d, 0, sizeof(CgState), c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

int ycn_precond_loss_grad(ycn_ctx *ctx, const double *phi, const double *r, const double *e, double *loss,
                          float *grad) {
    CHECK_CTX(ctx);
    if (!r || !e) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    int rc = need_cnn(c);
    if (rc) return rc;
    const size_t n = (size_t)8 * c.V;
    if (c.pairCap < 1) {
        if (c.pairR) (void)hipFree(c.pairR);
        if (c.pairX) (void)hipFree(c.pairX);
        YCN_HIP(hipMalloc(&c.pairR, sizeof(double) * n));
        YCN_HIP(hipMalloc(&c.pairX, sizeof(double) * n));
        c.pairCap = 1;
    }
    DevTmp aos;
    const double *phi_dev = nullptr;
    if ((rc = stage_phi(c, phi, &phi_dev))) return rc;
    YCN_HIP(hipMalloc(&aos.p, sizeof(double) * n));
    YCN_HIP(hipMemcpyAsync(aos.p, r, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    if ((rc = aos_to_soa(c, (const double *)aos.p, c.pairR, 1))) return rc;
    YCN_HIP(hipMemcpyAsync(aos.p, e, sizeof(double) * n, hipMemcpyHostToDevice, c.stream));
    if ((rc = aos_to_soa(c, (const double *)aos.p, c.pairX, 1))) return rc;
    if ((rc = field_to_f32(c, phi_dev, c.phis))) return rc;
    double L = 0;
    if ((rc = cnn_train_step(c, c.pairR, c.pairX, c.phis, 0.f, &L, false))) return rc;
    if (loss) *loss = L;
    if (grad && (rc = cnn_get_grad(c, grad))) return rc;
    YCN_HIP(hipMemsetAsync(c.cg_d, 0, sizeof(CgState), c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

int ycn_precond_train(ycn_ctx *ctx, uint64_t step0, int nsteps, int pairsPerSolve, double lr,
                      ycn_train_stats *stats) {
    CHECK_CTX(ctx);
    if (nsteps < 0 || pairsPerSolve < 1 || !(lr > 0)) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    int rc = need_cnn(c);
    if (rc) return rc;
    int used = 0;
    if ((rc = train_preconditioner(c, step0, nsteps, pairsPerSolve, (float)lr, &used))) return rc;
    if (stats) {
        StepState S{};
        YCN_HIP(hipMemcpyAsync(&S, c.step_d, sizeof(StepState), hipMemcpyDeviceToHost, c.stream));
        YCN_HIP(hipStreamSynchronize(c.stream));
        stats->pairs = used;
        stats->loss_first = c.trainLossFirst;
        stats->loss_last = c.trainLossLast;
        stats->eps_last = S.eps;
    }
    return YCN_OK;
}

int ycn_sync(ycn_ctx *ctx) {
    CHECK_CTX(ctx);
    Ctx &c = *ctx;
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

int ycn_profile_begin(ycn_ctx *ctx, const char *kernel) {
    CHECK_CTX(ctx);
    if (!kernel) return YCN_ERR_INVALID;
    Ctx &c = *ctx;
    static const char *names[PK_COUNT] = {"cg_normal_op", "cg_update", "dirac", "noise_cols", "bosonic", "cnn"};
    int k = PK_NONE;
    for (int i = 0; i < PK_COUNT; ++i)
        if (std::strcmp(kernel, names[i]) == 0) k = i;
    if (k == PK_NONE) return fail(c, YCN_ERR_INVALID, std::string("unknown kernel ") + kernel);
    YCN_HIP(hipStreamSynchronize(c.stream));
    if (c.evStart.empty()) {   // the timing events are made at the first profiled region
        constexpr int kEvents = 8192;
        c.evStart.reserve(kEvents);
        c.evStop.reserve(kEvents);
        for (int i = 0; i < kEvents; ++i) {
            hipEvent_t a = nullptr, b = nullptr;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
                if (a) (void)hipEventDestroy(a);
                return fail(c, YCN_ERR_HIP, "hipEventCreate failed");
            }
            c.evStart.push_back(a);
            c.evStop.push_back(b);
        }
        c.evWork.assign(kEvents, -1);
    }
    c.profTarget = k;
    c.evCount = 0;
    c.profBytes = 0;
    c.profLaunches = 0;
    return YCN_OK;
}

int ycn_profile_end(ycn_ctx *ctx, double *total_ms, int64_t *launches, double *bytes_per_launch) {
    CHECK_CTX(ctx);
    Ctx &c = *ctx;
    YCN_HIP(hipStreamSynchronize(c.stream));
    // Single launches: those replayed after every system converged exit at once (a few us
    // against tens for a working launch); those shorter than a quarter of the 90th
    // percentile of the single launches are not counted (a percentile, not the median: in
    // batched solves more than half of the launches can be ones that exit at once).  Chunks: the whole bracket over the launches
    // that did an iteration (launches after a pause count as time, not as launches).
    std::vector<float> d(c.evCount), single;
    for (int i = 0; i < c.evCount; ++i) {
        YCN_HIP(hipEventElapsedTime(&d[i], c.evStart[i], c.evStop[i]));
        if (c.evWork[i] < 0) single.push_back(d[i]);
    }
    std::sort(single.begin(), single.end());
    const float cut = single.empty() ? 0.f : 0.25f * single[(single.size() * 9) / 10];
    double tot = 0;
    int64_t n = 0;
    for (int i = 0; i < c.evCount; ++i) {
        if (c.evWork[i] < 0 && d[i] >= cut) {
            tot += d[i];
            ++n;
        } else if (c.evWork[i] > 0) {
            tot += d[i];
            n += c.evWork[i];
        }
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = n;
    if (bytes_per_launch) *bytes_per_launch = c.profLaunches ? c.profBytes / (double)c.profLaunches : 0.0;
    c.profTarget = PK_NONE;
    return YCN_OK;
}

}  // extern "C"
