// cg.hip -- Wilson-Dirac operator, CG on D^dag D / D D^dag and the stochastic
// fermion force on gfx950 (replaces DiracOP, CGsolver and FermionicDrift,
// code/src/Dirac.cu:23-190, CGsolver.cu:42-177, FermionicDrift.cu:57-110).
//
// All kernels work on `nsys` independent systems at once (grid.y = system): a
// system is one (replica, noise vector) pair, so replicas and several right-hand
// sides share every launch.  One thread owns one lattice site and all four spinor
// components (SoA planes, coalesced 16-byte loads).
//
// Global dot products: every block writes one partial; a finalize kernel (one block
// per system, launched after the producer) sums the partials in block order and
// updates the per-system CgState on the device.  The result is deterministic and
// needs neither atomics on data nor agent-scope release fences inside the
// producers (on the multi-XCD gfx950 such a fence writes back the XCD's L2).
//
// One FP64 CG iteration (standard CG dataflow of CGsolver.cu:107-166):
//   normal op : p' = r + beta p; Ap = D2 D1 p'; <p', Ap>   (fused LDS kernel, or the
//               two-stencil pair cg_a + cg_b when no tile shape divides the lattice)
//   fin_alpha : alpha = rr / <p', Ap>                        (complex, as the reference)
//   cg_c      : x += alpha p', r -= alpha Ap, |r|^2
//   fin_beta  : beta = |r'|^2 / |r|^2, iteration count, stop test
// Algorithmic HBM bytes per site (S = 64 B spinor, P = 8 B field): fused 4S + P,
// cg_c 6S -> 10S + P = 648 B per iteration (two-stencil: 13S + 2P = 848 B).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cg_fused.hpp"
#include "cg_host.hpp"
#include "ycn_device.hpp"
#include "ycn_internal.hpp"

namespace ycn {

constexpr int kNT = 256;        // threads per block for the stencil and BLAS-1 kernels

__device__ __forceinline__ WilsonCoef make_coef(double r, double M, double g) {
    WilsonCoef w;
    w.hp = 0.5 * (r + 1.0);
    w.hm = 0.5 * (r - 1.0);
    w.hr = 0.5 * r;
    w.h = 0.5;
    w.two_r_plus_m = 2.0 * r + M;
    w.g = g;
    return w;
}

struct LatticeArgs {
    int Nt, Nx, V, nNoise;
    double r, M, g;
};

// Full Wilson operator on one site for the four components of `v` (SoA planes).
// Load functor ld(i, c) returns component c at flat site i.
template <bool DAG, typename Load>
__device__ __forceinline__ void wilson_site(const LatticeArgs &L, const WilsonCoef &w, double phi, int s, Load ld,
                                            double2 o[4]) {
    const int t = s / L.Nx, x = s - t * L.Nx;
    const Nbr n = neighbours(t, x, L.Nt, L.Nx);
    const double d = w.two_r_plus_m + w.g * phi;
#pragma unroll
    for (int f = 0; f < 4; f += 2) {
        wilson_flavour<DAG, double2>(w, d, ld(s, f), ld(s, f + 1), ld(n.tf, f), ld(n.tf, f + 1), n.sf, ld(n.tb, f),
                                     ld(n.tb, f + 1), n.sb, ld(n.xf, f), ld(n.xf, f + 1), ld(n.xb, f),
                                     ld(n.xb, f + 1), o[f], o[f + 1]);
    }
}

// ---------------------------------------------------------------------------
// plain Dirac apply: out = D in / D^dag in (DiracOP::applyD, Dirac.cu:23-47)
// ---------------------------------------------------------------------------
template <bool DAG>
__global__ void __launch_bounds__(kNT)
k_dirac(LatticeArgs L, const double *__restrict__ phi, const double2 *__restrict__ in, double2 *__restrict__ out) {
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= L.V) return;
    const int sys = blockIdx.y;
    const size_t V = L.V;
    const double2 *v = in + (size_t)sys * 4 * V;
    const WilsonCoef w = make_coef(L.r, L.M, L.g);
    double2 o[4];
    wilson_site<DAG>(L, w, phi[(size_t)(sys / L.nNoise) * V + s], s, [&](int i, int c) { return v[c * V + i]; }, o);
    double2 *q = out + (size_t)sys * 4 * V;
#pragma unroll
    for (int c = 0; c < 4; ++c) q[c * V + s] = o[c];
}

__global__ void k_set_int(int *p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ---------------------------------------------------------------------------
// finalize kernels: one block per system, partials summed in block order
// ---------------------------------------------------------------------------
// Sum of nb block partials by one block, in a fixed order (thread t adds entries t,
// t + kNT, ... in sequence, then the block tree).  Loads are issued in batches of kFinU
// before they are added, so the partials' memory latency is paid nb / (kNT kFinU) times
// rather than nb / kNT times; the order of the additions is unchanged.
constexpr int kFinU = 8;
__device__ __forceinline__ double2 fin_sum(const double2 *part, int nb, double2 *lds) {
    double2 acc = make_double2(0.0, 0.0);
    int i = threadIdx.x;
    for (; i + (kFinU - 1) * kNT < nb; i += kFinU * kNT) {
        double2 v[kFinU];
#pragma unroll
        for (int u = 0; u < kFinU; ++u) v[u] = part[i + u * kNT];
#pragma unroll
        for (int u = 0; u < kFinU; ++u) acc = cadd(acc, v[u]);
    }
    for (; i < nb; i += kNT) acc = cadd(acc, part[i]);
    return block_sum2<kNT>(acc, lds);
}

// after setup: rr = |b|^2, stopping threshold, run flag
__global__ void __launch_bounds__(kNT)
k_fin_setup(const double2 *__restrict__ part, int nb, CgState *__restrict__ st, int *__restrict__ active, double tol,
            int mode, int maxIter, int maxRestarts) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.x;
    const double2 tot = fin_sum(part + (size_t)sys * nb, nb, lds);
    if (threadIdx.x) return;
    CgState &S = st[sys];
    S.rr = tot.x;
    S.rr_prev = tot.x;
    S.rrRel = tot.x;
    S.bnorm2 = tot.x;
    S.tol = tol;
    S.mode = mode;
    S.maxIter = maxIter;
    S.maxRestarts = maxRestarts;
    S.stop = (mode == YCN_TOL_ABS_RECURSIVE) ? tol : tol * sqrt(tot.x);
    S.beta = 0.0;
    S.alpha = make_double2(0.0, 0.0);
    S.iter = 0;
    S.restarts = 0;
    S.true_rel = 0.0;
    S.rz = 0.0;
    S.true_prev = 1.0;
    S.precOff = 0;
    S.restartNext = 0;
    const bool run = (0 < maxIter) && (sqrt(tot.x) > S.stop);
    S.done = run ? 0 : 1;
    S.final_ = run ? 0 : 1;
    S.converged = run ? 0 : 1;
    if (!run) atomicSub(active, 1);
}

// alpha = rr / <p', Ap> (complex division, CGsolver.cu:137)
__global__ void __launch_bounds__(kNT) k_fin_alpha(const double2 *__restrict__ part, int nb, CgState *__restrict__ st) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.x;
    const int skip = st[sys].done;   // loaded together with the partials
    const double2 dot = fin_sum(part + (size_t)sys * nb, nb, lds);
    if (skip || threadIdx.x) return;
    const double rr = st[sys].rr;
    const double den = dot.x * dot.x + dot.y * dot.y;
    st[sys].alpha = make_double2(rr * dot.x / den, -rr * dot.y / den);
}

// beta = |r'|^2 / |r|^2 and the reference stopping rule (CGsolver.cu:107,155-158).
// MIXED: the FP32 loop pauses at the stop or the cap and waits for a reliable update.
// It also pauses once the recursive |r| has fallen by kReliableDrop since the last anchor.
// EVEN: a system pauses only after an even number of iterations (or at the cap): chunks
// have even length and p ping-pongs between two buffers, so the last direction is then
// in p[0], where the iteration that resumes it after a reliable update or a tightened
// threshold reads it.
template <bool MIXED, bool EVEN = MIXED>
__global__ void __launch_bounds__(kNT)
k_fin_beta(const double2 *__restrict__ part, int nb, CgState *__restrict__ st, int *__restrict__ active) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.x;
    const int skip = st[sys].done;   // loaded together with the partials
    const double2 tot = fin_sum(part + (size_t)sys * nb, nb, lds);
    if (skip || threadIdx.x) return;
    CgState &S = st[sys];
    const double rrn = tot.x;
    S.rr_prev = S.rr;
    S.beta = rrn / S.rr;
    S.rr = rrn;
    S.iter += 1;
    const bool pause = !(sqrt(rrn) > S.stop) || (MIXED && rrn < kReliableDrop * kReliableDrop * S.rrRel);
    const bool run = (S.iter < S.maxIter) && (!pause || (EVEN && (S.iter & 1)));
    if (!run) {
        S.done = 1;
        if (!MIXED) S.converged = sqrt(rrn) <= S.stop ? 1 : 0;
        atomicSub(active, 1);
    }
}

// FP64 true residual |b - A x| / |b|; YCN_TOL_REL_TRUE restarts from x while above tol
__global__ void __launch_bounds__(kNT)
k_fin_true(const double2 *__restrict__ part, int nb, CgState *__restrict__ st, int *__restrict__ active) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.x;
    const double2 tot = fin_sum(part + (size_t)sys * nb, nb, lds);
    if (threadIdx.x) return;
    CgState &S = st[sys];
    const double bn = sqrt(S.bnorm2);
    S.true_rel = bn > 0.0 ? sqrt(tot.x) / bn : sqrt(tot.x);
    if (S.mode != YCN_TOL_REL_TRUE) return;
    if (S.true_rel > S.tol && S.iter < S.maxIter && S.restarts < S.maxRestarts) {
        S.restarts += 1;   // r already holds b - A x; p' = r on the next iteration
        S.rr = tot.x;
        S.beta = 0.0;
        S.done = 0;
        S.converged = 0;
        atomicAdd(active, 1);
    } else {
        S.converged = S.true_rel <= S.tol ? 1 : 0;
    }
}

// mixed precision reliable update decision: converged on the FP64 true residual, or
// resume the FP32 loop keeping the search direction.  Mixed CG: beta = |r_true|^2 /
// |r_prev|^2.  Flexible CG: beta follows from z = CNN(r_true) (k_fin_fcg); the network
// is switched off, and the direction restarted, when a reliable interval did not
// reduce the true residual.
// FCG guard: refRate (< 0) is the decades of true-residual reduction per iteration of
// the context's last plain solve (0: none known).  From max(100, 0.3 n_plain) iterations
// on, a flexible solve that has reduced its true residual at less than half that rate
// switches the network off: a preconditioner that slows CG down costs at most that
// stretch of iterations.
template <bool FCG>
__global__ void __launch_bounds__(kNT)
k_fin_mp_true(const double2 *__restrict__ part, int nb, CgState *__restrict__ st, int *__restrict__ active,
              double refRate) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.x;
    const int skip = st[sys].final_;   // loaded together with the partials
    const double2 tot = fin_sum(part + (size_t)sys * nb, nb, lds);
    if (skip || threadIdx.x) return;
    CgState &S = st[sys];
    const double rrt = tot.x;
    const double bn = sqrt(S.bnorm2);
    S.true_rel = bn > 0.0 ? sqrt(rrt) / bn : sqrt(rrt);
    S.restarts += 1;
    const bool conv = (S.mode == YCN_TOL_REL_TRUE) ? (S.true_rel <= S.tol) : (sqrt(rrt) <= S.stop);
    if (conv || S.iter >= S.maxIter) {
        S.converged = conv ? 1 : 0;
        S.final_ = 1;
        if (!S.done) { S.done = 1; atomicSub(active, 1); }
        return;
    }
    if (FCG) {
        // the direction is kept (beta follows from the new z in k_fin_fcg); a reliable
        // interval that did not reduce the true residual switches the network off and
        // restarts the direction (plain CG from here on)
        bool slow = false;
        if (refRate < 0.0 && S.true_rel > 0.0) {
            const double nPlain = log10(S.tol) / refRate;
            if (S.iter >= fmax(100.0, 0.3 * nPlain)) slow = log10(S.true_rel) / S.iter > 0.5 * refRate;
        }
        if ((!(S.true_rel < S.true_prev) || slow) && !S.precOff) {
            S.precOff = 1;
            S.restartNext = 1;
        }
        S.true_prev = S.true_rel;
    } else {
        S.beta = rrt / S.rr_prev;
    }
    S.rr = rrt;
    S.rrRel = rrt;
    if (S.done) { S.done = 0; atomicAdd(active, 1); }
}

// flexible CG: alpha = <r, z> / Re <p', Ap> (real, the system seen as real 8V-vectors)
__global__ void __launch_bounds__(kNT) k_fin_alpha_fcg(const double2 *__restrict__ part, int nb, CgState *__restrict__ st) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.x;
    const int skip = st[sys].done;   // loaded together with the partials
    const double2 dot = fin_sum(part + (size_t)sys * nb, nb, lds);
    if (skip || threadIdx.x) return;
    CgState &S = st[sys];
    if (!(dot.x > 0.0) || !isfinite(S.rz) || !(S.rz > 0.0)) {
        // the preconditioned direction is not a descent direction: no step, and plain
        // CG from the next reliable update on
        S.alpha = make_double2(0.0, 0.0);
        S.precOff = 1;
        return;
    }
    S.alpha = make_double2(S.rz / dot.x, 0.0);
}

// flexible CG after z = CNN(r'): partials hold (<z, r'>, <z, Ap>).  Polak-Ribiere form
// (Notay, SIAM J. Sci. Comput. 22 (2000)): beta = <z, r' - r> / <z_old, r> with
// r = r' + alpha Ap, i.e. beta = -alpha <z, Ap> / rz_old.  MODE 0: inside the loop;
// MODE 1: after a reliable update replaced r' by the true residual (rz_old is then the
// value before the last iteration, rz_prev); MODE 2: restart (beta = 0).
template <int MODE>
__global__ void __launch_bounds__(kNT) k_fin_fcg(const double2 *__restrict__ part, int nb, CgState *__restrict__ st) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.x;
    const int skip = st[sys].done;   // loaded together with the partials
    const double2 tot = fin_sum(part + (size_t)sys * nb, nb, lds);
    if (skip || threadIdx.x) return;
    CgState &S = st[sys];
    if (!isfinite(tot.x) || !isfinite(tot.y)) {
        S.precOff = 1;
        S.beta = 0.0;
        return;
    }
    const double rzOld = (MODE == 1) ? S.rz_prev : S.rz;
    S.beta = (MODE == 2 || S.restartNext || !(rzOld > 0.0)) ? 0.0 : -S.alpha.x * tot.y / rzOld;
    S.restartNext = 0;
    if (MODE != 1) S.rz_prev = S.rz;
    S.rz = tot.x;
}

// ---------------------------------------------------------------------------
// FP64 CG kernels
// ---------------------------------------------------------------------------
// setup: eta (generated or given) -> b; x = 0; r = b; p0 = p1 = b; partial |b|^2
template <bool GEN>
__global__ void __launch_bounds__(kNT)
k_cg_setup(LatticeArgs L, const uint64_t *__restrict__ seeds, uint32_t streamBase, uint64_t step,
           double2 *__restrict__ eta, double2 *__restrict__ x, double2 *__restrict__ r, double2 *__restrict__ p0,
           double2 *__restrict__ p1, double2 *__restrict__ part) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.y;
    const size_t V = L.V;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t off = (size_t)sys * 4 * V;
    double2 acc = make_double2(0.0, 0.0);
    if (s < L.V) {
        double2 b[4];
        if (GEN) {
            eta_site(seeds, streamBase, step, sys, L.nNoise, s, b);
#pragma unroll
            for (int c = 0; c < 4; ++c) eta[off + c * V + s] = b[c];
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = eta[off + c * V + s];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t i = off + c * V + s;
            x[i] = make_double2(0.0, 0.0);
            r[i] = b[c];
            p0[i] = b[c];
            p1[i] = b[c];
            acc.x += b[c].x * b[c].x + b[c].y * b[c].y;
        }
    }
    acc = block_sum2<kNT>(acc, lds);
    if (threadIdx.x == 0) part[(size_t)sys * gridDim.x + blockIdx.x] = acc;
}

// cg_a: p' = r + beta p_old (written at own site), t = D1 p'
template <bool DAG1>
__global__ void __launch_bounds__(kNT)
k_cg_a(LatticeArgs L, const double *__restrict__ phi, const CgState *__restrict__ st, const double2 *__restrict__ r,
       const double2 *__restrict__ pold, double2 *__restrict__ pnew, double2 *__restrict__ t) {
    const int sys = blockIdx.y;
    if (st[sys].done) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= L.V) return;
    const double beta = st[sys].beta;
    const size_t V = L.V, off = (size_t)sys * 4 * V;
    const double2 *rr = r + off, *po = pold + off;
    auto ld = [&](int i, int c) {
        const double2 a = rr[c * V + i], b = po[c * V + i];
        return make_double2(a.x + beta * b.x, a.y + beta * b.y);
    };
#pragma unroll
    for (int c = 0; c < 4; ++c) pnew[off + c * V + s] = ld(s, c);
    const WilsonCoef w = make_coef(L.r, L.M, L.g);
    double2 o[4];
    wilson_site<DAG1>(L, w, phi[(size_t)(sys / L.nNoise) * V + s], s, ld, o);
#pragma unroll
    for (int c = 0; c < 4; ++c) t[off + c * V + s] = o[c];
}

// FP32 two-stencil pair for lattices that no fused FP32 tile divides (mixed precision on
// any lattice).  cg_a32: p' = r + beta p_old in FP32 (written at the own site; the hop
// reads the same FP32 values at the neighbours), t = D1 p'; cg_b32: Ap = D2 t, partial
// <p', Ap>.  The hop arithmetic runs in FP64 on the FP32 vectors and FP32 field.
template <bool DAG1>
__global__ void __launch_bounds__(kNT)
k_cg_a32(LatticeArgs L, const float *__restrict__ phi, const CgState *__restrict__ st, const float2 *__restrict__ r,
         const float2 *__restrict__ pold, float2 *__restrict__ pnew, float2 *__restrict__ t) {
    const int sys = blockIdx.y;
    if (st[sys].done) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= L.V) return;
    const float beta = (float)st[sys].beta;
    const size_t V = L.V, off = (size_t)sys * 4 * V;
    const float2 *rr = r + off, *po = pold + off;
    auto pf = [&](int i, int c) {
        const float2 a = rr[c * V + i], b = po[c * V + i];
        return make_float2(a.x + beta * b.x, a.y + beta * b.y);
    };
    auto ld = [&](int i, int c) {
        const float2 v = pf(i, c);
        return make_double2((double)v.x, (double)v.y);
    };
#pragma unroll
    for (int c = 0; c < 4; ++c) pnew[off + c * V + s] = pf(s, c);
    const WilsonCoef w = make_coef(L.r, L.M, L.g);
    double2 o[4];
    wilson_site<DAG1>(L, w, (double)phi[(size_t)(sys / L.nNoise) * V + s], s, ld, o);
#pragma unroll
    for (int c = 0; c < 4; ++c) t[off + c * V + s] = make_float2((float)o[c].x, (float)o[c].y);
}

template <bool DAG2>
__global__ void __launch_bounds__(kNT)
k_cg_b32(LatticeArgs L, const float *__restrict__ phi, const CgState *__restrict__ st, const float2 *__restrict__ t,
         const float2 *__restrict__ p, float2 *__restrict__ Ap, double2 *__restrict__ part) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.y;
    if (st[sys].done) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t V = L.V, off = (size_t)sys * 4 * V;
    double2 acc = make_double2(0.0, 0.0);
    if (s < L.V) {
        const float2 *tv = t + off;
        const WilsonCoef w = make_coef(L.r, L.M, L.g);
        double2 o[4];
        wilson_site<DAG2>(L, w, (double)phi[(size_t)(sys / L.nNoise) * V + s], s,
                          [&](int i, int c) {
                              const float2 v = tv[c * V + i];
                              return make_double2((double)v.x, (double)v.y);
                          },
                          o);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float2 a = make_float2((float)o[c].x, (float)o[c].y);
            Ap[off + c * V + s] = a;
            const float2 pv = p[off + c * V + s];
            acc = cadd(acc, cdotc(make_double2((double)pv.x, (double)pv.y), make_double2((double)a.x, (double)a.y)));
        }
    }
    acc = block_sum2<kNT>(acc, lds);
    if (threadIdx.x == 0) part[(size_t)sys * gridDim.x + blockIdx.x] = acc;
}

// cg_b: Ap = D2 t; partial <p', Ap>// cg_b: Ap = D2 t; partial <p', Ap>
template <bool DAG2>
__global__ void __launch_bounds__(kNT)
k_cg_b(LatticeArgs L, const double *__restrict__ phi, const CgState *__restrict__ st, const double2 *__restrict__ t,
       const double2 *__restrict__ p, double2 *__restrict__ Ap, double2 *__restrict__ part) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.y;
    if (st[sys].done) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t V = L.V, off = (size_t)sys * 4 * V;
    double2 acc = make_double2(0.0, 0.0);
    if (s < L.V) {
        const double2 *tv = t + off;
        const WilsonCoef w = make_coef(L.r, L.M, L.g);
        double2 o[4];
        wilson_site<DAG2>(L, w, phi[(size_t)(sys / L.nNoise) * V + s], s,
                          [&](int i, int c) { return tv[c * V + i]; }, o);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            Ap[off + c * V + s] = o[c];
            acc = cadd(acc, cdotc(p[off + c * V + s], o[c]));
        }
    }
    acc = block_sum2<kNT>(acc, lds);
    if (threadIdx.x == 0) part[(size_t)sys * gridDim.x + blockIdx.x] = acc;
}

// cg_c: x += alpha p'; r -= alpha Ap; partial |r|^2.  All loads are issued before
// any store so each thread keeps 16 loads in flight.
__global__ void __launch_bounds__(kNT)
k_cg_c(int V_, const CgState *__restrict__ st, double2 *__restrict__ x, double2 *__restrict__ r,
       const double2 *__restrict__ p, const double2 *__restrict__ Ap, double2 *__restrict__ part) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.y;
    if (st[sys].done) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t V = V_, off = (size_t)sys * 4 * V;
    const double2 al = st[sys].alpha;
    double2 acc = make_double2(0.0, 0.0);
    if (s < V_) {
        double2 xv[4], rv[4], pv[4], av[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t i = off + c * V + s;
            xv[c] = x[i];
            rv[c] = r[i];
            pv[c] = p[i];
            av[c] = Ap[i];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t i = off + c * V + s;
            x[i] = cadd(xv[c], cmul(al, pv[c]));
            const double2 rn = csub(rv[c], cmul(al, av[c]));
            r[i] = rn;
            acc.x += rn.x * rn.x + rn.y * rn.y;
        }
    }
    acc = block_sum2<kNT>(acc, lds);
    if (threadIdx.x == 0) part[(size_t)sys * gridDim.x + blockIdx.x] = acc;
}

// true residual, part 1: t = D1 x
template <bool DAG1>
__global__ void __launch_bounds__(kNT)
k_true_a(LatticeArgs L, const double *__restrict__ phi, const double2 *__restrict__ x, double2 *__restrict__ t) {
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= L.V) return;
    const int sys = blockIdx.y;
    const size_t V = L.V, off = (size_t)sys * 4 * V;
    const double2 *xv = x + off;
    const WilsonCoef w = make_coef(L.r, L.M, L.g);
    double2 o[4];
    wilson_site<DAG1>(L, w, phi[(size_t)(sys / L.nNoise) * V + s], s, [&](int i, int c) { return xv[c * V + i]; }, o);
#pragma unroll
    for (int c = 0; c < 4; ++c) t[off + c * V + s] = o[c];
}

// true residual, part 2: res = b - D2 t -> r (FP64) or, for the mixed solver, rs
// (FP32); partial |res|^2.  The mixed solver skips systems that are already final.
template <bool DAG2, bool MIXED>
__global__ void __launch_bounds__(kNT)
k_true_b(LatticeArgs L, const double *__restrict__ phi, const CgState *__restrict__ st, const double2 *__restrict__ b,
         const double2 *__restrict__ t, double2 *__restrict__ r, float2 *__restrict__ rs, double2 *__restrict__ part) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.y;
    if (MIXED && st[sys].final_) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t V = L.V, off = (size_t)sys * 4 * V;
    double2 acc = make_double2(0.0, 0.0);
    if (s < L.V) {
        const double2 *tv = t + off;
        const WilsonCoef w = make_coef(L.r, L.M, L.g);
        double2 o[4];
        wilson_site<DAG2>(L, w, phi[(size_t)(sys / L.nNoise) * V + s], s,
                          [&](int i, int c) { return tv[c * V + i]; }, o);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double2 res = csub(b[off + c * V + s], o[c]);
            if (MIXED)
                rs[off + c * V + s] = make_float2((float)res.x, (float)res.y);
            else
                r[off + c * V + s] = res;
            acc.x += res.x * res.x + res.y * res.y;
        }
    }
    acc = block_sum2<kNT>(acc, lds);
    if (threadIdx.x == 0) part[(size_t)sys * gridDim.x + blockIdx.x] = acc;
}

// drift_x = fac/nNoise * sum_j Re sum_k conj((D y_j)_xk) eta_j,xk
// (FermionicDrift.cu:69 applyD Normal + computeDrift :83-100)
__global__ void __launch_bounds__(kNT)
k_drift(LatticeArgs L, const double *__restrict__ phi, const double2 *__restrict__ y, const double2 *__restrict__ eta,
        double *__restrict__ drift, double fac) {
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= L.V) return;
    const int rep = blockIdx.y;
    const size_t V = L.V;
    const WilsonCoef w = make_coef(L.r, L.M, L.g);
    const double ph = phi[(size_t)rep * V + s];
    double acc = 0.0;
    for (int j = 0; j < L.nNoise; ++j) {
        const size_t off = ((size_t)rep * L.nNoise + j) * 4 * V;
        const double2 *yv = y + off;
        double2 o[4];
        wilson_site<false>(L, w, ph, s, [&](int i, int c) { return yv[c * V + i]; }, o);
        double2 sm = cdotc(o[0], eta[off + s]);
        sm = cadd(sm, cdotc(o[1], eta[off + V + s]));
        sm = cadd(sm, cdotc(o[2], eta[off + 2 * V + s]));
        sm = cadd(sm, cdotc(o[3], eta[off + 3 * V + s]));
        acc += sm.x;
    }
    drift[(size_t)rep * V + s] = (L.nNoise == 1) ? fac * acc : fac * acc / (double)L.nNoise;
}

// corr[t] = sum_x sum_k Re psi (gpuTimeSlices_spinors, reductions.cu:176-191);
// one block per time slice, fixed-order block reduction.
__global__ void __launch_bounds__(kNT)
k_spinor_slices(const double2 *__restrict__ psi, int Nt, int Nx, double *__restrict__ out) {
    __shared__ double2 lds[kNT / kWave];
    const int t = blockIdx.x;
    const size_t V = (size_t)Nt * Nx;
    double2 acc = make_double2(0.0, 0.0);
    for (int x = threadIdx.x; x < Nx; x += kNT) {
        const size_t s = (size_t)t * Nx + x;
        acc.x += psi[s].x + psi[V + s].x + psi[2 * V + s].x + psi[3 * V + s].x;
    }
    acc = block_sum2<kNT>(acc, lds);
    if (threadIdx.x == 0) out[t] = acc.x;
}

// ---------------------------------------------------------------------------
// mixed-precision kernels
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kNT)
k_to_f32(const double *__restrict__ a, float *__restrict__ b, int64_t n, const int *__restrict__ halt,
         double *__restrict__ copy) {
    if (halt && *halt) return;
    const int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x;
    if (i >= n) return;
    const double v = a[i];
    b[i] = (float)v;
    if (copy) copy[i] = v;
}

template <bool GEN>
__global__ void __launch_bounds__(kNT)
k_mp_setup(LatticeArgs L, const uint64_t *__restrict__ seeds, uint32_t streamBase, uint64_t step,
           double2 *__restrict__ eta, double2 *__restrict__ x, float2 *__restrict__ xs, float2 *__restrict__ rs,
           float2 *__restrict__ p0, float2 *__restrict__ p1, double2 *__restrict__ part) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.y;
    const size_t V = L.V;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t off = (size_t)sys * 4 * V;
    double2 acc = make_double2(0.0, 0.0);
    if (s < L.V) {
        double2 b[4];
        if (GEN) {
            eta_site(seeds, streamBase, step, sys, L.nNoise, s, b);
#pragma unroll
            for (int c = 0; c < 4; ++c) eta[off + c * V + s] = b[c];
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = eta[off + c * V + s];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t i = off + c * V + s;
            const float2 bs = make_float2((float)b[c].x, (float)b[c].y);
            x[i] = make_double2(0.0, 0.0);
            xs[i] = make_float2(0.f, 0.f);
            rs[i] = bs;
            p0[i] = bs;
            p1[i] = bs;
            acc.x += b[c].x * b[c].x + b[c].y * b[c].y;
        }
    }
    acc = block_sum2<kNT>(acc, lds);
    if (threadIdx.x == 0) part[(size_t)sys * gridDim.x + blockIdx.x] = acc;
}

// FP32 update: xs += alpha p', rs -= alpha Ap; partial |rs|^2 accumulated in FP64
__global__ void __launch_bounds__(kNT)
k_mp_update(int V_, const CgState *__restrict__ st, float2 *__restrict__ xs, float2 *__restrict__ rs,
            const float2 *__restrict__ p, const float2 *__restrict__ Ap, double2 *__restrict__ part) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.y;
    if (st[sys].done) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t V = V_, off = (size_t)sys * 4 * V;
    const double2 ad = st[sys].alpha;
    const float ax = (float)ad.x, ay = (float)ad.y;
    double2 acc = make_double2(0.0, 0.0);
    if (s < V_) {
        float2 xv[4], rv[4], pv[4], av[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t i = off + c * V + s;
            xv[c] = xs[i];
            rv[c] = rs[i];
            pv[c] = p[i];
            av[c] = Ap[i];
        }
        float prr = 0.f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t i = off + c * V + s;
            float2 xn, rn;
            xn.x = xv[c].x + (ax * pv[c].x - ay * pv[c].y);
            xn.y = xv[c].y + (ax * pv[c].y + ay * pv[c].x);
            rn.x = rv[c].x - (ax * av[c].x - ay * av[c].y);
            rn.y = rv[c].y - (ax * av[c].y + ay * av[c].x);
            xs[i] = xn;
            rs[i] = rn;
            prr += rn.x * rn.x + rn.y * rn.y;
        }
        acc.x = (double)prr;
    }
    acc = block_sum2<kNT>(acc, lds);
    if (threadIdx.x == 0) part[(size_t)sys * gridDim.x + blockIdx.x] = acc;
}

// reliable update, part 1: x += xs (FP64), xs = 0
__global__ void __launch_bounds__(kNT)
k_mp_accum(int V_, const CgState *__restrict__ st, double2 *__restrict__ x, float2 *__restrict__ xs) {
    const int sys = blockIdx.y;
    if (st[sys].final_) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= V_) return;
    const size_t V = V_, off = (size_t)sys * 4 * V;
    float2 d[4];
    double2 xv[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        d[c] = xs[off + c * V + s];
        xv[c] = x[off + c * V + s];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        x[off + c * V + s] = make_double2(xv[c].x + (double)d[c].x, xv[c].y + (double)d[c].y);
        xs[off + c * V + s] = make_float2(0.f, 0.f);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static LatticeArgs largs(const Ctx &c) {
    LatticeArgs L;
    L.Nt = c.p.Nt;
    L.Nx = c.p.Nx;
    L.V = c.V;
    L.nNoise = c.nNoise;
    L.r = c.p.r;
    L.M = c.p.M;
    L.g = c.p.g;
    return L;
}

static dim3 site_grid(const Ctx &c, int nsys) { return dim3((c.V + kNT - 1) / kNT, nsys); }
static int site_blocks(const Ctx &c) { return (c.V + kNT - 1) / kNT; }

int cg_alloc(Ctx &c) {
    const size_t n = (size_t)c.nSys * 4 * c.V;
    // partial slots: two per block of the smallest (64-site) tiles, which also covers one
    // full-lattice plus two half-lattice site grids of 256-thread blocks (even-odd setup)
    c.maxBlocks = (c.V + 31) / 32 + 4;
    double2 **bufs[] = {&c.eta, &c.x, &c.r, &c.pv[0], &c.pv[1], &c.t, &c.Ap};
    for (double2 **b : bufs) {
        YCN_HIP(hipMalloc(b, sizeof(double2) * n));
        YCN_HIP(hipMemset(*b, 0, sizeof(double2) * n));
    }
    YCN_HIP(hipMalloc(&c.cg_d, sizeof(CgState) * c.nSys));
    YCN_HIP(hipMemset(c.cg_d, 0, sizeof(CgState) * c.nSys));
    YCN_HIP(hipMalloc(&c.partials, sizeof(double2) * (size_t)std::max(c.nSys, c.nRep) * c.maxBlocks));
    YCN_HIP(hipMalloc(&c.active, sizeof(int)));
    YCN_HIP(hipHostMalloc((void **)&c.active_h, sizeof(int), hipHostMallocDefault));
    YCN_HIP(hipHostMalloc((void **)&c.cg_h, sizeof(CgState) * c.nSys, hipHostMallocDefault));
    YCN_HIP(hipEventCreateWithFlags(&c.pollEv, hipEventDisableTiming));
    return YCN_OK;
}

void cg_free(Ctx &c) {
    for (const Ctx::GraphEntry &e : c.graphs) (void)hipGraphExecDestroy(e.exec);
    c.graphs.clear();
    for (const Ctx::StepGraph &e : c.stepGraphs) (void)hipGraphExecDestroy(e.exec);
    c.stepGraphs.clear();
    if (c.ctl_d) (void)hipFree(c.ctl_d);
    if (c.ctl_h) (void)hipHostFree(c.ctl_h);
    if (c.phiSolve) (void)hipFree(c.phiSolve);
    c.ctl_d = c.ctl_h = nullptr;
    c.phiSolve = nullptr;
    double2 **bufs[] = {&c.eta, &c.x, &c.r, &c.pv[0], &c.pv[1], &c.t, &c.Ap};
    for (double2 **b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    if (c.cg_d) (void)hipFree(c.cg_d);
    if (c.partials) (void)hipFree(c.partials);
    if (c.active) (void)hipFree(c.active);
    if (c.active_h) (void)hipHostFree(c.active_h);
    if (c.cg_h) (void)hipHostFree(c.cg_h);
    c.cg_h = nullptr;
    if (c.pollEv) (void)hipEventDestroy(c.pollEv);
    c.pollEv = nullptr;
    c.cg_d = nullptr;
    c.partials = nullptr;
    c.active = nullptr;
    c.active_h = nullptr;
}

int set_int(Ctx &c, int *ptr, int v) {
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, c.stream, ptr, v);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int dirac_apply(Ctx &c, const double *phi_dev, const double2 *in, double2 *out, int nsys, int dagger) {
    const LatticeArgs L = largs(c);
    prof_before(c, PK_DIRAC, (double)nsys * c.V * (2 * 64 + 8));
    if (dagger)
        hipLaunchKernelGGL(k_dirac<true>, site_grid(c, nsys), dim3(kNT), 0, c.stream, L, phi_dev, in, out);
    else
        hipLaunchKernelGGL(k_dirac<false>, site_grid(c, nsys), dim3(kNT), 0, c.stream, L, phi_dev, in, out);
    prof_after(c, PK_DIRAC);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

// p' = r + beta p_old, Ap = A p' and the partials of <p', Ap>: the fused LDS kernel
// with tile `tile` (>= 0) or the two-stencil pair cg_a + cg_b (tile = -1).  Returns
// the number of partials per system.
static int launch_normal_op(Ctx &c, const LatticeArgs &L, const double *phi, int nsys, int mtype, int parity,
                            int tile) {
    const double2 *pold = c.pv[parity];
    double2 *pnew = c.pv[parity ^ 1];
    const bool dag1 = (mtype != YCN_DAGGER);
    if (tile >= 0) {
        // compulsory traffic: read r, p_old, phi; write p', Ap  (4S + P per site)
        prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (4 * 64 + 8));
        launch_fused<double>(c, tile, nsys, dag1, phi, c.cg_d, (const double2 *)c.r, pold, pnew, c.Ap, c.partials);
        prof_after(c, PK_NORMAL_OP);
        return fused_args(c, kTiles64[tile].TT, kTiles64[tile].TX).tilesPerSys;
    }
    const dim3 g = site_grid(c, nsys);
    // cg_a + cg_b: read r, p_old, phi, write p', t; read t, p', phi, write Ap  (7S + 2P)
    prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (7 * 64 + 16));
    if (!dag1) {
        hipLaunchKernelGGL(k_cg_a<false>, g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                           (const double2 *)c.r, pold, pnew, c.t);
        hipLaunchKernelGGL(k_cg_b<true>, g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                           (const double2 *)c.t, (const double2 *)pnew, c.Ap, c.partials);
    } else {
        hipLaunchKernelGGL(k_cg_a<true>, g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                           (const double2 *)c.r, pold, pnew, c.t);
        hipLaunchKernelGGL(k_cg_b<false>, g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                           (const double2 *)c.t, (const double2 *)pnew, c.Ap, c.partials);
    }
    prof_after(c, PK_NORMAL_OP);
    return site_blocks(c);
}

static void launch_iteration(Ctx &c, const LatticeArgs &L, const double *phi, int nsys, int mtype, int parity) {
    const int nb = launch_normal_op(c, L, phi, nsys, mtype, parity, c.tile64);
    hipLaunchKernelGGL(k_fin_alpha, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb, c.cg_d);
    prof_before(c, PK_CG_UPDATE, (double)nsys * c.V * 6 * 64);
    hipLaunchKernelGGL(k_cg_c, site_grid(c, nsys), dim3(kNT), 0, c.stream, c.V, (const CgState *)c.cg_d, c.x, c.r,
                       (const double2 *)c.pv[parity ^ 1], (const double2 *)c.Ap, c.partials);
    prof_after(c, PK_CG_UPDATE);
    hipLaunchKernelGGL(k_fin_beta<false>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials,
                       site_blocks(c), c.cg_d, c.active);
}

void launch_fin_alpha(Ctx &c, int nsys, int nb) {
    hipLaunchKernelGGL(k_fin_alpha, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb, c.cg_d);
}

void launch_fin_beta(Ctx &c, int nsys, int nb, bool mixed, bool even) {
    if (mixed)
        hipLaunchKernelGGL(k_fin_beta<true>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb,
                           c.cg_d, c.active);
    else if (even)
        hipLaunchKernelGGL((k_fin_beta<false, true>), dim3(nsys), dim3(kNT), 0, c.stream,
                           (const double2 *)c.partials, nb, c.cg_d, c.active);
    else
        hipLaunchKernelGGL(k_fin_beta<false>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb,
                           c.cg_d, c.active);
}

int launch_cg_update(Ctx &c, int nsys, int n, const double2 *p) {
    const dim3 g((n + kNT - 1) / kNT, nsys);
    prof_before(c, PK_CG_UPDATE, (double)nsys * n * 6 * 64);
    hipLaunchKernelGGL(k_cg_c, g, dim3(kNT), 0, c.stream, n, (const CgState *)c.cg_d, c.x, c.r, p,
                       (const double2 *)c.Ap, c.partials);
    prof_after(c, PK_CG_UPDATE);
    return (int)g.x;
}

int launch_mp_update(Ctx &c, int nsys, int n, const float2 *p) {
    const dim3 g((n + kNT - 1) / kNT, nsys);
    prof_before(c, PK_CG_UPDATE, (double)nsys * n * 6 * 32);
    hipLaunchKernelGGL(k_mp_update, g, dim3(kNT), 0, c.stream, n, (const CgState *)c.cg_d, c.xs, c.rs, p,
                       (const float2 *)c.Aps, c.partials);
    prof_after(c, PK_CG_UPDATE);
    return (int)g.x;
}

int launch_mp_accum(Ctx &c, int nsys, int n) {
    hipLaunchKernelGGL(k_mp_accum, dim3((n + kNT - 1) / kNT, nsys), dim3(kNT), 0, c.stream, n,
                       (const CgState *)c.cg_d, c.x, c.xs);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int to_f32_n(Ctx &c, const double *a, float *b, int64_t n, const int *halt, double *copy) {
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + kNT - 1) / kNT)), dim3(kNT), 0, c.stream, a, b, n, halt, copy);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

// Wait for the stream by spinning on an event: a solve polls a few times per Langevin
// step, and each blocking wait leaves the GPU idle until the host thread is scheduled
// again.
int spin_sync(Ctx &c) {
    ++c.hostSyncs;
    YCN_HIP(hipEventRecord(c.pollEv, c.stream));
    hipError_t e;
    while ((e = hipEventQuery(c.pollEv)) == hipErrorNotReady) {
    }
    YCN_HIP(e);
    return YCN_OK;
}

int poll_active(Ctx &c, int &active) {
    YCN_HIP(hipMemcpyAsync(c.active_h, c.active, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    int rc = spin_sync(c);
    if (rc) return rc;
    active = *c.active_h;
    return YCN_OK;
}

// active count and the largest iteration count of the systems not yet final (the
// states stay in the pinned mirror c.cg_h)
int poll_iters(Ctx &c, int nsys, int &active, int &itMax) {
    YCN_HIP(hipMemcpyAsync(c.active_h, c.active, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipMemcpyAsync(c.cg_h, c.cg_d, sizeof(CgState) * nsys, hipMemcpyDeviceToHost, c.stream));
    int rc = spin_sync(c);
    if (rc) return rc;
    active = *c.active_h;
    itMax = 0;
    for (int k = 0; k < nsys; ++k)
        if (!c.cg_h[k].final_) itMax = std::max(itMax, c.cg_h[k].iter);
    return YCN_OK;
}

// Stencil-tile autotuner: times the fused normal operator for every tile shape that
// divides the lattice (and the two-stencil pair) on the context's own buffers and
// keeps the fastest; with mixed precision also the FP32 tiles.  Runs once per context
// before any solve; the next solve's setup overwrites what the trial launches wrote.
int cg_autotune(Ctx &c) {
    const LatticeArgs L = largs(c);
    char line[160];
    hipEvent_t e0, e1;
    YCN_HIP(hipEventCreate(&e0));
    YCN_HIP(hipEventCreate(&e1));
    const int reps = c.V >= (1 << 16) ? 10 : 3;
    auto time_it = [&](auto launch) -> double {
        launch(0);
        launch(1);
        (void)hipEventRecord(e0, c.stream);
        for (int i = 0; i < reps; ++i) launch(i & 1);
        (void)hipEventRecord(e1, c.stream);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return 1e3 * ms / reps;
    };
    std::vector<int> cands = {-1}, cands32;
    for (int i = 0; i < kNumTiles64; ++i)
        if (tile_fits(c, kTiles64[i])) cands.push_back(i);
    if (c.rs)
        for (int i = 0; i < kNumTiles32; ++i)
            if (tile_fits(c, kTiles32[i])) cands32.push_back(i);
    // measured times (us per launch) of cands, then cands32; reused from an earlier
    // context of the same shape when there is one
    std::vector<double> us;
    const bool memo = tune_lookup(0, c, us) && us.size() == cands.size() + cands32.size();
    if (!memo) {
        us.clear();
        for (int tile : cands)
            us.push_back(time_it(
                [&](int par) { (void)launch_normal_op(c, L, c.phi[c.cur], c.nSys, YCN_DAGGER, par, tile); }));
        for (int i : cands32)
            us.push_back(time_it([&](int par) {
                launch_fused<float>(c, i, c.nSys, false, c.phis, c.cg_d, (const float2 *)c.rs, c.ps[par],
                                    c.ps[par ^ 1], c.Aps, c.partials);
            }));
        tune_store(0, c, us);
    }
    c.tuneReport = memo ? "fp64 normal-operator tiles (us/launch, measured by an earlier context):"
                        : "fp64 normal-operator tiles (us/launch):";
    double best = 1e30;
    int bestTile = -1;
    for (size_t k = 0; k < cands.size(); ++k) {
        const int tile = cands[k];
        if (tile < 0)
            std::snprintf(line, sizeof line, " two-stencil %.1f;", us[k]);
        else
            std::snprintf(line, sizeof line, " %dx%d %.1f;", kTiles64[tile].TT, kTiles64[tile].TX, us[k]);
        c.tuneReport += line;
        if (us[k] < best) { best = us[k]; bestTile = tile; }
    }
    if (const char *force = std::getenv("YCN_TILE")) {
        const int f = std::atoi(force);
        bestTile = (f >= 0 && f < kNumTiles64 && tile_fits(c, kTiles64[f])) ? f : -1;
        c.tuneReport += " (forced by YCN_TILE)";
    }
    c.tile64 = bestTile;
    if (c.rs) {
        c.tuneReport += " | fp32 tiles (us/launch):";
        best = 1e30;
        int best32 = -1;
        for (size_t k = 0; k < cands32.size(); ++k) {
            const int i = cands32[k];
            const double t = us[cands.size() + k];
            std::snprintf(line, sizeof line, " %dx%d %.1f;", kTiles32[i].TT, kTiles32[i].TX, t);
            c.tuneReport += line;
            if (t < best) { best = t; best32 = i; }
        }
        if (const char *f32 = std::getenv("YCN_TILE32")) {
            const int f = std::atoi(f32);
            if (f >= 0 && f < kNumTiles32 && tile_fits(c, kTiles32[f])) best32 = f;
        }
        c.tile32 = best32;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    YCN_HIP(hipGetLastError());
    YCN_HIP(hipMemsetAsync(c.cg_d, 0, sizeof(CgState) * c.nSys, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

// the systems' final states into the pinned mirror c.cg_h, and their summary
int read_stats(Ctx &c, int nsys, SolveStats *st) {
    YCN_HIP(hipMemcpyAsync(c.cg_h, c.cg_d, sizeof(CgState) * nsys, hipMemcpyDeviceToHost, c.stream));
    int rc = spin_sync(c);
    if (rc) return rc;
    return stats_from_mirror(c, nsys, st);
}

// summary of the states already in the pinned mirror
int stats_from_mirror(Ctx &c, int nsys, SolveStats *st) {
    SolveStats out;
    out.systems = nsys;
    double sum = 0;
    for (int k = 0; k < nsys; ++k) {
        const CgState &s = c.cg_h[k];
        out.iters_max = std::max(out.iters_max, s.iter);
        out.restarts = std::max(out.restarts, s.restarts);
        out.converged &= s.converged;
        out.rr_max = std::max(out.rr_max, s.rr);
        out.true_rel_max = std::max(out.true_rel_max, s.true_rel);
        sum += s.iter;
    }
    out.iters_mean = sum / nsys;
    c.lastIters = std::max(kGraphIters, out.iters_max);
    if (st) *st = out;
    return YCN_OK;
}

int cg_solve(Ctx &c, const double *phi, int nsys, int mtype, bool gen, uint32_t streamBase, uint64_t step,
             SolveStats *st) {
    if (c.p.precond == YCN_PRECOND_CNN && mtype == YCN_DAGGER)
        return fcg_solve(c, phi, nsys, mtype, gen, streamBase, step, st);
    // mixed precision where the FP32 tiles fit the lattice (with even-odd preconditioning
    // the fermion force may run mixed on shapes the plain FP32 tiles do not cover)
    if ((c.p.cgPrecision == YCN_CG_MIXED && c.rs) || c.p.precond == YCN_PRECOND_CNN)
        return mp_solve(c, phi, nsys, mtype, gen, streamBase, step, st);
    const LatticeArgs L = largs(c);
    const dim3 g = site_grid(c, nsys);
    const int nb = site_blocks(c);
    int rc = set_int(c, c.active, nsys);
    if (rc) return rc;
    const int maxIter = c.p.cgMaxIter, maxRestarts = 4;
    if (gen)
        hipLaunchKernelGGL(k_cg_setup<true>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, streamBase,
                           step, c.eta, c.x, c.r, c.pv[0], c.pv[1], c.partials);
    else
        hipLaunchKernelGGL(k_cg_setup<false>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, streamBase,
                           step, c.eta, c.x, c.r, c.pv[0], c.pv[1], c.partials);
    hipLaunchKernelGGL(k_fin_setup, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb, c.cg_d,
                       c.active, c.p.cgTol, c.p.cgTolMode, maxIter, maxRestarts);
    YCN_HIP(hipGetLastError());

    int launched = 0, active = nsys;
    int chunk = std::max(kGraphIters, std::min(c.lastIters, maxIter));
    const int hardCap = maxIter * (maxRestarts + 1) + 2 * kGraphIters;
    for (int guard = 0; guard < maxRestarts + 2; ++guard) {
        while (active > 0 && launched < hardCap) {
            // about `chunk` iterations as captured graphs of kGraphIters, then poll
            const int n = std::min(chunk, hardCap - launched);
            for (int k = 0; k < n; k += kGraphIters) {
                rc = run_chunk(c, phi, nsys, mtype, kGraphIters, YCN_CG_FP64, c.tile64,
                               [&](int i) { launch_iteration(c, L, phi, nsys, mtype, i & 1); });
                if (rc) return rc;
            }
            launched += ((n + kGraphIters - 1) / kGraphIters) * kGraphIters;
            rc = poll_active(c, active);
            if (rc) return rc;
            chunk = std::max(kGraphIters, std::min(64, chunk / 2));
        }
        // FP64 true residual (restarts systems in YCN_TOL_REL_TRUE mode while above tol)
        if (mtype == YCN_DAGGER) {
            hipLaunchKernelGGL(k_true_a<false>, g, dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x, c.t);
            hipLaunchKernelGGL((k_true_b<true, false>), g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                               (const double2 *)c.eta, (const double2 *)c.t, c.r, (float2 *)nullptr, c.partials);
        } else {
            hipLaunchKernelGGL(k_true_a<true>, g, dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x, c.t);
            hipLaunchKernelGGL((k_true_b<false, false>), g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                               (const double2 *)c.eta, (const double2 *)c.t, c.r, (float2 *)nullptr, c.partials);
        }
        hipLaunchKernelGGL(k_fin_true, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb, c.cg_d,
                           c.active);
        YCN_HIP(hipGetLastError());
        rc = poll_active(c, active);
        if (rc) return rc;
        if (active == 0 || launched >= hardCap) break;
        chunk = kGraphIters;
    }
    SolveStats out;
    rc = read_stats(c, nsys, &out);
    if (rc) return rc;
    if (mtype == YCN_DAGGER && c.p.cgTolMode == YCN_TOL_REL_TRUE && out.converged)
        c.plainIters = out.iters_max;   // the flexible solver's yardstick
    if (st) *st = out;
    return YCN_OK;
}

// ===========================================================================
// Mixed-precision CG with reliable updates (Clark et al., "Solving lattice QCD
// systems of equations using mixed precision solvers on GPUs", CPC 181 (2010)):
// FP32 inner iterations (fused FP32 normal operator, FP32 x-increment and residual,
// dot products accumulated in FP64); every `reliableEvery` iterations, and once the
// FP32 recursive residual reaches the stop, the increment is folded into the FP64
// solution and the residual is recomputed in FP64, r = b - A x.  The search
// direction is kept across a reliable update (beta = |r_true|^2 / |r_prev|^2).
// Convergence is judged on that FP64 true residual only.
// ===========================================================================
int mp_alloc(Ctx &c) {
    const size_t n = (size_t)c.nSys * 4 * c.V;
    float2 **bufs[] = {&c.rs, &c.ps[0], &c.ps[1], &c.Aps, &c.xs};
    for (float2 **b : bufs) {
        YCN_HIP(hipMalloc(b, sizeof(float2) * n));
        YCN_HIP(hipMemset(*b, 0, sizeof(float2) * n));
    }
    YCN_HIP(hipMalloc(&c.phis, sizeof(float) * (size_t)c.nRep * c.V));
    YCN_HIP(hipMemset(c.phis, 0, sizeof(float) * (size_t)c.nRep * c.V));
    if (c.p.evenOdd) {   // second Ap buffer and third direction buffer of the pipelined even-odd iteration (half lattice)
        const size_t h = (size_t)c.nSys * 4 * (c.V / 2);
        float2 **pf[] = {&c.pfA1, &c.pfP2};
        for (float2 **b : pf) {
            YCN_HIP(hipMalloc(b, sizeof(float2) * h));
            YCN_HIP(hipMemset(*b, 0, sizeof(float2) * h));
        }
        int rc = eo_pf_alloc(c);
        if (rc) return rc;
    }
    return YCN_OK;
}

void mp_free(Ctx &c) {
    float2 **bufs[] = {&c.rs, &c.ps[0], &c.ps[1], &c.Aps, &c.xs, &c.pfA1, &c.pfP2};
    for (float2 **b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    if (c.phis) { (void)hipFree(c.phis); c.phis = nullptr; }
    for (double2 *&slot : c.pfSlot)
        if (slot) { (void)hipFree(slot); slot = nullptr; }
}

static void launch_mp_iteration(Ctx &c, int nsys, int mtype, int parity) {
    const float2 *pold = c.ps[parity];
    float2 *pnew = c.ps[parity ^ 1];
    int nbF;
    if (c.tile32 >= 0) {
        prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (4 * 32 + 4));
        launch_fused<float>(c, c.tile32, nsys, mtype != YCN_DAGGER, c.phis, c.cg_d, (const float2 *)c.rs, pold, pnew,
                            c.Aps, c.partials);
        prof_after(c, PK_NORMAL_OP);
        nbF = fused_args(c, kTiles32[c.tile32].TT, kTiles32[c.tile32].TX).tilesPerSys;
    } else {
        // no fused FP32 tile divides the lattice: the two-stencil pair, its intermediate
        // in the FP64 scratch c.t (free during the iteration chunks)
        const LatticeArgs L = largs(c);
        const dim3 g = site_grid(c, nsys);
        float2 *t = reinterpret_cast<float2 *>(c.t);
        // read r, p_old, phi, write p', t; read t, p', phi, write Ap  (7S + 2P, FP32)
        prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (7 * 32 + 8));
        if (mtype == YCN_DAGGER) {
            hipLaunchKernelGGL(k_cg_a32<false>, g, dim3(kNT), 0, c.stream, L, (const float *)c.phis,
                               (const CgState *)c.cg_d, (const float2 *)c.rs, pold, pnew, t);
            hipLaunchKernelGGL(k_cg_b32<true>, g, dim3(kNT), 0, c.stream, L, (const float *)c.phis,
                               (const CgState *)c.cg_d, (const float2 *)t, (const float2 *)pnew, c.Aps, c.partials);
        } else {
            hipLaunchKernelGGL(k_cg_a32<true>, g, dim3(kNT), 0, c.stream, L, (const float *)c.phis,
                               (const CgState *)c.cg_d, (const float2 *)c.rs, pold, pnew, t);
            hipLaunchKernelGGL(k_cg_b32<false>, g, dim3(kNT), 0, c.stream, L, (const float *)c.phis,
                               (const CgState *)c.cg_d, (const float2 *)t, (const float2 *)pnew, c.Aps, c.partials);
        }
        prof_after(c, PK_NORMAL_OP);
        nbF = site_blocks(c);
    }
    hipLaunchKernelGGL(k_fin_alpha, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nbF, c.cg_d);
    prof_before(c, PK_CG_UPDATE, (double)nsys * c.V * 6 * 32);
    hipLaunchKernelGGL(k_mp_update, site_grid(c, nsys), dim3(kNT), 0, c.stream, c.V, (const CgState *)c.cg_d, c.xs,
                       c.rs, (const float2 *)pnew, (const float2 *)c.Aps, c.partials);
    prof_after(c, PK_CG_UPDATE);
    hipLaunchKernelGGL(k_fin_beta<true>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials,
                       site_blocks(c), c.cg_d, c.active);
}

int mp_solve(Ctx &c, const double *phi, int nsys, int mtype, bool gen, uint32_t streamBase, uint64_t step,
             SolveStats *st) {
    if (!c.rs) {
        c.err = "mixed-precision CG: the context has no FP32 buffers";
        return YCN_ERR_UNSUPPORTED;
    }
    const LatticeArgs L = largs(c);
    const dim3 g = site_grid(c, nsys);
    const int nb = site_blocks(c);
    // FP32 copy of the field(s) these systems use
    const int nrepUsed = (nsys + c.nNoise - 1) / c.nNoise;
    const int64_t nconv = std::min<int64_t>((int64_t)c.nRep * c.V, (int64_t)nrepUsed * c.V);
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((nconv + kNT - 1) / kNT)), dim3(kNT), 0, c.stream, phi, c.phis,
                       nconv, (const int *)nullptr, (double *)nullptr);
    int rc = set_int(c, c.active, nsys);
    if (rc) return rc;
    const int maxIter = c.p.cgMaxIter;
    if (gen)
        hipLaunchKernelGGL(k_mp_setup<true>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, streamBase,
                           step, c.eta, c.x, c.xs, c.rs, c.ps[0], c.ps[1], c.partials);
    else
        hipLaunchKernelGGL(k_mp_setup<false>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, streamBase,
                           step, c.eta, c.x, c.xs, c.rs, c.ps[0], c.ps[1], c.partials);
    hipLaunchKernelGGL(k_fin_setup, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb, c.cg_d,
                       c.active, c.p.cgTol, c.p.cgTolMode, maxIter, 0);
    YCN_HIP(hipGetLastError());
    const int every = std::max(1, c.p.reliableEvery);
    const int nIt = every + (every & 1);   // even: keeps the p parity
    const int hardCap = maxIter + nIt + 8;
    int launched = 0, active = nsys;
    rc = poll_active(c, active);
    if (rc) return rc;
    while (active > 0 && launched < hardCap) {
        rc = run_chunk(c, nullptr, nsys, mtype, nIt, YCN_CG_MIXED, c.tile32,
                       [&](int i) { launch_mp_iteration(c, nsys, mtype, i & 1); });
        if (rc) return rc;
        launched += nIt;
        // reliable update: x += xs; rs = b - A x in FP64
        hipLaunchKernelGGL(k_mp_accum, g, dim3(kNT), 0, c.stream, c.V, (const CgState *)c.cg_d, c.x, c.xs);
        if (mtype == YCN_DAGGER) {
            hipLaunchKernelGGL(k_true_a<false>, g, dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x, c.t);
            hipLaunchKernelGGL((k_true_b<true, true>), g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                               (const double2 *)c.eta, (const double2 *)c.t, c.r, c.rs, c.partials);
        } else {
            hipLaunchKernelGGL(k_true_a<true>, g, dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x, c.t);
            hipLaunchKernelGGL((k_true_b<false, true>), g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                               (const double2 *)c.eta, (const double2 *)c.t, c.r, c.rs, c.partials);
        }
        hipLaunchKernelGGL(k_fin_mp_true<false>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials,
                           nb, c.cg_d, c.active, 0.0);
        YCN_HIP(hipGetLastError());
        rc = poll_active(c, active);
        if (rc) return rc;
    }
    return read_stats(c, nsys, st);
}

int fermion_drift(Ctx &c, const double *phi, double *drift) {
    const LatticeArgs L = largs(c);
    const double fac = -c.p.g * (c.p.driftMode == 1 ? c.p.cutFraction * c.p.cutFraction : 1.0);
    hipLaunchKernelGGL(k_drift, site_grid(c, c.nRep), dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x,
                       (const double2 *)c.eta, drift, fac);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int spinor_slices(Ctx &c, const double2 *psi, double *out_dev) {
    hipLaunchKernelGGL(k_spinor_slices, dim3(c.p.Nt), dim3(kNT), 0, c.stream, psi, c.p.Nt, c.p.Nx, out_dev);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

// ===========================================================================
// Flexible CG with the learned preconditioner (no reference
// counterpart): FP32 inner iterations as in the mixed solver, z = CNN(r_s) each
// iteration, FP64 reliable updates every `reliableEvery` iterations that restart the
// direction, and plain CG (identity preconditioner) once a reliable interval stalls.
// The FP64 true residual with the plain operator decides convergence.
// ===========================================================================
static void launch_fcg_iteration(Ctx &c, int nsys, int mtype, int parity) {
    const float2 *pold = c.ps[parity];
    float2 *pnew = c.ps[parity ^ 1];
    prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (4 * 32 + 4));
    launch_fused<float>(c, c.tile32, nsys, mtype != YCN_DAGGER, c.phis, c.cg_d, (const float2 *)c.zs, pold, pnew,
                        c.Aps, c.partials);
    prof_after(c, PK_NORMAL_OP);
    const int nbF = fused_args(c, kTiles32[c.tile32].TT, kTiles32[c.tile32].TX).tilesPerSys;
    hipLaunchKernelGGL(k_fin_alpha_fcg, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nbF, c.cg_d);
    prof_before(c, PK_CG_UPDATE, (double)nsys * c.V * 6 * 32);
    hipLaunchKernelGGL(k_mp_update, site_grid(c, nsys), dim3(kNT), 0, c.stream, c.V, (const CgState *)c.cg_d, c.xs,
                       c.rs, (const float2 *)pnew, (const float2 *)c.Aps, c.partials);
    prof_after(c, PK_CG_UPDATE);
    hipLaunchKernelGGL(k_fin_beta<true>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials,
                       site_blocks(c), c.cg_d, c.active);
    const int nbC = cnn_apply(c, nsys, c.cg_d, c.rs, c.Aps, c.zs, c.partials);
    hipLaunchKernelGGL(k_fin_fcg<0>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nbC, c.cg_d);
}

int fcg_solve(Ctx &c, const double *phi, int nsys, int mtype, bool gen, uint32_t streamBase, uint64_t step,
              SolveStats *st) {
    if (c.tile32 < 0 || !c.zs) {
        c.err = "the learned preconditioner needs Nt % 16 == 0 and Nx % 16 == 0";
        return YCN_ERR_UNSUPPORTED;
    }
    const LatticeArgs L = largs(c);
    const dim3 g = site_grid(c, nsys);
    const int nb = site_blocks(c);
    const int nrepUsed = (nsys + c.nNoise - 1) / c.nNoise;
    const int64_t nconv = std::min<int64_t>((int64_t)c.nRep * c.V, (int64_t)nrepUsed * c.V);
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((nconv + kNT - 1) / kNT)), dim3(kNT), 0, c.stream, phi, c.phis,
                       nconv, (const int *)nullptr, (double *)nullptr);
    int rc = cnn_gates(c, (int)std::min<int64_t>(nrepUsed, c.nRep));   // the network's field gates, once per solve
    if (rc) return rc;
    rc = set_int(c, c.active, nsys);
    if (rc) return rc;
    const int maxIter = c.p.cgMaxIter;
    if (gen)
        hipLaunchKernelGGL(k_mp_setup<true>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, streamBase,
                           step, c.eta, c.x, c.xs, c.rs, c.ps[0], c.ps[1], c.partials);
    else
        hipLaunchKernelGGL(k_mp_setup<false>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, streamBase,
                           step, c.eta, c.x, c.xs, c.rs, c.ps[0], c.ps[1], c.partials);
    hipLaunchKernelGGL(k_fin_setup, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb, c.cg_d,
                       c.active, c.p.cgTol, c.p.cgTolMode, maxIter, 0);
    // z = CNN(b), rz = <b, z>, p' = z on the first iteration
    int nbC = cnn_apply(c, nsys, c.cg_d, c.rs, c.Aps, c.zs, c.partials);
    if (nbC < 0) return nbC;
    hipLaunchKernelGGL(k_fin_fcg<2>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nbC, c.cg_d);
    YCN_HIP(hipGetLastError());
    const int every = std::max(1, c.p.reliableEvery);
    const int nIt = every + (every & 1);
    const int hardCap = maxIter + nIt + 8;
    int launched = 0, active = nsys;
    rc = poll_active(c, active);
    if (rc) return rc;
    while (active > 0 && launched < hardCap) {
        rc = run_chunk(c, nullptr, nsys, mtype, nIt, 2 /* flexible */, c.tile32,
                       [&](int i) { launch_fcg_iteration(c, nsys, mtype, i & 1); });
        if (rc) return rc;
        launched += nIt;
        hipLaunchKernelGGL(k_mp_accum, g, dim3(kNT), 0, c.stream, c.V, (const CgState *)c.cg_d, c.x, c.xs);
        hipLaunchKernelGGL(k_true_a<false>, g, dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x, c.t);
        hipLaunchKernelGGL((k_true_b<true, true>), g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                           (const double2 *)c.eta, (const double2 *)c.t, c.r, c.rs, c.partials);
        const double refRate = (c.plainIters > 0 && c.p.cgTol > 0.0) ? std::log10(c.p.cgTol) / c.plainIters : 0.0;
        hipLaunchKernelGGL(k_fin_mp_true<true>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb,
                           c.cg_d, c.active, refRate);
        nbC = cnn_apply(c, nsys, c.cg_d, c.rs, c.Aps, c.zs, c.partials);
        if (nbC < 0) return nbC;
        hipLaunchKernelGGL(k_fin_fcg<1>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nbC,
                           c.cg_d);
        YCN_HIP(hipGetLastError());
        rc = poll_active(c, active);
        if (rc) return rc;
    }
    return read_stats(c, nsys, st);
}

// e = x_final - x_snapshot
__global__ void __launch_bounds__(kNT) k_sub(const double2 *__restrict__ a, double2 *__restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x;
    if (i < n) b[i] = csub(a[i], b[i]);
}

// Online training on thermalisation: `nsteps` Langevin steps from counter step0 whose
// fermion solves (FP64 CG, one system) leave `pairsPerSolve` snapshots (r_k, x_k) at
// evenly spaced iterations; once a solve has converged each pair (r_k, e_k = x* - x_k)
// drives one Adam step of the network.  The chain itself advances as usual.
int train_preconditioner(Ctx &c, uint64_t step0, int nsteps, int pairsPerSolve, float lr, int *pairsUsed) {
    if (c.nSys != 1) { c.err = "training runs on a single system (nReplicas = nNoise = 1)"; return YCN_ERR_UNSUPPORTED; }
    if (!c.cnnW || c.tile32 < 0) { c.err = "the learned preconditioner is not available on this lattice"; return YCN_ERR_UNSUPPORTED; }
    const size_t sp = (size_t)4 * c.V;
    if (c.pairCap < pairsPerSolve) {
        if (c.pairR) (void)hipFree(c.pairR);
        if (c.pairX) (void)hipFree(c.pairX);
        YCN_HIP(hipMalloc(&c.pairR, sizeof(double2) * sp * pairsPerSolve));
        YCN_HIP(hipMalloc(&c.pairX, sizeof(double2) * sp * pairsPerSolve));
        c.pairCap = pairsPerSolve;
    }
    const LatticeArgs L = largs(c);
    const dim3 g = site_grid(c, 1);
    const int nb = site_blocks(c);
    int used = 0;
    c.trainLossFirst = -1;
    for (int k = 0; k < nsteps; ++k) {
        const uint64_t step = step0 + (uint64_t)k;
        int rc = noise_generate(c, step, nullptr, c.noise);
        if (rc) return rc;
        const double *phi = c.phi[c.cur];
        hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((c.V + kNT - 1) / kNT)), dim3(kNT), 0, c.stream, phi, c.phis,
                           (int64_t)c.V, (const int *)nullptr, (double *)nullptr);
        // FP64 CG (the reference recurrence) with snapshots
        rc = set_int(c, c.active, 1);
        if (rc) return rc;
        hipLaunchKernelGGL(k_cg_setup<true>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, 1u, step, c.eta,
                           c.x, c.r, c.pv[0], c.pv[1], c.partials);
        hipLaunchKernelGGL(k_fin_setup, dim3(1), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb, c.cg_d,
                           c.active, c.p.cgTol, c.p.cgTolMode, c.p.cgMaxIter, 4);
        const int pred = std::max(kGraphIters * (pairsPerSolve + 1), c.lastIters);
        int launched = 0, active = 1, taken = 0;
        while (active > 0 && launched < c.p.cgMaxIter) {
            rc = run_chunk(c, phi, 1, YCN_DAGGER, kGraphIters, YCN_CG_FP64, c.tile64,
                           [&](int i) { launch_iteration(c, L, phi, 1, YCN_DAGGER, i & 1); });
            if (rc) return rc;
            launched += kGraphIters;
            if (taken < pairsPerSolve && launched >= (taken + 1) * pred / (pairsPerSolve + 1)) {
                YCN_HIP(hipMemcpyAsync(c.pairR + taken * sp, c.r, sizeof(double2) * sp, hipMemcpyDeviceToDevice, c.stream));
                YCN_HIP(hipMemcpyAsync(c.pairX + taken * sp, c.x, sizeof(double2) * sp, hipMemcpyDeviceToDevice, c.stream));
                ++taken;
            }
            if (launched % (4 * kGraphIters) == 0) {
                rc = poll_active(c, active);
                if (rc) return rc;
            }
        }
        SolveStats st;
        rc = read_stats(c, 1, &st);
        if (rc) return rc;
        if (st.iters_max > 0) c.plainIters = st.iters_max;   // the flexible solver's yardstick
        // fermion drift and update: the chain advances exactly as a plain step
        rc = fermion_drift(c, phi, c.drift);
        if (rc) return rc;
        rc = bosonic_update(c, c.p.adaptive ? 0 : 1);
        if (rc) return rc;
        for (int i = 0; i < taken; ++i) {
            hipLaunchKernelGGL(k_sub, dim3((unsigned)((sp + kNT - 1) / kNT)), dim3(kNT), 0, c.stream,
                               (const double2 *)c.x, c.pairX + i * sp, (int64_t)sp);
            double loss = 0;
            rc = cnn_train_step(c, c.pairR + i * sp, c.pairX + i * sp, c.phis, lr, &loss, true);
            if (rc) return rc;
            if (c.trainLossFirst < 0) c.trainLossFirst = loss;
            c.trainLossLast = loss;
            ++used;
        }
    }
    YCN_HIP(hipMemsetAsync(c.cg_d, 0, sizeof(CgState), c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    c.trainPairs = used;
    if (pairsUsed) *pairsUsed = used;
    return YCN_OK;
}

}  // namespace ycn
