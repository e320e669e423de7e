// cg.hip -- Wilson-Dirac operator, CG on D^dag D / D D^dag and the stochastic
// fermion force on gfx950 (replaces DiracOP, CGsolver and FermionicDrift,
// code/src/Dirac.cu:23-190, CGsolver.cu:95-230, FermionicDrift.cu:96-149).
//
// All kernels work on `nsys` independent systems at once (grid.y = system): a
// system is one (replica, noise vector) pair, so replicas and several right-hand
// sides share every launch.  One thread owns one lattice site and all four spinor
// components (SoA planes, coalesced double2 loads).  Scalars of the recurrence live
// in CgState on the device; a block-partial + last-block scheme gives deterministic
// global dot products without atomics on the data and without host round trips.
//
// One CG iteration = three launches (standard CG dataflow of CGsolver.cu:160-219,
// with the p update folded into the first stencil):
//   cg_a : p' = r + beta p            (own site written, neighbours recomputed)
//          t  = D1 p'                 (D1 = D for A = D^dag D, D^dag for A = D D^dag)
//   cg_b : Ap = D2 t, <p', Ap>  -> alpha = rr / <p', Ap>            (complex, as ref)
//   cg_c : x += alpha p', r -= alpha Ap, |r|^2 -> beta, convergence test
// Algorithmic HBM bytes per site and iteration (S = 64 B spinor, P = 8 B field):
//   cg_a 4S+P, cg_b 3S+P, cg_c 6S  -> 13S + 2P = 848 B.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "cg_fused.hpp"
#include "ycn_device.hpp"
#include "ycn_internal.hpp"

namespace ycn {

constexpr int kNT = 256;  // threads per block for the stencil and BLAS-1 kernels
constexpr int kGraphIters = 8;  // CG iterations per captured graph (even: keeps the p parity)

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
// Load functor L(i, c) returns component c at flat site i.
template <bool DAG, typename Load>
__device__ __forceinline__ void wilson_site(const LatticeArgs &L, const WilsonCoef &w, double phi, int s,
                                            Load ld, double2 o[4]) {
    const int t = s / L.Nx, x = s - t * L.Nx;
    const Nbr n = neighbours(t, x, L.Nt, L.Nx);
    const double d = w.two_r_plus_m + w.g * phi;
#pragma unroll
    for (int f = 0; f < 4; f += 2) {
        wilson_flavour<DAG, double2>(w, d, ld(s, f), ld(s, f + 1), ld(n.tf, f), ld(n.tf, f + 1), n.sf,
                                     ld(n.tb, f), ld(n.tb, f + 1), n.sb, ld(n.xf, f), ld(n.xf, f + 1),
                                     ld(n.xb, f), ld(n.xb, f + 1), o[f], o[f + 1]);
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

// ---------------------------------------------------------------------------
// CG kernels
// ---------------------------------------------------------------------------
__global__ void k_set_int(int *p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// Finish a global sum over the partials of one system: called by the last block.
__device__ __forceinline__ double2 sum_partials(const double2 *part, int n, double2 *lds) {
    double2 acc = make_double2(0.0, 0.0);
    for (int i = threadIdx.x; i < n; i += kNT) {
        const double *pp = reinterpret_cast<const double *>(part + i);
        acc.x += load_coherent(pp);
        acc.y += load_coherent(pp + 1);
    }
    __syncthreads();
    return block_sum2<kNT>(acc, lds);
}

// setup: eta (generated or given) -> b; x = 0; r = b; p0 = p1 = b; rr = |b|^2
template <bool GEN>
__global__ void __launch_bounds__(kNT)
k_cg_setup(LatticeArgs L, const uint64_t *__restrict__ seeds, uint32_t streamBase, uint64_t step,
           double2 *__restrict__ eta, double2 *__restrict__ x, double2 *__restrict__ r, double2 *__restrict__ p0,
           double2 *__restrict__ p1, double2 *__restrict__ part, unsigned *__restrict__ counters,
           CgState *__restrict__ st, int *__restrict__ active, double tol, int mode, int maxIter, int maxRestarts) {
    __shared__ double2 lds[kNT / kWave];
    __shared__ bool last;
    const int sys = blockIdx.y;
    const size_t V = L.V;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t off = (size_t)sys * 4 * V;
    double2 acc = make_double2(0.0, 0.0);
    if (s < L.V) {
        double2 b[4];
        if (GEN) {
            // fillNormalRND (FermionicDrift.cu:141-149): component c of site s is
            // normal number 4s+c of the stream; imaginary part 0
            const int rep = sys / L.nNoise, j = sys - rep * L.nNoise;
            const double2 g0 = normal_pair(seeds[rep], streamBase + (uint32_t)j, step, (uint64_t)2 * s);
            const double2 g1 = normal_pair(seeds[rep], streamBase + (uint32_t)j, step, (uint64_t)2 * s + 1);
            b[0] = make_double2(g0.x, 0.0);
            b[1] = make_double2(g0.y, 0.0);
            b[2] = make_double2(g1.x, 0.0);
            b[3] = make_double2(g1.y, 0.0);
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
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
        part[(size_t)sys * nb + blockIdx.x] = acc;
        last = arrive_last(&counters[sys], nb);
    }
    __syncthreads();
    if (!last) return;
    const double2 tot = sum_partials(part + (size_t)sys * nb, nb, lds);
    if (threadIdx.x == 0) {
        CgState &S = st[sys];
        S.rr = tot.x;
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
        const bool run = (0 < maxIter) && (sqrt(tot.x) > S.stop);
        S.done = run ? 0 : 1;
        S.converged = run ? 0 : 1;
        if (!run) atomicSub(active, 1);
        counters[sys] = 0u;
    }
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

// cg_b: Ap = D2 t; <p', Ap> -> alpha = rr / <p', Ap>
template <bool DAG2>
__global__ void __launch_bounds__(kNT)
k_cg_b(LatticeArgs L, const double *__restrict__ phi, CgState *__restrict__ st, const double2 *__restrict__ t,
       const double2 *__restrict__ p, double2 *__restrict__ Ap, double2 *__restrict__ part,
       unsigned *__restrict__ counters) {
    __shared__ double2 lds[kNT / kWave];
    __shared__ bool last;
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
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
        part[(size_t)sys * nb + blockIdx.x] = acc;
        last = arrive_last(&counters[sys], nb);
    }
    __syncthreads();
    if (!last) return;
    const double2 dot = sum_partials(part + (size_t)sys * nb, nb, lds);
    if (threadIdx.x == 0) {
        // alpha = rr / dot (complex division, CGsolver.cu:190)
        const double rr = st[sys].rr;
        const double den = dot.x * dot.x + dot.y * dot.y;
        st[sys].alpha = make_double2(rr * dot.x / den, -rr * dot.y / den);
        counters[sys] = 0u;
    }
}

// cg_c: x += alpha p'; r -= alpha Ap; |r|^2 -> beta, iteration count, stop test
__global__ void __launch_bounds__(kNT)
k_cg_c(int V_, CgState *__restrict__ st, double2 *__restrict__ x, double2 *__restrict__ r,
       const double2 *__restrict__ p, const double2 *__restrict__ Ap, double2 *__restrict__ part,
       unsigned *__restrict__ counters, int *__restrict__ active) {
    __shared__ double2 lds[kNT / kWave];
    __shared__ bool last;
    const int sys = blockIdx.y;
    if (st[sys].done) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t V = V_, off = (size_t)sys * 4 * V;
    const double2 al = st[sys].alpha;
    double2 acc = make_double2(0.0, 0.0);
    if (s < V_) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t i = off + c * V + s;
            const double2 pv = p[i], av = Ap[i];
            x[i] = cadd(x[i], cmul(al, pv));
            const double2 rn = csub(r[i], cmul(al, av));
            r[i] = rn;
            acc.x += rn.x * rn.x + rn.y * rn.y;
        }
    }
    acc = block_sum2<kNT>(acc, lds);
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
        part[(size_t)sys * nb + blockIdx.x] = acc;
        last = arrive_last(&counters[sys], nb);
    }
    __syncthreads();
    if (!last) return;
    const double2 tot = sum_partials(part + (size_t)sys * nb, nb, lds);
    if (threadIdx.x == 0) {
        CgState &S = st[sys];
        const double rrn = tot.x;
        S.beta = rrn / S.rr;
        S.rr = rrn;
        S.iter += 1;
        const bool run = (S.iter < S.maxIter) && (sqrt(rrn) > S.stop);
        if (!run) {
            S.done = 1;
            S.converged = sqrt(rrn) <= S.stop ? 1 : 0;
            atomicSub(active, 1);
        }
        counters[sys] = 0u;
    }
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

// true residual, part 2: res = b - D2 t -> r; |res| / |b|; restart decision
template <bool DAG2>
__global__ void __launch_bounds__(kNT)
k_true_b(LatticeArgs L, const double *__restrict__ phi, CgState *__restrict__ st, const double2 *__restrict__ b,
         const double2 *__restrict__ t, double2 *__restrict__ r, double2 *__restrict__ part,
         unsigned *__restrict__ counters, int *__restrict__ active) {
    __shared__ double2 lds[kNT / kWave];
    __shared__ bool last;
    const int sys = blockIdx.y;
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
            r[off + c * V + s] = res;
            acc.x += res.x * res.x + res.y * res.y;
        }
    }
    acc = block_sum2<kNT>(acc, lds);
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
        part[(size_t)sys * nb + blockIdx.x] = acc;
        last = arrive_last(&counters[sys], nb);
    }
    __syncthreads();
    if (!last) return;
    const double2 tot = sum_partials(part + (size_t)sys * nb, nb, lds);
    if (threadIdx.x == 0) {
        CgState &S = st[sys];
        const double bn = sqrt(S.bnorm2);
        S.true_rel = bn > 0.0 ? sqrt(tot.x) / bn : sqrt(tot.x);
        if (S.mode == YCN_TOL_REL_TRUE) {
            if (S.true_rel > S.tol && S.iter < S.maxIter && S.restarts < S.maxRestarts) {
                // reliable restart from x: r = b - A x (already in r), p = r
                S.restarts += 1;
                S.rr = tot.x;
                S.beta = 0.0;
                S.done = 0;
                S.converged = 0;
                atomicAdd(active, 1);
            } else {
                S.converged = S.true_rel <= S.tol ? 1 : 0;
            }
        }
        counters[sys] = 0u;
    }
}

// drift_x = fac/nNoise * sum_j Re sum_k conj((D y_j)_xk) eta_j,xk
// (FermionicDrift.cu:108 applyD Normal + computeDrift :122-139)
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

int cg_alloc(Ctx &c) {
    const size_t n = (size_t)c.nSys * 4 * c.V;
    c.maxBlocks = (c.V + 63) / 64;   // partial slots for the smallest (64-site) blocks
    double2 **bufs[] = {&c.eta, &c.x, &c.r, &c.pv[0], &c.pv[1], &c.t, &c.Ap};
    for (double2 **b : bufs) YCN_HIP(hipMalloc(b, sizeof(double2) * n));
    for (double2 **b : bufs) YCN_HIP(hipMemset(*b, 0, sizeof(double2) * n));
    YCN_HIP(hipMalloc(&c.cg_d, sizeof(CgState) * c.nSys));
    YCN_HIP(hipMemset(c.cg_d, 0, sizeof(CgState) * c.nSys));
    YCN_HIP(hipMalloc(&c.partials, sizeof(double2) * (size_t)c.nSys * c.maxBlocks));
    YCN_HIP(hipMalloc(&c.counters, sizeof(unsigned) * (size_t)(c.nSys + c.nRep)));
    YCN_HIP(hipMemset(c.counters, 0, sizeof(unsigned) * (size_t)(c.nSys + c.nRep)));
    YCN_HIP(hipMalloc(&c.active, sizeof(int)));
    YCN_HIP(hipHostMalloc((void **)&c.active_h, sizeof(int), hipHostMallocDefault));
    return YCN_OK;
}

void cg_free(Ctx &c) {
    for (const Ctx::GraphEntry &e : c.graphs) (void)hipGraphExecDestroy(e.exec);
    c.graphs.clear();
    double2 **bufs[] = {&c.eta, &c.x, &c.r, &c.pv[0], &c.pv[1], &c.t, &c.Ap};
    for (double2 **b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    if (c.cg_d) (void)hipFree(c.cg_d);
    if (c.partials) (void)hipFree(c.partials);
    if (c.counters) (void)hipFree(c.counters);
    if (c.active) (void)hipFree(c.active);
    if (c.active_h) (void)hipHostFree(c.active_h);
    c.cg_d = nullptr;
    c.partials = nullptr;
    c.counters = nullptr;
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
    dim3 g((c.V + kNT - 1) / kNT, nsys);
    prof_before(c, PK_DIRAC, (double)nsys * c.V * (2 * 64 + 8));
    if (dagger)
        hipLaunchKernelGGL(k_dirac<true>, g, dim3(kNT), 0, c.stream, L, phi_dev, in, out);
    else
        hipLaunchKernelGGL(k_dirac<false>, g, dim3(kNT), 0, c.stream, L, phi_dev, in, out);
    prof_after(c, PK_DIRAC);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

static void launch_cg_c(Ctx &c, int nsys, const double2 *pnew) {
    dim3 g((c.V + kNT - 1) / kNT, nsys);
    prof_before(c, PK_CG_UPDATE, (double)nsys * c.V * 6 * 64);
    hipLaunchKernelGGL(k_cg_c, g, dim3(kNT), 0, c.stream, c.V, c.cg_d, c.x, c.r, pnew, (const double2 *)c.Ap,
                       c.partials, c.counters, c.active);
    prof_after(c, PK_CG_UPDATE);
}

// p' = r + beta p_old and Ap = A p' with <p', Ap>: the fused LDS kernel with tile `tile`
// (>= 0) or the two-stencil pair cg_a + cg_b (tile = -1).
static void launch_normal_op(Ctx &c, const LatticeArgs &L, const double *phi, int nsys, int mtype, int parity,
                             int tile) {
    const double2 *pold = c.pv[parity];
    double2 *pnew = c.pv[parity ^ 1];
    const bool dag1 = (mtype != YCN_DAGGER);
    if (tile >= 0) {
        // compulsory traffic: read r, p_old, phi; write p', Ap  (4S + P per site)
        prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (4 * 64 + 8));
        launch_fused<double>(c, tile, nsys, dag1, phi, c.cg_d, (const double2 *)c.r, pold, pnew, c.Ap, c.partials,
                             c.counters);
        prof_after(c, PK_NORMAL_OP);
        return;
    }
    dim3 g((c.V + kNT - 1) / kNT, nsys);
    // cg_a + cg_b: read r, p_old, phi, write p', t; read t, p', phi, write Ap  (7S + 2P)
    prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (7 * 64 + 16));
    if (!dag1)
        hipLaunchKernelGGL(k_cg_a<false>, g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                           (const double2 *)c.r, pold, pnew, c.t);
    else
        hipLaunchKernelGGL(k_cg_a<true>, g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                           (const double2 *)c.r, pold, pnew, c.t);
    if (!dag1)
        hipLaunchKernelGGL(k_cg_b<true>, g, dim3(kNT), 0, c.stream, L, phi, c.cg_d, (const double2 *)c.t,
                           (const double2 *)pnew, c.Ap, c.partials, c.counters);
    else
        hipLaunchKernelGGL(k_cg_b<false>, g, dim3(kNT), 0, c.stream, L, phi, c.cg_d, (const double2 *)c.t,
                           (const double2 *)pnew, c.Ap, c.partials, c.counters);
    prof_after(c, PK_NORMAL_OP);
}

static int launch_iteration(Ctx &c, const LatticeArgs &L, const double *phi, int nsys, int mtype, int parity) {
    launch_normal_op(c, L, phi, nsys, mtype, parity, c.tile64);
    launch_cg_c(c, nsys, c.pv[parity ^ 1]);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

// Launch n CG iterations (n even keeps the p ping-pong parity).  When graphs are on and
// no kernel is being profiled, the n iterations are captured once into a hipGraph per
// (field, systems, operator, n, precision, tile) and replayed; every kernel reads its
// scalars from device state and exits early for converged systems, so replays need no
// host-side parameters.
template <typename Launch>
static int run_chunk(Ctx &c, const void *phi, int nsys, int mtype, int n, int precision, int tile, Launch launch) {
    if (!c.useGraphs || c.profTarget != PK_NONE) {
        for (int i = 0; i < n; ++i) launch(i & 1);
        YCN_HIP(hipGetLastError());
        return YCN_OK;
    }
    for (const Ctx::GraphEntry &e : c.graphs)
        if (e.phi == phi && e.nsys == nsys && e.mtype == mtype && e.n == n && e.precision == precision &&
            e.tile == tile) {
            YCN_HIP(hipGraphLaunch(e.exec, c.stream));
            return YCN_OK;
        }
    hipGraph_t g;
    YCN_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < n; ++i) launch(i & 1);
    const hipError_t le = hipGetLastError();
    YCN_HIP(hipStreamEndCapture(c.stream, &g));
    if (le != hipSuccess) {
        (void)hipGraphDestroy(g);
        c.err = std::string("launch during capture: ") + hipGetErrorString(le);
        return YCN_ERR_HIP;
    }
    hipGraphExec_t exec;
    YCN_HIP(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    c.graphs.push_back({phi, nsys, mtype, n, precision, tile, exec});
    YCN_HIP(hipGraphLaunch(exec, c.stream));
    return YCN_OK;
}

// Stencil-tile autotuner: times the fused normal operator for every tile shape that
// divides the lattice (and the two-stencil pair) on the context's own buffers and
// keeps the fastest.  Runs once per context, before any solve; its writes to the CG
// state and partial buffers are overwritten by the next solve's setup.
int cg_autotune(Ctx &c) {
    const LatticeArgs L = largs(c);
    const char *force = std::getenv("YCN_TILE");
    std::vector<int> cands = {-1};
    for (int i = 0; i < kNumTiles64; ++i)
        if (tile_fits(c, kTiles64[i])) cands.push_back(i);
    char line[160];
    c.tuneReport = "fp64 normal-operator tiles (us/launch):";
    hipEvent_t e0, e1;
    YCN_HIP(hipEventCreate(&e0));
    YCN_HIP(hipEventCreate(&e1));
    double best = 1e30;
    int bestTile = -1;
    const int reps = c.V >= (1 << 16) ? 10 : 3;
    for (int tile : cands) {
        for (int w = 0; w < 2; ++w) launch_normal_op(c, L, c.phi[c.cur], c.nSys, YCN_DAGGER, w & 1, tile);
        YCN_HIP(hipEventRecord(e0, c.stream));
        for (int i = 0; i < reps; ++i) launch_normal_op(c, L, c.phi[c.cur], c.nSys, YCN_DAGGER, i & 1, tile);
        YCN_HIP(hipEventRecord(e1, c.stream));
        YCN_HIP(hipEventSynchronize(e1));
        float ms = 0;
        YCN_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double us = 1e3 * ms / reps;
        if (tile < 0)
            std::snprintf(line, sizeof line, " two-stencil %.1f;", us);
        else
            std::snprintf(line, sizeof line, " %dx%d %.1f;", kTiles64[tile].TT, kTiles64[tile].TX, us);
        c.tuneReport += line;
        if (us < best) { best = us; bestTile = tile; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (force) {
        const int f = std::atoi(force);
        bestTile = (f >= 0 && f < kNumTiles64 && tile_fits(c, kTiles64[f])) ? f : -1;
        c.tuneReport += " (forced by YCN_TILE)";
    }
    c.tile64 = bestTile;
    if (c.p.cgPrecision == YCN_CG_MIXED) {
        c.tuneReport += " | fp32 tiles (us/launch):";
        best = 1e30;
        int best32 = -1;
        YCN_HIP(hipEventCreate(&e0));
        YCN_HIP(hipEventCreate(&e1));
        for (int i = 0; i < kNumTiles32; ++i) {
            if (!tile_fits(c, kTiles32[i])) continue;
            for (int w = 0; w < reps + 2; ++w) {
                if (w == 2) YCN_HIP(hipEventRecord(e0, c.stream));
                launch_fused<float>(c, i, c.nSys, false, c.phis, c.cg_d, (const float2 *)c.rs, c.ps[w & 1],
                                    c.ps[(w & 1) ^ 1], c.Aps, c.partials, c.counters);
            }
            YCN_HIP(hipEventRecord(e1, c.stream));
            YCN_HIP(hipEventSynchronize(e1));
            float ms = 0;
            YCN_HIP(hipEventElapsedTime(&ms, e0, e1));
            const double us = 1e3 * ms / reps;
            std::snprintf(line, sizeof line, " %dx%d %.1f;", kTiles32[i].TT, kTiles32[i].TX, us);
            c.tuneReport += line;
            if (us < best) { best = us; best32 = i; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        const char *f32 = std::getenv("YCN_TILE32");
        if (f32) {
            const int f = std::atoi(f32);
            if (f >= 0 && f < kNumTiles32 && tile_fits(c, kTiles32[f])) best32 = f;
        }
        c.tile32 = best32;
    }
    // reset the scratch state the trial launches touched
    YCN_HIP(hipMemsetAsync(c.cg_d, 0, sizeof(CgState) * c.nSys, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

static int poll_active(Ctx &c, int &active) {
    YCN_HIP(hipMemcpyAsync(c.active_h, c.active, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    active = *c.active_h;
    return YCN_OK;
}

int cg_solve(Ctx &c, const double *phi, int nsys, int mtype, bool gen, uint32_t streamBase, uint64_t step,
             SolveStats *st) {
    if (c.p.cgPrecision == YCN_CG_MIXED) return mp_solve(c, phi, nsys, mtype, gen, streamBase, step, st);
    const LatticeArgs L = largs(c);
    dim3 g((c.V + kNT - 1) / kNT, nsys);
    int rc = set_int(c, c.active, nsys);
    if (rc) return rc;
    const int maxIter = c.p.cgMaxIter;
    const int maxRestarts = 4;
    if (gen)
        hipLaunchKernelGGL(k_cg_setup<true>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, streamBase,
                           step, c.eta, c.x, c.r, c.pv[0], c.pv[1], c.partials, c.counters, c.cg_d, c.active,
                           c.p.cgTol, c.p.cgTolMode, maxIter, maxRestarts);
    else
        hipLaunchKernelGGL(k_cg_setup<false>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, streamBase,
                           step, c.eta, c.x, c.r, c.pv[0], c.pv[1], c.partials, c.counters, c.cg_d, c.active,
                           c.p.cgTol, c.p.cgTolMode, maxIter, maxRestarts);
    YCN_HIP(hipGetLastError());

    int launched = 0, active = nsys;
    int chunk = std::max(kGraphIters, std::min(c.lastIters, maxIter));
    const int hardCap = maxIter * (maxRestarts + 1) + 8;
    for (int guard = 0; guard < 64 + hardCap; ++guard) {
        while (active > 0 && launched < hardCap) {
            // launch ~chunk iterations as graphs of kGraphIters, then poll
            const int n = std::min(chunk, hardCap - launched);
            for (int done_n = 0; done_n < n; done_n += kGraphIters) {
                rc = run_chunk(c, phi, nsys, mtype, kGraphIters, YCN_CG_FP64, c.tile64, [&](int par) {
                    (void)launch_iteration(c, L, phi, nsys, mtype, par);
                });
                if (rc) return rc;
            }
            launched += ((n + kGraphIters - 1) / kGraphIters) * kGraphIters;
            rc = poll_active(c, active);
            if (rc) return rc;
            chunk = std::max(kGraphIters, std::min(64, chunk / 2));
        }
        // true residual check (may restart systems in YCN_TOL_REL_TRUE mode)
        if (mtype == YCN_DAGGER) {
            hipLaunchKernelGGL(k_true_a<false>, g, dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x, c.t);
            hipLaunchKernelGGL(k_true_b<true>, g, dim3(kNT), 0, c.stream, L, phi, c.cg_d, (const double2 *)c.eta,
                               (const double2 *)c.t, c.r, c.partials, c.counters, c.active);
        } else {
            hipLaunchKernelGGL(k_true_a<true>, g, dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x, c.t);
            hipLaunchKernelGGL(k_true_b<false>, g, dim3(kNT), 0, c.stream, L, phi, c.cg_d, (const double2 *)c.eta,
                               (const double2 *)c.t, c.r, c.partials, c.counters, c.active);
        }
        YCN_HIP(hipGetLastError());
        rc = poll_active(c, active);
        if (rc) return rc;
        if (active == 0 || launched >= hardCap) break;
        chunk = kGraphIters;
    }
    std::vector<CgState> S(nsys);
    YCN_HIP(hipMemcpyAsync(S.data(), c.cg_d, sizeof(CgState) * nsys, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    SolveStats out;
    out.systems = nsys;
    double sum = 0;
    for (const CgState &s : S) {
        out.iters_max = std::max(out.iters_max, s.iter);
        out.restarts = std::max(out.restarts, s.restarts);
        out.converged &= s.converged;
        out.rr_max = std::max(out.rr_max, s.rr);
        out.true_rel_max = std::max(out.true_rel_max, s.true_rel);
        sum += s.iter;
    }
    out.iters_mean = sum / nsys;
    c.lastIters = std::max(4, out.iters_max);
    if (st) *st = out;
    return YCN_OK;
}

int fermion_drift(Ctx &c, const double *phi, double *drift) {
    const LatticeArgs L = largs(c);
    dim3 g((c.V + kNT - 1) / kNT, c.nRep);
    const double fac = -c.p.g * (c.p.driftMode == 1 ? c.p.cutFraction * c.p.cutFraction : 1.0);
    hipLaunchKernelGGL(k_drift, g, dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x, (const double2 *)c.eta,
                       drift, fac);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int spinor_slices(Ctx &c, const double2 *psi, double *out_dev) {
    hipLaunchKernelGGL(k_spinor_slices, dim3(c.p.Nt), dim3(kNT), 0, c.stream, psi, c.p.Nt, c.p.Nx, out_dev);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

}  // namespace ycn

// ===========================================================================
// Mixed-precision CG with reliable updates (Clark et al., "Solving lattice QCD
// systems of equations using mixed precision solvers on GPUs", CPC 181 (2010)):
// FP32 inner iterations (fused FP32 normal operator, FP32 x-increment and
// residual, dot products accumulated in FP64); every `reliableEvery` iterations
// and whenever the FP32 recursive residual reaches the stop, the increment is
// folded into the FP64 solution and the residual is recomputed in FP64,
// r = b - A x.  The search direction is kept across a reliable update
// (beta = |r_true|^2 / |r_prev|^2).  Convergence is judged on that FP64 true
// residual only.
// ===========================================================================
namespace ycn {

__global__ void __launch_bounds__(kNT) k_to_f32(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kNT + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

template <bool GEN>
__global__ void __launch_bounds__(kNT)
k_mp_setup(LatticeArgs L, const uint64_t *__restrict__ seeds, uint32_t streamBase, uint64_t step,
           double2 *__restrict__ eta, double2 *__restrict__ x, float2 *__restrict__ xs, float2 *__restrict__ rs,
           float2 *__restrict__ p0, float2 *__restrict__ p1, double2 *__restrict__ part,
           unsigned *__restrict__ counters, CgState *__restrict__ st, int *__restrict__ active, double tol, int mode,
           int maxIter) {
    __shared__ double2 lds[kNT / kWave];
    __shared__ bool last;
    const int sys = blockIdx.y;
    const size_t V = L.V;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t off = (size_t)sys * 4 * V;
    double2 acc = make_double2(0.0, 0.0);
    if (s < L.V) {
        double2 b[4];
        if (GEN) {
            const int rep = sys / L.nNoise, j = sys - rep * L.nNoise;
            const double2 g0 = normal_pair(seeds[rep], streamBase + (uint32_t)j, step, (uint64_t)2 * s);
            const double2 g1 = normal_pair(seeds[rep], streamBase + (uint32_t)j, step, (uint64_t)2 * s + 1);
            b[0] = make_double2(g0.x, 0.0);
            b[1] = make_double2(g0.y, 0.0);
            b[2] = make_double2(g1.x, 0.0);
            b[3] = make_double2(g1.y, 0.0);
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
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
        part[(size_t)sys * nb + blockIdx.x] = acc;
        last = arrive_last(&counters[sys], nb);
    }
    __syncthreads();
    if (!last) return;
    const double2 tot = sum_partials(part + (size_t)sys * nb, nb, lds);
    if (threadIdx.x == 0) {
        CgState &S = st[sys];
        S.rr = tot.x;
        S.rr_prev = tot.x;
        S.bnorm2 = tot.x;
        S.tol = tol;
        S.mode = mode;
        S.maxIter = maxIter;
        S.maxRestarts = 0;
        S.stop = (mode == YCN_TOL_ABS_RECURSIVE) ? tol : tol * sqrt(tot.x);
        S.beta = 0.0;
        S.alpha = make_double2(0.0, 0.0);
        S.iter = 0;
        S.restarts = 0;
        S.true_rel = 0.0;
        const bool run = (0 < maxIter) && (sqrt(tot.x) > S.stop);
        S.done = run ? 0 : 1;
        S.final_ = run ? 0 : 1;
        S.converged = run ? 0 : 1;
        if (!run) atomicSub(active, 1);
        counters[sys] = 0u;
    }
}

// FP32 update: xs += alpha p', rs -= alpha Ap; |rs|^2 in FP64 -> beta; pause the
// system when its recursive residual reaches the stop or the iteration cap.
__global__ void __launch_bounds__(kNT)
k_mp_update(int V_, CgState *__restrict__ st, float2 *__restrict__ xs, float2 *__restrict__ rs,
            const float2 *__restrict__ p, const float2 *__restrict__ Ap, double2 *__restrict__ part,
            unsigned *__restrict__ counters, int *__restrict__ active) {
    __shared__ double2 lds[kNT / kWave];
    __shared__ bool last;
    const int sys = blockIdx.y;
    if (st[sys].done) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    const size_t V = V_, off = (size_t)sys * 4 * V;
    const double2 ad = st[sys].alpha;
    const float ax = (float)ad.x, ay = (float)ad.y;
    double2 acc = make_double2(0.0, 0.0);
    if (s < V_) {
        float part_rr = 0.f;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const size_t i = off + c * V + s;
            const float2 pv = p[i], av = Ap[i];
            float2 xv = xs[i], rv = rs[i];
            xv.x += ax * pv.x - ay * pv.y;
            xv.y += ax * pv.y + ay * pv.x;
            rv.x -= ax * av.x - ay * av.y;
            rv.y -= ax * av.y + ay * av.x;
            xs[i] = xv;
            rs[i] = rv;
            part_rr += rv.x * rv.x + rv.y * rv.y;
        }
        acc.x = (double)part_rr;
    }
    acc = block_sum2<kNT>(acc, lds);
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
        part[(size_t)sys * nb + blockIdx.x] = acc;
        last = arrive_last(&counters[sys], nb);
    }
    __syncthreads();
    if (!last) return;
    const double2 tot = sum_partials(part + (size_t)sys * nb, nb, lds);
    if (threadIdx.x == 0) {
        CgState &S = st[sys];
        const double rrn = tot.x;
        S.rr_prev = S.rr;
        S.beta = rrn / S.rr;
        S.rr = rrn;
        S.iter += 1;
        if (S.iter >= S.maxIter || sqrt(rrn) <= S.stop) {
            S.done = 1;   // wait for the reliable update to decide
            atomicSub(active, 1);
        }
        counters[sys] = 0u;
    }
}

// reliable update, part 1: x += xs (FP64), xs = 0
__global__ void __launch_bounds__(kNT)
k_mp_accum(int V_, const CgState *__restrict__ st, double2 *__restrict__ x, float2 *__restrict__ xs) {
    const int sys = blockIdx.y;
    if (st[sys].final_) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= V_) return;
    const size_t V = V_, off = (size_t)sys * 4 * V;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const size_t i = off + c * V + s;
        const float2 d = xs[i];
        const double2 xv = x[i];
        x[i] = make_double2(xv.x + (double)d.x, xv.y + (double)d.y);
        xs[i] = make_float2(0.f, 0.f);
    }
}

// reliable update, part 2 (after k_true_a: t = D1 x): rs = b - D2 t in FP64, decide
template <bool DAG2>
__global__ void __launch_bounds__(kNT)
k_mp_true_b(LatticeArgs L, const double *__restrict__ phi, CgState *__restrict__ st, const double2 *__restrict__ b,
            const double2 *__restrict__ t, float2 *__restrict__ rs, double2 *__restrict__ part,
            unsigned *__restrict__ counters, int *__restrict__ active) {
    __shared__ double2 lds[kNT / kWave];
    __shared__ bool last;
    const int sys = blockIdx.y;
    if (st[sys].final_) return;
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
            rs[off + c * V + s] = make_float2((float)res.x, (float)res.y);
            acc.x += res.x * res.x + res.y * res.y;
        }
    }
    acc = block_sum2<kNT>(acc, lds);
    const int nb = gridDim.x;
    if (threadIdx.x == 0) {
        part[(size_t)sys * nb + blockIdx.x] = acc;
        last = arrive_last(&counters[sys], nb);
    }
    __syncthreads();
    if (!last) return;
    const double2 tot = sum_partials(part + (size_t)sys * nb, nb, lds);
    if (threadIdx.x == 0) {
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
        } else {
            S.beta = rrt / S.rr_prev;
            S.rr = rrt;
            if (S.done) { S.done = 0; atomicAdd(active, 1); }
        }
        counters[sys] = 0u;
    }
}

int mp_alloc(Ctx &c) {
    const size_t n = (size_t)c.nSys * 4 * c.V;
    float2 **bufs[] = {&c.rs, &c.ps[0], &c.ps[1], &c.Aps, &c.xs};
    for (float2 **b : bufs) {
        YCN_HIP(hipMalloc(b, sizeof(float2) * n));
        YCN_HIP(hipMemset(*b, 0, sizeof(float2) * n));
    }
    YCN_HIP(hipMalloc(&c.phis, sizeof(float) * (size_t)c.nRep * c.V));
    YCN_HIP(hipMemset(c.phis, 0, sizeof(float) * (size_t)c.nRep * c.V));
    return YCN_OK;
}

void mp_free(Ctx &c) {
    float2 **bufs[] = {&c.rs, &c.ps[0], &c.ps[1], &c.Aps, &c.xs};
    for (float2 **b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    if (c.phis) { (void)hipFree(c.phis); c.phis = nullptr; }
}

static void launch_mp_iteration(Ctx &c, int nsys, int mtype, int parity, const float *phis) {
    dim3 g((c.V + kNT - 1) / kNT, nsys);
    const float2 *pold = c.ps[parity];
    float2 *pnew = c.ps[parity ^ 1];
    prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (4 * 32 + 4));
    launch_fused<float>(c, c.tile32, nsys, mtype != YCN_DAGGER, phis, c.cg_d, (const float2 *)c.rs, pold, pnew,
                        c.Aps, c.partials, c.counters);
    prof_after(c, PK_NORMAL_OP);
    prof_before(c, PK_CG_UPDATE, (double)nsys * c.V * 6 * 32);
    hipLaunchKernelGGL(k_mp_update, g, dim3(kNT), 0, c.stream, c.V, c.cg_d, c.xs, c.rs, (const float2 *)pnew,
                       (const float2 *)c.Aps, c.partials, c.counters, c.active);
    prof_after(c, PK_CG_UPDATE);
}

int mp_solve(Ctx &c, const double *phi, int nsys, int mtype, bool gen, uint32_t streamBase, uint64_t step,
             SolveStats *st) {
    if (c.tile32 < 0) {
        c.err = "mixed-precision CG needs Nt % 4 == 0 and Nx % 16 == 0 (fused FP32 tiles)";
        return YCN_ERR_UNSUPPORTED;
    }
    const LatticeArgs L = largs(c);
    dim3 g((c.V + kNT - 1) / kNT, nsys);
    const int64_t nphi = (int64_t)c.nRep * c.V;
    const int nrep_used = (nsys + c.nNoise - 1) / c.nNoise;
    // FP32 copy of the field(s) used by these systems
    const int64_t nconv = std::min<int64_t>(nphi, (int64_t)nrep_used * c.V);
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((nconv + kNT - 1) / kNT)), dim3(kNT), 0, c.stream, phi, c.phis,
                       nconv);
    int rc = set_int(c, c.active, nsys);
    if (rc) return rc;
    const int maxIter = c.p.cgMaxIter;
    if (gen)
        hipLaunchKernelGGL(k_mp_setup<true>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, streamBase,
                           step, c.eta, c.x, c.xs, c.rs, c.ps[0], c.ps[1], c.partials, c.counters, c.cg_d, c.active,
                           c.p.cgTol, c.p.cgTolMode, maxIter);
    else
        hipLaunchKernelGGL(k_mp_setup<false>, g, dim3(kNT), 0, c.stream, L, (const uint64_t *)c.seed_d, streamBase,
                           step, c.eta, c.x, c.xs, c.rs, c.ps[0], c.ps[1], c.partials, c.counters, c.cg_d, c.active,
                           c.p.cgTol, c.p.cgTolMode, maxIter);
    YCN_HIP(hipGetLastError());
    const int every = std::max(1, c.p.reliableEvery);
    const int hardCap = maxIter + every + 8;
    int launched = 0, active = nsys;
    rc = poll_active(c, active);
    if (rc) return rc;
    while (active > 0 && launched < hardCap) {
        const int nIt = every + (every & 1);   // even: keeps the p parity
        rc = run_chunk(c, phi, nsys, mtype, nIt, YCN_CG_MIXED, c.tile32,
                       [&](int par) { launch_mp_iteration(c, nsys, mtype, par, c.phis); });
        if (rc) return rc;
        launched += nIt;
        // reliable update: x += xs; r = b - A x (FP64)
        hipLaunchKernelGGL(k_mp_accum, g, dim3(kNT), 0, c.stream, c.V, (const CgState *)c.cg_d, c.x, c.xs);
        if (mtype == YCN_DAGGER) {
            hipLaunchKernelGGL(k_true_a<false>, g, dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x, c.t);
            hipLaunchKernelGGL(k_mp_true_b<true>, g, dim3(kNT), 0, c.stream, L, phi, c.cg_d, (const double2 *)c.eta,
                               (const double2 *)c.t, c.rs, c.partials, c.counters, c.active);
        } else {
            hipLaunchKernelGGL(k_true_a<true>, g, dim3(kNT), 0, c.stream, L, phi, (const double2 *)c.x, c.t);
            hipLaunchKernelGGL(k_mp_true_b<false>, g, dim3(kNT), 0, c.stream, L, phi, c.cg_d,
                               (const double2 *)c.eta, (const double2 *)c.t, c.rs, c.partials, c.counters, c.active);
        }
        YCN_HIP(hipGetLastError());
        rc = poll_active(c, active);
        if (rc) return rc;
    }
    std::vector<CgState> S(nsys);
    YCN_HIP(hipMemcpyAsync(S.data(), c.cg_d, sizeof(CgState) * nsys, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    SolveStats out;
    out.systems = nsys;
    double sum = 0;
    for (const CgState &s : S) {
        out.iters_max = std::max(out.iters_max, s.iter);
        out.restarts = std::max(out.restarts, s.restarts);
        out.converged &= s.converged;
        out.rr_max = std::max(out.rr_max, s.rr);
        out.true_rel_max = std::max(out.true_rel_max, s.true_rel);
        sum += s.iter;
    }
    out.iters_mean = sum / nsys;
    if (st) *st = out;
    return YCN_OK;
}

}  // namespace ycn
