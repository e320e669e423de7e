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
#include <vector>

#include "ycn_device.hpp"
#include "ycn_internal.hpp"

namespace ycn {

constexpr int kNT = 256;  // threads per block for the stencil and BLAS-1 kernels

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
    c.maxBlocks = (c.V + kNT - 1) / kNT;
    double2 **bufs[] = {&c.eta, &c.x, &c.r, &c.pv[0], &c.pv[1], &c.t, &c.Ap};
    for (double2 **b : bufs) YCN_HIP(hipMalloc(b, sizeof(double2) * n));
    YCN_HIP(hipMalloc(&c.cg_d, sizeof(CgState) * c.nSys));
    YCN_HIP(hipMalloc(&c.partials, sizeof(double2) * (size_t)c.nSys * c.maxBlocks));
    YCN_HIP(hipMalloc(&c.counters, sizeof(unsigned) * (size_t)(c.nSys + c.nRep)));
    YCN_HIP(hipMemset(c.counters, 0, sizeof(unsigned) * (size_t)(c.nSys + c.nRep)));
    YCN_HIP(hipMalloc(&c.active, sizeof(int)));
    YCN_HIP(hipHostMalloc((void **)&c.active_h, sizeof(int), hipHostMallocDefault));
    return YCN_OK;
}

void cg_free(Ctx &c) {
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

static int launch_iteration(Ctx &c, const LatticeArgs &L, const double *phi, int nsys, int mtype, int parity) {
    dim3 g((c.V + kNT - 1) / kNT, nsys);
    const double2 *pold = c.pv[parity];
    double2 *pnew = c.pv[parity ^ 1];
    // cg_a compulsory traffic: read r, p_old, phi; write p', t  (4S + P per site)
    prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (4 * 64 + 8));
    if (mtype == YCN_DAGGER)
        hipLaunchKernelGGL(k_cg_a<false>, g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                           (const double2 *)c.r, pold, pnew, c.t);
    else
        hipLaunchKernelGGL(k_cg_a<true>, g, dim3(kNT), 0, c.stream, L, phi, (const CgState *)c.cg_d,
                           (const double2 *)c.r, pold, pnew, c.t);
    prof_after(c, PK_NORMAL_OP);
    if (mtype == YCN_DAGGER)
        hipLaunchKernelGGL(k_cg_b<true>, g, dim3(kNT), 0, c.stream, L, phi, c.cg_d, (const double2 *)c.t,
                           (const double2 *)pnew, c.Ap, c.partials, c.counters);
    else
        hipLaunchKernelGGL(k_cg_b<false>, g, dim3(kNT), 0, c.stream, L, phi, c.cg_d, (const double2 *)c.t,
                           (const double2 *)pnew, c.Ap, c.partials, c.counters);
    prof_before(c, PK_CG_UPDATE, (double)nsys * c.V * 6 * 64);
    hipLaunchKernelGGL(k_cg_c, g, dim3(kNT), 0, c.stream, c.V, c.cg_d, c.x, c.r, (const double2 *)pnew,
                       (const double2 *)c.Ap, c.partials, c.counters, c.active);
    prof_after(c, PK_CG_UPDATE);
    YCN_HIP(hipGetLastError());
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

    int parity = 0, launched = 0, active = nsys;
    int chunk = std::max(4, std::min(c.lastIters, maxIter));
    const int hardCap = maxIter * (maxRestarts + 1) + 8;
    for (int guard = 0; guard < 64 + hardCap; ++guard) {
        while (active > 0 && launched < hardCap) {
            const int n = std::min(chunk, hardCap - launched);
            for (int i = 0; i < n; ++i) {
                rc = launch_iteration(c, L, phi, nsys, mtype, parity);
                if (rc) return rc;
                parity ^= 1;
            }
            launched += n;
            rc = poll_active(c, active);
            if (rc) return rc;
            chunk = std::max(4, std::min(64, chunk / 2));
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
        chunk = 8;
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
