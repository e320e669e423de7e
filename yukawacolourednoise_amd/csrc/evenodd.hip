// evenodd.hip -- even-odd preconditioned fermion force on gfx950.
//
// The drift only needs chi = (D^dag)^-1 eta (chi = D y with D^dag D y = eta,
// FermionicDrift.cu:57-100).  With D^dag = A + K, A = diag(2r + M + g phi) and K the
// hopping, which connects even and odd sites, the Schur complement on the even sites
//   S = A_e - K_eo A_o^-1 K_oe,   b_e = eta_e - K_eo A_o^-1 eta_o,
//   S chi_e = b_e,   chi_o = A_o^-1 (eta_o - K_oe chi_e)
// is solved by CG on S^T S chi_e = S^T b_e; S^T = A_e - H_eo A_o^-1 H_oe is the Schur
// complement of D (H = K^T, the hopping of D).  |eta - D^dag chi| = |b_e - S chi_e|
// exactly (the odd rows vanish by construction), so the solve stops on the residual of
// the unpreconditioned system, the quantity the plain solver checks.
//
// Half-lattice vectors use the compact checkerboard order of the reference's even-odd
// blocks (commented out in Dirac.cu:256-379): site (t, x) of a parity at t * Nx/2 +
// (x >> 1), SoA planes [sys][c][Vh].  Needs even Nt and Nx.
//
// FP64: one CG iteration = k_eo_normal_fused (p' = r + beta p, v = S p', Ap = S^T v,
// |v|^2: four half-lattice hops through LDS) + the BLAS-1 update on Vh sites: 5S + P
// bytes per site of the full lattice, against 10S + P for the unpreconditioned
// iteration.  Mixed precision: one launch per pipelined FP32 iteration (132 B per site),
// its finalize in the next launch's prologue or a separate kernel, and FP64 reliable
// updates whose residual b - S x and S^T of it take one fused launch.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "cg_fused.hpp"
#include "cg_host.hpp"
#include "ycn_device.hpp"
#include "ycn_internal.hpp"

namespace ycn {

namespace {
constexpr int kNT = 256;
}

struct EoArgs {
    int Nt, Nx, Nh, V, Vh, nNoise;   // Nh = Nx / 2, Vh = V / 2
    double r, M, g;
    int tilesX, tilesPerSys, xcdRemap;
    float invNt, invNx;              // 1 / Nt, 1 / Nx in single precision (wrapf)
    // graph-captured steps (else null): the setup, drift kernels return at once while
    // *halt is set, and k_eo_eta takes its noise counter from *stepPtr
    const int *halt;
    const uint64_t *stepPtr;
};

__device__ __forceinline__ bool eo_halted(const EoArgs &A) { return A.halt && *A.halt; }

// site (t, x) of compact index j on parity par
__device__ __forceinline__ void eo_site(const EoArgs &A, int par, int j, int &t, int &x) {
    t = j / A.Nh;
    x = 2 * (j - t * A.Nh) + ((t + par) & 1);
}
__device__ __forceinline__ int eo_j(const EoArgs &A, int t, int x) { return t * A.Nh + (x >> 1); }

// a mod n in [0, n) without an integer division (the 4-site halo of a tile spans several
// periods when an extent is smaller than 4, Nt = 2).  For |a| < 2^22 and inv = fl(1 / n),
// (a + 1/2) / n lies at least 1 / (2n) from an integer and the two single-precision
// roundings move it by less than |a + 1/2| / n * 2^-23, so the floor is exact.
__device__ __forceinline__ int wrapf(int a, int n, float inv) {
    return a - n * (int)floorf(((float)a + 0.5f) * inv);
}
// element idx of a buffer through a 32-bit byte offset (the buffers of one system are
// below 4 GiB), so that the load or store takes a scalar base and a vector offset
template <typename T>
__device__ __forceinline__ T &at(T *base, unsigned idx) {
    using C = typename std::conditional<std::is_const<T>::value, const char, char>::type;
    return *reinterpret_cast<T *>(reinterpret_cast<C *>(base) + (unsigned)(idx * (unsigned)sizeof(T)));
}

// Hopping (K for DAG, H otherwise) at site (t, x); ld(t', x', c) returns component c of
// the input at neighbour (t', x') of the other parity.
template <bool DAG, typename R, typename V2, typename Ld>
__device__ __forceinline__ void eo_hop(const EoArgs &A, int t, int x, Ld ld, V2 o[4]) {
    const int tp = (t == A.Nt - 1) ? 0 : t + 1, tm = (t == 0) ? A.Nt - 1 : t - 1;
    const int xp = (x == A.Nx - 1) ? 0 : x + 1, xm = (x == 0) ? A.Nx - 1 : x - 1;
    const R sf = (t == A.Nt - 1) ? (R)-1 : (R)1, sb = (t == 0) ? (R)-1 : (R)1;
    const R hp = (R)(0.5 * (A.r + 1.0)), hm = (R)(0.5 * (A.r - 1.0)), hr = (R)(0.5 * A.r), h = (R)0.5;
#pragma unroll
    for (int f = 0; f < 4; f += 2)
        hop_flavour_r<DAG, R, V2>(hp, hm, hr, h, ld(tp, x, f), ld(tp, x, f + 1), sf, ld(tm, x, f), ld(tm, x, f + 1),
                                  sb, ld(t, xp, f), ld(t, xp, f + 1), ld(t, xm, f), ld(t, xm, f + 1), o[f], o[f + 1]);
}

// ---------------------------------------------------------------------------
// setup, Schur blocks for setup and reliable updates, drift (FP64, untiled)
// ---------------------------------------------------------------------------

// eta generated (GEN) or given; u_o = A_o^-1 eta_o on the odd sites (compact layout);
// partial |eta|^2 over the full lattice
template <bool GEN>
__global__ void __launch_bounds__(kNT)
k_eo_eta(EoArgs A, const uint64_t *__restrict__ seeds, uint32_t streamBase, uint64_t step, const double *__restrict__ phi,
         double2 *__restrict__ eta, double2 *__restrict__ u, double2 *__restrict__ part) {
    __shared__ double2 lds[kNT / kWave];
    if (eo_halted(A)) return;
    if (A.stepPtr) step = *A.stepPtr;
    const int sys = blockIdx.y, s = blockIdx.x * kNT + threadIdx.x;
    const size_t V = A.V, off = (size_t)sys * 4 * V;
    double2 acc = make_double2(0.0, 0.0);
    if (s < A.V) {
        double2 b[4];
        if (GEN) {
            eta_site(seeds, streamBase, step, sys, A.nNoise, s, b);
#pragma unroll
            for (int c = 0; c < 4; ++c) eta[off + c * V + s] = b[c];
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) b[c] = eta[off + c * V + s];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) acc.x += b[c].x * b[c].x + b[c].y * b[c].y;
        const int t = s / A.Nx, x = s - t * A.Nx;
        if ((t + x) & 1) {
            const double inv = 1.0 / (2.0 * A.r + A.M + A.g * phi[(size_t)(sys / A.nNoise) * V + s]);
            const size_t j = (size_t)sys * 4 * A.Vh + eo_j(A, t, x);
#pragma unroll
            for (int c = 0; c < 4; ++c) u[j + (size_t)c * A.Vh] = make_double2(b[c].x * inv, b[c].y * inv);
        }
    }
    acc = block_sum2<kNT>(acc, lds);
    if (threadIdx.x == 0) part[(size_t)sys * gridDim.x + blockIdx.x] = acc;
}

// b_e = eta_e - K_eo u_o with u_o = A_o^-1 eta_o (from k_eo_eta)
__global__ void __launch_bounds__(kNT)
k_eo_rhs(EoArgs A, const double2 *__restrict__ eta, const double2 *__restrict__ u, double2 *__restrict__ b) {
    if (eo_halted(A)) return;
    const int sys = blockIdx.y, j = blockIdx.x * kNT + threadIdx.x;
    if (j >= A.Vh) return;
    int t, x;
    eo_site(A, 0, j, t, x);
    const size_t V = A.V, Vh = A.Vh;
    const double2 *e = eta + (size_t)sys * 4 * V;
    const double2 *uo = u + (size_t)sys * 4 * Vh;
    double2 k[4];
    eo_hop<true, double, double2>(A, t, x, [&](int tn, int xn, int c) { return uo[c * Vh + eo_j(A, tn, xn)]; }, k);
    const size_t fs = (size_t)t * A.Nx + x;
    double2 *bo = b + (size_t)sys * 4 * Vh;
#pragma unroll
    for (int c = 0; c < 4; ++c) bo[c * Vh + j] = csub(e[c * V + fs], k[c]);
}

// U_o = A_o^-1 K_oe in_e (DAG) or A_o^-1 H_oe in_e
template <bool DAG>
__global__ void __launch_bounds__(kNT)
k_eo_hop_oe(EoArgs A, const double *__restrict__ phi, const double2 *__restrict__ in, double2 *__restrict__ U) {
    if (eo_halted(A)) return;
    const int sys = blockIdx.y, j = blockIdx.x * kNT + threadIdx.x;
    if (j >= A.Vh) return;
    int t, x;
    eo_site(A, 1, j, t, x);
    const size_t Vh = A.Vh, off = (size_t)sys * 4 * Vh;
    const double *ph = phi + (size_t)(sys / A.nNoise) * A.V;
    double2 k[4];
    eo_hop<DAG, double, double2>(A, t, x, [&](int tn, int xn, int c) { return in[off + c * Vh + eo_j(A, tn, xn)]; },
                                 k);
    const double inv = 1.0 / (2.0 * A.r + A.M + A.g * ph[(size_t)t * A.Nx + x]);
#pragma unroll
    for (int c = 0; c < 4; ++c) U[off + c * Vh + j] = make_double2(k[c].x * inv, k[c].y * inv);
}

// half-lattice vectors a setup launch also clears at its sites (any may be null)
struct EoZero {
    double2 *d[3];
    float2 *f[3];
};

// v_e = A_e in_e - K_eo U_o (DAG: S in) or A_e in_e - H_eo U_o (S^T in).
// MODE 0: out = v; MODE 1: out = b - v and partial |out|^2; MODE 2 (a solve's setup):
// out = v and partial |v|^2, and the vectors of z cleared.  out32 (optional) receives
// an FP32 copy of out.
template <bool DAG, int MODE>
__global__ void __launch_bounds__(kNT)
k_eo_hop_eo(EoArgs A, const double *__restrict__ phi, const double2 *__restrict__ in, const double2 *__restrict__ U,
            const double2 *__restrict__ b, double2 *__restrict__ out, float2 *__restrict__ out32,
            double2 *__restrict__ part, EoZero z) {
    __shared__ double2 lds[kNT / kWave];
    if (eo_halted(A)) return;
    const int sys = blockIdx.y, j = blockIdx.x * kNT + threadIdx.x;
    const size_t Vh = A.Vh, off = (size_t)sys * 4 * Vh;
    double2 acc = make_double2(0.0, 0.0);
    if (j < A.Vh) {
        int t, x;
        eo_site(A, 0, j, t, x);
        const double *ph = phi + (size_t)(sys / A.nNoise) * A.V;
        double2 k[4];
        eo_hop<DAG, double, double2>(A, t, x, [&](int tn, int xn, int c) { return U[off + c * Vh + eo_j(A, tn, xn)]; },
                                     k);
        const double d = 2.0 * A.r + A.M + A.g * ph[(size_t)t * A.Nx + x];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double2 iv = in[off + c * Vh + j];
            double2 v = make_double2(d * iv.x - k[c].x, d * iv.y - k[c].y);
            if (MODE == 1) v = csub(b[off + c * Vh + j], v);
            out[off + c * Vh + j] = v;
            if (out32) out32[off + c * Vh + j] = make_float2((float)v.x, (float)v.y);
            acc.x += v.x * v.x + v.y * v.y;
            if (MODE == 2) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    if (z.d[q]) z.d[q][off + c * Vh + j] = make_double2(0.0, 0.0);
                    if (z.f[q]) z.f[q][off + c * Vh + j] = make_float2(0.f, 0.f);
                }
            }
        }
    }
    if (MODE != 0) {
        acc = block_sum2<kNT>(acc, lds);
        if (threadIdx.x == 0) part[(size_t)sys * gridDim.x + blockIdx.x] = acc;
    }
}

// drift_x = fac Re sum_k conj(chi_xk) eta_xk (FermionicDrift.cu:83-100) with chi_e from
// the solve and chi_o = A_o^-1 (eta_o - K_oe chi_e) formed on the fly; mean over noise
// vectors.
__global__ void __launch_bounds__(kNT)
k_eo_drift(EoArgs A, const double *__restrict__ phi, const double2 *__restrict__ chi_e,
           const double2 *__restrict__ eta, double *__restrict__ drift, double fac) {
    if (eo_halted(A)) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= A.V) return;
    const int rep = blockIdx.y;
    const size_t V = A.V, Vh = A.Vh;
    const int t = s / A.Nx, x = s - t * A.Nx;
    const double *ph = phi + (size_t)rep * V;
    const bool odd = ((t + x) & 1) != 0;
    const double d = 2.0 * A.r + A.M + A.g * ph[s];
    double acc = 0.0;
    for (int jn = 0; jn < A.nNoise; ++jn) {
        const size_t sys = (size_t)rep * A.nNoise + jn;
        const double2 *ce = chi_e + sys * 4 * Vh;
        const double2 *e = eta + sys * 4 * V;
        double2 ch[4];
        if (odd) {
            double2 k[4];
            eo_hop<true, double, double2>(A, t, x, [&](int tn, int xn, int c) { return ce[c * Vh + eo_j(A, tn, xn)]; },
                                          k);
            const double inv = 1.0 / d;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double2 ev = e[c * V + s];
                ch[c] = make_double2((ev.x - k[c].x) * inv, (ev.y - k[c].y) * inv);
            }
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) ch[c] = ce[c * Vh + eo_j(A, t, x)];
        }
        double2 sm = cdotc(ch[0], e[s]);
        sm = cadd(sm, cdotc(ch[1], e[V + s]));
        sm = cadd(sm, cdotc(ch[2], e[2 * V + s]));
        sm = cadd(sm, cdotc(ch[3], e[3 * V + s]));
        acc += sm.x;
    }
    drift[(size_t)rep * V + s] = (A.nNoise == 1) ? fac * acc : fac * acc / (double)A.nNoise;
}

// gamma_5 = sigma_2 on each flavour's spin pair, (c0, c1) -> (-i c1, i c0); with it
// gamma_5 D gamma_5 = D^dag for the Wilson-Yukawa operator (Dirac.cu:144-190)
__device__ __forceinline__ void g5_pair(double2 a0, double2 a1, double2 &o0, double2 &o1) {
    o0 = make_double2(a1.y, -a1.x);
    o1 = make_double2(-a0.y, a0.x);
}

// v = gamma_5 v in place (system 0)
__global__ void __launch_bounds__(kNT) k_g5(int V, double2 *__restrict__ v) {
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= V) return;
#pragma unroll
    for (int f = 0; f < 4; f += 2) {
        double2 o0, o1;
        g5_pair(v[f * V + s], v[(f + 1) * V + s], o0, o1);
        v[f * V + s] = o0;
        v[(f + 1) * V + s] = o1;
    }
}

// The measurement's propagator from the even-odd solve of D^dag chi = eta with
// eta = gamma_5 src (system 0): chi on both parities (the odd sites rebuilt as in
// k_eo_drift) and out = gamma_5 chi = D^-1 src, full lattice, SoA.
__global__ void __launch_bounds__(kNT)
k_eo_propagator(EoArgs A, const double *__restrict__ phi, const double2 *__restrict__ chi_e,
                const double2 *__restrict__ eta, double2 *__restrict__ out) {
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= A.V) return;
    const size_t V = A.V, Vh = A.Vh;
    const int t = s / A.Nx, x = s - t * A.Nx;
    double2 ch[4];
    if ((t + x) & 1) {
        double2 k[4];
        eo_hop<true, double, double2>(A, t, x, [&](int tn, int xn, int c) { return chi_e[c * Vh + eo_j(A, tn, xn)]; },
                                      k);
        const double inv = 1.0 / (2.0 * A.r + A.M + A.g * phi[s]);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double2 ev = eta[c * V + s];
            ch[c] = make_double2((ev.x - k[c].x) * inv, (ev.y - k[c].y) * inv);
        }
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c) ch[c] = chi_e[c * Vh + eo_j(A, t, x)];
    }
#pragma unroll
    for (int f = 0; f < 4; f += 2) {
        double2 o0, o1;
        g5_pair(ch[f], ch[f + 1], o0, o1);
        out[f * V + s] = o0;
        out[(f + 1) * V + s] = o1;
    }
}

// ---------------------------------------------------------------------------
// finalize kernels (one block per system, fixed-order sums)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 eo_fin_sum(const double2 *part, int nb, double2 *lds) {
    double2 acc = make_double2(0.0, 0.0);
    int i = threadIdx.x;
    for (; i + 7 * kNT < nb; i += 8 * kNT) {
        double2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[i + u * kNT];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = cadd(acc, v[u]);
    }
    for (; i < nb; i += kNT) acc = cadd(acc, part[i]);
    return block_sum2<kNT>(acc, lds);
}

// |eta|^2 and |c|^2 = |S^T b|^2: rr, thresholds and run flags of a new solve
__global__ void __launch_bounds__(kNT)
k_fin_eo_setup(const double2 *__restrict__ partE, int nbE, const double2 *__restrict__ partC, int nbC,
               CgState *__restrict__ st, int *__restrict__ active, double tol, int mode, int maxIter, int maxRestarts,
               double stopScale, const StepCtl *__restrict__ ctl) {
    __shared__ double2 lds0[kNT / kWave], lds1[kNT / kWave];
    if (ctl) {   // graph-captured step: nothing while halted; the threshold learned on the device
        if (ctl->halt) return;
        stopScale = ctl->stopScale;
    }
    const int sys = blockIdx.x;
    const double2 e = eo_fin_sum(partE + (size_t)sys * nbE, nbE, lds0);
    const double2 cc = eo_fin_sum(partC + (size_t)sys * nbC, nbC, lds1);
    if (threadIdx.x) return;
    CgState &S = st[sys];
    S.enorm2 = e.x;
    S.rr = cc.x;
    S.rr_prev = cc.x;
    S.rrRel = cc.x;
    S.bnorm2 = cc.x;
    S.tol = tol;
    S.mode = mode;
    S.maxIter = maxIter;
    S.maxRestarts = maxRestarts;
    S.stop0 = (mode == YCN_TOL_ABS_RECURSIVE) ? tol : tol * sqrt(cc.x);
    S.stop = (mode == YCN_TOL_REL_TRUE) ? S.stop0 * stopScale : S.stop0;
    S.need = 0.0;
    S.nRu = 0;
    S.beta = 0.0;
    S.alpha = make_double2(0.0, 0.0);
    S.iter = 0;
    S.iter0 = 0;
    S.restarts = 0;
    S.true_rel = 0.0;
    S.rz = 0.0;
    S.true_prev = 1.0;
    S.precOff = 0;
    S.restartNext = 0;
    const bool run = (0 < maxIter) && (sqrt(cc.x) > S.stop);
    S.done = run ? 0 : 1;
    S.final_ = run ? 0 : 1;
    S.converged = run ? 0 : 1;
    if (!run) atomicSub(active, 1);
}

// FP64 solve: true residual |b_e - S x_e| / |eta|.  While it exceeds tol (relative mode)
// the iteration continues from its current state with the normal-equation threshold
// lowered by the missing factor (times 1/2).
__global__ void __launch_bounds__(kNT)
k_fin_eo_true(const double2 *__restrict__ part, int nb, CgState *__restrict__ st, int *__restrict__ active) {
    __shared__ double2 lds[kNT / kWave];
    const int sys = blockIdx.x;
    const double2 tot = eo_fin_sum(part + (size_t)sys * nb, nb, lds);
    if (threadIdx.x) return;
    CgState &S = st[sys];
    S.true_rel = S.enorm2 > 0.0 ? sqrt(tot.x / S.enorm2) : sqrt(tot.x);
    if (S.mode != YCN_TOL_REL_TRUE) return;
    if (S.done && S.true_rel > 0.0) S.need = sqrt(S.rr) * S.tol / S.true_rel;   // paused at the stop
    if (S.true_rel > S.tol && S.iter < S.maxIter && S.restarts < S.maxRestarts) {
        S.restarts += 1;
        S.stop = sqrt(S.rr) * 0.5 * S.tol / S.true_rel;
        S.converged = 0;
        if (S.done) {
            S.done = 0;
            atomicAdd(active, 1);
        }
    } else {
        S.converged = S.true_rel <= S.tol ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// finalize of the pipelined FP32 iteration
// ---------------------------------------------------------------------------
// Sums of one system's nb block partials of the pipelined iteration (pp[2 b] = (mu, nu),
// pp[2 b + 1] = (rho, 0)) by one block in a fixed order: thread t adds entries t, t + NT,
// ..., each round issuing its U loads before adding; valid in thread 0.
template <int NT>
__device__ void sum_pf_partials(const double2 *pp, int nb, double2 *lds0, double2 *lds1, double2 &mn, double2 &rh) {
    constexpr int U = 4;
    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
    for (int base = threadIdx.x; base < nb; base += U * NT) {
        double2 v[2 * U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = base + u * NT;
            v[2 * u] = k < nb ? pp[2 * k] : make_double2(0.0, 0.0);
            v[2 * u + 1] = k < nb ? pp[2 * k + 1] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            a0 = cadd(a0, v[2 * u]);
            a1 = cadd(a1, v[2 * u + 1]);
        }
    }
    mn = block_sum2<NT>(a0, lds0);
    rh = block_sum2<NT>(a1, lds1);
}

// The decision after iteration number `iter` of the pipelined iteration, from (mu, nu)
// and the exact rho of that iteration: alpha = rho / mu and beta = (alpha^2 nu - rho) /
// rho for the next one (a negative prediction, from cancellation, restarts the
// direction).  Pause at the stop after an even number of iterations (the last direction
// applied is then in p[1]), at the cap, or on a non-positive <p, Ap>; a reliable update
// follows a pause.
struct PfStep {
    double alpha, beta;
    bool stop;
};
__device__ __forceinline__ PfStep pf_decide(double mu, double nu, double rho, int iter, const CgState &S) {
    PfStep d;
    d.stop = (iter >= S.maxIter) || (sqrt(rho) <= S.stop && !(iter & 1)) || (!(mu > 0.0) && !(iter & 1)) ||
             (rho < kReliableDrop * kReliableDrop * S.rrRel && !(iter & 1)) || !isfinite(mu) || !isfinite(rho);
    d.alpha = (mu > 0.0) ? rho / mu : 0.0;
    double rhoNext = d.alpha * d.alpha * nu - rho;
    if (!(rhoNext > 0.0)) rhoNext = 0.0;
    d.beta = (rho > 0.0) ? rhoNext / rho : 0.0;
    return d;
}

// the finalize of one iteration from its sums (thread 0): iteration count, rho, pause or
// the next alpha, beta; commit = false only keeps the sums (timing launches)
__device__ __forceinline__ void pf_commit(CgState &S, double2 mn, double2 rh, int *active, int commit) {
    if (!commit) {
        S.rz = mn.x + mn.y + rh.x;   // a field the even-odd iteration never reads
        return;
    }
    S.iter += 1;
    const PfStep d = pf_decide(mn.x, mn.y, rh.x, S.iter, S);
    S.rr_prev = S.rr;
    S.rr = rh.x;
    if (d.stop) {
        S.done = 1;
        atomicSub(active, 1);
        return;
    }
    S.alpha = make_double2(d.alpha, 0.0);
    S.beta = d.beta;
}

// the separate finalize of one iteration (one block per system, after a finMode 0
// launch); commit = false only reduces (timing)
template <int NT>
__global__ void __launch_bounds__(NT)
k_fin_pf(const double2 *__restrict__ part, int nb, CgState *__restrict__ st, int *__restrict__ active, int commit) {
    __shared__ double2 lds0[NT / kWave], lds1[NT / kWave];
    CgState &S = st[blockIdx.x];
    const int skip = S.done;
    double2 mn, rh;
    sum_pf_partials<NT>(part + 2 * (size_t)blockIdx.x * nb, nb, lds0, lds1, mn, rh);
    if (skip || threadIdx.x) return;
    pf_commit(S, mn, rh, active, commit);
}

// the same with one wave per system (systems with at most kNT block partials: one or a
// few loads per lane and a shuffle reduction, no LDS or barriers)
__global__ void __launch_bounds__(kNT)
k_fin_pf_wave(const double2 *__restrict__ part, int nb, int nsys, CgState *__restrict__ st, int *__restrict__ active,
              int commit) {
    const int sys = blockIdx.x * (kNT / kWave) + (int)threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (sys >= nsys) return;
    CgState &S = st[sys];
    const int skip = S.done;
    const double2 *pp = part + 2 * (size_t)sys * nb;
    double2 a0 = make_double2(0.0, 0.0), a1 = make_double2(0.0, 0.0);
    for (int k = lane; k < nb; k += kWave) {
        a0 = cadd(a0, pp[2 * k]);
        a1 = cadd(a1, pp[2 * k + 1]);
    }
    const double2 mn = make_double2(wave_sum(a0.x), wave_sum(a0.y));
    const double2 rh = make_double2(wave_sum(a1.x), wave_sum(a1.y));
    if (skip || lane) return;
    pf_commit(S, mn, rh, active, commit);
}

// Finalize at the end of a chunk of n launches that finalized each other in their
// prologues (finMode 1): launch n - 1's partials in slotLast, launch n - 2's in slotPrev.
// Systems that paused inside the chunk already hold their state.
template <int NT>
__global__ void __launch_bounds__(NT)
k_fin_pf_chunk(const double2 *__restrict__ slotLast, const double2 *__restrict__ slotPrev, int nb, int n,
               CgState *__restrict__ st, int *__restrict__ active) {
    __shared__ double2 lds0[NT / kWave], lds1[NT / kWave];
    CgState &S = st[blockIdx.x];
    if (S.done) return;
    double2 mn, rh, mp, rp;
    sum_pf_partials<NT>(slotLast + 2 * (size_t)blockIdx.x * nb, nb, lds0, lds1, mn, rh);
    __syncthreads();
    sum_pf_partials<NT>(slotPrev + 2 * (size_t)blockIdx.x * nb, nb, lds0, lds1, mp, rp);
    if (threadIdx.x) return;
    S.iter = S.iter0 + n;
    const PfStep d = pf_decide(mn.x, mn.y, rh.x, S.iter, S);
    S.rr_prev = rp.x;
    S.rr = rh.x;
    if (d.stop) {
        S.done = 1;
        atomicSub(active, 1);
    }
}

// ---------------------------------------------------------------------------
// fused Schur normal operator: p' = r + beta p_old; v = S p'; Ap = S^T v; |v|^2
// ---------------------------------------------------------------------------
// One block owns a TT x TX tile of the full lattice (TT, TX even) and produces Ap on the
// tile's TT*TX/2 even sites.  LDS holds the tile plus a 4-site halo, (TT+8) x (TX+8)
// sites in full-lattice coordinates, split by lattice parity: E[c][row][k] and
// O[c][row][k] hold the site of that parity in array row `row`, column 2k or 2k+1
// (k = column >> 1), four SoA planes each.  Lanes of a wave take consecutive k, so every
// neighbour read (rows row +- 1 at the same k, the same row at k or k +- 1) is a
// contiguous run: no bank conflicts.  Each stage fills the parity it produces on a halo
// one site smaller than the stage before:
//   0: p'  on even sites, 4-site halo (from HBM/L2; p' of the tile also written out)
//   1: U = K p' / d on odd sites, 3-site halo          (= A_o^-1 K_oe p')
//   2: v = d p' - K U on even sites, 2-site halo        (= S p', overwrites p')
//   3: W = H v / d on odd sites, 1-site halo            (overwrites U)
//   4: Ap = d v - H W on the tile's even sites          (= S^T S p')
// The diagonal d = 2r + M + g phi is staged per parity for the whole array.  |v|^2 over
// the tile (= <p', Ap>) is the block's partial.  HBM traffic per launch: read r, p_old
// (Vh sites each) and phi, write p', Ap: 4 S/2 + P per site of the full lattice.
constexpr int eo_threads(int tt, int tx) { return (tt * tx / 2 < kWave) ? kWave : tt * tx / 2; }

// PF = false: the operator alone, p' = r + beta p_old, partial <p', Ap> = |v|^2.
// PF = true, the pipelined FP32 iteration with one reduction.  CG with the residual-norm
// recurrence of D'Azevedo, Eijkhout and Romine (LAPACK Working Note 56, 1993): with
// rho = |r|^2, mu = <p, Ap>, nu = |Ap|^2, alpha_k = rho_k / mu_k and <r_k, Ap_k> = mu_k,
// rho_{k+1} = alpha_k^2 nu_k - rho_k, so beta_{k+1} is known before r_{k+1} exists and
// the x, r updates move into the next operator application.  Stage 0 then forms, on its
// even sites, r' = r - alpha Ap, p' = r' + beta p and (tile only) the x_s update; the block
// partials are (mu, nu) and (rho', 0), rho' the exact |r'|^2 that re-anchors the
// recurrence every iteration.  Ap ping-pongs between two buffers and the directions rotate
// through three: blocks read the old ones on their halo while other blocks write the new
// ones on their tiles.
// PERSIST = true: the grid holds only as many blocks as are resident; each block walks
// the tiles blockIdx.x, blockIdx.x + gridDim.x, ... and loads the next tile's stage-0
// data into registers while it runs stages 1-4 of the current one, so global loads
// overlap the LDS stages.  One partial per block (its tiles in order).
// finMode (PF only): 0 = alpha and beta from the state, partials to part (k_fin_pf
// follows).  1 = launch idx of a chunk finalizes launch idx - 1 in its prologue: every
// block sums that launch's partials (prevPart) in the same fixed order and derives the
// same alpha, beta and stop decision, so no finalize launch runs between iterations; a
// stopping launch's block 0 records the pause and every block returns.  Launch 0 of a
// chunk starts from the state, k_fin_pf_chunk closes the chunk, and the state is not
// written otherwise during a chunk (blocks read iter0, not iter).  The per-block reads
// of prevPart suit grids of resident blocks (PERSIST).  2 = the prologue's reductions
// without pausing or writing the state (the autotuner's timing launches).
// RU = true (FP64, PF = false): the reliable update's residual instead of an iteration.
// Stage 0 stages x (r holds x, beta = 0), stage 2 forms res = b - S x with b (Ap_in) read
// at its own sites, stage 4 writes r = S^T res (Ap) and its FP32 copy (out32); the
// partials are |res|^2 (part) and |r|^2 (part2).  Systems already final are skipped.
// WR1 = true: Wilson parameter r = 1 (hop_flavour_r1), chosen at launch when A.r == 1 (every
// precision and mode).
// The pipelined iteration keeps neither r nor x_s in every launch.  Block 0 of each launch
// records the alpha, beta it used (from the prologue, or from the state that the separate
// finalize wrote) in the state as ab[idx & 1] for the next launch, and the directions
// rotate through three buffers.
// RM (PF only), where r_{idx-1} comes from:
//   1: read from r (launch 0 of a chunk, after the setup or a reliable update has put the
//      residual there);
//   2: rebuilt as r_{idx-1} = p_{idx-1} - beta_{idx-1} p_{idx-2}, the relation that formed
//      p_{idx-1} (same FP32 beta); r then carries p_{idx-2}, read in place of the residual.
//      The rebuild's rounding error is a few eps |p| against the eps |r| of a stored r
//      (|p| is a few |r| in CG), and the FP64 reliable updates bound its drift as they do
//      the recursion's.
// XM (PF only), the x_s update:
//   0: none (an even launch);
//   2: the odd launch after it adds both terms, alpha_e p_{idx-2} + alpha p_{idx-1}, with
//      p_{idx-2} the input RM = 2 reads anyway.
// Pairs never straddle a chunk (chunks have an even number of launches), and a system
// pauses only after an odd launch (after whole pairs), except at the iteration cap; the
// reliable update's k_pf_reliable then adds the deferred term.  Per even site and launch:
// 168 B (even) and 232 B (odd).  XM and RM do not apply to the FP64 forms (PF = false).
template <typename R, int TT, int TX, bool PF, bool PERSIST = false, bool RU = false, bool WR1 = false, int XM = 1,
          int RM = 0>
__global__ void __launch_bounds__(eo_threads(TT, TX))
k_eo_normal_fused(EoArgs A, const R *__restrict__ phi, CgState *st, const typename Vec2<R>::T *__restrict__ r,
                  const typename Vec2<R>::T *__restrict__ Ap_in, const typename Vec2<R>::T *__restrict__ pold,
                  typename Vec2<R>::T *__restrict__ pnew, typename Vec2<R>::T *__restrict__ Ap,
                  typename Vec2<R>::T *__restrict__ xs, double2 *__restrict__ part, int finMode, int idx,
                  const double2 *__restrict__ prevPart, int *active, float2 *__restrict__ out32,
                  double2 *__restrict__ part2) {
    static_assert(!(PF && RU), "the reliable-update residual runs in FP64 only");
    static_assert(!PF || ((XM == 0 || XM == 2) && (RM == 1 || RM == 2)), "pipelined forms: XM 0, 2 and RM 1, 2");
    static_assert(XM != 2 || RM == 2, "the pair's x_s update reads p_{idx-2}");
    using V2 = typename Vec2<R>::T;
    constexpr int H = TT + 8, K = (TX + 8) / 2, NT = eo_threads(TT, TX);
    constexpr int M0 = (H * K + NT - 1) / NT;
    // O and FO hold only array rows 1 .. H - 2 (odd sites are needed on the 3-site halo at
    // most): odd-site entry (row, k) sits at (row - 1) K + k, 7-8 % less LDS per block
    constexpr int HO = H - 2;
    __shared__ V2 E[4][H * K];
    __shared__ V2 O[4][HO * K];
    __shared__ R FE[H * K], FO[HO * K];
    __shared__ double2 red[NT / kWave > 0 ? NT / kWave : 1];
    __shared__ double2 red2[NT / kWave > 0 ? NT / kWave : 1];

    const int sys = blockIdx.y;
    if (RU ? st[sys].final_ : st[sys].done) return;
    const int nTiles = A.tilesPerSys;
    const int Nt = A.Nt, Nx = A.Nx, Nh = A.Nh;
    const unsigned Vh = (unsigned)A.Vh;
    const size_t off = (size_t)sys * 4 * A.Vh;   // this system's vectors
    r += off;
    if (!RU) {
        pold += off;
        pnew += off;
    }
    Ap += off;
    if (PF) xs += off;
    if (PF || RU) Ap_in += off;
    if (RU) out32 += off;
    R beta = RU ? (R)0 : (R)st[sys].beta;
    R alpha = PF ? (R)st[sys].alpha.x : (R)0;
    double alphaD = PF ? st[sys].alpha.x : 0.0, betaD = RU ? 0.0 : st[sys].beta;   // recorded in ab[idx & 1]
    // RM = 2, XM = 2: the alpha, beta of launch idx - 1
    const double2 prev = (PF && RM == 2) ? st[sys].ab[(idx - 1) & 1] : make_double2(0.0, 0.0);
    const R betaPrev = (R)prev.y, alphaPrev = (R)prev.x;
    const R *ph = phi + (size_t)(sys / A.nNoise) * A.V;
    const R hp = (R)(0.5 * (A.r + 1.0)), hm = (R)(0.5 * (A.r - 1.0)), hr = (R)(0.5 * A.r), h = (R)0.5;
    const R d0 = (R)(2.0 * A.r + A.M), g = (R)A.g;

    // tile origin of tile number t (contiguous bands of tiles per XCD)
    auto origin = [&](int t, int &t0, int &x0) {
        if (A.xcdRemap) t = (t % 8) * (nTiles / 8) + t / 8;
        const int ty = t / A.tilesX, tx = t - ty * A.tilesX;
        t0 = ty * TT;
        x0 = tx * TX;
    };

    // stage-0 inputs of one tile, loaded into registers: a thread handles M0 sites of the
    // array's even sites.  Array (row, col) is lattice site (t0 - 4 + row, x0 - 4 + col);
    // its parity is (row + col) & 1 because t0 + x0 is even.
    struct Regs {
        V2 ra[M0][4], pa[M0][4], aa[M0][4], xa[M0][4];
        R fe[M0], fo[M0];
        unsigned jj[M0];
    };
    auto load0 = [&](int t0, int x0, Regs &G) {
#pragma unroll
        for (int u = 0; u < M0; ++u) {
            const int w = threadIdx.x + u * NT;
            if (w < H * K) {
                const int row = w / K, k = w - row * K;
                const int gt = wrapf(t0 - 4 + row, Nt, A.invNt);
                const int colE = 2 * k + (row & 1), colO = 2 * k + 1 - (row & 1);
                const int gxE = wrapf(x0 - 4 + colE, Nx, A.invNx), gxO = wrapf(x0 - 4 + colO, Nx, A.invNx);
                G.jj[u] = (unsigned)(gt * Nh + (gxE >> 1));
                const bool inner = row >= 4 && row < 4 + TT && colE >= 4 && colE < 4 + TX;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const unsigned i = c * Vh + G.jj[u];
                    G.ra[u][c] = at(r, i);
                    if (!RU) G.pa[u][c] = at(pold, i);
                    if (PF) G.aa[u][c] = at(Ap_in, i);
                    if (PF && XM != 0 && inner) G.xa[u][c] = at(xs, i);
                }
                G.fe[u] = at(ph, (unsigned)(gt * Nx + gxE));
                G.fo[u] = at(ph, (unsigned)(gt * Nx + gxO));
            }
        }
    };
    // stage 0 proper: p' (PF: r' = r - alpha Ap, x_s += alpha p) into LDS and, on the
    // tile, to global memory; d on both parities
    double rho = 0.0;
    auto compute0 = [&](const Regs &G) {
#pragma unroll
        for (int u = 0; u < M0; ++u) {
            const int w = threadIdx.x + u * NT;
            if (w < H * K) {
                const int row = w / K, k = w - row * K;
                const int colE = 2 * k + (row & 1);
                const bool inner = row >= 4 && row < 4 + TT && colE >= 4 && colE < 4 + TX;
                R rs = 0;   // |r'|^2 of the site (4 terms in R, then accumulated in FP64)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const unsigned i = c * Vh + G.jj[u];
                    V2 a = G.ra[u][c];   // r_{idx-1}, or p_{idx-2} (RM = 2)
                    const V2 b = RU ? a : G.pa[u][c];
                    if (PF && RM == 2) a = mk2<R>(b.x - betaPrev * a.x, b.y - betaPrev * a.y);
                    if (PF) a = mk2<R>(a.x - alpha * G.aa[u][c].x, a.y - alpha * G.aa[u][c].y);
                    const V2 pv = RU ? a : mk2<R>(a.x + beta * b.x, a.y + beta * b.y);
                    E[c][w] = pv;
                    if (inner && !RU) {
                        at(pnew, i) = pv;
                        if (PF) {
                            const V2 xa = G.xa[u][c], p2 = G.ra[u][c];
                            if (XM == 2)
                                at(xs, i) = mk2<R>(xa.x + alphaPrev * p2.x + alpha * b.x,
                                                   xa.y + alphaPrev * p2.y + alpha * b.y);
                            rs += a.x * a.x + a.y * a.y;
                        }
                    }
                }
                if (PF && inner) rho += (double)rs;
                FE[w] = d0 + g * G.fe[u];
                if (row >= 1 && row <= HO) FO[w - K] = d0 + g * G.fo[u];
            }
        }
    };

    // the sites of parity `par` in rows/cols [4 - hh, 4 + T + hh): w -> (row, k) and the
    // array indices of the site and its four neighbours (other parity)
    struct Site {
        int row, col, i, itf, itb, ixf, ixb;
    };
    auto site_of = [](int w, int hh, int par) {
        constexpr int Kw = K;
        const int c0 = 4 - hh, half = (TX + 2 * hh) / 2;
        Site q;
        q.row = c0 + w / half;
        q.col = c0 + 2 * (w - (q.row - c0) * half) + (((par + q.row) ^ c0) & 1);
        const int k = q.col >> 1;
        q.i = q.row * Kw + k;
        q.itf = q.i + Kw;
        q.itb = q.i - Kw;
        q.ixf = q.row * Kw + ((q.col + 1) >> 1);
        q.ixb = q.row * Kw + ((q.col - 1) >> 1);
        return q;
    };

    double acc = 0.0, nu = 0.0;
    // stages 1-4 of the tile at (t0, x0); stage 0 must be in LDS
    auto stages = [&](int t0, int x0) {
        // hopping at site q from the other-parity array N (component stride NS, entries
        // shifted down by SH: 0 for E, K for O)
        auto hop_at = [&](auto dag, auto ns, auto sh, const Site &q, const V2 *N, V2 o[4]) {
            constexpr bool DAG = decltype(dag)::value;
            constexpr int NS = decltype(ns)::value, SH = decltype(sh)::value;
            const int gt = wrapf(t0 - 4 + q.row, Nt, A.invNt);
            const R sf = (gt == Nt - 1) ? (R)-1 : (R)1, sb = (gt == 0) ? (R)-1 : (R)1;
            const int itf = q.itf - SH, itb = q.itb - SH, ixf = q.ixf - SH, ixb = q.ixb - SH;
#pragma unroll
            for (int f = 0; f < 4; f += 2) {
                const V2 *N0 = N + f * NS, *N1 = N + (f + 1) * NS;
                if constexpr (WR1)
                    hop_flavour_r1<DAG, R, V2>(N0[itf], N1[itf], sf, N0[itb], N1[itb], sb, N0[ixf], N1[ixf], N0[ixb],
                                               N1[ixb], o[f], o[f + 1]);
                else
                    hop_flavour_r<DAG, R, V2>(hp, hm, hr, h, N0[itf], N1[itf], sf, N0[itb], N1[itb], sb, N0[ixf],
                                              N1[ixf], N0[ixb], N1[ixb], o[f], o[f + 1]);
            }
        };
        using FromE = std::integral_constant<int, H * K>;
        using FromO = std::integral_constant<int, HO * K>;
        using Sh0 = std::integral_constant<int, 0>;
        using ShK = std::integral_constant<int, K>;
        // stage 1: U = K p' / d on odd sites, 3-site halo
        for (int w = threadIdx.x; w < (TT + 6) * ((TX + 6) / 2); w += NT) {
            const Site q = site_of(w, 3, 1);
            V2 o[4];
            hop_at(std::true_type(), FromE(), Sh0(), q, &E[0][0], o);
            const R inv = (R)1 / FO[q.i - K];
#pragma unroll
            for (int c = 0; c < 4; ++c) O[c][q.i - K] = mk2<R>(o[c].x * inv, o[c].y * inv);
        }
        __syncthreads();
        // stage 2: v = d p' - K U on even sites, 2-site halo; |v|^2 over the tile
        for (int w = threadIdx.x; w < (TT + 4) * ((TX + 4) / 2); w += NT) {
            const Site q = site_of(w, 2, 0);
            V2 o[4];
            hop_at(std::true_type(), FromO(), ShK(), q, &O[0][0], o);
            const R d = FE[q.i];
            const bool inner = q.row >= 4 && q.row < 4 + TT && q.col >= 4 && q.col < 4 + TX;
            unsigned jb = 0;   // RU: this site's entry of b
            if (RU)
                jb = (unsigned)(wrapf(t0 - 4 + q.row, Nt, A.invNt) * Nh + (wrapf(x0 - 4 + q.col, Nx, A.invNx) >> 1));
            R vs = 0;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const V2 pv = E[c][q.i];
                V2 v = mk2<R>(d * pv.x - o[c].x, d * pv.y - o[c].y);
                if (RU) {
                    const V2 bb = at(Ap_in, c * Vh + jb);
                    v = mk2<R>(bb.x - v.x, bb.y - v.y);
                }
                E[c][q.i] = v;
                vs += v.x * v.x + v.y * v.y;
            }
            if (inner) acc += (double)vs;
        }
        __syncthreads();
        // stage 3: W = H v / d on odd sites, 1-site halo
        for (int w = threadIdx.x; w < (TT + 2) * ((TX + 2) / 2); w += NT) {
            const Site q = site_of(w, 1, 1);
            V2 o[4];
            hop_at(std::false_type(), FromE(), Sh0(), q, &E[0][0], o);
            const R inv = (R)1 / FO[q.i - K];
#pragma unroll
            for (int c = 0; c < 4; ++c) O[c][q.i - K] = mk2<R>(o[c].x * inv, o[c].y * inv);
        }
        __syncthreads();
        // stage 4: Ap = d v - H W on the tile's even sites (PF: and |Ap|^2)
        for (int w = threadIdx.x; w < TT * (TX / 2); w += NT) {
            const Site q = site_of(w, 0, 0);
            V2 o[4];
            hop_at(std::false_type(), FromO(), ShK(), q, &O[0][0], o);
            const R d = FE[q.i];
            // stage-4 sites are the tile's own: no wrap
            const unsigned j = (unsigned)((t0 - 4 + q.row) * Nh + ((x0 - 4 + q.col) >> 1));
            R as = 0;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const V2 v = E[c][q.i];
                const V2 a = mk2<R>(d * v.x - o[c].x, d * v.y - o[c].y);
                at(Ap, c * Vh + j) = a;
                if (RU) at(out32, c * Vh + j) = make_float2((float)a.x, (float)a.y);
                as += a.x * a.x + a.y * a.y;
            }
            if (PF || RU) nu += (double)as;
        }
    };

    // finMode 1, 2: finalize launch idx - 1 (its iteration is number iter0 + idx); false:
    // the system pauses here
    auto prologue = [&]() -> bool {
        if (!PF || finMode == 0 || idx == 0) return true;
        __shared__ double bc[3];   // alpha, beta, rho of launch idx - 1
        __shared__ int pause;
        const int nb = (int)gridDim.x;
        double2 mn, rh;
        sum_pf_partials<NT>(prevPart + 2 * (size_t)sys * nb, nb, red, red2, mn, rh);
        if (threadIdx.x == 0) {
            const PfStep d = pf_decide(mn.x, mn.y, rh.x, st[sys].iter0 + idx, st[sys]);
            bc[0] = d.alpha;
            bc[1] = d.beta;
            bc[2] = rh.x;
            pause = finMode == 1 && d.stop;
        }
        __syncthreads();
        if (!pause) {
            alpha = (R)bc[0];
            beta = (R)bc[1];
            alphaD = bc[0];
            betaD = bc[1];
            return true;
        }
        if (blockIdx.x == 0) {   // record the pause; rho of the iteration before it
            double rprev = st[sys].rr;   // iteration iter0 (chunk start) when idx == 1
            if (idx >= 2) {              // launch idx - 2's partials, in this launch's slot
                double2 m2, r2;
                sum_pf_partials<NT>(part + 2 * (size_t)sys * nb, nb, red, red2, m2, r2);
                rprev = r2.x;
            }
            if (threadIdx.x == 0) {
                CgState &S = st[sys];
                S.iter = S.iter0 + idx;
                S.rr_prev = rprev;
                S.rr = bc[2];
                S.done = 1;
                atomicSub(active, 1);
            }
        }
        return false;
    };

    // once alpha is fixed, a pipelined launch records its alpha, beta for the next one
    auto record_ab = [&]() {
        if constexpr (PF) {
            if (blockIdx.x == 0 && threadIdx.x == 0) st[sys].ab[idx & 1] = make_double2(alphaD, betaD);
        }
    };
    Regs G;
    int t0, x0;
    if (!PERSIST) {
        origin(blockIdx.x, t0, x0);
        load0(t0, x0, G);
        if (!prologue()) return;
        record_ab();
        compute0(G);
        __syncthreads();
        stages(t0, x0);
    } else {
        int t = blockIdx.x;
        origin(t, t0, x0);
        load0(t0, x0, G);
        if (!prologue()) return;
        record_ab();
        while (t < nTiles) {
            compute0(G);
            __syncthreads();
            const int tn = t + (int)gridDim.x;
            int n0 = 0, nx0 = 0;
            if (tn < nTiles) {   // next tile's loads in flight during stages 1-4
                origin(tn, n0, nx0);
                load0(n0, nx0, G);
            }
            stages(t0, x0);
            __syncthreads();     // stage 4 reads of E and O precede the next stage 0
            t = tn;
            t0 = n0;
            x0 = nx0;
        }
    }
    const size_t pidx = (size_t)sys * gridDim.x + blockIdx.x;
    if (PF) {
        const double2 t1 = block_sum2<NT>(make_double2(acc, nu), red);
        const double2 t2 = block_sum2<NT>(make_double2(rho, 0.0), red2);
        if (threadIdx.x == 0) {
            part[2 * pidx] = t1;
            part[2 * pidx + 1] = t2;
        }
    } else {
        const double2 tot = block_sum2<NT>(make_double2(acc, RU ? nu : 0.0), red);
        if (threadIdx.x == 0) {
            part[pidx] = make_double2(tot.x, 0.0);
            if (RU) part2[pidx] = make_double2(tot.y, 0.0);
        }
    }
}

// one wave per system for few partials, 1024 threads once a system has more partials
// than four rounds of 256 threads cover
static void launch_fin_pf(Ctx &c, int nsys, int nb, int commit) {
    if (nb <= kNT)
        hipLaunchKernelGGL(k_fin_pf_wave, dim3((nsys + kNT / kWave - 1) / (kNT / kWave)), dim3(kNT), 0, c.stream,
                           (const double2 *)c.partials, nb, nsys, c.cg_d, c.active, commit);
    else if (nb > 4 * kNT)
        hipLaunchKernelGGL(k_fin_pf<1024>, dim3(nsys), dim3(1024), 0, c.stream, (const double2 *)c.partials, nb,
                           c.cg_d, c.active, commit);
    else
        hipLaunchKernelGGL(k_fin_pf<kNT>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.partials, nb,
                           c.cg_d, c.active, commit);
}

// reliable update of the pipelined iteration: x += x_s (FP64), x_s = 0, and the last
// direction applied copied to p[0], where the next chunk's launch 0 reads it.  After m
// launches (m counted from the state per system) that is p_{m-2}, in rotation slot
// (m - 1) mod 3 of p[0], p[1], p2.  For odd m (a stop at the iteration cap after an even
// launch) the even launch's deferred term alpha_{m-1} p_{m-2} is added too, with the alpha
// it recorded in ab.
__global__ void __launch_bounds__(kNT)
k_pf_reliable(int n, const CgState *__restrict__ st, double2 *__restrict__ x, float2 *__restrict__ xs,
              const float2 *__restrict__ p1, const float2 *__restrict__ p2, float2 *__restrict__ p0) {
    const int sys = blockIdx.y;
    if (st[sys].final_) return;
    const int s = blockIdx.x * kNT + threadIdx.x;
    if (s >= n) return;
    const int m = st[sys].iter - st[sys].iter0, slot = m >= 1 ? (m - 1) % 3 : 0;
    const float2 *src = slot == 1 ? p1 : slot == 2 ? p2 : p0;
    const double a = (m & 1) ? st[sys].ab[(m - 1) & 1].x : 0.0;
    const size_t off = (size_t)sys * 4 * n;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const size_t i = off + (size_t)c * n + s;
        const float2 d = xs[i];
        const double2 xv = x[i];
        double2 xn = make_double2(xv.x + (double)d.x, xv.y + (double)d.y);
        const float2 pk = (slot != 0 || (m & 1)) ? src[i] : make_float2(0.f, 0.f);
        if (m & 1) xn = make_double2(xn.x + a * (double)pk.x, xn.y + a * (double)pk.y);
        x[i] = xn;
        xs[i] = make_float2(0.f, 0.f);
        if (slot != 0) p0[i] = pk;
    }
}

// after the reliable update of the pipelined iteration: partRes = |b_e - S x_e|^2 and
// partN = |S^T (b_e - S x_e)|^2 (the new normal residual).  Converged on the system residual, or resume with alpha = 0 (r already is the true
// residual) and beta = |r_true|^2 / |r_prev|^2 (the direction is kept)
__global__ void __launch_bounds__(kNT)
k_fin_eo_pf_true(const double2 *__restrict__ partRes, const double2 *__restrict__ partN, int nb,
                 CgState *__restrict__ st, int *__restrict__ active) {
    __shared__ double2 lds0[kNT / kWave], lds1[kNT / kWave];
    const int sys = blockIdx.x;
    const int skip = st[sys].final_;
    const double2 res = eo_fin_sum(partRes + (size_t)sys * nb, nb, lds0);
    const double2 rn = eo_fin_sum(partN + (size_t)sys * nb, nb, lds1);
    if (skip || threadIdx.x) return;
    CgState &S = st[sys];
    S.true_rel = S.enorm2 > 0.0 ? sqrt(res.x / S.enorm2) : sqrt(res.x);
    S.restarts += 1;
    if (S.nRu < kMaxRounds) S.ruIter[S.nRu] = S.iter;
    S.nRu += 1;
    // paused at the stop (not for the drop since the last anchor): the normal-equation
    // threshold the true residual asks for
    if (S.done && sqrt(S.rr) <= S.stop && S.true_rel > 0.0) S.need = sqrt(rn.x) * S.tol / S.true_rel;
    const bool conv = (S.mode == YCN_TOL_REL_TRUE) ? (S.true_rel <= S.tol) : (sqrt(rn.x) <= S.stop);
    if (conv || S.iter >= S.maxIter) {
        S.converged = conv ? 1 : 0;
        S.final_ = 1;
        if (!S.done) {
            S.done = 1;
            atomicSub(active, 1);
        }
        return;
    }
    S.alpha = make_double2(0.0, 0.0);
    S.beta = (S.rr_prev > 0.0) ? rn.x / S.rr_prev : 0.0;
    S.rr = rn.x;
    S.rrRel = rn.x;
    S.iter0 = S.iter;   // the next chunk's launches count from here
    if (S.mode == YCN_TOL_REL_TRUE && sqrt(rn.x) <= S.stop) S.stop = sqrt(rn.x) * 0.5 * S.tol / S.true_rel;
    if (S.done) {
        S.done = 0;
        atomicAdd(active, 1);
    }
}

// graph-captured Langevin steps: start of a step.  The count of active systems of its
// solve, and the halt flag and counter the step's coloured noise reads, copied into a
// second block (noise): the noise runs on a branch of the graph beside the solve, whose
// check kernel may change the first block while the noise kernels run.
__global__ void k_step_begin(const StepCtl *__restrict__ ctl, StepCtl *__restrict__ noise, int *__restrict__ active,
                             int nsys) {
    if (threadIdx.x) return;
    if (!ctl->halt) *active = nsys;
    noise->halt = ctl->halt;
    noise->step = ctl->step;
}

// graph-captured Langevin steps: after the step's chunk schedule.  Every system final:
// the step's statistics into the control block, the next solve's starting threshold
// (relative mode: 0.9 times the largest ratio need / stop0 the systems reported, as
// eo_learn_stop_scale does on the host) and the next step's counter.  A system still
// iterating halts the run at this step: the drift, the update and every later step's
// kernels outside the chunk schedule return at once.
__global__ void k_step_check(StepCtl *__restrict__ ctl, const CgState *__restrict__ st, int nsys, int relMode) {
    if (threadIdx.x || ctl->halt) return;
    int itMax = -1, kMax = 0, rsMax = 0, conv = 1;
    double itSum = 0.0, rrMax = 0.0, trMax = 0.0, ratio = 0.0;
    for (int k = 0; k < nsys; ++k) {
        const CgState &S = st[k];
        if (!S.final_) {
            ctl->halt = 1;
            ctl->haltStep = ctl->step;
            return;
        }
        if (S.iter > itMax) {
            itMax = S.iter;
            kMax = k;
        }
        rsMax = max(rsMax, S.restarts);
        conv &= S.converged;
        itSum += S.iter;
        rrMax = fmax(rrMax, S.rr);
        trMax = fmax(trMax, S.true_rel);
        if (S.need > 0.0 && S.stop0 > 0.0) ratio = fmax(ratio, S.need / S.stop0);
    }
    ctl->itersMaxSum += itMax;
    ctl->itersMeanSum += itSum / nsys;
    ctl->restarts += rsMax;
    ctl->converged &= conv;
    ctl->rrMax = fmax(ctl->rrMax, rrMax);
    ctl->trueRelMax = fmax(ctl->trueRelMax, trMax);
    ctl->lastIters = itMax;
    ctl->nRounds = min(st[kMax].nRu, kMaxRounds);
    for (int i = 0; i < ctl->nRounds; ++i) ctl->rounds[i] = st[kMax].ruIter[i];
    if (relMode && ratio > 0.0) ctl->stopScale = fmin(1.0, fmax(1e-3, 0.9 * ratio));
    ctl->steps += 1;
    ctl->step += 1;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// at least 64 sites per tile: the partials buffer holds 2 V/64 entries per system.
// FP32 entries with persist = true launch only the resident blocks, each walking several
// tiles with the next tile's loads in flight.
struct EoTileCfg {
    int TT, TX;
    bool persist;
};
constexpr EoTileCfg kEoTiles64[] = {{8, 32, false}, {16, 16, false}, {4, 64, false}, {4, 32, false}, {2, 32, false}};
constexpr EoTileCfg kEoTiles32[] = {{8, 32, false},  {16, 32, false}, {8, 64, false},  {16, 16, false},
                                    {4, 64, false},  {2, 32, false},  {32, 32, false}, {16, 64, false},
                                    {8, 32, true},   {16, 32, true},  {16, 16, true},  {8, 64, true}};
constexpr int kNumEo64 = sizeof(kEoTiles64) / sizeof(kEoTiles64[0]);
constexpr int kNumEo32 = sizeof(kEoTiles32) / sizeof(kEoTiles32[0]);

static EoArgs eo_args(const Ctx &c, int TT, int TX) {
    EoArgs A;
    A.Nt = c.p.Nt;
    A.Nx = c.p.Nx;
    A.Nh = c.p.Nx / 2;
    A.V = c.V;
    A.Vh = c.V / 2;
    A.nNoise = c.nNoise;
    A.r = c.p.r;
    A.M = c.p.M;
    A.g = c.p.g;
    A.tilesX = TX > 0 ? c.p.Nx / TX : 0;
    A.tilesPerSys = (TT > 0 && TX > 0) ? (c.p.Nt / TT) * (c.p.Nx / TX) : 0;
    A.xcdRemap = (A.tilesPerSys % 8 == 0) ? 1 : 0;
    A.invNt = 1.0f / (float)A.Nt;
    A.invNx = 1.0f / (float)A.Nx;
    A.halt = nullptr;
    A.stepPtr = nullptr;
    return A;
}

static bool eo_fits(const Ctx &c, EoTileCfg t) { return c.p.Nt % t.TT == 0 && c.p.Nx % t.TX == 0; }

// the two partial slots of prologue-finalized launches (launch idx writes slot idx & 1
// and reads the other), sized for the smallest tile (64 sites: at most V / 64 blocks)
int eo_pf_alloc(Ctx &c) {
    const size_t n = (size_t)c.nSys * ((c.V + 63) / 64) * 2;
    for (double2 *&slot : c.pfSlot) {
        YCN_HIP(hipMalloc(&slot, sizeof(double2) * n));
        YCN_HIP(hipMemset(slot, 0, sizeof(double2) * n));
    }
    return YCN_OK;
}

// even extents; the fused kernels address one system's four FP64 half-lattice
// components and phi through 32-bit byte offsets and wrap halo coordinates with wrapf
// (|a| < 2^22)
static bool eo_shape_ok(const Ctx &c) {
    return c.p.Nt % 2 == 0 && c.p.Nx % 2 == 0 && (size_t)4 * (c.V / 2) * sizeof(double2) < ((size_t)1 << 32) &&
           c.p.Nt < (1 << 21) && c.p.Nx < (1 << 21);
}

bool eo_available(const Ctx &c) { return eo_shape_ok(c) && c.tileEo64 >= 0; }

// buffers of one fused launch (PF: the pipelined iteration's Ap pair, directions and x_s)
template <typename R>
struct EoPtrs {
    using V2 = typename Vec2<R>::T;
    const V2 *r = nullptr;
    const V2 *Ap_in = nullptr;
    const V2 *pold = nullptr;
    V2 *pnew = nullptr, *Ap = nullptr, *xs = nullptr;
    double2 *part = nullptr;            // this launch's block partials
    const double2 *prevPart = nullptr;  // finMode 1, 2: the previous launch's
    int idx = 0;                        // launch index in the chunk
};

// resident blocks of a kernel on the whole device (persistent launches)
static int resident_blocks(const void *kernel, int threads) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess) return 0;
    return per * cus;
}

// returns the number of blocks per system (= partial slots per system)
// the pipelined FP32 iteration's launch (persistent or one block per tile), WR1 as in
// k_eo_normal_fused; returns the blocks per system
// XM, RM as in k_eo_normal_fused: (0, 1) for launch 0 of a chunk, then (2, 2) and (0, 2)
// alternating.  Persistent grids are sized by the form with the fewest resident blocks,
// so that consecutive launches leave the same number of partials.
template <int TT, int TX, bool WR1>
static int launch_eo_pf_tt(const Ctx &c, int nsys, const EoArgs &A, const float *phi, bool persist,
                           const EoPtrs<float> &b, int finMode) {
    const dim3 block(eo_threads(TT, TX));
    const int form = b.idx == 0 ? 0 : (b.idx & 1) ? 1 : 2;
    auto pick = [&](auto persistTag) {
        constexpr bool P = decltype(persistTag)::value;
        using K = decltype(&k_eo_normal_fused<float, TT, TX, true, P, false, WR1, 0, 1>);
        const K ks[3] = {&k_eo_normal_fused<float, TT, TX, true, P, false, WR1, 0, 1>,
                         &k_eo_normal_fused<float, TT, TX, true, P, false, WR1, 2, 2>,
                         &k_eo_normal_fused<float, TT, TX, true, P, false, WR1, 0, 2>};
        return std::make_pair(ks[form], ks);
    };
    if (persist) {
        const auto kk = pick(std::true_type());
        static const int resident = [&] {
            int m = 1 << 30;
            for (int i = 0; i < 3; ++i)
                m = std::min(m, resident_blocks(reinterpret_cast<const void *>(kk.second[i]), block.x));
            return m;
        }();
        int g = std::min(A.tilesPerSys, std::max(resident, 8));
        if (A.xcdRemap && g >= 8) g -= g % 8;   // keeps every block's tiles on one XCD's band
        hipLaunchKernelGGL(kk.first, dim3(g, nsys), block, 0, c.stream, A, phi, c.cg_d, b.r, b.Ap_in, b.pold,
                           b.pnew, b.Ap, b.xs, b.part, finMode, b.idx, b.prevPart, c.active, (float2 *)nullptr,
                           (double2 *)nullptr);
        return g;
    }
    hipLaunchKernelGGL(pick(std::false_type()).first, dim3(A.tilesPerSys, nsys), block, 0, c.stream, A, phi, c.cg_d,
                       b.r, b.Ap_in, b.pold, b.pnew, b.Ap, b.xs, b.part, finMode, b.idx, b.prevPart, c.active,
                       (float2 *)nullptr, (double2 *)nullptr);
    return A.tilesPerSys;
}

template <typename R, int TT, int TX>
static int launch_eo_tt(const Ctx &c, int nsys, const EoArgs &A, const R *phi, bool pf, bool persist,
                        const EoPtrs<R> &b, int finMode = 0) {
    const dim3 block(eo_threads(TT, TX));
    if constexpr (std::is_same<R, float>::value) {
        if (pf) {
            return A.r == 1.0 ? launch_eo_pf_tt<TT, TX, true>(c, nsys, A, phi, persist, b, finMode)
                              : launch_eo_pf_tt<TT, TX, false>(c, nsys, A, phi, persist, b, finMode);
        }
    }
    if (A.r == 1.0)
        hipLaunchKernelGGL((k_eo_normal_fused<R, TT, TX, false, false, false, true>), dim3(A.tilesPerSys, nsys), block,
                           0, c.stream, A, phi, c.cg_d, b.r, b.Ap_in, b.pold, b.pnew, b.Ap, b.xs, c.partials, 0,
                           0, (const double2 *)nullptr, (int *)nullptr, (float2 *)nullptr, (double2 *)nullptr);
    else
        hipLaunchKernelGGL((k_eo_normal_fused<R, TT, TX, false>), dim3(A.tilesPerSys, nsys), block, 0, c.stream, A,
                           phi, c.cg_d, b.r, b.Ap_in, b.pold, b.pnew, b.Ap, b.xs, c.partials, 0, 0,
                           (const double2 *)nullptr, (int *)nullptr, (float2 *)nullptr, (double2 *)nullptr);
    return A.tilesPerSys;
}

// launch the fused Schur normal operator (FP64) with tile `ti`; returns the partial count
static int launch_eo_fused(Ctx &c, int ti, int nsys, const double *phi, const double2 *r, const double2 *pold,
                           double2 *pnew, double2 *Ap) {
    const EoTileCfg t = kEoTiles64[ti];
    const EoArgs A = eo_args(c, t.TT, t.TX);
    EoPtrs<double> b;
    b.r = r;
    b.pold = pold;
    b.pnew = pnew;
    b.Ap = Ap;
    prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (2 * 64 + 8));
    switch (ti) {
    case 0: (void)launch_eo_tt<double, 8, 32>(c, nsys, A, phi, false, false, b); break;
    case 1: (void)launch_eo_tt<double, 16, 16>(c, nsys, A, phi, false, false, b); break;
    case 2: (void)launch_eo_tt<double, 4, 64>(c, nsys, A, phi, false, false, b); break;
    case 3: (void)launch_eo_tt<double, 4, 32>(c, nsys, A, phi, false, false, b); break;
    default: (void)launch_eo_tt<double, 2, 32>(c, nsys, A, phi, false, false, b); break;
    }
    prof_after(c, PK_NORMAL_OP);
    return A.tilesPerSys;
}

// the reliable update's residual in one fused FP64 launch (tile of the FP64 table):
// res = b - S x, r = S^T res (FP64 and FP32 copy), partials |res|^2 and |r|^2 of the
// systems not yet final; returns the partials per system
template <int TT, int TX>
static int launch_eo_ru_tt(Ctx &c, int nsys, const EoArgs &A, const double *phi, const double2 *x, const double2 *b,
                           double2 *r, float2 *r32, double2 *partRes, double2 *partN) {
    auto *k = A.r == 1.0 ? &k_eo_normal_fused<double, TT, TX, false, false, true, true>
                         : &k_eo_normal_fused<double, TT, TX, false, false, true, false>;
    hipLaunchKernelGGL(k, dim3(A.tilesPerSys, nsys), dim3(eo_threads(TT, TX)), 0, c.stream, A, phi, c.cg_d, x, b,
                       (const double2 *)nullptr, (double2 *)nullptr, r, (double2 *)nullptr, partRes, 0, 0,
                       (const double2 *)nullptr, (int *)nullptr, r32, partN);
    return A.tilesPerSys;
}
static int launch_eo_ru(Ctx &c, int nsys, const double *phi, const double2 *x, const double2 *b, double2 *r,
                        float2 *r32, double2 *partRes, double2 *partN) {
    const EoTileCfg t = kEoTiles64[c.tileEo64];
    const EoArgs A = eo_args(c, t.TT, t.TX);
    switch (c.tileEo64) {
    case 0: return launch_eo_ru_tt<8, 32>(c, nsys, A, phi, x, b, r, r32, partRes, partN);
    case 1: return launch_eo_ru_tt<16, 16>(c, nsys, A, phi, x, b, r, r32, partRes, partN);
    case 2: return launch_eo_ru_tt<4, 64>(c, nsys, A, phi, x, b, r, r32, partRes, partN);
    case 3: return launch_eo_ru_tt<4, 32>(c, nsys, A, phi, x, b, r, r32, partRes, partN);
    default: return launch_eo_ru_tt<2, 32>(c, nsys, A, phi, x, b, r, r32, partRes, partN);
    }
}

// direction p_k (k >= -1) of a chunk of pipelined launches: the three buffers in rotation,
// p_{-1} (the direction kept across the reliable update) in ps[0]
static float2 *eo_pdir(const Ctx &c, int k) {
    float2 *const rot[3] = {c.ps[0], c.ps[1], c.pfP2};
    return rot[(k + 1) % 3];
}

// the pipelined FP32 iteration's fused kernel (finMode as in k_eo_normal_fused) with
// tile `ti` as launch idx of a chunk (idx & 1: the Ap and partial-slot parity, the
// directions in rotation); returns the block partials per system
static int launch_eo_pf(Ctx &c, int ti, int nsys, int idx, int finMode) {
    const EoTileCfg t = kEoTiles32[ti];
    const EoArgs A = eo_args(c, t.TT, t.TX);
    const int par = idx & 1;
    EoPtrs<float> b;
    b.idx = idx;
    b.part = finMode ? c.pfSlot[par] : c.partials;
    b.prevPart = c.pfSlot[par ^ 1];
    b.Ap_in = par ? c.pfA1 : c.Aps;
    b.Ap = par ? c.Aps : c.pfA1;
    b.xs = c.xs;
    // direction p_k in eo_pdir(c, k); launch 0 reads r, the others p_{idx-2}
    b.r = idx == 0 ? c.rs : eo_pdir(c, idx - 2);
    b.pold = eo_pdir(c, idx - 1);
    b.pnew = eo_pdir(c, idx);
    // half-lattice FP32 vectors (16 B per site of the full lattice): read p_{idx-2} (or r),
    // Ap, p and write p, Ap, and x_s read and written in odd launches; and read phi
    const int vecs = (idx & 1) ? 7 : 5;
    prof_before(c, PK_NORMAL_OP, (double)nsys * c.V * (vecs * 16 + 4));
    const bool ps = t.persist;
    int nb = 0;
    switch (t.TT * 1000 + t.TX) {
    case 8032: nb = launch_eo_tt<float, 8, 32>(c, nsys, A, c.phis, true, ps, b, finMode); break;
    case 16032: nb = launch_eo_tt<float, 16, 32>(c, nsys, A, c.phis, true, ps, b, finMode); break;
    case 8064: nb = launch_eo_tt<float, 8, 64>(c, nsys, A, c.phis, true, ps, b, finMode); break;
    case 16016: nb = launch_eo_tt<float, 16, 16>(c, nsys, A, c.phis, true, ps, b, finMode); break;
    case 4064: nb = launch_eo_tt<float, 4, 64>(c, nsys, A, c.phis, true, ps, b, finMode); break;
    case 2032: nb = launch_eo_tt<float, 2, 32>(c, nsys, A, c.phis, true, ps, b, finMode); break;
    case 32032: nb = launch_eo_tt<float, 32, 32>(c, nsys, A, c.phis, true, ps, b, finMode); break;
    default: nb = launch_eo_tt<float, 16, 64>(c, nsys, A, c.phis, true, ps, b, finMode); break;
    }
    prof_after(c, PK_NORMAL_OP);
    return nb;
}

static dim3 half_grid(const Ctx &c, int nsys) { return dim3((c.V / 2 + kNT - 1) / kNT, nsys); }

// Times every even-odd tile that fits (FP64, and FP32 when the mixed buffers exist) on
// the context's buffers and keeps the fastest.  The trial launches only touch CG scratch.
int eo_autotune(Ctx &c) {
    c.tileEo64 = c.tileEo32 = -1;
    if (!eo_shape_ok(c)) return YCN_OK;
    hipEvent_t e0, e1;
    YCN_HIP(hipEventCreate(&e0));
    YCN_HIP(hipEventCreate(&e1));
    YCN_HIP(hipMemsetAsync(c.cg_d, 0, sizeof(CgState) * c.nSys, c.stream));
    const int reps = c.V >= (1 << 16) ? 10 : 3;
    auto time_it = [&](auto launch) -> double {   // launch(k), k = 0, 1, ... (k & 1: parity)
        launch(0);
        launch(1);
        (void)hipEventRecord(e0, c.stream);
        for (int i = 0; i < reps; ++i) launch(2 + i);
        (void)hipEventRecord(e1, c.stream);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return 1e3 * ms / reps;
    };
    char line[96];
    const bool pfOk = c.rs && c.phis && c.pfA1 && c.pfP2 && c.pfSlot[0] && c.pfSlot[1];
    // measured times: kNumEo64 FP64 tiles, then kNumEo32 x {separate, prologue} FP32
    // pipelined iterations (1e30 where a tile does not fit); reused from an earlier
    // context of the same shape when there is one
    std::vector<double> us;
    const bool memo = tune_lookup(1, c, us) && us.size() == (size_t)(kNumEo64 + 2 * kNumEo32);
    if (!memo) {
        us.assign(kNumEo64 + 2 * kNumEo32, 1e30);
        for (int i = 0; i < kNumEo64; ++i)
            if (eo_fits(c, kEoTiles64[i]))
                us[i] = time_it([&](int k) {
                    (void)launch_eo_fused(c, i, c.nSys, c.phi[c.cur], c.r, c.pv[k & 1], c.pv[(k & 1) ^ 1], c.Ap);
                });
        if (pfOk)
            for (int i = 0; i < kNumEo32; ++i) {
                if (!eo_fits(c, kEoTiles32[i])) continue;
                us[kNumEo64 + 2 * i] =
                    time_it([&](int k) { launch_fin_pf(c, c.nSys, launch_eo_pf(c, i, c.nSys, k, 0), 0); });
                us[kNumEo64 + 2 * i + 1] = time_it([&](int k) { (void)launch_eo_pf(c, i, c.nSys, k, 2); });
            }
        tune_store(1, c, us);
    }
    c.tuneReport += memo ? " | even-odd fp64 tiles (us/launch, measured by an earlier context):"
                         : " | even-odd fp64 tiles (us/launch):";
    double best = 1e30;
    for (int i = 0; i < kNumEo64; ++i) {
        if (!eo_fits(c, kEoTiles64[i])) continue;
        std::snprintf(line, sizeof line, " %dx%d %.1f;", kEoTiles64[i].TT, kEoTiles64[i].TX, us[i]);
        c.tuneReport += line;
        if (us[i] < best) { best = us[i]; c.tileEo64 = i; }
    }
    // FP32 pipelined iteration: each tile with the separate finalize and with the
    // finalize in the next launch's prologue; YCN_EO_FIN=separate|fused restricts the
    // choice to one form
    const char *fin = std::getenv("YCN_EO_FIN");
    const bool allowSep = !fin || std::strcmp(fin, "fused") != 0;
    const bool allowFused = !fin || std::strcmp(fin, "separate") != 0;
    double us32[kNumEo32][2] = {};
    for (int i = 0; i < kNumEo32; ++i) {
        us32[i][0] = us[kNumEo64 + 2 * i];
        us32[i][1] = us[kNumEo64 + 2 * i + 1];
    }
    if (pfOk) {
        c.tuneReport += " | even-odd fp32 pipelined tiles (us/iteration, separate | prologue finalize):";
        best = 1e30;
        for (int i = 0; i < kNumEo32; ++i) {
            if (!eo_fits(c, kEoTiles32[i])) continue;
            std::snprintf(line, sizeof line, " %dx%d%s %.1f|%.1f;", kEoTiles32[i].TT, kEoTiles32[i].TX,
                          kEoTiles32[i].persist ? "p" : "", us32[i][0], us32[i][1]);
            c.tuneReport += line;
            for (int m = 0; m < 2; ++m)
                if ((m ? allowFused : allowSep) && us32[i][m] < best) {
                    best = us32[i][m];
                    c.tileEo32 = i;
                    c.eoFinFused = m == 1;
                }
        }
    }
    // YCN_TILE_EO / YCN_TILE_EO32 force a tile of the tables (tests cover every shape)
    if (const char *f = std::getenv("YCN_TILE_EO")) {
        const int i = std::atoi(f);
        if (i >= 0 && i < kNumEo64 && eo_fits(c, kEoTiles64[i])) c.tileEo64 = i;
        c.tuneReport += " (even-odd fp64 tile forced by YCN_TILE_EO)";
    }
    if (const char *f = std::getenv("YCN_TILE_EO32")) {
        const int i = std::atoi(f);
        if (c.rs && c.phis && c.pfA1 && c.pfP2 && c.pfSlot[0] && i >= 0 && i < kNumEo32 && eo_fits(c, kEoTiles32[i])) {
            c.tileEo32 = i;
            c.eoFinFused = allowFused && (!allowSep || us32[i][1] < us32[i][0]);
        }
        c.tuneReport += " (even-odd fp32 tile forced by YCN_TILE_EO32)";
    }
    if (c.tileEo32 >= 0) {
        const EoTileCfg t = kEoTiles32[c.tileEo32];
        std::snprintf(line, sizeof line, " | fp32 pipelined iteration: %dx%d%s, %s finalize", t.TT, t.TX,
                      t.persist ? "p" : "", c.eoFinFused ? "prologue" : "separate");
        c.tuneReport += line;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    YCN_HIP(hipGetLastError());
    YCN_HIP(hipMemsetAsync(c.cg_d, 0, sizeof(CgState) * c.nSys, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

// scratch of a solve: b_e and U_o in c.t (first and second half), the residual in c.Ap
struct EoBufs {
    double2 *b, *U, *res;
};
static EoBufs eo_bufs(Ctx &c, int nsys) {
    const size_t h = (size_t)nsys * 4 * (c.V / 2);
    return {c.t, c.t + h, c.Ap};
}

static void launch_eo_iteration(Ctx &c, int nsys, const double *phi, int parity) {
    const int nb = launch_eo_fused(c, c.tileEo64, nsys, phi, c.r, c.pv[parity], c.pv[parity ^ 1], c.Ap);
    launch_fin_alpha(c, nsys, nb);
    const int nbu = launch_cg_update(c, nsys, c.V / 2, c.pv[parity ^ 1]);
    launch_fin_beta(c, nsys, nbu, false, true);
}

// launch i of a chunk of the pipelined FP32 iteration: one launch that finalizes its
// predecessor in the prologue, or the fused kernel and k_fin_pf (whichever the autotuner
// timed faster)
static void launch_eo_pf_iteration(Ctx &c, int nsys, int i) {
    if (c.eoFinFused) {
        c.eoPfBlocks = launch_eo_pf(c, c.tileEo32, nsys, i, 1);
        return;
    }
    launch_fin_pf(c, nsys, launch_eo_pf(c, c.tileEo32, nsys, i, 0), 1);
}
// after a chunk of n prologue-finalized launches (n even: launch n - 1 wrote slot 1)
static void launch_fin_pf_chunk(Ctx &c, int nsys, int n) {
    hipLaunchKernelGGL(k_fin_pf_chunk<kNT>, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)c.pfSlot[(n - 1) & 1],
                       (const double2 *)c.pfSlot[n & 1], c.eoPfBlocks, n, c.cg_d, c.active);
}

// Relative mode: the recursive (normal-equation) residual at which the FP64 true
// residual reaches tol varies little between consecutive solves, so the next solve
// starts from 0.9 times the largest ratio need / stop0 the systems of this one reported
// (the first solve starts from tol |r_0|).  A threshold that proves too loose is
// tightened by the true-residual check as before; one too tight costs a few iterations.
static void eo_learn_stop_scale(Ctx &c, int nsys) {
    if (c.p.cgTolMode != YCN_TOL_REL_TRUE) return;
    double r = 0.0;
    for (int k = 0; k < nsys; ++k)
        if (c.cg_h[k].need > 0.0 && c.cg_h[k].stop0 > 0.0) r = std::max(r, c.cg_h[k].need / c.cg_h[k].stop0);
    if (r > 0.0) c.eoStopScale = std::min(1.0, std::max(1e-3, 0.9 * r));
}

// the pipelined FP32 solve with FP64 reliable updates runs where its buffers exist
static bool eo_mixed(const Ctx &c) {
    return c.p.cgPrecision == YCN_CG_MIXED && c.tileEo32 >= 0 && c.rs && c.phis && c.pfA1 && c.pfP2 &&
           c.pfSlot[0] && c.pfSlot[1];
}

// Setup of a solve: eta (from the Philox streams when gen) and u_o = A_o^-1 eta_o,
// b_e = eta_e - K_eo u_o, r = S^T b (FP32 copy when mixed), x and the directions cleared,
// the FP32 copy of the field (mixed) and the per-system state.  ctl (a graph-captured
// step, else null): every launch returns at once while ctl->halt is set, the noise
// counter is ctl->step and the starting threshold ctl->stopScale.
static int eo_setup(Ctx &c, const double *phi, int nsys, bool gen, uint32_t streamBase, uint64_t step, bool mixed,
                    const StepCtl *ctl) {
    EoArgs A = eo_args(c, 0, 0);
    if (ctl) {
        A.halt = &ctl->halt;
        A.stepPtr = &ctl->step;
    }
    const dim3 gF((c.V + kNT - 1) / kNT, nsys), gH = half_grid(c, nsys);
    const int nbF = gF.x, nbH = gH.x;
    double2 *partE = c.partials, *partH = c.partials + (size_t)nsys * nbF;
    const EoBufs B = eo_bufs(c, nsys);
    if (!ctl) {   // (a captured step sets the count in k_step_begin)
        const int rc = set_int(c, c.active, nsys);
        if (rc) return rc;
    }
    const int maxIter = c.p.cgMaxIter, maxRestarts = mixed ? 1 << 20 : 4;
    if (gen)
        hipLaunchKernelGGL(k_eo_eta<true>, gF, dim3(kNT), 0, c.stream, A, (const uint64_t *)c.seed_d, streamBase, step,
                           phi, c.eta, B.U, partE);
    else
        hipLaunchKernelGGL(k_eo_eta<false>, gF, dim3(kNT), 0, c.stream, A, (const uint64_t *)c.seed_d, streamBase,
                           step, phi, c.eta, B.U, partE);
    hipLaunchKernelGGL(k_eo_rhs, gH, dim3(kNT), 0, c.stream, A, (const double2 *)c.eta, (const double2 *)B.U, B.b);
    // r = S^T b (and its FP32 copy); the same launch clears x and the directions (FP64
    // solver: p[0], p[1]; mixed: x_s, p[0], Ap)
    hipLaunchKernelGGL(k_eo_hop_oe<false>, gH, dim3(kNT), 0, c.stream, A, phi, (const double2 *)B.b, B.U);
    EoZero zero{};
    zero.d[0] = c.x;
    if (mixed) {
        zero.f[0] = c.xs;
        zero.f[1] = c.ps[0];
        zero.f[2] = c.Aps;
    } else {
        zero.d[1] = c.pv[0];
        zero.d[2] = c.pv[1];
    }
    hipLaunchKernelGGL((k_eo_hop_eo<false, 2>), gH, dim3(kNT), 0, c.stream, A, phi, (const double2 *)B.b,
                       (const double2 *)B.U, (const double2 *)nullptr, c.r, mixed ? c.rs : (float2 *)nullptr, partH,
                       zero);
    if (mixed) {
        const int nrepUsed = (nsys + c.nNoise - 1) / c.nNoise;
        // a captured step also keeps the FP64 field its solve reads in c.phiSolve: the
        // reliable updates of a halted solve that later steps' chunks run read it there
        const int rc = to_f32_n(c, phi, c.phis, (int64_t)std::min(nrepUsed, c.nRep) * c.V,
                                ctl ? &ctl->halt : nullptr, ctl ? c.phiSolve : nullptr);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_fin_eo_setup, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)partE, nbF,
                       (const double2 *)partH, nbH, c.cg_d, c.active, c.p.cgTol, c.p.cgTolMode, maxIter, maxRestarts,
                       c.eoStopScale, ctl);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

// One chunk of n pipelined FP32 iterations and the reliable update that closes it:
// x += x_s, p[0] = p[1]; res = b - S x; r = S^T res (FP64, FP32 copy into r[0]); then
// converged on the FP64 true residual, or resume.  Systems already final exit every
// kernel at once.
static int eo_mixed_chunk(Ctx &c, const double *phi, int nsys, int n, bool graphed) {
    const int Vh = c.V / 2;
    const dim3 gH = half_grid(c, nsys);
    const EoBufs B = eo_bufs(c, nsys);
    // reliable-update partials: at most V / 64 blocks per system each (smallest FP64 tile)
    double2 *partRu = c.partials, *partRu2 = c.partials + (size_t)nsys * ((c.V + 63) / 64);
    if (n & 1) {   // the pipelined launches defer x_s updates in pairs
        c.err = "internal: a chunk of the pipelined iteration has an odd number of launches";
        return YCN_ERR_INVALID;
    }
    if (graphed) {   // already inside a capture: the launches themselves
        for (int i = 0; i < n; ++i) launch_eo_pf_iteration(c, nsys, i);
    } else {
        const int rc = run_chunk(c, nullptr, nsys, YCN_DAGGER, n, 4, c.tileEo32,
                                 [&](int i) { launch_eo_pf_iteration(c, nsys, i); });
        if (rc) return rc;
    }
    if (c.eoFinFused) launch_fin_pf_chunk(c, nsys, n);
    hipLaunchKernelGGL(k_pf_reliable, gH, dim3(kNT), 0, c.stream, Vh, (const CgState *)c.cg_d, c.x, c.xs,
                       (const float2 *)c.ps[1], (const float2 *)c.pfP2, c.ps[0]);
    const int nbR = launch_eo_ru(c, nsys, phi, c.x, B.b, c.r, c.rs, partRu, partRu2);
    hipLaunchKernelGGL(k_fin_eo_pf_true, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)partRu,
                       (const double2 *)partRu2, nbR, c.cg_d, c.active);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

// Length of the next chunk: the reliable-update interval until the predicted iteration
// count is near, then just past it (a multiple of 8), so that few launches run after the
// systems have paused.
static int eo_chunk_len(const Ctx &c, int pred, int itDone) {
    const int every = std::max(1, c.p.reliableEvery);
    const int nIt = every + (every & 1);
    const int rem = pred - itDone + 4;
    if (pred > 0 && rem < nIt) return std::max(8, (rem + 7) / 8 * 8);
    return nIt;
}

// the eager loop polls once it is within this many iterations of the prediction
static int eo_poll_margin(int pred) { return std::max(8, pred / 8); }

// The pipelined FP32 solve from the state the setup (or an unfinished captured solve)
// left, until every system is final, polling the systems only where they may have
// converged; itDone: iterations already done (the chunk lengths count from there).
// The summary goes to *st, the iteration count and the starting threshold of the next
// solve are learned from it.
static int eo_mixed_loop(Ctx &c, const double *phi, int nsys, int itDone, SolveStats *st) {
    const int every = std::max(1, c.p.reliableEvery);
    const int nIt = every + (every & 1);
    const int hardCap = c.p.cgMaxIter + nIt + 8;
    int launched = 0, active = nsys, rc;
    // profiling the iteration with the prologue finalize: a chunk is nothing but
    // back-to-back iteration launches, so one event pair brackets it and its working
    // launches are the iterations the systems did in it (the most of any system)
    const bool chunkProf = c.profTarget == PK_NORMAL_OP && c.eoFinFused;
    std::vector<int> itPrev(chunkProf ? nsys : 0, 0);
    bool polled = true;
    while (active > 0 && launched < hardCap) {
        const int n = eo_chunk_len(c, c.eoPredIters, itDone);
        const int slot = chunkProf ? prof_chunk_begin(c, PK_NORMAL_OP) : -1;
        rc = eo_mixed_chunk(c, phi, nsys, n, false);
        prof_chunk_end(c, slot);
        if (rc) return rc;
        launched += n;
        // poll only where the systems may have converged: well short of the previous
        // solve's iteration count they cannot have, so the next chunk is enqueued at
        // once (its kernels exit at once for systems that are final after all)
        if (c.eoPredIters > 0 && itDone + n + eo_poll_margin(c.eoPredIters) < c.eoPredIters && !chunkProf) {
            itDone += n;
            polled = false;
            continue;
        }
        rc = poll_iters(c, nsys, active, itDone);
        if (rc) return rc;
        polled = true;
        if (chunkProf) {
            int work = 0;
            for (int k = 0; k < nsys; ++k) {
                work = std::max(work, c.cg_h[k].iter - itPrev[k]);
                itPrev[k] = c.cg_h[k].iter;
            }
            if (slot >= 0) c.evWork[slot] = work;
        }
    }
    if (!polled) {   // the launch cap ended the loop after a skipped poll
        rc = poll_iters(c, nsys, active, itDone);
        if (rc) return rc;
    }
    SolveStats out;   // the last poll left the final states in the mirror
    rc = stats_from_mirror(c, nsys, &out);
    if (rc) return rc;
    c.eoPredIters = out.iters_max;
    int kMax = 0;
    for (int k = 1; k < nsys; ++k)
        if (c.cg_h[k].iter > c.cg_h[kMax].iter) kMax = k;
    c.eoRounds.assign(c.cg_h[kMax].ruIter, c.cg_h[kMax].ruIter + std::min(c.cg_h[kMax].nRu, kMaxRounds));
    eo_learn_stop_scale(c, nsys);
    if (st) *st = out;
    return YCN_OK;
}

// Solve S chi_e = b_e for every system (eta from the Philox streams when gen); chi_e is
// left in c.x, eta in c.eta.  FP64 CG, or FP32 inner iterations with FP64 reliable
// updates when cgPrecision is mixed.
int eo_solve(Ctx &c, const double *phi, int nsys, bool gen, uint32_t streamBase, uint64_t step, SolveStats *st) {
    if (!eo_available(c)) {
        c.err = "even-odd preconditioning needs even Nt and Nx and a fused tile that divides the lattice";
        return YCN_ERR_UNSUPPORTED;
    }
    const bool mixed = eo_mixed(c);
    int rc = eo_setup(c, phi, nsys, gen, streamBase, step, mixed, nullptr);
    if (rc) return rc;
    if (mixed) return eo_mixed_loop(c, phi, nsys, 0, st);

    const EoArgs A = eo_args(c, 0, 0);
    const dim3 gH = half_grid(c, nsys);
    const int nbH = gH.x;
    double2 *partH = c.partials + (size_t)nsys * ((c.V + kNT - 1) / kNT);
    const EoBufs B = eo_bufs(c, nsys);
    const int maxIter = c.p.cgMaxIter, maxRestarts = 4;
    int launched = 0, active = nsys;
    int chunk = std::max(kGraphIters, std::min(c.lastIters, maxIter));
    const int hardCap = maxIter * (maxRestarts + 1) + 2 * kGraphIters;
    for (int guard = 0; guard < maxRestarts + 2; ++guard) {
        while (active > 0 && launched < hardCap) {
            const int n = std::min(chunk, hardCap - launched);
            for (int k = 0; k < n; k += kGraphIters) {
                rc = run_chunk(c, phi, nsys, YCN_DAGGER, kGraphIters, 3, c.tileEo64,
                               [&](int i) { launch_eo_iteration(c, nsys, phi, i & 1); });
                if (rc) return rc;
            }
            launched += ((n + kGraphIters - 1) / kGraphIters) * kGraphIters;
            rc = poll_active(c, active);
            if (rc) return rc;
            chunk = std::max(kGraphIters, std::min(64, chunk / 2));
        }
        // true residual of the unpreconditioned system, |b_e - S x_e| / |eta|
        hipLaunchKernelGGL(k_eo_hop_oe<true>, gH, dim3(kNT), 0, c.stream, A, phi, (const double2 *)c.x, B.U);
        hipLaunchKernelGGL((k_eo_hop_eo<true, 1>), gH, dim3(kNT), 0, c.stream, A, phi, (const double2 *)c.x,
                           (const double2 *)B.U, (const double2 *)B.b, B.res, (float2 *)nullptr, partH, EoZero{});
        hipLaunchKernelGGL(k_fin_eo_true, dim3(nsys), dim3(kNT), 0, c.stream, (const double2 *)partH, nbH, c.cg_d,
                           c.active);
        YCN_HIP(hipGetLastError());
        rc = poll_active(c, active);
        if (rc) return rc;
        if (active == 0 || launched >= hardCap) break;
        chunk = kGraphIters;
    }
    rc = read_stats(c, nsys, st);
    if (rc) return rc;
    eo_learn_stop_scale(c, nsys);
    return YCN_OK;
}

// D^-1 src for the measurement's point source (src in c.eta, system 0, c.nNoise = 1):
// gamma_5 D gamma_5 = D^dag, so D^-1 src = gamma_5 (D^dag)^-1 (gamma_5 src), and the
// fermion-force solve of D^dag chi = eta gives it.  The solve's iteration count, starting
// threshold and chunk length are not learned from this right-hand side: they size the
// fermion-force solves.
int eo_propagator(Ctx &c, const double *phi, double2 *out, SolveStats *st) {
    const unsigned nbV = (unsigned)((c.V + kNT - 1) / kNT);
    hipLaunchKernelGGL(k_g5, dim3(nbV), dim3(kNT), 0, c.stream, c.V, c.eta);
    const int pred = c.eoPredIters, last = c.lastIters;
    const double scale = c.eoStopScale;
    int rc = eo_solve(c, phi, 1, false, 0, 0, st);
    c.eoPredIters = pred;
    c.lastIters = last;
    c.eoStopScale = scale;
    if (rc) return rc;
    const EoArgs A = eo_args(c, 0, 0);
    hipLaunchKernelGGL(k_eo_propagator, dim3(nbV), dim3(kNT), 0, c.stream, A, phi, (const double2 *)c.x,
                       (const double2 *)c.eta, out);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int eo_drift(Ctx &c, const double *phi, double *drift, const StepCtl *ctl) {
    EoArgs A = eo_args(c, 0, 0);
    if (ctl) A.halt = &ctl->halt;
    const double fac = -c.p.g * (c.p.driftMode == 1 ? c.p.cutFraction * c.p.cutFraction : 1.0);
    hipLaunchKernelGGL(k_eo_drift, dim3((c.V + kNT - 1) / kNT, c.nRep), dim3(kNT), 0, c.stream, A, phi,
                       (const double2 *)c.x, (const double2 *)c.eta, drift, fac);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

// ---------------------------------------------------------------------------
// graph-captured Langevin steps
// ---------------------------------------------------------------------------
// A step replays one captured graph -- the solve's setup, a fixed chunk schedule of the
// pipelined FP32 iteration with its reliable updates, k_step_check and the drift --
// between k_step_begin and the coloured noise before it and the update after it, with no
// host wait.  The schedule repeats the previous solve's rounds (eo_schedule), then `tail`
// chunks of 8.  A solve that needs more than the schedule holds halts the run
// (k_step_check): the later steps' kernels return at once except their chunk schedules,
// which keep iterating the halted solve (its FP32 and FP64 fields stay in c.phis and
// c.phiSolve), and the host finishes the halted step when it reads the control block at
// the end of the segment, then resumes.  Every chunk past the solve's end costs its
// launches (their kernels exit at once) and a halted step the rest of its segment, so the
// tail starts at one chunk, grows by one after each halted step and shrinks by one after
// four segments without one.  The field ping-pongs between the context's two buffers with
// a fixed step size and is updated in place with an adaptive one, so each schedule has at
// most two executables (one per field buffer).
constexpr int kStepSegment = 16;   // captured steps between two reads of the control block
constexpr int kMaxTail = 8;
constexpr size_t kMaxStepGraphs = 48;

bool eo_graph_eligible(const Ctx &c) {
    return c.stepGraphsOn && c.useGraphs && c.profTarget == PK_NONE && c.p.evenOdd && eo_available(c) &&
           eo_mixed(c) && c.p.precond == YCN_PRECOND_NONE && c.eoPredIters > 0;
}

// chunks of the schedule's main part for a predicted count when no earlier solve's
// reliable updates are known: the chunks the eager loop launches up to its first poll
static std::vector<int> eo_main_chunks(const Ctx &c, int pred) {
    const int every = std::max(1, c.p.reliableEvery);
    const int cap = c.p.cgMaxIter + every + (every & 1) + 8;
    std::vector<int> v;
    int itDone = 0;
    do {
        v.push_back(eo_chunk_len(c, pred, itDone));
        itDone += v.back();
    } while (itDone + eo_poll_margin(pred) < pred && itDone < cap);
    return v;
}

// The chunk schedule of a captured step: one chunk per reliable update of the previous
// solve (its longest system), each as long as that solve's round (the iterations between
// two reliable updates: up to a pause for the residual drop, to the stop, or to the end of
// an eager chunk) plus 2, rounded up to a multiple of 8 (so that consecutive solves mostly
// share one schedule and its captured graphs), then `tail` chunks of 8.  A
// round of this solve that ends a little later than the previous one's ends at the chunk
// boundary instead and the rest moves to the next round; a solve that needs more than
// the schedule holds halts the run.
constexpr int kTailRound = 8;
static std::vector<int> eo_schedule(const Ctx &c, int tail) {
    std::vector<int> v;
    if (c.eoRounds.empty()) {
        v = eo_main_chunks(c, c.eoPredIters);
    } else {
        int prev = 0, total = 0;
        for (int r : c.eoRounds) {
            v.push_back(std::max(8, (r - prev + 2 + 7) / 8 * 8));
            total += v.back();
            prev = r;
        }
        // a solve with more reliable updates than CgState records ends after its last
        // recorded round: chunks of the reliable-update interval up to its iteration count
        if (prev < c.eoPredIters) {
            const int every = std::max(1, c.p.reliableEvery), nIt = every + (every & 1);
            const int want = std::min(c.eoPredIters, c.p.cgMaxIter) + 2;
            while (total < want) {
                v.push_back(nIt);
                total += nIt;
            }
        }
    }
    for (int k = 0; k < tail; ++k) v.push_back(kTailRound);
    return v;
}

// the captured part of a step: the solve's setup, its chunk schedule, k_step_check and
// the drift (the step's noise and update are launched around it by eo_enqueue_step)
static int eo_capture_step(Ctx &c, int parity, const std::vector<int> &chunks, hipGraphExec_t *out) {
    const double *phiIn = c.phi[parity];
    const int nsys = c.nSys;
    hipGraph_t g = nullptr;
    YCN_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    int rc = eo_setup(c, phiIn, nsys, true, 1u, 0, true, c.ctl_d);
    for (int n : chunks)
        if (!rc) rc = eo_mixed_chunk(c, c.phiSolve, nsys, n, true);
    if (!rc) {
        hipLaunchKernelGGL(k_step_check, dim3(1), dim3(kWave), 0, c.stream, c.ctl_d, (const CgState *)c.cg_d, nsys,
                           c.p.cgTolMode == YCN_TOL_REL_TRUE ? 1 : 0);
        rc = eo_drift(c, phiIn, c.drift, c.ctl_d);   // fDrift.getForce(drift)
    }
    const hipError_t le = hipGetLastError();
    const hipError_t ee = hipStreamEndCapture(c.stream, &g);
    if (rc || le != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (!rc) {
            c.err = std::string("step capture: ") + hipGetErrorString(le != hipSuccess ? le : ee);
            rc = YCN_ERR_HIP;
        }
        return rc;
    }
    const hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        c.err = std::string("step graph instantiation: ") + hipGetErrorString(ie);
        return YCN_ERR_HIP;
    }
    return YCN_OK;
}

// the executable of a schedule and field parity, captured at first use (the cache keeps
// the most recently used last); called only when no step graph is in flight: a segment
// fetches both of its executables before it launches a step
static int eo_step_graph(Ctx &c, int parity, const std::vector<int> &chunks, hipGraphExec_t *out) {
    for (size_t i = 0; i < c.stepGraphs.size(); ++i) {
        const Ctx::StepGraph e = c.stepGraphs[i];
        if (e.chunks == chunks && e.parity == parity) {
            c.stepGraphs.erase(c.stepGraphs.begin() + (long)i);
            c.stepGraphs.push_back(e);
            *out = e.exec;
            return YCN_OK;
        }
    }
    if (c.stepGraphs.size() >= kMaxStepGraphs) {
        (void)hipGraphExecDestroy(c.stepGraphs.front().exec);
        c.stepGraphs.erase(c.stepGraphs.begin());
    }
    const int rc = eo_capture_step(c, parity, chunks, out);
    if (rc) return rc;
    c.stepGraphs.push_back({chunks, parity, *out});
    return YCN_OK;
}

// While a segment runs on the GPU the host would only wait: it captures the schedules
// the next segment most likely needs, the current one with one chunk 8 launches shorter
// or longer (the rounds of consecutive solves move by a few iterations), for both field
// buffers, at most kPrecapture new ones per segment.  They go in at the cache's old end,
// so that they are evicted before anything the running segment uses.
constexpr int kPrecapture = 8;
static int eo_precapture(Ctx &c, const std::vector<int> &chunks, int nTail, bool bothParities) {
    int made = 0;
    const size_t nMain = chunks.size() - (size_t)nTail;
    for (size_t i = 0; i < nMain && made < kPrecapture; ++i)
        for (int d : {-8, 8}) {
            if (chunks[i] + d < 8) continue;
            std::vector<int> v = chunks;
            v[i] += d;
            for (int par = 0; par < (bothParities ? 2 : 1) && made < kPrecapture; ++par) {
                const int parity = bothParities ? par : c.cur;
                bool have = false;
                for (const Ctx::StepGraph &e : c.stepGraphs)
                    if (e.chunks == v && e.parity == parity) have = true;
                if (have) continue;
                if (c.stepGraphs.size() + 1 >= kMaxStepGraphs) return YCN_OK;   // never evict here
                hipGraphExec_t ex;
                const int rc = eo_capture_step(c, parity, v, &ex);
                if (rc) return rc;
                c.stepGraphs.insert(c.stepGraphs.begin(), {v, parity, ex});
                ++made;
            }
        }
    return YCN_OK;
}

// One captured step on the context stream, no host wait: k_step_begin, the coloured
// noise (on the noise stream beside the solve when the context has it, as a single step
// does: a graph with a second branch measured slower), the captured solve and drift, then
// the update once the noise is there.  The noise reads the copy of the halt flag and
// counter k_step_begin made, so the solve's check kernel cannot change them under it.
static int eo_enqueue_step(Ctx &c, int parity, hipGraphExec_t ex) {
    const bool fixed = c.p.adaptive == 0;
    double *phiIn = c.phi[parity], *phiOut = fixed ? c.phi[parity ^ 1] : c.phi[parity];
    StepCtl *ctlNoise = c.ctl_d + 1;
    const bool side = c.noiseSide && c.noiseStream;
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(kWave), 0, c.stream, (const StepCtl *)c.ctl_d, ctlNoise, c.active,
                       c.nSys);
    int rc;
    if (side) {
        YCN_HIP(hipEventRecord(c.evMain, c.stream));
        YCN_HIP(hipStreamWaitEvent(c.noiseStream, c.evMain, 0));
        rc = noise_generate_ctl(c, c.noiseStream, ctlNoise, c.noise);   // cn()
        if (rc) return rc;
        YCN_HIP(hipEventRecord(c.evNoise, c.noiseStream));
    } else {
        rc = noise_generate_ctl(c, c.stream, ctlNoise, c.noise);   // cn()
        if (rc) return rc;
    }
    YCN_HIP(hipGraphLaunch(ex, c.stream));   // fDrift.getForce(drift)
    if (side) YCN_HIP(hipStreamWaitEvent(c.stream, c.evNoise, 0));
    return bosonic_update_gated(c, fixed ? 1 : 0, phiIn, phiOut, &c.ctl_d->halt);   // kli.Run(kAll)
}

static void add_step_stats(SolveStats *tot, const SolveStats &st) {
    tot->iters_max += st.iters_max;
    tot->restarts += st.restarts;
    tot->converged &= st.converged;
    tot->rr_max = std::max(tot->rr_max, st.rr_max);
    tot->true_rel_max = std::max(tot->true_rel_max, st.true_rel_max);
    tot->iters_mean += st.iters_mean;
}

int eo_graph_run(Ctx &c, uint64_t step0, int nsteps, SolveStats *tot) {
    if (!c.ctl_d) {
        YCN_HIP(hipMalloc(&c.ctl_d, 2 * sizeof(StepCtl)));   // the control block and the noise's copy
        YCN_HIP(hipHostMalloc((void **)&c.ctl_h, sizeof(StepCtl), hipHostMallocDefault));
        YCN_HIP(hipMalloc(&c.phiSolve, sizeof(double) * (size_t)c.nRep * c.V));
    }
    const bool fixed = c.p.adaptive == 0;
    const int nsys = c.nSys;
    StepCtl &h = *c.ctl_h;
    h = StepCtl{};
    h.converged = 1;
    h.step = step0;
    h.stopScale = c.eoStopScale;
    h.lastIters = c.eoPredIters;
    h.nRounds = std::min((int)c.eoRounds.size(), kMaxRounds);
    for (int i = 0; i < h.nRounds; ++i) h.rounds[i] = c.eoRounds[i];
    YCN_HIP(hipMemcpyAsync(c.ctl_d, c.ctl_h, sizeof(StepCtl), hipMemcpyHostToDevice, c.stream));
    SolveStats host;   // figures of the steps the host finished
    int done = 0, rc;
    while (done < nsteps) {
        const int seg = std::min(kStepSegment, nsteps - done);
        const std::vector<int> chunks = eo_schedule(c, c.stepTail);
        const int p0 = c.cur;
        hipGraphExec_t ex[2] = {nullptr, nullptr};
        if ((rc = eo_step_graph(c, p0, chunks, &ex[p0]))) return rc;
        if (fixed && seg > 1 && (rc = eo_step_graph(c, p0 ^ 1, chunks, &ex[p0 ^ 1]))) return rc;
        for (int i = 0; i < seg; ++i) {
            const int parity = fixed ? p0 ^ (i & 1) : p0;
            if ((rc = eo_enqueue_step(c, parity, ex[parity]))) return rc;
        }
        YCN_HIP(hipMemcpyAsync(c.ctl_h, c.ctl_d, sizeof(StepCtl), hipMemcpyDeviceToHost, c.stream));
        if ((rc = eo_precapture(c, chunks, c.stepTail, fixed))) return rc;
        if ((rc = spin_sync(c))) return rc;
        if (!h.halt) {
            c.cur = fixed ? p0 ^ (seg & 1) : p0;
            done += seg;
            if (++c.stepClean >= 4 && c.stepTail > 1) {
                --c.stepTail;
                c.stepClean = 0;
            }
        } else {
            // the halted step's solve is still in the context's state (the later steps'
            // chunk kernels may have iterated it further): finish it, then its drift and
            // update, eagerly; the field it started from is the one that step read
            const uint64_t first = step0 + (uint64_t)done;
            if (h.haltStep < first || h.haltStep >= first + (uint64_t)seg) {
                c.err = "captured steps: the control block names a step outside the segment";
                return YCN_ERR_HIP;
            }
            const int k = (int)(h.haltStep - first);
            c.cur = fixed ? p0 ^ (k & 1) : p0;
            SolveStats st;
            if ((rc = eo_mixed_loop(c, c.phi[c.cur], nsys, c.eoPredIters, &st))) return rc;
            if ((rc = eo_drift(c, c.phi[c.cur], c.drift))) return rc;
            if ((rc = bosonic_update(c, fixed ? 1 : 0))) return rc;
            add_step_stats(&host, st);
            ++c.haltedSteps;
            c.stepTail = std::min(c.stepTail + 1, kMaxTail);
            c.stepClean = 0;
            h.halt = 0;
            h.step = h.haltStep + 1;
            h.stopScale = c.eoStopScale;
            h.lastIters = st.iters_max;
            h.nRounds = (int)c.eoRounds.size();
            for (int i = 0; i < h.nRounds; ++i) h.rounds[i] = c.eoRounds[i];
            YCN_HIP(hipMemcpyAsync(c.ctl_d, c.ctl_h, sizeof(StepCtl), hipMemcpyHostToDevice, c.stream));
            done += k + 1;
        }
        // the next segment's schedule from the last finished step
        if (h.lastIters > 0) {
            c.eoPredIters = h.lastIters;
            c.eoRounds.assign(h.rounds, h.rounds + std::min(h.nRounds, kMaxRounds));
        }
        c.eoStopScale = h.stopScale;
    }
    c.lastIters = std::max(kGraphIters, c.eoPredIters);
    // the captured steps' figures (the last read of the block holds them all)
    SolveStats dev;
    dev.iters_max = (int)h.itersMaxSum;
    dev.iters_mean = h.itersMeanSum;
    dev.restarts = h.restarts;
    dev.converged = h.converged;
    dev.rr_max = h.rrMax;
    dev.true_rel_max = h.trueRelMax;
    add_step_stats(tot, dev);
    add_step_stats(tot, host);
    return YCN_OK;
}

}  // namespace ycn
