// ycn_device.hpp -- device helpers shared by the gfx950 kernels of libycn.
//
// Layouts (device):
//   scalar field   double[rep][t][x]                       (x fastest)
//   spinor (SoA)   double2[sys][c][t][x], c = f1s1,f1s2,f2s1,f2s2
//                  sys = rep * nNoise + j  (j = noise vector)
// The AoS spinor of the reference (code/src/Spinor.cuh:7-16) is converted at the
// ABI boundary only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace ycn {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// complex helpers on double2 / float2
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscl(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cdotc(double2 a, double2 b) {
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11) and Box-Muller (1958)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k) {
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        if (i) { k.x += 0x9E3779B9u; k.y += 0xBB67AE85u; }
        const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    }
    return c;
}

// Two N(0,1) values from Philox counter (q, step) of key (seed, stream):
// u1 in (0,1], u2 in [0,1) with 53 bits each; (rho cos 2 pi u2, rho sin 2 pi u2).
__device__ __forceinline__ double2 normal_pair(uint64_t seed, uint32_t stream, uint64_t step, uint64_t q) {
    const uint4 o = philox4x32_10(make_uint4((uint32_t)q, (uint32_t)(q >> 32), (uint32_t)step, (uint32_t)(step >> 32)),
                                  make_uint2((uint32_t)seed, stream));
    const uint64_t a = ((((uint64_t)o.x) << 32) | o.y) >> 11;
    const uint64_t b = ((((uint64_t)o.z) << 32) | o.w) >> 11;
    const double u1 = (double)(a + 1) * 0x1p-53;
    const double u2 = (double)b * 0x1p-53;
    const double rho = sqrt(-2.0 * log(u1));
    const double th = (2.0 * M_PI) * u2;
    return make_double2(rho * cos(th), rho * sin(th));
}

// ---------------------------------------------------------------------------
// deterministic reductions: wave64 butterfly, then LDS across waves
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Sum over the block; valid in thread 0.  lds must hold NT/64 double2.
template <int NT>
__device__ __forceinline__ double2 block_sum2(double2 v, double2 *lds) {
    v.x = wave_sum(v.x);
    v.y = wave_sum(v.y);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (NT > kWave) {
        if (lane == 0) lds[wid] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double2 s = lds[0];
#pragma unroll
            for (int w = 1; w < NT / kWave; ++w) s = cadd(s, lds[w]);
            v = s;
        }
    }
    return v;
}
template <int NT>
__device__ __forceinline__ double block_max(double v, double *lds) {
    v = wave_max(v);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    if (NT > kWave) {
        if (lane == 0) lds[wid] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double s = lds[0];
#pragma unroll
            for (int w = 1; w < NT / kWave; ++w) s = fmax(s, lds[w]);
            v = s;
        }
    }
    return v;
}

// ---------------------------------------------------------------------------
// Wilson-Dirac hopping on one site (code/src/Dirac.cu:49-70,144-190).
//   D   : T_b = s1 psi(t-1), T_f = s0 psi(t+1), X_f = psi(x+1), X_b = psi(x-1)
//   D^+ : same with forward and backward exchanged (D is real, D^+ = D^T)
//   out0 = d in0 - [hp T_b0 + hm T_f0 + hr (X_f0 + X_b0) + h (X_b1 - X_f1)]
//   out1 = d in1 - [hm T_b1 + hp T_f1 + hr (X_f1 + X_b1) + h (X_b0 - X_f0)]
// with hp = (r+1)/2, hm = (r-1)/2, hr = r/2, h = 1/2, d = 2r + M + g phi,
// s0 = -1 on the t = Nt-1 -> 0 hop and s1 = -1 on the t = 0 -> Nt-1 hop
// (antiperiodic time), same for both flavours.
// ---------------------------------------------------------------------------
struct WilsonCoef {
    double hp, hm, hr, h, two_r_plus_m, g;
};

template <typename V2>
__device__ __forceinline__ V2 v2_fma(double a, V2 x, V2 acc) {
    acc.x += a * x.x;
    acc.y += a * x.y;
    return acc;
}

// Neighbour offsets for site (t, x) on an Nt x Nx periodic lattice.
struct Nbr {
    int tf, tb, xf, xb;   // flat indices
    double sf, sb;        // antiperiodic signs for the time hops (s0, s1)
};
__device__ __forceinline__ Nbr neighbours(int t, int x, int Nt, int Nx) {
    Nbr n;
    const int tp = (t == Nt - 1) ? 0 : t + 1, tm = (t == 0) ? Nt - 1 : t - 1;
    const int xp = (x == Nx - 1) ? 0 : x + 1, xm = (x == 0) ? Nx - 1 : x - 1;
    n.tf = tp * Nx + x;
    n.tb = tm * Nx + x;
    n.xf = t * Nx + xp;
    n.xb = t * Nx + xm;
    n.sf = (t == Nt - 1) ? -1.0 : 1.0;
    n.sb = (t == 0) ? -1.0 : 1.0;
    return n;
}

// hop for one flavour given the two spin components at own site and the
// (sign-folded) neighbour values.  DAG selects D^dagger.
template <bool DAG, typename V2>
__device__ __forceinline__ void wilson_flavour(const WilsonCoef &w, double d, V2 in0, V2 in1,
                                               V2 tf0, V2 tf1, double sf, V2 tb0, V2 tb1, double sb,
                                               V2 xf0, V2 xf1, V2 xb0, V2 xb1, V2 &o0, V2 &o1) {
    // Map to (T_b, T_f, X_f, X_b) of the operator.
    const V2 Tb0 = DAG ? tf0 : tb0, Tb1 = DAG ? tf1 : tb1;
    const V2 Tf0 = DAG ? tb0 : tf0, Tf1 = DAG ? tb1 : tf1;
    const double sTb = DAG ? sf : sb, sTf = DAG ? sb : sf;
    const V2 Xf0 = DAG ? xb0 : xf0, Xf1 = DAG ? xb1 : xf1;
    const V2 Xb0 = DAG ? xf0 : xb0, Xb1 = DAG ? xf1 : xb1;
    const double a_hp = w.hp * sTb, a_hm = w.hm * sTf;   // comp 0 time coefficients
    const double b_hm = w.hm * sTb, b_hp = w.hp * sTf;   // comp 1 time coefficients
    V2 s0, s1;
    s0.x = a_hp * Tb0.x + a_hm * Tf0.x + w.hr * (Xf0.x + Xb0.x) + w.h * (Xb1.x - Xf1.x);
    s0.y = a_hp * Tb0.y + a_hm * Tf0.y + w.hr * (Xf0.y + Xb0.y) + w.h * (Xb1.y - Xf1.y);
    s1.x = b_hm * Tb1.x + b_hp * Tf1.x + w.hr * (Xf1.x + Xb1.x) + w.h * (Xb0.x - Xf0.x);
    s1.y = b_hm * Tb1.y + b_hp * Tf1.y + w.hr * (Xf1.y + Xb1.y) + w.h * (Xb0.y - Xf0.y);
    o0.x = d * in0.x - s0.x;
    o0.y = d * in0.y - s0.y;
    o1.x = d * in1.x - s1.x;
    o1.y = d * in1.y - s1.y;
}

// Gaussian noise vector of system sys at site s: component c is normal number 4s + c
// of Philox stream (seed of the replica, streamBase + noise index), imaginary part 0
// (replaces fillNormalRND, FermionicDrift.cu:102-110).
__device__ __forceinline__ void eta_site(const uint64_t *seeds, uint32_t streamBase, uint64_t step, int sys,
                                         int nNoise, int s, double2 b[4]) {
    const int rep = sys / nNoise, j = sys - rep * nNoise;
    const double2 g0 = normal_pair(seeds[rep], streamBase + (uint32_t)j, step, (uint64_t)2 * s);
    const double2 g1 = normal_pair(seeds[rep], streamBase + (uint32_t)j, step, (uint64_t)2 * s + 1);
    b[0] = make_double2(g0.x, 0.0);
    b[1] = make_double2(g0.y, 0.0);
    b[2] = make_double2(g1.x, 0.0);
    b[3] = make_double2(g1.y, 0.0);
}

}  // namespace ycn
