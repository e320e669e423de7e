// cnn.hip -- the learned preconditioner on gfx950: a 3-layer periodic 3x3 CNN
// (8 -> 16 -> 16 -> 8 channels, no bias) with BF16 weights and activations and FP32
// accumulation on MFMA, trained online with Adam on (residual, error) pairs the CG
// solver produces.  No reference counterpart.
//
// The data path is linear in r; each hidden activation is gated by the field:
//   u1 = (W1 * u0) . (1 + tanh(G1 * q)),  u2 = (W2 * u1) . (1 + tanh(G2 * q)),
//   y = W3 * u2,  q = g phi,
// so the preconditioner is a linear map of r (flexible CG stays well behaved) whose
// stencil follows the local Wilson diagonal 2r + M + g phi.  Channel 2c / 2c+1 = Re / Im
// of spinor component c.  The network is a learned correction to an exact FP32 scaled
// identity: z = c0 r + s * y(r / s), with c0 = 1 / (2r + M)^2 and s = RMS of the 8V real
// entries of r, so a zero network is plain CG and the bf16 rounding only touches the
// learned part.
// Weight layout (FP32 master copy): W1[9][8][16], W2[9][16][16], W3[9][16][8],
// G1[9][16], G2[9][16], [tap][in][out], tap = (dt+1)*3 + (dx+1),
// out(site) = sum W * in(site + (dt, dx)).  The gates run in FP32 on the VALU.
//
// Inference (inside flexible CG): the gates depend on the field only, so k_cnn_gates
// computes both layers' gates once per solve into HBM (FP16, 64 B per site);
// k_cnn_fused keeps all three layers of a 16x16 tile (plus its 3-site halo) in LDS as
// BF16, multiplies the hidden layers by the gates it loads, and fuses the two dot
// products the flexible CG recurrence needs.  Compulsory HBM traffic per site: read r_s,
// Ap (32 + 32 B) and the gates (64 B), write z (32 B).
// Training: layer-wise MFMA kernels over channels-last FP32 activations in HBM,
// forward, data gradient (a forward convolution with mirrored, transposed weights),
// weight gradient (an MFMA GEMM whose reduction runs over lattice sites), then Adam.
// The loss is the residual the correction leaves, |A (e - z)|^2 / |r|^2 (= |r - A z|^2 /
// |r|^2 for r = A e): it weights every mode the way the next CG residual does.
#include <algorithm>
#include <cmath>
#include <vector>

#include "cg_fused.hpp"
#include "cnn_mfma.hpp"
#include "ycn_device.hpp"
#include "ycn_internal.hpp"

namespace ycn {

constexpr int kCT = 256;   // threads per CNN block (4 waves)

__device__ __forceinline__ int wrapn(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// ---------------------------------------------------------------------------
// weight fragments
// ---------------------------------------------------------------------------
// Forward fragment of a layer: B[k][n] = W[k][n] (k = tap*cin + ci), zero padded.
// Data-gradient fragment: the transposed convolution equals a forward convolution with
// W'[tap'][co][ci] = W[8 - tap'][ci][co] (mirrored taps), i.e. cin' = cout.
__global__ void k_make_frags(const float *__restrict__ W, bf16x8 *__restrict__ f1, bf16x8 *__restrict__ f2,
                             bf16x8 *__restrict__ f3, bf16x8 *__restrict__ d3, bf16x8 *__restrict__ d2) {
    const int lane = threadIdx.x & 63;
    const int job = blockIdx.x;   // 0..4
    const float *w;
    int cin, cout, ks;
    bool dgrad;
    bf16x8 *dst;
    switch (job) {
    case 0: w = W; cin = 8; cout = 16; dgrad = false; dst = f1; break;
    case 1: w = W + kCnnNW1; cin = 16; cout = 16; dgrad = false; dst = f2; break;
    case 2: w = W + kCnnNW1 + kCnnNW2; cin = 16; cout = 8; dgrad = false; dst = f3; break;
    case 3: w = W + kCnnNW1 + kCnnNW2; cin = 8; cout = 16; dgrad = true; dst = d3; break;   // layer 3: 16 -> 8
    default: w = W + kCnnNW1; cin = 16; cout = 16; dgrad = true; dst = d2; break;          // layer 2
    }
    ks = ksteps(cin);
    const int n = lane % 16, kg = lane / 16;
    for (int s = threadIdx.x / 64; s < ks; s += blockDim.x / 64) {
        bf16x8 v;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = s * 32 + kg * 8 + i;
            float val = 0.f;
            if (k < 9 * cin && n < cout) {
                const int tap = k / cin, ci = k % cin;
                // dgrad: this "layer" maps cin (= original cout) to cout (= original cin)
                val = dgrad ? w[((8 - tap) * cout + n) * cin + ci] : w[(tap * cin + ci) * cout + n];
            }
            v[i] = (__bf16)val;
        }
        dst[s * 64 + lane] = v;
    }
}

// ---------------------------------------------------------------------------
// fused inference: z = s * net(r_s / s) on a 16x16 tile, plus partial <z, r_s>, <z, Ap>
// ---------------------------------------------------------------------------
// One layer on the LDS-resident tile: out[m][n] = sum_k in[site(m) + off(tap)][ci] W[k][n]
// for the WOUT x WOUT output sites m (input array width WIN, one more site of halo on
// each side).  The MFMA runs transposed, D[n][m] = sum_k W[k][n] A[m][k]: the weight
// fragments are its first operand, the activations its second, so a lane ends with four
// consecutive channels 4 kg .. 4 kg + 3 of one site m = 16 g + col and stores them with
// one access.  Each lane's tap offsets are fixed for the layer (kg picks them); padded K
// steps load the centre tap (finite activations against zero weights) instead of
// branching.  pre(m, kg) runs before a group's MFMAs (its global loads overlap them) and
// store(m, kg, acc, pre's result) gets the four channels.
template <int CIN, int WOUT, int WIN, int KS, typename Pre, typename Store>
__device__ __forceinline__ void conv_lds(const __bf16 *__restrict__ in, const bf16x8 *__restrict__ frag, Pre pre,
                                         Store store) {
    constexpr int N = WOUT * WOUT;
    constexpr int G = (N + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kg = lane >> 4, col = lane & 15;
    bf16x8 B[KS];   // this lane's weight fragments, held for all of the wave's site groups
    int toff[KS];   // LDS offset of this lane's tap and channel block from its site's corner
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        B[ks] = frag[ks * 64 + lane];
        const int k0 = ks * 32 + kg * 8;
        const int tap = k0 / CIN, ci0 = k0 % CIN;
        toff[ks] = tap < 9 ? ((tap / 3) * WIN + tap % 3) * CIN + ci0 : (WIN + 1) * CIN;
    }
    for (int g = wave; g < G; g += kCT / 64) {
        const int m = g * 16 + col;
        const int mc = m < N ? m : N - 1;
        const int a = mc / WOUT, b = mc - a * WOUT;
        const __bf16 *base = in + (a * WIN + b) * CIN;
        const auto pl = pre(mc, kg);
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            acc = mfma_bf16(B[ks], *reinterpret_cast<const bf16x8 *>(base + toff[ks]), acc);
        if (m < N) store(m, kg, acc, pl);
    }
}

// The field gates of both hidden layers at every site, once per solve (they depend on
// the field and the gate weights only): gates[rep][site][32], halves 0-15 the layer-1
// gates 1 + tanh(sum_tap G1[tap][n] q(site + off(tap))), halves 16-31 the layer-2 ones,
// q = g phi.  One thread per site, nine q taps in registers, weights read as broadcast
// float4s; FP32 arithmetic, stored as FP16 (the precision the fused kernel multiplies by).
__global__ void __launch_bounds__(kCT)
k_cnn_gates(int Nt, int Nx, const float *__restrict__ phis, float g, const float4 *__restrict__ G,
            _Float16 *__restrict__ gates) {
    const int V = Nt * Nx, rep = blockIdx.y;
    const int s = blockIdx.x * kCT + threadIdx.x;
    if (s >= V) return;
    const float *ph = phis + (size_t)rep * V;
    const int t = s / Nx, x = s - t * Nx;
    float qn[9];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) qn[tap] = g * ph[wrapn(t + tap / 3 - 1, Nt) * Nx + wrapn(x + tap % 3 - 1, Nx)];
    _Float16 *out = gates + ((size_t)rep * V + s) * 32;
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        const float4 *Gl = G + layer * (kCnnNWG / 4);
        float tt[16];
#pragma unroll
        for (int n = 0; n < 16; ++n) tt[n] = 0.f;
#pragma unroll
        for (int tap = 0; tap < 9; ++tap)
#pragma unroll
            for (int n4 = 0; n4 < 4; ++n4) {
                const float4 w = Gl[tap * 4 + n4];
                tt[4 * n4] += w.x * qn[tap];
                tt[4 * n4 + 1] += w.y * qn[tap];
                tt[4 * n4 + 2] += w.z * qn[tap];
                tt[4 * n4 + 3] += w.w * qn[tap];
            }
        typedef _Float16 h8 __attribute__((ext_vector_type(8)));
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            h8 v;
#pragma unroll
            for (int i = 0; i < 8; ++i) v[i] = (_Float16)gate_fn(tt[8 * h + i]);
            *reinterpret_cast<h8 *>(out + layer * 16 + 8 * h) = v;
        }
    }
}

__global__ void __launch_bounds__(kCT)
k_cnn_fused(int Nt, int Nx, int tilesX, int nNoise, float c0, const CgState *__restrict__ st,
            const float2 *__restrict__ rs, const float2 *__restrict__ Ap, const _Float16 *__restrict__ gates,
            float2 *__restrict__ z, const bf16x8 *__restrict__ gf1, const bf16x8 *__restrict__ gf2,
            const bf16x8 *__restrict__ gf3, double2 *__restrict__ part) {
    constexpr int W0 = 22, W1 = 20, W2 = 18, W3 = 16;
    __shared__ __attribute__((aligned(16))) __bf16 U0[W0 * W0 * 8];
    __shared__ __attribute__((aligned(16))) __bf16 U1[W1 * W1 * 16];
    __shared__ __attribute__((aligned(16))) __bf16 U2[W2 * W2 * 16];
    __shared__ double2 red[kCT / kWave];
    const int sys = blockIdx.y;
    if (st[sys].done) return;
    const int V = Nt * Nx;
    const size_t off = (size_t)sys * 4 * V;
    const int tile = blockIdx.x;
    const int ty = tile / tilesX, tx = tile - ty * tilesX;
    const int t0 = ty * 16, x0 = tx * 16;
    const double s = sqrt(st[sys].rr / (8.0 * (double)V));
    const float inv = s > 0.0 ? (float)(1.0 / s) : 0.f, sf = (float)s;
    double2 acc = make_double2(0.0, 0.0);
    if (st[sys].precOff) {
        // plain CG fallback: identity preconditioner z = r_s
        const int lt = threadIdx.x / 16, lx = threadIdx.x % 16;
        const size_t gs = (size_t)(t0 + lt) * Nx + x0 + lx;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const float2 rv = rs[off + c * V + gs], av = Ap[off + c * V + gs];
            z[off + c * V + gs] = rv;
            acc.x += (double)(rv.x * rv.x + rv.y * rv.y);
            acc.y += (double)(rv.x * av.x + rv.y * av.y);
        }
    } else {
        const _Float16 *gr = gates + (size_t)(sys / nNoise) * V * 32;
        // all of a thread's loads are issued before the first conversion
        constexpr int L0 = (W0 * W0 + kCT - 1) / kCT;
        float2 rv0[L0][4];
#pragma unroll
        for (int u = 0; u < L0; ++u) {
            const int w = threadIdx.x + u * kCT;
            if (w < W0 * W0) {
                const int a = w / W0, b = w - a * W0;
                const size_t gs = (size_t)wrapn(t0 - 3 + a, Nt) * Nx + wrapn(x0 - 3 + b, Nx);
#pragma unroll
                for (int c = 0; c < 4; ++c) rv0[u][c] = rs[off + c * V + gs];
            }
        }
#pragma unroll
        for (int u = 0; u < L0; ++u) {
            const int w = threadIdx.x + u * kCT;
            if (w < W0 * W0) {
                bf16x8 q;
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    q[2 * c] = (__bf16)(rv0[u][c].x * inv);
                    q[2 * c + 1] = (__bf16)(rv0[u][c].y * inv);
                }
                *reinterpret_cast<bf16x8 *>(U0 + w * 8) = q;
            }
        }
        __syncthreads();
        typedef _Float16 h4 __attribute__((ext_vector_type(4)));
        typedef __bf16 b4 __attribute__((ext_vector_type(4)));
        // hidden layers: the lane's four gates of site m of a W x W array whose corner is
        // lattice site (t0 - o, x0 - o), read from the per-solve gate array (halves
        // `half` .. half + 15 of the site's 32) before the group's MFMAs; the four gated
        // channels go to LDS with one 8-byte store
        auto gate_of = [&](int W, int o, int half, int m, int kg) {
            const int a = m / W, b = m - a * W;
            const size_t gs = (size_t)wrapn(t0 - o + a, Nt) * Nx + wrapn(x0 - o + b, Nx);
            return *reinterpret_cast<const h4 *>(gr + gs * 32 + half + 4 * kg);
        };
        auto gated = [&](__bf16 *U, int m, int kg, f32x4 v, h4 gt) {
            b4 o;
#pragma unroll
            for (int i = 0; i < 4; ++i) o[i] = (__bf16)(v[i] * (float)gt[i]);
            *reinterpret_cast<b4 *>(U + m * 16 + 4 * kg) = o;
        };
        conv_lds<8, W1, W0, ksteps(8)>(
            U0, gf1, [&](int m, int kg) { return gate_of(W1, 2, 0, m, kg); },
            [&](int m, int kg, f32x4 v, h4 gt) { gated(U1, m, kg, v, gt); });
        __syncthreads();
        conv_lds<16, W2, W1, ksteps(16)>(
            U1, gf2, [&](int m, int kg) { return gate_of(W2, 1, 16, m, kg); },
            [&](int m, int kg, f32x4 v, h4 gt) { gated(U2, m, kg, v, gt); });
        __syncthreads();
        // output layer (8 channels: lanes kg = 0, 1 hold spinor components 2 kg, 2 kg + 1);
        // r_s and Ap of the lane's two components are loaded before the group's MFMAs
        struct RA {
            float2 r[2], a[2];
        };
        conv_lds<16, W3, W2, ksteps(16)>(
            U2, gf3,
            [&](int m, int kg) {
                RA q{};
                if (kg < 2) {
                    const int a = m / W3, b = m - a * W3;
                    const size_t gs = (size_t)(t0 + a) * Nx + x0 + b;
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const size_t idx = off + (size_t)(2 * kg + h) * V + gs;
                        q.r[h] = rs[idx];
                        q.a[h] = Ap[idx];
                    }
                }
                return q;
            },
            [&](int m, int kg, f32x4 v, const RA &q) {
                if (kg >= 2) return;
                const int a = m / W3, b = m - a * W3;
                const size_t gs = (size_t)(t0 + a) * Nx + x0 + b;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const size_t idx = off + (size_t)(2 * kg + h) * V + gs;
                    const float2 rv = q.r[h], av = q.a[h];
                    const float2 zv = make_float2(c0 * rv.x + sf * v[2 * h], c0 * rv.y + sf * v[2 * h + 1]);
                    z[idx] = zv;
                    acc.x += (double)(zv.x * rv.x) + (double)(zv.y * rv.y);
                    acc.y += (double)(zv.x * av.x) + (double)(zv.y * av.y);
                }
            });
    }
    acc = block_sum2<kCT>(acc, red);
    if (threadIdx.x == 0) part[(size_t)sys * gridDim.x + blockIdx.x] = acc;
}

// ---------------------------------------------------------------------------
// training kernels (layer-wise, channels-last FP32 activations [site][C])
// ---------------------------------------------------------------------------
// out[site][co] = sum_k in[site + off(tap)][ci] W[k][co]; one wave = 16 sites
template <int CIN, int COUT>
__global__ void __launch_bounds__(kCT)
k_conv_fwd(int Nt, int Nx, const float *__restrict__ in, const bf16x8 *__restrict__ frag, float *__restrict__ out) {
    constexpr int KS = ksteps(CIN);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kg = lane >> 4, col = lane & 15;
    const int V = Nt * Nx;
    const int g = blockIdx.x * (kCT / 64) + wave;
    if (g * 16 >= V) return;
    const int site = g * 16 + col;
    const int t = site / Nx, x = site - t * Nx;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const int k0 = ks * 32 + kg * 8;
        const int tap = k0 / CIN, ci0 = k0 % CIN;
        bf16x8 A = bf16x8_zero();
        if (tap < 9) {
            const int dt = tap / 3 - 1, dx = tap % 3 - 1;
            const int s2 = wrapn(t + dt, Nt) * Nx + wrapn(x + dx, Nx);
            const float4 lo = *reinterpret_cast<const float4 *>(in + (size_t)s2 * CIN + ci0);
            const float4 hi = *reinterpret_cast<const float4 *>(in + (size_t)s2 * CIN + ci0 + 4);
            A[0] = (__bf16)lo.x; A[1] = (__bf16)lo.y; A[2] = (__bf16)lo.z; A[3] = (__bf16)lo.w;
            A[4] = (__bf16)hi.x; A[5] = (__bf16)hi.y; A[6] = (__bf16)hi.z; A[7] = (__bf16)hi.w;
        }
        acc = mfma_bf16(A, frag[ks * 64 + lane], acc);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int so = g * 16 + kg * 4 + i;
        if (col < COUT) out[(size_t)so * COUT + col] = acc[i];
    }
}

// partial dW[m][n] = sum_{sites of this block} in[site + off(tap_m)][ci_m] * gy[site][n],
// m = tap * CIN + ci; MFMA with the lattice sites as the reduction dimension.
// Each block reduces kWgSites sites; partials go to wpart[block][9*CIN][16].
constexpr int kWgSites = 1024;
template <int CIN, int COUT>
__global__ void __launch_bounds__(kCT)
k_conv_wgrad(int Nt, int Nx, const float *__restrict__ in, const float *__restrict__ gy, float *__restrict__ wpart) {
    constexpr int M = 9 * CIN, MT = (M + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int kg = lane >> 4, col = lane & 15;
    const int V = Nt * Nx;
    const int s0 = blockIdx.x * kWgSites;
    float *dst = wpart + (size_t)blockIdx.x * M * 16;
    for (int mt = wave; mt < MT; mt += kCT / 64) {
        const int m = mt * 16 + col;          // A row held by this lane
        const bool mvalid = m < M;
        const int tap = mvalid ? m / CIN : 0, ci = mvalid ? m % CIN : 0;
        const int dt = tap / 3 - 1, dx = tap % 3 - 1;
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (int sb = s0; sb < s0 + kWgSites && sb < V; sb += 32) {
            bf16x8 A, B;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int site = sb + kg * 8 + i;
                float av = 0.f, bv = 0.f;
                if (site < V) {
                    const int t = site / Nx, x = site - t * Nx;
                    if (mvalid) av = in[(size_t)(wrapn(t + dt, Nt) * Nx + wrapn(x + dx, Nx)) * CIN + ci];
                    if (col < COUT) bv = gy[(size_t)site * COUT + col];
                }
                A[i] = (__bf16)av;
                B[i] = (__bf16)bv;
            }
            acc = mfma_bf16(A, B, acc);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int mm = mt * 16 + kg * 4 + i;
            if (mm < M) dst[mm * 16 + col] = acc[i];
        }
    }
}

// grad[w0 + m*COUT + n] = sum_blocks wpart[b][m][n] (block order, deterministic)
__global__ void k_wgrad_reduce(const float *__restrict__ wpart, int nblocks, int M, int COUT, float *__restrict__ grad) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * COUT) return;
    const int m = idx / COUT, n = idx - m * COUT;
    double acc = 0.0;
    for (int b = 0; b < nblocks; ++b) acc += (double)wpart[((size_t)b * M + m) * 16 + n];
    grad[idx] = (float)acc;
}

// x0[site][2c+k] = r (FP64 SoA) / s
__global__ void __launch_bounds__(kCT) k_cnn_prep(const double2 *__restrict__ r, int V, double inv, float *__restrict__ x0) {
    const int s = blockIdx.x * kCT + threadIdx.x;
    if (s >= V) return;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double2 v = r[(size_t)c * V + s];
        x0[(size_t)s * 8 + 2 * c] = (float)(v.x * inv);
        x0[(size_t)s * 8 + 2 * c + 1] = (float)(v.y * inv);
    }
}

// d = e - (c0 r + s * y3) (FP32 SoA for the FP32 normal operator)
__global__ void __launch_bounds__(kCT)
k_cnn_resid(const double2 *__restrict__ e, const double2 *__restrict__ r, const float *__restrict__ y3, int V,
            double s, double c0, float2 *__restrict__ d) {
    const int i = blockIdx.x * kCT + threadIdx.x;
    if (i >= V) return;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double2 ev = e[(size_t)c * V + i], rv = r[(size_t)c * V + i];
        d[(size_t)c * V + i] = make_float2((float)(ev.x - (c0 * rv.x + s * y3[(size_t)i * 8 + 2 * c])),
                                           (float)(ev.y - (c0 * rv.y + s * y3[(size_t)i * 8 + 2 * c + 1])));
    }
}

// g3[site][2c+k] = fac * (A d)[c][site].{x,y}
__global__ void __launch_bounds__(kCT)
k_cnn_g3(const float2 *__restrict__ Ad, int V, float fac, float *__restrict__ g3) {
    const int i = blockIdx.x * kCT + threadIdx.x;
    if (i >= V) return;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float2 v = Ad[(size_t)c * V + i];
        g3[(size_t)i * 8 + 2 * c] = fac * v.x;
        g3[(size_t)i * 8 + 2 * c + 1] = fac * v.y;
    }
}

// q[site] = g phi[site] (the gates' input, channels-last with one channel)
__global__ void __launch_bounds__(kCT) k_gate_input(const float *__restrict__ phis, int V, float g, float *__restrict__ q) {
    const int i = blockIdx.x * kCT + threadIdx.x;
    if (i < V) q[i] = g * phis[i];
}

// gm[site][n] = 1 + tanh(sum_tap G[tap][n] q[site + off(tap)]), u = y . gm (16 channels)
__global__ void __launch_bounds__(kCT)
k_gate_fwd(int Nt, int Nx, const float *__restrict__ y, const float *__restrict__ q, const float *__restrict__ G,
           float *__restrict__ gm, float *__restrict__ u) {
    const int s = blockIdx.x * kCT + threadIdx.x;
    if (s >= Nt * Nx) return;
    const int t = s / Nx, x = s - t * Nx;
    float qn[9];
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) qn[tap] = q[wrapn(t + tap / 3 - 1, Nt) * Nx + wrapn(x + tap % 3 - 1, Nx)];
#pragma unroll
    for (int n = 0; n < 16; ++n) {
        float a = 0.f;
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) a += G[tap * 16 + n] * qn[tap];
        const float gv = gate_fn(a);
        gm[(size_t)s * 16 + n] = gv;
        u[(size_t)s * 16 + n] = y[(size_t)s * 16 + n] * gv;
    }
}

// backward through u = y . gm, gm = 1 + tanh(t): from dL/du (in g) to dL/dy (in g) and
// dL/dt (gt)
__global__ void k_gate_bwd(float *__restrict__ g, const float *__restrict__ y, const float *__restrict__ gm,
                           float *__restrict__ gt, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float th = gm[i] - 1.f;
    gt[i] = g[i] * y[i] * (1.f - th * th);
    g[i] *= gm[i];
}

// <a, b> partials of FP32 SoA spinors, accumulated in FP64
__global__ void __launch_bounds__(kCT)
k_dot_f2(const float2 *__restrict__ a, const float2 *__restrict__ b, int n, double2 *__restrict__ part) {
    __shared__ double2 red[kCT / kWave];
    double2 acc = make_double2(0.0, 0.0);
    for (int i = blockIdx.x * kCT + threadIdx.x; i < n; i += gridDim.x * kCT) {
        const float2 u = a[i], v = b[i];
        acc.x += (double)u.x * v.x + (double)u.y * v.y;
    }
    acc = block_sum2<kCT>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// <a, b> and <a, a> partials of FP64 SoA spinors (for s and <e, r>)
__global__ void __launch_bounds__(kCT)
k_dot2(const double2 *__restrict__ a, const double2 *__restrict__ b, int n, double2 *__restrict__ part) {
    __shared__ double2 red[kCT / kWave];
    double2 acc = make_double2(0.0, 0.0);
    for (int i = blockIdx.x * kCT + threadIdx.x; i < n; i += gridDim.x * kCT) {
        const double2 u = a[i], v = b[i];
        acc.x += u.x * v.x + u.y * v.y;
        acc.y += u.x * u.x + u.y * u.y;
    }
    acc = block_sum2<kCT>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Adam (Kingma & Ba 2015) on the FP32 master weights
__global__ void k_adam(float *__restrict__ W, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    if (!isfinite(gi)) return;
    const float mi = b1 * m[i] + (1.f - b1) * gi, vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    W[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static bool cnn_fits(const Ctx &c) { return c.p.Nt % 16 == 0 && c.p.Nx % 16 == 0; }

// Skip-path weight c0 = 1 / (2r + M)^2 (the scale of the diagonal of D^dag D).
double cnn_c0(const Ctx &c) { return 1.0 / ((2.0 * c.p.r + c.p.M) * (2.0 * c.p.r + c.p.M)); }

// Default weights: identity-like centre taps in the hidden layers, zero gates (gates of
// 1) and a zero output layer, so the untrained preconditioner is z = c0 r (plain CG).
static std::vector<float> default_weights(const Ctx &) {
    std::vector<float> W(kCnnNW, 0.f);
    for (int ci = 0; ci < 8; ++ci) W[(4 * 8 + ci) * 16 + ci] = 0.1f;              // W1 centre
    for (int ci = 0; ci < 16; ++ci) W[kCnnNW1 + (4 * 16 + ci) * 16 + ci] = 1.0f; // W2 centre
    return W;                                                                      // W3 = 0
}

int cnn_make_frags(Ctx &c) {
    hipLaunchKernelGGL(k_make_frags, dim3(5), dim3(256), 0, c.stream, (const float *)c.cnnW, c.cnnF1, c.cnnF2,
                       c.cnnF3, c.cnnD3, c.cnnD2);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int cnn_alloc(Ctx &c) {
    if (!cnn_fits(c)) {
        c.err = "the learned preconditioner needs Nt % 16 == 0 and Nx % 16 == 0";
        return YCN_ERR_UNSUPPORTED;
    }
    YCN_HIP(hipMalloc(&c.cnnW, sizeof(float) * kCnnNW * 4));   // W, Adam m, Adam v, grad
    YCN_HIP(hipMemset(c.cnnW, 0, sizeof(float) * kCnnNW * 4));
    YCN_HIP(hipMalloc(&c.cnnF1, sizeof(bf16x8) * (kFrag1 + kFrag2 + kFrag3 + ksteps(8) * 64 + ksteps(16) * 64)));
    c.cnnF2 = c.cnnF1 + kFrag1;
    c.cnnF3 = c.cnnF2 + kFrag2;
    c.cnnD3 = c.cnnF3 + kFrag3;
    c.cnnD2 = c.cnnD3 + ksteps(8) * 64;
    YCN_HIP(hipMalloc(&c.zs, sizeof(float2) * (size_t)c.nSys * 4 * c.V));
    YCN_HIP(hipMemset(c.zs, 0, sizeof(float2) * (size_t)c.nSys * 4 * c.V));
    YCN_HIP(hipMalloc(&c.cnnGates, sizeof(_Float16) * 32 * (size_t)c.nRep * c.V));
    YCN_HIP(hipMemset(c.cnnGates, 0, sizeof(_Float16) * 32 * (size_t)c.nRep * c.V));
    const std::vector<float> W = default_weights(c);
    YCN_HIP(hipMemcpy(c.cnnW, W.data(), sizeof(float) * kCnnNW, hipMemcpyHostToDevice));
    c.adamStep = 0;
    return cnn_make_frags(c);
}

void cnn_free(Ctx &c) {
    void **bufs[] = {(void **)&c.cnnW, (void **)&c.cnnF1, (void **)&c.zs, (void **)&c.cnnAct, (void **)&c.cnnGates};
    for (void **b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    c.cnnActV = 0;
}

int cnn_set_weights(Ctx &c, const float *w) {
    YCN_HIP(hipMemcpyAsync(c.cnnW, w, sizeof(float) * kCnnNW, hipMemcpyHostToDevice, c.stream));
    YCN_HIP(hipMemsetAsync(c.cnnW + kCnnNW, 0, sizeof(float) * kCnnNW * 2, c.stream));   // reset Adam moments
    c.adamStep = 0;
    int rc = cnn_make_frags(c);
    if (rc) return rc;
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

int cnn_get_weights(Ctx &c, float *w) {
    YCN_HIP(hipMemcpyAsync(w, c.cnnW, sizeof(float) * kCnnNW, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

int cnn_gates(Ctx &c, int nrep) {
    hipLaunchKernelGGL(k_cnn_gates, dim3((unsigned)((c.V + kCT - 1) / kCT), nrep), dim3(kCT), 0, c.stream, c.p.Nt,
                       c.p.Nx, (const float *)c.phis, (float)c.p.g, (const float4 *)(c.cnnW + kCnnOffG1), c.cnnGates);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int cnn_apply(Ctx &c, int nsys, const CgState *st, const float2 *rs, const float2 *Ap, float2 *z, double2 *part) {
    const int tilesX = c.p.Nx / 16, tiles = (c.p.Nt / 16) * tilesX;
    // read r_s, Ap, the gates of both layers; write z
    prof_before(c, PK_CNN, (double)nsys * c.V * 160.0);
    hipLaunchKernelGGL(k_cnn_fused, dim3(tiles, nsys), dim3(kCT), 0, c.stream, c.p.Nt, c.p.Nx, tilesX, c.nNoise,
                       (float)cnn_c0(c), st, rs, Ap, (const _Float16 *)c.cnnGates, z, (const bf16x8 *)c.cnnF1,
                       (const bf16x8 *)c.cnnF2, (const bf16x8 *)c.cnnF3, part);
    prof_after(c, PK_CNN);
    YCN_HIP(hipGetLastError());
    return tiles;
}

// activations for training at lattice volume V: x0[8V] y1[16V] gm1[16V] u1[16V]
// y2[16V] gm2[16V] u2[16V] y3[8V] g3[8V] ga[16V] gb[16V] gt[16V] q[V]
// wpart[blocks*144*16] red[256]
static int ensure_act(Ctx &c) {
    const size_t V = c.V;
    const size_t nwb = (V + kWgSites - 1) / kWgSites;
    const size_t need = V * (8 + 16 * 6 + 8 + 8 + 16 * 3 + 1) + nwb * 144 * 16 + 512;
    if (c.cnnActV == need) return YCN_OK;
    if (c.cnnAct) (void)hipFree(c.cnnAct);
    YCN_HIP(hipMalloc(&c.cnnAct, sizeof(float) * need));
    c.cnnActV = need;
    return YCN_OK;
}

// One training step on the pair (r, e) (FP64 SoA, system layout of one system) at the
// field phis (FP32): forward, loss L = |A d|^2 / |r|^2 with d = e - z (the residual the
// correction leaves), backward, Adam.  Uses FP32 normal-operator scratch (rs, ps, Aps,
// zs of system 0) and the CgState slot cg_d[0].  Returns the loss (before the update)
// in *loss.
int cnn_train_step(Ctx &c, const double2 *r, const double2 *e, const float *phis, float lr, double *loss,
                   bool update) {
    int rc = ensure_act(c);
    if (rc) return rc;
    const int Nt = c.p.Nt, Nx = c.p.Nx, V = c.V;
    float *x0 = c.cnnAct, *y1 = x0 + (size_t)8 * V, *gm1 = y1 + (size_t)16 * V, *u1 = gm1 + (size_t)16 * V;
    float *y2 = u1 + (size_t)16 * V, *gm2 = y2 + (size_t)16 * V, *u2 = gm2 + (size_t)16 * V;
    float *y3 = u2 + (size_t)16 * V, *g3 = y3 + (size_t)8 * V, *ga = g3 + (size_t)8 * V, *gb = ga + (size_t)16 * V;
    float *gt = gb + (size_t)16 * V, *q = gt + (size_t)16 * V;
    float *wpart = q + (size_t)V;
    const int nwb = (V + kWgSites - 1) / kWgSites;
    double2 *red = reinterpret_cast<double2 *>(wpart + (size_t)nwb * 144 * 16);
    float *grad = c.cnnW + 3 * kCnnNW;
    const float *G1 = c.cnnW + kCnnOffG1, *G2 = c.cnnW + kCnnOffG2;
    // s = RMS(r), |r|^2
    const int nred = 64;
    hipLaunchKernelGGL(k_dot2, dim3(nred), dim3(kCT), 0, c.stream, r, r, 4 * V, red);
    std::vector<double2> hr(nred);
    YCN_HIP(hipMemcpyAsync(hr.data(), red, sizeof(double2) * nred, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    double rr = 0;
    for (int i = 0; i < nred; ++i) rr += hr[i].y;
    if (!(rr > 0)) { *loss = 0; return YCN_OK; }
    const double s = std::sqrt(rr / (8.0 * V));
    const dim3 gs((V + kCT - 1) / kCT), gc((V / 16 + 3) / 4);
    const dim3 g16((16 * V + 255) / 256);
    // forward
    hipLaunchKernelGGL(k_cnn_prep, gs, dim3(kCT), 0, c.stream, r, V, 1.0 / s, x0);
    hipLaunchKernelGGL(k_gate_input, gs, dim3(kCT), 0, c.stream, phis, V, (float)c.p.g, q);
    hipLaunchKernelGGL((k_conv_fwd<8, 16>), gc, dim3(kCT), 0, c.stream, Nt, Nx, (const float *)x0,
                       (const bf16x8 *)c.cnnF1, y1);
    hipLaunchKernelGGL(k_gate_fwd, gs, dim3(kCT), 0, c.stream, Nt, Nx, (const float *)y1, (const float *)q, G1, gm1, u1);
    hipLaunchKernelGGL((k_conv_fwd<16, 16>), gc, dim3(kCT), 0, c.stream, Nt, Nx, (const float *)u1,
                       (const bf16x8 *)c.cnnF2, y2);
    hipLaunchKernelGGL(k_gate_fwd, gs, dim3(kCT), 0, c.stream, Nt, Nx, (const float *)y2, (const float *)q, G2, gm2, u2);
    hipLaunchKernelGGL((k_conv_fwd<16, 8>), gc, dim3(kCT), 0, c.stream, Nt, Nx, (const float *)u2,
                       (const bf16x8 *)c.cnnF3, y3);
    // d = e - z; A d, then A (A d) with the FP32 fused normal operator (beta = 0: p' = input)
    hipLaunchKernelGGL(k_cnn_resid, gs, dim3(kCT), 0, c.stream, e, r, (const float *)y3, V, s, cnn_c0(c), c.rs);
    CgState S0{};
    S0.beta = 0.0;
    S0.done = 0;
    YCN_HIP(hipMemcpyAsync(c.cg_d, &S0, sizeof(CgState), hipMemcpyHostToDevice, c.stream));
    launch_fused<float>(c, c.tile32, 1, false, phis, c.cg_d, (const float2 *)c.rs, c.ps[0], c.ps[1], c.Aps,
                        c.partials);
    launch_fused<float>(c, c.tile32, 1, false, phis, c.cg_d, (const float2 *)c.Aps, c.ps[0], c.ps[1], c.zs,
                        c.partials);
    hipLaunchKernelGGL(k_dot_f2, dim3(nred), dim3(kCT), 0, c.stream, (const float2 *)c.Aps, (const float2 *)c.Aps,
                       4 * V, red);
    std::vector<double2> hp(nred);
    YCN_HIP(hipMemcpyAsync(hp.data(), red, sizeof(double2) * nred, hipMemcpyDeviceToHost, c.stream));
    // backward: dL/dy3 = s dL/dz = -2 s A (A d) / |r|^2
    hipLaunchKernelGGL(k_cnn_g3, gs, dim3(kCT), 0, c.stream, (const float2 *)c.zs, V, (float)(-2.0 * s / rr), g3);
    hipLaunchKernelGGL((k_conv_wgrad<16, 8>), dim3(nwb), dim3(kCT), 0, c.stream, Nt, Nx, (const float *)u2,
                       (const float *)g3, wpart);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((144 * 8 + 255) / 256), dim3(256), 0, c.stream, (const float *)wpart, nwb,
                       144, 8, grad + kCnnNW1 + kCnnNW2);
    hipLaunchKernelGGL((k_conv_fwd<8, 16>), gc, dim3(kCT), 0, c.stream, Nt, Nx, (const float *)g3,
                       (const bf16x8 *)c.cnnD3, ga);                                          // dL/du2
    hipLaunchKernelGGL(k_gate_bwd, g16, dim3(256), 0, c.stream, ga, (const float *)y2, (const float *)gm2, gt,
                       (int64_t)16 * V);                                                        // dL/dy2, dL/dt2
    hipLaunchKernelGGL((k_conv_wgrad<1, 16>), dim3(nwb), dim3(kCT), 0, c.stream, Nt, Nx, (const float *)q,
                       (const float *)gt, wpart);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((9 * 16 + 255) / 256), dim3(256), 0, c.stream, (const float *)wpart, nwb,
                       9, 16, grad + kCnnOffG2);
    hipLaunchKernelGGL((k_conv_wgrad<16, 16>), dim3(nwb), dim3(kCT), 0, c.stream, Nt, Nx, (const float *)u1,
                       (const float *)ga, wpart);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((144 * 16 + 255) / 256), dim3(256), 0, c.stream, (const float *)wpart, nwb,
                       144, 16, grad + kCnnNW1);
    hipLaunchKernelGGL((k_conv_fwd<16, 16>), gc, dim3(kCT), 0, c.stream, Nt, Nx, (const float *)ga,
                       (const bf16x8 *)c.cnnD2, gb);                                          // dL/du1
    hipLaunchKernelGGL(k_gate_bwd, g16, dim3(256), 0, c.stream, gb, (const float *)y1, (const float *)gm1, gt,
                       (int64_t)16 * V);                                                        // dL/dy1, dL/dt1
    hipLaunchKernelGGL((k_conv_wgrad<1, 16>), dim3(nwb), dim3(kCT), 0, c.stream, Nt, Nx, (const float *)q,
                       (const float *)gt, wpart);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((9 * 16 + 255) / 256), dim3(256), 0, c.stream, (const float *)wpart, nwb,
                       9, 16, grad + kCnnOffG1);
    hipLaunchKernelGGL((k_conv_wgrad<8, 16>), dim3(nwb), dim3(kCT), 0, c.stream, Nt, Nx, (const float *)x0,
                       (const float *)gb, wpart);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((72 * 16 + 255) / 256), dim3(256), 0, c.stream, (const float *)wpart, nwb,
                       72, 16, grad);
    YCN_HIP(hipGetLastError());
    if (update) {
        ++c.adamStep;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)c.adamStep), c2 = 1.f - std::pow(b2, (float)c.adamStep);
        hipLaunchKernelGGL(k_adam, dim3((kCnnNW + 255) / 256), dim3(256), 0, c.stream, c.cnnW, c.cnnW + kCnnNW,
                           c.cnnW + 2 * kCnnNW, (const float *)grad, kCnnNW, lr, b1, b2, 1e-8f, c1, c2);
        rc = cnn_make_frags(c);
        if (rc) return rc;
    }
    YCN_HIP(hipStreamSynchronize(c.stream));
    double add = 0;
    for (const double2 &v : hp) add += v.x;
    *loss = add / rr;
    return YCN_OK;
}

int cnn_get_grad(Ctx &c, float *g) {
    YCN_HIP(hipMemcpyAsync(g, c.cnnW + 3 * kCnnNW, sizeof(float) * kCnnNW, hipMemcpyDeviceToHost, c.stream));
    YCN_HIP(hipStreamSynchronize(c.stream));
    return YCN_OK;
}

}  // namespace ycn
