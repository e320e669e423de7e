// cg_fused.hpp -- LDS-tiled fused application of the CG normal operator on gfx950.
//
//   p' = r + beta p_old                          (over the tile plus a 2-site halo)
//   q  = D1 p'                                   (tile plus a 1-site halo, in LDS)
//   Ap = D2 q, <p', Ap>                          (tile; one partial per block, a
//                                                 finalize kernel forms alpha)
// with (D1, D2) = (D, D^dag) for A = D^dag D and (D^dag, D) for A = D D^dag.
//
// One block owns a TT x TX tile of one system; p' is staged once in LDS (SoA planes,
// conflict-free double2/float2 rows), q never leaves LDS.  Compulsory HBM traffic per
// site: read r, p_old, phi; write p', Ap = 4S + P (S = spinor bytes, P = field bytes),
// against 7S + 2P for the two-stencil dataflow.  Tiles are handed to the 8 XCDs in
// contiguous bands so halo re-reads stay in one XCD's L2.
#pragma once
#include "ycn_device.hpp"
#include "ycn_internal.hpp"

namespace ycn {

template <typename R> struct Vec2;
template <> struct Vec2<double> { using T = double2; };
template <> struct Vec2<float> { using T = float2; };

template <typename R>
__device__ __forceinline__ typename Vec2<R>::T mk2(R a, R b) {
    typename Vec2<R>::T v;
    v.x = a;
    v.y = b;
    return v;
}

// one flavour of the Wilson hop (see wilson_flavour in ycn_device.hpp), in precision R
template <bool DAG, typename R, typename V2>
__device__ __forceinline__ void wilson_flavour_r(R hp, R hm, R hr, R h, R d, V2 in0, V2 in1, V2 tf0, V2 tf1, R sf,
                                                 V2 tb0, V2 tb1, R sb, V2 xf0, V2 xf1, V2 xb0, V2 xb1, V2 &o0,
                                                 V2 &o1) {
    const V2 Tb0 = DAG ? tf0 : tb0, Tb1 = DAG ? tf1 : tb1;
    const V2 Tf0 = DAG ? tb0 : tf0, Tf1 = DAG ? tb1 : tf1;
    const R sTb = DAG ? sf : sb, sTf = DAG ? sb : sf;
    const V2 Xf0 = DAG ? xb0 : xf0, Xf1 = DAG ? xb1 : xf1;
    const V2 Xb0 = DAG ? xf0 : xb0, Xb1 = DAG ? xf1 : xb1;
    const R a_hp = hp * sTb, a_hm = hm * sTf, b_hm = hm * sTb, b_hp = hp * sTf;
    o0.x = d * in0.x - (a_hp * Tb0.x + a_hm * Tf0.x + hr * (Xf0.x + Xb0.x) + h * (Xb1.x - Xf1.x));
    o0.y = d * in0.y - (a_hp * Tb0.y + a_hm * Tf0.y + hr * (Xf0.y + Xb0.y) + h * (Xb1.y - Xf1.y));
    o1.x = d * in1.x - (b_hm * Tb1.x + b_hp * Tf1.x + hr * (Xf1.x + Xb1.x) + h * (Xb0.x - Xf0.x));
    o1.y = d * in1.y - (b_hm * Tb1.y + b_hp * Tf1.y + hr * (Xf1.y + Xb1.y) + h * (Xb0.y - Xf0.y));
}

// The hopping part alone (wilson_flavour_r with d = 0): o = H psi (DAG = false) or
// K psi = H^T psi (DAG = true), the even-odd blocks of D and D^dag.
template <bool DAG, typename R, typename V2>
__device__ __forceinline__ void hop_flavour_r(R hp, R hm, R hr, R h, V2 tf0, V2 tf1, R sf, V2 tb0, V2 tb1, R sb,
                                              V2 xf0, V2 xf1, V2 xb0, V2 xb1, V2 &o0, V2 &o1) {
    const V2 Tb0 = DAG ? tf0 : tb0, Tb1 = DAG ? tf1 : tb1;
    const V2 Tf0 = DAG ? tb0 : tf0, Tf1 = DAG ? tb1 : tf1;
    const R sTb = DAG ? sf : sb, sTf = DAG ? sb : sf;
    const V2 Xf0 = DAG ? xb0 : xf0, Xf1 = DAG ? xb1 : xf1;
    const V2 Xb0 = DAG ? xf0 : xb0, Xb1 = DAG ? xf1 : xb1;
    const R a_hp = hp * sTb, a_hm = hm * sTf, b_hm = hm * sTb, b_hp = hp * sTf;
    o0.x = -(a_hp * Tb0.x + a_hm * Tf0.x + hr * (Xf0.x + Xb0.x) + h * (Xb1.x - Xf1.x));
    o0.y = -(a_hp * Tb0.y + a_hm * Tf0.y + hr * (Xf0.y + Xb0.y) + h * (Xb1.y - Xf1.y));
    o1.x = -(b_hm * Tb1.x + b_hp * Tf1.x + hr * (Xf1.x + Xb1.x) + h * (Xb0.x - Xf0.x));
    o1.y = -(b_hm * Tb1.y + b_hp * Tf1.y + hr * (Xf1.y + Xb1.y) + h * (Xb0.y - Xf0.y));
}

// hop_flavour_r at Wilson parameter r = 1, where (r -+ gamma_t) / 2 are projectors: the
// (r - 1) / 2 terms vanish and (r + 1) / 2 = 1, r / 2 = 1 / 2, so each output component
// takes one time neighbour and half the x-neighbour combination
template <bool DAG, typename R, typename V2>
__device__ __forceinline__ void hop_flavour_r1(V2 tf0, V2 tf1, R sf, V2 tb0, V2 tb1, R sb, V2 xf0, V2 xf1, V2 xb0,
                                               V2 xb1, V2 &o0, V2 &o1) {
    const V2 Tb0 = DAG ? tf0 : tb0, Tf1 = DAG ? tb1 : tf1;
    const R sTb = DAG ? sf : sb, sTf = DAG ? sb : sf;
    const V2 Xf0 = DAG ? xb0 : xf0, Xf1 = DAG ? xb1 : xf1;
    const V2 Xb0 = DAG ? xf0 : xb0, Xb1 = DAG ? xf1 : xb1;
    const R hf = (R)0.5;
    o0.x = -(sTb * Tb0.x + hf * ((Xf0.x + Xb0.x) + (Xb1.x - Xf1.x)));
    o0.y = -(sTb * Tb0.y + hf * ((Xf0.y + Xb0.y) + (Xb1.y - Xf1.y)));
    o1.x = -(sTf * Tf1.x + hf * ((Xf1.x + Xb1.x) + (Xb0.x - Xf0.x)));
    o1.y = -(sTf * Tf1.y + hf * ((Xf1.y + Xb1.y) + (Xb0.y - Xf0.y)));
}

struct FusedArgs {
    int Nt, Nx, V, nNoise;
    double r, M, g;
    int tilesX, tilesPerSys;
    int xcdRemap;   // 1: hand contiguous tile bands to the 8 XCDs
};

__device__ __forceinline__ int wrapi(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

template <typename R, int TT, int TX, bool DAG1>
__global__ void __launch_bounds__(TT *TX)
k_cg_normal_fused(FusedArgs A, const R *__restrict__ phi, CgState *st,
                  const typename Vec2<R>::T *__restrict__ r, const typename Vec2<R>::T *__restrict__ pold,
                  typename Vec2<R>::T *__restrict__ pnew, typename Vec2<R>::T *__restrict__ Ap,
                  double2 *__restrict__ part) {
    using V2 = typename Vec2<R>::T;
    constexpr int NT = TT * TX;
    constexpr int W2 = TX + 4, H2 = TT + 4, N2 = W2 * H2;
    constexpr int W1 = TX + 2, H1 = TT + 2, N1 = W1 * H1;
    __shared__ V2 P[4][N2];
    __shared__ V2 Q[4][N1];
    __shared__ R F[N1];
    __shared__ double2 red[NT / kWave > 0 ? NT / kWave : 1];

    const int sys = blockIdx.y;
    if (st[sys].done) return;
    const int nb = gridDim.x;
    int tile = blockIdx.x;
    if (A.xcdRemap) tile = (tile % 8) * (nb / 8) + tile / 8;
    const int ty = tile / A.tilesX, tx = tile - ty * A.tilesX;
    const int t0 = ty * TT, x0 = tx * TX;
    const int Nt = A.Nt, Nx = A.Nx;
    const size_t V = A.V, off = (size_t)sys * 4 * V;
    const R beta = (R)st[sys].beta;
    const R *ph = phi + (size_t)(sys / A.nNoise) * V;

    for (int w = threadIdx.x; w < N2; w += NT) {
        const int row = w / W2, col = w - row * W2;
        const int gt = wrapi(t0 - 2 + row, Nt), gx = wrapi(x0 - 2 + col, Nx);
        const size_t gs = (size_t)gt * Nx + gx;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const V2 a = r[off + c * V + gs], b = pold[off + c * V + gs];
            P[c][w] = mk2<R>(a.x + beta * b.x, a.y + beta * b.y);
        }
    }
    for (int w = threadIdx.x; w < N1; w += NT) {
        const int row = w / W1, col = w - row * W1;
        const int gt = wrapi(t0 - 1 + row, Nt), gx = wrapi(x0 - 1 + col, Nx);
        F[w] = ph[(size_t)gt * Nx + gx];
    }
    __syncthreads();

    const R hp = (R)(0.5 * (A.r + 1.0)), hm = (R)(0.5 * (A.r - 1.0)), hr = (R)(0.5 * A.r), h = (R)0.5;
    const R d0 = (R)(2.0 * A.r + A.M), g = (R)A.g;
    // q = D1 p' on the tile plus its 1-site halo
    for (int w = threadIdx.x; w < N1; w += NT) {
        const int row = w / W1, col = w - row * W1;
        const int gt = wrapi(t0 - 1 + row, Nt);
        const int i2 = (row + 1) * W2 + col + 1;
        const R sf = (gt == Nt - 1) ? (R)-1 : (R)1, sb = (gt == 0) ? (R)-1 : (R)1;
        const R d = d0 + g * F[w];
#pragma unroll
        for (int f = 0; f < 4; f += 2) {
            V2 o0, o1;
            wilson_flavour_r<DAG1, R, V2>(hp, hm, hr, h, d, P[f][i2], P[f + 1][i2], P[f][i2 + W2], P[f + 1][i2 + W2],
                                          sf, P[f][i2 - W2], P[f + 1][i2 - W2], sb, P[f][i2 + 1], P[f + 1][i2 + 1],
                                          P[f][i2 - 1], P[f + 1][i2 - 1], o0, o1);
            Q[f][w] = o0;
            Q[f + 1][w] = o1;
        }
    }
    __syncthreads();

    // Ap = D2 q on the tile, <p', Ap>
    const int lt = threadIdx.x / TX, lx = threadIdx.x - lt * TX;
    const int gt = t0 + lt, gx = x0 + lx;
    const size_t gs = (size_t)gt * Nx + gx;
    const int i1 = (lt + 1) * W1 + lx + 1;
    const int i2 = (lt + 2) * W2 + lx + 2;
    const R sf = (gt == Nt - 1) ? (R)-1 : (R)1, sb = (gt == 0) ? (R)-1 : (R)1;
    const R d = d0 + g * F[i1];
    double2 acc = make_double2(0.0, 0.0);
#pragma unroll
    for (int f = 0; f < 4; f += 2) {
        V2 o0, o1;
        wilson_flavour_r<!DAG1, R, V2>(hp, hm, hr, h, d, Q[f][i1], Q[f + 1][i1], Q[f][i1 + W1], Q[f + 1][i1 + W1], sf,
                                       Q[f][i1 - W1], Q[f + 1][i1 - W1], sb, Q[f][i1 + 1], Q[f + 1][i1 + 1],
                                       Q[f][i1 - 1], Q[f + 1][i1 - 1], o0, o1);
        const V2 p0 = P[f][i2], p1 = P[f + 1][i2];
        pnew[off + f * V + gs] = p0;
        pnew[off + (f + 1) * V + gs] = p1;
        Ap[off + f * V + gs] = o0;
        Ap[off + (f + 1) * V + gs] = o1;
        // conj(p) * Ap accumulated in double
        acc.x += (double)(p0.x * o0.x + p0.y * o0.y) + (double)(p1.x * o1.x + p1.y * o1.y);
        acc.y += (double)(p0.x * o0.y - p0.y * o0.x) + (double)(p1.x * o1.y - p1.y * o1.x);
    }
    acc = block_sum2<NT>(acc, red);
    if (threadIdx.x == 0) part[(size_t)sys * nb + blockIdx.x] = acc;
}

}  // namespace ycn

namespace ycn {

// Tile shapes the autotuner chooses from (TT x TX sites, TT*TX threads); LDS per block
// stays under 64 KiB for the precision each table serves.
struct TileShape {
    int TT, TX;
};
constexpr TileShape kTiles64[] = {{8, 32}, {4, 64}, {16, 16}, {4, 32}};
constexpr TileShape kTiles32[] = {{8, 32}, {4, 64}, {16, 16}, {8, 64}, {16, 32}, {4, 16}};
constexpr int kNumTiles64 = sizeof(kTiles64) / sizeof(kTiles64[0]);
constexpr int kNumTiles32 = sizeof(kTiles32) / sizeof(kTiles32[0]);

template <typename R, int TT, int TX>
inline void launch_fused_tt(hipStream_t s, int nsys, bool dag1, const FusedArgs &A, const R *phi, CgState *st,
                            const typename Vec2<R>::T *r, const typename Vec2<R>::T *pold, typename Vec2<R>::T *pnew,
                            typename Vec2<R>::T *Ap, double2 *part) {
    dim3 grid(A.tilesPerSys, nsys), block(TT * TX);
    if (dag1)
        hipLaunchKernelGGL((k_cg_normal_fused<R, TT, TX, true>), grid, block, 0, s, A, phi, st, r, pold, pnew, Ap,
                           part);
    else
        hipLaunchKernelGGL((k_cg_normal_fused<R, TT, TX, false>), grid, block, 0, s, A, phi, st, r, pold, pnew, Ap,
                           part);
}

inline FusedArgs fused_args(const Ctx &c, int TT, int TX) {
    FusedArgs A;
    A.Nt = c.p.Nt;
    A.Nx = c.p.Nx;
    A.V = c.V;
    A.nNoise = c.nNoise;
    A.r = c.p.r;
    A.M = c.p.M;
    A.g = c.p.g;
    A.tilesX = c.p.Nx / TX;
    A.tilesPerSys = (c.p.Nt / TT) * (c.p.Nx / TX);
    A.xcdRemap = (A.tilesPerSys % 8 == 0) ? 1 : 0;
    return A;
}

inline bool tile_fits(const Ctx &c, TileShape t) { return c.p.Nt % t.TT == 0 && c.p.Nx % t.TX == 0; }

// launch the fused normal operator with tile index `ti` of the table for R
template <typename R>
inline void launch_fused(const Ctx &c, int ti, int nsys, bool dag1, const R *phi, CgState *st,
                         const typename Vec2<R>::T *r, const typename Vec2<R>::T *pold, typename Vec2<R>::T *pnew,
                         typename Vec2<R>::T *Ap, double2 *part);

template <>
inline void launch_fused<double>(const Ctx &c, int ti, int nsys, bool dag1, const double *phi, CgState *st,
                                 const double2 *r, const double2 *pold, double2 *pnew, double2 *Ap, double2 *part) {
    const TileShape t = kTiles64[ti];
    const FusedArgs A = fused_args(c, t.TT, t.TX);
    switch (ti) {
    case 0: launch_fused_tt<double, 8, 32>(c.stream, nsys, dag1, A, phi, st, r, pold, pnew, Ap, part); break;
    case 1: launch_fused_tt<double, 4, 64>(c.stream, nsys, dag1, A, phi, st, r, pold, pnew, Ap, part); break;
    case 2: launch_fused_tt<double, 16, 16>(c.stream, nsys, dag1, A, phi, st, r, pold, pnew, Ap, part); break;
    default: launch_fused_tt<double, 4, 32>(c.stream, nsys, dag1, A, phi, st, r, pold, pnew, Ap, part); break;
    }
}

template <>
inline void launch_fused<float>(const Ctx &c, int ti, int nsys, bool dag1, const float *phi, CgState *st,
                                const float2 *r, const float2 *pold, float2 *pnew, float2 *Ap, double2 *part) {
    const TileShape t = kTiles32[ti];
    const FusedArgs A = fused_args(c, t.TT, t.TX);
    switch (ti) {
    case 0: launch_fused_tt<float, 8, 32>(c.stream, nsys, dag1, A, phi, st, r, pold, pnew, Ap, part); break;
    case 1: launch_fused_tt<float, 4, 64>(c.stream, nsys, dag1, A, phi, st, r, pold, pnew, Ap, part); break;
    case 2: launch_fused_tt<float, 16, 16>(c.stream, nsys, dag1, A, phi, st, r, pold, pnew, Ap, part); break;
    case 3: launch_fused_tt<float, 8, 64>(c.stream, nsys, dag1, A, phi, st, r, pold, pnew, Ap, part); break;
    case 4: launch_fused_tt<float, 16, 32>(c.stream, nsys, dag1, A, phi, st, r, pold, pnew, Ap, part); break;
    default: launch_fused_tt<float, 4, 16>(c.stream, nsys, dag1, A, phi, st, r, pold, pnew, Ap, part); break;
    }
}

}  // namespace ycn
