// update.hip -- bosonic drift, adaptive step and Euler-Maruyama update on gfx950
// (replaces Run/PotentialAndMass/gpuSpMV/Evolve, code/src/langevin_gpu_v2.cu:9-117,
// gpuMaxAbsReduce, code/src/reductions.cu:13-63, and the CSR Laplacian of
// code/src/Laplace.cu:61-100, here an arithmetic 5-point stencil), plus the field
// reductions of the measurement block (reductions.cu:65-140) and layout conversion.
#include <algorithm>

#include "ycn_device.hpp"
#include "ycn_internal.hpp"

namespace ycn {

constexpr int kUT = 256;

// drift += phi (m2 + lambda/6 phi^2) - (sum_nbr phi - 4 phi)
__device__ __forceinline__ double bosonic_drift(const double *__restrict__ ph, int s, int Nt, int Nx, double m2,
                                                double lam6) {
    const int t = s / Nx, x = s - t * Nx;
    const Nbr n = neighbours(t, x, Nt, Nx);
    const double f = ph[s];
    const double pot = f * (m2 + lam6 * (f * f));
    const double lap = (ph[n.tf] + ph[n.tb] + ph[n.xf] + ph[n.xb]) - 4.0 * f;
    return pot - lap;
}

// adaptive path, pass 1: total drift, one partial max |drift| per block
__global__ void __launch_bounds__(kUT)
k_bosonic_max(const double *__restrict__ phi, double *__restrict__ drift, const double *__restrict__ m2, double lam6,
              int Nt, int Nx, double *__restrict__ part, const int *__restrict__ halt) {
    __shared__ double lds[kUT / kWave];
    if (halt && *halt) return;
    const int rep = blockIdx.y;
    const int V = Nt * Nx;
    const int s = blockIdx.x * kUT + threadIdx.x;
    const double *ph = phi + (size_t)rep * V;
    double mx = 0.0;
    if (s < V) {
        const double d = drift[(size_t)rep * V + s] + bosonic_drift(ph, s, Nt, Nx, m2[rep], lam6);
        drift[(size_t)rep * V + s] = d;
        mx = fabs(d);
    }
    mx = block_max<kUT>(mx, lds);
    if (threadIdx.x == 0) part[(size_t)rep * gridDim.x + blockIdx.x] = mx;
}

// adaptive path, finalize (one block per replica): clamp max |drift| to [1e-5, 1e5]
// (langevin_gpu_v2.cu:83-91), eps = epsBar * 30 / max (:112), Langevin time
__global__ void __launch_bounds__(kUT)
k_fin_eps(const double *__restrict__ part, int nb, double epsBar, int adaptive, StepState *__restrict__ st,
          const int *__restrict__ halt) {
    __shared__ double lds[kUT / kWave];
    if (halt && *halt) return;
    const int rep = blockIdx.x;
    double m = 0.0;
    for (int i = threadIdx.x; i < nb; i += kUT) m = fmax(m, part[(size_t)rep * nb + i]);
    m = block_max<kUT>(m, lds);
    if (threadIdx.x) return;
    StepState &S = st[rep];
    if (m > 1e5) { m = 1e5; S.clamps += 1; }
    else if (m < 1e-5) { m = 1e-5; S.clamps += 1; }
    const double e = adaptive ? epsBar * 30.0 / m : epsBar;
    S.maxDrift = m;
    S.eps = e;
    S.tL += e;
}

// adaptive path, pass 2: phi -= eps drift + sqrt(2 eps) noise (Evolve, :29-37)
__global__ void __launch_bounds__(kUT)
k_evolve(double *__restrict__ phi, const double *__restrict__ drift, const double *__restrict__ noise, int V,
         const StepState *__restrict__ st, const int *__restrict__ halt) {
    if (halt && *halt) return;
    const int rep = blockIdx.y;
    const int s = blockIdx.x * kUT + threadIdx.x;
    if (s >= V) return;
    const double e = st[rep].eps;
    const double fact = sqrt(2.0 * e);
    const size_t i = (size_t)rep * V + s;
    phi[i] -= e * drift[i] + fact * noise[i];
}

// fixed-eps path, one pass: phi' = phi - eps (drift_f + bosonic(phi)) - sqrt(2 eps) noise
__global__ void __launch_bounds__(kUT)
k_update_fixed(const double *__restrict__ phi, double *__restrict__ phin, const double *__restrict__ drift,
               const double *__restrict__ noise, const double *__restrict__ m2, double lam6, int Nt, int Nx,
               double eps, StepState *__restrict__ st, const int *__restrict__ halt) {
    if (halt && *halt) return;
    const int rep = blockIdx.y;
    const int V = Nt * Nx;
    const int s = blockIdx.x * kUT + threadIdx.x;
    if (s < V) {
        const size_t i = (size_t)rep * V + s;
        const double d = drift[i] + bosonic_drift(phi + (size_t)rep * V, s, Nt, Nx, m2[rep], lam6);
        phin[i] = phi[i] - (eps * d + sqrt(2.0 * eps) * noise[i]);
    }
    if (s == 0) {
        st[rep].eps = eps;
        st[rep].tL += eps;
    }
}

// per-time-slice sums of phi and phi^2 for one replica (gpuTimeSlices, reductions.cu:65-104)
__global__ void __launch_bounds__(kUT)
k_field_slices(const double *__restrict__ phi, int Nx, double *__restrict__ slices, double *__restrict__ sq) {
    __shared__ double2 lds[kUT / kWave];
    const int t = blockIdx.x;
    double2 acc = make_double2(0.0, 0.0);
    for (int x = threadIdx.x; x < Nx; x += kUT) {
        const double f = phi[(size_t)t * Nx + x];
        acc.x += f;
        acc.y += f * f;
    }
    acc = block_sum2<kUT>(acc, lds);
    if (threadIdx.x == 0) {
        slices[t] = acc.x;
        sq[t] = acc.y;
    }
}

__global__ void __launch_bounds__(kUT) k_hot_start(double *__restrict__ phi, uint64_t seed, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kUT + threadIdx.x;
    if (i >= n) return;
    const uint4 o = philox4x32_10(make_uint4((uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0u, 0u),
                                  make_uint2((uint32_t)seed, 99u));
    const double u = (double)(((((uint64_t)o.x) << 32) | o.y) >> 11) * 0x1p-53;
    phi[i] = 1.0 - 2.0 * u;
}

// reference AoS [site][4] complex <-> SoA [c][site]
__global__ void __launch_bounds__(kUT) k_aos_to_soa(const double2 *__restrict__ aos, double2 *__restrict__ soa, int V) {
    const int s = blockIdx.x * kUT + threadIdx.x;
    const int sys = blockIdx.y;
    if (s >= V) return;
    const size_t o = (size_t)sys * 4 * V;
#pragma unroll
    for (int c = 0; c < 4; ++c) soa[o + (size_t)c * V + s] = aos[o + 4 * (size_t)s + c];
}
__global__ void __launch_bounds__(kUT) k_soa_to_aos(const double2 *__restrict__ soa, double2 *__restrict__ aos, int V) {
    const int s = blockIdx.x * kUT + threadIdx.x;
    const int sys = blockIdx.y;
    if (s >= V) return;
    const size_t o = (size_t)sys * 4 * V;
#pragma unroll
    for (int c = 0; c < 4; ++c) aos[o + 4 * (size_t)s + c] = soa[o + (size_t)c * V + s];
}

__global__ void __launch_bounds__(kUT) k_d2f(const double2 *__restrict__ a, float2 *__restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kUT + threadIdx.x;
    if (i < n) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}
__global__ void __launch_bounds__(kUT) k_f2d(const float2 *__restrict__ a, double2 *__restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kUT + threadIdx.x;
    if (i < n) b[i] = make_double2((double)a[i].x, (double)a[i].y);
}
__global__ void __launch_bounds__(kUT) k_field_f32(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kUT + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------
int spinor_to_f32(Ctx &c, const double2 *a, float2 *b, int nsys) {
    const int64_t n = (int64_t)nsys * 4 * c.V;
    hipLaunchKernelGGL(k_d2f, dim3((unsigned)((n + kUT - 1) / kUT)), dim3(kUT), 0, c.stream, a, b, n);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}
int spinor_to_f64(Ctx &c, const float2 *a, double2 *b, int nsys) {
    const int64_t n = (int64_t)nsys * 4 * c.V;
    hipLaunchKernelGGL(k_f2d, dim3((unsigned)((n + kUT - 1) / kUT)), dim3(kUT), 0, c.stream, a, b, n);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}
int field_to_f32(Ctx &c, const double *a, float *b) {
    hipLaunchKernelGGL(k_field_f32, dim3((unsigned)((c.V + kUT - 1) / kUT)), dim3(kUT), 0, c.stream, a, b,
                       (int64_t)c.V);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int bosonic_update_gated(Ctx &c, int fixedEps, double *phi_in, double *phi_out, const int *halt) {
    const int Nt = c.p.Nt, Nx = c.p.Nx, V = c.V;
    dim3 g((V + kUT - 1) / kUT, c.nRep);
    const double lam6 = c.p.lambda / 6.0;
    if (fixedEps) {
        prof_before(c, PK_BOSONIC, (double)c.nRep * V * 32);
        hipLaunchKernelGGL(k_update_fixed, g, dim3(kUT), 0, c.stream, (const double *)phi_in, phi_out,
                           (const double *)c.drift, (const double *)c.noise, (const double *)c.m2_d, lam6, Nt, Nx,
                           c.epsBar, c.step_d, halt);
        prof_after(c, PK_BOSONIC);
        YCN_HIP(hipGetLastError());
        return YCN_OK;
    }
    // partials: the CG partial buffer holds >= nRep * blocks doubles
    double *part = reinterpret_cast<double *>(c.partials);
    const int nb = (int)g.x;
    prof_before(c, PK_BOSONIC, (double)c.nRep * V * 24);
    hipLaunchKernelGGL(k_bosonic_max, g, dim3(kUT), 0, c.stream, (const double *)phi_in, c.drift,
                       (const double *)c.m2_d, lam6, Nt, Nx, part, halt);
    prof_after(c, PK_BOSONIC);
    hipLaunchKernelGGL(k_fin_eps, dim3(c.nRep), dim3(kUT), 0, c.stream, (const double *)part, nb, c.epsBar,
                       c.p.adaptive, c.step_d, halt);
    hipLaunchKernelGGL(k_evolve, g, dim3(kUT), 0, c.stream, phi_in, (const double *)c.drift,
                       (const double *)c.noise, V, (const StepState *)c.step_d, halt);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int bosonic_update(Ctx &c, int fixedEps) {
    const int rc = bosonic_update_gated(c, fixedEps, c.phi[c.cur], c.phi[c.cur ^ 1], nullptr);
    if (rc) return rc;
    if (fixedEps) c.cur ^= 1;   // the fixed-eps update writes the other buffer of the pair
    return YCN_OK;
}

int field_slices(Ctx &c, int rep, double *slices_dev, double *sums_dev) {
    hipLaunchKernelGGL(k_field_slices, dim3(c.p.Nt), dim3(kUT), 0, c.stream,
                       (const double *)(c.phi[c.cur] + (size_t)rep * c.V), c.p.Nx, slices_dev, sums_dev);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int hot_start(Ctx &c, uint64_t initSeed) {
    const int64_t n = (int64_t)c.nRep * c.V;
    hipLaunchKernelGGL(k_hot_start, dim3((unsigned)((n + kUT - 1) / kUT)), dim3(kUT), 0, c.stream, c.phi[c.cur],
                       initSeed, n);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

int aos_to_soa(Ctx &c, const double *aos_dev, double2 *soa, int nsys) {
    dim3 g((c.V + kUT - 1) / kUT, nsys);
    hipLaunchKernelGGL(k_aos_to_soa, g, dim3(kUT), 0, c.stream, (const double2 *)aos_dev, soa, c.V);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}
int soa_to_aos(Ctx &c, const double2 *soa, double *aos_dev, int nsys) {
    dim3 g((c.V + kUT - 1) / kUT, nsys);
    hipLaunchKernelGGL(k_soa_to_aos, g, dim3(kUT), 0, c.stream, soa, (double2 *)aos_dev, c.V);
    YCN_HIP(hipGetLastError());
    return YCN_OK;
}

}  // namespace ycn
