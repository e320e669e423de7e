// noise.hip -- coloured noise on gfx950 (replaces ColouredNoise,
// code/src/colouredNoise.cpp:16-29,40-67,138-178, and its cuFFT D2Z/Z2D calls).
//
// Three LDS passes, all mixed-radix Stockham FFTs (radix 4/2 specialised, any other
// prime factor by a direct DFT):
//   rows_fwd : Philox N(0,1) field generated in registers, rows t and t+1 packed as
//              z = xi_t + i xi_{t+1}, forward FFT along x, untangled into the two
//              half spectra (kx = 0..Nx/2) -> spec[rep][t][kx]
//   cols     : each kx column: forward FFT along t, mask (kx <= kmax[kt] ? 1/V : 0),
//              inverse FFT along t, in place
//   rows_inv : rows t, t+1 re-packed with the Hermitian extension (C2R convention:
//              imaginary parts of the kx = 0 and kx = Nx/2 bins are ignored), inverse
//              FFT along x, real/imag parts -> noise rows t, t+1
// Algorithmic HBM traffic ~ 8 (rows_fwd write) + 16 (cols r/w) + 8 + 8 (rows_inv)
// = 40 bytes per site.
#include <math.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "ycn_device.hpp"
#include "ycn_internal.hpp"

namespace ycn {

constexpr int kFftThreads = 256;
constexpr int kFftLdsBytes = 65536;  // two ping-pong buffers per block

// ---------------------------------------------------------------------------
// LDS Stockham FFT over `lines` lines of length n stored at a[line*n + e].
// Returns the buffer holding the result (a or b).  sign -1 forward, +1 inverse.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 tw_get(const double2 *tw, int e, int sign) {
    const double2 w = tw[e];
    return sign < 0 ? w : make_double2(w.x, -w.y);
}

__device__ double2 *lds_fft(double2 *a, double2 *b, int lines, const FftPlan &P, int sign) {
    const int n = P.n;
    int Ns = 1;
    for (int f = 0; f < P.nf; ++f) {
        const int R = P.radix[f];
        const int nb = n / R;
        const int tstride = n / (Ns * R);
        for (int w = threadIdx.x; w < nb * lines; w += blockDim.x) {
            const int line = w / nb, j = w - line * nb;
            const double2 *xin = a + line * n;
            double2 *yout = b + line * n;
            const int jm = j % Ns;
            const int tstep = jm * tstride;
            const int idx = (j / Ns) * Ns * R + jm;
            if (R == 4) {
                double2 v0 = xin[j], v1 = xin[j + nb], v2 = xin[j + 2 * nb], v3 = xin[j + 3 * nb];
                if (tstep) {
                    v1 = cmul(v1, tw_get(P.tw, tstep, sign));
                    v2 = cmul(v2, tw_get(P.tw, 2 * tstep, sign));
                    v3 = cmul(v3, tw_get(P.tw, 3 * tstep, sign));
                }
                const double2 t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3);
                const double2 d = csub(v1, v3);
                // (v1 - v3) * (-i) forward, * (+i) inverse
                const double2 t3 = sign < 0 ? make_double2(d.y, -d.x) : make_double2(-d.y, d.x);
                yout[idx] = cadd(t0, t2);
                yout[idx + Ns] = cadd(t1, t3);
                yout[idx + 2 * Ns] = csub(t0, t2);
                yout[idx + 3 * Ns] = csub(t1, t3);
            } else if (R == 2) {
                double2 v0 = xin[j], v1 = xin[j + nb];
                if (tstep) v1 = cmul(v1, tw_get(P.tw, tstep, sign));
                yout[idx] = cadd(v0, v1);
                yout[idx + Ns] = csub(v0, v1);
            } else {
                // direct DFT of size R read straight from LDS
                const int nR = n / R;
                for (int k = 0; k < R; ++k) {
                    double2 acc = make_double2(0.0, 0.0);
                    for (int q = 0; q < R; ++q) {
                        double2 v = xin[j + q * nb];
                        const int e1 = q * tstep;                  // Stockham twiddle
                        const int e2 = ((q * k) % R) * nR;         // DFT kernel w_R^{qk}
                        const int e = (e1 + e2) % n;
                        v = cmul(v, tw_get(P.tw, e, sign));
                        acc = cadd(acc, v);
                    }
                    yout[idx + k * Ns] = acc;
                }
            }
        }
        __syncthreads();
        double2 *tmp = a;
        a = b;
        b = tmp;
        Ns *= R;
    }
    return a;
}

// ---------------------------------------------------------------------------
// pass 1: Philox field + forward row FFTs (two rows per complex transform)
// grid: (ceil(Nt / (2*lpb)), nRep)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kFftThreads)
k_noise_rows_fwd(const double *__restrict__ xi, const uint64_t *__restrict__ seeds, uint64_t step,
                 double2 *__restrict__ spec, int Nt, int Nx, int H, int lpb, FftPlan P, const StepCtl *ctl) {
    extern __shared__ double2 lds[];
    if (ctl) {   // a graph-captured step: counter from the control block, nothing while halted
        if (ctl->halt) return;
        step = ctl->step;
    }
    double2 *a = lds, *b = lds + lpb * Nx;
    const int rep = blockIdx.y;
    const int V = Nt * Nx;
    const uint64_t seed = seeds[rep];
    const int pair0 = blockIdx.x * lpb;
    for (int w = threadIdx.x; w < lpb * Nx; w += blockDim.x) {
        const int l = w / Nx, x = w - l * Nx;
        const int t0 = 2 * (pair0 + l), t1 = t0 + 1;
        double v0 = 0.0, v1 = 0.0;
        if (t0 < Nt) {
            const int64_t i0 = (int64_t)t0 * Nx + x;
            if (xi) {
                v0 = xi[(size_t)rep * V + i0];
                if (t1 < Nt) v1 = xi[(size_t)rep * V + i0 + Nx];
            } else {
                const double2 g0 = normal_pair(seed, 0u, step, (uint64_t)(i0 >> 1));
                v0 = (i0 & 1) ? g0.y : g0.x;
                if (t1 < Nt) {
                    const int64_t i1 = i0 + Nx;
                    const double2 g1 = normal_pair(seed, 0u, step, (uint64_t)(i1 >> 1));
                    v1 = (i1 & 1) ? g1.y : g1.x;
                }
            }
        }
        a[w] = make_double2(v0, v1);
    }
    __syncthreads();
    double2 *res = lds_fft(a, b, lpb, P, -1);
    double2 *out = spec + (size_t)rep * Nt * H;
    for (int w = threadIdx.x; w < lpb * H; w += blockDim.x) {
        const int l = w / H, k = w - l * H;
        const int t0 = 2 * (pair0 + l), t1 = t0 + 1;
        if (t0 >= Nt) continue;
        const double2 zk = res[l * Nx + k];
        const double2 zm = res[l * Nx + (k == 0 ? 0 : Nx - k)];
        // A = (Z_k + conj Z_-k)/2 ; B = (Z_k - conj Z_-k)/(2i)
        const double2 A = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
        const double2 dd = make_double2(zk.x - zm.x, zk.y + zm.y);
        const double2 B = make_double2(0.5 * dd.y, -0.5 * dd.x);
        out[(size_t)t0 * H + k] = A;
        if (t1 < Nt) out[(size_t)t1 * H + k] = B;
    }
}

// ---------------------------------------------------------------------------
// pass 2: column FFT along t, mask, inverse FFT along t (in place)
// grid: (ceil(H / lpb), nRep)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kFftThreads)
k_noise_cols(double2 *__restrict__ spec, const int *__restrict__ kmax, int kmaxAll, double invV, int Nt, int H,
             int lpb, FftPlan P, const StepCtl *ctl) {
    extern __shared__ double2 lds[];
    if (ctl && ctl->halt) return;
    double2 *a = lds, *b = lds + lpb * Nt;
    const int rep = blockIdx.y;
    const int k0 = blockIdx.x * lpb;
    const int nl = min(lpb, H - k0);
    double2 *s = spec + (size_t)rep * Nt * H;
    if (k0 > kmaxAll) {
        // every mode of these columns is cut
        for (int w = threadIdx.x; w < nl * Nt; w += blockDim.x) {
            const int t = w / nl, l = w - t * nl;
            s[(size_t)t * H + k0 + l] = make_double2(0.0, 0.0);
        }
        return;
    }
    for (int w = threadIdx.x; w < lpb * Nt; w += blockDim.x) {
        const int t = w / lpb, l = w - t * lpb;
        a[l * Nt + t] = (l < nl) ? s[(size_t)t * H + k0 + l] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    double2 *res = lds_fft(a, b, lpb, P, -1);
    double2 *oth = (res == a) ? b : a;
    for (int w = threadIdx.x; w < lpb * Nt; w += blockDim.x) {
        const int l = w / Nt, kt = w - l * Nt;
        const double m = (k0 + l <= kmax[kt]) ? invV : 0.0;
        res[w] = cscl(m, res[w]);
    }
    __syncthreads();
    double2 *fin = lds_fft(res, oth, lpb, P, +1);
    for (int w = threadIdx.x; w < lpb * Nt; w += blockDim.x) {
        const int t = w / lpb, l = w - t * lpb;
        if (l < nl) s[(size_t)t * H + k0 + l] = fin[l * Nt + t];
    }
}

// ---------------------------------------------------------------------------
// pass 3: inverse row FFTs (C2R), two rows per complex transform
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kFftThreads)
k_noise_rows_inv(const double2 *__restrict__ spec, double *__restrict__ noise, int Nt, int Nx, int H, int lpb,
                 FftPlan P, const StepCtl *ctl) {
    extern __shared__ double2 lds[];
    if (ctl && ctl->halt) return;
    double2 *a = lds, *b = lds + lpb * Nx;
    const int rep = blockIdx.y;
    const int pair0 = blockIdx.x * lpb;
    const double2 *s = spec + (size_t)rep * Nt * H;
    const bool evenNx = (Nx % 2) == 0;
    for (int w = threadIdx.x; w < lpb * Nx; w += blockDim.x) {
        const int l = w / Nx, k = w - l * Nx;
        const int t0 = 2 * (pair0 + l), t1 = t0 + 1;
        double2 A = make_double2(0.0, 0.0), B = make_double2(0.0, 0.0);
        if (t0 < Nt) {
            const bool lower = k < H;
            const int kk = lower ? k : Nx - k;
            A = s[(size_t)t0 * H + kk];
            if (t1 < Nt) B = s[(size_t)t1 * H + kk];
            if (!lower) { A.y = -A.y; B.y = -B.y; }
            if (k == 0 || (evenNx && k == Nx / 2)) { A.y = 0.0; B.y = 0.0; }
        }
        a[w] = make_double2(A.x - B.y, A.y + B.x);
    }
    __syncthreads();
    double2 *res = lds_fft(a, b, lpb, P, +1);
    double *o = noise + (size_t)rep * Nt * Nx;
    for (int w = threadIdx.x; w < lpb * Nx; w += blockDim.x) {
        const int l = w / Nx, x = w - l * Nx;
        const int t0 = 2 * (pair0 + l), t1 = t0 + 1;
        if (t0 >= Nt) continue;
        const double2 z = res[w];
        o[(size_t)t0 * Nx + x] = z.x;
        if (t1 < Nt) o[(size_t)t1 * Nx + x] = z.y;
    }
}

// twiddle table tw[e] = exp(-2 pi i e / n), e < n (both entry points build it this way,
// so their transforms agree bit for bit)
__global__ void __launch_bounds__(kFftThreads) k_twiddles(double2 *__restrict__ tw, int n) {
    const int e = blockIdx.x * kFftThreads + threadIdx.x;
    if (e >= n) return;
    double s, c;
    sincospi(-2.0 * (double)e / (double)n, &s, &c);
    tw[e] = make_double2(c, s);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// radix factors of n: 4s, then 2s, then odd primes (direct DFTs); false if more than
// kMaxFactors are needed
static bool factor_plan(int n, FftPlan &P) {
    P.n = n;
    P.nf = 0;
    int m = n;
    while (m % 4 == 0 && P.nf < kMaxFactors) { P.radix[P.nf++] = 4; m /= 4; }
    while (m % 2 == 0 && P.nf < kMaxFactors) { P.radix[P.nf++] = 2; m /= 2; }
    for (int q = 3; m > 1 && P.nf < kMaxFactors; q += 2)
        while (m % q == 0 && P.nf < kMaxFactors) { P.radix[P.nf++] = q; m /= q; }
    return m == 1;
}

static int make_plan(Ctx &c, int n, FftPlan &P) {
    if (!factor_plan(n, P)) { c.err = "FFT length has too many factors"; return YCN_ERR_UNSUPPORTED; }
    double2 *d = nullptr;
    YCN_HIP(hipMalloc(&d, sizeof(double2) * n));
    hipLaunchKernelGGL(k_twiddles, dim3((n + kFftThreads - 1) / kFftThreads), dim3(kFftThreads), 0, c.stream, d, n);
    YCN_HIP(hipGetLastError());
    P.tw = d;
    return YCN_OK;
}

// GlobalCoordinateToMomentumSquared (colouredNoise.cpp:16-29) and the mask build
// (:138-145), evaluated on the host in the same IEEE order (no contraction), reduced
// to kmax[kt] = largest kept kx of row kt (the kept set of a row is a prefix because
// p_x^2 grows with kx).
#pragma clang fp contract(off)
static double momentum_squared(int Nt, int Nx, int ct, int cx) {
    const int sizes[2] = {Nt, Nx}, coo[2] = {ct, cx};
    double msq = 0.0, sgn = 1.0;
    if (coo[1] > (int)std::floor(Nx / 2)) sgn = -1.0;
    for (int n = 0; n < 2; ++n) {
        const int full = sizes[n];
        const double tmp = (coo[n] <= (int)std::floor(full / 2) ? coo[n] : coo[n] - full) * 2.0 * M_PI / full;
        msq += tmp * tmp;
    }
    return msq * sgn;
}

// (s sqrt(d) pi)^2, the cutoff of colouredNoise.cpp:138-145 in its arithmetic order
static double cut_squared(double cutFraction) {
    const double cut = std::sqrt(2.0) * M_PI * cutFraction;
    return cut * cut;
}

static int build_mask(Ctx &c, std::vector<int> &kmax) {
    const int Nt = c.p.Nt, Nx = c.p.Nx, H = c.H;
    const double cutSq = cut_squared(c.p.cutFraction);
    kmax.assign(Nt, -1);
    for (int t = 0; t < Nt; ++t) {
        bool prefix = true;
        for (int k = 0; k < H; ++k) {
            const double m = momentum_squared(Nt, Nx, t, k);
            const bool keep = (m >= 0) && (m <= cutSq);
            if (keep) {
                if (kmax[t] != k - 1) prefix = false;
                kmax[t] = k;
            }
        }
        if (!prefix) { c.err = "coloured-noise mask row is not a prefix"; return YCN_ERR_UNSUPPORTED; }
    }
    return YCN_OK;
}
#pragma clang fp contract(on)

static int lines_per_block(int n) { return std::max(1, std::min(64, kFftLdsBytes / (2 * (int)sizeof(double2) * n))); }

int noise_init(Ctx &c) {
    const int Nt = c.p.Nt, Nx = c.p.Nx;
    if ((size_t)2 * sizeof(double2) * std::max(Nt, Nx) > (size_t)kFftLdsBytes) {
        c.err = "lattice extent too large for the LDS FFT (max 2048)";
        return YCN_ERR_UNSUPPORTED;
    }
    int rc = make_plan(c, Nx, c.planX);
    if (rc) return rc;
    rc = make_plan(c, Nt, c.planT);
    if (rc) return rc;
    std::vector<int> km;
    rc = build_mask(c, km);
    if (rc) return rc;
    c.kmaxAll = -1;
    for (int v : km) c.kmaxAll = std::max(c.kmaxAll, v);
    YCN_HIP(hipMalloc(&c.kmax, sizeof(int) * Nt));
    YCN_HIP(hipMemcpy(c.kmax, km.data(), sizeof(int) * Nt, hipMemcpyHostToDevice));
    YCN_HIP(hipMalloc(&c.spec, sizeof(double2) * (size_t)c.nRep * Nt * c.H));
    return YCN_OK;
}

void noise_free(Ctx &c) {
    if (c.planX.tw) (void)hipFree((void *)c.planX.tw);
    if (c.planT.tw) (void)hipFree((void *)c.planT.tw);
    if (c.kmax) (void)hipFree(c.kmax);
    if (c.spec) (void)hipFree(c.spec);
    c.planX.tw = c.planT.tw = nullptr;
    c.kmax = nullptr;
    c.spec = nullptr;
}

int noise_generate(Ctx &c, uint64_t step, const double *xi_dev, double *out_dev) {
    return noise_generate_on(c, c.stream, step, xi_dev, out_dev);
}

// the three noise kernels on stream s over explicit buffers (a context's, or the
// scratch array of the allocation-free entry point); prof (optional) brackets the
// column pass for a context's kernel timing
static int launch_noise(hipStream_t s, int Nt, int Nx, int nRep, const double *xi_dev, const uint64_t *seeds,
                        uint64_t step, double2 *spec, const int *kmax, int kmaxAll, const FftPlan &planX,
                        const FftPlan &planT, double *out_dev, Ctx *prof, const StepCtl *ctl = nullptr) {
    const int H = Nx / 2 + 1;
    const int lx = lines_per_block(Nx), lt = lines_per_block(Nt);
    const int npairs = (Nt + 1) / 2;
    dim3 g1((npairs + lx - 1) / lx, nRep), g2((H + lt - 1) / lt, nRep);
    const size_t sh1 = (size_t)2 * lx * Nx * sizeof(double2), sh2 = (size_t)2 * lt * Nt * sizeof(double2);
    hipLaunchKernelGGL(k_noise_rows_fwd, g1, dim3(kFftThreads), sh1, s, xi_dev, seeds, step, spec, Nt, Nx, H, lx,
                       planX, ctl);
    if (prof) prof_before(*prof, PK_NOISE_COLS, 32.0 * (double)nRep * Nt * H);
    hipLaunchKernelGGL(k_noise_cols, g2, dim3(kFftThreads), sh2, s, spec, kmax, kmaxAll,
                       1.0 / ((double)Nt * (double)Nx), Nt, H, lt, planT, ctl);
    if (prof) prof_after(*prof, PK_NOISE_COLS);
    hipLaunchKernelGGL(k_noise_rows_inv, g1, dim3(kFftThreads), sh1, s, (const double2 *)spec, out_dev, Nt, Nx, H,
                       lx, planX, ctl);
    return hipGetLastError() == hipSuccess ? YCN_OK : YCN_ERR_HIP;
}

int noise_generate_on(Ctx &c, hipStream_t s, uint64_t step, const double *xi_dev, double *out_dev) {
    const int rc = launch_noise(s, c.p.Nt, c.p.Nx, c.nRep, xi_dev, (const uint64_t *)c.seed_d, step, c.spec,
                                (const int *)c.kmax, c.kmaxAll, c.planX, c.planT, out_dev, &c);
    if (rc) c.err = "coloured-noise kernel launch failed";
    return rc;
}

int noise_generate_ctl(Ctx &c, hipStream_t s, const StepCtl *ctl, double *out_dev) {
    const int rc = launch_noise(s, c.p.Nt, c.p.Nx, c.nRep, nullptr, (const uint64_t *)c.seed_d, 0, c.spec,
                                (const int *)c.kmax, c.kmaxAll, c.planX, c.planT, out_dev, nullptr, ctl);
    if (rc) c.err = "coloured-noise kernel launch failed";
    return rc;
}

// ===========================================================================
// Coloured noise without a context and without allocations
// (ycn_coloured_noise_noalloc, for callers that cannot allocate device memory).
// Its scratch -- the half spectrum, the two twiddle tables, the mask rows and the
// replica seeds -- lives in one file-scope device array of fixed size.  Each device
// holds its own copy at its own address, looked up once per device with
// hipGetSymbolAddress; host code uses only that address, never the array's name.
// Every call on a device shares the array: calls that overlap on one device race.
// Nothing in it needs clearing: each call writes every entry it reads.
// ===========================================================================
constexpr size_t kNaSpec = (size_t)1024 * 513;   // half-spectrum entries (double2): one 1024^2 lattice
constexpr int kNaMaxN = 2048;                     // largest extent (the LDS FFT's limit)
constexpr int kNaMaxRep = 64;                     // replicas (their seeds pass by value)
constexpr int kNaMaxDevices = 64;
constexpr size_t kNaOffTwX = kNaSpec * sizeof(double2);
constexpr size_t kNaOffTwT = kNaOffTwX + kNaMaxN * sizeof(double2);
constexpr size_t kNaOffKmax = kNaOffTwT + kNaMaxN * sizeof(double2);
constexpr size_t kNaOffSeed = kNaOffKmax + kNaMaxN * sizeof(int);
constexpr size_t kNaBytes = kNaOffSeed + kNaMaxRep * sizeof(uint64_t);   // 8 479 232 B

static __device__ __align__(256) unsigned char scratch[kNaBytes];

static void *g_naScratch[kNaMaxDevices];   // the array's address on each device
static std::mutex g_naMutex;
static thread_local const char *g_naErr = nullptr;

// address of `scratch` on device dev, which must be current; looked up once per device
static int na_scratch(int dev, unsigned char **out) {
    std::lock_guard<std::mutex> lock(g_naMutex);
    if (!g_naScratch[dev]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return YCN_ERR_HIP;
        g_naScratch[dev] = ptr;
    }
    *out = (unsigned char *)g_naScratch[dev];
    return YCN_OK;
}

int noise_noalloc_setup(int dev) {
    unsigned char *p = nullptr;
    if (dev < 0 || dev >= kNaMaxDevices) return YCN_ERR_INVALID;
    return na_scratch(dev, &p);
}

const char *noise_noalloc_error() { return g_naErr; }

struct NaSeeds {
    uint64_t v[kNaMaxRep];
};
__global__ void k_na_seeds(NaSeeds s, uint64_t *__restrict__ dst, int n) {
    if ((int)threadIdx.x < n) dst[threadIdx.x] = s.v[threadIdx.x];
}

// kmax[kt] of the mask on the device, in the host's IEEE order (momentum_squared, no
// contraction; correctly rounded double multiply, divide and add), so both entry points
// keep the same modes.  The kept set of a row is a prefix (checked on the host path).
#pragma clang fp contract(off)
__device__ double na_coord_sq(int c, int n) {
    const double k = (double)(c <= n / 2 ? c : c - n);
    const double t = __ddiv_rn(__dmul_rn(__dmul_rn(k, 2.0), M_PI), (double)n);
    return __dmul_rn(t, t);
}
__global__ void __launch_bounds__(kFftThreads) k_na_mask(int *__restrict__ kmax, int Nt, int Nx, int H, double cutSq) {
    const int kt = blockIdx.x * kFftThreads + threadIdx.x;
    if (kt >= Nt) return;
    const double mt = __dadd_rn(0.0, na_coord_sq(kt, Nt));
    int km = -1;
    for (int k = 0; k < H; ++k) {
        const double m = __dadd_rn(mt, na_coord_sq(k, Nx));
        if (!(m >= 0.0 && m <= cutSq)) break;
        km = k;
    }
    kmax[kt] = km;
}
#pragma clang fp contract(on)

// largest kept kx over all rows: row kt = 0 (p_t = 0) keeps the most, since adding
// p_t^2 >= 0 and rounding cannot lower the momentum
static int na_kmax_all(int Nt, int Nx, double cutSq) {
    int km = -1;
    for (int k = 0; k < Nx / 2 + 1; ++k) {
        const double m = momentum_squared(Nt, Nx, 0, k);
        if (!(m >= 0 && m <= cutSq)) break;
        km = k;
    }
    return km;
}

static int na_fail(const char *msg, int code) {
    g_naErr = msg;
    return code;
}

static int na_run(const ycn_params *p, const uint64_t *seeds, uint64_t step, double *noise, hipStream_t s) {
    const int Nt = p->Nt, Nx = p->Nx, nRep = p->nReplicas, H = Nx / 2 + 1;
    FftPlan pX{}, pT{};
    if (!factor_plan(Nx, pX) || !factor_plan(Nt, pT)) return na_fail("FFT length has too many factors", YCN_ERR_UNSUPPORTED);
    unsigned char *base = nullptr;
    if (na_scratch(p->device, &base)) return na_fail("hipGetSymbolAddress of the scratch array failed", YCN_ERR_HIP);
    double2 *spec = (double2 *)base;
    double2 *twX = (double2 *)(base + kNaOffTwX), *twT = (double2 *)(base + kNaOffTwT);
    int *kmax = (int *)(base + kNaOffKmax);
    uint64_t *seed_d = (uint64_t *)(base + kNaOffSeed);
    NaSeeds sd{};
    for (int i = 0; i < nRep; ++i) sd.v[i] = seeds ? seeds[i] : p->seed;
    const double cutSq = cut_squared(p->cutFraction);
    hipLaunchKernelGGL(k_na_seeds, dim3(1), dim3(kNaMaxRep), 0, s, sd, seed_d, nRep);
    hipLaunchKernelGGL(k_twiddles, dim3((Nx + kFftThreads - 1) / kFftThreads), dim3(kFftThreads), 0, s, twX, Nx);
    hipLaunchKernelGGL(k_twiddles, dim3((Nt + kFftThreads - 1) / kFftThreads), dim3(kFftThreads), 0, s, twT, Nt);
    hipLaunchKernelGGL(k_na_mask, dim3((Nt + kFftThreads - 1) / kFftThreads), dim3(kFftThreads), 0, s, kmax, Nt, Nx,
                       H, cutSq);
    pX.tw = twX;
    pT.tw = twT;
    if (launch_noise(s, Nt, Nx, nRep, nullptr, seed_d, step, spec, kmax, na_kmax_all(Nt, Nx, cutSq), pX, pT, noise,
                     nullptr))
        return na_fail("coloured-noise kernel launch failed", YCN_ERR_HIP);
    return YCN_OK;
}

}  // namespace ycn

extern "C" int ycn_coloured_noise_noalloc(const ycn_params *p, const uint64_t *seeds, uint64_t step, double *noise,
                                          void *stream) {
    using namespace ycn;
    if (!p || !noise) return na_fail("ycn_coloured_noise_noalloc: null params or output", YCN_ERR_INVALID);
    if (p->Nt < 2 || p->Nx < 2 || p->nReplicas < 1)
        return na_fail("ycn_coloured_noise_noalloc: needs Nt, Nx >= 2 and nReplicas >= 1", YCN_ERR_INVALID);
    if (p->Nt > kNaMaxN || p->Nx > kNaMaxN)
        return na_fail("ycn_coloured_noise_noalloc: Nt and Nx are limited to 2048", YCN_ERR_INVALID);
    if (p->nReplicas > kNaMaxRep)
        return na_fail("ycn_coloured_noise_noalloc: nReplicas is limited to 64", YCN_ERR_INVALID);
    if ((size_t)p->nReplicas * p->Nt * (p->Nx / 2 + 1) > kNaSpec)
        return na_fail("ycn_coloured_noise_noalloc: nReplicas * Nt * (Nx/2 + 1) exceeds the scratch capacity "
                       "1024 * 513 (one 1024^2 lattice)",
                       YCN_ERR_INVALID);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return na_fail("ycn_coloured_noise_noalloc: no HIP device", YCN_ERR_NO_DEVICE);
    if (p->device < 0 || p->device >= ndev || p->device >= kNaMaxDevices)
        return na_fail("ycn_coloured_noise_noalloc: invalid device", YCN_ERR_INVALID);
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) return na_fail("ycn_coloured_noise_noalloc: hipGetDevice failed", YCN_ERR_HIP);
    if (prev != p->device && hipSetDevice(p->device) != hipSuccess)
        return na_fail("ycn_coloured_noise_noalloc: hipSetDevice failed", YCN_ERR_HIP);
    const int rc = na_run(p, seeds, step, noise, (hipStream_t)stream);
    if (prev != p->device) (void)hipSetDevice(prev);
    return rc;
}
