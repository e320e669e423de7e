// rocFFT cross-check of the coloured-noise filter (colouredNoise.cpp:155-178):
//   out = IFFT( mask * FFT(xi) )
// with rocFFT's double-precision real-to-complex forward and complex-to-real inverse
// transforms on the GPU; the half-spectrum mask (Nt x (Nx/2 + 1), which holds 1 / (Nt Nx)
// inside the cutoff and 0 outside, as the reference's does) is applied on the host.
//   rocfft_check Nt Nx xi.bin mask.bin out.bin
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(x)                                                           \
    do {                                                                   \
        if (!(x)) {                                                        \
            std::fprintf(stderr, "rocfft_check: %s failed (line %d)\n", #x, __LINE__); \
            return 1;                                                      \
        }                                                                  \
    } while (0)

static bool read_file(const char *path, std::vector<double> &v) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    const size_t n = std::fread(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
    return n == v.size();
}

int main(int argc, char **argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: rocfft_check Nt Nx xi.bin mask.bin out.bin\n");
        return 2;
    }
    const size_t Nt = std::strtoul(argv[1], nullptr, 10), Nx = std::strtoul(argv[2], nullptr, 10);
    const size_t H = Nx / 2 + 1, V = Nt * Nx, NH = Nt * H;
    CHECK(Nt >= 2 && Nx >= 2 && Nx % 2 == 0);
    std::vector<double> xi(V), mask(NH), spec(2 * NH), out(V);
    CHECK(read_file(argv[3], xi));
    CHECK(read_file(argv[4], mask));

    double *dReal = nullptr, *dSpec = nullptr;
    CHECK(hipMalloc(&dReal, sizeof(double) * V) == hipSuccess);
    CHECK(hipMalloc(&dSpec, sizeof(double) * 2 * NH) == hipSuccess);
    CHECK(hipMemcpy(dReal, xi.data(), sizeof(double) * V, hipMemcpyHostToDevice) == hipSuccess);

    CHECK(rocfft_setup() == rocfft_status_success);
    const size_t lengths[2] = {Nx, Nt};   // fastest dimension first
    rocfft_plan fwd = nullptr, inv = nullptr;
    CHECK(rocfft_plan_create(&fwd, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                             rocfft_precision_double, 2, lengths, 1, nullptr) == rocfft_status_success);
    CHECK(rocfft_plan_create(&inv, rocfft_placement_notinplace, rocfft_transform_type_real_inverse,
                             rocfft_precision_double, 2, lengths, 1, nullptr) == rocfft_status_success);
    size_t wf = 0, wi = 0;
    CHECK(rocfft_plan_get_work_buffer_size(fwd, &wf) == rocfft_status_success);
    CHECK(rocfft_plan_get_work_buffer_size(inv, &wi) == rocfft_status_success);
    void *work = nullptr;
    const size_t wsize = wf > wi ? wf : wi;
    rocfft_execution_info info = nullptr;
    CHECK(rocfft_execution_info_create(&info) == rocfft_status_success);
    if (wsize) {
        CHECK(hipMalloc(&work, wsize) == hipSuccess);
        CHECK(rocfft_execution_info_set_work_buffer(info, work, wsize) == rocfft_status_success);
    }

    void *in[1] = {dReal}, *sp[1] = {dSpec};
    CHECK(rocfft_execute(fwd, in, sp, info) == rocfft_status_success);
    CHECK(hipDeviceSynchronize() == hipSuccess);
    CHECK(hipMemcpy(spec.data(), dSpec, sizeof(double) * 2 * NH, hipMemcpyDeviceToHost) == hipSuccess);
    for (size_t k = 0; k < NH; ++k) {
        spec[2 * k] *= mask[k];
        spec[2 * k + 1] *= mask[k];
    }
    CHECK(hipMemcpy(dSpec, spec.data(), sizeof(double) * 2 * NH, hipMemcpyHostToDevice) == hipSuccess);
    CHECK(rocfft_execute(inv, sp, in, info) == rocfft_status_success);
    CHECK(hipDeviceSynchronize() == hipSuccess);
    CHECK(hipMemcpy(out.data(), dReal, sizeof(double) * V, hipMemcpyDeviceToHost) == hipSuccess);

    FILE *f = std::fopen(argv[5], "wb");
    CHECK(f != nullptr);
    CHECK(std::fwrite(out.data(), sizeof(double), V, f) == V);
    std::fclose(f);
    rocfft_execution_info_destroy(info);
    rocfft_plan_destroy(fwd);
    rocfft_plan_destroy(inv);
    rocfft_cleanup();
    if (work) (void)hipFree(work);
    (void)hipFree(dReal);
    (void)hipFree(dSpec);
    std::printf("rocfft_check: %zux%zu filtered\n", Nt, Nx);
    return 0;
}
