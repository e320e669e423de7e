// cg_host.hpp -- host-side launch helpers shared by the CG solvers.
#pragma once
#include "ycn_internal.hpp"

namespace ycn {

constexpr int kGraphIters = 8;  // CG iterations per captured graph (even: keeps the p parity)

// Launch n CG iterations, launch(i) for i = 0 .. n - 1 (i & 1 is the ping-pong parity; n
// even keeps it across chunks).  When graphs are on and
// no kernel is being profiled, the n iterations are captured once into a hipGraph per
// (field, systems, operator, n, precision, tile) and replayed; every kernel reads its
// scalars from device state and exits early for converged systems, so replays need no
// host-side parameters.  No memset nodes are captured: all state changes are kernels.
// `phi` is the field buffer the captured kernels read: one of the context's two
// ping-pong buffers or its staging buffer for FP64 iterations, nullptr for the FP32
// loops, which read the context's FP32 copy of the field; so each configuration has at
// most two executables and no key outlives the buffer it names.
template <typename Launch>
inline int run_chunk(Ctx &c, const void *phi, int nsys, int mtype, int n, int precision, int tile, Launch launch) {
    if (!c.useGraphs || c.profTarget != PK_NONE) {
        for (int i = 0; i < n; ++i) launch(i);
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
    for (int i = 0; i < n; ++i) launch(i);
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

}  // namespace ycn
