// Drives the reference-named C++ wrappers of host/ycn_backend.hpp (DiracOP, CGsolver,
// FermionicDrift, ColouredNoise) on a GPU context the way code/src/main.cu drives the
// reference's objects, and writes their outputs as raw doubles for comparison with the
// oracle.  Usage: wrappers_check outdir   (16x16, the 1024^2 workload's couplings)
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "ycn_backend.hpp"

static void dump(const std::string &path, const std::vector<double> &v) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(sizeof(double) * v.size()));
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "usage: wrappers_check outdir\n";
        return 2;
    }
    const std::string out = argv[1];
    ycn_params p;
    ycn_default_params(&p);
    p.Nt = p.Nx = 16;
    p.m2 = -0.4; p.lambda = 1.0; p.g = 0.5; p.M = 0.05; p.r = 1.0; p.cutFraction = 0.5;
    p.epsBar = 0.005; p.adaptive = 0; p.seed = 1234;
    p.cgTol = 1e-10; p.cgTolMode = YCN_TOL_REL_TRUE; p.cgMaxIter = 20000;
    const int V = p.Nt * p.Nx;
    try {
        ycn::Context ctx(p);
        ycn::check(ctx.get(), ycn_hot_start(ctx.get(), 7), "ycn_hot_start");
        std::vector<double> phi(V), noise(V), drift(V), psi(8 * V), Dpsi(8 * V), Ddpsi(8 * V), x(8 * V);
        ycn::check(ctx.get(), ycn_get_field(ctx.get(), phi.data()), "ycn_get_field");
        for (int i = 0; i < 8 * V; ++i) psi[i] = std::sin(0.37 * i) + 0.25 * std::cos(1.9 * i);

        ycn::ColouredNoise cn(ctx);
        cn(3, noise.data());                                     // cn() (main.cu:305)
        ycn::DiracOP D(ctx);
        D.setScalar(phi.data());                                 // DiracOP::setScalar
        D.applyD(psi.data(), Dpsi.data(), ycn::MatrixType::Normal);
        D.applyD(psi.data(), Ddpsi.data(), ycn::MatrixType::Dagger);
        ycn::CGsolver CG(ctx);
        const ycn_cg_stats st = CG.solve(psi.data(), x.data(), D, ycn::MatrixType::Dagger);
        ycn::FermionicDrift fd(ctx);
        const ycn_cg_stats sf = fd.getForce(drift.data(), 0);   // fDrift.getForce (main.cu:306)
        dump(out + "/phi.bin", phi);
        dump(out + "/noise.bin", noise);
        dump(out + "/psi.bin", psi);
        dump(out + "/Dpsi.bin", Dpsi);
        dump(out + "/Ddagpsi.bin", Ddpsi);
        dump(out + "/x.bin", x);
        dump(out + "/drift.bin", drift);
        std::printf("cg %d %.3e %d force %d %.3e %d\n", st.iters, st.true_rel_resid, st.converged, sf.iters,
                    sf.true_rel_resid, sf.converged);
    } catch (const ycn::Error &e) {
        std::cerr << e.what() << "\n";
        return e.code == YCN_ERR_NO_DEVICE ? 3 : 1;
    }
    return 0;
}
