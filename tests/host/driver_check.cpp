// Drives ycn::run_langevin (the main.cu control flow) with a deterministic backend so
// the loop, the epsBar calibration, the output files and SIGUSR2 handling are checked
// on the CPU.  Usage: driver_check input.toml outdir [usr2_at_counter]
#include <csignal>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "langevin_driver.hpp"

struct FakeBackend : ycn::StepBackend {
    double epsBar;
    bool adaptive;
    int Nt;
    long long usr2At;
    std::ofstream log;
    FakeBackend(double e, bool a, int nt, long long u, const std::string &dir)
        : epsBar(e), adaptive(a), Nt(nt), usr2At(u), log(dir + "/steps.log") { log << std::setprecision(17); }
    void start(bool hot, uint64_t seed) override { log << "start " << hot << " " << seed << "\n"; }
    double step(uint64_t c) override {
        if ((long long)c == usr2At) std::raise(SIGUSR2);
        const double f = adaptive ? 0.75 + 0.5 * std::fmod(0.6180339887498949 * (double)c, 1.0) : 1.0;
        const double e = epsBar * f;
        log << "step " << c << " " << e << "\n";
        return e;
    }
    void set_eps_bar(double e) override {
        epsBar = e;
        log << "epsbar " << e << "\n";
    }
    ycn::Measurement measure(uint64_t c) override {
        ycn::Measurement m;
        m.magnetisation = {0.5 * (double)c - 40.0};
        for (int t = 0; t < Nt; ++t) {
            m.slices.push_back(t + 0.25 * (double)c);
            m.corr.push_back(std::exp(-0.1 * t) + 1e-3 * (double)c);
        }
        m.trace = 2.0 * (double)c;
        m.phi2 = 3.0 * (double)c;
        m.cgSolves = 2;
        m.cgItersMean = 100.0 + (double)c;
        m.cgTrueRelMax = 1e-11 * (double)c;
        m.cgNotConverged = (int)(c % 2);
        log << "measure " << c << "\n";
        return m;
    }
    std::string describe() const override { return "fake"; }
};

int main(int argc, char **argv) {
    ycn::install_signal_handler();
    std::ifstream f(argv[1]);
    std::stringstream ss;
    ss << f.rdbuf();
    const ycn::RunConfig cfg = ycn::RunConfig::from_toml(ycn::Toml::parse_string(ss.str()));
    const long long usr2 = argc > 3 ? std::atoll(argv[3]) : -1;
    FakeBackend b(cfg.epsBar, cfg.adaptive, cfg.Nt, usr2, argv[2]);
    ycn::DriverIO io;
    std::ofstream out(std::string(argv[2]) + "/stdout.txt");
    io.out = &out;
    io.dir = argv[2];
    io.inputEcho = ss.str();
    const int n = ycn::run_langevin(cfg, b, io);
    std::cout << n << "\n";
    std::cout << "cfg " << cfg.m2 << " " << cfg.lambda << " " << cfg.kappa << " " << cfg.yukawa_coupling << " "
              << cfg.Nt << " " << cfg.Nx << " " << cfg.cgTol << " " << cfg.cgTolMode << " " << cfg.precond << " "
              << cfg.cnnTrainPairs << " " << cfg.cnnPairsPerSolve << " " << cfg.cnnLearningRate << " " << cfg.evenOdd << "\n";
    return 0;
}
