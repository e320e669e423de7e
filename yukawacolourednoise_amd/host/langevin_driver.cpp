// langevin_driver.cpp -- see langevin_driver.hpp.
#include "langevin_driver.hpp"

#include <chrono>
#include <cmath>
#include <csignal>
#include <fstream>
#include <iomanip>
#include <ostream>
#include <sstream>
#include <stdexcept>

namespace ycn {

constexpr int nDim = 2;

std::atomic<bool> g_early_finish{false};

static void on_usr2(int) { g_early_finish = true; }

void install_signal_handler() { std::signal(SIGUSR2, on_usr2); }

double find_kappa(double m2, double lambda0) {
    const double Delta = (2.0 * nDim + m2) * (2.0 * nDim + m2) + 4.0 * lambda0 / 3.0;
    return (-(2.0 * nDim + m2) + std::sqrt(Delta)) / (2.0 * lambda0 / 3.0);
}

static int parse_choice(const std::string &v, std::initializer_list<const char *> names, const char *key) {
    int i = 0;
    for (const char *n : names) {
        if (v == n) return i;
        ++i;
    }
    throw std::runtime_error(std::string("input: bad value '") + v + "' for " + key);
}

RunConfig RunConfig::from_toml(const Toml &t) {
    RunConfig c;
    // main.cu:114-131
    c.useMass = t.get_string("physics", "useMass");
    if (c.useMass == "true") {
        c.m2 = t.get_double("physics", "mass");
        c.lambda = t.get_double("physics", "g");
        c.kappa = find_kappa(c.m2, c.lambda);
        c.Lambda = c.kappa * c.kappa * c.lambda / 6.0;
    } else {
        c.kappa = t.get_double("physics", "kappa");
        c.Lambda = t.get_double("physics", "lambda");
        c.m2 = ((1.0 - 2.0 * c.Lambda) / c.kappa - 2.0 * nDim);
        c.lambda = (6.0 * c.Lambda / (c.kappa * c.kappa));
    }
    c.sq2Kappa = std::sqrt(2.0 * c.kappa);
    c.cutFraction = t.get_double("physics", "cutFraction");
    c.seed = t.get_int("random", "seed");
    // main.cu:136-141
    c.epsBar = t.get_double("langevin", "averageEpsilon");
    c.MaxLangevinTime = t.get_double("langevin", "MaxLangevinTime");
    c.ExportTime = t.get_double("langevin", "ExportTime");
    c.burnCount = (int)t.get_int("langevin", "burnCount");
    c.MeasureDriftCount = (int)t.get_int("langevin", "MeasureDriftCount");
    c.adaptive = t.opt_bool("langevin", "adaptive", true);
    // main.cu:143-157 (export = "false" is a string: find<bool> throws, stays false)
    c.configFileName = t.get_string("io", "configFileName");
    c.timeSliceFileName = t.get_string("io", "timeSliceFileName");
    c.startFileName = t.opt_string("io", "startFileName", "");
    c.exportHDF = t.opt_bool("io", "export", false);
    c.resume = t.opt_bool("io", "resume", false);
    // main.cu:162-170
    c.fermion_mass = t.get_double("fermions", "fermion_mass");
    c.WilsonParam = t.get_double("fermions", "WilsonParam");
    const double y = t.get_double("fermions", "yukawa_coupling");
    c.yukawa_coupling = (c.useMass == "true") ? y : y / c.sq2Kappa;
    c.driftMode = t.get_int("fermions", "driftMode") == 0 ? 0 : 1;
    // runtime lattice (params.h:29 is compile-time in the reference; runs.py:36 writes [lattice])
    c.Nt = (int)t.opt_int("lattice", "Nt", 16);
    c.Nx = (int)t.opt_int("lattice", "Nx", 16);
    // solver and extensions; defaults reproduce params.h:13-15
    c.cgTol = t.opt_double("cg", "tol", 1e-8);
    c.cgTolMode = parse_choice(t.opt_string("cg", "tolMode", "abs"), {"abs", "rel"}, "cg.tolMode");
    c.cgMaxIter = (int)t.opt_int("cg", "maxIter", 1000);
    c.cgPrecision = parse_choice(t.opt_string("cg", "precision", "fp64"), {"fp64", "mixed"}, "cg.precision");
    c.reliableEvery = (int)t.opt_int("cg", "reliableEvery", 50);
    c.precond = parse_choice(t.opt_string("cg", "precond", "none"), {"none", "cnn"}, "cg.precond");
    c.cnnTrainPairs = (int)t.opt_int("cg", "cnnTrainPairs", 2000);
    c.cnnPairsPerSolve = (int)t.opt_int("cg", "cnnPairsPerSolve", 16);
    c.cnnLearningRate = t.opt_double("cg", "cnnLearningRate", 1e-3);
    c.evenOdd = t.opt_bool("cg", "evenOdd", false);
    if (c.cnnTrainPairs < 0 || c.cnnPairsPerSolve < 1 || !(c.cnnLearningRate > 0))
        throw std::runtime_error("cg.cnnTrainPairs >= 0, cg.cnnPairsPerSolve >= 1 and cg.cnnLearningRate > 0 required");
    c.hotStart = parse_choice(t.opt_string("start", "mode", "hot"), {"cold", "hot"}, "start.mode") == 1;
    c.initSeed = t.opt_int("start", "initSeed", 0);
    c.nReplicas = (int)t.opt_int("batch", "nReplicas", 1);
    c.nNoise = (int)t.opt_int("batch", "nNoise", 1);
    if (c.Nt < 2 || c.Nx < 2) throw std::runtime_error("input: lattice extents must be >= 2");
    if (!(c.ExportTime > 0)) throw std::runtime_error("input: ExportTime must be > 0");
    return c;
}

int run_langevin(const RunConfig &cfg, StepBackend &B, DriverIO &io) {
    std::ostream &out = *io.out;
    const int Nt = cfg.Nt, Nx = cfg.Nx;
    const int N = Nt * Nx, vol = N;
    const int SpatialVolume = Nx;
    out << "Nt: " << Nt << " Nx: " << Nx << std::endl;
    if (!io.inputEcho.empty()) {
        std::istringstream in(io.inputEcho);
        std::string line;
        while (std::getline(in, line)) out << '#' << line << '\n';
        out << "\n\n";
    }
    out << "#backend: " << B.describe() << "\n";

    std::ofstream datafile(io.dir + "/data.csv"), tracefile(io.dir + "/traces.csv");
    datafile << "corr" << "\n";
    tracefile << "tr,sigma,phi" << "\n";

    uint64_t counter = 0;
    double epsBar = cfg.epsBar;
    double h_eps = epsBar;
    B.start(cfg.hotStart, (uint64_t)cfg.initSeed);
    out << std::endl;

    // burn-in (main.cu:302-317)
    for (int burn = 0; burn < cfg.burnCount; ++burn) {
        double t = 0.0;
        while (t < cfg.ExportTime) {
            h_eps = B.step(counter++);
            t += h_eps;
        }
    }
    out << "Thermalization done!" << std::endl;

    int nMeasurements = 0;
    const int oldMeasurements = 0;
    double elapsedLangevinTime = 0.0;
    // epsBar calibration (main.cu:325-352)
    if (cfg.MeasureDriftCount > 0) {
        double epsSum = 0.0;
        while (elapsedLangevinTime < cfg.MeasureDriftCount * cfg.ExportTime) {
            double t = 0.0;
            while (t < cfg.ExportTime) {
                h_eps = B.step(counter++);
                t += h_eps;
            }
            elapsedLangevinTime += t;
            epsSum += h_eps;
            nMeasurements++;
        }
        epsSum /= nMeasurements;
        out << "#Average eps during drift measurement = " << epsSum << std::endl;
        epsBar *= epsBar / epsSum;
    }
    B.set_eps_bar(epsBar);
    out << "Drift measuring done!" << std::endl;

    // main loop (main.cu:356-487)
    nMeasurements = oldMeasurements;
    elapsedLangevinTime = nMeasurements * cfg.ExportTime;
    std::ofstream timeSliceFile(io.dir + "/" + cfg.timeSliceFileName);
    std::ofstream statsFile(io.dir + "/stats.csv");
    statsFile << "time,eps,phi2,cg_solves,cg_iters_mean,cg_true_rel_max,cg_not_converged\n";
    const auto timerStart = std::chrono::high_resolution_clock::now();
    double avg_magnetisation = 0.0;
    double trace = 0.0;   // printed one measurement late, as in main.cu:438 vs :472-482
    while (elapsedLangevinTime < cfg.MaxLangevinTime) {
        double t = 0.0;
        while (t < cfg.ExportTime) {
            h_eps = B.step(counter++);
            t += h_eps;
        }
        elapsedLangevinTime += t;

        Measurement m = B.measure(counter);
        for (int tt = 0; tt < Nt; ++tt) timeSliceFile << m.slices[tt] / SpatialVolume << '\t';
        timeSliceFile << '\n';
        timeSliceFile << '\n';

        out << elapsedLangevinTime << '\t' << h_eps << '\t';
        double sum2 = 0.0;
        for (double &e : m.magnetisation) {
            if (cfg.useMass == "false") e /= cfg.sq2Kappa;
            out << e / N << '\t';
            sum2 += e * e;
        }
        if (g_early_finish) {
            out << "#Early termination signal received.\n#Wrapping up.\n";
            elapsedLangevinTime = cfg.MaxLangevinTime + 1.0;
        }
        nMeasurements++;
        avg_magnetisation += std::abs(m.magnetisation[0]);
        out << "abs. mag: " << (double)std::abs(m.magnetisation[0]) / N << " \t "
            << "cond: " << (double)std::abs(trace) / N << " \t "
            << "partial avg: " << (double)avg_magnetisation / (double)nMeasurements / (double)vol << "\n";

        for (int nt = 0; nt < Nt; nt++) datafile << m.corr[nt] << "\n";

        trace = (cfg.yukawa_coupling != 0.0) ? m.trace : 0.0;
        tracefile << (double)trace / vol << "," << (double)(m.magnetisation[0] / vol) << ","
                  << (double)(std::sqrt(sum2) / vol) << "\n";
        statsFile << elapsedLangevinTime << ',' << h_eps << ',' << m.phi2 / vol << ',' << m.cgSolves << ','
                  << m.cgItersMean << ',' << m.cgTrueRelMax << ',' << m.cgNotConverged << '\n';
        out.flush();
        datafile.flush();
        tracefile.flush();
        timeSliceFile.flush();
        statsFile.flush();
    }
    out << "Final abs. magnetisation: " << (double)avg_magnetisation / (double)nMeasurements / (double)vol << std::endl;
    const auto timerStop = std::chrono::high_resolution_clock::now();
    const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(timerStop - timerStart).count();
    out << "#Number of measurements: " << nMeasurements << '\n';
    out << "#Run time for main loop: " << ms / 1000.0 << "s\n";
    return nMeasurements;
}

}  // namespace ycn
