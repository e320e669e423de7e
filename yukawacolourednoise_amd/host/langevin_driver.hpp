// langevin_driver.hpp -- the host Langevin loop of code/src/main.cu:73-516 over an
// abstract step backend: input.toml parameters (main.cu:113-170), burn-in
// (:302-317), epsBar calibration (:325-352), main loop with measurements and the
// traces.csv / data.csv / slice.dat / stdout outputs (:356-502), and the SIGUSR2
// early finish (:65-71,425-428).
//
// The production backend (ycn_backend.cpp) drives libycn.so through the C-ABI; tests
// drive the same loop with a deterministic backend to pin the control flow and the
// file formats without a GPU.
#pragma once
#include <stdint.h>

#include <atomic>
#include <iosfwd>
#include <string>
#include <vector>

#include "toml_lite.hpp"

namespace ycn {

struct RunConfig {
    // [physics] (main.cu:114-131)
    std::string useMass = "false";
    double m2 = 0, lambda = 0, kappa = 0, Lambda = 0, sq2Kappa = 1, cutFraction = 1;
    // [random]
    long long seed = 1234;
    // [langevin]
    double epsBar = 0.02, MaxLangevinTime = 100, ExportTime = 0.5;
    int burnCount = 10, MeasureDriftCount = 10;
    bool adaptive = true;            // new key langevin.adaptive (reference: adaptive)
    // [io]
    std::string configFileName, timeSliceFileName = "slice.dat", startFileName;
    bool exportHDF = false, resume = false;
    // [fermions]
    double fermion_mass = 1, WilsonParam = 1, yukawa_coupling = 0;  // device value (after /sq2Kappa)
    int driftMode = 0;
    // [lattice] (runs.py:36 writes it; params.h:29 default 16x16)
    int Nt = 16, Nx = 16;
    // [cg]  (params.h:13-15 defaults)
    double cgTol = 1e-8;
    int cgTolMode = 0, cgMaxIter = 1000, cgPrecision = 0, reliableEvery = 50, precond = 0;
    // [cg] learned preconditioner: the first Langevin steps train it online until it has
    // seen cnnTrainPairs (residual, error) pairs, cnnPairsPerSolve per fermion solve
    int cnnTrainPairs = 2000, cnnPairsPerSolve = 16;
    double cnnLearningRate = 1e-3;
    // [cg] evenOdd: fermion force through the even-odd preconditioned solve
    bool evenOdd = false;
    // [start]
    bool hotStart = true;
    long long initSeed = 0;
    // [batch]
    int nReplicas = 1, nNoise = 1;

    static RunConfig from_toml(const Toml &t);
};

// FindKappa (main.cu:58-63)
double find_kappa(double m2, double lambda0);

struct Measurement {
    std::vector<double> magnetisation;   // sum_x phi per field component (one component here)
    std::vector<double> slices;          // Nt sums over x of phi
    std::vector<double> corr;            // Nt point-source correlator
    double trace = 0;                    // sum_x drift_x / g
    double phi2 = 0;                     // sum_x phi^2
    // fermion solves of the Langevin steps since the previous measurement
    int cgSolves = 0, cgNotConverged = 0;
    double cgItersMean = 0, cgTrueRelMax = 0;
};

class StepBackend {
public:
    virtual ~StepBackend() = default;
    virtual void start(bool hot, uint64_t initSeed) = 0;
    virtual double step(uint64_t counter) = 0;          // returns eps of the step
    virtual void set_eps_bar(double epsBar) = 0;
    virtual Measurement measure(uint64_t counter) = 0;
    virtual std::string describe() const = 0;
};

struct DriverIO {
    std::ostream *out;               // stdout of the reference
    std::string dir = ".";           // where data.csv / traces.csv / slice.dat go
    std::string inputEcho;           // text of input.toml, echoed with '#' (main.cu:106-111)
};

extern std::atomic<bool> g_early_finish;   // set by SIGUSR2
void install_signal_handler();

// Runs main.cu's control flow; returns the number of measurements.  Besides the
// reference's outputs it writes stats.csv: per measurement the Langevin time, eps,
// <phi^2> and the CG statistics of the solves since the previous measurement.
int run_langevin(const RunConfig &cfg, StepBackend &backend, DriverIO &io);

}  // namespace ycn
