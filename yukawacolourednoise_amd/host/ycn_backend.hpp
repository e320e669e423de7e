// ycn_backend.hpp -- StepBackend over the libycn C-ABI (include/ycn.h), and thin C++
// classes with the reference's host-side names (DiracOP, CGsolver, FermionicDrift,
// ColouredNoise; code/src/Dirac.cuh:21-22, CGsolver.cuh:36, FermionicDrift.cuh:13-15,
// colouredNoise.h:28-33) so code written against the reference's objects maps
// one-to-one onto the C-ABI.  Errors become exceptions here, on the host side of the
// ABI.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ycn.h"
#include "langevin_driver.hpp"

namespace ycn {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
    int code;
};

void check(ycn_ctx *ctx, int rc, const char *what, bool allowNotConverged = false);

// owning handle
class Context {
public:
    explicit Context(const ycn_params &p);
    ~Context();
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    ycn_ctx *get() const { return ctx_; }
    const ycn_params &params() const { return p_; }

private:
    ycn_ctx *ctx_ = nullptr;
    ycn_params p_;
};

ycn_params params_from_config(const RunConfig &cfg);

// --- reference-named wrappers ------------------------------------------------
enum class MatrixType { Normal = YCN_NORMAL, Dagger = YCN_DAGGER };

// DiracOP<double>::applyD(in, out, MType) (Dirac.cu:23-35), spinors in the reference AoS
class DiracOP {
public:
    explicit DiracOP(Context &c) : c_(c) {}
    void setScalar(const double *phi) { phi_ = phi; }   // host field; nullptr = context field
    const double *scalar() const { return phi_; }
    void applyD(const double *in, double *out, MatrixType t);

private:
    Context &c_;
    const double *phi_ = nullptr;
};

// CGsolver::solve(in, out, D, MType) (CGsolver.cu:42-59): the operator is D's, at D's
// scalar field (setScalar) unless phi is given
class CGsolver {
public:
    explicit CGsolver(Context &c) : c_(c) {}
    ycn_cg_stats solve(const double *in, double *out, DiracOP &D, MatrixType t, const double *phi = nullptr);

private:
    Context &c_;
};

// FermionicDrift::getForce(out, ...) (FermionicDrift.cu:57-80)
class FermionicDrift {
public:
    explicit FermionicDrift(Context &c) : c_(c) {}
    ycn_cg_stats getForce(double *out, uint64_t step);

private:
    Context &c_;
};

// ColouredNoise::operator() (colouredNoise.cpp:155-178)
class ColouredNoise {
public:
    explicit ColouredNoise(Context &c) : c_(c) {}
    void operator()(uint64_t step, double *noiseOut = nullptr);

private:
    Context &c_;
};

// --- the production step backend -------------------------------------------
class YcnBackend : public StepBackend {
public:
    explicit YcnBackend(const RunConfig &cfg);
    void start(bool hot, uint64_t initSeed) override;
    double step(uint64_t counter) override;
    void set_eps_bar(double epsBar) override;
    Measurement measure(uint64_t counter) override;
    std::string describe() const override;
    const ycn_cg_stats &last_stats() const { return last_; }

private:
    RunConfig cfg_;
    Context ctx_;
    ycn_cg_stats last_{};
    std::vector<double> eps_;
    int trainedPairs_ = 0;   // pairs the learned preconditioner has been trained on
    // solves since the last measurement
    int solves_ = 0, notConverged_ = 0;
    double itersSum_ = 0, trueRelMax_ = 0;
};

}  // namespace ycn
