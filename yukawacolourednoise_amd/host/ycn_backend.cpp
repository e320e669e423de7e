// ycn_backend.cpp -- see ycn_backend.hpp.
#include "ycn_backend.hpp"

#include <algorithm>

#include <sstream>

namespace ycn {

void check(ycn_ctx *ctx, int rc, const char *what, bool allowNotConverged) {
    if (rc == YCN_OK || (allowNotConverged && rc == YCN_ERR_NOT_CONVERGED)) return;
    std::ostringstream s;
    s << what << " failed (" << rc << ")";
    if (ctx) s << ": " << ycn_last_error(ctx);
    throw Error(rc, s.str());
}

Context::Context(const ycn_params &p) : p_(p) {
    const int rc = ycn_create(&p_, &ctx_);
    if (rc == YCN_ERR_NO_DEVICE) throw Error(rc, "ycn_create: no HIP device (there is no CPU fallback)");
    check(nullptr, rc, "ycn_create");
}

Context::~Context() {
    if (ctx_) ycn_destroy(ctx_);
}

ycn_params params_from_config(const RunConfig &cfg) {
    ycn_params p;
    ycn_default_params(&p);
    p.Nt = cfg.Nt;
    p.Nx = cfg.Nx;
    p.nReplicas = cfg.nReplicas;
    p.nNoise = cfg.nNoise;
    p.m2 = cfg.m2;
    p.lambda = cfg.lambda;
    p.g = cfg.yukawa_coupling;
    p.M = cfg.fermion_mass;
    p.r = cfg.WilsonParam;
    p.cutFraction = cfg.cutFraction;
    p.driftMode = cfg.driftMode;
    p.epsBar = cfg.epsBar;
    p.adaptive = cfg.adaptive ? 1 : 0;
    p.seed = (uint64_t)cfg.seed;
    p.cgTol = cfg.cgTol;
    p.cgTolMode = cfg.cgTolMode;
    p.cgMaxIter = cfg.cgMaxIter;
    p.cgPrecision = cfg.cgPrecision;
    p.reliableEvery = cfg.reliableEvery;
    p.precond = cfg.precond;
    p.evenOdd = cfg.evenOdd ? 1 : 0;
    return p;
}

void DiracOP::applyD(const double *in, double *out, MatrixType t) {
    check(c_.get(), ycn_dirac_apply(c_.get(), phi_, in, out, (int)t), "DiracOP::applyD");
}

ycn_cg_stats CGsolver::solve(const double *in, double *out, DiracOP &D, MatrixType t, const double *phi) {
    ycn_cg_stats st{};
    check(c_.get(), ycn_cg_solve(c_.get(), phi ? phi : D.scalar(), in, out, (int)t, &st), "CGsolver::solve", true);
    return st;
}

ycn_cg_stats FermionicDrift::getForce(double *out, uint64_t step) {
    ycn_cg_stats st{};
    check(c_.get(), ycn_fermion_force(c_.get(), step, out, &st), "FermionicDrift::getForce", true);
    return st;
}

void ColouredNoise::operator()(uint64_t step, double *noiseOut) {
    check(c_.get(), ycn_coloured_noise(c_.get(), step, noiseOut), "ColouredNoise");
}

YcnBackend::YcnBackend(const RunConfig &cfg) : cfg_(cfg), ctx_(params_from_config(cfg)), eps_(cfg.nReplicas) {}

void YcnBackend::start(bool hot, uint64_t initSeed) {
    if (hot) {
        check(ctx_.get(), ycn_hot_start(ctx_.get(), initSeed), "ycn_hot_start");
    } else {
        std::vector<double> zero((size_t)cfg_.nReplicas * cfg_.Nt * cfg_.Nx, 0.0);
        check(ctx_.get(), ycn_set_field(ctx_.get(), zero.data()), "ycn_set_field");
    }
}

// With the learned preconditioner the first steps are training steps: the same Langevin
// step, its fermion solve by plain FP64 CG, whose iterates give the (residual, error)
// pairs for online training, until cnnTrainPairs pairs have been used.
double YcnBackend::step(uint64_t counter) {
    if (cfg_.precond == YCN_PRECOND_CNN && trainedPairs_ < cfg_.cnnTrainPairs) {
        ycn_train_stats ts{};
        check(ctx_.get(),
              ycn_precond_train(ctx_.get(), counter, 1, cfg_.cnnPairsPerSolve, cfg_.cnnLearningRate, &ts),
              "ycn_precond_train");
        trainedPairs_ += ts.pairs > 0 ? ts.pairs : 1;
        return ts.eps_last;
    }
    check(ctx_.get(), ycn_langevin_step(ctx_.get(), counter, eps_.data(), &last_), "ycn_langevin_step", true);
    solves_ += 1;
    itersSum_ += last_.iters_mean;
    trueRelMax_ = std::max(trueRelMax_, last_.true_rel_resid);
    notConverged_ += last_.converged ? 0 : 1;
    return eps_[0];
}

void YcnBackend::set_eps_bar(double epsBar) {
    check(ctx_.get(), ycn_set_eps_bar(ctx_.get(), epsBar), "ycn_set_eps_bar");
}

Measurement YcnBackend::measure(uint64_t counter) {
    Measurement m;
    ycn_observables ob{};
    m.slices.resize(cfg_.Nt);
    m.corr.resize(cfg_.Nt);
    check(ctx_.get(), ycn_measure(ctx_.get(), 0, counter, &ob, m.slices.data(), m.corr.data()), "ycn_measure");
    m.magnetisation = {ob.magnetisation};
    m.trace = ob.trace;
    m.phi2 = ob.phi2;
    m.cgSolves = solves_;
    m.cgNotConverged = notConverged_;
    m.cgItersMean = solves_ ? itersSum_ / solves_ : 0.0;
    m.cgTrueRelMax = trueRelMax_;
    solves_ = notConverged_ = 0;
    itersSum_ = trueRelMax_ = 0.0;
    return m;
}

std::string YcnBackend::describe() const {
    std::ostringstream s;
    s << "libycn (gfx950 HIP), lattice " << cfg_.Nt << "x" << cfg_.Nx << ", CG "
      << (cfg_.cgPrecision ? "mixed FP32/FP64" : "FP64") << ", tol " << cfg_.cgTol
      << (cfg_.cgTolMode ? " relative (true residual)" : " absolute (recursive)");
    return s.str();
}

}  // namespace ycn
