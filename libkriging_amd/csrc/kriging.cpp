#include "kriging.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <limits>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>

#include "lbfgsb.hpp"

namespace lk {

// ---------------------------------------------------------------------------
// process-global state: optimiser knobs (Optim.cpp:16-141, with the LK_* env
// overrides of get_env_or_default), the RNG (Random.cpp) and one runtime per
// device
// ---------------------------------------------------------------------------
namespace {
template <class T>
T env_or(const char* name, T def) {
  const char* v = std::getenv(name);
  if (v == nullptr) return def;
  try {
    if constexpr (std::is_same<T, bool>::value) {
      std::string s(v);
      return s == "1" || s == "true" || s == "TRUE" || s == "True";
    } else if constexpr (std::is_integral<T>::value) {
      return (T)std::stol(v);
    } else {
      return (T)std::stod(v);
    }
  } catch (...) {
    return def;
  }
}

OptimConfig make_optim_config() {
  OptimConfig c;
  c.reparametrize = env_or("LK_REPARAMETRIZE", true);
  c.theta_lower_factor = env_or("LK_THETA_LOWER_FACTOR", 0.02);
  c.theta_upper_factor = env_or("LK_THETA_UPPER_FACTOR", 10.0);
  c.variogram_bounds_heuristic = env_or("LK_VARIOGRAM_BOUNDS_HEURISTIC", true);
  c.log_level = env_or("LK_LOG_LEVEL", 0);
  c.max_restart = env_or("LK_MAX_RESTART", 10);
  c.max_iteration = env_or("LK_MAX_ITERATION", 20);
  c.gradient_tolerance = env_or("LK_GRADIENT_TOLERANCE", 1e-3);
  c.objective_rel_tolerance = env_or("LK_OBJECTIVE_REL_TOLERANCE", 1e-3);
  return c;
}

unsigned g_seed = 123;
std::mt19937 g_engine(123);
std::mutex g_rt_mutex;
std::map<int, std::unique_ptr<Runtime>> g_runtimes;

constexpr double NUGGET_ALPHA_LOWER = 1e-3;  // Kriging.cpp:678
}  // namespace

OptimConfig& optim_config() {
  static OptimConfig c = make_optim_config();
  return c;
}

void random_reset_seed(unsigned seed) {
  g_seed = seed;
  g_engine.seed(seed);
}

int default_device() { return env_or("LKGPU_DEVICE", 0); }

Runtime& runtime_for(int device) {
  std::lock_guard<std::mutex> lock(g_rt_mutex);
  auto it = g_runtimes.find(device);
  if (it == g_runtimes.end()) it = g_runtimes.emplace(device, std::make_unique<Runtime>(device)).first;
  return *it->second;
}

RegModel regmodel_from_string(const std::string& s) {
  if (s == "none") return RegModel::None;
  if (s == "constant") return RegModel::Constant;
  if (s == "linear") return RegModel::Linear;
  throw std::invalid_argument("Unsupported regmodel on this path: " + s + " (constant, linear, none)");
}

int kernel_id(const std::string& name) {
  if (name == "gauss") return K_GAUSS;
  if (name == "exp") return K_EXP;
  if (name == "matern3_2") return K_MATERN32;
  if (name == "matern5_2") return K_MATERN52;
  throw std::invalid_argument("Unsupported covariance kernel: " + name);
}

// ---------------------------------------------------------------------------
Kriging::Kriging(const std::string& kernel, NoiseModel noise) : kernel_(kernel), noise_(noise) {
  kernel_id(kernel);  // validates the name (Covariance::resolve)
  device_ = default_device();
}

Kriging::~Kriging() {
  if (rt_ != nullptr) {
    hipSetDevice(device_);
    if (problem_ >= 0) {
      try {
        rt_->remove_problem(problem_);
      } catch (...) {
      }
    }
    if (d_L_) hipFree(d_L_);
    if (d_Linv_) hipFree(d_Linv_);
    if (d_vec_) hipFree(d_vec_);
  }
}

// fit_setup_impl (KrigingImpl.cpp:764-841): normalisation, trend matrix
void Kriging::setup(const double* y, const double* X, int n, int d, RegModel regmodel, bool normalize) {
  if (n <= 0 || d <= 0) throw std::invalid_argument("empty design");
  const int n_old = n_, d_old = d_;
  const RegModel regmodel_old = regmodel_;
  n_ = n;
  d_ = d;
  regmodel_ = regmodel;
  normalize_ = normalize;
  centerX_.assign(d, 0.0);
  scaleX_.assign(d, 1.0);
  centerY_ = 0.0;
  scaleY_ = 1.0;
  if (normalize) {
    for (int k = 0; k < d; ++k) {
      double mn = X[(size_t)k * n], mx = mn;
      for (int i = 1; i < n; ++i) {
        mn = std::min(mn, X[i + (size_t)k * n]);
        mx = std::max(mx, X[i + (size_t)k * n]);
      }
      centerX_[k] = mn;
      scaleX_[k] = mx - mn;
    }
    double mn = y[0], mx = y[0];
    for (int i = 1; i < n; ++i) {
      mn = std::min(mn, y[i]);
      mx = std::max(mx, y[i]);
    }
    centerY_ = mn;
    scaleY_ = mx - mn;
  }
  std::vector<double> Xn((size_t)n * d), yn(n);
  for (int k = 0; k < d; ++k)
    for (int i = 0; i < n; ++i) Xn[i + (size_t)k * n] = (X[i + (size_t)k * n] - centerX_[k]) / scaleX_[k];
  for (int i = 0; i < n; ++i) yn[i] = (y[i] - centerY_) / scaleY_;
  // the device problem already holds these data (a re-fit on the same
  // design, e.g. NestedKriging's common-prior pass): keep it
  const bool same = problem_ >= 0 && n == n_old && d == d_old && regmodel == regmodel_old && Xn == X_ && yn == y_;
  X_.swap(Xn);
  y_.swap(yn);
  if (!same) upload_data();
  is_empty_ = true;
}

// Trend::regressionModelMatrix (Trend.cpp:34-61) of the normalised design,
// then the design, y and F uploaded as this model's device problem
void Kriging::upload_data() {
  const int n = n_, d = d_;
  p_ = regmodel_ == RegModel::None ? 0 : (regmodel_ == RegModel::Constant ? 1 : 1 + d);
  F_.assign((size_t)n * std::max(p_, 1), 0.0);
  if (p_ > 0)
    for (int i = 0; i < n; ++i) F_[i] = 1.0;
  if (regmodel_ == RegModel::Linear)
    for (int k = 0; k < d; ++k)
      for (int i = 0; i < n; ++i) F_[i + (size_t)(k + 1) * n] = X_[i + (size_t)k * n];
  rt_ = &runtime_for(device_);
  if (problem_ >= 0) rt_->remove_problem(problem_);
  problem_ = -1;
  problem_ = rt_->add_problem(X_.data(), n, d, y_.data(), F_.data(), p_, kernel_id(kernel_), (int)noise_);
}

ModelFlags Kriging::flags_for_eval() const {
  ModelFlags f;
  f.est_sigma2 = est_sigma2_;
  f.est_nugget = est_nugget_;
  f.sigma2 = sigma2_;
  f.nugget = nugget_;
  f.alpha = alpha_;
  return f;
}

// Optim::theta_bounds (Optim.cpp:179-209)
void Kriging::theta_bounds(std::vector<double>& lo, std::vector<double>& up) {
  if (!pre_lo_.empty() && (int)pre_lo_.size() == d_ && (int)pre_up_.size() == d_ && !normalize_) {
    lo = pre_lo_;
    up = pre_up_;
    pre_lo_.clear();
    pre_up_.clear();
    return;
  }
  pre_lo_.clear();
  pre_up_.clear();
  theta_bounds_of(X_.data(), y_.data(), n_, d_, lo, up);
}

void Kriging::theta_bounds_of(const double* X_, const double* y_, int n, int d, std::vector<double>& lo,
                              std::vector<double>& up) {
  const OptimConfig& oc = optim_config();
  lo.assign(d, 0.0);
  up.assign(d, 0.0);
  for (int k = 0; k < d; ++k) {
    double mn = X_[(size_t)k * n], mx = mn;
    for (int i = 1; i < n; ++i) {
      mn = std::min(mn, X_[i + (size_t)k * n]);
      mx = std::max(mx, X_[i + (size_t)k * n]);
    }
    lo[k] = oc.theta_lower_factor * (mx - mn);
    up[k] = oc.theta_upper_factor * (mx - mn);
  }
  if (!oc.variogram_bounds_heuristic || n <= 0) return;
  // slope weights w_ij = (y_i - y_j)^2 / |x_i - x_j|^2 over ordered pairs
  // (NaN -> 0), normalised; lower bound raised to factor * sum w |dx|
  double wsum = 0.0;
  std::vector<double> acc(d, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k < d; ++k) {
        const double v = X_[i + (size_t)k * n] - X_[j + (size_t)k * n];
        s += v * v;
      }
      const double dy = y_[i] - y_[j];
      double w = (dy * dy) / s;
      if (std::isnan(w)) w = 0.0;
      wsum += 2.0 * w;
      for (int k = 0; k < d; ++k) acc[k] += 2.0 * w * std::fabs(X_[i + (size_t)k * n] - X_[j + (size_t)k * n]);
    }
  if (wsum > 0.0) {
    for (int k = 0; k < d; ++k) {
      lo[k] = std::max(lo[k], oc.theta_lower_factor * (acc[k] / wsum));
      lo[k] = std::min(lo[k], up[k]);
      up[k] = std::max(lo[k], up[k]);
    }
  }
}

// ---------------------------------------------------------------------------
// likelihood API
// ---------------------------------------------------------------------------
double Kriging::logLikelihoodFun(const double* gamma, int gamma_n, bool want_grad, double* grad) {
  double ll = 0.0;
  logLikelihoodFunBatch(1, gamma, gamma_n, want_grad, &ll, grad);
  return ll;
}

void Kriging::logLikelihoodFunBatch(int nb, const double* gamma, int gamma_n, bool want_grad, double* ll,
                                    double* grad) {
  if (rt_ == nullptr) throw std::runtime_error("model is not fitted (no data)");
  if (gamma_n != d_ && !(noise_ == NoiseModel::Nugget && gamma_n == d_ + 1))
    throw std::invalid_argument("theta must have length " + std::to_string(d_) +
                                (noise_ == NoiseModel::Nugget ? " or " + std::to_string(d_ + 1) : std::string()));
  std::vector<EvalRequest> reqs(nb);
  for (int i = 0; i < nb; ++i) {
    reqs[i].problem = problem_;
    reqs[i].gamma = gamma + (size_t)i * gamma_n;
    reqs[i].gamma_n = gamma_n;
    reqs[i].flags = flags_for_eval();
  }
  std::vector<EvalResult> res;
  rt_->evaluate(reqs, want_grad, res);
  for (int i = 0; i < nb; ++i) {
    if (res[i].status != EVAL_OK)
      throw std::runtime_error(chol_failure_message(res[i].status, res[i].inc, rt_->chol.num_nugget));
    ll[i] = res[i].ll;
    if (want_grad)
      for (int k = 0; k < gamma_n; ++k) grad[(size_t)i * gamma_n + k] = res[i].grad[k];
  }
}

// Kriging::leaveOneOutFun (Kriging.cpp:470-476): theta has d entries
double Kriging::leaveOneOutFun(const double* theta, int theta_n, bool want_grad, double* grad) {
  if (rt_ == nullptr) throw std::runtime_error("model is not fitted (no data)");
  if (noise_ != NoiseModel::None)
    throw std::invalid_argument("LOO objective not supported for Nugget/Heterogeneous noise modes");
  if (theta_n != d_) throw std::invalid_argument("theta must have length " + std::to_string(d_));
  std::vector<EvalRequest> reqs(1);
  reqs[0].problem = problem_;
  reqs[0].gamma = theta;
  reqs[0].gamma_n = theta_n;
  reqs[0].flags = flags_for_eval();
  std::vector<EvalResult> res;
  rt_->evaluate(reqs, want_grad, res, nullptr, OBJ_LOO);
  if (res[0].status != EVAL_OK)
    throw std::runtime_error(chol_failure_message(res[0].status, res[0].inc, rt_->chol.num_nugget));
  if (want_grad)
    for (int k = 0; k < theta_n; ++k) grad[k] = res[0].grad[k];
  return res[0].ll;
}

// Kriging::leaveOneOutVec (Kriging.cpp:478-485): yhat = y - errors and
// sqrt(sigma2LOO) * sqrt(sigma2), in the model's (normalised) units
void Kriging::leaveOneOutVec(const double* theta, int theta_n, double* yhat, double* stderr_out) {
  leaveOneOutFun(theta, theta_n, false, nullptr);
  std::vector<double> e(n_), s2(n_);
  rt_->read_slot(0, 7, e.data());
  rt_->read_slot(0, 8, s2.data());
  const double sd = std::sqrt(sigma2_);
  for (int i = 0; i < n_; ++i) {
    if (yhat) yhat[i] = y_[i] - e[i];
    if (stderr_out) stderr_out[i] = std::sqrt(s2[i]) * sd;
  }
}

double Kriging::leaveOneOut() {
  if (is_empty_) throw std::runtime_error("model is not fitted");
  return leaveOneOutFun(theta_.data(), (int)theta_.size(), false, nullptr);
}

// Kriging::logMargPostFun (Kriging.cpp:652-660): gamma = theta (d; the
// nugget model appends the current alpha) or [theta, alpha] (d + 1)
double Kriging::logMargPostFun(const double* gamma, int gamma_n, bool want_grad, double* grad) {
  if (rt_ == nullptr) throw std::runtime_error("model is not fitted (no data)");
  if (gamma_n != d_ && !(noise_ == NoiseModel::Nugget && gamma_n == d_ + 1))
    throw std::invalid_argument("theta must have length " + std::to_string(d_) +
                                (noise_ == NoiseModel::Nugget ? " or " + std::to_string(d_ + 1) : std::string()));
  std::vector<double> g(gamma, gamma + gamma_n);
  if (noise_ == NoiseModel::Nugget && gamma_n == d_) g.push_back(alpha_);
  std::vector<EvalRequest> reqs(1);
  reqs[0].problem = problem_;
  reqs[0].gamma = g.data();
  reqs[0].gamma_n = (int)g.size();
  reqs[0].flags = flags_for_eval();
  std::vector<EvalResult> res;
  rt_->evaluate(reqs, want_grad, res, nullptr, OBJ_LMP);
  if (res[0].status != EVAL_OK)
    throw std::runtime_error(chol_failure_message(res[0].status, res[0].inc, rt_->chol.num_nugget));
  if (want_grad)
    for (int k = 0; k < gamma_n; ++k) grad[k] = res[0].grad[k];
  return res[0].ll;
}

double Kriging::logMargPost() {
  if (is_empty_) throw std::runtime_error("model is not fitted");
  return logMargPostFun(theta_.data(), (int)theta_.size(), false, nullptr);
}

double Kriging::logLikelihood() {
  if (is_empty_) throw std::runtime_error("model is not fitted");
  std::vector<double> g(theta_);
  if (noise_ == NoiseModel::Nugget) g.push_back(alpha_);
  return logLikelihoodFun(g.data(), (int)g.size(), false, nullptr);
}

// ---------------------------------------------------------------------------
// fit
// ---------------------------------------------------------------------------
void Kriging::fit(const double* y, const double* X, int n, int d, RegModel regmodel, bool normalize,
                  const std::string& optim, const std::string& objective, const Parameters& P) {
  std::unique_ptr<FitJob> job = begin_fit(y, X, n, d, regmodel, normalize, optim, objective, P);
  run_fit_jobs({job.get()});
}

namespace {
// gamma <-> (theta, alpha) reparametrisations (Optim.cpp:49-67, Kriging.cpp:674-702)
struct Reparam {
  int d;
  bool nugget;
  bool log;
  void to_gamma(const double* theta, double alpha, double* g) const {
    for (int k = 0; k < d; ++k) g[k] = log ? std::log(theta[k]) : theta[k];
    if (nugget) g[d] = log ? -std::log(1.0 + NUGGET_ALPHA_LOWER - alpha) : alpha;
  }
  void from_gamma(const double* g, double* theta, double* alpha) const {
    for (int k = 0; k < d; ++k) theta[k] = log ? std::exp(g[k]) : g[k];
    if (nugget) *alpha = log ? 1.0 + NUGGET_ALPHA_LOWER - std::exp(-g[d]) : g[d];
  }
  // gradient in gamma of sign * objective from its gradient in (theta,
  // alpha): sign -1 for the minimised -LL, +1 for LOO
  void to_gamma_grad(const double* theta, double alpha, const double* gll, double* out, double sign) const {
    for (int k = 0; k < d; ++k) out[k] = sign * (log ? gll[k] * theta[k] : gll[k]);
    if (nugget) out[d] = sign * (log ? gll[d] * (1.0 + NUGGET_ALPHA_LOWER - alpha) : gll[d]);
  }
};

struct StartWorker {
  int idx = 0;
  std::vector<double> theta_start;
  double extra0 = 0.0;
  std::vector<double> glo, gup;  // local gamma bounds
  std::vector<double> gamma;     // current start / optimiser point
  std::unique_ptr<LBFGSB> opt;
  int retry = 0;
  double best_f = std::numeric_limits<double>::infinity();
  std::vector<double> best_gamma;
  bool done = false, success = true;
  std::string error;
  double objective = std::numeric_limits<double>::infinity();
};
}  // namespace

// State of one fit between begin_fit and the commit: the multistart
// workers of a BFGS fit (none for optim = "none") and the committed point.
struct FitState : FitJob {
  Kriging* m = nullptr;
  bool bfgs = false;
  int d = 0, gd = 0;
  bool nug = false;
  int obj = OBJ_LL;  // LL and LMP are maximised (the optimiser sees -f), LOO minimised
  bool no_restart = false;  // update's warm start: one optimiser run, no restart rules
  int n = 0;
  Reparam rp{0, false, true};
  double sign() const { return obj == OBJ_LOO ? 1.0 : -1.0; }
  std::vector<double> th_lo, th_up;
  std::vector<StartWorker> W;
  ModelFlags flags;
  // per-slot (theta, alpha) handed to the runtime for the current batch
  std::vector<std::vector<double>> th_alpha;
  // result: the point to commit
  std::vector<double> theta;
  double alpha = 1.0;

  int nslots() const { return bfgs ? (int)W.size() : 0; }

  // fill this job's requests; returns how many are active
  int requests(EvalRequest* reqs, bool final_pass) {
    int nact = 0;
    for (int s = 0; s < (int)W.size(); ++s) {
      StartWorker& w = W[s];
      const double* g = nullptr;
      if (final_pass)
        g = w.success ? w.best_gamma.data() : nullptr;
      else if (!w.done && w.opt->need_eval())
        g = w.opt->x().data();
      reqs[s].problem = m->problem_;
      reqs[s].gamma_n = gd;
      reqs[s].flags = flags;
      reqs[s].active = g != nullptr;
      if (g != nullptr) {
        double a = 1.0;
        rp.from_gamma(g, th_alpha[s].data(), &a);
        if (nug) th_alpha[s][d] = a;
        ++nact;
      }
      reqs[s].gamma = th_alpha[s].data();
    }
    m->fit_evals_ += nact;
    return nact;
  }

  // feed one batch of results to the optimisers; restart rules of
  // Kriging.cpp:1994-2041 when a run ends
  void consume(const EvalRequest* reqs, const EvalResult* res) {
    const OptimConfig& oc = optim_config();
    std::vector<double> gbuf(gd), th(d), theta_part(d);
    for (int s = 0; s < (int)W.size(); ++s) {
      StartWorker& w = W[s];
      if (!reqs[s].active) continue;
      if (res[s].status != EVAL_OK) {
        // the reference's objective throws here and the start is dropped
        // (Kriging.cpp:2075-2081)
        w.done = true;
        w.success = false;
        w.error = chol_failure_message(res[s].status, res[s].inc, m->rt_->chol.num_nugget);
        continue;
      }
      const double f = sign() * res[s].ll;
      double a = 1.0;
      rp.from_gamma(w.opt->x().data(), th.data(), &a);
      rp.to_gamma_grad(th.data(), a, res[s].grad, gbuf.data(), sign());
      w.opt->tell(f, gbuf);
      if (!w.opt->done()) continue;
      w.gamma = w.opt->x_opt();
      const double fopt = w.opt->f_opt();
      if (fopt < w.best_f) {
        w.best_f = fopt;
        w.best_gamma = w.gamma;
      }
      double a_tmp;
      rp.from_gamma(w.gamma.data(), theta_part.data(), &a_tmp);
      double sol_to_lb = std::numeric_limits<double>::infinity();
      for (int k = 0; k < d; ++k) sol_to_lb = std::min(sol_to_lb, std::fabs(theta_part[k] - th_lo[k]));
      const bool restart = !no_restart && (w.retry < oc.max_restart) &&
                           (w.opt->task().rfind("ABNORMAL_TERMINATION_IN_LNSRCH", 0) == 0 ||
                            w.opt->num_iters() <= 2 || sol_to_lb < std::numeric_limits<double>::epsilon() ||
                            fopt > w.best_f);
      if (restart) {
        std::vector<double> rth(d);
        const double div = std::pow(2.0, w.retry + 1);
        for (int k = 0; k < d; ++k) rth[k] = (w.theta_start[k] + th_lo[k]) / div;
        w.retry++;
        restart_point(w, rth);
      } else {
        w.done = true;
      }
    }
  }

  // best start from the final objective values (Kriging.cpp:2097-2119)
  void select_best(const EvalResult* res) {
    int best = -1, nsucc = 0;
    double min_ofn = std::numeric_limits<double>::infinity();
    for (int s = 0; s < (int)W.size(); ++s) {
      StartWorker& w = W[s];
      if (!w.success) continue;
      if (res[s].status != EVAL_OK) {
        w.success = false;
        continue;
      }
      w.objective = sign() * res[s].ll;
      ++nsucc;
      if (w.objective < min_ofn) {
        min_ofn = w.objective;
        best = s;
      }
    }
    if (nsucc == 0) throw std::runtime_error("All " + std::to_string(W.size()) + " optimization attempts failed");
    theta.assign(d, 0.0);
    alpha = 1.0;
    rp.from_gamma(W[best].best_gamma.data(), theta.data(), &alpha);
    m->est_theta_ = true;
    m->alpha_ = alpha;
  }

  void restart_point(StartWorker& w, const std::vector<double>& th) {
    const OptimConfig& oc = optim_config();
    w.gamma.assign(gd, 0.0);
    rp.to_gamma(th.data(), nug ? w.extra0 : 1.0, w.gamma.data());
    for (int k = 0; k < gd; ++k) {
      w.glo[k] = std::min(w.gamma[k], w.glo[k]);
      w.gup[k] = std::max(w.gamma[k], w.gup[k]);
    }
    w.opt = std::make_unique<LBFGSB>(gd, oc.lbfgs_memory);
    w.opt->max_iter = oc.max_iteration;
    // LOO tolerances scaled by 1 / n^2 (Kriging.cpp:1953-1958)
    const double scale = obj == OBJ_LOO ? 1.0 / ((double)n * n) : 1.0;
    w.opt->pgtol = oc.gradient_tolerance * scale;
    w.opt->factr = oc.objective_rel_tolerance / 1e-13 * scale;
    w.opt->start(w.gamma, w.glo, w.gup);
  }
};

std::unique_ptr<FitJob> Kriging::begin_fit(const double* y, const double* X, int n, int d, RegModel regmodel,
                                           bool normalize, const std::string& optim, const std::string& objective,
                                           const Parameters& P) {
  if (objective != "LL" && objective != "LOO" && objective != "LMP")
    throw std::invalid_argument("Unsupported objective on this path: " + objective + " (LL, LOO, LMP)");
  // Kriging::make_fit_objective (Kriging.cpp:1471-1474)
  if (objective == "LOO" && noise_ != NoiseModel::None)
    throw std::invalid_argument("LOO objective not supported for Nugget/Heterogeneous noise modes");
  int multistart = 1;
  const bool none = optim == "none";
  if (!none) {
    if (optim.rfind("BFGS", 0) != 0)
      throw std::runtime_error("Unsupported optim: " + optim + " (supported are: none, BFGS[#])");
    // Optim::parse_method("BFGS64") -> 64 starts
    const std::string rest = optim.substr(4);
    size_t k = 0;
    while (k < rest.size() && std::isdigit((unsigned char)rest[k])) ++k;
    if (k > 0) multistart = std::max(1, std::atoi(rest.substr(0, k).c_str()));
  }
  optim_ = optim;
  objective_ = objective;
  setup(y, X, n, d, regmodel, normalize);
  fit_evals_ = 0;
  fit_batches_ = 0;

  // beta fixed or estimated (KrigingImpl.cpp:816-826)
  est_beta_ = P.is_beta_estim && regmodel != RegModel::None;
  if (!est_beta_ && P.beta.has_value() && !P.beta->empty()) {
    est_beta_ = false;
    beta_ = *P.beta;
    if ((int)beta_.size() != p_) throw std::invalid_argument("beta must have length " + std::to_string(p_));
    if (normalize_)
      for (double& b : beta_) b /= scaleY_;
  } else {
    est_beta_ = true;
  }
  // theta0 rows (normalised)
  std::vector<double> theta0;
  int rows0 = 0;
  if (P.theta.has_value()) {
    theta0 = *P.theta;
    rows0 = P.theta_rows;
    if (rows0 <= 0 || (int)theta0.size() != rows0 * d)
      throw std::runtime_error("Dimension of theta should be nx" + std::to_string(d));
    if (normalize_)
      for (int r = 0; r < rows0; ++r)
        for (int k = 0; k < d; ++k) theta0[(size_t)r * d + k] /= scaleX_[k];
  }

  auto st = std::make_unique<FitState>();
  st->m = this;
  st->d = d;
  st->nug = noise_ == NoiseModel::Nugget;
  st->obj = objective == "LOO" ? OBJ_LOO : objective == "LMP" ? OBJ_LMP : OBJ_LL;
  st->n = n;
  st->gd = d + (st->nug ? 1 : 0);

  if (none) {
    if (!P.theta.has_value())
      throw std::runtime_error("Theta should be given (1x" + std::to_string(d) + ") matrix, when optim=none");
    std::vector<double> th(theta0.begin(), theta0.begin() + d);
    est_theta_ = false;
    double sigma2 = -1;
    est_sigma2_ = P.is_sigma2_estim;
    if (P.sigma2.has_value()) {
      sigma2 = *P.sigma2;
      if (normalize_) sigma2 /= scaleY_ * scaleY_;
    } else {
      est_sigma2_ = true;
    }
    double alpha = 1.0, nugget_param = 0.0;
    if (noise_ == NoiseModel::Nugget) {
      est_nugget_ = P.is_nugget_estim;
      if (P.nugget.has_value()) {
        nugget_param = *P.nugget;
        if (normalize_) nugget_param /= scaleY_ * scaleY_;
      }
      alpha = (sigma2 > 0 && (sigma2 + nugget_param) > 0) ? sigma2 / (sigma2 + nugget_param)
                                                           : 1.0 - NUGGET_ALPHA_LOWER;
      sigma2_ = sigma2;
      nugget_ = nugget_param;
    } else if (!est_sigma2_) {
      sigma2_ = sigma2;
    }
    alpha_ = alpha;
    st->theta = th;
    st->alpha = alpha;
    return st;
  }

  // ---- BFGS multistart (Kriging.cpp:1591-2120) ----
  const OptimConfig& oc = optim_config();
  const bool nug = st->nug;
  const int gd = st->gd;
  st->bfgs = true;
  theta_bounds(st->th_lo, st->th_up);
  const std::vector<double>& th_lo = st->th_lo;
  const std::vector<double>& th_up = st->th_up;

  // start points (Kriging.cpp:1708-1765): Random::init(), then
  // theta_lower + randu(multistart, d) % (upper - lower), column-major fill
  g_engine.seed(g_seed);
  std::uniform_real_distribution<double> unif{};
  std::vector<double> rand0((size_t)multistart * d);  // row-major rows x d
  for (int k = 0; k < d; ++k)
    for (int r = 0; r < multistart; ++r) rand0[(size_t)r * d + k] = th_lo[k] + unif(g_engine) * (th_up[k] - th_lo[k]);
  std::vector<double> thetas;
  if (P.theta.has_value()) {
    multistart = std::max(multistart, rows0);
    thetas = theta0;
    thetas.insert(thetas.end(), rand0.begin(), rand0.end());
    thetas.resize((size_t)multistart * d);
  } else {
    thetas = rand0;
  }
  // extra0 (alpha) for the nugget model (Kriging.cpp:1774-1783)
  std::vector<double> extra0;
  double ext_lo = 0.0, ext_up = 1.0;
  if (nug) {
    ext_lo = NUGGET_ALPHA_LOWER;
    ext_up = 1.0;
    if (P.sigma2.has_value() && P.nugget.has_value()) {
      const double s = *P.sigma2, nu = *P.nugget;
      extra0.push_back((s > 0 && (s + nu) > 0) ? s / (s + nu) : ext_lo + (ext_up - ext_lo) * 0.5);
    } else {
      for (int r = 0; r < multistart; ++r) {
        const double u = unif(g_engine);
        extra0.push_back(ext_lo + (ext_up - ext_lo) * (1.0 - u * u * u));
      }
    }
  }
  st->rp = Reparam{d, nug, oc.reparametrize};
  std::vector<double> glo(gd), gup(gd);
  st->rp.to_gamma(th_lo.data(), ext_lo, glo.data());
  st->rp.to_gamma(th_up.data(), ext_up, gup.data());
  // estimation flags (Kriging.cpp:1832-1849)
  est_sigma2_ = P.is_sigma2_estim;
  if (!est_sigma2_ && P.sigma2.has_value()) {
    sigma2_ = *P.sigma2;
    if (normalize_) sigma2_ /= scaleY_ * scaleY_;
  } else {
    est_sigma2_ = true;
  }
  if (nug) {
    est_nugget_ = P.is_nugget_estim;
    if (!est_nugget_ && P.nugget.has_value()) {
      nugget_ = *P.nugget;
      if (normalize_) nugget_ /= scaleY_ * scaleY_;
    } else {
      est_nugget_ = true;
    }
  }
  st->W.resize(multistart);
  st->th_alpha.assign(multistart, std::vector<double>(gd));
  for (int s = 0; s < multistart; ++s) {
    StartWorker& w = st->W[s];
    w.idx = s;
    w.theta_start.assign(thetas.begin() + (size_t)s * d, thetas.begin() + (size_t)(s + 1) * d);
    w.extra0 = nug ? extra0[s % extra0.size()] : 1.0;
    w.glo = glo;
    w.gup = gup;
    st->restart_point(w, w.theta_start);
    w.best_gamma = w.gamma;
  }
  st->flags = flags_for_eval();
  return st;
}

void Kriging::run_fit_jobs(const std::vector<FitJob*>& jobs_in) {
  std::vector<FitState*> jobs;
  for (FitJob* j : jobs_in) {
    FitState* f = dynamic_cast<FitState*>(j);
    if (f == nullptr) throw std::invalid_argument("run_fit_jobs: not a job of begin_fit");
    jobs.push_back(f);
  }
  if (jobs.empty()) return;
  Runtime* rt = jobs[0]->m->rt_;
  for (FitState* f : jobs)
    if (f->m->rt_ != rt) throw std::invalid_argument("run_fit_jobs: all models must use the same device");
  // every job's starts keep fixed slots for the whole optimisation (inactive
  // ones stay in place), so the pool layout and the captured graph are
  // reused from one step to the next
  std::vector<int> base(jobs.size() + 1, 0);
  for (size_t j = 0; j < jobs.size(); ++j) base[j + 1] = base[j] + jobs[j]->nslots();
  const int total = base.back();
  const bool batched = std::all_of(jobs.begin(), jobs.end(), [](FitState* f) { return f->m->batched_fit_; });
  const int obj = jobs[0]->obj;
  for (FitState* f : jobs)
    if (f->obj != obj) throw std::invalid_argument("run_fit_jobs: all models must share the objective");
  std::vector<EvalRequest> reqs(total);
  std::vector<EvalResult> res;
  auto evaluate = [&](bool want_grad) {
    if (batched) {
      rt->evaluate(reqs, want_grad, res, nullptr, obj);
      for (FitState* f : jobs)
        if (f->bfgs) f->m->fit_batches_++;
      return;
    }
    res.assign(total, EvalResult{});
    std::vector<EvalRequest> one(1);
    std::vector<EvalResult> r1;
    for (size_t j = 0; j < jobs.size(); ++j)
      for (int s = base[j]; s < base[j + 1]; ++s) {
        if (!reqs[s].active) continue;
        one[0] = reqs[s];
        rt->evaluate(one, want_grad, r1, nullptr, obj);
        res[s] = r1[0];
        jobs[j]->m->fit_batches_++;
      }
  };
  if (total > 0) {
    for (;;) {
      int nact = 0;
      for (size_t j = 0; j < jobs.size(); ++j) nact += jobs[j]->requests(reqs.data() + base[j], false);
      if (nact == 0) break;
      evaluate(true);
      for (size_t j = 0; j < jobs.size(); ++j) jobs[j]->consume(reqs.data() + base[j], res.data() + base[j]);
    }
    // final objective at each start's best point (fit_ofn(best_gamma, nullptr))
    for (size_t j = 0; j < jobs.size(); ++j) jobs[j]->requests(reqs.data() + base[j], true);
    evaluate(false);
    for (size_t j = 0; j < jobs.size(); ++j)
      if (jobs[j]->bfgs) jobs[j]->select_best(res.data() + base[j]);
  }
  // commit every model in one batch, one slot each
  std::vector<std::vector<double>> gam(jobs.size());
  std::vector<EvalRequest> creq(jobs.size());
  for (size_t j = 0; j < jobs.size(); ++j) {
    gam[j] = jobs[j]->theta;
    if (jobs[j]->nug) gam[j].push_back(jobs[j]->alpha);
    creq[j] = jobs[j]->m->commit_request(gam[j]);
    creq[j].flags.alpha = jobs[j]->alpha;
  }
  std::vector<EvalResult> cres;
  rt->evaluate(creq, false, cres);
  // every model's copies first, one synchronisation, then the host state;
  // a model that failed commits none of its state (the ones before it are
  // committed, as one commit after the other would leave them)
  size_t ncopied = 0;
  std::exception_ptr err;
  for (; ncopied < jobs.size(); ++ncopied) {
    try {
      jobs[ncopied]->m->commit_copy((int)ncopied, cres[ncopied]);
    } catch (...) {
      err = std::current_exception();
      break;
    }
  }
  hip_check(hipStreamSynchronize(rt->stream()), "sync");
  for (size_t j = 0; j < ncopied; ++j) jobs[j]->m->commit_finish(jobs[j]->theta, jobs[j]->alpha, (int)j, cres[j]);
  if (err) std::rethrow_exception(err);
}

// ---------------------------------------------------------------------------
// update (Kriging::update, Kriging.cpp:2433-2633): append m points
// ---------------------------------------------------------------------------
void Kriging::update(const double* y_u, const double* X_u, int m, bool refit) {
  if (is_empty_) throw std::runtime_error("model is not fitted");
  if (m < 0) throw std::invalid_argument("negative number of new points");
  const int n = n_, d = d_, N = n + m;
  if (refit && optim_ != "none" && noise_ == NoiseModel::Nugget) {
    // nugget model: a full re-fit on the de-normalised data, from the
    // current parameters unless every one is estimated (Kriging.cpp:2447-2467)
    std::vector<double> y_all(N), X_all((size_t)N * d);
    for (int i = 0; i < n; ++i) y_all[i] = y_[i] * scaleY_ + centerY_;
    for (int i = 0; i < m; ++i) y_all[n + i] = y_u[i];
    for (int k = 0; k < d; ++k) {
      for (int i = 0; i < n; ++i) X_all[i + (size_t)k * N] = X_[i + (size_t)k * n] * scaleX_[k] + centerX_[k];
      for (int i = 0; i < m; ++i) X_all[n + i + (size_t)k * N] = X_u[i + (size_t)k * m];
    }
    Parameters P;
    if (!(est_beta_ && est_nugget_ && est_sigma2_ && est_theta_)) {
      P.sigma2 = sigma2_ * scaleY_ * scaleY_;
      P.is_sigma2_estim = est_sigma2_;
      std::vector<double> th(d);
      for (int k = 0; k < d; ++k) th[k] = theta_[k] * scaleX_[k];
      P.theta = th;
      P.theta_rows = 1;
      P.is_theta_estim = est_theta_;
      P.nugget = nugget_ * scaleY_ * scaleY_;
      P.is_nugget_estim = est_nugget_;
      if (!est_beta_) {
        std::vector<double> b(beta_);
        for (double& v : b) v *= scaleY_;
        P.beta = b;
        P.is_beta_estim = false;
      }
    }
    const std::string optim = optim_, objective = objective_;
    fit(y_all.data(), X_all.data(), N, d, regmodel_, normalize_, optim, objective, P);
    return;
  }
  // the new points in the existing normalisation (KrigingImpl.cpp:580-596)
  std::vector<double> Xn((size_t)N * d), yn(N);
  for (int k = 0; k < d; ++k) {
    for (int i = 0; i < n; ++i) Xn[i + (size_t)k * N] = X_[i + (size_t)k * n];
    for (int i = 0; i < m; ++i) Xn[n + i + (size_t)k * N] = (X_u[i + (size_t)k * m] - centerX_[k]) / scaleX_[k];
  }
  for (int i = 0; i < n; ++i) yn[i] = y_[i];
  for (int i = 0; i < m; ++i) yn[n + i] = (y_u[i] - centerY_) / scaleY_;
  X_.swap(Xn);
  y_.swap(yn);
  n_ = N;
  upload_data();
  if (refit && optim_ != "none") {
    // warm restart: one L-BFGS-B run from the current theta within the
    // bounds of the extended design, no restarts (Kriging.cpp:2469-2633)
    const OptimConfig& oc = optim_config();
    auto st = std::make_unique<FitState>();
    st->m = this;
    st->d = d;
    st->gd = d;
    st->bfgs = true;
    st->no_restart = true;
    st->obj = objective_ == "LOO" ? OBJ_LOO : (objective_ == "LMP" ? OBJ_LMP : OBJ_LL);
    st->n = N;
    st->rp = Reparam{d, false, oc.reparametrize};
    theta_bounds(st->th_lo, st->th_up);
    std::vector<double> glo(d), gup(d);
    st->rp.to_gamma(st->th_lo.data(), 1.0, glo.data());
    st->rp.to_gamma(st->th_up.data(), 1.0, gup.data());
    st->W.resize(1);
    st->th_alpha.assign(1, std::vector<double>(d));
    StartWorker& w = st->W[0];
    w.theta_start = theta_;
    w.glo = glo;
    w.gup = gup;
    st->restart_point(w, theta_);
    w.best_gamma = w.gamma;
    st->flags = flags_for_eval();
    est_theta_ = true;
    run_fit_jobs({st.get()});
    return;
  }
  // no re-fit: the model at the current parameters on the extended design,
  // sigma2 = SSE / n when estimated (KrigingImpl.cpp:576-632).  With new
  // points, the factor is the block update of the committed one
  // (Kriging.cpp:172-187 update_eligible -> LinearAlgebra::update_cholCov,
  // LinearAlgebra.cpp:206-247): R on the extended design (factor * k off the
  // diagonal, 1 on it), L = chol_block(R, T_old); the model follows from L
  // in the factor mode of the pipeline
  std::vector<double> g(theta_);
  if (noise_ == NoiseModel::Nugget) g.push_back(alpha_);
  std::vector<EvalRequest> reqs{commit_request(g)};
  reqs[0].flags.alpha = alpha_;
  std::vector<EvalResult> res;
  double *dR = nullptr, *dLn = nullptr;
  struct Free {
    double** p[2];
    ~Free() {
      for (double** q : p)
        if (*q) hipFree(*q);
    }
  } fr{{&dR, &dLn}};
  if (m > 0 && d_L_ != nullptr && d_Linv_ != nullptr) {
    hip_check(hipSetDevice(device_), "hipSetDevice");
    const size_t bn = sizeof(double) * (size_t)N * N;
    hip_check(hipMalloc(&dR, bn), "hipMalloc");
    hip_check(hipMalloc(&dLn, bn), "hipMalloc");
    const DeviceProblem& P = rt_->problem(problem_);
    CovMatArgs c{};
    c.X1 = P.X;
    c.X2 = P.X;
    c.out = dR;
    for (int k = 0; k < d; ++k) c.theta[k] = theta_[k];
    c.scale = (noise_ == NoiseModel::Nugget) ? alpha_ : 1.0;
    c.n1 = c.n2 = N;
    c.d = d;
    c.kernel = kernel_id(kernel_);
    c.unit_diag = 1;
    hip_check(launch_cov_mat(c, rt_->stream()), "R");
    rt_->chol_block_device(dR, N, d_L_, n, d_Linv_, npad_, n, dLn);
    rt_->set_problem_factor(problem_, dLn);
    reqs[0].mode = MODE_FACTOR;
    try {
      rt_->evaluate(reqs, false, res);
    } catch (...) {
      rt_->set_problem_factor(problem_, nullptr);
      throw;
    }
    rt_->set_problem_factor(problem_, nullptr);
  } else {
    rt_->evaluate(reqs, false, res);
  }
  const double s2 = sigma2_, nug = nugget_;
  const bool est_theta = est_theta_;
  commit_apply(theta_, alpha_, 0, res[0]);
  est_theta_ = est_theta;
  sigma2_ = est_sigma2_ ? res[0].sse / N : s2;
  nugget_ = nug;
}

// ---------------------------------------------------------------------------
// predict (Kriging::predict, Kriging.cpp:2241-2285 -> predict_impl,
// KrigingImpl.cpp:145-243): mean and standard deviation at m points
// ---------------------------------------------------------------------------
void Kriging::predict(const double* Xn, int m, bool want_stdev, double* mean, double* stdev, bool want_cov,
                      double* cov, bool want_deriv, double* mean_deriv, double* stdev_deriv) {
  if (is_empty_) throw std::runtime_error("model is not fitted");
  const int d = d_, p = p_;
  if (m <= 0) return;
  hip_check(hipSetDevice(device_), "hipSetDevice");
  // points per chunk: the covariance needs every pair in one chunk (m x m,
  // 8 GiB at the 32768-point limit); the derivatives carry d extra
  // npad x chunk matrices (bounded to ~4 GiB)
  int chunk = std::min(m, 8192);
  if (want_cov) {
    if (m > 32768) throw std::invalid_argument("predict: return_cov is limited to 32768 points per call on this path");
    chunk = m;
  }
  if (want_deriv) {
    const double per_col = 16.0 * npad_ * d;
    const int cmax = std::max(64, (int)std::min<double>(8192.0, (4.0 * (1ull << 30)) / per_col) / 64 * 64);
    if (!want_cov) chunk = std::min(chunk, cmax);
  }
  const int mpad_max = (chunk + NB - 1) / NB * NB;
  double *dXn = nullptr, *dRon = nullptr, *dC = nullptr, *dout = nullptr, *dcoef = nullptr, *dEc = nullptr,
         *dCov = nullptr, *dDR = nullptr, *dDW = nullptr, *dDm = nullptr, *dDv = nullptr;
  auto cleanup = [&] {
    for (double* q : {dXn, dRon, dC, dout, dcoef, dEc, dCov, dDR, dDW, dDm, dDv})
      if (q) hipFree(q);
  };
  try {
    hip_check(hipMalloc(&dXn, sizeof(double) * (size_t)chunk * d), "hipMalloc");
    hip_check(hipMalloc(&dRon, sizeof(double) * (size_t)npad_ * mpad_max), "hipMalloc");
    hip_check(hipMalloc(&dC, sizeof(double) * (size_t)npad_ * mpad_max), "hipMalloc");
    hip_check(hipMalloc(&dout, sizeof(double) * 2 * (size_t)chunk), "hipMalloc");
    hip_check(hipMalloc(&dcoef, sizeof(double) * (size_t)(p + p * p + 1)), "hipMalloc");
    if ((want_cov || want_deriv) && p > 0)
      hip_check(hipMalloc(&dEc, sizeof(double) * (size_t)chunk * p), "hipMalloc");
    if (want_cov) hip_check(hipMalloc(&dCov, sizeof(double) * (size_t)m * m), "hipMalloc");
    if (want_deriv) {
      hip_check(hipMalloc(&dDR, sizeof(double) * (size_t)npad_ * mpad_max * d), "hipMalloc");
      hip_check(hipMalloc(&dDW, sizeof(double) * (size_t)npad_ * mpad_max * d), "hipMalloc");
      hip_check(hipMalloc(&dDm, sizeof(double) * (size_t)chunk * d), "hipMalloc");
      hip_check(hipMalloc(&dDv, sizeof(double) * (size_t)chunk * d), "hipMalloc");
    }
    std::vector<double> coef(p + p * p);
    for (int a = 0; a < p; ++a) coef[a] = beta_[a];
    for (int q = 0; q < p * p; ++q) coef[p + q] = Rstar_[q];
    if (p > 0) hip_check(hipMemcpy(dcoef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice), "upload");
    const bool nug = noise_ == NoiseModel::Nugget;
    const double var_scale = nug ? sigma2_ / alpha_ : sigma2_;
    const double vs = var_scale * scaleY_ * scaleY_;
    std::vector<double> xn((size_t)chunk * d), host_out(2 * (size_t)chunk), hdm, hdv;
    if (want_deriv) {
      hdm.resize((size_t)chunk * d);
      hdv.resize((size_t)chunk * d);
    }
    for (int s0 = 0; s0 < m; s0 += chunk) {
      const int mc = std::min(chunk, m - s0);
      for (int k = 0; k < d; ++k)
        for (int j = 0; j < mc; ++j) xn[j + (size_t)k * mc] = (Xn[s0 + j + (size_t)k * m] - centerX_[k]) / scaleX_[k];
      hip_check(hipMemcpy(dXn, xn.data(), sizeof(double) * (size_t)mc * d, hipMemcpyHostToDevice), "upload Xn");
      PredictArgs a{};
      a.X = rt_->problem(problem_).X;
      a.Linv = d_Linv_;
      a.MZ = d_vec_;
      a.coef = dcoef;
      a.Xn = dXn;
      a.Ron = dRon;
      a.C = dC;
      a.out = dout;
      a.ecirc = dEc;
      a.cov = dCov;
      a.nn_factor = nug ? alpha_ : 1.0;
      a.DR = dDR;
      a.DW = dDW;
      a.dmean = dDm;
      a.dvar = dDv;
      for (int k = 0; k < d; ++k) a.theta[k] = theta_[k];
      a.factor = nug ? alpha_ : 1.0;
      a.n = n_;
      a.d = d;
      a.npad = npad_;
      a.p = p;
      a.m = mc;
      a.mpad = (mc + NB - 1) / NB * NB;
      a.kernel = kernel_id(kernel_);
      a.linear_trend = regmodel_ == RegModel::Linear;
      a.want_var = want_stdev ? 1 : 0;
      hip_check(launch_predict(a, rt_->stream()), "predict");
      if (want_cov) hip_check(launch_predict_cov(a, rt_->stream()), "predict cov");
      if (want_deriv) hip_check(launch_predict_deriv(a, rt_->stream()), "predict deriv");
      hip_check(hipMemcpyAsync(host_out.data(), dout, sizeof(double) * 2 * mc, hipMemcpyDeviceToHost, rt_->stream()),
                "download");
      if (want_deriv) {
        hip_check(hipMemcpyAsync(hdm.data(), dDm, sizeof(double) * (size_t)mc * d, hipMemcpyDeviceToHost,
                                 rt_->stream()),
                  "download");
        if (want_stdev)
          hip_check(hipMemcpyAsync(hdv.data(), dDv, sizeof(double) * (size_t)mc * d, hipMemcpyDeviceToHost,
                                   rt_->stream()),
                    "download");
      }
      if (want_cov && cov != nullptr)
        hip_check(hipMemcpyAsync(cov, dCov, sizeof(double) * (size_t)m * m, hipMemcpyDeviceToHost, rt_->stream()),
                  "download");
      hip_check(hipStreamSynchronize(rt_->stream()), "sync");
      for (int j = 0; j < mc; ++j) {
        mean[s0 + j] = centerY_ + scaleY_ * host_out[j];
        if (want_stdev && stdev != nullptr) stdev[s0 + j] = std::sqrt(host_out[mc + j] * vs);
      }
      if (want_deriv) {
        // chain rule of the normalisation (KrigingImpl.cpp:289-293), then
        // d stdev = d var / (2 stdev)
        for (int k = 0; k < d; ++k)
          for (int j = 0; j < mc; ++j) {
            if (mean_deriv != nullptr) mean_deriv[s0 + j + (size_t)k * m] = hdm[j + (size_t)k * mc] * scaleY_ / scaleX_[k];
            if (stdev_deriv != nullptr) {
              const double sd = std::sqrt(host_out[mc + j] * vs);
              const double dv = want_stdev ? hdv[j + (size_t)k * mc] * vs / scaleX_[k] : 0.0;
              stdev_deriv[s0 + j + (size_t)k * m] = dv / (2.0 * (want_stdev ? sd : 0.0));
            }
          }
      }
    }
    if (want_cov && cov != nullptr)
      for (size_t q = 0; q < (size_t)m * m; ++q) cov[q] *= vs;
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

// Evaluate the model at (theta, alpha), keep L^-1, z, M, R* on the device
// for predict, and set sigma2 / nugget / beta as the reference's commit
// (Kriging.cpp:2156-2202 and the optim="none" branch 1681-1700).
EvalRequest Kriging::commit_request(const std::vector<double>& gamma) const {
  EvalRequest R;
  R.problem = problem_;
  R.gamma = gamma.data();
  R.gamma_n = (int)gamma.size();
  R.flags = flags_for_eval();
  return R;
}

void Kriging::commit(const std::vector<double>& theta, double alpha) {
  std::vector<double> g(theta);
  if (noise_ == NoiseModel::Nugget) g.push_back(alpha);
  std::vector<EvalRequest> reqs{commit_request(g)};
  reqs[0].flags.alpha = alpha;
  std::vector<EvalResult> res;
  rt_->evaluate(reqs, false, res);
  commit_apply(theta, alpha, 0, res[0]);
}

// a model buffer of `bytes`, reallocated only when its size changes
static void ensure_model_buffer(double*& ptr, size_t& cap, size_t bytes) {
  if (ptr != nullptr && cap == bytes) return;
  if (ptr != nullptr) hipFree(ptr);
  ptr = nullptr;
  cap = 0;
  hip_check(hipMalloc(&ptr, bytes), "hipMalloc model");
  cap = bytes;
}

void Kriging::commit_apply(const std::vector<double>& theta, double alpha, int slot, const EvalResult& r) {
  commit_copy(slot, r);
  hip_check(hipStreamSynchronize(rt_->stream()), "sync");
  commit_finish(theta, alpha, slot, r);
}

// the device copies of slot `slot`'s factors into the model's buffers,
// enqueued on the runtime's stream (no synchronisation)
void Kriging::commit_copy(int slot, const EvalResult& r) {
  const int p = p_;
  if (r.status != EVAL_OK) throw std::runtime_error(chol_failure_message(r.status, r.inc, rt_->chol.num_nugget));
  const Runtime::SlotView v = rt_->slot_view(slot);
  npad_ = v.npad;
  hip_check(hipSetDevice(device_), "hipSetDevice");
  ensure_model_buffer(d_L_, cap_L_, sizeof(double) * (size_t)n_ * n_);
  ensure_model_buffer(d_Linv_, cap_Linv_, sizeof(double) * (size_t)npad_ * npad_);
  ensure_model_buffer(d_vec_, cap_vec_, sizeof(double) * (size_t)npad_ * (p + 1));
  hip_check(hipMemcpy2DAsync(d_L_, sizeof(double) * n_, v.A, sizeof(double) * npad_, sizeof(double) * n_, n_,
                             hipMemcpyDeviceToDevice, rt_->stream()),
            "copy model");
  hip_check(hipMemcpyAsync(d_Linv_, v.W, sizeof(double) * (size_t)npad_ * npad_, hipMemcpyDeviceToDevice,
                           rt_->stream()),
            "copy model");
  if (p > 0)
    hip_check(hipMemcpyAsync(d_vec_, v.V, sizeof(double) * (size_t)npad_ * p, hipMemcpyDeviceToDevice, rt_->stream()),
              "copy model");
  if (est_beta_)
    hip_check(hipMemcpyAsync(d_vec_ + (size_t)npad_ * p, v.V + (size_t)(p + 2) * npad_, sizeof(double) * npad_,
                             hipMemcpyDeviceToDevice, rt_->stream()),
              "copy model");
}

// the host state of the commit, once commit_copy's copies are done
void Kriging::commit_finish(const std::vector<double>& theta, double alpha, int slot, const EvalResult& r) {
  const int p = p_;
  theta_ = theta;
  const double sse = r.sse;
  std::vector<double> bhat(r.beta, r.beta + p);
  if (est_beta_) {
    beta_ = bhat;
  } else {
    // z = y* - M beta (Kriging.cpp:2170-2172)
    std::vector<double> ystar(n_), M((size_t)n_ * std::max(p, 1)), z(npad_, 0.0);
    rt_->read_slot(slot, 3, ystar.data());
    if (p > 0) rt_->read_slot(slot, 2, M.data());
    for (int i = 0; i < n_; ++i) {
      double s = ystar[i];
      for (int a = 0; a < p; ++a) s -= M[i + (size_t)a * n_] * beta_[a];
      z[i] = s;
    }
    hip_check(hipMemcpy(d_vec_ + (size_t)npad_ * p, z.data(), sizeof(double) * npad_, hipMemcpyHostToDevice),
              "upload z");
  }
  Rstar_.assign((size_t)p * p, 0.0);
  if (p > 0) rt_->read_slot(slot, 6, Rstar_.data());
  const int n = n_;
  const double total_var = sse / n;
  if (noise_ == NoiseModel::Nugget) {
    if (!est_theta_) {
      // optim = "none" (Kriging.cpp:1688-1696); nugget_ holds the given value
      if (est_sigma2_) {
        sigma2_ = alpha * total_var;
        nugget_ = est_nugget_ ? (1.0 - alpha) * total_var : nugget_;
      } else {
        nugget_ = est_nugget_ ? 0.0 : nugget_;
      }
    } else if (est_sigma2_) {
      // BFGS commit (Kriging.cpp:2174-2189)
      if (est_nugget_) {
        sigma2_ = alpha * total_var;
        if (objective_ == "LMP") sigma2_ = sigma2_ * n / (n - p - 2);
        nugget_ = sigma2_ / alpha - sigma2_;
      } else {
        sigma2_ = nugget_ * alpha / (1.0 - alpha);
      }
    } else if (est_nugget_) {
      nugget_ = sigma2_ * (1.0 - alpha) / alpha;
    }
  } else if (est_sigma2_) {
    sigma2_ = total_var;
    // BFGS commit with the LMP objective (Kriging.cpp:2192-2195)
    if (est_theta_ && objective_ == "LMP") sigma2_ = sse / (n - p);
  }
  is_empty_ = false;
}

}  // namespace lk

namespace lk {

// KrigingImpl::summary_top / summary_bottom (KrigingImpl.cpp:713-762) and
// Kriging::summary (Kriging.cpp:2660-2670)
std::string Kriging::summary() const {
  std::ostringstream oss;
  auto vec = [&oss](const std::vector<double>& v) {
    for (size_t i = 0; i < v.size(); ++i) oss << (i ? ", " : "") << v[i];
  };
  if (is_empty_ || n_ == 0) {
    oss << "* covariance:\n";
    oss << "  * kernel: " << kernel_ << "\n";
    return oss.str();
  }
  oss << "* data" << (normalize_ ? " (normalized): " : ": ") << n_ << "x";
  for (int k = 0; k < d_; ++k) {
    // the displayed design is the original one
    double mn = X_[(size_t)k * n_], mx = mn;
    for (int i = 1; i < n_; ++i) {
      mn = std::min(mn, X_[i + (size_t)k * n_]);
      mx = std::max(mx, X_[i + (size_t)k * n_]);
    }
    oss << "[" << mn * scaleX_[k] + centerX_[k] << "," << mx * scaleX_[k] + centerX_[k] << "]" << (k + 1 < d_ ? "," : "");
  }
  const double ymn = *std::min_element(y_.begin(), y_.end()), ymx = *std::max_element(y_.begin(), y_.end());
  oss << " -> " << n_ << "x[" << ymn * scaleY_ + centerY_ << "," << ymx * scaleY_ + centerY_ << "]\n";
  oss << "* trend " << regmodel_name(regmodel_) << (est_beta_ ? " (est.): " : ": ");
  vec(beta_);
  oss << "\n* variance" << (est_sigma2_ ? " (est.): " : ": ") << sigma2_ << "\n";
  oss << "* covariance:\n";
  oss << "  * kernel: " << kernel_ << "\n";
  oss << "  * range" << (est_theta_ ? " (est.): " : ": ");
  vec(theta_);
  oss << "\n";
  if (noise_ == NoiseModel::Nugget) oss << "  * nugget" << (est_nugget_ ? " (est.): " : ": ") << nugget_ << "\n";
  oss << "  * fit:\n";
  oss << "    * objective: " << objective_ << "\n";
  oss << "    * optim: " << optim_ << "\n";
  return oss.str();
}

void Kriging::cov_mat(const double* X1, int n1, const double* X2, int n2, double* out) const {
  if (is_empty_) throw std::runtime_error("model is not fitted");
  if (n1 <= 0 || n2 <= 0) return;
  const int d = d_;
  // normalised inputs (KrigingImpl.cpp:61-66)
  std::vector<double> a((size_t)n1 * d), b((size_t)n2 * d);
  for (int k = 0; k < d; ++k) {
    for (int i = 0; i < n1; ++i) a[i + (size_t)k * n1] = (X1[i + (size_t)k * n1] - centerX_[k]) / scaleX_[k];
    for (int j = 0; j < n2; ++j) b[j + (size_t)k * n2] = (X2[j + (size_t)k * n2] - centerX_[k]) / scaleX_[k];
  }
  hip_check(hipSetDevice(device_), "hipSetDevice");
  double *dA = nullptr, *dB = nullptr, *dO = nullptr;
  auto cleanup = [&] {
    if (dA) hipFree(dA);
    if (dB) hipFree(dB);
    if (dO) hipFree(dO);
  };
  try {
    hip_check(hipMalloc(&dA, sizeof(double) * a.size()), "hipMalloc");
    hip_check(hipMalloc(&dB, sizeof(double) * b.size()), "hipMalloc");
    hip_check(hipMalloc(&dO, sizeof(double) * (size_t)n1 * n2), "hipMalloc");
    hip_check(hipMemcpy(dA, a.data(), sizeof(double) * a.size(), hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(dB, b.data(), sizeof(double) * b.size(), hipMemcpyHostToDevice), "upload");
    CovMatArgs c{};
    c.X1 = dA;
    c.X2 = dB;
    c.out = dO;
    for (int k = 0; k < d; ++k) c.theta[k] = theta_[k];
    c.scale = sigma2_;
    c.n1 = n1;
    c.n2 = n2;
    c.d = d;
    c.kernel = kernel_id(kernel_);
    hip_check(launch_cov_mat(c, rt_->stream()), "cov_mat");
    hip_check(hipStreamSynchronize(rt_->stream()), "sync");
    hip_check(hipMemcpy(out, dO, sizeof(double) * (size_t)n1 * n2, hipMemcpyDeviceToHost), "download");
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace lk
