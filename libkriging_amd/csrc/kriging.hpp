// Kriging model (ordinary / universal kriging, optional nugget) on the GPU
// likelihood pipeline.  Mirrors the reference class Kriging
// (src/lib/include/libKriging/Kriging.hpp:18-366): constructor, fit (BFGS
// multistart with restarts, or "none"), logLikelihoodFun, logLikelihood,
// predict and the getters, with the same argument meaning.  All the
// O(n^2)..O(n^3) work runs through lk::Runtime; the host keeps the
// optimiser state machines and the O(d) parameter transforms.
#pragma once
#include <cstdint>
#include <memory>
#include <optional>
#include <random>
#include <string>
#include <vector>

#include "lk_runtime.hpp"

namespace lk {

enum class NoiseModel { None = 0, Nugget = 1 };
enum class RegModel { None = 0, Constant = 1, Linear = 2 };

RegModel regmodel_from_string(const std::string& s);
// Trend::toString (Trend.cpp:29-32): "none", "constant", "linear"
const char* regmodel_name(RegModel r);
int kernel_id(const std::string& name);

// KrigingParameters (Kriging.hpp:18-28); theta is rows x d, row-major
struct Parameters {
  std::optional<double> sigma2;
  bool is_sigma2_estim = true;
  std::optional<std::vector<double>> theta;
  int theta_rows = 0;
  bool is_theta_estim = true;
  std::optional<std::vector<double>> beta;
  bool is_beta_estim = true;
  std::optional<double> nugget;
  bool is_nugget_estim = true;
};

// Process-global optimisation knobs (Optim statics, src/lib/Optim.cpp:16-141)
struct OptimConfig {
  bool reparametrize = true;
  double theta_lower_factor = 0.02;
  double theta_upper_factor = 10.0;
  bool variogram_bounds_heuristic = true;
  int max_restart = 10;
  int max_iteration = 20;
  double gradient_tolerance = 1e-3;
  double objective_rel_tolerance = 1e-3;
  int lbfgs_memory = 5;
  int log_level = 0;
};
OptimConfig& optim_config();

// Random (src/lib/Random.cpp): global mt19937 reseeded by every BFGS fit
void random_reset_seed(unsigned seed);

// Per-process runtime of one device (created on first use)
Runtime& runtime_for(int device);
int default_device();

// The optimisation part of one model's fit, run step by step: the pending
// evaluations of every job handed to Kriging::run_fit_jobs go to the device
// as one batched pipeline per optimiser step (several models' starts, e.g.
// the groups of a NestedKriging, side by side).
class FitJob {
 public:
  virtual ~FitJob() = default;
};

class Kriging {
 public:
  Kriging(const std::string& kernel, NoiseModel noise = NoiseModel::None);
  ~Kriging();
  Kriging(const Kriging&) = delete;
  Kriging& operator=(const Kriging&) = delete;

  // X: n x d column-major
  void fit(const double* y, const double* X, int n, int d, RegModel regmodel = RegModel::Constant,
           bool normalize = false, const std::string& optim = "BFGS", const std::string& objective = "LL",
           const Parameters& parameters = Parameters{});
  // fit() in two halves: set up data and start points, then optimise and
  // commit every job together (all models on the same device)
  std::unique_ptr<FitJob> begin_fit(const double* y, const double* X, int n, int d, RegModel regmodel,
                                    bool normalize, const std::string& optim, const std::string& objective,
                                    const Parameters& parameters);
  static void run_fit_jobs(const std::vector<FitJob*>& jobs);

  // gamma = theta (d) or [theta, alpha] (d + 1, nugget model); returns LL
  double logLikelihoodFun(const double* gamma, int gamma_n, bool want_grad, double* grad);
  // nb evaluations at once (rows of gamma, each gamma_n long), one batched pipeline
  void logLikelihoodFunBatch(int nb, const double* gamma, int gamma_n, bool want_grad, double* ll, double* grad);
  double logLikelihood();
  // leave-one-out objective (Kriging::leaveOneOutFun / leaveOneOut /
  // leaveOneOutVec, Kriging.cpp:470-485, 669-671); theta has d entries,
  // noise model none only (as the reference's LOO fit objective)
  double leaveOneOutFun(const double* theta, int theta_n, bool want_grad, double* grad);
  void leaveOneOutVec(const double* theta, int theta_n, double* yhat, double* stderr_out);
  double leaveOneOut();
  // Kriging::update (Kriging.cpp:2433-2633): append m points (X_u m x d
  // column-major); refit = warm-started re-optimisation (a full re-fit for
  // the nugget model), else the model at the current parameters
  void update(const double* y_u, const double* X_u, int m, bool refit);
  // model files in the reference's JSON version-2 format (Kriging::save /
  // Kriging::load, Kriging.cpp:2691-2750); load recomputes the factors on
  // the device at the saved parameters
  void save(const std::string& filename);
  static Kriging* load(const std::string& filename);
  std::string to_json() const;
  static Kriging* from_json(const std::string& text, const std::string& filename);
  // an independent model with this one's state (Kriging.cpp:164)
  Kriging* copy() const;
  // T = L, M = F*, z and R^-1 at the committed parameters (any may be null)
  void factors(std::vector<double>* T, std::vector<double>* M, std::vector<double>* z,
               std::vector<double>* Rinv) const;
  // Kriging::summary (Kriging.cpp:2660-2670, KrigingImpl.cpp:713-762)
  std::string summary() const;
  // KrigingImpl::covMat (KrigingImpl.cpp:60-71): sigma2 k(x1_i - x2_j) of
  // the normalised inputs, n1 x n2 (X1 n1 x d, X2 n2 x d column-major)
  void cov_mat(const double* X1, int n1, const double* X2, int n2, double* out) const;
  // log marginal posterior (Kriging::logMargPostFun / logMargPost,
  // Kriging.cpp:488-672); gamma as logLikelihoodFun
  double logMargPostFun(const double* gamma, int gamma_n, bool want_grad, double* grad);
  double logMargPost();

  // Xn: m x d column-major; mean[m], stdev[m] (stdev may be null).
  // cov: m x m; mean_deriv, stdev_deriv: m x d (column-major), when asked
  // (KrigingImpl::predict_impl, KrigingImpl.cpp:145-300)
  void predict(const double* Xn, int m, bool want_stdev, double* mean, double* stdev, bool want_cov = false,
               double* cov = nullptr, bool want_deriv = false, double* mean_deriv = nullptr,
               double* stdev_deriv = nullptr);

  // getters (normalised space, as the reference's members)
  const std::string& kernel() const { return kernel_; }
  NoiseModel noise() const { return noise_; }
  const std::vector<double>& theta() const { return theta_; }
  double sigma2() const { return sigma2_; }
  double nugget() const { return nugget_; }
  double alpha() const { return alpha_; }
  const std::vector<double>& beta() const { return beta_; }
  int n() const { return n_; }
  int d() const { return d_; }
  int p() const { return p_; }
  bool is_fitted() const { return !is_empty_; }
  const std::string& optim() const { return optim_; }
  const std::string& objective() const { return objective_; }
  double centerY() const { return centerY_; }
  double scaleY() const { return scaleY_; }
  const std::vector<double>& centerX() const { return centerX_; }
  const std::vector<double>& scaleX() const { return scaleX_; }
  bool normalize() const { return normalize_; }
  RegModel regmodel() const { return regmodel_; }
  const std::vector<double>& X() const { return X_; }
  const std::vector<double>& y() const { return y_; }
  const std::vector<double>& F() const { return F_; }
  bool is_beta_estim() const { return est_beta_; }
  bool is_theta_estim() const { return est_theta_; }
  bool is_sigma2_estim() const { return est_sigma2_; }
  bool is_nugget_estim() const { return est_nugget_; }
  // statistics of the last fit
  int fit_evaluations() const { return fit_evals_; }
  int fit_batches() const { return fit_batches_; }
  // force the plain one-evaluation-per-launch path in fit (comparison only)
  void set_batched_fit(bool b) { batched_fit_ = b; }
  // the uploaded design (device problem id) and the device runtime
  int problem_id() const { return problem_; }
  Runtime* runtime() const { return rt_; }

 private:
  friend struct FitState;
  void setup(const double* y, const double* X, int n, int d, RegModel regmodel, bool normalize);
  void upload_data();
  ModelFlags flags_for_eval() const;
  // commit in two halves so that many models commit in one batch: the
  // request, then the copy of slot `slot`'s factors once it is evaluated
  EvalRequest commit_request(const std::vector<double>& gamma) const;
  void commit_apply(const std::vector<double>& theta, double alpha, int slot, const EvalResult& r);
  // commit_apply in two halves, so that many models' copies go out before
  // one synchronisation: the device copies (enqueued), then the host state
  void commit_copy(int slot, const EvalResult& r);
  void commit_finish(const std::vector<double>& theta, double alpha, int slot, const EvalResult& r);
  void commit(const std::vector<double>& theta, double alpha);
  void theta_bounds(std::vector<double>& lo, std::vector<double>& up);

 public:
  // Optim::theta_bounds (Optim.cpp:179-209) of a design X (n x d,
  // column-major) and response y: [lower_factor, upper_factor] x the column
  // ranges, the lower bound raised by the variogram heuristic (an O(n^2 d)
  // pair loop, the order of the reference's)
  static void theta_bounds_of(const double* X, const double* y, int n, int d, std::vector<double>& lo,
                              std::vector<double>& up);
  // bounds the next begin_fit uses instead of computing them (a caller that
  // computed them for many models at once, in parallel); they must be those
  // of the data that begin_fit receives, without normalisation
  void use_theta_bounds(std::vector<double> lo, std::vector<double> up) {
    pre_lo_ = std::move(lo);
    pre_up_ = std::move(up);
  }

 private:
  std::vector<double> pre_lo_, pre_up_;
  std::string kernel_;
  NoiseModel noise_;
  int device_;
  Runtime* rt_ = nullptr;
  int problem_ = -1;

  // data (normalised), column-major
  int n_ = 0, d_ = 0, p_ = 0;
  std::vector<double> X_, y_, F_;
  RegModel regmodel_ = RegModel::Constant;
  bool normalize_ = false;
  std::vector<double> centerX_, scaleX_;
  double centerY_ = 0.0, scaleY_ = 1.0;
  std::string optim_ = "BFGS", objective_ = "LL";

  // fitted state
  bool is_empty_ = true;
  std::vector<double> theta_;
  double sigma2_ = 1.0, nugget_ = 0.0, alpha_ = 1.0;
  bool est_sigma2_ = true, est_nugget_ = true, est_beta_ = true, est_theta_ = true;
  std::vector<double> beta_;

  // committed model on the device: L (n x n, ld = n; only the lower
  // triangle is meaningful), L^-1 (npad x npad),
  // z = E*, M = F*, R*
  double* d_L_ = nullptr;
  double* d_Linv_ = nullptr;
  double* d_vec_ = nullptr;  // npad x (p + 1): [M | z]
  size_t cap_L_ = 0, cap_Linv_ = 0, cap_vec_ = 0;  // their sizes in bytes (reused when unchanged)
  int npad_ = 0;
  std::vector<double> Rstar_;  // p x p upper

  int fit_evals_ = 0, fit_batches_ = 0;
  bool batched_fit_ = true;
};

}  // namespace lk
