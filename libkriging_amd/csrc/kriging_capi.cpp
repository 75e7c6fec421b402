// extern "C" Kriging object API, mirroring the reference's own C ABI
// (bindings/Julia/jlibkriging/csrc/libkriging_c.h): same names, argument
// meaning and the 0 / -1 + lk_get_last_error() convention.
#include <cstring>
#include <exception>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lkgpu.h"
#include "kriging.hpp"
#include "lbfgsb.hpp"

namespace lk {
void set_last_error(const std::string& s);
}

namespace {
template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    lk::set_last_error(e.what());
  } catch (...) {
    lk::set_last_error("unknown error");
  }
  return -1;
}

lk::Kriging* K(void* p) {
  if (p == nullptr) throw std::invalid_argument("Kriging handle is NULL");
  return static_cast<lk::Kriging*>(p);
}

lk::NoiseModel parse_noise(const char* s) {
  if (s == nullptr || std::strcmp(s, "none") == 0 || s[0] == '\0') return lk::NoiseModel::None;
  if (std::strcmp(s, "nugget") == 0) return lk::NoiseModel::Nugget;
  throw std::invalid_argument(std::string("Unsupported noise model on this path: ") + s + " (none, nugget)");
}

lk::Parameters make_params(int d, const double* sigma2, int is_sigma2_estim, const double* theta, int theta_n,
                           int is_theta_estim, const double* beta, int beta_n, int is_beta_estim,
                           const double* nugget, int is_nugget_estim) {
  lk::Parameters P;
  if (sigma2) P.sigma2 = *sigma2;
  P.is_sigma2_estim = is_sigma2_estim != 0;
  if (theta && theta_n > 0) {
    if (theta_n % d != 0)
      throw std::runtime_error("Dimension of theta should be nx" + std::to_string(d) + " (got " +
                               std::to_string(theta_n) + " values)");
    P.theta = std::vector<double>(theta, theta + theta_n);  // rows x d, row-major
    P.theta_rows = theta_n / d;
  }
  P.is_theta_estim = is_theta_estim != 0;
  if (beta && beta_n > 0) P.beta = std::vector<double>(beta, beta + beta_n);
  P.is_beta_estim = is_beta_estim != 0;
  if (nugget) P.nugget = *nugget;
  P.is_nugget_estim = is_nugget_estim != 0;
  return P;
}

thread_local std::string g_str;
const char* ret_str(const std::string& s) {
  g_str = s;
  return g_str.c_str();
}

int copy_vec(const std::vector<double>& v, double* out, int* n) {
  if (n) *n = (int)v.size();
  if (out) std::memcpy(out, v.data(), sizeof(double) * v.size());
  return 0;
}
}  // namespace

extern "C" {

void* lk_kriging_new(const char* kernel, const char* noise_model) {
  void* out = nullptr;
  guarded([&] { out = new lk::Kriging(kernel ? kernel : "", parse_noise(noise_model)); });
  return out;
}

void* lk_kriging_new_fit(const double* y, int n, const double* noise, int noise_n, const double* X, int nX, int d,
                         const char* kernel, const char* noise_model, const char* regmodel, int normalize,
                         const char* optim, const char* objective, const double* sigma2, int is_sigma2_estim,
                         const double* theta, int theta_n, int is_theta_estim, const double* beta, int beta_n,
                         int is_beta_estim, const double* nugget, int is_nugget_estim) {
  void* out = nullptr;
  guarded([&] {
    if (noise != nullptr && noise_n > 0)
      throw std::invalid_argument("heterogeneous noise is not supported on this path");
    if (nX != n)
      throw std::runtime_error("Dimension of new data should be the same:\n X: (" + std::to_string(nX) + "x" +
                               std::to_string(d) + "), y: (" + std::to_string(n) + ")");
    auto k = std::make_unique<lk::Kriging>(kernel ? kernel : "", parse_noise(noise_model));
    lk::Parameters P = make_params(d, sigma2, is_sigma2_estim, theta, theta_n, is_theta_estim, beta, beta_n,
                                   is_beta_estim, nugget, is_nugget_estim);
    k->fit(y, X, n, d, lk::regmodel_from_string(regmodel ? regmodel : "constant"), normalize != 0,
           optim ? optim : "BFGS", objective ? objective : "LL", P);
    out = k.release();
  });
  return out;
}

void lk_kriging_delete(void* ptr) { delete static_cast<lk::Kriging*>(ptr); }

int lk_kriging_fit(void* ptr, const double* y, int n, const double* noise, int noise_n, const double* X, int nX, int d,
                   const char* regmodel, int normalize, const char* optim, const char* objective) {
  return guarded([&] {
    if (noise != nullptr && noise_n > 0)
      throw std::invalid_argument("heterogeneous noise is not supported on this path");
    if (nX != n) throw std::runtime_error("Dimension of new data should be the same");
    K(ptr)->fit(y, X, n, d, lk::regmodel_from_string(regmodel ? regmodel : "constant"), normalize != 0,
                optim ? optim : "BFGS", objective ? objective : "LL", lk::Parameters{});
  });
}

int lk_kriging_fit_parameters(void* ptr, const double* y, int n, const double* X, int nX, int d, const char* regmodel,
                              int normalize, const char* optim, const char* objective, const double* sigma2,
                              int is_sigma2_estim, const double* theta, int theta_n, int is_theta_estim,
                              const double* beta, int beta_n, int is_beta_estim, const double* nugget,
                              int is_nugget_estim) {
  return guarded([&] {
    if (nX != n) throw std::runtime_error("Dimension of new data should be the same");
    lk::Parameters P = make_params(d, sigma2, is_sigma2_estim, theta, theta_n, is_theta_estim, beta, beta_n,
                                   is_beta_estim, nugget, is_nugget_estim);
    K(ptr)->fit(y, X, n, d, lk::regmodel_from_string(regmodel ? regmodel : "constant"), normalize != 0,
                optim ? optim : "BFGS", objective ? objective : "LL", P);
  });
}

int lk_kriging_predict(void* ptr, const double* X_n, int m, int d, int return_stdev, int return_cov, int return_deriv,
                       double* mean_out, double* stdev_out, double* cov_out, double* mean_deriv_out,
                       double* stdev_deriv_out) {
  return guarded([&] {
    lk::Kriging* k = K(ptr);
    if (d != k->d())
      throw std::runtime_error("Predict locations have wrong dimension: " + std::to_string(d) + " instead of " +
                               std::to_string(k->d()));
    k->predict(X_n, m, return_stdev != 0, mean_out, return_stdev ? stdev_out : nullptr, return_cov != 0,
               return_cov ? cov_out : nullptr, return_deriv != 0, return_deriv ? mean_deriv_out : nullptr,
               return_deriv ? stdev_deriv_out : nullptr);
  });
}

int lk_kriging_log_likelihood_fun(void* ptr, const double* theta, int theta_n, int return_grad, int return_hess,
                                  double* ll_out, double* grad_out, double* hess_out) {
  return guarded([&] {
    if (return_hess) throw std::invalid_argument("the Hessian is not available on this path");
    (void)hess_out;
    if (ll_out == nullptr) throw std::invalid_argument("ll_out is NULL");
    if (return_grad && grad_out == nullptr) throw std::invalid_argument("grad_out is NULL");
    *ll_out = K(ptr)->logLikelihoodFun(theta, theta_n, return_grad != 0, grad_out);
  });
}

int lk_kriging_log_likelihood_fun_batch(void* ptr, int nb, const double* theta, int theta_n, int return_grad,
                                        double* ll_out, double* grad_out) {
  return guarded([&] {
    if (ll_out == nullptr || theta == nullptr) throw std::invalid_argument("NULL argument");
    if (return_grad && grad_out == nullptr) throw std::invalid_argument("grad_out is NULL");
    K(ptr)->logLikelihoodFunBatch(nb, theta, theta_n, return_grad != 0, ll_out, grad_out);
  });
}

int lk_kriging_leave_one_out_fun(void* ptr, const double* theta, int theta_n, int return_grad, double* loo_out,
                                 double* grad_out) {
  return guarded([&] {
    if (theta == nullptr) throw std::invalid_argument("theta is NULL");
    std::vector<double> g(theta_n);
    const double loo = K(ptr)->leaveOneOutFun(theta, theta_n, return_grad != 0, g.data());
    if (loo_out) *loo_out = loo;
    if (grad_out && return_grad) std::memcpy(grad_out, g.data(), sizeof(double) * theta_n);
  });
}

int lk_kriging_log_marg_post_fun(void* ptr, const double* theta, int theta_n, int return_grad, double* lmp_out,
                                 double* grad_out) {
  return guarded([&] {
    if (theta == nullptr) throw std::invalid_argument("theta is NULL");
    std::vector<double> g(theta_n + 1);
    const double v = K(ptr)->logMargPostFun(theta, theta_n, return_grad != 0, g.data());
    if (lmp_out) *lmp_out = v;
    if (grad_out && return_grad) std::memcpy(grad_out, g.data(), sizeof(double) * theta_n);
  });
}

double lk_kriging_log_marg_post(void* ptr) {
  double v = 0.0;
  if (guarded([&] { v = K(ptr)->logMargPost(); }) != 0) return std::numeric_limits<double>::quiet_NaN();
  return v;
}

double lk_kriging_leave_one_out(void* ptr) {
  double v = 0.0;
  if (guarded([&] { v = K(ptr)->leaveOneOut(); }) != 0) return std::numeric_limits<double>::quiet_NaN();
  return v;
}

int lk_kriging_leave_one_out_vec(void* ptr, const double* theta, int theta_n, double* yhat_out, double* stderr_out) {
  return guarded([&] {
    if (theta == nullptr) throw std::invalid_argument("theta is NULL");
    K(ptr)->leaveOneOutVec(theta, theta_n, yhat_out, stderr_out);
  });
}

int lk_kriging_update(void* ptr, const double* y_u, int n, const double* noise_u, int noise_n, const double* X_u,
                      int nX, int d, int refit) {
  return guarded([&] {
    (void)noise_u;
    if (noise_n > 0) throw std::invalid_argument("heterogeneous noise is not supported on this path");
    lk::Kriging* k = K(ptr);
    if (n != nX)
      throw std::runtime_error("Dimension of new data should be the same:\n X: (" + std::to_string(nX) + "x" +
                               std::to_string(d) + "), y: (" + std::to_string(n) + ")");
    if (d != k->d())
      throw std::runtime_error("Dimension of new data should be the same:\n X: (...x" + std::to_string(k->d()) +
                               "), new X: (...x" + std::to_string(d) + ")");
    k->update(y_u, X_u, n, refit != 0);
  });
}

int lk_kriging_save(void* ptr, const char* filename) {
  return guarded([&] {
    if (filename == nullptr) throw std::invalid_argument("filename is NULL");
    K(ptr)->save(filename);
  });
}

void* lk_kriging_load(const char* filename) {
  void* out = nullptr;
  guarded([&] {
    if (filename == nullptr) throw std::invalid_argument("filename is NULL");
    out = lk::Kriging::load(filename);
  });
  return out;
}

double lk_kriging_log_likelihood(void* ptr) {
  double v = 0.0;
  if (guarded([&] { v = K(ptr)->logLikelihood(); }) != 0) return std::numeric_limits<double>::quiet_NaN();
  return v;
}

const char* lk_kriging_kernel(void* ptr) { return ret_str(K(ptr)->kernel()); }
const char* lk_kriging_optim(void* ptr) { return ret_str(K(ptr)->optim()); }
const char* lk_kriging_objective(void* ptr) { return ret_str(K(ptr)->objective()); }
int lk_kriging_is_normalize(void* ptr) { return K(ptr)->normalize() ? 1 : 0; }
const char* lk_kriging_regmodel(void* ptr) {
  switch (K(ptr)->regmodel()) {
    case lk::RegModel::None:
      return "none";
    case lk::RegModel::Constant:
      return "constant";
    default:
      return "linear";
  }
}
const char* lk_kriging_noise_model(void* ptr) {
  return K(ptr)->noise() == lk::NoiseModel::Nugget ? "nugget" : "none";
}
int lk_kriging_get_X(void* ptr, double* out, int* n, int* d) {
  lk::Kriging* k = K(ptr);
  if (n) *n = k->n();
  if (d) *d = k->d();
  if (out) std::memcpy(out, k->X().data(), sizeof(double) * k->X().size());
  return 0;
}
int lk_kriging_get_centerX(void* ptr, double* out, int* d) { return copy_vec(K(ptr)->centerX(), out, d); }
int lk_kriging_get_scaleX(void* ptr, double* out, int* d) { return copy_vec(K(ptr)->scaleX(), out, d); }
int lk_kriging_get_y(void* ptr, double* out, int* n) { return copy_vec(K(ptr)->y(), out, n); }
double lk_kriging_get_centerY(void* ptr) { return K(ptr)->centerY(); }
double lk_kriging_get_scaleY(void* ptr) { return K(ptr)->scaleY(); }
int lk_kriging_get_beta(void* ptr, double* out, int* n) { return copy_vec(K(ptr)->beta(), out, n); }
int lk_kriging_get_theta(void* ptr, double* out, int* n) { return copy_vec(K(ptr)->theta(), out, n); }
double lk_kriging_get_sigma2(void* ptr) { return K(ptr)->sigma2(); }
double lk_kriging_get_nugget(void* ptr) { return K(ptr)->nugget(); }

void* lk_kriging_copy(void* ptr) {
  lk::Kriging* c = nullptr;
  guarded([&] { c = K(ptr)->copy(); });
  return c;
}

const char* lk_kriging_summary(void* ptr) {
  std::string s;
  if (guarded([&] { s = K(ptr)->summary(); }) != 0) return nullptr;
  return ret_str(s);
}

int lk_kriging_cov_mat(void* ptr, const double* X1, int n1, int d1, const double* X2, int n2, int d2, double* cov_out) {
  return guarded([&] {
    lk::Kriging* k = K(ptr);
    if (d1 != k->d() || d2 != k->d()) throw std::invalid_argument("X1 and X2 should have d columns");
    if (n1 < 0 || n2 < 0 || (n1 > 0 && X1 == nullptr) || (n2 > 0 && X2 == nullptr) || cov_out == nullptr)
      throw std::invalid_argument("NULL argument");
    k->cov_mat(X1, n1, X2, n2, cov_out);
  });
}

static int copy_mat(const std::vector<double>& v, int rows, int cols, double* out, int* n, int* d) {
  if (n) *n = rows;
  if (d) *d = cols;
  if (out) std::memcpy(out, v.data(), sizeof(double) * (size_t)rows * cols);
  return 0;
}

int lk_kriging_get_F(void* ptr, double* out, int* n, int* d) {
  return guarded([&] { copy_mat(K(ptr)->F(), K(ptr)->n(), K(ptr)->p(), out, n, d); });
}
int lk_kriging_get_T(void* ptr, double* out, int* n, int* d) {
  return guarded([&] {
    lk::Kriging* k = K(ptr);
    if (n) *n = k->n();
    if (d) *d = k->n();
    if (out == nullptr) return;
    std::vector<double> T;
    k->factors(&T, nullptr, nullptr, nullptr);
    copy_mat(T, k->n(), k->n(), out, n, d);
  });
}
int lk_kriging_get_M(void* ptr, double* out, int* n, int* d) {
  return guarded([&] {
    lk::Kriging* k = K(ptr);
    if (n) *n = k->n();
    if (d) *d = k->p();
    if (out == nullptr) return;
    std::vector<double> M;
    k->factors(nullptr, &M, nullptr, nullptr);
    copy_mat(M, k->n(), k->p(), out, n, d);
  });
}
int lk_kriging_get_z(void* ptr, double* out, int* n) {
  return guarded([&] {
    lk::Kriging* k = K(ptr);
    if (n) *n = k->n();
    if (out == nullptr) return;
    std::vector<double> z;
    k->factors(nullptr, nullptr, &z, nullptr);
    copy_vec(z, out, n);
  });
}
int lk_kriging_is_beta_estim(void* ptr) { return K(ptr)->is_beta_estim() ? 1 : 0; }
int lk_kriging_is_theta_estim(void* ptr) { return K(ptr)->is_theta_estim() ? 1 : 0; }
int lk_kriging_is_sigma2_estim(void* ptr) { return K(ptr)->is_sigma2_estim() ? 1 : 0; }
int lk_kriging_is_nugget_estim(void* ptr) { return K(ptr)->is_nugget_estim() ? 1 : 0; }
int lk_kriging_get_noise(void* ptr, double* out, int* n) {
  (void)out;
  return guarded([&] {
    K(ptr);
    if (n) *n = 0;
  });
}

int lk_kriging_fit_statistics(void* ptr, int* evaluations, int* batches) {
  return guarded([&] {
    lk::Kriging* k = K(ptr);
    if (evaluations) *evaluations = k->fit_evaluations();
    if (batches) *batches = k->fit_batches();
  });
}

int lk_kriging_set_batched_fit(void* ptr, int batched) {
  return guarded([&] { K(ptr)->set_batched_fit(batched != 0); });
}

// process-global knobs
int lk_optim_set(const char* name, double value) {
  return guarded([&] {
    lk::OptimConfig& c = lk::optim_config();
    const std::string s = name ? name : "";
    if (s == "reparametrize") c.reparametrize = value != 0;
    else if (s == "theta_lower_factor") c.theta_lower_factor = value;
    else if (s == "theta_upper_factor") c.theta_upper_factor = value;
    else if (s == "variogram_bounds_heuristic") c.variogram_bounds_heuristic = value != 0;
    else if (s == "log_level") c.log_level = (int)value;
    else if (s == "max_restart") c.max_restart = (int)value;
    else if (s == "max_iteration") c.max_iteration = (int)value;
    else if (s == "gradient_tolerance") c.gradient_tolerance = value;
    else if (s == "objective_rel_tolerance") c.objective_rel_tolerance = value;
    else if (s == "lbfgs_memory") c.lbfgs_memory = (int)value;
    else throw std::invalid_argument("unknown Optim setting: " + s);
  });
}

int lk_optim_get(const char* name, double* value) {
  return guarded([&] {
    const lk::OptimConfig& c = lk::optim_config();
    const std::string s = name ? name : "";
    if (value == nullptr) throw std::invalid_argument("value is NULL");
    if (s == "reparametrize") *value = c.reparametrize;
    else if (s == "theta_lower_factor") *value = c.theta_lower_factor;
    else if (s == "theta_upper_factor") *value = c.theta_upper_factor;
    else if (s == "variogram_bounds_heuristic") *value = c.variogram_bounds_heuristic;
    else if (s == "log_level") *value = c.log_level;
    else if (s == "max_restart") *value = c.max_restart;
    else if (s == "max_iteration") *value = c.max_iteration;
    else if (s == "gradient_tolerance") *value = c.gradient_tolerance;
    else if (s == "objective_rel_tolerance") *value = c.objective_rel_tolerance;
    else if (s == "lbfgs_memory") *value = c.lbfgs_memory;
    else throw std::invalid_argument("unknown Optim setting: " + s);
  });
}

int lk_linalg_set(const char* name, double value) {
  return guarded([&] {
    lk::Runtime& rt = lk::runtime_for(lk::default_device());
    const std::string s = name ? name : "";
    if (s == "num_nugget") rt.chol.num_nugget = value;
    else if (s == "chol_rcond_check") rt.chol.rcond_check = value != 0;
    else if (s == "max_inc_choldiag") rt.chol.max_inc = (int)value;
    else if (s == "min_rcond") rt.chol.min_rcond = value;
    else throw std::invalid_argument("unknown LinearAlgebra setting: " + s);
  });
}

int lk_linalg_get(const char* name, double* value) {
  return guarded([&] {
    if (value == nullptr) throw std::invalid_argument("NULL argument");
    lk::Runtime& rt = lk::runtime_for(lk::default_device());
    const std::string s = name ? name : "";
    if (s == "num_nugget") *value = rt.chol.num_nugget;
    else if (s == "chol_rcond_check") *value = rt.chol.rcond_check ? 1.0 : 0.0;
    else if (s == "max_inc_choldiag") *value = rt.chol.max_inc;
    else if (s == "min_rcond") *value = rt.chol.min_rcond;
    else throw std::invalid_argument("unknown LinearAlgebra setting: " + s);
  });
}

int lk_linalg_chol_safe(const double* X, int n, double* L_out, int* inc_out) {
  return guarded([&] {
    if (X == nullptr || L_out == nullptr) throw std::invalid_argument("NULL argument");
    if (n <= 0) throw std::invalid_argument("matrix must have at least one row");
    int inc = 0;
    lk::runtime_for(lk::default_device()).safe_chol(X, n, L_out, &inc);
    if (inc_out) *inc_out = inc;
  });
}

int lk_linalg_rcond_chol(const double* L, int n, double* rcond_out) {
  return guarded([&] {
    if (L == nullptr || rcond_out == nullptr) throw std::invalid_argument("NULL argument");
    if (n <= 0) throw std::invalid_argument("matrix must have at least one row");
    *rcond_out = lk::runtime_for(lk::default_device()).rcond_chol(L, n);
  });
}

int lk_linalg_rcond_approx_chol(const double* L, int n, double* rcond_out) {
  return guarded([&] {
    if (L == nullptr || rcond_out == nullptr) throw std::invalid_argument("NULL argument");
    if (n <= 0) throw std::invalid_argument("matrix must have at least one row");
    // (min / max of the diagonal)^2, the scan of LinearAlgebra.cpp:120-139
    double lo = L[0], hi = L[0];
    for (int i = 1; i < n; ++i) {
      const double c = L[i + (size_t)i * n];
      if (c < lo) lo = c;
      else if (c > hi) hi = c;
    }
    const double r = lo / hi;
    *rcond_out = r * r;
  });
}

int lk_linalg_chol_block(const double* C, int n, const double* Loo, int no, double* L_out) {
  return guarded([&] {
    if (C == nullptr || Loo == nullptr || L_out == nullptr) throw std::invalid_argument("NULL argument");
    if (n <= 0) throw std::invalid_argument("matrix must have at least one row");
    lk::runtime_for(lk::default_device()).chol_block(C, n, Loo, no, L_out);
  });
}

int lk_random_reset_seed(unsigned seed) {
  return guarded([&] { lk::random_reset_seed(seed); });
}

int lk_lbfgsb_minimize(int n, double* x, const double* lower, const double* upper, lk_objective_fn fn, void* user,
                       int max_iter, double pgtol, double factr, int m, double* f_out, int* iters_out,
                       int* evals_out, char* task_out, int task_len) {
  return guarded([&] {
    if (n <= 0 || x == nullptr || lower == nullptr || upper == nullptr || fn == nullptr)
      throw std::invalid_argument("NULL argument");
    lk::LBFGSB opt(n, m);
    opt.max_iter = max_iter;
    opt.pgtol = pgtol;
    opt.factr = factr;
    std::vector<double> x0(x, x + n), l(lower, lower + n), u(upper, upper + n), g(n);
    opt.start(x0, l, u);
    while (opt.need_eval()) {
      const double f = fn(opt.x().data(), g.data(), user);
      opt.tell(f, g);
    }
    for (int i = 0; i < n; ++i) x[i] = opt.x_opt()[i];
    if (f_out) *f_out = opt.f_opt();
    if (iters_out) *iters_out = opt.num_iters();
    if (evals_out) *evals_out = opt.num_evals();
    if (task_out && task_len > 0) {
      std::strncpy(task_out, opt.task().c_str(), task_len - 1);
      task_out[task_len - 1] = '\0';
    }
  });
}

}  // extern "C"
