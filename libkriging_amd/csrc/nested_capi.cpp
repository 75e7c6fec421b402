// extern "C" NestedKriging object API, mirroring the reference's C ABI
// (bindings/Julia/jlibkriging/csrc/libkriging_c.h:394-456,
// libkriging_c.cpp:1663-1870): same names, argument meaning and the 0 / -1 +
// lk_get_last_error() convention (NULL handle from new_fit on error).
#include <cmath>
#include <cstring>
#include <exception>
#include <limits>
#include <string>
#include <vector>

#include "../../include/lkgpu.h"
#include "nested.hpp"

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

lk::NestedKriging* NK(void* p) {
  if (p == nullptr) throw std::invalid_argument("NestedKriging handle is NULL");
  return static_cast<lk::NestedKriging*>(p);
}

// lk_nested_make_params (libkriging_c.cpp:1663-1682): theta is one row
lk::Parameters nested_params(int d, const double* sigma2, int is_sigma2_estim, const double* theta, int theta_n,
                             int is_theta_estim, const double* beta, int beta_n, int is_beta_estim) {
  lk::Parameters P;
  if (sigma2) P.sigma2 = *sigma2;
  P.is_sigma2_estim = is_sigma2_estim != 0;
  if (theta && theta_n > 0) {
    if (theta_n % d != 0) throw std::runtime_error("Dimension of theta should be nx" + std::to_string(d));
    P.theta = std::vector<double>(theta, theta + theta_n);
    P.theta_rows = theta_n / d;
  }
  P.is_theta_estim = is_theta_estim != 0;
  if (beta && beta_n > 0) P.beta = std::vector<double>(beta, beta + beta_n);
  P.is_beta_estim = is_beta_estim != 0;
  return P;
}

void check_no_warping(int n_warping) {
  if (n_warping > 0)
    throw std::invalid_argument("warping (WarpKriging submodels) is not supported on this path");
}

void check_objective(const char* objective) {
  if (objective != nullptr && std::string(objective).rfind("LLVecchia", 0) == 0)
    throw std::invalid_argument("LLVecchia objective is not supported on this path");
}

thread_local std::string g_str;
const char* ret_str(const std::string& s) {
  g_str = s;
  return g_str.c_str();
}
}  // namespace

extern "C" {

void* lk_nested_kriging_new(const char* kernel, const char* aggregation, const char* partition, int seed) {
  lk::NestedKriging* k = nullptr;
  const int rc = guarded([&] {
    auto* nk = new lk::NestedKriging(kernel ? kernel : "matern3_2");
    try {
      nk->set_aggregation(lk::NestedKriging::aggregation_from_string(aggregation ? aggregation : "NK"));
      nk->set_partition(lk::NestedKriging::partition_from_string(partition ? partition : "kmeans"));
      nk->set_seed(seed);
    } catch (...) {
      delete nk;
      throw;
    }
    k = nk;
  });
  return rc == 0 ? k : nullptr;
}

void* lk_nested_kriging_new_fit(const double* y, int n, const double* X, int nX, int d, const char* kernel,
                                int nb_groups, const char* aggregation, const char* partition, int seed,
                                const char** warping, int n_warping, const char* regmodel, const char* optim,
                                const char* objective, const double* sigma2, int is_sigma2_estim, const double* theta,
                                int theta_n, int is_theta_estim, const double* beta, int beta_n, int is_beta_estim) {
  (void)warping;
  lk::NestedKriging* k = nullptr;
  const int rc = guarded([&] {
    check_no_warping(n_warping);
    check_objective(objective);
    if (n != nX) throw std::invalid_argument("y and X should have the same number of rows");
    auto* nk = new lk::NestedKriging(kernel ? kernel : "matern3_2");
    try {
      nk->set_aggregation(lk::NestedKriging::aggregation_from_string(aggregation ? aggregation : "NK"));
      nk->set_partition(lk::NestedKriging::partition_from_string(partition ? partition : "kmeans"));
      nk->set_seed(seed);
      nk->fit(y, X, n, d, nb_groups, lk::regmodel_from_string(regmodel ? regmodel : "constant"),
              optim ? optim : "BFGS", objective ? objective : "LL",
              nested_params(d, sigma2, is_sigma2_estim, theta, theta_n, is_theta_estim, beta, beta_n, is_beta_estim));
    } catch (...) {
      delete nk;
      throw;
    }
    k = nk;
  });
  return rc == 0 ? k : nullptr;
}

void lk_nested_kriging_delete(void* ptr) { delete static_cast<lk::NestedKriging*>(ptr); }

int lk_nested_kriging_fit(void* ptr, const double* y, int n, const double* X, int nX, int d, int nb_groups,
                          const char** warping, int n_warping, const char* regmodel, const char* optim,
                          const char* objective, const double* sigma2, int is_sigma2_estim, const double* theta,
                          int theta_n, int is_theta_estim, const double* beta, int beta_n, int is_beta_estim) {
  (void)warping;
  return guarded([&] {
    check_no_warping(n_warping);
    check_objective(objective);
    if (n != nX) throw std::invalid_argument("y and X should have the same number of rows");
    NK(ptr)->fit(y, X, n, d, nb_groups, lk::regmodel_from_string(regmodel ? regmodel : "constant"),
                 optim ? optim : "BFGS", objective ? objective : "LL",
                 nested_params(d, sigma2, is_sigma2_estim, theta, theta_n, is_theta_estim, beta, beta_n,
                               is_beta_estim));
  });
}

int lk_nested_kriging_fit_groups(void* ptr, const double* y, int n, const double* X, int nX, int d,
                                 const int* group_of, int nb_groups, const char* regmodel, const char* optim,
                                 const char* objective, const double* sigma2, int is_sigma2_estim,
                                 const double* theta, int theta_n, int is_theta_estim, const double* beta,
                                 int beta_n, int is_beta_estim) {
  return guarded([&] {
    check_objective(objective);
    if (n != nX) throw std::invalid_argument("y and X should have the same number of rows");
    if (group_of == nullptr) throw std::invalid_argument("group_of is NULL");
    NK(ptr)->fit_groups(y, X, n, d, group_of, nb_groups, lk::regmodel_from_string(regmodel ? regmodel : "constant"),
                        optim ? optim : "BFGS", objective ? objective : "LL",
                        nested_params(d, sigma2, is_sigma2_estim, theta, theta_n, is_theta_estim, beta, beta_n,
                                      is_beta_estim));
  });
}

int lk_nested_kriging_predict(void* ptr, const double* X_n, int m, int d, int return_stdev, double* mean_out,
                              double* stdev_out) {
  return guarded([&] {
    lk::NestedKriging* k = NK(ptr);
    if (d != k->d()) throw std::invalid_argument("X_n should have the same number of columns as X");
    std::vector<double> mean(m), sd(m);
    k->predict(X_n, m, return_stdev != 0, mean.data(), sd.data());
    if (mean_out) std::memcpy(mean_out, mean.data(), sizeof(double) * m);
    if (stdev_out && return_stdev) std::memcpy(stdev_out, sd.data(), sizeof(double) * m);
  });
}

const char* lk_nested_kriging_summary(void* ptr) {
  std::string s;
  if (guarded([&] { s = NK(ptr)->summary(); }) != 0) return nullptr;
  return ret_str(s);
}
const char* lk_nested_kriging_kernel(void* ptr) {
  std::string s;
  if (guarded([&] { s = NK(ptr)->kernel(); }) != 0) return nullptr;
  return ret_str(s);
}
const char* lk_nested_kriging_aggregation(void* ptr) {
  std::string s;
  if (guarded([&] { s = lk::NestedKriging::aggregation_to_string(NK(ptr)->aggregation()); }) != 0) return nullptr;
  return ret_str(s);
}
int lk_nested_kriging_nb_groups(void* ptr) {
  int g = -1;
  guarded([&] { g = NK(ptr)->nb_groups(); });
  return g;
}
int lk_nested_kriging_get_theta(void* ptr, double* out, int* n) {
  return guarded([&] {
    const std::vector<double>& t = NK(ptr)->theta();
    if (n) *n = (int)t.size();
    if (out) std::memcpy(out, t.data(), sizeof(double) * t.size());
  });
}
double lk_nested_kriging_get_sigma2(void* ptr) {
  double v = std::numeric_limits<double>::quiet_NaN();
  guarded([&] { v = NK(ptr)->sigma2(); });
  return v;
}
double lk_nested_kriging_get_beta0(void* ptr) {
  double v = std::numeric_limits<double>::quiet_NaN();
  guarded([&] { v = NK(ptr)->beta0(); });
  return v;
}

int lk_nested_kriging_groups(void* ptr, int* group_of_out) {
  return guarded([&] {
    lk::NestedKriging* k = NK(ptr);
    const auto& G = k->groups();
    for (size_t g = 0; g < G.size(); ++g)
      for (int i : G[g]) group_of_out[i] = (int)g;
  });
}

void* lk_nested_kriging_submodel(void* ptr, int i) {
  void* s = nullptr;
  guarded([&] { s = &NK(ptr)->submodel(i); });
  return s;
}

int lk_nested_kriging_set_predict_chunk(void* ptr, int chunk) {
  return guarded([&] { NK(ptr)->set_predict_chunk(chunk); });
}

int lk_nested_kriging_fit_statistics(void* ptr, int* evaluations, int* batches) {
  return guarded([&] {
    lk::NestedKriging* k = NK(ptr);
    if (evaluations) *evaluations = k->fit_evaluations();
    if (batches) *batches = k->fit_batches();
  });
}

int lk_nested_kriging_group_fits(void* ptr, double* theta_out, double* sigma2_out, double* beta0_out) {
  return guarded([&] {
    lk::NestedKriging* k = NK(ptr);
    const auto& t = k->group_theta();
    if (t.empty()) throw std::runtime_error("NestedKriging is not fitted");
    if (theta_out) std::memcpy(theta_out, t.data(), sizeof(double) * t.size());
    if (sigma2_out) std::memcpy(sigma2_out, k->group_sigma2().data(), sizeof(double) * k->group_sigma2().size());
    if (beta0_out) std::memcpy(beta0_out, k->group_beta0().data(), sizeof(double) * k->group_beta0().size());
  });
}

int lk_nested_kriging_time_pairs(void* ptr, int enable) {
  return guarded([&] { NK(ptr)->set_pair_timing(enable != 0); });
}

int lk_nested_kriging_pair_time(void* ptr, double* ms_out, int* launches_out, int* points_out) {
  return guarded([&] { NK(ptr)->take_pair_time(ms_out, launches_out, points_out); });
}

int lk_nested_kriging_predict_fallbacks(void* ptr) {
  int v = -1;
  guarded([&] { v = NK(ptr)->predict_fallbacks(); });
  return v;
}

}  // extern "C"
