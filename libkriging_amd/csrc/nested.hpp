// NestedKriging (divide-and-conquer Kriging) on the GPU pipeline.  Mirrors
// the reference class NestedKriging (src/lib/include/libKriging/NestedKriging.hpp:
// 17-176, src/lib/NestedKriging.cpp) for plain Kriging submodels:
//   fit      partition -> one Kriging fit per group -> common prior (weighted
//            geometric-mean theta, weighted sigma2 and beta0) -> refit with
//            optim = "none" -> NK precomputation (chol(R_g + 1e-10 I), alpha_g)
//   predict  NK optimal aggregation, or PoE / gPoE / BCM / rBCM.
// Every group's fit runs in the same batched pipelines (uneven sizes side by
// side), and the NK aggregation's group-pair contractions run on FP64 MFMA.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "kriging.hpp"

namespace lk {

class NestedKriging {
 public:
  enum class Aggregation { PoE, gPoE, BCM, rBCM, NK };
  enum class Partition { Random, KMeans };
  static Aggregation aggregation_from_string(const std::string& s);
  static std::string aggregation_to_string(Aggregation a);
  static Partition partition_from_string(const std::string& s);

  explicit NestedKriging(const std::string& kernel);
  ~NestedKriging();
  NestedKriging(const NestedKriging&) = delete;
  NestedKriging& operator=(const NestedKriging&) = delete;

  void set_aggregation(Aggregation a) { aggregation_ = a; }
  void set_partition(Partition p) { partition_ = p; }
  void set_seed(int seed) { seed_ = seed; }
  void set_predict_chunk(int chunk) { chunk_ = chunk; }

  // X: n x d column-major.  Groups from the partition method (fit) or given
  // explicitly as a group index per point, 0..nb_groups-1 (fit_groups).
  void fit(const double* y, const double* X, int n, int d, int nb_groups, RegModel regmodel = RegModel::Constant,
           const std::string& optim = "BFGS", const std::string& objective = "LL",
           const Parameters& parameters = Parameters{});
  void fit_groups(const double* y, const double* X, int n, int d, const int* group_of, int nb_groups,
                  RegModel regmodel = RegModel::Constant, const std::string& optim = "BFGS",
                  const std::string& objective = "LL", const Parameters& parameters = Parameters{});

  // Xn: m x d column-major; stdev may be null
  void predict(const double* Xn, int m, bool want_stdev, double* mean, double* stdev);

  const std::string& kernel() const { return kernel_; }
  Aggregation aggregation() const { return aggregation_; }
  int nb_groups() const { return (int)groups_.size(); }
  const std::vector<std::vector<int>>& groups() const { return groups_; }
  const Kriging& submodel(int i) const;
  Kriging& submodel(int i);
  const std::vector<double>& theta() const { return theta_; }
  double sigma2() const { return sigma2_; }
  double beta0() const { return beta0_; }
  const std::vector<double>& X() const { return X_; }
  const std::vector<double>& y() const { return y_; }
  int n() const { return n_; }
  int d() const { return d_; }
  int predict_fallbacks() const { return fallbacks_; }  // points solved by LU / pinv in the last predict
  // HIP-event timing of the NK pair-contraction launches of later predicts
  void set_pair_timing(bool on);
  // (ms, launches, points) accumulated since the last call; resets them
  void take_pair_time(double* ms, int* launches, int* points);
  // evaluations and batched pipelines of the groups' optimisation in the last fit
  int fit_evaluations() const { return fit_evals_; }
  // the groups' own fitted hyper-parameters, before unification: theta
  // (nb_groups x d, group-major), sigma2 and the constant-trend beta0
  const std::vector<double>& group_theta() const { return gtheta_; }
  const std::vector<double>& group_sigma2() const { return gsigma2_; }
  const std::vector<double>& group_beta0() const { return gbeta0_; }
  int fit_batches() const { return fit_batches_; }
  std::string summary() const;

 private:
  void check_args(int n, int d, int nb_groups, RegModel regmodel) const;
  void make_partition(int nb_groups);
  void fit_impl(RegModel regmodel, const std::string& optim, const std::string& objective,
                const Parameters& parameters);
  void unify(const std::string& objective);
  void precompute_nk();
  void release_nk();
  void predict_nk(const double* Xn, int m, bool want_stdev, double* mean, double* stdev);
  void predict_poe(const double* Xn, int m, bool want_stdev, double* mean, double* stdev);

  std::string kernel_;
  Aggregation aggregation_ = Aggregation::NK;
  Partition partition_ = Partition::KMeans;
  int seed_ = 123;
  int chunk_ = 0;  // 0: sized from the device memory budget
  RegModel regmodel_ = RegModel::Constant;

  int n_ = 0, d_ = 0;
  std::vector<double> X_, y_;  // n x d column-major, n
  std::vector<std::vector<int>> groups_;
  std::vector<std::unique_ptr<Kriging>> sub_;
  std::vector<double> theta_;
  double sigma2_ = 0.0, beta0_ = 0.0;
  std::vector<double> gtheta_, gsigma2_, gbeta0_;
  bool fitted_ = false;
  int fallbacks_ = 0;
  int fit_evals_ = 0, fit_batches_ = 0;
  bool time_pairs_ = false;
  hipEvent_t pev_[2] = {nullptr, nullptr};
  double pair_ms_ = 0.0;
  int pair_launches_ = 0, pair_points_ = 0;

  // NK device state (group-ordered, padded to 64-row tiles)
  int device_ = 0;
  Runtime* rt_ = nullptr;
  int Npad_ = 0;
  std::vector<int> goff_, gpad_;
  double* d_Xg_ = nullptr;
  double* d_Linv_ = nullptr;
  double* d_alpha_ = nullptr;
  void* d_groups_ = nullptr;
  int* d_tile_group_ = nullptr;
  void* d_pairs_ = nullptr;
  int npairs_ = 0;
};

}  // namespace lk
