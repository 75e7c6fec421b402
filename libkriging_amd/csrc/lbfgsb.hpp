// L-BFGS-B, bound-constrained limited-memory BFGS (R. H. Byrd, P. Lu,
// J. Nocedal, C. Zhu, SIAM J. Sci. Comput. 16 (1995) 1190-1208, with the
// subspace-minimisation projection of J. L. Morales, J. Nocedal, ACM TOMS 38
// (2011) article 7, "version 3.0"), written as a reverse-communication state
// machine so that many independent optimisers (multistart fits) can share one
// batched likelihood evaluation per step.
//
// The reference drives lbfgsb_cpp (a C++ port of L-BFGS-B 3.0, pinned only by
// its git submodule URL, src/lib/Kriging.cpp:1950-1987) with max_iter, pgtol,
// factr and box bounds; this class keeps those semantics: projected-gradient
// and relative-reduction stopping tests, Moré-Thuente line search (ftol 1e-3,
// gtol 0.9, xtol 0.1), memory refresh on a failed line search, and the task
// strings the reference inspects ("ABNORMAL_TERMINATION_IN_LNSRCH", ...).
// The problem dimension here is small (d + 1 <= 33), so the limited-memory
// matrix B = theta I - W M W^T is formed densely and the generalized Cauchy
// point and the subspace step are computed on it directly.
#pragma once
#include <string>
#include <vector>

namespace lk {

class LBFGSB {
 public:
  explicit LBFGSB(int n, int m = 5);

  int max_iter = 20;     // Optim::max_iteration (Optim.cpp:111)
  double pgtol = 1e-3;   // Optim::gradient_tolerance (Optim.cpp:121)
  double factr = 1e10;   // objective_rel_tolerance / 1e-13 (Kriging.cpp:1958)

  // Start from x0 (projected onto [l, u]).  After this, need_eval() is true.
  void start(const std::vector<double>& x0, const std::vector<double>& l, const std::vector<double>& u);
  bool need_eval() const { return state_ == S_FG_START || state_ == S_FG_LNSRCH; }
  bool done() const { return state_ == S_DONE; }
  // point at which f and g are requested
  const std::vector<double>& x() const { return x_; }
  // supply f(x()) and g(x()); advances to the next request or to done()
  void tell(double f, const std::vector<double>& g);

  // results (valid once done())
  const std::string& task() const { return task_; }
  int num_iters() const { return iter_; }
  double f_opt() const { return f_; }
  const std::vector<double>& x_opt() const { return x_; }
  int num_evals() const { return nfgv_; }

 private:
  enum State { S_IDLE, S_FG_START, S_FG_LNSRCH, S_DONE };
  void iterate();            // from a new accepted point: next search direction
  void begin_line_search();  // lnsrlb setup
  void line_search_step();   // after an evaluation inside the line search
  void after_new_x();        // convergence tests + memory update
  void finish(const std::string& t) {
    task_ = t;
    state_ = S_DONE;
  }
  void form_B();             // dense B = theta I - W M W^T
  bool cauchy_and_subspace(std::vector<double>& z);
  double proj_grad_norm() const;

  // Moré-Thuente line search (MINPACK-2 dcsrch / dcstep)
  struct Dcsrch {
    bool brackt = false;
    int stage = 1;
    double ginit = 0, gtest = 0, gx = 0, gy = 0, finit = 0, fx = 0, fy = 0, stx = 0, sty = 0, stmin = 0, stmax = 0,
           width = 0, width1 = 0;
    std::string task;
  };
  void dcsrch(double f, double g, double& stp, double ftol, double gtol, double xtol, double stpmin, double stpmax,
              Dcsrch& s);
  static void dcstep(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy, double& stp,
                     double fp, double dp, bool& brackt, double stpmin, double stpmax);

  int n_, m_;
  State state_ = S_IDLE;
  std::string task_;
  std::vector<double> x_, l_, u_, g_;
  double f_ = 0;
  int iter_ = 0, nfgv_ = 0;
  // limited memory
  std::vector<std::vector<double>> S_, Y_;  // newest last
  double theta_ = 1.0;
  std::vector<double> B_;                   // n x n dense
  // line search
  std::vector<double> d_, t_, r_, z_;       // direction, saved x, saved g, subspace point
  double fold_ = 0, gd_ = 0, gdold_ = 0, stp_ = 0, stpmx_ = 0, dnorm_ = 0;
  int ifun_ = 0, iback_ = 0;
  Dcsrch ls_;
};

}  // namespace lk
