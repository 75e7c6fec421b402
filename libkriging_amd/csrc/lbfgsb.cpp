#include "lbfgsb.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

namespace lk {

namespace {
constexpr double EPSMCH = std::numeric_limits<double>::epsilon();
constexpr double FTOL = 1e-3, GTOL = 0.9, XTOL = 0.1;  // lnsrlb's dcsrch constants
constexpr double BIG = 1e10;

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// in-place Cholesky of an SPD k x k row-major matrix; false if not SPD
bool chol(std::vector<double>& A, int k) {
  for (int j = 0; j < k; ++j) {
    double s = A[j * k + j];
    for (int q = 0; q < j; ++q) s -= A[j * k + q] * A[j * k + q];
    if (!(s > 0.0)) return false;
    const double ljj = std::sqrt(s);
    A[j * k + j] = ljj;
    for (int i = j + 1; i < k; ++i) {
      double t = A[i * k + j];
      for (int q = 0; q < j; ++q) t -= A[i * k + q] * A[j * k + q];
      A[i * k + j] = t / ljj;
    }
  }
  return true;
}

void chol_solve(const std::vector<double>& Lc, int k, std::vector<double>& b) {
  for (int i = 0; i < k; ++i) {
    double s = b[i];
    for (int q = 0; q < i; ++q) s -= Lc[i * k + q] * b[q];
    b[i] = s / Lc[i * k + i];
  }
  for (int i = k - 1; i >= 0; --i) {
    double s = b[i];
    for (int q = i + 1; q < k; ++q) s -= Lc[q * k + i] * b[q];
    b[i] = s / Lc[i * k + i];
  }
}

// general k x k inverse by Gauss-Jordan with partial pivoting (k <= 2m)
bool invert(std::vector<double> A, int k, std::vector<double>& inv) {
  inv.assign(k * k, 0.0);
  for (int i = 0; i < k; ++i) inv[i * k + i] = 1.0;
  for (int c = 0; c < k; ++c) {
    int piv = c;
    for (int r = c + 1; r < k; ++r)
      if (std::fabs(A[r * k + c]) > std::fabs(A[piv * k + c])) piv = r;
    if (A[piv * k + c] == 0.0) return false;
    if (piv != c)
      for (int q = 0; q < k; ++q) {
        std::swap(A[c * k + q], A[piv * k + q]);
        std::swap(inv[c * k + q], inv[piv * k + q]);
      }
    const double p = A[c * k + c];
    for (int q = 0; q < k; ++q) {
      A[c * k + q] /= p;
      inv[c * k + q] /= p;
    }
    for (int r = 0; r < k; ++r) {
      if (r == c) continue;
      const double f = A[r * k + c];
      if (f == 0.0) continue;
      for (int q = 0; q < k; ++q) {
        A[r * k + q] -= f * A[c * k + q];
        inv[r * k + q] -= f * inv[c * k + q];
      }
    }
  }
  return true;
}
}  // namespace

LBFGSB::LBFGSB(int n, int m) : n_(n), m_(m) {}

void LBFGSB::start(const std::vector<double>& x0, const std::vector<double>& l, const std::vector<double>& u) {
  l_ = l;
  u_ = u;
  x_.resize(n_);
  for (int i = 0; i < n_; ++i) x_[i] = std::max(l_[i], std::min(u_[i], x0[i]));
  S_.clear();
  Y_.clear();
  theta_ = 1.0;
  iter_ = 0;
  nfgv_ = 0;
  task_ = "FG_START";
  state_ = S_FG_START;
}

double LBFGSB::proj_grad_norm() const {
  double nrm = 0.0;
  for (int i = 0; i < n_; ++i) {
    double gi = g_[i];
    if (gi < 0.0)
      gi = std::max(x_[i] - u_[i], gi);
    else
      gi = std::min(x_[i] - l_[i], gi);
    nrm = std::max(nrm, std::fabs(gi));
  }
  return nrm;
}

void LBFGSB::tell(double f, const std::vector<double>& g) {
  ++nfgv_;
  if (state_ == S_FG_START) {
    f_ = f;
    g_ = g;
    if (proj_grad_norm() <= pgtol) {
      finish("CONVERGENCE: NORM_OF_PROJECTED_GRADIENT_<=_PGTOL");
      return;
    }
    iterate();
  } else if (state_ == S_FG_LNSRCH) {
    f_ = f;
    g_ = g;
    line_search_step();
  }
}

// Dense B = theta I - W M W^T with W = [Y, theta S] and
// M = [[-D, L^T], [L, theta S^T S]]^-1 (Byrd-Nocedal-Schnabel compact form).
void LBFGSB::form_B() {
  const int n = n_, k = (int)S_.size();
  B_.assign(n * n, 0.0);
  for (int i = 0; i < n; ++i) B_[i * n + i] = theta_;
  if (k == 0) return;
  std::vector<double> Mi(4 * k * k, 0.0);
  const int K2 = 2 * k;
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      const double sy = dot(S_[i], Y_[j]);
      if (i == j) Mi[i * K2 + j] = -sy;                 // -D
      if (i > j) Mi[(k + i) * K2 + j] = sy;             // L (lower-left)
      if (j > i) Mi[i * K2 + (k + j)] = dot(S_[j], Y_[i]);  // L^T (upper-right)
      Mi[(k + i) * K2 + (k + j)] = theta_ * dot(S_[i], S_[j]);
    }
  std::vector<double> M;
  if (!invert(Mi, K2, M)) return;  // B stays theta I
  // W rows: w_r = [Y_1r .. Y_kr, theta S_1r .. theta S_kr]
  std::vector<double> W(n * K2);
  for (int r = 0; r < n; ++r)
    for (int i = 0; i < k; ++i) {
      W[r * K2 + i] = Y_[i][r];
      W[r * K2 + k + i] = theta_ * S_[i][r];
    }
  std::vector<double> WM(n * K2, 0.0);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < K2; ++c) {
      double s = 0.0;
      for (int q = 0; q < K2; ++q) s += W[r * K2 + q] * M[q * K2 + c];
      WM[r * K2 + c] = s;
    }
  for (int r = 0; r < n; ++r)
    for (int c = 0; c < n; ++c) {
      double s = 0.0;
      for (int q = 0; q < K2; ++q) s += WM[r * K2 + q] * W[c * K2 + q];
      B_[r * n + c] -= s;
    }
}

// Generalized Cauchy point along the projected steepest-descent path, then
// minimisation of the quadratic model over the variables left free, with the
// projection / backtracking of L-BFGS-B 3.0.  z receives the target point.
bool LBFGSB::cauchy_and_subspace(std::vector<double>& z) {
  const int n = n_;
  form_B();
  // breakpoints
  std::vector<double> d(n), t(n, std::numeric_limits<double>::infinity());
  std::vector<int> fixed(n, 0);
  for (int i = 0; i < n; ++i) {
    d[i] = -g_[i];
    if (g_[i] < 0.0 && u_[i] < std::numeric_limits<double>::infinity()) t[i] = (x_[i] - u_[i]) / g_[i];
    if (g_[i] > 0.0 && l_[i] > -std::numeric_limits<double>::infinity()) t[i] = (x_[i] - l_[i]) / g_[i];
    if (t[i] <= 0.0) {  // at a bound and pushed outward: fixed
      d[i] = 0.0;
      fixed[i] = 1;
    }
    if (g_[i] == 0.0) d[i] = 0.0;
  }
  std::vector<int> order;
  for (int i = 0; i < n; ++i)
    if (!fixed[i] && d[i] != 0.0) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int a, int b) { return t[a] < t[b]; });
  auto matvec = [&](const std::vector<double>& v, std::vector<double>& o) {
    for (int r = 0; r < n; ++r) {
      double s = 0.0;
      for (int c = 0; c < n; ++c) s += B_[r * n + c] * v[c];
      o[r] = s;
    }
  };
  // walk the piecewise-linear path x + z(t): on each segment the model
  // m(z) = g'z + z'Bz/2 is a quadratic in t with slope f1 and curvature f2
  std::vector<double> Bd(n), zc(n, 0.0);
  double told = 0.0;
  size_t ib = 0;
  for (;;) {
    bool any = false;
    for (int i = 0; i < n; ++i) any = any || d[i] != 0.0;
    if (!any) break;
    matvec(d, Bd);
    const double f1 = dot(g_, d) + dot(zc, Bd);
    if (f1 >= 0.0) break;
    double f2 = dot(d, Bd);
    f2 = std::max(f2, EPSMCH * std::fabs(f1));
    const double dtm = -f1 / f2;
    const double tnext = (ib < order.size()) ? t[order[ib]] : std::numeric_limits<double>::infinity();
    if (dtm < tnext - told) {  // minimiser inside this segment
      for (int i = 0; i < n; ++i) zc[i] += dtm * d[i];
      break;
    }
    // advance to the breakpoint and fix that variable at its bound
    const double dt = tnext - told;
    for (int i = 0; i < n; ++i) zc[i] += dt * d[i];
    told = tnext;
    const int b = order[ib++];
    zc[b] = ((d[b] > 0.0) ? u_[b] : l_[b]) - x_[b];
    d[b] = 0.0;
    fixed[b] = 1;
  }
  std::vector<double> xc(n);
  for (int i = 0; i < n; ++i) xc[i] = std::max(l_[i], std::min(u_[i], x_[i] + zc[i]));

  // subspace minimisation over the free variables (at least one correction)
  std::vector<int> freev;
  for (int i = 0; i < n; ++i)
    if (!fixed[i] && l_[i] < u_[i]) freev.push_back(i);
  z = xc;
  if (S_.empty() || freev.empty()) return true;
  const int nf = (int)freev.size();
  std::vector<double> zcp(n);
  for (int i = 0; i < n; ++i) zcp[i] = xc[i] - x_[i];
  std::vector<double> Bz2(n);
  matvec(zcp, Bz2);
  std::vector<double> rhs(nf), BF(nf * nf);
  for (int a = 0; a < nf; ++a) {
    rhs[a] = -(g_[freev[a]] + Bz2[freev[a]]);
    for (int b = 0; b < nf; ++b) BF[a * nf + b] = B_[freev[a] * n + freev[b]];
  }
  if (!chol(BF, nf)) return false;
  chol_solve(BF, nf, rhs);  // rhs <- newton step on the free variables
  // projection of the Newton point (3.0)
  std::vector<double> xp = xc;
  for (int a = 0; a < nf; ++a) {
    const int k = freev[a];
    xp[k] = std::max(l_[k], std::min(u_[k], xc[k] + rhs[a]));
  }
  double dd_p = 0.0;
  for (int i = 0; i < n; ++i) dd_p += (xp[i] - x_[i]) * g_[i];
  if (dd_p <= 0.0) {
    z = xp;
    return true;
  }
  // backtracking along the Newton direction from the Cauchy point
  double alpha = 1.0, temp1 = 1.0;
  int ibd = -1;
  for (int a = 0; a < nf; ++a) {
    const int k = freev[a];
    const double dk = rhs[a];
    if (dk < 0.0) {
      const double temp2 = l_[k] - xc[k];
      if (temp2 >= 0.0)
        temp1 = 0.0;
      else if (dk * alpha < temp2)
        temp1 = temp2 / dk;
    } else if (dk > 0.0) {
      const double temp2 = u_[k] - xc[k];
      if (temp2 <= 0.0)
        temp1 = 0.0;
      else if (dk * alpha > temp2)
        temp1 = temp2 / dk;
    }
    if (temp1 < alpha) {
      alpha = temp1;
      ibd = a;
    }
  }
  z = xc;
  if (alpha < 1.0 && ibd >= 0) {
    const int k = freev[ibd];
    z[k] = (rhs[ibd] > 0.0) ? u_[k] : l_[k];
    rhs[ibd] = 0.0;
  }
  for (int a = 0; a < nf; ++a) z[freev[a]] += alpha * rhs[a];
  return true;
}

void LBFGSB::iterate() {
  // new search direction from the current accepted point x_, f_, g_
  for (;;) {
    std::vector<double> z;
    if (!cauchy_and_subspace(z)) {
      // singular reduced system: refresh the memory and restart
      S_.clear();
      Y_.clear();
      theta_ = 1.0;
      continue;
    }
    z_ = z;
    d_.resize(n_);
    for (int i = 0; i < n_; ++i) d_[i] = z_[i] - x_[i];
    begin_line_search();
    return;
  }
}

// lnsrlb: maximal feasible step, initial step 1 (every variable is boxed)
void LBFGSB::begin_line_search() {
  dnorm_ = std::sqrt(dot(d_, d_));
  stpmx_ = BIG;
  if (iter_ == 0) {
    stpmx_ = 1.0;
  } else {
    for (int i = 0; i < n_; ++i) {
      const double a1 = d_[i];
      if (a1 < 0.0) {
        const double a2 = l_[i] - x_[i];
        if (a2 >= 0.0)
          stpmx_ = 0.0;
        else if (a1 * stpmx_ < a2)
          stpmx_ = a2 / a1;
      } else if (a1 > 0.0) {
        const double a2 = u_[i] - x_[i];
        if (a2 <= 0.0)
          stpmx_ = 0.0;
        else if (a1 * stpmx_ > a2)
          stpmx_ = a2 / a1;
      }
    }
  }
  stp_ = 1.0;
  t_ = x_;
  r_ = g_;
  fold_ = f_;
  ifun_ = 0;
  iback_ = 0;
  ls_ = Dcsrch{};
  ls_.task = "START";
  gd_ = dot(g_, d_);
  gdold_ = gd_;
  if (gd_ >= 0.0 || dnorm_ == 0.0) {
    // not a descent direction: the line search cannot proceed
    if (S_.empty()) {
      x_ = t_;
      g_ = r_;
      f_ = fold_;
      ++iter_;
      finish("ABNORMAL_TERMINATION_IN_LNSRCH");
      return;
    }
    S_.clear();
    Y_.clear();
    theta_ = 1.0;
    iterate();
    return;
  }
  dcsrch(f_, gd_, stp_, FTOL, GTOL, XTOL, 0.0, stpmx_, ls_);
  // first trial point
  ++ifun_;
  iback_ = ifun_ - 1;
  if (stp_ == 1.0)
    x_ = z_;
  else
    for (int i = 0; i < n_; ++i) x_[i] = stp_ * d_[i] + t_[i];
  state_ = S_FG_LNSRCH;
}

void LBFGSB::line_search_step() {
  gd_ = dot(g_, d_);
  dcsrch(f_, gd_, stp_, FTOL, GTOL, XTOL, 0.0, stpmx_, ls_);
  const bool converged = ls_.task.rfind("CONV", 0) == 0 || ls_.task.rfind("WARN", 0) == 0;
  if (!converged && ls_.task.rfind("ERROR", 0) != 0 && iback_ < 20) {
    ++ifun_;
    iback_ = ifun_ - 1;
    if (stp_ == 1.0)
      x_ = z_;
    else
      for (int i = 0; i < n_; ++i) x_[i] = stp_ * d_[i] + t_[i];
    state_ = S_FG_LNSRCH;
    return;
  }
  if (!converged) {
    // failed line search: restore the last accepted point
    x_ = t_;
    g_ = r_;
    f_ = fold_;
    if (S_.empty()) {
      ++iter_;
      finish("ABNORMAL_TERMINATION_IN_LNSRCH");
      return;
    }
    S_.clear();
    Y_.clear();
    theta_ = 1.0;
    iterate();
    return;
  }
  // new iterate
  ++iter_;
  after_new_x();
}

void LBFGSB::after_new_x() {
  // the driver stops at the iteration limit before the convergence tests
  if (iter_ >= max_iter) {
    finish("STOP: TOTAL NO. of ITERATIONS REACHED LIMIT");
    return;
  }
  if (proj_grad_norm() <= pgtol) {
    finish("CONVERGENCE: NORM_OF_PROJECTED_GRADIENT_<=_PGTOL");
    return;
  }
  const double ddum = std::max({std::fabs(fold_), std::fabs(f_), 1.0});
  if ((fold_ - f_) <= factr * EPSMCH * ddum) {
    finish("CONVERGENCE: REL_REDUCTION_OF_F_<=_FACTR*EPSMCH");
    return;
  }
  // correction pair
  std::vector<double> s(n_), y(n_);
  for (int i = 0; i < n_; ++i) {
    s[i] = stp_ * d_[i];
    y[i] = g_[i] - r_[i];
  }
  const double rr = dot(y, y);
  double dr, dd;
  if (stp_ == 1.0) {
    dr = gd_ - gdold_;
    dd = -gdold_;
  } else {
    dr = (gd_ - gdold_) * stp_;
    dd = -gdold_ * stp_;
  }
  if (dr > EPSMCH * dd) {
    if ((int)S_.size() == m_) {
      S_.erase(S_.begin());
      Y_.erase(Y_.begin());
    }
    S_.push_back(s);
    Y_.push_back(y);
    theta_ = rr / dr;
  }
  iterate();
}

// Moré-Thuente line search (MINPACK-2 dcsrch)
void LBFGSB::dcsrch(double f, double g, double& stp, double ftol, double gtol, double xtol, double stpmin,
                    double stpmax, Dcsrch& s) {
  constexpr double xtrapl = 1.1, xtrapu = 4.0;
  if (s.task == "START") {
    if (stp < stpmin) {
      s.task = "ERROR: STP .LT. STPMIN";
      return;
    }
    if (stp > stpmax) {
      s.task = "ERROR: STP .GT. STPMAX";
      return;
    }
    if (g >= 0.0) {
      s.task = "ERROR: INITIAL G .GE. ZERO";
      return;
    }
    s.brackt = false;
    s.stage = 1;
    s.finit = f;
    s.ginit = g;
    s.gtest = ftol * s.ginit;
    s.width = stpmax - stpmin;
    s.width1 = s.width / 0.5;
    s.stx = 0.0;
    s.fx = s.finit;
    s.gx = s.ginit;
    s.sty = 0.0;
    s.fy = s.finit;
    s.gy = s.ginit;
    s.stmin = 0.0;
    s.stmax = stp + xtrapu * stp;
    s.task = "FG";
    return;
  }
  const double ftest = s.finit + stp * s.gtest;
  if (s.stage == 1 && f <= ftest && g >= 0.0) s.stage = 2;
  if (s.brackt && (stp <= s.stmin || stp >= s.stmax)) s.task = "WARNING: ROUNDING ERRORS PREVENT PROGRESS";
  if (s.brackt && s.stmax - s.stmin <= xtol * s.stmax) s.task = "WARNING: XTOL TEST SATISFIED";
  if (stp == stpmax && f <= ftest && g <= s.gtest) s.task = "WARNING: STP = STPMAX";
  if (stp == stpmin && (f > ftest || g >= s.gtest)) s.task = "WARNING: STP = STPMIN";
  if (f <= ftest && std::fabs(g) <= gtol * (-s.ginit)) s.task = "CONVERGENCE";
  if (s.task.rfind("WARN", 0) == 0 || s.task.rfind("CONV", 0) == 0) return;
  if (s.stage == 1 && f <= s.fx && f > ftest) {
    double fm = f - stp * s.gtest;
    double fxm = s.fx - s.stx * s.gtest;
    double fym = s.fy - s.sty * s.gtest;
    double gm = g - s.gtest;
    double gxm = s.gx - s.gtest;
    double gym = s.gy - s.gtest;
    dcstep(s.stx, fxm, gxm, s.sty, fym, gym, stp, fm, gm, s.brackt, s.stmin, s.stmax);
    s.fx = fxm + s.stx * s.gtest;
    s.fy = fym + s.sty * s.gtest;
    s.gx = gxm + s.gtest;
    s.gy = gym + s.gtest;
  } else {
    dcstep(s.stx, s.fx, s.gx, s.sty, s.fy, s.gy, stp, f, g, s.brackt, s.stmin, s.stmax);
  }
  if (s.brackt) {
    if (std::fabs(s.sty - s.stx) >= 0.66 * s.width1) stp = s.stx + 0.5 * (s.sty - s.stx);
    s.width1 = s.width;
    s.width = std::fabs(s.sty - s.stx);
  }
  if (s.brackt) {
    s.stmin = std::min(s.stx, s.sty);
    s.stmax = std::max(s.stx, s.sty);
  } else {
    s.stmin = stp + xtrapl * (stp - s.stx);
    s.stmax = stp + xtrapu * (stp - s.stx);
  }
  stp = std::max(stp, stpmin);
  stp = std::min(stp, stpmax);
  if ((s.brackt && (stp <= s.stmin || stp >= s.stmax)) || (s.brackt && s.stmax - s.stmin <= xtol * s.stmax))
    stp = s.stx;
  s.task = "FG";
}

// safeguarded cubic / quadratic step (MINPACK-2 dcstep)
void LBFGSB::dcstep(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy, double& stp,
                    double fp, double dp, bool& brackt, double stpmin, double stpmax) {
  const double sgnd = dp * (dx / std::fabs(dx));
  double stpf, stpc, stpq, theta, s, gamma, p, q, r;
  if (fp > fx) {
    theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
    gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
    if (stp < stx) gamma = -gamma;
    p = (gamma - dx) + theta;
    q = ((gamma - dx) + gamma) + dp;
    r = p / q;
    stpc = stx + r * (stp - stx);
    stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
    if (std::fabs(stpc - stx) < std::fabs(stpq - stx))
      stpf = stpc;
    else
      stpf = stpc + (stpq - stpc) / 2.0;
    brackt = true;
  } else if (sgnd < 0.0) {
    theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
    gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
    if (stp > stx) gamma = -gamma;
    p = (gamma - dp) + theta;
    q = ((gamma - dp) + gamma) + dx;
    r = p / q;
    stpc = stp + r * (stx - stp);
    stpq = stp + (dp / (dp - dx)) * (stx - stp);
    if (std::fabs(stpc - stp) > std::fabs(stpq - stp))
      stpf = stpc;
    else
      stpf = stpq;
    brackt = true;
  } else if (std::fabs(dp) < std::fabs(dx)) {
    theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
    s = std::max({std::fabs(theta), std::fabs(dx), std::fabs(dp)});
    gamma = s * std::sqrt(std::max(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
    if (stp > stx) gamma = -gamma;
    p = (gamma - dp) + theta;
    q = (gamma + (dx - dp)) + gamma;
    r = p / q;
    if (r < 0.0 && gamma != 0.0)
      stpc = stp + r * (stx - stp);
    else if (stp > stx)
      stpc = stpmax;
    else
      stpc = stpmin;
    stpq = stp + (dp / (dp - dx)) * (stx - stp);
    if (brackt) {
      if (std::fabs(stpc - stp) < std::fabs(stpq - stp))
        stpf = stpc;
      else
        stpf = stpq;
      if (stp > stx)
        stpf = std::min(stp + 0.66 * (sty - stp), stpf);
      else
        stpf = std::max(stp + 0.66 * (sty - stp), stpf);
    } else {
      if (std::fabs(stpc - stp) > std::fabs(stpq - stp))
        stpf = stpc;
      else
        stpf = stpq;
      stpf = std::min(stpmax, stpf);
      stpf = std::max(stpmin, stpf);
    }
  } else {
    if (brackt) {
      theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
      s = std::max({std::fabs(theta), std::fabs(dy), std::fabs(dp)});
      gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
      if (stp > sty) gamma = -gamma;
      p = (gamma - dp) + theta;
      q = ((gamma - dp) + gamma) + dy;
      r = p / q;
      stpc = stp + r * (sty - stp);
      stpf = stpc;
    } else if (stp > stx) {
      stpf = stpmax;
    } else {
      stpf = stpmin;
    }
  }
  if (fp > fx) {
    sty = stp;
    fy = fp;
    dy = dp;
  } else {
    if (sgnd < 0.0) {
      sty = stx;
      fy = fx;
      dy = dx;
    }
    stx = stp;
    fx = fp;
    dx = dp;
  }
  stp = stpf;
}

}  // namespace lk
