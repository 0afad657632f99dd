#include "dlr.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "linalg.hpp"

namespace xca::dlr {

double kernel(double t, double om) {
  if (t >= 0) {
    if (om >= 0) return -std::exp(-t * om) / (1.0 + std::exp(-om));
    return -std::exp((1.0 - t) * om) / (1.0 + std::exp(om));
  }
  // t < 0 encodes tau/beta = 1 + t; -t = 1 - tau/beta is exact
  if (om >= 0) return -std::exp(-(1.0 + t) * om) / (1.0 + std::exp(-om));
  return -std::exp(-t * om) / (1.0 + std::exp(om));
}

namespace {

constexpr int kChebPerPanel = 24;

// Chebyshev points of the first kind mapped to [a, b], ascending.
void cheb_panel(double a, double b, std::vector<double> &out) {
  const int p = kChebPerPanel;
  for (int j = p - 1; j >= 0; --j) {
    double x = std::cos(M_PI * (2.0 * j + 1.0) / (2.0 * p));
    out.push_back(0.5 * (a + b) + 0.5 * (b - a) * x);
  }
}

int n_tau_panels(double lambda) { return std::max((int)std::ceil(std::log2(lambda)) - 2, 1); }
int n_om_panels(double lambda) { return std::max((int)std::ceil(std::log2(lambda)), 1); }

// Fine composite Chebyshev grid on t in [0, 1/2], panels refined dyadically towards t = 0,
// mirrored to the second half in relative format (negative values).
std::vector<double> fine_tau(double lambda) {
  int npt = n_tau_panels(lambda);
  std::vector<double> half;
  double a = 0.0;
  for (int k = 0; k < npt; ++k) {
    double b = std::ldexp(1.0, k - npt);  // 2^{k-npt}, last panel ends at 1/2
    cheb_panel(a, b, half);
    a = b;
  }
  std::vector<double> t(half);
  for (double x : half) t.push_back(-x);
  return t;
}

// Fine composite Chebyshev grid on om in [-lambda, lambda], panels refined dyadically towards 0.
std::vector<double> fine_om(double lambda) {
  int npo = n_om_panels(lambda);
  std::vector<double> pos;
  double a = 0.0;
  for (int k = 0; k < npo; ++k) {
    double b = lambda * std::ldexp(1.0, k - npo + 1);
    cheb_panel(a, b, pos);
    a = b;
  }
  std::vector<double> om;
  for (auto it = pos.rbegin(); it != pos.rend(); ++it) om.push_back(-*it);
  for (double x : pos) om.push_back(x);
  return om;
}

} // namespace

std::vector<double> build_rf(double lambda, double eps) {
  if (!(lambda > 0) || !(eps > 0)) throw std::invalid_argument("build_rf: lambda and eps must be positive");
  auto t = fine_tau(lambda);
  auto om = fine_om(lambda);
  int nt = (int)t.size(), no = (int)om.size();
  // rows = frequencies (columns of the kernel matrix), so the row Gram-Schmidt picks frequencies
  std::vector<double> A((size_t)no * nt);
  for (int j = 0; j < no; ++j)
    for (int i = 0; i < nt; ++i) A[(size_t)j * nt + i] = kernel(t[i], om[j]);
  auto sel = la::pivoted_gram_schmidt_rows(A, no, nt, eps, 0);
  std::vector<double> rf;
  for (int j : sel) rf.push_back(om[j]);
  std::sort(rf.begin(), rf.end());
  return rf;
}

std::vector<double> build_it(double lambda, const std::vector<double> &rf) {
  auto t = fine_tau(lambda);
  int nt = (int)t.size(), r = (int)rf.size();
  std::vector<double> A((size_t)nt * r);
  for (int i = 0; i < nt; ++i)
    for (int k = 0; k < r; ++k) A[(size_t)i * r + k] = kernel(t[i], rf[k]);
  auto sel = la::pivoted_gram_schmidt_rows(A, nt, r, 0.0, r);
  if ((int)sel.size() != r) throw std::runtime_error("build_it: rank deficient kernel matrix");
  std::vector<double> it;
  for (int i : sel) it.push_back(t[i]);
  std::sort(it.begin(), it.end(), [](double a, double b) { return rel2abs(a) < rel2abs(b); });
  return it;
}

void Ops::init(const std::vector<double> &it_nodes, const std::vector<double> &rf_nodes) {
  it = it_nodes;
  rf = rf_nodes;
  r = (int)rf.size();
  if ((int)it.size() != r) throw std::invalid_argument("Ops::init: it and rf node counts differ");
  size_t rr = (size_t)r * r;
  cf2it.assign(rr, 0.0);
  for (int i = 0; i < r; ++i)
    for (int k = 0; k < r; ++k) cf2it[(size_t)i * r + k] = kernel(it[i], rf[k]);
  lu = cf2it;
  la::getrf(lu.data(), r, piv);
  it2cf.assign(rr, 0.0);
  for (int i = 0; i < r; ++i) it2cf[(size_t)i * r + i] = 1.0;
  la::getrs(lu.data(), r, piv, it2cf.data(), r);
  k0.resize(r);
  for (int k = 0; k < r; ++k) k0[k] = kernel(0.0, rf[k]);
  W.assign(rr, 0.0);
  for (int j = 0; j < r; ++j)
    for (int k = 0; k < r; ++k)
      if (j != k) W[(size_t)j * r + k] = k0[k] / (rf[j] - rf[k]);
  tcf2it.assign(rr, 0.0);
  for (int i = 0; i < r; ++i)
    for (int k = 0; k < r; ++k) tcf2it[(size_t)i * r + k] = rel2abs(it[i]) * cf2it[(size_t)i * r + k] * k0[k];
  // reflection: values at -t_i from values at t_i, via coefficients
  std::vector<double> kr(rr);
  for (int i = 0; i < r; ++i)
    for (int k = 0; k < r; ++k) kr[(size_t)i * r + k] = kernel(-it[i], rf[k]);
  refl.assign(rr, 0.0);
  la::matmul(kr.data(), it2cf.data(), refl.data(), r, r, r);
}

void Ops::vals2coefs(double *g, int m) const { la::getrs(lu.data(), r, piv, g, m); }

} // namespace xca::dlr
