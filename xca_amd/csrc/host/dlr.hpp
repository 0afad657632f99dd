// Discrete Lehmann representation (DLR) in imaginary time, written from the published method
// (Kaye, Chen, Parcollet, PRB 105, 235115 (2022); Kaye, Strand, Wentzell, JOSS 9, 6297 (2024);
// Kaye, Strand, Adv. Comput. Math. 49, 63 (2023)). It replaces the cppdlr calls made by
// c++/triqs_xca/block_sparse_backbone.cpp:68-70,177 and hyb.cpp:20-41.
//
// Conventions (fixed by the reference):
//  * imaginary time t = tau/beta in "relative format": t in [0, 1/2] is stored as t, t in (1/2, 1]
//    is stored as t - 1 (negative), so that 1 - t is exact near tau = beta;
//  * fermionic kernel K(t, om) = -exp(-om t) / (1 + exp(-om)) with dimensionless om = beta * energy;
//  * data arrays are (r, ...) with the node index slowest (nda C order).
#pragma once
#include <vector>

namespace xca::dlr {

// K(t, om) in relative time format, evaluated without overflow for any sign of t and om.
double kernel(double t, double om);
inline double rel2abs(double t) { return t < 0 ? 1.0 + t : t; }

// Real-frequency DLR nodes (dimensionless, sorted ascending) for cutoff lambda and tolerance eps.
std::vector<double> build_rf(double lambda, double eps);
// Imaginary-time DLR nodes (relative format, sorted by absolute time) for the given rf nodes.
std::vector<double> build_it(double lambda, const std::vector<double> &rf);

// Operator matrices on one DLR grid, all r x r row-major.
struct Ops {
  int r = 0;
  std::vector<double> it, rf;  // nodes
  std::vector<double> cf2it;   // K(t_i, om_k)
  std::vector<double> lu;      // LU factors of cf2it (partial pivoting)
  std::vector<int> piv;
  std::vector<double> it2cf;   // explicit inverse of cf2it (columns from LU solves)
  std::vector<double> k0;      // K(0, om_k)
  std::vector<double> W;       // W(j,k) = k0_k / (om_j - om_k), zero diagonal
  std::vector<double> tcf2it;  // tau_i K(t_i, om_k) K(0, om_k)
  std::vector<double> refl;    // K(-t_i, om_k) applied after vals2coefs: values at reflected nodes
  void init(const std::vector<double> &it_nodes, const std::vector<double> &rf_nodes);
  // g: (r, m) values -> coefficients, LU solve in place
  void vals2coefs(double *g, int m) const;
};

} // namespace xca::dlr
