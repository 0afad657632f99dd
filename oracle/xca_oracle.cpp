// ORACLE — test infrastructure only. This is a CPU restatement of the reference's pseudo-particle
// self-energy path, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker. Nothing in the product (xca_amd/) may link, load or call it.
//
// It follows, function by function (citations are into the reference tree):
//   Backbone                      c++/triqs_xca/backbone.cpp:19-237
//   Hybridization                 c++/triqs_xca/hyb.cpp:15-113
//   BlockOpSymQuartet / bars      c++/triqs_xca/block_sparse.cpp:443-577
//   BlockDiagOpFun                c++/triqs_xca/block_sparse.cpp:21-147
//   DiagramEvaluator (sigma)      c++/triqs_xca/block_sparse_backbone.cpp:127-459
//   CorrelatorBackbone            c++/triqs_xca/backbone.cpp:239-292
//   DiagramEvaluator (correlator) c++/triqs_xca/block_sparse_backbone.cpp:459-706,732-902
//   symmetry-set labelling        c++/triqs_xca/atom_diag_utils.cpp:42-68,150-185
//   dyson_it_ppsc::solve          c++/triqs_xca/dlr_dyson_ppsc.hpp:92-168
// The DLR operations (vals2coefs by LU, time-ordered convolution, reflection, evaluation,
// convolution matrix) are cppdlr's, which is third-party and not in the reference tree; they are
// restated from the published method (Kaye, Strand, Adv. Comput. Math. 49, 63 (2023)):
//   h(t) = beta * sum_{k,l} f_k g_l I_kl(t),  I_kl = [K_l(t)K_k(0) - K_k(t)K_l(0)]/(om_k - om_l),
//   I_kk = t K_k(t) K_k(0),  giving  h = beta * [cf2it c + tcf2it (f o g)],
//   c_j = -[(W f)_j g_j + f_j (W g)_j],  W(j,k) = K_k(0)/(om_j - om_k).
//
// Parity pins (tests/): closed forms of the one- and two-fermion models and the n-fermion H = 0
// model, DLR closed forms (free-G interpolation, convolution of exponentials, Volterra tau), the
// reference's Backbone indexing known answers, the reference's hard-coded Fock-space matrices,
// and the dense == block-sparse and f_ix-decomposition invariants.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// ORC_REAL selects the arithmetic: double (liboracle.so) or long double (liboracle_ld.so, the
// extended-precision build used to pin ill-conditioned workloads). Inputs and outputs of the C
// functions below stay double / complex<double> in both builds.
#ifndef ORC_REAL
#define ORC_REAL double
#endif
using real = ORC_REAL;
using cplx = std::complex<real>;
using ioc = std::complex<double>;

inline real conj_if(real x) { return x; }
inline cplx conj_if(cplx x) { return std::conj(x); }
inline real absval(real x) { return std::fabs(x); }
inline real absval(cplx x) { return std::abs(x); }

// ---------------------------------------------------------------- DLR (cppdlr restated)

// fermionic kernel in relative time format (see the kernel sign pinned by the one-fermion NCA
// known answer, test_block_sparse_backbone_eval.cpp:142-172)
real k_it(real t, real om) {
  if (t >= 0) {
    if (om >= 0) return -std::exp(-t * om) / (1.0 + std::exp(-om));
    return -std::exp((1.0 - t) * om) / (1.0 + std::exp(om));
  }
  if (om >= 0) return -std::exp(-(1.0 + t) * om) / (1.0 + std::exp(-om));
  return -std::exp(-t * om) / (1.0 + std::exp(om));
}
inline real rel2abs(real t) { return t < 0 ? 1.0 + t : t; }

struct ImTimeOps {
  int r = 0;
  std::vector<real> it, rf, cf2it, lu, W, tcf2it, k0;
  std::vector<int> piv;

  template <class T> ImTimeOps(const T *it_, const T *rf_, int r_) : r(r_), it(it_, it_ + r_), rf(rf_, rf_ + r_) {
    cf2it.resize((size_t)r * r);
    for (int i = 0; i < r; ++i)
      for (int k = 0; k < r; ++k) cf2it[(size_t)i * r + k] = k_it(it[i], rf[k]);
    lu = cf2it;
    piv.assign(r, 0);
    for (int k = 0; k < r; ++k) {  // LU with partial pivoting
      int p = k;
      for (int i = k + 1; i < r; ++i)
        if (std::fabs(lu[(size_t)i * r + k]) > std::fabs(lu[(size_t)p * r + k])) p = i;
      piv[k] = p;
      if (lu[(size_t)p * r + k] == 0.0) throw std::runtime_error("ImTimeOps: singular cf2it");
      if (p != k)
        for (int j = 0; j < r; ++j) std::swap(lu[(size_t)k * r + j], lu[(size_t)p * r + j]);
      for (int i = k + 1; i < r; ++i) {
        real l = lu[(size_t)i * r + k] / lu[(size_t)k * r + k];
        lu[(size_t)i * r + k] = l;
        for (int j = k + 1; j < r; ++j) lu[(size_t)i * r + j] -= l * lu[(size_t)k * r + j];
      }
    }
    k0.resize(r);
    for (int k = 0; k < r; ++k) k0[k] = k_it(0.0, rf[k]);
    W.assign((size_t)r * r, 0.0);
    for (int j = 0; j < r; ++j)
      for (int k = 0; k < r; ++k)
        if (j != k) W[(size_t)j * r + k] = k0[k] / (rf[j] - rf[k]);
    tcf2it.resize((size_t)r * r);
    for (int i = 0; i < r; ++i)
      for (int k = 0; k < r; ++k) tcf2it[(size_t)i * r + k] = rel2abs(it[i]) * cf2it[(size_t)i * r + k] * k0[k];
  }

  // in place: g (r, m) values -> coefficients (LU solve, as cppdlr's vals2coefs)
  template <class S> void vals2coefs(S *g, int m) const {
    for (int k = 0; k < r; ++k)
      if (piv[k] != k)
        for (int j = 0; j < m; ++j) std::swap(g[(size_t)k * m + j], g[(size_t)piv[k] * m + j]);
    for (int i = 1; i < r; ++i)
      for (int k = 0; k < i; ++k) {
        real l = lu[(size_t)i * r + k];
        for (int j = 0; j < m; ++j) g[(size_t)i * m + j] -= l * g[(size_t)k * m + j];
      }
    for (int i = r - 1; i >= 0; --i) {
      for (int k = i + 1; k < r; ++k) {
        real u = lu[(size_t)i * r + k];
        for (int j = 0; j < m; ++j) g[(size_t)i * m + j] -= u * g[(size_t)k * m + j];
      }
      for (int j = 0; j < m; ++j) g[(size_t)i * m + j] /= lu[(size_t)i * r + i];
    }
  }

  // in place: solve cf2it^T y = b for one vector (LAPACK getrs with transposition)
  template <class S> void solve_transposed(S *y) const {
    for (int i = 0; i < r; ++i) {  // U^T z = b
      for (int k = 0; k < i; ++k) y[i] -= lu[(size_t)k * r + i] * y[k];
      y[i] /= lu[(size_t)i * r + i];
    }
    for (int i = r - 1; i >= 0; --i)  // L^T w = z (unit diagonal)
      for (int k = i + 1; k < r; ++k) y[i] -= lu[(size_t)k * r + i] * y[k];
    for (int k = r - 1; k >= 0; --k)  // undo the row interchanges in reverse order
      if (piv[k] != k) std::swap(y[k], y[piv[k]]);
  }

  // out (r, m) = M (r x r) * in (r, m)
  template <class S> void apply(const std::vector<real> &M, const S *in, S *out, int m) const {
    for (int i = 0; i < r; ++i) {
      S *o = out + (size_t)i * m;
      for (int j = 0; j < m; ++j) o[j] = S(0);
      for (int k = 0; k < r; ++k) {
        real a = M[(size_t)i * r + k];
        const S *x = in + (size_t)k * m;
        for (int j = 0; j < m; ++j) o[j] += a * x[j];
      }
    }
  }

  // time-ordered convolution h(t) = int_0^t f(t - t') g(t') dt' at the nodes, from coefficient
  // arrays fc (r, a, b) and gc (r, b, c); returns values (r, a, c)
  template <class S> std::vector<S> convolve_to(real beta, const std::vector<S> &fc, const std::vector<S> &gc, int a, int b, int c) const {
    std::vector<S> Wf((size_t)r * a * b), Wg((size_t)r * b * c), cc((size_t)r * a * c), dd((size_t)r * a * c);
    apply(W, fc.data(), Wf.data(), a * b);
    apply(W, gc.data(), Wg.data(), b * c);
    for (int k = 0; k < r; ++k) {
      const S *f = &fc[(size_t)k * a * b], *wf = &Wf[(size_t)k * a * b];
      const S *g = &gc[(size_t)k * b * c], *wg = &Wg[(size_t)k * b * c];
      S *ck = &cc[(size_t)k * a * c], *dk = &dd[(size_t)k * a * c];
      for (int i = 0; i < a; ++i)
        for (int j = 0; j < c; ++j) {
          S s1(0), s2(0);
          for (int l = 0; l < b; ++l) {
            s1 += wf[i * b + l] * g[l * c + j] + f[i * b + l] * wg[l * c + j];
            s2 += f[i * b + l] * g[l * c + j];
          }
          ck[i * c + j] = -s1;
          dk[i * c + j] = s2;
        }
    }
    std::vector<S> h1((size_t)r * a * c), h2((size_t)r * a * c);
    apply(cf2it, cc.data(), h1.data(), a * c);
    apply(tcf2it, dd.data(), h2.data(), a * c);
    for (size_t x = 0; x < h1.size(); ++x) h1[x] = beta * (h1[x] + h2[x]);
    return h1;
  }

  // values at the reflected nodes -t_i (cppdlr reflect)
  template <class S> std::vector<S> reflect(const std::vector<S> &vals, int m) const {
    std::vector<S> c(vals);
    vals2coefs(c.data(), m);
    std::vector<S> out((size_t)r * m, S(0));
    for (int i = 0; i < r; ++i)
      for (int k = 0; k < r; ++k) {
        real kk = k_it(-it[i], rf[k]);
        for (int j = 0; j < m; ++j) out[(size_t)i * m + j] += kk * c[(size_t)k * m + j];
      }
    return out;
  }

  // evaluate from coefficients at relative time t
  template <class S> std::vector<S> coefs2eval(const std::vector<S> &c, int m, real t) const {
    std::vector<S> out(m, S(0));
    for (int k = 0; k < r; ++k) {
      real kk = k_it(t, rf[k]);
      for (int j = 0; j < m; ++j) out[j] += kk * c[(size_t)k * m + j];
    }
    return out;
  }
};

// ---------------------------------------------------------------- containers

template <class S> struct BlockDiagOpFun {  // block_sparse.cpp:21-147
  int r = 0;
  std::vector<int> dims;
  std::vector<std::vector<S>> blocks;  // (r, b, b)
  std::vector<int> zero;               // -1 if block is zero
  // diagnostic, not part of the restated algorithm: per block, the sum of max |block| over the
  // contents that add_block discarded by replacing a block flagged zero (a running sum that fell
  // below 1e-16 between two flat indices of a list)
  std::vector<real> dropped;
  void set_zero_block_indices() {
    for (size_t i = 0; i < blocks.size(); ++i) {
      real mx = 0;
      for (auto &x : blocks[i]) mx = std::max(mx, absval(x));
      zero[i] = (mx < 1e-16) ? -1 : 0;
    }
  }
  void add_block(int i, const std::vector<S> &blk) {
    if (blk.size() != blocks[i].size()) {
      std::ostringstream msg;
      msg << "BlockDiagOpFun::add_block: block " << i << " has shape (" << r << ", " << dims[i] << ", " << dims[i]
          << ") but the contribution has a different shape";
      throw std::invalid_argument(msg.str());
    }
    if (zero[i] == -1) {
      real mx = 0;
      for (auto &x : blocks[i]) mx = std::max(mx, absval(x));
      if (dropped.size() == blocks.size()) dropped[i] += mx;
      blocks[i] = blk;
    } else
      for (size_t x = 0; x < blk.size(); ++x) blocks[i][x] += blk[x];
    zero[i] = 0;
  }
};

template <class S> struct BlockOpSymSet {  // (q, b_out, b_in) per block column
  std::vector<int> block_indices;
  std::vector<std::vector<S>> blocks;
  std::vector<int> bout, bin;
  int q = 0;
  int get_block_index(int i) const { return block_indices[i]; }
  int get_block_size(int i, int dim) const { return block_indices[i] == -1 ? -1 : (dim == 0 ? bout[i] : bin[i]); }
};

template <class S> struct BlockOpSymSetBar {  // (q, p, b_out, b_in)
  std::vector<int> block_indices;
  std::vector<std::vector<S>> blocks;
  std::vector<int> bout, bin;
  int q = 0, p = 0;
};

// one block operator: target block per source block (-1: none) and its (b_out, b_in) matrices
template <class S> struct BlockOp {
  std::vector<int> block_indices;
  std::vector<std::vector<S>> blocks;
  std::vector<int> bout, bin;
};

template <class S> struct BlockOpSymQuartet {  // block_sparse.cpp:496-577
  std::vector<BlockOpSymSet<S>> Fs, F_dags;
  std::vector<BlockOpSymSetBar<S>> F_dag_bars, F_bars_refl;
  std::vector<long> sym_set_labels, sym_set_inds, sym_set_sizes;
  std::vector<std::vector<long>> sym_set_to_orb;
  int p = 0;
};

// cross-set hybridization check, block_sparse.cpp:443-492 (message text kept)
void check_sym_set_block_diagonal(const std::vector<cplx> &hc, int p, int n, const std::vector<long> &labels, const std::string &who) {
  real max_abs = 0;
  for (auto &x : hc) max_abs = std::max(max_abs, std::abs(x));
  if (!std::isfinite(max_abs)) throw std::invalid_argument(who + ": hyb_coeffs contains a non-finite entry");
  real thresh = 1.0e-12 * max_abs, worst = 0;
  long n_bad = 0, n_cross = 0, wl = -1, wi = -1, wj = -1;
  for (int l = 0; l < p; ++l)
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        if (labels[i] == labels[j]) continue;
        n_cross++;
        real a = std::abs(hc[((size_t)l * n + i) * n + j]);
        if (a > thresh) {
          n_bad++;
          if (a > worst) { worst = a; wl = l; wi = i; wj = j; }
        }
      }
  if (n_bad > 0) {
    std::ostringstream msg;
    msg << std::scientific << who << ": hyb_coeffs couples different symmetry sets, which the symmetry-set storage cannot represent. "
        << n_bad << " of " << n_cross << " cross-set entries exceed the tolerance " << 1.0e-12 << " * max|hyb_coeffs| = " << thresh
        << "; the largest is |hyb_coeffs(" << wl << ", " << wi << ", " << wj << ")| = " << worst << " (relative " << worst / max_abs
        << "), coupling symmetry set " << labels[wi] << " to symmetry set " << labels[wj] << ".";
    throw std::invalid_argument(msg.str());
  }
}

// ---------------------------------------------------------------- Backbone (backbone.cpp:19-237)

struct Vertex {
  bool bar = false, dag = false;
  int hyb_ind = 0, Ksign = 0, orb = 0;
};

// correlator == true gives CorrelatorBackbone (backbone.cpp:239-292): no line on vertex 0, whose
// operator is a creator, the vertex paired with it an annihilator, 2^(m-1) direction patterns.
struct Backbone {
  int m, n, fb_ix_max, o_ix_max, prefactor_sign = 1;
  bool correlator = false;
  std::vector<int> topo;  // (m, 2)
  std::vector<int> prefactor_Ksigns, prefactor_Kexps, edges, pole_inds, orb_inds, fb;
  std::vector<Vertex> vertices;
  Backbone(const int *topology, int m_, int n_, bool corr = false)
      : m(m_), n(n_), correlator(corr), topo(topology, topology + 2 * m_) {
    fb_ix_max = corr ? 1 << (m - 1) : 1 << m;
    o_ix_max = (int)std::pow(n, m - 1);
    prefactor_Ksigns.assign(m - 1, 0);
    prefactor_Kexps.assign(m - 1, 0);
    vertices.assign(2 * m, Vertex());
    edges.assign((2 * m - 1) * std::max(m - 1, 1), 0);
    fb.assign(m, 0);
    pole_inds.assign(std::max(m - 1, 1), 0);
    orb_inds.assign(2 * m, 0);
  }
  int T(int i, int j) const { return topo[2 * i + j]; }
  int &edge(int e, int x) { return edges[e * std::max(m - 1, 1) + x]; }
  int get_edge(int e, int x) const { return edges[e * std::max(m - 1, 1) + x]; }

  void set_directions(const std::vector<int> &fbv) {
    fb = fbv;
    for (int i = 0; i < m; i++)
      if (T(i, 0) >= T(i, 1)) throw std::invalid_argument("first row of topology must contain smaller-numbered vertices");
    if (T(0, 0) != 0) throw std::invalid_argument("topology(0,0) must be 0");
    vertices[0].bar = false;
    vertices[T(0, 1)].bar = false;
    if (correlator) {  // backbone.cpp:253-265; fb[0] = -1 marks the missing line
      vertices[0].dag = true;
      vertices[T(0, 1)].dag = false;
    } else {
      vertices[0].dag = (fbv[0] != 1);
      vertices[T(0, 1)].dag = (fbv[0] == 1);
    }
    for (int i = 1; i < m; i++) {
      vertices[T(i, 0)].bar = false;
      vertices[T(i, 1)].bar = true;
      vertices[T(i, 0)].dag = (fbv[i] != 1);
      vertices[T(i, 1)].dag = (fbv[i] == 1);
    }
  }
  void set_directions(int fb_ix) {
    std::vector<int> v(m);
    if (correlator) {  // backbone.cpp:241-249: bits for lines 1..m-1
      v[0] = -1;
      for (int i = 1; i < m; i++) { v[i] = fb_ix % 2; fb_ix /= 2; }
    } else {
      for (int i = 0; i < m; i++) { v[i] = fb_ix % 2; fb_ix /= 2; }
    }
    set_directions(v);
  }
  void set_pole_inds(const std::vector<int> &pv, const std::vector<real> &poles) {
    pole_inds = pv;
    for (int i = 1; i < m; i++) {
      int a = T(i, 0), b = T(i, 1);
      real w = poles[pole_inds[i - 1]];
      if (fb[i] == 1) {
        if (w <= 0) {  // step 4(a)
          vertices[a].hyb_ind = i - 1; vertices[a].Ksign = -1;
          vertices[b].hyb_ind = i - 1; vertices[b].Ksign = 1;
          prefactor_Ksigns[i - 1] = -1; prefactor_Kexps[i - 1] = 1;
        } else {  // step 4(b)
          vertices[b].hyb_ind = i - 1;
          for (int j = a; j < b; j++) edge(j, i - 1) = 1;
          prefactor_Ksigns[i - 1] = 1; prefactor_Kexps[i - 1] = b - a - 1;
        }
      } else {
        if (w >= 0) {
          vertices[a].hyb_ind = i - 1; vertices[a].Ksign = 1;
          vertices[b].hyb_ind = i - 1; vertices[b].Ksign = -1;
          prefactor_sign *= -1;
          prefactor_Ksigns[i - 1] = 1; prefactor_Kexps[i - 1] = 1;
        } else {
          vertices[b].hyb_ind = i - 1;
          for (int j = a; j < b; j++) edge(j, i - 1) = -1;
          prefactor_sign *= -1;
          prefactor_Ksigns[i - 1] = -1; prefactor_Kexps[i - 1] = b - a - 1;
        }
      }
    }
  }
  void set_pole_inds(int p_ix, const std::vector<real> &poles) {
    int p = (int)poles.size();
    std::vector<int> v(std::max(m - 1, 1), 0);
    for (int i = 0; i < m - 1; i++) { v[i] = p_ix % p; p_ix /= p; }
    set_pole_inds(v, poles);
  }
  void set_orb_inds(const std::vector<int> &ov) {
    orb_inds = ov;
    for (int i = 0; i < 2 * m; i++)
      if (i != 0 && i != T(0, 1)) vertices[i].orb = ov[i];
  }
  void set_orb_inds(int o_ix) {
    std::vector<int> v(2 * m, 0);
    v[0] = -1;
    v[T(0, 1)] = -1;
    for (int i = 1; i < m; i++) {
      v[T(i, 0)] = o_ix % n;
      v[T(i, 1)] = o_ix % n;
      o_ix /= n;
    }
    set_orb_inds(v);
  }
  void set_flat_index(int f_ix, const std::vector<real> &poles) {
    int p = (int)poles.size();
    int o_ix = f_ix % o_ix_max;
    f_ix /= o_ix_max;
    int p_ix_max = (int)std::pow(p, m - 1);
    int p_ix = f_ix % p_ix_max;
    int fb_ix = f_ix / p_ix_max;
    set_directions(fb_ix);
    set_pole_inds(p_ix, poles);
    set_orb_inds(o_ix);
  }
  void reset_all_inds() {
    std::fill(fb.begin(), fb.end(), 0);
    for (auto &v : vertices) v = Vertex();
    std::fill(pole_inds.begin(), pole_inds.end(), 0);
    std::fill(edges.begin(), edges.end(), 0);
    prefactor_sign = 1;
    std::fill(prefactor_Ksigns.begin(), prefactor_Ksigns.end(), 0);
    std::fill(prefactor_Kexps.begin(), prefactor_Kexps.end(), 0);
  }
};

// ---------------------------------------------------------------- problem description

struct Problem {
  real beta;
  int r, p, n, nblocks;
  std::vector<real> it, rf, poles;  // poles in energy units
  std::vector<cplx> coeffs;           // (p, n, n)
  std::vector<int> dims;
  std::vector<int> c_conn, cdag_conn;             // (n, nblocks)
  std::vector<std::vector<cplx>> c_mats, cdag_mats;  // [o * nblocks + j] -> (b_out, b_in)
};

// ---------------------------------------------------------------- Hybridization (hyb.cpp:15-113)

template <class S> struct Hybridization {
  std::vector<real> poles;     // beta * hyb_poles
  std::vector<S> values, values_reflect;  // (r, n, n)
  std::vector<real> k_0_p, k_0_m, k_it_p, k_it_m;  // (p), (r, p)
  int r, p, n;
  Hybridization(const Problem &P, const ImTimeOps &ops) : r(P.r), p(P.p), n(P.n) {
    poles.resize(p);
    for (int l = 0; l < p; ++l) poles[l] = P.poles[l] * P.beta;
    // coefs2vals (hyb.cpp:101-113): values(t, i, j) = sum_l k_it(t_i, beta * pole_l) c(l, i, j)
    std::vector<cplx> vals((size_t)r * n * n, cplx(0));
    for (int t = 0; t < r; ++t)
      for (int l = 0; l < p; ++l) {
        real k = k_it(ops.it[t], P.beta * P.poles[l]);
        for (int x = 0; x < n * n; ++x) vals[(size_t)t * n * n + x] += k * P.coeffs[(size_t)l * n * n + x];
      }
    auto refl = ops.reflect(vals, n * n);
    values.resize(vals.size());
    values_reflect.resize(vals.size());
    for (size_t x = 0; x < vals.size(); ++x) {
      values[x] = to_s(vals[x]);
      values_reflect[x] = to_s(refl[x]);
    }
    k_0_p.resize(p); k_0_m.resize(p);
    k_it_p.resize((size_t)r * p); k_it_m.resize((size_t)r * p);
    for (int l = 0; l < p; ++l) {
      k_0_p[l] = k_it(0, +poles[l]);
      k_0_m[l] = k_it(0, -poles[l]);
      for (int t = 0; t < r; ++t) {
        k_it_p[(size_t)t * p + l] = k_it(ops.it[t], +poles[l]);
        k_it_m[(size_t)t * p + l] = k_it(ops.it[t], -poles[l]);
      }
    }
  }
  static S to_s(cplx x);
};
template <> real Hybridization<real>::to_s(cplx x) { return x.real(); }
template <> cplx Hybridization<cplx>::to_s(cplx x) { return x; }

// ---------------------------------------------------------------- DiagramEvaluator

template <class S> struct DiagramEvaluator {
  real beta;
  int r, n, q, Nmax;
  ImTimeOps ops;
  BlockOpSymQuartet<S> Fq;
  Hybridization<S> hyb;
  BlockDiagOpFun<S> Sigma;
  std::vector<int> dims;

  static S to_s(cplx x) { return Hybridization<S>::to_s(x); }

  DiagramEvaluator(const Problem &P)
      : beta(P.beta), r(P.r), n(P.n), ops(P.it.data(), P.rf.data(), P.r), hyb(P, ops), dims(P.dims) {
    build_quartet(P);
    q = (int)Fq.Fs.size();
    Nmax = *std::max_element(dims.begin(), dims.end());
  }

  // atom_diag_utils.cpp:42-68 (labels from identical c_connection rows) and :150-185 (blocks);
  // block_sparse.cpp:496-577 (barred operators)
  void build_quartet(const Problem &P) {
    int nb = P.nblocks;
    std::vector<long> labels(n, 0);
    int counter = 0;
    for (int o = 0; o < n; ++o) {
      bool found = false;
      for (int o2 = 0; o2 < o; ++o2) {
        bool same = true;
        for (int s = 0; s < nb; ++s)
          if (P.c_conn[o * nb + s] != P.c_conn[o2 * nb + s]) { same = false; break; }
        if (same) { labels[o] = labels[o2]; found = true; break; }
      }
      if (!found) labels[o] = counter++;
    }
    int nq = counter;
    Fq.sym_set_labels = labels;
    Fq.sym_set_inds.assign(n, 0);
    Fq.sym_set_sizes.assign(nq, 0);
    Fq.sym_set_to_orb.assign(nq, std::vector<long>(n, -1));
    for (int i = 0; i < n; ++i) {
      Fq.sym_set_inds[i] = Fq.sym_set_sizes[labels[i]];
      Fq.sym_set_sizes[labels[i]]++;
      Fq.sym_set_to_orb[labels[i]][Fq.sym_set_inds[i]] = i;
    }
    auto make_set = [&](const std::vector<int> &conn, const std::vector<std::vector<cplx>> &mats, int g) {
      BlockOpSymSet<S> F;
      F.q = (int)Fq.sym_set_sizes[g];
      F.block_indices.assign(nb, -1);
      F.blocks.resize(nb);
      F.bout.assign(nb, 0);
      F.bin.assign(nb, 0);
      int rep = (int)Fq.sym_set_to_orb[g][0];
      for (int s = 0; s < nb; ++s) {
        int t = conn[rep * nb + s];
        F.block_indices[s] = t;
        if (t < 0) continue;
        F.bout[s] = P.dims[t];
        F.bin[s] = P.dims[s];
        F.blocks[s].assign((size_t)F.q * P.dims[t] * P.dims[s], S(0));
        for (int k = 0; k < F.q; ++k) {
          int o = (int)Fq.sym_set_to_orb[g][k];
          if (conn[o * nb + s] != t) throw std::invalid_argument("symmetry set with inconsistent connections");
          auto &M = mats[(size_t)o * nb + s];
          for (size_t x = 0; x < M.size(); ++x) F.blocks[s][(size_t)k * P.dims[t] * P.dims[s] + x] = to_s(M[x]);
        }
      }
      return F;
    };
    // F_dag sets are labelled by the same groups (atom_diag_utils.cpp:162-175)
    for (int g = 0; g < nq; ++g) {
      Fq.Fs.push_back(make_set(P.c_conn, P.c_mats, g));
      Fq.F_dags.push_back(make_set(P.cdag_conn, P.cdag_mats, g));
    }
    check_sym_set_block_diagonal(P.coeffs, P.p, n, labels, "BlockOpSymQuartet");
    int p = P.p;
    Fq.p = p;
    for (int g = 0; g < nq; ++g) {
      auto &Fd = Fq.F_dags[g];
      auto &F = Fq.Fs[g];
      BlockOpSymSetBar<S> Bd, Bf;
      Bd.q = Fd.q; Bd.p = p; Bd.block_indices = Fd.block_indices; Bd.bout = Fd.bout; Bd.bin = Fd.bin;
      Bf.q = F.q; Bf.p = p; Bf.block_indices = F.block_indices; Bf.bout = F.bout; Bf.bin = F.bin;
      Bd.blocks.resize(nb);
      Bf.blocks.resize(nb);
      for (int b = 0; b < nb; ++b) {
        if (Fd.block_indices[b] != -1) Bd.blocks[b].assign((size_t)Fd.q * p * Fd.bout[b] * Fd.bin[b], S(0));
        if (F.block_indices[b] != -1) Bf.blocks[b].assign((size_t)F.q * p * F.bout[b] * F.bin[b], S(0));
      }
      for (int l = 0; l < p; l++)
        for (int lam = 0; lam < (int)Fq.sym_set_sizes[g]; lam++)
          for (int nu = 0; nu < (int)Fq.sym_set_sizes[g]; nu++) {
            long lam_orb = Fq.sym_set_to_orb[g][lam], nu_orb = Fq.sym_set_to_orb[g][nu];
            S c = to_s(P.coeffs[((size_t)l * n + nu_orb) * n + lam_orb]);
            for (int b = 0; b < nb; b++) {
              if (Fd.block_indices[b] != -1) {
                size_t sz = (size_t)Fd.bout[b] * Fd.bin[b];
                S *dst = &Bd.blocks[b][((size_t)lam * p + l) * sz];
                const S *src = &Fd.blocks[b][(size_t)nu * sz];
                for (size_t x = 0; x < sz; ++x) dst[x] += c * src[x];
              }
              if (F.block_indices[b] != -1) {
                size_t sz = (size_t)F.bout[b] * F.bin[b];
                S *dst = &Bf.blocks[b][((size_t)nu * p + l) * sz];
                const S *src = &F.blocks[b][(size_t)lam * sz];
                for (size_t x = 0; x < sz; ++x) dst[x] += -c * src[x];
              }
            }
          }
      Fq.F_dag_bars.push_back(Bd);
      Fq.F_bars_refl.push_back(Bf);
    }
  }

  // multiply_left_vertex (block_sparse_backbone.cpp:127-163); T (r, bout x cols) <- K F T
  void multiply_left_vertex(std::vector<S> &T, int &rows, int cols, Backbone &bb, int v, const std::vector<int> &ind_path,
                            const std::vector<int> &block_dims) {
    int o = bb.vertices[v].orb;
    int qi = (int)Fq.sym_set_labels[o], qo = (int)Fq.sym_set_inds[o];
    int l = bb.pole_inds[bb.vertices[v].hyb_ind];
    int b = ind_path[v - 1];
    int bin = block_dims[v], bout = block_dims[v + 1];
    const S *F;
    if (bb.vertices[v].bar) {
      auto &B = bb.vertices[v].dag ? Fq.F_dag_bars[qi] : Fq.F_bars_refl[qi];
      F = &B.blocks[b][((size_t)qo * B.p + l) * bout * bin];
    } else {
      auto &B = bb.vertices[v].dag ? Fq.F_dags[qi] : Fq.Fs[qi];
      F = &B.blocks[b][(size_t)qo * bout * bin];
    }
    if (bin != rows) throw std::logic_error("vertex shape");
    std::vector<S> out((size_t)r * bout * cols, S(0));
    for (int t = 0; t < r; ++t)
      for (int i = 0; i < bout; ++i)
        for (int j = 0; j < cols; ++j) {
          S s(0);
          for (int k = 0; k < bin; ++k) s += F[i * bin + k] * T[((size_t)t * bin + k) * cols + j];
          out[((size_t)t * bout + i) * cols + j] = s;
        }
    int Ks = bb.vertices[v].Ksign;  // hyb.cpp:46-57
    if (Ks != 0) {
      auto &kt = (Ks > 0) ? hyb.k_it_p : hyb.k_it_m;
      for (int t = 0; t < r; ++t)
        for (int x = 0; x < bout * cols; ++x) out[(size_t)t * bout * cols + x] *= kt[(size_t)t * hyb.p + l];
    }
    T.swap(out);
    rows = bout;
  }

  // integrate_left_edge (block_sparse_backbone.cpp:165-178)
  void integrate_left_edge(std::vector<S> &T, int rows, int cols, const BlockDiagOpFun<S> &Gt, Backbone &bb, int e,
                           const std::vector<int> &ind_path) {
    int b = ind_path[e];
    int d = Gt.dims[b];
    std::vector<S> GK = Gt.blocks[b];
    for (int x = 0; x < bb.m - 1; x++) {  // hyb.cpp:59-73
      int Ks = bb.get_edge(e, x);
      if (Ks != 0) {
        int l = bb.pole_inds[x];
        auto &kt = (Ks > 0) ? hyb.k_it_p : hyb.k_it_m;
        for (int t = 0; t < r; ++t)
          for (int y = 0; y < d * d; ++y) GK[(size_t)t * d * d + y] *= kt[(size_t)t * hyb.p + l];
      }
    }
    if (d != rows) throw std::logic_error("edge shape");
    ops.vals2coefs(GK.data(), d * d);
    ops.vals2coefs(T.data(), rows * cols);
    T = ops.convolve_to(beta, GK, T, d, d, cols);
  }

  // multiply_left_vertex_and_right_zero_vertex (block_sparse_backbone.cpp:187-227)
  void zero_vertex(std::vector<S> &T, int &rows, Backbone &bb, bool is_forward, int b_ix_0, int p_kap, int p_mu,
                   const std::vector<int> &ind_path, const std::vector<int> &block_dims) {
    int v = bb.T(0, 1);
    int b_mu = ind_path[v - 1];
    int bv = block_dims[v], b1 = block_dims[1], b0 = block_dims[0], bvp = block_dims[v + 1];
    auto &Fk = is_forward ? Fq.Fs[p_kap] : Fq.F_dags[p_kap];
    auto &Fm = (!is_forward) ? Fq.Fs[p_mu] : Fq.F_dags[p_mu];
    int qk = (int)Fq.sym_set_sizes[p_kap], qm = (int)Fq.sym_set_sizes[p_mu];
    std::vector<S> Tk((size_t)qk * r * bv * b0, S(0));
    for (int k = 0; k < qk; ++k) {
      const S *F = &Fk.blocks[b_ix_0][(size_t)k * b1 * b0];
      for (int t = 0; t < r; ++t)
        for (int i = 0; i < bv; ++i)
          for (int j = 0; j < b0; ++j) {
            S s(0);
            for (int a = 0; a < b1; ++a) s += T[((size_t)t * bv + i) * b1 + a] * F[a * b0 + j];
            Tk[(((size_t)k * r + t) * bv + i) * b0 + j] = s;
          }
    }
    std::vector<S> out((size_t)r * bvp * b0, S(0));
    real hyb_sign = is_forward ? 1.0 : -1.0;
    std::vector<S> Tmu((size_t)r * bv * b0);
    for (int mu = 0; mu < qm; ++mu) {
      std::fill(Tmu.begin(), Tmu.end(), S(0));
      for (int k = 0; k < qk; ++k) {
        long om = Fq.sym_set_to_orb[p_mu][mu], ok = Fq.sym_set_to_orb[p_kap][k];
        for (int t = 0; t < r; ++t) {
          S h = is_forward ? hyb.values[((size_t)t * n + om) * n + ok] : hyb.values_reflect[((size_t)t * n + ok) * n + om];
          S hs = hyb_sign * h;
          for (int x = 0; x < bv * b0; ++x) Tmu[(size_t)t * bv * b0 + x] += hs * Tk[(((size_t)k * r + t) * bv) * b0 + x];
        }
      }
      const S *F = &Fm.blocks[b_mu][(size_t)mu * bvp * bv];
      for (int t = 0; t < r; ++t)
        for (int i = 0; i < bvp; ++i)
          for (int j = 0; j < b0; ++j) {
            S s(0);
            for (int a = 0; a < bv; ++a) s += F[i * bv + a] * Tmu[((size_t)t * bv + a) * b0 + j];
            out[((size_t)t * bvp + i) * b0 + j] += s;
          }
    }
    T.swap(out);
    rows = bvp;
  }

  // eval_self_energy_fixed_indices (block_sparse_backbone.cpp:339-367)
  void eval_fixed(const BlockDiagOpFun<S> &Gt, Backbone &bb, int b_ix, int p_kap, int p_mu, const std::vector<int> &ind_path,
                  const std::vector<int> &block_dims) {
    int m = bb.m, vct0 = bb.T(0, 1);
    std::vector<S> T = Gt.blocks[ind_path[0]];
    int rows = block_dims[1], cols = block_dims[1];
    for (int v = 1; v < vct0; v++) {
      multiply_left_vertex(T, rows, cols, bb, v, ind_path, block_dims);
      integrate_left_edge(T, rows, cols, Gt, bb, v, ind_path);
    }
    zero_vertex(T, rows, bb, !bb.vertices[0].dag, b_ix, p_kap, p_mu, ind_path, block_dims);
    cols = block_dims[0];
    for (int v = vct0 + 1; v < 2 * m; v++) {
      integrate_left_edge(T, rows, cols, Gt, bb, v - 1, ind_path);
      multiply_left_vertex(T, rows, cols, bb, v, ind_path, block_dims);
    }
    // prefactor (hyb.cpp:75-92)
    real pref = 1.0;
    for (int x = 0; x < m - 1; x++) {
      int e = bb.prefactor_Kexps[x];
      if (e != 0) {
        int l = bb.pole_inds[x];
        real k0 = bb.prefactor_Ksigns[x] > 0 ? hyb.k_0_p[l] : hyb.k_0_m[l];
        pref *= std::pow(k0, -e);
      }
    }
    for (auto &x : T) x *= pref;
    int sgn = (m % 2 == 0) ? -1 : 1;
    if (bb.fb[0] == 0) sgn *= -1;
    sgn *= bb.prefactor_sign;
    real mx = 0;
    for (auto &x : T) {
      x *= (real)sgn;
      mx = std::max(mx, absval(x));
    }
    if (mx > 1e-16) Sigma.add_block(b_ix, T);
  }

  const BlockOpSymSet<S> &Fsel(int w, int set, const Backbone &bb) const { return bb.vertices[w].dag ? Fq.F_dags[set] : Fq.Fs[set]; }

  // find_path_self_energy (block_sparse_backbone.cpp:229-314)
  void find_path(const BlockDiagOpFun<S> &Gt, Backbone &bb, int f_ix) {
    int m = bb.m, vct0 = bb.T(0, 1);
    bb.set_flat_index(f_ix, hyb.poles);
    std::vector<int> ind_path(2 * m - 1, 0), block_dims(2 * m + 1, 0);
    auto Fw = [&](int w) -> const BlockOpSymSet<S> & { return Fsel(w, (int)Fq.sym_set_labels[bb.orb_inds[w]], bb); };
    auto incomplete = [&](int w, int ip) { return ip == -1 || (w < 2 * m - 1 && Gt.zero[ip] == -1); };
    int nb = (int)Gt.blocks.size();
    for (int b_ix = 0; b_ix < nb; b_ix++) {
      for (int p_kap = 0; p_kap < q; p_kap++) {
        bool inc = false;
        int ip = -1;
        for (int w = 0; w < vct0; w++) {
          ip = (w == 0) ? Fsel(w, p_kap, bb).get_block_index(b_ix) : Fw(w).get_block_index(ip);
          inc = incomplete(w, ip);
          if (inc) break;
          ind_path[w] = ip;
        }
        if (inc) continue;
        block_dims[0] = Fsel(0, p_kap, bb).get_block_size(b_ix, 1);
        block_dims[1] = Fsel(0, p_kap, bb).get_block_size(b_ix, 0);
        for (int w = 0; w < vct0 - 1; w++) block_dims[w + 2] = Fw(w + 1).get_block_size(ind_path[w], 0);
        for (int p_mu = 0; p_mu < q; p_mu++) {
          if (p_mu != p_kap) continue;
          ip = ind_path[vct0 - 1];
          for (int w = vct0; w < 2 * m; w++) {
            ip = (w == vct0) ? Fsel(w, p_mu, bb).get_block_index(ip) : Fw(w).get_block_index(ip);
            inc = incomplete(w, ip);
            if (inc) break;
            if (w < 2 * m - 1) ind_path[w] = ip;
          }
          if (inc) continue;
          block_dims[vct0 + 1] = Fsel(vct0, p_mu, bb).get_block_size(ind_path[vct0 - 1], 0);
          for (int w = vct0; w < 2 * m - 1; w++) block_dims[w + 2] = Fw(w + 1).get_block_size(ind_path[w], 0);
          eval_fixed(Gt, bb, b_ix, p_kap, p_mu, ind_path, block_dims);
        }
      }
    }
    bb.reset_all_inds();
  }

  long num_backbones(int m) const { return (long)(1 << m) * (long)std::pow(n, m - 1) * (long)std::pow(hyb.p, m - 1); }

  // ------------------------------------------------------------------ correlators

  // hyb.cpp:75-92 (multiply_kernels_prefactor)
  real prefactor(const Backbone &bb) const {
    real pref = 1.0;
    for (int x = 0; x < bb.m - 1; x++) {
      int e = bb.prefactor_Kexps[x];
      if (e != 0) {
        int l = bb.pole_inds[x];
        real k0 = bb.prefactor_Ksigns[x] > 0 ? hyb.k_0_p[l] : hyb.k_0_m[l];
        pref *= std::pow(k0, -e);
      }
    }
    return pref;
  }

  // multiply_right_vertex (block_sparse_backbone.cpp:459-493): U (rows x b_out) <- U F, then the
  // vertex kernel with the opposite sign (hyb.cpp:46-57 with sign -1)
  void multiply_right_vertex(std::vector<S> &U, int rows, int &cols, Backbone &bb, int v, const std::vector<int> &ind_path,
                             const std::vector<int> &block_dims) {
    int o = bb.vertices[v].orb;
    int qi = (int)Fq.sym_set_labels[o], qo = (int)Fq.sym_set_inds[o];
    int l = bb.pole_inds[bb.vertices[v].hyb_ind];
    int b = ind_path[v - 1];
    int bin = block_dims[v], bout = block_dims[v + 1];
    const S *F;
    if (bb.vertices[v].bar) {
      auto &B = bb.vertices[v].dag ? Fq.F_dag_bars[qi] : Fq.F_bars_refl[qi];
      F = &B.blocks[b][((size_t)qo * B.p + l) * bout * bin];
    } else {
      auto &B = bb.vertices[v].dag ? Fq.F_dags[qi] : Fq.Fs[qi];
      F = &B.blocks[b][(size_t)qo * bout * bin];
    }
    if (cols != bout) throw std::logic_error("right vertex shape");
    std::vector<S> out((size_t)r * rows * bin, S(0));
    for (int t = 0; t < r; ++t)
      for (int i = 0; i < rows; ++i)
        for (int j = 0; j < bin; ++j) {
          S s(0);
          for (int k = 0; k < bout; ++k) s += U[((size_t)t * rows + i) * bout + k] * F[k * bin + j];
          out[((size_t)t * rows + i) * bin + j] = s;
        }
    int Ks = bb.vertices[v].Ksign;
    if (Ks != 0) {
      auto &kt = (-Ks > 0) ? hyb.k_it_p : hyb.k_it_m;
      for (int t = 0; t < r; ++t)
        for (int x = 0; x < rows * bin; ++x) out[(size_t)t * rows * bin + x] *= kt[(size_t)t * hyb.p + l];
    }
    U.swap(out);
    cols = bin;
  }

  // integrate_right_edge (block_sparse_backbone.cpp:495-514): U <- convolve_TO(U, G_b K...)
  void integrate_right_edge(std::vector<S> &U, int rows, int cols, const BlockDiagOpFun<S> &Gt, Backbone &bb, int e,
                            const std::vector<int> &ind_path) {
    int b = ind_path[e];
    int d = Gt.dims[b];
    std::vector<S> GK = Gt.blocks[b];
    for (int x = 0; x < bb.m - 1; x++) {  // hyb.cpp:59-73
      int Ks = bb.get_edge(e, x);
      if (Ks != 0) {
        int l = bb.pole_inds[x];
        auto &kt = (Ks > 0) ? hyb.k_it_p : hyb.k_it_m;
        for (int t = 0; t < r; ++t)
          for (int y = 0; y < d * d; ++y) GK[(size_t)t * d * d + y] *= kt[(size_t)t * hyb.p + l];
      }
    }
    if (d != cols) throw std::logic_error("right edge shape");
    ops.vals2coefs(U.data(), rows * cols);
    ops.vals2coefs(GK.data(), d * d);
    U = ops.convolve_to(beta, U, GK, rows, d, d);
  }

  // eval_correlator for one flat index (block_sparse_backbone.cpp:526-706); adds into corr
  // (r, nmu, nkap): sum over Sigma-free paths of Tr[left(t) O_mu right(t) O_kap], where right is
  // the [tau, 0] side and left the reflected [beta, tau] side.
  void eval_correlator(const BlockDiagOpFun<S> &Gt, Backbone &bb, const std::vector<BlockOp<S>> &mu_ops,
                       const std::vector<BlockOp<S>> &kap_ops, int f_ix, std::vector<S> &corr) {
    const int m = bb.m, vct0 = bb.T(0, 1);
    const int nb = (int)Gt.blocks.size(), nmu = (int)mu_ops.size(), nkap = (int)kap_ops.size();
    bb.set_flat_index(f_ix, hyb.poles);
    std::vector<int> ind_path(2 * m, 0), block_dims(2 * m + 1, 0);
    std::vector<std::vector<S>> right(nb), left(nb);
    std::vector<int> right_inds(nb, -1), left_inds(nb, -1);
    auto Fw = [&](int w) -> const BlockOpSymSet<S> & { return Fsel(w, (int)Fq.sym_set_labels[bb.orb_inds[w]], bb); };
    auto incomplete = [&](int ip) { return ip == -1 || Gt.zero[ip] == -1; };
    const real pref = prefactor(bb) * (real)bb.prefactor_sign;
    // [tau, 0] side
    for (int b_ix = 0; b_ix < nb; ++b_ix) {
      bool inc = false;
      int ip = b_ix;
      ind_path[0] = ip;
      for (int w = 1; w < vct0; w++) {
        ip = Fw(w).get_block_index(ip);
        inc = incomplete(ip);
        if (inc) break;
        ind_path[w] = ip;
      }
      if (inc) continue;
      block_dims[1] = Gt.dims[b_ix];
      for (int w = 1; w < vct0; w++) block_dims[w + 1] = Fw(w).get_block_size(ind_path[w - 1], 0);
      std::vector<S> T = Gt.blocks[b_ix];
      int rows = block_dims[1], cols = block_dims[1];
      for (int v = 1; v < vct0; v++) {
        multiply_left_vertex(T, rows, cols, bb, v, ind_path, block_dims);
        integrate_left_edge(T, rows, cols, Gt, bb, v, ind_path);
      }
      for (auto &x : T) x *= pref;
      right[b_ix] = std::move(T);  // (block_dims[vct0], block_dims[1])
      right_inds[b_ix] = ind_path[vct0 - 1];
    }
    // [beta, tau] side
    for (int b_ix = 0; b_ix < nb; ++b_ix) {
      bool inc = false;
      int ip = b_ix;
      ind_path[vct0] = ip;
      for (int w = vct0 + 1; w < 2 * m; w++) {
        ip = Fw(w).get_block_index(ip);
        inc = incomplete(ip);
        if (inc) break;
        ind_path[w] = ip;
      }
      if (inc) continue;
      block_dims[vct0 + 1] = Gt.dims[b_ix];
      for (int w = vct0 + 1; w < 2 * m; w++) block_dims[w + 1] = Fw(w).get_block_size(ind_path[w - 1], 0);
      std::vector<S> U = Gt.blocks[ind_path[2 * m - 1]];
      const int rows = block_dims[2 * m];
      int cols = rows;
      for (int v = 2 * m - 1; v > vct0; v--) {
        multiply_right_vertex(U, rows, cols, bb, v, ind_path, block_dims);
        integrate_right_edge(U, rows, cols, Gt, bb, v - 1, ind_path);
      }
      left[b_ix] = ops.reflect(U, rows * cols);  // (block_dims[2m], block_dims[vct0 + 1])
      left_inds[b_ix] = ind_path[2 * m - 1];
    }
    // contraction
    std::vector<S> c((size_t)r * nmu * nkap, S(0));
    for (int mu = 0; mu < nmu; ++mu) {
      const BlockOp<S> &Om = mu_ops[mu];
      for (int b_ix = 0; b_ix < nb; ++b_ix) {
        int e0 = right_inds[b_ix];
        if (incomplete(e0)) continue;
        int e1 = Om.block_indices[e0];
        if (incomplete(e1)) continue;
        int cx = left_inds[e1];
        if (incomplete(cx)) continue;
        const int db = Gt.dims[b_ix], d0 = Gt.dims[e0], d1 = Gt.dims[e1], dc = Gt.dims[cx];
        const std::vector<S> &R = right[b_ix], &L = left[e1];
        const S *M = Om.blocks[e0].data();  // (d1, d0)
        // Tmuop(t) = L(t) M R(t): (dc, db)
        std::vector<S> Tm((size_t)r * dc * db, S(0)), MR((size_t)d1 * db);
        for (int t = 0; t < r; ++t) {
          for (int i = 0; i < d1; ++i)
            for (int j = 0; j < db; ++j) {
              S s(0);
              for (int k = 0; k < d0; ++k) s += M[i * d0 + k] * R[((size_t)t * d0 + k) * db + j];
              MR[(size_t)i * db + j] = s;
            }
          for (int i = 0; i < dc; ++i)
            for (int j = 0; j < db; ++j) {
              S s(0);
              for (int k = 0; k < d1; ++k) s += L[((size_t)t * dc + i) * d1 + k] * MR[(size_t)k * db + j];
              Tm[((size_t)t * dc + i) * db + j] = s;
            }
        }
        for (int kap = 0; kap < nkap; ++kap) {
          if (kap_ops[kap].block_indices[cx] != b_ix) continue;
          const S *K = kap_ops[kap].blocks[cx].data();  // (db, dc)
          for (int t = 0; t < r; ++t) {
            S tr(0);
            for (int i = 0; i < dc; ++i)
              for (int j = 0; j < db; ++j) tr += Tm[((size_t)t * dc + i) * db + j] * K[j * dc + i];
            c[((size_t)t * nmu + mu) * nkap + kap] += tr;
          }
        }
      }
    }
    for (size_t x = 0; x < c.size(); ++x) corr[x] += c[x];
    bb.reset_all_inds();
  }

  long num_correlator_backbones(int m) const { return (long)(1 << (m - 1)) * (long)std::pow(n, m - 1) * (long)std::pow(hyb.p, m - 1); }

  // setup_mu_ops / setup_kap_ops_for_single_ptcle_gf (block_sparse_backbone.cpp:737-777): one
  // operator per orbital, in orbital order
  std::vector<BlockOp<S>> orbital_ops(bool dag) const {
    std::vector<BlockOp<S>> out;
    for (int o = 0; o < n; ++o) {
      const auto &F = dag ? Fq.F_dags[Fq.sym_set_labels[o]] : Fq.Fs[Fq.sym_set_labels[o]];
      int i = (int)Fq.sym_set_inds[o];
      BlockOp<S> op;
      op.block_indices = F.block_indices;
      op.bout = F.bout;
      op.bin = F.bin;
      op.blocks.resize(F.blocks.size());
      for (size_t j = 0; j < F.blocks.size(); ++j) {
        if (F.block_indices[j] == -1) continue;
        size_t sz = (size_t)F.bout[j] * F.bin[j];
        op.blocks[j].assign(F.blocks[j].begin() + i * sz, F.blocks[j].begin() + (i + 1) * sz);
      }
      out.push_back(std::move(op));
    }
    return out;
  }

  // compute_one_time_correlator / compute_single_ptcle_gf (block_sparse_backbone.cpp:779-902):
  // sum over all flat indices (fix == nullptr) or the listed ones
  std::vector<S> compute_correlator(const BlockDiagOpFun<S> &Gt, const int *topo, int m, const int *fix, long nfix,
                                    const std::vector<BlockOp<S>> &mu_ops, const std::vector<BlockOp<S>> &kap_ops) {
    Backbone bb(topo, m, n, true);
    std::vector<S> corr((size_t)r * mu_ops.size() * kap_ops.size(), S(0));
    long fmax = num_correlator_backbones(m);
    if (fix == nullptr) {
      for (long f = 0; f < fmax; ++f) eval_correlator(Gt, bb, mu_ops, kap_ops, (int)f, corr);
    } else {
      for (long i = 0; i < nfix; ++i) {
        // the reference decodes any integer; an index past the end would alias another backbone
        if (fix[i] < 0 || fix[i] >= fmax) throw std::runtime_error("DiagramEvaluator::eval_correlator: f_ix out of range");
        eval_correlator(Gt, bb, mu_ops, kap_ops, fix[i], corr);
      }
    }
    return corr;
  }

  // compute_self_energy (block_sparse_backbone.cpp:376-459); fix == nullptr means all f_ix
  void compute_self_energy(const BlockDiagOpFun<S> &Gt, const int *topo, int m, const int *fix, long nfix) {
    Sigma = Gt;
    for (size_t i = 0; i < Sigma.blocks.size(); ++i)
      for (auto &x : Sigma.blocks[i]) x *= 0.0;
    Sigma.dropped.assign(Sigma.blocks.size(), 0);
    Backbone bb(topo, m, n);
    long fmax = num_backbones(m);
    if (fix == nullptr) {
      for (long f = 0; f < fmax; ++f) find_path(Gt, bb, (int)f);
      Sigma.set_zero_block_indices();
    } else {
      for (long i = 0; i < nfix; ++i) {
        if (fix[i] < 0 || fix[i] >= fmax) throw std::runtime_error("DiagramEvaluator::eval_self_energy: f_ix out of range");
        find_path(Gt, bb, fix[i]);
        Sigma.set_zero_block_indices();
      }
    }
  }
};

// ---------------------------------------------------------------- Dyson (dlr_dyson_ppsc.hpp:92-168)

// Matrix of X -> convolve_to(beta, fc, vals2coefs(X)) acting on one column of X(t) (b x b blocks),
// indexed (t, i) x (t', i'). The coefficient-to-value convolution matrix Mc (columns: unit coefficient
// vectors through convolve_to) is composed with vals2coefs from the right by a transposed LU solve,
// C^T = cf2it^{-T} Mc^T, as every other coefficient transform here goes through the LU factors.
template <class S>
std::vector<S> convmat(const ImTimeOps &ops, real beta, const std::vector<S> &fc, int b) {
  int r = ops.r, N = r * b;
  std::vector<S> Mc((size_t)N * N, S(0));
  std::vector<S> g((size_t)r * b), res;
  for (int kp = 0; kp < r; ++kp)
    for (int ip = 0; ip < b; ++ip) {
      std::fill(g.begin(), g.end(), S(0));
      g[(size_t)kp * b + ip] = S(1);
      res = ops.convolve_to(beta, fc, g, b, b, 1);
      for (int t = 0; t < r; ++t)
        for (int i = 0; i < b; ++i) Mc[((size_t)t * b + i) * N + (size_t)kp * b + ip] = res[(size_t)t * b + i];
    }
  // each row (t, i) and column orbital i': y(t') solves cf2it^T y = Mc[(t,i), (:, i')]
  std::vector<S> y(r);
  for (size_t row = 0; row < (size_t)N; ++row)
    for (int ip = 0; ip < b; ++ip) {
      for (int k = 0; k < r; ++k) y[k] = Mc[row * N + (size_t)k * b + ip];
      ops.solve_transposed(y.data());
      for (int k = 0; k < r; ++k) Mc[row * N + (size_t)k * b + ip] = y[k];
    }
  return Mc;
}

template <class S> void lu_solve_inplace(std::vector<S> &A, int N, std::vector<S> &B, int nrhs) {
  std::vector<int> piv(N);
  for (int k = 0; k < N; ++k) {
    int p = k;
    for (int i = k + 1; i < N; ++i)
      if (absval(A[(size_t)i * N + k]) > absval(A[(size_t)p * N + k])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < N; ++j) std::swap(A[(size_t)k * N + j], A[(size_t)p * N + j]);
    for (int i = k + 1; i < N; ++i) {
      S l = A[(size_t)i * N + k] / A[(size_t)k * N + k];
      A[(size_t)i * N + k] = l;
      for (int j = k + 1; j < N; ++j) A[(size_t)i * N + j] -= l * A[(size_t)k * N + j];
    }
  }
  for (int k = 0; k < N; ++k)
    if (piv[k] != k)
      for (int j = 0; j < nrhs; ++j) std::swap(B[(size_t)k * nrhs + j], B[(size_t)piv[k] * nrhs + j]);
  for (int i = 1; i < N; ++i)
    for (int k = 0; k < i; ++k)
      for (int j = 0; j < nrhs; ++j) B[(size_t)i * nrhs + j] -= A[(size_t)i * N + k] * B[(size_t)k * nrhs + j];
  for (int i = N - 1; i >= 0; --i) {
    for (int k = i + 1; k < N; ++k)
      for (int j = 0; j < nrhs; ++j) B[(size_t)i * nrhs + j] -= A[(size_t)i * N + k] * B[(size_t)k * nrhs + j];
    for (int j = 0; j < nrhs; ++j) B[(size_t)i * nrhs + j] /= A[(size_t)i * N + i];
  }
}

// G = (1 - convmat(G0 * (Sigma + eta)))^{-1} G0, column by column
template <class S>
std::vector<S> dyson_solve(const ImTimeOps &ops, real beta, const std::vector<S> &g0, const std::vector<S> &sig, real eta, int b) {
  int r = ops.r, N = r * b;
  std::vector<S> g0c(g0), sigc(sig);
  ops.vals2coefs(g0c.data(), b * b);
  ops.vals2coefs(sigc.data(), b * b);
  auto ggs = ops.convolve_to(beta, g0c, sigc, b, b, b);
  for (size_t x = 0; x < ggs.size(); ++x) ggs[x] = -(ggs[x] + eta * g0[x]);
  ops.vals2coefs(ggs.data(), b * b);
  auto M = convmat(ops, beta, ggs, b);
  for (int i = 0; i < N; ++i) M[(size_t)i * N + i] += S(1);
  // right-hand sides: column j of G0(t), as a vector over (t, i)
  std::vector<S> B((size_t)N * b);
  for (int t = 0; t < r; ++t)
    for (int i = 0; i < b; ++i)
      for (int j = 0; j < b; ++j) B[((size_t)t * b + i) * b + j] = g0[((size_t)t * b + i) * b + j];
  lu_solve_inplace(M, N, B, b);
  return B;  // (t, i, j) layout already
}

} // namespace orc

// ============================================================================ C ABI
using namespace orc;

struct orc_handle;
static ioc to_io(real x) { return ioc((double)x, 0.0); }
static ioc to_io(const cplx &x) { return ioc((double)x.real(), (double)x.imag()); }
template <class V> static void to_io(const V &src, double *out) {
  ioc *o = reinterpret_cast<ioc *>(out);
  for (size_t x = 0; x < src.size(); ++x) o[x] = to_io(src[x]);
}
static std::vector<cplx> from_io(const double *in, size_t n) {
  const ioc *v = reinterpret_cast<const ioc *>(in);
  return std::vector<cplx>(v, v + n);
}

template <class S> static BlockDiagOpFun<S> make_G(const Problem &P, const double *G) {
  BlockDiagOpFun<S> Gt;
  Gt.r = P.r;
  Gt.dims = P.dims;
  const ioc *g = reinterpret_cast<const ioc *>(G);
  for (int b = 0; b < P.nblocks; ++b) {
    size_t sz = (size_t)P.r * P.dims[b] * P.dims[b];
    std::vector<S> blk(sz);
    for (size_t x = 0; x < sz; ++x) blk[x] = DiagramEvaluator<S>::to_s(cplx(g[x]));
    g += sz;
    Gt.blocks.push_back(blk);
  }
  Gt.zero.assign(P.nblocks, 0);
  Gt.set_zero_block_indices();
  return Gt;
}

// block operators from (nops, nblocks) connections and matrices packed in (op, source block) order
template <class S> static std::vector<BlockOp<S>> make_ops(const Problem &P, int nops, const int *conn, const double *mats) {
  std::vector<BlockOp<S>> out(nops);
  const ioc *m = reinterpret_cast<const ioc *>(mats);
  for (int o = 0; o < nops; ++o) {
    BlockOp<S> &op = out[o];
    op.block_indices.assign(conn + (size_t)o * P.nblocks, conn + (size_t)(o + 1) * P.nblocks);
    op.blocks.resize(P.nblocks);
    op.bout.assign(P.nblocks, 0);
    op.bin.assign(P.nblocks, 0);
    for (int j = 0; j < P.nblocks; ++j) {
      int t = op.block_indices[j];
      if (t < -1 || t >= P.nblocks) throw std::invalid_argument("correlator operator: block index out of range");
      if (t < 0) continue;
      op.bout[j] = P.dims[t];
      op.bin[j] = P.dims[j];
      size_t sz = (size_t)P.dims[t] * P.dims[j];
      op.blocks[j].resize(sz);
      for (size_t x = 0; x < sz; ++x) op.blocks[j][x] = DiagramEvaluator<S>::to_s(cplx(m[x]));
      m += sz;
    }
  }
  return out;
}

template <class S> static int correlator_impl(DiagramEvaluator<S> &D, const Problem &P, const double *G, int m, const int *topo, const int *fix,
                                              long nfix, int nmu, const int *mu_conn, const double *mu_mats, int nkap, const int *kap_conn,
                                              const double *kap_mats, double *out) {
  auto Gt = make_G<S>(P, G);
  std::vector<BlockOp<S>> mu, kap;
  if (mu_conn == nullptr) {  // single-particle Green's function: c_o and c^dag_o per orbital
    mu = D.orbital_ops(false);
    kap = D.orbital_ops(true);
  } else {
    mu = make_ops<S>(P, nmu, mu_conn, mu_mats);
    kap = make_ops<S>(P, nkap, kap_conn, kap_mats);
  }
  auto c = D.compute_correlator(Gt, topo, m, fix, nfix, mu, kap);
  to_io(c, out);
  return 0;
}

template <class S> static int sigma_impl(DiagramEvaluator<S> &D, const Problem &P, const double *G, int m, const int *topo, const int *fix,
                                         long nfix, double *out) {
  auto Gt = make_G<S>(P, G);
  D.compute_self_energy(Gt, topo, m, fix, nfix);
  ioc *o = reinterpret_cast<ioc *>(out);
  for (int b = 0; b < P.nblocks; ++b) {
    size_t sz = (size_t)P.r * P.dims[b] * P.dims[b];
    for (size_t x = 0; x < sz; ++x) o[x] = (D.Sigma.zero[b] == -1) ? ioc(0) : to_io(D.Sigma.blocks[b][x]);
    o += sz;
  }
  return 0;
}


extern "C" {

struct orc_problem {
  double beta;
  int r;
  const double *it;
  const double *rf;
  int p;
  const double *poles;
  const double *coeffs;  // (p, n, n) complex, interleaved re/im
  int n;
  int nblocks;
  const int *dims;
  const int *c_conn;      // (n, nblocks)
  const int *cdag_conn;   // (n, nblocks)
  const double *c_mats;   // concatenated complex blocks in (o, j) order for connected j
  const double *cdag_mats;
};

static thread_local std::string g_err;
const char *orc_last_error() { return g_err.c_str(); }

struct orc_handle {
  Problem P;
  std::unique_ptr<DiagramEvaluator<cplx>> dc;
  std::unique_ptr<DiagramEvaluator<real>> dr;
};

void *orc_create(const orc_problem *d, int complex_arith) {
  try {
    auto h = new orc_handle();
    Problem &P = h->P;
    P.beta = d->beta; P.r = d->r; P.p = d->p; P.n = d->n; P.nblocks = d->nblocks;
    P.it.assign(d->it, d->it + d->r);
    P.rf.assign(d->rf, d->rf + d->r);
    P.poles.assign(d->poles, d->poles + d->p);
    const ioc *cc = reinterpret_cast<const ioc *>(d->coeffs);
    P.coeffs.assign(cc, cc + (size_t)d->p * d->n * d->n);
    P.dims.assign(d->dims, d->dims + d->nblocks);
    P.c_conn.assign(d->c_conn, d->c_conn + (size_t)d->n * d->nblocks);
    P.cdag_conn.assign(d->cdag_conn, d->cdag_conn + (size_t)d->n * d->nblocks);
    auto unpack = [&](const std::vector<int> &conn, const double *src, std::vector<std::vector<cplx>> &dst) {
      dst.assign((size_t)P.n * P.nblocks, {});
      const ioc *s = reinterpret_cast<const ioc *>(src);
      for (int o = 0; o < P.n; ++o)
        for (int j = 0; j < P.nblocks; ++j) {
          int t = conn[o * P.nblocks + j];
          if (t < 0) continue;
          size_t sz = (size_t)P.dims[t] * P.dims[j];
          dst[(size_t)o * P.nblocks + j].assign(s, s + sz);
          s += sz;
        }
    };
    unpack(P.c_conn, d->c_mats, P.c_mats);
    unpack(P.cdag_conn, d->cdag_mats, P.cdag_mats);
    if (complex_arith) h->dc.reset(new DiagramEvaluator<cplx>(P));
    else h->dr.reset(new DiagramEvaluator<real>(P));
    return h;
  } catch (std::exception &e) {
    g_err = e.what();
    return nullptr;
  }
}

void orc_destroy(void *h) { delete static_cast<orc_handle *>(h); }

long orc_num_backbones(void *hv, int m) {
  auto h = static_cast<orc_handle *>(hv);
  return h->dc ? h->dc->num_backbones(m) : h->dr->num_backbones(m);
}

int orc_num_sym_sets(void *hv) {
  auto h = static_cast<orc_handle *>(hv);
  return h->dc ? h->dc->q : h->dr->q;
}

// Sigma for one topology, all f_ix (fix == NULL) or the listed ones. G and out: concatenated complex
// blocks (r, b, b). Returns 0, or -1 with orc_last_error() set.
int orc_sigma(void *hv, const double *G, int m, const int *topo, const int *fix, long nfix, double *out) {
  try {
    auto h = static_cast<orc_handle *>(hv);
    if (h->dc) return sigma_impl(*h->dc, h->P, G, m, topo, fix, nfix, out);
    return sigma_impl(*h->dr, h->P, G, m, topo, fix, nfix, out);
  } catch (std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// Per block, the sum of max |.| over the running sums the last orc_sigma call discarded (a block whose
// sum fell below 1e-16 after one flat index of a list is replaced at its next contribution).
int orc_last_dropped(void *hv, double *out) {
  auto h = static_cast<orc_handle *>(hv);
  const auto &d = h->dc ? h->dc->Sigma.dropped : h->dr->Sigma.dropped;
  for (size_t i = 0; i < d.size(); ++i) out[i] = (double)d[i];
  return (int)d.size();
}

long orc_num_correlator_backbones(void *hv, int m) {
  auto h = static_cast<orc_handle *>(hv);
  return h->dc ? h->dc->num_correlator_backbones(m) : h->dr->num_correlator_backbones(m);
}

// One-time correlator <T O_mu(tau) O_kap(0)> diagrams of one topology, out (r, nmu, nkap) complex.
// mu_conn == NULL selects the single-particle Green's function (O_mu = c_mu, O_kap = c^dag_kap,
// nmu = nkap = n). Operators: (nops, nblocks) target blocks and matrices packed as for orc_problem.
int orc_correlator(void *hv, const double *G, int m, const int *topo, const int *fix, long nfix, int nmu, const int *mu_conn,
                   const double *mu_mats, int nkap, const int *kap_conn, const double *kap_mats, double *out) {
  try {
    auto h = static_cast<orc_handle *>(hv);
    if (h->dc) return correlator_impl(*h->dc, h->P, G, m, topo, fix, nfix, nmu, mu_conn, mu_mats, nkap, kap_conn, kap_mats, out);
    return correlator_impl(*h->dr, h->P, G, m, topo, fix, nfix, nmu, mu_conn, mu_mats, nkap, kap_conn, kap_mats, out);
  } catch (std::exception &e) {
    g_err = e.what();
    return -1;
  }
}

// DLR unit operations on a given grid (complex data, interleaved)
int orc_vals2coefs(int r, const double *it, const double *rf, double *data, int m) {
  try {
    ImTimeOps ops(it, rf, r);
    auto v = from_io(data, (size_t)r * m);
    ops.vals2coefs(v.data(), m);
    to_io(v, data);
    return 0;
  } catch (std::exception &e) { g_err = e.what(); return -1; }
}

int orc_convolve(double beta, int r, const double *it, const double *rf, const double *fvals, const double *gvals, int a, int b, int c,
                 double *out) {
  try {
    ImTimeOps ops(it, rf, r);
    auto fc = from_io(fvals, (size_t)r * a * b), gc = from_io(gvals, (size_t)r * b * c);
    ops.vals2coefs(fc.data(), a * b);
    ops.vals2coefs(gc.data(), b * c);
    auto h = ops.convolve_to(real(beta), fc, gc, a, b, c);
    to_io(h, out);
    return 0;
  } catch (std::exception &e) { g_err = e.what(); return -1; }
}

int orc_reflect(int r, const double *it, const double *rf, const double *vals, int m, double *out) {
  try {
    ImTimeOps ops(it, rf, r);
    auto o = ops.reflect(from_io(vals, (size_t)r * m), m);
    to_io(o, out);
    return 0;
  } catch (std::exception &e) { g_err = e.what(); return -1; }
}

// evaluate at relative times ts[nt] from values at the nodes: out (nt, m)
int orc_interp(int r, const double *it, const double *rf, const double *vals, int m, const double *ts, int nt, double *out) {
  try {
    ImTimeOps ops(it, rf, r);
    auto c = from_io(vals, (size_t)r * m);
    ops.vals2coefs(c.data(), m);
    for (int k = 0; k < nt; ++k) to_io(ops.coefs2eval(c, m, real(ts[k])), out + 2 * (size_t)k * m);
    return 0;
  } catch (std::exception &e) { g_err = e.what(); return -1; }
}

int orc_dyson(double beta, int r, const double *it, const double *rf, int b, const double *g0, const double *sig, double eta, double *out) {
  try {
    ImTimeOps ops(it, rf, r);
    size_t sz = (size_t)r * b * b;
    auto G = dyson_solve(ops, real(beta), from_io(g0, sz), from_io(sig, sz), real(eta), b);
    to_io(G, out);
    return 0;
  } catch (std::exception &e) { g_err = e.what(); return -1; }
}

// Backbone decode for known-answer tests: fills per-vertex flags and edges/prefactor.
// out_vert (2m, 5): bar, dag, hyb_ind, Ksign, orb; out_edges (2m-1, m-1); out_pref (m-1, 2): sign, exp;
// out_misc: prefactor_sign, fb[0..m)
int orc_backbone(const int *topo, int m, int n, const double *poles_dimless, int p, int f_ix, int *out_vert, int *out_edges, int *out_pref,
                 int *out_misc) {
  try {
    Backbone bb(topo, m, n);
    std::vector<real> poles(poles_dimless, poles_dimless + p);
    bb.set_flat_index(f_ix, poles);
    for (int v = 0; v < 2 * m; ++v) {
      out_vert[v * 5 + 0] = bb.vertices[v].bar;
      out_vert[v * 5 + 1] = bb.vertices[v].dag;
      out_vert[v * 5 + 2] = bb.vertices[v].hyb_ind;
      out_vert[v * 5 + 3] = bb.vertices[v].Ksign;
      out_vert[v * 5 + 4] = bb.vertices[v].orb;
    }
    for (int e = 0; e < 2 * m - 1; ++e)
      for (int x = 0; x < m - 1; ++x) out_edges[e * (m - 1) + x] = bb.get_edge(e, x);
    for (int x = 0; x < m - 1; ++x) {
      out_pref[2 * x] = bb.prefactor_Ksigns[x];
      out_pref[2 * x + 1] = bb.prefactor_Kexps[x];
    }
    out_misc[0] = bb.prefactor_sign;
    for (int i = 0; i < m; ++i) out_misc[1 + i] = bb.fb[i];
    for (int i = 0; i < m - 1; ++i) out_misc[1 + m + i] = bb.pole_inds[i];
    return 0;
  } catch (std::exception &e) { g_err = e.what(); return -1; }
}

}  // extern "C"
