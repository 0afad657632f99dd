#include "model.hpp"

#include <algorithm>
#include <cmath>
#include <map>
#include <sstream>
#include <stdexcept>

#include "linalg.hpp"

namespace xca {

static int64_t ipow(int64_t a, int e) {
  int64_t r = 1;
  for (int i = 0; i < e; ++i) r *= a;
  return r;
}

int64_t Model::num_backbones(int m) const { return ipow(2, m) * ipow(P.n, m - 1) * ipow(P.p, m - 1); }

void PoleSets::init(const std::vector<double> &w) {
  fwd_a.clear(); fwd_b.clear(); bwd_a.clear(); bwd_b.clear();
  for (int l = 0; l < (int)w.size(); ++l) {
    (w[l] <= 0 ? fwd_a : fwd_b).push_back(l);   // backbone.cpp:88
    (w[l] >= 0 ? bwd_a : bwd_b).push_back(l);   // backbone.cpp:110
  }
}

void Model::init(const Problem &Pin) {
  P = Pin;
  const int n = P.n, nb = P.nb, r = P.r, p = P.p;
  if ((int)P.it.size() != r || (int)P.rf.size() != r) throw std::invalid_argument("DiagramEvaluator: DLR node arrays must have r entries");
  if ((int)P.poles.size() != p) throw std::invalid_argument("DiagramEvaluator: hyb_poles.size() must equal p");
  if ((int64_t)P.coeffs.size() != (int64_t)p * n * n)
    throw std::invalid_argument("DiagramEvaluator: hyb_coeffs must have shape (p, n, n) with p = hyb_poles.size()");
  ops.init(P.it, P.rf);

  // symmetry sets from identical annihilator connection rows (atom_diag_utils.cpp:42-68)
  S.label.assign(n, 0);
  int counter = 0;
  for (int o = 0; o < n; ++o) {
    int found = -1;
    for (int o2 = 0; o2 < o && found < 0; ++o2)
      if (std::equal(P.c_conn.begin() + (size_t)o * nb, P.c_conn.begin() + (size_t)(o + 1) * nb, P.c_conn.begin() + (size_t)o2 * nb))
        found = o2;
    S.label[o] = (found >= 0) ? S.label[found] : counter++;
  }
  S.q = counter;
  S.ind.assign(n, 0);
  S.size.assign(S.q, 0);
  S.to_orb.assign(S.q, {});
  for (int o = 0; o < n; ++o) {
    S.ind[o] = S.size[S.label[o]]++;
    S.to_orb[S.label[o]].push_back(o);
  }
  for (int o = 0; o < n; ++o)
    for (int j = 0; j < nb; ++j) {
      int rep = S.to_orb[S.label[o]][0];
      if (P.cdag_conn[(size_t)o * nb + j] != P.cdag_conn[(size_t)rep * nb + j])
        throw std::invalid_argument("DiagramEvaluator: creation operators of one symmetry set have different block connections");
    }

  // cross-set coupling check (block_sparse.cpp:443-492)
  {
    double mx = 0;
    for (double x : P.coeffs) mx = std::max(mx, std::fabs(x));
    if (!std::isfinite(mx)) throw std::invalid_argument("DiagramEvaluator: hyb_coeffs contains a non-finite entry");
    double thresh = 1e-12 * mx, worst = 0;
    long nbad = 0, ncross = 0, wl = -1, wi = -1, wj = -1;
    for (int l = 0; l < p; ++l)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          if (S.label[i] == S.label[j]) continue;
          ncross++;
          double a = std::fabs(P.coeffs[((size_t)l * n + i) * n + j]);
          if (a > thresh) {
            nbad++;
            if (a > worst) { worst = a; wl = l; wi = i; wj = j; }
          }
        }
    if (nbad > 0) {
      std::ostringstream msg;
      msg << std::scientific << "DiagramEvaluator: hyb_coeffs couples different symmetry sets, which the symmetry-set storage cannot represent. "
          << nbad << " of " << ncross << " cross-set entries exceed the tolerance " << 1e-12 << " * max|hyb_coeffs| = " << thresh
          << "; the largest is |hyb_coeffs(" << wl << ", " << wi << ", " << wj << ")| = " << worst << " (relative " << worst / mx
          << "), coupling symmetry set " << S.label[wi] << " to symmetry set " << S.label[wj] << ".";
      throw std::invalid_argument(msg.str());
    }
  }

  // operator matrices in set layout and barred operators (block_sparse.cpp:556-576):
  //   F_dag_bar[s](lam, l) = sum_nu c(l, nu, lam) F^dag_nu ;  F_bar_refl[s](nu, l) = -sum_lam c(l, nu, lam) F_lam
  off_F.assign((size_t)S.q * 2 * nb, -1);
  off_bar.assign((size_t)S.q * 2 * nb, -1);
  conn_set.assign((size_t)S.q * 2 * nb, -1);
  opmats.clear();
  for (int s = 0; s < S.q; ++s) {
    int qs = S.size[s];
    for (int dag = 0; dag < 2; ++dag) {
      const auto &conn = dag ? P.cdag_conn : P.c_conn;
      const auto &mats = dag ? P.cdag_mats : P.c_mats;
      int rep = S.to_orb[s][0];
      for (int j = 0; j < nb; ++j) {
        int t = conn[(size_t)rep * nb + j];
        conn_set[((size_t)s * 2 + dag) * nb + j] = t;
        if (t < 0) continue;
        size_t sz = (size_t)P.dims[t] * P.dims[j];
        off_F[((size_t)s * 2 + dag) * nb + j] = (int64_t)opmats.size();
        for (int k = 0; k < qs; ++k) {
          const auto &M = mats[(size_t)S.to_orb[s][k] * nb + j];
          if (M.size() != sz) throw std::invalid_argument("DiagramEvaluator: operator block has the wrong shape");
          opmats.insert(opmats.end(), M.begin(), M.end());
        }
        off_bar[((size_t)s * 2 + dag) * nb + j] = (int64_t)opmats.size();
        size_t base = opmats.size();
        opmats.resize(base + (size_t)qs * p * sz, 0.0);
        for (int l = 0; l < p; ++l)
          for (int a = 0; a < qs; ++a)    // lam for dag, nu for refl
            for (int b = 0; b < qs; ++b) {  // nu for dag, lam for refl
              int oa = S.to_orb[s][a], ob = S.to_orb[s][b];
              double c, sg;
              const double *src;
              if (dag) {  // F_dag_bar(lam=a, l) += c(l, nu=b, lam=a) F^dag_nu
                c = P.coeffs[((size_t)l * n + ob) * n + oa];
                sg = 1.0;
                src = mats[(size_t)ob * nb + j].data();
              } else {    // F_bar_refl(nu=a, l) += -c(l, nu=a, lam=b) F_lam
                c = P.coeffs[((size_t)l * n + oa) * n + ob];
                sg = -1.0;
                src = mats[(size_t)ob * nb + j].data();
              }
              double *dst = &opmats[base + ((size_t)a * p + l) * sz];
              double cc = sg * c;
              for (size_t x = 0; x < sz; ++x) dst[x] += cc * src[x];
            }
      }
    }
  }

  // transposed copies (the [beta, tau] side of correlators multiplies from the right)
  off_FT.assign(off_F.size(), -1);
  off_barT.assign(off_bar.size(), -1);
  for (int s = 0; s < S.q; ++s)
    for (int dag = 0; dag < 2; ++dag)
      for (int j = 0; j < nb; ++j) {
        const size_t key = ((size_t)s * 2 + dag) * nb + j;
        const int t = conn_set[key];
        if (t < 0) continue;
        const int dout = P.dims[t], din = P.dims[j];
        const size_t sz = (size_t)dout * din;
        const int nmat[2] = {S.size[s], S.size[s] * p};
        const int64_t src[2] = {off_F[key], off_bar[key]};
        int64_t *dst[2] = {&off_FT[key], &off_barT[key]};
        for (int w = 0; w < 2; ++w) {
          *dst[w] = (int64_t)opmats.size();
          const size_t base = opmats.size();
          opmats.resize(base + nmat[w] * sz);
          for (int k = 0; k < nmat[w]; ++k)
            for (int a = 0; a < dout; ++a)
              for (int b = 0; b < din; ++b) opmats[base + k * sz + (size_t)b * dout + a] = opmats[src[w] + k * sz + (size_t)a * din + b];
        }
      }

  // hybridization tables (hyb.cpp:15-44, 94-113)
  kp.assign((size_t)p * r, 0); km.assign((size_t)p * r, 0);
  k0p.assign(p, 0); k0m.assign(p, 0);
  for (int l = 0; l < p; ++l) {
    double w = P.poles[l] * P.beta;
    k0p[l] = dlr::kernel(0.0, +w);
    k0m[l] = dlr::kernel(0.0, -w);
    for (int t = 0; t < r; ++t) {
      kp[(size_t)l * r + t] = dlr::kernel(P.it[t], +w);
      km[(size_t)l * r + t] = dlr::kernel(P.it[t], -w);
    }
  }
  delta.assign((size_t)r * n * n, 0.0);
  for (int t = 0; t < r; ++t)
    for (int l = 0; l < p; ++l) {
      double k = dlr::kernel(P.it[t], P.beta * P.poles[l]);
      for (int x = 0; x < n * n; ++x) delta[(size_t)t * n * n + x] += k * P.coeffs[(size_t)l * n * n + x];
    }
  // reflection through coefficients (LU solve), values at -t_i
  {
    std::vector<double> c(delta);
    ops.vals2coefs(c.data(), n * n);
    delta_refl.assign((size_t)r * n * n, 0.0);
    for (int t = 0; t < r; ++t)
      for (int k = 0; k < r; ++k) {
        double kk = dlr::kernel(-P.it[t], P.rf[k]);
        for (int x = 0; x < n * n; ++x) delta_refl[(size_t)t * n * n + x] += kk * c[(size_t)k * n * n + x];
      }
  }
  // zero-vertex kappa sums (the Delta contraction of the pair's two operators, independent of G
  // and of the pole tuple), accumulated in kappa order with fused multiply-adds
  off_Fh.assign(off_F.size(), -1);
  if (*std::max_element(P.dims.begin(), P.dims.end()) <= 12)
    for (int s = 0; s < S.q; ++s)
      for (int dag = 0; dag < 2; ++dag)
        for (int j = 0; j < nb; ++j) {
          const size_t key = ((size_t)s * 2 + dag) * nb + j;
          const int tb = conn_set[key];
          if (tb < 0) continue;
          const int qs = S.size[s], b1 = P.dims[tb], b0 = P.dims[j];
          const size_t sz = (size_t)b1 * b0;
          off_Fh[key] = (int64_t)opmats.size();
          const size_t base = opmats.size();
          opmats.resize(base + (size_t)r * qs * sz, 0.0);
          for (int t = 0; t < r; ++t)
            for (int mu = 0; mu < qs; ++mu)
              for (size_t x = 0; x < sz; ++x) {
                double acc = 0.0;
                for (int kk = 0; kk < qs; ++kk) {
                  const int om = S.to_orb[s][mu], ok = S.to_orb[s][kk];
                  const double h = dag == 0 ? delta[((size_t)t * n + om) * n + ok] : -delta_refl[((size_t)t * n + ok) * n + om];
                  acc = std::fma(h, opmats[off_F[key] + kk * sz + x], acc);
                }
                opmats[base + ((size_t)t * qs + mu) * sz + x] = acc;
              }
        }
  G_off.assign(nb, 0);
  G_size = 0;
  Nmax = 0;
  for (int b = 0; b < nb; ++b) {
    G_off[b] = (int)G_size;
    G_size += (int64_t)r * P.dims[b] * P.dims[b];
    Nmax = std::max(Nmax, P.dims[b]);
  }
}

// ----------------------------------------------------------------------------------------------

namespace {
struct Walk {
  std::vector<int> ind_path, block_dims;
};
}  // namespace

void build_programs(const Model &M, const PoleSets &ps, const int *topo, int m, int topo_id, const std::vector<int> &G_zero, Plan &plan) {
  const Problem &P = M.P;
  const int n = P.n, nb = P.nb;
  auto T = [&](int i, int j) { return topo[2 * i + j]; };
  for (int i = 0; i < m; ++i)
    if (T(i, 0) >= T(i, 1)) throw std::invalid_argument("first row of topology must contain smaller-numbered vertices");
  if (T(0, 0) != 0) throw std::invalid_argument("topology(0,0) must be 0");
  const int vct0 = T(0, 1);
  const int nlines = m - 1;
  // line x (0-based, = topology row x+1) endpoints
  std::vector<int> la(nlines), lb(nlines);
  for (int x = 0; x < nlines; ++x) { la[x] = T(x + 1, 0); lb[x] = T(x + 1, 1); }
  // which line's left vertex carries a bar: vertex -> line
  std::vector<int> bar_line(2 * m, -1), line_of_vertex(2 * m, -1);
  for (int x = 0; x < nlines; ++x) {
    bar_line[lb[x]] = x;
    line_of_vertex[la[x]] = x;
    line_of_vertex[lb[x]] = x;
  }
  std::map<std::vector<int>, int> table_index;
  for (size_t i = 0; i < plan.tables.size(); ++i) {
    const auto &tb = plan.tables[i];
    table_index[{tb.block, tb.nlines, tb.lsign[0], tb.lsign[1], tb.lsign[2]}] = (int)i;
  }
  auto get_table = [&](int block, int nl, const int *sg) {
    std::vector<int> key = {block, nl, nl > 0 ? sg[0] : 0, nl > 1 ? sg[1] : 0, nl > 2 ? sg[2] : 0};
    auto it = table_index.find(key);
    if (it != table_index.end()) return it->second;
    EdgeTable tb;
    tb.block = block;
    tb.nlines = nl;
    for (int i = 0; i < nl; ++i) tb.lsign[i] = sg[i];
    tb.nentries = ipow(P.p, nl);
    tb.offset = plan.ftab_size;
    plan.ftab_size += tb.nentries * 2 * (int64_t)P.r * P.dims[block] * P.dims[block];
    plan.tables.push_back(tb);
    int idx = (int)plan.tables.size() - 1;
    table_index[key] = idx;
    return idx;
  };

  const int64_t o_max = ipow(n, m - 1);
  for (int fb_ix = 0; fb_ix < (1 << m); ++fb_ix) {
    std::vector<int> fb(m), dag(2 * m, 0), bar(2 * m, 0);
    for (int i = 0, f = fb_ix; i < m; ++i, f /= 2) fb[i] = f % 2;
    // backbone.cpp:35-66
    dag[0] = (fb[0] != 1);
    dag[vct0] = (fb[0] == 1);
    for (int i = 1; i < m; ++i) {
      bar[T(i, 1)] = 1;
      dag[T(i, 0)] = (fb[i] != 1);
      dag[T(i, 1)] = (fb[i] == 1);
    }
    for (int64_t o_ix = 0; o_ix < o_max; ++o_ix) {
      std::vector<int> orb(2 * m, -1);
      for (int i = 1, oo = (int)o_ix; i < m; ++i, oo /= n) { orb[T(i, 0)] = oo % n; orb[T(i, 1)] = oo % n; }
      auto conn_w = [&](int w, int ip) { return M.conn(M.S.label[orb[w]], dag[w], ip); };
      for (int b_ix = 0; b_ix < nb; ++b_ix) {
        for (int s = 0; s < M.S.q; ++s) {
          // block walk (block_sparse_backbone.cpp:250-306), p_mu == p_kap == s
          std::vector<int> ind(2 * m - 1, -1), bd(2 * m + 1, 0);
          bool inc = false;
          int ip = -1;
          auto incomplete = [&](int w, int x) { return x == -1 || (w < 2 * m - 1 && G_zero[x]); };
          for (int w = 0; w < vct0 && !inc; ++w) {
            ip = (w == 0) ? M.conn(s, dag[0], b_ix) : conn_w(w, ip);
            inc = incomplete(w, ip);
            if (!inc) ind[w] = ip;
          }
          if (inc) continue;
          int last = -1;
          for (int w = vct0; w < 2 * m && !inc; ++w) {
            ip = (w == vct0) ? M.conn(s, dag[vct0], ip) : conn_w(w, ip);
            inc = incomplete(w, ip);
            if (!inc) {
              if (w < 2 * m - 1) ind[w] = ip;
              else last = ip;
            }
          }
          if (inc) continue;
          if (last != b_ix) {
            std::ostringstream msg;
            msg << "BlockDiagOpFun::add_block: backbone path for Sigma block " << b_ix << " ends in block " << last;
            throw std::invalid_argument(msg.str());
          }
          bd[0] = P.dims[b_ix];
          bd[1] = P.dims[ind[0]];
          for (int w = 1; w < 2 * m; ++w) bd[w + 1] = (w < 2 * m - 1) ? P.dims[ind[w]] : P.dims[last];

          // one program per non-empty rule class
          for (int cls = 0; cls < (1 << nlines); ++cls) {
            bool empty = false;
            for (int x = 0; x < nlines; ++x)
              if (ps.get(fb[x + 1], (cls >> x) & 1).empty()) empty = true;
            if (empty) continue;
            Program pg;
            pg.m = m; pg.topo_id = topo_id; pg.sigma_block = b_ix; pg.b0 = bd[0]; pg.b1 = bd[1];
            pg.init_block = ind[0]; pg.fb_ix = fb_ix; pg.o_ix = (int)o_ix; pg.set = s; pg.cls = cls;
            // K placement per line (backbone.cpp:86-139)
            std::vector<int> vK(2 * m, 0), vKx(2 * m, -1);
            std::vector<std::vector<int>> eK(2 * m - 1, std::vector<int>(std::max(nlines, 1), 0));
            int psign = 1;
            for (int x = 0; x < nlines; ++x) {
              bool rb = (cls >> x) & 1;
              bool fwd = fb[x + 1] == 1;
              if (!fwd) psign = -psign;
              if (!rb) {
                vK[la[x]] = fwd ? -1 : 1; vKx[la[x]] = x;
                vK[lb[x]] = fwd ? 1 : -1; vKx[lb[x]] = x;
                pg.pk_sign[x] = fwd ? -1 : 1;
                pg.pk_exp[x] = 1;
              } else {
                for (int j = la[x]; j < lb[x]; ++j) eK[j][x] = fwd ? 1 : -1;
                pg.pk_sign[x] = fwd ? 1 : -1;
                pg.pk_exp[x] = lb[x] - la[x] - 1;
              }
            }
            int sgn = (m % 2 == 0) ? -1 : 1;
            if (fb[0] == 0) sgn = -sgn;
            pg.sign = sgn * psign;

            auto vertex_op = [&](int v, int cols) {
              Op op;
              op.type = OP_V;
              op.bin = bd[v];
              op.bout = bd[v + 1];
              op.cols = cols;
              int set = M.S.label[orb[v]], qo = M.S.ind[orb[v]];
              int src = ind[v - 1];
              size_t sz = (size_t)op.bout * op.bin;
              if (bar[v]) {
                op.fbase = M.bar_offset(set, dag[v], src) + (int64_t)qo * P.p * sz;
                op.fstride_l = (int64_t)sz;
                op.bar_line = bar_line[v];
              } else {
                op.fbase = M.F_offset(set, dag[v], src) + (int64_t)qo * sz;
              }
              if (op.fbase < 0) throw std::logic_error("missing operator block");
              if (vK[v] != 0) { op.kx = vKx[v]; op.ks = vK[v]; }
              return op;
            };
            auto edge_op = [&](int e, int cols) {
              Op op;
              op.type = OP_E;
              op.block = ind[e];
              op.bin = op.bout = bd[e + 1];
              op.cols = cols;
              int sg[3] = {0, 0, 0};
              for (int x = 0; x < nlines; ++x)
                if (eK[e][x] != 0) { op.lines[op.nlines] = x; op.lsign[op.nlines] = eK[e][x]; sg[op.nlines] = eK[e][x]; op.nlines++; }
              op.etab = get_table(op.block, op.nlines, sg);
              return op;
            };
            for (int v = 1; v < vct0; ++v) {
              pg.ops.push_back(vertex_op(v, bd[1]));
              pg.ops.push_back(edge_op(v, bd[1]));
            }
            {
              Op z;
              z.type = OP_Z;
              z.forward = !dag[0];
              z.set = s;
              z.qs = M.S.size[s];
              z.bin = bd[vct0];
              z.bout = bd[vct0 + 1];
              z.b1 = bd[1];
              z.cols = bd[0];
              // kappa operators: (forward ? F : F^dag)[s] at block b_ix, (b1 x b0) each
              z.fbase = M.F_offset(s, z.forward ? 0 : 1, b_ix);
              z.fhbase = M.Fh_offset(s, z.forward ? 0 : 1, b_ix);
              // mu operators: (forward ? F^dag : F)[s] at block ind[vct0-1], (bvp x bv) each
              z.mubase = M.F_offset(s, z.forward ? 1 : 0, ind[vct0 - 1]);
              if (z.fbase < 0 || z.mubase < 0) throw std::logic_error("missing zero-vertex operator block");
              pg.ops.push_back(z);
            }
            for (int v = vct0 + 1; v < 2 * m; ++v) {
              pg.ops.push_back(edge_op(v - 1, bd[0]));
              pg.ops.push_back(vertex_op(v, bd[0]));
            }
            pg.max_w = bd[1] * bd[1];
            for (auto &op : pg.ops) {
              int win = (op.type == OP_Z) ? op.bin * op.b1 : op.bin * op.cols;
              pg.max_w = std::max(pg.max_w, std::max(win, op.bout * op.cols));
            }
            plan.progs.push_back(std::move(pg));
          }
        }
      }
    }
  }
}

// ----------------------------------------------------------------------------------------------

void build_corr_programs(const Model &M, const PoleSets &ps, const int *topo, int m, const std::vector<int> &G_zero,
                         const CorrOps &co, CorrPlanHost &plan) {
  const Problem &P = M.P;
  const int n = P.n, nb = P.nb;
  auto T = [&](int i, int j) { return topo[2 * i + j]; };
  for (int i = 0; i < m; ++i)
    if (T(i, 0) >= T(i, 1)) throw std::invalid_argument("first row of topology must contain smaller-numbered vertices");
  if (T(0, 0) != 0) throw std::invalid_argument("topology(0,0) must be 0");
  const int vct0 = T(0, 1);
  const int nlines = m - 1;
  std::vector<int> la(std::max(nlines, 1)), lb(std::max(nlines, 1));
  for (int x = 0; x < nlines; ++x) { la[x] = T(x + 1, 0); lb[x] = T(x + 1, 1); }
  std::vector<int> bar_line(2 * m, -1);
  for (int x = 0; x < nlines; ++x) bar_line[lb[x]] = x;
  auto get_table = [&](int block, int nl, const int *sg, int tr) {
    for (size_t i = 0; i < plan.tables.size(); ++i) {
      const auto &tb = plan.tables[i];
      if (tb.block == block && tb.tr == tr && tb.nlines == nl && std::equal(sg, sg + nl, tb.lsign)) return (int)i;
    }
    EdgeTable tb;
    tb.block = block;
    tb.tr = tr;
    tb.nlines = nl;
    for (int i = 0; i < nl; ++i) tb.lsign[i] = sg[i];
    tb.nentries = ipow(P.p, nl);
    tb.offset = plan.ftab_size;
    plan.ftab_size += tb.nentries * 2 * (int64_t)P.r * P.dims[block] * P.dims[block];
    plan.tables.push_back(tb);
    return (int)plan.tables.size() - 1;
  };
  auto zero = [&](int b) { return b < 0 || G_zero[b]; };

  const int64_t o_max = ipow(n, m - 1);
  for (int fb_ix = 0; fb_ix < (1 << nlines); ++fb_ix) {
    // backbone.cpp:253-265
    std::vector<int> fb(m, -1), dag(2 * m, 0), bar(2 * m, 0);
    for (int i = 1, f = fb_ix; i < m; ++i, f /= 2) fb[i] = f % 2;
    dag[0] = 1;
    dag[vct0] = 0;
    for (int i = 1; i < m; ++i) {
      bar[T(i, 1)] = 1;
      dag[T(i, 0)] = (fb[i] != 1);
      dag[T(i, 1)] = (fb[i] == 1);
    }
    for (int64_t o_ix = 0; o_ix < o_max; ++o_ix) {
      std::vector<int> orb(2 * m, -1);
      for (int i = 1, oo = (int)o_ix; i < m; ++i, oo /= n) { orb[T(i, 0)] = oo % n; orb[T(i, 1)] = oo % n; }
      auto conn_w = [&](int w, int ip) { return M.conn(M.S.label[orb[w]], dag[w], ip); };
      // block walks of both sides (block_sparse_backbone.cpp:565-575, 606-616)
      std::vector<std::vector<int>> rind(nb), lind(nb);
      std::vector<int> rend(nb, -1), lend(nb, -1);
      for (int b = 0; b < nb; ++b) {
        std::vector<int> ind(2 * m, -1);
        bool inc = false;
        int ip = b;
        ind[0] = b;
        for (int w = 1; w < vct0 && !inc; ++w) {
          ip = conn_w(w, ip);
          inc = zero(ip);
          if (!inc) ind[w] = ip;
        }
        if (!inc) { rind[b] = ind; rend[b] = ind[vct0 - 1]; }
        std::fill(ind.begin(), ind.end(), -1);
        inc = false;
        ip = b;
        ind[vct0] = b;
        for (int w = vct0 + 1; w < 2 * m && !inc; ++w) {
          ip = conn_w(w, ip);
          inc = zero(ip);
          if (!inc) ind[w] = ip;
        }
        if (!inc) { lind[b] = ind; lend[b] = ind[2 * m - 1]; }
      }
      // contraction terms (block_sparse_backbone.cpp:639-702), independent of the pole tuple
      std::vector<CorrTerm> terms;
      for (int mu = 0; mu < co.nmu; ++mu)
        for (int kap = 0; kap < co.nkap; ++kap)
          for (int b = 0; b < nb; ++b) {
            if (rind[b].empty()) continue;
            int e0 = rend[b];
            if (zero(e0)) continue;
            int e1 = co.mu_conn[(size_t)mu * nb + e0];
            if (zero(e1) || lind[e1].empty()) continue;
            int c = lend[e1];
            if (zero(c)) continue;
            if (co.kap_conn[(size_t)kap * nb + c] != b) continue;
            CorrTerm t;
            t.mu = mu; t.kap = kap; t.b = b; t.e0 = e0; t.e1 = e1; t.c = c;
            terms.push_back(t);
          }
      if (terms.empty()) continue;
      for (int cls = 0; cls < (1 << nlines); ++cls) {
        bool empty = false;
        for (int x = 0; x < nlines; ++x)
          if (ps.get(fb[x + 1], (cls >> x) & 1).empty()) empty = true;
        if (empty) continue;
        // K placement per line (backbone.cpp:86-139)
        std::vector<int> vK(2 * m, 0), vKx(2 * m, -1);
        std::vector<std::vector<int>> eK(2 * m - 1, std::vector<int>(std::max(nlines, 1), 0));
        int psign = 1, pks[3] = {0, 0, 0}, pke[3] = {0, 0, 0};
        for (int x = 0; x < nlines; ++x) {
          bool rb = (cls >> x) & 1;
          bool fwd = fb[x + 1] == 1;
          if (!fwd) psign = -psign;
          if (!rb) {
            vK[la[x]] = fwd ? -1 : 1; vKx[la[x]] = x;
            vK[lb[x]] = fwd ? 1 : -1; vKx[lb[x]] = x;
            pks[x] = fwd ? -1 : 1;
            pke[x] = 1;
          } else {
            for (int j = la[x]; j < lb[x]; ++j) eK[j][x] = fwd ? 1 : -1;
            pks[x] = fwd ? 1 : -1;
            pke[x] = lb[x] - la[x] - 1;
          }
        }
        // vertex v on a walk with block path ind and dimensions bd (F maps ind[v-1] -> ind[v])
        auto vertex_op = [&](int v, const std::vector<int> &ind, int cols, bool transposed) {
          Op op;
          op.type = OP_V;
          const int din = P.dims[ind[v - 1]], dout = P.dims[ind[v]];
          op.bin = transposed ? dout : din;
          op.bout = transposed ? din : dout;
          op.cols = cols;
          op.tr = transposed ? 1 : 0;
          int set = M.S.label[orb[v]], qo = M.S.ind[orb[v]];
          size_t sz = (size_t)din * dout;
          if (bar[v]) {
            op.fbase = (transposed ? M.bar_offset_T(set, dag[v], ind[v - 1]) : M.bar_offset(set, dag[v], ind[v - 1])) + (int64_t)qo * P.p * sz;
            op.fstride_l = (int64_t)sz;
            op.bar_line = bar_line[v];
          } else {
            op.fbase = (transposed ? M.F_offset_T(set, dag[v], ind[v - 1]) : M.F_offset(set, dag[v], ind[v - 1])) + (int64_t)qo * sz;
          }
          if (op.fbase < 0) throw std::logic_error("missing operator block");
          // [beta, tau] vertices take the kernel with the opposite sign (block_sparse_backbone.cpp:492)
          if (vK[v] != 0) { op.kx = vKx[v]; op.ks = transposed ? -vK[v] : vK[v]; }
          return op;
        };
        auto edge_op = [&](int e, const std::vector<int> &ind, int cols, bool transposed) {
          Op op;
          op.type = OP_E;
          op.block = ind[e];
          op.bin = op.bout = P.dims[ind[e]];
          op.cols = cols;
          op.tr = transposed ? 1 : 0;
          int sg[3] = {0, 0, 0};
          for (int x = 0; x < nlines; ++x)
            if (eK[e][x] != 0) { op.lines[op.nlines] = x; op.lsign[op.nlines] = eK[e][x]; sg[op.nlines] = eK[e][x]; op.nlines++; }
          op.etab = get_table(op.block, op.nlines, sg, op.tr);
          return op;
        };
        auto finish = [&](Program &pg) {
          int rows = P.dims[pg.init_block], cols = rows;
          pg.max_w = rows * cols;
          for (auto &op : pg.ops) {
            pg.max_w = std::max(pg.max_w, std::max(op.bin * cols, op.bout * cols));
            rows = op.bout;
          }
          pg.out_rows = rows;
          pg.out_cols = cols;
        };
        CorrGroup g;
        g.fb_ix = fb_ix; g.o_ix = (int)o_ix; g.cls = cls;
        g.prog0 = (int)plan.progs.size();
        std::vector<int> rprog(nb, -1), lprog(nb, -1);
        for (auto &t : terms) {
          if (rprog[t.b] < 0) {  // [tau, 0] side from G_b (block_sparse_backbone.cpp:577-591)
            const auto &ind = rind[t.b];
            Program pg;
            pg.m = m; pg.sigma_block = -1; pg.init_block = t.b; pg.fb_ix = fb_ix; pg.o_ix = (int)o_ix; pg.cls = cls;
            pg.b1 = P.dims[t.b];
            pg.sign = psign;
            for (int x = 0; x < nlines; ++x) { pg.pk_sign[x] = pks[x]; pg.pk_exp[x] = pke[x]; }
            for (int v = 1; v < vct0; ++v) {
              pg.ops.push_back(vertex_op(v, ind, pg.b1, false));
              pg.ops.push_back(edge_op(v, ind, pg.b1, false));
            }
            finish(pg);
            rprog[t.b] = (int)plan.progs.size();
            plan.progs.push_back(std::move(pg));
          }
          if (lprog[t.e1] < 0) {  // [beta, tau] side, transposed (block_sparse_backbone.cpp:618-633)
            const auto &ind = lind[t.e1];
            Program pg;
            pg.m = m; pg.sigma_block = -1; pg.init_block = ind[2 * m - 1]; pg.fb_ix = fb_ix; pg.o_ix = (int)o_ix; pg.cls = cls;
            pg.b1 = P.dims[pg.init_block];
            pg.gtr = 1;
            pg.sign = 1;
            for (int v = 2 * m - 1; v > vct0; --v) {
              pg.ops.push_back(vertex_op(v, ind, pg.b1, true));
              pg.ops.push_back(edge_op(v - 1, ind, pg.b1, true));
            }
            pg.refl = 1;
            finish(pg);
            lprog[t.e1] = (int)plan.progs.size();
            plan.progs.push_back(std::move(pg));
          }
          CorrTerm u = t;
          u.rprog = rprog[t.b];
          u.lprog = lprog[t.e1];
          g.terms.push_back(u);
        }
        g.prog1 = (int)plan.progs.size();
        plan.groups.push_back(std::move(g));
      }
    }
  }
}

} // namespace xca
