// Host-side set-up of one pseudo-particle self-energy problem for the device path: symmetry sets,
// barred operators, hybridization/kernel tables and the enumeration of backbone block paths into
// small "programs" that the device kernels interpret.
//
// The decomposition of the reference's loop over flat backbone indices
// (c++/triqs_xca/block_sparse_backbone.cpp:229-337) is re-grouped here:
//   flat index f_ix = o_ix + n^(m-1) (p_ix + p^(m-1) fb_ix)         (backbone.cpp:192-208)
// Block walks depend only on (fb_ix, o_ix, Sigma block, symmetry set), never on the pole tuple
// p_ix, so each walk becomes one "path"; pole tuples are a batch dimension of that path. The
// step-4(a)/4(b) choice of backbone.cpp:86-139 depends on the sign of each pole, so the pole tuples
// of a path split into 2^(m-1) rule classes, each with its own program.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "dlr.hpp"

namespace xca {

// Problem description with real arithmetic (all inputs real).
struct Problem {
  double beta = 0;
  int r = 0, p = 0, n = 0, nb = 0;
  std::vector<double> it, rf;       // DLR nodes: relative times, dimensionless frequencies
  std::vector<double> poles;        // hybridization poles (energy units)
  std::vector<double> coeffs;       // (p, n, n)
  std::vector<int> dims;            // block dimensions
  std::vector<int> c_conn, cdag_conn;                // (n, nb): target block or -1
  std::vector<std::vector<double>> c_mats, cdag_mats;  // [o*nb + j] -> (dims[t], dims[j])
};

// Symmetry sets: orbitals with identical annihilator block connections
// (c++/triqs_xca/atom_diag_utils.cpp:42-68).
struct SymSets {
  std::vector<int> label, ind, size;
  std::vector<std::vector<int>> to_orb;
  int q = 0;
};

// Vertex/edge/zero-vertex operations of one program.
enum OpType : int { OP_V = 0, OP_E = 1, OP_Z = 2 };

struct Op {
  int type = 0;
  int bin = 0, bout = 0;  // rows in / rows out (E: bin == bout == block dim)
  int cols = 0;           // columns of T during this op
  // V: matrix = opmats + fbase + (bar ? line pole * fstride_l : 0); K scaling on line kx with sign ks
  int64_t fbase = 0, fstride_l = 0;
  int bar_line = -1;
  int kx = -1, ks = 0;
  // E: table entry = ftab + ebase + sum_i lines[i] pole * estride[i]  (entry holds f and W f)
  int block = -1;
  int nlines = 0;
  int lines[3] = {-1, -1, -1};
  int lsign[3] = {0, 0, 0};
  int etab = -1;       // index into the table list
  // Z: kappa operators (b1 x b0) at fbase, mu operators (bvp x bv) at mubase, set size qs;
  // fhbase: the kappa sums Fh (r, qs, b1, b0) of the same (set, dag, block) (-1: not tabulated)
  int64_t mubase = 0, fhbase = -1;
  int qs = 0, forward = 0, set = -1, b1 = 0;
  int tr = 0;  // V, E: transposed operator / edge function (the [beta, tau] side of a correlator)
};

struct Program {
  int m = 0;
  int topo_id = 0;
  int sigma_block = 0, b0 = 0, b1 = 0;
  int init_block = 0;
  int fb_ix = 0, o_ix = 0, set = 0;
  int cls = 0;                  // rule class bit mask: bit x set = line x+1 uses step 4(b)
  int sign = 1;                 // diag order sign * prefactor sign
  int pk_sign[3] = {0, 0, 0};   // prefactor K sign per line
  int pk_exp[3] = {0, 0, 0};    // prefactor K exponent per line
  std::vector<Op> ops;
  int max_w = 0;                // max rows*cols over the program
  int gtr = 0;                  // start from the transposed G block
  int refl = 0;                 // reflect the result t -> 1 - t (after the prefactor)
  int out_rows = 0, out_cols = 0;  // shape of the final T
};

// Table of edge functions for one (block, K-line pattern): entries indexed by the poles of the
// lines in the pattern; each entry holds f = vals2coefs(G_b K...) and W f, both (r, b, b).
struct EdgeTable {
  int block = 0;
  int tr = 0;                   // table of G_block^T
  int nlines = 0;
  int lsign[3] = {0, 0, 0};
  int64_t nentries = 1;
  int64_t offset = 0;           // in doubles, into the f-table buffer
};

struct Model {
  Problem P;
  SymSets S;
  dlr::Ops ops;
  std::vector<double> kp, km;   // (p, r): K(t_i, +beta*pole_l), K(t_i, -beta*pole_l)
  std::vector<double> k0p, k0m; // (p)
  std::vector<double> delta, delta_refl;  // (r, n, n)
  std::vector<double> opmats;   // F, F^dag (q_s, bout, bin) and barred (q_s, p, bout, bin) blocks,
                                // then the same blocks transposed ((q_s, bin, bout) and (q_s, p, bin, bout))
  // offsets into opmats: [set][dag][block], -1 if absent
  std::vector<int64_t> off_F, off_bar, off_FT, off_barT;
  // zero-vertex kappa sums Fh(t, mu) = sum_kappa h(t, mu, kappa) F_kappa, (r, q_s, bout, bin) per
  // [set][dag][block] (dag 0: h = Delta(t)[mu][kappa]; dag 1: h = -Delta_refl(t)[kappa][mu]);
  // tabulated only when every block has at most 12 rows (the tile engine's operand)
  std::vector<int64_t> off_Fh;
  std::vector<int> conn_set;    // [set*2+dag][block] target block or -1
  std::vector<int> G_off;       // offsets (in doubles) of G blocks in a packed (r,b,b) array
  int64_t G_size = 0;
  int Nmax = 0;

  void init(const Problem &P);
  int64_t F_offset(int set, int dag, int block) const { return off_F[((size_t)set * 2 + dag) * P.nb + block]; }
  int64_t bar_offset(int set, int dag, int block) const { return off_bar[((size_t)set * 2 + dag) * P.nb + block]; }
  int64_t F_offset_T(int set, int dag, int block) const { return off_FT[((size_t)set * 2 + dag) * P.nb + block]; }
  int64_t bar_offset_T(int set, int dag, int block) const { return off_barT[((size_t)set * 2 + dag) * P.nb + block]; }
  int64_t Fh_offset(int set, int dag, int block) const { return off_Fh[((size_t)set * 2 + dag) * P.nb + block]; }
  int conn(int set, int dag, int block) const { return conn_set[((size_t)set * 2 + dag) * P.nb + block]; }
  int64_t num_backbones(int m) const;
};

// Pole lists per rule: for a forward line step 4(a) takes poles <= 0, for a backward line poles >= 0
struct PoleSets {
  std::vector<int> fwd_a, fwd_b, bwd_a, bwd_b;
  void init(const std::vector<double> &poles_dimless);
  const std::vector<int> &get(int forward, int rule_b) const {
    return forward ? (rule_b ? fwd_b : fwd_a) : (rule_b ? bwd_b : bwd_a);
  }
};

struct Plan {
  std::vector<Program> progs;
  std::vector<EdgeTable> tables;
  int64_t ftab_size = 0;
};

// Enumerate the complete block paths of one topology (m x 2, int) for G with the given zero-block
// flags, appending programs (one per path and non-empty rule class) and edge tables to plan.
void build_programs(const Model &M, const PoleSets &ps, const int *topo, int m, int topo_id, const std::vector<int> &G_zero,
                    Plan &plan);

// ---------------------------------------------------------------------------------------------
// One-time correlators <T O_mu(tau) O_kap(0)> (c++/triqs_xca/block_sparse_backbone.cpp:526-706,
// backbone.cpp:239-292). Vertex 0 carries O_kap, vertex vct0 = topology(0, 1) carries O_mu, and
// the flat index f_ix = o_ix + n^(m-1) (p_ix + p^(m-1) fb_ix) runs over 2^(m-1) direction patterns
// of the lines 1..m-1. For one (fb_ix, o_ix, rule class) "group" and every pole tuple:
//   right[b] = the [tau, 0] side started from G_b (vertices 1..vct0-1), times the prefactor;
//   left[b]  = the [beta, tau] side started from G at the path's end (vertices 2m-1..vct0+1,
//              multiplied from the right), then reflected; it is built transposed, as a left
//              product with transposed operators and edge functions;
//   C(t, mu, kap) += Tr[left[e1](t) O_mu right[b](t) O_kap] for e0 = end of right[b],
//              e1 = O_mu(e0), c = end of left[e1] and O_kap(c) = b.

// Block operators: target block per source block (-1: none), matrices at offsets into `mats`.
struct CorrOps {
  int nmu = 0, nkap = 0;
  std::vector<int> mu_conn, kap_conn;    // (nops, nb)
  std::vector<int64_t> mu_off, kap_off;  // (nops, nb), -1 if absent
  std::vector<double> mats;
};

struct CorrTerm {
  int mu = 0, kap = 0, b = 0, e0 = 0, e1 = 0, c = 0;
  int rprog = 0, lprog = 0;  // indices into CorrPlanHost::progs
};

struct CorrGroup {
  int fb_ix = 0, o_ix = 0, cls = 0;
  int prog0 = 0, prog1 = 0;  // programs [prog0, prog1) of the group (right and left)
  std::vector<CorrTerm> terms;  // sorted by (mu, kap)
};

struct CorrPlanHost {
  std::vector<Program> progs;
  std::vector<EdgeTable> tables;
  int64_t ftab_size = 0;
  std::vector<CorrGroup> groups;
};

void build_corr_programs(const Model &M, const PoleSets &ps, const int *topo, int m, const std::vector<int> &G_zero,
                         const CorrOps &co, CorrPlanHost &plan);

} // namespace xca
