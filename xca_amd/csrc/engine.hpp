// Device data structures shared by the host orchestration and the HIP kernels.
#pragma once
#include <cstdint>

namespace xca::dev {

constexpr int kMaxLines = 3;   // orders up to 4
constexpr int kMaxSet = 8;     // largest symmetry set handled by the zero-vertex op

// One vertex/edge/zero-vertex operation (see host/model.hpp for the semantics).
struct DOp {
  int type, bin, bout, cols;
  int64_t fbase, fstride_l;      // V: operator matrix, bar stride per pole; Z: kappa sums Fh (r, qs, b1, b0)
  int bar_line, kx, ks, nlines;  // V: bar line, K line and sign; E: number of K lines
  int lines[kMaxLines];
  int pad0;
  int64_t ebase;                 // E: edge-table entry 0 offset (doubles)
  int64_t estride[kMaxLines];    // E: entry stride per line pole (doubles)
  int64_t mubase;                // Z: mu operators
  int qs, forward, b1, set;      // Z
  int orb[kMaxSet];              // Z: orbitals of the set
};

struct DProg {
  int op0, nops, sigma_block, b0;
  int b1, g_init, sign, nlines;
  int pk_sign[kMaxLines], pk_exp[kMaxLines];
  int set_off[kMaxLines], set_len[kMaxLines];  // pole lists per line (into the pole-list array)
  int ntup;                                     // pole tuples (items) of the program
  int items_per_tile, max_w;
  int split;      // ops [0, split) depend on the first line's pole only (0: nothing shared)
  int share_len;  // consecutive items with the same first-line pole
  int gtr;        // start from the transposed G block
  int tup_off = -1;  // >= 0: item u of the program is pole tuple tuples[tup_off + u] (an explicit
                     // flat-index subset); < 0: item u is tuple u
  int64_t out_off;  // < 0: accumulate into the chunk's Sigma slot; else store item u's T (r, rows,
                    // cols) at store + out_off + u r rows cols
};

// A run of consecutive tiles [t0, t1) of one program; tile t holds items
// [t * items_per_tile, min((t + 1) * items_per_tile, ntup)). The chunk's results go to its own
// partial slot (offset in doubles).
struct DChunk {
  int prog, t0, t1, pad0;
  int64_t slot;
};

// Edge-function table job: entries [e0, e0 + ne) of one table
struct DTabJob {
  int block, b, nlines, e0;
  int ne, g_off, tr, pad1;  // tr: table of the transposed G block
  int lsign[kMaxLines], pad2;
  int64_t out_off;   // doubles, entry 0 of this table
};

struct I2 {
  int x, y;
};

// DLR operator fragments for the FP64 MFMA path, packed in lane order (see engine.hip).
struct DlrFrag {
  const double *LiP;   // (RB, KS, 64): A fragments of L^{-1} P, KS = RP/4
  const double *W;     // (RB, KS, 64)
  const double *M2;    // (RB, 2*KS, 64): beta [cf2it | tcf2it]
  const double *AU;    // upper-triangular trailing-update fragments, per (step, row block)
  const double *Dinv;  // (RP/4, 16): inverses of the 4x4 diagonal blocks of U
  const double *Upk;   // -U strictly above the diagonal, packed by rows (RP (RP - 1) / 2)
  const double *E;     // (RB, KS, 64): K(-t_i, om_k), values at the reflected nodes from coefficients
  int r, RP;
};

struct ChainArgs {
  DlrFrag D;
  const DOp *ops;
  const DProg *progs;
  const DChunk *chunks;
  const int *plist;            // pole lists
  const double *G, *opmats, *ftab, *kp, *km, *kpow;
  double *partial;
  double *store;               // per-item outputs of store-mode programs
  const int *tuples;           // tuple lists of explicit-subset programs (DProg::tup_off)
  int nchunks;
  int r, p, maxexp;
  int maxb, pad1;              // largest block dimension of the problem
};

// One chain launch drains the chunks of several segments (plans, possibly of different problems
// with the same padded DLR rank): global chunk k belongs to the segment s with
// prefix[s] <= k < prefix[s + 1].
struct ChainBatch {
  const ChainArgs *segs;
  const int *prefix;  // (nsegs + 1)
  int nsegs, total;
  int *queue;         // chunk counter (reset before each launch)
  double *scratch;    // per-workgroup cache of a shared prefix, scratch_stride doubles each
  int64_t scratch_stride;
  int maxb = 12;      // largest block dimension of the segments' problems (selects the kernel instance)
  int pad0 = 0;
};

// Edge-table work of one plan.
struct TabSeg {
  DlrFrag D;
  const DTabJob *jobs;
  const double *G, *kp, *km;
  double *ftab;
  int p, pad0;
};

struct TabBatch {
  const TabSeg *segs;
  const int *prefix;  // job prefix (nsegs + 1)
  int nsegs, total;
};

// Sigma of one problem = sum over its plans, in order, of scale * (sum over the plan's chunks of
// the partial slots), written (not accumulated) into the problem's Sigma buffer.
struct RedPlan {
  const int64_t *slot;
  const int *red_chunks;
  const I2 *red_range;
  const double *partial;
  double scale;
};

struct RedProb {
  const RedPlan *plans;
  const int64_t *sig_off;
  const int *sig_size;
  double *sigma;
  int nplans, nb;
};

struct RedBatch {
  const RedProb *probs;
  const int *prefix;  // (problem, block) prefix over problems (nprobs + 1)
  int nprobs, total;
};

// One-time correlator contraction. A term adds Tr[L(t) O_mu R(t) O_kap] for every pole tuple of its
// group, with R (d0 x db) the [tau, 0] side, L^T (d1 x dc) the reflected [beta, tau] side stored
// transposed, O_mu (d1 x d0) and O_kap (db x dc). Stored items are (rows, cols, r) with the time
// index fastest. Operators are sparse rows: row pointers (rows + 1) then column indices in cidx,
// values in cval.
struct CTerm {
  int64_t r_off, l_off;      // item 0 of the right / left program in the store (doubles)
  int mu_csr, kap_csr;       // row pointers of O_mu / O_kap in cidx
  int mu_val, kap_val;       // their values in cval
  int d0, d1, db, dc;
};

// Terms [t0, t1) of one group that add into output out = mu * nkap + kap.
struct CRun {
  int out, t0, t1, pad0;
};

// A run of pole tuples [k0, k1) of one group with runs [run0, run1): one workgroup each.
struct CChunk {
  int run0, run1, k0, k1;
};

// Columns [c0, c0 + ncol) of a stored program output at base; column c holds r doubles at
// base + c r (item c / w, element c % w).
struct RJob {
  int64_t base;
  int w, c0, ncol, pad0;
};

struct ReflArgs {
  DlrFrag D;
  const RJob *jobs;
  double *store;
  int njobs, pad0;
};

// Large-block engine, node-space part of an edge: G = vals2coefs(X), WG = W G for columns
// [0, ncol) of every item (r rows of ld doubles; item strides xsi, gsi, wsi).
struct WV2C {
  DlrFrag D;
  const double *X;
  double *G, *WG;
  int64_t xsi, gsi, wsi;
  int ld, ncol, nitem, pad0;
};

struct CorrArgs {
  const CTerm *terms;
  const CRun *runs;
  const CChunk *chunks;
  const double *store, *cval;
  const int *cidx;
  double *partial;   // (nchunks, r, nout)
  int nchunks, r, nout;
  int large;         // 1: blocks of any dimension (one product element at a time), 0: at most kMaxB
};

}  // namespace xca::dev
