// Device descriptors of the Dyson solve and of the self-consistency update, shared by the kernels
// (dyson.hip) and the host orchestration (ppsc.cpp, batch.cpp).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

namespace xca::dev {

// Launches whose largest padded system (r n rounded up to 16) has at most this many rows keep every
// block's system in LDS and never touch DysonArgs::sys.
constexpr int kDysonLdsSysMax = 112;

// Dyson solve of every block of one problem.
struct DysonArgs {
  int r, nb;
  double beta;
  const double *eta;        // device scalar: pseudo-particle chemical potential
  const int *dims;          // (nb)
  const int64_t *goff;      // (nb) offsets of (r, n, n) blocks
  const int64_t *soff;      // (nb) offsets of the Np x Np system matrices (Np = r n rounded up to 16)
  const int64_t *poff;      // (nb) offsets of the Np pivots of each block
  const int64_t *roff;      // (nb) offsets of the Np x n right-hand sides
  const double *lu;         // (r, r) LU factors of cf2it
  const int *piv;           // (r)
  const double *cf2it, *tcf2it, *W;           // (r, r)
  const double *G0, *g0c, *Wg0c;              // packed blocks
  const double *sigma;                        // packed blocks
  double *ggs, *wf, *sys, *rhs, *gout;       // scratch / output
  int *ipiv;                                  // pivots
  const int *large;                           // (nb) 1: block solved by the large-block path (skipped here), or null
};

// Many problems: global block k belongs to problem s with prefix[s] <= k < prefix[s + 1].
struct DysonBatch {
  const DysonArgs *probs;
  const int *prefix;
  int nprobs, total;
  int np_max;               // largest padded system dimension among the blocks (LDS panel rows)
  int mode = 0;             // dyson_kernel: 0 the whole solve; 1 up to the assembled system; 2 from the
                            // factorised system (set by launch_dyson)
};

// Launches whose largest padded system exceeds this factorise every block's system across
// workgroups: one launch per 16-column panel (one workgroup per block) and one per trailing update
// (one workgroup per 16-column strip of every block), between a launch that assembles the systems
// and one that solves with the factors.
constexpr int kDysonMultiNp = 128;

// Layout of the per-problem iteration status (doubles) read back after every iteration.
constexpr int kStDiff = 0;   // max |G_new - G| (bits of a non-negative double, atomic max)
constexpr int kStEta = 1;    // pseudo-particle chemical potential
constexpr int kStZ = 2;      // Z = -Tr G_new(beta) before the normalisation
constexpr int kStDeta = 3;   // ln|Z| / beta of this iteration
constexpr int kStConv = 4;   // 1 when this iteration met the tolerance (G is then left unchanged)
constexpr int kStBmax = 5;   // (nb) max |G| per block after the update
constexpr int status_stride(int nb) { return ((kStBmax + nb + 7) / 8) * 8; }

// Normalisation, convergence test and mixing of one problem (block_sparse_solver.py:295-346,427-449).
struct IterArgs {
  int r, nb, ndiag, pad0;
  double beta, tol, mix;
  int64_t total;            // size of G
  const int *dims;
  const int64_t *goff, *gend;
  const double *lu;
  const int *piv;
  const double *zw;         // (r) G(beta) = sum_t zw[t] G(t): zw = cf2it^{-T} K(1, w) (DLR evaluation at tau = beta)
  const double *tau;        // (r) absolute times
  double *G;                // current iterate
  double *Gnew;             // Dyson solution, normalised in place
  double *status;           // status_stride(nb) doubles
};

struct IterBatch {
  const IterArgs *probs;
  const int *prefix;        // (problem, block) prefix for the per-block kernels
  int nprobs, total_blocks;
  int64_t max_total;        // largest G among the problems
};

hipError_t launch_dyson(const DysonBatch &B, hipStream_t s);
// Z, eta and the rescaling of G_new with max |G_new - G| (then launch_block_max)
hipError_t launch_iterate(const IterBatch &B, hipStream_t s);
// the mixing (unless converged) and the block maxima of the new G with the convergence flag
hipError_t launch_block_max(const IterBatch &B, hipStream_t s);
hipError_t launch_fill(int64_t n, double v, double *p, hipStream_t s);

}  // namespace xca::dev
