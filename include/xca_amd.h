/*
 * xca_amd — C ABI of the MI355X pseudo-particle self-energy engine.
 *
 * Drop-in boundary for triqs_xca's DiagramEvaluator (c++/triqs_xca/block_sparse_backbone.hpp:23-165)
 * and dyson_it_ppsc (c++/triqs_xca/dlr_dyson_ppsc.hpp:45-230). Plain pointers and sizes only; no
 * exception crosses the ABI: every entry point returns a status (0 = success, < 0 = error) or a
 * handle (NULL on error), and xca_last_error() returns the message of the last failure on the
 * calling thread, with the reference's message text where the reference raises.
 *
 * Array layouts follow nda C order as in the reference: Green's-function blocks (r, b, b) indexed
 * t*b*b + i*b + j, hybridization coefficients (p, n, n), topology (m, 2) int32. Time nodes are in
 * relative format (t in [0, 1/2] as t, t in (1/2, 1] as t - 1). Poles are in energy units and are
 * multiplied by beta inside, as hyb.cpp:18 does. All device arithmetic is real FP64.
 */
#ifndef XCA_AMD_H
#define XCA_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failing call on this thread ("" if none). */
const char *xca_last_error(void);

/* DLR grid for cutoff lambda = beta * w_max and tolerance eps. Writes r real-frequency nodes
 * (dimensionless) and r imaginary-time nodes (relative format) and returns r, or < 0 if r > max_r.
 * Replaces cppdlr::build_dlr_rf + imtime_ops node selection used at
 * c++/triqs_xca/block_sparse_backbone.cpp:68-70 and python/triqs_xca/block_sparse_solver.py:105,136. */
int xca_dlr_build(double lambda, double eps, int max_r, double *rf_out, double *it_out);

/* Fermionic kernel K(t, om) = -exp(-om t)/(1 + exp(-om)), t in relative format (cppdlr::k_it as
 * called at c++/triqs_xca/hyb.cpp:37-41). */
double xca_dlr_kernel(double t, double om);

/* One impurity problem: DLR grid, hybridization poles/coefficients and the block-sparse
 * annihilation/creation operators in the Fock basis (what get_operators builds from atom_diag,
 * c++/triqs_xca/atom_diag_utils.cpp:39-185). c_conn[o*nblocks + j] is the block that c_o maps block
 * j to (-1: none); c_mats concatenates the (dims[c_conn], dims[j]) matrices in (o, j) order over the
 * connected j. The same for cdag. */
typedef struct xca_problem {
  double beta;
  int r;
  const double *it;      /* (r) relative times */
  const double *rf;      /* (r) dimensionless real frequencies */
  int p;
  const double *poles;   /* (p) energy units */
  const double *coeffs;  /* (p, n, n) */
  int n;                 /* number of flavours */
  int nblocks;
  const int *dims;       /* (nblocks) */
  const int *c_conn;     /* (n, nblocks) */
  const int *cdag_conn;  /* (n, nblocks) */
  const double *c_mats;
  const double *cdag_mats;
  /* Optional (3): the extents of coeffs as the caller holds it. NULL means (p, n, n). When given,
   * they are checked against (p, n, n) before coeffs is read, with the message of the reference's
   * Fq-only constructor check (c++/triqs_xca/block_sparse_backbone.cpp:28-40): "hyb_coeffs must have
   * shape (p, n, n) with p = hyb_poles.size() = .. and n = sum(Fq.sym_set_sizes) = .., got (..)". */
  const int64_t *coeffs_shape;
} xca_problem;

typedef struct xca_ctx xca_ctx;

/* Build an evaluator on GPU `device` and upload operators, barred operators, hybridization and
 * kernel tables; they stay resident until xca_ctx_destroy. Replaces
 * DiagramEvaluator(beta, Lambda, eps, hyb_poles, hyb_coeffs, Fq)
 * (c++/triqs_xca/block_sparse_backbone.cpp:63-85). NULL on error (shape errors, cross-set
 * hybridization: "couples different symmetry sets", no GPU). */
xca_ctx *xca_ctx_create(const xca_problem *desc, int device);
void xca_ctx_destroy(xca_ctx *ctx);

/* get_num_self_energy_backbones (c++/triqs_xca/block_sparse_backbone.cpp:721-724):
 * 2^m n^(m-1) p^(m-1). */
int64_t xca_num_backbones(xca_ctx *ctx, int m);
/* Number of symmetry sets found from the annihilator connections. */
int xca_num_sym_sets(xca_ctx *ctx);
/* Number of device plans the context holds. Plans are keyed by topology and by the zero-block pattern
 * of G (zero blocks prune paths, block_sparse.cpp:45-49); the plans of at most 4 patterns are kept,
 * the least recently used pattern's released first. No reference counterpart (the reference walks
 * block paths anew on every call). */
int64_t xca_num_cached_plans(xca_ctx *ctx);

/* compute_self_energy(G, topology[, f_ix_vec]) (c++/triqs_xca/block_sparse_backbone.cpp:369-459):
 * G and sigma_out are the concatenated (r, b, b) blocks (host memory). fix == NULL evaluates every
 * flat backbone index; otherwise the nfix listed ones (error "f_ix out of range" if any is outside
 * [0, xca_num_backbones)). The result carries the reference's per-backbone signs but not the
 * order/topology prefactor applied by the Python caller (block_sparse_solver.py:618,636). Blocks whose
 * max |.| is below 1e-16 are returned as zero, as compute_self_energy returns them. (With an explicit
 * list the reference also re-flags such blocks after every flat index, so that a running sum that
 * falls below 1e-16 part-way is dropped at the block's next contribution; this is not mirrored, a
 * difference of at most 1e-16 absolute per such event.) */
int xca_sigma(xca_ctx *ctx, const double *G, int m, const int *topo, const int32_t *fix, int64_t nfix, double *sigma_out);

/* get_num_single_ptcle_gf_backbones (c++/triqs_xca/block_sparse_backbone.cpp:732-735):
 * 2^(m-1) n^(m-1) p^(m-1), the flat indices of CorrelatorBackbone (backbone.cpp:239-292). */
int64_t xca_num_correlator_backbones(xca_ctx *ctx, int m);

/* compute_single_ptcle_gf(G, topology[, f_ix_vec]) (c++/triqs_xca/block_sparse_backbone.cpp:779-825,
 * eval_correlator :526-706 with O_mu = c_mu, O_kap = c^dag_kap per orbital, :737-777): out is
 * (r, n, n) real, indexed by orbital. fix == NULL evaluates every flat index; otherwise the nfix listed
 * ones (error "f_ix out of range" outside [0, xca_num_correlator_backbones); the reference decodes such
 * an index silently into another backbone). Carries the per-backbone signs but not the order/topology
 * factor (-1)^order s_topo applied by the caller (block_sparse_solver.py:667-677). G == NULL uses the
 * G last uploaded (xca_set_G, xca_sigma or a previous call). */
int xca_spgf(xca_ctx *ctx, const double *G, int m, const int *topo, const int32_t *fix, int64_t nfix, double *out);

/* compute_one_time_correlator(G, ops_tau, ops_0, topology, f_ix_vec)
 * (c++/triqs_xca/block_sparse_backbone.cpp:850-876): <T O_mu(tau) O_kap(0)> diagrams for nmu operators
 * at tau and nkap at 0, each a block operator: conn[(op, j)] is the block it maps block j to (-1: none)
 * and mats concatenate the (dims[conn], dims[j]) matrices in (op, j) order over the connected j (the
 * layout of xca_problem's c_mats). out is (r, nmu, nkap) real. G == NULL as for xca_spgf. */
int xca_correlator(xca_ctx *ctx, const double *G, int m, const int *topo, const int32_t *fix, int64_t nfix, int nmu, const int *mu_conn,
                   const double *mu_mats, int nkap, const int *kap_conn, const double *kap_mats, double *out);

/* Device-resident form for benchmarking and the iteration loop: upload G once, evaluate any number
 * of topologies into the device Sigma accumulator with weight `scale` (Sigma += scale * S_topo),
 * then download. xca_sigma_zero clears the accumulator. */
int xca_set_G(xca_ctx *ctx, const double *G);
int xca_sigma_zero(xca_ctx *ctx);
int xca_sigma_accumulate(xca_ctx *ctx, int m, const int *topo, double scale);
int xca_get_sigma(xca_ctx *ctx, double *sigma_out);
int xca_synchronize(xca_ctx *ctx);

/* The device Sigma accumulator <- Sigma of the uploaded G summed over orders 1..max_order and their
 * connected topologies with the solver's net factor -s_topo
 * (python/triqs_xca/block_sparse_solver.py:597-640, the loop of eval_pseudo_particle_self_energy),
 * in one launch sequence: the edge tables of every topology in one launch, every topology's chunks in
 * one persistent chain launch, one per-block reduction. The sequence is captured once as a HIP graph
 * and replayed while the zero-block pattern is unchanged; timed != 0 launches it eagerly with the
 * chain launches bracketed by events for xca_last_stats. Asynchronous: xca_get_sigma or
 * xca_synchronize waits for it. Equal, bit for bit, to xca_sigma_zero followed by
 * xca_sigma_accumulate over the same topologies with the same factors. */
int xca_sigma_order(xca_ctx *ctx, int max_order, int timed);

/* Pseudo-particle Dyson solver: one dyson_it_ppsc(beta, itops, G0_block) per block
 * (c++/triqs_xca/dlr_dyson_ppsc.hpp:92-117, constructed at python/triqs_xca/block_sparse_solver.py:140).
 * G0: concatenated (r, b, b) blocks of the atomic pseudo-particle propagator. */
int xca_dyson_init(xca_ctx *ctx, const double *G0);

/* dyson_it_ppsc::solve(sigma, eta) for every block followed by hermitization
 * (dlr_dyson_ppsc.hpp:135-168, block_sparse_solver.py:572-590). sigma == NULL takes the device Sigma
 * accumulator. G_out: concatenated blocks (host). */
int xca_dyson_solve(xca_ctx *ctx, const double *sigma, double eta, double *G_out);

/* dyson_it_ppsc::solve for every block, as xca_dyson_solve, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): it allocates nothing. No counterpart
 * in the reference. G0 is passed here instead of through xca_dyson_init (the G0 coefficients are
 * formed on the host); every device buffer the Dyson kernels use (descriptors, DLR factors, the
 * block systems, right-hand sides, pivots, staging) lives in one static device array of 32 MiB per
 * GPU, held until the process ends. A problem needing more is rejected ("static maximum"); 3-orbital
 * problems with systems up to r b = 528 need about 11 MiB. The array keeps the last call's contents
 * (nothing in it must start at zero); calls that overlap on one GPU, from several threads or
 * streams, race. */
int xca_dyson_solve_noalloc(xca_ctx *ctx, const double *G0, const double *sigma, double eta, double *G_out);

/* BlockSparseSolver.solve(max_order, tol, maxiter, mix) with classic normalization
 * (python/triqs_xca/block_sparse_solver.py:240-355, 427-449): Sigma over orders 1..max_order -> Dyson
 * -> Z = -Tr G(beta), eta += ln|Z|/beta, G_new *= exp(-tau deta) -> stop when max|G_new - G| < tol,
 * else G = mix G_new + (1 - mix) G. Starts from the G0 given to xca_dyson_init and pseudo-particle
 * chemical potential eta0. Outputs the solver's G, eta, the iteration count and the last difference.
 * Each iteration runs as one captured HIP graph (set XCA_NO_GRAPH to launch eagerly); a problem
 * with blocks above 12 rows runs each iteration as launches on the large-block engine and the
 * device-wide Dyson path, with one status read per iteration. */
int xca_ppsc_solve(xca_ctx *ctx, double eta0, int max_order, double tol, int maxiter, double mix, double *G_out, double *eta_out,
                   int *iters_out, double *diff_out);
/* The per-iteration report of the last xca_ppsc_solve (what BlockSparseSolver.solve prints each
 * iteration, block_sparse_solver.py:336-341): rows of (diff_G = max |G_new - G|, Z = -Tr G_new(beta)
 * before the normalisation, eta) for at most max_iters iterations; returns the iteration count. */
int xca_ppsc_history(xca_ctx *ctx, double *out, int max_iters);

/* Serving: many impurity problems solved to self-consistency on one GPU with continuous batching.
 * Every admitted problem keeps its operators, G, Sigma, plans and Dyson state resident; one
 * xca_server_step runs one iteration of every admitted problem in shared batched launches (captured
 * as a HIP graph while the set of problems is unchanged), retires the converged ones and admits
 * queued problems in arrival order as batch slots and the device-memory budget allow. A problem's
 * arithmetic does not depend on the other problems in the batch, so its result is the one
 * xca_ppsc_solve gives with the same tile variant.
 * No reference counterpart: the reference runs one BlockSparseSolver per process
 * (python/triqs_xca/block_sparse_solver.py:240-355). */
typedef struct xca_server xca_server;
/* mem_budget_bytes <= 0: 80 % of the free device memory. tile_variant: 0, 1, 2 (fixed, so results do
 * not depend on timing). */
xca_server *xca_server_create(int device, int64_t mem_budget_bytes, int max_active, int tile_variant);
void xca_server_destroy(xca_server *sv);
/* Queue a problem (G0: its atomic propagator blocks, eta0 the initial pseudo-particle chemical
 * potential); returns the request id, or < 0 on error (also when its estimated footprint exceeds the
 * budget). */
int64_t xca_server_submit(xca_server *sv, const xca_problem *desc, const double *G0, double eta0, int max_order, double tol, int maxiter,
                          double mix);
/* Admit, iterate once, retire. Outputs the number of active and queued problems and of problems that
 * finished in this step. */
int xca_server_step(xca_server *sv, int *n_active, int *n_queued, int *n_finished);
/* Ids of requests finished since the last poll (at most max_ids); returns how many. */
int xca_server_poll(xca_server *sv, int64_t *ids, int max_ids);
/* Result of a finished request (G blocks, eta, iterations, last difference, converged flag and,
 * optionally, 3 times in seconds: queued, admitted-to-finished, submitted-to-finished); releases it. */
int xca_server_result(xca_server *sv, int64_t id, double *G_out, double *eta_out, int *iters_out, double *diff_out, int *converged,
                      double *times_out);
/* Estimated device bytes of a submitted request: an upper bound of what it holds once admitted
 * (resident tables, Dyson state, and its plans at G0's block pattern for every tile variant they
 * allocate). The shared per-workgroup launch buffers are reserved from the budget at creation. */
int64_t xca_server_request_bytes(xca_server *sv, int64_t id);
/* Device bytes an admitted, unfinished request holds now (-1 otherwise). */
int64_t xca_server_request_device_bytes(xca_server *sv, int64_t id);
/* Bytes of mem_budget_bytes reserved at creation for the launch buffers all requests share; requests
 * are admitted against the rest. */
int64_t xca_server_shared_bytes(xca_server *sv);
/* Iterations run, device bytes held by admitted requests, and the budget. */
int xca_server_stats(xca_server *sv, int64_t *ticks, int64_t *bytes_in_use, int64_t *budget);

/* Tile width of the chain kernel: variant 0, 1, 2 = 128, 64, 32 columns per tile; -1 (the default)
 * lets the autotuner time every applicable variant on each new plan and keep the fastest. The
 * environment variable XCA_TILE_VARIANT sets the initial value. Changing it drops cached plans.
 * No reference counterpart (the reference has no device tiles). */
int xca_set_tile_variant(xca_ctx *ctx, int variant);
/* Launch form of xca_sigma / xca_sigma_accumulate on blocks of at most 12 rows: 0 (the default)
 * drains every chunk of a plan in one persistent chain launch; 1 launches the chain kernel once per
 * chunk, and a call with an explicit flat-index subset has one chunk per (path, pole tuple) item, so
 * mode 1 with a subset is the one-backbone-term-per-launch form. Results are identical in both
 * modes. xca_sigma_order always uses the combined launch. No reference counterpart (a GPU launch
 * form; the reference evaluates backbone terms one after another on the CPU). */
int xca_set_launch_mode(xca_ctx *ctx, int mode);
/* Variant and columns per tile used by the last accumulate call. */
int xca_last_tile(xca_ctx *ctx, int *variant, int *columns);

/* Statistics of the last accumulate call: number of (path, pole tuple) items evaluated, canonical
 * multiply-add count, and the device time of the chain kernel in milliseconds. */
int xca_last_stats(xca_ctx *ctx, int64_t *items, double *macs, double *kernel_ms);

#ifdef __cplusplus
}
#endif

#endif
