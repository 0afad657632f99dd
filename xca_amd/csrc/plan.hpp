// Device plans of the self-energy engine and the launch entry points of the HIP kernels, shared by
// the single-problem ABI (context.cpp) and the batched iteration engine (batch.cpp).
#pragma once
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "ctx.hpp"
#include "engine.hpp"
#include "wide.hpp"

namespace xca::dev {
int num_tile_variants();
int tile_columns(int variant);
hipError_t prepare_kernels(int RB, hipStream_t s);
hipError_t launch_tab(int RB, const TabBatch &B, hipStream_t s);
// a chain launch's queue: two ints (chunk counter, exit counter), zero before the first launch;
// every launch leaves them zero again
hipError_t init_queue(int *q, hipStream_t s);
hipError_t launch_chain(int RB, int variant, const ChainBatch &B, hipStream_t s);
// workgroups of one chain launch resident on the device at once (its grid never exceeds this)
int chain_resident(int RB, int variant);
hipError_t launch_fill(int64_t n, double v, double *p, hipStream_t s);
hipError_t launch_reduce(const int64_t *slot, const int *red_chunks, const I2 *red_range, const int64_t *sig_off, const int *sig_size,
                         int nblocks, int max_size, const double *partial, double *sigma, double scale, hipStream_t s);
hipError_t launch_reduce_assign(const RedBatch &B, int max_size, hipStream_t s);
hipError_t launch_corr(const CorrArgs &A, hipStream_t s);
hipError_t launch_refl(int RB, const ReflArgs &A, hipStream_t s);
int refl_columns();  // columns of one reflection job
hipError_t launch_corr_reduce(const double *partial, int nchunks, int n, double *out, int accumulate, hipStream_t s);
}  // namespace xca::dev

namespace xca {

using namespace xca::dev;

// Programs, chunks and partial slots of one tile width.
struct DevVariant {
  int NT = 0;
  DevBuf<DProg> progs;
  DevBuf<DChunk> chunks;
  DevBuf<int64_t> slot;
  DevBuf<int> red_chunks;
  DevBuf<I2> red_range;
  DevBuf<double> partial;
  DevBuf<double> scratch;  // per-workgroup shared-prefix caches
  int nchunks = 0;
  double ms = -1.0;  // autotuner estimate of the launch time
  std::vector<DProg> h_progs;
  std::vector<DChunk> h_chunks;
  std::vector<int64_t> h_slot;
  std::vector<I2> h_red_range;
  std::vector<int> h_red;
  // one-segment launch descriptor of this plan alone
  DevBuf<ChainArgs> seg;
  DevBuf<int> seg_prefix;
  std::vector<ChainArgs> h_seg;
  std::vector<int> h_seg_prefix;
  // per-chunk launch descriptors (launch mode 1): entry k covers chunk k alone
  DevBuf<ChainArgs> item_seg;
  DevBuf<int> item_prefix;
};

// Everything one (topology, zero-block pattern, flat-index subset) needs on the device: the op list
// and edge tables (independent of the tile width) and one DevVariant per candidate tile width.
struct DevPlan {
  DevBuf<DOp> ops;
  DevBuf<DTabJob> tabjobs;
  DevBuf<double> ftab;
  DevBuf<int> queue;
  int ntabjobs = 0;
  int64_t items = 0;
  double macs = 0;
  int64_t device_bytes = 0;  // device memory held by the plan (all variants)
  int chosen = -1;
  std::vector<std::unique_ptr<DevVariant>> var;  // index = tile variant (null if not applicable)
  std::vector<DOp> h_ops;
  std::vector<DTabJob> h_tab;
  DevBuf<int> tuples;          // pole-tuple lists of explicit-subset programs
  std::vector<int> h_tuples;
  // one-segment edge-table launch descriptor
  DevBuf<TabSeg> tseg;
  DevBuf<int> tseg_prefix;
  std::vector<TabSeg> h_tseg;
  std::vector<int> h_tseg_prefix;
};

// Correlator programs of one batch (a run of groups whose per-item outputs fit the store): chain
// programs and chunks (store mode) and the contraction terms, runs and chunks.
struct CorrBatchDev {
  DevBuf<DProg> progs;
  DevBuf<DChunk> chunks;
  DevBuf<CTerm> terms;
  DevBuf<CRun> runs;
  DevBuf<CChunk> cchunks;
  DevBuf<RJob> rjobs;
  DevBuf<ChainArgs> seg;
  DevBuf<int> seg_prefix;
  int nchunks = 0, ncchunks = 0, nrjobs = 0;
  int64_t store_size = 0;
  std::vector<DProg> h_progs;
  std::vector<DChunk> h_chunks;
};

// Large-block plan: its launch list (edge tables first, then every program's ops and its final
// accumulation into Σ or store into the correlator buffer), the per-item operand offsets and
// prefactors, and the workspace. seg (correlators): launches [0, seg[0]) are the tables and
// [seg[i], seg[i + 1]) the chains of batch i.
struct WidePlan {
  std::vector<WLaunch> launches;
  DevBuf<double> work, ftab, pref, vmax, zperm;
  DevBuf<int64_t> offs;
  std::vector<size_t> seg;
  int64_t items = 0;
  double macs = 0;
  int64_t device_bytes = 0;
};

// One correlator program over its items (pole tuples in store order), stored at out_off of the
// correlator buffer in batch `batch`.
struct WideCorrProg {
  const Program *pg;
  std::vector<std::vector<int>> poles;
  int64_t out_off;
  int batch;
};

// Large-block correlator chains of every batch (wide.cpp).
std::unique_ptr<WidePlan> build_wide_corr(xca_ctx *c, const std::vector<EdgeTable> &tables, int64_t ftab_size,
                                          const std::vector<WideCorrProg> &progs, int nbatch, double *store, int64_t store_n);
// Device bytes of the large-block Sigma plan of one topology (no allocation).
int64_t wide_plan_bytes_estimate(xca_ctx *c, int m, const int *topo);
// Launches [l0, l1) of a large-block plan on the context's stream.
void wide_launch_range(xca_ctx *c, const WidePlan &wp, size_t l0, size_t l1);

// Everything one correlator evaluation (topology, zero-block pattern, flat-index subset, operator
// connections) needs on the device.
struct CorrDevPlan {
  DevBuf<DOp> ops;
  DevBuf<DTabJob> tabjobs;
  DevBuf<double> ftab, store, partial, cval;
  DevBuf<int> queue, cidx;
  DevBuf<TabSeg> tseg;
  DevBuf<int> tseg_prefix;
  std::vector<std::unique_ptr<CorrBatchDev>> batches;
  int ntabjobs = 0, nmu = 0, nkap = 0, variant = 0;
  int64_t items = 0;       // (path, pole tuple) chain evaluations
  double macs = 0;         // canonical multiply-adds of the chains and the contraction
  int64_t device_bytes = 0;
  std::vector<DTabJob> h_tab;
  std::vector<int64_t> cval_src;  // operator value x sits at position cval_src[x] of the call's matrices
  std::unique_ptr<WidePlan> wide;  // blocks above the tile engine's: the chains on the large-block engine
};

// Problem-level DLR matrix of the large-block engine (plain row-major): beta [cf2it | tcf2it].
struct WideCtx {
  DevBuf<double> M2;
};

// Σ over orders 1..max_order of one context in one launch sequence: the edge tables of every plan
// in one launch, the chunks of every plan in one persistent chain launch per tile width, and one
// per-block reduction with the topology factors. Replayed as a HIP graph while the plan set holds.
struct SigmaAll {
  struct Group {
    int RB = 0, variant = 0;
    std::vector<ChainArgs> h_segs;
    std::vector<int> h_prefix;
    DevBuf<ChainArgs> segs;
    DevBuf<int> prefix, queue;
    DevBuf<double> scratch;
    ChainBatch B{};
  };
  std::vector<int> key;
  std::vector<DevPlan *> plans;
  std::vector<double> scales;
  std::vector<TabSeg> h_tab;
  std::vector<int> h_tab_prefix;
  DevBuf<TabSeg> tab;
  DevBuf<int> tab_prefix;
  std::vector<std::unique_ptr<Group>> groups;
  std::vector<RedPlan> h_red;
  DevBuf<RedPlan> red;
  DevBuf<RedProb> prob;
  DevBuf<int> prob_prefix;
  RedBatch R{};
  int max_sig = 0;
  int64_t items = 0;
  double macs = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~SigmaAll() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

DevPlan *get_plan(xca_ctx *c, int m, const int *topo, const int32_t *fix, int64_t nfix);
void accumulate(xca_ctx *c, int m, const int *topo, const int32_t *fix, int64_t nfix, double scale);
void wide_accumulate(xca_ctx *c, int m, const int *topo, const int32_t *fix, int64_t nfix, double scale);
int chunk_tiles(int64_t total_tiles);
void costliest_first(std::vector<DChunk> &chunks, const std::vector<DProg> &progs);
double canonical_macs(const Program &pg, int r);
DOp device_op(const Op &op, const std::vector<EdgeTable> &tables, const xca_ctx *c);
std::vector<DTabJob> tab_jobs(const std::vector<EdgeTable> &tables, const xca_ctx *c);
void corr_release(xca_ctx *c);
ChainArgs chain_args(xca_ctx *c, DevPlan *dp, DevVariant *V);
int ctx_max_block(const xca_ctx *c);
// largest block dimension over the segments of a chain launch
inline int max_block_of(const std::vector<ChainArgs> &segs) {
  int m = 1;
  for (const auto &a : segs) m = std::max(m, a.maxb);
  return m;
}
TabSeg tab_seg(xca_ctx *c, DevPlan *dp);
void autotune(xca_ctx *c, DevPlan *dp);
std::vector<std::pair<int, std::vector<int>>> connected_topologies(int order);

}  // namespace xca
