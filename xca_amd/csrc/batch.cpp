// Batched pseudo-particle self-consistency on one GPU.
//
// One iteration ("tick") of every active problem is a fixed sequence of batched launches:
// edge tables of every plan, the chain kernels (all chunks of all plans in one persistent queue per
// tile width), the per-problem Sigma reduction, the Dyson build and solve of every block, the
// normalisation / convergence / mixing update and the block maxima of the new G, followed by one
// copy of the per-problem status words to pinned host memory. The launch sequence depends only on
// the set of (problem, plan) pairs, so it is captured once as a HIP graph and replayed until a
// problem joins or leaves the batch.
//
// Per problem this follows BlockSparseSolver.solve with classic normalisation
// (python/triqs_xca/block_sparse_solver.py:240-355,427-449): Sigma over orders 1..max_order with the
// net factor -s per topology (:597-640), Dyson per block with hermitisation (:572-590),
// Z = -Tr G(beta), eta += ln|Z|/beta, G_new *= exp(-tau deta), stop when max|G_new - G| < tol, else
// mix. Every problem's arithmetic is independent of the other problems in the batch: chunks write
// their own partial slots, reductions run in a fixed order per problem, Dyson per block.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../../include/xca_amd.h"
#include "ctx.hpp"
#include "dyson.hpp"
#include "host/errors.hpp"
#include "plan.hpp"
#include "ppsc.hpp"

namespace xca {

xca_ctx *ctx_init_host(const xca_problem *desc);
void ctx_init_device(xca_ctx *c, int device);
int64_t ctx_bytes(const xca_ctx *c);
int64_t ctx_device_bytes(const xca_ctx *c);
int64_t plan_bytes_estimate(xca_ctx *c, int m, const int *topo);

constexpr int kStatusStride = 72;  // doubles per problem status slot (blocks <= 64)

// Descriptors of the Dyson and update launches of one problem whose Sigma runs on the large-block
// engine (a batch of one each).
struct WideIter {
  DevBuf<DysonArgs> dyson;
  DevBuf<IterArgs> iter;
  DevBuf<int> prefix;
  DysonBatch DB{};
  IterBatch IB{};
};

struct BatchItem {
  xca_ctx *c = nullptr;
  bool wide = false;            // blocks above 12 rows (or XCA_WIDE): its iteration runs on its own stream
  std::unique_ptr<WideIter> wi;
  int max_order = 1, maxiter = 100;
  double tol = 1e-10, mix = 1.0, eta0 = 0.0;
  int iters = 0;
  bool done = false, converged = false;
  double eta = 0.0, diff = 0.0;
  int slot = -1;
  uint64_t serial = 0;         // unique per admission into a batch (identifies the item in the plan key)
  std::vector<double> bmax;    // block maxima of the current G (zero-block flags of the next Sigma)
  std::vector<int> zero_seen;  // zero-block pattern of the plans in use
  std::vector<DevPlan *> plans;
  std::vector<double> scales;
  std::vector<double> history;  // per iteration: max |G_new - G|, Z before the normalisation, eta
};

class Batch {
 public:
  Batch(int device, hipStream_t stream, int capacity) : device_(device), stream_(stream), cap_(capacity) {
    HIP_CHECK(hipSetDevice(device));
    status_.alloc((size_t)cap_ * kStatusStride);
    HIP_CHECK(hipHostMalloc((void **)&host_status_, sizeof(double) * (size_t)cap_ * kStatusStride, hipHostMallocDefault));
    free_slots_.clear();
    for (int i = cap_ - 1; i >= 0; --i) free_slots_.push_back(i);
  }
  ~Batch() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graph();
    for (auto &q : queues_) q.reset();
    if (host_status_) (void)hipHostFree(host_status_);
  }

  bool full() const { return free_slots_.empty(); }
  int size() const { return (int)items_.size(); }

  // Start a problem: G = G0, eta = eta0, status cleared.
  void add(BatchItem *it) {
    if (free_slots_.empty()) throw std::runtime_error("xca: batch is full");
    if (it->c->M.P.nb + kStBmax > kStatusStride) throw std::invalid_argument("xca: more than 64 blocks");
    require_dyson(it->c);
    it->slot = free_slots_.back();
    free_slots_.pop_back();
    it->serial = ++serial_;
    xca_ctx *c = it->c;
    HIP_CHECK(hipMemcpyAsync(c->G.ptr, c->dyson->G0.ptr, c->M.G_size * sizeof(double), hipMemcpyDeviceToDevice, stream_));
    HIP_CHECK(launch_fill(kStatusStride, 0.0, status_.ptr + (size_t)it->slot * kStatusStride, stream_));
    HIP_CHECK(launch_fill(1, it->eta0, status_.ptr + (size_t)it->slot * kStatusStride + kStEta, stream_));
    // zero-block flags of G0
    const Problem &P = c->M.P;
    it->bmax.assign(P.nb, 0.0);
    for (int b = 0; b < P.nb; ++b)
      for (int64_t x = c->M.G_off[b]; x < c->M.G_off[b] + (int64_t)P.r * P.dims[b] * P.dims[b]; ++x)
        it->bmax[b] = std::max(it->bmax[b], std::fabs(c->dyson->G0_h[x]));
    it->iters = 0;
    it->done = it->converged = false;
    it->eta = it->eta0;
    it->wide = c->wide;
    if (it->wide) {
      auto w = std::make_unique<WideIter>();
      double *st = status_.ptr + (size_t)it->slot * kStatusStride;
      std::vector<DysonArgs> hd(1, dyson_args(c, st + kStEta, c->sigma.ptr));
      std::vector<IterArgs> hi(1, iter_args(c, st, it->tol, it->mix));
      std::vector<int> pre{0, P.nb};
      w->dyson.upload(hd, stream_);
      w->iter.upload(hi, stream_);
      w->prefix.upload(pre, stream_);
      w->DB = DysonBatch{w->dyson.ptr, w->prefix.ptr, 1, P.nb, dyson_np_max_small(c)};
      w->IB = IterBatch{w->iter.ptr, w->prefix.ptr, 1, P.nb, c->M.G_size};
      it->wi = std::move(w);
      // G = G0 and the cleared status precede the first launches on the problem's own stream
      HIP_CHECK(hipStreamSynchronize(stream_));
    }
    items_.push_back(it);
  }

  void remove(BatchItem *it) {
    auto f = std::find(items_.begin(), items_.end(), it);
    if (f == items_.end()) return;
    items_.erase(f);
    free_slots_.push_back(it->slot);
    it->slot = -1;
    it->wi.reset();
  }

  // One iteration of every active problem. use_graph: replay the captured launch sequence when the
  // composition is unchanged since the last tick.
  void tick(bool use_graph) {
    if (items_.empty()) return;
    HIP_CHECK(hipSetDevice(device_));
    prepare();
    bool any_tile = false;
    for (BatchItem *it : items_) any_tile = any_tile || !it->wide;
    if (!any_tile) {
      // only problems on the large-block engine: nothing for the shared launch sequence
    } else if (use_graph && graph_ok_) {
      HIP_CHECK(hipGraphLaunch(exec_, stream_));
    } else if (use_graph && eager_done_) {
      HIP_CHECK(hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed));
      hipError_t e = launch_all();
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(stream_, &g);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        throw std::runtime_error(std::string("xca: graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
      }
      graph_ = g;
      HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
      graph_ok_ = true;
      HIP_CHECK(hipGraphLaunch(exec_, stream_));
    } else {
      HIP_CHECK(launch_all());
      eager_done_ = true;
    }
    // problems on the large-block engine: their iteration on their own streams, beside the shared
    // sequence; their status words are copied once both have finished
    for (BatchItem *it : items_)
      if (it->wide) launch_wide_iteration(*it);
    HIP_CHECK(hipStreamSynchronize(stream_));
    for (BatchItem *it : items_)
      if (it->wide) {
        HIP_CHECK(hipStreamSynchronize(it->c->stream));
        HIP_CHECK(hipMemcpy(host_status_ + (size_t)it->slot * kStatusStride, status_.ptr + (size_t)it->slot * kStatusStride,
                            sizeof(double) * kStatusStride, hipMemcpyDeviceToHost));
      }
    for (BatchItem *it : items_) {
      const double *st = host_status_ + (size_t)it->slot * kStatusStride;
      it->iters += 1;
      it->eta = st[kStEta];
      it->diff = st[kStDiff];
      it->history.insert(it->history.end(), {st[kStDiff], st[kStZ], st[kStEta]});
      for (int b = 0; b < it->c->M.P.nb; ++b) it->bmax[b] = st[kStBmax + b];
      if (st[kStConv] != 0.0) {
        it->done = true;
        it->converged = true;
      } else if (it->iters >= it->maxiter) {
        it->done = true;
      }
    }
  }

  int64_t ticks_graph() const { return graph_ok_ ? 1 : 0; }

 private:
  // Plans of every problem for its current zero-block pattern; descriptors rebuilt when the
  // (problem, plan) composition changed.
  void prepare() {
    std::vector<uint64_t> key;
    bool evicted = false;
    for (BatchItem *it : items_) {
      xca_ctx *c = it->c;
      const Problem &P = c->M.P;
      c->G_zero.assign(P.nb, 0);
      for (int b = 0; b < P.nb; ++b) c->G_zero[b] = !(it->bmax[b] >= 1e-16);
      c->have_G = true;
      if (it->zero_seen != c->G_zero) {
        // a block crossed the zero threshold: the plans of the old pattern are released, so that a
        // problem never holds more than one pattern's plans (its admission estimate). Every launch
        // that used them has finished (each tick ends with a synchronisation). Contexts whose plans
        // a Sigma launch sequence also references keep them.
        if (!it->zero_seen.empty() && !c->sigall) {
          c->plans.clear();
          c->wide_plans.clear();
          evicted = true;
        }
        it->zero_seen = c->G_zero;
      }
      it->plans.clear();
      it->scales.clear();
      key.push_back(it->serial);
      if (it->wide) {  // Sigma on the large-block engine, planned by its own launches
        key.push_back(0);
        continue;
      }
      key.pop_back();
      for (int order = 1; order <= it->max_order; ++order)
        for (auto &st : connected_topologies(order)) {
          DevPlan *dp = get_plan(c, order, st.second.data(), nullptr, 0);
          if (dp->items == 0) continue;
          if (dp->chosen < 0) {
            // autotune on first use: tables for this G, then time the tile widths
            if (dp->ntabjobs > 0) {
              std::vector<TabSeg> ts(1, tab_seg(c, dp));
              std::vector<int> tp{0, dp->ntabjobs};
              DevBuf<TabSeg> dts;
              DevBuf<int> dtp;
              dts.upload(ts, c->stream);
              dtp.upload(tp, c->stream);
              HIP_CHECK(launch_tab(c->RB, TabBatch{dts.ptr, dtp.ptr, 1, dp->ntabjobs}, c->stream));
              autotune(c, dp);
            } else {
              autotune(c, dp);
            }
          }
          it->plans.push_back(dp);
          it->scales.push_back(-(double)st.first);
        }
      key.push_back(it->serial);
      for (DevPlan *dp : it->plans) key.push_back((uint64_t)(uintptr_t)dp);
      key.push_back(0);
    }
    if (key == key_ && !evicted) return;  // after an eviction, new plans may reuse old addresses
    key_ = key;
    drop_graph();
    eager_done_ = false;
    build_descriptors();
  }

  void build_descriptors() {
    hipStream_t s = stream_;
    std::vector<BatchItem *> tiles;  // the problems of the shared launch sequence
    for (BatchItem *it : items_)
      if (!it->wide) tiles.push_back(it);
    // edge tables and chain segments, grouped by padded rank and tile width
    std::map<int, std::vector<TabSeg>> tabs;                       // RB -> segments
    std::map<std::pair<int, int>, std::vector<ChainArgs>> chains;  // (RB, variant) -> segments
    h_redplans_.clear();
    h_redprobs_.clear();
    h_dyson_.clear();
    h_iter_.clear();
    std::vector<int> red_prefix{0}, dys_prefix{0};
    max_sig_ = 0;
    int64_t max_total = 0;
    for (BatchItem *it : tiles) {
      xca_ctx *c = it->c;
      const Problem &P = c->M.P;
      for (size_t q = 0; q < it->plans.size(); ++q) {
        DevPlan *dp = it->plans[q];
        if (dp->ntabjobs > 0) tabs[c->RB].push_back(tab_seg(c, dp));
        DevVariant *V = dp->var[dp->chosen].get();
        chains[{c->RB, dp->chosen}].push_back(chain_args(c, dp, V));
      }
    }
    // red plans after the loop above so the vector does not reallocate under stored pointers
    for (BatchItem *it : tiles)
      for (size_t q = 0; q < it->plans.size(); ++q) {
        DevVariant *V = it->plans[q]->var[it->plans[q]->chosen].get();
        h_redplans_.push_back(RedPlan{V->slot.ptr, V->red_chunks.ptr, V->red_range.ptr, V->partial.ptr, it->scales[q]});
      }
    redplans_.upload(h_redplans_, s);
    int rp = 0;
    for (BatchItem *it : tiles) {
      xca_ctx *c = it->c;
      const Problem &P = c->M.P;
      RedProb rq{};
      rq.plans = redplans_.ptr + rp;
      rq.nplans = (int)it->plans.size();
      rp += rq.nplans;
      rq.sig_off = c->sig_off.ptr;
      rq.sig_size = c->sig_size.ptr;
      rq.sigma = c->sigma.ptr;
      rq.nb = P.nb;
      h_redprobs_.push_back(rq);
      red_prefix.push_back(red_prefix.back() + P.nb);
      for (int b = 0; b < P.nb; ++b) max_sig_ = std::max(max_sig_, P.r * P.dims[b] * P.dims[b]);
      double *st = status_.ptr + (size_t)it->slot * kStatusStride;
      h_dyson_.push_back(dyson_args(c, st + kStEta, c->sigma.ptr));
      dys_prefix.push_back(dys_prefix.back() + P.nb);
      h_iter_.push_back(iter_args(c, st, it->tol, it->mix));
      max_total = std::max<int64_t>(max_total, c->M.G_size);
    }
    redprobs_.upload(h_redprobs_, s);
    h_red_prefix_ = red_prefix;
    red_prefix_.upload(h_red_prefix_, s);
    red_ = RedBatch{redprobs_.ptr, red_prefix_.ptr, (int)tiles.size(), red_prefix.back()};
    dyson_.upload(h_dyson_, s);
    h_dys_prefix_ = dys_prefix;
    dys_prefix_.upload(h_dys_prefix_, s);
    int np_max = 16;
    for (BatchItem *it : tiles) np_max = std::max(np_max, dyson_np_max(it->c));
    dys_ = DysonBatch{dyson_.ptr, dys_prefix_.ptr, (int)tiles.size(), dys_prefix.back(), np_max};
    iter_.upload(h_iter_, s);
    iter_b_ = IterBatch{iter_.ptr, dys_prefix_.ptr, (int)tiles.size(), dys_prefix.back(), max_total};
    // tab groups
    tab_groups_.clear();
    for (auto &kv : tabs) {
      auto g = std::make_unique<TabGroup>();
      g->RB = kv.first;
      g->h_segs = kv.second;
      g->h_prefix.assign(1, 0);
      for (auto &t : kv.second) g->h_prefix.push_back(g->h_prefix.back() + jobs_of(t));
      g->segs.upload(g->h_segs, s);
      g->prefix.upload(g->h_prefix, s);
      g->B = TabBatch{g->segs.ptr, g->prefix.ptr, (int)g->h_segs.size(), g->h_prefix.back()};
      tab_groups_.push_back(std::move(g));
    }
    chain_groups_.clear();
    size_t qi = 0;
    for (auto &kv : chains) {
      auto g = std::make_unique<ChainGroup>();
      g->RB = kv.first.first;
      g->variant = kv.first.second;
      g->h_segs = kv.second;
      g->h_prefix.assign(1, 0);
      for (auto &a : kv.second) g->h_prefix.push_back(g->h_prefix.back() + a.nchunks);
      g->segs.upload(g->h_segs, s);
      g->prefix.upload(g->h_prefix, s);
      if (qi >= queues_.size()) {
        queues_.push_back(std::make_unique<DevBuf<int>>());
        queues_.back()->alloc(2);
        HIP_CHECK(init_queue(queues_.back()->ptr, s));
      }
      const int64_t stride = 16LL * g->RB * tile_columns(g->variant);
      g->scratch.alloc((size_t)chain_resident(g->RB, g->variant) * stride);
      g->B = ChainBatch{g->segs.ptr, g->prefix.ptr, (int)g->h_segs.size(), g->h_prefix.back(), queues_[qi]->ptr, g->scratch.ptr, stride};
      g->B.maxb = max_block_of(g->h_segs);
      ++qi;
      chain_groups_.push_back(std::move(g));
    }
    HIP_CHECK(hipStreamSynchronize(s));
  }

  // jobs of a tab segment: recovered from the plan that produced it
  int jobs_of(const TabSeg &t) const {
    for (BatchItem *it : items_)
      for (DevPlan *dp : it->plans)
        if (dp->tabjobs.ptr == t.jobs) return dp->ntabjobs;
    return 0;
  }

  // One iteration of a problem on the large-block engine, on its context's stream: Sigma topology by
  // topology (net factor -s per topology), the Dyson solve (blocks up to 12 rows in the one-workgroup
  // kernel, larger ones device-wide, both hermitised; eta of the device-wide path from the last
  // status read), then the same update launches as the shared sequence.
  void launch_wide_iteration(BatchItem &it) {
    xca_ctx *c = it.c;
    hipStream_t s = c->stream;
    HIP_CHECK(launch_fill(c->M.G_size, 0.0, c->sigma.ptr, s));
    c->outer_events = true;
    try {
      for (int order = 1; order <= it.max_order; ++order)
        for (auto &tp : connected_topologies(order)) accumulate(c, order, tp.second.data(), nullptr, 0, -(double)tp.first);
    } catch (...) {
      c->outer_events = false;
      throw;
    }
    c->outer_events = false;
    HIP_CHECK(launch_dyson(it.wi->DB, s));
    if (c->dyson->large) large_dyson_solve(c, c->dyson->large, c->sigma.ptr, it.eta, c->dyson->gnew.ptr);
    HIP_CHECK(launch_iterate(it.wi->IB, s));
    HIP_CHECK(launch_block_max(it.wi->IB, s));
  }

  hipError_t launch_all() {
    hipError_t e;
    for (auto &g : tab_groups_)
      if ((e = launch_tab(g->RB, g->B, stream_)) != hipSuccess) return e;
    for (auto &g : chain_groups_)
      if ((e = launch_chain(g->RB, g->variant, g->B, stream_)) != hipSuccess) return e;
    if ((e = launch_reduce_assign(red_, max_sig_, stream_)) != hipSuccess) return e;
    if ((e = launch_dyson(dys_, stream_)) != hipSuccess) return e;
    if ((e = launch_iterate(iter_b_, stream_)) != hipSuccess) return e;
    if ((e = launch_block_max(iter_b_, stream_)) != hipSuccess) return e;
    return hipMemcpyAsync(host_status_, status_.ptr, sizeof(double) * (size_t)cap_ * kStatusStride, hipMemcpyDeviceToHost, stream_);
  }

  void drop_graph() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    graph_ok_ = false;
  }

  struct TabGroup {
    int RB = 0;
    std::vector<TabSeg> h_segs;
    std::vector<int> h_prefix;
    DevBuf<TabSeg> segs;
    DevBuf<int> prefix;
    TabBatch B{};
  };
  struct ChainGroup {
    int RB = 0, variant = 0;
    std::vector<ChainArgs> h_segs;
    std::vector<int> h_prefix;
    DevBuf<ChainArgs> segs;
    DevBuf<int> prefix;
    DevBuf<double> scratch;
    ChainBatch B{};
  };

  int device_;
  hipStream_t stream_;
  int cap_;
  std::vector<BatchItem *> items_;
  std::vector<int> free_slots_;
  DevBuf<double> status_;
  double *host_status_ = nullptr;
  std::vector<uint64_t> key_;
  uint64_t serial_ = 0;
  std::vector<std::unique_ptr<TabGroup>> tab_groups_;
  std::vector<std::unique_ptr<ChainGroup>> chain_groups_;
  std::vector<std::unique_ptr<DevBuf<int>>> queues_;
  std::vector<RedPlan> h_redplans_;
  std::vector<RedProb> h_redprobs_;
  std::vector<int> h_red_prefix_, h_dys_prefix_;
  std::vector<DysonArgs> h_dyson_;
  std::vector<IterArgs> h_iter_;
  DevBuf<RedPlan> redplans_;
  DevBuf<RedProb> redprobs_;
  DevBuf<int> red_prefix_, dys_prefix_;
  DevBuf<DysonArgs> dyson_;
  DevBuf<IterArgs> iter_;
  RedBatch red_{};
  DysonBatch dys_{};
  IterBatch iter_b_{};
  int max_sig_ = 0;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  bool graph_ok_ = false, eager_done_ = false;
};

}  // namespace xca

using namespace xca;

// Pseudo-particle self-consistency of one problem through the batched engine (a batch of one).
// On return G_out holds the solver's G (the last mixed iterate, block_sparse_solver.py:338-346),
// eta_out the pseudo-particle chemical potential.
extern "C" int xca_ppsc_solve(xca_ctx *c, double eta0, int max_order, double tol, int maxiter, double mix, double *G_out, double *eta_out,
                              int *iters_out, double *diff_out) {
  try {
    require_dyson(c);
    if (max_order < 1 || max_order > 4) throw std::invalid_argument("xca_ppsc_solve: max_order must be between 1 and 4");
    if (maxiter < 1) throw std::invalid_argument("xca_ppsc_solve: maxiter must be positive");
    HIP_CHECK(hipSetDevice(c->device));
    BatchItem it;
    it.c = c;
    it.max_order = max_order;
    it.maxiter = maxiter;
    it.tol = tol;
    it.mix = mix;
    it.eta0 = eta0;
    Batch batch(c->device, c->stream, 1);
    batch.add(&it);
    const bool graphs = !std::getenv("XCA_NO_GRAPH");
    while (!it.done) batch.tick(graphs);
    batch.remove(&it);
    HIP_CHECK(hipMemcpyAsync(G_out, c->G.ptr, c->M.G_size * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    if (eta_out) *eta_out = it.eta;
    if (iters_out) *iters_out = it.iters;
    if (diff_out) *diff_out = it.diff;
    c->ppsc_history = it.history;
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

// Per-iteration record of the last xca_ppsc_solve on this context: rows of (max |G_new - G|, Z before
// the normalisation, eta); returns the number of iterations.
extern "C" int xca_ppsc_history(xca_ctx *c, double *out, int max_iters) {
  if (!c) return -1;
  const int n = (int)(c->ppsc_history.size() / 3);
  if (out && max_iters > 0) std::copy(c->ppsc_history.begin(), c->ppsc_history.begin() + 3 * std::min(n, max_iters), out);
  return n;
}

// ======================================================================= full-order Sigma

namespace xca {

// Build (or reuse) the combined launch descriptors of Sigma over orders 1..max_order for the
// context's current zero-block pattern.
static SigmaAll *sigma_all(xca_ctx *c, int max_order) {
  std::vector<int> key(c->G_zero.begin(), c->G_zero.end());
  key.push_back(max_order);
  key.push_back(c->tile_variant);
  if (c->sigall && c->sigall->key == key) return c->sigall.get();
  c->sigall.reset();
  auto S = std::make_unique<SigmaAll>();
  S->key = key;
  hipStream_t s = c->stream;
  for (int order = 1; order <= max_order; ++order)
    for (auto &st : connected_topologies(order)) {
      DevPlan *dp = get_plan(c, order, st.second.data(), nullptr, 0);
      if (dp->items == 0) continue;
      if (dp->chosen < 0) {  // autotune on first use, with this G's tables
        if (dp->ntabjobs > 0) {
          std::vector<TabSeg> ts(1, tab_seg(c, dp));
          std::vector<int> tp{0, dp->ntabjobs};
          DevBuf<TabSeg> dts;
          DevBuf<int> dtp;
          dts.upload(ts, s);
          dtp.upload(tp, s);
          HIP_CHECK(launch_tab(c->RB, TabBatch{dts.ptr, dtp.ptr, 1, dp->ntabjobs}, s));
        }
        autotune(c, dp);
      }
      S->plans.push_back(dp);
      S->scales.push_back(-(double)st.first);
      S->items += dp->items;
      S->macs += dp->macs;
    }
  S->h_tab_prefix.assign(1, 0);
  // Every plan takes the tile width of the costliest one (by the autotuner's estimate) where it has
  // that variant, so the small plans' chunks fill the idle workgroups of the costliest plan's last
  // round instead of running as a launch of their own; the costliest plans are queued first.
  std::vector<size_t> order(S->plans.size());
  for (size_t q = 0; q < order.size(); ++q) order[q] = q;
  auto est = [&](size_t q) { return S->plans[q]->var[S->plans[q]->chosen]->ms; };
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return est(a) > est(b); });
  const int vdom = order.empty() ? -1 : S->plans[order[0]]->chosen;
  std::vector<DevVariant *> pv(S->plans.size());
  for (size_t q = 0; q < S->plans.size(); ++q) {
    DevPlan *dp = S->plans[q];
    pv[q] = (vdom >= 0 && dp->var[vdom]) ? dp->var[vdom].get() : dp->var[dp->chosen].get();
    if (dp->ntabjobs > 0) {
      S->h_tab.push_back(tab_seg(c, dp));
      S->h_tab_prefix.push_back(S->h_tab_prefix.back() + dp->ntabjobs);
    }
    DevVariant *V = pv[q];
    S->h_red.push_back(RedPlan{V->slot.ptr, V->red_chunks.ptr, V->red_range.ptr, V->partial.ptr, S->scales[q]});
  }
  std::map<int, size_t> group_of;  // tile variant -> group
  for (size_t q : order) {
    DevPlan *dp = S->plans[q];
    DevVariant *V = pv[q];
    const int v = (int)(std::find_if(dp->var.begin(), dp->var.end(), [&](const auto &x) { return x.get() == V; }) - dp->var.begin());
    auto g = group_of.find(v);
    if (g == group_of.end()) {
      group_of[v] = S->groups.size();
      S->groups.push_back(std::make_unique<SigmaAll::Group>());
      S->groups.back()->RB = c->RB;
      S->groups.back()->variant = v;
      S->groups.back()->h_prefix.assign(1, 0);
      g = group_of.find(v);
    }
    SigmaAll::Group &G = *S->groups[g->second];
    G.h_segs.push_back(chain_args(c, dp, V));
    G.h_prefix.push_back(G.h_prefix.back() + V->nchunks);
  }
  S->tab.upload(S->h_tab, s);
  S->tab_prefix.upload(S->h_tab_prefix, s);
  for (auto &gp : S->groups) {
    SigmaAll::Group &G = *gp;
    G.segs.upload(G.h_segs, s);
    G.prefix.upload(G.h_prefix, s);
    G.queue.alloc(2);
    HIP_CHECK(init_queue(G.queue.ptr, s));
    const int64_t stride = 16LL * G.RB * tile_columns(G.variant);
    G.scratch.alloc((size_t)chain_resident(G.RB, G.variant) * stride);
    G.B = ChainBatch{G.segs.ptr, G.prefix.ptr, (int)G.h_segs.size(), G.h_prefix.back(), G.queue.ptr, G.scratch.ptr, stride};
    G.B.maxb = max_block_of(G.h_segs);
  }
  S->red.upload(S->h_red, s);
  const Problem &P = c->M.P;
  RedProb rq{};
  rq.plans = S->red.ptr;
  rq.nplans = (int)S->plans.size();
  rq.sig_off = c->sig_off.ptr;
  rq.sig_size = c->sig_size.ptr;
  rq.sigma = c->sigma.ptr;
  rq.nb = P.nb;
  S->prob.upload(std::vector<RedProb>(1, rq), s);
  S->prob_prefix.upload(std::vector<int>{0, P.nb}, s);
  S->R = RedBatch{S->prob.ptr, S->prob_prefix.ptr, 1, P.nb};
  for (int b = 0; b < P.nb; ++b) S->max_sig = std::max(S->max_sig, P.r * P.dims[b] * P.dims[b]);
  HIP_CHECK(hipStreamSynchronize(s));
  c->sigall = std::move(S);
  return c->sigall.get();
}

// the launch sequence; events (when given) bracket the chain launches
static hipError_t sigma_all_launch(xca_ctx *c, SigmaAll &S, hipEvent_t e0, hipEvent_t e1) {
  hipError_t e;
  hipStream_t s = c->stream;
  if (S.plans.empty()) return launch_fill(c->M.G_size, 0.0, c->sigma.ptr, s);
  if (!S.h_tab.empty() &&
      (e = launch_tab(c->RB, TabBatch{S.tab.ptr, S.tab_prefix.ptr, (int)S.h_tab.size(), S.h_tab_prefix.back()}, s)) != hipSuccess)
    return e;
  if (e0 && (e = hipEventRecord(e0, s)) != hipSuccess) return e;
  for (auto &g : S.groups)
    if ((e = launch_chain(g->RB, g->variant, g->B, s)) != hipSuccess) return e;
  if (e1 && (e = hipEventRecord(e1, s)) != hipSuccess) return e;
  return launch_reduce_assign(S.R, S.max_sig, s);
}

}  // namespace xca

extern "C" int xca_sigma_order(xca_ctx *c, int max_order, int timed) {
  try {
    if (!c) throw std::invalid_argument("xca: null context");
    if (max_order < 1 || max_order > 4) throw std::invalid_argument("xca_sigma_order: max_order must be between 1 and 4");
    if (!c->have_G) throw std::runtime_error("xca: G has not been set");
    HIP_CHECK(hipSetDevice(c->device));
    if (c->wide) {  // large blocks: topology by topology on the large-block engine, timed as a whole
      HIP_CHECK(launch_fill(c->M.G_size, 0.0, c->sigma.ptr, c->stream));
      int64_t items = 0;
      double macs = 0;
      c->outer_events = true;
      HIP_CHECK(hipEventRecord(c->ev0, c->stream));
      try {
        for (int order = 1; order <= max_order; ++order)
          for (auto &st : connected_topologies(order)) {
            accumulate(c, order, st.second.data(), nullptr, 0, -(double)st.first);
            items += c->last_items;
            macs += c->last_macs;
          }
      } catch (...) {
        c->outer_events = false;
        throw;
      }
      c->outer_events = false;
      HIP_CHECK(hipEventRecord(c->ev1, c->stream));
      c->last_items = items;
      c->last_macs = macs;
      c->last_ms = 0;
      return 0;
    }
    SigmaAll *S = sigma_all(c, max_order);
    c->last_items = S->items;
    c->last_macs = S->macs;
    c->last_ms = 0;
    if (timed || std::getenv("XCA_NO_GRAPH")) {
      HIP_CHECK(sigma_all_launch(c, *S, timed ? c->ev0 : nullptr, timed ? c->ev1 : nullptr));
      return 0;
    }
    if (!S->exec) {
      HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
      hipError_t e = sigma_all_launch(c, *S, nullptr, nullptr);
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(c->stream, &g);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        throw std::runtime_error(std::string("xca: graph capture failed: ") + hipGetErrorString(e != hipSuccess ? e : e2));
      }
      S->graph = g;
      HIP_CHECK(hipGraphInstantiate(&S->exec, S->graph, nullptr, nullptr, 0));
    }
    HIP_CHECK(hipGraphLaunch(S->exec, c->stream));
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

// ================================================================================== server

struct XcaRequest {
  int64_t id = 0;
  std::unique_ptr<xca_ctx, void (*)(xca_ctx *)> ctx{nullptr, xca_ctx_destroy};
  std::vector<double> G0;
  BatchItem item;
  int64_t bytes = 0;
  bool admitted = false, finished = false;
  double t_submit = 0, t_admit = 0, t_finish = 0;
  std::vector<double> G_result;
};

struct xca_server {
  int device = 0;
  hipStream_t stream = nullptr;
  int64_t budget = 0, in_use = 0, reserved = 0;  // budget: for requests, after the shared launch buffers (reserved)
  int max_active = 0;
  int tile_variant = 0;
  bool graphs = true;
  int64_t next_id = 0;
  std::unique_ptr<Batch> batch;
  std::deque<XcaRequest *> queue;                    // submitted, not admitted (arrival order)
  std::vector<XcaRequest *> active;
  std::map<int64_t, std::unique_ptr<XcaRequest>> all;
  std::vector<int64_t> finished;                     // not yet polled
  int64_t ticks = 0;
};

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

extern "C" xca_server *xca_server_create(int device, int64_t mem_budget_bytes, int max_active, int tile_variant) {
  xca_server *sv = nullptr;
  try {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw std::runtime_error("xca_server_create: no GPU device available");
    if (device < 0 || device >= ndev) throw std::runtime_error("xca_server_create: device index out of range");
    if (max_active < 1) throw std::invalid_argument("xca_server_create: max_active must be positive");
    if (tile_variant < 0 || tile_variant >= num_tile_variants()) throw std::invalid_argument("xca_server_create: tile variant out of range");
    sv = new xca_server();
    sv->device = device;
    sv->max_active = max_active;
    sv->tile_variant = tile_variant;
    sv->graphs = !std::getenv("XCA_NO_GRAPH");
    HIP_CHECK(hipSetDevice(device));
    size_t fr = 0, tot = 0;
    HIP_CHECK(hipMemGetInfo(&fr, &tot));
    sv->budget = mem_budget_bytes > 0 ? mem_budget_bytes : (int64_t)(0.8 * (double)fr);
    // the batched chain launches share per-workgroup shared-prefix caches, one per (padded rank,
    // tile variant) group: the server's variant and the widest one, which plans fall back to
    for (int rb : {2, 4})
      for (int v : {tile_variant, 0}) {
        sv->reserved += (int64_t)chain_resident(rb, v) * 16 * rb * tile_columns(v) * (int64_t)sizeof(double);
        if (v == tile_variant && tile_variant == 0) break;
      }
    sv->budget -= sv->reserved;
    if (sv->budget <= 0) throw std::invalid_argument("xca_server_create: memory budget too small for the shared launch buffers");
    HIP_CHECK(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking));
    sv->batch = std::make_unique<Batch>(device, sv->stream, max_active);
    return sv;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    delete sv;
    return nullptr;
  }
}

extern "C" void xca_server_destroy(xca_server *sv) {
  if (!sv) return;
  (void)hipSetDevice(sv->device);
  if (sv->stream) (void)hipStreamSynchronize(sv->stream);
  sv->batch.reset();
  sv->all.clear();
  if (sv->stream) (void)hipStreamDestroy(sv->stream);
  delete sv;
}

// Queue a problem; its device footprint is estimated on the host now and it is admitted when both a
// batch slot and that much of the memory budget are free. Returns the request id (>= 0).
extern "C" int64_t xca_server_submit(xca_server *sv, const xca_problem *desc, const double *G0, double eta0, int max_order, double tol,
                                     int maxiter, double mix) {
  try {
    if (!sv) throw std::invalid_argument("xca_server_submit: null server");
    if (max_order < 1 || max_order > 4) throw std::invalid_argument("xca_server_submit: max_order must be between 1 and 4");
    auto rq = std::make_unique<XcaRequest>();
    rq->id = sv->next_id++;
    rq->t_submit = now_s();
    rq->ctx.reset(ctx_init_host(desc));
    xca_ctx *c = rq->ctx.get();
    c->tile_variant = sv->tile_variant;
    rq->G0.assign(G0, G0 + c->M.G_size);
    rq->item.max_order = max_order;
    rq->item.maxiter = maxiter;
    rq->item.tol = tol;
    rq->item.mix = mix;
    rq->item.eta0 = eta0;
    // footprint: resident tables, Dyson state and the plans at G0's (all non-zero) block pattern
    c->G_zero.assign(c->M.P.nb, 0);
    int64_t bytes = ctx_bytes(c) + dyson_bytes(c);
    for (int order = 1; order <= max_order; ++order)
      for (auto &st : connected_topologies(order)) bytes += plan_bytes_estimate(c, order, st.second.data());
    rq->bytes = bytes;
    if (bytes > sv->budget) throw std::runtime_error("xca_server_submit: problem needs more device memory than the server budget");
    XcaRequest *raw = rq.get();
    sv->all[raw->id] = std::move(rq);
    sv->queue.push_back(raw);
    return raw->id;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

// Admit queued requests in arrival order while a slot and the memory budget allow, run one
// iteration of every active problem, and retire the finished ones.
extern "C" int xca_server_step(xca_server *sv, int *n_active, int *n_queued, int *n_finished) {
  try {
    if (!sv) throw std::invalid_argument("xca_server_step: null server");
    HIP_CHECK(hipSetDevice(sv->device));
    while (!sv->queue.empty() && !sv->batch->full()) {
      XcaRequest *rq = sv->queue.front();
      if (sv->in_use + rq->bytes > sv->budget && !sv->active.empty()) break;
      sv->queue.pop_front();
      xca_ctx *c = rq->ctx.get();
      ctx_init_device(c, sv->device);
      if (xca_dyson_init(c, rq->G0.data()) != 0) throw std::runtime_error(xca_last_error());
      rq->item.c = c;
      sv->batch->add(&rq->item);
      rq->admitted = true;
      rq->t_admit = now_s();
      sv->in_use += rq->bytes;
      sv->active.push_back(rq);
    }
    int nfin = 0;
    if (!sv->active.empty()) {
      sv->batch->tick(sv->graphs);
      ++sv->ticks;
      for (size_t i = 0; i < sv->active.size();) {
        XcaRequest *rq = sv->active[i];
        if (!rq->item.done) {
          ++i;
          continue;
        }
        xca_ctx *c = rq->ctx.get();
        rq->G_result.resize(c->M.G_size);
        HIP_CHECK(hipMemcpy(rq->G_result.data(), c->G.ptr, c->M.G_size * sizeof(double), hipMemcpyDeviceToHost));
        sv->batch->remove(&rq->item);
        rq->finished = true;
        rq->t_finish = now_s();
        sv->in_use -= rq->bytes;
        rq->ctx.reset();  // frees the device state
        rq->item.c = nullptr;
        sv->finished.push_back(rq->id);
        sv->active.erase(sv->active.begin() + i);
        ++nfin;
      }
    }
    if (n_active) *n_active = (int)sv->active.size();
    if (n_queued) *n_queued = (int)sv->queue.size();
    if (n_finished) *n_finished = nfin;
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

// Ids of requests finished since the last poll (up to max_ids); returns how many were written.
extern "C" int xca_server_poll(xca_server *sv, int64_t *ids, int max_ids) {
  if (!sv) return -1;
  int n = std::min<int>(max_ids, (int)sv->finished.size());
  for (int i = 0; i < n; ++i) ids[i] = sv->finished[i];
  sv->finished.erase(sv->finished.begin(), sv->finished.begin() + n);
  return n;
}

// Result of a finished request; the request is released afterwards. times_out (optional, 3 doubles):
// seconds from submission to admission, admission to finish, and submission to finish.
extern "C" int xca_server_result(xca_server *sv, int64_t id, double *G_out, double *eta_out, int *iters_out, double *diff_out, int *converged,
                                 double *times_out) {
  try {
    if (!sv) throw std::invalid_argument("xca_server_result: null server");
    auto f = sv->all.find(id);
    if (f == sv->all.end()) throw std::invalid_argument("xca_server_result: unknown request id");
    XcaRequest *rq = f->second.get();
    if (!rq->finished) throw std::runtime_error("xca_server_result: request has not finished");
    if (G_out) std::memcpy(G_out, rq->G_result.data(), rq->G_result.size() * sizeof(double));
    if (eta_out) *eta_out = rq->item.eta;
    if (iters_out) *iters_out = rq->item.iters;
    if (diff_out) *diff_out = rq->item.diff;
    if (converged) *converged = rq->item.converged ? 1 : 0;
    if (times_out) {
      times_out[0] = rq->t_admit - rq->t_submit;
      times_out[1] = rq->t_finish - rq->t_admit;
      times_out[2] = rq->t_finish - rq->t_submit;
    }
    sv->all.erase(f);
    return 0;
  } catch (std::exception &e) {
    xca::set_error(e.what());
    return -1;
  }
}

// Estimated device bytes of a submitted request (the admission footprint).
extern "C" int64_t xca_server_request_bytes(xca_server *sv, int64_t id) {
  if (!sv) return -1;
  auto f = sv->all.find(id);
  return f == sv->all.end() ? -1 : f->second->bytes;
}

// Bytes of the budget reserved at creation for the launch buffers shared by all requests.
extern "C" int64_t xca_server_shared_bytes(xca_server *sv) { return sv ? sv->reserved : -1; }

// Device bytes an admitted request holds now (tables, Dyson state, plans); -1 if it is not active.
extern "C" int64_t xca_server_request_device_bytes(xca_server *sv, int64_t id) {
  if (!sv) return -1;
  auto f = sv->all.find(id);
  if (f == sv->all.end() || !f->second->admitted || f->second->finished) return -1;
  return ctx_device_bytes(f->second->ctx.get());
}

// Counters: iterations run (ticks) and device bytes in use by admitted requests.
extern "C" int xca_server_stats(xca_server *sv, int64_t *ticks, int64_t *bytes_in_use, int64_t *budget) {
  if (!sv) return -1;
  if (ticks) *ticks = sv->ticks;
  if (bytes_in_use) *bytes_in_use = sv->in_use;
  if (budget) *budget = sv->budget;
  return 0;
}
