// Device plans of the self-energy engine and the launch entry points of the HIP kernels, shared by
// the single-problem ABI (context.cpp) and the batched iteration engine (batch.cpp).
#pragma once
#include <hip/hip_runtime_api.h>

#include <memory>
#include <vector>

#include "ctx.hpp"
#include "engine.hpp"

namespace xca::dev {
int num_tile_variants();
int tile_columns(int variant);
hipError_t launch_tab(int RB, const TabBatch &B, hipStream_t s);
hipError_t launch_chain(int RB, int variant, const ChainBatch &B, hipStream_t s);
// workgroups of one chain launch resident on the device at once (its grid never exceeds this)
int chain_resident(int RB, int variant);
hipError_t launch_fill(int64_t n, double v, double *p, hipStream_t s);
hipError_t launch_reduce(const int64_t *slot, const int *red_chunks, const I2 *red_range, const int64_t *sig_off, const int *sig_size,
                         int nblocks, int max_size, const double *partial, double *sigma, double scale, hipStream_t s);
hipError_t launch_reduce_assign(const RedBatch &B, int max_size, hipStream_t s);
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
  // one-segment edge-table launch descriptor
  DevBuf<TabSeg> tseg;
  DevBuf<int> tseg_prefix;
  std::vector<TabSeg> h_tseg;
  std::vector<int> h_tseg_prefix;
};

DevPlan *get_plan(xca_ctx *c, int m, const int *topo, const int32_t *fix, int64_t nfix);
ChainArgs chain_args(xca_ctx *c, DevPlan *dp, DevVariant *V);
TabSeg tab_seg(xca_ctx *c, DevPlan *dp);
void autotune(xca_ctx *c, DevPlan *dp);
std::vector<std::pair<int, std::vector<int>>> connected_topologies(int order);

}  // namespace xca
