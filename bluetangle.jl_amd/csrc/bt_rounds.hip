// Lockstep ("rounds") executor: many states advance together.
//
// Each trajectory or request owns a device state and an interpreter that
// runs on a host worker thread.  A round takes the next pass of every active
// job, uploads the jobs' tile-local operator tables in one copy, and runs
// them all in ONE kernel launch (k_round: one workgroup per group of tiles of
// one job, states of any size and qubit count side by side), then one
// finishing launch that reduces each job's FP64 partials in a fixed order.
// Jobs leave as they finish and new ones are admitted into the freed HBM
// (continuous batching); bt_run_batch and the server are both built on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bt_circuit.hpp"
#include "bt_internal.hpp"
#include "bt_interp.hpp"
#include "bt_tile.hpp"

namespace bt {

int tile_bits_for(int prec, int nbits);
int coalesce_bits(int prec);
int tiles_per_block();
int item_tpb(int setting, uint64_t ntiles, int cap_div);

// ---------------------------------------------------------------------------
// device side
// ---------------------------------------------------------------------------
struct RoundItem {
  void* state;
  uint64_t table;     // byte offset of the job's table in the round's blob
  uint32_t ntiles;    // tiles of the pass
  uint32_t block0;    // first workgroup of this item
  uint32_t nblocks;
  uint32_t tpb;       // tiles per workgroup of this item
  int32_t fresh;      // state is |0..0> and not yet written
  int32_t pad;
};

// table layout: PassGeom, then int32 nmats (padded to 16 B), then nmats matrix entries
constexpr size_t kGeomBytes = (sizeof(PassGeom) + 15) / 16 * 16;


// One workgroup per `tpb` consecutive tiles of one item (tpb per item, from
// the item's own size: an item's reductions, summed per workgroup and then
// over its workgroups in a fixed order, do not depend on what else the
// launch holds).  nb_uniform != 0: every item has that many workgroups (the
// block -> item map is a division), else a binary search over the items'
// first blocks.
template <typename R>
__global__ void __launch_bounds__(TileCfg<R>::NT, TileCfg<R>::WAVES) k_round(const RoundItem* items, int nitems, uint32_t nb_uniform,
                                                          const unsigned char* blob, double* part) {
  constexpr int NT = TileCfg<R>::NT;
  extern __shared__ __align__(16) unsigned char smem[];
  C<R>* tile = (C<R>*)smem;
  __shared__ PassGeom geom;
  __shared__ C<R> mats[kMatEntries];
  __shared__ double red[(NT / 64) * kAcc];
  __shared__ double sums[kAcc];
  __shared__ int item_sh;
  int item;
  if (nb_uniform) {
    item = (int)(blockIdx.x / nb_uniform);
  } else {
    if (threadIdx.x == 0) {
      int lo = 0, hi = nitems - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].block0 <= blockIdx.x) lo = mid;
        else hi = mid - 1;
      }
      item_sh = lo;
    }
    __syncthreads();
    item = item_sh;
  }
  const RoundItem it = items[item];
  const unsigned char* tab = blob + it.table;
  load_geom((const PassGeom*)tab, &geom, threadIdx.x, NT);
  {
    const int nm = *(const int*)(tab + kGeomBytes);
    const C<R>* m = (const C<R>*)(tab + kGeomBytes + 16);
    for (int i = threadIdx.x; i < nm; i += NT) mats[i] = m[i];
  }
  if (threadIdx.x < kAcc) sums[threadIdx.x] = 0.0;
  __syncthreads();
  const uint32_t local = blockIdx.x - it.block0;
  const uint64_t t0 = (uint64_t)local * it.tpb;
  const uint64_t tend = std::min<uint64_t>((uint64_t)it.ntiles, t0 + it.tpb);
  tile_pass<R, uint32_t>((C<R>*)it.state, geom, mats, tile, t0, 1, tend, red, sums, it.fresh);
  if (threadIdx.x < kAcc) part[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = sums[threadIdx.x];
}

// out[item, j]: fixed-order sum of the item's partials (grid: items x nacc,
// the accumulators the round's reductions use)
__global__ void __launch_bounds__(256) k_round_finalize(const RoundItem* items, const double* part, uint64_t stride,
                                                        double* out) {
  const RoundItem it = items[blockIdx.x];
  const double t = sum_blocks(part + (uint64_t)blockIdx.y * stride + it.block0, it.nblocks);
  if (threadIdx.x == 0) out[(uint64_t)blockIdx.x * gridDim.y + blockIdx.y] = t;
}

template <typename R>
__global__ void k_export_state(const C<R>* s, uint64_t n, double* out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    C<R> v = s[i];
    out[2 * i] = (double)v.x;
    out[2 * i + 1] = (double)v.y;
  }
}

// ---------------------------------------------------------------------------
// host worker pool
// ---------------------------------------------------------------------------
class Pool {
 public:
  explicit Pool(int n) {
    for (int i = 0; i < n; i++) th_.emplace_back([this] { loop(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      gen_++;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  // run f(i) for i in [0, n) on the pool and the caller; returns when done
  void parallel_for(int n, const std::function<void(int)>& f) {
    if (n <= 0) return;
    if (th_.empty() || n == 1) {
      for (int i = 0; i < n; i++) f(i);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &f;
      n_ = n;
      next_.store(0);
      busy_ = (int)th_.size();
      gen_++;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [this] { return busy_ == 0; });
    fn_ = nullptr;
  }

 private:
  void work() {
    for (;;) {
      int i = next_.fetch_add(1);
      if (i >= n_) break;
      (*fn_)(i);
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        if (stop_) return;
      }
      work();
      {
        std::lock_guard<std::mutex> lk(mu_);
        busy_--;
      }
      done_cv_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* fn_ = nullptr;
  std::atomic<int> next_{0};
  int n_ = 0, busy_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

static int pool_threads() {
  unsigned h = std::thread::hardware_concurrency();
  int n = h ? (int)h : 8;
  return std::max(1, std::min(16, n) - 1);
}

// ---------------------------------------------------------------------------
// jobs and rounds
// ---------------------------------------------------------------------------
struct Job {
  int64_t id = 0;
  std::shared_ptr<const btc::CircuitHost> circ;
  int nbits = 0;
  int prec = BT_C64;
  int fold = 1;
  uint64_t seed = 0, traj = 0;
  void* state = nullptr;  // device
  size_t bytes = 0;
  std::unique_ptr<bti::Interp> ip;
  std::vector<QOp> slots;
  Plan plan;
  PassResult res;
  int has_plan = 0, fresh = 1, done = 0, err = 0, red_swap = 0;
  std::vector<int32_t> outs;
  std::vector<unsigned char> table;
  double t_submit = 0, t_admit = 0, t_done = 0;
  int64_t passes = 0, rpasses = 0, boundary = 0, draws = 0;
  double bytes_moved = 0;
};

static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

class RoundEngine {
 public:
  explicit RoundEngine(int prec) : prec_(prec), pool_(pool_threads()) {}
  ~RoundEngine() { release(); }

  // shared != nullptr: launch on that stream (not owned), e.g. two engines
  // whose rounds alternate on one stream so the GPU runs one cohort while
  // the host plans the other
  int init(hipStream_t shared = nullptr) {
    int rc = ensure_device();
    if (rc) return rc;
    if (shared) {
      stream_ = shared;
      own_stream_ = false;
    } else {
      BT_HIP(hipStreamCreate(&stream_));
      own_stream_ = true;
    }
    BT_HIP(hipEventCreate(&ev_a_));
    BT_HIP(hipEventCreate(&ev_b_));
    BT_HIP(hipEventCreateWithFlags(&ev_done_, hipEventDisableTiming));
    return BT_OK;
  }

  void release() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (h_blob_) (void)hipHostFree(h_blob_);
    if (d_blob_) (void)hipFree(d_blob_);
    if (h_items_) (void)hipHostFree(h_items_);
    if (d_items_) (void)hipFree(d_items_);
    if (d_part_) (void)hipFree(d_part_);
    if (d_out_) (void)hipFree(d_out_);
    if (h_out_) (void)hipHostFree(h_out_);
    if (ev_a_) (void)hipEventDestroy(ev_a_);
    if (ev_b_) (void)hipEventDestroy(ev_b_);
    if (ev_done_) (void)hipEventDestroy(ev_done_);
    if (stream_ && own_stream_) (void)hipStreamDestroy(stream_);
    ev_a_ = ev_b_ = ev_done_ = nullptr;
    h_blob_ = d_blob_ = nullptr;
    h_items_ = nullptr;
    d_items_ = nullptr;
    d_part_ = d_out_ = h_out_ = nullptr;
    stream_ = nullptr;
  }

  hipStream_t stream() const { return stream_; }
  int tile_T(int nbits) const { return tile_bits_for(prec_, nbits); }
  size_t elem() const { return prec_ == BT_C64 ? 8 : 16; }

  void start_job(Job* j) {
    j->ip.reset(new bti::Interp());
    j->slots.assign(bti::kQMax, QOp());
    bti::Program prog{j->circ->ops.data(), j->circ->ntop, j->circ->pool.data(), j->circ->chan.data(),
                      j->circ->chanw.data(), j->nbits};
    // density-matrix jobs (no norm folding): a superoperator spans up to
    // kMaxQK bits of vec(rho), so a smaller tile setting is raised to hold it
    const int T = j->fold ? tile_T(j->nbits) : dm_tile_bits_for(prec_, j->nbits);
    j->ip->init(prog, j->slots.data(), T, coalesce_bits(prec_));
    j->ip->seed = j->seed;
    j->ip->traj = j->traj;
    j->ip->fold = j->fold;
    j->ip->tol_exact = prec_ == BT_C64 ? 1e-6 : 1e-12;
    j->outs.assign(std::max<size_t>(16, 2 * j->circ->ops.size()), 0);
    j->ip->outs = j->outs.data();
    j->ip->outcap = (int64_t)j->outs.size();
    j->has_plan = 0;
    j->fresh = 1;
    j->done = 0;
    j->err = 0;
    memset(&j->res, 0, sizeof j->res);
  }

  // One round over `jobs` (all unfinished).  Returns BT_OK or an error.
  int round(std::vector<Job*>& jobs) {
    int rc = launch_round(jobs);
    if (rc) return rc;
    return finish_round();
  }

  // Plan and enqueue the next pass of every job; finish_round() waits for it
  // and hands the reductions back.  Returns BT_OK or an error.
  int launch_round(std::vector<Job*>& jobs) {
    // 1. advance every interpreter to its next pass and build its table
    pool_.parallel_for((int)jobs.size(), [&](int i) {
      Job* j = jobs[i];
      if (j->done) return;
      const int c = j->ip->step(&j->plan, &j->res);
      if (j->ip->err) {
        j->err = j->ip->err;
        j->done = 1;
        j->has_plan = 0;
        return;
      }
      if (!c) {
        j->done = 1;
        j->has_plan = 0;
        return;
      }
      j->has_plan = 1;
      const size_t need = kGeomBytes + 16 + sizeof(C<double>) * kMatEntries;
      j->table.resize(need);
      PassGeom* g = (PassGeom*)j->table.data();
      int nm = 0, grc;
      if (prec_ == BT_C64) {
        C<float>* m = (C<float>*)(j->table.data() + kGeomBytes + 16);
        grc = build_geom<float>(j->plan, j->slots.data(), j->nbits, g, m);
      } else {
        C<double>* m = (C<double>*)(j->table.data() + kGeomBytes + 16);
        grc = build_geom<double>(j->plan, j->slots.data(), j->nbits, g, m);
      }
      if (grc) {
        j->err = grc;
        j->done = 1;
        j->has_plan = 0;
        return;
      }
      j->red_swap = g->red_swap;
      nm = g->nmats;
      *(int*)(j->table.data() + kGeomBytes) = nm;
      j->table.resize(kGeomBytes + 16 + elem() * nm);
    });
    // 2. pack
    act_.clear();
    for (Job* j : jobs)
      if (j->has_plan) act_.push_back(j);
    std::vector<Job*>& act = act_;
    pending_ = false;
    if (act.empty()) return BT_OK;
    // tiles per workgroup: enough to amortise the workgroup's set-up, few
    // enough that the largest item still spreads over the chip and that the
    // launch has workgroups for several waves of the chip's slots — a
    // workgroup walks its tiles in turn, so in a launch with little total
    // work (a serving round) many tiles per workgroup only lengthen it
    const int setting = tiles_per_block();
    size_t blob = 0;
    uint32_t blocks = 0;
    std::vector<RoundItem> items(act.size());
    size_t maxT = 0;
    uint32_t nb_uniform = 0;
    for (size_t i = 0; i < act.size(); i++) {
      Job* j = act[i];
      const int T = j->plan.ntile;
      maxT = std::max<size_t>(maxT, (size_t)T);
      RoundItem& it = items[i];
      it.state = j->state;
      it.table = blob;
      it.ntiles = (uint32_t)(1ull << (j->nbits - T));
      it.tpb = (uint32_t)item_tpb(setting, it.ntiles, cap_div_);
      it.block0 = blocks;
      it.nblocks = (it.ntiles + it.tpb - 1) / it.tpb;
      it.fresh = j->fresh;
      if (i == 0) nb_uniform = it.nblocks;
      else if (nb_uniform != it.nblocks) nb_uniform = 0;
      blocks += it.nblocks;
      blob += (j->table.size() + 15) / 16 * 16;
    }
    nacc_ = kAccSmall;
    for (Job* j : act)
      if (j->plan.red == bti::RED_RHO && j->plan.red_k == 3) nacc_ = kAcc;
    int rc = reserve(blob, act.size(), (uint64_t)blocks * (uint64_t)kAcc);
    if (rc) return rc;
    pool_.parallel_for((int)act.size(), [&](int i) {
      memcpy((unsigned char*)h_blob_ + items[i].table, act[i]->table.data(), act[i]->table.size());
    });
    memcpy(h_items_, items.data(), sizeof(RoundItem) * items.size());
    BT_HIP(hipMemcpyAsync(d_blob_, h_blob_, blob, hipMemcpyHostToDevice, stream_));
    BT_HIP(hipMemcpyAsync(d_items_, h_items_, sizeof(RoundItem) * items.size(), hipMemcpyHostToDevice, stream_));
    // 3. launch
    const size_t lds = (size_t)(1ull << maxT) * elem();
    BT_HIP(hipEventRecord(ev_a_, stream_));
    if (prec_ == BT_C64) {
      (void)hipFuncSetAttribute((const void*)k_round<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL((k_round<float>), dim3(blocks), dim3(TileCfg<float>::NT), lds, stream_, d_items_,
                         (int)items.size(), nb_uniform, (const unsigned char*)d_blob_, d_part_);
    } else {
      (void)hipFuncSetAttribute((const void*)k_round<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL((k_round<double>), dim3(blocks), dim3(TileCfg<double>::NT), lds, stream_, d_items_,
                         (int)items.size(), nb_uniform, (const unsigned char*)d_blob_, d_part_);
    }
    BT_HIP(hipGetLastError());
    BT_HIP(hipEventRecord(ev_b_, stream_));
    hipLaunchKernelGGL(k_round_finalize, dim3((unsigned)items.size(), (unsigned)nacc_), dim3(256), 0, stream_,
                       d_items_, d_part_, (uint64_t)blocks, d_out_);
    BT_HIP(hipGetLastError());
    BT_HIP(hipMemcpyAsync(h_out_, d_out_, sizeof(double) * nacc_ * items.size(), hipMemcpyDeviceToHost, stream_));
    BT_HIP(hipEventRecord(ev_done_, stream_));
    pending_ = true;
    launched_blocks_ += blocks;
    return BT_OK;
  }

  int finish_round() {
    if (!pending_) return BT_OK;
    pending_ = false;
    BT_HIP(hipEventSynchronize(ev_done_));
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev_a_, ev_b_) == hipSuccess) pass_ms_ += ms;
    rounds_++;
    // 4. hand the results back
    for (size_t i = 0; i < act_.size(); i++) {
      Job* j = act_[i];
      to_result(j->plan, j->red_swap, h_out_ + (size_t)nacc_ * i, &j->res);
      if (j->plan.write) j->fresh = 0;
    }
    return BT_OK;
  }

  // collect a finished job's statistics
  static void finish_stats(Job* j, size_t elem) {
    j->passes = j->ip->passes;
    j->rpasses = j->ip->rpasses;
    j->boundary = j->ip->boundary;
    j->draws = j->ip->draw;
    // read+write per pass, read per reduction pass
    j->bytes_moved = j->ip->bytes_units * (double)(1ull << j->nbits) * (double)elem;
  }

  // items get at least ntiles / cap_div workgroups (see item_tpb)
  void set_cap_div(int d) { cap_div_ = d; }
  int64_t rounds() const { return rounds_; }
  double pass_ms() const { return pass_ms_; }
  void reset_counters() {
    rounds_ = 0;
    pass_ms_ = 0;
    launched_blocks_ = 0;
  }

 private:
  // pass tables (bytes), items, reduction slots (doubles)
  int reserve(size_t blob, size_t nitems, uint64_t part_doubles) {  // part: nacc rows x workgroups
    if (blob > cap_blob_) {
      if (h_blob_) (void)hipHostFree(h_blob_);
      if (d_blob_) (void)hipFree(d_blob_);
      h_blob_ = d_blob_ = nullptr;
      cap_blob_ = std::max(blob, cap_blob_ * 2);
      BT_HIP(hipHostMalloc(&h_blob_, cap_blob_));
      BT_HIP(hipMalloc(&d_blob_, cap_blob_));
    }
    if (nitems > cap_items_) {
      if (h_items_) (void)hipHostFree(h_items_);
      if (d_items_) (void)hipFree(d_items_);
      if (d_out_) (void)hipFree(d_out_);
      if (h_out_) (void)hipHostFree(h_out_);
      cap_items_ = std::max(nitems, cap_items_ * 2);
      BT_HIP(hipHostMalloc((void**)&h_items_, sizeof(RoundItem) * cap_items_));
      BT_HIP(hipMalloc((void**)&d_items_, sizeof(RoundItem) * cap_items_));
      BT_HIP(hipMalloc((void**)&d_out_, sizeof(double) * kAcc * cap_items_));
      BT_HIP(hipHostMalloc((void**)&h_out_, sizeof(double) * kAcc * cap_items_));
    }
    if (part_doubles > cap_part_) {
      if (d_part_) (void)hipFree(d_part_);
      cap_part_ = std::max<size_t>(part_doubles, cap_part_ * 2);
      BT_HIP(hipMalloc((void**)&d_part_, sizeof(double) * cap_part_));
    }
    return BT_OK;
  }

  int prec_;
  Pool pool_;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = true;
  hipEvent_t ev_done_ = nullptr;
  std::vector<Job*> act_;
  bool pending_ = false;
  void* h_blob_ = nullptr;
  void* d_blob_ = nullptr;
  size_t cap_blob_ = 0;
  RoundItem* h_items_ = nullptr;
  RoundItem* d_items_ = nullptr;
  size_t cap_items_ = 0;
  double* d_part_ = nullptr;
  size_t cap_part_ = 0;
  double* d_out_ = nullptr;
  double* h_out_ = nullptr;
  int64_t rounds_ = 0;
  int nacc_ = kAccSmall;
  int cap_div_ = 32;
  uint64_t launched_blocks_ = 0;
  hipEvent_t ev_a_ = nullptr, ev_b_ = nullptr;
  double pass_ms_ = 0;
};

// Export a finished job's state as complex128 into host memory.
static int export_state(Job* j, hipStream_t st, double* host, double* d_tmp) {
  const uint64_t n = 1ull << j->nbits;
  if (j->fresh) {  // never written: still |0..0>
    memset(host, 0, sizeof(double) * 2 * n);
    host[0] = 1.0;
    return BT_OK;
  }
  const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 16);
  if (j->prec == BT_C64)
    hipLaunchKernelGGL(k_export_state<float>, dim3(grid), dim3(256), 0, st, (const C<float>*)j->state, n, d_tmp);
  else
    hipLaunchKernelGGL(k_export_state<double>, dim3(grid), dim3(256), 0, st, (const C<double>*)j->state, n, d_tmp);
  BT_HIP(hipGetLastError());
  BT_HIP(hipMemcpyAsync(host, d_tmp, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, st));
  BT_HIP(hipStreamSynchronize(st));
  return BT_OK;
}

}  // namespace bt

namespace bt {

// What a batch needs besides its inputs and outputs, kept between
// bt_run_batch calls: the two cohort engines of each precision (their worker
// thread pools, stream, events and pinned staging) and the slab that holds
// the resident states.  Building them costs tens of milliseconds of host time
// (30 worker threads, a slab of up to 3/4 of HBM), which every call would
// otherwise pay.  One call at a time uses the cache; a call that finds it in
// use builds its own context.  The slab is released by bt_release_cache or
// when another allocation of the library runs out of HBM
// (release_cached_memory).  The cache is never destroyed: its threads and
// device memory last until the process ends.
struct BatchContext {
  std::unique_ptr<RoundEngine> e0, e1;
  void* slab = nullptr;
  size_t slab_bytes = 0;
  int init(int prec) {
    if (e0) return BT_OK;
    std::unique_ptr<RoundEngine> a(new RoundEngine(prec)), b(new RoundEngine(prec));
    int rc = a->init();
    if (rc) return rc;
    rc = b->init(a->stream());
    if (rc) return rc;
    e0 = std::move(a);
    e1 = std::move(b);
    return BT_OK;
  }
  void free_slab() {
    if (slab) {
      if (e0) (void)hipStreamSynchronize(e0->stream());
      (void)hipFree(slab);
    }
    slab = nullptr;
    slab_bytes = 0;
  }
};

struct BatchCache {
  std::mutex mu;  // held for the whole of a bt_run_batch call that uses the cache
  BatchContext ctx[2];
};

static BatchCache& batch_cache() {
  static BatchCache* c = new BatchCache();  // never destroyed (see BatchContext)
  return *c;
}

// Free the cached batch slabs unless a batch is running; returns the bytes
// released.
size_t release_cached_memory() {
  BatchCache& bc = batch_cache();
  std::unique_lock<std::mutex> lk(bc.mu, std::try_to_lock);
  if (!lk.owns_lock()) return 0;
  size_t n = 0;
  for (BatchContext& x : bc.ctx) {
    n += x.slab_bytes;
    x.free_slab();
  }
  return n;
}

}  // namespace bt

using namespace bt;

extern "C" int bt_release_cache(void) {
  (void)release_cached_memory();
  return BT_OK;
}

// The lockstep-rounds batch; samples (nullable): one basis index per
// trajectory drawn from its final state with u(seed, traj, draws used).
static int run_batch(bt_circuit c, int prec, uint64_t seed, int64_t traj0, int64_t ntraj, int* outcomes,
                     int64_t outcap, int64_t* nouts, double* final_states, uint64_t* samples,
                     int64_t* sample_boundary, bt_run_stats* stats) {
  if (!c) return fail(BT_ERR_ARG, "null circuit");
  if (ntraj <= 0 || !outcomes || !nouts || outcap < 0) return fail(BT_ERR_ARG, "bad batch arguments");
  if (prec != BT_C64 && prec != BT_C128) return fail(BT_ERR_ARG, "bad precision");
  int rc = ensure_device();
  if (rc) return rc;
  // two cohorts whose rounds alternate on one stream: while the GPU runs one
  // cohort's pass, the host plans the other's next pass; engines and slab
  // from the cache when it is free, else built for this call
  BatchCache& bc = batch_cache();
  std::unique_lock<std::mutex> cache_lock(bc.mu, std::try_to_lock);
  BatchContext own;
  BatchContext& X = cache_lock.owns_lock() ? bc.ctx[prec == BT_C64 ? 0 : 1] : own;
  rc = X.init(prec);
  if (rc) return rc;
  RoundEngine& eng0 = *X.e0;
  RoundEngine& eng1 = *X.e1;
  eng0.reset_counters();
  eng1.reset_counters();
  RoundEngine* E[2] = {&eng0, &eng1};
  hipStream_t stream = eng0.stream();
  auto circ = std::shared_ptr<const btc::CircuitHost>(&c->h, [](const btc::CircuitHost*) {});
  const int nbits = c->h.n;
  const size_t bytes = (size_t)(1ull << nbits) * eng0.elem();
  int64_t nslots = 0;
  if (X.slab && X.slab_bytes >= bytes * (size_t)ntraj) {
    nslots = ntraj;  // the cached slab holds every state
  } else {
    X.free_slab();
    size_t freeb = 0, totalb = 0;
    BT_HIP(hipMemGetInfo(&freeb, &totalb));
    // resident states: all of them when they fit in 3/4 of free HBM
    nslots = std::min<int64_t>(ntraj, (int64_t)(freeb / 4 * 3 / (bytes + 4096)));
    if (nslots < 1) return fail(BT_ERR_NOMEM, "bt_run_batch: a single state does not fit in HBM");
    // one allocation for all slots
    void* p = nullptr;
    while (nslots >= 1 && hipMalloc(&p, bytes * nslots) != hipSuccess) {
      (void)hipGetLastError();
      nslots /= 2;
    }
    if (nslots < 1) return fail(BT_ERR_NOMEM, "bt_run_batch: state allocation failed");
    X.slab = p;
    X.slab_bytes = bytes * (size_t)nslots;
  }
  std::vector<void*> slot_mem;  // per-call buffers, freed at the end
  double* d_tmp = nullptr;
  auto cleanup = [&]() {
    for (void* p : slot_mem) (void)hipFree(p);
    if (d_tmp) (void)hipFree(d_tmp);
    if (&X == &own) {
      for (int ch = 0; ch < 2; ch++) (void)E[ch]->finish_round();
      own.free_slab();
    }
  };
  if (final_states && hipMalloc((void**)&d_tmp, sizeof(double) * 2 * (1ull << nbits)) != hipSuccess) {
    cleanup();
    return fail(BT_ERR_NOMEM, "bt_run_batch: export buffer");
  }
  // final sampling: scratch for one state at a time (the stream orders them),
  // one index and one boundary flag per trajectory
  char* d_samp = nullptr;
  const size_t samp_scratch = samples ? sample_scratch_bytes(nbits) : 0;
  std::vector<uint8_t> fresh_end(samples ? ntraj : 0, 0);
  if (samples) {
    if (hipMalloc((void**)&d_samp, samp_scratch + 9 * (size_t)ntraj + 64) != hipSuccess) {
      (void)hipGetLastError();
      cleanup();
      return fail(BT_ERR_NOMEM, "bt_run_batch: sampling buffers");
    }
    slot_mem.push_back(d_samp);
  }
  uint64_t* d_sidx = samples ? (uint64_t*)(d_samp + samp_scratch) : nullptr;
  uint8_t* d_sflag = samples ? (uint8_t*)(d_samp + samp_scratch + 8 * (size_t)ntraj) : nullptr;
  std::vector<std::unique_ptr<Job>> jobs(ntraj);
  std::vector<int64_t> free_slots;
  for (int64_t s = nslots - 1; s >= 0; s--) free_slots.push_back(s);
  std::vector<Job*> active[2];
  const int64_t cap[2] = {(nslots + 1) / 2, nslots / 2};
  std::vector<int64_t> slot_of(ntraj, -1);
  int64_t next = 0, done = 0;
  bt_run_stats acc;
  memset(&acc, 0, sizeof acc);
  int first_err = 0;
  auto admit = [&](int ch) {
    while (next < ntraj && !free_slots.empty() && (int64_t)active[ch].size() < cap[ch]) {
      const int64_t s = free_slots.back();
      free_slots.pop_back();
      jobs[next].reset(new Job());
      Job* j = jobs[next].get();
      j->id = next;
      j->circ = circ;
      j->nbits = nbits;
      j->prec = prec;
      j->seed = seed;
      j->traj = (uint64_t)(traj0 + next);
      j->state = (char*)X.slab + bytes * s;
      j->bytes = bytes;
      slot_of[next] = s;
      E[ch]->start_job(j);
      active[ch].push_back(j);
      next++;
    }
  };
  auto retire = [&](int ch) -> int {
    std::vector<Job*> still;
    for (Job* j : active[ch]) {
      if (!j->done) {
        still.push_back(j);
        continue;
      }
      RoundEngine::finish_stats(j, eng0.elem());
      acc.passes += j->passes;
      acc.reduce_passes += j->rpasses;
      acc.boundary_draws += j->boundary;
      acc.draws += j->draws;
      acc.bytes += j->bytes_moved;
      if (j->err && !first_err) first_err = j->err;
      const int64_t t = j->id;
      const int64_t n = std::min<int64_t>(outcap, j->ip->nout);
      for (int64_t i = 0; i < n; i++) outcomes[t * outcap + i] = j->outs[i];
      nouts[t] = j->ip->nout;
      if (final_states) {
        const int erc = export_state(j, stream, final_states + (uint64_t)t * 2 * (1ull << nbits), d_tmp);
        if (erc) return erc;
      }
      if (samples) {
        if (j->fresh) {  // never written: |0..0>, index 0 with certainty
          fresh_end[t] = 1;
        } else {
          const int src = launch_sample(j->state, prec, nbits, stream, d_samp, 1, nullptr,
                                        uniform(seed, j->traj, (uint64_t)j->draws), d_sidx + t, d_sflag + t);
          if (src) return src;
        }
      }
      free_slots.push_back(slot_of[t]);
      jobs[t].reset();
      done++;
    }
    active[ch].swap(still);
    return BT_OK;
  };
  for (int ch = 0; ch < 2 && rc == BT_OK; ch++) {
    admit(ch);
    rc = E[ch]->launch_round(active[ch]);
  }
  while (rc == BT_OK && done < ntraj) {
    for (int ch = 0; ch < 2 && rc == BT_OK; ch++) {
      rc = E[ch]->finish_round();
      if (rc == BT_OK) rc = retire(ch);
      if (rc == BT_OK) {
        admit(ch);
        rc = E[ch]->launch_round(active[ch]);
      }
    }
  }
  for (int ch = 0; ch < 2; ch++) (void)E[ch]->finish_round();
  (void)hipStreamSynchronize(stream);
  if (rc == BT_OK && samples) {
    std::vector<uint8_t> flags(ntraj);
    if (hipMemcpy(samples, d_sidx, 8 * (size_t)ntraj, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(flags.data(), d_sflag, (size_t)ntraj, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = fail(BT_ERR_HIP, "bt_run_batch: sample download");
    } else {
      int64_t nb = 0;
      for (int64_t t = 0; t < ntraj; t++) {
        if (fresh_end[t]) {
          samples[t] = 0;
          flags[t] = 0;
        }
        nb += flags[t];
      }
      if (sample_boundary) *sample_boundary = nb;
    }
  }
  // device time of the pass kernels (HIP events around each launch)
  acc.kernel_ms = eng0.pass_ms() + eng1.pass_ms();
  acc.launches = eng0.rounds() + eng1.rounds();
  if (stats) *stats = acc;
  cleanup();
  if (rc) return rc;
  if (first_err) return fail(first_err, "bt_run_batch: a trajectory failed (interpreter error " +
                                            std::to_string(first_err) + ")");
  return BT_OK;
}

extern "C" int bt_run_batch(bt_circuit c, int prec, uint64_t seed, int64_t traj0, int64_t ntraj, int* outcomes,
                            int64_t outcap, int64_t* nouts, double* final_states, bt_run_stats* stats) {
  return run_batch(c, prec, seed, traj0, ntraj, outcomes, outcap, nouts, final_states, nullptr, nullptr, stats);
}

extern "C" int bt_run_batch_sample(bt_circuit c, int prec, uint64_t seed, int64_t traj0, int64_t ntraj,
                                   int* outcomes, int64_t outcap, int64_t* nouts, uint64_t* samples,
                                   int64_t* sample_boundary, bt_run_stats* stats) {
  if (!samples) return fail(BT_ERR_ARG, "bt_run_batch_sample: null samples");
  return run_batch(c, prec, seed, traj0, ntraj, outcomes, outcap, nouts, nullptr, samples, sample_boundary, stats);
}

// ---------------------------------------------------------------------------
// serving: continuous batching with HBM-budget admission
// ---------------------------------------------------------------------------
// A worker thread owns the round engines.  Each loop it admits queued
// requests in FIFO order while their states fit in the HBM budget (and the
// slot cap), runs one round over every resident request (one launch per
// precision), and retires finished ones.  A finished request keeps its state
// (and its share of the budget) until bt_server_fetch reads it.  Requests
// never influence each other's results: each has its own uniforms
// u(seed, traj, .) and its reductions are summed in a fixed order that does
// not depend on what else is resident.
namespace bt {

template <typename R>
__global__ void k_probs(const C<R>* s, uint64_t n, int dm, double* out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if (dm) {
      out[i] = (double)s[i * (n + 1)].x;  // rho[i, i] (column-major, row bits low)
    } else {
      const C<R> v = s[i];
      out[i] = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
    }
  }
}

struct Request {
  int64_t id = 0;
  int kind = BT_SV, prec = BT_C64, nq = 0;
  uint64_t seed = 0, traj = 0;
  size_t bytes = 0;
  std::shared_ptr<const btc::CircuitHost> circ;
  std::unique_ptr<Job> job;
  std::vector<int32_t> outs;
  int64_t nout = 0;
  int fresh = 1;
  double t_submit = 0, t_admit = 0;
  bt_result res{};
};

}  // namespace bt

struct bt_server_s {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  bool stop = false;
  int device = 0;
  uint64_t budget = 0;
  int max_slots = 0;
  int64_t next_id = 1;
  std::deque<std::shared_ptr<bt::Request>> queue;
  int nrunning = 0;
  std::deque<bt_result> unpolled;
  std::unordered_map<int64_t, std::shared_ptr<bt::Request>> finished;
  uint64_t in_use = 0;   // states of resident + finished-unfetched requests
  uint64_t cached = 0;   // bytes in the free lists
  std::unordered_map<size_t, std::vector<void*>> free_bufs;
  std::vector<void*> to_release;  // freed by the worker (it owns allocation)
  std::unordered_map<uint64_t, std::shared_ptr<const btc::CircuitHost>> sv_cache, dm_cache;
  std::unique_ptr<bt::RoundEngine> eng[2][2];  // [precision][cohort]; a precision's cohorts share one stream
  hipStream_t fetch_stream = nullptr;
  double* d_probs = nullptr;
  size_t cap_probs = 0;
  std::mutex fetch_mu;
};

namespace bt {

static size_t state_bytes(int kind, int prec, int nq) {
  const int nbits = kind == BT_DM ? 2 * nq : nq;
  return (size_t)(1ull << nbits) * (prec == BT_C64 ? 8 : 16);
}

// caller holds srv->mu; only the worker thread calls this
static void* take_buffer(bt_server_s* S, size_t bytes) {
  auto it = S->free_bufs.find(bytes);
  if (it != S->free_bufs.end() && !it->second.empty()) {
    void* p = it->second.back();
    it->second.pop_back();
    S->cached -= bytes;
    return p;
  }
  // release cached buffers of other sizes until the new one fits the budget
  for (auto& kv : S->free_bufs) {
    while (!kv.second.empty() && S->in_use + S->cached + bytes > S->budget) {
      (void)hipFree(kv.second.back());
      kv.second.pop_back();
      S->cached -= kv.first;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    if (!release_cached_memory() || hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  return p;
}

static void finish_request(bt_server_s* S, Request* r, int err) {
  Job* j = r->job.get();
  r->res.id = r->id;
  r->res.status = err ? -std::abs(err) : 0;
  if (j && j->ip) {
    r->nout = j->ip->nout;
    r->outs.assign(j->outs.begin(), j->outs.begin() + std::min<int64_t>((int64_t)j->outs.size(), r->nout));
    r->fresh = j->fresh;
  }
  r->res.nout = (int32_t)r->nout;
  const double t = now_ms();
  r->res.latency_ms = t - r->t_submit;
  r->res.queue_ms = r->t_admit - r->t_submit;
  (void)S;
}

// The worker: requests are split into two cohorts whose rounds alternate on
// each precision's stream, as in bt_run_batch: while the device runs one
// cohort's pass, the host plans the other's next pass (interpreters and pass
// tables on the engine's thread pool).  Each cycle, for cohort c: wait for
// its last round, retire what finished, admit queued requests FIFO while
// they fit the HBM budget and the slot cap (into cohort c), and launch its
// next round (one launch per precision over the cohort's requests).
static void server_loop(bt_server_s* S) {
  (void)hipSetDevice(S->device);
  std::vector<std::shared_ptr<Request>> running[2];  // per cohort
  auto engine = [&](int prec, int c) -> RoundEngine* {
    const int e = prec == BT_C64 ? 0 : 1;
    if (!S->eng[e][0]) {
      std::unique_ptr<RoundEngine> a(new RoundEngine(prec)), b(new RoundEngine(prec));
      if (a->init() || b->init(a->stream())) return nullptr;
      a->set_cap_div(256);  // latency: a lone request still spreads over the chip
      b->set_cap_div(256);
      S->eng[e][0] = std::move(a);
      S->eng[e][1] = std::move(b);
    }
    return S->eng[e][c].get();
  };
  auto reject = [&](const std::shared_ptr<Request>& r, int err) {
    r->t_admit = now_ms();
    finish_request(S, r.get(), err);
    S->finished[r->id] = r;
    S->unpolled.push_back(r->res);
    S->cv_done.notify_all();
  };
  for (;;) {
    bool idle = true;
    for (int c = 0; c < 2; c++) {
      // 1. the cohort's launched round: wait, hand back the reductions
      for (int e = 0; e < 2; e++)
        for (int cc = 0; cc < 2; cc++)
          if (cc == c && S->eng[e][cc]) {
            const int rc = S->eng[e][cc]->finish_round();
            if (rc)
              for (auto& r : running[c])
                if ((r->prec == BT_C64 ? 0 : 1) == e && !r->job->done) {
                  r->job->done = 1;
                  if (!r->job->err) r->job->err = rc;
                }
          }
      std::unique_lock<std::mutex> lk(S->mu);
      // 2. retire
      std::vector<std::shared_ptr<Request>> still;
      for (auto& r : running[c]) {
        if (!r->job->done) {
          still.push_back(r);
          continue;
        }
        finish_request(S, r.get(), r->job->err);
        r->job->ip.reset();
        S->finished[r->id] = r;
        S->unpolled.push_back(r->res);
      }
      if (still.size() != running[c].size()) S->cv_done.notify_all();
      running[c].swap(still);
      for (void* p : S->to_release) (void)hipFree(p);
      S->to_release.clear();
      // 3. admission, FIFO, into this cohort
      while (!S->stop && !S->queue.empty() && (int)(running[0].size() + running[1].size()) < S->max_slots) {
        std::shared_ptr<Request> r = S->queue.front();
        if (r->bytes > S->budget) {  // can never fit
          S->queue.pop_front();
          reject(r, BT_ERR_NOMEM);
          continue;
        }
        if (S->in_use + r->bytes > S->budget) break;
        void* mem = take_buffer(S, r->bytes);
        if (!mem) {
          if (!running[0].empty() || !running[1].empty()) break;  // retry when a resident request retires
          S->queue.pop_front();  // nothing resident: HBM is really short
          reject(r, BT_ERR_NOMEM);
          continue;
        }
        S->queue.pop_front();
        S->in_use += r->bytes;
        r->t_admit = now_ms();
        RoundEngine* E = engine(r->prec, c);
        r->job.reset(new Job());
        Job* j = r->job.get();
        j->id = r->id;
        j->circ = r->circ;
        j->nbits = r->kind == BT_DM ? 2 * r->nq : r->nq;
        j->prec = r->prec;
        j->fold = r->kind == BT_DM ? 0 : 1;
        j->seed = r->seed;
        j->traj = r->traj;
        j->state = mem;
        j->bytes = r->bytes;
        if (!E) {
          j->done = 1;
          j->err = BT_ERR_HIP;
        } else {
          E->start_job(j);
        }
        running[c].push_back(r);
      }
      S->nrunning = (int)(running[0].size() + running[1].size());
      if (S->stop) {
        lk.unlock();
        for (int e = 0; e < 2; e++)
          for (int cc = 0; cc < 2; cc++)
            if (S->eng[e][cc]) (void)S->eng[e][cc]->finish_round();
        lk.lock();
        for (int cc = 0; cc < 2; cc++) {
          for (auto& r : running[cc]) {
            S->free_bufs[r->bytes].push_back(r->job->state);
            S->cached += r->bytes;
            S->in_use -= r->bytes;
          }
          running[cc].clear();
        }
        S->nrunning = 0;
        return;
      }
      lk.unlock();
      // 4. the cohort's next round, one launch per precision
      for (int e = 0; e < 2; e++) {
        if (!S->eng[e][c]) continue;
        std::vector<Job*> jobs;
        for (auto& r : running[c])
          if ((r->prec == BT_C64 ? 0 : 1) == e && !r->job->done) jobs.push_back(r->job.get());
        if (jobs.empty()) continue;
        idle = false;
        const int rc = S->eng[e][c]->launch_round(jobs);
        if (rc)
          for (Job* j : jobs) {
            j->done = 1;
            if (!j->err) j->err = rc;
          }
      }
      if (!running[c].empty()) idle = false;
    }
    if (idle) {
      std::unique_lock<std::mutex> lk(S->mu);
      if (running[0].empty() && running[1].empty()) {
        if (S->stop) return;
        S->cv_work.wait(lk, [&] {
          return S->stop || (!S->queue.empty() &&
                             (S->queue.front()->bytes > S->budget || S->in_use + S->queue.front()->bytes <= S->budget));
        });
      }
    }
  }
}

}  // namespace bt

extern "C" int bt_server_create(uint64_t hbm_budget_bytes, int max_slots, bt_server* out) {
  if (!out) return fail(BT_ERR_ARG, "null output");
  *out = nullptr;
  int rc = ensure_device();
  if (rc) return rc;
  size_t freeb = 0, totalb = 0;
  if (!hbm_budget_bytes) (void)release_cached_memory();  // the default budget is measured against free HBM
  BT_HIP(hipMemGetInfo(&freeb, &totalb));
  auto* S = new bt_server_s();
  S->budget = hbm_budget_bytes ? hbm_budget_bytes : (uint64_t)(freeb / 4 * 3);
  S->max_slots = max_slots > 0 ? max_slots : 1 << 20;
  (void)hipGetDevice(&S->device);
  if (hipStreamCreate(&S->fetch_stream) != hipSuccess) {
    delete S;
    return fail(BT_ERR_HIP, "bt_server_create: stream");
  }
  S->worker = std::thread(server_loop, S);
  *out = S;
  return BT_OK;
}

extern "C" int bt_server_submit(bt_server S, bt_circuit c, int kind, int prec, uint64_t seed, uint64_t traj,
                                int64_t* id) {
  if (!S || !c || !id) return fail(BT_ERR_ARG, "null server, circuit or id");
  if (kind != BT_SV && kind != BT_DM) return fail(BT_ERR_ARG, "bad state kind");
  if (prec != BT_C64 && prec != BT_C128) return fail(BT_ERR_ARG, "bad precision");
  auto r = std::make_shared<Request>();
  r->kind = kind;
  r->prec = prec;
  r->nq = c->h.n;
  r->seed = seed;
  r->traj = traj;
  r->bytes = state_bytes(kind, prec, c->h.n);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto& cache = kind == BT_DM ? S->dm_cache : S->sv_cache;
    auto it = cache.find(c->uid);
    if (it != cache.end()) {
      r->circ = it->second;
    } else {
      auto h = std::make_shared<btc::CircuitHost>();
      if (kind == BT_DM) {
        std::string msg;
        const int rc = lower_dm(c->h, h.get(), &msg);
        if (rc) return fail(rc, msg);
      } else {
        *h = c->h;
      }
      cache[c->uid] = h;
      r->circ = h;
    }
    r->id = S->next_id++;
    r->t_submit = now_ms();
    *id = r->id;
    S->queue.push_back(r);
  }
  S->cv_work.notify_all();
  return BT_OK;
}

extern "C" int bt_server_poll(bt_server S, bt_result* results, int cap, int wait, int* n) {
  if (!S || !n || (cap > 0 && !results)) return fail(BT_ERR_ARG, "bad poll arguments");
  std::unique_lock<std::mutex> lk(S->mu);
  if (wait)
    S->cv_done.wait(lk, [&] { return !S->unpolled.empty() || (S->queue.empty() && S->nrunning == 0); });
  int k = 0;
  while (k < cap && !S->unpolled.empty()) {
    results[k++] = S->unpolled.front();
    S->unpolled.pop_front();
  }
  *n = k;
  return BT_OK;
}

extern "C" int bt_server_reserve(bt_server S, int kind, int prec, int nqubits, int count, int* added) {
  if (!S) return fail(BT_ERR_ARG, "null server");
  if (kind != BT_SV && kind != BT_DM) return fail(BT_ERR_ARG, "bad state kind");
  if (prec != BT_C64 && prec != BT_C128) return fail(BT_ERR_ARG, "bad precision");
  const int nbits = kind == BT_DM ? 2 * nqubits : nqubits;
  if (nqubits < 1 || nbits > 36 || count < 0) return fail(BT_ERR_ARG, "bt_server_reserve: bad size or count");
  const size_t bytes = state_bytes(kind, prec, nqubits);
  int n = 0;
  (void)hipSetDevice(S->device);
  for (; n < count; n++) {
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (S->in_use + S->cached + bytes > S->budget) break;
      S->cached += bytes;  // counted now so that concurrent reservations respect the budget
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      std::lock_guard<std::mutex> lk(S->mu);
      S->cached -= bytes;
      break;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    S->free_bufs[bytes].push_back(p);
  }
  if (added) *added = n;
  S->cv_work.notify_all();
  return BT_OK;
}

extern "C" int bt_server_fetch(bt_server S, int64_t id, int* outcomes, int64_t outcap, int64_t* nout, double* probs,
                               uint64_t nprobs) {
  if (!S) return fail(BT_ERR_ARG, "null server");
  std::shared_ptr<Request> r;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->finished.find(id);
    if (it == S->finished.end()) return fail(BT_ERR_ARG, "unknown or unfinished request");
    r = it->second;
  }
  const uint64_t n = 1ull << r->nq;
  if (probs && nprobs < n) return fail(BT_ERR_ARG, "probability buffer too small");
  if (nout) *nout = r->nout;
  if (outcomes)
    for (int64_t i = 0; i < std::min<int64_t>(outcap, (int64_t)r->outs.size()); i++) outcomes[i] = r->outs[i];
  int rc = BT_OK;
  if (probs && r->job && r->res.status == 0) {
    std::lock_guard<std::mutex> fl(S->fetch_mu);
    if (r->fresh) {  // never written: |0..0>
      for (uint64_t i = 0; i < n; i++) probs[i] = i == 0 ? 1.0 : 0.0;
    } else {
      if (n > S->cap_probs) {
        if (S->d_probs) (void)hipFree(S->d_probs);
        S->d_probs = nullptr;
        if (hipMalloc((void**)&S->d_probs, sizeof(double) * n) != hipSuccess) {
          S->cap_probs = 0;
          return fail(BT_ERR_NOMEM, "bt_server_fetch: probability buffer");
        }
        S->cap_probs = n;
      }
      const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 1u << 16);
      const int dm = r->kind == BT_DM ? 1 : 0;
      if (r->prec == BT_C64)
        hipLaunchKernelGGL(k_probs<float>, dim3(grid), dim3(256), 0, S->fetch_stream, (const C<float>*)r->job->state,
                           n, dm, S->d_probs);
      else
        hipLaunchKernelGGL(k_probs<double>, dim3(grid), dim3(256), 0, S->fetch_stream,
                           (const C<double>*)r->job->state, n, dm, S->d_probs);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(probs, S->d_probs, sizeof(double) * n, hipMemcpyDeviceToHost, S->fetch_stream) !=
              hipSuccess ||
          hipStreamSynchronize(S->fetch_stream) != hipSuccess)
        rc = fail(BT_ERR_HIP, "bt_server_fetch: probabilities");
    }
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (r->job && r->job->state) {
      S->free_bufs[r->bytes].push_back(r->job->state);
      S->cached += r->bytes;
      S->in_use -= r->bytes;
      r->job->state = nullptr;
    }
    S->finished.erase(id);
  }
  S->cv_work.notify_all();
  return rc;
}

extern "C" int bt_server_stats(bt_server S, int* active, int* queued, uint64_t* hbm_in_use) {
  if (!S) return fail(BT_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(S->mu);
  if (active) *active = S->nrunning;
  if (queued) *queued = (int)S->queue.size();
  if (hbm_in_use) *hbm_in_use = S->in_use;
  return BT_OK;
}

extern "C" int bt_server_destroy(bt_server S) {
  if (!S) return BT_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  for (auto& kv : S->finished)
    if (kv.second->job && kv.second->job->state) (void)hipFree(kv.second->job->state);
  for (auto& kv : S->free_bufs)
    for (void* p : kv.second) (void)hipFree(p);
  for (void* p : S->to_release) (void)hipFree(p);
  for (int e = 0; e < 2; e++)
    for (int c = 1; c >= 0; c--) S->eng[e][c].reset();  // cohort 1 uses cohort 0's stream
  if (S->d_probs) (void)hipFree(S->d_probs);
  if (S->fetch_stream) (void)hipStreamDestroy(S->fetch_stream);
  delete S;
  return BT_OK;
}
