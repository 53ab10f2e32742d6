// Circuit executors for gfx950.
//
//  * k_batch: persistent batched-trajectory kernel.  One workgroup owns one
//    trajectory at a time (its state in HBM, its interpreter driven by lane 0,
//    its reductions private to the workgroup: no grid-wide synchronisation),
//    and pulls the next trajectory index from an atomic counter until all are
//    done.  Every pass streams the state through LDS in tiles of 2^T
//    amplitudes and applies all of the pass's fused operators to the tile.
//  * k_grid_pass: the same tile pass spread over a grid, for one large state
//    (bt_run_circuit / bt_run_circuit_dm), driven by the interpreter on the
//    host; reductions finish in a fixed order (bit-reproducible).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "bt_circuit.hpp"
#include "bt_internal.hpp"
#include "bt_interp.hpp"
#include "bt_tile.hpp"

namespace bt {

// ---------------------------------------------------------------------------
// persistent batched trajectories
// ---------------------------------------------------------------------------
struct BatchArgs {
  void* states;       // one state per workgroup
  QOp* slots;         // kQMax per workgroup
  bti::Interp* interps;
  bti::Program prog;
  int32_t nbits, T, cmin, want_final;
  uint64_t seed;
  int64_t traj0, ntraj;
  unsigned long long* counter;
  int32_t* outcomes;
  int64_t outcap;
  int64_t* nouts;
  double* final_states;
  double tol_exact;
  unsigned long long* stats;  // passes, rpasses, boundary, draws, overflow, bytes_units*2
  int32_t* err;
  int64_t max_passes;
};

template <typename R>
__global__ void __launch_bounds__(TileCfg<R>::NT, TileCfg<R>::WAVES) k_batch(BatchArgs a) {
  constexpr int NT = TileCfg<R>::NT;
  extern __shared__ __align__(16) unsigned char smem[];
  C<R>* tile = (C<R>*)smem;
  __shared__ PassGeom geom;
  __shared__ C<R> mats[kMatEntries];
  __shared__ double red[(NT / 64) * kAcc];
  __shared__ double sums[kAcc];
  __shared__ Plan plan;
  __shared__ PassResult res;
  __shared__ int cmd;
  __shared__ unsigned long long traj_sh;

  const uint64_t len = 1ull << a.nbits;
  C<R>* state = (C<R>*)a.states + (uint64_t)blockIdx.x * len;
  bti::Interp* ip = a.interps + blockIdx.x;
  const uint64_t ntiles = 1ull << (a.nbits - (a.T < a.nbits ? a.T : a.nbits));

  for (;;) {
    if (threadIdx.x == 0) traj_sh = atomicAdd(a.counter, 1ull);
    __syncthreads();
    const unsigned long long t = traj_sh;
    if (t >= (unsigned long long)a.ntraj) break;
    if (threadIdx.x == 0) {
      ip->init(a.prog, a.slots + (uint64_t)blockIdx.x * bti::kQMax, a.T, a.cmin);
      ip->seed = a.seed;
      ip->traj = (uint64_t)(a.traj0 + (int64_t)t);
      ip->outs = a.outcomes + (int64_t)t * a.outcap;
      ip->outcap = a.outcap;
      ip->tol_exact = a.tol_exact;
      ip->normalize_end = a.want_final;
    }
    __syncthreads();
    int64_t npass = 0;
    int fresh = 1;
    for (;;) {
      if (threadIdx.x == 0) {
        int c = ip->step(&plan, &res);
        if (c && ++npass > a.max_passes) {
          ip->err = BT_ERR_STATE;
          c = 0;
        }
        if (c) {
          const int rc = build_geom<R>(plan, ip->slots, a.nbits, &geom, mats);
          if (rc) {
            ip->err = rc;
            c = 0;
          }
        }
        cmd = c;
      }
      __syncthreads();
      if (!cmd) break;
      if (threadIdx.x < kAcc) sums[threadIdx.x] = 0.0;
      __syncthreads();
      tile_pass<R, uint32_t>(state, geom, mats, tile, 0, 1, ntiles, red, sums, fresh);
      if (geom.write) fresh = 0;
      if (threadIdx.x == 0) to_result(plan, geom.red_swap, sums, &res);
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      a.nouts[t] = ip->nout;
      atomicAdd(&a.stats[0], (unsigned long long)ip->passes);
      atomicAdd(&a.stats[1], (unsigned long long)ip->rpasses);
      atomicAdd(&a.stats[2], (unsigned long long)ip->boundary);
      atomicAdd(&a.stats[3], (unsigned long long)ip->draw);
      atomicAdd(&a.stats[4], (unsigned long long)ip->overflow);
      if (ip->err) atomicCAS((int*)a.err, 0, ip->err);
    }
    if (a.want_final) {
      double* out = a.final_states + (uint64_t)t * 2 * len;
      for (uint64_t i = threadIdx.x; i < len; i += NT) {
        C<R> v = fresh ? cmake<C<R>>(i == 0 ? (R)1 : (R)0, (R)0) : state[i];
        out[2 * i] = (double)v.x;
        out[2 * i + 1] = (double)v.y;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// grid pass for one large state (host-driven interpreter)
// ---------------------------------------------------------------------------
template <typename R>
struct GridTable {
  PassGeom g;
  C<R> mats[kMatEntries];
};

// O: offset type of the tile pass (uint32_t up to 2^32 amplitudes)
template <typename R, typename O>
__global__ void __launch_bounds__(TileCfg<R>::NT, TileCfg<R>::WAVES) k_grid_pass(C<R>* state, const GridTable<R>* tab, uint64_t ntiles,
                                                              double* part) {
  constexpr int NT = TileCfg<R>::NT;
  extern __shared__ __align__(16) unsigned char smem[];
  C<R>* tile = (C<R>*)smem;
  __shared__ PassGeom geom;
  __shared__ C<R> mats[kMatEntries];
  __shared__ double red[(NT / 64) * kAcc];
  __shared__ double sums[kAcc];
  load_geom(&tab->g, &geom, threadIdx.x, NT);
  {
    const int nm = tab->g.nmats;
    for (int i = threadIdx.x; i < nm; i += NT) mats[i] = tab->mats[i];
  }
  if (threadIdx.x < kAcc) sums[threadIdx.x] = 0.0;
  __syncthreads();
  tile_pass<R, O>(state, geom, mats, tile, blockIdx.x, gridDim.x, ntiles, red, sums);
  if (threadIdx.x < kAcc) part[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x] = sums[threadIdx.x];
}

// out[j] = fixed-order sum of the grid's partials of accumulator j
__global__ void __launch_bounds__(256) k_finalize_acc(const double* part, int nblocks, double* out) {
  const double t = sum_blocks(part + (uint64_t)blockIdx.x * nblocks, (uint32_t)nblocks);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

}  // namespace bt

// ---------------------------------------------------------------------------
// circuits
// ---------------------------------------------------------------------------

namespace bt {

static int upload_circuit(bt_circuit_s* c) {
  std::lock_guard<std::mutex> lk(c->mu);
  if (c->d_ops) return BT_OK;
  const size_t nops = std::max<size_t>(1, c->h.ops.size());
  BT_HIP(hipMalloc(&c->d_ops, sizeof(bt_ir_op) * nops));
  BT_HIP(hipMalloc(&c->d_pool, sizeof(double) * std::max<size_t>(2, c->h.pool.size())));
  BT_HIP(hipMalloc(&c->d_chan, sizeof(bti::ChanDesc) * nops));
  BT_HIP(hipMalloc(&c->d_chanw, sizeof(double) * std::max<size_t>(1, c->h.chanw.size())));
  if (!c->h.ops.empty()) {
    BT_HIP(hipMemcpy(c->d_ops, c->h.ops.data(), sizeof(bt_ir_op) * c->h.ops.size(), hipMemcpyHostToDevice));
    BT_HIP(hipMemcpy(c->d_chan, c->h.chan.data(), sizeof(bti::ChanDesc) * c->h.ops.size(), hipMemcpyHostToDevice));
  }
  if (!c->h.pool.empty())
    BT_HIP(hipMemcpy(c->d_pool, c->h.pool.data(), sizeof(double) * c->h.pool.size(), hipMemcpyHostToDevice));
  BT_HIP(hipMemcpy(c->d_chanw, c->h.chanw.data(), sizeof(double) * c->h.chanw.size(), hipMemcpyHostToDevice));
  return BT_OK;
}

// Tile width for a state of nbits: the register tile of a 256-thread
// workgroup (2^(8 + register bits) amplitudes), or an override from the
// autotuner; likewise the coalescing reserve and the rounds executor's
// tiles per workgroup.
static int g_tile_override[2] = {0, 0};
static int g_coalesce_override[2] = {0, 0};
static int g_tpb_override = 0;

int tiles_per_block() { return g_tpb_override ? g_tpb_override : 8; }

// Tiles per workgroup of a rounds item of ntiles tiles: the setting, capped
// so that the item still spreads over at least ntiles / cap_div workgroups
// (a workgroup walks its tiles in turn).  It depends on the item alone, so
// an item's reductions, summed per workgroup and then over its workgroups in
// a fixed order, are the same whatever else a launch holds.  The batch
// executor uses cap_div = 32, the server 256 (a lone request still fills the
// chip: latency).
int item_tpb(int setting, uint64_t ntiles, int cap_div) {
  return (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)setting, ntiles / (uint64_t)cap_div));
}

int default_tile_bits(int prec) {
  return prec == BT_C64 ? TileCfg<float>::LOGNT + TileCfg<float>::RB : TileCfg<double>::LOGNT + TileCfg<double>::RB;
}

// Low global bits every tile keeps (runs of 2^c consecutive amplitudes per
// wavefront access): 128 B for either precision.
int coalesce_bits(int prec) {
  const int p = prec == BT_C64 ? 0 : 1;
  return g_coalesce_override[p] ? g_coalesce_override[p] : (prec == BT_C64 ? 4 : 3);
}

int tile_bits_for(int prec, int nbits) {
  int T = g_tile_override[prec] ? g_tile_override[prec] : default_tile_bits(prec);
  return std::min(T, nbits);
}


static_assert(kGridAccRows == kAcc, "grid pass partial rows");

// Workgroups of a grid pass over ntiles tiles: eight per CU at most.
static unsigned grid_pass_blocks(uint64_t ntiles) {
  const int cus = ctx().cus > 0 ? ctx().cus : 256;
  return (unsigned)std::min<uint64_t>(ntiles, (uint64_t)cus * 8);
}

constexpr int kRing = 8;  // pass tables in flight

// The circuit executor's staging of a state, made on its first run and kept
// until the state is destroyed (a state is used by one host thread at a
// time): pinned and device rings of pass tables, reduction partials for the
// widest grid, the finished sums and their pinned copy, events.
struct GridRes {
  void* h_ring = nullptr;
  void* d_ring = nullptr;
  double* d_part = nullptr;
  double* d_acc = nullptr;
  double* h_acc = nullptr;
  hipEvent_t ev[kRing] = {};
  hipEvent_t e0 = nullptr, e1 = nullptr;
};

void free_grid_res(GridRes* g) {
  if (!g) return;
  for (int i = 0; i < kRing; i++)
    if (g->ev[i]) (void)hipEventDestroy(g->ev[i]);
  if (g->e0) (void)hipEventDestroy(g->e0);
  if (g->e1) (void)hipEventDestroy(g->e1);
  if (g->h_ring) (void)hipHostFree(g->h_ring);
  if (g->h_acc) (void)hipHostFree(g->h_acc);
  if (g->d_ring) (void)hipFree(g->d_ring);
  if (g->d_part) (void)hipFree(g->d_part);
  if (g->d_acc) (void)hipFree(g->d_acc);
  delete g;
}

static int grid_res(bt_state_s* s, GridRes** out) {
  if (!s->grid) {
    const size_t tabsz = s->prec == BT_C64 ? sizeof(GridTable<float>) : sizeof(GridTable<double>);
    const int cus = ctx().cus > 0 ? ctx().cus : 256;
    auto* g = new GridRes();
    bool ok = hipHostMalloc(&g->h_ring, tabsz * kRing) == hipSuccess &&
              hipMalloc(&g->d_ring, tabsz * kRing) == hipSuccess &&
              hipMalloc((void**)&g->d_part, sizeof(double) * kAcc * (size_t)cus * 8) == hipSuccess &&
              hipMalloc((void**)&g->d_acc, sizeof(double) * kAcc) == hipSuccess &&
              hipHostMalloc((void**)&g->h_acc, sizeof(double) * kAcc) == hipSuccess;
    for (int i = 0; i < kRing && ok; i++) ok = hipEventCreateWithFlags(&g->ev[i], hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventCreate(&g->e0) == hipSuccess && hipEventCreate(&g->e1) == hipSuccess;
    if (!ok) {
      (void)hipGetLastError();
      free_grid_res(g);
      return fail(BT_ERR_NOMEM, "run_grid: allocation failed");
    }
    s->grid = g;
  }
  *out = s->grid;
  return BT_OK;
}

// Host driver of the interpreter over grid passes on state s (SV or DM).
static int run_grid(bt_state_s* s, const btc::CircuitHost& ch, uint64_t seed, uint64_t traj, const double* uni,
                    int64_t nuni, int32_t* outs, int64_t outcap, int64_t* nout, bt_run_stats* stats, int fold,
                    int prec_tile_T) {
  const int nbits = s->nbits;
  std::vector<QOp> slots(bti::kQMax);
  std::unique_ptr<bti::Interp> ip(new bti::Interp());
  bti::Program prog{ch.ops.data(), ch.ntop, ch.pool.data(), ch.chan.data(), ch.chanw.data(), nbits};
  const int T = prec_tile_T;
  ip->init(prog, slots.data(), T, coalesce_bits(s->prec));
  ip->seed = seed;
  ip->traj = traj;
  if (uni) {
    ip->uni = uni;
    ip->nuni = nuni;
  }
  ip->outs = outs;
  ip->outcap = outcap;
  ip->fold = fold;
  ip->tol_exact = s->prec == BT_C64 ? 1e-6 : 1e-12;
  const size_t tabsz = s->prec == BT_C64 ? sizeof(GridTable<float>) : sizeof(GridTable<double>);
  const uint64_t ntiles = 1ull << (nbits - std::min(T, nbits));
  const unsigned blocks = grid_pass_blocks(ntiles);
  const size_t lds = (size_t)(1u << std::min(T, nbits)) * s->elem();
  GridRes* gr = nullptr;
  int rc = grid_res(s, &gr);
  if (rc) return rc;
  void* const h_ring = gr->h_ring;
  void* const d_ring = gr->d_ring;
  double* const d_part = gr->d_part;
  double* const d_acc = gr->d_acc;
  double* const h_acc = gr->h_acc;
  hipEvent_t* const ev = gr->ev;
  hipEvent_t const e0 = gr->e0, e1 = gr->e1;
  grid_pass_lds_attr(s->prec, std::min(T, nbits));
  (void)hipEventRecord(e0, s->stream);
  Plan plan;
  PassResult res;
  memset(&res, 0, sizeof res);
  int64_t npass = 0;
  int red_swap = 0;
  const int64_t max_passes = 64 + 8 * (int64_t)ch.ops.size();
  while (rc == BT_OK && ip->step(&plan, &res)) {
    if (++npass > max_passes) {
      rc = fail(BT_ERR_STATE, "run_grid: pass limit exceeded");
      break;
    }
    // the slot's previous table copy (this run's or an earlier one's) is done
    const int slot = (int)(npass % kRing);
    if (hipEventSynchronize(ev[slot]) != hipSuccess) {
      rc = fail(BT_ERR_HIP, "run_grid: event sync");
      break;
    }
    char* hslot = (char*)h_ring + tabsz * slot;
    char* dslot = (char*)d_ring + tabsz * slot;
    int grc;
    if (s->prec == BT_C64) {
      auto* t = (GridTable<float>*)hslot;
      grc = build_geom<float>(plan, slots.data(), nbits, &t->g, t->mats);
      red_swap = t->g.red_swap;
    } else {
      auto* t = (GridTable<double>*)hslot;
      grc = build_geom<double>(plan, slots.data(), nbits, &t->g, t->mats);
      red_swap = t->g.red_swap;
    }
    if (grc) {
      rc = fail(grc, "run_grid: pass geometry");
      break;
    }
    if (hipMemcpyAsync(dslot, hslot, tabsz, hipMemcpyHostToDevice, s->stream) != hipSuccess) {
      rc = fail(BT_ERR_HIP, "run_grid: table copy");
      break;
    }
    (void)hipEventRecord(ev[slot], s->stream);
    // 32-bit in-tile offsets whenever every tile bit is below 32 (the tile
    // base, which holds the outside bits, stays 64-bit)
    const bool narrow = plan.ntile == 0 || plan.tile[plan.ntile - 1] < 32;
    if (s->prec == BT_C64) {
      if (narrow)
        hipLaunchKernelGGL((k_grid_pass<float, uint32_t>), dim3(blocks), dim3(TileCfg<float>::NT), lds, s->stream,
                           (C<float>*)s->d, (const GridTable<float>*)dslot, ntiles, d_part);
      else
        hipLaunchKernelGGL((k_grid_pass<float, uint64_t>), dim3(blocks), dim3(TileCfg<float>::NT), lds, s->stream,
                           (C<float>*)s->d, (const GridTable<float>*)dslot, ntiles, d_part);
    } else {
      if (narrow)
        hipLaunchKernelGGL((k_grid_pass<double, uint32_t>), dim3(blocks), dim3(TileCfg<double>::NT), lds,
                           s->stream, (C<double>*)s->d, (const GridTable<double>*)dslot, ntiles, d_part);
      else
        hipLaunchKernelGGL((k_grid_pass<double, uint64_t>), dim3(blocks), dim3(TileCfg<double>::NT), lds,
                           s->stream, (C<double>*)s->d, (const GridTable<double>*)dslot, ntiles, d_part);
    }
    if (hipGetLastError() != hipSuccess) {
      rc = fail(BT_ERR_HIP, "run_grid: pass launch");
      break;
    }
    // the host waits for a pass's sums only when a decision needs them or
    // the norm is folded into the next pass (fold = 0, a density matrix:
    // passes follow one another without a round trip)
    if (plan.red != bti::RED_NONE || (plan.aligned && fold)) {
      hipLaunchKernelGGL(k_finalize_acc, dim3(kAcc), dim3(256), 0, s->stream, (const double*)d_part, (int)blocks,
                         d_acc);
      if (hipMemcpyAsync(h_acc, d_acc, sizeof(double) * kAcc, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
          hipStreamSynchronize(s->stream) != hipSuccess) {
        rc = fail(BT_ERR_HIP, "run_grid: reduction copy");
        break;
      }
      to_result(plan, red_swap, h_acc, &res);
    } else {
      res.norm2 = 0.0;  // not used for an intermediate pass
    }
  }
  if (rc == BT_OK && ip->err) rc = fail(ip->err, "interpreter error (" + std::to_string(ip->err) + ")");
  if (rc == BT_OK) {
    (void)hipEventRecord(e1, s->stream);
    if (hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(BT_ERR_HIP, "run_grid: final sync");
  }
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->passes = ip->passes;
    stats->reduce_passes = ip->rpasses;
    stats->bytes = ip->bytes_units * (double)(1ull << nbits) * (double)s->elem();
    stats->boundary_draws = ip->boundary;
    stats->draws = ip->draw;
    stats->launches = ip->passes + ip->rpasses;
    float ms = 0;
    if (rc == BT_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) stats->kernel_ms = ms;
  }
  if (nout) *nout = ip->nout;
  return rc;
}

// Pass tables of a circuit whose pass plan does not depend on reductions
// (a density matrix: no sampling, no renormalisation, fold = 0), planned on
// the host once: tables gets npass GridTable entries of tabsz bytes each,
// narrow[i] whether pass i uses 32-bit in-tile offsets.
int plan_static_passes(bt_state_s* s, const btc::CircuitHost& ch, int T, std::vector<unsigned char>* tables,
                       std::vector<uint8_t>* narrow, size_t* tabsz) {
  const int nbits = s->nbits;
  std::vector<QOp> slots(bti::kQMax);
  std::unique_ptr<bti::Interp> ip(new bti::Interp());
  bti::Program prog{ch.ops.data(), ch.ntop, ch.pool.data(), ch.chan.data(), ch.chanw.data(), nbits};
  ip->init(prog, slots.data(), T, coalesce_bits(s->prec));
  ip->fold = 0;
  *tabsz = s->prec == BT_C64 ? sizeof(GridTable<float>) : sizeof(GridTable<double>);
  tables->clear();
  narrow->clear();
  Plan plan;
  PassResult res;
  memset(&res, 0, sizeof res);
  const int64_t max_passes = 64 + 8 * (int64_t)ch.ops.size();
  int64_t npass = 0;
  while (ip->step(&plan, &res)) {
    if (plan.red != bti::RED_NONE) return fail(BT_ERR_UNSUPPORTED, "plan_static_passes: a pass needs a reduction");
    if (++npass > max_passes) return fail(BT_ERR_STATE, "plan_static_passes: pass limit exceeded");
    tables->resize(tables->size() + *tabsz);
    unsigned char* t = tables->data() + tables->size() - *tabsz;
    memset(t, 0, *tabsz);
    const int grc = s->prec == BT_C64
                        ? build_geom<float>(plan, slots.data(), nbits, &((GridTable<float>*)t)->g, ((GridTable<float>*)t)->mats)
                        : build_geom<double>(plan, slots.data(), nbits, &((GridTable<double>*)t)->g,
                                             ((GridTable<double>*)t)->mats);
    if (grc) return fail(grc, "plan_static_passes: pass geometry");
    narrow->push_back(plan.ntile == 0 || plan.tile[plan.ntile - 1] < 32 ? 1 : 0);
    memset(&res, 0, sizeof res);
  }
  if (ip->err) return fail(ip->err, "plan_static_passes: interpreter error (" + std::to_string(ip->err) + ")");
  return BT_OK;
}

// Enqueue one grid pass of state s (tile bits T) with its table at dtab;
// d_part: kAcc x grid doubles of reduction partials (written, unused here).
int launch_grid_pass(bt_state_s* s, const void* dtab, bool narrow, int T, double* d_part) {
  const int nbits = s->nbits;
  const uint64_t ntiles = 1ull << (nbits - std::min(T, nbits));
  const unsigned blocks = grid_pass_blocks(ntiles);
  const size_t lds = (size_t)(1u << std::min(T, nbits)) * s->elem();
  if (s->prec == BT_C64) {
    if (narrow)
      hipLaunchKernelGGL((k_grid_pass<float, uint32_t>), dim3(blocks), dim3(TileCfg<float>::NT), lds, s->stream,
                         (C<float>*)s->d, (const GridTable<float>*)dtab, ntiles, d_part);
    else
      hipLaunchKernelGGL((k_grid_pass<float, uint64_t>), dim3(blocks), dim3(TileCfg<float>::NT), lds, s->stream,
                         (C<float>*)s->d, (const GridTable<float>*)dtab, ntiles, d_part);
  } else {
    if (narrow)
      hipLaunchKernelGGL((k_grid_pass<double, uint32_t>), dim3(blocks), dim3(TileCfg<double>::NT), lds, s->stream,
                         (C<double>*)s->d, (const GridTable<double>*)dtab, ntiles, d_part);
    else
      hipLaunchKernelGGL((k_grid_pass<double, uint64_t>), dim3(blocks), dim3(TileCfg<double>::NT), lds, s->stream,
                         (C<double>*)s->d, (const GridTable<double>*)dtab, ntiles, d_part);
  }
  BT_HIP(hipGetLastError());
  return BT_OK;
}

// The pass tables of circuit c (dm: its density-matrix form) at tile bits T
// on states like s: planned once per tiling and kept in the circuit's slot,
// copied out for this run.
static int static_plan(bt_state_s* s, bt_circuit_s* c, bool dm, int T, std::vector<unsigned char>* tables,
                       std::vector<uint8_t>* narrow, size_t* tabsz) {
  const int cb = coalesce_bits(s->prec);
  std::lock_guard<std::mutex> lk(c->mu);
  bt_circuit_s::StaticPlan* P = dm ? &c->dm_plan : &c->sv_plan;
  if (P->T != T || P->prec != s->prec || P->cb != cb) {
    P->T = -1;
    int rc;
    if (dm) {
      btc::CircuitHost lowered;
      std::string msg;
      rc = lower_dm(c->h, &lowered, &msg);
      if (rc) return fail(rc, msg);
      rc = plan_static_passes(s, lowered, T, &P->tables, &P->narrow, &P->tabsz);
    } else {
      rc = plan_static_passes(s, c->h, T, &P->tables, &P->narrow, &P->tabsz);
    }
    if (rc) return rc;
    P->T = T;
    P->prec = s->prec;
    P->cb = cb;
  }
  *tables = P->tables;
  *narrow = P->narrow;
  *tabsz = P->tabsz;
  return BT_OK;
}

// Passes whose plan depends on nothing they compute (a density matrix, a
// gate-only state-vector circuit) run eagerly from tables planned on the
// host beforehand: uploaded in one copy and launched back to back with no
// host round trip between them.
static int run_static(bt_state_s* s, const std::vector<unsigned char>& tables, const std::vector<uint8_t>& narrow,
                      size_t tabsz, int T, bt_run_stats* stats) {
  GridRes* gr = nullptr;
  int rc = grid_res(s, &gr);
  if (rc) return rc;
  void* d_tab = nullptr;
  if (!tables.empty() && hipMalloc(&d_tab, tables.size()) != hipSuccess) {
    (void)hipGetLastError();
    return fail(BT_ERR_NOMEM, "run: pass tables");
  }
  if (d_tab && hipMemcpyAsync(d_tab, tables.data(), tables.size(), hipMemcpyHostToDevice, s->stream) != hipSuccess)
    rc = fail(BT_ERR_HIP, "run: table upload");
  grid_pass_lds_attr(s->prec, std::min(T, s->nbits));
  if (rc == BT_OK) (void)hipEventRecord(gr->e0, s->stream);
  for (size_t i = 0; rc == BT_OK && i < narrow.size(); i++)
    rc = launch_grid_pass(s, (const unsigned char*)d_tab + i * tabsz, narrow[i] != 0, T, gr->d_part);
  if (rc == BT_OK) {
    (void)hipEventRecord(gr->e1, s->stream);
    if (hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(BT_ERR_HIP, "run: sync");
  } else {
    (void)hipStreamSynchronize(s->stream);  // the upload is done before the tables are freed
  }
  if (d_tab) (void)hipFree(d_tab);
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->passes = (int64_t)narrow.size();
    stats->bytes = 2.0 * (double)narrow.size() * (double)(1ull << s->nbits) * (double)s->elem();
    stats->launches = (int64_t)narrow.size();
    float ms = 0;
    if (rc == BT_OK && hipEventElapsedTime(&ms, gr->e0, gr->e1) == hipSuccess) stats->kernel_ms = ms;
  }
  return rc;
}

// Allow the grid pass's dynamic LDS tile of 2^T elements.
void grid_pass_lds_attr(int prec, int T) {
  const int lds = (int)((size_t)(1u << T) * (prec == BT_C64 ? 8 : 16));
  if (prec == BT_C64) {
    (void)hipFuncSetAttribute((const void*)k_grid_pass<float, uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void*)k_grid_pass<float, uint64_t>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  } else {
    (void)hipFuncSetAttribute((const void*)k_grid_pass<double, uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    (void)hipFuncSetAttribute((const void*)k_grid_pass<double, uint64_t>, hipFuncAttributeMaxDynamicSharedMemorySize, lds);
  }
}

// Tile bits of the density-matrix executor: a superoperator spans up to
// kMaxQK bits of vec(rho), so a smaller tile setting is raised to hold it.
// Small complex64 density matrices (vec(rho) of at most 2^20 entries,
// N <= 10) run on tiles of at most 2^12 entries: half the LDS per
// workgroup, so twice the workgroups per CU and per state, for the 16 x 16
// superoperators applied in LDS (measured on noisy brickworks, 16 layers:
// N = 8 2.99 vs 3.37 ms, N = 10 3.94 vs 4.57 ms per circuit; from N = 11 up
// 2^13-entry tiles are faster).  A smaller tile setting is kept.
int dm_tile_bits_for(int prec, int nbits) {
  int T = tile_bits_for(prec, nbits);
  if (prec == BT_C64 && nbits <= 20) T = std::min(T, 12);
  return std::max(T, std::min(nbits, bti::kMaxQK));
}

int dm_tile_bits(bt_state_s* rho) { return dm_tile_bits_for(rho->prec, rho->nbits); }

}  // namespace bt

using namespace bt;

extern "C" {

int bt_circuit_create(int nqubits, const bt_ir_op* ops, int nops, int ntop, const double* pool, int64_t npool,
                      bt_circuit* out) {
  if (!out) return fail(BT_ERR_ARG, "null output");
  *out = nullptr;
  static std::atomic<uint64_t> next_uid{1};
  auto* c = new bt_circuit_s();
  c->uid = next_uid.fetch_add(1);
  std::string msg;
  int rc = btc::analyze(nqubits, ops, nops, ntop, pool, npool, &c->h, &msg);
  if (rc) {
    delete c;
    return fail(rc, msg);
  }
  *out = c;
  return BT_OK;
}

int bt_circuit_destroy(bt_circuit c) {
  if (!c) return BT_OK;
  if (c->d_ops) (void)hipFree(c->d_ops);
  if (c->d_pool) (void)hipFree(c->d_pool);
  if (c->d_chan) (void)hipFree(c->d_chan);
  if (c->d_chanw) (void)hipFree(c->d_chanw);
  delete c;
  return BT_OK;
}

int64_t bt_circuit_max_draws(bt_circuit c) { return c ? c->h.max_draws : -1; }

int bt_run_circuit(bt_state s, bt_circuit c, uint64_t seed, uint64_t traj, const double* uniforms, int64_t nuni,
                   int* outcomes, int64_t outcap, int64_t* nout, bt_run_stats* stats) {
  if (!s || !c) return fail(BT_ERR_ARG, "null state or circuit");
  int rc = ensure_device();
  if (rc) return rc;
  if (s->kind != BT_SV) return fail(BT_ERR_ARG, "bt_run_circuit needs a state vector (use bt_run_circuit_dm)");
  if (s->n != c->h.n) return fail(BT_ERR_ARG, "circuit and state qubit counts differ");
  if (uniforms && nuni < c->h.max_draws) return fail(BT_ERR_DRAWS, "too few uniforms for the circuit");
  bool gates_only = true;
  for (int i = 0; i < c->h.ntop && gates_only; i++) gates_only = c->h.ops[i].kind == BT_OP_GATE;
  if (!gates_only)
    return run_grid(s, c->h, seed, traj, uniforms, nuni, outcomes, outcap, nout, stats, 1,
                    tile_bits_for(s->prec, s->nbits));
  // gates only: no decisions and no renormalisation, so the passes are
  // planned once per tiling and launched back to back
  const int T = tile_bits_for(s->prec, s->nbits);
  std::vector<unsigned char> tables;
  std::vector<uint8_t> narrow;
  size_t tabsz = 0;
  rc = static_plan(s, c, false, T, &tables, &narrow, &tabsz);
  if (rc) return rc;
  if (nout) *nout = 0;
  return run_static(s, tables, narrow, tabsz, T, stats);
}

int bt_run_circuit_dm(bt_state rho, bt_circuit c, bt_run_stats* stats) {
  if (!rho || !c) return fail(BT_ERR_ARG, "null state or circuit");
  int rc = ensure_device();
  if (rc) return rc;
  if (rho->kind != BT_DM) return fail(BT_ERR_ARG, "bt_run_circuit_dm needs a density matrix");
  if (rho->n != c->h.n) return fail(BT_ERR_ARG, "circuit and state qubit counts differ");
  const int T = dm_tile_bits(rho);
  std::vector<unsigned char> tables;
  std::vector<uint8_t> narrow;
  size_t tabsz = 0;
  rc = static_plan(rho, c, true, T, &tables, &narrow, &tabsz);
  if (rc) return rc;
  return run_static(rho, tables, narrow, tabsz, T, stats);
}

int bt_run_batch_persistent(bt_circuit c, int prec, uint64_t seed, int64_t traj0, int64_t ntraj, int* outcomes, int64_t outcap,
                 int64_t* nouts, double* final_states, bt_run_stats* stats) {
  if (!c) return fail(BT_ERR_ARG, "null circuit");
  if (ntraj <= 0 || !outcomes || !nouts || outcap < 0) return fail(BT_ERR_ARG, "bad batch arguments");
  if (prec != BT_C64 && prec != BT_C128) return fail(BT_ERR_ARG, "bad precision");
  int rc = ensure_device();
  if (rc) return rc;
  rc = upload_circuit(c);
  if (rc) return rc;
  const int nbits = c->h.n;
  const size_t elem = prec == BT_C64 ? 8 : 16;
  const uint64_t len = 1ull << nbits;
  const int T = tile_bits_for(prec, nbits);
  const int cus = ctx().cus > 0 ? ctx().cus : 256;
  // workgroups resident per CU (LDS-bound: tile + tables)
  int per_cu = 2;
  int64_t nwg = std::min<int64_t>(ntraj, (int64_t)cus * per_cu);
  // keep the per-workgroup states within a quarter of free HBM
  size_t freeb = 0, totalb = 0;
  (void)release_cached_memory();  // the batch executor's cached slab would otherwise hold the HBM
  BT_HIP(hipMemGetInfo(&freeb, &totalb));
  const size_t per_wg = len * elem + sizeof(QOp) * bti::kQMax + sizeof(bti::Interp);
  nwg = std::max<int64_t>(1, std::min<int64_t>(nwg, (int64_t)(freeb / 4 / per_wg)));
  void* d_states = nullptr;
  QOp* d_slots = nullptr;
  bti::Interp* d_interps = nullptr;
  unsigned long long* d_counter = nullptr;
  unsigned long long* d_stats = nullptr;
  int32_t* d_err = nullptr;
  int32_t* d_outs = nullptr;
  int64_t* d_nouts = nullptr;
  double* d_final = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    if (d_states) (void)hipFree(d_states);
    if (d_slots) (void)hipFree(d_slots);
    if (d_interps) (void)hipFree(d_interps);
    if (d_counter) (void)hipFree(d_counter);
    if (d_stats) (void)hipFree(d_stats);
    if (d_err) (void)hipFree(d_err);
    if (d_outs) (void)hipFree(d_outs);
    if (d_nouts) (void)hipFree(d_nouts);
    if (d_final) (void)hipFree(d_final);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  };
  const size_t outs_bytes = sizeof(int32_t) * (size_t)ntraj * (size_t)std::max<int64_t>(outcap, 1);
  if (hipMalloc(&d_states, len * elem * nwg) != hipSuccess ||
      hipMalloc((void**)&d_slots, sizeof(QOp) * bti::kQMax * nwg) != hipSuccess ||
      hipMalloc((void**)&d_interps, sizeof(bti::Interp) * nwg) != hipSuccess ||
      hipMalloc((void**)&d_counter, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc((void**)&d_stats, sizeof(unsigned long long) * 8) != hipSuccess ||
      hipMalloc((void**)&d_err, sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void**)&d_outs, outs_bytes) != hipSuccess ||
      hipMalloc((void**)&d_nouts, sizeof(int64_t) * ntraj) != hipSuccess ||
      (final_states && hipMalloc((void**)&d_final, sizeof(double) * 2 * len * ntraj) != hipSuccess)) {
    cleanup();
    (void)hipGetLastError();
    return fail(BT_ERR_NOMEM, "bt_run_batch_persistent: device allocation failed");
  }
  if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    cleanup();
    return fail(BT_ERR_HIP, "bt_run_batch_persistent: stream/event");
  }
  (void)hipMemsetAsync(d_counter, 0, sizeof(unsigned long long), st);
  (void)hipMemsetAsync(d_stats, 0, sizeof(unsigned long long) * 8, st);
  (void)hipMemsetAsync(d_err, 0, sizeof(int32_t), st);
  BatchArgs a;
  memset(&a, 0, sizeof a);
  a.states = d_states;
  a.slots = d_slots;
  a.interps = d_interps;
  a.prog = bti::Program{c->d_ops, c->h.ntop, c->d_pool, c->d_chan, c->d_chanw, nbits};
  a.nbits = nbits;
  a.T = T;
  a.cmin = coalesce_bits(prec);
  a.want_final = final_states ? 1 : 0;
  a.seed = seed;
  a.traj0 = traj0;
  a.ntraj = ntraj;
  a.counter = d_counter;
  a.outcomes = d_outs;
  a.outcap = outcap;
  a.nouts = d_nouts;
  a.final_states = d_final;
  a.tol_exact = prec == BT_C64 ? 1e-6 : 1e-12;
  a.stats = d_stats;
  a.err = d_err;
  a.max_passes = 64 + 8 * (int64_t)c->h.ops.size();
  const size_t lds = (size_t)(1u << T) * elem;
  (void)hipEventRecord(e0, st);
  if (prec == BT_C64) {
    (void)hipFuncSetAttribute((const void*)k_batch<float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_batch<float>), dim3((unsigned)nwg), dim3(TileCfg<float>::NT), lds, st, a);
  } else {
    (void)hipFuncSetAttribute((const void*)k_batch<double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_batch<double>), dim3((unsigned)nwg), dim3(TileCfg<double>::NT), lds, st, a);
  }
  hipError_t le = hipGetLastError();
  (void)hipEventRecord(e1, st);
  hipError_t se = hipStreamSynchronize(st);
  if (le != hipSuccess || se != hipSuccess) {
    std::string m = std::string("bt_run_batch_persistent: kernel failed: ") + hipGetErrorString(le != hipSuccess ? le : se);
    cleanup();
    return fail(BT_ERR_HIP, m);
  }
  int32_t herr = 0;
  unsigned long long hst[8];
  (void)hipMemcpy(&herr, d_err, sizeof herr, hipMemcpyDeviceToHost);
  (void)hipMemcpy(hst, d_stats, sizeof hst, hipMemcpyDeviceToHost);
  (void)hipMemcpy(outcomes, d_outs, outs_bytes, hipMemcpyDeviceToHost);
  (void)hipMemcpy(nouts, d_nouts, sizeof(int64_t) * ntraj, hipMemcpyDeviceToHost);
  if (final_states) (void)hipMemcpy(final_states, d_final, sizeof(double) * 2 * len * ntraj, hipMemcpyDeviceToHost);
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->passes = (int64_t)hst[0];
    stats->reduce_passes = (int64_t)hst[1];
    stats->boundary_draws = (int64_t)hst[2];
    stats->draws = (int64_t)hst[3];
    stats->launches = 1;
    // bytes: read+write per pass, read per reduction pass, one write to zero
    // each trajectory's state
    stats->bytes = ((double)hst[0] * 2.0 + (double)hst[1] + (double)ntraj) * (double)len * (double)elem;
    float ms = 0;
    if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) stats->kernel_ms = ms;
  }
  cleanup();
  if (herr) return fail(herr, "bt_run_batch_persistent: a trajectory failed (interpreter error " + std::to_string(herr) + ")");
  return BT_OK;
}

// Tile width override used by the autotuner (0 restores the default).
int bt_set_tile_bits(int prec, int T) {
  // (must not exceed log2(threads) + register bits of the tile pass)
  if (prec != BT_C64 && prec != BT_C128) return fail(BT_ERR_ARG, "bad precision");
  if (T != 0 && (T < 2 || T > default_tile_bits(prec))) return fail(BT_ERR_ARG, "tile bits out of range");
  g_tile_override[prec] = T;
  return BT_OK;
}

int bt_set_tuning(int prec, int tile_bits, int coalesce, int tiles_per_blk) {
  if (prec != BT_C64 && prec != BT_C128) return fail(BT_ERR_ARG, "bad precision");
  if (tile_bits != 0 && (tile_bits < 6 || tile_bits > default_tile_bits(prec)))
    return fail(BT_ERR_ARG, "tile bits out of range");
  if (coalesce < 0 || coalesce > 6) return fail(BT_ERR_ARG, "coalescing bits out of range");
  if (tiles_per_blk < 0 || tiles_per_blk > 64) return fail(BT_ERR_ARG, "tiles per workgroup out of range");
  g_tile_override[prec] = tile_bits;
  g_coalesce_override[prec] = coalesce;
  g_tpb_override = tiles_per_blk;
  return BT_OK;
}

int bt_get_tuning(int prec, int* tile_bits, int* coalesce, int* tiles_per_blk) {
  if (prec != BT_C64 && prec != BT_C128) return fail(BT_ERR_ARG, "bad precision");
  if (tile_bits) *tile_bits = g_tile_override[prec] ? g_tile_override[prec] : default_tile_bits(prec);
  if (coalesce) *coalesce = coalesce_bits(prec);
  if (tiles_per_blk) *tiles_per_blk = tiles_per_block();
  return BT_OK;
}

// Autotuner of the batched executor, in two stages: tile bits and
// coalescing reserve timed on `sample` trajectories of the circuit (device
// time of the round passes) at 8 tiles per workgroup, then the tiles per
// workgroup of the winner on 4 x `sample` trajectories, the largest within 1%
// of the fastest (a run's workgroup count sets what the per-workgroup costs
// weigh: pass tables staged into LDS, one partial sum each; a sample smaller
// than the run underweights them).  Stage 2 times each distinct effective
// value once (item_tpb: candidates the cap turns into the same launch are
// one candidate) and installs and reports the effective value.  The first
// failing run ends the search with its error and the previous tuning
// restored.
int bt_autotune_batch(bt_circuit c, int prec, int64_t sample, uint64_t seed, bt_tuning* best) {
  if (!c || !best) return fail(BT_ERR_ARG, "null circuit or result");
  if (prec != BT_C64 && prec != BT_C128) return fail(BT_ERR_ARG, "bad precision");
  if (sample < 1) return fail(BT_ERR_ARG, "sample must be positive");
  const int p = prec == BT_C64 ? 0 : 1;
  const int saved[3] = {g_tile_override[p], g_coalesce_override[p], g_tpb_override};
  const int Tmax = default_tile_bits(prec);
  const int nb = c->h.n;
  std::vector<int> Ts;
  for (int T = Tmax; T >= Tmax - 1; T--)
    if (T <= nb) Ts.push_back(T);
  if (Ts.empty()) Ts.push_back(std::min(Tmax, nb));
  const int outcap = (int)std::max<size_t>(16, 2 * c->h.ops.size());
  const int64_t sample2 = 4 * sample;
  std::vector<int> outs((size_t)sample2 * outcap);
  std::vector<int64_t> nouts(sample2);
  double best_ms = -1;
  bt_tuning b{};
  int rc = BT_OK;
  // rc keeps the first error: no candidate runs after it
  auto time_one = [&](int T, int cb, int tpb, int64_t n) -> double {
    if (rc) return -1.0;
    g_tile_override[p] = T;
    g_coalesce_override[p] = cb;
    g_tpb_override = tpb;
    bt_run_stats st;
    rc = bt_run_batch(c, prec, seed, 0, n, outs.data(), outcap, nouts.data(), nullptr, &st);
    return rc ? -1.0 : st.kernel_ms / (double)n;
  };
  for (int T : Ts)
    for (int cb : {3, 4, 5}) {
      if (rc) break;
      if (cb >= T) continue;
      const double ms = time_one(T, cb, 8, sample);
      if (rc) break;
      if (best_ms < 0 || ms < best_ms) {
        best_ms = ms;
        b = bt_tuning{T, cb, 8, ms};
      }
    }
  if (rc == BT_OK && best_ms >= 0) {
    const uint64_t ntiles = 1ull << (nb - std::min(nb, b.tile_bits));
    int cand[4], nc = 0;
    for (int want : {4, 8, 16, 32}) {
      const int eff = item_tpb(want, ntiles, 32);
      bool dup = false;
      for (int i = 0; i < nc; i++) dup = dup || cand[i] == eff;
      if (!dup) cand[nc++] = eff;
    }
    double t2[4];
    double best2 = -1;
    for (int i = 0; i < nc && rc == BT_OK; i++) {
      t2[i] = time_one(b.tile_bits, b.coalesce_bits, cand[i], sample2);
      if (rc == BT_OK && (best2 < 0 || t2[i] < best2)) best2 = t2[i];
    }
    // the largest tiles per workgroup within 1% of the best: the sample
    // still has fewer workgroups than a full run, whose per-workgroup costs
    // it underweights
    for (int i = 0; i < nc && rc == BT_OK; i++)
      if (t2[i] <= best2 * 1.01) {
        b.tiles_per_block = cand[i];
        b.ms_per_traj = t2[i];
      }
  }
  if (rc || best_ms < 0) {
    g_tile_override[p] = saved[0];
    g_coalesce_override[p] = saved[1];
    g_tpb_override = saved[2];
    return rc ? rc : fail(BT_ERR_STATE, "autotune: no candidate ran");
  }
  g_tile_override[p] = b.tile_bits;
  g_coalesce_override[p] = b.coalesce_bits;
  g_tpb_override = b.tiles_per_block;
  *best = b;
  return BT_OK;
}

}  // extern "C"
