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

using bti::Plan;
using bti::PassResult;
using bti::QOp;

namespace bt {

constexpr int kMatEntries = bti::Interp::kPassMatEntries;
constexpr int kAcc = 33;  // norm2, hist[16], packed rho[16]

struct LOp {
  int32_t k;
  int32_t lsorted[4];  // ascending local target positions
  uint32_t loff[16];   // tile offset of each local matrix index
  uint32_t lctrl;      // control bits inside the tile
  uint64_t gctrl;      // control bits outside the tile (global positions)
  int32_t moff;        // matrix offset in the pass's matrix table
};

struct PassGeom {
  int32_t T, c0, nA, nB, ncomp, nops, red, red_k, write, need_norm;
  int32_t tile[bti::kMaxT];
  int32_t comp[40];
  int32_t red_bits[4];  // RED_HIST: global bits; RED_RHO: local positions, MSB first
  double scale;
  LOp ops[bti::kPMax];
};

// Host/device: turn a plan + queue slots into the tile-local operator table.
template <typename R>
__host__ __device__ __attribute__((noinline)) void build_geom(const Plan& plan, const QOp* slots, int nbits, PassGeom* g, C<R>* mats) {
  g->T = plan.ntile;
  for (int i = 0; i < plan.ntile; i++) g->tile[i] = plan.tile[i];
  int c0 = 0;
  while (c0 < plan.ntile && plan.tile[c0] == c0) c0++;
  g->c0 = c0;
  const int nhigh = plan.ntile - c0;
  g->nA = nhigh < 7 ? nhigh : 7;
  g->nB = nhigh - g->nA;
  g->ncomp = 0;
  uint64_t tmask = 0;
  for (int i = 0; i < plan.ntile; i++) tmask |= 1ull << plan.tile[i];
  for (int b = 0; b < nbits; b++)
    if (!((tmask >> b) & 1)) g->comp[g->ncomp++] = b;
  int rank[64];
  for (int b = 0; b < 64; b++) rank[b] = -1;
  for (int i = 0; i < plan.ntile; i++) rank[plan.tile[i]] = i;
  g->nops = plan.nops;
  g->write = plan.write;
  g->need_norm = (plan.aligned || plan.red != bti::RED_NONE) ? 1 : 0;
  g->scale = plan.scale;
  int moff = 0;
  for (int o = 0; o < plan.nops; o++) {
    const QOp& q = slots[plan.qidx[o]];
    LOp& L = g->ops[o];
    L.k = q.k;
    int lb[4];
    for (int i = 0; i < q.k; i++) lb[i] = rank[q.bits[i]];
    for (int i = 0; i < q.k; i++) L.lsorted[i] = lb[i];
    for (int i = 0; i < q.k; i++)
      for (int j = i + 1; j < q.k; j++)
        if (L.lsorted[j] < L.lsorted[i]) {
          int t = L.lsorted[i];
          L.lsorted[i] = L.lsorted[j];
          L.lsorted[j] = t;
        }
    const int D = 1 << q.k;
    for (int l = 0; l < D; l++) {
      uint32_t o2 = 0;
      for (int i = 0; i < q.k; i++)
        if ((l >> (q.k - 1 - i)) & 1) o2 |= 1u << lb[i];
      L.loff[l] = o2;
    }
    L.lctrl = 0;
    L.gctrl = 0;
    for (int b = 0; b < nbits; b++)
      if ((q.ctrl >> b) & 1) {
        if (rank[b] >= 0) L.lctrl |= 1u << rank[b];
        else L.gctrl |= 1ull << b;
      }
    L.moff = moff;
    for (int i = 0; i < D * D; i++) {
      mats[moff + i].x = (R)q.m[2 * i];
      mats[moff + i].y = (R)q.m[2 * i + 1];
    }
    moff += D * D;
  }
  g->red = plan.red;
  g->red_k = plan.red_k;
  for (int i = 0; i < plan.red_k; i++)
    g->red_bits[i] = plan.red == bti::RED_RHO ? rank[plan.red_bits[i]] : plan.red_bits[i];
}

__device__ __forceinline__ uint32_t ins0(uint32_t x, int b) { return ((x >> b) << (b + 1)) | (x & ((1u << b) - 1u)); }

template <typename R, int K, int NT>
__device__ __forceinline__ void apply_lop(C<R>* tile, int T, const LOp& op, const C<R>* mats) {
  constexpr int D = 1 << K;
  const uint32_t ng = 1u << (T - K);
  const C<R>* m = mats + op.moff;
  for (uint32_t g = threadIdx.x; g < ng; g += NT) {
    uint32_t lb = g;
#pragma unroll
    for (int i = 0; i < K; i++) lb = ins0(lb, op.lsorted[i]);
    if ((lb & op.lctrl) != op.lctrl) continue;
    C<R> v[D];
#pragma unroll
    for (int l = 0; l < D; l++) v[l] = tile[lb + op.loff[l]];
    if (K <= 2) {
#pragma unroll
      for (int r = 0; r < D; r++) {
        C<R> acc = cmake<C<R>>(0, 0);
#pragma unroll
        for (int c = 0; c < D; c++) acc = cfma(acc, m[r + c * D], v[c]);
        tile[lb + op.loff[r]] = acc;
      }
    } else {
      // volatile reads keep the D*D matrix entries (LDS broadcasts) inside
      // the group loop instead of hoisting them all into registers; every v
      // is read before any write
      const volatile R* mv = (const volatile R*)m;
      C<R> w[D];
#pragma unroll
      for (int r = 0; r < D; r++) {
        C<R> acc = cmake<C<R>>(0, 0);
#pragma unroll
        for (int c = 0; c < D; c++) acc = cfma(acc, cmake<C<R>>(mv[2 * (r + c * D)], mv[2 * (r + c * D) + 1]), v[c]);
        w[r] = acc;
      }
#pragma unroll
      for (int r = 0; r < D; r++) tile[lb + op.loff[r]] = w[r];
    }
  }
}

// Block-wide sum of M per-thread values, added to sums[0..M) (LDS) in a
// fixed order (waves, then lanes by xor butterfly): bit-reproducible.
template <int M, int NT>
__device__ __forceinline__ void block_add(const double* v, double* red /*LDS [NT/64][M]*/, double* sums) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll 1
  for (int j = 0; j < M; j++) {
    double x = v[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lane == 0) red[wave * M + j] = x;
  }
  __syncthreads();
  if (threadIdx.x < M) {
    double t = 0;
    for (int w = 0; w < NT / 64; w++) t += red[w * M + threadIdx.x];
    sums[threadIdx.x] += t;
  }
  __syncthreads();
}

// Packed reduced rho over RK local bits (MSB first) of the tile, added to
// racc: diagonal, then Re/Im of rho[a,b] for a < b.
template <typename R, int RK, int NT>
__device__ __forceinline__ void rho_accumulate(const C<R>* tile, int T, const int32_t* rbits, double* racc) {
  constexpr int D = 1 << RK;
  int srt[RK];
#pragma unroll
  for (int i = 0; i < RK; i++) srt[i] = rbits[i];
  if (RK > 1 && srt[RK - 1] < srt[0]) {
    int x = srt[0];
    srt[0] = srt[RK - 1];
    srt[RK - 1] = x;
  }
  uint32_t roff[D];
#pragma unroll
  for (int a = 0; a < D; a++) {
    uint32_t o2 = 0;
#pragma unroll
    for (int i = 0; i < RK; i++)
      if ((a >> (RK - 1 - i)) & 1) o2 |= 1u << rbits[i];
    roff[a] = o2;
  }
  const uint32_t ng = 1u << (T - RK);
  for (uint32_t gi = threadIdx.x; gi < ng; gi += NT) {
    uint32_t lb = gi;
#pragma unroll
    for (int i = 0; i < RK; i++) lb = ins0(lb, srt[i]);
    double vr[D], vi[D];
#pragma unroll
    for (int a = 0; a < D; a++) {
      C<R> v = tile[lb + roff[a]];
      vr[a] = (double)v.x;
      vi[a] = (double)v.y;
    }
    int p = D;
#pragma unroll
    for (int a = 0; a < D; a++) {
      racc[a] += vr[a] * vr[a] + vi[a] * vi[a];
#pragma unroll
      for (int b = a + 1; b < D; b++) {
        racc[p] += vr[a] * vr[b] + vi[a] * vi[b];
        racc[p + 1] += vi[a] * vr[b] - vr[a] * vi[b];
        p += 2;
      }
    }
  }
}

// One pass over tiles t = t0, t0 + tstep, ... < ntiles of `state`.
// Reductions accumulate per tile and are added to sums[0..kAcc) (LDS, zeroed
// by the caller): [0] norm^2, [1..16] histogram, [17..32] packed rho.
template <typename R, int NT>
__device__ void tile_pass(C<R>* __restrict__ state, const PassGeom& g, const C<R>* mats, C<R>* tile,
                          const uint64_t* hA, const uint64_t* hB, uint64_t t0, uint64_t tstep, uint64_t ntiles,
                          double* red, double* sums) {
  const int T = g.T;
  const uint32_t TL = 1u << T;
  const uint32_t lowmask = (1u << g.c0) - 1u;
  const uint32_t maskA = (1u << g.nA) - 1u;
  const int shB = g.c0 + g.nA;
  const R scale = (R)g.scale;
  for (uint64_t t = t0; t < ntiles; t += tstep) {
    uint64_t base = 0, tt = t;
    for (int j = 0; j < g.ncomp; j++) {
      base |= (tt & 1ull) << g.comp[j];
      tt >>= 1;
    }
    for (uint32_t l = threadIdx.x; l < TL; l += NT) {
      const uint64_t off = (uint64_t)(l & lowmask) | hA[(l >> g.c0) & maskA] | hB[l >> shB];
      C<R> v = state[base + off];
      v.x *= scale;
      v.y *= scale;
      tile[l] = v;
    }
    __syncthreads();
    for (int o = 0; o < g.nops; o++) {
      const LOp& op = g.ops[o];
      if ((base & op.gctrl) != op.gctrl) continue;  // uniform over the block
      switch (op.k) {
        case 1: apply_lop<R, 1, NT>(tile, T, op, mats); break;
        case 2: apply_lop<R, 2, NT>(tile, T, op, mats); break;
        case 3: apply_lop<R, 3, NT>(tile, T, op, mats); break;
        default: apply_lop<R, 4, NT>(tile, T, op, mats); break;
      }
      __syncthreads();
    }
    if (g.red == bti::RED_RHO) {
      double racc[16];
#pragma unroll
      for (int j = 0; j < 16; j++) racc[j] = 0.0;
      if (g.red_k == 1) rho_accumulate<R, 1, NT>(tile, T, g.red_bits, racc);
      else rho_accumulate<R, 2, NT>(tile, T, g.red_bits, racc);
      block_add<16, NT>(racc, red, sums + 17);
    }
    if (g.red == bti::RED_HIST) {
      double hacc[17];
#pragma unroll
      for (int j = 0; j < 17; j++) hacc[j] = 0.0;
      for (uint32_t l = threadIdx.x; l < TL; l += NT) {
        const uint64_t off = (uint64_t)(l & lowmask) | hA[(l >> g.c0) & maskA] | hB[l >> shB];
        const C<R> v = tile[l];
        if (g.write) state[base + off] = v;
        const double w = (double)v.x * (double)v.x + (double)v.y * (double)v.y;
        hacc[0] += w;
        const uint64_t gi = base | off;
        int bin = 0;
        for (int j = 0; j < g.red_k; j++) bin |= (int)((gi >> g.red_bits[j]) & 1ull) << j;
#pragma unroll
        for (int b = 0; b < 16; b++) hacc[1 + b] += (bin == b) ? w : 0.0;
      }
      block_add<17, NT>(hacc, red, sums);
    } else if (g.need_norm) {
      double nacc[1] = {0.0};
      for (uint32_t l = threadIdx.x; l < TL; l += NT) {
        const uint64_t off = (uint64_t)(l & lowmask) | hA[(l >> g.c0) & maskA] | hB[l >> shB];
        const C<R> v = tile[l];
        if (g.write) state[base + off] = v;
        nacc[0] += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
      }
      block_add<1, NT>(nacc, red, sums);
    } else {
      for (uint32_t l = threadIdx.x; l < TL; l += NT) {
        const uint64_t off = (uint64_t)(l & lowmask) | hA[(l >> g.c0) & maskA] | hB[l >> shB];
        if (g.write) state[base + off] = tile[l];
      }
    }
    __syncthreads();
  }
}

// Fill the offset tables of the tile's high bits (all threads).
template <int NT>
__device__ __forceinline__ void build_tables(const PassGeom& g, uint64_t* hA, uint64_t* hB) {
  for (int h = threadIdx.x; h < (1 << g.nA); h += NT) {
    uint64_t o = 0;
    for (int j = 0; j < g.nA; j++)
      if ((h >> j) & 1) o |= 1ull << g.tile[g.c0 + j];
    hA[h] = o;
  }
  for (int h = threadIdx.x; h < (1 << g.nB); h += NT) {
    uint64_t o = 0;
    for (int j = 0; j < g.nB; j++)
      if ((h >> j) & 1) o |= 1ull << g.tile[g.c0 + g.nA + j];
    hB[h] = o;
  }
}

// Packed (diag, then a<b re/im) rho over k bits -> column-major complex.
__host__ __device__ inline void unpack_rho(int k, const double* packed, double* rho) {
  const int D = 1 << k;
  for (int i = 0; i < 2 * D * D; i++) rho[i] = 0.0;
  for (int a = 0; a < D; a++) rho[2 * (a + a * D)] = packed[a];
  int p = D;
  for (int a = 0; a < D; a++)
    for (int b = a + 1; b < D; b++) {
      rho[2 * (a + b * D)] = packed[p];
      rho[2 * (a + b * D) + 1] = packed[p + 1];
      rho[2 * (b + a * D)] = packed[p];
      rho[2 * (b + a * D) + 1] = -packed[p + 1];
      p += 2;
    }
}

__host__ __device__ inline void to_result(const Plan& plan, const double* sum, PassResult* res) {
  res->norm2 = sum[0];
  for (int j = 0; j < 16; j++) res->hist[j] = sum[1 + j];
  for (int j = 0; j < 32; j++) res->rho[j] = 0.0;
  if (plan.red == bti::RED_RHO) unpack_rho(plan.red_k, sum + 17, res->rho);
}

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

template <typename R, int NT>
__global__ void __launch_bounds__(NT) k_batch(BatchArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  C<R>* tile = (C<R>*)smem;
  __shared__ PassGeom geom;
  __shared__ C<R> mats[kMatEntries];
  __shared__ uint64_t hA[128], hB[256];
  __shared__ double red[(NT / 64) * 17];
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
    for (uint64_t i = threadIdx.x; i < len; i += NT) state[i] = cmake<C<R>>(i == 0 ? (R)1 : (R)0, (R)0);
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
    for (;;) {
      if (threadIdx.x == 0) {
        int c = ip->step(&plan, &res);
        if (c && ++npass > a.max_passes) {
          ip->err = BT_ERR_STATE;
          c = 0;
        }
        if (c) build_geom<R>(plan, ip->slots, a.nbits, &geom, mats);
        cmd = c;
      }
      __syncthreads();
      if (!cmd) break;
      build_tables<NT>(geom, hA, hB);
      if (threadIdx.x < kAcc) sums[threadIdx.x] = 0.0;
      __syncthreads();
      tile_pass<R, NT>(state, geom, mats, tile, hA, hB, 0, 1, ntiles, red, sums);
      if (threadIdx.x == 0) to_result(plan, sums, &res);
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
        C<R> v = state[i];
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

template <typename R, int NT>
__global__ void __launch_bounds__(NT) k_grid_pass(C<R>* state, const GridTable<R>* tab, uint64_t ntiles,
                                                  double* part) {
  extern __shared__ __align__(16) unsigned char smem[];
  C<R>* tile = (C<R>*)smem;
  __shared__ PassGeom geom;
  __shared__ C<R> mats[kMatEntries];
  __shared__ uint64_t hA[128], hB[256];
  __shared__ double red[(NT / 64) * 17];
  __shared__ double sums[kAcc];
  {
    const int* src = (const int*)&tab->g;
    int* dst = (int*)&geom;
    for (int i = threadIdx.x; i < (int)(sizeof(PassGeom) / sizeof(int)); i += NT) dst[i] = src[i];
    for (int i = threadIdx.x; i < kMatEntries; i += NT) mats[i] = tab->mats[i];
  }
  __syncthreads();
  build_tables<NT>(geom, hA, hB);
  if (threadIdx.x < kAcc) sums[threadIdx.x] = 0.0;
  __syncthreads();
  tile_pass<R, NT>(state, geom, mats, tile, hA, hB, blockIdx.x, gridDim.x, ntiles, red, sums);
  if (threadIdx.x < kAcc) part[(uint64_t)blockIdx.x * kAcc + threadIdx.x] = sums[threadIdx.x];
}

__global__ void k_finalize_acc(const double* part, int nblocks, double* out) {
  for (int j = threadIdx.x; j < kAcc; j += blockDim.x) {
    double t = 0.0;
    for (int b = 0; b < nblocks; b++) t += part[(size_t)b * kAcc + j];
    out[j] = t;
  }
}

}  // namespace bt

// ---------------------------------------------------------------------------
// circuits
// ---------------------------------------------------------------------------
struct bt_circuit_s {
  btc::CircuitHost h;
  // device copies (uploaded on first use)
  bt_ir_op* d_ops = nullptr;
  double* d_pool = nullptr;
  bti::ChanDesc* d_chan = nullptr;
  double* d_chanw = nullptr;
  std::mutex mu;
};

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

// Tile width for a state of nbits: the largest that fits the LDS budget of a
// 512-thread workgroup, or an override from the autotuner.
static int g_tile_override[2] = {0, 0};

int default_tile_bits(int prec) { return prec == BT_C64 ? 14 : 13; }

int tile_bits_for(int prec, int nbits) {
  int T = g_tile_override[prec] ? g_tile_override[prec] : default_tile_bits(prec);
  return std::min(T, nbits);
}

constexpr int kNT = 512;

// Host driver of the interpreter over grid passes on state s (SV or DM).
static int run_grid(bt_state_s* s, const btc::CircuitHost& ch, uint64_t seed, uint64_t traj, const double* uni,
                    int64_t nuni, int32_t* outs, int64_t outcap, int64_t* nout, bt_run_stats* stats, int fold,
                    int prec_tile_T) {
  const int nbits = s->nbits;
  std::vector<QOp> slots(bti::kQMax);
  std::unique_ptr<bti::Interp> ip(new bti::Interp());
  bti::Program prog{ch.ops.data(), ch.ntop, ch.pool.data(), ch.chan.data(), ch.chanw.data(), nbits};
  const int T = prec_tile_T;
  ip->init(prog, slots.data(), T, 3);
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
  const int cus = ctx().cus > 0 ? ctx().cus : 256;
  const unsigned blocks = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)cus * 2);
  const size_t lds = (size_t)(1u << std::min(T, nbits)) * s->elem();
  constexpr int kRing = 8;
  void* h_ring = nullptr;
  void* d_ring = nullptr;
  double* d_part = nullptr;
  double* d_acc = nullptr;
  double* h_acc = nullptr;
  hipEvent_t ev[kRing];
  hipEvent_t e0, e1;
  int rc = BT_OK;
  for (int i = 0; i < kRing; i++) ev[i] = nullptr;
  BT_HIP(hipHostMalloc(&h_ring, tabsz * kRing));
  if (hipMalloc(&d_ring, tabsz * kRing) != hipSuccess || hipMalloc(&d_part, sizeof(double) * kAcc * blocks) ||
      hipMalloc(&d_acc, sizeof(double) * kAcc) || hipHostMalloc(&h_acc, sizeof(double) * kAcc)) {
    rc = fail(BT_ERR_NOMEM, "run_grid: allocation failed");
  }
  for (int i = 0; i < kRing && rc == BT_OK; i++)
    if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) rc = fail(BT_ERR_HIP, "event");
  if (rc == BT_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(BT_ERR_HIP, "event");
  if (rc == BT_OK) {
    if (s->prec == BT_C64)
      (void)hipFuncSetAttribute((const void*)k_grid_pass<float, kNT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
    else
      (void)hipFuncSetAttribute((const void*)k_grid_pass<double, kNT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
    (void)hipEventRecord(e0, s->stream);
  }
  Plan plan;
  PassResult res;
  memset(&res, 0, sizeof res);
  int64_t npass = 0;
  const int64_t max_passes = 64 + 8 * (int64_t)ch.ops.size();
  while (rc == BT_OK && ip->step(&plan, &res)) {
    if (++npass > max_passes) {
      rc = fail(BT_ERR_STATE, "run_grid: pass limit exceeded");
      break;
    }
    const int slot = (int)(npass % kRing);
    if (npass > kRing && hipEventSynchronize(ev[slot]) != hipSuccess) {
      rc = fail(BT_ERR_HIP, "run_grid: event sync");
      break;
    }
    char* hslot = (char*)h_ring + tabsz * slot;
    char* dslot = (char*)d_ring + tabsz * slot;
    if (s->prec == BT_C64) {
      auto* t = (GridTable<float>*)hslot;
      build_geom<float>(plan, slots.data(), nbits, &t->g, t->mats);
    } else {
      auto* t = (GridTable<double>*)hslot;
      build_geom<double>(plan, slots.data(), nbits, &t->g, t->mats);
    }
    if (hipMemcpyAsync(dslot, hslot, tabsz, hipMemcpyHostToDevice, s->stream) != hipSuccess) {
      rc = fail(BT_ERR_HIP, "run_grid: table copy");
      break;
    }
    (void)hipEventRecord(ev[slot], s->stream);
    if (s->prec == BT_C64)
      hipLaunchKernelGGL((k_grid_pass<float, kNT>), dim3(blocks), dim3(kNT), lds, s->stream, (C<float>*)s->d,
                         (const GridTable<float>*)dslot, ntiles, d_part);
    else
      hipLaunchKernelGGL((k_grid_pass<double, kNT>), dim3(blocks), dim3(kNT), lds, s->stream, (C<double>*)s->d,
                         (const GridTable<double>*)dslot, ntiles, d_part);
    if (hipGetLastError() != hipSuccess) {
      rc = fail(BT_ERR_HIP, "run_grid: pass launch");
      break;
    }
    if (plan.aligned || plan.red != bti::RED_NONE) {
      hipLaunchKernelGGL(k_finalize_acc, dim3(1), dim3(64), 0, s->stream, (const double*)d_part, (int)blocks, d_acc);
      if (hipMemcpyAsync(h_acc, d_acc, sizeof(double) * kAcc, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
          hipStreamSynchronize(s->stream) != hipSuccess) {
        rc = fail(BT_ERR_HIP, "run_grid: reduction copy");
        break;
      }
      to_result(plan, h_acc, &res);
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
    float ms = 0;
    if (rc == BT_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) stats->kernel_ms = ms;
  }
  if (nout) *nout = ip->nout;
  for (int i = 0; i < kRing; i++)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (h_ring) (void)hipHostFree(h_ring);
  if (h_acc) (void)hipHostFree(h_acc);
  if (d_ring) (void)hipFree(d_ring);
  if (d_part) (void)hipFree(d_part);
  if (d_acc) (void)hipFree(d_acc);
  return rc;
}

using cd = std::complex<double>;

// Density-matrix lowering: every op becomes superoperators on vec(rho)
// (row bit b at b, column bit b at b + N).  A gate followed by its noise
// channel on the same bits becomes one superoperator sum_i (K_i U) (x)
// conj(K_i U); other gates become U on the row bits and conj(U) on the column
// bits (controls on both sides).
static int lower_dm(const btc::CircuitHost& c, btc::CircuitHost* out, std::string* msg) {
  const int N = c.n;
  out->n = 2 * N;
  out->ops.clear();
  out->pool.clear();
  out->chanw.assign(1, 0.0);
  auto add_gate = [&](int k, const int32_t* bits, uint64_t ctrl, const std::vector<cd>& m) {
    bt_ir_op o;
    memset(&o, 0, sizeof o);
    o.kind = BT_OP_GATE;
    o.k = k;
    for (int i = 0; i < k; i++) o.bits[i] = bits[i];
    o.ctrl_mask = ctrl;
    o.mat = (int64_t)(out->pool.size() / 2);
    for (const cd& v : m) {
      out->pool.push_back(v.real());
      out->pool.push_back(v.imag());
    }
    out->ops.push_back(o);
  };
  auto superop = [&](int k, const std::vector<std::vector<cd>>& A, const int32_t* bits) {
    const int D = 1 << k, S = D * D;
    std::vector<cd> sup(S * S, cd(0, 0));
    for (const auto& K : A)
      for (int r1 = 0; r1 < D; r1++)
        for (int c1 = 0; c1 < D; c1++)
          for (int r0 = 0; r0 < D; r0++)
            for (int c0 = 0; c0 < D; c0++) {
              cd v = K[r1 + r0 * D] * std::conj(K[c1 + c0 * D]);
              sup[((r1 << k) | c1) + ((r0 << k) | c0) * S] += v;
            }
    int32_t sb[4];
    for (int i = 0; i < k; i++) {
      sb[i] = bits[i];
      sb[k + i] = bits[i] + N;
    }
    add_gate(2 * k, sb, 0, sup);
  };
  auto mat = [&](const bt_ir_op& op, int idx) {
    const int D = 1 << op.k;
    std::vector<cd> m(D * D);
    const double* p = c.pool.data() + 2 * (op.mat + (int64_t)idx * D * D);
    for (int i = 0; i < D * D; i++) m[i] = cd(p[2 * i], p[2 * i + 1]);
    return m;
  };
  for (int i = 0; i < c.ntop; i++) {
    const bt_ir_op& op = c.ops[i];
    if (op.kind == BT_OP_MEASURE || op.kind == BT_OP_IF) {
      *msg = "fix this: density-matrix measurement is not implemented";
      return BT_ERR_UNSUPPORTED;
    }
    if (op.kind == BT_OP_RESET) {  // OpQC("RES") on rho: amplitude damping with gamma = 1
      std::vector<cd> k0 = {1, 0, 0, 0}, k1 = {0, 0, 1, 0};
      superop(1, {k0, k1}, op.bits);
      continue;
    }
    if (op.kind == BT_OP_CHANNEL) {
      if (op.k > 2) {
        *msg = "3-qubit channels have no density-matrix method";
        return BT_ERR_UNSUPPORTED;
      }
      std::vector<std::vector<cd>> A;
      for (int j = 0; j < op.nkraus; j++) A.push_back(mat(op, j));
      superop(op.k, A, op.bits);
      continue;
    }
    // gate
    if (op.k > 2) {
      *msg = "density-matrix gates take k <= 2";
      return BT_ERR_UNSUPPORTED;
    }
    const std::vector<cd> U = mat(op, 0);
    const int D = 1 << op.k;
    if (op.ctrl_mask == 0 && i + 1 < c.ntop && c.ops[i + 1].kind == BT_OP_CHANNEL && c.ops[i + 1].k == op.k &&
        std::equal(op.bits, op.bits + op.k, c.ops[i + 1].bits)) {
      const bt_ir_op& ch = c.ops[i + 1];
      std::vector<std::vector<cd>> A;
      for (int j = 0; j < ch.nkraus; j++) {
        std::vector<cd> K = mat(ch, j), KU(D * D, cd(0, 0));
        for (int r = 0; r < D; r++)
          for (int cc = 0; cc < D; cc++)
            for (int x = 0; x < D; x++) KU[r + cc * D] += K[r + x * D] * U[x + cc * D];
        A.push_back(KU);
      }
      superop(op.k, A, op.bits);
      i++;
      continue;
    }
    std::vector<cd> Uc(U.size());
    for (size_t j = 0; j < U.size(); j++) Uc[j] = std::conj(U[j]);
    int32_t cb[4];
    for (int j = 0; j < op.k; j++) cb[j] = op.bits[j] + N;
    add_gate(op.k, op.bits, op.ctrl_mask, U);
    add_gate(op.k, cb, op.ctrl_mask << N, Uc);
  }
  out->ntop = (int)out->ops.size();
  out->chan.assign(out->ops.size(), bti::ChanDesc{1, 0, 0});  // no renormalisation bookkeeping
  out->max_draws = 0;
  out->has_measure = false;
  return BT_OK;
}

}  // namespace bt

using namespace bt;

extern "C" {

int bt_circuit_create(int nqubits, const bt_ir_op* ops, int nops, int ntop, const double* pool, int64_t npool,
                      bt_circuit* out) {
  if (!out) return fail(BT_ERR_ARG, "null output");
  *out = nullptr;
  auto* c = new bt_circuit_s();
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
  return run_grid(s, c->h, seed, traj, uniforms, nuni, outcomes, outcap, nout, stats, 1,
                  tile_bits_for(s->prec, s->nbits));
}

int bt_run_circuit_dm(bt_state rho, bt_circuit c, bt_run_stats* stats) {
  if (!rho || !c) return fail(BT_ERR_ARG, "null state or circuit");
  int rc = ensure_device();
  if (rc) return rc;
  if (rho->kind != BT_DM) return fail(BT_ERR_ARG, "bt_run_circuit_dm needs a density matrix");
  if (rho->n != c->h.n) return fail(BT_ERR_ARG, "circuit and state qubit counts differ");
  btc::CircuitHost dm;
  std::string msg;
  rc = lower_dm(c->h, &dm, &msg);
  if (rc) return fail(rc, msg);
  int64_t nout = 0;
  return run_grid(rho, dm, 0, 0, nullptr, 0, nullptr, 0, &nout, stats, 0, tile_bits_for(rho->prec, rho->nbits));
}

int bt_run_batch(bt_circuit c, int prec, uint64_t seed, int64_t traj0, int64_t ntraj, int* outcomes, int64_t outcap,
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
  // one workgroup per CU for wide tiles (LDS-bound), two for narrow ones
  int per_cu = (size_t)(1u << T) * elem > 80 * 1024 ? 1 : 2;
  int64_t nwg = std::min<int64_t>(ntraj, (int64_t)cus * per_cu);
  // keep the per-workgroup states within a quarter of free HBM
  size_t freeb = 0, totalb = 0;
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
    return fail(BT_ERR_NOMEM, "bt_run_batch: device allocation failed");
  }
  if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    cleanup();
    return fail(BT_ERR_HIP, "bt_run_batch: stream/event");
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
  a.cmin = 3;
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
    (void)hipFuncSetAttribute((const void*)k_batch<float, kNT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_batch<float, kNT>), dim3((unsigned)nwg), dim3(kNT), lds, st, a);
  } else {
    (void)hipFuncSetAttribute((const void*)k_batch<double, kNT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL((k_batch<double, kNT>), dim3((unsigned)nwg), dim3(kNT), lds, st, a);
  }
  hipError_t le = hipGetLastError();
  (void)hipEventRecord(e1, st);
  hipError_t se = hipStreamSynchronize(st);
  if (le != hipSuccess || se != hipSuccess) {
    std::string m = std::string("bt_run_batch: kernel failed: ") + hipGetErrorString(le != hipSuccess ? le : se);
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
    // bytes: read+write per pass, read per reduction pass, one write to zero
    // each trajectory's state
    stats->bytes = ((double)hst[0] * 2.0 + (double)hst[1] + (double)ntraj) * (double)len * (double)elem;
    float ms = 0;
    if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) stats->kernel_ms = ms;
  }
  cleanup();
  if (herr) return fail(herr, "bt_run_batch: a trajectory failed (interpreter error " + std::to_string(herr) + ")");
  return BT_OK;
}

// Tile width override used by the autotuner (0 restores the default).
int bt_set_tile_bits(int prec, int T) {
  if (prec != BT_C64 && prec != BT_C128) return fail(BT_ERR_ARG, "bad precision");
  if (T != 0 && (T < 2 || T > default_tile_bits(prec))) return fail(BT_ERR_ARG, "tile bits out of range");
  g_tile_override[prec] = T;
  return BT_OK;
}

}  // extern "C"
