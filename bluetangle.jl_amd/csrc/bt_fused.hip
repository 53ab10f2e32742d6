// Dense 5-qubit blocks on FP32 MFMA (complex64 state vectors), and the fused
// circuit executor built on them.
//
// A 32 x 32 complex block M acting on 5 bits is, per group of 32 amplitudes
// x = xr + i xi, y = M x computed with three real 32 x 32 products (Gauss):
// t1 = Re M xr, t2 = Im M xi, t3 = (Re M + Im M)(xr + xi); then
// Re y = t1 - t2, Im y = t3 - t1 - t2.  k_fused5 streams the state in tiles
// of 2^12 amplitudes (the 5 block bits + the 7 lowest other bits): a
// coalesced load into LDS, then every wavefront multiplies 32 of the tile's
// 128 groups with v_mfma_f32_32x32x2f32 (the groups as the 32 columns; step s
// takes components 2s in lanes 0-31 and 2s+1 in lanes 32-63; the three real
// matrices are staged once per workgroup in LDS in lane order and read per
// step), writes its results back over its own groups, and the tile is stored
// coalesced.  44 KiB of LDS: three workgroups per CU.  Per amplitude: 96
// real MACs (three quarters of the 64 x 64 real form), 12 flop/byte of HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "bt_circuit.hpp"
#include "bt_fuse.hpp"
#include "bt_internal.hpp"

namespace bt {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kF5T = 12;  // tile bits
constexpr int kF5N = 1 << kF5T;
constexpr int kF5Lane = 3 * 16 * 64;  // floats of one block's MFMA operand (three real matrices)

struct F5Args {
  float2* s;
  uint64_t ntiles;
  int32_t tile[kF5T];   // local bit -> global bit, ascending
  int32_t nrun;         // tile index -> global base: runs of consecutive outside bits
  int32_t run_dst[40], run_len[40];
  int32_t vloc[7];      // local positions of the group-index bits, ascending
  int32_t cloc[5];      // local position of operand bit i (i = 0 is the MSB)
};

// LDS slot of local index l: XOR-swizzle of the low 5 bits (8-byte elements)
__device__ __forceinline__ uint32_t f5_swz(uint32_t l) { return l ^ (((l >> 5) ^ (l >> 10)) & 31u); }

__global__ void __launch_bounds__(256, 3) k_fused5(F5Args a, const float* __restrict__ Alane) {
  __shared__ float2 lds[kF5N];
  __shared__ float la[kF5Lane];  // la[(t * 16 + s) * 64 + lane], t: Re M, Im M, Re M + Im M
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int i = 0; i < kF5Lane / 256; i++) la[tid + 256 * i] = Alane[tid + 256 * i];
  // this lane's group: g = 32 * wave + lane % 32 -> local base
  const uint32_t grp = 32 * wave + (lane & 31);
  uint32_t gb = 0;
#pragma unroll
  for (int j = 0; j < 7; j++)
    if ((grp >> j) & 1) gb |= 1u << a.vloc[j];
  const uint32_t s0 = f5_swz(gb);
  // component c's local offset (operand 0 = MSB of c), swizzled (the swizzle is XOR-linear)
  uint32_t cb[5];
#pragma unroll
  for (int i = 0; i < 5; i++) cb[i] = __builtin_amdgcn_readfirstlane(f5_swz(1u << a.cloc[i]));
  // load / store: thread tid takes local indices tid + 256 i (i < 16)
  uint64_t dlo = 0;
#pragma unroll
  for (int b = 0; b < 8; b++)
    if ((tid >> b) & 1) dlo |= 1ull << a.tile[b];
  uint64_t dhi[4];
#pragma unroll
  for (int b = 0; b < 4; b++) dhi[b] = 1ull << a.tile[8 + b];
  const bool hi_half = lane >= 32;
  // B of step s is component 2s + (lane >= 32): bit 0 of the component is operand 4
  const uint32_t s1 = s0 ^ (hi_half ? cb[4] : 0u);
  auto tile_base = [&](uint64_t t) {
    uint64_t base = 0;
    for (int r = 0; r < a.nrun; r++) {
      base |= (t & ((1ull << a.run_len[r]) - 1ull)) << a.run_dst[r];
      t >>= a.run_len[r];
    }
    return base;
  };
  float2 x[16];
  auto fetch = [&](uint64_t base) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      uint64_t gi = base | dlo;
#pragma unroll
      for (int b = 0; b < 4; b++)
        if ((i >> b) & 1) gi |= dhi[b];
      x[i] = a.s[gi];
    }
  };
  if (blockIdx.x < a.ntiles) fetch(tile_base(blockIdx.x));
  for (uint64_t t = blockIdx.x; t < a.ntiles; t += gridDim.x) {
    const uint64_t base = tile_base(t);
#pragma unroll
    for (int i = 0; i < 16; i++) lds[f5_swz(tid + 256u * i)] = x[i];
    // the next tile's loads fly during this tile's MFMA work (tiles are disjoint)
    if (t + gridDim.x < a.ntiles) fetch(tile_base(t + gridDim.x));
    __syncthreads();
    floatx16 t1 = {}, t2 = {}, t3 = {};
#pragma unroll
    for (int st = 0; st < 16; st++) {
      uint32_t sc = 0;
#pragma unroll
      for (int k = 0; k < 4; k++)
        if ((st >> (3 - k)) & 1) sc ^= cb[k];
      const float2 e = lds[s1 ^ sc];
      t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(la[st * 64 + lane], e.x, t1, 0, 0, 0);
      t2 = __builtin_amdgcn_mfma_f32_32x32x2f32(la[(16 + st) * 64 + lane], e.y, t2, 0, 0, 0);
      t3 = __builtin_amdgcn_mfma_f32_32x32x2f32(la[(32 + st) * 64 + lane], e.x + e.y, t3, 0, 0, 0);
    }
    // each wavefront reads and then overwrites only its own 32 groups: no barrier here
    // accumulator j holds output component 8 (j / 4) + 4 (lane / 32) + j % 4
#pragma unroll
    for (int j = 0; j < 16; j++) {
      const int r = 8 * (j / 4) + (j % 4) + (hi_half ? 4 : 0);
      uint32_t sc = 0;
#pragma unroll
      for (int i = 0; i < 5; i++)
        if ((r >> (4 - i)) & 1) sc ^= cb[i];
      const float re = t1[j] - t2[j];
      lds[s0 ^ sc] = make_float2(re, t3[j] - t1[j] - t2[j]);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 16; i++) {
      uint64_t gi = base | dlo;
#pragma unroll
      for (int b = 0; b < 4; b++)
        if ((i >> b) & 1) gi |= dhi[b];
      a.s[gi] = lds[f5_swz(tid + 256u * i)];
    }
    __syncthreads();
  }
}

// The three real matrices Re M, Im M, Re M + Im M of a 32 x 32 complex128
// column-major block in MFMA lane order:
// out[(t * 16 + s) * 64 + lane] = A_t[lane % 32][2 s + lane / 32].
void fused5_lane_matrix(const double* mat, float* out) {
  for (int t = 0; t < 3; t++)
    for (int st = 0; st < 16; st++)
      for (int lane = 0; lane < 64; lane++) {
        const int r = lane % 32, c = 2 * st + lane / 32;
        const double re = mat[2 * (r + c * 32)], im = mat[2 * (r + c * 32) + 1];
        out[(t * 16 + st) * 64 + lane] = (float)(t == 0 ? re : (t == 1 ? im : re + im));
      }
}

// psi <- M psi on 5 bits (bits[0] = MSB); d_alane: M in MFMA lane order on
// the device (kF5Lane floats).  Enqueued on the state's stream.
int launch_fused5(bt_state_s* s, const int* bits, const float* d_alane) {
  const int N = s->n;
  if (s->kind != BT_SV || s->prec != BT_C64 || N < kF5T) return fail(BT_ERR_UNSUPPORTED, "fused5: c64 SV, N >= 12");
  F5Args a;
  memset(&a, 0, sizeof a);
  a.s = (float2*)s->d;
  uint64_t bm = 0;
  for (int i = 0; i < 5; i++) {
    if (bits[i] < 0 || bits[i] >= N) return fail(BT_ERR_ARG, "fused5: bit out of range");
    bm |= 1ull << bits[i];
  }
  if (__builtin_popcountll(bm) != 5) return fail(BT_ERR_ARG, "fused5: repeated bit");
  uint64_t tm = bm;
  for (int b = 0; b < N && __builtin_popcountll(tm) < kF5T; b++) tm |= 1ull << b;
  int n = 0;
  int rank[64];
  for (int b = 0; b < 64; b++) rank[b] = -1;
  for (int b = 0; b < N; b++)
    if ((tm >> b) & 1) {
      rank[b] = n;
      a.tile[n++] = b;
    }
  int nv = 0;
  for (int l = 0; l < kF5T; l++)
    if (!((bm >> a.tile[l]) & 1)) a.vloc[nv++] = l;
  for (int i = 0; i < 5; i++) a.cloc[i] = rank[bits[i]];
  a.nrun = 0;
  int prev = -2;
  for (int b = 0; b < N; b++)
    if (!((tm >> b) & 1)) {
      if (b != prev + 1) {
        a.run_dst[a.nrun] = b;
        a.run_len[a.nrun] = 0;
        a.nrun++;
      }
      a.run_len[a.nrun - 1]++;
      prev = b;
    }
  a.ntiles = 1ull << (N - kF5T);
  // one wave of resident workgroups (each strides over the tiles): a grid
  // of several partial waves would leave the last one running on part of
  // the chip
  static const int per_cu = [] {  // thread-safe one-time query (every device runs the same code object)
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused5, 256, 0) != hipSuccess || nb < 1) nb = 1;
    return nb;
  }();
  const int cus = ctx().cus > 0 ? ctx().cus : 256;
  const unsigned grid = (unsigned)std::min<uint64_t>(a.ntiles, (uint64_t)cus * per_cu);
  hipLaunchKernelGGL(k_fused5, dim3(grid), dim3(256), 0, s->stream, a, d_alane);
  BT_HIP(hipGetLastError());
  return BT_OK;
}

}  // namespace bt

using namespace bt;

extern "C" int bt_apply_fused5(bt_state s, const double* mat, const int* bits) {
  int rc = ensure_device();
  if (rc) return rc;
  if (!s || !mat || !bits) return fail(BT_ERR_ARG, "null argument");
  std::vector<float> al(kF5Lane);
  fused5_lane_matrix(mat, al.data());
  float* d = nullptr;
  BT_HIP(hipMalloc((void**)&d, sizeof(float) * al.size()));
  rc = hipMemcpyAsync(d, al.data(), sizeof(float) * al.size(), hipMemcpyHostToDevice, s->stream) == hipSuccess
           ? launch_fused5(s, bits, d)
           : fail(BT_ERR_HIP, "fused5: matrix upload");
  (void)hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  return rc;
}

extern "C" int bt_run_fused(bt_state s, bt_circuit c, int max_k, int norm_every, double* norms, int64_t nnorms,
                            int64_t* nblocks, bt_run_stats* stats) {
  int rc = ensure_device();
  if (rc) return rc;
  if (!s || !c) return fail(BT_ERR_ARG, "null state or circuit");
  if (s->kind != BT_SV) return fail(BT_ERR_ARG, "bt_run_fused: state vectors only");
  if (s->n != c->h.n) return fail(BT_ERR_ARG, "circuit and state qubit counts differ");
  if (max_k == 0) {  // the autotuner's choice (5 when not tuned)
    max_k = c->fused_k < 0 ? 5 : c->fused_k;
    if (max_k == 0) {  // the register-tile executor: one pass per tile plan, no norm checks in between
      int64_t nout = 0;
      rc = bt_run_circuit(s, c, 0, 0, nullptr, 0, nullptr, 0, &nout, stats);
      if (rc == BT_OK && nblocks) *nblocks = stats ? stats->passes : 0;
      if (rc == BT_OK && norms && norm_every > 0 && nnorms > 0) rc = bt_norm2(s, norms);
      return rc;
    }
  }
  const bool mfma = s->prec == BT_C64 && max_k == 5 && s->n >= kF5T;
  std::vector<btf::Block> blocks;
  std::string msg;
  rc = btf::fuse_gates(c->h, max_k, mfma ? 5 : 0, &blocks, &msg);
  if (rc) return fail(rc, msg);
  if (nblocks) *nblocks = (int64_t)blocks.size();
  // every 5-bit block's MFMA operand in one device buffer
  float* d_al = nullptr;
  float* h_al = nullptr;
  const size_t per = kF5Lane;
  if (mfma && !blocks.empty()) {
    BT_HIP(hipHostMalloc((void**)&h_al, sizeof(float) * per * blocks.size()));
    for (size_t b = 0; b < blocks.size(); b++)
      fused5_lane_matrix((const double*)blocks[b].m.data(), h_al + per * b);
    if (hipMalloc((void**)&d_al, sizeof(float) * per * blocks.size()) != hipSuccess) {
      (void)hipHostFree(h_al);
      return fail(BT_ERR_NOMEM, "bt_run_fused: matrix buffer");
    }
    if (hipMemcpyAsync(d_al, h_al, sizeof(float) * per * blocks.size(), hipMemcpyHostToDevice, s->stream) !=
        hipSuccess)
      rc = fail(BT_ERR_HIP, "bt_run_fused: matrix upload");
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (rc == BT_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = fail(BT_ERR_HIP, "event");
  if (rc == BT_OK) (void)hipEventRecord(e0, s->stream);
  int64_t nn = 0, reads = 0;
  for (size_t b = 0; rc == BT_OK && b < blocks.size(); b++) {
    const btf::Block& blk = blocks[b];
    if (mfma) {
      rc = launch_fused5(s, blk.bits, d_al + per * b);
    } else {
      Placement p;
      rc = make_placement(s->nbits, blk.k, blk.bits, 0, &p);
      if (rc == BT_OK) rc = launch_apply(s, s->nbits, blk.k, (const double*)blk.m.data(), p, 1.0, false);
    }
    if (rc == BT_OK && norms && norm_every > 0 && (int64_t)(b + 1) % norm_every == 0 && nn < nnorms) {
      rc = bt_norm2(s, norms + nn);
      nn++;
      reads++;
    }
  }
  if (rc == BT_OK) {
    (void)hipEventRecord(e1, s->stream);
    if (hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(BT_ERR_HIP, "bt_run_fused: sync");
  }
  if (stats) {
    memset(stats, 0, sizeof *stats);
    stats->passes = (int64_t)blocks.size();
    stats->reduce_passes = reads;
    stats->bytes = ((double)blocks.size() * 2.0 + (double)reads) * (double)s->len * (double)s->elem();
    stats->launches = (int64_t)blocks.size();
    float ms = 0;
    if (rc == BT_OK && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) stats->kernel_ms = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (d_al) (void)hipFree(d_al);
  if (h_al) (void)hipHostFree(h_al);
  return rc;
}

// ---------------------------------------------------------------------------
// HIP graphs of fused circuits: the whole pass sequence captured once and
// replayed with one launch (no per-pass host work).  The graph owns its
// matrices; it is bound to the state it was captured on.
// ---------------------------------------------------------------------------
struct bt_graph_s {
  bt_state_s* s = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  float* d_al = nullptr;
  void* d_tab = nullptr;    // density-matrix graphs: the passes' tables
  double* d_part = nullptr; // and their (unused) reduction partials
  std::vector<btf::Block> blocks;
  int64_t nblocks = 0;
};

extern "C" int bt_graph_create_fused(bt_state s, bt_circuit c, int max_k, bt_graph* out) {
  int rc = ensure_device();
  if (rc) return rc;
  if (!s || !c || !out) return fail(BT_ERR_ARG, "null argument");
  *out = nullptr;
  if (s->kind != BT_SV) return fail(BT_ERR_ARG, "bt_graph_create_fused: state vectors only");
  if (s->n != c->h.n) return fail(BT_ERR_ARG, "circuit and state qubit counts differ");
  const bool mfma = s->prec == BT_C64 && max_k == 5 && s->n >= kF5T;
  if (!mfma && max_k > 3)
    return fail(BT_ERR_UNSUPPORTED, "bt_graph_create_fused: complex64 with max_k = 5 (MFMA) or max_k <= 3");
  auto* g = new bt_graph_s();
  g->s = s;
  std::string msg;
  rc = btf::fuse_gates(c->h, max_k, mfma ? 5 : 0, &g->blocks, &msg);
  if (rc) {
    delete g;
    return fail(rc, msg);
  }
  g->nblocks = (int64_t)g->blocks.size();
  const size_t per = kF5Lane;
  auto cleanup = [&](int r) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->d_al) (void)hipFree(g->d_al);
    delete g;
    return r;
  };
  if (mfma && g->nblocks) {
    std::vector<float> al(per * g->blocks.size());
    for (size_t b = 0; b < g->blocks.size(); b++)
      fused5_lane_matrix((const double*)g->blocks[b].m.data(), al.data() + per * b);
    if (hipMalloc((void**)&g->d_al, sizeof(float) * al.size()) != hipSuccess)
      return cleanup(fail(BT_ERR_NOMEM, "bt_graph_create_fused: matrix buffer"));
    if (hipMemcpy(g->d_al, al.data(), sizeof(float) * al.size(), hipMemcpyHostToDevice) != hipSuccess)
      return cleanup(fail(BT_ERR_HIP, "bt_graph_create_fused: matrix upload"));
  }
  if (hipStreamSynchronize(s->stream) != hipSuccess) return cleanup(fail(BT_ERR_HIP, "sync"));
  if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return cleanup(fail(BT_ERR_HIP, "bt_graph_create_fused: begin capture"));
  int crc = BT_OK;
  for (size_t b = 0; crc == BT_OK && b < g->blocks.size(); b++) {
    const btf::Block& blk = g->blocks[b];
    if (mfma) {
      crc = launch_fused5(s, blk.bits, g->d_al + per * b);
    } else {
      Placement p;
      crc = make_placement(s->nbits, blk.k, blk.bits, 0, &p);
      if (crc == BT_OK) crc = launch_apply(s, s->nbits, blk.k, (const double*)blk.m.data(), p, 1.0, false);
    }
  }
  const hipError_t ec = hipStreamEndCapture(s->stream, &g->graph);
  if (crc) return cleanup(crc);
  if (ec != hipSuccess || !g->graph) return cleanup(fail(BT_ERR_HIP, "bt_graph_create_fused: end capture"));
  if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    return cleanup(fail(BT_ERR_HIP, "bt_graph_create_fused: instantiate"));
  *out = g;
  return BT_OK;
}

extern "C" int bt_graph_launch(bt_graph g) {
  if (!g || !g->exec) return fail(BT_ERR_ARG, "null graph");
  BT_HIP(hipGraphLaunch(g->exec, g->s->stream));
  return BT_OK;
}

extern "C" int64_t bt_graph_passes(bt_graph g) { return g ? g->nblocks : -1; }

extern "C" int bt_graph_destroy(bt_graph g) {
  if (!g) return BT_OK;
  if (g->s && g->s->stream) (void)hipStreamSynchronize(g->s->stream);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->d_al) (void)hipFree(g->d_al);
  if (g->d_tab) (void)hipFree(g->d_tab);
  if (g->d_part) (void)hipFree(g->d_part);
  delete g;
  return BT_OK;
}

// to_rho as one HIP graph: the circuit lowered to superoperator passes on
// vec(rho) (the density-matrix executor's plan, which depends on nothing the
// passes compute), every pass's table uploaded once and every launch
// recorded; bt_graph_launch replays the whole circuit on rho.
extern "C" int bt_graph_create_dm(bt_state rho, bt_circuit c, bt_graph* out) {
  int rc = ensure_device();
  if (rc) return rc;
  if (!rho || !c || !out) return fail(BT_ERR_ARG, "null argument");
  *out = nullptr;
  if (rho->kind != BT_DM) return fail(BT_ERR_ARG, "bt_graph_create_dm needs a density matrix");
  if (rho->n != c->h.n) return fail(BT_ERR_ARG, "circuit and state qubit counts differ");
  btc::CircuitHost dm;
  std::string msg;
  rc = btc::lower_dm(c->h, &dm, &msg);
  if (rc) return fail(rc, msg);
  const int T = dm_tile_bits(rho);
  std::vector<unsigned char> tables;
  std::vector<uint8_t> narrow;
  size_t tabsz = 0;
  rc = plan_static_passes(rho, dm, T, &tables, &narrow, &tabsz);
  if (rc) return rc;
  auto* g = new bt_graph_s();
  g->s = rho;
  g->nblocks = (int64_t)narrow.size();
  auto cleanup = [&](int r) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->d_tab) (void)hipFree(g->d_tab);
    if (g->d_part) (void)hipFree(g->d_part);
    delete g;
    return r;
  };
  const int cus = ctx().cus > 0 ? ctx().cus : 256;
  if (hipMalloc(&g->d_tab, std::max<size_t>(tables.size(), 16)) != hipSuccess ||
      hipMalloc((void**)&g->d_part, sizeof(double) * kGridAccRows * (size_t)cus * 8) != hipSuccess)
    return cleanup(fail(BT_ERR_NOMEM, "bt_graph_create_dm: tables"));
  if (!tables.empty() && hipMemcpy(g->d_tab, tables.data(), tables.size(), hipMemcpyHostToDevice) != hipSuccess)
    return cleanup(fail(BT_ERR_HIP, "bt_graph_create_dm: table upload"));
  grid_pass_lds_attr(rho->prec, std::min(T, rho->nbits));
  if (hipStreamSynchronize(rho->stream) != hipSuccess) return cleanup(fail(BT_ERR_HIP, "sync"));
  if (hipStreamBeginCapture(rho->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return cleanup(fail(BT_ERR_HIP, "bt_graph_create_dm: begin capture"));
  int crc = BT_OK;
  for (size_t i = 0; crc == BT_OK && i < narrow.size(); i++)
    crc = launch_grid_pass(rho, (const unsigned char*)g->d_tab + i * tabsz, narrow[i] != 0, T, g->d_part);
  const hipError_t ec = hipStreamEndCapture(rho->stream, &g->graph);
  if (crc) return cleanup(crc);
  if (ec != hipSuccess || !g->graph) return cleanup(fail(BT_ERR_HIP, "bt_graph_create_dm: end capture"));
  if (hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    return cleanup(fail(BT_ERR_HIP, "bt_graph_create_dm: instantiate"));
  *out = g;
  return BT_OK;
}

extern "C" int bt_autotune_fused(bt_state s, bt_circuit c, int* best_k, double* ms) {
  int rc = ensure_device();
  if (rc) return rc;
  if (!s || !c || !best_k) return fail(BT_ERR_ARG, "null argument");
  if (s->kind != BT_SV) return fail(BT_ERR_ARG, "bt_autotune_fused: state vectors only");
  if (s->n != c->h.n) return fail(BT_ERR_ARG, "circuit and state qubit counts differ");
  bt_state scratch = nullptr;
  rc = bt_state_create(s->n, BT_SV, s->prec, &scratch);
  if (rc) return rc;
  double t[4] = {-1, -1, -1, -1};
  int best = -1;
  double best_ms = 0;
  for (int i = 0; i < 4 && rc == BT_OK; i++) {
    const int k = i < 3 ? 3 + i : 0;
    if (k == 5 && !(s->prec == BT_C64 && s->n >= kF5T)) continue;  // MFMA blocks: complex64, N >= 12
    bt_run_stats st;
    memset(&st, 0, sizeof st);
    rc = bt_state_set_basis(scratch, 0);
    if (rc) break;
    if (k) {
      int64_t nb = 0;
      rc = bt_run_fused(scratch, c, k, 0, nullptr, 0, &nb, &st);
    } else {
      int64_t nout = 0;
      rc = bt_run_circuit(scratch, c, 0, 0, nullptr, 0, nullptr, 0, &nout, &st);
    }
    if (rc) break;
    t[i] = st.kernel_ms;
    if (best < 0 || t[i] < best_ms) {
      best = k;
      best_ms = t[i];
    }
  }
  (void)bt_state_destroy(scratch);
  if (rc) return rc;
  if (best < 0) return fail(BT_ERR_STATE, "bt_autotune_fused: no candidate ran");
  c->fused_k = best;
  *best_k = best;
  if (ms)
    for (int i = 0; i < 4; i++) ms[i] = t[i];
  return BT_OK;
}
