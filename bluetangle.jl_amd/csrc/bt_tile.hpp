// Tile pass shared by the circuit executors.
//
// A pass streams the state through the chip in tiles of 2^T amplitudes whose
// T local bits are the pass's tile bits.  A workgroup of NT threads keeps a
// tile in registers, 2^RB = 32 amplitudes per thread: RB "register bits" of
// the local index are held inside each thread and the log2(NT) others select
// the thread.  Operators are applied in registers with the matrix pattern
// fixed at compile time (one instantiation per set of register positions),
// so a pass with many fused operators costs FMAs, not memory traffic.  When
// the next operators need other bits in registers the tile is re-laid out
// through LDS (one swizzled write + read: a "phase" change).  Global
// addresses are bit deposits of (tile index, thread, register) - no tables -
// and the layouts used to load and store keep the low local bits on the
// threads, so each wavefront touches contiguous 64 B+ runs.  Reductions (norm,
// histogram over up to 4 bits, reduced rho over up to 3 bits) run in FP64
// and are added per tile in a fixed order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "bt_internal.hpp"
#include "bt_interp.hpp"

using bti::Plan;
using bti::PassResult;
using bti::QOp;

namespace bt {

constexpr int kMatEntries = bti::Interp::kPassMatEntries;
// accumulators of a pass: [0] norm^2, then either the histogram (16 bins) or
// the packed reduced rho (diagonal, then Re/Im of rho[a, b] for a < b: 16
// doubles over 2 bits, 64 over 3)
constexpr int kAcc = 65;
constexpr int kAccSmall = 17;  // enough unless a pass reduces rho over 3 bits
constexpr int kRB = 5;     // register bits (most): 32 amplitudes per thread
constexpr int kMaxPh = 28;

// 256 threads per workgroup; complex64 keeps 32 amplitudes per thread (T =
// 13), complex128 16 (T = 12): 64 KiB tiles, registers <= 256 per lane.
template <typename R>
struct TileCfg;
// WAVES: waves per SIMD the kernels are compiled for (launch-bounds
// minimum): complex64 runs two 64 KiB-tile workgroups per CU (registers
// <= 256), complex128 one (its tile + matrices exceed half the LDS).
template <>
struct TileCfg<float> {
  static constexpr int NT = 256;
  static constexpr int LOGNT = 8;
  static constexpr int RB = 5;
  static constexpr int WAVES = 2;
};
template <>
struct TileCfg<double> {
  static constexpr int NT = 256;
  static constexpr int LOGNT = 8;
  static constexpr int RB = 4;
  static constexpr int WAVES = 1;
};

// An operator of a pass.  k <= 2: applied in registers, `mask` = register
// positions, matrix index bit j <-> j-th lowest position.  k >= 3 (rare:
// 3-bit channels, doubly controlled gates, 2-qubit density-matrix
// superoperators): applied to the tile in LDS, matrix in operand order
// (operand 0 = MSB) at local offsets loff.
struct ROp {
  int32_t k;
  int32_t mask;
  int32_t moff;
  int32_t lsorted[4];  // k >= 3: ascending local bits
  uint32_t loff[16];   // k >= 3: local offset of each matrix index
  uint64_t gctrl;      // control bits outside the tile (global positions)
};

// LDS slots of a layout, precomputed per pass: the tile image is indexed by
// a swizzled local index and the swizzle is XOR-linear, so the slot of
// (thread base | register offset) is the XOR of the slots of its bits.
struct LSlots {
  uint16_t ts[8];    // slot of thread-index bit i (0 where the thread bit is unused)
  uint16_t rs[kRB];  // slot of register position j (0 for phantoms)
  uint16_t pad[3];
};

struct RPhase {
  int8_t rb[kRB];  // local tile bit held at register position j (-1: phantom)
  int8_t tmpl;     // > 0: the phase's operators match kPhaseTemplates[tmpl]
  int8_t pad[2];
  int32_t op0, nops;
};

// Register-mask sequences that start a phase and that the device applies as
// one straight-line block (no per-operator dispatch, whose branch chain
// copies the 2^RB live registers at every level); the phase's remaining
// operators follow one by one.  These are the openings layered brickworks
// produce most (two disjoint register pairs).  Phases with out-of-tile
// controls or LDS operators never match.
struct PhaseTemplate {
  int8_t n;
  int8_t mask[4];
};
constexpr PhaseTemplate kPhaseTemplates[] = {
    {0, {0, 0, 0, 0}},
    {2, {0x0c, 0x03, 0, 0}},
};
constexpr int kNumPhaseTemplates = (int)(sizeof(kPhaseTemplates) / sizeof(kPhaseTemplates[0]));

constexpr int kLayIn = kMaxPh, kLayOut = kMaxPh + 1;  // slot tables of the load / store layouts

struct PassGeom {
  int32_t T, nrb, tb, ncomp, write, need_norm, red, red_k, red_swap, nph, nops, nmats;
  int32_t hist_fast;         // RED_HIST with no histogram bit in the store layout's registers
  int32_t tile[bti::kMaxT];  // local bit -> global bit
  int32_t comp[40];          // global bits outside the tile, ascending
  int32_t nrun;              // tile index -> global base: runs of consecutive outside bits
  int8_t run_dst[bti::kMaxT + 1], run_len[bti::kMaxT + 1];
  int32_t red_bits[4];       // RED_HIST: global bits; RED_RHO: register positions (last phase), ascending
  int8_t io_in[kRB], io_out[kRB];
  int8_t pad[6];
  uint32_t red_loff[8];      // RED_RHO over 3 bits: local offset of each index (red_bits[0] = MSB)
  int32_t red_ls[4];         // RED_RHO over 3 bits: the three local bits, ascending
  LSlots lay[kMaxPh + 2];    // slot tables: phases, then kLayIn, kLayOut
  double scale;
  RPhase ph[kMaxPh];
  ROp ops[bti::kPMax];
};

// ---------------------------------------------------------------------------
// host/device: plan + queue slots -> register-tile geometry and matrices
// ---------------------------------------------------------------------------
__host__ __device__ inline int popc32(uint32_t x) {
  int c = 0;
  while (x) {
    x &= x - 1;
    c++;
  }
  return c;
}

// Layout with register positions 0..nrb-1 holding the given local bits
// (ascending) and the rest phantom.
__host__ __device__ inline void set_layout(int8_t* rb, const int* bits, int n) {
  for (int j = 0; j < kRB; j++) rb[j] = j < n ? (int8_t)bits[j] : (int8_t)-1;
}

// LDS element swizzle of local index l: XOR the low W bits with higher bits
// so strided layouts spread over the banks (W = 5 for 8-byte, 4 for 16-byte
// elements).  Linear over XOR.
template <typename R>
__host__ __device__ __forceinline__ uint32_t swz(uint32_t l) {
  constexpr int W = sizeof(C<R>) == 8 ? 5 : 4;
  return l ^ (((l >> W) ^ (l >> (2 * W))) & ((1u << W) - 1u));
}

template <typename R>
__host__ __device__ inline void make_slots(const int8_t* rb, int T, LSlots* s) {
  uint32_t regm = 0;
  for (int j = 0; j < kRB; j++) {
    if (rb[j] >= 0) regm |= 1u << rb[j];
    s->rs[j] = rb[j] >= 0 ? (uint16_t)swz<R>(1u << rb[j]) : (uint16_t)0;
  }
  int i = 0;
  for (int b = 0; b < T && i < 8; b++)
    if (!((regm >> b) & 1)) s->ts[i++] = (uint16_t)swz<R>(1u << b);
  for (; i < 8; i++) s->ts[i] = 0;
  s->pad[0] = s->pad[1] = s->pad[2] = 0;
}

template <typename R>
__host__ __device__ __attribute__((noinline)) int build_geom(const Plan& plan, const QOp* slots, int nbits, PassGeom* g,
                                                             C<R>* mats) {
  const int T = plan.ntile;
  g->T = T;
  constexpr int RB = TileCfg<R>::RB;
  g->nrb = T < RB ? T : RB;
  g->tb = T - g->nrb;
  for (int i = 0; i < T; i++) g->tile[i] = plan.tile[i];
  uint64_t tmask = 0;
  for (int i = 0; i < T; i++) tmask |= 1ull << plan.tile[i];
  g->ncomp = 0;
  for (int b = 0; b < nbits; b++)
    if (!((tmask >> b) & 1)) g->comp[g->ncomp++] = b;
  g->nrun = 0;
  for (int i = 0; i < g->ncomp; i++) {
    if (i == 0 || g->comp[i] != g->comp[i - 1] + 1) {
      if (g->nrun > bti::kMaxT) return BT_ERR_UNSUPPORTED;
      g->run_dst[g->nrun] = (int8_t)g->comp[i];
      g->run_len[g->nrun] = 0;
      g->nrun++;
    }
    g->run_len[g->nrun - 1]++;
  }
  int rank[64];
  for (int b = 0; b < 64; b++) rank[b] = -1;
  for (int i = 0; i < T; i++) rank[plan.tile[i]] = i;
  g->write = plan.write;
  g->need_norm = (plan.aligned || plan.red != bti::RED_NONE) ? 1 : 0;
  g->scale = plan.scale;
  g->red = plan.red;
  g->red_k = plan.red_k;
  g->red_swap = 0;
  const int nrb = g->nrb;
  // 1. operators in local bits; in-tile controls become target bits
  int lbits[bti::kPMax][kRB];
  int lk[bti::kPMax];
  for (int o = 0; o < plan.nops; o++) {
    const QOp& q = slots[plan.qidx[o]];
    int nb = 0;
    for (int b = 0; b < nbits; b++)
      if (((q.ctrl >> b) & 1) && rank[b] >= 0) {
        if (nb >= kRB) return BT_ERR_UNSUPPORTED;
        lbits[o][nb++] = rank[b];
      }
    for (int i = 0; i < q.k; i++) {
      if (nb >= kRB) return BT_ERR_UNSUPPORTED;
      lbits[o][nb++] = rank[q.bits[i]];
    }
    if (nb > 4) return BT_ERR_UNSUPPORTED;
    lk[o] = nb;
  }
  // 2. phases: sets of operators whose register bits fit together.
  //    Operators on disjoint qubits commute, so a phase may take any
  //    operator whose predecessors (earlier operators sharing a qubit) are
  //    already placed.  Greedy: each phase is grown from the seed that lets
  //    it take the most operators; within a phase the original order holds.
  const int nops = plan.nops;
  uint32_t opm[bti::kPMax];   // register bits an operator needs (k <= 2)
  uint32_t pred[bti::kPMax];  // operators that must run first
  for (int o = 0; o < nops; o++) {
    opm[o] = 0;
    if (lk[o] <= 2)
      for (int i = 0; i < lk[o]; i++) opm[o] |= 1u << lbits[o][i];
    const QOp& q = slots[plan.qidx[o]];
    uint64_t to = q.ctrl;
    for (int i = 0; i < q.k; i++) to |= 1ull << q.bits[i];
    pred[o] = 0;
    for (int a = 0; a < o; a++) {
      const QOp& qa = slots[plan.qidx[a]];
      uint64_t ta = qa.ctrl;
      for (int i = 0; i < qa.k; i++) ta |= 1ull << qa.bits[i];
      if (ta & to) pred[o] |= 1u << a;
    }
  }
  int nph = 0;
  uint32_t cur = 0;  // register-bit mask of the last phase
  int ph_of[bti::kPMax];
  int exec[bti::kPMax];  // execution order
  int ne = 0;
  uint32_t ph_mask[kMaxPh];
  uint32_t done = 0;
  const uint32_t all = nops >= 32 ? 0xffffffffu : (1u << nops) - 1u;
  while (done != all) {
    uint32_t best = 0, best_m = 0;
    int best_n = -1;
    for (int sd = 0; sd < nops; sd++) {
      if (((done >> sd) & 1) || (pred[sd] & ~done)) continue;
      uint32_t sel = 1u << sd, m = opm[sd];
      int n = 1;
      for (bool grew = true; grew;) {
        grew = false;
        for (int o = 0; o < nops; o++) {
          if (((done | sel) >> o) & 1) continue;
          if (pred[o] & ~(done | sel)) continue;
          if (popc32(m | opm[o]) > nrb) continue;
          sel |= 1u << o;
          m |= opm[o];
          n++;
          grew = true;
        }
      }
      if (n > best_n) {
        best_n = n;
        best = sel;
        best_m = m;
      }
    }
    if (best_n <= 0 || nph >= kMaxPh) return BT_ERR_UNSUPPORTED;
    for (int o = 0; o < nops; o++)
      if ((best >> o) & 1) {
        ph_of[o] = nph;
        exec[ne++] = o;
      }
    ph_mask[nph++] = best_m;
    cur = best_m;
    done |= best;
  }
  uint32_t redm = 0;
  if (plan.red == bti::RED_RHO && plan.red_k == 3) {
    // read from the tile image in LDS after the last phase: no register
    // positions needed, but the pass needs a phase to end
    if (nph == 0) ph_mask[nph++] = 0;
  } else if (plan.red == bti::RED_RHO) {
    for (int i = 0; i < plan.red_k; i++) redm |= 1u << rank[plan.red_bits[i]];
    if (nph == 0 || popc32(cur | redm) > nrb) {
      if (nph >= kMaxPh) return BT_ERR_UNSUPPORTED;
      ph_mask[nph] = 0;
      nph++;
      cur = 0;
    }
    cur |= redm;
    ph_mask[nph - 1] = cur;
  }
  // 3. fill every phase to nrb register bits with the highest free local bits
  //    (the low bits stay on the threads), positions = ascending local bits
  for (int p = 0; p < nph; p++) {
    uint32_t m = ph_mask[p];
    for (int b = T - 1; b >= 0 && popc32(m) < nrb; b--) m |= 1u << b;
    ph_mask[p] = m;
    int bits[kRB], n = 0;
    for (int b = 0; b < T; b++)
      if ((m >> b) & 1) bits[n++] = b;
    set_layout(g->ph[p].rb, bits, n);
    g->ph[p].op0 = 0;
    g->ph[p].nops = 0;
    // Register order: a phase with the same bits as the one before keeps its
    // order (no relayout between them).  Otherwise a relayout enters it
    // anyway, so the order is free: an opening pair of disjoint two-bit
    // operators goes to positions (2,3) then (0,1), the straight-line
    // template, and a lone opening pair to (1,2), which the per-operator
    // dispatch tests first.
    if (p > 0 && ph_mask[p] == ph_mask[p - 1]) {
      for (int j = 0; j < kRB; j++) g->ph[p].rb[j] = g->ph[p - 1].rb[j];
      continue;
    }
    int first[2], nf = 0;
    for (int e = 0; e < ne && nf < 2; e++)
      if (ph_of[exec[e]] == p) first[nf++] = exec[e];
    auto pair = [&](int o) { return lk[o] == 2 && popc32(opm[o]) == 2; };
    if (n < 4 || nf == 0 || !pair(first[0])) continue;
    int order[kRB], cnt = 0;
    uint32_t used = 0;
    auto put = [&](uint32_t m) {
      for (int b = 0; b < T; b++)
        if (((m >> b) & 1) && !((used >> b) & 1)) {
          order[cnt++] = b;
          used |= 1u << b;
        }
    };
    if (nf == 2 && pair(first[1]) && !(opm[first[0]] & opm[first[1]])) {
      put(opm[first[1]]);  // positions 0, 1
      put(opm[first[0]]);  // positions 2, 3
    } else {
      put(ph_mask[p] & ~opm[first[0]] & (0u - (ph_mask[p] & ~opm[first[0]])));  // lowest other bit: position 0
      put(opm[first[0]]);                                                       // positions 1, 2
    }
    put(ph_mask[p]);  // the rest, ascending
    if (cnt == n) set_layout(g->ph[p].rb, order, n);
  }
  g->nph = nph;
  // load/store layouts: the top local bits in registers, unless the first /
  // last phase's layout holds at most two of the lowest four local bits in
  // registers (the wavefront still reads whole 128 B lines within a few
  // instructions, and a relayout costs more than the extra requests)
  {
    int bits[kRB], n = 0;
    for (int b = T - nrb; b < T; b++) bits[n++] = b;
    set_layout(g->io_in, bits, n);
    set_layout(g->io_out, bits, n);
    const uint32_t low = (1u << (g->tb < 4 ? g->tb : 4)) - 1u;
    if (nph > 0 && popc32(ph_mask[0] & low) <= 2)
      for (int j = 0; j < kRB; j++) g->io_in[j] = g->ph[0].rb[j];
    if (nph > 0 && popc32(ph_mask[nph - 1] & low) <= 2)
      for (int j = 0; j < kRB; j++) g->io_out[j] = g->ph[nph - 1].rb[j];
    // a histogram is cheapest when its bits stay off the store layout's
    // registers: each thread's 2^RB amplitudes then share one bin per tile
    g->hist_fast = 0;
    if (plan.red == bti::RED_HIST) {
      uint32_t lm = 0;
      for (int i = 0; i < plan.red_k; i++)
        if (rank[plan.red_bits[i]] >= 0) lm |= 1u << rank[plan.red_bits[i]];
      uint32_t om = 0;
      for (int j = 0; j < kRB; j++)
        if (g->io_out[j] >= 0) om |= 1u << g->io_out[j];
      if (om & lm) {
        int hb[kRB], m = 0;
        for (int b = T - 1; b >= 0 && m < nrb; b--)
          if (!((lm >> b) & 1)) hb[m++] = b;
        if (m == nrb) {
          for (int i = 0; i < m / 2; i++) {
            const int t = hb[i];
            hb[i] = hb[m - 1 - i];
            hb[m - 1 - i] = t;
          }
          set_layout(g->io_out, hb, m);
          om = 0;
          for (int j = 0; j < kRB; j++)
            if (g->io_out[j] >= 0) om |= 1u << g->io_out[j];
        }
      }
      g->hist_fast = (om & lm) ? 0 : 1;
    }
  }
  // 4. operators in execution order, in register positions, matrices
  //    permuted to the ascending position convention (index bit j <-> j-th
  //    lowest position)
  int moff = 0;
  g->nops = plan.nops;
  for (int e = 0; e < plan.nops; e++) {
    const int o = exec[e];
    const QOp& q = slots[plan.qidx[o]];
    const RPhase& ph = g->ph[ph_of[o]];
    if (ph.nops == 0) g->ph[ph_of[o]].op0 = e;
    g->ph[ph_of[o]].nops++;
    ROp& r = g->ops[e];
    const int k2 = lk[o], D2 = 1 << k2;
    r.k = k2;
    int pos[kRB];  // register position of lbits[o][i] (k <= 2), else the operand order
    r.mask = 0;
    if (k2 <= 2) {
      for (int i = 0; i < k2; i++)
        for (int j = 0; j < kRB; j++)
          if (ph.rb[j] == lbits[o][i]) pos[i] = j;
      for (int i = 0; i < k2; i++) r.mask |= 1 << pos[i];
    } else {
      for (int i = 0; i < k2; i++) pos[i] = k2 - 1 - i;  // identity: operand order, operand 0 = MSB
      for (int i = 0; i < k2; i++) r.lsorted[i] = lbits[o][i];
      for (int i = 0; i < k2; i++)
        for (int j = i + 1; j < k2; j++)
          if (r.lsorted[j] < r.lsorted[i]) {
            int t = r.lsorted[i];
            r.lsorted[i] = r.lsorted[j];
            r.lsorted[j] = t;
          }
      for (int l = 0; l < D2; l++) {
        uint32_t off = 0;
        for (int i = 0; i < k2; i++)
          if ((l >> (k2 - 1 - i)) & 1) off |= 1u << lbits[o][i];
        r.loff[l] = off;
      }
    }
    r.gctrl = 0;
    for (int b = 0; b < nbits; b++)
      if (((q.ctrl >> b) & 1) && rank[b] < 0) r.gctrl |= 1ull << b;
    // rank of each operand bit among the ascending positions
    int asc[kRB];
    for (int i = 0; i < k2; i++) {
      int c = 0;
      for (int j = 0; j < k2; j++) c += pos[j] < pos[i];
      asc[i] = c;
    }
    const int nc = k2 - q.k;  // leading operand bits are in-tile controls
    const int D = 1 << q.k;
    if (moff + D2 * D2 > kMatEntries) return BT_ERR_UNSUPPORTED;
    r.moff = moff;
    for (int nr = 0; nr < D2; nr++)
      for (int ncl = 0; ncl < D2; ncl++) {
        // operand-order indices (operand 0 = MSB) of the new indices
        int orr = 0, occ = 0;
        for (int i = 0; i < k2; i++) {
          orr |= ((nr >> asc[i]) & 1) << (k2 - 1 - i);
          occ |= ((ncl >> asc[i]) & 1) << (k2 - 1 - i);
        }
        const int cr = orr >> q.k, cc = occ >> q.k;
        const int tr = orr & (D - 1), tc = occ & (D - 1);
        double vr = 0, vi = 0;
        if (cr == cc) {
          if (cr == (1 << nc) - 1) {
            vr = q.m[2 * (tr + tc * D)];
            vi = q.m[2 * (tr + tc * D) + 1];
          } else if (tr == tc) {
            vr = 1.0;
          }
        }
        mats[moff + nr + ncl * D2].x = (R)vr;
        mats[moff + nr + ncl * D2].y = (R)vi;
      }
    moff += D2 * D2;
  }
  g->nmats = moff;
  // straight-line phase templates
  for (int p = 0; p < nph; p++) {
    RPhase& ph = g->ph[p];
    ph.tmpl = 0;
    bool plain = ph.nops > 0;
    for (int e = ph.op0; e < ph.op0 + ph.nops && plain; e++)
      plain = g->ops[e].k <= 2 && g->ops[e].gctrl == 0;
    if (!plain) continue;
    for (int t = 1; t < kNumPhaseTemplates; t++) {
      const PhaseTemplate& pt = kPhaseTemplates[t];
      bool ok = pt.n <= ph.nops;  // a template covers the phase's first pt.n operators
      for (int i = 0; i < pt.n && ok; i++) {
        ok = g->ops[ph.op0 + i].mask == pt.mask[i];
        if (TileCfg<R>::RB < 5) ok = ok && pt.mask[i] < 16;
      }
      if (ok) {
        ph.tmpl = (int8_t)t;
        break;
      }
    }
  }
  // the input scale rides on the first operator that every amplitude of
  // every tile goes through (no out-of-tile control), sparing the per-tile
  // multiply of each register
  for (int e = 0; e < plan.nops && g->scale != 1.0; e++) {
    const ROp& r = g->ops[e];
    if (r.gctrl) continue;
    const int D2 = 1 << r.k;
    for (int i = 0; i < D2 * D2; i++) {
      mats[r.moff + i].x = (R)((double)mats[r.moff + i].x * g->scale);
      mats[r.moff + i].y = (R)((double)mats[r.moff + i].y * g->scale);
    }
    g->scale = 1.0;
  }
  for (int p = 0; p < nph; p++) make_slots<R>(g->ph[p].rb, T, &g->lay[p]);
  make_slots<R>(g->io_in, T, &g->lay[kLayIn]);
  make_slots<R>(g->io_out, T, &g->lay[kLayOut]);
  // 5. reductions
  if (plan.red == bti::RED_HIST) {
    for (int i = 0; i < plan.red_k; i++) g->red_bits[i] = plan.red_bits[i];
  } else if (plan.red == bti::RED_RHO && plan.red_k == 3) {
    for (int a = 0; a < 8; a++) {
      uint32_t off = 0;
      for (int i = 0; i < 3; i++)
        if ((a >> (2 - i)) & 1) off |= 1u << rank[plan.red_bits[i]];
      g->red_loff[a] = off;
    }
    int ls[3];
    for (int i = 0; i < 3; i++) ls[i] = rank[plan.red_bits[i]];
    for (int i = 0; i < 3; i++)
      for (int j = i + 1; j < 3; j++)
        if (ls[j] < ls[i]) {
          const int t = ls[i];
          ls[i] = ls[j];
          ls[j] = t;
        }
    for (int i = 0; i < 3; i++) g->red_ls[i] = ls[i];
    g->red_ls[3] = 0;
  } else if (plan.red == bti::RED_RHO) {
    const RPhase& ph = g->ph[nph - 1];
    int p[2] = {0, 0};
    for (int i = 0; i < plan.red_k; i++)
      for (int j = 0; j < kRB; j++)
        if (ph.rb[j] == rank[plan.red_bits[i]]) p[i] = j;
    if (plan.red_k == 1) {
      g->red_bits[0] = p[0];
    } else {
      g->red_bits[0] = p[0] < p[1] ? p[0] : p[1];
      g->red_bits[1] = p[0] < p[1] ? p[1] : p[0];
      // packed rho is in ascending-position index order (bit 0 = lower
      // position); the plan wants red_bits[0] as the MSB
      g->red_swap = p[0] < p[1] ? 1 : 0;
    }
  }
  return BT_OK;
}

// ---------------------------------------------------------------------------
// device side
// ---------------------------------------------------------------------------
template <typename R, int MASK>
__host__ __device__ __forceinline__ void apply_reg(C<R>* v, const C<R>* mats) {
  constexpr int K = __builtin_popcount(MASK);
  constexpr int D = 1 << K;
  constexpr int NR = 1 << TileCfg<R>::RB;
  int P[K > 0 ? K : 1];
  {
    int c = 0;
#pragma unroll
    for (int j = 0; j < kRB; j++)
      if ((MASK >> j) & 1) P[c++] = j;
  }
#pragma unroll
  for (int base = 0; base < NR; base++) {
    if (base & MASK) continue;
    C<R> x[D];
#pragma unroll
    for (int l = 0; l < D; l++) {
      int idx = base;
#pragma unroll
      for (int j = 0; j < K; j++)
        if ((l >> j) & 1) idx |= 1 << P[j];
      x[l] = v[idx];
    }
#pragma unroll
    for (int r = 0; r < D; r++) {
      C<R> acc = cmake<C<R>>(0, 0);
#pragma unroll
      for (int c = 0; c < D; c++) acc = cfma(acc, mats[r + c * D], x[c]);
      int idx = base;
#pragma unroll
      for (int j = 0; j < K; j++)
        if ((r >> j) & 1) idx |= 1 << P[j];
      v[idx] = acc;
    }
  }
}

#if defined(__HIP_DEVICE_COMPILE__)
typedef float floatx4_t __attribute__((ext_vector_type(4)));

// The same operator on a complex64 tile, on the FP32 matrix cores.
// v_mfma_f32_4x4x1f32 is 16 blocks of 4 x 4 x 1 products: lane l (block
// b = l / 4) gets D[i] += A(lane 4b + i) * B(lane l), i = 0..3.  With B = one
// real input of lane l's own group and A(lane 4b + i) = the operator's real
// form at output row i of a 4-row slice, every lane computes 4 real outputs
// of its own group from its own inputs: nothing moves between lanes, and the
// 2^K x 2^K complex product (2^(K+1) real inputs) costs 2^(K+1) instructions
// per 4 outputs, at twice the FP32 rate of the vector FMAs and beside them.
// Output slice h holds rows 2h, 2h + 1 of the complex matrix as
// (Re y_2h, Im y_2h, Re y_2h+1, Im y_2h+1); the lane with l % 4 = i holds
// the slice's row 2h + (i >> 1), real part if i is even, else imaginary,
// against the inputs (Re x_c, Im x_c): (Mr, -Mi) for a real part, (Mi, Mr)
// for an imaginary one.
template <int MASK>
__device__ __forceinline__ void apply_reg_mfma(C<float>* v, const C<float>* mats) {
  constexpr int K = __builtin_popcount(MASK);
  static_assert(K == 1 || K == 2, "matrix-core register operators cover one and two bits");
  constexpr int D = 1 << K, NS = D / 2;
  constexpr int NR = 1 << kRB;
  int P[K];
  {
    int c = 0;
#pragma unroll
    for (int j = 0; j < kRB; j++)
      if ((MASK >> j) & 1) P[c++] = j;
  }
  const int i = threadIdx.x & 3;
  const bool imag = (i & 1) != 0;
  float a[NS][2 * D];
#pragma unroll
  for (int h = 0; h < NS; h++) {
    const int r = 2 * h + (i >> 1);
#pragma unroll
    for (int c = 0; c < D; c++) {
      const C<float> m = mats[r + c * D];
      a[h][2 * c] = imag ? m.y : m.x;
      a[h][2 * c + 1] = imag ? m.x : -m.y;
    }
  }
#pragma unroll
  for (int base = 0; base < NR; base++) {
    if (base & MASK) continue;
    C<float> x[D];
    int idx[D];
#pragma unroll
    for (int l = 0; l < D; l++) {
      idx[l] = base;
#pragma unroll
      for (int j = 0; j < K; j++)
        if ((l >> j) & 1) idx[l] |= 1 << P[j];
      x[l] = v[idx[l]];
    }
#pragma unroll
    for (int h = 0; h < NS; h++) {
      floatx4_t d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < D; c++) {
        d = __builtin_amdgcn_mfma_f32_4x4x1f32(a[h][2 * c], x[c].x, d, 0, 0, 0);
        d = __builtin_amdgcn_mfma_f32_4x4x1f32(a[h][2 * c + 1], x[c].y, d, 0, 0, 0);
      }
      v[idx[2 * h]] = cmake<C<float>>(d[0], d[1]);
      v[idx[2 * h + 1]] = cmake<C<float>>(d[2], d[3]);
    }
  }
}
#endif

// Register operator with the mask a compile-time constant: complex64 on the
// device runs on the matrix cores, everything else (complex128, the host
// emulation) as vector FMAs.
template <typename R, int MASK>
__host__ __device__ __forceinline__ void apply_regk(C<R>* v, const C<R>* m) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(R) == 4) {
    apply_reg_mfma<MASK>(v, m);
    return;
  }
#endif
  apply_reg<R, MASK>(v, m);
}

template <typename R>
__host__ __device__ __forceinline__ void apply_mask(int mask, C<R>* v, const C<R>* m) {
  // adjacent register pairs are the common case (brickwork pairs after the
  // phase planner): test them first, every level of the dispatch costs moves
  if (__builtin_expect(mask == 0x06, 1)) {
    apply_regk<R, 0x06>(v, m);
    return;
  }
  if constexpr (TileCfg<R>::RB == 5) {
    if (__builtin_expect(mask == 0x18, 1)) {
      apply_regk<R, 0x18>(v, m);
      return;
    }
  }
  switch (mask) {
#define BT_CASE(M) \
  case M: apply_regk<R, M>(v, m); break;
#define BT_CASE5(M)                                   \
  case M:                                             \
    if constexpr (TileCfg<R>::RB == 5) apply_regk<R, M>(v, m); \
    break;
    BT_CASE(1) BT_CASE(2) BT_CASE(4) BT_CASE(8) BT_CASE(3) BT_CASE(5) BT_CASE(6) BT_CASE(9) BT_CASE(10) BT_CASE(12)
    BT_CASE5(16) BT_CASE5(17) BT_CASE5(18) BT_CASE5(20) BT_CASE5(24)
#undef BT_CASE
#undef BT_CASE5
    default: break;
  }
}

// A matched phase template: the operators' register applications in order,
// straight-line (the masks are compile-time constants of each case).
// Keeps the next operator's matrix loads from being hoisted above this
// operator (all of a template's matrices live at once would spill).
__host__ __device__ __forceinline__ void op_fence() {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" ::: "memory");
#endif
}

// A matched phase template: the operators' register applications in order,
// straight-line (the masks are compile-time constants of each case).
template <typename R>
__host__ __device__ __forceinline__ void apply_template(int tm, C<R>* v, const C<R>* mats, const ROp* ops) {
#define BT_A(M, I)                              \
  apply_regk<R, M>(v, mats + ops[I].moff); \
  op_fence();
  if (tm == 1) {
    BT_A(0x0c, 0)
    BT_A(0x03, 1)
  }
#undef BT_A
}

// reduced rho over register positions P0 < P1 (K = 1 or 2), packed as
// diagonal then (a < b) Re, Im of rho[a, b]; index bit j <-> P_j
template <typename R, int MASK>
__host__ __device__ __forceinline__ void rho_reg(const C<R>* v, int nvalid_mask, double* racc) {
  constexpr int K = __builtin_popcount(MASK);
  constexpr int D = 1 << K;
  constexpr int NR = 1 << TileCfg<R>::RB;
  int P[K];
  {
    int c = 0;
#pragma unroll
    for (int j = 0; j < kRB; j++)
      if ((MASK >> j) & 1) P[c++] = j;
  }
#pragma unroll
  for (int base = 0; base < NR; base++) {
    if (base & MASK) continue;
    if (base & nvalid_mask) continue;  // phantom registers
    double vr[D], vi[D];
#pragma unroll
    for (int l = 0; l < D; l++) {
      int idx = base;
#pragma unroll
      for (int j = 0; j < K; j++)
        if ((l >> j) & 1) idx |= 1 << P[j];
      vr[l] = (double)v[idx].x;
      vi[l] = (double)v[idx].y;
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

template <typename R>
__host__ __device__ __forceinline__ void rho_mask(int mask, const C<R>* v, int nvalid_mask, double* racc) {
  switch (mask) {
#define BT_CASE(M) \
  case M: rho_reg<R, M>(v, nvalid_mask, racc); break;
#define BT_CASE5(M)                                                \
  case M:                                                          \
    if constexpr (TileCfg<R>::RB == 5) rho_reg<R, M>(v, nvalid_mask, racc); \
    break;
    BT_CASE(1) BT_CASE(2) BT_CASE(4) BT_CASE(8) BT_CASE(3) BT_CASE(5) BT_CASE(6) BT_CASE(9) BT_CASE(10) BT_CASE(12)
    BT_CASE5(16) BT_CASE5(17) BT_CASE5(18) BT_CASE5(20) BT_CASE5(24)
#undef BT_CASE
#undef BT_CASE5
    default: break;
  }
}


__host__ __device__ __forceinline__ uint32_t ins0(uint32_t x, int b) { return ((x >> b) << (b + 1)) | (x & ((1u << b) - 1u)); }

// k >= 3 operator on the tile in LDS (swizzled), thread tid of the block
template <typename R, int K>
__host__ __device__ __forceinline__ void apply_lds(C<R>* lds, int T, const ROp& op, const C<R>* mats, uint32_t tid) {
  constexpr int NT = TileCfg<R>::NT;
  constexpr int D = 1 << K;
  const uint32_t ng = 1u << (T - K);
  for (uint32_t g = tid; g < ng; g += NT) {
    uint32_t lb = g;
#pragma unroll
    for (int i = 0; i < K; i++) lb = ins0(lb, op.lsorted[i]);
    C<R> x[D];
#pragma unroll
    for (int l = 0; l < D; l++) x[l] = lds[swz<R>(lb + op.loff[l])];
#pragma unroll 1
    for (int r = 0; r < D; r++) {  // rows one at a time: the D*D entries stay in LDS
      C<R> acc = cmake<C<R>>(0, 0);
#pragma unroll
      for (int c = 0; c < D; c++) acc = cfma(acc, mats[r + c * D], x[c]);
      lds[swz<R>(lb + op.loff[r])] = acc;
    }
  }
}

// Reduced rho over three tile bits, read from the tile image in LDS
// (swizzled): thread tid takes packed entry e = tid / 4 (the 8 diagonal
// entries, then the 28 pairs a < b in order (0,1), (0,2), ..., (6,7)) and the
// index groups tid % 4, tid % 4 + 4, ...; returns e (-1: no entry) and the
// thread's partial sum of Re (xr) and Im (xi) of rho[a, b].
template <typename R>
__host__ __device__ __forceinline__ int rho3_partial(const C<R>* lds, const PassGeom& g, uint32_t tid, double* xr,
                                                     double* xi) {
  const int e = (int)(tid >> 2), s = (int)(tid & 3);
  *xr = 0.0;
  *xi = 0.0;
  if (e >= 36) return -1;
  int a = e, b = e;
  if (e >= 8) {
    int p = e - 8;
    a = 0;
    while (p >= 7 - a) {
      p -= 7 - a;
      a++;
    }
    b = a + 1 + p;
  }
  const uint32_t oa = g.red_loff[a], ob = g.red_loff[b];
  const uint32_t ng = 1u << (g.T - 3);
  double sr = 0.0, si = 0.0;
  for (uint32_t grp = (uint32_t)s; grp < ng; grp += 4) {
    uint32_t lb = grp;
#pragma unroll
    for (int i = 0; i < 3; i++) lb = ((lb >> g.red_ls[i]) << (g.red_ls[i] + 1)) | (lb & ((1u << g.red_ls[i]) - 1u));
    const C<R> va = lds[swz<R>(lb + oa)], vb = lds[swz<R>(lb + ob)];
    sr += (double)va.x * (double)vb.x + (double)va.y * (double)vb.y;
    si += (double)va.y * (double)vb.x - (double)va.x * (double)vb.y;
  }
  *xr = sr;
  *xi = si;
  return e;
}

// packed index of entry e's real part (its imaginary part follows for e >= 8)
__host__ __device__ __forceinline__ int rho3_slot(int e) { return e < 8 ? e : 8 + 2 * (e - 8); }

#if defined(__HIPCC__)
// A double moved across lanes by a DPP pattern (both 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  // every lane reads a valid source under these permutations (full row and
  // bank masks), so no "old" value has to be materialised first
  const int lo = __builtin_amdgcn_mov_dpp((int)(uint32_t)u, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)u, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the 64 lanes of a wavefront in a fixed order (xor-1 and xor-2
// within quads, then the half-row and row mirrors, then the four rows in
// order); the same value in every lane.
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_f64<0xb1>(x);   // quad_perm [1,0,3,2]
  x += dpp_f64<0x4e>(x);   // quad_perm [2,3,0,1]
  x += dpp_f64<0x141>(x);  // row_half_mirror
  x += dpp_f64<0x140>(x);  // row_mirror
  return (readlane_f64(x, 0) + readlane_f64(x, 16)) + (readlane_f64(x, 32) + readlane_f64(x, 48));
}

// The quad's four partials of a 3-bit reduced-rho entry, summed in a fixed
// order ((s0 + s1) + (s2 + s3)), added by the quad's first lane to its
// wavefront's packed-rho slots (wrho: LDS; each entry has one writer).
template <typename R>
__device__ __forceinline__ void rho3_accumulate_t(const C<R>* lds, const PassGeom& g, double* wrho) {
  double xr, xi;
  const int e = rho3_partial<R>(lds, g, threadIdx.x, &xr, &xi);
  xr += dpp_f64<0xb1>(xr);
  xr += dpp_f64<0x4e>(xr);
  xi += dpp_f64<0xb1>(xi);
  xi += dpp_f64<0x4e>(xi);
  if (e >= 0 && (threadIdx.x & 3) == 0) {
    const int sl = rho3_slot(e);
    wrho[sl] += xr;
    if (e >= 8) wrho[sl + 1] += xi;
  }
}

// Add M per-thread values into this wavefront's accumulator slots
// (wred[0..M), LDS): one writer lane, tiles in order, no barrier.
template <int M>
__device__ __forceinline__ void wave_accumulate(const double* v, double* wred) {
  const bool lead = (threadIdx.x & 63) == 0;
#pragma unroll
  for (int j = 0; j < M; j++) {
    const double t = wave_sum(v[j]);
    if (lead) wred[j] += t;
  }
}
#endif

// Local index of thread t's register r under a layout: thread bits are the
// local bits not in registers, ascending, holding t's bits in order.
__host__ __device__ __forceinline__ uint32_t thread_base(const int8_t* rb, int T, uint32_t t) {
  uint32_t regm = 0;
#pragma unroll
  for (int j = 0; j < kRB; j++)
    if (rb[j] >= 0) regm |= 1u << rb[j];
  uint32_t base = 0;
  int i = 0;
  for (int b = 0; b < T; b++)
    if (!((regm >> b) & 1)) {
      base |= ((t >> i) & 1u) << b;
      i++;
    }
  return base;
}

// A value the whole workgroup holds in common, moved to a scalar register
// on the device (identity on the host).
__host__ __device__ __forceinline__ uint32_t uni32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_readfirstlane(x);
#else
  return x;
#endif
}

// Slot of thread tid's register 0 under a layout (XOR of its bits' slots).
__host__ __device__ __forceinline__ uint32_t slot_base(const LSlots& s, uint32_t tid) {
  uint32_t b = 0;
#pragma unroll
  for (int i = 0; i < 8; i++)
    if ((tid >> i) & 1) b ^= uni32(s.ts[i]);
  return b;
}

// XOR of the register-position slots of register r (uniform across the workgroup)
__host__ __device__ __forceinline__ uint32_t reg_slot(const LSlots& s, int r) {
  uint32_t u = 0;
#pragma unroll
  for (int j = 0; j < kRB; j++)
    if ((r >> j) & 1) u ^= uni32(s.rs[j]);
  return u;
}

template <typename R>
__host__ __device__ __forceinline__ void regs_to_lds(const C<R>* v, C<R>* lds, const LSlots& S, uint32_t tid,
                                                     bool active, int nvalid_mask) {
  constexpr int NR = 1 << TileCfg<R>::RB;
  const uint32_t s0 = slot_base(S, tid);
  if (active && nvalid_mask == 0) {
#pragma unroll
    for (int r = 0; r < NR; r++) lds[s0 ^ reg_slot(S, r)] = v[r];
  } else {
#pragma unroll
    for (int r = 0; r < NR; r++)
      if (active && !(r & nvalid_mask)) lds[s0 ^ reg_slot(S, r)] = v[r];
  }
}

// Reads are unconditional (straight-line, so they issue back to back):
// inactive threads pass a mirrored tid < 2^tb and phantom registers alias a
// valid one; such values are never stored or reduced.
template <typename R>
__host__ __device__ __forceinline__ void lds_to_regs(C<R>* v, const C<R>* lds, const LSlots& S, uint32_t ltid) {
  constexpr int NR = 1 << TileCfg<R>::RB;
  const uint32_t s0 = slot_base(S, ltid);
#pragma unroll
  for (int r = 0; r < NR; r++) v[r] = lds[s0 ^ reg_slot(S, r)];
}

// Per-thread constants of a pass: offsets (in amplitudes, from the tile's
// first amplitude) of the load / store layouts.  O = uint32_t when every
// tile bit is below 32 (states of up to 2^32 amplitudes), else uint64_t.
// Threads >= 2^tb are inactive; their loads mirror thread tid mod 2^tb, so
// every address is inside the state.
template <typename O>
struct ThreadIO {
  O gin[kRB], gout[kRB];
  O gthr_in, gthr_out;
  uint32_t tid, ltid;
  bool active;
  int nvalid_mask;  // register bits beyond nrb are phantom
};

template <typename R, typename O>
__host__ __device__ __forceinline__ void thread_io(const PassGeom& g, uint32_t tid, ThreadIO<O>& io) {
  constexpr int NR = 1 << TileCfg<R>::RB;
  io.tid = tid;
  io.active = tid < (1u << g.tb);
  io.ltid = tid & ((1u << g.tb) - 1u);
  io.nvalid_mask = (NR - 1) & ~((1 << g.nrb) - 1);
#pragma unroll
  for (int j = 0; j < kRB; j++) {
    io.gin[j] = g.io_in[j] >= 0 ? (O)1 << g.tile[g.io_in[j]] : 0;
    io.gout[j] = g.io_out[j] >= 0 ? (O)1 << g.tile[g.io_out[j]] : 0;
  }
  io.gthr_in = 0;
  io.gthr_out = 0;
  const uint32_t bi = thread_base(g.io_in, g.T, io.ltid), bo = thread_base(g.io_out, g.T, io.ltid);
  for (int b = 0; b < g.T; b++) {
    if ((bi >> b) & 1) io.gthr_in |= (O)1 << g.tile[b];
    if ((bo >> b) & 1) io.gthr_out |= (O)1 << g.tile[b];
  }
}

// Hide a value's origin from the optimiser (no instruction is emitted), so
// the per-register offsets built from it are recomputed per tile instead of
// being hoisted out of the tile loop as 64 live registers.
template <typename T>
__host__ __device__ __forceinline__ void opaque(T& x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x));
#else
  (void)x;
#endif
}

// The same for a value that is equal in every lane, kept in a scalar
// register (moved there first: a lane-uniform value the compiler cannot
// prove uniform would otherwise occupy a vector register per lane).
template <typename T>
__host__ __device__ __forceinline__ void opaque_uniform(T& x) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = (uint64_t)x;
    x = (T)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(u >> 32)) << 32) |
            (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u));
  } else {
    x = (T)__builtin_amdgcn_readfirstlane((int)x);
  }
  asm volatile("" : "+s"(x));
#else
  (void)x;
#endif
}

// gin / gout are the same in every lane (pass geometry); gthr_* are per thread.
template <typename O>
__host__ __device__ __forceinline__ void opaque_io(ThreadIO<O>& io) {
#pragma unroll
  for (int j = 0; j < kRB; j++) {
    opaque_uniform(io.gin[j]);
    opaque_uniform(io.gout[j]);
  }
  opaque(io.gthr_in);
  opaque(io.gthr_out);
}

// State loads and stores on the device go through the global address space
// (global_load / global_store, counted by vmcnt alone).  Through a generic
// pointer they compile to flat_* accesses, which also count against lgkmcnt,
// so every LDS or scalar-load wait in the pass would also wait for the
// in-flight prefetch of the next tile and the stores of the last one.
template <typename R>
__host__ __device__ __forceinline__ C<R> gload(const C<R>* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef R v2 __attribute__((ext_vector_type(2)));  // a builtin vector: no struct copy through a generic pointer
  const v2 x = *(const __attribute__((address_space(1))) v2*)p;
  return cmake<C<R>>(x.x, x.y);
#else
  return *p;
#endif
}
template <typename R>
__host__ __device__ __forceinline__ void gstore(C<R>* p, const C<R>& v) {
#if defined(__HIP_DEVICE_COMPILE__)
  typedef R v2 __attribute__((ext_vector_type(2)));
  v2 x;
  x.x = v.x;
  x.y = v.y;
  *(__attribute__((address_space(1))) v2*)p = x;
#else
  *p = v;
#endif
}

// A flag that is the same in every lane (a pass-geometry field), moved to a
// scalar register so branches on it are scalar branches.
__host__ __device__ __forceinline__ bool uflag(bool x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_readfirstlane((int)x) != 0;
#else
  return x;
#endif
}

// Wait until every vector-memory access this wave issued has completed
// (s_waitcnt vmcnt(0), the LDS and export counters left alone).
__host__ __device__ __forceinline__ void vm_wait_all() {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_s_waitcnt(0x0F70);
#endif
}

// global index of tile t's first amplitude
__host__ __device__ __forceinline__ uint64_t tile_base(const PassGeom& g, uint64_t t) {
  uint64_t base = 0;
  for (int i = 0; i < g.nrun; i++) {
    const int len = g.run_len[i];
    base |= (t & ((1ull << len) - 1ull)) << g.run_dst[i];
    t >>= len;
  }
  return base;
}

// Issue the loads of a tile's registers (io_in layout) into raw[].  The
// loads are unconditional and straight-line (phantom registers alias a
// valid address) so they are all in flight at once; the device loop issues
// them for the next tile before it computes on the current one.
template <typename R, typename O>
__host__ __device__ __forceinline__ void fetch_regs(C<R>* raw, const C<R>* __restrict__ state, const ThreadIO<O>& io,
                                                    uint64_t base) {
  constexpr int NR = 1 << TileCfg<R>::RB;
  const C<R>* __restrict__ tp = state + base;
#pragma unroll
  for (int r = 0; r < NR; r++) {
    O off = io.gthr_in;
#pragma unroll
    for (int j = 0; j < kRB; j++)
      if ((r >> j) & 1) off |= io.gin[j];
    raw[r] = gload<R>(tp + off);
  }
}

// Registers of a tile from its fetched values: scaled, and zero where the
// register is phantom or the thread inactive.  fresh != 0: the state is
// |0...0> and nothing was fetched.
template <typename R, typename O>
__host__ __device__ __forceinline__ void take_regs(C<R>* v, const C<R>* raw, const ThreadIO<O>& io, uint64_t base,
                                                   R scale, int fresh) {
  constexpr int NR = 1 << TileCfg<R>::RB;
  if (!fresh && io.active && io.nvalid_mask == 0) {  // every register real: no masks
#pragma unroll
    for (int r = 0; r < NR; r++) {
      v[r].x = raw[r].x * scale;
      v[r].y = raw[r].y * scale;
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < NR; r++) {
    const bool ok = io.active && !(r & io.nvalid_mask);
    if (fresh) {
      O off = io.gthr_in;
#pragma unroll
      for (int j = 0; j < kRB; j++)
        if ((r >> j) & 1) off |= io.gin[j];
      v[r] = cmake<C<R>>(ok && base == 0 && off == 0 ? scale : (R)0, (R)0);
    } else {
      v[r].x = ok ? raw[r].x * scale : (R)0;
      v[r].y = ok ? raw[r].y * scale : (R)0;
    }
  }
}

// Store a tile's registers (io_out layout) and accumulate FP64 norm^2 into
// acc[0] and the histogram over g.red_bits into acc[1..2^red_k].
template <typename R, typename O>
__host__ __device__ __forceinline__ void store_regs(const C<R>* v, C<R>* __restrict__ state, const PassGeom& g,
                                                    const ThreadIO<O>& io, uint64_t base, double* acc) {
  constexpr int NR = 1 << TileCfg<R>::RB;
  C<R>* __restrict__ tp = state + base;
  const bool slow_hist = uflag(g.red == bti::RED_HIST && !g.hist_fast);
  const bool write = uflag(g.write), need_norm = uflag(g.need_norm);
  double s = 0.0;
  if (io.active && io.nvalid_mask == 0 && !slow_hist) {  // every register real: no masks
    if (write) {
#pragma unroll
      for (int r = 0; r < NR; r++) {
        O off = io.gthr_out;
#pragma unroll
        for (int j = 0; j < kRB; j++)
          if ((r >> j) & 1) off |= io.gout[j];
        gstore<R>(tp + off, v[r]);
      }
    }
    if (need_norm) {
#pragma unroll
      for (int r = 0; r < NR; r++) {
        const double w = (double)v[r].x * (double)v[r].x + (double)v[r].y * (double)v[r].y;
        s += w;
      }
    }
  } else {
#pragma unroll
  for (int r = 0; r < NR; r++) {
    O off = io.gthr_out;
#pragma unroll
    for (int j = 0; j < kRB; j++)
      if ((r >> j) & 1) off |= io.gout[j];
    if (io.active && !(r & io.nvalid_mask)) {
      if (g.write) gstore<R>(tp + off, v[r]);
      if (g.need_norm) {
        const double w = (double)v[r].x * (double)v[r].x + (double)v[r].y * (double)v[r].y;
        s += w;
        if (slow_hist) {
          int bin = 0;
          const uint64_t gi = base | (uint64_t)off;
          for (int j = 0; j < g.red_k; j++) bin |= (int)((gi >> g.red_bits[j]) & 1ull) << j;
#pragma unroll
          for (int b = 0; b < 16; b++) acc[1 + b] += (bin == b) ? w : 0.0;
        }
      }
    }
  }
  }
  acc[0] += s;
  if (g.red == bti::RED_HIST && g.hist_fast) {  // one bin for all of this thread's amplitudes
    const uint64_t gi = base | (uint64_t)io.gthr_out;
    int bin = 0;
    for (int j = 0; j < g.red_k; j++) bin |= (int)((gi >> g.red_bits[j]) & 1ull) << j;
#pragma unroll
    for (int b = 0; b < 16; b++) acc[1 + b] += (bin == b) ? s : 0.0;
  }
}

__host__ __device__ __forceinline__ int rho_reg_mask(const PassGeom& g) {
  return (1 << g.red_bits[0]) | (g.red_k > 1 ? 1 << g.red_bits[1] : 0);
}

#if defined(__HIPCC__)
// A value the whole workgroup holds in common, moved to scalar registers.
__device__ __forceinline__ int uniform_i32(int x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ uint64_t uniform_u64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Copy a pass geometry into LDS: the header, the used phases and operators.
__device__ __forceinline__ void load_geom(const PassGeom* __restrict__ src, PassGeom* dst, int tid, int nt) {
  constexpr int hdr = (int)(offsetof(PassGeom, ph) / 4);
  const int nph = src->nph, nops = src->nops;
  const int* s = (const int*)src;
  int* d = (int*)dst;
  const int ph_end = hdr + nph * (int)(sizeof(RPhase) / 4);
  constexpr int ops0 = (int)(offsetof(PassGeom, ops) / 4);
  const int ops_end = ops0 + nops * (int)(sizeof(ROp) / 4);
  for (int i = tid; i < ph_end; i += nt) d[i] = s[i];
  for (int i = ops0 + tid; i < ops_end; i += nt) d[i] = s[i];
}

// Fixed-order sum of per-workgroup partials laid out accumulator-major
// (part[j * stride + block]): one workgroup of 256 threads per (item j).
__device__ __forceinline__ double sum_blocks(const double* __restrict__ p, uint32_t n) {
  __shared__ double sh[256];
  double t = 0.0;
  for (uint32_t b = threadIdx.x; b < n; b += 256) t += p[b];
  sh[threadIdx.x] = t;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  return sh[0];
}

// One pass over tiles t = t0, t0 + tstep, ... < tend.  Reductions are added
// to sums[0..kAcc) (LDS, zeroed by the caller): [0] norm^2, [1..16]
// histogram, [17..32] packed rho; each wavefront accumulates its tiles in
// red[wave * kAcc + j] (LDS, NT/64 * kAcc doubles) and the waves are summed
// in order at the end.  lds holds 2^T elements.
// fresh != 0: the state is |0...0> and is not read (its first pass).
// The host emulation of this loop (tests/emu) follows it step for step.
template <typename R, typename O>
__device__ __forceinline__ void tile_pass(C<R>* __restrict__ state, const PassGeom& g, const C<R>* mats, C<R>* lds,
                                          uint64_t t0, uint64_t tstep, uint64_t tend, double* red, double* sums,
                                          int fresh = 0) {
  constexpr int NT = TileCfg<R>::NT;
  constexpr int NR = 1 << TileCfg<R>::RB;
  const int T = g.T;
  const R scale = (R)g.scale;
  double* wred = red + (threadIdx.x >> 6) * kAcc;
  for (int i = threadIdx.x; i < (NT / 64) * kAcc; i += NT) red[i] = 0.0;
  __syncthreads();
  ThreadIO<O> io;
  thread_io<R, O>(g, threadIdx.x, io);
  C<R> v[NR], nx[NR];
  // tile t's base is t's bits deposited on the outside bits M (ascending), so
  // the next tile's base is a masked add: ((base | ~M) + deposit(tstep)) & M
  const uint64_t M = uniform_u64(tile_base(g, ~0ull)), S = uniform_u64(tile_base(g, tstep));
  uint64_t base = uniform_u64(tile_base(g, t0));
  // Software pipeline over the tiles (disjoint): tile t's operators run while
  // the loads of tile t+1 and the stores of tile t-1 are in flight.  The wait
  // for tile t+1's loads is placed before tile t's stores are issued: the
  // vector-memory counter does not order loads against stores, so a wait for
  // loads issued before stores also waits for the stores; waiting first lets
  // tile t's stores complete under tile t+1's operators instead.
  if (t0 < tend) {
    if (!fresh) fetch_regs<R, O>(nx, state, io, base);
    take_regs<R, O>(v, nx, io, base, scale, fresh);
    if (!fresh && t0 + tstep < tend) fetch_regs<R, O>(nx, state, io, ((base | ~M) + S) & M);
  }
  for (uint64_t t = t0; t < tend; t += tstep) {
    opaque_io(io);
    const uint64_t nbase = ((base | ~M) + S) & M;
    const bool more = t + tstep < tend;
    int8_t cur[kRB];
#pragma unroll
    for (int j = 0; j < kRB; j++) cur[j] = g.io_in[j];
    int ci = kLayIn;
    for (int p = 0; p <= g.nph; p++) {
      int8_t want[kRB];
      const int wi = p < g.nph ? p : kLayOut;
#pragma unroll
      for (int j = 0; j < kRB; j++) want[j] = p < g.nph ? g.ph[p].rb[j] : g.io_out[j];
      bool same = true;
#pragma unroll
      for (int j = 0; j < kRB; j++) same = same && (cur[j] == want[j]);
      if (!same) {
        regs_to_lds<R>(v, lds, g.lay[ci], io.tid, io.active, io.nvalid_mask);
        __syncthreads();
        lds_to_regs<R>(v, lds, g.lay[wi], io.ltid);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kRB; j++) cur[j] = want[j];
        ci = wi;
      }
      if (p == g.nph) break;
      const int op0 = g.ph[p].op0, op1 = op0 + g.ph[p].nops;
      const int tm = uniform_i32(g.ph[p].tmpl);
      if (tm > 0) apply_template<R>(tm, v, mats, g.ops + op0);
      for (int o = op0 + (tm > 0 ? 2 : 0); o < op1; o++) {  // every template covers two operators
        const ROp& op = g.ops[o];
        if ((base & op.gctrl) != op.gctrl) continue;
        if (op.k <= 2) {
          apply_mask<R>(op.mask, v, mats + op.moff);
        } else {
          regs_to_lds<R>(v, lds, g.lay[ci], io.tid, io.active, io.nvalid_mask);
          __syncthreads();
          if (op.k == 3) apply_lds<R, 3>(lds, T, op, mats + op.moff, io.tid);
          else apply_lds<R, 4>(lds, T, op, mats + op.moff, io.tid);
          __syncthreads();
          lds_to_regs<R>(v, lds, g.lay[ci], io.ltid);
          __syncthreads();
        }
      }
      if (p == g.nph - 1 && g.red == bti::RED_RHO) {
        if (uflag(g.red_k == 3)) {
          regs_to_lds<R>(v, lds, g.lay[ci], io.tid, io.active, io.nvalid_mask);
          __syncthreads();
          rho3_accumulate_t<R>(lds, g, wred + 1);
          __syncthreads();
        } else {
          double racc[16];
#pragma unroll
          for (int j = 0; j < 16; j++) racc[j] = 0.0;
          if (io.active) rho_mask<R>(rho_reg_mask(g), v, io.nvalid_mask, racc);
          wave_accumulate<16>(racc, wred + 1);
        }
      }
    }
    double acc[17];
#pragma unroll
    for (int j = 0; j < 17; j++) acc[j] = 0.0;
    {
      // tile t+1's loads (and tile t-1's stores) are done; prefetched values
      // the register allocator parked in scratch come back here, before the
      // stores, so that reloading them does not wait for the stores.  On
      // every path (also the last tile, or when nothing was fetched): the
      // compiler's wait bookkeeping merges paths, and one without this wait
      // would make it wait at the take below, after the stores.
#pragma unroll
      for (int r = 0; r < NR; r++) {
        opaque(nx[r].x);
        opaque(nx[r].y);
      }
      vm_wait_all();
    }
    store_regs<R, O>(v, state, g, io, base, acc);
    if (g.red == bti::RED_HIST) {
      switch (g.red_k) {  // norm + 2^red_k bins
        case 1: wave_accumulate<3>(acc, wred); break;
        case 2: wave_accumulate<5>(acc, wred); break;
        case 3: wave_accumulate<9>(acc, wred); break;
        default: wave_accumulate<17>(acc, wred); break;
      }
    } else if (g.need_norm) {
      wave_accumulate<1>(acc, wred);
    }
    if (more) {  // tile t+1 into registers, tile t+2's loads in flight
      take_regs<R, O>(v, nx, io, nbase, scale, fresh);
      if (!fresh && t + 2 * tstep < tend) fetch_regs<R, O>(nx, state, io, ((nbase | ~M) + S) & M);
    }
    base = nbase;
  }
  __syncthreads();
  if (threadIdx.x < kAcc) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; w++) t += red[w * kAcc + threadIdx.x];
    sums[threadIdx.x] += t;
  }
  __syncthreads();
}
#endif

// Packed reduced rho (diag, then a < b Re/Im) over k bits -> column-major
// complex; swap: the packed index's bit 0 is the rho's MSB.
__host__ __device__ inline void unpack_rho(int k, const double* packed, double* rho, int swap = 0) {
  const int D = 1 << k;
  for (int i = 0; i < 2 * D * D; i++) rho[i] = 0.0;
  auto mp = [&](int a) { return (swap && k == 2) ? (((a & 1) << 1) | (a >> 1)) : a; };
  for (int a = 0; a < D; a++) rho[2 * (mp(a) + mp(a) * D)] = packed[a];
  int p = D;
  for (int a = 0; a < D; a++)
    for (int b = a + 1; b < D; b++) {
      const int A = mp(a), B = mp(b);
      rho[2 * (A + B * D)] = packed[p];
      rho[2 * (A + B * D) + 1] = packed[p + 1];
      rho[2 * (B + A * D)] = packed[p];
      rho[2 * (B + A * D) + 1] = -packed[p + 1];
      p += 2;
    }
}

__host__ __device__ inline void to_result(const Plan& plan, int red_swap, const double* sum, PassResult* res) {
  res->norm2 = sum[0];
  for (int j = 0; j < 16; j++) res->hist[j] = plan.red == bti::RED_HIST && j < (1 << plan.red_k) ? sum[1 + j] : 0.0;
  for (int j = 0; j < 128; j++) res->rho[j] = 0.0;
  if (plan.red == bti::RED_RHO) unpack_rho(plan.red_k, sum + 1, res->rho, red_swap);
}

}  // namespace bt
