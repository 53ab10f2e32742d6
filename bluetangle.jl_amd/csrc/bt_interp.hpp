// Trajectory interpreter shared by the engine's executors.
//
// It walks a circuit's IR (gates, Kraus channels, measurements, resets,
// conditional branches) for one trajectory, keeps the not-yet-applied
// operators in a queue (fusing 1- and 2-bit operators on the same bits by
// matrix products), decides sampled branches as early as the mathematics
// allows, and emits "passes": one streaming sweep over the state in tiles
// of 2^T amplitudes whose T tile bits hold every bit the pass's operators
// touch.  A pass also returns the output's FP64 norm^2 and, when a decision
// needs it, a histogram of |amplitude|^2 over up to 4 bits or the reduced
// density matrix over up to 2 bits.
//
// Semantics follow the reference's trajectory rules:
//  - one uniform per sampled channel (hilbert.jl:810-819, u <= cumsum), per
//    Born measurement (hilbert.jl:682-696, u < prob0) and one extra for the
//    MR basis (struct.jl:565-571); channel probabilities use the reduced rho
//    in ascending qubit order (struct.jl:9-29);
//  - every channel / measurement branch renormalises (struct.jl:34,40);
//    plain gates never do (hilbert.jl:505), even non-unitary ones such as P0.
//    Here the device state is phi = c * psi_ref with a positive factor c:
//    lazily decided branches leave c unknown, every pass measures its output
//    norm^2 N = c^2 * |psi_ref|^2, and at the end of each flush (state aligned
//    with a point of the op stream) c is recovered and 1/c is folded into the
//    next pass's input scale.  Probabilities are ratios of reductions, so
//    they never need c.
//  - probabilities are those of the reference's unnormalised state: after a
//    non-unitary gate |psi_ref|^2 = r != 1 and p_i = r * (ratio), so r is
//    tracked too (and measured by a flush when a non-unitary gate made it
//    unknown);
//  - a channel whose Kraus operators all satisfy K_i'K_i = w_i I (Pauli-like
//    mixtures, e.g. depolarizing) has state-independent probabilities w_i and
//    is decided from u alone; other channels are decided from u alone when
//    u lies outside the interval the cumulative sums can reach given the
//    eigenvalue bounds of K_i'K_i (for amplitude damping: u <= 1 - gamma),
//    otherwise they wait for an exact reduction.
//
// Compiled for the device (one interpreter per workgroup, driven by lane 0)
// and for the host (the large-state executor and the CPU test emulator).
#pragma once

#include <stdint.h>
#include <math.h>

#include "bt.h"

#if defined(__HIPCC__)
#define BT_HD __host__ __device__
#define BT_NOINLINE __attribute__((noinline))
#else
#define BT_HD
#define BT_NOINLINE
#endif

namespace bti {

constexpr int kQMax = 40;     // queued operators (slots)
constexpr int kPMax = 24;     // operators per pass
constexpr int kMaxT = 15;     // tile bits
constexpr int kGroupMax = 4;  // decisions resolved from one histogram
constexpr int kMaxQK = 4;     // widest queued operator (16 x 16: density-matrix superoperators)
// a sampled Kraus branch whose norm^2 is below this fraction of the state's
// is taken as annihilating it (exactly zero up to rounding)
constexpr double kZeroBranch = 1e-10;
#if defined(__HIP_DEVICE_COMPILE__)
constexpr int kFuseMax = 2;   // fused products on the device stay 4 x 4 (no scratch arrays)
#else
constexpr int kFuseMax = 4;
#endif

enum { RED_NONE = 0, RED_HIST = 1, RED_RHO = 2 };
enum { CH_PAULI = 0, CH_BOUNDED = 1 };

// Per-channel facts computed on the host when the circuit is created.
struct ChanDesc {
  int32_t mode;      // channels: CH_PAULI / CH_BOUNDED; gates: 1 if the matrix is unitary
  int32_t monomial;  // every Kraus op has at most one nonzero per column
  int64_t w;         // offset into the weights array: nk weights (PAULI) or nk (lo, hi) pairs
};

struct QOp {
  int32_t k;
  int32_t bits[kMaxQK];  // bits[0] = MSB of the local index
  uint64_t ctrl;
  double m[2 * 256];     // column-major complex128, (1<<k)^2 entries used
};

struct Plan {
  int32_t ntile;
  int32_t tile[kMaxT];   // ascending
  int32_t nops;
  int32_t qidx[kPMax];   // queue slots of the pass's operators, in application order
  int32_t red;           // RED_*
  int32_t red_k;
  int32_t red_bits[kGroupMax];  // RED_HIST: bin bit j <-> red_bits[j]; RED_RHO: MSB first
  int32_t write;         // 0: read-only reduction pass
  int32_t aligned;       // last pass of a flush: the output is the op-stream state
  double scale;          // input scale
};

struct PassResult {
  double norm2;
  double hist[16];
  double rho[128];       // up to 8x8 complex, column-major
};

struct Program {
  const bt_ir_op* ops;
  int32_t ntop;
  const double* pool;    // complex128 interleaved
  const ChanDesc* chan;  // indexed by op index
  const double* chanw;
  int32_t nbits;
};

// ---------------------------------------------------------------------------
// small complex helpers on interleaved doubles
// ---------------------------------------------------------------------------
BT_HD inline void cmulacc(double& ar, double& ai, double br, double bi, double cr, double ci) {
  ar += br * cr - bi * ci;
  ai += br * ci + bi * cr;
}

BT_HD inline uint64_t splitmix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

struct Interp {
  // program and trajectory
  Program P;
  uint64_t seed, traj;
  const double* uni;  // optional caller uniforms
  int64_t nuni;
  int32_t T;          // tile bits
  int32_t cmin;       // low contiguous bits a tile keeps for coalescing
  double tol_exact, tol_pauli;
  int32_t normalize_end;
  // outcome record
  int32_t* outs;
  int64_t outcap, nout;
  // queue: operator slots (storage provided by the caller) in queue order
  QOp* slots;
  int32_t order[kQMax];
  int32_t nq;
  uint64_t used_slots;
  int32_t fold;        // 0: no renormalisation bookkeeping (density matrices)
  // control
  int32_t pc, end, sp, ret_pc, ret_end;
  int64_t draw;
  double norm2;        // measured norm^2 of the last aligned state
  double c2, r, pend_scale;
  int32_t c2_known, r_known;
  int32_t flushing, after, awaiting, done_program, final_stage;
  int32_t red_req, red_k, red_bits[kGroupMax], red_done;
  PassResult last;
  // pending decision
  int32_t g_n, g_pc[kGroupMax];
  double g_u[kGroupMax];
  int32_t d_pc;
  double d_u;
  int32_t d_basis;
  // bookkeeping
  int32_t pass_nops;
  int32_t pass_q[kPMax];
  int64_t passes, rpasses, boundary, overflow;
  double bytes_units;  // in units of "one read of the state"
  int32_t err;

  BT_HD void init(const Program& prog, QOp* queue, int32_t tile_bits, int32_t coalesce_bits) {
    P = prog;
    slots = queue;
    nq = 0;
    used_slots = 0;
    fold = 1;
    T = tile_bits < prog.nbits ? tile_bits : prog.nbits;
    cmin = coalesce_bits;
    if (cmin > T) cmin = T;
    pc = 0;
    end = prog.ntop;
    sp = 0;
    draw = 0;
    norm2 = 1.0;
    c2 = 1.0;
    r = 1.0;
    pend_scale = 1.0;
    c2_known = r_known = 1;
    flushing = after = awaiting = done_program = final_stage = 0;
    red_req = RED_NONE;
    red_done = 1;
    nout = 0;
    passes = rpasses = boundary = overflow = 0;
    bytes_units = 0;
    err = 0;
    uni = nullptr;
    nuni = 0;
    outs = nullptr;
    outcap = 0;
    tol_exact = 1e-6;
    tol_pauli = 1e-12;
    normalize_end = 1;
  }

  BT_HD double uniform_at(int64_t i) {
    if (uni) {
      if (i >= nuni) {
        err = BT_ERR_DRAWS;
        return 0.5;
      }
      return uni[i];
    }
    uint64_t z = splitmix(seed ^ splitmix(traj ^ splitmix((uint64_t)i)));
    return (double)(z >> 11) * 0x1.0p-53;
  }

  // ---- operator queue with fusion ----------------------------------------
  BT_HD static uint64_t maskof(const int32_t* bits, int k) {
    uint64_t m = 0;
    for (int i = 0; i < k; i++) m |= 1ull << bits[i];
    return m;
  }

  // Express an operator on bits `sb` (k_s) as an operator on bits `db` (k_d)
  // with sb a subset of db: out = M_s (x) I on the other bits.
  BT_HD static void expand(const double* ms, const int32_t* sb, int ks, const int32_t* db, int kd, double* out) {
    const int D = 1 << kd, Ds = 1 << ks;
    int pos[kMaxQK];  // position of sb[i] within db
    for (int i = 0; i < ks; i++)
      for (int j = 0; j < kd; j++)
        if (db[j] == sb[i]) pos[i] = j;
    uint64_t smask = 0;  // local-index bits of db covered by sb
    for (int i = 0; i < ks; i++) smask |= 1ull << (kd - 1 - pos[i]);
    for (int r = 0; r < D; r++)
      for (int c = 0; c < D; c++) {
        double vr = 0, vi = 0;
        if ((r & ~smask) == (c & ~smask)) {
          int rs = 0, cs = 0;
          for (int i = 0; i < ks; i++) {
            int b = kd - 1 - pos[i];
            rs |= ((r >> b) & 1) << (ks - 1 - i);
            cs |= ((c >> b) & 1) << (ks - 1 - i);
          }
          vr = ms[2 * (rs + cs * Ds)];
          vi = ms[2 * (rs + cs * Ds) + 1];
        }
        out[2 * (r + c * D)] = vr;
        out[2 * (r + c * D) + 1] = vi;
      }
  }

  // out = A * B (D x D), all column-major complex128
  BT_HD static void matmul(const double* A, const double* B, int D, double* out) {
    for (int r = 0; r < D; r++)
      for (int c = 0; c < D; c++) {
        double sr = 0, si = 0;
        for (int x = 0; x < D; x++)
          cmulacc(sr, si, A[2 * (r + x * D)], A[2 * (r + x * D) + 1], B[2 * (x + c * D)], B[2 * (x + c * D) + 1]);
        out[2 * (r + c * D)] = sr;
        out[2 * (r + c * D) + 1] = si;
      }
  }

  BT_HD QOp& Q(int i) { return slots[order[i]]; }

  BT_HD void queue_remove(int i) {
    used_slots &= ~(1ull << order[i]);
    for (int j = i; j + 1 < nq; j++) order[j] = order[j + 1];
    nq--;
  }

  BT_HD int alloc_slot() {
    for (int s = 0; s < kQMax; s++)
      if (!((used_slots >> s) & 1)) {
        used_slots |= 1ull << s;
        return s;
      }
    return -1;
  }

  // Push an operator (k <= 4, column-major complex128 matrix times `scale`).
  // Operators on the same bits are multiplied together: if the last queued
  // operator touching any of the new one's bits acts inside them, it is
  // absorbed (nothing after it touches those bits, so the product may move
  // to the end); if the new one acts inside the last one's bits, it is
  // absorbed into it in place.
  BT_HD void push(int k, const int32_t* bits, uint64_t ctrl, const double* m, double scale) {
    const int sa = alloc_slot();
    if (sa < 0) {
      err = BT_ERR_STATE;
      return;
    }
    QOp& a = slots[sa];
    a.k = k;
    for (int i = 0; i < k; i++) a.bits[i] = bits[i];
    a.ctrl = ctrl;
    const int D = 1 << k;
    for (int i = 0; i < 2 * D * D; i++) a.m[i] = m[i] * scale;
    const uint64_t am = maskof(a.bits, a.k) | a.ctrl;
    constexpr int FD = 1 << kFuseMax;
    for (;;) {
      int last = -1;
      for (int i = nq - 1; i >= 0; i--) {
        uint64_t bm = maskof(Q(i).bits, Q(i).k) | Q(i).ctrl;
        if (bm & am) {
          last = i;
          break;
        }
      }
      if (last < 0 || a.ctrl || Q(last).ctrl || a.k > kFuseMax || Q(last).k > kFuseMax) break;
      QOp& b = Q(last);
      const uint64_t amask = maskof(a.bits, a.k), bmask = maskof(b.bits, b.k);
      if ((bmask & ~amask) == 0) {
        // b acts inside a's bits: a <- a * expand(b); b leaves the queue
        double eb[2 * FD * FD], prod[2 * FD * FD];
        expand(b.m, b.bits, b.k, a.bits, a.k, eb);
        matmul(a.m, eb, D, prod);
        for (int i = 0; i < 2 * D * D; i++) a.m[i] = prod[i];
        queue_remove(last);
        continue;
      }
      if ((amask & ~bmask) == 0) {
        // a acts inside b's bits: b <- expand(a) * b, in place
        double ea[2 * FD * FD], prod[2 * FD * FD];
        const int Db = 1 << b.k;
        expand(a.m, a.bits, a.k, b.bits, b.k, ea);
        matmul(ea, b.m, Db, prod);
        for (int i = 0; i < 2 * Db * Db; i++) b.m[i] = prod[i];
        used_slots &= ~(1ull << sa);
        return;
      }
      break;
    }
    order[nq++] = sa;
  }

  BT_HD void push_pool(const bt_ir_op& op) {
    push(op.k, op.bits, op.ctrl_mask, P.pool + 2 * op.mat, 1.0);
  }

  // ---- flush planning --------------------------------------------------------
  BT_HD void start_flush(int red, int k, const int32_t* bits) {
    flushing = 1;
    red_req = red;
    red_k = k;
    for (int i = 0; i < k; i++) red_bits[i] = bits[i];
    red_done = (red == RED_NONE);
  }

  // choose the ops of the next pass for a candidate tile mask; returns count
  // matrix entries a pass can hold (the device keeps them in LDS)
  static constexpr int kPassMatEntries = 768;

  BT_HD int select_ops(uint64_t tile, int32_t* sel) {
    uint64_t blocked = 0;
    int n = 0, ents = 0;
    for (int i = 0; i < nq && n < kPMax; i++) {
      uint64_t b = maskof(Q(i).bits, Q(i).k);
      uint64_t touch = b | Q(i).ctrl;
      if ((touch & blocked) == 0 && (b & ~tile) == 0) {
        // in-tile controls become operand bits of the pass's operator
        int kc = Q(i).k + popc(Q(i).ctrl);
        if (kc > 4) kc = 4;
        const int e = 1 << (2 * kc);
        if (ents + e > kPassMatEntries) break;
        ents += e;
        sel[n++] = i;
      } else {
        blocked |= touch;
      }
    }
    return n;
  }

  BT_HD uint64_t fill_low(uint64_t tile) {
    int cnt = 0;
    for (int b = 0; b < P.nbits; b++) cnt += (tile >> b) & 1;
    for (int b = 0; b < P.nbits && cnt < T; b++)
      if (!((tile >> b) & 1)) {
        tile |= 1ull << b;
        cnt++;
      }
    return tile;
  }

  BT_HD static int popc(uint64_t x) {
    int c = 0;
    while (x) {
      x &= x - 1;
      c++;
    }
    return c;
  }

  // Greedy over the queue, seeded with the first operator: the target bits
  // of the operators that fit in `cap` bits, in queue order, skipping those
  // that touch a bit an earlier skipped operator touches.
  BT_HD uint64_t greedy_bits(int cap) {
    uint64_t g = 0, blocked = 0;
    for (int i = 0; i < nq; i++) {
      uint64_t b = maskof(Q(i).bits, Q(i).k);
      uint64_t touch = b | Q(i).ctrl;
      if (touch & blocked) {
        blocked |= touch;
        continue;
      }
      // the first operator always seeds the tile, even past the coalescing reserve
      if (popc(g | b) <= cap || (g == 0 && popc(b) <= T)) g |= b;
      else blocked |= touch;
    }
    return g;
  }

  BT_HD void make_plan(Plan* plan) {
    uint64_t need = 0;  // reduction bits that must be inside the tile
    if (red_req == RED_RHO && !red_done)
      for (int i = 0; i < red_k; i++) need |= 1ull << red_bits[i];
    uint64_t tile;
    int32_t sel[kPMax];
    int n = 0;
    if (nq == 0) {
      tile = fill_low(need);
    } else {
      // candidate 1: greedy over the queue, seeded with the first operator
      uint64_t t1 = fill_low(greedy_bits(T - cmin));
      int32_t s1[kPMax];
      int n1 = select_ops(t1, s1);
      // candidate 1b: the same with one more low bit reserved (twice the
      // contiguous run per access, 128 B for complex64 at 3 reserved bits),
      // taken when it selects as many operators
      if (cmin + 1 < T) {
        const uint64_t t1b = fill_low(greedy_bits(T - cmin - 1));
        if (t1b != t1) {
          int32_t s1b[kPMax];
          const int n1b = select_ops(t1b, s1b);
          if (n1b >= n1) {
            t1 = t1b;
            n1 = n1b;
            for (int i = 0; i < n1; i++) s1[i] = s1b[i];
          }
        }
      }
      // candidate 2: the low tile [0, T)
      uint64_t t2 = fill_low(0);
      int32_t s2[kPMax];
      int n2 = select_ops(t2, s2);
      if (n2 > n1) {
        tile = t2;
        n = n2;
        for (int i = 0; i < n; i++) sel[i] = s2[i];
      } else {
        tile = t1;
        n = n1;
        for (int i = 0; i < n; i++) sel[i] = s1[i];
      }
      if (n == 0) {  // cannot happen with k <= 3 <= T; guard against a stall
        err = BT_ERR_UNSUPPORTED;
        tile = t2;
      }
      // the last pass of a flush should hold the reduction bits
      if (n == nq && need && (need & ~tile)) {
        uint64_t used = 0;
        for (int i = 0; i < n; i++) used |= maskof(Q(sel[i]).bits, Q(sel[i]).k);
        if (popc(used | need) <= T) {
          uint64_t t3 = fill_low(used | need);
          int32_t s3[kPMax];
          int n3 = select_ops(t3, s3);
          if (n3 == n) {
            tile = t3;
            for (int i = 0; i < n; i++) sel[i] = s3[i];
          }
        }
      }
    }
    plan->ntile = 0;
    for (int b = 0; b < P.nbits; b++)
      if ((tile >> b) & 1) plan->tile[plan->ntile++] = b;
    plan->nops = n;
    for (int i = 0; i < n; i++) plan->qidx[i] = order[sel[i]];
    const bool last = (n == nq);
    plan->red = RED_NONE;
    plan->red_k = 0;
    if (last && !red_done) {
      if (red_req == RED_HIST || (need & ~tile) == 0) {
        plan->red = red_req;
        plan->red_k = red_k;
        for (int i = 0; i < red_k; i++) plan->red_bits[i] = red_bits[i];
      }
    }
    plan->write = n > 0 ? 1 : 0;
    plan->aligned = last ? 1 : 0;
    // the pending scale rides on the next pass that writes the state; a
    // reduction-only pass leaves it pending (its reductions are of the
    // stored state, and flush_end accounts for the scale still to come)
    if (plan->write) {
      plan->scale = pend_scale;
      pend_scale = 1.0;
    } else {
      plan->scale = 1.0;
    }
    pass_nops = n;
    for (int i = 0; i < n; i++) pass_q[i] = sel[i];
  }

  // The state is aligned with the op stream: N is the measured norm^2 of the
  // stored state phi, and pend_scale^2 N = c^2 |psi_ref|^2 (the pending
  // scale, 1 after a writing pass, is still to be applied to phi).
  BT_HD void flush_end(double N) {
    norm2 = N;
    if (!fold) return;
    const double Ne = N * pend_scale * pend_scale;
    if (r_known) {
      c2 = Ne / r;
    } else if (c2_known) {
      r = Ne / c2;
      r_known = 1;
    } else {
      err = BT_ERR_STATE;
      return;
    }
    c2_known = 1;
    pend_scale *= 1.0 / sqrt(c2);
    c2 = 1.0;
  }

  // Bookkeeping of c and |psi_ref| when a renormalising branch with
  // (ratio) probability p is pushed; p < 0 means "not known".
  BT_HD void renormalised(double p) {
    if (p >= 0 && c2_known && r_known) c2 *= p * r;
    else c2_known = 0;
    r = 1.0;
    r_known = 1;
  }

  BT_HD void consume(const Plan& plan, const PassResult& r) {
    // drop the applied operators (indices ascending)
    for (int i = pass_nops - 1; i >= 0; i--) queue_remove(pass_q[i]);
    pass_nops = 0;
    if (plan.aligned) flush_end(r.norm2);
    if (plan.red != RED_NONE) {
      last = r;
      red_done = 1;
    }
    if (plan.write) {
      passes++;
      bytes_units += 2.0;
    } else {
      rpasses++;
      bytes_units += 1.0;
    }
  }

  // ---- decisions ----------------------------------------------------------------
  BT_HD int weighted(const double* p, int n, double u, double tol) {
    double c = 0;
    int chosen = -1;
    for (int i = 0; i < n; i++) {
      c += p[i];
      if (fabs(u - c) <= tol && i < n - 1) boundary++;
      if (chosen < 0 && u <= c) chosen = i;
    }
    if (chosen < 0) {
      overflow++;
      chosen = n - 1;
      while (chosen > 0 && !(p[chosen] > 0)) chosen--;
    }
    return chosen;
  }

  // Pick from bounds alone (probabilities p_i = r * ratio_i with ratio_i in
  // [lo_i, hi_i] summing to 1); returns -1 when undecided.
  BT_HD int bounded_pick(const double* lohi, int n, double u) {
    double slo = 0, shi = 0, tlo = 0, thi = 0;
    for (int i = 0; i < n; i++) {
      tlo += lohi[2 * i];
      thi += lohi[2 * i + 1];
    }
    for (int i = 0; i < n; i++) {
      slo += lohi[2 * i];
      shi += lohi[2 * i + 1];
      double rest_lo = tlo - slo, rest_hi = thi - shi;
      double clo = r * fmax(slo, 1.0 - rest_hi);  // the cumulative sum c_i lies in [clo, chi]
      double chi = r * fmin(shi, 1.0 - rest_lo);
      if (u <= clo - tol_exact) return i;  // u <= c_i for sure, and u > c_{i-1} was established
      if (u <= chi + tol_exact) return -1; // could go either way
    }
    return -1;
  }

  BT_HD static bool is_group_member(const bt_ir_op& op) {
    return (op.kind == BT_OP_MEASURE && op.basis == BT_BASIS_Z) || op.kind == BT_OP_RESET;
  }

  // Apply a decided monomial 2x2 operator on histogram bit j (classical update).
  BT_HD static void hist_apply(double* h, int nb, int j, const double* m /*2x2 col-major complex*/) {
    const int nbin = 1 << nb;
    double out[16];
    for (int i = 0; i < nbin; i++) out[i] = 0;
    for (int i = 0; i < nbin; i++) {
      int v = (i >> j) & 1;
      for (int r = 0; r < 2; r++) {
        double re = m[2 * (r + 2 * v)], im = m[2 * (r + 2 * v) + 1];
        double w = re * re + im * im;
        if (w != 0) out[(i & ~(1 << j)) | (r << j)] += w * h[i];
      }
    }
    for (int i = 0; i < nbin; i++) h[i] = out[i];
  }

  BT_HD void decide_group() {
    double h[16];
    for (int i = 0; i < 16; i++) h[i] = last.hist[i];
    const int nb = g_n;
    for (int j = 0; j < g_n; j++) {
      const bt_ir_op& op = P.ops[g_pc[j]];
      double tot = 0, z = 0;
      for (int i = 0; i < (1 << nb); i++) {
        tot += h[i];
        if (!((i >> j) & 1)) z += h[i];
      }
      const double q0 = z / tot;       // ratio
      const double p0 = q0 * r;        // the reference's unnormalised prob0
      const double u = g_u[j];
      if (fabs(u - p0) <= tol_exact) boundary++;
      const int ind = u < p0 ? 0 : 1;
      renormalised(ind == 0 ? q0 : 1.0 - q0);
      double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (ind == 0) m[0] = 1.0;            // P0
      else if (op.kind == BT_OP_RESET) m[4] = 1.0;  // |0><1|
      else m[6] = 1.0;                     // P1
      if (op.kind == BT_OP_MEASURE && op.record) record(ind);
      hist_apply(h, nb, j, m);
      push(1, op.bits, 0, m, 1.0);
    }
  }

  BT_HD void record(int ind) {
    if (outs && nout < outcap) outs[nout] = ind;
    nout++;
  }

  // rho decision: measurement in X/Y basis or a channel needing exact rho
  BT_HD void decide_rho() {
    const bt_ir_op& op = P.ops[d_pc];
    if (op.kind == BT_OP_CHANNEL) {
      const int k = op.k, D = 1 << k;
      double tr = 0;
      for (int a = 0; a < D; a++) tr += last.rho[2 * (a + a * D)];
      double p[64];
      for (int i = 0; i < op.nkraus; i++) {
        const double* K = P.pool + 2 * (op.mat + (int64_t)i * D * D);
        double t = 0;
        for (int r = 0; r < D; r++)
          for (int a = 0; a < D; a++) {
            double kr = K[2 * (r + a * D)], ki = K[2 * (r + a * D) + 1];
            if (kr == 0 && ki == 0) continue;
            for (int b = 0; b < D; b++) {
              // K[r,a] rho[a,b] conj(K[r,b])
              double xr = 0, xi = 0;
              cmulacc(xr, xi, kr, ki, last.rho[2 * (a + b * D)], last.rho[2 * (a + b * D) + 1]);
              t += xr * K[2 * (r + b * D)] + xi * K[2 * (r + b * D) + 1];
            }
          }
        p[i] = t / tr * r;
      }
      int ind = weighted(p, op.nkraus, d_u, tol_exact);
      const double* Kc = P.pool + 2 * (op.mat + (int64_t)ind * D * D);
      if (op.bits[0] == op.prob_bits[0]) {
        renormalised(p[ind] / r);
        push(k, op.bits, 0, Kc, 1.0);
        return;
      }
      // qubit > target ordering quirk: |K psi|^2 in application order is not
      // p[ind] (c stays unknown until the next flush), and the sampled branch
      // can annihilate the state, which the reference's normalize turns into
      // NaN everywhere; the application-order norm from the same reduced rho
      // (bit order swapped) tells such a branch from a small one
      double psum = 0, q = 0;
      for (int i = 0; i < op.nkraus; i++) psum += p[i];
      auto sw = [](int a) { return ((a & 1) << 1) | (a >> 1); };
      for (int r2 = 0; r2 < D; r2++)
        for (int a = 0; a < D; a++) {
          const double kr = Kc[2 * (r2 + a * D)], ki = Kc[2 * (r2 + a * D) + 1];
          if (kr == 0 && ki == 0) continue;
          for (int b = 0; b < D; b++) {
            const int sa = k == 2 ? sw(a) : a, sb = k == 2 ? sw(b) : b;
            double xr = 0, xi = 0;
            cmulacc(xr, xi, kr, ki, last.rho[2 * (sa + sb * D)], last.rho[2 * (sa + sb * D) + 1]);
            q += xr * Kc[2 * (r2 + b * D)] + xi * Kc[2 * (r2 + b * D) + 1];
          }
        }
      q = q / tr * r;
      renormalised(-1.0);
      push(k, op.bits, 0, Kc, q <= kZeroBranch * psum ? NAN : 1.0);
      return;
    }
    // measurement / if: rho over one bit; rotate by R
    const double s = 0.70710678118654752440;
    double R[8];
    if (d_basis == BT_BASIS_X) {
      double t[8] = {s, 0, s, 0, s, 0, -s, 0};
      for (int i = 0; i < 8; i++) R[i] = t[i];
    } else if (d_basis == BT_BASIS_Y) {
      double t[8] = {s, 0, s, 0, 0, -s, 0, s};
      for (int i = 0; i < 8; i++) R[i] = t[i];
    } else {
      double t[8] = {1, 0, 0, 0, 0, 0, 1, 0};
      for (int i = 0; i < 8; i++) R[i] = t[i];
    }
    double tr = last.rho[0] + last.rho[6];
    // (R rho R')[0,0] = sum_xy R[0,x] rho[x,y] conj(R[0,y])
    double v00 = 0;
    for (int x = 0; x < 2; x++)
      for (int y = 0; y < 2; y++) {
        double ar = 0, ai = 0;
        cmulacc(ar, ai, R[2 * (0 + 2 * x)], R[2 * (0 + 2 * x) + 1], last.rho[2 * (x + 2 * y)],
                last.rho[2 * (x + 2 * y) + 1]);
        v00 += ar * R[2 * (0 + 2 * y)] + ai * R[2 * (0 + 2 * y) + 1];
      }
    const double q0 = v00 / tr;
    const double p0 = q0 * r;
    if (fabs(d_u - p0) <= tol_exact) boundary++;
    const int ind = d_u < p0 ? 0 : 1;
    renormalised(ind == 0 ? q0 : 1.0 - q0);
    // G = R' P_ind R : G[a,b] = conj(R[ind,a]) R[ind,b]
    double G[8];
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) {
        double xr = R[2 * (ind + 2 * a)], xi = -R[2 * (ind + 2 * a) + 1];
        double yr = R[2 * (ind + 2 * b)], yi = R[2 * (ind + 2 * b) + 1];
        G[2 * (a + 2 * b)] = xr * yr - xi * yi;
        G[2 * (a + 2 * b) + 1] = xr * yi + xi * yr;
      }
    if (op.record) record(ind);
    push(1, op.bits, 0, G, 1.0);
    if (op.kind == BT_OP_IF) enter_branch(op, ind);
  }

  BT_HD void enter_branch(const bt_ir_op& op, int ind) {
    // pc already points past the IF op
    ret_pc = pc;
    ret_end = end;
    sp = 1;
    pc = ind == 0 ? op.br0_begin : op.br1_begin;
    end = ind == 0 ? op.br0_end : op.br1_end;
  }

  // ---- the driver -------------------------------------------------------------
  // Returns 1 with a pass in *plan (run it, then call again with its
  // result), or 0 when the trajectory is complete.
  BT_HD BT_NOINLINE int step(Plan* plan, const PassResult* res) {
    if (awaiting) {
      consume(*plan, *res);
      awaiting = 0;
    }
    for (int guard = 0; guard < (1 << 30); guard++) {
      if (err) return 0;
      if (flushing) {
        if (nq > 0 || !red_done) {
          make_plan(plan);
          if (err) return 0;
          awaiting = 1;
          return 1;
        }
        flushing = 0;
        int a = after;
        after = 0;
        if (a == 1) decide_group();
        else if (a == 2) decide_rho();
        continue;
      }
      if (done_program) {
        if (final_stage == 0) {
          final_stage = 1;
          if (nq > 0) {
            start_flush(RED_NONE, 0, nullptr);
            continue;
          }
        }
        if (final_stage == 1) {
          final_stage = 2;
          if (normalize_end && fabs(pend_scale - 1.0) > 1e-15) {
            plan->ntile = 0;
            uint64_t t = fill_low(0);
            for (int b = 0; b < P.nbits; b++)
              if ((t >> b) & 1) plan->tile[plan->ntile++] = b;
            plan->nops = 0;
            plan->red = RED_NONE;
            plan->red_k = 0;
            plan->write = 1;
            plan->aligned = 0;
            plan->scale = pend_scale;
            pend_scale = 1.0;
            pass_nops = 0;
            awaiting = 1;
            return 1;
          }
        }
        return 0;
      }
      if (pc >= end) {
        if (sp > 0) {
          sp = 0;
          pc = ret_pc;
          end = ret_end;
          continue;
        }
        done_program = 1;
        continue;
      }
      if (nq >= kQMax - 2) {
        start_flush(RED_NONE, 0, nullptr);
        continue;
      }
      const bt_ir_op& op = P.ops[pc];
      switch (op.kind) {
        case BT_OP_GATE: {
          if (op.k > kMaxQK) {
            err = BT_ERR_UNSUPPORTED;
            return 0;
          }
          if (!P.chan[pc].mode) {
            // a non-unitary gate changes |psi_ref| (gates never renormalise):
            // c must be known before it, so settle it with a flush first
            if (!c2_known) {
              start_flush(RED_NONE, 0, nullptr);
              break;
            }
            r_known = 0;
          }
          push_pool(op);
          pc++;
          break;
        }
        case BT_OP_CHANNEL: {
          const ChanDesc& cd = P.chan[pc];
          const int D = 1 << op.k;
          const double u = uniform_at(draw);
          if (!r_known) {  // |psi_ref| must be known to sample: measure it
            start_flush(RED_NONE, 0, nullptr);
            break;
          }
          if (cd.mode == CH_PAULI) {
            const double* w = P.chanw + cd.w;
            double pw[64];
            for (int i = 0; i < op.nkraus; i++) pw[i] = w[i] * r;
            const int ind = weighted(pw, op.nkraus, u, tol_pauli);
            // K/sqrt(w) is unitary: phi' = c |psi_ref| psi_ref'
            if (r_known) c2 *= r;
            else c2_known = 0;
            r = 1.0;
            r_known = 1;
            push(op.k, op.bits, 0, P.pool + 2 * (op.mat + (int64_t)ind * D * D), 1.0 / sqrt(w[ind]));
            draw++;
            pc++;
            break;
          }
          const int ind = bounded_pick(P.chanw + cd.w, op.nkraus, u);
          if (ind >= 0) {
            renormalised(-1.0);
            push(op.k, op.bits, 0, P.pool + 2 * (op.mat + (int64_t)ind * D * D), 1.0);
            draw++;
            pc++;
            break;
          }
          // exact: reduced rho over the probability bits (up to 3)
          d_pc = pc;
          d_u = u;
          draw++;
          pc++;
          after = 2;
          start_flush(RED_RHO, op.k, op.prob_bits);
          break;
        }
        case BT_OP_MEASURE:
        case BT_OP_RESET:
        case BT_OP_IF: {
          if (is_group_member(op)) {
            // gather consecutive Z-basis measurements / resets on distinct bits
            g_n = 0;
            uint64_t used = 0;
            int j = pc;
            int64_t dj = draw;
            while (j < end && g_n < kGroupMax && is_group_member(P.ops[j]) && !((used >> P.ops[j].bits[0]) & 1)) {
              g_pc[g_n] = j;
              g_u[g_n] = uniform_at(dj);
              used |= 1ull << P.ops[j].bits[0];
              g_n++;
              dj++;
              j++;
            }
            int32_t hb[kGroupMax];
            for (int i = 0; i < g_n; i++) hb[i] = P.ops[g_pc[i]].bits[0];
            draw = dj;
            pc = j;
            after = 1;
            start_flush(RED_HIST, g_n, hb);
            break;
          }
          int basis = op.basis;
          if (basis == BT_BASIS_R) {
            double ub = uniform_at(draw);
            draw++;
            int pick = (int)floor(ub * 3.0);
            if (pick > 2) pick = 2;
            basis = pick == 0 ? BT_BASIS_X : (pick == 1 ? BT_BASIS_Y : BT_BASIS_Z);
          }
          d_basis = basis;
          d_pc = pc;
          d_u = uniform_at(draw);
          draw++;
          pc++;
          after = 2;
          start_flush(RED_RHO, 1, op.bits);
          break;
        }
        default:
          err = BT_ERR_ARG;
          return 0;
      }
    }
    err = BT_ERR_STATE;
    return 0;
  }
};

}  // namespace bti
