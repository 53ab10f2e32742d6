// Host-side circuit validation and analysis (no HIP dependencies): checks the
// IR, counts draws, and classifies every channel for the interpreter
// (Pauli-like weights, or eigenvalue bounds of K_i'K_i) and every gate
// (unitary or not).
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bt.h"
#include "bt_interp.hpp"

namespace btc {

struct CircuitHost {
  int n = 0;
  std::vector<bt_ir_op> ops;
  int ntop = 0;
  std::vector<double> pool;
  std::vector<bti::ChanDesc> chan;
  std::vector<double> chanw;
  int64_t max_draws = 0;
  bool has_measure = false;
};

using cd = std::complex<double>;

inline cd at(const double* m, int D, int r, int c) { return cd(m[2 * (r + c * D)], m[2 * (r + c * D) + 1]); }

// eigenvalues of a Hermitian D x D matrix via cyclic Jacobi on the real
// symmetric 2D x 2D embedding [[Re, -Im], [Im, Re]] (each eigenvalue twice)
inline void herm_eig_bounds(const std::vector<cd>& H, int D, double* lo, double* hi) {
  const int n = 2 * D;
  std::vector<double> A(n * n);
  for (int r = 0; r < D; r++)
    for (int c = 0; c < D; c++) {
      cd v = H[r + c * D];
      A[r + c * n] = v.real();
      A[(r + D) + (c + D) * n] = v.real();
      A[(r + D) + c * n] = v.imag();
      A[r + (c + D) * n] = -v.imag();
    }
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) off += A[p + q * n] * A[p + q * n];
    if (off < 1e-30) break;
    for (int p = 0; p < n; p++)
      for (int q = p + 1; q < n; q++) {
        double apq = A[p + q * n];
        if (std::fabs(apq) < 1e-300) continue;
        double theta = (A[q + q * n] - A[p + p * n]) / (2 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < n; k++) {
          double akp = A[k + p * n], akq = A[k + q * n];
          A[k + p * n] = c * akp - s * akq;
          A[k + q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {
          double apk = A[p + k * n], aqk = A[q + k * n];
          A[p + k * n] = c * apk - s * aqk;
          A[q + k * n] = s * apk + c * aqk;
        }
      }
  }
  double mn = 1e300, mx = -1e300;
  for (int i = 0; i < n; i++) {
    mn = std::fmin(mn, A[i + i * n]);
    mx = std::fmax(mx, A[i + i * n]);
  }
  *lo = mn;
  *hi = mx;
}

inline int64_t draws_of(const std::vector<bt_ir_op>& ops, int i) {
  const bt_ir_op& op = ops[i];
  switch (op.kind) {
    case BT_OP_CHANNEL:
    case BT_OP_RESET: return 1;
    case BT_OP_MEASURE: return op.basis == BT_BASIS_R ? 2 : 1;
    case BT_OP_IF: {
      int64_t a = 0, b = 0;
      for (int j = op.br0_begin; j < op.br0_end; j++) a += draws_of(ops, j);
      for (int j = op.br1_begin; j < op.br1_end; j++) b += draws_of(ops, j);
      return (op.basis == BT_BASIS_R ? 2 : 1) + (a > b ? a : b);
    }
    default: return 0;
  }
}

// Validate and analyse.  Returns BT_OK or an error code with *msg set.
inline int analyze(int nbits, const bt_ir_op* ops, int nops, int ntop, const double* pool, int64_t npool,
                   CircuitHost* out, std::string* msg) {
  if (nbits < 1 || nbits > 36) {
    *msg = "qubit count out of range";
    return BT_ERR_ARG;
  }
  if (nops < 0 || ntop < 0 || ntop > nops || (nops > 0 && !ops) || npool < 0 || (npool > 0 && !pool)) {
    *msg = "bad circuit arrays";
    return BT_ERR_ARG;
  }
  out->n = nbits;
  out->ops.assign(ops, ops + nops);
  out->ntop = ntop;
  out->pool.assign(pool, pool + 2 * npool);
  out->chan.assign(nops, bti::ChanDesc{1, 0, 0});
  out->chanw.clear();
  out->has_measure = false;
  for (int i = 0; i < nops; i++) {
    const bt_ir_op& op = ops[i];
    const int maxk = op.kind == BT_OP_GATE ? 5 : (op.kind == BT_OP_CHANNEL ? 3 : 1);
    if (op.k < 1 || op.k > maxk) {
      *msg = "op " + std::to_string(i) + ": width k out of range";
      return BT_ERR_ARG;
    }
    uint64_t used = 0;
    for (int j = 0; j < op.k; j++) {
      if (op.bits[j] < 0 || op.bits[j] >= nbits || ((used >> op.bits[j]) & 1)) {
        *msg = "op " + std::to_string(i) + ": bad bit position";
        return BT_ERR_ARG;
      }
      used |= 1ull << op.bits[j];
    }
    if (op.kind == BT_OP_GATE) {
      if ((op.ctrl_mask & used) || (nbits < 64 && (op.ctrl_mask >> nbits))) {
        *msg = "op " + std::to_string(i) + ": bad control mask";
        return BT_ERR_ARG;
      }
    } else if (op.ctrl_mask) {
      *msg = "op " + std::to_string(i) + ": only gates take controls";
      return BT_ERR_ARG;
    }
    const int D = 1 << op.k;
    if (op.kind == BT_OP_GATE || op.kind == BT_OP_CHANNEL) {
      int64_t need = (int64_t)D * D * (op.kind == BT_OP_CHANNEL ? op.nkraus : 1);
      if (op.mat < 0 || op.mat + need > npool || (op.kind == BT_OP_CHANNEL && (op.nkraus < 1 || op.nkraus > 64))) {
        *msg = "op " + std::to_string(i) + ": matrix pool range";
        return BT_ERR_ARG;
      }
    }
    if (op.kind == BT_OP_CHANNEL) {
      uint64_t pu = 0;
      for (int j = 0; j < op.k; j++) pu |= 1ull << op.prob_bits[j];
      if (pu != used) {
        *msg = "op " + std::to_string(i) + ": prob_bits must permute bits";
        return BT_ERR_ARG;
      }
    }
    if (op.kind == BT_OP_MEASURE || op.kind == BT_OP_IF || op.kind == BT_OP_RESET) {
      out->has_measure = true;
      if (op.basis < 0 || op.basis > 3) {
        *msg = "op " + std::to_string(i) + ": bad basis";
        return BT_ERR_ARG;
      }
    }
    if (op.kind == BT_OP_IF) {
      if (i >= ntop || op.br0_begin < ntop || op.br1_begin < ntop || op.br0_end < op.br0_begin ||
          op.br1_end < op.br1_begin || op.br0_end > nops || op.br1_end > nops) {
        *msg = "op " + std::to_string(i) + ": bad branch ranges";
        return BT_ERR_ARG;
      }
      for (int j = op.br0_begin; j < op.br0_end; j++)
        if (ops[j].kind == BT_OP_IF) {
          *msg = "nested conditional branches are not supported";
          return BT_ERR_ARG;
        }
      for (int j = op.br1_begin; j < op.br1_end; j++)
        if (ops[j].kind == BT_OP_IF) {
          *msg = "nested conditional branches are not supported";
          return BT_ERR_ARG;
        }
    }
    if (op.kind < 0 || op.kind > BT_OP_IF) {
      *msg = "op " + std::to_string(i) + ": unknown kind";
      return BT_ERR_ARG;
    }
  }
  // classification
  for (int i = 0; i < nops; i++) {
    const bt_ir_op& op = ops[i];
    const int D = 1 << op.k;
    if (op.kind == BT_OP_GATE) {
      const double* M = pool + 2 * op.mat;
      bool unitary = true;
      for (int a = 0; a < D && unitary; a++)
        for (int b = 0; b < D; b++) {
          cd s = 0;
          for (int r = 0; r < D; r++) s += std::conj(at(M, D, r, a)) * at(M, D, r, b);
          if (std::abs(s - (a == b ? 1.0 : 0.0)) > 1e-12) {
            unitary = false;
            break;
          }
        }
      out->chan[i].mode = unitary ? 1 : 0;
      continue;
    }
    if (op.kind != BT_OP_CHANNEL) continue;
    bool pauli = true, mono = true;
    std::vector<double> w(op.nkraus), lo(op.nkraus), hi(op.nkraus);
    for (int k = 0; k < op.nkraus; k++) {
      const double* K = pool + 2 * (op.mat + (int64_t)k * D * D);
      std::vector<cd> H(D * D);
      for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++) {
          cd s = 0;
          for (int r = 0; r < D; r++) s += std::conj(at(K, D, r, a)) * at(K, D, r, b);
          H[a + b * D] = s;
        }
      for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++) {
          if (a != b && std::abs(H[a + b * D]) > 1e-13) pauli = false;
          if (a == b && std::abs(H[a + a * D] - H[0]) > 1e-13) pauli = false;
        }
      w[k] = H[0].real();
      herm_eig_bounds(H, D, &lo[k], &hi[k]);
      for (int c = 0; c < D; c++) {
        int nz = 0;
        for (int r = 0; r < D; r++) nz += at(K, D, r, c) != cd(0, 0);
        if (nz > 1) mono = false;
      }
    }
    bti::ChanDesc d;
    d.monomial = mono;
    d.w = (int64_t)out->chanw.size();
    if (pauli) {
      d.mode = bti::CH_PAULI;
      for (double x : w) out->chanw.push_back(x);
    } else {
      d.mode = bti::CH_BOUNDED;
      for (int k = 0; k < op.nkraus; k++) {
        out->chanw.push_back(lo[k]);
        out->chanw.push_back(hi[k]);
      }
    }
    out->chan[i] = d;
  }
  if (out->chanw.empty()) out->chanw.push_back(0.0);
  int64_t total = 0;
  for (int i = 0; i < ntop; i++) total += draws_of(out->ops, i);
  out->max_draws = total;
  return BT_OK;
}

using cd = std::complex<double>;

// Density-matrix lowering: every op becomes superoperators on vec(rho)
// (row bit b at b, column bit b at b + N).  A gate followed by its noise
// channel on the same bits becomes one superoperator sum_i (K_i U) (x)
// conj(K_i U); other gates become U on the row bits and conj(U) on the column
// bits (controls on both sides).
inline int lower_dm(const CircuitHost& c, CircuitHost* out, std::string* msg) {
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

}  // namespace btc

// A validated circuit and its device copies (uploaded on first use).
struct bt_circuit_s {
  btc::CircuitHost h;
  uint64_t uid = 0;  // process-unique, never reused (server caches key on it)
  // device copies (uploaded on first use)
  bt_ir_op* d_ops = nullptr;
  double* d_pool = nullptr;
  bti::ChanDesc* d_chan = nullptr;
  double* d_chanw = nullptr;
  std::mutex mu;
  int fused_k = -1;  // bt_autotune_fused's choice (0: register-tile executor), -1: not tuned
  // pass tables of this circuit whose plan depends on nothing the passes
  // compute (its density-matrix form; its state-vector form when it holds
  // gates only), kept for the tiling they were planned with; guarded by mu
  struct StaticPlan {
    int T = -1, prec = -1, cb = -1;
    size_t tabsz = 0;
    std::vector<unsigned char> tables;
    std::vector<uint8_t> narrow;
  } dm_plan, sv_plan;
};
