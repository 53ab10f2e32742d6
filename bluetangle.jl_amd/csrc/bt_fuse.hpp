// Gate fusion for noiseless gate-only circuits (host, no HIP dependencies):
// consecutive gates are multiplied into dense blocks on at most max_k bits,
// which one pass applies at once (the k = 5 blocks of complex64 states run
// on FP32 MFMA).  Gates on disjoint qubits commute, so a block also takes a
// later gate when no skipped gate in between shares a qubit with it.
#pragma once

#include <algorithm>
#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "bt.h"
#include "bt_circuit.hpp"

namespace btf {

using cd = std::complex<double>;

struct Block {
  int k = 0;
  int bits[5] = {0, 0, 0, 0, 0};  // bits[0] = MSB of the block's matrix index
  std::vector<cd> m;              // 2^k x 2^k, column-major
  int ngates = 0;
};

// Expanded gate g (targets t[0..gk), t[0] = MSB; controls ctrl) applied on
// the left of the block matrix M over bits B (B[0] = MSB): M <- G M.
inline void left_multiply(std::vector<cd>& M, int k, const int* B, int gk, const int32_t* t, uint64_t ctrl,
                          const double* g) {
  const int D = 1 << k, GD = 1 << gk;
  int tpos[5], cpos[5], nc = 0;
  for (int i = 0; i < gk; i++)
    for (int j = 0; j < k; j++)
      if (B[j] == t[i]) tpos[i] = k - 1 - j;  // index bit of target i
  for (int j = 0; j < k; j++)
    if ((ctrl >> B[j]) & 1) cpos[nc++] = k - 1 - j;
  std::vector<cd> out(M.size());
  for (int c = 0; c < D; c++) {  // column c of M
    for (int r = 0; r < D; r++) {  // out[r][c] = sum_x G_exp[r][x] M[x][c]
      bool on = true;
      for (int i = 0; i < nc; i++) on = on && ((r >> cpos[i]) & 1);
      if (!on) {
        out[r + c * D] = M[r + c * D];
        continue;
      }
      int gr = 0, rest = r;
      for (int i = 0; i < gk; i++) {
        gr |= ((r >> tpos[i]) & 1) << (gk - 1 - i);
        rest &= ~(1 << tpos[i]);
      }
      cd acc = 0;
      for (int gc = 0; gc < GD; gc++) {
        int x = rest;
        for (int i = 0; i < gk; i++)
          if ((gc >> (gk - 1 - i)) & 1) x |= 1 << tpos[i];
        acc += btc::at(g, GD, gr, gc) * M[x + c * D];
      }
      out[r + c * D] = acc;
    }
  }
  M.swap(out);
}

// Plan the circuit into blocks of at most max_k bits (2 <= max_k <= 5);
// pad_to > 0 widens every block to that many bits with identity on extra
// low bits (so every pass can use the same k-bit kernel).
inline int fuse_gates(const btc::CircuitHost& c, int max_k, int pad_to, std::vector<Block>* out, std::string* msg) {
  out->clear();
  if (max_k < 2 || max_k > 5 || pad_to > 5) {
    *msg = "fusion width must be 2..5";
    return BT_ERR_ARG;
  }
  const int nops = c.ntop;
  std::vector<uint64_t> touch(nops);
  for (int o = 0; o < nops; o++) {
    const bt_ir_op& op = c.ops[o];
    if (op.kind != BT_OP_GATE) {
      *msg = "fused runs take gate-only circuits (no channels, noise or measurements)";
      return BT_ERR_UNSUPPORTED;
    }
    uint64_t t = op.ctrl_mask;
    for (int i = 0; i < op.k; i++) t |= 1ull << op.bits[i];
    if (__builtin_popcountll(t) > max_k) {
      *msg = "a gate touches more bits than the fusion width";
      return BT_ERR_ARG;
    }
    touch[o] = t;
  }
  std::vector<char> done(nops, 0);
  int first = 0;
  while (true) {
    while (first < nops && done[first]) first++;
    if (first >= nops) break;
    uint64_t B = touch[first], blocked = 0;
    std::vector<int> sel;
    for (int o = first; o < nops; o++) {
      if (done[o]) continue;
      if ((touch[o] & blocked) == 0 && __builtin_popcountll(B | touch[o]) <= max_k) {
        B |= touch[o];
        sel.push_back(o);
      } else {
        blocked |= touch[o];
      }
    }
    // pad with the lowest free bits
    const int want = pad_to > 0 ? std::min(pad_to, c.n) : 0;
    for (int b = 0; b < c.n && __builtin_popcountll(B) < want; b++) B |= 1ull << b;
    Block blk;
    blk.k = __builtin_popcountll(B);
    int j = 0;
    for (int b = 63; b >= 0; b--)
      if ((B >> b) & 1) blk.bits[j++] = b;  // descending: bits[0] = MSB
    const int D = 1 << blk.k;
    blk.m.assign((size_t)D * D, cd(0, 0));
    for (int i = 0; i < D; i++) blk.m[i + i * D] = 1.0;
    for (int o : sel) {
      const bt_ir_op& op = c.ops[o];
      left_multiply(blk.m, blk.k, blk.bits, op.k, op.bits, op.ctrl_mask, c.pool.data() + 2 * op.mat);
      done[o] = 1;
    }
    blk.ngates = (int)sel.size();
    out->push_back(std::move(blk));
  }
  return BT_OK;
}

}  // namespace btf
