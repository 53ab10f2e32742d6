// CPU emulation of the engine's trajectory executor, for tests only.
//
// It drives the product's interpreter (bt_interp.hpp) and circuit analysis
// (bt_circuit.hpp) exactly as the GPU executors do, but runs each planned
// pass on a complex128 vector on the host.  It checks that every operator of
// a pass lies inside the pass's tile bits, so planner mistakes fail here
// before any GPU time is spent.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bt_circuit.hpp"
#include "bt_interp.hpp"

using cd = std::complex<double>;

static std::string g_msg;

static void apply_op(std::vector<cd>& s, int nbits, const bti::QOp& op) {
  const int k = op.k, D = 1 << k;
  uint64_t tmask = 0;
  for (int i = 0; i < k; i++) tmask |= 1ull << op.bits[i];
  std::vector<cd> v(D), w(D);
  for (uint64_t base = 0; base < (1ull << nbits); base++) {
    if (base & tmask) continue;
    if ((base & op.ctrl) != op.ctrl) continue;
    for (int l = 0; l < D; l++) {
      uint64_t o = 0;
      for (int i = 0; i < k; i++)
        if ((l >> (k - 1 - i)) & 1) o |= 1ull << op.bits[i];
      v[l] = s[base | o];
    }
    for (int r = 0; r < D; r++) {
      cd a = 0;
      for (int c = 0; c < D; c++) a += cd(op.m[2 * (r + c * D)], op.m[2 * (r + c * D) + 1]) * v[c];
      w[r] = a;
    }
    for (int l = 0; l < D; l++) {
      uint64_t o = 0;
      for (int i = 0; i < k; i++)
        if ((l >> (k - 1 - i)) & 1) o |= 1ull << op.bits[i];
      s[base | o] = w[l];
    }
  }
}

extern "C" const char* emu_error(void) { return g_msg.c_str(); }

// stats: passes, reduce passes, boundary draws, draws used, overflow
extern "C" int emu_run(int nbits, const bt_ir_op* ops, int nops, int ntop, const double* pool, int64_t npool,
                       uint64_t seed, uint64_t traj, int T, int cmin, int fold, double* state, int* outs,
                       int64_t outcap, int64_t* nout, int64_t* stats, double tol_exact) {
  btc::CircuitHost c;
  int rc = btc::analyze(nbits, ops, nops, ntop, pool, npool, &c, &g_msg);
  if (rc) return rc;
  std::vector<bti::QOp> slots(bti::kQMax);
  bti::Interp ip;
  bti::Program prog{c.ops.data(), c.ntop, c.pool.data(), c.chan.data(), c.chanw.data(), nbits};
  ip.init(prog, slots.data(), T, cmin);
  ip.seed = seed;
  ip.traj = traj;
  ip.outs = outs;
  ip.outcap = outcap;
  ip.fold = fold;
  ip.tol_exact = tol_exact;
  const uint64_t len = 1ull << nbits;
  std::vector<cd> s(len);
  for (uint64_t i = 0; i < len; i++) s[i] = cd(state[2 * i], state[2 * i + 1]);
  bti::Plan plan;
  bti::PassResult res;
  int guard = 0;
  while (ip.step(&plan, &res)) {
    if (++guard > 1000000) {
      g_msg = "emulator: too many passes";
      return BT_ERR_STATE;
    }
    uint64_t tile = 0;
    for (int i = 0; i < plan.ntile; i++) tile |= 1ull << plan.tile[i];
    const int want = T < nbits ? T : nbits;
    if (plan.ntile != want) {
      g_msg = "emulator: tile has " + std::to_string(plan.ntile) + " bits, expected " + std::to_string(want);
      return BT_ERR_STATE;
    }
    for (uint64_t i = 0; i < len; i++) s[i] *= plan.scale;
    for (int j = 0; j < plan.nops; j++) {
      const bti::QOp& op = slots[plan.qidx[j]];
      for (int b = 0; b < op.k; b++)
        if (!((tile >> op.bits[b]) & 1)) {
          g_msg = "emulator: operator bit outside the tile";
          return BT_ERR_STATE;
        }
      if (plan.write) apply_op(s, nbits, op);
    }
    memset(&res, 0, sizeof res);
    for (uint64_t i = 0; i < len; i++) res.norm2 += std::norm(s[i]);
    if (plan.red == bti::RED_HIST) {
      for (uint64_t i = 0; i < len; i++) {
        int bin = 0;
        for (int j = 0; j < plan.red_k; j++) bin |= (int)((i >> plan.red_bits[j]) & 1) << j;
        res.hist[bin] += std::norm(s[i]);
      }
    } else if (plan.red == bti::RED_RHO) {
      const int k = plan.red_k, D = 1 << k;
      uint64_t m = 0;
      for (int j = 0; j < k; j++) {
        m |= 1ull << plan.red_bits[j];
        if (!((tile >> plan.red_bits[j]) & 1)) {
          g_msg = "emulator: reduction bit outside the tile";
          return BT_ERR_STATE;
        }
      }
      for (uint64_t base = 0; base < len; base++) {
        if (base & m) continue;
        cd v[4];
        for (int a = 0; a < D; a++) {
          uint64_t o = 0;
          for (int j = 0; j < k; j++)
            if ((a >> (k - 1 - j)) & 1) o |= 1ull << plan.red_bits[j];
          v[a] = s[base | o];
        }
        for (int a = 0; a < D; a++)
          for (int b = 0; b < D; b++) {
            cd t = v[a] * std::conj(v[b]);
            res.rho[2 * (a + b * D)] += t.real();
            res.rho[2 * (a + b * D) + 1] += t.imag();
          }
      }
    }
  }
  if (ip.err) {
    g_msg = "interpreter error " + std::to_string(ip.err);
    return ip.err;
  }
  for (uint64_t i = 0; i < len; i++) {
    state[2 * i] = s[i].real();
    state[2 * i + 1] = s[i].imag();
  }
  *nout = ip.nout;
  stats[0] = ip.passes;
  stats[1] = ip.rpasses;
  stats[2] = ip.boundary;
  stats[3] = ip.draw;
  stats[4] = ip.overflow;
  return BT_OK;
}
