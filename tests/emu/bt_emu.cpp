// CPU emulation of the engine's trajectory executor, for tests only.
//
// It drives the product's interpreter (bt_interp.hpp) and circuit analysis
// (bt_circuit.hpp) exactly as the GPU executors do, but runs each planned
// pass on the host.  It checks that every operator of a pass lies inside the
// pass's tile bits, so planner mistakes fail here before any GPU time is
// spent.  Two pass executors:
//  * exec 0: every operator of the plan on a complex128 vector (planner check);
//  * exec 1 / 2: the register-tile pass of bt_tile.hpp in complex128 /
//    complex64, thread by thread: the product's geometry (build_geom) and
//    per-thread load / relayout / operator / store routines, with each
//    barrier of the device loop becoming a loop over the workgroup.
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bt_circuit.hpp"
#include "bt_fuse.hpp"
#include "bt_interp.hpp"
#include "bt_tile.hpp"

using cd = std::complex<double>;

static std::string g_msg;

static int emu_fail(const char* m) {
  g_msg = m;
  return BT_ERR_STATE;
}

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

static int g_exec = 0;
// geometry statistics over all register-tile passes: passes, operators,
// operators with k = 1, 2, 3, 4 (in-tile controls counted), phases, relayouts
// per tile, operators applied through LDS
static int64_t g_gstat[9];

static FILE* g_geom_dump = nullptr;
static int64_t g_tmpl_phases = 0;  // phases run through a straight-line template
// passes whose tile holds local bit `g_deep_bit` (the bit right after the
// coalescing reserve: twice the contiguous run per access)
static int g_deep_bit = -1;
static int64_t g_deep_passes = 0;

extern "C" void emu_set_exec(int mode) { g_exec = mode; }
// Append every register-tile pass geometry (PassGeom, then nmats complex
// entries of the pass precision) to a file; NULL closes it.
extern "C" int emu_dump_geometry(const char* path) {
  if (g_geom_dump) fclose(g_geom_dump);
  g_geom_dump = path ? fopen(path, "wb") : nullptr;
  return path && !g_geom_dump ? -1 : 0;
}
extern "C" int64_t emu_template_phases(int reset) {
  const int64_t v = g_tmpl_phases;
  if (reset) g_tmpl_phases = 0;
  return v;
}

extern "C" int64_t emu_deep_coalesced_passes(int reset) {
  const int64_t v = g_deep_passes;
  if (reset) g_deep_passes = 0;
  return v;
}

extern "C" void emu_geom_stats(int64_t* out, int reset) {
  for (int i = 0; i < 9; i++) out[i] = g_gstat[i];
  if (reset) memset(g_gstat, 0, sizeof g_gstat);
}

template <typename R, typename O>
static int tile_pass_emu(std::vector<bt::C<R>>& st, int nbits, const bti::Plan& plan, const bti::QOp* slots,
                         bti::PassResult* res, int fresh) {
  using namespace bt;
  constexpr int NT = TileCfg<R>::NT;
  constexpr int NR = 1 << TileCfg<R>::RB;
  PassGeom g;
  memset(&g, 0, sizeof g);
  std::vector<C<R>> mats(kMatEntries);
  int rc = build_geom<R>(plan, slots, nbits, &g, mats.data());
  if (rc) {
    g_msg = "emulator: build_geom failed";
    return rc;
  }
  g_gstat[0]++;
  if (g_geom_dump) {
    fwrite(&g, sizeof g, 1, g_geom_dump);
    fwrite(mats.data(), sizeof(C<R>), (size_t)g.nmats, g_geom_dump);
  }
  g_gstat[1] += g.nops;
  for (int o = 0; o < g.nops; o++) g_gstat[1 + g.ops[o].k]++;
  g_gstat[6] += g.nph;
  for (int i = 0; i < g.T; i++)
    if (g.tile[i] == g_deep_bit) g_deep_passes++;
  const int T = g.T;
  const R scale = (R)g.scale;
  const uint64_t ntiles = 1ull << (nbits - T);
  std::vector<C<R>> lds((size_t)1 << T);
  std::vector<std::array<C<R>, NR>> v(NT);
  std::vector<ThreadIO<O>> io(NT);
  for (int tid = 0; tid < NT; tid++) thread_io<R, O>(g, tid, io[tid]);
  double sums[kAcc];
  for (int j = 0; j < kAcc; j++) sums[j] = 0.0;
  // the device advances tile bases by a masked add of the deposited step;
  // check it against the direct deposit for a few steps
  const uint64_t M = tile_base(g, ~0ull);
  uint64_t base = tile_base(g, 0);
  for (uint64_t t = 0; t < ntiles; t++) {
    if (base != tile_base(g, t)) return emu_fail("masked-add tile base mismatch");
    for (uint64_t step : {3ull, 7ull, 64ull})
      if (t + step < ntiles && (((base | ~M) + tile_base(g, step)) & M) != tile_base(g, t + step))
        return emu_fail("masked-add tile base mismatch (step)");
    for (int tid = 0; tid < NT; tid++) {
      std::array<C<R>, NR> raw{};
      if (!fresh) fetch_regs<R, O>(raw.data(), st.data(), io[tid], base);
      take_regs<R, O>(v[tid].data(), raw.data(), io[tid], base, scale, fresh);
    }
    int8_t cur[kRB];
    for (int j = 0; j < kRB; j++) cur[j] = g.io_in[j];
    LSlots curS = g.lay[kLayIn];
    for (int p = 0; p <= g.nph; p++) {
      int8_t want[kRB];
      for (int j = 0; j < kRB; j++) want[j] = p < g.nph ? g.ph[p].rb[j] : g.io_out[j];
      const LSlots wantS = g.lay[p < g.nph ? p : kLayOut];
      bool same = true;
      for (int j = 0; j < kRB; j++) same = same && (cur[j] == want[j]);
      if (!same) {
        if (t == 0) g_gstat[7]++;
        for (int tid = 0; tid < NT; tid++)
          regs_to_lds<R>(v[tid].data(), lds.data(), curS, tid, io[tid].active, io[tid].nvalid_mask);
        for (int tid = 0; tid < NT; tid++) lds_to_regs<R>(v[tid].data(), lds.data(), wantS, io[tid].ltid);
        for (int j = 0; j < kRB; j++) cur[j] = want[j];
        curS = wantS;
      }
      if (p == g.nph) break;
      const int op0 = g.ph[p].op0, op1 = op0 + g.ph[p].nops;
      if (g.ph[p].tmpl > 0) {  // as the device: the matched template, straight-line
        for (int tid = 0; tid < NT; tid++) apply_template<R>(g.ph[p].tmpl, v[tid].data(), mats.data(), g.ops + op0);
        g_tmpl_phases++;
      }
      for (int o = op0 + (g.ph[p].tmpl > 0 ? kPhaseTemplates[g.ph[p].tmpl].n : 0); o < op1; o++) {
        const ROp& op = g.ops[o];
        if ((base & op.gctrl) != op.gctrl) continue;
        if (op.k <= 2) {
          for (int tid = 0; tid < NT; tid++) apply_mask<R>(op.mask, v[tid].data(), mats.data() + op.moff);
        } else {
          if (t == 0) g_gstat[8]++;
          for (int tid = 0; tid < NT; tid++)
            regs_to_lds<R>(v[tid].data(), lds.data(), curS, tid, io[tid].active, io[tid].nvalid_mask);
          for (int tid = 0; tid < NT; tid++) {
            if (op.k == 3) apply_lds<R, 3>(lds.data(), T, op, mats.data() + op.moff, tid);
            else apply_lds<R, 4>(lds.data(), T, op, mats.data() + op.moff, tid);
          }
          for (int tid = 0; tid < NT; tid++) lds_to_regs<R>(v[tid].data(), lds.data(), curS, io[tid].ltid);
        }
      }
      if (p == g.nph - 1 && g.red == bti::RED_RHO && g.red_k == 3) {
        // as the device: the tile image in LDS, per-thread partials, the
        // quad's four summed (s0 + s1) + (s2 + s3)
        for (int tid = 0; tid < NT; tid++)
          regs_to_lds<R>(v[tid].data(), lds.data(), curS, tid, io[tid].active, io[tid].nvalid_mask);
        for (int q = 0; q < NT / 4; q++) {
          double xr[4], xi[4];
          int e = -1;
          for (int s = 0; s < 4; s++) e = rho3_partial<R>(lds.data(), g, (uint32_t)(4 * q + s), &xr[s], &xi[s]);
          if (e < 0) continue;
          const int sl = rho3_slot(e);
          sums[1 + sl] += (xr[0] + xr[1]) + (xr[2] + xr[3]);
          if (e >= 8) sums[2 + sl] += (xi[0] + xi[1]) + (xi[2] + xi[3]);
        }
      } else if (p == g.nph - 1 && g.red == bti::RED_RHO) {
        for (int tid = 0; tid < NT; tid++) {
          double racc[16];
          for (int j = 0; j < 16; j++) racc[j] = 0.0;
          if (io[tid].active) rho_mask<R>(rho_reg_mask(g), v[tid].data(), io[tid].nvalid_mask, racc);
          for (int j = 0; j < 16; j++) sums[1 + j] += racc[j];
        }
      }
    }
    for (int tid = 0; tid < NT; tid++) {
      double acc[17];
      for (int j = 0; j < 17; j++) acc[j] = 0.0;
      store_regs<R, O>(v[tid].data(), st.data(), g, io[tid], base, acc);
      if (g.red == bti::RED_HIST)
        for (int j = 0; j < 17; j++) sums[j] += acc[j];
      else if (g.need_norm)
        sums[0] += acc[0];
    }
    base = ((base | ~M) + tile_base(g, 1)) & M;
  }
  to_result(plan, g.red_swap, sums, res);
  return BT_OK;
}

template <typename R>
static int run_tiles(bti::Interp& ip, const bti::QOp* slots, int nbits, std::vector<std::complex<double>>& s) {
  const uint64_t len = 1ull << nbits;
  std::vector<bt::C<R>> st(len);
  bool zero = true;
  for (uint64_t i = 0; i < len; i++) {
    st[i].x = (R)s[i].real();
    st[i].y = (R)s[i].imag();
    if (s[i] != (i == 0 ? std::complex<double>(1, 0) : std::complex<double>(0, 0))) zero = false;
  }
  int fresh = zero ? 1 : 0;  // as the batched executors: |0..0> is not read on its first pass
  bti::Plan plan;
  bti::PassResult res;
  memset(&res, 0, sizeof res);
  int guard = 0;
  while (ip.step(&plan, &res)) {
    if (++guard > 1000000) {
      g_msg = "emulator: too many passes";
      return BT_ERR_STATE;
    }
    memset(&res, 0, sizeof res);
    // offsets as the device: 32-bit up to 2^32 amplitudes
    const int rc = nbits <= 32 ? tile_pass_emu<R, uint32_t>(st, nbits, plan, slots, &res, fresh)
                               : tile_pass_emu<R, uint64_t>(st, nbits, plan, slots, &res, fresh);
    if (rc) return rc;
    if (plan.write) fresh = 0;
  }
  for (uint64_t i = 0; i < len; i++) s[i] = std::complex<double>(st[i].x, st[i].y);
  return BT_OK;
}

// stats: passes, reduce passes, boundary draws, draws used, overflow
static int run_host(const btc::CircuitHost& c, int nbits, uint64_t seed, uint64_t traj, int T, int cmin, int fold,
                    double* state, int* outs, int64_t outcap, int64_t* nout, int64_t* stats, double tol_exact) {
  int rc = BT_OK;
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
  if (g_exec) {
    rc = g_exec == 1 ? run_tiles<double>(ip, slots.data(), nbits, s) : run_tiles<float>(ip, slots.data(), nbits, s);
    if (rc) return rc;
  }
  while (!g_exec && ip.step(&plan, &res)) {
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
        cd v[8];
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

extern "C" int emu_run(int nbits, const bt_ir_op* ops, int nops, int ntop, const double* pool, int64_t npool,
                       uint64_t seed, uint64_t traj, int T, int cmin, int fold, double* state, int* outs,
                       int64_t outcap, int64_t* nout, int64_t* stats, double tol_exact) {
  btc::CircuitHost c;
  int rc = btc::analyze(nbits, ops, nops, ntop, pool, npool, &c, &g_msg);
  if (rc) return rc;
  g_deep_bit = cmin;
  return run_host(c, nbits, seed, traj, T, cmin, fold, state, outs, outcap, nout, stats, tol_exact);
}

// to_rho as the density-matrix executor runs it: the circuit lowered to
// superoperators on the 2N-bit vector of rho (column-major: bits 0..N-1 the
// row, N..2N-1 the column), then the same pass machinery.
extern "C" int emu_run_dm(int nqubits, const bt_ir_op* ops, int nops, int ntop, const double* pool, int64_t npool,
                          int T, int cmin, double* rho, int64_t* stats) {
  btc::CircuitHost c, dm;
  int rc = btc::analyze(nqubits, ops, nops, ntop, pool, npool, &c, &g_msg);
  if (rc) return rc;
  rc = btc::lower_dm(c, &dm, &g_msg);
  if (rc) return rc;
  int64_t nout = 0;
  int dummy = 0;
  return run_host(dm, 2 * nqubits, 0, 0, T, cmin, 0, rho, &dummy, 0, &nout, stats, 1e-9);  // no norm folding: tr rho, not |vec rho|, is 1
}

// The fusion planner of the fused executor: blocks (k, bits MSB first,
// 2^k x 2^k complex128 column-major) written to out_k[b], out_bits[5 b ..],
// out_mats (1024 complex per block).  Returns the block count or < 0.
extern "C" int64_t emu_fuse(int nbits, const bt_ir_op* ops, int nops, int ntop, const double* pool, int64_t npool,
                            int max_k, int pad_to, int* out_k, int* out_bits, double* out_mats, int64_t cap) {
  btc::CircuitHost c;
  int rc = btc::analyze(nbits, ops, nops, ntop, pool, npool, &c, &g_msg);
  if (rc) return -rc;
  std::vector<btf::Block> blocks;
  rc = btf::fuse_gates(c, max_k, pad_to, &blocks, &g_msg);
  if (rc) return -rc;
  if ((int64_t)blocks.size() > cap) return -1000;
  for (size_t b = 0; b < blocks.size(); b++) {
    out_k[b] = blocks[b].k;
    for (int i = 0; i < 5; i++) out_bits[5 * b + i] = blocks[b].bits[i];
    for (size_t e = 0; e < blocks[b].m.size(); e++) {
      out_mats[2 * (1024 * b + e)] = blocks[b].m[e].real();
      out_mats[2 * (1024 * b + e) + 1] = blocks[b].m[e].imag();
    }
  }
  return (int64_t)blocks.size();
}
