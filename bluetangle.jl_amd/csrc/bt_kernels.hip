// Streaming HIP kernels for gfx950: in-place k-bit operators on a state
// vector (or on vec(rho)), FP64 reductions, probabilities and helpers.
//
// Every operator pass is one read and one write of the 2^K amplitudes a
// thread owns (64-wide wavefronts, consecutive lanes on consecutive groups so
// that the free low bits give coalesced addresses); reductions accumulate in
// FP64 and finish in a fixed order, so they are bit-reproducible.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "bt_internal.hpp"

namespace bt {

template <typename R, int K>
struct ApplyArgs {
  static constexpr int D = 1 << K;
  static constexpr int kInline = (K <= 3) ? D * D : 1;
  C<R>* s;
  uint64_t ngroups;
  uint64_t ctrl;
  uint64_t off[D];
  int nfixed;
  int fixed[kMaxFixed];
  R scale;
  const C<R>* mptr;  // K >= 4: device copy of the matrix (column-major)
  C<R> m[kInline];   // K <= 3: matrix inline in the kernel arguments
};

template <typename R, int K>
__global__ void __launch_bounds__(256) k_apply(ApplyArgs<R, K> a) {
  constexpr int D = 1 << K;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < a.ngroups; g += stride) {
    uint64_t base = g;
    for (int i = 0; i < a.nfixed; i++) base = insert0(base, a.fixed[i]);
    base |= a.ctrl;
    C<R> v[D];
#pragma unroll
    for (int l = 0; l < D; l++) v[l] = a.s[base + a.off[l]];
#pragma unroll
    for (int r = 0; r < D; r++) {
      C<R> acc = cmake<C<R>>(0, 0);
#pragma unroll
      for (int c = 0; c < D; c++) {
        const C<R> mrc = (K <= 3) ? a.m[r + c * D] : a.mptr[r + c * D];
        acc = cfma(acc, mrc, v[c]);
      }
      acc.x *= a.scale;
      acc.y *= a.scale;
      a.s[base + a.off[r]] = acc;
    }
  }
}

// One-qubit operator on a low target bit b <= 6 of a complex64 state: every
// lane takes two consecutive amplitudes with one 16-byte access, so a
// wavefront reads and writes 1 KiB contiguous per instruction whatever the
// target (the generic kernel above gives each lane one amplitude of the pair,
// 8 bytes every 2^(b+1) x 8 bytes, for low b).  b = 0: the pair is in the
// lane; b >= 1: the partner amplitudes are in lane ^ 2^(b-1) of the same
// wavefront, exchanged with cross-lane shuffles.
__global__ void __launch_bounds__(256) k_apply1_low(float4* __restrict__ s, uint64_t nvec, int b, float2 u00,
                                                    float2 u01, float2 u10, float2 u11, float scale) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  // every lane of a wavefront runs the same trips (nvec is a multiple of 64),
  // so the shuffles below always find their partner lane active
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride) {
    const float4 v = s[i];
    const float2 x0 = make_float2(v.x, v.y), x1 = make_float2(v.z, v.w);
    float2 y0, y1;
    if (b == 0) {
      y0 = cfma(cmul(u00, x0), u01, x1);
      y1 = cfma(cmul(u10, x0), u11, x1);
    } else {
      const int m = 1 << (b - 1);
      const float2 p0 = make_float2(__shfl_xor(x0.x, m, 64), __shfl_xor(x0.y, m, 64));
      const float2 p1 = make_float2(__shfl_xor(x1.x, m, 64), __shfl_xor(x1.y, m, 64));
      // this lane holds the amplitudes with target bit = (i >> (b - 1)) & 1
      const bool hi = (i >> (b - 1)) & 1;
      const float2 cs = hi ? u11 : u00, co = hi ? u10 : u01;
      y0 = cfma(cmul(cs, x0), co, p0);
      y1 = cfma(cmul(cs, x1), co, p1);
    }
    s[i] = make_float4(y0.x * scale, y0.y * scale, y1.x * scale, y1.y * scale);
  }
}

// Uncontrolled two-qubit operator on low target bits (both <= 6) of a
// complex64 state: a wavefront reads 64 consecutive 16-byte pairs of
// amplitudes (a 128-amplitude segment that holds whole groups) into LDS,
// every lane forms its two outputs from the segment, and writes them back
// with one 16-byte access, so reads and writes stay 1 KiB contiguous per
// wavefront whatever the targets.  The workgroup walks 256-pair blocks in a
// uniform trip count (nvec is a multiple of 256).
__global__ void __launch_bounds__(256) k_apply2_low(float4* __restrict__ s, uint64_t nvec, ApplyArgs<float, 2> a) {
  __shared__ float2 seg[4][128];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t mask = (uint32_t)a.off[3];
  for (uint64_t b0 = (uint64_t)blockIdx.x * 256; b0 < nvec; b0 += (uint64_t)gridDim.x * 256) {
    const float4 v = s[b0 + threadIdx.x];
    seg[w][2 * lane] = make_float2(v.x, v.y);
    seg[w][2 * lane + 1] = make_float2(v.z, v.w);
    __syncthreads();
    float2 y[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const uint32_t i = 2u * (uint32_t)lane + (uint32_t)e;  // index in the segment
      const uint32_t rest = i & ~mask;
      int r = 0;
#pragma unroll
      for (int l = 1; l < 4; l++)
        if ((i & mask) == (uint32_t)a.off[l]) r = l;
      float2 acc = cmake<float2>(0, 0);
#pragma unroll
      for (int c = 0; c < 4; c++) acc = cfma(acc, a.m[r + c * 4], seg[w][rest | (uint32_t)a.off[c]]);
      y[e] = acc;
    }
    s[b0 + threadIdx.x] = make_float4(y[0].x * a.scale, y[0].y * a.scale, y[1].x * a.scale, y[1].y * a.scale);
    __syncthreads();  // the segment is rewritten on the next trip
  }
}

// Uncontrolled two-qubit operator on a complex64 state whose target bits
// are both >= 1: every lane takes the two groups that differ in amplitude
// bit 0 (consecutive amplitudes) with one 16-byte access per matrix index,
// so a wavefront moves 1 KiB contiguous per access instead of 512 B.  The
// launcher uses it where that measured faster (both targets in [9, 24]:
// 5.4-5.9 TB/s against 4.3-5.2 for one group per lane on a 2^30-amplitude
// state; below and above that range one group per lane is as fast or
// faster).
__global__ void __launch_bounds__(256) k_apply2_pair(ApplyArgs<float, 2> a) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t nh = a.ngroups / 2;
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < nh; h += stride) {
    uint64_t base = 2 * h;  // bit 0 stays clear: every inserted bit is >= 1
    for (int i = 0; i < a.nfixed; i++) base = insert0(base, a.fixed[i]);
    float2 x[4], y[4];
#pragma unroll
    for (int l = 0; l < 4; l++) {
      const float4 w = *(const float4*)(a.s + base + a.off[l]);
      x[l] = make_float2(w.x, w.y);
      y[l] = make_float2(w.z, w.w);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float2 u = cmake<float2>(0, 0), v = cmake<float2>(0, 0);
#pragma unroll
      for (int c = 0; c < 4; c++) {
        u = cfma(u, a.m[r + c * 4], x[c]);
        v = cfma(v, a.m[r + c * 4], y[c]);
      }
      *(float4*)(a.s + base + a.off[r]) = make_float4(u.x * a.scale, u.y * a.scale, v.x * a.scale, v.y * a.scale);
    }
  }
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block-level reduction of M per-thread values; thread j < M writes the
// block sum of slot j to out[j].
template <int M>
__device__ __forceinline__ void block_reduce_store(double (&acc)[M], double* out) {
  __shared__ double red[4][M];  // blockDim.x == 256 -> 4 waves
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < M; j++) {
    double v = wave_sum(acc[j]);
    if (lane == 0) red[wave][j] = v;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < M; j += blockDim.x) out[j] = red[0][j] + red[1][j] + red[2][j] + red[3][j];
}

template <typename R, int K>
struct ReduceArgs {
  static constexpr int D = 1 << K;
  const C<R>* s;
  uint64_t ngroups;
  uint64_t off[D];
  int nfixed;
  int fixed[kMaxFixed];
  double* part;
};

// Packed layout of a reduced rho (D*D reals): [0, D) diagonal, then for
// each a < b (row-major over a): Re, Im of rho[a, b] = sum v_a conj(v_b).
// Items per loop trip of the reductions: their loads are issued together
// (more bytes in flight per wave); an item past the end re-reads the trip's
// first item and contributes nothing.
template <int K>
constexpr int red_unroll() { return K == 0 ? 4 : K == 1 ? 2 : 1; }

// The work items [lo, hi) of this workgroup in a reduction over n items:
// one contiguous range per workgroup, a whole number of loop trips of
// U * blockDim.x items each (the last ranges may be short or empty).  A trip
// reads U * blockDim.x consecutive items, and the workgroups that run at
// once stream separate long runs of the state: 6.1-6.2 TB/s for one read of
// a 2^30-amplitude complex64 state, against 5.5-5.6 TB/s for the same loads
// strided over the whole grid.
__device__ __forceinline__ void red_range(uint64_t n, int U, uint64_t& lo, uint64_t& hi) {
  const uint64_t trip = (uint64_t)U * blockDim.x;
  const uint64_t per = ((n + gridDim.x - 1) / gridDim.x + trip - 1) / trip * trip;
  lo = std::min<uint64_t>(n, (uint64_t)blockIdx.x * per);
  hi = std::min<uint64_t>(n, lo + per);
}

template <typename R, int K>
__global__ void __launch_bounds__(256) k_reduce(ReduceArgs<R, K> a) {
  constexpr int D = 1 << K;
  constexpr int M = D * D;
  constexpr int U = red_unroll<K>();
  double acc[M];
#pragma unroll
  for (int j = 0; j < M; j++) acc[j] = 0.0;
  uint64_t lo, hi;
  red_range(a.ngroups, U, lo, hi);
  const uint64_t stride = blockDim.x;
  for (uint64_t g0 = lo + threadIdx.x; g0 < hi; g0 += U * stride) {
    C<R> v[U][D];
    bool live[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t g = g0 + u * stride;
      live[u] = g < hi;
      uint64_t base = live[u] ? g : g0;
      for (int i = 0; i < a.nfixed; i++) base = insert0(base, a.fixed[i]);
#pragma unroll
      for (int l = 0; l < D; l++) v[u][l] = a.s[base + a.off[l]];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (!live[u]) continue;
      double vr[D], vi[D];
#pragma unroll
      for (int l = 0; l < D; l++) {
        vr[l] = (double)v[u][l].x;
        vi[l] = (double)v[u][l].y;
      }
      int p = D;
#pragma unroll
      for (int x = 0; x < D; x++) {
        acc[x] = fma(vr[x], vr[x], fma(vi[x], vi[x], acc[x]));
#pragma unroll
        for (int y = x + 1; y < D; y++) {
          // v_x conj(v_y)
          acc[p] += vr[x] * vr[y] + vi[x] * vi[y];
          acc[p + 1] += vi[x] * vr[y] - vr[x] * vi[y];
          p += 2;
        }
      }
    }
  }
  block_reduce_store<M>(acc, a.part + (size_t)blockIdx.x * M);
}

// out[j] = sum over blocks of part[b * M + j], one workgroup per j: each
// thread sums a fixed stride of blocks, then a fixed tree (bit-reproducible).
__global__ void __launch_bounds__(256) k_finalize(const double* part, int nblocks, int M, double* out) {
  __shared__ double sh[256];
  const int j = blockIdx.x;
  double t = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += 256) t += part[(size_t)b * M + j];
  sh[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = sh[0];
}

// <a|b> = sum conj(a_i) b_i in FP64 (either precision on either side):
// per-workgroup partials (re, im)
template <typename RA, typename RB>
__global__ void __launch_bounds__(256) k_inner(const C<RA>* a, const C<RB>* b, uint64_t n, double* part) {
  constexpr int U = 4;
  double acc[2] = {0.0, 0.0};
  uint64_t lo, hi;
  red_range(n, U, lo, hi);
  const uint64_t stride = blockDim.x;
  for (uint64_t i0 = lo + threadIdx.x; i0 < hi; i0 += U * stride) {
    C<RA> x[U];
    C<RB> y[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t i = i0 + u * stride < hi ? i0 + u * stride : i0;
      x[u] = a[i];
      y[u] = b[i];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i0 + u * stride >= hi) continue;
      acc[0] = fma((double)x[u].x, (double)y[u].x, fma((double)x[u].y, (double)y[u].y, acc[0]));
      acc[1] = fma((double)x[u].x, (double)y[u].y, fma(-(double)x[u].y, (double)y[u].x, acc[1]));
    }
  }
  block_reduce_store<2>(acc, part + (size_t)blockIdx.x * 2);
}

// Expectation of O = P(x, z) G(B): P(x, z)[a, b] = (-1)^popcount(a & z) when
// b = a ^ x, G a dense 2^K x 2^K block on the bits B (disjoint from x and z),
// identity elsewhere.  Work item g enumerates the indices a0 with the B bits
// clear and sums over a = a0 | off(i), b = (a0 ^ x) | off(j):
//   SV: conj(psi_a) O[a, b] psi_b       DM: O[a, b] rho[b + a 2^n]
// On a state vector with x != 0, a0 also has the top bit of x clear and the
// item takes both a0 and a0 ^ x, so each amplitude is read once.
// Per-workgroup FP64 partials (re, im).
template <typename R, int D>
__device__ __forceinline__ void expect_quad(const C<R> (&u)[D], const C<R> (&v)[D], const double* g, double& sr,
                                            double& si) {
#pragma unroll
  for (int i = 0; i < D; i++) {
    double tr = 0.0, ti = 0.0;  // (G v)_i
#pragma unroll
    for (int j = 0; j < D; j++) {
      const double gr = g[2 * (i + j * D)], gi = g[2 * (i + j * D) + 1];
      tr = fma(gr, (double)v[j].x, fma(-gi, (double)v[j].y, tr));
      ti = fma(gr, (double)v[j].y, fma(gi, (double)v[j].x, ti));
    }
    sr = fma((double)u[i].x, tr, fma((double)u[i].y, ti, sr));
    si = fma((double)u[i].x, ti, fma(-(double)u[i].y, tr, si));
  }
}

template <typename R, int K, bool DM>
__global__ void __launch_bounds__(256) k_expect(const C<R>* s, ExpectArgs a, double* part) {
  constexpr int D = 1 << K;
  constexpr int U = DM ? 1 : red_unroll<K>();
  double acc[2] = {0.0, 0.0};
  uint64_t lo, hi;
  red_range(a.ngroups, U, lo, hi);
  const uint64_t stride = blockDim.x;
  for (uint64_t g0 = lo + threadIdx.x; g0 < hi; g0 += U * stride) {
    if constexpr (!DM) {
      C<R> u[U][D], v[U][D];
      uint64_t a0[U];
#pragma unroll
      for (int t = 0; t < U; t++) {
        const uint64_t g = g0 + t * stride < hi ? g0 + t * stride : g0;
        a0[t] = g;
        for (int j = 0; j < a.nfixed; j++) a0[t] = insert0(a0[t], a.fixed[j]);
        const uint64_t b0 = a0[t] ^ a.xmask;
#pragma unroll
        for (int j = 0; j < D; j++) v[t][j] = s[b0 | a.off[j]];
        if (a.xmask) {
#pragma unroll
          for (int j = 0; j < D; j++) u[t][j] = s[a0[t] | a.off[j]];
        }
      }
#pragma unroll
      for (int t = 0; t < U; t++) {
        if (g0 + t * stride >= hi) continue;
        if (!a.xmask) {
#pragma unroll
          for (int j = 0; j < D; j++) u[t][j] = v[t][j];
        }
        double sr = 0.0, si = 0.0;
        expect_quad<R, D>(u[t], v[t], a.g, sr, si);
        const bool neg = __popcll(a0[t] & a.zmask) & 1;
        acc[0] += neg ? -sr : sr;
        acc[1] += neg ? -si : si;
        if (a.xmask) {  // the partner item a0 ^ x
          double tr = 0.0, ti = 0.0;
          expect_quad<R, D>(v[t], u[t], a.g, tr, ti);
          const bool negb = __popcll((a0[t] ^ a.xmask) & a.zmask) & 1;
          acc[0] += negb ? -tr : tr;
          acc[1] += negb ? -ti : ti;
        }
      }
    } else {
      uint64_t a0 = g0;
      for (int j = 0; j < a.nfixed; j++) a0 = insert0(a0, a.fixed[j]);
      const uint64_t b0 = a0 ^ a.xmask;
      double sr = 0.0, si = 0.0;
#pragma unroll
      for (int i = 0; i < D; i++) {
        const uint64_t col = (a0 | a.off[i]) << a.n;
#pragma unroll
        for (int j = 0; j < D; j++) {
          const C<R> r = s[(b0 | a.off[j]) + col];
          const double gr = a.g[2 * (i + j * D)], gi = a.g[2 * (i + j * D) + 1];
          sr = fma(gr, (double)r.x, fma(-gi, (double)r.y, sr));
          si = fma(gr, (double)r.y, fma(gi, (double)r.x, si));
        }
      }
      const bool neg = __popcll(a0 & a.zmask) & 1;
      acc[0] += neg ? -sr : sr;
      acc[1] += neg ? -si : si;
    }
  }
  block_reduce_store<2>(acc, part + (size_t)blockIdx.x * 2);
}

// sum over i of Re(s[i * step]) for i < count (trace of vec(rho): step = D+1)
template <typename R>
__global__ void __launch_bounds__(256) k_strided_re_sum(const C<R>* s, uint64_t count, uint64_t step, double* part) {
  double acc[1] = {0.0};
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
    acc[0] += (double)s[i * step].x;
  block_reduce_store<1>(acc, part + blockIdx.x);
}

template <typename R>
__global__ void k_probabilities(const C<R>* s, uint64_t count, uint64_t step, double* out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    C<R> v = s[i * step];
    out[i] = step == 1 ? (double)v.x * (double)v.x + (double)v.y * (double)v.y : (double)v.x;
  }
}

template <typename R>
__global__ void k_scale(C<R>* s, uint64_t n, R f) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    C<R> v = s[i];
    v.x *= f;
    v.y *= f;
    s[i] = v;
  }
}

template <typename R>
__global__ void k_fill_basis(C<R>* s, uint64_t n, uint64_t index) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    s[i] = cmake<C<R>>(i == index ? (R)1 : (R)0, (R)0);
}

template <typename R>
__global__ void k_gather(const C<R>* s, uint64_t len, const uint64_t* idx, uint64_t n, double* out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint64_t j = idx[i];
    C<R> v = j < len ? s[j] : cmake<C<R>>(0, 0);
    out[2 * i] = (double)v.x;
    out[2 * i + 1] = (double)v.y;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// Workgroups of a reduction with M partial values each over `work` items:
// 16384 for large states (each a contiguous range, see red_range: 6.2 TB/s
// against 6.1 at 4096 and 6.0 at 2048 for one read of 2^30 complex64
// amplitudes); kRedBlocks below 2^26 items, where the launch and the final
// sum of the partials dominate; always within the partial buffer of
// kRedBlocks * kRedMax values.
static uint64_t red_cap(int M, uint64_t work) {
  const uint64_t c = (uint64_t)kRedBlocks * kRedMax / (uint64_t)M;
  const uint64_t want = work >= (1ull << 26) ? 16384 : kRedBlocks;
  return c < want ? c : want;
}

static unsigned grid_for(uint64_t work, unsigned block = 256, uint64_t cap = (1ull << 20)) {
  uint64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

int make_placement(int nbits, int k, const int* bits, uint64_t ctrl_mask, Placement* p) {
  if (k < 0 || k > kMaxK) return fail(BT_ERR_ARG, "operator width k out of range");
  uint64_t used = 0;
  for (int i = 0; i < k; i++) {
    if (bits[i] < 0 || bits[i] >= nbits) return fail(BT_ERR_ARG, "bit position out of range");
    if (used & (1ull << bits[i])) return fail(BT_ERR_ARG, "repeated bit position");
    used |= 1ull << bits[i];
  }
  if (nbits < 64 && (ctrl_mask >> nbits)) return fail(BT_ERR_ARG, "control bit out of range");
  if (ctrl_mask & used) return fail(BT_ERR_ARG, "control bit equals a target bit");
  uint64_t all = used | ctrl_mask;
  p->nfixed = 0;
  for (int b = 0; b < nbits; b++)
    if (all & (1ull << b)) {
      if (p->nfixed >= kMaxFixed) return fail(BT_ERR_ARG, "too many target/control bits");
      p->fixed[p->nfixed++] = b;
    }
  p->ctrl = ctrl_mask;
  for (int l = 0; l < (1 << k); l++) {
    uint64_t o = 0;
    for (int i = 0; i < k; i++)
      if ((l >> (k - 1 - i)) & 1) o |= 1ull << bits[i];
    p->off[l] = o;
  }
  return BT_OK;
}

template <typename R, int K>
static int apply_t(bt_state_s* s, int nbits, const double* m, const Placement& p, double scale, bool conj_mat) {
  constexpr int D = 1 << K;
  ApplyArgs<R, K> a;
  memset(&a, 0, sizeof a);
  a.s = (C<R>*)s->d;
  a.ngroups = 1ull << (nbits - p.nfixed);
  a.ctrl = p.ctrl;
  for (int l = 0; l < D; l++) a.off[l] = p.off[l];
  a.nfixed = p.nfixed;
  for (int i = 0; i < p.nfixed; i++) a.fixed[i] = p.fixed[i];
  a.scale = (R)scale;
  const double sgn = conj_mat ? -1.0 : 1.0;
  if (K <= 3) {
    for (int i = 0; i < D * D; i++) a.m[i] = cmake<C<R>>((R)m[2 * i], (R)(sgn * m[2 * i + 1]));
    a.mptr = nullptr;
  } else {
    std::vector<C<R>> hm(D * D);
    for (int i = 0; i < D * D; i++) hm[i] = cmake<C<R>>((R)m[2 * i], (R)(sgn * m[2 * i + 1]));
    if (!s->d_mat) BT_HIP(hipMalloc(&s->d_mat, sizeof(C<double>) * 1024));
    // the matrix must not change while an earlier launch may still read it
    BT_HIP(hipStreamSynchronize(s->stream));
    BT_HIP(hipMemcpy(s->d_mat, hm.data(), sizeof(C<R>) * D * D, hipMemcpyHostToDevice));
    a.mptr = (const C<R>*)s->d_mat;
  }
  if constexpr (K == 1 && sizeof(R) == 4) {
    // uncontrolled one-qubit operator on a low bit: the coalesced kernel
    if (p.ctrl == 0 && p.nfixed == 1 && p.fixed[0] <= 6 && nbits >= 7) {
      const uint64_t nvec = 1ull << (nbits - 1);  // two amplitudes per lane
      hipLaunchKernelGGL(k_apply1_low, dim3(grid_for(nvec)), dim3(256), 0, s->stream, (float4*)s->d, nvec,
                         p.fixed[0], a.m[0], a.m[2], a.m[1], a.m[3], (float)scale);
      BT_HIP(hipGetLastError());
      return BT_OK;
    }
  }
  if constexpr (K == 2 && sizeof(R) == 4) {
    // uncontrolled, both targets low: staged through LDS per wavefront
    if (p.ctrl == 0 && p.nfixed == 2 && p.fixed[1] <= 6 && nbits >= 9) {
      const uint64_t nvec = 1ull << (nbits - 1);  // 16-byte pairs, a multiple of 256
      hipLaunchKernelGGL(k_apply2_low, dim3(grid_for(nvec)), dim3(256), 0, s->stream, (float4*)s->d, nvec, a);
      BT_HIP(hipGetLastError());
      return BT_OK;
    }
    // uncontrolled, both targets in the middle bit range: two groups per lane
    if (p.ctrl == 0 && p.nfixed == 2 && p.fixed[0] >= 9 && p.fixed[1] <= 24) {
      hipLaunchKernelGGL(k_apply2_pair, dim3(grid_for(a.ngroups / 2)), dim3(256), 0, s->stream, a);
      BT_HIP(hipGetLastError());
      return BT_OK;
    }
  }
  hipLaunchKernelGGL((k_apply<R, K>), dim3(grid_for(a.ngroups)), dim3(256), 0, s->stream, a);
  BT_HIP(hipGetLastError());
  return BT_OK;
}

int launch_apply(bt_state_s* s, int nbits, int k, const double* m, const Placement& p, double scale,
                 bool conj_mat) {
  if (nbits - p.nfixed < 0) return fail(BT_ERR_ARG, "placement wider than the state");
  const bool f = s->prec == BT_C64;
  switch (k) {
    case 1: return f ? apply_t<float, 1>(s, nbits, m, p, scale, conj_mat) : apply_t<double, 1>(s, nbits, m, p, scale, conj_mat);
    case 2: return f ? apply_t<float, 2>(s, nbits, m, p, scale, conj_mat) : apply_t<double, 2>(s, nbits, m, p, scale, conj_mat);
    case 3: return f ? apply_t<float, 3>(s, nbits, m, p, scale, conj_mat) : apply_t<double, 3>(s, nbits, m, p, scale, conj_mat);
    case 4: return f ? apply_t<float, 4>(s, nbits, m, p, scale, conj_mat) : apply_t<double, 4>(s, nbits, m, p, scale, conj_mat);
    case 5: return f ? apply_t<float, 5>(s, nbits, m, p, scale, conj_mat) : apply_t<double, 5>(s, nbits, m, p, scale, conj_mat);
    default: return fail(BT_ERR_ARG, "operator width k out of range");
  }
}

template <typename R, int K>
static int reduce_t(bt_state_s* s, const Placement& p) {
  constexpr int D = 1 << K;
  constexpr int M = D * D;
  ReduceArgs<R, K> a;
  memset(&a, 0, sizeof a);
  a.s = (const C<R>*)s->d;
  a.ngroups = 1ull << (s->nbits - p.nfixed);
  for (int l = 0; l < D; l++) a.off[l] = p.off[l];
  a.nfixed = p.nfixed;
  for (int i = 0; i < p.nfixed; i++) a.fixed[i] = p.fixed[i];
  a.part = s->d_part;
  const unsigned blocks = grid_for(a.ngroups, 256, red_cap(M, a.ngroups));
  hipLaunchKernelGGL((k_reduce<R, K>), dim3(blocks), dim3(256), 0, s->stream, a);
  BT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_finalize, dim3(M), dim3(256), 0, s->stream, (const double*)s->d_part, (int)blocks, M, s->d_out);
  BT_HIP(hipGetLastError());
  BT_HIP(hipMemcpyAsync(s->h_out, s->d_out, sizeof(double) * M, hipMemcpyDeviceToHost, s->stream));
  BT_HIP(hipStreamSynchronize(s->stream));
  return BT_OK;
}

int launch_inner(bt_state_s* a, bt_state_s* b, double* re, double* im) {
  const uint64_t n = a->len;
  const unsigned blocks = grid_for(n, 256, red_cap(2, n));
  const bool fa = a->prec == BT_C64, fb = b->prec == BT_C64;
  if (fa && fb)
    hipLaunchKernelGGL((k_inner<float, float>), dim3(blocks), dim3(256), 0, a->stream, (const C<float>*)a->d,
                       (const C<float>*)b->d, n, a->d_part);
  else if (fa)
    hipLaunchKernelGGL((k_inner<float, double>), dim3(blocks), dim3(256), 0, a->stream, (const C<float>*)a->d,
                       (const C<double>*)b->d, n, a->d_part);
  else if (fb)
    hipLaunchKernelGGL((k_inner<double, float>), dim3(blocks), dim3(256), 0, a->stream, (const C<double>*)a->d,
                       (const C<float>*)b->d, n, a->d_part);
  else
    hipLaunchKernelGGL((k_inner<double, double>), dim3(blocks), dim3(256), 0, a->stream, (const C<double>*)a->d,
                       (const C<double>*)b->d, n, a->d_part);
  BT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_finalize, dim3(2), dim3(256), 0, a->stream, (const double*)a->d_part, (int)blocks, 2, a->d_out);
  BT_HIP(hipGetLastError());
  BT_HIP(hipMemcpyAsync(a->h_out, a->d_out, sizeof(double) * 2, hipMemcpyDeviceToHost, a->stream));
  BT_HIP(hipStreamSynchronize(a->stream));
  *re = a->h_out[0];
  *im = a->h_out[1];
  return BT_OK;
}

int launch_reduce(bt_state_s* s, int k, const int* bits) {
  Placement p;
  int rc = make_placement(s->nbits, k, bits, 0, &p);
  if (rc) return rc;
  const bool f = s->prec == BT_C64;
  switch (k) {
    case 0: return f ? reduce_t<float, 0>(s, p) : reduce_t<double, 0>(s, p);
    case 1: return f ? reduce_t<float, 1>(s, p) : reduce_t<double, 1>(s, p);
    case 2: return f ? reduce_t<float, 2>(s, p) : reduce_t<double, 2>(s, p);
    case 3: return f ? reduce_t<float, 3>(s, p) : reduce_t<double, 3>(s, p);
    default: return fail(BT_ERR_ARG, "reduction width k out of range");
  }
}

void unpack_reduced(int k, const double* packed, double* rho) {
  const int D = 1 << k;
  for (int i = 0; i < 2 * D * D; i++) rho[i] = 0.0;
  for (int a = 0; a < D; a++) rho[2 * (a + a * D)] = packed[a];
  int p = D;
  for (int a = 0; a < D; a++)
    for (int b = a + 1; b < D; b++) {
      // rho[a, b] = re + i im ; rho[b, a] = conj
      rho[2 * (a + b * D)] = packed[p];
      rho[2 * (a + b * D) + 1] = packed[p + 1];
      rho[2 * (b + a * D)] = packed[p];
      rho[2 * (b + a * D) + 1] = -packed[p + 1];
      p += 2;
    }
}

int launch_norm2(bt_state_s* s, double* out) {
  int rc = launch_reduce(s, 0, nullptr);
  if (rc) return rc;
  *out = s->h_out[0];
  return BT_OK;
}

int launch_trace(bt_state_s* s, double* out) {
  const uint64_t D = 1ull << s->n;
  const unsigned blocks = grid_for(D, 256, kRedBlocks);
  if (s->prec == BT_C64)
    hipLaunchKernelGGL(k_strided_re_sum<float>, dim3(blocks), dim3(256), 0, s->stream, (const C<float>*)s->d, D, D + 1,
                       s->d_part);
  else
    hipLaunchKernelGGL(k_strided_re_sum<double>, dim3(blocks), dim3(256), 0, s->stream, (const C<double>*)s->d, D,
                       D + 1, s->d_part);
  BT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s->stream, (const double*)s->d_part, (int)blocks, 1, s->d_out);
  BT_HIP(hipGetLastError());
  BT_HIP(hipMemcpyAsync(s->h_out, s->d_out, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  BT_HIP(hipStreamSynchronize(s->stream));
  *out = s->h_out[0];
  return BT_OK;
}

int launch_probabilities(bt_state_s* s, double* d_out) {
  const uint64_t D = 1ull << s->n;
  const uint64_t step = s->kind == BT_DM ? D + 1 : 1;
  if (s->prec == BT_C64)
    hipLaunchKernelGGL(k_probabilities<float>, dim3(grid_for(D)), dim3(256), 0, s->stream, (const C<float>*)s->d, D,
                       step, d_out);
  else
    hipLaunchKernelGGL(k_probabilities<double>, dim3(grid_for(D)), dim3(256), 0, s->stream, (const C<double>*)s->d, D,
                       step, d_out);
  BT_HIP(hipGetLastError());
  return BT_OK;
}

int launch_scale(bt_state_s* s, double f) {
  if (s->prec == BT_C64)
    hipLaunchKernelGGL(k_scale<float>, dim3(grid_for(s->len)), dim3(256), 0, s->stream, (C<float>*)s->d, s->len,
                       (float)f);
  else
    hipLaunchKernelGGL(k_scale<double>, dim3(grid_for(s->len)), dim3(256), 0, s->stream, (C<double>*)s->d, s->len, f);
  BT_HIP(hipGetLastError());
  return BT_OK;
}

int launch_set_basis(bt_state_s* s, uint64_t index) {
  if (s->prec == BT_C64)
    hipLaunchKernelGGL(k_fill_basis<float>, dim3(grid_for(s->len)), dim3(256), 0, s->stream, (C<float>*)s->d, s->len,
                       index);
  else
    hipLaunchKernelGGL(k_fill_basis<double>, dim3(grid_for(s->len)), dim3(256), 0, s->stream, (C<double>*)s->d,
                       s->len, index);
  BT_HIP(hipGetLastError());
  return BT_OK;
}

int launch_gather(bt_state_s* s, const uint64_t* d_idx, uint64_t n, double* d_out) {
  if (s->prec == BT_C64)
    hipLaunchKernelGGL(k_gather<float>, dim3(grid_for(n)), dim3(256), 0, s->stream, (const C<float>*)s->d, s->len,
                       d_idx, n, d_out);
  else
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n)), dim3(256), 0, s->stream, (const C<double>*)s->d, s->len,
                       d_idx, n, d_out);
  BT_HIP(hipGetLastError());
  return BT_OK;
}

template <typename R, int K>
static void expect_t(bt_state_s* s, const ExpectArgs& a, unsigned blocks) {
  if (s->kind == BT_DM)
    hipLaunchKernelGGL((k_expect<R, K, true>), dim3(blocks), dim3(256), 0, s->stream, (const C<R>*)s->d, a, s->d_part);
  else
    hipLaunchKernelGGL((k_expect<R, K, false>), dim3(blocks), dim3(256), 0, s->stream, (const C<R>*)s->d, a,
                       s->d_part);
}

int launch_expect(bt_state_s* s, int k, const double* mat, const int* bits, uint64_t xmask, uint64_t zmask,
                  double* re, double* im) {
  if (k < 0 || k > 3) return fail(BT_ERR_ARG, "expectation: dense block width k must be 0..3");
  const int n = s->n;
  const uint64_t all = n == 64 ? ~0ull : ((1ull << n) - 1ull);
  if ((xmask | zmask) & ~all) return fail(BT_ERR_ARG, "expectation: Pauli mask outside the register");
  ExpectArgs a;
  std::memset(&a, 0, sizeof(a));
  a.n = n;
  a.xmask = xmask;
  a.zmask = zmask;
  uint64_t used = 0;
  for (int i = 0; i < k; i++) {
    if (bits[i] < 0 || bits[i] >= n) return fail(BT_ERR_ARG, "expectation: bit out of range");
    if ((used >> bits[i]) & 1) return fail(BT_ERR_ARG, "expectation: repeated bit");
    used |= 1ull << bits[i];
  }
  if (used & (xmask | zmask)) return fail(BT_ERR_ARG, "expectation: dense bits overlap the Pauli masks");
  // inserted bits, ascending: the block bits, and on a state vector the top bit of x
  uint64_t ins = used;
  if (s->kind == BT_SV && xmask) ins |= 1ull << (63 - __builtin_clzll(xmask));
  int j = 0;
  for (int b = 0; b < 64; b++)
    if ((ins >> b) & 1) a.fixed[j++] = b;
  a.nfixed = j;
  const int D = 1 << k;
  for (int l = 0; l < D; l++) {
    uint64_t o = 0;
    for (int i = 0; i < k; i++)
      if ((l >> (k - 1 - i)) & 1) o |= 1ull << bits[i];
    a.off[l] = o;
  }
  if (k == 0) {
    a.g[0] = 1.0;
  } else {
    for (int i = 0; i < 2 * D * D; i++) a.g[i] = mat[i];
  }
  a.ngroups = 1ull << (n - a.nfixed);
  const unsigned blocks = grid_for(a.ngroups, 256, red_cap(2, a.ngroups));
  const bool f = s->prec == BT_C64;
  switch (k) {
    case 0: f ? expect_t<float, 0>(s, a, blocks) : expect_t<double, 0>(s, a, blocks); break;
    case 1: f ? expect_t<float, 1>(s, a, blocks) : expect_t<double, 1>(s, a, blocks); break;
    case 2: f ? expect_t<float, 2>(s, a, blocks) : expect_t<double, 2>(s, a, blocks); break;
    default: f ? expect_t<float, 3>(s, a, blocks) : expect_t<double, 3>(s, a, blocks); break;
  }
  BT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_finalize, dim3(2), dim3(256), 0, s->stream, (const double*)s->d_part, (int)blocks, 2, s->d_out);
  BT_HIP(hipGetLastError());
  BT_HIP(hipMemcpyAsync(s->h_out, s->d_out, sizeof(double) * 2, hipMemcpyDeviceToHost, s->stream));
  BT_HIP(hipStreamSynchronize(s->stream));
  *re = s->h_out[0];
  *im = s->h_out[1];
  return BT_OK;
}

}  // namespace bt
