// Shot sampling from a state vector's basis-state probabilities p_i = |psi_i|^2
// (the final sampling of measure / sample, ops.jl:46-60,599-646), with
// explicit uniforms in place of StatsBase's RNG.  Rule: shot s returns the
// smallest i with u_s T < C_i, where C_i is the FP64 cumulative sum in index
// order and T = C_{2^N - 1} (so the state need not be normalised).
//
// Three kernels on the caller's stream:
//   k_chunk_sums   FP64 sum of each chunk of 2048 probabilities;
//   k_chunk_scan   exclusive prefix of the chunk sums (one workgroup, fixed
//                  order), prefix[nchunks] = T;
//   k_sample_find  one wavefront per shot: binary search of the chunk prefix,
//                  then a wavefront scan of the chunk's probabilities (32
//                  consecutive entries per lane).
// Sums are formed in a different order than a sequential cumsum; a shot whose
// u_s T lies within 1e-12 T of a cumulative boundary is flagged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "bt_internal.hpp"

namespace bt {

constexpr int kChunk = 2048;       // probabilities per chunk
constexpr int kLaneRun = kChunk / 64;  // consecutive entries per lane in the search
constexpr double kBoundaryTol = 1e-12;

template <typename R>
__device__ __forceinline__ double prob(const C<R>* s, uint64_t i) {
  const C<R> v = s[i];
  return (double)v.x * (double)v.x + (double)v.y * (double)v.y;
}

// One wavefront per chunk (no workgroup barriers): lane l sums entries
// l, l + 64, ... of the chunk in order, then a fixed butterfly over the
// lanes.  The workgroup's four wavefronts take four consecutive chunks.
template <typename R>
__global__ void __launch_bounds__(256) k_chunk_sums(const C<R>* s, uint64_t n, uint64_t nchunks, double* sums) {
  const int lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t)gridDim.x * 4;
  for (uint64_t c = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6); c < nchunks; c += nw) {
    const uint64_t base = c * kChunk + (uint64_t)lane;
    double t = 0.0;
#pragma unroll
    for (int j = 0; j < kChunk / 64; j++) {
      const uint64_t i = base + 64u * j;
      if (i < n) t += prob<R>(s, i);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (lane == 0) sums[c] = t;
  }
}

// prefix[k] = sum of sums[0..k) for k <= nchunks; one workgroup of 1024 threads,
// thread t owns a contiguous run of chunks.
__global__ void __launch_bounds__(1024) k_chunk_scan(const double* sums, uint64_t nchunks, double* prefix) {
  __shared__ double part[1024];
  const uint64_t per = (nchunks + 1023) / 1024;
  const uint64_t b = threadIdx.x * per, e = std::min<uint64_t>(b + per, nchunks);
  double t = 0.0;
  for (uint64_t k = b; k < e; k++) t += sums[k];
  part[threadIdx.x] = t;
  __syncthreads();
  if (threadIdx.x == 0) {  // exclusive scan of the 1024 run totals, in order
    double run = 0.0;
    for (int i = 0; i < 1024; i++) {
      const double v = part[i];
      part[i] = run;
      run += v;
    }
    prefix[nchunks] = run;
  }
  __syncthreads();
  double run = part[threadIdx.x];
  for (uint64_t k = b; k < e; k++) {
    prefix[k] = run;
    run += sums[k];
  }
}

// One wavefront per shot.  u: device uniforms (or null: every shot takes u0).
template <typename R>
__global__ void __launch_bounds__(256) k_sample_find(const C<R>* s, uint64_t n, const double* prefix,
                                                     uint64_t nchunks, const double* u, double u0, int64_t shots,
                                                     uint64_t* out, uint8_t* flag) {
  const int lane = threadIdx.x & 63;
  const int64_t shot = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (shot >= shots) return;  // whole wavefronts leave together
  const double T = prefix[nchunks];
  const double tgt = (u ? u[shot] : u0) * T;
  // largest k < nchunks with prefix[k] <= tgt
  uint64_t lo = 0, hi = nchunks;
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (prefix[mid] <= tgt) lo = mid;
    else hi = mid;
  }
  const uint64_t base = lo * kChunk + (uint64_t)lane * kLaneRun;
  double p[kLaneRun];
  double mine = 0.0;
#pragma unroll
  for (int j = 0; j < kLaneRun; j++) {
    const uint64_t i = base + j;
    p[j] = i < n ? prob<R>(s, i) : 0.0;
    mine += p[j];
  }
  // inclusive scan of the lane totals in lane order (Hillis-Steele)
  double incl = mine;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double v = __shfl_up(incl, o, 64);
    if (lane >= o) incl += v;
  }
  const double excl = incl - mine;
  const double c0 = prefix[lo];
  // first lane whose run ends past the target; if rounding leaves none, the
  // last lane holding probability
  const uint64_t hit = __ballot(c0 + incl > tgt && mine > 0.0);
  const uint64_t nz = __ballot(mine > 0.0);
  int src;
  if (hit) src = __builtin_ctzll(hit);
  else if (nz) src = 63 - __builtin_clzll(nz);
  else src = -1;
  if (lane == (src < 0 ? 0 : src)) {
    uint64_t idx = 0;
    double lo_c = c0 + excl, hi_c = lo_c;
    if (src >= 0) {
      int pick = -1, last = -1;
      double run = c0 + excl, last_lo = run, last_hi = run;
#pragma unroll
      for (int j = 0; j < kLaneRun; j++) {
        const double nxt = run + p[j];
        if (p[j] > 0.0) {
          last = j;
          last_lo = run;
          last_hi = nxt;
          if (pick < 0 && tgt < nxt) {
            pick = j;
            lo_c = run;
            hi_c = nxt;
          }
        }
        run = nxt;
      }
      if (pick < 0) {  // past the end by rounding: the last entry with probability
        pick = last;
        lo_c = last_lo;
        hi_c = last_hi;
      }
      idx = base + pick;
    }
    out[shot] = idx;
    flag[shot] = (src < 0 || tgt - lo_c < kBoundaryTol * T || hi_c - tgt < kBoundaryTol * T) ? 1 : 0;
  }
}

size_t sample_scratch_bytes(int nbits) {
  const uint64_t n = 1ull << nbits;
  const uint64_t nchunks = (n + kChunk - 1) / kChunk;
  return sizeof(double) * (2 * nchunks + 1);
}

int launch_sample(const void* state, int prec, int nbits, hipStream_t st, void* scratch, int64_t shots,
                  const double* d_u, double u0, uint64_t* d_out, uint8_t* d_flag) {
  const uint64_t n = 1ull << nbits;
  const uint64_t nchunks = (n + kChunk - 1) / kChunk;
  double* sums = (double*)scratch;
  double* prefix = sums + nchunks;
  const unsigned g1 = (unsigned)std::min<uint64_t>((nchunks + 3) / 4, 1u << 16);
  const unsigned g3 = (unsigned)((shots + 3) / 4);
  if (prec == BT_C64) {
    hipLaunchKernelGGL(k_chunk_sums<float>, dim3(g1), dim3(256), 0, st, (const C<float>*)state, n, nchunks, sums);
    hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(1024), 0, st, (const double*)sums, nchunks, prefix);
    hipLaunchKernelGGL(k_sample_find<float>, dim3(g3), dim3(256), 0, st, (const C<float>*)state, n,
                       (const double*)prefix, nchunks, d_u, u0, shots, d_out, d_flag);
  } else {
    hipLaunchKernelGGL(k_chunk_sums<double>, dim3(g1), dim3(256), 0, st, (const C<double>*)state, n, nchunks, sums);
    hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(1024), 0, st, (const double*)sums, nchunks, prefix);
    hipLaunchKernelGGL(k_sample_find<double>, dim3(g3), dim3(256), 0, st, (const C<double>*)state, n,
                       (const double*)prefix, nchunks, d_u, u0, shots, d_out, d_flag);
  }
  BT_HIP(hipGetLastError());
  return BT_OK;
}

// ---------------------------------------------------------------------------
// the no-allocation variant's scratch: one static device array per GPU
// ---------------------------------------------------------------------------
// Holds the chunk sums and prefix, the shots' uniforms, indices and boundary
// flags, and the boundary count: 32 MiB covers a 2^30-amplitude state (8 MiB
// of sums and prefix) with about 1.4 million shots.  In host code the array's
// name is not a device address: kernels and copies take the address that
// hipGetSymbolAddress returned for the current device (noalloc_setup).
constexpr size_t kNoAllocScratch = 32ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratch];
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current (the library's device set-up).
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(BT_ERR_ARG, "noalloc_setup: device index out of range");
  if (g_scratch_ptr[device]) return BT_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(BT_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return BT_OK;
}

// Number of flagged shots, written (not accumulated) to *cnt: one workgroup.
__global__ void __launch_bounds__(256) k_count_flags(const uint8_t* flag, int64_t shots, int64_t* cnt) {
  __shared__ int64_t red[256];
  int64_t t = 0;
  for (int64_t i = threadIdx.x; i < shots; i += 256) t += flag[i];
  red[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *cnt = red[0];
}

}  // namespace bt

using namespace bt;

extern "C" int bt_sample_noalloc(bt_state s, int64_t shots, const double* u, uint64_t* out, int64_t* nboundary) {
  if (!s) return fail(BT_ERR_ARG, "null state");
  if (s->kind != BT_SV) return fail(BT_ERR_UNSUPPORTED, "bt_sample_noalloc: state vectors only");
  if (shots < 0 || (shots > 0 && (!u || !out))) return fail(BT_ERR_ARG, "bt_sample_noalloc: bad arguments");
  if (shots == 0) {
    if (nboundary) *nboundary = 0;
    return BT_OK;
  }
  for (int64_t i = 0; i < shots; i++)
    if (!(u[i] >= 0.0 && u[i] < 1.0)) return fail(BT_ERR_ARG, "bt_sample_noalloc: uniforms must lie in [0, 1)");
  const size_t sc = (sample_scratch_bytes(s->nbits) + 255) / 256 * 256;
  const size_t ub = sizeof(double) * (size_t)shots, ob = sizeof(uint64_t) * (size_t)shots, fb = (size_t)shots;
  const size_t need = sc + ub + ob + (fb + 7) / 8 * 8 + sizeof(int64_t);
  if (need > kNoAllocScratch)
    return fail(BT_ERR_ARG, "bt_sample_noalloc: state and shots need " + std::to_string(need) +
                                " bytes of scratch, more than its " + std::to_string(kNoAllocScratch) +
                                " (use bt_sample)");
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices || !g_scratch_ptr[dev])
    return fail(BT_ERR_STATE, "bt_sample_noalloc: the device's scratch is not set up (bt_init)");
  char* base = (char*)g_scratch_ptr[dev];
  double* d_u = (double*)(base + sc);
  uint64_t* d_out = (uint64_t*)(base + sc + ub);
  uint8_t* d_flag = (uint8_t*)(base + sc + ub + ob);
  int64_t* d_cnt = (int64_t*)(base + sc + ub + ob + (fb + 7) / 8 * 8);
  int64_t nb = 0;
  int rc = BT_OK;
  if (hipMemcpyAsync(d_u, u, ub, hipMemcpyHostToDevice, s->stream) != hipSuccess)
    rc = fail(BT_ERR_HIP, "bt_sample_noalloc: upload");
  if (rc == BT_OK) rc = launch_sample(s->d, s->prec, s->nbits, s->stream, base, shots, d_u, 0.0, d_out, d_flag);
  if (rc == BT_OK) {
    hipLaunchKernelGGL(k_count_flags, dim3(1), dim3(256), 0, s->stream, (const uint8_t*)d_flag, shots, d_cnt);
    if (hipGetLastError() != hipSuccess) rc = fail(BT_ERR_HIP, "bt_sample_noalloc: count launch");
  }
  if (rc == BT_OK && (hipMemcpyAsync(out, d_out, ob, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
                      hipMemcpyAsync(&nb, d_cnt, sizeof nb, hipMemcpyDeviceToHost, s->stream) != hipSuccess))
    rc = fail(BT_ERR_HIP, "bt_sample_noalloc: download");
  if (hipStreamSynchronize(s->stream) != hipSuccess && rc == BT_OK) rc = fail(BT_ERR_HIP, "bt_sample_noalloc: sync");
  if (rc) return rc;
  if (nboundary) *nboundary = nb;
  return BT_OK;
}

extern "C" int bt_sample(bt_state s, int64_t shots, const double* u, uint64_t* out, int64_t* nboundary) {
  if (!s) return fail(BT_ERR_ARG, "null state");
  if (s->kind != BT_SV) return fail(BT_ERR_UNSUPPORTED, "bt_sample: state vectors only (sample, ops.jl:46-60)");
  if (shots < 0 || (shots > 0 && (!u || !out))) return fail(BT_ERR_ARG, "bt_sample: bad arguments");
  if (shots == 0) {
    if (nboundary) *nboundary = 0;
    return BT_OK;
  }
  for (int64_t i = 0; i < shots; i++)
    if (!(u[i] >= 0.0 && u[i] < 1.0)) return fail(BT_ERR_ARG, "bt_sample: uniforms must lie in [0, 1)");
  const size_t sc = sample_scratch_bytes(s->nbits);
  const size_t ub = sizeof(double) * shots, ob = sizeof(uint64_t) * shots, fb = (size_t)shots;
  char* buf = nullptr;
  if (hipMalloc((void**)&buf, sc + ub + ob + fb + 64) != hipSuccess) {
    (void)hipGetLastError();
    return fail(BT_ERR_NOMEM, "bt_sample: device buffers");
  }
  double* d_u = (double*)(buf + sc);
  uint64_t* d_out = (uint64_t*)(buf + sc + ub);
  uint8_t* d_flag = (uint8_t*)(buf + sc + ub + ob);
  std::vector<uint8_t> flags(shots);
  int rc = BT_OK;
  if (hipMemcpyAsync(d_u, u, ub, hipMemcpyHostToDevice, s->stream) != hipSuccess) rc = fail(BT_ERR_HIP, "bt_sample: upload");
  if (rc == BT_OK) rc = launch_sample(s->d, s->prec, s->nbits, s->stream, buf, shots, d_u, 0.0, d_out, d_flag);
  if (rc == BT_OK && (hipMemcpyAsync(out, d_out, ob, hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
                      hipMemcpyAsync(flags.data(), d_flag, fb, hipMemcpyDeviceToHost, s->stream) != hipSuccess))
    rc = fail(BT_ERR_HIP, "bt_sample: download");
  if (hipStreamSynchronize(s->stream) != hipSuccess && rc == BT_OK) rc = fail(BT_ERR_HIP, "bt_sample: sync");
  (void)hipFree(buf);
  if (rc) return rc;
  int64_t nb = 0;
  for (int64_t i = 0; i < shots; i++) nb += flags[i];
  if (nboundary) *nboundary = nb;
  return BT_OK;
}
