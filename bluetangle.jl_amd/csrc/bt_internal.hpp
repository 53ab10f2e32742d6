// Internal declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "bt.h"

namespace bt {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define BT_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) return ::bt::fail(BT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ---------------------------------------------------------------------------
// complex helpers (device and host)
// ---------------------------------------------------------------------------
template <typename R>
struct cx;
template <>
struct cx<float> {
  using t = float2;
};
template <>
struct cx<double> {
  using t = double2;
};
template <typename R>
using C = typename cx<R>::t;

template <typename T>
__host__ __device__ __forceinline__ T cmake(decltype(T().x) a, decltype(T().x) b) {
  T r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__host__ __device__ __forceinline__ T cmul(T a, T b) {
  return cmake<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// acc + a*b
template <typename T>
__host__ __device__ __forceinline__ T cfma(T acc, T a, T b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
  return acc;
}

__host__ __device__ __forceinline__ uint64_t insert0(uint64_t x, int b) {
  uint64_t lo = x & ((1ull << b) - 1ull);
  return ((x >> b) << (b + 1)) | lo;
}

// Counter-based uniform stream: u(seed, traj, draw) in [0, 1) with 53 bits.
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__host__ __device__ __forceinline__ double uniform(uint64_t seed, uint64_t traj, uint64_t draw) {
  uint64_t z = splitmix64(seed ^ splitmix64(traj ^ splitmix64(draw)));
  return (double)(z >> 11) * 0x1.0p-53;
}

// ---------------------------------------------------------------------------
// placement of a k-bit operator: fixed bit positions and per-element offsets
// ---------------------------------------------------------------------------
constexpr int kMaxK = 5;      // dense SV blocks up to 32 x 32
constexpr int kMaxFixed = 12; // targets + controls

struct Placement {
  int nfixed;                   // number of inserted (zeroed) positions, ascending
  int fixed[kMaxFixed];
  uint64_t ctrl;                // OR-ed into every base index
  uint64_t off[1 << kMaxK];     // offset of local index l
};

// Build from target bits (bits[0] = MSB of local index) and a control mask.
int make_placement(int nbits, int k, const int* bits, uint64_t ctrl_mask, Placement* p);

// ---------------------------------------------------------------------------
// state handle
// ---------------------------------------------------------------------------
struct DeviceCtx {
  int device = -1;
  int cus = 0;
  bool ready = false;
};
DeviceCtx& ctx();
int ensure_device();
// address of the no-allocation entry point's static scratch on `device`
// (hipGetSymbolAddress, once per device, with the device current)
int noalloc_setup(int device);
// free the batch executor's cached state slab (unless a batch is running);
// returns the bytes released
size_t release_cached_memory();
}  // namespace bt

namespace btc {
struct CircuitHost;
}

namespace bt {
// density-matrix executor pieces shared with the graph capture; a grid
// pass writes kGridAccRows rows of reduction partials per workgroup
constexpr int kGridAccRows = 65;
int plan_static_passes(bt_state_s* s, const btc::CircuitHost& ch, int T, std::vector<unsigned char>* tables,
                       std::vector<uint8_t>* narrow, size_t* tabsz);
int launch_grid_pass(bt_state_s* s, const void* dtab, bool narrow, int T, double* d_part);
void grid_pass_lds_attr(int prec, int T);
int dm_tile_bits(bt_state_s* rho);              // tile bits of a density matrix's passes
int dm_tile_bits_for(int prec, int nbits);      // the same for vec(rho) of nbits index bits

}  // namespace bt

namespace bt {
struct GridRes;  // the circuit executor's staging of a state (bt_exec.hip)
void free_grid_res(GridRes* g);
}  // namespace bt

struct bt_state_s {
  int n = 0;         // qubits
  int kind = BT_SV;  // BT_SV / BT_DM
  int prec = BT_C64; // BT_C64 / BT_C128
  int nbits = 0;     // n (SV) or 2n (DM)
  uint64_t len = 0;  // 2^nbits complex entries
  void* d = nullptr; // device entries
  double* d_part = nullptr;  // reduction partials (device)
  double* d_out = nullptr;   // reduction result (device)
  double* h_out = nullptr;   // pinned host copy of the result
  void* d_mat = nullptr;     // device matrix staging (k >= 4 blocks)
  bt::GridRes* grid = nullptr;  // circuit executor staging, made on the first run, kept with the state
  hipStream_t stream = nullptr;
  size_t elem() const { return prec == BT_C64 ? 8 : 16; }
};

namespace bt {

// ---------------------------------------------------------------------------
// kernel launchers (bt_kernels.hip)
// ---------------------------------------------------------------------------
constexpr int kRedBlocks = 1024;  // reduction partial blocks
constexpr int kRedMax = 64 + 2;   // max reals per reduction (8x8 Hermitian + extras)

// psi <- scale * M psi on a placement; M column-major complex128 (conj_mat: conj(M)).
int launch_apply(bt_state_s* s, int nbits, int k, const double* mat_colmajor, const Placement& p, double scale,
                 bool conj_mat);
// FP64 reduced density matrix over k bits (k <= 3) into s->h_out (d*d reals:
// diag then upper triangle re/im) -- see unpack_reduced.
int launch_reduce(bt_state_s* s, int k, const int* bits);
void unpack_reduced(int k, const double* packed, double* rho_colmajor);
int launch_norm2(bt_state_s* s, double* out);
int launch_inner(bt_state_s* a, bt_state_s* b, double* re, double* im);  // <a|b>, FP64
// Expectation of a Pauli string (x, z masks) times a dense k-bit block (k <= 3)
// on a state vector (<psi|O|psi>) or density matrix (tr(rho O)), FP64.
struct ExpectArgs {
  uint64_t ngroups;  // 2^(n - nfixed)
  uint64_t xmask, zmask;
  uint64_t off[8];   // offset of block index l (bits[0] = MSB)
  int32_t n;         // qubits
  int32_t nfixed;    // inserted bits: the block's, plus the top bit of x on a state vector
  int32_t fixed[4];  // ascending
  double g[2 * 64];  // block, complex128 column-major
};
int launch_expect(bt_state_s* s, int k, const double* mat, const int* bits, uint64_t xmask, uint64_t zmask,
                  double* re, double* im);
// Shot sampling of a state vector (bt_sample.hip): shot s takes the smallest i
// with u_s T < C_i (FP64 cumulative sums in index order, T the total);
// d_u null -> every shot takes u0.  d_flag[s] = 1 when u_s T lies within 1e-12 T
// of a boundary.  scratch: sample_scratch_bytes(nbits) device bytes.
size_t sample_scratch_bytes(int nbits);
int launch_sample(const void* state, int prec, int nbits, hipStream_t st, void* scratch, int64_t shots,
                  const double* d_u, double u0, uint64_t* d_out, uint8_t* d_flag);
int launch_trace(bt_state_s* s, double* out);
int launch_probabilities(bt_state_s* s, double* d_out);
int launch_scale(bt_state_s* s, double f);
int launch_set_basis(bt_state_s* s, uint64_t index);
int launch_gather(bt_state_s* s, const uint64_t* d_idx, uint64_t n, double* d_out);
int launch_convert_upload(bt_state_s* s, const double* d_src);   // complex128 staging -> state
int launch_convert_download(bt_state_s* s, double* d_dst);       // state -> complex128 staging

}  // namespace bt
