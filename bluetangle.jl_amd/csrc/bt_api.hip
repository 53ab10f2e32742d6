// C ABI: library, state handles and the per-op entry points that stand in
// for BlueTangle.jl's apply(state, op; noise) (src/hilbert.jl:469-666).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bt_internal.hpp"
#include "bt_interp.hpp"

namespace bt {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

DeviceCtx& ctx() {
  static DeviceCtx c;
  return c;
}

static std::mutex g_init_mu;

int ensure_device() {
  std::lock_guard<std::mutex> lk(g_init_mu);
  DeviceCtx& c = ctx();
  if (c.ready) {
    BT_HIP(hipSetDevice(c.device));
    return BT_OK;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(BT_ERR_NODEVICE, "no HIP device visible");
  c.device = c.device < 0 ? 0 : c.device;
  BT_HIP(hipSetDevice(c.device));
  hipDeviceProp_t prop;
  BT_HIP(hipGetDeviceProperties(&prop, c.device));
  c.cus = prop.multiProcessorCount;
  const int rc = noalloc_setup(c.device);  // the no-allocation sampler's scratch address on this device
  if (rc) return rc;
  c.ready = true;
  return BT_OK;
}

using cd = std::complex<double>;

static inline cd mget(const double* m, int dim, int r, int c) {
  return cd(m[2 * (r + c * dim)], m[2 * (r + c * dim) + 1]);
}
static inline void mset(double* m, int dim, int r, int c, cd v) {
  m[2 * (r + c * dim)] = v.real();
  m[2 * (r + c * dim) + 1] = v.imag();
}

// Reduced rho (column-major d x d complex) of the SV state over k bits.
static int reduced(bt_state_s* s, int k, const int* bits, std::vector<cd>& rho, double* tr) {
  int rc = launch_reduce(s, k, bits);
  if (rc) return rc;
  const int D = 1 << k;
  std::vector<double> r(2 * D * D);
  unpack_reduced(k, s->h_out, r.data());
  rho.assign(D * D, cd(0, 0));
  double t = 0;
  for (int i = 0; i < D * D; i++) rho[i] = cd(r[2 * i], r[2 * i + 1]);
  for (int a = 0; a < D; a++) t += rho[a + a * D].real();
  *tr = t;
  return BT_OK;
}

// Re tr(K rho K')
static double sandwich_trace(const double* K, const std::vector<cd>& rho, int D) {
  double t = 0;
  for (int i = 0; i < D; i++)
    for (int a = 0; a < D; a++) {
      cd kia = mget(K, D, i, a);
      if (kia == cd(0, 0)) continue;
      for (int b = 0; b < D; b++) t += (kia * rho[a + b * D] * std::conj(mget(K, D, i, b))).real();
    }
  return t;
}

// _weighted_sample (hilbert.jl:810-819): first i with u <= cumsum; past the
// FP64 total the reference returns `nothing` -- here the last branch with
// positive probability.
static int weighted_pick(const std::vector<double>& p, double u) {
  double c = 0;
  for (size_t i = 0; i < p.size(); i++) {
    c += p[i];
    if (u <= c) return (int)i;
  }
  int i = (int)p.size() - 1;
  while (i > 0 && !(p[i] > 0)) i--;
  return i;
}

static const double kSqrtHalf = 0.70710678118654752440;

// Measurement basis rotation R (gates.jl:22-59: I, H, HSP; struct.jl:554-563)
static void basis_rotation(int basis, cd R[4]) {
  if (basis == BT_BASIS_X) {
    R[0] = kSqrtHalf; R[1] = kSqrtHalf; R[2] = kSqrtHalf; R[3] = -kSqrtHalf;  // col-major
  } else if (basis == BT_BASIS_Y) {
    R[0] = kSqrtHalf; R[1] = kSqrtHalf; R[2] = cd(0, -kSqrtHalf); R[3] = cd(0, kSqrtHalf);
  } else {
    R[0] = 1; R[1] = 0; R[2] = 0; R[3] = 1;
  }
}

static int check_state(bt_state_s* s) {
  if (!s || !s->d) return fail(BT_ERR_ARG, "null or destroyed state");
  return ensure_device();
}

}  // namespace bt

using namespace bt;

namespace bt {
// y = M x, one thread per row; M column-major complex128 (a wavefront reads a
// contiguous piece of each column), FP64 accumulation
template <typename R>
__global__ void k_matvec(const double2* __restrict__ M, const C<R>* __restrict__ x, C<R>* __restrict__ y,
                         uint64_t D) {
  const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= D) return;
  double ar = 0.0, ai = 0.0;
  for (uint64_t c = 0; c < D; c++) {
    const double2 m = M[r + c * D];
    const C<R> v = x[c];
    ar += m.x * (double)v.x - m.y * (double)v.y;
    ai += m.x * (double)v.y + m.y * (double)v.x;
  }
  y[r].x = (R)ar;
  y[r].y = (R)ai;
}
}  // namespace bt

extern "C" {

const char* bt_last_error(void) { return g_err.c_str(); }
int bt_version(void) { return 10000; }

double bt_uniform(uint64_t seed, uint64_t traj, uint64_t draw) { return bt::uniform(seed, traj, draw); }

int bt_init(int device) {
  {
    std::lock_guard<std::mutex> lk(g_init_mu);
    DeviceCtx& c = ctx();
    if (c.ready && c.device != device) return fail(BT_ERR_STATE, "bt_init: already bound to another device");
    c.device = device;
  }
  return ensure_device();
}

uint64_t bt_state_bytes(int nqubits, int kind, int prec) {
  if (nqubits < 1 || nqubits > 40) return 0;
  int nbits = kind == BT_DM ? 2 * nqubits : nqubits;
  if (nbits > 40) return 0;
  return (1ull << nbits) * (prec == BT_C64 ? 8ull : 16ull);
}

int bt_state_create(int nqubits, int kind, int prec, bt_state* out) {
  if (!out) return fail(BT_ERR_ARG, "null output");
  *out = nullptr;
  if (kind != BT_SV && kind != BT_DM) return fail(BT_ERR_ARG, "kind must be BT_SV or BT_DM");
  if (prec != BT_C64 && prec != BT_C128) return fail(BT_ERR_ARG, "prec must be BT_C64 or BT_C128");
  const int nbits = kind == BT_DM ? 2 * nqubits : nqubits;
  if (nqubits < 1 || nbits > 36) return fail(BT_ERR_ARG, "qubit count out of range");
  int rc = ensure_device();
  if (rc) return rc;
  auto* s = new bt_state_s();
  s->n = nqubits;
  s->kind = kind;
  s->prec = prec;
  s->nbits = nbits;
  s->len = 1ull << nbits;
  if (hipMalloc(&s->d, s->len * s->elem()) != hipSuccess) {
    (void)hipGetLastError();
    // the batch executor's cached slab may hold the memory: release it, try again
    if (!release_cached_memory() || hipMalloc(&s->d, s->len * s->elem()) != hipSuccess) {
      delete s;
      (void)hipGetLastError();
      return fail(BT_ERR_NOMEM, "hipMalloc of the state failed (out of HBM)");
    }
  }
  if (hipMalloc(&s->d_part, sizeof(double) * kRedBlocks * kRedMax) != hipSuccess ||
      hipMalloc(&s->d_out, sizeof(double) * kRedMax) != hipSuccess ||
      hipHostMalloc(&s->h_out, sizeof(double) * kRedMax) != hipSuccess ||
      hipStreamCreate(&s->stream) != hipSuccess) {
    bt_state_destroy(s);
    return fail(BT_ERR_NOMEM, "allocation of state scratch failed");
  }
  rc = launch_set_basis(s, 0);
  if (rc) {
    bt_state_destroy(s);
    return rc;
  }
  *out = s;
  return BT_OK;
}

int bt_state_destroy(bt_state s) {
  if (!s) return BT_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->d) (void)hipFree(s->d);
  if (s->d_part) (void)hipFree(s->d_part);
  if (s->d_out) (void)hipFree(s->d_out);
  if (s->d_mat) (void)hipFree(s->d_mat);
  if (s->h_out) (void)hipHostFree(s->h_out);
  if (s->grid) bt::free_grid_res(s->grid);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return BT_OK;
}

int bt_state_info(bt_state s, int* nqubits, int* kind, int* prec) {
  if (!s) return fail(BT_ERR_ARG, "null state");
  if (nqubits) *nqubits = s->n;
  if (kind) *kind = s->kind;
  if (prec) *prec = s->prec;
  return BT_OK;
}

int bt_state_device_ptr(bt_state s, void** ptr, void** stream) {
  if (!s) return fail(BT_ERR_ARG, "null state");
  if (ptr) *ptr = s->d;
  if (stream) *stream = (void*)s->stream;
  return BT_OK;
}

int bt_state_sync(bt_state s) {
  int rc = check_state(s);
  if (rc) return rc;
  BT_HIP(hipStreamSynchronize(s->stream));
  return BT_OK;
}

int bt_state_set_basis(bt_state s, uint64_t index) {
  int rc = check_state(s);
  if (rc) return rc;
  uint64_t D = 1ull << s->n;
  if (index >= D) return fail(BT_ERR_ARG, "basis index out of range");
  uint64_t pos = s->kind == BT_DM ? index + index * D : index;
  return launch_set_basis(s, pos);
}

int bt_state_copy(bt_state dst, bt_state src) {
  int rc = check_state(dst);
  if (rc) return rc;
  rc = check_state(src);
  if (rc) return rc;
  if (dst->len != src->len || dst->prec != src->prec || dst->kind != src->kind)
    return fail(BT_ERR_ARG, "bt_state_copy: shapes differ");
  BT_HIP(hipStreamSynchronize(src->stream));
  BT_HIP(hipMemcpyAsync(dst->d, src->d, dst->len * dst->elem(), hipMemcpyDeviceToDevice, dst->stream));
  BT_HIP(hipStreamSynchronize(dst->stream));
  return BT_OK;
}

static const uint64_t kChunk = 1ull << 22;  // complex entries per staging chunk

int bt_state_upload(bt_state s, const double* data, uint64_t n) {
  int rc = check_state(s);
  if (rc) return rc;
  if (n != s->len || !data) return fail(BT_ERR_ARG, "bt_state_upload: wrong length");
  BT_HIP(hipStreamSynchronize(s->stream));
  if (s->prec == BT_C128) {
    BT_HIP(hipMemcpy(s->d, data, n * 16, hipMemcpyHostToDevice));
    return BT_OK;
  }
  float* stage = nullptr;
  BT_HIP(hipHostMalloc((void**)&stage, std::min(n, kChunk) * 8));
  for (uint64_t o = 0; o < n; o += kChunk) {
    uint64_t m = std::min(kChunk, n - o);
    for (uint64_t i = 0; i < 2 * m; i++) stage[i] = (float)data[2 * o + i];
    hipError_t e = hipMemcpy((char*)s->d + o * 8, stage, m * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipHostFree(stage);
      return fail(BT_ERR_HIP, std::string("upload: ") + hipGetErrorString(e));
    }
  }
  BT_HIP(hipHostFree(stage));
  return BT_OK;
}

int bt_state_download(bt_state s, double* data, uint64_t n) {
  int rc = check_state(s);
  if (rc) return rc;
  if (n != s->len || !data) return fail(BT_ERR_ARG, "bt_state_download: wrong length");
  BT_HIP(hipStreamSynchronize(s->stream));
  if (s->prec == BT_C128) {
    BT_HIP(hipMemcpy(data, s->d, n * 16, hipMemcpyDeviceToHost));
    return BT_OK;
  }
  float* stage = nullptr;
  BT_HIP(hipHostMalloc((void**)&stage, std::min(n, kChunk) * 8));
  for (uint64_t o = 0; o < n; o += kChunk) {
    uint64_t m = std::min(kChunk, n - o);
    hipError_t e = hipMemcpy(stage, (char*)s->d + o * 8, m * 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      (void)hipHostFree(stage);
      return fail(BT_ERR_HIP, std::string("download: ") + hipGetErrorString(e));
    }
    for (uint64_t i = 0; i < 2 * m; i++) data[2 * o + i] = (double)stage[i];
  }
  BT_HIP(hipHostFree(stage));
  return BT_OK;
}

// ---- gates ------------------------------------------------------------------
int bt_apply_matrix(bt_state s, int k, const double* mat, const int* bits, uint64_t ctrl_mask) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!mat || !bits) return fail(BT_ERR_ARG, "null matrix or bits");
  if (s->kind == BT_SV) {
    Placement p;
    rc = make_placement(s->n, k, bits, ctrl_mask, &p);
    if (rc) return rc;
    return launch_apply(s, s->n, k, mat, p, 1.0, false);
  }
  // DM: rho <- E rho E' : E on row bits, conj(E) on column bits
  if (k > 3) return fail(BT_ERR_ARG, "density-matrix gates take k <= 3");
  Placement pr, pc;
  rc = make_placement(s->nbits, k, bits, ctrl_mask, &pr);
  if (rc) return rc;
  for (int i = 0; i < k; i++)
    if (bits[i] >= s->n) return fail(BT_ERR_ARG, "bit position out of range");
  if (s->n < 64 && (ctrl_mask >> s->n)) return fail(BT_ERR_ARG, "control bit out of range");
  int cb[kMaxK];
  for (int i = 0; i < k; i++) cb[i] = bits[i] + s->n;
  rc = make_placement(s->nbits, k, cb, ctrl_mask << s->n, &pc);
  if (rc) return rc;
  rc = launch_apply(s, s->nbits, k, mat, pr, 1.0, false);
  if (rc) return rc;
  return launch_apply(s, s->nbits, k, mat, pc, 1.0, true);
}


int bt_apply_full(bt_state s, const double* mat) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!mat) return fail(BT_ERR_ARG, "null matrix");
  if (s->kind != BT_SV) return fail(BT_ERR_UNSUPPORTED, "bt_apply_full: state-vector only");
  if (s->n > 15) return fail(BT_ERR_ARG, "bt_apply_full: a full-space matrix needs N <= 15");
  const uint64_t D = s->len;
  void* dM = nullptr;
  void* dy = nullptr;
  if (hipMalloc(&dM, sizeof(double2) * D * D) != hipSuccess || hipMalloc(&dy, s->elem() * D) != hipSuccess) {
    if (dM) (void)hipFree(dM);
    return fail(BT_ERR_NOMEM, "bt_apply_full: device buffers");
  }
  auto done = [&](int r) {
    (void)hipStreamSynchronize(s->stream);
    (void)hipFree(dM);
    (void)hipFree(dy);
    return r;
  };
  if (hipMemcpyAsync(dM, mat, sizeof(double2) * D * D, hipMemcpyHostToDevice, s->stream) != hipSuccess)
    return done(fail(BT_ERR_HIP, "bt_apply_full: matrix upload"));
  const unsigned grid = (unsigned)((D + 255) / 256);
  if (s->prec == BT_C64)
    hipLaunchKernelGGL(k_matvec<float>, dim3(grid), dim3(256), 0, s->stream, (const double2*)dM,
                       (const C<float>*)s->d, (C<float>*)dy, D);
  else
    hipLaunchKernelGGL(k_matvec<double>, dim3(grid), dim3(256), 0, s->stream, (const double2*)dM,
                       (const C<double>*)s->d, (C<double>*)dy, D);
  if (hipGetLastError() != hipSuccess) return done(fail(BT_ERR_HIP, "bt_apply_full: launch"));
  if (hipMemcpyAsync(s->d, dy, s->elem() * D, hipMemcpyDeviceToDevice, s->stream) != hipSuccess)
    return done(fail(BT_ERR_HIP, "bt_apply_full: copy back"));
  return done(BT_OK);
}

int bt_reduced_rho(bt_state s, int k, const int* bits, double* rho_out, double* norm2_out) {
  int rc = check_state(s);
  if (rc) return rc;
  if (s->kind != BT_SV) return fail(BT_ERR_UNSUPPORTED, "bt_reduced_rho: state-vector only");
  if (k < 1 || k > 3) return fail(BT_ERR_ARG, "bt_reduced_rho: k must be 1..3");
  rc = launch_reduce(s, k, bits);
  if (rc) return rc;
  const int D = 1 << k;
  if (rho_out) unpack_reduced(k, s->h_out, rho_out);
  if (norm2_out) {
    double t = 0;
    for (int a = 0; a < D; a++) t += s->h_out[a];
    *norm2_out = t;
  }
  return BT_OK;
}

// ---- trajectory channels ------------------------------------------------------
int bt_apply_kraus_traj(bt_state s, int k, int nkraus, const double* kraus, const int* apply_bits,
                        const int* prob_bits, double u, int* chosen, double* probs_out) {
  int rc = check_state(s);
  if (rc) return rc;
  if (s->kind != BT_SV) return fail(BT_ERR_UNSUPPORTED, "trajectory channels need a state vector");
  if (k < 1 || k > 3 || nkraus < 1 || !kraus || !apply_bits) return fail(BT_ERR_ARG, "bad channel arguments");
  if (!prob_bits) prob_bits = apply_bits;
  const int D = 1 << k;
  // the probability bits must be a permutation of the apply bits
  int perm[3];  // prob position j holds apply position perm[j]
  for (int j = 0; j < k; j++) {
    perm[j] = -1;
    for (int i = 0; i < k; i++)
      if (prob_bits[j] == apply_bits[i]) perm[j] = i;
    if (perm[j] < 0) return fail(BT_ERR_ARG, "prob_bits must permute apply_bits");
  }
  std::vector<cd> rho;
  double tr = 0;
  rc = reduced(s, k, prob_bits, rho, &tr);
  if (rc) return rc;
  // the reference samples from the state as it is (unnormalised after a
  // non-unitary gate): p_i = Re tr(K_i rho_A K_i') without dividing by tr
  (void)tr;
  std::vector<double> p(nkraus);
  for (int i = 0; i < nkraus; i++) p[i] = sandwich_trace(kraus + (size_t)i * 2 * D * D, rho, D);
  int ind = weighted_pick(p, u);
  if (chosen) *chosen = ind;
  if (probs_out)
    for (int i = 0; i < nkraus; i++) probs_out[i] = p[i];
  // ||K psi||^2 with K acting in apply order: permute rho from prob order
  std::vector<cd> rho_a(D * D);
  for (int a = 0; a < D; a++)
    for (int b = 0; b < D; b++) {
      // local index in prob order -> apply order
      int aa = 0, bb = 0;
      for (int j = 0; j < k; j++) {
        if ((a >> (k - 1 - j)) & 1) aa |= 1 << (k - 1 - perm[j]);
        if ((b >> (k - 1 - j)) & 1) bb |= 1 << (k - 1 - perm[j]);
      }
      rho_a[aa + bb * D] = rho[a + b * D];
    }
  const double* K = kraus + (size_t)ind * 2 * D * D;
  double nk2 = sandwich_trace(K, rho_a, D);
  // under the qubit > target ordering quirk the sampled branch can annihilate
  // the state: the reference's normalize then gives NaN everywhere, and a
  // norm left over from rounding must not be renormalised into a state
  double psum = 0;
  for (int i = 0; i < nkraus; i++) psum += p[i];
  const double scale = nk2 <= bti::kZeroBranch * psum ? std::nan("") : 1.0 / std::sqrt(nk2);
  Placement pl;
  rc = make_placement(s->n, k, apply_bits, 0, &pl);
  if (rc) return rc;
  return launch_apply(s, s->n, k, K, pl, scale, false);
}

int bt_apply_kraus_dm(bt_state s, int k, int nkraus, const double* kraus, const int* bits) {
  int rc = check_state(s);
  if (rc) return rc;
  if (s->kind != BT_DM) return fail(BT_ERR_UNSUPPORTED, "bt_apply_kraus_dm needs a density matrix");
  if (k < 1 || k > 2 || nkraus < 1 || !kraus || !bits) return fail(BT_ERR_ARG, "bad channel arguments");
  const int D = 1 << k, S = D * D;
  for (int i = 0; i < k; i++)
    if (bits[i] < 0 || bits[i] >= s->n) return fail(BT_ERR_ARG, "bit position out of range");
  // superoperator sum_i K_i (x) conj(K_i) on [row bits, column bits]
  std::vector<double> sup(2 * S * S, 0.0);
  for (int i = 0; i < nkraus; i++) {
    const double* K = kraus + (size_t)i * 2 * D * D;
    for (int r1 = 0; r1 < D; r1++)
      for (int c1 = 0; c1 < D; c1++)
        for (int r0 = 0; r0 < D; r0++)
          for (int c0 = 0; c0 < D; c0++) {
            cd v = mget(K, D, r1, r0) * std::conj(mget(K, D, c1, c0));
            int row = (r1 << k) | c1, col = (r0 << k) | c0;
            sup[2 * (row + col * S)] += v.real();
            sup[2 * (row + col * S) + 1] += v.imag();
          }
  }
  int sb[2 * 3];
  for (int i = 0; i < k; i++) {
    sb[i] = bits[i];
    sb[k + i] = bits[i] + s->n;
  }
  Placement pl;
  rc = make_placement(s->nbits, 2 * k, sb, 0, &pl);
  if (rc) return rc;
  return launch_apply(s, s->nbits, 2 * k, sup.data(), pl, 1.0, false);
}

// ---- measurement -------------------------------------------------------------
int bt_measure(bt_state s, int bit, int basis, const double* u, int* outcome, double* prob0) {
  int rc = check_state(s);
  if (rc) return rc;
  if (s->kind != BT_SV) return fail(BT_ERR_UNSUPPORTED, "fix this: density-matrix measurement is not implemented");
  if (!u) return fail(BT_ERR_ARG, "null uniforms");
  if (bit < 0 || bit >= s->n) return fail(BT_ERR_ARG, "bit position out of range");
  double ud = u[0];
  if (basis == BT_BASIS_R) {
    int pick = (int)std::floor(u[0] * 3.0);
    if (pick > 2) pick = 2;
    static const int order[3] = {BT_BASIS_X, BT_BASIS_Y, BT_BASIS_Z};
    basis = order[pick];
    ud = u[1];
  } else if (basis < 0 || basis > 3) {
    return fail(BT_ERR_ARG, "bad basis");
  }
  std::vector<cd> rho;
  double tr = 0;
  rc = reduced(s, 1, &bit, rho, &tr);
  if (rc) return rc;
  cd R[4];
  basis_rotation(basis, R);
  // rho' = R rho R'
  cd Rr[4];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      cd v = 0;
      for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++) v += R[a + 2 * x] * rho[x + 2 * y] * std::conj(R[b + 2 * y]);
      Rr[a + 2 * b] = v;
    }
  (void)tr;
  const double p0 = Rr[0].real();  // prob0 of the state as it is (hilbert.jl:690)
  const int ind = ud < p0 ? 0 : 1;
  if (outcome) *outcome = ind;
  if (prob0) *prob0 = p0;
  // G = R' P_ind R, scaled by 1 / ||P_ind R psi||
  double g[8];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      cd v = std::conj(R[ind + 2 * a]) * R[ind + 2 * b];
      mset(g, 2, a, b, v);
    }
  Placement pl;
  rc = make_placement(s->n, 1, &bit, 0, &pl);
  if (rc) return rc;
  return launch_apply(s, s->n, 1, g, pl, 1.0 / std::sqrt(Rr[ind + 2 * ind].real()), false);
}

int bt_reset(bt_state s, int bit, double u, int* outcome) {
  int rc = check_state(s);
  if (rc) return rc;
  if (s->kind != BT_SV) return fail(BT_ERR_UNSUPPORTED, "bt_reset: state-vector only (rho uses the channel)");
  if (bit < 0 || bit >= s->n) return fail(BT_ERR_ARG, "bit position out of range");
  std::vector<cd> rho;
  double tr = 0;
  rc = reduced(s, 1, &bit, rho, &tr);
  if (rc) return rc;
  (void)tr;
  const double p0 = rho[0].real();
  const int ind = u < p0 ? 0 : 1;
  if (outcome) *outcome = ind;
  // P0 (outcome 0) or X P1 = |0><1| (outcome 1), renormalised
  double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (ind == 0) g[0] = 1.0;
  else g[4] = 1.0;  // (0,1) entry, column-major index 2
  Placement pl;
  rc = make_placement(s->n, 1, &bit, 0, &pl);
  if (rc) return rc;
  return launch_apply(s, s->n, 1, g, pl, 1.0 / std::sqrt(rho[ind * 3].real()), false);
}

// ---- outputs -----------------------------------------------------------------
int bt_probabilities(bt_state s, double* out) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!out) return fail(BT_ERR_ARG, "null output");
  const uint64_t D = 1ull << s->n;
  double* d = nullptr;
  BT_HIP(hipMalloc(&d, D * sizeof(double)));
  rc = launch_probabilities(s, d);
  if (rc) {
    (void)hipFree(d);
    return rc;
  }
  hipError_t e = hipMemcpyAsync(out, d, D * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  hipError_t e2 = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  if (e != hipSuccess || e2 != hipSuccess) return fail(BT_ERR_HIP, "bt_probabilities: copy failed");
  return BT_OK;
}

int bt_norm2(bt_state s, double* out) {
  int rc = check_state(s);
  if (rc) return rc;
  return launch_norm2(s, out);
}

int bt_inner(bt_state a, bt_state b, double* re, double* im) {
  int rc = check_state(a);
  if (rc) return rc;
  rc = check_state(b);
  if (rc) return rc;
  if (a->kind != b->kind || a->n != b->n) return fail(BT_ERR_ARG, "bt_inner: states of different shapes");
  if (!re || !im) return fail(BT_ERR_ARG, "null output");
  // b's queued work must be done before a's stream reads it
  BT_HIP(hipStreamSynchronize(b->stream));
  return launch_inner(a, b, re, im);
}

int bt_expect(bt_state s, int k, const double* mat, const int* bits, uint64_t xmask, uint64_t zmask, double* re,
              double* im) {
  int rc = check_state(s);
  if (rc) return rc;
  if (!re || !im) return fail(BT_ERR_ARG, "null output");
  if (k > 0 && (!mat || !bits)) return fail(BT_ERR_ARG, "bt_expect: null block");
  return launch_expect(s, k, mat, bits, xmask, zmask, re, im);
}

int bt_trace_purity(bt_state s, double* trace, double* purity) {
  int rc = check_state(s);
  if (rc) return rc;
  if (s->kind != BT_DM) return fail(BT_ERR_UNSUPPORTED, "bt_trace_purity: density matrix only");
  if (trace) {
    rc = launch_trace(s, trace);
    if (rc) return rc;
  }
  if (purity) {
    rc = launch_norm2(s, purity);
    if (rc) return rc;
  }
  return BT_OK;
}

int bt_gather(bt_state s, const uint64_t* idx, uint64_t n, double* out) {
  int rc = check_state(s);
  if (rc) return rc;
  if (n == 0) return BT_OK;
  if (!idx || !out) return fail(BT_ERR_ARG, "null index or output");
  uint64_t* di = nullptr;
  double* dout = nullptr;
  BT_HIP(hipMalloc(&di, n * sizeof(uint64_t)));
  if (hipMalloc(&dout, n * 16) != hipSuccess) {
    (void)hipFree(di);
    return fail(BT_ERR_NOMEM, "bt_gather: allocation failed");
  }
  hipError_t e = hipMemcpyAsync(di, idx, n * sizeof(uint64_t), hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) rc = launch_gather(s, di, n, dout);
  if (e == hipSuccess && rc == BT_OK) e = hipMemcpyAsync(out, dout, n * 16, hipMemcpyDeviceToHost, s->stream);
  hipError_t e2 = hipStreamSynchronize(s->stream);
  (void)hipFree(di);
  (void)hipFree(dout);
  if (rc) return rc;
  if (e != hipSuccess || e2 != hipSuccess) return fail(BT_ERR_HIP, "bt_gather: copy failed");
  return BT_OK;
}

}  // extern "C"
