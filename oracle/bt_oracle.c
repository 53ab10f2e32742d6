/*
 * bt_oracle.c -- FP64 CPU restatement of BlueTangle.jl's apply path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it.  The product
 * (libbt.so) never links, loads or calls it.
 *
 * Everything is complex128 / float64, as in the reference.  Two modes:
 *   mode 0 ("fast"):      in-place strided updates (OpenMP over amplitude
 *                         groups for N >= 14).
 *   mode 1 ("reference"): the reference's own algorithm -- every gate or
 *                         Kraus operator is built as a full 2^N x 2^N sparse
 *                         CSC matrix by folding kron over per-qubit factors
 *                         (src/hilbert.jl:18-70,106-128,143-159), then
 *                         multiplied into the state (hilbert.jl:505) or into
 *                         rho from both sides (hilbert.jl:655-656,
 *                         struct.jl:58-76).  Single-threaded, like Julia.
 *
 * Conventions (src/bit.jl:9-15): qubit q (1-based) is bit (N-q) of the
 * 0-based basis index; qubit 1 is the most significant bit.  Matrices are
 * column-major, complex interleaved (re,im), like Julia's Matrix{ComplexF64}.
 * A 2-qubit matrix's first tensor factor acts on `qubit`
 * (hilbert.jl:30-36 with _swap_control_target, hilbert.jl:215-237).
 *
 * Uniforms: the reference draws rand() from Julia's global RNG.  The oracle
 * instead consumes a pre-drawn uniform stream in the reference's draw order
 * (one draw per sampled channel, hilbert.jl:812; one per Born measurement,
 * hilbert.jl:693; one extra for the MR basis, struct.jl:568).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define BTO_OK 0
#define BTO_ERR_ARG -1
#define BTO_ERR_UNSUPPORTED -2
#define BTO_ERR_DRAWS -3
#define BTO_ERR_NOMEM -4

static char g_err[256];
const char* bto_last_error(void) { return g_err; }
static int fail(int code, const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

/* ------------------------------------------------------------------ */
/* Counter-based uniform stream (the engine's contract; restated here). */
/* SplitMix64 finaliser (Steele, Lea, Flood 2014).                       */
/* ------------------------------------------------------------------ */
uint64_t bto_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
double bto_uniform(uint64_t seed, uint64_t traj, uint64_t draw) {
  uint64_t z = bto_splitmix64(seed ^ bto_splitmix64(traj ^ bto_splitmix64(draw)));
  return (double)(z >> 11) * 0x1.0p-53;
}
void bto_uniforms(uint64_t seed, uint64_t traj, int64_t n, double* out) {
  for (int64_t i = 0; i < n; i++) out[i] = bto_uniform(seed, traj, (uint64_t)i);
}

/* ------------------------------------------------------------------ */
/* Reference-level op description                                      */
/* ------------------------------------------------------------------ */
enum { BTO_GATE = 0, BTO_CHANNEL = 1, BTO_MEASURE = 2, BTO_RESET = 3, BTO_IFOP = 4 };
enum { BTO_MZ = 0, BTO_MX = 1, BTO_MY = 2, BTO_MR = 3 };

typedef struct {
  int32_t kind;    /* BTO_* */
  int32_t q;       /* qubits spanned by mat / kraus: 1, 2 or 3 */
  int32_t qubit;   /* 1-based */
  int32_t target;  /* 1-based or -1 */
  int32_t control; /* 1-based or -2 */
  int32_t noisy;   /* Op.noisy (struct.jl:372) */
  int32_t nkraus;  /* channels */
  int32_t basis;   /* measurement / ifOp basis */
  const double* mat;   /* gate matrix, col-major interleaved, 2^q x 2^q */
  const double* kraus; /* nkraus matrices, col-major interleaved */
  int32_t if0_begin, if0_end, if1_begin, if1_end; /* ifOp branches (indices into ops) */
  int32_t is_cx, is_cz; /* mat == gate.CX / gate.CZ (hilbert.jl:58-63) */
} bto_op;

typedef struct {
  int32_t n1; const double* k1; /* q1 channel: n1 2x2 Kraus ops, n1 = 0 -> `false` */
  int32_t n2; const double* k2; /* q2 channel: n2 4x4 Kraus ops, n2 = 0 -> `false` */
} bto_noise;

typedef struct {
  const double* u; int64_t n; int64_t pos;
  int64_t boundary;      /* draws within `tol` of a cumulative boundary */
  double tol;
  int64_t overflow;      /* _weighted_sample draws beyond the FP64 total (Q5) */
} draws_t;

static int next_draw(draws_t* d, double* u) {
  if (d->pos >= d->n) return fail(BTO_ERR_DRAWS, "uniform stream exhausted");
  *u = d->u[d->pos++];
  return BTO_OK;
}

/* ------------------------------------------------------------------ */
/* Small dense helpers (col-major)                                     */
/* ------------------------------------------------------------------ */
static inline cplx mget(const double* m, int dim, int r, int c) {
  return m[2 * (r + c * dim)] + I * m[2 * (r + c * dim) + 1];
}

/* gate table values used internally (gates.jl:22-59) */
static const double M_I[8] = {1, 0, 0, 0, 0, 0, 1, 0};
static const double M_X[8] = {0, 0, 1, 0, 1, 0, 0, 0};
static const double M_P0[8] = {1, 0, 0, 0, 0, 0, 0, 0};
static const double M_P1[8] = {0, 0, 0, 0, 0, 0, 1, 0};
static double M_H[8], M_HSP[8];
static void init_tables(void) {
  static int done = 0;
  if (done) return;
  double s = 1.0 / sqrt(2.0);
  /* H = (1/sqrt2) [1 1; 1 -1] */
  double h[8] = {s, 0, s, 0, s, 0, -s, 0};
  memcpy(M_H, h, sizeof h);
  /* HSP = (1/sqrt2) [1 -i; 1 i]  (gates.jl:34), col-major: (0,0)=1,(1,0)=1,(0,1)=-i,(1,1)=i */
  double hs[8] = {s, 0, s, 0, 0, -s, 0, s};
  memcpy(M_HSP, hs, sizeof hs);
  done = 1;
}

static const double* measurement_mat(int basis) {
  init_tables();
  switch (basis) {
    case BTO_MX: return M_H;
    case BTO_MY: return M_HSP;
    default: return M_I;
  }
}

/* ------------------------------------------------------------------ */
/* Fast mode: in-place strided kernels on the state vector             */
/* ------------------------------------------------------------------ */
static inline uint64_t insert0(uint64_t x, int b) {
  uint64_t lo = x & ((1ull << b) - 1);
  return ((x >> b) << (b + 1)) | lo;
}

/* Apply a dense 2^k x 2^k matrix on bits[0..k-1] (bits[0] = MSB of the local
 * index) where all bits in cmask are 1.  k <= 3. */
static void sv_apply_k(cplx* s, int N, int k, const double* m, const int* bits, uint64_t cmask) {
  int dim = 1 << k;
  cplx M[64];
  for (int c = 0; c < dim; c++)
    for (int r = 0; r < dim; r++) M[r + c * dim] = mget(m, dim, r, c);
  int sorted[3];
  for (int i = 0; i < k; i++) sorted[i] = bits[i];
  for (int i = 0; i < k; i++)
    for (int j = i + 1; j < k; j++)
      if (sorted[j] < sorted[i]) { int t = sorted[i]; sorted[i] = sorted[j]; sorted[j] = t; }
  uint64_t off[8];
  for (int l = 0; l < dim; l++) {
    uint64_t o = 0;
    for (int i = 0; i < k; i++)
      if ((l >> (k - 1 - i)) & 1) o |= 1ull << bits[i];
    off[l] = o;
  }
  int64_t ngroups = (int64_t)1 << (N - k);
#pragma omp parallel for schedule(static) if (N >= 14)
  for (int64_t g = 0; g < ngroups; g++) {
    uint64_t base = (uint64_t)g;
    for (int i = 0; i < k; i++) base = insert0(base, sorted[i]);
    if ((base & cmask) != cmask) continue;
    cplx v[8], w[8];
    for (int l = 0; l < dim; l++) v[l] = s[base | off[l]];
    for (int r = 0; r < dim; r++) {
      cplx acc = 0;
      for (int c = 0; c < dim; c++) acc += M[r + c * dim] * v[c];
      w[r] = acc;
    }
    for (int l = 0; l < dim; l++) s[base | off[l]] = w[l];
  }
}

static double sv_norm2(const cplx* s, int N) {
  double acc = 0;
  int64_t n = (int64_t)1 << N;
#pragma omp parallel for reduction(+ : acc) schedule(static) if (N >= 14)
  for (int64_t i = 0; i < n; i++) acc += creal(s[i]) * creal(s[i]) + cimag(s[i]) * cimag(s[i]);
  return acc;
}

static void sv_scale(cplx* s, int N, double f) {
  int64_t n = (int64_t)1 << N;
#pragma omp parallel for schedule(static) if (N >= 14)
  for (int64_t i = 0; i < n; i++) s[i] *= f;
}

/* LinearAlgebra.normalize(v) = v / norm(v) (struct.jl:34, hilbert.jl:694) */
static void sv_normalize(cplx* s, int N) {
  double nrm = sqrt(sv_norm2(s, N));
  sv_scale(s, N, 1.0 / nrm);
}

/* Reduced density matrix over `keep` qubits (1-based), kept in ascending
 * qubit order (linalg.jl:83-140,167-230): rho[a,b] = sum_rest psi[a,rest]
 * conj(psi[b,rest]); a's MSB is the smallest kept qubit.  out is col-major
 * 2^k x 2^k. */
static void sv_reduced(const cplx* s, int N, int k, const int* keep_sorted, cplx* out) {
  int dim = 1 << k;
  int bits[3];
  for (int i = 0; i < k; i++) bits[i] = N - keep_sorted[i];
  int asc[3];
  for (int i = 0; i < k; i++) asc[i] = bits[i];
  for (int i = 0; i < k; i++)
    for (int j = i + 1; j < k; j++)
      if (asc[j] < asc[i]) { int t = asc[i]; asc[i] = asc[j]; asc[j] = t; }
  uint64_t off[8];
  for (int l = 0; l < dim; l++) {
    uint64_t o = 0;
    for (int i = 0; i < k; i++)
      if ((l >> (k - 1 - i)) & 1) o |= 1ull << bits[i];
    off[l] = o;
  }
  double re[64] = {0}, im[64] = {0};
  int64_t ngroups = (int64_t)1 << (N - k);
  for (int64_t g = 0; g < ngroups; g++) {
    uint64_t base = (uint64_t)g;
    for (int i = 0; i < k; i++) base = insert0(base, asc[i]);
    cplx v[8];
    for (int l = 0; l < dim; l++) v[l] = s[base | off[l]];
    for (int a = 0; a < dim; a++)
      for (int b = 0; b < dim; b++) {
        cplx t = v[a] * conj(v[b]);
        re[a + b * dim] += creal(t);
        im[a + b * dim] += cimag(t);
      }
  }
  for (int i = 0; i < dim * dim; i++) out[i] = re[i] + I * im[i];
}

/* p_k = real(tr(K rho K')) (struct.jl:9-29,44-49) */
static void kraus_probs(const cplx* rho, int dim, int nk, const double* kraus, double* probs) {
  for (int kk = 0; kk < nk; kk++) {
    const double* K = kraus + (size_t)kk * 2 * dim * dim;
    double tr = 0;
    /* tr(K rho K^dag) = sum_{i} sum_{a,b} K[i,a] rho[a,b] conj(K[i,b]) */
    for (int i = 0; i < dim; i++)
      for (int a = 0; a < dim; a++) {
        cplx kia = mget(K, dim, i, a);
        if (kia == 0) continue;
        for (int b = 0; b < dim; b++) tr += creal(kia * rho[a + b * dim] * conj(mget(K, dim, i, b)));
      }
    probs[kk] = tr;
  }
}

/* _weighted_sample (hilbert.jl:810-819): first i with u <= cumsum[i].  The
 * reference returns `nothing` when u exceeds the FP64 total (and then fails);
 * the oracle clamps to the last branch with positive probability and counts
 * the event (documented divergence). */
static int weighted_sample(const double* probs, int n, double u, draws_t* d) {
  double cum = 0;
  int chosen = -1;
  for (int i = 0; i < n; i++) {
    cum += probs[i];
    if (d && fabs(u - cum) <= d->tol) d->boundary++;
    if (chosen < 0 && u <= cum) chosen = i;
  }
  if (chosen < 0) {
    /* clamp to the last branch that has positive probability */
    if (d) d->overflow++;
    chosen = n - 1;
    while (chosen > 0 && !(probs[chosen] > 0)) chosen--;
  }
  return chosen;
}

/* ------------------------------------------------------------------ */
/* Reference-algorithm mode: sparse full-space operators               */
/* ------------------------------------------------------------------ */
typedef struct {
  int64_t n;      /* square dimension */
  int64_t nnz;
  int64_t* colptr; /* n+1 */
  int64_t* rowval;
  cplx* nzval;
} csc_t;

static void csc_free(csc_t* a) {
  free(a->colptr); free(a->rowval); free(a->nzval);
  memset(a, 0, sizeof *a);
}

/* sa.sparse(dense): stores structural nonzeros (entries != 0) */
static csc_t csc_from_dense(const double* m, int dim) {
  csc_t a;
  a.n = dim;
  a.colptr = (int64_t*)malloc(sizeof(int64_t) * (dim + 1));
  a.rowval = (int64_t*)malloc(sizeof(int64_t) * dim * dim);
  a.nzval = (cplx*)malloc(sizeof(cplx) * dim * dim);
  int64_t p = 0;
  for (int c = 0; c < dim; c++) {
    a.colptr[c] = p;
    for (int r = 0; r < dim; r++) {
      cplx v = mget(m, dim, r, c);
      if (v != 0) { a.rowval[p] = r; a.nzval[p] = v; p++; }
    }
  }
  a.colptr[dim] = p;
  a.nnz = p;
  return a;
}

/* kron(A, B) for CSC (SparseArrays.kron) */
static csc_t csc_kron(const csc_t* A, const csc_t* B) {
  csc_t C;
  C.n = A->n * B->n;
  C.nnz = A->nnz * B->nnz;
  C.colptr = (int64_t*)malloc(sizeof(int64_t) * (C.n + 1));
  C.rowval = (int64_t*)malloc(sizeof(int64_t) * (C.nnz ? C.nnz : 1));
  C.nzval = (cplx*)malloc(sizeof(cplx) * (C.nnz ? C.nnz : 1));
  int64_t p = 0;
  for (int64_t ca = 0; ca < A->n; ca++)
    for (int64_t cb = 0; cb < B->n; cb++) {
      C.colptr[ca * B->n + cb] = p;
      for (int64_t pa = A->colptr[ca]; pa < A->colptr[ca + 1]; pa++)
        for (int64_t pb = B->colptr[cb]; pb < B->colptr[cb + 1]; pb++) {
          C.rowval[p] = A->rowval[pa] * B->n + B->rowval[pb];
          C.nzval[p] = A->nzval[pa] * B->nzval[pb];
          p++;
        }
    }
  C.colptr[C.n] = p;
  return C;
}

/* A + B (same dimension), dropping nothing (Julia keeps computed zeros only
 * when both operands store them; exact structure does not affect values). */
static csc_t csc_add(const csc_t* A, const csc_t* B) {
  csc_t C;
  C.n = A->n;
  C.colptr = (int64_t*)malloc(sizeof(int64_t) * (C.n + 1));
  C.rowval = (int64_t*)malloc(sizeof(int64_t) * (A->nnz + B->nnz + 1));
  C.nzval = (cplx*)malloc(sizeof(cplx) * (A->nnz + B->nnz + 1));
  int64_t p = 0;
  for (int64_t c = 0; c < C.n; c++) {
    C.colptr[c] = p;
    int64_t pa = A->colptr[c], pb = B->colptr[c];
    int64_t ea = A->colptr[c + 1], eb = B->colptr[c + 1];
    while (pa < ea || pb < eb) {
      if (pb >= eb || (pa < ea && A->rowval[pa] < B->rowval[pb])) {
        C.rowval[p] = A->rowval[pa]; C.nzval[p] = A->nzval[pa]; pa++;
      } else if (pa >= ea || B->rowval[pb] < A->rowval[pa]) {
        C.rowval[p] = B->rowval[pb]; C.nzval[p] = B->nzval[pb]; pb++;
      } else {
        C.rowval[p] = A->rowval[pa]; C.nzval[p] = A->nzval[pa] + B->nzval[pb]; pa++; pb++;
      }
      p++;
    }
  }
  C.colptr[C.n] = p;
  C.nnz = p;
  return C;
}

static csc_t csc_scale(const csc_t* A, cplx f) {
  csc_t C;
  C.n = A->n; C.nnz = A->nnz;
  C.colptr = (int64_t*)malloc(sizeof(int64_t) * (C.n + 1));
  C.rowval = (int64_t*)malloc(sizeof(int64_t) * (C.nnz + 1));
  C.nzval = (cplx*)malloc(sizeof(cplx) * (C.nnz + 1));
  memcpy(C.colptr, A->colptr, sizeof(int64_t) * (C.n + 1));
  memcpy(C.rowval, A->rowval, sizeof(int64_t) * C.nnz);
  for (int64_t i = 0; i < C.nnz; i++) C.nzval[i] = A->nzval[i] * f;
  return C;
}

/* foldl(kron, factors) */
static csc_t kron_fold(csc_t* factors, int nf) {
  csc_t acc = csc_kron(&factors[0], &factors[1]);
  for (int i = 2; i < nf; i++) {
    csc_t nxt = csc_kron(&acc, &factors[i]);
    csc_free(&acc);
    acc = nxt;
  }
  return acc;
}

static csc_t id2(void) {
  return csc_from_dense(M_I, 2);
}

/* The 4x4 with its two tensor factors swapped (hilbert.jl:215-237);
 * permutation [1,3,2,4] applied to rows and columns. */
static void swap_factors4(const double* m, double* out) {
  static const int perm[4] = {0, 2, 1, 3};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      int si = perm[i], sj = perm[j];
      out[2 * (i + 4 * j)] = m[2 * (si + 4 * sj)];
      out[2 * (i + 4 * j) + 1] = m[2 * (si + 4 * sj) + 1];
    }
}

/* hilbert(N, mat, qubit; control) (hilbert.jl:143-159) */
static csc_t hilbert1(int N, const double* mat, int qubit, int control) {
  init_tables();
  csc_t* f = (csc_t*)malloc(sizeof(csc_t) * N);
  for (int x = 1; x <= N; x++) {
    if (x == qubit) f[x - 1] = csc_from_dense(mat, 2);
    else if (control != -2 && x == control) f[x - 1] = csc_from_dense(M_P1, 2);
    else f[x - 1] = id2();
  }
  csc_t out;
  if (N == 1) { out = f[0]; free(f); return out; }
  out = kron_fold(f, N);
  for (int i = 0; i < N; i++) csc_free(&f[i]);
  if (control != -2) {
    for (int x = 1; x <= N; x++) f[x - 1] = (x == control) ? csc_from_dense(M_P0, 2) : id2();
    csc_t second = kron_fold(f, N);
    for (int i = 0; i < N; i++) csc_free(&f[i]);
    csc_t sum = csc_add(&out, &second);
    csc_free(&out); csc_free(&second);
    out = sum;
  }
  free(f);
  return out;
}

/* Pauli tensors for the non-adjacent reconstruction (decompose.jl:5-100). */
static const double PAULI[4][8] = {
    {1, 0, 0, 0, 0, 0, 1, 0},  /* I */
    {0, 0, 1, 0, 1, 0, 0, 0},  /* X */
    {0, 0, 0, 1, 0, -1, 0, 0}, /* Y = [0 -i; i 0] col-major: (0,0)=0,(1,0)=i,(0,1)=-i,(1,1)=0 */
    {1, 0, 0, 0, 0, 0, -1, 0}, /* Z */
};

static void kron_dense2(const double* a, const double* b, double* out4) {
  for (int ra = 0; ra < 2; ra++)
    for (int ca = 0; ca < 2; ca++)
      for (int rb = 0; rb < 2; rb++)
        for (int cb = 0; cb < 2; cb++) {
          cplx v = mget(a, 2, ra, ca) * mget(b, 2, rb, cb);
          int r = ra * 2 + rb, c = ca * 2 + cb;
          out4[2 * (r + 4 * c)] = creal(v);
          out4[2 * (r + 4 * c) + 1] = cimag(v);
        }
}

/* CCZX (hilbert.jl:73-103) */
static csc_t cczx(int is_ccz, int N, int i, int j, int k) {
  i = N - i; j = N - j; k = N - k;
  int64_t D = (int64_t)1 << N;
  csc_t H;
  H.n = D; H.nnz = D;
  H.colptr = (int64_t*)malloc(sizeof(int64_t) * (D + 1));
  H.rowval = (int64_t*)malloc(sizeof(int64_t) * D);
  H.nzval = (cplx*)malloc(sizeof(cplx) * D);
  /* one entry per column: row a -> column b; build as column-sorted. */
  int64_t* row_of_col = (int64_t*)malloc(sizeof(int64_t) * D);
  cplx* val_of_col = (cplx*)malloc(sizeof(cplx) * D);
  for (int64_t a = 0; a < D; a++) {
    if (is_ccz) {
      row_of_col[a] = a;
      val_of_col[a] = (((a >> i) & (a >> j) & (a >> k) & 1) == 1) ? -1.0 : 1.0;
    } else {
      if (((a >> i) & (a >> j) & 1) == 1) {
        int64_t b = a ^ (1ll << k);
        row_of_col[b] = a; val_of_col[b] = 1.0;
      } else {
        row_of_col[a] = a; val_of_col[a] = 1.0;
      }
    }
  }
  for (int64_t c = 0; c < D; c++) {
    H.colptr[c] = c; H.rowval[c] = row_of_col[c]; H.nzval[c] = val_of_col[c];
  }
  H.colptr[D] = D;
  free(row_of_col); free(val_of_col);
  return H;
}

/* hilbert(N, mat, qubit, target; control) (hilbert.jl:18-70) */
static int hilbert2(int N, const double* mat, int qubit, int target, int control, int is_cx, int is_cz, csc_t* out) {
  init_tables();
  if (N < qubit || N < target || N < control) return fail(BTO_ERR_ARG, "N must be larger than qubits");
  int distance = abs(qubit - target);
  double fm[32];
  if (qubit > target) swap_factors4(mat, fm); else memcpy(fm, mat, sizeof fm);
  if (control == -2) {
    if (distance == 1) {
      int nf = 0;
      csc_t* f = (csc_t*)malloc(sizeof(csc_t) * N);
      for (int x = 1; x <= N; x++) {
        if (x == target) continue;
        f[nf++] = (x == qubit) ? csc_from_dense(fm, 4) : id2();
      }
      /* when qubit > target the 4x4 sits at qubit's slot, which is the
       * (target, qubit) position once target is removed */
      if (nf == 1) { *out = f[0]; free(f); return BTO_OK; }
      *out = kron_fold(f, nf);
      for (int i = 0; i < nf; i++) csc_free(&f[i]);
      free(f);
      return BTO_OK;
    }
    /* non-adjacent: Pauli decomposition then reconstruction with identities
     * between (decompose.jl:71-100) */
    int lo = qubit < target ? qubit : target;
    int hi = qubit < target ? target : qubit;
    csc_t acc; int have = 0;
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        double P[32];
        kron_dense2(PAULI[a], PAULI[b], P);
        /* coefficient tr(A * P) / 4 */
        cplx tr = 0;
        for (int i2 = 0; i2 < 4; i2++)
          for (int k2 = 0; k2 < 4; k2++) tr += mget(fm, 4, i2, k2) * mget(P, 4, k2, i2);
        cplx coef = tr / 4.0;
        if (coef == 0) continue;
        csc_t* f = (csc_t*)malloc(sizeof(csc_t) * N);
        for (int x = 1; x <= N; x++) {
          if (x == lo) f[x - 1] = csc_from_dense(PAULI[a], 2);
          else if (x == hi) f[x - 1] = csc_from_dense(PAULI[b], 2);
          else f[x - 1] = id2();
        }
        csc_t term = kron_fold(f, N);
        for (int i = 0; i < N; i++) csc_free(&f[i]);
        free(f);
        csc_t sterm = csc_scale(&term, coef);
        csc_free(&term);
        if (!have) { acc = sterm; have = 1; }
        else { csc_t s2 = csc_add(&acc, &sterm); csc_free(&acc); csc_free(&sterm); acc = s2; }
      }
    *out = acc;
    return BTO_OK;
  }
  if (distance == 1) {
    int nf = 0;
    csc_t* f = (csc_t*)malloc(sizeof(csc_t) * N);
    for (int x = 1; x <= N; x++) {
      if (x == target) continue;
      if (x == qubit) f[nf++] = csc_from_dense(fm, 4);
      else if (x == control) f[nf++] = csc_from_dense(M_P1, 2);
      else f[nf++] = id2();
    }
    csc_t first = kron_fold(f, nf);
    for (int i = 0; i < nf; i++) csc_free(&f[i]);
    free(f);
    f = (csc_t*)malloc(sizeof(csc_t) * N);
    for (int x = 1; x <= N; x++) f[x - 1] = (x == control) ? csc_from_dense(M_P0, 2) : id2();
    csc_t second = kron_fold(f, N);
    for (int i = 0; i < N; i++) csc_free(&f[i]);
    free(f);
    *out = csc_add(&first, &second);
    csc_free(&first); csc_free(&second);
    return BTO_OK;
  }
  if (is_cx) { *out = cczx(0, N, qubit, control, target); return BTO_OK; }
  if (is_cz) { *out = cczx(1, N, qubit, control, target); return BTO_OK; }
  return fail(BTO_ERR_UNSUPPORTED, "Unsupported operation. Use 'CCZ' or 'CCX'.");
}

/* hilbert3(N, mat, first) (hilbert.jl:106-128) */
static csc_t hilbert3(int N, const double* mat, int first) {
  int nf = 0;
  csc_t* f = (csc_t*)malloc(sizeof(csc_t) * N);
  for (int x = 1; x <= N; x++) {
    if (x == first + 1 || x == first + 2) continue;
    f[nf++] = (x == first) ? csc_from_dense(mat, 8) : id2();
  }
  csc_t out;
  if (nf == 1) { out = f[0]; free(f); return out; }
  out = kron_fold(f, nf);
  for (int i = 0; i < nf; i++) csc_free(&f[i]);
  free(f);
  return out;
}

/* y = A x (dense vector), like SparseMatrixCSC * vector */
static void csc_matvec(const csc_t* A, const cplx* x, cplx* y) {
  memset(y, 0, sizeof(cplx) * A->n);
  for (int64_t c = 0; c < A->n; c++) {
    cplx xc = x[c];
    if (xc == 0) continue;
    for (int64_t p = A->colptr[c]; p < A->colptr[c + 1]; p++) y[A->rowval[p]] += A->nzval[p] * xc;
  }
}

/* rho <- A rho A' for dense col-major rho (two products, hilbert.jl:655-656) */
static void csc_sandwich_add(const csc_t* A, const cplx* rho, cplx* out, cplx* tmp) {
  int64_t D = A->n;
  /* tmp = A * rho  (column by column) */
  for (int64_t j = 0; j < D; j++) csc_matvec(A, rho + j * D, tmp + j * D);
  /* out += tmp * A' ; (tmp A')[r, c] = sum_k tmp[r,k] conj(A[c,k]) */
  for (int64_t k = 0; k < D; k++)
    for (int64_t p = A->colptr[k]; p < A->colptr[k + 1]; p++) {
      int64_t c = A->rowval[p];
      cplx a = conj(A->nzval[p]);
      const cplx* tk = tmp + k * D;
      cplx* oc = out + c * D;
      for (int64_t r = 0; r < D; r++) oc[r] += tk[r] * a;
    }
}

/* ------------------------------------------------------------------ */
/* Full-space operator of a k-qubit matrix placement in either mode    */
/* ------------------------------------------------------------------ */
static int build_full(int N, int q, const double* mat, int qubit, int target, int control, int is_cx, int is_cz,
                      csc_t* out) {
  if (q == 1) {
    if (N < qubit || N < control) return fail(BTO_ERR_ARG, "N must be larger than qubit");
    *out = hilbert1(N, mat, qubit, control);
    return BTO_OK;
  }
  if (q == 2) return hilbert2(N, mat, qubit, target, control, is_cx, is_cz, out);
  if (q == 3) {
    if (N < qubit + 2) return fail(BTO_ERR_ARG, "N must be larger than all three qubits");
    *out = hilbert3(N, mat, qubit);
    return BTO_OK;
  }
  return fail(BTO_ERR_ARG, "bad q");
}

/* Apply mat at (qubit[,target]; control) to the state, either mode. */
static int sv_place(cplx* s, int N, int q, const double* mat, int qubit, int target, int control, int is_cx,
                    int is_cz, int mode) {
  if (mode == 1) {
    csc_t A;
    int rc = build_full(N, q, mat, qubit, target, control, is_cx, is_cz, &A);
    if (rc) return rc;
    int64_t D = (int64_t)1 << N;
    cplx* y = (cplx*)malloc(sizeof(cplx) * D);
    csc_matvec(&A, s, y);
    memcpy(s, y, sizeof(cplx) * D);
    free(y);
    csc_free(&A);
    return BTO_OK;
  }
  uint64_t cmask = 0;
  if (q == 1) {
    if (N < qubit || N < control || qubit < 1) return fail(BTO_ERR_ARG, "N must be larger than qubit");
    if (control != -2) cmask = 1ull << (N - control);
    int bits[1] = {N - qubit};
    sv_apply_k(s, N, 1, mat, bits, cmask);
    return BTO_OK;
  }
  if (q == 2) {
    if (N < qubit || N < target || N < control) return fail(BTO_ERR_ARG, "N must be larger than qubits");
    if (control != -2) {
      if (abs(qubit - target) != 1 && !is_cx && !is_cz)
        return fail(BTO_ERR_UNSUPPORTED, "Unsupported operation. Use 'CCZ' or 'CCX'.");
      cmask = 1ull << (N - control);
    }
    int bits[2] = {N - qubit, N - target};
    sv_apply_k(s, N, 2, mat, bits, cmask);
    return BTO_OK;
  }
  if (q == 3) {
    if (N < qubit + 2) return fail(BTO_ERR_ARG, "N must be larger than all three qubits");
    int bits[3] = {N - qubit, N - qubit - 1, N - qubit - 2};
    sv_apply_k(s, N, 3, mat, bits, 0);
    return BTO_OK;
  }
  return fail(BTO_ERR_ARG, "bad q");
}

/* ------------------------------------------------------------------ */
/* Trajectory channel and Born measurement                             */
/* ------------------------------------------------------------------ */

/* __QuantumChannel_new_apply (struct.jl:31-41,51-55): probabilities from the
 * reduced rho with qubits in ascending order (Q3), sample, apply K at
 * (qubit[, target]) with its first factor on `qubit`, normalise. */
static int sv_channel(cplx* s, int N, int q, int nk, const double* kraus, int qubit, int target, draws_t* d,
                      int mode, int* chosen_out, double* probs_out) {
  int dim = 1 << q;
  int keep[3];
  if (q == 1) keep[0] = qubit;
  else if (q == 2) { keep[0] = qubit < target ? qubit : target; keep[1] = qubit < target ? target : qubit; }
  else { keep[0] = qubit; keep[1] = qubit + 1; keep[2] = qubit + 2; }
  for (int i = 0; i < q; i++)
    if (keep[i] < 1 || keep[i] > N) return fail(BTO_ERR_ARG, "qubit out of range");
  cplx rho[64];
  sv_reduced(s, N, q, keep, rho);
  double probs[64];
  kraus_probs(rho, dim, nk, kraus, probs);
  double u;
  int rc = next_draw(d, &u);
  if (rc) return rc;
  int ind = weighted_sample(probs, nk, u, d);
  if (chosen_out) *chosen_out = ind;
  if (probs_out) memcpy(probs_out, probs, sizeof(double) * nk);
  const double* K = kraus + (size_t)ind * 2 * dim * dim;
  rc = sv_place(s, N, q, K, qubit, q == 2 ? target : -1, -2, 0, 0, mode);
  if (rc) return rc;
  sv_normalize(s, N);
  return BTO_OK;
}

/* born_measure_Z (hilbert.jl:682-696): prob0 = rho_q[1,1];
 * ind = rand() < prob0 ? 0 : 1; normalize(P_ind psi). */
static int sv_born_z(cplx* s, int N, int qubit, draws_t* d, int mode, int* ind_out, double* prob0_out) {
  init_tables();
  cplx rho[4];
  int keep[1] = {qubit};
  if (qubit < 1 || qubit > N) return fail(BTO_ERR_ARG, "qubit out of range");
  sv_reduced(s, N, 1, keep, rho);
  double prob0 = creal(rho[0]);
  double u;
  int rc = next_draw(d, &u);
  if (rc) return rc;
  if (fabs(u - prob0) <= d->tol) d->boundary++;
  int ind = (u < prob0) ? 0 : 1;
  rc = sv_place(s, N, 1, ind == 0 ? M_P0 : M_P1, qubit, -1, -2, 0, 0, mode);
  if (rc) return rc;
  sv_normalize(s, N);
  *ind_out = ind;
  if (prob0_out) *prob0_out = prob0;
  return BTO_OK;
}

static void adjoint2(const double* m, double* out) {
  for (int r = 0; r < 2; r++)
    for (int c = 0; c < 2; c++) {
      cplx v = conj(mget(m, 2, c, r));
      out[2 * (r + 2 * c)] = creal(v);
      out[2 * (r + 2 * c) + 1] = cimag(v);
    }
}

/* _born_measure / __ifOp_apply rotation + Z measurement + rotate back
 * (hilbert.jl:669-679, struct.jl:565-594).  MR draws its basis first
 * (struct.jl:568): basis = [MX, MY, MZ][floor(3u)]. */
static int sv_measure(cplx* s, int N, int basis, int qubit, draws_t* d, int mode, int* ind_out) {
  if (basis == BTO_MR) {
    double ub;
    int rc = next_draw(d, &ub);
    if (rc) return rc;
    int pick = (int)floor(ub * 3.0);
    if (pick > 2) pick = 2;
    static const int order[3] = {BTO_MX, BTO_MY, BTO_MZ};
    basis = order[pick];
  }
  const double* R = measurement_mat(basis);
  int rc = sv_place(s, N, 1, R, qubit, -1, -2, 0, 0, mode);
  if (rc) return rc;
  rc = sv_born_z(s, N, qubit, d, mode, ind_out, NULL);
  if (rc) return rc;
  double Rd[8];
  adjoint2(R, Rd);
  return sv_place(s, N, 1, Rd, qubit, -1, -2, 0, 0, mode);
}

/* ------------------------------------------------------------------ */
/* apply(state, op; noise, track_measurements) (hilbert.jl:469-515)   */
/* ------------------------------------------------------------------ */
typedef struct {
  int* outcomes; int64_t cap; int64_t n;
} outs_t;

static int sv_apply_noise(cplx* s, int N, const bto_op* op, const bto_noise* nm, draws_t* d, int mode) {
  /* apply_noise (hilbert.jl:322-340); the SV method does not guard a `false`
   * channel, which raises in the reference. */
  if (!nm || !op->noisy) return BTO_OK;
  if (op->q == 1) {
    if (op->control == -2) {
      if (nm->n1 == 0) return fail(BTO_ERR_UNSUPPORTED, "noise model has no 1-qubit channel");
      return sv_channel(s, N, 1, nm->n1, nm->k1, op->qubit, -1, d, mode, NULL, NULL);
    }
    if (nm->n2 == 0) return fail(BTO_ERR_UNSUPPORTED, "noise model has no 2-qubit channel");
    return sv_channel(s, N, 2, nm->n2, nm->k2, op->control, op->qubit, d, mode, NULL, NULL);
  }
  if (op->q == 2) {
    if (nm->n2 == 0) return fail(BTO_ERR_UNSUPPORTED, "noise model has no 2-qubit channel");
    return sv_channel(s, N, 2, nm->n2, nm->k2, op->qubit, op->target, d, mode, NULL, NULL);
  }
  return BTO_OK;
}

static int sv_apply_one(cplx* s, int N, const bto_op* ops, int idx, const bto_noise* nm, draws_t* d, int mode,
                        outs_t* outs, int track) {
  const bto_op* op = &ops[idx];
  int rc = BTO_OK;
  switch (op->kind) {
    case BTO_GATE:
      rc = sv_place(s, N, op->q, op->mat, op->qubit, op->q >= 2 ? op->target : -1, op->control, op->is_cx, op->is_cz,
                    mode);
      if (rc) return rc;
      return sv_apply_noise(s, N, op, nm, d, mode);
    case BTO_CHANNEL: /* OpQC: noisy == false, no noise follows (struct.jl:214,226) */
      return sv_channel(s, N, op->q, op->nkraus, op->kraus, op->qubit, op->target, d, mode, NULL, NULL);
    case BTO_RESET: { /* _reset_Z (hilbert.jl:752-759); outcome not tracked */
      int ind;
      rc = sv_born_z(s, N, op->qubit, d, mode, &ind, NULL);
      if (rc) return rc;
      if (ind == 1) rc = sv_place(s, N, 1, M_X, op->qubit, -1, -2, 0, 0, mode);
      return rc;
    }
    case BTO_MEASURE: {
      int ind;
      rc = sv_measure(s, N, op->basis, op->qubit, d, mode, &ind);
      if (rc) return rc;
      if (track) {
        if (outs->n >= outs->cap) return fail(BTO_ERR_ARG, "outcome buffer too small");
        outs->outcomes[outs->n++] = ind;
      }
      return BTO_OK;
    }
    case BTO_IFOP: {
      int ind;
      rc = sv_measure(s, N, op->basis, op->qubit, d, mode, &ind);
      if (rc) return rc;
      if (track) {
        if (outs->n >= outs->cap) return fail(BTO_ERR_ARG, "outcome buffer too small");
        outs->outcomes[outs->n++] = ind;
      }
      int b = ind == 0 ? op->if0_begin : op->if1_begin;
      int e = ind == 0 ? op->if0_end : op->if1_end;
      /* branch ops: apply(state, ifop; noise=noise) without tracking (struct.jl:587-590) */
      for (int j = b; j < e; j++) {
        rc = sv_apply_one(s, N, ops, j, nm, d, mode, outs, 0);
        if (rc) return rc;
      }
      return BTO_OK;
    }
    default:
      return fail(BTO_ERR_ARG, "unknown op kind");
  }
}

/* Run ops[0..ntop) on state s (length 2^N, complex128 interleaved).
 * Returns the number of uniforms consumed in *used and outcomes tracked in
 * outcomes[0..*nout). */
int bto_run_sv(int N, double* state, const bto_op* ops, int ntop, const bto_noise* nm, const double* uniforms,
               int64_t nuni, double tol, int mode, int* outcomes, int64_t outcap, int64_t* nout, int64_t* used,
               int64_t* boundary, int64_t* overflow) {
  draws_t d = {uniforms, nuni, 0, 0, tol, 0};
  outs_t o = {outcomes, outcap, 0};
  cplx* s = (cplx*)state;
  for (int i = 0; i < ntop; i++) {
    int rc = sv_apply_one(s, N, ops, i, nm, &d, mode, &o, 1);
    if (rc) return rc;
  }
  if (nout) *nout = o.n;
  if (used) *used = d.pos;
  if (boundary) *boundary = d.boundary;
  if (overflow) *overflow = d.overflow;
  return BTO_OK;
}

/* Many trajectories from |0...0>, uniform stream bto_uniform(seed, t, .) for
 * t in [traj0, traj0+ntraj).  outcomes: ntraj x outcap.  Optionally keeps the
 * final states (ntraj x 2^N complex128).  Trajectories run one after the other
 * with OpenMP inside each op (fast mode) so memory stays at one state. */
int bto_run_traj(int N, const bto_op* ops, int ntop, const bto_noise* nm, uint64_t seed, int64_t traj0,
                 int64_t ntraj, int64_t max_draws, double tol, int mode, int* outcomes, int64_t outcap,
                 int64_t* nouts, double* final_states, int64_t* boundary_total) {
  int64_t D = (int64_t)1 << N;
  cplx* s = (cplx*)malloc(sizeof(cplx) * D);
  double* u = (double*)malloc(sizeof(double) * (max_draws > 0 ? max_draws : 1));
  if (!s || !u) { free(s); free(u); return fail(BTO_ERR_NOMEM, "out of memory"); }
  int64_t btot = 0;
  for (int64_t t = 0; t < ntraj; t++) {
    memset(s, 0, sizeof(cplx) * D);
    s[0] = 1.0;
    bto_uniforms(seed, (uint64_t)(traj0 + t), max_draws, u);
    int64_t nout = 0, used = 0, b = 0, ov = 0;
    int rc = bto_run_sv(N, (double*)s, ops, ntop, nm, u, max_draws, tol, mode, outcomes + t * outcap, outcap, &nout,
                        &used, &b, &ov);
    if (rc) { free(s); free(u); return rc; }
    if (nouts) nouts[t] = nout;
    btot += b;
    if (final_states) memcpy(final_states + 2 * t * D, s, sizeof(cplx) * D);
  }
  if (boundary_total) *boundary_total = btot;
  free(s); free(u);
  return BTO_OK;
}

/* ------------------------------------------------------------------ */
/* Density matrix path (hilbert.jl:639-666, struct.jl:58-76)           */
/* rho is column-major 2^N x 2^N: rho[r + c*D]                          */
/* ------------------------------------------------------------------ */

/* rho <- E rho E' in place, fast mode: rows with mat, columns with conj(mat) */
static int dm_place_sandwich(cplx* rho, int N, int q, const double* mat, int qubit, int target, int control,
                             int is_cx, int is_cz, cplx* acc /* NULL: in place; else acc += E rho E' */,
                             int mode) {
  int64_t D = (int64_t)1 << N;
  if (mode == 1) {
    csc_t A;
    int rc = build_full(N, q, mat, qubit, target, control, is_cx, is_cz, &A);
    if (rc) return rc;
    cplx* tmp = (cplx*)malloc(sizeof(cplx) * D * D);
    if (acc) {
      csc_sandwich_add(&A, rho, acc, tmp);
    } else {
      cplx* out = (cplx*)calloc((size_t)(D * D), sizeof(cplx));
      csc_sandwich_add(&A, rho, out, tmp);
      memcpy(rho, out, sizeof(cplx) * D * D);
      free(out);
    }
    free(tmp);
    csc_free(&A);
    return BTO_OK;
  }
  /* fast: treat vec(rho) as a 2N-qubit vector; row bit b <-> bit b, column
   * bit b <-> bit b+N.  E acts on row bits, conj(E) on column bits. */
  int dim = 1 << q;
  double cm[128];
  for (int i = 0; i < dim * dim; i++) { cm[2 * i] = mat[2 * i]; cm[2 * i + 1] = -mat[2 * i + 1]; }
  int bits[3];
  uint64_t cmask_r = 0, cmask_c = 0;
  if (q == 1) {
    if (N < qubit || N < control || qubit < 1) return fail(BTO_ERR_ARG, "N must be larger than qubit");
    bits[0] = N - qubit;
  } else if (q == 2) {
    if (N < qubit || N < target || N < control) return fail(BTO_ERR_ARG, "N must be larger than qubits");
    if (control != -2 && abs(qubit - target) != 1 && !is_cx && !is_cz)
      return fail(BTO_ERR_UNSUPPORTED, "Unsupported operation. Use 'CCZ' or 'CCX'.");
    bits[0] = N - qubit; bits[1] = N - target;
  } else {
    if (N < qubit + 2) return fail(BTO_ERR_ARG, "N must be larger than all three qubits");
    bits[0] = N - qubit; bits[1] = N - qubit - 1; bits[2] = N - qubit - 2;
  }
  if (control != -2) { cmask_r = 1ull << (N - control); cmask_c = 1ull << (2 * N - control); }
  cplx* work = rho;
  if (acc) {
    work = (cplx*)malloc(sizeof(cplx) * D * D);
    memcpy(work, rho, sizeof(cplx) * D * D);
  }
  int cbits[3];
  for (int i = 0; i < q; i++) cbits[i] = bits[i] + N;
  sv_apply_k(work, 2 * N, q, mat, bits, cmask_r);
  sv_apply_k(work, 2 * N, q, cm, cbits, cmask_c);
  if (acc) {
    for (int64_t i = 0; i < D * D; i++) acc[i] += work[i];
    free(work);
  }
  return BTO_OK;
}

static int dm_channel(cplx* rho, int N, int q, int nk, const double* kraus, int qubit, int target, int mode) {
  if (q == 3) return fail(BTO_ERR_UNSUPPORTED, "3-qubit channels have no density-matrix method");
  int64_t D = (int64_t)1 << N;
  int dim = 1 << q;
  cplx* acc = (cplx*)calloc((size_t)(D * D), sizeof(cplx));
  for (int k = 0; k < nk; k++) {
    int rc = dm_place_sandwich(rho, N, q, kraus + (size_t)k * 2 * dim * dim, qubit, q == 2 ? target : -1, -2, 0, 0,
                               acc, mode);
    if (rc) { free(acc); return rc; }
  }
  memcpy(rho, acc, sizeof(cplx) * D * D);
  free(acc);
  return BTO_OK;
}

static int dm_apply_noise(cplx* rho, int N, const bto_op* op, const bto_noise* nm, int mode) {
  if (!nm || !op->noisy) return BTO_OK;
  if (op->q == 1) {
    if (op->control == -2) {
      if (nm->n1 == 0) return fail(BTO_ERR_UNSUPPORTED, "noise model has no 1-qubit channel");
      return dm_channel(rho, N, 1, nm->n1, nm->k1, op->qubit, -1, mode);
    }
    if (nm->n2 == 0) return fail(BTO_ERR_UNSUPPORTED, "noise model has no 2-qubit channel");
    return dm_channel(rho, N, 2, nm->n2, nm->k2, op->control, op->qubit, mode);
  }
  if (op->q == 2) {
    if (nm->n2 == 0) return fail(BTO_ERR_UNSUPPORTED, "noise model has no 2-qubit channel");
    return dm_channel(rho, N, 2, nm->n2, nm->k2, op->qubit, op->target, mode);
  }
  return BTO_OK;
}

int bto_run_dm(int N, double* rho_d, const bto_op* ops, int ntop, const bto_noise* nm, int mode) {
  cplx* rho = (cplx*)rho_d;
  for (int i = 0; i < ntop; i++) {
    const bto_op* op = &ops[i];
    int rc;
    switch (op->kind) {
      case BTO_GATE:
        rc = dm_place_sandwich(rho, N, op->q, op->mat, op->qubit, op->q >= 2 ? op->target : -1, op->control,
                               op->is_cx, op->is_cz, NULL, mode);
        if (rc) return rc;
        rc = dm_apply_noise(rho, N, op, nm, mode);
        if (rc) return rc;
        break;
      case BTO_CHANNEL:
      case BTO_RESET: /* OpQC("RES") is amplitude damping with gamma = 1 on rho (struct.jl:423) */
        rc = dm_channel(rho, N, op->q, op->nkraus, op->kraus, op->qubit, op->target, mode);
        if (rc) return rc;
        break;
      case BTO_MEASURE:
      case BTO_IFOP:
        return fail(BTO_ERR_UNSUPPORTED, "fix this: density-matrix measurement is not implemented");
      default:
        return fail(BTO_ERR_ARG, "unknown op kind");
    }
  }
  return BTO_OK;
}

/* ------------------------------------------------------------------ */
/* Exposed primitives for pinning tests                                 */
/* ------------------------------------------------------------------ */
int bto_apply_gate(int N, double* state, int q, const double* mat, int qubit, int target, int control, int is_cx,
                   int is_cz, int mode) {
  return sv_place((cplx*)state, N, q, mat, qubit, target, control, is_cx, is_cz, mode);
}

/* Full-space operator as a dense col-major matrix (reference mode build). */
int bto_full_operator(int N, int q, const double* mat, int qubit, int target, int control, int is_cx, int is_cz,
                      double* out) {
  csc_t A;
  int rc = build_full(N, q, mat, qubit, target, control, is_cx, is_cz, &A);
  if (rc) return rc;
  int64_t D = (int64_t)1 << N;
  memset(out, 0, sizeof(double) * 2 * D * D);
  for (int64_t c = 0; c < D; c++)
    for (int64_t p = A.colptr[c]; p < A.colptr[c + 1]; p++) {
      out[2 * (A.rowval[p] + c * D)] += creal(A.nzval[p]);
      out[2 * (A.rowval[p] + c * D) + 1] += cimag(A.nzval[p]);
    }
  csc_free(&A);
  return BTO_OK;
}

int bto_partial_trace(int N, const double* state, int k, const int* keep_sorted, double* out) {
  sv_reduced((const cplx*)state, N, k, keep_sorted, (cplx*)out);
  return BTO_OK;
}

int bto_calc_prob(int N, const double* state, int q, int nk, const double* kraus, int qubit, int target,
                  double* probs) {
  int dim = 1 << q;
  int keep[3];
  if (q == 1) keep[0] = qubit;
  else if (q == 2) { keep[0] = qubit < target ? qubit : target; keep[1] = qubit < target ? target : qubit; }
  else { keep[0] = qubit; keep[1] = qubit + 1; keep[2] = qubit + 2; }
  cplx rho[64];
  sv_reduced((const cplx*)state, N, q, keep, rho);
  kraus_probs(rho, dim, nk, kraus, probs);
  return BTO_OK;
}

int bto_weighted_sample(const double* probs, int n, double u) {
  return weighted_sample(probs, n, u, NULL);
}

void bto_probabilities(int N, const double* state, double* out) {
  int64_t D = (int64_t)1 << N;
  for (int64_t i = 0; i < D; i++) out[i] = state[2 * i] * state[2 * i] + state[2 * i + 1] * state[2 * i + 1];
}

void bto_dm_trace_purity(int N, const double* rho, double* trace, double* purity) {
  int64_t D = (int64_t)1 << N;
  double t = 0, p = 0;
  for (int64_t i = 0; i < D; i++) t += rho[2 * (i + i * D)];
  for (int64_t i = 0; i < D * D; i++) p += rho[2 * i] * rho[2 * i] + rho[2 * i + 1] * rho[2 * i + 1];
  *trace = t;
  *purity = p;
}

int bto_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Thread count of the fast mode's parallel loops (the reference algorithm
 * mode is timed on one thread, as the reference runs). */
void bto_set_num_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}
