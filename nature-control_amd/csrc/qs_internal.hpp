// qs_internal.hpp -- shared host/device definitions of libqsolver (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdio>
#include <string>
#include <type_traits>

#include "qsolver.h"

// ---------------------------------------------------------------------------
// Error handling: thread-local message + status codes (never abort).
// ---------------------------------------------------------------------------
namespace qs {

void set_error(const char* fmt, ...);
int fail(int status, const char* fmt, ...);

#define QS_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ::qs::fail(QS_ERR_HIP, "%s:%d %s -> %s", __FILE__, __LINE__, #call,       \
                        hipGetErrorString(e_));                                        \
  } while (0)

#define QS_CHECK_LAUNCH() QS_HIP(hipGetLastError())

#define QS_TRY(expr)              \
  do {                            \
    int s_ = (expr);              \
    if (s_ != QS_OK) return s_;   \
  } while (0)

// Host range for rocprofv3 --marker-trace (roctx): the span of a call on the
// host (launch and enqueue), nested as the calls nest.  Without a profiler
// attached a push / pop is a call into an idle stub.
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  Range(const char* fmt, int a) {
    char b[64];
    std::snprintf(b, sizeof b, fmt, a);
    roctxRangePushA(b);
  }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

inline bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
// Mesh sides the per-op transforms support: 2^a in [16, 2048] and 3 * 2^a in
// [48, 1536].  The fused kernels take the powers of two up to 1024.
inline bool supported_side(int n) {
  return (is_pow2(n) && n >= 16 && n <= 2048) || (n % 3 == 0 && is_pow2(n / 3) && n >= 48 && n <= 1536);
}
inline bool fused_side(int n) { return is_pow2(n) && n >= 16 && n <= 1024; }
inline bool fused_tri_side(int n) { return n % 3 == 0 && is_pow2(n / 3) && n >= 48 && n <= 768; }
inline int ilog2(int n) { int l = 0; while ((1 << l) < n) ++l; return l; }

}  // namespace qs

// Build-private part of the Mesh.
struct qs_mesh_priv {
  hipStream_t stream = nullptr;
  double2* twx = nullptr;   // W_Nx^m = exp(-2 pi i m/Nx), m < Nx
  double2* twy = nullptr;   // W_Ny^m
  double* kx1 = nullptr;    // kx(i), i < Nxh   (1-D, separable tables)
  double* ky1 = nullptr;    // ky(j), j < Ny
  double2* tw1024 = nullptr; // W_1024^m, m < 1024 (fused kernels use the first 256)
  double2* tw768 = nullptr;  // W_768^m, m < 768 (fused kernels of 3 * 2^a sides use the first 192)
  int ncu = 0;               // compute units of the device the mesh was created on
  double* diag = nullptr;    // partial sums of the diagnostics (qs_diag.hip), grown on demand
  size_t diag_n = 0;
};

namespace qs {

inline hipStream_t stream_of(const Mesh* m) { return m->priv->stream; }
int check_mesh(const Mesh* m);

// Per-op 2-D transforms in the reference layouts (qs_fft.hip).
int fft_r2c_2d(const Mesh* m, const double* f, double2* mix, double2* ft);
int fft_c2r_2d(const Mesh* m, double2* ft, double2* mix, double* f, bool symmetrise_input);

// Coefficient-aware elementwise forms (qs_fldops.hip).
int qlin_c(double* IF, double* IFh, double dt, double DQ, double aQ, const Mesh* mesh);
int single1_c(double* s1, const double* r, const double* S, double lambda, double cQ, double aQ,
              const Mesh* mesh);
int S2r_c(double* rs, const double* r, const double* S, double cQ, const Mesh* mesh);

// Generalised Q-equation coefficients; all 1 (and advect_w) = reference.
struct Coef {
  double DQ = 1.0, aQ = 1.0, cQ = 1.0;
  bool advect_w = true;
};

// Coefficient-aware composites of friction/Qsolver.cuh (qs_physics.cpp).
int single_c(Field* s0, Field* r, Field* S, Qreal lambda, const Coef& c);
int cross_c(Field* cs, Field* r1, Field* r2, Field* aux, const Coef& c);
int p11func_c(Field* p11, Field* r1, Field* S, Field* Ra, Qreal lambda, const Coef& c);
int p12func_c(Field* p12, Field* r1, Field* r2, Field* S, Field* Ra, Qreal lambda, Field* aux,
              Field* aux1, const Coef& c);
int p21func_c(Field* p21, Field* r1, Field* r2, Field* S, Field* Ra, Qreal lambda, Field* aux,
              Field* aux1, const Coef& c);
int wnonl_func_c(Field* wnonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr,
                 Field* Ra, Field* S, Field* p11, Field* p12, Field* p21, Qreal Re, Qreal Er,
                 Qreal lambda, Field* aux, Field* aux1, const Coef& c);
int r1nonl_func_c(Field* r1nonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr,
                  Field* S, Qreal lambda, Field* aux, Field* aux1, const Coef& c);
int r2nonl_func_c(Field* r2nonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr,
                  Field* S, Qreal lambda, Field* aux, Field* aux1, const Coef& c);

// Host twiddle table W_n^m = exp(-2 pi i m / n), accurate to ~0.5 ulp.
void host_twiddles(int n, double2* out);

// ---------------------------------------------------------------------------
// Device complex helpers (double2 = {re, im}).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by i (SIGN>0) or -i (SIGN<0)
template <int SIGN>
__device__ __forceinline__ double2 mul_pmi(double2 a) {
  return SIGN > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// Small DFTs in registers.  SIGN = -1 forward (e^{-i}), +1 inverse.
// ---------------------------------------------------------------------------
template <int SIGN>
__device__ __forceinline__ void dft2(double2& a, double2& b) {
  double2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

template <int SIGN>
__device__ __forceinline__ void dft4(double2* v) {
  double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
  double2 s13 = cadd(v[1], v[3]), d13 = mul_pmi<SIGN>(csub(v[1], v[3]));
  v[0] = cadd(s02, s13);
  v[2] = csub(s02, s13);
  v[1] = cadd(d02, d13);
  v[3] = csub(d02, d13);
}

// radix-8 as 2 x 4 with the W8 twiddles applied between the stages.
template <int SIGN>
__device__ __forceinline__ void dft8(double2* v) {
  constexpr double H = 0.70710678118654752440;  // cos(pi/4)
  double2 a[4] = {v[0], v[2], v[4], v[6]};
  double2 b[4] = {v[1], v[3], v[5], v[7]};
  dft4<SIGN>(a);
  dft4<SIGN>(b);
  // b[k] *= W8^k (SIGN)
  b[1] = SIGN < 0 ? make_double2(H * (b[1].x + b[1].y), H * (b[1].y - b[1].x))
                  : make_double2(H * (b[1].x - b[1].y), H * (b[1].y + b[1].x));
  b[2] = mul_pmi<SIGN>(b[2]);
  b[3] = SIGN < 0 ? make_double2(H * (b[3].y - b[3].x), -H * (b[3].x + b[3].y))
                  : make_double2(-H * (b[3].x + b[3].y), H * (b[3].x - b[3].y));
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[k] = cadd(a[k], b[k]);
    v[k + 4] = csub(a[k], b[k]);
  }
}

// radix-3: y0 = a + b + c, y1,2 = a - (b + c)/2 -/+ SIGN' i (sqrt(3)/2)(b - c)
// with W3 = exp(SIGN 2 pi i / 3).
template <int SIGN>
__device__ __forceinline__ void dft3(double2* v) {
  constexpr double S3 = 0.86602540378443864676;  // sin(pi/3)
  const double2 s = cadd(v[1], v[2]);
  const double2 t = make_double2(v[0].x - 0.5 * s.x, v[0].y - 0.5 * s.y);
  const double2 d = make_double2(S3 * (v[1].x - v[2].x), S3 * (v[1].y - v[2].y));
  const double2 id = mul_pmi<SIGN>(d);   // SIGN i d
  v[0] = cadd(v[0], s);
  v[1] = cadd(t, id);
  v[2] = csub(t, id);
}

template <int R, int SIGN>
__device__ __forceinline__ void dft_small(double2* v) {
  if constexpr (R == 2) dft2<SIGN>(v[0], v[1]);
  else if constexpr (R == 3) dft3<SIGN>(v);
  else if constexpr (R == 4) dft4<SIGN>(v);
  else if constexpr (R == 8) dft8<SIGN>(v);
  else static_assert(R == 2 || R == 3 || R == 4 || R == 8, "radix");
}

// Radix of the Stockham pass that meets M = N / P remaining points: RMAX
// while it divides M, then a power of two below RMAX, then 4, 2, and finally
// 3 (N = 2^a or 3 * 2^a).
constexpr int stockham_radix(int M, int RMAX) {
  return M % RMAX == 0 ? RMAX : ((M & (M - 1)) == 0 ? M : (M % 4 == 0 ? 4 : (M % 2 == 0 ? 2 : 3)));
}

// ---------------------------------------------------------------------------
// Twiddle accessors: W_N^m = exp(-2 pi i m / N), 0 <= m < N.
//   TwFull:    table of all N entries (global or LDS).
//   TwQuarter: table of the first N/4 entries; W^m = (-i)^q W^r with
//              q = m / (N/4), r = m % (N/4) (exact quadrant rotation).
// ---------------------------------------------------------------------------
struct TwFull {
  const double2* p;
  __device__ __forceinline__ double2 operator()(int m) const { return p[m]; }
};

template <int N>
struct TwQuarter {
  const double2* p;
  __device__ __forceinline__ double2 operator()(int m) const {
    constexpr int Q4 = N / 4;
    const int q = (m / Q4) & 3, r = m % Q4;
    const double2 w = p[r];
    // (-i)^q w without branches: q odd swaps the parts; signs
    // q=0 (+,+), q=1 (+,-), q=2 (-,-), q=3 (-,+).
    const bool odd = q & 1;
    const double a = odd ? w.y : w.x;
    const double b = odd ? w.x : w.y;
    const double sa = (q & 2) ? -1.0 : 1.0;
    const double sb = ((q ^ (q >> 1)) & 1) ? -1.0 : 1.0;
    return make_double2(sa * a, sb * b);
  }
};

// Cooperative copy of the first n twiddles into LDS (caller synchronises).
__device__ __forceinline__ void load_twiddles(double2* dst, const double2* __restrict__ src, int n,
                                              int tid, int nt) {
  for (int m = tid; m < n; m += nt) dst[m] = src[m];
}

// ---------------------------------------------------------------------------
// In-LDS Stockham FFT of NV vectors of length N (element e of vector v at
// s[v*LD + e]), executed by NT threads.  Radix RMAX passes while they fit,
// then one smaller pass.  Output in natural order, in place.  Ends with a
// barrier.
// ---------------------------------------------------------------------------
template <int N, int R, int P, int SIGN, int NV, int LD, int NT, class TW>
__device__ __forceinline__ void stockham_pass(double2* s, const TW& tw, int tid) {
  constexpr int Q = N / R;             // butterflies per vector
  constexpr int NB = NV * Q;           // butterflies in total
  constexpr int PER = (NB + NT - 1) / NT;
  constexpr int STEP = N / (P * R);
  double2 v[PER][R];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = tid + q * NT;
    if (NB % NT == 0 || b < NB) {
      const int vec = b / Q, i = b % Q;
      const int k = i & (P - 1);
      const double2* src = s + vec * LD + i;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = src[r * Q];
      if (P > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          double2 w = tw(r * k * STEP);
          if (SIGN > 0) w.y = -w.y;
          v[q][r] = cmul(v[q][r], w);
        }
      }
      dft_small<R, SIGN>(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int b = tid + q * NT;
    if (NB % NT == 0 || b < NB) {
      const int vec = b / Q, i = b % Q;
      const int k = i & (P - 1);
      double2* dst = s + vec * LD + (i - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) dst[r * P] = v[q][r];
    }
  }
  __syncthreads();
}

template <int N, int RMAX, int P, int SIGN, int NV, int LD, int NT, class TW>
__device__ __forceinline__ void stockham_from(double2* s, const TW& tw, int tid) {
  if constexpr (P < N) {
    constexpr int R = stockham_radix(N / P, RMAX);
    stockham_pass<N, R, P, SIGN, NV, LD, NT>(s, tw, tid);
    stockham_from<N, RMAX, P * R, SIGN, NV, LD, NT>(s, tw, tid);
  }
}

template <int N, int RMAX, int SIGN, int NV, int LD, int NT, class TW>
__device__ __forceinline__ void fft_lds(double2* s, const TW& tw, int tid) {
  stockham_from<N, RMAX, 1, SIGN, NV, LD, NT>(s, tw, tid);
}

// WG id -> logical id such that the WGs one XCD runs (ids congruent mod 8,
// dispatched round-robin over the 8 XCDs) get a contiguous logical range, so
// neighbouring column tiles share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int G) {
  const int per = G / 8, rem = G % 8;
  const int x = b % 8, slot = b / 8;
  return x * per + (x < rem ? x : rem) + slot;
}

// Global-memory accessors.  Pointers read from a descriptor are generic to the
// compiler, which then emits flat instructions: they count against both the
// vector-memory and the LDS counters, so every LDS wait also waits for them.
// Casting to the global address space gives global_load / global_store.
#define QS_GLOBAL __attribute__((address_space(1)))
typedef double qs_d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 gload(const double2* p) {
  const qs_d2v v = *(const QS_GLOBAL qs_d2v*)p;
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ double gload(const double* p) { return *(const QS_GLOBAL double*)p; }
__device__ __forceinline__ void gstore(double2* p, double2 v) {
  const qs_d2v t = {v.x, v.y};
  *(QS_GLOBAL qs_d2v*)p = t;
}

}  // namespace qs

namespace qs {

// ---------------------------------------------------------------------------
// Radix-16 DFT in registers: 4 x 4 with the W16 twiddles between the stages.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 w16(int m, int sign) {
  constexpr double C1 = 0.92387953251128675613;  // cos(pi/8)
  constexpr double S1 = 0.38268343236508977173;  // sin(pi/8)
  constexpr double H = 0.70710678118654752440;   // cos(pi/4)
  double c = 1.0, s = 0.0;                        // W16^m = c - i s (forward)
  switch (m) {
    case 1: c = C1; s = S1; break;
    case 2: c = H; s = H; break;
    case 3: c = S1; s = C1; break;
    case 4: c = 0.0; s = 1.0; break;
    case 6: c = -H; s = H; break;
    case 9: c = -C1; s = -S1; break;
    default: break;
  }
  return make_double2(c, sign < 0 ? -s : s);
}

template <int SIGN>
__device__ __forceinline__ void dft16(double2* v) {
  double2 a[4][4];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    double2 t[4] = {v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]};
    dft4<SIGN>(t);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) a[n2][k1] = (n2 * k1 == 0) ? t[k1] : cmul(t[k1], w16(n2 * k1, SIGN));
  }
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    double2 t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    dft4<SIGN>(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
  }
}

// ---------------------------------------------------------------------------
// Radix-12 DFT in registers: 3 x 4 with the W12 twiddles between the stages
// (n = n2 + 4 n1, k = k1 + 3 k2).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 w12(int m, int sign) {
  constexpr double C = 0.86602540378443864676;   // cos(pi/6)
  double c = 1.0, s = 0.0;                        // W12^m = c - i s (forward)
  switch (m) {
    case 1: c = C; s = 0.5; break;
    case 2: c = 0.5; s = C; break;
    case 3: c = 0.0; s = 1.0; break;
    case 4: c = -0.5; s = C; break;
    case 6: c = -1.0; s = 0.0; break;
    default: break;
  }
  return make_double2(c, sign < 0 ? -s : s);
}

template <int SIGN>
__device__ __forceinline__ void dft12(double2* v) {
  double2 a[4][3];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    double2 t[3] = {v[n2], v[n2 + 4], v[n2 + 8]};
    dft3<SIGN>(t);
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) a[n2][k1] = (n2 * k1 == 0) ? t[k1] : cmul(t[k1], w12(n2 * k1, SIGN));
  }
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    double2 t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
    dft4<SIGN>(t);
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) v[k1 + 3 * k2] = t[k2];
  }
}

// ---------------------------------------------------------------------------
// Lane layout of the per-wavefront FFT of length-N vectors: E = wave_e(N)
// elements per lane, LPC = N/E lanes per vector, NV = 64/LPC vectors per wave
// (one wave holds NV whole vectors), vector stride VS in a wave buffer
// (swizzled).  The first pass is a radix-E DFT in registers, then Stockham
// passes through LDS of radix RMAX (16 for E = 16, E for E < 16, 4 for
// 3 * 2^a) while it divides the remaining length:
//   N = 2^a, N != 256:  E = 16 (radix-16 passes);
//   N = 256:            E = 8 (radix-8 first pass, then radix-8 and radix-4);
//   N = 3 * 2^a:        E = 12 (radix-12 = 3 x 4 first pass, radix-4/2 passes).
// ---------------------------------------------------------------------------
// Why 8 at 256: a single 256^2 run's 256 rows then make 128 row jobs (64 at
// 16 per lane) and its columns 51 column jobs (27): 1.6x the steps/s of one
// 256^2 run and 1.55x its adjoint, for 8 % less aggregate rate in ensembles
// that already fill the device (4 per lane: 2.25x one run, but ensembles
// 18 % slower).
constexpr int wave_e(int N) { return N % 3 == 0 ? 12 : (N == 256 ? 8 : 16); }
// EL: elements per lane of a non-default layout (powers of two 4..16 with
// N/EL <= 64), chosen per size class by the ensemble autotuner.
template <int N, int EL = wave_e(N)>
struct WaveLay {
  static constexpr bool TRI = N % 3 == 0;
  static constexpr int E = EL;
  static constexpr int LPC = N / E;
  static constexpr int NV = 64 / LPC;
  static constexpr int VS = N + N / 16;
  static constexpr int RMAX = TRI ? 4 : (E < 16 ? E : 16);   // radix of the LDS passes
  static_assert(LPC * NV == 64 && LPC * E == N, "one wave holds NV whole vectors");
};
constexpr int wave_nv(int N, int el = 0) { return 64 * (el > 0 ? el : wave_e(N)) / N; }

template <int R, int SIGN>
__device__ __forceinline__ void dft_any(double2* v) {
  if constexpr (R == 16) dft16<SIGN>(v);
  else if constexpr (R == 12) dft12<SIGN>(v);
  else dft_small<R, SIGN>(v);
}

// Orders LDS traffic of one wavefront: prior LDS writes of the wave have
// completed (s_waitcnt lgkmcnt(0)) before its later LDS reads, and neither the
// IR optimiser (memory clobber) nor the machine scheduler (wave barrier) moves
// an access across this point.  Lanes of one wave execute LDS instructions in
// order, so no workgroup barrier is needed for wave-private LDS regions.
__device__ __forceinline__ void wave_sync() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xc07f);
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// LDS index with one pad slot per 16 elements (breaks the stride-16 pattern
// of the first Stockham write across the 64 lanes).
__device__ __forceinline__ int swz(int e) { return e + (e >> 4); }

// One Stockham pass P of the per-wave FFT (radix R = min(RMAX, N/P), Q = N/R
// butterflies per vector, BPL per lane) in its three steps: the lane's E
// inputs from `buf`, twiddles and radix-R butterflies in registers, outputs
// to `buf` (or into v after the last pass).  t holds butterfly q's inputs at
// t[q R + r].
template <int N, int P, int EL = wave_e(N)>
struct PassShape {
  using L = WaveLay<N, EL>;
  static constexpr int R = (N / P >= L::RMAX) ? L::RMAX : N / P;
  static constexpr int Q = N / R;
  static constexpr int BPL = Q / L::LPC;   // butterflies per lane
  static constexpr bool LAST = (P * R == N);
  static constexpr int STEP = N / (P * R);
  static_assert(BPL * R == L::E && N % (P * R) == 0, "pass shape");
};

template <int N, int P, int EL = wave_e(N)>
__device__ __forceinline__ void pass_load(double2 (&t)[16], const double2* buf, int lp) {
  using S = PassShape<N, P, EL>;
#pragma unroll
  for (int q = 0; q < S::BPL; ++q) {
    const int b = lp + WaveLay<N, EL>::LPC * q;
#pragma unroll
    for (int r = 0; r < S::R; ++r) t[q * S::R + r] = buf[swz(b + r * S::Q)];
  }
}

// A twiddle accessor with kPre16: it also holds the lane's factors of the
// radix-16 pass P = 16 (w16[r-1] = w^r, computed once per kernel as below).
template <class T, class = void>
struct tw_has_pre16 : std::false_type {};
template <class T>
struct tw_has_pre16<T, std::void_t<decltype(T::kPre16)>> : std::integral_constant<bool, T::kPre16> {};

// The powers w^r, r = 1 .. R-1, from w^1, w^2, w^4, w^8 (wp): the power of
// the lowest set bit times those of the others, in rising order.
template <int R, int NPOW>
__device__ __forceinline__ double2 tw_power(const double2 (&wp)[NPOW], int r) {
  const int lo = __builtin_ctz(r);
  double2 w = wp[lo];
#pragma unroll
  for (int e = lo + 1; e < NPOW; ++e)
    if (r & (1 << e)) w = cmul(w, wp[e]);
  return w;
}

template <int N, int P, int SIGN, int EL = wave_e(N), class TW>
__device__ __forceinline__ void pass_compute(double2 (&t)[16], int lp, const TW& tw) {
  using S = PassShape<N, P, EL>;
  constexpr int R = S::R, STEP = S::STEP;
#pragma unroll
  for (int q = 0; q < S::BPL; ++q) {
    const int b = lp + WaveLay<N, EL>::LPC * q;
    const int k = (int)((unsigned)b % (unsigned)P);   // b >= 0: an AND for P = 2^a
    double2* tq = t + q * R;
    // Twiddles w^r, w = W^(k STEP).  With a per-pass table (TwT): one LDS read
    // each.  Otherwise the powers w^1, w^2, w^4, w^8 come from the quarter
    // table and every other w^r is the product of the powers of its set bits
    // (at most three extra products, a few ulp): 4 table reads per butterfly
    // instead of R - 1, and only 4 twiddles live in registers.
    if constexpr (tw_has_pre16<TW>::value && P == 16 && R == 16 && S::BPL == 1) {
      // precomputed (the same products as below)
#pragma unroll
      for (int r = 1; r < R; ++r) {
        double2 w = tw.w16[r - 1];
        if (SIGN > 0) w.y = -w.y;
        tq[r] = cmul(tq[r], w);
      }
    } else if constexpr (TW::kTable) {
      // per-pass table: W^(r k STEP) for r = 1 .. R-1 at tw.pass<P>(k)
      const double2* tp = tw.template pass<P>(k);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        double2 w = tp[r - 1];
        if (SIGN > 0) w.y = -w.y;
        tq[r] = cmul(tq[r], w);
      }
    } else {
      constexpr int NPOW = R >= 16 ? 4 : (R >= 8 ? 3 : (R >= 4 ? 2 : 1));
      double2 wp[NPOW];
#pragma unroll
      for (int e = 0; e < NPOW; ++e) {
        double2 w = tw((1 << e) * k * STEP);
        if (SIGN > 0) w.y = -w.y;
        wp[e] = w;
      }
#pragma unroll
      for (int r = 1; r < R; ++r) tq[r] = cmul(tq[r], tw_power<R, NPOW>(wp, r));
    }
    dft_any<R, SIGN>(tq);
  }
}

template <int N, int P, int EL = wave_e(N)>
__device__ __forceinline__ void pass_store(const double2 (&t)[16], double2* buf, int lp) {
  using S = PassShape<N, P, EL>;
#pragma unroll
  for (int q = 0; q < S::BPL; ++q) {
    const int b = lp + WaveLay<N, EL>::LPC * q;
    const int k = (int)((unsigned)b % (unsigned)P);
#pragma unroll
    for (int r = 0; r < S::R; ++r) buf[swz((b - k) * S::R + k + r * P)] = t[q * S::R + r];
  }
}

template <int N, int P, int EL = wave_e(N)>
__device__ __forceinline__ void pass_final(const double2 (&t)[16], double2 (&v)[16]) {
  using S = PassShape<N, P, EL>;
#pragma unroll
  for (int q = 0; q < S::BPL; ++q)
#pragma unroll
    for (int r = 0; r < S::R; ++r) v[q + r * S::BPL] = t[q * S::R + r];
}

// ---------------------------------------------------------------------------
// Per-wavefront FFT of NV independent length-N vectors (WaveLay<N>).  Lane
// (lp, g) with lp < LPC holds elements lp + LPC*m, m < E, of vector g in v[];
// on return it holds the transform at the same positions.  Radix-E first pass
// in registers, then Stockham passes of radix min(RMAX, N/P) through `buf`
// (this lane's vector in its wave's LDS buffer, VS entries).  Only wave-level
// ordering is needed: no workgroup barrier.
// ---------------------------------------------------------------------------
template <int N, int P, int SIGN, int EL = wave_e(N), class TW>
__device__ __forceinline__ void wave_fft_pass(double2 (&v)[16], double2* buf, int lp, const TW& tw) {
  double2 t[16];
  pass_load<N, P, EL>(t, buf, lp);
  pass_compute<N, P, SIGN, EL>(t, lp, tw);
  if constexpr (PassShape<N, P, EL>::LAST) {
    pass_final<N, P, EL>(t, v);
  } else {
    wave_sync();
    pass_store<N, P, EL>(t, buf, lp);
    wave_sync();
    wave_fft_pass<N, P * PassShape<N, P, EL>::R, SIGN, EL>(v, buf, lp, tw);
  }
}

// Cross-lane swaps of gfx950 on a double2 (four dwords):
//   pl32: lanes 32-63 of a <-> lanes 0-31 of b   (v_permlane32_swap)
//   pl16: lanes 16-31 of a <-> lanes 0-15 of b, lanes 48-63 of a <-> 32-47 of b
//         (v_permlane16_swap)
__device__ __forceinline__ void swap_dw(double& a, double& b, bool rows32) {
  const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  unsigned alo = (unsigned)ua, ahi = (unsigned)(ua >> 32), blo = (unsigned)ub, bhi = (unsigned)(ub >> 32);
  if (rows32) {
    const auto lo = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
    alo = lo[0]; blo = lo[1]; ahi = hi[0]; bhi = hi[1];
  } else {
    const auto lo = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
    alo = lo[0]; blo = lo[1]; ahi = hi[0]; bhi = hi[1];
  }
  a = __longlong_as_double((long long)(((unsigned long long)ahi << 32) | alo));
  b = __longlong_as_double((long long)(((unsigned long long)bhi << 32) | blo));
}
__device__ __forceinline__ void swap_c(double2& a, double2& b, bool rows32) {
  swap_dw(a.x, b.x, rows32);
  swap_dw(a.y, b.y, rows32);
}

// The exchange between the two LDS passes of the 1024-point per-wave FFT
// (P = 16, radix 16, then P = 256, radix 4), in registers.  Lane (a, c) =
// 16 a + c holds in t[4 g + s] the pass-16 output of index 256 a + 64 g +
// 16 s + c; the P = 256 pass wants in lane (s, c), at t[4 g + a], the element
// of index (16 s + c) + 64 g + 256 a.  For each g that is the transpose of a
// 4 x 4 block between the lane row (a) and the register (s), done as a swap
// of lane-row bit 1 with register bit 1 (pl32 on the pairs (0, 2), (1, 3))
// and then of bit 0 with bit 0 (pl16 on the pairs (0, 1), (2, 3)).  It takes
// the place of one LDS store and load of the whole vector (16 KB per wave).
__device__ __forceinline__ void transpose_1024_p16_p256(double2 (&t)[16]) {
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    swap_c(t[4 * g + 0], t[4 * g + 2], true);
    swap_c(t[4 * g + 1], t[4 * g + 3], true);
    swap_c(t[4 * g + 0], t[4 * g + 1], false);
    swap_c(t[4 * g + 2], t[4 * g + 3], false);
  }
}

template <int N, int SIGN, int EL = wave_e(N), class TW>
__device__ __forceinline__ void wave_fft(double2 (&v)[16], double2* buf, int lp, const TW& tw) {
  constexpr int E = WaveLay<N, EL>::E;
  dft_any<E, SIGN>(v);                     // pass 1 (P = 1): butterfly lp
  if constexpr (N == 1024 && E == 16) {
    // one LDS exchange (after the register pass); the second in registers
#pragma unroll
    for (int r = 0; r < E; ++r) buf[swz(lp * E + r)] = v[r];
    wave_sync();
    pass_load<N, 16>(v, buf, lp);
    pass_compute<N, 16, SIGN>(v, lp, tw);
    transpose_1024_p16_p256(v);
    pass_compute<N, 256, SIGN>(v, lp, tw);
    double2 t[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) t[r] = v[r];
    pass_final<N, 256>(t, v);
    wave_sync();                           // buf may be rewritten by the caller
  } else if constexpr (N > E) {
#pragma unroll
    for (int r = 0; r < E; ++r) buf[swz(lp * E + r)] = v[r];
    wave_sync();
    wave_fft_pass<N, E, SIGN, EL>(v, buf, lp, tw);
    wave_sync();                           // buf may be rewritten by the caller
  }
}

}  // namespace qs
