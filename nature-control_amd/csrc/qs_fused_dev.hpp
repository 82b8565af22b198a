// qs_fused_dev.hpp -- device building blocks shared by the fused IFRK2 kernels
// (qs_fused.hip) and their discrete adjoint (qs_adjoint.hip).
#pragma once

#include <cmath>

#include "qs_fused.hpp"

namespace qs {
namespace {

constexpr int kBuf = 1088;           // LDS entries of one wave buffer (1024 + 1024/16)
constexpr int kTwq = 256;            // staged twiddles: the quarter table of W_1024 (or of W_768)

// threadIdx.x as a value the compiler treats as new on every call: lane
// offsets derived from it are recomputed per job instead of being hoisted out
// of a persistent job loop and held (or spilled) across it.
__device__ __forceinline__ int job_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}
// x as a value the compiler treats as new at every call: what is derived from
// it (wavenumbers, integrating factors) is recomputed where it is used instead
// of being kept live, or spilled, across the transforms of a column job.
__device__ __forceinline__ int opq(int x) {
  asm volatile("" : "+v"(x));
  return x;
}


// W_N^m from the staged quarter table: of W_1024 for N = 2^a <= 1024, of W_768
// for N = 3 * 2^a <= 768 (the kernel is handed the matching table).
template <int N>
struct TwN {
  static constexpr bool kTable = false;
  const double2* p;
  __device__ __forceinline__ double2 operator()(int m) const {
    if constexpr (N % 3 == 0) return TwQuarter<768>{p}(m * (768 / N));
    else return TwQuarter<1024>{p}(m * (1024 / N));
  }
};

// The accessor of N = 2^a >= 256 (16 elements per lane) with the lane's
// factors of the radix-16 pass P = 16 held in registers (kPre16): built once
// per kernel by the row kernels, whose waves run two such transforms per
// job, instead of four table reads and eleven products per transform.  Same
// values as computed in pass_compute (one butterfly per lane, b = lp,
// k = lp % 16, STEP = N / 256).
template <int N>
struct TwNP {
  static_assert(N >= 256 && (N & (N - 1)) == 0, "radix-16 pass with one butterfly per lane");
  static constexpr bool kTable = false;
  static constexpr bool kPre16 = true;
  const double2* p;
  double2 w16[15];
  __device__ __forceinline__ double2 operator()(int m) const { return TwN<N>{p}(m); }
  __device__ __forceinline__ TwNP(const double2* q, int lp) : p(q) {
    const int k = lp & 15;
    double2 wp[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) wp[e] = TwN<N>{q}((1 << e) * k * (N / 256));
#pragma unroll
    for (int r = 1; r < 16; ++r) w16[r - 1] = tw_power<16, 4>(wp, r);
  }
};

// Per-pass twiddle tables of the per-wave FFT of length N: for each LDS pass
// (P = E, E R_1, ...; radix R = min(RMAX, N/P)) the factors W_N^(r k STEP),
// STEP = N/(P R), at twtab_off(P) + k (R-1) + r - 1, for k < P, 1 <= r < R.
// Built once per block from the quarter table (build_twtab), they replace the
// per-butterfly quadrant selects and power products: one LDS read per factor.
template <int N>
constexpr int twtab_radix(int P) { return N / P >= WaveLay<N>::RMAX ? WaveLay<N>::RMAX : N / P; }
template <int N>
constexpr int twtab_off(int P) {
  int s = 0;
  for (int Q = WaveLay<N>::E; Q < P; Q *= twtab_radix<N>(Q)) s += Q * (twtab_radix<N>(Q) - 1);
  return s;
}
template <int N>
constexpr int twtab_size() { return twtab_off<N>(N); }
static_assert(twtab_size<768>() == 756, "768: 12*3 + 48*3 + 192*3");
template <int N>
struct TwT {
  static constexpr bool kTable = true;
  const double2* q;   // staged quarter table
  const double2* t;   // the pass tables
  __device__ __forceinline__ double2 operator()(int m) const { return TwN<N>{q}(m); }
  template <int P>
  __device__ __forceinline__ const double2* pass(int k) const {
    return t + twtab_off<N>(P) + k * (twtab_radix<N>(P) - 1);
  }
};

// Fills the pass tables from the staged quarter table q (which must be
// complete: a barrier between the two); the caller synchronises after.
template <int N>
__device__ __forceinline__ void build_twtab(double2* t, const double2* q, int tid, int nt) {
  const TwN<N> tw{q};
  for (int P = WaveLay<N>::E; P < N; P *= twtab_radix<N>(P)) {
    const int R = twtab_radix<N>(P), STEP = N / (P * R), base = twtab_off<N>(P);
    for (int e = tid; e < P * (R - 1); e += nt) {
      const int k = e / (R - 1), r = e % (R - 1) + 1;
      t[base + e] = tw(r * k * STEP);
    }
  }
}

// The fused kernels of size NFIX use the tables for N = 3 * 2^a, whose
// radix-4 pass chain is long (768: 4% faster per step); for N = 2^a the
// computed factors are as fast or faster (1024: the row kernel 2.5 us slower
// with tables).
template <int NFIX>
constexpr bool use_twt() { return NFIX % 3 == 0; }
// LDS entries such a kernel reserves for its tables (1 when it has none).
template <int NFIX>
constexpr int twt_len() {
  if constexpr (use_twt<NFIX>()) return twtab_size<NFIX>();
  else return 1;
}
// The twiddle accessor of a fixed-size fused kernel (tables built if used).
template <int NFIX>
__device__ __forceinline__ auto fused_tw(const double2* twq, const double2* twt) {
  if constexpr (use_twt<NFIX>()) return TwT<NFIX>{twq, twt};
  else return TwN<NFIX>{twq};
}

// ===========================================================================
// Row job
// ===========================================================================

// Z_k of a two-for-one C2R (Hermitian extensions of X and Y; imaginary parts
// of the DC and Nyquist bins discarded).
template <int N>
__device__ __forceinline__ double2 herm_pair(int k, double2 X, double2 Y) {
  if (k == 0 || 2 * k == N) return make_double2(X.x, Y.x);
  if (2 * k < N) return make_double2(X.x - Y.y, X.y + Y.x);
  return make_double2(X.x + Y.y, Y.x - X.y);
}

__device__ __forceinline__ void stage_twiddles(double2* twq, const double2* tw1024, int tid, int nt) {
  for (int m = tid; m < kTwq; m += nt) twq[m] = tw1024[m];
}

// 1/x for x > 0 (hardware reciprocal + two Newton steps, within an ulp or
// two of the quotient): far fewer live registers than an IEEE division.
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = r * (2.0 - x * r);
  return r * (2.0 - x * r);
}

// Integrating factors exp(L dt) and exp(L dt / 2) of the column kernels
// (friction/Qsolver.cuh:20-54 with the D_Q, a_Q coefficients), separably:
// L = c2 (kx^2 + ky^2) + c0 with c2 = -1/Re, c0 = -Rf/Re for w and c2 = -D_Q,
// c0 = a_Q for r1, r2, so exp(L h) = exp((c2 kx^2 + c0) h) exp(c2 ky^2 h), the
// product of a column factor and a row factor from the run's tables a.ify
// (ifac_row, ifac_col; built once per solver).  Within a few ulp of exp(L h).
// No exp is evaluated in the kernels: its constants, hoisted out of the job
// loop, had been spilled.
struct IFac {
  double fx, fxh;
  const double* fy;
  const double* fyh;
  __device__ __forceinline__ IFac(const RunDesc& a, int FLD, int i, int NY) {
    const double* cx = a.ify + 4 * NY;
    fx = gload(cx + (FLD == 0 ? 0 : a.Nxh) + i);
    fxh = gload(cx + (FLD == 0 ? 2 : 3) * a.Nxh + i);
    fy = a.ify + (FLD == 0 ? 0 : NY);
    fyh = a.ify + (FLD == 0 ? 2 * NY : 3 * NY);
  }
  __device__ __forceinline__ double row(int j) const { return gload(fy + j); }   // exp(c2 ky_j^2 dt)
  __device__ __forceinline__ double IF(int j) const { return fx * gload(fy + j); }
  __device__ __forceinline__ double IFh(int j) const { return fxh * gload(fyh + j); }
};
// Row staging: rows row0 .. row0+NV-1 of the 7 arrays arr[0..6] (x-spectral,
// H = N/2+1 bins each) into sm[t*RH + vec*H + k].  RH = NV*H <= 576 < 2*384:
// two slots per thread and array, all 14 loads in flight before the first LDS
// write.  Ends with a barrier.
template <int N, int EL = wave_e(N)>
__device__ __forceinline__ void stage_rows(double2* const* arr, const RunDesc& a, int row0, double2* sm) {
  constexpr int NV = WaveLay<N, EL>::NV, H = N / 2 + 1, RH = NV * H;
  static_assert(RH <= 2 * 384, "row staging assumes two slots per thread");
  const int tid = threadIdx.x;
  double2 ld[7][2];
  unsigned off[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int r = tid + 384 * h;
    const int rr = r < RH ? r : RH - 1;
    const int vec = rr / H, k = rr % H;
    const int y = row0 + vec;
    off[h] = (unsigned)((y < a.Ny ? y : a.Ny - 1) * a.pitch + k);
  }
#pragma unroll
  for (int t = 0; t < 7; ++t)
#pragma unroll
    for (int h = 0; h < 2; ++h) ld[t][h] = gload(arr[t] + off[h]);
  __builtin_amdgcn_sched_barrier(0);
  __syncthreads();   // the staging area overlaps wave buffers still in use before this call
#pragma unroll
  for (int t = 0; t < 7; ++t)
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (tid + 384 * h < RH) sm[t * RH + tid + 384 * h] = ld[t][h];
  __syncthreads();
}

// Row loader of the fixed-size row kernels: two extra waves (threads
// 384..511, loader thread lt = tid - 384) hold the next job's inverse-set rows
// in registers while the six compute waves run the current job, and write
// them to the staging area (the layout of stage_rows) once that job has
// released the wave buffers.  Only these waves wait on the loads; a compute
// wave's wait would also drain its own pending stores (loads and stores share
// one counter).  S = ceil(RH / 128) slots per array and thread.
template <int N, int EL = wave_e(N)>
struct RowLoader {
  static constexpr int NV = WaveLay<N, EL>::NV, H = N / 2 + 1, RH = NV * H, S = (RH + 127) / 128;
  double2 ld[7][S];
  double w00, alu;   // the run's W(0,0) and the step's uniform activity
  // rows row0 .. row0+NV-1 of a's inverse set (out-of-range slots read a
  // clamped, valid address and are not committed)
  __device__ __forceinline__ void issue(const RunDesc& a, int row0, int lt) {
    w00 = gload(a.w00);
    alu = a.alpha_p ? gload(a.alpha_p) : a.alpha;
    unsigned off[S];
#pragma unroll
    for (int h = 0; h < S; ++h) {
      const int r = lt + 128 * h;
      const int rr = r < RH ? r : RH - 1;
      const int vec = rr / H, k = rr % H;
      const int y = row0 + vec;
      off[h] = (unsigned)((y < a.Ny ? y : a.Ny - 1) * a.pitch + k);
    }
#pragma unroll
    for (int t = 0; t < 7; ++t)
#pragma unroll
      for (int h = 0; h < S; ++h) ld[t][h] = gload(a.minp[t] + off[h]);
  }
  // rows into the staging area, (w00, alu) into sc[0..1]
  __device__ __forceinline__ void commit(double2* sm, double* sc, int lt) const {
    commit_part(sm, sc, lt, 0, 7);
  }
  // arrays t0 .. t1-1 only, at stg + t RH ((w00, alu) with the part from t0 = 0)
  __device__ __forceinline__ void commit_part(double2* stg, double* sc, int lt, int t0, int t1) const {
    if (t0 == 0 && lt == 0) {
      sc[0] = w00;
      sc[1] = alu;
    }
#pragma unroll
    for (int t = 0; t < 7; ++t)
#pragma unroll
      for (int h = 0; h < S; ++h)
        if (t >= t0 && t < t1 && lt + 128 * h < RH) stg[t * RH + lt + 128 * h] = ld[t][h];
  }
};

// Physical fields of a row job from the staged inverse set (after stage_rows
// of the inverse-set arrays).  Waves 0-4 build two-for-one C2R inputs, pairing
// fields of like magnitude, and run the inverse x FFTs; the fields, scaled by
// 1/(Nx Ny), land in wave buffers 0-4 as pairs
//   0: (w, dy u + dx v)   1: (u, v)   2: (r1, r2)   3: (L1, L2)   4: (dx u, 0)
// (vector g of wave buffer q at sm[q*kBuf + g*VS + swz(x)]).  v: scratch
// registers of the caller; w00: the run's W(0,0) (*a.w00).  Ends with a
// barrier.
// NOWN: staged arrays t < NOWN at sm + 6 kBuf + t RH (an area of their own
// after the six wave buffers), the others at sm + t RH (in the buffers).
// EARLY: the others in wave buffer 5 instead (at sm + 5 kBuf + (t - NOWN) RH;
// buffer 5 is not written before the products), so no staged array overlaps
// buffers 0-4 and each wave transforms as soon as it has built its inputs:
// one barrier (after the transforms) instead of two.
template <int N, int EL = wave_e(N), bool OPQ = false, int NOWN = 0, bool EARLY = false, class TW>
__device__ __forceinline__ void inverse_fields(const RunDesc& a, int row0, double2* sm, const TW& tw,
                                               double2 (&v)[16], double w00) {
  using L = WaveLay<N, EL>;
  constexpr int NV = L::NV, LPC = L::LPC, VS = L::VS, E = L::E, H = N / 2 + 1, RH = NV * H;
  const int tid = OPQ ? job_tid() : (int)threadIdx.x;   // OPQ: see row_body
  // the wave index as a scalar: the branches below are wave-uniform, so each
  // wave runs only its own unrolled loop, with all its LDS reads in flight
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int lp = lane / NV, g = lane % NV;
  double2* vb = sm + wave * kBuf + g * VS;
  // (the own area's base as a per-job value: beyond the 64 KB reach of an LDS
  // instruction's offset field, its constant addresses would be kept, and
  // spilled, across the job loop)
  const double2* stg = sm + (NOWN > 0 ? opq(6 * kBuf) : 0);
  const double2* stg5 = sm + (EARLY ? opq(5 * kBuf) : 0);
  static_assert(!EARLY || (7 - NOWN) * RH <= kBuf, "the other arrays fit wave buffer 5");
  auto arr = [&](int t) -> const double2* {   // staged array t, vector g
    return (t < NOWN ? stg + t * RH : (EARLY ? stg5 + (t - NOWN) * RH : sm + t * RH)) + g * H;
  };
  auto build = [&](auto&& pair) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int k = lp + LPC * m;
      const int bin = k > N / 2 ? N - k : k;
      double2 X, Y;
      pair(bin, a.dkx * bin, X, Y);
      v[m] = herm_pair<N>(k, X, Y);
    }
  };
  if (wave == 0) {          // (w, dy u + dx v = -W + 2 kx^2 psi + W00 [kx = 0])
    build([&](int bin, double kx, double2& X, double2& Y) {
      const double2 W = arr(0)[bin], ps = arr(1)[bin];
      const double t2 = 2.0 * kx * kx;
      X = W;
      Y = make_double2(-W.x + t2 * ps.x + (bin == 0 ? w00 : 0.0), -W.y + t2 * ps.y);
    });
  } else if (wave == 1) {   // (u = i ky psi, v = -i kx psi)
    build([&](int bin, double kx, double2& X, double2& Y) {
      const double2 ps = arr(1)[bin];
      X = arr(2)[bin];
      Y = make_double2(kx * ps.y, -kx * ps.x);
    });
  } else if (wave == 2) {   // (r1, r2)
    build([&](int bin, double, double2& X, double2& Y) {
      X = arr(3)[bin];
      Y = arr(5)[bin];
    });
  } else if (wave == 3) {   // (L1, L2) = -kx^2 R + R_yy
    build([&](int bin, double kx, double2& X, double2& Y) {
      const double2 R1 = arr(3)[bin], R1yy = arr(4)[bin];
      const double2 R2 = arr(5)[bin], R2yy = arr(6)[bin];
      X = make_double2(-kx * kx * R1.x + R1yy.x, -kx * kx * R1.y + R1yy.y);
      Y = make_double2(-kx * kx * R2.x + R2yy.x, -kx * kx * R2.y + R2yy.y);
    });
  } else if (wave == 4) {   // (dx u = i kx u, -)
    build([&](int bin, double kx, double2& X, double2& Y) {
      const double2 U = arr(2)[bin];
      X = make_double2(-kx * U.y, kx * U.x);
      Y = make_double2(0.0, 0.0);
    });
  }
  if (!EARLY) __syncthreads();   // staging area free
  if (wave < 5) {
    const double sc = a.scale;
    wave_fft<N, +1, EL>(v, vb, lp, tw);
#pragma unroll
    for (int m = 0; m < E; ++m) vb[swz(lp + LPC * m)] = make_double2(v[m].x * sc, v[m].y * sc);
  }
  __syncthreads();
}

// Column tile of TC = 4 * CW columns starting at i0 (CW = WaveLay<NY>::NV):
// wave w owns columns i0 + w*CW + g (g < CW).  Global traffic is staged
// through the wave buffers (stride WS) so every load / store instruction
// covers whole row segments of the tile, while the FFTs run per wavefront on
// the lane layout "lane (lp, g) holds rows lp + LPC*m, m < E".  A tile holds
// TC * NY = 256 * E elements: E per thread.
template <int NY, int EL = wave_e(NY)>
struct ColTile {
  static constexpr int CW = WaveLay<NY, EL>::NV, TC = 4 * CW, LPC = WaveLay<NY, EL>::LPC, VS = WaveLay<NY, EL>::VS;
  static constexpr int E = WaveLay<NY, EL>::E;
  static constexpr int WS = kBuf + 1;   // odd wave stride: spreads LDS banks
  const RunDesc& a;
  double2* sm;
  int i0, tid, lp, g;
  double2* vb;
  __device__ __forceinline__ ColTile(const RunDesc& a_, double2* sm_, int i0_) : a(a_), sm(sm_), i0(i0_) {
    tid = job_tid();
    const int wave = tid >> 6, lane = tid & 63;
    lp = lane / CW;
    g = lane % CW;
    vb = sm + wave * WS + g * VS;
  }
  __device__ __forceinline__ int slot(int j, int c) const { return (c / CW) * WS + (c % CW) * VS + swz(j); }
  // cooperative coalesced tile load into the wave buffers (ends with a barrier);
  // all E loads are in flight before the first LDS write (the scheduling
  // barrier keeps the compiler from pairing each load with its write)
  // Thread t moves column c = t % TC of rows t / TC + q * (256 / TC): one
  // base offset plus a uniform stride per instruction, recomputed per call
  // (job_tid) so that no per-load address stays live between calls.
  __device__ __forceinline__ void load(const double2* __restrict__ src) {
    static_assert(256 % TC == 0, "tile width divides the block");
    __builtin_amdgcn_sched_barrier(0);   // no load hoisted above the previous phase
    const int t = job_tid();
    const int c = t % TC, j0 = t / TC;
    const int ii = i0 + c;
    const bool in = ii < a.Nxh;
    // 32-bit element offsets (SGPR base + VGPR offset form)
    const unsigned base = (unsigned)(j0 * a.pitch + (in ? ii : a.Nxh - 1));
    const unsigned stride = (unsigned)((256 / TC) * a.pitch);
    double2 ld[E];
#pragma unroll
    for (int q = 0; q < E; ++q) ld[q] = gload(src + (base + q * stride));
    __builtin_amdgcn_sched_barrier(0);
    __syncthreads();
    double2* s0 = sm + (c / CW) * WS + (c % CW) * VS;
#pragma unroll
    for (int q = 0; q < E; ++q) s0[swz(j0 + q * (256 / TC))] = in ? ld[q] : make_double2(0.0, 0.0);
    __syncthreads();
  }
  // cooperative coalesced tile store from the wave buffers.  Every thread reads
  // slots of every wave's buffer, so the closing barrier keeps a wave from
  // rewriting its buffer (its next FFT) while others still read it.
  __device__ __forceinline__ void store(double2* __restrict__ dst) {
    const int t = job_tid();
    const int c = t % TC, j0 = t / TC;
    const int ii = i0 + c;
    const unsigned base = (unsigned)(j0 * a.pitch + ii);
    const unsigned stride = (unsigned)((256 / TC) * a.pitch);
    const double2* s0 = sm + (c / CW) * WS + (c % CW) * VS;
    __syncthreads();
    if (ii < a.Nxh) {
#pragma unroll
      for (int q = 0; q < E; ++q) gstore(dst + (base + q * stride), s0[swz(j0 + q * (256 / TC))]);
    }
    __syncthreads();
  }
  // (the lane's LDS slots are recomputed per call from an opaque lane index, not
  // kept live across the job's transforms)
  __device__ __forceinline__ void stage_in(const double2* __restrict__ src, double2 (&v)[16]) {
    load(src);
    const int l = opq(lp);
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = vb[swz(l + LPC * m)];
    wave_sync();
  }
  __device__ __forceinline__ void stage_out(double2* __restrict__ dst, const double2 (&v)[16]) {
    const int l = opq(lp);
#pragma unroll
    for (int m = 0; m < E; ++m) vb[swz(l + LPC * m)] = v[m];
    store(dst);
  }
};

// Column job index -> (field, tile).  The column kernels remap block ids so
// that each XCD takes a contiguous range of about 3 * ntiles / 8 jobs; jobs
// are therefore numbered in groups of G = ntiles / 8 tiles, each group listing
// the vorticity jobs of its tiles (the heaviest: three transforms in and
// three out) before those of r1 and r2.  Every XCD then gets the same mix of
// heavy and light jobs, heaviest first (a field-major numbering put all the
// vorticity jobs on the first three XCDs).
__device__ __forceinline__ void col_job_map(int job, int ntiles, int& fld, int& tile) {
  const int G = ntiles >= 8 ? ntiles / 8 : 1;
  const int full = ntiles / G;
  const int grp = job / (3 * G);
  if (grp < full) {
    const int r = job - grp * 3 * G;
    fld = r / G;
    tile = grp * G + r % G;
  } else {   // the last, partial group: tiles full*G .. ntiles-1
    const int g2 = ntiles - full * G;
    const int r = job - full * 3 * G;
    fld = r / g2;
    tile = full * G + r % g2;
  }
}

// The final column stages (IFRK2 stage 2, RK4 stage 4, IF-Euler) end with the
// dealias mask (Qsolver.cuh:314-322), which zeroes every column i >= kcut =
// (Nxh - 1) * 2 / 3 + 1.  A tile whose first column is >= kcut therefore holds
// only zeros after the stage, state and next inverse set alike: its job
// stores zeros and runs no transform (col_zero_job).
__device__ __forceinline__ int col_kcut(const RunDesc& a) { return (a.Nxh - 1) * 2 / 3 + 1; }
// x-bins of the forward set that a final column stage reads: the columns of
// its tiles that start below kcut.  The tile width of every column layout of
// side Ny divides 4096 / Ny (powers of two) or 3072 / Ny (3 * 2^a), so kcut
// rounded up to that multiple covers them all.
__device__ __forceinline__ int col_kept_cols(const RunDesc& a) {
  const int T = (a.Ny % 3 == 0 ? 3072 : 4096) / a.Ny;
  const int kc = T > 0 ? (col_kcut(a) + T - 1) / T * T : a.Nxh;
  return kc < a.Nxh ? kc : a.Nxh;
}

// Column job index -> (field, tile) of a final stage, nkeep = tiles that start
// below kcut.  The eight XCD ranges of the numbering (xcd_remap) each get an
// even share of the transform-carrying tiles and of the zero tiles, each share
// contiguous in tile index: group x lists the vorticity, r1 and r2 jobs of its
// kept tiles, then those of its zero tiles.  (The stage-1 numbering would
// leave the last XCDs only zero tiles.)
__device__ __forceinline__ void col_job_map_final(int job, int ntiles, int nkeep, int& fld, int& tile) {
  const int nz = ntiles - nkeep;
  int x = 0;   // the last group starting at or before job (groups may be empty)
  for (int g = 1; g < 8; ++g)
    if (job >= 3 * ((g * nkeep) / 8 + (g * nz) / 8)) x = g;
  const int h0 = (x * nkeep) / 8, hc = ((x + 1) * nkeep) / 8 - h0;
  const int z0 = (x * nz) / 8, zc = ((x + 1) * nz) / 8 - z0;
  int r = job - 3 * (h0 + z0);
  if (r < 3 * hc) {
    fld = r / hc;
    tile = h0 + r % hc;
  } else {
    r -= 3 * hc;
    fld = r / zc;
    tile = nkeep + z0 + r % zc;
  }
}

// Column job placement over the XCDs of an MI355X (8 XCDs of 32 CUs, two
// column blocks per CU).  The kernels run one block per job; xcd_remap gives
// XCD x the contiguous job range [J_x, J_x + n_x), J_x = x per + min(x, rem)
// for n jobs, per = n / 8, rem = n % 8.  Of an XCD's blocks in dispatch order,
// the first 32 take one CU each and block k >= 32 joins the CU of block
// k - 32.  A vorticity job (six transforms) sharing a CU runs 20-30 % longer
// than alone, two r jobs (four each) sharing one end together about when a
// lone vorticity job does, and a final-stage zero job (stores only) slows its
// partner little.  So each XCD lists its r jobs first, then its vorticity
// jobs in the slots that end the first round (no partner), then the rest of
// its r jobs, the job of the partial tile and the zero jobs, which take the
// second round beside the r jobs.
// Applies when the full tiles (nfull = Nxh / tc) split evenly over the XCDs,
// with at most one partial tile (every fused size from 256 up); returns false
// otherwise.  nk: full tiles that start below the dealias cut in a final stage
// (nfull in stage 1); the partial tile, the last, is a zero tile when the
// caller's kept count ends before it.
constexpr int kCusPerXcd = 32;
__host__ __device__ __forceinline__ bool col_job_place(int job, int nxh, int tc, int nk, int& fld, int& tile) {
  const int nfull = nxh / tc, hasp = nxh % tc != 0;
  if (nfull < 8 || nfull % 8 != 0 || 3 * (nfull / 8) + 1 > 2 * kCusPerXcd) return false;
  const int g = nfull / 8, n = 3 * (nfull + hasp), per = n / 8, rem = n % 8;   // per = 3 g
  const int x = job < rem * (per + 1) ? job / (per + 1) : rem + (job - rem * (per + 1)) / per;
  const int k = job - (x * per + (x < rem ? x : rem));
  const int e = x < rem ? 1 : 0;                      // this XCD's job of the partial tile
  const int nx = 3 * g + e;
  const int K0 = (x * nk) / 8, hk = ((x + 1) * nk) / 8 - K0;   // kept full tiles K0 .. K0+hk-1
  const int Z0 = nk + x * g - K0;                     // zero full tiles Z0 .. Z0+g-hk-1
  // slot of the first vorticity job: the end of the first round, or the front
  // when every job has a CU of its own
  const int w0 = nx > kCusPerXcd ? kCusPerXcd - hk : 0;
  if (k >= w0 && k < w0 + hk) {
    fld = 0;
    tile = K0 + (k - w0);
    return true;
  }
  int o = k < w0 ? k : k - hk;                        // rank among the other jobs
  if (o < 2 * hk) {                                   // r1, then r2, of the kept tiles
    fld = 1 + o / hk;
    tile = K0 + o % hk;
    return true;
  }
  o -= 2 * hk;
  if (o < e) {                                        // the partial tile: field x
    fld = x;
    tile = nfull;
    return true;
  }
  o -= e;                                             // zero jobs: 3 per zero full tile
  const int zc = g - hk > 0 ? g - hk : 1;             // (>= 1 whenever a zero job exists)
  fld = o / zc;
  tile = Z0 + o % zc;
  return true;
}

}  // namespace
}  // namespace qs
