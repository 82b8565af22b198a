// qs_fused.hip -- fused IFRK2 right-hand-side kernels for gfx950.
//
// Reference: InertialFriction/friction/Qsolver.cuh:63-322 (right-hand side),
// TimeIntegration/IFRK2.cuh:16-51 (stages, IFh := exp(L dt)),
// InertialFriction/friction/friction.cu:212-260 (call order).
//
// Two kernels per right-hand side, both built on per-wavefront FFTs (radix-16
// first pass in registers, then LDS Stockham passes ordered by wave-level
// synchronisation only):
//
//  row job  (6 waves, one job = 1024/N rows):
//    7 mixed arrays (x-spectral, y-physical) -> five two-for-one x C2R
//    (waves 0-4) -> nine physical fields in LDS -> pointwise products in place
//    -> six two-for-one x R2C (waves 0-5) -> 7 mixed arrays.
//  column job (4 waves, one job = one field x 4*1024/Ny columns):
//    y FFT of the field's forward arrays -> N_X -> IFRK2 stage update of the
//    state -> the next stage's derived spectra -> y inverse FFT -> mixed arrays.
//    Stage 1 reads old and writes new (IF old + dt IF N/2); stage 2 reads new
//    and writes the dealiased next state over old.
//
// Mixed arrays (row pitch `pitch` >= Nxh, 8-aligned):
//  inverse set: M0 = W, M1 = psi = W/k^2, M2 = i ky psi,
//               M3 = R1, M4 = -ky^2 R1, M5 = R2, M6 = -ky^2 R2
//  forward set (s = 1/(Er Re)):
//    F0 = i kx A [kx != Nyquist] - kx^2 s P12,  A = R(-u w)
//    F1 = B - 2 i kx s P11,                     B = R(-v w)
//    F2 = s P21
//    F3 = R(loc1) + i kx R(-u r1),  F4 = R(-v r1)
//    F5 = R(loc2) + i kx R(-u r2),  F6 = R(-v r2)
//  N_w = F0 + i ky F1 + ky^2 F2 (real part at the four self-conjugate bins:
//  the projection the reference's BwdTrans/FwdTrans round trip applies to the
//  advection term, friction/Qsolver.cuh:179-193), N_r1 = F3 + i ky F4,
//  N_r2 = F5 + i ky F6.
#include "qs_fused.hpp"

namespace qs {
namespace {

constexpr int kBuf = 1088;           // LDS entries of one wave buffer (1024 + 1024/16)
constexpr int kTwq = 256;            // quarter twiddle table of W_1024

// W_N^m from the quarter table of W_1024 (N <= 1024, power of two).
template <int N>
struct TwN {
  const double2* p;
  __device__ __forceinline__ double2 operator()(int m) const {
    return TwQuarter<1024>{p}(m * (1024 / N));
  }
};

// ===========================================================================
// Row job
// ===========================================================================
template <int N>
__device__ __forceinline__ void row_job(const RunDesc& a, int row0, double2* sm, const double2* twq) {
  constexpr int NV = 1024 / N;          // rows of this job (= vectors per wave FFT)
  constexpr int LPC = N / 16;
  constexpr int VS = N + N / 16;        // padded vector stride inside a wave buffer
  const TwN<N> tw{twq};
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int lp = lane / NV, g = lane % NV;     // lane = lp * NV + g
  const int y = row0 + g;
  const bool yok = y < a.Ny;
  const size_t rowoff = (size_t)y * a.pitch;
  const double sc = a.scale;
  double2* wb = sm + wave * kBuf;              // this wave's buffer
  double2* vb = wb + g * VS;                   // this lane's vector

  // ---- x C2R of the nine fields (waves 0..4, one two-for-one FFT each) ----
  if (wave < 5) {
    const double w00 = *a.w00;
    double2 v[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int k = lp + LPC * m;
      const bool hi = k > N / 2;
      const int bin = hi ? N - k : k;
      const double kx = a.kx1[bin];
      double2 X = make_double2(0.0, 0.0), Y = make_double2(0.0, 0.0);
      if (yok) {
        const double2* M[7];
#pragma unroll
        for (int t = 0; t < 7; ++t) M[t] = a.mix[t] + rowoff + bin;
        switch (wave) {
          case 0: X = *M[0]; Y = *M[3]; break;                       // (w, r1)
          case 1: X = *M[5]; Y = *M[2]; break;                       // (r2, u)
          case 2: {                                                  // (v, dy u + dx v)
            const double2 ps = *M[1], W = *M[0];
            X = make_double2(kx * ps.y, -kx * ps.x);                 // v = -i kx psi
            const double t2 = 2.0 * kx * kx;
            Y = make_double2(-W.x + t2 * ps.x + (bin == 0 ? w00 : 0.0), -W.y + t2 * ps.y);
          } break;
          case 3: {                                                  // (L1, L2)
            const double kx2 = kx * kx;
            const double2 R1 = *M[3], R1yy = *M[4], R2 = *M[5], R2yy = *M[6];
            X = make_double2(-kx2 * R1.x + R1yy.x, -kx2 * R1.y + R1yy.y);
            Y = make_double2(-kx2 * R2.x + R2yy.x, -kx2 * R2.y + R2yy.y);
          } break;
          default: {                                                 // (dx u, -)
            const double2 U = *M[2];
            X = make_double2(-kx * U.y, kx * U.x);
          } break;
        }
      }
      if (bin == 0 || 2 * bin == N) v[m] = make_double2(X.x, Y.x);
      else if (!hi) v[m] = make_double2(X.x - Y.y, X.y + Y.x);
      else v[m] = make_double2(X.x + Y.y, Y.x - X.y);
    }
    wave_fft<N, +1>(v, vb, lp, tw);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int x = lp + LPC * m;
      vb[swz(x)] = make_double2(v[m].x * sc, v[m].y * sc);
    }
  }
  __syncthreads();

  // ---- pointwise products, in place (friction/Qsolver.cuh:63-164,206-300) -
  for (int p = tid; p < 1024; p += 384) {
    const int gg = p / N, x = p % N;
    const int o = gg * VS + swz(x);
    const int yy = row0 + gg;
    const double2 f0 = sm[0 * kBuf + o], f1 = sm[1 * kBuf + o], f2 = sm[2 * kBuf + o],
                  f3 = sm[3 * kBuf + o], f4 = sm[4 * kBuf + o];
    const double w = f0.x, r1 = f0.y, r2 = f1.x, u = f1.y, v = f2.x, uyvx = f2.y;
    const double L1 = f3.x, L2 = f3.y, ux = f4.x;
    const double al = a.Ra ? (yy < a.Ny ? a.Ra[(size_t)yy * N + x] : 0.0) : a.alpha;
    const double S = 2 * sqrt(r1 * r1 + r2 * r2);
    const double cS2 = a.cQ * S * S;
    const double gg1 = a.lambda * S * (cS2 - a.aQ);
    const double mlD = -1.0 * a.lambda * a.DQ;
    const double s1 = mlD * L1 * S + gg1 * r1;                          // single(r1)
    const double s2 = mlD * L2 * S + gg1 * r2;                          // single(r2)
    const double cr = 2.0 * a.DQ * L1 * r2 + (-2.0 * a.DQ) * r1 * L2;   // cross(r1, r2)
    const double p11 = s1 - al * r1;
    const double p12 = (s2 + cr) - al * r2;
    const double p21 = (s2 - cr) - al * r2;
    const double pA = a.advect_w ? -1.0 * u * w : 0.0;
    const double pB = a.advect_w ? -1.0 * v * w : 0.0;
    const double loc1 = (-1.0 * w * r2 - cS2 * r1) + a.lambda * S * ux;
    const double loc2 = (1.0 * w * r1 - cS2 * r2) + 0.5 * a.lambda * S * uyvx;
    sm[0 * kBuf + o] = make_double2(pA, p12);
    sm[1 * kBuf + o] = make_double2(pB, p11);
    sm[2 * kBuf + o] = make_double2(p21, 0.0);
    sm[3 * kBuf + o] = make_double2(loc1, -1.0 * u * r1);
    sm[4 * kBuf + o] = make_double2(loc2, -1.0 * u * r2);
    sm[5 * kBuf + o] = make_double2(-1.0 * v * r1, -1.0 * v * r2);
  }
  __syncthreads();

  // ---- x R2C of the products (waves 0..5) and the spectral combinations ---
  {
    double2 v[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) v[m] = vb[swz(lp + LPC * m)];
    wave_fft<N, -1>(v, vb, lp, tw);
#pragma unroll
    for (int m = 0; m < 16; ++m) vb[swz(lp + LPC * m)] = v[m];
    wave_sync();
    if (yok) {
      const double sE = a.s_ErRe;
#pragma unroll
      for (int m = 0; m < 9; ++m) {
        const int k = lp + LPC * m;
        if (k <= N / 2) {
          const double2 zk = vb[swz(k)];
          const double2 zc = cconj(vb[swz((N - k) & (N - 1))]);
          const double2 A = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
          const double2 d = csub(zk, zc);
          const double2 B = make_double2(0.5 * d.y, -0.5 * d.x);
          const double kx = a.kx1[k];
          const size_t o = rowoff + k;
          switch (wave) {
            case 0: {   // (A, P12) -> F0
              const double kxa = (2 * k == N) ? 0.0 : kx;
              const double c = -kx * kx * sE;
              a.mix[0][o] = make_double2(-kxa * A.y + c * B.x, kxa * A.x + c * B.y);
            } break;
            case 1: {   // (B, P11) -> F1
              const double t = 2.0 * kx * sE;
              a.mix[1][o] = make_double2(A.x + t * B.y, A.y - t * B.x);
            } break;
            case 2:     // (P21, -) -> F2
              a.mix[2][o] = make_double2(sE * A.x, sE * A.y);
              break;
            case 3:     // (loc1, -u r1) -> F3
              a.mix[3][o] = make_double2(A.x - kx * B.y, A.y + kx * B.x);
              break;
            case 4:     // (loc2, -u r2) -> F5
              a.mix[5][o] = make_double2(A.x - kx * B.y, A.y + kx * B.x);
              break;
            default:    // (-v r1, -v r2) -> F4, F6
              a.mix[4][o] = A;
              a.mix[6][o] = B;
              break;
          }
        }
      }
    }
  }
}

// ===========================================================================
// Column job
// ===========================================================================
template <int NY, int MODE>
__device__ __forceinline__ void col_job(const RunDesc& a, int fld, int i0, double2* sm, const double2* twq) {
  constexpr int CW = 1024 / NY;          // columns per wave
  constexpr int LPC = NY / 16;
  constexpr int VS = NY + NY / 16;
  const TwN<NY> tw{twq};
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int lp = lane / CW, g = lane % CW;     // lane = lp * CW + g
  const int i = i0 + wave * CW + g;
  const bool iok = i < a.Nxh;
  const int ic = iok ? i : 0;
  double2* vb = sm + wave * kBuf + g * VS;
  const double kx = a.kx1[ic];
  const int nyq_x = a.Nx / 2;
  bool bad = false;

  double2 st[16];   // the state of the next stage (curr / new / old)
  if (MODE == MODE_PRO) {
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int j = lp + LPC * m;
      st[m] = iok ? a.old_[fld][(size_t)j * a.pitch + i] : make_double2(0.0, 0.0);
    }
  } else {
    // ---- N_X from the forward set ----------------------------------------
    double2 nl[16];
    const int f0 = fld == 0 ? 0 : (fld == 1 ? 3 : 5);
    const int nterm = fld == 0 ? 3 : 2;
#pragma unroll 1
    for (int t = 0; t < nterm; ++t) {
      double2 v[16];
      const double2* src = a.mix[f0 + t];
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const int j = lp + LPC * m;
        v[m] = iok ? src[(size_t)j * a.pitch + i] : make_double2(0.0, 0.0);
      }
      wave_fft<NY, -1>(v, vb, lp, tw);
#pragma unroll
      for (int m = 0; m < 16; ++m) {
        const double ky = a.ky1[lp + LPC * m];
        const double2 z = v[m];
        if (t == 0) nl[m] = z;
        else if (t == 1) nl[m] = make_double2(nl[m].x - ky * z.y, nl[m].y + ky * z.x);
        else nl[m] = make_double2(nl[m].x + ky * ky * z.x, nl[m].y + ky * ky * z.y);
      }
    }
    // ---- IFRK2 stage update -------------------------------------------------
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int j = lp + LPC * m;
      if (!iok) { st[m] = make_double2(0.0, 0.0); continue; }
      double2 n = nl[m];
      if (fld == 0 && (i == 0 || i == nyq_x) && (j == 0 || 2 * j == NY)) n.y = 0.0;
      const size_t o = (size_t)j * a.pitch + i;
      const double ky = a.ky1[j];
      const double k2 = kx * kx + ky * ky;
      const double Lk = fld == 0 ? 1.0 / a.Re * (-1.0 * k2 - a.Rf) : (-1.0 * a.DQ * k2 + a.aQ);
      const double IF = exp(Lk * a.dt);
      const double2 an = make_double2(n.x * a.dt, n.y * a.dt);
      if (MODE == MODE_S1) {
        const double2 old = a.old_[fld][o];
        const double2 nw0 = make_double2(old.x * IF, old.y * IF);              // IFRK2.cuh:16-26
        const double h = 1.0 / 2.0 * IF;
        a.new_[fld][o] = make_double2(nw0.x + h * an.x, nw0.y + h * an.y);     // IFRK2.cuh:28-39
        st[m] = make_double2((old.x + an.x) * IF, (old.y + an.y) * IF);      // curr (IFh = IF)
      } else {
        const double2 nw = a.new_[fld][o];
        double2 r = make_double2(nw.x + 1.0 / 2.0 * an.x, nw.y + 1.0 / 2.0 * an.y);  // IFRK2.cuh:41-51
        const bool keep = (i < ((a.Nxh - 1) * 2 / 3 + 1)) && ((j < NY / 3 + 1) || ((NY - j) < NY / 3));
        const double cut = keep ? 1.0 : 0.0;                                  // Qsolver.cuh:314-322
        r = make_double2(r.x * cut, r.y * cut);
        bad |= !isfinite(r.x) || !isfinite(r.y);
        a.old_[fld][o] = r;                                                   // next state
        st[m] = r;
      }
    }
  }

  // ---- inverse set of the next stage from st --------------------------------
  const int nout = fld == 0 ? 3 : 2;
  const int m0 = fld == 0 ? 0 : (fld == 1 ? 3 : 5);
  if (fld == 0 && i == 0 && lp == 0 && a.w00) *a.w00 = st[0].x;   // element (j=0, i=0)
#pragma unroll 1
  for (int t = 0; t < nout; ++t) {
    double2 v[16];
#pragma unroll
    for (int m = 0; m < 16; ++m) {
      const int j = lp + LPC * m;
      const double ky = a.ky1[j];
      const double2 s = st[m];
      if (fld == 0) {   // W, psi = W/k^2, i ky psi  (vel_func, QFldfuncs.cuh:204-224)
        if (t == 0) {
          v[m] = s;
        } else {
          const double k2 = kx * kx + ky * ky;
          const double2 ps = (i == 0 && j == 0) ? make_double2(0.0, 0.0) : make_double2(s.x / k2, s.y / k2);
          v[m] = t == 1 ? ps : make_double2(-ky * ps.y, ky * ps.x);
        }
      } else {          // R, -ky^2 R
        v[m] = t == 0 ? s : make_double2(-ky * ky * s.x, -ky * ky * s.y);
      }
    }
    wave_fft<NY, +1>(v, vb, lp, tw);
    if (iok) {
      double2* dst = a.mix[m0 + t];
#pragma unroll
      for (int m = 0; m < 16; ++m) dst[(size_t)(lp + LPC * m) * a.pitch + i] = v[m];
    }
  }
  if (MODE == MODE_S2 && bad) atomicOr(a.flag, 1);
}

// ===========================================================================
// Kernels: one run (grid over jobs) and an ensemble (work list over runs).
// ===========================================================================
__device__ __forceinline__ void stage_twiddles(double2* twq, const double2* tw1024, int tid, int nt) {
  for (int m = tid; m < kTwq; m += nt) twq[m] = tw1024[m];
}

template <int N>
__device__ __forceinline__ void row_dispatch_n(const RunDesc& a, int job, double2* sm, const double2* twq) {
  row_job<N>(a, job * (1024 / N), sm, twq);
}

__device__ void row_dispatch(const RunDesc& a, int job, double2* sm, const double2* twq) {
  switch (a.Nx) {
    case 16: row_dispatch_n<16>(a, job, sm, twq); break;
    case 32: row_dispatch_n<32>(a, job, sm, twq); break;
    case 64: row_dispatch_n<64>(a, job, sm, twq); break;
    case 128: row_dispatch_n<128>(a, job, sm, twq); break;
    case 256: row_dispatch_n<256>(a, job, sm, twq); break;
    case 512: row_dispatch_n<512>(a, job, sm, twq); break;
    default: row_dispatch_n<1024>(a, job, sm, twq); break;
  }
}

template <int MODE>
__device__ void col_dispatch(const RunDesc& a, int job, double2* sm, const double2* twq) {
  const int tc = 4 * (1024 / a.Ny);
  const int ntiles = (a.Nxh + tc - 1) / tc;
  const int fld = job / ntiles, i0 = (job % ntiles) * tc;
  switch (a.Ny) {
    case 16: col_job<16, MODE>(a, fld, i0, sm, twq); break;
    case 32: col_job<32, MODE>(a, fld, i0, sm, twq); break;
    case 64: col_job<64, MODE>(a, fld, i0, sm, twq); break;
    case 128: col_job<128, MODE>(a, fld, i0, sm, twq); break;
    case 256: col_job<256, MODE>(a, fld, i0, sm, twq); break;
    case 512: col_job<512, MODE>(a, fld, i0, sm, twq); break;
    default: col_job<1024, MODE>(a, fld, i0, sm, twq); break;
  }
}

// NFIX > 0: every run of the work list has Nx = Ny = NFIX (one instantiation);
// NFIX = 0: per-run dispatch on the mesh size (ensembles of uneven runs).
template <int NFIX>
__global__ __launch_bounds__(384) void k_row(const RunDesc* __restrict__ runs, const int2* __restrict__ work,
                                             int nwork, const double2* __restrict__ tw1024) {
  __shared__ double2 sm[6 * kBuf];
  __shared__ double2 twq[kTwq];
  stage_twiddles(twq, tw1024, threadIdx.x, 384);
  __syncthreads();
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  for (int w = L; w < nwork; w += gridDim.x) {
    const int2 it = work[w];          // (run, job)
    if constexpr (NFIX > 0) row_dispatch_n<NFIX>(runs[it.x], it.y, sm, twq);
    else row_dispatch(runs[it.x], it.y, sm, twq);
    __syncthreads();
  }
}

template <int NY, int MODE>
__device__ __forceinline__ void col_dispatch_n(const RunDesc& a, int job, double2* sm, const double2* twq) {
  constexpr int tc = 4 * (1024 / NY);
  const int ntiles = (a.Nxh + tc - 1) / tc;
  col_job<NY, MODE>(a, job / ntiles, (job % ntiles) * tc, sm, twq);
}

template <int NFIX, int MODE>
__global__ __launch_bounds__(256) void k_col(const RunDesc* __restrict__ runs, const int2* __restrict__ work,
                                             int nwork, const double2* __restrict__ tw1024) {
  __shared__ double2 sm[4 * kBuf];
  __shared__ double2 twq[kTwq];
  stage_twiddles(twq, tw1024, threadIdx.x, 256);
  __syncthreads();
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  for (int w = L; w < nwork; w += gridDim.x) {
    const int2 it = work[w];
    if constexpr (NFIX > 0) col_dispatch_n<NFIX, MODE>(runs[it.x], it.y, sm, twq);
    else col_dispatch<MODE>(runs[it.x], it.y, sm, twq);
    __syncthreads();
  }
}

template <int NFIX>
int launch_row_n(const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw, hipStream_t st) {
  hipLaunchKernelGGL(k_row<NFIX>, dim3(grid), dim3(384), 0, st, runs, work, nwork, tw);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

template <int NFIX>
int launch_col_n(int mode, const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw,
                 hipStream_t st) {
  if (mode == MODE_PRO) hipLaunchKernelGGL((k_col<NFIX, MODE_PRO>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw);
  else if (mode == MODE_S1) hipLaunchKernelGGL((k_col<NFIX, MODE_S1>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw);
  else hipLaunchKernelGGL((k_col<NFIX, MODE_S2>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

}  // namespace

int row_jobs(const RunDesc& a) { return (a.Ny + 1024 / a.Nx - 1) / (1024 / a.Nx); }
int col_jobs(const RunDesc& a) {
  const int tc = 4 * (1024 / a.Ny);
  return 3 * ((a.Nxh + tc - 1) / tc);
}

#define QS_NFIX_SWITCH(n, FN, ...)                 \
  switch (n) {                                      \
    case 16: return FN<16>(__VA_ARGS__);            \
    case 32: return FN<32>(__VA_ARGS__);            \
    case 64: return FN<64>(__VA_ARGS__);            \
    case 128: return FN<128>(__VA_ARGS__);          \
    case 256: return FN<256>(__VA_ARGS__);          \
    case 512: return FN<512>(__VA_ARGS__);          \
    case 1024: return FN<1024>(__VA_ARGS__);        \
    default: return FN<0>(__VA_ARGS__);             \
  }

int launch_fused_row(int nfix, const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw1024,
                     hipStream_t st) {
  QS_NFIX_SWITCH(nfix, launch_row_n, runs, work, nwork, grid, tw1024, st);
}

int launch_fused_col(int nfix, int mode, const RunDesc* runs, const int2* work, int nwork, int grid,
                     const double2* tw1024, hipStream_t st) {
  QS_NFIX_SWITCH(nfix, launch_col_n, mode, runs, work, nwork, grid, tw1024, st);
}

}  // namespace qs

// ---------------------------------------------------------------------------
// Test hook: per-wavefront FFT of CW = 1024/N vectors (one wave, one block).
// in/out: CW vectors of length N, contiguous.
// ---------------------------------------------------------------------------
namespace qs {
namespace {
template <int N, int SIGN>
__global__ __launch_bounds__(64) void k_wave_fft_test(const double2* in, double2* out, const double2* tw1024) {
  __shared__ double2 sm[kBuf];
  __shared__ double2 twq[kTwq];
  stage_twiddles(twq, tw1024, threadIdx.x, 64);
  __syncthreads();
  constexpr int CW = 1024 / N, LPC = N / 16, VS = N + N / 16;
  const int lane = threadIdx.x, lp = lane / CW, g = lane % CW;
  double2 v[16];
  for (int m = 0; m < 16; ++m) v[m] = in[g * N + lp + LPC * m];
  wave_fft<N, SIGN>(v, sm + g * VS, lp, TwN<N>{twq});
  for (int m = 0; m < 16; ++m) out[g * N + lp + LPC * m] = v[m];
}
template <int N>
int launch_wave_fft_test(int sign, const double2* in, double2* out, const double2* tw, hipStream_t st) {
  if (sign < 0) hipLaunchKernelGGL((k_wave_fft_test<N, -1>), dim3(1), dim3(64), 0, st, in, out, tw);
  else hipLaunchKernelGGL((k_wave_fft_test<N, +1>), dim3(1), dim3(64), 0, st, in, out, tw);
  QS_CHECK_LAUNCH();
  return QS_OK;
}
}  // namespace

int wave_fft_test(int n, int sign, const double2* in, double2* out, const double2* tw1024, hipStream_t st) {
  switch (n) {
    case 16: return launch_wave_fft_test<16>(sign, in, out, tw1024, st);
    case 32: return launch_wave_fft_test<32>(sign, in, out, tw1024, st);
    case 64: return launch_wave_fft_test<64>(sign, in, out, tw1024, st);
    case 128: return launch_wave_fft_test<128>(sign, in, out, tw1024, st);
    case 256: return launch_wave_fft_test<256>(sign, in, out, tw1024, st);
    case 512: return launch_wave_fft_test<512>(sign, in, out, tw1024, st);
    case 1024: return launch_wave_fft_test<1024>(sign, in, out, tw1024, st);
    default: return fail(QS_ERR_UNSUPPORTED_SIZE, "wave fft %d", n);
  }
}
}  // namespace qs
