// qs_fused.hip -- fused IFRK2 right-hand-side kernels for gfx950.
//
// Reference: InertialFriction/friction/Qsolver.cuh:63-322 (right-hand side),
// TimeIntegration/IFRK2.cuh:16-51 (stages, IFh := exp(L dt)),
// InertialFriction/friction/friction.cu:212-260 (call order).
//
// Two kernels per right-hand side, both built on per-wavefront FFTs (radix-E
// first pass in registers, E = wave_e(N) elements per lane, then LDS Stockham
// passes ordered by wave-level synchronisation only):
//
//  row job  (6 compute waves, one job = NV = 64 E / N rows, E elements per
//    lane (wave_e); 2 loader waves fetch the next job's rows):
//    7 mixed arrays (x-spectral, y-physical) -> five two-for-one x C2R
//    (waves 0-4) -> nine physical fields in LDS -> pointwise products in place
//    -> six two-for-one x R2C (waves 0-5) -> 7 mixed arrays.
//  column job (4 waves, one job = one field x 4 NV columns):
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
#include "qs_fused_kern.hpp"

namespace qs {

namespace {
__global__ __launch_bounds__(256) void k_ifac(double* ify, int Nxh, int Ny, double dkx, double dky, double Re,
                                              double Rf, double DQ, double aQ, double dt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < Ny) ifac_row(t, Ny, dky, Re, DQ, dt, ify);
  if (t < Nxh) ifac_col(t, Nxh, Ny, dkx, Re, Rf, DQ, aQ, dt, ify);
}
}  // namespace

int ifac_launch(double* ify, int Nxh, int Ny, double dkx, double dky, double Re, double Rf, double DQ, double aQ,
                double dt, hipStream_t st) {
  const int n = Ny > Nxh ? Ny : Nxh;
  hipLaunchKernelGGL(k_ifac, dim3((n + 255) / 256), dim3(256), 0, st, ify, Nxh, Ny, dkx, dky, Re, Rf, DQ, aQ, dt);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

int row_jobs(const RunDesc& a, int el) { return (a.Ny + wave_nv(a.Nx, el) - 1) / wave_nv(a.Nx, el); }
int col_jobs(const RunDesc& a, int el) {
  const int tc = 4 * wave_nv(a.Ny, el);
  return 3 * ((a.Nxh + tc - 1) / tc);
}

#define QS_NFIX_SWITCH(n, FN, ...)                 \
  switch (n) {                                      \
    case 48: return FN<48>(__VA_ARGS__);            \
    case 96: return FN<96>(__VA_ARGS__);            \
    case 192: return FN<192>(__VA_ARGS__);          \
    case 384: return FN<384>(__VA_ARGS__);          \
    case 768: return FN<768>(__VA_ARGS__);          \
    case 16: return FN<16>(__VA_ARGS__);            \
    case 32: return FN<32>(__VA_ARGS__);            \
    case 64: return FN<64>(__VA_ARGS__);            \
    case 128: return FN<128>(__VA_ARGS__);          \
    case 256: return FN<256>(__VA_ARGS__);          \
    case 512: return FN<512>(__VA_ARGS__);          \
    case 1024: return FN<1024>(__VA_ARGS__);        \
    default: return fail(QS_ERR_UNSUPPORTED_SIZE,   \
                         "fused kernels: side %d", n); \
  }

int launch_fused_row(int nfix, const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw1024,
                     hipStream_t st, int fin) {
  QS_NFIX_SWITCH(nfix, launch_row_n, runs, work, nwork, grid, tw1024, st, fin);
}

int launch_fused_col(int nfix, int mode, const RunDesc* runs, const int2* work, int nwork, int grid,
                     const double2* tw1024, hipStream_t st) {
  QS_NFIX_SWITCH(nfix, launch_col_n, mode, runs, work, nwork, grid, tw1024, st);
}

}  // namespace qs

// ---------------------------------------------------------------------------
// Test hook: per-wavefront FFT of NVEC = 1024/N (768/N for N = 3 * 2^a)
// contiguous vectors of length N, transformed CW = WaveLay<N>::NV vectors (one
// wave's worth) at a time.
// ---------------------------------------------------------------------------
namespace qs {
namespace {
template <int N, int SIGN>
__global__ __launch_bounds__(64) void k_wave_fft_test(const double2* in, double2* out, const double2* tw1024) {
  __shared__ double2 sm[kBuf];
  __shared__ double2 twq[kTwq];
  stage_twiddles(twq, tw1024, threadIdx.x, 64);
  __syncthreads();
  using L = WaveLay<N>;
  constexpr int CW = L::NV, LPC = L::LPC, VS = L::VS, E = L::E;
  constexpr int NVEC = (N % 3 == 0 ? 768 : 1024) / N;   // vectors per call, CW per pass
  const int lane = threadIdx.x, lp = lane / CW, g = lane % CW;
  for (int g0 = 0; g0 < NVEC; g0 += CW) {
    double2 v[16];
    for (int m = 0; m < E; ++m) v[m] = in[(g0 + g) * N + lp + LPC * m];
    wave_fft<N, SIGN>(v, sm + g * VS, lp, TwN<N>{twq});
    for (int m = 0; m < E; ++m) out[(g0 + g) * N + lp + LPC * m] = v[m];
  }
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
    case 48: return launch_wave_fft_test<48>(sign, in, out, tw1024, st);
    case 96: return launch_wave_fft_test<96>(sign, in, out, tw1024, st);
    case 192: return launch_wave_fft_test<192>(sign, in, out, tw1024, st);
    case 384: return launch_wave_fft_test<384>(sign, in, out, tw1024, st);
    case 768: return launch_wave_fft_test<768>(sign, in, out, tw1024, st);
    default: return fail(QS_ERR_UNSUPPORTED_SIZE, "wave fft %d", n);
  }
}
}  // namespace qs
