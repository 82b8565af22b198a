// qs_fft.hip -- per-op 2-D real transforms in the reference layouts.
//
// Replaces the cuFFT plans of Basic/Mesh.cuh:72-73 as used by FwdTrans /
// BwdTrans (FldOp/QFldfuncs.cuh:45-55) for Nx, Ny = 2^a (16..2048) and
// 3 * 2^a (48..1536, the reference driver's 768 = 512 * 3/2, friction.cu:109).
// Stockham passes of radix 4, then 2 and a last radix-3 pass where needed.
// A 2-D transform is two kernels:
//   R2C: row pass  (x-direction real FFT, two rows per complex FFT)  -> mixed
//        column pass (y-direction complex FFT on LDS tiles of columns) -> ft
//   C2R: column pass (inverse; also writes BwdTrans's column-0 symmetrisation
//        of the caller's spectrum) -> mixed; row pass (C2R, x 1/(Nx*Ny)) -> f
// The C2R row pass discards the imaginary parts of the DC and Nyquist bins,
// which fixes the semantics for non-Hermitian input (see qsolver.h).
#include "qs_fused_dev.hpp"

namespace qs {
namespace {

constexpr int kRadix = 4;

template <int N>
struct RowCfg {
  static constexpr int NV = (512 / N) > 1 ? 512 / N : 1;     // row pairs per block
  static constexpr int NT0 = NV * N / kRadix;
  static constexpr int NT = NT0 > 256 ? 256 : NT0;
};

template <int NY>
struct ColCfg {
  static constexpr int TC0 = 4096 / NY;
  static constexpr int TC = TC0 < 1 ? 1 : (TC0 > 32 ? 32 : TC0);   // columns per block
  static constexpr int LD = NY + 1;
  static constexpr int NT0 = TC * NY / kRadix;
  static constexpr int NT = NT0 > 256 ? 256 : NT0;
};

// x-direction R2C of row pairs: z = f[y0] + i f[y1]; A = (Z_k + conj Z_{N-k})/2,
// B = (Z_k - conj Z_{N-k})/(2i), k <= N/2.
template <int N>
__global__ __launch_bounds__(RowCfg<N>::NT) void k_row_r2c(const double* __restrict__ f,
                                                           double2* __restrict__ mix, int Ny,
                                                           const double2* __restrict__ tw) {
  using C = RowCfg<N>;
  constexpr int H = N / 2 + 1;
  __shared__ double2 s[C::NV * N];
  const int tid = threadIdx.x;
  const int pair0 = blockIdx.x * C::NV;
  for (int e = tid; e < C::NV * N; e += C::NT) {
    const int vec = e / N, x = e % N;
    const int y0 = 2 * (pair0 + vec), y1 = y0 + 1;
    const double a = y0 < Ny ? f[(size_t)y0 * N + x] : 0.0;
    const double b = y1 < Ny ? f[(size_t)y1 * N + x] : 0.0;
    s[vec * N + x] = make_double2(a, b);
  }
  __syncthreads();
  fft_lds<N, kRadix, -1, C::NV, N, C::NT>(s, TwFull{tw}, tid);
  for (int e = tid; e < C::NV * H; e += C::NT) {
    const int vec = e / H, k = e % H;
    const int y0 = 2 * (pair0 + vec), y1 = y0 + 1;
    const double2 zk = s[vec * N + k];
    const double2 zc = cconj(s[vec * N + (N - k) % N]);
    const double2 A = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
    const double2 d = csub(zk, zc);
    const double2 B = make_double2(0.5 * d.y, -0.5 * d.x);
    if (y0 < Ny) mix[(size_t)y0 * H + k] = A;
    if (y1 < Ny) mix[(size_t)y1 * H + k] = B;
  }
}

// x-direction C2R of row pairs with the Hermitian extension
// Z_k = Ah_k + i Bh_k (imaginary parts of bins 0 and N/2 discarded).
template <int N>
__global__ __launch_bounds__(RowCfg<N>::NT) void k_row_c2r(const double2* __restrict__ mix,
                                                           double* __restrict__ f, int Ny,
                                                           double scale,
                                                           const double2* __restrict__ tw) {
  using C = RowCfg<N>;
  constexpr int H = N / 2 + 1;
  __shared__ double2 s[C::NV * N];
  const int tid = threadIdx.x;
  const int pair0 = blockIdx.x * C::NV;
  for (int e = tid; e < C::NV * H; e += C::NT) {
    const int vec = e / H, k = e % H;
    const int y0 = 2 * (pair0 + vec), y1 = y0 + 1;
    const double2 a = y0 < Ny ? mix[(size_t)y0 * H + k] : make_double2(0.0, 0.0);
    const double2 b = y1 < Ny ? mix[(size_t)y1 * H + k] : make_double2(0.0, 0.0);
    double2* z = s + vec * N;
    if (k == 0 || k == N / 2) {
      z[k] = make_double2(a.x, b.x);
    } else {
      z[k] = make_double2(a.x - b.y, a.y + b.x);
      z[N - k] = make_double2(a.x + b.y, b.x - a.y);
    }
  }
  __syncthreads();
  fft_lds<N, kRadix, +1, C::NV, N, C::NT>(s, TwFull{tw}, tid);
  for (int e = tid; e < C::NV * N; e += C::NT) {
    const int vec = e / N, x = e % N;
    const int y0 = 2 * (pair0 + vec), y1 = y0 + 1;
    const double2 z = s[vec * N + x];
    if (y0 < Ny) f[(size_t)y0 * N + x] = z.x * scale;
    if (y1 < Ny) f[(size_t)y1 * N + x] = z.y * scale;
  }
}

// y-direction complex FFT of TC columns per block.  With SYM the block that
// owns column 0 also writes the reference's symmetry_func result
// (FldOp/QFldfuncs.cuh:18-44) into `sym` (the caller's spectrum), while the
// transform itself uses the unsymmetrised values, as BwdTrans does.
template <int NY, int SIGN, bool SYM>
__global__ __launch_bounds__(ColCfg<NY>::NT) void k_col_fft(const double2* in, double2* out, int Nxh,
                                                            const double2* __restrict__ tw,
                                                            double2* sym) {
  using C = ColCfg<NY>;
  __shared__ double2 s[C::TC * C::LD];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * C::TC;
  for (int e = tid; e < C::TC * NY; e += C::NT) {
    const int c = e % C::TC, j = e / C::TC, i = i0 + c;
    s[c * C::LD + j] = i < Nxh ? in[(size_t)j * Nxh + i] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  if (SYM && i0 == 0) {
    for (int j = 1 + tid; j < NY / 2; j += C::NT) {
      const double2 a = s[j], b = cconj(s[NY - j]);
      const double2 m = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
      sym[(size_t)j * Nxh] = m;
      sym[(size_t)(NY - j) * Nxh] = cconj(m);
    }
  }
  fft_lds<NY, kRadix, SIGN, C::TC, C::LD, C::NT>(s, TwFull{tw}, tid);
  for (int e = tid; e < C::TC * NY; e += C::NT) {
    const int c = e % C::TC, j = e / C::TC, i = i0 + c;
    if (i < Nxh) out[(size_t)j * Nxh + i] = s[c * C::LD + j];
  }
}

// y-direction FFT with the per-wavefront FFT of the fused kernels
// (WaveLay<NY>): a block of WPB wavefronts transforms TC = WPB * CW adjacent
// columns (CW = the vectors one wave holds), loaded and stored as whole row
// segments through LDS.  About Nxh / TC blocks, so the launch fills the device
// (the block-wide Stockham kernel above gives one block per 4096/NY columns),
// and the passes need only wave-level ordering.  SYM as in k_col_fft.
template <int NY>
struct WColCfg {
  static constexpr int WPB = NY >= 768 ? 2 : 1;            // wavefronts per block
  static constexpr int CW = WaveLay<NY>::NV, TC = WPB * CW, NT = 64 * WPB;
  static constexpr int WS = kBuf + 1;                       // odd wave stride
};

template <int NY, int SIGN, bool SYM>
__global__ __launch_bounds__(WColCfg<NY>::NT) void k_col_wfft(const double2* in, double2* out, int Nxh,
                                                              const double2* __restrict__ twsrc, double2* sym) {
  using C = WColCfg<NY>;
  using L = WaveLay<NY>;
  constexpr int CW = C::CW, TC = C::TC, NT = C::NT, WS = C::WS, LPC = L::LPC, VS = L::VS, E = L::E;
  __shared__ double2 sm[C::WPB * WS];
  __shared__ double2 twq[kTwq];
  const int tid = threadIdx.x;
  stage_twiddles(twq, twsrc, tid, NT);
  const int i0 = blockIdx.x * TC;
  auto slot = [&](int j, int c) { return (c / CW) * WS + (c % CW) * VS + swz(j); };
  // TC * NY = 64 * WPB * E elements: E per thread, all loads in flight first
  double2 ld[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = tid + NT * q, c = e % TC, j = e / TC, i = i0 + c;
    ld[q] = i < Nxh ? in[(size_t)j * Nxh + i] : make_double2(0.0, 0.0);
  }
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = tid + NT * q;
    sm[slot(e / TC, e % TC)] = ld[q];
  }
  __syncthreads();
  if (SYM && i0 == 0) {   // block-uniform: column 0 is tile column 0
    for (int j = 1 + tid; j < NY / 2; j += NT) {
      const double2 a = sm[slot(j, 0)], b = cconj(sm[slot(NY - j, 0)]);
      const double2 mm = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y + b.y));
      sym[(size_t)j * Nxh] = mm;
      sym[(size_t)(NY - j) * Nxh] = cconj(mm);
    }
    __syncthreads();   // the FFT below rewrites column 0's buffer
  }
  const int wave = tid >> 6, lane = tid & 63, lp = lane / CW, g = lane % CW;
  double2* vb = sm + wave * WS + g * VS;
  double2 v[16];
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = vb[swz(lp + LPC * m)];
  wave_sync();
  wave_fft<NY, SIGN>(v, vb, lp, TwN<NY>{twq});
#pragma unroll
  for (int m = 0; m < E; ++m) vb[swz(lp + LPC * m)] = v[m];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int e = tid + NT * q, c = e % TC, j = e / TC, i = i0 + c;
    if (i < Nxh) out[(size_t)j * Nxh + i] = sm[slot(j, c)];
  }
}

template <int N>
int launch_row_r2c(const Mesh* m, const double* f, double2* mix) {
  using C = RowCfg<N>;
  const int grid = (m->Ny / 2 + C::NV - 1) / C::NV;
  hipLaunchKernelGGL(k_row_r2c<N>, dim3(grid), dim3(C::NT), 0, stream_of(m), f, mix, m->Ny,
                     m->priv->twx);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

template <int N>
int launch_row_c2r(const Mesh* m, const double2* mix, double* f) {
  using C = RowCfg<N>;
  const int grid = (m->Ny / 2 + C::NV - 1) / C::NV;
  const double scale = 1.0 / ((double)m->Nx * (double)m->Ny);
  hipLaunchKernelGGL(k_row_c2r<N>, dim3(grid), dim3(C::NT), 0, stream_of(m), mix, f, m->Ny, scale,
                     m->priv->twx);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

template <int NY, int SIGN, bool SYM>
int launch_col(const Mesh* m, const double2* in, double2* out, double2* sym) {
  if constexpr ((NY % 3 == 0) ? NY <= 768 : NY <= 1024) {   // sizes the per-wave FFT takes
    using W = WColCfg<NY>;
    const int grid = (m->Nxh + W::TC - 1) / W::TC;
    const double2* tw = NY % 3 == 0 ? m->priv->tw768 : m->priv->tw1024;
    hipLaunchKernelGGL((k_col_wfft<NY, SIGN, SYM>), dim3(grid), dim3(W::NT), 0, stream_of(m), in, out, m->Nxh,
                       tw, sym);
    QS_CHECK_LAUNCH();
    return QS_OK;
  }
  using C = ColCfg<NY>;
  const int grid = (m->Nxh + C::TC - 1) / C::TC;
  hipLaunchKernelGGL((k_col_fft<NY, SIGN, SYM>), dim3(grid), dim3(C::NT), 0, stream_of(m), in, out,
                     m->Nxh, m->priv->twy, sym);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

// Powers of two 16..2048 and 3 * 2^a 48..1536 (radix-3 last pass).
#define QS_DISPATCH_N(n, FN, ...)                                   \
  switch (n) {                                                          \
    case 48: return FN<48>(__VA_ARGS__);                                \
    case 96: return FN<96>(__VA_ARGS__);                                \
    case 192: return FN<192>(__VA_ARGS__);                              \
    case 384: return FN<384>(__VA_ARGS__);                              \
    case 768: return FN<768>(__VA_ARGS__);                              \
    case 1536: return FN<1536>(__VA_ARGS__);                            \
    case 16: return FN<16>(__VA_ARGS__);                                \
    case 32: return FN<32>(__VA_ARGS__);                                \
    case 64: return FN<64>(__VA_ARGS__);                                \
    case 128: return FN<128>(__VA_ARGS__);                              \
    case 256: return FN<256>(__VA_ARGS__);                              \
    case 512: return FN<512>(__VA_ARGS__);                              \
    case 1024: return FN<1024>(__VA_ARGS__);                            \
    case 2048: return FN<2048>(__VA_ARGS__);                            \
    default: return fail(QS_ERR_UNSUPPORTED_SIZE, "unsupported size %d", n); \
  }

template <int NY> int col_fwd(const Mesh* m, const double2* in, double2* out) {
  return launch_col<NY, -1, false>(m, in, out, nullptr);
}
template <int NY> int col_inv(const Mesh* m, const double2* in, double2* out, double2* sym) {
  return sym ? launch_col<NY, +1, true>(m, in, out, sym) : launch_col<NY, +1, false>(m, in, out, nullptr);
}

int row_r2c(const Mesh* m, const double* f, double2* mix) { QS_DISPATCH_N(m->Nx, launch_row_r2c, m, f, mix); }
int row_c2r(const Mesh* m, const double2* mix, double* f) { QS_DISPATCH_N(m->Nx, launch_row_c2r, m, mix, f); }
int colf(const Mesh* m, const double2* in, double2* out) { QS_DISPATCH_N(m->Ny, col_fwd, m, in, out); }
int coli(const Mesh* m, const double2* in, double2* out, double2* sym) {
  QS_DISPATCH_N(m->Ny, col_inv, m, in, out, sym);
}

}  // namespace

int fft_r2c_2d(const Mesh* m, const double* f, double2* mix, double2* ft) {
  QS_TRY(row_r2c(m, f, mix));
  return colf(m, mix, ft);
}

int fft_c2r_2d(const Mesh* m, double2* ft, double2* mix, double* f, bool symmetrise_input) {
  QS_TRY(coli(m, ft, mix, symmetrise_input ? ft : nullptr));
  return row_c2r(m, mix, f);
}

}  // namespace qs
