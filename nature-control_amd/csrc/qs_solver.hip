// qs_solver.hip -- the IFRK2 step of the Qsolver, fused for gfx950.
//
// Reference: InertialFriction/friction/friction.cu:212-260 (call order),
// friction/Qsolver.cuh:166-322 (right-hand side), TimeIntegration/IFRK2.cuh
// (stages, with IFh := exp(L dt)).
//
// Fused step (per right-hand side, two per step):
//   row kernel   k_row_rhs : 7 mixed arrays (x-spectral, y-physical) -> x C2R
//                            of 9 fields -> pointwise products -> x R2C of 11
//                            products -> 7 mixed arrays (in place).
//   column kernel k_col_stage: y FFT of the 7 arrays -> N_w, N_r1, N_r2 ->
//                            IFRK2 stage update of the state -> the next
//                            stage's fields -> y inverse FFT -> 7 mixed arrays.
// Mixed arrays M[a][y][i] use the spectral layout [Ny][Nxh].
//
//  Inverse set (column kernel -> row kernel):
//    M0 = W       M1 = psi = W/k^2      M2 = i ky psi
//    M3 = R1      M4 = -ky^2 R1         M5 = R2      M6 = -ky^2 R2
//  Forward set (row kernel -> column kernel), s = 1/(Er Re):
//    F0 = i kx A [kx != Nyquist] - kx^2 s P12     A = R(-u w)
//    F1 = B - 2 i kx s P11                        B = R(-v w)
//    F2 = s P21
//    F3 = R(loc1) + i kx R(-u r1)   F4 = R(-v r1)
//    F5 = R(loc2) + i kx R(-u r2)   F6 = R(-v r2)
//  N_w = F0 + i ky F1 + ky^2 F2 (real part at the four self-conjugate bins),
//  N_r1 = F3 + i ky F4,  N_r2 = F5 + i ky F6.
// The projections that the reference's BwdTrans/FwdTrans round trips apply to
// the advection term of N_w (friction/Qsolver.cuh:179-193) are the dropped
// Nyquist-column x-derivative and the real parts at self-conjugate bins.
#include <vector>

#include "qs_internal.hpp"

struct qs_solver {
  Mesh* mesh = nullptr;
  qs_params p{};
  qs::Coef coef;
  // state (reference layout [Ny][Nxh])
  double2* old_[3] = {nullptr, nullptr, nullptr};   // w, r1, r2
  double2* new_[3] = {nullptr, nullptr, nullptr};
  double2* mix[7] = {};                              // mixed arrays
  double* Ra = nullptr;                              // activity field or null
  double* w00 = nullptr;                             // W(0,0) of the stage
  int* flag = nullptr;                               // non-finite flag
  bool mixed_valid = false;
  // unfused path (reference call granularity)
  Field* F[21] = {};
  double* IFt[4] = {};   // IFw, IFwh, IFq, IFqh
  bool unfused_ready = false;
  // per-kernel HIP-event timing (qs_solver_profile)
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
};

namespace qs {
namespace {

constexpr int kRadixS = 4;

// --------------------------------------------------------------------------
// Row kernel
// --------------------------------------------------------------------------
template <int N>
struct RowRhsCfg {
  static constexpr int NV = (1024 / N) > 1 ? 1024 / N : 1;   // rows per block
  static constexpr int NT = 256;
  static constexpr int H = N / 2 + 1;
  static constexpr int PTS = NV * N / NT;                    // physical points per thread
  static constexpr int BINS = (NV * H + NT - 1) / NT;        // spectral bins per thread
};

struct RowArgs {
  double2* mix[7];
  const double* Ra;     // nullptr: uniform alpha
  const double* w00;
  const double* kx1;
  const double2* tw;
  int Ny;
  double alpha, lambda, DQ, aQ, cQ, s_ErRe, scale;
  int advect_w;
};

// Builds Z = Xh + i Yh (Hermitian extensions, DC/Nyquist imaginary parts
// discarded) from two half spectra given by functor g(vec, k) -> (X, Y).
template <int N, class G>
__device__ __forceinline__ void build_c2r_pair(double2* s, int tid, G g) {
  using C = RowRhsCfg<N>;
#pragma unroll
  for (int q = 0; q < C::BINS; ++q) {
    const int e = tid + q * C::NT;
    if (e < C::NV * C::H) {
      const int vec = e / C::H, k = e % C::H;
      double2 a, b;
      g(vec, k, a, b);
      double2* z = s + vec * N;
      if (k == 0 || k == N / 2) {
        z[k] = make_double2(a.x, b.x);
      } else {
        z[k] = make_double2(a.x - b.y, a.y + b.x);
        z[N - k] = make_double2(a.x + b.y, b.x - a.y);
      }
    }
  }
}

// Unpacks a two-for-one R2C: A_k = (Z_k + conj Z_{N-k})/2, B_k = (Z_k - conj Z_{N-k})/(2i).
template <int N>
__device__ __forceinline__ void unpack_r2c(const double2* s, int vec, int k, double2& A, double2& B) {
  const double2 zk = s[vec * N + k];
  const double2 zc = cconj(s[vec * N + ((N - k) & (N - 1))]);
  A = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
  const double2 d = csub(zk, zc);
  B = make_double2(0.5 * d.y, -0.5 * d.x);
}

template <int N>
__global__ __launch_bounds__(256) void k_row_rhs(RowArgs a) {
  using C = RowRhsCfg<N>;
  __shared__ double2 s[C::NV * N];
  const int tid = threadIdx.x;
  const int row0 = blockIdx.x * C::NV;
  const double w00 = *a.w00;
  const double sc = a.scale;

  auto rowof = [&](int vec) { return row0 + vec; };
  auto ld = [&](int arr, int vec, int k) -> double2 {
    const int y = rowof(vec);
    return y < a.Ny ? a.mix[arr][(size_t)y * C::H + k] : make_double2(0.0, 0.0);
  };

  // ---- x C2R of the nine fields (five complex FFTs) -----------------------
  double w[C::PTS], r1[C::PTS], r2[C::PTS], u[C::PTS], v[C::PTS], uyvx[C::PTS], L1[C::PTS],
      L2[C::PTS], ux[C::PTS];
  auto fft_out = [&](double* X, double* Y) {
    __syncthreads();
    fft_lds<N, kRadixS, +1, C::NV, N, C::NT>(s, a.tw, tid);
#pragma unroll
    for (int q = 0; q < C::PTS; ++q) {
      const double2 z = s[tid + q * C::NT];
      X[q] = z.x * sc;
      if (Y) Y[q] = z.y * sc;
    }
    __syncthreads();
  };
  // (w, r1)
  build_c2r_pair<N>(s, tid, [&](int vec, int k, double2& X, double2& Y) {
    X = ld(0, vec, k);
    Y = ld(3, vec, k);
  });
  fft_out(w, r1);
  // (r2, u)
  build_c2r_pair<N>(s, tid, [&](int vec, int k, double2& X, double2& Y) {
    X = ld(5, vec, k);
    Y = ld(2, vec, k);
  });
  fft_out(r2, u);
  // (v, uyvx):  v = -i kx psi;  dy u + dx v = -W + 2 kx^2 psi (+ W00 at kx = 0)
  build_c2r_pair<N>(s, tid, [&](int vec, int k, double2& X, double2& Y) {
    const double kx = a.kx1[k];
    const double2 ps = ld(1, vec, k), W = ld(0, vec, k);
    X = make_double2(kx * ps.y, -kx * ps.x);
    const double t = 2.0 * kx * kx;
    Y = make_double2(-W.x + t * ps.x + (k == 0 ? w00 : 0.0), -W.y + t * ps.y);
  });
  fft_out(v, uyvx);
  // (L1, L2):  -kx^2 R + R_yy
  build_c2r_pair<N>(s, tid, [&](int vec, int k, double2& X, double2& Y) {
    const double kx2 = a.kx1[k] * a.kx1[k];
    const double2 R1 = ld(3, vec, k), R1yy = ld(4, vec, k), R2 = ld(5, vec, k), R2yy = ld(6, vec, k);
    X = make_double2(-kx2 * R1.x + R1yy.x, -kx2 * R1.y + R1yy.y);
    Y = make_double2(-kx2 * R2.x + R2yy.x, -kx2 * R2.y + R2yy.y);
  });
  fft_out(L1, L2);
  // (ux, -):  i kx u
  build_c2r_pair<N>(s, tid, [&](int vec, int k, double2& X, double2& Y) {
    const double kx = a.kx1[k];
    const double2 U = ld(2, vec, k);
    X = make_double2(-kx * U.y, kx * U.x);
    Y = make_double2(0.0, 0.0);
  });
  fft_out(ux, nullptr);

  // ---- pointwise products (friction/Qsolver.cuh:63-164, 206-300) ----------
  double pA[C::PTS], pB[C::PTS], p11[C::PTS], p12[C::PTS], p21[C::PTS], loc1[C::PTS], loc2[C::PTS],
      U1[C::PTS], V1[C::PTS], U2[C::PTS], V2[C::PTS];
#pragma unroll
  for (int q = 0; q < C::PTS; ++q) {
    const int e = tid + q * C::NT;
    const int vec = e / N, x = e % N;
    const int y = rowof(vec);
    const double al = a.Ra ? (y < a.Ny ? a.Ra[(size_t)y * N + x] : 0.0) : a.alpha;
    const double S = 2 * sqrt(r1[q] * r1[q] + r2[q] * r2[q]);
    const double cS2 = a.cQ * S * S;
    const double g = a.lambda * S * (cS2 - a.aQ);
    const double mlD = -1.0 * a.lambda * a.DQ;
    const double s1 = mlD * L1[q] * S + g * r1[q];
    const double s2 = mlD * L2[q] * S + g * r2[q];
    const double cr = 2.0 * a.DQ * L1[q] * r2[q] + (-2.0 * a.DQ) * r1[q] * L2[q];
    p11[q] = s1 - al * r1[q];
    p12[q] = (s2 + cr) - al * r2[q];
    p21[q] = (s2 - cr) - al * r2[q];
    pA[q] = a.advect_w ? -1.0 * u[q] * w[q] : 0.0;
    pB[q] = a.advect_w ? -1.0 * v[q] * w[q] : 0.0;
    loc1[q] = (-1.0 * w[q] * r2[q] - cS2 * r1[q]) + a.lambda * S * ux[q];
    loc2[q] = (1.0 * w[q] * r1[q] - cS2 * r2[q]) + 0.5 * a.lambda * S * uyvx[q];
    U1[q] = -1.0 * u[q] * r1[q];
    V1[q] = -1.0 * v[q] * r1[q];
    U2[q] = -1.0 * u[q] * r2[q];
    V2[q] = -1.0 * v[q] * r2[q];
  }

  // ---- x R2C of the products (six complex FFTs) and the combinations ------
  auto r2c = [&](const double* X, const double* Y, auto emit) {
#pragma unroll
    for (int q = 0; q < C::PTS; ++q) s[tid + q * C::NT] = make_double2(X[q], Y ? Y[q] : 0.0);
    __syncthreads();
    fft_lds<N, kRadixS, -1, C::NV, N, C::NT>(s, a.tw, tid);
#pragma unroll
    for (int q = 0; q < C::BINS; ++q) {
      const int e = tid + q * C::NT;
      if (e < C::NV * C::H) {
        const int vec = e / C::H, k = e % C::H;
        const int y = rowof(vec);
        double2 A, B;
        unpack_r2c<N>(s, vec, k, A, B);
        if (y < a.Ny) emit((size_t)y * C::H + k, k, A, B);
      }
    }
    __syncthreads();
  };
  const double sE = a.s_ErRe;
  // (A, P12) -> F0 = i kx A [k != N/2] - kx^2 s P12
  r2c(pA, p12, [&](size_t o, int k, double2 A, double2 P) {
    const double kx = a.kx1[k];
    const double kxa = (k == N / 2) ? 0.0 : kx;
    const double c = -kx * kx * sE;
    a.mix[0][o] = make_double2(-kxa * A.y + c * P.x, kxa * A.x + c * P.y);
  });
  // (B, P11) -> F1 = B - 2 i kx s P11
  r2c(pB, p11, [&](size_t o, int k, double2 B, double2 P) {
    const double t = 2.0 * a.kx1[k] * sE;
    a.mix[1][o] = make_double2(B.x + t * P.y, B.y - t * P.x);
  });
  // (P21, -) -> F2 = s P21
  r2c(p21, nullptr, [&](size_t o, int k, double2 P, double2) {
    a.mix[2][o] = make_double2(sE * P.x, sE * P.y);
  });
  // (loc1, U1) -> F3 = loc1 + i kx U1
  r2c(loc1, U1, [&](size_t o, int k, double2 Lc, double2 U) {
    const double kx = a.kx1[k];
    a.mix[3][o] = make_double2(Lc.x - kx * U.y, Lc.y + kx * U.x);
  });
  // (loc2, U2) -> F5 = loc2 + i kx U2
  r2c(loc2, U2, [&](size_t o, int k, double2 Lc, double2 U) {
    const double kx = a.kx1[k];
    a.mix[5][o] = make_double2(Lc.x - kx * U.y, Lc.y + kx * U.x);
  });
  // (V1, V2) -> F4, F6
  r2c(V1, V2, [&](size_t o, int, double2 X, double2 Y) {
    a.mix[4][o] = X;
    a.mix[6][o] = Y;
  });
}

// --------------------------------------------------------------------------
// Column kernel
// --------------------------------------------------------------------------
enum { MODE_PRO = 0, MODE_S1 = 1, MODE_S2 = 2 };

template <int NY>
struct ColStageCfg {
  static constexpr int TC0 = 4096 / NY;
  static constexpr int TC = TC0 < 1 ? 1 : (TC0 > 32 ? 32 : TC0);
  static constexpr int LD = NY + 1;
  static constexpr int NT = 256;
  static constexpr int EPT = TC * NY / NT;   // elements per thread
};

struct ColArgs {
  double2* mix[7];
  double2* old_[3];
  double2* new_[3];
  double* w00;
  int* flag;
  const double* kx1;
  const double* ky1;
  const double2* tw;
  int Nx, Nxh;
  double dt, Re, Rf, DQ, aQ;
};

template <int NY>
__device__ __forceinline__ void col_load_fft(double2* s, const double2* src, int i0, int Nxh,
                                             const double2* tw, int tid) {
  using C = ColStageCfg<NY>;
  for (int e = tid; e < C::TC * NY; e += C::NT) {
    const int c = e % C::TC, j = e / C::TC, i = i0 + c;
    s[c * C::LD + j] = i < Nxh ? src[(size_t)j * Nxh + i] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  fft_lds<NY, kRadixS, -1, C::TC, C::LD, C::NT>(s, tw, tid);
}

template <int NY>
__device__ __forceinline__ void col_ifft_store(double2* s, double2* dst, int i0, int Nxh,
                                               const double2* tw, int tid) {
  using C = ColStageCfg<NY>;
  __syncthreads();
  fft_lds<NY, kRadixS, +1, C::TC, C::LD, C::NT>(s, tw, tid);
  for (int e = tid; e < C::TC * NY; e += C::NT) {
    const int c = e % C::TC, j = e / C::TC, i = i0 + c;
    if (i < Nxh) dst[(size_t)j * Nxh + i] = s[c * C::LD + j];
  }
  __syncthreads();
}

template <int NY, int MODE>
__global__ __launch_bounds__(256) void k_col_stage(ColArgs a) {
  using C = ColStageCfg<NY>;
  __shared__ double2 s[C::TC * C::LD];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * C::TC;
  const int nyq_x = a.Nx / 2;
  bool bad = false;

#pragma unroll 1
  for (int fld = 0; fld < 3; ++fld) {
    double2 st[C::EPT];   // the state of the next stage (curr / new / old)
    if (MODE == MODE_PRO) {
#pragma unroll
      for (int q = 0; q < C::EPT; ++q) {
        const int e = tid + q * C::NT, c = e % C::TC, j = e / C::TC, i = i0 + c;
        st[q] = i < a.Nxh ? a.old_[fld][(size_t)j * a.Nxh + i] : make_double2(0.0, 0.0);
      }
    } else {
      // ---- N_X from the forward set -------------------------------------
      double2 nl[C::EPT];
      const int f0 = fld == 0 ? 0 : (fld == 1 ? 3 : 5);
      const int nterm = fld == 0 ? 3 : 2;
#pragma unroll 1
      for (int t = 0; t < nterm; ++t) {
        col_load_fft<NY>(s, a.mix[f0 + t], i0, a.Nxh, a.tw, tid);
#pragma unroll
        for (int q = 0; q < C::EPT; ++q) {
          const int e = tid + q * C::NT, c = e % C::TC, j = e / C::TC;
          const double2 z = s[c * C::LD + j];
          const double ky = a.ky1[j];
          if (t == 0) nl[q] = z;
          else if (t == 1) nl[q] = make_double2(nl[q].x - ky * z.y, nl[q].y + ky * z.x);
          else nl[q] = make_double2(nl[q].x + ky * ky * z.x, nl[q].y + ky * ky * z.y);
        }
        __syncthreads();
      }
      // ---- IFRK2 stage update --------------------------------------------
#pragma unroll
      for (int q = 0; q < C::EPT; ++q) {
        const int e = tid + q * C::NT, c = e % C::TC, j = e / C::TC, i = i0 + c;
        if (i >= a.Nxh) { st[q] = make_double2(0.0, 0.0); continue; }
        double2 n = nl[q];
        if (fld == 0 && (i == 0 || i == nyq_x) && (j == 0 || 2 * j == NY)) n.y = 0.0;
        const size_t o = (size_t)j * a.Nxh + i;
        const double kx = a.kx1[i], ky = a.ky1[j];
        const double k2 = kx * kx + ky * ky;
        const double L = fld == 0 ? 1.0 / a.Re * (-1.0 * k2 - a.Rf) : (-1.0 * a.DQ * k2 + a.aQ);
        const double IF = exp(L * a.dt);
        const double2 an = make_double2(n.x * a.dt, n.y * a.dt);
        if (MODE == MODE_S1) {
          const double2 old = a.old_[fld][o];
          const double2 nw0 = make_double2(old.x * IF, old.y * IF);               // IFRK2.cuh:16-26
          const double h = 1.0 / 2.0 * IF;
          a.new_[fld][o] = make_double2(nw0.x + h * an.x, nw0.y + h * an.y);      // IFRK2.cuh:28-39
          st[q] = make_double2((old.x + an.x) * IF, (old.y + an.y) * IF);       // curr (IFh = IF)
        } else {
          const double2 nw = a.new_[fld][o];
          double2 r = make_double2(nw.x + 1.0 / 2.0 * an.x, nw.y + 1.0 / 2.0 * an.y);  // IFRK2.cuh:41-51
          const bool keep = (i < ((a.Nxh - 1) * 2 / 3 + 1)) && ((j < NY / 3 + 1) || ((NY - j) < NY / 3));
          const double cut = keep ? 1.0 : 0.0;                                   // Qsolver.cuh:314-322
          r = make_double2(r.x * cut, r.y * cut);
          bad |= !isfinite(r.x) || !isfinite(r.y);
          a.new_[fld][o] = r;
          st[q] = r;
        }
      }
    }
    // ---- inverse set of the next stage from st --------------------------
    if (fld == 0) {
      // M0 = W, M1 = psi = W/k^2, M2 = i ky psi   (vel_func, QFldfuncs.cuh:204-224)
      if (i0 == 0 && tid == 0) *a.w00 = st[0].x;   // element (j=0, i=0)
#pragma unroll 1
      for (int t = 0; t < 3; ++t) {
#pragma unroll
        for (int q = 0; q < C::EPT; ++q) {
          const int e = tid + q * C::NT, c = e % C::TC, j = e / C::TC, i = i0 + c;
          double2 val = st[q];
          if (t > 0) {
            const double kx = i < a.Nxh ? a.kx1[i] : 0.0, ky = a.ky1[j];
            const double k2 = kx * kx + ky * ky;
            const double2 ps = (i == 0 && j == 0) ? make_double2(0.0, 0.0)
                                                  : make_double2(val.x / k2, val.y / k2);
            val = t == 1 ? ps : make_double2(-ky * ps.y, ky * ps.x);
          }
          s[c * C::LD + j] = val;
        }
        col_ifft_store<NY>(s, a.mix[t], i0, a.Nxh, a.tw, tid);
      }
    } else {
      const int m0 = fld == 1 ? 3 : 5;
#pragma unroll 1
      for (int t = 0; t < 2; ++t) {
#pragma unroll
        for (int q = 0; q < C::EPT; ++q) {
          const int e = tid + q * C::NT, c = e % C::TC, j = e / C::TC;
          const double ky = a.ky1[j];
          s[c * C::LD + j] = t == 0 ? st[q] : make_double2(-ky * ky * st[q].x, -ky * ky * st[q].y);
        }
        col_ifft_store<NY>(s, a.mix[m0 + t], i0, a.Nxh, a.tw, tid);
      }
    }
  }
  if (MODE == MODE_S2 && bad) atomicOr(a.flag, 1);
}

template <int N>
int launch_row_rhs(qs_solver* S) {
  using C = RowRhsCfg<N>;
  const Mesh* m = S->mesh;
  RowArgs a;
  for (int k = 0; k < 7; ++k) a.mix[k] = S->mix[k];
  a.Ra = S->Ra;
  a.w00 = S->w00;
  a.kx1 = m->priv->kx1;
  a.tw = m->priv->twx;
  a.Ny = m->Ny;
  a.alpha = S->p.alpha;
  a.lambda = S->p.lambda;
  a.DQ = S->p.D_Q; a.aQ = S->p.a_Q; a.cQ = S->p.c_Q;
  a.s_ErRe = 1.0 / (S->p.Er * S->p.Re);
  a.scale = 1.0 / ((double)m->Nx * (double)m->Ny);
  a.advect_w = S->p.advect_w;
  const int grid = (m->Ny + C::NV - 1) / C::NV;
  hipLaunchKernelGGL(k_row_rhs<N>, dim3(grid), dim3(C::NT), 0, stream_of(m), a);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

template <int NY, int MODE>
int launch_col_stage_m(qs_solver* S) {
  using C = ColStageCfg<NY>;
  const Mesh* m = S->mesh;
  ColArgs a;
  for (int k = 0; k < 7; ++k) a.mix[k] = S->mix[k];
  for (int k = 0; k < 3; ++k) { a.old_[k] = S->old_[k]; a.new_[k] = S->new_[k]; }
  a.w00 = S->w00;
  a.flag = S->flag;
  a.kx1 = m->priv->kx1;
  a.ky1 = m->priv->ky1;
  a.tw = m->priv->twy;
  a.Nx = m->Nx;
  a.Nxh = m->Nxh;
  a.dt = S->p.dt; a.Re = S->p.Re; a.Rf = S->p.Rf; a.DQ = S->p.D_Q; a.aQ = S->p.a_Q;
  const int grid = (m->Nxh + C::TC - 1) / C::TC;
  hipLaunchKernelGGL((k_col_stage<NY, MODE>), dim3(grid), dim3(C::NT), 0, stream_of(m), a);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

template <int NY>
int launch_col_stage(qs_solver* S, int mode) {
  if (mode == MODE_PRO) return launch_col_stage_m<NY, MODE_PRO>(S);
  if (mode == MODE_S1) return launch_col_stage_m<NY, MODE_S1>(S);
  return launch_col_stage_m<NY, MODE_S2>(S);
}

#define QS_DISPATCH_POW2_S(n, FN, ...)                                  \
  switch (n) {                                                          \
    case 16: return FN<16>(__VA_ARGS__);                                \
    case 32: return FN<32>(__VA_ARGS__);                                \
    case 64: return FN<64>(__VA_ARGS__);                                \
    case 128: return FN<128>(__VA_ARGS__);                              \
    case 256: return FN<256>(__VA_ARGS__);                              \
    case 512: return FN<512>(__VA_ARGS__);                              \
    case 1024: return FN<1024>(__VA_ARGS__);                            \
    default: return fail(QS_ERR_UNSUPPORTED_SIZE, "fused step: unsupported size %d", n); \
  }

int row_rhs_raw(qs_solver* S) { QS_DISPATCH_POW2_S(S->mesh->Nx, launch_row_rhs, S); }
int col_stage_raw(qs_solver* S, int mode) { QS_DISPATCH_POW2_S(S->mesh->Ny, launch_col_stage, S, mode); }

hipEvent_t take_event(qs_solver* S) {
  if (!S->ev_pool.empty()) {
    hipEvent_t e = S->ev_pool.back();
    S->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Kernel classes of the fused step, for qs_solver_profile_read.
enum { K_COL_PRO = 0, K_ROW = 1, K_COL_S1 = 2, K_COL_S2 = 3, K_NCLASS = 4 };

template <class F>
int timed(qs_solver* S, int cls, F launch) {
  if (!S->prof) return launch();
  hipEvent_t a = take_event(S), b = take_event(S);
  if (!a || !b) return fail(QS_ERR_HIP, "hipEventCreate");
  QS_HIP(hipEventRecord(a, stream_of(S->mesh)));
  QS_TRY(launch());
  QS_HIP(hipEventRecord(b, stream_of(S->mesh)));
  S->ev_used.push_back({cls, {a, b}});
  return QS_OK;
}

int row_rhs(qs_solver* S) { return timed(S, K_ROW, [&] { return row_rhs_raw(S); }); }
int col_stage(qs_solver* S, int mode) {
  const int cls = mode == MODE_PRO ? K_COL_PRO : (mode == MODE_S1 ? K_COL_S1 : K_COL_S2);
  return timed(S, cls, [&] { return col_stage_raw(S, mode); });
}

// --------------------------------------------------------------------------
// Unfused step: the reference's call sequence through the per-op entry points.
// --------------------------------------------------------------------------
enum { F_WCURR, F_WNONL, F_U, F_V, F_R1CURR, F_R1NONL, F_R2CURR, F_R2NONL, F_S, F_P11, F_P12, F_P21,
       F_RA, F_AUX, F_AUX1, F_NFIELDS };

int unfused_setup(qs_solver* S) {
  if (S->unfused_ready) return QS_OK;
  Mesh* m = S->mesh;
  for (int k = 0; k < F_NFIELDS; ++k) QS_TRY(qs_field_create(m, &S->F[k]));
  const size_t ws = (size_t)m->wavesize;
  for (int k = 0; k < 4; ++k)
    if (hipMalloc((void**)&S->IFt[k], ws) != hipSuccess) return fail(QS_ERR_OOM, "IF tables");
  // friction.cu:208-210 (IF tables; Q tables with the D_Q, a_Q coefficients)
  QS_TRY(qs_wlin_func(S->IFt[0], S->IFt[1], S->p.Re, S->p.Rf, S->p.dt, m));
  QS_TRY(qlin_c(S->IFt[2], S->IFt[3], S->p.dt, S->p.D_Q, S->p.a_Q, m));
  S->unfused_ready = true;
  return QS_OK;
}

int unfused_activity(qs_solver* S) {
  Mesh* m = S->mesh;
  if (S->Ra) return qs_FldCopy(S->F[F_RA]->phys, S->Ra, m);
  return qs_FldSet(S->F[F_RA]->phys, S->p.alpha, m);
}

int unfused_rhs(qs_solver* S) {
  Field** F = S->F;
  const qs_params& p = S->p;
  QS_TRY(qs_curr_func(F[F_R1CURR], F[F_R2CURR], F[F_WCURR], F[F_U], F[F_V], F[F_S]));
  QS_TRY(wnonl_func_c(F[F_WNONL], F[F_WCURR], F[F_U], F[F_V], F[F_R1CURR], F[F_R2CURR], F[F_RA],
                      F[F_S], F[F_P11], F[F_P12], F[F_P21], p.Re, p.Er, p.lambda, F[F_AUX],
                      F[F_AUX1], S->coef));
  QS_TRY(r1nonl_func_c(F[F_R1NONL], F[F_WCURR], F[F_U], F[F_V], F[F_R1CURR], F[F_R2CURR], F[F_S],
                       p.lambda, F[F_AUX], F[F_AUX1], S->coef));
  return r2nonl_func_c(F[F_R2NONL], F[F_WCURR], F[F_U], F[F_V], F[F_R1CURR], F[F_R2CURR], F[F_S],
                       p.lambda, F[F_AUX], F[F_AUX1], S->coef);
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

int qs_solver_create(Mesh* mesh, const qs_params* params, qs_solver** out) {
  if (!out || !params) return fail(QS_ERR_INVALID_ARG, "qs_solver_create: null argument");
  *out = nullptr;
  QS_TRY(check_mesh(mesh));
  if (mesh->Nx > 1024 || mesh->Ny > 1024)
    return fail(QS_ERR_UNSUPPORTED_SIZE, "qs_solver_create: fused step supports N <= 1024");
  if (!(params->Re != 0) || !(params->Er != 0) || !(params->dt > 0))
    return fail(QS_ERR_INVALID_ARG, "qs_solver_create: Re, Er must be non-zero and dt > 0");
  qs_solver* S = new qs_solver();
  S->mesh = mesh;
  S->p = *params;
  S->coef.DQ = params->D_Q;
  S->coef.aQ = params->a_Q;
  S->coef.cQ = params->c_Q;
  S->coef.advect_w = params->advect_w != 0;
  const size_t hs = (size_t)mesh->specsize;
  bool ok = true;
  for (int k = 0; k < 3 && ok; ++k)
    ok = hipMalloc((void**)&S->old_[k], hs) == hipSuccess && hipMalloc((void**)&S->new_[k], hs) == hipSuccess;
  for (int k = 0; k < 7 && ok; ++k) ok = hipMalloc((void**)&S->mix[k], hs) == hipSuccess;
  ok = ok && hipMalloc((void**)&S->w00, sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&S->flag, sizeof(int)) == hipSuccess;
  if (!ok) {
    qs_solver_destroy(S);
    return fail(QS_ERR_OOM, "qs_solver_create: hipMalloc");
  }
  for (int k = 0; k < 3; ++k) {
    if (hipMemsetAsync(S->old_[k], 0, hs, stream_of(mesh)) != hipSuccess ||
        hipMemsetAsync(S->new_[k], 0, hs, stream_of(mesh)) != hipSuccess) {
      qs_solver_destroy(S);
      return fail(QS_ERR_HIP, "qs_solver_create: memset");
    }
  }
  if (hipMemsetAsync(S->flag, 0, sizeof(int), stream_of(mesh)) != hipSuccess) {
    qs_solver_destroy(S);
    return fail(QS_ERR_HIP, "qs_solver_create: memset");
  }
  *out = S;
  return QS_OK;
}

int qs_solver_destroy(qs_solver* S) {
  if (!S) return QS_OK;
  for (int k = 0; k < 3; ++k) {
    if (S->old_[k]) (void)hipFree(S->old_[k]);
    if (S->new_[k]) (void)hipFree(S->new_[k]);
  }
  for (int k = 0; k < 7; ++k)
    if (S->mix[k]) (void)hipFree(S->mix[k]);
  if (S->Ra) (void)hipFree(S->Ra);
  if (S->w00) (void)hipFree(S->w00);
  if (S->flag) (void)hipFree(S->flag);
  for (auto* f : S->F) qs_field_destroy(f);
  for (auto* t : S->IFt)
    if (t) (void)hipFree(t);
  for (auto& u : S->ev_used) {
    (void)hipEventDestroy(u.second.first);
    (void)hipEventDestroy(u.second.second);
  }
  for (auto e : S->ev_pool) (void)hipEventDestroy(e);
  delete S;
  return QS_OK;
}

int qs_solver_set_activity(qs_solver* S, const Qreal* Ra) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  const Mesh* m = S->mesh;
  if (!Ra) {
    if (S->Ra) (void)hipFree(S->Ra);
    S->Ra = nullptr;
    return QS_OK;
  }
  if (!S->Ra && hipMalloc((void**)&S->Ra, (size_t)m->physsize) != hipSuccess)
    return fail(QS_ERR_OOM, "activity field");
  QS_HIP(hipMemcpyAsync(S->Ra, Ra, (size_t)m->physsize, hipMemcpyDeviceToDevice, stream_of(m)));
  return QS_OK;
}

int qs_solver_set_state(qs_solver* S, const Qcomp* w, const Qcomp* r1, const Qcomp* r2) {
  if (!S || !w || !r1 || !r2) return fail(QS_ERR_INVALID_ARG, "qs_solver_set_state: null argument");
  const Mesh* m = S->mesh;
  const Qcomp* src[3] = {w, r1, r2};
  for (int k = 0; k < 3; ++k)
    QS_HIP(hipMemcpyAsync(S->old_[k], src[k], (size_t)m->specsize, hipMemcpyDeviceToDevice, stream_of(m)));
  S->mixed_valid = false;
  return QS_OK;
}

int qs_solver_get_state(qs_solver* S, Qcomp* w, Qcomp* r1, Qcomp* r2) {
  if (!S || !w || !r1 || !r2) return fail(QS_ERR_INVALID_ARG, "qs_solver_get_state: null argument");
  const Mesh* m = S->mesh;
  Qcomp* dst[3] = {w, r1, r2};
  for (int k = 0; k < 3; ++k)
    QS_HIP(hipMemcpyAsync(dst[k], S->old_[k], (size_t)m->specsize, hipMemcpyDeviceToDevice, stream_of(m)));
  return QS_OK;
}

int qs_step_ifrk2(qs_solver* S, int nsteps) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  if (nsteps < 0) return fail(QS_ERR_INVALID_ARG, "nsteps < 0");
  for (int n = 0; n < nsteps; ++n) {
    if (!S->mixed_valid) {
      QS_TRY(col_stage(S, MODE_PRO));
      S->mixed_valid = true;
    }
    QS_TRY(row_rhs(S));
    QS_TRY(col_stage(S, MODE_S1));
    QS_TRY(row_rhs(S));
    QS_TRY(col_stage(S, MODE_S2));
    for (int k = 0; k < 3; ++k) std::swap(S->old_[k], S->new_[k]);
  }
  return QS_OK;
}

// friction.cu:212-260 with the IFRK2 stage kernels and IFh := IF.
int qs_step_ifrk2_unfused(qs_solver* S, int nsteps) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  QS_TRY(unfused_setup(S));
  QS_TRY(unfused_activity(S));
  Mesh* m = S->mesh;
  Field** F = S->F;
  const double dt = S->p.dt;
  Field* curr[3] = {F[F_WCURR], F[F_R1CURR], F[F_R2CURR]};
  Field* nonl[3] = {F[F_WNONL], F[F_R1NONL], F[F_R2NONL]};
  for (int n = 0; n < nsteps; ++n) {
    const double* IF[3] = {S->IFt[0], S->IFt[2], S->IFt[2]};
    for (int k = 0; k < 3; ++k)
      QS_TRY(qs_ifrk2_func0D((Qcomp*)S->old_[k], curr[k]->spec, (Qcomp*)S->new_[k], IF[k], IF[k], dt, m));
    QS_TRY(unfused_rhs(S));
    for (int k = 0; k < 3; ++k)
      QS_TRY(qs_ifrk2_func1D((Qcomp*)S->old_[k], curr[k]->spec, (Qcomp*)S->new_[k], nonl[k]->spec, IF[k],
                             IF[k], dt, m));
    QS_TRY(unfused_rhs(S));
    for (int k = 0; k < 3; ++k)
      QS_TRY(qs_ifrk2_func2D((Qcomp*)S->old_[k], curr[k]->spec, (Qcomp*)S->new_[k], nonl[k]->spec, IF[k],
                             IF[k], dt, m));
    for (int k = 0; k < 3; ++k) QS_TRY(qs_dealias((Qcomp*)S->new_[k], m));
    for (int k = 0; k < 3; ++k) QS_TRY(qs_SpecSet((Qcomp*)S->old_[k], (Qcomp*)S->new_[k], m));
  }
  S->mixed_valid = false;
  return QS_OK;
}

int qs_solver_profile(qs_solver* S, int enable) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  S->prof = enable != 0;
  return QS_OK;
}

int qs_solver_profile_read(qs_solver* S, double* avg_ms, int* counts) {
  if (!S || !avg_ms || !counts) return fail(QS_ERR_INVALID_ARG, "null argument");
  double sum[K_NCLASS] = {0, 0, 0, 0};
  int cnt[K_NCLASS] = {0, 0, 0, 0};
  QS_HIP(hipStreamSynchronize(stream_of(S->mesh)));
  for (auto& u : S->ev_used) {
    float ms = 0.f;
    QS_HIP(hipEventElapsedTime(&ms, u.second.first, u.second.second));
    sum[u.first] += ms;
    cnt[u.first] += 1;
    S->ev_pool.push_back(u.second.first);
    S->ev_pool.push_back(u.second.second);
  }
  S->ev_used.clear();
  for (int k = 0; k < K_NCLASS; ++k) {
    counts[k] = cnt[k];
    avg_ms[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
  }
  return QS_OK;
}

int qs_solver_nonfinite(qs_solver* S, int* flag) {
  if (!S || !flag) return fail(QS_ERR_INVALID_ARG, "null argument");
  const Mesh* m = S->mesh;
  QS_HIP(hipMemcpyAsync(flag, S->flag, sizeof(int), hipMemcpyDeviceToHost, stream_of(m)));
  QS_HIP(hipStreamSynchronize(stream_of(m)));
  QS_HIP(hipMemsetAsync(S->flag, 0, sizeof(int), stream_of(m)));
  return QS_OK;
}

}  // extern "C"
