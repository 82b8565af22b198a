// qs_adjoint.hip -- discrete adjoint kernels of the fused IFRK2 step (gfx950).
//
// No reference counterpart: the reference (whpy/nature-control) has no
// adjoint.  These kernels are the exact transpose, in the real inner products
// of the stored arrays (a half-spectrum entry is two real coordinates), of the
// fused forward kernels in qs_fused.hip, which compute the step of
// InertialFriction/friction/Qsolver.cuh:63-322 and TimeIntegration/IFRK2.cuh.
//
// Forward step (per field, E = exp(L dt), M = dealias mask):
//   c1 = u_n,  c2 = E (u_n + dt N(c1)),  u_{n+1} = M (E u_n + dt/2 E N(c1) + dt/2 N(c2)).
// Reverse step, given g = dL/du_{n+1}:
//   lambda = M g;  cbar2 = N'(c2)^T (dt/2 lambda);  acc = E (lambda + cbar2);
//   cbar1 = N'(c1)^T (E (dt/2 lambda + dt cbar2));  dL/du_n = acc + cbar1.
// N'(c)^T is a column kernel (transposed y combinations -> Fbar, y inverse
// FFTs), a row kernel (transposed x R2C -> product adjoints, the stage's
// physical fields, the pointwise Jacobian transposed, transposed x C2R ->
// Mbar) and the y FFTs of Mbar at the start of the next column kernel, which
// also applies the IFRK2 combinations above:
//   AMODE_START: lambda = M g (g in lam_), Fbar of dt/2 lambda
//   AMODE_B:     cbar2 from Mbar; acc_ = E (lambda + cbar2); Fbar of E (dt/2 lambda + dt cbar2)
//   AMODE_A:     cbar1 from Mbar; lambda' = M (acc_ + cbar1) -> lam_; Fbar of dt/2 lambda'
//   AMODE_END:   cbar1 from Mbar; acc_ + cbar1 -> uout (dL/du_n at a segment start)
// Transposes: (x R2C)^T G = C2R(Wt G) with Wt = 1 on bins 0 and N/2, 1/2
// elsewhere; (x C2R)^T f = c R2C(f), c = 1 on bins 0 and N/2, 2 elsewhere;
// unnormalised y FFT <-> y inverse FFT; a complex multiplier transposes to its
// conjugate; the real-part projection at the self-conjugate bins is self-adjoint.
#include <vector>

#include "qs_fused_dev.hpp"

namespace qs {
namespace {

__device__ __forceinline__ double wave_sum(double s) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
  return s;
}

// ===========================================================================
// Row adjoint job (one job = NV = WaveLay<N>::NV rows, 6 waves):
//  A. Fbar rows -> transposed x combinations -> (x R2C)^T: physical adjoints
//     of the 11 products, as 6 pairs (waves 0-5), kept in the block's scratch
//     slot (global memory: they do not fit LDS beside the fields);
//  B. the stage's physical fields from its inverse set (as the forward row);
//  C. pointwise Jacobian transposed -> field adjoints in LDS; the activity
//     gradient (field and per-row sums);
//  D. (x C2R)^T of the field adjoints (waves 0-4) -> Mbar over the Fbar rows,
//     and the per-row parts of dL/dRe W(0,0).
// ===========================================================================
template <int N>
__device__ __forceinline__ void arow_job(const RunDesc& a, int row0, double2* sm, const double2* twq,
                                         double2* scr) {
  using L = WaveLay<N>;
  constexpr int NV = L::NV, LPC = L::LPC, VS = L::VS, E = L::E, H = N / 2 + 1, RH = NV * H;
  const TwN<N> tw{twq};
  // at 512 (two rows per job) the lane offsets, held across the job, spilled
  // (236 B of scratch per lane): there they are recomputed per phase from an
  // opaque thread index (job_tid); the other sides keep them (faster)
  constexpr bool OPQ = N == 512;
  const int tid = OPQ ? job_tid() : (int)threadIdx.x;
  // the wave index as a scalar: wave-uniform branches, one unrolled loop each
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int lp = lane / NV, g = lane % NV;
  double2* vb = sm + wave * kBuf + g * VS;
  const double sE = a.s_ErRe;
  double2 v[16];

  // ---- A. product adjoints ------------------------------------------------
  stage_rows<N>(a.mix, a, row0, sm);
  {
    const double2* st = sm + g * H;
    auto build = [&](auto&& pair) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int k = lp + LPC * m;
        const int bin = k > N / 2 ? N - k : k;
        const double kx = a.dkx * bin;
        const double wt = (bin == 0 || 2 * bin == N) ? 1.0 : 0.5;
        double2 X = make_double2(0.0, 0.0), Y = make_double2(0.0, 0.0);
        pair(bin, kx, X, Y);
        v[m] = herm_pair<N>(k, make_double2(wt * X.x, wt * X.y), make_double2(wt * Y.x, wt * Y.y));
      }
    };
    if (wave == 0) {            // (Abar, Bbar) = (-i kx' F0bar, F1bar)
      const bool adv = a.advect_w;
      build([&](int bin, double kx, double2& X, double2& Y) {
        if (adv) {
          const double kxa = (2 * bin == N) ? 0.0 : kx;
          const double2 F0 = st[0 * RH + bin];
          X = make_double2(kxa * F0.y, -kxa * F0.x);
          Y = st[1 * RH + bin];
        }
      });
    } else if (wave == 1) {     // (U1bar, V1bar) = (-i kx F3bar, F4bar)
      build([&](int bin, double kx, double2& X, double2& Y) {
        const double2 F3 = st[3 * RH + bin];
        X = make_double2(kx * F3.y, -kx * F3.x);
        Y = st[4 * RH + bin];
      });
    } else if (wave == 2) {     // (U2bar, V2bar) = (-i kx F5bar, F6bar)
      build([&](int bin, double kx, double2& X, double2& Y) {
        const double2 F5 = st[5 * RH + bin];
        X = make_double2(kx * F5.y, -kx * F5.x);
        Y = st[6 * RH + bin];
      });
    } else if (wave == 3) {     // (P11bar, P12bar) = (2 i kx s F1bar, -kx^2 s F0bar)
      build([&](int bin, double kx, double2& X, double2& Y) {
        const double2 F0 = st[0 * RH + bin], F1 = st[1 * RH + bin];
        const double t = 2.0 * kx * sE, c = -kx * kx * sE;
        X = make_double2(-t * F1.y, t * F1.x);
        Y = make_double2(c * F0.x, c * F0.y);
      });
    } else if (wave == 4) {     // (P21bar, -) = (s F2bar, 0)
      build([&](int bin, double, double2& X, double2&) {
        const double2 F2 = st[2 * RH + bin];
        X = make_double2(sE * F2.x, sE * F2.y);
      });
    } else {                    // (loc1bar, loc2bar) = (F3bar, F5bar)
      build([&](int bin, double, double2& X, double2& Y) {
        X = st[3 * RH + bin];
        Y = st[5 * RH + bin];
      });
    }
  }
  __syncthreads();   // staging area free
  wave_fft<N, +1>(v, vb, lp, tw);
#pragma unroll
  for (int m = 0; m < E; ++m) gstore(scr + wave * 1024 + g * N + lp + LPC * m, v[m]);

  // ---- B. the stage's physical fields (buffers 0-4) -------------------------
  stage_rows<N>(a.minp, a, row0, sm);   // its barrier also publishes the scratch writes
  inverse_fields<N, wave_e(N), OPQ>(a, row0, sm, tw, v, gload(a.w00));

  // ---- C. pointwise Jacobian transposed (friction/Qsolver.cuh:63-300) -------
  const double alu = a.alpha_p ? gload(a.alpha_p) : a.alpha;
  const double lam = a.lambda, DQ = a.DQ, aQ = a.aQ, cQ = a.cQ;
  for (int p = tid; p < NV * N; p += 384) {
    const int gg = p / N, x = p % N;
    const int o = gg * VS + swz(x);
    const int yy = row0 + gg;
    const bool yok = yy < a.Ny;
    double2 pb[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) pb[q] = gload(scr + q * 1024 + gg * N + x);
    const double2 f0 = sm[0 * kBuf + o], f1 = sm[1 * kBuf + o], f2 = sm[2 * kBuf + o],
                  f3 = sm[3 * kBuf + o], f4 = sm[4 * kBuf + o];
    const double w = f0.x, uyvx = f0.y, u = f1.x, vv = f1.y, r1 = f2.x, r2 = f2.y;
    const double L1 = f3.x, L2 = f3.y, ux = f4.x;
    const double al = a.Ra ? gload(a.Ra + (size_t)(yok ? yy : a.Ny - 1) * N + x) : alu;
    const double rr = sqrt(r1 * r1 + r2 * r2), S = 2.0 * rr, cS2 = cQ * S * S;
    const double gs = lam * S * (cS2 - aQ);
    double wb = 0, uyvxb = 0, ub = 0, vbb = 0, uxb = 0, r1b = 0, r2b = 0, L1b = 0, L2b = 0;
    double Sb = 0, cS2b = 0, gsb = 0;
    if (a.advect_w) {             // A = -u w, B = -v w
      ub += -w * pb[0].x; wb += -u * pb[0].x;
      vbb += -w * pb[0].y; wb += -vv * pb[0].y;
    }
    ub += -r1 * pb[1].x; r1b += -u * pb[1].x;      // -u r1
    vbb += -r1 * pb[1].y; r1b += -vv * pb[1].y;    // -v r1
    ub += -r2 * pb[2].x; r2b += -u * pb[2].x;      // -u r2
    vbb += -r2 * pb[2].y; r2b += -vv * pb[2].y;    // -v r2
    const double p11b = pb[3].x, p12b = pb[3].y, p21b = pb[4].x;
    const double s1b = p11b, s2b = p12b + p21b, crb = p12b - p21b;
    r1b += -al * p11b;
    r2b += -al * (p12b + p21b);
    const double abar = -r1 * p11b - r2 * (p12b + p21b);
    {   // loc1 = -w r2 - cS2 r1 + lam S ux
      const double t = pb[5].x;
      wb += -r2 * t; r2b += -w * t; cS2b += -r1 * t; r1b += -cS2 * t; Sb += lam * ux * t; uxb += lam * S * t;
    }
    {   // loc2 = w r1 - cS2 r2 + 1/2 lam S uyvx
      const double t = pb[5].y;
      wb += r1 * t; r1b += w * t; cS2b += -r2 * t; r2b += -cS2 * t; Sb += 0.5 * lam * uyvx * t;
      uyvxb += 0.5 * lam * S * t;
    }
    // cross = 2 DQ (L1 r2 - r1 L2)
    L1b += 2.0 * DQ * r2 * crb; r2b += 2.0 * DQ * L1 * crb; r1b += -2.0 * DQ * L2 * crb; L2b += -2.0 * DQ * r1 * crb;
    // single_a = -lam DQ L_a S + gs r_a
    L1b += -lam * DQ * S * s1b; Sb += -lam * DQ * L1 * s1b; gsb += r1 * s1b; r1b += gs * s1b;
    L2b += -lam * DQ * S * s2b; Sb += -lam * DQ * L2 * s2b; gsb += r2 * s2b; r2b += gs * s2b;
    // gs = lam S (cS2 - aQ), cS2 = cQ S^2, S = 2 sqrt(r1^2 + r2^2)
    Sb += gsb * lam * (cS2 - aQ);
    cS2b += gsb * lam * S;
    Sb += cS2b * 2.0 * cQ * S;
    if (rr > 0.0) {
      r1b += Sb * 2.0 * r1 / rr;
      r2b += Sb * 2.0 * r2 / rr;
    }
    if (yok && a.abar) {
      double* ap = a.abar + (size_t)yy * N + x;
      *(QS_GLOBAL double*)ap = gload(ap) + abar;
    }
    sm[0 * kBuf + o] = make_double2(wb, uyvxb);
    sm[1 * kBuf + o] = make_double2(ub, vbb);
    sm[2 * kBuf + o] = make_double2(r1b, r2b);
    sm[3 * kBuf + o] = make_double2(L1b, L2b);
    sm[4 * kBuf + o] = make_double2(uxb, 0.0);
    sm[5 * kBuf + o] = make_double2(yok ? abar : 0.0, 0.0);
  }
  __syncthreads();
  // per-row sums of the activity gradient (fixed order: deterministic)
  if (a.arow) {
    for (int gg = wave; gg < NV; gg += 6) {
      double s = 0.0;
      for (int x = lane; x < N; x += 64) s += sm[5 * kBuf + gg * VS + swz(x)].x;
      s = wave_sum(s);
      if (lane == 0 && row0 + gg < a.Ny) {
        double* ap = a.arow + row0 + gg;
        *(QS_GLOBAL double*)ap = gload(ap) + s;
      }
    }
  }

  // ---- D. (x C2R)^T of the field adjoints -> Mbar ---------------------------
  if (wave < 5) {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = vb[swz(lp + LPC * m)];
    wave_sync();
    wave_fft<N, -1>(v, vb, lp, tw);
#pragma unroll
    for (int m = 0; m < E; ++m) vb[swz(lp + LPC * m)] = v[m];
  }
  __syncthreads();
  const double sc = a.scale;
  for (int e = tid; e < RH; e += 384) {
    const int vec = e / H, k = e % H;
    const int y = row0 + vec;
    if (y >= a.Ny) continue;
    const int o = vec * VS + swz(k), oc = vec * VS + swz((N - k) & (N - 1));
    const double cw = (k == 0 || 2 * k == N) ? sc : 2.0 * sc;
    double2 Sp[5][2];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      const double2 zk = sm[q * kBuf + o], zc = cconj(sm[q * kBuf + oc]);
      Sp[q][0] = make_double2(0.5 * cw * (zk.x + zc.x), 0.5 * cw * (zk.y + zc.y));
      const double2 d = csub(zk, zc);
      Sp[q][1] = make_double2(0.5 * cw * d.y, -0.5 * cw * d.x);
    }
    const double2 wb = Sp[0][0], uyvxb = Sp[0][1], ub = Sp[1][0], vbb = Sp[1][1];
    const double2 r1b = Sp[2][0], r2b = Sp[2][1], L1b = Sp[3][0], L2b = Sp[3][1], uxb = Sp[4][0];
    const double kx = a.dkx * k, t2 = 2.0 * kx * kx, k2x = kx * kx;
    const unsigned ro = (unsigned)(y * a.pitch + k);
    gstore(a.mix[0] + ro, make_double2(wb.x - uyvxb.x, wb.y - uyvxb.y));                        // w, -uyvx
    gstore(a.mix[1] + ro, make_double2(t2 * uyvxb.x - kx * vbb.y, t2 * uyvxb.y + kx * vbb.x));   // 2kx^2 uyvx + i kx v
    gstore(a.mix[2] + ro, make_double2(ub.x + kx * uxb.y, ub.y - kx * uxb.x));                  // u - i kx ux
    gstore(a.mix[3] + ro, make_double2(r1b.x - k2x * L1b.x, r1b.y - k2x * L1b.y));
    gstore(a.mix[4] + ro, L1b);
    gstore(a.mix[5] + ro, make_double2(r2b.x - k2x * L2b.x, r2b.y - k2x * L2b.y));
    gstore(a.mix[6] + ro, L2b);
    if (k == 0) *(QS_GLOBAL double*)(a.w00bar + y) = uyvxb.x;   // W(0,0) enters dy u + dx v at bin 0
  }
}

// ===========================================================================
// Column adjoint job: one field on a tile of columns (as the forward).
// ===========================================================================
template <int NY, int AMODE>
__device__ __forceinline__ void acol_job(const RunDesc& a, const int FLD, int i0, double2* sm, const double2* twq) {
  constexpr int LPC = WaveLay<NY>::LPC, E = WaveLay<NY>::E;
  const TwN<NY> tw{twq};
  ColTile<NY> T(a, sm, i0);
  const int lp = T.lp;
  // sides >= 512: wavenumbers and LDS slots from an opaque lane index,
  // recomputed per use (held across the job they spilled); 256 keeps them
  auto lpo = [&]() { return NY >= 512 ? opq(lp) : lp; };
  const int wave = T.tid >> 6;
  const int i = i0 + wave * ColTile<NY>::CW + T.g;
  const bool iok = i < a.Nxh;
  const int ic = iok ? i : 0;
  double2* vb = T.vb;
  const double kx = a.dkx * ic;
  const double dt = a.dt;
  auto kyj = [&](int j) { return a.dky * (j < NY / 2 + 1 ? j : j - NY); };
  auto keep = [&](int j) { return (i < ((a.Nxh - 1) * 2 / 3 + 1)) && ((j < NY / 3 + 1) || ((NY - j) < NY / 3)); };
  const IFac ifc(a, FLD, ic, NY);   // the forward's integrating factor, exp(L dt)
  auto IFof = [&](int j) { return ifc.IF(j); };

  // W(0,0) enters every row's dy u + dx v at bin 0: its adjoint is the sum of
  // the row adjoints' parts (fixed-order block reduction; thread 0 owns the
  // element i = 0, j = 0 of the first tile)
  double w00sum = 0.0;
  if (AMODE != AMODE_START && FLD == 0 && i0 == 0) {   // block-uniform
    double s = 0.0;
    for (int y = threadIdx.x; y < NY; y += 256) s += gload(a.w00bar + y);
    sm[threadIdx.x] = make_double2(s, 0.0);
    __syncthreads();
    if (threadIdx.x == 0)
      for (int t = 0; t < 256; ++t) w00sum += sm[t].x;
    __syncthreads();
  }

  double2 st[16];
  if (AMODE == AMODE_START) {
    T.load(a.lam_[FLD]);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int j = lpo() + LPC * m;
      double2* sl = vb + swz(j);
      const double c = keep(j) ? 1.0 : 0.0;
      const double2 lamv = make_double2(sl->x * c, sl->y * c);
      *sl = lamv;
      st[m] = make_double2(0.5 * dt * lamv.x, 0.5 * dt * lamv.y);
    }
    T.store(a.lam_[FLD]);
  } else {
    // ---- cbar = (y inverse set)^T Mbar: y FFTs of the field's Mbar arrays ----
    // (tiles staged through the wave buffers; the last one loaded is the
    // state tile of the combination below)
    const double2* state_in = AMODE == AMODE_B ? a.lam_[FLD] : a.acc_[FLD];
    if (FLD == 0) {
      T.stage_in(a.mix[0], st);
      wave_fft<NY, -1>(st, vb, lp, tw);
      {
        double2 v[16];
        T.stage_in(a.mix[1], v);
        wave_fft<NY, -1>(v, vb, lp, tw);
#pragma unroll
        for (int m = 0; m < E; ++m) {   // psi = W / k^2
          const int j = lpo() + LPC * m;
          const double ky = kyj(j);
          const double k2 = kx * kx + ky * ky;
          const double ik2 = (i == 0 && j == 0) || !iok ? 0.0 : recip(k2);
          st[m] = make_double2(st[m].x + ik2 * v[m].x, st[m].y + ik2 * v[m].y);
        }
      }
      {
        double2 v[16];
        T.stage_in(a.mix[2], v);
        wave_fft<NY, -1>(v, vb, lp, tw);
#pragma unroll
        for (int m = 0; m < E; ++m) {   // i ky psi -> conj: -i ky / k^2
          const int j = lpo() + LPC * m;
          const double ky = kyj(j);
          const double k2 = kx * kx + ky * ky;
          const double t = (i == 0 && j == 0) || !iok ? 0.0 : ky * recip(k2);
          st[m] = make_double2(st[m].x + t * v[m].y, st[m].y - t * v[m].x);
        }
      }
      if (i0 == 0 && threadIdx.x == 0) st[0].x += w00sum;   // element (i = 0, j = 0)
    } else {
      const int m0 = FLD == 1 ? 3 : 5;
      T.stage_in(a.mix[m0], st);
      wave_fft<NY, -1>(st, vb, lp, tw);
      double2 v[16];
      T.stage_in(a.mix[m0 + 1], v);
      wave_fft<NY, -1>(v, vb, lp, tw);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double ky = kyj(lpo() + LPC * m);
        st[m] = make_double2(st[m].x - ky * ky * v[m].x, st[m].y - ky * ky * v[m].y);
      }
    }
    // ---- IFRK2 reverse combinations ----------------------------------------
    // (the state tile, lambda or acc, staged in the wave buffers and
    // overwritten in place by the results for the store)
    T.load(state_in);
    if (AMODE == AMODE_B) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int j = lpo() + LPC * m;
        double2* sl = vb + swz(j);
        const double2 lamv = *sl;
        const double IFj = IFof(j);
        *sl = make_double2(IFj * (lamv.x + st[m].x), IFj * (lamv.y + st[m].y));
        st[m] = make_double2(IFj * (0.5 * dt * lamv.x + dt * st[m].x), IFj * (0.5 * dt * lamv.y + dt * st[m].y));
      }
      T.store(a.acc_[FLD]);
    } else if (AMODE == AMODE_A) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int j = lpo() + LPC * m;
        double2* sl = vb + swz(j);
        const double c = keep(j) ? 1.0 : 0.0;
        const double2 acc = *sl;
        const double2 lamv = make_double2(c * (acc.x + st[m].x), c * (acc.y + st[m].y));
        *sl = lamv;
        st[m] = make_double2(0.5 * dt * lamv.x, 0.5 * dt * lamv.y);
      }
      T.store(a.lam_[FLD]);
    } else {   // AMODE_END
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int j = lpo() + LPC * m;
        const double2 acc = vb[swz(j)];
        vb[swz(j)] = make_double2(acc.x + st[m].x, acc.y + st[m].y);
      }
      if (a.uout[FLD]) T.store(a.uout[FLD]);
      return;
    }
  }

  // ---- Fbar = (y forward combinations)^T Nbar: y inverse FFTs ---------------
  if (FLD == 0) {
    const int nyq_x = a.Nx / 2;
#pragma unroll
    for (int m = 0; m < E; ++m) {   // real part at the self-conjugate bins
      const int j = lpo() + LPC * m;
      if ((i == 0 || i == nyq_x) && (j == 0 || 2 * j == NY)) st[m].y = 0.0;
    }
    {
      double2 v[16];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = st[m];
      wave_fft<NY, +1>(v, vb, lp, tw);
      T.stage_out(a.mix[0], v);
    }
    {
      double2 v[16];
#pragma unroll
      for (int m = 0; m < E; ++m) {   // -i ky
        const double ky = kyj(lpo() + LPC * m);
        v[m] = make_double2(ky * st[m].y, -ky * st[m].x);
      }
      wave_fft<NY, +1>(v, vb, lp, tw);
      T.stage_out(a.mix[1], v);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {     // ky^2
      const double ky = kyj(lpo() + LPC * m);
      st[m] = make_double2(ky * ky * st[m].x, ky * ky * st[m].y);
    }
    wave_fft<NY, +1>(st, vb, lp, tw);
    T.stage_out(a.mix[2], st);
  } else {
    const int m0 = FLD == 1 ? 3 : 5;
    {
      double2 v[16];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = st[m];
      wave_fft<NY, +1>(v, vb, lp, tw);
      T.stage_out(a.mix[m0], v);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {     // -i ky
      const double ky = kyj(lpo() + LPC * m);
      st[m] = make_double2(ky * st[m].y, -ky * st[m].x);
    }
    wave_fft<NY, +1>(st, vb, lp, tw);
    T.stage_out(a.mix[m0 + 1], st);
  }
}

// ===========================================================================
// Kernels (fixed mesh size per launch; grid loops over the work list).
// ===========================================================================
template <int N>
__global__ __launch_bounds__(384) void k_arow(const RunDesc* __restrict__ runs, const int2* __restrict__ work,
                                              int nwork, const double2* __restrict__ tw1024) {
  __shared__ double2 sm[6 * kBuf];
  __shared__ double2 twq[kTwq];
  stage_twiddles(twq, tw1024, threadIdx.x, 384);
  __syncthreads();
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  for (int w = L; w < nwork; w += gridDim.x) {
    const int2 it = work[w];
    const RunDesc& a = runs[it.x];
    arow_job<N>(a, it.y * WaveLay<N>::NV, sm, twq, a.scratch + (size_t)blockIdx.x * 6 * 1024);
    __syncthreads();
  }
}

// Two blocks per CU (72 KB of LDS, at most 256 registers per lane), as the
// forward column kernels.
template <int NY, int AMODE>
__global__ __launch_bounds__(256, 2) void k_acol(const RunDesc* __restrict__ runs, const int2* __restrict__ work,
                                                 int nwork, const double2* __restrict__ tw1024) {
  __shared__ double2 sm[4 * (kBuf + 1)];
  __shared__ double2 twq[kTwq];
  stage_twiddles(twq, tw1024, threadIdx.x, 256);
  __syncthreads();
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  constexpr int tc = 4 * WaveLay<NY>::NV;
  for (int w = L; w < nwork; w += gridDim.x) {
    const int2 it = work[w];
    const RunDesc& a = runs[it.x];
    const int ntiles = (a.Nxh + tc - 1) / tc;
    int fld, tile;
    col_job_map(it.y, ntiles, fld, tile);   // XCD-balanced numbering, as the forward column kernels
    acol_job<NY, AMODE>(a, fld, tile * tc, sm, twq);
    __syncthreads();
  }
}

template <int N>
int launch_arow_n(const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw, hipStream_t st) {
  hipLaunchKernelGGL(k_arow<N>, dim3(grid), dim3(384), 0, st, runs, work, nwork, tw);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

template <int N>
int launch_acol_n(int amode, const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw,
                  hipStream_t st) {
  switch (amode) {
    case AMODE_START: hipLaunchKernelGGL((k_acol<N, AMODE_START>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw); break;
    case AMODE_B: hipLaunchKernelGGL((k_acol<N, AMODE_B>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw); break;
    case AMODE_A: hipLaunchKernelGGL((k_acol<N, AMODE_A>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw); break;
    default: hipLaunchKernelGGL((k_acol<N, AMODE_END>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw); break;
  }
  QS_CHECK_LAUNCH();
  return QS_OK;
}

}  // namespace

#define QS_ADJ_SWITCH(n, FN, ...)                                         \
  switch (n) {                                                            \
    case 16: return FN<16>(__VA_ARGS__);                                  \
    case 32: return FN<32>(__VA_ARGS__);                                  \
    case 64: return FN<64>(__VA_ARGS__);                                  \
    case 128: return FN<128>(__VA_ARGS__);                                \
    case 256: return FN<256>(__VA_ARGS__);                                \
    case 512: return FN<512>(__VA_ARGS__);                                \
    case 1024: return FN<1024>(__VA_ARGS__);                              \
    default: return fail(QS_ERR_UNSUPPORTED_SIZE, "adjoint: mesh %d", n); \
  }

int launch_adj_row(int nfix, const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw1024,
                   hipStream_t st) {
  QS_ADJ_SWITCH(nfix, launch_arow_n, runs, work, nwork, grid, tw1024, st);
}

int launch_adj_col(int nfix, int amode, const RunDesc* runs, const int2* work, int nwork, int grid,
                   const double2* tw1024, hipStream_t st) {
  QS_ADJ_SWITCH(nfix, launch_acol_n, amode, runs, work, nwork, grid, tw1024, st);
}


// ===========================================================================
// Reductions and the loss (small kernels, not on the step's critical path).
// ===========================================================================
namespace {
__global__ __launch_bounds__(256) void k_row_sums(const double* arow, int Ny, double* out) {
  __shared__ double red[256];
  const int n = blockIdx.x, t = threadIdx.x;
  double s = 0.0;
  for (int y = t; y < Ny; y += 256) s += arow[(size_t)n * Ny + y];
  red[t] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  if (t == 0) out[n] = red[0];
}

// Loss of a vorticity spectrum and its gradient.  With B the C2R/(Nx Ny) of the
// A10 semantics and F the R2C, F(B(G)) = P(G): the Hermitian part of columns 0
// and Nx/2 (real at the self-conjugate bins), G elsewhere.  Parseval gives
//   1/2 mean(B(G)^2) = sum c |P(G)|^2 / (2 NN^2),   B^T(B(G)/NN) = c P(G) / NN^2
// with c = 1 on columns 0 and Nx/2 and 2 elsewhere.  kind 0: G = W;  kind 1:
// G = i ky psi and -i kx psi (psi = W/k^2), chained through the conjugate
// multipliers.  partial[block] = the block's sum of c |P|^2 (fixed order).
__global__ __launch_bounds__(256) void k_loss(int Nx, int Ny, double dkx, double dky, int kind, const double2* W,
                                              int pin, double2* Wbar, int pout, double* partial) {
  __shared__ double red[256];
  const int Nxh = Nx / 2 + 1, t = threadIdx.x;
  const long e = (long)blockIdx.x * 256 + t;
  double contrib = 0.0;
  if (e < (long)Nxh * Ny) {
    const int j = (int)(e / Nxh), i = (int)(e % Nxh);
    const bool edge = i == 0 || 2 * i == Nx;
    const double c = edge ? 1.0 : 2.0;
    const double nn = (double)Nx * Ny, s = c / (nn * nn);
    auto kyof = [&](int jj) { return dky * (jj < Ny / 2 + 1 ? jj : jj - Ny); };
    const double kx = dkx * i;
    // G_kind(jj) at column i: W, i ky psi, or -i kx psi
    auto G = [&](int which, int jj) {
      const double2 w = W[(size_t)jj * pin + i];
      if (which == 0) return w;
      const double ky = kyof(jj), k2 = kx * kx + ky * ky;
      const double ip = (i == 0 && jj == 0) ? 0.0 : 1.0 / k2;
      if (which == 1) return make_double2(-ky * ip * w.y, ky * ip * w.x);
      return make_double2(kx * ip * w.y, -kx * ip * w.x);
    };
    auto P = [&](int which) {
      const double2 g0 = G(which, j);
      if (!edge) return g0;
      if (j == 0 || 2 * j == Ny) return make_double2(g0.x, 0.0);
      const double2 g1 = G(which, (Ny - j) % Ny);
      return make_double2(0.5 * (g0.x + g1.x), 0.5 * (g0.y - g1.y));
    };
    double2 wb;
    if (kind == 0) {
      const double2 p = P(0);
      contrib = c * (p.x * p.x + p.y * p.y);
      wb = make_double2(s * p.x, s * p.y);
    } else {
      const double2 pu = P(1), pv = P(2);
      contrib = c * (pu.x * pu.x + pu.y * pu.y + pv.x * pv.x + pv.y * pv.y);
      const double ky = kyof(j), k2 = kx * kx + ky * ky;
      const double ip = (i == 0 && j == 0) ? 0.0 : 1.0 / k2;
      // conj(i ky ip) pu + conj(-i kx ip) pv = -i ky ip pu + i kx ip pv
      wb = make_double2(s * (ky * ip * pu.y - kx * ip * pv.y), s * (-ky * ip * pu.x + kx * ip * pv.x));
    }
    Wbar[(size_t)j * pout + i] = wb;
  }
  red[t] = contrib;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  if (t == 0) partial[blockIdx.x] = red[0];
}
}  // namespace

int row_sums_launch(const double* arow, int nsteps, int Ny, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_row_sums, dim3(nsteps), dim3(256), 0, st, arow, Ny, out);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

int loss_launch(const Mesh* m, int kind, const double2* W, int pin, double* loss, double2* Wbar, int pout) {
  const long n = (long)m->Nxh * m->Ny;
  const int blocks = (int)((n + 255) / 256);
  const hipStream_t st = stream_of(m);
  double* partial = m->phys;   // Nx*Ny >= blocks
  hipLaunchKernelGGL(k_loss, dim3(blocks), dim3(256), 0, st, m->Nx, m->Ny, 2 * M_PI / m->Lx, 2 * M_PI / m->Ly, kind,
                     W, pin, Wbar, pout, partial);
  QS_CHECK_LAUNCH();
  std::vector<double> h(blocks);
  QS_HIP(hipMemcpyAsync(h.data(), partial, sizeof(double) * blocks, hipMemcpyDeviceToHost, st));
  QS_HIP(hipStreamSynchronize(st));
  double s = 0.0;
  for (double v : h) s += v;
  const double nn = (double)m->Nx * m->Ny;
  *loss = s / (2.0 * nn * nn);
  return QS_OK;
}

}  // namespace qs
