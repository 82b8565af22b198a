// qs_fused_kern.hpp -- the fused right-hand-side kernels (row job, column job,
// their kernels and launchers), templated on the mesh side and on the lane
// layout of the per-wave FFT (elements per lane EL, default wave_e(N)).
// Included by qs_fused.hip (the default layouts, every mode and size) and by
// qs_fused_alt.hip (the other layouts the ensemble autotuner may choose).
#pragma once

#include "qs_fused_dev.hpp"

namespace qs {
namespace {

// Row job layout (one job = NV = WaveLay<N>::NV rows, 6 compute waves, 6 LDS
// buffers):
//  1. the 7 mixed rows are loaded cooperatively (coalesced) into a staging
//     view of the buffers (stage_rows, or the loader waves' RowLoader);
//  2. waves 0-4 build their two-for-one C2R inputs from the staging area,
//     pairing fields of like magnitude: (w, dy u + dx v), (u, v), (r1, r2),
//     (L1, L2), (dx u, -);
//  3. inverse x FFTs (per wave) -> physical fields in buffers 0-4;
//  4. pointwise products (all threads) -> buffers 0-5 as R2C pairs
//     (A, B), (-u r1, -v r1), (-u r2, -v r2), (P11, P12), (P21, -), (loc1, loc2);
//  5. forward x FFTs (per wave), spectra back into the buffers;
//  6. cooperative unpack + spectral combination -> F0..F6 (coalesced stores).
// The part of a row job after staging: phases 3-6, run by threads 0..383.
// w00 = *a.w00; alu: the step's uniform activity (*a.alpha_p, or a.alpha).
// Block barriers: kRowBodyBarriers (two in inverse_fields, one with EARLY;
// one after the products, one after the forward FFTs).
constexpr int kRowBodyBarriers = 4;
// OPQ: lane offsets from an opaque thread index (job_tid), recomputed per job:
// the merged kernels need it (their several sizes' job code in one function
// spills otherwise); the fixed-size kernels keep the precomputed offsets
// (faster there).
// Pointwise products of a row job, in place (friction/Qsolver.cuh:63-164,
// 206-300): points p0, p0 + pstep, ... of the job's NV rows; physical fields
// in wave buffers 0-4, the R2C pairs out to buffers 0-5.
template <int N, int EL>
__device__ __forceinline__ void row_products(const RunDesc& a, int row0, double2* sm, double alu, int p0, int pstep) {
  using L = WaveLay<N, EL>;
  constexpr int NV = L::NV, VS = L::VS;
  for (int p = p0; p < NV * N; p += pstep) {
    const int gg = p / N, x = p % N;
    const int o = gg * VS + swz(x);
    const int yy = row0 + gg;
    const double2 f0 = sm[0 * kBuf + o], f1 = sm[1 * kBuf + o], f2 = sm[2 * kBuf + o],
                  f3 = sm[3 * kBuf + o], f4 = sm[4 * kBuf + o];
    const double w = f0.x, uyvx = f0.y, u = f1.x, v = f1.y, r1 = f2.x, r2 = f2.y;
    const double L1 = f3.x, L2 = f3.y, ux = f4.x;
    const double al = a.Ra ? gload(a.Ra + (size_t)(yy < a.Ny ? yy : a.Ny - 1) * N + x) : alu;
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
    sm[0 * kBuf + o] = make_double2(pA, pB);
    sm[1 * kBuf + o] = make_double2(-1.0 * u * r1, -1.0 * v * r1);
    sm[2 * kBuf + o] = make_double2(-1.0 * u * r2, -1.0 * v * r2);
    sm[3 * kBuf + o] = make_double2(p11, p12);
    sm[4 * kBuf + o] = make_double2(p21, 0.0);
    sm[5 * kBuf + o] = make_double2(loc1, loc2);
  }
}

// NOWN, EARLY: the staging layout of inverse_fields (EARLY: one barrier
// fewer).  PT: threads sharing the products (384, or 512 with the loader
// waves, which call row_products themselves between the same barriers).
template <int N, int EL = wave_e(N), bool OPQ = false, int NOWN = 0, int PT = 384, bool EARLY = false, class TW>
__device__ __forceinline__ void row_body(const RunDesc& a, int row0, double2* sm, const TW& tw, double w00,
                                         double alu, int kout) {
  using L = WaveLay<N, EL>;
  constexpr int NV = L::NV;             // rows of this job (= vectors per wave FFT)
  constexpr int LPC = L::LPC, E = L::E;
  constexpr int VS = L::VS;             // padded vector stride inside a wave buffer
  const int tid = OPQ ? job_tid() : (int)threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int lp = lane / NV, g = lane % NV;     // lane = lp * NV + g
  double2* vb = sm + wave * kBuf + g * VS;     // this lane's vector in its wave buffer

  double2 v[16];
  inverse_fields<N, EL, OPQ, NOWN, EARLY>(a, row0, sm, tw, v, w00);

  // ---- 4. pointwise products (the loader waves take their share: PT threads)
  row_products<N, EL>(a, row0, sm, alu, tid, PT);
  __syncthreads();

  // ---- 5. forward x FFTs ------------------------------------------------------
#pragma unroll
  for (int m = 0; m < E; ++m) v[m] = vb[swz(lp + LPC * m)];
  wave_sync();
  wave_fft<N, -1, EL>(v, vb, lp, tw);
#pragma unroll
  for (int m = 0; m < E; ++m) vb[swz(lp + LPC * m)] = v[m];
  __syncthreads();

  // ---- 6. unpack the two-for-one spectra and form F0..F6 ---------------------
  // (bins k < kout: all H, or those a final column stage reads)
  const double sE = a.s_ErRe;
  for (int e = tid; e < NV * kout; e += 384) {
    const int vec = e / kout, k = e % kout;
    const int y = row0 + vec;
    if (y >= a.Ny) continue;
    constexpr bool POW2 = (N & (N - 1)) == 0;
    const int o = vec * VS + swz(k), oc = vec * VS + swz(POW2 ? ((N - k) & (N - 1)) : (k == 0 ? 0 : N - k));
    double2 Sp[6][2];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const double2 zk = sm[q * kBuf + o], zc = cconj(sm[q * kBuf + oc]);
      Sp[q][0] = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
      const double2 d = csub(zk, zc);
      Sp[q][1] = make_double2(0.5 * d.y, -0.5 * d.x);
    }
    const double2 A = Sp[0][0], B = Sp[0][1], U1 = Sp[1][0], V1 = Sp[1][1], U2 = Sp[2][0], V2 = Sp[2][1];
    const double2 P11 = Sp[3][0], P12 = Sp[3][1], P21 = Sp[4][0], Lc1 = Sp[5][0], Lc2 = Sp[5][1];
    const double kx = a.dkx * k;
    const double kxa = (2 * k == N) ? 0.0 : kx;      // projected x-derivative of the advection
    const double c12 = -kx * kx * sE, t11 = 2.0 * kx * sE;
    const size_t ro = (size_t)y * a.pitch + k;
    gstore(a.mix[0] + ro, make_double2(-kxa * A.y + c12 * P12.x, kxa * A.x + c12 * P12.y));  // i kx A - kx^2 s P12
    gstore(a.mix[1] + ro, make_double2(B.x + t11 * P11.y, B.y - t11 * P11.x));             // B - 2 i kx s P11
    gstore(a.mix[2] + ro, make_double2(sE * P21.x, sE * P21.y));                             // s P21
    gstore(a.mix[3] + ro, make_double2(Lc1.x - kx * U1.y, Lc1.y + kx * U1.x));               // loc1 + i kx U1
    gstore(a.mix[4] + ro, V1);
    gstore(a.mix[5] + ro, make_double2(Lc2.x - kx * U2.y, Lc2.y + kx * U2.x));               // loc2 + i kx U2
    gstore(a.mix[6] + ro, V2);
  }
}


// ===========================================================================
// Column job
// ===========================================================================
// Column job: one field (FLD 0 = w, 1 = r1, 2 = r2) on a tile of 4 NV
// columns (ColTile), one column per lane group, FFTs along y per wavefront.
// OPQ: each transform derives its lane offsets from an opaque lane index (for
// the merged kernels, as in row_body).
// Every input tile is staged through the wave buffers (ColTile::load) and
// every output tile leaves through them (ColTile::store).
template <int NY, int MODE, int EL = wave_e(NY), bool OPQ = false, class TW>
__device__ __forceinline__ void col_job(const RunDesc& a, const int FLD, int i0, double2* sm, const TW& tw,
                                        bool pro_arg) {
  // prologue: final-stage kernels only (see is_prologue)
  const bool pro = (MODE == MODE_S2 || MODE == MODE_R4 || MODE == MODE_E) && pro_arg;
  constexpr int LPC = WaveLay<NY, EL>::LPC, E = WaveLay<NY, EL>::E;
  ColTile<NY, EL> T(a, sm, i0);
  const int lp = T.lp;
  const int i = i0 + (T.tid >> 6) * ColTile<NY, EL>::CW + T.g;
  const bool iok = i < a.Nxh;
  const int ic = iok ? i : 0;
  double2* vb = T.vb;
  // a wave whose columns all lie beyond Nxh (the partial last tile) takes part
  // in the tile loads and stores only: its transforms are skipped (their
  // results are never stored).  (Stage-1 modes with one column per wave run
  // such tiles as col_job_split; the kernels with opaque lane indices, which
  // spill with the branch, run the idle transforms.)
  constexpr bool SPLIT1 = (MODE == MODE_S1 || MODE == MODE_R1) && WaveLay<NY, EL>::NV == 1;
  const bool widle = !SPLIT1 && !OPQ &&
                     __builtin_amdgcn_readfirstlane(i0 + (T.tid >> 6) * ColTile<NY, EL>::CW) >= a.Nxh;
  auto fwd = [&](double2 (&x)[16]) {
    if (!widle) wave_fft<NY, -1, EL>(x, vb, OPQ ? opq(lp) : lp, tw);
  };
  auto inv = [&](double2 (&x)[16]) {
    if (!widle) wave_fft<NY, +1, EL>(x, vb, OPQ ? opq(lp) : lp, tw);
  };
  const double kx = a.dkx * ic;
  // ky(j) of the mesh tables (Basic/Mesh.cuh:6-21)
  auto kyj = [&](int j) { return a.dky * (j < NY / 2 + 1 ? j : j - NY); };
  const int nyq_x = a.Nx / 2;
  bool bad = false;
  auto stage_out = [&](double2* __restrict__ dst, const double2 (&v)[16]) { T.stage_out(dst, v); };
  auto tile_store = [&](double2* __restrict__ dst) { T.store(dst); };

  double2 st[16];   // the state of the next stage (curr / new / old)
  if (pro) {
    // straight from HBM in the lane layout (a prologue runs once per restart:
    // no staging registers, which the stage path needs for its own tiles)
    const double2* src = a.old_[FLD] + ic;
#pragma unroll
    for (int m = 0; m < E; ++m) st[m] = gload(src + (unsigned)((lp + LPC * m) * a.pitch));
  } else {
    // ---- N_X from the forward set ------------------------------------------
    //   N_w = F0 + i ky F1 + ky^2 F2,  N_r = F_a + i ky F_b   (in st)
    const bool from_old = MODE == MODE_S1 || MODE == MODE_R1 || MODE == MODE_E;
    const double2* state_in = from_old ? a.old_[FLD] : a.new_[FLD];   // R2, R3: new first
    const int f0 = FLD == 0 ? 0 : (FLD == 1 ? 3 : 5);
    T.stage_in(a.mix[f0], st);
    fwd(st);
    {
      double2 v[16];
      T.stage_in(a.mix[f0 + 1], v);
      fwd(v);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double ky = kyj(opq(lp) + LPC * m);
        st[m] = make_double2(st[m].x - ky * v[m].y, st[m].y + ky * v[m].x);
      }
    }
    if (FLD == 0) {
      double2 v[16];
      T.stage_in(a.mix[2], v);
      fwd(v);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double ky = kyj(opq(lp) + LPC * m);
        st[m] = make_double2(st[m].x + ky * ky * v[m].x, st[m].y + ky * ky * v[m].y);
      }
    }
    // ---- stage update (state tiles read and rewritten in LDS) ---------------
    // N_X is in st.  Per element: IF = exp(L dt), IFh = exp(L dt / 2) (IFac).
    const IFac ifc(a, FLD, ic, NY);
    auto keep = [&](int j) {   // dealias mask, Qsolver.cuh:314-322
      return (i < ((a.Nxh - 1) * 2 / 3 + 1)) && ((j < NY / 3 + 1) || ((NY - j) < NY / 3));
    };
#pragma unroll
    for (int m = 0; m < E; ++m) {   // real part of N_w at the self-conjugate bins
      const int j = opq(lp) + LPC * m;
      if (FLD == 0 && (i == 0 || i == nyq_x) && (j == 0 || 2 * j == NY)) st[m].y = 0.0;
    }
    // the row factors of exp(L dt) for S1 / E, fetched with the state tile (one
    // memory round trip for both instead of a second one in the update)
    constexpr bool PRE_IF = MODE == MODE_S1 || MODE == MODE_E;
    double fyr[16];
    if (PRE_IF) {
#pragma unroll
      for (int m = 0; m < E; ++m) fyr[m] = ifc.row(opq(lp) + LPC * m);
    }
    // the state tile, staged in the wave buffers and updated in place
    T.load(state_in);
    if (MODE == MODE_S1 || MODE == MODE_S2 || MODE == MODE_R1 || MODE == MODE_R4 || MODE == MODE_E) {
      // one state tile: S1 / R1 / E read old, S2 / R4 read new
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int j = opq(lp) + LPC * m;
        double2* sl = vb + swz(j);
        const double2 in = *sl;
        const double2 n = st[m];
        const double2 an = make_double2(n.x * a.dt, n.y * a.dt);
        if (MODE == MODE_S1) {
          const double IF = PRE_IF ? ifc.fx * fyr[m] : ifc.IF(j);
          const double2 old = in;
          const double2 nw0 = make_double2(old.x * IF, old.y * IF);              // IFRK2.cuh:16-26
          const double h = 1.0 / 2.0 * IF;
          *sl = make_double2(nw0.x + h * an.x, nw0.y + h * an.y);                // IFRK2.cuh:28-39
          st[m] = make_double2((old.x + an.x) * IF, (old.y + an.y) * IF);      // curr (IFh = IF)
        } else if (MODE == MODE_R1) {
          const double IF = ifc.IF(j), IFh = ifc.IFh(j);
          const double2 old = in;
          const double2 nw0 = make_double2(old.x * IF, old.y * IF);              // RK4.cuh:12-22
          const double h = 1.0 / 6.0 * IF;
          *sl = make_double2(nw0.x + h * an.x, nw0.y + h * an.y);                // RK4.cuh:25-36
          st[m] = make_double2((old.x + an.x / 2.0) * IFh, (old.y + an.y / 2.0) * IFh);
        } else if (MODE == MODE_E) {
          const double IF = PRE_IF ? ifc.fx * fyr[m] : ifc.IF(j);
          const double2 old = in;                                                // IFE.cuh:15-25
          double2 r = make_double2((old.x + an.x) * IF, (old.y + an.y) * IF);
          const double cut = keep(j) ? 1.0 : 0.0;
          r = make_double2(r.x * cut, r.y * cut);
          if (iok) bad |= !isfinite(r.x) || !isfinite(r.y);
          *sl = r;
          st[m] = r;
        } else {
          const double2 nw = in;
          double2 r = MODE == MODE_S2
                          ? make_double2(nw.x + 1.0 / 2.0 * an.x, nw.y + 1.0 / 2.0 * an.y)   // IFRK2.cuh:41-51
                          : make_double2(nw.x + 1.0 / 6.0 * an.x, nw.y + 1.0 / 6.0 * an.y);  // RK4.cuh:63-74
          const double cut = keep(j) ? 1.0 : 0.0;
          r = make_double2(r.x * cut, r.y * cut);
          if (iok) bad |= !isfinite(r.x) || !isfinite(r.y);
          *sl = r;
          st[m] = r;
        }
      }
      // S1, R1 -> new; S2, R4, E -> the next state over old
      tile_store((MODE == MODE_S1 || MODE == MODE_R1) ? a.new_[FLD] : a.old_[FLD]);
    } else {
      // R2, R3: new += 1/3 IFh dt N (new tile), then curr from old (old tile)
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int j = opq(lp) + LPC * m;
        double2* sl = vb + swz(j);
        const double2 nw = *sl;
        const double IFh = ifc.IFh(j);
        const double2 bn = make_double2(st[m].x * a.dt, st[m].y * a.dt);
        const double h = 1.0 / 3.0 * IFh;                                          // RK4.cuh:37-61
        *sl = make_double2(nw.x + h * bn.x, nw.y + h * bn.y);
      }
      tile_store(a.new_[FLD]);
      T.load(a.old_[FLD]);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int j = opq(lp) + LPC * m;
        const double2 old = vb[swz(j)];
        const double IFh = ifc.IFh(j);
        const double2 bn = make_double2(st[m].x * a.dt, st[m].y * a.dt);
        if (MODE == MODE_R2) {
          st[m] = make_double2(old.x * IFh + bn.x / 2.0, old.y * IFh + bn.y / 2.0);
        } else {
          const double IF = ifc.IF(j);
          st[m] = make_double2(old.x * IF + bn.x * IFh, old.y * IF + bn.y * IFh);
        }
      }
    }
  }

  // ---- inverse set of the next stage from st ----------------------------------
  if (FLD == 0) {
    // M0 = W, M1 = psi = W/k^2, M2 = i ky psi   (vel_func, QFldfuncs.cuh:204-224)
    if (i == 0 && lp == 0 && a.w00o) *(QS_GLOBAL double*)a.w00o = st[0].x;   // element (j=0, i=0)
    {
      double2 v[16];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = st[m];
      inv(v);
      stage_out(a.mout[0], v);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {   // st <- psi
      const int j = opq(lp) + LPC * m;
      const double ky = kyj(j);
      const double r = recip(kx * kx + ky * ky);   // one reciprocal for both parts (k^2 = 0: not selected)
      st[m] = (i == 0 && j == 0) || !iok ? make_double2(0.0, 0.0) : make_double2(st[m].x * r, st[m].y * r);
    }
    {
      double2 v[16];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = st[m];
      inv(v);
      stage_out(a.mout[1], v);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {   // st <- i ky psi
      const double ky = kyj(opq(lp) + LPC * m);
      st[m] = make_double2(-ky * st[m].y, ky * st[m].x);
    }
    inv(st);
    stage_out(a.mout[2], st);
  } else {
    const int m0 = FLD == 1 ? 3 : 5;
    {
      double2 v[16];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = st[m];
      inv(v);
      stage_out(a.mout[m0], v);
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {   // st <- -ky^2 R
      const double ky = kyj(opq(lp) + LPC * m);
      st[m] = make_double2(-ky * ky * st[m].x, -ky * ky * st[m].y);
    }
    inv(st);
    stage_out(a.mout[m0 + 1], st);
  }
  if ((MODE == MODE_S2 || MODE == MODE_R4 || MODE == MODE_E) && bad) atomicOr(a.flag, 1);
}


// ===========================================================================
// Kernels: one run (grid over jobs) and an ensemble (work list over runs).
// ===========================================================================

// Every run of a row kernel's work list has Nx = NFIX (its x FFTs; the row
// count is a runtime value): six compute waves and two loader waves
// (RowLoader) per block.
constexpr int kRowThreads = 512;

// fin != 0: a final column stage follows, which reads only the x-bins below
// col_kept_cols of the forward set; the others are not formed.
template <int NFIX, int EL = wave_e(NFIX)>
__global__ __launch_bounds__(kRowThreads) void k_row(const RunDesc* __restrict__ runs, const int2* __restrict__ work,
                                                    int nwork, const double2* __restrict__ tw1024, int fin) {
  static_assert(NFIX > 0, "one instantiation per x size");
  constexpr int NT = kRowThreads;
  // Staging: arrays 0 .. NOWN-1 in an area of their own after the six wave
  // buffers (as many as the 160 KB of LDS holds beside them: six, five for
  // N = 16), the others in the wave buffers.
  constexpr int RH = WaveLay<NFIX, EL>::NV * (NFIX / 2 + 1);
  constexpr int FREE = 10240 - 6 * kBuf - kTwq - twt_len<NFIX>() - 2;   // 16-byte entries
  constexpr int NOWN = FREE / RH < 6 ? FREE / RH : 6;
  static_assert(NOWN >= 1, "own staging area");
  // the other arrays in wave buffer 5 where they fit (all sides but 16)
  constexpr bool EARLY = (7 - NOWN) * RH <= kBuf;
  constexpr bool PRE16 = EL == 16 && (NFIX == 256 || NFIX == 1024);   // (512 spills with them)
  __shared__ double2 sm[6 * kBuf + NOWN * RH];
  double2* const late = EARLY ? sm + 5 * kBuf - NOWN * RH : sm;   // array t >= NOWN at late + t RH
  __shared__ double2 twq[kTwq];
  __shared__ double2 twt[twt_len<NFIX>()];   // per-pass twiddles (use_twt)
  __shared__ double rsc[2];                  // the job's (w00, alu), from the loader
  __shared__ int2 rjob;                      // the job's (run, job), from the loader (OWN)
  stage_twiddles(twq, tw1024, threadIdx.x, NT);
  __syncthreads();
  if constexpr (use_twt<NFIX>()) {
    build_twtab<NFIX>(twt, twq, threadIdx.x, NT);
    __syncthreads();
  }
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  constexpr int NV = WaveLay<NFIX, EL>::NV;
  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) >= 6) {   // wave-uniform
    // Loader waves.  Per job they pass exactly the compute waves' block
    // barriers: staging ready, the kRowBodyBarriers of row_body, job done.
    const int lt = threadIdx.x - 384;
    RowLoader<NFIX, EL> rl;
    if (L < nwork) {
      const int2 it = work[L];
      rl.issue(runs[it.x], it.y * NV, lt);
    }
    // Arrays 0 .. NOWN-1 of the next job (with its w00, alu and (run, job))
    // are committed as soon as this job's inputs are built (row_body's first
    // barrier), while its transforms run; only the others wait for the job's
    // end.  The compute waves read what they need of the handoff before that
    // barrier.  The next job's loads are issued as this one starts.
    // They also take a quarter of the products (row_body's third phase): by
    // then they hold no load in flight.
    int2 itc = make_int2(0, 0);
    if (L < nwork) {
      rl.commit_part(sm + 6 * kBuf, rsc, lt, 0, NOWN);
      itc = work[L];
      if (lt == 0) rjob = itc;
    }
    for (int w = L; w < nwork; w += gridDim.x) {
      rl.commit_part(late, rsc, lt, NOWN, 7);   // the rest into the wave buffers
      __syncthreads();   // staging ready
      const int wn = w + gridDim.x;
      int2 itn = make_int2(0, 0);
      if (wn < nwork) {
        itn = work[wn];
        rl.issue(runs[itn.x], itn.y * NV, lt);
      }
      __syncthreads();   // own staging area free (row_body's first barrier)
      if (wn < nwork) {
        rl.commit_part(sm + 6 * kBuf, rsc, lt, 0, NOWN);
        if (lt == 0) rjob = itn;
      }
      if (!EARLY) __syncthreads();   // inverse transforms done (row_body's second barrier)
      {
        const RunDesc& a = runs[itc.x];
        const double alu = a.alpha_p ? gload(a.alpha_p) : a.alpha;
        row_products<NFIX, EL>(a, itc.y * NV, sm, alu, threadIdx.x, kRowThreads);
      }
      __syncthreads();   // products done
      __syncthreads();   // forward transforms done
      __syncthreads();   // job done
      itc = itn;
    }
    static_assert(kRowBodyBarriers == 4, "the loader waves pass row_body's four barriers (three when EARLY)");
  } else {
    // sides 2^a with 16 elements per lane: the radix-16 pass factors in
    // registers for the whole kernel (where they fit without spilling)
    const auto tw = [&]() {
      if constexpr (PRE16) return TwNP<NFIX>(twq, (threadIdx.x & 63) / NV);
      else return fused_tw<NFIX>(twq, twt);
    }();
    for (int w = L; w < nwork; w += gridDim.x) {
      __syncthreads();   // staging ready
      const int2 it = rjob;   // (run, job)
      const RunDesc& a = runs[it.x];
      row_body<NFIX, EL, false, NOWN, kRowThreads, EARLY>(a, it.y * NV, sm, tw, rsc[0], rsc[1],
                                                          fin ? col_kept_cols(a) : NFIX / 2 + 1);
      __syncthreads();   // job done: wave buffers free for the next staging
    }
  }
}

// Job of stage 1 (IFRK2 S1, RK4 R1) on a partial last tile with one valid
// column i (every side with one vector per wave: Nxh = nfull * tc + 1), its
// transforms spread over the block's waves instead of run by one wave with
// three idle: wave q < nf transforms forward set array f0 + q (vorticity:
// F0, F1, F2; r: F_a, F_b), wave 0 combines them into N_X, updates the state
// column and hands the next stage's state to the others through the idle
// wave's buffer, and wave q < nf then forms and transforms inverse set array
// m0 + q.  The job's chain is two transforms instead of six (four), so that
// it no longer outlasts the full tiles' jobs.  Columns move between HBM and
// registers directly (one column, 16 B per row).  Arithmetic per element as
// in col_job.
template <int NY, int MODE, int EL, class TW>
__device__ __forceinline__ void col_job_split(const RunDesc& a, const int FLD, int i, double2* sm, const TW& tw) {
  static_assert(MODE == MODE_S1 || MODE == MODE_R1, "stage-1 modes");
  using L = WaveLay<NY, EL>;
  static_assert(L::NV == 1, "one column per wave");
  constexpr int LPC = L::LPC, E = L::E, WS = ColTile<NY, EL>::WS;
  const int tid = job_tid();
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lp = tid & 63;
  double2* vb = sm + wave * WS;
  double2* vst = sm + 3 * WS;   // the fourth wave's buffer: the updated state
  const int nf = FLD == 0 ? 3 : 2;
  const int f0 = FLD == 0 ? 0 : (FLD == 1 ? 3 : 5);
  const double kx = a.dkx * i;
  auto kyj = [&](int j) { return a.dky * (j < NY / 2 + 1 ? j : j - NY); };
  auto col_ld = [&](const double2* base, double2 (&v)[16]) {
#pragma unroll
    for (int m = 0; m < E; ++m) v[m] = gload(base + i + (unsigned)((opq(lp) + LPC * m) * a.pitch));
  };
  auto col_st = [&](double2* base, const double2 (&v)[16]) {
#pragma unroll
    for (int m = 0; m < E; ++m) gstore(base + i + (unsigned)((opq(lp) + LPC * m) * a.pitch), v[m]);
  };
  // ---- forward transforms, one array per wave --------------------------------
  double2 st[16];
  if (wave < nf) {
    col_ld(a.mix[f0 + wave], st);
    wave_fft<NY, -1, EL>(st, vb, lp, tw);
    if (wave > 0) {
#pragma unroll
      for (int m = 0; m < E; ++m) vb[swz(opq(lp) + LPC * m)] = st[m];
    }
  }
  __syncthreads();
  if (wave == 0) {
    // N_w = F0 + i ky F1 + ky^2 F2,  N_r = F_a + i ky F_b
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const double2 v = sm[WS + swz(opq(lp) + LPC * m)];
      const double ky = kyj(opq(lp) + LPC * m);
      st[m] = make_double2(st[m].x - ky * v.y, st[m].y + ky * v.x);
    }
    if (FLD == 0) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const double2 v = sm[2 * WS + swz(opq(lp) + LPC * m)];
        const double ky = kyj(opq(lp) + LPC * m);
        st[m] = make_double2(st[m].x + ky * ky * v.x, st[m].y + ky * ky * v.y);
      }
    }
    const IFac ifc(a, FLD, i, NY);
    const int nyq_x = a.Nx / 2;
#pragma unroll
    for (int m = 0; m < E; ++m) {   // real part of N_w at the self-conjugate bins
      const int j = opq(lp) + LPC * m;
      if (FLD == 0 && (i == 0 || i == nyq_x) && (j == 0 || 2 * j == NY)) st[m].y = 0.0;
    }
    double2 in[16];
    col_ld(a.old_[FLD], in);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int j = opq(lp) + LPC * m;
      const double2 n = st[m];
      const double2 an = make_double2(n.x * a.dt, n.y * a.dt);
      const double2 old = in[m];
      if (MODE == MODE_S1) {
        const double IF = ifc.IF(j);
        const double2 nw0 = make_double2(old.x * IF, old.y * IF);              // IFRK2.cuh:16-26
        const double h = 1.0 / 2.0 * IF;
        in[m] = make_double2(nw0.x + h * an.x, nw0.y + h * an.y);              // IFRK2.cuh:28-39
        st[m] = make_double2((old.x + an.x) * IF, (old.y + an.y) * IF);      // curr (IFh = IF)
      } else {
        const double IF = ifc.IF(j), IFh = ifc.IFh(j);
        const double2 nw0 = make_double2(old.x * IF, old.y * IF);              // RK4.cuh:12-22
        const double h = 1.0 / 6.0 * IF;
        in[m] = make_double2(nw0.x + h * an.x, nw0.y + h * an.y);              // RK4.cuh:25-36
        st[m] = make_double2((old.x + an.x / 2.0) * IFh, (old.y + an.y / 2.0) * IFh);
      }
      vst[swz(j)] = st[m];
    }
    col_st(a.new_[FLD], in);
  }
  __syncthreads();
  // ---- inverse set of the next stage, one array per wave ----------------------
  if (wave < nf) {
    if (wave > 0) {
#pragma unroll
      for (int m = 0; m < E; ++m) st[m] = vst[swz(opq(lp) + LPC * m)];
    }
    if (FLD == 0) {
      // M0 = W, M1 = psi = W/k^2, M2 = i ky psi   (vel_func, QFldfuncs.cuh:204-224)
      if (wave == 0 && i == 0 && lp == 0 && a.w00o) *(QS_GLOBAL double*)a.w00o = st[0].x;
      if (wave > 0) {
#pragma unroll
        for (int m = 0; m < E; ++m) {   // st <- psi
          const int j = opq(lp) + LPC * m;
          const double ky = kyj(j);
          const double r = recip(kx * kx + ky * ky);
          st[m] = (i == 0 && j == 0) ? make_double2(0.0, 0.0) : make_double2(st[m].x * r, st[m].y * r);
        }
      }
      if (wave == 2) {
#pragma unroll
        for (int m = 0; m < E; ++m) {   // st <- i ky psi
          const double ky = kyj(opq(lp) + LPC * m);
          st[m] = make_double2(-ky * st[m].y, ky * st[m].x);
        }
      }
      wave_fft<NY, +1, EL>(st, vb, lp, tw);
      col_st(a.mout[wave], st);
    } else {
      const int m0 = FLD == 1 ? 3 : 5;
      if (wave == 1) {
#pragma unroll
        for (int m = 0; m < E; ++m) {   // st <- -ky^2 R
          const double ky = kyj(opq(lp) + LPC * m);
          st[m] = make_double2(-ky * ky * st[m].x, -ky * ky * st[m].y);
        }
      }
      wave_fft<NY, +1, EL>(st, vb, lp, tw);
      col_st(a.mout[m0 + wave], st);
    }
  }
}

// Job of a final stage on a tile that starts at or beyond kcut (every column
// masked): the state and the field's inverse-set arrays get zeros, which is
// what the stage computes there from finite values (and all that later
// kernels read).  Coalesced as ColTile::store: thread t writes column
// i0 + t % TC of rows t / TC + q (256 / TC).
template <int NY, int EL = wave_e(NY)>
__device__ __forceinline__ void col_zero_job(const RunDesc& a, int fld, int i0) {
  using T = ColTile<NY, EL>;
  const int t = job_tid();
  const int c = t % T::TC, j0 = t / T::TC, ii = i0 + c;
  if (ii >= a.Nxh) return;
  const unsigned base = (unsigned)(j0 * a.pitch + ii);
  const unsigned stride = (unsigned)((256 / T::TC) * a.pitch);
  const int m0 = fld == 0 ? 0 : (fld == 1 ? 3 : 5), nm = fld == 0 ? 3 : 2;
  const double2 z = make_double2(0.0, 0.0);
#pragma unroll
  for (int q = 0; q < T::E; ++q) gstore(a.old_[fld] + (base + q * stride), z);
  for (int k = 0; k < nm; ++k) {
    double2* dst = a.mout[m0 + k];
#pragma unroll
    for (int q = 0; q < T::E; ++q) gstore(dst + (base + q * stride), z);
  }
}

// MIX: a job of the merged ensemble kernels, whose work list mixes runs (the
// XCD placement, made for one run's jobs, and the split partial-tile job are
// left out: one kernel holds every side's job code).
template <int NY, int MODE, int EL = wave_e(NY), bool OPQ = false, bool MIX = false, class TW>
__device__ __forceinline__ void col_dispatch_n(const RunDesc& a, int job, double2* sm, const TW& tw, bool pro) {
  constexpr int tc = 4 * WaveLay<NY, EL>::NV;
  const int ntiles = (a.Nxh + tc - 1) / tc;
  int fld, tile;
  const int nfull = a.Nxh / tc;
  if ((MODE == MODE_S2 || MODE == MODE_R4 || MODE == MODE_E) && !pro) {
    const int nk = (col_kcut(a) + tc - 1) / tc;
    const int nkeep = nk < ntiles ? nk : ntiles;
    if (MIX || !col_job_place(job, a.Nxh, tc, nkeep < nfull ? nkeep : nfull, fld, tile))
      col_job_map_final(job, ntiles, nkeep, fld, tile);
    if (tile >= nkeep) {
      col_zero_job<NY, EL>(a, fld, tile * tc);
      return;
    }
  } else {
    if (MIX || !col_job_place(job, a.Nxh, tc, nfull, fld, tile)) col_job_map(job, ntiles, fld, tile);
    if constexpr (!MIX && (MODE == MODE_S1 || MODE == MODE_R1) && WaveLay<NY, EL>::NV == 1) {
      if (tile == nfull && a.Nxh == nfull * tc + 1) {   // one valid column
        col_job_split<NY, MODE, EL>(a, fld, tile * tc, sm, tw);
        return;
      }
    }
  }
  col_job<NY, MODE, EL, OPQ>(a, fld, tile * tc, sm, tw, pro);
}

// Every run of a column kernel's work list has Ny = NFIX (its y FFTs; the
// column count is a runtime value).  pro != 0: the prologue of the
// final-stage modes (S2, R4, E).
// Two blocks per CU, so that one block's FFTs run while the other waits on
// its loads, stores and barriers.  Tiles are staged
// through the wave buffers (no DMA buffer), the twiddles are computed from
// the quarter table (no per-pass table), so a block takes 72 KB of LDS, and
// registers are capped at 256 per lane.
template <int NFIX, int MODE, int EL = wave_e(NFIX)>
__global__ __launch_bounds__(256, 2) void k_col(const RunDesc* __restrict__ runs, const int2* __restrict__ work,
                                                int nwork, const double2* __restrict__ tw1024, int pro) {
  __shared__ double2 sm[4 * (kBuf + 1)];
  __shared__ double2 twq[kTwq];
  stage_twiddles(twq, tw1024, threadIdx.x, 256);
  __syncthreads();
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  for (int w = L; w < nwork; w += gridDim.x) {
    const int2 it = work[w];
    col_dispatch_n<NFIX, MODE, EL, (NFIX == 512)>(runs[it.x], it.y, sm, TwN<NFIX>{twq}, pro != 0);
    __syncthreads();
  }
}

template <int NFIX, int EL = wave_e(NFIX)>
int launch_row_n(const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw, hipStream_t st,
                 int fin) {
  hipLaunchKernelGGL((k_row<NFIX, EL>), dim3(grid), dim3(kRowThreads), 0, st, runs, work, nwork, tw, fin);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

template <int NFIX>
int launch_col_n(int mode, const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw,
                 hipStream_t st) {
  switch (mode) {
#define QS_COL_MODE(M, K, P) \
    case M: hipLaunchKernelGGL((k_col<NFIX, K>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw, P); break;
    QS_COL_MODE(MODE_PRO, MODE_S2, 1)
    QS_COL_MODE(MODE_PRO_R4, MODE_R4, 1)
    QS_COL_MODE(MODE_PRO_E, MODE_E, 1)
    QS_COL_MODE(MODE_S1, MODE_S1, 0)
    QS_COL_MODE(MODE_S2, MODE_S2, 0)
    QS_COL_MODE(MODE_R1, MODE_R1, 0)
    QS_COL_MODE(MODE_R2, MODE_R2, 0)
    QS_COL_MODE(MODE_R3, MODE_R3, 0)
    QS_COL_MODE(MODE_R4, MODE_R4, 0)
    QS_COL_MODE(MODE_E, MODE_E, 0)
#undef QS_COL_MODE
    default: return fail(QS_ERR_INVALID_ARG, "column mode %d", mode);
  }
  QS_CHECK_LAUNCH();
  return QS_OK;
}

}  // namespace
}  // namespace qs
