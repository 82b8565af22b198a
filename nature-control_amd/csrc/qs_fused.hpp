// qs_fused.hpp -- descriptors and launchers of the fused IFRK2 kernels.
#pragma once

#include <cmath>

#include "qs_internal.hpp"

namespace qs {

// Column kernel modes: prologue (inverse set of the state), IFRK2 stages 1-2,
// IF-RK4 stages 1-4, IF-Euler.
enum { MODE_PRO = 0, MODE_S1 = 1, MODE_S2 = 2, MODE_R1 = 3, MODE_R2 = 4, MODE_R3 = 5, MODE_R4 = 6, MODE_E = 7,
       MODE_PRO_R4 = 8, MODE_PRO_E = 9 };
// Prologue modes (the inverse set of the state in old_, no stage update) run
// the kernel of the integrator's final stage with its update skipped:
// MODE_PRO that of IFRK2 stage 2, MODE_PRO_R4 that of RK4 stage 4, MODE_PRO_E
// that of IF-Euler.  The inverse set is then computed by the same code from
// the same values as at the end of a step, so a restart from a stored state
// continues the run bit for bit.
inline bool is_prologue(int mode) { return mode == MODE_PRO || mode == MODE_PRO_R4 || mode == MODE_PRO_E; }
// Column kernels of the adjoint step (see qs_adjoint.hip).
enum { AMODE_START = 0, AMODE_B = 1, AMODE_A = 2, AMODE_END = 3 };

// Everything a fused job needs about one run (device-resident array).
struct RunDesc {
  int Nx, Ny, Nxh, pitch;         // pitch: row stride (complex) of state and mixed arrays
  double2* mix[7];                // forward set [Ny][pitch] (row pass writes, column reads);
                                  // adjoint: Fbar (column writes) then Mbar (row adjoint, in place)
  double2* minp[7];               // inverse set read by the row pass
  double2* mout[7];               // inverse set written by the column kernels
  double2* old_[3];               // state w, r1, r2 [Ny][pitch]
  double2* new_[3];               // stage-1 partial sums
  const double* Ra;               // activity field [Ny][Nx] or null (uniform activity)
  const double* alpha_p;          // uniform activity of this step (device scalar) or null (alpha)
  double* w00;                    // W(0,0) of the stage the row pass reads
  double* w00o;                   // W(0,0) of the stage the column kernel writes
  int* flag;                      // non-finite flag
  const double* ify;              // integrating-factor tables (IFac): [4][Ny] rows, then [4][Nxh] columns
  double dkx, dky;                // 2 pi / Lx, 2 pi / Ly: kx(i) = dkx i, ky(j) = dky (j or j - Ny)
  double alpha, lambda, DQ, aQ, cQ, s_ErRe, scale, dt, Re, Rf;
  int advect_w;
  // ---- discrete adjoint -------------------------------------------------
  double2* lam_[3];               // masked dL/du_{n+1} (lambda) / incoming gradient
  double2* acc_[3];               // E (lambda + cbar2): the stage-2 part of dL/du_n
  double2* uout[3];               // dL/du_0 [Ny][pitch] or null
  double* abar;                   // dL/dRa [Ny][Nx] (accumulated) or null
  double* arow;                   // per-row sums of dL/dalpha of this step [Ny] (accumulated) or null
  double* w00bar;                 // per-row parts of dL/dRe W(0,0) [Ny] (row adjoint -> column)
  double2* scratch;               // row adjoint: [grid][6][1024] physical product adjoints
};

// Jobs of one run per launch (row: WaveLay<Nx, el>::NV rows each; column: 3
// fields x tiles of 4 NV columns); el = 0: the default layout wave_e.
int row_jobs(const RunDesc& a, int el = 0);
int col_jobs(const RunDesc& a, int el = 0);

// work[k] = (run index, job index).  grid <= nwork; blocks loop over the list.
// nfix: the size common to all listed runs in the kernel's FFT direction (Nx
// for the row pass, Ny for the column stages): a power of two 16..1024 or
// 3 * 2^a 48..768; any other value is QS_ERR_UNSUPPORTED_SIZE.
// fin != 0: the launch is followed by a final column stage (IFRK2 stage 2, RK4
// stage 4, IF-Euler), so only the forward-set bins that stage reads are formed.
int launch_fused_row(int nfix, const RunDesc* runs, const int2* work, int nwork, int grid,
                     const double2* tw1024, hipStream_t st, int fin = 0);
int launch_fused_col(int nfix, int mode, const RunDesc* runs, const int2* work, int nwork, int grid,
                     const double2* tw1024, hipStream_t st);

// The same kernels with another lane layout of the per-wave FFT: el elements
// per lane (0 or wave_e(nfix): the default).  Other layouts exist for the
// sides 64, 128, 256 and 512 (lane_layouts) and the modes of the IFRK2 step
// (MODE_PRO, MODE_S1, MODE_S2); anything else is QS_ERR_UNSUPPORTED_SIZE.
int launch_fused_row_l(int nfix, int el, const RunDesc* runs, const int2* work, int nwork, int grid,
                       const double2* tw1024, hipStream_t st, int fin = 0);
int launch_fused_col_l(int nfix, int el, int mode, const RunDesc* runs, const int2* work, int nwork, int grid,
                       const double2* tw1024, hipStream_t st);
// Layouts (elements per lane) available for side n, the default first; returns
// their count (at most 4).
int lane_layouts(int n, int* el);

// One kernel per phase over runs of every power-of-two side 16..1024 (mix_side)
// at 16 elements per lane (qs_fused_mix.hip): the ensemble's merged launches.
// Job counts are row_jobs / col_jobs with el = 16; modes MODE_PRO, MODE_S1,
// MODE_S2.
bool mix_side(int n);
int launch_mix_row(const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw1024,
                   hipStream_t st, int fin);
int launch_mix_col(int mode, const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw1024,
                   hipStream_t st);

// Discrete adjoint kernels (qs_adjoint.hip).  nfix > 0 only.
int launch_adj_row(int nfix, const RunDesc* runs, const int2* work, int nwork, int grid,
                   const double2* tw1024, hipStream_t st);
int launch_adj_col(int nfix, int amode, const RunDesc* runs, const int2* work, int nwork, int grid,
                   const double2* tw1024, hipStream_t st);

// out[n] = sum_y arow[n*Ny + y] (fixed-order reduction), n < nsteps.
int row_sums_launch(const double* arow, int nsteps, int Ny, double* out, hipStream_t st);
// Loss of a vorticity spectrum W (row pitch pin) and its gradient (row pitch
// pout); kind 0: 1/2 mean(omega^2), kind 1: 1/2 mean(u^2 + v^2).  Synchronises
// the mesh stream.
int loss_launch(const Mesh* m, int kind, const double2* W, int pin, double* loss, double2* Wbar, int pout);

int wave_fft_test(int n, int sign, const double2* in, double2* out, const double2* tw1024, hipStream_t st);

inline int pitch_of(int Nxh) { return (Nxh + 7) & ~7; }

// The integrating factors exp(L h) of IFac as separable tables, L = c2 (kx^2 +
// ky^2) + c0 (c2 = -1/Re, c0 = -Rf/Re for w; c2 = -D_Q, c0 = a_Q for Q):
//   rows    ify[q Ny + j]           = exp(c2 ky(j)^2 h),
//   columns ify[4 Ny + q Nxh + i]   = exp((c2 kx(i)^2 + c0) h),
// q = 0: w with h = dt, 1: Q with dt, 2: w with dt/2, 3: Q with dt/2.  One
// device routine per table, so that runs built by different paths compute
// identical factors.
__device__ __forceinline__ void ifac_row(int j, int Ny, double dky, double Re, double DQ, double dt, double* ify) {
  const double ky = dky * (j < Ny / 2 + 1 ? j : j - Ny);
  const double cw = -1.0 / Re * ky * ky, cq = -DQ * ky * ky;
  ify[j] = exp(cw * dt);
  ify[Ny + j] = exp(cq * dt);
  ify[2 * Ny + j] = exp(cw * (dt * 0.5));
  ify[3 * Ny + j] = exp(cq * (dt * 0.5));
}
__device__ __forceinline__ void ifac_col(int i, int Nxh, int Ny, double dkx, double Re, double Rf, double DQ,
                                         double aQ, double dt, double* ify) {
  const double kx = dkx * i;
  const double lw = -1.0 / Re * kx * kx + -Rf / Re;
  const double lq = -DQ * kx * kx + aQ;
  double* t = ify + 4 * Ny;
  t[i] = exp(lw * dt);
  t[Nxh + i] = exp(lq * dt);
  t[2 * Nxh + i] = exp(lw * (dt * 0.5));
  t[3 * Nxh + i] = exp(lq * (dt * 0.5));
}
// Length of the tables of an Nxh x Ny half spectrum (doubles).
inline size_t ifac_len(int Nxh, int Ny) { return 4 * ((size_t)Ny + (size_t)Nxh); }
// Fills a solver's tables on stream st.
int ifac_launch(double* ify, int Nxh, int Ny, double dkx, double dky, double Re, double Rf, double DQ, double aQ,
                double dt, hipStream_t st);


}  // namespace qs
