// qs_fused_alt.hip -- the IFRK2 row and column kernels with the non-default
// lane layouts of the per-wave FFT, for the ensemble autotuner.
//
// Elements per lane E set the first-pass radix (a radix-E DFT in registers),
// the later Stockham radix (min(E, 16)), the rows per row job (NV = 64 E / N)
// and the column tile width (4 NV columns).  Fewer elements per lane give a
// run more, smaller jobs (a lone small run fills more of the device); more
// give fewer, larger ones with less LDS traffic per point (better for
// ensembles that fill the device anyway).  Results differ from the default
// layout's by rounding only (another FFT factorisation).
#include "qs_fused_kern.hpp"

namespace qs {
namespace {

template <int NFIX, int EL>
int alt_col(int mode, const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw, hipStream_t st) {
  switch (mode) {
    case MODE_PRO:
      hipLaunchKernelGGL((k_col<NFIX, MODE_S2, EL>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw, 1);
      break;
    case MODE_S1:
      hipLaunchKernelGGL((k_col<NFIX, MODE_S1, EL>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw, 0);
      break;
    case MODE_S2:
      hipLaunchKernelGGL((k_col<NFIX, MODE_S2, EL>), dim3(grid), dim3(256), 0, st, runs, work, nwork, tw, 0);
      break;
    default: return fail(QS_ERR_UNSUPPORTED_SIZE, "column mode %d with lane layout %d at side %d", mode, EL, NFIX);
  }
  QS_CHECK_LAUNCH();
  return QS_OK;
}

// (side, elements per lane) of the non-default layouts.
#define QS_ALT_LAYOUTS(X) X(64, 4) X(64, 8) X(128, 4) X(128, 8) X(256, 4) X(256, 16) X(512, 8)

}  // namespace

int lane_layouts(int n, int* el) {
  int k = 0;
  el[k++] = wave_e(n);
#define QS_ALT_LIST(NN, EE) \
  if (n == NN) el[k++] = EE;
  QS_ALT_LAYOUTS(QS_ALT_LIST)
#undef QS_ALT_LIST
  return k;
}

int launch_fused_row_l(int nfix, int el, const RunDesc* runs, const int2* work, int nwork, int grid,
                       const double2* tw1024, hipStream_t st, int fin) {
  if (el == 0 || el == wave_e(nfix)) return launch_fused_row(nfix, runs, work, nwork, grid, tw1024, st, fin);
#define QS_ALT_ROW(NN, EE) \
  if (nfix == NN && el == EE) return launch_row_n<NN, EE>(runs, work, nwork, grid, tw1024, st, fin);
  QS_ALT_LAYOUTS(QS_ALT_ROW)
#undef QS_ALT_ROW
  return fail(QS_ERR_UNSUPPORTED_SIZE, "row kernel: lane layout %d at side %d", el, nfix);
}

int launch_fused_col_l(int nfix, int el, int mode, const RunDesc* runs, const int2* work, int nwork, int grid,
                       const double2* tw1024, hipStream_t st) {
  if (el == 0 || el == wave_e(nfix)) return launch_fused_col(nfix, mode, runs, work, nwork, grid, tw1024, st);
#define QS_ALT_COL(NN, EE) \
  if (nfix == NN && el == EE) return alt_col<NN, EE>(mode, runs, work, nwork, grid, tw1024, st);
  QS_ALT_LAYOUTS(QS_ALT_COL)
#undef QS_ALT_COL
  return fail(QS_ERR_UNSUPPORTED_SIZE, "column kernel: lane layout %d at side %d", el, nfix);
}

}  // namespace qs
