// qs_fused_mix.hip -- the fused IFRK2 kernels of an ensemble as one persistent
// kernel per phase over every size class.
//
// The work list holds (run, job) pairs of runs of any power-of-two side
// 16..1024; each job is dispatched on its run's side (uniform per block) to the
// fixed-size job code of qs_fused_kern.hpp, with 16 elements per lane for
// every side (rows per row job 1024 / N, column tiles of 4096 / N columns).
// One launch per phase covers all classes, so the tail of one class's jobs is
// filled by the jobs of the others, and there is no fork / join between
// classes.  Per job the arithmetic is that of k_row<N, 16> / k_col<N, M, 16>.
#include "qs_fused_kern.hpp"

namespace qs {
namespace {

constexpr int kMixEL = 16;

// The row loader (RowLoader) with the side known at run time: NV = 1024 / N
// rows of H = N/2 + 1 bins, RH = NV H <= 576 staged entries per array (N = 16),
// so five slots per array and loader thread.
struct MixRowLoader {
  static constexpr int S = 5;
  double2 ld[7][S];
  double w00, alu;
  int rh;
  __device__ __forceinline__ void issue(const RunDesc& a, int row0, int lt) {
    const int H = a.Nx / 2 + 1, NV = 64 * kMixEL / a.Nx;
    rh = NV * H;
    w00 = gload(a.w00);
    alu = a.alpha_p ? gload(a.alpha_p) : a.alpha;
    unsigned off[S];
#pragma unroll
    for (int h = 0; h < S; ++h) {
      const int r = lt + 128 * h;
      const int rr = r < rh ? r : rh - 1;
      const int vec = rr / H, k = rr % H;
      const int y = row0 + vec;
      off[h] = (unsigned)((y < a.Ny ? y : a.Ny - 1) * a.pitch + k);
    }
#pragma unroll
    for (int t = 0; t < 7; ++t)
#pragma unroll
      for (int h = 0; h < S; ++h) ld[t][h] = gload(a.minp[t] + off[h]);
  }
  // arrays t0 .. t1-1 at stg + t rh ((w00, alu) into sc with the part from t0 = 0)
  __device__ __forceinline__ void commit_part(double2* stg, double* sc, int lt, int t0, int t1) const {
    if (t0 == 0 && lt == 0) {
      sc[0] = w00;
      sc[1] = alu;
    }
#pragma unroll
    for (int t = 0; t < 7; ++t)
#pragma unroll
      for (int h = 0; h < S; ++h)
        if (t >= t0 && t < t1 && lt + 128 * h < rh) stg[t * rh + lt + 128 * h] = ld[t][h];
  }
};
// Staging: arrays 0-4 in an area of their own after the six wave buffers
// (committed while the previous job's transforms run), arrays 5-6 in the
// buffers (committed at the previous job's end).  rh <= 576 (N = 16).
constexpr int kMixOwn = 5, kMixRh = 64 * 9;
static_assert(7 * kMixRh <= 6 * kBuf, "the staging area of the smallest side fits the wave buffers");

// Row pass over every class: the block layout of k_row (six compute waves, two
// loader waves, the same barriers per job).  fin as in k_row.
__global__ __launch_bounds__(kRowThreads) void k_row_mix(const RunDesc* __restrict__ runs,
                                                        const int2* __restrict__ work, int nwork,
                                                        const double2* __restrict__ tw1024, int fin) {
  __shared__ double2 sm[6 * kBuf + kMixOwn * kMixRh];
  __shared__ double2 twq[kTwq];
  __shared__ double rsc[2];
  __shared__ int2 rjob;   // the job's (run, job), from the loader
  stage_twiddles(twq, tw1024, threadIdx.x, kRowThreads);
  __syncthreads();
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  if (__builtin_amdgcn_readfirstlane(threadIdx.x >> 6) >= 6) {   // loader waves (see k_row)
    const int lt = threadIdx.x - 384;
    MixRowLoader rl;
    if (L < nwork) {
      const int2 it = work[L];
      const RunDesc& a = runs[it.x];
      rl.issue(a, it.y * (64 * kMixEL / a.Nx), lt);
      rl.commit_part(sm + 6 * kBuf, rsc, lt, 0, kMixOwn);
      if (lt == 0) rjob = it;
    }
    for (int w = L; w < nwork; w += gridDim.x) {
      rl.commit_part(sm, rsc, lt, kMixOwn, 7);
      __syncthreads();   // staging ready
      const int wn = w + gridDim.x;
      int2 itn = make_int2(0, 0);
      if (wn < nwork) {
        itn = work[wn];
        const RunDesc& a = runs[itn.x];
        rl.issue(a, itn.y * (64 * kMixEL / a.Nx), lt);
      }
      __syncthreads();   // own staging area free (row_body's first barrier)
      if (wn < nwork) {
        rl.commit_part(sm + 6 * kBuf, rsc, lt, 0, kMixOwn);
        if (lt == 0) rjob = itn;
      }
#pragma unroll 1
      for (int b = 1; b < kRowBodyBarriers; ++b) __syncthreads();
      __syncthreads();   // job done
    }
  } else {
    for (int w = L; w < nwork; w += gridDim.x) {
      __syncthreads();   // staging ready
      const int2 it = rjob;
      const RunDesc& a = runs[it.x];
      const int kout = fin ? col_kept_cols(a) : a.Nxh;
      const double w00 = rsc[0], alu = rsc[1];
      switch (a.Nx) {   // uniform across the block
#define QS_MIX_ROW(NN) \
        case NN: row_body<NN, kMixEL, true, kMixOwn>(a, it.y * (64 * kMixEL / NN), sm, TwN<NN>{twq}, w00, alu, kout); break;
        QS_MIX_ROW(16) QS_MIX_ROW(32) QS_MIX_ROW(64) QS_MIX_ROW(128) QS_MIX_ROW(256) QS_MIX_ROW(512)
        QS_MIX_ROW(1024)
#undef QS_MIX_ROW
        default: break;   // rejected by the host
      }
      __syncthreads();   // job done
    }
  }
}

// Column stage (or prologue, pro != 0) over every class: the block layout of
// k_col (four waves, two blocks per CU).
template <int MODE>
__global__ __launch_bounds__(256, 2) void k_col_mix(const RunDesc* __restrict__ runs, const int2* __restrict__ work,
                                                   int nwork, const double2* __restrict__ tw1024, int pro) {
  __shared__ double2 sm[4 * (kBuf + 1)];
  __shared__ double2 twq[kTwq];
  stage_twiddles(twq, tw1024, threadIdx.x, 256);
  __syncthreads();
  const int L = xcd_remap(blockIdx.x, gridDim.x);
  for (int w = L; w < nwork; w += gridDim.x) {
    const int2 it = work[w];
    const RunDesc& a = runs[it.x];
    switch (a.Ny) {   // uniform across the block
#define QS_MIX_COL(NN) \
      case NN: col_dispatch_n<NN, MODE, kMixEL, true, true>(a, it.y, sm, TwN<NN>{twq}, pro != 0); break;
      QS_MIX_COL(16) QS_MIX_COL(32) QS_MIX_COL(64) QS_MIX_COL(128) QS_MIX_COL(256) QS_MIX_COL(512)
      QS_MIX_COL(1024)
#undef QS_MIX_COL
      default: break;
    }
    __syncthreads();
  }
}

}  // namespace

bool mix_side(int n) { return n >= 16 && n <= 1024 && (n & (n - 1)) == 0; }

int launch_mix_row(const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw1024,
                   hipStream_t st, int fin) {
  hipLaunchKernelGGL(k_row_mix, dim3(grid), dim3(kRowThreads), 0, st, runs, work, nwork, tw1024, fin);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

int launch_mix_col(int mode, const RunDesc* runs, const int2* work, int nwork, int grid, const double2* tw1024,
                   hipStream_t st) {
  switch (mode) {
    case MODE_PRO:
      hipLaunchKernelGGL(k_col_mix<MODE_S2>, dim3(grid), dim3(256), 0, st, runs, work, nwork, tw1024, 1);
      break;
    case MODE_S1:
      hipLaunchKernelGGL(k_col_mix<MODE_S1>, dim3(grid), dim3(256), 0, st, runs, work, nwork, tw1024, 0);
      break;
    case MODE_S2:
      hipLaunchKernelGGL(k_col_mix<MODE_S2>, dim3(grid), dim3(256), 0, st, runs, work, nwork, tw1024, 0);
      break;
    default: return fail(QS_ERR_INVALID_ARG, "merged column kernel: mode %d", mode);
  }
  QS_CHECK_LAUNCH();
  return QS_OK;
}

}  // namespace qs
