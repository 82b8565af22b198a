// qs_noalloc.hip -- the fused IFRK2 step for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
//
// qs_step_ifrk2_noalloc runs the kernels of qs_step_ifrk2 (column prologue,
// then row pass / column stage 1 / row pass / column stage 2 per step) on the
// caller's state spectra, in place, in the reference layout (row pitch Nxh).
// Everything else those kernels need -- the stage-1 partial sums (3 spectra),
// the seven mixed arrays, the run descriptor, the work list, W(0,0) and the
// non-finite flag -- lives in `scratch`, a fixed device array of this file.
// Each GPU holds its own copy until the process ends; its address on the
// current device comes from hipGetSymbolAddress at the first call there.
//
// Nothing here allocates: no hipMalloc / hipFree, no host-to-device copy
// (the descriptor travels as a kernel argument), no memset node (a small
// kernel clears the flag), so the whole call can be captured into a graph.
// Every call on one GPU shares that GPU's array: calls that overlap on one GPU
// (two streams, two host threads) race and give wrong results.
#include <mutex>

#include "qs_fused.hpp"

namespace qs {
namespace {

// Largest spectrum (Ny * Nxh entries) the scratch array holds: 512 x 512
// (512 * 257), or any other fused mesh of at most that many entries.
constexpr long kMaxSpec = 512L * 257;
constexpr int kArrays = 10;                  // new_[3] + mix[7]
constexpr long kArraySlots = kArrays * kMaxSpec;
constexpr int kMaxWork = 1024;               // work-list entries (row and column jobs)
constexpr int kMaxSide = 1024;               // longest Ny (integrating-factor rows)
// descriptor, work list, W(0,0), flag; integrating-factor tables (rows and columns)
constexpr long kTailBytes = 16384 + 4 * (kMaxSide + kMaxSide / 2 + 1) * 8;
static_assert(((sizeof(RunDesc) + 255) & ~(size_t)255) + sizeof(int2) * kMaxWork + 16 + 4 * kMaxSide * 8 <=
                  (size_t)kTailBytes,
              "scratch tail");
constexpr long kScratchSlots = kArraySlots + kTailBytes / (long)sizeof(double2);

// 21,102,592 bytes per GPU.
__device__ __align__(256) double2 scratch[kScratchSlots];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Address of this device's scratch array (hipGetSymbolAddress once per
// device, with that device current).
int scratch_ptr(double2** out) {
  int dev = -1;
  QS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(QS_ERR_INVALID_ARG, "noalloc: device %d", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(QS_ERR_HIP, "noalloc: hipGetSymbolAddress(scratch) -> %s", hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = (double2*)g_scratch[dev];
  return QS_OK;
}

// Descriptor into the scratch tail, the identity work list (run 0, job k),
// the integrating-factor tables (those of ifac_launch), and the non-finite
// flag cleared.
__global__ __launch_bounds__(256) void k_noalloc_setup(RunDesc d, RunDesc* dd, int2* work, int nwork, int* flag,
                                                       double* ify) {
  if (threadIdx.x == 0) {
    *dd = d;
    *flag = 0;
  }
  for (int k = threadIdx.x; k < nwork; k += blockDim.x) work[k] = make_int2(0, k);
  for (int j = threadIdx.x; j < d.Ny; j += blockDim.x) ifac_row(j, d.Ny, d.dky, d.Re, d.DQ, d.dt, ify);
  for (int i = threadIdx.x; i < d.Nxh; i += blockDim.x)
    ifac_col(i, d.Nxh, d.Ny, d.dkx, d.Re, d.Rf, d.DQ, d.aQ, d.dt, ify);
}

__global__ void k_noalloc_flag(const int* flag, int* out) { *out = *flag; }

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" int qs_step_ifrk2_noalloc(const Mesh* mesh, const qs_params* p, const Qreal* Ra, Qcomp* w, Qcomp* r1,
                                     Qcomp* r2, int nsteps, int* nonfinite) {
  QS_TRY(check_mesh(mesh));
  if (!p || !w || !r1 || !r2) return fail(QS_ERR_INVALID_ARG, "qs_step_ifrk2_noalloc: null argument");
  if (nsteps < 0) return fail(QS_ERR_INVALID_ARG, "qs_step_ifrk2_noalloc: nsteps < 0");
  if (!(p->Re != 0) || !(p->Er != 0) || !(p->dt > 0))
    return fail(QS_ERR_INVALID_ARG, "qs_step_ifrk2_noalloc: Re, Er must be non-zero and dt > 0");
  const int Nx = mesh->Nx, Ny = mesh->Ny, Nxh = mesh->Nxh;
  const bool pow2 = fused_side(Nx) && fused_side(Ny);
  const bool tri = Nx == Ny && fused_tri_side(Nx);
  if (!pow2 && !tri)
    return fail(QS_ERR_UNSUPPORTED_SIZE, "qs_step_ifrk2_noalloc: %dx%d mesh (powers of two up to 1024, or square "
                "3*2^a up to 768)", Nx, Ny);
  if ((long)Ny * Nxh > kMaxSpec)
    return fail(QS_ERR_UNSUPPORTED_SIZE, "qs_step_ifrk2_noalloc: %dx%d mesh needs %ld spectral entries per array; "
                "its fixed scratch holds %ld (512x512); use qs_solver_create / qs_step_ifrk2", Nx, Ny,
                (long)Ny * Nxh, kMaxSpec);
  double2* base = nullptr;
  QS_TRY(scratch_ptr(&base));

  // Arrays of the fused kernels at row pitch Nxh (the caller's layout).
  const long hs = (long)Ny * Nxh;
  RunDesc d{};
  d.Nx = Nx; d.Ny = Ny; d.Nxh = Nxh; d.pitch = Nxh;
  double2* st[3] = {(double2*)w, (double2*)r1, (double2*)r2};
  for (int k = 0; k < 3; ++k) {
    d.old_[k] = st[k];
    d.new_[k] = base + k * hs;
  }
  for (int k = 0; k < 7; ++k) d.mix[k] = d.minp[k] = d.mout[k] = base + (3 + k) * hs;
  char* tail = (char*)(base + kArraySlots);
  RunDesc* dd = (RunDesc*)tail;
  int2* work = (int2*)(tail + ((sizeof(RunDesc) + 255) & ~(size_t)255));
  double* w00 = (double*)(work + kMaxWork);
  int* flag = (int*)(w00 + 1);
  double* ify = w00 + 2;   // ifac_len(Nxh, Ny)
  d.Ra = Ra;
  d.alpha_p = nullptr;
  d.w00 = d.w00o = w00;
  d.flag = flag;
  d.ify = ify;
  d.dkx = 2 * M_PI / mesh->Lx;   // the mesh tables' own expressions (Basic/Mesh.cuh:6-21)
  d.dky = 2 * M_PI / mesh->Ly;
  d.alpha = p->alpha; d.lambda = p->lambda;
  d.DQ = p->D_Q; d.aQ = p->a_Q; d.cQ = p->c_Q;
  d.s_ErRe = 1.0 / (p->Er * p->Re);
  d.scale = 1.0 / ((double)Nx * (double)Ny);
  d.dt = p->dt; d.Re = p->Re; d.Rf = p->Rf;
  d.advect_w = p->advect_w;

  const int nrow = row_jobs(d), ncol = col_jobs(d);
  const int nwork = nrow > ncol ? nrow : ncol;
  if (nwork > kMaxWork) return fail(QS_ERR_UNSUPPORTED_SIZE, "qs_step_ifrk2_noalloc: %d jobs", nwork);
  const int ncu = mesh->priv->ncu;
  const int rgrid = ncu > 0 && ncu < nrow ? ncu : nrow;
  const double2* tw = Nx % 3 == 0 ? mesh->priv->tw768 : mesh->priv->tw1024;
  const hipStream_t s = stream_of(mesh);

  hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(256), 0, s, d, dd, work, nwork, flag, ify);
  QS_CHECK_LAUNCH();
  // The inverse set of the caller's state (the scratch keeps nothing from an
  // earlier call), then the steps: the same kernels and arithmetic as
  // qs_solver_set_state + qs_step_ifrk2.
  QS_TRY(launch_fused_col(Ny, MODE_PRO, dd, work, ncol, ncol, tw, s));
  for (int n = 0; n < nsteps; ++n) {
    QS_TRY(launch_fused_row(Nx, dd, work, nrow, rgrid, tw, s));
    QS_TRY(launch_fused_col(Ny, MODE_S1, dd, work, ncol, ncol, tw, s));
    QS_TRY(launch_fused_row(Nx, dd, work, nrow, rgrid, tw, s, 1));
    QS_TRY(launch_fused_col(Ny, MODE_S2, dd, work, ncol, ncol, tw, s));
  }
  if (nonfinite) {
    hipLaunchKernelGGL(k_noalloc_flag, dim3(1), dim3(1), 0, s, (const int*)flag, nonfinite);
    QS_CHECK_LAUNCH();
  }
  return QS_OK;
}
