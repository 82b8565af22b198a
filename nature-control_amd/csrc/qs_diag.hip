// qs_diag.hip -- diagnostics of the post-processing drivers as GPU reductions:
// the kinetic energy Etot_func (InertialFriction/post_process/postprocess.cu:
// 54-65) and the shell energy spectrum Espec_func (post_process/EnergySpec.cu:
// 81-96).  Both reductions are deterministic: per-row partial sums in a fixed
// order, then one block per output summing the rows in a fixed order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "qs_internal.hpp"

namespace qs {
namespace {

// Block of 256 threads: sum of x over the block in a fixed tree order.
__device__ double block_sum(double x, double* red) {
  const int t = threadIdx.x;
  red[t] = x;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] += red[t + o];
    __syncthreads();
  }
  const double s = red[0];
  __syncthreads();
  return s;
}

// Row j of Etot_func: sum_i (u^2 + v^2) 0.5 dx dy (postprocess.cu:57-62).
__global__ __launch_bounds__(256) void k_etot_rows(const double* __restrict__ u, const double* __restrict__ v,
                                                   int Nx, double dx, double dy, double* __restrict__ part) {
  __shared__ double red[256];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < Nx; i += 256) {
    const size_t e = (size_t)j * Nx + i;
    s += (u[e] * u[e] + v[e] * v[e]) * 0.5 * dx * dy;
  }
  const double r = block_sum(s, red);
  if (threadIdx.x == 0) part[j] = r;
}

// out[b] = (sum_{j < n} part[b * n + j]) / div, one block per output.
__global__ __launch_bounds__(256) void k_sum_rows(const double* __restrict__ part, int n, double div,
                                                  double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += 256) s += part[(size_t)blockIdx.x * n + j];
  const double r = block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = r / div;
}

// |K|^2 of bin (i, row with ky) as Espec_func forms it (EnergySpec.cu:88):
// kx*kx + ky*ky with each product rounded (no fused multiply-add), so that a
// mode on a shell edge falls in the same shell as on the host.
__device__ __forceinline__ double mod2(double kx, double ky2) { return __dadd_rn(__dmul_rn(kx, kx), ky2); }

// First i in [0, n) with mod2(kx[i], ky2) >= e (mod2 is nondecreasing in i).
__device__ __forceinline__ int first_at_least(const double* kx, int n, double ky2, double e) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (mod2(kx[mid], ky2) >= e) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// Thread (shell s, row j): the half-spectrum bins of row j with
// lo[s] <= |K|^2 < hi[s] (EnergySpec.cu:89) form one run of i; their
// |ut|^2 + |vt|^2 (EnergySpec.cu:90) summed into part[s * Ny + j].
__global__ __launch_bounds__(256) void k_espec_rows(const double2* __restrict__ ut, const double2* __restrict__ vt,
                                                    const double* __restrict__ kx, const double* __restrict__ ky,
                                                    int Nxh, int Ny, const double* __restrict__ lo,
                                                    const double* __restrict__ hi, int nshell,
                                                    double* __restrict__ part) {
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long)nshell * Ny) return;
  const int s = (int)(t / Ny), j = (int)(t % Ny);
  const double ky2 = __dmul_rn(ky[j], ky[j]);
  const int i0 = first_at_least(kx, Nxh, ky2, lo[s]);
  const int i1 = first_at_least(kx, Nxh, ky2, hi[s]);
  double E = 0.0;
  for (int i = i0; i < i1; ++i) {
    const size_t e = (size_t)j * Nxh + i;
    const double2 a = ut[e], b = vt[e];
    E += a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y;
  }
  part[t] = E;
}

int diag_scratch(const Mesh* m, size_t n) {
  qs_mesh_priv* p = m->priv;
  if (p->diag_n >= n) return QS_OK;
  if (p->diag) {
    QS_HIP(hipStreamSynchronize(p->stream));
    (void)hipFree(p->diag);
    p->diag = nullptr;
    p->diag_n = 0;
  }
  if (hipMalloc((void**)&p->diag, sizeof(double) * n) != hipSuccess)
    return fail(QS_ERR_OOM, "diagnostics: hipMalloc of %zu doubles", n);
  p->diag_n = n;
  return QS_OK;
}

// Shells with bounds lo[s] <= |K|^2 < hi[s]; E[s] = sum / (2 alpha Nx^2 Ny^2)
// with the normaliser evaluated as EnergySpec.cu:95 writes it.
int espec_launch(const Mesh* m, const Qcomp* ut, const Qcomp* vt, const std::vector<double>& lo,
                 const std::vector<double>& hi, double alpha, Qreal* E) {
  const int ns = (int)lo.size(), Ny = m->Ny, Nxh = m->Nxh;
  const size_t npart = (size_t)ns * Ny;
  QS_TRY(diag_scratch(m, npart + 2 * (size_t)ns));
  double* part = m->priv->diag;
  double* dlo = part + npart;
  double* dhi = dlo + ns;
  const hipStream_t st = stream_of(m);
  QS_HIP(hipMemcpyAsync(dlo, lo.data(), sizeof(double) * ns, hipMemcpyHostToDevice, st));
  QS_HIP(hipMemcpyAsync(dhi, hi.data(), sizeof(double) * ns, hipMemcpyHostToDevice, st));
  const int blocks = (int)((npart + 255) / 256);
  hipLaunchKernelGGL(k_espec_rows, dim3(blocks), dim3(256), 0, st, (const double2*)ut, (const double2*)vt,
                     m->priv->kx1, m->priv->ky1, Nxh, Ny, dlo, dhi, ns, part);
  QS_CHECK_LAUNCH();
  const double norm = 2.0 * alpha * (Nxh - 1) * 2 * Ny * (Nxh - 1) * 2 * Ny;
  hipLaunchKernelGGL(k_sum_rows, dim3(ns), dim3(256), 0, st, part, Ny, norm, E);
  QS_CHECK_LAUNCH();
  // the host bound arrays are freed on return: wait for their copies
  QS_HIP(hipStreamSynchronize(st));
  return QS_OK;
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

int qs_Etot_func(const Mesh* m, const Qreal* u, const Qreal* v, Qreal* E) {
  QS_TRY(check_mesh(m));
  if (!u || !v || !E) return fail(QS_ERR_INVALID_ARG, "qs_Etot_func: null argument");
  QS_TRY(diag_scratch(m, (size_t)m->Ny));
  const hipStream_t st = stream_of(m);
  hipLaunchKernelGGL(k_etot_rows, dim3(m->Ny), dim3(256), 0, st, u, v, m->Nx, m->dx, m->dy, m->priv->diag);
  QS_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, st, m->priv->diag, m->Ny, 1.0, E);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

int qs_Espec_func(const Mesh* m, Qreal k, const Qcomp* ut, const Qcomp* vt, Qreal alpha, Qreal* E) {
  QS_TRY(check_mesh(m));
  if (!ut || !vt || !E) return fail(QS_ERR_INVALID_ARG, "qs_Espec_func: null argument");
  const std::vector<double> lo{k * k}, hi{(k + alpha) * (k + alpha)};
  return espec_launch(m, ut, vt, lo, hi, alpha, E);
}

int qs_Espec_shells(const Mesh* m, const Qcomp* ut, const Qcomp* vt, Qreal dk, Qreal kmax, int nmax, Qreal* E,
                    int* nshell) {
  QS_TRY(check_mesh(m));
  if (!ut || !vt || !E || !nshell) return fail(QS_ERR_INVALID_ARG, "qs_Espec_shells: null argument");
  if (!(dk > 0)) return fail(QS_ERR_INVALID_ARG, "qs_Espec_shells: dk must be > 0");
  // the shells of EnergySpec.cu:198-204: k = 0, dk, 2 dk, ... accumulated by
  // repeated addition while k < kmax; shell n is k_n^2 <= |K|^2 < (k_n + dk)^2
  std::vector<double> lo, hi;
  for (double k = 0.0; k < kmax; k += dk) {
    if ((int)lo.size() >= nmax)
      return fail(QS_ERR_INVALID_ARG, "qs_Espec_shells: more than nmax=%d shells below kmax", nmax);
    lo.push_back(k * k);
    hi.push_back((k + dk) * (k + dk));
  }
  *nshell = (int)lo.size();
  if (lo.empty()) return QS_OK;
  return espec_launch(m, ut, vt, lo, hi, dk, E);
}

}  // extern "C"
