// qs_adjoint_host.hip -- checkpointed discrete adjoint of the fused IFRK2 step:
// forward with checkpoints, segment recompute recording every stage's inverse
// set, reverse sweep (kernels in qs_adjoint.hip), loss and its gradient.
//
// No reference counterpart (the reference has no adjoint).  The forward that
// is differentiated is the fused step of qs_solver.hip (friction.cu:212-260
// with TimeIntegration/IFRK2.cuh:16-51, IFh = IF).
//
// Memory (H = Ny * pitch * 16 B): checkpoints 3 H each (one per segment plus
// the final state), stage slots 7 H each for 2k+1 stages of one segment,
// adjoint state 6 H, activity gradient Nx*Ny*8 B, per-row sums nsteps*Ny*8 B,
// row-adjoint scratch grid * 96 KiB.
#include <algorithm>
#include <cstring>

#include "qs_solver.hpp"

namespace qs {
namespace {

constexpr int kAdjRowGrid = 256;   // row adjoint blocks (one per CU; LDS-bound)

int nfix_of(const Mesh* m) { return m->Nx == m->Ny ? m->Nx : 0; }

template <class T>
int dalloc(AdjState* A, T** p, size_t bytes) {
  if (hipMalloc((void**)p, bytes) != hipSuccess) return fail(QS_ERR_OOM, "adjoint: hipMalloc of %zu bytes", bytes);
  A->bytes += bytes;
  return QS_OK;
}

void adj_free(AdjState* A) {
  if (!A) return;
  for (auto p : A->ckpt) (void)hipFree(p);
  for (auto p : A->slots) (void)hipFree(p);
  for (void* p : {(void*)A->alpha_t, (void*)A->w00s, (void*)A->abar, (void*)A->arow, (void*)A->w00bar,
                  (void*)A->scratch, (void*)A->d_desc})
    if (p) (void)hipFree(p);
  for (int k = 0; k < 3; ++k) {
    if (A->lam_[k]) (void)hipFree(A->lam_[k]);
    if (A->acc_[k]) (void)hipFree(A->acc_[k]);
  }
  delete A;
}

size_t hbytes(const qs_solver* S) { return sizeof(double2) * (size_t)S->pitch * S->mesh->Ny; }

// (Re)allocates the adjoint's memory for nsteps with checkpoint interval k.
int adj_prepare(qs_solver* S, int nsteps, int k) {
  const Mesh* m = S->mesh;
  AdjState* A = S->adj;
  const int nseg = (nsteps + k - 1) / k;
  const bool fits = A && A->kcap >= k && (int)A->ckpt.size() >= 3 * (nseg + 1) && A->ncap >= nsteps;
  if (!fits) {
    adj_free(A);
    S->adj = A = new AdjState();
    A->kcap = k;
    const size_t hs = hbytes(S);
    A->ckpt.assign(3 * (size_t)(nseg + 1), nullptr);
    for (auto& p : A->ckpt) QS_TRY(dalloc(A, &p, hs));
    A->slots.assign(7 * (size_t)(2 * k + 1), nullptr);
    for (auto& p : A->slots) QS_TRY(dalloc(A, &p, hs));
    QS_TRY(dalloc(A, &A->w00s, sizeof(double) * (2 * k + 1)));
    for (int c = 0; c < 3; ++c) {
      QS_TRY(dalloc(A, &A->lam_[c], hs));
      QS_TRY(dalloc(A, &A->acc_[c], hs));
    }
    QS_TRY(dalloc(A, &A->abar, sizeof(double) * (size_t)m->Nx * m->Ny));
    QS_TRY(dalloc(A, &A->arow, sizeof(double) * (size_t)std::max(nsteps, 1) * m->Ny));
    QS_TRY(dalloc(A, &A->w00bar, sizeof(double) * m->Ny));
    QS_TRY(dalloc(A, &A->alpha_t, sizeof(double) * std::max(nsteps, 1)));
    A->grid = std::min(kAdjRowGrid, S->nrow);
    QS_TRY(dalloc(A, &A->scratch, sizeof(double2) * 6 * 1024 * (size_t)A->grid));
    A->ndesc = std::max(4 * k + 1, nsteps);
    QS_TRY(dalloc(A, &A->d_desc, sizeof(RunDesc) * A->ndesc));
    A->h_desc.resize(A->ndesc);
    A->ncap = nsteps;
  }
  A->k = k;
  return QS_OK;
}

int copy_pitched(qs_solver* S, double2* dst, const double2* src) {
  const size_t row = sizeof(double2) * S->pitch;
  QS_HIP(hipMemcpyAsync(dst, src, row * S->mesh->Ny, hipMemcpyDeviceToDevice, stream_of(S->mesh)));
  return QS_OK;
}

int launch_row(qs_solver* S, const RunDesc* d) {
  return launch_fused_row(nfix_of(S->mesh), d, S->d_work, S->nrow, S->nrow, S->mesh->priv->tw1024,
                          stream_of(S->mesh));
}
int launch_col(qs_solver* S, const RunDesc* d, int mode) {
  return launch_fused_col(nfix_of(S->mesh), mode, d, S->d_work, S->ncol, S->ncol, S->mesh->priv->tw1024,
                          stream_of(S->mesh));
}
int launch_arow(qs_solver* S, const RunDesc* d) {
  return launch_adj_row(nfix_of(S->mesh), d, S->d_work, S->nrow, S->adj->grid, S->mesh->priv->tw1024,
                        stream_of(S->mesh));
}
int launch_acol(qs_solver* S, const RunDesc* d, int amode) {
  return launch_adj_col(nfix_of(S->mesh), amode, d, S->d_work, S->ncol, S->ncol, S->mesh->priv->tw1024,
                        stream_of(S->mesh));
}

int upload(qs_solver* S, int n) {
  AdjState* A = S->adj;
  QS_HIP(hipMemcpyAsync(A->d_desc, A->h_desc.data(), sizeof(RunDesc) * n, hipMemcpyHostToDevice,
                        stream_of(S->mesh)));
  // the host array is rewritten for the next segment: wait for the copy
  QS_HIP(hipStreamSynchronize(stream_of(S->mesh)));
  return QS_OK;
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

int qs_adj_release(qs_solver* S) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  if (S->adj) {
    (void)hipStreamSynchronize(stream_of(S->mesh));
    adj_free(S->adj);
    S->adj = nullptr;
  }
  return QS_OK;
}

int qs_adj_memory(qs_solver* S, size_t* bytes) {
  if (!S || !bytes) return fail(QS_ERR_INVALID_ARG, "null argument");
  *bytes = S->adj ? S->adj->bytes : 0;
  return QS_OK;
}

int qs_adj_forward(qs_solver* S, int nsteps, int ckpt_every, const Qreal* alpha_t) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  if (nsteps <= 0) return fail(QS_ERR_INVALID_ARG, "qs_adj_forward: nsteps must be > 0");
  const Mesh* m = S->mesh;
  if (!nfix_of(m)) return fail(QS_ERR_UNSUPPORTED_SIZE, "qs_adj_forward: square meshes only");
  const int k = ckpt_every <= 0 ? nsteps : std::min(ckpt_every, nsteps);
  QS_TRY(adj_prepare(S, nsteps, k));
  AdjState* A = S->adj;
  A->nsteps = nsteps;
  A->sched = alpha_t != nullptr;
  const hipStream_t st = stream_of(m);
  if (alpha_t)
    QS_HIP(hipMemcpyAsync(A->alpha_t, alpha_t, sizeof(double) * nsteps, hipMemcpyDeviceToDevice, st));
  // one descriptor per step (the activity of the step), the solver's own arrays
  const RunDesc base = make_desc(S);
  for (int n = 0; n < nsteps; ++n) {
    RunDesc d = base;
    d.alpha_p = alpha_t ? A->alpha_t + n : nullptr;
    A->h_desc[n] = d;
  }
  QS_TRY(upload(S, nsteps));
  const int nseg = (nsteps + k - 1) / k;
  for (int n = 0; n < nsteps; ++n) {
    if (n % k == 0)
      for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, A->ckpt[3 * (n / k) + c], S->old_[c]));
    const RunDesc* d = A->d_desc + n;
    if (!S->mixed_valid) {
      QS_TRY(launch_col(S, d, MODE_PRO));
      S->mixed_valid = true;
    }
    QS_TRY(launch_row(S, d));
    QS_TRY(launch_col(S, d, MODE_S1));
    QS_TRY(launch_row(S, d));
    QS_TRY(launch_col(S, d, MODE_S2));
  }
  for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, A->ckpt[3 * nseg + c], S->old_[c]));
  return QS_OK;
}

int qs_adj_backward(qs_solver* S, const Qcomp* gW, const Qcomp* gR1, const Qcomp* gR2, Qreal* abar,
                    Qreal* abar_t, Qcomp* uW, Qcomp* uR1, Qcomp* uR2) {
  if (!S || !gW || !gR1 || !gR2) return fail(QS_ERR_INVALID_ARG, "qs_adj_backward: null argument");
  AdjState* A = S->adj;
  if (!A || A->nsteps <= 0) return fail(QS_ERR_STATE, "qs_adj_backward: no qs_adj_forward to reverse");
  const Mesh* m = S->mesh;
  const hipStream_t st = stream_of(m);
  const int nsteps = A->nsteps, k = A->k, nseg = (nsteps + k - 1) / k;
  const size_t prow = sizeof(double2) * S->pitch, row = sizeof(Qcomp) * m->Nxh;
  const Qcomp* g[3] = {gW, gR1, gR2};
  for (int c = 0; c < 3; ++c)
    QS_HIP(hipMemcpy2DAsync(A->lam_[c], prow, g[c], row, row, m->Ny, hipMemcpyDeviceToDevice, st));
  QS_HIP(hipMemsetAsync(A->abar, 0, sizeof(double) * (size_t)m->Nx * m->Ny, st));
  QS_HIP(hipMemsetAsync(A->arow, 0, sizeof(double) * (size_t)nsteps * m->Ny, st));
  const RunDesc base = make_desc(S);
  auto slot = [&](int s, int q) { return A->slots[7 * (size_t)s + q]; };

  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int s0 = seg * k, len = std::min(k, nsteps - s0);
    // ---- recompute the segment, recording every stage's inverse set -------
    // descriptors: [0] prologue, [1 + s] stage s (row input slot s, column
    // output slot s + 1), [2k + 1 + s] adjoint of stage s
    for (int s = 0; s <= 2 * len; ++s) {
      RunDesc d = base;
      const int step = s0 + (s > 0 ? (s - 1) / 2 : 0);
      d.alpha_p = A->sched ? A->alpha_t + step : nullptr;
      if (s == 0) {
        for (int q = 0; q < 7; ++q) d.mout[q] = slot(0, q);
        d.w00o = A->w00s;
      } else {
        const int t = s - 1;
        for (int q = 0; q < 7; ++q) { d.minp[q] = slot(t, q); d.mout[q] = slot(t + 1, q); }
        d.w00 = A->w00s + t;
        d.w00o = A->w00s + t + 1;
      }
      A->h_desc[s] = d;
    }
    for (int s = 0; s < 2 * len; ++s) {
      RunDesc d = base;
      const int step = s0 + s / 2;
      d.alpha_p = A->sched ? A->alpha_t + step : nullptr;
      for (int q = 0; q < 7; ++q) d.minp[q] = slot(s, q);
      d.w00 = A->w00s + s;
      for (int c = 0; c < 3; ++c) { d.lam_[c] = A->lam_[c]; d.acc_[c] = A->acc_[c]; d.uout[c] = A->lam_[c]; }
      d.abar = A->abar;
      d.arow = A->arow + (size_t)step * m->Ny;
      d.w00bar = A->w00bar;
      d.scratch = A->scratch;
      A->h_desc[2 * k + 1 + s] = d;
    }
    QS_TRY(upload(S, 2 * k + 1 + 2 * len));
    const RunDesc* D = A->d_desc;
    const RunDesc* AD = A->d_desc + 2 * k + 1;
    for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, S->old_[c], A->ckpt[3 * seg + c]));
    QS_TRY(launch_col(S, D, MODE_PRO));
    for (int n = 0; n < len; ++n) {
      QS_TRY(launch_row(S, D + 1 + 2 * n));
      QS_TRY(launch_col(S, D + 1 + 2 * n, MODE_S1));
      QS_TRY(launch_row(S, D + 2 + 2 * n));
      QS_TRY(launch_col(S, D + 2 + 2 * n, MODE_S2));
    }
    // ---- reverse sweep over the segment -------------------------------------
    QS_TRY(launch_acol(S, AD + 2 * len - 1, AMODE_START));
    for (int n = len - 1; n >= 0; --n) {
      QS_TRY(launch_arow(S, AD + 2 * n + 1));                   // N'(c2)^T
      QS_TRY(launch_acol(S, AD + 2 * n + 1, AMODE_B));
      QS_TRY(launch_arow(S, AD + 2 * n));                       // N'(c1)^T
      QS_TRY(launch_acol(S, AD + 2 * n, n > 0 ? AMODE_A : AMODE_END));
    }
  }
  // outputs
  if (abar) QS_HIP(hipMemcpyAsync(abar, A->abar, sizeof(double) * (size_t)m->Nx * m->Ny, hipMemcpyDeviceToDevice, st));
  if (abar_t) QS_TRY(row_sums_launch(A->arow, nsteps, m->Ny, abar_t, st));
  Qcomp* u[3] = {uW, uR1, uR2};
  for (int c = 0; c < 3; ++c)
    if (u[c]) QS_HIP(hipMemcpy2DAsync(u[c], row, A->lam_[c], prow, row, m->Ny, hipMemcpyDeviceToDevice, st));
  // leave the solver at the final state of the horizon
  for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, S->old_[c], A->ckpt[3 * nseg + c]));
  S->mixed_valid = false;
  return QS_OK;
}

int qs_adj_loss(const Mesh* m, int kind, const Qcomp* W, Qreal* loss, Qcomp* Wbar) {
  QS_TRY(check_mesh(m));
  if (!W || !loss || !Wbar) return fail(QS_ERR_INVALID_ARG, "qs_adj_loss: null argument");
  if (kind != 0 && kind != 1) return fail(QS_ERR_INVALID_ARG, "qs_adj_loss: kind %d", kind);
  return loss_launch(m, kind, (const double2*)W, loss, (double2*)Wbar);
}

}  // extern "C"
