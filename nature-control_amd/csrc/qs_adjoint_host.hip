// qs_adjoint_host.hip -- checkpointed discrete adjoint of the fused IFRK2 step
// for a group of runs of one mesh size (one run: qs_adj_*; an ensemble of one
// size: qs_ensemble_adj_*): forward with checkpoints, segment recompute
// recording every stage's inverse set, reverse sweep (kernels in
// qs_adjoint.hip), losses and their gradients.
//
// No reference counterpart (the reference has no adjoint).  The forward that
// is differentiated is the fused step of qs_solver.hip (friction.cu:212-260
// with TimeIntegration/IFRK2.cuh:16-51, IFh = IF).
//
// Every launch covers all runs of the group: work lists of (run, job) pairs
// over a descriptor array laid out [stage][run].  Memory per run (H = Ny *
// pitch * 16 B): checkpoints 3 H each (one per segment plus the final state),
// stage slots 7 H each for 2k+1 stages of one segment, adjoint state 6 H,
// activity gradient Nx*Ny*8 B, per-row sums nsteps*Ny*8 B; per group the
// row-adjoint scratch grid * 96 KiB and the descriptors.
#include <algorithm>
#include <cstring>

#include "qs_solver.hpp"

namespace qs {

struct AdjGroup {
  std::vector<qs_solver*> runs;   // one mesh size, one stream
  int R = 0;
  int nsteps = 0, k = 0;          // horizon and checkpoint interval of the last forward
  int kcap = 0, ncap = 0;
  bool sched = false;
  double* alpha_t = nullptr;      // [R][ncap] uniform activity per run and step
  RunDesc* d_desc = nullptr;      // [stage][run]
  int ndesc = 0;
  std::vector<RunDesc> h_desc;
  int2* d_row = nullptr;          // (run, job)
  int2* d_col = nullptr;
  int nrow = 0, ncol = 0;
  double2* scratch = nullptr;     // [grid][6][1024]
  int grid = 0;
  size_t bytes = 0;
};

namespace {

constexpr int kAdjRowGrid = 256;   // row adjoint blocks (one per CU; LDS-bound)

template <class T>
int dalloc(size_t* acc, T** p, size_t bytes) {
  if (hipMalloc((void**)p, bytes) != hipSuccess) return fail(QS_ERR_OOM, "adjoint: hipMalloc of %zu bytes", bytes);
  *acc += bytes;
  return QS_OK;
}

void run_free(AdjState* A);
void group_free(AdjGroup* G) {
  if (!G) return;
  for (void* p : {(void*)G->alpha_t, (void*)G->d_desc, (void*)G->d_row, (void*)G->d_col, (void*)G->scratch})
    if (p) (void)hipFree(p);
  delete G;
}

void run_free(AdjState* A) {
  if (!A) return;
  if (A->arena) (void)hipFree(A->arena);
  group_free(A->solo);
  delete A;
}

size_t hbytes(const qs_solver* S) { return sizeof(double2) * (size_t)S->pitch * S->mesh->Ny; }
const Mesh* mesh_of(const AdjGroup* G) { return G->runs[0]->mesh; }
hipStream_t stream_g(const AdjGroup* G) { return stream_of(mesh_of(G)); }

// Per-run memory for a horizon of nsteps with checkpoint interval k: one
// allocation carved into 256-byte aligned arrays (many separate allocations of
// a few MB each round up to the allocator's granularity).
int run_prepare(qs_solver* S, int nsteps, int k) {
  const Mesh* m = S->mesh;
  AdjState* A = S->adj;
  const int nseg = (nsteps + k - 1) / k;
  if (A && A->kcap >= k && A->nckpt >= nseg + 1 && A->ncap >= nsteps) return QS_OK;
  AdjGroup* solo = A ? A->solo : nullptr;
  if (A) A->solo = nullptr;
  run_free(A);
  S->adj = A = new AdjState();
  A->solo = solo;
  A->kcap = k;
  A->ncap = nsteps;
  A->nckpt = nseg + 1;
  const size_t hs = (hbytes(S) + 255) & ~(size_t)255;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t P = (size_t)m->Nx * m->Ny;
  const size_t total = hs * (3 * (size_t)(nseg + 1) + 7 * (size_t)(2 * k + 1) + 6) +
                       al(sizeof(double) * (2 * k + 1)) + al(sizeof(double) * P) +
                       al(sizeof(double) * (size_t)nsteps * m->Ny) + al(sizeof(double) * m->Ny);
  QS_TRY(dalloc(&A->bytes, &A->arena, total));
  char* p = (char*)A->arena;
  auto take = [&](size_t b) { char* q = p; p += b; return q; };
  A->ckpt.assign(3 * (size_t)(nseg + 1), nullptr);
  for (auto& q : A->ckpt) q = (double2*)take(hs);
  A->slots.assign(7 * (size_t)(2 * k + 1), nullptr);
  for (auto& q : A->slots) q = (double2*)take(hs);
  for (int c = 0; c < 3; ++c) {
    A->lam_[c] = (double2*)take(hs);
    A->acc_[c] = (double2*)take(hs);
  }
  A->w00s = (double*)take(al(sizeof(double) * (2 * k + 1)));
  A->abar = (double*)take(al(sizeof(double) * P));
  A->arow = (double*)take(al(sizeof(double) * (size_t)nsteps * m->Ny));
  A->w00bar = (double*)take(al(sizeof(double) * m->Ny));
  return QS_OK;
}

int group_prepare(AdjGroup* G, int nsteps, int k) {
  for (auto* S : G->runs) QS_TRY(run_prepare(S, nsteps, k));
  if (G->kcap >= k && G->ncap >= nsteps && G->d_desc) return QS_OK;
  for (void* p : {(void*)G->alpha_t, (void*)G->d_desc, (void*)G->d_row, (void*)G->d_col, (void*)G->scratch})
    if (p) (void)hipFree(p);
  G->bytes = 0;
  G->kcap = k;
  G->ncap = nsteps;
  const int R = G->R;
  QS_TRY(dalloc(&G->bytes, &G->alpha_t, sizeof(double) * (size_t)R * nsteps));
  G->ndesc = std::max(4 * k + 1, nsteps) * R;
  QS_TRY(dalloc(&G->bytes, &G->d_desc, sizeof(RunDesc) * G->ndesc));
  G->h_desc.resize(G->ndesc);
  // work lists (run, job), run-major
  const RunDesc d0 = make_desc(G->runs[0]);
  const int nr = row_jobs(d0), nc = col_jobs(d0);
  std::vector<int2> row, col;
  for (int r = 0; r < R; ++r) {
    for (int j = 0; j < nr; ++j) row.push_back(make_int2(r, j));
    for (int j = 0; j < nc; ++j) col.push_back(make_int2(r, j));
  }
  G->nrow = (int)row.size();
  G->ncol = (int)col.size();
  QS_TRY(dalloc(&G->bytes, &G->d_row, sizeof(int2) * row.size()));
  QS_TRY(dalloc(&G->bytes, &G->d_col, sizeof(int2) * col.size()));
  QS_HIP(hipMemcpy(G->d_row, row.data(), sizeof(int2) * row.size(), hipMemcpyHostToDevice));
  QS_HIP(hipMemcpy(G->d_col, col.data(), sizeof(int2) * col.size(), hipMemcpyHostToDevice));
  G->grid = std::min(kAdjRowGrid, G->nrow);
  QS_TRY(dalloc(&G->bytes, &G->scratch, sizeof(double2) * 6 * 1024 * (size_t)G->grid));
  return QS_OK;
}

int copy_pitched(const qs_solver* S, double2* dst, const double2* src) {
  QS_HIP(hipMemcpyAsync(dst, src, hbytes(S), hipMemcpyDeviceToDevice, stream_of(S->mesh)));
  return QS_OK;
}

int launch_row(AdjGroup* G, const RunDesc* d) {   // persistent: one block per CU
  const Mesh* m = mesh_of(G);
  const int grid = m->priv->ncu > 0 ? std::min(G->nrow, m->priv->ncu) : G->nrow;
  return launch_fused_row(m->Nx, d, G->d_row, G->nrow, grid, m->priv->tw1024, stream_g(G));
}
int launch_col(AdjGroup* G, const RunDesc* d, int mode) {
  const Mesh* m = mesh_of(G);
  return launch_fused_col(m->Nx, mode, d, G->d_col, G->ncol, G->ncol, m->priv->tw1024, stream_g(G));
}
int launch_arow(AdjGroup* G, const RunDesc* d) {
  const Mesh* m = mesh_of(G);
  return launch_adj_row(m->Nx, d, G->d_row, G->nrow, G->grid, m->priv->tw1024, stream_g(G));
}
int launch_acol(AdjGroup* G, const RunDesc* d, int amode) {
  const Mesh* m = mesh_of(G);
  return launch_adj_col(m->Nx, amode, d, G->d_col, G->ncol, G->ncol, m->priv->tw1024, stream_g(G));
}

int upload(AdjGroup* G, int n) {
  QS_HIP(hipMemcpyAsync(G->d_desc, G->h_desc.data(), sizeof(RunDesc) * n, hipMemcpyHostToDevice, stream_g(G)));
  // the host array is rewritten for the next segment: wait for the copy
  QS_HIP(hipStreamSynchronize(stream_g(G)));
  return QS_OK;
}

// Forward: nsteps steps of every run with checkpoints every k steps (and the
// final state).  alpha_t: device [R][nsteps] or null.
int group_forward(AdjGroup* G, int nsteps, int k, const double* alpha_t) {
  QS_TRY(group_prepare(G, nsteps, k));
  const int R = G->R;
  G->nsteps = nsteps;
  G->k = k;
  G->sched = alpha_t != nullptr;
  const hipStream_t st = stream_g(G);
  if (alpha_t)
    QS_HIP(hipMemcpyAsync(G->alpha_t, alpha_t, sizeof(double) * (size_t)R * nsteps, hipMemcpyDeviceToDevice, st));
  for (int n = 0; n < nsteps; ++n)
    for (int r = 0; r < R; ++r) {
      RunDesc d = make_desc(G->runs[r]);
      d.alpha_p = alpha_t ? G->alpha_t + (size_t)r * nsteps + n : nullptr;
      G->h_desc[(size_t)n * R + r] = d;
    }
  QS_TRY(upload(G, nsteps * R));
  const int nseg = (nsteps + k - 1) / k;
  for (int n = 0; n < nsteps; ++n) {
    if (n % k == 0)
      for (auto* S : G->runs)
        for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, S->adj->ckpt[3 * (n / k) + c], S->old_[c]));
    const RunDesc* d = G->d_desc + (size_t)n * R;
    if (n == 0) QS_TRY(launch_col(G, d, MODE_PRO));
    QS_TRY(launch_row(G, d));
    QS_TRY(launch_col(G, d, MODE_S1));
    QS_TRY(launch_row(G, d));
    QS_TRY(launch_col(G, d, MODE_S2));
  }
  for (auto* S : G->runs) {
    for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, S->adj->ckpt[3 * nseg + c], S->old_[c]));
    S->mixed_valid = true;   // the mixed arrays hold the final state's inverse set
  }
  return QS_OK;
}

// Reverse sweep.  The seeds dL/d(final state) are in every run's lam_.
int group_backward(AdjGroup* G) {
  const int R = G->R, nsteps = G->nsteps, k = G->k, nseg = (nsteps + k - 1) / k;
  const Mesh* m = mesh_of(G);
  const hipStream_t st = stream_g(G);
  for (auto* S : G->runs) {
    QS_HIP(hipMemsetAsync(S->adj->abar, 0, sizeof(double) * (size_t)m->Nx * m->Ny, st));
    QS_HIP(hipMemsetAsync(S->adj->arow, 0, sizeof(double) * (size_t)nsteps * m->Ny, st));
  }
  auto slot = [&](const qs_solver* S, int s, int q) { return S->adj->slots[7 * (size_t)s + q]; };
  const size_t adj0 = (size_t)(2 * k + 1) * R;   // first adjoint descriptor
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int s0 = seg * k, len = std::min(k, nsteps - s0);
    // descriptors [stage][run]: [0] prologue, [1 + s] stage s (row input slot
    // s, column output slot s + 1); from adj0: adjoint of stage s
    for (int s = 0; s <= 2 * len; ++s)
      for (int r = 0; r < R; ++r) {
        const qs_solver* S = G->runs[r];
        const AdjState* A = S->adj;
        RunDesc d = make_desc(G->runs[r]);
        const int step = s0 + (s > 0 ? (s - 1) / 2 : 0);
        d.alpha_p = G->sched ? G->alpha_t + (size_t)r * nsteps + step : nullptr;
        if (s == 0) {
          for (int q = 0; q < 7; ++q) d.mout[q] = slot(S, 0, q);
          d.w00o = A->w00s;
        } else {
          const int t = s - 1;
          for (int q = 0; q < 7; ++q) { d.minp[q] = slot(S, t, q); d.mout[q] = slot(S, t + 1, q); }
          d.w00 = A->w00s + t;
          d.w00o = A->w00s + t + 1;
        }
        G->h_desc[(size_t)s * R + r] = d;
      }
    for (int s = 0; s < 2 * len; ++s)
      for (int r = 0; r < R; ++r) {
        const qs_solver* S = G->runs[r];
        const AdjState* A = S->adj;
        RunDesc d = make_desc(G->runs[r]);
        const int step = s0 + s / 2;
        d.alpha_p = G->sched ? G->alpha_t + (size_t)r * nsteps + step : nullptr;
        for (int q = 0; q < 7; ++q) d.minp[q] = slot(S, s, q);
        d.w00 = A->w00s + s;
        for (int c = 0; c < 3; ++c) { d.lam_[c] = A->lam_[c]; d.acc_[c] = A->acc_[c]; d.uout[c] = A->lam_[c]; }
        d.abar = A->abar;
        d.arow = A->arow + (size_t)step * m->Ny;
        d.w00bar = A->w00bar;
        d.scratch = G->scratch;
        G->h_desc[adj0 + (size_t)s * R + r] = d;
      }
    QS_TRY(upload(G, (int)(adj0 + (size_t)2 * len * R)));
    const RunDesc* D = G->d_desc;
    const RunDesc* AD = G->d_desc + adj0;
    auto at = [&](const RunDesc* base, int s) { return base + (size_t)s * R; };
    for (auto* S : G->runs)
      for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, S->old_[c], S->adj->ckpt[3 * seg + c]));
    // ---- recompute the segment, recording every stage's inverse set -------
    QS_TRY(launch_col(G, at(D, 0), MODE_PRO));
    for (int n = 0; n < len; ++n) {
      QS_TRY(launch_row(G, at(D, 1 + 2 * n)));
      QS_TRY(launch_col(G, at(D, 1 + 2 * n), MODE_S1));
      QS_TRY(launch_row(G, at(D, 2 + 2 * n)));
      QS_TRY(launch_col(G, at(D, 2 + 2 * n), MODE_S2));
    }
    // ---- reverse sweep over the segment -------------------------------------
    QS_TRY(launch_acol(G, at(AD, 2 * len - 1), AMODE_START));
    for (int n = len - 1; n >= 0; --n) {
      QS_TRY(launch_arow(G, at(AD, 2 * n + 1)));                 // N'(c2)^T
      QS_TRY(launch_acol(G, at(AD, 2 * n + 1), AMODE_B));
      QS_TRY(launch_arow(G, at(AD, 2 * n)));                     // N'(c1)^T
      QS_TRY(launch_acol(G, at(AD, 2 * n), n > 0 ? AMODE_A : AMODE_END));
    }
  }
  // leave every run at the final state of the horizon
  for (auto* S : G->runs) {
    for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, S->old_[c], S->adj->ckpt[3 * nseg + c]));
    S->mixed_valid = false;
  }
  return QS_OK;
}

int square_check(const Mesh* m) {
  if (m->Nx != m->Ny) return fail(QS_ERR_UNSUPPORTED_SIZE, "adjoint: square meshes only");
  return QS_OK;
}

}  // namespace

// ---- used by qs_ensemble.hip --------------------------------------------------
AdjGroup* adj_group_new(const std::vector<qs_solver*>& runs) {
  AdjGroup* G = new AdjGroup();
  G->runs = runs;
  G->R = (int)runs.size();
  return G;
}
void adj_group_free(AdjGroup* G) { group_free(G); }
int adj_group_forward(AdjGroup* G, int nsteps, int k, const double* alpha_t) {
  QS_TRY(square_check(mesh_of(G)));
  return group_forward(G, nsteps, k, alpha_t);
}
int adj_group_backward(AdjGroup* G) {
  if (G->nsteps <= 0) return fail(QS_ERR_STATE, "adjoint: no forward to reverse");
  return group_backward(G);
}
int adj_group_steps(const AdjGroup* G) { return G->nsteps; }
size_t adj_group_bytes(const AdjGroup* G) {
  size_t b = G->bytes;
  for (auto* S : G->runs) b += S->adj ? S->adj->bytes : 0;
  return b;
}
// Loss of a run's current vorticity and its gradient as the run's seed
// (dL/dW in lam_[0], zero for r1, r2).
int adj_run_seed(qs_solver* S, int kind, double* loss) {
  const Mesh* m = S->mesh;
  const hipStream_t st = stream_of(m);
  AdjState* A = S->adj;
  if (!A) return fail(QS_ERR_STATE, "adjoint: no forward");
  QS_TRY(loss_launch(m, kind, S->old_[0], S->pitch, loss, A->lam_[0], S->pitch));
  QS_HIP(hipMemsetAsync(A->lam_[1], 0, hbytes(S), st));
  QS_HIP(hipMemsetAsync(A->lam_[2], 0, hbytes(S), st));
  return QS_OK;
}

}  // namespace qs

using namespace qs;

extern "C" {

int qs_adj_release(qs_solver* S) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  if (S->adj) {
    (void)hipStreamSynchronize(stream_of(S->mesh));
    run_free(S->adj);
    S->adj = nullptr;
  }
  return QS_OK;
}

int qs_adj_memory(qs_solver* S, size_t* bytes) {
  if (!S || !bytes) return fail(QS_ERR_INVALID_ARG, "null argument");
  *bytes = 0;
  if (S->adj) *bytes = S->adj->bytes + (S->adj->solo ? S->adj->solo->bytes : 0);
  return QS_OK;
}

int qs_adj_forward(qs_solver* S, int nsteps, int ckpt_every, const Qreal* alpha_t) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  if (nsteps <= 0) return fail(QS_ERR_INVALID_ARG, "qs_adj_forward: nsteps must be > 0");
  QS_TRY(square_check(S->mesh));
  const int k = ckpt_every <= 0 ? nsteps : std::min(ckpt_every, nsteps);
  QS_TRY(run_prepare(S, nsteps, k));
  if (!S->adj->solo) S->adj->solo = adj_group_new({S});
  return group_forward(S->adj->solo, nsteps, k, alpha_t);
}

int qs_adj_backward(qs_solver* S, const Qcomp* gW, const Qcomp* gR1, const Qcomp* gR2, Qreal* abar,
                    Qreal* abar_t, Qcomp* uW, Qcomp* uR1, Qcomp* uR2) {
  if (!S || !gW || !gR1 || !gR2) return fail(QS_ERR_INVALID_ARG, "qs_adj_backward: null argument");
  if (!S->adj || !S->adj->solo || S->adj->solo->nsteps <= 0)
    return fail(QS_ERR_STATE, "qs_adj_backward: no qs_adj_forward to reverse");
  AdjState* A = S->adj;
  AdjGroup* G = A->solo;
  const Mesh* m = S->mesh;
  const hipStream_t st = stream_of(m);
  const size_t prow = sizeof(double2) * S->pitch, row = sizeof(Qcomp) * m->Nxh;
  const Qcomp* g[3] = {gW, gR1, gR2};
  for (int c = 0; c < 3; ++c)
    QS_HIP(hipMemcpy2DAsync(A->lam_[c], prow, g[c], row, row, m->Ny, hipMemcpyDeviceToDevice, st));
  QS_TRY(group_backward(G));
  if (abar) QS_HIP(hipMemcpyAsync(abar, A->abar, sizeof(double) * (size_t)m->Nx * m->Ny, hipMemcpyDeviceToDevice, st));
  if (abar_t) QS_TRY(row_sums_launch(A->arow, G->nsteps, m->Ny, abar_t, st));
  Qcomp* u[3] = {uW, uR1, uR2};
  for (int c = 0; c < 3; ++c)
    if (u[c]) QS_HIP(hipMemcpy2DAsync(u[c], row, A->lam_[c], prow, row, m->Ny, hipMemcpyDeviceToDevice, st));
  return QS_OK;
}

int qs_adj_loss(const Mesh* m, int kind, const Qcomp* W, Qreal* loss, Qcomp* Wbar) {
  QS_TRY(check_mesh(m));
  if (!W || !loss || !Wbar) return fail(QS_ERR_INVALID_ARG, "qs_adj_loss: null argument");
  if (kind != 0 && kind != 1) return fail(QS_ERR_INVALID_ARG, "qs_adj_loss: kind %d", kind);
  return loss_launch(m, kind, (const double2*)W, m->Nxh, loss, (double2*)Wbar, m->Nxh);
}

}  // extern "C"
