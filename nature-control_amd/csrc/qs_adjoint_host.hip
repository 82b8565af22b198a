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
// over a descriptor array.  The forward records the last segment's stages as
// it runs them, so the reverse sweep starts without a recompute.  When the
// group's row jobs leave compute units idle, the stage slots are
// double-buffered and segment s-1 is recomputed on a second stream while
// segment s is reversed on the group's stream; otherwise the two alternate on
// the group's stream.
//
// Memory per run (H = Ny * pitch * 16 B): checkpoints 3 H each (one per
// segment plus the final state), stage slots 7 H each for 2k+1 stages of one
// segment per buffer, adjoint state 6 H (and 7 H of mixed arrays when double
// buffered), activity gradient Nx*Ny*8 B, per-row sums nsteps*Ny*8 B; per
// group the row-adjoint scratch grid * 96 KiB and (5 nsteps + segments)
// descriptors per run.
#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>

#include "qs_solver.hpp"

namespace qs {

struct AdjGroup {
  std::vector<qs_solver*> runs;   // one mesh size, one stream
  int R = 0;
  unsigned long id = 0;           // identifies the group in AdjState::recorded_by
  int nsteps = 0, k = 0;          // horizon and checkpoint interval of the last forward
  int nbuf = 1;                   // stage-slot buffers in use
  bool sched = false;
  double* alpha_t = nullptr;      // [R][nsteps] uniform activity per run and step
  int acap = 0;                   // steps alpha_t holds per run
  RunDesc* d_desc = nullptr;      // descriptor tables of the horizon (build_tables)
  size_t dcap = 0;
  std::vector<RunDesc> h_desc;
  std::vector<size_t> seg_base;   // first descriptor of each segment's recompute table
  int2* d_row = nullptr;          // (run, job)
  int2* d_col = nullptr;
  int nrow = 0, ncol = 0;
  double2* scratch = nullptr;     // [grid][6][1024]
  int grid = 0;
  hipStream_t side = nullptr;     // recompute stream (nbuf == 2)
  hipEvent_t fork = nullptr, rec_done[2] = {}, rev_done[2] = {};
  size_t bytes = 0;
};

namespace {

constexpr int kAdjRowGrid = 256;   // row adjoint blocks (one per CU; LDS-bound)

std::atomic<unsigned long> g_group_ids{0};

template <class T>
int dalloc(size_t* acc, T** p, size_t bytes) {
  if (hipMalloc((void**)p, bytes) != hipSuccess) return fail(QS_ERR_OOM, "adjoint: hipMalloc of %zu bytes", bytes);
  *acc += bytes;
  return QS_OK;
}

template <class T>
void dfree(size_t* acc, T** p, size_t bytes) {
  if (*p) {
    (void)hipFree(*p);
    *acc -= bytes;
  }
  *p = nullptr;
}

void run_free(AdjState* A);
void group_free(AdjGroup* G) {
  if (!G) return;
  if (G->side) (void)hipStreamSynchronize(G->side);
  for (void* p : {(void*)G->alpha_t, (void*)G->d_desc, (void*)G->d_row, (void*)G->d_col, (void*)G->scratch})
    if (p) (void)hipFree(p);
  for (hipEvent_t e : {G->fork, G->rec_done[0], G->rec_done[1], G->rev_done[0], G->rev_done[1]})
    if (e) (void)hipEventDestroy(e);
  if (G->side) (void)hipStreamDestroy(G->side);
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
int nseg_of(const AdjGroup* G) { return (G->nsteps + G->k - 1) / G->k; }
int seg_len(const AdjGroup* G, int seg) { return std::min(G->k, G->nsteps - seg * G->k); }

// Per-run memory for a horizon of nsteps with checkpoint interval k and nbuf
// stage-slot buffers: one allocation carved into 256-byte aligned arrays (many
// separate allocations of a few MB each round up to the allocator's
// granularity).
int run_prepare(qs_solver* S, int nsteps, int k, int nbuf) {
  const Mesh* m = S->mesh;
  AdjState* A = S->adj;
  const int nseg = (nsteps + k - 1) / k;
  if (A && A->kcap >= k && A->nckpt >= nseg + 1 && A->ncap >= nsteps && A->nbuf >= nbuf) return QS_OK;
  AdjGroup* solo = A ? A->solo : nullptr;
  if (A) A->solo = nullptr;
  run_free(A);
  S->adj = A = new AdjState();
  A->solo = solo;
  A->kcap = k;
  A->ncap = nsteps;
  A->nckpt = nseg + 1;
  A->nbuf = nbuf;
  const size_t hs = (hbytes(S) + 255) & ~(size_t)255;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t P = (size_t)m->Nx * m->Ny;
  const size_t ns = 2 * (size_t)k + 1;   // stages per buffer
  const size_t total = hs * (3 * (size_t)(nseg + 1) + 7 * ns * nbuf + 6 + (nbuf == 2 ? 7 : 0)) +
                       al(sizeof(double) * ns * nbuf) + al(sizeof(double) * P) +
                       al(sizeof(double) * (size_t)nsteps * m->Ny) + al(sizeof(double) * m->Ny);
  QS_TRY(dalloc(&A->bytes, &A->arena, total));
  char* p = (char*)A->arena;
  auto take = [&](size_t b) { char* q = p; p += b; return q; };
  A->ckpt.assign(3 * (size_t)(nseg + 1), nullptr);
  for (auto& q : A->ckpt) q = (double2*)take(hs);
  A->slots.assign(7 * ns * nbuf, nullptr);
  for (auto& q : A->slots) q = (double2*)take(hs);
  for (int c = 0; c < 3; ++c) {
    A->lam_[c] = (double2*)take(hs);
    A->acc_[c] = (double2*)take(hs);
  }
  if (nbuf == 2)
    for (auto& q : A->amix) q = (double2*)take(hs);
  A->w00s = (double*)take(al(sizeof(double) * ns * nbuf));
  A->abar = (double*)take(al(sizeof(double) * P));
  A->arow = (double*)take(al(sizeof(double) * (size_t)nsteps * m->Ny));
  A->w00bar = (double*)take(al(sizeof(double) * m->Ny));
  return QS_OK;
}

// Overlapped recompute pays off only while the reverse sweep leaves compute
// units idle: fewer row jobs per launch than CUs.  QS_ADJ_OVERLAP=0 / 1
// forces the choice.
int choose_nbuf(const AdjGroup* G) {
  if (const char* e = std::getenv("QS_ADJ_OVERLAP")) return std::atoi(e) != 0 ? 2 : 1;
  const int ncu = mesh_of(G)->priv->ncu;
  return ncu > 0 && G->nrow < ncu ? 2 : 1;
}

int group_prepare(AdjGroup* G, int nsteps, int k) {
  const int R = G->R;
  if (!G->d_row) {   // work lists (run, job), run-major, and the row-adjoint scratch
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
  }
  G->nbuf = choose_nbuf(G);
  for (auto* S : G->runs) QS_TRY(run_prepare(S, nsteps, k, G->nbuf));
  if (G->nbuf == 2 && !G->side) {
    QS_HIP(hipStreamCreateWithFlags(&G->side, hipStreamNonBlocking));
    for (hipEvent_t* e : {&G->fork, &G->rec_done[0], &G->rec_done[1], &G->rev_done[0], &G->rev_done[1]})
      QS_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
  }
  const int nseg = (nsteps + k - 1) / k;
  const size_t nd = (size_t)(5 * nsteps + nseg) * R;
  if (G->dcap < nd) {
    dfree(&G->bytes, &G->d_desc, sizeof(RunDesc) * G->dcap);
    G->dcap = 0;
    QS_TRY(dalloc(&G->bytes, &G->d_desc, sizeof(RunDesc) * nd));
    G->dcap = nd;
    G->h_desc.resize(nd);
  }
  if (G->acap < nsteps) {
    dfree(&G->bytes, &G->alpha_t, sizeof(double) * (size_t)R * G->acap);
    G->acap = 0;
    QS_TRY(dalloc(&G->bytes, &G->alpha_t, sizeof(double) * (size_t)R * nsteps));
    G->acap = nsteps;
  }
  return QS_OK;
}

int copy_pitched(const qs_solver* S, double2* dst, const double2* src, hipStream_t st) {
  QS_HIP(hipMemcpyAsync(dst, src, hbytes(S), hipMemcpyDeviceToDevice, st));
  return QS_OK;
}

// fin != 0: stage 2 follows (see launch_fused_row)
int launch_row(AdjGroup* G, const RunDesc* d, hipStream_t st, int fin) {   // persistent: one block per CU
  const Mesh* m = mesh_of(G);
  const int grid = m->priv->ncu > 0 ? std::min(G->nrow, m->priv->ncu) : G->nrow;
  return launch_fused_row(m->Nx, d, G->d_row, G->nrow, grid, m->priv->tw1024, st, fin);
}
int launch_col(AdjGroup* G, const RunDesc* d, int mode, hipStream_t st) {
  const Mesh* m = mesh_of(G);
  return launch_fused_col(m->Ny, mode, d, G->d_col, G->ncol, G->ncol, m->priv->tw1024, st);
}
int launch_arow(AdjGroup* G, const RunDesc* d, hipStream_t st) {
  const Mesh* m = mesh_of(G);
  return launch_adj_row(m->Nx, d, G->d_row, G->nrow, G->grid, m->priv->tw1024, st);
}
int launch_acol(AdjGroup* G, const RunDesc* d, int amode, hipStream_t st) {
  const Mesh* m = mesh_of(G);
  return launch_adj_col(m->Ny, amode, d, G->d_col, G->ncol, G->ncol, m->priv->tw1024, st);
}

double2* slot(const qs_solver* S, int b, int s, int q) {
  const AdjState* A = S->adj;
  return A->slots[7 * ((size_t)(2 * A->kcap + 1) * b + s) + q];
}
double* w00_slot(const qs_solver* S, int b, int s) {
  const AdjState* A = S->adj;
  return A->w00s + (size_t)(2 * A->kcap + 1) * b + s;
}

// Descriptor tables of the horizon, laid out [stage][run]:
//   [0, nsteps R)         the plain forward steps;
//   from seg_base[seg]    segment seg's recompute table ([0] prologue, [1 + t]
//                         stage t: row input slot t, column output slot t + 1,
//                         slots of buffer seg % nbuf), then its reverse table
//                         (adjoint of stage t, reading slot t).
void build_tables(AdjGroup* G) {
  const int R = G->R, nsteps = G->nsteps, k = G->k, nseg = nseg_of(G);
  const Mesh* m = mesh_of(G);
  auto alpha_at = [&](int r, int step) -> const double* {
    return G->sched ? G->alpha_t + (size_t)r * nsteps + step : nullptr;
  };
  for (int n = 0; n < nsteps; ++n)
    for (int r = 0; r < R; ++r) {
      RunDesc d = make_desc(G->runs[r]);
      d.alpha_p = alpha_at(r, n);
      G->h_desc[(size_t)n * R + r] = d;
    }
  G->seg_base.assign(nseg, 0);
  size_t base = (size_t)nsteps * R;
  for (int seg = 0; seg < nseg; ++seg) {
    const int s0 = seg * k, len = seg_len(G, seg), b = seg % G->nbuf;
    G->seg_base[seg] = base;
    for (int s = 0; s <= 2 * len; ++s)
      for (int r = 0; r < R; ++r) {
        const qs_solver* S = G->runs[r];
        RunDesc d = make_desc(G->runs[r]);
        d.alpha_p = alpha_at(r, s0 + (s > 0 ? (s - 1) / 2 : 0));
        if (s == 0) {
          for (int q = 0; q < 7; ++q) d.mout[q] = slot(S, b, 0, q);
          d.w00o = w00_slot(S, b, 0);
        } else {
          const int t = s - 1;
          for (int q = 0; q < 7; ++q) { d.minp[q] = slot(S, b, t, q); d.mout[q] = slot(S, b, t + 1, q); }
          d.w00 = w00_slot(S, b, t);
          d.w00o = w00_slot(S, b, t + 1);
        }
        G->h_desc[base + (size_t)s * R + r] = d;
      }
    base += (size_t)(2 * len + 1) * R;
    for (int s = 0; s < 2 * len; ++s)
      for (int r = 0; r < R; ++r) {
        const qs_solver* S = G->runs[r];
        const AdjState* A = S->adj;
        RunDesc d = make_desc(G->runs[r]);
        const int step = s0 + s / 2;
        d.alpha_p = alpha_at(r, step);
        if (G->nbuf == 2)   // the recompute running beside this sweep uses the run's own mixed arrays
          for (int q = 0; q < 7; ++q) d.mix[q] = A->amix[q];
        for (int q = 0; q < 7; ++q) d.minp[q] = slot(S, b, s, q);
        d.w00 = w00_slot(S, b, s);
        for (int c = 0; c < 3; ++c) { d.lam_[c] = A->lam_[c]; d.acc_[c] = A->acc_[c]; d.uout[c] = A->lam_[c]; }
        d.abar = A->abar;
        d.arow = A->arow + (size_t)step * m->Ny;
        d.w00bar = A->w00bar;
        d.scratch = G->scratch;
        G->h_desc[base + (size_t)s * R + r] = d;
      }
    base += (size_t)(2 * len) * R;
  }
}

int upload(AdjGroup* G, size_t n) {
  QS_HIP(hipMemcpyAsync(G->d_desc, G->h_desc.data(), sizeof(RunDesc) * n, hipMemcpyHostToDevice, stream_g(G)));
  // the host tables are rewritten by the next forward: wait for the copy
  QS_HIP(hipStreamSynchronize(stream_g(G)));
  return QS_OK;
}

// Segment seg from the state in every run's old_, recording every stage's
// inverse set in the slots of buffer seg % nbuf.
int record_segment(AdjGroup* G, int seg, hipStream_t st) {
  const RunDesc* D = G->d_desc + G->seg_base[seg];
  auto at = [&](int s) { return D + (size_t)s * G->R; };
  QS_TRY(launch_col(G, at(0), MODE_PRO, st));
  for (int n = 0; n < seg_len(G, seg); ++n) {
    QS_TRY(launch_row(G, at(1 + 2 * n), st, 0));
    QS_TRY(launch_col(G, at(1 + 2 * n), MODE_S1, st));
    QS_TRY(launch_row(G, at(2 + 2 * n), st, 1));
    QS_TRY(launch_col(G, at(2 + 2 * n), MODE_S2, st));
  }
  return QS_OK;
}

// Reverse sweep over segment seg (its stages recorded in buffer seg % nbuf):
// lam_ holds dL/d(state at the segment's end) on entry, dL/d(state at its
// start) on return.
int reverse_segment(AdjGroup* G, int seg, hipStream_t st) {
  const int len = seg_len(G, seg);
  const RunDesc* AD = G->d_desc + G->seg_base[seg] + (size_t)(2 * len + 1) * G->R;
  auto at = [&](int s) { return AD + (size_t)s * G->R; };
  QS_TRY(launch_acol(G, at(2 * len - 1), AMODE_START, st));
  for (int n = len - 1; n >= 0; --n) {
    QS_TRY(launch_arow(G, at(2 * n + 1), st));                 // N'(c2)^T
    QS_TRY(launch_acol(G, at(2 * n + 1), AMODE_B, st));
    QS_TRY(launch_arow(G, at(2 * n), st));                     // N'(c1)^T
    QS_TRY(launch_acol(G, at(2 * n), n > 0 ? AMODE_A : AMODE_END, st));
  }
  return QS_OK;
}

// Forward: nsteps steps of every run with checkpoints every k steps (and the
// final state); the last segment records its stages.  alpha_t: device
// [R][nsteps] or null.
int group_forward(AdjGroup* G, int nsteps, int k, const double* alpha_t) {
  Range rg("adjoint forward");
  QS_TRY(group_prepare(G, nsteps, k));
  const int R = G->R;
  G->nsteps = nsteps;
  G->k = k;
  G->sched = alpha_t != nullptr;
  const hipStream_t st = stream_g(G);
  if (alpha_t)
    QS_HIP(hipMemcpyAsync(G->alpha_t, alpha_t, sizeof(double) * (size_t)R * nsteps, hipMemcpyDeviceToDevice, st));
  build_tables(G);
  const int nseg = nseg_of(G), last = nseg - 1;
  QS_TRY(upload(G, G->seg_base[last] + (size_t)(4 * seg_len(G, last) + 1) * R));
  auto checkpoint = [&](int c0) -> int {
    for (auto* S : G->runs)
      for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, S->adj->ckpt[3 * c0 + c], S->old_[c], st));
    return QS_OK;
  };
  for (int n = 0; n < last * k; ++n) {
    if (n % k == 0) QS_TRY(checkpoint(n / k));
    const RunDesc* d = G->d_desc + (size_t)n * R;
    if (n == 0) QS_TRY(launch_col(G, d, MODE_PRO, st));
    QS_TRY(launch_row(G, d, st, 0));
    QS_TRY(launch_col(G, d, MODE_S1, st));
    QS_TRY(launch_row(G, d, st, 1));
    QS_TRY(launch_col(G, d, MODE_S2, st));
  }
  QS_TRY(checkpoint(last));
  QS_TRY(record_segment(G, last, st));
  QS_TRY(checkpoint(nseg));
  for (auto* S : G->runs) {
    S->mixed_valid = false;   // the last segment's inverse sets went to the stage slots
    S->adj->recorded_by = G->id;
  }
  return QS_OK;
}

// Reverse sweep.  The seeds dL/d(final state) are in every run's lam_.
int group_backward(AdjGroup* G) {
  Range rg("adjoint backward");
  const int nsteps = G->nsteps, nseg = nseg_of(G), nbuf = G->nbuf;
  const Mesh* m = mesh_of(G);
  const hipStream_t st = stream_g(G);
  const hipStream_t side = nbuf == 2 ? G->side : st;
  bool have_last = true;   // the slots still hold the last segment the forward recorded
  for (auto* S : G->runs) {
    QS_HIP(hipMemsetAsync(S->adj->abar, 0, sizeof(double) * (size_t)m->Nx * m->Ny, st));
    QS_HIP(hipMemsetAsync(S->adj->arow, 0, sizeof(double) * (size_t)nsteps * m->Ny, st));
    have_last = have_last && S->adj->recorded_by == G->id;
  }
  if (nbuf == 2) {
    QS_HIP(hipEventRecord(G->fork, st));
    QS_HIP(hipStreamWaitEvent(side, G->fork, 0));
  }
  bool rec_pending[2] = {false, false}, rev_pending[2] = {false, false};
  auto recompute = [&](int seg) -> int {
    Range r("adjoint recompute segment %d", seg);
    const int b = seg % nbuf;
    if (nbuf == 2 && rev_pending[b]) QS_HIP(hipStreamWaitEvent(side, G->rev_done[b], 0));
    for (auto* S : G->runs)
      for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, S->old_[c], S->adj->ckpt[3 * seg + c], side));
    QS_TRY(record_segment(G, seg, side));
    if (nbuf == 2) {
      QS_HIP(hipEventRecord(G->rec_done[b], side));
      rec_pending[b] = true;
    }
    return QS_OK;
  };
  auto reverse = [&](int seg) -> int {
    Range r("adjoint reverse segment %d", seg);
    const int b = seg % nbuf;
    if (nbuf == 2 && rec_pending[b]) QS_HIP(hipStreamWaitEvent(st, G->rec_done[b], 0));
    QS_TRY(reverse_segment(G, seg, st));
    if (nbuf == 2) {
      QS_HIP(hipEventRecord(G->rev_done[b], st));
      rev_pending[b] = true;
    }
    return QS_OK;
  };
  if (!have_last) QS_TRY(recompute(nseg - 1));
  for (int seg = nseg - 1; seg >= 0; --seg) {
    if (nbuf == 2 && seg > 0) QS_TRY(recompute(seg - 1));   // overlaps the reverse of seg
    QS_TRY(reverse(seg));
    if (nbuf == 1 && seg > 0) QS_TRY(recompute(seg - 1));
  }
  // leave every run at the final state of the horizon (every recompute on the
  // side stream precedes a reverse sweep that waited for it)
  for (auto* S : G->runs) {
    for (int c = 0; c < 3; ++c) QS_TRY(copy_pitched(S, S->old_[c], S->adj->ckpt[3 * nseg + c], st));
    S->mixed_valid = false;
    S->adj->recorded_by = nseg <= nbuf ? G->id : 0;
  }
  return QS_OK;
}

// The adjoint's row kernels are instantiated by Nx, its column kernels by Ny:
// both sides powers of two, 16..1024 (not necessarily equal).
int size_check(const Mesh* m) {
  if (!fused_side(m->Nx) || !fused_side(m->Ny))
    return fail(QS_ERR_UNSUPPORTED_SIZE, "adjoint: mesh %dx%d (sides powers of two, 16..1024)", m->Nx, m->Ny);
  return QS_OK;
}

}  // namespace

// ---- used by qs_ensemble.hip --------------------------------------------------
AdjGroup* adj_group_new(const std::vector<qs_solver*>& runs) {
  AdjGroup* G = new AdjGroup();
  G->id = ++g_group_ids;
  G->runs = runs;
  G->R = (int)runs.size();
  return G;
}
void adj_group_free(AdjGroup* G) { group_free(G); }
int adj_group_forward(AdjGroup* G, int nsteps, int k, const double* alpha_t) {
  QS_TRY(size_check(mesh_of(G)));
  return group_forward(G, nsteps, k, alpha_t);
}
int adj_group_backward(AdjGroup* G) {
  if (G->nsteps <= 0) return fail(QS_ERR_STATE, "adjoint: no forward to reverse");
  return group_backward(G);
}
int adj_group_steps(const AdjGroup* G) { return G->nsteps; }
// Frees a run's adjoint memory without touching its Mesh (the solver's
// destroy path: the device is synchronised first, the Mesh may be gone).
void adj_state_free(qs_solver* S) {
  run_free(S->adj);
  S->adj = nullptr;
}
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
  QS_TRY(size_check(S->mesh));
  const int k = ckpt_every <= 0 ? nsteps : std::min(ckpt_every, nsteps);
  if (!S->adj) S->adj = new AdjState();   // sized by the group's first forward
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
