// qs_ensemble.hip -- many independent IFRK2 runs of uneven mesh sizes and
// parameters advanced together; a captured hipGraph per block of steps and a
// project-written autotuner for the launch shapes.
//
// No reference counterpart: the reference advances one run per process
// (InertialFriction/friction/friction.cu).  Each run is a qs_solver (its own
// state, mixed arrays, parameters); runs of one mesh size form a size class.
// One ensemble step launches, per class, the fused kernels of qs_fused.hip
// over a work list of (run, job) pairs -- row pass, column stage 1, row pass,
// column stage 2 -- so every kernel of a class covers all its runs.  The
// classes are independent: with `concurrent` they are parallel branches of the
// captured graph (fork / join through events), so small classes fill the CUs
// that large ones leave idle.  The kernels are persistent: a launch of B blocks
// loops over the work list (B = one per job, or fewer).  With `merged`, one
// persistent kernel per phase (qs_fused_mix.hip) covers the jobs of every class
// through a single work list of (run, job) pairs, largest side first.
#include <algorithm>
#include <chrono>
#include <cstring>

#include "qs_solver.hpp"

namespace qs {
namespace {

struct EnsClass {
  int N = 0;
  Mesh* mesh = nullptr;
  std::vector<int> members;
  int2* d_row = nullptr;
  int2* d_col = nullptr;
  int nrow = 0, ncol = 0;
  int cap_row = 0, cap_col = 0;     // entries allocated in d_row, d_col
  hipStream_t side = nullptr;
  hipEvent_t done = nullptr;       // join event of direct launches
  hipEvent_t cap_done = nullptr;   // join event used only inside graph captures
  bool valid = false;
};

}  // namespace
}  // namespace qs

struct qs_ensemble {
  std::vector<qs_solver*> runs;
  std::vector<qs::EnsClass> cls;
  std::vector<int> cls_of;                 // run -> class index
  qs::RunDesc* d_desc = nullptr;           // [n] run descriptors
  hipStream_t stream = nullptr;            // the ensemble's stream (null = default)
  hipEvent_t fork = nullptr;               // fork event of direct launches
  hipEvent_t cap_fork = nullptr;           // fork event used only inside graph captures
  hipStream_t cap = nullptr;               // graph capture stream
  qs_ensemble_config cfg{};
  hipGraphExec_t exec = nullptr;
  int exec_steps = 0;
  int order_built = -1;
  int merged_built = -1;
  int lanes_built[8] = {-1, -1, -1, -1, -1, -1, -1, -1};   // elements per lane of each class's inverse sets
  int2* d_row_all = nullptr;               // merged work lists (every class)
  int2* d_col_all = nullptr;
  int nrow_all = 0, ncol_all = 0, cap_row_all = 0, cap_col_all = 0;
  qs::AdjGroup* adj = nullptr;             // batched adjoint (one mesh size)
};

namespace qs {
namespace {

int upload_descs(qs_ensemble* E) {
  std::vector<RunDesc> h(E->runs.size());
  for (size_t r = 0; r < E->runs.size(); ++r) h[r] = make_desc(E->runs[r]);
  QS_HIP(hipMemcpyAsync(E->d_desc, h.data(), sizeof(RunDesc) * h.size(), hipMemcpyHostToDevice, E->stream));
  QS_HIP(hipStreamSynchronize(E->stream));   // h is a local
  return QS_OK;
}

// Elements per lane that class ci's kernels use under the current config.
int class_el(const qs_ensemble* E, int ci) {
  if (E->cfg.merged) return 16;
  return E->cfg.lanes[ci] ? E->cfg.lanes[ci] : wave_e(E->cls[ci].N);
}

int upload_list(std::vector<int2>& v, int2** d, int* cap, hipStream_t st) {
  if ((int)v.size() > *cap) {
    (void)hipStreamSynchronize(st);
    if (*d) (void)hipFree(*d);
    *d = nullptr;
    *cap = 0;
    if (hipMalloc((void**)d, sizeof(int2) * v.size()) != hipSuccess) return fail(QS_ERR_OOM, "ensemble work lists");
    *cap = (int)v.size();
  }
  QS_HIP(hipMemcpy(*d, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice));
  return QS_OK;
}

// Work lists of every class in the configured order (and, merged, one list of
// all classes, largest side first).
int build_work(qs_ensemble* E) {
  bool same = E->order_built == E->cfg.order && E->merged_built == E->cfg.merged;
  for (size_t ci = 0; ci < E->cls.size(); ++ci) same = same && E->lanes_built[ci] == class_el(E, (int)ci);
  if (same) return QS_OK;
  std::vector<int2> row_all, col_all;
  for (int ci = (int)E->cls.size() - 1; ci >= 0; --ci) {
    EnsClass& c = E->cls[ci];
    std::vector<int2> row, col;
    const RunDesc d0 = make_desc(E->runs[c.members[0]]);
    const int el = class_el(E, ci);
    const int nr = row_jobs(d0, el), nc = col_jobs(d0, el);
    if (E->cfg.order == 0) {
      for (int r : c.members) {
        for (int j = 0; j < nr; ++j) row.push_back(make_int2(r, j));
        for (int j = 0; j < nc; ++j) col.push_back(make_int2(r, j));
      }
    } else {
      for (int j = 0; j < nr; ++j)
        for (int r : c.members) row.push_back(make_int2(r, j));
      for (int j = 0; j < nc; ++j)
        for (int r : c.members) col.push_back(make_int2(r, j));
    }
    QS_TRY(upload_list(row, &c.d_row, &c.cap_row, E->stream));
    QS_TRY(upload_list(col, &c.d_col, &c.cap_col, E->stream));
    if (E->lanes_built[ci] != el) c.valid = false;   // inverse sets from this layout's prologue
    c.nrow = (int)row.size();
    c.ncol = (int)col.size();
    E->lanes_built[ci] = el;
    row_all.insert(row_all.end(), row.begin(), row.end());
    col_all.insert(col_all.end(), col.begin(), col.end());
  }
  if (E->cfg.merged) {
    QS_TRY(upload_list(row_all, &E->d_row_all, &E->cap_row_all, E->stream));
    QS_TRY(upload_list(col_all, &E->d_col_all, &E->cap_col_all, E->stream));
    E->nrow_all = (int)row_all.size();
    E->ncol_all = (int)col_all.size();
  }
  E->order_built = E->cfg.order;
  E->merged_built = E->cfg.merged;
  return QS_OK;
}

int blocks(int want, int nwork) { return want > 0 ? std::min(want, nwork) : nwork; }

// One IFRK2 step of class c on stream st.
int class_step(qs_ensemble* E, int ci, hipStream_t st) {
  EnsClass& c = E->cls[ci];
  const double2* tw = c.mesh->priv->tw1024;
  const int rb = blocks(E->cfg.row_blocks[ci], c.nrow), cb = blocks(E->cfg.col_blocks[ci], c.ncol);
  const int el = E->cfg.lanes[ci];
  QS_TRY(launch_fused_row_l(c.N, el, E->d_desc, c.d_row, c.nrow, rb, tw, st));
  QS_TRY(launch_fused_col_l(c.N, el, MODE_S1, E->d_desc, c.d_col, c.ncol, cb, tw, st));
  QS_TRY(launch_fused_row_l(c.N, el, E->d_desc, c.d_row, c.nrow, rb, tw, st, 1));
  QS_TRY(launch_fused_col_l(c.N, el, MODE_S2, E->d_desc, c.d_col, c.ncol, cb, tw, st));
  return QS_OK;
}

// The merged launches' block counts: row pass one block per CU (its LDS
// admits one) unless configured, column stages one per job unless configured.
int merged_row_grid(const qs_ensemble* E) {
  const int ncu = E->cls[0].mesh->priv->ncu;
  const int want = E->cfg.row_blocks[0] > 0 ? E->cfg.row_blocks[0] : (ncu > 0 ? ncu : E->nrow_all);
  return std::min(want, E->nrow_all);
}
int merged_col_grid(const qs_ensemble* E) { return blocks(E->cfg.col_blocks[0], E->ncol_all); }

// One step of every class as four launches on `st` (merged).
int merged_step(qs_ensemble* E, hipStream_t st) {
  const double2* tw = E->cls[0].mesh->priv->tw1024;
  const int rb = merged_row_grid(E), cb = merged_col_grid(E);
  QS_TRY(launch_mix_row(E->d_desc, E->d_row_all, E->nrow_all, rb, tw, st, 0));
  QS_TRY(launch_mix_col(MODE_S1, E->d_desc, E->d_col_all, E->ncol_all, cb, tw, st));
  QS_TRY(launch_mix_row(E->d_desc, E->d_row_all, E->nrow_all, rb, tw, st, 1));
  QS_TRY(launch_mix_col(MODE_S2, E->d_desc, E->d_col_all, E->ncol_all, cb, tw, st));
  return QS_OK;
}

// One step of every class on stream `main`: merged launches, parallel
// branches (fork / join through events) or all on `main`.  Captures use their
// own events, so no event is shared between a graph and directly launched
// work.
int ensemble_step_direct(qs_ensemble* E, hipStream_t main, bool capturing) {
  if (E->cfg.merged) return merged_step(E, main);
  const int nc = (int)E->cls.size();
  if (!E->cfg.concurrent || nc == 1) {
    for (int ci = 0; ci < nc; ++ci) QS_TRY(class_step(E, ci, main));
    return QS_OK;
  }
  hipEvent_t fork = capturing ? E->cap_fork : E->fork;
  QS_HIP(hipEventRecord(fork, main));
  for (int ci = 0; ci < nc; ++ci) {
    EnsClass& c = E->cls[ci];
    QS_HIP(hipStreamWaitEvent(c.side, fork, 0));
    QS_TRY(class_step(E, ci, c.side));
    QS_HIP(hipEventRecord(capturing ? c.cap_done : c.done, c.side));
  }
  for (auto& c : E->cls) QS_HIP(hipStreamWaitEvent(main, capturing ? c.cap_done : c.done, 0));
  return QS_OK;
}

// Destroys the executable graph once no launch of it can still be running.
void drop_graph(qs_ensemble* E) {
  if (E->exec) {
    (void)hipStreamSynchronize(E->stream);
    (void)hipGraphExecDestroy(E->exec);
  }
  E->exec = nullptr;
  E->exec_steps = 0;
}

// Captured on the ensemble's own capture stream (the null stream cannot be
// captured); the graph is then launched on the ensemble's stream.
int capture_graph(qs_ensemble* E, int G) {
  drop_graph(E);
  hipGraph_t g = nullptr;
  QS_HIP(hipStreamBeginCapture(E->cap, hipStreamCaptureModeThreadLocal));
  int st = QS_OK;
  for (int k = 0; k < G && st == QS_OK; ++k) st = ensemble_step_direct(E, E->cap, true);
  const hipError_t e = hipStreamEndCapture(E->cap, &g);
  if (st != QS_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess) return fail(QS_ERR_HIP, "ensemble graph capture: %s", hipGetErrorString(e));
  const hipError_t ei = hipGraphInstantiate(&E->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(QS_ERR_HIP, "ensemble graph instantiate: %s", hipGetErrorString(ei));
  E->exec_steps = G;
  return QS_OK;
}

int prologues(qs_ensemble* E) {
  if (E->cfg.merged) {
    // one merged prologue when any class is stale: recomputing a valid class's
    // inverse sets from its state gives the same values again
    bool stale = false;
    for (auto& c : E->cls) stale = stale || !c.valid;
    if (!stale) return QS_OK;
    QS_TRY(launch_mix_col(MODE_PRO, E->d_desc, E->d_col_all, E->ncol_all, merged_col_grid(E),
                          E->cls[0].mesh->priv->tw1024, E->stream));
    for (auto& c : E->cls) c.valid = true;
    return QS_OK;
  }
  for (size_t ci = 0; ci < E->cls.size(); ++ci) {
    EnsClass& c = E->cls[ci];
    if (c.valid) continue;
    QS_TRY(launch_fused_col_l(c.N, E->cfg.lanes[ci], MODE_PRO, E->d_desc, c.d_col, c.ncol,
                              blocks(E->cfg.col_blocks[ci], c.ncol), c.mesh->priv->tw1024, E->stream));
    c.valid = true;
  }
  return QS_OK;
}

int ensemble_steps(qs_ensemble* E, int nsteps) {
  QS_TRY(build_work(E));
  QS_TRY(prologues(E));
  const int G = E->cfg.graph_steps;
  int done = 0;
  if (G > 0 && nsteps >= G) {
    if (E->exec_steps != G) QS_TRY(capture_graph(E, G));
    for (; done + G <= nsteps; done += G) QS_HIP(hipGraphLaunch(E->exec, E->stream));
  }
  for (; done < nsteps; ++done) QS_TRY(ensemble_step_direct(E, E->stream, false));
  return QS_OK;
}

void ensemble_free(qs_ensemble* E) {
  if (!E) return;
  (void)hipDeviceSynchronize();
  drop_graph(E);
  adj_group_free(E->adj);
  for (auto* s : E->runs) qs_solver_destroy(s);
  if (E->d_row_all) (void)hipFree(E->d_row_all);
  if (E->d_col_all) (void)hipFree(E->d_col_all);
  for (auto& c : E->cls) {
    if (c.d_row) (void)hipFree(c.d_row);
    if (c.d_col) (void)hipFree(c.d_col);
    if (c.side) (void)hipStreamDestroy(c.side);
    if (c.done) (void)hipEventDestroy(c.done);
    if (c.cap_done) (void)hipEventDestroy(c.cap_done);
    qs_mesh_destroy(c.mesh);
  }
  if (E->fork) (void)hipEventDestroy(E->fork);
  if (E->cap_fork) (void)hipEventDestroy(E->cap_fork);
  if (E->cap) (void)hipStreamDestroy(E->cap);
  if (E->d_desc) (void)hipFree(E->d_desc);
  delete E;
}

int check_run(const qs_ensemble* E, int r) {
  if (!E) return fail(QS_ERR_INVALID_ARG, "null ensemble");
  if (r < 0 || r >= (int)E->runs.size()) return fail(QS_ERR_INVALID_ARG, "run %d of %zu", r, E->runs.size());
  return QS_OK;
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

int qs_ensemble_create(int n, const int* N, const qs_params* params, qs_ensemble** out) {
  if (!out || !N || !params || n <= 0) return fail(QS_ERR_INVALID_ARG, "qs_ensemble_create: bad arguments");
  *out = nullptr;
  qs_ensemble* E = new qs_ensemble();
  auto bail = [&](int st) { ensemble_free(E); return st; };
  // size classes in ascending mesh size
  std::vector<int> sizes(N, N + n);
  std::sort(sizes.begin(), sizes.end());
  sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
  if (sizes.size() > 8) return bail(fail(QS_ERR_INVALID_ARG, "qs_ensemble_create: more than 8 mesh sizes"));
  for (int s : sizes)
    if (!fused_side(s)) return bail(fail(QS_ERR_UNSUPPORTED_SIZE, "qs_ensemble_create: mesh %d (powers of two, 16..1024)", s));
  for (int s : sizes) {
    EnsClass c;
    c.N = s;
    const int st = qs_mesh_create(s, s, 2 * M_PI, 2 * M_PI, 16, &c.mesh);
    E->cls.push_back(c);
    if (st != QS_OK) return bail(st);
    if (hipStreamCreateWithFlags(&E->cls.back().side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&E->cls.back().done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&E->cls.back().cap_done, hipEventDisableTiming) != hipSuccess)
      return bail(fail(QS_ERR_HIP, "qs_ensemble_create: streams"));
  }
  if (hipEventCreateWithFlags(&E->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&E->cap_fork, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&E->cap, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(QS_ERR_HIP, "qs_ensemble_create: events"));
  E->runs.assign(n, nullptr);
  E->cls_of.assign(n, 0);
  for (int r = 0; r < n; ++r) {
    const int ci = (int)(std::lower_bound(sizes.begin(), sizes.end(), N[r]) - sizes.begin());
    E->cls_of[r] = ci;
    E->cls[ci].members.push_back(r);
    const int st = qs_solver_create(E->cls[ci].mesh, &params[r], &E->runs[r]);
    if (st != QS_OK) return bail(st);
  }
  if (hipMalloc((void**)&E->d_desc, sizeof(RunDesc) * n) != hipSuccess)
    return bail(fail(QS_ERR_OOM, "qs_ensemble_create: descriptors"));
  E->cfg.graph_steps = 10;
  E->cfg.concurrent = 1;
  E->cfg.order = 0;
  int st = upload_descs(E);
  if (st == QS_OK) st = build_work(E);
  if (st != QS_OK) return bail(st);
  *out = E;
  return QS_OK;
}

int qs_ensemble_destroy(qs_ensemble* E) {
  ensemble_free(E);
  return QS_OK;
}

int qs_ensemble_set_stream(qs_ensemble* E, void* stream) {
  if (!E) return fail(QS_ERR_INVALID_ARG, "null ensemble");
  E->stream = (hipStream_t)stream;
  for (auto& c : E->cls) QS_TRY(qs_mesh_set_stream(c.mesh, stream));
  drop_graph(E);
  return QS_OK;
}

int qs_ensemble_size(const qs_ensemble* E, int* n) {
  if (!E || !n) return fail(QS_ERR_INVALID_ARG, "null argument");
  *n = (int)E->runs.size();
  return QS_OK;
}

int qs_ensemble_solver(qs_ensemble* E, int r, qs_solver** out) {
  QS_TRY(check_run(E, r));
  if (!out) return fail(QS_ERR_INVALID_ARG, "null out");
  *out = E->runs[r];
  return QS_OK;
}

int qs_ensemble_set_state(qs_ensemble* E, int r, const Qcomp* w, const Qcomp* r1, const Qcomp* r2) {
  QS_TRY(check_run(E, r));
  QS_TRY(qs_solver_set_state(E->runs[r], w, r1, r2));
  E->cls[E->cls_of[r]].valid = false;
  return QS_OK;
}

int qs_ensemble_get_state(qs_ensemble* E, int r, Qcomp* w, Qcomp* r1, Qcomp* r2) {
  QS_TRY(check_run(E, r));
  return qs_solver_get_state(E->runs[r], w, r1, r2);
}

int qs_ensemble_set_activity(qs_ensemble* E, int r, const Qreal* Ra) {
  QS_TRY(check_run(E, r));
  QS_TRY(qs_solver_set_activity(E->runs[r], Ra));
  return upload_descs(E);
}

int qs_ensemble_step(qs_ensemble* E, int nsteps) {
  if (!E) return fail(QS_ERR_INVALID_ARG, "null ensemble");
  if (nsteps < 0) return fail(QS_ERR_INVALID_ARG, "nsteps < 0");
  Range rg("qs_ensemble_step");
  return ensemble_steps(E, nsteps);
}

int qs_ensemble_get_config(const qs_ensemble* E, qs_ensemble_config* cfg) {
  if (!E || !cfg) return fail(QS_ERR_INVALID_ARG, "null argument");
  *cfg = E->cfg;
  return QS_OK;
}

int qs_ensemble_set_config(qs_ensemble* E, const qs_ensemble_config* cfg) {
  if (!E || !cfg) return fail(QS_ERR_INVALID_ARG, "null argument");
  if (cfg->graph_steps < 0 || (cfg->order != 0 && cfg->order != 1) || (cfg->merged != 0 && cfg->merged != 1))
    return fail(QS_ERR_INVALID_ARG, "qs_ensemble_set_config: graph_steps %d order %d merged %d", cfg->graph_steps,
                cfg->order, cfg->merged);
  if (cfg->merged)
    for (auto& c : E->cls)
      if (!mix_side(c.N)) return fail(QS_ERR_UNSUPPORTED_SIZE, "qs_ensemble_set_config: merged kernels, side %d", c.N);
  for (size_t ci = 0; ci < E->cls.size(); ++ci) {
    if (cfg->lanes[ci] == 0) continue;
    int el[4];
    const int n = lane_layouts(E->cls[ci].N, el);
    bool ok = false;
    for (int k = 0; k < n; ++k) ok = ok || el[k] == cfg->lanes[ci];
    if (!ok)
      return fail(QS_ERR_UNSUPPORTED_SIZE, "qs_ensemble_set_config: %d elements per lane at side %d", cfg->lanes[ci],
                  E->cls[ci].N);
  }
  E->cfg = *cfg;
  drop_graph(E);
  return build_work(E);
}

// Coordinate-descent autotuner over the launch configuration: merged or
// per-class kernels, work-list order, branch concurrency, per-class lane
// layout of the per-wave FFT (its first-pass radix, rows per row job and
// column-tile width) and block counts (per class, or of the merged kernels),
// each candidate timed (HIP events on the ensemble's stream) over about
// `trial_steps` directly launched steps after a warm-up; then the graph depth,
// each candidate captured and timed through its graph launches.  The states
// are saved first and restored afterwards.
int qs_ensemble_tune(qs_ensemble* E, int trial_steps, qs_ensemble_config* best_out, double* best_ms) {
  if (!E) return fail(QS_ERR_INVALID_ARG, "null ensemble");
  if (trial_steps <= 0) return fail(QS_ERR_INVALID_ARG, "trial_steps must be > 0");
  const int n = (int)E->runs.size();
  // save states
  std::vector<double2*> saved(3 * (size_t)n, nullptr);
  int st = QS_OK;
  for (int r = 0; r < n && st == QS_OK; ++r) {
    qs_solver* S = E->runs[r];
    const size_t hs = sizeof(double2) * (size_t)S->pitch * S->mesh->Ny;
    for (int c = 0; c < 3 && st == QS_OK; ++c) {
      if (hipMalloc((void**)&saved[3 * r + c], hs) != hipSuccess) st = fail(QS_ERR_OOM, "tune: state copy");
      else if (hipMemcpyAsync(saved[3 * r + c], S->old_[c], hs, hipMemcpyDeviceToDevice, E->stream) != hipSuccess)
        st = fail(QS_ERR_HIP, "tune: state copy");
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (st == QS_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    st = fail(QS_ERR_HIP, "tune: events");
  // One candidate: steps through the candidate's captured graph (G > 0,
  // captured at its first launch, G steps per launch) or direct launches
  // (G = 0), after one warm-up launch.  Re-capturing is safe: set_config drops
  // the previous graph only once no launch of it can still run (drop_graph).
  auto measure = [&](qs_ensemble_config c, double* ms) -> int {
    QS_TRY(qs_ensemble_set_config(E, &c));
    const int G = c.graph_steps;
    const int n = G > 0 ? std::max(1, (trial_steps + G - 1) / G) * G : trial_steps;
    QS_TRY(ensemble_steps(E, G > 0 ? G : 1));                 // warm-up (and capture)
    QS_HIP(hipEventRecord(e0, E->stream));
    QS_TRY(ensemble_steps(E, n));
    QS_HIP(hipEventRecord(e1, E->stream));
    QS_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    QS_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / n;
    return QS_OK;
  };
  // Two descents, one over the per-class kernels and one over the merged
  // kernels, each from its own start; the faster result wins.  (A single
  // descent that switched between the two early was led by the noise of short
  // trials into the other family's untuned corner.)
  qs_ensemble_config best = E->cfg;
  best.graph_steps = 0;                                       // launch shapes first, directly launched
  double bms = 0.0;
  const int nc = (int)E->cls.size();
  auto try_cfg = [&](const qs_ensemble_config& c) -> int {
    double ms = 0.0;
    const int s = measure(c, &ms);
    if (s != QS_OK) return s;
    if (ms < bms) { bms = ms; best = c; }
    return QS_OK;
  };
  auto descend = [&](int merged) -> int {
    best.merged = merged;
    QS_TRY(measure(best, &bms));
    for (int sweep = 0; sweep < 2; ++sweep) {
      for (int v : {0, 1}) {
        qs_ensemble_config c = best; c.order = v;
        QS_TRY(try_cfg(c));
      }
      if (merged) {   // one kernel per phase: its block counts
        for (int b : {0, 128, 256, 512, 1024, 2048}) {
          qs_ensemble_config c = best; c.row_blocks[0] = b;
          QS_TRY(try_cfg(c));
          c = best; c.col_blocks[0] = b;
          QS_TRY(try_cfg(c));
        }
        continue;
      }
      for (int v : {0, 1}) {
        qs_ensemble_config c = best; c.concurrent = v;
        QS_TRY(try_cfg(c));
      }
      for (int ci = 0; ci < nc; ++ci) {
        int el[4];
        const int nl = lane_layouts(E->cls[ci].N, el);
        for (int k = 0; k < nl; ++k) {
          qs_ensemble_config c = best; c.lanes[ci] = el[k];
          QS_TRY(try_cfg(c));
        }
        for (int b : {0, 128, 256, 512, 1024, 2048}) {
          qs_ensemble_config c = best; c.row_blocks[ci] = b;
          QS_TRY(try_cfg(c));
          c = best; c.col_blocks[ci] = b;
          QS_TRY(try_cfg(c));
        }
      }
    }
    return QS_OK;
  };
  bool can_merge = true;
  for (auto& c : E->cls) can_merge = can_merge && mix_side(c.N);
  const qs_ensemble_config start = best;
  if (st == QS_OK) st = descend(0);
  if (st == QS_OK && can_merge) {
    const qs_ensemble_config b0 = best;
    const double ms0 = bms;
    best = start;
    st = descend(1);
    if (st == QS_OK && ms0 <= bms) { best = b0; bms = ms0; }
  }
  // graph depth, each candidate captured and launched as a graph
  for (int G : {5, 10, 20}) {
    qs_ensemble_config c = best; c.graph_steps = G;
    if (st == QS_OK) st = try_cfg(c);
  }
  if (st == QS_OK) st = qs_ensemble_set_config(E, &best);
  // restore states
  for (int r = 0; r < n; ++r) {
    qs_solver* S = E->runs[r];
    const size_t hs = sizeof(double2) * (size_t)S->pitch * S->mesh->Ny;
    for (int c = 0; c < 3; ++c)
      if (saved[3 * r + c]) {
        if (st == QS_OK && hipMemcpyAsync(S->old_[c], saved[3 * r + c], hs, hipMemcpyDeviceToDevice, E->stream) != hipSuccess)
          st = fail(QS_ERR_HIP, "tune: state restore");
      }
  }
  (void)hipStreamSynchronize(E->stream);
  for (auto p : saved)
    if (p) (void)hipFree(p);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (auto& c : E->cls) c.valid = false;
  if (st == QS_OK) {
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = bms;
  }
  return st;
}

// ---- batched discrete adjoint (one mesh size) ------------------------------
int qs_ensemble_adj_forward(qs_ensemble* E, int nsteps, int ckpt_every, const Qreal* alpha_t) {
  if (!E) return fail(QS_ERR_INVALID_ARG, "null ensemble");
  if (E->cls.size() != 1) return fail(QS_ERR_UNSUPPORTED_SIZE, "ensemble adjoint: runs of one mesh size only");
  if (nsteps <= 0) return fail(QS_ERR_INVALID_ARG, "nsteps must be > 0");
  if (!E->adj) E->adj = adj_group_new(E->runs);
  const int k = ckpt_every <= 0 ? nsteps : std::min(ckpt_every, nsteps);
  QS_TRY(adj_group_forward(E->adj, nsteps, k, alpha_t));
  E->cls[0].valid = false;   // the last segment's inverse sets went to the stage slots
  return QS_OK;
}

int qs_ensemble_adj_loss(qs_ensemble* E, int kind, Qreal* losses) {
  if (!E || !losses) return fail(QS_ERR_INVALID_ARG, "null argument");
  if (kind != 0 && kind != 1) return fail(QS_ERR_INVALID_ARG, "loss kind %d", kind);
  if (!E->adj || adj_group_steps(E->adj) <= 0) return fail(QS_ERR_STATE, "ensemble adjoint: no forward");
  for (size_t r = 0; r < E->runs.size(); ++r) QS_TRY(adj_run_seed(E->runs[r], kind, &losses[r]));
  return QS_OK;
}

int qs_ensemble_adj_backward(qs_ensemble* E, Qreal* abar_t, Qreal* abar) {
  if (!E) return fail(QS_ERR_INVALID_ARG, "null ensemble");
  if (!E->adj || adj_group_steps(E->adj) <= 0) return fail(QS_ERR_STATE, "ensemble adjoint: no forward");
  QS_TRY(adj_group_backward(E->adj));
  const int nsteps = adj_group_steps(E->adj);
  for (size_t r = 0; r < E->runs.size(); ++r) {
    qs_solver* S = E->runs[r];
    const Mesh* m = S->mesh;
    const size_t P = (size_t)m->Nx * m->Ny;
    if (abar_t) QS_TRY(row_sums_launch(S->adj->arow, nsteps, m->Ny, abar_t + r * (size_t)nsteps, stream_of(m)));
    if (abar) QS_HIP(hipMemcpyAsync(abar + r * P, S->adj->abar, sizeof(double) * P, hipMemcpyDeviceToDevice, stream_of(m)));
  }
  E->cls[0].valid = false;
  return QS_OK;
}

int qs_ensemble_adj_memory(qs_ensemble* E, size_t* bytes) {
  if (!E || !bytes) return fail(QS_ERR_INVALID_ARG, "null argument");
  *bytes = E->adj ? adj_group_bytes(E->adj) : 0;
  return QS_OK;
}

}  // extern "C"
