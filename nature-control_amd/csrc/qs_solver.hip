// qs_solver.hip -- the Qsolver's IFRK2 solver object: state, fused step, the
// unfused reference-structure step, per-kernel timing.
//
// Reference: InertialFriction/friction/friction.cu:150-260 (fields, IF tables,
// time loop), TimeIntegration/IFRK2.cuh:16-51 (stages; IFh := exp(L dt)).
//
// Fused step: [column prologue when the mixed arrays are stale] then
//   row pass -> column stage 1 -> row pass -> column stage 2
// (kernels in qs_fused.hip).  State and mixed arrays live in HBM with a row
// pitch padded to a multiple of 8 complex values (128-byte aligned rows); the
// reference layout is used only at qs_solver_set_state / get_state.
#include <algorithm>
#include <utility>
#include <vector>

#include "qs_solver.hpp"


namespace qs {
// Descriptor of this solver's run (device copy in S->d_desc) and the work
// lists of its row / column launches.
RunDesc make_desc(qs_solver* S) {
  const Mesh* m = S->mesh;
  RunDesc d{};
  d.Nx = m->Nx; d.Ny = m->Ny; d.Nxh = m->Nxh; d.pitch = S->pitch;
  for (int k = 0; k < 7; ++k) d.mix[k] = d.minp[k] = d.mout[k] = S->mix[k];
  for (int k = 0; k < 3; ++k) { d.old_[k] = S->old_[k]; d.new_[k] = S->new_[k]; }
  d.Ra = S->Ra;
  d.w00 = d.w00o = S->w00;
  d.flag = S->flag;
  d.ify = S->ify;
  d.dkx = 2 * M_PI / m->Lx;   // the mesh tables' own expressions (Basic/Mesh.cuh:6-21)
  d.dky = 2 * M_PI / m->Ly;
  d.alpha = S->p.alpha; d.lambda = S->p.lambda;
  d.DQ = S->p.D_Q; d.aQ = S->p.a_Q; d.cQ = S->p.c_Q;
  d.s_ErRe = 1.0 / (S->p.Er * S->p.Re);
  d.scale = 1.0 / ((double)m->Nx * (double)m->Ny);
  d.dt = S->p.dt; d.Re = S->p.Re; d.Rf = S->p.Rf;
  d.advect_w = S->p.advect_w;
  return d;
}

}  // namespace qs

namespace qs {
namespace {

int upload_desc(qs_solver* S) {
  const RunDesc d = make_desc(S);
  // d is a local: the copy completes before it goes out of scope
  QS_HIP(hipMemcpyAsync(S->d_desc, &d, sizeof(RunDesc), hipMemcpyHostToDevice, stream_of(S->mesh)));
  QS_HIP(hipStreamSynchronize(stream_of(S->mesh)));
  S->nrow = row_jobs(d);
  S->ncol = col_jobs(d);
  return QS_OK;
}

// The row pass's kernels are instantiated by Nx (their x FFTs), the column
// stages' by Ny: a non-square mesh runs the same fixed-size kernels.
// Row pass blocks: one per compute unit (the kernel's LDS admits one block
// per CU), each looping over its jobs.
int row_grid(const qs_solver* S) {
  const int ncu = S->mesh->priv->ncu;
  return ncu > 0 ? std::min(S->nrow, ncu) : S->nrow;
}
// The fused kernels are instantiated for power-of-two sides up to 1024 and
// for square meshes of side 3 * 2^a up to 768; other meshes run the per-op
// path (qs_step_ifrk2_unfused, qs_step_rk4_unfused).
int fused_check(const qs_solver* S) {
  const Mesh* m = S->mesh;
  const bool pow2 = fused_side(m->Nx) && fused_side(m->Ny);
  const bool tri = m->Nx == m->Ny && fused_tri_side(m->Nx);
  if (!pow2 && !tri)
    return fail(QS_ERR_UNSUPPORTED_SIZE, "fused step: %dx%d mesh (powers of two up to 1024, or square 3*2^a up to "
                "768; the unfused steps run any supported mesh)", m->Nx, m->Ny);
  return QS_OK;
}
// The staged twiddle table of the fused kernels: W_768 for 3 * 2^a sides.
const double2* fused_tw(const Mesh* m) { return m->Nx % 3 == 0 ? m->priv->tw768 : m->priv->tw1024; }
// fin != 0: a final column stage follows (see launch_fused_row).
int row_rhs_raw(qs_solver* S, int fin) {
  QS_TRY(fused_check(S));
  return launch_fused_row(S->mesh->Nx, S->d_desc, S->d_work, S->nrow, row_grid(S), fused_tw(S->mesh),
                          stream_of(S->mesh), fin);
}
int col_stage_raw(qs_solver* S, int mode) {
  QS_TRY(fused_check(S));
  return launch_fused_col(S->mesh->Ny, mode, S->d_desc, S->d_work, S->ncol, S->ncol, fused_tw(S->mesh),
                          stream_of(S->mesh));
}

// Non-finite state (friction.cu:280 checks the S field of each snapshot and
// exits): the final column stages raise the device flag; each fused step call
// ends with an asynchronous copy of it into pinned host memory, and a later
// call that finds the copy set refuses to run with QS_ERR_NONFINITE.  Nothing
// synchronises on the way: the refusal comes at the first call after the copy
// has landed.  qs_solver_nonfinite synchronises, reports and clears the flag.
int nonfinite_gate(const qs_solver* S) {
  if (*(volatile const int*)S->hflag)
    return fail(QS_ERR_NONFINITE, "non-finite value in the state of an earlier step (qs_solver_nonfinite "
                "reports and clears the flag)");
  return QS_OK;
}
int nonfinite_publish(qs_solver* S) {
  QS_HIP(hipMemcpyAsync(S->hflag, S->flag, sizeof(int), hipMemcpyDeviceToHost, stream_of(S->mesh)));
  return QS_OK;
}

hipEvent_t take_event(qs_solver* S) {
  if (!S->ev_pool.empty()) {
    hipEvent_t e = S->ev_pool.back();
    S->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Kernel classes of the fused step, for qs_solver_profile_read.
enum { K_COL_PRO = 0, K_ROW = 1, K_COL_S1 = 2, K_COL_S2 = 3, K_NCLASS = 4 };

template <class F>
int timed(qs_solver* S, int cls, F launch) {
  if (!S->prof) return launch();
  hipEvent_t a = take_event(S), b = take_event(S);
  if (!a || !b) return fail(QS_ERR_HIP, "hipEventCreate");
  QS_HIP(hipEventRecord(a, stream_of(S->mesh)));
  QS_TRY(launch());
  QS_HIP(hipEventRecord(b, stream_of(S->mesh)));
  S->ev_used.push_back({cls, {a, b}});
  return QS_OK;
}

int row_rhs(qs_solver* S, int fin = 0) { return timed(S, K_ROW, [&] { return row_rhs_raw(S, fin); }); }
int col_stage(qs_solver* S, int mode) {
  // intermediate stages (IFRK2 1, RK4 1-3) and final stages (IFRK2 2, RK4 4, IFE)
  const int cls = is_prologue(mode) ? K_COL_PRO
                  : (mode == MODE_S1 || mode == MODE_R1 || mode == MODE_R2 || mode == MODE_R3) ? K_COL_S1
                                                                                            : K_COL_S2;
  return timed(S, cls, [&] { return col_stage_raw(S, mode); });
}

// --------------------------------------------------------------------------
// Unfused step: the reference's call sequence through the per-op entry points.
// --------------------------------------------------------------------------
enum { F_WCURR, F_WNONL, F_U, F_V, F_R1CURR, F_R1NONL, F_R2CURR, F_R2NONL, F_S, F_P11, F_P12, F_P21,
       F_RA, F_AUX, F_AUX1, F_NFIELDS };

int unfused_setup(qs_solver* S) {
  if (S->unfused_ready) return QS_OK;
  Mesh* m = S->mesh;
  for (int k = 0; k < F_NFIELDS; ++k) QS_TRY(qs_field_create(m, &S->F[k]));
  const size_t ws = (size_t)m->wavesize;
  for (int k = 0; k < 4; ++k)
    if (hipMalloc((void**)&S->IFt[k], ws) != hipSuccess) return fail(QS_ERR_OOM, "IF tables");
  for (int k = 0; k < 3; ++k)
    if (hipMalloc((void**)&S->uold[k], (size_t)m->specsize) != hipSuccess ||
        hipMalloc((void**)&S->unew[k], (size_t)m->specsize) != hipSuccess)
      return fail(QS_ERR_OOM, "unfused state");
  // friction.cu:208-210 (IF tables; Q tables with the D_Q, a_Q coefficients)
  QS_TRY(qs_wlin_func(S->IFt[0], S->IFt[1], S->p.Re, S->p.Rf, S->p.dt, m));
  QS_TRY(qlin_c(S->IFt[2], S->IFt[3], S->p.dt, S->p.D_Q, S->p.a_Q, m));
  S->unfused_ready = true;
  return QS_OK;
}

int unfused_activity(qs_solver* S) {
  Mesh* m = S->mesh;
  if (S->Ra) return qs_FldCopy(S->F[F_RA]->phys, S->Ra, m);
  return qs_FldSet(S->F[F_RA]->phys, S->p.alpha, m);
}

int unfused_rhs(qs_solver* S) {
  Field** F = S->F;
  const qs_params& p = S->p;
  QS_TRY(qs_curr_func(F[F_R1CURR], F[F_R2CURR], F[F_WCURR], F[F_U], F[F_V], F[F_S]));
  QS_TRY(wnonl_func_c(F[F_WNONL], F[F_WCURR], F[F_U], F[F_V], F[F_R1CURR], F[F_R2CURR], F[F_RA],
                      F[F_S], F[F_P11], F[F_P12], F[F_P21], p.Re, p.Er, p.lambda, F[F_AUX],
                      F[F_AUX1], S->coef));
  QS_TRY(r1nonl_func_c(F[F_R1NONL], F[F_WCURR], F[F_U], F[F_V], F[F_R1CURR], F[F_R2CURR], F[F_S],
                       p.lambda, F[F_AUX], F[F_AUX1], S->coef));
  return r2nonl_func_c(F[F_R2NONL], F[F_WCURR], F[F_U], F[F_V], F[F_R1CURR], F[F_R2CURR], F[F_S],
                       p.lambda, F[F_AUX], F[F_AUX1], S->coef);
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

int qs_solver_create(Mesh* mesh, const qs_params* params, qs_solver** out) {
  if (!out || !params) return fail(QS_ERR_INVALID_ARG, "qs_solver_create: null argument");
  *out = nullptr;
  QS_TRY(check_mesh(mesh));
  if (mesh->Nx > 1024 || mesh->Ny > 1024)
    return fail(QS_ERR_UNSUPPORTED_SIZE, "qs_solver_create: fused step supports N <= 1024");
  if (!(params->Re != 0) || !(params->Er != 0) || !(params->dt > 0))
    return fail(QS_ERR_INVALID_ARG, "qs_solver_create: Re, Er must be non-zero and dt > 0");
  qs_solver* S = new qs_solver();
  S->mesh = mesh;
  S->p = *params;
  S->coef.DQ = params->D_Q;
  S->coef.aQ = params->a_Q;
  S->coef.cQ = params->c_Q;
  S->coef.advect_w = params->advect_w != 0;
  S->pitch = pitch_of(mesh->Nxh);
  const size_t hs = sizeof(double2) * (size_t)S->pitch * mesh->Ny;
  const int nwork = 3 * (mesh->Nxh + 3) + mesh->Ny;   // >= row and column job counts
  bool ok = true;
  for (int k = 0; k < 3 && ok; ++k)
    ok = hipMalloc((void**)&S->old_[k], hs) == hipSuccess && hipMalloc((void**)&S->new_[k], hs) == hipSuccess;
  for (int k = 0; k < 7 && ok; ++k) ok = hipMalloc((void**)&S->mix[k], hs) == hipSuccess;
  ok = ok && hipMalloc((void**)&S->w00, sizeof(double)) == hipSuccess;
  ok = ok && hipMalloc((void**)&S->flag, sizeof(int)) == hipSuccess;
  ok = ok && hipMalloc((void**)&S->ify, sizeof(double) * ifac_len(mesh->Nxh, mesh->Ny)) == hipSuccess;
  ok = ok && hipHostMalloc((void**)&S->hflag, sizeof(int), hipHostMallocDefault) == hipSuccess;
  if (S->hflag) *S->hflag = 0;
  ok = ok && hipMalloc((void**)&S->d_desc, sizeof(RunDesc)) == hipSuccess;
  ok = ok && hipMalloc((void**)&S->d_work, sizeof(int2) * nwork) == hipSuccess;
  if (!ok) {
    qs_solver_destroy(S);
    return fail(QS_ERR_OOM, "qs_solver_create: hipMalloc");
  }
  std::vector<int2> work(nwork);
  for (int k = 0; k < nwork; ++k) work[k] = make_int2(0, k);
  int st = QS_OK;
  const hipStream_t sm = stream_of(mesh);
  for (int k = 0; k < 3 && st == QS_OK; ++k)
    if (hipMemsetAsync(S->old_[k], 0, hs, sm) != hipSuccess || hipMemsetAsync(S->new_[k], 0, hs, sm) != hipSuccess)
      st = fail(QS_ERR_HIP, "qs_solver_create: memset");
  for (int k = 0; k < 7 && st == QS_OK; ++k)
    if (hipMemsetAsync(S->mix[k], 0, hs, sm) != hipSuccess) st = fail(QS_ERR_HIP, "qs_solver_create: memset");
  if (st == QS_OK && (hipMemsetAsync(S->flag, 0, sizeof(int), sm) != hipSuccess ||
                      hipMemsetAsync(S->w00, 0, sizeof(double), sm) != hipSuccess))
    st = fail(QS_ERR_HIP, "qs_solver_create: memset");
  if (st == QS_OK && hipMemcpy(S->d_work, work.data(), sizeof(int2) * nwork, hipMemcpyHostToDevice) != hipSuccess)
    st = fail(QS_ERR_HIP, "qs_solver_create: work list");
  if (st == QS_OK)
    st = ifac_launch(S->ify, mesh->Nxh, mesh->Ny, 2 * M_PI / mesh->Lx, 2 * M_PI / mesh->Ly, params->Re,
                     params->Rf, params->D_Q, params->a_Q, params->dt, sm);
  if (st == QS_OK) st = upload_desc(S);
  if (st == QS_OK && hipStreamSynchronize(sm) != hipSuccess) st = fail(QS_ERR_HIP, "qs_solver_create: sync");
  if (st != QS_OK) {
    qs_solver_destroy(S);
    return st;
  }
  *out = S;
  return QS_OK;
}

int qs_solver_destroy(qs_solver* S) {
  if (!S) return QS_OK;
  (void)hipDeviceSynchronize();   // the Mesh may already be gone: do not touch it
  adj_state_free(S);
  for (int k = 0; k < 3; ++k) {
    if (S->old_[k]) (void)hipFree(S->old_[k]);
    if (S->new_[k]) (void)hipFree(S->new_[k]);
    if (S->uold[k]) (void)hipFree(S->uold[k]);
    if (S->unew[k]) (void)hipFree(S->unew[k]);
  }
  for (int k = 0; k < 7; ++k)
    if (S->mix[k]) (void)hipFree(S->mix[k]);
  for (void* p : {(void*)S->Ra, (void*)S->w00, (void*)S->flag, (void*)S->d_desc, (void*)S->d_work, (void*)S->ify})
    if (p) (void)hipFree(p);
  if (S->hflag) (void)hipHostFree(S->hflag);
  for (auto* f : S->F) qs_field_destroy(f);
  for (auto* t : S->IFt)
    if (t) (void)hipFree(t);
  for (auto& u : S->ev_used) {
    (void)hipEventDestroy(u.second.first);
    (void)hipEventDestroy(u.second.second);
  }
  for (auto e : S->ev_pool) (void)hipEventDestroy(e);
  delete S;
  return QS_OK;
}

int qs_solver_set_activity(qs_solver* S, const Qreal* Ra) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  const Mesh* m = S->mesh;
  if (!Ra) {
    if (S->Ra) {
      QS_HIP(hipStreamSynchronize(stream_of(m)));
      (void)hipFree(S->Ra);
    }
    S->Ra = nullptr;
    return upload_desc(S);
  }
  if (!S->Ra) {
    if (hipMalloc((void**)&S->Ra, (size_t)m->physsize) != hipSuccess) return fail(QS_ERR_OOM, "activity field");
    QS_TRY(upload_desc(S));
  }
  QS_HIP(hipMemcpyAsync(S->Ra, Ra, (size_t)m->physsize, hipMemcpyDeviceToDevice, stream_of(m)));
  return QS_OK;
}

int qs_solver_set_state(qs_solver* S, const Qcomp* w, const Qcomp* r1, const Qcomp* r2) {
  if (!S || !w || !r1 || !r2) return fail(QS_ERR_INVALID_ARG, "qs_solver_set_state: null argument");
  const Mesh* m = S->mesh;
  const Qcomp* src[3] = {w, r1, r2};
  const size_t row = sizeof(Qcomp) * m->Nxh;
  for (int k = 0; k < 3; ++k)
    QS_HIP(hipMemcpy2DAsync(S->old_[k], sizeof(double2) * S->pitch, src[k], row, row, m->Ny,
                            hipMemcpyDeviceToDevice, stream_of(m)));
  S->mixed_valid = false;
  return QS_OK;
}

int qs_solver_get_state(qs_solver* S, Qcomp* w, Qcomp* r1, Qcomp* r2) {
  if (!S || !w || !r1 || !r2) return fail(QS_ERR_INVALID_ARG, "qs_solver_get_state: null argument");
  const Mesh* m = S->mesh;
  Qcomp* dst[3] = {w, r1, r2};
  const size_t row = sizeof(Qcomp) * m->Nxh;
  for (int k = 0; k < 3; ++k)
    QS_HIP(hipMemcpy2DAsync(dst[k], row, S->old_[k], sizeof(double2) * S->pitch, row, m->Ny,
                            hipMemcpyDeviceToDevice, stream_of(m)));
  return QS_OK;
}

int qs_step_ifrk2(qs_solver* S, int nsteps) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  if (nsteps < 0) return fail(QS_ERR_INVALID_ARG, "nsteps < 0");
  QS_TRY(nonfinite_gate(S));
  Range rg("qs_step_ifrk2");
  for (int n = 0; n < nsteps; ++n) {
    Range rs("ifrk2 step %d", n);
    if (!S->mixed_valid) {
      Range rp("column prologue");
      QS_TRY(col_stage(S, MODE_PRO));
      S->mixed_valid = true;
    }
    {
      Range r1("ifrk2 stage 1");
      QS_TRY(row_rhs(S));
      QS_TRY(col_stage(S, MODE_S1));
    }
    Range r2("ifrk2 stage 2");
    QS_TRY(row_rhs(S, 1));
    QS_TRY(col_stage(S, MODE_S2));
  }
  return nonfinite_publish(S);
}

// friction.cu:212-260 verbatim (IF-RK4, RK4.cuh:12-74 with IF = exp(L dt),
// IFh = exp(L dt / 2)): four right-hand sides per step, each a row pass and a
// column stage.
int qs_step_rk4(qs_solver* S, int nsteps) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  if (nsteps < 0) return fail(QS_ERR_INVALID_ARG, "nsteps < 0");
  QS_TRY(nonfinite_gate(S));
  Range rg("qs_step_rk4");
  for (int n = 0; n < nsteps; ++n) {
    Range rs("rk4 step %d", n);
    if (!S->mixed_valid) {
      Range rp("column prologue");
      QS_TRY(col_stage(S, MODE_PRO_R4));
      S->mixed_valid = true;
    }
    for (int mode : {MODE_R1, MODE_R2, MODE_R3, MODE_R4}) {
      Range rq("rk4 stage %d", mode - MODE_R1 + 1);
      QS_TRY(row_rhs(S, mode == MODE_R4));
      QS_TRY(col_stage(S, mode));
    }
  }
  return nonfinite_publish(S);
}

// IF-Euler (TimeIntegration/IFE.cuh:15-25), then dealias: one right-hand side
// per step.
int qs_step_ife(qs_solver* S, int nsteps) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  if (nsteps < 0) return fail(QS_ERR_INVALID_ARG, "nsteps < 0");
  QS_TRY(nonfinite_gate(S));
  Range rg("qs_step_ife");
  for (int n = 0; n < nsteps; ++n) {
    Range rs("ife step %d", n);
    if (!S->mixed_valid) {
      Range rp("column prologue");
      QS_TRY(col_stage(S, MODE_PRO_E));
      S->mixed_valid = true;
    }
    QS_TRY(row_rhs(S, 1));
    QS_TRY(col_stage(S, MODE_E));
  }
  return nonfinite_publish(S);
}

// friction.cu:212-260 with the IFRK2 stage kernels and IFh := IF.
int qs_step_ifrk2_unfused(qs_solver* S, int nsteps) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  QS_TRY(unfused_setup(S));
  QS_TRY(unfused_activity(S));
  Mesh* m = S->mesh;
  Field** F = S->F;
  const double dt = S->p.dt;
  Field* curr[3] = {F[F_WCURR], F[F_R1CURR], F[F_R2CURR]};
  Field* nonl[3] = {F[F_WNONL], F[F_R1NONL], F[F_R2NONL]};
  const size_t row = sizeof(Qcomp) * m->Nxh;
  const size_t prow = sizeof(double2) * S->pitch;
  for (int k = 0; k < 3; ++k)
    QS_HIP(hipMemcpy2DAsync(S->uold[k], row, S->old_[k], prow, row, m->Ny, hipMemcpyDeviceToDevice,
                            stream_of(m)));
  for (int n = 0; n < nsteps; ++n) {
    const double* IF[3] = {S->IFt[0], S->IFt[2], S->IFt[2]};
    for (int k = 0; k < 3; ++k)
      QS_TRY(qs_ifrk2_func0D(S->uold[k], curr[k]->spec, S->unew[k], IF[k], IF[k], dt, m));
    QS_TRY(unfused_rhs(S));
    for (int k = 0; k < 3; ++k)
      QS_TRY(qs_ifrk2_func1D(S->uold[k], curr[k]->spec, S->unew[k], nonl[k]->spec, IF[k], IF[k], dt, m));
    QS_TRY(unfused_rhs(S));
    for (int k = 0; k < 3; ++k)
      QS_TRY(qs_ifrk2_func2D(S->uold[k], curr[k]->spec, S->unew[k], nonl[k]->spec, IF[k], IF[k], dt, m));
    for (int k = 0; k < 3; ++k) QS_TRY(qs_dealias(S->unew[k], m));
    for (int k = 0; k < 3; ++k) QS_TRY(qs_SpecSet(S->uold[k], S->unew[k], m));
  }
  for (int k = 0; k < 3; ++k)
    QS_HIP(hipMemcpy2DAsync(S->old_[k], prow, S->uold[k], row, row, m->Ny, hipMemcpyDeviceToDevice,
                            stream_of(m)));
  S->mixed_valid = false;
  return QS_OK;
}

// The reference driver's own step (friction.cu:212-260 with RK4.cuh:12-74:
// IF = exp(L dt), IFh = exp(L dt/2)) through the per-op entry points.
int qs_step_rk4_unfused(qs_solver* S, int nsteps) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  if (nsteps < 0) return fail(QS_ERR_INVALID_ARG, "nsteps < 0");
  QS_TRY(unfused_setup(S));
  QS_TRY(unfused_activity(S));
  Mesh* m = S->mesh;
  Field** F = S->F;
  const double dt = S->p.dt;
  Field* curr[3] = {F[F_WCURR], F[F_R1CURR], F[F_R2CURR]};
  Field* nonl[3] = {F[F_WNONL], F[F_R1NONL], F[F_R2NONL]};
  const double* IF[3] = {S->IFt[0], S->IFt[2], S->IFt[2]};
  const double* IFh[3] = {S->IFt[1], S->IFt[3], S->IFt[3]};
  const size_t row = sizeof(Qcomp) * m->Nxh;
  const size_t prow = sizeof(double2) * S->pitch;
  for (int k = 0; k < 3; ++k)
    QS_HIP(hipMemcpy2DAsync(S->uold[k], row, S->old_[k], prow, row, m->Ny, hipMemcpyDeviceToDevice,
                            stream_of(m)));
  using StageFn = int (*)(const Qcomp*, Qcomp*, Qcomp*, const Qcomp*, const Qreal*, const Qreal*, Qreal,
                          const Mesh*);
  const StageFn stage[4] = {qs_rk4_func1D, qs_rk4_func2D, qs_rk4_func3D, qs_rk4_func4D};
  for (int n = 0; n < nsteps; ++n) {
    for (int k = 0; k < 3; ++k) QS_TRY(qs_rk4_func0D(S->uold[k], curr[k]->spec, S->unew[k], IF[k], IFh[k], dt, m));
    for (int q = 0; q < 4; ++q) {
      QS_TRY(unfused_rhs(S));
      for (int k = 0; k < 3; ++k)
        QS_TRY(stage[q](S->uold[k], curr[k]->spec, S->unew[k], nonl[k]->spec, IF[k], IFh[k], dt, m));
    }
    for (int k = 0; k < 3; ++k) QS_TRY(qs_dealias(S->unew[k], m));
    for (int k = 0; k < 3; ++k) QS_TRY(qs_SpecSet(S->uold[k], S->unew[k], m));
  }
  for (int k = 0; k < 3; ++k)
    QS_HIP(hipMemcpy2DAsync(S->old_[k], prow, S->uold[k], row, row, m->Ny, hipMemcpyDeviceToDevice,
                            stream_of(m)));
  S->mixed_valid = false;
  return QS_OK;
}

// Test hooks: run one kernel of the fused step / read one internal array.
int qs_solver_debug_phase(qs_solver* S, int phase) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  switch (phase) {
    case 0: QS_TRY(col_stage(S, MODE_PRO)); S->mixed_valid = true; return QS_OK;
    case 1: return row_rhs(S);
    case 2: return col_stage(S, MODE_S1);
    case 3: return col_stage(S, MODE_S2);
    default: return fail(QS_ERR_INVALID_ARG, "phase %d", phase);
  }
}

int qs_solver_debug_array(qs_solver* S, int which, Qcomp* out) {
  if (!S || !out) return fail(QS_ERR_INVALID_ARG, "null argument");
  const Mesh* m = S->mesh;
  const double2* src = nullptr;
  if (which >= 0 && which < 7) src = S->mix[which];
  else if (which >= 7 && which < 10) src = S->old_[which - 7];
  else if (which >= 10 && which < 13) src = S->new_[which - 10];
  else return fail(QS_ERR_INVALID_ARG, "array %d", which);
  const size_t row = sizeof(Qcomp) * m->Nxh;
  QS_HIP(hipMemcpy2DAsync(out, row, src, sizeof(double2) * S->pitch, row, m->Ny, hipMemcpyDeviceToDevice,
                          stream_of(m)));
  return QS_OK;
}

int qs_debug_wave_fft(const Mesh* mesh, int n, int sign, const Qcomp* in, Qcomp* out) {
  QS_TRY(check_mesh(mesh));
  if (!in || !out) return fail(QS_ERR_INVALID_ARG, "null argument");
  return wave_fft_test(n, sign, (const double2*)in, (double2*)out,
                       n % 3 == 0 ? mesh->priv->tw768 : mesh->priv->tw1024, stream_of(mesh));
}

int qs_solver_profile(qs_solver* S, int enable) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  S->prof = enable != 0;
  return QS_OK;
}

int qs_solver_profile_read(qs_solver* S, double* avg_ms, int* counts) {
  if (!S || !avg_ms || !counts) return fail(QS_ERR_INVALID_ARG, "null argument");
  double sum[K_NCLASS] = {0, 0, 0, 0};
  int cnt[K_NCLASS] = {0, 0, 0, 0};
  QS_HIP(hipStreamSynchronize(stream_of(S->mesh)));
  for (auto& u : S->ev_used) {
    float ms = 0.f;
    QS_HIP(hipEventElapsedTime(&ms, u.second.first, u.second.second));
    sum[u.first] += ms;
    cnt[u.first] += 1;
    S->ev_pool.push_back(u.second.first);
    S->ev_pool.push_back(u.second.second);
  }
  S->ev_used.clear();
  for (int k = 0; k < K_NCLASS; ++k) {
    counts[k] = cnt[k];
    avg_ms[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
  }
  return QS_OK;
}

int qs_solver_nonfinite(qs_solver* S, int* flag) {
  if (!S || !flag) return fail(QS_ERR_INVALID_ARG, "null argument");
  const Mesh* m = S->mesh;
  QS_HIP(hipStreamSynchronize(stream_of(m)));   // every earlier copy into hflag has landed
  QS_HIP(hipMemcpy(S->hflag, S->flag, sizeof(int), hipMemcpyDeviceToHost));
  *flag = *S->hflag;
  QS_HIP(hipMemsetAsync(S->flag, 0, sizeof(int), stream_of(m)));
  QS_HIP(hipStreamSynchronize(stream_of(m)));
  *S->hflag = 0;
  return QS_OK;
}

}  // extern "C"
