/*
 * qsolver.h -- C-ABI of the MI355X (gfx950) Qsolver library, libqsolver.so.
 *
 * Drop-in boundary for the reference's single-GPU Qsolver code
 * (whpy/nature-control).  Every entry point replaces one reference call; the
 * comment on each names the reference file:line it stands in for.  Argument
 * order is the reference's, with the CUDA launch configuration
 * (<<<grid, block>>>, Nx|Nxh, Ny, BSZ) replaced by the Mesh that carries it.
 *
 * Conventions
 *  - Every pointer to field data is caller-owned DEVICE memory in the
 *    reference layouts: physical  Qreal  phys[j*Nx + i]   (Nx*Ny values)
 *                       spectral  Qcomp  spec[j*Nxh + i]  (Nxh*Ny, Nxh=Nx/2+1)
 *                       wave table Qreal w[j*Nxh + i]     (Nxh*Ny)
 *  - All calls are asynchronous on the Mesh's HIP stream (qs_mesh_set_stream)
 *    except the create/destroy calls, qs_solver_get_state's completion and the
 *    error query.  A Mesh/solver is not thread-safe; distinct objects are.
 *  - Every function returns a qs_status; nothing terminates the process or
 *    throws across the boundary (the reference's cudaErr.h:54-72 aborts).
 *    qs_last_error() returns the message of the calling thread's last failure.
 *  - In-place aliasing is allowed where the reference aliases (elementwise and
 *    derivative ops, e.g. friction/Qsolver.cuh:181,238).  FwdTrans/BwdTrans
 *    need f != ft and preserve their input, except BwdTrans's documented
 *    column-0 symmetrisation of ft (FldOp/QFldfuncs.cuh:50-55).
 *  - Supported meshes: Nx, Ny = 2^a in [16, 2048] or 3 * 2^a in [48, 1536]
 *    (the per-op calls and the unfused steps); the fused steps take powers of
 *    two up to 1024 and square 3 * 2^a meshes up to 768 (the reference
 *    driver's grid); the adjoint takes power-of-two sides up to 1024 (Nx and
 *    Ny independently), ensembles square power-of-two runs up to 1024.
 */
#ifndef QSOLVER_H
#define QSOLVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double Qreal;                  /* Basic/QActFlowDef.cuh:6 */
typedef struct { double x, y; } Qcomp; /* Basic/QActFlowDef.cuh:7 (cuDoubleComplex layout) */

typedef enum qs_status {
  QS_OK = 0,
  QS_ERR_INVALID_ARG = 1,
  QS_ERR_UNSUPPORTED_SIZE = 2,
  QS_ERR_HIP = 3,
  QS_ERR_NONFINITE = 4,
  QS_ERR_OOM = 5,
  QS_ERR_STATE = 6,
  QS_ERR_IO = 7
} qs_status;

typedef struct qs_mesh_priv qs_mesh_priv;

/* Basic/Mesh.cuh:48-113.  Same field names and order; the dim3 launch shapes
 * and cufftHandle plans are replaced by the build-private trailing pointer. */
typedef struct Mesh {
  int Nx, Nxh, Ny;
  Qreal Lx, Ly, dx, dy;
  int specsize, physsize, wavesize;
  int BSZ;
  Qcomp* spec;  /* device scratch, Nxh*Ny */
  Qreal* phys;  /* device scratch, Nx*Ny */
  Qreal *kx, *ky, *k_squared;
  Qreal* cutoff;
  qs_mesh_priv* priv;
} Mesh;

/* Basic/Mesh.cuh:115-129 */
typedef struct Field {
  Qcomp* spec;
  Qreal* phys;
  Mesh* mesh;
} Field;

/* ---- lifecycle --------------------------------------------------------- */
/* Mesh::Mesh (Basic/Mesh.cuh:61-107): tables (waveinit :6-21, cutoff_func
 * :23-46), scratch and FFT set-up.  BSZ is kept for the mirror only. */
int qs_mesh_create(int Nx, int Ny, Qreal Lx, Qreal Ly, int BSZ, Mesh** out);
/* Mesh::~Mesh (Basic/Mesh.cuh:109-112); also frees what the reference leaks. */
int qs_mesh_destroy(Mesh* mesh);
/* No reference counterpart (the reference uses the legacy default stream). */
int qs_mesh_set_stream(Mesh* mesh, void* hip_stream);
int qs_mesh_synchronize(const Mesh* mesh);
/* Field::Field / ~Field (Basic/Mesh.cuh:120-128). */
int qs_field_create(Mesh* mesh, Field** out);
int qs_field_destroy(Field* field);
const char* qs_last_error(void);
const char* qs_status_string(int status);

/* ---- transforms: FldOp/QFldfuncs.cuh:8-55 ------------------------------- */
int qs_FwdTrans(Qcomp* ft, const Qreal* f, const Mesh* mesh);     /* :45-48 */
int qs_BwdTrans(Qreal* f, Qcomp* ft, const Mesh* mesh);           /* :50-55 */
int qs_coeff(Qreal* f, const Mesh* mesh);                         /* :8-16 */
int qs_symmetry_func(Qcomp* f_spec, const Mesh* mesh);            /* :18-44 */

/* ---- field algebra: FldOp/QFldfuncs.cuh:63-224 -------------------------- */
int qs_FldMul(Qreal* pc, const Qreal* pa, const Qreal* pb, Qreal C, const Mesh* mesh);  /* :63-71 */
int qs_FldAdd(Qreal* pc, Qreal a, const Qreal* pa, Qreal b, const Qreal* pb,
              const Mesh* mesh);                                                      /* :74-82 */
int qs_FldAddScalar(Qreal* pc, Qreal a, const Qreal* pa, Qreal b, const Mesh* mesh);   /* :85-93 */
int qs_FldSet(Qreal* pf, Qreal c, const Mesh* mesh);                                   /* :96-104 */
int qs_FldCopy(Qreal* pf, const Qreal* c, const Mesh* mesh);                           /* :107-115 */
int qs_SpecSet(Qcomp* pa, const Qcomp* pb, const Mesh* mesh);                          /* :118-126 */
int qs_SpecSetConst(Qcomp* pa, Qcomp c, const Mesh* mesh);                             /* :129-137 */
int qs_SpecAdd(Qcomp* spc, Qreal a, const Qcomp* spa, Qreal b, const Qcomp* spb,
               const Mesh* mesh);                                                     /* :141-149 */
int qs_SpecMul(Qcomp* spc, const Qcomp* spa, Qreal C, const Mesh* mesh);               /* :151-159 */
int qs_xDeriv(Qcomp* fx, const Qcomp* f, const Mesh* mesh);                            /* :165-176 */
int qs_yDeriv(Qcomp* fy, const Qcomp* f, const Mesh* mesh);                            /* :178-189 */
int qs_Laplacian(Qcomp* fL, const Qcomp* f, const Mesh* mesh);                         /* :191-202 */
int qs_vel_func(const Qcomp* w_c, Qcomp* u_c, Qcomp* v_c, const Mesh* mesh);           /* :204-224 */

/* ---- physics: InertialFriction/friction/Qsolver.cuh --------------------- */
int qs_wlin_func(Qreal* IFw, Qreal* IFwh, Qreal Re, Qreal Rf, Qreal dt, const Mesh* mesh); /* :20-30 */
int qs_r1lin_func(Qreal* IFr1, Qreal* IFr1h, Qreal dt, const Mesh* mesh);                 /* :32-42 */
int qs_r2lin_func(Qreal* IFr2, Qreal* IFr2h, Qreal dt, const Mesh* mesh);                 /* :44-54 */
int qs_S_func(Qreal* S, const Qreal* r1, const Qreal* r2, const Mesh* mesh);              /* :63-71 */
int qs_Rar(Qreal* rar, const Qreal* r, const Qreal* Ra, const Mesh* mesh);                /* :75-83 */
int qs_single1(Qreal* s1, const Qreal* r, const Qreal* S, Qreal lambda, const Mesh* mesh); /* :86-94 */
int qs_S2r(Qreal* rs, const Qreal* r, const Qreal* S, const Mesh* mesh);                  /* :206-215 */
int qs_PiDeriv(Qcomp* dP, const Qcomp* p12, const Qcomp* p11, const Qcomp* p21, Qreal Er,
               Qreal Re, const Mesh* mesh);                                               /* :154-164 */
int qs_dealias(Qcomp* ft, const Mesh* mesh);                                              /* :314-322 */
int qs_single(Field* s0, Field* r, Field* S, Qreal lambda);                                /* :96-105 */
int qs_cross(Field* cs, Field* r1, Field* r2, Field* aux);                                 /* :107-122 */
int qs_p11func(Field* p11, Field* r1, Field* S, Field* Ra, Qreal lambda);                  /* :124-131 */
int qs_p12func(Field* p12, Field* r1, Field* r2, Field* S, Field* Ra, Qreal lambda,
               Field* aux, Field* aux1);                                                  /* :133-141 */
int qs_p21func(Field* p21, Field* r1, Field* r2, Field* S, Field* Ra, Qreal lambda,
               Field* aux, Field* aux1);                                                  /* :143-151 */
int qs_wnonl_func(Field* wnonl, Field* wcurr, Field* u, Field* v, Field* r1curr,
                  Field* r2curr, Field* Ra, Field* S, Field* p11, Field* p12, Field* p21,
                  Qreal Re, Qreal Er, Qreal lambda, Field* aux, Field* aux1);             /* :166-201 */
int qs_r1nonl_func(Field* r1nonl, Field* wcurr, Field* u, Field* v, Field* r1curr,
                   Field* r2curr, Field* S, Qreal lambda, Field* aux, Field* aux1);       /* :217-255 */
int qs_r2nonl_func(Field* r2nonl, Field* wcurr, Field* u, Field* v, Field* r1curr,
                   Field* r2curr, Field* S, Qreal lambda, Field* aux, Field* aux1);       /* :257-300 */
int qs_curr_func(Field* r1curr, Field* r2curr, Field* wcurr, Field* u, Field* v,
                 Field* S);                                                               /* :302-312 */

/* ---- time integration: TimeIntegration/{IFE,IFRK2,RK4}.cuh ------------- */
int qs_ife_func1D(const Qcomp* old_spec, Qcomp* new_spec, const Qcomp* nonl_spec,
                  const Qreal* IF, const Qreal* IFh, Qreal dt, const Mesh* mesh);   /* IFE.cuh:15-25 */
int qs_ifrk2_func0D(const Qcomp* old_spec, Qcomp* curr_spec, Qcomp* new_spec,
                    const Qreal* IF, const Qreal* IFh, Qreal dt, const Mesh* mesh); /* IFRK2.cuh:16-26 */
int qs_ifrk2_func1D(const Qcomp* old_spec, Qcomp* curr_spec, Qcomp* new_spec,
                    const Qcomp* nonl_spec, const Qreal* IF, const Qreal* IFh, Qreal dt,
                    const Mesh* mesh);                                              /* IFRK2.cuh:28-39 */
int qs_ifrk2_func2D(const Qcomp* old_spec, Qcomp* curr_spec, Qcomp* new_spec,
                    const Qcomp* nonl_spec, const Qreal* IF, const Qreal* IFh, Qreal dt,
                    const Mesh* mesh);                                              /* IFRK2.cuh:41-51 */
int qs_rk4_func0D(const Qcomp* old_spec, Qcomp* curr_spec, Qcomp* new_spec,
                  const Qreal* IF, const Qreal* IFh, Qreal dt, const Mesh* mesh);   /* RK4.cuh:12-22 */
int qs_rk4_func1D(const Qcomp* old_spec, Qcomp* curr_spec, Qcomp* new_spec,
                  const Qcomp* nonl_spec, const Qreal* IF, const Qreal* IFh, Qreal dt,
                  const Mesh* mesh);                                                /* RK4.cuh:24-35 */
int qs_rk4_func2D(const Qcomp* old_spec, Qcomp* curr_spec, Qcomp* new_spec,
                  const Qcomp* nonl_spec, const Qreal* IF, const Qreal* IFh, Qreal dt,
                  const Mesh* mesh);                                                /* RK4.cuh:37-48 */
int qs_rk4_func3D(const Qcomp* old_spec, Qcomp* curr_spec, Qcomp* new_spec,
                  const Qcomp* nonl_spec, const Qreal* IF, const Qreal* IFh, Qreal dt,
                  const Mesh* mesh);                                                /* RK4.cuh:50-61 */
int qs_rk4_func4D(const Qcomp* old_spec, Qcomp* curr_spec, Qcomp* new_spec,
                  const Qcomp* nonl_spec, const Qreal* IF, const Qreal* IFh, Qreal dt,
                  const Mesh* mesh);                                                /* RK4.cuh:63-74 */

/* ---- fused solver ------------------------------------------------------- */
/* Parameters of InertialFriction/friction/friction.cu:100-105 plus the
 * Q-equation coefficients D_Q (Laplacian), a_Q (linear) and c_Q (S^2 term);
 * D_Q = a_Q = c_Q = 1 and advect_w = 1 is exactly the reference physics
 * (friction/Qsolver.cuh:38,50,213; advect_w = 0 is beginning/Qsolver.cuh:192). */
typedef struct qs_params {
  Qreal lambda, Re, Er, Rf, alpha;
  Qreal D_Q, a_Q, c_Q;
  Qreal dt;
  int advect_w;
  int reserved;
} qs_params;

typedef struct qs_solver qs_solver;

/* Owns the state (old/new/curr spectra of omega, r1, r2), the IF tables and
 * the workspaces of the fused kernels.  Replaces the Field set-up and IF
 * tables of friction.cu:150-210. */
int qs_solver_create(Mesh* mesh, const qs_params* params, qs_solver** out);
int qs_solver_destroy(qs_solver* s);
/* friction.cu:191-192 (Ra field).  Ra: device Nx*Ny, or NULL = uniform alpha. */
int qs_solver_set_activity(qs_solver* s, const Qreal* Ra);
/* Device spectra in the reference layout (copied in / out). */
int qs_solver_set_state(qs_solver* s, const Qcomp* w, const Qcomp* r1, const Qcomp* r2);
int qs_solver_get_state(qs_solver* s, Qcomp* w, Qcomp* r1, Qcomp* r2);
/* nsteps IFRK2 steps (friction.cu:212-260 with TimeIntegration/IFRK2.cuh:16-51,
 * IFh := exp(L dt)), fused into four kernel launches per step (row pass and
 * column stage, twice) plus one column prologue after set_state. */
int qs_step_ifrk2(qs_solver* s, int nsteps);
/* nsteps IF-RK4 steps: friction.cu:212-260 with TimeIntegration/RK4.cuh:12-74
 * (IF = exp(L dt), IFh = exp(L dt/2), as the reference's driver), fused into
 * four row passes and four column stages per step. */
int qs_step_rk4(qs_solver* s, int nsteps);
/* nsteps IF-Euler steps (TimeIntegration/IFE.cuh:15-25, then dealias), fused
 * into one row pass and one column stage per step. */
int qs_step_ife(qs_solver* s, int nsteps);
/* Same step built from the per-op entry points above (reference call
 * granularity); for parity and for timing the unfused structure.  Runs on
 * every supported mesh. */
int qs_step_ifrk2_unfused(qs_solver* s, int nsteps);
/* The reference driver's RK4 step loop itself (friction.cu:212-260 with
 * TimeIntegration/RK4.cuh:12-74) through the per-op entry points. */
int qs_step_rk4_unfused(qs_solver* s, int nsteps);
/* 1 if a non-finite value reached the state since the last query (the
 * reference checks the snapshot's S for NaN, friction.cu:280).  Synchronises,
 * then clears the flag.  While it is set, qs_step_ifrk2 / _rk4 / _ife return
 * QS_ERR_NONFINITE without running (they notice it without synchronising,
 * from an asynchronous copy each of them ends with). */
int qs_solver_nonfinite(qs_solver* s, int* flag);
/* No reference counterpart (the reference has no timers).  While enabled,
 * every kernel of qs_step_ifrk2 is bracketed by HIP events on the Mesh's
 * stream; qs_solver_profile_read synchronises and returns the average
 * duration in ms and the launch count of each kernel class
 * [0] column prologue, [1] row pass, [2] intermediate column stages (IFRK2
 * stage 1, RK4 stages 1-3), [3] final column stages (IFRK2 stage 2, RK4
 * stage 4, IFE),
 * then forgets the recorded events. */
int qs_solver_profile(qs_solver* s, int enable);
int qs_solver_profile_read(qs_solver* s, double* avg_ms, int* counts);
/* Test hooks (no reference counterpart): run one kernel of the fused step
 * (0 column prologue, 1 row pass, 2 column stage 1, 3 column stage 2) and
 * copy an internal array out in the reference layout (0..6 mixed arrays,
 * 7..9 state w/r1/r2, 10..12 stage-1 partial sums). */
int qs_solver_debug_phase(qs_solver* s, int phase);
int qs_solver_debug_array(qs_solver* s, int which, Qcomp* out);
/* Test hook: the fused kernels' per-wavefront FFT (length n = 2^a in
 * [16, 1024] or 3 * 2^a in [48, 768], sign -1 forward / +1 inverse,
 * unnormalised) on 1024/n (resp. 768/n) contiguous vectors, transformed one
 * wave's worth of vectors at a time. */
int qs_debug_wave_fft(const Mesh* mesh, int n, int sign, const Qcomp* in, Qcomp* out);

/* No reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops): nsteps fused IFRK2
 * steps, the kernels and results of qs_solver_set_state + qs_step_ifrk2, on
 * the caller's state spectra w, r1, r2 (device, reference layout, updated in
 * place), with parameters *p and activity field Ra (device Nx*Ny, or NULL for
 * the uniform p->alpha), on the Mesh's stream.  nonfinite: device int set to
 * 1 (0) if a non-finite value reached (did not reach) the state, or NULL.
 * The working arrays live in one fixed device array per GPU (21 MB, held
 * until the process ends): meshes of at most 512 x 257 spectral entries per
 * array (512 x 512, 1024 x 256, ...), larger ones are QS_ERR_UNSUPPORTED_SIZE.
 * Calls no allocating routine.  All calls on one GPU share that array: calls
 * that overlap on one GPU (several streams or threads) race. */
int qs_step_ifrk2_noalloc(const Mesh* mesh, const qs_params* p, const Qreal* Ra, Qcomp* w, Qcomp* r1, Qcomp* r2,
                          int nsteps, int* nonfinite);

/* ---- discrete adjoint of the IFRK2 step ----------------------------------
 * No reference counterpart (the reference has no adjoint).  Exact transpose of
 * qs_step_ifrk2 in the real coordinates of the stored arrays (a spectral entry
 * is two reals; gradients of spectra are returned as d/dRe + i d/dIm).  The
 * control is the activity: the solver's activity field (qs_solver_set_activity;
 * the uniform alpha when none is set) or, with alpha_t, a uniform activity per
 * step.  All pointers are device memory; spectra in the reference layout. */
/* nsteps fused steps from the current state, keeping the state every
 * ckpt_every steps (<= 0: only the initial state) and the final state.
 * alpha_t: [nsteps] uniform activity of each step, or NULL. */
int qs_adj_forward(qs_solver* s, int nsteps, int ckpt_every, const Qreal* alpha_t);
/* Reverse sweep of the last qs_adj_forward.  gW, gR1, gR2: dL/d(final state).
 * Outputs (each may be NULL): abar = dL/dRa [Ny*Nx]; abar_t = dL/dalpha_t
 * [nsteps]; uW, uR1, uR2 = dL/d(initial state).  Segments are reversed last
 * to first, each with every stage's inverse set (7 spectra per stage) at hand:
 * qs_adj_forward records the last segment's, the others are recomputed from
 * their checkpoints (beside the reverse sweep, on a second stream, when one
 * launch leaves compute units idle; QS_ADJ_OVERLAP=0|1 in the environment
 * forces the choice).  Leaves the solver at the final state. */
int qs_adj_backward(qs_solver* s, const Qcomp* gW, const Qcomp* gR1, const Qcomp* gR2,
                    Qreal* abar, Qreal* abar_t, Qcomp* uW, Qcomp* uR1, Qcomp* uR2);
/* Loss of a vorticity spectrum W and dL/dW: kind 0 = 1/2 mean(omega^2) with
 * omega = BwdTrans(W); kind 1 = 1/2 mean(u^2 + v^2) with (u, v) from vel_func
 * (QFldfuncs.cuh:204-224).  *loss is host memory; synchronises the stream. */
int qs_adj_loss(const Mesh* mesh, int kind, const Qcomp* W, Qreal* loss, Qcomp* Wbar);
/* Device bytes held by the adjoint of s. */
int qs_adj_memory(qs_solver* s, size_t* bytes);
/* Frees the adjoint's device memory (qs_solver_destroy also does). */
int qs_adj_release(qs_solver* s);

/* ---- ensembles of independent runs ---------------------------------------
 * No reference counterpart (the reference runs one mesh per process).  n runs
 * of square meshes N[r] (powers of two, 16..1024) with their own parameters
 * (activity alpha, friction Rf, ...), advanced together: per mesh size, every
 * fused kernel covers all runs of that size; sizes run as parallel branches of
 * a captured hipGraph.  Each run's results equal those of its own qs_solver. */
typedef struct qs_ensemble qs_ensemble;
typedef struct qs_ensemble_config {
  int graph_steps;     /* steps per captured graph launch (0: direct launches) */
  int concurrent;      /* 1: size classes as parallel branches */
  int order;           /* work lists: 0 run-major, 1 job-major (interleaved runs) */
  int row_blocks[8];   /* persistent blocks of the row kernel per size class, 0 = one per job */
  int col_blocks[8];   /* same for the column kernels (classes in ascending size) */
  int lanes[8];        /* elements per lane of the per-wave FFT per size class (its first-pass radix;
                          rows per row job and column-tile width follow), 0 = the library default;
                          sides 64, 128, 256 also take 4 or 8 (256: 4 or 16), 512 also takes 8 */
  int merged;          /* 1: one persistent kernel per phase over the jobs of every class (16 elements
                          per lane for every side; lanes and concurrent are then unused, row_blocks[0] /
                          col_blocks[0] give its block counts, 0 = one per CU / one per job) */
} qs_ensemble_config;
int qs_ensemble_create(int n, const int* N, const qs_params* params, qs_ensemble** out);
int qs_ensemble_destroy(qs_ensemble* e);
int qs_ensemble_set_stream(qs_ensemble* e, void* hip_stream);
int qs_ensemble_size(const qs_ensemble* e, int* n);
/* The run's solver (for qs_adj_* or per-run stepping); owned by the ensemble. */
int qs_ensemble_solver(qs_ensemble* e, int run, qs_solver** out);
int qs_ensemble_set_state(qs_ensemble* e, int run, const Qcomp* w, const Qcomp* r1, const Qcomp* r2);
int qs_ensemble_get_state(qs_ensemble* e, int run, Qcomp* w, Qcomp* r1, Qcomp* r2);
int qs_ensemble_set_activity(qs_ensemble* e, int run, const Qreal* Ra);
/* nsteps IFRK2 steps of every run. */
int qs_ensemble_step(qs_ensemble* e, int nsteps);
int qs_ensemble_get_config(const qs_ensemble* e, qs_ensemble_config* cfg);
int qs_ensemble_set_config(qs_ensemble* e, const qs_ensemble_config* cfg);
/* Autotuner: coordinate descent over branch concurrency, work-list order,
 * per-class FFT lane layout (radix / tile width) and block counts, each
 * candidate timed over trial_steps directly launched steps, then over the
 * graph depth (0, 5, 10, 20 steps per captured graph, each candidate captured
 * and timed through graph launches); the states are saved and restored.  Sets
 * and returns the fastest configuration and its time per ensemble step in ms. */
int qs_ensemble_tune(qs_ensemble* e, int trial_steps, qs_ensemble_config* best, double* best_ms);

/* Batched discrete adjoint of an ensemble whose runs share one mesh size: the
 * qs_adj_* sequence for every run at once (each launch covers all runs).
 * alpha_t: device [n][nsteps] uniform activity per run and step, or NULL. */
int qs_ensemble_adj_forward(qs_ensemble* e, int nsteps, int ckpt_every, const Qreal* alpha_t);
/* After qs_ensemble_adj_forward: losses[r] (host) of every run's final state
 * (kind as qs_adj_loss) and their gradients as the seeds of the reverse sweep.
 * Synchronises. */
int qs_ensemble_adj_loss(qs_ensemble* e, int kind, Qreal* losses);
/* Reverse sweep for the sum of the losses.  Outputs (device, may be NULL):
 * abar_t [n][nsteps] = dL/dalpha_t, abar [n][Ny*Nx] = dL/dRa. */
int qs_ensemble_adj_backward(qs_ensemble* e, Qreal* abar_t, Qreal* abar);
int qs_ensemble_adj_memory(qs_ensemble* e, size_t* bytes);

/* ---- diagnostics (device reductions; results in device memory) ----------- */
/* Etot_func (InertialFriction/post_process/postprocess.cu:54-65):
 * *E = sum (u^2 + v^2) * 0.5 * dx * dy over the physical arrays u, v; the
 * driver divides by Lx*Ly (:161).  Deterministic (fixed summation order). */
int qs_Etot_func(const Mesh* mesh, const Qreal* u, const Qreal* v, Qreal* E);
/* Espec_func (post_process/EnergySpec.cu:81-96): *E = sum of |ut|^2 + |vt|^2
 * over the half-spectrum bins with k^2 <= kx^2 + ky^2 < (k + alpha)^2, divided
 * by 2 alpha Nx^2 Ny^2.  ut, vt: velocity spectra (qs_vel_func). */
int qs_Espec_func(const Mesh* mesh, Qreal k, const Qcomp* ut, const Qcomp* vt, Qreal alpha, Qreal* E);
/* Every shell of the driver loop EnergySpec.cu:198-204 (k = 0, dk, 2 dk, ...
 * by repeated addition while k < kmax) in one pass: E[0..*nshell) (device),
 * QS_ERR_INVALID_ARG when there are more than nmax shells. */
int qs_Espec_shells(const Mesh* mesh, const Qcomp* ut, const Qcomp* vt, Qreal dk, Qreal kmax, int nmax,
                    Qreal* E, int* nshell);

/* ---- snapshot and restart I/O -------------------------------------------
 * CSV text in the reference's format (C++ stream defaults, 6 significant
 * digits, one comma-terminated line per row j).  f arrays are HOST memory
 * here, as in the reference. */
int qs_field_visual(const Qreal* f, const char* name, int Nx, int Ny);   /* BasicUtils.cuh:38-50 */
int qs_wave_visual(const Qreal* f, const char* name, int Nxh, int Ny);   /* EnergySpec.cu:8-20 */
int qs_spec_visual(const Qcomp* f, const char* name, int Nxh, int Ny);   /* EnergySpec.cu:22-34 */
/* BasicUtils.cuh:52-67 (x = dx i, y = dy j as float); NULL names = "x.csv", "y.csv". */
int qs_coord(Qreal dx, Qreal dy, int Nx, int Ny, const char* xname, const char* yname);
/* friction.cu:62-97.  A missing file or a row count / row length other than
 * the mesh's is an error (the reference exits or returns leaving f as it was). */
int qs_file_init(const char* filename, Qreal* f, const Mesh* mesh);
/* friction.cu:262-279: <prefix><step>w.csv, r1.csv, r2.csv (BwdTrans of the
 * state, whose column-0 symmetrisation lands in the state as there) and
 * S.csv; QS_ERR_NONFINITE when S[0] is NaN. */
int qs_solver_snapshot_csv(qs_solver* s, const char* prefix, int step);
/* friction.cu:163-185: state from three physical CSVs (file_init, FwdTrans,
 * the BwdTrans round trip that symmetrises column 0). */
int qs_solver_init_csv(qs_solver* s, const char* w_csv, const char* r1_csv, const char* r2_csv);
/* No reference counterpart (the reference restarts from its 6-digit CSVs):
 * exact binary checkpoint of the state spectra and the activity field, with
 * the mesh, the parameters, a caller's step counter and a checksum.  load
 * requires the same mesh; it restores state and activity and returns the
 * stored step and parameters (each may be NULL) without applying the latter. */
int qs_solver_save(qs_solver* s, const char* path, int64_t step);
int qs_solver_load(qs_solver* s, const char* path, int64_t* step, qs_params* params);

/* ---- host helpers (no GPU) ---------------------------------------------- */
/* splitmix64 uniform [0,1) stream: u_k = (x_k >> 11) * 2^-53. */
int qs_synth_uniform(uint64_t seed, Qreal lo, Qreal hi, size_t n, Qreal* host_out);

#ifdef __cplusplus
}
#endif
#endif /* QSOLVER_H */
