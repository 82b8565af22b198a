/*
 * qs_oracle.h -- FP64 CPU oracle of the Qsolver step (TEST INFRASTRUCTURE).
 *
 * This library is the correctness checker for the HIP implementation.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it,
 * and only as the checker or the timed CPU baseline -- never as the product
 * path.  The product library (libqsolver.so) does not link or call it.
 *
 * It restates, call by call, the reference's single-GPU Qsolver code:
 *   Basic/Mesh.cuh:6-46                 wave tables and 2/3-rule mask
 *   FldOp/QFldfuncs.cuh:8-224           transforms and field operations
 *   InertialFriction/friction/Qsolver.cuh:20-322   physics (canonical copy)
 *   TimeIntegration/IFE.cuh:15-25, IFRK2.cuh:16-51, RK4.cuh:12-74
 *   InertialFriction/friction/friction.cu:212-260  per-step call order
 * with cuFFT replaced by an own mixed-radix FFT (forward e^{-i}, inverse
 * e^{+i}, both unnormalised, as cuFFT's D2Z / Z2D).
 *
 * C2R semantics (cuFFT's treatment of a non-Hermitian half spectrum is not
 * documented, so the semantics are fixed here and shared with the HIP path):
 *   f = Re( IDFT_x( G ) ) / (Nx*Ny),  G = IDFT_y(F) column by column, where
 *   the x-direction C2R discards the imaginary parts of its DC and Nyquist
 *   bins.  Equivalently: the Hermitian part of column 0 and column Nx/2 is
 *   taken before a Hermitian C2R.
 */
#ifndef QS_ORACLE_H
#define QS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y; } or_comp;

/* Mesh + Field mirrors (Basic/Mesh.cuh:48-129).  Opaque handles. */
void*    or_mesh_new(int Nx, int Ny, double Lx, double Ly);
void     or_mesh_free(void* mesh);
void     or_tables(void* mesh, double* kx, double* ky, double* k2, double* cutoff);
void*    or_field_new(void* mesh);
void     or_field_free(void* field);
or_comp* or_field_spec(void* field);
double*  or_field_phys(void* field);

/* Raw 1-D FFT of the oracle (for the brute-force DFT check). sign=-1 fwd. */
void or_fft1d(int n, int sign, const or_comp* in, or_comp* out);

/* FldOp/QFldfuncs.cuh:45-55 */
void or_FwdTrans(or_comp* ft, const double* f, void* mesh);
void or_BwdTrans(double* f, or_comp* ft, void* mesh);

/* InertialFriction/friction/Qsolver.cuh composites, reference signatures
 * (Field handles in the reference's argument order). */
void or_single(void* s0, void* r, void* S, double lambda);
void or_cross(void* cs, void* r1, void* r2, void* aux);
void or_p11func(void* p11, void* r1, void* S, void* Ra, double lambda);
void or_p12func(void* p12, void* r1, void* r2, void* S, void* Ra, double lambda,
                void* aux, void* aux1);
void or_p21func(void* p21, void* r1, void* r2, void* S, void* Ra, double lambda,
                void* aux, void* aux1);
void or_wnonl_func(void* wnonl, void* wcurr, void* u, void* v, void* r1curr,
                   void* r2curr, void* Ra, void* S, void* p11, void* p12,
                   void* p21, double Re, double Er, double lambda, void* aux,
                   void* aux1);
void or_r1nonl_func(void* r1nonl, void* wcurr, void* u, void* v, void* r1curr,
                    void* r2curr, void* S, double lambda, void* aux, void* aux1);
void or_r2nonl_func(void* r2nonl, void* wcurr, void* u, void* v, void* r1curr,
                    void* r2curr, void* S, double lambda, void* aux, void* aux1);
void or_curr_func(void* r1curr, void* r2curr, void* wcurr, void* u, void* v,
                  void* S);

/*
 * Solver: the whole step with the generalised coefficients.
 * p[] = {lambda, Re, Er, Rf, alpha, D_Q, a_Q, c_Q, dt, advect_w}
 * D_Q scales the Laplacian of the Q equation (IF table, single, cross);
 * a_Q is the Q-linear coefficient; c_Q scales S^2.  D_Q=a_Q=c_Q=1 and
 * advect_w=1 is exactly the reference (friction/Qsolver.cuh).
 */
enum { OR_IFE = 0, OR_IFRK2 = 1, OR_RK4 = 2 };
void* or_solver_new(void* mesh, const double* p);
void  or_solver_free(void* s);
/* Ra: physical activity field (Nx*Ny) or NULL for the uniform alpha. */
void  or_solver_set_activity(void* s, const double* Ra);
void  or_solver_set_state(void* s, const or_comp* w, const or_comp* r1, const or_comp* r2);
void  or_solver_get_state(void* s, or_comp* w, or_comp* r1, or_comp* r2);
void  or_solver_step(void* s, int method, int nsteps);
/* One RHS evaluation on given stage spectra (curr_func + w/r1/r2 nonl). The
 * inputs are Hermitian-symmetrised in column 0 in place (BwdTrans side effect). */
void  or_solver_rhs(void* s, or_comp* w, or_comp* r1, or_comp* r2,
                    or_comp* wn, or_comp* r1n, or_comp* r2n);
/* Integrating-factor tables of B1/B2 (IF = exp(L dt), IFh = exp(L dt/2)). */
void  or_solver_lin(void* s, double* IFw, double* IFwh, double* IFr1,
                    double* IFr1h, double* IFr2, double* IFr2h);

/*
 * Discrete adjoint of the IFRK2 step (no reference counterpart).  Exact
 * transposes in the real inner products of the stored arrays:
 *   B = A10 C2R / (Nx Ny), B^T;  F = unnormalised R2C, F^T;
 *   rhs_vjp: (X, Nbar) -> Xbar = (dN/dX)^T Nbar, abar += (dN/dalpha)^T Nbar
 *   step_vjp: (u_n, dL/du_{n+1}) -> dL/du_n, abar += ...
 *   loss_grad: kind 0 = 1/2 mean(w^2), kind 1 = 1/2 mean(u^2 + v^2); returns L.
 *   gradient: nsteps from the solver's state; alpha_t (nsteps, or NULL) gives
 *   a uniform activity per step and abar_t its per-step gradient sums.
 * The solver's activity field (or_solver_set_activity) is the control alpha(x).
 */
void   or_adj_B(void* s, const or_comp* G, double* f);
void   or_adj_BT(void* s, const double* f, or_comp* G);
void   or_adj_F(void* s, const double* f, or_comp* G);
void   or_adj_FT(void* s, const or_comp* G, double* f);
void   or_adj_rhs_vjp(void* s, const or_comp* W, const or_comp* R1, const or_comp* R2,
                      const or_comp* Nbw, const or_comp* Nbr1, const or_comp* Nbr2,
                      or_comp* Wb, or_comp* R1b, or_comp* R2b, double* abar);
void   or_adj_step_vjp(void* s, const or_comp* W, const or_comp* R1, const or_comp* R2,
                       const or_comp* gW, const or_comp* gR1, const or_comp* gR2,
                       or_comp* uW, or_comp* uR1, or_comp* uR2, double* abar);
double or_adj_loss_grad(void* s, const or_comp* W, int kind, or_comp* Wb);
double or_adj_gradient(void* s, int nsteps, int kind, const double* alpha_t,
                       double* abar_field, double* abar_t);

/* Threads used by the oracle's OpenMP loops (0 = library default). */
/* Diagnostics: Etot_func (post_process/postprocess.cu:54-65) and Espec_func
 * (post_process/EnergySpec.cu:81-96, one shell; shells: its driver loop). */
double or_Etot_func(const double* u, const double* v, int Nx, int Ny, double dx, double dy);
double or_Espec_func(void* mesh, double k, const or_comp* ut, const or_comp* vt, double alpha);
int    or_Espec_shells(void* mesh, const or_comp* ut, const or_comp* vt, double dk, double kmax, int nmax,
                       double* E);

void  or_set_threads(int n);
int   or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
