// qsolver.hpp -- reference-named C++ wrappers over the C-ABI of libqsolver.so.
//
// Reference host code (whpy/nature-control) ports call by call: delete the
// launch configuration `<<<mesh->dimGrid.., mesh->dimBlock..>>>` and the
// trailing `Nx|Nxh, Ny, BSZ` size arguments, and pass the Mesh instead.  For
// example the reference's
//     FldMul<<<mesh->dimGridp, mesh->dimBlockp>>>(pc, pa, pb, C, Nx, Ny, BSZ);
//     IFRK2::integrate_func1D<<<...>>>(old, curr, new_, nonl, IF, IFh, Nxh, Ny, BSZ, dt);
// become
//     qsolver::FldMul(pc, pa, pb, C, mesh);
//     qsolver::IFRK2::integrate_func1D(old, curr, new_, nonl, IF, IFh, dt, mesh);
// The wrappers throw qsolver::error (with qs_last_error()'s message) where the
// reference's cudaErr.h:54-72 would abort; the C-ABI underneath never throws.
#pragma once

#include <stdexcept>
#include <string>

#include "qsolver.h"

namespace qsolver {

struct error : std::runtime_error {
  int status;
  error(int s, const std::string& where)
      : std::runtime_error(where + ": " + qs_status_string(s) + ": " + qs_last_error()), status(s) {}
};

inline void check(int s, const char* where) {
  if (s != QS_OK) throw error(s, where);
}

// ---- Basic/Mesh.cuh:48-129 ----------------------------------------------
inline Mesh* make_mesh(int Nx, int Ny, Qreal Lx, Qreal Ly, int BSZ = 16) {
  Mesh* m = nullptr;
  check(qs_mesh_create(Nx, Ny, Lx, Ly, BSZ, &m), "Mesh");
  return m;
}
inline Field* make_field(Mesh* m) {
  Field* f = nullptr;
  check(qs_field_create(m, &f), "Field");
  return f;
}

// ---- FldOp/QFldfuncs.cuh ------------------------------------------------
inline void coeff(Qreal* f, const Mesh* m) { check(qs_coeff(f, m), "coeff"); }
inline void symmetry_func(Qcomp* f, const Mesh* m) { check(qs_symmetry_func(f, m), "symmetry_func"); }
inline void FwdTrans(Qcomp* ft, const Qreal* f, const Mesh* m) { check(qs_FwdTrans(ft, f, m), "FwdTrans"); }
inline void BwdTrans(Qreal* f, Qcomp* ft, const Mesh* m) { check(qs_BwdTrans(f, ft, m), "BwdTrans"); }
inline void FldMul(Qreal* pc, const Qreal* pa, const Qreal* pb, Qreal C, const Mesh* m) {
  check(qs_FldMul(pc, pa, pb, C, m), "FldMul");
}
inline void FldAdd(Qreal* pc, Qreal a, const Qreal* pa, Qreal b, const Qreal* pb, const Mesh* m) {
  check(qs_FldAdd(pc, a, pa, b, pb, m), "FldAdd");
}
inline void FldAdd(Qreal* pc, Qreal a, const Qreal* pa, Qreal b, const Mesh* m) {
  check(qs_FldAddScalar(pc, a, pa, b, m), "FldAdd");
}
inline void FldSet(Qreal* pf, Qreal c, const Mesh* m) { check(qs_FldSet(pf, c, m), "FldSet"); }
inline void FldSet(Qreal* pf, const Qreal* c, const Mesh* m) { check(qs_FldCopy(pf, c, m), "FldSet"); }
inline void SpecSet(Qcomp* pa, const Qcomp* pb, const Mesh* m) { check(qs_SpecSet(pa, pb, m), "SpecSet"); }
inline void SpecSet(Qcomp* pa, Qcomp c, const Mesh* m) { check(qs_SpecSetConst(pa, c, m), "SpecSet"); }
inline void SpecAdd(Qcomp* spc, Qreal a, const Qcomp* spa, Qreal b, const Qcomp* spb, const Mesh* m) {
  check(qs_SpecAdd(spc, a, spa, b, spb, m), "SpecAdd");
}
inline void SpecMul(Qcomp* spc, const Qcomp* spa, Qreal C, const Mesh* m) {
  check(qs_SpecMul(spc, spa, C, m), "SpecMul");
}
inline void xDeriv(Qcomp* fx, const Qcomp* f, const Mesh* m) { check(qs_xDeriv(fx, f, m), "xDeriv"); }
inline void yDeriv(Qcomp* fy, const Qcomp* f, const Mesh* m) { check(qs_yDeriv(fy, f, m), "yDeriv"); }
inline void Laplacian(Qcomp* fL, const Qcomp* f, const Mesh* m) { check(qs_Laplacian(fL, f, m), "Laplacian"); }
inline void vel_func(const Qcomp* w, Qcomp* u, Qcomp* v, const Mesh* m) { check(qs_vel_func(w, u, v, m), "vel_func"); }

// ---- InertialFriction/friction/Qsolver.cuh -------------------------------
inline void wlin_func(Qreal* IFw, Qreal* IFwh, Qreal Re, Qreal Rf, Qreal dt, const Mesh* m) {
  check(qs_wlin_func(IFw, IFwh, Re, Rf, dt, m), "wlin_func");
}
inline void r1lin_func(Qreal* IF, Qreal* IFh, Qreal dt, const Mesh* m) { check(qs_r1lin_func(IF, IFh, dt, m), "r1lin_func"); }
inline void r2lin_func(Qreal* IF, Qreal* IFh, Qreal dt, const Mesh* m) { check(qs_r2lin_func(IF, IFh, dt, m), "r2lin_func"); }
inline void S_func(Qreal* S, const Qreal* r1, const Qreal* r2, const Mesh* m) { check(qs_S_func(S, r1, r2, m), "S_func"); }
inline void Rar(Qreal* rar, const Qreal* r, const Qreal* Ra, const Mesh* m) { check(qs_Rar(rar, r, Ra, m), "Rar"); }
inline void single1(Qreal* s1, const Qreal* r, const Qreal* S, Qreal lambda, const Mesh* m) {
  check(qs_single1(s1, r, S, lambda, m), "single1");
}
inline void S2r(Qreal* rs, const Qreal* r, const Qreal* S, const Mesh* m) { check(qs_S2r(rs, r, S, m), "S2r"); }
inline void PiDeriv(Qcomp* dP, const Qcomp* p12, const Qcomp* p11, const Qcomp* p21, Qreal Er, Qreal Re,
                    const Mesh* m) {
  check(qs_PiDeriv(dP, p12, p11, p21, Er, Re, m), "PiDeriv");
}
inline void dealias(Qcomp* ft, const Mesh* m) { check(qs_dealias(ft, m), "dealias"); }
inline void single(Field* s0, Field* r, Field* S, Qreal lambda) { check(qs_single(s0, r, S, lambda), "single"); }
inline void cross(Field* cs, Field* r1, Field* r2, Field* aux) { check(qs_cross(cs, r1, r2, aux), "cross"); }
inline void p11func(Field* p11, Field* r1, Field* S, Field* Ra, Qreal lambda) {
  check(qs_p11func(p11, r1, S, Ra, lambda), "p11func");
}
inline void p12func(Field* p12, Field* r1, Field* r2, Field* S, Field* Ra, Qreal lambda, Field* aux, Field* aux1) {
  check(qs_p12func(p12, r1, r2, S, Ra, lambda, aux, aux1), "p12func");
}
inline void p21func(Field* p21, Field* r1, Field* r2, Field* S, Field* Ra, Qreal lambda, Field* aux, Field* aux1) {
  check(qs_p21func(p21, r1, r2, S, Ra, lambda, aux, aux1), "p21func");
}
inline void wnonl_func(Field* wnonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr, Field* Ra,
                       Field* S, Field* p11, Field* p12, Field* p21, Qreal Re, Qreal Er, Qreal lambda, Field* aux,
                       Field* aux1) {
  check(qs_wnonl_func(wnonl, wcurr, u, v, r1curr, r2curr, Ra, S, p11, p12, p21, Re, Er, lambda, aux, aux1),
        "wnonl_func");
}
inline void r1nonl_func(Field* r1nonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr, Field* S,
                        Qreal lambda, Field* aux, Field* aux1) {
  check(qs_r1nonl_func(r1nonl, wcurr, u, v, r1curr, r2curr, S, lambda, aux, aux1), "r1nonl_func");
}
inline void r2nonl_func(Field* r2nonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr, Field* S,
                        Qreal lambda, Field* aux, Field* aux1) {
  check(qs_r2nonl_func(r2nonl, wcurr, u, v, r1curr, r2curr, S, lambda, aux, aux1), "r2nonl_func");
}
inline void curr_func(Field* r1curr, Field* r2curr, Field* wcurr, Field* u, Field* v, Field* S) {
  check(qs_curr_func(r1curr, r2curr, wcurr, u, v, S), "curr_func");
}

// ---- TimeIntegration/{IFE,IFRK2,RK4}.cuh ---------------------------------
namespace IFE {
inline void integrate_func1D(const Qcomp* old_spec, Qcomp* new_spec, const Qcomp* nonl, const Qreal* IF,
                             const Qreal* IFh, Qreal dt, const Mesh* m) {
  check(qs_ife_func1D(old_spec, new_spec, nonl, IF, IFh, dt, m), "IFE::integrate_func1D");
}
}  // namespace IFE

namespace IFRK2 {
inline void integrate_func0D(const Qcomp* old_spec, Qcomp* curr, Qcomp* new_spec, const Qreal* IF, const Qreal* IFh,
                             Qreal dt, const Mesh* m) {
  check(qs_ifrk2_func0D(old_spec, curr, new_spec, IF, IFh, dt, m), "IFRK2::integrate_func0D");
}
inline void integrate_func1D(const Qcomp* old_spec, Qcomp* curr, Qcomp* new_spec, const Qcomp* nonl, const Qreal* IF,
                             const Qreal* IFh, Qreal dt, const Mesh* m) {
  check(qs_ifrk2_func1D(old_spec, curr, new_spec, nonl, IF, IFh, dt, m), "IFRK2::integrate_func1D");
}
inline void integrate_func2D(const Qcomp* old_spec, Qcomp* curr, Qcomp* new_spec, const Qcomp* nonl, const Qreal* IF,
                             const Qreal* IFh, Qreal dt, const Mesh* m) {
  check(qs_ifrk2_func2D(old_spec, curr, new_spec, nonl, IF, IFh, dt, m), "IFRK2::integrate_func2D");
}
}  // namespace IFRK2

namespace RK4 {
inline void integrate_func0D(const Qcomp* old_spec, Qcomp* curr, Qcomp* new_spec, const Qreal* IF, const Qreal* IFh,
                             Qreal dt, const Mesh* m) {
  check(qs_rk4_func0D(old_spec, curr, new_spec, IF, IFh, dt, m), "RK4::integrate_func0D");
}
#define QSOLVER_RK4_STAGE(n)                                                                              \
  inline void integrate_func##n##D(const Qcomp* old_spec, Qcomp* curr, Qcomp* new_spec, const Qcomp* nonl, \
                                   const Qreal* IF, const Qreal* IFh, Qreal dt, const Mesh* m) {          \
    check(qs_rk4_func##n##D(old_spec, curr, new_spec, nonl, IF, IFh, dt, m), "RK4::integrate_func" #n "D"); \
  }
QSOLVER_RK4_STAGE(1)
QSOLVER_RK4_STAGE(2)
QSOLVER_RK4_STAGE(3)
QSOLVER_RK4_STAGE(4)
#undef QSOLVER_RK4_STAGE
}  // namespace RK4

// ---- fused solver (no single reference counterpart: friction.cu:150-260) --
class Solver {
 public:
  Solver(Mesh* m, const qs_params& p) { check(qs_solver_create(m, &p, &s_), "qs_solver_create"); }
  ~Solver() { qs_solver_destroy(s_); }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;
  void set_activity(const Qreal* Ra) { check(qs_solver_set_activity(s_, Ra), "qs_solver_set_activity"); }
  void set_state(const Qcomp* w, const Qcomp* r1, const Qcomp* r2) {
    check(qs_solver_set_state(s_, w, r1, r2), "qs_solver_set_state");
  }
  void get_state(Qcomp* w, Qcomp* r1, Qcomp* r2) { check(qs_solver_get_state(s_, w, r1, r2), "qs_solver_get_state"); }
  void step_ifrk2(int nsteps) { check(qs_step_ifrk2(s_, nsteps), "qs_step_ifrk2"); }
  qs_solver* handle() { return s_; }

 private:
  qs_solver* s_ = nullptr;
};

}  // namespace qsolver
