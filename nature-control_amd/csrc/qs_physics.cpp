// qs_physics.cpp -- host-side composites of InertialFriction/friction/Qsolver.cuh.
//
// Same call sequence as the reference wrappers, each launch replaced by the
// corresponding qs_* entry point (all stream-ordered on the Mesh's stream).
#include "qs_internal.hpp"

using namespace qs;

#define FLD(f) \
  if (!(f) || !(f)->mesh) return fail(QS_ERR_INVALID_ARG, "%s: null Field " #f, __func__)

namespace qs {

// friction/Qsolver.cuh:96-105
int single_c(Field* s0, Field* r, Field* S, Qreal lambda, const Coef& c) {
  FLD(s0); FLD(r); FLD(S);
  Mesh* mesh = s0->mesh;
  QS_TRY(qs_Laplacian(s0->spec, r->spec, mesh));
  QS_TRY(qs_BwdTrans(s0->phys, s0->spec, mesh));
  QS_TRY(qs_FldMul(s0->phys, s0->phys, S->phys, -1.0 * lambda * c.DQ, mesh));
  return single1_c(s0->phys, r->phys, S->phys, lambda, c.cQ, c.aQ, mesh);
}

// friction/Qsolver.cuh:107-122
int cross_c(Field* cs, Field* r1, Field* r2, Field* aux, const Coef& c) {
  FLD(cs); FLD(r1); FLD(r2); FLD(aux);
  Mesh* mesh = cs->mesh;
  QS_TRY(qs_Laplacian(cs->spec, r1->spec, mesh));
  QS_TRY(qs_BwdTrans(cs->phys, cs->spec, mesh));
  QS_TRY(qs_FldMul(cs->phys, cs->phys, r2->phys, 2.0 * c.DQ, mesh));
  QS_TRY(qs_Laplacian(aux->spec, r2->spec, mesh));
  QS_TRY(qs_BwdTrans(aux->phys, aux->spec, mesh));
  QS_TRY(qs_FldMul(aux->phys, r1->phys, aux->phys, -2.0 * c.DQ, mesh));
  return qs_FldAdd(cs->phys, 1.0, cs->phys, 1.0, aux->phys, mesh);
}

// friction/Qsolver.cuh:124-131
int p11func_c(Field* p11, Field* r1, Field* S, Field* Ra, Qreal lambda, const Coef& c) {
  FLD(p11); FLD(r1); FLD(S); FLD(Ra);
  Mesh* mesh = p11->mesh;
  QS_TRY(single_c(p11, r1, S, lambda, c));
  QS_TRY(qs_Rar(p11->phys, r1->phys, Ra->phys, mesh));
  return qs_FwdTrans(p11->spec, p11->phys, mesh);
}

// friction/Qsolver.cuh:133-141
int p12func_c(Field* p12, Field* r1, Field* r2, Field* S, Field* Ra, Qreal lambda, Field* aux,
              Field* aux1, const Coef& c) {
  FLD(p12); FLD(r1); FLD(r2); FLD(S); FLD(Ra); FLD(aux); FLD(aux1);
  Mesh* mesh = p12->mesh;
  QS_TRY(single_c(p12, r2, S, lambda, c));
  QS_TRY(cross_c(aux, r1, r2, aux1, c));
  QS_TRY(qs_FldAdd(p12->phys, 1.0, p12->phys, 1.0, aux->phys, mesh));
  QS_TRY(qs_Rar(p12->phys, r2->phys, Ra->phys, mesh));
  return qs_FwdTrans(p12->spec, p12->phys, mesh);
}

// friction/Qsolver.cuh:143-151
int p21func_c(Field* p21, Field* r1, Field* r2, Field* S, Field* Ra, Qreal lambda, Field* aux,
              Field* aux1, const Coef& c) {
  FLD(p21); FLD(r1); FLD(r2); FLD(S); FLD(Ra); FLD(aux); FLD(aux1);
  Mesh* mesh = p21->mesh;
  QS_TRY(single_c(p21, r2, S, lambda, c));
  QS_TRY(cross_c(aux, r2, r1, aux1, c));
  QS_TRY(qs_FldAdd(p21->phys, 1.0, p21->phys, 1.0, aux->phys, mesh));
  QS_TRY(qs_Rar(p21->phys, r2->phys, Ra->phys, mesh));
  return qs_FwdTrans(p21->spec, p21->phys, mesh);
}

// friction/Qsolver.cuh:166-201
int wnonl_func_c(Field* wnonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr,
                  Field* Ra, Field* S, Field* p11, Field* p12, Field* p21, Qreal Re, Qreal Er,
                 Qreal lambda, Field* aux, Field* aux1, const Coef& c) {
  FLD(wnonl); FLD(wcurr); FLD(u); FLD(v); FLD(r1curr); FLD(r2curr); FLD(Ra); FLD(S);
  FLD(p11); FLD(p12); FLD(p21); FLD(aux); FLD(aux1);
  Mesh* mesh = wnonl->mesh;
  if (!c.advect_w) {  // InertialFriction/beginning/Qsolver.cuh:179-192
    QS_TRY(qs_FldSet(wnonl->phys, 0.0, mesh));
  } else {
  QS_TRY(qs_FldMul(wnonl->phys, u->phys, wcurr->phys, -1.0, mesh));
  QS_TRY(qs_FwdTrans(wnonl->spec, wnonl->phys, mesh));
  QS_TRY(qs_xDeriv(wnonl->spec, wnonl->spec, mesh));
  QS_TRY(qs_BwdTrans(wnonl->phys, wnonl->spec, mesh));

  QS_TRY(qs_FldMul(aux->phys, v->phys, wcurr->phys, -1.0, mesh));
  QS_TRY(qs_FwdTrans(aux->spec, aux->phys, mesh));
  QS_TRY(qs_yDeriv(aux->spec, aux->spec, mesh));
  QS_TRY(qs_BwdTrans(aux->phys, aux->spec, mesh));

  QS_TRY(qs_FldAdd(wnonl->phys, 1.0, wnonl->phys, 1.0, aux->phys, mesh));
  }
  QS_TRY(qs_FwdTrans(wnonl->spec, wnonl->phys, mesh));

  QS_TRY(p11func_c(p11, r1curr, S, Ra, lambda, c));
  QS_TRY(p12func_c(p12, r1curr, r2curr, S, Ra, lambda, aux, aux1, c));
  QS_TRY(p21func_c(p21, r1curr, r2curr, S, Ra, lambda, aux, aux1, c));
  return qs_PiDeriv(wnonl->spec, p12->spec, p11->spec, p21->spec, Er, Re, mesh);
}

// friction/Qsolver.cuh:217-255
int r1nonl_func_c(Field* r1nonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr,
                  Field* S, Qreal lambda, Field* aux, Field* aux1, const Coef& c) {
  FLD(r1nonl); FLD(wcurr); FLD(u); FLD(v); FLD(r1curr); FLD(r2curr); FLD(S); FLD(aux); FLD(aux1);
  Mesh* mesh = r1nonl->mesh;
  QS_TRY(qs_FldMul(r1nonl->phys, wcurr->phys, r2curr->phys, -1.0, mesh));
  QS_TRY(S2r_c(r1nonl->phys, r1curr->phys, S->phys, c.cQ, mesh));
  QS_TRY(qs_FwdTrans(r1nonl->spec, r1nonl->phys, mesh));

  QS_TRY(qs_FldMul(aux->phys, u->phys, r1curr->phys, -1.0, mesh));
  QS_TRY(qs_FwdTrans(aux->spec, aux->phys, mesh));
  QS_TRY(qs_xDeriv(aux->spec, aux->spec, mesh));

  QS_TRY(qs_FldMul(aux1->phys, v->phys, r1curr->phys, -1.0, mesh));
  QS_TRY(qs_FwdTrans(aux1->spec, aux1->phys, mesh));
  QS_TRY(qs_yDeriv(aux1->spec, aux1->spec, mesh));
  QS_TRY(qs_SpecAdd(aux->spec, 1.0, aux->spec, 1.0, aux1->spec, mesh));
  QS_TRY(qs_SpecAdd(r1nonl->spec, 1.0, r1nonl->spec, 1.0, aux->spec, mesh));

  QS_TRY(qs_xDeriv(aux->spec, u->spec, mesh));
  QS_TRY(qs_BwdTrans(aux->phys, aux->spec, mesh));
  QS_TRY(qs_FldMul(aux->phys, S->phys, aux->phys, lambda, mesh));
  QS_TRY(qs_FwdTrans(aux->spec, aux->phys, mesh));
  return qs_SpecAdd(r1nonl->spec, 1.0, r1nonl->spec, 1.0, aux->spec, mesh);
}

// friction/Qsolver.cuh:257-300
int r2nonl_func_c(Field* r2nonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr,
                  Field* S, Qreal lambda, Field* aux, Field* aux1, const Coef& c) {
  FLD(r2nonl); FLD(wcurr); FLD(u); FLD(v); FLD(r1curr); FLD(r2curr); FLD(S); FLD(aux); FLD(aux1);
  Mesh* mesh = r2nonl->mesh;
  QS_TRY(qs_FldMul(r2nonl->phys, wcurr->phys, r1curr->phys, 1.0, mesh));
  QS_TRY(S2r_c(r2nonl->phys, r2curr->phys, S->phys, c.cQ, mesh));
  QS_TRY(qs_FwdTrans(r2nonl->spec, r2nonl->phys, mesh));

  QS_TRY(qs_FldMul(aux->phys, u->phys, r2curr->phys, -1.0, mesh));
  QS_TRY(qs_FwdTrans(aux->spec, aux->phys, mesh));
  QS_TRY(qs_xDeriv(aux->spec, aux->spec, mesh));

  QS_TRY(qs_FldMul(aux1->phys, v->phys, r2curr->phys, -1.0, mesh));
  QS_TRY(qs_FwdTrans(aux1->spec, aux1->phys, mesh));
  QS_TRY(qs_yDeriv(aux1->spec, aux1->spec, mesh));
  QS_TRY(qs_SpecAdd(aux->spec, 1.0, aux->spec, 1.0, aux1->spec, mesh));
  QS_TRY(qs_SpecAdd(r2nonl->spec, 1.0, r2nonl->spec, 1.0, aux->spec, mesh));

  QS_TRY(qs_yDeriv(aux->spec, u->spec, mesh));
  QS_TRY(qs_xDeriv(aux1->spec, v->spec, mesh));
  QS_TRY(qs_SpecAdd(aux->spec, 1.0, aux->spec, 1.0, aux1->spec, mesh));
  QS_TRY(qs_BwdTrans(aux->phys, aux->spec, mesh));
  QS_TRY(qs_FldMul(aux->phys, S->phys, aux->phys, 0.5 * lambda, mesh));
  QS_TRY(qs_FwdTrans(aux->spec, aux->phys, mesh));
  return qs_SpecAdd(r2nonl->spec, 1.0, r2nonl->spec, 1.0, aux->spec, mesh);
}

}  // namespace qs

extern "C" {

int qs_single(Field* s0, Field* r, Field* S, Qreal lambda) { return single_c(s0, r, S, lambda, Coef()); }
int qs_cross(Field* cs, Field* r1, Field* r2, Field* aux) { return cross_c(cs, r1, r2, aux, Coef()); }
int qs_p11func(Field* p11, Field* r1, Field* S, Field* Ra, Qreal lambda) {
  return p11func_c(p11, r1, S, Ra, lambda, Coef());
}
int qs_p12func(Field* p12, Field* r1, Field* r2, Field* S, Field* Ra, Qreal lambda, Field* aux, Field* aux1) {
  return p12func_c(p12, r1, r2, S, Ra, lambda, aux, aux1, Coef());
}
int qs_p21func(Field* p21, Field* r1, Field* r2, Field* S, Field* Ra, Qreal lambda, Field* aux, Field* aux1) {
  return p21func_c(p21, r1, r2, S, Ra, lambda, aux, aux1, Coef());
}
int qs_wnonl_func(Field* wnonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr, Field* Ra,
                  Field* S, Field* p11, Field* p12, Field* p21, Qreal Re, Qreal Er, Qreal lambda, Field* aux,
                  Field* aux1) {
  return wnonl_func_c(wnonl, wcurr, u, v, r1curr, r2curr, Ra, S, p11, p12, p21, Re, Er, lambda, aux, aux1,
                      Coef());
}
int qs_r1nonl_func(Field* r1nonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr, Field* S,
                   Qreal lambda, Field* aux, Field* aux1) {
  return r1nonl_func_c(r1nonl, wcurr, u, v, r1curr, r2curr, S, lambda, aux, aux1, Coef());
}
int qs_r2nonl_func(Field* r2nonl, Field* wcurr, Field* u, Field* v, Field* r1curr, Field* r2curr, Field* S,
                   Qreal lambda, Field* aux, Field* aux1) {
  return r2nonl_func_c(r2nonl, wcurr, u, v, r1curr, r2curr, S, lambda, aux, aux1, Coef());
}

// friction/Qsolver.cuh:302-312
int qs_curr_func(Field* r1curr, Field* r2curr, Field* wcurr, Field* u, Field* v, Field* S) {
  FLD(r1curr); FLD(r2curr); FLD(wcurr); FLD(u); FLD(v); FLD(S);
  Mesh* mesh = r1curr->mesh;
  QS_TRY(qs_BwdTrans(r1curr->phys, r1curr->spec, mesh));
  QS_TRY(qs_BwdTrans(r2curr->phys, r2curr->spec, mesh));
  QS_TRY(qs_BwdTrans(wcurr->phys, wcurr->spec, mesh));
  QS_TRY(qs_S_func(S->phys, r1curr->phys, r2curr->phys, mesh));
  QS_TRY(qs_vel_func(wcurr->spec, u->spec, v->spec, mesh));
  QS_TRY(qs_BwdTrans(u->phys, u->spec, mesh));
  return qs_BwdTrans(v->phys, v->spec, mesh);
}

}  // extern "C"
