// qs_fldops.hip -- per-op entry points at the reference's call granularity.
//
// Each qs_* function replaces one reference kernel or host wrapper; the
// arithmetic of each kernel body is the reference's (FldOp/QFldfuncs.cuh,
// InertialFriction/friction/Qsolver.cuh, TimeIntegration/*.cuh), launched as
// a flat grid-stride loop (64-wide waves, 256 threads) instead of the
// reference's 16x16 blocks.
#include "qs_internal.hpp"

namespace qs {
namespace {

template <class F>
__global__ __launch_bounds__(256) void k_each(long n, F fn) {
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (long)gridDim.x * blockDim.x)
    fn(idx);
}

template <class F>
int each(const Mesh* m, long n, F fn) {
  QS_TRY(check_mesh(m));
  long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_each<F>, dim3((unsigned)blocks), dim3(256), 0, stream_of(m), n, fn);
  QS_CHECK_LAUNCH();
  return QS_OK;
}

inline long np(const Mesh* m) { return (long)m->Nx * m->Ny; }
inline long ns(const Mesh* m) { return (long)m->Nxh * m->Ny; }

// Basic/cuComplexBinOp.cuh operators on {x, y}.
__device__ __forceinline__ double2 d2(Qcomp c) { return make_double2(c.x, c.y); }
__device__ __forceinline__ Qcomp qc(double2 c) { Qcomp r; r.x = c.x; r.y = c.y; return r; }
__device__ __forceinline__ double2 im_times(double k, double2 f) {
  // im() * k * f  ==  (0, k) * f
  return make_double2(0.0 * f.x - k * f.y, 0.0 * f.y + k * f.x);
}

}  // namespace

int check_mesh(const Mesh* m) {
  if (!m || !m->priv) return fail(QS_ERR_INVALID_ARG, "null mesh");
  return QS_OK;
}

// Coefficient-aware forms of r1lin/r2lin_func (friction/Qsolver.cuh:32-54),
// single1 (:86-94) and S2r (:206-215): L_q = -DQ k^2 + aQ,
// s1 += lambda S (cQ S^2 - aQ) r,  rs -= cQ S^2 r.  All 1 = reference.
int qlin_c(double* IF, double* IFh, double dt, double DQ, double aQ, const Mesh* mesh) {
  if (!IF || !IFh) return fail(QS_ERR_INVALID_ARG, "qlin: null table");
  QS_TRY(check_mesh(mesh));
  const double* k2 = mesh->k_squared;
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double alpha1 = (-1.0 * DQ * k2[i] + aQ);
    IF[i] = exp(alpha1 * dt);
    IFh[i] = exp(alpha1 * dt * 0.5);
  });
}
int single1_c(double* s1, const double* r, const double* S, double lambda, double cQ, double aQ,
              const Mesh* mesh) {
  if (!s1 || !r || !S) return fail(QS_ERR_INVALID_ARG, "single1: null pointer");
  QS_TRY(check_mesh(mesh));
  return each(mesh, np(mesh), [=] __device__(long i) {
    s1[i] = s1[i] + lambda * S[i] * (cQ * S[i] * S[i] - aQ) * r[i];
  });
}
int S2r_c(double* rs, const double* r, const double* S, double cQ, const Mesh* mesh) {
  if (!rs || !r || !S) return fail(QS_ERR_INVALID_ARG, "S2r: null pointer");
  QS_TRY(check_mesh(mesh));
  return each(mesh, np(mesh), [=] __device__(long i) { rs[i] = rs[i] - cQ * S[i] * S[i] * r[i]; });
}

}  // namespace qs

using namespace qs;

#define NN(p) \
  if (!(p)) return fail(QS_ERR_INVALID_ARG, "%s: null pointer argument " #p, __func__)

extern "C" {

// FldOp/QFldfuncs.cuh:45-48
int qs_FwdTrans(Qcomp* ft, const Qreal* f, const Mesh* mesh) {
  QS_TRY(check_mesh(mesh));
  NN(ft); NN(f);
  if ((const void*)ft == (const void*)mesh->spec)
    return fail(QS_ERR_INVALID_ARG, "qs_FwdTrans: ft must not be the mesh scratch");
  return fft_r2c_2d(mesh, f, (double2*)mesh->spec, (double2*)ft);
}

// FldOp/QFldfuncs.cuh:50-55
int qs_BwdTrans(Qreal* f, Qcomp* ft, const Mesh* mesh) {
  QS_TRY(check_mesh(mesh));
  NN(ft); NN(f);
  if ((const void*)ft == (const void*)mesh->spec)
    return fail(QS_ERR_INVALID_ARG, "qs_BwdTrans: ft must not be the mesh scratch");
  return fft_c2r_2d(mesh, (double2*)ft, (double2*)mesh->spec, f, true);
}

// FldOp/QFldfuncs.cuh:8-16
int qs_coeff(Qreal* f, const Mesh* mesh) {
  NN(f); QS_TRY(check_mesh(mesh));
  const int NxNy = mesh->Nx * mesh->Ny;
  return each(mesh, np(mesh), [=] __device__(long i) { f[i] = f[i] / NxNy; });
}

// FldOp/QFldfuncs.cuh:18-44
int qs_symmetry_func(Qcomp* fs, const Mesh* mesh) {
  NN(fs); QS_TRY(check_mesh(mesh));
  const int Nxh = mesh->Nxh, Ny = mesh->Ny;
  return each(mesh, Ny / 2, [=] __device__(long j) {
    if (j > 0 && j < Ny / 2) {
      const long index = j * Nxh, index2 = (Ny - j) * (long)Nxh;
      const double2 a = d2(fs[index]), b = d2(fs[index2]);
      const double2 mv = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
      fs[index] = qc(mv);
      fs[index2] = qc(cconj(mv));
    }
  });
}

int qs_FldMul(Qreal* pc, const Qreal* pa, const Qreal* pb, Qreal C, const Mesh* mesh) {
  NN(pc); NN(pa); NN(pb); QS_TRY(check_mesh(mesh));
  return each(mesh, np(mesh), [=] __device__(long i) { pc[i] = C * pa[i] * pb[i]; });
}
int qs_FldAdd(Qreal* pc, Qreal a, const Qreal* pa, Qreal b, const Qreal* pb, const Mesh* mesh) {
  NN(pc); NN(pa); NN(pb); QS_TRY(check_mesh(mesh));
  return each(mesh, np(mesh), [=] __device__(long i) { pc[i] = a * pa[i] + b * pb[i]; });
}
int qs_FldAddScalar(Qreal* pc, Qreal a, const Qreal* pa, Qreal b, const Mesh* mesh) {
  NN(pc); NN(pa); QS_TRY(check_mesh(mesh));
  return each(mesh, np(mesh), [=] __device__(long i) { pc[i] = a * pa[i] + b; });
}
int qs_FldSet(Qreal* pf, Qreal c, const Mesh* mesh) {
  NN(pf); QS_TRY(check_mesh(mesh));
  return each(mesh, np(mesh), [=] __device__(long i) { pf[i] = c; });
}
int qs_FldCopy(Qreal* pf, const Qreal* c, const Mesh* mesh) {
  NN(pf); NN(c); QS_TRY(check_mesh(mesh));
  return each(mesh, np(mesh), [=] __device__(long i) { pf[i] = c[i]; });
}
int qs_SpecSet(Qcomp* pa, const Qcomp* pb, const Mesh* mesh) {
  NN(pa); NN(pb); QS_TRY(check_mesh(mesh));
  return each(mesh, ns(mesh), [=] __device__(long i) { pa[i] = pb[i]; });
}
int qs_SpecSetConst(Qcomp* pa, Qcomp c, const Mesh* mesh) {
  NN(pa); QS_TRY(check_mesh(mesh));
  return each(mesh, ns(mesh), [=] __device__(long i) { pa[i] = c; });
}
int qs_SpecAdd(Qcomp* spc, Qreal a, const Qcomp* spa, Qreal b, const Qcomp* spb, const Mesh* mesh) {
  NN(spc); NN(spa); NN(spb); QS_TRY(check_mesh(mesh));
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double2 x = d2(spa[i]), y = d2(spb[i]);
    spc[i] = qc(make_double2(a * x.x + b * y.x, a * x.y + b * y.y));
  });
}
int qs_SpecMul(Qcomp* spc, const Qcomp* spa, Qreal C, const Mesh* mesh) {
  NN(spc); NN(spa); QS_TRY(check_mesh(mesh));
  return each(mesh, ns(mesh), [=] __device__(long i) { spc[i] = qc(cscale(C, d2(spa[i]))); });
}
int qs_xDeriv(Qcomp* fx, const Qcomp* f, const Mesh* mesh) {
  NN(fx); NN(f); QS_TRY(check_mesh(mesh));
  const double* kx = mesh->kx;
  return each(mesh, ns(mesh), [=] __device__(long i) { fx[i] = qc(im_times(kx[i], d2(f[i]))); });
}
int qs_yDeriv(Qcomp* fy, const Qcomp* f, const Mesh* mesh) {
  NN(fy); NN(f); QS_TRY(check_mesh(mesh));
  const double* ky = mesh->ky;
  return each(mesh, ns(mesh), [=] __device__(long i) { fy[i] = qc(im_times(ky[i], d2(f[i]))); });
}
int qs_Laplacian(Qcomp* fL, const Qcomp* f, const Mesh* mesh) {
  NN(fL); NN(f); QS_TRY(check_mesh(mesh));
  const double* k2 = mesh->k_squared;
  return each(mesh, ns(mesh), [=] __device__(long i) { fL[i] = qc(cscale(-1.0 * k2[i], d2(f[i]))); });
}
int qs_vel_func(const Qcomp* w_c, Qcomp* u_c, Qcomp* v_c, const Mesh* mesh) {
  NN(w_c); NN(u_c); NN(v_c); QS_TRY(check_mesh(mesh));
  const double *kx = mesh->kx, *ky = mesh->ky, *k2 = mesh->k_squared;
  return each(mesh, ns(mesh), [=] __device__(long i) {
    if (i == 0) {
      u_c[i] = qc(make_double2(0.0, 0.0));
      v_c[i] = qc(make_double2(0.0, 0.0));
    } else {
      const double2 w = d2(w_c[i]);
      const double2 u = im_times(ky[i], w);
      const double2 v = im_times(-1.0 * kx[i], w);
      u_c[i] = qc(make_double2(u.x / k2[i], u.y / k2[i]));
      v_c[i] = qc(make_double2(v.x / k2[i], v.y / k2[i]));
    }
  });
}

// ---- InertialFriction/friction/Qsolver.cuh --------------------------------
int qs_wlin_func(Qreal* IFw, Qreal* IFwh, Qreal Re, Qreal Rf, Qreal dt, const Mesh* mesh) {
  NN(IFw); NN(IFwh); QS_TRY(check_mesh(mesh));
  const double* k2 = mesh->k_squared;
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double alpha = 1.0 / Re * (-1.0 * k2[i] - Rf);
    IFw[i] = exp(alpha * dt);
    IFwh[i] = exp(alpha * dt * 0.5);
  });
}
int qs_r1lin_func(Qreal* IFr1, Qreal* IFr1h, Qreal dt, const Mesh* mesh) {
  return qs::qlin_c(IFr1, IFr1h, dt, 1.0, 1.0, mesh);
}
int qs_r2lin_func(Qreal* IFr2, Qreal* IFr2h, Qreal dt, const Mesh* mesh) {
  return qs::qlin_c(IFr2, IFr2h, dt, 1.0, 1.0, mesh);
}

int qs_S_func(Qreal* S, const Qreal* r1, const Qreal* r2, const Mesh* mesh) {
  NN(S); NN(r1); NN(r2); QS_TRY(check_mesh(mesh));
  return each(mesh, np(mesh), [=] __device__(long i) { S[i] = 2 * sqrt(r1[i] * r1[i] + r2[i] * r2[i]); });
}
int qs_Rar(Qreal* rar, const Qreal* r, const Qreal* Ra, const Mesh* mesh) {
  NN(rar); NN(r); NN(Ra); QS_TRY(check_mesh(mesh));
  return each(mesh, np(mesh), [=] __device__(long i) { rar[i] = rar[i] - Ra[i] * r[i]; });
}
int qs_single1(Qreal* s1, const Qreal* r, const Qreal* S, Qreal lambda, const Mesh* mesh) {
  return qs::single1_c(s1, r, S, lambda, 1.0, 1.0, mesh);
}
int qs_S2r(Qreal* rs, const Qreal* r, const Qreal* S, const Mesh* mesh) {
  return qs::S2r_c(rs, r, S, 1.0, mesh);
}
int qs_PiDeriv(Qcomp* dP, const Qcomp* p12, const Qcomp* p11, const Qcomp* p21, Qreal Er, Qreal Re,
               const Mesh* mesh) {
  NN(dP); NN(p12); NN(p11); NN(p21); QS_TRY(check_mesh(mesh));
  const double *kx = mesh->kx, *ky = mesh->ky;
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double c12 = -1.0 * kx[i] * kx[i], c11 = 2.0 * kx[i] * ky[i], c21 = ky[i] * ky[i];
    const double2 a = d2(p12[i]), b = d2(p11[i]), c = d2(p21[i]);
    const double2 t = make_double2(c12 * a.x + c11 * b.x + c21 * c.x, c12 * a.y + c11 * b.y + c21 * c.y);
    const double2 d = d2(dP[i]);
    const double s = 1.0 / (Er * Re);
    dP[i] = qc(make_double2(d.x + s * t.x, d.y + s * t.y));
  });
}
int qs_dealias(Qcomp* ft, const Mesh* mesh) {
  NN(ft); QS_TRY(check_mesh(mesh));
  const double* cut = mesh->cutoff;
  return each(mesh, ns(mesh), [=] __device__(long i) { ft[i] = qc(cscale(cut[i], d2(ft[i]))); });
}

// ---- TimeIntegration ---------------------------------------------------------
int qs_ife_func1D(const Qcomp* o, Qcomp* nw, const Qcomp* nl, const Qreal* IF, const Qreal* IFh, Qreal dt,
                  const Mesh* mesh) {
  NN(o); NN(nw); NN(nl); NN(IF); QS_TRY(check_mesh(mesh));
  (void)IFh;
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double2 an = cscale(dt, d2(nl[i]));
    nw[i] = qc(cscale(IF[i], cadd(d2(o[i]), an)));
  });
}
static int func0D(const Qcomp* o, Qcomp* cu, Qcomp* nw, const Qreal* IF, const Mesh* mesh) {
  NN(o); NN(cu); NN(nw); NN(IF); QS_TRY(check_mesh(mesh));
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double2 old = d2(o[i]);
    nw[i] = qc(cscale(IF[i], old));
    cu[i] = qc(old);
  });
}
int qs_ifrk2_func0D(const Qcomp* o, Qcomp* cu, Qcomp* nw, const Qreal* IF, const Qreal* IFh, Qreal dt,
                    const Mesh* mesh) {
  (void)IFh; (void)dt;
  return func0D(o, cu, nw, IF, mesh);
}
int qs_ifrk2_func1D(const Qcomp* o, Qcomp* cu, Qcomp* nw, const Qcomp* nl, const Qreal* IF, const Qreal* IFh,
                    Qreal dt, const Mesh* mesh) {
  NN(o); NN(cu); NN(nw); NN(nl); NN(IF); NN(IFh); QS_TRY(check_mesh(mesh));
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double2 an = cscale(dt, d2(nl[i]));
    nw[i] = qc(cadd(d2(nw[i]), cscale(1.0 / 2.0 * IF[i], an)));
    cu[i] = qc(cscale(IFh[i], cadd(d2(o[i]), an)));
  });
}
int qs_ifrk2_func2D(const Qcomp* o, Qcomp* cu, Qcomp* nw, const Qcomp* nl, const Qreal* IF, const Qreal* IFh,
                    Qreal dt, const Mesh* mesh) {
  NN(nw); NN(nl); QS_TRY(check_mesh(mesh));
  (void)o; (void)cu; (void)IF; (void)IFh;
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double2 bn = cscale(dt, d2(nl[i]));
    nw[i] = qc(cadd(d2(nw[i]), cscale(1.0 / 2.0, bn)));
  });
}
int qs_rk4_func0D(const Qcomp* o, Qcomp* cu, Qcomp* nw, const Qreal* IF, const Qreal* IFh, Qreal dt,
                  const Mesh* mesh) {
  (void)IFh; (void)dt;
  return func0D(o, cu, nw, IF, mesh);
}
int qs_rk4_func1D(const Qcomp* o, Qcomp* cu, Qcomp* nw, const Qcomp* nl, const Qreal* IF, const Qreal* IFh,
                  Qreal dt, const Mesh* mesh) {
  NN(o); NN(cu); NN(nw); NN(nl); NN(IF); NN(IFh); QS_TRY(check_mesh(mesh));
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double2 an = cscale(dt, d2(nl[i]));
    nw[i] = qc(cadd(d2(nw[i]), cscale(1.0 / 6.0 * IF[i], an)));
    const double2 half = make_double2(an.x / 2.0, an.y / 2.0);
    cu[i] = qc(cscale(IFh[i], cadd(d2(o[i]), half)));
  });
}
int qs_rk4_func2D(const Qcomp* o, Qcomp* cu, Qcomp* nw, const Qcomp* nl, const Qreal* IF, const Qreal* IFh,
                  Qreal dt, const Mesh* mesh) {
  NN(o); NN(cu); NN(nw); NN(nl); NN(IFh); QS_TRY(check_mesh(mesh));
  (void)IF;
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double2 bn = cscale(dt, d2(nl[i]));
    nw[i] = qc(cadd(d2(nw[i]), cscale(1.0 / 3.0 * IFh[i], bn)));
    cu[i] = qc(cadd(cscale(IFh[i], d2(o[i])), make_double2(bn.x / 2.0, bn.y / 2.0)));
  });
}
int qs_rk4_func3D(const Qcomp* o, Qcomp* cu, Qcomp* nw, const Qcomp* nl, const Qreal* IF, const Qreal* IFh,
                  Qreal dt, const Mesh* mesh) {
  NN(o); NN(cu); NN(nw); NN(nl); NN(IF); NN(IFh); QS_TRY(check_mesh(mesh));
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double2 cn = cscale(dt, d2(nl[i]));
    nw[i] = qc(cadd(d2(nw[i]), cscale(1.0 / 3.0 * IFh[i], cn)));
    cu[i] = qc(cadd(cscale(IF[i], d2(o[i])), cscale(IFh[i], cn)));
  });
}
int qs_rk4_func4D(const Qcomp* o, Qcomp* cu, Qcomp* nw, const Qcomp* nl, const Qreal* IF, const Qreal* IFh,
                  Qreal dt, const Mesh* mesh) {
  NN(nw); NN(nl); QS_TRY(check_mesh(mesh));
  (void)o; (void)cu; (void)IF; (void)IFh;
  return each(mesh, ns(mesh), [=] __device__(long i) {
    const double2 dn = cscale(dt, d2(nl[i]));
    nw[i] = qc(cadd(d2(nw[i]), cscale(1.0 / 6.0, dn)));
  });
}

}  // extern "C"
