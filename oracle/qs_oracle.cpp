// qs_oracle.cpp -- FP64 CPU oracle of the Qsolver step (TEST INFRASTRUCTURE).
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
// this file, as the checker / CPU baseline.  The HIP product never calls it.
//
// Every function names the reference file:line it restates.  The reference is
// CUDA; here each __global__ kernel becomes a loop over the same index space
// and each host wrapper keeps its call sequence, so the arithmetic of the
// reference is followed operation by operation (apart from the FFT, which is
// this file's own mixed-radix Stockham in place of cuFFT).
#include "qs_oracle.h"

#include <cmath>
#include <cstring>
#include <memory>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = or_comp;

inline cd mk(double x, double y) { cd c; c.x = x; c.y = y; return c; }
// Basic/cuComplexBinOp.cuh: naive complex arithmetic.
inline cd operator+(cd a, cd b) { return mk(a.x + b.x, a.y + b.y); }
inline cd operator-(cd a, cd b) { return mk(a.x - b.x, a.y - b.y); }
inline cd operator*(cd a, cd b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
inline cd operator*(double a, cd b) { return mk(a * b.x, a * b.y); }
inline cd operator*(cd a, double b) { return mk(a.x * b, a.y * b); }
inline cd operator/(cd a, double b) { return mk(a.x / b, a.y / b); }
inline cd conj(cd a) { return mk(a.x, -a.y); }
inline cd im() { return mk(0.0, 1.0); }   // Basic/cuComplexBinOp.cuh:14-18

int g_threads = 0;
inline int nthreads() {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}
// Parallel regions only pay off above this many elements.
constexpr long kParMin = 1 << 15;

// ---------------------------------------------------------------------------
// 1-D complex FFT: iterative Stockham autosort, mixed radix (4, 2, 3, 5, and a
// direct DFT for any other prime factor).  Forward uses e^{-2 pi i nk/N}.
// Twiddles come from a table W[m] = exp(-2 pi i m / N) evaluated in long
// double (x87, 64-bit mantissa), so each entry is within ~0.5 ulp in double.
// ---------------------------------------------------------------------------
struct Fft1 {
  int n = 0;
  std::vector<int> radix;
  std::vector<cd> tw;   // tw[m] = exp(-2 pi i m / n)

  explicit Fft1(int n_ = 1) : n(n_) {
    int m = n;
    for (int r : {4, 2, 3, 5}) while (m % r == 0) { radix.push_back(r); m /= r; }
    for (int p = 7; m > 1; p += 2) while (m % p == 0) { radix.push_back(p); m /= p; }
    tw.resize(n);
    for (int k = 0; k < n; ++k) {
      long double two_pi = 6.283185307179586476925286766559005768L;
      long double c = cosl(two_pi * (long double)k / (long double)n);
      long double s = sinl(two_pi * (long double)k / (long double)n);
      tw[k] = mk((double)c, -(double)s);
    }
  }

  // x: data (length n), y: scratch (length n).  Result written to x.
  void exec(cd* x, cd* y, int sign) const {
    if (n == 1) return;
    cd* a = x;
    cd* b = y;
    int p = 1;
    cd v[64];
    for (int R : radix) {
      const int T = n / R;
      for (int i = 0; i < T; ++i) {
        const int k = i % p;
        // twiddle exp(sign * 2 pi i r k / (p R)) = tw[r k (n/(pR))] (conj if sign>0)
        const int step = n / (p * R);
        for (int r = 0; r < R; ++r) {
          cd t = a[i + r * T];
          if (r && k) {
            cd w = tw[(r * k * step) % n];
            if (sign > 0) w = conj(w);
            t = t * w;
          }
          v[r] = t;
        }
        dft_small(v, R, sign);
        const int j = (i - k) * R + k;
        for (int r = 0; r < R; ++r) b[j + r * p] = v[r];
      }
      p *= R;
      cd* t = a; a = b; b = t;
    }
    if (a != x) std::memcpy(x, a, sizeof(cd) * n);
  }

  void dft_small(cd* v, int R, int sign) const {
    if (R == 2) {
      cd a = v[0], b = v[1];
      v[0] = a + b; v[1] = a - b;
    } else if (R == 4) {
      cd a0 = v[0], a1 = v[1], a2 = v[2], a3 = v[3];
      cd s02 = a0 + a2, d02 = a0 - a2, s13 = a1 + a3, d13 = a1 - a3;
      // multiply d13 by -i (forward) or +i (inverse)
      cd jd13 = sign < 0 ? mk(d13.y, -d13.x) : mk(-d13.y, d13.x);
      v[0] = s02 + s13; v[2] = s02 - s13;
      v[1] = d02 + jd13; v[3] = d02 - jd13;
    } else {
      cd out[64];
      const int step = n / R;
      for (int s = 0; s < R; ++s) {
        cd acc = v[0];
        for (int r = 1; r < R; ++r) {
          cd w = tw[((r * s) % R) * step];
          if (sign > 0) w = conj(w);
          acc = acc + v[r] * w;
        }
        out[s] = acc;
      }
      for (int s = 0; s < R; ++s) v[s] = out[s];
    }
  }
};

// ---------------------------------------------------------------------------
// Mesh and Field (Basic/Mesh.cuh:48-129)
// ---------------------------------------------------------------------------
struct OMesh {
  int Nx, Nxh, Ny;
  double Lx, Ly, dx, dy;
  std::vector<double> kx, ky, k_squared, cutoff;
  std::vector<cd> spec;      // scratch, as Mesh::spec
  std::vector<double> phys;  // scratch, as Mesh::phys
  Fft1 fx, fy;
  OMesh(int nx, int ny, double lx, double ly)
      : Nx(nx), Nxh(nx / 2 + 1), Ny(ny), Lx(lx), Ly(ly), dx(lx / nx), dy(ly / ny),
        fx(nx), fy(ny) {
    const size_t ws = (size_t)Nxh * Ny;
    kx.resize(ws); ky.resize(ws); k_squared.resize(ws); cutoff.resize(ws);
    spec.resize(ws); phys.resize((size_t)Nx * Ny);
    waveinit();
    cutoff_func();
  }
  size_t ns() const { return (size_t)Nxh * Ny; }
  size_t np() const { return (size_t)Nx * Ny; }

  // Basic/Mesh.cuh:6-21 -- Nyquist row (j == Ny/2) gets the positive wavenumber.
  // Host code in the reference: k^2 from two rounded products (no fused
  // multiply-add), as an x86-64 host compiler evaluates it.
  __attribute__((optimize("fp-contract=off"))) void waveinit() {
    for (int i = 0; i < Nxh; i++)
      for (int j = 0; j < Ny; j++) {
        int index = i + j * Nxh;
        if (j < Ny / 2 + 1) ky[index] = 2 * M_PI / Ly * j;
        else ky[index] = 2 * M_PI / Ly * (j - Ny);
        kx[index] = 2 * M_PI / Lx * i;
        k_squared[index] = kx[index] * kx[index] + ky[index] * ky[index];
      }
  }
  // Basic/Mesh.cuh:23-46 (effective rule :36-42).
  void cutoff_func() {
    for (int i = 0; i < Nxh; i++)
      for (int j = 0; j < Ny; j++) {
        int index = i + j * Nxh;
        cutoff[index] = ((i < ((Nxh - 1) * 2 / 3 + 1)) && ((j < Ny / 3 + 1) || ((Ny - j) < Ny / 3))) ? 1.0 : 0.0;
      }
  }
};

struct OField {
  OMesh* mesh;
  std::vector<cd> spec;
  std::vector<double> phys;
  explicit OField(OMesh* m) : mesh(m), spec(m->ns(), mk(0, 0)), phys(m->np(), 0.0) {}
};

// ---------------------------------------------------------------------------
// Transforms (FldOp/QFldfuncs.cuh:8-55)
// ---------------------------------------------------------------------------

// Unnormalised 2-D R2C: x-direction real FFT of every row, then y-direction
// complex FFT of every column i < Nxh.  Pairs of rows share one complex FFT
// (z = a + i b;  A = (Z_k + conj Z_{N-k})/2,  B = (Z_k - conj Z_{N-k})/(2i)).
void r2c2d(OMesh* m, const double* f, cd* ft) {
  const int Nx = m->Nx, Ny = m->Ny, Nxh = m->Nxh;
  const int npairs = (Ny + 1) / 2;
#pragma omp parallel num_threads(nthreads()) if ((long)Nx * Ny > kParMin)
  {
    std::vector<cd> z(Nx), s(std::max(Nx, Ny)), col(Ny);
#pragma omp for schedule(static)
    for (int pr = 0; pr < npairs; ++pr) {
      const int y0 = 2 * pr, y1 = y0 + 1;
      for (int x = 0; x < Nx; ++x)
        z[x] = mk(f[(size_t)y0 * Nx + x], y1 < Ny ? f[(size_t)y1 * Nx + x] : 0.0);
      m->fx.exec(z.data(), s.data(), -1);
      for (int k = 0; k < Nxh; ++k) {
        cd zk = z[k], zc = conj(z[(Nx - k) % Nx]);
        cd A = mk(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
        cd d = zk - zc;                       // 2i * B
        cd B = mk(0.5 * d.y, -0.5 * d.x);
        ft[(size_t)y0 * Nxh + k] = A;
        if (y1 < Ny) ft[(size_t)y1 * Nxh + k] = B;
      }
    }
#pragma omp for schedule(static)
    for (int i = 0; i < Nxh; ++i) {
      for (int j = 0; j < Ny; ++j) col[j] = ft[(size_t)j * Nxh + i];
      m->fy.exec(col.data(), s.data(), -1);
      for (int j = 0; j < Ny; ++j) ft[(size_t)j * Nxh + i] = col[j];
    }
  }
}

// Unnormalised 2-D C2R with the fixed semantics: y-direction inverse complex
// FFT of every column, then x-direction C2R of every row that discards the
// imaginary parts of the DC and Nyquist bins.  Input is destroyed.
void c2r2d(OMesh* m, cd* ft, double* f) {
  const int Nx = m->Nx, Ny = m->Ny, Nxh = m->Nxh;
  const int npairs = (Ny + 1) / 2;
#pragma omp parallel num_threads(nthreads()) if ((long)Nx * Ny > kParMin)
  {
    std::vector<cd> z(Nx), s(std::max(Nx, Ny)), col(Ny);
#pragma omp for schedule(static)
    for (int i = 0; i < Nxh; ++i) {
      for (int j = 0; j < Ny; ++j) col[j] = ft[(size_t)j * Nxh + i];
      m->fy.exec(col.data(), s.data(), +1);
      for (int j = 0; j < Ny; ++j) ft[(size_t)j * Nxh + i] = col[j];
    }
#pragma omp for schedule(static)
    for (int pr = 0; pr < npairs; ++pr) {
      const int y0 = 2 * pr, y1 = y0 + 1;
      const cd* a = ft + (size_t)y0 * Nxh;
      const cd* b = y1 < Ny ? ft + (size_t)y1 * Nxh : nullptr;
      auto A = [&](int k) {
        cd v = a[k];
        if (k == 0 || 2 * k == Nx) v.y = 0.0;
        return v;
      };
      auto B = [&](int k) {
        if (!b) return mk(0.0, 0.0);
        cd v = b[k];
        if (k == 0 || 2 * k == Nx) v.y = 0.0;
        return v;
      };
      // Z_k = Ah_k + i Bh_k with Ah, Bh the Hermitian extensions.
      for (int k = 0; k < Nx; ++k) {
        cd ah, bh;
        if (k < Nxh) { ah = A(k); bh = B(k); }
        else { ah = conj(A(Nx - k)); bh = conj(B(Nx - k)); }
        z[k] = mk(ah.x - bh.y, ah.y + bh.x);
      }
      m->fx.exec(z.data(), s.data(), +1);
      for (int x = 0; x < Nx; ++x) {
        f[(size_t)y0 * Nx + x] = z[x].x;
        if (y1 < Ny) f[(size_t)y1 * Nx + x] = z[x].y;
      }
    }
  }
}

// FldOp/QFldfuncs.cuh:8-16
void coeff(double* f, int Nx, int Ny) {
  const long n = (long)Nx * Ny;
#pragma omp parallel for num_threads(nthreads()) if (n > kParMin)
  for (long idx = 0; idx < n; ++idx) f[idx] = f[idx] / (Nx * Ny);
}

// FldOp/QFldfuncs.cuh:18-44
void symmetry_func(cd* f_spec, int Nxh, int Ny) {
  for (int j = 1; j < Ny / 2; ++j) {
    int index = j * Nxh, index2 = (Ny - j) * Nxh;
    cd mean_value = 0.5 * (f_spec[index] + conj(f_spec[index2]));
    f_spec[index] = mean_value;
    f_spec[index2] = conj(mean_value);
  }
}

// FldOp/QFldfuncs.cuh:45-48 (input f preserved: copy to mesh->phys first)
void FwdTrans(cd* ft, const double* f, OMesh* mesh) {
  std::memcpy(mesh->phys.data(), f, sizeof(double) * mesh->np());
  r2c2d(mesh, mesh->phys.data(), ft);
}

// FldOp/QFldfuncs.cuh:50-55: copy ft -> scratch; symmetrise the caller's ft
// in place (side effect); C2R the unsymmetrised copy; divide by Nx*Ny.
void BwdTrans(double* f, cd* ft, OMesh* mesh) {
  std::memcpy(mesh->spec.data(), ft, sizeof(cd) * mesh->ns());
  symmetry_func(ft, mesh->Nxh, mesh->Ny);
  c2r2d(mesh, mesh->spec.data(), f);
  coeff(f, mesh->Nx, mesh->Ny);
}

// ---------------------------------------------------------------------------
// Elementwise field operations (FldOp/QFldfuncs.cuh:63-224)
// ---------------------------------------------------------------------------
#define OR_LOOP(N)                                                           \
  const long or_n_ = (long)(N);                                              \
  _Pragma("omp parallel for num_threads(nthreads()) if (or_n_ > kParMin)") \
  for (long index = 0; index < or_n_; ++index)

void FldMul(double* pc, const double* pa, const double* pb, double C, long n) {  // :63-71
  OR_LOOP(n) pc[index] = C * pa[index] * pb[index];
}
void FldAdd(double* pc, double a, const double* pa, double b, const double* pb, long n) {  // :74-82
  OR_LOOP(n) pc[index] = a * pa[index] + b * pb[index];
}
void FldSet(double* pf, double c, long n) {  // :96-104
  OR_LOOP(n) pf[index] = c;
}
void SpecAdd(cd* spc, double a, const cd* spa, double b, const cd* spb, long n) {  // :141-149
  OR_LOOP(n) spc[index] = a * spa[index] + b * spb[index];
}
void xDeriv(cd* fx, const cd* f, const OMesh* m) {  // :165-176
  const double* kx = m->kx.data();
  OR_LOOP(m->ns()) fx[index] = im() * kx[index] * f[index];
}
void yDeriv(cd* fy, const cd* f, const OMesh* m) {  // :178-189
  const double* ky = m->ky.data();
  OR_LOOP(m->ns()) fy[index] = im() * ky[index] * f[index];
}
void Laplacian(cd* fL, const cd* f, const OMesh* m) {  // :191-202
  const double* k2 = m->k_squared.data();
  OR_LOOP(m->ns()) fL[index] = -1.0 * k2[index] * f[index];
}
void vel_func(const cd* w_c, cd* u_c, cd* v_c, const OMesh* m) {  // :204-224
  const double *kx = m->kx.data(), *ky = m->ky.data(), *k2 = m->k_squared.data();
  OR_LOOP(m->ns()) {
    if (index == 0) {
      u_c[index] = mk(0.0, 0.0);
      v_c[index] = mk(0.0, 0.0);
    } else {
      u_c[index] = ky[index] * im() * w_c[index] / (k2[index]);
      v_c[index] = -1.0 * kx[index] * im() * w_c[index] / (k2[index]);
    }
  }
}

// ---------------------------------------------------------------------------
// Physics (InertialFriction/friction/Qsolver.cuh).  Coef carries the
// generalised Q-equation coefficients; all 1 reproduces the reference.
// ---------------------------------------------------------------------------
struct Coef { double DQ = 1.0, aQ = 1.0, cQ = 1.0; };

void S_func(double* S, const double* r1, const double* r2, long n) {  // :63-71
  OR_LOOP(n) S[index] = 2 * std::sqrt(r1[index] * r1[index] + r2[index] * r2[index]);
}
void Rar(double* rar, const double* r, const double* Ra, long n) {  // :75-83
  OR_LOOP(n) rar[index] = rar[index] - Ra[index] * r[index];
}
void single1(double* s1, const double* r, const double* S, double lambda, const Coef& c, long n) {  // :86-94
  OR_LOOP(n) s1[index] = s1[index] + lambda * S[index] * (c.cQ * S[index] * S[index] - c.aQ) * r[index];
}
void S2r(double* rs, const double* r, const double* S, const Coef& c, long n) {  // :206-215
  OR_LOOP(n) rs[index] = rs[index] - c.cQ * S[index] * S[index] * r[index];
}

// :96-105
void single(OField* s0, OField* r, OField* S, double lambda, const Coef& c) {
  OMesh* mesh = s0->mesh;
  const long n = mesh->np();
  Laplacian(s0->spec.data(), r->spec.data(), mesh);
  BwdTrans(s0->phys.data(), s0->spec.data(), mesh);
  FldMul(s0->phys.data(), s0->phys.data(), S->phys.data(), -1.0 * lambda * c.DQ, n);
  single1(s0->phys.data(), r->phys.data(), S->phys.data(), lambda, c, n);
}

// :107-122
void cross(OField* cs, OField* r1, OField* r2, OField* aux, const Coef& c) {
  OMesh* mesh = cs->mesh;
  const long n = mesh->np();
  Laplacian(cs->spec.data(), r1->spec.data(), mesh);
  BwdTrans(cs->phys.data(), cs->spec.data(), mesh);
  FldMul(cs->phys.data(), cs->phys.data(), r2->phys.data(), 2.0 * c.DQ, n);
  Laplacian(aux->spec.data(), r2->spec.data(), mesh);
  BwdTrans(aux->phys.data(), aux->spec.data(), mesh);
  FldMul(aux->phys.data(), r1->phys.data(), aux->phys.data(), -2.0 * c.DQ, n);
  FldAdd(cs->phys.data(), 1.0, cs->phys.data(), 1.0, aux->phys.data(), n);
}

// :124-131
void p11func(OField* p11, OField* r1, OField* S, OField* Ra, double lambda, const Coef& c) {
  single(p11, r1, S, lambda, c);
  Rar(p11->phys.data(), r1->phys.data(), Ra->phys.data(), p11->mesh->np());
  FwdTrans(p11->spec.data(), p11->phys.data(), p11->mesh);
}

// :133-141
void p12func(OField* p12, OField* r1, OField* r2, OField* S, OField* Ra, double lambda,
             OField* aux, OField* aux1, const Coef& c) {
  const long n = p12->mesh->np();
  single(p12, r2, S, lambda, c);
  cross(aux, r1, r2, aux1, c);
  FldAdd(p12->phys.data(), 1.0, p12->phys.data(), 1.0, aux->phys.data(), n);
  Rar(p12->phys.data(), r2->phys.data(), Ra->phys.data(), n);
  FwdTrans(p12->spec.data(), p12->phys.data(), p12->mesh);
}

// :143-151
void p21func(OField* p21, OField* r1, OField* r2, OField* S, OField* Ra, double lambda,
             OField* aux, OField* aux1, const Coef& c) {
  const long n = p21->mesh->np();
  single(p21, r2, S, lambda, c);
  cross(aux, r2, r1, aux1, c);
  FldAdd(p21->phys.data(), 1.0, p21->phys.data(), 1.0, aux->phys.data(), n);
  Rar(p21->phys.data(), r2->phys.data(), Ra->phys.data(), n);
  FwdTrans(p21->spec.data(), p21->phys.data(), p21->mesh);
}

// :154-164
void PiDeriv(cd* dP, const cd* p12, const cd* p11, const cd* p21, const OMesh* m, double Er, double Re) {
  const double *kx = m->kx.data(), *ky = m->ky.data();
  OR_LOOP(m->ns()) {
    dP[index] = dP[index] + 1.0 / (Er * Re) *
        (-1.0 * kx[index] * kx[index] * p12[index] + 2.0 * kx[index] * ky[index] * p11[index] +
         ky[index] * ky[index] * p21[index]);
  }
}

// :166-201 (advect_w == 0 is InertialFriction/beginning/Qsolver.cuh:179-192)
void wnonl_func(OField* wnonl, OField* wcurr, OField* u, OField* v, OField* r1curr, OField* r2curr,
                OField* Ra, OField* S, OField* p11, OField* p12, OField* p21, double Re, double Er,
                double lambda, OField* aux, OField* aux1, const Coef& c, bool advect_w) {
  OMesh* mesh = wnonl->mesh;
  const long n = mesh->np();
  if (advect_w) {
    FldMul(wnonl->phys.data(), u->phys.data(), wcurr->phys.data(), -1.0, n);
    FwdTrans(wnonl->spec.data(), wnonl->phys.data(), mesh);
    xDeriv(wnonl->spec.data(), wnonl->spec.data(), mesh);
    BwdTrans(wnonl->phys.data(), wnonl->spec.data(), mesh);

    FldMul(aux->phys.data(), v->phys.data(), wcurr->phys.data(), -1.0, n);
    FwdTrans(aux->spec.data(), aux->phys.data(), mesh);
    yDeriv(aux->spec.data(), aux->spec.data(), mesh);
    BwdTrans(aux->phys.data(), aux->spec.data(), mesh);

    FldAdd(wnonl->phys.data(), 1.0, wnonl->phys.data(), 1.0, aux->phys.data(), n);
  } else {
    FldSet(wnonl->phys.data(), 0.0, n);
  }
  FwdTrans(wnonl->spec.data(), wnonl->phys.data(), mesh);

  p11func(p11, r1curr, S, Ra, lambda, c);
  p12func(p12, r1curr, r2curr, S, Ra, lambda, aux, aux1, c);
  p21func(p21, r1curr, r2curr, S, Ra, lambda, aux, aux1, c);
  PiDeriv(wnonl->spec.data(), p12->spec.data(), p11->spec.data(), p21->spec.data(), mesh, Er, Re);
}

// :217-255
void r1nonl_func(OField* r1nonl, OField* wcurr, OField* u, OField* v, OField* r1curr, OField* r2curr,
                 OField* S, double lambda, OField* aux, OField* aux1, const Coef& c) {
  OMesh* mesh = r1nonl->mesh;
  const long n = mesh->np(), ns = mesh->ns();
  FldMul(r1nonl->phys.data(), wcurr->phys.data(), r2curr->phys.data(), -1.0, n);
  S2r(r1nonl->phys.data(), r1curr->phys.data(), S->phys.data(), c, n);
  FwdTrans(r1nonl->spec.data(), r1nonl->phys.data(), mesh);

  FldMul(aux->phys.data(), u->phys.data(), r1curr->phys.data(), -1.0, n);
  FwdTrans(aux->spec.data(), aux->phys.data(), mesh);
  xDeriv(aux->spec.data(), aux->spec.data(), mesh);

  FldMul(aux1->phys.data(), v->phys.data(), r1curr->phys.data(), -1.0, n);
  FwdTrans(aux1->spec.data(), aux1->phys.data(), mesh);
  yDeriv(aux1->spec.data(), aux1->spec.data(), mesh);
  SpecAdd(aux->spec.data(), 1.0, aux->spec.data(), 1.0, aux1->spec.data(), ns);
  SpecAdd(r1nonl->spec.data(), 1.0, r1nonl->spec.data(), 1.0, aux->spec.data(), ns);

  xDeriv(aux->spec.data(), u->spec.data(), mesh);
  BwdTrans(aux->phys.data(), aux->spec.data(), mesh);
  FldMul(aux->phys.data(), S->phys.data(), aux->phys.data(), lambda, n);
  FwdTrans(aux->spec.data(), aux->phys.data(), mesh);
  SpecAdd(r1nonl->spec.data(), 1.0, r1nonl->spec.data(), 1.0, aux->spec.data(), ns);
}

// :257-300
void r2nonl_func(OField* r2nonl, OField* wcurr, OField* u, OField* v, OField* r1curr, OField* r2curr,
                 OField* S, double lambda, OField* aux, OField* aux1, const Coef& c) {
  OMesh* mesh = r2nonl->mesh;
  const long n = mesh->np(), ns = mesh->ns();
  FldMul(r2nonl->phys.data(), wcurr->phys.data(), r1curr->phys.data(), 1.0, n);
  S2r(r2nonl->phys.data(), r2curr->phys.data(), S->phys.data(), c, n);
  FwdTrans(r2nonl->spec.data(), r2nonl->phys.data(), mesh);

  FldMul(aux->phys.data(), u->phys.data(), r2curr->phys.data(), -1.0, n);
  FwdTrans(aux->spec.data(), aux->phys.data(), mesh);
  xDeriv(aux->spec.data(), aux->spec.data(), mesh);

  FldMul(aux1->phys.data(), v->phys.data(), r2curr->phys.data(), -1.0, n);
  FwdTrans(aux1->spec.data(), aux1->phys.data(), mesh);
  yDeriv(aux1->spec.data(), aux1->spec.data(), mesh);
  SpecAdd(aux->spec.data(), 1.0, aux->spec.data(), 1.0, aux1->spec.data(), ns);
  SpecAdd(r2nonl->spec.data(), 1.0, r2nonl->spec.data(), 1.0, aux->spec.data(), ns);

  yDeriv(aux->spec.data(), u->spec.data(), mesh);
  xDeriv(aux1->spec.data(), v->spec.data(), mesh);
  SpecAdd(aux->spec.data(), 1.0, aux->spec.data(), 1.0, aux1->spec.data(), ns);
  BwdTrans(aux->phys.data(), aux->spec.data(), mesh);
  FldMul(aux->phys.data(), S->phys.data(), aux->phys.data(), 0.5 * lambda, n);
  FwdTrans(aux->spec.data(), aux->phys.data(), mesh);
  SpecAdd(r2nonl->spec.data(), 1.0, r2nonl->spec.data(), 1.0, aux->spec.data(), ns);
}

// :302-312
void curr_func(OField* r1curr, OField* r2curr, OField* wcurr, OField* u, OField* v, OField* S) {
  OMesh* mesh = r1curr->mesh;
  BwdTrans(r1curr->phys.data(), r1curr->spec.data(), mesh);
  BwdTrans(r2curr->phys.data(), r2curr->spec.data(), mesh);
  BwdTrans(wcurr->phys.data(), wcurr->spec.data(), mesh);
  S_func(S->phys.data(), r1curr->phys.data(), r2curr->phys.data(), mesh->np());
  vel_func(wcurr->spec.data(), u->spec.data(), v->spec.data(), mesh);
  BwdTrans(u->phys.data(), u->spec.data(), mesh);
  BwdTrans(v->phys.data(), v->spec.data(), mesh);
}

// :314-322
void dealias(cd* ft, const OMesh* m) {
  const double* cutoff = m->cutoff.data();
  OR_LOOP(m->ns()) ft[index] = ft[index] * cutoff[index];
}

// ---------------------------------------------------------------------------
// Time integration (TimeIntegration/*.cuh)
// ---------------------------------------------------------------------------
// IFRK2.cuh:16-26 and RK4.cuh:12-22 (identical)
void func0D(const cd* old_spec, cd* curr_spec, cd* new_spec, const double* IF, long n) {
  OR_LOOP(n) {
    new_spec[index] = old_spec[index] * IF[index];
    curr_spec[index] = old_spec[index];
  }
}
// IFE.cuh:15-25
void ife_func1D(const cd* old_spec, cd* new_spec, const cd* nonl_spec, const double* IF, double dt, long n) {
  OR_LOOP(n) {
    cd an = nonl_spec[index] * dt;
    new_spec[index] = (old_spec[index] + an) * IF[index];
  }
}
// IFRK2.cuh:28-39
void ifrk2_func1D(const cd* old_spec, cd* curr_spec, cd* new_spec, const cd* nonl_spec,
                  const double* IF, const double* IFh, double dt, long n) {
  OR_LOOP(n) {
    cd an = nonl_spec[index] * dt;
    new_spec[index] = new_spec[index] + 1.0 / 2.0 * (IF[index]) * an;
    curr_spec[index] = (old_spec[index] + an) * (IFh[index]);
  }
}
// IFRK2.cuh:41-51
void ifrk2_func2D(cd* new_spec, const cd* nonl_spec, double dt, long n) {
  OR_LOOP(n) {
    cd bn = nonl_spec[index] * dt;
    new_spec[index] = new_spec[index] + 1.0 / 2.0 * bn;
  }
}
// RK4.cuh:25-74
void rk4_func1D(const cd* old_spec, cd* curr_spec, cd* new_spec, const cd* nonl_spec,
                const double* IF, const double* IFh, double dt, long n) {
  OR_LOOP(n) {
    cd an = nonl_spec[index] * dt;
    new_spec[index] = new_spec[index] + 1.0 / 6.0 * (IF[index]) * an;
    curr_spec[index] = (old_spec[index] + an / 2.0) * (IFh[index]);
  }
}
void rk4_func2D(const cd* old_spec, cd* curr_spec, cd* new_spec, const cd* nonl_spec,
                const double* IFh, double dt, long n) {
  OR_LOOP(n) {
    cd bn = nonl_spec[index] * dt;
    new_spec[index] = new_spec[index] + 1.0 / 3.0 * IFh[index] * bn;
    curr_spec[index] = old_spec[index] * (IFh[index]) + bn / 2.0;
  }
}
void rk4_func3D(const cd* old_spec, cd* curr_spec, cd* new_spec, const cd* nonl_spec,
                const double* IF, const double* IFh, double dt, long n) {
  OR_LOOP(n) {
    cd cn = nonl_spec[index] * dt;
    new_spec[index] = new_spec[index] + 1.0 / 3.0 * IFh[index] * cn;
    curr_spec[index] = old_spec[index] * IF[index] + cn * IFh[index];
  }
}
void rk4_func4D(cd* new_spec, const cd* nonl_spec, double dt, long n) {
  OR_LOOP(n) {
    cd dn = nonl_spec[index] * dt;
    new_spec[index] = new_spec[index] + 1.0 / 6.0 * dn;
  }
}

// ---------------------------------------------------------------------------
// Solver: fields of InertialFriction/friction/friction.cu:150-162, IF tables
// of :196-210, time loop of :212-260 with the IFRK2 stage kernels.
// ---------------------------------------------------------------------------
struct OSolver {
  OMesh* mesh;
  double lambda, Re, Er, Rf, alpha, dt;
  Coef c;
  bool advect_w;
  std::unique_ptr<OField> wold, wcurr, wnew, wnonl, u, v, r1old, r1curr, r1new, r1nonl, r2old,
      r2curr, r2new, r2nonl, S, p11, p12, p21, Ra, aux, aux1;
  std::vector<double> IFw, IFwh, IFr1, IFr1h, IFr2, IFr2h;   // exp(L dt), exp(L dt/2)

  OSolver(OMesh* m, const double* p) : mesh(m) {
    lambda = p[0]; Re = p[1]; Er = p[2]; Rf = p[3]; alpha = p[4];
    c.DQ = p[5]; c.aQ = p[6]; c.cQ = p[7]; dt = p[8]; advect_w = p[9] != 0.0;
    for (auto* f : {&wold, &wcurr, &wnew, &wnonl, &u, &v, &r1old, &r1curr, &r1new, &r1nonl, &r2old,
                    &r2curr, &r2new, &r2nonl, &S, &p11, &p12, &p21, &Ra, &aux, &aux1})
      f->reset(new OField(m));
    FldSet(Ra->phys.data(), alpha, m->np());
    const size_t ns = m->ns();
    for (auto* t : {&IFw, &IFwh, &IFr1, &IFr1h, &IFr2, &IFr2h}) t->resize(ns);
    for (size_t index = 0; index < ns; ++index) {
      const double k2 = m->k_squared[index];
      // friction/Qsolver.cuh:20-30 (wlin_func)
      double a = 1.0 / Re * (-1.0 * k2 - Rf);
      IFw[index] = std::exp(a * dt);
      IFwh[index] = std::exp(a * dt * 0.5);
      // friction/Qsolver.cuh:32-54 (r1lin_func, r2lin_func)
      double a1 = (-1.0 * c.DQ * k2 + c.aQ);
      IFr1[index] = std::exp(a1 * dt);
      IFr1h[index] = std::exp(a1 * dt * 0.5);
      IFr2[index] = IFr1[index];
      IFr2h[index] = IFr1h[index];
    }
  }

  void rhs() {
    curr_func(r1curr.get(), r2curr.get(), wcurr.get(), u.get(), v.get(), S.get());
    wnonl_func(wnonl.get(), wcurr.get(), u.get(), v.get(), r1curr.get(), r2curr.get(), Ra.get(),
               S.get(), p11.get(), p12.get(), p21.get(), Re, Er, lambda, aux.get(), aux1.get(), c,
               advect_w);
    r1nonl_func(r1nonl.get(), wcurr.get(), u.get(), v.get(), r1curr.get(), r2curr.get(), S.get(),
                lambda, aux.get(), aux1.get(), c);
    r2nonl_func(r2nonl.get(), wcurr.get(), u.get(), v.get(), r1curr.get(), r2curr.get(), S.get(),
                lambda, aux.get(), aux1.get(), c);
  }

  struct Trip { OField *old_, *curr, *nw, *nonl; const double *IF, *IFh; };
  std::vector<Trip> trips() {
    return {{wold.get(), wcurr.get(), wnew.get(), wnonl.get(), IFw.data(), IFwh.data()},
            {r1old.get(), r1curr.get(), r1new.get(), r1nonl.get(), IFr1.data(), IFr1h.data()},
            {r2old.get(), r2curr.get(), r2new.get(), r2nonl.get(), IFr2.data(), IFr2h.data()}};
  }

  void finish_step() {
    const long ns = mesh->ns();
    for (auto& t : trips()) dealias(t.nw->spec.data(), mesh);          // friction.cu:253-255
    for (auto& t : trips()) std::memcpy(t.old_->spec.data(), t.nw->spec.data(), sizeof(cd) * ns);  // :257-259
  }

  // IFRK2 step: friction.cu:212-260 with TimeIntegration/IFRK2.cuh:16-51 and
  // IFh := exp(L dt) (the full-step factor that makes the scheme 2nd order).
  void step_ifrk2() {
    const long ns = mesh->ns();
    for (auto& t : trips()) func0D(t.old_->spec.data(), t.curr->spec.data(), t.nw->spec.data(), t.IF, ns);
    rhs();
    for (auto& t : trips())
      ifrk2_func1D(t.old_->spec.data(), t.curr->spec.data(), t.nw->spec.data(), t.nonl->spec.data(),
                   t.IF, t.IF, dt, ns);
    rhs();
    for (auto& t : trips()) ifrk2_func2D(t.nw->spec.data(), t.nonl->spec.data(), dt, ns);
    finish_step();
  }

  // IF-RK4 step: friction.cu:212-260 verbatim (IF = exp(L dt), IFh = exp(L dt/2)).
  void step_rk4() {
    const long ns = mesh->ns();
    for (auto& t : trips()) func0D(t.old_->spec.data(), t.curr->spec.data(), t.nw->spec.data(), t.IF, ns);
    rhs();
    for (auto& t : trips())
      rk4_func1D(t.old_->spec.data(), t.curr->spec.data(), t.nw->spec.data(), t.nonl->spec.data(), t.IF,
                 t.IFh, dt, ns);
    rhs();
    for (auto& t : trips())
      rk4_func2D(t.old_->spec.data(), t.curr->spec.data(), t.nw->spec.data(), t.nonl->spec.data(), t.IFh,
                 dt, ns);
    rhs();
    for (auto& t : trips())
      rk4_func3D(t.old_->spec.data(), t.curr->spec.data(), t.nw->spec.data(), t.nonl->spec.data(), t.IF,
                 t.IFh, dt, ns);
    rhs();
    for (auto& t : trips()) rk4_func4D(t.nw->spec.data(), t.nonl->spec.data(), dt, ns);
    finish_step();
  }

  // IF-Euler step: curr <- old, one RHS, TimeIntegration/IFE.cuh:15-25.
  void step_ife() {
    const long ns = mesh->ns();
    for (auto& t : trips()) std::memcpy(t.curr->spec.data(), t.old_->spec.data(), sizeof(cd) * ns);
    rhs();
    for (auto& t : trips())
      ife_func1D(t.old_->spec.data(), t.nw->spec.data(), t.nonl->spec.data(), t.IF, dt, ns);
    finish_step();
  }
};


// ---------------------------------------------------------------------------
// Discrete adjoint of the IFRK2 step (no reference counterpart: the reference
// has no adjoint).  Exact transpose, in the real inner products
//   <F, G> = sum_{j, i < Nxh} Re(F conj G)   (half spectra as stored)
//   <f, g> = sum_{y, x} f g                   (physical fields),
// of the right-hand side the functions above compute.  That right-hand side
// equals, for every state X = (W, R1, R2) (BwdTrans's column-0 side effect
// changes nothing downstream of a C2R), the compact form
//   w = B W, uyvx = B((kx^2-ky^2) psi), u = B(i ky psi), v = B(-i kx psi),
//   ux = B(-kx ky psi), r_a = B R_a, L_a = B(-k^2 R_a),  psi = W/k^2 (0 at k=0)
//   N_w  = Re4[i kx' F(-u w) + i ky F(-v w)] + s(-kx^2 F P12 + 2 kx ky F P11 + ky^2 F P21)
//   N_ra = F(loc_a) + i kx F(-u r_a) + i ky F(-v r_a)
// with B the A10 C2R / (Nx Ny), F the unnormalised R2C, kx' = kx except 0 on
// the Nyquist column (the x-derivative that BwdTrans/FwdTrans projects away,
// friction/Qsolver.cuh:179-193), Re4 the real part at the four self-conjugate
// bins, s = 1/(Er Re), and the pointwise terms of Qsolver.cuh:63-300.
// Transposes used:  F^T G = NN B(Wt G) (Wt = 1 on columns 0 and Nx/2, 1/2
// elsewhere);  B^T f = c F(f) / NN (c = 1 on those columns, 2 elsewhere);
// multiplication by a complex c transposes to multiplication by conj(c).
// ---------------------------------------------------------------------------
struct OAdj {
  OSolver* s;
  OMesh* m;
  std::vector<cd> tmp;
  explicit OAdj(OSolver* sv) : s(sv), m(sv->mesh), tmp(sv->mesh->ns()) {}
  long ns() const { return (long)m->ns(); }
  long np() const { return (long)m->np(); }
  double NN() const { return (double)m->Nx * m->Ny; }
  bool edge(long index) const { const int i = (int)(index % m->Nxh); return i == 0 || 2 * i == m->Nx; }

  void B(const cd* G, double* f) {   // A10 C2R / NN, input preserved
    std::memcpy(tmp.data(), G, sizeof(cd) * ns());
    c2r2d(m, tmp.data(), f);
    coeff(f, m->Nx, m->Ny);
  }
  void F(const double* f, cd* G) {   // unnormalised R2C, input preserved
    std::memcpy(m->phys.data(), f, sizeof(double) * np());
    r2c2d(m, m->phys.data(), G);
  }
  void FT(const cd* G, double* f) {  // F^T
    std::vector<cd> g(G, G + ns());
    for (long k = 0; k < ns(); ++k) if (!edge(k)) g[k] = 0.5 * g[k];
    B(g.data(), f);
    const double nn = NN();
    for (long k = 0; k < np(); ++k) f[k] *= nn;
  }
  void BT(const double* f, cd* G) {  // B^T
    F(f, G);
    const double nn = NN();
    for (long k = 0; k < ns(); ++k) G[k] = (edge(k) ? 1.0 : 2.0) / nn * G[k];
  }

  struct Fields { std::vector<double> w, uyvx, u, v, ux, r1, r2, L1, L2; };
  void forward_fields(const cd* W, const cd* R1, const cd* R2, Fields& f) {
    const long n = ns();
    const double *kx = m->kx.data(), *ky = m->ky.data(), *k2 = m->k_squared.data();
    for (auto* v : {&f.w, &f.uyvx, &f.u, &f.v, &f.ux, &f.r1, &f.r2, &f.L1, &f.L2}) v->assign(np(), 0.0);
    std::vector<cd> g(n);
    auto from = [&](std::vector<double>& out, auto coef) {
      for (long k = 0; k < n; ++k) g[k] = coef(k);
      B(g.data(), out.data());
    };
    auto psi = [&](long k) { return k == 0 ? mk(0, 0) : W[k] / k2[k]; };
    from(f.w, [&](long k) { return W[k]; });
    from(f.uyvx, [&](long k) { return (kx[k] * kx[k] - ky[k] * ky[k]) * psi(k); });
    from(f.u, [&](long k) { return im() * ky[k] * psi(k); });
    from(f.v, [&](long k) { return -1.0 * kx[k] * im() * psi(k); });
    from(f.ux, [&](long k) { return -kx[k] * ky[k] * psi(k); });
    from(f.r1, [&](long k) { return R1[k]; });
    from(f.r2, [&](long k) { return R2[k]; });
    from(f.L1, [&](long k) { return -k2[k] * R1[k]; });
    from(f.L2, [&](long k) { return -k2[k] * R2[k]; });
  }

  // (Xbar) += VJP of the right-hand side at X applied to Nbar; abar += the
  // activity field's part (alpha enters P11, P12, P21 as -alpha r).
  void rhs_vjp(const cd* W, const cd* R1, const cd* R2, const double* Ra, const cd* Nbw, const cd* Nbr1,
               const cd* Nbr2, cd* Wb, cd* R1b, cd* R2b, double* abar) {
    const long n = ns(), P = np();
    const int Nx = m->Nx, Ny = m->Ny, Nxh = m->Nxh;
    const double *kx = m->kx.data(), *ky = m->ky.data(), *k2 = m->k_squared.data();
    const double sE = 1.0 / (s->Er * s->Re), lam = s->lambda, DQ = s->c.DQ, aQ = s->c.aQ, cQ = s->c.cQ;
    Fields f;
    forward_fields(W, R1, R2, f);
    // spectral adjoints of the products
    std::vector<cd> Ob(Nbw, Nbw + n);
    for (int j : {0, Ny / 2})
      for (int i : {0, Nx / 2}) Ob[(size_t)j * Nxh + i].y = 0.0;
    enum { A_, B_, P11, P12, P21, L1c, L2c, U1, V1, U2, V2, NPROD };
    std::vector<std::vector<double>> pb(NPROD, std::vector<double>(P));
    std::vector<cd> g(n);
    auto adj = [&](int which, auto coef) {
      for (long k = 0; k < n; ++k) g[k] = coef(k);
      FT(g.data(), pb[which].data());
    };
    auto ikx_nyq0 = [&](long k) { return (2 * (k % Nxh) == Nx) ? 0.0 : kx[k]; };
    if (s->advect_w) {
      adj(A_, [&](long k) { return mk(0.0, -ikx_nyq0(k)) * Ob[k]; });
      adj(B_, [&](long k) { return mk(0.0, -ky[k]) * Ob[k]; });
    }
    adj(P12, [&](long k) { return (-sE * kx[k] * kx[k]) * Ob[k]; });
    adj(P11, [&](long k) { return (2.0 * sE * kx[k] * ky[k]) * Ob[k]; });
    adj(P21, [&](long k) { return (sE * ky[k] * ky[k]) * Ob[k]; });
    adj(L1c, [&](long k) { return Nbr1[k]; });
    adj(U1, [&](long k) { return mk(0.0, -kx[k]) * Nbr1[k]; });
    adj(V1, [&](long k) { return mk(0.0, -ky[k]) * Nbr1[k]; });
    adj(L2c, [&](long k) { return Nbr2[k]; });
    adj(U2, [&](long k) { return mk(0.0, -kx[k]) * Nbr2[k]; });
    adj(V2, [&](long k) { return mk(0.0, -ky[k]) * Nbr2[k]; });
    // pointwise adjoint
    Fields b;
    for (auto* v : {&b.w, &b.uyvx, &b.u, &b.v, &b.ux, &b.r1, &b.r2, &b.L1, &b.L2}) v->assign(P, 0.0);
    for (long x = 0; x < P; ++x) {
      const double w = f.w[x], uyvx = f.uyvx[x], u = f.u[x], v = f.v[x], ux = f.ux[x];
      const double r1 = f.r1[x], r2 = f.r2[x], L1 = f.L1[x], L2 = f.L2[x];
      const double al = Ra[x];
      const double rr = std::sqrt(r1 * r1 + r2 * r2), S = 2 * rr, cS2 = cQ * S * S;
      const double gg = lam * S * (cS2 - aQ);
      double wb = 0, uyvxb = 0, ub = 0, vb = 0, uxb = 0, r1b = 0, r2b = 0, L1b = 0, L2b = 0;
      double Sb = 0, cS2b = 0, ggb = 0;
      if (s->advect_w) {
        const double Ab = pb[A_][x], Bb = pb[B_][x];
        ub += -w * Ab; wb += -u * Ab;
        vb += -w * Bb; wb += -v * Bb;
      }
      { const double t = pb[U1][x]; ub += -r1 * t; r1b += -u * t; }
      { const double t = pb[V1][x]; vb += -r1 * t; r1b += -v * t; }
      { const double t = pb[U2][x]; ub += -r2 * t; r2b += -u * t; }
      { const double t = pb[V2][x]; vb += -r2 * t; r2b += -v * t; }
      const double p11b = pb[P11][x], p12b = pb[P12][x], p21b = pb[P21][x];
      const double s1b = p11b, s2b = p12b + p21b, crb = p12b - p21b;
      r1b += -al * p11b;
      r2b += -al * (p12b + p21b);
      abar[x] += -r1 * p11b - r2 * (p12b + p21b);
      { const double t = pb[L1c][x];   // loc1 = -w r2 - cS2 r1 + lam S ux
        wb += -r2 * t; r2b += -w * t; cS2b += -r1 * t; r1b += -cS2 * t; Sb += lam * ux * t; uxb += lam * S * t; }
      { const double t = pb[L2c][x];   // loc2 = w r1 - cS2 r2 + 1/2 lam S uyvx
        wb += r1 * t; r1b += w * t; cS2b += -r2 * t; r2b += -cS2 * t; Sb += 0.5 * lam * uyvx * t;
        uyvxb += 0.5 * lam * S * t; }
      // cr = 2 DQ (L1 r2 - r1 L2)
      L1b += 2 * DQ * r2 * crb; r2b += 2 * DQ * L1 * crb; r1b += -2 * DQ * L2 * crb; L2b += -2 * DQ * r1 * crb;
      // s_a = -lam DQ L_a S + gg r_a
      L1b += -lam * DQ * S * s1b; Sb += -lam * DQ * L1 * s1b; ggb += r1 * s1b; r1b += gg * s1b;
      L2b += -lam * DQ * S * s2b; Sb += -lam * DQ * L2 * s2b; ggb += r2 * s2b; r2b += gg * s2b;
      // gg = lam S (cS2 - aQ), cS2 = cQ S^2, S = 2 sqrt(r1^2 + r2^2)
      Sb += ggb * lam * (cS2 - aQ);
      cS2b += ggb * lam * S;
      Sb += cS2b * 2.0 * cQ * S;
      if (rr > 0) { r1b += Sb * 2.0 * r1 / rr; r2b += Sb * 2.0 * r2 / rr; }
      b.w[x] = wb; b.uyvx[x] = uyvxb; b.u[x] = ub; b.v[x] = vb; b.ux[x] = uxb;
      b.r1[x] = r1b; b.r2[x] = r2b; b.L1[x] = L1b; b.L2[x] = L2b;
    }
    // spectral adjoint of the inverse maps
    std::vector<cd> t(n);
    auto acc = [&](cd* out, const std::vector<double>& fb, auto coef) {
      BT(fb.data(), t.data());
      for (long k = 0; k < n; ++k) out[k] = out[k] + coef(k) * t[k];
    };
    auto ipsi = [&](long k) { return k == 0 ? 0.0 : 1.0 / k2[k]; };
    acc(Wb, b.w, [&](long) { return mk(1, 0); });
    acc(Wb, b.uyvx, [&](long k) { return mk((kx[k] * kx[k] - ky[k] * ky[k]) * ipsi(k), 0); });
    acc(Wb, b.u, [&](long k) { return mk(0, -ky[k] * ipsi(k)); });
    acc(Wb, b.v, [&](long k) { return mk(0, kx[k] * ipsi(k)); });
    acc(Wb, b.ux, [&](long k) { return mk(-kx[k] * ky[k] * ipsi(k), 0); });
    acc(R1b, b.r1, [&](long) { return mk(1, 0); });
    acc(R1b, b.L1, [&](long k) { return mk(-k2[k], 0); });
    acc(R2b, b.r2, [&](long) { return mk(1, 0); });
    acc(R2b, b.L2, [&](long k) { return mk(-k2[k], 0); });
  }

  // One IFRK2 step's VJP: (u_n, gbar = dL/du_{n+1}) -> dL/du_n, abar += ...
  // Forward (IFRK2.cuh:16-51, IFh = IF, then dealias):
  //   c2 = E (u + dt N(u)),  u_{n+1} = M (E u + dt/2 E N(u) + dt/2 N(c2)).
  void step_vjp(const cd* const* u, const cd* const* gbar, cd* const* ubar, double* abar) {
    const long n = ns();
    const double dt = s->dt;
    const double* E[3] = {s->IFw.data(), s->IFr1.data(), s->IFr2.data()};
    const double* Ra = s->Ra->phys.data();
    std::vector<cd> N1[3], c2[3], lamv[3], nb[3], cb[3];
    for (int a = 0; a < 3; ++a) { N1[a].resize(n); c2[a].resize(n); lamv[a].resize(n); nb[a].resize(n); cb[a].assign(n, mk(0, 0)); }
    // forward recompute: N(u) through the oracle's own right-hand side
    {
      std::vector<cd> w(u[0], u[0] + n), r1(u[1], u[1] + n), r2(u[2], u[2] + n);
      or_rhs_into(w.data(), r1.data(), r2.data(), N1[0].data(), N1[1].data(), N1[2].data());
    }
    for (int a = 0; a < 3; ++a)
      for (long k = 0; k < n; ++k) c2[a][k] = (u[a][k] + dt * N1[a][k]) * E[a][k];
    const double* cut = m->cutoff.data();
    for (int a = 0; a < 3; ++a)
      for (long k = 0; k < n; ++k) { lamv[a][k] = cut[k] * gbar[a][k]; nb[a][k] = 0.5 * dt * lamv[a][k]; }
    rhs_vjp(c2[0].data(), c2[1].data(), c2[2].data(), Ra, nb[0].data(), nb[1].data(), nb[2].data(),
            cb[0].data(), cb[1].data(), cb[2].data(), abar);
    for (int a = 0; a < 3; ++a)
      for (long k = 0; k < n; ++k) {
        ubar[a][k] = E[a][k] * (lamv[a][k] + cb[a][k]);
        nb[a][k] = E[a][k] * (0.5 * dt * lamv[a][k] + dt * cb[a][k]);
        cb[a][k] = mk(0, 0);
      }
    rhs_vjp(u[0], u[1], u[2], Ra, nb[0].data(), nb[1].data(), nb[2].data(), cb[0].data(), cb[1].data(),
            cb[2].data(), abar);
    for (int a = 0; a < 3; ++a)
      for (long k = 0; k < n; ++k) ubar[a][k] = ubar[a][k] + cb[a][k];
  }

  void or_rhs_into(cd* w, cd* r1, cd* r2, cd* wn, cd* r1n, cd* r2n) {
    const size_t bb = sizeof(cd) * ns();
    std::memcpy(s->wcurr->spec.data(), w, bb);
    std::memcpy(s->r1curr->spec.data(), r1, bb);
    std::memcpy(s->r2curr->spec.data(), r2, bb);
    s->rhs();
    std::memcpy(wn, s->wnonl->spec.data(), bb);
    std::memcpy(r1n, s->r1nonl->spec.data(), bb);
    std::memcpy(r2n, s->r2nonl->spec.data(), bb);
  }

  // Loss at a state and dL/dW (kind 0: 1/2 mean(w^2); kind 1: 1/2 mean(u^2 + v^2)).
  double loss_grad(const cd* W, int kind, cd* Wb) {
    const long n = ns(), P = np();
    const double *kx = m->kx.data(), *ky = m->ky.data(), *k2 = m->k_squared.data();
    const double nn = NN();
    std::vector<cd> g(n), t(n);
    std::vector<double> f(P), h(P);
    double L = 0;
    for (long k = 0; k < n; ++k) Wb[k] = mk(0, 0);
    if (kind == 0) {
      B(W, f.data());
      for (long x = 0; x < P; ++x) { L += 0.5 * f[x] * f[x] / nn; f[x] /= nn; }
      BT(f.data(), Wb);
      return L;
    }
    for (int c = 0; c < 2; ++c) {
      for (long k = 0; k < n; ++k) {
        const double ip = k == 0 ? 0.0 : 1.0 / k2[k];
        g[k] = c == 0 ? mk(0, ky[k] * ip) * W[k] : mk(0, -kx[k] * ip) * W[k];
      }
      B(g.data(), f.data());
      for (long x = 0; x < P; ++x) { L += 0.5 * f[x] * f[x] / nn; f[x] /= nn; }
      BT(f.data(), t.data());
      for (long k = 0; k < n; ++k) {
        const double ip = k == 0 ? 0.0 : 1.0 / k2[k];
        const cd c2 = c == 0 ? mk(0, -ky[k] * ip) : mk(0, kx[k] * ip);
        Wb[k] = Wb[k] + c2 * t[k];
      }
    }
    return L;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
extern "C" {

void* or_mesh_new(int Nx, int Ny, double Lx, double Ly) { return new OMesh(Nx, Ny, Lx, Ly); }
void or_mesh_free(void* m) { delete static_cast<OMesh*>(m); }
void or_tables(void* mv, double* kx, double* ky, double* k2, double* cutoff) {
  OMesh* m = static_cast<OMesh*>(mv);
  const size_t b = sizeof(double) * m->ns();
  if (kx) std::memcpy(kx, m->kx.data(), b);
  if (ky) std::memcpy(ky, m->ky.data(), b);
  if (k2) std::memcpy(k2, m->k_squared.data(), b);
  if (cutoff) std::memcpy(cutoff, m->cutoff.data(), b);
}
void* or_field_new(void* m) { return new OField(static_cast<OMesh*>(m)); }
void or_field_free(void* f) { delete static_cast<OField*>(f); }
or_comp* or_field_spec(void* f) { return static_cast<OField*>(f)->spec.data(); }
double* or_field_phys(void* f) { return static_cast<OField*>(f)->phys.data(); }

void or_fft1d(int n, int sign, const or_comp* in, or_comp* out) {
  Fft1 f(n);
  std::vector<cd> s(n);
  std::memcpy(out, in, sizeof(cd) * n);
  f.exec(out, s.data(), sign);
}

void or_FwdTrans(or_comp* ft, const double* f, void* m) { FwdTrans(ft, f, static_cast<OMesh*>(m)); }
void or_BwdTrans(double* f, or_comp* ft, void* m) { BwdTrans(f, ft, static_cast<OMesh*>(m)); }

#define F(p) static_cast<OField*>(p)
void or_single(void* s0, void* r, void* S, double lambda) { single(F(s0), F(r), F(S), lambda, Coef()); }
void or_cross(void* cs, void* r1, void* r2, void* aux) { cross(F(cs), F(r1), F(r2), F(aux), Coef()); }
void or_p11func(void* p11, void* r1, void* S, void* Ra, double lambda) {
  p11func(F(p11), F(r1), F(S), F(Ra), lambda, Coef());
}
void or_p12func(void* p12, void* r1, void* r2, void* S, void* Ra, double lambda, void* aux, void* aux1) {
  p12func(F(p12), F(r1), F(r2), F(S), F(Ra), lambda, F(aux), F(aux1), Coef());
}
void or_p21func(void* p21, void* r1, void* r2, void* S, void* Ra, double lambda, void* aux, void* aux1) {
  p21func(F(p21), F(r1), F(r2), F(S), F(Ra), lambda, F(aux), F(aux1), Coef());
}
void or_wnonl_func(void* wnonl, void* wcurr, void* u, void* v, void* r1curr, void* r2curr, void* Ra,
                   void* S, void* p11, void* p12, void* p21, double Re, double Er, double lambda,
                   void* aux, void* aux1) {
  wnonl_func(F(wnonl), F(wcurr), F(u), F(v), F(r1curr), F(r2curr), F(Ra), F(S), F(p11), F(p12), F(p21),
             Re, Er, lambda, F(aux), F(aux1), Coef(), true);
}
void or_r1nonl_func(void* r1nonl, void* wcurr, void* u, void* v, void* r1curr, void* r2curr, void* S,
                    double lambda, void* aux, void* aux1) {
  r1nonl_func(F(r1nonl), F(wcurr), F(u), F(v), F(r1curr), F(r2curr), F(S), lambda, F(aux), F(aux1),
              Coef());
}
void or_r2nonl_func(void* r2nonl, void* wcurr, void* u, void* v, void* r1curr, void* r2curr, void* S,
                    double lambda, void* aux, void* aux1) {
  r2nonl_func(F(r2nonl), F(wcurr), F(u), F(v), F(r1curr), F(r2curr), F(S), lambda, F(aux), F(aux1),
              Coef());
}
void or_curr_func(void* r1curr, void* r2curr, void* wcurr, void* u, void* v, void* S) {
  curr_func(F(r1curr), F(r2curr), F(wcurr), F(u), F(v), F(S));
}
#undef F

void* or_solver_new(void* m, const double* p) { return new OSolver(static_cast<OMesh*>(m), p); }
void or_solver_free(void* s) { delete static_cast<OSolver*>(s); }
void or_solver_set_activity(void* sv, const double* Ra) {
  OSolver* s = static_cast<OSolver*>(sv);
  if (Ra) std::memcpy(s->Ra->phys.data(), Ra, sizeof(double) * s->mesh->np());
  else FldSet(s->Ra->phys.data(), s->alpha, s->mesh->np());
}
void or_solver_set_state(void* sv, const or_comp* w, const or_comp* r1, const or_comp* r2) {
  OSolver* s = static_cast<OSolver*>(sv);
  const size_t b = sizeof(cd) * s->mesh->ns();
  std::memcpy(s->wold->spec.data(), w, b);
  std::memcpy(s->r1old->spec.data(), r1, b);
  std::memcpy(s->r2old->spec.data(), r2, b);
}
void or_solver_get_state(void* sv, or_comp* w, or_comp* r1, or_comp* r2) {
  OSolver* s = static_cast<OSolver*>(sv);
  const size_t b = sizeof(cd) * s->mesh->ns();
  std::memcpy(w, s->wold->spec.data(), b);
  std::memcpy(r1, s->r1old->spec.data(), b);
  std::memcpy(r2, s->r2old->spec.data(), b);
}
void or_solver_step(void* sv, int method, int nsteps) {
  OSolver* s = static_cast<OSolver*>(sv);
  for (int k = 0; k < nsteps; ++k) {
    if (method == OR_IFRK2) s->step_ifrk2();
    else if (method == OR_RK4) s->step_rk4();
    else s->step_ife();
  }
}
void or_solver_rhs(void* sv, or_comp* w, or_comp* r1, or_comp* r2, or_comp* wn, or_comp* r1n,
                   or_comp* r2n) {
  OSolver* s = static_cast<OSolver*>(sv);
  const size_t b = sizeof(cd) * s->mesh->ns();
  std::memcpy(s->wcurr->spec.data(), w, b);
  std::memcpy(s->r1curr->spec.data(), r1, b);
  std::memcpy(s->r2curr->spec.data(), r2, b);
  s->rhs();
  std::memcpy(w, s->wcurr->spec.data(), b);
  std::memcpy(r1, s->r1curr->spec.data(), b);
  std::memcpy(r2, s->r2curr->spec.data(), b);
  std::memcpy(wn, s->wnonl->spec.data(), b);
  std::memcpy(r1n, s->r1nonl->spec.data(), b);
  std::memcpy(r2n, s->r2nonl->spec.data(), b);
}
void or_solver_lin(void* sv, double* IFw, double* IFwh, double* IFr1, double* IFr1h, double* IFr2,
                   double* IFr2h) {
  OSolver* s = static_cast<OSolver*>(sv);
  const size_t b = sizeof(double) * s->mesh->ns();
  std::memcpy(IFw, s->IFw.data(), b);
  std::memcpy(IFwh, s->IFwh.data(), b);
  std::memcpy(IFr1, s->IFr1.data(), b);
  std::memcpy(IFr1h, s->IFr1h.data(), b);
  std::memcpy(IFr2, s->IFr2.data(), b);
  std::memcpy(IFr2h, s->IFr2h.data(), b);
}


// ---- discrete adjoint (test infrastructure, no reference counterpart) ----
void or_adj_B(void* sv, const or_comp* G, double* f) { OAdj(static_cast<OSolver*>(sv)).B(G, f); }
void or_adj_BT(void* sv, const double* f, or_comp* G) { OAdj(static_cast<OSolver*>(sv)).BT(f, G); }
void or_adj_F(void* sv, const double* f, or_comp* G) { OAdj(static_cast<OSolver*>(sv)).F(f, G); }
void or_adj_FT(void* sv, const or_comp* G, double* f) { OAdj(static_cast<OSolver*>(sv)).FT(G, f); }
void or_adj_rhs_vjp(void* sv, const or_comp* W, const or_comp* R1, const or_comp* R2, const or_comp* Nbw,
                    const or_comp* Nbr1, const or_comp* Nbr2, or_comp* Wb, or_comp* R1b, or_comp* R2b,
                    double* abar) {
  OSolver* s = static_cast<OSolver*>(sv);
  const size_t n = s->mesh->ns();
  for (or_comp* p : {Wb, R1b, R2b}) std::memset(p, 0, sizeof(or_comp) * n);
  OAdj(s).rhs_vjp(W, R1, R2, s->Ra->phys.data(), Nbw, Nbr1, Nbr2, Wb, R1b, R2b, abar);
}
void or_adj_step_vjp(void* sv, const or_comp* W, const or_comp* R1, const or_comp* R2, const or_comp* gW,
                     const or_comp* gR1, const or_comp* gR2, or_comp* uW, or_comp* uR1, or_comp* uR2,
                     double* abar) {
  const or_comp* u[3] = {W, R1, R2};
  const or_comp* g[3] = {gW, gR1, gR2};
  or_comp* ub[3] = {uW, uR1, uR2};
  OAdj(static_cast<OSolver*>(sv)).step_vjp(u, g, ub, abar);
}
double or_adj_loss_grad(void* sv, const or_comp* W, int kind, or_comp* Wb) {
  return OAdj(static_cast<OSolver*>(sv)).loss_grad(W, kind, Wb);
}
// Gradient of the loss after nsteps IFRK2 steps from the solver's state with
// respect to the activity: the field (abar_field, Nx*Ny, may be NULL) and, when
// alpha_t is given (step n then uses the uniform activity alpha_t[n]), its
// per-step sums (abar_t[n]).  Stores every state (small sizes only).  The
// solver's state is left at the final step.  Returns the loss.
double or_adj_gradient(void* sv, int nsteps, int kind, const double* alpha_t, double* abar_field,
                       double* abar_t) {
  OSolver* s = static_cast<OSolver*>(sv);
  OMesh* m = s->mesh;
  const size_t n = m->ns(), P = m->np();
  OAdj A(s);
  std::vector<std::vector<cd>> st(3 * (size_t)(nsteps + 1), std::vector<cd>(n));
  auto save = [&](int k) {
    std::memcpy(st[3 * k + 0].data(), s->wold->spec.data(), sizeof(cd) * n);
    std::memcpy(st[3 * k + 1].data(), s->r1old->spec.data(), sizeof(cd) * n);
    std::memcpy(st[3 * k + 2].data(), s->r2old->spec.data(), sizeof(cd) * n);
  };
  std::vector<double> Ra_keep(s->Ra->phys);
  save(0);
  for (int k = 0; k < nsteps; ++k) {
    if (alpha_t) FldSet(s->Ra->phys.data(), alpha_t[k], (long)P);
    s->step_ifrk2();
    save(k + 1);
  }
  std::vector<cd> g[3], ub[3];
  for (int a = 0; a < 3; ++a) { g[a].assign(n, mk(0, 0)); ub[a].assign(n, mk(0, 0)); }
  const double L = A.loss_grad(st[3 * nsteps].data(), kind, g[0].data());
  std::vector<double> ab(P, 0.0), abk(P);
  for (int k = nsteps - 1; k >= 0; --k) {
    if (alpha_t) FldSet(s->Ra->phys.data(), alpha_t[k], (long)P);
    std::fill(abk.begin(), abk.end(), 0.0);
    const cd* u[3] = {st[3 * k].data(), st[3 * k + 1].data(), st[3 * k + 2].data()};
    const cd* gg[3] = {g[0].data(), g[1].data(), g[2].data()};
    cd* uu[3] = {ub[0].data(), ub[1].data(), ub[2].data()};
    A.step_vjp(u, gg, uu, abk.data());
    for (int a = 0; a < 3; ++a) g[a].swap(ub[a]);
    double sum = 0;
    for (size_t x = 0; x < P; ++x) { ab[x] += abk[x]; sum += abk[x]; }
    if (abar_t) abar_t[k] = sum;
  }
  if (abar_field) std::memcpy(abar_field, ab.data(), sizeof(double) * P);
  s->Ra->phys = Ra_keep;
  return L;
}

// InertialFriction/post_process/postprocess.cu:54-65, same summation order;
// products rounded one by one (no fused multiply-add), as on the host.
__attribute__((optimize("fp-contract=off"))) double or_Etot_func(const double* u, const double* v, int Nx,
                                                                  int Ny, double dx, double dy) {
  double E = 0.0;
  for (int j = 0; j < Ny; j++)
    for (int i = 0; i < Nx; i++) {
      const int index = j * Nx + i;
      E += (u[index] * u[index] + v[index] * v[index]) * 0.5 * dx * dy;
    }
  return E;
}

// post_process/EnergySpec.cu:81-96 over the mesh's wave tables.
__attribute__((optimize("fp-contract=off"))) double or_Espec_func(void* mv, double k, const or_comp* ut,
                                                                   const or_comp* vt, double alpha) {
  const OMesh* m = static_cast<OMesh*>(mv);
  const int Nxh = m->Nxh, Ny = m->Ny;
  double E = 0.0;
  for (int j = 0; j < Ny; j++)
    for (int i = 0; i < Nxh; i++) {
      const int index = j * Nxh + i;
      const double mod = m->kx[index] * m->kx[index] + m->ky[index] * m->ky[index];
      if ((mod < ((k + alpha) * (k + alpha))) && (mod >= (k * k)))
        E += ut[index].x * ut[index].x + ut[index].y * ut[index].y + vt[index].x * vt[index].x +
             vt[index].y * vt[index].y;
    }
  return E / (2.0 * alpha * (Nxh - 1) * 2 * Ny * (Nxh - 1) * 2 * Ny);
}

// The driver loop of EnergySpec.cu:194-204: k = 0, dk, ... while k < kmax.
// Returns the number of shells written to E (at most nmax).
int or_Espec_shells(void* mv, const or_comp* ut, const or_comp* vt, double dk, double kmax, int nmax,
                    double* E) {
  int n = 0;
  for (double k = 0.0; k < kmax && n < nmax; k += dk) E[n++] = or_Espec_func(mv, k, ut, vt, dk);
  return n;
}

void or_set_threads(int n) { g_threads = n; }
int or_get_threads(void) { return nthreads(); }

}  // extern "C"
