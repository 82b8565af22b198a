// Host AddressSanitizer / UndefinedBehaviorSanitizer run of the oracle
// (oracle/qs_oracle.cpp, test infrastructure): every entry point on small
// square, non-square and 3 * 2^a meshes.  Exits non-zero (the sanitizers
// abort on the first finding) or prints "oracle sanitizer run: ok".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "qs_oracle.h"

namespace {
unsigned long long g_state = 12345;
double urand() {   // splitmix64, uniform [-0.5, 0.5)
  unsigned long long z = (g_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * 0x1.0p-53 - 0.5;
}
void fill(std::vector<double>& v, double s) {
  for (auto& x : v) x = s * urand();
}

int run_mesh(int Nx, int Ny) {
  const int Nxh = Nx / 2 + 1;
  const size_t P = (size_t)Nx * Ny, H = (size_t)Nxh * Ny;
  void* m = or_mesh_new(Nx, Ny, 2 * M_PI, 2 * M_PI);
  std::vector<double> kx(H), ky(H), k2(H), cut(H);
  or_tables(m, kx.data(), ky.data(), k2.data(), cut.data());
  std::vector<double> f(P), g(P), Ra(P);
  std::vector<or_comp> F(H), W(H), R1(H), R2(H), N0(H), N1(H), N2(H), B0(H), B1(H), B2(H);
  fill(f, 0.02);
  or_FwdTrans(F.data(), f.data(), m);
  or_BwdTrans(g.data(), F.data(), m);
  for (size_t k = 0; k < P; ++k)
    if (std::fabs(f[k] - g[k]) > 1e-12) { std::printf("round trip %dx%d\n", Nx, Ny); return 1; }
  std::vector<or_comp> in(Nx), out(Nx);
  for (auto& z : in) z = {urand(), urand()};
  or_fft1d(Nx, -1, in.data(), out.data());
  or_fft1d(Nx, +1, out.data(), in.data());
  // fields for the composites and the solver
  void* fld[15];
  for (auto& q : fld) q = or_field_new(m);
  for (int q = 0; q < 3; ++q) {
    fill(f, 0.02);
    or_FwdTrans(or_field_spec(fld[q]), f.data(), m);
  }
  for (size_t k = 0; k < P; ++k) or_field_phys(fld[9])[k] = 1.0 + 0.1 * urand();
  or_curr_func(fld[1], fld[2], fld[0], fld[3], fld[4], fld[5]);
  or_single(fld[6], fld[1], fld[5], 0.1);
  or_cross(fld[7], fld[1], fld[2], fld[10]);
  or_p11func(fld[6], fld[1], fld[5], fld[9], 0.1);
  or_p12func(fld[7], fld[1], fld[2], fld[5], fld[9], 0.1, fld[10], fld[11]);
  or_p21func(fld[8], fld[1], fld[2], fld[5], fld[9], 0.1, fld[10], fld[11]);
  or_wnonl_func(fld[12], fld[0], fld[3], fld[4], fld[1], fld[2], fld[9], fld[5], fld[6], fld[7], fld[8], 0.1,
                0.1, 0.1, fld[10], fld[11]);
  or_r1nonl_func(fld[13], fld[0], fld[3], fld[4], fld[1], fld[2], fld[5], 0.1, fld[10], fld[11]);
  or_r2nonl_func(fld[14], fld[0], fld[3], fld[4], fld[1], fld[2], fld[5], 0.1, fld[10], fld[11]);
  for (auto q : fld) or_field_free(q);
  // whole steps, right-hand side, tables, adjoint, diagnostics
  const double p[10] = {1.0, 1.0, 1.0, 0.1, 1.0, 0.05, 1.0, 1.0, 1e-3, 1.0};
  void* s = or_solver_new(m, p);
  fill(f, 0.02);
  or_FwdTrans(R1.data(), f.data(), m);
  fill(f, 0.02);
  or_FwdTrans(R2.data(), f.data(), m);
  fill(f, 0.1);
  or_FwdTrans(W.data(), f.data(), m);
  for (int meth : {OR_IFE, OR_IFRK2, OR_RK4}) {
    or_solver_set_state(s, W.data(), R1.data(), R2.data());
    or_solver_step(s, meth, 2);
  }
  fill(Ra, 0.1);
  for (auto& a : Ra) a += 1.0;
  or_solver_set_activity(s, Ra.data());
  or_solver_step(s, OR_IFRK2, 1);
  or_solver_get_state(s, B0.data(), B1.data(), B2.data());
  or_solver_rhs(s, B0.data(), B1.data(), B2.data(), N0.data(), N1.data(), N2.data());
  std::vector<double> t6(6 * H);
  or_solver_lin(s, &t6[0], &t6[H], &t6[2 * H], &t6[3 * H], &t6[4 * H], &t6[5 * H]);
  or_adj_B(s, W.data(), g.data());
  or_adj_BT(s, g.data(), B0.data());
  or_adj_F(s, g.data(), B1.data());
  or_adj_FT(s, B1.data(), f.data());
  std::vector<double> abar(P, 0.0);
  or_adj_rhs_vjp(s, W.data(), R1.data(), R2.data(), W.data(), R1.data(), R2.data(), B0.data(), B1.data(),
                 B2.data(), abar.data());
  or_adj_step_vjp(s, W.data(), R1.data(), R2.data(), W.data(), R1.data(), R2.data(), B0.data(), B1.data(),
                  B2.data(), abar.data());
  or_adj_loss_grad(s, W.data(), 1, B0.data());
  or_solver_set_state(s, W.data(), R1.data(), R2.data());
  std::vector<double> at(3, 1.0), abt(3);
  const double L = or_adj_gradient(s, 3, 0, at.data(), abar.data(), abt.data());
  if (!std::isfinite(L)) { std::printf("adjoint loss %dx%d\n", Nx, Ny); return 1; }
  or_solver_free(s);
  or_Etot_func(f.data(), g.data(), Nx, Ny, 0.1, 0.1);
  or_Espec_func(m, 1.0, W.data(), R1.data(), 1.0);
  std::vector<double> E(256);
  or_Espec_shells(m, W.data(), R1.data(), 0.5, 20.0, 256, E.data());
  or_mesh_free(m);
  return 0;
}
}  // namespace

int main() {
  or_set_threads(2);
  const int meshes[][2] = {{16, 16}, {32, 16}, {16, 64}, {48, 48}, {96, 48}, {64, 64}};
  for (const auto& n : meshes)
    if (run_mesh(n[0], n[1])) return 1;
  std::printf("oracle sanitizer run: ok\n");
  return 0;
}
