// qs_api.cpp -- Mesh / Field lifecycle, error reporting and host helpers.
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "qs_internal.hpp"

namespace qs {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return status;
}

// W_n^m = exp(-2 pi i m/n) in long double (64-bit mantissa), rounded once.
void host_twiddles(int n, double2* out) {
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < n; ++m) {
    const long double a = two_pi * (long double)m / (long double)n;
    out[m] = make_double2((double)cosl(a), -(double)sinl(a));
  }
}

}  // namespace qs

using namespace qs;

extern "C" {

const char* qs_last_error(void) { return g_err; }

const char* qs_status_string(int s) {
  switch (s) {
    case QS_OK: return "ok";
    case QS_ERR_INVALID_ARG: return "invalid argument";
    case QS_ERR_UNSUPPORTED_SIZE: return "unsupported size";
    case QS_ERR_HIP: return "HIP runtime error";
    case QS_ERR_NONFINITE: return "non-finite value in state";
    case QS_ERR_OOM: return "out of device memory";
    case QS_ERR_STATE: return "invalid object state";
    case QS_ERR_IO: return "file I/O error";
    default: return "unknown status";
  }
}

// Basic/Mesh.cuh:61-107 (tables from waveinit :6-21 and cutoff_func :23-46).
int qs_mesh_create(int Nx, int Ny, Qreal Lx, Qreal Ly, int BSZ, Mesh** out) {
  if (!out) return fail(QS_ERR_INVALID_ARG, "qs_mesh_create: null out");
  *out = nullptr;
  if (!supported_side(Nx) || !supported_side(Ny))
    return fail(QS_ERR_UNSUPPORTED_SIZE,
                "qs_mesh_create: Nx=%d Ny=%d (2^a in [16,2048] or 3*2^a in [48,1536])", Nx, Ny);
  if (!(Lx > 0) || !(Ly > 0)) return fail(QS_ERR_INVALID_ARG, "qs_mesh_create: Lx, Ly must be > 0");
  Mesh* m = (Mesh*)std::calloc(1, sizeof(Mesh));
  if (!m) return fail(QS_ERR_OOM, "host allocation");
  m->priv = new qs_mesh_priv();
  m->Nx = Nx; m->Ny = Ny; m->Nxh = Nx / 2 + 1; m->BSZ = BSZ;
  m->Lx = Lx; m->Ly = Ly; m->dx = Lx / Nx; m->dy = Ly / Ny;
  m->specsize = m->Nxh * Ny * (int)sizeof(Qcomp);
  m->physsize = Nx * Ny * (int)sizeof(Qreal);
  m->wavesize = m->Nxh * Ny * (int)sizeof(Qreal);
  const int Nxh = m->Nxh;
  std::vector<double> kx((size_t)Nxh * Ny), ky(kx.size()), k2(kx.size()), cut(kx.size());
  for (int i = 0; i < Nxh; i++)
    for (int j = 0; j < Ny; j++) {
      const int index = i + j * Nxh;
      ky[index] = (j < Ny / 2 + 1) ? 2 * M_PI / Ly * j : 2 * M_PI / Ly * (j - Ny);
      kx[index] = 2 * M_PI / Lx * i;
      k2[index] = kx[index] * kx[index] + ky[index] * ky[index];
      cut[index] = ((i < ((Nxh - 1) * 2 / 3 + 1)) && ((j < Ny / 3 + 1) || ((Ny - j) < Ny / 3))) ? 1.0 : 0.0;
    }
  std::vector<double> kx1(Nxh), ky1(Ny);
  for (int i = 0; i < Nxh; ++i) kx1[i] = kx[i];
  for (int j = 0; j < Ny; ++j) ky1[j] = ky[(size_t)j * Nxh];
  std::vector<double2> twx(Nx), twy(Ny), tw1024(1024), tw768(768);
  host_twiddles(Nx, twx.data());
  host_twiddles(Ny, twy.data());
  host_twiddles(1024, tw1024.data());
  host_twiddles(768, tw768.data());

  auto cleanup = [&]() { qs_mesh_destroy(m); };
  {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
      m->priv->ncu = ncu;
  }
#define MALLOC(ptr, bytes)                                                   \
  if (hipMalloc((void**)&(ptr), (bytes)) != hipSuccess) {                   \
    cleanup();                                                               \
    return fail(QS_ERR_OOM, "qs_mesh_create: hipMalloc of %zu bytes", (size_t)(bytes)); \
  }
  const size_t ws = sizeof(double) * kx.size();
  MALLOC(m->spec, sizeof(Qcomp) * kx.size());
  MALLOC(m->phys, sizeof(double) * (size_t)Nx * Ny);
  MALLOC(m->kx, ws);
  MALLOC(m->ky, ws);
  MALLOC(m->k_squared, ws);
  MALLOC(m->cutoff, ws);
  MALLOC(m->priv->twx, sizeof(double2) * Nx);
  MALLOC(m->priv->twy, sizeof(double2) * Ny);
  MALLOC(m->priv->kx1, sizeof(double) * Nxh);
  MALLOC(m->priv->ky1, sizeof(double) * Ny);
  MALLOC(m->priv->tw1024, sizeof(double2) * 1024);
  MALLOC(m->priv->tw768, sizeof(double2) * 768);
#undef MALLOC
  struct { void* d; const void* h; size_t b; } cp[] = {
      {m->kx, kx.data(), ws}, {m->ky, ky.data(), ws}, {m->k_squared, k2.data(), ws},
      {m->cutoff, cut.data(), ws}, {m->priv->twx, twx.data(), sizeof(double2) * Nx},
      {m->priv->twy, twy.data(), sizeof(double2) * Ny}, {m->priv->kx1, kx1.data(), sizeof(double) * Nxh},
      {m->priv->ky1, ky1.data(), sizeof(double) * Ny},
      {m->priv->tw1024, tw1024.data(), sizeof(double2) * 1024},
      {m->priv->tw768, tw768.data(), sizeof(double2) * 768}};
  for (auto& c : cp) {
    hipError_t e = hipMemcpy(c.d, c.h, c.b, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      cleanup();
      return fail(QS_ERR_HIP, "qs_mesh_create: hipMemcpy -> %s", hipGetErrorString(e));
    }
  }
  *out = m;
  return QS_OK;
}

int qs_mesh_destroy(Mesh* m) {
  if (!m) return QS_OK;
  void* ptrs[] = {m->spec, m->phys, m->kx, m->ky, m->k_squared, m->cutoff,
                  m->priv ? m->priv->twx : nullptr, m->priv ? m->priv->twy : nullptr,
                  m->priv ? m->priv->kx1 : nullptr, m->priv ? m->priv->ky1 : nullptr,
                  m->priv ? (void*)m->priv->tw1024 : nullptr, m->priv ? (void*)m->priv->tw768 : nullptr,
                  m->priv ? (void*)m->priv->diag : nullptr};
  if (m->priv && m->priv->diag) (void)hipStreamSynchronize(m->priv->stream);
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete m->priv;
  std::free(m);
  return QS_OK;
}

int qs_mesh_set_stream(Mesh* m, void* stream) {
  QS_TRY(check_mesh(m));
  m->priv->stream = (hipStream_t)stream;
  return QS_OK;
}

int qs_mesh_synchronize(const Mesh* m) {
  QS_TRY(check_mesh(m));
  QS_HIP(hipStreamSynchronize(stream_of(m)));
  return QS_OK;
}

// Basic/Mesh.cuh:120-128
int qs_field_create(Mesh* m, Field** out) {
  if (!out) return fail(QS_ERR_INVALID_ARG, "qs_field_create: null out");
  *out = nullptr;
  QS_TRY(check_mesh(m));
  Field* f = (Field*)std::calloc(1, sizeof(Field));
  if (!f) return fail(QS_ERR_OOM, "host allocation");
  f->mesh = m;
  if (hipMalloc((void**)&f->spec, (size_t)m->specsize) != hipSuccess ||
      hipMalloc((void**)&f->phys, (size_t)m->physsize) != hipSuccess) {
    qs_field_destroy(f);
    return fail(QS_ERR_OOM, "qs_field_create: hipMalloc");
  }
  *out = f;
  return QS_OK;
}

int qs_field_destroy(Field* f) {
  if (!f) return QS_OK;
  if (f->spec) (void)hipFree(f->spec);
  if (f->phys) (void)hipFree(f->phys);
  std::free(f);
  return QS_OK;
}

int qs_synth_uniform(uint64_t seed, Qreal lo, Qreal hi, size_t n, Qreal* out) {
  if (!out && n) return fail(QS_ERR_INVALID_ARG, "qs_synth_uniform: null out");
  uint64_t state = seed;
  for (size_t k = 0; k < n; ++k) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    out[k] = lo + (hi - lo) * u;
  }
  return QS_OK;
}

}  // extern "C"
