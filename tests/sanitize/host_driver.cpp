// Host AddressSanitizer / UndefinedBehaviorSanitizer run of libqsolver's
// host-only code (qs_api.cpp: error strings, Mesh set-up and teardown,
// splitmix64; qs_io.cpp: the reference's CSV text formats and their parser),
// built without device code and run without a GPU: a call that would need one
// must fail with a status, not crash.  Exits non-zero on a wrong result (the sanitizers abort on their
// first finding) or prints "host sanitizer run: ok".
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "qsolver.h"

namespace {
int fails = 0;
void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAILED: %s (last error: %s)\n", what, qs_last_error());
    ++fails;
  }
}
void write_text(const std::string& path, const std::string& text) {
  std::ofstream o(path);
  o << text;
}
}  // namespace

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  for (int s = 0; s <= 8; ++s) expect(qs_status_string(s) != nullptr, "status string");
  // splitmix64 stream
  std::vector<double> u(1000);
  expect(qs_synth_uniform(1, -0.01, 0.01, u.size(), u.data()) == QS_OK, "synth");
  for (double x : u) expect(x >= -0.01 && x < 0.01, "synth range");
  expect(qs_synth_uniform(1, 0, 1, 0, nullptr) == QS_OK, "synth n=0");
  expect(qs_synth_uniform(1, 0, 1, 5, nullptr) == QS_ERR_INVALID_ARG, "synth null");
  // CSV round trip through the reference's text format
  const int Nx = 12, Ny = 7;
  std::vector<double> f(Nx * Ny), g(Nx * Ny, -1.0);
  for (int k = 0; k < Nx * Ny; ++k) f[k] = std::sin(0.37 * k) * std::pow(10.0, (k % 9) - 4);
  const std::string csv = dir + "/f.csv";
  expect(qs_field_visual(f.data(), csv.c_str(), Nx, Ny) == QS_OK, "field_visual");
  Mesh m{};
  m.Nx = Nx; m.Ny = Ny; m.Nxh = Nx / 2 + 1;
  expect(qs_file_init(csv.c_str(), g.data(), &m) == QS_OK, "file_init");
  for (int k = 0; k < Nx * Ny; ++k) expect(std::fabs(g[k] - f[k]) <= 5e-6 * std::fabs(f[k]), "6 digits");
  // malformed files: wrong row length, wrong row count, missing file, junk tokens
  write_text(dir + "/short.csv", "1,2,3,\n");
  expect(qs_file_init((dir + "/short.csv").c_str(), g.data(), &m) == QS_ERR_INVALID_ARG, "short row");
  std::string rows;
  for (int j = 0; j < Ny + 2; ++j) {
    for (int i = 0; i < Nx; ++i) rows += "1,";
    rows += "\n";
  }
  write_text(dir + "/long.csv", rows);
  expect(qs_file_init((dir + "/long.csv").c_str(), g.data(), &m) == QS_ERR_INVALID_ARG, "extra rows");
  expect(qs_file_init((dir + "/missing.csv").c_str(), g.data(), &m) == QS_ERR_IO, "missing file");
  std::string junk;
  for (int j = 0; j < Ny; ++j) {
    for (int i = 0; i < Nx; ++i) junk += (i % 3 == 0 ? "x," : (i % 3 == 1 ? "," : "2.5e-3,"));
    junk += "\n";
  }
  write_text(dir + "/junk.csv", junk);
  expect(qs_file_init((dir + "/junk.csv").c_str(), g.data(), &m) == QS_OK, "junk tokens parse as 0");
  expect(g[0] == 0.0 && g[1] == 0.0 && g[2] == 2.5e-3, "junk values");
  expect(qs_file_init(nullptr, g.data(), &m) == QS_ERR_INVALID_ARG, "null name");
  expect(qs_file_init(csv.c_str(), g.data(), nullptr) == QS_ERR_INVALID_ARG, "null mesh");
  std::vector<Qcomp> z(4 * 3);
  for (auto& c : z) c = {0.5, -0.25};
  expect(qs_wave_visual(f.data(), (dir + "/w.csv").c_str(), 4, 3) == QS_OK, "wave_visual");
  expect(qs_spec_visual(z.data(), (dir + "/s.csv").c_str(), 4, 3) == QS_OK, "spec_visual");
  expect(qs_coord(0.1, 0.2, 5, 4, (dir + "/x.csv").c_str(), (dir + "/y.csv").c_str()) == QS_OK, "coord");
  expect(qs_field_visual(nullptr, csv.c_str(), Nx, Ny) == QS_ERR_INVALID_ARG, "null field");
  expect(qs_field_visual(f.data(), "/nonexistent-dir/f.csv", Nx, Ny) == QS_ERR_IO, "unwritable");
  // teardown of nothing
  expect(qs_mesh_destroy(nullptr) == QS_OK, "mesh destroy null");
  expect(qs_field_destroy(nullptr) == QS_OK, "field destroy null");
  expect(qs_field_create(nullptr, nullptr) != QS_OK, "field create null");
  Mesh* mp = nullptr;
  expect(qs_mesh_create(100, 64, 1.0, 1.0, 16, &mp) == QS_ERR_UNSUPPORTED_SIZE && mp == nullptr, "odd side");
  expect(qs_mesh_create(64, 64, 1.0, 1.0, 16, nullptr) != QS_OK, "mesh out null");
  expect(qs_mesh_create(64, 64, -1.0, 1.0, 16, &mp) == QS_ERR_INVALID_ARG && mp == nullptr, "negative length");
  // Mesh set-up (host tables, twiddles) up to the first device allocation:
  // without a GPU it must fail cleanly and free what it built
  const int st = qs_mesh_create(768, 48, 66 * M_PI, 1.0, 16, &mp);
  expect(st == QS_OK || st == QS_ERR_OOM || st == QS_ERR_HIP, "mesh create status");
  if (st == QS_OK) expect(qs_mesh_destroy(mp) == QS_OK, "mesh destroy");
  else expect(mp == nullptr, "no mesh on failure");
  expect(std::strlen(qs_last_error()) > 0, "last error");
  if (fails) return 1;
  std::printf("host sanitizer run: ok\n");
  return 0;
}
