// qs_io.cpp -- snapshot and restart I/O of the reference drivers:
//  - CSV fields: field_visual (BasicUtils/BasicUtils.cuh:38-50), coord (:52-67),
//    wave_visual / spec_visual (InertialFriction/post_process/EnergySpec.cu:8-34)
//    and file_init (InertialFriction/friction/friction.cu:62-97);
//  - the solver-level blocks built from them: the CSV snapshot every 1000
//    steps (friction.cu:262-279) and the CSV initialisation (friction.cu:163-185);
//  - exact binary checkpoints of a solver's state and activity (no reference
//    counterpart: the reference restarts only from its 6-digit CSVs).
// CSV text goes through the C++ streams with their default formatting (6
// significant digits), as the reference writes and reads it.
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "qs_solver.hpp"

namespace qs {
namespace {

template <class F>
int write_rows(const char* name, int nrow, int ncol, F&& put) {
  if (!name) return fail(QS_ERR_INVALID_ARG, "null file name");
  std::ofstream out(name);
  if (!out) return fail(QS_ERR_IO, "cannot open %s for writing", name);
  for (int j = 0; j < nrow; ++j) {
    for (int i = 0; i < ncol; ++i) put(out, j, i);
    out << std::endl;
  }
  if (!out) return fail(QS_ERR_IO, "write error on %s", name);
  return QS_OK;
}

// One CSV line split as friction.cu:79-84 splits it: getline(ss, str, ',')
// yields every comma-terminated token, and the text after the last comma
// only when it is not empty.
void split_line(const std::string& line, std::vector<std::string>& tok) {
  tok.clear();
  size_t a = 0;
  while (a <= line.size()) {
    const size_t b = line.find(',', a);
    if (b == std::string::npos) {
      if (a < line.size()) tok.emplace_back(line, a, std::string::npos);
      break;
    }
    tok.emplace_back(line, a, b - a);
    a = b + 1;
  }
}

int read_csv(const char* filename, int Nx, int Ny, double* f) {
  if (!filename || !f) return fail(QS_ERR_INVALID_ARG, "qs_file_init: null argument");
  std::ifstream in(filename);
  if (!in.good()) return fail(QS_ERR_IO, "qs_file_init: no such file: %s", filename);
  std::vector<double> vals;
  vals.reserve((size_t)Nx * Ny);
  std::vector<std::string> tok;
  std::string line;
  std::istringstream iss;
  int rows = 0;
  while (std::getline(in, line)) {
    split_line(line, tok);
    if (rows < Ny && (int)tok.size() != Nx)
      return fail(QS_ERR_INVALID_ARG, "qs_file_init: %s row %d has %zu values, the mesh has Nx=%d", filename, rows,
                  tok.size(), Nx);
    if (rows < Ny)
      for (const auto& t : tok) {   // string2num<double> (friction.cu:54-60): 0 when nothing converts
        iss.clear();
        iss.str(t);
        double v = 0.0;
        iss >> v;
        vals.push_back(v);
      }
    ++rows;
  }
  if (rows != Ny) return fail(QS_ERR_INVALID_ARG, "qs_file_init: %s has %d rows, the mesh has Ny=%d", filename, rows, Ny);
  std::memcpy(f, vals.data(), sizeof(double) * vals.size());
  return QS_OK;
}

std::string snap_name(const char* prefix, int step, const char* what) {
  return std::string(prefix ? prefix : "") + std::to_string(step) + what;
}

// Device temporaries of the solver-level I/O calls.
struct Scratch {
  std::vector<void*> p;
  ~Scratch() {
    for (void* q : p) (void)hipFree(q);
  }
  template <class T>
  int get(T** out, size_t bytes) {
    *out = nullptr;
    if (hipMalloc((void**)out, bytes) != hipSuccess) return fail(QS_ERR_OOM, "I/O scratch of %zu bytes", bytes);
    p.push_back(*out);
    return QS_OK;
  }
};

int state_to_packed(qs_solver* S, int c, Qcomp* dst) {
  const Mesh* m = S->mesh;
  const size_t row = sizeof(Qcomp) * m->Nxh;
  QS_HIP(hipMemcpy2DAsync(dst, row, S->old_[c], sizeof(double2) * S->pitch, row, m->Ny, hipMemcpyDeviceToDevice,
                          stream_of(m)));
  return QS_OK;
}
int packed_to_state(qs_solver* S, int c, const Qcomp* src) {
  const Mesh* m = S->mesh;
  const size_t row = sizeof(Qcomp) * m->Nxh;
  QS_HIP(hipMemcpy2DAsync(S->old_[c], sizeof(double2) * S->pitch, src, row, row, m->Ny, hipMemcpyDeviceToDevice,
                          stream_of(m)));
  S->mixed_valid = false;
  return QS_OK;
}

// ---- binary checkpoint -----------------------------------------------------
constexpr char kMagic[8] = {'Q', 'S', 'S', 'N', 'A', 'P', '0', '1'};
struct SnapHeader {
  char magic[8];
  int32_t Nx, Ny;
  int32_t has_activity, reserved;
  double Lx, Ly;
  qs_params params;
  int64_t step;
  uint64_t payload_bytes;
};

uint64_t fnv1a(const void* p, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
  const unsigned char* b = (const unsigned char*)p;
  for (size_t k = 0; k < n; ++k) h = (h ^ b[k]) * 0x100000001b3ull;
  return h;
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" {

int qs_field_visual(const Qreal* f, const char* name, int Nx, int Ny) {
  if (!f || Nx <= 0 || Ny <= 0) return fail(QS_ERR_INVALID_ARG, "qs_field_visual: bad argument");
  return write_rows(name, Ny, Nx, [&](std::ofstream& o, int j, int i) { o << f[(size_t)j * Nx + i] << ","; });
}

int qs_wave_visual(const Qreal* f, const char* name, int Nxh, int Ny) {
  if (!f || Nxh <= 0 || Ny <= 0) return fail(QS_ERR_INVALID_ARG, "qs_wave_visual: bad argument");
  return write_rows(name, Ny, Nxh, [&](std::ofstream& o, int j, int i) { o << f[(size_t)j * Nxh + i] << ","; });
}

int qs_spec_visual(const Qcomp* f, const char* name, int Nxh, int Ny) {
  if (!f || Nxh <= 0 || Ny <= 0) return fail(QS_ERR_INVALID_ARG, "qs_spec_visual: bad argument");
  return write_rows(name, Ny, Nxh, [&](std::ofstream& o, int j, int i) {
    const Qcomp z = f[(size_t)j * Nxh + i];
    o << "(" << z.x << "," << z.y << ")" << ",";
  });
}

int qs_coord(Qreal dx, Qreal dy, int Nx, int Ny, const char* xname, const char* yname) {
  if (Nx <= 0 || Ny <= 0) return fail(QS_ERR_INVALID_ARG, "qs_coord: bad size");
  // single-precision coordinates, as BasicUtils.cuh:57-60 writes them
  QS_TRY(write_rows(xname ? xname : "x.csv", Ny, Nx, [&](std::ofstream& o, int, int i) {
    const float x = dx * i;
    o << x << ",";
  }));
  return write_rows(yname ? yname : "y.csv", Ny, Nx, [&](std::ofstream& o, int j, int) {
    const float y = dy * j;
    o << y << ",";
  });
}

int qs_file_init(const char* filename, Qreal* f, const Mesh* m) {
  if (!m) return fail(QS_ERR_INVALID_ARG, "qs_file_init: null mesh");
  return read_csv(filename, m->Nx, m->Ny, f);
}

int qs_solver_snapshot_csv(qs_solver* S, const char* prefix, int step) {
  if (!S) return fail(QS_ERR_INVALID_ARG, "null solver");
  Mesh* m = S->mesh;
  const hipStream_t st = stream_of(m);
  const size_t P = (size_t)m->Nx * m->Ny;
  Scratch tmp;
  Qcomp* spec;
  Qreal* ph[4];
  QS_TRY(tmp.get(&spec, (size_t)m->specsize));
  for (auto& q : ph) QS_TRY(tmp.get(&q, sizeof(double) * P));
  std::vector<double> h(P);
  const char* names[4] = {"w.csv", "r1.csv", "r2.csv", "S.csv"};
  for (int c = 0; c < 3; ++c) {
    // BwdTrans of the state spectrum itself: its column-0 symmetrisation
    // (QFldfuncs.cuh:50-55) lands in the state, as in friction.cu:265-275
    QS_TRY(state_to_packed(S, c, spec));
    QS_TRY(qs_BwdTrans(ph[c], spec, m));
    QS_TRY(packed_to_state(S, c, spec));
    QS_HIP(hipMemcpyAsync(h.data(), ph[c], sizeof(double) * P, hipMemcpyDeviceToHost, st));
    QS_HIP(hipStreamSynchronize(st));
    QS_TRY(qs_field_visual(h.data(), snap_name(prefix, step, names[c]).c_str(), m->Nx, m->Ny));
  }
  QS_TRY(qs_S_func(ph[3], ph[1], ph[2], m));
  QS_HIP(hipMemcpyAsync(h.data(), ph[3], sizeof(double) * P, hipMemcpyDeviceToHost, st));
  QS_HIP(hipStreamSynchronize(st));
  QS_TRY(qs_field_visual(h.data(), snap_name(prefix, step, names[3]).c_str(), m->Nx, m->Ny));
  if (std::isnan(h[0])) return fail(QS_ERR_NONFINITE, "snapshot %d: S is NaN", step);   // friction.cu:279
  return QS_OK;
}

int qs_solver_init_csv(qs_solver* S, const char* w_csv, const char* r1_csv, const char* r2_csv) {
  if (!S || !w_csv || !r1_csv || !r2_csv) return fail(QS_ERR_INVALID_ARG, "qs_solver_init_csv: null argument");
  Mesh* m = S->mesh;
  const hipStream_t st = stream_of(m);
  const size_t P = (size_t)m->Nx * m->Ny;
  Scratch tmp;
  Qcomp* spec[3];
  Qreal* ph;
  for (auto& q : spec) QS_TRY(tmp.get(&q, (size_t)m->specsize));
  QS_TRY(tmp.get(&ph, sizeof(double) * P));
  std::vector<double> h(P);
  const char* files[3] = {w_csv, r1_csv, r2_csv};
  for (int c = 0; c < 3; ++c) {
    // friction.cu:164-167: file_init, FwdTrans, then the BwdTrans round trip
    // that symmetrises column 0 of the spectrum the run starts from
    QS_TRY(read_csv(files[c], m->Nx, m->Ny, h.data()));
    QS_HIP(hipMemcpyAsync(ph, h.data(), sizeof(double) * P, hipMemcpyHostToDevice, st));
    QS_TRY(qs_FwdTrans(spec[c], ph, m));
    QS_TRY(qs_BwdTrans(ph, spec[c], m));
    QS_HIP(hipStreamSynchronize(st));   // h and ph are reused
  }
  QS_TRY(qs_solver_set_state(S, spec[0], spec[1], spec[2]));
  QS_HIP(hipStreamSynchronize(st));     // the temporaries are freed on return
  return QS_OK;
}

int qs_solver_save(qs_solver* S, const char* path, int64_t step) {
  if (!S || !path) return fail(QS_ERR_INVALID_ARG, "qs_solver_save: null argument");
  const Mesh* m = S->mesh;
  const hipStream_t st = stream_of(m);
  const size_t H = (size_t)m->Nxh * m->Ny, P = (size_t)m->Nx * m->Ny;
  std::vector<Qcomp> spec(3 * H);
  std::vector<double> act(S->Ra ? P : 0);
  const size_t row = sizeof(Qcomp) * m->Nxh;
  for (int c = 0; c < 3; ++c)
    QS_HIP(hipMemcpy2DAsync(spec.data() + c * H, row, S->old_[c], sizeof(double2) * S->pitch, row, m->Ny,
                            hipMemcpyDeviceToHost, st));
  if (S->Ra) QS_HIP(hipMemcpyAsync(act.data(), S->Ra, sizeof(double) * P, hipMemcpyDeviceToHost, st));
  QS_HIP(hipStreamSynchronize(st));
  SnapHeader hd{};
  std::memcpy(hd.magic, kMagic, 8);
  hd.Nx = m->Nx;
  hd.Ny = m->Ny;
  hd.has_activity = S->Ra ? 1 : 0;
  hd.Lx = m->Lx;
  hd.Ly = m->Ly;
  hd.params = S->p;
  hd.step = step;
  hd.payload_bytes = sizeof(Qcomp) * spec.size() + sizeof(double) * act.size();
  uint64_t hash = fnv1a(spec.data(), sizeof(Qcomp) * spec.size());
  hash = fnv1a(act.data(), sizeof(double) * act.size(), hash);
  std::ofstream out(path, std::ios::binary);
  if (!out) return fail(QS_ERR_IO, "qs_solver_save: cannot open %s", path);
  out.write((const char*)&hd, sizeof(hd));
  out.write((const char*)spec.data(), (std::streamsize)(sizeof(Qcomp) * spec.size()));
  out.write((const char*)act.data(), (std::streamsize)(sizeof(double) * act.size()));
  out.write((const char*)&hash, sizeof(hash));
  if (!out) return fail(QS_ERR_IO, "qs_solver_save: write error on %s", path);
  return QS_OK;
}

int qs_solver_load(qs_solver* S, const char* path, int64_t* step, qs_params* params) {
  if (!S || !path) return fail(QS_ERR_INVALID_ARG, "qs_solver_load: null argument");
  const Mesh* m = S->mesh;
  const hipStream_t st = stream_of(m);
  std::ifstream in(path, std::ios::binary);
  if (!in) return fail(QS_ERR_IO, "qs_solver_load: cannot open %s", path);
  SnapHeader hd{};
  in.read((char*)&hd, sizeof(hd));
  if (!in || std::memcmp(hd.magic, kMagic, 8) != 0) return fail(QS_ERR_IO, "qs_solver_load: %s is not a snapshot", path);
  if (hd.Nx != m->Nx || hd.Ny != m->Ny || hd.Lx != m->Lx || hd.Ly != m->Ly)
    return fail(QS_ERR_INVALID_ARG, "qs_solver_load: %s holds a %dx%d mesh (Lx=%g, Ly=%g), the solver's is %dx%d "
                "(Lx=%g, Ly=%g)", path, hd.Nx, hd.Ny, hd.Lx, hd.Ly, m->Nx, m->Ny, m->Lx, m->Ly);
  const size_t H = (size_t)m->Nxh * m->Ny, P = (size_t)m->Nx * m->Ny;
  const size_t want = sizeof(Qcomp) * 3 * H + (hd.has_activity ? sizeof(double) * P : 0);
  if (hd.payload_bytes != want) return fail(QS_ERR_IO, "qs_solver_load: %s has a bad payload size", path);
  std::vector<Qcomp> spec(3 * H);
  std::vector<double> act(hd.has_activity ? P : 0);
  uint64_t hash = 0;
  in.read((char*)spec.data(), (std::streamsize)(sizeof(Qcomp) * spec.size()));
  in.read((char*)act.data(), (std::streamsize)(sizeof(double) * act.size()));
  in.read((char*)&hash, sizeof(hash));
  if (!in) return fail(QS_ERR_IO, "qs_solver_load: %s is truncated", path);
  uint64_t h = fnv1a(spec.data(), sizeof(Qcomp) * spec.size());
  h = fnv1a(act.data(), sizeof(double) * act.size(), h);
  if (h != hash) return fail(QS_ERR_IO, "qs_solver_load: checksum mismatch in %s", path);
  const size_t row = sizeof(Qcomp) * m->Nxh;
  for (int c = 0; c < 3; ++c)
    QS_HIP(hipMemcpy2DAsync(S->old_[c], sizeof(double2) * S->pitch, spec.data() + c * H, row, row, m->Ny,
                            hipMemcpyHostToDevice, st));
  S->mixed_valid = false;
  if (hd.has_activity) {
    Scratch tmp;
    double* d;
    QS_TRY(tmp.get(&d, sizeof(double) * P));
    QS_HIP(hipMemcpyAsync(d, act.data(), sizeof(double) * P, hipMemcpyHostToDevice, st));
    QS_TRY(qs_solver_set_activity(S, d));
    QS_HIP(hipStreamSynchronize(st));
  } else {
    QS_TRY(qs_solver_set_activity(S, nullptr));
  }
  QS_HIP(hipStreamSynchronize(st));   // the host copies are freed on return
  if (step) *step = hd.step;
  if (params) *params = hd.params;
  return QS_OK;
}

}  // extern "C"
