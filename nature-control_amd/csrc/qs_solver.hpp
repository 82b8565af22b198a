// qs_solver.hpp -- the solver object shared by the forward (qs_solver.hip) and
// the discrete adjoint (qs_adjoint_host.hip).
#pragma once

#include <utility>
#include <vector>

#include "qs_fused.hpp"

namespace qs {

struct AdjGroup;

// Device memory of the checkpointed discrete adjoint of one run.
struct AdjState {
  int kcap = 0;          // segment length the stage slots are sized for
  int ncap = 0;          // horizon the per-step arrays are sized for
  int nckpt = 0;         // checkpoints held
  int nbuf = 0;          // stage-slot buffers (2: a segment is recomputed while another is reversed)
  std::vector<double2*> ckpt;                // 3 per checkpoint, [Ny][pitch]
  std::vector<double2*> slots;               // [buffer][stage][7], (2 kcap + 1) stages per buffer
  double* w00s = nullptr;                    // W(0,0) per recorded stage, [buffer][2 kcap + 1]
  double2* amix[7] = {};                     // the reverse sweep's own mixed arrays (nbuf == 2)
  unsigned long recorded_by = 0;             // id of the group whose last segment the slots still hold
  double2* lam_[3] = {};                     // adjoint state
  double2* acc_[3] = {};
  double* abar = nullptr;                    // [Ny][Nx]
  double* arow = nullptr;                    // [ncap][Ny]
  double* w00bar = nullptr;                  // [Ny]
  void* arena = nullptr;                     // the one allocation holding all of the above
  size_t bytes = 0;                          // device bytes held
  AdjGroup* solo = nullptr;                  // the group of this run alone (qs_adj_*)
};

}  // namespace qs

struct qs_solver {
  Mesh* mesh = nullptr;
  qs_params p{};
  qs::Coef coef;
  int pitch = 0;
  // state [Ny][pitch]
  double2* old_[3] = {nullptr, nullptr, nullptr};   // w, r1, r2
  double2* new_[3] = {nullptr, nullptr, nullptr};   // stage-1 partial sums
  double2* mix[7] = {};                              // mixed arrays [Ny][pitch]
  double* Ra = nullptr;                              // activity field or null
  double* w00 = nullptr;                             // W(0,0) of the stage
  int* flag = nullptr;                               // non-finite flag (device)
  int* hflag = nullptr;                              // its host copy (pinned), refreshed after each fused call
  double* ify = nullptr;                             // integrating-factor tables (IFac, ifac_len doubles)
  qs::RunDesc* d_desc = nullptr;
  int2* d_work = nullptr;
  int nrow = 0, ncol = 0;
  bool mixed_valid = false;
  // unfused path (reference call granularity, reference layout)
  Field* F[21] = {};
  double* IFt[4] = {};   // IFw, IFwh, IFq, IFqh
  Qcomp* uold[3] = {};
  Qcomp* unew[3] = {};
  bool unfused_ready = false;
  // per-kernel HIP-event timing (qs_solver_profile)
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
  // discrete adjoint (qs_adj_*): checkpoints, recorded stages, adjoint state
  qs::AdjState* adj = nullptr;
};

namespace qs {
// Descriptor of a solver's run with the solver's own arrays (the plain step).
RunDesc make_desc(qs_solver* S);
}  // namespace qs

namespace qs {
// Adjoint of a group of runs of one mesh size (qs_adjoint_host.hip).
AdjGroup* adj_group_new(const std::vector<qs_solver*>& runs);
void adj_group_free(AdjGroup* G);
int adj_group_forward(AdjGroup* G, int nsteps, int k, const double* alpha_t);
int adj_group_backward(AdjGroup* G);
int adj_group_steps(const AdjGroup* G);
void adj_state_free(qs_solver* S);
size_t adj_group_bytes(const AdjGroup* G);
int adj_run_seed(qs_solver* S, int kind, double* loss);
}  // namespace qs
