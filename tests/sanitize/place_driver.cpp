// Host check of the column-job placement (col_job_place in
// nature-control_amd/csrc/qs_fused_dev.hpp): for every fused side, column
// tile width and stage kind, the jobs 0 .. 3 ntiles - 1 map one to one onto
// the (field, tile) pairs, and at 1024^2 (tile width 4) each XCD range lists
// its r jobs first, its vorticity jobs in the slots that end the first round
// of 32 CUs, then the rest.  Built by tests/sanitize/Makefile (host code
// only, under the sanitizers) and run by tests/test_sanitizers.py.
#include <cstdio>
#include <vector>

#include "qs_fused_dev.hpp"

using qs::col_job_place;

static int check_bijection(int side, int tc, bool final_stage) {
  const int nxh = side / 2 + 1, ntiles = (nxh + tc - 1) / tc, nfull = nxh / tc, n = 3 * ntiles;
  const int kcut = (nxh - 1) * 2 / 3 + 1, nkc = (kcut + tc - 1) / tc;
  const int nkeep = nkc < ntiles ? nkc : ntiles;
  const int nk = final_stage ? (nkeep < nfull ? nkeep : nfull) : nfull;
  std::vector<int> seen(n, 0);
  for (int j = 0; j < n; ++j) {
    int f = -1, t = -1;
    if (!col_job_place(j, nxh, tc, nk, f, t)) return j == 0 ? 0 : -1;   // not applicable: falls back
    if (f < 0 || f > 2 || t < 0 || t >= ntiles) {
      std::printf("side %d tc %d final %d: job %d -> (%d, %d) out of range\n", side, tc, final_stage, j, f, t);
      return -1;
    }
    ++seen[3 * t + f];
  }
  for (int i = 0; i < n; ++i)
    if (seen[i] != 1) {
      std::printf("side %d tc %d final %d: (field %d, tile %d) taken %d times\n", side, tc, final_stage, i % 3,
                  i / 3, seen[i]);
      return -1;
    }
  return 1;
}

int main() {
  int checked = 0, bad = 0;
  for (int side : {16, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024})
    for (int tc : {1, 2, 4, 8, 12, 16, 24, 32, 48, 64, 128, 256}) {
      if (tc > side / 2 + 1) continue;
      for (int fin = 0; fin < 2; ++fin) {
        const int r = check_bijection(side, tc, fin != 0);
        if (r < 0) ++bad;
        if (r > 0) ++checked;
      }
    }
  // 1024^2, tile width 4, stage 1: XCD 0 lists r1 (16 jobs), vorticity (16),
  // r2 (16), then the partial tile's vorticity job
  const int nxh = 513;
  for (int k = 0; k < 49; ++k) {
    int f, t;
    col_job_place(k, nxh, 4, 128, f, t);
    const int want_f = k < 16 ? 1 : (k < 32 ? 0 : (k < 48 ? 2 : 0));
    const int want_t = k < 48 ? k % 16 : 128;
    if (f != want_f || t != want_t) {
      std::printf("1024 stage 1: job %d -> (%d, %d), expected (%d, %d)\n", k, f, t, want_f, want_t);
      ++bad;
    }
  }
  // final stage: the vorticity jobs end the first round (slots 22..31 of XCD 0)
  for (int k = 22; k < 32; ++k) {
    int f, t;
    col_job_place(k, nxh, 4, 86, f, t);
    if (f != 0) {
      std::printf("1024 final stage: job %d -> field %d, expected vorticity\n", k, f);
      ++bad;
    }
  }
  std::printf("placement cases %d, failures %d\n", checked, bad);
  if (bad == 0 && checked > 0) std::printf("placement run: ok\n");
  return bad != 0;
}
