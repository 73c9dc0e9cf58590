// Host-only check of the K offset tables' pad: replays the GEMM gather's K
// loop (prologue and main loop, both pipelines, every split of a split-K
// launch) for every stage depth the tile shapes give, and reports any table
// index at or past K + kKTablePad.  Built and run by tests/test_tiles_host.py.
#include <cstdio>
#include <initializer_list>

#include "../../quasistrided.jl_amd/csrc/qs_tiles.h"

using namespace qs;

static long long worst = -1;   // largest index - K seen
static int failures = 0;

// A fetch of one stage's K offsets starting at `k`: lanes read k .. k + bk - 1.
static void fetch(long long k, int bk, long long K, const char *what, int rs, int pl, int bm, int bn, int akf, int bkf) {
  const long long hi = k + bk - 1;
  if (hi - K > worst) worst = hi - K;
  if (hi >= K + kKTablePad) {
    if (failures++ < 10)
      std::printf("OVERRUN %s: rsize %d planes %d tile %dx%d akf %d bkf %d bk %d K %lld reads index %lld (table %lld)\n",
                  what, rs, pl, bm, bn, akf, bkf, bk, K, hi, K + kKTablePad);
  }
}

int main() {
  const int tiles[6][2] = {{128, 128}, {64, 64}, {128, 64}, {64, 128}, {32, 64}, {32, 32}};
  int configs = 0;
  for (int rs : {4, 8})
    for (int pl : {1, 2})
      for (auto &t : tiles)
        for (int akf = 0; akf < 2; ++akf)
          for (int bkf = 0; bkf < 2; ++bkf) {
            const int bm = t[0], bn = t[1];
            const int bk = pick_bk_for(rs, pl, bm, bn, akf, bkf);
            ++configs;
            for (long long K = 1; K <= 1200; ++K)
              for (int splits = 1; splits <= 4; ++splits) {
                // Split-K chunks are whole stage pairs (qs_api.cpp configure_gemm).
                const long long chunk = ((K + splits - 1) / splits + 2 * bk - 1) / (2 * bk) * (2 * bk);
                for (long long kbeg = 0; kbeg < K; kbeg += chunk) {
                  const long long kend = kbeg + chunk < K ? kbeg + chunk : K;
                  // Two-stage pipeline (NS = 2): prologue, then NS steps per trip.
                  const int NS = 2;
                  fetch(kbeg, bk, K, "deep prologue", rs, pl, bm, bn, akf, bkf);
                  for (int S = 0; S < NS; ++S) {
                    fetch(kbeg + (long long)(S + NS - 1) * bk, bk, K, "deep prologue lane", rs, pl, bm, bn, akf, bkf);
                    fetch(kbeg + (long long)(S + 1) * bk, bk, K, "deep prologue uniform", rs, pl, bm, bn, akf, bkf);
                  }
                  for (long long k0 = kbeg; k0 < kend; k0 += NS * bk)
                    for (int u = 0; u < NS; ++u) {
                      const long long ks = k0 + (long long)u * bk;
                      fetch(ks + (long long)deep_lane_fetch_ahead(NS) * bk, bk, K, "deep lane", rs, pl, bm, bn, akf, bkf);
                      fetch(ks + (long long)deep_uniform_fetch_ahead(NS) * bk, bk, K, "deep uniform", rs, pl, bm, bn,
                            akf, bkf);
                    }
                  // One-stage pipeline.
                  fetch(kbeg, bk, K, "shallow prologue", rs, pl, bm, bn, akf, bkf);
                  fetch(kbeg + bk, bk, K, "shallow prologue", rs, pl, bm, bn, akf, bkf);
                  for (long long k0 = kbeg; k0 < kend; k0 += bk)
                    fetch(k0 + (long long)shallow_fetch_ahead() * bk, bk, K, "shallow", rs, pl, bm, bn, akf, bkf);
                }
              }
          }
  std::printf("configs %d worst_past_K %lld pad %lld failures %d\n", configs, worst, (long long)kKTablePad, failures);
  return failures ? 1 : 0;
}
