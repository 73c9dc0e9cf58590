// Tile geometry shared by the GEMM kernels, the host planner and the host-only
// tests: LDS pitches, the K depth of a pipeline stage, and how far ahead of
// the stage it multiplies the gather reads the K offset tables.  Plain C++
// (no HIP), so a host compiler can include it on its own.
#pragma once

#include <stdint.h>

namespace qs {

// Row pitch, in elements, of an LDS tile stored [row][k]: the smallest pitch
// >= bk that keeps a fragment read free of bank conflicts.
constexpr int lds_pitch_rowmajor(int rsize, int bk) { return rsize == 8 ? bk + 1 : ((bk / 4) % 2 ? bk : bk + 4); }

// K depth of one LDS stage: the largest power of two in [4, 32] (4-byte
// elements: [4, 64]) whose double-buffered A and B tiles fit the LDS budget
// (76 KiB: two workgroups per CU).  FP32 MFMAs run at twice the FP64 rate, so
// a stage's fixed costs weigh twice as much; a deeper stage halves them.
// Shared by the kernels and the host planner, which rounds split-K chunks to it.
constexpr int pick_bk_for(int rsize, int planes, int bm, int bn, bool akf, bool bkf) {
  int best = 4;
  for (int bk = 4; bk <= (rsize == 4 ? 64 : 32); bk *= 2) {
    const int a = akf ? bm * lds_pitch_rowmajor(rsize, bk) : bk * (bm + 16);
    const int b = bkf ? bn * lds_pitch_rowmajor(rsize, bk) : bk * (bn + 16);
    if (2 * planes * (a + b) * rsize <= 76 * 1024) best = bk;
  }
  return best;
}

// How many stages ahead of the stage an iteration issues loads for the
// gather fetches K offsets.  The two-stage pipeline (NS register sets)
// issues the loads of stage j + NS in iteration j and fetches per-lane
// offsets NS - 1 iterations before their loads, i.e. for stage
// j + 2 NS - 1; wave-uniform offsets (scalar loads) for stage j + NS + 1.
// The one-stage pipeline fetches the offsets of stage j + 2.
constexpr int deep_lane_fetch_ahead(int ns) { return 2 * ns - 1; }
constexpr int deep_uniform_fetch_ahead(int ns) { return ns + 1; }
constexpr int shallow_fetch_ahead() { return 2; }

// Entries past K that a gather reads in its K tables: the last iteration
// starts at a stage k0 <= K - 1 and, in its last step (k0 + (NS - 1) BK),
// fetches a whole stage deep_lane_fetch_ahead(NS) stages on.
constexpr int64_t k_table_reach(int bk, int ns) {
  return ns >= 2 ? (int64_t)(ns - 1 + deep_lane_fetch_ahead(ns) + 1) * bk - 1
                 : (int64_t)(shallow_fetch_ahead() + 1) * bk - 1;
}

// Largest operand span, in bytes, that the buffer-load gather addresses: a
// K-table pad entry moves an offset 2^31 bytes on, past any such span.
constexpr int64_t kBufferSpanMax = 0x7FFFFF00LL;
// Entries past K in the K offset tables: enough for the deepest stage (64 K
// steps of a 4-byte tile) on the two-stage pipeline.
constexpr int64_t kKTablePad = 320;
static_assert(k_table_reach(64, 2) <= kKTablePad, "K table pad shorter than the two-stage gather's reach");
static_assert(k_table_reach(64, 1) <= kKTablePad, "K table pad shorter than the one-stage gather's reach");

}  // namespace qs
