// Internal declarations shared by the C-ABI layer (qs_api.cpp) and the
// gfx950 kernels (qs_gemm.hip, qs_bw.hip).  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/qs_amd.h"
#include "qs_tiles.h"

namespace qs {

// Thread-local error text behind qs_last_error().
void set_error(const std::string &msg);
qs_status fail(qs_status code, const std::string &msg);

// Every device and pinned-host allocation the library makes goes through
// these, which count them (qs_alloc_stats_get): a steady-state execute must
// leave the counts unchanged.
hipError_t dev_malloc(void **p, size_t bytes);
template <typename T> hipError_t dev_malloc(T **p, size_t bytes) { return dev_malloc((void **)p, bytes); }
hipError_t dev_free(void *p);
hipError_t pinned_malloc(void **p, size_t bytes, unsigned int flags);
template <typename T> hipError_t pinned_malloc(T **p, size_t bytes, unsigned int flags) {
  return pinned_malloc((void **)p, bytes, flags);
}
hipError_t pinned_free(void *p);
// Workspace blocks taken from / handed back to a caller's qs_allocator.
void count_bridge(bool alloc);

inline bool is_complex(int t) { return t == QS_C32 || t == QS_C64; }
inline bool is_double(int t) { return t == QS_F64 || t == QS_C64; }
inline int elem_bytes(int t) { return t == QS_F32 ? 4 : (t == QS_F64 || t == QS_C32) ? 8 : 16; }

// Which tile shape a GEMM launch uses; every variant is one template
// instantiation in qs_gemm.hip.
enum TileId : int {
  TILE_128x128 = 0,   // 4 waves, 64x64 per wave (4x4 MFMA 16x16 blocks)
  TILE_64x64 = 1,     // 4 waves, 32x32 per wave (2x2 blocks)
  TILE_128x64 = 2,    // 4 waves, 64x32 per wave
  TILE_64x128 = 3,    // 4 waves, 32x64 per wave
  TILE_32x64 = 4,     // 4 waves, 16x32 per wave (small M)
  TILE_32x32 = 5,     // 4 waves, 16x16 per wave (four workgroups per CU where the grid is small)
  TILE_COUNT = 6
};
struct TileShape { int bm, bn, bk; };
TileShape tile_shape(int tile, int compute_type, bool a_kfast, bool b_kfast);

// Everything one GEMM launch reads.  Offsets are in elements of each
// storage; the six tables are the full-extent fill_offsets! of each map.
struct GemmArgs {
  const void *A;
  const void *B;
  void *C;
  int64_t a_base, b_base, c_base;
  const int64_t *mA, *kA, *nB, *kB, *mC, *nC;
  int64_t M, N, K;
  int a_type, b_type, c_type, compute_type;
  int conj_a, conj_b;
  double alpha_re, alpha_im, beta_re, beta_im;
  int beta_case;          // 0: beta == 0, 1: beta == 1, 2: general
  int k_splits;           // > 1: partial sums into `partials`, reduced in a fixed order
  int64_t k_chunk;        // K steps per split (a multiple of twice the tile's BK)
  void *partials;         // [k_splits][planes][M][N] of the real compute type
  int tile;               // TileId
  int a_kfast, b_kfast;   // gather lanes along K (1) or along M / N (0)
  int mixed;              // the generic gather: some storage type differs from the compute
                          // type, or an operand span exceeds the 32-bit buffer range
  int64_t a_lo, b_lo;     // accessed spans of A and B (elements from the storage start) ...
  uint32_t a_bytes, b_bytes;   // ... and their lengths in bytes, for the buffer-load gather
  // A batched launch (k_splits == 1, mixed == 0): batch contractions, batch
  // b with batch_off[3b .. 3b+2] elements added to the A, B and C bases.
  // batch_off null: one contraction.
  const int64_t *batch_off;
  int64_t batch;
  // Every row pair (K pair) the gathers read in twos is adjacent in memory:
  // the tiles on the two-stage pipeline load two elements per lane.
  int vec = 0;
  // Mixed domains (complex compute type, mixed storage types): 1 = A complex
  // and B real, 2 = A real and B complex; 0 otherwise.
  int dom = 0;
  // dom != 0 with the complex operand and C in the compute type and the real
  // operand in its real type, spans below 2 GiB: both gathers take buffer
  // loads (a_lo / a_bytes and b_lo / b_bytes are set).
  int dom_buf = 0;
  // Workgroup -> tile order of a one-contraction launch.  The hardware deals
  // workgroups round-robin over `xcds` XCDs (one L2 each).  map_gx > 0: the
  // tile grid is cut into xcds patches of map_pm x map_pn tiles (map_gx of
  // them along M), one per XCD, so each L2 holds a slice of A's rows and of
  // B's columns; inside a patch, groups of map_gh tile rows are walked along
  // N.  map_gx == 0: column-major order, renumbered so each XCD walks a
  // contiguous run.
  int xcds = 8;
  int map_gx = 0, map_pm = 0, map_pn = 0, map_gh = 0;
  // Maps that are a single ramp (offset of coordinate q = q * step,
  // affine_ramp of src/layout/axis_group.jl:149-180): the kernels compute
  // those offsets instead of reading the table where the read would sit on
  // a tile's critical path (row offsets, the first stages' K offsets, C's
  // offsets in the epilogue).  Bits RAMP_*; kpad_*: the K tables' pad value.
  int ramps = 0;
  int64_t step_mA = 0, step_kA = 0, step_nB = 0, step_kB = 0, step_mC = 0, step_nC = 0;
  int64_t kpad_a = 0, kpad_b = 0;
};
enum RampBits { RAMP_MA = 1, RAMP_KA = 2, RAMP_NB = 4, RAMP_KB = 8, RAMP_MC = 16, RAMP_NC = 32 };

// Facts of the current device that planning and launches use
// (hipGetDeviceProperties, cached per device): the gfx950 TargetProfile's
// counterpart of src/hardware/target.jl:26-47.
struct DeviceFacts {
  int cus = 256;          // compute units
  int xcds = 8;           // XCDs (one L2 each); workgroups are dealt round-robin over them
  int wave = 64;          // wavefront size
  int lds_per_cu = 160 * 1024;
  int lds_per_block = 64 * 1024;
  char arch[64] = {0};
};
const DeviceFacts &device_facts();

// Launch the tiled MFMA GEMM (with k_splits > 1, a second kernel reduces the
// partials in a fixed order).  Shapes and tables must already be validated.
hipError_t launch_gemm(const GemmArgs &g, hipStream_t stream);

// C = beta * C over the M x N rectangle (alpha == 0 or K == 0).
hipError_t launch_scale(const GemmArgs &g, hipStream_t stream);

// Bandwidth-bound paths.
hipError_t launch_dot(const GemmArgs &g, hipStream_t stream);
hipError_t launch_outer(const GemmArgs &g, hipStream_t stream);

}  // namespace qs
