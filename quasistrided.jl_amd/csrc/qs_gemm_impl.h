// The tiled block-scatter GEMM for gfx950, templated on the real compute type
// R (double / float) and on complex-ness.  Included by one .hip file per
// compute type, so the instantiations compile in parallel.
//
// Replaces the reference's five-loop nest (src/execution/nest.jl:7-122), its
// packers (src/packing/pack.jl:43-99) and microkernels
// (src/microkernels/steps.jl:73-167, stores.jl:150-278):
//   * one workgroup (4 waves, 2x2) per BM x BN tile of C, full K (or one K
//     split) accumulated in MFMA registers -- beta applied once, in the
//     epilogue, by construction (nest.jl:83 applies it on the first K block);
//     with split-K every split stores its partial sums and a second kernel
//     adds them in a fixed order and runs the epilogue;
//   * the pack is a gather fused into the global->LDS load: address =
//     base + rowoffset[i] + koffset[k] (Tile, src/layout/tiles.jl:102-103),
//     lanes laid along whichever of the free or the K group is unit-stride in
//     that operand, conj applied and storage type converted on load;
//   * FP64 / FP32 MFMA 16x16x4 register tiles; complex T is planar ("1r"):
//     four real MFMAs per block, re = (-ai)bi + (ar br + c), im = ai br + (ar bi + c)
//     (steps.jl:121-167);
//   * epilogue C = alpha*acc + beta*C in the compute type with the reference's
//     expression trees (interface.jl:99-104, stores.jl:91-124), rounded once to
//     C's storage type and scatter-stored through C's offset tables; beta == 0
//     never reads C.
#pragma once

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "qs_internal.h"

namespace qs {

typedef double qs_d4 __attribute__((ext_vector_type(4)));
typedef float qs_f4 __attribute__((ext_vector_type(4)));

template <typename R> struct Acc4;
template <> struct Acc4<double> { typedef qs_d4 type; };
template <> struct Acc4<float> { typedef qs_f4 type; };

// v_mfma_f64_16x16x4_f64 / v_mfma_f32_16x16x4_f32.  Lane l holds A[l%16][l/16]
// and B[l/16][l%16].  D register r of lane l is D[4*(l/16)+r][l%16] for FP32
// but D[l/16+4r][l%16] for FP64 (checked on gfx950 with one-hot operands).
__device__ __forceinline__ qs_d4 mfma16(double a, double b, qs_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ qs_f4 mfma16(float a, float b, qs_f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row of D held in register r by a lane of group fk = lane/16.
template <typename R> __device__ __forceinline__ int mfma_row(int fk, int r);
template <> __device__ __forceinline__ int mfma_row<double>(int fk, int r) { return fk + 4 * r; }
template <> __device__ __forceinline__ int mfma_row<float>(int fk, int r) { return 4 * fk + r; }

__device__ __forceinline__ double qs_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float qs_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Operand storages live in device global memory; addressing them through
// address-space-1 pointers keeps every access a global_load/global_store
// (a generic pointer would turn them into flat accesses, which also count
// against the LDS wait counter).
#define QS_GLOBAL __attribute__((address_space(1)))

// Element `idx` of storage `p` as the compute type.  MIXED: the storage type
// `t` is a runtime value (uniform across the workgroup); otherwise it is the
// compute type itself.
// NT: a non-temporal load, for operands streamed once (a dot product's
// matrix operand, far larger than the caches).
template <typename R, bool CPLX, bool MIXED, bool NT = false>
__device__ __forceinline__ void load_elem(const void *p, int t, int64_t idx, R &re, R &im) {
  if (!MIXED) {
    const QS_GLOBAL R *q = (const QS_GLOBAL R *)p;
    if constexpr (NT) {
      if (!CPLX) {
        re = __builtin_nontemporal_load(q + idx);
        im = 0;
      } else {
        re = __builtin_nontemporal_load(q + 2 * idx);
        im = __builtin_nontemporal_load(q + 2 * idx + 1);
      }
    } else if (!CPLX) {
      re = q[idx];
      im = 0;
    } else {
      re = q[2 * idx];
      im = q[2 * idx + 1];
    }
    return;
  }
  const QS_GLOBAL float *f = (const QS_GLOBAL float *)p;
  const QS_GLOBAL double *d = (const QS_GLOBAL double *)p;
  switch (t) {
    case QS_F32: re = (R)f[idx]; im = 0; break;
    case QS_F64: re = (R)d[idx]; im = 0; break;
    case QS_C32: re = (R)f[2 * idx]; im = (R)f[2 * idx + 1]; break;
    default: re = (R)d[2 * idx]; im = (R)d[2 * idx + 1]; break;
  }
}

// NT: a non-temporal store, for outputs written once and not read back by
// the same kernel (streams of C much larger than the caches: 112 -> 82 us
// for an 8192 x 8192 outer product).
template <typename R, bool CPLX, bool MIXED, bool NT = false>
__device__ __forceinline__ void store_elem(void *p, int t, int64_t idx, R re, R im) {
  if (!MIXED) {
    QS_GLOBAL R *q = (QS_GLOBAL R *)p;
    if constexpr (NT) {
      if (!CPLX) {
        __builtin_nontemporal_store(re, q + idx);
      } else {
        __builtin_nontemporal_store(re, q + 2 * idx);
        __builtin_nontemporal_store(im, q + 2 * idx + 1);
      }
    } else if (!CPLX) {
      q[idx] = re;
    } else {
      q[2 * idx] = re;
      q[2 * idx + 1] = im;
    }
    return;
  }
  QS_GLOBAL float *f = (QS_GLOBAL float *)p;
  QS_GLOBAL double *d = (QS_GLOBAL double *)p;
  switch (t) {
    case QS_F32: f[idx] = (float)re; break;
    case QS_F64: d[idx] = (double)re; break;
    case QS_C32: f[2 * idx] = (float)re; f[2 * idx + 1] = (float)im; break;
    default: d[2 * idx] = (double)re; d[2 * idx + 1] = (double)im; break;
  }
}

// The epilogue at one element of C: `alpha*x + beta*C` for the beta case,
// with Base's Complex expression trees (stores.jl:91-124; the library is
// built with -ffp-contract=off so the unfused products stay unfused).  Split
// in two so callers can issue every old-C load of a batch before the first
// store (loads after a store to C cannot be hoisted above it).  BC: the beta
// case when the caller has already branched on it (-1: read it from g).
template <typename R, bool CPLX, bool MIXED, int BC = -1>
__device__ __forceinline__ void load_c_old(const GemmArgs &g, int64_t idx, R &o, R &oi) {
  o = 0;
  oi = 0;
  if ((BC < 0 ? g.beta_case : BC) != 0) load_elem<R, CPLX, MIXED>(g.C, g.c_type, idx, o, oi);
}

template <typename R, bool CPLX, bool MIXED, bool NT = false, int BC = -1>
__device__ __forceinline__ void store_c_at(const GemmArgs &g, int64_t idx, R xr, R xi, R o, R oi) {
  const R ar = (R)g.alpha_re, ai = (R)g.alpha_im;
  const int bc = BC < 0 ? g.beta_case : BC;
  R outr, outi = 0;
  if (!CPLX) {
    if (bc == 0) outr = ar * xr;
    else outr = bc == 1 ? qs_fma(ar, xr, o) : qs_fma(ar, xr, (R)g.beta_re * o);
  } else {
    if (bc == 0) {
      R t1 = ar * xr, t2 = ai * xi, t3 = ar * xi, t4 = ai * xr;
      outr = t1 - t2;
      outi = t3 + t4;
    } else {
      R tr = o, ti = oi;
      if (bc == 2) {
        const R br = (R)g.beta_re, bi = (R)g.beta_im;
        R u1 = br * o, u2 = bi * oi, u3 = br * oi, u4 = bi * o;
        tr = u1 - u2;
        ti = u3 + u4;
      }
      outr = qs_fma(ar, xr, -qs_fma(ai, xi, -tr));
      outi = qs_fma(ar, xi, qs_fma(ai, xr, ti));
    }
  }
  store_elem<R, CPLX, MIXED, NT>(g.C, g.c_type, idx, outr, outi);
}

// Both halves at one element (old C loaded only when beta != 0).
template <typename R, bool CPLX, bool MIXED>
__device__ __forceinline__ void store_c(const GemmArgs &g, int64_t row, int64_t col, R xr, R xi) {
  const int64_t idx = g.c_base + g.mC[row] + g.nC[col];
  const R ar = (R)g.alpha_re, ai = (R)g.alpha_im;
  R outr, outi = 0;
  if (!CPLX) {
    if (g.beta_case == 0) {
      outr = ar * xr;
    } else {
      R o, oi;
      load_elem<R, false, MIXED>(g.C, g.c_type, idx, o, oi);
      outr = g.beta_case == 1 ? qs_fma(ar, xr, o) : qs_fma(ar, xr, (R)g.beta_re * o);
    }
  } else {
    if (g.beta_case == 0) {
      R t1 = ar * xr, t2 = ai * xi, t3 = ar * xi, t4 = ai * xr;
      outr = t1 - t2;
      outi = t3 + t4;
    } else {
      R orr, ori;
      load_elem<R, true, MIXED>(g.C, g.c_type, idx, orr, ori);
      R tr = orr, ti = ori;
      if (g.beta_case == 2) {
        const R br = (R)g.beta_re, bi = (R)g.beta_im;
        R u1 = br * orr, u2 = bi * ori, u3 = br * ori, u4 = bi * orr;
        tr = u1 - u2;
        ti = u3 + u4;
      }
      outr = qs_fma(ar, xr, -qs_fma(ai, xi, -tr));
      outi = qs_fma(ar, xi, qs_fma(ai, xr, ti));
    }
  }
  store_elem<R, CPLX, MIXED>(g.C, g.c_type, idx, outr, outi);
}

template <typename R, bool CPLX, int BM, int BN, bool AKF, bool BKF>
constexpr int pick_bk() { return pick_bk_for((int)sizeof(R), CPLX ? 2 : 1, BM, BN, AKF, BKF); }

// LDS geometry of one operand tile (ROWS x BK, P planes, double-buffered).
// A gather laid along the free group (KF == false) stores the tile K-major,
// [k][row] with a row pitch of ROWS + 16; one laid along K (KF == true)
// stores it row-major, [row][k], with the smallest pitch >= BK that keeps a
// fragment read free of bank conflicts (a quarter-wave's 16 rows land on
// distinct banks).
template <typename R, int ROWS, int BK, bool KF>
struct LdsTile {
  static constexpr int kpitch = ROWS + 16;
  static constexpr int rpitch = lds_pitch_rowmajor((int)sizeof(R), BK);
  static constexpr int plane = KF ? ROWS * rpitch : BK * kpitch;
  __device__ static __forceinline__ int at(int row, int k) { return KF ? row * rpitch + k : k * kpitch + row; }
};


// XCD-aware workgroup order: the hardware deals workgroups round-robin over
// the X XCDs, so consecutive ids land on different L2s.  Renumber so each XCD
// walks a contiguous run of tiles (which share B columns and A rows).
__device__ __forceinline__ int64_t xcd_swizzle(int64_t bid, int64_t nwg, int X) {
  if (X <= 1 || nwg % X != 0) return bid;
  return (bid % X) * (nwg / X) + bid / X;
}

// The tile (tm, tn) of workgroup `bid` of a one-contraction launch of `nwg`
// workgroups (GemmArgs::map_*): XCD x = bid % X takes patch x of the tile
// grid, and its j-th workgroup the j-th tile of that patch, walked in groups
// of map_gh tile rows along N.  Without a patch map, column-major order
// through the contiguous-run renumbering.
__device__ __forceinline__ void tile_of(const GemmArgs &g, int64_t bid, int64_t nwg, int64_t tiles_m, int64_t &tm,
                                        int64_t &tn) {
  const int X = g.xcds;
  if (g.map_gx > 0 && X > 1 && nwg % X == 0) {
    const int64_t x = bid % X, j = bid / X;
    const int64_t pm = g.map_pm, pn = g.map_pn, gh = g.map_gh;
    const int64_t grp = j / (gh * pn), r = j % (gh * pn);
    const int64_t h = min(gh, pm - grp * gh);
    tm = (x % g.map_gx) * pm + grp * gh + r % h;
    tn = (x / g.map_gx) * pn + r / h;
  } else {
    const int64_t v = xcd_swizzle(bid, nwg, X);
    tm = v % tiles_m;
    tn = v / tiles_m;
  }
}

// Host: the patch map of a BM x BN launch over tiles_m x tiles_n tiles.  The
// grid is cut into X = gx * gy patches (gx along M) that divide it exactly;
// among those, the one whose patch needs the fewest A rows plus B columns
// (what each XCD's L2 must fetch per K step).  Inside a patch, groups of gh
// tile rows make the workgroups one XCD runs at once (two per CU) about
// square in rows and columns.  No exact cut: map_gx stays 0.
inline void set_tile_map(GemmArgs &g, int bm, int bn, int64_t tiles_m, int64_t tiles_n, int cus) {
  g.map_gx = g.map_pm = g.map_pn = g.map_gh = 0;
  const int X = g.xcds;
  if (X <= 1 || (tiles_m * tiles_n) % X != 0) return;
  int64_t best = -1;
  for (int gx = 1; gx <= X; ++gx) {
    if (X % gx != 0 || tiles_m % gx != 0 || tiles_n % (X / gx) != 0) continue;
    const int64_t pm = tiles_m / gx, pn = tiles_n / (X / gx);
    const int64_t cost = pm * bm + pn * bn;
    if (best < 0 || cost < best) {
      best = cost;
      g.map_gx = gx;
      g.map_pm = (int)pm;
      g.map_pn = (int)pn;
    }
  }
  if (g.map_gx == 0) return;
  // Concurrent workgroups per XCD (two per CU): gh rows x (conc / gh) columns
  // of tiles with gh * bm close to (conc / gh) * bn.
  const double conc = 2.0 * cus / X;
  const int target = std::max(1, (int)(std::sqrt(conc * bn / bm) + 0.5));
  int gh = 1;
  for (int d = 1; d <= g.map_pm; ++d)
    if (g.map_pm % d == 0 && d <= target) gh = d;
  g.map_gh = gh;
}

typedef const __attribute__((address_space(4))) int64_t *qs_const_i64;

// Sixteen zero bytes: what a gather lane outside the tile reads.
__device__ const double qs_zero16[2] = {0.0, 0.0};

// One operand's share of the gather.  Thread t loads NL elements of the
// ROWS x BK stage.  Laid along the free group (KF == false), its row is fixed
// (t % ROWS) and its K steps are kk0 + r*(256/ROWS) -- uniform over a wave
// when ROWS >= 64, so their offsets come through the scalar cache.  Laid
// along K (KF == true), its K step is fixed (t % BK) and its rows are
// t/BK + r*(256/BK).  Row offsets are read once; K offsets are fetched one
// stage ahead of the data that uses them.  Loads are unconditional at
// clamped in-range coordinates; out-of-tile values become zero.
//
// VW == 2 (real, one storage type): each load takes two consecutive elements
// -- two rows of one K step (KF == false) or two K steps of one row -- that
// the host has found adjacent in memory for every such pair of the operand's
// tables, in one 2 x ESZ-byte buffer load.  Half the load instructions.
//
// AUX: the buffer loads' cache-policy bits (2: non-temporal, for an operand
// that only one workgroup reads).
template <typename R, bool CPLX, bool MIXED, int ROWS, int BK, bool KF, int VS = 2, int VW = 1, int AUX = 0>
struct Gather {
  static constexpr int P = CPLX ? 2 : 1;
  static constexpr int NL = (ROWS * BK + 256 * VW - 1) / (256 * VW);   // loads per thread
  static constexpr int RU = KF ? ROWS : ROWS / VW;   // row units (a unit is VW rows when KF == false)
  static constexpr int KU = KF ? BK / VW : BK;       // K units (VW K steps when KF == true)
  static constexpr int NR = KF ? NL : 1;           // distinct rows per thread
  static constexpr int NK = KF ? 1 : NL;           // distinct K steps per thread
  static constexpr int RSTEP = KF ? 256 / KU : 0;  // row step between loads
  static constexpr int KSTEP = KF ? 0 : 256 / RU;
  static constexpr bool UNIFORM_K = !KF && RU >= 64;
  static constexpr bool PARTIAL = ROWS * BK % (256 * VW) != 0;
  static_assert(VW == 1 || (!CPLX && !MIXED && !PARTIAL), "paired loads: real, one storage type, whole stages");
  // One storage type: buffer loads over the operand's accessed span, with
  // 32-bit byte offsets and no bounds tests.  Rows past M (N) repeat the last
  // row: they only reach accumulators of rows (columns) that are never
  // stored.  K past the range reads the K tables' pad entries, whose offsets
  // land past the span, which the hardware returns as zero.  A split's K
  // range is a whole number of stage pairs, so no multiplied stage reaches
  // into the next split.  Mixed types: 64-bit element indices, bounds tests
  // and a pointer select (the device zero for outside elements).
  static constexpr bool BUF = !MIXED;
  static constexpr int ESZ = P * (int)sizeof(R);
  typedef typename std::conditional<BUF, uint32_t, int64_t>::type Off;
  // K offsets of VS stages (set = stage mod VS).  Wave-uniform offsets are
  // scalar loads, outside the in-order vector-load count: one set, fetched
  // after the loads that use it.
  static constexpr int KS = UNIFORM_K ? 1 : VS;
  __amdgpu_buffer_rsrc_t rsrc;
  Off rowoff[NR];
  bool rowok[NR];
  int row0, kk0;
  Off koff[KS][NK];
  bool kok[KS][NK];
  R val[VS][NL][P * VW];   // VS stages in flight

  // `lo` / `bytes`: the operand's accessed span (elements from its storage
  // start, and its length in bytes) for the buffer loads.
  // rramp: the row map is a single ramp of step rstep (no table read).
  __device__ __forceinline__ void init(int t, int64_t r0, int64_t rows_total, int64_t base,
                                       const int64_t *__restrict__ rowtab, bool rramp, int64_t rstep, const void *p,
                                       int64_t lo, uint32_t bytes) {
    row0 = KF ? t / KU : VW * (t % RU);
    kk0 = KF ? VW * (t % KU) : t / RU;
    if (UNIFORM_K) kk0 = __builtin_amdgcn_readfirstlane(kk0);
    if constexpr (BUF)
      rsrc = __builtin_amdgcn_make_buffer_rsrc((void *)((const char *)p + lo * ESZ), (short)0, (int)bytes, 0x00020000);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int64_t gi = r0 + row0 + r * RSTEP;
      rowok[r] = gi < rows_total && (!PARTIAL || t + 256 * r < ROWS * BK);
      // Past the last row: the last one (the last pair; rows_total is even when VW == 2).
      const int64_t gc = gi < rows_total ? gi : rows_total - (KF ? 1 : VW);
      const int64_t o = base + (rramp ? gc * rstep : rowtab[gc]);
      // Row and K byte offsets add up, modulo 2^32, to the element's offset in the span.
      rowoff[r] = BUF ? (Off)(uint32_t)((o - lo) * ESZ) : (Off)o;
    }
  }
  template <int S>
  __device__ __forceinline__ void fetch_k(int64_t k0, int64_t kend, const int64_t *__restrict__ ktab) {
#pragma unroll
    for (int r = 0; r < NK; ++r) {
      const int64_t gk = k0 + kk0 + r * KSTEP;
      if constexpr (BUF) {
        // The element offset as loaded: scaling it here would make the
        // compiler wait for the table load before the stage's own loads.
        const int64_t o = UNIFORM_K ? ((qs_const_i64)ktab)[gk] : ktab[gk];
        koff[S % KS][r] = (Off)(uint32_t)o;
      } else {
        kok[S % KS][r] = gk < kend;
        const int64_t gkc = gk < kend ? gk : kend - 1;
        koff[S % KS][r] = UNIFORM_K ? ((qs_const_i64)ktab)[gkc] : ktab[gkc];
      }
    }
  }
  // fetch_k for a stage on a tile's critical path (the first ones): with a
  // single-ramp K map (kramp) the offsets are computed, not read; K steps
  // past kend take the tables' pad value.
  template <int S>
  __device__ __forceinline__ void fetch_k_first(int64_t k0, int64_t kend, const int64_t *__restrict__ ktab, bool kramp,
                                                int64_t kstep, int64_t kpad) {
    if (!kramp) {
      fetch_k<S>(k0, kend, ktab);
      return;
    }
#pragma unroll
    for (int r = 0; r < NK; ++r) {
      const int64_t gk = k0 + kk0 + r * KSTEP;
      if constexpr (BUF) {
        koff[S % KS][r] = (Off)(uint32_t)(gk < kend ? gk * kstep : kpad);
      } else {
        kok[S % KS][r] = gk < kend;
        koff[S % KS][r] = (gk < kend ? gk : kend - 1) * kstep;
      }
    }
  }
  // Issue the stage's loads; nothing waits on them until `stage` consumes them.
  template <int S>
  __device__ __forceinline__ void load(const void *__restrict__ p, int type) {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      const int ir = KF ? r : 0, kr = KF ? 0 : r;
      if constexpr (BUF) {
        const uint32_t off = rowoff[ir] + koff[S % KS][kr] * (uint32_t)ESZ;
        // The loaded vector is cast whole to a scalar or a struct before any
        // part is taken: extracting its lanes directly lets the compiler
        // narrow the load to one dword (and then read the wrong lane).
        if constexpr (VW == 2) {
          struct Two { R x, y; };
          Two v;
          if constexpr (ESZ == 4)
            v = __builtin_bit_cast(Two, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, AUX));
          else
            v = __builtin_bit_cast(Two, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, AUX));
          val[S][r][0] = v.x;
          val[S][r][1] = v.y;
        } else if constexpr (ESZ == 4) {
          val[S][r][0] = __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b32(rsrc, off, 0, AUX));
        } else if constexpr (ESZ == 8 && !CPLX) {
          val[S][r][0] = __builtin_bit_cast(R, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, AUX));
        } else if constexpr (ESZ == 8) {
          const uint64_t u = __builtin_bit_cast(uint64_t, __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, AUX));
          val[S][r][0] = __builtin_bit_cast(R, (uint32_t)u);
          val[S][r][P - 1] = __builtin_bit_cast(R, (uint32_t)(u >> 32));
        } else {
          struct Pair { R re, im; };
          const Pair v = __builtin_bit_cast(Pair, __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, AUX));
          val[S][r][0] = v.re;
          val[S][r][P - 1] = v.im;
        }
      } else {
        const bool ok = rowok[ir] && kok[S % KS][kr];
        const void *src = ok ? p : (const void *)qs_zero16;
        const int64_t idx = ok ? rowoff[ir] + koff[S % KS][kr] : 0;
        R im;
        load_elem<R, CPLX, MIXED>(src, type, idx, val[S][r][0], im);
        if (CPLX) val[S][r][P - 1] = im;
      }
    }
  }
  template <int S, int PLANE>
  __device__ __forceinline__ void stage(R *lds, int t, R sgn) {
    typedef LdsTile<R, ROWS, BK, KF> L;
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      if (!PARTIAL || t + 256 * r < ROWS * BK) {
        const int row = row0 + (KF ? r * RSTEP : 0);
        const int k = kk0 + (KF ? 0 : r * KSTEP);
        if constexpr (VW == 2) {
          lds[L::at(row, k)] = val[S][r][0];
          lds[L::at(KF ? row : row + 1, KF ? k + 1 : k)] = val[S][r][1];
        } else {
          lds[L::at(row, k)] = val[S][r][0];
          if (CPLX) lds[PLANE + L::at(row, k)] = sgn * val[S][r][P - 1];
        }
      }
    }
  }
};

// Scheduling of one K stage's basic block: its NMF MFMAs with the stage's
// other work spread between them -- after FIRST LDS fragment reads, each MFMA
// is followed by one global load, PER fragment reads and one LDS write of the
// next stage.  Issued as a burst, the gather loads stalled the wave in order
// at the memory unit and held up the MFMAs behind them: on the 1024^3 FP64
// shape with 32 x 64 tiles, 61 -> 51 us; without any in-loop loads the same
// kernel took 44 us.  Waves with four or more MFMA blocks read six fragments
// ahead and one per MFMA (64 x 64 tiles 55 -> 52 us, 64 x 128 at 4096^3
// 2.10 -> 2.05 ms); smaller ones three ahead and two per MFMA.
template <int NMF, int FIRST, int PER>
__device__ __forceinline__ void sched_interleave() {
  __builtin_amdgcn_sched_group_barrier(0x100, FIRST, 0);   // LDS reads
#pragma unroll
  for (int i = 0; i < NMF; ++i) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);     // MFMA
    __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);     // global load
    __builtin_amdgcn_sched_group_barrier(0x100, PER, 0);   // LDS reads
    __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);     // LDS write
  }
}

// One BM x BN tile of C over the K range of split `split`: the whole
// workgroup's work, shared by the one-contraction kernel and the grouped
// persistent kernel.  `smem` holds SMEM elements of R.
//
// DOM (complex compute type, mixed storage types): 1 = complex A times real
// B, 2 = real A times complex B -- the reference's ComplexRealKernel and
// RealComplexKernel (src/microkernels/kernels.jl:125-160).  The real operand
// keeps one LDS plane and a real gather, and each block takes two MFMAs
// (re += ar br, im += ai br; or re += ar br, im += ar bi) instead of four.
//
// BKO: the K stage depth, if not the deepest that fits the LDS budget
// (pick_bk).
template <typename R, bool CPLX, int BM, int BN, bool AKF, bool BKF, bool MIXED, bool VEC = false, int DOM = 0,
          int BAUX = 0, int BKO = 0>
struct TileGemm {
  static constexpr int P = CPLX ? 2 : 1;
  static constexpr bool CA = CPLX && DOM != 2, CB = CPLX && DOM != 1;   // complex A / B
  static constexpr int PA = CA ? 2 : 1, PB = CB ? 2 : 1;
  static constexpr int NMFMA = CPLX ? (DOM ? 2 : 4) : 1;   // MFMAs per 16 x 16 block and k step
  static_assert(DOM == 0 || CPLX, "mixed domains: complex compute type");
  static constexpr int BK = BKO ? BKO : pick_bk<R, CPLX, BM, BN, AKF, BKF>();
  static constexpr int WTM = BM / 2, WTN = BN / 2;       // 2 x 2 waves
  static constexpr int FM = WTM / 16, FN = WTN / 16;     // MFMA blocks per wave
  typedef LdsTile<R, BM, BK, AKF> LA;
  typedef LdsTile<R, BN, BK, BKF> LB;
  static constexpr int SA = PA * LA::plane, SB = PB * LB::plane;   // one stage of A / B
  static constexpr int SMEM = 2 * (SA + SB);
  static_assert(256 % BM == 0 && 256 % BN == 0, "tile rows");
  static_assert(BK % 4 == 0, "BK");

  // DEEPEN: keep two stages of gathered operands in flight where registers
  // allow (one storage type; real tiles up to 64 x 128, complex up to
  // 32 x 64); the grouped kernel, whose tile classes share one register
  // allocation, passes false.
  // DEEPEN 2: two stages for any tile (a caller whose register budget is
  // known to hold them).
  template <int DEEPEN = 1>
  __device__ static void run(const GemmArgs &g, int64_t tm, int64_t tn, int split, R *smem) {
    // Complex tiles up to 32 x 64 too: their gathers are small (ComplexF64
    // 32 x 64 x 16: 166 VGPRs, 139 -> 130 us at 1024^3).
    constexpr bool DEEP =
        !MIXED && (DEEPEN == 2 || (DEEPEN == 1 && (CPLX ? BM * BN <= 32 * 64 : BM * BN <= 64 * 128)));
    typedef typename Acc4<R>::type V;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t m0 = tm * BM, n0 = tn * BN;
    const int64_t kbeg = (int64_t)split * g.k_chunk;
    const int64_t kend = min(g.K, kbeg + g.k_chunk);

    const R sa = g.conj_a ? (R)-1 : (R)1, sb = g.conj_b ? (R)-1 : (R)1;

    V acc[P][FM][FN];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) acc[p][a][b] = V{0, 0, 0, 0};

    const int fr = lane & 15, fk = lane >> 4;
    if constexpr (DEEP) {
      // NS gather register sets (2, or 4 for tiles whose gathers are small).
      // Stage j of the K range lives in LDS buffer j & 1 and, while in
      // flight, in gather register set j mod NS (its K offsets in offset set
      // j mod NS).  Iteration j issues the loads of stage j + NS
      // (unconditionally: past kend every lane reads the device zero, so the
      // compiler's wait counts stay exact), multiplies stage j and stores
      // stage j + 1 into LDS: each stage's loads have NS - 1 iterations and a
      // part to land.  Per-lane K offsets are fetched NS - 1 iterations
      // before their loads (stage j + 2 NS - 1 in iteration j) and issued
      // ahead of the iteration's loads: vector loads complete in order, so
      // the wait for an offset covers every load issued before it, and this
      // way it never reaches a load younger than the one being staged.
      // Wave-uniform offsets (scalar loads) are fetched one stage ahead.
      auto deep = [&](auto ns) {
        constexpr int NS = decltype(ns)::value;
        // The buffer gather reads the K tables' pad entries past K unguarded.
        static_assert(k_table_reach(BK, NS) <= kKTablePad, "K table pad shorter than the gather's reach");
        Gather<R, CA, MIXED, BM, BK, AKF, NS, VEC ? 2 : 1> ga;
        Gather<R, CB, MIXED, BN, BK, BKF, NS, VEC ? 2 : 1, BAUX> gb;
        ga.init(t, m0, g.M, g.a_base, g.mA, g.ramps & RAMP_MA, g.step_mA, g.A, g.a_lo, g.a_bytes);
        gb.init(t, n0, g.N, g.b_base, g.nB, g.ramps & RAMP_NB, g.step_nB, g.B, g.b_lo, g.b_bytes);
        typedef decltype(ga) GA;
        typedef decltype(gb) GB;
        auto step = [&](int64_t k0, auto u) {
          constexpr int U = decltype(u)::value, NX = (U + 1) % NS, KL = (U + NS - 1) % NS;
          if constexpr (!GA::UNIFORM_K) ga.template fetch_k<KL>(k0 + deep_lane_fetch_ahead(NS) * BK, kend, g.kA);
          if constexpr (!GB::UNIFORM_K) gb.template fetch_k<KL>(k0 + deep_lane_fetch_ahead(NS) * BK, kend, g.kB);
          __builtin_amdgcn_sched_barrier(0);   // offsets issued before the loads
          ga.template load<U>(g.A, g.a_type);
          gb.template load<U>(g.B, g.b_type);
          if constexpr (GA::UNIFORM_K) ga.template fetch_k<0>(k0 + deep_uniform_fetch_ahead(NS) * BK, kend, g.kA);
          if constexpr (GB::UNIFORM_K) gb.template fetch_k<0>(k0 + deep_uniform_fetch_ahead(NS) * BK, kend, g.kB);
          const R *As = smem + (U & 1) * (SA + SB);
          const R *Bs = As + SA;
          R *Ws = smem + ((U + 1) & 1) * (SA + SB);
  #pragma unroll
          for (int ks = 0; ks < BK / 4; ++ks) {
            R af[PA][FM], bf[PB][FN];
  #pragma unroll
            for (int p = 0; p < PA; ++p)
  #pragma unroll
              for (int a = 0; a < FM; ++a) af[p][a] = As[p * LA::plane + LA::at(wm * WTM + a * 16 + fr, ks * 4 + fk)];
  #pragma unroll
            for (int p = 0; p < PB; ++p)
  #pragma unroll
              for (int b = 0; b < FN; ++b) bf[p][b] = Bs[p * LB::plane + LB::at(wn * WTN + b * 16 + fr, ks * 4 + fk)];
  #pragma unroll
            for (int a = 0; a < FM; ++a)
  #pragma unroll
              for (int b = 0; b < FN; ++b) {
                if constexpr (!CPLX) {
                  acc[0][a][b] = mfma16(af[0][a], bf[0][b], acc[0][a][b]);
                } else if constexpr (DOM == 1) {   // complex A, real B
                  acc[0][a][b] = mfma16(af[0][a], bf[0][b], acc[0][a][b]);
                  acc[1][a][b] = mfma16(af[PA - 1][a], bf[0][b], acc[1][a][b]);
                } else if constexpr (DOM == 2) {   // real A, complex B
                  acc[0][a][b] = mfma16(af[0][a], bf[0][b], acc[0][a][b]);
                  acc[1][a][b] = mfma16(af[0][a], bf[PB - 1][b], acc[1][a][b]);
                } else {
                  acc[0][a][b] = mfma16(af[0][a], bf[0][b], acc[0][a][b]);
                  acc[0][a][b] = mfma16(-af[PA - 1][a], bf[PB - 1][b], acc[0][a][b]);
                  acc[P - 1][a][b] = mfma16(af[0][a], bf[PB - 1][b], acc[P - 1][a][b]);
                  acc[P - 1][a][b] = mfma16(af[PA - 1][a], bf[0][b], acc[P - 1][a][b]);
                }
              }
          }
          ga.template stage<NX, LA::plane>(Ws, t, sa);
          gb.template stage<NX, LB::plane>(Ws + SA, t, sb);
          // The stage's loads, fragment reads and LDS writes spread between
          // its MFMAs (the loads are not pinned ahead of them).
          sched_interleave<(BK / 4) * FM * FN * NMFMA, FM * FN >= 4 ? 6 : 3, FM * FN >= 4 ? 1 : 2>();
          __syncthreads();
        };
        if (kbeg < kend) {
          // The same issue order as the loop ([per-lane offsets of stage
          // s + NS - 1][loads of stage s]), so the wait counts agree on entry
          // and on the back edge.
          const bool kra = g.ramps & RAMP_KA, krb = g.ramps & RAMP_KB;
          ga.template fetch_k_first<0>(kbeg, kend, g.kA, kra, g.step_kA, g.kpad_a);
          gb.template fetch_k_first<0>(kbeg, kend, g.kB, krb, g.step_kB, g.kpad_b);
          if constexpr (NS == 4) {
            if constexpr (!GA::UNIFORM_K) {
              ga.template fetch_k<1>(kbeg + BK, kend, g.kA);
              ga.template fetch_k<2>(kbeg + 2 * BK, kend, g.kA);
            }
            if constexpr (!GB::UNIFORM_K) {
              gb.template fetch_k<1>(kbeg + BK, kend, g.kB);
              gb.template fetch_k<2>(kbeg + 2 * BK, kend, g.kB);
            }
          }
          auto pro = [&](auto sc) {
            constexpr int S = decltype(sc)::value;
            if constexpr (!GA::UNIFORM_K)
              ga.template fetch_k_first<(S + NS - 1) % NS>(kbeg + (S + NS - 1) * BK, kend, g.kA, kra, g.step_kA, g.kpad_a);
            if constexpr (!GB::UNIFORM_K)
              gb.template fetch_k_first<(S + NS - 1) % NS>(kbeg + (S + NS - 1) * BK, kend, g.kB, krb, g.step_kB, g.kpad_b);
            __builtin_amdgcn_sched_barrier(0);
            ga.template load<S>(g.A, g.a_type);
            gb.template load<S>(g.B, g.b_type);
            if constexpr (GA::UNIFORM_K) ga.template fetch_k_first<0>(kbeg + (S + 1) * BK, kend, g.kA, kra, g.step_kA, g.kpad_a);
            if constexpr (GB::UNIFORM_K) gb.template fetch_k_first<0>(kbeg + (S + 1) * BK, kend, g.kB, krb, g.step_kB, g.kpad_b);
            __builtin_amdgcn_sched_barrier(0);
          };
          pro(std::integral_constant<int, 0>());
          pro(std::integral_constant<int, 1>());
          if constexpr (NS == 4) {
            pro(std::integral_constant<int, 2>());
            pro(std::integral_constant<int, 3>());
          }
          ga.template stage<0, LA::plane>(smem, t, sa);
          gb.template stage<0, LB::plane>(smem + SA, t, sb);
        }
        __syncthreads();
        // NS stages per trip, all unconditional: stages past kend are all
        // zeros (every lane past kend reads the device zero).
        if constexpr (NS == 2 && BM * BN <= 32 * 32) {
          // The smallest tiles serve short K ranges: whole trips while both
          // stages hold K, then a last stage alone (K not a multiple of two
          // stages) rather than a stage of zeros.  48/96 launch (K = 96,
          // stages of 32): 7.9-8.0 -> 7.6-7.7 us.  The same loop form on
          // every tile cost the rank-4 4096^3 and the permuted all-32
          // kernels 0.5-1% (their K is a multiple of two stages).
          int64_t k0 = kbeg;
          for (; k0 + BK < kend; k0 += 2 * BK) {
            step(k0, std::integral_constant<int, 0>());
            step(k0 + BK, std::integral_constant<int, 1>());
          }
          if (k0 < kend) step(k0, std::integral_constant<int, 0>());
        } else {
          for (int64_t k0 = kbeg; k0 < kend; k0 += NS * BK) {
            step(k0, std::integral_constant<int, 0>());
            step(k0 + BK, std::integral_constant<int, 1>());
            if constexpr (NS == 4) {
              step(k0 + 2 * BK, std::integral_constant<int, 2>());
              step(k0 + 3 * BK, std::integral_constant<int, 3>());
            }
          }
        }
      };
      // Two sets.  Four (loads four stages ahead) measured slower on the
      // 32 x 64 FP64 tile at M = N = K = 1024: 59-60 against 56 us.
      deep(std::integral_constant<int, 2>());
    } else {
      // One stage in flight: stage j + 1 is loaded while stage j multiplies.
      static_assert(k_table_reach(BK, 1) <= kKTablePad, "K table pad shorter than the gather's reach");
      Gather<R, CA, MIXED, BM, BK, AKF> ga;
      Gather<R, CB, MIXED, BN, BK, BKF> gb;
      ga.init(t, m0, g.M, g.a_base, g.mA, g.ramps & RAMP_MA, g.step_mA, g.A, g.a_lo, g.a_bytes);
      gb.init(t, n0, g.N, g.b_base, g.nB, g.ramps & RAMP_NB, g.step_nB, g.B, g.b_lo, g.b_bytes);
      int buf = 0;
      if (kbeg < kend) {
        const bool kra = g.ramps & RAMP_KA, krb = g.ramps & RAMP_KB;
        ga.template fetch_k_first<0>(kbeg, kend, g.kA, kra, g.step_kA, g.kpad_a);
        gb.template fetch_k_first<0>(kbeg, kend, g.kB, krb, g.step_kB, g.kpad_b);
        ga.template load<0>(g.A, g.a_type);
        gb.template load<0>(g.B, g.b_type);
        ga.template fetch_k_first<0>(kbeg + BK, kend, g.kA, kra, g.step_kA, g.kpad_a);
        gb.template fetch_k_first<0>(kbeg + BK, kend, g.kB, krb, g.step_kB, g.kpad_b);
        ga.template stage<0, LA::plane>(smem, t, sa);
        gb.template stage<0, LB::plane>(smem + SA, t, sb);
      }
      __syncthreads();
      for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
        // The next stage's loads and the offsets of the one after, issued
        // unconditionally so that they share the MFMAs' basic block: past
        // the K range the buffer gather reads the K tables' pad entries
        // (zeros) and the generic gather reads the device zero; a stage
        // past the range is staged but never multiplied.
        ga.template load<0>(g.A, g.a_type);
        gb.template load<0>(g.B, g.b_type);
        ga.template fetch_k<0>(k0 + shallow_fetch_ahead() * BK, kend, g.kA);
        gb.template fetch_k<0>(k0 + shallow_fetch_ahead() * BK, kend, g.kB);
        const R *As = smem + buf * (SA + SB);
        const R *Bs = As + SA;
  #pragma unroll
        for (int ks = 0; ks < BK / 4; ++ks) {
          R af[PA][FM], bf[PB][FN];
  #pragma unroll
          for (int p = 0; p < PA; ++p)
  #pragma unroll
            for (int a = 0; a < FM; ++a) af[p][a] = As[p * LA::plane + LA::at(wm * WTM + a * 16 + fr, ks * 4 + fk)];
  #pragma unroll
          for (int p = 0; p < PB; ++p)
  #pragma unroll
            for (int b = 0; b < FN; ++b) bf[p][b] = Bs[p * LB::plane + LB::at(wn * WTN + b * 16 + fr, ks * 4 + fk)];
  #pragma unroll
          for (int a = 0; a < FM; ++a)
  #pragma unroll
            for (int b = 0; b < FN; ++b) {
              if constexpr (!CPLX) {
                acc[0][a][b] = mfma16(af[0][a], bf[0][b], acc[0][a][b]);
              } else if constexpr (DOM == 1) {   // complex A, real B
                acc[0][a][b] = mfma16(af[0][a], bf[0][b], acc[0][a][b]);
                acc[1][a][b] = mfma16(af[PA - 1][a], bf[0][b], acc[1][a][b]);
              } else if constexpr (DOM == 2) {   // real A, complex B
                acc[0][a][b] = mfma16(af[0][a], bf[0][b], acc[0][a][b]);
                acc[1][a][b] = mfma16(af[0][a], bf[PB - 1][b], acc[1][a][b]);
              } else {
                acc[0][a][b] = mfma16(af[0][a], bf[0][b], acc[0][a][b]);
                acc[0][a][b] = mfma16(-af[PA - 1][a], bf[PB - 1][b], acc[0][a][b]);
                acc[P - 1][a][b] = mfma16(af[0][a], bf[PB - 1][b], acc[P - 1][a][b]);
                acc[P - 1][a][b] = mfma16(af[PA - 1][a], bf[0][b], acc[P - 1][a][b]);
              }
            }
        }
        R *Ws = smem + (buf ^ 1) * (SA + SB);
        ga.template stage<0, LA::plane>(Ws, t, sa);
        gb.template stage<0, LB::plane>(Ws + SA, t, sb);
        sched_interleave<(BK / 4) * FM * FN * NMFMA, FM * FN >= 4 ? 6 : 3, FM * FN >= 4 ? 1 : 2>();
        __syncthreads();
        buf ^= 1;
      }
    }

    // Epilogue.  Complex, mixed-type and 128 x 128 accumulators leave no
    // registers to spare: one element at a time.  Otherwise every old-C load
    // of a 16-row block row is issued before its first store.
    constexpr bool HEAVY = CPLX || MIXED || FM * FN * 4 * (int)sizeof(R) / 4 >= 128;
    if (g.k_splits > 1) {
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = m0 + wm * WTM + a * 16 + mfma_row<R>(fk, r);
          if (row >= g.M) continue;
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            const int64_t col = n0 + wn * WTN + b * 16 + fr;
            if (col >= g.N) continue;
            R *part = (R *)g.partials;
            const int64_t plane = g.M * g.N;
            part[((int64_t)split * P) * plane + row * g.N + col] = acc[0][a][b][r];
            if (CPLX) part[((int64_t)split * P + 1) * plane + row * g.N + col] = acc[P - 1][a][b][r];
          }
        }
    } else if (HEAVY) {
      // A block row's C offsets read once; each element's old-C load and
      // store one after the other (registers are short here).
#pragma unroll
      for (int a = 0; a < FM; ++a) {
        int64_t moff[4], noff[FN];
        bool rok[4], cok[FN];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = m0 + wm * WTM + a * 16 + mfma_row<R>(fk, r);
          rok[r] = row < g.M;
          moff[r] = !rok[r] ? 0 : (g.ramps & RAMP_MC) ? row * g.step_mC : g.mC[row];
        }
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          const int64_t col = n0 + wn * WTN + b * 16 + fr;
          cok[b] = col < g.N;
          noff[b] = !cok[b] ? 0 : (g.ramps & RAMP_NC) ? col * g.step_nC : g.nC[col];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int b = 0; b < FN; ++b)
            if (rok[r] && cok[b]) {
              const int64_t idx = g.c_base + moff[r] + noff[b];
              R o, oi;
              load_c_old<R, CPLX, MIXED>(g, idx, o, oi);
              store_c_at<R, CPLX, MIXED>(g, idx, acc[0][a][b][r], CPLX ? acc[P - 1][a][b][r] : (R)0, o, oi);
            }
      }
    } else {
      // Every old-C load of a block row before its first store (none when
      // beta == 0).
#pragma unroll
      for (int a = 0; a < FM; ++a) {
        int64_t moff[4], noff[FN];
        bool rok[4], cok[FN];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t row = m0 + wm * WTM + a * 16 + mfma_row<R>(fk, r);
          rok[r] = row < g.M;
          moff[r] = !rok[r] ? 0 : (g.ramps & RAMP_MC) ? row * g.step_mC : g.mC[row];
        }
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          const int64_t col = n0 + wn * WTN + b * 16 + fr;
          cok[b] = col < g.N;
          noff[b] = !cok[b] ? 0 : (g.ramps & RAMP_NC) ? col * g.step_nC : g.nC[col];
        }
        R o[4][FN], oi[4][FN];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int b = 0; b < FN; ++b) {
            o[r][b] = 0;
            oi[r][b] = 0;
            if (rok[r] && cok[b]) load_c_old<R, CPLX, MIXED>(g, g.c_base + moff[r] + noff[b], o[r][b], oi[r][b]);
          }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int b = 0; b < FN; ++b)
            if (rok[r] && cok[b])
              store_c_at<R, CPLX, MIXED>(g, g.c_base + moff[r] + noff[b], acc[0][a][b][r], (R)0, o[r][b], oi[r][b]);
      }
    }
  }
};

// Split-K: every split of a tile stores its partial sums; this kernel, one
// workgroup per tile launched after the GEMM on the same stream, adds them in
// split order 0, 1, ..., S-1 (a fixed order, so results do not depend on
// scheduling) and applies the epilogue.  The kernel boundary orders the
// partial stores before these loads.  (A last-arriver reduction inside the
// GEMM needs an agent-scope release and acquire per workgroup, which on a
// part with one L2 per XCD writes back and invalidates L2 each time.)
template <typename R, bool CPLX, bool MIXED, int BM, int BN>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const GemmArgs g) {
  constexpr int P = CPLX ? 2 : 1;
  const int64_t tiles_m = (g.M + BM - 1) / BM;
  const int64_t tm = blockIdx.x % tiles_m, tn = blockIdx.x / tiles_m;
  const int64_t plane = g.M * g.N;
  const QS_GLOBAL R *part = (const QS_GLOBAL R *)g.partials;
  // Batches of elements per thread: every partial and old-C load of a batch
  // is issued before its sums and stores.
  constexpr int EPT = BM * BN / 256, CW = (CPLX || MIXED || BM * BN >= 128 * 128) ? 2 : 8, CH = EPT < CW ? EPT : CW;
#pragma unroll 1
  for (int c0 = 0; c0 < EPT; c0 += CH) {
    int64_t idx[CH], cidx[CH];
    bool ok[CH];
    R xr[CH], xi[CH], o[CH], oi[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      const int e = threadIdx.x + 256 * (c0 + j);
      const int64_t row = tm * BM + e / BN, col = tn * BN + e % BN;
      ok[j] = row < g.M && col < g.N;
      idx[j] = ok[j] ? row * g.N + col : 0;
      cidx[j] = ok[j] ? g.c_base + g.mC[row] + g.nC[col] : 0;
      xr[j] = part[idx[j]];
      xi[j] = CPLX ? part[plane + idx[j]] : (R)0;
    }
    for (int s = 1; s < g.k_splits; ++s) {
      R yr[CH], yi[CH];
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        yr[j] = part[((int64_t)s * P) * plane + idx[j]];
        yi[j] = CPLX ? part[((int64_t)s * P + 1) * plane + idx[j]] : (R)0;
      }
#pragma unroll
      for (int j = 0; j < CH; ++j) {
        xr[j] += yr[j];
        if (CPLX) xi[j] += yi[j];
      }
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
      o[j] = 0;
      oi[j] = 0;
      if (ok[j]) load_c_old<R, CPLX, MIXED>(g, cidx[j], o[j], oi[j]);
    }
#pragma unroll
    for (int j = 0; j < CH; ++j)
      if (ok[j]) store_c_at<R, CPLX, MIXED>(g, cidx[j], xr[j], xi[j], o[j], oi[j]);
  }
}

template <typename R, bool CPLX, int BM, int BN, bool AKF, bool BKF, bool MIXED, bool VEC = false, int DOM = 0,
          int BAUX = 0>
__global__ __launch_bounds__(256, 2) void gemm_kernel(const GemmArgs g) {
  typedef TileGemm<R, CPLX, BM, BN, AKF, BKF, MIXED, VEC, DOM, BAUX> TG;
  __shared__ R smem[TG::SMEM];
  const int64_t tiles_m = (g.M + BM - 1) / BM;
  const int64_t tiles_n = (g.N + BN - 1) / BN;
  int64_t tm, tn;
  tile_of(g, blockIdx.x, (int64_t)gridDim.x, tiles_m, tm, tn);
  if (tn >= tiles_n || tm >= tiles_m) return;
  TG::template run<true>(g, tm, tn, blockIdx.y, smem);
}

// A batched launch (single storage type, no split-K) over a grid of
// (tiles, 1, batches): each workgroup runs one tile of one batch coordinate,
// whose offsets move the three bases.  With several tiles per coordinate the
// XCD renumbering runs over the whole grid, so the tiles of one coordinate
// (which share its A rows or B columns) are dealt to the same XCD and its L2
// (256 x 256^3 F64: 192 -> 182 us).  With one tile per coordinate the
// coordinates keep the hardware's order, which measured faster there
// (2048 x 64^3: 56.8 against 58.9 us).
//
// BKO / OCC: K stage depth (0: pick_bk) and workgroups per CU
// (batched_shape).
template <typename R, bool CPLX, int BM, int BN, bool AKF, bool BKF, bool VEC = false, int BKO = 0, int OCC = 2>
__global__ __launch_bounds__(256, OCC) void gemm_batched_kernel(const GemmArgs g) {
  typedef TileGemm<R, CPLX, BM, BN, AKF, BKF, false, VEC, 0, 0, BKO> TG;
  __shared__ R smem[TG::SMEM];
  const int64_t tiles_m = (g.M + BM - 1) / BM;
  const int64_t tiles_n = (g.N + BN - 1) / BN;
  const int64_t tiles = gridDim.x;
  const int64_t lin = tiles > 1 ? xcd_swizzle((int64_t)blockIdx.z * tiles + blockIdx.x, tiles * (int64_t)gridDim.z,
                                             g.xcds)
                                 : (int64_t)blockIdx.z;
  const int64_t bid = lin % tiles, z = lin / tiles;
  const int64_t tm = bid % tiles_m, tn = bid / tiles_m;
  if (tn >= tiles_n) return;
  GemmArgs gz = g;
  const int64_t *o = g.batch_off + 3 * z;
  gz.a_base += o[0];
  gz.a_lo += o[0];
  gz.b_base += o[1];
  gz.b_lo += o[1];
  gz.c_base += o[2];
  TG::template run<true>(gz, tm, tn, 0, smem);
}

// K stage depth and occupancy of batched launches.  Float64 64 x 64 tiles
// take K stages of 16 (37-41 KB of LDS instead of 73-75) and four
// workgroups per CU (at most 128 VGPRs) instead of two: a batched launch's
// tiles are all alike, so two per CU ran their prologues and epilogues in
// step, with nothing to hide them.  Float64 C[i,j,b] := A[i,k,b] B[k,j,b],
// column-major: 2048 x 64^3 58.9 -> 47.8 us; 256 x 256^3 201 -> 175 us on
// this tile, now the best of any tile there (64 x 128 at two per CU: 183).
// Measured and not kept: 64 x 64 with stages of 8 at four or six per CU
// (2048 x 64^3: 49.9-50 us); 128 x 64 and 64 x 128 at three per CU with
// stages of 16 or four per CU with stages of 8 (256 x 256^3: 174-193 us).
template <typename R, bool CPLX, int BM, int BN>
struct batched_shape {
  static constexpr bool SMALL = !CPLX && sizeof(R) == 8 && BM == 64 && BN == 64;
  static constexpr int BKO = SMALL ? 16 : 0;
  static constexpr int OCC = SMALL ? 4 : 2;
};

// A mixed-domain GEMM (complex compute type R, DOM 1 or 2) over the four
// gather directions.
template <typename R, int BM, int BN, bool MIXED, int DOM>
static void launch_dom(const GemmArgs &g, dim3 grid, hipStream_t s) {
  const dim3 block(256);
  if (g.a_kfast && g.b_kfast)
    hipLaunchKernelGGL((gemm_kernel<R, true, BM, BN, true, true, MIXED, false, DOM>), grid, block, 0, s, g);
  else if (g.a_kfast)
    hipLaunchKernelGGL((gemm_kernel<R, true, BM, BN, true, false, MIXED, false, DOM>), grid, block, 0, s, g);
  else if (g.b_kfast)
    hipLaunchKernelGGL((gemm_kernel<R, true, BM, BN, false, true, MIXED, false, DOM>), grid, block, 0, s, g);
  else
    hipLaunchKernelGGL((gemm_kernel<R, true, BM, BN, false, false, MIXED, false, DOM>), grid, block, 0, s, g);
}

template <typename R, bool CPLX, int BM, int BN, bool MIXED>
static hipError_t launch_tile_grid(const GemmArgs &g, dim3 grid, hipStream_t s);

template <typename R, bool CPLX, int BM, int BN, bool MIXED>
static hipError_t launch_tile_mixed(const GemmArgs &g, hipStream_t s) {
  const int64_t tiles = ((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
  if constexpr (!MIXED) {
    if (g.batch_off) {   // batched: at most 65535 batches per launch
      for (int64_t z0 = 0; z0 < g.batch; z0 += 65535) {
        GemmArgs part = g;
        part.batch_off = g.batch_off + 3 * z0;
        const dim3 bgrid((unsigned)tiles, 1, (unsigned)std::min<int64_t>(65535, g.batch - z0));
        constexpr int BKO = batched_shape<R, CPLX, BM, BN>::BKO, OCC = batched_shape<R, CPLX, BM, BN>::OCC;
        // Paired loads as in the one-contraction kernel: the batch offsets
        // move the bases, not the tables, so the pairs stay adjacent.
        if constexpr (!CPLX && BM * BN <= 64 * 128) {
          if (g.vec) {
            if (g.a_kfast && g.b_kfast)
              hipLaunchKernelGGL((gemm_batched_kernel<R, CPLX, BM, BN, true, true, true, BKO, OCC>), bgrid, dim3(256), 0,
                                 s, part);
            else if (g.a_kfast)
              hipLaunchKernelGGL((gemm_batched_kernel<R, CPLX, BM, BN, true, false, true, BKO, OCC>), bgrid, dim3(256), 0,
                                 s, part);
            else if (g.b_kfast)
              hipLaunchKernelGGL((gemm_batched_kernel<R, CPLX, BM, BN, false, true, true, BKO, OCC>), bgrid, dim3(256), 0,
                                 s, part);
            else
              hipLaunchKernelGGL((gemm_batched_kernel<R, CPLX, BM, BN, false, false, true, BKO, OCC>), bgrid, dim3(256), 0,
                                 s, part);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            continue;
          }
        }
        if (g.a_kfast && g.b_kfast)
          hipLaunchKernelGGL((gemm_batched_kernel<R, CPLX, BM, BN, true, true, false, BKO, OCC>), bgrid, dim3(256), 0, s,
                             part);
        else if (g.a_kfast)
          hipLaunchKernelGGL((gemm_batched_kernel<R, CPLX, BM, BN, true, false, false, BKO, OCC>), bgrid, dim3(256), 0, s,
                             part);
        else if (g.b_kfast)
          hipLaunchKernelGGL((gemm_batched_kernel<R, CPLX, BM, BN, false, true, false, BKO, OCC>), bgrid, dim3(256), 0, s,
                             part);
        else
          hipLaunchKernelGGL((gemm_batched_kernel<R, CPLX, BM, BN, false, false, false, BKO, OCC>), bgrid, dim3(256), 0, s,
                             part);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
      }
      return hipSuccess;
    }
  }
  dim3 grid((unsigned)tiles, (unsigned)g.k_splits), block(256);
  GemmArgs gm = g;
  set_tile_map(gm, BM, BN, (g.M + BM - 1) / BM, (g.N + BN - 1) / BN, device_facts().cus);
  return launch_tile_grid<R, CPLX, BM, BN, MIXED>(gm, grid, s);
}

template <typename R, bool CPLX, int BM, int BN, bool MIXED>
static hipError_t launch_tile_grid(const GemmArgs &g, dim3 grid, hipStream_t s) {
  const int64_t tiles = grid.x;
  const dim3 block(256);
  // Paired loads: real, one storage type, tiles on the two-stage pipeline.
  if constexpr (!MIXED && !CPLX && BM * BN <= 64 * 128) {
    if (g.vec) {
      // One row of 32-row tiles (M <= 32, B read in place): each B element
      // is read by one workgroup, so B's loads are non-temporal.
      const bool streamed = BM == 32 && g.M <= 32 && g.k_splits == 1;
      if (streamed) {
        if constexpr (BM == 32) {
          if (g.a_kfast && g.b_kfast)
            hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, true, true, false, true, 0, 2>), grid, block, 0, s, g);
          else if (g.a_kfast)
            hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, true, false, false, true, 0, 2>), grid, block, 0, s, g);
          else if (g.b_kfast)
            hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, false, true, false, true, 0, 2>), grid, block, 0, s, g);
          else
            hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, false, false, false, true, 0, 2>), grid, block, 0, s, g);
        }
      } else if (g.a_kfast && g.b_kfast)
        hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, true, true, false, true>), grid, block, 0, s, g);
      else if (g.a_kfast)
        hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, true, false, false, true>), grid, block, 0, s, g);
      else if (g.b_kfast)
        hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, false, true, false, true>), grid, block, 0, s, g);
      else
        hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, false, false, false, true>), grid, block, 0, s, g);
      if (g.k_splits > 1) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((splitk_reduce_kernel<R, CPLX, MIXED, BM, BN>), dim3((unsigned)tiles), block, 0, s, g);
      }
      return hipGetLastError();
    }
  }
  // Mixed domains: the real operand's plane and half the MFMAs left out; with
  // every storage type native (dom_buf), both gathers take buffer loads.
  if constexpr (MIXED && CPLX) {
    if (g.dom == 1 || g.dom == 2) {
      if (g.dom == 1 && g.dom_buf)
        launch_dom<R, BM, BN, false, 1>(g, grid, s);
      else if (g.dom == 1)
        launch_dom<R, BM, BN, true, 1>(g, grid, s);
      else if (g.dom_buf)
        launch_dom<R, BM, BN, false, 2>(g, grid, s);
      else
        launch_dom<R, BM, BN, true, 2>(g, grid, s);
      if (g.k_splits > 1) {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        if (g.dom_buf)
          hipLaunchKernelGGL((splitk_reduce_kernel<R, CPLX, false, BM, BN>), dim3((unsigned)tiles), block, 0, s, g);
        else
          hipLaunchKernelGGL((splitk_reduce_kernel<R, CPLX, true, BM, BN>), dim3((unsigned)tiles), block, 0, s, g);
      }
      return hipGetLastError();
    }
  }
  if (g.a_kfast && g.b_kfast)
    hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, true, true, MIXED>), grid, block, 0, s, g);
  else if (g.a_kfast)
    hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, true, false, MIXED>), grid, block, 0, s, g);
  else if (g.b_kfast)
    hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, false, true, MIXED>), grid, block, 0, s, g);
  else
    hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, false, false, MIXED>), grid, block, 0, s, g);
  if (g.k_splits > 1) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((splitk_reduce_kernel<R, CPLX, MIXED, BM, BN>), dim3((unsigned)tiles), block, 0, s, g);
  }
  return hipGetLastError();
}

template <typename R, bool CPLX, int BM, int BN>
static hipError_t launch_tile(const GemmArgs &g, hipStream_t s) {
  return g.mixed ? launch_tile_mixed<R, CPLX, BM, BN, true>(g, s) : launch_tile_mixed<R, CPLX, BM, BN, false>(g, s);
}

// One compute type's launcher: every tile shape of TileId.
template <typename R, bool CPLX>
static hipError_t launch_gemm_typed(const GemmArgs &g, hipStream_t s) {
  switch (g.tile) {
    case TILE_128x128: return launch_tile<R, CPLX, 128, 128>(g, s);
    case TILE_128x64: return launch_tile<R, CPLX, 128, 64>(g, s);
    case TILE_64x128: return launch_tile<R, CPLX, 64, 128>(g, s);
    case TILE_32x64: return launch_tile<R, CPLX, 32, 64>(g, s);
    case TILE_32x32: return launch_tile<R, CPLX, 32, 32>(g, s);
    default: return launch_tile<R, CPLX, 64, 64>(g, s);
  }
}

}  // namespace qs
