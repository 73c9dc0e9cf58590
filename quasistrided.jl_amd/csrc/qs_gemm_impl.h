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

// Element `idx` of storage `p` as the compute type.  MIXED: the storage type
// `t` is a runtime value (uniform across the workgroup); otherwise it is the
// compute type itself.
template <typename R, bool CPLX, bool MIXED>
__device__ __forceinline__ void load_elem(const void *__restrict__ p, int t, int64_t idx, R &re, R &im) {
  if (!MIXED) {
    if (!CPLX) {
      re = ((const R *)p)[idx];
      im = 0;
    } else {
      re = ((const R *)p)[2 * idx];
      im = ((const R *)p)[2 * idx + 1];
    }
    return;
  }
  switch (t) {
    case QS_F32: re = (R)((const float *)p)[idx]; im = 0; break;
    case QS_F64: re = (R)((const double *)p)[idx]; im = 0; break;
    case QS_C32: re = (R)((const float *)p)[2 * idx]; im = (R)((const float *)p)[2 * idx + 1]; break;
    default: re = (R)((const double *)p)[2 * idx]; im = (R)((const double *)p)[2 * idx + 1]; break;
  }
}

template <typename R, bool CPLX, bool MIXED>
__device__ __forceinline__ void store_elem(void *__restrict__ p, int t, int64_t idx, R re, R im) {
  if (!MIXED) {
    if (!CPLX) {
      ((R *)p)[idx] = re;
    } else {
      ((R *)p)[2 * idx] = re;
      ((R *)p)[2 * idx + 1] = im;
    }
    return;
  }
  switch (t) {
    case QS_F32: ((float *)p)[idx] = (float)re; break;
    case QS_F64: ((double *)p)[idx] = (double)re; break;
    case QS_C32: ((float *)p)[2 * idx] = (float)re; ((float *)p)[2 * idx + 1] = (float)im; break;
    default: ((double *)p)[2 * idx] = (double)re; ((double *)p)[2 * idx + 1] = (double)im; break;
  }
}

// The epilogue at one element of C: `alpha*x + beta*C` for the beta case,
// with Base's Complex expression trees (stores.jl:91-124; the library is
// built with -ffp-contract=off so the unfused products stay unfused).
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

// K depth of one LDS stage: the largest power of two in [4, 32] whose double-
// buffered A and B tiles fit 48 KiB.
template <typename R, bool CPLX, int BM, int BN>
constexpr int pick_bk() {
  int planes = CPLX ? 2 : 1;
  int bk = (48 * 1024) / (planes * (BM + BN + 32) * (int)sizeof(R) * 2);
  int p = 4;
  while (p * 2 <= bk && p * 2 <= 32) p *= 2;
  return p;
}

// XCD-aware workgroup order: the hardware deals workgroups round-robin over
// the 8 XCDs, so consecutive ids land on different L2s.  Renumber so each XCD
// walks a contiguous run of tiles (which share B columns and A rows).
__device__ __forceinline__ int64_t xcd_swizzle(int64_t bid, int64_t nwg) {
  if (nwg % 8 != 0) return bid;
  return (bid % 8) * (nwg / 8) + bid / 8;
}

template <typename R, bool CPLX, int BM, int BN, bool AKF, bool BKF, bool MIXED>
__global__ __launch_bounds__(256) void gemm_kernel(const GemmArgs g) {
  constexpr int P = CPLX ? 2 : 1;
  constexpr int BK = pick_bk<R, CPLX, BM, BN>();
  constexpr int PAD = 16;
  constexpr int LDA = BM + PAD, LDB = BN + PAD;
  constexpr int WTM = BM / 2, WTN = BN / 2;       // 2 x 2 waves
  constexpr int FM = WTM / 16, FN = WTN / 16;     // MFMA blocks per wave
  constexpr int NLA = (BM * BK + 255) / 256, NLB = (BN * BK + 255) / 256;
  static_assert(256 % BM == 0 || BM % 256 == 0, "BM");
  static_assert(BK % 4 == 0, "BK");
  typedef typename Acc4<R>::type V;

  __shared__ R As[2][P][BK][LDA];
  __shared__ R Bs[2][P][BK][LDB];

  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  const int64_t tiles_m = (g.M + BM - 1) / BM;
  const int64_t tiles_n = (g.N + BN - 1) / BN;
  const int64_t bid = xcd_swizzle(blockIdx.x, (int64_t)gridDim.x);
  const int64_t tm = bid % tiles_m, tn = bid / tiles_m;
  if (tn >= tiles_n) return;
  const int64_t m0 = tm * BM, n0 = tn * BN;
  const int split = blockIdx.y;
  const int64_t kbeg = (int64_t)split * g.k_chunk;
  const int64_t kend = min(g.K, kbeg + g.k_chunk);

  // Per-thread gather coordinates, fixed over the K loop.
  int64_t a_row[NLA];
  int a_kk[NLA];
  bool a_ok[NLA];
#pragma unroll
  for (int r = 0; r < NLA; ++r) {
    const int e = t + 256 * r;
    const int i = AKF ? e / BK : e % BM;
    a_kk[r] = AKF ? e % BK : e / BM;
    const int64_t gi = m0 + i;
    a_ok[r] = (e < BM * BK) && gi < g.M;
    a_row[r] = a_ok[r] ? g.a_base + g.mA[gi] : 0;
  }
  int64_t b_col[NLB];
  int b_kk[NLB];
  bool b_ok[NLB];
#pragma unroll
  for (int r = 0; r < NLB; ++r) {
    const int e = t + 256 * r;
    const int j = BKF ? e / BK : e % BN;
    b_kk[r] = BKF ? e % BK : e / BN;
    const int64_t gj = n0 + j;
    b_ok[r] = (e < BN * BK) && gj < g.N;
    b_col[r] = b_ok[r] ? g.b_base + g.nB[gj] : 0;
  }

  R ra[NLA][P], rb[NLB][P];
  const R sa = g.conj_a ? (R)-1 : (R)1, sb = g.conj_b ? (R)-1 : (R)1;

  auto gather = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < NLA; ++r) {
      const int64_t gk = k0 + a_kk[r];
      R re = 0, im = 0;
      if (a_ok[r] && gk < kend) load_elem<R, CPLX, MIXED>(g.A, g.a_type, a_row[r] + g.kA[gk], re, im);
      ra[r][0] = re;
      if (CPLX) ra[r][P - 1] = sa * im;
    }
#pragma unroll
    for (int r = 0; r < NLB; ++r) {
      const int64_t gk = k0 + b_kk[r];
      R re = 0, im = 0;
      if (b_ok[r] && gk < kend) load_elem<R, CPLX, MIXED>(g.B, g.b_type, b_col[r] + g.kB[gk], re, im);
      rb[r][0] = re;
      if (CPLX) rb[r][P - 1] = sb * im;
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int r = 0; r < NLA; ++r) {
      const int e = t + 256 * r;
      if (e < BM * BK) {
        const int i = AKF ? e / BK : e % BM;
#pragma unroll
        for (int p = 0; p < P; ++p) As[buf][p][a_kk[r]][i] = ra[r][p];
      }
    }
#pragma unroll
    for (int r = 0; r < NLB; ++r) {
      const int e = t + 256 * r;
      if (e < BN * BK) {
        const int j = BKF ? e / BK : e % BN;
#pragma unroll
        for (int p = 0; p < P; ++p) Bs[buf][p][b_kk[r]][j] = rb[r][p];
      }
    }
  };

  V acc[P][FM][FN];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
      for (int b = 0; b < FN; ++b) acc[p][a][b] = V{0, 0, 0, 0};

  const int fr = lane & 15, fk = lane >> 4;
  int buf = 0;
  if (kbeg < kend) {
    gather(kbeg);
    stage(0);
  }
  __syncthreads();
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    const bool more = k0 + BK < kend;
    if (more) gather(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      R af[P][FM], bf[P][FN];
#pragma unroll
      for (int p = 0; p < P; ++p) {
#pragma unroll
        for (int a = 0; a < FM; ++a) af[p][a] = As[buf][p][ks * 4 + fk][wm * WTM + a * 16 + fr];
#pragma unroll
        for (int b = 0; b < FN; ++b) bf[p][b] = Bs[buf][p][ks * 4 + fk][wn * WTN + b * 16 + fr];
      }
#pragma unroll
      for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int b = 0; b < FN; ++b) {
          if (!CPLX) {
            acc[0][a][b] = mfma16(af[0][a], bf[0][b], acc[0][a][b]);
          } else {
            acc[0][a][b] = mfma16(af[0][a], bf[0][b], acc[0][a][b]);
            acc[0][a][b] = mfma16(-af[P - 1][a], bf[P - 1][b], acc[0][a][b]);
            acc[P - 1][a][b] = mfma16(af[0][a], bf[P - 1][b], acc[P - 1][a][b]);
            acc[P - 1][a][b] = mfma16(af[P - 1][a], bf[0][b], acc[P - 1][a][b]);
          }
        }
    }
    if (more) stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  // Epilogue.
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
        const R xr = acc[0][a][b][r];
        const R xi = CPLX ? acc[P - 1][a][b][r] : (R)0;
        if (g.k_splits == 1) {
          store_c<R, CPLX, MIXED>(g, row, col, xr, xi);
        } else {
          R *part = (R *)g.partials;
          const int64_t plane = g.M * g.N;
          part[((int64_t)split * P) * plane + row * g.N + col] = xr;
          if (CPLX) part[((int64_t)split * P + 1) * plane + row * g.N + col] = xi;
        }
      }
    }
}

// Split-K reduction: partial sums added in split order 0, 1, ..., S-1 (a
// fixed order, so results do not depend on scheduling), then the epilogue.
template <typename R, bool CPLX, bool MIXED>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const GemmArgs g) {
  constexpr int P = CPLX ? 2 : 1;
  const int64_t plane = g.M * g.N;
  const R *part = (const R *)g.partials;
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < plane; e += (int64_t)gridDim.x * 256) {
    R xr = part[e], xi = CPLX ? part[plane + e] : (R)0;
    for (int s = 1; s < g.k_splits; ++s) {
      xr += part[((int64_t)s * P) * plane + e];
      if (CPLX) xi += part[((int64_t)s * P + 1) * plane + e];
    }
    const int64_t row = e / g.N, col = e % g.N;
    store_c<R, CPLX, MIXED>(g, row, col, xr, xi);
  }
}

template <typename R, bool CPLX, int BM, int BN, bool MIXED>
static hipError_t launch_tile_mixed(const GemmArgs &g, hipStream_t s) {
  const int64_t tiles = ((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
  dim3 grid((unsigned)tiles, (unsigned)g.k_splits), block(256);
  if (g.a_kfast && g.b_kfast)
    hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, true, true, MIXED>), grid, block, 0, s, g);
  else if (g.a_kfast)
    hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, true, false, MIXED>), grid, block, 0, s, g);
  else if (g.b_kfast)
    hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, false, true, MIXED>), grid, block, 0, s, g);
  else
    hipLaunchKernelGGL((gemm_kernel<R, CPLX, BM, BN, false, false, MIXED>), grid, block, 0, s, g);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess || g.k_splits == 1) return err;
  const int64_t n = g.M * g.N;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL((splitk_reduce_kernel<R, CPLX, MIXED>), dim3((unsigned)blocks), dim3(256), 0, s, g);
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
    default: return launch_tile<R, CPLX, 64, 64>(g, s);
  }
}

}  // namespace qs
