// Split-precision FP64 / FP32 contraction on the INT8 MFMA units by the
// Chinese remainder theorem (Ozaki scheme II: Ozaki, Uchino, Imamura, "Ozaki
// Scheme II: A GEMM-oriented emulation of floating-point matrix
// multiplication using an integer modular technique", 2025).
//
//   1. Scales.  For every row i of A (over K) and column j of B the 2-norm
//      n_i (scaled sum of squares, full double range) gives a scale
//      s_i = T / n_i, kept as a mantissa q_i in (0.5, 2] and a power of two.
//      A'_ik = rint(A_ik s_i) and B'_kj = rint(B_kj t_j) are integers whose
//      row / column norms are at most T (+ sqrt(K) / 2 from the rounding), so
//      by Cauchy-Schwarz every integer product C'_ij = sum_k A'_ik B'_kj has
//      |C'_ij| < P / 2, with P = m_1 ... m_n a product of pairwise coprime
//      moduli <= 256.  No bound depends on K.
//   2. Residues.  For every modulus m_l: A'_ik y_l mod m_l and B'_kj mod m_l
//      in the symmetric range [-m/2, m/2), which is INT8 (y_l = (P/m_l)^-1
//      mod m_l, folded into A).  A' is split into two base-2^32 digits so
//      every residue is computed exactly in FP64 with a correctly rounded
//      symmetric quotient (no fix-up steps).  Residues are
//      stored in MFMA fragment order: 16 rows x 64 K per 1 KiB block, lane l
//      holding row l % 16, K bytes 16 (l / 16) .. +15.
//   3. n INT8 GEMMs on v_mfma_i32_16x16x64_i8 with exact INT32 accumulation
//      (up to a padded K of 130816 the sums stay below 2^31; beyond it they
//      are reduced mod m every 4096 K, so K is unbounded); the epilogue stores
//      U_l = acc mod m_l as a byte.
//   4. Reconstruction.  C' = sum_l U_l Q_l - q P with Q_l = P / m_l and
//      q = rint(sum_l U_l Q_l / P).  Q_l = Q1_l + Q2_l where Q1_l is a
//      multiple of 2^g short enough that sum_l U_l Q1_l and q P1 are exact in
//      FP64, so C' comes out to within a few units of its last bit.  Then
//      C = alpha C' / (s_i t_j) + beta C through C's offset tables, one
//      rounding to C's type.
//
// The number of moduli follows from the requested relative error (default
// 1e-13 for F64, 1e-6 for F32) and K; the sampled error report (Dot2 of
// Ogita-Rump-Oishi 2005 on the same gathered operands) measures what was
// delivered.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "qs_crt.h"
#include "qs_gemm_impl.h"
#include "qs_plan.h"

namespace qs {

typedef int qs_v4i __attribute__((ext_vector_type(4)));

static constexpr int kMaxModuli = 15;    // (n + 1) P < 2^128
static constexpr int kTile = 256;        // GEMM workgroup tile (M and N)

// Scaled sum of squares: value s * 2^(2e).
struct Ssq {
  double s;
  int e, pad;
};

struct RowScale {
  double q;    // scale mantissa: A'_ik = rint(A_ik 2^-sh q)  (0 for a zero or non-finite row)
  double iq;   // 1 / q_i (NaN for a row holding Inf / NaN)
  int sh;      // power of two that undoes the scale
  int pad;
};

// Residue weights and reconstruction constants (host-computed, passed by value).
struct CrtConst {
  int n;
  double mf[kMaxModuli], invm[kMaxModuli];
  double cst[kMaxModuli];                 // 0.5 - 128 / m
  double w[2][kMaxModuli][2];             // digit weights 1, 2^32 mod m: [0] A (y folded in), [1] B
  double q1[kMaxModuli], q2[kMaxModuli];
  double p1, p2, invp;
};

struct CrtOperand {
  const void *X;
  int type, kfast;
  int64_t base;
  const int64_t *rowtab, *ktab;
  int64_t R, K, Rp, Kp;   // rows (M of A, N of B), K, padded to 256 and 128
  int side;               // 0: A, 1: B
  RowScale *scale;        // [R]
  int8_t *res;            // [n][Rp/16][Kp/64][1024]
  double Tm;              // T = Tm * 2^eT
  int eT;
};

__device__ __forceinline__ Ssq ssq_merge(Ssq a, Ssq b) {
  Ssq r;
  r.e = max(a.e, b.e);
  r.s = ldexp(a.s, 2 * (a.e - r.e)) + ldexp(b.s, 2 * (b.e - r.e));
  r.pad = 0;
  return r;
}

__device__ __forceinline__ int ssq_expo(double x) {
  const double ax = fabs(x);
  return (ax > 0.0 && ax <= DBL_MAX) ? ilogb(ax) : -2000;
}

// The row's scale from its norm (scaled sum of squares): q = T / f and 2^sh
// with norm = f 2^(e + en0).
__device__ __forceinline__ RowScale row_scale_of(const CrtOperand &o, const Ssq a) {
  RowScale rs{0.0, 1.0, 0, 0};
  if (!(a.s <= DBL_MAX)) {
    rs.iq = __builtin_nan("");
  } else if (a.s > 0.0) {
    int en0;
    const double f = frexp(sqrt(a.s), &en0);   // f in [0.5, 1)
    rs.q = o.Tm / f;
    rs.iq = f / o.Tm;
    rs.sh = (en0 + a.e) - o.eT;
  }
  return rs;
}

// The row norms and scales in one pass (no partial sums, no separate merge):
// each workgroup of four waves takes a few rows over all of K, with its loads
// batched ahead of their arithmetic, and writes the rows' scales itself.
// Each batch's squares are scaled by the batch's largest exponent (per lane)
// and merged into the lane's running sum; lanes and waves are merged by fixed
// trees, so the scales are reproducible bit for bit.
//   K fast: one row per wave (4 per workgroup), lanes along K, 32 K steps of
//           64 per batch.
//   Otherwise: 16 rows per workgroup; lanes 16 rows x 4 K offsets (a load
//           covers 4 runs of 16 elements along the rows), wave w takes K steps
//           4 w + 16 i, 32 per batch.
constexpr int kRowsKFast = 4, kRowsMFast = 16;

template <typename T>
__global__ __launch_bounds__(256) void crt_rownorm_kernel(const CrtOperand o) {
  constexpr int NB = 32;
  __shared__ Ssq red[4][16];
  const QS_GLOBAL T *X = (const QS_GLOBAL T *)o.X;
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t safe = o.base + o.rowtab[0] + o.ktab[0];
  auto batch = [](const double (&x)[NB], Ssq &acc) {
    int e = -2000;
#pragma unroll
    for (int j = 0; j < NB; ++j) e = max(e, ssq_expo(x[j]));
    double sq = 0.0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const double y = ldexp(fabs(x[j]), -e);
      sq = __builtin_fma(y, y, sq);
    }
    acc = ssq_merge(acc, Ssq{sq, e, 0});
  };
  Ssq acc{0.0, -2000, 0};
  if (o.kfast) {
    const int64_t r = (int64_t)blockIdx.x * kRowsKFast + wave;
    const bool rin = r < o.R;
    const int64_t rb = o.base + o.rowtab[rin ? r : 0];
    for (int64_t k0 = 0; k0 < o.K; k0 += 64 * NB) {
      int64_t ko[NB];
      bool kin[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int64_t k = k0 + 64 * j + lane;
        kin[j] = k < o.K;
        ko[j] = o.ktab[kin[j] ? k : 0];
      }
      double x[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) x[j] = (double)X[rin && kin[j] ? rb + ko[j] : safe];
#pragma unroll
      for (int j = 0; j < NB; ++j) x[j] = rin && kin[j] ? x[j] : 0.0;
      batch(x, acc);
    }
#pragma unroll
    for (int off = 1; off <= 32; off <<= 1) {
      const Ssq b{__shfl_xor(acc.s, off, 64), __shfl_xor(acc.e, off, 64), 0};
      acc = ssq_merge(acc, b);   // commutative: both lanes of a pair form the same value
    }
    if (lane == 0 && rin) o.scale[r] = row_scale_of(o, acc);
  } else {
    const int64_t r0 = (int64_t)blockIdx.x * kRowsMFast;
    const int rr = lane & 15, ks = lane >> 4;
    const int64_t r = r0 + rr;
    const bool rin = r < o.R;
    const int64_t rbase = o.base + o.rowtab[rin ? r : 0];
    for (int64_t k0 = 0; k0 < o.K; k0 += 16 * NB) {
      int64_t ko[NB];
      bool kin[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int64_t k = k0 + 16 * j + 4 * wave + ks;
        kin[j] = k < o.K;
        ko[j] = o.ktab[kin[j] ? k : 0];
      }
      double x[NB];
#pragma unroll
      for (int j = 0; j < NB; ++j) x[j] = (double)X[rin && kin[j] ? rbase + ko[j] : safe];
#pragma unroll
      for (int j = 0; j < NB; ++j) x[j] = rin && kin[j] ? x[j] : 0.0;
      batch(x, acc);
    }
    // lanes rr, rr + 16, rr + 32, rr + 48 hold one row
#pragma unroll
    for (int off = 16; off <= 32; off <<= 1) {
      const Ssq b{__shfl_xor(acc.s, off, 64), __shfl_xor(acc.e, off, 64), 0};
      acc = ssq_merge(acc, b);
    }
    if (lane < 16) red[wave][lane] = acc;
    __syncthreads();
    if (t < 16) {
      Ssq a = red[0][t];
#pragma unroll
      for (int w = 1; w < 4; ++w) a = ssq_merge(a, red[w][t]);
      if (r0 + t < o.R) o.scale[r0 + t] = row_scale_of(o, a);
    }
  }
}

// Pass 2: scale, round and reduce one 64-row x 64-K tile to one residue
// fragment column per modulus.
struct CrtResidueLds {
  double tile[64][65];
  double sq[64];
  int ssh[64];
};

template <typename T>
__device__ __forceinline__ void crt_residue_body(const CrtOperand &o, const CrtConst &c, const T *__restrict__ X,
                                                 const int64_t *__restrict__ rowtab, const int64_t *__restrict__ ktab,
                                                 const RowScale *__restrict__ scale, int8_t *__restrict__ res,
                                                 CrtResidueLds &sm, int64_t bx) {
  const QS_GLOBAL T *Xg = (const QS_GLOBAL T *)X;
  auto &tile = sm.tile;
  auto &sq = sm.sq;
  auto &ssh = sm.ssh;
  const int t = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t r0 = bx * 64, k0 = (int64_t)blockIdx.y * 64;
  // 16 independent loads per thread, issued before the row scales are read.
  double x[16];
  if (o.kfast) {
    // Lanes along K (tk = lane), rows wave + 4 it.
    const int64_t k = k0 + (t & 63);
    const bool kin = k < o.K;
    const int64_t ko = kin ? ktab[k] : 0;
#pragma unroll
    for (int it = 0; it < 16; ++it) {
      const int64_t r = r0 + wave + 4 * it;
      x[it] = 0.0;
      if (kin && r < o.R) x[it] = (double)Xg[o.base + rowtab[r] + ko];
    }
  } else {
    // Lanes along the rows (tr = lane), K steps wave + 4 it.
    const int64_t r = r0 + (t & 63);
    const bool rin = r < o.R;
    const int64_t rbase = rin ? o.base + rowtab[r] : 0;
#pragma unroll
    for (int it = 0; it < 16; ++it) {
      const int64_t k = k0 + wave + 4 * it;
      x[it] = 0.0;
      if (rin && k < o.K) x[it] = (double)Xg[rbase + ktab[k]];
    }
  }
  if (t < 64) {
    // Rows past R (padding) get q = 0: their integers are zero.
    const RowScale rs = r0 + t < o.R ? scale[r0 + t] : RowScale{0.0, 1.0, 0, 0};
    sq[t] = rs.q;
    ssh[t] = -rs.sh;
  }
  __syncthreads();
  // |x 2^sh| <= f 2^eT < T: no overflow; q = 0 zeroes empty and padding rows.
  if (o.kfast) {
#pragma unroll
    for (int it = 0; it < 16; ++it) {
      const int tr = wave + 4 * it;
      tile[tr][t & 63] = rint(ldexp(x[it], ssh[tr]) * sq[tr]);
    }
  } else {
    const int tr = t & 63;
    const double qv = sq[tr];
    const int sh = ssh[tr];
#pragma unroll
    for (int it = 0; it < 16; ++it) tile[tr][wave + 4 * it] = rint(ldexp(x[it], sh) * qv);
  }
  __syncthreads();
  const int f = t >> 6, lane = t & 63;
  const int row = 16 * f + (lane & 15), kk = 16 * (lane >> 4);
  // Two digits: x = d1 2^32 + d0, d0 in [0, 2^32), |d1| < 2^31.
  double d0[16], d1[16];
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const double xv = tile[row][kk + v];
    d1[v] = floor(xv * 0x1p-32);
    d0[v] = __builtin_fma(d1[v], -0x1p32, xv);
  }
  const int64_t RB = o.Rp / 16, KB = o.Kp / 64;
  const int64_t frag = (r0 / 16 + f) * KB + k0 / 64;
  for (int l = 0; l < c.n; ++l) {
    const double m = c.mf[l], im = c.invm[l], cst = c.cst[l];
    const double w0 = c.w[o.side][l][0], w1 = c.w[o.side][l][1];
    int word[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      unsigned int packed = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int v = 4 * q + b;
        // s = x w + 128 exactly (|s| < 2^41).  The symmetric quotient
        // floor((s - 128) / m + 1/2) is exact: for odd m that value is never
        // within 1 / (2m) of an integer, and for m = 256 every step is exact.
        const double sv = __builtin_fma(d1[v], w1, __builtin_fma(d0[v], w0, 128.0));
        const double qq = floor(__builtin_fma(sv, im, cst));
        const double u = __builtin_fma(-qq, m, sv);    // residue + 128, in [0, 256)
        packed = __builtin_amdgcn_cvt_pk_u8_f32((float)u, b, packed);
      }
      word[q] = (int)(packed ^ 0x80808080u);            // two's complement of the residue
    }
    const qs_v4i out = {word[0], word[1], word[2], word[3]};
    __builtin_nontemporal_store(out, (QS_GLOBAL qs_v4i *)(res + ((int64_t)l * RB * KB + frag) * 1024 + lane * 16));
  }
}

template <typename T>
__global__ __launch_bounds__(256) void crt_residue_kernel(const CrtOperand o, const CrtConst c) {
  __shared__ CrtResidueLds sm;
  crt_residue_body<T>(o, c, (const T *)o.X, o.rowtab, o.ktab, o.scale, o.res, sm, blockIdx.x);
}

// The residue GEMMs, for padded K above kCrtMaxKpUnreduced: one 256 x 256
// tile of one modulus per workgroup, 8 waves (2 x 4, each 128 x 64 = 8 x 4
// MFMA blocks of v_mfma_i32_16x16x64_i8), K in stages of one fragment (64
// bytes).  A's 16 row blocks of a stage (16 KiB) go through LDS,
// double-buffered, loaded two stages ahead; each wave reads its 4 B
// fragments straight from global memory into registers one stage ahead (a
// contiguous 1 KiB per wave in the residue layout, wanted by the 2 waves of a
// column, so L1 absorbs the repeat).  The INT32 accumulators are reduced mod
// m every 4096 K.
struct CrtGemmArgs {
  const int8_t *ra, *rb;
  unsigned int *u;           // [Mp/16][Np/16][n][64]
  int64_t MB, NB, KB;        // Mp/16, Np/16, Kp/64
  int64_t tiles_m, tiles_n;
  int n;
  int xcds;
  int m[kMaxModuli];
  float invm[kMaxModuli];
  double invmd[kMaxModuli];
};

// Longest padded K the four-wave kernel takes: every INT32 sum of Kp
// products of two residues (|r| <= 128) stays below 2^31, so no reduction
// is needed inside the K loop.
static constexpr int64_t kCrtMaxKpUnreduced = 130816;

__device__ __forceinline__ int crt_mod(int x, int m, float invm) {
  // |x| < 2^27: the float quotient is off by at most one.
  const int q = (int)floorf((float)x * invm);
  int r = x - q * m;
  r = r < 0 ? r + m : r;
  r = r >= m ? r - m : r;
  return r;
}

// Padded K above kCrtMaxKpUnreduced.  WM x WN waves, each TI x TJ blocks of
// 16 x 16: a (16 WM TI) x (16 WN TJ) tile of one modulus per workgroup.  B is
// loaded BS - 1 stages ahead into BS rotating register sets (2 measured ahead
// of 4).
template <int WM, int WN, int TI, int TJ, int BS>
__global__ __launch_bounds__(64 * WM * WN, 1) void crt_gemm_longk_kernel(const CrtGemmArgs a) {
  constexpr int NW = WM * WN, FA = WM * TI, NLA = FA / NW;
  static_assert(FA % NW == 0, "A stage slots must split evenly over the waves");
  static_assert(BS == 2 || BS == 4, "B register sets: 2 or 4 (the loop is unrolled by 4)");
  __shared__ __attribute__((aligned(16))) int8_t smem[2][FA * 1024];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = wave % WM, wn = wave / WM;
  const int64_t tiles = a.tiles_m * a.tiles_n;
  const int64_t bid = xcd_swizzle(blockIdx.x, (int64_t)gridDim.x, a.xcds);
  const int l = (int)(bid / tiles);
  const int64_t tt = bid % tiles;
  const int64_t grp = tt / (4 * a.tiles_n), in = tt % (4 * a.tiles_n);
  const int64_t rows = min((int64_t)4, a.tiles_m - 4 * grp);
  const int64_t tm = 4 * grp + in % rows, tn = in / rows;
  const int64_t mb0 = tm * FA, nb0 = tn * (WN * TJ);
  const int64_t KB = a.KB;
  const int8_t *A = a.ra + ((int64_t)l * a.MB + mb0) * KB * 1024 + lane * 16;
  const int8_t *Bw = a.rb + ((int64_t)l * a.NB + nb0 + wn * TJ) * KB * 1024 + lane * 16;
  const int m = a.m[l];
  const float invm = a.invm[l];

  // Stage indices past the end are clamped to the last stage (loaded, never used).
  auto clampk = [&](int64_t kb) { return kb < KB ? kb : KB - 1; };
  auto loadA = [&](qs_v4i (&r)[NLA], int64_t kb) {
    kb = clampk(kb);
#pragma unroll
    for (int i = 0; i < NLA; ++i) r[i] = *(const QS_GLOBAL qs_v4i *)(A + ((int64_t)(wave + NW * i) * KB + kb) * 1024);
  };
  auto putA = [&](const qs_v4i (&r)[NLA], int buf) {
#pragma unroll
    for (int i = 0; i < NLA; ++i) *(qs_v4i *)(smem[buf] + (wave + NW * i) * 1024 + lane * 16) = r[i];
  };
  auto loadB = [&](qs_v4i (&r)[TJ], int64_t kb) {
    kb = clampk(kb);
#pragma unroll
    for (int j = 0; j < TJ; ++j) r[j] = *(const QS_GLOBAL qs_v4i *)(Bw + ((int64_t)j * KB + kb) * 1024);
  };

  qs_v4i acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = qs_v4i{0, 0, 0, 0};

  qs_v4i as[2][NLA], bs[BS][TJ];
  // Stage kb (U = kb mod 4): A stage kb + 2 into the A set stage kb left
  // (already in LDS), B stage kb + BS - 1 into the B set stage kb - 1 left;
  // multiply stage kb (A from LDS buffer kb & 1, B from set kb mod BS); A
  // stage kb + 1 to the other LDS buffer.
  auto step = [&](int64_t kb, auto U) {
    constexpr int u = decltype(U)::value;
    loadA(as[u & 1], kb + 2);
    loadB(bs[(u + BS - 1) % BS], kb + BS - 1);
    const int8_t *st = smem[u & 1];
    qs_v4i af[TI];
#pragma unroll
    for (int i = 0; i < TI; ++i) af[i] = *(const qs_v4i *)(st + (wm * TI + i) * 1024 + lane * 16);
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[i], bs[u % BS][j], acc[i][j], 0, 0, 0);
    putA(as[(u + 1) & 1], (u + 1) & 1);
    // The stage's loads, fragment reads and LDS writes spread between its
    // MFMAs rather than issued as a burst ahead of them (about 2 % faster).
    __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
  #pragma unroll
    for (int i = 0; i < TI * TJ; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);
    }
    __syncthreads();
  };

  loadA(as[0], 0);
  loadA(as[1], 1);
#pragma unroll
  for (int q = 0; q < BS - 1; ++q) loadB(bs[q], q);
  putA(as[0], 0);
  __syncthreads();
  // Chunks of 64 stages (4096 K) with no branch inside, so no load is moved
  // past the MFMAs of its stage; the accumulators are reduced mod m between
  // chunks (|acc| < 4096 * 128 * 128 = 2^26 within one).  KB is a multiple of 4.
  for (int64_t kc = 0; kc < KB; kc += 64) {
    const int64_t ke = kc + 64 < KB ? kc + 64 : KB;
    for (int64_t kb = kc; kb < ke; kb += 4) {
      step(kb, std::integral_constant<int, 0>());
      step(kb + 1, std::integral_constant<int, 1>());
      step(kb + 2, std::integral_constant<int, 2>());
      step(kb + 3, std::integral_constant<int, 3>());
    }
    if (ke < KB) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] = crt_mod(acc[i][j][r], m, invm);
    }
  }
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      unsigned int w = 0;
#pragma unroll
      for (int r = 0; r < 4; ++r) w |= (unsigned int)crt_mod(acc[i][j][r], m, invm) << (8 * r);
      const int64_t blk = (mb0 + wm * TI + i) * a.NB + nb0 + wn * TJ + j;
      ((QS_GLOBAL unsigned int *)a.u)[(blk * a.n + l) * 64 + lane] = w;
    }
}

typedef __attribute__((address_space(3))) void qs_lds_void;

// One MFMA whose accumulator is tied to an AGPR quad: the compiler otherwise
// rotates 256 accumulator registers through copies every iteration.  Issued
// in program order; the loads between them are placed by hand.
__device__ __forceinline__ void mfma_i8_agpr(qs_v4i &acc, const qs_v4i &a, const qs_v4i &b) {
  asm volatile("v_mfma_i32_16x16x64_i8 %0, %1, %2, %0" : "+a"(acc) : "v"(a), "v"(b));
}

// The residue GEMMs (padded K up to kCrtMaxKpUnreduced): one 256 x 256 tile
// of one modulus per workgroup, 4 waves (2 x 2) of 128 x 128 = 8 x 8 MFMA
// blocks, one wave per SIMD with its 256 accumulators in AGPRs.  K runs in
// stages of one 64-byte fragment.  A stage's 16 A and 16 B fragment blocks
// (32 KiB) are copied by buffer loads straight into LDS (4 + 4 blocks per
// wave), three stages ahead, through a ring of four buffers; the fragments of
// stage kb + 1 are read from LDS during the first half of stage kb's 64
// MFMAs, into the second of two register sets.  One barrier per stage.  The
// epilogue reduces each sum mod m exactly in FP64 (x / m is never within
// 2^-20 below an integer unless it is one, for |x| < 2^31: crt_mod_exact)
// and stores the bytes.  Against eight waves of 128 x 64 with B read straight into
// registers: 32 against 48 KiB of vector-memory reads and 96 against 80 KiB
// of LDS traffic per CU and stage, 8-9 % faster at M = N = K = 4096.
__global__ __launch_bounds__(256, 1) void crt_gemm_kernel(const CrtGemmArgs a) {
  constexpr int TI = 8, TJ = 8, SB = 32 * 1024;
  __shared__ __attribute__((aligned(1024))) int8_t smem[4 * SB];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t tiles = a.tiles_m * a.tiles_n;
  const int64_t bid = xcd_swizzle(blockIdx.x, (int64_t)gridDim.x, a.xcds);
  const int l = (int)(bid / tiles);
  const int64_t tt = bid % tiles;
  // groups of 4 tile rows, walked along N
  const int64_t grp = tt / (4 * a.tiles_n), in = tt % (4 * a.tiles_n);
  const int64_t rows = min((int64_t)4, a.tiles_m - 4 * grp);
  const int64_t mb0 = (4 * grp + in % rows) * 16, nb0 = (in / rows) * 16;
  const int KB = (int)a.KB;
  // the tile's 16 row blocks of one operand over all of K: contiguous in the residue layout
  const int panel = 16 * KB * 1024;
  const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(a.ra + ((int64_t)l * a.MB + mb0) * (int64_t)KB * 1024), (short)0, panel, 0x00020000);
  const __amdgpu_buffer_rsrc_t rb = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(a.rb + ((int64_t)l * a.NB + nb0) * (int64_t)KB * 1024), (short)0, panel, 0x00020000);
  const unsigned vo = (unsigned)lane * 16u;
  const double md = (double)a.m[l], imd = a.invmd[l];

  // copy q (0..7) of stage kb into LDS buffer buf: block wave + 4 (q / 2) of A (q even) or B (q odd)
  auto dcopy = [&](int kb, int buf, int q) {
    const int blk = wave + 4 * (q >> 1);
    if ((q & 1) == 0)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (qs_lds_void *)(smem + buf * SB + blk * 1024), 16, vo,
                                               (blk * KB + kb) * 1024, 0, 0);
    else
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (qs_lds_void *)(smem + buf * SB + (16 + blk) * 1024), 16, vo,
                                               (blk * KB + kb) * 1024, 0, 0);
  };
  auto rdA = [&](int buf, int i) { return *(const qs_v4i *)(smem + buf * SB + (wm * TI + i) * 1024 + vo); };
  auto rdB = [&](int buf, int j) { return *(const qs_v4i *)(smem + buf * SB + (16 + wn * TJ + j) * 1024 + vo); };

  qs_v4i acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = qs_v4i{0, 0, 0, 0};
  qs_v4i af[2][TI], bf[2][TJ];

  // Stage kb (u = kb & 1): multiply register set u; copy stage kb + 3 (the
  // last stage again past the end) into buffer (kb + 3) & 3, which held stage
  // kb - 1, read during stage kb - 2; read stage kb + 1 from buffer
  // (kb + 1) & 3 into set u ^ 1.
  auto step = [&](int kb, auto U) {
    constexpr int u = decltype(U)::value;
    const int nb = (kb + 1) & 3, cb = (kb + 3) & 3;
    const int kc = kb + 3 < KB ? kb + 3 : KB - 1;
#pragma unroll
    for (int i = 0; i < TI; ++i) {
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        mfma_i8_agpr(acc[i][j], af[u][i], bf[u][j]);
        if (j == 1) dcopy(kc, cb, i);
        if (i < 4) {
          if (j == 2) bf[u ^ 1][2 * i] = rdB(nb, 2 * i);
          if (j == 4) bf[u ^ 1][2 * i + 1] = rdB(nb, 2 * i + 1);
          if (j == 6) af[u ^ 1][2 * i] = rdA(nb, 2 * i);
          if (j == 7) af[u ^ 1][2 * i + 1] = rdA(nb, 2 * i + 1);
        }
      }
    }
    // Stage kb + 2 landed in every wave (stage kb + 3's 8 copies may still be
    // in flight; vector memory counts retire in order), and every wave's
    // fragment reads are done before any buffer is refilled.
    asm volatile("s_waitcnt vmcnt(8) lgkmcnt(0)\n\ts_barrier" ::: "memory");
  };

  for (int q = 0; q < 8; ++q) dcopy(0, 0, q);
  for (int q = 0; q < 8; ++q) dcopy(1 < KB ? 1 : KB - 1, 1, q);
  for (int q = 0; q < 8; ++q) dcopy(2 < KB ? 2 : KB - 1, 2, q);
  asm volatile("s_waitcnt vmcnt(16)\n\ts_barrier" ::: "memory");
#pragma unroll
  for (int i = 0; i < TI; ++i) {
    af[0][i] = rdA(0, i);
    bf[0][i] = rdB(0, i);
  }
  // stage 1 landed; the accumulators' zeroing writes settled before the first MFMA reads them
  asm volatile("s_waitcnt vmcnt(8) lgkmcnt(0)\n\ts_barrier\n\ts_nop 7\n\ts_nop 7" ::: "memory");
  for (int kb = 0; kb < KB; kb += 2) {   // KB is a multiple of 4
    step(kb, std::integral_constant<int, 0>());
    step(kb + 1, std::integral_constant<int, 1>());
  }
  // No copy outstanding at exit; the last MFMAs' passes done before the AGPRs are read.
  asm volatile("s_waitcnt vmcnt(0)\n\ts_nop 15\n\ts_nop 15\n\ts_nop 15\n\ts_nop 15" ::: "memory");
  QS_GLOBAL unsigned int *U = (QS_GLOBAL unsigned int *)a.u;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      unsigned int w = 0;
#pragma unroll
      for (int r = 0; r < 4; ++r) w |= crt_mod_exact(acc[i][j][r], md, imd) << (8 * r);
      const int64_t blk = (mb0 + wm * TI + i) * a.NB + nb0 + wn * TJ + j;
      U[(blk * a.n + l) * 64 + lane] = w;
    }
}

// Reconstruction and the alpha / beta store.  One wave per strip of four
// 16 x 16 blocks laid along C's unit-stride side (64 x 16 when M is fast,
// 16 x 64 when N is), turned through LDS so each store covers 64 consecutive
// elements of C.
struct CrtRecArgs {
  GemmArgs g;
  int64_t MB, NB;
};

// Lane `l`'s value, as a wave-uniform one (l a compile-time constant).
__device__ __forceinline__ double lane_f64(double x, int l) {
  const int64_t b = __builtin_bit_cast(int64_t, x);
  const int lo = __builtin_amdgcn_readlane((int)b, l), hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, (int64_t)(((uint64_t)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ int64_t lane_i64(int64_t x, int l) {
  const int lo = __builtin_amdgcn_readlane((int)x, l), hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
  return (int64_t)(((uint64_t)(unsigned)hi << 32) | (unsigned)lo);
}

// R: the compute type; MIXED: C stored in the other real type.
template <bool MFAST, typename R, bool MIXED>
__global__ __launch_bounds__(256) void crt_rec_kernel(const CrtRecArgs a, const CrtConst c,
                                                      const unsigned int *__restrict__ ures,
                                                      const RowScale *__restrict__ sa, const RowScale *__restrict__ sb,
                                                      const int64_t *__restrict__ mC, const int64_t *__restrict__ nC) {
  __shared__ double tile[4][MFAST ? 64 : 16][MFAST ? 17 : 65];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t strip = (int64_t)blockIdx.x * 4 + wave;
  int64_t mb0, nb0;
  if (MFAST) {
    mb0 = 4 * (strip % (a.MB / 4));
    nb0 = strip / (a.MB / 4);
  } else {
    mb0 = strip % a.MB;
    nb0 = 4 * (strip / a.MB);
  }
  const GemmArgs &g = a.g;
  // The strip's scales and C offsets, loaded ahead of the residue words: lane
  // `lane` finishes C's element `fixed` along the unit-stride side, and lane e
  // (mod 16) holds the other side's element obase + e, which the store phase
  // broadcasts.  Indices past C's extents are clamped for the loads; those
  // elements are never stored.
  const int64_t flim = MFAST ? g.M : g.N, olim = MFAST ? g.N : g.M;
  const int64_t fixed = (MFAST ? mb0 : nb0) * 16 + lane;
  const int64_t obase = (MFAST ? nb0 : mb0) * 16;
  const RowScale *__restrict__ fsc = MFAST ? sa : sb;
  const RowScale *__restrict__ osc = MFAST ? sb : sa;
  const int64_t *__restrict__ foffs = MFAST ? mC : nC;
  const int64_t *__restrict__ ooffs = MFAST ? nC : mC;
  const int64_t fcl = fixed < flim ? fixed : flim - 1;
  const int64_t ocl = obase + (lane & 15) < olim ? obase + (lane & 15) : olim - 1;
  const double f_iq = fsc[fcl].iq, o_iq = osc[ocl].iq;
  const int f_sh = fsc[fcl].sh, o_sh = osc[ocl].sh;
  const int64_t foff = foffs[fcl], ooff = ooffs[ocl];
  const QS_GLOBAL unsigned int *u = (const QS_GLOBAL unsigned int *)ures;
  // The residue words of RB blocks are loaded before any arithmetic on them.
  // Words past n re-read word n - 1 (their constants are zero), so the loads
  // carry no branch.
  constexpr int RB = 2;   // four measured slower (102 against 95 us at 4096 x 4096, 13 moduli)
#pragma unroll 1
  for (int b0 = 0; b0 < 4; b0 += RB) {
  unsigned int wb[RB][kMaxModuli];
#pragma unroll
  for (int h = 0; h < RB; ++h) {
    const int b = b0 + h;
    const int64_t blk = (MFAST ? mb0 + b : mb0) * a.NB + (MFAST ? nb0 : nb0 + b);
    const QS_GLOBAL unsigned int *up = u + blk * c.n * 64 + lane;
#pragma unroll
    for (int l = 0; l < kMaxModuli; ++l) wb[h][l] = up[(l < c.n ? l : c.n - 1) * 64];
  }
#pragma unroll
  for (int h = 0; h < RB; ++h) {
    const int b = b0 + h;
    const unsigned int (&w)[kMaxModuli] = wb[h];
    double s1[4] = {0, 0, 0, 0}, s2[4] = {0, 0, 0, 0};
#pragma unroll
    for (int l = 0; l < kMaxModuli; ++l) {
      // The constants of words past n are zero: adding their products is exact.
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double uv = (double)((w[l] >> (8 * r)) & 0xffu);
        s1[r] = __builtin_fma(uv, c.q1[l], s1[r]);   // exact
        s2[r] = __builtin_fma(uv, c.q2[l], s2[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double q = rint((s1[r] + s2[r]) * c.invp);
      const double cp = __builtin_fma(-q, c.p1, s1[r]) + __builtin_fma(-q, c.p2, s2[r]);
      // Lane l holds row 4 (l / 16) + r, column l % 16 of the block.
      const int lr = 4 * (lane >> 4) + r, lc = lane & 15;
      if (MFAST) tile[wave][16 * b + lr][lc] = cp;
      else tile[wave][lr][16 * b + lc] = cp;
    }
  }
  }
  // Each wave reads back only its own tile: a wavefront-scope fence orders
  // its LDS writes before its reads, and the waves run on independently.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // Lane `lane` finishes 16 elements along C's unit-stride side: every
  // old-C load is issued before the first store, and no store waits on a
  // load of the other side's scales or offsets.
  if (fixed >= flim) return;
  const int nvalid = olim - obase < 16 ? (int)(olim - obase) : 16;   // wave-uniform
  int64_t idx[16];
  double v[16];
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    // Elements past C's extent take element 0's address for the old-C load
    // (in bounds; never stored), so the loads carry no branch.
    idx[e] = g.c_base + foff + lane_i64(ooff, e < nvalid ? e : 0);
    v[e] = ldexp((MFAST ? tile[wave][lane][e] : tile[wave][e][lane]) * f_iq * lane_f64(o_iq, e),
                 f_sh + __builtin_amdgcn_readlane(o_sh, e));
  }
  auto finish = [&](auto bcase) {
    constexpr int BC = decltype(bcase)::value;
    R o[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      R oi;
      load_c_old<R, false, MIXED, BC>(g, idx[e], o[e], oi);
    }
#pragma unroll
    for (int e = 0; e < 16; ++e)
      if (e < nvalid) store_c_at<R, false, MIXED, true, BC>(g, idx[e], (R)v[e], (R)0, o[e], (R)0);
  };
  if (g.beta_case == 0) finish(std::integral_constant<int, 0>());
  else if (g.beta_case == 1) finish(std::integral_constant<int, 1>());
  else finish(std::integral_constant<int, 2>());
}

// Error report: for sampled (i, j), the FP64 result against a Dot2
// (compensated, ~twice the working precision) evaluation of the same
// alpha * sum_k A[i,k] B[k,j] + beta * C_old.
struct SplitSample {
  int64_t i, j;
  double c_old;
};

__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e) {
  s = a + b;
  const double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__global__ __launch_bounds__(64) void split_sample_old_kernel(GemmArgs g, SplitSample *smp, int n) {
  const int s = blockIdx.x * 64 + threadIdx.x;
  if (s >= n) return;
  double re, im;
  load_elem<double, false, true>(g.C, g.c_type, g.c_base + g.mC[smp[s].i] + g.nC[smp[s].j], re, im);
  smp[s].c_old = g.beta_case == 0 ? 0.0 : re;
}

// One wave per sample; out[s] = {delivered, reference hi, reference lo}.
__global__ __launch_bounds__(64) void split_report_kernel(GemmArgs g, const SplitSample *smp, double *out) {
  const SplitSample sm = smp[blockIdx.x];
  const int lane = threadIdx.x;
  double s = 0.0, c = 0.0;
  for (int64_t k = lane; k < g.K; k += 64) {
    double x, y, im;
    load_elem<double, false, true>(g.A, g.a_type, g.a_base + g.mA[sm.i] + g.kA[k], x, im);
    load_elem<double, false, true>(g.B, g.b_type, g.b_base + g.kB[k] + g.nB[sm.j], y, im);
    const double p = x * y;
    const double pe = __builtin_fma(x, y, -p);
    double e;
    two_sum(s, p, s, e);
    c += pe + e;
  }
  for (int off = 32; off >= 1; off >>= 1) {   // combine lanes, still compensated
    const double so = __shfl_xor(s, off, 64), co = __shfl_xor(c, off, 64);
    double e;
    two_sum(s, so, s, e);
    c += co + e;
  }
  if (lane == 0) {
    // alpha * (s + c) + beta * c_old, compensated
    const double ar = g.alpha_re, br = g.beta_re;
    const double p1 = ar * s, p1e = __builtin_fma(ar, s, -p1);
    const double p2 = br * sm.c_old, p2e = __builtin_fma(br, sm.c_old, -p2);
    double h, e;
    two_sum(p1, p2, h, e);
    const double lo = e + p1e + p2e + ar * c;
    double delivered, im;
    load_elem<double, false, true>(g.C, g.c_type, g.c_base + g.mC[sm.i] + g.nC[sm.j], delivered, im);
    out[3 * blockIdx.x] = delivered;
    out[3 * blockIdx.x + 1] = h;
    out[3 * blockIdx.x + 2] = lo;
  }
}

// ------------------------------------------------------------------ host

typedef unsigned __int128 u128;

static int bitlen(u128 x) {
  int b = 0;
  while (x) { ++b; x >>= 1; }
  return b;
}

static double u128_to_double(u128 x) {
  return std::ldexp((double)(uint64_t)(x >> 64), 64) + (double)(uint64_t)x;
}

// Pairwise coprime moduli, largest first: 256, 255, 253, 251, 247, 241, ...
const std::vector<int> &split_moduli_table() {
  static const std::vector<int> tab = [] {
    std::vector<int> v;
    for (int m = 256; m >= 2 && (int)v.size() < kMaxModuli; --m) {
      bool ok = true;
      for (int x : v) {
        int a = m, b = x;
        while (b) { const int r = a % b; a = b; b = r; }
        if (a != 1) { ok = false; break; }
      }
      if (ok) v.push_back(m);
    }
    return v;
  }();
  return tab;
}

static u128 moduli_product(int n) {
  u128 p = 1;
  for (int i = 0; i < n; ++i) p *= (u128)split_moduli_table()[i];
  return p;
}

// Bound on the integer row norms: T = sqrt(P / 2) (1 - 2^-20) - sqrt(K) - 1,
// so (T + sqrt(K) / 2)^2 stays below P / 2 by a margin the rounding of
// rint(S / P) cannot cross.
static double crt_T(int n, int64_t K) {
  return std::sqrt(u128_to_double(moduli_product(n)) / 2.0) * (1.0 - 0x1p-20) - std::sqrt((double)K) - 1.0;
}

int split_moduli_for(double target, int ct, int64_t K) {
  if (!(target > 0)) target = ct == QS_F64 ? 1e-13 : 1e-6;
  // Rounding A and B to integers of row norm T leaves a relative error of
  // about 0.58 sqrt(K) / T in C (entries of random sign); a factor of two
  // margin on top.
  const double need = 1.16 * std::sqrt((double)std::max<int64_t>(K, 1)) / target;
  for (int n = 2; n <= kMaxModuli; ++n)
    if (crt_T(n, K) >= need) return n;
  return kMaxModuli;
}

static CrtConst crt_constants(int n) {
  CrtConst c{};
  const std::vector<int> &mt = split_moduli_table();
  const u128 P = moduli_product(n);
  c.n = n;
  u128 sum = 0;
  std::vector<u128> Q(n);
  for (int l = 0; l < n; ++l) {
    Q[l] = P / (u128)mt[l];
    sum += (u128)(mt[l] - 1) * Q[l];
  }
  const u128 top = std::max(sum, (u128)(n + 1) * P);
  const int g = std::max(0, bitlen(top) - 53);
  const u128 low = ((u128)1 << g) - 1;
  for (int l = 0; l < n; ++l) {
    const int m = mt[l];
    c.mf[l] = (double)m;
    c.invm[l] = 1.0 / (double)m;
    c.cst[l] = 0.5 - 128.0 / (double)m;
    const int qmod = (int)(Q[l] % (u128)m);
    int y = 1;
    while ((qmod * y) % m != 1) ++y;                   // (P/m)^-1 mod m
    const int p32 = (int)(((u128)1 << 32) % (u128)m);   // 2^32 mod m
    c.w[1][l][0] = (double)(1 % m);
    c.w[1][l][1] = (double)p32;
    c.w[0][l][0] = (double)(y % m);
    c.w[0][l][1] = (double)((p32 * y) % m);
    c.q1[l] = u128_to_double(Q[l] & ~low);              // exact: fewer than 53 significant bits
    c.q2[l] = u128_to_double(Q[l] & low);
  }
  c.p1 = u128_to_double(P & ~low);
  c.p2 = u128_to_double(P & low);
  c.invp = 1.0 / u128_to_double(P);
  return c;
}

static int64_t rup(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

struct SplitLayout {
  size_t ra, rb, u, scA, scB, smp, rep, total;
};

static SplitLayout split_layout(const qs_plan *p) {
  SplitLayout L;
  const int n = p->moduli;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off = (off + bytes + 255) / 256 * 256;
    return at;
  };
  L.ra = take((size_t)n * p->split_Mp * p->split_Kp);
  L.rb = take((size_t)n * p->split_Np * p->split_Kp);
  L.u = take((size_t)n * p->split_Mp * p->split_Np);
  L.scA = take(sizeof(RowScale) * p->M);
  L.scB = take(sizeof(RowScale) * p->N);
  L.smp = take(sizeof(SplitSample) * kSplitSamples);
  L.rep = take(sizeof(double) * 3 * kSplitSamples);
  L.total = off;
  return L;
}

qs_status split_prepare(qs_plan *p) {
  p->split_Mp = rup(p->M, kTile);
  p->split_Np = rup(p->N, kTile);
  p->split_Kp = rup(std::max<int64_t>(p->K, 1), 256);   // a multiple of four 64-K stages (the K loop's unroll)
  const SplitLayout L = split_layout(p);
  p->d_split = ws_alloc(p, L.total);
  if (!p->d_split) return fail(QS_ERR_MEMORY, "plan_contract: split workspace");
  p->split_bytes = (int64_t)L.total;
  if (hipStreamCreateWithFlags(&p->split_side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&p->split_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&p->split_join, hipEventDisableTiming) != hipSuccess)
    return fail(QS_ERR_DEVICE, "plan_contract: split side stream");
  return QS_OK;
}

// Error report, part 1 (before C is written): the sample positions and the
// old C at each of them, in the workspace.
static hipError_t split_sample(qs_plan *p, const GemmArgs &g, hipStream_t s, int &nsmp) {
  const SplitLayout L = split_layout(p);
  char *base = (char *)p->d_split;
  SplitSample *smp = (SplitSample *)(base + L.smp);
  nsmp = 0;
  {
    nsmp = (int)std::min<int64_t>(kSplitSamples, p->M * p->N);
    // Deterministic sample positions: a fixed stride walk over the M x N grid.
    std::vector<SplitSample> h(nsmp);
    const int64_t tot = p->M * p->N;
    for (int i = 0; i < nsmp; ++i) {
      const int64_t e = (int64_t)((double)i * (double)tot / (double)nsmp + 0.5) % tot;
      h[i].i = e % p->M;
      h[i].j = e / p->M;
      h[i].c_old = 0;
    }
    hipError_t e = hipMemcpyAsync(smp, h.data(), sizeof(SplitSample) * nsmp, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(split_sample_old_kernel, dim3((nsmp + 63) / 64), dim3(64), 0, s, g, smp, nsmp);
    e = hipStreamSynchronize(s);   // `h` goes out of scope
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// The split-precision contraction itself, with the plan's moduli.
static hipError_t split_compute(qs_plan *p, const GemmArgs &g, hipStream_t s) {
  const SplitLayout L = split_layout(p);
  char *base = (char *)p->d_split;
  const int n = p->moduli;
  const CrtConst c = crt_constants(n);
  int eT;
  const double Tm = std::frexp(crt_T(n, p->K), &eT);
  CrtOperand oa;
  oa.X = g.A; oa.type = g.a_type; oa.kfast = p->a_kfast; oa.base = g.a_base; oa.rowtab = g.mA; oa.ktab = g.kA;
  oa.R = p->M; oa.K = p->K; oa.Rp = p->split_Mp; oa.Kp = p->split_Kp; oa.side = 0;
  oa.scale = (RowScale *)(base + L.scA); oa.res = (int8_t *)(base + L.ra);
  oa.Tm = Tm; oa.eT = eT;
  CrtOperand ob = oa;
  ob.X = g.B; ob.type = g.b_type; ob.kfast = p->b_kfast; ob.base = g.b_base; ob.rowtab = g.nB; ob.ktab = g.kB;
  ob.R = p->N; ob.Rp = p->split_Np; ob.side = 1;
  ob.scale = (RowScale *)(base + L.scB); ob.res = (int8_t *)(base + L.rb);
  // A's passes on the caller's stream, B's on the side stream, forked and
  // joined by events.  Measured against one launch per pass covering both
  // operands (M = N = K = 4096): 226 us before the GEMM against 240 us with
  // the operands' workgroups in turn and 276 us with them interleaved.
  hipError_t fe = hipEventRecord(p->split_fork, s);
  if (fe == hipSuccess) fe = hipStreamWaitEvent(p->split_side, p->split_fork, 0);
  if (fe != hipSuccess) return fe;
  for (int side = 0; side < 2; ++side) {
    const CrtOperand &o = side ? ob : oa;
    hipStream_t os = side ? p->split_side : s;
    const dim3 rgrid((unsigned)(o.Rp / 64), (unsigned)(o.Kp / 64));
    const int rpw = o.kfast ? kRowsKFast : kRowsMFast;
    const dim3 sgrid((unsigned)((o.R + rpw - 1) / rpw));
    if (o.type == QS_F32) hipLaunchKernelGGL(crt_rownorm_kernel<float>, sgrid, dim3(256), 0, os, o);
    else hipLaunchKernelGGL(crt_rownorm_kernel<double>, sgrid, dim3(256), 0, os, o);
    if (o.type == QS_F32) hipLaunchKernelGGL(crt_residue_kernel<float>, rgrid, dim3(256), 0, os, o, c);
    else hipLaunchKernelGGL(crt_residue_kernel<double>, rgrid, dim3(256), 0, os, o, c);
  }
  fe = hipEventRecord(p->split_join, p->split_side);
  if (fe == hipSuccess) fe = hipStreamWaitEvent(s, p->split_join, 0);
  if (fe != hipSuccess) return fe;

  CrtGemmArgs ga;
  ga.ra = oa.res;
  ga.rb = ob.res;
  ga.u = (unsigned int *)(base + L.u);
  ga.MB = p->split_Mp / 16;
  ga.NB = p->split_Np / 16;
  ga.KB = p->split_Kp / 64;
  ga.n = n;
  ga.xcds = device_facts().xcds;
  for (int l = 0; l < kMaxModuli; ++l) {
    ga.m[l] = l < n ? split_moduli_table()[l] : 1;
    ga.invm[l] = 1.0f / (float)ga.m[l];
    ga.invmd[l] = 1.0 / (double)ga.m[l];
  }
  // 256 x 256 tiles, one workgroup per CU.  Up to kCrtMaxKpUnreduced: four
  // waves of 128 x 128 fed through LDS by buffer-to-LDS copies.  Beyond it,
  // eight waves of 128 x 64 with mid-K reductions (measured ahead of 128 x 128
  // tiles with 4 waves, 3 per CU, and 256 x 128 with 8 waves, 2 per CU).
  ga.tiles_m = p->split_Mp / kTile;
  ga.tiles_n = p->split_Np / kTile;
  const dim3 ggrid((unsigned)(ga.tiles_m * ga.tiles_n * n));
  if (p->split_Kp <= kCrtMaxKpUnreduced)
    hipLaunchKernelGGL(crt_gemm_kernel, ggrid, dim3(256), 0, s, ga);
  else
    hipLaunchKernelGGL((crt_gemm_longk_kernel<2, 4, 8, 4, 2>), ggrid, dim3(512), 0, s, ga);

  CrtRecArgs ra;
  ra.g = g;
  ra.MB = ga.MB;
  ra.NB = ga.NB;
  // Strips of four blocks, four per workgroup: Mp and Np are multiples of 256.
  const dim3 rgrid((unsigned)(ga.MB * ga.NB / 16));
  const bool f64 = g.compute_type == QS_F64, mixed = g.c_type != g.compute_type;
#define QS_REC_LAUNCH(MF, R, MX) \
  hipLaunchKernelGGL((crt_rec_kernel<MF, R, MX>), rgrid, dim3(256), 0, s, ra, c, ga.u, oa.scale, ob.scale, g.mC, g.nC)
  if (p->c_mfast) {
    if (f64) { if (mixed) QS_REC_LAUNCH(true, double, true); else QS_REC_LAUNCH(true, double, false); }
    else { if (mixed) QS_REC_LAUNCH(true, float, true); else QS_REC_LAUNCH(true, float, false); }
  } else {
    if (f64) { if (mixed) QS_REC_LAUNCH(false, double, true); else QS_REC_LAUNCH(false, double, false); }
    else { if (mixed) QS_REC_LAUNCH(false, float, true); else QS_REC_LAUNCH(false, float, false); }
  }
#undef QS_REC_LAUNCH
  return hipGetLastError();
}

static double split_target(const qs_plan *p) {
  const double t = p->desc.target_rel_err;
  return t > 0 ? t : (p->desc.compute_type == QS_F64 ? 1e-13 : 1e-6);
}

// Error report, part 2 (after C is written): Dot2 at the samples, reduced on
// the host into the plan's report.
static hipError_t split_report(qs_plan *p, const GemmArgs &g, hipStream_t s, int nsmp) {
  const SplitLayout L = split_layout(p);
  char *base = (char *)p->d_split;
  SplitSample *smp = (SplitSample *)(base + L.smp);
  double *rep = (double *)(base + L.rep);
  hipLaunchKernelGGL(split_report_kernel, dim3(nsmp), dim3(64), 0, s, g, smp, rep);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  std::vector<double> out(3 * nsmp);
  e = hipMemcpyAsync(out.data(), rep, sizeof(double) * 3 * nsmp, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  double num = 0, den = 0, mx = 0;
  for (int i = 0; i < nsmp; ++i) {
    const double ref = out[3 * i + 1] + out[3 * i + 2];
    const double d = std::fabs((out[3 * i] - out[3 * i + 1]) - out[3 * i + 2]);
    num += d * d;
    den += ref * ref;
    if (ref != 0) mx = std::max(mx, d / std::fabs(ref));
  }
  p->split_report.samples = nsmp;
  p->split_report.rel_fro = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
  p->split_report.max_rel = mx;
  p->split_report.exceeded = p->split_report.rel_fro > split_target(p) ? 1 : 0;
  return hipSuccess;
}

hipError_t launch_split(qs_plan *p, const GemmArgs &g, bool report, hipStream_t s) {
  int nsmp = 0;
  hipError_t e = hipSuccess;
  if (report) {
    p->split_report.escalations = 0;
    p->split_report.fallback = 0;
    e = split_sample(p, g, s, nsmp);
  }
  if (e == hipSuccess) e = split_compute(p, g, s);
  if (e == hipSuccess && report) e = split_report(p, g, s, nsmp);
  return e;
}

// The workspace for `n` moduli (its layout depends on n).
static bool split_resize(qs_plan *p, int n, hipStream_t s) {
  (void)hipStreamSynchronize(s);
  (void)hipStreamSynchronize(p->split_side);
  ws_free(p, p->d_split, (size_t)p->split_bytes);
  p->d_split = nullptr;
  p->split_bytes = 0;
  p->moduli = n;
  const SplitLayout L = split_layout(p);
  p->d_split = ws_alloc(p, L.total);
  if (!p->d_split) return false;
  p->split_bytes = (int64_t)L.total;
  return true;
}

// Guarded execute: the sampled error relative to C decides.  Above the
// target, the execute is redone from C as it was with more moduli (each adds
// about 3.9 bits to T, which the error scales with), up to kMaxModuli, and
// then with the plain MFMA kernel in the compute type.
qs_status run_split_guarded(qs_plan *p, const GemmArgs &g, hipStream_t s) {
  auto dev_err = [](hipError_t e) { return fail(QS_ERR_DEVICE, std::string("execute!: split: ") + hipGetErrorString(e)); };
  const double target = split_target(p);
  p->split_report.escalations = 0;
  p->split_report.fallback = 0;
  const int eb = elem_bytes(g.c_type);
  const int64_t clo = g.c_base + p->mC_lo + p->nC_lo, chi = g.c_base + p->mC_hi + p->nC_hi;
  const size_t cbytes = (size_t)(chi - clo + 1) * eb;
  char *cspan = (char *)g.C + clo * eb;
  const bool keep = g.beta_case != 0;
  if (keep && (int64_t)cbytes > p->cbackup_bytes) {
    ws_free(p, p->d_cbackup, (size_t)p->cbackup_bytes);
    p->cbackup_bytes = 0;
    p->d_cbackup = ws_alloc(p, cbytes);
    if (!p->d_cbackup) return fail(QS_ERR_MEMORY, "execute!: split: buffer for C's span");
    p->cbackup_bytes = (int64_t)cbytes;
  }
  hipError_t e = hipSuccess;
  if (keep) e = hipMemcpyAsync(p->d_cbackup, cspan, cbytes, hipMemcpyDeviceToDevice, s);
  auto restore = [&]() {
    return keep ? hipMemcpyAsync(cspan, p->d_cbackup, cbytes, hipMemcpyDeviceToDevice, s) : hipSuccess;
  };
  int nsmp = 0;
  if (e == hipSuccess) e = split_sample(p, g, s, nsmp);
  if (e == hipSuccess) e = split_compute(p, g, s);
  if (e == hipSuccess) e = split_report(p, g, s, nsmp);
  if (e != hipSuccess) return dev_err(e);
  while (p->split_report.rel_fro > target && p->moduli < kMaxModuli) {
    const double f = p->split_report.rel_fro / target;
    const int extra = std::max(1, (int)std::ceil(std::log2(2.0 * f) / 3.9));
    if (!split_resize(p, std::min(kMaxModuli, p->moduli + extra), s)) break;
    ++p->split_report.escalations;
    e = restore();
    if (e == hipSuccess) e = split_sample(p, g, s, nsmp);
    if (e == hipSuccess) e = split_compute(p, g, s);
    if (e == hipSuccess) e = split_report(p, g, s, nsmp);
    if (e != hipSuccess) return dev_err(e);
  }
  if (!p->d_split) {   // a resize failed: the workspace is gone; the plain kernel needs none
    if (!split_resize(p, p->moduli, s)) return fail(QS_ERR_MEMORY, "execute!: split workspace");
    e = restore();
    if (e == hipSuccess) e = split_sample(p, g, s, nsmp);
    if (e != hipSuccess) return dev_err(e);
    p->split_report.rel_fro = 2 * target + 1;
  }
  if (p->split_report.rel_fro > target) {
    GemmArgs gp = g;
    gp.tile = choose_tile(p->M, p->N, p->desc.compute_type);
    gp.k_splits = 1;
    gp.k_chunk = p->K;
    gp.partials = nullptr;
    e = restore();
    if (e == hipSuccess) e = launch_gemm(gp, s);
    if (e == hipSuccess) e = split_report(p, g, s, nsmp);
    if (e != hipSuccess) return dev_err(e);
    p->split_report.fallback = 1;
  }
  return QS_OK;
}

}  // namespace qs
