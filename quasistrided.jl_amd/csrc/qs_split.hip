// Split-precision FP64 / FP32 contraction on the INT8 MFMA units (Ozaki
// scheme, Ozaki-Ogita-Oishi-Rump 2012, with integer slices as in the
// int8-emulation literature).
//
//   1. Exact power-of-two scales: e_i = exponent of max_k |A[i,k]| (rows of
//      A), f_j = exponent of max_k |B[k,j]| (columns of B), so every scaled
//      entry lies in (-1, 1).  Scaling by 2^-e is exact.
//   2. S slices per operand: t = x * 2^7, sigma = trunc(t), x <- t - sigma,
//      repeated S times; x = sum_p sigma_p 2^(-7p) + O(2^(-7S)), |sigma| <= 127.
//      Every step is exact in FP64.  Slices are stored dense, K contiguous,
//      rows and K zero-padded to multiples of 64.
//   3. C_scaled = sum_{p+q <= S+1} 2^(-7(p+q)) (A_p B_q): each slice product
//      on v_mfma_i32_16x16x64_i8 with exact INT32 accumulation (the pairs of
//      one diagonal d = p+q share an INT32 accumulator; flushed to FP64 every
//      8192 K steps so it never overflows), folded into FP64 registers.
//   4. Epilogue: C = alpha * 2^(e_i + f_j) C_scaled + beta * C through C's
//      offset tables, one rounding to C's type.
//
// S is chosen from the requested relative error (qs_contract_desc
// target_rel_err); the sampled error report (Dot2 of Ogita-Rump-Oishi 2005
// on the same gathered operands) measures what was delivered.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "qs_gemm_impl.h"
#include "qs_plan.h"

#include <vector>

namespace qs {

typedef int qs_v4i __attribute__((ext_vector_type(4)));

// One operand's gather for the prep kernels: R rows (M of A, N of B) by K.
struct SplitOperand {
  const void *X;
  int type;
  int kfast;
  int64_t base;
  const int64_t *rowtab, *ktab;
  int64_t R, K;
  unsigned long long *rowmax;   // bit pattern of max |x| per row (>= 0 doubles order as integers)
  int *exps;                    // per row: e with max * 2^-e in [0.5, 1)
  int8_t *slices;               // [S][Rp][Kp]
  int64_t Rp, Kp;
  int S;
};

__global__ __launch_bounds__(256) void split_zero_kernel(unsigned long long *p, int64_t n) {
  for (int64_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0ull;
}

__device__ __forceinline__ double split_load(const SplitOperand &o, int64_t r, int64_t k) {
  double re, im;
  load_elem<double, false, true>(o.X, o.type, o.base + o.rowtab[r] + o.ktab[k], re, im);
  return re;
}

// Tile of 64 rows x 64 K per workgroup; lanes along the operand's
// unit-stride side.  Element (tr, tk) of the tile for pass `it` (0..15).
__device__ __forceinline__ void split_coord(int kfast, int t, int it, int &tr, int &tk) {
  const int e = t + 256 * it;
  if (kfast) { tk = e % 64; tr = e / 64; } else { tr = e % 64; tk = e / 64; }
}

__global__ __launch_bounds__(256) void split_rowmax_kernel(SplitOperand o) {
  __shared__ unsigned long long smax[64];
  const int t = threadIdx.x;
  if (t < 64) smax[t] = 0ull;
  __syncthreads();
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  for (int it = 0; it < 16; ++it) {
    int tr, tk;
    split_coord(o.kfast, t, it, tr, tk);
    const int64_t r = r0 + tr, k = k0 + tk;
    if (r < o.R && k < o.K) {
      const double v = fabs(split_load(o, r, k));
      atomicMax(&smax[tr], (unsigned long long)__double_as_longlong(v));
    }
  }
  __syncthreads();
  if (t < 64 && r0 + t < o.R && smax[t]) atomicMax(&o.rowmax[r0 + t], smax[t]);
}

template <int S>
__global__ __launch_bounds__(256) void split_slice_kernel(SplitOperand o) {
  __shared__ int8_t sl[S][64][64 + 4];
  __shared__ int sexp[64];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  if (t < 64) {
    const int64_t r = r0 + t;
    int e = 0;
    if (r < o.R) {
      const double m = __longlong_as_double((long long)o.rowmax[r]);
      if (m > 0) (void)frexp(m, &e);
      if (blockIdx.y == 0) o.exps[r] = e;
    }
    sexp[t] = e;
  }
  __syncthreads();
  for (int it = 0; it < 16; ++it) {
    int tr, tk;
    split_coord(o.kfast, t, it, tr, tk);
    const int64_t r = r0 + tr, k = k0 + tk;
    double x = (r < o.R && k < o.K) ? ldexp(split_load(o, r, k), -sexp[tr]) : 0.0;
#pragma unroll
    for (int p = 0; p < S; ++p) {
      const double tt = x * 128.0;
      const double sg = trunc(tt);
      sl[p][tr][tk] = (int8_t)(int)sg;
      x = tt - sg;
    }
  }
  __syncthreads();
  // Write out K-contiguous: each thread stores 4 consecutive bytes of a row.
  for (int p = 0; p < S; ++p)
    for (int it = 0; it < 4; ++it) {
      const int e = t + 256 * it;          // 1024 words: 64 rows x 16 words
      const int tr = e / 16, w = e % 16;
      const int64_t r = r0 + tr;
      if (r >= o.Rp) continue;
      const int64_t kk = k0 + 4 * w;
      int word;
      __builtin_memcpy(&word, &sl[p][tr][4 * w], 4);
      *(QS_GLOBAL int *)(o.slices + ((int64_t)p * o.Rp + r) * o.Kp + kk) = word;
    }
}

// The slice GEMM: 64 x 64 tile per workgroup (4 waves, 2 x 2, each 32 x 32 =
// 2 x 2 MFMA blocks), K in stages of 64 bytes, all S slices of A and B of a
// stage in LDS (row pitch 80 B keeps the 16-byte fragment reads conflict
// free), double-buffered with a register prefetch of the next stage.
struct SplitGemmArgs {
  GemmArgs g;                       // C, its tables, M, N, alpha/beta, types
  const int8_t *as, *bs;            // [S][Mp][Kp], [S][Np][Kp]
  const int *ea, *eb;
  int64_t Mp, Np, Kp;
};

template <int S>
struct SplitTile {
  static constexpr int BM = 64, BN = 64, BK = 64, PITCH = 80;
  static constexpr int STAGE = S * (BM + BN) * PITCH;    // bytes
  static constexpr int WORDS = S * (BM + BN) * BK / 16;  // 16-byte words per stage
  static constexpr int NW = WORDS / 256;                 // per thread
  // Two LDS stages while they fit beside the runtime's reservation; from 8
  // slices on, one stage refilled from the register prefetch.
  static constexpr int STAGES = 2 * STAGE <= 150 * 1024 ? 2 : 1;
};

template <int S>
__global__ __launch_bounds__(256, 1) void split_gemm_kernel(const SplitGemmArgs a) {
  typedef SplitTile<S> T;
  __shared__ __attribute__((aligned(16))) int8_t smem[T::STAGES * T::STAGE];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int fr = lane & 15, fk = lane >> 4;
  const int64_t tiles_m = a.Mp / 64;
  const int64_t bid = xcd_swizzle(blockIdx.x, (int64_t)gridDim.x);
  const int64_t m0 = (bid % tiles_m) * 64, n0 = (bid / tiles_m) * 64;

  // Stage word w (16 bytes): operand (A rows 0..63, then B rows), slice, row, k-word.
  auto src = [&](int w, int64_t k0) -> const QS_GLOBAL qs_v4i * {
    const int kw = w & 3;               // 4 words of 16 bytes per 64-byte row
    const int rowid = (w >> 2) % 128;   // 0..63 A, 64..127 B
    const int p = (w >> 2) / 128;
    const int8_t *base = rowid < 64 ? a.as + ((int64_t)p * a.Mp + m0 + rowid) * a.Kp
                                    : a.bs + ((int64_t)p * a.Np + n0 + rowid - 64) * a.Kp;
    return (const QS_GLOBAL qs_v4i *)(base + k0 + 16 * kw);
  };
  auto dst = [&](int8_t *stage, int w) -> qs_v4i * {
    const int kw = w & 3, rowid = (w >> 2) % 128, p = (w >> 2) / 128;
    return (qs_v4i *)(stage + ((p * 128) + rowid) * T::PITCH + 16 * kw);
  };

  qs_v4i pre[T::NW];
  qs_v4i acc[S][2][2];
  double accd[2][2][4];
#pragma unroll
  for (int d = 0; d < S; ++d)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[d][i][j] = qs_v4i{0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) accd[i][j][r] = 0.0;

  auto flush = [&]() {
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          double s = 0.0;
#pragma unroll
          for (int d = S - 1; d >= 0; --d) s += ldexp((double)acc[d][i][j][r], -7 * (d + 2));   // small first
          accd[i][j][r] += s;
        }
#pragma unroll
    for (int d = 0; d < S; ++d)
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[d][i][j] = qs_v4i{0, 0, 0, 0};
  };

  const int64_t nk = a.Kp / 64;
#pragma unroll
  for (int i = 0; i < T::NW; ++i) pre[i] = *src(t + 256 * i, 0);
#pragma unroll
  for (int i = 0; i < T::NW; ++i) *dst(smem, t + 256 * i) = pre[i];
  __syncthreads();
  int buf = 0;
  for (int64_t kc = 0; kc < nk; ++kc) {
    const bool more = kc + 1 < nk;
    if (more) {
#pragma unroll
      for (int i = 0; i < T::NW; ++i) pre[i] = *src(t + 256 * i, (kc + 1) * 64);
    }
    const int8_t *st = smem + buf * T::STAGE;
    // All A fragments of the stage stay in registers; B fragments are read
    // one slice at a time and used by every A slice it pairs with.
    qs_v4i af[S][2];
#pragma unroll
    for (int p = 0; p < S; ++p)
#pragma unroll
      for (int i = 0; i < 2; ++i)
        af[p][i] = *(const qs_v4i *)(st + ((p * 128) + wm * 32 + i * 16 + fr) * T::PITCH + 16 * fk);
#pragma unroll
    for (int q = 0; q < S; ++q) {
      qs_v4i bf[2];
#pragma unroll
      for (int j = 0; j < 2; ++j)
        bf[j] = *(const qs_v4i *)(st + ((q * 128) + 64 + wn * 32 + j * 16 + fr) * T::PITCH + 16 * fk);
#pragma unroll
      for (int p = 0; p + q < S; ++p)   // one-based p + q <= S + 1
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
            acc[p + q][i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(af[p][i], bf[j], acc[p + q][i][j], 0, 0, 0);
    }
    if (((kc + 1) & 127) == 0) flush();   // every 8192 K steps: INT32 never overflows
    if (more) {
      if (T::STAGES == 1) __syncthreads();   // everyone is done reading the only stage
      int8_t *ws = smem + (T::STAGES == 2 ? (buf ^ 1) : 0) * T::STAGE;
#pragma unroll
      for (int i = 0; i < T::NW; ++i) *dst(ws, t + 256 * i) = pre[i];
    }
    __syncthreads();
    if (T::STAGES == 2) buf ^= 1;
  }
  flush();

  // Epilogue: undo the exact scales, then alpha / beta through C's tables.
  const GemmArgs &g = a.g;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t row = m0 + wm * 32 + i * 16 + 4 * fk + r;
      if (row >= g.M) continue;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int64_t col = n0 + wn * 32 + j * 16 + fr;
        if (col >= g.N) continue;
        const double v = ldexp(accd[i][j][r], a.ea[row] + a.eb[col]);
        if (g.compute_type == QS_F64) store_c<double, false, true>(g, row, col, v, 0.0);
        else store_c<float, false, true>(g, row, col, (float)v, 0.0f);
      }
    }
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

int split_slices_for(double target, int ct) {
  if (!(target > 0)) target = ct == QS_F64 ? 1e-13 : 1e-6;
  // Relative error of the truncated product about 2^(6 - 7S): the residual
  // slices and the dropped pairs each contribute a few units of 2^-7S of the
  // row / column maxima, and N(0,1) rows sit about 2 bits below their
  // maximum (measured: 1.1e-13 at S = 7, M = N = K = 4096).
  int s = (int)std::ceil((6.0 - std::log2(target)) / 7.0);
  return std::max(2, std::min(8, s));
}

static int64_t rup(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

qs_status split_prepare(qs_plan *p) {
  const int64_t M = p->M, N = p->N, K = p->K;
  p->split_Mp = rup(M, 64);
  p->split_Np = rup(N, 64);
  p->split_Kp = rup(std::max<int64_t>(K, 1), 64);
  const int S = p->slices;
  const size_t sa = (size_t)S * p->split_Mp * p->split_Kp, sb = (size_t)S * p->split_Np * p->split_Kp;
  const size_t bytes = sa + sb + sizeof(unsigned long long) * (M + N) + sizeof(int) * (M + N) +
                       sizeof(SplitSample) * kSplitSamples + sizeof(double) * 3 * kSplitSamples + 256;
  if (hipMalloc(&p->d_split, bytes) != hipSuccess) return fail(QS_ERR_MEMORY, "plan_contract: split workspace");
  // Padding rows and K stay zero for the plan's lifetime (written here once,
  // outside any stream the caller may capture).
  if (hipMemset(p->d_split, 0, sa + sb) != hipSuccess) return fail(QS_ERR_DEVICE, "plan_contract: split workspace");
  p->split_bytes = (int64_t)bytes;
  p->workspace_bytes += (int64_t)bytes;
  return QS_OK;
}

struct SplitPtrs {
  int8_t *as, *bs;
  unsigned long long *rmA, *rmB;
  int *eA, *eB;
  SplitSample *smp;
  double *rep;
};

static SplitPtrs split_ptrs(qs_plan *p) {
  SplitPtrs q;
  char *c = (char *)p->d_split;
  const size_t sa = (size_t)p->slices * p->split_Mp * p->split_Kp, sb = (size_t)p->slices * p->split_Np * p->split_Kp;
  q.as = (int8_t *)c;
  q.bs = (int8_t *)(c + sa);
  size_t off = (sa + sb + 15) / 16 * 16;
  q.rmA = (unsigned long long *)(c + off);
  q.rmB = q.rmA + p->M;
  off += sizeof(unsigned long long) * (p->M + p->N);
  q.eA = (int *)(c + off);
  q.eB = q.eA + p->M;
  off += sizeof(int) * (p->M + p->N);
  off = (off + 15) / 16 * 16;
  q.smp = (SplitSample *)(c + off);
  off += sizeof(SplitSample) * kSplitSamples;
  q.rep = (double *)(c + off);
  return q;
}

template <int S>
static void launch_slicing(SplitOperand &o, hipStream_t s) {
  dim3 grid((unsigned)((o.R + 63) / 64), (unsigned)((o.K + 63) / 64));
  hipLaunchKernelGGL(split_rowmax_kernel, grid, dim3(256), 0, s, o);
  hipLaunchKernelGGL(split_slice_kernel<S>, grid, dim3(256), 0, s, o);
}

template <int S>
static hipError_t launch_split_s(qs_plan *p, const GemmArgs &g, const SplitPtrs &q, hipStream_t s) {
  SplitOperand oa, ob;
  oa.X = g.A; oa.type = g.a_type; oa.kfast = p->a_kfast; oa.base = g.a_base; oa.rowtab = g.mA; oa.ktab = g.kA;
  oa.R = p->M; oa.K = p->K; oa.rowmax = q.rmA; oa.exps = q.eA; oa.slices = q.as; oa.Rp = p->split_Mp;
  oa.Kp = p->split_Kp; oa.S = S;
  ob.X = g.B; ob.type = g.b_type; ob.kfast = p->b_kfast; ob.base = g.b_base; ob.rowtab = g.nB; ob.ktab = g.kB;
  ob.R = p->N; ob.K = p->K; ob.rowmax = q.rmB; ob.exps = q.eB; ob.slices = q.bs; ob.Rp = p->split_Np;
  ob.Kp = p->split_Kp; ob.S = S;
  const int64_t nmax = p->M + p->N;
  hipLaunchKernelGGL(split_zero_kernel, dim3((unsigned)std::min<int64_t>((nmax + 255) / 256, 1024)), dim3(256), 0, s,
                     q.rmA, nmax);
  launch_slicing<S>(oa, s);
  launch_slicing<S>(ob, s);
  SplitGemmArgs a;
  a.g = g;
  a.as = q.as;
  a.bs = q.bs;
  a.ea = q.eA;
  a.eb = q.eB;
  a.Mp = p->split_Mp;
  a.Np = p->split_Np;
  a.Kp = p->split_Kp;
  const int64_t tiles = (a.Mp / 64) * (a.Np / 64);
  hipLaunchKernelGGL(split_gemm_kernel<S>, dim3((unsigned)tiles), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_split(qs_plan *p, const GemmArgs &g, bool report, hipStream_t s) {
  const SplitPtrs q = split_ptrs(p);
  int n = 0;
  if (report) {
    n = (int)std::min<int64_t>(kSplitSamples, p->M * p->N);
    // Deterministic sample positions: a fixed stride walk over the M x N grid.
    std::vector<SplitSample> h(n);
    const int64_t tot = p->M * p->N;
    for (int i = 0; i < n; ++i) {
      const int64_t e = (int64_t)((double)i * (double)tot / (double)n + 0.5) % tot;
      h[i].i = e % p->M;
      h[i].j = e / p->M;
      h[i].c_old = 0;
    }
    hipError_t e = hipMemcpyAsync(q.smp, h.data(), sizeof(SplitSample) * n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(split_sample_old_kernel, dim3((n + 63) / 64), dim3(64), 0, s, g, q.smp, n);
    e = hipStreamSynchronize(s);   // `h` goes out of scope
    if (e != hipSuccess) return e;
  }
  hipError_t e;
  switch (p->slices) {
    case 2: e = launch_split_s<2>(p, g, q, s); break;
    case 3: e = launch_split_s<3>(p, g, q, s); break;
    case 4: e = launch_split_s<4>(p, g, q, s); break;
    case 5: e = launch_split_s<5>(p, g, q, s); break;
    case 6: e = launch_split_s<6>(p, g, q, s); break;
    case 7: e = launch_split_s<7>(p, g, q, s); break;
    default: e = launch_split_s<8>(p, g, q, s); break;
  }
  if (e != hipSuccess || !report) return e;
  hipLaunchKernelGGL(split_report_kernel, dim3(n), dim3(64), 0, s, g, q.smp, q.rep);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  std::vector<double> out(3 * n);
  e = hipMemcpyAsync(out.data(), q.rep, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  double num = 0, den = 0, mx = 0;
  for (int i = 0; i < n; ++i) {
    const double ref = out[3 * i + 1] + out[3 * i + 2];
    const double d = std::fabs((out[3 * i] - out[3 * i + 1]) - out[3 * i + 2]);
    num += d * d;
    den += ref * ref;
    if (ref != 0) mx = std::max(mx, d / std::fabs(ref));
  }
  p->split_report.samples = n;
  p->split_report.rel_fro = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
  p->split_report.max_rel = mx;
  return hipSuccess;
}

}  // namespace qs
