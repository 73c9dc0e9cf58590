// Bandwidth-bound paths and the GEMM dispatcher.
//
//   scale  -- C = beta * C with A and B never read: alpha == 0 or K == 0
//             (src/execution/execute.jl:34-65, interface.jl:80-95 scale_tile!).
//   dot    -- M == 1 or N == 1: a matrix operand reduced against a vector
//             (src/execution/dot.jl:13-243); HBM-bound, roofline 8 TB/s.
//   outer  -- K == 1: C = alpha * a b^T + beta * C streamed
//             (src/execution/outer.jl:10-69); HBM-bound.
#include "qs_gemm_impl.h"

namespace qs {

hipError_t launch_gemm_f64(const GemmArgs &g, hipStream_t s);
hipError_t launch_gemm_f32(const GemmArgs &g, hipStream_t s);
hipError_t launch_gemm_c64(const GemmArgs &g, hipStream_t s);
hipError_t launch_gemm_c32(const GemmArgs &g, hipStream_t s);

hipError_t launch_gemm(const GemmArgs &g, hipStream_t s) {
  switch (g.compute_type) {
    case QS_F64: return launch_gemm_f64(g, s);
    case QS_F32: return launch_gemm_f32(g, s);
    case QS_C64: return launch_gemm_c64(g, s);
    default: return launch_gemm_c32(g, s);
  }
}

// ------------------------------------------------------------------ scale
template <typename R, bool CPLX, bool MIXED>
__global__ __launch_bounds__(256) void scale_kernel(const GemmArgs g) {
  const int64_t n = g.M * g.N;
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int64_t i = e % g.M, j = e / g.M;
    const int64_t idx = g.c_base + g.mC[i] + g.nC[j];
    if (g.beta_case == 0) {
      store_elem<R, CPLX, MIXED>(g.C, g.c_type, idx, (R)0, (R)0);
    } else {
      R cr, ci;
      load_elem<R, CPLX, MIXED>(g.C, g.c_type, idx, cr, ci);
      const R br = (R)g.beta_re, bi = (R)g.beta_im;
      if (!CPLX) {
        store_elem<R, CPLX, MIXED>(g.C, g.c_type, idx, cr * br, (R)0);
      } else {
        R u1 = cr * br, u2 = ci * bi, u3 = cr * bi, u4 = ci * br;
        store_elem<R, CPLX, MIXED>(g.C, g.c_type, idx, u1 - u2, u3 + u4);
      }
    }
  }
}

// -------------------------------------------------------------------- dot
// The matrix operand X (free extent Q, tables fX/kX, base xbase, conj sx)
// against the vector v (table kv, base vbase, conj sv); output coordinate q
// goes to C through (fC[q] + cfixed).
struct DotArgs {
  const void *X; const void *V;
  int x_type, v_type;
  int64_t xbase, vbase;
  const int64_t *fX, *kX, *kv;
  int64_t Q, K;
  int conj_x, conj_v;
  bool rows_are_m;   // q indexes M (N == 1) or N (M == 1)
};

template <typename R, bool CPLX, bool MIXED>
__device__ __forceinline__ void dot_step(const DotArgs &d, int64_t xrow, int64_t k, R &accr, R &acci) {
  R xr, xi, vr, vi;
  load_elem<R, CPLX, MIXED>(d.X, d.x_type, xrow + d.kX[k], xr, xi);
  load_elem<R, CPLX, MIXED>(d.V, d.v_type, d.vbase + d.kv[k], vr, vi);
  if (!CPLX) {
    accr = qs_fma(xr, vr, accr);
  } else {
    if (d.conj_x) xi = -xi;
    if (d.conj_v) vi = -vi;
    const R nr = qs_fma(-xi, vi, qs_fma(xr, vr, accr));
    const R ni = qs_fma(xi, vr, qs_fma(xr, vi, acci));
    accr = nr;
    acci = ni;
  }
}

template <typename R, bool CPLX, bool MIXED>
__device__ __forceinline__ void dot_store(const GemmArgs &g, const DotArgs &d, int64_t q, R xr, R xi) {
  if (d.rows_are_m) store_c<R, CPLX, MIXED>(g, q, 0, xr, xi);
  else store_c<R, CPLX, MIXED>(g, 0, q, xr, xi);
}

// One wave per output, lanes along K (the matrix is K-contiguous): each lane
// sums its own K residue class in order, then a fixed butterfly reduction.
template <typename R, bool CPLX, bool MIXED>
__global__ __launch_bounds__(256) void dot_klanes_kernel(const GemmArgs g, const DotArgs d) {
  const int lane = threadIdx.x & 63;
  const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= d.Q) return;
  const int64_t xrow = d.xbase + d.fX[q];
  R accr = 0, acci = 0;
  for (int64_t k = lane; k < d.K; k += 64) dot_step<R, CPLX, MIXED>(d, xrow, k, accr, acci);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    accr += __shfl_xor(accr, off, 64);
    if (CPLX) acci += __shfl_xor(acci, off, 64);
  }
  if (lane == 0) dot_store<R, CPLX, MIXED>(g, d, q, accr, acci);
}

// One thread per output, lanes along the free extent (the matrix is
// free-contiguous): every thread walks K in order, as the scalar kernel does.
template <typename R, bool CPLX, bool MIXED>
__global__ __launch_bounds__(256) void dot_qlanes_kernel(const GemmArgs g, const DotArgs d) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= d.Q) return;
  const int64_t xrow = d.xbase + d.fX[q];
  R accr = 0, acci = 0;
  for (int64_t k = 0; k < d.K; ++k) dot_step<R, CPLX, MIXED>(d, xrow, k, accr, acci);
  dot_store<R, CPLX, MIXED>(g, d, q, accr, acci);
}

static DotArgs make_dot(const GemmArgs &g) {
  DotArgs d;
  if (g.N == 1) {   // matrix A (M x K), vector B
    d.X = g.A; d.V = g.B; d.x_type = g.a_type; d.v_type = g.b_type;
    d.xbase = g.a_base; d.vbase = g.b_base + 0;
    d.fX = g.mA; d.kX = g.kA; d.kv = g.kB; d.Q = g.M;
    d.conj_x = g.conj_a; d.conj_v = g.conj_b; d.rows_are_m = true;
  } else {          // M == 1: matrix B (N x K), vector A
    d.X = g.B; d.V = g.A; d.x_type = g.b_type; d.v_type = g.a_type;
    d.xbase = g.b_base; d.vbase = g.a_base + 0;
    d.fX = g.nB; d.kX = g.kB; d.kv = g.kA; d.Q = g.N;
    d.conj_x = g.conj_b; d.conj_v = g.conj_a; d.rows_are_m = false;
  }
  d.K = g.K;
  return d;
}

template <typename R, bool CPLX, bool MIXED>
static hipError_t launch_dot_typed(const GemmArgs &g, hipStream_t s) {
  DotArgs d = make_dot(g);
  // The vector's own free offset (N == 1: nB[0]; M == 1: mA[0]) is added on
  // the device side through the base, read back here from the fixed tables:
  // the caller folded it into `vbase` (see qs_api.cpp).
  const bool klanes = g.N == 1 ? g.a_kfast : g.b_kfast;
  if (klanes) {
    const int64_t blocks = (d.Q + 3) / 4;
    hipLaunchKernelGGL((dot_klanes_kernel<R, CPLX, MIXED>), dim3((unsigned)blocks), dim3(256), 0, s, g, d);
  } else {
    const int64_t blocks = (d.Q + 255) / 256;
    hipLaunchKernelGGL((dot_qlanes_kernel<R, CPLX, MIXED>), dim3((unsigned)blocks), dim3(256), 0, s, g, d);
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------ outer
// K == 1: one thread per C element, threads laid along C's unit-stride side
// (`a_kfast` carries "N fastest" for this path).
template <typename R, bool CPLX, bool MIXED, bool NFAST>
__global__ __launch_bounds__(256) void outer_kernel(const GemmArgs g) {
  const int64_t n = g.M * g.N;
  for (int64_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int64_t i = NFAST ? e / g.N : e % g.M;
    const int64_t j = NFAST ? e % g.N : e / g.M;
    R ar, ai, br, bi;
    load_elem<R, CPLX, MIXED>(g.A, g.a_type, g.a_base + g.mA[i] + g.kA[0], ar, ai);
    load_elem<R, CPLX, MIXED>(g.B, g.b_type, g.b_base + g.nB[j] + g.kB[0], br, bi);
    R xr, xi = 0;
    if (!CPLX) {
      xr = qs_fma(ar, br, (R)0);
    } else {
      if (g.conj_a) ai = -ai;
      if (g.conj_b) bi = -bi;
      xr = qs_fma(-ai, bi, qs_fma(ar, br, (R)0));
      xi = qs_fma(ai, br, qs_fma(ar, bi, (R)0));
    }
    store_c<R, CPLX, MIXED>(g, i, j, xr, xi);
  }
}

template <typename R, bool CPLX, bool MIXED>
static hipError_t launch_outer_typed(const GemmArgs &g, hipStream_t s) {
  int64_t blocks = (g.M * g.N + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (g.a_kfast)
    hipLaunchKernelGGL((outer_kernel<R, CPLX, MIXED, true>), dim3((unsigned)blocks), dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL((outer_kernel<R, CPLX, MIXED, false>), dim3((unsigned)blocks), dim3(256), 0, s, g);
  return hipGetLastError();
}

template <typename R, bool CPLX, bool MIXED>
static hipError_t launch_scale_typed(const GemmArgs &g, hipStream_t s) {
  int64_t blocks = (g.M * g.N + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((scale_kernel<R, CPLX, MIXED>), dim3((unsigned)blocks), dim3(256), 0, s, g);
  return hipGetLastError();
}

#define QS_DISPATCH(FN, g, s)                                                        \
  do {                                                                               \
    switch ((g).compute_type) {                                                      \
      case QS_F64: return (g).mixed ? FN<double, false, true>(g, s) : FN<double, false, false>(g, s); \
      case QS_F32: return (g).mixed ? FN<float, false, true>(g, s) : FN<float, false, false>(g, s);   \
      case QS_C64: return (g).mixed ? FN<double, true, true>(g, s) : FN<double, true, false>(g, s);   \
      default: return (g).mixed ? FN<float, true, true>(g, s) : FN<float, true, false>(g, s);         \
    }                                                                                \
  } while (0)

hipError_t launch_scale(const GemmArgs &g, hipStream_t s) { QS_DISPATCH(launch_scale_typed, g, s); }
hipError_t launch_dot(const GemmArgs &g, hipStream_t s) { QS_DISPATCH(launch_dot_typed, g, s); }
hipError_t launch_outer(const GemmArgs &g, hipStream_t s) { QS_DISPATCH(launch_outer_typed, g, s); }

}  // namespace qs
