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
__device__ __forceinline__ void dot_store(const GemmArgs &g, const DotArgs &d, int64_t q, R xr, R xi) {
  if (d.rows_are_m) store_c<R, CPLX, MIXED>(g, q, 0, xr, xi);
  else store_c<R, CPLX, MIXED>(g, 0, q, xr, xi);
}

// Lanes along K (the matrix is K-contiguous): one wave per 4 outputs.  Each
// lane sums its own K residue class (k = lane mod 64) of each of the 4 rows
// in order, with 4 k-steps of loads issued together; the K offsets and the
// vector elements of a step are loaded once for the 4 rows.  Then a fixed
// butterfly reduction per row.  The matrix is read with non-temporal loads:
// it streams through once, and the vector and K table stay cached
// (65536 x 2048 doubles: 217 -> 208 us).
template <typename R, bool CPLX, bool MIXED>
__global__ __launch_bounds__(256) void dot_klanes_kernel(const GemmArgs g, const DotArgs d) {
  constexpr int RW = 4, U = 4;
  const int lane = threadIdx.x & 63;
  const int64_t q0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RW;
  if (q0 >= d.Q) return;
  int64_t xrow[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) xrow[r] = d.xbase + d.fX[q0 + r < d.Q ? q0 + r : d.Q - 1];
  R accr[RW], acci[RW];
#pragma unroll
  for (int r = 0; r < RW; ++r) accr[r] = acci[r] = 0;
  for (int64_t k0 = 0; k0 < d.K; k0 += 64 * U) {
    R xr[U][RW], xi[U][RW], vr[U], vi[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = k0 + 64 * u + lane;
      ok[u] = k < d.K;
      const int64_t kc = ok[u] ? k : 0;
      const int64_t kx = d.kX[kc];
      load_elem<R, CPLX, MIXED>(d.V, d.v_type, d.vbase + d.kv[kc], vr[u], vi[u]);
#pragma unroll
      for (int r = 0; r < RW; ++r) load_elem<R, CPLX, MIXED, true>(d.X, d.x_type, xrow[r] + kx, xr[u][r], xi[u][r]);
    }
    __builtin_amdgcn_sched_barrier(0);   // every load of the step is in flight
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const R v_r = vr[u], v_i = d.conj_v ? -vi[u] : vi[u];
#pragma unroll
      for (int r = 0; r < RW; ++r) {
        if (!CPLX) {
          const R nr = qs_fma(xr[u][r], v_r, accr[r]);
          accr[r] = ok[u] ? nr : accr[r];
        } else {
          const R x_i = d.conj_x ? -xi[u][r] : xi[u][r];
          const R nr = qs_fma(-x_i, v_i, qs_fma(xr[u][r], v_r, accr[r]));
          const R ni = qs_fma(x_i, v_r, qs_fma(xr[u][r], v_i, acci[r]));
          accr[r] = ok[u] ? nr : accr[r];
          acci[r] = ok[u] ? ni : acci[r];
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < RW; ++r) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      accr[r] += __shfl_xor(accr[r], off, 64);
      if (CPLX) acci[r] += __shfl_xor(acci[r], off, 64);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r < RW; ++r)
      if (q0 + r < d.Q) dot_store<R, CPLX, MIXED>(g, d, q0 + r, accr[r], acci[r]);
  }
}

// Lanes along the free extent (the matrix is free-contiguous): 64 outputs per
// workgroup, each wave one contiguous quarter of K.  A thread walks its
// quarter in order with 8 k-steps of loads issued together (the K offsets
// and vector elements are wave-uniform); the four quarters are then added in
// order 0..3.
template <typename R, bool CPLX, bool MIXED>
__global__ __launch_bounds__(256) void dot_qlanes_kernel(const GemmArgs g, const DotArgs d) {
  constexpr int S = 4, U = 8;
  __shared__ R part[S][64][CPLX ? 2 : 1];
  const int lane = threadIdx.x & 63;
  const int s = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t q = (int64_t)blockIdx.x * 64 + lane;
  const bool qok = q < d.Q;
  const int64_t xrow = d.xbase + d.fX[qok ? q : d.Q - 1];
  const int64_t chunk = (d.K + S - 1) / S;
  const int64_t kb = s * chunk, ke = kb + chunk < d.K ? kb + chunk : d.K;
  R accr = 0, acci = 0;
  for (int64_t k0 = kb; k0 < ke; k0 += U) {
    R xr[U], xi[U], vr[U], vi[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t k = k0 + u < ke ? k0 + u : ke - 1;
      load_elem<R, CPLX, MIXED>(d.V, d.v_type, d.vbase + ((qs_const_i64)d.kv)[k], vr[u], vi[u]);
      load_elem<R, CPLX, MIXED>(d.X, d.x_type, xrow + ((qs_const_i64)d.kX)[k], xr[u], xi[u]);
    }
    __builtin_amdgcn_sched_barrier(0);   // every load of the step is in flight
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool in = k0 + u < ke;
      if (!CPLX) {
        const R nr = qs_fma(xr[u], vr[u], accr);
        accr = in ? nr : accr;
      } else {
        const R x_i = d.conj_x ? -xi[u] : xi[u], v_i = d.conj_v ? -vi[u] : vi[u];
        const R nr = qs_fma(-x_i, v_i, qs_fma(xr[u], vr[u], accr));
        const R ni = qs_fma(x_i, vr[u], qs_fma(xr[u], v_i, acci));
        accr = in ? nr : accr;
        acci = in ? ni : acci;
      }
    }
  }
  part[s][lane][0] = accr;
  if (CPLX) part[s][lane][CPLX ? 1 : 0] = acci;
  __syncthreads();
  if (s == 0 && qok) {
    R tr = part[0][lane][0], ti = CPLX ? part[0][lane][CPLX ? 1 : 0] : (R)0;
#pragma unroll
    for (int w = 1; w < S; ++w) {
      tr += part[w][lane][0];
      if (CPLX) ti += part[w][lane][CPLX ? 1 : 0];
    }
    dot_store<R, CPLX, MIXED>(g, d, q, tr, ti);
  }
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
    const int64_t blocks = (d.Q + 15) / 16;
    hipLaunchKernelGGL((dot_klanes_kernel<R, CPLX, MIXED>), dim3((unsigned)blocks), dim3(256), 0, s, g, d);
  } else {
    const int64_t blocks = (d.Q + 63) / 64;
    hipLaunchKernelGGL((dot_qlanes_kernel<R, CPLX, MIXED>), dim3((unsigned)blocks), dim3(256), 0, s, g, d);
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------ outer
// K == 1, in strips of 256 x 16 elements of C: threads laid along C's
// unit-stride side (`a_kfast` carries "N fastest" for this path), each
// thread owning one row (column) and walking 16 columns (rows) whose offsets
// and operand elements are workgroup-uniform.  Per thread, the 16 old-C
// loads are issued before the first store.
template <typename R, bool CPLX, bool MIXED, bool NFAST>
__global__ __launch_bounds__(256) void outer_kernel(const GemmArgs g) {
  // Slow indices per thread.  8 measured best at M = N = 8192, beta = 0:
  // 112 us against 124 (16), 160 (32), 120 (4) and 150 (2).
  constexpr int W = 8;
  const int64_t fast = NFAST ? g.N : g.M, slow = NFAST ? g.M : g.N;
  const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (f >= fast) return;
  // The fast side's operand element and C offset, once per thread.
  R fr, fi;
  int64_t foff;
  if (NFAST) {
    load_elem<R, CPLX, MIXED>(g.B, g.b_type, g.b_base + g.nB[f] + g.kB[0], fr, fi);
    if (CPLX && g.conj_b) fi = -fi;
    foff = g.nC[f];
  } else {
    load_elem<R, CPLX, MIXED>(g.A, g.a_type, g.a_base + g.mA[f] + g.kA[0], fr, fi);
    if (CPLX && g.conj_a) fi = -fi;
    foff = g.mC[f];
  }
  for (int64_t s0 = (int64_t)blockIdx.y * W; s0 < slow; s0 += (int64_t)gridDim.y * W) {
  int64_t idx[W];
  bool ok[W];
  R xr[W], xi[W], o[W], oi[W];
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const int64_t sl = s0 + w;
    ok[w] = sl < slow;
    const int64_t sc = ok[w] ? sl : slow - 1;
    R sr, si;
    // sc is workgroup-uniform: its table entries come through the scalar cache.
    if (NFAST) {
      load_elem<R, CPLX, MIXED>(g.A, g.a_type, g.a_base + ((qs_const_i64)g.mA)[sc] + ((qs_const_i64)g.kA)[0], sr, si);
      if (CPLX && g.conj_a) si = -si;
      idx[w] = g.c_base + ((qs_const_i64)g.mC)[sc] + foff;
    } else {
      load_elem<R, CPLX, MIXED>(g.B, g.b_type, g.b_base + ((qs_const_i64)g.nB)[sc] + ((qs_const_i64)g.kB)[0], sr, si);
      if (CPLX && g.conj_b) si = -si;
      idx[w] = g.c_base + foff + ((qs_const_i64)g.nC)[sc];
    }
    // a_i * b_j in the order the scalar kernel forms it (A's element first).
    const R ar = NFAST ? sr : fr, ai = NFAST ? si : fi, br = NFAST ? fr : sr, bi = NFAST ? fi : si;
    if (!CPLX) {
      xr[w] = qs_fma(ar, br, (R)0);
      xi[w] = 0;
    } else {
      xr[w] = qs_fma(-ai, bi, qs_fma(ar, br, (R)0));
      xi[w] = qs_fma(ai, br, qs_fma(ar, bi, (R)0));
    }
    load_c_old<R, CPLX, MIXED>(g, idx[w], o[w], oi[w]);   // idx is valid: clamped
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int w = 0; w < W; ++w)
    if (ok[w]) store_c_at<R, CPLX, MIXED, true>(g, idx[w], xr[w], xi[w], o[w], oi[w]);
  }
}

template <typename R, bool CPLX, bool MIXED>
static hipError_t launch_outer_typed(const GemmArgs &g, hipStream_t s) {
  const int64_t fast = g.a_kfast ? g.N : g.M, slow = g.a_kfast ? g.M : g.N;
  const int64_t gy = (slow + 7) / 8;
  const dim3 grid((unsigned)((fast + 255) / 256), (unsigned)(gy < 65535 ? gy : 65535));
  if (g.a_kfast)
    hipLaunchKernelGGL((outer_kernel<R, CPLX, MIXED, true>), grid, dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL((outer_kernel<R, CPLX, MIXED, false>), grid, dim3(256), 0, s, g);
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
