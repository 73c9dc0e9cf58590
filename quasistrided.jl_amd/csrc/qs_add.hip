// Permuted add and partial trace: the device counterparts of TensorOperations'
// tensoradd! / tensortrace!, which the reference forwards to StridedNative
// (src/integrations/tensoroperations.jl:90-107).
//
//   C[m, n] = alpha * sum_k op(A)[m, n, k] + beta * C[m, n]
//
// M and N are groups of C's axes (two maps each: A, C), K the traced pairs
// (map 0: the sum of the pair's strides in A).  With both M and N longer than
// one, each workgroup owns a 64 x 64 tile: it reads the tile with lanes laid
// along A's unit-stride side (the leading N dim when A's unit-stride axis is
// there), sums K in table order, turns the tile through LDS and writes it
// with lanes along C's unit-stride side (the leading M dim), so a transposing
// permutation is coalesced on both sides.  Row and column offsets are read
// once per lane or through the scalar cache; every load of a batch is issued
// before its arithmetic and every old-C load before the first store.  A 1-D
// strip kernel covers M == 1 or N == 1.
#include <hip/hip_runtime.h>

#include <vector>

#include "qs_gemm_impl.h"
#include "qs_plan.h"

namespace qs {

struct AddArgs {
  GemmArgs g;   // the epilogue's fields: C, its type and base, alpha / beta, mC / nC, M, N
  const void *A;
  int a_type, conj_a;
  int64_t a_base;
  const int64_t *mA, *nA, *kA;
  int64_t K;
};

// 64 x 64 tiles.  ANF: A's unit-stride side is N (read lanes along N).
template <typename R, bool CPLX, bool MIXED, bool ANF>
__global__ __launch_bounds__(256) void add_tile_kernel(const AddArgs a) {
  constexpr int P = CPLX ? 2 : 1;
  __shared__ R tile[P][64][65];
  const GemmArgs &g = a.g;
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t tiles_m = (g.M + 63) / 64;
  const int64_t m0 = ((int64_t)blockIdx.x % tiles_m) * 64, n0 = ((int64_t)blockIdx.x / tiles_m) * 64;

  // Read phase: lane-fixed coordinate f (N when ANF, else M) and 16
  // wave-uniform coordinates v = w + 4 it of the other group.
  const int64_t f = (ANF ? n0 : m0) + lane;
  const int64_t flen = ANF ? g.N : g.M, vlen = ANF ? g.M : g.N;
  const int64_t fA = a.a_base + (ANF ? a.nA : a.mA)[f < flen ? f : 0];
  int64_t vA[16];
#pragma unroll
  for (int it = 0; it < 16; ++it) {
    const int64_t v = (ANF ? m0 : n0) + w + 4 * it;
    vA[it] = ((qs_const_i64)(ANF ? a.mA : a.nA))[v < vlen ? v : 0];
  }
  R xr[16], xi[16];
#pragma unroll
  for (int it = 0; it < 16; ++it) xr[it] = xi[it] = (R)0;
  // A is read once: non-temporal loads keep it from evicting the offset
  // tables and C's lines (8192 x 8192 transpose: 227 -> 209 us).
  for (int64_t k = 0; k < a.K; ++k) {
    const int64_t ko = ((qs_const_i64)a.kA)[k];
    R yr[16], yi[16];
#pragma unroll
    for (int it = 0; it < 16; ++it) load_elem<R, CPLX, MIXED, true>(a.A, a.a_type, fA + vA[it] + ko, yr[it], yi[it]);
#pragma unroll
    for (int it = 0; it < 16; ++it) {
      xr[it] += yr[it];
      if (CPLX) xi[it] += a.conj_a ? -yi[it] : yi[it];
    }
  }
#pragma unroll
  for (int it = 0; it < 16; ++it) {
    const int i = ANF ? w + 4 * it : lane, j = ANF ? lane : w + 4 * it;
    tile[0][i][j] = xr[it];
    if (CPLX) tile[P - 1][i][j] = xi[it];
  }
  __syncthreads();

  // Write phase: lanes along M (C's unit-stride side), columns w + 4 it.
  const int64_t row = m0 + lane;
  if (row >= g.M) return;
  const int64_t rowC = g.c_base + g.mC[row];
  int64_t idx[16];
  bool ok[16];
  R o[16], oi[16];
#pragma unroll
  for (int it = 0; it < 16; ++it) {
    const int64_t col = n0 + w + 4 * it;
    ok[it] = col < g.N;
    idx[it] = rowC + ((qs_const_i64)g.nC)[ok[it] ? col : 0];
    load_c_old<R, CPLX, MIXED>(g, idx[it], o[it], oi[it]);
  }
#pragma unroll
  for (int it = 0; it < 16; ++it)
    if (ok[it])
      store_c_at<R, CPLX, MIXED, true>(g, idx[it], tile[0][lane][w + 4 * it], CPLX ? tile[P - 1][lane][w + 4 * it] : (R)0,
                                 o[it], oi[it]);
}

// M == 1 or N == 1: one element per thread along the longer group, four
// elements per thread in flight.
template <typename R, bool CPLX, bool MIXED>
__global__ __launch_bounds__(256) void add_line_kernel(const AddArgs a) {
  const GemmArgs &g = a.g;
  const bool along_m = g.N == 1;
  const int64_t len = along_m ? g.M : g.N;
  const int64_t e0 = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  const int64_t *tA = along_m ? a.mA : a.nA;
  const int64_t *tC = along_m ? g.mC : g.nC;
  const int64_t fixA = a.a_base + (along_m ? a.nA[0] : a.mA[0]);
  const int64_t fixC = g.c_base + (along_m ? g.nC[0] : g.mC[0]);
  int64_t offA[4], idx[4];
  bool ok[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int64_t e = e0 + 256 * u;
    ok[u] = e < len;
    offA[u] = fixA + tA[ok[u] ? e : 0];
    idx[u] = fixC + tC[ok[u] ? e : 0];
  }
  R xr[4], xi[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) xr[u] = xi[u] = (R)0;
  for (int64_t k = 0; k < a.K; ++k) {
    const int64_t ko = ((qs_const_i64)a.kA)[k];
    R yr[4], yi[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) load_elem<R, CPLX, MIXED>(a.A, a.a_type, offA[u] + ko, yr[u], yi[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xr[u] += yr[u];
      if (CPLX) xi[u] += a.conj_a ? -yi[u] : yi[u];
    }
  }
  R o[4], oi[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) load_c_old<R, CPLX, MIXED>(g, idx[u], o[u], oi[u]);
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (ok[u]) store_c_at<R, CPLX, MIXED, true>(g, idx[u], xr[u], xi[u], o[u], oi[u]);
}

template <typename R, bool CPLX, bool MIXED>
static hipError_t launch_add_typed(const AddArgs &a, bool anf, hipStream_t s) {
  const GemmArgs &g = a.g;
  if (g.M == 1 || g.N == 1) {
    const int64_t len = g.N == 1 ? g.M : g.N;
    hipLaunchKernelGGL((add_line_kernel<R, CPLX, MIXED>), dim3((unsigned)((len + 1023) / 1024)), dim3(256), 0, s, a);
  } else {
    const int64_t tiles = ((g.M + 63) / 64) * ((g.N + 63) / 64);
    if (anf)
      hipLaunchKernelGGL((add_tile_kernel<R, CPLX, MIXED, true>), dim3((unsigned)tiles), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((add_tile_kernel<R, CPLX, MIXED, false>), dim3((unsigned)tiles), dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

static hipError_t launch_add(const AddArgs &a, bool anf, bool mixed, hipStream_t s) {
  switch (a.g.compute_type) {
    case QS_F64: return mixed ? launch_add_typed<double, false, true>(a, anf, s) : launch_add_typed<double, false, false>(a, anf, s);
    case QS_F32: return mixed ? launch_add_typed<float, false, true>(a, anf, s) : launch_add_typed<float, false, false>(a, anf, s);
    case QS_C64: return mixed ? launch_add_typed<double, true, true>(a, anf, s) : launch_add_typed<double, true, false>(a, anf, s);
    default: return mixed ? launch_add_typed<float, true, true>(a, anf, s) : launch_add_typed<float, true, false>(a, anf, s);
  }
}

}  // namespace qs

using namespace qs;

struct qs_add_plan {
  qs_add_desc desc;
  int64_t M = 0, N = 0, K = 0;
  int anf = 0, mixed = 0;
  int64_t *d_tables = nullptr;   // mA mC nA nC kA
  int64_t mA_lo = 0, mA_hi = 0, mC_lo = 0, mC_hi = 0, nA_lo = 0, nA_hi = 0, nC_lo = 0, nC_hi = 0, kA_lo = 0,
          kA_hi = 0;
};

static double add_scalar_at(const void *p, int ct, int part) {
  if (ct == QS_F32 || ct == QS_C32) return (double)((const float *)p)[part];
  return ((const double *)p)[part];
}

extern "C" {

qs_status qs_add_plan_create(const qs_add_desc *desc, qs_add_plan **out) {
  if (!desc || !out) return fail(QS_ERR_ARGUMENT, "tensoradd!: null argument");
  *out = nullptr;
  const qs_add_desc &d = *desc;
  for (int t : {d.a_type, d.c_type, d.compute_type})
    if (t < QS_F32 || t > QS_C64)
      return fail(QS_ERR_ARGUMENT, "tensoradd!: eltypes must be Float32, Float64, ComplexF32 or ComplexF64");
  if (is_complex(d.a_type) && !is_complex(d.c_type))
    return fail(QS_ERR_ARGUMENT, "tensoradd!: a complex A needs a complex C");
  if (is_complex(d.compute_type) != (is_complex(d.a_type) || is_complex(d.c_type)))
    return fail(QS_ERR_ARGUMENT, "tensoradd!: the compute type's domain must be the promoted domain of A and C");
  qs_status st;
  if ((st = validate_group(d.m, "M")) != QS_OK) return st;
  if ((st = validate_group(d.n, "N")) != QS_OK) return st;
  if ((st = validate_group(d.k, "K")) != QS_OK) return st;
  if (d.a_len < 0 || d.c_len < 0) return fail(QS_ERR_ARGUMENT, "tensoradd!: negative storage length");

  qs_add_plan *p = new qs_add_plan();
  p->desc = d;
  qs_group mg = d.m, ng = d.n, kg = d.k;
  p->M = group_length(mg);
  p->N = group_length(ng);
  p->K = group_length(kg);
  map_range(mg, 0, p->mA_lo, p->mA_hi);
  map_range(mg, 1, p->mC_lo, p->mC_hi);
  map_range(ng, 0, p->nA_lo, p->nA_hi);
  map_range(ng, 1, p->nC_lo, p->nC_hi);
  map_range(kg, 0, p->kA_lo, p->kA_hi);
  const int ct = d.compute_type;
  p->mixed = !(d.a_type == ct && d.c_type == ct);
  // Coalescing: C's unit-stride dim leads M; A's leads N if it is there
  // (then the tile is read along N), else M.
  const int cm = unit_dim(mg, 1);
  if (cm >= 0) move_to_front(mg, cm);
  const int an = unit_dim(ng, 0);
  if (an >= 0) move_to_front(ng, an);
  p->anf = an >= 0 && !(mg.lengths[0] > 1 && (mg.strides[0][0] == 1 || mg.strides[0][0] == -1));

  const int64_t M = p->M, N = p->N, K = p->K;
  const int64_t nt = 2 * M + 2 * N + K;
  if (nt > 0 && M > 0 && N > 0) {
    std::vector<int64_t> h((size_t)nt);
    int64_t *mA = h.data(), *mC = mA + M, *nA = mC + M, *nC = nA + N, *kA = nC + N;
    fill_offsets(mg, 0, M, mA, mC);
    fill_offsets(ng, 0, N, nA, nC);
    std::vector<int64_t> unused((size_t)(K > 0 ? K : 1));
    if (K > 0) fill_offsets(kg, 0, K, kA, unused.data());
    if (dev_malloc(&p->d_tables, sizeof(int64_t) * (size_t)nt) != hipSuccess) {
      (void)hipGetLastError();
      delete p;
      return fail(QS_ERR_MEMORY, "tensoradd!: device tables");
    }
    if (hipMemcpy(p->d_tables, h.data(), sizeof(int64_t) * (size_t)nt, hipMemcpyHostToDevice) != hipSuccess) {
      (void)dev_free(p->d_tables);
      delete p;
      return fail(QS_ERR_DEVICE, "tensoradd!: table upload");
    }
  }
  *out = p;
  return QS_OK;
}

qs_status qs_add_execute(qs_add_plan *p, const void *alpha, const void *beta, void *stream) {
  if (!p || !alpha || !beta) return fail(QS_ERR_ARGUMENT, "tensoradd!: null argument");
  const qs_add_desc &d = p->desc;
  const int ct = d.compute_type;
  const double al[2] = {add_scalar_at(alpha, ct, 0), is_complex(ct) ? add_scalar_at(alpha, ct, 1) : 0.0};
  const double be[2] = {add_scalar_at(beta, ct, 0), is_complex(ct) ? add_scalar_at(beta, ct, 1) : 0.0};
  const int64_t M = p->M, N = p->N, K = p->K;
  if (M == 0 || N == 0) return QS_OK;
  const bool reads_a = !(al[0] == 0 && al[1] == 0) && K > 0;
  // Exact spans before any write (tiles.jl:120-134).
  const __int128 clo = (__int128)d.c_base + p->mC_lo + p->nC_lo, chi = (__int128)d.c_base + p->mC_hi + p->nC_hi;
  if (clo < 0 || chi > (__int128)d.c_len - 1)
    return fail(QS_ERR_BOUNDS, "tensoradd!: C addresses [" + std::to_string((long long)clo) + ", " +
                                   std::to_string((long long)chi) + "] exceed storage bounds [0, " +
                                   std::to_string(d.c_len - 1) + "]");
  if (reads_a) {
    const __int128 alo = (__int128)d.a_base + p->mA_lo + p->nA_lo + p->kA_lo;
    const __int128 ahi = (__int128)d.a_base + p->mA_hi + p->nA_hi + p->kA_hi;
    if (alo < 0 || ahi > (__int128)d.a_len - 1)
      return fail(QS_ERR_BOUNDS, "tensoradd!: A addresses [" + std::to_string((long long)alo) + ", " +
                                     std::to_string((long long)ahi) + "] exceed storage bounds [0, " +
                                     std::to_string(d.a_len - 1) + "]");
  }
  AddArgs a{};
  GemmArgs &g = a.g;
  g.C = d.c;
  g.c_type = d.c_type;
  g.c_base = d.c_base;
  g.compute_type = ct;
  g.M = M;
  g.N = N;
  g.K = K;
  g.mC = p->d_tables + M;
  g.nC = p->d_tables + 2 * M + N;
  g.alpha_re = al[0];
  g.alpha_im = al[1];
  g.beta_re = be[0];
  g.beta_im = be[1];
  g.beta_case = (be[0] == 0 && be[1] == 0) ? 0 : (be[0] == 1 && be[1] == 0) ? 1 : 2;
  g.mixed = p->mixed;
  g.k_splits = 1;
  a.A = d.a;
  a.a_type = d.a_type;
  a.conj_a = d.conj_a;
  a.a_base = d.a_base;
  a.mA = p->d_tables;
  a.nA = p->d_tables + 2 * M;
  a.kA = p->d_tables + 2 * M + 2 * N;
  a.K = K;
  hipStream_t s = (hipStream_t)stream;
  hipError_t e;
  if (!reads_a) {
    if (g.beta_case == 1) return QS_OK;   // C = 1 * C
    e = launch_scale(g, s);               // alpha == 0 or K == 0: C = beta * C, A never read
  } else {
    e = launch_add(a, p->anf != 0, p->mixed != 0, s);
  }
  if (e != hipSuccess) return fail(QS_ERR_DEVICE, std::string("tensoradd!: launch: ") + hipGetErrorString(e));
  return QS_OK;
}

qs_status qs_add_plan_rebind(qs_add_plan *p, const void *a, int64_t a_base, int64_t a_len, void *c, int64_t c_base,
                             int64_t c_len) {
  if (!p || !c || (!a && a_len > 0)) return fail(QS_ERR_ARGUMENT, "tensoradd!: null argument");
  if (a_len < 0 || c_len < 0) return fail(QS_ERR_ARGUMENT, "tensoradd!: negative storage length");
  p->desc.a = a; p->desc.a_base = a_base; p->desc.a_len = a_len;
  p->desc.c = c; p->desc.c_base = c_base; p->desc.c_len = c_len;
  return QS_OK;
}

qs_status qs_add_plan_destroy(qs_add_plan *p) {
  if (!p) return QS_OK;
  if (p->d_tables) (void)dev_free(p->d_tables);
  delete p;
  return QS_OK;
}

}  // extern "C"
