// A contraction that allocates nothing: plan and execute in one call, with
// the plan's offset tables in a fixed device array instead of a library
// allocation.  For callers that cannot allocate (code under HIP graph
// capture, serving loops that forbid hipMalloc): the success path makes no
// device or host allocation.  The plan is built on the stack, the tables are
// written by a kernel on the caller's stream into `scratch`, and the
// contraction runs unsplit (no split-K partial sums), with the same kernels
// as qs_execute.
//
// `scratch` is one array per GPU, shared by every call on that GPU and
// holding the last call's tables: calls that overlap on one GPU (on two
// streams, or from two threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>

#include "qs_plan.h"

namespace qs {

// Table entries the array holds: the six tables of one plan take
// 2 (M + N + K + kKTablePad) entries.  4 MiB per GPU.
static constexpr int64_t kNoallocEntries = int64_t(1) << 19;
static constexpr int kMaxDevices = 64;

static __device__ __align__(256) int64_t scratch[kNoallocEntries];

// The array's device address on each GPU (the host-side name is only a
// stand-in), taken once per device with that device current.
static std::mutex g_scratch_mu;
static int64_t *g_scratch_ptr[kMaxDevices];

static qs_status scratch_for_device(int64_t **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail_fmt(QS_ERR_DEVICE, "qs_contract_noalloc: hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail_fmt(QS_ERR_UNSUPPORTED, "qs_contract_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail_fmt(QS_ERR_DEVICE, "qs_contract_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = (int64_t *)ptr;
  }
  *out = g_scratch_ptr[dev];
  return QS_OK;
}

// The six tables of one plan, in the layout set_table_pointers expects
// (mA mC nB nC kA kB; the K tables kKTablePad entries longer than K).  One
// thread per entry of the two maps: entry j of a group is the mixed-radix
// decode of j (first dim fastest), as fill_offsets! enumerates it
// (src/layout/axis_group.jl:87-145).
struct FillArgs {
  qs_group m, n, k;
  int64_t M, N, K;
  int64_t padA, padB;
  int64_t *t;
};

__device__ __forceinline__ void decode(const qs_group &g, int64_t j, int64_t &o0, int64_t &o1) {
  o0 = 0;
  o1 = 0;
  for (int d = 0; d < g.ndim; ++d) {
    const int64_t L = g.lengths[d];
    if (L == 1) continue;
    const int64_t x = j % L;
    j /= L;
    o0 += x * g.strides[0][d];
    o1 += x * g.strides[1][d];
  }
}

__global__ __launch_bounds__(256) void noalloc_fill_kernel(const FillArgs a) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t KP = a.K + kKTablePad;
  int64_t *mA = a.t, *mC = mA + a.M, *nB = mC + a.M, *nC = nB + a.N, *kA = nC + a.N, *kB = kA + KP;
  int64_t o0, o1;
  if (j < a.M) {
    decode(a.m, j, o0, o1);
    mA[j] = o0;
    mC[j] = o1;
  } else if (j < a.M + a.N) {
    decode(a.n, j - a.M, o0, o1);
    nB[j - a.M] = o0;
    nC[j - a.M] = o1;
  } else if (j < a.M + a.N + KP) {
    const int64_t k = j - a.M - a.N;
    if (k < a.K) {
      decode(a.k, k, o0, o1);
    } else {
      o0 = a.padA;
      o1 = a.padB;
    }
    kA[k] = o0;
    kB[k] = o1;
  }
}

}  // namespace qs

using namespace qs;

extern "C" {

qs_status qs_contract_noalloc(const qs_contract_desc *desc, const void *alpha, const void *beta, void *stream) {
  if (!desc || !alpha || !beta) return fail_fmt(QS_ERR_ARGUMENT, "qs_contract_noalloc: null argument");
  if (desc->memory != QS_MEM_DEVICE)
    return fail_fmt(QS_ERR_UNSUPPORTED, "qs_contract_noalloc: device memory only (host operands need staging buffers)");
  qs_plan p;   // on the stack: no tables, no workspace of its own
  qs_status st = plan_setup(*desc, &p);
  if (st != QS_OK) return st;
  if (p.path == QS_PATH_SPLIT)
    return fail_fmt(QS_ERR_UNSUPPORTED, "qs_contract_noalloc: split precision needs a workspace (use qs_plan_create)");
  const int64_t entries = 2 * (p.M + p.N + p.K + kKTablePad);
  if (entries > kNoallocEntries)
    return fail_fmt(QS_ERR_UNSUPPORTED,
                    "qs_contract_noalloc: the offset tables need %lld entries (2 (M + N + K + %lld)); the static "
                    "scratch holds %lld (M + N + K <= %lld)",
                    (long long)entries, (long long)kKTablePad, (long long)kNoallocEntries,
                    (long long)(kNoallocEntries / 2 - kKTablePad));
  const int ct = p.desc.compute_type;
  const bool cplx = is_complex(ct);
  const double al[2] = {ct == QS_F32 || ct == QS_C32 ? (double)((const float *)alpha)[0] : ((const double *)alpha)[0],
                        !cplx ? 0.0 : ct == QS_C32 ? (double)((const float *)alpha)[1] : ((const double *)alpha)[1]};
  const double be[2] = {ct == QS_F32 || ct == QS_C32 ? (double)((const float *)beta)[0] : ((const double *)beta)[0],
                        !cplx ? 0.0 : ct == QS_C32 ? (double)((const float *)beta)[1] : ((const double *)beta)[1]};
  // Exact span checks before anything is launched.
  st = check_spans(&p, !(al[0] == 0 && al[1] == 0));
  if (st != QS_OK) return st;
  if (p.path == QS_PATH_EMPTY) return QS_OK;
  int64_t *tab = nullptr;
  if ((st = scratch_for_device(&tab)) != QS_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  FillArgs fa;
  fa.m = p.mg;
  fa.n = p.ng;
  fa.k = p.kg;
  fa.M = p.M;
  fa.N = p.N;
  fa.K = p.K;
  fa.padA = k_pad_offset(&p, 0);
  fa.padB = k_pad_offset(&p, 1);
  fa.t = tab;
  const int64_t rows = p.M + p.N + p.K + kKTablePad;
  hipLaunchKernelGGL(noalloc_fill_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, fa);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail_fmt(QS_ERR_DEVICE, "qs_contract_noalloc: table fill: %s", hipGetErrorString(e));
  set_table_pointers(&p, tab);
  if (p.path == QS_PATH_GEMM || p.path == QS_PATH_SKINNY) {
    // The tile override or the heuristic; K unsplit (no partial sums).
    p.tile = (desc->tile >= 0 && desc->tile < TILE_COUNT) ? desc->tile : choose_tile(p.M, p.N, ct);
    p.k_splits = 1;
    p.k_chunk = p.K;
  }
  return run_device(&p, p.desc.a, p.desc.a_base, p.desc.b, p.desc.b_base, p.desc.c, p.desc.c_base, al, be, s);
}

}  // extern "C"
