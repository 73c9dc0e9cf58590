// The C ABI (include/qs_amd.h): validation, plans, path and tile choice,
// launches.  Host code only; kernels live in the .hip files.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "qs_internal.h"
#include "qs_plan.h"

namespace qs {

// A fixed buffer: setting the message allocates nothing.
static thread_local char g_last_error[1024];

static void set_error_text(const char *msg) { std::snprintf(g_last_error, sizeof g_last_error, "%s", msg); }

void set_error(const std::string &msg) { set_error_text(msg.c_str()); }

qs_status fail(qs_status code, const std::string &msg) {
  set_error(msg);
  return code;
}

qs_status fail_fmt(qs_status code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
  va_end(ap);
  return code;
}

static std::atomic<int64_t> g_dev_allocs{0}, g_dev_frees{0}, g_dev_bytes{0}, g_pin_allocs{0}, g_pin_frees{0};
static std::atomic<int64_t> g_bridge_allocs{0}, g_bridge_frees{0};

// Armed one-shot failures (qs_fault_inject): the number of allocations of the
// kind still to succeed before one fails; negative when disarmed.
static std::atomic<int64_t> g_fault_left[2] = {{-1}, {-1}};

static bool fault_fires(int kind) {
  std::atomic<int64_t> &c = g_fault_left[kind];
  int64_t v = c.load();
  while (v >= 0)
    if (c.compare_exchange_weak(v, v - 1)) return v == 0;
  return false;
}

hipError_t dev_malloc(void **p, size_t bytes) {
  if (fault_fires(QS_FAULT_DEVICE_ALLOC)) {
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  const hipError_t e = hipMalloc(p, bytes);
  if (e == hipSuccess) {
    g_dev_allocs.fetch_add(1, std::memory_order_relaxed);
    g_dev_bytes.fetch_add((int64_t)bytes, std::memory_order_relaxed);
  }
  return e;
}
hipError_t dev_free(void *p) {
  if (p) g_dev_frees.fetch_add(1, std::memory_order_relaxed);
  return hipFree(p);
}
hipError_t pinned_malloc(void **p, size_t bytes, unsigned int flags) {
  if (fault_fires(QS_FAULT_PINNED_ALLOC)) {
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  const hipError_t e = hipHostMalloc(p, bytes, flags);
  if (e == hipSuccess) g_pin_allocs.fetch_add(1, std::memory_order_relaxed);
  return e;
}
hipError_t pinned_free(void *p) {
  if (p) g_pin_frees.fetch_add(1, std::memory_order_relaxed);
  return hipHostFree(p);
}
void count_bridge(bool alloc) { (alloc ? g_bridge_allocs : g_bridge_frees).fetch_add(1, std::memory_order_relaxed); }

static qs_status hip_fail(hipError_t e, const char *what) {
  return fail(QS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// ------------------------------------------------------------------ groups

int64_t group_length(const qs_group &g) {
  int64_t q = 1;
  for (int d = 0; d < g.ndim; ++d) q *= g.lengths[d];
  return q;
}

// AxisGroup validation, src/layout/axis_group.jl:9-41.
qs_status validate_group(const qs_group &g, const char *name) {
  if (g.ndim < 0 || g.ndim > QS_MAX_DIMS)
    return fail(QS_ERR_ARGUMENT, std::string(name) + " group: ndim out of range [0, 16]");
  for (int d = 0; d < g.ndim; ++d)
    if (g.lengths[d] < 0)
      return fail(QS_ERR_ARGUMENT, std::string(name) + " group: AxisGroup lengths must be nonnegative, got lengths[" +
                                       std::to_string(d + 1) + "] = " + std::to_string(g.lengths[d]));
  for (int d = 0; d < g.ndim; ++d)
    if (g.lengths[d] == 0) return QS_OK;
  int64_t card = 1;
  for (int d = 0; d < g.ndim; ++d)
    if (__builtin_mul_overflow(card, g.lengths[d], &card))
      return fail(QS_ERR_OVERFLOW, std::string(name) + " group: cardinality overflows Int64");
  for (int p = 0; p < 2; ++p) {
    int64_t acc = 0;
    for (int d = 0; d < g.ndim; ++d) {
      if (g.lengths[d] <= 1) continue;
      const int64_t s = g.strides[p][d];
      if (s == INT64_MIN) return fail(QS_ERR_OVERFLOW, std::string(name) + " group: |stride| overflows Int64");
      int64_t term;
      if (__builtin_mul_overflow(g.lengths[d] - 1, s < 0 ? -s : s, &term) || __builtin_add_overflow(acc, term, &acc))
        return fail(QS_ERR_OVERFLOW, std::string(name) + " group: offset range overflows Int64");
    }
  }
  return QS_OK;
}

// fill_offsets! (axis_group.jl:87-145) for both maps of a two-map group.
void fill_offsets(const qs_group &g, int64_t first, int64_t count, int64_t *o0, int64_t *o1) {
  if (count <= 0) return;
  int64_t x[QS_MAX_DIMS] = {0};
  int64_t a = 0, b = 0, r = first;
  for (int d = 0; d < g.ndim; ++d) {
    const int64_t L = g.lengths[d];
    x[d] = r % L;
    r /= L;
    a += x[d] * g.strides[0][d];
    b += x[d] * g.strides[1][d];
  }
  for (int64_t t = 0;;) {
    o0[t] = a;
    o1[t] = b;
    if (++t == count) break;
    for (int d = 0; d < g.ndim; ++d) {
      const int64_t L = g.lengths[d];
      if (x[d] < L - 1) {
        ++x[d];
        a += g.strides[0][d];
        b += g.strides[1][d];
        break;
      }
      x[d] = 0;
      a -= (L - 1) * g.strides[0][d];
      b -= (L - 1) * g.strides[1][d];
    }
  }
}

// Exact offset range of map p (per dim min(0,(L-1)s) / max(0,(L-1)s)).
void map_range(const qs_group &g, int p, int64_t &lo, int64_t &hi) {
  lo = hi = 0;
  for (int d = 0; d < g.ndim; ++d) {
    if (g.lengths[d] <= 1) continue;
    const int64_t e = (g.lengths[d] - 1) * g.strides[p][d];
    if (e < 0) lo += e; else hi += e;
  }
}

// Move dim `d` of group `g` to the front (the enumeration order changes; the
// set of (map0, map1) offset pairs does not).
void move_to_front(qs_group &g, int d) {
  if (d <= 0) return;
  const int64_t L = g.lengths[d], s0 = g.strides[0][d], s1 = g.strides[1][d];
  for (int e = d; e > 0; --e) {
    g.lengths[e] = g.lengths[e - 1];
    g.strides[0][e] = g.strides[0][e - 1];
    g.strides[1][e] = g.strides[1][e - 1];
  }
  g.lengths[0] = L;
  g.strides[0][0] = s0;
  g.strides[1][0] = s1;
}

// A non-singleton dim of map p with |stride| == 1, or -1.
int unit_dim(const qs_group &g, int p) {
  for (int d = 0; d < g.ndim; ++d)
    if (g.lengths[d] > 1 && (g.strides[p][d] == 1 || g.strides[p][d] == -1)) return d;
  return -1;
}

// --------------------------------------------------------- device facts

// XCDs of a part: not a device property; the CDNA3/4 parts deal workgroups
// over 8 XCDs (gfx950: 8 x 32 CUs; gfx942: 38 CUs each, 6 on MI300A).
static int xcds_of(const char *arch, int cus) {
  if (std::strncmp(arch, "gfx950", 6) == 0) return 8;
  if (std::strncmp(arch, "gfx942", 6) == 0) return cus % 38 == 0 ? cus / 38 : 8;
  return 1;
}

const DeviceFacts &device_facts() {
  static std::mutex mu;
  static DeviceFacts facts[64];
  static bool have[64] = {false};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  std::lock_guard<std::mutex> lk(mu);
  if (!have[dev]) {
    DeviceFacts f;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, dev) == hipSuccess) {
      std::snprintf(f.arch, sizeof f.arch, "%s", pr.gcnArchName);
      if (pr.multiProcessorCount > 0) f.cus = pr.multiProcessorCount;
      if (pr.warpSize > 0) f.wave = pr.warpSize;
      if (pr.maxSharedMemoryPerMultiProcessor > 0) f.lds_per_cu = (int)pr.maxSharedMemoryPerMultiProcessor;
      if (pr.sharedMemPerBlock > 0) f.lds_per_block = (int)pr.sharedMemPerBlock;
      f.xcds = xcds_of(f.arch, f.cus);
    } else {
      (void)hipGetLastError();
    }
    facts[dev] = f;
    have[dev] = true;
  }
  return facts[dev];
}

// ------------------------------------------------------------------ tiles

// Mirrors pick_bk_for in qs_gemm_impl.h.
static int pick_bk_host(int bm, int bn, int ct, bool akf, bool bkf) {
  return pick_bk_for(is_double(ct) ? 8 : 4, is_complex(ct) ? 2 : 1, bm, bn, akf, bkf);
}

TileShape tile_shape(int tile, int ct, bool akf, bool bkf) {
  int bm = 64, bn = 64;
  switch (tile) {
    case TILE_128x128: bm = 128; bn = 128; break;
    case TILE_128x64: bm = 128; bn = 64; break;
    case TILE_64x128: bm = 64; bn = 128; break;
    case TILE_32x64: bm = 32; bn = 64; break;
    case TILE_32x32: bm = 32; bn = 32; break;
    default: break;
  }
  return TileShape{bm, bn, pick_bk_host(bm, bn, ct, akf, bkf)};
}

static int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Workgroup tile for an M x N x K contraction of compute type ct: enough
// tiles to cover the device's CUs, the largest register tile that does so.
int choose_tile(int64_t M, int64_t N, int ct) {
  const bool cplx = is_complex(ct);
  const int64_t cus = device_facts().cus;
  if (M <= 32) return TILE_32x64;
  const int64_t t128 = cdiv(M, 128) * cdiv(N, 128);
  if (!cplx && t128 >= cus) return TILE_128x128;
  const int64_t t12864 = cdiv(M, 128) * cdiv(N, 64);
  if (t12864 >= cus) return TILE_128x64;
  return TILE_64x64;
}

// Fixed-order split-K when the tile grid alone leaves CUs idle: up to two
// workgroups per CU in all.
static int choose_splits(int64_t tiles, int64_t K, int bk, int64_t &k_chunk) {
  const int64_t cus = device_facts().cus;
  int s = 1;
  if (tiles < cus && K >= 512) {
    s = (int)std::min<int64_t>(cdiv(2 * cus, tiles), K / 256);
    s = std::max(s, 1);
    s = std::min(s, 16);
  }
  k_chunk = cdiv(cdiv(K, s), 2 * bk) * 2 * bk;   // whole stage pairs (the gather's loop step)
  s = (int)cdiv(K, k_chunk);
  return s;
}

// ------------------------------------------------------------------- plan

static int bytes_of_compute_real(int ct) { return is_double(ct) ? 8 : 4; }

static double scalar_at(const void *p, int ct, int part) {
  if (ct == QS_F32 || ct == QS_C32) return (double)((const float *)p)[part];
  return ((const double *)p)[part];
}

// Offset of entry 1 minus entry 0 of map p's table: the stride of the
// group's first dim longer than one (0 when there is none).
static int64_t first_step(const qs_group &g, int p) {
  for (int d = 0; d < g.ndim; ++d)
    if (g.lengths[d] > 1) return g.strides[p][d];
  return 0;
}

// Whether entries 2i and 2i + 1 of map p's table are adjacent in memory for
// every i (n even): walks the enumeration without storing it.
static bool pairs_adjacent(const qs_group &g, int p, int64_t n) {
  if (n % 2 != 0) return false;
  if (n == 0) return true;
  int64_t x[QS_MAX_DIMS] = {0};
  int64_t o = 0, prev = 0;
  for (int64_t t = 0;;) {
    if (t % 2 == 1 && o != prev + 1) return false;
    prev = o;
    if (++t == n) break;
    for (int d = 0; d < g.ndim; ++d) {
      const int64_t L = g.lengths[d];
      if (x[d] < L - 1) {
        ++x[d];
        o += g.strides[p][d];
        break;
      }
      x[d] = 0;
      o -= (L - 1) * g.strides[p][d];
    }
  }
  return true;
}

// The step of map p if it is a single ramp, offset(q) == q * step for every
// coordinate q (map_ramp_step, src/layout/axis_group.jl:149-168; an empty
// group is a ramp of step 0).
static bool map_ramp_step(const qs_group &g, int p, int64_t &step) {
  step = 0;
  __int128 run = 1;
  bool started = false;
  for (int d = 0; d < g.ndim; ++d) {
    const int64_t L = g.lengths[d];
    if (L == 0) { step = 0; return true; }
    if (L == 1) continue;
    const int64_t S = g.strides[p][d];
    if (!started) {
      step = S;
      run = L;
      started = true;
    } else {
      if (run * (__int128)step != (__int128)S) return false;
      run *= L;
    }
  }
  return true;
}

qs_status plan_setup(const qs_contract_desc &d, qs_plan *p) {
  for (int t : {d.a_type, d.b_type, d.c_type, d.compute_type})
    if (t < QS_F32 || t > QS_C64)
      return fail(QS_ERR_ARGUMENT, "plan_contract: eltypes must be Float32, Float64, ComplexF32 or ComplexF64");
  if (!is_complex(d.c_type) && (is_complex(d.a_type) || is_complex(d.b_type)))
    return fail(QS_ERR_ARGUMENT, "plan_contract: a complex operand needs a complex C");
  if (is_complex(d.compute_type) != (is_complex(d.a_type) || is_complex(d.b_type) || is_complex(d.c_type)))
    return fail(QS_ERR_ARGUMENT, "plan_contract: the compute type's domain must be the promoted domain of A, B, C");
  if (d.memory != QS_MEM_DEVICE && d.memory != QS_MEM_HOST)
    return fail(QS_ERR_ARGUMENT, "plan_contract: memory must be QS_MEM_DEVICE or QS_MEM_HOST");
  if (d.precision != QS_PLAIN && d.precision != QS_SPLIT_INT8)
    return fail(QS_ERR_ARGUMENT, "plan_contract: unknown precision mode");
  qs_status st;
  if ((st = validate_group(d.m, "M")) != QS_OK) return st;
  if ((st = validate_group(d.n, "N")) != QS_OK) return st;
  if ((st = validate_group(d.k, "K")) != QS_OK) return st;
  if (d.a_len < 0 || d.b_len < 0 || d.c_len < 0)
    return fail(QS_ERR_ARGUMENT, "plan_contract: negative storage length");

  p->desc = d;
  p->M = group_length(d.m);
  p->N = group_length(d.n);
  p->K = group_length(d.k);
  p->mg = d.m;
  p->ng = d.n;
  p->kg = d.k;

  // Spans for the exact bounds checks (tiles.jl:120-134), in the caller's
  // group order (the ranges do not depend on the order).
  map_range(d.m, 0, p->mA_lo, p->mA_hi);
  map_range(d.m, 1, p->mC_lo, p->mC_hi);
  map_range(d.n, 0, p->nB_lo, p->nB_hi);
  map_range(d.n, 1, p->nC_lo, p->nC_hi);
  map_range(d.k, 0, p->kA_lo, p->kA_hi);
  map_range(d.k, 1, p->kB_lo, p->kB_hi);

  const int ct = d.compute_type;
  p->mixed = !(d.a_type == ct && d.b_type == ct && d.c_type == ct);

  // Path choice (select_path, plan.jl:395-412, with GPU paths).
  const int64_t M = p->M, N = p->N, K = p->K;
  if (M == 0 || N == 0) p->path = QS_PATH_EMPTY;
  else if (K == 0) p->path = QS_PATH_SCALE;
  else if (d.path_mode == 1) p->path = QS_PATH_GEMM;
  else if ((M == 1 || N == 1)) p->path = QS_PATH_DOT;
  else if (K == 1) p->path = QS_PATH_OUTER;
  else if (M <= 32) p->path = QS_PATH_SKINNY;
  else p->path = QS_PATH_GEMM;
  const bool split = d.precision == QS_SPLIT_INT8 && (p->path == QS_PATH_GEMM || p->path == QS_PATH_SKINNY);
  if (split && is_complex(ct))
    return fail(QS_ERR_UNSUPPORTED, "plan_contract: split precision is implemented for real compute types only");

  // Coalescing: put a unit-stride dim first in each operand's gather
  // direction.  Reordering M or N dims changes no arithmetic; reordering K
  // changes only the summation order.
  qs_group &mg = p->mg, &ng = p->ng, &kg = p->kg;
  if (p->path == QS_PATH_OUTER) {
    const int cm = unit_dim(mg, 1), cn = unit_dim(ng, 1);
    if (cm >= 0) move_to_front(mg, cm);
    else if (cn >= 0) move_to_front(ng, cn);
    p->a_kfast = (cm < 0 && cn >= 0) ? 1 : 0;   // "N fastest" for the outer kernel
    p->b_kfast = 0;
  } else {
    const int am = unit_dim(mg, 0), ak = unit_dim(kg, 0);
    const int bn = unit_dim(ng, 0), bk = unit_dim(kg, 1);
    bool akf = am < 0 && ak >= 0, bkf = bn < 0 && bk >= 0;
    if (p->path == QS_PATH_DOT) {   // only the matrix operand's choice matters
      if (N == 1) { bkf = false; } else { akf = false; }
    }
    if (akf && bkf && ak != bk) {
      // Both operands want a different K dim first: the larger one wins.
      if (M >= N) bkf = false; else akf = false;
    }
    if (am >= 0) move_to_front(mg, am);
    if (bn >= 0) move_to_front(ng, bn);
    if (akf) move_to_front(kg, ak);
    else if (bkf) move_to_front(kg, bk);
    p->a_kfast = akf;
    p->b_kfast = bkf;
  }
  p->c_mfast = (M > 1 && first_step(mg, 1) == 1) || !(N > 1 && first_step(ng, 1) == 1);
  // Paired loads read entries 2i and 2i + 1 of the table each gather is
  // laid along; both must be adjacent in memory for every i.
  const bool va = p->a_kfast ? pairs_adjacent(kg, 0, K) : pairs_adjacent(mg, 0, M);
  const bool vb = p->b_kfast ? pairs_adjacent(kg, 1, K) : pairs_adjacent(ng, 0, N);
  p->vec = (va && vb) ? 1 : 0;
  // Single-ramp maps, in the device enumeration order.
  {
    const qs_group *grp[6] = {&mg, &kg, &ng, &kg, &mg, &ng};
    const int map[6] = {0, 0, 0, 1, 1, 1};
    const int bit[6] = {RAMP_MA, RAMP_KA, RAMP_NB, RAMP_KB, RAMP_MC, RAMP_NC};
    p->ramps = 0;
    for (int i = 0; i < 6; ++i)
      if (map_ramp_step(*grp[i], map[i], p->steps[i])) p->ramps |= bit[i];
      else p->steps[i] = 0;
  }
  p->tile = TILE_64x64;
  p->k_splits = 1;
  p->k_chunk = K;
  if (split) p->path = QS_PATH_SPLIT;
  return QS_OK;
}

// The value of the K tables' pad entries: an offset that puts any row of the
// operand 2^31 bytes past the start of its span, which the buffer-load
// gather reads as zero.
int64_t k_pad_offset(const qs_plan *p, int side) {
  return side == 0 ? p->kA_lo + (int64_t(1) << 31) / elem_bytes(p->desc.a_type)
                   : p->kB_lo + (int64_t(1) << 31) / elem_bytes(p->desc.b_type);
}

qs_status plan_build(const qs_contract_desc *desc, const qs_allocator *al, qs_plan **out) {
  if (!desc || !out) return fail(QS_ERR_ARGUMENT, "qs_plan_create: null argument");
  qs_plan *p = new qs_plan();
  if (al) {
    p->al = *al;
    p->has_al = 1;
  }
  qs_status st = plan_setup(*desc, p);
  if (st != QS_OK) {
    delete p;
    return st;
  }
  const qs_contract_desc &d = p->desc;
  const int64_t M = p->M, N = p->N, K = p->K;
  const int ct = d.compute_type;

  // Offset tables, built once and kept on the device.  The K tables carry
  // kKTablePad pad entries past K (k_pad_offset): the buffer-load gather
  // reads them as zero, with no bounds test (its spans stay below 2^31 bytes).
  const int64_t KP = K + kKTablePad;
  const int64_t nt = 2 * (M + N + KP);
  std::vector<int64_t> h((size_t)std::max<int64_t>(nt, 1));
  int64_t *mA = h.data(), *mC = mA + M, *nB = mC + M, *nC = nB + N, *kA = nC + N, *kB = kA + KP;
  fill_offsets(p->mg, 0, M, mA, mC);
  fill_offsets(p->ng, 0, N, nB, nC);
  fill_offsets(p->kg, 0, K, kA, kB);
  for (int64_t k = K; k < KP; ++k) {
    kA[k] = k_pad_offset(p, 0);
    kB[k] = k_pad_offset(p, 1);
  }
  hipError_t e = hipSuccess;
  if (nt > 0) {
    e = dev_malloc(&p->d_tables, sizeof(int64_t) * (size_t)nt);
    if (e == hipSuccess) e = hipMemcpy(p->d_tables, h.data(), sizeof(int64_t) * (size_t)nt, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      qs_status r = e == hipErrorOutOfMemory ? fail(QS_ERR_MEMORY, "plan_contract: device tables") : hip_fail(e, "plan_contract: tables");
      qs_plan_destroy(p);
      return r;
    }
  }
  set_table_pointers(p, p->d_tables);

  // Tile, splits, workspace: an explicit override, else the autotune cache's
  // choice for this shape class, else the heuristic.
  if (p->path == QS_PATH_SPLIT) {
    p->moduli = split_moduli_for(d.target_rel_err, ct, K);
    qs_status ss = split_prepare(p);
    if (ss != QS_OK) { qs_plan_destroy(p); return ss; }
  } else if (p->path == QS_PATH_GEMM || p->path == QS_PATH_SKINNY) {
    int tile = (d.tile >= 0 && d.tile < TILE_COUNT) ? d.tile : -1;
    int splits = d.k_splits > 0 ? d.k_splits : 0;
    if (tile < 0 && splits == 0) {
      int t2, s2;
      if (tune_lookup(p, t2, s2)) { tile = t2; splits = s2; }
    }
    if (tile < 0) tile = choose_tile(M, N, ct);
    qs_status cs = configure_gemm(p, tile, splits);
    if (cs != QS_OK) { qs_plan_destroy(p); return cs; }
  }
  *out = p;
  return QS_OK;
}

// The six tables' positions inside one block: mA mC nB nC kA kB.
void set_table_pointers(qs_plan *p, const int64_t *t) {
  const int64_t KP = p->K + kKTablePad;
  p->t_mA = t;
  p->t_mC = p->t_mA + p->M;
  p->t_nB = p->t_mC + p->M;
  p->t_nC = p->t_nB + p->N;
  p->t_kA = p->t_nC + p->N;
  p->t_kB = p->t_kA + KP;
}

void *ws_alloc(qs_plan *p, size_t bytes) {
  if (p->has_al) {
    count_bridge(true);
    void *q = p->al.alloc(p->al.ctx, bytes);
    if (q && (uintptr_t)q % 256 != 0) {   // misaligned: handed back unused
      p->al.free(p->al.ctx, q, bytes);
      q = nullptr;
    }
    return q;
  }
  void *q = nullptr;
  if (dev_malloc(&q, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return q;
}
void ws_free(qs_plan *p, void *ptr, size_t bytes) {
  if (!ptr) return;
  if (p->has_al) {
    count_bridge(false);
    p->al.free(p->al.ctx, ptr, bytes);
    return;
  }
  (void)dev_free(ptr);
}

qs_status configure_gemm(qs_plan *p, int tile, int splits) {
  const int ct = p->desc.compute_type;
  const TileShape ts = tile_shape(tile, ct, p->a_kfast, p->b_kfast);
  const int64_t tiles = cdiv(p->M, ts.bm) * cdiv(p->N, ts.bn);
  if (tiles > INT32_MAX) return fail(QS_ERR_UNSUPPORTED, "plan_contract: too many tiles");
  int ks;
  int64_t chunk;
  if (splits <= 0) {
    ks = choose_splits(tiles, p->K, ts.bk, chunk);
  } else {
    chunk = cdiv(cdiv(p->K, splits), 2 * ts.bk) * 2 * ts.bk;
    ks = (int)cdiv(p->K, chunk);
  }
  // Partial sums of every split.
  const size_t need = ks > 1 ? (size_t)ks * (is_complex(ct) ? 2 : 1) * (size_t)p->M * (size_t)p->N *
                                   bytes_of_compute_real(ct)
                             : 0;
  if (need > (size_t)p->partials_bytes) {
    if (p->d_partials) ws_free(p, p->d_partials, (size_t)p->partials_bytes);
    p->d_partials = nullptr;
    p->partials_bytes = 0;
    p->d_partials = ws_alloc(p, need);
    if (!p->d_partials) {
      ks = 1;   // no room for partial sums: run unsplit
      chunk = p->K;
    } else {
      p->partials_bytes = (int64_t)need;
    }
  }
  p->tile = tile;
  p->k_splits = ks;
  p->k_chunk = chunk;
  return QS_OK;
}

static bool span_ok(int64_t base, int64_t rlo, int64_t rhi, int64_t clo, int64_t chi, int64_t len, int64_t &lo_out,
                    int64_t &hi_out) {
  const __int128 lo = (__int128)base + rlo + clo, hi = (__int128)base + rhi + chi;
  lo_out = (int64_t)lo;
  hi_out = (int64_t)hi;
  return lo >= 0 && hi <= (__int128)len - 1;
}

qs_status check_spans(qs_plan *p, bool reads_ab) {
  const qs_contract_desc &d = p->desc;
  if (p->M == 0 || p->N == 0) return QS_OK;
  if (!span_ok(d.c_base, p->mC_lo, p->mC_hi, p->nC_lo, p->nC_hi, d.c_len, p->c_lo, p->c_hi))
    return fail(QS_ERR_BOUNDS, "execute!: C tile addresses [" + std::to_string(p->c_lo) + ", " + std::to_string(p->c_hi) +
                                   "] exceed storage bounds [0, " + std::to_string(d.c_len - 1) + "]");
  if (reads_ab && p->K > 0) {
    if (!span_ok(d.a_base, p->mA_lo, p->mA_hi, p->kA_lo, p->kA_hi, d.a_len, p->a_lo, p->a_hi))
      return fail(QS_ERR_BOUNDS, "execute!: A tile addresses [" + std::to_string(p->a_lo) + ", " +
                                     std::to_string(p->a_hi) + "] exceed storage bounds [0, " +
                                     std::to_string(d.a_len - 1) + "]");
    if (!span_ok(d.b_base, p->kB_lo, p->kB_hi, p->nB_lo, p->nB_hi, d.b_len, p->b_lo, p->b_hi))
      return fail(QS_ERR_BOUNDS, "execute!: B tile addresses [" + std::to_string(p->b_lo) + ", " +
                                     std::to_string(p->b_hi) + "] exceed storage bounds [0, " +
                                     std::to_string(d.b_len - 1) + "]");
  }
  return QS_OK;
}

// Complex A times real B (1) or real A times complex B (2) on the generic
// gather; 0 otherwise.
static int mixed_domain(const qs_contract_desc &d, bool generic) {
  if (!generic || !is_complex(d.compute_type)) return 0;
  const bool ac = is_complex(d.a_type), bc = is_complex(d.b_type);
  return (ac && !bc) ? 1 : (!ac && bc) ? 2 : 0;
}

GemmArgs make_args(const qs_plan *p, const void *A, int64_t a_base, const void *B, int64_t b_base, void *C,
                   int64_t c_base) {
  const qs_contract_desc &d = p->desc;
  GemmArgs g;
  g.A = A;
  g.B = B;
  g.C = C;
  g.a_base = a_base;
  g.b_base = b_base;
  g.c_base = c_base;
  g.mA = p->t_mA;
  g.kA = p->t_kA;
  g.nB = p->t_nB;
  g.kB = p->t_kB;
  g.mC = p->t_mC;
  g.nC = p->t_nC;
  g.M = p->M;
  g.N = p->N;
  g.K = p->K;
  g.a_type = d.a_type;
  g.b_type = d.b_type;
  g.c_type = d.c_type;
  g.compute_type = d.compute_type;
  g.conj_a = is_complex(d.compute_type) ? d.conj_a : 0;
  g.conj_b = is_complex(d.compute_type) ? d.conj_b : 0;
  g.alpha_re = 1;
  g.alpha_im = 0;
  g.beta_re = 0;
  g.beta_im = 0;
  g.beta_case = 0;
  g.k_splits = p->k_splits;
  g.k_chunk = p->k_chunk;
  g.partials = p->d_partials;
  g.tile = p->tile;
  g.a_kfast = p->a_kfast;
  g.b_kfast = p->b_kfast;
  // Spans of A and B as the tables address them; one storage type and spans
  // within 32-bit byte offsets take the buffer-load gather.
  g.a_lo = a_base + p->mA_lo + p->kA_lo;
  g.b_lo = b_base + p->kB_lo + p->nB_lo;
  const int64_t a_bytes = (p->mA_hi - p->mA_lo + p->kA_hi - p->kA_lo + 1) * elem_bytes(d.a_type);
  const int64_t b_bytes = (p->kB_hi - p->kB_lo + p->nB_hi - p->nB_lo + 1) * elem_bytes(d.b_type);
  const bool spans_ok = a_bytes <= kBufferSpanMax && b_bytes <= kBufferSpanMax;
  const bool buf = !p->mixed && spans_ok;
  g.mixed = buf ? 0 : 1;
  g.vec = buf && !is_complex(d.compute_type) ? p->vec : 0;
  g.dom = mixed_domain(d, !buf);
  if (g.dom) {
    // The complex operand and C in the compute type, the real one in its real type.
    const int ct = d.compute_type, rt = ct == QS_C64 ? QS_F64 : QS_F32;
    const int at = g.dom == 1 ? ct : rt, bt = g.dom == 1 ? rt : ct;
    g.dom_buf = spans_ok && d.a_type == at && d.b_type == bt && d.c_type == ct ? 1 : 0;
  }
  g.a_bytes = buf || g.dom_buf ? (uint32_t)a_bytes : 0;
  g.b_bytes = buf || g.dom_buf ? (uint32_t)b_bytes : 0;
  g.batch_off = nullptr;
  g.batch = 1;
  g.xcds = device_facts().xcds;
  g.ramps = p->ramps;
  g.step_mA = p->steps[0];
  g.step_kA = p->steps[1];
  g.step_nB = p->steps[2];
  g.step_kB = p->steps[3];
  g.step_mC = p->steps[4];
  g.step_nC = p->steps[5];
  g.kpad_a = k_pad_offset(p, 0);
  g.kpad_b = k_pad_offset(p, 1);
  return g;
}

// Run the plan's path on device storages (bases already shifted for staging).
qs_status run_device(qs_plan *p, const void *A, int64_t a_base, const void *B, int64_t b_base, void *C, int64_t c_base,
                     const double *al, const double *be, hipStream_t s) {
  GemmArgs g = make_args(p, A, a_base, B, b_base, C, c_base);
  g.alpha_re = al[0];
  g.alpha_im = al[1];
  g.beta_re = be[0];
  g.beta_im = be[1];
  g.beta_case = (be[0] == 0 && be[1] == 0) ? 0 : (be[0] == 1 && be[1] == 0) ? 1 : 2;
  const bool alpha_zero = al[0] == 0 && al[1] == 0;
  hipError_t e = hipSuccess;
  if (p->path == QS_PATH_EMPTY) return QS_OK;
  if (alpha_zero || p->path == QS_PATH_SCALE) {
    if (g.beta_case == 1) return QS_OK;   // scale_tile!: beta == 1 is a no-op
    e = launch_scale(g, s);
  } else if (p->path == QS_PATH_DOT) {
    e = launch_dot(g, s);
  } else if (p->path == QS_PATH_OUTER) {
    e = launch_outer(g, s);
  } else if (p->path == QS_PATH_SPLIT) {
    if (p->split_report_on == 2) return run_split_guarded(p, g, s);
    e = launch_split(p, g, p->split_report_on != 0, s);
  } else {
    e = launch_gemm(g, s);
  }
  if (e != hipSuccess) return hip_fail(e, "execute!: launch");
  return QS_OK;
}

}  // namespace qs

using namespace qs;

extern "C" {

const char *qs_last_error(void) { return g_last_error; }

const char *qs_version(void) { return "qs_amd 0.1.0 (gfx950)"; }

qs_status qs_target_profile_get(qs_target_profile *out) {
  if (!out) return fail(QS_ERR_ARGUMENT, "qs_target_profile_get: null argument");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  hipDeviceProp_t pr;
  e = hipGetDeviceProperties(&pr, dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  std::memset(out, 0, sizeof *out);
  std::snprintf(out->name, sizeof out->name, "%s", pr.gcnArchName);
  out->compute_units = pr.multiProcessorCount;
  out->wavefront_size = pr.warpSize;
  out->lds_bytes_per_cu = (int32_t)pr.maxSharedMemoryPerMultiProcessor;
  out->lds_bytes_per_block = (int32_t)pr.sharedMemPerBlock;
  out->clock_khz = pr.clockRate;
  out->memory_clock_khz = pr.memoryClockRate;
  out->memory_bus_bits = pr.memoryBusWidth;
  out->l2_bytes = pr.l2CacheSize;
  out->hbm_bytes = (int64_t)pr.totalGlobalMem;
  out->xcds = device_facts().xcds;
  out->mfma_f64_m = 16; out->mfma_f64_n = 16; out->mfma_f64_k = 4;
  out->mfma_f32_m = 16; out->mfma_f32_n = 16; out->mfma_f32_k = 4;
  out->mfma_i8_m = 16; out->mfma_i8_n = 16; out->mfma_i8_k = 64;
  out->peak_fp64_tflops = 78.6;
  out->peak_fp32_tflops = 157.3;
  out->peak_int8_tops = 5000.0;
  out->peak_hbm_gbs = 8000.0;
  return QS_OK;
}

qs_status qs_plan_create(const qs_contract_desc *desc, qs_plan **out) {
  if (out) *out = nullptr;
  return plan_build(desc, nullptr, out);
}

qs_status qs_plan_create_with_allocator(const qs_contract_desc *desc, const qs_allocator *al, qs_plan **out) {
  if (out) *out = nullptr;
  if (al && (!al->alloc || !al->free)) return fail(QS_ERR_ARGUMENT, "plan_contract: allocator without alloc / free");
  return plan_build(desc, al, out);
}


qs_status qs_plan_release(qs_plan *p) {
  if (!p) return fail(QS_ERR_ARGUMENT, "release!: null plan");
  if (p->released) return QS_OK;
  // Work already queued on the workspace finishes before it is handed back.
  if (hipDeviceSynchronize() != hipSuccess) return fail(QS_ERR_DEVICE, "release!: device sync");
  ws_free(p, p->d_cbackup, (size_t)p->cbackup_bytes);
  ws_free(p, p->d_split, (size_t)p->split_bytes);
  ws_free(p, p->d_partials, (size_t)p->partials_bytes);
  p->d_cbackup = nullptr;
  p->d_split = nullptr;
  p->d_partials = nullptr;
  p->released = 1;
  return QS_OK;
}

qs_status qs_plan_get_info(const qs_plan *p, qs_plan_info *out) {
  if (!p || !out) return fail(QS_ERR_ARGUMENT, "qs_plan_get_info: null argument");
  std::memset(out, 0, sizeof *out);
  out->path = p->path;
  const TileShape ts = tile_shape(p->tile, p->desc.compute_type, p->a_kfast, p->b_kfast);
  out->tile_m = ts.bm;
  out->tile_n = ts.bn;
  out->tile_k = ts.bk;
  out->k_splits = p->k_splits;
  out->a_kfast = p->a_kfast;
  out->b_kfast = p->b_kfast;
  out->moduli = p->moduli;
  out->m = p->M;
  out->n = p->N;
  out->k = p->K;
  out->workspace_bytes = p->released ? 0 : p->partials_bytes + p->split_bytes;
  out->paired_loads =
      p->path == QS_PATH_GEMM && !p->mixed && !is_complex(p->desc.compute_type) && ts.bm * ts.bn <= 64 * 128 ? p->vec : 0;
  out->mixed_domain = p->path == QS_PATH_GEMM || p->path == QS_PATH_SKINNY ? mixed_domain(p->desc, p->mixed != 0) : 0;
  return QS_OK;
}

qs_status qs_execute(qs_plan *p, const void *alpha, const void *beta, void *stream) {
  if (!p || !alpha || !beta) return fail(QS_ERR_ARGUMENT, "execute!: null argument");
  if (p->released) return fail(QS_ERR_ARGUMENT, "execute!: the plan's workspace was released");
  const int ct = p->desc.compute_type;
  double al[2] = {scalar_at(alpha, ct, 0), is_complex(ct) ? scalar_at(alpha, ct, 1) : 0.0};
  double be[2] = {scalar_at(beta, ct, 0), is_complex(ct) ? scalar_at(beta, ct, 1) : 0.0};
  const bool reads_ab = !(al[0] == 0 && al[1] == 0);
  qs_status st = check_spans(p, reads_ab);
  if (st != QS_OK) return st;
  if (p->path == QS_PATH_EMPTY) return QS_OK;
  hipStream_t s = (hipStream_t)stream;
  const qs_contract_desc &d = p->desc;
  if (d.memory == QS_MEM_DEVICE)
    return run_device(p, d.a, d.a_base, d.b, d.b_base, d.c, d.c_base, al, be, s);
  return run_staged(p, al, be, reads_ab, s);
}

qs_status qs_plan_rebind(qs_plan *p, const void *a, int64_t a_base, int64_t a_len, const void *b, int64_t b_base,
                         int64_t b_len, void *c, int64_t c_base, int64_t c_len) {
  if (!p) return fail(QS_ERR_ARGUMENT, "qs_plan_rebind: null plan");
  if (p->desc.memory != QS_MEM_DEVICE) return fail(QS_ERR_ARGUMENT, "qs_plan_rebind: device plans only");
  if (a_len < 0 || b_len < 0 || c_len < 0) return fail(QS_ERR_ARGUMENT, "qs_plan_rebind: negative storage length");
  p->desc.a = a; p->desc.a_base = a_base; p->desc.a_len = a_len;
  p->desc.b = b; p->desc.b_base = b_base; p->desc.b_len = b_len;
  p->desc.c = c; p->desc.c_base = c_base; p->desc.c_len = c_len;
  return QS_OK;
}

qs_status qs_plan_destroy(qs_plan *p) {
  if (!p) return QS_OK;
  if (!p->released && (p->d_split || p->d_partials || p->d_cbackup)) {
    // Queued work may still read the workspace; an allocator may hand it on at once.
    if (p->has_al) (void)hipDeviceSynchronize();
    ws_free(p, p->d_cbackup, (size_t)p->cbackup_bytes);
    ws_free(p, p->d_split, (size_t)p->split_bytes);
    ws_free(p, p->d_partials, (size_t)p->partials_bytes);
  }
  if (p->d_tables) (void)dev_free(p->d_tables);
  if (p->split_side) (void)hipStreamDestroy(p->split_side);
  if (p->split_fork) (void)hipEventDestroy(p->split_fork);
  if (p->split_join) (void)hipEventDestroy(p->split_join);
  for (int i = 0; i < 3; ++i)
    if (p->stage[i]) (void)dev_free(p->stage[i]);
  delete p;
  return QS_OK;
}

qs_status qs_plan_set_error_report(qs_plan *p, int32_t on) {
  if (!p) return fail(QS_ERR_ARGUMENT, "qs_plan_set_error_report: null plan");
  if (on < 0 || on > 2) return fail(QS_ERR_ARGUMENT, "qs_plan_set_error_report: mode must be 0, 1 or 2");
  p->split_report_on = on;
  return QS_OK;
}

qs_status qs_plan_get_error_report(const qs_plan *p, qs_error_report *out) {
  if (!p || !out) return fail(QS_ERR_ARGUMENT, "qs_plan_get_error_report: null argument");
  std::memset(out, 0, sizeof *out);
  out->moduli = p->moduli;
  out->samples = p->split_report.samples;
  out->rel_fro = p->split_report.rel_fro;
  out->max_rel = p->split_report.max_rel;
  out->exceeded = p->split_report.exceeded;
  out->escalations = p->split_report.escalations;
  out->fallback = p->split_report.fallback;
  out->target = p->desc.target_rel_err;
  return QS_OK;
}

qs_status qs_alloc_stats_get(qs_alloc_stats *out) {
  if (!out) return fail(QS_ERR_ARGUMENT, "qs_alloc_stats_get: null argument");
  out->device_allocs = g_dev_allocs.load();
  out->device_frees = g_dev_frees.load();
  out->device_bytes = g_dev_bytes.load();
  out->pinned_allocs = g_pin_allocs.load();
  out->pinned_frees = g_pin_frees.load();
  out->bridge_allocs = g_bridge_allocs.load();
  out->bridge_frees = g_bridge_frees.load();
  return QS_OK;
}

qs_status qs_fault_inject(int32_t kind, int64_t after) {
  if (kind != QS_FAULT_DEVICE_ALLOC && kind != QS_FAULT_PINNED_ALLOC)
    return fail_fmt(QS_ERR_ARGUMENT, "qs_fault_inject: unknown kind %d", (int)kind);
  g_fault_left[kind].store(after < 0 ? -1 : after);
  return QS_OK;
}

qs_status qs_fill_offsets(const qs_group *g, int64_t first, int64_t count, int64_t *out0, int64_t *out1) {
  if (!g) return fail(QS_ERR_ARGUMENT, "fill_offsets!: null group");
  qs_status st = validate_group(*g, "fill_offsets!");
  if (st != QS_OK) return st;
  const int64_t Q = group_length(*g);
  if (count < 0) return fail(QS_ERR_ARGUMENT, "count must be nonnegative, got " + std::to_string(count));
  if (!(0 <= first && first <= Q))
    return fail(QS_ERR_BOUNDS, "AxisGroup interval start " + std::to_string(first) + " out of range [0, " +
                                   std::to_string(Q) + "]");
  if (count > Q - first)
    return fail(QS_ERR_BOUNDS, "AxisGroup interval (first=" + std::to_string(first) + ", count=" +
                                   std::to_string(count) + ") exceeds domain size " + std::to_string(Q));
  if (count > 0 && (!out0 || !out1)) return fail(QS_ERR_DIMENSION, "fill_offsets!: null buffer");
  fill_offsets(*g, first, count, out0, out1);
  return QS_OK;
}

}  // extern "C"
