// Grouped persistent execution: many differently shaped and strided
// contractions in one launch.
//
// The tensor-network sweeps this engine serves (MPS/DMRG environments, bond
// dimensions from 16 to 512) issue many small pairwise contractions; one
// launch per contraction leaves the GPU idle between them.  A group is built
// from existing plans: their GEMM descriptors (offset tables, storages, tile
// class) go into one device table; a persistent grid walks the union of their
// tiles, each workgroup taking tiles t, t + grid, t + 2 grid, ... and finding
// the owning contraction by binary search over the tile prefix sums.  Every
// tile computes all of K in registers (no split-K), so a contraction's result
// does not depend on what else is in the group.  Contractions whose compute
// type or gather directions differ go to separate launches (one kernel
// instantiation each) on the same stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "qs_gemm_impl.h"
#include "qs_plan.h"

namespace qs {

struct GroupItem {
  GemmArgs g;           // alpha / beta are taken from the coefficient table
  int64_t tile_start;   // first global tile index of this contraction
  int64_t tiles_m;
  int32_t cls;          // TileId of its workgroup tile
  int32_t index;        // position in the caller's plan array
};

// The tile classes a grouped launch holds: 128x64, 64x64, 32x64 and 32x32
// (the bodies share one register allocation, and 128x128 alone needs ~230
// VGPRs).  Their largest LDS footprint, or the one body's (ONLY):
template <typename R, bool CPLX, bool AKF, bool BKF, int ONLY = -1, int DOM = 0>
struct GroupSmem {
  static constexpr int a = TileGemm<R, CPLX, 128, 64, AKF, BKF, false, false, DOM>::SMEM;
  static constexpr int c = TileGemm<R, CPLX, 64, 64, AKF, BKF, false, false, DOM>::SMEM;
  static constexpr int d = TileGemm<R, CPLX, 32, 64, AKF, BKF, false, false, DOM>::SMEM;
  static constexpr int e = TileGemm<R, CPLX, 32, 32, AKF, BKF, false, false, DOM>::SMEM;
  static constexpr int cd = c > d ? (c > e ? c : e) : (d > e ? d : e);
  static constexpr int value = ONLY == TILE_128x64 ? a : (a > cd ? a : cd);
};

// coef: per item {alpha_re, alpha_im, beta_re, beta_im}, indexed by item.index.
// cycles (the autotuner's measurement, null otherwise): per member, the
// workgroup time its tiles took, in ticks of the constant 100 MHz clock.
//
// Round r of the persistent loop covers tiles [r G, (r + 1) G) of the G
// workgroups; within a round each XCD takes a contiguous run of them
// (xcd_swizzle), so the tiles one XCD runs at once are neighbours of one
// contraction, which share its A rows or B columns in that XCD's L2.  Odd
// rounds run their slots in reverse: the items come longest K first, so the
// workgroup that took a round's longest tile takes the next round's
// shortest, which evens out the workgroups' totals.
//
// ONLY == TILE_128x64: a launch whose items all run 128 x 64 tiles (the
// padding heuristic's pick for every member of the 512-member MPS sweeps).
// Float64 sweep 380-385k -> 401-402k contractions per second, two runs each
// alternating on one box (the one body keeps two stages in flight);
// ComplexF64, one stage either way, 152.1-152.3k -> 152.5-153.5k.
//
// DOM (complex compute types): the mixed-domain tiles, 1 = complex A times
// real B, 2 = real A times complex B, for members whose storage types are
// native (GemmArgs::dom_buf); 0 otherwise.
template <typename R, bool CPLX, bool AKF, bool BKF, int ONLY = -1, int DOM = 0>
__global__ __launch_bounds__(256, 2) void grouped_kernel(const GroupItem *__restrict__ items,
                                                          const int64_t *__restrict__ starts, int n_items,
                                                          int64_t total_tiles, const double *__restrict__ coef,
                                                          unsigned long long *cycles, int xcds) {
  __shared__ R smem[GroupSmem<R, CPLX, AKF, BKF, ONLY, DOM>::value];
  const int64_t G = gridDim.x, slot = xcd_swizzle(blockIdx.x, G, xcds);
  bool odd = false;
  for (int64_t r0 = 0; r0 < total_tiles; r0 += G, odd = !odd) {
    const int64_t t = r0 + (odd ? G - 1 - slot : slot);
    if (t >= total_tiles) break;   // only the last round is partial
    uint64_t t0 = 0;
    if (cycles) {
      __syncthreads();
      t0 = wall_clock64();
    }
    // Last item with tile_start <= t, searched in the items' dense copy of
    // tile_start (the probes of a search over the items themselves each
    // touch another cache line).
    int lo = 0, hi = n_items - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (starts[mid] <= t) lo = mid; else hi = mid - 1;
    }
    lo = __builtin_amdgcn_readfirstlane(lo);
    GemmArgs g = items[lo].g;
    const int64_t local = t - items[lo].tile_start;
    const int64_t tm = local % items[lo].tiles_m, tn = local / items[lo].tiles_m;
    const double *c = coef + 4 * items[lo].index;
    g.alpha_re = c[0];
    g.alpha_im = c[1];
    g.beta_re = c[2];
    g.beta_im = c[3];
    g.beta_case = (c[2] == 0 && c[3] == 0) ? 0 : (c[2] == 1 && c[3] == 0) ? 1 : 2;
    if (c[0] == 0 && c[1] == 0) g.K = 0;   // alpha == 0: C = beta C, A and B never read
    // One stage in flight: the three tile bodies share one register
    // allocation.  (Complex types with 64 x 64 and 32 x 64 tiles only, each
    // with two stages in flight, measured slower on the 512-member ComplexF64
    // sweep: 131-132k against 147k contractions per second.)
    if constexpr (ONLY == TILE_128x64) {
      // Every item of the launch on 128 x 64 tiles: one tile body, which
      // has the registers for two stages in flight where its type allows.
      TileGemm<R, CPLX, 128, 64, AKF, BKF, false, false, DOM>::template run<1>(g, tm, tn, 0, smem);
    } else {
    typedef TileGemm<R, CPLX, 128, 64, AKF, BKF, false, false, DOM> T128x64;
    typedef TileGemm<R, CPLX, 64, 64, AKF, BKF, false, false, DOM> T64x64;
    typedef TileGemm<R, CPLX, 32, 64, AKF, BKF, false, false, DOM> T32x64;
    typedef TileGemm<R, CPLX, 32, 32, AKF, BKF, false, false, DOM> T32x32;
    switch (items[lo].cls) {
      case TILE_128x64: T128x64::template run<0>(g, tm, tn, 0, smem); break;
      case TILE_32x64: T32x64::template run<0>(g, tm, tn, 0, smem); break;
      case TILE_32x32:
        // Float64 only: other types never carry this class (grouped_class).
        // A fourth tile body cost the complex kernels spilled registers and
        // the Float32 kernel with both operands K fast a stack array.
        if constexpr (!CPLX && sizeof(R) == 8)
          T32x32::template run<0>(g, tm, tn, 0, smem);
        break;
      default: T64x64::template run<0>(g, tm, tn, 0, smem); break;
    }
    }
    if (cycles) {
      __syncthreads();
      if (threadIdx.x == 0) atomicAdd(cycles + items[lo].index, (unsigned long long)(wall_clock64() - t0));
    }
  }
}

template <typename R, bool CPLX, bool AKF, bool BKF, int DOM>
static hipError_t launch_grouped_t(const GroupItem *items, const int64_t *starts, int n, int64_t total,
                                   const double *coef, int grid, hipStream_t s, unsigned long long *cycles,
                                   int only) {
  if (only == TILE_128x64) {
    hipLaunchKernelGGL((grouped_kernel<R, CPLX, AKF, BKF, TILE_128x64, DOM>), dim3(grid), dim3(256), 0, s, items,
                       starts, n, total, coef, cycles, device_facts().xcds);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((grouped_kernel<R, CPLX, AKF, BKF, -1, DOM>), dim3(grid), dim3(256), 0, s, items, starts, n,
                     total, coef, cycles, device_facts().xcds);
  return hipGetLastError();
}

template <typename R, bool CPLX, int DOM>
static hipError_t launch_grouped_d(int akf, int bkf, const GroupItem *items, const int64_t *starts, int n,
                                   int64_t total, const double *coef, int grid, hipStream_t s,
                                   unsigned long long *cy, int only) {
  switch ((akf ? 2 : 0) | (bkf ? 1 : 0)) {
    case 0: return launch_grouped_t<R, CPLX, false, false, DOM>(items, starts, n, total, coef, grid, s, cy, only);
    case 1: return launch_grouped_t<R, CPLX, false, true, DOM>(items, starts, n, total, coef, grid, s, cy, only);
    case 2: return launch_grouped_t<R, CPLX, true, false, DOM>(items, starts, n, total, coef, grid, s, cy, only);
    default: return launch_grouped_t<R, CPLX, true, true, DOM>(items, starts, n, total, coef, grid, s, cy, only);
  }
}

template <typename R, bool CPLX>
static hipError_t launch_grouped_rc(int akf, int bkf, int dom, const GroupItem *items, const int64_t *starts, int n,
                                    int64_t total, const double *coef, int grid, hipStream_t s,
                                    unsigned long long *cy, int only) {
  if constexpr (CPLX) {
    if (dom == 1) return launch_grouped_d<R, CPLX, 1>(akf, bkf, items, starts, n, total, coef, grid, s, cy, only);
    if (dom == 2) return launch_grouped_d<R, CPLX, 2>(akf, bkf, items, starts, n, total, coef, grid, s, cy, only);
  }
  return launch_grouped_d<R, CPLX, 0>(akf, bkf, items, starts, n, total, coef, grid, s, cy, only);
}

// items: n GroupItems, each launch's table followed by the n tile_start
// values (item_starts).
static const int64_t *item_starts(const GroupItem *items, int n) {
  return reinterpret_cast<const int64_t *>(items + n);
}

// only: the tile class every item of the launch runs, or -1.  dom: the
// items' GemmArgs::dom (0 for real compute types).
static hipError_t launch_grouped(int ct, int akf, int bkf, int dom, const GroupItem *items, const int64_t *starts,
                                 int n, int64_t total, const double *coef, int grid, hipStream_t s,
                                 unsigned long long *cy, int only) {
  switch (ct) {
    case QS_F64: return launch_grouped_rc<double, false>(akf, bkf, 0, items, starts, n, total, coef, grid, s, cy, only);
    case QS_F32: return launch_grouped_rc<float, false>(akf, bkf, 0, items, starts, n, total, coef, grid, s, cy, only);
    case QS_C64:
      return launch_grouped_rc<double, true>(akf, bkf, dom, items, starts, n, total, coef, grid, s, cy, only);
    default: return launch_grouped_rc<float, true>(akf, bkf, dom, items, starts, n, total, coef, grid, s, cy, only);
  }
}


// The grouped kernel's tile classes: 128 x 64, 64 x 64, 32 x 64, and 32 x 32
// for Float64 (other compute types take 32 x 64 in its place).
static int grouped_class(int cls, int ct) {
  if (cls == TILE_32x32) return ct == QS_F64 ? TILE_32x32 : TILE_32x64;
  return (cls == TILE_128x64 || cls == TILE_32x64) ? cls : TILE_64x64;
}

// Workgroup tile of one contraction inside a group: the grid is already
// full of other contractions' tiles, so take the tile that wastes the least
// padded work, weighted by how much slower the smaller register tiles run
// per padded element.  The weights are measured: on the 512-contraction MPS
// sweep (FP64 and complex FP64) they give the whole group the time of
// all-128x64 tiles, which a 64x64 weight of 1.2 and a 32x64 weight of 1.7
// did not (1-8% slower), and they keep the smaller tiles for small members.
int grouped_tile(int64_t M, int64_t N, int ct) {
  struct Cand { int id, bm, bn; double cost; };
  const Cand cands[] = {{TILE_128x64, 128, 64, 1.0}, {TILE_64x64, 64, 64, 1.35}, {TILE_32x64, 32, 64, 2.0}};
  int best = TILE_64x64;
  double best_cost = 1e300;
  for (const Cand &c : cands) {
    const double padded = (double)((M + c.bm - 1) / c.bm * c.bm) * (double)((N + c.bn - 1) / c.bn * c.bn);
    const double cost = padded * c.cost;
    if (cost < best_cost) { best_cost = cost; best = c.id; }
  }
  return best;
}

struct Launch {   // one kernel instantiation's share of the group
  int ct, akf, bkf, dom;
  int only = -1;    // the tile class of every item, or -1
  GroupItem *d_items = nullptr;
  int n = 0;
  int64_t total = 0;
  int grid = 0;
};

}  // namespace qs

struct qs_group_plan {
  std::vector<qs_plan *> plans;
  std::vector<int> tile_override;       // per plan, -1 = heuristic
  // Batched groups: element offsets added to each member's A / B / C bases
  // (empty for a group of separate plans).  With equal coefficients for all
  // members they run as one batched launch of the plan's own kernel
  // (batch_args, offsets in d_batch_off); otherwise as a grouped launch.
  std::vector<int64_t> a_off, b_off, c_off;
  int64_t *d_batch_off = nullptr;
  qs::GemmArgs batch_args;
  std::vector<qs::Launch> launches;
  std::vector<int> standalone;          // members on the generic gather, run one by one
  double *d_coef = nullptr;             // 4 doubles per plan
  double *h_coef = nullptr;             // pinned host staging
  std::vector<double> last_coef;
  int64_t flops = 0;
  // Each member's storages as the device tables hold them (snapshotted when
  // the tables were built): a member rebound since then has its tables
  // rebuilt before the next execute or capture.
  struct Storages {
    const void *a, *b;
    void *c;
    int64_t a_base, a_len, b_base, b_len, c_base, c_len;
    bool operator==(const Storages &o) const {
      return a == o.a && b == o.b && c == o.c && a_base == o.a_base && a_len == o.a_len && b_base == o.b_base &&
             b_len == o.b_len && c_base == o.c_base && c_len == o.c_len;
    }
  };
  std::vector<Storages> held;
  // The launch tables match `held` and every member is in them.  Cleared
  // while they are rebuilt, so a rebuild that failed (an allocation) is
  // retried, or its error returned, by the next execute, capture or autotune
  // instead of running with some members missing.
  bool tables_ok = false;
};

struct qs_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace qs {

// Two workgroups per CU: the persistent grid's width.
static int cu_count() { return device_facts().cus; }

static void free_launches(qs_group_plan *g) {
  for (auto &l : g->launches)
    if (l.d_items) (void)dev_free(l.d_items);
  g->launches.clear();
  g->standalone.clear();
}

// Member i's storage bases.
static int64_t base_a(const qs_group_plan *g, size_t i) {
  return g->plans[i]->desc.a_base + (g->a_off.empty() ? 0 : g->a_off[i]);
}
static int64_t base_b(const qs_group_plan *g, size_t i) {
  return g->plans[i]->desc.b_base + (g->b_off.empty() ? 0 : g->b_off[i]);
}
static int64_t base_c(const qs_group_plan *g, size_t i) {
  return g->plans[i]->desc.c_base + (g->c_off.empty() ? 0 : g->c_off[i]);
}

static qs_group_plan::Storages storages_of(const qs_plan *p) {
  const qs_contract_desc &d = p->desc;
  return qs_group_plan::Storages{d.a, d.b, d.c, d.a_base, d.a_len, d.b_base, d.b_len, d.c_base, d.c_len};
}

// The tile class a member takes: the caller's override, else the autotune
// cache's choice for its grouped shape class, else the padding heuristic.
static int member_class(const qs_group_plan *g, size_t i) {
  const qs_plan *p = g->plans[i];
  int cls = g->tile_override[i], splits = 1;
  if (cls < 0 && !tune_find(group_shape_key(p), cls, splits)) cls = -1;
  if (cls < 0) cls = grouped_tile(p->M, p->N, p->desc.compute_type);
  return grouped_class(cls, p->desc.compute_type);
}

// Device tables of grouped launches for the members with cls[i] >= 0 (one
// launch per compute type and gather directions).  scratch != null: every
// member writes C into `scratch` instead (the autotuner's runs).
static qs_status make_launches(qs_group_plan *g, const std::vector<int> &cls, void *scratch, std::vector<Launch> &out) {
  std::map<std::tuple<int, int, int, int>, std::vector<GroupItem>> by_launch;
  for (size_t i = 0; i < g->plans.size(); ++i) {
    if (cls[i] < 0) continue;
    qs_plan *p = g->plans[i];
    GemmArgs a = scratch ? make_args(p, p->desc.a, base_a(g, i), p->desc.b, base_b(g, i), scratch,
                                     -(p->mC_lo + p->nC_lo))
                         : make_args(p, p->desc.a, base_a(g, i), p->desc.b, base_b(g, i), p->desc.c, base_c(g, i));
    a.k_splits = 1;
    a.k_chunk = p->K;
    a.partials = nullptr;
    GroupItem it;
    std::memset(&it, 0, sizeof it);
    it.g = a;
    it.cls = grouped_class(cls[i], p->desc.compute_type);
    const TileShape ts = tile_shape(it.cls, p->desc.compute_type, p->a_kfast, p->b_kfast);
    it.tiles_m = (p->M + ts.bm - 1) / ts.bm;
    it.tile_start = it.tiles_m * ((p->N + ts.bn - 1) / ts.bn);   // tile count, prefix-summed below
    it.index = (int32_t)i;
    by_launch[std::make_tuple(p->desc.compute_type, p->a_kfast, p->b_kfast, a.dom)].push_back(it);
  }
  for (auto &kv : by_launch) {
    std::vector<GroupItem> &v = kv.second;
    // Longest K first: a tile's cost grows with K, so the persistent grid's
    // round-robin hands out the long tiles first.
    std::stable_sort(v.begin(), v.end(), [](const GroupItem &x, const GroupItem &y) { return x.g.K > y.g.K; });
    int64_t start = 0;
    for (auto &it : v) {
      const int64_t n = it.tile_start;
      it.tile_start = start;
      start += n;
    }
    Launch l;
    l.ct = std::get<0>(kv.first);
    l.akf = std::get<1>(kv.first);
    l.bkf = std::get<2>(kv.first);
    l.dom = std::get<3>(kv.first);
    l.n = (int)v.size();
    l.total = start;
    l.only = v[0].cls;
    for (auto &it : v)
      if (it.cls != l.only) l.only = -1;
    l.grid = (int)std::min<int64_t>(start, 2 * (int64_t)cu_count());
    if (l.grid < 1) l.grid = 1;
    std::vector<int64_t> st(v.size());
    for (size_t j = 0; j < v.size(); ++j) st[j] = v[j].tile_start;
    if (dev_malloc(&l.d_items, (sizeof(GroupItem) + sizeof(int64_t)) * v.size()) != hipSuccess)
      return fail(QS_ERR_MEMORY, "qs_group_create: device table");
    out.push_back(l);
    if (hipMemcpy(l.d_items, v.data(), sizeof(GroupItem) * v.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy((void *)item_starts(l.d_items, l.n), st.data(), sizeof(int64_t) * v.size(),
                  hipMemcpyHostToDevice) != hipSuccess)
      return fail(QS_ERR_DEVICE, "qs_group_create: table upload");
  }
  return QS_OK;
}

static void free_launch_list(std::vector<Launch> &ls) {
  for (auto &l : ls)
    if (l.d_items) (void)dev_free(l.d_items);
  ls.clear();
}

// (Re)build the device tables from the plans and their tile classes.
static qs_status build_tables(qs_group_plan *g) {
  g->tables_ok = false;
  free_launches(g);
  g->held.clear();
  for (const qs_plan *p : g->plans) g->held.push_back(storages_of(p));
  std::vector<int> cls(g->plans.size(), -1);
  for (size_t i = 0; i < g->plans.size(); ++i) {
    qs_plan *p = g->plans[i];
    GemmArgs a = make_args(p, p->desc.a, base_a(g, i), p->desc.b, base_b(g, i), p->desc.c, base_c(g, i));
    if (!grouped_member(a)) {   // the generic gather runs standalone, on the same stream
      g->standalone.push_back((int)i);
      continue;
    }
    cls[i] = member_class(g, i);
  }
  qs_status st = make_launches(g, cls, nullptr, g->launches);
  if (st != QS_OK) free_launches(g);
  else g->tables_ok = true;
  return st;
}

static qs_status set_coefficients(qs_group_plan *g, const double *alphas, const double *betas, hipStream_t s) {
  const size_t n = g->plans.size();
  std::vector<double> c(4 * n);
  for (size_t i = 0; i < n; ++i) {
    c[4 * i] = alphas ? alphas[2 * i] : 1.0;
    c[4 * i + 1] = alphas ? alphas[2 * i + 1] : 0.0;
    c[4 * i + 2] = betas ? betas[2 * i] : 0.0;
    c[4 * i + 3] = betas ? betas[2 * i + 1] : 0.0;
  }
  if (c == g->last_coef) return QS_OK;
  // The previous copy out of the staging buffer may still be in flight.
  if (hipStreamSynchronize(s) != hipSuccess) return fail(QS_ERR_DEVICE, "group: coefficient sync");
  std::memcpy(g->h_coef, c.data(), sizeof(double) * c.size());
  if (hipMemcpyAsync(g->d_coef, g->h_coef, sizeof(double) * c.size(), hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(QS_ERR_DEVICE, "group: coefficient upload");
  g->last_coef.swap(c);
  return QS_OK;
}

static qs_status launch_all(qs_group_plan *g, hipStream_t s) {
  if (g->d_batch_off && !g->plans.empty() && !g->batch_args.mixed) {
    // One batched launch when every member has the same coefficients (and
    // the plan takes the buffer-load gather).
    const std::vector<double> &c = g->last_coef;
    bool same = true;
    for (size_t i = 4; i < c.size() && same; ++i) same = c[i] == c[i % 4];
    if (same) {
      GemmArgs a = g->batch_args;
      a.alpha_re = c[0];
      a.alpha_im = c[1];
      a.beta_re = c[2];
      a.beta_im = c[3];
      a.beta_case = (c[2] == 0 && c[3] == 0) ? 0 : (c[2] == 1 && c[3] == 0) ? 1 : 2;
      if (c[0] == 0 && c[1] == 0) a.K = 0;   // alpha == 0: C = beta C, A and B never read
      hipError_t e = launch_gemm(a, s);
      if (e != hipSuccess) return fail(QS_ERR_DEVICE, std::string("group: batched launch: ") + hipGetErrorString(e));
      return QS_OK;
    }
  }
  for (const Launch &l : g->launches) {
    if (l.total == 0) continue;
    hipError_t e = launch_grouped(l.ct, l.akf, l.bkf, l.dom, l.d_items, item_starts(l.d_items, l.n), l.n, l.total,
                                  g->d_coef, l.grid, s, nullptr, l.only);
    if (e != hipSuccess) return fail(QS_ERR_DEVICE, std::string("group: launch: ") + hipGetErrorString(e));
  }
  for (int i : g->standalone) {
    qs_plan *p = g->plans[i];
    const double *c = g->last_coef.data() + 4 * i;
    qs_status st = run_device(p, p->desc.a, base_a(g, i), p->desc.b, base_b(g, i), p->desc.c, base_c(g, i), c, c + 2, s);
    if (st != QS_OK) return st;
  }
  return QS_OK;
}

// Exact span checks of every member (its plan's storages plus the member's
// base offsets) before anything is launched or captured.
static qs_status check_member_spans(const qs_group_plan *g, const char *who) {
  for (size_t i = 0; i < g->plans.size(); ++i) {
    const qs_plan *p = g->plans[i];
    const qs_contract_desc &d = p->desc;
    const __int128 cb = base_c(g, i), ab = base_a(g, i), bb = base_b(g, i);
    const __int128 clo = cb + p->mC_lo + p->nC_lo, chi = cb + p->mC_hi + p->nC_hi;
    const __int128 alo = ab + p->mA_lo + p->kA_lo, ahi = ab + p->mA_hi + p->kA_hi;
    const __int128 blo = bb + p->kB_lo + p->nB_lo, bhi = bb + p->kB_hi + p->nB_hi;
    if (clo < 0 || chi >= d.c_len || alo < 0 || ahi >= d.a_len || blo < 0 || bhi >= d.b_len)
      return fail(QS_ERR_BOUNDS, std::string(who) + ": member " + std::to_string(i) +
                                     "'s operand span exceeds its storage");
  }
  return QS_OK;
}

static void build_batch_args(qs_group_plan *g);

// Rebuild the device tables when a member plan was rebound since they were
// built (the tables hold storage pointers and bases).  Work already queued
// may still read the old tables, so the device is synchronized first.
static qs_status refresh_tables(qs_group_plan *g) {
  bool same = g->held.size() == g->plans.size();
  for (size_t i = 0; same && i < g->plans.size(); ++i) same = g->held[i] == storages_of(g->plans[i]);
  if (same && g->tables_ok) return QS_OK;
  if (hipDeviceSynchronize() != hipSuccess) return fail(QS_ERR_DEVICE, "group: sync before rebuilding tables");
  qs_status st = build_tables(g);
  if (st != QS_OK) return st;
  if (g->d_batch_off) build_batch_args(g);
  return QS_OK;
}

// The one batched launch of a batched group: the plan's own kernel over all
// coordinates, with the offsets table moving the bases.
static void build_batch_args(qs_group_plan *g) {
  qs_plan *plan = g->plans[0];
  GemmArgs a = make_args(plan, plan->desc.a, plan->desc.a_base, plan->desc.b, plan->desc.b_base, plan->desc.c,
                         plan->desc.c_base);
  a.k_splits = 1;
  a.k_chunk = plan->K;
  a.partials = nullptr;
  if (g->tile_override[0] >= 0) a.tile = g->tile_override[0];
  a.batch_off = g->d_batch_off;
  a.batch = (int64_t)g->plans.size();
  g->batch_args = a;
}

// One batch of members for the serving loop: descriptors already built,
// staged through pinned host buffers owned by the caller's slot.
qs_status launch_member_batch(const std::vector<GemmArgs> &args, const std::vector<int> &cls,
                              const std::vector<double> &coef4, BatchBuffers &B, hipStream_t s) {
  const size_t n = args.size();
  if (n == 0) return QS_OK;
  if (n > (size_t)B.cap) {
    if (B.d_items) (void)dev_free(B.d_items);
    if (B.h_items) (void)pinned_free(B.h_items);
    if (B.d_coef) (void)dev_free(B.d_coef);
    if (B.h_coef) (void)pinned_free(B.h_coef);
    B.d_items = nullptr; B.h_items = nullptr; B.d_coef = nullptr; B.h_coef = nullptr; B.cap = 0;
    const size_t cap = std::max<size_t>(n, 64);
    // cap GroupItems, then cap tile_start values.
    if (dev_malloc((void **)&B.d_items, (sizeof(GroupItem) + sizeof(int64_t)) * cap) != hipSuccess ||
        pinned_malloc((void **)&B.h_items, (sizeof(GroupItem) + sizeof(int64_t)) * cap, hipHostMallocDefault) !=
            hipSuccess ||
        dev_malloc((void **)&B.d_coef, sizeof(double) * 4 * cap) != hipSuccess ||
        pinned_malloc((void **)&B.h_coef, sizeof(double) * 4 * cap, hipHostMallocDefault) != hipSuccess)
      return fail(QS_ERR_MEMORY, "serving: batch buffers");
    B.cap = (int)cap;
  }
  GroupItem *items = (GroupItem *)B.h_items;
  int64_t *starts = reinterpret_cast<int64_t *>(items + B.cap);
  std::map<std::tuple<int, int, int, int>, std::vector<GroupItem>> classes;
  for (size_t i = 0; i < n; ++i) {
    GroupItem it;
    std::memset(&it, 0, sizeof it);
    it.g = args[i];
    it.cls = grouped_class(cls[i], it.g.compute_type);
    const TileShape ts = tile_shape(it.cls, it.g.compute_type, it.g.a_kfast, it.g.b_kfast);
    it.tiles_m = (it.g.M + ts.bm - 1) / ts.bm;
    it.tile_start = it.tiles_m * ((it.g.N + ts.bn - 1) / ts.bn);
    it.index = (int32_t)i;
    classes[std::make_tuple(it.g.compute_type, it.g.a_kfast, it.g.b_kfast, it.g.dom)].push_back(it);
  }
  std::memcpy(B.h_coef, coef4.data(), sizeof(double) * 4 * n);
  struct L { size_t off; int n; int64_t total; int ct, akf, bkf, dom, only; };
  std::vector<L> ls;
  size_t off = 0;
  for (auto &kv : classes) {
    std::vector<GroupItem> &v = kv.second;
    std::stable_sort(v.begin(), v.end(), [](const GroupItem &x, const GroupItem &y) { return x.g.K > y.g.K; });
    int64_t start = 0;
    for (auto &it : v) {
      const int64_t c = it.tile_start;
      it.tile_start = start;
      start += c;
    }
    std::memcpy(items + off, v.data(), sizeof(GroupItem) * v.size());
    for (size_t j = 0; j < v.size(); ++j) starts[off + j] = v[j].tile_start;
    int only = v[0].cls;
    for (auto &it : v)
      if (it.cls != only) only = -1;
    ls.push_back(
        L{off, (int)v.size(), start, std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first),
          std::get<3>(kv.first), only});
    off += v.size();
  }
  GroupItem *d_items = (GroupItem *)B.d_items;
  int64_t *d_starts = reinterpret_cast<int64_t *>(d_items + B.cap);
  if (hipMemcpyAsync(d_items, items, sizeof(GroupItem) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(d_starts, starts, sizeof(int64_t) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(B.d_coef, B.h_coef, sizeof(double) * 4 * n, hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(QS_ERR_DEVICE, "serving: batch upload");
  for (const L &l : ls) {
    if (l.total == 0) continue;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(l.total, 2 * (int64_t)cu_count()));
    hipError_t e = launch_grouped(l.ct, l.akf, l.bkf, l.dom, d_items + l.off, d_starts + l.off, l.n, l.total,
                                  B.d_coef, grid, s, nullptr, l.only);
    if (e != hipSuccess) return fail(QS_ERR_DEVICE, std::string("serving: launch: ") + hipGetErrorString(e));
  }
  return QS_OK;
}

}  // namespace qs

using namespace qs;

extern "C" {

static qs_status check_member(const qs_plan *p, int32_t i, int32_t tile) {
  if (!p) return fail(QS_ERR_ARGUMENT, "qs_group_create: null plan " + std::to_string(i));
  if (p->released) return fail(QS_ERR_ARGUMENT, "qs_group_create: plan " + std::to_string(i) + " was released");
  if (p->desc.memory != QS_MEM_DEVICE)
    return fail(QS_ERR_ARGUMENT, "qs_group_create: plan " + std::to_string(i) + " is not in device memory");
  if (p->path != QS_PATH_GEMM && p->path != QS_PATH_SKINNY)
    return fail(QS_ERR_ARGUMENT, "qs_group_create: plan " + std::to_string(i) +
                                     " does not run the tiled GEMM path (degenerate extents)");
  if (tile < -1 || tile >= TILE_COUNT)
    return fail(QS_ERR_ARGUMENT, "qs_group_create: bad tile for plan " + std::to_string(i));
  return QS_OK;
}

qs_status qs_group_create(qs_plan *const *plans, int32_t n, const int32_t *tiles, qs_group_plan **out) {
  if (!out) return fail(QS_ERR_ARGUMENT, "qs_group_create: null output");
  *out = nullptr;
  if (n < 0 || (n > 0 && !plans)) return fail(QS_ERR_ARGUMENT, "qs_group_create: bad plan array");
  for (int32_t i = 0; i < n; ++i) {
    const qs_status st = check_member(plans[i], i, tiles ? tiles[i] : -1);
    if (st != QS_OK) return st;
  }
  qs_group_plan *g = new qs_group_plan();
  g->plans.assign(plans, plans + n);
  g->tile_override.assign(n, -1);
  if (tiles) g->tile_override.assign(tiles, tiles + n);
  for (qs_plan *p : g->plans) g->flops += (is_complex(p->desc.compute_type) ? 8 : 2) * p->M * p->N * p->K;
  const size_t cb = sizeof(double) * 4 * (size_t)std::max<int32_t>(n, 1);
  if (dev_malloc(&g->d_coef, cb) != hipSuccess || pinned_malloc(&g->h_coef, cb, hipHostMallocDefault) != hipSuccess) {
    qs_group_destroy(g);
    return fail(QS_ERR_MEMORY, "qs_group_create: coefficient buffers");
  }
  qs_status st = build_tables(g);
  if (st != QS_OK) {
    qs_group_destroy(g);
    return st;
  }
  *out = g;
  return QS_OK;
}

qs_status qs_group_create_batched(qs_plan *plan, int32_t n, const int64_t *a_offsets, const int64_t *b_offsets,
                                  const int64_t *c_offsets, int32_t tile, qs_group_plan **out) {
  if (!out) return fail(QS_ERR_ARGUMENT, "qs_group_create_batched: null output");
  *out = nullptr;
  if (!plan || n < 0 || (n > 0 && (!a_offsets || !b_offsets || !c_offsets)))
    return fail(QS_ERR_ARGUMENT, "qs_group_create_batched: null plan or offset array");
  qs_status st = check_member(plan, 0, tile);
  if (st != QS_OK) return st;
  qs_group_plan *g = new qs_group_plan();
  g->plans.assign((size_t)n, plan);
  g->tile_override.assign((size_t)n, tile);
  g->a_off.assign(a_offsets, a_offsets + n);
  g->b_off.assign(b_offsets, b_offsets + n);
  g->c_off.assign(c_offsets, c_offsets + n);
  g->flops = (int64_t)n * (is_complex(plan->desc.compute_type) ? 8 : 2) * plan->M * plan->N * plan->K;
  const size_t cb = sizeof(double) * 4 * (size_t)std::max<int32_t>(n, 1);
  if (dev_malloc(&g->d_coef, cb) != hipSuccess || pinned_malloc(&g->h_coef, cb, hipHostMallocDefault) != hipSuccess) {
    qs_group_destroy(g);
    return fail(QS_ERR_MEMORY, "qs_group_create_batched: coefficient buffers");
  }
  // Every coordinate's spans are checked before anything is built.
  st = check_member_spans(g, "qs_group_create_batched");
  if (st != QS_OK) {
    qs_group_destroy(g);
    return st;
  }
  st = build_tables(g);
  if (st != QS_OK) {
    qs_group_destroy(g);
    return st;
  }
  if (n > 0) {
    std::vector<int64_t> off(3 * (size_t)n);
    for (int32_t i = 0; i < n; ++i) {
      off[3 * i] = a_offsets[i];
      off[3 * i + 1] = b_offsets[i];
      off[3 * i + 2] = c_offsets[i];
    }
    if (dev_malloc(&g->d_batch_off, sizeof(int64_t) * off.size()) != hipSuccess ||
        hipMemcpy(g->d_batch_off, off.data(), sizeof(int64_t) * off.size(), hipMemcpyHostToDevice) != hipSuccess) {
      qs_group_destroy(g);
      return fail(QS_ERR_MEMORY, "qs_group_create_batched: offset table");
    }
    build_batch_args(g);
  }
  *out = g;
  return QS_OK;
}

qs_status qs_group_execute(qs_group_plan *g, const double *alphas, const double *betas, void *stream) {
  if (!g) return fail(QS_ERR_ARGUMENT, "qs_group_execute: null group");
  hipStream_t s = (hipStream_t)stream;
  // Exact span checks for every member before anything is launched.
  qs_status st = check_member_spans(g, "qs_group_execute");
  if (st != QS_OK) return st;
  if ((st = refresh_tables(g)) != QS_OK) return st;
  st = set_coefficients(g, alphas, betas, s);
  if (st != QS_OK) return st;
  return launch_all(g, s);
}

// Per-shape-class tiles for grouped launches, timed in grouped context.  The
// whole group runs (writing a scratch C, with alpha = 1 and beta = 0) once
// with every tunable member on each grouped tile class, the kernel recording
// the workgroup time each member's tiles take inside that mixed launch.  A
// class's cost for a tile is the sum over its members (group_shape_key:
// types, gather directions, octave buckets of M, N, K); the cheapest wins.
// The tuned assignment is then timed against the padding heuristic's, whole
// group against whole group: it is kept only if it is faster, so tuning
// never slows the group down.  The result goes to the autotune cache (saved
// by qs_autotune_save) and the group's tables are rebuilt.  Every tile sums
// K in the same order, so results do not change.  Members with a tile
// override or on the generic gather are left alone.
qs_status qs_group_autotune(qs_group_plan *g, void *stream, int32_t reps, qs_group_tune_result *out) {
  if (!g) return fail(QS_ERR_ARGUMENT, "qs_group_autotune: null group");
  if (out) std::memset(out, 0, sizeof *out);
  if (reps <= 0) reps = 3;
  hipStream_t s = (hipStream_t)stream;
  qs_status st = check_member_spans(g, "qs_group_autotune");
  if (st != QS_OK) return st;
  if ((st = refresh_tables(g)) != QS_OK) return st;
  const size_t n = g->plans.size();
  std::vector<int> fixed(n, -1), heur(n, -1), before_cls(n, -1);
  std::map<std::string, std::vector<size_t>> classes;
  size_t scratch_bytes = 16;
  for (size_t i = 0; i < n; ++i) {
    const qs_plan *p = g->plans[i];
    GemmArgs a = make_args(p, p->desc.a, base_a(g, i), p->desc.b, base_b(g, i), p->desc.c, base_c(g, i));
    if (!grouped_member(a)) continue;   // standalone members take no part
    scratch_bytes = std::max(scratch_bytes, (size_t)(p->mC_hi + p->nC_hi - p->mC_lo - p->nC_lo + 1) *
                                                (size_t)elem_bytes(p->desc.c_type));
    before_cls[i] = member_class(g, i);
    if (g->tile_override[i] >= 0) {
      fixed[i] = before_cls[i];
      continue;
    }
    heur[i] = grouped_class(grouped_tile(p->M, p->N, p->desc.compute_type), p->desc.compute_type);
    classes[group_shape_key(p)].push_back(i);
  }
  if (classes.empty()) return QS_OK;
  void *scratch = nullptr;
  double *d_coef = nullptr;
  unsigned long long *d_cyc = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (dev_malloc(&scratch, scratch_bytes) != hipSuccess || dev_malloc(&d_coef, sizeof(double) * 4 * n) != hipSuccess ||
      dev_malloc(&d_cyc, sizeof(unsigned long long) * n) != hipSuccess)
    st = fail(QS_ERR_MEMORY, "qs_group_autotune: buffers");
  if (st == QS_OK) {
    std::vector<double> coef(4 * n, 0.0);
    for (size_t j = 0; j < n; ++j) coef[4 * j] = 1.0;   // alpha = 1, beta = 0
    if (hipMemcpy(d_coef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice) != hipSuccess)
      st = fail(QS_ERR_DEVICE, "qs_group_autotune: coefficient upload");
  }
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  // One assignment of classes: a warm-up run, a run recording each member's
  // tile time (cyc, when given), then `reps` timed runs (ms, when given).
  auto run = [&](const std::vector<int> &cls, std::vector<unsigned long long> *cyc, double *ms) -> qs_status {
    std::vector<Launch> ls;
    qs_status r = make_launches(g, cls, scratch, ls);
    hipError_t e = hipSuccess;
    auto all = [&](unsigned long long *cy) {
      for (const Launch &l : ls)
        if (e == hipSuccess && l.total > 0)
          e = launch_grouped(l.ct, l.akf, l.bkf, l.dom, l.d_items, item_starts(l.d_items, l.n), l.n, l.total, d_coef,
                             l.grid,
                             s, cy, l.only);
    };
    if (r == QS_OK) all(nullptr);
    if (r == QS_OK && cyc) {
      if (e == hipSuccess) e = hipMemsetAsync(d_cyc, 0, sizeof(unsigned long long) * n, s);
      all(d_cyc);
      cyc->assign(n, 0);
      if (e == hipSuccess) e = hipMemcpyAsync(cyc->data(), d_cyc, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (r == QS_OK && ms) {
      (void)hipEventRecord(e0, s);
      for (int k = 0; k < reps; ++k) all(nullptr);
      (void)hipEventRecord(e1, s);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float f = 0;
      (void)hipEventElapsedTime(&f, e0, e1);
      *ms = (double)f / reps;
    }
    if (r == QS_OK && e != hipSuccess) r = fail(QS_ERR_DEVICE, std::string("qs_group_autotune: ") + hipGetErrorString(e));
    if (hipStreamSynchronize(s) != hipSuccess && r == QS_OK) r = fail(QS_ERR_DEVICE, "qs_group_autotune: sync");
    free_launch_list(ls);
    return r;
  };
  constexpr int NC = 4;
  const int cands[NC] = {TILE_128x64, TILE_64x64, TILE_32x64, TILE_32x32};
  std::vector<unsigned long long> cyc[NC];
  for (int c = 0; c < NC && st == QS_OK; ++c) {
    std::vector<int> cls = fixed;
    for (auto &kv : classes)
      for (size_t i : kv.second) cls[i] = cands[c];
    st = run(cls, &cyc[c], nullptr);
  }
  std::vector<int> tuned = fixed, heuristic = fixed;
  std::map<std::string, int> best_of;
  for (auto &kv : classes) {
    int best = 0;
    unsigned long long best_cost = 0;
    for (int c = 0; c < NC && st == QS_OK; ++c) {
      unsigned long long cost = 0;
      for (size_t i : kv.second) cost += cyc[c][i];
      if (c == 0 || cost < best_cost) { best = c; best_cost = cost; }
    }
    best_of[kv.first] = cands[best];
    for (size_t i : kv.second) {
      tuned[i] = cands[best];
      heuristic[i] = heur[i];
    }
  }
  // Whole group against whole group, alternating.
  double t_tuned = 1e300, t_heur = 1e300;
  for (int k = 0; k < 2 && st == QS_OK; ++k) {
    double a = 0, b = 0;
    st = run(tuned, nullptr, &a);
    if (st == QS_OK) st = run(heuristic, nullptr, &b);
    t_tuned = std::min(t_tuned, a);
    t_heur = std::min(t_heur, b);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)dev_free(scratch);
  (void)dev_free(d_coef);
  (void)dev_free(d_cyc);
  if (st != QS_OK) return st;
  const bool keep = t_tuned < t_heur;
  for (auto &kv : classes)
    tune_store(kv.first, keep ? best_of[kv.first] : heur[kv.second[0]], 1, keep ? t_tuned : t_heur);
  if (!keep) {
    // The heuristic's per-member tiles: a class may hold members the
    // heuristic treats differently, so its cache entry is dropped.
    for (auto &kv : classes) tune_erase(kv.first);
  }
  st = build_tables(g);
  if (st != QS_OK) return st;
  int changed = 0;
  for (size_t i = 0; i < n; ++i)
    if (before_cls[i] >= 0 && member_class(g, i) != before_cls[i]) ++changed;
  if (out) {
    out->classes = (int32_t)classes.size();
    out->candidates = NC;
    out->members_changed = changed;
    out->before_ms = t_heur;
    out->after_ms = keep ? t_tuned : t_heur;
  }
  return QS_OK;
}

qs_status qs_group_flops(const qs_group_plan *g, int64_t *flops) {
  if (!g || !flops) return fail(QS_ERR_ARGUMENT, "qs_group_flops: null argument");
  *flops = g->flops;
  return QS_OK;
}

qs_status qs_group_stats(qs_group_plan *g, int32_t *launches, int32_t *standalone) {
  if (!g || !launches || !standalone) return fail(QS_ERR_ARGUMENT, "qs_group_stats: null argument");
  const qs_status st = refresh_tables(g);
  if (st != QS_OK) return st;
  int32_t n = 0;
  for (const Launch &l : g->launches) n += l.total > 0 ? 1 : 0;
  *launches = n;
  *standalone = (int32_t)g->standalone.size();
  return QS_OK;
}

qs_status qs_group_destroy(qs_group_plan *g) {
  if (!g) return QS_OK;
  free_launches(g);
  if (g->d_coef) (void)dev_free(g->d_coef);
  if (g->h_coef) (void)pinned_free(g->h_coef);
  if (g->d_batch_off) (void)dev_free(g->d_batch_off);
  delete g;
  return QS_OK;
}

qs_status qs_group_capture(qs_group_plan *g, const double *alphas, const double *betas, int32_t repeats, void *stream,
                           qs_graph **out) {
  if (!g || !out) return fail(QS_ERR_ARGUMENT, "qs_group_capture: null argument");
  *out = nullptr;
  if (repeats < 1) return fail(QS_ERR_ARGUMENT, "qs_group_capture: repeats must be >= 1");
  hipStream_t s = (hipStream_t)stream;
  // The same exact span checks as an execute, before anything is captured:
  // the graph replays these storages (a member rebound later needs a new
  // capture).
  qs_status st = check_member_spans(g, "qs_group_capture");
  if (st != QS_OK) return st;
  if ((st = refresh_tables(g)) != QS_OK) return st;
  // Coefficients are uploaded before capture: the graph holds kernel nodes only.
  st = set_coefficients(g, alphas, betas, s);
  if (st != QS_OK) return st;
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
    return fail(QS_ERR_DEVICE, "qs_group_capture: stream");
  if (hipStreamSynchronize(s) != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return fail(QS_ERR_DEVICE, "qs_group_capture: sync");
  }
  qs_graph *gr = new qs_graph();
  hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  for (int r = 0; r < repeats && e == hipSuccess && st == QS_OK; ++r) st = launch_all(g, cs);
  hipError_t e2 = hipStreamEndCapture(cs, &gr->graph);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess && st == QS_OK) e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess || st != QS_OK) {
    qs_graph_destroy(gr);
    return st != QS_OK ? st : fail(QS_ERR_DEVICE, std::string("qs_group_capture: ") + hipGetErrorString(e));
  }
  *out = gr;
  return QS_OK;
}

qs_status qs_graph_launch(qs_graph *gr, void *stream) {
  if (!gr || !gr->exec) return fail(QS_ERR_ARGUMENT, "qs_graph_launch: null graph");
  hipError_t e = hipGraphLaunch(gr->exec, (hipStream_t)stream);
  if (e != hipSuccess) return fail(QS_ERR_DEVICE, std::string("qs_graph_launch: ") + hipGetErrorString(e));
  return QS_OK;
}

qs_status qs_graph_destroy(qs_graph *gr) {
  if (!gr) return QS_OK;
  if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
  if (gr->graph) (void)hipGraphDestroy(gr->graph);
  delete gr;
  return QS_OK;
}

}  // extern "C"
