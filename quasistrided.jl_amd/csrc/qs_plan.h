// The plan object behind the opaque `qs_plan*` (counterpart of ContractPlan,
// src/planning/plan.jl:16-45, and its ContractWorkspace, workspace.jl:24-52).
#pragma once

#include <string>
#include <vector>

#include "qs_internal.h"

struct qs_plan {
  qs_contract_desc desc;      // the caller's descriptor (storages may be rebound)
  int64_t M = 0, N = 0, K = 0;
  qs_group mg, ng, kg;        // groups in the device enumeration order
  int path = QS_PATH_EMPTY;
  int tile = 0;
  int k_splits = 1;
  int64_t k_chunk = 0;
  int a_kfast = 0, b_kfast = 0;
  int c_mfast = 1;             // C's unit-stride side: M (1) or N (0)
  int mixed = 0;
  int vec = 0;                // the gathers' element pairs are adjacent (GemmArgs::vec)
  int ramps = 0;              // RampBits of the maps that are single ramps, and their steps
  int64_t steps[6] = {0, 0, 0, 0, 0, 0};   // mA kA nB kB mC nC
  int moduli = 0;             // split precision: CRT moduli (INT8 GEMMs)
  int64_t partials_bytes = 0;   // capacity of d_partials (split-K partial sums)
  // Device offset tables: mA mC nB nC kA kB, one allocation.
  int64_t *d_tables = nullptr;
  const int64_t *t_mA = nullptr, *t_mC = nullptr, *t_nB = nullptr, *t_nC = nullptr, *t_kA = nullptr, *t_kB = nullptr;
  void *d_partials = nullptr;
  // The workspace's allocator (the library's when `has_al` is 0), and whether
  // the workspace was handed back (qs_plan_release).
  qs_allocator al = {nullptr, nullptr, nullptr};
  int has_al = 0;
  int released = 0;
  // Map ranges (exact min / max offsets of each map).
  int64_t mA_lo = 0, mA_hi = 0, mC_lo = 0, mC_hi = 0, nB_lo = 0, nB_hi = 0, nC_lo = 0, nC_hi = 0;
  int64_t kA_lo = 0, kA_hi = 0, kB_lo = 0, kB_hi = 0;
  // Absolute element spans of the last bounds check.
  int64_t a_lo = 0, a_hi = -1, b_lo = 0, b_hi = -1, c_lo = 0, c_hi = -1;
  // Split precision: residues, scales and the sampled error report.
  void *d_split = nullptr;
  int64_t split_bytes = 0, split_Mp = 0, split_Np = 0, split_Kp = 0;
  // B's norm / residue passes run on a side stream beside A's (default
  // priority), forked from and joined back to the caller's stream by events.
  hipStream_t split_side = nullptr;
  hipEvent_t split_fork = nullptr, split_join = nullptr;
  // 0: no report; 1: report; 2: guarded -- an execute whose report misses
  // the target is redone with more moduli, then with the plain FP64 kernel.
  int split_report_on = 0;
  struct {
    int32_t samples;
    double rel_fro, max_rel;
    int32_t exceeded, escalations, fallback;
  } split_report = {0, 0.0, 0.0, 0, 0, 0};
  // Guarded mode with beta != 0: C's span as it was before the execute.
  void *d_cbackup = nullptr;
  int64_t cbackup_bytes = 0;
  // Host-memory staging buffers (A, B, C spans) and their capacities in bytes.
  void *stage[3] = {nullptr, nullptr, nullptr};
  size_t stage_cap[3] = {0, 0, 0};
};

namespace qs {
// Everything a plan decides without allocating: validation, groups in device
// order, path, gather directions, span ranges, paired-load flag.
qs_status plan_setup(const qs_contract_desc &d, qs_plan *p);
// Point the plan's six table pointers into one block laid out mA mC nB nC kA kB.
void set_table_pointers(qs_plan *p, const int64_t *tables);
// Value of the K tables' pad entries for A (side 0) or B (side 1).
int64_t k_pad_offset(const qs_plan *p, int side);
// The heuristic workgroup tile.
int choose_tile(int64_t M, int64_t N, int ct);
// Exact span checks of the plan's current storages (tiles.jl:123-134).
qs_status check_spans(qs_plan *p, bool reads_ab);
// An error message formatted into the thread's fixed buffer (no allocation).
qs_status fail_fmt(qs_status code, const char *fmt, ...);
// The plan's device workspace, from its allocator (or hipMalloc / hipFree).
void *ws_alloc(qs_plan *p, size_t bytes);
void ws_free(qs_plan *p, void *ptr, size_t bytes);
// Device and pinned host buffers of one in-flight grouped batch.
struct BatchBuffers {
  void *d_items = nullptr, *h_items = nullptr;
  double *d_coef = nullptr, *h_coef = nullptr;
  int cap = 0;
};
// A member the grouped kernel can run: one storage type, or mixed domains
// (complex times real) with every storage type native.  The generic gather
// (other mixed types, spans past 2 GiB) runs standalone.
inline bool grouped_member(const GemmArgs &a) { return !a.mixed || (a.dom && a.dom_buf); }
// Launch `args` (grouped_member GEMM descriptors, K unsplit) as grouped launches
// with tile classes `cls` and coefficients {are, aim, bre, bim} per member.
qs_status launch_member_batch(const std::vector<GemmArgs> &args, const std::vector<int> &cls,
                              const std::vector<double> &coef4, BatchBuffers &B, hipStream_t s);
int64_t group_length(const qs_group &g);
qs_status validate_group(const qs_group &g, const char *name);
void fill_offsets(const qs_group &g, int64_t first, int64_t count, int64_t *o0, int64_t *o1);
void map_range(const qs_group &g, int p, int64_t &lo, int64_t &hi);
// Move dim d of a group to the front; a non-singleton dim with |stride| 1 in map p, or -1.
void move_to_front(qs_group &g, int d);
int unit_dim(const qs_group &g, int p);
qs_status run_device(qs_plan *p, const void *A, int64_t a_base, const void *B, int64_t b_base, void *C, int64_t c_base,
                     const double *al, const double *be, hipStream_t s);
qs_status run_staged(qs_plan *p, const double *al, const double *be, bool reads_ab, hipStream_t s);
// Set a GEMM plan's tile and split-K factor (0 splits: the heuristic),
// reallocating the partial-sum workspace as needed.
qs_status configure_gemm(qs_plan *p, int tile, int splits);
// The autotune cache's choice for the plan's shape class, if any.
bool tune_lookup(const qs_plan *p, int &tile, int &splits);
int grouped_tile(int64_t M, int64_t N, int ct);
// Grouped-launch shape classes and the autotune cache's raw entries.
std::string group_shape_key(const qs_plan *p);
bool tune_find(const std::string &key, int &tile, int &splits);
void tune_store(const std::string &key, int tile, int splits, double ms);
void tune_erase(const std::string &key);
// Split precision (qs_split.hip).
static constexpr int kSplitSamples = 256;
int split_moduli_for(double target, int ct, int64_t K);
const std::vector<int> &split_moduli_table();
qs_status split_prepare(qs_plan *p);
hipError_t launch_split(qs_plan *p, const GemmArgs &g, bool report, hipStream_t s);
// Guarded split execute (split_report_on == 2): report, escalate, fall back.
qs_status run_split_guarded(qs_plan *p, const GemmArgs &g, hipStream_t s);
GemmArgs make_args(const qs_plan *p, const void *A, int64_t a_base, const void *B, int64_t b_base, void *C,
                   int64_t c_base);
}  // namespace qs
