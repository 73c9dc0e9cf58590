// Tile / split-K autotuner with a process-wide cache keyed by shape class.
//
// The reference chooses its register tile from ISA menus and its cache
// blocking from an analytic model (src/planning/kernel_selection.jl:232-256,
// src/planning/blocking.jl:66-77).  On the GPU the choice that matters is the
// workgroup tile and the split-K factor, which set how many workgroups cover
// the 256 CUs and how long each runs; it is measured, not modelled: every
// candidate is timed on the plan's own operands and the fastest is kept for
// the plan and for its shape class.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "qs_plan.h"

namespace qs {

struct TuneEntry {
  int tile, splits;
  double ms;
};

static std::mutex g_tune_mu;
static std::map<std::string, TuneEntry> g_tune;

static int bucket(int64_t x) { return x <= 1 ? 0 : (int)std::floor(2.0 * std::log2((double)x)); }

std::string shape_key(const qs_plan *p) {
  std::ostringstream os;
  const qs_contract_desc &d = p->desc;
  os << d.compute_type << ',' << d.a_type << ',' << d.b_type << ',' << d.c_type << ',' << p->a_kfast << ','
     << p->b_kfast << ',' << bucket(p->M) << ',' << bucket(p->N) << ',' << bucket(p->K);
  return os.str();
}

// The shape class of a member of a grouped launch: types, gather
// directions and octave buckets of M, N, K (coarser than a one-contraction
// plan's: a group holds many shapes, and each class is timed on its own).
static int octave(int64_t x) { return x <= 1 ? 0 : (int)std::floor(std::log2((double)x)); }

std::string group_shape_key(const qs_plan *p) {
  std::ostringstream os;
  const qs_contract_desc &d = p->desc;
  os << "g:" << d.compute_type << ',' << d.a_type << ',' << d.b_type << ',' << d.c_type << ',' << p->a_kfast << ','
     << p->b_kfast << ',' << octave(p->M) << ',' << octave(p->N) << ',' << octave(p->K);
  return os.str();
}

bool tune_find(const std::string &key, int &tile, int &splits) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(key);
  if (it == g_tune.end()) return false;
  tile = it->second.tile;
  splits = it->second.splits;
  return true;
}

void tune_store(const std::string &key, int tile, int splits, double ms) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = TuneEntry{tile, splits, ms};
}

void tune_erase(const std::string &key) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.erase(key);
}

bool tune_lookup(const qs_plan *p, int &tile, int &splits) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = g_tune.find(shape_key(p));
  if (it == g_tune.end()) return false;
  tile = it->second.tile;
  splits = it->second.splits;
  return true;
}

static double time_config(qs_plan *p, void *scratch, int64_t c_base, hipStream_t s, int reps, hipEvent_t e0,
                          hipEvent_t e1, qs_status &st) {
  const double al[2] = {1.0, 0.0}, be[2] = {0.0, 0.0};
  st = run_device(p, p->desc.a, p->desc.a_base, p->desc.b, p->desc.b_base, scratch, c_base, al, be, s);
  if (st != QS_OK) return 0;
  (void)hipEventRecord(e0, s);
  for (int r = 0; r < reps && st == QS_OK; ++r)
    st = run_device(p, p->desc.a, p->desc.a_base, p->desc.b, p->desc.b_base, scratch, c_base, al, be, s);
  (void)hipEventRecord(e1, s);
  if (hipEventSynchronize(e1) != hipSuccess) st = fail(QS_ERR_DEVICE, "autotune: event");
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  return (double)ms / reps;
}

}  // namespace qs

using namespace qs;

extern "C" {

qs_status qs_autotune(qs_plan *p, void *stream, int32_t reps, qs_tune_result *out) {
  if (!p) return fail(QS_ERR_ARGUMENT, "autotune: null plan");
  if (p->released) return fail(QS_ERR_ARGUMENT, "autotune: the plan's workspace was released");
  if (out) std::memset(out, 0, sizeof *out);
  if (p->path != QS_PATH_GEMM && p->path != QS_PATH_SKINNY) {
    if (out) { out->tile = p->tile; out->k_splits = p->k_splits; }
    return QS_OK;   // nothing to choose on the other paths
  }
  if (p->desc.memory != QS_MEM_DEVICE) return fail(QS_ERR_ARGUMENT, "autotune: device-memory plans only");
  if (reps <= 0) reps = 5;
  // Spans (A, B read; C's span sizes the scratch).
  int64_t clo, chi, alo, ahi, blo, bhi;
  clo = p->desc.c_base + p->mC_lo + p->nC_lo;
  chi = p->desc.c_base + p->mC_hi + p->nC_hi;
  alo = p->desc.a_base + p->mA_lo + p->kA_lo;
  ahi = p->desc.a_base + p->mA_hi + p->kA_hi;
  blo = p->desc.b_base + p->kB_lo + p->nB_lo;
  bhi = p->desc.b_base + p->kB_hi + p->nB_hi;
  if (clo < 0 || chi >= p->desc.c_len || alo < 0 || ahi >= p->desc.a_len || blo < 0 || bhi >= p->desc.b_len)
    return fail(QS_ERR_BOUNDS, "autotune: operand spans exceed their storages");
  void *scratch = nullptr;
  const size_t bytes = (size_t)(chi - clo + 1) * elem_bytes(p->desc.c_type);
  hipError_t e = dev_malloc(&scratch, bytes);
  if (e != hipSuccess) return fail(QS_ERR_MEMORY, "autotune: scratch C");
  const int64_t c_base = p->desc.c_base - clo;
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);

  const int tile0 = p->tile, splits0 = p->k_splits;
  qs_status st = QS_OK;
  const double default_ms = time_config(p, scratch, c_base, s, reps, e0, e1, st);
  int best_tile = tile0, best_splits = splits0, count = 1;
  double best_ms = default_ms;
  const bool cplx = is_complex(p->desc.compute_type);
  const int tiles[] = {TILE_128x128, TILE_128x64, TILE_64x128, TILE_64x64, TILE_32x64, TILE_32x32};
  const int splits_menu[] = {1, 2, 4, 8};
  for (int tile : tiles) {
    if (st != QS_OK) break;
    if (cplx && tile == TILE_128x128) continue;   // two accumulator planes do not fit
    const TileShape ts = tile_shape(tile, p->desc.compute_type, p->a_kfast, p->b_kfast);
    const int64_t ntiles = ((p->M + ts.bm - 1) / ts.bm) * ((p->N + ts.bn - 1) / ts.bn);
    for (int sp : splits_menu) {
      if (sp > 1 && (p->K / sp < 128 || ntiles * sp > 8192)) continue;
      if (tile == tile0 && sp == splits0) continue;
      if ((st = configure_gemm(p, tile, sp)) != QS_OK) break;
      const double ms = time_config(p, scratch, c_base, s, reps, e0, e1, st);
      if (st != QS_OK) break;
      ++count;
      // A candidate with more K partitions must win by 3 %: the extra
      // reduction pass and workspace are not worth timing noise.
      const double need = sp > best_splits ? 0.97 * best_ms : best_ms;
      if (ms < need) { best_ms = ms; best_tile = tile; best_splits = sp; }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)dev_free(scratch);
  if (st != QS_OK) {
    configure_gemm(p, tile0, splits0);
    return st;
  }
  st = configure_gemm(p, best_tile, best_splits);
  if (st != QS_OK) return st;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[shape_key(p)] = TuneEntry{best_tile, best_splits, best_ms};
  }
  if (out) {
    out->tile = best_tile;
    out->k_splits = p->k_splits;
    out->best_ms = best_ms;
    out->default_ms = default_ms;
    out->candidates = count;
  }
  return QS_OK;
}

qs_status qs_autotune_save(const char *path) {
  if (!path) return fail(QS_ERR_ARGUMENT, "autotune_save: null path");
  std::ofstream f(path);
  if (!f) return fail(QS_ERR_ARGUMENT, std::string("autotune_save: cannot open ") + path);
  std::lock_guard<std::mutex> lk(g_tune_mu);
  f << "# qs_amd autotune cache: compute,a,b,c,a_kfast,b_kfast,mbucket,nbucket,kbucket tile splits ms"
       " (g: grouped-launch classes, octave buckets)\n";
  for (const auto &kv : g_tune) f << kv.first << ' ' << kv.second.tile << ' ' << kv.second.splits << ' ' << kv.second.ms << '\n';
  return QS_OK;
}

qs_status qs_autotune_load(const char *path) {
  if (!path) return fail(QS_ERR_ARGUMENT, "autotune_load: null path");
  std::ifstream f(path);
  if (!f) return fail(QS_ERR_ARGUMENT, std::string("autotune_load: cannot open ") + path);
  std::string line;
  std::map<std::string, TuneEntry> loaded;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    std::string key;
    TuneEntry e{};
    if (!(is >> key >> e.tile >> e.splits >> e.ms) || e.tile < 0 || e.tile >= TILE_COUNT || e.splits < 1)
      return fail(QS_ERR_ARGUMENT, "autotune_load: malformed line: " + line);
    loaded[key] = e;
  }
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (const auto &kv : loaded) g_tune[kv.first] = kv.second;
  return QS_OK;
}

qs_status qs_autotune_clear(void) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune.clear();
  return QS_OK;
}

}  // extern "C"
