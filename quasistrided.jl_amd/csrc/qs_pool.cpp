// A stream-ordered device memory pool (hipMemPool_t) as a qs_allocator, for
// callers that bring no allocator of their own: plan workspaces come from
// the pool and go back to it, and the pool keeps up to its release threshold
// reserved, so building and destroying plans in a loop reaches the runtime's
// allocator only while the pool grows.
#include <atomic>

#include "qs_internal.h"

namespace qs {
namespace {

struct PoolCtx {
  hipMemPool_t pool = nullptr;
  int device = 0;
  std::atomic<int64_t> live{0};   // blocks handed out and not yet returned
};

// Blocks are taken on the null stream and the allocation is waited for, so
// any stream may use them at once; they go back on the null stream, which
// is safe because the library synchronizes the device before it hands a
// workspace back (qs_plan_release, qs_plan_destroy).
void *pool_alloc(void *ctx, size_t bytes) {
  PoolCtx *c = (PoolCtx *)ctx;
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess) return nullptr;
  if (prev != c->device && hipSetDevice(c->device) != hipSuccess) return nullptr;
  void *p = nullptr;
  hipError_t e = hipMallocFromPoolAsync(&p, bytes, c->pool, nullptr);
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (p) (void)hipFreeAsync(p, nullptr);
    p = nullptr;
  } else {
    c->live.fetch_add(1, std::memory_order_relaxed);
  }
  if (prev != c->device) (void)hipSetDevice(prev);
  return p;
}

void pool_free(void *ctx, void *ptr, size_t) {
  PoolCtx *c = (PoolCtx *)ctx;
  if (!ptr) return;
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess) return;
  if (prev != c->device) (void)hipSetDevice(c->device);
  if (hipFreeAsync(ptr, nullptr) == hipSuccess) c->live.fetch_sub(1, std::memory_order_relaxed);
  else (void)hipGetLastError();
  if (prev != c->device) (void)hipSetDevice(prev);
}

PoolCtx *pool_of(const qs_allocator *a) {
  return (a && a->alloc == pool_alloc && a->ctx) ? (PoolCtx *)a->ctx : nullptr;
}

}  // namespace
}  // namespace qs

using namespace qs;

extern "C" qs_status qs_pool_allocator_create(int device, uint64_t release_threshold, qs_allocator *out) {
  if (!out) return fail(QS_ERR_ARGUMENT, "qs_pool_allocator_create: out is NULL");
  *out = qs_allocator{nullptr, nullptr, nullptr};
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return fail(QS_ERR_ARGUMENT, "qs_pool_allocator_create: no device " + std::to_string(device));
  }
  hipMemPoolProps props = {};
  props.allocType = hipMemAllocationTypePinned;
  props.handleTypes = hipMemHandleTypeNone;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = device;
  PoolCtx *c = new PoolCtx();
  c->device = device;
  if (hipMemPoolCreate(&c->pool, &props) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return fail(QS_ERR_MEMORY, "qs_pool_allocator_create: hipMemPoolCreate failed");
  }
  uint64_t thr = release_threshold;
  if (hipMemPoolSetAttribute(c->pool, hipMemPoolAttrReleaseThreshold, &thr) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipMemPoolDestroy(c->pool);
    delete c;
    return fail(QS_ERR_DEVICE, "qs_pool_allocator_create: setting the release threshold failed");
  }
  *out = qs_allocator{pool_alloc, pool_free, c};
  return QS_OK;
}

extern "C" qs_status qs_pool_allocator_stats(const qs_allocator *a, uint64_t *reserved_bytes, uint64_t *used_bytes,
                                             int64_t *live_blocks) {
  PoolCtx *c = pool_of(a);
  if (!c) return fail(QS_ERR_ARGUMENT, "qs_pool_allocator_stats: not a pool allocator");
  uint64_t r = 0, u = 0;
  if (hipMemPoolGetAttribute(c->pool, hipMemPoolAttrReservedMemCurrent, &r) != hipSuccess ||
      hipMemPoolGetAttribute(c->pool, hipMemPoolAttrUsedMemCurrent, &u) != hipSuccess) {
    (void)hipGetLastError();
    return fail(QS_ERR_DEVICE, "qs_pool_allocator_stats: hipMemPoolGetAttribute failed");
  }
  if (reserved_bytes) *reserved_bytes = r;
  if (used_bytes) *used_bytes = u;
  if (live_blocks) *live_blocks = c->live.load();
  return QS_OK;
}

extern "C" qs_status qs_pool_allocator_destroy(qs_allocator *a) {
  PoolCtx *c = pool_of(a);
  if (!c) return fail(QS_ERR_ARGUMENT, "qs_pool_allocator_destroy: not a pool allocator");
  if (c->live.load() != 0)
    return fail(QS_ERR_ARGUMENT, "qs_pool_allocator_destroy: " + std::to_string(c->live.load()) +
                                     " blocks still held by plans (destroy or release them first)");
  if (hipStreamSynchronize(nullptr) != hipSuccess || hipMemPoolDestroy(c->pool) != hipSuccess) {
    (void)hipGetLastError();
    return fail(QS_ERR_DEVICE, "qs_pool_allocator_destroy: hipMemPoolDestroy failed");
  }
  delete c;
  *a = qs_allocator{nullptr, nullptr, nullptr};
  return QS_OK;
}
