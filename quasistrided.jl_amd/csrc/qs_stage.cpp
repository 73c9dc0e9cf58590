// Host-memory operands (QS_MEM_HOST): the drop-in for a caller whose arrays
// live in host memory (a Julia `Array`).  Each execute copies the touched
// spans of A and B, and C's span, to device buffers owned by the plan, runs
// the device path with the bases shifted into those buffers, and copies C's
// span back.  Elements inside C's span but outside the view round-trip
// unchanged.
#include <hip/hip_runtime.h>

#include <string>

#include "qs_plan.h"

namespace qs {

static qs_status ensure(qs_plan *p, int i, size_t bytes) {
  if (bytes <= p->stage_cap[i]) return QS_OK;
  if (p->stage[i]) (void)dev_free(p->stage[i]);
  p->stage[i] = nullptr;
  p->stage_cap[i] = 0;
  hipError_t e = dev_malloc(&p->stage[i], bytes);
  if (e != hipSuccess) return fail(QS_ERR_MEMORY, std::string("execute!: staging buffer: ") + hipGetErrorString(e));
  p->stage_cap[i] = bytes;
  return QS_OK;
}

qs_status run_staged(qs_plan *p, const double *al, const double *be, bool reads_ab, hipStream_t s) {
  const qs_contract_desc &d = p->desc;
  const bool need_ab = reads_ab && p->K > 0 && p->path != QS_PATH_SCALE;
  const size_t ea = elem_bytes(d.a_type), eb = elem_bytes(d.b_type), ec = elem_bytes(d.c_type);
  const size_t na = need_ab ? (size_t)(p->a_hi - p->a_lo + 1) : 0;
  const size_t nb = need_ab ? (size_t)(p->b_hi - p->b_lo + 1) : 0;
  const size_t nc = (size_t)(p->c_hi - p->c_lo + 1);
  qs_status st;
  if ((st = ensure(p, 0, na * ea + 16)) != QS_OK) return st;
  if ((st = ensure(p, 1, nb * eb + 16)) != QS_OK) return st;
  if ((st = ensure(p, 2, nc * ec + 16)) != QS_OK) return st;
  hipError_t e = hipSuccess;
  if (need_ab) {
    e = hipMemcpyAsync(p->stage[0], (const char *)d.a + p->a_lo * ea, na * ea, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(p->stage[1], (const char *)d.b + p->b_lo * eb, nb * eb, hipMemcpyHostToDevice, s);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(p->stage[2], (const char *)d.c + p->c_lo * ec, nc * ec, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return fail(QS_ERR_DEVICE, std::string("execute!: staging copy: ") + hipGetErrorString(e));
  st = run_device(p, p->stage[0], d.a_base - p->a_lo, p->stage[1], d.b_base - p->b_lo, p->stage[2], d.c_base - p->c_lo,
                  al, be, s);
  if (st != QS_OK) return st;
  e = hipMemcpyAsync((char *)d.c + p->c_lo * ec, p->stage[2], nc * ec, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return fail(QS_ERR_DEVICE, std::string("execute!: result copy: ") + hipGetErrorString(e));
  return QS_OK;
}

}  // namespace qs
