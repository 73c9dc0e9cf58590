/*
 * ORACLE -- test infrastructure only (see qs_oracle.h).  Runtime ISA choice
 * for the BLIS baseline, as src/hardware/target.jl:73-78 detects the ISA and
 * rule_mv/profile_mv (src/planning/kernel_selection.jl) pick the tile.
 */
#include "qs_oracle.h"

#include <stdio.h>
#include <string.h>

int qso_blis_avx512(const qs_contract_desc *, const void *, const void *, int64_t, int64_t, int64_t, int);
int qso_blis_avx2(const qs_contract_desc *, const void *, const void *, int64_t, int64_t, int64_t, int);
int qso_lanes_avx512(int);
int qso_lanes_avx2(int);
int qso_check_desc(const qs_contract_desc *d);

static int has_avx512(void) {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq");
}

/* AMD's AVX-512 cores double-pump 512-bit FMAs (target.jl DOUBLE_PUMPED_CPUS),
 * where the reference keeps the MV = 2 tile. */
static int double_pumped(void) { return __builtin_cpu_is("amd"); }

int qso_contract_blis(const qs_contract_desc *d, const void *alpha, const void *beta,
                      int64_t mc, int64_t kc, int64_t nc) {
  int r = qso_check_desc(d);
  if (r) return r;
  if (d->a_type != d->compute_type || d->b_type != d->compute_type || d->c_type != d->compute_type)
    return QS_ERR_UNSUPPORTED;
  if (has_avx512()) return qso_blis_avx512(d, alpha, beta, mc, kc, nc, !double_pumped());
  return qso_blis_avx2(d, alpha, beta, mc, kc, nc, 0);
}

const char *qso_blis_kernel(int dtype) {
  static char buf[64];
  int wide = has_avx512();
  int W = wide ? qso_lanes_avx512(dtype) : qso_lanes_avx2(dtype);
  int cplx = dtype == QS_C32 || dtype == QS_C64;
  int mv = (wide && !cplx && !double_pumped()) ? 4 : 2;
  snprintf(buf, sizeof buf, "%s %s %dx6 (W=%d)", wide ? "avx512" : "avx2",
           cplx ? "PlanarKernel" : "SIMDKernel", mv * W, W);
  return buf;
}
