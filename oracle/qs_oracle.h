/*
 * ORACLE -- test infrastructure only.  A CPU restatement of QuasiStrided.jl's
 * block-scatter contraction path, used as the checker by tests/, by
 * __graft_entry__.smoke() and as bench.py's cpu_baseline.  Nothing in the
 * product links, loads or calls it.
 *
 * Restated from (paths relative to the reference checkout):
 *   src/layout/axis_group.jl   AxisGroup validation, fill_offsets!, affine_ramp
 *   src/layout/tiles.jl        describe_block, checked_span_bounds
 *   src/microkernels/steps.jl  ScalarKernel/SIMDKernel/PlanarKernel K steps
 *   src/microkernels/stores.jl alpha/beta store expression trees
 *   src/microkernels/interface.jl  scale_tile!, axpby, execute_tile! order
 *   src/execution/nest.jl      BLIS five-loop nest, beta on the first K block
 *   src/execution/execute.jl   execute! short-circuits
 *   src/planning/blocking.jl   modelled_blocking / l1_k_block
 *   src/hardware/target.jl     cache topology and core shares
 *
 * Pins (tests/test_oracle.py): the reference's known-answer tests
 * (test/layout/test_axis_group.jl, test/helpers.jl _worked_fixture,
 * test/execution/test_c_panel.jl), numpy.einsum in float64/complex128 and
 * brute force on tiny inputs.
 */
#ifndef QS_ORACLE_H
#define QS_ORACLE_H

#include <stdint.h>
#include "../include/qs_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* fill_offsets! for P maps: strides[p*ndim + d]; writes out[p*count + t].
 * Returns 0, or 1 (ArgumentError) / 2 (DimensionMismatch) / 3 (BoundsError)
 * with `out` untouched. */
int qso_fill_offsets(int ndim, const int64_t *lengths, int P, const int64_t *strides,
                     int64_t first, int64_t count, int64_t *out);

/* AxisGroup constructor checks (axis_group.jl:9-41): 0 ok, 1 ArgumentError,
 * 4 OverflowError. */
int qso_validate_group(int ndim, const int64_t *lengths, int P, const int64_t *strides);

/* affine_ramp (axis_group.jl:149-180): returns 1 and the per-map steps if every
 * map is a single ramp, else 0. */
int qso_affine_ramp(int ndim, const int64_t *lengths, int P, const int64_t *strides,
                    int64_t *steps);

/* describe_block (tiles.jl:59-79): out = {base, stride, count, regular}.
 * Returns 0, 1 (ArgumentError) or 2 (DimensionMismatch). */
int qso_describe_block(const int64_t *buf, int64_t buflen, int64_t first, int64_t count,
                       int64_t *out);

/* The ScalarKernel contraction on host storages (desc->memory ignored): for
 * every C element, acc = 0, acc = muladd(a, b, acc) along K in K-group order
 * (planar complex step for a complex T), then the store of stores.jl over
 * blocks of `k_block` K steps with beta on the first block only (nest.jl:83);
 * k_block <= 0 means one block of all of K.  Returns a qs_status. */
int qso_contract(const qs_contract_desc *d, const void *alpha, const void *beta,
                 int64_t k_block);

/* The SIMDKernel/PlanarKernel BLIS nest (nest.jl:7-122) with packing
 * (pack.jl:43-99) and the modelled cache blocking (blocking.jl:66-77), as the
 * timed single-thread CPU baseline.  Compute type == every storage type.
 * mc/kc/nc <= 0 take the modelled values.  Returns a qs_status. */
int qso_contract_blis(const qs_contract_desc *d, const void *alpha, const void *beta,
                      int64_t mc, int64_t kc, int64_t nc);

/* What qso_contract_blis runs on this host: "avx512 32x6" etc. */
const char *qso_blis_kernel(int dtype);

/* Detected caches (bytes): L1d, core share of L2, core share of L3. */
void qso_cache_shares(int64_t *l1d, int64_t *l2core, int64_t *l3core);

#ifdef __cplusplus
}
#endif
#endif
