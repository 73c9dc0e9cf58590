/*
 * ORACLE -- test infrastructure only (see qs_oracle.h).  Never linked into
 * the product.  Offsets, group checks and the ScalarKernel contraction.
 */
#include "qs_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- groups */

/* AxisGroup validation: src/layout/axis_group.jl:9-41.  Lengths >= 0, P >= 1,
 * then (only when no length is zero) the cardinality and every map's offset
 * range sum((L-1)*|S|) must fit Int64. */
int qso_validate_group(int ndim, const int64_t *lengths, int P, const int64_t *strides) {
  if (P < 1) return 1;
  for (int d = 0; d < ndim; ++d)
    if (lengths[d] < 0) return 1;
  for (int d = 0; d < ndim; ++d)
    if (lengths[d] == 0) return 0;
  int64_t card = 1;
  for (int d = 0; d < ndim; ++d)
    if (__builtin_mul_overflow(card, lengths[d], &card)) return 4;
  for (int p = 0; p < P; ++p) {
    int64_t acc = 0;
    for (int d = 0; d < ndim; ++d) {
      if (lengths[d] <= 1) continue;
      int64_t s = strides[p * ndim + d];
      if (s == INT64_MIN) return 4;              /* checked_abs(typemin) */
      int64_t as = s < 0 ? -s : s, term;
      if (__builtin_mul_overflow(lengths[d] - 1, as, &term)) return 4;
      if (__builtin_add_overflow(acc, term, &acc)) return 4;
    }
  }
  return 0;
}

static int64_t group_length(int ndim, const int64_t *lengths) {
  int64_t q = 1;
  for (int d = 0; d < ndim; ++d) q *= lengths[d];
  return q;
}

/* fill_offsets!: src/layout/axis_group.jl:87-145.  Decode `first` by mixed
 * radix, then step with reset-before-carry increments. */
int qso_fill_offsets(int ndim, const int64_t *lengths, int P, const int64_t *strides,
                     int64_t first, int64_t count, int64_t *out) {
  int64_t Q = group_length(ndim, lengths);
  if (count < 0) return 1;
  if (!(0 <= first && first <= Q)) return 3;
  if (count > Q - first) return 3;
  if (count == 0) return 0;
  int64_t x[QS_MAX_DIMS] = {0};
  int64_t offs[4] = {0, 0, 0, 0};
  if (P > 4 || ndim > QS_MAX_DIMS) return 1;
  int64_t r = first;
  for (int d = 0; d < ndim; ++d) {
    int64_t L = lengths[d];
    int64_t xd = r % L;
    r /= L;
    x[d] = xd;
    if (xd != 0)
      for (int p = 0; p < P; ++p) offs[p] += xd * strides[p * ndim + d];
  }
  int64_t t = 0;
  for (;;) {
    for (int p = 0; p < P; ++p) out[p * count + t] = offs[p];
    ++t;
    if (t == count) break;
    int d = 0;
    while (d < ndim) {
      int64_t L = lengths[d];
      if (x[d] < L - 1) {
        x[d] += 1;
        for (int p = 0; p < P; ++p) offs[p] += strides[p * ndim + d];
        break;
      }
      x[d] = 0;
      for (int p = 0; p < P; ++p) offs[p] -= (L - 1) * strides[p * ndim + d];
      ++d;
    }
  }
  return 0;
}

/* map_ramp_step / affine_ramp: src/layout/axis_group.jl:149-180. */
static int map_ramp_step(int ndim, const int64_t *lengths, const int64_t *S, int64_t *step_out) {
  int64_t step = 0, run = 1;
  int started = 0;
  for (int d = 0; d < ndim; ++d) {
    int64_t L = lengths[d];
    if (L == 0) { *step_out = 0; return 1; }
    if (L == 1) continue;
    if (!started) {
      step = S[d];
      run = L;
      started = 1;
    } else {
      __int128 prod = (__int128)run * (__int128)step;
      if (prod != (__int128)S[d]) return 0;
      run *= L;
    }
  }
  *step_out = step;
  return 1;
}

int qso_affine_ramp(int ndim, const int64_t *lengths, int P, const int64_t *strides,
                    int64_t *steps) {
  int64_t tmp[4];
  for (int p = 0; p < P && p < 4; ++p) {
    if (!map_ramp_step(ndim, lengths, strides + p * ndim, &tmp[p])) {
      for (int q = 0; q < P; ++q) steps[q] = 0;
      return 0;
    }
  }
  for (int p = 0; p < P; ++p) steps[p] = tmp[p];
  return 1;
}

/* describe_block: src/layout/tiles.jl:59-79. */
int qso_describe_block(const int64_t *buf, int64_t buflen, int64_t first, int64_t count,
                       int64_t *out) {
  if (first < 0) return 1;
  if (count < 0) return 1;
  if (count > buflen - first) return 2;
  if (count == 0) { out[0] = 0; out[1] = 0; out[2] = 0; out[3] = 1; return 0; }
  int64_t base = buf[first];
  if (count == 1) { out[0] = base; out[1] = 0; out[2] = 1; out[3] = 1; return 0; }
  int64_t stride;
  if (__builtin_sub_overflow(buf[first + 1], buf[first], &stride)) {
    out[0] = base; out[1] = 0; out[2] = count; out[3] = 0; return 0;
  }
  for (int64_t t = 1; t < count - 1; ++t) {
    int64_t diff;
    if (__builtin_sub_overflow(buf[first + t + 1], buf[first + t], &diff) || diff != stride) {
      out[0] = base; out[1] = 0; out[2] = count; out[3] = 0; return 0;
    }
  }
  out[0] = base; out[1] = stride; out[2] = count; out[3] = 1;
  return 0;
}

/* ---------------------------------------------------------- contraction */

static int is_complex(int t) { return t == QS_C32 || t == QS_C64; }
static int is_double(int t) { return t == QS_F64 || t == QS_C64; }

/* compute_type rules, plan.jl:278-285 (the caller passes T; this only checks
 * what the device library also checks). */
int qso_check_types(const qs_contract_desc *d) {
  for (int i = 0; i < 4; ++i) {
    int t = i == 0 ? d->a_type : i == 1 ? d->b_type : i == 2 ? d->c_type : d->compute_type;
    if (t < 0 || t > 3) return 1;
  }
  if (!is_complex(d->c_type) && (is_complex(d->a_type) || is_complex(d->b_type))) return 1;
  return 0;
}

/* Exact offset range of one map of a group: per dim min(0,(L-1)s), max(0,(L-1)s). */
static void map_range(const qs_group *g, int p, int64_t *lo, int64_t *hi) {
  *lo = 0; *hi = 0;
  for (int d = 0; d < g->ndim; ++d) {
    int64_t L = g->lengths[d];
    if (L <= 1) continue;
    int64_t e = (L - 1) * g->strides[p][d];
    if (e < 0) *lo += e; else *hi += e;
  }
}

/* checked_span_bounds (tiles.jl:123-134) over the whole operand. */
static int span_ok(int64_t base, const qs_group *r, int pr, const qs_group *c, int pc, int64_t len) {
  int64_t rlo, rhi, clo, chi;
  map_range(r, pr, &rlo, &rhi);
  map_range(c, pc, &clo, &chi);
  __int128 lo = (__int128)base + rlo + clo, hi = (__int128)base + rhi + chi;
  return lo >= 0 && hi <= (__int128)len - 1;
}

static int64_t glen(const qs_group *g) { return group_length(g->ndim, g->lengths); }

static int64_t *group_table(const qs_group *g, int p) {
  int64_t Q = glen(g);
  int64_t *both = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * (Q > 0 ? Q : 1)));
  int64_t st[2 * QS_MAX_DIMS] = {0};
  for (int d = 0; d < g->ndim; ++d) { st[d] = g->strides[0][d]; st[g->ndim + d] = g->strides[1][d]; }
  qso_fill_offsets(g->ndim, g->lengths, 2, st, 0, Q, both);
  int64_t *one = (int64_t *)malloc(sizeof(int64_t) * (size_t)(Q > 0 ? Q : 1));
  if (Q > 0) memcpy(one, both + p * Q, sizeof(int64_t) * (size_t)Q);
  free(both);
  return one;
}

#define GEN_SCALAR(R, FMA, SUF)                                                        \
  static inline void load_##SUF(const void *p, int t, int64_t idx, int cj, R *re, R *im) { \
    switch (t) {                                                                       \
    case QS_F32: *re = (R)((const float *)p)[idx]; *im = 0; break;                     \
    case QS_F64: *re = (R)((const double *)p)[idx]; *im = 0; break;                    \
    case QS_C32: *re = (R)((const float *)p)[2 * idx];                                 \
      *im = (R)((const float *)p)[2 * idx + 1]; if (cj) *im = -*im; break;             \
    default: *re = (R)((const double *)p)[2 * idx];                                    \
      *im = (R)((const double *)p)[2 * idx + 1]; if (cj) *im = -*im; break;            \
    }                                                                                  \
  }                                                                                    \
  static inline void store_##SUF(void *p, int t, int64_t idx, R re, R im) {            \
    switch (t) {                                                                       \
    case QS_F32: ((float *)p)[idx] = (float)re; break;                                 \
    case QS_F64: ((double *)p)[idx] = (double)re; break;                               \
    case QS_C32: ((float *)p)[2 * idx] = (float)re; ((float *)p)[2 * idx + 1] = (float)im; break; \
    default: ((double *)p)[2 * idx] = (double)re; ((double *)p)[2 * idx + 1] = (double)im; break; \
    }                                                                                  \
  }                                                                                    \
  /* axpby for the three beta cases (interface.jl:99-104, stores.jl:91-124). */        \
  static inline void axpby_##SUF(int cplx, R ar, R ai, R rr, R ri, R br, R bi,        \
                                 int bcase, R orr, R ori, R *outr, R *outi) {         \
    if (!cplx) {                                                                       \
      if (bcase == 0) *outr = ar * rr;                                                 \
      else if (bcase == 1) *outr = FMA(ar, rr, orr);                                   \
      else *outr = FMA(ar, rr, br * orr);                                              \
      *outi = 0;                                                                       \
      return;                                                                          \
    }                                                                                  \
    if (bcase == 0) {                                                                  \
      R t1 = ar * rr, t2 = ai * ri, t3 = ar * ri, t4 = ai * rr;                        \
      *outr = t1 - t2; *outi = t3 + t4;                                                \
      return;                                                                          \
    }                                                                                  \
    R xr = orr, xi = ori;                                                              \
    if (bcase == 2) {                                                                  \
      R u1 = br * orr, u2 = bi * ori, u3 = br * ori, u4 = bi * orr;                    \
      xr = u1 - u2; xi = u3 + u4;                                                      \
    }                                                                                  \
    *outr = FMA(ar, rr, -FMA(ai, ri, -xr));                                            \
    *outi = FMA(ar, ri, FMA(ai, rr, xi));                                              \
  }                                                                                    \
  static int scalar_contract_##SUF(const qs_contract_desc *d, const R *alpha, const R *beta, \
                                   int64_t k_block) {                                  \
    int cplx = is_complex(d->compute_type);                                            \
    R ar = alpha[0], ai = cplx ? alpha[1] : 0, br = beta[0], bi = cplx ? beta[1] : 0;  \
    int64_t M = glen(&d->m), N = glen(&d->n), K = glen(&d->k);                         \
    if (M == 0 || N == 0) return 0;                                                    \
    int64_t *mC = group_table(&d->m, 1), *nC = group_table(&d->n, 1);                 \
    int scale_only = (ar == 0 && ai == 0) || K == 0;                                  \
    if (scale_only) {                                                                  \
      /* scale_tile!: beta==1 no-op, beta==0 zeros without reading C. */               \
      int one = (br == 1 && bi == 0), zero = (br == 0 && bi == 0);                     \
      if (!one)                                                                        \
        for (int64_t j = 0; j < N; ++j)                                                \
          for (int64_t i = 0; i < M; ++i) {                                            \
            int64_t idx = d->c_base + mC[i] + nC[j];                                   \
            if (zero) { store_##SUF(d->c, d->c_type, idx, 0, 0); continue; }           \
            R cr, ci; load_##SUF(d->c, d->c_type, idx, 0, &cr, &ci);                  \
            if (!cplx) store_##SUF(d->c, d->c_type, idx, cr * br, 0);                  \
            else { R u1 = cr * br, u2 = ci * bi, u3 = cr * bi, u4 = ci * br;           \
                   store_##SUF(d->c, d->c_type, idx, u1 - u2, u3 + u4); }              \
          }                                                                            \
      free(mC); free(nC);                                                              \
      return 0;                                                                        \
    }                                                                                  \
    int64_t *mA = group_table(&d->m, 0), *nB = group_table(&d->n, 0);                 \
    int64_t *kA = group_table(&d->k, 0), *kB = group_table(&d->k, 1);                 \
    if (k_block <= 0 || k_block > K) k_block = K;                                      \
    for (int64_t k0 = 0; k0 < K; k0 += k_block) {                                      \
      int64_t k1 = k0 + k_block < K ? k0 + k_block : K;                                \
      int bcase;                                                                       \
      R ebr = br, ebi = bi;                                                            \
      if (k0 > 0) { ebr = 1; ebi = 0; }                                                \
      bcase = (ebr == 0 && ebi == 0) ? 0 : (ebr == 1 && ebi == 0) ? 1 : 2;             \
      for (int64_t j = 0; j < N; ++j)                                                  \
        for (int64_t i = 0; i < M; ++i) {                                              \
          R accr = 0, acci = 0;                                                        \
          for (int64_t p = k0; p < k1; ++p) {                                          \
            R xr, xi, yr, yi;                                                          \
            load_##SUF(d->a, d->a_type, d->a_base + mA[i] + kA[p], d->conj_a, &xr, &xi); \
            load_##SUF(d->b, d->b_type, d->b_base + kB[p] + nB[j], d->conj_b, &yr, &yi); \
            if (!cplx) accr = FMA(xr, yr, accr);                                       \
            else {                                                                     \
              /* planar step, steps.jl:121-167 */                                      \
              R nre = FMA(-xi, yi, FMA(xr, yr, accr));                                 \
              R nim = FMA(xi, yr, FMA(xr, yi, acci));                                  \
              accr = nre; acci = nim;                                                  \
            }                                                                          \
          }                                                                            \
          int64_t idx = d->c_base + mC[i] + nC[j];                                     \
          R orr = 0, ori = 0;                                                          \
          if (bcase != 0) load_##SUF(d->c, d->c_type, idx, 0, &orr, &ori);             \
          R outr, outi;                                                                \
          axpby_##SUF(cplx, ar, ai, accr, acci, ebr, ebi, bcase, orr, ori, &outr, &outi); \
          store_##SUF(d->c, d->c_type, idx, outr, outi);                               \
        }                                                                              \
    }                                                                                  \
    free(mA); free(nB); free(kA); free(kB); free(mC); free(nC);                        \
    return 0;                                                                          \
  }

GEN_SCALAR(double, fma, d)
GEN_SCALAR(float, fmaf, f)

int qso_check_desc(const qs_contract_desc *d) {
  if (qso_check_types(d)) return QS_ERR_ARGUMENT;
  const qs_group *gs[3] = {&d->m, &d->n, &d->k};
  for (int i = 0; i < 3; ++i) {
    const qs_group *g = gs[i];
    if (g->ndim < 0 || g->ndim > QS_MAX_DIMS) return QS_ERR_ARGUMENT;
    int64_t st[2 * QS_MAX_DIMS] = {0};
    for (int dd = 0; dd < g->ndim; ++dd) { st[dd] = g->strides[0][dd]; st[g->ndim + dd] = g->strides[1][dd]; }
    int r = qso_validate_group(g->ndim, g->lengths, 2, st);
    if (r) return r;
  }
  int64_t M = glen(&d->m), N = glen(&d->n), K = glen(&d->k);
  if (M == 0 || N == 0) return 0;
  if (!span_ok(d->c_base, &d->m, 1, &d->n, 1, d->c_len)) return QS_ERR_BOUNDS;
  if (K > 0) {
    if (!span_ok(d->a_base, &d->m, 0, &d->k, 0, d->a_len)) return QS_ERR_BOUNDS;
    if (!span_ok(d->b_base, &d->k, 1, &d->n, 0, d->b_len)) return QS_ERR_BOUNDS;
  }
  return 0;
}

int qso_contract(const qs_contract_desc *d, const void *alpha, const void *beta, int64_t k_block) {
  int r = qso_check_desc(d);
  if (r) return r;
  if (is_double(d->compute_type))
    return scalar_contract_d(d, (const double *)alpha, (const double *)beta, k_block);
  return scalar_contract_f(d, (const float *)alpha, (const float *)beta, k_block);
}

/* ------------------------------------------------------- cache topology */

/* "0-7,16-23" -> 16 (count_cpu_list in src/hardware/target.jl). */
static int count_cpu_list(const char *s) {
  int n = 0;
  while (*s) {
    char *end;
    long a = strtol(s, &end, 10);
    if (end == s) break;
    long b = a;
    s = end;
    if (*s == '-') { b = strtol(s + 1, &end, 10); s = end; }
    n += (int)(b - a + 1);
    if (*s == ',') ++s; else break;
  }
  return n;
}

static int64_t parse_size(const char *s) {
  char *end;
  long long v = strtoll(s, &end, 10);
  if (*end == 'K') v *= 1024; else if (*end == 'M') v *= 1024 * 1024; else if (*end == 'G') v *= 1024LL * 1024 * 1024;
  return v;
}

static int read_line(const char *path, char *buf, int n) {
  FILE *f = fopen(path, "r");
  if (!f) return 0;
  int ok = fgets(buf, n, f) != NULL;
  fclose(f);
  if (ok) { size_t l = strlen(buf); while (l && (buf[l - 1] == '\n' || buf[l - 1] == ' ')) buf[--l] = 0; }
  return ok;
}

/* core_bytes (target.jl:50): a level's bytes over the cores sharing it. */
void qso_cache_shares(int64_t *l1d, int64_t *l2core, int64_t *l3core) {
  int64_t sz[4] = {0, 0, 0, 0};
  int sh[4] = {0, 0, 0, 0};
  for (int idx = 0; idx < 8; ++idx) {
    char path[256], buf[256];
    snprintf(path, sizeof path, "/sys/devices/system/cpu/cpu0/cache/index%d/level", idx);
    if (!read_line(path, buf, sizeof buf)) break;
    int level = atoi(buf);
    snprintf(path, sizeof path, "/sys/devices/system/cpu/cpu0/cache/index%d/type", idx);
    if (!read_line(path, buf, sizeof buf)) continue;
    if (strcmp(buf, "Instruction") == 0) continue;
    if (level < 1 || level > 3) continue;
    snprintf(path, sizeof path, "/sys/devices/system/cpu/cpu0/cache/index%d/size", idx);
    if (read_line(path, buf, sizeof buf)) sz[level] = parse_size(buf);
    snprintf(path, sizeof path, "/sys/devices/system/cpu/cpu0/cache/index%d/shared_cpu_list", idx);
    if (read_line(path, buf, sizeof buf)) sh[level] = count_cpu_list(buf);
  }
  int l1s = sh[1] > 0 ? sh[1] : 1;
  *l1d = sz[1];
  *l2core = sz[2] / (sh[2] / l1s > 0 ? sh[2] / l1s : 1);
  *l3core = sz[3] / (sh[3] / l1s > 0 ? sh[3] / l1s : 1);
}
