/*
 * ORACLE -- test infrastructure only (see qs_oracle.h).  The reference's
 * SIMDKernel / PlanarKernel BLIS nest, restated as the timed CPU baseline.
 * Compiled once per ISA (QSO_ISA = avx512 | avx2) with -O3 -ffp-contract=fast
 * so the vector muladds become FMAs, as SIMD.jl's `muladd` does.
 *
 *   nest.jl:7-122        loop order N block -> K block -> M block -> n_tile -> m_tile,
 *                        beta only on the first K block (nest.jl:83)
 *   pack.jl:43-99        slivers of MR (A) / NR (B) lanes per K step, zero padding
 *   format.jl            real and planar ("1r": [re.. | im..] per K step) formats
 *   steps.jl:73-106      SIMD step: MR/W A vectors x NR broadcast B scalars
 *   steps.jl:121-167     planar step: four real FMAs per complex MAC
 *   stores.jl:150-278    alpha/beta store over the valid rectangle
 */
#include "qs_oracle.h"

#include <stdlib.h>
#include <string.h>

#ifndef QSO_ISA
#error "define QSO_ISA"
#endif
#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define ISAFN(name) CAT(name, QSO_ISA)

static int64_t glen(const qs_group *g) {
  int64_t q = 1;
  for (int d = 0; d < g->ndim; ++d) q *= g->lengths[d];
  return q;
}

/* Both maps of a group, full extent (fill_offsets! from 0). */
static int64_t *tables(const qs_group *g) {
  int64_t Q = glen(g);
  int64_t *t = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * (Q > 0 ? Q : 1)));
  int64_t st[2 * QS_MAX_DIMS] = {0};
  for (int d = 0; d < g->ndim; ++d) { st[d] = g->strides[0][d]; st[g->ndim + d] = g->strides[1][d]; }
  qso_fill_offsets(g->ndim, g->lengths, 2, st, 0, Q, t);
  return t;
}

static int64_t roundup(int64_t x, int64_t n) { return (x + n - 1) / n * n; }

/* modelled_blocking / l1_k_block (blocking.jl:66-77), fallback (:48-49). */
static void blocking(int is_f64, int cplx, int MR, int NR, int64_t *mc, int64_t *kc, int64_t *nc) {
  int64_t l1, l2, l3;
  qso_cache_shares(&l1, &l2, &l3);
  int64_t R = is_f64 ? 8 : 4, reals = cplx ? 2 : 1;
  if (l1 > 0 && l2 > 0) {
    int64_t k = (l1 / 2) / (6 * R);
    if (k < 1) k = 1;
    int64_t m = ((l2 / 2) / (k * R * reals)) / MR * MR;
    int64_t n = ((l2 + l3) / (k * R * reals)) / NR * NR;
    *kc = k; *mc = m > MR ? m : MR; *nc = n > NR ? n : NR;
  } else {
    *mc = (is_f64 ? 128 : 96) / reals; *kc = is_f64 ? 256 : 768; *nc = (is_f64 ? 768 : 1152) / reals;
    if (*mc < 1) *mc = 1;
  }
}

/* -------------------------------------------------------------- real */
#define GEN_REAL(R, W, MV, NR, SUF)                                                     \
  typedef R CAT(vec, SUF) __attribute__((vector_size((W) * sizeof(R))));                \
  static void CAT(mk, SUF)(int64_t kc, const R *Ap, const R *Bp, R *out) {              \
    enum { MR = (MV) * (W) };                                                           \
    CAT(vec, SUF) acc[MV][NR];                                                          \
    memset(acc, 0, sizeof acc);                                                         \
    for (int64_t p = 0; p < kc; ++p) {                                                  \
      CAT(vec, SUF) a[MV];                                                              \
      for (int v = 0; v < (MV); ++v) memcpy(&a[v], Ap + p * MR + v * (W), sizeof a[v]); \
      for (int j = 0; j < (NR); ++j) {                                                  \
        R b = Bp[p * (NR) + j];                                                         \
        for (int v = 0; v < (MV); ++v) acc[v][j] += a[v] * b;                           \
      }                                                                                 \
    }                                                                                   \
    for (int j = 0; j < (NR); ++j)                                                      \
      for (int v = 0; v < (MV); ++v) memcpy(out + j * MR + v * (W), &acc[v][j], sizeof acc[v][j]); \
  }                                                                                     \
  static int CAT(nest, SUF)(const qs_contract_desc *d, R alpha, R beta,                 \
                            int64_t mc, int64_t kc, int64_t nc) {                       \
    enum { MR = (MV) * (W) };                                                           \
    const R *A = (const R *)d->a, *B = (const R *)d->b;                                 \
    R *C = (R *)d->c;                                                                   \
    int64_t M = glen(&d->m), N = glen(&d->n), K = glen(&d->k);                          \
    int64_t *tm = tables(&d->m), *tn = tables(&d->n), *tk = tables(&d->k);              \
    const int64_t *mA = tm, *mC = tm + M, *nB = tn, *nC = tn + N, *kA = tk, *kB = tk + K; \
    int64_t dm, dk, dn;                                                                 \
    blocking(sizeof(R) == 8, 0, MR, NR, &dm, &dk, &dn);                                 \
    if (mc <= 0) mc = dm; if (kc <= 0) kc = dk; if (nc <= 0) nc = dn;                   \
    mc = roundup(mc, MR); nc = roundup(nc, NR);                                         \
    if (mc > roundup(M, MR)) mc = roundup(M, MR);                                       \
    if (nc > roundup(N, NR)) nc = roundup(N, NR);                                       \
    if (kc > K) kc = K;                                                                 \
    R *Ap = (R *)aligned_alloc(64, sizeof(R) * (size_t)roundup(mc * kc + 64, 64));      \
    R *Bp = (R *)aligned_alloc(64, sizeof(R) * (size_t)roundup(nc * kc + 64, 64));      \
    R out[MR * (NR)];                                                                   \
    for (int64_t jc = 0; jc < N; jc += nc) {                                            \
      int64_t nl = N - jc < nc ? N - jc : nc;                                           \
      for (int64_t pc = 0; pc < K; pc += kc) {                                          \
        int64_t kl = K - pc < kc ? K - pc : kc;                                         \
        R be = pc == 0 ? beta : (R)1;                                                   \
        for (int64_t jr = 0; jr < nl; jr += (NR)) {                                     \
          R *s = Bp + (jr / (NR)) * (NR) * kl;                                          \
          for (int64_t p = 0; p < kl; ++p)                                              \
            for (int j = 0; j < (NR); ++j)                                              \
              s[p * (NR) + j] = jr + j < nl ? B[d->b_base + kB[pc + p] + nB[jc + jr + j]] : (R)0; \
        }                                                                               \
        for (int64_t ic = 0; ic < M; ic += mc) {                                        \
          int64_t ml = M - ic < mc ? M - ic : mc;                                       \
          for (int64_t ir = 0; ir < ml; ir += MR) {                                     \
            R *s = Ap + (ir / MR) * MR * kl;                                            \
            for (int64_t p = 0; p < kl; ++p)                                            \
              for (int i = 0; i < MR; ++i)                                              \
                s[p * MR + i] = ir + i < ml ? A[d->a_base + mA[ic + ir + i] + kA[pc + p]] : (R)0; \
          }                                                                             \
          for (int64_t jr = 0; jr < nl; jr += (NR))                                     \
            for (int64_t ir = 0; ir < ml; ir += MR) {                                   \
              CAT(mk, SUF)(kl, Ap + (ir / MR) * MR * kl, Bp + (jr / (NR)) * (NR) * kl, out); \
              int64_t mv = ml - ir < MR ? ml - ir : MR, nv = nl - jr < (NR) ? nl - jr : (NR); \
              for (int64_t j = 0; j < nv; ++j)                                          \
                for (int64_t i = 0; i < mv; ++i) {                                      \
                  int64_t idx = d->c_base + mC[ic + ir + i] + nC[jc + jr + j];          \
                  R r = out[j * MR + i];                                                \
                  C[idx] = be == 0 ? alpha * r : be == 1 ? alpha * r + C[idx]           \
                                                          : alpha * r + be * C[idx];    \
                }                                                                       \
            }                                                                           \
        }                                                                               \
      }                                                                                 \
    }                                                                                   \
    free(Ap); free(Bp); free(tm); free(tn); free(tk);                                   \
    return 0;                                                                           \
  }

/* ----------------------------------------------------------- planar complex */
#define GEN_CPLX(R, W, MV, NR, SUF)                                                     \
  typedef R CAT(vec, SUF) __attribute__((vector_size((W) * sizeof(R))));                \
  static void CAT(mk, SUF)(int64_t kc, const R *Ap, const R *Bp, R *outr, R *outi) {    \
    enum { MR = (MV) * (W) };                                                           \
    CAT(vec, SUF) cr[MV][NR], ci[MV][NR];                                               \
    memset(cr, 0, sizeof cr); memset(ci, 0, sizeof ci);                                 \
    for (int64_t p = 0; p < kc; ++p) {                                                  \
      CAT(vec, SUF) ar[MV], ai[MV], nai[MV];                                            \
      for (int v = 0; v < (MV); ++v) {                                                  \
        memcpy(&ar[v], Ap + 2 * p * MR + v * (W), sizeof ar[v]);                        \
        memcpy(&ai[v], Ap + (2 * p + 1) * MR + v * (W), sizeof ai[v]);                  \
        nai[v] = -ai[v];                                                                \
      }                                                                                 \
      for (int j = 0; j < (NR); ++j) {                                                  \
        R br = Bp[2 * p * (NR) + j], bi = Bp[(2 * p + 1) * (NR) + j];                   \
        for (int v = 0; v < (MV); ++v) {                                                \
          cr[v][j] += ar[v] * br; cr[v][j] += nai[v] * bi;                              \
          ci[v][j] += ar[v] * bi; ci[v][j] += ai[v] * br;                               \
        }                                                                               \
      }                                                                                 \
    }                                                                                   \
    for (int j = 0; j < (NR); ++j)                                                      \
      for (int v = 0; v < (MV); ++v) {                                                  \
        memcpy(outr + j * MR + v * (W), &cr[v][j], sizeof cr[v][j]);                    \
        memcpy(outi + j * MR + v * (W), &ci[v][j], sizeof ci[v][j]);                    \
      }                                                                                 \
  }                                                                                     \
  static int CAT(nest, SUF)(const qs_contract_desc *d, const R *alpha, const R *beta,   \
                            int64_t mc, int64_t kc, int64_t nc) {                       \
    enum { MR = (MV) * (W) };                                                           \
    const R *A = (const R *)d->a, *B = (const R *)d->b;                                 \
    R *C = (R *)d->c;                                                                   \
    R sa = d->conj_a ? (R)-1 : (R)1, sb = d->conj_b ? (R)-1 : (R)1;                     \
    R ar = alpha[0], ai = alpha[1];                                                     \
    int64_t M = glen(&d->m), N = glen(&d->n), K = glen(&d->k);                          \
    int64_t *tm = tables(&d->m), *tn = tables(&d->n), *tk = tables(&d->k);              \
    const int64_t *mA = tm, *mC = tm + M, *nB = tn, *nC = tn + N, *kA = tk, *kB = tk + K; \
    int64_t dm, dk, dn;                                                                 \
    blocking(sizeof(R) == 8, 1, MR, NR, &dm, &dk, &dn);                                 \
    if (mc <= 0) mc = dm; if (kc <= 0) kc = dk; if (nc <= 0) nc = dn;                   \
    mc = roundup(mc, MR); nc = roundup(nc, NR);                                         \
    if (mc > roundup(M, MR)) mc = roundup(M, MR);                                       \
    if (nc > roundup(N, NR)) nc = roundup(N, NR);                                       \
    if (kc > K) kc = K;                                                                 \
    R *Ap = (R *)aligned_alloc(64, sizeof(R) * (size_t)roundup(2 * mc * kc + 64, 64));  \
    R *Bp = (R *)aligned_alloc(64, sizeof(R) * (size_t)roundup(2 * nc * kc + 64, 64));  \
    R outr[MR * (NR)], outi[MR * (NR)];                                                 \
    for (int64_t jc = 0; jc < N; jc += nc) {                                            \
      int64_t nl = N - jc < nc ? N - jc : nc;                                           \
      for (int64_t pc = 0; pc < K; pc += kc) {                                          \
        int64_t kl = K - pc < kc ? K - pc : kc;                                         \
        R br = pc == 0 ? beta[0] : (R)1, bi = pc == 0 ? beta[1] : (R)0;                 \
        int bcase = (br == 0 && bi == 0) ? 0 : (br == 1 && bi == 0) ? 1 : 2;            \
        for (int64_t jr = 0; jr < nl; jr += (NR)) {                                     \
          R *s = Bp + (jr / (NR)) * 2 * (NR) * kl;                                      \
          for (int64_t p = 0; p < kl; ++p)                                              \
            for (int j = 0; j < (NR); ++j) {                                            \
              int ok = jr + j < nl;                                                     \
              int64_t e = ok ? d->b_base + kB[pc + p] + nB[jc + jr + j] : 0;            \
              s[2 * p * (NR) + j] = ok ? B[2 * e] : (R)0;                               \
              s[(2 * p + 1) * (NR) + j] = ok ? sb * B[2 * e + 1] : (R)0;                \
            }                                                                           \
        }                                                                               \
        for (int64_t ic = 0; ic < M; ic += mc) {                                        \
          int64_t ml = M - ic < mc ? M - ic : mc;                                       \
          for (int64_t ir = 0; ir < ml; ir += MR) {                                     \
            R *s = Ap + (ir / MR) * 2 * MR * kl;                                        \
            for (int64_t p = 0; p < kl; ++p)                                            \
              for (int i = 0; i < MR; ++i) {                                            \
                int ok = ir + i < ml;                                                   \
                int64_t e = ok ? d->a_base + mA[ic + ir + i] + kA[pc + p] : 0;          \
                s[2 * p * MR + i] = ok ? A[2 * e] : (R)0;                               \
                s[(2 * p + 1) * MR + i] = ok ? sa * A[2 * e + 1] : (R)0;                \
              }                                                                         \
          }                                                                             \
          for (int64_t jr = 0; jr < nl; jr += (NR))                                     \
            for (int64_t ir = 0; ir < ml; ir += MR) {                                   \
              CAT(mk, SUF)(kl, Ap + (ir / MR) * 2 * MR * kl, Bp + (jr / (NR)) * 2 * (NR) * kl, outr, outi); \
              int64_t mv = ml - ir < MR ? ml - ir : MR, nv = nl - jr < (NR) ? nl - jr : (NR); \
              for (int64_t j = 0; j < nv; ++j)                                          \
                for (int64_t i = 0; i < mv; ++i) {                                      \
                  int64_t idx = d->c_base + mC[ic + ir + i] + nC[jc + jr + j];          \
                  R rr = outr[j * MR + i], ri = outi[j * MR + i];                       \
                  R xr = 0, xi = 0;                                                     \
                  if (bcase == 1) { xr = C[2 * idx]; xi = C[2 * idx + 1]; }             \
                  if (bcase == 2) { R o1 = C[2 * idx], o2 = C[2 * idx + 1];             \
                                    xr = br * o1 - bi * o2; xi = br * o2 + bi * o1; }   \
                  C[2 * idx] = ar * rr - ai * ri + xr;                                  \
                  C[2 * idx + 1] = ar * ri + ai * rr + xi;                              \
                }                                                                       \
            }                                                                           \
        }                                                                               \
      }                                                                                 \
    }                                                                                   \
    free(Ap); free(Bp); free(tm); free(tn); free(tk);                                   \
    return 0;                                                                           \
  }

#if defined(QSO_WIDE)
/* 512-bit vectors: F64 W=8, F32 W=16 (kernel_selection.jl:31-36 menus). */
GEN_REAL(double, 8, 4, 6, d4)
GEN_REAL(double, 8, 2, 6, d2)
GEN_REAL(float, 16, 4, 6, f4)
GEN_REAL(float, 16, 2, 6, f2)
GEN_CPLX(double, 8, 2, 6, z2)
GEN_CPLX(float, 16, 2, 6, c2)
#define LANES_D 8
#define LANES_F 16
#else
/* 256-bit vectors: F64 W=4, F32 W=8, MV = 2 (rule_mv). */
GEN_REAL(double, 4, 2, 6, d4)
GEN_REAL(double, 4, 2, 6, d2)
GEN_REAL(float, 8, 2, 6, f4)
GEN_REAL(float, 8, 2, 6, f2)
GEN_CPLX(double, 4, 2, 6, z2)
GEN_CPLX(float, 8, 2, 6, c2)
#define LANES_D 4
#define LANES_F 8
#endif

/* mv4: the MV = 4 real tile (rule_mv on a non-double-pumped AVX-512 core). */
int ISAFN(qso_blis)(const qs_contract_desc *d, const void *alpha, const void *beta,
                    int64_t mc, int64_t kc, int64_t nc, int mv4) {
  int64_t M = glen(&d->m), N = glen(&d->n), K = glen(&d->k);
  if (M == 0 || N == 0 || K == 0) return qso_contract(d, alpha, beta, 0);
  switch (d->compute_type) {
  case QS_F64:
    return mv4 ? CAT(nest, d4)(d, *(const double *)alpha, *(const double *)beta, mc, kc, nc)
               : CAT(nest, d2)(d, *(const double *)alpha, *(const double *)beta, mc, kc, nc);
  case QS_F32:
    return mv4 ? CAT(nest, f4)(d, *(const float *)alpha, *(const float *)beta, mc, kc, nc)
               : CAT(nest, f2)(d, *(const float *)alpha, *(const float *)beta, mc, kc, nc);
  case QS_C64:
    return CAT(nest, z2)(d, (const double *)alpha, (const double *)beta, mc, kc, nc);
  default:
    return CAT(nest, c2)(d, (const float *)alpha, (const float *)beta, mc, kc, nc);
  }
}

int ISAFN(qso_lanes)(int dtype) { return (dtype == QS_F64 || dtype == QS_C64) ? LANES_D : LANES_F; }
