/*
 * skyorc.c -- CPU oracle for the skycube hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product library (libskycube_amd.so) never links, loads or calls it.
 *
 * It restates in plain C (OpenMP for speed) what the reference fixes for a
 * skycube over FP32 points, smaller is better in every dimension:
 *
 *   dominance      q dominates p in subspace s  iff  q_i <= p_i for all i in s
 *                  and q_i < p_i for some i in s; points equal on all of s do
 *                  not dominate each other.  DominateLeft_array returns 1 / -1
 *                  / -2 for "not dominated or equal" / "dominated" / "strictly
 *                  dominated"              (src/utilities/dt_avx2.h:550-640)
 *   strict         q_i < p_i for all i in s (src/utilities/dt_avx2.h:756-866)
 *   cuboid result  sky(s) = not dominated; ext(s) = dominated but not strictly
 *                                           (src/hybrid/mommyq.cpp:50-61)
 *   MDMC bitmap    one bit per subspace s, word s/32, bit s%32, 1 = p is
 *                  dominated in s     (src/skycube/NodeAuxilliary.h:29-60)
 *
 * Three algorithms compute the same skycube and are checked against each
 * other by the tests:
 *   orc_brute      all pairs x all subspaces (tiny inputs)
 *   orc_stsc       STSC's top-down lattice walk: every cuboid from the
 *                  smallest parent's sky+ext      (src/stsc/HybridCube.cpp:30-136)
 *   orc_mdmc       per-point closure of the pair masks over the lattice
 *                  (what the MDMC bitmap means, src/utilities/kernels.cuh:415-908)
 * plus orc_skyalign, the one-cuboid sky/ext split that run_skyalign returns
 * (src/sdsc/skyAlign.cu:9-53), and orc_score_tie_count /
 * orc_score_tie_unresolved, which find the inputs where STSC's FP32
 * Manhattan-norm early exit (src/hybrid/mommyq.cpp:104) could make STSC
 * differ from the definition.
 *
 * Planted-mistake switch: orc_set_mutation() changes the pair test so the
 * tests can show that their checks catch a wrong oracle.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#define ORC_OK 0
#define ORC_ERR_ARG (-1)
#define ORC_ERR_NAN (-2)
#define ORC_ERR_OOM (-3)

static int g_mutation = 0;

void orc_set_mutation(int m) { g_mutation = m; }
int orc_get_mutation(void) { return g_mutation; }

int orc_words_per_point(int d) { return d >= 5 ? 1 << (d - 5) : 1; }

/* Pair masks of dominator candidate q against p: bit i of *lt is q_i < p_i,
 * bit i of *eq is q_i == p_i (IEEE ordered compares, as the AVX tests use). */
static inline void pair_masks(const float *q, const float *p, int d,
                              uint32_t *lt, uint32_t *eq) {
  uint32_t l = 0, e = 0;
  for (int i = 0; i < d; ++i) {
    l |= (uint32_t)(q[i] < p[i]) << i;
    e |= (uint32_t)(q[i] == p[i]) << i;
  }
  switch (g_mutation) {
  case 1: /* planted: ties count as "better" */
    l |= e;
    e = 0;
    break;
  case 2: /* planted: the last dimension is ignored */
    l &= ~(1u << (d - 1));
    e &= ~(1u << (d - 1));
    break;
  case 3: /* planted: ties count as "worse" (equality dropped) */
    e = 0;
    break;
  default:
    break;
  }
  *lt = l;
  *eq = e;
}

/* DominateLeft_array (src/utilities/dt_avx2.h:629-640). */
static inline int dominate_left(const float *q, const float *p, uint32_t s,
                                int d) {
  uint32_t lt, eq;
  pair_masks(q, p, d, &lt, &eq);
  if ((((lt | eq) & s) != s) || ((eq & s) == s))
    return 1;
  if ((lt & s) == s)
    return -2;
  return -1;
}

static int has_nan(const float *rows, uint64_t n, int d) {
  for (uint64_t i = 0; i < n * (uint64_t)d; ++i)
    if (isnan(rows[i]))
      return 1;
  return 0;
}

/* Manhattan norm over the dimensions of s, accumulated in ascending
 * dimension order in FP32, as MommyQ and PQFilter score a tuple
 * (src/hybrid/mommyq.cpp:25-33, src/utilities/pq_filter.h:446-450).  It
 * only orders the candidates (a dominator never has a larger norm).  A row
 * holding both +inf and -inf over s has a NaN norm and is ordered as -inf:
 * it has a -inf there, so does every row that dominates it, and their norms
 * are -inf or NaN too. */
static inline float score_of(const float *p, uint32_t s, int d) {
  float sum = 0.0f;
  for (int i = 0; i < d; ++i)
    if (s >> i & 1u)
      sum += p[i];
  return sum == sum ? sum : -INFINITY;
}

/* ----------------------------------------------------------------------- */
/* Lattice closure over 2^d bits stored as words (bit s = word s>>5, s&31).  */

static const uint32_t kHiHalf[5] = {0xAAAAAAAAu, 0xCCCCCCCCu, 0xF0F0F0F0u,
                                    0xFF00FF00u, 0xFFFF0000u};

/* After this, bit s is set iff some set bit t has s a subset of t. */
static void down_closure(uint32_t *w, int d) {
  const int W = orc_words_per_point(d);
  const int lo = d < 5 ? d : 5;
  for (int k = 0; k < W; ++k) {
    uint32_t x = w[k];
    for (int j = 0; j < lo; ++j)
      x |= (x & kHiHalf[j]) >> (1 << j);
    w[k] = x;
  }
  for (int j = 5; j < d; ++j) {
    const int step = 1 << (j - 5);
    for (int k = 0; k < W; ++k)
      if (k & step)
        w[k ^ step] |= w[k];
  }
}

/* Subspaces outside the scope -- bit 0 (the empty subspace), popcount >
 * max_d, and for d < 5 the unused bits s >= 2^d -- have no computed skyline;
 * they are reported as "not a member" (bit set). */
static void mask_scope(uint32_t *w, int d, int max_d) {
  const int W = orc_words_per_point(d);
  for (int k = 0; k < W; ++k) {
    uint32_t keep = 0;
    for (int b = 0; b < 32; ++b) {
      const uint32_t s = ((uint32_t)k << 5) | (uint32_t)b;
      if (s == 0 || s >= (1u << d))
        continue;
      if (__builtin_popcount(s) <= max_d)
        keep |= 1u << b;
    }
    w[k] = (w[k] & keep) | ~keep;
  }
}

static inline void set_bit(uint32_t *w, uint32_t s) { w[s >> 5] |= 1u << (s & 31); }
static inline int get_bit(const uint32_t *w, uint32_t s) { return (w[s >> 5] >> (s & 31)) & 1u; }

/* ----------------------------------------------------------------------- */
/* One-cuboid sky/ext split (the contract of run_skyalign,                 */
/* src/sdsc/skyAlign.cuh:27-318, and of MommyQ::Execute,                  */
/* src/hybrid/mommyq.cpp:50-61).                                           */

typedef struct {
  float score;
  uint32_t idx; /* position in the candidate list */
} scored_t;

static int cmp_scored(const void *a, const void *b) {
  const scored_t *x = (const scored_t *)a, *y = (const scored_t *)b;
  if (x->score < y->score) return -1;
  if (x->score > y->score) return 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/*
 * status[k] for candidate cand[k] (or row k when cand == NULL):
 *   0 = in sky(s), 1 = in ext(s) \ sky(s), 2 = strictly dominated in s.
 * Only the candidates are compared with each other (STSC and SDSC pass the
 * smallest parent's sky+ext, which holds a dominator of every dominated
 * candidate: src/stsc/HybridCube.cpp:76-100).
 *
 * Method: sort by FP32 score.  A dominator never has a larger score (FP
 * addition is monotone), so p is tested only against q with score <= p's.
 * A first pass against the `npr` lowest-score candidates removes most
 * strictly dominated points (the role of PQFilter, src/utilities/pq_filter.h:
 * 418-505); the survivors are then tested against each other only, which is
 * exact: a survivor's dominators are survivors themselves, or their strict
 * dominator in the pruner set would have removed it.
 */
int orc_skyalign(const float *rows, uint32_t n_rows, int d, uint32_t sub,
                 const uint32_t *cand, uint32_t n_cand, uint8_t *status) {
  if (d < 1 || d > 16 || sub == 0 || (sub >> d) != 0 || !rows || !status)
    return ORC_ERR_ARG;
  if (has_nan(rows, n_rows, d))
    return ORC_ERR_NAN;
  const uint32_t c = cand ? n_cand : n_rows;
  if (c == 0)
    return ORC_OK;
  scored_t *sc = (scored_t *)malloc(sizeof(scored_t) * c);
  uint32_t *surv = (uint32_t *)malloc(sizeof(uint32_t) * c);
  if (!sc || !surv) {
    free(sc);
    free(surv);
    return ORC_ERR_OOM;
  }
#define ROW(k) (rows + (uint64_t)(cand ? cand[k] : (k)) * (uint64_t)d)
  for (uint32_t k = 0; k < c; ++k) {
    if (cand && cand[k] >= n_rows) {
      free(sc);
      free(surv);
      return ORC_ERR_ARG;
    }
    sc[k].score = score_of(ROW(k), sub, d);
    sc[k].idx = k;
  }
  qsort(sc, c, sizeof(scored_t), cmp_scored);
  const uint32_t npr = c < 1024 ? c : 1024;

#pragma omp parallel for schedule(dynamic, 256)
  for (uint32_t r = 0; r < c; ++r) {
    const float *p = ROW(sc[r].idx);
    uint8_t st = 0;
    for (uint32_t t = 0; t < npr; ++t) {
      if (sc[t].score > sc[r].score)
        break;
      if (t == r)
        continue;
      if (dominate_left(ROW(sc[t].idx), p, sub, d) == -2) {
        st = 2;
        break;
      }
    }
    status[sc[r].idx] = st;
  }
  uint32_t ns = 0;
  for (uint32_t r = 0; r < c; ++r)
    if (status[sc[r].idx] != 2)
      surv[ns++] = r; /* rank in score order */

#pragma omp parallel for schedule(dynamic, 64)
  for (uint32_t a = 0; a < ns; ++a) {
    const uint32_t r = surv[a];
    const float *p = ROW(sc[r].idx);
    uint8_t st = 0;
    for (uint32_t b = 0; b < ns; ++b) {
      const uint32_t t = surv[b];
      if (sc[t].score > sc[r].score)
        break;
      if (t == r)
        continue;
      const int res = dominate_left(ROW(sc[t].idx), p, sub, d);
      if (res == -2) {
        st = 2;
        break;
      }
      if (res == -1)
        st = 1;
    }
    status[sc[r].idx] = st;
  }
#undef ROW
  free(sc);
  free(surv);
  return ORC_OK;
}

/* Extended skyline of the full space: ascending row ids not strictly
 * dominated (sky + ext of src/sdsc/skyAlign.cuh as used by MDMC,
 * src/mdmc/MDMC.cu:504-522).  Same method as orc_skyalign, with the strict
 * test alone, which stops at the first column where q is not better. */
static inline int strictly_dominates(const float *q, const float *p, int d) {
  for (int i = 0; i < d; ++i)
    if (!(q[i] < p[i]))
      return 0;
  return 1;
}

/* The strict-only split behind orc_ext_skyline.  With stride > 1 only every
 * stride-th survivor of the prefilter gets the survivor pass (for timing a
 * bounded part of a large input; dead[] is then incomplete).  Wall times of
 * the two passes go to *t_prune / *t_self when not NULL. */
static int ext_split(const float *rows, uint32_t n, int d, uint32_t stride,
                     uint8_t *in, double *t_prune, double *t_self,
                     uint32_t *n_surv) {
  const uint32_t full = (1u << d) - 1u;
  scored_t *sc = (scored_t *)malloc(sizeof(scored_t) * n);
  uint8_t *dead = (uint8_t *)calloc(n, 1);
  uint32_t *surv = (uint32_t *)malloc(sizeof(uint32_t) * n);
  if (!sc || !dead || !surv) {
    free(sc);
    free(dead);
    free(surv);
    return ORC_ERR_OOM;
  }
  double t0 = omp_get_wtime();
  for (uint32_t k = 0; k < n; ++k) {
    sc[k].score = score_of(rows + (size_t)k * d, full, d);
    sc[k].idx = k;
  }
  qsort(sc, n, sizeof(scored_t), cmp_scored);
  const uint32_t npr = n < 1024 ? n : 1024;
#pragma omp parallel for schedule(dynamic, 256)
  for (uint32_t r = 0; r < n; ++r) {
    const float *p = rows + (size_t)sc[r].idx * d;
    for (uint32_t t = 0; t < npr; ++t) {
      if (sc[t].score > sc[r].score)
        break;
      if (t != r && strictly_dominates(rows + (size_t)sc[t].idx * d, p, d)) {
        dead[r] = 1;
        break;
      }
    }
  }
  uint32_t ns = 0;
  for (uint32_t r = 0; r < n; ++r)
    if (!dead[r])
      surv[ns++] = r;
  double t1 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 64)
  for (uint32_t a = 0; a < ns; a += stride) {
    const uint32_t r = surv[a];
    const float *p = rows + (size_t)sc[r].idx * d;
    for (uint32_t b = 0; b < ns; ++b) {
      const uint32_t t = surv[b];
      if (sc[t].score > sc[r].score)
        break;
      if (t != r && strictly_dominates(rows + (size_t)sc[t].idx * d, p, d)) {
        dead[r] = 1;
        break;
      }
    }
  }
  double t2 = omp_get_wtime();
  if (t_prune)
    *t_prune = t1 - t0;
  if (t_self)
    *t_self = t2 - t1;
  if (n_surv)
    *n_surv = ns;
  if (in)
    for (uint32_t r = 0; r < n; ++r)
      in[sc[r].idx] = !dead[r];
  free(sc);
  free(dead);
  free(surv);
  return ORC_OK;
}

int orc_ext_skyline(const float *rows, uint32_t n, int d, uint32_t *ext_idx,
                    uint32_t *n_ext) {
  if (d < 1 || d > 16)
    return ORC_ERR_ARG;
  if (n == 0) {
    *n_ext = 0;
    return ORC_OK;
  }
  if (has_nan(rows, n, d))
    return ORC_ERR_NAN;
  uint8_t *in = (uint8_t *)calloc(n, 1);
  if (!in)
    return ORC_ERR_OOM;
  int rc;
  if (g_mutation != 0) { /* planted mistakes live in pair_masks: use it */
    rc = orc_skyalign(rows, n, d, (1u << d) - 1u, NULL, 0, in);
    for (uint32_t i = 0; i < n; ++i)
      in[i] = in[i] != 2;
  } else {
    rc = ext_split(rows, n, d, 1, in, NULL, NULL, NULL);
  }
  uint32_t m = 0;
  for (uint32_t i = 0; rc == ORC_OK && i < n; ++i)
    if (in[i])
      ext_idx[m++] = i;
  *n_ext = m;
  free(in);
  return rc;
}

/* Times the extended-skyline split with the survivor pass run on every
 * stride-th survivor only (a bounded part of a large input). */
int orc_ext_skyline_timed(const float *rows, uint32_t n, int d, uint32_t stride,
                          double *t_prune, double *t_self, uint32_t *n_surv) {
  if (d < 1 || d > 16 || stride == 0 || n == 0)
    return ORC_ERR_ARG;
  return ext_split(rows, n, d, stride, NULL, t_prune, t_self, n_surv);
}

/* ----------------------------------------------------------------------- */
/* Per-point membership (the MDMC bitmap).                                 */
/*
 * For p and a candidate q let L = {i : q_i <= p_i}, E = {i : q_i == p_i},
 * T = L \ E.  q dominates p exactly in the subspaces s with s subset of L and
 * s not a subset of E, i.e. s subset of L and s meeting T.  Every dominator of
 * an extended-skyline point has a representative in the extended skyline, so
 * q ranges over the extended skyline only.
 *
 *   E empty : all non-empty subsets of L                -> mark L in A0
 *   E = L   : nothing
 *   else    : for each j in T, the subsets of L holding j -> mark L in plane j
 * D = closure(A0)  union over j of ({s : j in s} intersect closure(plane j)).
 * Bits of subspaces outside the scope (bit 0, popcount > max_d) are set.
 */
#define ORC_BLOCK 32 /* points per block of membership_block */

/* Membership of np <= ORC_BLOCK points (rows pts[b]) against the candidates
 * cand[0 .. n_cand-1] (n_cand x d rows, contiguous).  The candidates are the
 * outer loop, so each candidate row is read once per block rather than once
 * per point.  out: np x W words; planes: np x d x W words of scratch. */
static void membership_block(const float *rows, int d, const float *cand,
                             uint32_t n_cand, const uint32_t *pts, int np,
                             int max_d, uint32_t *out, uint32_t *planes) {
  const int W = orc_words_per_point(d);
  uint32_t used[ORC_BLOCK];
  memset(out, 0, sizeof(uint32_t) * W * (size_t)np);
  for (int b = 0; b < np; ++b)
    used[b] = 0;
  for (uint32_t k = 0; k < n_cand; ++k) {
    const float *q = cand + (uint64_t)k * d;
    for (int b = 0; b < np; ++b) {
      const float *p = rows + (uint64_t)pts[b] * d;
      uint32_t lt, eq;
      pair_masks(q, p, d, &lt, &eq);
      const uint32_t le = lt | eq;
      if (lt == 0)
        continue;
      uint32_t *ob = out + (size_t)b * W;
      if (eq == 0) {
        set_bit(ob, le);
      } else {
        uint32_t *pb = planes + (size_t)b * d * W;
        for (uint32_t t = lt; t; t &= t - 1) {
          const int j = __builtin_ctz(t);
          if (!(used[b] >> j & 1u)) {
            memset(pb + (size_t)j * W, 0, sizeof(uint32_t) * W);
            used[b] |= 1u << j;
          }
          set_bit(pb + (size_t)j * W, le);
        }
      }
    }
  }
  for (int b = 0; b < np; ++b) {
    uint32_t *ob = out + (size_t)b * W;
    uint32_t *pb = planes + (size_t)b * d * W;
    down_closure(ob, d);
    for (uint32_t t = used[b]; t; t &= t - 1) {
      const int j = __builtin_ctz(t);
      uint32_t *pl = pb + (size_t)j * W;
      down_closure(pl, d);
      for (int k = 0; k < W; ++k) {
        uint32_t m;
        if (j < 5)
          m = kHiHalf[j];
        else
          m = (k >> (j - 5) & 1) ? 0xFFFFFFFFu : 0u;
        ob[k] |= pl[k] & m;
      }
    }
    mask_scope(ob, d, max_d);
  }
}

/* The candidate rows (the extended skyline) gathered contiguously. */
static float *gather_rows(const float *rows, int d, const uint32_t *ext, uint32_t m) {
  float *c = (float *)malloc(sizeof(float) * (size_t)d * (m ? m : 1));
  if (!c)
    return NULL;
  for (uint32_t k = 0; k < m; ++k)
    memcpy(c + (size_t)k * d, rows + (uint64_t)ext[k] * d, sizeof(float) * d);
  return c;
}

/*
 * Membership of every extended-skyline point.  ext_idx receives the row ids
 * in ascending order, dom_words receives words_per_point(d) words per
 * extended-skyline point in that order; bit s = 1 iff dominated in s.
 * dom_words must hold n * words_per_point(d) words (the worst case).
 */
int orc_mdmc(const float *rows, uint32_t n, int d, int max_d, uint32_t *ext_idx,
             uint32_t *n_ext, uint32_t *dom_words) {
  if (d < 1 || d > 16 || max_d < 1)
    return ORC_ERR_ARG;
  if (max_d > d)
    max_d = d;
  int rc = orc_ext_skyline(rows, n, d, ext_idx, n_ext);
  if (rc != ORC_OK)
    return rc;
  const int W = orc_words_per_point(d);
  const uint32_t m = *n_ext;
  float *cand = gather_rows(rows, d, ext_idx, m);
  if (!cand)
    return ORC_ERR_OOM;
  int oom = 0;
#pragma omp parallel
  {
    uint32_t *planes = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)W * d * ORC_BLOCK);
    if (!planes) {
#pragma omp atomic write
      oom = 1;
    }
#pragma omp for schedule(dynamic, 1)
    for (uint32_t a = 0; a < m; a += ORC_BLOCK) {
      const int np = (int)(m - a < ORC_BLOCK ? m - a : ORC_BLOCK);
      if (planes)
        membership_block(rows, d, cand, m, ext_idx + a, np, max_d,
                         dom_words + (size_t)a * W, planes);
    }
    free(planes);
  }
  free(cand);
  return oom ? ORC_ERR_OOM : ORC_OK;
}

/* Membership bitmaps of the extended-skyline points ext_idx[sample[k]] only
 * (k < n_sample), each against the whole extended skyline: the per-point
 * cost of orc_mdmc on a sample, for timing the CPU path on a bounded part of
 * a large input. */
int orc_mdmc_sample(const float *rows, int d, int max_d, const uint32_t *ext_idx,
                    uint32_t n_ext, const uint32_t *sample, uint32_t n_sample,
                    uint32_t *dom_words) {
  if (d < 1 || d > 16 || max_d < 1)
    return ORC_ERR_ARG;
  if (max_d > d)
    max_d = d;
  const int W = orc_words_per_point(d);
  for (uint32_t k = 0; k < n_sample; ++k)
    if (sample[k] >= n_ext)
      return ORC_ERR_ARG;
  float *cand = gather_rows(rows, d, ext_idx, n_ext);
  if (!cand)
    return ORC_ERR_OOM;
  int bad = 0;
#pragma omp parallel
  {
    uint32_t *planes = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)W * d * ORC_BLOCK);
    uint32_t pts[ORC_BLOCK];
#pragma omp for schedule(dynamic, 1)
    for (uint32_t k0 = 0; k0 < n_sample; k0 += ORC_BLOCK) {
      const int np = (int)(n_sample - k0 < ORC_BLOCK ? n_sample - k0 : ORC_BLOCK);
      if (!planes) {
#pragma omp atomic write
        bad = 1;
        continue;
      }
      for (int b = 0; b < np; ++b)
        pts[b] = ext_idx[sample[k0 + b]];
      membership_block(rows, d, cand, n_ext, pts, np, max_d, dom_words + (size_t)k0 * W,
                       planes);
    }
    free(planes);
  }
  free(cand);
  return bad ? ORC_ERR_OOM : ORC_OK;
}

/* ----------------------------------------------------------------------- */
/* Brute force: every point, every subspace, every other point.            */
/* dom bit s: p is dominated in s;  sdom bit s: p is strictly dominated.    */
int orc_brute(const float *rows, uint32_t n, int d, uint32_t *dom,
              uint32_t *sdom) {
  if (d < 1 || d > 16)
    return ORC_ERR_ARG;
  const int W = orc_words_per_point(d);
  const uint32_t S = 1u << d;
#pragma omp parallel for schedule(dynamic, 4)
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t *dw = dom + (size_t)i * W, *sw = sdom + (size_t)i * W;
    memset(dw, 0, sizeof(uint32_t) * W);
    memset(sw, 0, sizeof(uint32_t) * W);
    for (uint32_t s = 1; s < S; ++s) {
      for (uint32_t j = 0; j < n; ++j) {
        if (j == i)
          continue;
        const int r = dominate_left(rows + (size_t)j * d, rows + (size_t)i * d, s, d);
        if (r <= -1)
          set_bit(dw, s);
        if (r == -2) {
          set_bit(sw, s);
          break;
        }
      }
    }
  }
  return ORC_OK;
}

/* ----------------------------------------------------------------------- */
/* STSC: top-down lattice (src/stsc/HybridCube.cpp:30-136).                */

/* parents of s in the order generate_lattice appends them
 * (src/utilities/lattice_utilities.cpp:52-79): descending parent id. */
static int parents_of(uint32_t s, int d, uint32_t *par) {
  int k = 0;
  for (int j = d - 1; j >= 0; --j)
    if (!(s >> j & 1u))
      par[k++] = s | (1u << j);
  return k;
}

/*
 * For every point and every subspace s with popcount <= max_d (plus the full
 * space, which STSC always computes): dom bit s = not in sky(s), sdom bit s =
 * not in sky(s) or ext(s).  Bits of skipped levels stay clear.  The cuboids of
 * a level are computed from the smallest parent's sky+ext, or from the full
 * space's when the parents were skipped (HybridCube.cpp:76-100); each cuboid
 * is a pairwise DominateLeft split over its candidates (HybridCube.cpp:
 * 113-131).  sky/ext sizes per cuboid go to sky_n / ext_n (2^d entries).
 */
int orc_stsc(const float *rows, uint32_t n, int d, int max_d, uint32_t *dom,
             uint32_t *sdom, uint32_t *sky_n, uint32_t *ext_n) {
  if (d < 1 || d > 16 || max_d < 1)
    return ORC_ERR_ARG;
  if (has_nan(rows, n, d))
    return ORC_ERR_NAN;
  const int W = orc_words_per_point(d);
  const uint32_t S = 1u << d, full = S - 1u;
  uint32_t **lists = (uint32_t **)calloc(S, sizeof(uint32_t *)); /* sky then ext */
  uint32_t *lsky = (uint32_t *)calloc(S, sizeof(uint32_t));
  uint32_t *lext = (uint32_t *)calloc(S, sizeof(uint32_t));
  if (!lists || !lsky || !lext)
    return ORC_ERR_OOM;
  memset(dom, 0, sizeof(uint32_t) * (size_t)W * n);
  memset(sdom, 0, sizeof(uint32_t) * (size_t)W * n);

  /* levels from d down to 1; the full space first */
  for (int lev = d; lev >= 1; --lev) {
    if (lev != d && lev > max_d)
      continue;
    /* the cuboids of this level */
    uint32_t *subs = (uint32_t *)malloc(sizeof(uint32_t) * S);
    uint32_t ns = 0;
    for (uint32_t s = 1; s < S; ++s)
      if (__builtin_popcount(s) == lev)
        subs[ns++] = s;
#pragma omp parallel for schedule(dynamic, 1)
    for (uint32_t a = 0; a < ns; ++a) {
      const uint32_t s = subs[a];
      const uint32_t *cand_sky, *cand_ext;
      uint32_t n_sky = 0, n_ext = 0;
      uint32_t *all = NULL;
      if (s == full) {
        all = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
        for (uint32_t i = 0; i < n; ++i)
          all[i] = i;
        cand_sky = all;
        n_sky = n;
        cand_ext = NULL;
      } else {
        uint32_t par[16] = {0};
        const int np = parents_of(s, d, par);
        uint32_t src = full;
        if (lsky[par[0]] != 0) {
          uint32_t best = lsky[par[0]] + lext[par[0]];
          src = par[0];
          for (int k = 1; k < np; ++k)
            if (lsky[par[k]] + lext[par[k]] < best) {
              best = lsky[par[k]] + lext[par[k]];
              src = par[k];
            }
        }
        cand_sky = lists[src];
        n_sky = lsky[src];
        cand_ext = lists[src] + lsky[src];
        n_ext = lext[src];
      }
      const uint32_t c = n_sky + n_ext;
      uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (c ? c : 1));
      memcpy(cand, cand_sky, sizeof(uint32_t) * n_sky);
      if (n_ext)
        memcpy(cand + n_sky, cand_ext, sizeof(uint32_t) * n_ext);
      int8_t *res = (int8_t *)calloc(c ? c : 1, 1);
      for (uint32_t i = 0; i < c; ++i) {
        for (uint32_t j = 0; j < c; ++j) {
          if (j == i)
            continue;
          const int r = dominate_left(rows + (size_t)cand[j] * d,
                                      rows + (size_t)cand[i] * d, s, d);
          if (r == -2) {
            res[i] = -2;
            break;
          }
          if (r == -1)
            res[i] = -1;
        }
      }
      uint32_t *out = (uint32_t *)malloc(sizeof(uint32_t) * (c ? c : 1));
      uint32_t ks = 0, ke = 0;
      for (uint32_t i = 0; i < c; ++i)
        if (res[i] == 0)
          out[ks++] = cand[i];
      for (uint32_t i = 0; i < c; ++i)
        if (res[i] == -1)
          out[ks + ke++] = cand[i];
      lists[s] = out;
      lsky[s] = ks;
      lext[s] = ke;
      free(res);
      free(cand);
      free(all);
    }
    free(subs);
  }

  /* Points outside sky(s) are dominated; outside sky(s)+ext(s), strictly. */
  for (uint32_t s = 1; s < S; ++s) {
    const int lev = __builtin_popcount(s);
    if (lev != d && lev > max_d)
      continue;
    uint8_t *in = (uint8_t *)calloc(n ? n : 1, 1);
    for (uint32_t k = 0; k < lsky[s]; ++k)
      in[lists[s][k]] = 1;
    for (uint32_t k = 0; k < lext[s]; ++k)
      in[lists[s][lsky[s] + k]] = 2;
    for (uint32_t i = 0; i < n; ++i) {
      if (in[i] != 1)
        set_bit(dom + (size_t)i * W, s);
      if (in[i] == 0)
        set_bit(sdom + (size_t)i * W, s);
    }
    if (sky_n)
      sky_n[s] = lsky[s];
    if (ext_n)
      ext_n[s] = lext[s];
    free(in);
  }
  for (uint32_t s = 0; s < S; ++s)
    free(lists[s]);
  free(lists);
  free(lsky);
  free(lext);
  return ORC_OK;
}

/* ----------------------------------------------------------------------- */
/* STSC's score early exit (src/hybrid/mommyq.cpp:104) skips a same-block,  */
/* same-partition dominator whose FP32 Manhattan norm over s equals p's.   */
/* This counts the (q, p, s) triples among the extended skyline where that */
/* can happen: q dominates p in s (popcount(s) <= max_d) and the ascending- */
/* order FP32 sums over s are equal.  A count of 0 means STSC computes the  */
/* definition on this input.  Pairs whose smallest strict gap exceeds the  */
/* worst rounding of two d-term sums cannot tie and are skipped quickly.   */
uint64_t orc_score_tie_count(const float *rows, int d, int max_d,
                             const uint32_t *ext, uint32_t n_ext) {
  float vmax = 0.0f;
  for (uint32_t a = 0; a < n_ext; ++a)
    for (int i = 0; i < d; ++i) {
      const float v = fabsf(rows[(size_t)ext[a] * d + i]);
      if (v > vmax)
        vmax = v;
    }
  /* |fl(sum) - sum| <= (k-1) u sum|x| for k terms, u = 2^-24; two sums. */
  const double thr = 2.02 * (d - 1) * (double)d * (double)vmax * ldexp(1.0, -24);
  /* points a in blocks of ORC_BLOCK, the candidates b outer: each
   * candidate row is read once per block (from a contiguous copy) */
  float *cand = gather_rows(rows, d, ext, n_ext);
  if (!cand)
    return UINT64_MAX;
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
  for (uint32_t a0 = 0; a0 < n_ext; a0 += ORC_BLOCK) {
    const uint32_t na = n_ext - a0 < ORC_BLOCK ? n_ext - a0 : ORC_BLOCK;
    for (uint32_t b = 0; b < n_ext; ++b) {
      const float *q = cand + (size_t)b * d;
      for (uint32_t k = 0; k < na; ++k) {
        const uint32_t a = a0 + k;
        if (a == b)
          continue;
        const float *p = cand + (size_t)a * d;
        uint32_t lt, eq;
        pair_masks(q, p, d, &lt, &eq);
        if (lt == 0)
          continue;
        double gap = INFINITY;
        for (uint32_t t = lt; t; t &= t - 1) {
          const int i = __builtin_ctz(t);
          const double g = (double)p[i] - (double)q[i];
          if (g < gap)
            gap = g;
        }
        if (gap > thr)
          continue;
        const uint32_t le = lt | eq;
        /* enumerate the subsets s of le that meet lt */
        for (uint32_t s = le; s; s = (s - 1) & le) {
          if (!(s & lt) || __builtin_popcount(s) > max_d)
            continue;
          if (score_of(q, s, d) == score_of(p, s, d))
            total++;
        }
      }
    }
  }
  free(cand);
  return total;
}

/* The triples counted by orc_score_tie_count for which p has no other
 * dominator r in s (among the extended skyline) whose FP32 norm over s is
 * strictly smaller than p's.  STSC scans candidates in norm order, so such a
 * smaller-norm dominator is met before the exit and the tie is harmless; 0
 * means STSC computes the definition on this input.  (If some dominator of p
 * in s exists outside ext(s), one inside ext(s) strictly dominates it and has
 * a norm no larger, so searching the extended skyline suffices.) */
uint64_t orc_score_tie_unresolved(const float *rows, int d, int max_d,
                                  const uint32_t *ext, uint32_t n_ext) {
  float vmax = 0.0f;
  for (uint32_t a = 0; a < n_ext; ++a)
    for (int i = 0; i < d; ++i) {
      const float v = fabsf(rows[(size_t)ext[a] * d + i]);
      if (v > vmax)
        vmax = v;
    }
  const double thr = 2.02 * (d - 1) * (double)d * (double)vmax * ldexp(1.0, -24);
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
  for (uint32_t a = 0; a < n_ext; ++a) {
    const float *p = rows + (size_t)ext[a] * d;
    for (uint32_t b = 0; b < n_ext; ++b) {
      if (a == b)
        continue;
      const float *q = rows + (size_t)ext[b] * d;
      uint32_t lt, eq;
      pair_masks(q, p, d, &lt, &eq);
      if (lt == 0)
        continue;
      double gap = INFINITY;
      for (uint32_t t = lt; t; t &= t - 1) {
        const int i = __builtin_ctz(t);
        const double g = (double)p[i] - (double)q[i];
        if (g < gap)
          gap = g;
      }
      if (gap > thr)
        continue;
      const uint32_t le = lt | eq;
      for (uint32_t s = le; s; s = (s - 1) & le) {
        if (!(s & lt) || __builtin_popcount(s) > max_d)
          continue;
        const float sp = score_of(p, s, d);
        if (score_of(q, s, d) != sp)
          continue;
        int found = 0;
        for (uint32_t c = 0; c < n_ext && !found; ++c) {
          if (c == a)
            continue;
          const float *r = rows + (size_t)ext[c] * d;
          uint32_t rl, re;
          pair_masks(r, p, d, &rl, &re);
          if ((s & ~(rl | re)) == 0 && (s & rl) && score_of(r, s, d) < sp)
            found = 1;
        }
        if (!found)
          total++;
      }
    }
  }
  return total;
}
