/*
 * oracle.c -- CPU restatement of the VRAD radiosity-bake hot path.
 *
 * TEST INFRASTRUCTURE ONLY (checker for tests/, smoke() and bench.py's
 * cpu_baseline leg).  Nothing in the product links or calls it.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math -msse2 -mfpmath=sse
 *        (see oracle/Makefile).  All FP32 expressions below are written with
 *        explicit left-to-right association; with contraction off every
 *        operation rounds once, exactly as on the GPU build.
 *
 * Pinning (what each function is checked against, tests/test_oracle.py):
 *   or_precompute      closed-form triangles (tests/golden/tri_precompute.json),
 *                      edge-equation invariants E.(p,1) in {0,1}
 *   or_trace_brute     analytic axis-aligned rays (exact t), FP64 brute force
 *   or_kd_build        node encode/decode, every triangle reachable, the
 *                      RefineNode cost function on hand-worked 3-4 triangle sets
 *   or_trace_kd        equality with or_trace_brute (bit-exact) on random rays
 *   or_testline        closed-form occluder scenes (box between two points)
 *   or_make_transfers  parallel / perpendicular unit-square form factors
 *                      (Howell catalogue), reciprocity, enclosure row sums
 *   or_gather_light    unoccluded closed form I*cos/d^2, occluded -> 0
 *   or_bounce_*        furnace closed form E*sum(rho^k) (dyadic, exact),
 *                      scipy sparse solve of (I - rho F) B = E on tiny N
 *   or_energy_reduce   exact integer-valued sums, long-double pairwise sum
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small helpers                                                             */
/* ------------------------------------------------------------------------ */
static inline float f_min(float a, float b) { return a < b ? a : b; }
static inline float f_max(float a, float b) { return a > b ? a : b; }

typedef struct {
  void (*fn)(void *ctx, int64_t begin, int64_t end);
  void *ctx;
  int64_t begin, end;
} or_job_t;

static void *or_job_run(void *arg) {
  or_job_t *j = (or_job_t *)arg;
  j->fn(j->ctx, j->begin, j->end);
  return NULL;
}

/* run fn over [0,n) split into nthreads contiguous ranges */
static void or_parallel(int64_t n, int nthreads,
                        void (*fn)(void *, int64_t, int64_t), void *ctx) {
  if (nthreads <= 1 || n < 64) {
    fn(ctx, 0, n);
    return;
  }
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  or_job_t jobs[256];
  int64_t chunk = (n + nthreads - 1) / nthreads;
  int started = 0;
  for (int t = 0; t < nthreads; ++t) {
    int64_t b = (int64_t)t * chunk, e = b + chunk;
    if (b >= n) break;
    if (e > n) e = n;
    jobs[t].fn = fn;
    jobs[t].ctx = ctx;
    jobs[t].begin = b;
    jobs[t].end = e;
    if (pthread_create(&th[t], NULL, or_job_run, &jobs[t]) != 0) {
      fn(ctx, b, e); /* run inline if a thread cannot start */
      jobs[t].fn = NULL;
    }
    started = t + 1;
  }
  for (int t = 0; t < started; ++t)
    if (jobs[t].fn) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------ */
/* ChangeIntoIntersectionFormat: raytracer/cache/optimisedtriangle.go:30-78, */
/* GetEdgeEquation: vmath/polygon/edge.go:5-25.  mgl32 Sub/Cross/Normalize/  */
/* Dot op order (go-gl/mathgl mgl32, pinned 5ab0e04e, Gopkg.lock:65-67).     */
/* ------------------------------------------------------------------------ */
static void or_edge(const float p1[3], const float p2[3], int c0, int c1,
                    const float in[3], float out[3]) {
  float nx = p1[c1] - p2[c1];                   /* edge.go:6 */
  float ny = p2[c0] - p1[c0];                   /* edge.go:7 */
  float d = -(nx * p1[c0] + ny * p1[c1]);       /* edge.go:8 */
  float tr = (in[c0] * nx + in[c1] * ny) + d;   /* edge.go:13 */
  if (tr < 0) {                                 /* edge.go:14-19 */
    nx = -nx;
    ny = -ny;
    d = -d;
    tr = -tr;
  }
  out[0] = nx / tr;                             /* edge.go:20-22 */
  out[1] = ny / tr;
  out[2] = d / tr;
}

void or_precompute(const float *v9, const int32_t *ids, const uint8_t *flags,
                   int64_t n, or_tri_t *out) {
  for (int64_t i = 0; i < n; ++i) {
    const float *v = v9 + 9 * i;
    float p1[3] = {v[0], v[1], v[2]}, p2[3] = {v[3], v[4], v[5]},
          p3[3] = {v[6], v[7], v[8]};
    float e1[3], e2[3], N[3];
    for (int c = 0; c < 3; ++c) {
      e1[c] = p2[c] - p1[c];                    /* optimisedtriangle.go:43 */
      e2[c] = p3[c] - p1[c];                    /* :44 */
    }
    N[0] = e1[1] * e2[2] - e1[2] * e2[1];       /* mgl32 Cross, :46 */
    N[1] = e1[2] * e2[0] - e1[0] * e2[2];
    N[2] = e1[0] * e2[1] - e1[1] * e2[0];
    /* mgl32 Normalize: l = 1/float32(sqrt(float64(x*x+y*y+z*z))), :47 */
    float len2 = (N[0] * N[0] + N[1] * N[1]) + N[2] * N[2];
    float len = (float)sqrt((double)len2);
    float l = 1.0f / len;
    N[0] = N[0] * l;
    N[1] = N[1] * l;
    N[2] = N[2] * l;
    int drop = 0;                               /* :50-55, strict > */
    for (int c = 1; c < 3; ++c)
      if (fabs((double)N[c]) > fabs((double)N[drop])) drop = c;
    or_tri_t *t = out + i;
    memset(t, 0, sizeof *t);
    t->d = (N[0] * p1[0] + N[1] * p1[1]) + N[2] * p1[2]; /* mgl32 Dot, :57 */
    t->nx = N[0];
    t->ny = N[1];
    t->nz = N[2];
    int c0 = (drop + 1) % 3, c1 = (drop + 2) % 3;        /* :63-64 */
    t->c0 = (uint8_t)c0;
    t->c1 = (uint8_t)c1;
    or_edge(p1, p2, c0, c1, p3, t->e + 0);               /* :69-72 */
    or_edge(p2, p3, c0, c1, p1, t->e + 3);               /* :74-77 */
    t->id = ids ? ids[i] : 0;
    t->flags = flags ? flags[i] : 0;
  }
}

/* ------------------------------------------------------------------------ */
/* Per-triangle hit test (frozen; the reference's Trace4Rays is a stub,     */
/* raytracer/environment.go:140-145).  Projected-barycentric test that the   */
/* TriIntersectData layout encodes.  Accepts t in [tmin, tbest) and resolves */
/* equal t to the smaller triangle index, so the result is the lexicographic */
/* minimum of (t, index) and does not depend on visiting order.              */
/* ------------------------------------------------------------------------ */
static inline void or_hit(const or_tri_t *T, int32_t idx, const float o[3],
                          const float d[3], float tmin, float *tbest,
                          int32_t *best) {
  float den = (T->nx * d[0] + T->ny * d[1]) + T->nz * d[2];
  if (!(den != 0.0f)) return;
  float num = T->d - ((T->nx * o[0] + T->ny * o[1]) + T->nz * o[2]);
  float t = num / den;
  if (!(t >= tmin && (t < *tbest || (t == *tbest && idx < *best)))) return;
  float h0 = o[T->c0] + t * d[T->c0];
  float h1 = o[T->c1] + t * d[T->c1];
  float b0 = (T->e[0] * h0 + T->e[1] * h1) + T->e[2];
  float b1 = (T->e[3] * h0 + T->e[4] * h1) + T->e[5];
  if (b0 >= 0.0f && b1 >= 0.0f && (b0 + b1) <= 1.0f) {
    *tbest = t;
    *best = idx;
  }
}

void or_trace_brute(const or_tri_t *tris, int64_t ntri, const or_rays_t *r,
                    int64_t n, int32_t skip_id, int32_t *hit, float *hit_t) {
  for (int64_t i = 0; i < n; ++i) {
    float o[3] = {r->ox[i], r->oy[i], r->oz[i]};
    float d[3] = {r->dx[i], r->dy[i], r->dz[i]};
    float tmin = r->tmin[i], tbest = r->tmax[i];
    int32_t best = -1;
    for (int64_t k = 0; k < ntri; ++k) {
      if (tris[k].id == skip_id) continue;
      or_hit(tris + k, (int32_t)k, o, d, tmin, &tbest, &best);
    }
    hit[i] = best;
    hit_t[i] = best >= 0 ? tbest : 0.0f;
  }
}

/* FP64 version: same construction and test, all in double. */
void or_trace_brute_f64(const float *v9, const int32_t *ids, int64_t ntri,
                        const or_rays_t *r, int64_t n, int32_t skip_id,
                        int32_t *hit, double *hit_t) {
  typedef struct { double n[3], d, e[6]; int c0, c1; } T64;
  T64 *T = (T64 *)malloc(sizeof(T64) * (size_t)(ntri > 0 ? ntri : 1));
  for (int64_t i = 0; i < ntri; ++i) {
    const float *v = v9 + 9 * i;
    double p1[3] = {v[0], v[1], v[2]}, p2[3] = {v[3], v[4], v[5]},
           p3[3] = {v[6], v[7], v[8]}, e1[3], e2[3], N[3];
    for (int c = 0; c < 3; ++c) { e1[c] = p2[c] - p1[c]; e2[c] = p3[c] - p1[c]; }
    N[0] = e1[1] * e2[2] - e1[2] * e2[1];
    N[1] = e1[2] * e2[0] - e1[0] * e2[2];
    N[2] = e1[0] * e2[1] - e1[1] * e2[0];
    double l = 1.0 / sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
    for (int c = 0; c < 3; ++c) N[c] *= l;
    int drop = 0;
    for (int c = 1; c < 3; ++c) if (fabs(N[c]) > fabs(N[drop])) drop = c;
    int c0 = (drop + 1) % 3, c1 = (drop + 2) % 3;
    T[i].c0 = c0; T[i].c1 = c1;
    memcpy(T[i].n, N, sizeof N);
    T[i].d = N[0] * p1[0] + N[1] * p1[1] + N[2] * p1[2];
    const double *P[3] = {p1, p2, p3};
    for (int ed = 0; ed < 2; ++ed) {
      const double *a = P[ed], *b = P[ed + 1], *in = P[(ed + 2) % 3];
      double nx = a[c1] - b[c1], ny = b[c0] - a[c0], d = -(nx * a[c0] + ny * a[c1]);
      double tr = in[c0] * nx + in[c1] * ny + d;
      if (tr < 0) { nx = -nx; ny = -ny; d = -d; tr = -tr; }
      T[i].e[3 * ed + 0] = nx / tr;
      T[i].e[3 * ed + 1] = ny / tr;
      T[i].e[3 * ed + 2] = d / tr;
    }
  }
  for (int64_t i = 0; i < n; ++i) {
    double o[3] = {r->ox[i], r->oy[i], r->oz[i]}, d[3] = {r->dx[i], r->dy[i], r->dz[i]};
    double tmin = r->tmin[i], tbest = r->tmax[i];
    int32_t best = -1;
    for (int64_t k = 0; k < ntri; ++k) {
      if (ids && ids[k] == skip_id) continue;
      const T64 *t = T + k;
      double den = t->n[0] * d[0] + t->n[1] * d[1] + t->n[2] * d[2];
      if (!(den != 0.0)) continue;
      double tt = (t->d - (t->n[0] * o[0] + t->n[1] * o[1] + t->n[2] * o[2])) / den;
      if (!(tt >= tmin && (tt < tbest || (tt == tbest && k < best)))) continue;
      double h0 = o[t->c0] + tt * d[t->c0], h1 = o[t->c1] + tt * d[t->c1];
      double b0 = t->e[0] * h0 + t->e[1] * h1 + t->e[2];
      double b1 = t->e[3] * h0 + t->e[4] * h1 + t->e[5];
      if (b0 >= 0 && b1 >= 0 && b0 + b1 <= 1) { tbest = tt; best = (int32_t)k; }
    }
    hit[i] = best;
    hit_t[i] = best >= 0 ? tbest : 0.0;
  }
  free(T);
}

/* ------------------------------------------------------------------------ */
/* SAH kd-tree, following the intent of RefineNode / CalculateCostsOfSplit   */
/* (raytracer/environment.go:181-387) with the documented defects repaired:  */
/* real triangle bounds (trianglelist.go:9-21 never writes its outputs),     */
/* real vertices (optimisedtriangle.go:22-28 returns one coordinate thrice), */
/* midpoint 0.5*(min+max) (environment.go:274), split value and counts      */
/* returned from the cost function (environment.go:181-236 takes them by    */
/* value), and the partition writes before incrementing (:347-349).         */
/* ------------------------------------------------------------------------ */
enum { PC_POS = 1, PC_NEG = -1, PC_STRAD = 0 };

typedef struct {
  const float *v9;
  float *tmin3, *tmax3; /* per-triangle per-axis vertex extents */
  uint32_t *children;
  float *split;
  int64_t n_nodes, cap_nodes;
  int32_t *idx;
  int64_t n_idx, cap_idx;
  int max_depth;
  signed char *lab;     /* per-triangle scratch label */
} or_kdb_t;

static int kdb_push_nodes(or_kdb_t *b, int k) {
  if (b->n_nodes + k > b->cap_nodes) {
    int64_t nc = b->cap_nodes ? b->cap_nodes * 2 : 1024;
    while (nc < b->n_nodes + k) nc *= 2;
    uint32_t *c = (uint32_t *)realloc(b->children, sizeof(uint32_t) * nc);
    if (!c) return -1;
    b->children = c;
    float *s = (float *)realloc(b->split, sizeof(float) * nc);
    if (!s) return -1;
    b->split = s;
    b->cap_nodes = nc;
  }
  for (int i = 0; i < k; ++i) {
    b->children[b->n_nodes + i] = 0;
    b->split[b->n_nodes + i] = 0.0f;
  }
  b->n_nodes += k;
  return 0;
}

/* ClassifyAgainstAxisSplit, optimisedtriangle.go:83-103 (real vertices) */
static inline int kdb_classify(const or_kdb_t *b, int32_t t, int axis, float sv) {
  float mn = b->tmin3[3 * t + axis], mx = b->tmax3[3 * t + axis];
  if (mn >= sv) return PC_POS;
  if (mx <= sv) return PC_NEG;
  if (mn == mx) return PC_POS;
  return PC_STRAD;
}

/* BoxSurfaceArea, vmath/polygon/surface.go:5-8 */
static inline float box_sa(const float mn[3], const float mx[3]) {
  float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
  return 2.0f * ((d0 * d2 + d0 * d1) + d1 * d2);
}

/* CalculateCostsOfSplit, environment.go:181-236 */
static float kdb_cost(const or_kdb_t *b, int axis, const int32_t *list, int64_t n,
                      const float mn[3], const float mx[3], float *split_value,
                      int64_t *nl, int64_t *nr, int64_t *nb) {
  int64_t l = 0, r = 0, both = 0;
  double minc = 1.0e23, maxc = -1.0e23;
  float sv = *split_value;
  for (int64_t i = 0; i < n; ++i) {
    int32_t t = list[i];
    double a = b->tmin3[3 * t + axis], c = b->tmax3[3 * t + axis];
    if (a < minc) minc = a;
    if (c > maxc) maxc = c;
    switch (kdb_classify(b, t, axis, sv)) {
      case PC_NEG: ++l; break;
      case PC_POS: ++r; break;
      default: ++both; break;
    }
  }
  if (l != 0 && both == 0 && r == 0) sv = (float)maxc; /* :216-218 */
  if (r != 0 && both == 0 && l == 0) sv = (float)minc; /* :219-221 */
  float lmx[3] = {mx[0], mx[1], mx[2]}, rmn[3] = {mn[0], mn[1], mn[2]};
  lmx[axis] = sv;
  rmn[axis] = sv;
  float sal = box_sa(mn, lmx), sar = box_sa(rmn, mx);
  float isa = 1.0f / box_sa(mn, mx);
  float cost = (float)OR_COST_OF_TRAVERSAL +
               (float)OR_COST_OF_INTERSECTION *
                   (((float)both + (sal * isa) * (float)l) + (sar * isa) * (float)r);
  *split_value = sv;
  *nl = l;
  *nr = r;
  *nb = both;
  return cost;
}

static int kdb_leaf(or_kdb_t *b, int64_t node, const int32_t *list, int64_t n,
                    int depth) {
  b->children[node] = 3u + ((uint32_t)b->n_idx << 2);  /* :243 / :313 */
  b->split[node] = (float)n;                           /* :244 / :314 */
  if (b->n_idx + n > b->cap_idx) {
    int64_t nc = b->cap_idx ? b->cap_idx * 2 : 4096;
    while (nc < b->n_idx + n) nc *= 2;
    int32_t *p = (int32_t *)realloc(b->idx, sizeof(int32_t) * nc);
    if (!p) return -1;
    b->idx = p;
    b->cap_idx = nc;
  }
  memcpy(b->idx + b->n_idx, list, sizeof(int32_t) * (size_t)n);
  b->n_idx += n;
  if (depth > b->max_depth) b->max_depth = depth;
  return 0;
}

/* RefineNode, environment.go:238-387.  `level` is the true tree depth;
 * `depth` is the reference's counter that jumps by 100. */
static int kdb_refine(or_kdb_t *b, int64_t node, const int32_t *list, int64_t n,
                      const float mn[3], const float mx[3], int depth, int level) {
  if (n < 3 || depth > OR_MAX_TREE_DEPTH)  /* :241, :311 (search result unused) */
    return kdb_leaf(b, node, list, n, level);
  float best_cost = 1.0e23f, best_moved = 0.0f, best_trial = 0.0f;
  int64_t bnl = 0, bnr = 0, bnb = 0;
  int best_axis = 0;
  int64_t skip = 1 + n / 10;                                     /* :265 */
  for (int axis = 0; axis < 3; ++axis) {
    for (int64_t ts = -1; ts < n; ts += skip) {
      for (int tv = 0; tv < 3; ++tv) {
        float trial;
        if (ts == -1) {
          trial = 0.5f * (mn[axis] + mx[axis]);
        } else {
          trial = b->v9[9 * (int64_t)list[ts] + 3 * tv + axis];
          if (trial > mx[axis] || trial < mn[axis]) continue;    /* :279-281 */
        }
        float sv = trial;
        int64_t nl, nr, nb;
        float cost = kdb_cost(b, axis, list, n, mn, mx, &sv, &nl, &nr, &nb);
        if (cost < best_cost) {                                  /* :288-300 */
          best_axis = axis;
          best_cost = cost;
          bnl = nl; bnr = nr; bnb = nb;
          best_moved = sv;
          best_trial = trial;
        }
        if (ts == -1) break;
      }
    }
  }
  float no_split = (float)((int64_t)OR_COST_OF_INTERSECTION * n); /* :310 */
  if (no_split <= best_cost) return kdb_leaf(b, node, list, n, level);
  /* labels are those of the winning trial value (saved at :296-299) */
  int32_t *nlist = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
  if (!nlist) return -1;
  int64_t lo = 0, ro = 0, bo = 0;
  for (int64_t i = 0; i < n; ++i) {                              /* :343-359 */
    int32_t t = list[i];
    switch (kdb_classify(b, t, best_axis, best_trial)) {
      case PC_NEG: nlist[lo++] = t; break;
      case PC_POS: ++ro; nlist[n - ro] = t; break;
      default: nlist[bnl + bo] = t; ++bo; break;
    }
  }
  int64_t left = b->n_nodes;                                     /* :361-376 */
  if (kdb_push_nodes(b, 2)) { free(nlist); return -1; }
  b->children[node] = (uint32_t)best_axis + ((uint32_t)left << 2);
  b->split[node] = best_moved;
  if (n < 20 && (bnl == 0 || bnr == 0)) depth += 100;            /* :378-380 */
  float lmx[3] = {mx[0], mx[1], mx[2]}, rmn[3] = {mn[0], mn[1], mn[2]};
  lmx[best_axis] = best_moved;
  rmn[best_axis] = best_moved;
  int rc = kdb_refine(b, left, nlist, bnl + bnb, mn, lmx, depth + 1, level + 1);
  if (!rc)
    rc = kdb_refine(b, left + 1, nlist + bnl, bnr + bnb, rmn, mx, depth + 1,
                    level + 1);
  free(nlist);
  return rc;
}

int or_kd_build(const float *v9, int64_t ntri, or_kdtree_t *out) {
  memset(out, 0, sizeof *out);
  or_kdb_t b;
  memset(&b, 0, sizeof b);
  b.v9 = v9;
  int64_t m = ntri > 0 ? ntri : 1;
  b.tmin3 = (float *)malloc(sizeof(float) * 3 * m);
  b.tmax3 = (float *)malloc(sizeof(float) * 3 * m);
  int32_t *root = (int32_t *)malloc(sizeof(int32_t) * m);
  if (!b.tmin3 || !b.tmax3 || !root) return -1;
  float mn[3] = {1.0e23f, 1.0e23f, 1.0e23f}, mx[3] = {-1.0e23f, -1.0e23f, -1.0e23f};
  for (int64_t t = 0; t < ntri; ++t) {
    root[t] = (int32_t)t;
    for (int c = 0; c < 3; ++c) {
      float a = v9[9 * t + c], bb = v9[9 * t + 3 + c], cc = v9[9 * t + 6 + c];
      float lo = f_min(a, f_min(bb, cc)), hi = f_max(a, f_max(bb, cc));
      b.tmin3[3 * t + c] = lo;
      b.tmax3[3 * t + c] = hi;
      if (lo < mn[c]) mn[c] = lo;
      if (hi > mx[c]) mx[c] = hi;
    }
  }
  if (ntri == 0)
    for (int c = 0; c < 3; ++c) mn[c] = mx[c] = 0.0f;
  int rc = kdb_push_nodes(&b, 1);
  if (!rc) rc = kdb_refine(&b, 0, root, ntri, mn, mx, 0, 0);
  free(root);
  free(b.tmin3);
  free(b.tmax3);
  if (rc) {
    free(b.children); free(b.split); free(b.idx);
    return rc;
  }
  out->n_nodes = b.n_nodes;
  out->children = b.children;
  out->split = b.split;
  out->n_idx = b.n_idx;
  out->idx = b.idx ? b.idx : (int32_t *)malloc(sizeof(int32_t));
  memcpy(out->bmin, mn, sizeof mn);
  memcpy(out->bmax, mx, sizeof mx);
  out->max_depth = b.max_depth;
  return 0;
}

void or_kd_free(or_kdtree_t *t) {
  free(t->children);
  free(t->split);
  free(t->idx);
  memset(t, 0, sizeof *t);
}

/* ------------------------------------------------------------------------ */
/* kd traversal with a conservative slab margin.  Every decision only culls: */
/* a cell is visited over the ray interval where the ray lies within dl of   */
/* it, dl = 2^-15 * (scene max |coord| + max |origin|), which exceeds the    */
/* rounding of the hit test and of the interval arithmetic by > 100x.  Leaf  */
/* tests use the global [tmin, tbest) window, so results equal brute force.  */
/* ------------------------------------------------------------------------ */
static float kd_scene_maxabs(const or_kdtree_t *kd) {
  float m = 1.0f;
  for (int c = 0; c < 3; ++c) {
    m = f_max(m, fabsf(kd->bmin[c]));
    m = f_max(m, fabsf(kd->bmax[c]));
  }
  return m;
}

#define OR_KD_STACK 64
/* nv / nt (nullable) count node visits and triangle tests (the work measure
 * of a traversal; not part of any result) */
static void kd_trace_one(const or_kdtree_t *kd, float M, const or_tri_t *tris,
                         const float o[3], const float d[3], float tmin, float tmax,
                         int32_t skip, int anyhit, int32_t *best_out, float *t_out,
                         int64_t *nv, int64_t *nt) {
  float tbest = tmax;
  int32_t best = -1;
  float om = f_max(fabsf(o[0]), f_max(fabsf(o[1]), fabsf(o[2])));
  float dl = (M + om) * 0x1p-15f;
  float inv[3];
  for (int c = 0; c < 3; ++c) {
    float dd = fabsf(d[c]) < 1e-20f ? copysignf(1e-20f, d[c]) : d[c];
    inv[c] = 1.0f / dd;
  }
  float ta = tmin, tb = tmax;
  for (int c = 0; c < 3; ++c) {
    float t1 = ((kd->bmin[c] - dl) - o[c]) * inv[c];
    float t2 = ((kd->bmax[c] + dl) - o[c]) * inv[c];
    float lo = t1 < t2 ? t1 : t2, hi = t1 < t2 ? t2 : t1;
    ta = ta > lo ? ta : lo;
    tb = tb < hi ? tb : hi;
  }
  if (!(ta <= tb) || kd->n_nodes == 0) {
    *best_out = -1;
    *t_out = 0.0f;
    return;
  }
  uint32_t st_node[OR_KD_STACK];
  float st_a[OR_KD_STACK], st_b[OR_KD_STACK];
  int sp = 0;
  uint32_t node = 0;
  for (;;) {
    uint32_t ch = kd->children[node];
    uint32_t type = ch & 3u;
    if (nv) ++*nv;
    if (type != 3u) {
      float sv = kd->split[node];
      float lo = ((sv - dl) - o[type]) * inv[type];
      float hi = ((sv + dl) - o[type]) * inv[type];
      uint32_t left = ch >> 2, near_n, far_n;
      float near_end, far_start;
      if (inv[type] >= 0.0f) { near_n = left; far_n = left + 1; near_end = hi; far_start = lo; }
      else { near_n = left + 1; far_n = left; near_end = lo; far_start = hi; }
      float tbe = tb < tbest ? tb : tbest;
      float ne = near_end < tbe ? near_end : tbe;
      float fs = far_start > ta ? far_start : ta;
      int near_ok = ta <= ne, far_ok = fs <= tbe;
      if (near_ok && far_ok && sp < OR_KD_STACK) {
        st_node[sp] = far_n; st_a[sp] = fs; st_b[sp] = tb; ++sp;
        node = near_n; tb = near_end < tb ? near_end : tb;
        continue;
      }
      if (near_ok && far_ok) { /* stack full (cannot happen for depth<=64) */
        node = near_n; tb = near_end < tb ? near_end : tb;
        continue;
      }
      if (near_ok) { node = near_n; tb = near_end < tb ? near_end : tb; continue; }
      if (far_ok) { node = far_n; ta = fs; continue; }
    } else {
      uint32_t start = ch >> 2, cnt = (uint32_t)kd->split[node];
      for (uint32_t k = start; k < start + cnt; ++k) {
        int32_t t = kd->idx[k];
        if (tris[t].id == skip) continue;
        if (nt) ++*nt;
        or_hit(tris + t, t, o, d, tmin, &tbest, &best);
      }
      if (anyhit && best >= 0) break;
    }
    /* pop */
    int found = 0;
    while (sp > 0) {
      --sp;
      if (st_a[sp] > tbest) continue;
      node = st_node[sp]; ta = st_a[sp]; tb = st_b[sp];
      found = 1;
      break;
    }
    if (!found) break;
  }
  *best_out = best;
  *t_out = best >= 0 ? tbest : 0.0f;
}

typedef struct {
  const or_kdtree_t *kd;
  const or_tri_t *tris;
  const or_rays_t *r;
  int32_t skip;
  int32_t *hit;
  float *hit_t;
  float M;
} kd_ctx_t;

static void kd_trace_range(void *vc, int64_t b, int64_t e) {
  kd_ctx_t *c = (kd_ctx_t *)vc;
  for (int64_t i = b; i < e; ++i) {
    float o[3] = {c->r->ox[i], c->r->oy[i], c->r->oz[i]};
    float d[3] = {c->r->dx[i], c->r->dy[i], c->r->dz[i]};
    kd_trace_one(c->kd, c->M, c->tris, o, d, c->r->tmin[i], c->r->tmax[i], c->skip,
                 0, c->hit + i, c->hit_t + i, NULL, NULL);
  }
}

void or_trace_kd(const or_kdtree_t *kd, const or_tri_t *tris, const or_rays_t *r,
                 int64_t n, int32_t skip_id, int32_t *hit, float *hit_t,
                 int nthreads) {
  kd_ctx_t c = {kd, tris, r, skip_id, hit, hit_t, kd_scene_maxabs(kd)};
  or_parallel(n, nthreads, kd_trace_range, &c);
}

/* ------------------------------------------------------------------------ */
/* TestLineDoesHitSky intent, raytracer/trace/testline.go:22-51,91-93:       */
/*   dir = (stop-start) * (1/|stop-start|), trace [0, len), occluded iff the */
/*   closest hit is not a TRACE_ID_SKY triangle.  The 3D-skybox recursion    */
/*   (:60-89) and texture coverage (:53-55) are off by default and absent.  */
/* ------------------------------------------------------------------------ */
static void testline_one(const or_kdtree_t *kd, float M, const or_tri_t *tris,
                         int64_t ntri, const float s[3], const float e[3],
                         int32_t skip, uint8_t *vis, uint8_t *sky, int64_t *nv,
                         int64_t *nt) {
  float dv[3] = {e[0] - s[0], e[1] - s[1], e[2] - s[2]};
  float len2 = (dv[0] * dv[0] + dv[1] * dv[1]) + dv[2] * dv[2];
  float len = sqrtf(len2);
  float inv = 1.0f / len;
  float d[3] = {dv[0] * inv, dv[1] * inv, dv[2] * inv};
  int32_t best = -1;
  float tb = 0.0f;
  if (kd) {
    kd_trace_one(kd, M, tris, s, d, 0.0f, len, skip, 0, &best, &tb, nv, nt);
  } else {
    float tbest = len;
    for (int64_t k = 0; k < ntri; ++k) {
      if (tris[k].id == skip) continue;
      or_hit(tris + k, (int32_t)k, s, d, 0.0f, &tbest, &best);
    }
    tb = tbest;
  }
  int hit = best >= 0 && tb < len;
  int is_sky = hit && (tris[best].id & OR_TRACE_ID_SKY) != 0;
  *vis = (uint8_t)!(hit && !is_sky);
  if (sky) *sky = (uint8_t)is_sky;
}

typedef struct {
  const or_kdtree_t *kd;
  const or_tri_t *tris;
  int64_t ntri;
  const or_segs_t *s;
  int32_t skip;
  uint8_t *vis, *sky;
  float M;
  int64_t *visits, *tests; /* nullable: totals of the work measure */
} tl_ctx_t;

static void tl_range(void *vc, int64_t b, int64_t e) {
  tl_ctx_t *c = (tl_ctx_t *)vc;
  int64_t nv = 0, nt = 0;
  for (int64_t i = b; i < e; ++i) {
    float s[3] = {c->s->sx[i], c->s->sy[i], c->s->sz[i]};
    float en[3] = {c->s->ex[i], c->s->ey[i], c->s->ez[i]};
    testline_one(c->kd, c->M, c->tris, c->ntri, s, en, c->skip, c->vis + i,
                 c->sky ? c->sky + i : NULL, c->visits ? &nv : NULL, c->tests ? &nt : NULL);
  }
  if (c->visits) __atomic_fetch_add(c->visits, nv, __ATOMIC_RELAXED);
  if (c->tests) __atomic_fetch_add(c->tests, nt, __ATOMIC_RELAXED);
}

void or_testline(const or_kdtree_t *kd, const or_tri_t *tris, int64_t ntri,
                 const or_segs_t *s, int64_t n, int32_t skip_id, uint8_t *visible,
                 uint8_t *hit_sky, int nthreads) {
  tl_ctx_t c = {kd, tris, ntri, s, skip_id, visible, hit_sky,
                kd ? kd_scene_maxabs(kd) : 1.0f, NULL, NULL};
  or_parallel(n, nthreads, tl_range, &c);
}

void or_testline_stats(const or_kdtree_t *kd, const or_tri_t *tris, const or_segs_t *s,
                       int64_t n, int32_t skip_id, uint8_t *visible, int nthreads,
                       int64_t *visits, int64_t *tests) {
  *visits = 0;
  *tests = 0;
  tl_ctx_t c = {kd, tris, 0, s, skip_id, visible, NULL, kd_scene_maxabs(kd), visits, tests};
  or_parallel(n, nthreads, tl_range, &c);
}

/* ------------------------------------------------------------------------ */
/* MakeTransfers (absent in the reference; output record Transfer,          */
/* common/types/transfer.go:3-6, Patch.NumTransfers/Transfers,              */
/* common/types/patch.go:60-61).  Frozen semantics:                         */
/*   unordered pair i<j, facing cull n_i.(o_j-o_i) > 0 && -(n_j.(o_j-o_i))>0 */
/*   visibility TestLine(o_i + eps n_i -> o_j + eps n_j) from the lower index */
/*   F_ij = ((di*dj)*A_j) / ((pi_f*r2)*r2)                                   */
/*   rows ascending in column; FP64 row sum in kernel wave order; rows with   */
/*   sum > 1 scaled by 1/sum in FP64 when normalize != 0.                    */
/* ------------------------------------------------------------------------ */
#define OR_PI_F 3.14159265358979323846f

double or_row_sum_f64(const float *coef, int64_t len) {
  /* lane l accumulates positions p with (p/4) % 64 == l, then butterfly */
  double acc[64];
  for (int l = 0; l < 64; ++l) acc[l] = 0.0;
  for (int64_t p = 0; p < len; ++p) {
    int l = (int)((p / 4) % 64);
    acc[l] = acc[l] + (double)coef[p];
  }
  for (int off = 32; off >= 1; off >>= 1) {
    double nw[64];
    for (int l = 0; l < 64; ++l) nw[l] = acc[l] + acc[l ^ off];
    memcpy(acc, nw, sizeof acc);
  }
  return acc[0];
}

typedef struct {
  const or_kdtree_t *kd;
  const or_tri_t *tris;
  int64_t ntri;
  const or_patches_t *p;
  int64_t np;
  float eps;
  int32_t skip;
  uint32_t *bits; /* np x W words, symmetric */
  int64_t W;
  float M;
  const int32_t *cl;   /* PVS: cluster per patch, NULL = every pair */
  const uint8_t *pvs;  /* n_clusters rows of (n_clusters + 7) >> 3 bytes */
  int64_t pvs_row;
} mt_ctx_t;

/* PVSCheck (rad/lightmap/lightmap.go:413-422) on the row of cluster a; the
 * row of a negative cluster is all ones (GetVisCache, rad/lightmap/vis.go:
 * 11-32) and a negative target passes. */
static int pvs_check(const mt_ctx_t *c, int32_t a, int32_t b) {
  if (a < 0 || b < 0) return 1;
  return (c->pvs[(int64_t)a * c->pvs_row + (b >> 3)] >> (b & 7)) & 1;
}

static void mt_range(void *vc, int64_t b, int64_t e) {
  mt_ctx_t *c = (mt_ctx_t *)vc;
  const or_patches_t *p = c->p;
  for (int64_t i = b; i < e; ++i) {
    for (int64_t j = i + 1; j < c->np; ++j) {
      if (c->cl && !pvs_check(c, c->cl[i], c->cl[j]) && !pvs_check(c, c->cl[j], c->cl[i]))
        continue;
      float dx = p->ox[j] - p->ox[i], dy = p->oy[j] - p->oy[i], dz = p->oz[j] - p->oz[i];
      float di = (p->nx[i] * dx + p->ny[i] * dy) + p->nz[i] * dz;
      float dj = -((p->nx[j] * dx + p->ny[j] * dy) + p->nz[j] * dz);
      if (!(di > 0.0f && dj > 0.0f)) continue;
      float s[3] = {p->ox[i] + c->eps * p->nx[i], p->oy[i] + c->eps * p->ny[i],
                    p->oz[i] + c->eps * p->nz[i]};
      float en[3] = {p->ox[j] + c->eps * p->nx[j], p->oy[j] + c->eps * p->ny[j],
                     p->oz[j] + c->eps * p->nz[j]};
      uint8_t vis;
      testline_one(c->kd, c->M, c->tris, c->ntri, s, en, c->skip, &vis, NULL, NULL, NULL);
      if (vis) {
        /* row i written only by this thread; (j,i) bit set under a lock-free
         * atomic since rows j belong to other threads */
        __atomic_fetch_or(&c->bits[i * c->W + (j >> 5)], 1u << (j & 31), __ATOMIC_RELAXED);
        __atomic_fetch_or(&c->bits[j * c->W + (i >> 5)], 1u << (i & 31), __ATOMIC_RELAXED);
      }
    }
  }
}

static inline float or_ff(const or_patches_t *p, int64_t i, int64_t j) {
  float dx = p->ox[j] - p->ox[i], dy = p->oy[j] - p->oy[i], dz = p->oz[j] - p->oz[i];
  float di = (p->nx[i] * dx + p->ny[i] * dy) + p->nz[i] * dz;
  float dj = -((p->nx[j] * dx + p->ny[j] * dy) + p->nz[j] * dz);
  float r2 = (dx * dx + dy * dy) + dz * dz;
  return ((di * dj) * p->area[j]) / ((OR_PI_F * r2) * r2);
}

int or_make_transfers(const or_kdtree_t *kd, const or_tri_t *tris, int64_t ntri,
                      const or_patches_t *p, int64_t np, float normal_offset,
                      int32_t skip_id, int normalize, int nthreads, or_csr_t *out) {
  return or_make_transfers_pvs(kd, tris, ntri, p, np, normal_offset, skip_id, normalize, nthreads,
                               NULL, NULL, 0, out);
}

int or_make_transfers_pvs(const or_kdtree_t *kd, const or_tri_t *tris, int64_t ntri,
                          const or_patches_t *p, int64_t np, float normal_offset,
                          int32_t skip_id, int normalize, int nthreads, const int32_t *cluster,
                          const uint8_t *pvs_rows, int32_t n_clusters, or_csr_t *out) {
  memset(out, 0, sizeof *out);
  int64_t W = (np + 31) / 32;
  uint32_t *bits = (uint32_t *)calloc((size_t)(np * W > 0 ? np * W : 1), sizeof(uint32_t));
  if (!bits) return -1;
  mt_ctx_t c = {kd, tris, ntri, p, np, normal_offset, skip_id, bits, W,
                kd ? kd_scene_maxabs(kd) : 1.0f, n_clusters > 0 ? cluster : NULL, pvs_rows,
                ((int64_t)n_clusters + 7) >> 3};
  /* rows have unequal work (j > i); interleave by splitting into many ranges */
  int64_t chunks = np;
  (void)chunks;
  or_parallel(np, nthreads, mt_range, &c);
  int64_t *rp = (int64_t *)malloc(sizeof(int64_t) * (size_t)(np + 1));
  if (!rp) { free(bits); return -1; }
  rp[0] = 0;
  for (int64_t i = 0; i < np; ++i) {
    int64_t cnt = 0;
    for (int64_t w = 0; w < W; ++w) cnt += __builtin_popcount(bits[i * W + w]);
    rp[i + 1] = rp[i] + cnt;
  }
  int64_t nnz = rp[np];
  int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
  float *coef = (float *)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
  double *rs = (double *)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
  if (!col || !coef || !rs) { free(bits); free(rp); free(col); free(coef); free(rs); return -1; }
  for (int64_t i = 0; i < np; ++i) {
    int64_t k = rp[i];
    for (int64_t w = 0; w < W; ++w) {
      uint32_t m = bits[i * W + w];
      while (m) {
        int bit = __builtin_ctz(m);
        m &= m - 1;
        int64_t j = w * 32 + bit;
        col[k] = (int32_t)j;
        coef[k] = or_ff(p, i, j);
        ++k;
      }
    }
    int64_t len = rp[i + 1] - rp[i];
    double s = or_row_sum_f64(coef + rp[i], len);
    rs[i] = s;
    if (normalize && s > 1.0)
      for (int64_t q = rp[i]; q < rp[i + 1]; ++q) coef[q] = (float)((double)coef[q] / s);
  }
  free(bits);
  out->n_rows = np;
  out->nnz = nnz;
  out->row_ptr = rp;
  out->col = col;
  out->coef = coef;
  out->row_sum = rs;
  return 0;
}

void or_csr_free(or_csr_t *c) {
  free(c->row_ptr);
  free(c->col);
  free(c->coef);
  free(c->row_sum);
  memset(c, 0, sizeof *c);
}

/* ------------------------------------------------------------------------ */
/* GatherSampleLight (absent; light records common/types/light.go:10-44,     */
/* falloff parameters rad/lightmap/lights.go:259-341).  Frozen semantics in  */
/* gather_one below; lights are summed in list order.                       */
/* ------------------------------------------------------------------------ */
/* The reference's Anorms table (vmath/constants.go:15-184), 162 float32
 * vectors, in table order: the directions CanLeafTraceToSky walks
 * (rad/lightmap/lightmap.go:435-448).  Literals round to nearest float32,
 * as Go's conversion of an untyped constant to float32 does. */
static const float or_anorms[162][3] = {
    {-0.525731f, 0.000000f, 0.850651f}, {-0.442863f, 0.238856f, 0.864188f}, {-0.295242f, 0.000000f, 0.955423f},
    {-0.309017f, 0.500000f, 0.809017f}, {-0.162460f, 0.262866f, 0.951056f}, {0.000000f, 0.000000f, 1.000000f},
    {0.000000f, 0.850651f, 0.525731f}, {-0.147621f, 0.716567f, 0.681718f}, {0.147621f, 0.716567f, 0.681718f},
    {0.000000f, 0.525731f, 0.850651f}, {0.309017f, 0.500000f, 0.809017f}, {0.525731f, 0.000000f, 0.850651f},
    {0.295242f, 0.000000f, 0.955423f}, {0.442863f, 0.238856f, 0.864188f}, {0.162460f, 0.262866f, 0.951056f},
    {-0.681718f, 0.147621f, 0.716567f}, {-0.809017f, 0.309017f, 0.500000f}, {-0.587785f, 0.425325f, 0.688191f},
    {-0.850651f, 0.525731f, 0.000000f}, {-0.864188f, 0.442863f, 0.238856f}, {-0.716567f, 0.681718f, 0.147621f},
    {-0.688191f, 0.587785f, 0.425325f}, {-0.500000f, 0.809017f, 0.309017f}, {-0.238856f, 0.864188f, 0.442863f},
    {-0.425325f, 0.688191f, 0.587785f}, {-0.716567f, 0.681718f, -0.147621f}, {-0.500000f, 0.809017f, -0.309017f},
    {-0.525731f, 0.850651f, 0.000000f}, {0.000000f, 0.850651f, -0.525731f}, {-0.238856f, 0.864188f, -0.442863f},
    {0.000000f, 0.955423f, -0.295242f}, {-0.262866f, 0.951056f, -0.162460f}, {0.000000f, 1.000000f, 0.000000f},
    {0.000000f, 0.955423f, 0.295242f}, {-0.262866f, 0.951056f, 0.162460f}, {0.238856f, 0.864188f, 0.442863f},
    {0.262866f, 0.951056f, 0.162460f}, {0.500000f, 0.809017f, 0.309017f}, {0.238856f, 0.864188f, -0.442863f},
    {0.262866f, 0.951056f, -0.162460f}, {0.500000f, 0.809017f, -0.309017f}, {0.850651f, 0.525731f, 0.000000f},
    {0.716567f, 0.681718f, 0.147621f}, {0.716567f, 0.681718f, -0.147621f}, {0.525731f, 0.850651f, 0.000000f},
    {0.425325f, 0.688191f, 0.587785f}, {0.864188f, 0.442863f, 0.238856f}, {0.688191f, 0.587785f, 0.425325f},
    {0.809017f, 0.309017f, 0.500000f}, {0.681718f, 0.147621f, 0.716567f}, {0.587785f, 0.425325f, 0.688191f},
    {0.955423f, 0.295242f, 0.000000f}, {1.000000f, 0.000000f, 0.000000f}, {0.951056f, 0.162460f, 0.262866f},
    {0.850651f, -0.525731f, 0.000000f}, {0.955423f, -0.295242f, 0.000000f}, {0.864188f, -0.442863f, 0.238856f},
    {0.951056f, -0.162460f, 0.262866f}, {0.809017f, -0.309017f, 0.500000f}, {0.681718f, -0.147621f, 0.716567f},
    {0.850651f, 0.000000f, 0.525731f}, {0.864188f, 0.442863f, -0.238856f}, {0.809017f, 0.309017f, -0.500000f},
    {0.951056f, 0.162460f, -0.262866f}, {0.525731f, 0.000000f, -0.850651f}, {0.681718f, 0.147621f, -0.716567f},
    {0.681718f, -0.147621f, -0.716567f}, {0.850651f, 0.000000f, -0.525731f}, {0.809017f, -0.309017f, -0.500000f},
    {0.864188f, -0.442863f, -0.238856f}, {0.951056f, -0.162460f, -0.262866f}, {0.147621f, 0.716567f, -0.681718f},
    {0.309017f, 0.500000f, -0.809017f}, {0.425325f, 0.688191f, -0.587785f}, {0.442863f, 0.238856f, -0.864188f},
    {0.587785f, 0.425325f, -0.688191f}, {0.688191f, 0.587785f, -0.425325f}, {-0.147621f, 0.716567f, -0.681718f},
    {-0.309017f, 0.500000f, -0.809017f}, {0.000000f, 0.525731f, -0.850651f}, {-0.525731f, 0.000000f, -0.850651f},
    {-0.442863f, 0.238856f, -0.864188f}, {-0.295242f, 0.000000f, -0.955423f}, {-0.162460f, 0.262866f, -0.951056f},
    {0.000000f, 0.000000f, -1.000000f}, {0.295242f, 0.000000f, -0.955423f}, {0.162460f, 0.262866f, -0.951056f},
    {-0.442863f, -0.238856f, -0.864188f}, {-0.309017f, -0.500000f, -0.809017f}, {-0.162460f, -0.262866f, -0.951056f},
    {0.000000f, -0.850651f, -0.525731f}, {-0.147621f, -0.716567f, -0.681718f}, {0.147621f, -0.716567f, -0.681718f},
    {0.000000f, -0.525731f, -0.850651f}, {0.309017f, -0.500000f, -0.809017f}, {0.442863f, -0.238856f, -0.864188f},
    {0.162460f, -0.262866f, -0.951056f}, {0.238856f, -0.864188f, -0.442863f}, {0.500000f, -0.809017f, -0.309017f},
    {0.425325f, -0.688191f, -0.587785f}, {0.716567f, -0.681718f, -0.147621f}, {0.688191f, -0.587785f, -0.425325f},
    {0.587785f, -0.425325f, -0.688191f}, {0.000000f, -0.955423f, -0.295242f}, {0.000000f, -1.000000f, 0.000000f},
    {0.262866f, -0.951056f, -0.162460f}, {0.000000f, -0.850651f, 0.525731f}, {0.000000f, -0.955423f, 0.295242f},
    {0.238856f, -0.864188f, 0.442863f}, {0.262866f, -0.951056f, 0.162460f}, {0.500000f, -0.809017f, 0.309017f},
    {0.716567f, -0.681718f, 0.147621f}, {0.525731f, -0.850651f, 0.000000f}, {-0.238856f, -0.864188f, -0.442863f},
    {-0.500000f, -0.809017f, -0.309017f}, {-0.262866f, -0.951056f, -0.162460f}, {-0.850651f, -0.525731f, 0.000000f},
    {-0.716567f, -0.681718f, -0.147621f}, {-0.716567f, -0.681718f, 0.147621f}, {-0.525731f, -0.850651f, 0.000000f},
    {-0.500000f, -0.809017f, 0.309017f}, {-0.238856f, -0.864188f, 0.442863f}, {-0.262866f, -0.951056f, 0.162460f},
    {-0.864188f, -0.442863f, 0.238856f}, {-0.809017f, -0.309017f, 0.500000f}, {-0.688191f, -0.587785f, 0.425325f},
    {-0.681718f, -0.147621f, 0.716567f}, {-0.442863f, -0.238856f, 0.864188f}, {-0.587785f, -0.425325f, 0.688191f},
    {-0.309017f, -0.500000f, 0.809017f}, {-0.147621f, -0.716567f, 0.681718f}, {-0.425325f, -0.688191f, 0.587785f},
    {-0.162460f, -0.262866f, 0.951056f}, {0.442863f, -0.238856f, 0.864188f}, {0.162460f, -0.262866f, 0.951056f},
    {0.309017f, -0.500000f, 0.809017f}, {0.147621f, -0.716567f, 0.681718f}, {0.000000f, -0.525731f, 0.850651f},
    {0.425325f, -0.688191f, 0.587785f}, {0.587785f, -0.425325f, 0.688191f}, {0.688191f, -0.587785f, 0.425325f},
    {-0.955423f, 0.295242f, 0.000000f}, {-0.951056f, 0.162460f, 0.262866f}, {-1.000000f, 0.000000f, 0.000000f},
    {-0.850651f, 0.000000f, 0.525731f}, {-0.955423f, -0.295242f, 0.000000f}, {-0.951056f, -0.162460f, 0.262866f},
    {-0.864188f, 0.442863f, -0.238856f}, {-0.951056f, 0.162460f, -0.262866f}, {-0.809017f, 0.309017f, -0.500000f},
    {-0.864188f, -0.442863f, -0.238856f}, {-0.951056f, -0.162460f, -0.262866f}, {-0.809017f, -0.309017f, -0.500000f},
    {-0.681718f, 0.147621f, -0.716567f}, {-0.681718f, -0.147621f, -0.716567f}, {-0.850651f, 0.000000f, -0.525731f},
    {-0.688191f, 0.587785f, -0.425325f}, {-0.587785f, 0.425325f, -0.688191f}, {-0.425325f, 0.688191f, -0.587785f},
    {-0.425325f, -0.688191f, -0.587785f}, {-0.587785f, -0.425325f, -0.688191f}, {-0.688191f, -0.587785f, -0.425325f},
};

void or_sky_dirs(float *dirs) { memcpy(dirs, or_anorms, sizeof or_anorms); }

/* CanLeafTraceToSky (rad/lightmap/lightmap.go:425-451), with the intent of
 * its broken SIMD loads: center = float32(mins + maxs) * 0.5 per axis (the
 * sum taken in 32-bit integers; Go's int16 sum could wrap at the coordinate
 * limit); for each of the 162 Anorms directions a, TestLineDoesHitSky from
 * center to (a * -MAX_TRACE_LENGTH) + center with skipId
 * TRACE_ID_STATICPROP | -1 = -1; true iff some lane's fractionVisible > 0,
 * i.e. the segment misses every non-sky triangle before its closest sky
 * hit.  "Any" does not depend on the 4-lane grouping of the directions. */
typedef struct {
  const or_kdtree_t *kd;
  const or_tri_t *tris;
  int64_t ntri;
  const int16_t *mins, *maxs;
  uint8_t *out;
  float M;
} sky_ctx_t;

static void leaf_sky_range(void *vc, int64_t b, int64_t e) {
  const sky_ctx_t *c = (const sky_ctx_t *)vc;
  const float MTL = (float)OR_MAX_TRACE_LENGTH;
  for (int64_t i = b; i < e; ++i) {
    float ctr[3];
    for (int a = 0; a < 3; ++a)
      ctr[a] = (float)((int32_t)c->mins[3 * i + a] + (int32_t)c->maxs[3 * i + a]) * 0.5f;
    uint8_t any = 0;
    for (int k = 0; k < 162 && !any; ++k) {
      float en[3];
      for (int a = 0; a < 3; ++a) en[a] = or_anorms[k][a] * -MTL + ctr[a];
      uint8_t v;
      testline_one(c->kd, c->M, c->tris, c->ntri, ctr, en, -1, &v, NULL, NULL, NULL);
      any = v;
    }
    c->out[i] = any;
  }
}

void or_can_leaf_trace_to_sky(const or_kdtree_t *kd, const or_tri_t *tris, int64_t ntri,
                              const int16_t *mins, const int16_t *maxs, int64_t n,
                              uint8_t *can_see_sky, int nthreads) {
  sky_ctx_t c = {kd, tris, ntri, mins, maxs, can_see_sky, kd ? kd_scene_maxabs(kd) : 1.0f};
  or_parallel(n, nthreads, leaf_sky_range, &c);
}

typedef struct {
  const or_kdtree_t *kd;
  const or_tri_t *tris;
  int64_t ntri;
  const or_samples_t *s;
  const or_light_t *L;
  int32_t nl;
  float eps;
  int32_t skip;
  float *rgb;
  float M;
  float dirs[162 * 3];
} gl_ctx_t;

static int seg_visible(const gl_ctx_t *c, const float s[3], const float e[3]) {
  uint8_t v;
  testline_one(c->kd, c->M, c->tris, c->ntri, s, e, c->skip, &v, NULL, NULL, NULL);
  return v;
}

static void gather_one(const gl_ctx_t *c, int64_t i, float out[3]) {
  const float p[3] = {c->s->px[i], c->s->py[i], c->s->pz[i]};
  const float n[3] = {c->s->nx[i], c->s->ny[i], c->s->nz[i]};
  const float st[3] = {p[0] + c->eps * n[0], p[1] + c->eps * n[1], p[2] + c->eps * n[2]};
  const float MTL = (float)OR_MAX_TRACE_LENGTH;
  float acc[3] = {0.0f, 0.0f, 0.0f};
  for (int32_t li = 0; li < c->nl; ++li) {
    const or_light_t *L = c->L + li;
    float scale = 0.0f;
    if (L->type == 0) {                          /* EMIT_SURFACE */
      float lx = L->origin[0] - p[0], ly = L->origin[1] - p[1], lz = L->origin[2] - p[2];
      float dist2 = (lx * lx + ly * ly) + lz * lz;
      if (!(dist2 > 0.0f)) continue;
      float dist = sqrtf(dist2);
      float inv = 1.0f / dist;
      float ux = lx * inv, uy = ly * inv, uz = lz * inv;
      float dot = (n[0] * ux + n[1] * uy) + n[2] * uz;
      if (!(dot > 0.0f)) continue;
      float dot2 = -((L->normal[0] * ux + L->normal[1] * uy) + L->normal[2] * uz);
      if (!(dot2 > 0.0f)) continue;
      if (!seg_visible(c, st, L->origin)) continue;
      scale = (dot * dot2) / dist2;
    } else if (L->type == 1 || L->type == 2) {   /* EMIT_POINT / EMIT_SPOTLIGHT */
      float lx = L->origin[0] - p[0], ly = L->origin[1] - p[1], lz = L->origin[2] - p[2];
      float dist2 = (lx * lx + ly * ly) + lz * lz;
      if (!(dist2 > 0.0f)) continue;
      float dist = sqrtf(dist2);
      float inv = 1.0f / dist;
      float ux = lx * inv, uy = ly * inv, uz = lz * inv;
      float dot = (n[0] * ux + n[1] * uy) + n[2] * uz;
      if (!(dot > 0.0f)) continue;
      float fd = dist < L->cap_distance ? dist : L->cap_distance;
      float den = (L->constant_attn + L->linear_attn * fd) + L->quadratic_attn * (fd * fd);
      if (!(den > 0.0f)) continue;
      float fall = 1.0f / den;
      if (L->type == 2) {
        float dot2 = -((L->normal[0] * ux + L->normal[1] * uy) + L->normal[2] * uz);
        if (!(dot2 > L->stopdot2)) continue;
        if (dot2 <= L->stopdot)
          fall = fall * ((dot2 - L->stopdot2) / (L->stopdot - L->stopdot2));
        if (L->exponent != 0.0f && L->exponent != 1.0f) /* float64 power, rounded once */
          fall = fall * (float)pow((double)dot2, (double)L->exponent);
      }
      if (L->end_fade > L->start_fade) {
        if (dist >= L->end_fade) continue;
        if (dist > L->start_fade) {
          float x = (dist - L->start_fade) / (L->end_fade - L->start_fade);
          float sm = (x * x) * (3.0f - 2.0f * x);
          fall = fall * (1.0f - sm);
        }
      }
      if (!seg_visible(c, st, L->origin)) continue;
      scale = dot * fall;
    } else if (L->type == 3) {                   /* EMIT_SKYLIGHT (sun) */
      float dot = -((n[0] * L->normal[0] + n[1] * L->normal[1]) + n[2] * L->normal[2]);
      if (!(dot > 0.0f)) continue;
      float en[3] = {st[0] - MTL * L->normal[0], st[1] - MTL * L->normal[1],
                     st[2] - MTL * L->normal[2]};
      if (!seg_visible(c, st, en)) continue;
      scale = dot;
    } else if (L->type == 5) {                   /* EMIT_SKYAMBIENT */
      float a = 0.0f, w = 0.0f;
      for (int k = 0; k < 162; ++k) {
        const float *dk = c->dirs + 3 * k;
        float ck = (n[0] * dk[0] + n[1] * dk[1]) + n[2] * dk[2];
        if (!(ck > 0.0f)) continue;
        w = w + ck;
        float en[3] = {st[0] + MTL * dk[0], st[1] + MTL * dk[1], st[2] + MTL * dk[2]};
        if (seg_visible(c, st, en)) a = a + ck;
      }
      if (!(w > 0.0f)) continue;
      scale = a / w;
    } else {
      continue;
    }
    acc[0] = acc[0] + L->intensity[0] * scale;
    acc[1] = acc[1] + L->intensity[1] * scale;
    acc[2] = acc[2] + L->intensity[2] * scale;
  }
  out[0] = acc[0];
  out[1] = acc[1];
  out[2] = acc[2];
}

static void gl_range(void *vc, int64_t b, int64_t e) {
  gl_ctx_t *c = (gl_ctx_t *)vc;
  for (int64_t i = b; i < e; ++i) gather_one(c, i, c->rgb + 3 * i);
}

void or_gather_light(const or_kdtree_t *kd, const or_tri_t *tris, int64_t ntri,
                     const or_samples_t *s, int64_t n, const or_light_t *lights,
                     int32_t n_lights, float normal_offset, int32_t skip_id,
                     float *rgb, int nthreads) {
  gl_ctx_t *c = (gl_ctx_t *)malloc(sizeof(gl_ctx_t));
  c->kd = kd; c->tris = tris; c->ntri = ntri; c->s = s; c->L = lights; c->nl = n_lights;
  c->eps = normal_offset; c->skip = skip_id; c->rgb = rgb;
  c->M = kd ? kd_scene_maxabs(kd) : 1.0f;
  or_sky_dirs(c->dirs);
  or_parallel(n, nthreads, gl_range, c);
  free(c);
}

/* ------------------------------------------------------------------------ */
/* FP16 / BF16 conversion, round to nearest even (IEEE 754 binary16, bfloat16) */
/* ------------------------------------------------------------------------ */
static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

uint16_t or_f32_to_bf16(float f) {
  uint32_t u = f2u(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
float or_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

uint16_t or_f32_to_f16(float f) {
  uint32_t u = f2u(f);
  uint32_t sign = (u >> 16) & 0x8000u;
  uint32_t a = u & 0x7fffffffu;
  if (a > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);       /* NaN */
  if (a >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);      /* overflow -> inf */
  if (a < 0x38800000u) {                                        /* subnormal half */
    if (a < 0x33000000u) return (uint16_t)sign;                 /* < 2^-25: +-0 */
    uint32_t e = a >> 23, m = (a & 0x7fffffu) | 0x800000u;
    /* half subnormal unit is 2^-24: value = m * 2^(e-150) = m*2^(e-126) units */
    int sh = 126 - (int)e;              /* 14..24 */
    uint32_t v = m >> sh;
    uint32_t rem = m & ((1u << sh) - 1u), halfu = 1u << (sh - 1);
    if (rem > halfu || (rem == halfu && (v & 1u))) ++v;
    return (uint16_t)(sign | v);
  }
  uint32_t e = (a >> 23) - 112u, m = a & 0x7fffffu;
  uint32_t v = (e << 10) | (m >> 13);
  uint32_t rem = m & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (v & 1u))) ++v;
  return (uint16_t)(sign | v);
}

float or_f16_to_f32(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  if (e == 0) {
    if (m == 0) return u2f(sign);
    float v = (float)m * 0x1p-24f;
    return sign ? -v : v;
  }
  if (e == 31) return u2f(sign | 0x7f800000u | (m << 13));
  return u2f(sign | ((e + 112u) << 23) | (m << 13));
}

/* ------------------------------------------------------------------------ */
/* BounceLight (absent; state Patch.TotalLight/Reflectivity/Area,            */
/* common/types/patch.go:33-42; -bounce cmd/args.go:66).  Frozen recurrence: */
/*   add_i  = sum_t coef_t * emit[col_t]        (FP32, kernel order)         */
/*   total_i += add_i ; emit_i = add_i * refl_i                              */
/*   E_k = sum_i area_i * ((add.r + add.g) + add.b)   (FP64, fixed tree)    */
/*   stop after bounce k when E_k <= stop_frac * sum_{m<=k} E_m, or k = max  */
/* Kernel orders (G = 4 for FP32, 8 for FP16/BF16; FP16/BF16 rows multiply   */
/* the reduced sum by the row scale):                                        */
/*  0 wave per row: lane l owns positions p with (p / G) % 64 == l, fmaf in  */
/*    increasing p, xor-butterfly 32..1;                                     */
/*  1 sequential (one thread per row);                                       */
/*  2 LDS-tiled: the row splits into segments by column tile of width C; in  */
/*    a segment [sa, sb) lane l of a GS-lane group owns positions p with     */
/*    ((p - sa) / G) % GS == l, butterfly GS/2..1, and segment sums are      */
/*    added to the row total in tile order (empty ones add 0).              */
/*  3 stream: rows are dealt to chunks and 64-row waves (see bg_stream);    */
/*    per tile a wave's segments (each padded to G) lie back to back and    */
/*    are read in passes of 64 G-entry chunks; a pass's chunk sums are      */
/*    combined by a segmented scan of fixed stages, and the value at each   */
/*    segment's last chunk is added to its row total in tile, pass order.   */
/* ------------------------------------------------------------------------ */
double or_energy_reduce(const double *e, int64_t n) {
  if (n <= 0) return 0.0;
  int64_t m = n;
  double *cur = (double *)malloc(sizeof(double) * (size_t)m);
  memcpy(cur, e, sizeof(double) * (size_t)m);
  while (m > 1) {
    int64_t nb = (m + 2047) / 2048;
    double *nx = (double *)malloc(sizeof(double) * (size_t)nb);
    for (int64_t b = 0; b < nb; ++b) {
      double v[256];
      for (int t = 0; t < 256; ++t) {
        double a = 0.0;
        for (int j = 0; j < 8; ++j) {
          int64_t k = b * 2048 + t + 256 * j;
          if (k < m) a = a + cur[k];
        }
        v[t] = a;
      }
      for (int s = 128; s >= 1; s >>= 1)
        for (int t = 0; t < s; ++t) v[t] = v[t] + v[t + s];
      nx[b] = v[0];
    }
    free(cur);
    cur = nx;
    m = nb;
  }
  double r = cur[0];
  free(cur);
  return r;
}

typedef struct {
  int64_t n;
  const int64_t *rp;
  const int32_t *col;
  const void *coef;
  int kind;
  int order;         /* 0 wave per row, 1 sequential, 2 LDS-tiled, 3 stream */
  int tile_cols, group;
  const float *scale;
  const float *emit; /* 3n, current */
  float *add;        /* 3n */
  const int32_t *rows_by_len; /* stream: rows longest first (ties by index) */
  int64_t n_chunks;           /* stream: row chunks */
} bg_ctx_t;

static inline float coef_at(const bg_ctx_t *c, int64_t k) {
  if (c->kind == 0) return ((const float *)c->coef)[k];
  uint16_t h = ((const uint16_t *)c->coef)[k];
  return c->kind == 1 ? or_f16_to_f32(h) : or_bf16_to_f32(h);
}

/* xor butterfly over the first `lanes` accumulators */
static void butterfly3(float (*acc)[3], int lanes) {
  for (int off = lanes / 2; off >= 1; off >>= 1) {
    float nw[64][3];
    for (int l = 0; l < lanes; ++l)
      for (int q = 0; q < 3; ++q) nw[l][q] = acc[l][q] + acc[l ^ off][q];
    for (int l = 0; l < lanes; ++l)
      for (int q = 0; q < 3; ++q) acc[l][q] = nw[l][q];
  }
}

static void bg_range(void *vc, int64_t b, int64_t e) {
  bg_ctx_t *c = (bg_ctx_t *)vc;
  const int G = c->kind == 0 ? 4 : 8;
  for (int64_t i = b; i < e; ++i) {
    const int64_t s0 = c->rp[i], len = c->rp[i + 1] - s0;
    float out[3] = {0.0f, 0.0f, 0.0f};
    if (c->order == 2) {
      /* LDS-tiled: segments per column tile, GS-lane groups of G-entry
       * chunks aligned to G, butterfly per segment, tile-ordered row sum */
      const int64_t C = c->tile_cols;
      const int GS = c->group;
      const int64_t nt = (c->n + C - 1) / C;
      int64_t q = 0;
      for (int64_t t = 0; t < nt; ++t) {
        const int64_t sa = q;
        while (q < len && (int64_t)c->col[s0 + q] < (t + 1) * C) ++q;
        const int64_t sb = q, base = sa;
        float acc[64][3];
        memset(acc, 0, sizeof acc);
        for (int64_t p = sa; p < sb; ++p) {
          const int l = (int)(((p - base) / G) % GS);
          const float w = coef_at(c, s0 + p);
          const float *em = c->emit + 3 * (int64_t)c->col[s0 + p];
          acc[l][0] = fmaf(w, em[0], acc[l][0]);
          acc[l][1] = fmaf(w, em[1], acc[l][1]);
          acc[l][2] = fmaf(w, em[2], acc[l][2]);
        }
        butterfly3(acc, GS);
        for (int k = 0; k < 3; ++k) out[k] = out[k] + acc[0][k];
      }
    } else {
      float acc[64][3];
      memset(acc, 0, sizeof acc);
      for (int64_t p = 0; p < len; ++p) {
        const int l = c->order == 1 ? 0 : (int)((p / G) % 64);
        const float w = coef_at(c, s0 + p);
        const float *em = c->emit + 3 * (int64_t)c->col[s0 + p];
        acc[l][0] = fmaf(w, em[0], acc[l][0]);
        acc[l][1] = fmaf(w, em[1], acc[l][1]);
        acc[l][2] = fmaf(w, em[2], acc[l][2]);
      }
      butterfly3(acc, 64);
      for (int k = 0; k < 3; ++k) out[k] = acc[0][k];
    }
    const float sc = c->kind == 0 ? 1.0f : c->scale[i];
    for (int q = 0; q < 3; ++q) c->add[3 * i + q] = c->kind == 0 ? out[q] : out[q] * sc;
  }
}

/* Stream order, one chunk of rows per call (chunks are independent).
 * Chunking: n_chunks = min(256, ceil(n / 256)) for n <= 256 x 1024, else
 * 256 ceil(n / (256 x 1024)) (at most 1024 rows a chunk either way); chunk c
 * holds rows rows_by_len[c + r n_chunks]; wave w (0..15), slot j (0..63)
 * takes rank r = 16 j + (j odd ? 15 - w : w).  In tile t a slot's segment
 * is its row's entries with columns in [t C, (t + 1) C), padded with zero
 * entries (column t C) to a multiple of G = 4; the wave's segments follow
 * each other in slot order from stream position 0, as G-entry chunks.  Pass i
 * covers chunks [128 i, 128 (i + 1)); lane l takes chunks A = 2 l and
 * B = 2 l + 1 (a chunk past the wave's last one is absent):
 *   pA = fmaf over A's entries from 0;
 *   v  = fmaf over B's entries from (B starts a segment ? 0 : pA), or pA
 *        without B;
 *   s(l) = the lane holding the first chunk, in this pass, of the segment of
 *        the lane's last chunk (0 if that segment began in an earlier pass);
 *   scan stages, each reading the values before the stage:
 *     d = 1, 2, 4, 8: v[l] += v[l - d] if l - d >= max(s(l), l & ~15);
 *     v[l] += v[(l & ~15) - 1] if (l & 16) and s(l) < (l & ~15);
 *     v[l] += v[31]             if l >= 32 and s(l) < 32;
 *   carry(l) = v[l - 1] after the scan (0 for lane 0);
 *   if B starts a segment, A's segment adds (A starts a segment ? pA :
 *   carry + pA) to its row total; if the lane's last chunk ends its segment
 *   or l = 63, v[l] is added to that segment's row total.
 * Row totals take these additions in tile and pass order. */
static void bg_stream_range(void *vc, int64_t cb, int64_t ce) {
  bg_ctx_t *c = (bg_ctx_t *)vc;
  const int G = 4; /* entries per chunk, every precision */
  const int64_t PW = 128 * (int64_t)G;
  const int64_t C = c->tile_cols;
  const int64_t nt = (c->n + C - 1) / C;
  for (int64_t ck = cb; ck < ce; ++ck) {
    for (int w = 0; w < 16; ++w) {
      int64_t row[64], cur[64], start[64], len[64], sa[65];
      float acc[64][3];
      for (int j = 0; j < 64; ++j) {
        const int64_t rank = (int64_t)j * 16 + ((j & 1) ? 15 - w : w);
        const int64_t oi = ck + rank * c->n_chunks;
        row[j] = oi < c->n ? c->rows_by_len[oi] : -1;
        cur[j] = 0;
        acc[j][0] = acc[j][1] = acc[j][2] = 0.0f;
      }
      for (int64_t t = 0; t < nt; ++t) {
        int64_t pos = 0;
        for (int j = 0; j < 64; ++j) {
          sa[j] = pos;
          len[j] = 0;
          start[j] = cur[j];
          if (row[j] < 0) continue;
          const int64_t s0 = c->rp[row[j]], L = c->rp[row[j] + 1] - s0;
          int64_t q = cur[j];
          while (q < L && (int64_t)c->col[s0 + q] < (t + 1) * C) ++q;
          len[j] = q - cur[j];
          cur[j] = q;
          pos += (len[j] + G - 1) / G * G;
        }
        sa[64] = pos;
        const float *pad = c->emit + 3 * t * C; /* LDS offset 0 of the tile */
        for (int64_t p0 = 0; p0 < pos; p0 += PW) {
          /* per chunk of the pass: value (from 0 or chained), segment, flags */
          float pa[64][3], v[64][3];
          int s[64], seg_a[64], seg_x[64], st_a[64], st_b[64], end_x[64], has_a[64], has_b[64];
          int js = 0;
          for (int l = 0; l < 64; ++l) {
            float acc_c[3] = {0.0f, 0.0f, 0.0f};
            has_a[l] = has_b[l] = 0;
            s[l] = 0;
            v[l][0] = v[l][1] = v[l][2] = 0.0f;
            pa[l][0] = pa[l][1] = pa[l][2] = 0.0f;
            for (int h = 0; h < 2; ++h) {
              const int64_t q = p0 + (int64_t)G * (2 * l + h);
              if (q >= pos) break;
              while (sa[js + 1] <= q) ++js;
              const int64_t off = q - sa[js];
              const int is_start = off == 0, is_last = q + G == sa[js + 1];
              if (h == 0) {
                has_a[l] = 1;
                st_a[l] = is_start;
                seg_a[l] = js;
              } else {
                has_b[l] = 1;
                st_b[l] = is_start;
                if (is_start) acc_c[0] = acc_c[1] = acc_c[2] = 0.0f;
              }
              const int64_t s0 = c->rp[row[js]] + start[js];
              for (int m = 0; m < G; ++m) {
                const int64_t e = off + m;
                float wgt;
                const float *em;
                if (e < len[js]) {
                  wgt = coef_at(c, s0 + e);
                  em = c->emit + 3 * (int64_t)c->col[s0 + e];
                } else {
                  wgt = 0.0f;
                  em = pad;
                }
                acc_c[0] = fmaf(wgt, em[0], acc_c[0]);
                acc_c[1] = fmaf(wgt, em[1], acc_c[1]);
                acc_c[2] = fmaf(wgt, em[2], acc_c[2]);
              }
              if (h == 0) memcpy(pa[l], acc_c, sizeof acc_c);
              memcpy(v[l], acc_c, sizeof acc_c);
              /* the lane's last chunk so far: its segment, end flag, s */
              seg_x[l] = js;
              end_x[l] = is_last;
              s[l] = sa[js] >= p0 ? (int)(((sa[js] - p0) / G) >> 1) : 0;
            }
            if (!has_b[l]) st_b[l] = 0;
          }
          float nv[64][3];
          for (int d = 1; d <= 8; d <<= 1) {
            for (int l = 0; l < 64; ++l) {
              const int lo = s[l] > (l & ~15) ? s[l] : (l & ~15);
              for (int q = 0; q < 3; ++q)
                nv[l][q] = l - d >= lo ? v[l][q] + v[l - d][q] : v[l][q];
            }
            memcpy(v, nv, sizeof v);
          }
          for (int l = 0; l < 64; ++l)
            for (int q = 0; q < 3; ++q)
              nv[l][q] = ((l & 16) && s[l] < (l & ~15)) ? v[l][q] + v[(l & ~15) - 1][q] : v[l][q];
          memcpy(v, nv, sizeof v);
          for (int l = 0; l < 64; ++l)
            for (int q = 0; q < 3; ++q)
              nv[l][q] = (l >= 32 && s[l] < 32) ? v[l][q] + v[31][q] : v[l][q];
          memcpy(v, nv, sizeof v);
          for (int l = 0; l < 64; ++l) {
            if (!has_a[l]) continue;
            if (st_b[l]) {
              const int j = seg_a[l];
              for (int q = 0; q < 3; ++q) {
                const float carry = l > 0 ? v[l - 1][q] : 0.0f;
                const float val = st_a[l] ? pa[l][q] : carry + pa[l][q];
                acc[j][q] = acc[j][q] + val;
              }
            }
          }
          for (int l = 0; l < 64; ++l) {
            if (!has_a[l] || !(end_x[l] || l == 63)) continue;
            const int j = seg_x[l];
            for (int q = 0; q < 3; ++q) acc[j][q] = acc[j][q] + v[l][q];
          }
        }
      }
      for (int j = 0; j < 64; ++j) {
        if (row[j] < 0) continue;
        const float sc = c->kind == 0 ? 1.0f : c->scale[row[j]];
        for (int q = 0; q < 3; ++q)
          c->add[3 * row[j] + q] = c->kind == 0 ? acc[j][q] : acc[j][q] * sc;
      }
    }
  }
}

static int cmp_i32(const void *a, const void *b) {
  const int32_t i = *(const int32_t *)a, j = *(const int32_t *)b;
  return i < j ? -1 : (i > j);
}

static int cmp_len_desc(const void *a, const void *b, void *rp) {
  const int64_t *r = (const int64_t *)rp;
  const int32_t i = *(const int32_t *)a, j = *(const int32_t *)b;
  const int64_t li = r[i + 1] - r[i], lj = r[j + 1] - r[j];
  if (li != lj) return li > lj ? -1 : 1;
  return i < j ? -1 : (i > j); /* ties by index: a stable order */
}

int or_bounce_f32(int64_t n, const int64_t *row_ptr, const int32_t *col,
                  const void *coef, int coef_kind, const float *row_scale,
                  const float *emit0, const float *refl, const float *area,
                  float *total, int max_bounces, double stop_frac, double *energy,
                  int order, int tile_cols, int group, int nthreads) {
  float *emit = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
  float *add = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
  double *e = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  memcpy(emit, emit0, sizeof(float) * 3 * (size_t)n);
  bg_ctx_t c = {n, row_ptr, col, coef, coef_kind, order, tile_cols, group, row_scale, emit, add,
                NULL, 0};
  int32_t *by_len = NULL;
  const int runs = order == 4;   /* stream order with rows dealt in runs */
  const int by_index = order == 5; /* stream order, chunks' rows ranked by index */
  if (runs || by_index) c.order = order = 3;
  if (order == 3 && n > 0) {
    by_len = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    /* at most 1024 rows a chunk: min(256, ceil(n / 256)) chunks up to 256
       full chunks, beyond that a whole multiple of 256 */
    const int64_t b = (n + 255) / 256;
    c.n_chunks = n > 256 * 1024 ? 256 * ((n + 256 * 1024 - 1) / (256 * 1024)) : (b < 256 ? b : 256);
    if (!runs) {
      /* every row longest first; chunk c takes by_len[c + k n_chunks] */
      for (int64_t i = 0; i < n; ++i) by_len[i] = (int32_t)i;
      qsort_r(by_len, (size_t)n, sizeof(int32_t), cmp_len_desc, (void *)row_ptr);
      if (by_index) {
        /* chunk c keeps its rows, ranked again by row index */
        int32_t *mine = (int32_t *)malloc(sizeof(int32_t) * 1024);
        for (int64_t ck = 0; ck < c.n_chunks; ++ck) {
          int64_t m = 0;
          for (int64_t k = ck; k < n; k += c.n_chunks) mine[m++] = by_len[k];
          qsort(mine, (size_t)m, sizeof(int32_t), cmp_i32);
          for (int64_t k = 0; k < m; ++k) by_len[ck + k * c.n_chunks] = mine[k];
        }
        free(mine);
      }
    } else {
      /* chunk c's rows are the c-th run of consecutive rows, runs of
         ceil((n - c) / n_chunks) rows, longest first inside the run, and
         its k-th row sits at by_len[c + k n_chunks] as above */
      int32_t *run = (int32_t *)malloc(sizeof(int32_t) * 1024);
      int64_t first = 0;
      for (int64_t ck = 0; ck < c.n_chunks && first < n; ++ck) {
        const int64_t size = (n - ck + c.n_chunks - 1) / c.n_chunks;
        for (int64_t k = 0; k < size; ++k) run[k] = (int32_t)(first + k);
        qsort_r(run, (size_t)size, sizeof(int32_t), cmp_len_desc, (void *)row_ptr);
        for (int64_t k = 0; k < size; ++k) by_len[ck + k * c.n_chunks] = run[k];
        first += size;
      }
      free(run);
    }
    c.rows_by_len = by_len;
  }
  double cum = 0.0;
  int k = 0;
  while (k < max_bounces) {
    if (order == 3) or_parallel(c.n_chunks, nthreads, bg_stream_range, &c);
    else or_parallel(n, nthreads, bg_range, &c);
    for (int64_t i = 0; i < n; ++i) {
      float a0 = add[3 * i], a1 = add[3 * i + 1], a2 = add[3 * i + 2];
      total[3 * i + 0] = total[3 * i + 0] + a0;
      total[3 * i + 1] = total[3 * i + 1] + a1;
      total[3 * i + 2] = total[3 * i + 2] + a2;
      emit[3 * i + 0] = a0 * refl[3 * i + 0];
      emit[3 * i + 1] = a1 * refl[3 * i + 1];
      emit[3 * i + 2] = a2 * refl[3 * i + 2];
      e[i] = (double)area[i] * (((double)a0 + (double)a1) + (double)a2);
    }
    double Ek = or_energy_reduce(e, n);
    energy[k] = Ek;
    cum = cum + Ek;
    ++k;
    if (Ek <= stop_frac * cum) break;
  }
  free(emit);
  free(add);
  free(e);
  free(by_len);
  return k;
}

int or_bounce_f64(int64_t n, const int64_t *row_ptr, const int32_t *col,
                  const float *coef, const float *emit0, const float *refl,
                  const float *area, double *total, int max_bounces,
                  double stop_frac, double *energy) {
  double *emit = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  double *add = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
  for (int64_t i = 0; i < 3 * n; ++i) emit[i] = emit0[i];
  double cum = 0.0;
  int k = 0;
  while (k < max_bounces) {
    for (int64_t i = 0; i < n; ++i) {
      double a[3] = {0, 0, 0};
      for (int64_t t = row_ptr[i]; t < row_ptr[i + 1]; ++t) {
        const double *em = emit + 3 * (int64_t)col[t];
        for (int q = 0; q < 3; ++q) a[q] += (double)coef[t] * em[q];
      }
      for (int q = 0; q < 3; ++q) add[3 * i + q] = a[q];
    }
    double Ek = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      for (int q = 0; q < 3; ++q) {
        total[3 * i + q] += add[3 * i + q];
        emit[3 * i + q] = add[3 * i + q] * (double)refl[3 * i + q];
      }
      Ek += (double)area[i] * (add[3 * i] + add[3 * i + 1] + add[3 * i + 2]);
    }
    energy[k] = Ek;
    cum += Ek;
    ++k;
    if (Ek <= stop_frac * cum) break;
  }
  free(emit);
  free(add);
  return k;
}

/* ---------------------------------------------------------------------------
 * Patch generation.  MakePatchForFace (rad/patches/face.go:30-197), then
 * SubdividePatches (rad/patches/subdivide.go:22-128): SubdividePatch
 * (:166-267), ClipWindingEpsilon (:269-347), CreateChildPatch (:353-410),
 * WindingAreaAndBalancePoint (:412-438), WindingArea / WindingBounds
 * (vmath/polygon/winding.go:217-260), with mathgl's float32 vector ops
 * (Sub, Cross, Len, Dot; vector.MA / Scale, vmath/vector) rounding once per
 * operation.  Where the Go code does not do what it intends, the intent:
 *   - WindingCenter (winding.go:231-239) writes a local pointer and divides by
 *     an integer: the centre is the points' average, scaled by 1 / n;
 *   - WindingAreaAndBalancePoint writes its centre to a local pointer: the
 *     area-weighted centre is returned.
 * A child's normal is its face's normal: GetPhongNormal
 * (rad/lightmap/normallist.go:51-144) with no neighbouring face within the
 * smoothing threshold.  "Zero base light" is mathgl's ApproxEqual to 0
 * (every component of magnitude below 1e-20).
 * ------------------------------------------------------------------------- */
typedef struct {
  int n;
  float p[OR_FACE_MAX_POINTS][3];
} or_wind_t;

static float pw_len3(const float v[3]) { return sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); }
static void pw_cross(const float a[3], const float b[3], float o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static float pw_area(const or_wind_t *w) {
  float total = 0.0f;
  for (int i = 2; i < w->n; ++i) {
    float d1[3], d2[3], c[3];
    for (int k = 0; k < 3; ++k) {
      d1[k] = w->p[i - 1][k] - w->p[0][k];
      d2[k] = w->p[i][k] - w->p[0][k];
    }
    pw_cross(d1, d2, c);
    total += pw_len3(c);
  }
  return total * 0.5f;
}
static void pw_bounds(const or_wind_t *w, float mn[3], float mx[3]) {
  for (int k = 0; k < 3; ++k) {
    mn[k] = 99999.0f;
    mx[k] = -99999.0f;
  }
  for (int i = 0; i < w->n; ++i)
    for (int k = 0; k < 3; ++k) {
      const float v = w->p[i][k];
      if (v < mn[k]) mn[k] = v;
      if (v > mx[k]) mx[k] = v;
    }
}
static float pw_area_balance(const or_wind_t *w, float c[3]) {
  float total = 0.0f;
  c[0] = c[1] = c[2] = 0.0f;
  if (!w) return 0.0f;
  for (int i = 2; i < w->n; ++i) {
    float d1[3], d2[3], x[3];
    for (int k = 0; k < 3; ++k) {
      d1[k] = w->p[i - 1][k] - w->p[0][k];
      d2[k] = w->p[i][k] - w->p[0][k];
    }
    pw_cross(d1, d2, x);
    const float area = pw_len3(x);
    total += area;
    const float s = area / 3.0f;
    for (int k = 0; k < 3; ++k) c[k] = c[k] + s * w->p[i - 1][k];
    for (int k = 0; k < 3; ++k) c[k] = c[k] + s * w->p[i][k];
    for (int k = 0; k < 3; ++k) c[k] = c[k] + s * w->p[0][k];
  }
  if (total != 0.0f) {
    const float inv = 1.0f / total;
    for (int k = 0; k < 3; ++k) c[k] = c[k] * inv;
  }
  return total * 0.5f;
}
/* ClipWindingEpsilon against the axis plane x[axis] = dist: has_f / has_b
   say whether front / back windings exist.  Returns -1 past the point cap. */
static int pw_clip(const or_wind_t *in, int axis, float dist, float eps, or_wind_t *f, int *has_f,
                   or_wind_t *b, int *has_b) {
  float dists[OR_FACE_MAX_POINTS + 4];
  int sides[OR_FACE_MAX_POINTS + 4], counts[3] = {0, 0, 0};
  const float nrm[3] = {axis == 0 ? 1.0f : 0.0f, axis == 1 ? 1.0f : 0.0f, axis == 2 ? 1.0f : 0.0f};
  int i;
  for (i = 0; i < in->n; ++i) {
    float dot = (in->p[i][0] * nrm[0] + in->p[i][1] * nrm[1]) + in->p[i][2] * nrm[2];
    dot = dot - dist;
    dists[i] = dot;
    sides[i] = dot > eps ? 0 : (dot < -eps ? 1 : 2);
    counts[sides[i]]++;
  }
  sides[i] = sides[0];
  dists[i] = dists[0];
  *has_f = *has_b = 0;
  if (counts[0] == 0) {
    *b = *in;
    *has_b = 1;
    return 0;
  }
  if (counts[1] == 0) {
    *f = *in;
    *has_f = 1;
    return 0;
  }
  f->n = b->n = 0;
  *has_f = *has_b = 1;
  for (i = 0; i < in->n; ++i) {
    const float *p1 = in->p[i];
    if (f->n + 2 > OR_FACE_MAX_POINTS || b->n + 2 > OR_FACE_MAX_POINTS) return -1;
    if (sides[i] == 2) {
      memcpy(f->p[f->n++], p1, 12);
      memcpy(b->p[b->n++], p1, 12);
      continue;
    }
    if (sides[i] == 0) memcpy(f->p[f->n++], p1, 12);
    if (sides[i] == 1) memcpy(b->p[b->n++], p1, 12);
    if (sides[i + 1] == 2 || sides[i + 1] == sides[i]) continue;
    const float *p2 = in->p[(i + 1) % in->n];
    const float dot = dists[i] / (dists[i] - dists[i + 1]);
    float mid[3];
    for (int j = 0; j < 3; ++j) {
      if (nrm[j] == 1.0f) mid[j] = dist;
      else if (nrm[j] == -1.0f) mid[j] = -dist;
      else mid[j] = p1[j] + dot * (p2[j] - p1[j]);
    }
    memcpy(f->p[f->n++], mid, 12);
    memcpy(b->p[b->n++], mid, 12);
  }
  return 0;
}

typedef struct {
  or_patch_t *out;
  int64_t cap, n;
  const or_face_t *faces;
  float face_mn[3], face_mx[3]; /* of the face being subdivided */
  float min_chop;
  const float *vn;     /* the face's vertex normals (NULL: children keep the face normal) */
  float centroid[3];   /* the face's base-patch centre */
  int err;
} pw_ctx_t;

/* GetPhongNormal (rad/lightmap/normallist.go:51-144): the face normal, or,
   for the first edge j whose fan triangle (centroid, vertex j, vertex j+1)
   holds spot in barycentric terms (a1 >= 0 and a2 >= 0), the normalised
   (1 - a1 - a2) face normal + a1 vn[j] + a2 vn[j+1].  mgl32 Dot is
   left to right; Len is float32(math.Sqrt(float64(.))), a correctly rounded
   float sqrt; Normalize multiplies by 1 / Len.  Below 1e-20 the reference
   stops the bake; this keeps the face normal. */
void or_phong_normal(const or_face_t *fc, const float c[3], const float *vn, const float spot[3],
                     float out[3]) {
#define PDOT(a, b) (((a)[0] * (b)[0] + (a)[1] * (b)[1]) + (a)[2] * (b)[2])
  memcpy(out, fc->normal, 12);
  const int E = fc->n_points;
  for (int j = 0; j < E; ++j) {
    const float *p1 = fc->points[j], *p2 = fc->points[(j + 1) % E];
    const float *n1 = vn + 3 * j, *n2 = vn + 3 * ((j + 1) % E);
    float v1[3], v2[3], vs[3];
    for (int k = 0; k < 3; ++k) {
      v1[k] = p1[k] - c[k];
      v2[k] = p2[k] - c[k];
      vs[k] = spot[k] - c[k];
    }
    const float aa = PDOT(v1, v1), bb = PDOT(v2, v2), ab = PDOT(v1, v2);
    const float num = bb * PDOT(v1, vs) - ab * PDOT(vs, v2);
    const float a1 = num / (aa * bb - ab * ab);
    const float a2 = (PDOT(vs, v2) - a1 * ab) / bb;
    if (a1 >= 0.0f && a2 >= 0.0f) {
      const float scale = (1.0f - a1) - a2;
      float ph[3], t[3];
      for (int k = 0; k < 3; ++k) ph[k] = fc->normal[k] * scale;
      for (int k = 0; k < 3; ++k) t[k] = n1[k] * a1;
      for (int k = 0; k < 3; ++k) ph[k] = ph[k] + t[k];
      for (int k = 0; k < 3; ++k) t[k] = n2[k] * a2;
      for (int k = 0; k < 3; ++k) ph[k] = ph[k] + t[k];
      const float len = sqrtf(PDOT(ph, ph));
      if (len < 1.0e-20f) return;
      const float inv = 1.0f / len;
      for (int k = 0; k < 3; ++k) out[k] = ph[k] * inv;
      return;
    }
  }
#undef PDOT
}

/* PairEdges (rad/lightmap/lightmap.go:37-228) with smoothing groups 0 and
   no displacements: vertex j of face i sums, in face order, the normals of
   every other face with a point at the same position (bit for bit) whose
   normal's dot with face i's is >= threshold, then adds face i's normal and
   normalises.  Brute force over face pairs (the reference indexes shared
   vertices; the reference's loop at :74 runs to len(Edges), not
   f.NumEdges: this follows the per-face vertex loop it stands for). */
void or_pair_edges(const or_face_t *faces, int64_t n, float threshold, float *out) {
  for (int64_t i = 0; i < n; ++i) {
    const float *fn = faces[i].normal;
    float *o = out + (size_t)i * OR_FACE_MAX_POINTS * 3;
    memset(o, 0, sizeof(float) * OR_FACE_MAX_POINTS * 3);
    for (int j = 0; j < faces[i].n_points; ++j) {
      const float *v = faces[i].points[j];
      float sum[3] = {0.0f, 0.0f, 0.0f};
      for (int64_t k = 0; k < n; ++k) {
        if (k == i) continue;
        int shares = 0;
        for (int q = 0; q < faces[k].n_points && !shares; ++q)
          shares = memcmp(faces[k].points[q], v, 12) == 0;
        if (!shares) continue;
        const float *nn = faces[k].normal;
        const float cosang = (nn[0] * fn[0] + nn[1] * fn[1]) + nn[2] * fn[2];
        if (cosang >= threshold)
          for (int q = 0; q < 3; ++q) sum[q] = sum[q] + nn[q];
      }
      for (int q = 0; q < 3; ++q) sum[q] = sum[q] + fn[q];
      const float len = sqrtf((sum[0] * sum[0] + sum[1] * sum[1]) + sum[2] * sum[2]);
      const float inv = 1.0f / len;
      for (int q = 0; q < 3; ++q) o[3 * j + q] = sum[q] * inv;
    }
  }
}

static int pw_zero_light(const float v[3]) {
  return fabsf(v[0]) < 1e-20f && fabsf(v[1]) < 1e-20f && fabsf(v[2]) < 1e-20f;
}

/* the patch record at index k is kept in *P (out may be too small) */
static void pw_put(pw_ctx_t *c, int64_t k, const or_patch_t *P) {
  if (k < c->cap) c->out[k] = *P;
}

static void pw_subdivide(pw_ctx_t *c, int64_t k, or_patch_t *P) {
  const or_face_t *fc = &c->faces[P->face];
  if (fc->flags & 1u) return; /* never subdivide sky patches */
  float total[3];
  float widest = -1.0f;
  int wa = -1, should = 0;
  for (int i = 0; i < 3; ++i) total[i] = (P->maxs[i] - P->mins[i]) * P->lux_scale;
  for (int i = 0; i < 3; ++i) {
    if (total[i] > widest) {
      wa = i;
      widest = total[i];
    }
    if (total[i] >= P->chop && total[i] >= c->min_chop) should = 1;
  }
  if (!should && wa != -1) {
    if (total[wa] > total[(wa + 1) % 3] * 2.0f && total[wa] > total[(wa + 2) % 3] * 2.0f) {
      if (P->chop > c->min_chop) {
        should = 1;
        const float h = P->chop / 2.0f;
        P->chop = c->min_chop > h ? c->min_chop : h;
        pw_put(c, k, P);
      }
    }
  }
  if (!should) return;
  const float dist = (P->mins[wa] + P->maxs[wa]) * 0.5f;
  or_wind_t w, o1, o2;
  w.n = P->n_points;
  memcpy(w.p, P->points, sizeof w.p);
  int h1, h2;
  if (pw_clip(&w, wa, dist, 0.1f, &o1, &h1, &o2, &h2)) {
    c->err = 1;
    return;
  }
  float c1[3], c2[3];
  const float a1 = pw_area_balance(h1 ? &o1 : NULL, c1);
  const float a2 = pw_area_balance(h2 ? &o2 : NULL, c2);
  if (a1 == 0.0f || a2 == 0.0f) return;
  or_patch_t ch[2];
  const or_wind_t *ow[2] = {&o1, &o2};
  const float *ca[2] = {c1, c2};
  const float ar[2] = {a1, a2};
  int64_t idx[2];
  for (int q = 0; q < 2; ++q) {
    or_patch_t *C = &ch[q];
    *C = *P;
    idx[q] = c->n++;
    C->child1 = C->child2 = -1;
    C->parent = (int32_t)k;
    C->n_points = ow[q]->n;
    memset(C->points, 0, sizeof C->points);
    memcpy(C->points, ow[q]->p, 12 * (size_t)ow[q]->n);
    C->area = ar[q];
    memcpy(C->origin, ca[q], 12);
    if (c->vn)
      or_phong_normal(fc, c->centroid, c->vn, C->origin, C->normal);
    else
      memcpy(C->normal, fc->normal, 12);
    pw_bounds(ow[q], C->mins, C->maxs);
    if (pw_zero_light(fc->base_light)) {
      float t[3];
      for (int i = 0; i < 3; ++i) t[i] = (C->maxs[i] - C->mins[i]) * C->lux_scale;
      if (C->chop > c->min_chop && t[0] < C->chop && t[1] < C->chop && t[2] < C->chop) {
        for (int i = 0; i < 3; ++i) {
          if ((c->face_mx[i] == C->maxs[i] || c->face_mn[i] == C->mins[i]) && t[i] > c->min_chop) {
            const float h = C->chop / 2.0f;
            C->chop = c->min_chop > h ? c->min_chop : h;
            break;
          }
        }
      }
    }
    pw_put(c, idx[q], C);
  }
  P->child1 = (int32_t)idx[0];
  P->child2 = (int32_t)idx[1];
  pw_put(c, k, P);
  pw_subdivide(c, idx[0], &ch[0]);
  pw_subdivide(c, idx[1], &ch[1]);
}

int or_make_patches(const or_face_t *faces, int64_t n_faces, float max_chop, float min_chop,
                    or_patch_t *out, int64_t cap, int64_t *n_out) {
  return or_make_patches_phong(faces, n_faces, max_chop, min_chop, NULL, out, cap, n_out);
}

int or_make_patches_phong(const or_face_t *faces, int64_t n_faces, float max_chop,
                          float min_chop, const float *vnorm, or_patch_t *out, int64_t cap,
                          int64_t *n_out) {
  pw_ctx_t c;
  memset(&c, 0, sizeof c);
  c.out = out;
  c.cap = out ? cap : 0;
  c.faces = faces;
  c.min_chop = min_chop;
  /* base patches, one per face of positive area */
  or_patch_t *base = (or_patch_t *)malloc(sizeof(or_patch_t) * (size_t)(n_faces > 0 ? n_faces : 1));
  int64_t nb = 0;
  for (int64_t f = 0; f < n_faces; ++f) {
    const or_face_t *fc = &faces[f];
    if (fc->n_points > OR_FACE_MAX_POINTS || fc->n_points < 0) {
      free(base);
      return -1;
    }
    or_wind_t w;
    w.n = fc->n_points;
    memcpy(w.p, fc->points, sizeof w.p);
    const float area = pw_area(&w);
    if (area <= 0.0f) continue;
    or_patch_t *P = &base[nb];
    memset(P, 0, sizeof *P);
    P->face = (int32_t)f;
    P->parent = P->child1 = P->child2 = -1;
    float cs[2] = {0.0f, 0.0f};
    for (int j = 0; j < 3; ++j) cs[0] += fc->lux_s[j] * fc->lux_s[j];
    for (int j = 0; j < 3; ++j) cs[1] += fc->lux_t[j] * fc->lux_t[j];
    cs[0] = sqrtf(cs[0]);
    cs[1] = sqrtf(cs[1]);
    P->lux_scale = (cs[0] + cs[1]) / 2.0f;
    P->chop = max_chop;
    P->area = area;
    float ctr[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < w.n; ++i)
      for (int k = 0; k < 3; ++k) ctr[k] = w.p[i][k] + ctr[k];
    const float sc = 1.0f / (float)w.n;
    for (int k = 0; k < 3; ++k) P->origin[k] = ctr[k] * sc;
    memcpy(P->normal, fc->normal, 12);
    pw_bounds(&w, P->mins, P->maxs);
    P->n_points = w.n;
    memcpy(P->points, w.p, 12 * (size_t)w.n);
    ++nb;
  }
  c.n = nb;
  for (int64_t i = 0; i < nb; ++i) pw_put(&c, i, &base[i]);
  for (int64_t i = 0; i < nb; ++i) {
    if (faces[base[i].face].flags & 2u) continue; /* PreventSubdivision */
    memcpy(c.face_mn, base[i].mins, 12); /* the face's bounds (FaceMins / FaceMaxs) */
    memcpy(c.face_mx, base[i].maxs, 12);
    memcpy(c.centroid, base[i].origin, 12); /* face.go:153 (face offsets 0) */
    c.vn = vnorm ? vnorm + (size_t)base[i].face * OR_FACE_MAX_POINTS * 3 : NULL;
    pw_subdivide(&c, i, &base[i]);
    if (c.err) break;
  }
  free(base);
  *n_out = c.n;
  return c.err ? -1 : 0;
}

/* ------------------------------------------------------------------------ */
/* BSP clusters and PVS.                                                     */
/* ------------------------------------------------------------------------ */
/* PointInLeaf (rad/clustertable/point.go:14-36): dist = point.Dot(normal) -
 * distance (mgl32 Vec3.Dot, left to right); in front of the plane by more
 * than TEST_EPSILON (vmath/constants.go:7) -> children[0], behind by more
 * -> children[1], otherwise children[0]'s leaf unless its cluster is -1. */
#define OR_TEST_EPSILON 0.03125f
static int16_t point_in_leaf(const or_plane_t *planes, const or_bspnode_t *nodes,
                             const int16_t *leaf_cluster, int node, const float p[3]) {
  if (node < 0) return leaf_cluster[-1 - node];
  const or_bspnode_t *n = &nodes[node];
  const or_plane_t *pl = &planes[n->planenum];
  float dot = p[0] * pl->normal[0] + p[1] * pl->normal[1];
  dot = dot + p[2] * pl->normal[2];
  const float dist = dot - pl->dist;
  if (dist > OR_TEST_EPSILON) return point_in_leaf(planes, nodes, leaf_cluster, n->children[0], p);
  if (dist < -OR_TEST_EPSILON) return point_in_leaf(planes, nodes, leaf_cluster, n->children[1], p);
  int16_t t = point_in_leaf(planes, nodes, leaf_cluster, n->children[0], p);
  if (t != -1) return t;
  return point_in_leaf(planes, nodes, leaf_cluster, n->children[1], p);
}

int16_t or_cluster_from_point(const or_plane_t *planes, const or_bspnode_t *nodes,
                              const int16_t *leaf_cluster, const float p[3]) {
  return point_in_leaf(planes, nodes, leaf_cluster, 0, p);
}

void or_patch_clusters(const or_plane_t *planes, const or_bspnode_t *nodes,
                       const int16_t *leaf_cluster, const or_patch_t *patches, int64_t n,
                       int32_t *out) {
  for (int64_t i = 0; i < n; ++i) {
    const or_patch_t *q = &patches[i];
    int32_t c = or_cluster_from_point(planes, nodes, leaf_cluster, q->origin);
    if (c == -1) {
      for (int j = 0; j < q->n_points && j < OR_FACE_MAX_POINTS; ++j) {
        const int32_t cj = or_cluster_from_point(planes, nodes, leaf_cluster, q->points[j]);
        if (cj != -1) {
          c = cj;
          break;
        }
      }
    }
    out[i] = c;
  }
}

/* DecompressVis (rad/lightmap/vis.go:53-94) as the Source run-length format
 * defines it: the Go loop copies its input byte for byte, and its bound
 * compares a byte value against the row length; the restatement takes the
 * format its comments name. */
int or_decompress_vis(const uint8_t *vis, int64_t vis_len, const int32_t *ofs, int32_t n_clusters,
                      uint8_t *rows, int64_t *n_overrun) {
  const int64_t row = ((int64_t)n_clusters + 7) >> 3;
  int64_t over = 0;
  int rc = 0;
  memset(rows, 0, (size_t)(row * n_clusters));
  for (int32_t cl = 0; cl < n_clusters && rc == 0; ++cl) {
    uint8_t *out = rows + (int64_t)cl * row;
    int64_t in = ofs[cl], o = 0;
    do {
      if (in >= vis_len) { rc = -2; break; }
      if (vis[in]) {
        out[o++] = vis[in++];
        continue;
      }
      if (in + 1 >= vis_len) { rc = -2; break; }
      int64_t c = vis[in + 1];
      if (c == 0) { rc = -1; break; }
      in += 2;
      if (o + c > row) {
        c = row - o;
        ++over;
      }
      while (c) {
        out[o++] = 0;
        --c;
      }
    } while (o < row);
  }
  if (n_overrun) *n_overrun = over;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Luxel stage: CalcFaceExtents (rad/world/face.go:14-90; the Go function    */
/* takes its two output arrays by value, so its results never reach the      */
/* face: the restatement returns them) and the luxels of each face           */
/* (FinalLightFace is absent in the reference; the rule is the build's).     */
/* ------------------------------------------------------------------------ */
static float lm_coord(const float *p, const float *l, float o) {
  float v = p[0] * l[0];
  v = v + p[1] * l[1];
  v = v + p[2] * l[2];
  return v + o;
}

int64_t or_calc_face_extents(const or_face_t *faces, int64_t n, const float *lux_ofs,
                             or_face_lightmap_t *out) {
  int64_t at = 0;
  for (int64_t f = 0; f < n; ++f) {
    const or_face_t *fc = &faces[f];
    or_face_lightmap_t *r = &out[f];
    memset(r, 0, sizeof *r);
    r->luxel_ofs = at;
    if (fc->flags & (1u | 4u)) {
      r->flags = 1;
      continue;
    }
    if (fc->n_points < 1 || fc->n_points > OR_FACE_MAX_POINTS) {
      r->flags = 4;
      continue;
    }
    const float o[2] = {lux_ofs ? lux_ofs[2 * f] : 0.0f, lux_ofs ? lux_ofs[2 * f + 1] : 0.0f};
    float mn[2] = {1e24f, 1e24f}, mx[2] = {-1e24f, -1e24f};
    for (int i = 0; i < fc->n_points; ++i)
      for (int j = 0; j < 2; ++j) {
        float v = lm_coord(fc->points[i], j ? fc->lux_t : fc->lux_s, o[j]);
        if (v < mn[j]) mn[j] = v;
        if (v > mx[j]) mx[j] = v;
      }
    int big = 0;
    for (int i = 0; i < 2; ++i) {
      float lo = (float)floor((double)mn[i]), hi = (float)ceil((double)mx[i]);
      if (!(lo >= -2147483648.0f && hi <= 2147483520.0f)) {
        big = 1;
        continue;
      }
      r->mins[i] = (int32_t)lo;
      float sz = hi - lo;
      if (sz > 126.0f) big = 1; /* MAX_LIGHTMAP_DIM_WITHOUT_BORDER + 1 */
      r->size[i] = sz < 2147483520.0f ? (int32_t)sz : 0;
    }
    if (big) {
      r->flags = 2;
      continue;
    }
    r->n_luxels = (r->size[0] + 1) * (r->size[1] + 1);
    at += r->n_luxels;
  }
  return at;
}

/* LuxelSpaceToWorld by Cramer's rule in double, in this order. */
static int lm_to_world(const or_face_t *fc, const float *o, float us, float ut, float *P) {
  double a0 = fc->lux_s[0], a1 = fc->lux_s[1], a2 = fc->lux_s[2];
  double b0 = fc->lux_t[0], b1 = fc->lux_t[1], b2 = fc->lux_t[2];
  double n0 = fc->normal[0], n1 = fc->normal[1], n2 = fc->normal[2];
  double bn0 = b1 * n2 - b2 * n1, bn1 = b2 * n0 - b0 * n2, bn2 = b0 * n1 - b1 * n0;
  double na0 = n1 * a2 - n2 * a1, na1 = n2 * a0 - n0 * a2, na2 = n0 * a1 - n1 * a0;
  double ab0 = a1 * b2 - a2 * b1, ab1 = a2 * b0 - a0 * b2, ab2 = a0 * b1 - a1 * b0;
  double det = (a0 * bn0 + a1 * bn1) + a2 * bn2;
  if (det == 0.0) {
    P[0] = P[1] = P[2] = 0.0f;
    return 0;
  }
  double d = (n0 * (double)fc->points[0][0] + n1 * (double)fc->points[0][1]) +
             n2 * (double)fc->points[0][2];
  double r0 = (double)us - (double)o[0], r1 = (double)ut - (double)o[1];
  P[0] = (float)(((r0 * bn0 + r1 * na0) + d * ab0) / det);
  P[1] = (float)(((r0 * bn1 + r1 * na1) + d * ab1) / det);
  P[2] = (float)(((r0 * bn2 + r1 * na2) + d * ab2) / det);
  return 1;
}

void or_face_luxels(const or_face_t *faces, int64_t n_faces, const float *lux_ofs,
                    const or_face_lightmap_t *lm, const or_patch_t *patches, int64_t n_patches,
                    float *pos, int32_t *patch) {
  /* each face's leaf patches in patch order */
  int64_t *lo = (int64_t *)calloc((size_t)n_faces + 1, sizeof(int64_t));
  int64_t *li = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_patches > 0 ? n_patches : 1));
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n_faces > 0 ? n_faces : 1));
  for (int64_t p = 0; p < n_patches; ++p)
    if (patches[p].child1 == -1 && patches[p].face >= 0 && patches[p].face < n_faces)
      ++lo[patches[p].face + 1];
  for (int64_t f = 0; f < n_faces; ++f) lo[f + 1] += lo[f];
  for (int64_t f = 0; f < n_faces; ++f) fill[f] = lo[f];
  for (int64_t p = 0; p < n_patches; ++p)
    if (patches[p].child1 == -1 && patches[p].face >= 0 && patches[p].face < n_faces)
      li[fill[patches[p].face]++] = p;
  for (int64_t f = 0; f < n_faces; ++f) {
    const or_face_t *fc = &faces[f];
    const or_face_lightmap_t *L = &lm[f];
    const float o[2] = {lux_ofs ? lux_ofs[2 * f] : 0.0f, lux_ofs ? lux_ofs[2 * f + 1] : 0.0f};
    const int w = L->size[0] + 1;
    for (int l = 0; l < L->n_luxels; ++l) {
      const float us = (float)(L->mins[0] + l % w), ut = (float)(L->mins[1] + l / w);
      const int64_t at = L->luxel_ofs + l;
      if (pos) lm_to_world(fc, o, us, ut, pos + 3 * at);
      int64_t found = -1, best = -1;
      float bestd = 0.0f;
      for (int64_t k = lo[f]; k < lo[f + 1] && found < 0; ++k) {
        const int64_t p = li[k];
        const or_patch_t *q = &patches[p];
        int np = q->n_points < 0 ? 0 : (q->n_points > OR_FACE_MAX_POINTS ? OR_FACE_MAX_POINTS
                                                                           : q->n_points);
        if (np >= 3) {
          int ps = 0, ns = 0;
          for (int e = 0; e < np; ++e) {
            int e1 = e + 1 == np ? 0 : e + 1;
            float u0 = lm_coord(q->points[e], fc->lux_s, o[0]);
            float v0 = lm_coord(q->points[e], fc->lux_t, o[1]);
            float u1 = lm_coord(q->points[e1], fc->lux_s, o[0]);
            float v1 = lm_coord(q->points[e1], fc->lux_t, o[1]);
            float c1 = (u1 - u0) * (ut - v0);
            float c2 = (v1 - v0) * (us - u0);
            float cr = c1 - c2;
            ps |= cr > 0.0f;
            ns |= cr < 0.0f;
          }
          if ((ps || ns) && !(ps && ns)) {
            found = p;
            break;
          }
        }
        float du = us - lm_coord(q->origin, fc->lux_s, o[0]);
        float dv = ut - lm_coord(q->origin, fc->lux_t, o[1]);
        float d2 = du * du + dv * dv;
        if (best < 0 || d2 < bestd) {
          best = p;
          bestd = d2;
        }
      }
      if (patch) patch[at] = (int32_t)(found >= 0 ? found : best);
    }
  }
  free(lo);
  free(li);
  free(fill);
}

/* ------------------------------------------------------------------------ */
/* Lighting lump: ColorRGBExp32 luxels and per-face offsets.  The reference's */
/* writer is a stub (cmd/tasks/finish/main.go:15-18); the encoding is the     */
/* Source engine's VectorToColorRGBExp32 (named in the reference's            */
/* compute/leafambient/lighting.go:67), restated from its definition: an     */
/* exponent that puts the largest channel in [128, 256), channels truncated.  */
/* ------------------------------------------------------------------------ */
uint32_t or_rgbexp32(float r, float g, float b) {
  float m = r > g ? r : g;
  m = m > b ? m : b;
  if (!(m > 0.0f) || !(m <= 3.402823466e38f)) return 0u;
  int ex;
  (void)frexpf(m, &ex);         /* m = f 2^ex, f in [0.5, 1): floor(log2 m) = ex - 1 */
  const int e = (ex - 1) - 7;
  if (e < -127 || e > 127) return 0u;
  const float scale = ldexpf(1.0f, -e);
  const float c[3] = {r, g, b};
  uint32_t out = 0u;
  for (int k = 0; k < 3; ++k) {
    float v = c[k] > 0.0f ? c[k] * scale : 0.0f;
    int q = (int)v;               /* truncation toward zero */
    if (q > 255) q = 255;
    out |= (uint32_t)q << (8 * k);
  }
  return out | ((uint32_t)(uint8_t)(int8_t)e << 24);
}

int64_t or_lighting_lump(const float *luxel_rgb, const or_face_lightmap_t *lm, int64_t n_faces,
                         uint32_t *lump, int32_t *lightofs) {
  int64_t at = 0; /* in 4-byte entries */
  for (int64_t f = 0; f < n_faces; ++f) {
    const int64_t n = lm[f].n_luxels;
    if (n <= 0) {
      if (lightofs) lightofs[f] = -1;
      continue;
    }
    if (lightofs) lightofs[f] = (int32_t)(4 * (at + 1));
    if (lump) {
      const float *v = luxel_rgb + 3 * lm[f].luxel_ofs;
      double part[64][3];
      for (int l = 0; l < 64; ++l) part[l][0] = part[l][1] = part[l][2] = 0.0;
      for (int64_t k = 0; k < n; ++k)
        for (int c = 0; c < 3; ++c) part[k % 64][c] = part[k % 64][c] + (double)v[3 * k + c];
      for (int off = 32; off >= 1; off >>= 1) {
        double nx[64][3];
        for (int l = 0; l < 64; ++l)
          for (int c = 0; c < 3; ++c) nx[l][c] = part[l][c] + part[l ^ off][c];
        memcpy(part, nx, sizeof part);
      }
      const float avg[3] = {(float)(part[0][0] / (double)n), (float)(part[0][1] / (double)n),
                            (float)(part[0][2] / (double)n)};
      lump[at] = or_rgbexp32(avg[0], avg[1], avg[2]);
      for (int64_t k = 0; k < n; ++k)
        lump[at + 1 + k] = or_rgbexp32(v[3 * k], v[3 * k + 1], v[3 * k + 2]);
    }
    at += 1 + n;
  }
  return 4 * at;
}
