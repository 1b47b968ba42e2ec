/*
 * contour3d_oracle.c -- FP64 CPU restatement of the reference's
 * tessellate -> smooth -> render path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it (through
 * oracle/oracle.py).  The product library (contour3d_amd/) never links,
 * loads or calls it, and has no CPU fallback to it.
 *
 * Build: -std=c99 -O2 -ffp-contract=off, no -march (oracle/Makefile).  The
 * reference is built -std=c99 (Makefile:2), under which gcc does not fuse
 * multiply-adds; keeping contraction off keeps every rounding step of the
 * expressions below identical to the reference's.
 *
 * Structure: the streaming execute (o_execute) walks z-slabs with a 3-slab
 * ring exactly like src/contour3d/contour/main.c:49-113, so its traversal,
 * its neighbour-cell normal gather and its serial render order are those of
 * the reference; the GPU library uses different algorithms (global scans,
 * row-indexed gathers, packed atomic z-buffer) and is checked against this.
 *
 * Pinning (tests/test_oracle_*.py):
 *   tessellation  exhaustive 256 corner codes vs the closed-form per-tetra
 *                 count; every vertex on a straddling lattice edge named by
 *                 its edge id; face normal == -grad f / |grad f| on linear
 *                 fields (orientation); closed surfaces are watertight with
 *                 every mesh edge used once in each direction.
 *   normals       brute-force global-edge dictionary on random fields (checks
 *                 the 19-case neighbour table and its summation order);
 *                 plane -> face normal; sphere -> radial.
 *   projection    analytic screen points, depth == n.(p - camera).
 *   raster        exact-rational coverage of large triangles, z-order and
 *                 tie rule, shading closed forms, bbox clamping.
 *   lines, PPM    geometric coverage; header and row order.
 * Parity unpinned (no reference build is possible here: every reference
 * translation unit includes the absent SimpleDecomp header):
 *   - which quad diagonal and which tail->head direction each case uses
 *     (triangulate.c:204-270) is transcribed, not independently derivable;
 *     it decides the last bits of vertex positions and which pixels a quad
 *     covers.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include "contour3d.h"

typedef contour3d_vector_t o_vec;

/* ---------------------------------------------------------------------- */
/* vector algebra, src/contour3d/vector.c:5-76                             */

static o_vec o_add(o_vec a, o_vec b) {
  o_vec r = { a.x + b.x, a.y + b.y, a.z + b.z };
  return r;
}

static o_vec o_sub(o_vec a, o_vec b) {
  o_vec r = { a.x - b.x, a.y - b.y, a.z - b.z };
  return r;
}

static o_vec o_scale(double w, o_vec a) {
  o_vec r = { w * a.x, w * a.y, w * a.z };
  return r;
}

/* (x*x' + y*y') + z*z', vector.c:41-49 */
static double o_dot(o_vec a, o_vec b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}

/* vector.c:51-64 */
static o_vec o_cross(o_vec a, o_vec b) {
  o_vec r = {
    a.y * b.z - a.z * b.y,
    a.z * b.x - a.x * b.z,
    a.x * b.y - a.y * b.x,
  };
  return r;
}

/* reciprocal of the norm, then a multiply per component, vector.c:66-76 */
static o_vec o_unit(o_vec a) {
  const double s = 1. / sqrt(o_dot(a, a));
  o_vec r = { a.x * s, a.y * s, a.z * s };
  return r;
}

/* ---------------------------------------------------------------------- */
/* tessellation tables                                                      */

/* corner c of a cell is lattice point (i + (c&1), j + (c>>1&1), k + (c>>2)),
 * triangulate.c:160-162.  The six tetrahedra share the 0-7 body diagonal,
 * triangulate.c:145-152. */
static const int kTet[6][4] = {
  {0, 7, 3, 1}, {0, 7, 1, 5}, {0, 7, 5, 4},
  {0, 7, 4, 6}, {0, 7, 6, 2}, {0, 7, 2, 3},
};

/* Reduced tetra mask m (1..7, bit n = tetra vertex n above the threshold)
 * -> its 1 or 2 triangles, each as three (tail, head) tetra-vertex pairs,
 * triangulate.c:204-270. */
static const int kCaseNtri[8] = { 0, 1, 1, 2, 1, 2, 2, 1 };
static const int kCase[8][2][3][2] = {
  { {{0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}} },
  { {{0, 1}, {0, 3}, {0, 2}}, {{0, 0}, {0, 0}, {0, 0}} },
  { {{1, 0}, {1, 2}, {1, 3}}, {{0, 0}, {0, 0}, {0, 0}} },
  { {{0, 2}, {1, 3}, {0, 3}}, {{0, 2}, {1, 2}, {1, 3}} },
  { {{2, 0}, {2, 3}, {2, 1}}, {{0, 0}, {0, 0}, {0, 0}} },
  { {{2, 1}, {0, 3}, {2, 3}}, {{2, 1}, {0, 1}, {0, 3}} },
  { {{0, 2}, {2, 3}, {1, 3}}, {{0, 2}, {1, 3}, {0, 1}} },
  { {{3, 0}, {3, 2}, {3, 1}}, {{0, 0}, {0, 0}, {0, 0}} },
};

/* Cell-local id (0..18) of the lattice edge joining cube corners a and b,
 * the numbering of triangulate.c:23-32:
 *   x edges 0-3  = j + 2k      y edges 4-7  = 4 + i + 2k
 *   z edges 8-11 = 8 + i + 2j  face diagonals 12/13 (x faces),
 *   14/15 (y faces), 16/17 (z faces); body diagonal 18.
 * Only the diagonals that the six tetrahedra use are meaningful. */
static int o_edge_id(int a, int b) {
  const int lo = a < b ? a : b;
  const int d = a ^ b;
  switch (d) {
    case 1: return lo >> 1;
    case 2: return 4 + (lo & 1) + 2 * (lo >> 2);
    case 4: return 8 + (lo & 3);
    case 6: return 12 + (lo & 1);
    case 5: return 14 + ((lo >> 1) & 1);
    case 3: return 16 + (lo >> 2);
    case 7: return 18;
    default: return -1;
  }
}

/* Cells sharing lattice edge e of cell (0,0,0), as (dk, dj, di, local id
 * in that cell), in increasing linear cell order.  This is the content of
 * the 19-case table of normal.c:78-158; it is derived here from the
 * geometry (every cell within one step that contains both endpoints as a
 * tetra edge) and pinned by a brute-force global-edge test. */
typedef struct { int dk, dj, di, e; } o_nbr;
static o_nbr g_nbr[19][4];
static int g_nbr_n[19];
static int g_tables_ready = 0;

static int o_is_tet_edge(int a, int b) {
  for (int t = 0; t < 6; t++) {
    int ha = 0, hb = 0;
    for (int n = 0; n < 4; n++) {
      ha |= kTet[t][n] == a;
      hb |= kTet[t][n] == b;
    }
    if (ha && hb && a != b) return 1;
  }
  return 0;
}

static void o_build_tables(void) {
  if (g_tables_ready) return;
  for (int e = 0; e < 19; e++) g_nbr_n[e] = 0;
  for (int a = 0; a < 8; a++) {
    for (int b = a + 1; b < 8; b++) {
      if (!o_is_tet_edge(a, b)) continue;
      const int e = o_edge_id(a, b);
      const int A[3] = { a & 1, (a >> 1) & 1, a >> 2 };
      const int B[3] = { b & 1, (b >> 1) & 1, b >> 2 };
      for (int dk = -1; dk <= 1; dk++) {
        for (int dj = -1; dj <= 1; dj++) {
          for (int di = -1; di <= 1; di++) {
            const int la[3] = { A[0] - di, A[1] - dj, A[2] - dk };
            const int lb[3] = { B[0] - di, B[1] - dj, B[2] - dk };
            int ok = 1;
            for (int d = 0; d < 3; d++) {
              ok &= la[d] >= 0 && la[d] <= 1 && lb[d] >= 0 && lb[d] <= 1;
            }
            if (!ok) continue;
            const int ca = la[0] + 2 * la[1] + 4 * la[2];
            const int cb = lb[0] + 2 * lb[1] + 4 * lb[2];
            if (!o_is_tet_edge(ca, cb)) continue;
            o_nbr nb = { dk, dj, di, o_edge_id(ca, cb) };
            g_nbr[e][g_nbr_n[e]++] = nb;
          }
        }
      }
    }
  }
  g_tables_ready = 1;
}

/* ---------------------------------------------------------------------- */
/* triangles                                                                */

typedef struct {
  o_vec p[3];      /* vertices */
  o_vec fn;        /* unit face normal */
  o_vec vn[3];     /* vertex normals */
  double area;     /* |(p1-p0) x (p2-p0)| (twice the geometric area) */
  int edge[3];     /* cell-local lattice edge id of each vertex */
} o_tri;

typedef struct {
  int n;
  o_tri t[12];
} o_cell;

/* linear interpolation along a tetra edge, triangulate.c:40-73 */
static o_vec o_interp(double thr, o_vec p0, double v0, o_vec p1, double v1) {
  if (fabs(v1 - v0) < 1.e-8) return p0;
  const double f = (thr - v0) / (v1 - v0);
  return o_add(p0, o_scale(f, o_sub(p1, p0)));
}

/* face normal and area, triangulate.c:75-93 */
static void o_face(o_tri *tr) {
  const o_vec n = o_cross(o_sub(tr->p[1], tr->p[0]), o_sub(tr->p[2], tr->p[0]));
  tr->area = sqrt(o_dot(n, n));
  tr->fn = o_scale(1. / tr->area, n);
}

/* all triangles of one cell in slot order (tetra 0..5, then triangle
 * within the tetra), triangulate.c:133-276 and :325-338 */
static void o_cell_tessellate(const double val[8], const o_vec pos[8], double thr, o_cell *out) {
  out->n = 0;
  for (int t = 0; t < 6; t++) {
    int m = 0;
    for (int n = 0; n < 4; n++) m |= (thr < val[kTet[t][n]]) << n;
    const int rev = m >= 8;
    if (rev) m = 15 - m;
    for (int s = 0; s < kCaseNtri[m]; s++) {
      o_tri *tr = &out->t[out->n++];
      for (int v = 0; v < 3; v++) {
        const int q = rev ? 2 - v : v;   /* triangulate.c:108-112 */
        const int tail = kTet[t][kCase[m][s][q][0]];
        const int head = kTet[t][kCase[m][s][q][1]];
        tr->p[v] = o_interp(thr, pos[tail], val[tail], pos[head], val[head]);
        tr->edge[v] = o_edge_id(tail, head);
      }
      o_face(tr);
    }
  }
}

static int o_corner_code(const double val[8], double thr) {
  int code = 0;
  for (int c = 0; c < 8; c++) code |= (thr < val[c]) << c;
  return code;
}

/* ---------------------------------------------------------------------- */
/* vertex normals, normal.c:6-165                                          */

typedef struct { const o_tri *t; int n; } o_span;
typedef o_span (*o_fetch_fn)(const void *ctx, long i, long j, long k);

/* area-weighted sum over every triangle of the cells sharing edge e of
 * cell (i,j,k), in increasing cell order and slot order, then normalised */
static o_vec o_vertex_normal(const void *ctx, o_fetch_fn fetch, long i, long j, long k, int e) {
  o_vec s = { 0., 0., 0. };
  for (int q = 0; q < g_nbr_n[e]; q++) {
    const o_nbr *nb = &g_nbr[e][q];
    const o_span sp = fetch(ctx, i + nb->di, j + nb->dj, k + nb->dk);
    for (int t = 0; t < sp.n; t++) {
      const o_tri *tr = &sp.t[t];
      if (tr->edge[0] == nb->e || tr->edge[1] == nb->e || tr->edge[2] == nb->e) {
        s = o_add(s, o_scale(tr->area, tr->fn));
      }
    }
  }
  return o_unit(s);
}

/* ---------------------------------------------------------------------- */
/* view, projection and rasterisation                                       */

typedef struct {
  o_vec cam, center, lx, ly, normal, light;
  size_t W, H;
} o_view;

typedef struct {
  double depth;
  contour3d_color_t c;
} o_pixel;

/* src/contour3d/main.c:25-63 */
static void o_view_init(o_view *v, const o_vec *cam, const o_vec *light, const size_t sizes[2],
                        const o_vec *center, const o_vec local[2]) {
  v->cam = *cam;
  v->light = o_unit(*light);
  v->center = *center;
  v->lx = local[0];
  v->ly = local[1];
  v->normal = o_unit(o_cross(local[0], local[1]));
  v->W = sizes[0];
  v->H = sizes[1];
}

/* perspective projection to screen coordinates in [-0.5, 0.5] with depth
 * n.(p - camera); project.c:9-67.  Returns nonzero when rejected. */
static int o_project(const o_view *v, o_vec p, o_vec *q) {
  const o_vec ray = o_sub(p, v->cam);
  const double nc = o_dot(v->normal, v->cam);
  const double nr = o_dot(v->normal, ray);
  const double d = o_dot(v->normal, v->center);
  if (-1.e-8 < nr && nr < 1.e-8) return 1;
  const double t = (d - nc) / nr;
  if (t <= 0. || 1. <= t) return 1;
  const o_vec delta = {
    v->cam.x + t * ray.x - v->center.x,
    v->cam.y + t * ray.y - v->center.y,
    v->cam.z + t * ray.z - v->center.z,
  };
  q->x = o_dot(delta, v->lx) / o_dot(v->lx, v->lx);
  q->y = o_dot(delta, v->ly) / o_dot(v->ly, v->ly);
  q->z = nr;
  return 0;
}

/* (c - a) x (b - a) in 2-D, render.c:38-46 */
static double o_edge_fn(const o_vec *a, const o_vec *b, const o_vec *c) {
  return (c->x - a->x) * (b->y - a->y) - (c->y - a->y) * (b->x - a->x);
}

static long o_lmin(long a, long b) { return a < b ? a : b; }
static long o_lmax(long a, long b) { return b < a ? a : b; }

/* one triangle into the canvas, render.c:48-163 */
static void o_raster_tri(const o_view *v, const contour3d_color_t *fg, const o_tri *tr, o_pixel *canvas) {
  o_vec q[3];
  for (int n = 0; n < 3; n++) {
    if (o_project(v, tr->p[n], &q[n])) return;
  }
  const long W = (long)v->W, H = (long)v->H;
  /* int_fast32_t is a 64-bit long on the reference's x86-64 target */
  const long xmin = (long)((0.5 + fmin(q[0].x, fmin(q[1].x, q[2].x))) * v->W - 1);
  const long xmax = (long)((0.5 + fmax(q[0].x, fmax(q[1].x, q[2].x))) * v->W + 1);
  const long ymin = (long)((0.5 + fmin(q[0].y, fmin(q[1].y, q[2].y))) * v->H - 1);
  const long ymax = (long)((0.5 + fmax(q[0].y, fmax(q[1].y, q[2].y))) * v->H + 1);
  const long i0 = o_lmin(W - 1, o_lmax(0, xmin));
  const long i1 = o_lmin(W - 1, o_lmax(0, xmax));
  const long j0 = o_lmin(H - 1, o_lmax(0, ymin));
  const long j1 = o_lmin(H - 1, o_lmax(0, ymax));
  for (long j = j0; j <= j1; j++) {
    for (long i = i0; i <= i1; i++) {
      const o_vec pc = {
        1. * ((size_t)i + 0.5) / v->W - 0.5,
        1. * ((size_t)j + 0.5) / v->H - 0.5,
        0.,
      };
      const double area = o_edge_fn(&q[0], &q[1], &q[2]);
      const double ainv = 1. / (area < 0. ? fmin(area, -1.e-16) : fmax(area, 1.e-16));
      const double w0 = ainv * o_edge_fn(&q[1], &q[2], &pc);
      const double w1 = ainv * o_edge_fn(&q[2], &q[0], &pc);
      const double w2 = ainv * o_edge_fn(&q[0], &q[1], &pc);
      if (w0 < -1.e-8 || w1 < -1.e-8 || w2 < -1.e-8) continue;
      o_pixel *px = &canvas[(size_t)j * v->W + (size_t)i];
      const double depth = 1. / (w0 / q[0].z + w1 / q[1].z + w2 / q[2].z);
      if (depth < px->depth) continue;
      px->depth = depth;
      const o_vec n = o_unit(o_add(o_scale(w0, tr->vn[0]),
                                   o_add(o_scale(w1, tr->vn[1]), o_scale(w2, tr->vn[2]))));
      const double f = fmax(0.15, fabs(o_dot(n, v->light)));
      px->c.r = (uint8_t)(f * fg->r);
      px->c.g = (uint8_t)(f * fg->g);
      px->c.b = (uint8_t)(f * fg->b);
    }
  }
}

/* one line piece, line.c:26-119 */
static void o_raster_segment(const o_view *v, const contour3d_line_obj_t *ln, o_vec s, o_vec e,
                             o_pixel *canvas) {
  const o_vec ps = ln->converter(s);
  const o_vec pe = ln->converter(e);
  o_vec a, b;
  if (o_project(v, ps, &a)) return;
  if (o_project(v, pe, &b)) return;
  const long W = (long)v->W, H = (long)v->H;
  a.x = (0.5 + a.x) * v->W;
  a.y = (0.5 + a.y) * v->H;
  b.x = (0.5 + b.x) * v->W;
  b.y = (0.5 + b.y) * v->H;
  const long xmin = (long)(fmin(a.x, b.x) - 1);
  const long xmax = (long)(fmax(a.x, b.x) + 1);
  const long ymin = (long)(fmin(a.y, b.y) - 1);
  const long ymax = (long)(fmax(a.y, b.y) + 1);
  const long i0 = o_lmin(W - 1, o_lmax(0, xmin));
  const long i1 = o_lmin(W - 1, o_lmax(0, xmax));
  const long j0 = o_lmin(H - 1, o_lmax(0, ymin));
  const long j1 = o_lmin(H - 1, o_lmax(0, ymax));
  const double la = b.y - a.y;
  const double lb = -b.x + a.x;
  const double lc = (b.x - a.x) * a.y - (b.y - a.y) * a.x;
  for (long j = j0; j <= j1; j++) {
    for (long i = i0; i <= i1; i++) {
      const double di = (double)(size_t)i, dj = (double)(size_t)j;
      const double lim = pow(0.5 * ln->width, 2.);
      const double dist = pow(la * di + lb * dj + lc, 2.) / (la * la + lb * lb);
      if (lim < dist) continue;
      const double prm = -1. * ((b.x - a.x) * (a.x - di) + (b.y - a.y) * (a.y - dj))
                         / (pow(b.x - a.x, 2.) + pow(b.y - a.y, 2.));
      const double depth = 1. / (prm / b.z + (1. - prm) / a.z);
      o_pixel *px = &canvas[(size_t)j * v->W + (size_t)i];
      if (depth < px->depth) continue;
      px->c = ln->color;
      px->depth = depth;
    }
  }
}

/* line.c:121-170; returns nonzero for nitems < 2 */
static int o_raster_line(const o_view *v, const contour3d_line_obj_t *ln, o_pixel *canvas) {
  const size_t np = ln->nitems;
  if (np < 2) return 1;
  const o_vec head = ln->edges[0], tail = ln->edges[1];
  const o_vec delta = o_scale(1. / (np - 1), o_sub(tail, head));
  for (size_t n = 0; n < np - 1; n++) {
    const o_vec s = o_add(head, o_scale((double)(n + 0), delta));
    const o_vec e = o_add(head, o_scale((double)(n + 1), delta));
    o_raster_segment(v, ln, s, e, canvas);
  }
  return 0;
}

/* ---------------------------------------------------------------------- */
/* one contour object, streamed through a 3-slab ring (contour/main.c)     */

typedef struct {
  o_cell *slab[3];
  long ncx, ncy;
} o_ring;

static o_span o_ring_fetch(const void *ctx, long i, long j, long k) {
  const o_ring *r = (const o_ring *)ctx;
  const o_cell *c = &r->slab[k % 3][j * r->ncx + i];
  o_span sp = { c->t, c->n };
  return sp;
}

static void o_cell_corners(const contour3d_contour_obj_t *obj, long i, long j, long k,
                           double val[8], o_vec pos[8]) {
  const size_t nx = obj->glsizes[0], ny = obj->glsizes[1];
  for (int c = 0; c < 8; c++) {
    const long ii = i + (c & 1), jj = j + ((c >> 1) & 1), kk = k + (c >> 2);
    val[c] = obj->array[((size_t)kk * ny + (size_t)jj) * nx + (size_t)ii];
    const o_vec g = { obj->grids[0][ii], obj->grids[1][jj], obj->grids[2][kk] };
    pos[c] = obj->converter(g);
  }
}

static int o_contour(const o_view *v, const contour3d_contour_obj_t *obj, o_pixel *canvas) {
  const long nx = (long)obj->glsizes[0], ny = (long)obj->glsizes[1], nz = (long)obj->glsizes[2];
  if (nx < 2 || ny < 2 || nz < 2) {
    fprintf(stderr, "[CONTOUR3D ERROR] every glsizes entry must be at least 2\n");
    return 1;
  }
  o_ring ring;
  ring.ncx = nx - 1;
  ring.ncy = ny - 1;
  const size_t ncell = (size_t)ring.ncx * (size_t)ring.ncy;
  for (int s = 0; s < 3; s++) {
    ring.slab[s] = (o_cell *)malloc(ncell * sizeof(o_cell));
    if (!ring.slab[s]) {
      for (int q = 0; q < s; q++) free(ring.slab[q]);
      fprintf(stderr, "[CONTOUR3D ERROR] oracle slab allocation failed\n");
      return 1;
    }
  }
  for (long k = 0; k < nz - 1; k++) {
    o_cell *slab = ring.slab[k % 3];
    for (long j = 0; j < ring.ncy; j++) {
      for (long i = 0; i < ring.ncx; i++) {
        double val[8];
        o_vec pos[8];
        o_cell_corners(obj, i, j, k, val, pos);
        o_cell_tessellate(val, pos, obj->threshold, &slab[j * ring.ncx + i]);
      }
    }
    if (k < 2) continue;
    const long kc = k - 1;
    o_cell *mid = ring.slab[kc % 3];
    for (long j = 1; j < ring.ncy - 1; j++) {
      for (long i = 1; i < ring.ncx - 1; i++) {
        o_cell *c = &mid[j * ring.ncx + i];
        for (int t = 0; t < c->n; t++) {
          for (int m = 0; m < 3; m++) {
            c->t[t].vn[m] = o_vertex_normal(&ring, o_ring_fetch, i, j, kc, c->t[t].edge[m]);
          }
        }
      }
    }
    for (long j = 1; j < ring.ncy - 1; j++) {
      for (long i = 1; i < ring.ncx - 1; i++) {
        const o_cell *c = &mid[j * ring.ncx + i];
        for (int t = 0; t < c->n; t++) o_raster_tri(v, &obj->color, &c->t[t], canvas);
      }
    }
  }
  for (int s = 0; s < 3; s++) free(ring.slab[s]);
  return 0;
}

/* ---------------------------------------------------------------------- */
/* exported entry points                                                    */

/* identity coordinate map, src/main.c:12-21 */
contour3d_vector_t oracle_identity_converter(const contour3d_vector_t p) {
  return p;
}

/* contour3d_execute (src/contour3d/main.c:68-146) on one rank.  The image
 * goes to rgb_out (W*H*3, PPM row order: top row = y index H-1) when not
 * NULL, to the file fname when not NULL, and the final z-buffer (row j =
 * y index j) to depth_out when not NULL. */
int oracle_execute(
    const contour3d_vector_t *camera_position,
    const contour3d_vector_t *light_direction,
    const size_t screen_sizes[2],
    const contour3d_vector_t *screen_center,
    const contour3d_vector_t screen_local[2],
    const contour3d_color_t *bg_color,
    size_t num_contours, const contour3d_contour_obj_t *contours,
    size_t num_lines, const contour3d_line_obj_t *lines,
    uint8_t *rgb_out, double *depth_out, const char *fname) {
  o_build_tables();
  o_view v;
  o_view_init(&v, camera_position, light_direction, screen_sizes, screen_center, screen_local);
  const size_t npx = v.W * v.H;
  o_pixel *canvas = (o_pixel *)malloc(npx * sizeof(o_pixel));
  if (!canvas) return 1;
  for (size_t n = 0; n < npx; n++) {
    canvas[n].c = *bg_color;
    canvas[n].depth = -1. * DBL_MAX;
  }
  for (size_t n = 0; n < num_contours; n++) {
    if (o_contour(&v, &contours[n], canvas)) {
      free(canvas);
      return 1;
    }
  }
  for (size_t n = 0; n < num_lines; n++) {
    if (o_raster_line(&v, &lines[n], canvas)) {
      fprintf(stderr, "[CONTOUR3D ERROR] line object needs nitems >= 2\n");
      free(canvas);
      return 1;
    }
  }
  uint8_t *rgb = rgb_out ? rgb_out : (uint8_t *)malloc(npx * 3);
  if (!rgb) {
    free(canvas);
    return 1;
  }
  /* y-flip pack, output.c:113-123 */
  for (size_t j = 0, q = 0; j < v.H; j++) {
    for (size_t i = 0; i < v.W; i++) {
      const o_pixel *px = &canvas[(v.H - j - 1) * v.W + i];
      rgb[q++] = px->c.r;
      rgb[q++] = px->c.g;
      rgb[q++] = px->c.b;
    }
  }
  if (depth_out) {
    for (size_t n = 0; n < npx; n++) depth_out[n] = canvas[n].depth;
  }
  int rc = 0;
  if (fname) {
    FILE *fp = fopen(fname, "w");
    if (!fp) {
      fprintf(stderr, "[CONTOUR3D ERROR] %s: %s\n", fname, strerror(errno));
      rc = 1;
    } else {
      fprintf(fp, "P6\n%zu %zu\n255\n", v.W, v.H);
      if (fwrite(rgb, 3, npx, fp) != npx) rc = 1;
      fclose(fp);
    }
  }
  if (!rgb_out) free(rgb);
  free(canvas);
  return rc;
}

/* ---------------------------------------------------------------------- */
/* whole-field mesh dump for stage parity tests                             */

typedef struct {
  const o_tri *tri;
  const long *start;   /* ncells + 1 */
  long ncx, ncy, ncz;
} o_flat;

static o_span o_flat_fetch(const void *ctx, long i, long j, long k) {
  const o_flat *f = (const o_flat *)ctx;
  const long c = (k * f->ncy + j) * f->ncx + i;
  o_span sp = { f->tri + f->start[c], (int)(f->start[c + 1] - f->start[c]) };
  return sp;
}

/* Tessellate every cell of one field for one threshold and compute vertex
 * normals for every triangle of an interior cell (index 1..nc-2 on each
 * axis; other triangles get NaN vertex normals).  Lattice positions come
 * from the SoA table (cx, cy, cz) when cx is not NULL, else from the grids
 * (identity map).  Triangles are in serial order (cell linear index, then
 * slot).  Any output pointer may be NULL.  Returns the triangle count, or
 * -1 when it exceeds cap (outputs then hold the first cap triangles) or on
 * allocation failure. */
long oracle_mesh(size_t nx, size_t ny, size_t nz,
                 const double *gx, const double *gy, const double *gz,
                 const double *cx, const double *cy, const double *cz,
                 const double *field, double thr,
                 uint8_t *codes, uint8_t *counts, long cap,
                 double *pos, double *fn, double *area, double *vn,
                 int32_t *edge, int64_t *cell, int32_t *slot, uint8_t *interior) {
  o_build_tables();
  if (nx < 2 || ny < 2 || nz < 2) return -1;
  const long ncx = (long)nx - 1, ncy = (long)ny - 1, ncz = (long)nz - 1;
  const long ncells = ncx * ncy * ncz;
  long *start = (long *)malloc((size_t)(ncells + 1) * sizeof(long));
  long tcap = 1024, tn = 0;
  o_tri *tri = (o_tri *)malloc((size_t)tcap * sizeof(o_tri));
  if (!start || !tri) {
    free(start);
    free(tri);
    return -1;
  }
  for (long k = 0; k < ncz; k++) {
    for (long j = 0; j < ncy; j++) {
      for (long i = 0; i < ncx; i++) {
        double val[8];
        o_vec p[8];
        for (int c = 0; c < 8; c++) {
          const size_t ii = (size_t)(i + (c & 1)), jj = (size_t)(j + ((c >> 1) & 1));
          const size_t kk = (size_t)(k + (c >> 2));
          const size_t q = (kk * ny + jj) * nx + ii;
          val[c] = field[q];
          if (cx) {
            o_vec g = { cx[q], cy[q], cz[q] };
            p[c] = g;
          } else {
            o_vec g = { gx[ii], gy[jj], gz[kk] };
            p[c] = g;
          }
        }
        const long cid = (k * ncy + j) * ncx + i;
        o_cell oc;
        o_cell_tessellate(val, p, thr, &oc);
        if (codes) codes[cid] = (uint8_t)o_corner_code(val, thr);
        if (counts) counts[cid] = (uint8_t)oc.n;
        start[cid] = tn;
        if (tn + oc.n > tcap) {
          tcap *= 2;
          o_tri *nt = (o_tri *)realloc(tri, (size_t)tcap * sizeof(o_tri));
          if (!nt) {
            free(start);
            free(tri);
            return -1;
          }
          tri = nt;
        }
        for (int t = 0; t < oc.n; t++) tri[tn++] = oc.t[t];
      }
    }
  }
  start[ncells] = tn;
  o_flat flat = { tri, start, ncx, ncy, ncz };
  const double nan = NAN;
  for (long k = 0; k < ncz; k++) {
    for (long j = 0; j < ncy; j++) {
      for (long i = 0; i < ncx; i++) {
        const long cid = (k * ncy + j) * ncx + i;
        const int in = i >= 1 && i <= ncx - 2 && j >= 1 && j <= ncy - 2 && k >= 1 && k <= ncz - 2;
        for (long t = start[cid]; t < start[cid + 1]; t++) {
          for (int m = 0; m < 3; m++) {
            if (in) {
              tri[t].vn[m] = o_vertex_normal(&flat, o_flat_fetch, i, j, k, tri[t].edge[m]);
            } else {
              o_vec bad = { nan, nan, nan };
              tri[t].vn[m] = bad;
            }
          }
          if (t < cap) {
            if (pos) {
              for (int m = 0; m < 3; m++) {
                pos[t * 9 + m * 3 + 0] = tri[t].p[m].x;
                pos[t * 9 + m * 3 + 1] = tri[t].p[m].y;
                pos[t * 9 + m * 3 + 2] = tri[t].p[m].z;
              }
            }
            if (vn) {
              for (int m = 0; m < 3; m++) {
                vn[t * 9 + m * 3 + 0] = tri[t].vn[m].x;
                vn[t * 9 + m * 3 + 1] = tri[t].vn[m].y;
                vn[t * 9 + m * 3 + 2] = tri[t].vn[m].z;
              }
            }
            if (fn) {
              fn[t * 3 + 0] = tri[t].fn.x;
              fn[t * 3 + 1] = tri[t].fn.y;
              fn[t * 3 + 2] = tri[t].fn.z;
            }
            if (area) area[t] = tri[t].area;
            if (edge) {
              for (int m = 0; m < 3; m++) edge[t * 3 + m] = tri[t].edge[m];
            }
            if (cell) cell[t] = cid;
            if (slot) slot[t] = (int32_t)(t - start[cid]);
            if (interior) interior[t] = (uint8_t)in;
          }
        }
      }
    }
  }
  free(start);
  free(tri);
  return tn > cap ? -1 : tn;
}

/* ---------------------------------------------------------------------- */
/* small kernels exposed for closed-form tests                              */

/* projection of n points; returns the number rejected, rejected entries get
 * NaN.  view = cam(3) center(3) lx(3) ly(3) */
long oracle_project(const double *view, size_t n, const double *pts, double *out) {
  o_view v;
  const o_vec cam = { view[0], view[1], view[2] };
  const o_vec ctr = { view[3], view[4], view[5] };
  const o_vec loc[2] = { { view[6], view[7], view[8] }, { view[9], view[10], view[11] } };
  const o_vec light = { 0., 0., 1. };
  const size_t sz[2] = { 1, 1 };
  o_view_init(&v, &cam, &light, sz, &ctr, loc);
  long rej = 0;
  for (size_t q = 0; q < n; q++) {
    const o_vec p = { pts[3 * q], pts[3 * q + 1], pts[3 * q + 2] };
    o_vec r;
    if (o_project(&v, p, &r)) {
      out[3 * q] = out[3 * q + 1] = out[3 * q + 2] = NAN;
      rej++;
    } else {
      out[3 * q] = r.x;
      out[3 * q + 1] = r.y;
      out[3 * q + 2] = r.z;
    }
  }
  return rej;
}

/* Raster a list of triangles (positions and vertex normals given directly)
 * in order into a fresh canvas; rgb/depth are row j = y index j (unflipped).
 * view = cam(3) center(3) lx(3) ly(3) light(3); colors = 3 bytes per tri. */
void oracle_raster(const double *view, size_t W, size_t H, const uint8_t *bg,
                   size_t ntri, const double *pos, const double *vn, const uint8_t *colors,
                   uint8_t *rgb, double *depth) {
  o_view v;
  const o_vec cam = { view[0], view[1], view[2] };
  const o_vec ctr = { view[3], view[4], view[5] };
  const o_vec loc[2] = { { view[6], view[7], view[8] }, { view[9], view[10], view[11] } };
  const o_vec light = { view[12], view[13], view[14] };
  const size_t sz[2] = { W, H };
  o_view_init(&v, &cam, &light, sz, &ctr, loc);
  o_pixel *canvas = (o_pixel *)malloc(W * H * sizeof(o_pixel));
  if (!canvas) return;
  for (size_t n = 0; n < W * H; n++) {
    canvas[n].c.r = bg[0];
    canvas[n].c.g = bg[1];
    canvas[n].c.b = bg[2];
    canvas[n].depth = -1. * DBL_MAX;
  }
  for (size_t t = 0; t < ntri; t++) {
    o_tri tr;
    for (int m = 0; m < 3; m++) {
      tr.p[m].x = pos[t * 9 + m * 3 + 0];
      tr.p[m].y = pos[t * 9 + m * 3 + 1];
      tr.p[m].z = pos[t * 9 + m * 3 + 2];
      tr.vn[m].x = vn[t * 9 + m * 3 + 0];
      tr.vn[m].y = vn[t * 9 + m * 3 + 1];
      tr.vn[m].z = vn[t * 9 + m * 3 + 2];
    }
    const contour3d_color_t fg = { colors[3 * t], colors[3 * t + 1], colors[3 * t + 2] };
    o_raster_tri(&v, &fg, &tr, canvas);
  }
  for (size_t n = 0; n < W * H; n++) {
    rgb[3 * n + 0] = canvas[n].c.r;
    rgb[3 * n + 1] = canvas[n].c.g;
    rgb[3 * n + 2] = canvas[n].c.b;
    depth[n] = canvas[n].depth;
  }
  free(canvas);
}

/* the derived neighbour table, for the brute-force test: out[e*16 + q*4 + 0..3]
 * = (dk, dj, di, e'), returns counts in n_out[19] */
void oracle_neighbour_table(int32_t *out, int32_t *n_out) {
  o_build_tables();
  for (int e = 0; e < 19; e++) {
    n_out[e] = g_nbr_n[e];
    for (int q = 0; q < g_nbr_n[e]; q++) {
      out[e * 16 + q * 4 + 0] = g_nbr[e][q].dk;
      out[e * 16 + q * 4 + 1] = g_nbr[e][q].dj;
      out[e * 16 + q * 4 + 2] = g_nbr[e][q].di;
      out[e * 16 + q * 4 + 3] = g_nbr[e][q].e;
    }
  }
}

/* ---------------------------------------------------------------------- */
/* user-style coordinate maps for tests of non-identity converters          */

/* cylindrical (r, theta, z) -> Cartesian */
contour3d_vector_t oracle_conv_cylinder(const contour3d_vector_t p) {
  const contour3d_vector_t r = { p.x * cos(p.y), p.x * sin(p.y), p.z };
  return r;
}

/* a shear that moves every lattice point */
contour3d_vector_t oracle_conv_shear(const contour3d_vector_t p) {
  const contour3d_vector_t r = { p.x + 0.25 * p.y, p.y, p.z - 0.125 * p.x };
  return r;
}

/* identity on every point except the last lattice corner of a grid whose
 * last x coordinate is 0.5 (exercises a map that departs late) */
contour3d_vector_t oracle_conv_late(const contour3d_vector_t p) {
  contour3d_vector_t r = p;
  if (p.x == 0.5 && p.y == 0.5 && p.z == 0.5) r.z = 0.5 + 1e-3;
  return r;
}
