/* Drives the CPU oracle (oracle/contour3d_oracle.c) through its two entry
 * points on small fields, lines and edge cases, for a build with the
 * address and undefined-behaviour sanitizers: any out-of-bounds access,
 * leak or undefined operation aborts with the sanitizer's report.  Exits 0
 * after printing the triangle and covered-pixel counts. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "contour3d.h"

long oracle_mesh(size_t nx, size_t ny, size_t nz, const double *gx, const double *gy, const double *gz,
                 const double *cx, const double *cy, const double *cz, const double *field, double thr,
                 uint8_t *codes, uint8_t *counts, long cap, double *pos, double *fn, double *area, double *vn,
                 int32_t *edge, int64_t *cell, int32_t *slot, uint8_t *interior);
int oracle_execute(const contour3d_vector_t *camera_position, const contour3d_vector_t *light_direction,
                   const size_t screen_sizes[2], const contour3d_vector_t *screen_center,
                   const contour3d_vector_t screen_local[2], const contour3d_color_t *bg_color,
                   size_t num_contours, const contour3d_contour_obj_t *contours, size_t num_lines,
                   const contour3d_line_obj_t *lines, uint8_t *rgb, double *depth, const char *fname);
contour3d_vector_t oracle_identity_converter(const contour3d_vector_t p);

static double *grid(size_t n) {
  double *g = malloc(n * sizeof(double));
  for (size_t m = 0; m < n; m++) g[m] = -0.5 + (m + 0.5) / (double)n;
  return g;
}

static long mesh_case(size_t nx, size_t ny, size_t nz, const double *f, double thr) {
  double *gx = grid(nx), *gy = grid(ny), *gz = grid(nz);
  const size_t nc = (nx - 1) * (ny - 1) * (nz - 1);
  uint8_t *codes = malloc(nc), *counts = malloc(nc);
  oracle_mesh(nx, ny, nz, gx, gy, gz, NULL, NULL, NULL, f, thr, codes, counts, 0, NULL, NULL, NULL, NULL, NULL,
              NULL, NULL, NULL);
  long T = 0;
  for (size_t c = 0; c < nc; c++) T += counts[c];
  double *pos = malloc((T + 1) * 9 * sizeof(double)), *fn = malloc((T + 1) * 3 * sizeof(double));
  double *area = malloc((T + 1) * sizeof(double)), *vn = malloc((T + 1) * 9 * sizeof(double));
  int32_t *edge = malloc((T + 1) * 3 * sizeof(int32_t)), *slot = malloc((T + 1) * sizeof(int32_t));
  int64_t *cell = malloc((T + 1) * sizeof(int64_t));
  uint8_t *in = malloc(T + 1);
  const long r = oracle_mesh(nx, ny, nz, gx, gy, gz, NULL, NULL, NULL, f, thr, codes, counts, T, pos, fn, area, vn,
                             edge, cell, slot, in);
  if (r != T) {
    printf("mesh returned %ld, expected %ld\n", r, T);
    exit(1);
  }
  free(gx), free(gy), free(gz), free(codes), free(counts), free(pos), free(fn), free(area), free(vn), free(edge);
  free(slot), free(cell), free(in);
  return T;
}

int main(void) {
  const size_t n = 14;
  double *f = malloc(n * n * n * sizeof(double));
  for (size_t k = 0; k < n; k++)
    for (size_t j = 0; j < n; j++)
      for (size_t i = 0; i < n; i++)
        f[(k * n + j) * n + i] = sin(0.9 * i + 0.3) * cos(0.7 * j) + 0.5 * sin(0.5 * k + 0.2 * i);
  long T = 0;
  T += mesh_case(n, n, n, f, 0.1);
  T += mesh_case(n, n, n, f, -0.4);
  T += mesh_case(2, n, n, f, 0.0);  /* one cell thick */
  double *t2 = calloc(8, sizeof(double));
  T += mesh_case(2, 2, 2, t2, 0.0); /* one cell, all corners equal: no crossing */
  /* a full frame with two contours and lines */
  double *gx = grid(n), *gy = grid(n), *gz = grid(n);
  contour3d_contour_obj_t objs[2];
  for (int q = 0; q < 2; q++) {
    objs[q].pencil = SDECOMP_X1PENCIL;
    objs[q].glsizes[0] = objs[q].glsizes[1] = objs[q].glsizes[2] = n;
    objs[q].grids[0] = gx;
    objs[q].grids[1] = gy;
    objs[q].grids[2] = gz;
    objs[q].converter = oracle_identity_converter;
    objs[q].threshold = q ? 0.3 : -0.2;
    objs[q].color = (contour3d_color_t){(uint8_t)(200 - 100 * q), 100, (uint8_t)(50 + 150 * q)};
    objs[q].array = f;
  }
  contour3d_line_obj_t lines[2] = {
      {{{-0.5, -0.5, -0.5}, {-0.5, 0.5, -0.5}}, 12, oracle_identity_converter, {255, 255, 255}, 3.0},
      {{{-0.5, -0.5, -0.5}, {0.5, 0.5, 0.5}}, 2, oracle_identity_converter, {255, 0, 0}, 1.5}};
  const contour3d_vector_t cam = {0.866, -1.5, 1.0}, light = {0.0, 2.0, -1.0}, ctr = {0.433, -0.75, 0.5};
  const contour3d_vector_t loc[2] = {{1.1547, 0.6667, 0.0}, {-0.25, 0.433, 0.866}};
  const size_t sizes[2] = {97, 61};
  const contour3d_color_t bg = {1, 2, 3};
  uint8_t *rgb = malloc(sizes[0] * sizes[1] * 3);
  double *dep = malloc(sizes[0] * sizes[1] * sizeof(double));
  if (oracle_execute(&cam, &light, sizes, &ctr, loc, &bg, 2, objs, 2, lines, rgb, dep, NULL)) {
    printf("execute failed\n");
    return 1;
  }
  long covered = 0;
  for (size_t p = 0; p < sizes[0] * sizes[1]; p++) covered += dep[p] > -1e300;
  printf("triangles %ld covered %ld\n", T, covered);
  free(f), free(t2), free(gx), free(gy), free(gz), free(rgb), free(dep);
  return 0;
}
