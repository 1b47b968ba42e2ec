/* A solver-style client of the drop-in API: builds a blob field on a
 * uniform lattice and calls contour3d_execute exactly as the reference's
 * example does (src/main.c:310-326), then exits with its return code.
 * usage: solver <n> <out.ppm> <field.raw (n^3 float64, x fastest)> */
#include <stdio.h>
#include <stdlib.h>
#include "contour3d.h"

static contour3d_vector_t identity(const contour3d_vector_t p) { return p; }

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  const size_t n = (size_t)atoi(argv[1]);
  double *f = malloc(n * n * n * sizeof(double));
  double *g = malloc(n * sizeof(double));
  FILE *fp = fopen(argv[3], "rb");
  if (!f || !g || !fp || fread(f, sizeof(double), n * n * n, fp) != n * n * n) return 3;
  fclose(fp);
  for (size_t q = 0; q < n; q++) g[q] = -0.5 + (q + 0.5) / n;
  const contour3d_vector_t cam = { 8.660254037844386e-1, -1.5, 1.0 };
  const contour3d_vector_t look = { 0., 0., 0. };
  const contour3d_vector_t ctr = { 4.330127018922194e-1, -7.500000000000003e-1, 5.000000000000003e-1 };
  const contour3d_vector_t local[2] = {
    { 1.154700538379251e+0, 6.666666666666665e-1, 0. },
    { -2.5e-1, 4.330127018922195e-1, 8.660254037844386e-1 },
  };
  const contour3d_vector_t light = { 0., 2., -1. };
  const size_t sizes[2] = { 160, 120 };
  const contour3d_color_t bg = { 0, 0, 0 };
  contour3d_contour_obj_t objs[2];
  for (int q = 0; q < 2; q++) {
    objs[q].pencil = SDECOMP_X1PENCIL;
    objs[q].glsizes[0] = objs[q].glsizes[1] = objs[q].glsizes[2] = n;
    objs[q].grids[0] = objs[q].grids[1] = objs[q].grids[2] = g;
    objs[q].converter = identity;
    objs[q].array = f;
  }
  objs[0].threshold = 0.4;
  objs[0].color = (contour3d_color_t){ 0, 255, 255 };
  objs[1].threshold = 0.9;
  objs[1].color = (contour3d_color_t){ 255, 255, 0 };
  const contour3d_line_obj_t line = { { { -0.5, -0.5, -0.5 }, { 0.5, -0.5, -0.5 } }, 32, identity,
                                      { 255, 255, 255 }, 3.0 };
  const int rc = contour3d_execute(NULL, &cam, &look, &light, sizes, &ctr, local, &bg, 2, objs, 1, &line,
                                   argv[2]);
  free(f);
  free(g);
  return rc;
}
