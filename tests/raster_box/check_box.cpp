// Brute-force check of c3d::tight_box: for random, thin, degenerate and
// huge triangles, every pixel of the reference box that the coverage test
// (c3d::pixel_weights) accepts must lie in the narrowed box.  Runs the same
// header the device kernels use, in FP64 and FP32.
// Usage: check_box <seed> <triangles per kind>; prints stats, exits 1 on a miss.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "c3d_geom.h"

using namespace c3d;

template <typename R>
struct Stats {
  long tris = 0, narrowed = 0, ref_cand = 0, tight_cand = 0, accepted = 0, misses = 0;
};

template <typename R>
static void check_one(const ViewDev<R>& v, const std::vector<R>& ax, const V3<R> q[3], Stats<R>& st) {
  Projected<R> P;
  for (int n = 0; n < 3; n++) P.q[n] = q[n];
  const R area = raster_box(v, P);
  Projected<R> T = P;
  tight_box(v, area, T);
  st.tris++;
  const long refc = long(P.i1 - P.i0 + 1) * long(P.j1 - P.j0 + 1);
  const long tc = (T.i1 >= T.i0 && T.j1 >= T.j0) ? long(T.i1 - T.i0 + 1) * long(T.j1 - T.j0 + 1) : 0;
  st.ref_cand += refc;
  st.tight_cand += tc;
  if (tc < refc) st.narrowed++;
  if (refc > 4096) return;  // brute force only boxes of reasonable size
  for (int32_t j = P.j0; j <= P.j1; j++) {
    for (int32_t i = P.i0; i <= P.i1; i++) {
      R w0, w1, w2;
      if (!pixel_weights(P, ax[i], ax[v.W + j], w0, w1, w2)) continue;
      st.accepted++;
      if (i < T.i0 || i > T.i1 || j < T.j0 || j > T.j1) {
        if (st.misses < 10)
          std::printf("MISS %s pixel (%d,%d) tight [%d,%d]x[%d,%d] q=(%.17g,%.17g) (%.17g,%.17g) (%.17g,%.17g)\n",
                      sizeof(R) == 8 ? "f64" : "f32", i, j, T.i0, T.i1, T.j0, T.j1, double(q[0].x),
                      double(q[0].y), double(q[1].x), double(q[1].y), double(q[2].x), double(q[2].y));
        st.misses++;
      }
    }
  }
}

template <typename R>
static Stats<R> run(unsigned seed, long per_kind) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  Stats<R> st;
  Stats<R> kinds[7];
  for (int view = 0; view < 6; view++) {
    ViewDev<R> v{};
    const int32_t sizes[6][2] = {{1920, 1080}, {64, 48}, {97, 211}, {800, 600}, {1, 1}, {3840, 2160}};
    v.W = sizes[view][0];
    v.H = sizes[view][1];
    v.fw = R(double(v.W));
    v.fh = R(double(v.H));
    set_pitch(v);
    std::vector<R> ax(v.W + v.H);
    for (int32_t x = 0; x < v.W + v.H; x++)
      ax[x] = x < v.W ? pixel_centre<R>(x, v.fw) : pixel_centre<R>(x - v.W, v.fh);
    const double pw = 1.0 / v.W, ph = 1.0 / v.H;
    for (int kind = 0; kind < 7; kind++) {
      for (long n = 0; n < per_kind; n++) {
        V3<R> q[3];
        const double cx = -0.6 + 1.2 * U(rng), cy = -0.6 + 1.2 * U(rng);
        double s = pw;
        switch (kind) {
          case 0: s = pw * (0.05 + 4 * U(rng)); break;             // pixel-sized
          case 1: s = pw * std::pow(10.0, -3 + 6 * U(rng)); break; // any scale
          case 2: s = pw * (0.5 + 3 * U(rng)); break;              // thin, below
          case 3: s = pw * (0.5 + 3 * U(rng)); break;              // collinear
          case 4: s = pw * (0.5 + 3 * U(rng)); break;              // repeated vertex
          case 5: s = pw * (0.3 + 2 * U(rng)); break;              // on pixel centres
          default: s = pw * (0.5 + 8 * U(rng)); break;             // far-away huge values
        }
        double p[3][2];
        for (int k = 0; k < 3; k++) {
          p[k][0] = cx + s * (U(rng) - 0.5) * 2;
          p[k][1] = cy + s * (U(rng) - 0.5) * 2 * (ph / pw);
        }
        if (kind == 2) {  // p2 pushed towards the line p0-p1 by a random power of ten
          const double t = U(rng), f = std::pow(10.0, -14 * U(rng));
          const double lx = p[0][0] + t * (p[1][0] - p[0][0]), ly = p[0][1] + t * (p[1][1] - p[0][1]);
          p[2][0] = lx + f * (p[2][0] - lx);
          p[2][1] = ly + f * (p[2][1] - ly);
        } else if (kind == 3) {
          const double t = -0.5 + 2 * U(rng);
          p[2][0] = p[0][0] + t * (p[1][0] - p[0][0]);
          p[2][1] = p[0][1] + t * (p[1][1] - p[0][1]);
        } else if (kind == 4) {
          p[2][0] = p[0][0];
          p[2][1] = p[0][1];
        } else if (kind == 5) {  // vertices exactly on pixel centres
          for (int k = 0; k < 3; k++) {
            const int32_t i = int32_t((p[k][0] + 0.5) * v.W), j = int32_t((p[k][1] + 0.5) * v.H);
            if (i >= 0 && i < v.W) p[k][0] = double(ax[i]);
            if (j >= 0 && j < v.H) p[k][1] = double(ax[v.W + j]);
          }
        } else if (kind == 6) {
          const double big = std::pow(10.0, 1 + 8 * U(rng));
          p[0][0] = cx + big * (U(rng) - 0.5);
        }
        for (int k = 0; k < 3; k++) {
          q[k].x = R(p[k][0]);
          q[k].y = R(p[k][1]);
          q[k].z = R(1);
        }
        check_one(v, ax, q, kinds[kind]);
      }
    }
  }
  static const char* names[7] = {"pixel-sized", "any scale", "thin", "collinear", "repeated vertex",
                                 "on pixel centres", "huge coordinate"};
  for (int k = 0; k < 7; k++) {
    const Stats<R>& s = kinds[k];
    std::printf("  %s %-17s candidates per triangle %7.2f -> %7.2f, accepted %.2f, misses %ld\n",
                sizeof(R) == 8 ? "f64" : "f32", names[k], double(s.ref_cand) / s.tris,
                double(s.tight_cand) / s.tris, double(s.accepted) / s.tris, s.misses);
    st.tris += s.tris;
    st.narrowed += s.narrowed;
    st.ref_cand += s.ref_cand;
    st.tight_cand += s.tight_cand;
    st.accepted += s.accepted;
    st.misses += s.misses;
  }
  return st;
}

int main(int argc, char** argv) {
  const unsigned seed = argc > 1 ? unsigned(std::atoi(argv[1])) : 1u;
  const long per = argc > 2 ? std::atol(argv[2]) : 20000;
  const Stats<double> a = run<double>(seed, per);
  const Stats<float> b = run<float>(seed + 1000, per);
  std::printf("f64 tris %ld narrowed %ld candidates %ld -> %ld accepted %ld misses %ld\n", a.tris, a.narrowed,
              a.ref_cand, a.tight_cand, a.accepted, a.misses);
  std::printf("f32 tris %ld narrowed %ld candidates %ld -> %ld accepted %ld misses %ld\n", b.tris, b.narrowed,
              b.ref_cand, b.tight_cand, b.accepted, b.misses);
  return (a.misses || b.misses) ? 1 : 0;
}
