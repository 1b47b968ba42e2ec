// Vector, projection and pixel-coverage arithmetic of the render stage,
// shared by the device kernels and host-side checks (compiled by hipcc for
// the device and by a plain C++ compiler for the CPU tests).  Every
// expression keeps the reference's operation order; builds use
// -ffp-contract=off, so no multiply-add is fused.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__) || defined(__HIP__)
#define C3D_HD __host__ __device__ __forceinline__
#else
#define C3D_HD inline
#endif

namespace c3d {

// Every triangle of a frame takes a 32-bit id (1..ntri, its serial position
// plus one) and every line fragment one after them (ntri + 1 + fragment);
// 0 marks an empty pixel.  A frame holds fewer than 2^32 - 2^24 triangles.
constexpr uint64_t kMaxFrameTriangles = (uint64_t(1) << 32) - (uint64_t(1) << 24);
C3D_HD bool frame_ids_fit(uint64_t ntri, uint64_t nline) {
  return ntri < kMaxFrameTriangles && ntri + nline <= uint64_t(0xffffffffu);
}

using std::ceil;
using std::fabs;
using std::floor;
using std::fmax;
using std::fmin;
using std::sqrt;

template <typename R>
struct V3 {
  R x, y, z;
};

template <typename R>
C3D_HD V3<R> vadd(V3<R> a, V3<R> b) {
  return {a.x + b.x, a.y + b.y, a.z + b.z};
}
template <typename R>
C3D_HD V3<R> vsub(V3<R> a, V3<R> b) {
  return {a.x - b.x, a.y - b.y, a.z - b.z};
}
template <typename R>
C3D_HD V3<R> vscale(R w, V3<R> a) {
  return {w * a.x, w * a.y, w * a.z};
}
template <typename R>
C3D_HD R vdot(V3<R> a, V3<R> b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename R>
C3D_HD V3<R> vcross(V3<R> a, V3<R> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// reciprocal of the norm, then one multiply per component (vector.c:66-76)
template <typename R>
C3D_HD V3<R> vunit(V3<R> a) {
  const R s = R(1) / sqrt(vdot(a, a));
  return {a.x * s, a.y * s, a.z * s};
}

// Camera/screen constants; nc = n.camera, d = n.center, lxx = lx.lx,
// lyy = ly.ly are formed on the host in FP64 exactly as project.c:34-36,
// 62-63 form them per call.
template <typename R>
struct ViewDev {
  R cam[3], ctr[3], lx[3], ly[3], nrm[3], light[3];
  R nc, d, lxx, lyy;
  R fw, fh;   // width, height as reals
  R pitch;    // max(1 / fw, 1 / fh): the larger pixel spacing (set_pitch)
  int32_t W, H;
};

// fills the derived pixel spacing of a view whose W, H, fw, fh are set
template <typename R>
C3D_HD void set_pitch(ViewDev<R>& v) {
  v.pitch = fmax(R(1) / v.fw, R(1) / v.fh);
}

#if defined(__HIP_DEVICE_COMPILE__)
// a / b in FP64 for a divisor b shared by many numerators (a view
// constant).  The compiler's IEEE division is v_div_scale of the
// denominator, v_rcp and two Newton steps on it, then v_div_scale of the
// numerator, a product, an FMA, v_div_fmas and v_div_fixup.  Here the
// denominator half is recip_shared(b), computed once for all numerators,
// and the rest is the same instructions.  The denominator's v_div_scale
// depends on a too: it leaves b unchanged except at extreme exponent
// ratios, which is checked per numerator, and any other case takes the
// plain division.  Every result therefore equals a / b bit for bit.
__device__ __forceinline__ double recip_shared(double b) {
  const double r0 = __builtin_amdgcn_rcp(b);
  const double e0 = __builtin_fma(-b, r0, 1.0);
  const double r1 = __builtin_fma(r0, e0, r0);
  const double e1 = __builtin_fma(-b, r1, 1.0);
  return __builtin_fma(r1, e1, r1);
}

__device__ __forceinline__ double div_shared(double a, double b, double rb) {
  bool fd, fn;
  if (__builtin_amdgcn_div_scale(a, b, false, &fd) != b) return a / b;
  const double n = __builtin_amdgcn_div_scale(a, b, true, &fn);
  const double q0 = n * rb;
  const double rem = __builtin_fma(-b, q0, n);
  return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(rem, rb, q0, fn), b, a);
}
#endif

// a / b, b a view constant (see div_shared)
template <typename R>
C3D_HD R div_view(R a, R b) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(R) == 8) return div_shared(a, b, recip_shared(b));
#endif
  return a / b;
}

// perspective projection (project.c:9-67); false when rejected
template <typename R>
C3D_HD bool project(const ViewDev<R>& v, V3<R> p, V3<R>& q) {
  const V3<R> cam = {v.cam[0], v.cam[1], v.cam[2]};
  const V3<R> nrm = {v.nrm[0], v.nrm[1], v.nrm[2]};
  const V3<R> ray = vsub(p, cam);
  const R nr = vdot(nrm, ray);
  if (R(-1.e-8) < nr && nr < R(1.e-8)) return false;
  const R t = (v.d - v.nc) / nr;
  if (t <= R(0) || R(1) <= t) return false;
  const V3<R> delta = {v.cam[0] + t * ray.x - v.ctr[0], v.cam[1] + t * ray.y - v.ctr[1],
                       v.cam[2] + t * ray.z - v.ctr[2]};
  const V3<R> lx = {v.lx[0], v.lx[1], v.lx[2]};
  const V3<R> ly = {v.ly[0], v.ly[1], v.ly[2]};
  q.x = div_view(vdot(delta, lx), v.lxx);
  q.y = div_view(vdot(delta, ly), v.lyy);
  q.z = nr;
  return true;
}

// 2-D edge function (render.c:38-46)
template <typename R>
C3D_HD R edge_fn(const V3<R>& a, const V3<R>& b, R cx, R cy) {
  return (cx - a.x) * (b.y - a.y) - (cy - a.y) * (b.x - a.x);
}

// truncation toward zero of a bbox bound; the bound only matters after
// clamping to [0, W-1], so it is first clamped into int range
template <typename R>
C3D_HD int32_t trunc_i(R x) {
  x = fmin(fmax(x, R(-1.e9)), R(1.e9));
  return int32_t(x);
}

C3D_HD int32_t clampi(int32_t lo, int32_t x, int32_t hi) { return x < lo ? lo : (x > hi ? hi : x); }

// pixel-centre coordinate of column i (render.c:91-94)
template <typename R>
C3D_HD R pixel_centre(int32_t i, R extent) {
  return R(1) * (R(i) + R(0.5)) / extent - R(0.5);
}

// A projected triangle: vertices, 1 / (clamped) doubled area and the pixel
// box its coverage test runs over.
template <typename R>
struct Projected {
  V3<R> q[3];
  R ainv;
  int32_t i0, i1, j0, j1;
};

// The reference's bounding box and area of projected vertices P.q
// (render.c:60-89); returns the unclamped doubled area.
template <typename R>
C3D_HD R raster_box(const ViewDev<R>& v, Projected<R>& P) {
  const V3<R>* q = P.q;
  const int32_t xmin = trunc_i<R>((R(0.5) + fmin(q[0].x, fmin(q[1].x, q[2].x))) * v.fw - R(1));
  const int32_t xmax = trunc_i<R>((R(0.5) + fmax(q[0].x, fmax(q[1].x, q[2].x))) * v.fw + R(1));
  const int32_t ymin = trunc_i<R>((R(0.5) + fmin(q[0].y, fmin(q[1].y, q[2].y))) * v.fh - R(1));
  const int32_t ymax = trunc_i<R>((R(0.5) + fmax(q[0].y, fmax(q[1].y, q[2].y))) * v.fh + R(1));
  P.i0 = clampi(0, xmin, v.W - 1);
  P.i1 = clampi(0, xmax, v.W - 1);
  P.j0 = clampi(0, ymin, v.H - 1);
  P.j1 = clampi(0, ymax, v.H - 1);
  const R area = edge_fn(q[0], q[1], q[2].x, q[2].y);
  P.ainv = R(1) / (area < R(0) ? fmin(area, R(-1.e-16)) : fmax(area, R(1.e-16)));
  return area;
}

// barycentric weights at the pixel centre (cx, cy); false when outside
// (render.c:91-115)
template <typename R>
C3D_HD bool pixel_weights(const Projected<R>& P, R cx, R cy, R& w0, R& w1, R& w2) {
  w0 = P.ainv * edge_fn(P.q[1], P.q[2], cx, cy);
  w1 = P.ainv * edge_fn(P.q[2], P.q[0], cx, cy);
  w2 = P.ainv * edge_fn(P.q[0], P.q[1], cx, cy);
  return !(w0 < R(-1.e-8) || w1 < R(-1.e-8) || w2 < R(-1.e-8));
}

// harmonic depth (render.c:120-124)
template <typename R>
C3D_HD R pixel_depth(const Projected<R>& P, R w0, R w1, R w2) {
  return R(1) / (w0 / P.q[0].z + w1 / P.q[1].z + w2 / P.q[2].z);
}

// Narrows P's box to the pixels whose centres pixel_weights can accept,
// so the coverage loop skips the reference box's margin.  Needs the
// P.ainv of raster_box and v.pitch (set_pitch).
//
// Bound (u the unit roundoff, g = 4.5u >= 4u / (1 - 4u)).  The computed
// test accepts p only if, for every edge i, ainv * e_i >= -1e-8 with e_i
// its computed edge function; with |A| the computed doubled area (not
// clamped: |A| > 1e-14) that gives s * e_i >= -1e-8 |A| (1 + 4u), s the
// sign of A.  Each e_i is a difference of two products of differences,
// so it is within g (|cx - ax| |by - ay| + |cy - ay| |bx - ax|) <=
// g sqrt(2) D L_i of the exact value (D >= the distance of any reference
// box centre from a vertex, L_i the edge's length), and so is the area
// within 2 g emax^2 (emax the largest coordinate difference along an
// edge).  When |A| > 4 g emax^2 the computed area has the exact sign, and
// an accepted p lies within d_i = 1e-8 |A| (1 + 4u) / L_i + g sqrt(2) D
// outside every edge: inside the triangle whose edges are pushed out by
// d_i, whose vertex between edges j and k moves by at most
// (d_j + d_k) L_j L_k / |A| <= (4.04e-8 Lmax |A| + 2 sqrt(2) g D Lmax^2) / |A|,
// Lmax <= sqrt(2) emax.  So with M = 5.8e-8 emax + 12 g D emax^2 / |A|
// (+ the rounding of the comparison bounds), every accepted centre lies
// within M of the vertices' box.  Tiny, collinear and non-finite
// triangles keep the reference box.
//
// Columns: an accepted centre has pixel_centre(i) >= xlo, and with
// u = (xlo + 0.5) fw - 0.5 the rounding of u and of pixel_centre stays
// below delta = 4 eps (|u| + fw + 2) index units, so i >= u - delta: the
// box starts at ceil(u - delta) (when u lies within delta of an integer it
// may hold one column more than needed, never one less).  Same for the
// other bounds.  No division is needed.
template <typename R>
C3D_HD void tight_box(const ViewDev<R>& v, R area, Projected<R>& P) {
  const R eps = sizeof(R) == 8 ? R(2.220446049250313e-16) : R(1.1920929e-7);
  const R g = R(2.25) * eps;  // 4.5 u
  const V3<R>* q = P.q;
  const R minx = fmin(q[0].x, fmin(q[1].x, q[2].x)), maxx = fmax(q[0].x, fmax(q[1].x, q[2].x));
  const R miny = fmin(q[0].y, fmin(q[1].y, q[2].y)), maxy = fmax(q[0].y, fmax(q[1].y, q[2].y));
  const R pitch = v.pitch;
  const R D = fmax(maxx - minx, maxy - miny) + R(4) * pitch;
  const R emax = fmax(fmax(fmax(fabs(q[1].x - q[0].x), fabs(q[1].y - q[0].y)),
                           fmax(fabs(q[2].x - q[1].x), fabs(q[2].y - q[1].y))),
                      fmax(fabs(q[0].x - q[2].x), fabs(q[0].y - q[2].y)));
  const R A = fabs(area);
  if (!(A > R(1.e-14)) || !(emax > R(0))) return;
  const R e2 = emax * emax;
  if (!(A > R(4.1) * g * e2)) return;
  // |P.ainv| is 1 / A rounded once; 1 + 2^-20 covers the products' rounding
  const R qmax = fmax(fmax(fmax(fabs(minx), fabs(maxx)), fmax(fabs(miny), fabs(maxy))), R(1));
  const R M = (R(5.8e-8) * emax + R(12) * g * D * e2 * fabs(P.ainv)) * R(1.00000095367431640625) + R(16) * eps * qmax;
  if (!(M < pitch)) return;
  const R ulo = (minx - M + R(0.5)) * v.fw - R(0.5), uhi = (maxx + M + R(0.5)) * v.fw - R(0.5);
  const R vlo = (miny - M + R(0.5)) * v.fh - R(0.5), vhi = (maxy + M + R(0.5)) * v.fh - R(0.5);
  const R dx = R(4) * eps * (v.fw + R(2)), dy = R(4) * eps * (v.fh + R(2));
  const int32_t a = clampi(P.i0, trunc_i<R>(ceil(ulo - (dx + R(4) * eps * fabs(ulo)))), P.i1 + 1);
  const int32_t b = clampi(a - 1, trunc_i<R>(floor(uhi + (dx + R(4) * eps * fabs(uhi)))), P.i1);
  const int32_t c = clampi(P.j0, trunc_i<R>(ceil(vlo - (dy + R(4) * eps * fabs(vlo)))), P.j1 + 1);
  const int32_t d = clampi(c - 1, trunc_i<R>(floor(vhi + (dy + R(4) * eps * fabs(vhi)))), P.j1);
  P.i0 = a;
  P.i1 = b;
  P.j0 = c;
  P.j1 = d;
}

}  // namespace c3d
