// GPU check of div_shared (contour3d_amd/csrc/c3d_geom.h): the FP64 division
// by a view constant that shares the denominator half of the division
// sequence must equal the compiler's a / b bit for bit.  2^25 numerators
// (any bit pattern; the projections' range; every exponent from the
// subnormals to the largest; zeros, subnormals, infinities, NaN, extremes)
// against 24 divisors (view-like values, random values over 2^+-100 and
// values near the ends of the range).  Prints the counts; exit status 1 on
// any mismatch.  Built by __graft_entry__.build() next to this file.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>
#include "c3d_geom.h"

__global__ void kcheck(const double* a, size_t n, const double* bs, int nb, unsigned long long* bad,
                       unsigned long long* fast) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
#if defined(__HIP_DEVICE_COMPILE__)
  const double x = a[i];
  for (int k = 0; k < nb; k++) {
    const double b = bs[k];
    const double r = c3d::recip_shared(b);
    const double q1 = c3d::div_shared(x, b, r);
    const double q2 = x / b;
    uint64_t u1, u2;
    memcpy(&u1, &q1, 8);
    memcpy(&u2, &q2, 8);
    const bool nan1 = q1 != q1, nan2 = q2 != q2;
    if (!(u1 == u2 || (nan1 && nan2))) atomicAdd(bad, 1ull);
    bool fd;
    if (__builtin_amdgcn_div_scale(x, b, false, &fd) == b) atomicAdd(fast, 1ull);
  }
#endif
}

int main() {
  const size_t n = size_t(1) << 25;
  std::vector<double> a(n);
  std::mt19937_64 g(7);
  for (size_t i = 0; i < n; i++) {
    uint64_t u = g();
    double x;
    switch (i % 4) {
      case 0: memcpy(&x, &u, 8); break;                                        // any bit pattern
      case 1: x = std::ldexp(double(u >> 11) / 9007199254740992.0, int(g() % 120) - 60); break;  // view range
      case 2: x = (u & 1 ? -1.0 : 1.0) * std::ldexp(1.0 + double(u >> 12) / 4503599627370496.0, int(g() % 2098) - 1074); break;
      default: { const double sp[] = {0.0, -0.0, 1e-310, -4.9e-324, 1e308, -1.7e308, INFINITY, -INFINITY, NAN, 1.0, 3.0};
                 x = sp[u % 11]; }
    }
    a[i] = x;
  }
  std::vector<double> bs = {1.0, 3.0, 0.7071067811865476, 1.3333333333333333, 1920.0, 1080.0, 2.5e-5, 7.3e4,
                            std::ldexp(1.0, -1000), std::ldexp(1.5, 1000), std::ldexp(1.0, -1022), 1e-300, 1e300};
  for (int k = 0; k < 11; k++) bs.push_back(std::ldexp(1.0 + double(g() >> 12) / 4503599627370496.0, int(g() % 200) - 100));
  double *da, *db;
  unsigned long long *dbad, *dfast;
  if (hipMalloc(&da, n * 8) || hipMalloc(&db, bs.size() * 8) || hipMalloc(&dbad, 8) || hipMalloc(&dfast, 8)) return 2;
  hipMemcpy(da, a.data(), n * 8, hipMemcpyHostToDevice);
  hipMemcpy(db, bs.data(), bs.size() * 8, hipMemcpyHostToDevice);
  hipMemset(dbad, 0, 8);
  hipMemset(dfast, 0, 8);
  const unsigned grid = unsigned((n + 255) / 256);
  hipLaunchKernelGGL(kcheck, dim3(grid), dim3(256), 0, 0, da, n, db, int(bs.size()), dbad, dfast);
  unsigned long long bad = 0, fast = 0;
  if (hipDeviceSynchronize() != hipSuccess) return 3;
  hipMemcpy(&bad, dbad, 8, hipMemcpyDeviceToHost);
  hipMemcpy(&fast, dfast, 8, hipMemcpyDeviceToHost);
  printf("divisions %zu mismatches %llu shared-path %llu\n", n * bs.size(), bad, fast);
  return bad ? 1 : 0;
}
