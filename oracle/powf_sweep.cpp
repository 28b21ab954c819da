// powf_sweep.cpp — TEST INFRASTRUCTURE ONLY: an exhaustive sweep of the Lp
// kernels' power (turbotops_amd/csrc/powf.hpp, host build of the same
// source) against glibc's powf, which the reference calls
// (include/turbotops/metric.hpp:204, :206), and against powl rounded to
// float: every float x in (0, 1] as a term x^p and every s in [1, 16] as a
// root s^(1/p), for p in {1.5, 1.25, 2.5, 3.7, 7.5, 50, 1.1}; and the
// half-integral terms pow_half_f32 (x^(n + 1/2), n = 1, 2, 3). Prints the
// largest difference in ulps and how many results differ.
//   make -C oracle powf-sweep && oracle/_ref/powf_sweep     (8 threads: ~7 min)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <omp.h>
#include "powf.hpp"
static inline int64_t ulpd(float a, float b) { int32_t ia, ib; memcpy(&ia,&a,4); memcpy(&ib,&b,4); return ia > ib ? ia - ib : ib - ia; }
int main() {
  const float ps[] = {1.5f, 1.25f, 2.5f, 3.7f, 7.5f, 50.0f, 1.1f};
  for (float p : ps) {
    for (int root = 0; root < 2; ++root) {
      const float q = root ? 1.0f / p : p;
      uint32_t lo = root ? 0x3f800000u : 1u, hi = root ? 0x41800000u : 0x3f800000u;
      int64_t maxg = 0, maxc = 0, ng = 0, nc = 0, ngc = 0;
      #pragma omp parallel for reduction(max:maxg,maxc) reduction(+:ng,nc,ngc) schedule(static, 65536)
      for (int64_t t = lo; t <= (int64_t)hi; ++t) {
        float x; uint32_t u = (uint32_t)t; memcpy(&x, &u, 4);
        const float a = ttg::pow_f32(x, q);
        const float g = powf(x, q);
        const float c = (float)powl((long double)x, (long double)q);
        const int64_t dg = ulpd(a, g), dc = ulpd(a, c), dgc = ulpd(g, c);
        if (dg > maxg) maxg = dg; if (dc > maxc) maxc = dc;
        ng += dg != 0; nc += dc != 0; ngc += dgc != 0;
      }
      printf("p=%g %s: vs glibc max %lld ulp, %lld differ; vs powl-rounded max %lld, %lld differ; glibc vs powl %lld differ (of %u)\n",
             p, root ? "root s^(1/p), s in [1,16]" : "terms x^p, x in (0,1]", (long long)maxg, (long long)ng, (long long)maxc, (long long)nc, (long long)ngc, hi - lo + 1);
      fflush(stdout);
    }
  }
  for (unsigned nh = 1; nh <= 3; ++nh) {
    const float p = float(nh) + 0.5f;
    int64_t maxg = 0, ng = 0, nc = 0;
    #pragma omp parallel for reduction(max:maxg) reduction(+:ng,nc) schedule(static, 65536)
    for (int64_t t = 1; t <= 0x3f800000; ++t) {
      float x; uint32_t u = (uint32_t)t; memcpy(&x, &u, 4);
      const float a = ttg::pow_half_f32(x, nh), g = powf(x, p), c = (float)powl((long double)x, (long double)p);
      const int64_t dg = ulpd(a, g);
      if (dg > maxg) maxg = dg;
      ng += dg != 0; nc += ulpd(a, c) != 0;
    }
    printf("p=%g half-integral terms x^p, x in (0,1]: vs glibc max %lld ulp, %lld differ; vs powl-rounded %lld differ\n",
           p, (long long)maxg, (long long)ng, (long long)nc);
    fflush(stdout);
  }
}
