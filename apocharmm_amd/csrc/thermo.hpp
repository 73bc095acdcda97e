// Kinetic-energy / temperature reduction of the integrator's step, shared by
// the integrator's own launches and the coordinate sort of the next force
// call, which runs it in one extra block (the step's partials are complete
// by then in stream order, so no fence or extra launch is needed).
#pragma once

#include "common.hpp"

namespace apo {

constexpr double kBoltz = 1.987191e-03;  // kcal/mol/K (include/Constants.h:17)

// Fixed-order kinetic energy sum, temperatures and their running means
// (src/CudaLangevinThermostatIntegrator.cu:684-701), step counter + 1.
// thermo: [0] ke (three-estimate mean) [1] ke (on-step) [2] T [3] T on-step
//         [4] <T> [5] <T on-step> [6] max displacement since the last list
//         build (A) [7] max displacement of one step since that build (A);
// counters: [0] step [1] averaging window [2] number of the list build that
// [6], [7] refer to (the build's meta[16]).  [6], [7], the step count and the
// build number also go to the host's pinned hdisp[0..3] (step count last).
__device__ __forceinline__ void lt_finalize_body(const double2 *__restrict__ part, int npart,
                                                 double ndof, double *__restrict__ thermo,
                                                 unsigned long long *__restrict__ counters,
                                                 const float2 *__restrict__ dpart,
                                                 const int *__restrict__ meta,
                                                 double *__restrict__ hdisp,
                                                 unsigned long long *__restrict__ pend) {
  // pend (the integrator's own steps): the step kernel raises it when it has
  // written partials, and only a raised flag is reduced (then lowered), so a
  // reduction merged into a later launch never counts a step twice
  if (pend && *pend == 0ull) return;
  __shared__ double s0[256], s1[256];
  __shared__ float m0[256], m1[256];
  double a = 0.0, b = 0.0;
  float d2 = 0.f, s2 = 0.f;
  for (int k = threadIdx.x; k < npart; k += 256) {
    a += part[k].x;
    b += part[k].y;
    d2 = fmaxf(d2, dpart[k].x);
    s2 = fmaxf(s2, dpart[k].y);
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = b;
  m0[threadIdx.x] = d2;
  m1[threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
      m0[threadIdx.x] = fmaxf(m0[threadIdx.x], m0[threadIdx.x + w]);
      m1[threadIdx.x] = fmaxf(m1[threadIdx.x], m1[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const unsigned long long seq = (unsigned long long)meta[16];
    if (counters[2] != seq) {
      counters[2] = seq;
      thermo[6] = 0.0;
      thermo[7] = 0.0;
    }
    thermo[6] = fmax(thermo[6], sqrt((double)m0[0]));
    thermo[7] = fmax(thermo[7], sqrt((double)m1[0]));
    hdisp[0] = thermo[6];
    hdisp[1] = thermo[7];
    hdisp[3] = (double)seq;
    __threadfence_system();
    hdisp[2] = (double)(counters[0] + 1);
    __threadfence_system();
    const double t0 = s0[0] / (0.5 * ndof * kBoltz), t1 = s1[0] / (0.5 * ndof * kBoltz);
    const double cnt = (double)(counters[1] + 1);
    thermo[0] = s0[0];
    thermo[1] = s1[0];
    thermo[2] = t0;
    thermo[3] = t1;
    thermo[4] += (t0 - thermo[4]) / cnt;
    thermo[5] += (t1 - thermo[5]) / cnt;
    counters[0] += 1;
    counters[1] += 1;
    if (pend) *pend = 0ull;
  }
}
struct LtFinArgs {
  const double2 *part;
  int npart;
  double ndof;
  double *thermo;
  unsigned long long *counters;
  const float2 *dpart;
  const int *meta;
  double *hdisp;
  unsigned long long *pend;  // NULL: reduce unconditionally
};

}  // namespace apo
