// Threshold-crossing events of Vm snapshots on the device: the activation /
// deactivation tables of compute_activation_maps
// (hrdayapy/simulation/functions/activation_maps.py:63-413 and the dV/dt gate
// of hrdayapy/anisotropic_simulation/functions/activation_maps.py:105-119).
//
// One thread per node scans the frames in order (consecutive threads read
// consecutive nodes of a frame: coalesced).  A crossing between frames f and
// f+1 is interpolated linearly in FP64 and stored as FP32; its |dV/dt| over
// the interval is rounded to FP32 and compared with the gate; an event closer
// than min_gap (FP32 difference) to the previous kept one is dropped.  These
// are the host tables' operations in the same precision and order (explicitly
// rounded operations, no fused multiply-add), so both give identical bits.
#include <math.h>

#include "common.h"
#include "hrdaya_hip.h"

namespace hp {

__global__ void __launch_bounds__(256) crossing_events_kernel(
    const float* __restrict__ Vm, const double* __restrict__ t, int64_t n_frames, int64_t n,
    float thr, double thr64, int rising, float min_vel, float min_gap,
    int32_t* __restrict__ counts, float* __restrict__ table, int32_t max_events) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int kept = 0;
  float last = -INFINITY;
  float v0 = Vm[i];
  for (int64_t f = 0; f + 1 < n_frames; ++f) {
    const float v1 = Vm[(f + 1) * n + i];
    const bool a0 = v0 >= thr, a1 = v1 >= thr;
    if (rising ? (!a0 && a1) : (a0 && !a1)) {
      const double d0 = (double)v0, dv = __dsub_rn((double)v1, d0);
      const double t0 = t[f], span = __dsub_rn(t[f + 1], t[f]);
      double frac = __ddiv_rn(__dsub_rn(thr64, d0), dv == 0.0 ? 1e-9 : dv);
      frac = fmin(fmax(frac, 0.0), 1.0);
      const float vel = (float)__ddiv_rn(dv, span == 0.0 ? 1e-9 : span);
      if (!(min_vel > 0.f) || fabsf(vel) >= min_vel) {
        const float tc = (float)__dadd_rn(t0, __dmul_rn(frac, span));
        if (!(min_gap > 0.f) || __fsub_rn(tc, last) >= min_gap) {
          if (table && kept < max_events) table[i * max_events + kept] = tc;
          ++kept;
          last = tc;
        }
      }
    }
    v0 = v1;
  }
  if (counts) counts[i] = kept;
  if (table)
    for (int k = kept; k < max_events; ++k) table[i * max_events + k] = __builtin_nanf("");
}

}  // namespace hp

using namespace hp;

extern "C" {

int hp_crossing_events_f32(const float* Vm, const double* t, int64_t n_frames, int64_t n,
                           double threshold, int rising, double min_vel, double min_gap,
                           int32_t* counts, float* table, int32_t max_events, void* stream) {
  if (!Vm || !t || n_frames < 0 || n < 0 || (!counts && !table) || (table && max_events < 0)) {
    set_error("hp_crossing_events_f32: bad arguments");
    return -1;
  }
  if (n == 0) return 0;
  hipLaunchKernelGGL(crossing_events_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, Vm, t, n_frames, n, (float)threshold, threshold, rising,
                     (float)min_vel, (float)min_gap, counts, table, max_events);
  return check_launch("hp_crossing_events_f32");
}

}  // extern "C"
