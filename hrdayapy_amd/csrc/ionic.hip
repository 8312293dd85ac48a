// TTP06 reaction step over a strided node set (the ionic plugin seam).
//
// Replaces the ~150-launch eager op chain of _currents_and_gates
// (hrdayapy/simulation/functions/ionic_models/gpu_ttp06.py:198-412) by one
// kernel: each lane loads its node's 18 state variables, runs the whole cell
// model in registers and writes the state back once.
#include "common.h"
#include "hrdaya_hip.h"
#include "ttp06_cell.h"

namespace hp {

template <typename T>
__global__ void __launch_bounds__(256) ttp06_strided_kernel(
    const T* __restrict__ V, T* __restrict__ state, T* __restrict__ I_ion,
    const uint8_t* __restrict__ celltype, int fixed_celltype, const T* __restrict__ gks_scale,
    const T* __restrict__ gto_scale, int64_t n, int64_t sv, int64_t sn, T dt, T ki_base,
    T nai_base) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ct = celltype ? (int)celltype[i] : fixed_celltype;
  T g_Ks = T(kGKs[ct]), g_to = T(kGto[ct]);
  if (gks_scale) g_Ks *= gks_scale[i];
  if (gto_scale) g_to *= gto_scale[i];
  T s[N_STATE];
  T* base = state + i * sn;
#pragma unroll
  for (int k = 0; k < N_STATE; ++k) s[k] = base[k * sv];
  constexpr bool exact = sizeof(T) == 8;
  const T I = ttp06_cell<T, exact>(V[i], s, dt, g_Ks, g_to, ct != 2, ki_base, nai_base);
#pragma unroll
  for (int k = 0; k < N_STATE; ++k) base[k * sv] = s[k];
  I_ion[i] = I;
}

template <typename T>
static int ttp06_launch(const T* V, T* state, T* I_ion, const uint8_t* celltype,
                        int fixed_celltype, const T* gks, const T* gto, int64_t n, int64_t sv,
                        int64_t sn, T dt, T ki_base, T nai_base, void* stream) {
  if (n < 0 || (!celltype && (fixed_celltype < 0 || fixed_celltype > 2)) ||
      (n > 0 && (!V || !state))) {
    set_error("hp_ttp06_step: bad arguments (n=%lld, fixed_celltype=%d; V and state required)",
              (long long)n, fixed_celltype);
    return -1;
  }
  if (n == 0) return 0;
  const int nt = 256;
  hipLaunchKernelGGL(ttp06_strided_kernel<T>, dim3(ceil_div(n, nt)), dim3(nt), 0,
                     (hipStream_t)stream, V, state, I_ion, celltype, fixed_celltype, gks, gto, n,
                     sv, sn, dt, ki_base, nai_base);
  return check_launch("hp_ttp06_step");
}

// The same cell algebra on the host, for checking the kernels' arithmetic
// form on a CPU (C library exp/log/division in place of the FP32 hardware
// instructions).
template <typename T>
static int ttp06_host(const T* V, T* state, T* I_ion, const uint8_t* celltype, int fixed_celltype,
                      const T* gks, const T* gto, int64_t n, int64_t sv, int64_t sn, T dt,
                      T ki_base, T nai_base) {
  if (n < 0 || (!celltype && (fixed_celltype < 0 || fixed_celltype > 2))) {
    set_error("hp_ttp06_step_host: bad arguments (n=%lld, fixed_celltype=%d)", (long long)n,
              fixed_celltype);
    return -1;
  }
  constexpr bool exact = sizeof(T) == 8;
  for (int64_t i = 0; i < n; ++i) {
    const int ct = celltype ? (int)celltype[i] : fixed_celltype;
    if (ct < 0 || ct > 2) {
      set_error("hp_ttp06_step_host: celltype %d at node %lld", ct, (long long)i);
      return -1;
    }
    T g_Ks = T(kGKsHost[ct]), g_to = T(kGtoHost[ct]);
    if (gks) g_Ks *= gks[i];
    if (gto) g_to *= gto[i];
    T s[N_STATE];
    T* base = state + i * sn;
    for (int k = 0; k < N_STATE; ++k) s[k] = base[k * sv];
    I_ion[i] = ttp06_cell<T, exact>(V[i], s, dt, g_Ks, g_to, ct != 2, ki_base, nai_base);
    for (int k = 0; k < N_STATE; ++k) base[k * sv] = s[k];
  }
  return 0;
}

}  // namespace hp

extern "C" int hp_ttp06_step_host_f32(const float* V, float* state, float* I_ion,
                                      const uint8_t* celltype, int fixed_celltype,
                                      const float* gks_scale, const float* gto_scale, int64_t n,
                                      int64_t stride_var, int64_t stride_node, float dt,
                                      float ki_base, float nai_base) {
  return hp::ttp06_host<float>(V, state, I_ion, celltype, fixed_celltype, gks_scale, gto_scale, n,
                               stride_var, stride_node, dt, ki_base, nai_base);
}

extern "C" int hp_ttp06_step_host_f64(const double* V, double* state, double* I_ion,
                                      const uint8_t* celltype, int fixed_celltype,
                                      const double* gks_scale, const double* gto_scale, int64_t n,
                                      int64_t stride_var, int64_t stride_node, double dt,
                                      double ki_base, double nai_base) {
  return hp::ttp06_host<double>(V, state, I_ion, celltype, fixed_celltype, gks_scale, gto_scale,
                                n, stride_var, stride_node, dt, ki_base, nai_base);
}

extern "C" int hp_ttp06_step_f32(const float* V, float* state, float* I_ion,
                                 const uint8_t* celltype, int fixed_celltype,
                                 const float* gks_scale, const float* gto_scale, int64_t n,
                                 int64_t stride_var, int64_t stride_node, float dt,
                                 float ki_base, float nai_base, void* stream) {
  return hp::ttp06_launch<float>(V, state, I_ion, celltype, fixed_celltype, gks_scale, gto_scale,
                                 n, stride_var, stride_node, dt, ki_base, nai_base, stream);
}

extern "C" int hp_ttp06_step_f64(const double* V, double* state, double* I_ion,
                                 const uint8_t* celltype, int fixed_celltype,
                                 const double* gks_scale, const double* gto_scale, int64_t n,
                                 int64_t stride_var, int64_t stride_node, double dt,
                                 double ki_base, double nai_base, void* stream) {
  return hp::ttp06_launch<double>(V, state, I_ion, celltype, fixed_celltype, gks_scale,
                                  gto_scale, n, stride_var, stride_node, dt, ki_base, nai_base,
                                  stream);
}
