// Heart-to-torso registration and projection on the device.
//
// * Voxel Dice of the 48 signed permutations of the rotation search
//   (hrdayapy/ecg/functions/register_heart_to_torso.py:146-224, the Dice of
//   _voxel_dice :109-138): for every permutation r the candidate cloud
//   s_i * src_c[:, perm_i] + c_tgt_i and the target subsample are voxelised on
//   one lattice anchored at their common minimum, and the occupied voxels of
//   both and of their intersection are counted.  One bitmap pair per
//   permutation; a thread per (point, permutation) sets its bit with an atomic
//   OR, then one pass per permutation counts the bits.  The lattice minima and
//   sizes come from the caller (they are 48 x 3 extremes, exact on the host).
// * Projection map of the heart nodes into the coarse torso rows
//   (hrdayapy/ecg/functions/compute_bspm.py:126-157).
//
// Coordinates are FP64 with every operation rounded as numpy rounds it
// (explicit _rn intrinsics, no fused multiply-add): a signed permutation
// applied to a vector is exact, so the voxel indices, the counts and the
// projected rows are the host's bit for bit.
#include <math.h>

#include <algorithm>

#include "common.h"
#include "hrdaya_hip.h"

namespace hp {

__global__ void __launch_bounds__(256) zero_words_kernel(uint32_t* __restrict__ w, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = 0u;
}

// blockIdx.y = permutation r.  Points [0, n_src) are the candidate cloud,
// [n_src, n_src + n_tgt) the target.  perm_sign[r] = {p0, p1, p2, s0, s1, s2}.
__global__ void __launch_bounds__(256) dice_mark_kernel(
    const double* __restrict__ src_c, int64_t n_src, const double* __restrict__ tgt,
    int64_t n_tgt, const int32_t* __restrict__ perm_sign, const double* __restrict__ c_tgt,
    const double* __restrict__ mn, const int32_t* __restrict__ dims,
    const int64_t* __restrict__ word_off, double vox, uint32_t* __restrict__ bits,
    unsigned long long* __restrict__ counts) {
  const int r = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_src + n_tgt) return;
  const int32_t* ps = perm_sign + 6 * r;
  double c[3];
  const bool is_src = i < n_src;
  if (is_src) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double x = src_c[3 * i + ps[a]];
      c[a] = __dadd_rn(ps[3 + a] > 0 ? x : -x, c_tgt[a]);
    }
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) c[a] = tgt[3 * (i - n_src) + a];
  }
  int64_t lin = 0;
  bool in = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double f = floor(__ddiv_rn(__dsub_rn(c[a], mn[3 * r + a]), vox));
    in = in && f >= 0.0 && f < (double)dims[3 * r + a];
    lin = lin * dims[3 * r + a] + (in ? (int64_t)f : 0);
  }
  if (!in) {  // a lattice that does not hold every point: counted, never written
    atomicAdd(counts + 3 * gridDim.y, 1ull);
    return;
  }
  const int64_t half = (word_off[r + 1] - word_off[r]) >> 1;
  uint32_t* w = bits + word_off[r] + (is_src ? 0 : half);
  atomicOr(w + (lin >> 5), 1u << (uint32_t)(lin & 31));
}

// counts[r] = {occupied voxels of the candidate, of the target, of both};
// counts[3 * n_perm] = points that fell outside their lattice (0 unless the
// caller's minima or sizes are wrong).
__global__ void __launch_bounds__(256) dice_count_kernel(const uint32_t* __restrict__ bits,
                                                         const int64_t* __restrict__ word_off,
                                                         unsigned long long* __restrict__ counts) {
  const int r = blockIdx.y;
  const int64_t half = (word_off[r + 1] - word_off[r]) >> 1;
  const uint32_t* a = bits + word_off[r];
  const uint32_t* b = a + half;
  uint32_t na = 0, nb = 0, nab = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < half;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t x = a[i], y = b[i];
    na += __popc(x);
    nb += __popc(y);
    nab += __popc(x & y);
  }
  // wave sums, then one atomic per wave (integer sums: order-independent)
  for (int o = 32; o > 0; o >>= 1) {
    na += __shfl_xor(na, o);
    nb += __shfl_xor(nb, o);
    nab += __shfl_xor(nab, o);
  }
  if ((threadIdx.x & 63) == 0 && (na | nb | nab)) {
    atomicAdd(counts + 3 * r + 0, (unsigned long long)na);
    atomicAdd(counts + 3 * r + 1, (unsigned long long)nb);
    atomicAdd(counts + 3 * r + 2, (unsigned long long)nab);
  }
}

__global__ void __launch_bounds__(64) zero_counts_kernel(unsigned long long* c, int n) {
  const int i = threadIdx.x + blockIdx.x * 64;
  if (i < n) c[i] = 0ull;
}

// proj[i] = id_map[flat(floor((R (x_i scale) + t - origin) / dx))] or -1.
// R is applied as written (three products summed left to right); for the
// registration's signed permutations every product and sum but the one that
// adds t is exact.
__global__ void __launch_bounds__(256) project_rows_kernel(
    const double* __restrict__ xyz, int64_t n, const double* __restrict__ Rt, double scale,
    const double* __restrict__ origin, double dx, const int32_t* __restrict__ id_map, int32_t ni,
    int32_t nj, int32_t nk, int64_t* __restrict__ proj) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double y[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) y[a] = __dmul_rn(xyz[3 * i + a], scale);
  int32_t ijk[3];
  bool in = true;
  const int32_t lim[3] = {ni, nj, nk};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double s = __dmul_rn(Rt[3 * a + 0], y[0]);
    s = __dadd_rn(s, __dmul_rn(Rt[3 * a + 1], y[1]));
    s = __dadd_rn(s, __dmul_rn(Rt[3 * a + 2], y[2]));
    const double f = floor(__ddiv_rn(__dsub_rn(__dadd_rn(s, Rt[9 + a]), origin[a]), dx));
    // numpy's astype(int32) of an out-of-range float is not defined; the
    // bounds test below rejects every such value first
    ijk[a] = (f >= -2147483648.0 && f < 2147483648.0) ? (int32_t)f : -1;
    in = in && f >= 0.0 && f < (double)lim[a];
  }
  int64_t out = -1;
  if (in) {
    const int64_t flat = ((int64_t)ijk[0] * nj + ijk[1]) * nk + ijk[2];
    out = id_map[flat];
  }
  proj[i] = out;
}

}  // namespace hp

using namespace hp;

extern "C" {

int hp_register_dice_f64(const double* src_c, int64_t n_src, const double* tgt, int64_t n_tgt,
                         const int32_t* perm_sign, int32_t n_perm, const double* c_tgt,
                         const double* mn, const int32_t* dims, const int64_t* word_off,
                         double vox_mm, uint32_t* bits, int64_t n_words,
                         unsigned long long* counts, void* stream) {
  if (!src_c || !tgt || !perm_sign || !c_tgt || !mn || !dims || !word_off || !bits || !counts ||
      n_src <= 0 || n_tgt <= 0 || n_perm <= 0 || n_perm > 65535 || !(vox_mm > 0.0) ||
      n_words <= 0 || n_src + n_tgt > ((int64_t)1 << 40)) {
    set_error("hp_register_dice_f64: bad arguments");
    return -1;
  }
  hipStream_t s = (hipStream_t)stream;
  const int zb = (int)std::min<int64_t>(ceil_div(n_words, 256), 4096);
  hipLaunchKernelGGL(zero_words_kernel, dim3(zb), dim3(256), 0, s, bits, n_words);
  hipLaunchKernelGGL(zero_counts_kernel, dim3(ceil_div(3 * n_perm + 1, 64)), dim3(64), 0, s,
                     counts, 3 * n_perm + 1);
  hipLaunchKernelGGL(dice_mark_kernel, dim3(ceil_div(n_src + n_tgt, 256), n_perm), dim3(256), 0,
                     s, src_c, n_src, tgt, n_tgt, perm_sign, c_tgt, mn, dims, word_off, vox_mm,
                     bits, counts);
  hipLaunchKernelGGL(dice_count_kernel, dim3(64, n_perm), dim3(256), 0, s, bits, word_off,
                     counts);
  return check_launch("hp_register_dice_f64");
}

int hp_project_rows_f64(const double* xyz, int64_t n, const double* Rt, double scale,
                        const double* origin, double dx_mm, const int32_t* id_map, int32_t ni,
                        int32_t nj, int32_t nk, int64_t* proj, void* stream) {
  if (!xyz || !Rt || !origin || !id_map || !proj || n < 0 || ni <= 0 || nj <= 0 || nk <= 0 ||
      !(dx_mm > 0.0)) {
    set_error("hp_project_rows_f64: bad arguments");
    return -1;
  }
  if (n == 0) return 0;
  hipLaunchKernelGGL(project_rows_kernel, dim3(ceil_div(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, xyz, n, Rt, scale, origin, dx_mm, id_map, ni, nj, nk,
                     proj);
  return check_launch("hp_project_rows_f64");
}

}  // extern "C"
