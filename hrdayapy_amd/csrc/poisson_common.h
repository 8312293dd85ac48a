// Shared pieces of the torso Poisson kernels (column-run layout, reductions,
// PCG scalar slots) used by poisson.hip and multigrid.hip.
#pragma once
#include "common.h"
#include "hrdaya_hip.h"

namespace hp {

constexpr int PWAVES = 4;          // waves per block
constexpr int PBLOCK = 64 * PWAVES;

struct ColMeta {
  int k0, len, off;
};

__device__ __forceinline__ ColMeta col_meta(const hp_torso& t, int i, int j) {
  ColMeta m;
  if (i < 0 || i >= t.ni || j < 0 || j >= t.nj) {
    m.k0 = 0;
    m.len = 0;
    m.off = 0;
    return m;
  }
  const int c = i * t.nj + j;
  m.k0 = t.col_k0[c];
  m.len = t.col_len[c];
  m.off = t.col_off[c];
  return m;
}

__device__ __forceinline__ int cell_at(const ColMeta& m, int k) {
  const int dk = k - m.k0;
  return (dk >= 0 && dk < m.len) ? m.off + dk : -1;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block reduction of NV doubles followed by the last-block finish: the block
// partials go to part[NV][gridDim.x]; the last block to arrive sums them in
// block order into out[0..NV) and resets the arrival counter.
template <int NV>
__device__ __forceinline__ bool block_reduce_last(double (&v)[NV], double* part, int* counter,
                                                  double* out) {
  __shared__ double sh[NV][PWAVES];
  __shared__ int is_last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double s = wave_sum(v[q]);
    if (lane == 0) sh[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int w = 0; w < PWAVES; ++w) s += sh[q][w];
      part[q * gridDim.x + blockIdx.x] = s;
    }
    __threadfence();
    const int old = atomicAdd(counter, 1);
    is_last = (old == (int)gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += PBLOCK) s += part[q * gridDim.x + b];
    s = wave_sum(s);
    if (lane == 0) sh[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int w = 0; w < PWAVES; ++w) s += sh[q][w];
      out[q] = s;
    }
    *counter = 0;
  }
  return true;
}

// The metadata of a column and of its four x/y neighbours, loaded once per
// column (wave-uniform: scalar loads).
struct ColNbr {
  ColMeta c, xp, xm, yp, ym;
};

__device__ __forceinline__ ColNbr col_nbr(const hp_torso& t, int i, int j) {
  ColNbr m;
  m.c = col_meta(t, i, j);
  m.xp = col_meta(t, i + 1, j);
  m.xm = col_meta(t, i - 1, j);
  m.yp = col_meta(t, i, j + 1);
  m.ym = col_meta(t, i, j - 1);
  return m;
}

// One wave per non-empty column (t.col_list), grid-stride over the list; the
// wave walks the column's run in 64-cell chunks.  F(cell, i, j, k, nbr) is
// called for the lanes inside the run.
template <typename F>
__device__ __forceinline__ void for_cells(const hp_torso& t, F&& f) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int64_t w = (int64_t)blockIdx.x * PWAVES + wave; w < t.n_col_list;
       w += (int64_t)gridDim.x * PWAVES) {
    const int col = t.col_list[w];
    const int i = col / t.nj, j = col % t.nj;
    const ColNbr m = col_nbr(t, i, j);
    for (int dk0 = 0; dk0 < m.c.len; dk0 += 64) {
      const int dk = dk0 + lane;
      if (dk < m.c.len) f(m.c.off + dk, i, j, m.c.k0 + dk, m);
    }
  }
}

// A x at one cell in FP32 with the stored diagonal (the inner PCG operator).
// (A x)_c in FP32 in zero-row-sum form, sum_f g_f (x_c - x_nb) with the
// neighbour values from xat(n): exact for a constant x, so the large
// near-constant component of a torso potential measured against one pinned
// cell does not leak into the product through a rounded diagonal.  The pinned
// cell and cells without any conducting face are identity rows.
template <typename F>
__device__ __forceinline__ float row_f32(const hp_torso& t, int c, int k, const ColNbr& m,
                                         float xc, F&& xat) {
  if (c == t.pin) return xc;
  float acc = 0.f, gs = 0.f, g;
  int n;
  if (k + 1 < m.c.k0 + m.c.len) { g = t.gz[c]; acc += g * (xc - xat(c + 1)); gs += g; }
  if (k - 1 >= m.c.k0) { g = t.gz[c - 1]; acc += g * (xc - xat(c - 1)); gs += g; }
  if ((n = cell_at(m.xp, k)) >= 0) { g = t.gx[c]; acc += g * (xc - xat(n)); gs += g; }
  if ((n = cell_at(m.xm, k)) >= 0) { g = t.gx[n]; acc += g * (xc - xat(n)); gs += g; }
  if ((n = cell_at(m.yp, k)) >= 0) { g = t.gy[c]; acc += g * (xc - xat(n)); gs += g; }
  if ((n = cell_at(m.ym, k)) >= 0) { g = t.gy[n]; acc += g * (xc - xat(n)); gs += g; }
  return gs == 0.f ? xc : acc;
}

__device__ __forceinline__ float apply_row_f32(const hp_torso& t, const float* __restrict__ x, int c,
                                               int i, int j, int k, const ColNbr& m) {
  return row_f32(t, c, k, m, x[c], [&](int n) { return x[n]; });
}

// ---------------------------------------------------------------------------
// PCG state in device memory (FP64 scalars)
// ---------------------------------------------------------------------------
enum { S_RZ = 0, S_PAP, S_ALPHA, S_BETA, S_RR0, S_RR, S_TOL2, S_DONE, S_ITER, S_MAXIT, S_ZZ,
       S_MODE, S_NCELL, S_NSC };

inline int grid_blocks(const hp_torso* t) {
  int64_t nb = (t->n_col_list + PWAVES - 1) / PWAVES;
  if (nb > 4096) nb = 4096;
  return (int)(nb < 1 ? 1 : nb);
}

inline bool torso_ok(const hp_torso* t, const char* what) {
  if (!t || t->ni <= 0 || t->nj <= 0 || t->nk <= 0 || t->nk > (1 << 20) || !t->col_list ||
      t->n_col_list < 0 || t->n_col_list > (int64_t)t->ni * t->nj) {
    set_error("%s: bad torso grid", what);
    return false;
  }
  return true;
}

}  // namespace hp
