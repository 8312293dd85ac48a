// Coarse torso grid on the device: the 7-point stiffness in the column-run
// layout (sharp and smoothed boundary), the system's row numbering, the
// surface cells and the pinned cell, and the smoothed-boundary domain
// function psi from the fine segmentation.
//
// Reference: hrdayapy/ecg/functions/build_torso_grid.py
//   _assemble_stiffness          :276-397  (harmonic face means of sigma and,
//                                           for the SBM, of psi; diagonal summed
//                                           over the six faces in the order
//                                           +i, -i, +j, -j, +k, -k)
//   compute_smoothed_domain      :130-186  (signed Euclidean distance of the
//                                           fine segmentation, tanh profile,
//                                           linear zoom to the coarse grid,
//                                           crop/pad, cutoff)
//   extend_conductivity          :189-206  (sigma of the nearest tissue cell)
//   _find_surface_ids            :469-488; pinned cell :616-623
//
// Distances come from a feature transform: for every voxel the coordinates
// of its nearest feature voxel, built one axis at a time (axis 0, then 1,
// then 2) by the linear-time lower-envelope method of Maurer, Qi and
// Raghavan (IEEE PAMI 25(2), 2003), with anisotropic voxel spacing.  On
// each line the sites are the previous pass's features; a site is dropped
// from the envelope when it is strictly hidden by its neighbours, and a
// query keeps the earlier site on an exact tie.  Distances are then
// sqrt(sum_a ((f_a - x_a) s_a)^2) summed over axes 0, 1, 2.
//
// Every count, scan and list is integer work; the conductances are FP64 with
// explicitly rounded operations (no fused multiply-add) in the reference's
// order, stored as FP32 for the solver and optionally as FP64.
#include <math.h>

#include <algorithm>

#include "common.h"
#include "hrdaya_hip.h"

namespace hp {

// ---------------------------------------------------------------------------
// Feature transform
// ---------------------------------------------------------------------------

struct LineGeom {
  int64_t base, stride;
  int32_t p[3];  // the line's fixed coordinates (entry d unused)
};

__device__ __forceinline__ LineGeom line_geom(int d, int64_t line, int32_t n0, int32_t n1,
                                              int32_t n2) {
  LineGeom g;
  if (d == 0) {  // lines along i at (j, k)
    const int32_t j = (int32_t)(line / n2), k = (int32_t)(line % n2);
    g.base = (int64_t)j * n2 + k;
    g.stride = (int64_t)n1 * n2;
    g.p[0] = 0; g.p[1] = j; g.p[2] = k;
  } else if (d == 1) {  // lines along j at (i, k)
    const int32_t i = (int32_t)(line / n2), k = (int32_t)(line % n2);
    g.base = (int64_t)i * n1 * n2 + k;
    g.stride = n2;
    g.p[0] = i; g.p[1] = 0; g.p[2] = k;
  } else {  // lines along k at (i, j)
    const int32_t i = (int32_t)(line / n1), j = (int32_t)(line % n1);
    g.base = ((int64_t)i * n1 + j) * n2;
    g.stride = 1;
    g.p[0] = i; g.p[1] = j; g.p[2] = 0;
  }
  return g;
}

// squared distance from grid point q to site f, summed over axes 0, 1, 2
__device__ __forceinline__ double dist2(const int32_t f[3], const int32_t q[3], const double s[3]) {
  double acc = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double t = __dmul_rn((double)(f[a] - q[a]), s[a]);
    acc = __dadd_rn(acc, __dmul_rn(t, t));
  }
  return acc;
}

// One pass of the feature transform along axis d for lines [line0, line0 +
// n_lines).  ft_in / ft_out: [3][N] int32, -1 in component 0 = no site.  For
// d == 0 the sites come from the mask (feature where mask == feature_value).
// stack: n_lines * len int32, entry m of line t at stack[m * n_lines + t].
__global__ void __launch_bounds__(256) ft_pass_kernel(
    int d, const uint8_t* __restrict__ mask, uint8_t feature_value,
    const int32_t* __restrict__ ft_in, int32_t* __restrict__ ft_out, int32_t n0, int32_t n1,
    int32_t n2, int64_t N, double s0, double s1, double s2, int64_t line0, int64_t n_lines,
    int32_t* __restrict__ stack) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_lines) return;
  const int64_t line = line0 + t;
  const LineGeom g = line_geom(d, line, n0, n1, n2);
  const int32_t len = d == 0 ? n0 : (d == 1 ? n1 : n2);
  const double s[3] = {s0, s1, s2};
  const double sd = s[d];
  // site coordinates at line position x
  auto site = [&](int32_t x, int32_t f[3]) -> bool {
    const int64_t v = g.base + (int64_t)x * g.stride;
    if (d == 0) {
      if (mask[v] != feature_value) return false;
      f[0] = x; f[1] = g.p[1]; f[2] = g.p[2];
      return true;
    }
    f[0] = ft_in[v];
    if (f[0] < 0) return false;
    f[1] = ft_in[N + v];
    f[2] = ft_in[2 * N + v];
    return true;
  };
  auto perp = [&](const int32_t f[3]) -> double {  // squared distance off the line
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (a == d) continue;
      const double q = __dmul_rn((double)(f[a] - g.p[a]), s[a]);
      acc = __dadd_rn(acc, __dmul_rn(q, q));
    }
    return acc;
  };
  // lower envelope of the sites (stack of line positions)
  int32_t top = -1;
  for (int32_t x = 0; x < len; ++x) {
    int32_t fw[3];
    if (!site(x, fw)) continue;
    const double wR = perp(fw);
    while (top >= 1) {
      int32_t fu[3], fv[3];
      site(stack[(int64_t)(top - 1) * n_lines + t], fu);
      site(stack[(int64_t)top * n_lines + t], fv);
      const double a = __dmul_rn((double)(fv[d] - fu[d]), sd);
      const double b = __dmul_rn((double)(fw[d] - fv[d]), sd);
      const double c = __dadd_rn(a, b);
      const double uR = perp(fu), vR = perp(fv);
      // v is hidden when its interval between u and w is empty
      double h = __dmul_rn(c, vR);
      h = __dsub_rn(h, __dmul_rn(b, uR));
      h = __dsub_rn(h, __dmul_rn(a, wR));
      h = __dsub_rn(h, __dmul_rn(__dmul_rn(a, b), c));
      if (!(h > 0.0)) break;
      --top;
    }
    ++top;
    stack[(int64_t)top * n_lines + t] = x;
  }
  // nearest site of every position, the earlier one on a tie
  int32_t l = 0;
  for (int32_t x = 0; x < len; ++x) {
    const int64_t v = g.base + (int64_t)x * g.stride;
    if (top < 0) {
      ft_out[v] = -1;
      ft_out[N + v] = -1;
      ft_out[2 * N + v] = -1;
      continue;
    }
    int32_t q[3] = {g.p[0], g.p[1], g.p[2]};
    q[d] = x;
    int32_t f1[3];
    site(stack[(int64_t)l * n_lines + t], f1);
    double d1 = dist2(f1, q, s);
    while (l < top) {
      int32_t f2[3];
      site(stack[(int64_t)(l + 1) * n_lines + t], f2);
      const double d2 = dist2(f2, q, s);
      if (d1 <= d2) break;
      d1 = d2;
      f1[0] = f2[0]; f1[1] = f2[1]; f1[2] = f2[2];
      ++l;
    }
    ft_out[v] = f1[0];
    ft_out[N + v] = f1[1];
    ft_out[2 * N + v] = f1[2];
  }
}

// psi_fine = (1 + tanh(3 d / w)) / 2 with d = +dist to the nearest outside
// voxel inside, -dist to the nearest inside voxel outside (ft_out_of_in /
// ft_in_of_out: the feature transforms of the two sets).
__global__ void __launch_bounds__(256) sbm_psi_kernel(
    const uint8_t* __restrict__ inside, const int32_t* __restrict__ ft_nearest_out,
    const int32_t* __restrict__ ft_nearest_in, int32_t n0, int32_t n1, int32_t n2, int64_t N,
    double s0, double s1, double s2, double width, double* __restrict__ psi) {
  const double s[3] = {s0, s1, s2};
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int32_t q[3] = {(int32_t)(v / ((int64_t)n1 * n2)), (int32_t)((v / n2) % n1),
                          (int32_t)(v % n2)};
    const bool in = inside[v] != 0;
    const int32_t* ft = in ? ft_nearest_out : ft_nearest_in;
    const int32_t f[3] = {ft[v], ft[N + v], ft[2 * N + v]};
    const double dist = sqrt(dist2(f, q, s));
    const double ds = in ? dist : -dist;  // d_in - 0 or 0 - d_out
    psi[v] = __dmul_rn(0.5, __dadd_rn(1.0, tanh(__ddiv_rn(__dmul_rn(3.0, ds), width))));
  }
}

// Linear zoom of a dense FP64 volume (n0, n1, n2) to the (o0, o1, o2) grid
// that maps corner to corner (coordinate kk * (n - 1) / (o - 1)), written into
// the (m0, m1, m2) output cropped / padded with 0; 0 at sample points beyond
// the input's edges and where the value falls below cutoff.  Weights
// w0 = 1 - x, w1 = 1 - w0 per axis, corners summed with the last axis
// fastest, each corner value times its weights in axis order.
__global__ void __launch_bounds__(256) zoom_linear_kernel(
    const double* __restrict__ in, int32_t n0, int32_t n1, int32_t n2, int32_t o0, int32_t o1,
    int32_t o2, double z0, double z1, double z2, int32_t m0, int32_t m1, int32_t m2,
    double cutoff, double* __restrict__ out, uint8_t* __restrict__ positive) {
  const int64_t M = (int64_t)m0 * m1 * m2;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < M;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c[3] = {(int32_t)(v / ((int64_t)m1 * m2)), (int32_t)((v / m2) % m1),
                          (int32_t)(v % m2)};
    const int32_t o[3] = {o0, o1, o2}, n[3] = {n0, n1, n2};
    const double z[3] = {z0, z1, z2};
    bool ok = true;
    int32_t st[3];
    double w[3][2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      if (c[a] >= o[a]) { ok = false; break; }  // padding beyond the zoomed grid
      const double cc = __dmul_rn((double)c[a], z[a]);
      if (cc < 0.0 || cc > (double)(n[a] - 1)) { ok = false; break; }
      const double fl = floor(cc);
      const double x = __dsub_rn(cc, fl);
      st[a] = (int32_t)fl;
      w[a][0] = __dsub_rn(1.0, x);
      w[a][1] = __dsub_rn(1.0, w[a][0]);
    }
    double val = 0.0;
    if (ok) {
      for (int h = 0; h < 8; ++h) {
        const int b0 = (h >> 2) & 1, b1 = (h >> 1) & 1, b2 = h & 1;
        // a knot past the last sample carries weight 0; its index is clamped
        const int32_t i0 = min(st[0] + b0, n0 - 1), i1 = min(st[1] + b1, n1 - 1),
                      i2 = min(st[2] + b2, n2 - 1);
        double t = in[((int64_t)i0 * n1 + i1) * n2 + i2];
        t = __dmul_rn(t, w[0][b0]);
        t = __dmul_rn(t, w[1][b1]);
        t = __dmul_rn(t, w[2][b2]);
        val = __dadd_rn(val, t);
      }
      if (val < cutoff) val = 0.0;
    }
    out[v] = val;
    if (positive) positive[v] = val > 0.0 ? 1 : 0;
  }
}

// Nearest-neighbour zoom of an int32 label volume to the (o0, o1, o2) grid
// that maps corner to corner: sample floor(kk * (n - 1) / (o - 1) + 0.5),
// 0 at sample points beyond the input's edges.
__global__ void __launch_bounds__(256) zoom_nearest_kernel(
    const int32_t* __restrict__ in, int32_t n0, int32_t n1, int32_t n2, int32_t o0, int32_t o1,
    int32_t o2, double z0, double z1, double z2, int32_t* __restrict__ out) {
  const int64_t M = (int64_t)o0 * o1 * o2;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < M;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int32_t c[3] = {(int32_t)(v / ((int64_t)o1 * o2)), (int32_t)((v / o2) % o1),
                          (int32_t)(v % o2)};
    const int32_t n[3] = {n0, n1, n2};
    const double z[3] = {z0, z1, z2};
    int32_t idx[3];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double cc = __dmul_rn((double)c[a], z[a]);
      if (cc < 0.0 || cc > (double)(n[a] - 1)) { ok = false; break; }
      idx[a] = min((int32_t)floor(__dadd_rn(cc, 0.5)), n[a] - 1);
    }
    out[v] = ok ? in[((int64_t)idx[0] * n1 + idx[1]) * n2 + idx[2]] : 0;
  }
}

// sigma from the label map (labels without an entry: 0) and the tissue flag
// (label != 0)
__global__ void __launch_bounds__(256) label_sigma_kernel(
    const int32_t* __restrict__ lab, int64_t N, const int32_t* __restrict__ keys,
    const double* __restrict__ vals, int32_t n_keys, double* __restrict__ sigma,
    uint8_t* __restrict__ tissue) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int32_t l = lab[v];
    double s = 0.0;
    for (int q = 0; q < n_keys; ++q)
      if (keys[q] == l) s = vals[q];  // the map's last entry for a label wins
    sigma[v] = s;
    tissue[v] = l != 0 ? 1 : 0;
  }
}

// sigma of the nearest tissue cell on the band (active, not tissue).
__global__ void __launch_bounds__(256) extend_sigma_kernel(
    const int32_t* __restrict__ ft_tissue, const uint8_t* __restrict__ tissue,
    const uint8_t* __restrict__ active, const double* __restrict__ sigma_in, int32_t n1,
    int32_t n2, int64_t N, double* __restrict__ sigma_out) {
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    double s = sigma_in[v];
    if (active[v] && !tissue[v] && ft_tissue[v] >= 0) {
      const int64_t f = ((int64_t)ft_tissue[v] * n1 + ft_tissue[N + v]) * n2 + ft_tissue[2 * N + v];
      s = sigma_in[f];
    }
    sigma_out[v] = s;
  }
}

// ---------------------------------------------------------------------------
// Columns, scans and lists
// ---------------------------------------------------------------------------

// Per column (i, j): first and last k with flag set (run k0, length), and the
// number of set flags.
__global__ void __launch_bounds__(256) column_stats_kernel(const uint8_t* __restrict__ flag,
                                                           int32_t ni, int32_t nj, int32_t nk,
                                                           int32_t* __restrict__ k0,
                                                           int32_t* __restrict__ run,
                                                           int32_t* __restrict__ cnt) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (int64_t)ni * nj) return;
  const uint8_t* f = flag + col * nk;
  int32_t first = -1, last = -1, n = 0;
  for (int32_t k = 0; k < nk; ++k) {
    if (f[k]) {
      if (first < 0) first = k;
      last = k;
      ++n;
    }
  }
  if (k0) k0[col] = first < 0 ? 0 : first;
  if (run) run[col] = first < 0 ? 0 : last - first + 1;
  if (cnt) cnt[col] = n;
}

// out[0..n] = exclusive prefix sums of in[0..n), out[n] = total; one block.
__global__ void __launch_bounds__(1024) scan_kernel(const int32_t* __restrict__ in, int64_t n,
                                                    int64_t* __restrict__ out) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (n + 1023) / 1024;
  const int64_t lo = std::min<int64_t>(n, t * per), hi = std::min<int64_t>(n, lo + per);
  int64_t s = 0;
  for (int64_t i = lo; i < hi; ++i) s += in[i];
  part[t] = s;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan of the parts
    const int64_t add = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += add;
    __syncthreads();
  }
  int64_t run = part[t] - s;
  for (int64_t i = lo; i < hi; ++i) {
    out[i] = run;
    run += in[i];
  }
  if (t == 1023) out[n] = part[1023];
}

// For every set flag of column col, in k order: list[base[col] + rank] = flat
// index; id_map (optional) gets the rank's row or -1 for every cell of the
// column; cell_of_row (optional) the stored cell col_off[col] + k - k0[col].
__global__ void __launch_bounds__(256) column_emit_kernel(
    const uint8_t* __restrict__ flag, int32_t ni, int32_t nj, int32_t nk,
    const int64_t* __restrict__ base, int32_t* __restrict__ list, int32_t* __restrict__ id_map,
    const int32_t* __restrict__ k0, const int64_t* __restrict__ col_off,
    int64_t* __restrict__ cell_of_row) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= (int64_t)ni * nj) return;
  int64_t r = base[col];
  for (int32_t k = 0; k < nk; ++k) {
    const int64_t flat = col * nk + k;
    if (flag[flat]) {
      if (list) list[r] = (int32_t)flat;
      if (id_map) id_map[flat] = (int32_t)r;
      if (cell_of_row) cell_of_row[r] = col_off[col] + (k - k0[col]);
      ++r;
    } else if (id_map) {
      id_map[flat] = -1;
    }
  }
}

// surface = tissue cells with an in-box face neighbour outside the tissue
__global__ void __launch_bounds__(256) surface_flag_kernel(const uint8_t* __restrict__ tissue,
                                                           int32_t ni, int32_t nj, int32_t nk,
                                                           uint8_t* __restrict__ surf) {
  const int64_t N = (int64_t)ni * nj * nk;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int32_t i = (int32_t)(v / ((int64_t)nj * nk)), j = (int32_t)((v / nk) % nj),
                  k = (int32_t)(v % nk);
    bool s = false;
    if (tissue[v]) {
      s = (i > 0 && !tissue[v - (int64_t)nj * nk]) || (i + 1 < ni && !tissue[v + (int64_t)nj * nk]) ||
          (j > 0 && !tissue[v - nk]) || (j + 1 < nj && !tissue[v + nk]) ||
          (k > 0 && !tissue[v - 1]) || (k + 1 < nk && !tissue[v + 1]);
    }
    surf[v] = s ? 1 : 0;
  }
}

// flat index of the tissue cell of rank `rank` (the scan over per-column
// counts locates its column, the column walk its k); one thread.
__global__ void select_rank_kernel(const uint8_t* __restrict__ flag, const int64_t* __restrict__ scan,
                                   int64_t n_cols, int32_t nk, int64_t rank,
                                   int64_t* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t lo = 0, hi = n_cols;  // last column with scan[col] <= rank
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) / 2;
    if (scan[mid] <= rank) lo = mid; else hi = mid;
  }
  int64_t left = rank - scan[lo];
  int64_t res = -1;
  for (int32_t k = 0; k < nk; ++k) {
    const int64_t flat = lo * nk + k;
    if (flag[flat]) {
      if (left == 0) { res = flat; break; }
      --left;
    }
  }
  out[0] = res;
}

// ---------------------------------------------------------------------------
// Face conductances and diagonal
// ---------------------------------------------------------------------------

// conductance of the face between cells a (row) and b, the reference's rules
__device__ __forceinline__ double face_cond(double sa, double sb, const double* psi, int64_t a,
                                            int64_t b, double h2) {
  double pf = 1.0;
  if (psi) {
    const double pa = psi[a], pb = psi[b];
    if (pa <= 0.0 && pb <= 0.0) return 0.0;
    if (__dadd_rn(sa, sb) < 1e-30) return 0.0;
    const double ps = __dadd_rn(pa, pb);
    pf = ps < 1e-30 ? 0.0 : __ddiv_rn(__dmul_rn(__dmul_rn(2.0, pa), pb), ps);
  } else if (__dadd_rn(sa, sb) < 1e-30) {
    return 0.0;
  }
  const double sf = __ddiv_rn(__dmul_rn(__dmul_rn(2.0, sa), sb), __dadd_rn(sa, sb));
  return __ddiv_rn(__dmul_rn(pf, sf), h2);
}

// One thread per dense cell; cells inside their column's run get the +i/+j/+k
// face conductances (0 unless both cells are in the system) and the diagonal
// (the six faces summed in the reference's order; 1 for cells outside the
// system and rows with no conductance; the FP32 diagonal also 1 on the pinned
// cell, the FP64 copy keeps the unpinned value).  bad[0] counts faces
// to a cell outside the system with a nonzero conductance (sigma on air),
// which the column-run operator cannot represent.
__global__ void __launch_bounds__(256) torso_faces_kernel(
    const double* __restrict__ sigma, const double* __restrict__ psi,
    const uint8_t* __restrict__ active, int32_t ni, int32_t nj, int32_t nk, double h2,
    const int32_t* __restrict__ col_k0, const int32_t* __restrict__ col_len,
    const int64_t* __restrict__ col_off, int64_t pin_flat, float* __restrict__ gx,
    float* __restrict__ gy, float* __restrict__ gz, float* __restrict__ diag,
    double* __restrict__ g64, int64_t n_cells, unsigned long long* __restrict__ bad) {
  const int64_t N = (int64_t)ni * nj * nk;
  const int64_t sI = (int64_t)nj * nk, sJ = nk;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < N;
       v += (int64_t)gridDim.x * blockDim.x) {
    const int64_t col = v / nk;
    const int32_t k = (int32_t)(v % nk);
    const int32_t kk = k - col_k0[col];
    if (kk < 0 || kk >= col_len[col]) continue;
    const int64_t cell = col_off[col] + kk;
    const int32_t i = (int32_t)(v / sI), j = (int32_t)((v / nk) % nj);
    const bool act = active[v] != 0;
    const double sv = sigma[v];
    double g[6];
    const int64_t nb[6] = {v + sI, v - sI, v + sJ, v - sJ, v + 1, v - 1};
    const bool inb[6] = {i + 1 < ni, i > 0, j + 1 < nj, j > 0, k + 1 < nk, k > 0};
    double dg = 0.0;
    for (int f = 0; f < 6; ++f) {
      g[f] = 0.0;
      if (!act || !inb[f]) continue;
      const double c = face_cond(sv, sigma[nb[f]], psi, v, nb[f], h2);
      dg = __dadd_rn(dg, c);
      if (active[nb[f]]) g[f] = c;
      else if (c != 0.0) atomicAdd(bad, 1ull);
    }
    if (!act || !(dg > 0.0)) dg = 1.0;
    gx[cell] = (float)g[0];
    gy[cell] = (float)g[2];
    gz[cell] = (float)g[4];
    diag[cell] = v == pin_flat ? 1.0f : (float)dg;  // the solver's pinned row
    if (g64) {  // the unpinned matrix
      g64[cell] = g[0];
      g64[n_cells + cell] = g[2];
      g64[2 * n_cells + cell] = g[4];
      g64[3 * n_cells + cell] = dg;
    }
  }
}

__global__ void zero_u64_kernel(unsigned long long* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = 0ull;
}

inline int grid_of(int64_t n, int cap = 8192) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, 256), cap));
}

}  // namespace hp

using namespace hp;

extern "C" {

int hp_feature_transform(const uint8_t* mask, uint8_t feature_value, int32_t n0, int32_t n1,
                         int32_t n2, const double* spacing, int32_t* ft, int32_t* ft_tmp,
                         int32_t* stack, int64_t stack_len, void* stream) {
  if (!mask || !spacing || !ft || !ft_tmp || !stack || n0 <= 0 || n1 <= 0 || n2 <= 0) {
    set_error("hp_feature_transform: bad arguments");
    return -1;
  }
  const int64_t N = (int64_t)n0 * n1 * n2;
  const int32_t lens[3] = {n0, n1, n2};
  const int32_t lmax = std::max(n0, std::max(n1, n2));
  if (stack_len < lmax || N >= ((int64_t)1 << 40)) {
    set_error("hp_feature_transform: stack of %lld entries holds no line of %d",
              (long long)stack_len, lmax);
    return -1;
  }
  if (!(spacing[0] > 0.0 && spacing[1] > 0.0 && spacing[2] > 0.0)) {
    set_error("hp_feature_transform: spacing must be positive");
    return -1;
  }
  hipStream_t s = (hipStream_t)stream;
  // pass 0 writes ft_tmp, pass 1 ft, pass 2 ft_tmp -> copied to ft below
  int32_t* bufs[3] = {ft_tmp, ft, ft_tmp};
  const int32_t* src[3] = {nullptr, ft_tmp, ft};
  for (int d = 0; d < 3; ++d) {
    const int64_t n_lines = N / lens[d];
    const int64_t batch = std::max<int64_t>(1, std::min<int64_t>(n_lines, stack_len / lens[d]));
    for (int64_t l0 = 0; l0 < n_lines; l0 += batch) {
      const int64_t nl = std::min<int64_t>(batch, n_lines - l0);
      hipLaunchKernelGGL(ft_pass_kernel, dim3(ceil_div(nl, 256)), dim3(256), 0, s, d, mask,
                         feature_value, src[d], bufs[d], n0, n1, n2, N, spacing[0], spacing[1],
                         spacing[2], l0, nl, stack);
    }
  }
  HP_CHECK(hipMemcpyAsync(ft, ft_tmp, (size_t)(3 * N) * sizeof(int32_t), hipMemcpyDeviceToDevice,
                          s));
  return check_launch("hp_feature_transform");
}

int hp_sbm_psi_f64(const uint8_t* inside, const int32_t* ft_nearest_out,
                   const int32_t* ft_nearest_in, int32_t n0, int32_t n1, int32_t n2,
                   const double* spacing, double width_mm, double* psi, void* stream) {
  if (!inside || !ft_nearest_out || !ft_nearest_in || !spacing || !psi || n0 <= 0 || n1 <= 0 ||
      n2 <= 0 || !(width_mm > 0.0)) {
    set_error("hp_sbm_psi_f64: bad arguments");
    return -1;
  }
  const int64_t N = (int64_t)n0 * n1 * n2;
  hipLaunchKernelGGL(sbm_psi_kernel, dim3(grid_of(N)), dim3(256), 0, (hipStream_t)stream, inside,
                     ft_nearest_out, ft_nearest_in, n0, n1, n2, N, spacing[0], spacing[1],
                     spacing[2], width_mm, psi);
  return check_launch("hp_sbm_psi_f64");
}

int hp_zoom_linear_f64(const double* in, int32_t n0, int32_t n1, int32_t n2, const int32_t* oshape,
                       const int32_t* mshape, double cutoff, double* out, uint8_t* positive,
                       void* stream) {
  if (!in || !oshape || !mshape || !out || n0 <= 0 || n1 <= 0 || n2 <= 0) {
    set_error("hp_zoom_linear_f64: bad arguments");
    return -1;
  }
  const int32_t n[3] = {n0, n1, n2};
  double z[3];
  for (int a = 0; a < 3; ++a) {
    if (oshape[a] <= 0 || mshape[a] <= 0) {
      set_error("hp_zoom_linear_f64: empty output shape");
      return -1;
    }
    z[a] = oshape[a] > 1 ? (double)(n[a] - 1) / (double)(oshape[a] - 1) : 1.0;
  }
  const int64_t M = (int64_t)mshape[0] * mshape[1] * mshape[2];
  hipLaunchKernelGGL(zoom_linear_kernel, dim3(grid_of(M)), dim3(256), 0, (hipStream_t)stream, in,
                     n0, n1, n2, oshape[0], oshape[1], oshape[2], z[0], z[1], z[2], mshape[0],
                     mshape[1], mshape[2], cutoff, out, positive);
  return check_launch("hp_zoom_linear_f64");
}

int hp_zoom_nearest_i32(const int32_t* in, int32_t n0, int32_t n1, int32_t n2,
                        const int32_t* oshape, int32_t* out, void* stream) {
  if (!in || !oshape || !out || n0 <= 0 || n1 <= 0 || n2 <= 0 || oshape[0] <= 0 ||
      oshape[1] <= 0 || oshape[2] <= 0) {
    set_error("hp_zoom_nearest_i32: bad arguments");
    return -1;
  }
  const int32_t n[3] = {n0, n1, n2};
  double z[3];
  for (int a = 0; a < 3; ++a)
    z[a] = oshape[a] > 1 ? (double)(n[a] - 1) / (double)(oshape[a] - 1) : 1.0;
  const int64_t M = (int64_t)oshape[0] * oshape[1] * oshape[2];
  hipLaunchKernelGGL(zoom_nearest_kernel, dim3(grid_of(M)), dim3(256), 0, (hipStream_t)stream, in,
                     n0, n1, n2, oshape[0], oshape[1], oshape[2], z[0], z[1], z[2], out);
  return check_launch("hp_zoom_nearest_i32");
}

int hp_label_sigma_f64(const int32_t* labels, int64_t n, const int32_t* keys, const double* vals,
                       int32_t n_keys, double* sigma, uint8_t* tissue, void* stream) {
  if (!labels || !sigma || !tissue || n < 0 || n_keys < 0 || (n_keys > 0 && (!keys || !vals))) {
    set_error("hp_label_sigma_f64: bad arguments");
    return -1;
  }
  if (n == 0) return 0;
  hipLaunchKernelGGL(label_sigma_kernel, dim3(grid_of(n)), dim3(256), 0, (hipStream_t)stream,
                     labels, n, keys, vals, n_keys, sigma, tissue);
  return check_launch("hp_label_sigma_f64");
}

int hp_extend_sigma_f64(const int32_t* ft_tissue, const uint8_t* tissue, const uint8_t* active,
                        const double* sigma_in, int32_t n0, int32_t n1, int32_t n2,
                        double* sigma_out, void* stream) {
  if (!ft_tissue || !tissue || !active || !sigma_in || !sigma_out || sigma_in == sigma_out ||
      n0 <= 0 || n1 <= 0 || n2 <= 0) {
    set_error("hp_extend_sigma_f64: bad arguments");
    return -1;
  }
  const int64_t N = (int64_t)n0 * n1 * n2;
  hipLaunchKernelGGL(extend_sigma_kernel, dim3(grid_of(N)), dim3(256), 0, (hipStream_t)stream,
                     ft_tissue, tissue, active, sigma_in, n1, n2, N, sigma_out);
  return check_launch("hp_extend_sigma_f64");
}

int hp_column_stats(const uint8_t* flag, int32_t ni, int32_t nj, int32_t nk, int32_t* k0,
                    int32_t* run, int32_t* cnt, void* stream) {
  if (!flag || ni <= 0 || nj <= 0 || nk <= 0 || (!k0 && !run && !cnt)) {
    set_error("hp_column_stats: bad arguments");
    return -1;
  }
  hipLaunchKernelGGL(column_stats_kernel, dim3(ceil_div((int64_t)ni * nj, 256)), dim3(256), 0,
                     (hipStream_t)stream, flag, ni, nj, nk, k0, run, cnt);
  return check_launch("hp_column_stats");
}

int hp_scan_i32(const int32_t* in, int64_t n, int64_t* out, void* stream) {
  if (!in || !out || n <= 0) {
    set_error("hp_scan_i32: bad arguments");
    return -1;
  }
  hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, (hipStream_t)stream, in, n, out);
  return check_launch("hp_scan_i32");
}

int hp_column_emit(const uint8_t* flag, int32_t ni, int32_t nj, int32_t nk, const int64_t* base,
                   int32_t* list, int32_t* id_map, const int32_t* k0, const int64_t* col_off,
                   int64_t* cell_of_row, void* stream) {
  if (!flag || !base || ni <= 0 || nj <= 0 || nk <= 0 || (cell_of_row && (!k0 || !col_off))) {
    set_error("hp_column_emit: bad arguments");
    return -1;
  }
  hipLaunchKernelGGL(column_emit_kernel, dim3(ceil_div((int64_t)ni * nj, 256)), dim3(256), 0,
                     (hipStream_t)stream, flag, ni, nj, nk, base, list, id_map, k0, col_off,
                     cell_of_row);
  return check_launch("hp_column_emit");
}

int hp_surface_flags(const uint8_t* tissue, int32_t ni, int32_t nj, int32_t nk, uint8_t* surf,
                     void* stream) {
  if (!tissue || !surf || ni <= 0 || nj <= 0 || nk <= 0) {
    set_error("hp_surface_flags: bad arguments");
    return -1;
  }
  hipLaunchKernelGGL(surface_flag_kernel, dim3(grid_of((int64_t)ni * nj * nk)), dim3(256), 0,
                     (hipStream_t)stream, tissue, ni, nj, nk, surf);
  return check_launch("hp_surface_flags");
}

int hp_select_rank(const uint8_t* flag, const int64_t* scan, int64_t n_cols, int32_t nk,
                   int64_t rank, int64_t* out, void* stream) {
  if (!flag || !scan || !out || n_cols <= 0 || nk <= 0 || rank < 0) {
    set_error("hp_select_rank: bad arguments");
    return -1;
  }
  hipLaunchKernelGGL(select_rank_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, flag, scan,
                     n_cols, nk, rank, out);
  return check_launch("hp_select_rank");
}

int hp_torso_faces_f64(const double* sigma, const double* psi, const uint8_t* active, int32_t ni,
                       int32_t nj, int32_t nk, double h_m, const int32_t* col_k0,
                       const int32_t* col_len, const int64_t* col_off, int64_t pin_flat,
                       float* gx, float* gy, float* gz, float* diag, double* g64,
                       int64_t n_cells, unsigned long long* bad, void* stream) {
  if (!sigma || !active || !col_k0 || !col_len || !col_off || !gx || !gy || !gz || !diag ||
      !bad || ni <= 0 || nj <= 0 || nk <= 0 || !(h_m > 0.0) || n_cells < 0) {
    set_error("hp_torso_faces_f64: bad arguments");
    return -1;
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, s, bad);
  const double h2 = h_m * h_m;
  hipLaunchKernelGGL(torso_faces_kernel, dim3(grid_of((int64_t)ni * nj * nk)), dim3(256), 0, s,
                     sigma, psi, active, ni, nj, nk, h2, col_k0, col_len, col_off, pin_flat, gx,
                     gy, gz, diag, g64, n_cells, bad);
  return check_launch("hp_torso_faces_f64");
}

}  // extern "C"
