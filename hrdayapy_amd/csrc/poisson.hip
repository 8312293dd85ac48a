// Torso Poisson forward solve: structured 7-point operator, Jacobi-PCG with
// FP32 SpMV and FP64 inner products, FP64 residual for iterative refinement,
// the heart source term and the FP64 electrode/lead accumulation.
//
// Operator (reference: _assemble_stiffness, hrdayapy/ecg/functions/
// build_torso_grid.py:276-397, pinned as in compute_bspm.py:327-331):
//   (A x)_c = diag_c x_c - sum_faces g_f x_nb,  g_f = harm(sigma_c, sigma_nb)/h^2,
//   diag_c = sum of g_f over the six faces (Neumann at the box edge and air),
//   row `pin` is the identity (reference potential).
// Both the FP32 and the FP64 products are evaluated as sum_f g_f (x_c - x_nb),
// exact for constants: with the reference cell on the body surface the
// solution carries a large near-constant part, and a rounded FP32 diagonal
// would turn it into an error that FP64 refinement cannot remove.
// Layout: every (i,j) column of the box stores the contiguous z-run
// [k0, k0+len) that covers its active cells; cell (i,j,k) of the run has index
// off[i*nj+j] + k - k0.  Cells inside a run that are not active (holes) have
// zero conductances and diag 1, so they stay at zero.  g_x/g_y/g_z hold the
// conductance of the +x/+y/+z face of each stored cell.
//
// The PCG loop runs without host round trips: alpha/beta, the stopping test
// and an iteration counter live in device memory, every kernel returns at
// once when the `done` flag is set, and reductions are finished by the last
// block to arrive (fixed summation order: deterministic results).
#include "poisson_common.h"

namespace hp {

__global__ void __launch_bounds__(PBLOCK) spmv_f32_kernel(hp_torso t, const float* __restrict__ x,
                                                          float* __restrict__ y) {
  for_cells(t, [&](int c, int i, int j, int k, const ColNbr& m) {
    y[c] = apply_row_f32(t, x, c, i, j, k, m);
  });
}

// FP64 row of the operator in zero-row-sum form: sum_f g_f (x_c - x_nb); a
// cell without any conducting face (hole, isolated voxel) and the pinned cell
// are identity rows.  The diagonal is the exact FP64 sum of the FP32 face
// conductances, so every interior row sums to zero exactly.
__device__ __forceinline__ double apply_row_f64(const hp_torso& t, const double* __restrict__ x,
                                                int c, int i, int j, int k, const ColNbr& m) {
  const double xc = x[c];
  if (c == t.pin) return xc;
  double acc = 0.0, gs = 0.0;
  double g;
  if (k + 1 < m.c.k0 + m.c.len && (g = t.gz[c]) != 0.0) { acc += g * (xc - x[c + 1]); gs += g; }
  if (k - 1 >= m.c.k0 && (g = t.gz[c - 1]) != 0.0) { acc += g * (xc - x[c - 1]); gs += g; }
  int n;
  if ((n = cell_at(m.xp, k)) >= 0 && (g = t.gx[c]) != 0.0) { acc += g * (xc - x[n]); gs += g; }
  if ((n = cell_at(m.xm, k)) >= 0 && (g = t.gx[n]) != 0.0) { acc += g * (xc - x[n]); gs += g; }
  if ((n = cell_at(m.yp, k)) >= 0 && (g = t.gy[c]) != 0.0) { acc += g * (xc - x[n]); gs += g; }
  if ((n = cell_at(m.ym, k)) >= 0 && (g = t.gy[n]) != 0.0) { acc += g * (xc - x[n]); gs += g; }
  return gs == 0.0 ? xc : acc;
}

// r = b - A x in FP64 and ||r||^2, ||b||^2 -> out[0], out[1]
__global__ void __launch_bounds__(PBLOCK) residual_f64_kernel(hp_torso t, const double* __restrict__ b,
                                                              const double* __restrict__ x,
                                                              double* __restrict__ r,
                                                              double* part, int* counter,
                                                              double* out) {
  double v[2] = {0.0, 0.0};
  for_cells(t, [&](int c, int i, int j, int k, const ColNbr& m) {
    const double rc = b[c] - apply_row_f64(t, x, c, i, j, k, m);
    r[c] = rc;
    v[0] += rc * rc;
    v[1] += b[c] * b[c];
  });
  block_reduce_last<2>(v, part, counter, out);
}

// PCG init for A d = r (FP32), d = 0: z = r/diag (stored over r: the solve
// carries z = M^-1 r instead of r, so the product kernel reads two arrays per
// neighbour and divides nothing), p = z, rz, rr0 = ||r||^2.
__global__ void __launch_bounds__(PBLOCK) pcg_init_kernel(hp_torso t, float* __restrict__ r,
                                                          float* __restrict__ d,
                                                          float* __restrict__ p, double* part,
                                                          int* counter, double* sc, double tol,
                                                          int maxit, int mode) {
  double v[3] = {0.0, 0.0, 0.0};
  for_cells(t, [&](int c, int i, int j, int k, const ColNbr& m) {
    const float rc = r[c];
    const float z = rc / t.diag[c];
    r[c] = z;
    d[c] = 0.0f;
    p[c] = z;
    v[0] += (double)rc * (double)z;
    v[1] += (double)rc * (double)rc;
    v[2] += (double)z * (double)z;
  });
  double red[3];
  if (block_reduce_last<3>(v, part, counter, red) && threadIdx.x == 0) {
    sc[S_RZ] = red[0];
    sc[S_RR0] = red[1];
    sc[S_RR] = red[1];
    sc[S_ZZ] = red[2];
    sc[S_MODE] = (double)mode;
    sc[S_NCELL] = (double)t.n_cells;
    sc[S_TOL2] = tol * tol;
    sc[S_ITER] = 0.0;
    sc[S_MAXIT] = (double)maxit;
    sc[S_BETA] = 0.0;
    const bool conv = mode == 0 ? red[1] == 0.0 : red[2] <= tol * tol * (double)t.n_cells;
    sc[S_DONE] = (conv || maxit <= 0) ? 1.0 : 0.0;
  }
}

// p = z + beta p (skipped on the first iteration: beta = 0 and p = z already),
// Ap = A p, pAp.  z (held in the r array) and the old p of a neighbour give
// its new p on the fly.
__global__ void __launch_bounds__(PBLOCK) pcg_spmv_kernel(hp_torso t, const float* __restrict__ z,
                                                          float* __restrict__ p,
                                                          float* __restrict__ Ap, double* part,
                                                          int* counter, double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const float beta = (float)sc[S_BETA];
  const bool first = sc[S_ITER] == 0.0;
  double v[1] = {0.0};
  auto pnew = [&](int n) -> float { return first ? p[n] : z[n] + beta * p[n]; };
  for_cells(t, [&](int c, int i, int j, int k, const ColNbr& m) {
    const float pc = pnew(c);
    const float acc = row_f32(t, c, k, m, pc, pnew);
    Ap[c] = acc;
    v[0] += (double)pc * (double)acc;
  });
  // p itself is not written here (other blocks still read the old p of their
  // neighbours); pcg_update materialises the same z + beta*p next.
  double red[1];
  if (block_reduce_last<1>(v, part, counter, red) && threadIdx.x == 0) {
    sc[S_PAP] = red[0];
    if (!(fabs(red[0]) > 1e-300)) sc[S_DONE] = 2.0;
    else sc[S_ALPHA] = sc[S_RZ] / red[0];
  }
}

// p <- z + beta p (materialised), d += alpha p, z -= alpha Ap/diag (z = M^-1 r
// held in the r array), rz, rr with r = diag z; finish: beta, stopping test,
// iteration count.
__global__ void __launch_bounds__(PBLOCK) pcg_update_kernel(hp_torso t, float* __restrict__ z,
                                                            float* __restrict__ d,
                                                            float* __restrict__ p,
                                                            const float* __restrict__ Ap,
                                                            double* part, int* counter,
                                                            double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const float alpha = (float)sc[S_ALPHA];
  const float beta = (float)sc[S_BETA];
  const bool first = sc[S_ITER] == 0.0;
  double v[3] = {0.0, 0.0, 0.0};
  for_cells(t, [&](int c, int i, int j, int k, const ColNbr& m) {
    const float zc = z[c];
    const float pc = first ? p[c] : zc + beta * p[c];
    p[c] = pc;
    d[c] += alpha * pc;
    const float dg = t.diag[c];
    const float zn = zc - alpha * (Ap[c] / dg);
    z[c] = zn;
    const double rn = (double)dg * (double)zn;
    v[0] += rn * (double)zn;
    v[1] += rn * rn;
    v[2] += (double)zn * (double)zn;
  });
  double red[3];
  if (block_reduce_last<3>(v, part, counter, red) && threadIdx.x == 0) {
    sc[S_BETA] = red[0] / sc[S_RZ];
    sc[S_RZ] = red[0];
    sc[S_RR] = red[1];
    sc[S_ZZ] = red[2];
    sc[S_ITER] += 1.0;
    // mode 0: relative residual; mode 1: rms of z = M^-1 r (compute_bspm.py:222-256)
    const bool conv = sc[S_MODE] == 0.0 ? red[1] <= sc[S_TOL2] * sc[S_RR0]
                                        : red[2] <= sc[S_TOL2] * sc[S_NCELL];
    if (conv || sc[S_ITER] >= sc[S_MAXIT]) sc[S_DONE] = 1.0;
  }
}

// x64 += scale * d
__global__ void axpy_f64_kernel(hp_torso t, double* __restrict__ x, const float* __restrict__ d,
                                double scale) {
  for_cells(t, [&](int c, int, int, int, const ColNbr&) { x[c] += scale * (double)d[c]; });
}

// r32 = r64 * scale
__global__ void cast_scale_kernel(hp_torso t, const double* __restrict__ r64, float* __restrict__ r32,
                                  double scale) {
  for_cells(t, [&](int c, int, int, int, const ColNbr&) { r32[c] = (float)(r64[c] * scale); });
}

// ---------------------------------------------------------------------------
// Heart source: 6-neighbour unit-weight graph Laplacian of Vm on the brick
// layout (build_heart_laplacian, compute_bspm.py:88-119), then coarse rows
// gather their fine voxels (fixed order) minus the mean leak
// (compute_bspm.py:386-419).
// ---------------------------------------------------------------------------
template <int BX, int BY, int BZ, typename TV>
__global__ void __launch_bounds__(BX * BY * BZ) heart_lap6_kernel(hp_bricks g, const TV* __restrict__ V,
                                                                  double* __restrict__ lap,
                                                                  double vscale) {
  constexpr int BV = BX * BY * BZ;
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t slot = b * BV + tid;
  if (g.label[slot] == 0xFF) {
    lap[slot] = 0.0;
    return;
  }
  const int lz = tid % BZ, ly = (tid / BZ) % BY, lx = tid / (BY * BZ);
  const double vc = (double)V[slot];
  double acc = 0.0;
  const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    int x = lx + off[q][0], y = ly + off[q][1], z = lz + off[q][2];
    const int dbx = x < 0 ? -1 : (x >= BX ? 1 : 0);
    const int dby = y < 0 ? -1 : (y >= BY ? 1 : 0);
    const int dbz = z < 0 ? -1 : (z >= BZ ? 1 : 0);
    int64_t nb = b;
    if (dbx | dby | dbz) nb = g.nbr[b * 27 + (dbx + 1) * 9 + (dby + 1) * 3 + (dbz + 1)];
    if (nb < 0) continue;
    x -= dbx * BX;
    y -= dby * BY;
    z -= dbz * BZ;
    const int64_t s = nb * BV + (x * BY + y) * BZ + z;
    if (g.label[s] != 0xFF) acc += (double)V[s] - vc;
  }
  lap[slot] = acc * vscale;
}

// mean over the valid fine voxels of lap -> out[0]
__global__ void __launch_bounds__(PBLOCK) masked_sum_kernel(const double* __restrict__ lap,
                                                            const int32_t* __restrict__ slots,
                                                            int64_t n, double* part, int* counter,
                                                            double* out) {
  double v[1] = {0.0};
  for (int64_t q = (int64_t)blockIdx.x * PBLOCK + threadIdx.x; q < n; q += (int64_t)gridDim.x * PBLOCK)
    v[0] += lap[slots[q]];
  block_reduce_last<1>(v, part, counter, out);
}

// b[row] = coeff*(sum_{fine in row} lap - cnt*mean)/n_expected; b[pin] = 0
__global__ void source_gather_kernel(const double* __restrict__ lap, const int32_t* __restrict__ rows,
                                     const int32_t* __restrict__ ptr, const int32_t* __restrict__ fine,
                                     int64_t n_rows_hit, const double* __restrict__ sum_n,
                                     double inv_nvalid, double coeff, double inv_nexp,
                                     double* __restrict__ b, int64_t pin) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_rows_hit) return;
  const double mean = sum_n[0] * inv_nvalid;
  double s = 0.0;
  for (int e = ptr[q]; e < ptr[q + 1]; ++e) s += lap[fine[e]] - mean;
  const int row = rows[q];
  b[row] = row == pin ? 0.0 : coeff * s * inv_nexp;
}

// Lead-field ECG: out[f][e] = sum_r W[r][e] * V[f][slot[r]] over the rows r
// (heart voxels), FP32 products accumulated in FP64.  Pass 1: block q sums
// its row range into part[q][f][e]; pass 2 adds the block partials in block
// order (deterministic).  Threads of a block take consecutive electrodes, so
// a row of W is read coalesced; the block's V values are staged in LDS.
constexpr int LF_ROWS = 256, LF_MAXF = 16;
__global__ void __launch_bounds__(1024) leadfield_partial_kernel(
    const float* __restrict__ W, const int32_t* __restrict__ row_slot, int64_t n_rows, int ne,
    const float* __restrict__ V, int64_t ldv, int n_frames, int64_t rows_per_block,
    double* __restrict__ part) {
  __shared__ double vs[LF_MAXF][LF_ROWS];
  const int e = blockIdx.y * blockDim.x + threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < n_rows ? r0 + rows_per_block : n_rows;
  double acc[LF_MAXF];
#pragma unroll
  for (int f = 0; f < LF_MAXF; ++f) acc[f] = 0.0;
  for (int64_t c0 = r0; c0 < r1; c0 += LF_ROWS) {
    const int nr = (int)(r1 - c0 < LF_ROWS ? r1 - c0 : LF_ROWS);
    __syncthreads();
    for (int q = threadIdx.x; q < nr * n_frames; q += blockDim.x) {
      const int f = q / nr, rr = q % nr;
      vs[f][rr] = (double)V[(int64_t)f * ldv + row_slot[c0 + rr]];
    }
    __syncthreads();
    if (e < ne) {
      for (int q = 0; q < nr; ++q) {
        const double w = (double)W[(c0 + q) * ne + e];
#pragma unroll
        for (int f = 0; f < LF_MAXF; ++f)
          if (f < n_frames) acc[f] = fma(w, vs[f][q], acc[f]);
      }
    }
  }
  if (e < ne)
    for (int f = 0; f < n_frames; ++f)
      part[((int64_t)blockIdx.x * n_frames + f) * ne + e] = acc[f];
}

__global__ void leadfield_reduce_kernel(const double* __restrict__ part, int n_blocks, int ne,
                                        int n_frames, double* __restrict__ out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;  // f*ne + e
  if (q >= ne * n_frames) return;
  double s = 0.0;
  for (int b = 0; b < n_blocks; ++b) s += part[(int64_t)b * ne * n_frames + q];
  out[q] = s;
}

// out[e] = sum_k w[k] * phi[idx[k]] over k in [ptr[e], ptr[e+1])  (FP64)
__global__ void leads_kernel(const double* __restrict__ phi, const int32_t* __restrict__ ptr,
                             const int32_t* __restrict__ idx, const double* __restrict__ w,
                             double* __restrict__ out, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double s = 0.0;
  for (int k = ptr[e]; k < ptr[e + 1]; ++k) s += w[k] * phi[idx[k]];
  out[e] = s;
}

}  // namespace hp

namespace hp {
template <typename TV>
static int heart_lap6(const hp_bricks* g, const TV* V, double* lap, double vscale,
                      void* stream) {
  if (!g || g->n_bricks <= 0) {
    set_error("hp_heart_lap6: empty brick grid");
    return -1;
  }
  const int key = g->bx * 100 + g->by * 10 + g->bz;
  const dim3 grid((unsigned)g->n_bricks);
  hipStream_t st = (hipStream_t)stream;
  switch (key) {
    case 444: hipLaunchKernelGGL((heart_lap6_kernel<4, 4, 4, TV>), grid, dim3(64), 0, st, *g, V, lap, vscale); break;
    case 844: hipLaunchKernelGGL((heart_lap6_kernel<8, 4, 4, TV>), grid, dim3(128), 0, st, *g, V, lap, vscale); break;
    case 884: hipLaunchKernelGGL((heart_lap6_kernel<8, 8, 4, TV>), grid, dim3(256), 0, st, *g, V, lap, vscale); break;
    case 888: hipLaunchKernelGGL((heart_lap6_kernel<8, 8, 8, TV>), grid, dim3(512), 0, st, *g, V, lap, vscale); break;
    default:
      set_error("hp_heart_lap6: unsupported brick shape");
      return -1;
  }
  return check_launch("hp_heart_lap6");
}
}  // namespace hp

using namespace hp;

extern "C" {

int hp_poisson_spmv_f32(const hp_torso* t, const float* x, float* y, void* stream) {
  if (!torso_ok(t, "hp_poisson_spmv_f32")) return -1;
  hipLaunchKernelGGL(spmv_f32_kernel, dim3(grid_blocks(t)), dim3(PBLOCK), 0, (hipStream_t)stream,
                     *t, x, y);
  return check_launch("hp_poisson_spmv_f32");
}

int hp_poisson_residual_f64(const hp_torso* t, const double* b, const double* x, double* r,
                            hp_pcg_ws* ws, void* stream) {
  if (!torso_ok(t, "hp_poisson_residual_f64") || !ws) return -1;
  hipLaunchKernelGGL(residual_f64_kernel, dim3(grid_blocks(t)), dim3(PBLOCK), 0,
                     (hipStream_t)stream, *t, b, x, r, ws->part, ws->counter, ws->out);
  return check_launch("hp_poisson_residual_f64");
}

int hp_pcg_begin_f32(const hp_torso* t, float* r, float* d, float* p, hp_pcg_ws* ws,
                     double tol, int max_iter, int mode, void* stream) {
  if (!torso_ok(t, "hp_pcg_begin_f32") || !ws) return -1;
  hipLaunchKernelGGL(pcg_init_kernel, dim3(grid_blocks(t)), dim3(PBLOCK), 0, (hipStream_t)stream,
                     *t, r, d, p, ws->part, ws->counter, ws->scalars, tol, max_iter, mode);
  return check_launch("hp_pcg_begin_f32");
}

int hp_pcg_iterate_f32(const hp_torso* t, float* r, float* d, float* p, float* Ap, hp_pcg_ws* ws,
                       int n_iter, void* stream) {
  if (!torso_ok(t, "hp_pcg_iterate_f32") || !ws || n_iter < 0) return -1;
  const int nb = grid_blocks(t);
  hipStream_t st = (hipStream_t)stream;
  for (int it = 0; it < n_iter; ++it) {
    hipLaunchKernelGGL(pcg_spmv_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, r, p, Ap, ws->part,
                       ws->counter, ws->scalars);
    hipLaunchKernelGGL(pcg_update_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, r, d, p, Ap,
                       ws->part, ws->counter, ws->scalars);
  }
  return check_launch("hp_pcg_iterate_f32");
}

int hp_refine_axpy_f64(const hp_torso* t, double* x, const float* d, double scale, void* stream) {
  if (!torso_ok(t, "hp_refine_axpy_f64")) return -1;
  hipLaunchKernelGGL(axpy_f64_kernel, dim3(grid_blocks(t)), dim3(PBLOCK), 0, (hipStream_t)stream,
                     *t, x, d, scale);
  return check_launch("hp_refine_axpy_f64");
}

int hp_cast_scale_f32(const hp_torso* t, const double* r64, float* r32, double scale,
                      void* stream) {
  if (!torso_ok(t, "hp_cast_scale_f32")) return -1;
  hipLaunchKernelGGL(cast_scale_kernel, dim3(grid_blocks(t)), dim3(PBLOCK), 0,
                     (hipStream_t)stream, *t, r64, r32, scale);
  return check_launch("hp_cast_scale_f32");
}

int hp_heart_lap6_f32(const hp_bricks* g, const float* V, double* lap, double vscale,
                      void* stream) {
  return heart_lap6<float>(g, V, lap, vscale, stream);
}

int hp_heart_lap6_f64(const hp_bricks* g, const double* V, double* lap, double vscale,
                      void* stream) {
  return heart_lap6<double>(g, V, lap, vscale, stream);
}

int hp_leadfield_apply_f32(const float* W, const int32_t* row_slot, int64_t n_rows, int ne,
                           const float* V, int64_t ldv, int n_frames, double* part,
                           int64_t part_len, double* out, void* stream) {
  if (!W || !row_slot || !V || !part || !out || n_rows <= 0 || ne <= 0 || n_frames < 1 ||
      n_frames > LF_MAXF) {
    set_error("hp_leadfield_apply_f32: bad arguments (1..%d frames)", LF_MAXF);
    return -1;
  }
  int nb = 1024;
  const int64_t rpb = (n_rows + nb - 1) / nb;
  nb = (int)((n_rows + rpb - 1) / rpb);
  if ((int64_t)nb * n_frames * ne > part_len) {
    set_error("hp_leadfield_apply_f32: partial buffer needs %lld doubles",
              (long long)nb * n_frames * ne);
    return -1;
  }
  hipStream_t st = (hipStream_t)stream;
  // one thread per electrode (up to 1024 per block), whole waves
  const int bt = ne <= 1024 ? (ne + 63) / 64 * 64 : 1024;
  hipLaunchKernelGGL(leadfield_partial_kernel, dim3(nb, (ne + bt - 1) / bt), dim3(bt), 0, st, W,
                     row_slot, n_rows, ne, V, ldv, n_frames, rpb, part);
  const int rc = check_launch("hp_leadfield_apply_f32/partial");
  if (rc) return rc;
  hipLaunchKernelGGL(leadfield_reduce_kernel, dim3(ceil_div((int64_t)ne * n_frames, 256)),
                     dim3(256), 0, st, part, nb, ne, n_frames, out);
  return check_launch("hp_leadfield_apply_f32/reduce");
}

int hp_bspm_source_f64(const double* lap, const int32_t* valid_slots, int64_t n_valid,
                       const int32_t* rows, const int32_t* ptr, const int32_t* fine,
                       int64_t n_rows_hit, double coeff, double n_expected, double* b,
                       int64_t pin, hp_pcg_ws* ws, void* stream) {
  if (!ws || n_valid <= 0 || n_expected <= 0.0) {
    set_error("hp_bspm_source_f64: bad arguments");
    return -1;
  }
  hipStream_t st = (hipStream_t)stream;
  int nb = (int)((n_valid + PBLOCK - 1) / PBLOCK);
  if (nb > 1024) nb = 1024;
  hipLaunchKernelGGL(masked_sum_kernel, dim3(nb), dim3(PBLOCK), 0, st, lap, valid_slots, n_valid,
                     ws->part, ws->counter, ws->out);
  int rc = check_launch("hp_bspm_source_f64/sum");
  if (rc) return rc;
  if (n_rows_hit > 0)
    hipLaunchKernelGGL(source_gather_kernel, dim3(ceil_div(n_rows_hit, 256)), dim3(256), 0, st,
                       lap, rows, ptr, fine, n_rows_hit, ws->out, 1.0 / (double)n_valid, coeff,
                       1.0 / n_expected, b, pin);
  return check_launch("hp_bspm_source_f64/gather");
}

int hp_leads_accumulate_f64(const double* phi, const int32_t* ptr, const int32_t* idx,
                            const double* w, double* out, int64_t n_electrodes, void* stream) {
  if (n_electrodes <= 0) return 0;
  hipLaunchKernelGGL(leads_kernel, dim3(ceil_div(n_electrodes, 256)), dim3(256), 0,
                     (hipStream_t)stream, phi, ptr, idx, w, out, n_electrodes);
  return check_launch("hp_leads_accumulate_f64");
}

}  // extern "C"
