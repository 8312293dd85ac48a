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
// once when the `done` flag is set.  A reducing kernel writes one partial sum
// per block and a one-block finish kernel after it sums them in block order
// and updates the scalars (fixed order: deterministic results; no fences).
#include <mutex>

#include "poisson_common.h"

namespace hp {

// Scratch of the allocation-free residual (hp_poisson_residual_f64_noalloc):
// the block partial sums of residual_f64_kernel, two per block for at most
// PMAX_BLOCKS blocks (64 KiB on every GPU the library runs on, held until the
// process ends).  Host code never passes `scratch` itself to a kernel: on the
// host the name is only a stand-in for the symbol; the device address comes
// from hipGetSymbolAddress, once per device (scratch_address below).
static __device__ __align__(256) double scratch[2 * PMAX_BLOCKS];
constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[SCRATCH_MAX_DEVICES] = {};

// Device address of `scratch` on the current device (looked up on the first
// call on each device, with that device current); nullptr with the error set.
static double* scratch_address(const char* what) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
    set_error("%s: no current HIP device (or device index >= %d)", what, SCRATCH_MAX_DEVICES);
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("%s: hipGetSymbolAddress(scratch) on device %d: %s", what, dev,
                hipGetErrorString(e));
      return nullptr;
    }
    g_scratch_dev[dev] = (double*)p;
  }
  return g_scratch_dev[dev];
}

__global__ void __launch_bounds__(PBLOCK) spmv_f32_kernel(hp_torso t, const float* __restrict__ x,
                                                          float* __restrict__ y) {
  const uint32_t vb = (uint32_t)t.n_cells * 4u;
  const buf_t X = buf_rsrc(x, vb), Y = buf_rsrc(y, vb);
  struct L {
    float g[6], x[6], xc;
  };
  for_cells_u<2>(
      t,
      [&](uint32_t c, int k, const Chunk& ch) {
        L l;
        const RowIdx ix = row_idx(c, k, ch);
        load_g(t, ix, l.g);
#pragma unroll
        for (int f = 0; f < 6; ++f) l.x[f] = buf_ld<float>(X, ix.n[f]);
        l.xc = buf_ld<float>(X, c);
        return l;
      },
      [&](uint32_t c, int, const Chunk&, const L& l) { buf_st(Y, c, row_sum(t, c, l.xc, l.g, l.x)); });
}

// FP64 row of the operator in zero-row-sum form: sum_f g_f (x_c - x_nb), faces
// in the order +z, -z, +x, -x, +y, -y; a cell without any conducting face
// (hole, isolated voxel) and the pinned cell are identity rows.  The diagonal
// is the exact FP64 sum of the FP32 face conductances, so every interior row
// sums to zero exactly.  Absent neighbours read g = 0 and x = 0 (T_OOB).
__device__ __forceinline__ double apply_row_f64(const hp_torso& t, const double* __restrict__ x,
                                                uint32_t c, int k, const Chunk& nb) {
  const uint32_t gb = (uint32_t)t.n_cells * 4u;
  const buf_t Gx = buf_rsrc(t.gx, gb), Gy = buf_rsrc(t.gy, gb), Gz = buf_rsrc(t.gz, gb);
  const buf_t X = buf_rsrc(x, (uint32_t)t.n_cells * 8u);
  const uint32_t n0 = nb(0, k + 1), n1 = nb(0, k - 1), n2 = nb(1, k), n3 = nb(2, k),
                 n4 = nb(3, k), n5 = nb(4, k);
  const double g0 = buf_ld<float>(Gz, n0 == T_OOB ? T_OOB : c), g1 = buf_ld<float>(Gz, n1),
               g2 = buf_ld<float>(Gx, n2 == T_OOB ? T_OOB : c), g3 = buf_ld<float>(Gx, n3),
               g4 = buf_ld<float>(Gy, n4 == T_OOB ? T_OOB : c), g5 = buf_ld<float>(Gy, n5);
  const double xc = buf_ld<double>(X, c);
  const double x0 = buf_ld<double>(X, n0), x1 = buf_ld<double>(X, n1), x2 = buf_ld<double>(X, n2),
               x3 = buf_ld<double>(X, n3), x4 = buf_ld<double>(X, n4), x5 = buf_ld<double>(X, n5);
  double acc = 0.0, gs = 0.0;
  acc += g0 * (xc - x0); gs += g0;
  acc += g1 * (xc - x1); gs += g1;
  acc += g2 * (xc - x2); gs += g2;
  acc += g3 * (xc - x3); gs += g3;
  acc += g4 * (xc - x4); gs += g4;
  acc += g5 * (xc - x5); gs += g5;
  return ((int64_t)c == t.pin || gs == 0.0) ? xc : acc;
}

// r = b - A x in FP64 and ||r||^2, ||b||^2 -> out[0], out[1]
__global__ void __launch_bounds__(PBLOCK) residual_f64_kernel(hp_torso t, const double* __restrict__ b,
                                                              const double* __restrict__ x,
                                                              double* __restrict__ r,
                                                              double* part) {
  double v[2] = {0.0, 0.0};
  for_cells(t, [&](int c, int k, const Chunk& ch) {
    const double rc = b[c] - apply_row_f64(t, x, c, k, ch);
    r[c] = rc;
    v[0] += rc * rc;
    v[1] += b[c] * b[c];
  });
  block_partials<2>(v, part);
}

// out[0..NV) = sums of the nb block partials of the kernel before
template <int NV>
__global__ void __launch_bounds__(PBLOCK) finish_sums_kernel(const double* __restrict__ part, int nb,
                                                             double* __restrict__ out) {
  double red[NV];
  sum_partials<NV>(part, nb, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < NV; ++q) out[q] = red[q];
}

// PCG init for A d = r (FP32), d = 0: z = r/diag (stored over r: the solve
// carries z = M^-1 r instead of r, so the product kernel reads two arrays per
// neighbour and divides nothing), p = z, rz, rr0 = ||r||^2.
__global__ void __launch_bounds__(PBLOCK) pcg_init_kernel(hp_torso t, float* __restrict__ r,
                                                          float* __restrict__ d,
                                                          float* __restrict__ p, double* part) {
  double v[3] = {0.0, 0.0, 0.0};
  for_cells(t, [&](int c, int, const Chunk&) {
    const float rc = r[c];
    const float z = rc / t.diag[c];
    r[c] = z;
    d[c] = 0.0f;
    p[c] = z;
    v[0] += (double)rc * (double)z;
    v[1] += (double)rc * (double)rc;
    v[2] += (double)z * (double)z;
  });
  block_partials<3>(v, part);
}

__global__ void __launch_bounds__(PBLOCK) pcg_init_finish_kernel(const double* __restrict__ part,
                                                                 int nb, double* sc, double tol,
                                                                 int maxit, int mode,
                                                                 int64_t n_cells) {
  double red[3];
  sum_partials<3>(part, nb, red);
  if (threadIdx.x == 0) {
    sc[S_RZ] = red[0];
    sc[S_RR0] = red[1];
    sc[S_RR] = red[1];
    sc[S_ZZ] = red[2];
    sc[S_MODE] = (double)mode;
    sc[S_NCELL] = (double)n_cells;
    sc[S_TOL2] = tol * tol;
    sc[S_ITER] = 0.0;
    sc[S_MAXIT] = (double)maxit;
    sc[S_BETA] = 0.0;
    sc[S_ZSTAT] = 0.0;
    // the reference tests its metric only after an iteration (compute_bspm.py:
    // 240-252); a zero residual ends a relative-residual solve at once
    const bool conv = mode == 0 && red[1] == 0.0;
    sc[S_DONE] = (conv || maxit <= 0) ? 1.0 : 0.0;
  }
}

// p = z + beta p (on the first iteration beta = 0 and p = z: p stays z),
// Ap = A p, pAp.  z (held in the r array) and the old p of a neighbour give
// its new p on the fly.
__global__ void __launch_bounds__(PBLOCK) pcg_spmv_kernel(hp_torso t, const float* __restrict__ z,
                                                          float* __restrict__ p,
                                                          float* __restrict__ Ap, double* part,
                                                          const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const float beta = (float)sc[S_BETA];
  double v[1] = {0.0};
  const uint32_t vb = (uint32_t)t.n_cells * 4u;
  const buf_t Z = buf_rsrc(z, vb), P = buf_rsrc(p, vb), AP = buf_rsrc(Ap, vb);
  // on the first iteration beta = 0 and p = z, so the direction is z (exactly)
  for_cells_u<2>(
      t, [&](uint32_t c, int k, const Chunk& ch) { return dir_loads(t, Z, P, c, k, ch); },
      [&](uint32_t c, int, const Chunk&, const DirLoads& l) {
        float pc;
        const float acc = dir_product(t, c, beta, l, pc);
        buf_st(AP, c, acc);
        v[0] = c != T_OOB ? v[0] + (double)pc * (double)acc : v[0];
      });
  // p itself is not written here (other blocks still read the old p of their
  // neighbours); pcg_update materialises the same z + beta*p next.
  block_partials<1>(v, part);
}

// pAp from the partials of the product kernel; alpha = rz / pAp (a vanishing
// pAp ends the solve: done = 2).  Shared by both PCG forms.
__global__ void __launch_bounds__(PBLOCK) pap_finish_kernel(const double* __restrict__ part, int nb,
                                                            double* sc) {
  if (sc[S_DONE] != 0.0) return;
  double red[1];
  sum_partials<1>(part, nb, red);
  if (threadIdx.x == 0) {
    sc[S_PAP] = red[0];
    if (!(fabs(red[0]) > 1e-300)) sc[S_DONE] = 2.0;
    else sc[S_ALPHA] = sc[S_RZ] / red[0];
  }
}

void launch_pap_finish(double* part, int nb, double* sc, hipStream_t st) {
  hipLaunchKernelGGL(pap_finish_kernel, dim3(1), dim3(PBLOCK), 0, st, part, nb, sc);
}

// p <- z + beta p (materialised), d += alpha p, z -= alpha Ap/diag (z = M^-1 r
// held in the r array), rz, rr with r = diag z; finish: beta, stopping test,
// iteration count.
// Extreme values of |z| for the 'linf' and 'p99' stopping metrics
// (compute_bspm.py:222-237): per block, the largest |z| below tolz, the
// smallest |z| at or above it, and the count at or above it (linf uses the
// largest |z| overall, kept in the first slot with tolz = +inf).
__device__ __forceinline__ void zstat_partials(double below_max, double above_min, double count,
                                               double* part) {
  __shared__ double sh[3][PWAVES];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    below_max = fmax(below_max, __shfl_xor(below_max, o, 64));
    above_min = fmin(above_min, __shfl_xor(above_min, o, 64));
    count += __shfl_xor(count, o, 64);
  }
  if (lane == 0) {
    sh[0][wave] = below_max;
    sh[1][wave] = above_min;
    sh[2][wave] = count;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < PWAVES; ++w) {
      below_max = fmax(below_max, sh[0][w]);
      above_min = fmin(above_min, sh[1][w]);
      count += sh[2][w];
    }
    part[3 * gridDim.x + blockIdx.x] = below_max;
    part[4 * gridDim.x + blockIdx.x] = above_min;
    part[5 * gridDim.x + blockIdx.x] = count;
  }
}

__global__ void __launch_bounds__(PBLOCK) pcg_update_kernel(hp_torso t, float* __restrict__ z,
                                                            float* __restrict__ d,
                                                            float* __restrict__ p,
                                                            const float* __restrict__ Ap,
                                                            double* part, const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const float alpha = (float)sc[S_ALPHA];
  const float beta = (float)sc[S_BETA];
  const int mode = (int)sc[S_MODE];
  const double tolz = mode == 3 ? sqrt(sc[S_TOL2]) : HUGE_VAL;
  double bmax = 0.0, amin = HUGE_VAL, cnt = 0.0;
  double v[3] = {0.0, 0.0, 0.0};
  const uint32_t vb = (uint32_t)t.n_cells * 4u;
  const buf_t Z = buf_rsrc(z, vb), D = buf_rsrc(d, vb), P = buf_rsrc(p, vb),
              AP = buf_rsrc(Ap, vb), DG = buf_rsrc(t.diag, vb);
  struct L {
    float z, p, d, ap, dg;
  };
  for_cells_u<4>(
      t,
      [&](uint32_t c, int, const Chunk&) {
        return L{buf_ld<float>(Z, c), buf_ld<float>(P, c), buf_ld<float>(D, c),
                 buf_ld<float>(AP, c), buf_ld<float>(DG, c)};
      },
      [&](uint32_t c, int, const Chunk&, const L& l) {
        const float pc = l.z + beta * l.p;
        buf_st(P, c, pc);
        buf_st(D, c, l.d + alpha * pc);
        const float zn = l.z - alpha * (l.ap / l.dg);
        buf_st(Z, c, zn);
        const bool in = c != T_OOB;
        const double rn = (double)l.dg * (double)zn;
        v[0] = in ? v[0] + rn * (double)zn : v[0];
        v[1] = in ? v[1] + rn * rn : v[1];
        v[2] = in ? v[2] + (double)zn * (double)zn : v[2];
        if (mode >= 2 && in) {
          const double a = fabs((double)zn);
          if (a < tolz) bmax = fmax(bmax, a);
          else {
            amin = fmin(amin, a);
            cnt += 1.0;
          }
        }
      });
  block_partials<3>(v, part);
  if (mode >= 2) zstat_partials(bmax, amin, cnt, part);
}

// rz, ||r||^2, ||z||^2 from the update's partials: beta, stopping test,
// iteration count
__global__ void __launch_bounds__(PBLOCK) pcg_update_finish_kernel(const double* __restrict__ part,
                                                                   int nb, double* sc) {
  if (sc[S_DONE] != 0.0) return;
  double red[3];
  sum_partials<3>(part, nb, red);
  const int mode = (int)sc[S_MODE];
  double bmax = 0.0, amin = HUGE_VAL, cnt = 0.0;
  if (mode >= 2) {  // the |z| extremes over the blocks
    __shared__ double sh[3][PBLOCK / 64];
    for (int b = threadIdx.x; b < nb; b += PBLOCK) {
      bmax = fmax(bmax, part[3 * nb + b]);
      amin = fmin(amin, part[4 * nb + b]);
      cnt += part[5 * nb + b];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      bmax = fmax(bmax, __shfl_xor(bmax, o, 64));
      amin = fmin(amin, __shfl_xor(amin, o, 64));
      cnt += __shfl_xor(cnt, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
      sh[0][threadIdx.x >> 6] = bmax;
      sh[1][threadIdx.x >> 6] = amin;
      sh[2][threadIdx.x >> 6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 1; w < PBLOCK / 64; ++w) {
        bmax = fmax(bmax, sh[0][w]);
        amin = fmin(amin, sh[1][w]);
        cnt += sh[2][w];
      }
  }
  if (threadIdx.x == 0) {
    sc[S_BETA] = red[0] / sc[S_RZ];
    sc[S_RZ] = red[0];
    sc[S_RR] = red[1];
    sc[S_ZZ] = red[2];
    sc[S_ITER] += 1.0;
    const double tol2 = sc[S_TOL2], n = sc[S_NCELL];
    bool conv;
    if (mode == 0) {  // relative residual
      conv = red[1] <= tol2 * sc[S_RR0];
    } else if (mode == 1) {  // rms(z) < tol (compute_bspm.py:222-223)
      conv = red[2] < tol2 * n;
      sc[S_ZSTAT] = sqrt(red[2] / n);
    } else if (mode == 2) {  // max |z| < tol
      sc[S_ZSTAT] = bmax;
      conv = bmax < sqrt(tol2);
    } else {  // 99th percentile of |z| (linear interpolation) < tol
      // sorted |z| of the n rows: the cnt largest are >= tol; position
      // q = 0.99 (n-1) = k + frac lies below them, straddles them, or not
      const double tol = sqrt(tol2), q = 0.99 * (n - 1.0), k = floor(q), frac = q - k;
      const double first_above = n - cnt;  // index of the smallest value >= tol
      if (k + 1.0 < first_above) conv = true;  // both neighbours of q below tol
      else if (k >= first_above) conv = false;
      else conv = bmax + frac * (amin - bmax) < tol;  // s[k] < tol <= s[k+1]
    }
    if (conv || sc[S_ITER] >= sc[S_MAXIT]) sc[S_DONE] = 1.0;
  }
}

// x64 += scale * d
__global__ void axpy_f64_kernel(hp_torso t, double* __restrict__ x, const float* __restrict__ d,
                                double scale) {
  for_cells(t, [&](int c, int, const Chunk&) { x[c] += scale * (double)d[c]; });
}

// r32 = r64 * scale
__global__ void cast_scale_kernel(hp_torso t, const double* __restrict__ r64, float* __restrict__ r32,
                                  double scale) {
  for_cells(t, [&](int c, int, const Chunk&) { r32[c] = (float)(r64[c] * scale); });
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
                                                            int64_t n, double* part) {
  double v[1] = {0.0};
  for (int64_t q = (int64_t)blockIdx.x * PBLOCK + threadIdx.x; q < n; q += (int64_t)gridDim.x * PBLOCK)
    v[0] += lap[slots[q]];
  block_partials<1>(v, part);
}

// b[row] = coeff*(sum_{fine in row} lap - cnt*mean)/n_expected; b[pin] = 0
__global__ void source_gather_kernel(const double* __restrict__ lap, const int32_t* __restrict__ rows,
                                     const int32_t* __restrict__ ptr, const int32_t* __restrict__ fine,
                                     int64_t n_rows_hit, const double* __restrict__ sum_n,
                                     double inv_nvalid, double coeff, double inv_nexp,
                                     const double* __restrict__ psi, double* __restrict__ b,
                                     int64_t pin) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_rows_hit) return;
  const double mean = sum_n[0] * inv_nvalid;
  double s = 0.0;
  for (int e = ptr[q]; e < ptr[q + 1]; ++e) s += lap[fine[e]] - mean;
  const int row = rows[q];
  double v = coeff * s * inv_nexp;
  if (psi) v *= psi[row];  // smoothed boundary: the row's psi scales its source
  b[row] = row == pin ? 0.0 : v;
}

// Lead-field ECG: out[f][e] = sum_r W[r][e] * V[f][slot[r]] over the rows r
// (heart voxels), FP32 products accumulated in FP64.  Pass 1: block q sums
// its row range into part[q][f][e]; pass 2 adds the block partials in block
// order (deterministic).
constexpr int LF_MAXF = 16;  // frames per call (the M = 16 of the MFMA tile)
// FP64 MFMA form: v_mfma_f64_16x16x4f64 over 4 rows per instruction (A = V^T:
// 16 frames x 4 rows, B = W: 4 rows x 16 electrodes, C = the 16 x 16 FP64
// partial sums in 4 registers per lane).
// Operand layout of the instruction: lane l holds A[l%16][l/16] and
// B[l/16][l%16]; C[4*i + l/16][l%16] in its i-th register (the FP64 form; the
// FP32 16x16x4 form keeps rows 4*(l/16)+i instead).  Eight K-steps
// of loads are issued before their MFMAs.  Frames past n_frames, electrodes
// past ne and rows past the block's range enter as 0.
constexpr int LF_MAXW = 16, LF_UNROLL = 8;
// One wave per 32 electrodes (two 16-electrode MFMA tiles sharing the rows'
// slot and V loads); one workgroup of up to LF_MAXW waves covers every
// electrode of its row range, so its waves read the same rows of W at about
// the same time and each 128-byte line of W is fetched once.
__global__ void __launch_bounds__(LF_MAXW * 64) leadfield_partial_kernel(
    const float* __restrict__ W, const int32_t* __restrict__ row_slot, int64_t n_rows, int ne,
    const float* __restrict__ V, int64_t ldv, int n_frames, int64_t rows_per_block,
    double* __restrict__ part) {
  typedef double d4 __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int e0 = (blockIdx.y * (blockDim.x >> 6) + wave) * 32;
  if (e0 >= ne) return;  // whole wave
  const int mi = lane & 15, kq = lane >> 4;  // A row (frame) / B column (electrode); K index
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < n_rows ? r0 + rows_per_block : n_rows;
  const bool frame_ok = mi < n_frames, ok_a = e0 + mi < ne, ok_b = e0 + 16 + mi < ne;
  // every load reads an in-range element (clamped row, frame, electrode) and
  // is zeroed afterwards, so the loads issue without branches
  const int mc = frame_ok ? mi : 0, ea = ok_a ? e0 + mi : ne - 1, eb = ok_b ? e0 + 16 + mi : ne - 1;
  d4 acc_a = {0.0, 0.0, 0.0, 0.0}, acc_b = {0.0, 0.0, 0.0, 0.0};
  if (r0 >= r1) return;
  for (int64_t r = r0; r < r1; r += 4 * LF_UNROLL) {
    int32_t sl[LF_UNROLL];
    float wa[LF_UNROLL], wb[LF_UNROLL];
#pragma unroll
    for (int u = 0; u < LF_UNROLL; ++u) {
      const int64_t row = r + 4 * u + kq, rc = row < r1 ? row : r1 - 1;
      sl[u] = row_slot[rc];
      wa[u] = W[rc * ne + ea];
      wb[u] = W[rc * ne + eb];
      if (!(row < r1 && ok_a)) wa[u] = 0.f;
      if (!(row < r1 && ok_b)) wb[u] = 0.f;
    }
    float av[LF_UNROLL];
#pragma unroll
    for (int u = 0; u < LF_UNROLL; ++u) {
      av[u] = V[(int64_t)mc * ldv + sl[u]];
      if (!(r + 4 * u + kq < r1 && frame_ok)) av[u] = 0.f;
    }
#pragma unroll
    for (int u = 0; u < LF_UNROLL; ++u) {
      acc_a = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[u], (double)wa[u], acc_a, 0, 0, 0);
      acc_b = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[u], (double)wb[u], acc_b, 0, 0, 0);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int f = 4 * i + kq;
    if (f >= n_frames) continue;
    double* dst = part + ((int64_t)blockIdx.x * n_frames + f) * ne;
    if (ok_a) dst[e0 + mi] = acc_a[i];
    if (ok_b) dst[e0 + 16 + mi] = acc_b[i];
  }
}

// out[q] (q = f*ne + e) = sum over the row blocks of part: one wave per
// output, lane l adding blocks l, l+64, ... in order (eight loads in flight),
// then a fixed shuffle tree: a fixed summation order.
__global__ void __launch_bounds__(256) leadfield_reduce_kernel(const double* __restrict__ part,
                                                               int n_blocks, int ne, int n_frames,
                                                               double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= ne * n_frames) return;  // whole wave
  const int64_t stride = (int64_t)ne * n_frames;
  double s = 0.0;
  int b = lane;
  for (; b + 7 * 64 < n_blocks; b += 8 * 64) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = part[(int64_t)(b + 64 * u) * stride + q];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += x[u];
  }
  for (; b < n_blocks; b += 64) s += part[(int64_t)b * stride + q];
  s = wave_sum(s);
  if (lane == 0) out[q] = s;
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

int hp_torso_chunks(int32_t ni, int32_t nj, int32_t nk, const int32_t* col_k0,
                    const int32_t* col_len, const int32_t* col_off, int32_t* chunks,
                    int64_t capacity, int64_t* n_chunks) {
  if (ni <= 0 || nj <= 0 || nk <= 0 || nk > 32767 || !col_k0 || !col_len || !col_off ||
      !n_chunks || (chunks && capacity < 0)) {
    set_error("hp_torso_chunks: bad arguments (nk <= 32767)");
    return -1;
  }
  // column (i, j): base = off - k0, span = k0 | (k0 + len) << 16
  auto col = [&](int i, int j, int32_t& base, int32_t& span) {
    base = 0;
    span = 0;
    if (i < 0 || i >= ni || j < 0 || j >= nj) return;
    const int64_t c = (int64_t)i * nj + j;
    if (col_len[c] <= 0) return;
    base = col_off[c] - col_k0[c];
    span = col_k0[c] | ((col_k0[c] + col_len[c]) << 16);
  };
  int64_t n = 0;
  for (int i = 0; i < ni; ++i)
    for (int j = 0; j < nj; ++j) {
      const int64_t c = (int64_t)i * nj + j;
      const int k0 = col_k0[c], len = col_len[c];
      if (len <= 0) continue;
      if (k0 < 0 || k0 + len > nk) {
        set_error("hp_torso_chunks: column (%d, %d) run [%d, %d) outside [0, %d)", i, j, k0,
                  k0 + len, nk);
        return -1;
      }
      for (int dk = 0; dk < len; dk += 64, ++n) {
        if (!chunks) continue;
        if (n >= capacity) {
          set_error("hp_torso_chunks: capacity %lld too small", (long long)capacity);
          return -1;
        }
        int32_t* r = chunks + n * HP_TORSO_CHUNK_INTS;
        r[0] = k0 + dk;
        r[1] = len - dk < 64 ? len - dk : 64;
        r[2] = i;
        r[3] = j;
        col(i, j, r[4], r[9]);
        col(i + 1, j, r[5], r[10]);
        col(i - 1, j, r[6], r[11]);
        col(i, j + 1, r[7], r[12]);
        col(i, j - 1, r[8], r[13]);
        r[14] = 0;
        r[15] = 0;
      }
    }
  *n_chunks = n;
  return 0;
}

int hp_poisson_spmv_f32(const hp_torso* t, const float* x, float* y, void* stream) {
  if (!torso_ok(t, "hp_poisson_spmv_f32")) return -1;
  hipLaunchKernelGGL(spmv_f32_kernel, dim3(grid_blocks(t)), dim3(PBLOCK), 0, (hipStream_t)stream,
                     *t, x, y);
  return check_launch("hp_poisson_spmv_f32");
}

int hp_poisson_residual_f64(const hp_torso* t, const double* b, const double* x, double* r,
                            hp_pcg_ws* ws, void* stream) {
  if (!torso_ok(t, "hp_poisson_residual_f64") ||
      !ws_ok(ws, 0, false, "hp_poisson_residual_f64")) return -1;
  const int nb = grid_blocks(t);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(residual_f64_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, b, x, r, ws->part);
  hipLaunchKernelGGL(finish_sums_kernel<2>, dim3(1), dim3(PBLOCK), 0, st, ws->part, nb, ws->out);
  return check_launch("hp_poisson_residual_f64");
}

int hp_poisson_residual_f64_noalloc(const hp_torso* t, const double* b, const double* x,
                                    double* r, double* out, void* stream) {
  if (!torso_ok(t, "hp_poisson_residual_f64_noalloc")) return -1;
  if (!b || !x || !r || !out) {
    set_error("hp_poisson_residual_f64_noalloc: b, x, r and out are required");
    return -1;
  }
  const int nb = grid_blocks(t);
  if (nb > PMAX_BLOCKS) {  // grid_blocks caps the grid; the scratch holds 2 partials per block
    set_error("hp_poisson_residual_f64_noalloc: %d blocks exceed the scratch of %d", nb,
              PMAX_BLOCKS);
    return -1;
  }
  double* part = scratch_address("hp_poisson_residual_f64_noalloc");
  if (!part) return -1;
  // every block writes both of its partials before the finish kernel reads
  // them, so nothing left in the scratch by an earlier call is read
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(residual_f64_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, b, x, r, part);
  hipLaunchKernelGGL(finish_sums_kernel<2>, dim3(1), dim3(PBLOCK), 0, st, part, nb, out);
  return check_launch("hp_poisson_residual_f64_noalloc");
}

static int pcg_begin(const hp_torso* t, float* r, float* d, float* p, hp_pcg_ws* ws, double tol,
                     int max_iter, int mode, int64_t n_rows, void* stream);

int hp_pcg_begin_f32(const hp_torso* t, float* r, float* d, float* p, hp_pcg_ws* ws,
                     double tol, int max_iter, int mode, void* stream) {
  if (mode < 0 || mode > 1) {
    set_error("hp_pcg_begin_f32: mode 0 or 1 (hp_pcg_begin_metric_f32 has the others)");
    return -1;
  }
  return pcg_begin(t, r, d, p, ws, tol, max_iter, mode, t ? t->n_cells : 0, stream);
}

int hp_pcg_begin_metric_f32(const hp_torso* t, float* r, float* d, float* p, hp_pcg_ws* ws,
                            double tol, int max_iter, int mode, int64_t n_rows, void* stream) {
  if (mode < 0 || mode > 3 || n_rows < 1 || (t && n_rows > t->n_cells)) {
    set_error("hp_pcg_begin_metric_f32: mode 0..3 and 1 <= n_rows <= n_cells required");
    return -1;
  }
  return pcg_begin(t, r, d, p, ws, tol, max_iter, mode, n_rows, stream);
}

static int pcg_begin(const hp_torso* t, float* r, float* d, float* p, hp_pcg_ws* ws, double tol,
                     int max_iter, int mode, int64_t n_rows, void* stream) {
  if (!torso_ok(t, "hp_pcg_begin_f32") ||
      !ws_ok(ws, 0, true, "hp_pcg_begin_f32")) return -1;
  const int nb = grid_blocks(t);
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(pcg_init_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, r, d, p, ws->part);
  hipLaunchKernelGGL(pcg_init_finish_kernel, dim3(1), dim3(PBLOCK), 0, st, ws->part, nb,
                     ws->scalars, tol, max_iter, mode, n_rows);
  return check_launch("hp_pcg_begin_f32");
}

int hp_pcg_iterate_f32(const hp_torso* t, float* r, float* d, float* p, float* Ap, hp_pcg_ws* ws,
                       int n_iter, void* stream) {
  if (!torso_ok(t, "hp_pcg_iterate_f32") ||
      !ws_ok(ws, n_iter, true, "hp_pcg_iterate_f32")) return -1;
  const int nb = grid_blocks(t);
  hipStream_t st = (hipStream_t)stream;
  for (int it = 0; it < n_iter; ++it) {
    hipLaunchKernelGGL(pcg_spmv_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, r, p, Ap, ws->part,
                       ws->scalars);
    launch_pap_finish(ws->part, nb, ws->scalars, st);
    hipLaunchKernelGGL(pcg_update_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, r, d, p, Ap,
                       ws->part, ws->scalars);
    hipLaunchKernelGGL(pcg_update_finish_kernel, dim3(1), dim3(PBLOCK), 0, st, ws->part, nb,
                       ws->scalars);
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
  // row blocks: as many as the partial buffer holds, up to 4096
  int64_t nbl = part_len / ((int64_t)n_frames * ne);
  int nb = (int)(nbl > 4096 ? 4096 : nbl);
  if (nb < 1) {
    set_error("hp_leadfield_apply_f32: partial buffer needs %lld doubles", (long long)n_frames * ne);
    return -1;
  }
  const int64_t rpb = (n_rows + nb - 1) / nb;
  nb = (int)((n_rows + rpb - 1) / rpb);
  if ((int64_t)nb * n_frames * ne > part_len) {
    set_error("hp_leadfield_apply_f32: partial buffer needs %lld doubles",
              (long long)nb * n_frames * ne);
    return -1;
  }
  hipStream_t st = (hipStream_t)stream;
  // one wave per 32 electrodes; one workgroup covers all of them (more than
  // LF_MAXW * 32 electrodes spill into further workgroups along y)
  const int tiles = (ne + 31) / 32, waves = tiles < LF_MAXW ? tiles : LF_MAXW;
  hipLaunchKernelGGL(leadfield_partial_kernel, dim3(nb, (tiles + waves - 1) / waves),
                     dim3(waves * 64), 0, st, W, row_slot, n_rows, ne, V, ldv, n_frames, rpb,
                     part);
  const int rc = check_launch("hp_leadfield_apply_f32/partial");
  if (rc) return rc;
  hipLaunchKernelGGL(leadfield_reduce_kernel, dim3(ceil_div((int64_t)ne * n_frames, 4)),
                     dim3(256), 0, st, part, nb, ne, n_frames, out);
  return check_launch("hp_leadfield_apply_f32/reduce");
}

int hp_bspm_source_f64(const double* lap, const int32_t* valid_slots, int64_t n_valid,
                       const int32_t* rows, const int32_t* ptr, const int32_t* fine,
                       int64_t n_rows_hit, double coeff, double n_expected, const double* psi,
                       double* b, int64_t pin, hp_pcg_ws* ws, void* stream) {
  if (!ws || !ws->part || !ws->out || n_valid <= 0 || n_expected <= 0.0) {
    set_error("hp_bspm_source_f64: bad arguments");
    return -1;
  }
  hipStream_t st = (hipStream_t)stream;
  int nb = (int)((n_valid + PBLOCK - 1) / PBLOCK);
  if (nb > 1024) nb = 1024;
  hipLaunchKernelGGL(masked_sum_kernel, dim3(nb), dim3(PBLOCK), 0, st, lap, valid_slots, n_valid,
                     ws->part);
  hipLaunchKernelGGL(finish_sums_kernel<1>, dim3(1), dim3(PBLOCK), 0, st, ws->part, nb, ws->out);
  int rc = check_launch("hp_bspm_source_f64/sum");
  if (rc) return rc;
  if (n_rows_hit > 0)
    hipLaunchKernelGGL(source_gather_kernel, dim3(ceil_div(n_rows_hit, 256)), dim3(256), 0, st,
                       lap, rows, ptr, fine, n_rows_hit, ws->out, 1.0 / (double)n_valid, coeff,
                       1.0 / n_expected, psi, b, pin);
  return check_launch("hp_bspm_source_f64/gather");
}

int hp_leads_accumulate_f64(const double* phi, const int32_t* ptr, const int32_t* idx,
                            const double* w, double* out, int64_t n_electrodes, void* stream) {
  if (n_electrodes <= 0) return 0;
  if (!phi || !ptr || !idx || !w || !out) {
    set_error("hp_leads_accumulate_f64: bad arguments (null operand)");
    return -1;
  }
  hipLaunchKernelGGL(leads_kernel, dim3(ceil_div(n_electrodes, 256)), dim3(256), 0,
                     (hipStream_t)stream, phi, ptr, idx, w, out, n_electrodes);
  return check_launch("hp_leads_accumulate_f64");
}

}  // extern "C"
