// Geometric multigrid preconditioner for the torso operator, and the PCG that
// uses it.
//
// The reference preconditions its torso CG with the diagonal
// (_pcg_gpu, hrdayapy/ecg/functions/compute_bspm.py:164-263); on the
// 384x384x256 torso that needs thousands of iterations per solve.  Here the
// preconditioner is one symmetric V-cycle:
//   level 0      the column-run torso grid itself (poisson_common.h);
//   levels 1..L  dense boxes from 2x2x2 cell agglomeration.  The face
//                conductance between two coarse cells is half the sum of the
//                four fine faces it covers (the operator rediscretised at 2h;
//                the Galerkin product with piecewise-constant transfer would
//                be twice as stiff).  Inactive cells and the cell holding the
//                pinned reference have inv_diag = 0 and stay at zero.
//   smoothing    nu damped-Jacobi sweeps before and nu after each coarse
//                correction (nu even, so the sweeps ping-pong back into place);
//   transfer     restriction = sum of the children's residuals, prolongation =
//                add the parent's correction (restriction = prolongation^T);
//   coarsest     exact: a dense inverse computed on the host.
// With a symmetric smoother, R = P^T and SPD coarse operators the V-cycle is a
// symmetric positive definite preconditioner, so plain PCG applies.
#include "poisson_common.h"

namespace hp {

__device__ __forceinline__ bool mg_done(const double* sc) { return sc && sc[S_DONE] != 0.0; }

__device__ __forceinline__ void dense_ijk(const hp_mg_level& L, int64_t c, int& i, int& j, int& k) {
  k = (int)(c % L.nk);
  const int64_t q = c / L.nk;
  j = (int)(q % L.nj);
  i = (int)(q / L.nj);
}

// sum over the faces of g_f (x_c - x_nb); faces to inactive cells have g = 0,
// and the grounded cell holds x = 0
__device__ __forceinline__ float dense_Ax(const hp_mg_level& L, const float* __restrict__ x,
                                          int64_t c, int i, int j, int k) {
  const int64_t sx = (int64_t)L.nj * L.nk, sy = L.nk;
  const float xc = x[c];
  float acc = 0.f;
  if (i + 1 < L.ni) acc += L.gx[c] * (xc - x[c + sx]);
  if (i > 0) acc += L.gx[c - sx] * (xc - x[c - sx]);
  if (j + 1 < L.nj) acc += L.gy[c] * (xc - x[c + sy]);
  if (j > 0) acc += L.gy[c - sy] * (xc - x[c - sy]);
  if (k + 1 < L.nk) acc += L.gz[c] * (xc - x[c + 1]);
  if (k > 0) acc += L.gz[c - 1] * (xc - x[c - 1]);
  return acc;
}

#define GRID_STRIDE(c, n) \
  for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < (n); c += (int64_t)gridDim.x * blockDim.x)

// x_out = x_in + omega D^-1 (b - A x_in)   (zero = 1: x_in = 0)
__global__ void __launch_bounds__(256) mg_smooth_dense_kernel(hp_mg_level L,
                                                             const float* __restrict__ xin,
                                                             float* __restrict__ xout,
                                                             const float* __restrict__ b,
                                                             float omega, int zero,
                                                             const double* sc) {
  if (mg_done(sc)) return;
  GRID_STRIDE(c, L.n) {
    const float inv = L.inv_diag[c];
    float v = 0.f;
    if (inv != 0.f) {
      if (zero) {
        v = omega * inv * b[c];
      } else {
        int i, j, k;
        dense_ijk(L, c, i, j, k);
        v = xin[c] + omega * inv * (b[c] - dense_Ax(L, xin, c, i, j, k));
      }
    }
    xout[c] = v;
  }
}

// b_c[C] = sum over the children of C of (b_f - A_f x_f)
__global__ void __launch_bounds__(256) mg_restrict_dense_kernel(hp_mg_level F, hp_mg_level Cl,
                                                               const double* sc) {
  if (mg_done(sc)) return;
  GRID_STRIDE(C, Cl.n) {
    float s = 0.f;
    if (Cl.inv_diag[C] != 0.f) {
      int I, J, K;
      dense_ijk(Cl, C, I, J, K);
#pragma unroll
      for (int d = 0; d < 8; ++d) {
        const int i = 2 * I + (d >> 2), j = 2 * J + ((d >> 1) & 1), k = 2 * K + (d & 1);
        if (i >= F.ni || j >= F.nj || k >= F.nk) continue;
        const int64_t c = ((int64_t)i * F.nj + j) * F.nk + k;
        if (F.inv_diag[c] != 0.f) s += F.b[c] - dense_Ax(F, F.x, c, i, j, k);
      }
    }
    Cl.b[C] = s;
  }
}

// x_f += x_c[parent]
__global__ void __launch_bounds__(256) mg_prolong_dense_kernel(hp_mg_level F, hp_mg_level Cl,
                                                              const double* sc) {
  if (mg_done(sc)) return;
  GRID_STRIDE(c, F.n) {
    if (F.inv_diag[c] != 0.f) {
      int i, j, k;
      dense_ijk(F, c, i, j, k);
      F.x[c] += Cl.x[((int64_t)(i >> 1) * Cl.nj + (j >> 1)) * Cl.nk + (k >> 1)];
    }
  }
}

// coarsest level: x = A^-1 b with the dense inverse (rows of inactive cells are 0)
__global__ void __launch_bounds__(256) mg_coarse_kernel(hp_mg_level L, const float* __restrict__ inv,
                                                       const double* sc) {
  if (mg_done(sc)) return;
  __shared__ float bs[1024];
  const int n = (int)L.n;
  for (int q = threadIdx.x; q < n; q += blockDim.x) bs[q] = L.b[q];
  __syncthreads();
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    float s = 0.f;
    const float* row = inv + (int64_t)c * n;
    for (int q = 0; q < n; ++q) s += row[q] * bs[q];
    L.x[c] = s;
  }
}

// level 0 (column-run): damped Jacobi with the stored diagonal
__global__ void __launch_bounds__(PBLOCK) mg_smooth0_kernel(hp_torso t, const float* __restrict__ inv0,
                                                           const float* __restrict__ xin,
                                                           float* __restrict__ xout,
                                                           const float* __restrict__ b, float omega,
                                                           int zero, const double* sc) {
  if (mg_done(sc)) return;
  const uint32_t vb = (uint32_t)t.n_cells * 4u;
  // with zero = 1 there is no input iterate (xin may be null): an empty
  // resource makes its loads read 0 without touching memory
  const buf_t X = buf_rsrc(xin, zero ? 0u : vb), XO = buf_rsrc(xout, vb), B = buf_rsrc(b, vb),
              INV = buf_rsrc(inv0, vb);
  struct L {
    float g[6], x[6], xc, b, inv;
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
        l.b = buf_ld<float>(B, c);
        l.inv = buf_ld<float>(INV, c);
        return l;
      },
      [&](uint32_t c, int, const Chunk&, const L& l) {
        const float v = zero ? omega * l.inv * l.b
                             : l.xc + omega * l.inv * (l.b - row_sum(t, c, l.xc, l.g, l.x));
        buf_st(XO, c, l.inv != 0.f ? v : 0.f);
      });
}

// level 0 -> level 1: b1[C] = sum over the stored children of (r - A x)
__global__ void __launch_bounds__(256) mg_restrict0_kernel(hp_torso t, const float* __restrict__ inv0,
                                                          const float* __restrict__ x,
                                                          const float* __restrict__ r,
                                                          hp_mg_level Cl, const double* sc) {
  if (mg_done(sc)) return;
  GRID_STRIDE(C, Cl.n) {
    float s = 0.f;
    if (Cl.inv_diag[C] != 0.f) {
      int I, J, K;
      dense_ijk(Cl, C, I, J, K);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = 2 * I + (q >> 1), j = 2 * J + (q & 1);
        if (i >= t.ni || j >= t.nj) continue;
        const ColNbr m = col_nbr(t, i, j);
#pragma unroll
        for (int dk = 0; dk < 2; ++dk) {
          const int k = 2 * K + dk;
          const int c = cell_at(m.c, k);
          if (c >= 0 && inv0[c] != 0.f) s += r[c] - apply_row_f32(t, x, c, k, ColLookup{m});
        }
      }
    }
    Cl.b[C] = s;
  }
}

// level 1 -> level 0: z += x1[parent]
__global__ void __launch_bounds__(PBLOCK) mg_prolong0_kernel(hp_torso t, const float* __restrict__ inv0,
                                                            float* __restrict__ z, hp_mg_level Cl,
                                                            const double* sc) {
  if (mg_done(sc)) return;
  for_cells(t, [&](int c, int k, const Chunk& ch) {
    if (inv0[c] != 0.f) z[c] += Cl.x[((int64_t)(ch.i >> 1) * Cl.nj + (ch.j >> 1)) * Cl.nk + (k >> 1)];
  });
}

static int dense_blocks(int64_t n) {
  int64_t nb = (n + 255) / 256;
  if (nb > 8192) nb = 8192;
  return (int)(nb < 1 ? 1 : nb);
}

// One V-cycle z = B r (z and every level's x overwritten).
static void vcycle(const hp_torso* t, const hp_mg* mg, const float* r, float* z, const double* sc,
                   hipStream_t st) {
  const int nb0 = grid_blocks(t);
  const int nu = mg->nu;
  const float w = mg->omega;
  const int L = mg->n_levels;
  // level 0 pre-smoothing from zero: sweeps alternate tmp0 / z, ending in z
  for (int s = 0; s < nu; ++s) {
    float* out = ((nu - s) & 1) ? z : mg->tmp0;
    const float* in = s == 0 ? nullptr : (((nu - s) & 1) ? mg->tmp0 : z);
    hipLaunchKernelGGL(mg_smooth0_kernel, dim3(nb0), dim3(PBLOCK), 0, st, *t, mg->inv_diag0, in,
                       out, r, w, s == 0 ? 1 : 0, sc);
  }
  hipLaunchKernelGGL(mg_restrict0_kernel, dim3(dense_blocks(mg->level[0].n)), dim3(256), 0, st, *t,
                     mg->inv_diag0, z, r, mg->level[0], sc);
  // down the dense levels
  for (int l = 0; l < L - 1; ++l) {
    const hp_mg_level& Lv = mg->level[l];
    const int nb = dense_blocks(Lv.n);
    for (int s = 0; s < nu; ++s) {
      float* out = ((nu - s) & 1) ? Lv.x : Lv.tmp;
      const float* in = s == 0 ? nullptr : (((nu - s) & 1) ? Lv.tmp : Lv.x);
      hipLaunchKernelGGL(mg_smooth_dense_kernel, dim3(nb), dim3(256), 0, st, Lv, in, out, Lv.b, w,
                         s == 0 ? 1 : 0, sc);
    }
    hipLaunchKernelGGL(mg_restrict_dense_kernel, dim3(dense_blocks(mg->level[l + 1].n)), dim3(256),
                       0, st, Lv, mg->level[l + 1], sc);
  }
  hipLaunchKernelGGL(mg_coarse_kernel, dim3(1), dim3(256), 0, st, mg->level[L - 1], mg->coarse_inv,
                     sc);
  // back up
  for (int l = L - 2; l >= 0; --l) {
    const hp_mg_level& Lv = mg->level[l];
    const int nb = dense_blocks(Lv.n);
    hipLaunchKernelGGL(mg_prolong_dense_kernel, dim3(nb), dim3(256), 0, st, Lv, mg->level[l + 1], sc);
    for (int s = 0; s < nu; ++s) {
      const bool to_x = ((nu - s) & 1) != 0;
      hipLaunchKernelGGL(mg_smooth_dense_kernel, dim3(nb), dim3(256), 0, st, Lv,
                         to_x ? Lv.tmp : Lv.x, to_x ? Lv.x : Lv.tmp, Lv.b, w, 0, sc);
    }
  }
  hipLaunchKernelGGL(mg_prolong0_kernel, dim3(nb0), dim3(PBLOCK), 0, st, *t, mg->inv_diag0, z,
                     mg->level[0], sc);
  for (int s = 0; s < nu; ++s) {
    const bool to_z = ((nu - s) & 1) != 0;
    hipLaunchKernelGGL(mg_smooth0_kernel, dim3(nb0), dim3(PBLOCK), 0, st, *t, mg->inv_diag0,
                       to_z ? mg->tmp0 : z, to_z ? z : mg->tmp0, r, w, 0, sc);
  }
}

// ---------------------------------------------------------------------------
// PCG with the V-cycle (FP32 vectors, FP64 inner products, device-side state
// in the scalar slots of poisson_common.h)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(PBLOCK) mgpcg_init_kernel(hp_torso t, const float* __restrict__ r,
                                                           const float* __restrict__ z,
                                                           float* __restrict__ d,
                                                           float* __restrict__ p, double* part) {
  double v[2] = {0.0, 0.0};
  for_cells(t, [&](int c, int, const Chunk&) {
    const float rc = r[c], zc = z[c];
    d[c] = 0.f;
    p[c] = zc;
    v[0] += (double)rc * (double)zc;
    v[1] += (double)rc * (double)rc;
  });
  block_partials<2>(v, part);
}

__global__ void __launch_bounds__(PBLOCK) mgpcg_init_finish_kernel(const double* __restrict__ part,
                                                                   int nb, double* sc, double tol,
                                                                   int maxit, int64_t n_cells) {
  double red[2];
  sum_partials<2>(part, nb, red);
  if (threadIdx.x == 0) {
    sc[S_RZ] = red[0];
    sc[S_RR0] = red[1];
    sc[S_RR] = red[1];
    sc[S_ZZ] = 0.0;
    sc[S_MODE] = 0.0;
    sc[S_NCELL] = (double)n_cells;
    sc[S_TOL2] = tol * tol;
    sc[S_ITER] = 0.0;
    sc[S_MAXIT] = (double)maxit;
    sc[S_BETA] = 0.0;
    sc[S_DONE] = (red[1] == 0.0 || maxit <= 0) ? 1.0 : 0.0;
  }
}

// Ap = A (z + beta p), pAp, alpha   (beta = 0 and p = z on the first iteration)
__global__ void __launch_bounds__(PBLOCK) mgpcg_spmv_kernel(hp_torso t, const float* __restrict__ z,
                                                           const float* __restrict__ p,
                                                           float* __restrict__ Ap, double* part,
                                                           const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const float beta = (float)sc[S_BETA];
  const uint32_t vb = (uint32_t)t.n_cells * 4u;
  const buf_t Z = buf_rsrc(z, vb), P = buf_rsrc(p, vb), AP = buf_rsrc(Ap, vb);
  double v[1] = {0.0};
  for_cells_u<2>(
      t, [&](uint32_t c, int k, const Chunk& ch) { return dir_loads(t, Z, P, c, k, ch); },
      [&](uint32_t c, int, const Chunk&, const DirLoads& l) {
        float pc;
        const float acc = dir_product(t, c, beta, l, pc);
        buf_st(AP, c, acc);
        v[0] = c != T_OOB ? v[0] + (double)pc * (double)acc : v[0];
      });
  block_partials<1>(v, part);
}

// p = z + beta p (materialised), d += alpha p, r -= alpha Ap, ||r||^2, stopping test
__global__ void __launch_bounds__(PBLOCK) mgpcg_update_kernel(hp_torso t, const float* __restrict__ z,
                                                             float* __restrict__ r,
                                                             float* __restrict__ d,
                                                             float* __restrict__ p,
                                                             const float* __restrict__ Ap,
                                                             double* part,
                                                             const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  const float alpha = (float)sc[S_ALPHA];
  const float beta = (float)sc[S_BETA];
  double v[1] = {0.0};
  const uint32_t vb = (uint32_t)t.n_cells * 4u;
  const buf_t Z = buf_rsrc(z, vb), R = buf_rsrc(r, vb), D = buf_rsrc(d, vb), P = buf_rsrc(p, vb),
              AP = buf_rsrc(Ap, vb);
  struct L {
    float z, p, d, r, ap;
  };
  for_cells_u<4>(
      t,
      [&](uint32_t c, int, const Chunk&) {
        return L{buf_ld<float>(Z, c), buf_ld<float>(P, c), buf_ld<float>(D, c),
                 buf_ld<float>(R, c), buf_ld<float>(AP, c)};
      },
      [&](uint32_t c, int, const Chunk&, const L& l) {
        const float pc = l.z + beta * l.p;
        buf_st(P, c, pc);
        buf_st(D, c, l.d + alpha * pc);
        const float rc = l.r - alpha * l.ap;
        buf_st(R, c, rc);
        v[0] = c != T_OOB ? v[0] + (double)rc * (double)rc : v[0];
      });
  block_partials<1>(v, part);
}

__global__ void __launch_bounds__(PBLOCK) mgpcg_update_finish_kernel(const double* __restrict__ part,
                                                                     int nb, double* sc) {
  if (sc[S_DONE] != 0.0) return;
  double red[1];
  sum_partials<1>(part, nb, red);
  if (threadIdx.x == 0) {
    sc[S_RR] = red[0];
    sc[S_ITER] += 1.0;
    if (red[0] <= sc[S_TOL2] * sc[S_RR0] || sc[S_ITER] >= sc[S_MAXIT]) sc[S_DONE] = 1.0;
  }
}

// rz = r.z (new z), beta = rz / rz_old
__global__ void __launch_bounds__(PBLOCK) mgpcg_rz_kernel(hp_torso t, const float* __restrict__ r,
                                                         const float* __restrict__ z, double* part,
                                                         const double* sc) {
  if (sc[S_DONE] != 0.0) return;
  double v[1] = {0.0};
  const uint32_t vb = (uint32_t)t.n_cells * 4u;
  const buf_t R = buf_rsrc(r, vb), Z = buf_rsrc(z, vb);
  struct L {
    float r, z;
  };
  for_cells_u<4>(
      t, [&](uint32_t c, int, const Chunk&) { return L{buf_ld<float>(R, c), buf_ld<float>(Z, c)}; },
      [&](uint32_t, int, const Chunk&, const L& l) { v[0] += (double)l.r * (double)l.z; });
  block_partials<1>(v, part);
}

__global__ void __launch_bounds__(PBLOCK) mgpcg_rz_finish_kernel(const double* __restrict__ part,
                                                                 int nb, double* sc) {
  if (sc[S_DONE] != 0.0) return;
  double red[1];
  sum_partials<1>(part, nb, red);
  if (threadIdx.x == 0) {
    sc[S_BETA] = red[0] / sc[S_RZ];
    sc[S_RZ] = red[0];
  }
}

static bool mg_ok(const hp_torso* t, const hp_mg* mg, const char* what) {
  if (!torso_ok(t, what)) return false;
  if (!mg || mg->n_levels < 1 || mg->n_levels > HP_MG_MAX_LEVELS || mg->nu < 2 || (mg->nu & 1) ||
      !mg->inv_diag0 || !mg->tmp0 || !mg->coarse_inv) {
    set_error("%s: bad multigrid hierarchy (levels %d, nu %d: nu must be even and >= 2)", what,
              mg ? mg->n_levels : -1, mg ? mg->nu : -1);
    return false;
  }
  const hp_mg_level& c = mg->level[mg->n_levels - 1];
  if (c.n > 1024) {
    set_error("%s: coarsest level has %lld cells (at most 1024)", what, (long long)c.n);
    return false;
  }
  const hp_mg_level& l1 = mg->level[0];
  if (l1.ni != (t->ni + 1) / 2 || l1.nj != (t->nj + 1) / 2 || l1.nk != (t->nk + 1) / 2) {
    set_error("%s: level 1 is %dx%dx%d, expected half of %dx%dx%d", what, l1.ni, l1.nj, l1.nk,
              t->ni, t->nj, t->nk);
    return false;
  }
  for (int l = 1; l < mg->n_levels; ++l) {
    const hp_mg_level &f = mg->level[l - 1], &g = mg->level[l];
    if (g.ni != (f.ni + 1) / 2 || g.nj != (f.nj + 1) / 2 || g.nk != (f.nk + 1) / 2 ||
        g.n != (int64_t)g.ni * g.nj * g.nk) {
      set_error("%s: level %d does not halve level %d", what, l + 1, l);
      return false;
    }
  }
  return true;
}

}  // namespace hp

using namespace hp;

extern "C" {

int hp_mg_vcycle_f32(const hp_torso* t, const hp_mg* mg, const float* r, float* z,
                     const double* scalars, void* stream) {
  if (!mg_ok(t, mg, "hp_mg_vcycle_f32")) return -1;
  vcycle(t, mg, r, z, scalars, (hipStream_t)stream);
  return check_launch("hp_mg_vcycle_f32");
}

int hp_mgpcg_begin_f32(const hp_torso* t, const hp_mg* mg, const float* r, float* d, float* z,
                       float* p, hp_pcg_ws* ws, double tol, int max_iter, void* stream) {
  if (!mg_ok(t, mg, "hp_mgpcg_begin_f32") ||
      !ws_ok(ws, 0, true, "hp_mgpcg_begin_f32")) return -1;
  hipStream_t st = (hipStream_t)stream;
  vcycle(t, mg, r, z, nullptr, st);
  const int nb = grid_blocks(t);
  hipLaunchKernelGGL(mgpcg_init_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, r, z, d, p, ws->part);
  hipLaunchKernelGGL(mgpcg_init_finish_kernel, dim3(1), dim3(PBLOCK), 0, st, ws->part, nb,
                     ws->scalars, tol, max_iter, t->n_cells);
  return check_launch("hp_mgpcg_begin_f32");
}

int hp_mgpcg_iterate_f32(const hp_torso* t, const hp_mg* mg, float* r, float* d, float* z,
                         float* p, float* Ap, hp_pcg_ws* ws, int n_iter, void* stream) {
  if (!mg_ok(t, mg, "hp_mgpcg_iterate_f32") ||
      !ws_ok(ws, n_iter, true, "hp_mgpcg_iterate_f32")) return -1;
  const int nb = grid_blocks(t);
  hipStream_t st = (hipStream_t)stream;
  for (int it = 0; it < n_iter; ++it) {
    hipLaunchKernelGGL(mgpcg_spmv_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, z, p, Ap, ws->part,
                       ws->scalars);
    launch_pap_finish(ws->part, nb, ws->scalars, st);
    hipLaunchKernelGGL(mgpcg_update_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, z, r, d, p, Ap,
                       ws->part, ws->scalars);
    hipLaunchKernelGGL(mgpcg_update_finish_kernel, dim3(1), dim3(PBLOCK), 0, st, ws->part, nb,
                       ws->scalars);
    vcycle(t, mg, r, z, ws->scalars, st);
    hipLaunchKernelGGL(mgpcg_rz_kernel, dim3(nb), dim3(PBLOCK), 0, st, *t, r, z, ws->part,
                       ws->scalars);
    hipLaunchKernelGGL(mgpcg_rz_finish_kernel, dim3(1), dim3(PBLOCK), 0, st, ws->part, nb,
                       ws->scalars);
  }
  return check_launch("hp_mgpcg_iterate_f32");
}

}  // extern "C"
