// Geometry set-up on the device: the transmural coordinate's Laplace solve
// (generate_transmural_coordinate.py:35-186) on the brick layout, and the
// rule-based fibre frame (generate_fibre_direction.py:135-281) on a dense box.
//
// Laplace rows of the myocardial slots: label bit 0 set = a Dirichlet row
// (identity; the caller sets the value in x), otherwise the 6-neighbour graph
// Laplacian deg*x_c - sum of the tissue neighbours (Neumann at the mask
// boundary).  The brick grid's D6 holds the unit tensor (rows 0-2 = 1, rows
// 3-5 = 0), so the 19-point stencil sum is exactly sum_nb (x_nb - x_c).
// CG from x0 with Dirichlet values in place: r = b - A x0 vanishes on the
// Dirichlet rows and stays 0 there, so the iteration is CG on the interior
// block with the boundary values moved to the right-hand side, as the
// reference's CG on its unsymmetrised matrix is.  FP32 vectors, FP64 inner
// products summed per block and then in block order by a one-block finish
// kernel that also applies the reference's rule ||r|| / ||r0|| < tol.
#include "brick_stencil.h"

namespace hp {

constexpr int LAP_FIN_THREADS = 1024;
// scalars: 0 r.r, 1 ||r0||, 2 alpha, 3 beta, 4 done (1 converged, 2 at
// max_iter), 5 iterations, 6 last ||r||/||r0||
enum { LC_RR = 0, LC_R0, LC_ALPHA, LC_BETA, LC_DONE, LC_ITERS, LC_REL };
enum { LAP_INIT = 0, LAP_AP = 1 };

__device__ __forceinline__ double lap_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NW>
__device__ __forceinline__ void lap_block_sum(double a, double* part) {
  __shared__ double sh[NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  a = lap_wave_sum(a);
  if (lane == 0) sh[wave] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += sh[w];
    part[blockIdx.x] = s;
  }
}

// INIT: r = sum_nb (x_nb - x_c) on the Laplace rows (b = 0 there), 0 on the
// Dirichlet rows; p = r; partial r.r.  AP: z = r + beta p_in formed for every
// tile cell while staging; p_out = z, ap = A z; partial z.Az.  One brick per
// workgroup.
template <int BX, int BY, int BZ, int MODE>
__global__ void __launch_bounds__(BX * BY * BZ) lap_pass_kernel(
    hp_bricks g, const float* __restrict__ x, float* __restrict__ r, const float* __restrict__ p_in,
    float* __restrict__ p_out, float* __restrict__ ap, const double* __restrict__ sc,
    double* __restrict__ part) {
  constexpr int BVOL = BX * BY * BZ;
  typedef Tile<float, BX, BY, BZ> TT;
  __shared__ TT t;
  if (MODE == LAP_AP && sc[LC_DONE] != 0.0) return;
  const float beta = MODE == LAP_AP ? (float)sc[LC_BETA] : 0.f;
  const uint32_t b = blockIdx.x, n = (uint32_t)g.n_bricks * BVOL;
  const buf_t rX = buf_rsrc(x, MODE == LAP_INIT ? n * 4u : 0u);
  const buf_t rR = buf_rsrc(r, MODE == LAP_AP ? n * 4u : 0u);
  const buf_t rP = buf_rsrc(p_in, MODE == LAP_AP ? n * 4u : 0u);
  const int tid = threadIdx.x;
  {
    HaloNbr<BX, BY, BZ> hn;
    halo_nbr<BX, BY, BZ>(hn, g, b);
    HaloRegs<float, BX, BY, BZ> hr;
    halo_load_fn<float, BX, BY, BZ>(hr, hn, g, g.label[b * BVOL + tid], [&](uint32_t q) {
#pragma clang fp contract(off)
      if (MODE == LAP_INIT) return buf_ld<float>(rX, q);
      return buf_ld<float>(rR, q) + beta * buf_ld<float>(rP, q);
    });
    halo_store_nan<float, BX, BY, BZ>(t, hr);
    __syncthreads();
  }
  const int lz = tid % BZ, ly = (tid / BZ) % BY, lx = tid / (BY * BZ);
  const int c = ((lx + 1) * TT::TY + (ly + 1)) * TT::TZ + (lz + 1);
  const uint32_t slot = b * BVOL + tid;
  double acc = 0.0;
  const uint8_t lab = t.L[c];
  if (lab != 0xFF) {
    float ws;
    const float zc = t.V[c];
    const float lsum = stencil19_wsum_nan<float, BX, BY, BZ>(t, c, ws);
    const bool dir = (lab & 1) != 0;
    if (MODE == LAP_INIT) {
      const float rv = dir ? 0.f : lsum;
      r[slot] = rv;
      p_out[slot] = rv;
      acc = (double)rv * (double)rv;
    } else {
      const float av = dir ? zc : -lsum;
      p_out[slot] = zc;
      ap[slot] = av;
      acc = (double)zc * (double)av;
    }
  }
  lap_block_sum<(BVOL + 63) / 64>(acc, part);
}

// x += alpha p, r -= alpha Ap; partial r.r
template <int BVOL>
__global__ void __launch_bounds__(BVOL) lap_update_kernel(hp_bricks g, float* __restrict__ x,
                                                          float* __restrict__ r,
                                                          const float* __restrict__ p,
                                                          const float* __restrict__ ap,
                                                          const double* __restrict__ sc,
                                                          double* __restrict__ part) {
#pragma clang fp contract(off)
  if (sc[LC_DONE] != 0.0) return;
  const float alpha = (float)sc[LC_ALPHA];
  const uint32_t slot = blockIdx.x * BVOL + threadIdx.x;
  double acc = 0.0;
  if (g.label[slot] != 0xFF) {
    x[slot] = x[slot] + alpha * p[slot];
    const float rv = r[slot] - alpha * ap[slot];
    r[slot] = rv;
    acc = (double)rv * (double)rv;
  }
  lap_block_sum<(BVOL + 63) / 64>(acc, part);
}

// Sum of nb partials in a fixed order (thread 0 holds it).
__device__ __forceinline__ double lap_sum_part(const double* __restrict__ part, int nb) {
  __shared__ double sh[LAP_FIN_THREADS / 64];
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += LAP_FIN_THREADS) a += part[i];
  a = lap_wave_sum(a);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < LAP_FIN_THREADS / 64; ++w) s += sh[w];
  return s;
}

enum { LF_BEGIN = 0, LF_ALPHA, LF_UPDATE };

// BEGIN: r.r and ||r0|| (a zero residual: x0 solves the system, done);
// ALPHA: alpha = r.r / p.Ap; UPDATE: the reference's stopping rule on the new
// r.r, else beta = r.r_new / r.r.
template <int STAGE>
__global__ void __launch_bounds__(LAP_FIN_THREADS) lap_finish_kernel(double* sc,
                                                                     const double* part, int nb,
                                                                     double tol, int64_t max_iter) {
  if (STAGE != LF_BEGIN && sc[LC_DONE] != 0.0) return;
  const double s = lap_sum_part(part, nb);
  if (threadIdx.x != 0) return;
  if (STAGE == LF_BEGIN) {
    sc[LC_RR] = s;
    sc[LC_R0] = sqrt(s);
    sc[LC_BETA] = 0.0;
    sc[LC_ITERS] = 0.0;
    sc[LC_REL] = 1.0;
    sc[LC_DONE] = s == 0.0 ? 1.0 : 0.0;
    return;
  }
  if (STAGE == LF_ALPHA) {
    sc[LC_ALPHA] = sc[LC_RR] / s;
    return;
  }
  const double it = sc[LC_ITERS] + 1.0;
  sc[LC_ITERS] = it;
  const double rel = sqrt(s) / (sc[LC_R0] + 1e-30);
  sc[LC_REL] = rel;
  if (rel < tol) {
    sc[LC_DONE] = 1.0;
  } else if (it >= (double)max_iter) {
    sc[LC_DONE] = 2.0;
  } else {
    sc[LC_BETA] = s / sc[LC_RR];
    sc[LC_RR] = s;
  }
}

template <int MODE>
static void lap_pass(int key, int64_t nb, hipStream_t st, const hp_bricks& g, const float* x,
                     float* r, const float* pin, float* pout, float* ap, const double* sc,
                     double* part) {
  const dim3 grid((unsigned)nb);
  switch (key) {
    case 444: hipLaunchKernelGGL((lap_pass_kernel<4, 4, 4, MODE>), grid, dim3(64), 0, st, g, x, r, pin, pout, ap, sc, part); break;
    case 844: hipLaunchKernelGGL((lap_pass_kernel<8, 4, 4, MODE>), grid, dim3(128), 0, st, g, x, r, pin, pout, ap, sc, part); break;
    case 884: hipLaunchKernelGGL((lap_pass_kernel<8, 8, 4, MODE>), grid, dim3(256), 0, st, g, x, r, pin, pout, ap, sc, part); break;
    default:  hipLaunchKernelGGL((lap_pass_kernel<8, 8, 8, MODE>), grid, dim3(512), 0, st, g, x, r, pin, pout, ap, sc, part); break;
  }
}

static int lap_check(const hp_bricks* g, const hp_lap_cg* cg, const char* what, int* key) {
  if (!g || !cg || !g->label || !g->D6 || !g->nbr || g->n_bricks <= 0 || !cg->x || !cg->r ||
      !cg->p[0] || !cg->p[1] || !cg->ap || !cg->scalars || !cg->part) {
    set_error("%s: brick grid (label, D6, nbr, n_bricks > 0) and the CG vectors x, r, p[2], "
              "ap, scalars, part are required",
              what);
    return -1;
  }
  const int64_t bvol = (int64_t)g->bx * g->by * g->bz;
  if ((uint64_t)g->n_bricks * bvol * 6 * sizeof(float) >= (1ull << 32) ||
      g->n_bricks > 0x7fffffff) {
    set_error("%s: %lld bricks exceed the 32-bit offset range", what, (long long)g->n_bricks);
    return -1;
  }
  if (cg->part_len < g->n_bricks) {
    set_error("%s: part_len >= n_bricks (%lld) doubles required", what, (long long)g->n_bricks);
    return -1;
  }
  *key = g->bx * 100 + g->by * 10 + g->bz;
  if (*key != 444 && *key != 844 && *key != 884 && *key != 888) {
    set_error("%s: unsupported brick shape %dx%dx%d", what, g->bx, g->by, g->bz);
    return -1;
  }
  return 0;
}

// Fibre frame at one voxel of the dense box: the gradients of the filled
// fields by np.gradient's rule (central differences inside, one-sided at the
// box faces), e_t, e_l, e_c and the helix rotation, in FP32 in the
// reference's order of operations.
__device__ __forceinline__ float fb_grad(const float* __restrict__ F, int64_t idx, int i, int n,
                                         int64_t stride, float h, float h2) {
#pragma clang fp contract(off)
  if (i == 0) return (F[idx + stride] - F[idx]) / h;
  if (i == n - 1) return (F[idx] - F[idx - stride]) / h;
  return (F[idx + stride] - F[idx - stride]) / h2;
}

__device__ __forceinline__ void fb_unit(float& x, float& y, float& z) {
#pragma clang fp contract(off)
  float nrm = sqrtf((x * x + y * y) + z * z);
  nrm = fmaxf(nrm, 1e-8f);
  x = x / nrm;
  y = y / nrm;
  z = z / nrm;
}

__global__ void __launch_bounds__(256) fibre_rule_kernel(const float* __restrict__ Fphi,
                                                         const float* __restrict__ Fpsi, int nx,
                                                         int ny, int nz, float hx, float hy,
                                                         float hz, float hx2, float hy2,
                                                         float hz2, float a0, float da,
                                                         float* __restrict__ f) {
#pragma clang fp contract(off)
  const int64_t n = (int64_t)nx * ny * nz;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * 256) {
    const int k = (int)(idx % nz), j = (int)((idx / nz) % ny), i = (int)(idx / ((int64_t)nz * ny));
    const int64_t sx = (int64_t)ny * nz, sy = nz;
    float tx = fb_grad(Fphi, idx, i, nx, sx, hx, hx2);
    float ty = fb_grad(Fphi, idx, j, ny, sy, hy, hy2);
    float tz = fb_grad(Fphi, idx, k, nz, 1, hz, hz2);
    fb_unit(tx, ty, tz);
    const float gx = fb_grad(Fpsi, idx, i, nx, sx, hx, hx2);
    const float gy = fb_grad(Fpsi, idx, j, ny, sy, hy, hy2);
    const float gz = fb_grad(Fpsi, idx, k, nz, 1, hz, hz2);
    const float d = (gx * tx + gy * ty) + gz * tz;
    float lx = gx - d * tx, ly = gy - d * ty, lz = gz - d * tz;
    fb_unit(lx, ly, lz);
    float cx = ty * lz - tz * ly, cy = tz * lx - tx * lz, cz = tx * ly - ty * lx;
    fb_unit(cx, cy, cz);
    const float a = a0 + da * Fphi[idx];
    const float ca = cosf(a), sa = sinf(a);
    float fx = ca * cx + sa * lx, fy = ca * cy + sa * ly, fz = ca * cz + sa * lz;
    fb_unit(fx, fy, fz);
    f[3 * idx + 0] = fx;
    f[3 * idx + 1] = fy;
    f[3 * idx + 2] = fz;
  }
}

}  // namespace hp

using namespace hp;

extern "C" {

int hp_laplace_cg_begin_f32(const hp_bricks* g, hp_lap_cg* cg, void* stream) {
  int key;
  if (lap_check(g, cg, "hp_laplace_cg_begin_f32", &key)) return -1;
  hipStream_t st = (hipStream_t)stream;
  lap_pass<LAP_INIT>(key, g->n_bricks, st, *g, cg->x, cg->r, nullptr, cg->p[0], nullptr,
                     cg->scalars, cg->part);
  if (int rc = check_launch("hp_laplace_cg_begin_f32")) return rc;
  hipLaunchKernelGGL((lap_finish_kernel<LF_BEGIN>), dim3(1), dim3(LAP_FIN_THREADS), 0, st,
                     cg->scalars, cg->part, (int)g->n_bricks, cg->tol, (int64_t)1);
  cg->launched = 0;
  return check_launch("hp_laplace_cg_begin_f32");
}

int hp_laplace_cg_iterate_f32(const hp_bricks* g, hp_lap_cg* cg, int32_t n_iter, int64_t max_iter,
                              void* stream) {
  int key;
  if (lap_check(g, cg, "hp_laplace_cg_iterate_f32", &key)) return -1;
  if (n_iter < 0 || max_iter < 1) {
    set_error("hp_laplace_cg_iterate_f32: n_iter >= 0 and max_iter >= 1 required");
    return -1;
  }
  hipStream_t st = (hipStream_t)stream;
  const int nb = (int)g->n_bricks;
  const int bvol = g->bx * g->by * g->bz;
  for (int i = 0; i < n_iter; ++i) {
    const int64_t k = cg->launched + i;
    float* pin = cg->p[k & 1];
    float* pout = cg->p[(k + 1) & 1];
    lap_pass<LAP_AP>(key, nb, st, *g, nullptr, cg->r, pin, pout, cg->ap, cg->scalars, cg->part);
    hipLaunchKernelGGL((lap_finish_kernel<LF_ALPHA>), dim3(1), dim3(LAP_FIN_THREADS), 0, st,
                       cg->scalars, cg->part, nb, cg->tol, max_iter);
    switch (bvol) {
      case 64: hipLaunchKernelGGL((lap_update_kernel<64>), dim3(nb), dim3(64), 0, st, *g, cg->x, cg->r, pout, cg->ap, cg->scalars, cg->part); break;
      case 128: hipLaunchKernelGGL((lap_update_kernel<128>), dim3(nb), dim3(128), 0, st, *g, cg->x, cg->r, pout, cg->ap, cg->scalars, cg->part); break;
      case 256: hipLaunchKernelGGL((lap_update_kernel<256>), dim3(nb), dim3(256), 0, st, *g, cg->x, cg->r, pout, cg->ap, cg->scalars, cg->part); break;
      default: hipLaunchKernelGGL((lap_update_kernel<512>), dim3(nb), dim3(512), 0, st, *g, cg->x, cg->r, pout, cg->ap, cg->scalars, cg->part); break;
    }
    hipLaunchKernelGGL((lap_finish_kernel<LF_UPDATE>), dim3(1), dim3(LAP_FIN_THREADS), 0, st,
                       cg->scalars, cg->part, nb, cg->tol, max_iter);
    if (int rc = check_launch("hp_laplace_cg_iterate_f32")) return rc;
  }
  cg->launched += n_iter;
  return 0;
}

int hp_fibre_rule_f32(const float* phi_filled, const float* psi_filled, int32_t nx, int32_t ny,
                      int32_t nz, double hx, double hy, double hz, double alpha_endo,
                      double alpha_epi, float* f, void* stream) {
  if (nx < 0 || ny < 0 || nz < 0 || (int64_t)nx * ny * nz > ((int64_t)1 << 40)) {
    set_error("hp_fibre_rule_f32: bad box %d x %d x %d", nx, ny, nz);
    return -1;
  }
  const int64_t n = (int64_t)nx * ny * nz;
  if (n == 0) return 0;
  if (!phi_filled || !psi_filled || !f || nx < 2 || ny < 2 || nz < 2 || !(hx > 0) ||
      !(hy > 0) || !(hz > 0)) {
    set_error("hp_fibre_rule_f32: phi, psi and f are required, every box side >= 2 and the "
              "spacings > 0");
    return -1;
  }
  const int64_t blocks = n / 256 + 1 < 65536 ? n / 256 + 1 : 65536;
  hipLaunchKernelGGL(fibre_rule_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     phi_filled, psi_filled, nx, ny, nz, (float)hx, (float)hy, (float)hz,
                     (float)(2.0 * hx), (float)(2.0 * hy), (float)(2.0 * hz), (float)alpha_endo,
                     (float)(alpha_epi - alpha_endo), f);
  return check_launch("hp_fibre_rule_f32");
}

}  // extern "C"
