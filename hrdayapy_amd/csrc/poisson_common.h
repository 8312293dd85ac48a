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

// Reductions without fences.  Each block of a producer kernel writes its
// partial sums to part[q * gridDim.x + blockIdx.x] with plain stores; a
// one-block finish kernel launched after it on the same stream (the kernel
// boundary makes the partials visible) sums them in block order.  A
// device-scope fence per block (the "last block finishes" pattern) would
// write back and invalidate the XCD's L2 at every block end.
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double* part) {
  __shared__ double sh[NV][PWAVES];
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
  }
}

// In a one-block (PBLOCK threads) finish kernel: the sums of the nb partials
// of each value, in a fixed order (eight independent loads per thread and
// pass), returned to every thread.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* __restrict__ part, int nb,
                                             double (&out)[NV]) {
  __shared__ double sh[NV][PWAVES];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int b0 = 0; b0 < nb; b0 += 8 * PBLOCK) {
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int b = b0 + (int)threadIdx.x + u * PBLOCK;
        const double x = part[q * nb + (b < nb ? b : 0)];
        s += b < nb ? x : 0.0;
      }
    }
    s = wave_sum(s);
    if (lane == 0) sh[q][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int w = 0; w < PWAVES; ++w) s += sh[q][w];
    out[q] = s;
  }
}

// The metadata of a column and of its four x/y neighbours (used where a
// kernel visits cells by coarse-grid position rather than by chunk).
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

constexpr uint32_t T_OOB = 0x1FFFFFFFu;  // element index that reads 0 / drops stores

// Neighbour lookup in the column-meta form: q = 0 own column (k is then the
// neighbour's k), 1..4 = +x, -x, +y, -y; T_OOB where there is no stored cell.
struct ColLookup {
  const ColNbr& m;
  __device__ __forceinline__ uint32_t operator()(int q, int k) const {
    const ColMeta& c = q == 0 ? m.c : q == 1 ? m.xp : q == 2 ? m.xm : q == 3 ? m.yp : m.ym;
    const int n = cell_at(c, k);
    return n >= 0 ? (uint32_t)n : T_OOB;
  }
};

// One 64-cell chunk of a column run, as stored in the chunk table (16 int32
// per chunk, hp_torso_chunks): k of lane 0, number of cells, column (i, j),
// and for the own column and its +x/-x/+y/-y neighbours the index base
// (stored index of cell k = base + k) and the stored k-range [lo, hi) packed
// as lo | hi << 16 (0 for an absent column).
struct Chunk {
  int k_start, n, i, j;
  int base[5];
  int span[5];
  __device__ __forceinline__ uint32_t operator()(int q, int k) const {
    const int lo = span[q] & 0xffff, hi = span[q] >> 16;
    return (k >= lo && k < hi) ? (uint32_t)(base[q] + k) : T_OOB;
  }
};

__device__ __forceinline__ Chunk load_chunk(const int32_t* __restrict__ tab, int64_t w) {
  // w is wave-uniform: one 64-byte scalar load of the whole record (read as
  // one value, so no part of it is deferred into the lane branch)
  typedef int v16i __attribute__((ext_vector_type(16)));
  const v16i r = *reinterpret_cast<const v16i*>(tab + w * HP_TORSO_CHUNK_INTS);
  Chunk ch;
  ch.k_start = r[0];
  ch.n = r[1];
  ch.i = r[2];
  ch.j = r[3];
  ch.base[0] = r[4]; ch.base[1] = r[5]; ch.base[2] = r[6]; ch.base[3] = r[7]; ch.base[4] = r[8];
  ch.span[0] = r[9]; ch.span[1] = r[10]; ch.span[2] = r[11]; ch.span[3] = r[12]; ch.span[4] = r[13];
  return ch;
}

// One wave per chunk, grid-stride over the chunk table.  With a grid of 8·m
// blocks, the blocks of one XCD (dispatch is round-robin over the 8 XCDs)
// walk one contiguous eighth of the table, so the neighbour columns that a
// chunk reads were recently read by the same XCD and sit in its L2.
// F(cell, k, chunk) is called for the lanes that hold a cell.
template <typename F>
__device__ __forceinline__ void for_cells(const hp_torso& t, F&& f) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t nch = t.n_chunks;
  int64_t w, hi, stride;
  if (gridDim.x >= 8 && (gridDim.x & 7) == 0) {
    const int xcd = blockIdx.x & 7;
    w = nch * xcd / 8 + (int64_t)(blockIdx.x >> 3) * PWAVES + wave;
    hi = nch * (xcd + 1) / 8;
    stride = (int64_t)(gridDim.x >> 3) * PWAVES;
  } else {
    w = (int64_t)blockIdx.x * PWAVES + wave;
    hi = nch;
    stride = (int64_t)gridDim.x * PWAVES;
  }
  for (; w < hi; w += stride) {
    const Chunk ch = load_chunk(t.chunks, w);
    if (lane < ch.n) {
      const int k = ch.k_start + lane;
      f(ch.base[0] + k, k, ch);
    }
  }
}

// Unrolled walk for the hot kernels: each wave takes U chunks per pass
// (stride apart), loads everything for all U cells first (ld), then computes
// and stores (fin), so U chunks' loads are in flight together.  Lanes without
// a cell (past the chunk's end, or a chunk past the range) get c = T_OOB:
// their buffer loads read 0 and their buffer stores are dropped, so ld and
// fin run unconditionally and only reductions need the c != T_OOB test.
template <int U, typename LD, typename FIN>
__device__ __forceinline__ void for_cells_u(const hp_torso& t, LD&& ld, FIN&& fin) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t nch = t.n_chunks;
  int64_t w, hi, stride;
  if (gridDim.x >= 8 && (gridDim.x & 7) == 0) {
    const int xcd = blockIdx.x & 7;
    w = nch * xcd / 8 + (int64_t)(blockIdx.x >> 3) * PWAVES + wave;
    hi = nch * (xcd + 1) / 8;
    stride = (int64_t)(gridDim.x >> 3) * PWAVES;
  } else {
    w = (int64_t)blockIdx.x * PWAVES + wave;
    hi = nch;
    stride = (int64_t)gridDim.x * PWAVES;
  }
  using R = decltype(ld(0u, 0, Chunk{}));
  for (; w < hi; w += U * stride) {
    Chunk ch[U];
    uint32_t c[U];
    int k[U];
    R r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t wu = w + u * stride;
      ch[u] = load_chunk(t.chunks, wu < hi ? wu : hi - 1);  // clamped: no branch
      const int n = wu < hi ? ch[u].n : 0;
      k[u] = ch[u].k_start + lane;
      c[u] = lane < n ? (uint32_t)(ch[u].base[0] + k[u]) : T_OOB;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) r[u] = ld(c[u], k[u], ch[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) fin(c[u], k[u], ch[u], r[u]);
  }
}

// The six faces of a row in the order +z, -z, +x, -x, +y, -y: index of the
// neighbour value (n) and of the face conductance (gi: the cell's own +z/+x/+y
// face, the neighbour's for -z/-x/-y); T_OOB for an absent neighbour, whose
// conductance and value then read as 0 and whose term vanishes.
struct RowIdx {
  uint32_t n[6], gi[6];
};

template <typename NB>
__device__ __forceinline__ RowIdx row_idx(uint32_t c, int k, const NB& nb) {
  RowIdx r;
  r.n[0] = nb(0, k + 1);
  r.n[1] = nb(0, k - 1);
  r.n[2] = nb(1, k);
  r.n[3] = nb(2, k);
  r.n[4] = nb(3, k);
  r.n[5] = nb(4, k);
#pragma unroll
  for (int f = 0; f < 6; f += 2) {
    r.gi[f] = r.n[f] == T_OOB ? T_OOB : c;
    r.gi[f + 1] = r.n[f + 1];
  }
  return r;
}

__device__ __forceinline__ void load_g(const hp_torso& t, const RowIdx& ix, float (&g)[6]) {
  const uint32_t gb = (uint32_t)t.n_cells * 4u;
  const buf_t Gx = buf_rsrc(t.gx, gb), Gy = buf_rsrc(t.gy, gb), Gz = buf_rsrc(t.gz, gb);
  g[0] = buf_ld<float>(Gz, ix.gi[0]);
  g[1] = buf_ld<float>(Gz, ix.gi[1]);
  g[2] = buf_ld<float>(Gx, ix.gi[2]);
  g[3] = buf_ld<float>(Gx, ix.gi[3]);
  g[4] = buf_ld<float>(Gy, ix.gi[4]);
  g[5] = buf_ld<float>(Gy, ix.gi[5]);
}

// (A x)_c in FP32 in zero-row-sum form, sum_f g_f (x_c - x_nb): exact for a
// constant x, so the large near-constant component of a torso potential
// measured against one pinned cell does not leak into the product through a
// rounded diagonal.  The pinned cell and cells without any conducting face
// are identity rows.
__device__ __forceinline__ float row_sum(const hp_torso& t, uint32_t c, float xc,
                                         const float (&g)[6], const float (&x)[6]) {
  float acc = 0.f, gs = 0.f;
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    acc += g[f] * (xc - x[f]);
    gs += g[f];
  }
  return ((int64_t)c == t.pin || gs == 0.f) ? xc : acc;
}

// The same row with the neighbour values from xat(index); every load is
// issued unconditionally.
template <typename NB, typename F>
__device__ __forceinline__ float row_f32(const hp_torso& t, uint32_t c, int k, const NB& nb,
                                         float xc, F&& xat) {
  const RowIdx ix = row_idx(c, k, nb);
  float g[6], x[6];
  load_g(t, ix, g);
#pragma unroll
  for (int f = 0; f < 6; ++f) x[f] = xat(ix.n[f]);
  return row_sum(t, c, xc, g, x);
}

template <typename NB>
__device__ __forceinline__ float apply_row_f32(const hp_torso& t, const float* __restrict__ x,
                                               uint32_t c, int k, const NB& nb) {
  const buf_t X = buf_rsrc(x, (uint32_t)t.n_cells * 4u);
  return row_f32(t, c, k, nb, buf_ld<float>(X, c), [&](uint32_t n) { return buf_ld<float>(X, n); });
}

// Loads and product of Ap = A (z + beta p) at one cell (the PCG direction
// update folded into the product): the new direction of a neighbour comes
// from its z and old p.
struct DirLoads {
  float g[6], z[7], p[7];  // [6] = the cell itself
};

__device__ __forceinline__ DirLoads dir_loads(const hp_torso& t, buf_t Z, buf_t P, uint32_t c,
                                              int k, const Chunk& ch) {
  DirLoads l;
  const RowIdx ix = row_idx(c, k, ch);
  load_g(t, ix, l.g);
#pragma unroll
  for (int f = 0; f < 6; ++f) {
    l.z[f] = buf_ld<float>(Z, ix.n[f]);
    l.p[f] = buf_ld<float>(P, ix.n[f]);
  }
  l.z[6] = buf_ld<float>(Z, c);
  l.p[6] = buf_ld<float>(P, c);
  return l;
}

// returns A (z + beta p) at c; pc = the cell's own new direction
__device__ __forceinline__ float dir_product(const hp_torso& t, uint32_t c, float beta,
                                             const DirLoads& l, float& pc) {
  float x[6];
#pragma unroll
  for (int f = 0; f < 6; ++f) x[f] = l.z[f] + beta * l.p[f];
  pc = l.z[6] + beta * l.p[6];
  return row_sum(t, c, pc, l.g, x);
}

// ---------------------------------------------------------------------------
// PCG state in device memory (FP64 scalars)
// ---------------------------------------------------------------------------
enum { S_RZ = 0, S_PAP, S_ALPHA, S_BETA, S_RR0, S_RR, S_TOL2, S_DONE, S_ITER, S_MAXIT, S_ZZ,
       S_MODE, S_NCELL, S_ZSTAT, S_NSC };

// Blocks of PBLOCK threads over the chunk table: at most 4096 (16 per CU;
// measured 6 % faster for the SpMV than 2048 on the 18.6 M-cell torso), a
// multiple of 8 when >= 8 (for_cells splits the table between the XCDs);
// partials of up to 3 values then fit the 3*4096-double reduction workspace.
constexpr int PMAX_BLOCKS = 4096;
inline int grid_blocks(const hp_torso* t) {
  int64_t nb = (t->n_chunks + PWAVES - 1) / PWAVES;
  if (nb > PMAX_BLOCKS) nb = PMAX_BLOCKS;
  if (nb >= 8) nb &= ~(int64_t)7;
  return (int)(nb < 1 ? 1 : nb);
}

// The kernels address the cell arrays with 32-bit byte offsets (FP64 vectors
// included) and pack k into 16 bits in the chunk table.
// pAp from the partials of a product kernel, alpha = rz / pAp (poisson.hip)
void launch_pap_finish(double* part, int nb, double* sc, hipStream_t st);

// the workspace: partials plus the PCG scalar slots (pcg) or the two sums
// (residual); a non-negative iteration count
inline bool ws_ok(const hp_pcg_ws* ws, int64_t n_iter, bool pcg, const char* what) {
  if (!ws || !ws->part || !(pcg ? (void*)ws->scalars : (void*)ws->out) || n_iter < 0) {
    set_error("%s: workspace (part and %s) required, n_iter >= 0", what, pcg ? "scalars" : "out");
    return false;
  }
  return true;
}

inline bool torso_ok(const hp_torso* t, const char* what) {
  if (!t || t->ni <= 0 || t->nj <= 0 || t->nk <= 0 || t->nk > 32767 || !t->chunks ||
      t->n_chunks < 0 || t->n_cells < 0 || t->n_cells > ((int64_t)1 << 28)) {
    set_error("%s: bad torso grid (nk <= 32767, n_cells <= 2^28, chunk table required)", what);
    return false;
  }
  return true;
}

}  // namespace hp
