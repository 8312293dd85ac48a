// Brick-layout building blocks shared by the monodomain kernels: the
// XCD-aware brick order, the LDS halo tile of a brick, its compile-time
// staging plan, 32-bit-offset loads/stores, and the 19-point operator of the
// anisotropic monodomain (purkinje_myocardium_anisotropic_solver.py:886-1074:
// face weight harm(D_kk,a, D_kk,b)*h, edge-diagonal weight +-0.5*mean(D_ab)*h/2).
#pragma once
#include "common.h"
#include "hrdaya_hip.h"

namespace hp {

__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  // Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
  // contiguous run of bricks so that halo re-reads stay in its own L2.
  const int xcd = bid & 7, k = bid >> 3;
  const int base = nblocks >> 3, rem = nblocks & 7;
  return xcd * base + (xcd < rem ? xcd : rem) + k;
}

template <typename T, int BX, int BY, int BZ>
struct Tile {
  static constexpr int TX = BX + 2, TY = BY + 2, TZ = BZ + 2;
  static constexpr int N = TX * TY * TZ;
  static constexpr int SX = TY * TZ, SY = TZ;
  static constexpr int BVOL = BX * BY * BZ;
  // one spare cell at index N takes the writes of lanes without a tile cell
  T V[N + 1];
  T D[6][N + 1];
  uint8_t L[N + 1];
};

// Cells of the halo tile that the 19-point stencil reads, enumerated so that
// consecutive lanes read consecutive slots of one neighbour brick:
//   [0, NI)           the brick itself (lane = own slot), all six D components;
//   then the -x/+x face layers (BY*BZ cells each, contiguous slots), the -y/+y
//   layers (runs of BZ), the -z/+z layers (stride BZ): three D components each
//   (the face-axis diagonal and the two off-diagonals of the planes through the
//   face); then the four z-, y- and x-parallel edge lines of the tile: one
//   off-diagonal each.  Tile corners are never read.
template <int BX, int BY, int BZ>
struct HaloPlan {
  static constexpr int NI = BX * BY * BZ;
  static constexpr int E0 = NI, E1 = E0 + BY * BZ, E2 = E1 + BY * BZ, E3 = E2 + BX * BZ,
                       E4 = E3 + BX * BZ, E5 = E4 + BX * BY, E6 = E5 + BX * BY,
                       E7 = E6 + 4 * BZ, E8 = E7 + 4 * BY, NT = E8 + 4 * BX;
  static constexpr int NIT = (NT + NI - 1) / NI;
};

struct HaloCell {
  int tile;   // index into the LDS tile, -1 = no cell for this lane
  int dir;    // neighbour-table entry of the brick holding the cell
  int local;  // slot of the cell inside that brick
  int kind;   // 0 none, 1 own brick (six D components), 2 face layer (three), 3 edge line (one)
  int c0;     // first D component of a face layer / the component of an edge line
};

template <int BX, int BY, int BZ>
constexpr HaloCell halo_cell(int e) {
  typedef HaloPlan<BX, BY, BZ> H;
  constexpr int TY = BY + 2, TZ = BZ + 2;
  HaloCell h{-1, 13, 0, 0, 0};
  int gx = 0, gy = 0, gz = 0, dx = 0, dy = 0, dz = 0, lx = 0, ly = 0, lz = 0;
  h.kind = 2;
  if (e < H::E0) {  // own brick
    lx = e / (BY * BZ); ly = (e / BZ) % BY; lz = e % BZ;
    gx = lx; gy = ly; gz = lz;
    h.kind = 1;
  } else if (e < H::E2) {  // x layers: D00, D01, D02
    const int q = e - H::E0, sd = q / (BY * BZ), r = q % (BY * BZ);
    ly = r / BZ; lz = r % BZ; dx = sd ? 1 : -1; lx = sd ? 0 : BX - 1;
    gx = sd ? BX : -1; gy = ly; gz = lz;
    h.c0 = 0;
  } else if (e < H::E4) {  // y layers: D11, D01, D12
    const int q = e - H::E2, sd = q / (BX * BZ), r = q % (BX * BZ);
    lx = r / BZ; lz = r % BZ; dy = sd ? 1 : -1; ly = sd ? 0 : BY - 1;
    gx = lx; gy = sd ? BY : -1; gz = lz;
    h.c0 = 1;
  } else if (e < H::E6) {  // z layers: D22, D02, D12
    const int q = e - H::E4, sd = q / (BX * BY), r = q % (BX * BY);
    lx = r / BY; ly = r % BY; dz = sd ? 1 : -1; lz = sd ? 0 : BZ - 1;
    gx = lx; gy = ly; gz = sd ? BZ : -1;
    h.c0 = 2;
  } else if (e < H::E7) {  // z-parallel edges: D01
    const int q = e - H::E6, ed = q / BZ;
    const int sx = ed >> 1, sy = ed & 1;
    lz = q % BZ; dx = sx ? 1 : -1; dy = sy ? 1 : -1;
    lx = sx ? 0 : BX - 1; ly = sy ? 0 : BY - 1;
    gx = sx ? BX : -1; gy = sy ? BY : -1; gz = lz;
    h.kind = 3; h.c0 = 3;
  } else if (e < H::E8) {  // y-parallel edges: D02
    const int q = e - H::E7, ed = q / BY;
    const int sx = ed >> 1, sz = ed & 1;
    ly = q % BY; dx = sx ? 1 : -1; dz = sz ? 1 : -1;
    lx = sx ? 0 : BX - 1; lz = sz ? 0 : BZ - 1;
    gx = sx ? BX : -1; gy = ly; gz = sz ? BZ : -1;
    h.kind = 3; h.c0 = 4;
  } else if (e < H::NT) {  // x-parallel edges: D12
    const int q = e - H::E8, ed = q / BX;
    const int sy = ed >> 1, sz = ed & 1;
    lx = q % BX; dy = sy ? 1 : -1; dz = sz ? 1 : -1;
    ly = sy ? 0 : BY - 1; lz = sz ? 0 : BZ - 1;
    gx = lx; gy = sy ? BY : -1; gz = sz ? BZ : -1;
    h.kind = 3; h.c0 = 5;
  } else {
    h.kind = 0;
    return h;
  }
  h.tile = ((gx + 1) * TY + (gy + 1)) * TZ + (gz + 1);
  h.dir = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
  h.local = (lx * BY + ly) * BZ + lz;
  return h;
}

// The staging plan of every (iteration, lane), packed as
// tile | dir << 10 | local << 15 | kind << 24 | c0 << 26, built at compile time
// and read from constant memory: the kernels decode one word per iteration
// instead of evaluating the case analysis above.
template <int BX, int BY, int BZ>
struct PlanTable {
  static constexpr int N = HaloPlan<BX, BY, BZ>::NIT * BX * BY * BZ;
  uint32_t w[N];
  constexpr PlanTable() : w() {
    for (int e = 0; e < N; ++e) {
      const HaloCell h = halo_cell<BX, BY, BZ>(e);
      w[e] = h.kind == 0 ? 0u
                         : (uint32_t)h.tile | ((uint32_t)h.dir << 10) | ((uint32_t)h.local << 15) |
                               ((uint32_t)h.kind << 24) | ((uint32_t)h.c0 << 26);
    }
  }
};
template <int BX, int BY, int BZ>
__constant__ PlanTable<BX, BY, BZ> kPlan = PlanTable<BX, BY, BZ>();

// Element idx of a uniform base pointer, addressed as base + 32-bit byte
// offset (the global "saddr + voffset" form: no 64-bit address arithmetic per
// lane).  Callers keep idx * sizeof(T) below 2^32 (checked at launch).
template <typename T>
__device__ __forceinline__ T ld32(const T* base, uint32_t idx) {
  return *(const T*)((const char*)base + (uint32_t)(idx * (uint32_t)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st32(T* base, uint32_t idx, T v) {
  *(T*)((char*)base + (uint32_t)(idx * (uint32_t)sizeof(T))) = v;
}

// Halo staging, split into a load half (global -> registers) and a store half
// (registers -> LDS) so that a persistent kernel can issue the loads of its
// next brick before it computes the current one.  Per lane and plan
// iteration: the plan word, the slot of the cell (OOB for cells without a
// brick: their loads return 0 without memory traffic), label, V and up to
// three D components.  The lane's own cell (iteration 0) loads all six D
// components, padding voxels of the brick too (their label masks them).
// Padding slots of neighbour bricks are read (they hold V0 and D = 0) and
// masked by their label.
constexpr uint32_t OOB = 0x1FFFFFFFu;  // element index past every bound
constexpr int MAX_GROUPS = 8192;       // stimulus groups (group_amp entries) at most

template <int BX, int BY, int BZ>
struct HaloNbr {
  static constexpr int NIT = HaloPlan<BX, BY, BZ>::NIT;
  uint32_t pw[NIT];  // plan words of this lane
  int nb[NIT];       // brick holding each cell, -1 = none (iteration 0: the brick itself)
  bool valid;        // false: no brick (its loads read out of range)
};

// The lane's plan words and the neighbour-table entries they name (read from
// the brick's 27-entry row; one cache line per brick).  Nothing here consumes
// a loaded value, so a persistent kernel can issue these loads early and use
// them an iteration later; valid = false issues out-of-range reads only.
template <int BX, int BY, int BZ>
__device__ __forceinline__ void halo_nbr(HaloNbr<BX, BY, BZ>& n, const hp_bricks& g, uint32_t b,
                                         bool valid = true) {
  constexpr int BVOL = BX * BY * BZ;
  const buf_t rN = buf_rsrc(g.nbr, (uint32_t)g.n_bricks * 27u * 4u);
  n.valid = valid;
#pragma unroll
  for (int it = 0; it < HaloNbr<BX, BY, BZ>::NIT; ++it) {
    n.pw[it] = kPlan<BX, BY, BZ>.w[it * BVOL + threadIdx.x];
    const int kind = (n.pw[it] >> 24) & 3;
    // every lane loads (out of range when it names no neighbour): no branch
    // around the load, so the loads that follow are issued without waiting
    const int raw = __builtin_bit_cast(
        int, buf_ld<float>(rN, (valid && kind) ? b * 27u + ((n.pw[it] >> 10) & 31) : OOB));
    n.nb[it] = it == 0 ? (int)b : (kind ? raw : -1);
  }
}

template <typename T, int BX, int BY, int BZ>
struct HaloRegs {
  static constexpr int NIT = HaloPlan<BX, BY, BZ>::NIT;
  uint32_t pw[NIT], q[NIT], lab[NIT];
  T v[NIT], d[NIT][3], dself[3];
};

// mylab: the lane's own label in brick n (unused: the own slot is read as a
// neighbour slot is, its label staged with it).
// value(q) gives the tile value of slot q (q = OOB for cells without a
// brick: it must then read out of range only); halo_load stages V itself.
template <typename T, int BX, int BY, int BZ, typename F>
__device__ __forceinline__ void halo_load_fn(HaloRegs<T, BX, BY, BZ>& h,
                                             const HaloNbr<BX, BY, BZ>& n, const hp_bricks& g,
                                             uint8_t mylab, F&& value) {
  constexpr int NIT = HaloPlan<BX, BY, BZ>::NIT, BVOL = BX * BY * BZ;
  const uint32_t n_slots = (uint32_t)g.n_bricks * BVOL, ld = n_slots;  // D6 row stride
  const buf_t rD = buf_rsrc(g.D6, 6u * ld * (uint32_t)sizeof(T));
  const buf_t rL = buf_rsrc(g.label, n_slots);
#pragma unroll
  for (int it = 0; it < NIT; ++it) {
    h.pw[it] = n.pw[it];
    // the brick's own slots are read unconditionally (padding slots hold V0, D = 0
    // and label 0xFF, which the staging masks), so these loads wait for nothing
    const bool ok = n.valid && (it == 0 || n.nb[it] >= 0);
    h.q[it] = ok ? (uint32_t)n.nb[it] * BVOL + ((n.pw[it] >> 15) & 511) : OOB;
  }
  {  // the brick's own slots first: their addresses need no neighbour-table entry
    const uint32_t q = h.q[0];
    h.lab[0] = buf_ld_u8(rL, q);
    h.v[0] = value(q);
#pragma unroll
    for (int k = 0; k < 3; ++k) h.d[0][k] = buf_ld<T>(rD, q == OOB ? OOB : k * ld + q);
#pragma unroll
    for (int k = 0; k < 3; ++k) h.dself[k] = buf_ld<T>(rD, q == OOB ? OOB : (3 + k) * ld + q);
  }
#pragma unroll
  for (int it = 1; it < NIT; ++it) {
    const uint32_t q = h.q[it];
    h.lab[it] = buf_ld_u8(rL, q);
    h.v[it] = value(q);
    {
      const uint32_t c0 = (h.pw[it] >> 26) & 7;
      const bool face = ((h.pw[it] >> 24) & 3) == 2 && q != OOB;
      const uint32_t c1 = c0 == 2 ? 4u : 3u, c2 = c0 == 0 ? 4u : 5u;
      h.d[it][0] = buf_ld<T>(rD, q == OOB ? OOB : c0 * ld + q);
      h.d[it][1] = buf_ld<T>(rD, face ? c1 * ld + q : OOB);  // face layers: the two
      h.d[it][2] = buf_ld<T>(rD, face ? c2 * ld + q : OOB);  // off-diagonals of its planes
    }
  }
}

template <typename T, int BX, int BY, int BZ>
__device__ __forceinline__ void halo_load(HaloRegs<T, BX, BY, BZ>& h, const HaloNbr<BX, BY, BZ>& n,
                                          const hp_bricks& g, const T* __restrict__ V,
                                          uint8_t mylab) {
  const buf_t rV = buf_rsrc(V, (uint32_t)g.n_bricks * (BX * BY * BZ) * (uint32_t)sizeof(T));
  halo_load_fn<T, BX, BY, BZ>(h, n, g, mylab, [&](uint32_t q) { return buf_ld<T>(rV, q); });
}

// Registers -> LDS tile; lanes without a tile cell write the spare cell N.
// The caller places the barriers.
template <typename T, int BX, int BY, int BZ>
__device__ __forceinline__ void halo_store(Tile<T, BX, BY, BZ>& t,
                                           const HaloRegs<T, BX, BY, BZ>& h) {
  typedef Tile<T, BX, BY, BZ> TT;
#pragma unroll
  for (int it = 0; it < HaloPlan<BX, BY, BZ>::NIT; ++it) {
    const int kind = (h.pw[it] >> 24) & 3;
    const int i = kind ? (int)(h.pw[it] & 1023) : TT::N;
    t.L[i] = h.q[it] == OOB ? (uint8_t)0xFF : (uint8_t)h.lab[it];
    t.V[i] = h.v[it];
    if (it == 0) {
#pragma unroll
      for (int k = 0; k < 3; ++k) t.D[k][i] = h.d[0][k];
#pragma unroll
      for (int k = 0; k < 3; ++k) t.D[3 + k][i] = h.dself[k];
    } else {
      const int c0 = (h.pw[it] >> 26) & 7;
      const int i12 = kind == 2 ? i : TT::N;
      t.D[c0 == 2 ? 4 : 3][i12] = h.d[it][1];
      t.D[c0 == 0 ? 4 : 5][i12] = h.d[it][2];
      t.D[c0][i] = h.d[it][0];
    }
  }
}

// Whole staging of brick b for the one-brick-per-workgroup kernels.
template <typename T, int BX, int BY, int BZ>
__device__ __forceinline__ void stage_cells(Tile<T, BX, BY, BZ>& t, const hp_bricks& g,
                                            const T* __restrict__ V, uint32_t b, uint8_t mylab) {
  HaloNbr<BX, BY, BZ> n;
  halo_nbr<BX, BY, BZ>(n, g, b);
  HaloRegs<T, BX, BY, BZ> h;
  halo_load<T, BX, BY, BZ>(h, n, g, V, mylab);
  halo_store<T, BX, BY, BZ>(t, h);
  __syncthreads();
}

// Sum over the 18 neighbours of w_ab*(V_b - V_a)/h, in units of D (mS/mm)*mV.
// Terms of absent neighbours (label 0xFF) are skipped.
template <typename T, int BX, int BY, int BZ>
__device__ __forceinline__ T stencil19(const Tile<T, BX, BY, BZ>& t, int c) {
  typedef Tile<T, BX, BY, BZ> TT;
  const T Vc = t.V[c];
  T acc = T(0);
  // face neighbours: harmonic mean of the axis component (aniso.py:996-1004)
  const int fo[3] = {TT::SX, TT::SY, 1};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T Da = t.D[a][c];
#pragma unroll
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      const int nb = c + sgn * fo[a];
      if (t.L[nb] != 0xFF) {
        const T Db = t.D[a][nb];
        const T s = Da + Db;
        const T dh = T(2) * Da * Db * M<T>::rcp(s > T(1e-30) ? s : T(1e-30));
        acc += dh * (t.V[nb] - Vc);
      }
    }
  }
  // edge-diagonal neighbours: +-0.5*mean(D_ij) (aniso.py:1006-1029); in
  // units of D*mV/h this is +-0.25*(D_ij,a + D_ij,b)*(V_b - V_a)
  const int po[3][2] = {{TT::SX, TT::SY}, {TT::SX, 1}, {TT::SY, 1}};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T Da = t.D[3 + q][c];
    const int o1 = po[q][0], o2 = po[q][1];
    const int co[2] = {o1 + o2, -(o1 + o2)};
    const int ct[2] = {o1 - o2, -(o1 - o2)};
    T sum = T(0);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int nb = c + co[k];
      if (t.L[nb] != 0xFF) sum += (Da + t.D[3 + q][nb]) * (t.V[nb] - Vc);
      const int nm = c + ct[k];
      if (t.L[nm] != 0xFF) sum -= (Da + t.D[3 + q][nm]) * (t.V[nm] - Vc);
    }
    acc += T(0.25) * sum;
  }
  return acc;
}

// halo_store with every cell that holds no tissue (label 0xFF, or no brick)
// marked by V = NaN, so that a stencil can test presence on the value it
// reads anyway instead of reading the label (stencil19_wsum_nan).
template <typename T, int BX, int BY, int BZ>
__device__ __forceinline__ void halo_store_nan(Tile<T, BX, BY, BZ>& t,
                                               const HaloRegs<T, BX, BY, BZ>& h) {
  typedef Tile<T, BX, BY, BZ> TT;
  const T qnan = __builtin_nan("");
#pragma unroll
  for (int it = 0; it < HaloPlan<BX, BY, BZ>::NIT; ++it) {
    const int kind = (h.pw[it] >> 24) & 3;
    const int i = kind ? (int)(h.pw[it] & 1023) : TT::N;
    const bool none = h.q[it] == OOB || (uint8_t)h.lab[it] == 0xFF;
    t.L[i] = none ? (uint8_t)0xFF : (uint8_t)h.lab[it];
    t.V[i] = none ? qnan : h.v[it];
    if (it == 0) {
#pragma unroll
      for (int k = 0; k < 3; ++k) t.D[k][i] = h.d[0][k];
#pragma unroll
      for (int k = 0; k < 3; ++k) t.D[3 + k][i] = h.dself[k];
    } else {
      const int c0 = (h.pw[it] >> 26) & 7;
      const int i12 = kind == 2 ? i : TT::N;
      t.D[c0 == 2 ? 4 : 3][i12] = h.d[it][1];
      t.D[c0 == 0 ? 4 : 5][i12] = h.d[it][2];
      t.D[c0][i] = h.d[it][0];
    }
  }
}

// stencil19_wsum on a tile staged by halo_store_nan: an absent neighbour is a
// NaN value (the comparison v == v is false for it), no label reads.  Every
// LDS value is read unconditionally and an absent neighbour's term is
// selected away (no branch per neighbour, so the reads are issued together);
// the present neighbours' terms are added in the same order as before.
template <typename T, int BX, int BY, int BZ>
__device__ __forceinline__ T stencil19_wsum_nan(const Tile<T, BX, BY, BZ>& t, int c, T& wsum) {
  typedef Tile<T, BX, BY, BZ> TT;
  const T Vc = t.V[c];
  T acc = T(0), ws = T(0);
  const int fo[3] = {TT::SX, TT::SY, 1};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T Da = t.D[a][c];
#pragma unroll
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      const int nb = c + sgn * fo[a];
      const T vn = t.V[nb];
      const T Db = t.D[a][nb];
      const T s = Da + Db;
      const T dh = T(2) * Da * Db * M<T>::rcp(s > T(1e-30) ? s : T(1e-30));
      const bool on = vn == vn;
      acc = on ? acc + dh * (vn - Vc) : acc;
      ws = on ? ws + dh : ws;
    }
  }
  const int po[3][2] = {{TT::SX, TT::SY}, {TT::SX, 1}, {TT::SY, 1}};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T Da = t.D[3 + q][c];
    const int o1 = po[q][0], o2 = po[q][1];
    const int co[2] = {o1 + o2, -(o1 + o2)};
    const int ct[2] = {o1 - o2, -(o1 - o2)};
    T sum = T(0), wq = T(0);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int nb = c + co[k];
      const T vb = t.V[nb];
      const T wb = Da + t.D[3 + q][nb];
      const bool onb = vb == vb;
      sum = onb ? sum + wb * (vb - Vc) : sum;
      wq = onb ? wq + wb : wq;
      const int nm = c + ct[k];
      const T vm = t.V[nm];
      const T wm = Da + t.D[3 + q][nm];
      const bool onm = vm == vm;
      sum = onm ? sum - wm * (vm - Vc) : sum;
      wq = onm ? wq - wm : wq;
    }
    acc += T(0.25) * sum;
    ws += T(0.25) * wq;
  }
  wsum = ws;
  return acc;
}

// Voxel of a brick that lane `tid` computes after staging.  Default: the
// slot order (z fastest).  8x4x4 bricks: each half-wave takes one y-row
// (all x, all z): with the tile's x-stride of 36 = 4 (mod 32) its 32 LDS
// addresses fall into 32 different banks, where the slot order leaves six
// banks with two lanes each.
template <int BX, int BY, int BZ>
__device__ __forceinline__ void compute_voxel(int tid, int& lx, int& ly, int& lz) {
  if (BX == 8 && BY == 4 && BZ == 4) {
    ly = tid >> 5;
    lx = (tid >> 2) & 7;
    lz = tid & 3;
  } else {
    lz = tid % BZ;
    ly = (tid / BZ) % BY;
    lx = tid / (BY * BZ);
  }
}

// stencil19 together with the sum of the 18 weights that enter it (the
// negated diagonal of the operator row, in the same units): the reference's
// diagonal is minus the sum of the signed weights of the node's edges
// (purkinje_myocardium_anisotropic_solver.py:713-717).
template <typename T, int BX, int BY, int BZ>
__device__ __forceinline__ T stencil19_wsum(const Tile<T, BX, BY, BZ>& t, int c, T& wsum) {
  typedef Tile<T, BX, BY, BZ> TT;
  const T Vc = t.V[c];
  T acc = T(0), ws = T(0);
  const int fo[3] = {TT::SX, TT::SY, 1};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T Da = t.D[a][c];
#pragma unroll
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      const int nb = c + sgn * fo[a];
      if (t.L[nb] != 0xFF) {
        const T Db = t.D[a][nb];
        const T s = Da + Db;
        const T dh = T(2) * Da * Db * M<T>::rcp(s > T(1e-30) ? s : T(1e-30));
        acc += dh * (t.V[nb] - Vc);
        ws += dh;
      }
    }
  }
  const int po[3][2] = {{TT::SX, TT::SY}, {TT::SX, 1}, {TT::SY, 1}};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const T Da = t.D[3 + q][c];
    const int o1 = po[q][0], o2 = po[q][1];
    const int co[2] = {o1 + o2, -(o1 + o2)};
    const int ct[2] = {o1 - o2, -(o1 - o2)};
    T sum = T(0), wq = T(0);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int nb = c + co[k];
      if (t.L[nb] != 0xFF) {
        const T w = Da + t.D[3 + q][nb];
        sum += w * (t.V[nb] - Vc);
        wq += w;
      }
      const int nm = c + ct[k];
      if (t.L[nm] != 0xFF) {
        const T w = Da + t.D[3 + q][nm];
        sum -= w * (t.V[nm] - Vc);
        wq -= w;
      }
    }
    acc += T(0.25) * sum;
    ws += T(0.25) * wq;
  }
  wsum = ws;
  return acc;
}

}  // namespace hp
