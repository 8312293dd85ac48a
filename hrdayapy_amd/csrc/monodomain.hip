// Monodomain time step on the brick layout: fused TTP06 reaction + 19-point
// anisotropic diffusion + stimulus + clamp + activation detection, the
// Purkinje cable step, and the PMJ exchange.
//
// Scheme (one call per 0.02 ms step, all terms evaluated at V^n):
//   V^{n+1} = clamp(V^n - dt*I_ion(V^n, s^n) + (dt/Cm)*(I_stim + (L V^n)), -150, 80)
//   s^{n+1} = TTP06 Rush-Larsen / forward-Euler update from (V^n, s^n)
// followed by the exact PMJ exchange on V^{n+1}.  This is the fast form; the
// reference's own scheme (Crank-Nicolson diffusion after the reaction,
// purkinje_myocardium_anisotropic_solver.py:1746-1791) is hp_cn_step in
// cn.hip.  At dt = 0.02 ms the explicit operator is stable (dt*|diag L|/Cm <
// 0.2 for the shipped conductivities); it differs from the CN scheme by the
// schemes' truncation error (upstroke timing).
//
// Memory plan per workgroup (one brick of BX*BY*BZ voxels, one lane each):
// the (BX+2)(BY+2)(BZ+2) halo tile of V, label and the six D components is
// staged in LDS once; each lane then reads its 18 neighbours from LDS and
// streams its own 18 state variables from HBM (coalesced SoA).
#include "brick_stencil.h"
#include "ttp06_cell.h"

namespace hp {

// do_reaction 1 (TTP06 every voxel) and 2 (TTP06 where rate_ext is NaN, else
// dV/dt = rate_ext); do_reaction 0 runs diff_step_kernel instead.
// FP32: at least six waves per SIMD (at most 80 VGPRs), so that the state and
// halo loads of several bricks are in flight together.
template <typename T, int BX, int BY, int BZ>
__device__ __forceinline__ void myo_brick(Tile<T, BX, BY, BZ>& t, const hp_bricks& g,
                                          const T* __restrict__ V_in, T* __restrict__ V_out,
                                          T* __restrict__ state, int64_t ld64,
                                          const hp_mono_params& p, uint32_t b) {
  typedef Tile<T, BX, BY, BZ> TT;
  const uint32_t ld = (uint32_t)ld64;
  const int tid = threadIdx.x;
  const uint32_t slot = b * TT::BVOL + tid;
  // every-voxel reaction: the 18 state loads of a tissue voxel are issued
  // before the halo staging loads, so that both are in flight together
  const bool pre = p.do_reaction == 1;
  const uint8_t mylab = g.label[slot];
  // compact state entry of this lane (padding lanes: out of range, so their
  // loads read 0 without memory traffic and their stores are dropped)
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t sidx = state_index_wave(g.groups, (b * TT::BVOL >> 6) + wave);
  T s[N_STATE];
  const uint32_t row_bytes = ld * (uint32_t)sizeof(T);
  if (pre) {
#pragma unroll
    for (int k = 0; k < N_STATE; ++k)
      s[k] = buf_ld<T>(buf_rsrc(state + k * (size_t)ld, row_bytes), sidx);
  }
  stage_cells<T, BX, BY, BZ>(t, g, V_in, b, mylab);
  const int lz = tid % BZ, ly = (tid / BZ) % BY, lx = tid / (BY * BZ);
  const int c = ((lx + 1) * TT::TY + (ly + 1)) * TT::TZ + (lz + 1);
  const uint8_t lab = t.L[c];
  if (lab == 0xFF) return;
  const T Vc = t.V[c];
  const T lv = T(p.dscale) * stencil19<T, BX, BY, BZ>(t, c);
  const T dt = T(p.dt);
  T dvdt = T(0);  // reaction part of dV/dt
  bool full = p.do_reaction == 1;
  if (p.do_reaction == 2) {  // surrogate form: NaN rate marks a fallback voxel
    const T r = ld32((const T*)p.rate_ext, slot);
    full = r != r;
    if (!full) dvdt = r;
  }
  if (full) {
    const int ct = lab & 3;
    T gks = T(kGKs[ct]), gto = T(kGto[ct]);
    if (p.inst_params && g.inst) {
      const T* ip = (const T*)p.inst_params + 4 * (int)g.inst[slot];
      gks *= ip[0];
      gto *= ip[1];
    }
    if (!pre) {
#pragma unroll
      for (int k = 0; k < N_STATE; ++k)
        s[k] = buf_ld<T>(buf_rsrc(state + k * (size_t)ld, row_bytes), sidx);
    }
    constexpr bool exact = sizeof(T) == 8;
    dvdt = -ttp06_cell<T, exact>(Vc, s, dt, gks, gto, ct != 2, T(p.ki_base), T(p.nai_base));
#pragma unroll
    for (int k = 0; k < N_STATE; ++k) buf_st(buf_rsrc(state + k * (size_t)ld, row_bytes), sidx, s[k]);
  }
  int grp = (lab >> 2) & 31;
  if (grp && g.inst) grp = ((int)g.inst[slot] << 3) | (grp & 7);  // per-instance groups
  const T stim = grp ? ((const T*)p.group_amp)[grp] : T(0);
  T Vn = Vc + dt * dvdt + dt * T(p.inv_cm) * (stim + lv);
  Vn = M<T>::mn(M<T>::mx(Vn, T(p.vmin)), T(p.vmax));
  buf_st(buf_rsrc(V_out, (uint32_t)g.n_bricks * TT::BVOL * (uint32_t)sizeof(T)), slot, Vn);
  // PMJ target voxels (label bit 7) are tested after the PMJ exchange instead
  if (p.at && !(lab & 0x80)) {
    const T thr = T(p.at_thr);
    if (Vc < thr && Vn >= thr) {
      T* at = (T*)p.at;
      const T a = ld32((const T*)at, slot);
      if (a != a) {
        const int64_t step = *p.step_ctr;
        st32(at, slot, (T)((double)(step - 1) * p.dt) + dt * (thr - Vc) / (Vn - Vc));
      }
    }
  }
}

template <typename T, int BX, int BY, int BZ>
__global__ void __launch_bounds__(BX * BY * BZ)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 1))) myo_step_kernel(
    hp_bricks g, const T* __restrict__ V_in, T* __restrict__ V_out, T* __restrict__ state,
    int64_t ld64, hp_mono_params p) {
  __shared__ Tile<T, BX, BY, BZ> t;
  myo_brick<T, BX, BY, BZ>(t, g, V_in, V_out, state, ld64, p, xcd_remap(blockIdx.x, gridDim.x));
}

// Multirate substep (do_reaction = 0): diffusion plus a given reaction rate
// (dV/dt = rate_ext, or 0 without it), the same update as myo_step_kernel
// without the cell model.  Persistent over bricks: the halo loads of the next
// brick are issued before the current brick is stored and computed, and the
// neighbour rows of the brick after that are fetched one iteration earlier,
// so the two dependent memory round trips of a brick overlap the work on the
// previous one.  Each XCD (blockIdx & 7) works through a contiguous run of
// bricks (compact in Morton order); the blocks of one XCD take its bricks
// round robin.
template <typename T>
__device__ __forceinline__ void cable_node(const hp_cable& c, const T* __restrict__ V_in,
                                           T* __restrict__ V_out, T* __restrict__ state,
                                           int64_t ld, const hp_mono_params& p, int64_t i);

// The first ncb workgroups (a multiple of 8, so the XCD of every brick
// workgroup is unchanged) step the Purkinje cable instead (full TTP06 with
// its own parameters), as in myo_cable_kernel.
template <typename T, int BX, int BY, int BZ, bool kInst>
__global__ void __launch_bounds__(BX * BY * BZ) diff_step_kernel(
    hp_bricks g, const T* __restrict__ V_in, T* __restrict__ V_out, hp_mono_params p,
    hp_cable cb, const T* __restrict__ Vpk_in, T* __restrict__ Vpk_out, T* __restrict__ state_pk,
    int64_t ld_pk, hp_mono_params p_pk, uint32_t ncb) {
  typedef Tile<T, BX, BY, BZ> TT;
  typedef HaloNbr<BX, BY, BZ> HN;
  typedef HaloRegs<T, BX, BY, BZ> HR;
  __shared__ TT t;
  if (blockIdx.x < ncb) {
    const int64_t i = (int64_t)blockIdx.x * (BX * BY * BZ) + threadIdx.x;
    if (i < cb.n) cable_node<T>(cb, Vpk_in, Vpk_out, state_pk, ld_pk, p_pk, i);
    return;
  }
  const uint32_t bid = blockIdx.x - ncb;
  const uint32_t NB = (uint32_t)g.n_bricks;
  const uint32_t xcd = bid & 7, G = gridDim.x - ncb;
  const uint32_t m = (G - xcd + 7) >> 3;  // blocks with index = xcd (mod 8)
  const uint32_t base = NB >> 3, rem = NB & 7;
  const uint32_t lo = xcd * base + (xcd < rem ? xcd : rem);
  const uint32_t hi = lo + base + (xcd < rem ? 1u : 0u);
  uint32_t b = lo + (bid >> 3);
  if (b >= hi) return;  // whole workgroup: uniform
  const int tid = threadIdx.x;
  const int lz = tid % BZ, ly = (tid / BZ) % BY, lx = tid / (BY * BZ);
  const int c = ((lx + 1) * TT::TY + (ly + 1)) * TT::TZ + (lz + 1);
  const uint32_t n_slots = NB * TT::BVOL;
  const buf_t rR = buf_rsrc(p.rate_ext, p.rate_ext ? n_slots * (uint32_t)sizeof(T) : 0u);
  const buf_t rG = buf_rsrc(p.group_amp, (uint32_t)(MAX_GROUPS * sizeof(T)));
  const T dt = T(p.dt), inv_cm = T(p.inv_cm), dscale = T(p.dscale);
  // prologue: brick b loaded, neighbour rows of the next brick requested.
  // Every load of an iteration is issued unconditionally (out of range past
  // the last brick), so the wait counts stay exact: memory returns in issue
  // order, and a load issued after the prefetch, or a path with fewer loads,
  // would make the compute wait for the whole prefetch.
  const buf_t rLab = buf_rsrc(g.label, n_slots);
  uint8_t mylab = g.label[b * TT::BVOL + tid];
  HN n;
  halo_nbr<BX, BY, BZ>(n, g, b);
  HR h;
  halo_load<T, BX, BY, BZ>(h, n, g, V_in, mylab);
  T r = buf_ld<T>(rR, mylab != 0xFF ? b * TT::BVOL + tid : OOB);
  uint32_t bn = b + m;
  uint8_t labn = (uint8_t)buf_ld_u8(rLab, bn < hi ? bn * TT::BVOL + tid : OOB);
  HN nn;
  halo_nbr<BX, BY, BZ>(nn, g, bn, bn < hi);
  for (;;) {
    // this brick's stimulus amplitude, before the next brick's loads
    const uint32_t slot = b * TT::BVOL + tid;
    const uint32_t lab0 = h.q[0] == OOB ? 0xFFu : h.lab[0];
    int grp = lab0 != 0xFF ? (int)((lab0 >> 2) & 31) : 0;
    if (kInst && grp) grp = ((int)g.inst[slot] << 3) | (grp & 7);  // per-instance groups
    const T stim = buf_ld<T>(rG, (uint32_t)grp);
    // next brick: halo and rate; the brick after: label and neighbour rows
    const bool more = bn < hi;
    HR hn;
    halo_load<T, BX, BY, BZ>(hn, nn, g, V_in, labn);
    const T rn = buf_ld<T>(rR, more && labn != 0xFF ? bn * TT::BVOL + tid : OOB);
    const uint32_t bnn = bn + m;
    const bool more2 = more && bnn < hi;
    const uint8_t labnn = (uint8_t)buf_ld_u8(rLab, more2 ? bnn * TT::BVOL + tid : OOB);
    HN nnn;
    halo_nbr<BX, BY, BZ>(nnn, g, bnn, more2);
    __syncthreads();  // every lane is done with the previous tile
    halo_store<T, BX, BY, BZ>(t, h);
    __syncthreads();
    if (lab0 != 0xFF) {
      const T Vc = t.V[c];
      const T lv = dscale * stencil19<T, BX, BY, BZ>(t, c);
      T Vn = Vc + dt * r + dt * inv_cm * (stim + lv);
      Vn = M<T>::mn(M<T>::mx(Vn, T(p.vmin)), T(p.vmax));
      st32(V_out, slot, Vn);
      const T thr = T(p.at_thr);
      if (p.at && !(lab0 & 0x80) && Vc < thr && Vn >= thr) {
        T* at = (T*)p.at;
        if (ld32((const T*)at, slot) != ld32((const T*)at, slot)) {
          const int64_t step = *p.step_ctr;
          st32(at, slot, (T)((double)(step - 1) * p.dt) + dt * (thr - Vc) / (Vn - Vc));
        }
      }
    }
    if (!more) break;
    b = bn;
    bn = bnn;
    h = hn;
    r = rn;
    nn = nnn;
    labn = labnn;
  }
}

template <typename T, int BX, int BY, int BZ>
__global__ void __launch_bounds__(BX * BY * BZ) stencil19_kernel(hp_bricks g,
                                                                 const T* __restrict__ V,
                                                                 T* __restrict__ out,
                                                                 double dscale) {
  typedef Tile<T, BX, BY, BZ> TT;
  __shared__ TT t;
  const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t slot = b * TT::BVOL + threadIdx.x;
  const uint8_t mylab = g.label[slot];
  stage_cells<T, BX, BY, BZ>(t, g, V, b, mylab);
  const int tid = threadIdx.x;
  const int lz = tid % BZ, ly = (tid / BZ) % BY, lx = tid / (BY * BZ);
  const int c = ((lx + 1) * TT::TY + (ly + 1)) * TT::TZ + (lz + 1);
  st32(out, slot, t.L[c] == 0xFF ? T(0) : T(dscale) * stencil19<T, BX, BY, BZ>(t, c));
}

template <typename T, template <typename, int, int, int> class K, typename... A>
static int launch_bricks(const hp_bricks* g, const char* what, hipStream_t st, A... args) {
  if (!g || g->n_bricks < 0 || g->n_bricks > 0x7fffffff) {
    set_error("%s: bad brick grid", what);
    return -1;
  }
  if (g->n_bricks == 0) return 0;
  // the kernels address every per-slot array (and the six D rows together)
  // with 32-bit byte offsets
  if ((uint64_t)g->n_bricks * g->bx * g->by * g->bz * 6 * sizeof(T) >= (1ull << 32)) {
    set_error("%s: %lld bricks exceed the 32-bit offset range of one launch", what,
              (long long)g->n_bricks);
    return -1;
  }
  const dim3 grid((unsigned)g->n_bricks);
  const int key = g->bx * 100 + g->by * 10 + g->bz;
  switch (key) {
    case 444: hipLaunchKernelGGL((K<T, 4, 4, 4>::fn), grid, dim3(64), 0, st, *g, args...); break;
    case 844: hipLaunchKernelGGL((K<T, 8, 4, 4>::fn), grid, dim3(128), 0, st, *g, args...); break;
    case 884: hipLaunchKernelGGL((K<T, 8, 8, 4>::fn), grid, dim3(256), 0, st, *g, args...); break;
    case 888: hipLaunchKernelGGL((K<T, 8, 8, 8>::fn), grid, dim3(512), 0, st, *g, args...); break;
    default:
      set_error("%s: unsupported brick shape %dx%dx%d", what, g->bx, g->by, g->bz);
      return -1;
  }
  return check_launch(what);
}

template <typename T, int BX, int BY, int BZ>
struct MyoK {
  static constexpr auto fn = myo_step_kernel<T, BX, BY, BZ>;
};
// The persistent diffusion kernel: 16 one-wave workgroups per CU.
static int compute_units() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n <= 0)
      n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

template <typename T>
static int launch_diff(const hp_bricks* g, const T* V_in, T* V_out, const hp_mono_params* p,
                       hipStream_t st, const hp_cable* c = nullptr, const T* Vpk_in = nullptr,
                       T* Vpk_out = nullptr, T* state_pk = nullptr, int64_t ld_pk = 0,
                       const hp_mono_params* p_pk = nullptr) {
  const hp_cable c0{};
  const hp_mono_params p0{};
  if (!c) c = &c0;
  if (!p_pk) p_pk = &p0;
  const int64_t bvol = (int64_t)g->bx * g->by * g->bz;
  const int64_t ncb = (((c->n + bvol - 1) / bvol) + 7) / 8 * 8;
  if (g->n_bricks == 0 && ncb == 0) return 0;
  const int64_t want = (int64_t)compute_units() * 16;
  const dim3 grid((unsigned)((g->n_bricks < want ? g->n_bricks : want) + ncb));
  const int key = g->bx * 100 + g->by * 10 + g->bz;
  switch (key) {
#define HP_DIFF(BX, BY, BZ)                                                                  \
  if (g->inst)                                                                               \
    hipLaunchKernelGGL((diff_step_kernel<T, BX, BY, BZ, true>), grid, dim3(BX * BY * BZ), 0, \
                       st, *g, V_in, V_out, *p, *c, Vpk_in, Vpk_out, state_pk, ld_pk, *p_pk,   \
                       (uint32_t)ncb);                                                       \
  else                                                                                       \
    hipLaunchKernelGGL((diff_step_kernel<T, BX, BY, BZ, false>), grid, dim3(BX * BY * BZ),   \
                       0, st, *g, V_in, V_out, *p, *c, Vpk_in, Vpk_out, state_pk, ld_pk,      \
                       *p_pk, (uint32_t)ncb)
    case 444: HP_DIFF(4, 4, 4); break;
    case 844: HP_DIFF(8, 4, 4); break;
    case 884: HP_DIFF(8, 8, 4); break;
    case 888: HP_DIFF(8, 8, 8); break;
#undef HP_DIFF
    default:
      set_error("hp_myo_step: unsupported brick shape %dx%dx%d", g->bx, g->by, g->bz);
      return -1;
  }
  return check_launch("hp_myo_step");
}
template <typename T, int BX, int BY, int BZ>
struct StencilK {
  static constexpr auto fn = stencil19_kernel<T, BX, BY, BZ>;
};

// ---------------------------------------------------------------------------
// Purkinje cable
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void cable_node(const hp_cable& c, const T* __restrict__ V_in,
                                           T* __restrict__ V_out, T* __restrict__ state,
                                           int64_t ld, const hp_mono_params& p, int64_t i) {
  const T Vc = V_in[i];
  T acc = T(0);
  for (int e = c.indptr[i]; e < c.indptr[i + 1]; ++e) acc += V_in[c.indices[e]] - Vc;
  const T lv = T(p.dscale) * T(c.w) * acc;
  const T dt = T(p.dt);
  const uint8_t lab = c.label[i];
  T dvdt = T(0);
  if (p.do_reaction) {
    const int ct = lab & 3;
    const uint32_t row_bytes = (uint32_t)ld * (uint32_t)sizeof(T);
    T s[N_STATE];
#pragma unroll
    for (int k = 0; k < N_STATE; ++k)
      s[k] = buf_ld<T>(buf_rsrc(state + k * (size_t)ld, row_bytes), (uint32_t)i);
    constexpr bool exact = sizeof(T) == 8;
    dvdt = -ttp06_cell<T, exact>(Vc, s, dt, T(kGKs[ct]), T(kGto[ct]), ct != 2, T(p.ki_base),
                                 T(p.nai_base));
#pragma unroll
    for (int k = 0; k < N_STATE; ++k)
      buf_st(buf_rsrc(state + k * (size_t)ld, row_bytes), (uint32_t)i, s[k]);
  }
  const int grp = (lab >> 2) & 31;
  const T stim = grp ? ((const T*)p.group_amp)[grp] : T(0);
  T Vn = Vc + dt * dvdt + dt * T(p.inv_cm) * (stim + lv);
  V_out[i] = M<T>::mn(M<T>::mx(Vn, T(p.vmin)), T(p.vmax));
}

template <typename T>
__global__ void __launch_bounds__(256) cable_step_kernel(hp_cable c, const T* __restrict__ V_in,
                                                         T* __restrict__ V_out,
                                                         T* __restrict__ state, int64_t ld,
                                                         hp_mono_params p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < c.n) cable_node<T>(c, V_in, V_out, state, ld, p, i);
}

// The coupled step in one launch: the first ncb workgroups (a multiple of 8,
// so the XCD of every brick workgroup is unchanged) update the cable nodes,
// the others the bricks.  Within a step the two read V^n and write V^{n+1} of
// their own domains only; the PMJ exchange that couples them follows in the
// tail kernel.  Running the cable here removes a kernel and two of the three
// kernel boundaries of a step.
template <typename T, int BX, int BY, int BZ>
__global__ void __launch_bounds__(BX * BY * BZ)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 6 : 1))) myo_cable_kernel(
    hp_bricks g, const T* __restrict__ V_in, T* __restrict__ V_out, T* __restrict__ state,
    int64_t ld64, hp_mono_params p, hp_cable c, const T* __restrict__ Vpk_in,
    T* __restrict__ Vpk_out, T* __restrict__ state_pk, int64_t ld_pk, hp_mono_params p_pk,
    uint32_t ncb) {
  __shared__ Tile<T, BX, BY, BZ> t;
  if (blockIdx.x < ncb) {
    const int64_t i = (int64_t)blockIdx.x * (BX * BY * BZ) + threadIdx.x;
    if (i < c.n) cable_node<T>(c, Vpk_in, Vpk_out, state_pk, ld_pk, p_pk, i);
    return;
  }
  myo_brick<T, BX, BY, BZ>(t, g, V_in, V_out, state, ld64, p,
                           xcd_remap(blockIdx.x - ncb, gridDim.x - ncb));
}

// ---------------------------------------------------------------------------
// PMJ exchange: exact integration of the junction conductances over dt.
// ---------------------------------------------------------------------------
// First activation of a PMJ target voxel between V^n (start of step) and the
// final V^{n+1} of step `step`
template <typename T>
__device__ __forceinline__ void pmj_at_update(const hp_pmj_at& at, int s, T v, int64_t step) {
  const T v0 = ((const T*)at.V_prev)[s];
  const T thr = T(at.at_thr);
  T* a = (T*)at.at;
  if (v0 < thr && v >= thr && a[s] != a[s])
    a[s] = (T)((double)(step - 1) * at.dt) + T(at.dt) * (thr - v0) / (v - v0);
}

// One junction's exchange.  kDirect (patches exclusive): the junction applies
// its increments to the patch voxels itself; else it writes them to scratch
// for pmj_target.  Both round the increment to T and add it to V_myo in T.
template <typename T, bool kDirect = false>
__device__ __forceinline__ void pmj_junction(const hp_pmj& j, int64_t q, T* __restrict__ V_pk,
                                             T* __restrict__ V_myo, T* __restrict__ scratch,
                                             double dt, double vmin, double vmax,
                                             const hp_pmj_at* at = nullptr, int64_t step = 0) {
  const int e0 = j.ptr[q], e1 = j.ptr[q + 1];
  const int n = e1 - e0;
  const int p = j.pk[q];
  const double Vp = (double)V_pk[p];
  double mean = 0.0;
  for (int e = e0; e < e1; ++e) mean += (double)V_myo[j.patch[e]];
  double wp, r1, r2;  // Cp/(Cp+Cq), relaxation factors of the mean and the deviations
  if (j.coef) {
    const double* k = j.coef + 4 * q;
    wp = k[0];
    r1 = k[1];
    r2 = k[2];
    mean *= k[3];
  } else {
    mean /= n;
    const double Cp = j.ai_p * j.cm, Cq = n * j.ai_m * j.cm;
    wp = Cp / (Cp + Cq);
    r1 = exp(-j.c_pmj * dt * (1.0 / Cp + 1.0 / Cq));
    r2 = exp(-(j.c_pmj / n) * dt / (j.ai_m * j.cm));
  }
  const double vbar = wp * Vp + (1.0 - wp) * mean;
  const double vp_new = vbar + (Vp - vbar) * r1;
  const double mean_new = vbar + (mean - vbar) * r1;
  for (int e = e0; e < e1; ++e) {
    const int s = j.patch[e];
    const T vt = V_myo[s];
    const double vm = (double)vt;
    const T delta = (T)(mean_new + (vm - mean) * r2 - vm);
    if (kDirect) {
      const T v = M<T>::mn(M<T>::mx(vt + delta, T(vmin)), T(vmax));
      V_myo[s] = v;
      if (at && at->at) pmj_at_update<T>(*at, s, v, step);
    } else {
      scratch[e] = delta;
    }
  }
  V_pk[p] = (T)fmin(fmax(vp_new, vmin), vmax);
}

template <typename T>
__device__ __forceinline__ void pmj_target(const hp_pmj& j, int64_t q, T* __restrict__ V_myo,
                                           const T* __restrict__ scratch, double vmin, double vmax,
                                           const hp_pmj_at& at, int64_t step) {
  T acc = T(0);
  for (int k = j.tgt_ptr[q]; k < j.tgt_ptr[q + 1]; ++k) acc += scratch[j.tgt_entry[k]];
  const int s = j.tgt_slot[q];
  T v = V_myo[s] + acc;
  v = M<T>::mn(M<T>::mx(v, T(vmin)), T(vmax));
  V_myo[s] = v;
  if (at.at) pmj_at_update<T>(at, s, v, step);
}

template <typename T>
__global__ void pmj_junction_kernel(hp_pmj j, T* __restrict__ V_pk, T* __restrict__ V_myo,
                                    T* __restrict__ scratch, double dt, double vmin,
                                    double vmax) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < j.n_junctions) pmj_junction<T>(j, q, V_pk, V_myo, scratch, dt, vmin, vmax);
}

template <typename T>
__global__ void pmj_target_kernel(hp_pmj j, T* __restrict__ V_myo, const T* __restrict__ scratch,
                                  double vmin, double vmax, hp_pmj_at at) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < j.n_targets) pmj_target<T>(j, q, V_myo, scratch, vmin, vmax, at, at.at ? *at.step_ctr : 0);
}

// ---------------------------------------------------------------------------
// Stimulus schedule and step counter
// ---------------------------------------------------------------------------
// Floor division of doubles as Python's float `//` defines it (the reference
// computes the pulse number with it): the remainder from fmod, the quotient of
// the remainder-free dividend, rounded to the nearest integer when it lies
// more than 0.5 above its floor.  Plain floor(a/b) can differ from it where
// a/b rounds up to an integer.
__device__ __forceinline__ double py_floordiv(double a, double b) {
#pragma clang fp contract(off)
  double mod = fmod(a, b);
  double div = (a - mod) / b;
  if (mod != 0.0 && ((b < 0.0) != (mod < 0.0))) div -= 1.0;
  if (div == 0.0) return copysign(0.0, a / b);
  double fl = floor(div);
  if (div - fl > 0.5) fl += 1.0;
  return fl;
}

// _stim_active (purkinje_myocardium_anisotropic_solver.py:1290-1301), each
// operation rounded on its own as Python evaluates it (no fused multiply-add:
// a phase equal to the pulse duration must compare as it does there)
__device__ __forceinline__ bool site_active(const double* sp, double t) {
#pragma clang fp contract(off)
  const double onset = sp[0], dur = sp[1], period = sp[2], npulse = sp[3];
  if (t < onset) return false;
  if (period <= 0.0) return t < onset + dur;
  const double k = py_floordiv(t - onset, period);
  if (npulse >= 0.0 && k >= npulse) return false;
  return (t - onset) - k * period < dur;
}

// Amplitude of stimulus group g at time t: sum over the sites named by the
// group's bitmask, relative to its site base (gs[2g] = base, gs[2g+1] = mask).
__device__ __forceinline__ double group_amplitude(int g, const double* __restrict__ sp,
                                                  const uint32_t* __restrict__ gs, double t) {
  const uint32_t base = gs[2 * g], mask = gs[2 * g + 1];
  double a = 0.0;
  for (int b = 0; b < 32; ++b)
    if ((mask >> b) & 1u) {
      const double* q = sp + 5 * (int64_t)(base + b);
      if (site_active(q, t)) a += q[4];
    }
  return a;
}

constexpr int STIM_THREADS = 1024;

template <typename T>
__global__ void __launch_bounds__(STIM_THREADS) stim_kernel(const double* __restrict__ sp,
                                                            const uint32_t* __restrict__ gs,
                                                            int n_groups, int64_t* step_ctr,
                                                            double dt, T* __restrict__ amp,
                                                            int advance) {
  const int64_t step = *step_ctr + advance;
  __syncthreads();
  const double t = (double)step * dt;
  for (int g = threadIdx.x; g < n_groups; g += STIM_THREADS)
    amp[g] = g == 0 ? T(0) : (T)group_amplitude(g, sp, gs, t);
  if (advance && threadIdx.x == 0) *step_ctr = step;
}

// One workgroup finishes the step: the PMJ junctions and, concurrently in the
// last TAIL_STIM_THREADS lanes, the stimulus amplitudes of the next step; then
// the PMJ targets and the step counter (replaces three latency-bound launches;
// used while the junction count fits one block well).
constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_STIM_THREADS = 128;
constexpr int64_t TAIL_MAX_JUNCTIONS = 65536;

template <typename T>
__global__ void __launch_bounds__(TAIL_THREADS) step_tail_kernel(
    hp_pmj j, int has_pmj, T* __restrict__ V_pk, T* __restrict__ V_myo, T* __restrict__ scratch,
    double dt, double vmin, double vmax, hp_pmj_at at, const double* __restrict__ sp,
    const uint32_t* __restrict__ gs, int n_groups, int64_t* step_ctr, T* __restrict__ amp) {
  constexpr int JT = TAIL_THREADS - TAIL_STIM_THREADS;
  const int tid = threadIdx.x;
  // every lane reads the counter before the barrier below; lane 0 advances it
  // only after that barrier
  const int64_t step = *step_ctr;
  if (tid < JT) {
    if (has_pmj && j.exclusive)
      for (int64_t q = tid; q < j.n_junctions; q += JT)
        pmj_junction<T, true>(j, q, V_pk, V_myo, scratch, dt, vmin, vmax, &at, step);
    else if (has_pmj)
      for (int64_t q = tid; q < j.n_junctions; q += JT)
        pmj_junction<T>(j, q, V_pk, V_myo, scratch, dt, vmin, vmax);
  } else {
    const double t = (double)(step + 1) * dt;
    for (int g = tid - JT; g < n_groups; g += TAIL_STIM_THREADS)
      amp[g] = g == 0 ? T(0) : (T)group_amplitude(g, sp, gs, t);
  }
  __syncthreads();
  if (has_pmj && !j.exclusive)
    for (int64_t q = tid; q < j.n_targets; q += TAIL_THREADS)
      pmj_target<T>(j, q, V_myo, scratch, vmin, vmax, at, step);
  if (tid == 0) *step_ctr = step + 1;
}

template <typename T>
__global__ void gather_kernel(const T* __restrict__ src, const int32_t* __restrict__ idx,
                              T* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

template <typename T>
static int myo_step(const hp_bricks* g, const T* V_in, T* V_out, T* state, int64_t ld,
                    const hp_mono_params* p, void* stream) {
  if (!p || !p->step_ctr || !p->group_amp) {
    set_error("hp_myo_step: params, step_ctr and group_amp are required");
    return -1;
  }
  if (p->do_reaction < 0 || p->do_reaction > 2 || (p->do_reaction == 2 && !p->rate_ext)) {
    set_error("hp_myo_step: do_reaction must be 0, 1 or 2 (2 needs rate_ext)");
    return -1;
  }
  if (!g) {
    set_error("hp_myo_step: brick grid required");
    return -1;
  }
  if (p->do_reaction != 0 &&
      (!g->groups || !state || ld < 0 || (uint64_t)ld * sizeof(T) >= (1ull << 32))) {
    set_error("hp_myo_step: the reaction needs groups and state (ld*sizeof(T) < 2^32)");
    return -1;
  }
  if (p->do_reaction == 0) {
    // launch_bricks checks the grid for the other kernels
    if (g->n_bricks < 0 || g->n_bricks > 0x7fffffff ||
        (uint64_t)g->n_bricks * g->bx * g->by * g->bz * 6 * sizeof(T) >= (1ull << 32)) {
      set_error("hp_myo_step: %lld bricks exceed the 32-bit offset range of one launch",
                (long long)g->n_bricks);
      return -1;
    }
    return launch_diff<T>(g, V_in, V_out, p, (hipStream_t)stream);
  }
  return launch_bricks<T, MyoK>(g, "hp_myo_step", (hipStream_t)stream, V_in, V_out, state, ld,
                                *p);
}

template <typename T>
static int myo_cable_step(const hp_bricks* g, const T* V_in, T* V_out, T* state, int64_t ld,
                          const hp_mono_params* p, const hp_cable* c, const T* Vpk_in,
                          T* Vpk_out, T* state_pk, int64_t ld_pk, const hp_mono_params* p_pk,
                          void* stream) {
  if (!g || !p || !p->step_ctr || !p->group_amp || p->do_reaction < 0 || p->do_reaction > 2 ||
      (p->do_reaction == 2 && !p->rate_ext) ||
      (p->do_reaction != 0 &&
       (!g->groups || !state || ld < 0 || (uint64_t)ld * sizeof(T) >= (1ull << 32)))) {
    set_error("hp_myo_cable_step: myocardium arguments as hp_myo_step");
    return -1;
  }
  if (!c || !p_pk || c->n < 0 || !p_pk->group_amp || ld_pk < c->n ||
      (uint64_t)ld_pk * sizeof(T) >= (1ull << 32) || (c->n > 0 && (!Vpk_in || !Vpk_out || !state_pk))) {
    set_error("hp_myo_cable_step: cable arguments as hp_cable_step");
    return -1;
  }
  const int64_t bvol = (int64_t)g->bx * g->by * g->bz;
  if (g->n_bricks < 0 || (uint64_t)g->n_bricks * bvol * 6 * sizeof(T) >= (1ull << 32)) {
    set_error("hp_myo_cable_step: %lld bricks exceed the 32-bit offset range of one launch",
              (long long)g->n_bricks);
    return -1;
  }
  // the diffusion-only substep: the persistent kernel, with the cable blocks
  if (p->do_reaction == 0)
    return launch_diff<T>(g, V_in, V_out, p, (hipStream_t)stream, c, Vpk_in, Vpk_out, state_pk,
                          ld_pk, p_pk);
  const int64_t ncb = (((c->n + bvol - 1) / bvol) + 7) / 8 * 8;
  if (g->n_bricks + ncb == 0) return 0;
  if (g->n_bricks + ncb > 0x7fffffff) {
    set_error("hp_myo_cable_step: grid too large");
    return -1;
  }
  const dim3 grid((unsigned)(g->n_bricks + ncb));
  hipStream_t st = (hipStream_t)stream;
  const int key = g->bx * 100 + g->by * 10 + g->bz;
  switch (key) {
#define HP_MC(BX, BY, BZ)                                                                      \
  hipLaunchKernelGGL((myo_cable_kernel<T, BX, BY, BZ>), grid, dim3(BX * BY * BZ), 0, st, *g, \
                     V_in, V_out, state, ld, *p, *c, Vpk_in, Vpk_out, state_pk, ld_pk, *p_pk, \
                     (uint32_t)ncb)
    case 444: HP_MC(4, 4, 4); break;
    case 844: HP_MC(8, 4, 4); break;
    case 884: HP_MC(8, 8, 4); break;
    case 888: HP_MC(8, 8, 8); break;
#undef HP_MC
    default:
      set_error("hp_myo_cable_step: unsupported brick shape %dx%dx%d", g->bx, g->by, g->bz);
      return -1;
  }
  return check_launch("hp_myo_cable_step");
}

// ---------------------------------------------------------------------------
// Multirate substep: the full TTP06 step for the compacted list of fallback
// voxels, writing rate[slot] = -I_ion(V^n) for the diffusion kernel
// (do_reaction = 0) that follows.  Grid-stride over *count entries.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) fallback_react_kernel(
    hp_bricks g, const float* __restrict__ V, float* __restrict__ state, int64_t ld,
    float* __restrict__ rate, const int32_t* __restrict__ list, const int32_t* __restrict__ count,
    hp_mono_params p) {
  const int n = *count;
  const uint32_t row_bytes = (uint32_t)ld * 4u;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t slot = (uint32_t)list[i];
    const uint32_t si = state_index(g.groups, slot);
    const int ct = g.label[slot] & 3;
    float gks = (float)kGKs[ct], gto = (float)kGto[ct];
    if (p.inst_params && g.inst) {
      const float* ip = (const float*)p.inst_params + 4 * (int)g.inst[slot];
      gks *= ip[0];
      gto *= ip[1];
    }
    float s[N_STATE];
#pragma unroll
    for (int k = 0; k < N_STATE; ++k)
      s[k] = buf_ld<float>(buf_rsrc(state + k * (size_t)ld, row_bytes), si);
    const float I = ttp06_cell<float, false>(ld32(V, slot), s, (float)p.dt, gks, gto, ct != 2,
                                             (float)p.ki_base, (float)p.nai_base);
#pragma unroll
    for (int k = 0; k < N_STATE; ++k) buf_st(buf_rsrc(state + k * (size_t)ld, row_bytes), si, s[k]);
    st32(rate, slot, -I);
  }
}

template <typename T>
static int cable_step(const hp_cable* c, const T* V_in, T* V_out, T* state, int64_t ld,
                      const hp_mono_params* p, void* stream) {
  if (!c || !p || c->n < 0 || !p->group_amp || ld < c->n ||
      (uint64_t)ld * sizeof(T) >= (1ull << 32)) {
    set_error("hp_cable_step: bad arguments (ld must cover the nodes and fit 32-bit offsets)");
    return -1;
  }
  if (c->n == 0) return 0;
  hipLaunchKernelGGL(cable_step_kernel<T>, dim3(ceil_div(c->n, 256)), dim3(256), 0,
                     (hipStream_t)stream, *c, V_in, V_out, state, ld, *p);
  return check_launch("hp_cable_step");
}

template <typename T>
static int pmj_couple(const hp_pmj* j, T* V_pk, T* V_myo, T* scratch, double dt, double vmin,
                      double vmax, const hp_pmj_at* at, void* stream) {
  if (!j || j->n_junctions < 0) {
    set_error("hp_pmj_couple: bad arguments");
    return -1;
  }
  if (j->n_junctions == 0 || j->c_pmj == 0.0) return 0;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(pmj_junction_kernel<T>, dim3(ceil_div(j->n_junctions, 128)), dim3(128), 0,
                     st, *j, V_pk, V_myo, scratch, dt, vmin, vmax);
  int rc = check_launch("hp_pmj_couple/junction");
  if (rc) return rc;
  if (j->n_targets > 0)
    hipLaunchKernelGGL(pmj_target_kernel<T>, dim3(ceil_div(j->n_targets, 128)), dim3(128), 0, st,
                       *j, V_myo, scratch, vmin, vmax, at ? *at : hp_pmj_at{});
  return check_launch("hp_pmj_couple/target");
}

template <typename T>
static int stim(const double* sp, int n_sites, const uint32_t* gs, int n_groups,
                const int64_t* ctr, double dt, T* amp, int advance, void* stream) {
  if (n_sites < 0 || n_groups < 1 || n_groups > MAX_GROUPS || !ctr || !amp || !gs) {
    set_error("hp_stim: 1..%d groups required (got %d)", MAX_GROUPS, n_groups);
    return -1;
  }
  hipLaunchKernelGGL(stim_kernel<T>, dim3(1), dim3(STIM_THREADS), 0, (hipStream_t)stream, sp, gs,
                     n_groups, (int64_t*)ctr, dt, amp, advance);
  return check_launch("hp_stim");
}

template <typename T>
static int step_tail(const hp_pmj* j, T* V_pk, T* V_myo, T* scratch, double dt, double vmin,
                     double vmax, const hp_pmj_at* at, const double* sp, int n_sites,
                     const uint32_t* gs, int n_groups, int64_t* ctr, T* amp, void* stream) {
  if (n_sites < 0 || n_groups < 1 || n_groups > MAX_GROUPS || !ctr || !amp || !gs) {
    set_error("hp_step_tail: 1..%d groups required (got %d)", MAX_GROUPS, n_groups);
    return -1;
  }
  const bool has = j && j->n_junctions > 0 && j->c_pmj != 0.0;
  if (has && j->n_junctions > TAIL_MAX_JUNCTIONS) {  // large trees: multi-block path
    int rc = pmj_couple<T>(j, V_pk, V_myo, scratch, dt, vmin, vmax, at, stream);
    if (rc) return rc;
    return stim<T>(sp, n_sites, gs, n_groups, ctr, dt, amp, 1, stream);
  }
  hipLaunchKernelGGL(step_tail_kernel<T>, dim3(1), dim3(TAIL_THREADS), 0, (hipStream_t)stream,
                     has ? *j : hp_pmj{}, has ? 1 : 0, V_pk, V_myo, scratch, dt, vmin, vmax,
                     at ? *at : hp_pmj_at{}, sp, gs, n_groups, ctr, amp);
  return check_launch("hp_step_tail");
}

// Plain streaming copy (16-byte vectors, grid-stride): the reference point for
// the HBM bandwidth and traffic counters of the other kernels.
__global__ void __launch_bounds__(256) stream_copy_kernel(const float4* __restrict__ src,
                                                          float4* __restrict__ dst, int64_t n4) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

template <typename T>
static int gather(const T* src, const int32_t* idx, T* out, int64_t n, void* stream) {
  if (n < 0 || (n > 0 && (!src || !idx || !out))) {
    set_error("hp_gather: bad arguments (n=%lld)", (long long)n);
    return -1;
  }
  if (n == 0) return 0;
  hipLaunchKernelGGL(gather_kernel<T>, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     src, idx, out, n);
  return check_launch("hp_gather");
}

}  // namespace hp

using namespace hp;

extern "C" {
int hp_myo_step_f32(const hp_bricks* g, const float* V_in, float* V_out, float* state, int64_t ld,
                    const hp_mono_params* p, void* stream) {
  return myo_step<float>(g, V_in, V_out, state, ld, p, stream);
}
int hp_myo_step_f64(const hp_bricks* g, const double* V_in, double* V_out, double* state,
                    int64_t ld, const hp_mono_params* p, void* stream) {
  return myo_step<double>(g, V_in, V_out, state, ld, p, stream);
}
int hp_myo_cable_step_f32(const hp_bricks* g, const float* V_in, float* V_out, float* state,
                          int64_t ld, const hp_mono_params* p, const hp_cable* c,
                          const float* Vpk_in, float* Vpk_out, float* state_pk, int64_t ld_pk,
                          const hp_mono_params* p_pk, void* stream) {
  return myo_cable_step<float>(g, V_in, V_out, state, ld, p, c, Vpk_in, Vpk_out, state_pk, ld_pk,
                               p_pk, stream);
}
int hp_myo_cable_step_f64(const hp_bricks* g, const double* V_in, double* V_out, double* state,
                          int64_t ld, const hp_mono_params* p, const hp_cable* c,
                          const double* Vpk_in, double* Vpk_out, double* state_pk, int64_t ld_pk,
                          const hp_mono_params* p_pk, void* stream) {
  return myo_cable_step<double>(g, V_in, V_out, state, ld, p, c, Vpk_in, Vpk_out, state_pk,
                                ld_pk, p_pk, stream);
}
int hp_fallback_react_f32(const hp_bricks* g, const float* V, float* state, int64_t ld,
                          float* rate, const int32_t* list, const int32_t* count,
                          int64_t max_count, const hp_mono_params* p, void* stream) {
  if (!g || !g->groups || !p || !list || !count || !rate || max_count < 0 || ld < 0 ||
      (uint64_t)ld * 4 >= (1ull << 32)) {
    set_error("hp_fallback_react: bad arguments (groups required, ld*4 < 2^32)");
    return -1;
  }
  if (max_count == 0) return 0;
  const int blocks = (int)(max_count < 1024 * 256 ? ceil_div(max_count, 256) : 1024);
  hipLaunchKernelGGL(fallback_react_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, *g,
                     V, state, ld, rate, list, count, *p);
  return check_launch("hp_fallback_react");
}
int hp_stencil19_apply_f32(const hp_bricks* g, const float* V, float* out, double dscale,
                           void* stream) {
  return launch_bricks<float, StencilK>(g, "hp_stencil19_apply", (hipStream_t)stream, V, out,
                                        dscale);
}
int hp_stencil19_apply_f64(const hp_bricks* g, const double* V, double* out, double dscale,
                           void* stream) {
  return launch_bricks<double, StencilK>(g, "hp_stencil19_apply", (hipStream_t)stream, V, out,
                                         dscale);
}
int hp_cable_step_f32(const hp_cable* c, const float* V_in, float* V_out, float* state,
                      int64_t ld, const hp_mono_params* p, void* stream) {
  return cable_step<float>(c, V_in, V_out, state, ld, p, stream);
}
int hp_cable_step_f64(const hp_cable* c, const double* V_in, double* V_out, double* state,
                      int64_t ld, const hp_mono_params* p, void* stream) {
  return cable_step<double>(c, V_in, V_out, state, ld, p, stream);
}
int hp_pmj_couple_f32(const hp_pmj* j, float* V_pk, float* V_myo, float* scratch, double dt,
                      double vmin, double vmax, const hp_pmj_at* at, void* stream) {
  return pmj_couple<float>(j, V_pk, V_myo, scratch, dt, vmin, vmax, at, stream);
}
int hp_pmj_couple_f64(const hp_pmj* j, double* V_pk, double* V_myo, double* scratch, double dt,
                      double vmin, double vmax, const hp_pmj_at* at, void* stream) {
  return pmj_couple<double>(j, V_pk, V_myo, scratch, dt, vmin, vmax, at, stream);
}
int hp_stim_eval_f32(const double* sp, int32_t ns, const uint32_t* gs, int32_t ng,
                     const int64_t* ctr, double dt, float* amp, void* stream) {
  return stim<float>(sp, ns, gs, ng, ctr, dt, amp, 0, stream);
}
int hp_stim_eval_f64(const double* sp, int32_t ns, const uint32_t* gs, int32_t ng,
                     const int64_t* ctr, double dt, double* amp, void* stream) {
  return stim<double>(sp, ns, gs, ng, ctr, dt, amp, 0, stream);
}
int hp_step_advance_f32(const double* sp, int32_t ns, const uint32_t* gs, int32_t ng,
                        int64_t* ctr, double dt, float* amp, void* stream) {
  return stim<float>(sp, ns, gs, ng, ctr, dt, amp, 1, stream);
}
int hp_step_advance_f64(const double* sp, int32_t ns, const uint32_t* gs, int32_t ng,
                        int64_t* ctr, double dt, double* amp, void* stream) {
  return stim<double>(sp, ns, gs, ng, ctr, dt, amp, 1, stream);
}
int hp_step_tail_f32(const hp_pmj* j, float* V_pk, float* V_myo, float* scratch, double dt,
                     double vmin, double vmax, const hp_pmj_at* at, const double* sp,
                     int32_t ns, const uint32_t* gs, int32_t ng, int64_t* ctr, float* amp,
                     void* stream) {
  return step_tail<float>(j, V_pk, V_myo, scratch, dt, vmin, vmax, at, sp, ns, gs, ng, ctr, amp,
                          stream);
}
int hp_step_tail_f64(const hp_pmj* j, double* V_pk, double* V_myo, double* scratch, double dt,
                     double vmin, double vmax, const hp_pmj_at* at, const double* sp,
                     int32_t ns, const uint32_t* gs, int32_t ng, int64_t* ctr, double* amp,
                     void* stream) {
  return step_tail<double>(j, V_pk, V_myo, scratch, dt, vmin, vmax, at, sp, ns, gs, ng, ctr, amp,
                           stream);
}
int hp_gather_f32(const float* src, const int32_t* idx, float* out, int64_t n, void* stream) {
  return gather<float>(src, idx, out, n, stream);
}
int hp_stream_copy_f32(const float* src, float* dst, int64_t n, void* stream) {
  if (n < 0 || (n & 3) || (((uintptr_t)src | (uintptr_t)dst) & 15)) {
    set_error("hp_stream_copy_f32: n must be a multiple of 4 and pointers 16-byte aligned");
    return -1;
  }
  if (n == 0) return 0;
  const int64_t n4 = n / 4;
  // 1024 workgroups (4 per CU) streaming long grid-stride loops measured the
  // highest copy rate on MI355X (5.6 TB/s read+write for 1 GiB)
  const int blocks = (int)(n4 / 256 < 1024 ? ceil_div(n4, 256) : 1024);
  hipLaunchKernelGGL(stream_copy_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                     (const float4*)src, (float4*)dst, n4);
  return check_launch("hp_stream_copy_f32");
}
int hp_gather_f64(const double* src, const int32_t* idx, double* out, int64_t n, void* stream) {
  return gather<double>(src, idx, out, n, stream);
}
}
