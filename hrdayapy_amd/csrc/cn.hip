// Crank-Nicolson diffusion of the coupled Purkinje + myocardium monodomain:
// the reference's time discretisation (purkinje_myocardium_anisotropic_solver.py
// :757-849 and the loop :1746-1791), matrix-free on the brick layout.
//
// One time step (Lie splitting, as the reference):
//   V*  = V^n - dt*I_ion(V^n, s^n) + (dt/Cm)*I_stim            cn_react_kernel
//   b   = (Cm/dt) V* + (1-theta) L V*
//   A x = b,  A = (Cm/dt) I - theta L,  from x0 = V*            cn_sweep_kernel
//   V^{n+1} = clamp(x, -150, 80)                               cn_final_kernel
// L = D^-1 K of the whole system: the 19-point anisotropic stencil on the
// myocardial rows (weights computed from D in-kernel), the cable edges w_pp on
// the Purkinje rows, and the PMJ couplings c_pmj/|patch| on both (:606-746).
//
// The linear solve replaces the reference's Jacobi-preconditioned BiCGSTAB
// (:786-849) by Jacobi sweeps x <- x + (b - A x)/diag(A) with the same stopping
// rule, ||b - A x|| <= tol*||b||, bounded by a sweep count.  A is strongly
// diagonally dominant at the reference's step sizes (the Jacobi contraction
// is ~0.05-0.2 on the tissue and ~0.4 per sweep on a junction pair), and each
// sweep is one pass over V and D with no reduction other than ||r||: 2-4
// sweeps meet tol = 1e-5 where BiCGSTAB needs 1-2 iterations of two products
// and five reductions each.  Both are approximations of the same CN solution
// within the tolerance.
//
// Each sweep evaluates the residual of its input iterate and writes the next
// iterate; FP64 block partials of ||r||^2 (and of ||b||^2 on the first sweep)
// are summed in block order by a one-block finish kernel, which sets the
// device-side `done` flag; later sweeps of the step return at once.  The
// residual of sweep k belongs to x_{k-1}, so the iterate returned is the one
// after the sweep whose input met the rule (one sweep past it).
//
// Junction rows: the PMJ conductance is orders of magnitude above the cable's
// own (c_pmj/Ai_P ~ 100/ms against ~0.1/ms), so a plain Jacobi sweep moves a
// junction node and its patch against each other with a contraction of ~0.5
// per sweep.  The finish kernel therefore re-solves each junction node's row
// after the sweep from the patch voxels' new values (and its cable
// neighbours' old ones): a Gauss-Seidel step on the stiff coupling, which
// halves the sweeps a step needs on trees with junctions, for no extra
// launch.
#include "brick_stencil.h"
#include "ttp06_cell.h"

namespace hp {

constexpr int CN_MAX_CABLE_BLOCKS = 512;
constexpr int64_t CN_SWEEP_BLOCKS = 8192;  // brick workgroups of a sweep at most
constexpr int CN_FIN_THREADS = 1024;

// scalars[]: 0 ||b||^2, 1 ||r||^2 of the last sweep, 2 done, 3 sweeps this step,
// 4 steps that stopped at max_sweeps, 5 largest ||r||/||b|| of those, 6 total
// sweeps, 7 steps solved, 16 the largest sweep count of any step, 17 steps
// begun with the warm start (capped at 2), 18 the largest sweep count since
// the caller last zeroed it (solver 0)
enum { SC_BB = 0, SC_RR, SC_DONE, SC_SWEEPS, SC_NCONV, SC_WORST, SC_TOTAL, SC_STEPS,
       SC_MAXIT = 16, SC_WARM = 17, SC_CMAX = 18 };

// Warm start of the sweeps (solver 0): x_0 = V* + e, e the increment x - V*
// of the last step extrapolated linearly from the last two steps (the last
// one alone after the first step, 0 before it).  The first sweep then forms
// b from V* as before and the residual of x_0 as
//   r_0 = b - A x_0 = L V* + theta L e - (Cm/dt) e,
// which is L V* exactly when e = 0.
template <typename T>
struct CnWarm {
  T* d;       // [n_slots] x - V* of the last step
  T* e;       // [n_slots] the extrapolated increment of this step
  T* d_pk;    // cable nodes
  T* e_pk;
};

__device__ __forceinline__ double cn_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block sums of (a, b) -> part[blockIdx], part[G + blockIdx] (NW waves per block).
template <int NW>
__device__ __forceinline__ void cn_block_partials(double a, double b, double* part) {
  __shared__ double sh[2][NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  a = cn_wave_sum(a);
  b = cn_wave_sum(b);
  if (lane == 0) {
    sh[0][wave] = a;
    sh[1][wave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      sa += sh[0][w];
      sb += sh[1][w];
    }
    part[blockIdx.x] = sa;
    part[gridDim.x + blockIdx.x] = sb;
  }
}

// g_Ks and g_to of a myocardial slot (phenotype, instance scales)
template <typename T>
__device__ __forceinline__ void cell_conductances(const hp_bricks& g, const hp_mono_params& p,
                                                  uint32_t slot, int ct, T& gks, T& gto) {
  gks = T(kGKs[ct]);
  gto = T(kGto[ct]);
  if (p.inst_params && g.inst) {
    const T* ip = (const T*)p.inst_params + 4 * (int)g.inst[slot];
    gks *= ip[0];
    gto *= ip[1];
  }
}

// ---------------------------------------------------------------------------
// Reaction: V* for every myocardial slot (one lane per slot, 256 per block, so
// each wave holds the 64 slots of one state group) and, in the first ncb
// blocks, for every cable node.  No clamp: the reference clamps after the
// diffusion solve only (:1791).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) cn_react_kernel(
    hp_bricks g, const T* __restrict__ V, T* __restrict__ Vs, T* __restrict__ state, int64_t ld,
    hp_mono_params p, hp_cable c, const T* __restrict__ Vpk, T* __restrict__ Vspk,
    T* __restrict__ state_pk, int64_t ld_pk, hp_mono_params p_pk, uint32_t ncb) {
  constexpr bool exact = sizeof(T) == 8;
  if (blockIdx.x < ncb) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= c.n) return;
    const uint8_t lab = c.label[i];
    const int ct = lab & 3;
    const uint32_t row_bytes = (uint32_t)ld_pk * (uint32_t)sizeof(T);
    T s[N_STATE];
#pragma unroll
    for (int k = 0; k < N_STATE; ++k)
      s[k] = buf_ld<T>(buf_rsrc(state_pk + k * (size_t)ld_pk, row_bytes), (uint32_t)i);
    const T Vc = Vpk[i];
    const T dt = T(p_pk.dt);
    const T I = ttp06_cell<T, exact>(Vc, s, dt, T(kGKs[ct]), T(kGto[ct]), ct != 2,
                                     T(p_pk.ki_base), T(p_pk.nai_base));
#pragma unroll
    for (int k = 0; k < N_STATE; ++k)
      buf_st(buf_rsrc(state_pk + k * (size_t)ld_pk, row_bytes), (uint32_t)i, s[k]);
    const int grp = (lab >> 2) & 31;
    const T stim = grp ? ((const T*)p_pk.group_amp)[grp] : T(0);
    Vspk[i] = Vc - dt * I + dt * T(p_pk.inv_cm) * stim;
    return;
  }
  uint32_t slot, sidx;
  if (p.cell_slot) {  // one lane per cell (compact state order)
    sidx = (blockIdx.x - ncb) * 256u + threadIdx.x;
    if (sidx >= (uint32_t)p.n_cells) return;
    slot = (uint32_t)p.cell_slot[sidx];
  } else {  // one lane per slot: waves of 64 slots, padding lanes idle
    slot = (blockIdx.x - ncb) * 256u + threadIdx.x;
    const uint32_t n_slots = (uint32_t)g.n_bricks * (uint32_t)(g.bx * g.by * g.bz);
    if (slot >= n_slots) return;
    sidx = state_index_wave(g.groups, slot >> 6);
  }
  const uint8_t lab = g.label[slot];
  if (lab == 0xFF) return;
  const T Vc = ld32(V, slot);
  const T dt = T(p.dt);
  T I;
  // do_reaction 2 (multirate forms): dV/dt = rate_ext where it is a number,
  // the full cell step where it is NaN (a fallback voxel)
  const T rate = p.do_reaction == 2 ? ld32((const T*)p.rate_ext, slot) : T(0);
  if (p.do_reaction == 2 && rate == rate) {
    I = -rate;
  } else {
    const uint32_t row_bytes = (uint32_t)ld * (uint32_t)sizeof(T);
    T s[N_STATE];
#pragma unroll
    for (int k = 0; k < N_STATE; ++k)
      s[k] = buf_ld<T>(buf_rsrc(state + k * (size_t)ld, row_bytes), sidx);
    const int ct = lab & 3;
    T gks, gto;
    cell_conductances<T>(g, p, slot, ct, gks, gto);
    I = ttp06_cell<T, exact>(Vc, s, dt, gks, gto, ct != 2, T(p.ki_base), T(p.nai_base));
#pragma unroll
    for (int k = 0; k < N_STATE; ++k)
      buf_st(buf_rsrc(state + k * (size_t)ld, row_bytes), sidx, s[k]);
  }
  int grp = (lab >> 2) & 31;
  if (grp && g.inst) grp = ((int)g.inst[slot] << 3) | (grp & 7);  // per-instance groups
  const T stim = grp ? ((const T*)p.group_amp)[grp] : T(0);
  st32(Vs, slot, Vc - dt * I + dt * T(p.inv_cm) * stim);
}

// Index of `slot` in the ascending target list (the slot is known to be there).
__device__ __forceinline__ int cn_find_target(const int32_t* __restrict__ tgt, int n, int slot) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (tgt[mid] < slot) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One Jacobi update of a row: r = b - A x (first sweep: b is formed, and
// r = L x exactly), x_new = x + r / (Cm/dt - theta*L_ii).  lx = (L x)_i,
// ld = -L_ii.  Accumulates r^2 (and b^2) into rr (bb).
template <typename T, bool kFirst>
__device__ __forceinline__ T cn_row(T xc, T lx, T ld, T& bv, const hp_cn_params& cp, double& rr,
                                    double& bb, T ec = T(0), T le = T(0)) {
  const T cm_dt = T(cp.cm_dt), th = T(cp.theta);
  T r;
  if (kFirst) {
    bv = cm_dt * xc + (T(1) - th) * lx;
    r = lx + th * le - cm_dt * ec;  // the residual of x_0 = xc + ec (lx when ec = le = 0)
    bb += (double)bv * (double)bv;
    rr += (double)r * (double)r;
    return (xc + ec) + r / (cm_dt + th * ld);
  }
  r = bv - cm_dt * xc + th * lx;
  rr += (double)r * (double)r;
  return xc + r / (cm_dt + th * ld);
}

// stencil19_wsum_nan applied to V and, with the same weights, to a second
// field E of the tile (absent neighbours: NaN in V; branch-free as there)
template <typename T, int BX, int BY, int BZ>
__device__ __forceinline__ T stencil19_wsum2_nan(const Tile<T, BX, BY, BZ>& t,
                                                 const T* __restrict__ E, int c, T& wsum,
                                                 T& acc_e) {
  typedef Tile<T, BX, BY, BZ> TT;
  const T Vc = t.V[c], Ec = E[c];
  T acc = T(0), ws = T(0), ae = T(0);
  const int fo[3] = {TT::SX, TT::SY, 1};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const T Da = t.D[a][c];
#pragma unroll
    for (int sgn = -1; sgn <= 1; sgn += 2) {
      const int nb = c + sgn * fo[a];
      const T vn = t.V[nb];
      const T en = E[nb];
      const T Db = t.D[a][nb];
      const T s = Da + Db;
      const T dh = T(2) * Da * Db * M<T>::rcp(s > T(1e-30) ? s : T(1e-30));
      const bool on = vn == vn;
      acc = on ? acc + dh * (vn - Vc) : acc;
      ae = on ? ae + dh * (en - Ec) : ae;
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
    T sum = T(0), wq = T(0), se = T(0);
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int nb = c + co[k];
      const T vb = t.V[nb], eb = E[nb];
      const T wb = Da + t.D[3 + q][nb];
      const bool onb = vb == vb;
      sum = onb ? sum + wb * (vb - Vc) : sum;
      se = onb ? se + wb * (eb - Ec) : se;
      wq = onb ? wq + wb : wq;
      const int nm = c + ct[k];
      const T vm = t.V[nm], em = E[nm];
      const T wm = Da + t.D[3 + q][nm];
      const bool onm = vm == vm;
      sum = onm ? sum - wm * (vm - Vc) : sum;
      se = onm ? se - wm * (em - Ec) : se;
      wq = onm ? wq - wm : wq;
    }
    acc += T(0.25) * sum;
    ae += T(0.25) * se;
    ws += T(0.25) * wq;
  }
  wsum = ws;
  acc_e = ae;
  return acc;
}

// One sweep: bricks in XCD-aware order (each XCD takes a contiguous run of
// bricks, so halo re-reads stay in its L2), every myocardial row of a brick
// per workgroup iteration, and, in the first ncb workgroups, the cable rows
// (grid-stride).  Each workgroup writes its partial sums of r^2 (and
// b^2) to part[blockIdx] and part[gridDim + blockIdx].
template <typename T, int BX, int BY, int BZ, bool kFirst, bool kWarm = false>
__global__ void __launch_bounds__(BX * BY * BZ) cn_sweep_kernel(
    hp_bricks g, const T* __restrict__ x_in, T* __restrict__ x_out, T* __restrict__ bvec,
    hp_cable c, const T* __restrict__ xpk_in, T* __restrict__ xpk_out, T* __restrict__ bpk,
    hp_cn_pmj j, hp_cn_params cp, uint32_t ncb, CnWarm<T> wm) {
  static_assert(kFirst || !kWarm, "the warm start enters the first sweep only");
  constexpr int BVOL = BX * BY * BZ;
  typedef Tile<T, BX, BY, BZ> TT;
  __shared__ TT t;
  __shared__ T Et[kWarm ? TT::N + 1 : 1];  // the extrapolated increment e on the tile
  // block-uniform; a sweep outside a step (hp_cn_sweep) has no scalars
  if (!kFirst && cp.scalars && cp.scalars[SC_DONE] != 0.0) return;
  double rr = 0.0, bb = 0.0;
  const int tid = threadIdx.x;
  if (blockIdx.x < ncb) {
    const T dsw = T(cp.dscale_pk * c.w);
    for (int64_t i = (int64_t)blockIdx.x * BVOL + tid; i < c.n; i += (int64_t)ncb * BVOL) {
      const T xc = xpk_in[i];
      const T ec = kWarm ? wm.e_pk[i] : T(0);
      T acc = T(0), acce = T(0);
      const int e0 = c.indptr[i], e1 = c.indptr[i + 1];
      for (int e = e0; e < e1; ++e) {
        acc += xpk_in[c.indices[e]] - xc;
        if (kWarm) acce += wm.e_pk[c.indices[e]] - ec;
      }
      T lx = dsw * acc, ldg = dsw * T(e1 - e0), le = dsw * acce;
      if (j.node_ptr) {
        for (int e = j.node_ptr[i]; e < j.node_ptr[i + 1]; ++e) {
          const T w = T(j.node_w[e]);
          lx += w * (ld32(x_in, (uint32_t)j.node_slot[e]) - xc);
          if (kWarm) le += w * (ld32((const T*)wm.e, (uint32_t)j.node_slot[e]) - ec);
          ldg += w;
        }
      }
      T bv = kFirst ? T(0) : bpk[i];
      xpk_out[i] = cn_row<T, kFirst>(xc, lx, ldg, bv, cp, rr, bb, ec, le);
      if (kFirst) bpk[i] = bv;
    }
  } else {
   // each XCD (blockIdx & 7) takes a contiguous run of bricks, its
   // workgroups in turn; with at most CN_SWEEP_BLOCKS workgroups the sweeps
   // launched after convergence cost little
   const uint32_t bid = blockIdx.x - ncb;
   const uint32_t NB = (uint32_t)g.n_bricks;
   const uint32_t xcd = bid & 7, G = gridDim.x - ncb;
   const uint32_t m = (G - xcd + 7) >> 3;  // workgroups with index = xcd (mod 8)
   const uint32_t base = NB >> 3, rem = NB & 7;
   const uint32_t lo = xcd * base + (xcd < rem ? xcd : rem);
   const uint32_t hi = lo + base + (xcd < rem ? 1u : 0u);
   int vx, vy, vz;  // the voxel this lane computes (staging uses the slot order)
   compute_voxel<BX, BY, BZ>(tid, vx, vy, vz);
   const int cc = ((vx + 1) * TT::TY + (vy + 1)) * TT::TZ + (vz + 1);
   const uint32_t vloc = (uint32_t)((vx * BY + vy) * BZ + vz);
   for (uint32_t b = lo + (bid >> 3); b < hi; b += m) {
    const uint32_t slot = b * BVOL + vloc;
    // the right-hand side is read before the tile is staged, so its latency
    // overlaps the halo loads (its lanes without a cell read an unused value)
    T bv = kFirst ? T(0) : ld32((const T*)bvec, slot);
    __syncthreads();  // every lane is done with the previous tile
    {
      HaloNbr<BX, BY, BZ> hn;
      halo_nbr<BX, BY, BZ>(hn, g, b);
      HaloRegs<T, BX, BY, BZ> hr;
      halo_load<T, BX, BY, BZ>(hr, hn, g, x_in, g.label[b * BVOL + tid]);
      if (kWarm) {  // e at the same tile cells (0 for cells without a brick)
        constexpr int NIT = HaloPlan<BX, BY, BZ>::NIT;
        const buf_t rE = buf_rsrc(wm.e, NB * (uint32_t)BVOL * (uint32_t)sizeof(T));
        T ev[NIT];
#pragma unroll
        for (int it = 0; it < NIT; ++it) ev[it] = buf_ld<T>(rE, hr.q[it]);
#pragma unroll
        for (int it = 0; it < NIT; ++it) {
          const int kind = (hr.pw[it] >> 24) & 3;
          Et[kind ? (int)(hr.pw[it] & 1023) : TT::N] = ev[it];
        }
      }
      halo_store_nan<T, BX, BY, BZ>(t, hr);
      __syncthreads();
    }
    const uint8_t mylab = t.L[cc];
    if (mylab != 0xFF) {
      const T dscale = T(cp.dscale);
      const T xc = t.V[cc];
      T ws, acce = T(0);
      const T acc = kWarm ? stencil19_wsum2_nan<T, BX, BY, BZ>(t, Et, cc, ws, acce)
                          : stencil19_wsum_nan<T, BX, BY, BZ>(t, cc, ws);
      T lx = dscale * acc, ldg = dscale * ws, le = dscale * acce;
      const T ec = kWarm ? Et[cc] : T(0);
      if ((mylab & 0x80) && j.n_targets > 0) {  // PMJ target voxel: its junction couplings
        const int q = cn_find_target(j.tgt_slot, (int)j.n_targets, (int)slot);
        const int e1 = j.tgt_slot[q] == (int)slot ? j.tgt_ptr[q + 1] : 0;  // listed slots only
        for (int e = j.tgt_ptr[q]; e < e1; ++e) {
          const T w = T(j.tgt_w[e]);
          lx += w * (xpk_in[j.tgt_node[e]] - xc);
          if (kWarm) le += w * (wm.e_pk[j.tgt_node[e]] - ec);
          ldg += w;
        }
      }
      st32(x_out, slot, cn_row<T, kFirst>(xc, lx, ldg, bv, cp, rr, bb, ec, le));
      if (kFirst) st32(bvec, slot, bv);
    }
   }
  }
  cn_block_partials<(BVOL + 63) / 64>(rr, bb, cp.part);
}

// Sums the partials of the sweep before it (fixed block order) and applies
// the stopping rule; re-solves the junction nodes' rows (see the top of the
// file): x_p = (b_p + theta*(w_pp/Ai_P * sum of its cable neighbours' input
// values + sum_m w_pm * x_m new)) / (Cm/dt + theta * (-L_pp)).  Returns at
// once when the step's solve was already done before the sweep.
// With the warm start (epk_in: e of the cable nodes, first sweep only) the
// sweep's input on the cable is x_0 = V* + e.
template <typename T, bool kFirst>
__global__ void __launch_bounds__(CN_FIN_THREADS) cn_finish_kernel(
    hp_cn_params cp, int nb, hp_cable c, hp_cn_pmj j, const T* __restrict__ xpk_in,
    T* __restrict__ xpk_out, const T* __restrict__ x_out, const T* __restrict__ bpk,
    const T* __restrict__ epk_in, int warm) {
  double* sc = cp.scalars;
  if (!kFirst && sc[SC_DONE] != 0.0) return;
  // the scalars the stopping rule updates are read now (only this kernel's
  // last lines write them), so their latency overlaps the work below
  const double s_bb = sc[SC_BB], s_sweeps = sc[SC_SWEEPS], s_warm = sc[SC_WARM],
               s_total = sc[SC_TOTAL], s_steps = sc[SC_STEPS], s_maxit = sc[SC_MAXIT],
               s_cmax = sc[SC_CMAX], s_nconv = sc[SC_NCONV], s_worst = sc[SC_WORST];
  // the first batch of block partials is loaded before the junction rows'
  // dependent loads (independent of them), so both latencies overlap
  double xr0[8], xb0[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int b = u * CN_FIN_THREADS + (int)threadIdx.x;
    xr0[u] = b < nb ? cp.part[b] : 0.0;
    xb0[u] = (kFirst && b < nb) ? cp.part[nb + b] : 0.0;
  }
  if (j.n_jnodes > 0) {
    const T cm_dt = T(cp.cm_dt), th = T(cp.theta), dsw = T(cp.dscale_pk * c.w);
    for (int64_t q = threadIdx.x; q < j.n_jnodes; q += CN_FIN_THREADS) {
      const int pn = j.jnode[q];
      T nb_sum = T(0);
      const int e0 = c.indptr[pn], e1 = c.indptr[pn + 1];
      for (int e = e0; e < e1; ++e)
        nb_sum += epk_in ? xpk_in[c.indices[e]] + epk_in[c.indices[e]] : xpk_in[c.indices[e]];
      T m_sum = T(0), wsum = T(0);
      for (int e = j.node_ptr[pn]; e < j.node_ptr[pn + 1]; ++e) {
        const T w = T(j.node_w[e]);
        m_sum += w * ld32(x_out, (uint32_t)j.node_slot[e]);
        wsum += w;
      }
      const T diag = cm_dt + th * (dsw * T(e1 - e0) + wsum);
      xpk_out[pn] = (bpk[pn] + th * (dsw * nb_sum + m_sum)) / diag;
    }
  }
  __shared__ double sh[2][CN_FIN_THREADS / 64];
  double rr = 0.0, bb = 0.0;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    rr += xr0[u];
    bb += xb0[u];
  }
  // eight independent loads in flight per array before the adds
  for (int b0 = 8 * CN_FIN_THREADS; b0 < nb; b0 += 8 * CN_FIN_THREADS) {
    double xr[8], xb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int b = b0 + u * CN_FIN_THREADS + (int)threadIdx.x;
      xr[u] = b < nb ? cp.part[b] : 0.0;
      xb[u] = (kFirst && b < nb) ? cp.part[nb + b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      rr += xr[u];
      bb += xb[u];
    }
  }
  // per-thread strided sums, then a fixed-order combination: deterministic
  rr = cn_wave_sum(rr);
  bb = cn_wave_sum(bb);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sh[0][wave] = rr;
    sh[1][wave] = bb;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  rr = 0.0;
  bb = 0.0;
  for (int w = 0; w < CN_FIN_THREADS / 64; ++w) {
    rr += sh[0][w];
    bb += sh[1][w];
  }
  if (kFirst) {
    sc[SC_BB] = bb;
    if (warm && s_warm < 2.0) sc[SC_WARM] = s_warm + 1.0;  // steps begun warm, capped at 2
  } else {
    bb = s_bb;
  }
  const double sweeps = kFirst ? 1.0 : s_sweeps + 1.0;
  sc[SC_RR] = rr;
  sc[SC_SWEEPS] = sweeps;
  const bool conv = rr <= cp.tol * cp.tol * bb;  // also when ||b|| = 0 (then r = 0)
  const bool done = conv || sweeps >= (double)cp.max_sweeps;
  sc[SC_DONE] = done ? 1.0 : 0.0;
  if (done) {
    sc[SC_TOTAL] = s_total + sweeps;
    sc[SC_STEPS] = s_steps + 1.0;
    if (sweeps > s_maxit) sc[SC_MAXIT] = sweeps;
    if (sweeps > s_cmax) sc[SC_CMAX] = sweeps;
    if (!conv) {
      sc[SC_NCONV] = s_nconv + 1.0;
      const double rel = bb > 0.0 ? sqrt(rr / bb) : 0.0;
      if (rel > s_worst) sc[SC_WORST] = rel;
    }
  }
}

// V^{n+1} = clamp(x) from the buffer of the last sweep (odd sweep count: x,
// even: V_out itself; BiCGSTAB: always x), first activation between V^n and V^{n+1} (every
// myocardial slot, junction voxels included: the PMJ coupling is part of the
// solve); cable nodes in the first ncb blocks.
// With the warm start (wm.d set): d = x - V* of this step (before the clamp)
// and the next step's e = d + (d - d_last), or d after the first warm step.
template <typename T>
__device__ __forceinline__ void cn_warm_update(T* __restrict__ d, T* __restrict__ e, int64_t i,
                                               T v, T vs, bool linear) {
  const T dn = v - vs;
  const T dl = d[i];
  d[i] = dn;
  e[i] = linear ? dn + (dn - dl) : dn;
}

template <typename T>
__global__ void __launch_bounds__(256) cn_final_kernel(
    hp_bricks g, const T* __restrict__ V_in, const T* __restrict__ x, T* __restrict__ V_out,
    const T* __restrict__ xpk, T* __restrict__ Vpk_out, int64_t n_cable, hp_mono_params p,
    const double* __restrict__ sc, uint32_t ncb, bool by_parity, const T* __restrict__ Vs,
    const T* __restrict__ Vspk, CnWarm<T> wm) {
  const bool odd = !by_parity || ((int64_t)sc[SC_SWEEPS] & 1) != 0;
  const bool linear = wm.d && sc[SC_WARM] >= 2.0;
  const T vmin = T(p.vmin), vmax = T(p.vmax);
  if (blockIdx.x < ncb) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n_cable) {
      const T v = odd ? xpk[i] : Vpk_out[i];
      Vpk_out[i] = M<T>::mn(M<T>::mx(v, vmin), vmax);
      if (wm.d) cn_warm_update<T>(wm.d_pk, wm.e_pk, i, v, Vspk[i], linear);
    }
    return;
  }
  uint32_t slot = (blockIdx.x - ncb) * 256u + threadIdx.x;
  if (p.cell_slot) {  // one lane per cell
    if (slot >= (uint32_t)p.n_cells) return;
    slot = (uint32_t)p.cell_slot[slot];
  } else {
    const uint32_t n_slots = (uint32_t)g.n_bricks * (uint32_t)(g.bx * g.by * g.bz);
    if (slot >= n_slots || g.label[slot] == 0xFF) return;
  }
  const T v = odd ? ld32(x, slot) : ld32((const T*)V_out, slot);
  const T Vn = M<T>::mn(M<T>::mx(v, vmin), vmax);
  st32(V_out, slot, Vn);
  if (wm.d) cn_warm_update<T>(wm.d, wm.e, slot, v, ld32(Vs, slot), linear);
  if (p.at) {
    const T Vc = ld32(V_in, slot), thr = T(p.at_thr);
    if (Vc < thr && Vn >= thr) {
      T* at = (T*)p.at;
      if (ld32((const T*)at, slot) != ld32((const T*)at, slot)) {
        const int64_t step = *p.step_ctr;
        st32(at, slot, (T)((double)(step - 1) * p.dt) + T(p.dt) * (thr - Vc) / (Vn - Vc));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// BiCGSTAB (solver 1): the reference's Jacobi-preconditioned BiCGSTAB
// (:786-849) from x0 = V*, with its stopping rules and breakdown guards, on
// the same matrix-free rows.  Iteration k is three passes, each followed by a
// one-block finish kernel that sums the passes' FP64 block partials in block
// order and takes the iteration's scalar decisions on the device:
//   A(k): p = r + beta (p - omega v) (k = 0: p = r), v = A M^-1 p,  r^.v
//   S(k): s = r - alpha v, t = A M^-1 s,  s.s, t.t, t.s
//   X(k): x += alpha M^-1 p (+ omega M^-1 s), r = s - omega t,  r.r, r^.r
// The stencil passes form M^-1 p (M^-1 s) of every tile cell from r, p, v
// and M^-1 as they stage the halo, so no preconditioned vector is stored; p
// and v alternate between two buffers (a pass reads its neighbours' old
// values while writing its own new ones).  The elementwise updates are
// evaluated as the reference writes them, without fused multiply-adds.
// The first pass (INIT) forms r = b - A x0 = L V*, r^ = r, ||b|| and M^-1.
// Passes after the step's solve ended return at once.
// ---------------------------------------------------------------------------
enum { SC_RHO0 = 8, SC_RHO1, SC_ALPHA, SC_OMEGA, SC_BETA, SC_RESET, SC_SEXIT, SC_CONV };
enum { BI_INIT = 0, BI_P0, BI_P, BI_S };

// One side of the solve (myocardial slots or cable nodes): n entries each.
template <typename T>
struct BiSide {
  T* x;        // the iterate, in place of V*
  T* r;
  T* rh;       // r^ (shadow residual)
  T* minv;     // M^-1 = 1/(Cm/dt - theta L_ii)
  T* t;
  const T* pin;
  const T* vin;
  T* pout;
  T* vout;
  uint32_t n;
};

template <typename T>
__device__ __forceinline__ T bi_pdir(T r, T p, T v, T beta, T omega) {
#pragma clang fp contract(off)
  return r + beta * (p - omega * v);
}
template <typename T>
__device__ __forceinline__ T bi_sdir(T r, T v, T alpha) {
#pragma clang fp contract(off)
  return r - alpha * v;
}

// The vector the pass applies A to, at entry q of one side (q may be OOB:
// every load then reads out of range and returns 0).
template <typename T, int MODE>
struct BiZ {
  buf_t x, r, p, v, mi;
  T alpha, beta, omega;
  __device__ __forceinline__ BiZ(const BiSide<T>& s, T a, T b, T w) : alpha(a), beta(b), omega(w) {
    const uint32_t bytes = s.n * (uint32_t)sizeof(T);
    x = buf_rsrc(s.x, MODE == BI_INIT ? bytes : 0u);
    r = buf_rsrc(s.r, MODE == BI_INIT ? 0u : bytes);
    p = buf_rsrc(s.pin, MODE == BI_P ? bytes : 0u);
    v = buf_rsrc(s.vin, MODE >= BI_P ? bytes : 0u);
    mi = buf_rsrc(s.minv, MODE == BI_INIT ? 0u : bytes);
  }
  __device__ __forceinline__ T operator()(uint32_t q) const {
#pragma clang fp contract(off)
    if (MODE == BI_INIT) return buf_ld<T>(x, q);
    const T mq = buf_ld<T>(mi, q), rq = buf_ld<T>(r, q);
    if (MODE == BI_P0) return mq * rq;
    if (MODE == BI_P) return mq * bi_pdir(rq, buf_ld<T>(p, q), buf_ld<T>(v, q), beta, omega);
    return mq * bi_sdir(rq, buf_ld<T>(v, q), alpha);
  }
};

// The row's own entries a pass reads besides z: r (all but INIT), p, v, r^
// (P), v (S).
template <typename T>
struct BiOwn {
  T r, p, v, rh;
};
template <typename T, int MODE>
__device__ __forceinline__ BiOwn<T> bi_own(const BiSide<T>& s, uint32_t i) {
  BiOwn<T> o{T(0), T(0), T(0), T(0)};
  if (MODE == BI_INIT) return o;
  o.r = ld32((const T*)s.r, i);
  if (MODE == BI_P || MODE == BI_S) o.v = ld32(s.vin, i);
  if (MODE == BI_P) {
    o.p = ld32(s.pin, i);
    o.rh = ld32((const T*)s.rh, i);
  }
  if (MODE == BI_P0) o.rh = ld32((const T*)s.rh, i);
  return o;
}

// Outputs of one row i of a pass given z_i, (L z)_i, -L_ii and the row's own
// entries; accumulates the pass's inner products in FP64.
template <typename T, int MODE>
__device__ __forceinline__ void bi_row(const BiSide<T>& s, uint32_t i, T zc, T lz, T ldg,
                                       const BiOwn<T>& o, const hp_cn_params& cp, T alpha,
                                       T beta, T omega, bool reset, double (&q)[3]) {
#pragma clang fp contract(off)
  const T cm_dt = T(cp.cm_dt), th = T(cp.theta);
  if (MODE == BI_INIT) {
    const T bv = cm_dt * zc + (T(1) - th) * lz;
    st32(s.r, i, lz);
    st32(s.rh, i, lz);
    st32(s.minv, i, T(1) / (cm_dt + th * ldg));
    q[0] += (double)bv * (double)bv;
    q[1] += (double)lz * (double)lz;
    return;
  }
  const T az = cm_dt * zc - th * lz;
  if (MODE == BI_S) {
    const T sv = bi_sdir(o.r, o.v, alpha);
    st32(s.t, i, az);
    q[0] += (double)sv * (double)sv;
    q[1] += (double)az * (double)az;
    q[2] += (double)az * (double)sv;
    return;
  }
  const T pn = MODE == BI_P0 ? o.r : bi_pdir(o.r, o.p, o.v, beta, omega);
  st32(s.pout, i, pn);
  st32(s.vout, i, az);
  T rh = o.rh;
  if (reset) {
    rh = o.r;
    st32(s.rh, i, o.r);
  }
  q[0] += (double)rh * (double)az;
}

// Block sums of q[0..NV) -> part[v * gridDim.x + blockIdx.x] (NW waves per block).
template <int NW, int NV>
__device__ __forceinline__ void bi_block_partials(const double (&q)[3], double* part) {
  __shared__ double sh[NV][NW];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double w = cn_wave_sum(q[v]);
    if (lane == 0) sh[v][wave] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double a = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) a += sh[v][w];
      part[v * gridDim.x + blockIdx.x] = a;
    }
  }
}

// The XCD-contiguous brick range of a capped grid (as the sweeps).
struct BrickRange {
  uint32_t first, end, stride;
};
__device__ __forceinline__ BrickRange brick_range(uint32_t bid, uint32_t G, uint32_t NB) {
  const uint32_t xcd = bid & 7;
  const uint32_t m = (G - xcd + 7) >> 3;
  const uint32_t base = NB >> 3, rem = NB & 7;
  const uint32_t lo = xcd * base + (xcd < rem ? xcd : rem);
  const uint32_t hi = lo + base + (xcd < rem ? 1u : 0u);
  return BrickRange{lo + (bid >> 3), hi, m};
}

// Stencil pass of the solve (INIT, P0/P = A(k), S = S(k)): cable rows in the
// first ncb workgroups (grid-stride), bricks after them.
template <typename T, int BX, int BY, int BZ, int MODE>
__global__ void __launch_bounds__(BX * BY * BZ) cn_bi_pass_kernel(
    hp_bricks g, BiSide<T> m, hp_cable c, BiSide<T> cv, hp_cn_pmj j, hp_cn_params cp,
    double* part, uint32_t ncb) {
  constexpr int BVOL = BX * BY * BZ;
  constexpr int NV = MODE == BI_INIT ? 2 : MODE == BI_S ? 3 : 1;
  typedef Tile<T, BX, BY, BZ> TT;
  __shared__ TT t;
  const double* sc = cp.scalars;
  if ((MODE == BI_P || MODE == BI_S) && sc[SC_DONE] != 0.0) return;
  const T alpha = MODE == BI_S ? T(sc[SC_ALPHA]) : T(0);
  const T beta = MODE == BI_P ? T(sc[SC_BETA]) : T(0);
  const T omega = MODE == BI_P ? T(sc[SC_OMEGA]) : T(0);
  const bool reset = MODE == BI_P && sc[SC_RESET] != 0.0;
  const BiZ<T, MODE> zm(m, alpha, beta, omega), zp(cv, alpha, beta, omega);
  double q[3] = {0.0, 0.0, 0.0};
  const int tid = threadIdx.x;
  if (blockIdx.x < ncb) {
    const T dsw = T(cp.dscale_pk * c.w);
    for (int64_t i = (int64_t)blockIdx.x * BVOL + tid; i < c.n; i += (int64_t)ncb * BVOL) {
      const T zc = zp((uint32_t)i);
      T acc = T(0);
      const int e0 = c.indptr[i], e1 = c.indptr[i + 1];
      for (int e = e0; e < e1; ++e) acc += zp((uint32_t)c.indices[e]) - zc;
      T lz = dsw * acc, ldg = dsw * T(e1 - e0);
      if (j.node_ptr) {
        for (int e = j.node_ptr[i]; e < j.node_ptr[i + 1]; ++e) {
          const T w = T(j.node_w[e]);
          lz += w * (zm((uint32_t)j.node_slot[e]) - zc);
          ldg += w;
        }
      }
      bi_row<T, MODE>(cv, (uint32_t)i, zc, lz, ldg, bi_own<T, MODE>(cv, (uint32_t)i), cp, alpha,
                      beta, omega, reset, q);
    }
  } else {
    const BrickRange br = brick_range(blockIdx.x - ncb, gridDim.x - ncb, (uint32_t)g.n_bricks);
    int vx, vy, vz;
    compute_voxel<BX, BY, BZ>(tid, vx, vy, vz);
    const int cc = ((vx + 1) * TT::TY + (vy + 1)) * TT::TZ + (vz + 1);
    const uint32_t vloc = (uint32_t)((vx * BY + vy) * BZ + vz);
    const T dscale = T(cp.dscale);
    // one brick per workgroup iteration (a workgroup takes ~2 bricks at the
    // default grid; prefetching the next brick's halo measured slower)
    for (uint32_t b = br.first; b < br.end; b += br.stride) {
      __syncthreads();  // every lane is done with the previous tile
      {
        HaloNbr<BX, BY, BZ> hn;
        halo_nbr<BX, BY, BZ>(hn, g, b);
        HaloRegs<T, BX, BY, BZ> hr;
        halo_load_fn<T, BX, BY, BZ>(hr, hn, g, g.label[b * BVOL + tid], zm);
        halo_store_nan<T, BX, BY, BZ>(t, hr);
        __syncthreads();
      }
      const uint32_t slot = b * BVOL + vloc;
      const uint8_t mylab = t.L[cc];
      if (mylab == 0xFF) continue;
      const T zc = t.V[cc];
      T ws;
      const T acc = stencil19_wsum_nan<T, BX, BY, BZ>(t, cc, ws);
      T lz = dscale * acc, ldg = dscale * ws;
      if ((mylab & 0x80) && j.n_targets > 0) {  // PMJ target voxel: its junction couplings
        const int qi = cn_find_target(j.tgt_slot, (int)j.n_targets, (int)slot);
        const int e1 = j.tgt_slot[qi] == (int)slot ? j.tgt_ptr[qi + 1] : 0;  // listed slots only
        for (int e = j.tgt_ptr[qi]; e < e1; ++e) {
          const T w = T(j.tgt_w[e]);
          lz += w * (zp((uint32_t)j.tgt_node[e]) - zc);
          ldg += w;
        }
      }
      bi_row<T, MODE>(m, slot, zc, lz, ldg, bi_own<T, MODE>(m, slot), cp, alpha, beta, omega,
                      reset, q);
    }
  }
  bi_block_partials<(BVOL + 63) / 64, NV>(q, part);
}

template <typename T>
__device__ __forceinline__ void bi_update(const BiSide<T>& s, uint32_t i, T alpha, T omega,
                                          bool sexit, double (&q)[3]) {
#pragma clang fp contract(off)
  const T ph = ld32((const T*)s.minv, i) * ld32(s.pin, i);
  const T x = ld32((const T*)s.x, i);
  if (sexit) {
    st32(s.x, i, x + alpha * ph);
    return;
  }
  const T sv = bi_sdir(ld32((const T*)s.r, i), ld32(s.vin, i), alpha);
  const T sh = ld32((const T*)s.minv, i) * sv;
  st32(s.x, i, (x + alpha * ph) + omega * sh);
  const T rn = sv - omega * ld32((const T*)s.t, i);
  st32(s.r, i, rn);
  q[0] += (double)rn * (double)rn;
  q[1] += (double)ld32((const T*)s.rh, i) * (double)rn;
}

// X(k): the iterate and residual update (pointwise), same grid as the passes.
template <typename T, int BVOL>
__global__ void __launch_bounds__(BVOL) cn_bi_update_kernel(hp_bricks g, BiSide<T> m, hp_cable c,
                                                            BiSide<T> cv, hp_cn_params cp,
                                                            double* part, uint32_t ncb) {
  const double* sc = cp.scalars;
  if (sc[SC_DONE] != 0.0) return;
  const T alpha = T(sc[SC_ALPHA]), omega = T(sc[SC_OMEGA]);
  const bool sexit = sc[SC_SEXIT] != 0.0;
  double q[3] = {0.0, 0.0, 0.0};
  if (blockIdx.x < ncb) {
    for (int64_t i = (int64_t)blockIdx.x * BVOL + threadIdx.x; i < c.n; i += (int64_t)ncb * BVOL)
      bi_update<T>(cv, (uint32_t)i, alpha, omega, sexit, q);
  } else {
    const BrickRange br = brick_range(blockIdx.x - ncb, gridDim.x - ncb, (uint32_t)g.n_bricks);
    for (uint32_t b = br.first; b < br.end; b += br.stride) {
      const uint32_t slot = b * BVOL + threadIdx.x;
      if (g.label[slot] != 0xFF) bi_update<T>(m, slot, alpha, omega, sexit, q);
    }
  }
  if (!sexit) bi_block_partials<(BVOL + 63) / 64, 2>(q, part);
}

// Sums of NV partial columns of nb workgroups (fixed order), in thread 0.
template <int NV>
__device__ __forceinline__ void bi_sum_cols(const double* __restrict__ part, int nb,
                                            double (&out)[NV]) {
  __shared__ double sh[NV][CN_FIN_THREADS / 64];
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  for (int b0 = 0; b0 < nb; b0 += 4 * CN_FIN_THREADS) {
    double x[NV][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + u * CN_FIN_THREADS + (int)threadIdx.x;
#pragma unroll
      for (int v = 0; v < NV; ++v) x[v][u] = b < nb ? part[v * nb + b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] += x[v][u];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double w = cn_wave_sum(acc[v]);
    if (lane == 0) sh[v][wave] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double a = 0.0;
      for (int w = 0; w < CN_FIN_THREADS / 64; ++w) a += sh[v][w];
      out[v] = a;
    }
  }
}

// The step's solve ends after `iters` iterations begun; statistics as for
// the sweeps (the iteration count in place of the sweep count).
__device__ __forceinline__ void bi_done(double* sc, bool conv, int iters, double rel) {
  sc[SC_DONE] = 1.0;
  sc[SC_CONV] = conv ? 1.0 : 0.0;
  sc[SC_SWEEPS] = (double)iters;
  sc[SC_TOTAL] += (double)iters;
  sc[SC_STEPS] += 1.0;
  if ((double)iters > sc[SC_MAXIT]) sc[SC_MAXIT] = (double)iters;
  if (!conv) {
    sc[SC_NCONV] += 1.0;
    if (rel > sc[SC_WORST]) sc[SC_WORST] = rel;
  }
}

enum { BI_FIN_A0 = 0, BI_FIN_A, BI_FIN_S, BI_FIN_X };

// The scalar part of iteration k (reference :786-849, in its order):
//   A0: ||b|| (0: x0 returned), rho = r^.r = r.r, da = r^.v (|da| < 1e-300:
//       stop), alpha = rho/da;   A(k >= 1): the same from da and the stored rho
//   S:  ||s|| < tol ||b||: the update takes x += alpha p^ and the solve ends;
//       omega = t.s/t.t (0 when |t.t| <= 1e-300)
//   X:  ||r||/||b|| < tol: converged; |omega| < 1e-300 or the last iteration:
//       stop; rho' = r^.r (|rho'| < 1e-300: r^ := r, rho' = r.r),
//       beta = (rho'/rho)(alpha/omega)
// part points at the stage's columns: A0 (bb, rr, da), A (da), S (ss, tt, ts),
// X (rr, r^.r); nb = workgroups of the passes.
template <int STAGE>
__global__ void __launch_bounds__(CN_FIN_THREADS) cn_bi_finish_kernel(hp_cn_params cp,
                                                                      const double* part, int nb,
                                                                      int k) {
  double* sc = cp.scalars;
  if (STAGE != BI_FIN_A0 && sc[SC_DONE] != 0.0) return;
  constexpr int NV = STAGE == BI_FIN_A ? 1 : STAGE == BI_FIN_X ? 2 : 3;
  double v[NV];
  bi_sum_cols<NV>(part, nb, v);
  if (threadIdx.x != 0) return;
  const double tol = cp.tol;
  if (STAGE == BI_FIN_A0) {
    const double bb = v[0], rr = v[1], da = v[2 % NV];
    sc[SC_BB] = bb;
    sc[SC_RR] = rr;
    sc[SC_SEXIT] = 0.0;
    sc[SC_RESET] = 0.0;
    sc[SC_DONE] = 0.0;
    const double bn = sqrt(bb);
    if (bn == 0.0) {
      bi_done(sc, true, 0, 0.0);
      return;
    }
    sc[SC_RHO0] = rr;
    if (fabs(da) < 1e-300) {
      const double rel = sqrt(rr) / bn;
      bi_done(sc, rel < tol, 1, rel);
      return;
    }
    sc[SC_ALPHA] = rr / da;
    return;
  }
  const double bn = sqrt(sc[SC_BB]);
  if (STAGE == BI_FIN_A) {
    const double da = v[0];
    if (fabs(da) < 1e-300) {
      const double rel = sqrt(sc[SC_RR]) / bn;
      bi_done(sc, rel < tol, k + 1, rel);
      return;
    }
    sc[SC_ALPHA] = sc[SC_RHO0 + (k & 1)] / da;
    return;
  }
  if (STAGE == BI_FIN_S) {
    const double ss = v[0], tt = v[1 % NV], ts = v[2 % NV];
    if (sqrt(ss) < tol * bn) {
      sc[SC_SEXIT] = 1.0;
      return;
    }
    sc[SC_OMEGA] = fabs(tt) > 1e-300 ? ts / tt : 0.0;
    return;
  }
  // BI_FIN_X
  if (sc[SC_SEXIT] != 0.0) {
    sc[SC_SEXIT] = 0.0;
    bi_done(sc, true, k + 1, 0.0);
    return;
  }
  const double rr = v[0], rhr = v[1 % NV];
  sc[SC_RR] = rr;
  const double rel = sqrt(rr) / bn;
  if (rel < tol) {
    bi_done(sc, true, k + 1, rel);
    return;
  }
  const double omega = sc[SC_OMEGA];
  if (fabs(omega) < 1e-300 || k + 1 >= cp.max_sweeps) {
    bi_done(sc, false, k + 1, rel);
    return;
  }
  double rho = rhr;
  const bool reset = fabs(rho) < 1e-300;
  if (reset) rho = rr;
  sc[SC_RESET] = reset ? 1.0 : 0.0;
  sc[SC_BETA] = (rho / sc[SC_RHO0 + (k & 1)]) * (sc[SC_ALPHA] / omega);
  sc[SC_RHO0 + ((k + 1) & 1)] = rho;
}

template <typename T, int BX, int BY, int BZ>
static void launch_bi_pass(int mode, dim3 grid, hipStream_t st, const hp_bricks& g,
                           const BiSide<T>& m, const hp_cable& c, const BiSide<T>& cv,
                           const hp_cn_pmj& j, const hp_cn_params& cp, double* part,
                           uint32_t ncb) {
  const dim3 blk(BX * BY * BZ);
  switch (mode) {
    case BI_INIT:
      hipLaunchKernelGGL((cn_bi_pass_kernel<T, BX, BY, BZ, BI_INIT>), grid, blk, 0, st, g, m, c,
                         cv, j, cp, part, ncb);
      break;
    case BI_P0:
      hipLaunchKernelGGL((cn_bi_pass_kernel<T, BX, BY, BZ, BI_P0>), grid, blk, 0, st, g, m, c, cv,
                         j, cp, part, ncb);
      break;
    case BI_P:
      hipLaunchKernelGGL((cn_bi_pass_kernel<T, BX, BY, BZ, BI_P>), grid, blk, 0, st, g, m, c, cv,
                         j, cp, part, ncb);
      break;
    default:
      hipLaunchKernelGGL((cn_bi_pass_kernel<T, BX, BY, BZ, BI_S>), grid, blk, 0, st, g, m, c, cv,
                         j, cp, part, ncb);
      break;
  }
}

template <typename T>
static void bi_pass(int key, int mode, dim3 grid, hipStream_t st, const hp_bricks& g,
                    const BiSide<T>& m, const hp_cable& c, const BiSide<T>& cv,
                    const hp_cn_pmj& j, const hp_cn_params& cp, double* part, uint32_t ncb) {
  switch (key) {
    case 444: launch_bi_pass<T, 4, 4, 4>(mode, grid, st, g, m, c, cv, j, cp, part, ncb); break;
    case 844: launch_bi_pass<T, 8, 4, 4>(mode, grid, st, g, m, c, cv, j, cp, part, ncb); break;
    case 884: launch_bi_pass<T, 8, 8, 4>(mode, grid, st, g, m, c, cv, j, cp, part, ncb); break;
    default:  launch_bi_pass<T, 8, 8, 8>(mode, grid, st, g, m, c, cv, j, cp, part, ncb); break;
  }
}

template <typename T>
static void bi_update_launch(int key, dim3 grid, hipStream_t st, const hp_bricks& g,
                             const BiSide<T>& m, const hp_cable& c, const BiSide<T>& cv,
                             const hp_cn_params& cp, double* part, uint32_t ncb) {
  const int bvol = key == 444 ? 64 : key == 844 ? 128 : key == 884 ? 256 : 512;
  switch (bvol) {
    case 64: hipLaunchKernelGGL((cn_bi_update_kernel<T, 64>), grid, dim3(64), 0, st, g, m, c, cv, cp, part, ncb); break;
    case 128: hipLaunchKernelGGL((cn_bi_update_kernel<T, 128>), grid, dim3(128), 0, st, g, m, c, cv, cp, part, ncb); break;
    case 256: hipLaunchKernelGGL((cn_bi_update_kernel<T, 256>), grid, dim3(256), 0, st, g, m, c, cv, cp, part, ncb); break;
    default: hipLaunchKernelGGL((cn_bi_update_kernel<T, 512>), grid, dim3(512), 0, st, g, m, c, cv, cp, part, ncb); break;
  }
}

template <typename T, int BX, int BY, int BZ>
static void launch_sweep(bool first, dim3 grid, hipStream_t st, const hp_bricks& g, const T* xi,
                         T* xo, T* bvec, const hp_cable& c, const T* xpi, T* xpo, T* bpk,
                         const hp_cn_pmj& j, const hp_cn_params& cp, uint32_t ncb,
                         const CnWarm<T>& wm = CnWarm<T>{}) {
  if (first && wm.d)
    hipLaunchKernelGGL((cn_sweep_kernel<T, BX, BY, BZ, true, true>), grid, dim3(BX * BY * BZ), 0,
                       st, g, xi, xo, bvec, c, xpi, xpo, bpk, j, cp, ncb, wm);
  else if (first)
    hipLaunchKernelGGL((cn_sweep_kernel<T, BX, BY, BZ, true>), grid, dim3(BX * BY * BZ), 0, st, g,
                       xi, xo, bvec, c, xpi, xpo, bpk, j, cp, ncb, wm);
  else
    hipLaunchKernelGGL((cn_sweep_kernel<T, BX, BY, BZ, false>), grid, dim3(BX * BY * BZ), 0, st,
                       g, xi, xo, bvec, c, xpi, xpo, bpk, j, cp, ncb, wm);
}

template <typename T>
static int cn_step(const hp_bricks* g, const T* V_in, T* V_out, T* state, int64_t ld,
                   const hp_mono_params* p, const hp_cable* c, const T* Vpk_in, T* Vpk_out,
                   T* state_pk, int64_t ld_pk, const hp_mono_params* p_pk, const hp_cn_pmj* j,
                   const hp_cn_params* cp, const hp_cn_work* w, void* stream) {
  const char* what = "hp_cn_step";
  if (!g || !p || !cp || !w || !p->step_ctr || !p->group_amp ||
      !(p->do_reaction == 1 || (p->do_reaction == 2 && p->rate_ext)) ||
      !g->groups || !g->label || !g->D6 || !g->nbr || !state || !V_in || !V_out || ld < 0 ||
      (uint64_t)ld * sizeof(T) >= (1ull << 32)) {
    set_error("%s: brick grid (groups, label, D6, nbr), V_in, V_out, state, params (step_ctr, "
              "group_amp, do_reaction 1, or 2 with rate_ext), CN params and work vectors are "
              "required",
              what);
    return -1;
  }
  const bool bicg = cp->solver == 1;
  bool kr_ok = true;
  for (int q = 0; q < 8; ++q) kr_ok = kr_ok && w->kr[q];
  if (!cp->scalars || !cp->part || !w->Vs || cp->max_sweeps < 1 ||
      !(cp->solver == 0 || cp->solver == 1) || (!bicg && (!w->b || !w->x)) || (bicg && !kr_ok) ||
      !(cp->cm_dt > 0.0) || !(cp->theta >= 0.0 && cp->theta <= 1.0) || !(cp->tol >= 0.0)) {
    set_error("%s: CN params need scalars, part, Cm/dt > 0, 0 <= theta <= 1, tol >= 0, "
              "max_sweeps >= 1, solver 0 (work vectors Vs, b, x) or 1 (Vs and kr[0..8))",
              what);
    return -1;
  }
  const int64_t ncab = c ? c->n : 0;
  if (ncab < 0 || (ncab > 0 && (!p_pk || !p_pk->group_amp || !Vpk_in || !Vpk_out || !state_pk ||
                                ld_pk < ncab || (uint64_t)ld_pk * sizeof(T) >= (1ull << 32) ||
                                !c->indptr || !c->indices || !c->label || !w->Vs_pk ||
                                (!bicg && (!w->b_pk || !w->x_pk)) ||
                                (bicg && !(w->kr_pk[0] && w->kr_pk[1] && w->kr_pk[2] &&
                                           w->kr_pk[3] && w->kr_pk[4] && w->kr_pk[5] &&
                                           w->kr_pk[6] && w->kr_pk[7]))))) {
    set_error("%s: a cable needs its params, V, state (ld_pk >= n), CSR, labels and the work "
              "vectors Vs_pk with b_pk, x_pk (solver 0) or kr_pk[0..8) (solver 1)",
              what);
    return -1;
  }
  if (w->warm[0] && (bicg || !w->warm[1] || (ncab > 0 && !(w->warm[2] && w->warm[3])))) {
    set_error("%s: the warm start (solver 0 only) needs d and e over the slots, and d_pk and "
              "e_pk with a cable", what);
    return -1;
  }
  const hp_cn_pmj j0{};
  if (!j) j = &j0;
  if (j->n_targets < 0 || (j->n_targets > 0 && (!j->tgt_slot || !j->tgt_ptr || !j->tgt_node ||
                                                !j->tgt_w || ncab == 0)) ||
      (j->node_ptr && (!j->node_slot || !j->node_w)) || j->n_jnodes < 0 ||
      (j->n_jnodes > 0 && (!j->jnode || !j->node_ptr))) {
    set_error("%s: PMJ rows need tgt_slot/tgt_ptr/tgt_node/tgt_w (and a cable), node_slot and "
              "node_w with node_ptr, and node_ptr with the junction node list",
              what);
    return -1;
  }
  const int64_t bvol = (int64_t)g->bx * g->by * g->bz;
  if (g->n_bricks < 0 || (uint64_t)g->n_bricks * bvol * 6 * sizeof(T) >= (1ull << 32)) {
    set_error("%s: %lld bricks exceed the 32-bit offset range of one launch", what,
              (long long)g->n_bricks);
    return -1;
  }
  const int key = g->bx * 100 + g->by * 10 + g->bz;
  if (key != 444 && key != 844 && key != 884 && key != 888) {
    set_error("%s: unsupported brick shape %dx%dx%d", what, g->bx, g->by, g->bz);
    return -1;
  }
  const hp_cable c0{};
  const hp_mono_params p0{};
  if (!c) c = &c0;
  if (!p_pk) p_pk = &p0;
  hipStream_t st = (hipStream_t)stream;
  const int64_t n_slots = g->n_bricks * bvol;
  T* Vs = (T*)w->Vs;
  T* bvec = (T*)w->b;
  T* x = (T*)w->x;
  T* Vspk = (T*)w->Vs_pk;
  T* bpk = (T*)w->b_pk;
  T* xpk = (T*)w->x_pk;

  // reaction
  const uint32_t ncb_r = (uint32_t)ceil_div(ncab, 256);
  if (p->cell_slot && (p->n_cells < 0 || p->n_cells > n_slots || p->n_cells > ld)) {
    set_error("%s: cell_slot needs 0 <= n_cells <= min(slots, ld)", what);
    return -1;
  }
  const int64_t nblk_r = ncb_r + ceil_div(p->cell_slot ? p->n_cells : n_slots, 256);
  if (nblk_r > 0) {
    hipLaunchKernelGGL(cn_react_kernel<T>, dim3((unsigned)nblk_r), dim3(256), 0, st, *g, V_in, Vs,
                       state, ld, *p, *c, Vpk_in, Vspk, state_pk, ld_pk, *p_pk, ncb_r);
    if (int rc = check_launch("hp_cn_step/react")) return rc;
  }
  int64_t ncb = ceil_div(ncab, bvol);
  if (ncb > CN_MAX_CABLE_BLOCKS) ncb = CN_MAX_CABLE_BLOCKS;
  ncb = (ncb + 7) / 8 * 8;
  const int64_t cap = cp->grid_cap > 0 ? cp->grid_cap : CN_SWEEP_BLOCKS;
  const int64_t G = ncb + (g->n_bricks < cap ? g->n_bricks : cap);
  const int64_t cols = cp->solver == 1 ? 8 : 2;
  if (G > 0x7fffffff || cols * G > cp->part_len) {
    set_error("%s: the passes' %lld workgroups need part_len >= %lld doubles (got %lld)", what,
              (long long)G, (long long)(cols * G), (long long)cp->part_len);
    return -1;
  }
  const uint32_t ncb_f = (uint32_t)ceil_div(ncab, 256);
  const int64_t nblk_f = ncb_f + ceil_div(p->cell_slot ? p->n_cells : n_slots, 256);
  if (cp->solver == 1) {
    // BiCGSTAB from x0 = V* (the iterate overwrites Vs)
    BiSide<T> m{}, cv{};
    T* const* kr = (T* const*)w->kr;
    T* const* kp = (T* const*)w->kr_pk;
    m = BiSide<T>{Vs, kr[0], kr[1], kr[2], kr[3], nullptr, nullptr, nullptr, nullptr,
                  (uint32_t)n_slots};
    if (ncab > 0)
      cv = BiSide<T>{Vspk, kp[0], kp[1], kp[2], kp[3], nullptr, nullptr, nullptr, nullptr,
                     (uint32_t)ncab};
    double* part = cp->part;
    const dim3 grid((unsigned)G);
    if (G > 0) {
      bi_pass<T>(key, BI_INIT, grid, st, *g, m, *c, cv, *j, *cp, part, (uint32_t)ncb);
      if (int rc = check_launch("hp_cn_step/bicgstab init")) return rc;
    }
    for (int k = 0; k < cp->max_sweeps; ++k) {
      const int a = k & 1, b = a ^ 1;
      m.pin = kr[4 + a];
      m.vin = kr[6 + a];
      m.pout = kr[4 + b];
      m.vout = kr[6 + b];
      if (ncab > 0) {
        cv.pin = kp[4 + a];
        cv.vin = kp[6 + a];
        cv.pout = kp[4 + b];
        cv.vout = kp[6 + b];
      }
      if (G > 0) bi_pass<T>(key, k == 0 ? BI_P0 : BI_P, grid, st, *g, m, *c, cv, *j, *cp,
                            part + 2 * G, (uint32_t)ncb);
      if (k == 0)
        hipLaunchKernelGGL((cn_bi_finish_kernel<BI_FIN_A0>), dim3(1), dim3(CN_FIN_THREADS), 0, st,
                           *cp, part, (int)G, k);
      else
        hipLaunchKernelGGL((cn_bi_finish_kernel<BI_FIN_A>), dim3(1), dim3(CN_FIN_THREADS), 0, st,
                           *cp, part + 2 * G, (int)G, k);
      if (int rc = check_launch("hp_cn_step/bicgstab A")) return rc;
      // S(k) and X(k) read the new direction and product
      m.pin = m.pout;
      m.vin = m.vout;
      cv.pin = cv.pout;
      cv.vin = cv.vout;
      if (G > 0) bi_pass<T>(key, BI_S, grid, st, *g, m, *c, cv, *j, *cp, part + 3 * G,
                            (uint32_t)ncb);
      hipLaunchKernelGGL((cn_bi_finish_kernel<BI_FIN_S>), dim3(1), dim3(CN_FIN_THREADS), 0, st,
                         *cp, part + 3 * G, (int)G, k);
      if (int rc = check_launch("hp_cn_step/bicgstab S")) return rc;
      if (G > 0) bi_update_launch<T>(key, grid, st, *g, m, *c, cv, *cp, part + 6 * G,
                                     (uint32_t)ncb);
      hipLaunchKernelGGL((cn_bi_finish_kernel<BI_FIN_X>), dim3(1), dim3(CN_FIN_THREADS), 0, st,
                         *cp, part + 6 * G, (int)G, k);
      if (int rc = check_launch("hp_cn_step/bicgstab X")) return rc;
    }
    if (nblk_f > 0)
      hipLaunchKernelGGL(cn_final_kernel<T>, dim3((unsigned)nblk_f), dim3(256), 0, st, *g, V_in,
                         Vs, V_out, Vspk, Vpk_out, ncab, *p, cp->scalars, ncb_f, false, Vs, Vspk,
                         CnWarm<T>{});
    return check_launch("hp_cn_step/final");
  }
  // sweeps: x_0 = V* (+ e with the warm start), sweep k writes x (k odd) or
  // V_out (k even)
  const CnWarm<T> wm{(T*)w->warm[0], (T*)w->warm[1], (T*)w->warm[2], (T*)w->warm[3]};
  for (int k = 1; k <= cp->max_sweeps; ++k) {
    const T* xi = k == 1 ? Vs : (k % 2 == 0 ? x : V_out);
    T* xo = k % 2 == 1 ? x : V_out;
    const T* xpi = k == 1 ? Vspk : (k % 2 == 0 ? xpk : Vpk_out);
    T* xpo = k % 2 == 1 ? xpk : Vpk_out;
    if (G > 0) {
      switch (key) {
        case 444: launch_sweep<T, 4, 4, 4>(k == 1, dim3(G), st, *g, xi, xo, bvec, *c, xpi, xpo, bpk, *j, *cp, (uint32_t)ncb, wm); break;
        case 844: launch_sweep<T, 8, 4, 4>(k == 1, dim3(G), st, *g, xi, xo, bvec, *c, xpi, xpo, bpk, *j, *cp, (uint32_t)ncb, wm); break;
        case 884: launch_sweep<T, 8, 8, 4>(k == 1, dim3(G), st, *g, xi, xo, bvec, *c, xpi, xpo, bpk, *j, *cp, (uint32_t)ncb, wm); break;
        default:  launch_sweep<T, 8, 8, 8>(k == 1, dim3(G), st, *g, xi, xo, bvec, *c, xpi, xpo, bpk, *j, *cp, (uint32_t)ncb, wm); break;
      }
      if (int rc = check_launch("hp_cn_step/sweep")) return rc;
    }
    if (k == 1)
      hipLaunchKernelGGL((cn_finish_kernel<T, true>), dim3(1), dim3(CN_FIN_THREADS), 0, st, *cp,
                         (int)G, *c, *j, xpi, xpo, xo, bpk, (const T*)wm.e_pk,
                         wm.d ? 1 : 0);
    else
      hipLaunchKernelGGL((cn_finish_kernel<T, false>), dim3(1), dim3(CN_FIN_THREADS), 0, st, *cp,
                         (int)G, *c, *j, xpi, xpo, xo, bpk, (const T*)nullptr, 0);
    if (int rc = check_launch("hp_cn_step/finish")) return rc;
  }
  // clamp + activation times
  if (nblk_f > 0)
    hipLaunchKernelGGL(cn_final_kernel<T>, dim3((unsigned)nblk_f), dim3(256), 0, st, *g, V_in, x,
                       V_out, xpk, Vpk_out, ncab, *p, cp->scalars, ncb_f, true, Vs, Vspk, wm);
  return check_launch("hp_cn_step/final");
}

// One sweep outside a step (the solve's dominant kernel alone, for
// measurement): x_out from x_in and b as the step's sweeps after the first;
// the partial sums land in cp->part, the done flag is not read or written.
template <typename T>
static int cn_sweep_only(const hp_bricks* g, const T* x_in, T* x_out, const T* bvec,
                         const hp_cn_params* cp, void* stream) {
  if (!g || !cp || !cp->part || !x_in || !x_out || !bvec || !g->label || !g->D6 || !g->nbr ||
      g->n_bricks < 0) {
    set_error("hp_cn_sweep: brick grid, x_in, x_out, b and CN params (part) are required");
    return -1;
  }
  const int64_t bvol = (int64_t)g->bx * g->by * g->bz;
  if ((uint64_t)g->n_bricks * bvol * 6 * sizeof(T) >= (1ull << 32)) {
    set_error("hp_cn_sweep: %lld bricks exceed the 32-bit offset range", (long long)g->n_bricks);
    return -1;
  }
  const int64_t G = g->n_bricks < CN_SWEEP_BLOCKS ? g->n_bricks : CN_SWEEP_BLOCKS;
  if (G == 0) return 0;
  if (2 * G > cp->part_len) {
    set_error("hp_cn_sweep: part_len >= %lld doubles required", (long long)(2 * G));
    return -1;
  }
  // without the scalars the sweep reads no done flag: it always runs
  hp_cn_params c2 = *cp;
  c2.scalars = nullptr;
  const hp_cable c0{};
  const hp_cn_pmj j0{};
  const int key = g->bx * 100 + g->by * 10 + g->bz;
  hipStream_t st = (hipStream_t)stream;
  switch (key) {
    case 444: launch_sweep<T, 4, 4, 4>(false, dim3(G), st, *g, x_in, x_out, (T*)bvec, c0, nullptr, nullptr, nullptr, j0, c2, 0u); break;
    case 844: launch_sweep<T, 8, 4, 4>(false, dim3(G), st, *g, x_in, x_out, (T*)bvec, c0, nullptr, nullptr, nullptr, j0, c2, 0u); break;
    case 884: launch_sweep<T, 8, 8, 4>(false, dim3(G), st, *g, x_in, x_out, (T*)bvec, c0, nullptr, nullptr, nullptr, j0, c2, 0u); break;
    case 888: launch_sweep<T, 8, 8, 8>(false, dim3(G), st, *g, x_in, x_out, (T*)bvec, c0, nullptr, nullptr, nullptr, j0, c2, 0u); break;
    default:
      set_error("hp_cn_sweep: unsupported brick shape %dx%dx%d", g->bx, g->by, g->bz);
      return -1;
  }
  return check_launch("hp_cn_sweep");
}

// One BiCGSTAB product pass outside a step (the solve's stencil kernel alone,
// for measurement): t = A M^-1 (r - alpha v) over the myocardial rows, with the
// inner products' partials in cp->part.
template <typename T>
static int cn_apply_only(const hp_bricks* g, const hp_cn_params* cp, const hp_cn_work* w,
                         void* stream) {
  if (!g || !cp || !w || !cp->part || !cp->scalars || !g->label || !g->D6 || !g->nbr ||
      g->n_bricks < 0 || !w->kr[0] || !w->kr[2] || !w->kr[3] || !w->kr[6]) {
    set_error("hp_cn_apply: brick grid, CN params (scalars, part) and the work vectors r, M^-1, "
              "t, v are required");
    return -1;
  }
  const int64_t bvol = (int64_t)g->bx * g->by * g->bz;
  if ((uint64_t)g->n_bricks * bvol * 6 * sizeof(T) >= (1ull << 32)) {
    set_error("hp_cn_apply: %lld bricks exceed the 32-bit offset range", (long long)g->n_bricks);
    return -1;
  }
  const int64_t cap = cp->grid_cap > 0 ? cp->grid_cap : CN_SWEEP_BLOCKS;
  const int64_t G = g->n_bricks < cap ? g->n_bricks : cap;
  if (G == 0) return 0;
  if (3 * G > cp->part_len) {
    set_error("hp_cn_apply: part_len >= %lld doubles required", (long long)(3 * G));
    return -1;
  }
  const int key = g->bx * 100 + g->by * 10 + g->bz;
  if (key != 444 && key != 844 && key != 884 && key != 888) {
    set_error("hp_cn_apply: unsupported brick shape %dx%dx%d", g->bx, g->by, g->bz);
    return -1;
  }
  T* const* kr = (T* const*)w->kr;
  const BiSide<T> m{(T*)w->Vs, kr[0], kr[1], kr[2], kr[3], kr[4], kr[6], nullptr, nullptr,
                    (uint32_t)(g->n_bricks * bvol)};
  const BiSide<T> cv{};
  const hp_cable c0{};
  const hp_cn_pmj j0{};
  bi_pass<T>(key, BI_S, dim3((unsigned)G), (hipStream_t)stream, *g, m, c0, cv, j0, *cp, cp->part,
             0u);
  return check_launch("hp_cn_apply");
}

}  // namespace hp

using namespace hp;

extern "C" {
int hp_cn_apply_f32(const hp_bricks* g, const hp_cn_params* cp, const hp_cn_work* w,
                    void* stream) {
  return cn_apply_only<float>(g, cp, w, stream);
}
int hp_cn_sweep_f32(const hp_bricks* g, const float* x_in, float* x_out, const float* b,
                    const hp_cn_params* cp, void* stream) {
  return cn_sweep_only<float>(g, x_in, x_out, b, cp, stream);
}
int hp_cn_step_f32(const hp_bricks* g, const float* V_in, float* V_out, float* state, int64_t ld,
                   const hp_mono_params* p, const hp_cable* c, const float* Vpk_in, float* Vpk_out,
                   float* state_pk, int64_t ld_pk, const hp_mono_params* p_pk,
                   const hp_cn_pmj* j, const hp_cn_params* cp, const hp_cn_work* w,
                   void* stream) {
  return cn_step<float>(g, V_in, V_out, state, ld, p, c, Vpk_in, Vpk_out, state_pk, ld_pk, p_pk,
                        j, cp, w, stream);
}
int hp_cn_step_f64(const hp_bricks* g, const double* V_in, double* V_out, double* state,
                   int64_t ld, const hp_mono_params* p, const hp_cable* c, const double* Vpk_in,
                   double* Vpk_out, double* state_pk, int64_t ld_pk, const hp_mono_params* p_pk,
                   const hp_cn_pmj* j, const hp_cn_params* cp, const hp_cn_work* w,
                   void* stream) {
  return cn_step<double>(g, V_in, V_out, state, ld, p, c, Vpk_in, Vpk_out, state_pk, ld_pk, p_pk,
                         j, cp, w, stream);
}
}
