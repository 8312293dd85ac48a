/* hrdaya_hip.h -- C ABI of libhrdaya_hip.so, the MI355X (gfx950) kernels behind
 * the hrdayapy simulation and ECG hot path.
 *
 * Conventions
 *   - Every pointer is a device pointer owned by the caller (torch tensors on
 *     the Python side); the library allocates nothing per call.
 *   - `stream` is a hipStream_t passed as void* (torch's current stream).
 *   - Every function returns 0 on success or a negative hipError_t; the
 *     message of the last failure on the calling thread is available from
 *     hp_last_error().
 *   - Suffix _f32 / _f64 selects the arithmetic type of the state and of the
 *     voltages.  FP64 kernels exist so that the GPU path can be compared with
 *     the FP64 oracle at round-off level; FP32 is the production form.
 *
 * Reference interfaces replaced (paths relative to the HrdayaPy tree):
 *   ionic step      hrdayapy/simulation/functions/ionic_models/gpu_ttp06.py:198-412
 *                   (RegionIonicModel.step, protocol.py:72-121;
 *                    HeteroIonicSolver.step, hetero_solver.py:141-194)
 *   diffusion       torch.sparse.mm(L, V*) + _bicgstab,
 *                   hrdayapy/anisotropic_simulation/functions/
 *                   purkinje_myocardium_anisotropic_solver.py:1787-1791, 786-849
 *   stimulus        _stim_active / I_stim_buf, same file :1290-1301, :1754-1759
 *   ECG source      make_gpu_solver.solve_frame, hrdayapy/ecg/functions/compute_bspm.py:381-419
 *   torso PCG       _pcg_gpu, compute_bspm.py:164-263
 */
#ifndef HRDAYA_HIP_H
#define HRDAYA_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HP_ABI_VERSION 6

/* ---------------------------------------------------------------------------
 * Library
 * ------------------------------------------------------------------------- */
int hp_abi_version(void);
/* Copies the last error message of the calling thread into buf (NUL-terminated). */
int hp_last_error(char* buf, size_t size);
/* Device memory for the resident per-slot arrays (no counterpart in the
 * reference, whose tensors come from torch's allocator): one allocation,
 * optionally physically contiguous (hipDeviceMallocContiguous), so the arrays
 * of a long-running engine do not inherit the fragmentation of earlier frees.
 * Free with hp_device_free after the last kernel that uses it has finished. */
int hp_device_alloc(void** ptr, int64_t bytes, int contiguous);
int hp_device_free(void* ptr);

/* ---------------------------------------------------------------------------
 * TTP06 reaction step (ionic plugin seam).
 * Replaces _currents_and_gates (gpu_ttp06.py:198-412) as called through
 * EndoTTP06/MidTTP06/EpiTTP06.step (gpu_ttp06.py:419-513).
 *
 * state element (node i, variable k) lives at state[k*stride_var + i*stride_node]:
 * (N,18) row-major tensors use stride_var=1, stride_node=18; the engine's SoA
 * buffers use stride_var=ld, stride_node=1.  Variables are in the reference's
 * column order (Ki, Nai, Cai, Xr1, Xr2, Xs, m, h, j, d, f, f2, fCass, s, r,
 * CaSR, Rprime, Cass).  Ki/Nai are stored as (value - ki_base/nai_base).
 * celltype: per-node phenotype code (0 endo, 1 mid, 2 epi) or NULL, in which
 * case fixed_celltype applies to every node.  gks_scale/gto_scale: optional
 * per-node multipliers of g_Ks/g_to (NULL = 1).
 * ------------------------------------------------------------------------- */
int hp_ttp06_step_f32(const float* V, float* state, float* I_ion, const uint8_t* celltype,
                      int fixed_celltype, const float* gks_scale, const float* gto_scale,
                      int64_t n, int64_t stride_var, int64_t stride_node, float dt,
                      float ki_base, float nai_base, void* stream);
int hp_ttp06_step_f64(const double* V, double* state, double* I_ion, const uint8_t* celltype,
                      int fixed_celltype, const double* gks_scale, const double* gto_scale,
                      int64_t n, int64_t stride_var, int64_t stride_node, double dt,
                      double ki_base, double nai_base, void* stream);
/* The same step on HOST memory, computed on the CPU with the kernels' own
 * arithmetic form (C library exp/log/division standing in for the FP32
 * hardware instructions).  For checking that form without a GPU; same
 * arguments as above, no stream. */
int hp_ttp06_step_host_f32(const float* V, float* state, float* I_ion, const uint8_t* celltype,
                           int fixed_celltype, const float* gks_scale, const float* gto_scale,
                           int64_t n, int64_t stride_var, int64_t stride_node, float dt,
                           float ki_base, float nai_base);
int hp_ttp06_step_host_f64(const double* V, double* state, double* I_ion,
                           const uint8_t* celltype, int fixed_celltype, const double* gks_scale,
                           const double* gto_scale, int64_t n, int64_t stride_var,
                           int64_t stride_node, double dt, double ki_base, double nai_base);

/* ---------------------------------------------------------------------------
 * Myocardium on a brick layout.
 *
 * The myocardial voxels of the (Nx,Ny,Nz) mask are stored in bricks of
 * bx*by*bz voxels (bx,by,bz in {4,8}); only bricks that hold at least one
 * myocardial voxel are stored.  Voxel (lx,ly,lz) of brick b has slot
 * b*bvol + (lx*by + ly)*bz + lz.  Per slot:
 *   label  uint8: 0xFF = not myocardium; else bits 0-1 phenotype code,
 *          bits 2-6 stimulus group (0 = none), bit 7 = PMJ target voxel
 *   D6     conductivity tensor D00,D11,D22,D01,D02,D12 [mS/mm], SoA [6][n_slots]
 *          (row stride n_slots = n_bricks*bvol)
 *   inst   uint16 ensemble instance (or NULL: instance 0)
 * Padding slots carry label 0xFF and D6 = 0.  Every per-slot array, and the
 * six D6 rows together, must fit 32-bit byte offsets (6*n_slots*sizeof(T) <
 * 2^32); the entry points reject larger grids.
 *
 * Cell state is stored for the myocardial slots only (no padding): SoA
 * [18][ld], entry q of a row = the q-th myocardial slot in slot order, so a
 * wave's loads cover only cache lines that hold tissue.  groups[g] (4 int32
 * per 64 consecutive slots, g = slot >> 6) = {number of myocardial slots
 * before slot 64*g, 0, myocardial-slot mask bits 0-31, bits 32-63}: the
 * state entry of slot s is groups[s>>6].base + popcount(mask below s&63).
 * The mask must agree with label != 0xFF.  The kernels that touch the state
 * (hp_myo_step, hp_fallback_react, hp_surrogate_step, hp_bigstep_reaction)
 * require groups; ld >= the myocardial slot count and ld*sizeof(T) < 2^32.
 * nbr[b*27 + (dx+1)*9 + (dy+1)*3 + (dz+1)] is the brick at offset (dx,dy,dz),
 * or -1.
 *
 * The 19-point operator is the reference's build_myo_mesh_anisotropic
 * (purkinje_myocardium_anisotropic_solver.py:886-1074) with L = D^-1 K
 * (:606-746): face weight harm(D_kk,a, D_kk,b)*h, edge-diagonal weight
 * +-0.5*(D_ij,a + D_ij,b)*h/2, rows divided by A_M*h^3.
 * ------------------------------------------------------------------------- */
typedef struct hp_bricks {
  int64_t n_bricks;
  int32_t bx, by, bz;
  int32_t pad_;
  const int32_t* nbr;    /* [n_bricks][27] */
  const uint8_t* label;  /* [n_slots] */
  const void* D6;        /* T[6][n_slots] */
  const uint16_t* inst;  /* [n_slots] or NULL */
  const int32_t* groups; /* [n_slots/64][4] state index of each 64-slot group (see above) */
} hp_bricks;

/* Per-step scalars of the monodomain update.
 *   V_new = clamp(V + dt*(-I_ion) + (dt/Cm)*(I_stim + L V), vmin, vmax)
 * with L V computed on V^n (explicit, fused with the reaction step). */
typedef struct hp_mono_params {
  double dt;          /* ms */
  double inv_cm;      /* 1/Cm, mm^2/uF */
  double dscale;      /* 1/(A_M*h^2) for the myocardium, 1/Ai_P for the cable */
  double ki_base, nai_base;
  double vmin, vmax;  /* clamp, reference: [-150, 80] */
  double at_thr;      /* activation threshold, reference: -40 mV */
  int32_t do_reaction;   /* 1: TTP06 step; 0: no TTP06, dV/dt = rate_ext (if given); 2: TTP06 only where
                            rate_ext is NaN, else dV/dt = rate_ext (surrogate form) */
  int32_t n_inst;        /* number of rows of inst_params */
  const void* inst_params;  /* T[n_inst][4]: g_Ks scale, g_to scale, 2 spare (conductivity
                               scaling of an instance is folded into its D6 at set-up) */
  const int64_t* step_ctr;  /* device step counter; time label t = step*dt */
  const void* group_amp;    /* T[n_groups] stimulus amplitude of each group this step [uA/mm^2] */
  void* at;                 /* T[n] first activation time (NaN until crossed) or NULL */
  const void* rate_ext;     /* T[n] extra dV/dt [mV/ms] (surrogate reaction) or NULL */
  const int32_t* cell_slot; /* [n_cells] slot of each compact state entry, ascending, or NULL:
                               the CN step's reaction and final (clamp, activation time)
                               passes then run one lane per cell instead of one per slot
                               (no idle padding lanes); the explicit kernels ignore it */
  int64_t n_cells;
} hp_mono_params;

/* One fused explicit step of the myocardium: TTP06 + 19-point diffusion +
 * stimulus + clamp + first-activation detection.  Reads V_in, writes V_out,
 * updates state (compact SoA [18][ld], see groups above) in place. */
int hp_myo_step_f32(const hp_bricks* g, const float* V_in, float* V_out, float* state,
                    int64_t ld, const hp_mono_params* p, void* stream);
int hp_myo_step_f64(const hp_bricks* g, const double* V_in, double* V_out, double* state,
                    int64_t ld, const hp_mono_params* p, void* stream);

/* Multirate substep, fallback part: one full TTP06 step of dt = p->dt for
 * each of the *count slots in list (the fallback voxels compacted by the
 * reaction call), at V^n = V, writing rate[slot] = -I_ion so that the
 * diffusion step that follows (hp_myo_step with do_reaction = 0 and
 * rate_ext = rate) applies it.  max_count bounds *count (sizes the grid).
 * Uses p->dt, ki_base, nai_base and inst_params. */
int hp_fallback_react_f32(const hp_bricks* g, const float* V, float* state, int64_t ld,
                          float* rate, const int32_t* list, const int32_t* count,
                          int64_t max_count, const hp_mono_params* p, void* stream);

/* out = L V on the brick layout (uA/mm^2); out is 0 on non-myocardial slots.
 * Replaces torch.sparse.mm(L_global, V) restricted to the myocardium
 * (purkinje_myocardium_anisotropic_solver.py:1787). */
int hp_stencil19_apply_f32(const hp_bricks* g, const float* V, float* out, double dscale,
                           void* stream);
int hp_stencil19_apply_f64(const hp_bricks* g, const double* V, double* out, double dscale,
                           void* stream);

/* ---------------------------------------------------------------------------
 * Purkinje cable (replaces the Purkinje rows of L_global,
 * purkinje_myocardium_anisotropic_solver.py:678-681, 721-724, and their part in
 * the ionic step, hetero_solver.py:163-168).  Nodes are stored compactly;
 * neighbours in CSR (indptr/indices), all edges weight w_pp.
 * label: bits 0-1 phenotype (always endo in the reference), bits 2-6 stim group.
 * ------------------------------------------------------------------------- */
typedef struct hp_cable {
  int64_t n;
  const int32_t* indptr;   /* [n+1] */
  const int32_t* indices;  /* [nnz] */
  const uint8_t* label;    /* [n] */
  double w;                /* w_pp = sigma_P*pi*R_P^2/dx_p [mS] */
} hp_cable;

int hp_cable_step_f32(const hp_cable* c, const float* V_in, float* V_out, float* state,
                      int64_t ld, const hp_mono_params* p, void* stream);
int hp_cable_step_f64(const hp_cable* c, const double* V_in, double* V_out, double* state,
                      int64_t ld, const hp_mono_params* p, void* stream);

/* hp_myo_step and hp_cable_step of one time step in a single launch (the
 * cable nodes are extra workgroups of the fused myocardium kernel; the two
 * domains share no data within a step, the PMJ exchange follows in
 * hp_step_tail).  Same arguments and results as the two calls, for every
 * do_reaction (0: the cable blocks join the persistent diffusion kernel).
 * No counterpart in the reference (which assembles both
 * domains into one sparse system, purkinje_myocardium_anisotropic_solver.py:
 * 1787-1791): it saves one kernel boundary per step. */
int hp_myo_cable_step_f32(const hp_bricks* g, const float* V_in, float* V_out, float* state,
                          int64_t ld, const hp_mono_params* p, const hp_cable* c,
                          const float* Vpk_in, float* Vpk_out, float* state_pk, int64_t ld_pk,
                          const hp_mono_params* p_pk, void* stream);
int hp_myo_cable_step_f64(const hp_bricks* g, const double* V_in, double* V_out, double* state,
                          int64_t ld, const hp_mono_params* p, const hp_cable* c,
                          const double* Vpk_in, double* Vpk_out, double* state_pk, int64_t ld_pk,
                          const hp_mono_params* p_pk, void* stream);

/* ---------------------------------------------------------------------------
 * Purkinje-myocardium junctions (replaces the PMJ entries of L_global,
 * purkinje_myocardium_anisotropic_solver.py:684-692).  Junction j couples
 * cable node pk[j] to the myocardial slots patch[ptr[j]..ptr[j+1]) with total
 * conductance c_pmj (c_pmj/|patch| per slot).  Over dt the exchange is
 * integrated exactly (two-compartment exponential relaxation of the cable
 * node against the patch mean, plus relaxation of the patch deviations),
 * conserving charge Ai_P*Vp + Ai_M*sum(Vm).  Increments on slots shared by
 * several junctions are summed in a fixed order (slot-major CSR
 * tgt_ptr/tgt_entry over the patch entries), so results are deterministic.
 * scratch: T[n_entries + n_junctions].
 * ------------------------------------------------------------------------- */
typedef struct hp_pmj {
  int64_t n_junctions;
  int64_t n_entries;
  const int32_t* pk;        /* [n_junctions] cable node */
  const int32_t* ptr;       /* [n_junctions+1] */
  const int32_t* patch;     /* [n_entries] myocardial slot */
  int64_t n_targets;
  const int32_t* tgt_slot;  /* [n_targets] distinct slots */
  const int32_t* tgt_ptr;   /* [n_targets+1] */
  const int32_t* tgt_entry; /* entries of each target slot */
  double c_pmj;             /* mS */
  double ai_p, ai_m, cm;    /* mm^2, mm^2, uF/mm^2 */
  /* optional [n_junctions][4] per-junction constants for the step dt:
   * Cp/(Cp+Cq), exp(-c dt (1/Cp + 1/Cq)), exp(-(c/n) dt/(Ai_M Cm)), 1/n
   * (Cp = Ai_P Cm, Cq = n Ai_M Cm, n = patch size); NULL = computed per step */
  const double* coef;
  /* 1 when no two entries name the same slot (every target has one entry):
   * hp_step_tail then updates the patch voxels in the junction pass, without
   * the scratch array and the target pass (same arithmetic, same result) */
  int32_t exclusive;
  int32_t pad_;
} hp_pmj;

/* Optional first-activation bookkeeping for the junction voxels: these carry
 * label bit 7, the fused myocardium step skips them, and the exchange tests
 * V_prev (V^n) against the final V^{n+1} instead (NULL at = off). */
typedef struct hp_pmj_at {
  void* at;                 /* T[n_slots] */
  const void* V_prev;       /* T[n_slots], V^n */
  const int64_t* step_ctr;
  double dt, at_thr;
} hp_pmj_at;

int hp_pmj_couple_f32(const hp_pmj* j, float* V_pk, float* V_myo, float* scratch, double dt,
                      double vmin, double vmax, const hp_pmj_at* at, void* stream);
int hp_pmj_couple_f64(const hp_pmj* j, double* V_pk, double* V_myo, double* scratch, double dt,
                      double vmin, double vmax, const hp_pmj_at* at, void* stream);

/* ---------------------------------------------------------------------------
 * Crank-Nicolson step of the coupled system: the reference's scheme
 * (purkinje_myocardium_anisotropic_solver.py:757-849, loop :1746-1791):
 *   V*  = V^n - dt*I_ion + (dt/Cm)*I_stim                (Lie splitting)
 *   ((Cm/dt) I - theta L) V^{n+1} = (Cm/dt) V* + (1-theta) L V*
 *   V^{n+1} = clamp(V^{n+1}, vmin, vmax)
 * with L = D^-1 K of the whole system (19-point stencil on the myocardial
 * rows, cable edges and PMJ couplings, :606-746).  The linear system is solved
 * matrix-free by Jacobi sweeps from x0 = V* (the junction nodes' rows
 * re-solved after each sweep from their patches' new values), stopping when ||b - A x|| <=
 * tol*||b|| (the rule of the reference's _bicgstab, :786-849, which this
 * replaces) or after max_sweeps sweeps; the iterate returned is the one after
 * the sweep whose input met the rule.  All decisions stay on the device (no
 * host synchronisation), so a step can be captured in a HIP graph.
 * hp_cn_step replaces the reaction + diffusion part of one step (the
 * explicit path's hp_myo_cable_step); hp_step_tail (without junctions, j =
 * NULL) then advances the step counter and the stimulus, as for that path.
 * ------------------------------------------------------------------------- */
/* PMJ couplings as rows of L (purkinje_myocardium_anisotropic_solver.py:684-692,
 * 721-724): each (junction node p, patch voxel m) pair has conductance
 * w = c_pmj/|patch|.  Myocardial side, per target slot (label bit 7, the slots
 * listed in ascending order): the cable node of each coupling and w/Ai_M.
 * Cable side (node_ptr may be NULL: no coupling), per cable node: the coupled
 * slots and w/Ai_P.  jnode lists the cable nodes that have couplings (their
 * rows are re-solved after each sweep from the patch's new values). */
typedef struct hp_cn_pmj {
  int64_t n_targets;
  const int32_t* tgt_slot;   /* [n_targets], ascending */
  const int32_t* tgt_ptr;    /* [n_targets+1] */
  const int32_t* tgt_node;   /* cable node of each coupling */
  const double* tgt_w;       /* w/Ai_M of each coupling */
  const int32_t* node_ptr;   /* [n_cable+1] or NULL */
  const int32_t* node_slot;  /* myocardial slot of each coupling */
  const double* node_w;      /* w/Ai_P of each coupling */
  int64_t n_jnodes;
  const int32_t* jnode;      /* [n_jnodes] distinct cable nodes with couplings */
} hp_cn_pmj;

/* solver 0: Jacobi sweeps x <- x + (b - A x)/diag(A) to the stopping rule
 * ||b - A x|| <= tol ||b|| (an approximation of the exact CN solution well
 * inside the tolerance); solver 1: the reference's Jacobi-preconditioned
 * BiCGSTAB (:786-849) with its stopping rules and breakdown guards, which
 * reproduces the reference's iterates (its error against the exact solution
 * included).  max_sweeps bounds the sweeps (BiCGSTAB iterations) per step.
 * scalars (device, >= 24 doubles): 0 ||b||^2, 1 ||r||^2 last evaluated,
 * 2 done flag, 3 sweeps (iterations) of the current step, 4 steps that stopped
 * at max_sweeps (BiCGSTAB: or at a breakdown) without meeting the rule, 5 the
 * largest ||r||/||b|| of those, 6 total sweeps (iterations), 7 steps solved,
 * 16 the largest sweep (iteration) count of any step (4-7 and 16 accumulate
 * over calls; the caller zeroes them); 8-15 BiCGSTAB state; 17 the warm
 * start's step count (hp_cn_work.warm); 18 (solver 0) the largest sweep count
 * since the caller last zeroed it (for a caller that bounds the sweeps of a
 * batch of steps by what recent steps needed and repeats the batch with a
 * higher bound when a step stopped at it: slot 4 tells).
 * part: device, doubles, part_len >= 2*G (solver 0) or 8*G (solver 1) with
 * G = min(n_bricks, grid_cap) + ncb the workgroups of a pass (ncb <= 520 for
 * the cable rows). */
typedef struct hp_cn_params {
  double cm_dt;        /* Cm/dt */
  double theta;        /* 0 <= theta <= 1 (reference default 0.5) */
  double dscale;       /* 1/(A_M h^2): L of the myocardial rows = dscale * stencil sum */
  double dscale_pk;    /* 1/Ai_P: L of the cable rows = dscale_pk * w_pp * sum */
  double tol;          /* relative residual of the stopping rule (reference cg_tol) */
  int32_t max_sweeps;  /* >= 1, sweeps (iterations) per step (launched every step; idle after done) */
  int32_t solver;      /* 0 Jacobi sweeps, 1 BiCGSTAB */
  double* scalars;
  double* part;
  int64_t part_len;    /* doubles in part */
  int64_t grid_cap;    /* brick workgroups of a pass at most (0: 8192) */
} hp_cn_params;

/* Work vectors (device, the arithmetic type of the call): Vs, b, x of n_slots
 * entries; Vs_pk, b_pk, x_pk of n_cable entries (may be NULL without a cable).
 * Solver 0 uses Vs, b, x and V_out as the second iterate buffer; solver 1
 * uses Vs (the iterate) and kr[8] = r, r^, M^-1, t, p (two buffers), v (two
 * buffers), n_slots entries each, kr_pk[8] likewise over the cable nodes.
 * warm (solver 0, optional; no counterpart in the reference, whose BiCGSTAB
 * starts from V*): d, e over the slots and d_pk, e_pk over the cable nodes,
 * zeroed by the caller before the first step.  The sweeps then start from
 * x_0 = V* + e, e the last step's increment x - V* (d) extrapolated linearly
 * from the last two steps; the stopping rule is unchanged.  scalars[17]
 * counts the steps begun warm (capped at 2; zero it with d and e). */
typedef struct hp_cn_work {
  void* Vs;
  void* b;
  void* x;
  void* Vs_pk;
  void* b_pk;
  void* x_pk;
  void* kr[8];
  void* kr_pk[8];
  void* warm[4];
} hp_cn_work;

/* p / p_pk: dt, inv_cm, ki/nai bases, instance scales, group_amp (stimulus of
 * this step), step_ctr, vmin/vmax, at (first activation times, optional) and
 * at_thr, as for hp_myo_step; do_reaction 1 (the cell model everywhere) or 2
 * (the multirate forms: dV/dt = rate_ext where it is a number, the full cell
 * step where it is NaN; the cable always runs the cell model).  c may be NULL
 * (no cable); j may be NULL (no junctions).  Reads V_in / Vpk_in (V^n), writes
 * V_out / Vpk_out (V^{n+1}), updates the cell states in place. */
int hp_cn_step_f32(const hp_bricks* g, const float* V_in, float* V_out, float* state, int64_t ld,
                   const hp_mono_params* p, const hp_cable* c, const float* Vpk_in,
                   float* Vpk_out, float* state_pk, int64_t ld_pk, const hp_mono_params* p_pk,
                   const hp_cn_pmj* j, const hp_cn_params* cp, const hp_cn_work* w,
                   void* stream);
/* One Jacobi sweep of the myocardial rows alone (no cable, no junctions):
 * x_out = x_in + (b - A x_in)/diag(A) with the step's kernel, the residual's
 * block partial sums into cp->part; cp->scalars is not used.  No counterpart in
 * the reference: the solve's dominant kernel by itself, for measurement. */
int hp_cn_sweep_f32(const hp_bricks* g, const float* x_in, float* x_out, const float* b,
                    const hp_cn_params* cp, void* stream);
/* One BiCGSTAB product pass of the myocardial rows alone (no cable, no
 * junctions): t = A M^-1 (r - alpha v) with the step's kernel, from the work
 * vectors r = kr[0], M^-1 = kr[2], v = kr[6] into t = kr[3], the partial sums
 * of s.s, t.t, t.s into cp->part (>= 3*min(n_bricks, 8192) doubles);
 * cp->scalars: 16 doubles with entry 2 (done) 0 and entry 10 = alpha.  No
 * counterpart in the reference: the BiCGSTAB solve's stencil kernel by itself,
 * for measurement. */
int hp_cn_apply_f32(const hp_bricks* g, const hp_cn_params* cp, const hp_cn_work* w,
                    void* stream);
int hp_cn_step_f64(const hp_bricks* g, const double* V_in, double* V_out, double* state,
                   int64_t ld, const hp_mono_params* p, const hp_cable* c, const double* Vpk_in,
                   double* Vpk_out, double* state_pk, int64_t ld_pk, const hp_mono_params* p_pk,
                   const hp_cn_pmj* j, const hp_cn_params* cp, const hp_cn_work* w,
                   void* stream);

/* ---------------------------------------------------------------------------
 * Stimulus schedule on the device (replaces _stim_active, :1290-1301, and the
 * host loop :1754-1759).  site_params[s*5 + 0..4] = onset, dur, period (0 =
 * single pulse), n_pulses (< 0 = unlimited), amp [uA/mm^2].  Stimulus group
 * g (1 <= g < n_groups <= 8192; group 0 = none) is the set of sites
 * group_sites[2g] + b for every bit b of the mask group_sites[2g+1];
 * group_amp[g] receives the summed amplitude of its active sites.  A voxel's
 * group is its label bits 2-6, or (inst << 3) | (bits 2-4) when the brick grid
 * carries instance ids (ensembles: up to 7 site combinations per instance).
 * hp_step_advance increments *step_ctr and writes group_amp for the new time
 * step*dt; call hp_stim_eval once after setting the counter.
 * ------------------------------------------------------------------------- */
int hp_stim_eval_f32(const double* site_params, int32_t n_sites, const uint32_t* group_sites,
                     int32_t n_groups, const int64_t* step_ctr, double dt, float* group_amp,
                     void* stream);
int hp_stim_eval_f64(const double* site_params, int32_t n_sites, const uint32_t* group_sites,
                     int32_t n_groups, const int64_t* step_ctr, double dt, double* group_amp,
                     void* stream);
int hp_step_advance_f32(const double* site_params, int32_t n_sites, const uint32_t* group_sites,
                        int32_t n_groups, int64_t* step_ctr, double dt, float* group_amp,
                        void* stream);
int hp_step_advance_f64(const double* site_params, int32_t n_sites, const uint32_t* group_sites,
                        int32_t n_groups, int64_t* step_ctr, double dt, double* group_amp,
                        void* stream);

/* End of a time step in one launch: the PMJ exchange (if j is non-NULL and
 * c_pmj != 0) followed by the step-counter advance and the stimulus
 * amplitudes of the next step.  Same results as hp_pmj_couple_* followed by
 * hp_step_advance_*; trees with more than 65536 junctions take that path. */
int hp_step_tail_f32(const hp_pmj* j, float* V_pk, float* V_myo, float* scratch, double dt,
                     double vmin, double vmax, const hp_pmj_at* at, const double* site_params,
                     int32_t n_sites, const uint32_t* group_sites, int32_t n_groups,
                     int64_t* step_ctr, float* group_amp, void* stream);
int hp_step_tail_f64(const hp_pmj* j, double* V_pk, double* V_myo, double* scratch, double dt,
                     double vmin, double vmax, const hp_pmj_at* at, const double* site_params,
                     int32_t n_sites, const uint32_t* group_sites, int32_t n_groups,
                     int64_t* step_ctr, double* group_amp, void* stream);

/* Gather slot values into reference node order: out[i] = src[slot_of_node[i]]
 * (Vm snapshots, vm_snapshots.npz schema, :1975-1984). */
int hp_gather_f32(const float* src, const int32_t* idx, float* out, int64_t n, void* stream);
int hp_gather_f64(const double* src, const int32_t* idx, double* out, int64_t n, void* stream);

/* dst[0:n] = src[0:n], a plain streaming copy (n a multiple of 4, 16-byte
 * aligned pointers).  No counterpart in the reference: the calibration point
 * for the HBM bandwidth and the traffic counters of the kernels above. */
int hp_stream_copy_f32(const float* src, float* dst, int64_t n, void* stream);


/* ---------------------------------------------------------------------------
 * Torso Poisson forward solve (replaces _assemble_stiffness's operator,
 * build_torso_grid.py:276-397 with the pin of compute_bspm.py:327-331, and
 * the _pcg_gpu solve, compute_bspm.py:164-263).
 *
 * Column-run layout of the (ni,nj,nk) box: column (i,j) stores the z-run
 * [col_k0, col_k0+col_len) at cells col_off.. (holes inside a run have zero
 * conductance and diag 1).  gx/gy/gz: conductance of the +x/+y/+z face of
 * each stored cell, harm(sigma_a, sigma_b)/h^2 [S/m^3]; diag: sum of the six
 * face conductances; row `pin` is the identity.
 * ------------------------------------------------------------------------- */
typedef struct hp_torso {
  int32_t ni, nj, nk, pad_;
  int64_t n_cells;
  const int32_t* col_k0;   /* [ni*nj] */
  const int32_t* col_len;  /* [ni*nj] */
  const int32_t* col_off;  /* [ni*nj] */
  const float* gx;
  const float* gy;
  const float* gz;
  const float* diag;
  int64_t pin;             /* stored index of the pinned cell, or -1 */
  const int32_t* chunks;   /* chunk table (device memory), from hp_torso_chunks */
  int64_t n_chunks;
} hp_torso;

/* The kernels walk the stored cells in chunks of up to 64 consecutive cells
 * of one column run, one wave per chunk.  Each chunk is a record of
 * HP_TORSO_CHUNK_INTS int32: k of its first cell, its cell count, i, j, then
 * for the own column and the +x, -x, +y, -y neighbour columns the index base
 * (stored index of cell k = base + k; 5 values) and the stored k-range
 * lo | hi << 16 (5 values; 0 for a column outside the box or without cells),
 * then two zeros.  Chunks follow the stored order.
 * hp_torso_chunks fills the table on the host from the column arrays (host
 * memory); with chunks == NULL it only counts.  Requires nk <= 32767.
 * Returns 0, or -1 if the capacity is too small or an argument is bad.
 * No counterpart in the reference (a layout helper of this library). */
#define HP_TORSO_CHUNK_INTS 16
int hp_torso_chunks(int32_t ni, int32_t nj, int32_t nk, const int32_t* col_k0,
                    const int32_t* col_len, const int32_t* col_off, int32_t* chunks,
                    int64_t capacity, int64_t* n_chunks);

/* Reduction workspace: part >= 3*4096 doubles (block partial sums; 6*4096 for
 * the linf and p99 metrics of hp_pcg_begin_metric_f32), counter: not used (kept
 * for the layout; may be NULL), out >= 2 doubles, scalars >= 16 doubles (PCG
 * state: rz, pAp, alpha, beta, rr0, rr, tol^2, done, iter, maxit, ||z||^2,
 * mode, rows, the metric's last value). */
typedef struct hp_pcg_ws {
  double* part;
  int* counter;
  double* out;
  double* scalars;
} hp_pcg_ws;

/* y = A x in FP32. */
int hp_poisson_spmv_f32(const hp_torso* t, const float* x, float* y, void* stream);
/* r = b - A x in FP64; ws->out[0] = ||r||^2, ws->out[1] = ||b||^2. */
int hp_poisson_residual_f64(const hp_torso* t, const double* b, const double* x, double* r,
                            hp_pcg_ws* ws, void* stream);
/* hp_poisson_residual_f64 without a workspace argument and without any
 * allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  out >= 2 doubles (device): ||r||^2,
 * ||b||^2.  The block partial sums go to a scratch array inside the library,
 * 2*4096 doubles (64 KiB) per GPU, held until the process ends; every call
 * writes the partials it reads, so nothing needs clearing between calls.
 * Calls that overlap on one GPU (two streams, two host threads) share that
 * array and race: serialise them.  Same results as hp_poisson_residual_f64. */
int hp_poisson_residual_f64_noalloc(const hp_torso* t, const double* b, const double* x,
                                    double* r, double* out, void* stream);
/* Jacobi-PCG for A d = r (FP32 vectors, FP64 inner products), d = 0 initially.
 * begin: sets p = M^-1 r and the device-side state; iterate: n_iter iterations,
 * each a no-op once converged or max_iter was reached.  mode 0 stops at
 * ||r|| <= tol*||r0||; mode 1 at rms(M^-1 r) <= tol (the reference's 'rms'
 * metric, compute_bspm.py:222-256).  ws->scalars then holds rz, pAp, alpha,
 * beta, ||r0||^2, ||r||^2, tol^2, done, iterations, max_iter, ||M^-1 r||^2.
 * r is overwritten: begin replaces it by z = M^-1 r, and the iterations carry
 * z (not r) in that array. */
int hp_pcg_begin_f32(const hp_torso* t, float* r, float* d, float* p, hp_pcg_ws* ws,
                     double tol, int max_iter, int mode, void* stream);
int hp_pcg_iterate_f32(const hp_torso* t, float* r, float* d, float* p, float* Ap,
                       hp_pcg_ws* ws, int n_iter, void* stream);
/* hp_pcg_begin_f32 with every stopping metric of _pcg_gpu (compute_bspm.py:
 * 164-263) on z = M^-1 r over the n_rows rows of the reference system (the
 * stored cells minus the holes of the column runs): mode 1 rms(z) < tol,
 * 2 max|z| < tol, 3 the 99th percentile of |z| (linear interpolation, as
 * torch.quantile) < tol, each tested after every iteration and never before
 * the first; mode 0 as in hp_pcg_begin_f32.  scalars[13] holds the last value
 * of the rms / linf statistic (p99: evaluate the quantile from z afterwards).
 * The p99 test needs no sort: from the count of |z| >= tol and the values
 * next to tol on either side it decides exactly where the quantile falls. */
int hp_pcg_begin_metric_f32(const hp_torso* t, float* r, float* d, float* p, hp_pcg_ws* ws,
                            double tol, int max_iter, int mode, int64_t n_rows, void* stream);
/* ---------------------------------------------------------------------------
 * Geometric multigrid preconditioner for the torso operator.  No counterpart
 * in the reference, whose _pcg_gpu preconditions with the diagonal
 * (bspm_fn.py:164-263); it solves the same pinned system.  Level 0 is the
 * column-run grid above; level[l] (l = 0..n_levels-1) is the dense box of grid
 * level l+1, from 2x2x2 agglomeration of the level below, with face
 * conductance 1/2 of the sum of the four fine faces it covers.  Cells with
 * inv_diag = 0 (inactive, or holding the pinned cell) stay zero.  One V-cycle:
 * nu damped-Jacobi sweeps (nu even) before and after each coarse correction,
 * restriction = sum of child residuals, prolongation = parent value, exact
 * solve on the coarsest level (<= 1024 cells) with a dense inverse.
 * ------------------------------------------------------------------------- */
typedef struct hp_mg_level {
  int32_t ni, nj, nk, pad_;
  int64_t n;              /* ni*nj*nk, cell (i,j,k) at (i*nj + j)*nk + k */
  const float* gx;        /* conductance of the +x face of each cell (0 = none) */
  const float* gy;
  const float* gz;
  const float* inv_diag;  /* 1/(sum of the six faces); 0 = inactive or holding the pinned cell */
  float* x;               /* work vectors, n each */
  float* tmp;
  float* b;
} hp_mg_level;

#define HP_MG_MAX_LEVELS 12
typedef struct hp_mg {
  int32_t n_levels;       /* dense levels used, 1..HP_MG_MAX_LEVELS */
  int32_t nu;             /* Jacobi sweeps before and after each coarse correction (even) */
  float omega;            /* Jacobi damping */
  int32_t pad_;
  const float* inv_diag0; /* level 0: 1/diag per stored cell; 0 = pinned, hole, isolated */
  float* tmp0;            /* level-0 work vector (n_cells) */
  const float* coarse_inv;  /* [n][n] inverse on the coarsest level's active cells, 0 elsewhere */
  hp_mg_level level[HP_MG_MAX_LEVELS];
} hp_mg;

/* z = B r, one V-cycle (all level work vectors overwritten).  scalars may be
 * NULL; otherwise every kernel returns at once when the PCG done flag is set. */
int hp_mg_vcycle_f32(const hp_torso* t, const hp_mg* mg, const float* r, float* z,
                     const double* scalars, void* stream);
/* PCG for A d = r with the V-cycle as preconditioner (FP32 vectors, FP64 inner
 * products), stopping at ||r|| <= tol*||r0||: the same protocol and scalar
 * slots as hp_pcg_begin/iterate (mode 0); z holds the preconditioned residual. */
int hp_mgpcg_begin_f32(const hp_torso* t, const hp_mg* mg, const float* r, float* d, float* z,
                       float* p, hp_pcg_ws* ws, double tol, int max_iter, void* stream);
int hp_mgpcg_iterate_f32(const hp_torso* t, const hp_mg* mg, float* r, float* d, float* z,
                         float* p, float* Ap, hp_pcg_ws* ws, int n_iter, void* stream);

/* Iterative refinement helpers: x += scale*d (FP64), r32 = scale*r64. */
int hp_refine_axpy_f64(const hp_torso* t, double* x, const float* d, double scale, void* stream);
int hp_cast_scale_f32(const hp_torso* t, const double* r64, float* r32, double scale,
                      void* stream);

/* Heart source (replaces build_heart_laplacian + solve_frame's source,
 * compute_bspm.py:88-119, 381-419): lap = vscale * (6-neighbour unit-weight
 * graph Laplacian of V) on the brick layout (FP64 out); then
 * b[rows[q]] = coeff * sum_{e in ptr[q]..} (lap[fine[e]] - mean) / n_expected
 * with mean over valid_slots, times psi[rows[q]] when psi (the smoothed-
 * boundary domain function per stored cell, :413-414) is given, and
 * b[pin] = 0.  Rows not listed are untouched. */
int hp_heart_lap6_f32(const hp_bricks* g, const float* V, double* lap, double vscale,
                      void* stream);
int hp_heart_lap6_f64(const hp_bricks* g, const double* V, double* lap, double vscale,
                      void* stream);

/* Lead-field ECG (no counterpart in the reference, which solves the torso for
 * every frame): the electrode potentials are a fixed linear functional of Vm,
 * phi[e] = sum_r W[r][e] Vm[slot[r]] (reciprocity with the pinned operator),
 * so out[f*ne + e] = sum_r W[r*ne + e] * V[f*ldv + row_slot[r]] for n_frames
 * (<= 16) frames, FP32 products summed in FP64 in a fixed order (FP64 MFMA
 * over row blocks, then the blocks).  part holds part_len >= n_frames*ne
 * doubles per row block; the row count is split into min(4096,
 * part_len/(n_frames*ne)) blocks. */
int hp_leadfield_apply_f32(const float* W, const int32_t* row_slot, int64_t n_rows, int ne,
                           const float* V, int64_t ldv, int n_frames, double* part,
                           int64_t part_len, double* out, void* stream);
int hp_bspm_source_f64(const double* lap, const int32_t* valid_slots, int64_t n_valid,
                       const int32_t* rows, const int32_t* ptr, const int32_t* fine,
                       int64_t n_rows_hit, double coeff, double n_expected, const double* psi,
                       double* b, int64_t pin, hp_pcg_ws* ws, void* stream);
/* Electrode potentials, FP64: out[e] = sum_k w[k]*phi[idx[k]], k in [ptr[e], ptr[e+1]).
 * No counterpart in the reference (it outputs raw surface potentials only). */
int hp_leads_accumulate_f64(const double* phi, const int32_t* ptr, const int32_t* idx,
                            const double* w, double* out, int64_t n_electrodes, void* stream);

/* ---------------------------------------------------------------------------
 * Threshold crossings of Vm snapshots (replaces the host tables of
 * compute_activation_maps, activation_maps.py:63-413, with the dV/dt gate of
 * anisotropic_simulation/functions/activation_maps.py:105-119).
 * Vm: [n_frames][n] FP32 mV on the device, t: [n_frames] FP64 ms.  rising = 1
 * counts upward crossings of `threshold` (activations), 0 downward ones.  A
 * crossing whose |dV/dt| over its frame interval is below min_vel (mV/ms; <= 0:
 * no gate) is dropped, then one closer than min_gap (ms; <= 0: none) to the
 * previous kept event of the node.  counts[i] (optional) receives the number
 * of kept events of node i; table (optional) [n][max_events] their times in
 * order, NaN padded (events beyond max_events are counted, not stored).
 * ------------------------------------------------------------------------- */
int hp_crossing_events_f32(const float* Vm, const double* t, int64_t n_frames, int64_t n,
                           double threshold, int rising, double min_vel, double min_gap,
                           int32_t* counts, float* table, int32_t max_events, void* stream);

/* ---------------------------------------------------------------------------
 * Learned ionic surrogate (no counterpart in the reference).  An MLP
 * 22 -> 128 -> 128 -> 128 -> 19 (SiLU), BF16 weights, FP32 accumulation on
 * MFMA, predicts the increments of V and of the 18 TTP06 variables over one
 * macro-step of 10 reaction substeps.
 *
 * hp_surrogate_step takes the weights in a scaled form that saves one
 * multiplication per activation: with c = log2(e), w1 and b1, b2, b3 are
 * multiplied by -c and w4 by -1/c (= -ln 2); b4 is unchanged.  Each hidden
 * layer then stores g = u/(1 + 2^u) = -c*SiLU(z) for its scaled
 * pre-activation u = -c*z, which the next layer's unscaled weights consume.
 * ------------------------------------------------------------------------- */
typedef struct hp_mlp_weights {
  const void* w1;  /* bf16 [128][22]  (out, in) */
  const void* w2;  /* bf16 [128][128] */
  const void* w3;  /* bf16 [128][128] */
  const void* w4;  /* bf16 [19][128]  */
  const float* b1; const float* b2; const float* b3; const float* b4;
} hp_mlp_weights;

/* Feature normalisation: x_in = (f(x) - in_mean)/in_std * ... with f = log for
 * in_log features; outputs y = out_mean + out_std * net (increments of f(x)). */
typedef struct hp_mlp_norm {
  float in_mean[22], in_rstd[22];
  int32_t in_log[22];
  float out_mean[19], out_std[19];
} hp_mlp_norm;

typedef struct hp_surrogate_params {
  float t_macro;        /* macro-step length, ms (10 substeps) */
  float dvdt_max;       /* fallback if |dV/dt| above, mV/ms (reference choice: 1) */
  float v_lo, v_hi;     /* fallback if v_lo < V < v_hi (-70, -40) */
  float ki_base, nai_base;
  int32_t force_fallback;
  int32_t variant;      /* inference kernel: 0 = 16 voxels per wave, 768 threads;
                           1 = 32 voxels per wave, 512 threads; 2 = 32 voxels per
                           wave, 1024 threads */
  /* Optional compaction of the fallback voxels (both NULL = off): their slots
   * are appended to fb_list (capacity n_slots) and counted in *fb_count, which
   * the caller zeroes before the reaction call.  List order is unspecified. */
  int32_t* fb_list;
  int32_t* fb_count;
  /* Optional list of the slots the inference kernel visits (NULL = every
   * slot): n_list entries, a multiple of 32, -1 = no voxel.  Listing only the
   * tissue slots skips the padding of partly filled bricks. */
  const int32_t* slot_list;
  int64_t n_list;
  /* Optional, with slot_list: the compact state index of each listed slot
   * (any value at -1 entries), which spares the kernel the group-table
   * lookup; NULL = looked up in g->groups. */
  const int32_t* state_list;
  /* 0: the network's output is the macro-step increment itself.  1: it is a
   * correction of the big step: hp_bigstep_reaction ran first in this
   * macro-step (state advanced by one t_macro step, rate = -I_ion, V_prev
   * updated, fallback voxels listed and given rate = NaN), and the network,
   * fed V and the advanced state, adds its output to that state and rate.
   * The fallback voxels are then the ones whose rate is NaN; V_prev and the
   * fallback list are left to the big step. */
  int32_t correct;
  /* Correction mode only: voxels at rest, V < rest_v and |rate| < rest_rate
   * (the big step's rate), keep the big step alone, whose error there is
   * below the network's; a wave tile of such voxels skips the network.
   * rest_rate = 0: the network corrects every non-fallback voxel. */
  float rest_v, rest_rate;
  /* Correction mode only: a voxel whose V correction exceeds corr_max mV in
   * magnitude keeps the big step alone (a trust bound on the network; the big
   * step's own error stays below ~0.1 mV per 0.2 ms on the plateau). */
  float corr_max;
} hp_surrogate_params;

/* Training primitives.  C[m,n] = alpha*sum_k A(m,k)B(k,n) + beta*C[m,n] (+ bias[n]),
 * A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn], bf16 operands, FP32 C.
 * ksplit > 1 splits K over that many workgroup slices that add into C
 * atomically (C zeroed by the caller, beta = 0, no bias). */
int hp_gemm_bf16(int M, int N, int K, const void* A, int64_t sam, int64_t sak, const void* B,
                 int64_t sbk, int64_t sbn, float* C, int64_t ldc, float alpha, float beta,
                 const float* bias, int ksplit, void* stream);
/* In the helpers below the FP32 gradient copies dxf and dYf are optional (NULL:
 * not written); every other pointer is required when the size is non-zero
 * (NULL returns -1). */
int hp_silu_bf16(const float* x, void* y, int64_t rows, int cols, int ld_in, int ld_out,
                 void* stream);
int hp_silu_bwd_bf16(const float* x, const float* dy, void* dx, float* dxf, int64_t n,
                     void* stream);
int hp_cast_bf16(const float* x, void* y, int64_t n, void* stream);
/* masked MSE over the first `used` of `cols` columns; loss accumulated into *loss. */
int hp_mse_grad(const float* Y, const float* T, void* dY, float* dYf, int rows, int cols,
                int used, float* loss, void* stream);
int hp_colsum(const float* X, int rows, int cols, int ld, float* out, void* stream);
int hp_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1,
            float b2, float eps, int step, float wd, void* stream);

/* One surrogate macro-step decision + update over the brick layout (FP32
 * engine state): per voxel, fallback if |V - V_prev|/t_macro > dvdt_max, or
 * v_lo < V < v_hi, or V_prev is NaN (first call); fallback voxels get
 * rate = NaN (the fused step then runs the full TTP06 for them), the others
 * get their 18 state variables advanced by the network and rate = dV/t_macro
 * (applied by hp_myo_step with do_reaction = 2).  V_prev <- V.  With
 * sp->correct = 1 the network corrects the big step that hp_bigstep_reaction
 * has just taken (see hp_surrogate_params.correct). */
int hp_surrogate_step(const hp_bricks* g, const float* V, float* V_prev, float* state,
                      int64_t ld, float* rate, const hp_mlp_weights* w, const hp_mlp_norm* n,
                      const hp_surrogate_params* sp, int n_blocks, void* stream);

/* The same decision rule without the network (the multirate comparison form):
 * non-fallback voxels take one t_macro Rush-Larsen/forward-Euler TTP06 step
 * and rate = -I_ion; fallback voxels get rate = NaN.  V_prev <- V. */
int hp_bigstep_reaction(const hp_bricks* g, const float* V, float* V_prev, float* state,
                        int64_t ld, float* rate, const hp_surrogate_params* sp, void* stream);

/* ---------------------------------------------------------------------------
 * Geometry set-up: the transmural coordinate and the
 * rule-based fibre field.
 *
 * Laplace solve of generate_transmural_coordinate.py:35-186 on a brick grid
 * of the myocardium: label bit 0 set = a Dirichlet row (identity; its value
 * is already in x), otherwise the 6-neighbour graph Laplacian (Neumann at the
 * mask boundary); label 0xFF = padding; D6 rows 0-2 = 1 and rows 3-5 = 0 on
 * the tissue slots.  CG in FP32 with FP64 inner products and the reference's
 * rule ||r|| / ||r0|| < tol, its decisions on the device: begin forms r and p
 * from x; iterate launches n_iter iterations (4 kernels each; those after
 * the stop return at once) and advances `launched`.  scalars (device, >= 8
 * doubles): 0 r.r, 1 ||r0||, 2 alpha, 3 beta, 4 done (1 converged, 2 stopped
 * at max_iter), 5 iterations, 6 last ||r||/||r0||.  part: device, >= n_bricks
 * doubles.  Replaces _conjugate_gradient and the torch sparse Laplacian of
 * generate_transmural_coordinate.py:35-78, 85-183.
 * ------------------------------------------------------------------------- */
typedef struct hp_lap_cg {
  float* x;          /* n_slots: the iterate */
  float* r;          /* n_slots */
  float* p[2];       /* n_slots each: the direction, alternating per iteration */
  float* ap;         /* n_slots */
  double* scalars;
  double* part;
  int64_t part_len;
  double tol;
  int64_t launched;  /* iterations launched since begin (host side) */
} hp_lap_cg;
int hp_laplace_cg_begin_f32(const hp_bricks* g, hp_lap_cg* cg, void* stream);
int hp_laplace_cg_iterate_f32(const hp_bricks* g, hp_lap_cg* cg, int32_t n_iter,
                              int64_t max_iter, void* stream);

/* Rule-based fibre frame (generate_fibre_direction.py:135-281) on a dense
 * nx*ny*nz box (C order) of the filled phi and psi fields (the nearest
 * in-mask value in a 2-voxel halo, 0 beyond: _fill_nan_nearest, :90-111):
 * gradients by np.gradient's rule at spacings (hx, hy, hz), e_t, e_l, e_c,
 * the helix angle alpha_endo + (alpha_epi - alpha_endo) phi (radians) and
 * f = cos(a) e_c + sin(a) e_l, normalised, into f[3*idx + c], in FP32 in the
 * reference's order of operations.  Every side >= 2. */
int hp_fibre_rule_f32(const float* phi_filled, const float* psi_filled, int32_t nx, int32_t ny,
                      int32_t nz, double hx, double hy, double hz, double alpha_endo,
                      double alpha_epi, float* f, void* stream);

/* ---------------------------------------------------------------------------
 * Heart-to-torso registration: the voxel Dice of the 48 signed permutations
 * of the rotation search (register_heart_to_torso.py:146-224, its Dice
 * _voxel_dice :109-138).  Candidate r is s_i * src_c[:, perm_i] + c_tgt[i]
 * with perm_sign[6r..6r+5] = {perm_0, perm_1, perm_2, s_0, s_1, s_2}; it and
 * the target tgt (n_tgt x 3) are voxelised at floor((x - mn[3r+a]) / vox_mm)
 * on a dims[3r..3r+2] lattice whose bits live at bits[word_off[r]] (candidate,
 * then target, (word_off[r+1] - word_off[r]) / 2 words each; word_off[n_perm]
 * <= n_words).  counts (device, 3 n_perm + 1 uint64): per permutation the
 * occupied voxels of the candidate, of the target and of both; the last
 * entry counts points outside their lattice (0 for correct mn / dims, and
 * such points are never written).  FP64 arithmetic rounded as numpy's.
 * ------------------------------------------------------------------------- */
int hp_register_dice_f64(const double* src_c, int64_t n_src, const double* tgt, int64_t n_tgt,
                         const int32_t* perm_sign, int32_t n_perm, const double* c_tgt,
                         const double* mn, const int32_t* dims, const int64_t* word_off,
                         double vox_mm, uint32_t* bits, int64_t n_words,
                         unsigned long long* counts, void* stream);

/* ---------------------------------------------------------------------------
 * Coarse torso grid on the device (build_torso_grid.py).  Dense volumes are
 * C order (n0, n1, n2); flags are uint8 0/1.
 * ------------------------------------------------------------------------- */
/* Feature transform: for every voxel the coordinates of its nearest voxel
 * with mask == feature_value (Euclidean, voxel spacing `spacing`, a host
 * array of 3), into ft[3][N] int32 (-1 where there is none), by lower
 * envelopes along axis 0, then 1, then 2 (Maurer, Qi, Raghavan 2003).  The
 * nearest-feature search of scipy.ndimage.distance_transform_edt, which
 * compute_smoothed_domain (:169-171) and extend_conductivity (:202) call.
 * ft_tmp: 3N int32; stack: stack_len int32 (>= the longest side; lines are
 * processed in batches of stack_len / side). */
int hp_feature_transform(const uint8_t* mask, uint8_t feature_value, int32_t n0, int32_t n1,
                         int32_t n2, const double* spacing, int32_t* ft, int32_t* ft_tmp,
                         int32_t* stack, int64_t stack_len, void* stream);
/* psi = (1 + tanh(3 d / width_mm)) / 2 of the signed distance d (+ inside,
 * mm) from the two feature transforms (nearest outside voxel, nearest inside
 * voxel): compute_smoothed_domain :169-173. */
int hp_sbm_psi_f64(const uint8_t* inside, const int32_t* ft_nearest_out,
                   const int32_t* ft_nearest_in, int32_t n0, int32_t n1, int32_t n2,
                   const double* spacing, double width_mm, double* psi, void* stream);
/* Linear zoom of a dense FP64 volume to oshape (host int32[3], corner to
 * corner, 0 beyond the input's edges), cropped / padded with 0 to mshape
 * (host int32[3]), values below cutoff set to 0 (compute_smoothed_domain
 * :175-186, scipy.ndimage.zoom order 1); positive (optional) = out > 0. */
int hp_zoom_linear_f64(const double* in, int32_t n0, int32_t n1, int32_t n2, const int32_t* oshape,
                       const int32_t* mshape, double cutoff, double* out, uint8_t* positive,
                       void* stream);
/* Nearest-neighbour zoom of a label volume to oshape (host int32[3]):
 * _resample_labels :260-269 (scipy.ndimage.zoom order 0). */
int hp_zoom_nearest_i32(const int32_t* in, int32_t n0, int32_t n1, int32_t n2,
                        const int32_t* oshape, int32_t* out, void* stream);
/* sigma[v] = vals[q] for keys[q] == labels[v] (0 without an entry), tissue =
 * label != 0: build_torso_grid :554-576. */
int hp_label_sigma_f64(const int32_t* labels, int64_t n, const int32_t* keys, const double* vals,
                       int32_t n_keys, double* sigma, uint8_t* tissue, void* stream);
/* sigma_out = sigma_in, with the band cells (active, not tissue) taking the
 * sigma of their nearest tissue cell from ft_tissue: extend_conductivity
 * :189-206. */
int hp_extend_sigma_f64(const int32_t* ft_tissue, const uint8_t* tissue, const uint8_t* active,
                        const double* sigma_in, int32_t n0, int32_t n1, int32_t n2,
                        double* sigma_out, void* stream);
/* Per column (i, j) of a flag volume: first set k (k0), run length to the last
 * set k (run) and number of set flags (cnt); each output optional. */
int hp_column_stats(const uint8_t* flag, int32_t ni, int32_t nj, int32_t nk, int32_t* k0,
                    int32_t* run, int32_t* cnt, void* stream);
/* out[0..n]: exclusive prefix sums of in, out[n] = total (one workgroup). */
int hp_scan_i32(const int32_t* in, int64_t n, int64_t* out, void* stream);
/* Walk each column's set flags in k order from row base[col]: list[row] =
 * flat index (optional), id_map[flat] = row or -1 (optional, every cell),
 * cell_of_row[row] = col_off[col] + k - k0[col] (optional). */
int hp_column_emit(const uint8_t* flag, int32_t ni, int32_t nj, int32_t nk, const int64_t* base,
                   int32_t* list, int32_t* id_map, const int32_t* k0, const int64_t* col_off,
                   int64_t* cell_of_row, void* stream);
/* surface = tissue cells with an in-box face neighbour outside the tissue:
 * _find_surface_ids :469-488 (6-connected dilation of the air). */
int hp_surface_flags(const uint8_t* tissue, int32_t ni, int32_t nj, int32_t nk, uint8_t* surf,
                     void* stream);
/* out[0] = flat index of the set flag of the given rank in flat order (scan:
 * the exclusive sums of the per-column counts, n_cols + 1 entries): the
 * pinned cell interior[len // 2] of :616-623. */
int hp_select_rank(const uint8_t* flag, const int64_t* scan, int64_t n_cols, int32_t nk,
                   int64_t rank, int64_t* out, void* stream);
/* Column-run stiffness of _assemble_stiffness (:276-397): per stored cell of
 * the runs (col_k0, col_len, col_off), the +i/+j/+k face conductances
 * harm(sigma)/h^2 (times harm(psi) when psi is given; 0 unless both cells
 * are active) and the diagonal summed over the six faces in the reference's
 * order (1 for inactive cells and rows without conductance), FP32 into gx,
 * gy, gz, diag (diag 1 on pin_flat: the pinned row) and, when g64 is given,
 * FP64 into g64[4][n_cells] (the unpinned matrix).
 * bad (device uint64) counts nonzero conductances to inactive cells. */
int hp_torso_faces_f64(const double* sigma, const double* psi, const uint8_t* active, int32_t ni,
                       int32_t nj, int32_t nk, double h_m, const int32_t* col_k0,
                       const int32_t* col_len, const int64_t* col_off, int64_t pin_flat,
                       float* gx, float* gy, float* gz, float* diag, double* g64,
                       int64_t n_cells, unsigned long long* bad, void* stream);

/* Projection of heart nodes into the coarse torso rows
 * (compute_bspm.py:126-157, build_projection_map): xyz (n x 3, mm), Rt =
 * {R row-major (9), t (3)}, proj[i] = id_map[flat(floor((R (x scale) + t -
 * origin) / dx_mm))] or -1 outside the grid (id_map: -1 for cells outside
 * the system).  R is applied as three products summed in order; exact for
 * the signed permutations the registration returns. */
int hp_project_rows_f64(const double* xyz, int64_t n, const double* Rt, double scale,
                        const double* origin, double dx_mm, const int32_t* id_map, int32_t ni,
                        int32_t nj, int32_t nk, int64_t* proj, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* HRDAYA_HIP_H */
