// Learned ionic surrogate: BF16 MFMA kernels (gfx950 v_mfma_f32_16x16x32_bf16,
// FP32 accumulation) for training and for the fused in-tissue inference.
//
// Network: 22 inputs (V, the 18 TTP06 state variables, phenotype one-hot) ->
// 128 -> 128 -> 128 (SiLU) -> 19 outputs (increments of V and of the 18
// state variables over one macro-step of 10 reaction substeps).  Inputs and
// outputs are normalised per feature; gates/concentrations with a wide
// dynamic range travel as logarithms (feature flags in hp_mlp_norm).
//
// MFMA operand layout (16x16x32): lane l holds 8 bf16 of A row (l%16) at
// k = 8*(l/16)..+8, 8 bf16 of B column (l%16) at the same k, and 4 floats of
// C column (l%16) at rows 4*(l/16)..+4.
//
// Inference computes H^T = W^T X^T with neurons as rows and voxels as
// columns, so the C registers of one layer become the B operand of the next
// without leaving registers: the 8 bf16 of a B fragment at K-step j are the 4
// rows of C tile 2j followed by the 4 rows of C tile 2j+1.  The weight
// matrices of layers 2-4 are therefore stored in LDS with their input index
// permuted the same way (perm_k below); the host keeps them in natural order.
#include "common.h"
#include "hrdaya_hip.h"
#include "ttp06_cell.h"

namespace hp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ floatx4 mfma16(const bf16x8& a, const bf16x8& b, const floatx4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ __bf16 f2bf(float x) { return (__bf16)x; }

// ---------------------------------------------------------------------------
// Generic GEMM for training: C[m,n] = alpha * sum_k A(m,k) B(k,n) + beta * C[m,n]
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] (bf16), C fp32 row-major (ldc).
// Optional bias[n] added (bias_mode 1) -- used for the forward layers.
// Block tile 64x64, 4 waves of 32x32, K step 32, operands staged in LDS.
// ---------------------------------------------------------------------------
constexpr int GT = 64, GK = 32;

__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const __bf16* __restrict__ A,
                                                        int64_t sam, int64_t sak,
                                                        const __bf16* __restrict__ B, int64_t sbk,
                                                        int64_t sbn, float* __restrict__ C,
                                                        int64_t ldc, float alpha, float beta,
                                                        const float* __restrict__ bias) {
  __shared__ __attribute__((aligned(16))) __bf16 As[GT][GK + 8];
  __shared__ __attribute__((aligned(16))) __bf16 Bt[GT][GK + 8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  // split-K: slice z of gridDim.z handles K range [kb, ke) and adds atomically
  const int ksl = (K + (int)gridDim.z - 1) / (int)gridDim.z;
  const int kb = (int)blockIdx.z * ((ksl + GK - 1) / GK * GK);
  const int ke = min(K, kb + (ksl + GK - 1) / GK * GK);
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  const __bf16 zero = f2bf(0.f);
  for (int k0 = kb; k0 < ke; k0 += GK) {
    for (int e = tid; e < GT * GK; e += 256) {
      const int r = e / GK, kk = e % GK;
      const int m = m0 + r, k = k0 + kk;
      As[r][kk] = (m < M && k < ke) ? A[(int64_t)m * sam + (int64_t)k * sak] : zero;
      const int n = n0 + r;
      Bt[r][kk] = (n < N && k < ke) ? B[(int64_t)k * sbk + (int64_t)n * sbn] : zero;
    }
    __syncthreads();
    const int row = lane & 15, kg = (lane >> 4) * 8;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const bf16x8 a = *(const bf16x8*)&As[wm + i * 16 + row][kg];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const bf16x8 b = *(const bf16x8*)&Bt[wn + j * 16 + row][kg];
        acc[i][j] = mfma16(a, b, acc[i][j]);
      }
    }
    __syncthreads();
  }
  const int col = lane & 15, rg = (lane >> 4) * 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + i * 16 + rg + q, n = n0 + wn + j * 16 + col;
        if (m < M && n < N) {
          float v = alpha * acc[i][j][q];
          if (gridDim.z > 1) {
            atomicAdd(&C[(int64_t)m * ldc + n], v);  // caller zeroes C; no beta/bias
          } else {
            if (beta != 0.f) v += beta * C[(int64_t)m * ldc + n];
            if (bias) v += bias[n];
            C[(int64_t)m * ldc + n] = v;
          }
        }
      }
}

// y = silu(x) as bf16 (and keep x), or the backward dx = dy * silu'(x)
__global__ void silu_kernel(const float* __restrict__ x, __bf16* __restrict__ y, int64_t n,
                            int ld_in, int ld_out, int cols) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t r = i / cols, c = i % cols;
  const float v = x[r * ld_in + c];
  y[r * ld_out + c] = f2bf(v / (1.f + __expf(-v)));
}

__global__ void silu_bwd_kernel(const float* __restrict__ x, const float* __restrict__ dy,
                                __bf16* __restrict__ dx, float* __restrict__ dxf, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = x[i];
  const float s = 1.f / (1.f + __expf(-v));
  const float g = dy[i] * s * (1.f + v * (1.f - s));
  dx[i] = f2bf(g);
  if (dxf) dxf[i] = g;
}

__global__ void cast_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// masked MSE: loss partial and dY = 2 (Y - T) / (rows * used) in bf16, over `used` of `cols`
__global__ void mse_grad_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                __bf16* __restrict__ dY, float* __restrict__ dYf, int rows,
                                int cols, int used, float* __restrict__ loss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < (int64_t)rows * cols) {
    const int c = (int)(i % cols);
    float g = 0.f;
    if (c < used) {
      const float d = Y[i] - T[(i / cols) * used + c];
      l = d * d;
      g = 2.f * d / ((float)rows * (float)used);
    }
    dY[i] = f2bf(g);
    if (dYf) dYf[i] = g;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l / ((float)rows * (float)used));
}

// column sums (bias gradient): out[c] = sum_r X[r*ld + c]
__global__ void colsum_kernel(const float* __restrict__ X, int rows, int cols, int ld,
                              float* __restrict__ out) {
  const int c = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x; r < rows; r += blockDim.x) s += X[(int64_t)r * ld + c];
  __shared__ float sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sh[0];
}

__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * w[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// ---------------------------------------------------------------------------
// Fused in-tissue inference.
// ---------------------------------------------------------------------------
constexpr int NIN = 22, NHID = 128, NOUT = 19;

// position p (0..31) of a 32-wide K block -> neuron index inside the block
__device__ __forceinline__ int perm_k(int p) {
  const int g = p >> 3, q = p & 7;
  return q < 4 ? 4 * g + q : 16 + 4 * g + (q - 4);
}
// its inverse: neuron n (0..31) -> position
__device__ __forceinline__ int inv_perm_k(int n) {
  return n < 16 ? 8 * (n >> 2) + (n & 3) : 8 * ((n - 16) >> 2) + 4 + ((n - 16) & 3);
}

// Per-feature (input f) and per-output (o) constants of the inference kernel.
// kind: input 0 = V, 1 = state row, 2 = state row as a logarithm, 3..5 =
// phenotype one-hot 0..2, 6 = padding; output 0 = V (the rate), 1 = state
// row, 2 = state row as a log-ratio, 6 = padding.  row: byte offset of the
// state row in the SoA state array, or kNoRow (past the buffer: loads read 0,
// stores are dropped).
struct MlpFeat {
  float rstd, nmb, base;  // x -> (x + base) * rstd + nmb  (nmb = -mean * rstd)
  int32_t kind;
  uint32_t row;
};
struct MlpOut {
  float std, mean, base;  // d = y * std + mean; stored value relative to base
  int32_t kind;
  uint32_t row;
};
constexpr uint32_t kNoRow = 0xF0000000u;

struct MlpSmem {
  __bf16 w1[NHID][32];          // [out][in], natural input order (padded 22 -> 32)
  __bf16 w2[NHID][NHID];        // [out][perm(in)]
  __bf16 w3[NHID][NHID];
  __bf16 w4[32][NHID];          // 19 -> 32 rows
  float b1[NHID], b2[NHID], b3[NHID], b4[32];
  MlpFeat fin[32];
  MlpOut fout[32];
};

// The bias of rows 4*grp..+4 of output tile m: the accumulator's initial value
__device__ __forceinline__ floatx4 bias4(const float* bias, int m, int grp) {
  const float* b = bias + m * 16 + grp * 4;
  return floatx4{b[0], b[1], b[2], b[3]};
}

// Scaled SiLU of a C tile (rows 4*grp..+4 of output tile m; bias already in
// the accumulator) into the B fragment positions (m&1)*4..+4 of K-step m/2.
// The weights are pre-scaled (see hp_mlp_weights): the accumulator holds
// u = -log2(e)*z and the stored activation is g = u/(1 + 2^u) = -log2(e)*SiLU(z),
// with the hardware exp2 and reciprocal (~1 ulp each, far below the bf16
// rounding of g); 2^u -> inf gives g = 0 for z << 0.
__device__ __forceinline__ void silu_pack(const floatx4& c, int m, bf16x8& dst) {
  // the four activations' steps interleaved (exp2, add, rcp, multiply), so
  // that no instruction waits on the one before it
  float e[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) e[q] = __builtin_amdgcn_exp2f(c[q]);
#pragma unroll
  for (int q = 0; q < 4; ++q) e[q] = __builtin_amdgcn_rcpf(1.f + e[q]);
#pragma unroll
  for (int q = 0; q < 4; ++q) dst[(m & 1) * 4 + q] = f2bf(c[q] * e[q]);
}


// Append the slots of the lanes with fb set to the fallback list: one atomic
// per wave.  Called by every lane of the wave.
__device__ __forceinline__ void append_fallback(bool fb, int64_t slot, int32_t* __restrict__ list,
                                                int32_t* __restrict__ count) {
  const uint64_t m = __ballot(fb);
  if (m == 0) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, leader);
  if (fb) list[base + __popcll(m & ((1ull << lane) - 1ull))] = (int32_t)slot;
}

// NC MFMA columns of 16 voxels per wave, THREADS per workgroup.  NC = 1 with
// 768 threads fits 170 registers (three waves per SIMD, the weights in LDS
// shared by 12 waves); NC = 2 with 512 threads reuses each A fragment for two
// columns at two waves per SIMD, with 1024 threads (within 128 registers) at
// four waves per SIMD sharing one copy of the weights.
template <int NC, int THREADS>
__global__ void __launch_bounds__(THREADS) surrogate_kernel(hp_bricks g, const float* __restrict__ V,
                                                        float* __restrict__ V_prev,
                                                        float* __restrict__ state, int64_t ld,
                                                        float* __restrict__ rate,
                                                        hp_mlp_weights wts, hp_mlp_norm nrm,
                                                        hp_surrogate_params sp) {
  constexpr int WAVES = THREADS / 64, TV = 16 * NC;  // voxels per wave tile
  extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
  MlpSmem& S = *reinterpret_cast<MlpSmem*>(smem_raw);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const __bf16* __restrict__ W1 = (const __bf16*)wts.w1;
  const __bf16* __restrict__ W2 = (const __bf16*)wts.w2;
  const __bf16* __restrict__ W3 = (const __bf16*)wts.w3;
  const __bf16* __restrict__ W4 = (const __bf16*)wts.w4;
  // stage weights (natural order in global) and the normalisation into LDS
  for (int e = tid; e < NHID * 32; e += THREADS) {
    const int o = e / 32, i = e % 32;
    S.w1[o][i] = i < NIN ? W1[o * NIN + i] : f2bf(0.f);
  }
  // layers 2-4: 16-byte loads of 8 natural inputs n0..n0+7 (n0 a multiple of
  // 8); under perm_k they land as two runs of 4 consecutive LDS positions
  // (inputs n0..n0+3 at inv(n0), n0+4..n0+7 at inv(n0+4))
  for (int e = tid; e < NHID * NHID / 8; e += THREADS) {
    const int o = e / (NHID / 8), n0 = (e % (NHID / 8)) * 8;
    const int pa = (n0 & ~31) + inv_perm_k(n0 & 31), pb = (n0 & ~31) + inv_perm_k((n0 & 31) + 4);
    const uint4 a = *reinterpret_cast<const uint4*>(W2 + o * NHID + n0);
    const uint4 b = *reinterpret_cast<const uint4*>(W3 + o * NHID + n0);
    *reinterpret_cast<uint2*>(&S.w2[o][pa]) = uint2{a.x, a.y};
    *reinterpret_cast<uint2*>(&S.w2[o][pb]) = uint2{a.z, a.w};
    *reinterpret_cast<uint2*>(&S.w3[o][pa]) = uint2{b.x, b.y};
    *reinterpret_cast<uint2*>(&S.w3[o][pb]) = uint2{b.z, b.w};
  }
  for (int e = tid; e < 32 * NHID / 8; e += THREADS) {
    const int o = e / (NHID / 8), n0 = (e % (NHID / 8)) * 8;
    const int pa = (n0 & ~31) + inv_perm_k(n0 & 31), pb = (n0 & ~31) + inv_perm_k((n0 & 31) + 4);
    const uint4 a = o < NOUT ? *reinterpret_cast<const uint4*>(W4 + o * NHID + n0) : uint4{0, 0, 0, 0};
    *reinterpret_cast<uint2*>(&S.w4[o][pa]) = uint2{a.x, a.y};
    *reinterpret_cast<uint2*>(&S.w4[o][pb]) = uint2{a.z, a.w};
  }
  for (int e = tid; e < NHID; e += THREADS) {
    S.b1[e] = wts.b1[e];
    S.b2[e] = wts.b2[e];
    S.b3[e] = wts.b3[e];
  }
  if (tid < 32) {
    const int f = tid;
    S.b4[f] = f < NOUT ? wts.b4[f] : 0.f;
    // the FP32 engine stores Ki and Nai relative to their bases
    const float base = f == 1 ? sp.ki_base : (f == 2 ? sp.nai_base : 0.f);
    const bool row = f >= 1 && f < 19;
    MlpFeat fi;
    fi.rstd = f < NIN ? nrm.in_rstd[f] : 0.f;
    fi.nmb = f < NIN ? -nrm.in_mean[f] * fi.rstd : 0.f;
    fi.base = row ? base : 0.f;
    fi.kind = f == 0 ? 0 : row ? (nrm.in_log[f] ? 2 : 1) : (f < NIN ? 3 + (f - 19) : 6);
    fi.row = row ? (uint32_t)(f - 1) * (uint32_t)ld * 4u : kNoRow;
    S.fin[f] = fi;
    MlpOut fo;
    fo.std = f < NOUT ? nrm.out_std[f] : 0.f;
    fo.mean = f < NOUT ? nrm.out_mean[f] : 0.f;
    fo.base = row ? base : 0.f;
    fo.kind = f == 0 ? 0 : row ? (nrm.in_log[f] ? 2 : 1) : 6;
    fo.row = row ? (uint32_t)(f - 1) * (uint32_t)ld * 4u : kNoRow;
    S.fout[f] = fo;
  }
  __syncthreads();
  // the whole SoA state array as one buffer (byte offsets; checked on the host)
  const buf_t st_r = buf_rsrc(state, (uint32_t)(ld * N_STATE * 4));

  const int64_t n_items = sp.slot_list ? sp.n_list : g.n_bricks * (int64_t)(g.bx * g.by * g.bz);
  const int col = lane & 15, grp = lane >> 4;
  const float dtm = sp.t_macro, inv_dtm = 1.f / sp.t_macro;
  // persistent loop: one TV-voxel tile (NC 16-voxel MFMA columns) per wave,
  // over the slots or over the listed slots
  for (int64_t tile = (int64_t)blockIdx.x * WAVES + wave; tile * TV < n_items;
       tile += (int64_t)gridDim.x * WAVES) {
    const int64_t base = tile * TV;
    int64_t slot[NC], si[NC];
    uint8_t lab[NC];
    float r0[NC];
    bool use[NC];  // the network's output is applied to this voxel
    bool any = false;
#pragma unroll
    for (int t = 0; t < NC; ++t) {
      const int64_t item = base + 16 * t + col;
      slot[t] = sp.slot_list ? (int64_t)sp.slot_list[item] : item;
      lab[t] = slot[t] >= 0 ? g.label[slot[t]] : (uint8_t)0xFF;
      if (slot[t] < 0) slot[t] = 0;  // no voxel: a valid address, never written (label 0xFF)
      // compact state entry (read and written only where lab != 0xFF)
      si[t] = lab[t] == 0xFF ? 0
              : (sp.slot_list && sp.state_list)
                  ? (int64_t)sp.state_list[item]
                  : (int64_t)state_index(g.groups, (uint32_t)slot[t]);
      // correction mode: the big step has decided (rate NaN = fallback); at
      // rest its own step stands
      r0[t] = sp.correct ? rate[slot[t]] : 0.f;
      use[t] = lab[t] != 0xFF;
      if (sp.correct) {
        const float v = V[slot[t]];
        use[t] = use[t] && r0[t] == r0[t] && !(v < sp.rest_v && fabsf(r0[t]) < sp.rest_rate);
      }
      any |= use[t];
    }
    // skip tiles without a voxel to update (wave-uniform)
    if (!__any(any)) continue;
    // ---- inputs: B fragments (32 features x 16 voxels) of the columns.
    // Every lane loads its 8 features' state rows at once (byte offsets into
    // the state buffer; rows that are not state variables read 0), then
    // forms the features without branches from the per-feature table.
    float vv[NC], raw[NC][8];
    uint32_t sbyte[NC];
#pragma unroll
    for (int t = 0; t < NC; ++t) {
      vv[t] = V[slot[t]];
      sbyte[t] = (uint32_t)si[t] * 4u;
#pragma unroll
      for (int q = 0; q < 8; ++q)
        raw[t][q] = __builtin_bit_cast(
            float, __builtin_amdgcn_raw_buffer_load_b32(st_r, S.fin[grp * 8 + q].row + sbyte[t], 0,
                                                        0));
    }
    bf16x8 xin[NC];
#pragma unroll
    for (int t = 0; t < NC; ++t) {
      const int ct = lab[t] & 3;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const MlpFeat c = S.fin[grp * 8 + q];
        float x = raw[t][q] + c.base;
        if (q == 0) x = c.kind == 0 ? vv[t] : x;  // V is feature 0 (group 0, q 0)
        if (q >= 3 && q <= 5)                      // the one-hot is features 19-21
          x = c.kind >= 3 ? (c.kind - 3 == ct ? 1.f : 0.f) : x;
        // hardware log2 (the argument is clamped above the denormal range)
        const float lx = __builtin_amdgcn_logf(fmaxf(x, 1e-30f)) * 0.69314718f;
        x = c.kind == 2 ? lx : x;
        x = fmaf(x, c.rstd, c.nmb);
        xin[t][q] = f2bf(lab[t] != 0xFF ? x : 0.f);
      }
    }
    // ---- layer 1 (one K step) -> hidden activations in B-fragment form
    bf16x8 ha[4][NC], hb[4][NC];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const bf16x8 a = *(const bf16x8*)&S.w1[m * 16 + col][grp * 8];
      const floatx4 z = bias4(S.b1, m, grp);
#pragma unroll
      for (int t = 0; t < NC; ++t) silu_pack(mfma16(a, xin[t], z), m, ha[m >> 1][t]);
    }
    // ---- layers 2 and 3
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
      const __bf16(*W)[NHID] = layer == 0 ? S.w2 : S.w3;
      const float* bias = layer == 0 ? S.b2 : S.b3;
      bf16x8(&hin)[4][NC] = layer == 0 ? ha : hb;
      bf16x8(&hout)[4][NC] = layer == 0 ? hb : ha;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        floatx4 c[NC];
#pragma unroll
        for (int t = 0; t < NC; ++t) c[t] = bias4(bias, m, grp);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const bf16x8 a = *(const bf16x8*)&W[m * 16 + col][j * 32 + grp * 8];
#pragma unroll
          for (int t = 0; t < NC; ++t) c[t] = mfma16(a, hin[j][t], c[t]);
        }
#pragma unroll
        for (int t = 0; t < NC; ++t) silu_pack(c[t], m, hout[m >> 1][t]);
      }
    }
    // ---- output layer (19 of 32 rows used); input is `ha` after two swaps
    floatx4 out[2][NC];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
      floatx4 c[NC];
#pragma unroll
      for (int t = 0; t < NC; ++t) c[t] = bias4(S.b4, m, grp);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const bf16x8 a = *(const bf16x8*)&S.w4[m * 16 + col][j * 32 + grp * 8];
#pragma unroll
        for (int t = 0; t < NC; ++t) c[t] = mfma16(a, ha[j][t], c[t]);
      }
#pragma unroll
      for (int t = 0; t < NC; ++t) out[m][t] = c[t];
    }
    // ---- mode decision and update, per voxel (this lane's 4+4 outputs).
    // The state values this lane updates are loaded first (rows that are not
    // state variables read 0); stores of rows that are not updated go past
    // the buffer and are dropped.
    float sx[NC][2][4];
#pragma unroll
    for (int t = 0; t < NC; ++t)
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          sx[t][m][q] = __builtin_bit_cast(
              float, __builtin_amdgcn_raw_buffer_load_b32(
                         st_r, S.fout[m * 16 + grp * 4 + q].row + sbyte[t], 0, 0));
#pragma unroll
    for (int t = 0; t < NC; ++t) {
      const int64_t sl = slot[t];
      // correction mode: the big step has set V_prev and listed the fallback
      // voxels; its rate is the base of ours (use[t] decided above)
      // the V correction (output 0, held by the lane of group 0 of this
      // column) to every lane of the column, for the trust bound
      const float dv0 = __shfl(out[0][t][0] * S.fout[0].std + S.fout[0].mean, col);
      bool fallback;
      if (sp.correct) {
        fallback = !(use[t] && fabsf(dv0) <= sp.corr_max);
      } else {
        const float v = vv[t], vp = V_prev[sl];
        fallback = sp.force_fallback || fabsf((v - vp) / dtm) > sp.dvdt_max ||
                   (v > sp.v_lo && v < sp.v_hi) || (vp != vp);
      }
      if (sp.fb_list && !sp.correct)
        append_fallback(lab[t] != 0xFF && grp == 0 && fallback, sl, sp.fb_list, sp.fb_count);
      const bool upd = lab[t] != 0xFF && !fallback;
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const MlpOut c = S.fout[m * 16 + grp * 4 + q];
          const float d = fmaf(out[m][t][q], c.std, c.mean);
          const float x0 = sx[t][m][q] + c.base;
          const float x1 =
              c.kind == 2 ? fmaxf(x0, 1e-30f) * __builtin_amdgcn_exp2f(d * 1.44269504f) : x0 + d;
          __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x1 - c.base), st_r,
                                                (upd ? c.row : kNoRow) + sbyte[t], 0, 0);
        }
      // the rate: output 0 (m = 0, q = 0) lives in the lanes of group 0
      if (grp == 0 && lab[t] != 0xFF) {
        const float d = fmaf(out[0][t][0], S.fout[0].std, S.fout[0].mean);
        if (!sp.correct) rate[sl] = fallback ? __builtin_nanf("") : d * inv_dtm;
        else if (!fallback) rate[sl] = r0[t] + d * inv_dtm;
      }
    }
    // V_prev of this tile was consumed above (the mode decision depends on it),
    // so the loads have completed before these stores are issued
    if (grp == 0 && !sp.correct) {
#pragma unroll
      for (int t = 0; t < NC; ++t)
        if (lab[t] != 0xFF) V_prev[slot[t]] = V[slot[t]];
    }
  }
}

// Multirate alternative without the network, same decision rule: one
// Rush-Larsen/forward-Euler step of t_macro for the non-fallback voxels.
__global__ void __launch_bounds__(256) bigstep_kernel(hp_bricks g, const float* __restrict__ V,
                                                      float* __restrict__ V_prev,
                                                      float* __restrict__ state, int64_t ld,
                                                      float* __restrict__ rate,
                                                      hp_surrogate_params sp) {
  const int64_t slot = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n_slots = g.n_bricks * (int64_t)(g.bx * g.by * g.bz);
  const uint8_t lab = slot < n_slots ? g.label[slot] : (uint8_t)0xFF;
  const float v = lab != 0xFF ? V[slot] : 0.f, vp = lab != 0xFF ? V_prev[slot] : 0.f;
  const float dtm = sp.t_macro;
  const bool fallback = sp.force_fallback || fabsf((v - vp) / dtm) > sp.dvdt_max ||
                        (v > sp.v_lo && v < sp.v_hi) || (vp != vp);
  if (sp.fb_list) append_fallback(lab != 0xFF && fallback, slot, sp.fb_list, sp.fb_count);
  if (lab == 0xFF) return;
  V_prev[slot] = v;
  if (fallback) {
    rate[slot] = __builtin_nanf("");
    return;
  }
  const int ct = lab & 3;
  const uint32_t row_bytes = (uint32_t)ld * 4u, sl = state_index(g.groups, (uint32_t)slot);
  float s[N_STATE];
#pragma unroll
  for (int k = 0; k < N_STATE; ++k) s[k] = buf_ld<float>(buf_rsrc(state + k * (size_t)ld, row_bytes), sl);
  const float I = ttp06_cell<float, false>(v, s, dtm, (float)kGKs[ct], (float)kGto[ct], ct != 2,
                                           sp.ki_base, sp.nai_base);
#pragma unroll
  for (int k = 0; k < N_STATE; ++k) buf_st(buf_rsrc(state + k * (size_t)ld, row_bytes), sl, s[k]);
  rate[slot] = -I;
}

}  // namespace hp

using namespace hp;

extern "C" {

int hp_gemm_bf16(int M, int N, int K, const void* A, int64_t sam, int64_t sak, const void* B,
                 int64_t sbk, int64_t sbn, float* C, int64_t ldc, float alpha, float beta,
                 const float* bias, int ksplit, void* stream) {
  if (M <= 0 || N <= 0 || K <= 0 || ksplit < 1 || ksplit > 1024 || !A || !B || !C) {
    set_error("hp_gemm_bf16: bad shape %d x %d x %d (ksplit %d) or null operand", M, N, K, ksplit);
    return -1;
  }
  if (ksplit > 1 && (beta != 0.f || bias)) {
    set_error("hp_gemm_bf16: split-K accumulates atomically; beta and bias must be 0");
    return -1;
  }
  dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, ksplit);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, (hipStream_t)stream, M, N, K,
                     (const __bf16*)A, sam, sak, (const __bf16*)B, sbk, sbn, C, ldc, alpha, beta,
                     bias);
  return check_launch("hp_gemm_bf16");
}

int hp_silu_bf16(const float* x, void* y, int64_t rows, int cols, int ld_in, int ld_out,
                 void* stream) {
  const int64_t n = rows * cols;
  if (n <= 0) return 0;
  if (!x || !y) {
    set_error("hp_silu_bf16: bad arguments (null operand)");
    return -1;
  }
  hipLaunchKernelGGL(silu_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, x,
                     (__bf16*)y, n, ld_in, ld_out, cols);
  return check_launch("hp_silu_bf16");
}

int hp_silu_bwd_bf16(const float* x, const float* dy, void* dx, float* dxf, int64_t n,
                     void* stream) {
  if (n <= 0) return 0;
  if (!x || !dy || !dx) {
    set_error("hp_silu_bwd_bf16: bad arguments (null operand)");
    return -1;
  }
  hipLaunchKernelGGL(silu_bwd_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     x, dy, (__bf16*)dx, dxf, n);
  return check_launch("hp_silu_bwd_bf16");
}

int hp_cast_bf16(const float* x, void* y, int64_t n, void* stream) {
  if (n <= 0) return 0;
  if (!x || !y) {
    set_error("hp_cast_bf16: bad arguments (null operand)");
    return -1;
  }
  hipLaunchKernelGGL(cast_bf16_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     x, (__bf16*)y, n);
  return check_launch("hp_cast_bf16");
}

int hp_mse_grad(const float* Y, const float* T, void* dY, float* dYf, int rows, int cols, int used,
                float* loss, void* stream) {
  const int64_t n = (int64_t)rows * cols;
  if (n <= 0) return 0;
  if (!Y || !T || !dY || !loss || used < 1 || used > cols) {
    set_error("hp_mse_grad: bad arguments (null operand, or used outside 1..cols)");
    return -1;
  }
  hipLaunchKernelGGL(mse_grad_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, Y,
                     T, (__bf16*)dY, dYf, rows, cols, used, loss);
  return check_launch("hp_mse_grad");
}

int hp_colsum(const float* X, int rows, int cols, int ld, float* out, void* stream) {
  if (rows <= 0 || cols <= 0) return 0;
  if (!X || !out || ld < cols) {
    set_error("hp_colsum: bad arguments (null operand, or ld < cols)");
    return -1;
  }
  hipLaunchKernelGGL(colsum_kernel, dim3(cols), dim3(256), 0, (hipStream_t)stream, X, rows, cols,
                     ld, out);
  return check_launch("hp_colsum");
}

int hp_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
            float eps, int step, float wd, void* stream) {
  if (n <= 0 || step <= 0) return 0;
  if (!w || !g || !m || !v) {
    set_error("hp_adam: bad arguments (null operand)");
    return -1;
  }
  const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, (hipStream_t)stream, w, g, m,
                     v, n, lr, b1, b2, eps, bc1, bc2, wd);
  return check_launch("hp_adam");
}

int hp_surrogate_step(const hp_bricks* g, const float* V, float* V_prev, float* state, int64_t ld,
                      float* rate, const hp_mlp_weights* w, const hp_mlp_norm* n,
                      const hp_surrogate_params* sp, int n_blocks, void* stream) {
  if (!g || !g->groups || !w || !n || !sp || g->n_bricks <= 0 || ld < 0) {
    set_error("hp_surrogate_step: bad arguments (groups required)");
    return -1;
  }
  const int64_t n_slots = g->n_bricks * (int64_t)(g->bx * g->by * g->bz);
  const int64_t n_items = sp->slot_list ? sp->n_list : n_slots;
  if ((uint64_t)ld * N_STATE * 4 >= (uint64_t)kNoRow) {
    set_error("hp_surrogate_step: the state array (18 x ld floats) must be below 3.75 GiB");
    return -1;
  }
  if (sp->variant < 0 || sp->variant > 2) {
    set_error("hp_surrogate_step: variant must be 0, 1 or 2");
    return -1;
  }
  if (n_items % 32 || n_items <= 0) {
    set_error("hp_surrogate_step: slot (or slot_list) count must be a positive multiple of 32");
    return -1;
  }
  const size_t smem = sizeof(MlpSmem);
  static bool attr_set = false;
  if (!attr_set) {
    HP_CHECK(hipFuncSetAttribute((const void*)surrogate_kernel<1, 768>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    HP_CHECK(hipFuncSetAttribute((const void*)surrogate_kernel<2, 512>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    HP_CHECK(hipFuncSetAttribute((const void*)surrogate_kernel<2, 1024>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    attr_set = true;
  }
  if (n_blocks <= 0) n_blocks = 256;
  const int nc = sp->variant >= 1 ? 2 : 1, waves = sp->variant == 2 ? 16 : nc == 2 ? 8 : 12;
  const int64_t tiles = n_items / (16 * nc);
  if ((int64_t)n_blocks * waves > tiles) n_blocks = (int)((tiles + waves - 1) / waves);
  if (nc == 1)
    hipLaunchKernelGGL((surrogate_kernel<1, 768>), dim3(n_blocks), dim3(768), smem,
                       (hipStream_t)stream, *g, V, V_prev, state, ld, rate, *w, *n, *sp);
  else if (sp->variant == 2)
    hipLaunchKernelGGL((surrogate_kernel<2, 1024>), dim3(n_blocks), dim3(1024), smem,
                       (hipStream_t)stream, *g, V, V_prev, state, ld, rate, *w, *n, *sp);
  else
    hipLaunchKernelGGL((surrogate_kernel<2, 512>), dim3(n_blocks), dim3(512), smem,
                       (hipStream_t)stream, *g, V, V_prev, state, ld, rate, *w, *n, *sp);
  return check_launch("hp_surrogate_step");
}

int hp_bigstep_reaction(const hp_bricks* g, const float* V, float* V_prev, float* state,
                        int64_t ld, float* rate, const hp_surrogate_params* sp, void* stream) {
  const int64_t n_slots = g ? g->n_bricks * (int64_t)(g->bx * g->by * g->bz) : 0;
  if (!g || !g->groups || !sp || g->n_bricks <= 0 || ld < 0 || (uint64_t)ld * 4 >= (1ull << 32)) {
    set_error("hp_bigstep_reaction: bad arguments (groups required, ld*4 < 2^32)");
    return -1;
  }
  hipLaunchKernelGGL(bigstep_kernel, dim3(ceil_div(n_slots, 256)), dim3(256), 0,
                     (hipStream_t)stream, *g, V, V_prev, state, ld, rate, *sp);
  return check_launch("hp_bigstep_reaction");
}

}  // extern "C"
