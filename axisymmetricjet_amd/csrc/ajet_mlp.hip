// gfx950 kernels of the learned LAD coefficient: a 104-128-128-128-2 MLP with GELU, BF16
// weights and activations, FP32 accumulation on the matrix cores (v_mfma_f32_32x32x16_bf16 for
// training, v_mfma_f32_16x16x32_bf16 for inference).  The network stands in for the costly part
// of ComputeArtTransCoeffs (ModShockCapturing.f90:5-106): it maps a 5 x 5 stencil of the state
// to beta* and kappa*.
//
//   mlp_gemm_kernel<EPI>   C = A B^T (both operands K-contiguous), one 32 x 32 tile per wave,
//                          with the training epilogues (bias + GELU, GELU', split-K partials)
//   mlp_loss_kernel        dy = (y - t) / n and the loss
//   mlp_bias_grad_kernel   column sums of dz
//   mlp_adam_kernel        split-K reduction of dW and the Adam update (torch.optim.Adam's
//                          arithmetic), writing the BF16 weights in both layouts
//   mlp_infer_kernel       features, normalisation and the four layers for groups of 16
//                          points, weights resident in LDS, beta* and kappa* out
//
// Build with -ffp-contract=off: the normalisation and epilogues follow the oracle's order.
#include <hip/hip_runtime.h>

#include "ajet_kernels.h"

namespace ajet {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float K_GELU = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float A_GELU = 0.044715f;
constexpr float LOG2E = 1.4426950408889634f;

// GELU in its tanh form, evaluated as x s with s = 1 / (1 + exp(-2u)) = (1 + tanh u) / 2,
// u = sqrt(2/pi) (x + 0.044715 x^3), so -2u log2(e) = x (G1 + G2 x^2); the hardware exp2 and
// reciprocal.
constexpr float G1 = -2.0f * LOG2E * K_GELU;
constexpr float G2 = -2.0f * LOG2E * K_GELU * A_GELU;
__device__ __forceinline__ float gelu_sig(float x) {
  const float w = x * __builtin_fmaf(G2, x * x, G1);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(w));
}
__device__ __forceinline__ float gelu_f(float x) { return x * gelu_sig(x); }
// d/dx [x s(x)] = s + 2 x s (1 - s) sqrt(2/pi) (1 + 3 0.044715 x^2)
__device__ __forceinline__ float gelu_grad_f(float x) {
  const float s = gelu_sig(x);
  return s + 2.0f * x * s * (1.0f - s) * K_GELU * (1.0f + 3.0f * A_GELU * x * x);
}

// ---- training GEMM: C[M x N] = A[M x K] B[N x K]^T over K-slice s of `splits`.
template <int EPI>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(MlpGemm g) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mt = g.M / 32, nt = g.N / 32;
  const int tile = (int)blockIdx.x * 4 + wave;
  if (tile >= mt * nt * g.splits) return;
  const int s = tile / (mt * nt), rem = tile - s * (mt * nt);
  const int m0 = (rem / nt) * 32, n0 = (rem % nt) * 32;
  const long long kb = (long long)s * g.K;
  // lane l supplies row (l mod 32) of the A tile and column (l mod 32) of the B^T tile, eight
  // K values each; both operands take the same K split, so the products pair up
  const __bf16* pa = g.A + (long long)(m0 + (lane & 31)) * g.lda + kb + 8 * (lane >> 5);
  const __bf16* pb = g.B + (long long)(n0 + (lane & 31)) * g.ldb + kb + 8 * (lane >> 5);
  f32x16 acc = {};
  for (int k = 0; k < g.K; k += 16) {
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(pa + k);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(pb + k);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  }
  // accumulator v of lane l: row 8 (v / 4) + 4 (l / 32) + v mod 4, column l mod 32
  const int col = n0 + (lane & 31);
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const int row = m0 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
    const long long rc = (long long)row * g.N + col;
    if constexpr (EPI == MLP_EPI_FWD_GELU) {
      const float x = acc[v] + g.bias[col];
      g.Z[rc] = x;
      const __bf16 h = (__bf16)gelu_f(x);
      g.O[rc] = h;
      g.OT[(long long)col * g.M + row] = h;
    } else if constexpr (EPI == MLP_EPI_FWD_LAST) {
      g.Z[rc] = acc[v] + g.bias[col];
    } else if constexpr (EPI == MLP_EPI_BWD_GELU) {
      const __bf16 h = (__bf16)(acc[v] * gelu_grad_f(g.Zin[rc]));
      g.O[rc] = h;
      g.OT[(long long)col * g.M + row] = h;
    } else {
      g.Z[(long long)s * g.M * g.N + rc] = acc[v];
    }
  }
}

// dy = (y - t) / n for the two outputs (the other 30 columns of the padded output are zero);
// the loss sum((y - t)^2) / (2 n) accumulates into *loss.
__global__ __launch_bounds__(256) void mlp_loss_kernel(const float* __restrict__ y, const float* __restrict__ t, int n,
                                                       __bf16* __restrict__ dz, __bf16* __restrict__ dzT,
                                                       double* __restrict__ loss) {
  __shared__ double part[256];
  const int b = (int)blockIdx.x * 256 + (int)threadIdx.x;
  double l = 0.0;
  if (b < n) {
    for (int c = 0; c < MLP_NOUT_PAD; ++c) {
      float d = 0.0f;
      if (c < MLP_NOUT) {
        const float yv = y[(long long)b * MLP_NOUT_PAD + c], tv = t[2 * b + c];
        d = (yv - tv) / (float)n;
        const double e = (double)yv - (double)tv;
        l += e * e;
      }
      const __bf16 h = (__bf16)d;
      dz[(long long)b * MLP_NOUT_PAD + c] = h;
      dzT[(long long)c * n + b] = h;
    }
  }
  part[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, part[0] / (2.0 * n));
}

// db[c] = sum_b dzT[c][b] (the BF16-rounded dz), one block per column
__global__ __launch_bounds__(256) void mlp_bias_grad_kernel(const __bf16* __restrict__ dzT, int n, float* __restrict__ db) {
  __shared__ float part[256];
  const int c = (int)blockIdx.x;
  float s = 0.0f;
  for (int b = (int)threadIdx.x; b < n; b += 256) s += (float)dzT[(long long)c * n + b];
  part[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = part[0];
}

// One layer of Adam: parameters W (N x K, row major) then b (N), master copies in FP32.
// g = sum of the split-K partials (weights) or db (biases); torch.optim.Adam:
//   m = lerp(m, g, 1 - beta1); v = v beta2 + (1 - beta2) g g;
//   p += (-step m) / (sqrt(v) / sqrt(bc2) + eps),  step = lr / bc1.
// grad_only: only write the reduced gradient into m (test hook, no update).
__global__ __launch_bounds__(256) void mlp_adam_kernel(MlpAdam a) {
  const int idx = (int)blockIdx.x * 256 + (int)threadIdx.x;
  const int nw = a.N * a.K;
  if (idx >= nw + a.N) return;
  float g;
  if (idx < nw) {
    g = 0.0f;
    for (int s = 0; s < a.splits; ++s) g += a.partial[(long long)s * nw + idx];
  } else {
    g = a.db[idx - nw];
  }
  if (a.grad_only) {
    a.grad[idx] = g;
    return;
  }
  float m = a.m[idx], v = a.v[idx];
  m = m + a.one_m_beta1 * (g - m);
  v = v * a.beta2 + a.one_m_beta2 * g * g;
  const float denom = sqrtf(v) / a.bc2_sqrt + a.eps;
  const float p = a.p[idx] + ((-a.step) * m) / denom;
  a.m[idx] = m;
  a.v[idx] = v;
  a.p[idx] = p;
  const __bf16 h = (__bf16)p;
  if (idx < nw) {
    const int r = idx / a.K, k = idx - r * a.K;
    a.W[idx] = h;
    a.WT[(long long)k * a.N + r] = h;
  } else {
    a.b[idx - nw] = p;
  }
}

// FP32 rows x (n x 104, already normalised) -> BF16 rows (n x 128, zero padded) and the
// transpose (128 x n); with idx, row b takes x[idx[b]] (a minibatch) and t[idx[b]].
__global__ __launch_bounds__(256) void mlp_rows_kernel(const float* __restrict__ x, const float* __restrict__ t,
                                                       const long long* __restrict__ idx, int n,
                                                       __bf16* __restrict__ a0, __bf16* __restrict__ a0T,
                                                       float* __restrict__ tb) {
  const int b = (int)blockIdx.x;
  if (b >= n) return;
  const long long src = idx ? idx[b] : b;
  for (int k = (int)threadIdx.x; k < MLP_K; k += 256) {
    const __bf16 h = (__bf16)(k < MLP_NIN ? x[src * MLP_NIN + k] : 0.0f);
    a0[(long long)b * MLP_K + k] = h;
    a0T[(long long)k * n + b] = h;
  }
  if (threadIdx.x < 2 && tb) tb[2 * b + threadIdx.x] = t[2 * src + threadIdx.x];
}

// minibatch indices: splitmix64 of (seed, step, b) modulo the number of samples
__global__ void mlp_sample_kernel(unsigned long long seed, long long step, long long nsamples, int n,
                                  long long* __restrict__ idx) {
  const int b = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (b >= n) return;
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(step * 1048576 + b + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  idx[b] = (long long)(z % (unsigned long long)nsamples);
}

// ---- features: the 5 x 5 stencil of rho, Vr, Vz, T (indices clamped to the grid), centre as
// its value, neighbours as differences from it, then dz(i), dr(j), ramp(i), r(j).
__device__ __forceinline__ float mlp_raw_feature(const MlpInfer& a, int i, int j, int n) {
  const Geometry& g = a.g;
  if (n < 100) {
    const int vi = n / 25, rem = n - 25 * vi, b = rem / 5 - 2, c = rem % 5 - 2;
    const int v = vi == 0 ? 0 : (vi == 1 ? 1 : (vi == 2 ? 3 : 4));
    const double* f = a.q + (long long)v * g.fstride;
    const double ctr = f[(long long)i + (long long)g.pitch * j];
    const int ii = min(max(i + c, 0), g.nz - 1), jj = min(max(j + b, 0), g.nr - 1);
    const double nb = f[(long long)ii + (long long)g.pitch * jj];
    return (c == 0 && b == 0) ? (float)ctr : (float)(nb - ctr);
  }
  if (n == 100) return (float)a.dzl[i];
  if (n == 101) return (float)a.drl[j];
  if (n == 102) return (float)a.ramp[i];
  return (float)a.rc[j];
}

__global__ __launch_bounds__(256) void mlp_features_kernel(MlpInfer a, const int* __restrict__ pts, int n,
                                                           float* __restrict__ out) {
  const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
  if (k >= (long long)n * MLP_NIN) return;
  const int p = (int)(k / MLP_NIN), f = (int)(k - (long long)p * MLP_NIN);
  out[k] = mlp_raw_feature(a, pts[2 * p], pts[2 * p + 1], f);
}

// ---- inference.  A wave takes 16 points (i0 .. i0+15, j) at a time.  Lane (p, q) = (l mod 16,
// l / 16) loads the 5 x 5 stencil of variable q around point p and writes features
// 25 q .. 25 q + 24 of that point (and metric 100 + q) into the wave's activation rows.  Each
// layer is 8 tiles of 16 outputs x 16 points: A = 16 weight rows, B = the activation rows of the
// 16 points, K = 128 in four 32-deep MFMA steps; the accumulator of lane (p, q) then holds
// outputs 16 t + 4 q .. + 3 of point p, which the epilogue (bias, GELU, BF16) writes back over
// the layer's input as one 8-byte store.
constexpr int WROW = MLP_K + 8;  // padded LDS row (BF16 elements): 16-byte reads conflict-free
constexpr int INFER_WAVES = 12;  // 3 waves per SIMD: 160 KB of LDS holds the weights and 12 activation tiles
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_order() {
  // the activation rows pass between lanes of one wave only: LDS keeps the wave's order, the
  // barrier keeps the compiler's
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64 * INFER_WAVES) void mlp_infer_kernel(MlpInfer a) {
  __shared__ __bf16 sW[3 * MLP_H + MLP_NOUT][WROW];      // W1 (K padded), W2, W3, the two rows of W4
  __shared__ __bf16 sAct[INFER_WAVES][16][WROW];
  __shared__ __attribute__((aligned(16))) float sBias[3 * MLP_H + MLP_NOUT];
  __shared__ float sShift[MLP_NIN], sScale[MLP_NIN];
  const Geometry& g = a.g;
  const int nz = g.nz, nr = g.nr, pitch = g.pitch;
  const long long fs = g.fstride;
  for (int k = (int)threadIdx.x; k < (3 * MLP_H + MLP_NOUT) * (MLP_K / 8); k += blockDim.x) {
    const int r = k / (MLP_K / 8), c = (k % (MLP_K / 8)) * 8;
    *reinterpret_cast<bf16x8*>(&sW[r][c]) = *reinterpret_cast<const bf16x8*>(a.W + (long long)r * MLP_K + c);
  }
  for (int k = (int)threadIdx.x; k < 3 * MLP_H + MLP_NOUT; k += blockDim.x) sBias[k] = a.bias[k];
  for (int k = (int)threadIdx.x; k < MLP_NIN; k += blockDim.x) {
    sShift[k] = a.shift[k];
    sScale[k] = a.scale[k];
  }
  __syncthreads();
  const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int p = lane & 15, q = lane >> 4;
  __bf16(*act)[WROW] = sAct[wave];
  const double* __restrict__ fq = a.q + (long long)(q < 2 ? q : q + 1) * fs;  // rho, Vr, Vz, T
  const int ngz = (nz + 15) / 16;
  const long long ngroups = (long long)ngz * nr;
  for (long long grp = (long long)blockIdx.x * INFER_WAVES + wave; grp < ngroups;
       grp += (long long)gridDim.x * INFER_WAVES) {
    const int j = (int)(grp / ngz), i0 = (int)(grp % ngz) * 16;
    const int ilast = min(i0 + 15, nz - 1);
    if (a.ramp[ilast] < a.ramp_min) {
      // the ramp (increasing in z) is below the cut-off over the whole group: no LAD here
      for (int r = q; r < 16; r += 4) {
        const int i = i0 + r;
        if (p < 3 && i < nz) {
          a.out[(long long)p * fs + i + (long long)pitch * j] = 0.0;
          if (a.out2) a.out2[(long long)p * fs + i + (long long)pitch * j] = 0.0;
        }
      }
      continue;
    }
    {
      const int i = min(i0 + p, nz - 1);
      int col[5];
      long long row[5];
#pragma unroll
      for (int c = 0; c < 5; ++c) col[c] = min(max(i + c - 2, 0), nz - 1);
#pragma unroll
      for (int b = 0; b < 5; ++b) row[b] = (long long)pitch * min(max(j + b - 2, 0), nr - 1);
      double v[25];
#pragma unroll
      for (int b = 0; b < 5; ++b)
#pragma unroll
        for (int c = 0; c < 5; ++c) v[5 * b + c] = fq[col[c] + row[b]];
      const double ctr = v[12];
#pragma unroll
      for (int s = 0; s < 25; ++s) {
        const int n = 25 * q + s;
        const float x = s == 12 ? (float)ctr : (float)(v[s] - ctr);
        act[p][n] = (__bf16)((x - sShift[n]) * sScale[n]);
      }
      const float m = q == 0 ? (float)a.dzl[i] : (q == 1 ? (float)a.drl[j] : (q == 2 ? (float)a.ramp[i] : (float)a.rc[j]));
      act[p][100 + q] = (__bf16)((m - sShift[100 + q]) * sScale[100 + q]);
#pragma unroll
      for (int t = 0; t < 6; ++t) act[p][104 + q + 4 * t] = (__bf16)0.0f;
    }
    wave_order();
    for (int layer = 0; layer < 3; ++layer) {
      f32x4 acc[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) acc[t] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const bf16x8 xf = *reinterpret_cast<const bf16x8*>(&act[p][32 * kk + 8 * q]);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const bf16x8 wf = *reinterpret_cast<const bf16x8*>(&sW[layer * MLP_H + 16 * t + p][32 * kk + 8 * q]);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, xf, acc[t], 0, 0, 0);
        }
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int o = 16 * t + 4 * q;
        const f32x4 bias = *reinterpret_cast<const f32x4*>(&sBias[layer * MLP_H + o]);
        bf16x4 h;
#pragma unroll
        for (int v = 0; v < 4; ++v) h[v] = (__bf16)gelu_f(acc[t][v] + bias[v]);
        *reinterpret_cast<bf16x4*>(&act[p][o]) = h;
      }
      wave_order();
    }
    f32x4 y = {0.0f, 0.0f, 0.0f, 0.0f};
    const bf16x8 zero8 = {};
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const bf16x8 xf = *reinterpret_cast<const bf16x8*>(&act[p][32 * kk + 8 * q]);
      // rows 2 .. 15 of the output tile are zero weights
      const bf16x8 wf = p < MLP_NOUT ? *reinterpret_cast<const bf16x8*>(&sW[3 * MLP_H + p][32 * kk + 8 * q]) : zero8;
      y = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wf, xf, y, 0, 0, 0);
    }
    // lanes q = 0 hold outputs 0 (beta*) and 1 (kappa*) of point p; mu* is zero
    const int i = i0 + p;
    if (q == 0 && i < nz) {
      const long long gi = (long long)i + (long long)pitch * j;
      const double r = a.ramp[i];
      const double b = r * (double)a.out_scale[0] * (double)fmaxf(y[0] + sBias[3 * MLP_H], 0.0f);
      const double k = r * (double)a.out_scale[1] * (double)fmaxf(y[1] + sBias[3 * MLP_H + 1], 0.0f);
      a.out[gi] = 0.0;
      a.out[fs + gi] = b;
      a.out[2 * fs + gi] = k;
      if (a.out2) {
        a.out2[gi] = 0.0;
        a.out2[fs + gi] = b;
        a.out2[2 * fs + gi] = k;
      }
    }
    wave_order();
  }
}

hipError_t launch_mlp_gemm(int epi, const MlpGemm& g, hipStream_t s) {
  const int tiles = (g.M / 32) * (g.N / 32) * g.splits;
  const dim3 grid((unsigned)((tiles + 3) / 4)), block(256);
  switch (epi) {
    case MLP_EPI_FWD_GELU: hipLaunchKernelGGL(mlp_gemm_kernel<MLP_EPI_FWD_GELU>, grid, block, 0, s, g); break;
    case MLP_EPI_FWD_LAST: hipLaunchKernelGGL(mlp_gemm_kernel<MLP_EPI_FWD_LAST>, grid, block, 0, s, g); break;
    case MLP_EPI_BWD_GELU: hipLaunchKernelGGL(mlp_gemm_kernel<MLP_EPI_BWD_GELU>, grid, block, 0, s, g); break;
    default: hipLaunchKernelGGL(mlp_gemm_kernel<MLP_EPI_PARTIAL>, grid, block, 0, s, g); break;
  }
  return hipGetLastError();
}

hipError_t launch_mlp_loss(const float* y, const float* t, int n, __bf16* dz, __bf16* dzT, double* loss, hipStream_t s) {
  hipLaunchKernelGGL(mlp_loss_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y, t, n, dz, dzT, loss);
  return hipGetLastError();
}

hipError_t launch_mlp_bias_grad(const __bf16* dzT, int n, int ncols, float* db, hipStream_t s) {
  hipLaunchKernelGGL(mlp_bias_grad_kernel, dim3((unsigned)ncols), dim3(256), 0, s, dzT, n, db);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(const MlpAdam& a, hipStream_t s) {
  const int n = a.N * a.K + a.N;
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_mlp_rows(const float* x, const float* t, const long long* idx, int n, __bf16* a0, __bf16* a0T,
                           float* tb, hipStream_t s) {
  hipLaunchKernelGGL(mlp_rows_kernel, dim3((unsigned)n), dim3(256), 0, s, x, t, idx, n, a0, a0T, tb);
  return hipGetLastError();
}

hipError_t launch_mlp_sample(unsigned long long seed, long long step, long long nsamples, int n, long long* idx,
                             hipStream_t s) {
  hipLaunchKernelGGL(mlp_sample_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, seed, step, nsamples, n, idx);
  return hipGetLastError();
}

hipError_t launch_mlp_features(const MlpInfer& a, const int* pts, int n, float* out, hipStream_t s) {
  const long long total = (long long)n * MLP_NIN;
  if (total == 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_features_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, pts, n, out);
  return hipGetLastError();
}

hipError_t launch_mlp_infer(const MlpInfer& a, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(mlp_infer_kernel, dim3((unsigned)nblocks), dim3(64 * INFER_WAVES), 0, s, a);
  return hipGetLastError();
}

}  // namespace ajet
