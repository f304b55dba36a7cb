// Pointwise subgrid-stress closure network on BF16 MFMA (gfx950).
//
// Network: 10 -> 64 -> 64 -> 64 -> 6, GELU (the tanh form, below) on the
// hidden layers.  Inputs: the nine resolved velocity gradients over a fixed
// gradient scale G and the filter width over its training value; outputs:
// the six deviatoric stresses T_ij over delta^2 G^2, in the order
// xx, yy, zz, xy, xz, yz.  Weights are stored as BF16 in MFMA operand order,
// every product accumulates in FP32; FP32 master weights and Adam moments
// live beside them for training.
//
// MFMA tiling (v_mfma_f32_16x16x16_bf16, one wavefront per 16 points):
//   A (16 x 16, weights):  lane l holds A[l%16][4(l/16) .. 4(l/16)+3]
//   B (16 x 16, points):   lane l holds B[4(l/16) .. +3][l%16]
//   D (16 x 16, FP32):     lane l holds D[4(l/16) .. +3][l%16]
// D's row index runs like B's K index, so a layer's output tile is the next
// layer's input tile with no data movement: the forward pass of 16 points is
// 4 + 16 + 16 + 4 MFMAs on registers.  The backward pass chains the same way
// through the transposed weights; the weight gradients (K = points) go
// through point-major BF16 activations in HBM.
#include <cstring>
#include <vector>

#include "x3d_internal.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace x3d {

constexpr int MLP_NIN = 10, MLP_NH = 64, MLP_NOUT = 6;
// flat FP32 parameter vector: W1 (64x10), b1, W2 (64x64), b2, W3, b3, W4 (6x64), b4
constexpr int OFF_W1 = 0, OFF_B1 = 640, OFF_W2 = 704, OFF_B2 = 4800, OFF_W3 = 4864, OFF_B3 = 8960,
              OFF_W4 = 9024, OFF_B4 = 9408, MLP_NPARAM = 9414;
static_assert(MLP_NPARAM == MLP_PARAMS, "parameter count");
// operand tiles (64 lanes x 4 BF16 each): forward A tiles, then transposed
// tiles for the backward pass
constexpr int T_A1 = 0, T_A2 = 4, T_A3 = 20, T_A4 = 36, T_AT4 = 40, T_AT3 = 44, T_AT2 = 60, MLP_NTILES = 76;
constexpr int MLP_WG_BLOCKS = 256;   // blocks of the weight-gradient kernel

// ---------------------------------------------------------------- BF16 helpers
// round to nearest even: gfx950's v_cvt_pk_bf16_f32
__device__ __forceinline__ short to_bf16(float f) { return __builtin_bit_cast(short, (__bf16)f); }
__device__ __forceinline__ float from_bf16(short h) { return __uint_as_float(((unsigned)(unsigned short)h) << 16); }

// GELU in the tanh form of Hendrycks & Gimpel (2016),
//   0.5 z (1 + tanh(k (z + 0.044715 z^3))) = z sigma(2 k (z + 0.044715 z^3)),
// k = sqrt(2/pi), sigma the logistic function, evaluated as
// z / (1 + 2^u), u = -2 k log2(e) (z + 0.044715 z^3): five VALU operations,
// one exp2 and one reciprocal, no branches.  It is the network's activation
// in training and inference alike (the oracle's too); it differs from the
// erf form z Phi(z) by at most 4.8e-4, 0.12 BF16 ulp at unit magnitude.
constexpr float GELU_C0 = -2.302208198144325f;    // -2 k log2(e)
constexpr float GELU_C1 = -0.1029432395800235f;   // -2 k log2(e) 0.044715
__device__ __forceinline__ float gelu(float z) {
  const float u = z * fmaf(z * z, GELU_C1, GELU_C0);
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(u));
}
// d/dz [z s(u)] = s + z s (1 - s) (-ln 2) du/dz, du/dz = C0 + 3 C1 z^2
__device__ __forceinline__ float gelu_grad(float z) {
  const float z2 = z * z;
  const float sg = __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * fmaf(z2, GELU_C1, GELU_C0)));
  const float du = fmaf(3.0f * GELU_C1, z2, GELU_C0) * -0.69314718055994531f;
  return fmaf(z * sg * (1.0f - sg), du, sg);
}

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct MlpDev {
  bf16x4 *tiles;     // MLP_NTILES * 64
  float *bias;       // b1 b2 b3 (64 each), b4 padded to 16
  float *master;     // MLP_NPARAM FP32
  float *m, *v;      // Adam moments
  float *grad;       // MLP_NPARAM, reduced gradient
  float *partial;    // MLP_WG_BLOCKS * MLP_NPARAM
  float *loss;       // per-block loss partials (MLP_WG_BLOCKS) + total
  bf16x4 *act;       // point-major activations H0..H3 and dZ1..dZ4 of a batch
  int act_cap;       // points the act buffer holds
  int step;
};

// ---------------------------------------------------------------- packing
// Builds every operand tile and the bias array from the FP32 master weights.
__global__ void mlp_pack_kernel(const float *__restrict__ p, bf16x4 *__restrict__ tiles, float *__restrict__ bias) {
  const int t = blockIdx.x, lane = threadIdx.x;
  const int r16 = lane & 15, kb = lane >> 4;
  int M, K, off, mt, kt;
  bool trans;
  if (t < T_A2) { M = 64; K = 10; off = OFF_W1; mt = t - T_A1; kt = 0; trans = false; }
  else if (t < T_A3) { M = 64; K = 64; off = OFF_W2; mt = (t - T_A2) >> 2; kt = (t - T_A2) & 3; trans = false; }
  else if (t < T_A4) { M = 64; K = 64; off = OFF_W3; mt = (t - T_A3) >> 2; kt = (t - T_A3) & 3; trans = false; }
  else if (t < T_AT4) { M = 6; K = 64; off = OFF_W4; mt = 0; kt = t - T_A4; trans = false; }
  else if (t < T_AT3) { M = 6; K = 64; off = OFF_W4; mt = t - T_AT4; kt = 0; trans = true; }
  else if (t < T_AT2) { M = 64; K = 64; off = OFF_W3; mt = (t - T_AT3) >> 2; kt = (t - T_AT3) & 3; trans = true; }
  else { M = 64; K = 64; off = OFF_W2; mt = (t - T_AT2) >> 2; kt = (t - T_AT2) & 3; trans = true; }
  bf16x4 v;
  for (int j = 0; j < 4; ++j) {
    const int row = 16 * mt + r16, col = 16 * kt + 4 * kb + j;
    float w = 0.0f;
    if (!trans) {
      if (row < M && col < K) w = p[off + row * K + col];   // W[row][col]
    } else {
      if (row < K && col < M) w = p[off + col * K + row];   // W^T[row][col] = W[col][row]
    }
    v[j] = to_bf16(w);
  }
  tiles[t * 64 + lane] = v;
  if (t == 0) {
    for (int i = lane; i < 208; i += 64) {
      float b = 0.0f;
      if (i < 64) b = p[OFF_B1 + i];
      else if (i < 128) b = p[OFF_B2 + i - 64];
      else if (i < 192) b = p[OFF_B3 + i - 128];
      else if (i - 192 < MLP_NOUT) b = p[OFF_B4 + i - 192];
      bias[i] = b;
    }
  }
}

// ---------------------------------------------------------------- forward of 16 points
// x: layer-1 B operand (features 4kb..4kb+3 of point l%16).  Fills the
// pre-activations z[layer][mt] and the BF16 layer inputs h[layer][kt]
// (h[0][0] = x), returns the output tile (rows 4kb..4kb+3 = outputs).
struct Fwd16 {
  f32x4 z[3][4];
  bf16x4 h[3][4];   // inputs of layers 2, 3, 4
};

// At[t]: this lane's operand of tile t (a register array, or GlobalTiles)
struct GlobalTiles {
  const bf16x4 *p;
  int lane;
  __device__ __forceinline__ bf16x4 operator[](int t) const { return p[t * 64 + lane]; }
};

template <typename ACC>
__device__ __forceinline__ f32x4 forward16(const ACC &At, const float *__restrict__ bias, bf16x4 x, int kb,
                                           Fwd16 &F) {
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  // the bias enters as the accumulator input of each tile's first MFMA
#pragma unroll
  for (int mt = 0; mt < 4; ++mt) {
    const f32x4 acc = mfma(At[T_A1 + mt], x, *reinterpret_cast<const f32x4 *>(bias + 16 * mt + 4 * kb));
    F.z[0][mt] = acc;
    bf16x4 h;
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = to_bf16(gelu(acc[i]));
    F.h[0][mt] = h;
  }
#pragma unroll
  for (int layer = 1; layer < 3; ++layer) {
    const int base = layer == 1 ? T_A2 : T_A3;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 acc = *reinterpret_cast<const f32x4 *>(bias + 64 * layer + 16 * mt + 4 * kb);
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) acc = mfma(At[base + 4 * mt + kt], F.h[layer - 1][kt], acc);
      F.z[layer][mt] = acc;
      bf16x4 h;
#pragma unroll
      for (int i = 0; i < 4; ++i) h[i] = to_bf16(gelu(acc[i]));
      F.h[layer][mt] = h;
    }
  }
  f32x4 y = *reinterpret_cast<const f32x4 *>(bias + 192 + 4 * kb);
#pragma unroll
  for (int kt = 0; kt < 4; ++kt) y = mfma(At[T_A4 + kt], F.h[2][kt], y);
  return y;
}

// feature tile from a point-major (N x 10) FP32 array
__device__ __forceinline__ bf16x4 load_features(const float *__restrict__ x, int pt, int n, int kb) {
  bf16x4 v;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int f = 4 * kb + j;
    v[j] = to_bf16((pt < n && f < MLP_NIN) ? x[(size_t)pt * MLP_NIN + f] : 0.0f);
  }
  return v;
}

// ---------------------------------------------------------------- inference on feature arrays
__global__ void __launch_bounds__(256) mlp_infer_kernel(const bf16x4 *__restrict__ tiles, const float *__restrict__ bias,
                                                        const float *__restrict__ x, float *__restrict__ y, int n) {
  __shared__ __align__(16) float sb[208];
  for (int i = threadIdx.x; i < 208; i += blockDim.x) sb[i] = bias[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, kb = lane >> 4;
  bf16x4 At[40];
#pragma unroll
  for (int t = 0; t < 40; ++t) At[t] = tiles[t * 64 + lane];
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwaves = (gridDim.x * blockDim.x) >> 6;
  const int ntiles = (n + 15) / 16;
  for (int tile = wave; tile < ntiles; tile += nwaves) {
    const int pt = tile * 16 + (lane & 15);
    Fwd16 F;
    const f32x4 out = forward16(At, sb, load_features(x, pt, n, kb), kb, F);
    if (pt < n) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int o = 4 * kb + i;
        if (o < MLP_NOUT) y[(size_t)pt * MLP_NOUT + o] = out[i];
      }
    }
  }
}

// ---------------------------------------------------------------- training: forward + backward chain
// Per point (point-major BF16 rows in act): H0 (16: features, padded),
// H1, H2, H3 (64 each), dZ1, dZ2, dZ3 (64 each), dZ4 (16, outputs padded):
// 416 BF16 = 104 bf16x4 per point.  Loss: mean over N*6 of (y - t)^2.
constexpr int ACT_H0 = 0, ACT_H1 = 4, ACT_H2 = 20, ACT_H3 = 36, ACT_D1 = 52, ACT_D2 = 68, ACT_D3 = 84,
              ACT_D4 = 100, ACT_ROW = 104;

__global__ void __launch_bounds__(256) mlp_fwdbwd_kernel(const bf16x4 *__restrict__ tiles,
                                                         const float *__restrict__ bias, const float *__restrict__ x,
                                                         const float *__restrict__ tgt, int n, bf16x4 *__restrict__ act,
                                                         float *__restrict__ loss_part) {
  __shared__ __align__(16) float sb[208];
  __shared__ float red[256];
  for (int i = threadIdx.x; i < 208; i += blockDim.x) sb[i] = bias[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, kb = lane >> 4;
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, nwaves = (gridDim.x * blockDim.x) >> 6;
  const int ntiles = (n + 15) / 16;
  const float gscale = 2.0f / (6.0f * (float)n);
  float lsum = 0.0f;
  for (int tile = wave; tile < ntiles; tile += nwaves) {
    const int pt = tile * 16 + (lane & 15);
    const bool valid = pt < n;
    const bf16x4 xin = load_features(x, pt, n, kb);
    Fwd16 F;
    const f32x4 y = forward16(GlobalTiles{tiles, lane}, sb, xin, kb, F);
    // dL/dy (rows 4kb..4kb+3 = outputs)
    f32x4 dy;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int o = 4 * kb + i;
      float d = 0.0f;
      if (valid && o < MLP_NOUT) {
        const float e = y[i] - tgt[(size_t)pt * MLP_NOUT + o];
        lsum += e * e;
        d = gscale * e;
      }
      dy[i] = d;
    }
    bf16x4 g;
#pragma unroll
    for (int i = 0; i < 4; ++i) g[i] = to_bf16(dy[i]);
    bf16x4 *row = act + (size_t)pt * ACT_ROW;
    if (valid) {
      row[ACT_H0 + kb] = xin;
#pragma unroll
      for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) row[ACT_H1 + 16 * l + 4 * mt + kb] = F.h[l][mt];
      row[ACT_D4 + kb] = g;
    }
    // layer 4 -> dH3 = W4^T dY, dZ3 = dH3 * gelu'(z3)
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    bf16x4 gz[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      f32x4 dh = mfma(tiles[(T_AT4 + mt) * 64 + lane], g, zero);
      bf16x4 q;
#pragma unroll
      for (int i = 0; i < 4; ++i) q[i] = to_bf16(dh[i] * gelu_grad(F.z[2][mt][i]));
      gz[mt] = q;
      if (valid) row[ACT_D3 + 4 * mt + kb] = q;
    }
    // layers 3, 2: dH_{l-1} = W_l^T dZ_l, dZ_{l-1} = dH * gelu'(z_{l-1})
#pragma unroll
    for (int layer = 2; layer >= 1; --layer) {
      const int base = layer == 2 ? T_AT3 : T_AT2;
      bf16x4 nz[4];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        f32x4 dh = zero;
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) dh = mfma(tiles[(base + 4 * mt + kt) * 64 + lane], gz[kt], dh);
        bf16x4 q;
#pragma unroll
        for (int i = 0; i < 4; ++i) q[i] = to_bf16(dh[i] * gelu_grad(F.z[layer - 1][mt][i]));
        nz[mt] = q;
        if (valid) row[(layer == 2 ? ACT_D2 : ACT_D1) + 4 * mt + kb] = q;
      }
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) gz[mt] = nz[mt];
    }
  }
  red[threadIdx.x] = lsum;
  __syncthreads();
  if (threadIdx.x == 0) {
    float s = 0.0f;
    for (int i = 0; i < (int)blockDim.x; ++i) s += red[i];
    loss_part[blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------- training: weight gradients
// dW_l[m][k] = sum_pt dZ_l[pt][m] H_{l-1}[pt][k], db_l[m] = sum_pt dZ_l[pt][m]
// as MFMAs with K = 16 points: A = dZ^T tile, B = H tile.  Each block sums a
// strided set of 16-point chunks (one layer per blockIdx.y) and writes its
// partial gradient; mlp_adam_kernel adds the partials in block order.
__device__ __forceinline__ short act_at(const bf16x4 *__restrict__ act, int pt, int col) {
  const short *r = reinterpret_cast<const short *>(act + (size_t)pt * ACT_ROW);
  return r[col];
}

__global__ void __launch_bounds__(256) mlp_wgrad_kernel(const bf16x4 *__restrict__ act, int n,
                                                        float *__restrict__ partial) {
  const int layer = blockIdx.y;   // 0..3
  // BF16 column offsets of dZ_l and H_{l-1} in a row, their widths
  const int dcol = layer == 0 ? 4 * ACT_D1 : layer == 1 ? 4 * ACT_D2 : layer == 2 ? 4 * ACT_D3 : 4 * ACT_D4;
  const int hcol = layer == 0 ? 4 * ACT_H0 : layer == 1 ? 4 * ACT_H1 : layer == 2 ? 4 * ACT_H2 : 4 * ACT_H3;
  const int M = layer == 3 ? MLP_NOUT : MLP_NH, K = layer == 0 ? MLP_NIN : MLP_NH;
  const int MT = layer == 3 ? 1 : 4, KT = layer == 0 ? 1 : 4;
  const int woff = layer == 0 ? OFF_W1 : layer == 1 ? OFF_W2 : layer == 2 ? OFF_W3 : OFF_W4;
  const int boff = layer == 0 ? OFF_B1 : layer == 1 ? OFF_B2 : layer == 2 ? OFF_B3 : OFF_B4;
  const int lane = threadIdx.x & 63, kb = lane >> 4, r16 = lane & 15;
  const int wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  f32x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = zero;
  float bacc[4] = {0.f, 0.f, 0.f, 0.f};
  const int nch = (n + 15) / 16;
  for (int ch = blockIdx.x * nwv + wv; ch < nch; ch += gridDim.x * nwv) {
    const int p0 = ch * 16;
    bf16x4 A[4], B[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int pt = p0 + 4 * kb + j;
        A[mt][j] = (mt < MT && pt < n) ? act_at(act, pt, dcol + 16 * mt + r16) : (short)0;
        B[mt][j] = (mt < KT && pt < n) ? act_at(act, pt, hcol + 16 * mt + r16) : (short)0;
      }
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int kt = 0; kt < 4; ++kt)
        if (mt < MT && kt < KT) acc[mt][kt] = mfma(A[mt], B[kt], acc[mt][kt]);
    // bias: lanes with kb == 0 sum dZ[pt][16 mt + r16] over the chunk
    if (kb == 0) {
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        if (mt < MT) {
          float s = 0.f;
          for (int q = 0; q < 16; ++q) {
            const int pt = p0 + q;
            if (pt < n) s += from_bf16(act_at(act, pt, dcol + 16 * mt + r16));
          }
          bacc[mt] += s;
        }
      }
    }
  }
  // sum the block's waves through LDS, write the block's partial
  __shared__ float red[4][64 * 64 + 64];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int m = 16 * mt + 4 * kb + i, k = 16 * kt + r16;
        if (mt < MT && kt < KT) red[wv][m * 64 + k] = acc[mt][kt][i];
      }
  if (kb == 0) {
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
      if (mt < MT) red[wv][4096 + 16 * mt + r16] = bacc[mt];
  }
  __syncthreads();
  float *out = partial + (size_t)blockIdx.x * MLP_NPARAM;
  for (int idx = threadIdx.x; idx < M * K + M; idx += blockDim.x) {
    int src, dst;
    if (idx < M * K) {
      const int m = idx / K, k = idx - m * K;
      src = m * 64 + k;
      dst = woff + idx;
    } else {
      src = 4096 + (idx - M * K);
      dst = boff + (idx - M * K);
    }
    float s = 0.f;
    for (int w = 0; w < nwv; ++w) s += red[w][src];
    out[dst] = s;
  }
}

// ---------------------------------------------------------------- Adam
__global__ void mlp_adam_kernel(const float *__restrict__ partial, int nblk, float *__restrict__ grad,
                                float *__restrict__ p, float *__restrict__ m, float *__restrict__ v, float lr,
                                float b1, float b2, float eps, float c1, float c2, int apply) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  float g = 0.f;
  for (int b = 0; b < nblk; ++b) g += partial[(size_t)b * MLP_NPARAM + i];
  grad[i] = g;
  if (!apply) return;
  const float mm = b1 * m[i] + (1.f - b1) * g;
  const float vv = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mm;
  v[i] = vv;
  p[i] = p[i] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void mlp_loss_kernel(const float *__restrict__ part, int nblk, float *__restrict__ out, float inv) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.f;
    for (int b = 0; b < nblk; ++b) s += part[b];
    out[0] = s * inv;
  }
}

// ---------------------------------------------------------------- closure on the LES fields
template <typename T>
struct ClosureArgs {
  const T *g[9];         // gxx gyx gzx gxy gyy gzy gxz gyz gzz (les_models.f90:208-211)
  T *t[6];               // T_xx T_yy T_zz T_xy T_xz T_yz
  float inv_gscale, dfeat, out_scale;
  double cs_d2;          // (Cs*delta)^2 of the fallback
  double beta;           // fallback where T:S < -beta * 2 nut S:S
  int project;           // 1: keep the dissipative projection 2 nu_t S, nu_t = T:S / (2 S:S)
  size_t n;
};

// A wavefront takes 64 points per iteration, one per lane: the nine
// gradients are read once (coalesced), cross-lane reads build the four
// 16-point MFMA input tiles, the outputs come back to the point's own lane,
// and the fallback / projection epilogue runs on all 64 lanes.
template <typename T>
__global__ void __launch_bounds__(256) closure_kernel(const bf16x4 *__restrict__ tiles, const float *__restrict__ bias,
                                                     const ClosureArgs<T> A, int *__restrict__ nfallback) {
  __shared__ __align__(16) float sb[208];
  __shared__ int cnt;
  for (int i = threadIdx.x; i < 208; i += blockDim.x) sb[i] = bias[i];
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, kb = lane >> 4, r16 = lane & 15;
  bf16x4 At[40];
#pragma unroll
  for (int t = 0; t < 40; ++t) At[t] = tiles[t * 64 + lane];
  const long long wave = (blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 6;
  const long long nwaves = (gridDim.x * (long long)blockDim.x) >> 6;
  const long long nchunks = (long long)((A.n + 63) / 64);
  int local = 0;
  for (long long chunk = wave; chunk < nchunks; chunk += nwaves) {
    const size_t pt = (size_t)chunk * 64 + lane;
    const bool valid = pt < A.n;
    T g[9];
    float xf[9];
#pragma unroll
    for (int f = 0; f < 9; ++f) {
      g[f] = valid ? A.g[f][pt] : T(0);
      xf[f] = (float)g[f] * A.inv_gscale;
    }
    float yo[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma nounroll
    for (int t = 0; t < 4; ++t) {
      // input tile t: column r16 is point 16t + r16, rows 4kb..4kb+3 its features
      const int src = 16 * t + r16;
      bf16x4 x;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float c0 = __shfl(xf[j], src), c1 = __shfl(xf[4 + j], src);
        const float c2 = j == 0 ? __shfl(xf[8], src) : (j == 1 ? A.dfeat : 0.f);
        x[j] = to_bf16(kb == 0 ? c0 : (kb == 1 ? c1 : (kb == 2 ? c2 : 0.f)));
      }
      Fwd16 F;
      const f32x4 y = forward16(At, sb, x, kb, F);
      // outputs 0..3 of column r16 sit in lane r16, outputs 4..5 in lane 16 + r16
      const float o0 = __shfl(y[0], r16), o1 = __shfl(y[1], r16), o2 = __shfl(y[2], r16), o3 = __shfl(y[3], r16);
      const float o4 = __shfl(y[0], 16 + r16), o5 = __shfl(y[1], 16 + r16);
      if (kb == t) {
        yo[0] = o0; yo[1] = o1; yo[2] = o2; yo[3] = o3; yo[4] = o4; yo[5] = o5;
      }
    }
    if (valid) {
      // S_ij as les_models.f90:214-269
      const T sxx = g[0], syy = g[4], szz = g[8];
      const T sxy = T(0.5) * (g[3] + g[1]), sxz = T(0.5) * (g[2] + g[6]), syz = T(0.5) * (g[7] + g[5]);
      const T ss = sxx * sxx + syy * syy + szz * szz + T(2) * sxy * sxy + T(2) * sxz * sxz + T(2) * syz * syz;
      const T nut = (T)A.cs_d2 * sqrt(T(2) * ss);
      T t[6];
#pragma unroll
      for (int c = 0; c < 6; ++c) t[c] = (T)yo[c] * (T)A.out_scale;
      const T pi = t[0] * sxx + t[1] * syy + t[2] * szz + T(2) * (t[3] * sxy + t[4] * sxz + t[5] * syz);
      if (pi < -(T)A.beta * T(2) * nut * ss) {
        t[0] = T(2) * nut * sxx; t[1] = T(2) * nut * syy; t[2] = T(2) * nut * szz;
        t[3] = T(2) * nut * sxy; t[4] = T(2) * nut * sxz; t[5] = T(2) * nut * syz;
        ++local;
      } else if (A.project) {
        // the part of T along S: an eddy viscosity nu_t = max(T:S, 0) / (2 S:S)
        const T nt = (ss > T(0) && pi > T(0)) ? pi / (T(2) * ss) : T(0);
        t[0] = T(2) * nt * sxx; t[1] = T(2) * nt * syy; t[2] = T(2) * nt * szz;
        t[3] = T(2) * nt * sxy; t[4] = T(2) * nt * sxz; t[5] = T(2) * nt * syz;
      }
#pragma unroll
      for (int c = 0; c < 6; ++c) A.t[c][pt] = t[c];
    }
  }
  if (nfallback) {
    atomicAdd(&cnt, local);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(nfallback, cnt);
  }
}

// ---------------------------------------------------------------- host side
int mlp_alloc(MlpDev *&out) {
  auto *M = new MlpDev();
  std::memset(M, 0, sizeof(MlpDev));
  hipError_t e = hipSuccess;
  e = e == hipSuccess ? hipMalloc(&M->tiles, sizeof(bf16x4) * MLP_NTILES * 64) : e;
  e = e == hipSuccess ? hipMalloc(&M->bias, sizeof(float) * 208) : e;
  e = e == hipSuccess ? hipMalloc(&M->master, sizeof(float) * MLP_NPARAM) : e;
  e = e == hipSuccess ? hipMalloc(&M->m, sizeof(float) * MLP_NPARAM) : e;
  e = e == hipSuccess ? hipMalloc(&M->v, sizeof(float) * MLP_NPARAM) : e;
  e = e == hipSuccess ? hipMalloc(&M->grad, sizeof(float) * MLP_NPARAM) : e;
  e = e == hipSuccess ? hipMalloc(&M->partial, sizeof(float) * MLP_NPARAM * MLP_WG_BLOCKS) : e;
  e = e == hipSuccess ? hipMalloc(&M->loss, sizeof(float) * (MLP_WG_BLOCKS + 1)) : e;
  if (e == hipSuccess) e = hipMemset(M->m, 0, sizeof(float) * MLP_NPARAM);
  if (e == hipSuccess) e = hipMemset(M->v, 0, sizeof(float) * MLP_NPARAM);
  if (e == hipSuccess) e = hipMemset(M->master, 0, sizeof(float) * MLP_NPARAM);
  if (e != hipSuccess) {
    mlp_free(M);
    return fail(X3D_ERR_HIP, std::string("mlp alloc: ") + hipGetErrorString(e));
  }
  out = M;
  return X3D_OK;
}

void mlp_free(MlpDev *M) {
  if (!M) return;
  void *ps[] = {M->tiles, M->bias, M->master, M->m, M->v, M->grad, M->partial, M->loss, M->act};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  delete M;
}

int mlp_repack(MlpDev *M, hipStream_t st) {
  hipLaunchKernelGGL(mlp_pack_kernel, dim3(MLP_NTILES), dim3(64), 0, st, M->master, M->tiles, M->bias);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

int mlp_set_params(MlpDev *M, const float *host, hipStream_t st) {
  X3D_HIP(hipMemcpyAsync(M->master, host, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, st));
  X3D_HIP(hipMemsetAsync(M->m, 0, sizeof(float) * MLP_NPARAM, st));
  X3D_HIP(hipMemsetAsync(M->v, 0, sizeof(float) * MLP_NPARAM, st));
  M->step = 0;
  X3D_TRY(mlp_repack(M, st));
  X3D_HIP(hipStreamSynchronize(st));
  return X3D_OK;
}

int mlp_get_params(const MlpDev *M, float *host, hipStream_t st) {
  X3D_HIP(hipMemcpyAsync(host, M->master, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, st));
  X3D_HIP(hipStreamSynchronize(st));
  return X3D_OK;
}

int mlp_forward(const MlpDev *M, const float *x, float *y, int n, hipStream_t st) {
  if (n <= 0) return fail(X3D_ERR_ARG, "mlp: n must be positive");
  const int ntiles = (n + 15) / 16;
  const int blocks = std::min((ntiles + 3) / 4, 2048);
  hipLaunchKernelGGL(mlp_infer_kernel, dim3(blocks), dim3(256), 0, st, M->tiles, M->bias, x, y, n);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}

// One batch: forward + backward, gradient reduction, optionally Adam.
int mlp_train_batch(MlpDev *M, const float *x, const float *t, int n, int apply, float lr, float *loss_host,
                    hipStream_t st) {
  if (n <= 0) return fail(X3D_ERR_ARG, "mlp: n must be positive");
  if (M->act_cap < n) {
    if (M->act) X3D_HIP(hipFree(M->act));
    M->act = nullptr;
    X3D_HIP(hipMalloc(&M->act, sizeof(bf16x4) * ACT_ROW * (size_t)n));
    M->act_cap = n;
  }
  const int ntiles = (n + 15) / 16;
  const int fb = std::min((ntiles + 3) / 4, MLP_WG_BLOCKS);
  hipLaunchKernelGGL(mlp_fwdbwd_kernel, dim3(fb), dim3(256), 0, st, M->tiles, M->bias, x, t, n, M->act, M->loss);
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(1), dim3(64), 0, st, M->loss, fb, M->loss + MLP_WG_BLOCKS,
                     1.0f / (6.0f * (float)n));
  hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(MLP_WG_BLOCKS, 4), dim3(256), 0, st, M->act, n, M->partial);
  float c1 = 1.f, c2 = 1.f;
  const float b1 = 0.9f, b2 = 0.999f;
  if (apply) {
    M->step += 1;
    c1 = 1.f - powf(b1, (float)M->step);
    c2 = 1.f - powf(b2, (float)M->step);
  }
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, M->partial, MLP_WG_BLOCKS,
                     M->grad, M->master, M->m, M->v, lr, b1, b2, 1e-8f, c1, c2, apply);
  X3D_HIP(hipGetLastError());
  if (apply) X3D_TRY(mlp_repack(M, st));
  if (loss_host) {
    X3D_HIP(hipMemcpyAsync(loss_host, M->loss + MLP_WG_BLOCKS, sizeof(float), hipMemcpyDeviceToHost, st));
    X3D_HIP(hipStreamSynchronize(st));
  }
  return X3D_OK;
}

int mlp_get_grad(const MlpDev *M, float *host, hipStream_t st) {
  X3D_HIP(hipMemcpyAsync(host, M->grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, st));
  X3D_HIP(hipStreamSynchronize(st));
  return X3D_OK;
}

template <typename T>
int launch_closure(const MlpDev *M, const T *const g[9], T *const t[6], size_t n, double gscale, double dfeat,
                   double delta, double cs, double beta, int project, int *nfallback, hipStream_t st) {
  ClosureArgs<T> A{};
  A.project = project;
  for (int i = 0; i < 9; ++i) A.g[i] = g[i];
  for (int i = 0; i < 6; ++i) A.t[i] = t[i];
  A.inv_gscale = (float)(1.0 / gscale);
  A.dfeat = (float)dfeat;
  A.out_scale = (float)(delta * delta * gscale * gscale);
  A.cs_d2 = (cs * delta) * (cs * delta);
  A.beta = beta;
  A.n = n;
  const size_t nchunks = (n + 63) / 64;
  const int blocks = (int)std::min<size_t>((nchunks + 3) / 4, 4096);
  hipLaunchKernelGGL((closure_kernel<T>), dim3(blocks), dim3(256), 0, st, M->tiles, M->bias, A, nfallback);
  X3D_HIP(hipGetLastError());
  return X3D_OK;
}
template int launch_closure<double>(const MlpDev *, const double *const[9], double *const[6], size_t, double, double,
                                    double, double, double, int, int *, hipStream_t);
template int launch_closure<float>(const MlpDev *, const float *const[9], float *const[6], size_t, double, double,
                                   double, double, double, int, int *, hipStream_t);

}  // namespace x3d
