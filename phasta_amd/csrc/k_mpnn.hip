// Learned preconditioner on gfx950: a 3-layer message-passing network over
// the periodic node graph (hidden width 64), bf16 operands on MFMA with
// float accumulation, and its hand-written backward pass for training.
//
//   x_i  = [Dn_i / dscale (25), v_i * s (5), 0, 0],   s = sqrt(n_act)/||v||
//   Dn_i = diag(d)^-1 BDiag_i diag(d)^-1,  d_r = sqrt|BDiag_i,rr|
//   h0   = bf(silu(x W0 + b0))
//   A_i  = bf([h_i, mean_j h_j, (1/deg_i) sum_j (h_j - h_i) (x) e_ij])    (320)
//   h    = bf(h + silu(A W_l + b_l)),  l = 1..3
//   y_i  = h_i Wout + bout,            z = v + y / s
//
// The network acts on vectors of the block-diagonally scaled system, where
// the block-Jacobi preconditioner is the identity; it learns the correction
// to it.  Rows of periodic slaves have no edges and stay zero.
//
// Layouts (row-major, node-major): h, A in bf16 (n x 64, n x 320); weights
// transposed, Wt[col][k] (64 x K) bf16, so one lane reads 8 consecutive k
// of one column with a 16-byte load, which is the MFMA B-operand layout.
#include <hip/hip_bf16.h>

#include "phg_internal.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kH = 64;
constexpr int kNin = 32;
constexpr int kNagg = 5 * kH;

__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// round to nearest even (NaN kept quiet)
__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline __bf16 u2bf(uint16_t u) { return __builtin_bit_cast(__bf16, u); }

__device__ inline float silu(float u) { return u / (1.0f + __expf(-u)); }
__device__ inline float dsilu(float u) {
  const float s = 1.0f / (1.0f + __expf(-u));
  return s * (1.0f + u * (1.0f - s));
}

// ---------------- inputs ----------------
// x (n x 32 bf16) from BDiag (25 planes, double) and v (5 planes) with scale s
__global__ void k_mp_features(const double* __restrict__ bdiag, const double* __restrict__ v,
                              const float* __restrict__ dscale, const int* __restrict__ ptr,
                              const double* __restrict__ sptr, uint16_t* __restrict__ x,
                              long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool act = ptr[i + 1] > ptr[i];
  const float s = (float)sptr[0];
  // symmetric diagonal scaling: Dn = diag(d)^-1 D diag(d)^-1, d_r = sqrt|D_rr|
  double d[5];
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    const double t = sqrt(fabs(bdiag[(size_t)(6 * r) * n + i]));
    d[r] = t > 0.0 ? t : 1.0;
  }
  uint16_t xv[kNin];
#pragma unroll
  for (int cc = 0; cc < 5; ++cc)
#pragma unroll
    for (int r = 0; r < 5; ++r) {
      const int e = r + 5 * cc;
      xv[e] = act ? f2bf((float)(bdiag[(size_t)e * n + i] / (d[r] * d[cc])) / dscale[e]) : 0;
    }
#pragma unroll
  for (int f = 0; f < 5; ++f) xv[25 + f] = act ? f2bf((float)v[(size_t)f * n + i] * s) : 0;
  xv[30] = 0;
  xv[31] = 0;
  // the node's 64-byte input row as four 16-byte stores
  uint4* xo = reinterpret_cast<uint4*>(x + (size_t)i * kNin);
#pragma unroll
  for (int t = 0; t < kNin / 8; ++t) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      w[q] = (uint32_t)xv[8 * t + 2 * q] | ((uint32_t)xv[8 * t + 2 * q + 1] << 16);
    xo[t] = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// ---------------- aggregation ----------------
// Sixteen lanes per node, four channels per lane (8-byte loads), so one
// neighbour row of 128 bytes is one coalesced load of the node's lanes;
// four nodes per wavefront.
__device__ inline void ld4(const uint16_t* p, float v[4]) {
  const uint2 w = *reinterpret_cast<const uint2*>(p);
  v[0] = __uint_as_float(w.x << 16);
  v[1] = __uint_as_float(w.x & 0xffff0000u);
  v[2] = __uint_as_float(w.y << 16);
  v[3] = __uint_as_float(w.y & 0xffff0000u);
}
__device__ inline void st4(uint16_t* p, const float v[4]) {
  uint2 w;
  w.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
  w.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
  *reinterpret_cast<uint2*>(p) = w;
}

__device__ inline void ld8(const uint16_t* p, float v[8]) {
  const uint4 w = *reinterpret_cast<const uint4*>(p);
  const uint32_t x[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[2 * k] = __uint_as_float(x[k] << 16);
    v[2 * k + 1] = __uint_as_float(x[k] & 0xffff0000u);
  }
}
__device__ inline void st8(uint16_t* p, const float v[8]) {
  uint32_t x[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) x[k] = (uint32_t)f2bf(v[2 * k]) | ((uint32_t)f2bf(v[2 * k + 1]) << 16);
  *reinterpret_cast<uint4*>(p) = make_uint4(x[0], x[1], x[2], x[3]);
}

__global__ void __launch_bounds__(256) k_mp_agg(const uint16_t* __restrict__ h,
                                                const int* __restrict__ ptr,
                                                const int* __restrict__ col,
                                                const float4* __restrict__ ev,
                                                const float* __restrict__ inv,
                                                uint16_t* __restrict__ A, long long n) {
  const long long i = blockIdx.x * 16LL + threadIdx.x / 16;
  const int c0 = 4 * (threadIdx.x & 15);
  if (i >= n) return;
  float hi[4], m[4] = {0.f, 0.f, 0.f, 0.f}, g0[4] = {0.f, 0.f, 0.f, 0.f},
               g1[4] = {0.f, 0.f, 0.f, 0.f}, g2[4] = {0.f, 0.f, 0.f, 0.f};
  ld4(h + (size_t)i * kH + c0, hi);
  const int k0 = ptr[i], k1 = ptr[i + 1];
#pragma unroll 2
  for (int k = k0; k < k1; ++k) {
    const int j = col[k];
    const float4 e = ev[k];
    float hj[4];
    ld4(h + (size_t)j * kH + c0, hj);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float d = hj[t] - hi[t];
      m[t] += hj[t];
      g0[t] += d * e.x;
      g1[t] += d * e.y;
      g2[t] += d * e.z;
    }
  }
  const float w = inv[i];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    m[t] *= w;
    g0[t] *= w;
    g1[t] *= w;
    g2[t] *= w;
  }
  uint16_t* a = A + (size_t)i * kNagg;
  *reinterpret_cast<uint2*>(a + c0) = *reinterpret_cast<const uint2*>(h + (size_t)i * kH + c0);
  st4(a + kH + c0, m);
  st4(a + 2 * kH + c0, g0);
  st4(a + 3 * kH + c0, g1);
  st4(a + 4 * kH + c0, g2);
}

// Fused inference layer: the aggregation of a 32-row tile goes to LDS and
// h_out = silu(A W + b) + h_in is formed on MFMA from there, so A never
// travels through HBM (training keeps the separate k_mp_agg / k_mp_dense
// pair, whose A and pre-activations the backward pass reads).  All four
// wavefronts aggregate (16 lanes per node, 4 channels per lane, 16 nodes
// per pass, as k_mp_agg); wavefront w then forms output columns 16w..16w+15
// of the 32 rows with its slice of W held in registers, and the output tile
// is staged in LDS and written as full 128-byte rows.  Same arithmetic, in
// the same order, as k_mp_agg followed by k_mp_dense<kNagg>.
constexpr int kLayerRows = 32;  // 25 KB of LDS per workgroup (64 rows: 1.25x slower)
constexpr int kLdsRow = kNagg + 8;  // padded bf16 row of the A tile: 656 bytes

__global__ void __launch_bounds__(256) k_mp_layer(const uint16_t* __restrict__ h,
                                                  const int* __restrict__ ptr,
                                                  const int* __restrict__ col,
                                                  const float4* __restrict__ ev,
                                                  const float* __restrict__ inv,
                                                  const uint16_t* __restrict__ Wt,
                                                  const float* __restrict__ bias,
                                                  uint16_t* __restrict__ hout, long long n) {
  __shared__ __align__(16) uint16_t At[kLayerRows * kLdsRow];
  __shared__ __align__(16) uint16_t Ot[kLayerRows * kH];
  // XCD-strided tiles (grid a multiple of 8): each XCD sweeps a contiguous
  // range of rows, so the neighbour rows it gathers stay in its L2
  const long long per = (gridDim.x + 7) / 8;
  const long long r0 = ((blockIdx.x % 8) * per + blockIdx.x / 8) * (long long)kLayerRows;
  if (r0 >= n) return;  // whole workgroup
  // eight lanes per row, eight channels each (16-byte reads of the
  // neighbour rows: 1.27 ms per layer at 128^3 with sixteen lanes of four)
  static_assert(kLayerRows % 32 == 0, "32 rows per pass");
  const int c0 = 8 * (threadIdx.x & 7);
  for (int pass = 0; pass < kLayerRows / 32; ++pass) {
    const int lr = pass * 32 + (int)(threadIdx.x >> 3);
    const long long i = r0 + lr;
    uint16_t* a = At + lr * kLdsRow;
    float hi[8], m[8], g0[8], g1[8], g2[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) hi[t] = m[t] = g0[t] = g1[t] = g2[t] = 0.f;
    if (i < n) {
      ld8(h + (size_t)i * kH + c0, hi);
      const int k0 = ptr[i], k1 = ptr[i + 1];
      auto acc_nb = [&](const float hj[8], const float4 e) {
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          const float d = hj[t] - hi[t];
          m[t] += hj[t];
          g0[t] += d * e.x;
          g1[t] += d * e.y;
          g2[t] += d * e.z;
        }
      };
      int k = k0;
      // four neighbour rows in flight, accumulated in edge order
      for (; k + 4 <= k1; k += 4) {
        float hj[4][8];
        float4 e[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          e[u] = ev[k + u];
          ld8(h + (size_t)col[k + u] * kH + c0, hj[u]);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc_nb(hj[u], e[u]);
      }
      for (; k < k1; ++k) {
        float hj[8];
        ld8(h + (size_t)col[k] * kH + c0, hj);
        acc_nb(hj, ev[k]);
      }
      const float w = inv[i];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        m[t] *= w;
        g0[t] *= w;
        g1[t] *= w;
        g2[t] *= w;
      }
      *reinterpret_cast<uint4*>(a + c0) = *reinterpret_cast<const uint4*>(h + (size_t)i * kH + c0);
    } else {
      *reinterpret_cast<uint4*>(a + c0) = make_uint4(0u, 0u, 0u, 0u);
    }
    st8(a + kH + c0, m);
    st8(a + 2 * kH + c0, g0);
    st8(a + 3 * kH + c0, g1);
    st8(a + 4 * kH + c0, g2);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lrow = lane & 15, lk = (lane >> 4) * 8;
  bf16x8 b[kNagg / 32];
#pragma unroll
  for (int ks = 0; ks < kNagg / 32; ++ks)
    b[ks] = *reinterpret_cast<const bf16x8*>(Wt + (size_t)(16 * wave + lrow) * kNagg + ks * 32 + lk);
  const int cc = 16 * wave + lrow;
  const float bb = bias[cc];
#pragma unroll
  for (int rt = 0; rt < kLayerRows / 16; ++rt) {
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kNagg / 32; ++ks) {
      const bf16x8 av =
          *reinterpret_cast<const bf16x8*>(At + (rt * 16 + lrow) * kLdsRow + ks * 32 + lk);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b[ks], acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int lr2 = rt * 16 + (lane >> 4) * 4 + q;
      const long long row = r0 + lr2;
      const float u = acc[q] + bb;
      const float hv = silu(u) + bf2f(At[lr2 * kLdsRow + cc]);  // A[., 0:64] is h_in
      const bool act = row < n && ptr[row + 1] > ptr[row];
      Ot[lr2 * kH + cc] = act ? f2bf(hv) : (uint16_t)0;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < kLayerRows * kH / 8; t += blockDim.x) {
    const int lr2 = t >> 3, ch = t & 7;
    const long long row = r0 + lr2;
    if (row < n)
      *reinterpret_cast<uint4*>(hout + (size_t)row * kH + ch * 8) =
          *reinterpret_cast<const uint4*>(Ot + lr2 * kH + ch * 8);
  }
}

// Adjoint of k_mp_agg:
//   dh_m += dA_m[0:64] - inv_m sum_c esum_m[c] dgrad_m[c]
//           + sum_{j in N(m)} inv_j (dmean_j + sum_c e_jm[c] dgrad_j[c])
// with e_jm the vector of the reverse edge (rev[k]) and esum_m = sum_j e_mj.
// Rows without edges get dh = 0.  Same lane layout as k_mp_agg.
__global__ void __launch_bounds__(256) k_mp_agg_bwd(const float* __restrict__ dA,
                                                    const int* __restrict__ ptr,
                                                    const int* __restrict__ col,
                                                    const int* __restrict__ rev,
                                                    const float4* __restrict__ ev,
                                                    const float4* __restrict__ esum,
                                                    const float* __restrict__ inv,
                                                    float* __restrict__ dh, long long n) {
  const long long m = blockIdx.x * 16LL + threadIdx.x / 16;
  const int c0 = 4 * (threadIdx.x & 15);
  if (m >= n) return;
  float4* out = reinterpret_cast<float4*>(dh + (size_t)m * kH + c0);
  const int k0 = ptr[m], k1 = ptr[m + 1];
  if (k1 == k0) {
    *out = make_float4(0.f, 0.f, 0.f, 0.f);
    return;
  }
  const float* am = dA + (size_t)m * kNagg;
  const float4 es = esum[m];
  const float wm = inv[m];
  const float4 d0 = *out;
  const float4 s0 = *reinterpret_cast<const float4*>(am + c0);
  const float4 a2 = *reinterpret_cast<const float4*>(am + 2 * kH + c0);
  const float4 a3 = *reinterpret_cast<const float4*>(am + 3 * kH + c0);
  const float4 a4 = *reinterpret_cast<const float4*>(am + 4 * kH + c0);
  float acc[4];
  acc[0] = d0.x + s0.x - wm * (es.x * a2.x + es.y * a3.x + es.z * a4.x);
  acc[1] = d0.y + s0.y - wm * (es.x * a2.y + es.y * a3.y + es.z * a4.y);
  acc[2] = d0.z + s0.z - wm * (es.x * a2.z + es.y * a3.z + es.z * a4.z);
  acc[3] = d0.w + s0.w - wm * (es.x * a2.w + es.y * a3.w + es.z * a4.w);
#pragma unroll 2
  for (int k = k0; k < k1; ++k) {
    const int j = col[k];
    const float4 e = ev[rev[k]];
    const float wj = inv[j];
    const float* aj = dA + (size_t)j * kNagg;
    const float4 b1 = *reinterpret_cast<const float4*>(aj + kH + c0);
    const float4 b2 = *reinterpret_cast<const float4*>(aj + 2 * kH + c0);
    const float4 b3 = *reinterpret_cast<const float4*>(aj + 3 * kH + c0);
    const float4 b4 = *reinterpret_cast<const float4*>(aj + 4 * kH + c0);
    acc[0] += wj * (b1.x + e.x * b2.x + e.y * b3.x + e.z * b4.x);
    acc[1] += wj * (b1.y + e.x * b2.y + e.y * b3.y + e.z * b4.y);
    acc[2] += wj * (b1.z + e.x * b2.z + e.y * b3.z + e.z * b4.z);
    acc[3] += wj * (b1.w + e.x * b2.w + e.y * b3.w + e.z * b4.w);
  }
  *out = make_float4(acc[0], acc[1], acc[2], acc[3]);
}

// ---------------- dense layers on MFMA ----------------
// U = A W + b for a 64-row tile per block (4 waves x 16 rows), 64 columns.
// v_mfma_f32_16x16x32_bf16: lane l holds A[row l%16][k 8(l/16)..+7],
// B[k 8(l/16)..+7][col l%16]; the result lane holds rows 4(l/16)..+3 of
// column l%16.
// mode 0: h = bf(silu(U))          (encoder)
// mode 1: h = bf(h_in + silu(U))   (message-passing layer)
// Inactive rows (no edges) produce h = 0.  U is kept (float) when Uout != 0.
template <int K>
__global__ void __launch_bounds__(256) k_mp_dense(const uint16_t* __restrict__ A,
                                                  const uint16_t* __restrict__ Wt,
                                                  const float* __restrict__ bias,
                                                  const uint16_t* __restrict__ hin,
                                                  const int* __restrict__ ptr,
                                                  uint16_t* __restrict__ hout,
                                                  float* __restrict__ Uout, int mode,
                                                  long long n) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long r0 = blockIdx.x * 64LL + wave * 16;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  const long long arow = r0 + lr;
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < K / 32; ++ks) {
    bf16x8 a;
    if (arow < n)
      a = *reinterpret_cast<const bf16x8*>(A + (size_t)arow * K + ks * 32 + lk);
    else
      a = bf16x8{};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(Wt + (size_t)(16 * t + lr) * K + ks * 32 + lk);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int cc = 16 * t + lr;
    const float bb = bias[cc];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long row = r0 + (lane >> 4) * 4 + q;
      if (row >= n) continue;
      const float u = acc[t][q] + bb;
      if (Uout) Uout[(size_t)row * kH + cc] = u;
      const bool act = ptr[row + 1] > ptr[row];
      float hv = silu(u);
      if (mode == 1) hv += bf2f(hin[(size_t)row * kH + cc]);
      hout[(size_t)row * kH + cc] = act ? f2bf(hv) : (uint16_t)0;
    }
  }
}

// dA = bf(dU) W^T for one layer: (n x 64) x (64 x 320), float result.  The
// B operand is B[c][f] = W[f][c]: lane reads W[f][c..c+7] from the row-major
// bf16 copy Wrow (K x 64).
__global__ void __launch_bounds__(256) k_mp_dA(const float* __restrict__ dU,
                                               const uint16_t* __restrict__ Wrow,
                                               float* __restrict__ dA, long long n) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long r0 = blockIdx.x * 64LL + wave * 16;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  const long long arow = r0 + lr;
  bf16x8 a[2];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int e = 0; e < 8; ++e)
      a[ks][e] = u2bf(arow < n ? f2bf(dU[(size_t)arow * kH + ks * 32 + lk + e]) : (uint16_t)0);
  for (int t = 0; t < kNagg / 16; ++t) {
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const bf16x8 b =
          *reinterpret_cast<const bf16x8*>(Wrow + (size_t)(16 * t + lr) * kH + ks * 32 + lk);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[ks], b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const long long row = r0 + (lane >> 4) * 4 + q;
      if (row < n) dA[(size_t)row * kNagg + 16 * t + lr] = acc[q];
    }
  }
}

// dU = dh * silu'(U) (zero on inactive rows), bias gradient partials.
__global__ void k_mp_dact(const float* __restrict__ dh, const float* __restrict__ U,
                          const int* __restrict__ ptr, float* __restrict__ dU, long long n) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= n * kH) return;
  const long long row = t / kH;
  const bool act = ptr[row + 1] > ptr[row];
  dU[t] = act ? dh[t] * dsilu(U[t]) : 0.f;
}

// Weight gradient partials: dW[k][c] = sum_rows A[row][k] dU[row][c] over a
// block's share of rows, MFMA with the rows as the reduction dimension.
// Tiles of 32 rows are staged in LDS (A transposed on the way in).
template <int K>
__global__ void __launch_bounds__(256) k_mp_dW(const uint16_t* __restrict__ A,
                                               const float* __restrict__ dU,
                                               float* __restrict__ part, long long n) {
  constexpr int NT = (K / 16) * 4;      // 16x16 output tiles
  constexpr int TPW = NT / 4;           // tiles per wave
  static_assert(NT % 4 == 0, "tiles must split evenly over the 4 waves");
  __shared__ uint16_t sA[K][32 + 8];    // A^T tile: [k][row]
  __shared__ uint16_t sU[kH][32 + 8];   // dU^T tile: [c][row]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  f32x4 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long long base = blockIdx.x * 32LL; base < n; base += (long long)gridDim.x * 32) {
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * K; e += 256) {
      const int r = e / K, k = e % K;
      const long long row = base + r;
      sA[k][r] = row < n ? A[(size_t)row * K + k] : (uint16_t)0;
    }
    for (int e = threadIdx.x; e < 32 * kH; e += 256) {
      const int r = e / kH, c = e % kH;
      const long long row = base + r;
      sU[c][r] = row < n ? f2bf(dU[(size_t)row * kH + c]) : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
      const int tile = wave * TPW + t;
      const int tk = tile / 4, tc = tile % 4;  // output rows k = 16 tk.., cols c = 16 tc..
      bf16x8 a, b;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        a[e] = u2bf(sA[16 * tk + lr][lk + e]);
        b[e] = u2bf(sU[16 * tc + lr][lk + e]);
      }
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
  }
  float* out = part + (size_t)blockIdx.x * K * kH;
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = wave * TPW + t;
    const int tk = tile / 4, tc = tile % 4;
#pragma unroll
    for (int q = 0; q < 4; ++q) out[(size_t)(16 * tk + (lane >> 4) * 4 + q) * kH + 16 * tc + lr] = acc[t][q];
  }
}

// sum of nb partial arrays of m floats, fixed order
__global__ void k_mp_sum_parts(const float* __restrict__ part, int nb, int m, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * m + i];
  out[i] = s;
}

// column sums of dU (bias gradient), fixed order per column
__global__ void k_mp_colsum(const float* __restrict__ dU, long long n, float* __restrict__ out) {
  __shared__ float red[256];
  const int c = blockIdx.x;  // column
  float s = 0.f;
  for (long long r = threadIdx.x; r < n; r += blockDim.x) s += dU[(size_t)r * kH + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = red[0];
}

// ---------------- output ----------------
// y = h Wout + bout (one thread per node); z = v + y / s.  With ystar given,
// also dy = 2 (y - ystar)/n_act into dY and the squared error partials.
__global__ void __launch_bounds__(256) k_mp_out(const uint16_t* __restrict__ h,
                                                const float* __restrict__ Wout,  // 64 x 5
                                                const float* __restrict__ bout,
                                                const int* __restrict__ ptr,
                                                const double* __restrict__ v,
                                                const double* __restrict__ sptr,
                                                double* __restrict__ z, float* __restrict__ Y,
                                                long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool act = ptr[i + 1] > ptr[i];
  float y[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) y[f] = bout[f];
  // the node's 128-byte hidden row as eight 16-byte loads, k ascending
  const uint4* hp = reinterpret_cast<const uint4*>(h + (size_t)i * kH);
#pragma unroll
  for (int t = 0; t < kH / 8; ++t) {
    const uint4 u = hp[t];
    const uint32_t ws[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int k = 8 * t + q;
      const uint32_t wq = ws[q >> 1];
      const float hv = __uint_as_float((q & 1) ? (wq & 0xffff0000u) : (wq << 16));
#pragma unroll
      for (int f = 0; f < 5; ++f) y[f] += hv * Wout[k * 5 + f];
    }
  }
  if (!act) {
#pragma unroll
    for (int f = 0; f < 5; ++f) y[f] = 0.f;
  }
  if (Y) {
#pragma unroll
    for (int f = 0; f < 5; ++f) Y[(size_t)i * 5 + f] = y[f];
  }
  if (z) {
    const double is = sptr[0] > 0.0 ? 1.0 / sptr[0] : 0.0;
#pragma unroll
    for (int f = 0; f < 5; ++f) z[(size_t)f * n + i] = v[(size_t)f * n + i] + is * (double)y[f];
  }
}

// dy = 2 (y - ystar)/na on active rows; dh3 = dy Wout^T; dWout, dbout partials
// are formed by k_mp_out_grad_w.
__global__ void k_mp_out_bwd(const float* __restrict__ Y, const float* __restrict__ Ystar,
                             const int* __restrict__ ptr, const float* __restrict__ Wout,
                             float inv_na, float* __restrict__ dY, float* __restrict__ dh,
                             float* __restrict__ sq, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool act = ptr[i + 1] > ptr[i];
  float d[5];
  float s = 0.f;
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    const float e = act ? Y[i * 5 + f] - Ystar[i * 5 + f] : 0.f;
    s += e * e;
    d[f] = 2.f * inv_na * e;
    dY[i * 5 + f] = d[f];
  }
  sq[i] = s;
  for (int k = 0; k < kH; ++k) {
    float a = 0.f;
#pragma unroll
    for (int f = 0; f < 5; ++f) a += d[f] * Wout[k * 5 + f];
    dh[(size_t)i * kH + k] = act ? a : 0.f;
  }
}

// dY = scale * dz / s on active rows (dz: 5 planes), scale = -2 / ||v||^2
// for the residual loss ||K z - v||^2 / ||v||^2 with dz = K^T (v - K z)
__global__ void k_mp_dy_from_dz(const double* __restrict__ dz, const double* __restrict__ sbuf,
                                const int* __restrict__ ptr, float* __restrict__ dY, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool act = ptr[i + 1] > ptr[i];
  const double vv = sbuf[0], sc = sbuf[1];
  const double f = (vv > 0.0 && sc > 0.0) ? -2.0 / (vv * sc) : 0.0;
#pragma unroll
  for (int k = 0; k < 5; ++k) dY[i * 5 + k] = act ? (float)(f * dz[(size_t)k * n + i]) : 0.f;
}

// dh = dY Wout^T (given dY)
__global__ void k_mp_dh_from_dy(const float* __restrict__ dY, const int* __restrict__ ptr,
                                const float* __restrict__ Wout, float* __restrict__ dh, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool act = ptr[i + 1] > ptr[i];
  float d[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) d[f] = dY[i * 5 + f];
  for (int k = 0; k < kH; ++k) {
    float a = 0.f;
#pragma unroll
    for (int f = 0; f < 5; ++f) a += d[f] * Wout[k * 5 + f];
    dh[(size_t)i * kH + k] = act ? a : 0.f;
  }
}

// loss = ||r||^2 / ||v||^2 (r: sbuf[3] holds ||r||^2) accumulated into sbuf[4]
__global__ void k_mp_res_loss(double* sbuf) {
  const double l = sbuf[0] > 0.0 ? sbuf[3] / sbuf[0] : 0.0;
  sbuf[2] = l;
  sbuf[4] += l;
}

// dWout[k][f] = sum_i h[i][k] dY[i][f]; dbout[f] = sum_i dY[i][f]  (one block
// per k, k = 64 for the bias)
__global__ void k_mp_out_grad_w(const uint16_t* __restrict__ h, const float* __restrict__ dY,
                                long long n, float* __restrict__ out /* 65 x 5 */) {
  __shared__ float red[5][256];
  const int k = blockIdx.x;
  float s[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
  for (long long i = threadIdx.x; i < n; i += blockDim.x) {
    const float hv = k < kH ? bf2f(h[(size_t)i * kH + k]) : 1.0f;
#pragma unroll
    for (int f = 0; f < 5; ++f) s[f] += hv * dY[i * 5 + f];
  }
#pragma unroll
  for (int f = 0; f < 5; ++f) red[f][threadIdx.x] = s[f];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
#pragma unroll
      for (int f = 0; f < 5; ++f) red[f][threadIdx.x] += red[f][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x < 5) out[k * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// sum of a float array in fixed order (one block)
__global__ void k_mp_sum(const float* __restrict__ a, long long n, float* __restrict__ out) {
  __shared__ float red[256];
  float s = 0.f;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) s += a[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// dh (float) -> add the encoder: dU0 = dh * silu'(U0) uses k_mp_dact.

// Adam on float master weights; writes the bf16 copies used by the forward
// pass (transposed, Wt[c][k], and row-major W[k][c] for the backward dA).
__global__ void k_mp_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                          float* __restrict__ m2, int count, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

// W (K x 64 float, row-major W[k][c]) -> Wt[c][k] bf16 and Wr[c][k]... the
// backward dA kernel reads B[k=channel][col=feature] = W[feature][channel],
// i.e. the same W[k][c] rows with k the feature: Wrow = bf(W) row-major.
__global__ void k_mp_pack_w(const float* __restrict__ W, int K, uint16_t* __restrict__ Wt,
                            uint16_t* __restrict__ Wrow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * kH) return;
  const int k = i / kH, c = i % kH;
  const uint16_t b = f2bf(W[i]);
  Wt[(size_t)c * K + k] = b;
  if (Wrow) Wrow[i] = b;
}

// targets y* = (t - v) s on active rows (n x 5, float)
__global__ void k_mp_target(const double* __restrict__ t, const double* __restrict__ v,
                            const double* __restrict__ sptr, const int* __restrict__ ptr,
                            float* __restrict__ ystar, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const bool act = ptr[i + 1] > ptr[i];
  const double s = sptr[0];
#pragma unroll
  for (int f = 0; f < 5; ++f)
    ystar[i * 5 + f] = act ? (float)((t[(size_t)f * n + i] - v[(size_t)f * n + i]) * s) : 0.f;
}

// pair slot (35 floats per node, node-major planes: 25 BDiag, 5 b, 5 z)
__global__ void k_mp_store(const double* __restrict__ bd, const double* __restrict__ b,
                           const double* __restrict__ z, float* __restrict__ slot, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int e = 0; e < 25; ++e) slot[(size_t)e * n + i] = (float)bd[(size_t)e * n + i];
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    slot[(size_t)(25 + f) * n + i] = (float)b[(size_t)f * n + i];
    slot[(size_t)(30 + f) * n + i] = (float)z[(size_t)f * n + i];
  }
}

__global__ void k_mp_load(const float* __restrict__ slot, double* __restrict__ bd,
                          double* __restrict__ b, double* __restrict__ z, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int e = 0; e < 25; ++e) bd[(size_t)e * n + i] = (double)slot[(size_t)e * n + i];
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    b[(size_t)f * n + i] = (double)slot[(size_t)(25 + f) * n + i];
    z[(size_t)f * n + i] = (double)slot[(size_t)(30 + f) * n + i];
  }
}

// s = sqrt(n_act / ||v||^2) (0 for a zero vector)
__global__ void k_mp_scale_from_sq(const double* sq, double n_act, double* s) {
  s[0] = sq[0] > 0.0 ? sqrt(n_act / sq[0]) : 0.0;
}

// loss = sum(sq) / n_act, in double from the float partial sums
__global__ void k_mp_loss(const float* part, double inv_na, double* out) { out[0] = part[0] * inv_na; }

inline int grid_for(long long n, int threads) {
  long long b = (n + threads - 1) / threads;
  return (int)(b < 1 ? 1 : b);
}

}  // namespace

// ---------------- host side ----------------

#include <algorithm>
#include <cmath>
#include <numeric>

namespace {

// offsets (floats) of the parameter block: W0 b0 W1 b1 W2 b2 W3 b3 Wout bout
struct Layout {
  static constexpr int W0 = 0, b0 = W0 + kNin * kH;
  static constexpr int W1 = b0 + kH, b1 = W1 + kNagg * kH;
  static constexpr int W2 = b1 + kH, b2 = W2 + kNagg * kH;
  static constexpr int W3 = b2 + kH, b3 = W3 + kNagg * kH;
  static constexpr int Wo = b3 + kH, bo = Wo + kH * 5;
  static constexpr int total = bo + 5;
};
constexpr int kWoff[4] = {Layout::W0, Layout::W1, Layout::W2, Layout::W3};
constexpr int kBoff[4] = {Layout::b0, Layout::b1, Layout::b2, Layout::b3};
constexpr int kK[4] = {kNin, kNagg, kNagg, kNagg};
constexpr int kPartBlocks = 256;

template <typename T>
int mp_alloc(phg_ctx* c, T** p, size_t count) {
  PHG_HIP(hipMalloc((void**)p, sizeof(T) * std::max<size_t>(count, 1)));
  c->dev_bytes += (long long)(sizeof(T) * count);
  return 0;
}

}  // namespace

int mp_param_count() { return Layout::total; }

// Periodic node graph of the element adjacency: edge (I, J) = (iper(i),
// iper(j)) for every CSR entry (i, j) with I != J, first occurrence kept
// (entries in row order, stable), vector x_j - x_i scaled by the RMS edge
// length of row I; rev[k] is the index of the reverse edge.
int mp_init(phg_ctx* c, const phg_mesh* m, const float* dscale) {
  MpNet& M = c->mp;
  if (M.ready) return 0;
  const long long n = c->nshg;
  struct E {
    int I, J;
    double v[3];
  };
  std::vector<E> e;
  e.reserve((size_t)m->nnz_tot);
  for (long long i = 0; i < n; ++i) {
    const int I = m->iper[i] - 1;
    for (long long k = m->colm[i] - 1; k < m->colm[i + 1] - 1; ++k) {
      const long long j = m->rowp[k] - 1;
      const int J = m->iper[j] - 1;
      if (I == J) continue;
      E t{I, J, {m->x[j] - m->x[i], m->x[j + n] - m->x[i + n], m->x[j + 2 * n] - m->x[i + 2 * n]}};
      e.push_back(t);
    }
  }
  std::stable_sort(e.begin(), e.end(),
                   [](const E& a, const E& b) { return a.I != b.I ? a.I < b.I : a.J < b.J; });
  std::vector<E> u;
  u.reserve(e.size());
  for (size_t k = 0; k < e.size(); ++k)
    if (k == 0 || e[k].I != e[k - 1].I || e[k].J != e[k - 1].J) u.push_back(e[k]);
  e.clear();
  e.shrink_to_fit();
  const long long ne = (long long)u.size();
  if (ne >= (1LL << 31)) {
    phg_set_error("mp_init: %lld graph edges exceed int32", ne);
    return PHG_EINVAL;
  }
  std::vector<int> ptr(n + 1, 0), col(ne), rev(ne);
  for (const E& t : u) ptr[t.I + 1]++;
  for (long long i = 0; i < n; ++i) ptr[i + 1] += ptr[i];
  std::vector<double> l2(n, 0.0);
  for (const E& t : u) l2[t.I] += t.v[0] * t.v[0] + t.v[1] * t.v[1] + t.v[2] * t.v[2];
  std::vector<float> inv(n, 0.f);
  std::vector<double> lr(n, 1.0);
  long long nact = 0;
  for (long long i = 0; i < n; ++i) {
    const int d = ptr[i + 1] - ptr[i];
    if (d > 0) {
      lr[i] = std::sqrt(l2[i] / d);
      inv[i] = 1.0f / (float)d;
      ++nact;
    }
  }
  std::vector<float4> ev(ne);
  std::vector<float4> es(n, float4{0.f, 0.f, 0.f, 0.f});
  for (long long k = 0; k < ne; ++k) {
    const E& t = u[k];
    col[k] = t.J;
    const double s = 1.0 / lr[t.I];
    ev[k] = float4{(float)(t.v[0] * s), (float)(t.v[1] * s), (float)(t.v[2] * s), 0.f};
  }
  for (long long i = 0; i < n; ++i) {
    double sx = 0, sy = 0, sz = 0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
      sx += ev[k].x;
      sy += ev[k].y;
      sz += ev[k].z;
    }
    es[i] = float4{(float)sx, (float)sy, (float)sz, 0.f};
  }
  for (long long i = 0; i < n; ++i)
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int j = col[k];
      const int* b = col.data() + ptr[j];
      const int* e_end = col.data() + ptr[j + 1];
      const int* f = std::lower_bound(b, e_end, (int)i);
      if (f == e_end || *f != i) {
        phg_set_error("mp_init: graph is not symmetric at (%lld,%d)", i, j);
        return PHG_EINVAL;
      }
      rev[k] = (int)(f - col.data());
    }
  M.n_act = nact;
  M.nedge = ne;
  PHG_TRY(mp_alloc(c, &M.ptr, n + 1));
  PHG_TRY(mp_alloc(c, &M.col, ne));
  PHG_TRY(mp_alloc(c, &M.rev, ne));
  PHG_TRY(mp_alloc(c, &M.ev, ne));
  PHG_TRY(mp_alloc(c, &M.esum, n));
  PHG_TRY(mp_alloc(c, &M.inv, n));
  PHG_TRY(mp_alloc(c, &M.dscale, 25));
  PHG_TRY(mp_alloc(c, &M.bd_raw, 25 * n));
  PHG_TRY(mp_alloc(c, &M.zhat, 5 * n));
  PHG_TRY(mp_alloc(c, &M.x, (size_t)kNin * n));
  PHG_TRY(mp_alloc(c, &M.h[0], (size_t)kH * n));
  PHG_TRY(mp_alloc(c, &M.h[1], (size_t)kH * n));
  PHG_TRY(mp_alloc(c, &M.w, Layout::total));
  for (int l = 0; l < 4; ++l) {
    PHG_TRY(mp_alloc(c, &M.wt[l], (size_t)kK[l] * kH));
    PHG_TRY(mp_alloc(c, &M.wr[l], (size_t)kK[l] * kH));
  }
  PHG_TRY(mp_alloc(c, &M.sbuf, 12));
  PHG_TRY(mp_alloc(c, &M.vin, 5 * n));
  PHG_TRY(mp_alloc(c, &M.zout, 5 * n));
  PHG_TRY(mp_alloc(c, &M.rbuf, 5 * n));
  PHG_TRY(mp_alloc(c, &M.gbuf, 5 * n));
  PHG_TRY(mp_alloc(c, &M.zbasis, (size_t)std::max(c->kcap, 1) * 5 * n));
  PHG_TRY(mp_alloc(c, &M.wplain, 5 * n));
  PHG_HIP(hipMemcpyAsync(M.ptr, ptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, c->stream));
  PHG_HIP(hipMemcpyAsync(M.col, col.data(), sizeof(int) * ne, hipMemcpyHostToDevice, c->stream));
  PHG_HIP(hipMemcpyAsync(M.rev, rev.data(), sizeof(int) * ne, hipMemcpyHostToDevice, c->stream));
  PHG_HIP(hipMemcpyAsync(M.ev, ev.data(), sizeof(float4) * ne, hipMemcpyHostToDevice, c->stream));
  PHG_HIP(hipMemcpyAsync(M.esum, es.data(), sizeof(float4) * n, hipMemcpyHostToDevice, c->stream));
  PHG_HIP(hipMemcpyAsync(M.inv, inv.data(), sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
  PHG_HIP(hipMemcpyAsync(M.dscale, dscale, sizeof(float) * 25, hipMemcpyHostToDevice, c->stream));
  PHG_HIP(hipStreamSynchronize(c->stream));
  PHG_TRY(launch_fill(c, M.w, sizeof(float) * Layout::total));
  PHG_TRY(mp_pack(c));
  M.bd_src = M.bd_raw;
  M.ready = true;
  return 0;
}

int mp_alloc_train(phg_ctx* c) {
  MpNet& M = c->mp;
  if (M.train_ready) return 0;
  const long long n = c->nshg;
  for (int l = 2; l < 4; ++l) PHG_TRY(mp_alloc(c, &M.h[l], (size_t)kH * n));
  for (int l = 0; l < 3; ++l) PHG_TRY(mp_alloc(c, &M.A[l], (size_t)kNagg * n));
  for (int l = 0; l < 4; ++l) PHG_TRY(mp_alloc(c, &M.U[l], (size_t)kH * n));
  PHG_TRY(mp_alloc(c, &M.Y, 5 * n));
  PHG_TRY(mp_alloc(c, &M.Ystar, 5 * n));
  PHG_TRY(mp_alloc(c, &M.dh, (size_t)kH * n));
  PHG_TRY(mp_alloc(c, &M.dU, (size_t)kH * n));
  PHG_TRY(mp_alloc(c, &M.dA, (size_t)kNagg * n));
  PHG_TRY(mp_alloc(c, &M.dY, 5 * n));
  PHG_TRY(mp_alloc(c, &M.sq, n));
  M.nparts = (int)std::min<long long>(kPartBlocks, (n + 31) / 32);
  PHG_TRY(mp_alloc(c, &M.part, (size_t)M.nparts * kNagg * kH));
  if (!M.param_owner) {
    PHG_TRY(mp_alloc(c, &M.g, Layout::total));
    PHG_TRY(mp_alloc(c, &M.m1, Layout::total));
    PHG_TRY(mp_alloc(c, &M.m2, Layout::total));
    PHG_TRY(launch_fill(c, M.m1, sizeof(float) * Layout::total));
    PHG_TRY(launch_fill(c, M.m2, sizeof(float) * Layout::total));
    M.adam_t = 0;
  }
  M.train_ready = true;
  return 0;
}

int mp_free(phg_ctx* c) {
  MpNet& M = c->mp;
  void* ptrs[] = {M.ptr,  M.col,  M.rev,  M.ev,   M.esum,  M.inv,   M.dscale, M.bd_raw, M.zhat,
                  M.x,    M.h[0], M.h[1], M.h[2], M.h[3],  M.A[0],  M.A[1],   M.A[2],   M.U[0],
                  M.U[1], M.U[2], M.U[3], M.Y,    M.Ystar, M.dh,    M.dU,     M.dA,     M.dY,
                  M.sq,   M.part, M.sbuf, M.vin,  M.zout,  M.zbasis, M.wplain, M.ds,
                  M.rbuf, M.gbuf};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (!M.param_owner) {
    void* own[] = {M.w,     M.g,     M.m1,    M.m2,    M.wt[0], M.wt[1],
                   M.wt[2], M.wt[3], M.wr[0], M.wr[1], M.wr[2], M.wr[3]};
    for (void* p : own)
      if (p) (void)hipFree(p);
  }
  M = MpNet{};
  return 0;
}

// dst trains and runs with src's parameters (and Adam state); both contexts
// live on the same device.  dst's own parameter buffers are released.
int mp_share(phg_ctx* dst, phg_ctx* src) {
  MpNet& D = dst->mp;
  MpNet& S = src->mp;
  if (!D.ready || !S.ready || S.param_owner || dst == src) {
    phg_set_error("phg_mpnn_share: both contexts need phg_mpnn_init; src must own its parameters");
    return PHG_EINVAL;
  }
  PHG_TRY(mp_alloc_train(src));
  PHG_TRY(mp_alloc_train(dst));
  PHG_HIP(hipDeviceSynchronize());
  if (!D.param_owner) {
    void* own[] = {D.w,     D.g,     D.m1,    D.m2,    D.wt[0], D.wt[1],
                   D.wt[2], D.wt[3], D.wr[0], D.wr[1], D.wr[2], D.wr[3]};
    for (void* p : own)
      if (p) (void)hipFree(p);
  }
  D.w = S.w;
  D.g = S.g;
  D.m1 = S.m1;
  D.m2 = S.m2;
  for (int l = 0; l < 4; ++l) {
    D.wt[l] = S.wt[l];
    D.wr[l] = S.wr[l];
  }
  D.param_owner = src;
  return 0;
}

int mp_dataset(phg_ctx* c, int capacity) {
  MpNet& M = c->mp;
  if (M.ds) {
    (void)hipFree(M.ds);
    M.ds = nullptr;
    M.ds_cap = 0;
  }
  if (capacity <= 0) return 0;
  PHG_TRY(mp_alloc(c, &M.ds, (size_t)capacity * 35 * c->nshg));
  M.ds_cap = capacity;
  return 0;
}

int mp_store_pair(phg_ctx* c, int slot) {
  MpNet& M = c->mp;
  if (slot < 0 || slot >= M.ds_cap) {
    phg_set_error("phg_mpnn_store_pair: slot %d outside 0..%d", slot, M.ds_cap - 1);
    return PHG_EINVAL;
  }
  const long long n = c->nshg;
  k_mp_store<<<grid_for(n, 256), 256, 0, c->stream>>>(M.bd_raw, c->res, M.zhat,
                                                      M.ds + (size_t)slot * 35 * n, n);
  PHG_HIP(hipGetLastError());
  return 0;
}

int mp_load_pair(phg_ctx* c, int slot) {
  MpNet& M = c->mp;
  if (slot < 0 || slot >= M.ds_cap) {
    phg_set_error("phg_mpnn_train_pair: slot %d outside 0..%d", slot, M.ds_cap - 1);
    return PHG_EINVAL;
  }
  const long long n = c->nshg;
  k_mp_load<<<grid_for(n, 256), 256, 0, c->stream>>>(M.ds + (size_t)slot * 35 * n, M.bd_raw,
                                                     M.vin, M.zout, n);
  PHG_HIP(hipGetLastError());
  return 0;
}

int mp_pack(phg_ctx* c) {
  MpNet& M = c->mp;
  for (int l = 0; l < 4; ++l) {
    k_mp_pack_w<<<grid_for(kK[l] * kH, 256), 256, 0, c->stream>>>(M.w + kWoff[l], kK[l], M.wt[l],
                                                                  M.wr[l]);
    PHG_HIP(hipGetLastError());
  }
  return 0;
}

int mp_forward(phg_ctx* c, const double* v, double* z, bool train) {
  MpNet& M = c->mp;
  if (!M.ready || (train && !M.train_ready)) {
    phg_set_error("learned preconditioner not initialised");
    return PHG_EINVAL;
  }
  const long long n = c->nshg;
  PHG_TRY(launch_dot(c, v, v, 5 * n, M.sbuf));
  k_mp_scale_from_sq<<<1, 1, 0, c->stream>>>(M.sbuf, (double)M.n_act, M.sbuf + 1);
  PHG_HIP(hipGetLastError());
  k_mp_features<<<grid_for(n, 256), 256, 0, c->stream>>>(M.bd_src, v, M.dscale, M.ptr, M.sbuf + 1,
                                                         M.x, n);
  PHG_HIP(hipGetLastError());
  const int gd = grid_for(n, 64);
  k_mp_dense<kNin><<<gd, 256, 0, c->stream>>>(M.x, M.wt[0], M.w + kBoff[0], nullptr, M.ptr,
                                             M.h[0], train ? M.U[0] : nullptr, 0, n);
  PHG_HIP(hipGetLastError());
  for (int l = 1; l <= 3; ++l) {
    // training keeps h0..h3 and A1..A3; inference ping-pongs h[0], h[1]
    uint16_t* hin = M.h[train ? l - 1 : (l - 1) & 1];
    uint16_t* hout = M.h[train ? l : l & 1];
    if (!train) {
      const int gl = (grid_for(n, kLayerRows) + 7) / 8 * 8;
      k_mp_layer<<<gl, 256, 0, c->stream>>>(hin, M.ptr, M.col, M.ev, M.inv, M.wt[l],
                                            M.w + kBoff[l], hout, n);
      PHG_HIP(hipGetLastError());
      continue;
    }
    uint16_t* A = M.A[l - 1];
    k_mp_agg<<<grid_for(n, 16), 256, 0, c->stream>>>(hin, M.ptr, M.col, M.ev, M.inv, A, n);
    PHG_HIP(hipGetLastError());
    k_mp_dense<kNagg><<<gd, 256, 0, c->stream>>>(A, M.wt[l], M.w + kBoff[l], hin, M.ptr, hout,
                                                M.U[l], 1, n);
    PHG_HIP(hipGetLastError());
  }
  const uint16_t* h3 = M.h[train ? 3 : 1];
  k_mp_out<<<grid_for(n, 256), 256, 0, c->stream>>>(h3, M.w + Layout::Wo, M.w + Layout::bo, M.ptr,
                                                    v, M.sbuf + 1, z, train ? M.Y : nullptr, n);
  PHG_HIP(hipGetLastError());
  return 0;
}

int mp_backward(phg_ctx* c, float lr, double* loss_out, bool given_dy) {
  MpNet& M = c->mp;
  if (!M.train_ready) {
    phg_set_error("learned preconditioner: training buffers not allocated");
    return PHG_EINVAL;
  }
  const long long n = c->nshg;
  if (given_dy) {
    k_mp_dh_from_dy<<<grid_for(n, 256), 256, 0, c->stream>>>(M.dY, M.ptr, M.w + Layout::Wo, M.dh,
                                                             n);
    PHG_HIP(hipGetLastError());
  } else {
    const float inv_na = 1.0f / (float)std::max<long long>(1, M.n_act);
    k_mp_out_bwd<<<grid_for(n, 256), 256, 0, c->stream>>>(M.Y, M.Ystar, M.ptr, M.w + Layout::Wo,
                                                          inv_na, M.dY, M.dh, M.sq, n);
    PHG_HIP(hipGetLastError());
    // loss: fixed-order float sum of the per-node squared errors
    k_mp_sum<<<1, 256, 0, c->stream>>>(M.sq, n, M.part);
    k_mp_loss<<<1, 1, 0, c->stream>>>(M.part, 1.0 / (double)std::max<long long>(1, M.n_act),
                                      M.sbuf + 2);
    PHG_HIP(hipGetLastError());
  }
  k_mp_out_grad_w<<<kH + 1, 256, 0, c->stream>>>(M.h[3], M.dY, n, M.g + Layout::Wo);
  PHG_HIP(hipGetLastError());
  const int gd = grid_for(n, 64);
  const long long nH = n * kH;
  for (int l = 3; l >= 0; --l) {
    k_mp_dact<<<grid_for(nH, 256), 256, 0, c->stream>>>(M.dh, M.U[l], M.ptr, M.dU, n);
    PHG_HIP(hipGetLastError());
    const int K = kK[l];
    if (l > 0)
      k_mp_dW<kNagg><<<M.nparts, 256, 0, c->stream>>>(M.A[l - 1], M.dU, M.part, n);
    else
      k_mp_dW<kNin><<<M.nparts, 256, 0, c->stream>>>(M.x, M.dU, M.part, n);
    PHG_HIP(hipGetLastError());
    k_mp_sum_parts<<<grid_for(K * kH, 256), 256, 0, c->stream>>>(M.part, M.nparts, K * kH,
                                                                 M.g + kWoff[l]);
    k_mp_colsum<<<kH, 256, 0, c->stream>>>(M.dU, n, M.g + kBoff[l]);
    PHG_HIP(hipGetLastError());
    if (l == 0) break;
    k_mp_dA<<<gd, 256, 0, c->stream>>>(M.dU, M.wr[l], M.dA, n);
    k_mp_agg_bwd<<<grid_for(n, 16), 256, 0, c->stream>>>(M.dA, M.ptr, M.col, M.rev, M.ev, M.esum,
                                                        M.inv, M.dh, n);
    PHG_HIP(hipGetLastError());
  }
  if (lr > 0.f) {
    int& step = M.param_owner ? M.param_owner->mp.adam_t : M.adam_t;
    step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f / (1.f - std::pow(b1, (float)step));
    const float c2 = 1.f / (1.f - std::pow(b2, (float)step));
    k_mp_adam<<<grid_for(Layout::total, 256), 256, 0, c->stream>>>(M.w, M.g, M.m1, M.m2,
                                                                    Layout::total, lr, b1, b2, eps,
                                                                    c1, c2);
    PHG_HIP(hipGetLastError());
    PHG_TRY(mp_pack(c));
  }
  if (loss_out) {
    PHG_HIP(hipMemcpyAsync(loss_out, M.sbuf + 2, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PHG_HIP(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int mp_targets(phg_ctx* c, const double* t, const double* v) {
  MpNet& M = c->mp;
  const long long n = c->nshg;
  k_mp_target<<<grid_for(n, 256), 256, 0, c->stream>>>(t, v, M.sbuf + 1, M.ptr, M.Ystar, n);
  PHG_HIP(hipGetLastError());
  return 0;
}

// One training step on the residual of the preconditioned direction:
// z = v + y(v)/s, loss = ||K z - v||^2 / ||v||^2 with K the scaled matrix
// of the last solve, gradient dz = 2 K^T (K z - v) / ||v||^2.
int mp_train_residual(phg_ctx* c, const double* v, float lr) {
  MpNet& M = c->mp;
  const long long n = c->nshg, V = 5 * n;
  PHG_TRY(mp_forward(c, v, M.zout, true));              // sbuf[0] = ||v||^2, sbuf[1] = s
  PHG_TRY(launch_spmv(c, M.zout, M.rbuf));
  PHG_TRY(launch_bc3per(c, M.rbuf, 5));
  PHG_TRY(launch_sub(c, M.gbuf, v, M.rbuf, V));         // gbuf = v - K z
  PHG_TRY(launch_dot(c, M.gbuf, M.gbuf, V, M.sbuf + 3));
  k_mp_res_loss<<<1, 1, 0, c->stream>>>(M.sbuf);
  PHG_HIP(hipGetLastError());
  PHG_TRY(launch_spmv_t(c, M.gbuf, M.rbuf));            // rbuf = K^T (v - K z)
  k_mp_dy_from_dz<<<grid_for(n, 256), 256, 0, c->stream>>>(M.rbuf, M.sbuf, M.ptr, M.dY, n);
  PHG_HIP(hipGetLastError());
  return mp_backward(c, lr, nullptr, true);
}
