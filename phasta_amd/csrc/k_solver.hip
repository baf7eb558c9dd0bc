// Linear-solver kernels on gfx950 for the block-diagonal-preconditioned
// GMRES of SolGMRs (solgmr.f:368-744): 5x5 block LU of BDiag (i3lu.f),
// two-sided block scaling of the matrix (spsi3pre.f), the block-CSR matvec
// with the periodic gather (sparseap.f), the periodic fold (bc3per.f),
// Krylov-vector kernels with deterministic wavefront/grid reductions
// (sumgat, mpitools.f:104-138) and the itrCorrect/itrBC state update
// (itrPC.f:128-148, itrbc.f).
//
// Matrix layout ("slots"): rows are grouped by 64, one wavefront per group.
// Group g owns slots [grp_slot[g], grp_slot[g+1]); slot s holds one 5x5 block
// for every row of the group, entry e (reference order f + 5 g) of row r at
// vals[(s*25 + e)*64 + r].  Each matrix load of a wavefront is therefore one
// contiguous 512-byte (FP64) or 256-byte (FP32) segment.  Rows shorter than
// the group's slot count are padded with zero blocks.
#include "phg_internal.h"

namespace {

constexpr int kRedThreads = 256;

// ---------------- block LU (i3lu.f) ----------------
// 25 planes, D[(i + 5 j) * n + node] = Diag(node, i+1, j+1)
__global__ void k_lu_fact(double* D, long long n) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  double a[5][5];
#pragma unroll
  for (int j = 0; j < 5; ++j)
#pragma unroll
    for (int i = 0; i < 5; ++i) a[i][j] = D[(i + 5 * j) * n + k];
  // Doolittle: unit-lower L, upper U with its diagonal stored inverted.
  // As in the reference (i3lu.f:82-97) the entry U(4,5) is never reduced:
  // row 4 of U only gets its diagonal, so (4,5) keeps the raw BDiag value
  // and the forward/backward sweeps use it as such.
#pragma unroll
  for (int p = 0; p < 5; ++p) {
#pragma unroll
    for (int j = p; j < 5; ++j) {
      if (p == 3 && j == 4) continue;
      double s = a[p][j];
#pragma unroll
      for (int m = 0; m < p; ++m) s -= a[p][m] * a[m][j];
      a[p][j] = s;
    }
    a[p][p] = 1.0 / a[p][p];
#pragma unroll
    for (int i = p + 1; i < 5; ++i) {
      double s = a[i][p];
#pragma unroll
      for (int m = 0; m < p; ++m) s -= a[i][m] * a[m][p];
      a[i][p] = a[p][p] * s;
    }
  }
#pragma unroll
  for (int j = 0; j < 5; ++j)
#pragma unroll
    for (int i = 0; i < 5; ++i) D[(i + 5 * j) * n + k] = a[i][j];
}

__global__ void k_lu_forward(const double* D, double* r, long long n) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) v[i] = r[i * n + k];
#pragma unroll
  for (int i = 1; i < 5; ++i) {
    double s = v[i];
#pragma unroll
    for (int m = 0; m < i; ++m) s -= D[(i + 5 * m) * n + k] * v[m];
    v[i] = s;
  }
#pragma unroll
  for (int i = 1; i < 5; ++i) r[i * n + k] = v[i];
}

__global__ void k_lu_backward(const double* D, double* r, long long n) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) v[i] = r[i * n + k];
#pragma unroll
  for (int i = 4; i >= 0; --i) {
    double s = v[i];
#pragma unroll
    for (int m = 4; m > i; --m) s -= v[m] * D[(i + 5 * m) * n + k];
    v[i] = D[(i + 5 * i) * n + k] * s;
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) r[i * n + k] = v[i];
}

// i3LU 'product' (i3lu.f:150-167): r <- U r with U's diagonal stored
// inverted (so r_i / D_ii), rows in ascending order (each row reads the
// entries below it before they change)
__global__ void k_lu_product(const double* D, double* r, long long n) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) v[i] = r[i * n + k];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    double s = v[i] / D[(i + 5 * i) * n + k];
#pragma unroll
    for (int m = i + 1; m < 5; ++m) s += v[m] * D[(i + 5 * m) * n + k];
    v[i] = s;
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) r[i * n + k] = v[i];
}

// yshuffle (shuffle.f:1-30) from src into dst: new2old (u,v,w,p,T) ->
// (p,u,v,w,T), else back
__global__ void k_yshuffle(const double* src, double* dst, long long n, int new2old) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) v[i] = src[i * n + k];
  if (new2old) {
    dst[k] = v[3];
    dst[n + k] = v[0];
    dst[2 * n + k] = v[1];
    dst[3 * n + k] = v[2];
  } else {
    dst[k] = v[1];
    dst[n + k] = v[2];
    dst[2 * n + k] = v[3];
    dst[3 * n + k] = v[0];
  }
  dst[4 * n + k] = v[4];
}

// ---------------- two-sided block scaling (spsi3pre.f) ----------------
// lhsK(i,j) <- L_i^{-1} lhsK(i,j) U_j^{-1} with the factors of BDiag.  Reads
// the assembled blocks (reference order, one padded record per block, 16-byte
// loads) and writes the 64-row interleaved matvec layout.
// One workgroup of kPreWaves wavefronts per 64-row group; wavefront w takes
// the group's slots w, w + kPreWaves, ... (more independent loads in flight
// than one wavefront walking all slots of its group).
constexpr int kPreWaves = 4;

template <typename TI, typename TO>
__global__ void __launch_bounds__(kPreWaves * kGroup) k_spsi3pre(
    const TI* __restrict__ aos, TO* __restrict__ vals, const int* __restrict__ s2c,
    const int* __restrict__ scol, const long long* __restrict__ grp_slot,
    const double* __restrict__ D, long long n) {
  const long long g = blockIdx.x;
  const int w = threadIdx.x / kGroup, r = threadIdx.x % kGroup;
  const long long row = g * kGroup + r;
  if (row >= n) return;
  double lo[5][5];
#pragma unroll
  for (int i = 1; i < 5; ++i)
#pragma unroll
    for (int m = 0; m < i; ++m) lo[i][m] = D[(i + 5 * m) * n + row];
  for (long long s = grp_slot[g] + w; s < grp_slot[g + 1]; s += kPreWaves) {
    const long long k = s2c[s * kGroup + r];
    if (k < 0) continue;  // padding stays zero
    const long long j = scol[s * kGroup + r];
    double a[28];
    if constexpr (sizeof(TI) == 4) {
      const float4* v = reinterpret_cast<const float4*>(aos + (size_t)aos_stride<TI>() * k);
#pragma unroll
      for (int t = 0; t < 7; ++t) {
        const float4 x = v[t];
        a[4 * t] = x.x;
        a[4 * t + 1] = x.y;
        a[4 * t + 2] = x.z;
        a[4 * t + 3] = x.w;
      }
    } else {
      const double2* v = reinterpret_cast<const double2*>(aos + (size_t)aos_stride<TI>() * k);
#pragma unroll
      for (int t = 0; t < 13; ++t) {
        const double2 x = v[t];
        a[2 * t] = x.x;
        a[2 * t + 1] = x.y;
      }
    }
    double b[5][5];
#pragma unroll
    for (int c = 0; c < 5; ++c)
#pragma unroll
      for (int f = 0; f < 5; ++f) b[f][c] = a[f + 5 * c];
    // left: L_i^{-1} on every column
#pragma unroll
    for (int c = 0; c < 5; ++c)
#pragma unroll
      for (int f = 1; f < 5; ++f) {
        double acc = b[f][c];
#pragma unroll
        for (int m = 0; m < f; ++m) acc -= lo[f][m] * b[m][c];
        b[f][c] = acc;
      }
    // right: U_j^{-1} on every row (columns in order)
    double up[5][5];
#pragma unroll
    for (int c = 0; c < 5; ++c)
#pragma unroll
      for (int m = 0; m <= c; ++m) up[m][c] = D[(m + 5 * c) * n + j];
#pragma unroll
    for (int c = 0; c < 5; ++c)
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        double acc = b[f][c];
#pragma unroll
        for (int m = 0; m < c; ++m) acc -= up[m][c] * b[f][m];
        b[f][c] = up[c][c] * acc;
      }
#pragma unroll
    for (int c = 0; c < 5; ++c)
#pragma unroll
      for (int f = 0; f < 5; ++f) a[f + 5 * c] = b[f][c];
    store_slot(vals + s * (25 * kGroup), r, a);
  }
}

// ---------------- block-CSR matvec (sparseap.f) ----------------
// q = K p(iper(:)); one lane per row of a 64-row group in each wavefront.
// T: stored value type; Acc: type of the products and row sums.  The mixed
// path stores float values and sums rows in double: float row sums of the
// scaled system lose digits to cancellation (on the 16^3 box the refinement
// loop then needed 12 instead of 10 Krylov iterations), while the kernel is
// bound by the matrix bytes, which float storage halves either way.
// One workgroup of kSpmvWaves wavefronts per 64-row group: wavefront w takes
// the group's slots w, w + kSpmvWaves, ..., so the wavefronts of a workgroup
// stream adjacent slots together and each runs a short loop; the partial row
// sums are added in wavefront order through LDS (fixed order, deterministic).
// The grid is a multiple of 8: consecutive workgroups go to different XCDs
// (round robin), so workgroup b takes group (b % 8) * per + b / 8 and each
// XCD sweeps one contiguous eighth of the rows, whose column neighbours
// (rows +-1, +-(N+1), +-(N+1)^2 on the box) then stay in that XCD's L2.
constexpr int kSpmvWaves = 4;

template <typename T, typename Acc>
__global__ void __launch_bounds__(kSpmvWaves * kGroup) k_spmv(
    const T* __restrict__ vals, const int* __restrict__ scolp,
    const long long* __restrict__ grp_slot, const double* __restrict__ p, double* __restrict__ q,
    long long n, long long ngroups) {
  __shared__ Acc part[kSpmvWaves][5][kGroup];
  const long long per = (gridDim.x + 7) / 8;
  const long long g = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (g >= ngroups) return;  // whole workgroup
  const int w = threadIdx.x / kGroup, r = threadIdx.x % kGroup;
  const long long row = g * kGroup + r;
  const long long s0 = grp_slot[g] + w, s1 = grp_slot[g + 1];
  Acc t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0;
  // software pipeline: the gathers of p for the next slot and the column
  // index of the one after are issued before the block of this slot is
  // consumed (rows past n in the last group read padding: column 0, zero
  // blocks)
  constexpr int W = kSpmvWaves;
  Acc pc[5] = {0, 0, 0, 0, 0};
  if (s0 < s1) {
    const long long j0 = scolp[s0 * kGroup + r];
#pragma unroll
    for (int f = 0; f < 5; ++f) pc[f] = (Acc)p[j0 + f * n];
  }
  long long jn = s0 + W < s1 ? scolp[(s0 + W) * kGroup + r] : 0;
  for (long long s = s0; s < s1; s += W) {
    Acc pn[5] = {0, 0, 0, 0, 0};
    if (s + W < s1) {
#pragma unroll
      for (int f = 0; f < 5; ++f) pn[f] = (Acc)p[jn + f * n];
      if (s + 2 * W < s1) jn = scolp[(s + 2 * W) * kGroup + r];
    }
    Acc a[25];
    load_slot(vals + s * (25 * kGroup), r, a);
    t0 += a[0] * pc[0] + a[5] * pc[1] + a[10] * pc[2] + a[15] * pc[3] + a[20] * pc[4];
    t1 += a[1] * pc[0] + a[6] * pc[1] + a[11] * pc[2] + a[16] * pc[3] + a[21] * pc[4];
    t2 += a[2] * pc[0] + a[7] * pc[1] + a[12] * pc[2] + a[17] * pc[3] + a[22] * pc[4];
    t3 += a[3] * pc[0] + a[8] * pc[1] + a[13] * pc[2] + a[18] * pc[3] + a[23] * pc[4];
    t4 += a[4] * pc[0] + a[9] * pc[1] + a[14] * pc[2] + a[19] * pc[3] + a[24] * pc[4];
#pragma unroll
    for (int f = 0; f < 5; ++f) pc[f] = pn[f];
  }
  part[w][0][r] = t0;
  part[w][1][r] = t1;
  part[w][2][r] = t2;
  part[w][3][r] = t3;
  part[w][4][r] = t4;
  __syncthreads();
  if (w != 0 || row >= n) return;
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    Acc v = part[0][f][r];
#pragma unroll
    for (int k = 1; k < W; ++k) v += part[k][f][r];
    q[row + f * n] = (double)v;
  }
}

// Transpose matvec without the periodic maps: q_j = sum_i B_ij^T x_i over
// the blocks of column j, B_ij read at the slot tslot of (i, j); scol gives
// the raw columns i of row j (the graph is structurally symmetric).
template <typename T>
__global__ void __launch_bounds__(256) k_spmv_t(const T* __restrict__ vals, const int* __restrict__ scol,
                                                const uint32_t* __restrict__ tslot,
                                                const long long* __restrict__ grp_slot,
                                                const double* __restrict__ x, double* __restrict__ q,
                                                long long n) {
  const long long row = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long g = row / kGroup;
  const int r = (int)(row % kGroup);
  if (row >= n) return;
  double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (long long s = grp_slot[g]; s < grp_slot[g + 1]; ++s) {
    const long long i = scol[s * kGroup + r];
    const uint32_t ts = tslot[s * kGroup + r];
    const T* blk = vals + (size_t)(ts >> 6) * (25 * kGroup);
    const int tr = (int)(ts & 63);
    double xi[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) xi[f] = x[i + f * n];
#pragma unroll
    for (int gg = 0; gg < 5; ++gg) {
      double a = 0.0;
#pragma unroll
      for (int f = 0; f < 5; ++f) a += (double)blk[slot_pos<T>(f + 5 * gg, tr)] * xi[f];
      t[gg] += a;
    }
  }
#pragma unroll
  for (int f = 0; f < 5; ++f) q[row + f * n] = t[f];
}

// slaves take their master's value (the transpose of the periodic fold)
__global__ void k_per_copy_vec(double* v, const int* master, const int* ptr, const int* slave,
                               int nfold, long long n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfold) return;
  const long long m = master[i];
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const long long s = slave[k];
#pragma unroll
    for (int f = 0; f < 5; ++f) v[s + f * n] = v[m + f * n];
  }
}

// r = b - q (true residual of the refinement loop)
__global__ void k_sub(double* r, const double* b, const double* q, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    r[i] = b[i] - q[i];
}

// plain copy, 16 B per lane (calibration of the memory counters)
__global__ void k_copy(double2* __restrict__ dst, const double2* __restrict__ src, long long n2) {
  // one 16-byte element per lane, no loop: the fastest plain copy measured
  // on MI355X (6.2 TB/s for 859 MB; grid-stride loops reached 4.6-5.2)
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n2) dst[i] = src[i];
}

// padded assembled blocks -> lhsK(25,nnz) (double) and back
template <typename T>
__global__ void k_aos_to_ref(const T* __restrict__ aos, double* __restrict__ ref, long long nnz) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= 25 * nnz) return;
  const long long k = i / 25;
  ref[i] = (double)aos[(size_t)aos_stride<T>() * k + (i - 25 * k)];
}
__global__ void k_ref_to_aos(const double* __restrict__ ref, double* __restrict__ aos,
                             long long nnz) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= (long long)aos_stride<double>() * nnz) return;
  const long long k = i / aos_stride<double>(), e = i - (long long)aos_stride<double>() * k;
  aos[i] = e < 25 ? ref[25 * k + e] : 0.0;
}


// bc3per: master += slaves (ascending), slaves zeroed
__global__ void k_fold_vec(double* v, const int* master, const int* ptr, const int* slave,
                           int nfold, long long n, int nq) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfold) return;
  const long long m = master[i];
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const long long s = slave[k];
    for (int f = 0; f < nq; ++f) {
      v[m + f * n] += v[s + f * n];
      v[s + f * n] = 0.0;
    }
  }
}

// ---------------- deterministic reductions ----------------
// A reduction kernel writes one partial per block (wavefront shuffle, then
// the block's waves in order) and never synchronises with other blocks;
// the partials are summed in a fixed order by k_finish or, for the MGS
// chain, by every block of the next step (same code, same data: the same
// value in every block).  No atomics, so results do not depend on timing.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ inline void block_partial(double v, double* part) {
  __shared__ double ws[kRedThreads / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) ws[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) s += ws[w];
    part[blockIdx.x] = s;
  }
}

// sum of np partials in a fixed order, by a whole block; every thread
// returns the total
__device__ inline double block_sum_partials(const double* part, int np) {
  __shared__ double ws2[kRedThreads / 64];
  __shared__ double total;
  double s = 0.0;
  for (int b = threadIdx.x; b < np; b += blockDim.x) s += part[b];
  s = wave_sum(s);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) ws2[wid] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) t += ws2[w];
    total = t;
  }
  __syncthreads();
  return total;
}

// out[s] = sum of the partials of channel s (stride between channels):
// np0 partials in channel 0, np partials in the others
__global__ void __launch_bounds__(kRedThreads) k_finish(const double* part, int np0, int np,
                                                        long long stride, int nsums, double* out) {
  for (int s0 = 0; s0 < nsums; ++s0) {
    const double t = block_sum_partials(part + (size_t)s0 * stride, s0 == 0 ? np0 : np);
    if (threadIdx.x == 0) out[s0] = t;
    __syncthreads();
  }
}

// Reductions take one pass of kUnroll independent, coalesced loads per
// thread (grid sized by red_grid), so every load of a thread is in flight
// at once; a grid-stride tail covers vectors larger than the grid.
constexpr int kUnroll = 8;

__global__ void __launch_bounds__(kRedThreads) k_dot(const double* a, const double* b, long long n,
                                                     double* part) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double v[kUnroll];
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) v[u] = 0.0;
  for (; i + (kUnroll - 1) * stride < n; i += kUnroll * stride) {
    double x[kUnroll], y[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      x[u] = a[i + u * stride];
      y[u] = b[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) v[u] += x[u] * y[u];
  }
  for (; i < n; i += stride) v[0] += a[i] * b[i];
  double t = 0.0;
#pragma unroll
  for (int u = 0; u < kUnroll; ++u) t += v[u];
  block_partial(t, part);
}

// MGS step j >= 2: u <- u - beta_{j-1} u_{j-1};  beta_j = (u, u_j).
// beta_{j-1} is the sum of the previous step's np_prev partials.
__global__ void __launch_bounds__(kRedThreads) k_mgs(double* u, const double* ujm1, const double* uj,
                                                     const double* part_prev, int np_prev,
                                                     long long n, double* part) {
  const double beta = block_sum_partials(part_prev, np_prev);
  const bool self = uj == u;
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  constexpr int U = kUnroll / 2;
  double v[U];
#pragma unroll
  for (int k = 0; k < U; ++k) v[k] = 0.0;
  for (; i + (U - 1) * stride < n; i += U * stride) {
    double x[U], y[U], z[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      x[k] = u[i + k * stride];
      y[k] = ujm1[i + k * stride];
      z[k] = self ? 0.0 : uj[i + k * stride];
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const double w = x[k] - beta * y[k];
      u[i + k * stride] = w;
      v[k] += w * (self ? w : z[k]);
    }
  }
  for (; i < n; i += stride) {
    const double w = u[i] - beta * ujm1[i];
    u[i] = w;
    v[0] += w * (self ? w : uj[i]);
  }
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < U; ++k) t += v[k];
  block_partial(t, part);
}

// Classical Gram-Schmidt for the mixed-precision path: all dot products of
// u with the basis in one pass, then the update u -= sum_j h_j u_j fused
// with (u, u).  Up to kCgsMax basis vectors; partial sums per workgroup in
// channels 0..nv-1 (dots) and nv (norm), summed in a fixed order.
constexpr int kCgsMax = 32;

__device__ inline void block_partials(double* v, int nv, double* part, long long chan_stride) {
  __shared__ double ws[kCgsMax][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // compile-time indices keep v in registers
#pragma unroll
  for (int j = 0; j < kCgsMax; ++j) {
    if (j < nv) {
      const double t = wave_sum(v[j]);
      if (lane == 0) ws[j][wid] = t;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    double s = 0.0;
    for (int w = 0; w < kRedThreads / 64; ++w) s += ws[threadIdx.x][w];
    part[(size_t)threadIdx.x * chan_stride + blockIdx.x] = s;
  }
}

__global__ void __launch_bounds__(kRedThreads) k_cgs_dots(const double* __restrict__ u,
                                                          const double* __restrict__ basis, int nv,
                                                          long long n, double* part,
                                                          long long chan_stride) {
  double v[kCgsMax];
#pragma unroll
  for (int j = 0; j < kCgsMax; ++j) v[j] = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double x = u[i];
#pragma unroll
    for (int j = 0; j < kCgsMax; ++j)
      if (j < nv) v[j] += x * basis[(size_t)j * n + i];
  }
  block_partials(v, nv, part, chan_stride);
}

__global__ void __launch_bounds__(kRedThreads) k_cgs_update(double* __restrict__ u,
                                                            const double* __restrict__ basis,
                                                            int nv, const double* __restrict__ h,
                                                            long long n, double* part) {
  __shared__ double hs[kCgsMax];
  if ((int)threadIdx.x < nv) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double x = u[i];
    for (int j = 0; j < nv; ++j) x -= hs[j] * basis[(size_t)j * n + i];
    u[i] = x;
    acc += x * x;
  }
  block_partial(acc, part);
}

// u = src / sqrt(*sq); norm_out <- sqrt(*sq) (written by one lane)
__global__ void k_scale(double* u, const double* src, long long n, const double* sq,
                        double* norm_out) {
  const double nrm = sqrt(*sq);
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i == 0 && norm_out) *norm_out = nrm;
  for (; i < n; i += (long long)gridDim.x * blockDim.x) u[i] = src[i] / nrm;
}

struct Coef {
  double c[64];
};

// dy += sum_k c_k basis_k, k ascending (solgmr.f:690-692)
__global__ void k_combine(double* dy, const double* basis, int k, Coef cf, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    double v = dy[i];
    for (int j = 0; j < k; ++j) v = v + cf.c[j] * basis[(size_t)j * n + i];
    dy[i] = v;
  }
}

// stats: [sum res^2, sum rmes^2, nan/large counts of res, BDiag]
__global__ void __launch_bounds__(kRedThreads) k_stat_bad(const double* a, long long n, double* part) {
  double v = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double x = a[i];
    if (x != x || fabs(x) > 1.0e10) v += 1.0;
  }
  block_partial(v, part);
}

// ---------------- layout conversion ----------------
template <typename T>
__global__ void k_csr_to_slots(const double* ref, T* slots, const uint32_t* c2s, long long nnz) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint32_t qs = c2s[k];
  T* dst = slots + (size_t)(qs >> 6) * (25 * kGroup);
#pragma unroll
  for (int e = 0; e < 25; ++e) dst[slot_pos<T>(e, (int)(qs & 63))] = (T)ref[k * 25 + e];
}

template <typename T>
__global__ void k_slots_to_csr(const T* slots, double* ref, const uint32_t* c2s, long long nnz) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint32_t qs = c2s[k];
  const T* src = slots + (size_t)(qs >> 6) * (25 * kGroup);
#pragma unroll
  for (int e = 0; e < 25; ++e) ref[k * 25 + e] = (double)src[slot_pos<T>(e, (int)(qs & 63))];
}

// ---------------- itrPredict / itrCorrect / itrBC ----------------
// itrPredict, same-velocity predictor (itrPC.f:67-73): y = yold,
// ac = acold (1 - almi/gami)
__global__ void k_predict(const double* yold, const double* acold, double* y, double* ac,
                          long long n5, double fac) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n5) return;
  y[i] = yold[i];
  ac[i] = acold[i] * fac;
}

__global__ void k_correct(double* y, double* ac, const double* yold, const double* acold,
                          const double* dy, long long n, double fct1, double fct2) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  // y(:,1:3) -= Dy(:,2:4); y(:,4) -= Dy(:,1); y(:,5) -= Dy(:,5)
  const int from[5] = {1, 2, 3, 0, 4};
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    const double v = y[k + f * n] - dy[k + from[f] * n];
    y[k + f * n] = v;
    ac[k + f * n] = acold[k + f * n] * fct1 + (v - yold[k + f * n]) * fct2;
  }
}

// Dirichlet values on y (temperature, no-slip velocity, pressure)
__global__ void k_bc_y(double* y, const uint8_t* bcm, const double* BC, long long n) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint8_t m = bcm[k];
  if (!m) return;
  if (m & 16) y[k + 4 * n] = BC[k + n];
  if (m & 2) {
    y[k] = BC[k + 2 * n];
    y[k + n] = BC[k + 3 * n];
    y[k + 2 * n] = BC[k + 4 * n];
  }
  if (m & 1) y[k + 3 * n] = BC[k];
}

// periodic copy y(slave) = y(master), ac likewise
__global__ void k_per_copy(double* y, double* ac, const int* master, const int* ptr,
                           const int* slave, int nfold, long long n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfold) return;
  const long long m = master[i];
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const long long s = slave[k];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      y[s + f * n] = y[m + f * n];
      ac[s + f * n] = ac[m + f * n];
    }
  }
}

// itrUpdate (itrPC.f:204-207): acold += (ac-acold)/almi, yold += (y-yold)/alfi
__global__ void k_update(double* yold, double* acold, const double* y, const double* ac,
                         long long n5, double fct2, double fct3) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n5) return;
  acold[i] = acold[i] + (ac[i] - acold[i]) * fct2;
  yold[i] = yold[i] + (y[i] - yold[i]) * fct3;
}

inline int grid_for(long long n, int threads) {
  long long b = (n + threads - 1) / threads;
  return (int)(b < 1 ? 1 : b);
}

}  // namespace

int launch_i3lu(phg_ctx* c, double* diag, double* r, int code) {
  const long long n = c->nshg;
  const int g = grid_for(n, 256);
  if (code == 0)
    k_lu_fact<<<g, 256, 0, c->stream>>>(diag, n);
  else if (code == 1)
    k_lu_forward<<<g, 256, 0, c->stream>>>(diag, r, n);
  else if (code == 2)
    k_lu_backward<<<g, 256, 0, c->stream>>>(diag, r, n);
  else if (code == 3)
    k_lu_product<<<g, 256, 0, c->stream>>>(diag, r, n);
  else {
    phg_set_error("i3LU: unsupported code %d", code);
    return PHG_EINVAL;
  }
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_spsi3pre(phg_ctx* c) {
  PHG_TRY(require_matrix(c, "Spsi3pre"));
  const int g = (int)c->ngroups, nt = kPreWaves * kGroup;
  if (c->precision == PHG_MIXED)
    k_spsi3pre<float, float><<<g, nt, 0, c->stream>>>(
        c->lhs_aos32, c->lhs32, c->slot2csr, c->scol, c->grp_slot, c->bdiag, c->nshg);
  else
    k_spsi3pre<double, double><<<g, nt, 0, c->stream>>>(
        c->lhs_aos, c->lhs64, c->slot2csr, c->scol, c->grp_slot, c->bdiag, c->nshg);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_spmv(phg_ctx* c, const double* p, double* q) {
  PHG_TRY(require_matrix(c, "SparseAp"));
  const int g = (int)((c->ngroups + 7) / 8 * 8);  // one workgroup per group, XCD-strided
  const int nt = kSpmvWaves * kGroup;
  if (c->precision == PHG_MIXED)
    k_spmv<float, double><<<g, nt, 0, c->stream>>>(c->lhs32, c->scolp, c->grp_slot, p, q, c->nshg,
                                                   c->ngroups);
  else
    k_spmv<double, double><<<g, nt, 0, c->stream>>>(c->lhs64, c->scolp, c->grp_slot, p, q,
                                                    c->nshg, c->ngroups);
  PHG_HIP(hipGetLastError());
  return 0;
}

// q = K^T x for the periodic-folded operator K = F A G of the matvec:
// K^T = F A^T G (G copies masters to slaves, F folds slaves into masters).
// x is overwritten by G x.
int launch_spmv_t(phg_ctx* c, double* x, double* q) {
  if (!c->tslot) {
    phg_set_error("transpose matvec: slot map not built");
    return PHG_EINVAL;
  }
  if (c->nfold > 0) {
    k_per_copy_vec<<<(c->nfold + 255) / 256, 256, 0, c->stream>>>(x, c->fold_master, c->fold_ptr,
                                                                  c->fold_slave, c->nfold, c->nshg);
    PHG_HIP(hipGetLastError());
  }
  const int g = grid_for(c->nshg, 256);
  if (c->precision == PHG_MIXED)
    k_spmv_t<float><<<g, 256, 0, c->stream>>>(c->lhs32, c->scol, c->tslot, c->grp_slot, x, q,
                                              c->nshg);
  else
    k_spmv_t<double><<<g, 256, 0, c->stream>>>(c->lhs64, c->scol, c->tslot, c->grp_slot, x, q,
                                               c->nshg);
  PHG_HIP(hipGetLastError());
  return launch_bc3per(c, q, 5);
}

int launch_sub(phg_ctx* c, double* r, const double* b, const double* q, long long n) {
  int g = (int)std::min<long long>(grid_for(n, 256), 16384);
  k_sub<<<g, 256, 0, c->stream>>>(r, b, q, n);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_copy(phg_ctx* c, double* dst, const double* src, long long n) {
  if (n & 1) {  // copied as 16-byte pairs
    phg_set_error("copy: odd length %lld", n);
    return PHG_EINVAL;
  }
  const long long n2 = n / 2;
  const int g = grid_for(n2, 256);
  k_copy<<<g, 256, 0, c->stream>>>(reinterpret_cast<double2*>(dst),
                                   reinterpret_cast<const double2*>(src), n2);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_aos_to_ref(phg_ctx* c, double* ref) {
  const long long n = 25 * c->nnz;
  if (c->precision == PHG_MIXED)
    k_aos_to_ref<float><<<grid_for(n, 256), 256, 0, c->stream>>>(c->lhs_aos32, ref, c->nnz);
  else
    k_aos_to_ref<double><<<grid_for(n, 256), 256, 0, c->stream>>>(c->lhs_aos, ref, c->nnz);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_ref_to_aos(phg_ctx* c, const double* ref) {
  const long long n = (long long)aos_stride<double>() * c->nnz;
  k_ref_to_aos<<<grid_for(n, 256), 256, 0, c->stream>>>(ref, c->lhs_aos, c->nnz);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_bc3per(phg_ctx* c, double* v, int nq) {
  if (c->nfold <= 0) return 0;
  k_fold_vec<<<(c->nfold + 255) / 256, 256, 0, c->stream>>>(v, c->fold_master, c->fold_ptr,
                                                            c->fold_slave, c->nfold, c->nshg, nq);
  PHG_HIP(hipGetLastError());
  return 0;
}

// blocks so that one unrolled pass covers n (at most red_blocks)
inline int red_grid(const phg_ctx* c, long long n, int per_thread) {
  const long long b = (n + (long long)kRedThreads * per_thread - 1) / ((long long)kRedThreads * per_thread);
  return (int)std::max<long long>(1, std::min<long long>(b, c->red_blocks));
}

// partial channel k of the reduction scratch (red_blocks doubles each)
inline double* red_chan(phg_ctx* c, int k) { return c->red_part + (size_t)k * c->red_blocks; }

int launch_finish(phg_ctx* c, int chan0, int np0, int np, int nsums, double* out) {
  k_finish<<<1, kRedThreads, 0, c->stream>>>(red_chan(c, chan0), np0, np, c->red_blocks, nsums,
                                             out);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_dot_part(phg_ctx* c, const double* a, const double* b, long long n, int chan, int* np) {
  const int g = red_grid(c, n, kUnroll);
  k_dot<<<g, kRedThreads, 0, c->stream>>>(a, b, n, red_chan(c, chan));
  PHG_HIP(hipGetLastError());
  *np = g;
  return 0;
}

int launch_dot(phg_ctx* c, const double* a, const double* b, long long n, double* out) {
  int np = 0;
  PHG_TRY(launch_dot_part(c, a, b, n, 0, &np));
  return launch_finish(c, 0, np, np, 1, out);
}

int launch_mgs_part(phg_ctx* c, double* u, const double* ujm1, const double* uj, int chan_prev,
                    int np_prev, long long n, int chan, int* np) {
  const int g = red_grid(c, n, kUnroll / 2);
  k_mgs<<<g, kRedThreads, 0, c->stream>>>(u, ujm1, uj, red_chan(c, chan_prev), np_prev, n,
                                          red_chan(c, chan));
  PHG_HIP(hipGetLastError());
  *np = g;
  return 0;
}

int launch_cgs(phg_ctx* c, double* u, const double* basis, int nv, long long n, double* h) {
  if (nv < 1 || nv > kCgsMax) {
    phg_set_error("classical Gram-Schmidt: %d basis vectors (1..%d)", nv, kCgsMax);
    return PHG_EINVAL;
  }
  const int g = red_grid(c, n, 4);
  k_cgs_dots<<<g, kRedThreads, 0, c->stream>>>(u, basis, nv, n, red_chan(c, 0), c->red_blocks);
  PHG_HIP(hipGetLastError());
  PHG_TRY(launch_finish(c, 0, g, g, nv, h));
  k_cgs_update<<<g, kRedThreads, 0, c->stream>>>(u, basis, nv, h, n, red_chan(c, nv));
  PHG_HIP(hipGetLastError());
  return launch_finish(c, nv, g, g, 1, h + nv);
}

int launch_scale(phg_ctx* c, double* u, const double* src, long long n, const double* sq_ptr,
                 double* norm_out) {
  int g = (int)std::min<long long>(grid_for(n, 256), 16384);
  k_scale<<<g, 256, 0, c->stream>>>(u, src, n, sq_ptr, norm_out);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_combine(phg_ctx* c, double* dy, const double* basis, int k, const double* coef,
                   long long n) {
  Coef cf{};
  if (k > 64) {
    phg_set_error("combine: Kspace %d > 64", k);
    return PHG_EINVAL;
  }
  for (int j = 0; j < k; ++j) cf.c[j] = coef[j];
  int g = (int)std::min<long long>(grid_for(n, 256), 16384);
  k_combine<<<g, 256, 0, c->stream>>>(dy, basis, k, cf, n);
  PHG_HIP(hipGetLastError());
  return 0;
}

// out[0] = sum res^2, out[1] = bad entries of res, out[2] = bad entries of BDiag
int launch_stats(phg_ctx* c, double* out) {
  const long long n5 = 5 * c->nshg;
  int np = 0;
  PHG_TRY(launch_dot_part(c, c->res, c->res, n5, 0, &np));
  PHG_TRY(launch_finish(c, 0, np, np, 1, out));
  k_stat_bad<<<c->red_blocks, kRedThreads, 0, c->stream>>>(c->res, n5, red_chan(c, 1));
  k_stat_bad<<<c->red_blocks, kRedThreads, 0, c->stream>>>(c->bdiag, 5 * n5, red_chan(c, 2));
  PHG_HIP(hipGetLastError());
  return launch_finish(c, 1, c->red_blocks, c->red_blocks, 2, out + 1);
}

int launch_csr_to_slots(phg_ctx* c, const double* ref, double* slots) {
  k_csr_to_slots<double><<<grid_for(c->nnz, 256), 256, 0, c->stream>>>(ref, slots, c->csr2slot,
                                                                       c->nnz);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_slots_to_csr(phg_ctx* c, const double* slots, double* ref) {
  k_slots_to_csr<double><<<grid_for(c->nnz, 256), 256, 0, c->stream>>>(slots, ref, c->csr2slot,
                                                                       c->nnz);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_slots_to_csr32(phg_ctx* c, const float* slots, double* ref) {
  k_slots_to_csr<float><<<grid_for(c->nnz, 256), 256, 0, c->stream>>>(slots, ref, c->csr2slot,
                                                                      c->nnz);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_predict(phg_ctx* c) {
  // The reference's other predictors (ipred 2-4, itrPC.f:75-119) call
  // itrBC with iBC, BC, iper and ilwork undeclared in itrPredict, so they
  // are implicitly typed uninitialised scalars passed where arrays belong:
  // the reference defines no result for them.
  if (c->prm.ipred != 1) {
    phg_set_error("phg_predict: ipred %d; only the same-velocity predictor (ipred 1, "
                  "itrPC.f:67-73) has a defined result in the reference", c->prm.ipred);
    return PHG_EINVAL;
  }
  const long long n5 = 5 * c->nshg;
  const double fac = 1.0 - c->prm.almi / c->prm.gami;
  k_predict<<<grid_for(n5, 256), 256, 0, c->stream>>>(c->yold, c->acold, c->y, c->ac, n5, fac);
  PHG_HIP(hipGetLastError());
  return launch_itrbc(c, c->y, c->ac);
}

int launch_correct(phg_ctx* c) {
  const double fct1 = 1.0 - c->prm.almi / c->prm.gami;
  const double fct2 = c->prm.almi * c->prm.Dtgl / c->prm.gami / c->prm.alfi;
  k_correct<<<grid_for(c->nshg, 256), 256, 0, c->stream>>>(c->y, c->ac, c->yold, c->acold, c->dy,
                                                           c->nshg, fct1, fct2);
  PHG_HIP(hipGetLastError());
  return launch_itrbc(c, c->y, c->ac);
}

int launch_update(phg_ctx* c) {
  const long long n5 = 5 * c->nshg;
  k_update<<<grid_for(n5, 256), 256, 0, c->stream>>>(c->yold, c->acold, c->y, c->ac, n5,
                                                     1.0 / c->prm.almi, 1.0 / c->prm.alfi);
  PHG_HIP(hipGetLastError());
  return launch_itrbc(c, c->yold, c->acold);
}

int launch_itrbc(phg_ctx* c, double* y, double* ac) {
  k_bc_y<<<grid_for(c->nshg, 256), 256, 0, c->stream>>>(y, c->bcm, c->BCv, c->nshg);
  PHG_HIP(hipGetLastError());
  if (c->nfold > 0) {
    k_per_copy<<<(c->nfold + 255) / 256, 256, 0, c->stream>>>(y, ac, c->fold_master,
                                                              c->fold_ptr, c->fold_slave, c->nfold,
                                                              c->nshg);
    PHG_HIP(hipGetLastError());
  }
  return 0;
}

int launch_yshuffle(phg_ctx* c, const double* src, double* dst, bool new2old) {
  const long long n = c->nshg;
  k_yshuffle<<<grid_for(n, 256), 256, 0, c->stream>>>(src, dst, n, new2old ? 1 : 0);
  PHG_HIP(hipGetLastError());
  return 0;
}
