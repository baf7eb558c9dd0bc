// C-ABI of libphgpu (include/phgpu.h): set-up of the device-resident mesh,
// matrix layout and colouring, and the host-side sequencing of the
// SolGMRs/ElmGMRs work on one HIP stream.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "phg_internal.h"

// Staging memory of phg_spsi3pre_noalloc (lhsK in the reference layout on
// its way in and out), so that entry point calls no allocator.  Every GPU
// the library runs on holds one copy, from the module load to the end of
// the process.  Host code never passes `scratch` itself (its host-side name
// is not a device address): set-up looks up each device's address with
// hipGetSymbolAddress and the kernels get that pointer.
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

namespace {

char g_err[1024] = "";

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

// Address of the current device's scratch, looked up once per device.
int scratch_address(void** out) {
  *out = nullptr;
  int dev = -1;
  PHG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    phg_set_error("device %d outside the scratch table (%d devices)", dev, kMaxDevices);
    return PHG_EINVAL;
  }
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    PHG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) {
      phg_set_error("hipGetSymbolAddress(scratch) returned no address");
      return PHG_EHIP;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

template <typename T>
int dev_alloc(phg_ctx* c, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return 0;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    phg_set_error("hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
    return PHG_ENOMEM;
  }
  c->dev_bytes += (long long)(count * sizeof(T));
  return 0;
}

template <typename T>
int upload(phg_ctx* c, T* dst, const T* src, size_t count) {
  if (count == 0) return 0;
  PHG_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return 0;
}

template <typename T>
int download(phg_ctx* c, T* dst, const T* src, size_t count) {
  if (count == 0) return 0;
  PHG_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  return 0;
}

int sync(phg_ctx* c) {
  PHG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// Greedy first-fit colouring: no two elements of a colour share a node.
// Returns colour offsets and the element order (ascending inside a colour).
int colour_elements(long long nnode, long long nel, const int* ien /*1-based (nel,4)*/,
                    std::vector<int>& offsets, std::vector<int>& order) {
  std::vector<uint64_t> used(nnode, 0);
  std::vector<uint8_t> col(nel);
  int ncol = 0;
  for (long long e = 0; e < nel; ++e) {
    uint64_t m = 0;
    for (int a = 0; a < 4; ++a) m |= used[ien[e + nel * a] - 1];
    if (~m == 0) return PHG_ECOLOR;
    int k = __builtin_ctzll(~m);
    col[e] = (uint8_t)k;
    for (int a = 0; a < 4; ++a) used[ien[e + nel * a] - 1] |= (1ull << k);
    ncol = std::max(ncol, k + 1);
  }
  offsets.assign(ncol + 1, 0);
  for (long long e = 0; e < nel; ++e) offsets[col[e] + 1]++;
  for (int k = 0; k < ncol; ++k) offsets[k + 1] += offsets[k];
  order.resize(nel);
  std::vector<int> pos(offsets.begin(), offsets.end() - 1);
  for (long long e = 0; e < nel; ++e) order[pos[col[e]]++] = (int)e;
  return 0;
}

bool bc_code_supported(int ibc) {
  const int v = (ibc >> 3) & 7;
  if (ibc & 1) return false;          // density constraint
  if (ibc & (1 << 11)) return false;  // scaled plane extraction
  return v == 0 || v == 7;
}

int setup(phg_ctx* c, const phg_mesh* m) {
  const long long nshg = m->nshg, nel = m->nel, nelb = m->nelb, nnz = m->nnz_tot;
  // ---- validation ----
  for (long long n = 0; n < nshg; ++n) {
    if (!bc_code_supported(m->iBC[n])) {
      phg_set_error("iBC(%lld)=%d: only pressure, temperature, no-slip velocity and periodic "
                    "codes are supported on this path",
                    n + 1, m->iBC[n]);
      return PHG_EINVAL;
    }
    const int ip = m->iper[n] - 1;
    if (ip < 0 || ip >= nshg || m->iper[ip] - 1 != ip) {
      phg_set_error("iper(%lld) must point to a final master", n + 1);
      return PHG_EINVAL;
    }
  }
  for (long long i = 0; i < 4 * nel; ++i)
    if (m->ien[i] < 1 || m->ien[i] > nshg) {
      phg_set_error("ien entry %lld out of range", i);
      return PHG_EINVAL;
    }
  for (long long i = 0; i < 4 * nelb; ++i)
    if (m->ienb[i] < 1 || m->ienb[i] > nshg) {
      phg_set_error("ienb entry %lld out of range", i);
      return PHG_EINVAL;
    }
  for (long long k = 0; k < nnz; ++k)
    if (m->rowp[k] < 1 || m->rowp[k] > nshg) {
      phg_set_error("rowp entry %lld out of range", k);
      return PHG_EINVAL;
    }
  for (long long r = 0; r < nshg; ++r)
    if (m->colm[r + 1] < m->colm[r]) {
      phg_set_error("colm not monotone at %lld", r);
      return PHG_EINVAL;
    }
  if (m->colm[0] != 1 || m->colm[nshg] - 1 != nnz) {
    phg_set_error("colm inconsistent with nnz_tot");
    return PHG_EINVAL;
  }
  if (nshg >= (1LL << 31) || nnz >= (1LL << 31)) {
    phg_set_error("mesh too large for 32-bit indices");
    return PHG_EINVAL;
  }

  // ---- colouring of the wall elements (their residual is scattered) ----
  std::vector<int> orderb;
  if (colour_elements(nshg, nelb, m->ienb, c->colb_off, orderb)) {
    phg_set_error("wall element colouring needs more than 64 colours");
    return PHG_ECOLOR;
  }

  // ---- slot layout of the block-CSR matrix ----
  c->ngroups = (nshg + kGroup - 1) / kGroup;
  c->h_grp_slot.assign(c->ngroups + 1, 0);
  for (long long g = 0; g < c->ngroups; ++g) {
    int smax = 0;
    for (long long r = g * kGroup; r < std::min(nshg, (g + 1) * kGroup); ++r)
      smax = std::max(smax, m->colm[r + 1] - m->colm[r]);
    c->h_grp_slot[g + 1] = c->h_grp_slot[g] + smax;
  }
  c->nslots = c->h_grp_slot[c->ngroups];
  if (c->nslots * kGroup >= (1LL << 32)) {
    phg_set_error("matrix too large for 32-bit slot indices");
    return PHG_EINVAL;
  }
  std::vector<int> scol(c->nslots * kGroup), scolp(c->nslots * kGroup);
  std::vector<int> s2c(c->nslots * kGroup, -1);
  std::vector<uint32_t> c2s(nnz);
#pragma omp parallel for schedule(static)
  for (long long g = 0; g < c->ngroups; ++g) {
    const long long S = c->h_grp_slot[g + 1] - c->h_grp_slot[g];
    for (int r = 0; r < kGroup; ++r) {
      const long long row = g * kGroup + r;
      const long long b0 = row < nshg ? m->colm[row] - 1 : 0;
      const long long len = row < nshg ? m->colm[row + 1] - m->colm[row] : 0;
      for (long long s = 0; s < S; ++s) {
        const long long q = (c->h_grp_slot[g] + s) * kGroup + r;
        int col = row < nshg ? (int)row : 0;
        if (s < len) {
          col = m->rowp[b0 + s] - 1;
          c2s[b0 + s] = (uint32_t)q;
          s2c[q] = (int)(b0 + s);
        }
        scol[q] = col;
        scolp[q] = m->iper[col] - 1;
      }
    }
  }

  // ---- slot of the transposed block (i, j) -> (j, i) of every slot
  // position (transpose matvec); padding positions point to themselves ----
  std::vector<uint32_t> tsl(c->nslots * kGroup);
  for (long long q = 0; q < c->nslots * kGroup; ++q) tsl[q] = (uint32_t)q;
  {
    int bad_t = 0;
#pragma omp parallel for schedule(static) reduction(| : bad_t)
    for (long long i = 0; i < nshg; ++i) {
      const long long base = m->colm[i] - 1, len = m->colm[i + 1] - m->colm[i];
      for (long long t = 0; t < len; ++t) {
        const long long j = m->rowp[base + t] - 1;
        const int* lst = m->rowp + (m->colm[j] - 1);
        const int lenj = m->colm[j + 1] - m->colm[j];
        const int* f = std::lower_bound(lst, lst + lenj, (int)(i + 1));
        if (f == lst + lenj || *f != i + 1) {
          bad_t = 1;
          continue;
        }
        const long long qj = (c->h_grp_slot[j / kGroup] + (f - lst)) * kGroup + j % kGroup;
        tsl[(c->h_grp_slot[i / kGroup] + t) * kGroup + i % kGroup] = (uint32_t)qj;
      }
    }
    if (bad_t) {
      phg_set_error("colm/rowp is not structurally symmetric");
      return PHG_EINVAL;
    }
  }

  // ---- element -> block map (replaces the per-entry sparseloc search) ----
  std::vector<int> ienc(4 * nel);
  std::vector<uint32_t> bmap(16 * nel);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (long long e = 0; e < nel; ++e) {
    int nd[4];
    for (int a = 0; a < 4; ++a) {
      nd[a] = m->ien[e + nel * a] - 1;
      ienc[a * nel + e] = nd[a];
    }
    for (int a = 0; a < 4; ++a) {
      const int* lst = m->rowp + (m->colm[nd[a]] - 1);
      const int len = m->colm[nd[a] + 1] - m->colm[nd[a]];
      for (int b = 0; b < 4; ++b) {
        const int* it = std::lower_bound(lst, lst + len, nd[b] + 1);
        if (it == lst + len || *it != nd[b] + 1) {
          bad = 1;
          continue;
        }
        bmap[(4 * a + b) * nel + e] = (uint32_t)((m->colm[nd[a]] - 1) + (it - lst));
      }
    }
  }
  if (bad) {
    phg_set_error("element adjacency missing from colm/rowp");
    return PHG_EINVAL;
  }
  if (16 * nel >= (1LL << 31)) {
    phg_set_error("too many elements for 32-bit block lists");
    return PHG_EINVAL;
  }
  // ---- block owners: the element blocks of every CSR entry, in element
  // order, cut into chunks of at most kChunk; the chunks of one entry are
  // consecutive lanes of one 64-lane group ----
  std::vector<int> bcount(nnz + 1, 0);
  for (long long i = 0; i < 16 * nel; ++i) bcount[bmap[i] + 1]++;
  for (long long k = 0; k < nnz; ++k) bcount[k + 1] += bcount[k];
  std::vector<int> blocks(16 * nel);
  {
    std::vector<int> pos(bcount.begin(), bcount.end() - 1);
    for (long long e = 0; e < nel; ++e)
      for (int ab = 0; ab < 16; ++ab) blocks[pos[bmap[ab * nel + e]]++] = (int)(e * 16 + ab);
  }
  std::vector<int4> lanes;
  lanes.reserve(nnz + 4 * nshg);
  // the same chunks for the diagonal entries alone: the pass without lhsK
  // (residual and BDiag only) runs these lanes
  std::vector<int4> lanes_diag;
  lanes_diag.reserve(4 * nshg);
  // the pass of k_asm_pairs: each off-diagonal pair (A,B), (B,A) once, on
  // the lanes of the entry whose row comes first in the visiting order,
  // with the transposed entry beside.  Pair lanes and diagonal lanes fill
  // 64-lane wavefronts of their own (pw, dw), each appended to the list
  // when full, so both kinds follow the row order together and share the
  // point states in L2
  std::vector<int4> lanes_pair, pw, dw;
  std::vector<int> lanes_k2, pk;
  lanes_pair.reserve(nnz / 2 + 4 * nshg);
  lanes_k2.reserve(nnz / 2 + 4 * nshg);
  auto flush = [&](std::vector<int4>& wv, std::vector<int>* k2v) {
    if (wv.empty()) return;
    while (wv.size() % 64) {
      wv.push_back(make_int4(0, 0, 0, 0));
      if (k2v) k2v->push_back(-1);
    }
    lanes_pair.insert(lanes_pair.end(), wv.begin(), wv.end());
    if (k2v)
      lanes_k2.insert(lanes_k2.end(), k2v->begin(), k2v->end());
    else
      lanes_k2.resize(lanes_pair.size(), -1);
    wv.clear();
    if (k2v) k2v->clear();
  };
  std::vector<int> blk_lane;  // blocks in lane order
  blk_lane.reserve(16 * nel);
  // rows in Morton order of their node coordinates, so the lanes that run
  // together touch a compact group of elements (cache reuse of the point
  // states)
  std::vector<int> rows(nshg);
  {
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
      lo[d] = hi[d] = m->x[d * nshg];
      for (long long r = 0; r < nshg; ++r) {
        lo[d] = std::min(lo[d], m->x[r + d * nshg]);
        hi[d] = std::max(hi[d], m->x[r + d * nshg]);
      }
    }
    std::vector<uint64_t> key(nshg);
    for (long long r = 0; r < nshg; ++r) {
      uint64_t k = 0;
      uint32_t q[3];
      for (int d = 0; d < 3; ++d) {
        const double span = hi[d] > lo[d] ? hi[d] - lo[d] : 1.0;
        q[d] = (uint32_t)std::min(1023.0, std::max(0.0, (m->x[r + d * nshg] - lo[d]) / span * 1023.0));
      }
      for (int b = 9; b >= 0; --b)
        for (int d = 2; d >= 0; --d) k = (k << 1) | ((q[d] >> b) & 1u);
      key[r] = k;
    }
    for (long long r = 0; r < nshg; ++r) rows[r] = (int)r;
    std::stable_sort(rows.begin(), rows.end(), [&](int a, int b) { return key[a] < key[b]; });
  }
  std::vector<int> rank(nshg);
  for (long long ri = 0; ri < nshg; ++ri) rank[rows[ri]] = (int)ri;
  for (long long ri = 0; ri < nshg; ++ri) {
    const long long row = rows[ri];
    for (long long k = m->colm[row] - 1; k < m->colm[row + 1] - 1; ++k) {
      const int n = bcount[k + 1] - bcount[k];
      if (n == 0) {
        phg_set_error("CSR entry %lld has no element block", k);
        return PHG_EINVAL;
      }
      const int nch = (n + kChunk - 1) / kChunk;
      if (nch > 64 || n >= (1 << 15)) {
        phg_set_error("CSR entry %lld has %d element blocks", k, n);
        return PHG_EINVAL;
      }
      if ((long long)(lanes.size() % 64) + nch > 64)
        while (lanes.size() % 64) lanes.push_back(make_int4(0, 0, 0, 0));
      const bool dg = (m->rowp[k] - 1) == row;
      if (dg && (long long)(lanes_diag.size() % 64) + nch > 64)
        while (lanes_diag.size() % 64) lanes_diag.push_back(make_int4(0, 0, 0, 0));
      const int col = m->rowp[k] - 1;
      const bool first = !dg && rank[col] > rank[row];
      int kt = -1;
      if (first) {
        const int* lst = m->rowp + (m->colm[col] - 1);
        const int len = m->colm[col + 1] - m->colm[col];
        const int* it = std::lower_bound(lst, lst + len, (int)row + 1);
        if (it == lst + len || *it != (int)row + 1) {
          phg_set_error("colm/rowp is not structurally symmetric");
          return PHG_EINVAL;
        }
        kt = (m->colm[col] - 1) + (int)(it - lst);
        if ((long long)pw.size() + nch > 64) flush(pw, &pk);
      }
      if (dg && (long long)dw.size() + nch > 64) flush(dw, nullptr);
      for (int ch = 0; ch < nch; ++ch) {
        const int b0 = (int)((long long)n * ch / nch), b1 = (int)((long long)n * (ch + 1) / nch);
        int flags = (dg ? 2 : 0) | (ch == 0 ? 1 : 0);
        if (ch == 0) flags |= (nch << 8) | (n << 16);
        const int4 ln = make_int4((int)blk_lane.size(), (int)k, b1 - b0, flags);
        lanes.push_back(ln);
        if (dg) lanes_diag.push_back(ln);
        if (first) {
          pw.push_back(ln);
          pk.push_back(kt);
        }
        if (dg) dw.push_back(ln);
        for (int b = b0; b < b1; ++b) blk_lane.push_back(blocks[bcount[k] + b]);
      }
    }
  }
  c->nlanes = (long long)lanes.size();
  c->nlanes_diag = (long long)lanes_diag.size();
  flush(pw, &pk);
  flush(dw, nullptr);
  c->nlanes_pair = (long long)lanes_pair.size();
  blocks.clear();
  blocks.shrink_to_fit();
  // ---- node -> (element, local node) incidence, element order per node
  // (the scatter of AsIq's element flux residuals, asiq.f:62-63), sliced
  // by 64-node groups: entry j of node 64 g + l at goff[g] + 64 j + l ----
  const long long ngq = (nshg + 63) / 64;
  std::vector<long long> goff(ngq + 1, 0);
  std::vector<int> incs;
  std::vector<int> incp(nshg + 1, 0), incl(4 * nel);
  {
    for (long long i = 0; i < 4 * nel; ++i) incp[ienc[i] + 1]++;
    for (long long n = 0; n < nshg; ++n) incp[n + 1] += incp[n];
    std::vector<int> pos(incp.begin(), incp.end() - 1);
    for (long long e = 0; e < nel; ++e)
      for (int a = 0; a < 4; ++a) incl[pos[ienc[a * nel + e]]++] = (int)(e * 4 + a);
    for (long long g = 0; g < ngq; ++g) {
      int deg = 0;
      for (long long n = 64 * g; n < std::min(nshg, 64 * g + 64); ++n)
        deg = std::max(deg, incp[n + 1] - incp[n]);
      goff[g + 1] = goff[g] + 64LL * deg;
    }
    incs.assign(goff[ngq], -1);
    for (long long n = 0; n < nshg; ++n) {
      const long long g = n / 64, l = n % 64;
      for (int j = 0; j < incp[n + 1] - incp[n]; ++j) incs[goff[g] + 64LL * j + l] = incl[incp[n] + j];
    }
  }
  incl.clear();
  incl.shrink_to_fit();
  // the slot of each (element, node) in that layout: the per-element
  // products of the tangent and of the modified residual are written there,
  // so the node sums read whole rows
  if (goff[ngq] >= (1LL << 31)) {
    phg_set_error("too many node incidences for 32-bit slots");
    return PHG_EINVAL;
  }
  std::vector<int> eslot(4 * nel, -1);
  for (long long p = 0; p < goff[ngq]; ++p)
    if (incs[p] >= 0) eslot[incs[p]] = (int)p;
  std::vector<int> ienbc(4 * std::max<long long>(nelb, 1));
  for (long long p = 0; p < nelb; ++p)
    for (int a = 0; a < 4; ++a) ienbc[a * nelb + p] = m->ienb[orderb[p] + nelb * a] - 1;

  // ---- periodic fold lists ----
  std::vector<std::vector<int>> sl(nshg);
  std::vector<int> masters;
  for (long long j = 0; j < nshg; ++j)
    if (m->iBC[j] & 1024) {
      const int i = m->iper[j] - 1;
      if (i == j) continue;
      if (sl[i].empty()) masters.push_back(i);
      sl[i].push_back((int)j);
    }
  std::sort(masters.begin(), masters.end());
  std::vector<int> fptr(1, 0), fsl;
  for (int mm : masters) {
    for (int s : sl[mm]) fsl.push_back(s);
    fptr.push_back((int)fsl.size());
  }
  c->nfold = (int)masters.size();

  std::vector<uint8_t> bcm(nshg);
  std::vector<int> iper0(nshg);
  for (long long n = 0; n < nshg; ++n) {
    bcm[n] = constraint_mask(m->iBC[n]);
    iper0[n] = m->iper[n] - 1;
  }

  // ---- device allocation ----
  const int K = std::max(1, c->prm.Kspace);
  c->kcap = K;
  const bool mixed = c->precision == PHG_MIXED;
  PHG_TRY(dev_alloc(c, &c->x, 3 * nshg));
  PHG_TRY(dev_alloc(c, &c->ienc, 4 * nel));
  // a matrix-free context keeps no lhsK, no matvec layout and no owner
  // lanes of the off-diagonal entries (itrdrv.f:180-185 allocates BDiag
  // alone for SolMFG)
  const bool mat = !c->matrix_free;
  if (mat) PHG_TRY(dev_alloc(c, &c->lanes, c->nlanes));
  if (mat) PHG_TRY(dev_alloc(c, &c->lanes_pair, c->nlanes_pair));
  if (mat) PHG_TRY(dev_alloc(c, &c->lanes_k2, c->nlanes_pair));
  PHG_TRY(dev_alloc(c, &c->lanes_diag, c->nlanes_diag));
  PHG_TRY(dev_alloc(c, &c->blocks, 16 * nel));
  // point records: 12 values per point; geometry: 13 values padded to 16
  // floats or 14 doubles (16-byte loads)
  if (mixed) {
    PHG_TRY(dev_alloc(c, &c->qps32, (size_t)nel * 4 * 12));
    PHG_TRY(dev_alloc(c, &c->geo32, (size_t)nel * 16));
  } else {
    PHG_TRY(dev_alloc(c, &c->qps, (size_t)nel * 4 * 12));
    PHG_TRY(dev_alloc(c, &c->geo, (size_t)nel * 14));
  }
  // element residuals (element order), then the element products in the
  // incidence slots (k_jv, k_mres)
  PHG_TRY(dev_alloc(c, &c->resid, (size_t)std::max<long long>(nel * 4, goff[ngq]) * 5));
  PHG_TRY(dev_alloc(c, &c->eslot, (size_t)nel * 4));
  PHG_TRY(dev_alloc(c, &c->ienbc, ienbc.size()));
  PHG_TRY(dev_alloc(c, &c->bcm, nshg));
  PHG_TRY(dev_alloc(c, &c->iper, nshg));
  PHG_TRY(dev_alloc(c, &c->BCv, (size_t)nshg * m->ndofBC));
  PHG_TRY(dev_alloc(c, &c->fold_master, std::max(1, c->nfold)));
  PHG_TRY(dev_alloc(c, &c->fold_ptr, c->nfold + 1));
  PHG_TRY(dev_alloc(c, &c->fold_slave, std::max<size_t>(1, fsl.size())));
  PHG_TRY(dev_alloc(c, &c->grp_slot, c->ngroups + 1));
  if (mat) {
    PHG_TRY(dev_alloc(c, &c->scol, c->nslots * kGroup));
    PHG_TRY(dev_alloc(c, &c->scolp, c->nslots * kGroup));
    PHG_TRY(dev_alloc(c, &c->tslot, c->nslots * kGroup));
    PHG_TRY(dev_alloc(c, &c->csr2slot, nnz));
    PHG_TRY(dev_alloc(c, &c->slot2csr, c->nslots * kGroup));
  }
  PHG_TRY(dev_alloc(c, &c->acc_res, 5 * nshg));
  if (mat && mixed)
    PHG_TRY(dev_alloc(c, &c->lhs_aos32, (size_t)aos_stride<float>() * nnz));
  else if (mat)
    PHG_TRY(dev_alloc(c, &c->lhs_aos, (size_t)aos_stride<double>() * nnz));
  PHG_TRY(dev_alloc(c, &c->acc_bd, 25 * nshg));
  PHG_TRY(dev_alloc(c, &c->nodepack, 16 * nshg));
  PHG_TRY(dev_alloc(c, &c->inc_goff, goff.size()));
  PHG_TRY(dev_alloc(c, &c->inc, incs.size()));
  if (mat && mixed)
    PHG_TRY(dev_alloc(c, &c->lhs32, c->nslots * kGroup * 25));
  else if (mat)
    PHG_TRY(dev_alloc(c, &c->lhs64, c->nslots * kGroup * 25));
  PHG_TRY(dev_alloc(c, &c->res, 5 * nshg));
  PHG_TRY(dev_alloc(c, &c->rmes, 5 * nshg));
  PHG_TRY(dev_alloc(c, &c->bdiag, 25 * nshg));
  PHG_TRY(dev_alloc(c, &c->y, 5 * nshg));
  PHG_TRY(dev_alloc(c, &c->ac, 5 * nshg));
  PHG_TRY(dev_alloc(c, &c->yold, 5 * nshg));
  PHG_TRY(dev_alloc(c, &c->acold, 5 * nshg));
  PHG_TRY(dev_alloc(c, &c->dy, 5 * nshg));
  // K+1 basis vectors; at least 3 (the refinement loop uses 1 and 2 as scratch)
  PHG_TRY(dev_alloc(c, &c->krylov, (size_t)(std::max(K, 2) + 1) * 5 * nshg));
  c->red_blocks = 1024;
  PHG_TRY(dev_alloc(c, &c->red_part, (size_t)(K + 3) * c->red_blocks));
  PHG_TRY(dev_alloc(c, &c->scal, 256));
  PHG_HIP(hipHostMalloc((void**)&c->h_scal, 256 * sizeof(double), hipHostMallocDefault));

  std::vector<double> xs(3 * nshg);
  std::copy(m->x, m->x + 3 * nshg, xs.begin());
  PHG_TRY(upload(c, c->x, xs.data(), xs.size()));
  PHG_TRY(upload(c, c->ienc, ienc.data(), ienc.size()));
  PHG_TRY(upload(c, c->inc_goff, goff.data(), goff.size()));
  PHG_TRY(upload(c, c->inc, incs.data(), incs.size()));
  PHG_TRY(upload(c, c->eslot, eslot.data(), eslot.size()));
  if (mat) PHG_TRY(upload(c, c->lanes, lanes.data(), lanes.size()));
  if (mat) PHG_TRY(upload(c, c->lanes_pair, lanes_pair.data(), lanes_pair.size()));
  if (mat) PHG_TRY(upload(c, c->lanes_k2, lanes_k2.data(), lanes_k2.size()));
  PHG_TRY(upload(c, c->lanes_diag, lanes_diag.data(), lanes_diag.size()));
  PHG_TRY(upload(c, c->blocks, blk_lane.data(), blk_lane.size()));
  PHG_TRY(upload(c, c->ienbc, ienbc.data(), ienbc.size()));
  PHG_TRY(upload(c, c->bcm, bcm.data(), bcm.size()));
  PHG_TRY(upload(c, c->iper, iper0.data(), iper0.size()));
  PHG_TRY(upload(c, c->BCv, m->BC, (size_t)nshg * m->ndofBC));
  if (c->nfold > 0) {
    PHG_TRY(upload(c, c->fold_master, masters.data(), masters.size()));
    PHG_TRY(upload(c, c->fold_slave, fsl.data(), fsl.size()));
  }
  PHG_TRY(upload(c, c->fold_ptr, fptr.data(), fptr.size()));
  PHG_TRY(upload(c, c->grp_slot, c->h_grp_slot.data(), c->h_grp_slot.size()));
  if (mat) {
    PHG_TRY(upload(c, c->scol, scol.data(), scol.size()));
    PHG_TRY(upload(c, c->scolp, scolp.data(), scolp.size()));
    PHG_TRY(upload(c, c->tslot, tsl.data(), tsl.size()));
    PHG_TRY(upload(c, c->csr2slot, c2s.data(), c2s.size()));
    PHG_TRY(upload(c, c->slot2csr, s2c.data(), s2c.size()));
    if (mixed)
      PHG_TRY(launch_fill(c, c->lhs32, sizeof(float) * 25 * kGroup * c->nslots));
    else
      PHG_TRY(launch_fill(c, c->lhs64, sizeof(double) * 25 * kGroup * c->nslots));
  }
  PHG_TRY(launch_fill(c, c->yold, sizeof(double) * 5 * nshg));
  PHG_TRY(launch_fill(c, c->acold, sizeof(double) * 5 * nshg));
  PHG_TRY(sync(c));
  return 0;
}

int ensure_tmp(phg_ctx* c, size_t bytes) {
  if (c->tmp_bytes >= bytes) return 0;
  if (c->tmp) PHG_HIP(hipFree(c->tmp));
  c->tmp = nullptr;
  c->tmp_bytes = 0;
  PHG_HIP(hipMalloc((void**)&c->tmp, bytes));
  c->tmp_bytes = bytes;
  return 0;
}

double* kvec(phg_ctx* c, int k) { return c->krylov + (size_t)k * 5 * c->nshg; }

// ElmGMRs on the resident y/ac: res, rmes(=0), BDiag, lhsK (elmgmr.f:395-610)
// idiff 0, 1 or 3; the flux buffers are allocated when first needed
int check_idiff(phg_ctx* c) {
  const int d = c->prm.idiff;
  if (d != 0 && d != 1 && d != 3) {
    phg_set_error("idiff %d: 0, 1 (row-sum) or 3 (special lumping) supported; with 2 the "
                  "reference never forms qres (elmgmr.f:361)", d);
    return PHG_EINVAL;
  }
  if (c->prm.idc < 0 || c->prm.idc > 3) {
    // e3dc.f:246-260 defines DC for iDC 1, 2, 3 only
    phg_set_error("iDC %d: 0 (off), 1, 2 or 3 supported", c->prm.idc);
    return PHG_EINVAL;
  }
  if (c->prm.idc != 0) {
    // DC factor per point; with the matrix tau DC giju per point (8 values)
    const size_t need = (size_t)c->nel * 4 * (c->prm.itau >= 10 ? 8 : 1);
    if (c->dcw_len < need) {
      if (c->dcw) {
        PHG_HIP(hipStreamSynchronize(c->stream));
        PHG_HIP(hipFree(c->dcw));
        c->dev_bytes -= (long long)(sizeof(double) * c->dcw_len);
        c->dcw = nullptr;
      }
      PHG_TRY(dev_alloc(c, &c->dcw, need));
      c->dcw_len = need;
    }
  }
  // itau 2 and 3 have no branch in the compressible e3tau (tau left unset);
  // 11 (Matrix-Modal) reaches the same e3tau_nd as 10 (e3ls.f:188-214)
  if (c->prm.itau != 0 && c->prm.itau != 1 && c->prm.itau != 10 && c->prm.itau != 11) {
    phg_set_error("itau %d: the diagonal taus of Shakib (0) and Franca (1) and the matrix "
                  "tau of e3tau_nd (10, 11) are built", c->prm.itau);
    return PHG_EINVAL;
  }
  if (c->prm.itau >= 10 && !c->mtau) {
    // W PTau per point: 25 values padded to 28 floats or 26 doubles
    const size_t rec = c->precision == PHG_MIXED ? 28 * sizeof(float) : 26 * sizeof(double);
    uint8_t* p = nullptr;
    PHG_TRY(dev_alloc(c, &p, (size_t)c->nel * 4 * rec));
    c->mtau = p;
  }
  if (d != 0 && !c->qnode) {
    PHG_TRY(dev_alloc(c, &c->qrec, (size_t)c->nel * 16));
    PHG_TRY(dev_alloc(c, &c->qnode, (size_t)c->nshg * 16));
  }
  return 0;
}

int elm_core(phg_ctx* c) {
  const long long n = c->nshg;
  const bool lhs = c->prm.lhs == 1, prec = c->prm.iprec != 0;
  PHG_TRY(check_idiff(c));
  // every node owns its diagonal block and every CSR entry has an owner, so
  // the assembly stores res, BDiag and lhsK completely: no clearing
  PHG_TRY(launch_assembly(c, lhs, prec));
  PHG_TRY(launch_bc_res_bdiag(c, prec));
  return 0;
}

// ---------------- matrix-free GMRES (SolMFG, solmfg.f) ----------------
int mfg_alloc(phg_ctx* c) {
  const long long n = c->nshg;
  if (!c->mfg.ypre) PHG_TRY(dev_alloc(c, &c->mfg.ypre, 5 * n));
  if (!c->mfg.t1) PHG_TRY(dev_alloc(c, &c->mfg.t1, 5 * n));
  if (!c->mfg.t2) PHG_TRY(dev_alloc(c, &c->mfg.t2, 5 * n));
  if (!c->mfg.pack) PHG_TRY(dev_alloc(c, &c->mfg.pack, 8 * n));
  return 0;
}

// Au1MFG (au1mfg.f:56-83): w = (L^-1 Rmod(itrBC(U^-1 (ypre + e v))) - rmes) / e
// with the forward-reduced modified residual rmes of the base point
int mfg_apply(phg_ctx* c, const double* v, double* w) {
  PHG_TRY(launch_mfg_point(c, c->mfg.ypre, v, c->mfg.eg, c->mfg.pack, true, true));
  PHG_TRY(launch_mres(c, c->mfg.pack, c->mfg.t1, false, false));
  return launch_mfg_fwd_diff(c, c->mfg.t1, c->rmes, c->mfg.eg, w);
}

// Au2MFG (au2mfg.f:54-112): out = res - (L^-1 Rmod(y+) - L^-1 Rmod(y-)) / (2 eps),
// y+- = itrBC(U^-1 (ypre +- eps Dy)), eps = epsM^(2/3) / ||Dy||
int mfg_au2(phg_ctx* c, double* out) {
  const long long V = 5 * c->nshg;
  double* S = c->scal + 200;
  PHG_TRY(launch_dot(c, c->dy, c->dy, V, S));
  PHG_TRY(download(c, c->h_scal + 200, S, 1));
  PHG_TRY(sync(c));
  const double eps = std::pow(c->prm.epsM, 0.6666666666666666666666666666667) /
                     std::sqrt(c->h_scal[200]);
  PHG_TRY(launch_mfg_point(c, c->mfg.ypre, c->dy, eps, c->mfg.pack, true, true));
  PHG_TRY(launch_mres(c, c->mfg.pack, c->mfg.t1, false, false));
  PHG_TRY(launch_i3lu(c, c->bdiag, c->mfg.t1, 1));
  PHG_TRY(launch_mfg_point(c, c->mfg.ypre, c->dy, -eps, c->mfg.pack, true, true));
  PHG_TRY(launch_mres(c, c->mfg.pack, c->mfg.t2, false, false));
  PHG_TRY(launch_i3lu(c, c->bdiag, c->mfg.t2, 1));
  return launch_mfg_combine(c, c->res, c->mfg.t1, c->mfg.t2, 2.0 * eps, out, 0);
}

// itrFDI (itrfdi.f:38-141): eGMRES = 2 sqrt(epsA / SDnrm) with the
// cancellation error epsA = epsM^2 ||L^-1 |Rmod|(ypre)|| and the second
// difference SDnrm = ||(L^-1 (Rmod(ypre + h u) + Rmod(ypre - h u)) -
// 2 rmes) / epsM||, h = sqrt(epsM); no itrBC on these points
int mfg_fdi(phg_ctx* c, const double* u, double* eg) {
  const long long V = 5 * c->nshg;
  const double epsM = c->prm.epsM;
  double* S = c->scal + 200;
  PHG_TRY(launch_mfg_point(c, c->mfg.ypre, nullptr, 0.0, c->mfg.pack, false, true));
  PHG_TRY(launch_mres(c, c->mfg.pack, c->mfg.t1, false, true));
  PHG_TRY(launch_i3lu(c, c->bdiag, c->mfg.t1, 1));
  PHG_TRY(launch_dot(c, c->mfg.t1, c->mfg.t1, V, S));
  const double h = std::sqrt(epsM);
  PHG_TRY(launch_mfg_point(c, c->mfg.ypre, u, h, c->mfg.pack, false, true));
  PHG_TRY(launch_mres(c, c->mfg.pack, c->mfg.t1, false, false));
  PHG_TRY(launch_mfg_point(c, c->mfg.ypre, u, -h, c->mfg.pack, false, true));
  PHG_TRY(launch_mres(c, c->mfg.pack, c->mfg.t1, true, false));
  PHG_TRY(launch_i3lu(c, c->bdiag, c->mfg.t1, 1));
  PHG_TRY(launch_mfg_combine(c, c->rmes, c->mfg.t1, nullptr, epsM, c->mfg.t2, 1));
  PHG_TRY(launch_dot(c, c->mfg.t2, c->mfg.t2, V, S + 1));
  PHG_TRY(download(c, c->h_scal + 200, S, 2));
  PHG_TRY(sync(c));
  const double epsA = (epsM * epsM) * std::sqrt(c->h_scal[200]);
  const double SDnrm = std::sqrt(c->h_scal[201]);
  *eg = 2.0 * std::sqrt(epsA / SDnrm);
  return 0;
}

// Flags the matrix-free path restates: the frozen set, idiff 0/1/3, lhs = 0
// (itrdrv.f:494), itau 0 or 10 (e3bdg or e3bdg_nd for BDiag); iDC is
// refused (e3dc's ires = 2 branch reads giju, rTLS and raLS that e3tau forms
// only for ires 1 and 3, e3tau.f:185-262)
int mfg_check(phg_ctx* c) {
  if (c->precision != PHG_FP64) {
    phg_set_error("SolMFG: finite-difference products need a PHG_FP64 context");
    return PHG_EINVAL;
  }
  if (c->prm.lhs != 0) {
    phg_set_error("SolMFG: lhs must be 0 (itrdrv.f:494); got %d", c->prm.lhs);
    return PHG_EINVAL;
  }
  if (c->prm.idc != 0) {
    phg_set_error("SolMFG: iDC %d not supported by the matrix-free path", c->prm.idc);
    return PHG_EINVAL;
  }
  const int K = c->prm.Kspace;
  if (K < 1 || K > c->kcap || K > 60 || c->prm.nGMRES < 1) {
    phg_set_error("SolMFG: Kspace %d outside 1..%d or nGMRES %d < 1", K, std::min(c->kcap, 60),
                  c->prm.nGMRES);
    return PHG_EINVAL;
  }
  return check_idiff(c);
}

// ElmMFG (elmmfg.f:128-256) on the resident y/ac: res and BDiag as the
// assembly forms them without lhsK (the tangent's diagonal blocks, which
// e3bdg.f writes out term by term; with the matrix tau e3bdg_nd's, the
// block pass adding its mis-indexed entries), and the modified residual
// rmes.  The point pass also leaves the state's W PTau records that the
// modified residual reads with the matrix tau.
int elm_mfg(phg_ctx* c) {
  const bool prec = c->prm.iprec != 0;
  PHG_TRY(mfg_check(c));
  PHG_TRY(mfg_alloc(c));
  PHG_TRY(launch_assembly(c, false, prec));
  PHG_TRY(launch_bc_res_bdiag(c, prec));
  return launch_mres(c, nullptr, c->rmes, false, false);
}

// One GMRES cycle on the scaled system (solgmr.f:522-692): kvec(0) holds the
// normalised start vector, unorm its norm.  Adds sum_k y_k u_k to c->dy.
struct CycleOut {
  int iKs = 0;
  double echeck = 0.0;
  bool rejected = false;  // the network's first direction lost the acceptance test
};

// flexible: precondition each direction with the network (FGMRES, right
// preconditioning) and combine the preconditioned directions.
// basis sizes the classical Gram-Schmidt kernels take (launch_cgs)
constexpr int kCgsVectors = 32;

int gmres_cycle(phg_ctx* c, double unorm, double epsnrm, phg_stats* st, CycleOut* out,
                bool flexible = false) {
  const long long V = 5 * c->nshg;
  auto zvec = [&](int k) { return c->mp.zbasis + (size_t)k * V; };
  const int K = c->prm.Kspace;
  double* H = c->scal + 16;
  int iKs = 0;
  double echeck = 0.0;
  std::vector<double> Hh((size_t)(K + 2) * (K + 1), 0.0), e(K + 2, 0.0), Rc(K + 2, 0.0),
      Rs(K + 2, 0.0), yB(K + 2, 0.0);
  auto Hm = [&](int a, int b) -> double& { return Hh[(size_t)(a - 1) + (size_t)(K + 1) * (b - 1)]; };
  e[1] = unorm;
  bool use_net = flexible && c->mp.use_pc >= 1;
  for (int iK = 1; iK <= K; ++iK) {
    iKs = iK;
    double* w = kvec(c, iK);
    const double* dir = kvec(c, iK - 1);
    bool have_w = false;
    if (flexible) {
      // the network preconditions the first use_pc directions of the cycle,
      // the block-diagonal preconditioner (identity here) the rest
      if (use_net && iK <= c->mp.use_pc) {
        PHG_TRY(mp_forward(c, kvec(c, iK - 1), zvec(iK - 1), false));
        if (iK == 1) {
          // acceptance test, block-diagonal GMRES as the yardstick: the
          // one-step residual min_a ||u_1 - a K z||^2 = 1 - (u_1.Kz)^2/|Kz|^2
          // of the network's direction z = N(u_1) against that of z = u_1;
          // the network's direction is kept unless it is larger
          double* wp = c->mp.wplain;
          double* sb = c->mp.sbuf + 8;
          PHG_HIP(hipEventRecord(c->kev[0], c->stream));
          PHG_TRY(launch_spmv(c, zvec(0), w));
          PHG_TRY(launch_spmv(c, kvec(c, 0), wp));
          PHG_HIP(hipEventRecord(c->kev[1], c->stream));
          PHG_TRY(launch_bc3per(c, w, 5));
          PHG_TRY(launch_bc3per(c, wp, 5));
          PHG_TRY(launch_dot(c, kvec(c, 0), w, V, sb));
          PHG_TRY(launch_dot(c, w, w, V, sb + 1));
          PHG_TRY(launch_dot(c, kvec(c, 0), wp, V, sb + 2));
          PHG_TRY(launch_dot(c, wp, wp, V, sb + 3));
          double d[4] = {0.0, 0.0, 0.0, 0.0};
          PHG_HIP(hipMemcpyAsync(d, sb, sizeof(d), hipMemcpyDeviceToHost, c->stream));
          PHG_TRY(sync(c));
          float a1 = 0;
          PHG_HIP(hipEventElapsedTime(&a1, c->kev[0], c->kev[1]));
          st->ms_spmv += a1;
          st->n_spmv += 2;
          const double r_net = d[1] > 0.0 ? 1.0 - d[0] * d[0] / d[1] : 1.0;
          const double r_blk = d[3] > 0.0 ? 1.0 - d[2] * d[2] / d[3] : 1.0;
          if (!(r_net <= r_blk)) {
            use_net = false;
            out->rejected = true;
            PHG_HIP(hipMemcpyAsync(zvec(0), kvec(c, 0), sizeof(double) * V,
                                   hipMemcpyDeviceToDevice, c->stream));
            PHG_HIP(hipMemcpyAsync(w, wp, sizeof(double) * V, hipMemcpyDeviceToDevice,
                                   c->stream));
          }
          have_w = true;
        }
      } else {
        PHG_HIP(hipMemcpyAsync(zvec(iK - 1), kvec(c, iK - 1), sizeof(double) * V,
                               hipMemcpyDeviceToDevice, c->stream));
      }
      dir = zvec(iK - 1);
    }
    if (!have_w) {
      PHG_HIP(hipEventRecord(c->kev[0], c->stream));
      if (c->mfg.on) {
        PHG_TRY(mfg_apply(c, dir, w));  // Au1MFG: bc3Res folds inside
      } else {
        PHG_TRY(launch_spmv(c, dir, w));
      }
      PHG_HIP(hipEventRecord(c->kev[1], c->stream));
      if (!c->mfg.on) PHG_TRY(launch_bc3per(c, w, 5));
    }
    PHG_HIP(hipEventRecord(c->kev[2], c->stream));
    if (c->precision == PHG_MIXED && iK <= kCgsVectors) {
      // the mixed path orthogonalises by classical Gram-Schmidt: two passes
      // over the basis instead of iK + 1 (its error is stated against FP64)
      PHG_TRY(launch_cgs(c, w, kvec(c, 0), iK, V, H));
    } else {
      // H(1..iK+1, iK) by MGS (solgmr.f:598-618): channel j-1 holds the
      // partials of H(j, iK)
      int np0 = 0, np = 0;
      PHG_TRY(launch_dot_part(c, w, kvec(c, 0), V, 0, &np0));
      int np_prev = np0;
      for (int jK = 2; jK <= iK + 1; ++jK) {
        PHG_TRY(launch_mgs_part(c, w, kvec(c, jK - 2), jK == iK + 1 ? w : kvec(c, jK - 1), jK - 2,
                                np_prev, V, jK - 1, &np));
        np_prev = np;
      }
      PHG_TRY(launch_finish(c, 0, np0, np, iK + 1, H));
    }
    PHG_HIP(hipEventRecord(c->kev[3], c->stream));
    PHG_TRY(launch_scale(c, w, w, V, H + iK, H + iK + 1));
    PHG_TRY(download(c, c->h_scal + 16, H, iK + 2));
    PHG_TRY(sync(c));
    {
      float a1 = 0, a2 = 0;
      if (!have_w) PHG_HIP(hipEventElapsedTime(&a1, c->kev[0], c->kev[1]));
      PHG_HIP(hipEventElapsedTime(&a2, c->kev[2], c->kev[3]));
      st->ms_spmv += a1;
      st->ms_mgs += a2;
      if (!have_w) st->n_spmv += 1;
      st->n_mgs += iK + 1;
    }
    for (int jK = 1; jK <= iK; ++jK) Hm(jK, iK) = c->h_scal[16 + jK - 1];
    Hm(iK + 1, iK) = c->h_scal[16 + iK + 1];
    // Givens rotations on the new column (solgmr.f:645-663)
    for (int jK = 1; jK <= iK - 1; ++jK) {
      const double t = Rc[jK] * Hm(jK, iK) + Rs[jK] * Hm(jK + 1, iK);
      Hm(jK + 1, iK) = -Rs[jK] * Hm(jK, iK) + Rc[jK] * Hm(jK + 1, iK);
      Hm(jK, iK) = t;
    }
    const double t = std::sqrt(Hm(iK, iK) * Hm(iK, iK) + Hm(iK + 1, iK) * Hm(iK + 1, iK));
    Rc[iK] = Hm(iK, iK) / t;
    Rs[iK] = Hm(iK + 1, iK) / t;
    Hm(iK, iK) = t;
    Hm(iK + 1, iK) = 0.0;
    const double t2 = Rc[iK] * e[iK] + Rs[iK] * e[iK + 1];
    e[iK + 1] = -Rs[iK] * e[iK] + Rc[iK] * e[iK + 1];
    e[iK] = t2;
    c->ntotGM++;
    echeck = std::fabs(e[iK + 1]);
    if (echeck <= epsnrm && iK >= c->prm.minIters) break;
  }
  if (c->diag_flags & PHG_DIAG_ORTH) {
    // orthogonality of the cycle's basis u_0..u_iKs: max |u_i . u_j - delta_ij|
    double* d = c->scal + 128;
    double loss = 0.0;
    for (int i = 0; i <= iKs; ++i) {
      for (int j = 0; j <= i; ++j) PHG_TRY(launch_dot(c, kvec(c, i), kvec(c, j), V, d + j));
      PHG_TRY(download(c, c->h_scal + 128, d, i + 1));
      PHG_TRY(sync(c));
      for (int j = 0; j <= i; ++j)
        loss = std::max(loss, std::fabs(c->h_scal[128 + j] - (i == j ? 1.0 : 0.0)));
    }
    st->orth_loss = std::max(st->orth_loss, loss);
  }
  // back substitution (solgmr.f:681-686), Dy += y_k u_k (solgmr.f:690-692)
  for (int jK = iKs; jK >= 1; --jK) {
    yB[jK] = e[jK] / Hm(jK, jK);
    for (int lK = 1; lK <= jK - 1; ++lK) e[lK] = e[lK] - yB[jK] * Hm(lK, jK);
  }
  PHG_TRY(launch_combine(c, c->dy, flexible ? c->mp.zbasis : c->krylov, iKs, yB.data() + 1, V));
  out->iKs = iKs;
  out->echeck = std::fabs(e[iKs + 1]);
  return 0;
}

// SolMFG (solmfg.f:84-381) after ElmMFG.  fdi: recompute the interval
// (the reference does when iter = 1 and mod(istep, 20) = 0); otherwise the
// context's interval (phg_solmfg's eGMRES) is used.  Leaves Dy in c->dy.
int mfg_core(phg_ctx* c, int fdi, phg_stats* st) {
  const long long n = c->nshg, V = 5 * n;
  PHG_HIP(hipEventRecord(c->ev[0], c->stream));
  PHG_TRY(elm_mfg(c));
  PHG_HIP(hipEventRecord(c->ev[1], c->stream));
  double* S = c->scal;  // [0] sum res^2, [3] unorm, [4] sum rmes^2
  if (c->prm.iprec != 0) PHG_TRY(launch_i3lu(c, c->bdiag, nullptr, 0));
  PHG_TRY(launch_i3lu(c, c->bdiag, c->res, 1));
  PHG_TRY(launch_i3lu(c, c->bdiag, c->rmes, 1));
  PHG_TRY(launch_dot(c, c->res, c->res, V, S));
  PHG_TRY(launch_dot(c, c->rmes, c->rmes, V, S + 4));
  PHG_TRY(launch_fill(c, c->dy, sizeof(double) * V));
  PHG_TRY(launch_scale(c, kvec(c, 0), c->res, V, S, S + 3));
  // ypre = U new2old(y) (solmfg.f:124-133)
  PHG_TRY(launch_yshuffle(c, c->y, c->mfg.ypre, true));
  PHG_TRY(launch_i3lu(c, c->bdiag, c->mfg.ypre, 3));
  PHG_TRY(download(c, c->h_scal, S, 8));
  PHG_TRY(sync(c));
  const double sumsq = c->h_scal[0], sumsq_m = c->h_scal[4];
  double unorm = c->h_scal[3];
  st->unorm0 = unorm;
  st->ms_spmv = st->ms_mgs = 0.0;
  st->n_spmv = st->n_mgs = 0;
  st->orth_loss = 0.0;
  st->refine_cycles = 0;
  st->pc_fallback = 0;
  st->nnan = st->nlarge = 0;
  st->lGMRES = 0;
  int iKs = 0;
  double echeck = 0.0, epsnrm = 0.0;
  if (!(unorm < 100. * c->prm.epsM * c->prm.epsM)) {
    epsnrm = c->prm.etol * unorm;
    if (fdi) PHG_TRY(mfg_fdi(c, c->res, &c->mfg.eg));
    if (!(c->mfg.eg > 0.0) || !std::isfinite(c->mfg.eg)) {
      phg_set_error("SolMFG: finite-difference interval eGMRES = %g (compute it with fdi = 1)",
                    c->mfg.eg);
      return PHG_EINVAL;
    }
    c->mfg.on = true;
    int rc = 0;
    for (int mG = 1; mG <= c->prm.nGMRES && rc == 0; ++mG) {
      st->lGMRES = mG - 1;
      if (mG > 1) {  // solmfg.f:181-202: R - A Dy, normalised into kvec(0)
        rc = mfg_au2(c, kvec(c, 1));
        if (!rc) rc = launch_dot(c, kvec(c, 1), kvec(c, 1), V, S + 5);
        if (!rc) rc = launch_scale(c, kvec(c, 0), kvec(c, 1), V, S + 5, S + 6);
        if (!rc) rc = download(c, c->h_scal + 6, S + 6, 1);
        if (!rc) rc = sync(c);
        if (rc) break;
        unorm = c->h_scal[6];
      }
      CycleOut co;
      rc = gmres_cycle(c, unorm, epsnrm, st, &co, false);
      if (rc) break;
      iKs = co.iKs;
      echeck = co.echeck;
      st->refine_cycles += 1;
      if (echeck <= epsnrm) break;
    }
    c->mfg.on = false;
    if (rc) return rc;
  }
  PHG_TRY(launch_i3lu(c, c->bdiag, c->dy, 2));
  PHG_HIP(hipEventRecord(c->ev[2], c->stream));
  PHG_TRY(sync(c));
  // rstat.f:95-116 on res; solmfg.f:368 passes rstat no second array, so
  // totres[1] reports the forward-reduced modified residual
  st->totres[0] = std::sqrt(sumsq) / (double)(float)n;
  st->totres[1] = std::sqrt(sumsq_m) / (double)(float)n;
  st->iKs = iKs;
  st->ntotGM = c->ntotGM;
  st->epsnrm = epsnrm;
  st->echeck = echeck;
  st->converged = echeck <= epsnrm;
  st->reslast = c->reslast;
  float ms1 = 0, ms2 = 0;
  PHG_HIP(hipEventElapsedTime(&ms1, c->ev[0], c->ev[1]));
  PHG_HIP(hipEventElapsedTime(&ms2, c->ev[1], c->ev[2]));
  st->ms_assembly = ms1;
  st->ms_solve = ms2;
  float ma = 0, mb = 0;
  PHG_HIP(hipEventElapsedTime(&ma, c->aev[0], c->aev[1]));
  PHG_HIP(hipEventElapsedTime(&mb, c->aev[1], c->aev[2]));
  st->ms_point = ma;
  st->ms_blocks = mb;
  st->ms_total = ms1 + ms2;
  return 0;
}

// Largest number of GMRES cycles of the mixed-precision refinement loop.
constexpr int kMaxRefine = 4;

// SolGMRs on the resident y/ac (solgmr.f:449-744).  Leaves Dy in c->dy.
//
// Mixed precision: the tangent is assembled, scaled and stored in float;
// the Krylov basis, Gram-Schmidt, Givens and norms stay in double.  After
// each cycle the true residual of the scaled system, b - K z, is formed with
// the FP64 tangent (launch_jv) and its norm is tested against the same
// epsnrm as the FP64 path; cycles restart from that residual until it
// passes (at most kMaxRefine cycles).
int sol_core(phg_ctx* c, phg_stats* st) {
  const long long n = c->nshg, V = 5 * n;
  const int K = c->prm.Kspace;
  const bool mixed = c->precision == PHG_MIXED;
  if (K < 1 || K > c->kcap || K > 60) {
    phg_set_error("Kspace %d outside 1..%d", K, std::min(c->kcap, 60));
    return PHG_EINVAL;
  }
  if (c->prm.nGMRES != 1) {
    phg_set_error("nGMRES must be 1 (restart branch of solgmr.f:526 is inactive)");
    return PHG_EINVAL;
  }
  PHG_HIP(hipEventRecord(c->ev[0], c->stream));
  PHG_TRY(elm_core(c));
  PHG_HIP(hipEventRecord(c->ev[1], c->stream));
  PHG_HIP(hipMemcpyAsync(c->rmes, c->res, sizeof(double) * V, hipMemcpyDeviceToDevice, c->stream));
  double* S = c->scal;  // [0] sum res^2 (after forward), [1],[2] bad counts, [3] unorm, [4] sum rmes^2
  // tnanq on res and BDiag before factorisation (solgmr.f:459-460)
  PHG_TRY(launch_stats(c, S));  // S[0] provisional, overwritten below
  const bool set_aside = c->mp.in_step && c->mp.step_off;
  const bool learned = c->mp.ready && c->mp.use_pc > 0 && !set_aside;
  if ((learned || c->mp.keep_pair) && c->mp.ready) {
    // the network's BDiag input is the block diagonal before factorisation:
    // a call with iprec = 0 reuses the factored BDiag of an earlier call
    // (solgmr.f:471), so the raw copy of that call is kept
    if (c->prm.iprec != 0) {
      PHG_HIP(hipMemcpyAsync(c->mp.bd_raw, c->bdiag, sizeof(double) * 25 * n,
                             hipMemcpyDeviceToDevice, c->stream));
      c->mp.bd_raw_ok = true;
    } else if (!c->mp.bd_raw_ok) {
      phg_set_error("learned preconditioner with iprec = 0: no earlier call formed BDiag");
      return PHG_EINVAL;
    }
  }
  if (c->prm.iprec != 0) PHG_TRY(launch_i3lu(c, c->bdiag, nullptr, 0));
  PHG_TRY(launch_i3lu(c, c->bdiag, c->res, 1));
  PHG_TRY(launch_dot(c, c->res, c->res, V, S));
  PHG_TRY(launch_dot(c, c->rmes, c->rmes, V, S + 4));
  PHG_TRY(launch_fill(c, c->dy, sizeof(double) * V));
  if (c->prm.lhs == 1) PHG_TRY(launch_spsi3pre(c));
  PHG_TRY(launch_scale(c, kvec(c, 0), c->res, V, S, S + 3));
  PHG_TRY(download(c, c->h_scal, S, 8));
  PHG_TRY(sync(c));
  const double sumsq = c->h_scal[0];
  st->nnan = 0;
  st->nlarge = (long long)(c->h_scal[1] + c->h_scal[2]);
  if (st->nlarge > 0) {
    phg_set_error("tnanq: %lld NaN/large entries in res/BDiag", st->nlarge);
    return PHG_ENAN;
  }
  // rstatCheck.f:36-79: from the second Newton iteration of a time step on
  const double totres_chk = std::sqrt(sumsq / (double)(float)n);
  if (c->prm.iter > 1 && totres_chk > 10000.0 * c->reslast) {
    phg_set_error("rstatCheck: residual diverging (%g > 1e4 * %g)", totres_chk, c->reslast);
    return PHG_EDIVERGE;
  }
  c->reslast = totres_chk;
  const double unorm = c->h_scal[3];
  st->unorm0 = unorm;
  const double epsnrm = c->prm.etol * unorm;
  st->epsnrm = epsnrm;
  st->ms_spmv = st->ms_mgs = 0.0;
  st->n_spmv = st->n_mgs = 0;
  st->orth_loss = 0.0;
  st->refine_cycles = 0;
  int iKs = 0;
  double echeck = 0.0;
  st->pc_fallback = set_aside && c->mp.ready && c->mp.use_pc > 0 ? 4 : 0;
  if (!(unorm < 100. * c->prm.epsM * c->prm.epsM)) {
    double rnorm = unorm;
    // true residual of the scaled system, r = b - K dy, normalised into
    // kvec(0); returns its norm.  FP64 path: the stored (FP64) matrix.
    // Mixed path: the FP64 tangent without the stored float matrix,
    // r = L^-1 (rmes - A U^-1 dy) with A applied element by element in
    // double (launch_jv), so the refinement converges on the FP64 system.
    auto true_residual = [&](double* norm) -> int {
      if (mixed) {
        PHG_HIP(hipMemcpyAsync(kvec(c, 1), c->dy, sizeof(double) * V, hipMemcpyDeviceToDevice,
                               c->stream));
        PHG_TRY(launch_i3lu(c, c->bdiag, kvec(c, 1), 2));
        PHG_TRY(launch_jv(c, kvec(c, 1), kvec(c, 2)));
        PHG_TRY(launch_sub(c, kvec(c, 2), c->rmes, kvec(c, 2), V));
        PHG_TRY(launch_i3lu(c, c->bdiag, kvec(c, 2), 1));
      } else {
        PHG_TRY(launch_spmv(c, c->dy, kvec(c, 1)));
        PHG_TRY(launch_bc3per(c, kvec(c, 1), 5));
        PHG_TRY(launch_sub(c, kvec(c, 2), c->res, kvec(c, 1), V));
      }
      PHG_TRY(launch_dot(c, kvec(c, 2), kvec(c, 2), V, S + 5));
      PHG_TRY(launch_scale(c, kvec(c, 0), kvec(c, 2), V, S + 5, S + 6));
      PHG_TRY(download(c, c->h_scal + 5, S + 5, 2));
      PHG_TRY(sync(c));
      *norm = c->h_scal[6];
      return 0;
    };
    bool plain_from_start = !learned;
    if (learned) {
      // FGMRES cycle with the network; block-diagonal GMRES is its fallback and yardstick
      CycleOut co;
      PHG_TRY(gmres_cycle(c, rnorm, epsnrm, st, &co, true));
      iKs += co.iKs;
      st->refine_cycles = 1;
      PHG_TRY(true_residual(&rnorm));
      echeck = rnorm;
      if (co.rejected) {
        st->pc_fallback = 3;  // the cycle ran as block-diagonal GMRES
        if (c->mp.in_step) c->mp.step_off = true;
      } else if (!(rnorm < unorm)) {
        // no reduction: discard and solve with block-diagonal GMRES
        st->pc_fallback = 1;
        PHG_TRY(launch_fill(c, c->dy, sizeof(double) * V));
        PHG_TRY(launch_scale(c, kvec(c, 0), c->res, V, S, S + 3));
        rnorm = unorm;
        plain_from_start = true;
      } else if (rnorm > epsnrm) {
        st->pc_fallback = 2;  // reduced, finished by block-diagonal cycles
      }
    }
    const bool need_plain = plain_from_start || rnorm > epsnrm;
    for (int cyc = 0; need_plain; ++cyc) {
      CycleOut co;
      PHG_TRY(gmres_cycle(c, rnorm, epsnrm, st, &co));
      iKs += co.iKs;
      echeck = co.echeck;
      st->refine_cycles += 1;
      if (!mixed && !learned) break;  // the reference's single cycle
      PHG_TRY(true_residual(&rnorm));
      echeck = rnorm;
      if (rnorm <= epsnrm || cyc + 1 >= kMaxRefine) break;
    }
  }
  if (c->mp.ready && c->mp.keep_pair)
    PHG_HIP(hipMemcpyAsync(c->mp.zhat, c->dy, sizeof(double) * V, hipMemcpyDeviceToDevice,
                           c->stream));
  PHG_TRY(launch_i3lu(c, c->bdiag, c->dy, 2));
  PHG_HIP(hipEventRecord(c->ev[2], c->stream));
  PHG_TRY(download(c, c->h_scal + 4, S + 4, 1));
  PHG_TRY(sync(c));
  // rstat.f:95-116
  st->totres[0] = std::sqrt(sumsq) / (double)(float)n;
  st->totres[1] = std::sqrt(c->h_scal[4]) / (double)(float)n;
  st->iKs = iKs;
  st->lGMRES = 0;
  st->ntotGM = c->ntotGM;
  st->echeck = echeck;
  st->converged = echeck <= epsnrm;
  st->reslast = c->reslast;
  float ms1 = 0, ms2 = 0;
  PHG_HIP(hipEventElapsedTime(&ms1, c->ev[0], c->ev[1]));
  PHG_HIP(hipEventElapsedTime(&ms2, c->ev[1], c->ev[2]));
  st->ms_assembly = ms1;
  st->ms_solve = ms2;
  float ma = 0, mb = 0;
  PHG_HIP(hipEventElapsedTime(&ma, c->aev[0], c->aev[1]));
  PHG_HIP(hipEventElapsedTime(&mb, c->aev[1], c->aev[2]));
  st->ms_point = ma;
  st->ms_blocks = mb;
  st->ms_total = ms1 + ms2;
  return 0;
}

int check_ctx(phg_ctx* c) {
  if (!c) {
    phg_set_error("null context");
    return PHG_EINVAL;
  }
  return 0;
}

}  // namespace

int require_matrix(phg_ctx* c, const char* who) {
  if (c->matrix_free) {
    phg_set_error("%s: a PHG_FP64_MATRIX_FREE context holds no lhsK (SolMFG only)", who);
    return PHG_EINVAL;
  }
  return 0;
}

namespace {

int require_fp64(phg_ctx* c, const char* who) {
  if (c->precision != PHG_FP64) {
    phg_set_error("%s: host-array matrix entry points need a PHG_FP64 context", who);
    return PHG_EINVAL;
  }
  return 0;
}

}  // namespace

void phg_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

extern "C" const char* phg_last_error(void) { return g_err; }

extern "C" int phg_init(const phg_mesh* mesh, const phg_params* prm, int precision, phg_ctx** out) {
  if (!mesh || !prm || !out) {
    phg_set_error("phg_init: null argument");
    return PHG_EINVAL;
  }
  *out = nullptr;
  if (precision != PHG_FP64 && precision != PHG_MIXED && precision != PHG_FP64_MATRIX_FREE) {
    phg_set_error("phg_init: unknown precision %d", precision);
    return PHG_EINVAL;
  }
  if (mesh->nshg <= 0 || mesh->nel <= 0 || mesh->nnz_tot <= 0 || mesh->ndofBC < 5) {
    phg_set_error("phg_init: bad sizes");
    return PHG_EINVAL;
  }
  phg_ctx* c = new phg_ctx();
  c->nshg = mesh->nshg;
  c->nel = mesh->nel;
  c->nelb = mesh->nelb;
  c->nnz = mesh->nnz_tot;
  c->ndofBC = mesh->ndofBC;
  c->matrix_free = precision == PHG_FP64_MATRIX_FREE;
  c->precision = c->matrix_free ? PHG_FP64 : precision;
  c->prm = *prm;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    phg_set_error("hipStreamCreate: %s", hipGetErrorString(e));
    delete c;
    return PHG_EHIP;
  }
  for (auto& ev : c->ev) {
    if (hipEventCreate(&ev) != hipSuccess) ev = nullptr;
  }
  for (auto& ev : c->kev) {
    if (hipEventCreate(&ev) != hipSuccess) ev = nullptr;
  }
  for (auto& ev : c->aev) {
    if (hipEventCreate(&ev) != hipSuccess) ev = nullptr;
  }
  int rc = setup(c, mesh);
  if (rc == 0) rc = scratch_address(&c->scratch);
  if (rc == 0) {
    for (auto ev : c->ev)
      if (!ev) rc = PHG_EHIP;
    for (auto ev : c->kev)
      if (!ev) rc = PHG_EHIP;
    for (auto ev : c->aev)
      if (!ev) rc = PHG_EHIP;
    if (rc) phg_set_error("hipEventCreate failed");
  }
  if (rc) {
    std::string keep = g_err;
    phg_finalize(c);
    phg_set_error("%s", keep.c_str());
    return rc;
  }
  *out = c;
  return 0;
}

extern "C" int phg_finalize(phg_ctx* c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  mp_free(c);
  void* ptrs[] = {c->x,        c->ienc,     c->lanes,  c->lanes_diag, c->lanes_pair,
                  c->lanes_k2, c->blocks, c->qps, c->qps32, c->geo,
                  c->geo32,    c->resid,    c->tslot,    c->ienbc,  c->bcm,       c->iper,
                  c->BCv,      c->fold_master, c->fold_ptr, c->fold_slave, c->grp_slot,
                  c->scol,     c->scolp,    c->csr2slot, c->lhs64, c->lhs32,     c->res,
                  c->rmes,     c->bdiag,    c->bdiag32, c->y,      c->ac,        c->yold,
                  c->acold,    c->dy,       c->krylov, c->red_part, c->scal,
                  c->tmp,      c->slot2csr, c->lhs_aos, c->lhs_aos32, c->acc_res, c->acc_bd,
                  c->nodepack, c->inc_goff, c->inc,   c->eslot,   c->qrec,    c->qnode,
                  c->dcw,      c->mtau,     c->mfg.ypre, c->mfg.t1, c->mfg.t2,  c->mfg.pack};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_scal) (void)hipHostFree(c->h_scal);
  for (auto ev : c->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto ev : c->kev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto ev : c->aev)
    if (ev) (void)hipEventDestroy(ev);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

extern "C" int phg_set_params(phg_ctx* c, const phg_params* prm) {
  PHG_TRY(check_ctx(c));
  if (!prm || prm->Kspace > c->kcap) {
    phg_set_error("phg_set_params: Kspace may not grow beyond the value given to phg_init");
    return PHG_EINVAL;
  }
  c->prm = *prm;
  return 0;
}

extern "C" int phg_info(phg_ctx* c, int* ncolors, int* ncolors_b, long long* dev_bytes) {
  PHG_TRY(check_ctx(c));
  if (ncolors) *ncolors = 0;  // interior elements are not coloured
  if (ncolors_b) *ncolors_b = (int)c->colb_off.size() - 1;
  if (dev_bytes) *dev_bytes = c->dev_bytes;
  return 0;
}

extern "C" int phg_asiq(phg_ctx* c, const double* y, double* qres, double* rmass) {
  PHG_TRY(check_ctx(c));
  if (c->prm.idiff != 1 && c->prm.idiff != 3) {
    phg_set_error("phg_asiq: idiff must be 1 or 3");
    return PHG_EINVAL;
  }
  PHG_TRY(check_idiff(c));
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->y, y, 5 * n));
  PHG_TRY(launch_pack_nodes(c));
  PHG_TRY(launch_diffusive_flux(c));
  std::vector<double> h((size_t)16 * n);
  PHG_TRY(download(c, h.data(), c->qnode, h.size()));
  PHG_TRY(sync(c));
  for (long long i = 0; i < n; ++i) {
    for (int k = 0; k < 12; ++k) qres[i + n * k] = h[16 * i + k];
    rmass[i] = h[16 * i + 12];
  }
  return 0;
}

extern "C" int phg_elmgmrs(phg_ctx* c, const double* y, const double* ac, double* res, double* rmes,
                           double* BDiag, double* lhsK) {
  PHG_TRY(check_ctx(c));
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->y, y, 5 * n));
  PHG_TRY(upload(c, c->ac, ac, 5 * n));
  PHG_TRY(elm_core(c));
  PHG_TRY(download(c, res, c->res, 5 * n));
  if (rmes) std::memset(rmes, 0, sizeof(double) * 5 * n);  // elmgmr.f:400
  if (BDiag && c->prm.iprec != 0) PHG_TRY(download(c, BDiag, c->bdiag, 25 * n));
  if (lhsK && c->prm.lhs == 1) {
    PHG_TRY(ensure_tmp(c, sizeof(double) * 25 * c->nnz));
    PHG_TRY(launch_aos_to_ref(c, c->tmp));
    PHG_TRY(download(c, lhsK, c->tmp, 25 * c->nnz));
  }
  return sync(c);
}

extern "C" int phg_solgmrs(phg_ctx* c, const double* y, const double* ac, double* res, double* rmes,
                           double* BDiag, double* lhsK, double* Dy, phg_stats* st) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(require_matrix(c, "phg_solgmrs"));
  phg_stats local{};
  if (!st) st = &local;
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->y, y, 5 * n));
  PHG_TRY(upload(c, c->ac, ac, 5 * n));
  PHG_TRY(sol_core(c, st));
  if (res) PHG_TRY(download(c, res, c->res, 5 * n));
  if (rmes) PHG_TRY(download(c, rmes, c->rmes, 5 * n));
  if (BDiag) PHG_TRY(download(c, BDiag, c->bdiag, 25 * n));
  if (Dy) PHG_TRY(download(c, Dy, c->dy, 5 * n));
  if (lhsK) {
    PHG_TRY(ensure_tmp(c, sizeof(double) * 25 * c->nnz));
    if (c->precision == PHG_MIXED)
      PHG_TRY(launch_slots_to_csr32(c, c->lhs32, c->tmp));
    else
      PHG_TRY(launch_slots_to_csr(c, c->lhs64, c->tmp));
    PHG_TRY(download(c, lhsK, c->tmp, 25 * c->nnz));
  }
  return sync(c);
}

extern "C" int phg_state_upload(phg_ctx* c, const double* yold, const double* acold) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(upload(c, c->yold, yold, 5 * c->nshg));
  PHG_TRY(upload(c, c->acold, acold, 5 * c->nshg));
  return sync(c);
}

extern "C" int phg_state_download(phg_ctx* c, double* y, double* ac, double* Dy) {
  PHG_TRY(check_ctx(c));
  if (y) PHG_TRY(download(c, y, c->y, 5 * c->nshg));
  if (ac) PHG_TRY(download(c, ac, c->ac, 5 * c->nshg));
  if (Dy) PHG_TRY(download(c, Dy, c->dy, 5 * c->nshg));
  return sync(c);
}

extern "C" int phg_state_upload_cur(phg_ctx* c, const double* y, const double* ac) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(upload(c, c->y, y, 5 * c->nshg));
  PHG_TRY(upload(c, c->ac, ac, 5 * c->nshg));
  return sync(c);
}

extern "C" int phg_state_download_old(phg_ctx* c, double* yold, double* acold) {
  PHG_TRY(check_ctx(c));
  if (yold) PHG_TRY(download(c, yold, c->yold, 5 * c->nshg));
  if (acold) PHG_TRY(download(c, acold, c->acold, 5 * c->nshg));
  return sync(c);
}

extern "C" int phg_predict(phg_ctx* c) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(launch_predict(c));
  c->prm.iter = 0;       // itrdrv.f:409
  c->mp.in_step = true;  // a new time step: the network is tried again
  c->mp.step_off = false;
  return 0;
}

extern "C" int phg_newton_iteration(phg_ctx* c, phg_stats* st) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(require_matrix(c, "phg_newton_iteration"));
  phg_stats local{};
  if (!st) st = &local;
  c->prm.iter += 1;  // itrdrv.f:434
  PHG_TRY(sol_core(c, st));
  PHG_TRY(launch_correct(c));
  return 0;
}

extern "C" int phg_update(phg_ctx* c) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(launch_update(c));
  c->mp.in_step = false;
  c->mp.step_off = false;
  return 0;
}

// ---------------- matrix-free GMRES ----------------
extern "C" int phg_elmmfg(phg_ctx* c, const double* y, const double* ac, double* res, double* rmes,
                          double* BDiag) {
  PHG_TRY(check_ctx(c));
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->y, y, 5 * n));
  PHG_TRY(upload(c, c->ac, ac, 5 * n));
  PHG_TRY(elm_mfg(c));
  if (res) PHG_TRY(download(c, res, c->res, 5 * n));
  if (rmes) PHG_TRY(download(c, rmes, c->rmes, 5 * n));
  if (BDiag && c->prm.iprec != 0) PHG_TRY(download(c, BDiag, c->bdiag, 25 * n));
  return sync(c);
}

extern "C" int phg_itrres(phg_ctx* c, const double* yp, const double* yc, const double* ac,
                          double* rmes, int iabres) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(mfg_check(c));
  PHG_TRY(mfg_alloc(c));
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->y, yc, 5 * n));
  PHG_TRY(upload(c, c->ac, ac, 5 * n));
  PHG_TRY(launch_pack_nodes(c));
  // the state's point records (and W PTau with the matrix tau), which the
  // modified residual reads
  PHG_TRY(launch_assembly(c, false, false));
  PHG_TRY(upload(c, c->mfg.t2, yp, 5 * n));
  PHG_TRY(launch_yshuffle(c, c->mfg.t2, c->mfg.t2, true));
  PHG_TRY(launch_mfg_point(c, c->mfg.t2, nullptr, 0.0, c->mfg.pack, false, false));
  PHG_TRY(upload(c, c->mfg.t1, rmes, 5 * n));
  PHG_TRY(launch_mres(c, c->mfg.pack, c->mfg.t1, true, iabres != 0));
  PHG_TRY(download(c, rmes, c->mfg.t1, 5 * n));
  return sync(c);
}

extern "C" int phg_au1mfg(phg_ctx* c, const double* ypre, const double* y, const double* ac,
                          const double* rmes, const double* BDiag, double eGMRES, double* u) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(mfg_check(c));
  PHG_TRY(mfg_alloc(c));
  if (!(eGMRES != 0.0) || !std::isfinite(eGMRES)) {
    phg_set_error("Au1MFG: eGMRES = %g", eGMRES);
    return PHG_EINVAL;
  }
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->y, y, 5 * n));
  PHG_TRY(upload(c, c->ac, ac, 5 * n));
  PHG_TRY(launch_pack_nodes(c));
  PHG_TRY(launch_assembly(c, false, false));  // the state's point records
  PHG_TRY(upload(c, c->bdiag, BDiag, 25 * n));
  PHG_TRY(upload(c, c->mfg.ypre, ypre, 5 * n));
  PHG_TRY(upload(c, c->rmes, rmes, 5 * n));
  PHG_TRY(upload(c, c->mfg.t2, u, 5 * n));
  const double keep = c->mfg.eg;
  c->mfg.eg = eGMRES;
  const int rc = mfg_apply(c, c->mfg.t2, c->mfg.t2);
  c->mfg.eg = keep;
  PHG_TRY(rc);
  PHG_TRY(download(c, u, c->mfg.t2, 5 * n));
  return sync(c);
}

extern "C" int phg_solmfg(phg_ctx* c, const double* y, const double* ac, double* res, double* rmes,
                          double* BDiag, double* Dy, double* eGMRES, int fdi, phg_stats* st) {
  PHG_TRY(check_ctx(c));
  phg_stats local{};
  if (!st) st = &local;
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->y, y, 5 * n));
  PHG_TRY(upload(c, c->ac, ac, 5 * n));
  if (eGMRES) c->mfg.eg = *eGMRES;
  PHG_TRY(mfg_core(c, fdi, st));
  if (eGMRES) *eGMRES = c->mfg.eg;
  if (res) PHG_TRY(download(c, res, c->res, 5 * n));
  if (rmes) PHG_TRY(download(c, rmes, c->rmes, 5 * n));
  if (BDiag) PHG_TRY(download(c, BDiag, c->bdiag, 25 * n));
  if (Dy) PHG_TRY(download(c, Dy, c->dy, 5 * n));
  return sync(c);
}

extern "C" int phg_newton_iteration_mfg(phg_ctx* c, double* eGMRES, int fdi, phg_stats* st) {
  PHG_TRY(check_ctx(c));
  phg_stats local{};
  if (!st) st = &local;
  c->prm.iter += 1;  // itrdrv.f:434
  if (eGMRES) c->mfg.eg = *eGMRES;
  PHG_TRY(mfg_core(c, fdi, st));
  if (eGMRES) *eGMRES = c->mfg.eg;
  PHG_TRY(launch_correct(c));
  return 0;
}

extern "C" int phg_sparse_ap(phg_ctx* c, const double* lhsK, double* p) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(require_matrix(c, "phg_sparse_ap"));
  PHG_TRY(require_fp64(c, "phg_sparse_ap"));
  const long long n = c->nshg;
  if (lhsK) {
    PHG_TRY(ensure_tmp(c, sizeof(double) * 25 * c->nnz));
    PHG_TRY(upload(c, c->tmp, lhsK, 25 * c->nnz));
    PHG_TRY(launch_fill(c, c->lhs64, sizeof(double) * 25 * kGroup * c->nslots));
    PHG_TRY(launch_csr_to_slots(c, c->tmp, c->lhs64));
  }
  PHG_TRY(upload(c, kvec(c, 0), p, 5 * n));
  PHG_TRY(launch_spmv(c, kvec(c, 0), kvec(c, 1)));
  PHG_TRY(download(c, p, kvec(c, 1), 5 * n));
  return sync(c);
}

extern "C" int phg_tangent_ap(phg_ctx* c, const double* y, double* p) {
  PHG_TRY(check_ctx(c));
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->y, y, 5 * n));
  PHG_TRY(launch_pack_nodes(c));
  PHG_TRY(upload(c, kvec(c, 0), p, 5 * n));
  PHG_TRY(launch_jv(c, kvec(c, 0), kvec(c, 1)));
  PHG_TRY(download(c, p, kvec(c, 1), 5 * n));
  return sync(c);
}

extern "C" int phg_sparse_ap_t(phg_ctx* c, const double* lhsK, double* p) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(require_matrix(c, "phg_sparse_ap_t"));
  PHG_TRY(require_fp64(c, "phg_sparse_ap_t"));
  const long long n = c->nshg;
  if (lhsK) {
    PHG_TRY(ensure_tmp(c, sizeof(double) * 25 * c->nnz));
    PHG_TRY(upload(c, c->tmp, lhsK, 25 * c->nnz));
    PHG_TRY(launch_fill(c, c->lhs64, sizeof(double) * 25 * kGroup * c->nslots));
    PHG_TRY(launch_csr_to_slots(c, c->tmp, c->lhs64));
  }
  PHG_TRY(upload(c, kvec(c, 0), p, 5 * n));
  PHG_TRY(launch_spmv_t(c, kvec(c, 0), kvec(c, 1)));
  PHG_TRY(download(c, p, kvec(c, 1), 5 * n));
  return sync(c);
}

extern "C" int phg_i3lu(phg_ctx* c, double* Diag, double* r, const char* code) {
  PHG_TRY(check_ctx(c));
  int k = -1;
  if (code && !std::strncmp(code, "LU_Fact ", 8)) k = 0;
  if (code && !std::strncmp(code, "forward ", 8)) k = 1;
  if (code && !std::strncmp(code, "backward", 8)) k = 2;
  if (code && !std::strncmp(code, "product ", 8)) k = 3;
  if (k < 0) {
    phg_set_error("i3LU: unknown code");  // i3lu.f:175 calls error()
    return PHG_EINVAL;
  }
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->bdiag, Diag, 25 * n));
  if (k > 0) PHG_TRY(upload(c, c->res, r, 5 * n));
  PHG_TRY(launch_i3lu(c, c->bdiag, c->res, k));
  if (k == 0) PHG_TRY(download(c, Diag, c->bdiag, 25 * n));
  if (k > 0) PHG_TRY(download(c, r, c->res, 5 * n));
  return sync(c);
}

extern "C" int phg_spsi3pre(phg_ctx* c, const double* BDiag, double* lhsK) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(require_matrix(c, "phg_spsi3pre"));
  PHG_TRY(require_fp64(c, "phg_spsi3pre"));
  PHG_TRY(upload(c, c->bdiag, BDiag, 25 * c->nshg));
  PHG_TRY(ensure_tmp(c, sizeof(double) * 25 * c->nnz));
  PHG_TRY(upload(c, c->tmp, lhsK, 25 * c->nnz));
  PHG_TRY(launch_ref_to_aos(c, c->tmp));
  PHG_TRY(launch_spsi3pre(c));
  PHG_TRY(launch_slots_to_csr(c, c->lhs64, c->tmp));
  PHG_TRY(download(c, lhsK, c->tmp, 25 * c->nnz));
  return sync(c);
}

extern "C" int phg_spsi3pre_noalloc(phg_ctx* c, const double* BDiag, double* lhsK) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(require_matrix(c, "phg_spsi3pre_noalloc"));
  PHG_TRY(require_fp64(c, "phg_spsi3pre_noalloc"));
  const size_t need = sizeof(double) * 25 * (size_t)c->nnz;
  if (need > kScratchBytes) {
    phg_set_error("phg_spsi3pre_noalloc: lhsK of %lld blocks needs %zu bytes of staging, "
                  "above the %zu-byte scratch (use phg_spsi3pre)",
                  c->nnz, need, kScratchBytes);
    return PHG_EINVAL;
  }
  if (!c->scratch) {
    phg_set_error("phg_spsi3pre_noalloc: scratch address not set up");
    return PHG_EINVAL;
  }
  // the same steps as phg_spsi3pre with the staging in the device's scratch
  double* stage = static_cast<double*>(c->scratch);
  PHG_TRY(upload(c, c->bdiag, BDiag, 25 * c->nshg));
  PHG_TRY(upload(c, stage, lhsK, 25 * c->nnz));
  PHG_TRY(launch_ref_to_aos(c, stage));
  PHG_TRY(launch_spsi3pre(c));
  PHG_TRY(launch_slots_to_csr(c, c->lhs64, stage));
  PHG_TRY(download(c, lhsK, stage, 25 * c->nnz));
  return sync(c);
}

extern "C" int phg_bc3per(phg_ctx* c, double* res, int nQs) {
  PHG_TRY(check_ctx(c));
  if (nQs < 1 || nQs > 25) {
    phg_set_error("bc3per: nQs %d", nQs);
    return PHG_EINVAL;
  }
  const long long n = c->nshg;
  PHG_TRY(ensure_tmp(c, sizeof(double) * nQs * n));
  PHG_TRY(upload(c, c->tmp, res, nQs * n));
  PHG_TRY(launch_bc3per(c, c->tmp, nQs));
  PHG_TRY(download(c, res, c->tmp, nQs * n));
  return sync(c);
}

extern "C" int phg_bench_spmv(phg_ctx* c, int reps, double* ms_per_launch, double* bytes) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(require_matrix(c, "phg_bench_spmv"));
  PHG_TRY(launch_spmv(c, kvec(c, 0), kvec(c, 1)));  // warm
  PHG_HIP(hipEventRecord(c->ev[0], c->stream));
  for (int i = 0; i < reps; ++i) PHG_TRY(launch_spmv(c, kvec(c, i & 1), kvec(c, (i + 1) & 1)));
  PHG_HIP(hipEventRecord(c->ev[1], c->stream));
  PHG_HIP(hipEventSynchronize(c->ev[1]));
  float ms = 0;
  PHG_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  if (ms_per_launch) *ms_per_launch = ms / std::max(1, reps);
  // values (float or double) + column index per block, row pointers, and
  // the double vectors read and written
  const double sA = c->precision == PHG_MIXED ? 4.0 : 8.0;
  if (bytes)
    *bytes = (double)c->nnz * (25.0 * sA + 4.0) + 4.0 * (double)(c->nshg + 1) +
             2.0 * (double)c->nshg * 5.0 * 8.0;
  return 0;
}

extern "C" int phg_bench_copy(phg_ctx* c, int reps, double* ms_per_launch, double* bytes) {
  PHG_TRY(check_ctx(c));
  // first half of the Krylov space -> second half
  const long long nvec = std::max(c->kcap, 2) + 1;  // vectors allocated by setup
  const long long half = ((nvec / 2) * 5 * c->nshg) & ~1LL;
  double* src = c->krylov;
  double* dst = c->krylov + half;
  PHG_TRY(launch_copy(c, dst, src, half));  // warm
  PHG_HIP(hipEventRecord(c->ev[0], c->stream));
  for (int i = 0; i < reps; ++i) PHG_TRY(launch_copy(c, dst, src, half));
  PHG_HIP(hipEventRecord(c->ev[1], c->stream));
  PHG_HIP(hipEventSynchronize(c->ev[1]));
  float ms = 0;
  PHG_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  if (ms_per_launch) *ms_per_launch = ms / std::max(1, reps);
  if (bytes) *bytes = 2.0 * 8.0 * (double)half;
  return 0;
}

// ---------------- learned preconditioner ----------------

extern "C" int phg_mpnn_param_count(void) { return mp_param_count(); }

extern "C" int phg_mpnn_init(phg_ctx* c, const phg_mesh* mesh, const float* dscale) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(require_matrix(c, "phg_mpnn_init"));
  if (!mesh || !dscale || mesh->nshg != c->nshg || mesh->nnz_tot != c->nnz) {
    phg_set_error("phg_mpnn_init: mesh does not match the context");
    return PHG_EINVAL;
  }
  return mp_init(c, mesh, dscale);
}

static int mp_check(phg_ctx* c) {
  PHG_TRY(check_ctx(c));
  if (!c->mp.ready) {
    phg_set_error("learned preconditioner: call phg_mpnn_init first");
    return PHG_EINVAL;
  }
  return 0;
}

extern "C" int phg_mpnn_set_weights(phg_ctx* c, const float* w, int count) {
  PHG_TRY(mp_check(c));
  if (!w || count != mp_param_count()) {
    phg_set_error("phg_mpnn_set_weights: expected %d floats", mp_param_count());
    return PHG_EINVAL;
  }
  PHG_HIP(hipMemcpyAsync(c->mp.w, w, sizeof(float) * count, hipMemcpyHostToDevice, c->stream));
  PHG_TRY(mp_pack(c));
  c->mp.adam_t = 0;
  if (c->mp.train_ready) {
    PHG_TRY(launch_fill(c, c->mp.m1, sizeof(float) * count));
    PHG_TRY(launch_fill(c, c->mp.m2, sizeof(float) * count));
  }
  return sync(c);
}

extern "C" int phg_mpnn_get_weights(phg_ctx* c, float* w, int count) {
  PHG_TRY(mp_check(c));
  if (!w || count != mp_param_count()) {
    phg_set_error("phg_mpnn_get_weights: expected %d floats", mp_param_count());
    return PHG_EINVAL;
  }
  PHG_HIP(hipMemcpyAsync(w, c->mp.w, sizeof(float) * count, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

extern "C" int phg_mpnn_enable(phg_ctx* c, int use_pc, int keep_pair) {
  PHG_TRY(mp_check(c));
  c->mp.use_pc = use_pc < 0 ? 0 : use_pc;
  c->mp.step_off = false;
  c->mp.keep_pair = keep_pair != 0;
  return 0;
}

extern "C" int phg_mpnn_apply(phg_ctx* c, const double* bdiag_raw, const double* v, double* z) {
  PHG_TRY(mp_check(c));
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->mp.bd_raw, bdiag_raw, 25 * n));
  PHG_TRY(upload(c, c->mp.vin, v, 5 * n));
  PHG_TRY(mp_forward(c, c->mp.vin, c->mp.zout, false));
  PHG_TRY(download(c, z, c->mp.zout, 5 * n));
  return sync(c);
}

extern "C" int phg_mpnn_train_step(phg_ctx* c, const double* bdiag_raw, const double* v,
                                   const double* t, float lr, double* loss, float* grad) {
  PHG_TRY(mp_check(c));
  PHG_TRY(mp_alloc_train(c));
  const long long n = c->nshg;
  PHG_TRY(upload(c, c->mp.bd_raw, bdiag_raw, 25 * n));
  PHG_TRY(upload(c, c->mp.vin, v, 5 * n));
  PHG_TRY(upload(c, c->mp.zout, t, 5 * n));
  PHG_TRY(mp_forward(c, c->mp.vin, nullptr, true));
  PHG_TRY(mp_targets(c, c->mp.zout, c->mp.vin));
  PHG_TRY(mp_backward(c, lr, loss, false));
  if (grad)
    PHG_HIP(hipMemcpyAsync(grad, c->mp.g, sizeof(float) * mp_param_count(), hipMemcpyDeviceToHost,
                           c->stream));
  return sync(c);
}

extern "C" int phg_mpnn_share(phg_ctx* dst, phg_ctx* src) {
  PHG_TRY(mp_check(dst));
  PHG_TRY(mp_check(src));
  return mp_share(dst, src);
}

extern "C" int phg_mpnn_dataset(phg_ctx* c, int capacity) {
  PHG_TRY(mp_check(c));
  PHG_TRY(mp_dataset(c, capacity));
  return sync(c);
}

extern "C" int phg_mpnn_store_pair(phg_ctx* c, int slot) {
  PHG_TRY(mp_check(c));
  if (!c->mp.keep_pair) {
    phg_set_error("phg_mpnn_store_pair: enable keep_pair before the solve");
    return PHG_EINVAL;
  }
  PHG_TRY(mp_store_pair(c, slot));
  return 0;
}

extern "C" int phg_mpnn_train_pair(phg_ctx* c, int slot, float lr, double* loss) {
  PHG_TRY(mp_check(c));
  PHG_TRY(mp_alloc_train(c));
  PHG_TRY(mp_load_pair(c, slot));
  PHG_TRY(mp_forward(c, c->mp.vin, nullptr, true));
  PHG_TRY(mp_targets(c, c->mp.zout, c->mp.vin));
  PHG_TRY(mp_backward(c, lr, loss, false));
  return 0;
}

extern "C" int phg_mpnn_train_krylov(phg_ctx* c, int nvec, float lr, double* mean_loss,
                                     float* grad) {
  PHG_TRY(mp_check(c));
  if (!c->mp.keep_pair) {
    phg_set_error("phg_mpnn_train_krylov: enable keep_pair before the solve");
    return PHG_EINVAL;
  }
  if (nvec < 1 || nvec > std::max(c->kcap, 2) + 1) {
    phg_set_error("phg_mpnn_train_krylov: nvec %d outside 1..%d", nvec, std::max(c->kcap, 2) + 1);
    return PHG_EINVAL;
  }
  PHG_TRY(mp_alloc_train(c));
  PHG_TRY(launch_fill(c, c->mp.sbuf + 4, sizeof(double)));
  for (int k = 0; k < nvec; ++k) PHG_TRY(mp_train_residual(c, kvec(c, k), lr));
  if (grad)
    PHG_HIP(hipMemcpyAsync(grad, c->mp.g, sizeof(float) * mp_param_count(), hipMemcpyDeviceToHost,
                           c->stream));
  if (mean_loss) {
    double s4 = 0.0;
    PHG_HIP(hipMemcpyAsync(&s4, c->mp.sbuf + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PHG_TRY(sync(c));
    *mean_loss = s4 / nvec;
  }
  return sync(c);
}

extern "C" int phg_mpnn_last_pair(phg_ctx* c, double* bdiag_raw, double* b, double* zhat) {
  PHG_TRY(mp_check(c));
  const long long n = c->nshg;
  if (!c->mp.keep_pair) {
    phg_set_error("phg_mpnn_last_pair: enable keep_pair before the solve");
    return PHG_EINVAL;
  }
  if (bdiag_raw) PHG_TRY(download(c, bdiag_raw, c->mp.bd_raw, 25 * n));
  if (b) PHG_TRY(download(c, b, c->res, 5 * n));
  if (zhat) PHG_TRY(download(c, zhat, c->mp.zhat, 5 * n));
  return sync(c);
}

extern "C" int phg_diagnostics(phg_ctx* c, int flags) {
  PHG_TRY(check_ctx(c));
  if (flags & ~PHG_DIAG_ORTH) {
    phg_set_error("phg_diagnostics: unknown flags %d", flags);
    return PHG_EINVAL;
  }
  if ((flags & PHG_DIAG_ORTH) && c->kcap > 127) {
    phg_set_error("phg_diagnostics: Kspace %d too large for the orthogonality check", c->kcap);
    return PHG_EINVAL;
  }
  c->diag_flags = flags;
  return 0;
}

extern "C" int phg_sync(phg_ctx* c) {
  PHG_TRY(check_ctx(c));
  return sync(c);
}

extern "C" int phg_bench_assembly(phg_ctx* c, int reps, double* ms_per_launch) {
  PHG_TRY(check_ctx(c));
  PHG_TRY(elm_core(c));
  PHG_HIP(hipEventRecord(c->ev[0], c->stream));
  for (int i = 0; i < reps; ++i) PHG_TRY(elm_core(c));
  PHG_HIP(hipEventRecord(c->ev[1], c->stream));
  PHG_HIP(hipEventSynchronize(c->ev[1]));
  float ms = 0;
  PHG_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
  if (ms_per_launch) *ms_per_launch = ms / std::max(1, reps);
  return 0;
}
