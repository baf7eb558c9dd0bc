// Box-mesh fixture generator: N^3 hexes, Kuhn 6-tet split, periodic x/z,
// no-slip walls at y=0,1.  Arrays are produced in the reference's layouts
// (see phmesh.h); this replaces the reference's geombc reading and
// preprocessing (readnblk.f, genblk.f, genbc*.f, genadj.f) for synthetic
// cases only.
#include "phmesh.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Counter-based uniform in [0,1): independent of evaluation order.
inline double u01(uint64_t seed, uint64_t stream, uint64_t idx) {
  uint64_t h = mix64(seed ^ mix64(stream * 0x632BE59BD9B4E019ull + idx));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

struct Box {
  int N;
  long long np1;
  long long node(int i, int j, int k) const {
    return (long long)i + np1 * ((long long)j + np1 * (long long)k);
  }
};

// y coordinate of grid line j (geometric stretching from both walls).
double ygrid(int N, double grade, int j) {
  if (grade == 1.0) return (double)j / N;
  int half = N / 2;
  double h0 = 0.5 * (grade - 1.0) / (std::pow(grade, half) - 1.0);
  if (j <= half) return h0 * (std::pow(grade, j) - 1.0) / (grade - 1.0);
  return 1.0 - h0 * (std::pow(grade, N - j) - 1.0) / (grade - 1.0);
}

// y coordinates of the grid lines with a capped geometric boundary layer at
// both walls: cell k from a wall (k = 0 at the wall) has height
// min(d0 grade^k, dc) with d0 = first / N, and the core height dc is found by
// bisection so that each half of the N cells spans 1/2.  Returns false if
// the layer cannot fill the half (grade^k growth too slow).
bool ygrid_layer(int N, double grade, double first, std::vector<double>& yg) {
  const int half = N / 2;
  const double d0 = first / N;
  auto span = [&](double dc) {
    double s = 0.0, d = d0;
    for (int k = 0; k < half; ++k, d *= grade) s += std::min(d, dc);
    return s;
  };
  if (span(0.5) < 0.5 || half * d0 > 0.5) return false;
  double lo = d0, hi = 0.5;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    (span(mid) < 0.5 ? lo : hi) = mid;
  }
  const double dc = 0.5 * (lo + hi);
  yg.assign(N + 1, 0.0);
  double d = d0;
  for (int k = 0; k < half; ++k, d *= grade) {
    yg[k + 1] = yg[k] + std::min(d, dc);
    yg[N - k - 1] = 1.0 - yg[k + 1];
  }
  yg[half] = 0.5;
  yg[N] = 1.0;
  return true;
}

const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                         {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// Corner offsets (dx,dy,dz) of the four Kuhn-tet vertices, reference
// numbering, already oriented so that det[x1-x4,x2-x4,x3-x4] > 0.
void kuhn_tet(int t, int off[4][3]) {
  int v[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {1, 1, 1}};
  v[1][kPerm[t][0]] = 1;
  v[2][kPerm[t][0]] = 1;
  v[2][kPerm[t][1]] = 1;
  double a[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) a[r][c] = v[r][c] - v[3][c];
  double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
               a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
               a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
  if (det < 0) {
    for (int c = 0; c < 3; ++c) std::swap(v[0][c], v[1][c]);
  }
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 3; ++c) off[r][c] = v[r][c];
}

}  // namespace

extern "C" int phm_box_counts(int N, long long* nshg, long long* nel,
                              long long* nelb) {
  if (N < 2) return 1;
  long long np1 = N + 1;
  if (nshg) *nshg = np1 * np1 * np1;
  if (nel) *nel = 6LL * N * N * N;
  if (nelb) *nelb = 4LL * N * N;
  return 0;
}

extern "C" int phm_box_mesh(int N, double jitter, double grade,
                            unsigned long long seed, double* x, int* ien,
                            int* ienb, int* iBC, int* iper) {
  return phm_box_mesh_layer(N, jitter, grade, 0.0, seed, x, ien, ienb, iBC, iper);
}

extern "C" int phm_box_mesh_layer(int N, double jitter, double grade, double first,
                                  unsigned long long seed, double* x, int* ien,
                                  int* ienb, int* iBC, int* iper) {
  if (N < 2 || (grade != 1.0 && (N % 2) != 0) || grade < 1.0 || first < 0.0) return 1;
  Box b{N, N + 1LL};
  const long long nshg = b.np1 * b.np1 * b.np1;
  const long long nel = 6LL * N * N * N;
  const long long nelb = 4LL * N * N;
  std::vector<double> yg(N + 1);
  if (first > 0.0 && grade != 1.0) {
    if (!ygrid_layer(N, grade, first, yg)) return 1;
  } else {
    for (int j = 0; j <= N; ++j) yg[j] = ygrid(N, grade, j);
  }
  const double h = 1.0 / N;

#pragma omp parallel for schedule(static)
  for (int k = 0; k <= N; ++k)
    for (int j = 0; j <= N; ++j)
      for (int i = 0; i <= N; ++i) {
        long long n = b.node(i, j, k);
        double px = i * h, py = yg[j], pz = k * h;
        bool interior = i > 0 && i < N && j > 0 && j < N && k > 0 && k < N;
        if (interior && jitter > 0) {
          double hy = std::min(yg[j + 1] - yg[j], yg[j] - yg[j - 1]);
          px += jitter * h * (2.0 * u01(seed, 0, 3 * n + 0) - 1.0);
          py += jitter * hy * (2.0 * u01(seed, 0, 3 * n + 1) - 1.0);
          pz += jitter * h * (2.0 * u01(seed, 0, 3 * n + 2) - 1.0);
        }
        x[n] = px;
        x[n + nshg] = py;
        x[n + 2 * nshg] = pz;
        int code = 0;
        if (j == 0 || j == N) code |= 56;
        if (i == N || k == N) code |= 1024;
        iBC[n] = code;
        int im = (i == N) ? 0 : i, km = (k == N) ? 0 : k;
        iper[n] = (int)(b.node(im, j, km) + 1);
      }

  int off[6][4][3];
  for (int t = 0; t < 6; ++t) kuhn_tet(t, off[t]);

#pragma omp parallel for schedule(static)
  for (int k = 0; k < N; ++k)
    for (int j = 0; j < N; ++j)
      for (int i = 0; i < N; ++i) {
        long long hx = (long long)i + (long long)N * (j + (long long)N * k);
        for (int t = 0; t < 6; ++t) {
          long long e = 6 * hx + t;
          for (int a = 0; a < 4; ++a)
            ien[e + nel * a] = (int)(b.node(i + off[t][a][0], j + off[t][a][1],
                                            k + off[t][a][2]) + 1);
        }
      }

  // Wall tets in element order: face nodes first, outward orientation.
  long long nb = 0;
  for (long long e = 0; e < nel && nb < nelb; ++e) {
    long long hx = e / 6;
    int t = (int)(e % 6);
    int j = (int)((hx / N) % N);
    if (j != 0 && j != N - 1) continue;
    for (int wall = 0; wall < 2; ++wall) {
      int jw = wall == 0 ? 0 : N;
      if ((wall == 0 && j != 0) || (wall == 1 && j != N - 1)) continue;
      int face[4], nf = 0, other = -1;
      for (int a = 0; a < 4; ++a) {
        if (j + off[t][a][1] == jw)
          face[nf++] = a;
        else
          other = a;
      }
      if (nf != 3) continue;
      int nd[4];
      for (int a = 0; a < 3; ++a) nd[a] = ien[e + nel * face[a]] - 1;
      nd[3] = ien[e + nel * other] - 1;
      double v1[3], v2[3];
      for (int c = 0; c < 3; ++c) {
        v1[c] = x[nd[1] + c * nshg] - x[nd[0] + c * nshg];
        v2[c] = x[nd[2] + c * nshg] - x[nd[0] + c * nshg];
      }
      double ny = v2[0] * v1[2] - v1[0] * v2[2];
      double outward = wall == 0 ? -1.0 : 1.0;
      if (ny * outward < 0) std::swap(nd[1], nd[2]);
      for (int a = 0; a < 4; ++a) ienb[nb + nelb * a] = nd[a] + 1;
      ++nb;
    }
  }
  return nb == nelb ? 0 : 2;
}

namespace {

// node -> incident elements (0-based), counting sort in element order.
void node_elements(long long nshg, long long nel, const int* ien,
                   std::vector<long long>& ptr, std::vector<int>& lst) {
  ptr.assign(nshg + 1, 0);
  for (long long a = 0; a < 4; ++a)
    for (long long e = 0; e < nel; ++e) ptr[ien[e + nel * a]]++;
  for (long long n = 0; n < nshg; ++n) ptr[n + 1] += ptr[n];
  lst.resize(ptr[nshg]);
  std::vector<long long> pos(ptr.begin(), ptr.end() - 1);
  for (long long e = 0; e < nel; ++e)
    for (long long a = 0; a < 4; ++a) lst[pos[ien[e + nel * a] - 1]++] = (int)e;
}

template <class F>
void for_each_row(long long nshg, long long nel, const int* ien, F&& f) {
  std::vector<long long> ptr;
  std::vector<int> lst;
  node_elements(nshg, nel, ien, ptr, lst);
#pragma omp parallel
  {
    std::vector<int> nb;
#pragma omp for schedule(dynamic, 4096)
    for (long long n = 0; n < nshg; ++n) {
      nb.clear();
      for (long long p = ptr[n]; p < ptr[n + 1]; ++p)
        for (long long a = 0; a < 4; ++a) nb.push_back(ien[lst[p] + nel * a]);
      std::sort(nb.begin(), nb.end());
      nb.erase(std::unique(nb.begin(), nb.end()), nb.end());
      f(n, nb);
    }
  }
}

}  // namespace

extern "C" int phm_csr_count(long long nshg, long long nel, const int* ien,
                             int* colm, long long* nnz_tot) {
  std::vector<long long> cnt(nshg, 0);
  for_each_row(nshg, nel, ien,
               [&](long long n, const std::vector<int>& nb) { cnt[n] = (long long)nb.size(); });
  long long acc = 1;
  for (long long n = 0; n < nshg; ++n) {
    colm[n] = (int)acc;
    acc += cnt[n];
  }
  colm[nshg] = (int)acc;
  if (nnz_tot) *nnz_tot = acc - 1;
  return acc - 1 > 2147483647LL ? 1 : 0;
}

extern "C" int phm_csr_fill(long long nshg, long long nel, const int* ien,
                            const int* colm, int* rowp) {
  int bad = 0;
  for_each_row(nshg, nel, ien, [&](long long n, const std::vector<int>& nb) {
    long long b0 = colm[n] - 1;
    if ((long long)nb.size() != colm[n + 1] - colm[n]) {
#pragma omp atomic write
      bad = 1;
      return;
    }
    for (size_t q = 0; q < nb.size(); ++q) rowp[b0 + q] = nb[q];
  });
  return bad;
}

extern "C" int phm_box_state(long long nshg, const double* x, double mach,
                             unsigned long long field_seed, double Rgas,
                             double gamma, double* y) {
  const double twopi = 6.283185307179586476925286766559;
  double kw[8][3], amp[8], phs[8];
  for (int m = 0; m < 8; ++m) {
    for (int c = 0; c < 3; ++c)
      kw[m][c] = (double)(int)(3.0 * u01(field_seed, 1, 5 * m + c));
    amp[m] = (2.0 * u01(field_seed, 1, 5 * m + 3) - 1.0) / std::sqrt(8.0);
    phs[m] = twopi * u01(field_seed, 1, 5 * m + 4);
  }
  const double c0 = std::sqrt(gamma * Rgas * 300.0);
#pragma omp parallel for schedule(static)
  for (long long n = 0; n < nshg; ++n) {
    double px = x[n], py = x[n + nshg], pz = x[n + 2 * nshg];
    double s = 0.0;
    for (int m = 0; m < 8; ++m)
      s += amp[m] * std::sin(twopi * (kw[m][0] * px + kw[m][1] * py + kw[m][2] * pz) + phs[m]);
    y[n] = c0 * (mach + 0.05 * s);
    y[n + nshg] = c0 * 0.05 * s;
    y[n + 2 * nshg] = c0 * 0.05 * s;
    y[n + 3 * nshg] = 1.0e5 * (1.0 + 0.01 * s);
    y[n + 4 * nshg] = 300.0 * (1.0 + 0.01 * s);
  }
  return 0;
}
