// count_flops.cpp -- freezes the algorithmic flop count of one element's e3
// (test infrastructure, not the product).
//
// Compiles the oracle's element routines (phoracle_e3.c, which follow e3.f,
// e3ivar.f, e3mtrx.f, e3conv.f, e3visc.f, e3ls.f, e3tau.f, e3massr.f,
// e3massl.f and e3wmlt.f operation for operation) with `double` replaced by
// a number type that counts every add/subtract, multiply, divide and square
// root it performs, then runs e3 on one element with the residual only and
// with the residual and tangent.  Values built from a literal 0 (memset,
// 0.0, exact-zero A-matrix entries) are tracked as structural zeros: an
// operation with such an operand does no arithmetic work and is counted
// apart.  Prints one line per mode:
//   lhs=<0|1> add A mul M div D sqrt S work W zero Z executed E
#include <cmath>
#include <cstdio>
#include <cstring>

#include "phoracle.h"

namespace cnt {
struct Counts {
  long long add = 0, mul = 0, div = 0, sqrt = 0, zero = 0;
};
static Counts C;

struct fd {
  double v;
  unsigned char nz;  // 0: structural zero (all-zero bytes, as memset leaves it)
  fd() = default;
  fd(double x) : v(x), nz(x != 0.0) {}
  operator double() const { return v; }
};
inline fd val(double v) {
  fd r;
  r.v = v;
  r.nz = 1;
  return r;
}
inline fd szero() {
  fd r;
  r.v = 0.0;
  r.nz = 0;
  return r;
}
inline fd operator+(fd a, fd b) {
  if (!a.nz || !b.nz) {
    C.zero++;
    return a.nz ? a : b;
  }
  C.add++;
  return val(a.v + b.v);
}
inline fd operator-(fd a) {
  a.v = -a.v;
  return a;
}
inline fd operator-(fd a, fd b) {
  if (!a.nz || !b.nz) {
    C.zero++;
    return b.nz ? -b : a;
  }
  C.add++;
  return val(a.v - b.v);
}
inline fd operator*(fd a, fd b) {
  if (!a.nz || !b.nz) {
    C.zero++;
    return szero();
  }
  C.mul++;
  return val(a.v * b.v);
}
inline fd operator/(fd a, fd b) {
  if (!a.nz) {
    C.zero++;
    return szero();
  }
  C.div++;
  return val(a.v / b.v);
}
#define CNT_MIXED(op)                                             \
  inline fd operator op(fd a, double b) { return a op fd(b); } \
  inline fd operator op(double a, fd b) { return fd(a) op b; } \
  inline fd operator op(fd a, int b) { return a op fd((double)b); } \
  inline fd operator op(int a, fd b) { return fd((double)a) op b; }
CNT_MIXED(+)
CNT_MIXED(-)
CNT_MIXED(*)
CNT_MIXED(/)
#undef CNT_MIXED
inline fd sqrt(fd a) {
  C.sqrt++;
  return val(std::sqrt(a.v));
}

#define double fd
#include "phoracle_e3.c"
#undef double
}  // namespace cnt

int main() {
  using cnt::fd;
  orc_params P{};
  P.Rgas = 288.3;
  P.gamma = 1.4;
  P.gamma1 = 0.4;
  P.pr = 0.72;
  P.mu = 1.8e-5;
  P.dtsfct = P.taucfct = P.temper = 1.0;
  P.almi = P.alfi = P.gami = 1.0;
  P.Dtgl = 1e4;
  // a general tet and a state with no zero entry (data zeros would count as
  // structural ones)
  const double xs[4][3] = {{0.1, 0.02, 0.03}, {1.05, 0.04, 0.01}, {0.03, 0.97, 0.06},
                           {0.02, 0.05, 1.1}};
  fd yl[4][5], acl[4][5], xl[4][3], rl[4][5], EG[20][20];
  for (int a = 0; a < 4; ++a) {
    for (int i = 0; i < 3; ++i) xl[a][i] = fd(0.01 * xs[a][i]);
    const double y0[5] = {1e5 * (1 + 0.01 * a), 100.0 + a, 7.0 - a, 3.0 + 0.5 * a, 300.0 + a};
    for (int f = 0; f < 5; ++f) {
      yl[a][f] = fd(y0[f]);
      acl[a][f] = fd(0.1 * (f + 1) + a);
    }
  }
  (void)cnt::quad();  // one-time quadrature set-up, not per element
  for (int lhs = 0; lhs < 2; ++lhs) {
    cnt::C = cnt::Counts{};
    cnt::orc_e3(&P, yl, acl, xl, rl, EG, lhs);
    const cnt::Counts& c = cnt::C;
    const long long work = c.add + c.mul + c.div + c.sqrt;
    std::printf("lhs=%d add %lld mul %lld div %lld sqrt %lld work %lld zero %lld executed %lld\n",
                lhs, c.add, c.mul, c.div, c.sqrt, work, c.zero, work + c.zero);
  }
  return 0;
}
