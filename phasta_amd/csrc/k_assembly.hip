// Element assembly on gfx950: residual, 5x5 block tangent and block
// diagonal for linear tets (the work of ElmGMRs -> AsIGMR -> e3 and
// AsBMFG -> e3b, elmgmr.f:280-612), plus the nodal Dirichlet/periodic
// constraints on res and BDiag (bc3res.f, bc3bdg.f).
//
// Two passes, every output written exactly once:
//
//  k_asm_qp      four lanes per element, lane q = quadrature point q: the
//                element metrics, the interpolated state, thermodynamics and
//                tau at q (12 values, kept for the tangent pass) and q's
//                residual contribution to the element's four nodes, summed
//                over q across the four lanes and written per (element,
//                node).
//  k_asm_blocks  one lane per chunk of (element, a, b) contributions to one
//                block of the block-CSR matrix ("owner" lanes, lists built
//                once on the host from the element-to-block map): the lane
//                forms each 5x5 tangent block from the stored point states,
//                sums them, the chunks of one block are combined in a fixed
//                order, Dirichlet rows/columns are applied once per block,
//                and the block is stored.  The owner of a diagonal block
//                also sums the node's residual and its BDiag.
//  k_asm_pairs   the same with one lane forming both blocks (A,B) and (B,A)
//                of an off-diagonal pair from one read of the element
//                records (Shakib's or Franca's tau without DC).
//
// No read-modify-write of the matrix, no atomics, no element colouring, and
// the summation order is fixed, so the result is deterministic.
//
// The tangent block (a,b) is formed per quadrature point as
//     WdetJ [ P Q + K_ab ],   P = A(s_a) diag(tau) + N_a I,
//                             Q = A(s_b) + c N_b A0,   c = almi/(gami alfi) Dtgl
// with A(v) = sum_i v_i A_i the flux Jacobian along v, s_a = grad N_a and
// K_ab = sum_ij s_a,i K_ij s_b,j the viscous diffusivity contracted with the
// shape gradients.  This is the sum of the reference's four contributions
// (e3conv N_a A_i N_b,i; e3LS time term N_a,i A_i tau A0 N_b; e3massl
// N_a N_b A0; e3wmlt N_a,i (K_ij + A_i tau A_j) N_b,j), regrouped so that no
// 15x15 stiffness array is ever formed.
#include "phg_internal.h"

namespace {

struct Quad {
  double N[4][4];   // interior shape values N[q][a]
  double Nb[3][4];  // wall-face shape values
  double Qwt, Qwtb;
  // diffusive flux reconstruction: with M_ab = sum_q N_a N_b, Md[a] =
  // M_aa - Mo[a] and Mo[a] the mean of M_ab over b != a (the four-point
  // rule makes the off-diagonal entries equal), so that sum_b M_ab u_b =
  // Md[a] u_a + Mo[a] sum_b u_b; sN[a] = sum_q N_a; lump3[a] = sum_q N_a^2
  // * 4 / sum_q sum_b N_b^2 (special lumping, e3q.f:193-205, the equal
  // point weights cancelled)
  double Md[4], Mo[4], sN[4], lump3[4];
};

__constant__ Quad cQ;

// Geometry records of 13 values (shape gradients (4 x 3) and WdetJ), padded
// so that a record is read with 16-byte loads: 14 doubles or 16 floats.
template <typename S>
constexpr int rec_stride() { return sizeof(S) == 8 ? 14 : 16; }

// Point records of 12 values, what the tangent pass needs at a point:
//   0..2 u, 3 rho, 4 drdp, 5 drdT, 6 e1p, 7 e3p, 8 e4p (e2p = e1p + 1),
//   9..11 WdetJ tau (the tangent pass builds P scaled by WdetJ)
// read with three 16-byte loads in float (six in double).  The pass is
// bound by the vector-memory address path (one tag lookup per lane and
// load instruction), so the record is kept to the values it needs.
template <typename S>
constexpr int pt_stride() { return 12; }
enum PtField { kPtRho = 3, kPtDrdp, kPtDrdT, kPtE1p, kPtE3p, kPtE4p, kPtT0, kPtT1, kPtT2 };

template <typename S>
__device__ inline void load_pt(const S* p, S v[12]) {
  if constexpr (sizeof(S) == 8) {
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const double2 t = q[i];
      v[2 * i] = t.x;
      v[2 * i + 1] = t.y;
    }
  } else {
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float4 t = q[i];
      v[4 * i] = t.x;
      v[4 * i + 1] = t.y;
      v[4 * i + 2] = t.z;
      v[4 * i + 3] = t.w;
    }
  }
}

template <typename S>
__device__ inline void load_rec(const S* p, S v[16]) {
  if constexpr (sizeof(S) == 8) {
    const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const double2 t = q[i];
      v[2 * i] = t.x;
      v[2 * i + 1] = t.y;
    }
  } else {
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float4 t = q[i];
      v[4 * i] = t.x;
      v[4 * i + 1] = t.y;
      v[4 * i + 2] = t.z;
      v[4 * i + 3] = t.w;
    }
  }
}

struct QpArgs {
  const int* ien;          // 4 planes of nel (0-based)
  const double* node;      // 16 per node: y (p,u,v,w,T), ac (same order), x
  const double* qnode;     // 16 per node: reconstructed diffusive flux (12)
  void* qps;               // (nel*4) x 12 point states (S)
  void* geo;               // nel x 13 (S), padded to 16 floats / 14 doubles
  double* resid;           // (nel*4) x 5 residual of element e at its node a
  double* dcw;             // nel*4 discontinuity-capturing factor per point (iDC != 0)
  void* mtau;              // nel*4 x mt_stride W PTau per point (itau = 10), S
  long long nel;
  PhysParams ph;
  double epsM;             // /precis/ epsM, regularises the DC norm
  int idc;                 // iDC 1, 2, 3 (only read when the kernel has kDC)
};

__device__ inline double sel4(double v0, double v1, double v2, double v3, int i) {
  return i == 0 ? v0 : (i == 1 ? v1 : (i == 2 ? v2 : v3));
}
__device__ inline int sel4i(int v0, int v1, int v2, int v3, int i) {
  return i == 0 ? v0 : (i == 1 ? v1 : (i == 2 ? v2 : v3));
}
template <typename T>
__device__ inline T sel4t(T v0, T v1, T v2, T v3, int i) {
  return i == 0 ? v0 : (i == 1 ? v1 : (i == 2 ? v2 : v3));
}

struct Thermo {
  double u[3], rho, drdp, drdT, e1p, e2p, e3p, e4p, tau[3];
};

__device__ inline void gather_local(const double* g, long long nshg, int n, double* v) {
  // global (u,v,w,p,T) -> local (p,u,v,w,T)   (localy.f:53-73)
  v[1] = g[n];
  v[2] = g[n + nshg];
  v[3] = g[n + 2 * nshg];
  v[0] = g[n + 3 * nshg];
  v[4] = g[n + 4 * nshg];
}

// A_i v along coordinate direction i (v in local order)
__device__ inline void Ai_times(const Thermo& t, int i, const double v[5], double r[5]) {
  const double ui = t.u[i];
  const double uv = t.u[0] * v[1] + t.u[1] * v[2] + t.u[2] * v[3];
  r[0] = t.drdp * ui * v[0] + t.rho * v[1 + i] + t.drdT * ui * v[4];
#pragma unroll
  for (int m = 0; m < 3; ++m)
    r[1 + m] = (t.drdp * t.u[m] * ui + (m == i ? 1.0 : 0.0)) * v[0] +
               t.rho * (t.u[m] * v[1 + i] + ui * v[1 + m]) + t.drdT * t.u[m] * ui * v[4];
  r[4] = ui * t.e2p * v[0] + t.e3p * v[1 + i] + t.rho * ui * uv + ui * t.e4p * v[4];
}

__device__ inline void A0_times(const Thermo& t, const double a[5], double r[5]) {
  r[0] = t.drdp * a[0] + t.drdT * a[4];
#pragma unroll
  for (int m = 0; m < 3; ++m) r[1 + m] = t.drdp * t.u[m] * a[0] + t.rho * a[1 + m] + t.drdT * t.u[m] * a[4];
  r[4] = t.e1p * a[0] + t.rho * (t.u[0] * a[1] + t.u[1] * a[2] + t.u[2] * a[3]) + t.e4p * a[4];
}

// Inverse Jacobian of a linear tet from its corner coordinates: dx/dxi =
// (x_j - x_4)/2 for j = 1..3 (the shape-function derivatives are halved,
// genshp.f:34-35) inverted by cofactors (e3metric.f:22-63); det is
// det(dx/dxi).
__device__ inline void tet_inverse(const double xl[4][3], double ix[3][3], double& det) {
  double d[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) d[i][j] = 0.5 * (xl[j][i] - xl[3][i]);
  ix[0][0] = d[1][1] * d[2][2] - d[2][1] * d[1][2];
  ix[0][1] = d[2][1] * d[0][2] - d[0][1] * d[2][2];
  ix[0][2] = d[0][1] * d[1][2] - d[0][2] * d[1][1];
  det = ix[0][0] * d[0][0] + ix[0][1] * d[1][0] + ix[0][2] * d[2][0];
  const double idet = 1.0 / det;
  ix[0][0] *= idet;
  ix[0][1] *= idet;
  ix[0][2] *= idet;
  ix[1][0] = (d[1][2] * d[2][0] - d[1][0] * d[2][2]) * idet;
  ix[1][1] = (d[0][0] * d[2][2] - d[2][0] * d[0][2]) * idet;
  ix[1][2] = (d[1][0] * d[0][2] - d[0][0] * d[1][2]) * idet;
  ix[2][0] = (d[1][0] * d[2][1] - d[1][1] * d[2][0]) * idet;
  ix[2][1] = (d[2][0] * d[0][1] - d[0][0] * d[2][1]) * idet;
  ix[2][2] = (d[0][0] * d[1][1] - d[0][1] * d[1][0]) * idet;
}

// shape gradients: node j<3 -> 0.5 * row j of dxi/dx, node 3 -> -(sum)
__device__ inline void tet_shg(const double ix[3][3], double shg[4][3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    shg[0][k] = 0.5 * ix[0][k];
    shg[1][k] = 0.5 * ix[1][k];
    shg[2][k] = 0.5 * ix[2][k];
    shg[3][k] = -0.5 * (ix[0][k] + ix[1][k] + ix[2][k]);
  }
}

// metric tensor G = dxi/dx^T M dxi/dx with the tet invariance correction
// (e3tau.f:1442-1474), its trace and G:G.  G's order: 11, 12, 22, 13, 23, 33.
template <typename T>
__device__ inline void tet_metric_tensor(const T ix[3][3], T G[6], T& trG, T& GG) {
  const T c1 = (T)1.259921049894873e+00, c2 = (T)6.299605249474365e-01;
  T t[3][3];  // t[k][j] = sum_l M_kl ix[l][j], M = c2 + (c1-c2) I
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    t[0][j] = c1 * ix[0][j] + c2 * (ix[1][j] + ix[2][j]);
    t[1][j] = c1 * ix[1][j] + c2 * (ix[0][j] + ix[2][j]);
    t[2][j] = c1 * ix[2][j] + c2 * (ix[0][j] + ix[1][j]);
  }
  G[0] = ix[0][0] * t[0][0] + ix[1][0] * t[1][0] + ix[2][0] * t[2][0];
  G[1] = ix[0][0] * t[0][1] + ix[1][0] * t[1][1] + ix[2][0] * t[2][1];
  G[2] = ix[0][1] * t[0][1] + ix[1][1] * t[1][1] + ix[2][1] * t[2][1];
  G[3] = ix[0][0] * t[0][2] + ix[1][0] * t[1][2] + ix[2][0] * t[2][2];
  G[4] = ix[0][1] * t[0][2] + ix[1][1] * t[1][2] + ix[2][1] * t[2][2];
  G[5] = ix[0][2] * t[0][2] + ix[1][2] * t[1][2] + ix[2][2] * t[2][2];
  trG = G[0] + G[2] + G[5];
  GG = G[0] * G[0] + G[2] * G[2] + G[5] * G[5] + (T)2 * (G[1] * G[1] + G[3] * G[3] + G[4] * G[4]);
}

// giju, the inverse of the metric tensor G by cofactors (e3tau.f:248-268),
// in the order 11, 22, 33, 12, 13, 23
template <typename T>
__device__ inline void metric_inverse(const T G[6], T gu[6]) {
  const T m1 = G[0], m2 = G[2], m3 = G[5], m4 = G[1], m5 = G[3], m6 = G[4];
  const T detI = (T)1 / (m1 * m2 * m3 - m1 * m6 * m6 - m4 * m4 * m3 + m4 * m5 * m6 * (T)2 -
                         m5 * m5 * m2);
  gu[0] = detI * (m2 * m3 - m6 * m6);
  gu[1] = detI * (m1 * m3 - m5 * m5);
  gu[2] = detI * (m1 * m2 - m4 * m4);
  gu[3] = detI * (m5 * m6 - m4 * m3);
  gu[4] = detI * (m4 * m6 - m5 * m2);
  gu[5] = detI * (m4 * m5 - m1 * m6);
}

// giju row i as (column 0, 1, 2) indices into gu
__device__ inline int gu_idx(int i, int j) {
  return i == j ? i : (i + j == 1 ? 3 : (i + j == 2 ? 4 : 5));
}

// Thermodynamic state at a point from Y = (p, u, v, w, T) (getthm.f:106-172
// with ipress = 0; e3mtrx.f's A0/A_i entries) and the diagonal tau: Shakib's
// (itau = 0, e3tau.f:140-176) or Franca's (itau = 1, e3tau.f:54-139 with
// ipord = 1).  ei: internal energy, rk: kinetic energy.
__device__ inline void point_thermo(const double Y[5], const PhysParams& ph, const double G[6],
                                    double trG, double GG, Thermo& t, double& ei, double& rk) {
  const double cv = ph.Rgas / ph.gamma1;
  const double cp = ph.Rgas * ph.gamma / ph.gamma1;
  const double mu = ph.mu;
  const double dts = ph.dtsfct * ph.Dtgl;
  const double pres = Y[0], Tq = Y[4];
  t.u[0] = Y[1];
  t.u[1] = Y[2];
  t.u[2] = Y[3];
  rk = 0.5 * (Y[1] * Y[1] + Y[2] * Y[2] + Y[3] * Y[3]);
  t.rho = pres / (ph.Rgas * Tq);
  ei = Tq * (ph.Rgas / ph.gamma1);
  const double h = Tq * (ph.Rgas * ph.gamma / ph.gamma1);
  const double alfap = 1.0 / Tq, betaT = 1.0 / pres;
  t.drdp = t.rho * betaT;
  t.drdT = -t.rho * alfap;
  t.e1p = t.drdp * (h + rk) - alfap * Tq;
  t.e2p = t.e1p + 1.0;
  t.e3p = t.rho * (h + rk);
  t.e4p = t.drdT * (h + rk) + t.rho * cp;
  const double uGu = t.u[0] * (t.u[0] * G[0] + 2.0 * (t.u[1] * G[1] + t.u[2] * G[3])) +
                     t.u[1] * (t.u[1] * G[2] + 2.0 * t.u[2] * G[4]) + t.u[2] * t.u[2] * G[5];
  const double tt = t.rho * t.rho * ((2.0 * dts) * (2.0 * dts) + uGu) + 36.0 * (mu * mu) * GG;
  const double sq = sqrt(tt);
  t.tau[0] = 0.125 * sq / (t.rho * trG) * ph.taucfct;
  t.tau[1] = 1.0 / sq;
  t.tau[2] = t.tau[1] / cv * ph.temper;
  if (ph.itau == 1) {
    // the diffusive length |u| h1 from the inverse metric, the advective
    // h2 / (2|u|), the cell Reynolds number limiting tau_c, the time-step
    // and viscous limits of tau_m and tau_e (celt = 2/3); the reference's
    // 1.0e-15 is a single-precision literal
    const double g1 = G[0], g2 = G[1], g3 = G[2], g4 = G[3], g5 = G[4], g6 = G[5];
    const double u1 = t.u[0], u2 = t.u[1], u3 = t.u[2];
    const double det = g1 * g3 * g6 - g1 * g5 * g5 - g3 * g4 * g4 - g6 * g2 * g2 + g2 * g4 * g5 * 2.0;
    const double adj = u1 * u1 * (g3 * g6 - g5 * g5) + u2 * u2 * (g1 * g6 - g4 * g4) +
                       u3 * u3 * (g1 * g3 - g2 * g2) +
                       2.0 * (u1 * u2 * (g4 * g5 - g2 * g6) + u1 * u3 * (g2 * g5 - g4 * g3) +
                              u2 * u3 * (g4 * g2 - g1 * g5));
    const double uh1 = 2.0 * sqrt(adj / det);
    const double h2 = 1.0 / sqrt(u1 * u1 * g1 + u2 * u2 * g3 + u3 * u3 * g6 +
                                 (u1 * u2 * g2 + u1 * u3 * g4 + u2 * u3 * g5) * 2.0 +
                                 (double)1.0e-15f);
    const double re = 0.5 * t.rho * uh1 / mu;
    t.tau[0] = 0.5 * uh1 * fmin(1.0, re) * ph.taucfct;
    double tm = fmin(1.0 / (ph.Dtgl * ph.dtsfct), h2) / t.rho;
    double te = tm / cv;
    const double lim = h2 * h2 * rk * (2.0 / 3.0) / mu;
    const double con = mu * cp / ph.pr;
    t.tau[1] = fmin(tm, lim);
    t.tau[2] = fmin(te, lim * mu / con) * ph.temper;
  }
}

// Discontinuity-capturing factor at a point (iDC 1-3, ipord = 1): e3mtrx.f's
// A0DC, A0inv and dV/dY of the state (ideal gas, ipress = 0), e3tau.f's rTLS
// = L.(dV/dY tau L) over its upper triangle and raLS = L.A0inv L
// (e3tau.f:198-247), and e3dc.f's norm of the state gradient in the metric
// giju with A0DC (e3dc.f:118-260).  Lt: the least-squares residual before
// tau, TL: after it, g: the Y gradients, gu: giju.
__device__ inline double dc_factor(int idc, double epsM, const Thermo& t, double pres,
                                   double Tq, double rk, double cp, const double Lt[5],
                                   const double TL[5], const double g[3][5], const double gu[6]) {
  const double rho = t.rho, u1 = t.u[0], u2 = t.u[1], u3 = t.u[2];
  const double alfap = 1.0 / Tq, betaT = 1.0 / pres;
  const double h = Tq * cp;
  const double s1 = 1.0 / (rho * rho * betaT * Tq);
  const double cvl = cp - (alfap * alfap * Tq / rho / betaT);
  const double D0 = (rho * betaT) * (rho * betaT) * s1;
  const double D1 = -rho * alfap * rho * betaT * s1;
  const double D2 = rho / Tq;
  const double D3 = (-rho * alfap) * (-rho * alfap) * s1 + (rho * cvl / (Tq * Tq));
  // A0inv, symmetric: 11 22 33 44 55 12 13 14 15 23 24 25 34 35 45
  const double fact1 = 1.0 / (rho * cvl * (Tq * Tq));
  const double d = alfap * Tq / rho / betaT;
  const double e1bar = h - rk, e2bar = e1bar - d, e3bar = e2bar - cvl * Tq;
  const double e5bar = e1bar * e1bar - 2.0 * e1bar * d + 2.0 * rk * cvl * Tq + cp * Tq / rho / betaT;
  double Ai[15];
  Ai[0] = e5bar * fact1;
  Ai[1] = (u1 * u1 + cvl * Tq) * fact1;
  Ai[2] = (u2 * u2 + cvl * Tq) * fact1;
  Ai[3] = (u3 * u3 + cvl * Tq) * fact1;
  Ai[4] = fact1;
  Ai[5] = u1 * e3bar * fact1;
  Ai[6] = u2 * e3bar * fact1;
  Ai[7] = u3 * e3bar * fact1;
  Ai[8] = -e2bar * fact1;
  Ai[9] = u1 * u2 * fact1;
  Ai[10] = u3 * u1 * fact1;
  Ai[11] = -u1 * fact1;
  Ai[12] = u2 * u3 * fact1;
  Ai[13] = -u2 * fact1;
  Ai[14] = -u3 * fact1;
  // dV/dY, upper triangle by columns: 11 12 22 13 23 33 14 24 34 44 15 25 35 45 55
  const double f1 = 1.0 / Tq, f2 = f1 / Tq;
  const double V0 = f1 / rho, V1 = -f1 * u1, V2 = f1, V3 = -f1 * u2, V5 = f1, V6 = -f1 * u3;
  const double V9 = f1, V10 = -(h - rk) * f2, V11 = -f2 * u1, V12 = -f2 * u2, V13 = -f2 * u3,
               V14 = f2;
  const double* L = Lt;
  const double* r = TL;
  const double rTLS = L[0] * (r[0] * V0 + V1 * r[1] + V3 * r[2] + r[3] * V6 + V10 * r[4]) +
                      L[1] * (r[1] * V2 + V11 * r[4]) + L[2] * (r[2] * V5 + V12 * r[4]) +
                      L[3] * (r[3] * V9 + V13 * r[4]) + L[4] * (V14 * r[4]);
  const double raLS = 2.0 * L[3] * L[4] * Ai[14] + 2.0 * L[2] * L[4] * Ai[13] +
                      2.0 * L[0] * L[1] * Ai[5] + 2.0 * L[1] * L[2] * Ai[9] +
                      2.0 * L[1] * L[3] * Ai[10] + 2.0 * L[0] * L[2] * Ai[6] +
                      2.0 * L[2] * L[3] * Ai[12] + 2.0 * L[1] * L[4] * Ai[11] +
                      2.0 * L[0] * L[3] * Ai[7] + 2.0 * L[0] * L[4] * Ai[8] + L[0] * L[0] * Ai[0] +
                      L[1] * L[1] * Ai[1] + L[2] * L[2] * Ai[2] + L[3] * L[3] * Ai[3] +
                      L[4] * L[4] * Ai[4];
  // Y_,i A0DC Y_,j for (ij) = 11 22 33 12 13 23
  double yAy[6];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double* a = g[j];
    yAy[j] = D0 * (a[0] * a[0]) + 2.0 * a[0] * D1 * a[4] + D2 * (a[1] * a[1]) +
             D2 * (a[2] * a[2]) + D2 * (a[3] * a[3]) + D3 * (a[4] * a[4]);
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double* a = g[k == 2 ? 1 : 0];
    const double* b = g[k == 0 ? 1 : 2];
    yAy[3 + k] = a[0] * D0 * b[0] + a[0] * D1 * b[4] + a[1] * D2 * b[1] + a[2] * D2 * b[2] +
                 a[3] * D2 * b[3] + a[4] * D1 * b[0] + a[4] * D3 * b[4];
  }
  const double gnorm = 1.0 / (gu[0] * yAy[0] + 2.0 * gu[3] * yAy[3] + 2.0 * gu[4] * yAy[4] +
                              gu[1] * yAy[1] + 2.0 * gu[5] * yAy[5] + gu[2] * yAy[2] + epsM);
  if (idc == 1) return fmax(0.0, sqrt(raLS * gnorm) - rTLS * gnorm);
  if (idc == 2) return 2.0 * rTLS * gnorm;
  return fmin(fmax(0.0, sqrt(raLS * gnorm) - rTLS * gnorm), 2.0 * rTLS * gnorm);
}

// DC flux of direction i: DC sum_j giju(i,j) A0 v_,j (e3dc.f:262-300 with
// ires = 1); A0g[j] = A0 v_,j
__device__ inline void dc_flux(double DC, const double gu[6], const double A0g[3][5], int i,
                               double out[5]) {
  const double w0 = gu[gu_idx(i, 0)], w1 = gu[gu_idx(i, 1)], w2 = gu[gu_idx(i, 2)];
#pragma unroll
  for (int f = 0; f < 5; ++f) out[f] = DC * (w0 * A0g[0][f] + w1 * A0g[1][f] + w2 * A0g[2][f]);
}

// Matrix tau (itau = 10, e3tau.f:284-518 "e3tau_nd" with e3eig1.f's
// e3eig1/e3eig2 and e3visc.f:143-268's compK, ipord = 1, Navier = 1, the
// stabilisation time term kept): PTau = dY/dV Q diag(lambda) Q^T in the
// primitive variables (p,u,v,w,T).  Q are the eigenvectors of the Euler
// Jacobian in the streamline frame (entropy variables; the frame's second
// axis is the acceleration projected off the velocity), lambda^-1/2 of the
// metric-weighted eigenvalues (one Jacobi rotation diagonalises the
// acoustic pair) reduced by the mode's Peclet number.  ix is rotated into
// the streamline frame in place, as the reference rotates dxidx.
__device__ inline void matrix_tau(const PhysParams& ph, double epsM, double rho, double cp,
                                  double T, double u1, double u2, double u3, double a1,
                                  double a2, double a3, double ix[3][3], const double compK[10],
                                  double PT[5][5]) {
  const double pt39 = 0.3968502629920498686879264098181;
  const double sqt2 = 1.4142135623730950488016887242097;
  const double pt33 = 0.3333333333333333333333333333333;
  const double gamb = ph.gamma1;
  double c = sqrt((ph.gamma * ph.Rgas) * T);
  const double u = sqrt(u1 * u1 + u2 * u2 + u3 * u3);
  const double eb1 = cp * T - 0.5 * (u1 * u1 + u2 * u2 + u3 * u3);
  // e3eig1: streamline frame R, rotated metric column norms AR, eigenvectors Q
  double R[3][3], fact;
  if (u != 0.0) {
    fact = 1.0 / u;
    R[0][0] = u1 * fact;
    R[0][1] = u2 * fact;
    R[0][2] = u3 * fact;
  } else {
    R[0][0] = 1.0;
    R[0][1] = 0.0;
    R[0][2] = 0.0;
  }
  fact = a1 * R[0][0] + a2 * R[0][1] + a3 * R[0][2];
  a1 = a1 - fact * R[0][0];
  a2 = a2 - fact * R[0][1];
  a3 = a3 - fact * R[0][2];
  fact = a1 * a1 + a2 * a2 + a3 * a3;
  if (fact > epsM) {
    R[1][0] = a1;
    R[1][1] = a2;
    R[1][2] = a3;
  } else {
    const double sg = copysign(1.0, R[0][1] * R[0][2]);
    R[1][0] = R[0][1] + sg * R[0][2];
    R[1][1] = -R[0][0] - sg * R[0][2];
    R[1][2] = sg * (R[0][1] - R[0][0]);
    fact = R[1][0] * R[1][0] + R[1][1] * R[1][1] + R[1][2] * R[1][2];
  }
  fact = 1.0 / sqrt(fact);
  R[1][0] *= fact;
  R[1][1] *= fact;
  R[1][2] *= fact;
  R[2][0] = R[0][1] * R[1][2] - R[0][2] * R[1][1];
  R[2][1] = R[0][2] * R[1][0] - R[0][0] * R[1][2];
  R[2][2] = R[0][0] * R[1][1] - R[0][1] * R[1][0];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double b1 = R[0][0] * ix[k][0] + R[0][1] * ix[k][1] + R[0][2] * ix[k][2];
    const double b2 = R[1][0] * ix[k][0] + R[1][1] * ix[k][1] + R[1][2] * ix[k][2];
    const double b3 = R[2][0] * ix[k][0] + R[2][1] * ix[k][1] + R[2][2] * ix[k][2];
    ix[k][0] = b1;
    ix[k][1] = b2;
    ix[k][2] = b3;
  }
  double AR[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double sq = ix[0][j] * ix[0][j] + ix[1][j] * ix[1][j] + ix[2][j] * ix[2][j];
    const double sm = ix[0][j] + ix[1][j] + ix[2][j];
    AR[j] = (sq + sm * sm) * pt39;
  }
  double Q[5][5];
  fact = (1.0 / sqrt(2.0 * rho * T)) / c;
  Q[0][4] = fact * ((c + u) * c - eb1 * gamb);
  Q[0][0] = fact * ((c - u) * c - eb1 * gamb);
  Q[4][4] = fact * gamb;
  Q[4][0] = fact * gamb;
  Q[0][1] = Q[0][2] = Q[4][1] = Q[4][2] = 0.0;
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    Q[1 + n][4] = -fact * (c + u * gamb) * R[0][n];
    Q[1 + n][0] = fact * (c - u * gamb) * R[0][n];
    Q[1 + n][2] = -fact * c * sqt2 * R[1][n];
    Q[1 + n][1] = fact * c * sqt2 * R[2][n];
  }
  fact = (1.0 / sqrt(cp * rho)) / T;
  Q[0][3] = fact * eb1;
#pragma unroll
  for (int n = 0; n < 3; ++n) Q[1 + n][3] = fact * u * R[0][n];
  Q[4][3] = -fact;
  // e3eig2
  double lam[5];
  const double offd = (AR[1] + AR[2]) * 0.5 * (c * c);
  lam[0] = AR[0] * ((u + c) * (u + c)) + offd;
  lam[1] = AR[0] * (u * u) + AR[2] * (c * c);
  lam[2] = AR[0] * (u * u) + AR[1] * (c * c);
  lam[3] = AR[0] * (u * u);
  lam[4] = AR[0] * ((u - c) * (u - c)) + offd;
  const double tt = ph.dtsfct * 4.0 * (ph.Dtgl * ph.Dtgl);
#pragma unroll
  for (int i = 0; i < 5; ++i) lam[i] += tt;
  double t = 0.5 * (lam[0] - lam[4]) / offd;
  t = copysign(1.0, t) / (fabs(t) + sqrt(1.0 + t * t));
  const double Rc = 1.0 / sqrt(1.0 + t * t), Rs = t * Rc;
  lam[0] = lam[0] - t * offd;
  lam[4] = lam[4] + t * offd;
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    const double q0 = Rc * Q[r][0] - Rs * Q[r][4];
    Q[r][4] = Rs * Q[r][0] + Rc * Q[r][4];
    Q[r][0] = q0;
  }
  double emax = lam[0];
#pragma unroll
  for (int i = 1; i < 5; ++i) emax = fmax(emax, lam[i]);
  // lambda^-1/2, then the diffusive (Peclet) reduction of every mode
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    lam[i] = lam[i] > (epsM * epsM) * emax ? 1.0 / sqrt(lam[i]) : 0.0;
    const double q2 = Q[1][i], q3 = Q[2][i], q4 = Q[3][i], q5 = Q[4][i];
    c = pt33 * (q2 * (compK[0] * q2 + compK[1] * q3 + compK[3] * q4 + compK[6] * q5) +
                q3 * (compK[1] * q2 + compK[2] * q3 + compK[4] * q4 + compK[7] * q5) +
                q4 * (compK[3] * q2 + compK[4] * q3 + compK[5] * q4 + compK[8] * q5) +
                q5 * (compK[6] * q2 + compK[7] * q3 + compK[8] * q4 + compK[9] * q5));
    const double Pe = 1.0 / (c * lam[i]);
    lam[i] = lam[i] * fmin(pt33 * Pe, 1.0);
  }
  // STau = Q diag(lambda) Q^T (entropy variables), then dY/dV STau
  double S[5][5];
#pragma unroll
  for (int r = 0; r < 5; ++r)
#pragma unroll
    for (int cc = r; cc < 5; ++cc) {
      const double v = lam[0] * Q[r][0] * Q[cc][0] + lam[1] * Q[r][1] * Q[cc][1] +
                       lam[2] * Q[r][2] * Q[cc][2] + lam[3] * Q[r][3] * Q[cc][3] +
                       lam[4] * Q[r][4] * Q[cc][4];
      S[r][cc] = v;
      S[cc][r] = v;
    }
  const double b = cp * T + 0.5 * (u1 * u1 + u2 * u2 + u3 * u3);
  const double uu[3] = {u1, u2, u3};
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    PT[0][j] = rho * T * (S[0][j] + u1 * S[1][j] + u2 * S[2][j] + u3 * S[3][j] + b * S[4][j]);
#pragma unroll
    for (int m = 0; m < 3; ++m) PT[1 + m][j] = T * (S[1 + m][j] + uu[m] * S[4][j]);
    PT[4][j] = (T * T) * S[4][j];
  }
}

// e3visc.f:143-268 (itau >= 10, tets): compact diffusivity in parent
// coordinates from the (unrotated) inverse Jacobian ix
__device__ inline void compact_k(const double ix[3][3], double T, double mu, double lam,
                                 double kap, double u1, double u2, double u3, double K[10]) {
  const double pt39 = 0.3968502629920498686879264098181;
  double f[6];
  const int pa[6] = {0, 0, 1, 0, 1, 2}, pb[6] = {0, 1, 1, 2, 2, 2};
  double sm[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) sm[j] = ix[0][j] + ix[1][j] + ix[2][j];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int a = pa[k], bb = pb[k];
    f[k] = (ix[0][a] * ix[0][bb] + ix[1][a] * ix[1][bb] + ix[2][a] * ix[2][bb] + sm[a] * sm[bb]) *
           pt39;
  }
  const double l2m = lam + 2.0 * mu, lm = lam + mu;
  K[0] = f[0] * T * l2m + f[2] * T * mu + f[5] * T * mu;
  K[1] = f[1] * T * lm;
  K[2] = f[0] * T * mu + f[2] * T * l2m + f[5] * T * mu;
  K[3] = f[3] * T * lm;
  K[4] = f[4] * T * lm;
  K[5] = f[0] * T * mu + f[2] * T * mu + f[5] * T * l2m;
  K[6] = f[0] * T * l2m * u1 + f[1] * T * lm * u2 + f[2] * T * mu * u1 + f[3] * T * lm * u3 +
         f[5] * T * mu * u1;
  K[7] = f[0] * T * mu * u2 + f[1] * T * lm * u1 + f[2] * T * l2m * u2 + f[4] * T * lm * u3 +
         f[5] * T * mu * u2;
  K[8] = f[0] * T * mu * u3 + f[2] * T * mu * u3 + f[3] * T * lm * u1 + f[4] * T * lm * u2 +
         f[5] * T * l2m * u3;
  const double rk = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3);
  K[9] = f[0] * T * (kap * T + 2.0 * mu * rk + lm * (u1 * u1)) + f[1] * T * lm * 2.0 * u1 * u2 +
         f[2] * T * (kap * T + 2.0 * mu * rk + lm * (u2 * u2)) + f[3] * T * lm * 2.0 * u1 * u3 +
         f[4] * T * lm * 2.0 * u2 * u3 + f[5] * T * (kap * T + 2.0 * mu * rk + lm * (u3 * u3));
}

// W PTau per point for the tangent pass: 25 values padded to 16-byte loads
template <typename S>
constexpr int mt_stride() { return sizeof(S) == 8 ? 26 : 28; }

// ---------------- pass 1: point states and residual ----------------
// S: storage type of the point states and geometry (double, or float for
// the mixed-precision path whose tangent is formed in float).  kDiff: the
// divergence of the reconstructed diffusive flux is subtracted from the
// least-squares residual (idiff 1 or 3; e3ivar.f:361-395, e3ls.f:148-155).
// kDC: discontinuity capturing (iDC 1-3, e3.f:217-222): the DC flux joins
// the diffusive flux and the point's factor is kept for the tangent pass.
// kMT: the matrix tau of e3tau_nd (itau = 10) replaces Shakib's diagonal
// tau; W PTau is kept for the tangent pass.
template <typename S, bool kDiff, bool kDC, bool kMT>
__global__ void __launch_bounds__(256) k_asm_qp(QpArgs A) {
  const long long e_raw = blockIdx.x * 64LL + threadIdx.x / 4;
  const int q = threadIdx.x & 3;
  const bool valid = e_raw < A.nel;
  const long long e = valid ? e_raw : A.nel - 1;  // idle lanes load a valid element
  const long long nel = A.nel;
  int nd[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) nd[k] = A.ien[k * nel + e];
  double xl[4][3];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    // node lines are 128-byte records: 16-byte loads
    const double2* v = reinterpret_cast<const double2*>(A.node + 16 * (size_t)nd[k]);
    const double2 x01 = v[5], x2 = v[6];
    xl[k][0] = x01.x;
    xl[k][1] = x01.y;
    xl[k][2] = x2.x;
  }
  const PhysParams& ph = A.ph;

  // ---- metrics, shape gradients and the tau metric tensor ----
  double ix[3][3], det;
  tet_inverse(xl, ix, det);
  const double WdetJ = cQ.Qwt * det;
  double shg[4][3];
  tet_shg(ix, shg);
  double G[6], trG, GG;
  tet_metric_tensor(ix, G, trG, GG);
  if (valid && q == 0) {
    S* go = reinterpret_cast<S*>(A.geo) + (size_t)e * rec_stride<S>();
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int c = 0; c < 3; ++c) go[3 * k + c] = (S)shg[k][c];
    go[12] = (S)WdetJ;
  }

  // nodal y and ac streamed into the point values and the (element-
  // constant) Y gradients
  double Nq[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) Nq[k] = sel4(cQ.N[0][k], cQ.N[1][k], cQ.N[2][k], cQ.N[3][k], q);
  double g[3][5], Y[5], aci[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    Y[f] = 0.0;
    aci[f] = 0.0;
    g[0][f] = g[1][f] = g[2][f] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double2* vp = reinterpret_cast<const double2*>(A.node + 16 * (size_t)nd[k]);
    double v[10];
#pragma unroll
    for (int t = 0; t < 5; ++t) {
      const double2 u = vp[t];
      v[2 * t] = u.x;
      v[2 * t + 1] = u.y;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      const double yv = v[f], av = v[5 + f];
      Y[f] += Nq[k] * yv;
      aci[f] += Nq[k] * av;
#pragma unroll
      for (int i = 0; i < 3; ++i) g[i][f] += shg[k][i] * yv;
    }
  }

  const double cp = ph.Rgas * ph.gamma / ph.gamma1;
  const double mu = ph.mu, lam = -(2.0 / 3.0) * mu, kap = mu * cp / ph.pr;
  const double divu = g[0][1] + g[1][2] + g[2][3];
  const double pres = Y[0];
  Thermo t;
  double ei, rk;
  point_thermo(Y, ph, G, trG, GG, t, ei, rk);
  if (valid) {
    S v[12];
    v[0] = (S)t.u[0];
    v[1] = (S)t.u[1];
    v[2] = (S)t.u[2];
    v[kPtRho] = (S)t.rho;
    v[kPtDrdp] = (S)t.drdp;
    v[kPtDrdT] = (S)t.drdT;
    v[kPtE1p] = (S)t.e1p;
    v[kPtE3p] = (S)t.e3p;
    v[kPtE4p] = (S)t.e4p;
    v[kPtT0] = (S)(WdetJ * t.tau[0]);
    v[kPtT1] = (S)(WdetJ * t.tau[1]);
    v[kPtT2] = (S)(WdetJ * t.tau[2]);
    S* o = reinterpret_cast<S*>(A.qps) + ((size_t)e * 4 + q) * pt_stride<S>();
    if constexpr (sizeof(S) == 8) {
      double2* o2 = reinterpret_cast<double2*>(o);
#pragma unroll
      for (int i = 0; i < 6; ++i) o2[i] = make_double2(v[2 * i], v[2 * i + 1]);
    } else {
      float4* o4 = reinterpret_cast<float4*>(o);
#pragma unroll
      for (int i = 0; i < 3; ++i) o4[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    }
  }
  double PT[5][5], ixr[3][3];
  if constexpr (kMT) {
    // e3visc's compK from the unrotated ix, then e3tau_nd (rotates a copy)
    double K[10];
    compact_k(ix, Y[4], mu, lam, kap, t.u[0], t.u[1], t.u[2], K);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) ixr[i][j] = ix[i][j];
    matrix_tau(ph, A.epsM, t.rho, cp, Y[4], t.u[0], t.u[1], t.u[2], aci[1], aci[2], aci[3], ixr,
               K, PT);
    if (valid) {
      S* mo = reinterpret_cast<S*>(A.mtau) + ((size_t)e * 4 + q) * mt_stride<S>();
#pragma unroll
      for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) mo[5 * i + j] = (S)(WdetJ * PT[i][j]);
#pragma unroll
      for (int k = 25; k < mt_stride<S>(); ++k) mo[k] = (S)0;
    }
  }

  // time term A0 aci and the least-squares residual L = A_i Y,i + A0 Y,t
  double m0[5], L[5], tmp[5];
  A0_times(t, aci, m0);
  Ai_times(t, 0, g[0], L);
#pragma unroll
  for (int i = 1; i < 3; ++i) {
    Ai_times(t, i, g[i], tmp);
#pragma unroll
    for (int f = 0; f < 5; ++f) L[f] += tmp[f];
  }
#pragma unroll
  for (int f = 0; f < 5; ++f) L[f] += m0[f];
  if constexpr (kDiff) {
    // div q = sum_n shg_n,i q_n(i-th direction block) is constant on the
    // element: lane q takes node q's term, the four lanes are summed
    const int nq = sel4i(nd[0], nd[1], nd[2], nd[3], q);
    const double2* qv = reinterpret_cast<const double2*>(A.qnode + 16 * (size_t)nq);
    double qq[12];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double2 t2 = qv[k];
      qq[2 * k] = t2.x;
      qq[2 * k + 1] = t2.y;
    }
    const double s0 = sel4(shg[0][0], shg[1][0], shg[2][0], shg[3][0], q);
    const double s1 = sel4(shg[0][1], shg[1][1], shg[2][1], shg[3][1], q);
    const double s2 = sel4(shg[0][2], shg[1][2], shg[2][2], shg[3][2], q);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      double dq = s0 * qq[m] + s1 * qq[4 + m] + s2 * qq[8 + m];
      dq += __shfl_xor(dq, 1, 64);
      dq += __shfl_xor(dq, 2, 64);
      L[1 + m] -= dq;
    }
  }
  double DC = 0.0, gu[6], A0g[3][5];
  if constexpr (kDC) {
    double TL[5];
    if constexpr (kMT) {
      // with the matrix tau, giju comes from the inverse Jacobian rotated
      // into the point's streamline frame (e3tau.f:683-710 after e3eig1)
      double Gr[6], trGr, GGr;
      tet_metric_tensor(ixr, Gr, trGr, GGr);
      metric_inverse(Gr, gu);
#pragma unroll
      for (int i = 0; i < 5; ++i)
        TL[i] = PT[i][0] * L[0] + PT[i][1] * L[1] + PT[i][2] * L[2] + PT[i][3] * L[3] +
                PT[i][4] * L[4];
    } else {
      metric_inverse(G, gu);
      TL[0] = t.tau[0] * L[0];
      TL[1] = t.tau[1] * L[1];
      TL[2] = t.tau[1] * L[2];
      TL[3] = t.tau[1] * L[3];
      TL[4] = t.tau[2] * L[4];
    }
    DC = dc_factor(A.idc, A.epsM, t, pres, Y[4], rk, cp, L, TL, g, gu);
#pragma unroll
    for (int j = 0; j < 3; ++j) A0_times(t, g[j], A0g[j]);
    if (valid) {
      if constexpr (kMT) {
        // the point's DC giju (6 values, padded to 8) for the tangent pass
        double* d = A.dcw + ((size_t)e * 4 + q) * 8;
#pragma unroll
        for (int k = 0; k < 6; ++k) d[k] = DC * gu[k];
        d[6] = d[7] = 0.0;
      } else {
        A.dcw[(size_t)e * 4 + q] = DC;
      }
    }
  }
  if constexpr (kMT) {
    // e3tau.f:633-643: tau L with the full matrix
    double TL[5];
#pragma unroll
    for (int i = 0; i < 5; ++i)
      TL[i] = PT[i][0] * L[0] + PT[i][1] * L[1] + PT[i][2] * L[2] + PT[i][3] * L[3] +
              PT[i][4] * L[4];
#pragma unroll
    for (int i = 0; i < 5; ++i) L[i] = TL[i];
  } else {
    L[0] *= t.tau[0];
    L[1] *= t.tau[1];
    L[2] *= t.tau[1];
    L[3] *= t.tau[1];
    L[4] *= t.tau[2];
  }
  // per direction i: F_i = -F_i(conv) + F^v_i + A_i tau L (+ the DC flux)
  const double E = ei + rk;
  double F[3][5];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double mui = -t.u[i] * t.rho;
    F[i][0] = mui;
    double vf[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      vf[m] = mu * (g[i][1 + m] + g[m][1 + i]) + (m == i ? lam * divu : 0.0);
      F[i][1 + m] = mui * t.u[m] - (m == i ? pres : 0.0) + vf[m];
    }
    F[i][4] = mui * E - t.u[i] * pres +
              (t.u[0] * vf[0] + t.u[1] * vf[1] + t.u[2] * vf[2] + kap * g[i][4]);
    Ai_times(t, i, L, tmp);
#pragma unroll
    for (int f = 0; f < 5; ++f) F[i][f] += tmp[f];
    if constexpr (kDC) {
      dc_flux(DC, gu, A0g, i, tmp);
#pragma unroll
      for (int f = 0; f < 5; ++f) F[i][f] += tmp[f];
    }
  }
  // contributions of point q to node a, summed over the element's four
  // lanes (q = 0..3); lane q keeps node a = q
  double mine[5];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double r5[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      r5[f] = WdetJ * (shg[a][0] * F[0][f] + shg[a][1] * F[1][f] + shg[a][2] * F[2][f]) +
              Nq[a] * WdetJ * m0[f];
      r5[f] += __shfl_xor(r5[f], 1, 64);
      r5[f] += __shfl_xor(r5[f], 2, 64);
      if (a == q) mine[f] = r5[f];
    }
  }
  if (valid) {
    double* ro = A.resid + ((size_t)e * 4 + q) * 5;
#pragma unroll
    for (int f = 0; f < 5; ++f) ro[f] = mine[f];
  }
}

// ---------------- FP64 product with the assembled tangent ----------------
// w_a = sum_b EGmass(a,b) v_b for every element, without forming EGmass:
// with EGmass(a,b) = sum_q WdetJ [P_a Q_b] + K_ab, P_a = A(s_a) diag(tau) +
// N_a I and Q_b = A(s_b) + c N_b A0 (the blocks the assembly forms),
//     sum_b Q_b v_b = sum_i A_i (grad v)_i + c A0 v(x_q) =: L_v
// and sum_b K_ab v_b is the viscous flux of v with the velocity frozen at
// the point.  So lane q forms L_v, tau L_v, the fluxes A_i tau L_v plus the
// viscous flux, and its contribution WdetJ [sum_i s_a,i F_i + N_a L_v] to
// the four nodes, summed over the element's four lanes as the residual is.
// The point states are recomputed in double from the node lines, so the
// product is the FP64 tangent whatever precision the stored matrix has.
// bc3LHS: v's constrained components at an element node are not seen (zero
// columns) and constrained rows are zeroed here; k_jv_nodes puts the unit
// diagonals of the constrained rows back.  v is read at the periodic master
// of each node (the p(iper) gather of sparseap.f:33-35).
struct JvArgs {
  const int* ien;          // 4 planes of nel
  const double* node;      // 16 per node: y (local order), ac, x
  const double* v;         // 5 planes, local order
  const int* iper;         // 0-based master
  const uint8_t* bcm;
  double* out;             // element e's product at its node a, 5 doubles at
                           // slot eslot[4 e + a] of the node incidence layout
  const int* eslot;
  const double* dcw;       // nel*4 DC factor per point of the last assembly (kDC)
  long long nel, nshg;
  PhysParams ph;
  double cfac;             // almi/(gami alfi) Dtgl
  double epsM;             // /precis/ epsM (kMT)
};

// kDC: with discontinuity capturing the product adds the DC flux of v,
// DC sum_j giju(i,j) A0 v_,j, with the point's factor of the last assembly.
// kMT: the matrix tau of the point (itau = 10) in place of the diagonal one.
template <bool kDC, bool kMT>
__global__ void __launch_bounds__(256) k_jv(JvArgs A) {
  const long long e_raw = blockIdx.x * 64LL + threadIdx.x / 4;
  const int q = threadIdx.x & 3;
  const bool valid = e_raw < A.nel;
  const long long e = valid ? e_raw : A.nel - 1;
  const long long nel = A.nel, nshg = A.nshg;
  const PhysParams& ph = A.ph;
  int nd[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) nd[k] = A.ien[k * nel + e];
  double xl[4][3], yl[4][5], al[4][3];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double2* v = reinterpret_cast<const double2*>(A.node + 16 * (size_t)nd[k]);
    const double2 y01 = v[0], y23 = v[1], y4a = v[2], x01 = v[5], x2 = v[6];
    yl[k][0] = y01.x;
    yl[k][1] = y01.y;
    yl[k][2] = y23.x;
    yl[k][3] = y23.y;
    yl[k][4] = y4a.x;
    xl[k][0] = x01.x;
    xl[k][1] = x01.y;
    xl[k][2] = x2.x;
    if constexpr (kMT) {  // velocity accelerations: ac in local order at 5..9
      const double2 a12 = v[3], a3 = v[4];
      al[k][0] = a12.x;
      al[k][1] = a12.y;
      al[k][2] = a3.x;
    }
  }
  double ix[3][3], det;
  tet_inverse(xl, ix, det);
  const double W = cQ.Qwt * det;
  double shg[4][3];
  tet_shg(ix, shg);
  double G[6], trG, GG;
  tet_metric_tensor(ix, G, trG, GG);
  double Nq[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) Nq[k] = sel4(cQ.N[0][k], cQ.N[1][k], cQ.N[2][k], cQ.N[3][k], q);
  double Y[5], gv[3][5], vq[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) Y[f] = vq[f] = gv[0][f] = gv[1][f] = gv[2][f] = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const long long m = A.iper[nd[k]];
    const uint8_t cm = A.bcm[nd[k]];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      Y[f] += Nq[k] * yl[k][f];
      const double vv = ((cm >> f) & 1) ? 0.0 : A.v[m + f * nshg];
      vq[f] += Nq[k] * vv;
#pragma unroll
      for (int i = 0; i < 3; ++i) gv[i][f] += shg[k][i] * vv;
    }
  }
  Thermo t;
  double ei, rk;
  point_thermo(Y, ph, G, trG, GG, t, ei, rk);
  const double cp = ph.Rgas * ph.gamma / ph.gamma1;
  const double mu = ph.mu, lam = -(2.0 / 3.0) * mu, kap = mu * cp / ph.pr;
  // L_v = sum_i A_i (grad v)_i + c A0 v(x_q)
  double L[5], tmp[5], m0[5];
  A0_times(t, vq, m0);
  Ai_times(t, 0, gv[0], L);
#pragma unroll
  for (int i = 1; i < 3; ++i) {
    Ai_times(t, i, gv[i], tmp);
#pragma unroll
    for (int f = 0; f < 5; ++f) L[f] += tmp[f];
  }
#pragma unroll
  for (int f = 0; f < 5; ++f) L[f] += A.cfac * m0[f];
  double TL[5] = {t.tau[0] * L[0], t.tau[1] * L[1], t.tau[1] * L[2], t.tau[1] * L[3],
                  t.tau[2] * L[4]};
  if constexpr (kMT) {
    double acq[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int c3 = 0; c3 < 3; ++c3) acq[c3] += Nq[k] * al[k][c3];
    double K[10], ixr[3][3], PT[5][5];
    compact_k(ix, Y[4], mu, lam, kap, t.u[0], t.u[1], t.u[2], K);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) ixr[i][j] = ix[i][j];
    matrix_tau(ph, A.epsM, t.rho, cp, Y[4], t.u[0], t.u[1], t.u[2], acq[0], acq[1], acq[2], ixr, K,
               PT);
#pragma unroll
    for (int i = 0; i < 5; ++i)
      TL[i] = PT[i][0] * L[0] + PT[i][1] * L[1] + PT[i][2] * L[2] + PT[i][3] * L[3] +
              PT[i][4] * L[4];
  }
  const double divv = gv[0][1] + gv[1][2] + gv[2][3];
  double F[3][5];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Ai_times(t, i, TL, F[i]);
    double tv[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      tv[m] = mu * (gv[i][1 + m] + gv[m][1 + i]) + (m == i ? lam * divv : 0.0);
      F[i][1 + m] += tv[m];
    }
    F[i][4] += t.u[0] * tv[0] + t.u[1] * tv[1] + t.u[2] * tv[2] + kap * gv[i][4];
  }
  if constexpr (kDC) {
    double DC, gu[6], A0g[3][5], dv[5];
    if constexpr (kMT) {
      // the point's DC giju of the last assembly (DC folded in)
      const double* d = A.dcw + ((size_t)e * 4 + q) * 8;
#pragma unroll
      for (int k = 0; k < 6; ++k) gu[k] = d[k];
      DC = 1.0;
    } else {
      DC = A.dcw[(size_t)e * 4 + q];
      metric_inverse(G, gu);
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) A0_times(t, gv[j], A0g[j]);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      dc_flux(DC, gu, A0g, i, dv);
#pragma unroll
      for (int f = 0; f < 5; ++f) F[i][f] += dv[f];
    }
  }
  double mine[5];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      double r = W * (shg[a][0] * F[0][f] + shg[a][1] * F[1][f] + shg[a][2] * F[2][f]) +
                 Nq[a] * W * L[f];
      r += __shfl_xor(r, 1, 64);
      r += __shfl_xor(r, 2, 64);
      if (a == q) mine[f] = r;
    }
  }
  if (valid) {
    const int nq = sel4i(nd[0], nd[1], nd[2], nd[3], q);
    const uint8_t rm = A.bcm[nq];
    double* o = A.out + (size_t)A.eslot[e * 4 + q] * 5;
#pragma unroll
    for (int f = 0; f < 5; ++f) o[f] = ((rm >> f) & 1) ? 0.0 : mine[f];
  }
}

// w_n = sum of the node's element products in element order (the node's
// incidence list, sliced by 64-node groups; k_jv writes each product to its
// slot there, so a wavefront reads whole rows); a constrained dof f of node n
// takes the unit diagonals of its element blocks, cnt(n) v(iper(n))_f
// (bc3lhs.f:55-235: one per element containing the node).
__global__ void __launch_bounds__(256) k_jv_nodes(const long long* __restrict__ goff,
                                                  const int* __restrict__ inc,
                                                  const double* __restrict__ out,
                                                  const double* __restrict__ v,
                                                  const int* __restrict__ iper,
                                                  const uint8_t* __restrict__ bcm,
                                                  double* __restrict__ w, long long nshg) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= nshg) return;
  const long long g = n >> 6;
  const int lane = (int)(n & 63);
  const long long o0 = goff[g];
  const int deg = (int)((goff[g + 1] - o0) >> 6);
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int cnt = 0;
  for (int j = 0; j < deg; ++j) {
    const long long slot = o0 + 64LL * j + lane;
    if (inc[slot] < 0) continue;
    ++cnt;
    const double* r = out + slot * 5;
#pragma unroll
    for (int f = 0; f < 5; ++f) acc[f] += r[f];
  }
  const uint8_t m = bcm[n];
  if (m) {
    const long long mm = iper[n];
#pragma unroll
    for (int f = 0; f < 5; ++f)
      if ((m >> f) & 1) acc[f] = (double)cnt * v[mm + f * nshg];
  }
#pragma unroll
  for (int f = 0; f < 5; ++f) w[n + f * nshg] = acc[f];
}

// ---------------- pass 2: block owners ----------------
// One point's WdetJ P Q with Shakib's diagonal tau, added to ac (row + 5
// column): P = A(sP) diag(W tau) + NaP I (NaP = WdetJ N_a), Q = A(sQ) +
// c NbQ A0 (NbQ = N_b).  With the products of the state with W tau, P's
// row 0 is v + N_a e_0, v = (W t0 drdp una, W t1 rho s, W t2 drdT una);
// rows 1+k are u_k v + (W t0 s_k) e_0 + dg e_{1+k}, dg = W t1 rho una + N_a;
// row 4 is (W t0 e2p una, W t1 e3p s + (W t1 rho una) u, W t2 e4p una +
// N_a).  Q's rows 1+k are u_k Q0 + r_k e_0 + rho w e_{1+k} (Q0 its row 0,
// r = sQ, w = u.r + c N_b), so with vQ = v Q
//   (P Q)(0)   = vQ + N_a Q0
//   (P Q)(1+k) = u_k (vQ + dg Q0) + (W t0 s_k) Q0 + dg r_k e_0 + dg rho w e_{1+k}
// and row 4 through the same factored Q.
template <typename R>
__device__ __forceinline__ void shakib_pq(const R uq[3], R rho, R drdp, R drdT, R e1p, R e3p,
                                          R e4p, R t0, R wt1, R wt2, R cfacR, const R sP[3],
                                          const R sQ[3], R NaP, R NbQ, R ac[25]) {
  const R u0 = uq[0], u1 = uq[1], u2 = uq[2];
  const R e2p = e1p + (R)1;
  const R una = u0 * sP[0] + u1 * sP[1] + u2 * sP[2];
  const R unb = u0 * sQ[0] + u1 * sQ[1] + u2 * sQ[2];
  const R w = unb + cfacR * NbQ;
  const R q00 = drdp * w, q04 = drdT * w, rw = rho * w;
  R q0[3], q4[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    q0[n] = rho * sQ[n];
    q4[n] = e3p * sQ[n] + rw * uq[n];
  }
  const R q40 = e2p * unb + cfacR * NbQ * e1p, q44 = e4p * w;
  const R rt1 = wt1 * rho, pt0 = t0 * drdp, tt2 = wt2 * drdT;
  const R et0 = t0 * e2p, et1 = wt1 * e3p, et2 = wt2 * e4p;
  const R a0 = pt0 * una, r1 = rt1 * una, p04 = tt2 * una, dg = r1 + NaP;
  const R v1[3] = {rt1 * sP[0], rt1 * sP[1], rt1 * sP[2]};
  const R Q0[5] = {q00, q0[0], q0[1], q0[2], q04};
  const R sg = a0 + (v1[0] * u0 + v1[1] * u1 + v1[2] * u2);
  const R rg = v1[0] * sQ[0] + v1[1] * sQ[1] + v1[2] * sQ[2];
  R vQ[5];
  vQ[0] = sg * q00 + rg + p04 * q40;
#pragma unroll
  for (int n = 0; n < 3; ++n) vQ[1 + n] = sg * q0[n] + rw * v1[n] + p04 * q4[n];
  vQ[4] = sg * q04 + p04 * q44;
  // row 0
#pragma unroll
  for (int c = 0; c < 5; ++c) ac[5 * c] += vQ[c] + NaP * Q0[c];
  // rows 1..3
  R Z[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) Z[c] = vQ[c] + dg * Q0[c];
  const R dgrw = dg * rw;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const R ts = t0 * sP[k];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      R& ck = ac[(1 + k) + 5 * c];
      ck += uq[k] * Z[c];
      ck += ts * Q0[c];
    }
    ac[1 + k] += dg * sQ[k];
    ac[(1 + k) + 5 * (1 + k)] += dgrw;
  }
  // row 4
  const R p40 = et0 * una, p44 = et2 * una + NaP;
  R p4[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) p4[n] = et1 * sP[n] + r1 * uq[n];
  const R s4 = p40 + (p4[0] * u0 + p4[1] * u1 + p4[2] * u2);
  ac[4] += s4 * q00 + (p4[0] * sQ[0] + p4[1] * sQ[1] + p4[2] * sQ[2]) + p44 * q40;
#pragma unroll
  for (int n = 0; n < 3; ++n) ac[4 + 5 * (1 + n)] += s4 * q0[n] + rw * p4[n] + p44 * q4[n];
  ac[24] += s4 * q04 + p44 * q44;
}

// sum_q WdetJ K_ab of one element block (e3visc.f's stiffness through
// e3wmlt): the momentum rows and the conduction term do not depend on the
// point (4 equal weights), the energy row only through the velocity,
// linearly (ub = the sum of the point velocities)
template <typename R>
__device__ __forceinline__ void visc_add(R W, const R sP[3], const R sQ[3], const R ub[3],
                                         R muR, R lamR, R kapR, R ac[25]) {
  const R W4 = (R)4 * W;
  const R st = sP[0] * sQ[0] + sP[1] * sQ[1] + sP[2] * sQ[2];
  const R ubs = ub[0] * sP[0] + ub[1] * sP[1] + ub[2] * sP[2];
  const R ubr = ub[0] * sQ[0] + ub[1] * sQ[1] + ub[2] * sQ[2];
#pragma unroll
  for (int m = 0; m < 3; ++m)
#pragma unroll
    for (int n = 0; n < 3; ++n)
      ac[(1 + m) + 5 * (1 + n)] +=
          W4 * (muR * ((m == n ? st : (R)0) + sP[n] * sQ[m]) + lamR * sP[m] * sQ[n]);
#pragma unroll
  for (int n = 0; n < 3; ++n)
    ac[4 + 5 * (1 + n)] += W * (muR * (st * ub[n] + sP[n] * ubr) + lamR * ubs * sQ[n]);
  ac[24] += W4 * kapR * st;
}

// One point's WdetJ P Q with the matrix tau, added to ac (row + 5
// column): P = A(sP) (W PTau) + NaP I, Q = A(sQ) + c NbQ A0 in the factored
// form of shakib_pq (Q's rows 0 and 4 and rho w; with pu = sum_k P(m,k) u_k
// and Sm = P(m,0) + pu, row m of P Q is
//   column 0:    Sm q00 + sum_k P(m,k) r_k + P(m,4) q40
//   column 1+n:  Sm q0n + rho w P(m,1+n) + P(m,4) q4n
//   column 4:    Sm q04 + P(m,4) q44).
template <typename R>
__device__ __forceinline__ void mtau_pq(const R uq[3], R rho, R drdp, R drdT, R e1p, R e3p,
                                        R e4p, const R WPT[25], R cfacR, const R sP[3],
                                        const R sQ[3], R NaP, R NbQ, R ac[25]) {
  const R u0 = uq[0], u1 = uq[1], u2 = uq[2];
  const R e2p = e1p + (R)1;
  const R una = u0 * sP[0] + u1 * sP[1] + u2 * sP[2];
  const R unb = u0 * sQ[0] + u1 * sQ[1] + u2 * sQ[2];
  const R w = unb + cfacR * NbQ;
  const R q00 = drdp * w, q04 = drdT * w, rw = rho * w;
  R q0[3], q4[3];
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    q0[n] = rho * sQ[n];
    q4[n] = e3p * sQ[n] + rw * uq[n];
  }
  const R q40 = e2p * unb + cfacR * NbQ * e1p, q44 = e4p * w;
#pragma unroll
  for (int m = 0; m < 5; ++m) {
    // row m of A(sP), then P(m, :) = A(sP)(m, :) W PTau + N_a e_m
    R Am[5];
    if (m == 0) {
      Am[0] = drdp * una;
#pragma unroll
      for (int n = 0; n < 3; ++n) Am[1 + n] = rho * sP[n];
      Am[4] = drdT * una;
    } else if (m < 4) {
      const int mm = m - 1;
      Am[0] = drdp * uq[mm] * una + sP[mm];
#pragma unroll
      for (int n = 0; n < 3; ++n) Am[1 + n] = rho * (uq[mm] * sP[n] + (mm == n ? una : (R)0));
      Am[4] = drdT * uq[mm] * una;
    } else {
      Am[0] = e2p * una;
#pragma unroll
      for (int n = 0; n < 3; ++n) Am[1 + n] = e3p * sP[n] + rho * uq[n] * una;
      Am[4] = e4p * una;
    }
    R Pr[5];
#pragma unroll
    for (int j = 0; j < 5; ++j)
      Pr[j] = Am[0] * WPT[j] + Am[1] * WPT[5 + j] + Am[2] * WPT[10 + j] + Am[3] * WPT[15 + j] +
              Am[4] * WPT[20 + j] + (m == j ? NaP : (R)0);
    const R Sm = Pr[0] + (Pr[1] * u0 + Pr[2] * u1 + Pr[3] * u2);
    R& c0 = ac[m];
    c0 += Sm * q00;
    c0 += Pr[1] * sQ[0];
    c0 += Pr[2] * sQ[1];
    c0 += Pr[3] * sQ[2];
    c0 += Pr[4] * q40;
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      R& cn = ac[m + 5 * (1 + n)];
      cn += Sm * q0[n];
      cn += rw * Pr[1 + n];
      cn += Pr[4] * q4[n];
    }
    R& c4 = ac[m + 20];
    c4 += Sm * q04;
    c4 += Pr[4] * q44;
  }
}

// Discontinuity capturing in the tangent (e3dc.f:410-446 through e3wmlt):
// per point, sum_q dq A0(q) through A0's distinct entries (dA: drdp, drdT,
// drdp u (3), drdT u (3), rho, rho u (3), e1p, e4p), dq the point's DC
// weight; per element block, c (sum_q dq A0(q)) with c = WdetJ s_a.giju s_b
// (or WdetJ when the weight already holds the metric).
template <typename R>
__device__ __forceinline__ void dc_point(R dq, R drdp, R drdT, R rho, const R uq[3], R e1p,
                                         R e4p, R dA[14]) {
  const R dp = dq * drdp, dT = dq * drdT, dr = dq * rho;
  dA[0] += dp;
  dA[1] += dT;
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    dA[2 + n] += dp * uq[n];
    dA[5 + n] += dT * uq[n];
    dA[9 + n] += dr * uq[n];
  }
  dA[8] += dr;
  dA[12] += dq * e1p;
  dA[13] += dq * e4p;
}
template <typename R>
__device__ __forceinline__ void dc_block(R c, const R dA[14], R ac[25]) {
  ac[0] += c * dA[0];
  ac[20] += c * dA[1];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    ac[1 + m] += c * dA[2 + m];
    ac[(1 + m) + 5 * (1 + m)] += c * dA[8];
    ac[(1 + m) + 20] += c * dA[5 + m];
    ac[4 + 5 * (1 + m)] += c * dA[9 + m];
  }
  ac[4] += c * dA[12];
  ac[24] += c * dA[13];
}

struct BlkArgs {
  const int4* lanes;       // per lane: block-list offset, CSR entry, count,
                           // flags (bit 0 leader, bit 1 diagonal, bits 8..15
                           // chunks of the entry, bits 16.. total count)
  const int* lanes_k2;     // k_asm_pairs: the transposed entry of the lane's
                           // (off-diagonal) entry, -1 for diagonal lanes
  const int* blocks;       // e * 16 + a * 4 + b, grouped per lane
  const int* ien;          // 4 planes of nel
  const uint8_t* bcm;
  const void* qps;
  const void* geo;
  const double* resid;
  const double* dcw;       // nel*4 DC factor per point (kDC)
  const void* mtau;        // nel*4 x mt_stride W PTau per point (kMT), S
  double* res;             // 5 per node
  double* bdiag;           // 25 per node
  void* lhs;               // one padded block per CSR entry (T), lhsK order
  long long nel, nlanes;
  int lhs_on, prec_on;
  double mu, lam, kap, cfac;
};

// T: stored block type and arithmetic type of the per-point tangent (float
// on the mixed path); S: storage type of the point states.  Sums over
// points and elements are taken in double.  kDC: discontinuity capturing
// adds WdetJ DC (s_a . giju s_b) A0 per point (e3dc.f:410-446 through
// e3wmlt), with giju recomputed from the stored shape gradients.  kMT: the
// matrix tau (itau = 10), P = A(s_a) (W PTau) + N_a I with the stored
// W PTau (e3ls.f:477-486 Atau = A_i PTau).  kNd (with kMT, the pass
// without lhsK for the matrix-free solver): the block diagonal is e3bdg_nd's
// (e3.f:273-283), which is the tangent's but for eight of e3bdg_nd.f's
// statements that read another row of Atau_k = A_k PTau than the entry's
// (e3bdg_nd.f:539, 678, 716, 754, 902, 940, 971, 978); the lane adds their
// difference, W s_k^2 (slip - own), per point.
template <typename T, typename S, bool kDC, bool kMT, bool kNd = false>
__global__ void __launch_bounds__(64) k_asm_blocks(BlkArgs A) {
  using R = T;
  // consecutive workgroups run on different XCDs (round robin over 8): give
  // each XCD a contiguous range of lanes so its L2 sees neighbouring blocks
  const long long per = (gridDim.x + 7) / 8;
  const long long wg = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const long long l = wg * 64 + threadIdx.x;
  int4 L = make_int4(0, 0, 0, 0);
  if (l < A.nlanes) L = A.lanes[l];
  const bool diag = (L.w >> 1) & 1;
  const int cnt = (A.lhs_on || diag) ? L.z : 0;
  // block sums in the tangent's type (float accumulation on the mixed
  // path); the chunks of one block are combined in that type too
  R acc[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) acc[i] = (R)0;
  double racc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const S* qps = reinterpret_cast<const S*>(A.qps);
  const S* geo = reinterpret_cast<const S*>(A.geo);
  const R muR = (R)A.mu, lamR = (R)A.lam, kapR = (R)A.kap, cfacR = (R)A.cfac;
  for (int k = 0; k < cnt; ++k) {
    const int code = A.blocks[L.x + k];
    const long long e = code >> 4;
    const int a = (code >> 2) & 3, b = code & 3;
    // the two shape gradients and WdetJ of the element's geometry record
    // the element's geometry record with 16-byte loads, then the two rows
    S g16[16];
    load_rec<S>(geo + (size_t)e * rec_stride<S>(), g16);
    R sR[3], rR[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sR[c] = (R)sel4t(g16[c], g16[3 + c], g16[6 + c], g16[9 + c], a);
      rR[c] = (R)sel4t(g16[c], g16[3 + c], g16[6 + c], g16[9 + c], b);
    }
    const R W = (R)g16[12];
    // acc += WdetJ [P Q] per point: the stored tau is WdetJ tau and N_a is
    // scaled here, so P's rows come out multiplied by WdetJ
    R ub[3] = {(R)0, (R)0, (R)0};  // sum of the point velocities
    // kDC: sum_q DC_q A0(q) through its distinct entries: drdp, drdT,
    // drdp u (3), drdT u (3), rho, rho u (3), e1p, e4p
    R dA[14];
    if constexpr (kDC) {
#pragma unroll
      for (int i = 0; i < 14; ++i) dA[i] = (R)0;
    }
    const S* qrec = qps + (size_t)e * 4 * pt_stride<S>();
    // float: unrolled so the four points' record loads issue together
    // (2.6% at 128^3); double: the unrolled body spills
    constexpr int kPtUnroll = sizeof(S) == 4 ? 4 : 1;
#pragma unroll kPtUnroll
    for (int q = 0; q < 4; ++q) {
      S o[12];
      load_pt<S>(qrec + q * pt_stride<S>(), o);
      const R u0 = (R)o[0], u1 = (R)o[1], u2 = (R)o[2];
      const R uq[3] = {u0, u1, u2};
      ub[0] += u0;
      ub[1] += u1;
      ub[2] += u2;
      const R rho = (R)o[kPtRho], drdp = (R)o[kPtDrdp], drdT = (R)o[kPtDrdT];
      const R e1p = (R)o[kPtE1p], e2p = e1p + (R)1, e3p = (R)o[kPtE3p], e4p = (R)o[kPtE4p];
      if constexpr (kDC) {
        // weight of the point's A0: DC (the element's giju is applied after
        // the loop), or with the matrix tau s_a . (DC giju_q) s_b of the
        // point's rotated giju
        R dq;
        if constexpr (kMT) {
          const double* d = A.dcw + ((size_t)e * 4 + q) * 8;
          dq = (R)0;
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) dq += sR[i] * (R)d[gu_idx(i, j)] * rR[j];
        } else {
          dq = (R)A.dcw[(size_t)e * 4 + q];
        }
        dc_point<R>(dq, drdp, drdT, rho, uq, e1p, e4p, dA);
      }
      const R Na = W * (R)sel4(cQ.N[q][0], cQ.N[q][1], cQ.N[q][2], cQ.N[q][3], a);
      const R Nb = (R)sel4(cQ.N[q][0], cQ.N[q][1], cQ.N[q][2], cQ.N[q][3], b);
      if constexpr (kMT) {
        // rows of P = A(s_a) W PTau + N_a I from the point's W PTau
        R WPT[25];
        const S* mp = reinterpret_cast<const S*>(A.mtau) + ((size_t)e * 4 + q) * mt_stride<S>();
#pragma unroll
        for (int k = 0; k < 25; ++k) WPT[k] = (R)mp[k];
        mtau_pq<R>(uq, rho, drdp, drdT, e1p, e3p, e4p, WPT, cfacR, sR, rR, Na, Nb, acc);
        if constexpr (kNd) {
          // row m of A_k (the flux Jacobian along x_k, pressure-primitive)
          auto arow = [&](int k, int m, R r[5]) {
            const R uk = uq[k];
            if (m == 0) {
              r[0] = drdp * uk;
#pragma unroll
              for (int n = 0; n < 3; ++n) r[1 + n] = n == k ? rho : (R)0;
              r[4] = drdT * uk;
            } else if (m < 4) {
              const int mm = m - 1;
              r[0] = drdp * uq[mm] * uk + (mm == k ? (R)1 : (R)0);
#pragma unroll
              for (int n = 0; n < 3; ++n)
                r[1 + n] = rho * ((n == k ? uq[mm] : (R)0) + (mm == n ? uk : (R)0));
              r[4] = drdT * uq[mm] * uk;
            } else {
              r[0] = e2p * uk;
#pragma unroll
              for (int n = 0; n < 3; ++n) r[1 + n] = (n == k ? e3p : (R)0) + rho * uq[n] * uk;
              r[4] = e4p * uk;
            }
          };
          // W Atau_k rows 1..4: row i of A_k times W PTau
          auto wat = [&](int k, R W4[4][5]) {
#pragma unroll
            for (int i = 1; i < 5; ++i) {
              R r[5];
              arow(k, i, r);
#pragma unroll
              for (int j = 0; j < 5; ++j)
                W4[i - 1][j] = r[0] * WPT[j] + r[1] * WPT[5 + j] + r[2] * WPT[10 + j] +
                               r[3] * WPT[15 + j] + r[4] * WPT[20 + j];
            }
          };
          R Wt[4][5], col[5], r4[5];
          // block 11, entry (5,3): Atau1(3,5) for Atau1(5,5)
          wat(0, Wt);
          arow(0, 4, r4);
          acc[4 + 10] += sR[0] * sR[0] * ((Wt[1][4] - Wt[3][4]) * r4[2]);
          // block 22, entries (3,3), (4,3), (5,3): row 2 of Atau2 for m = 1, 2, 4;
          // (5,3) also Atau2(3,5) for Atau2(5,5)
          wat(1, Wt);
#pragma unroll
          for (int m = 0; m < 5; ++m) {
            R r[5];
            arow(1, m, r);
            col[m] = r[2];
          }
          const R s1 = sR[1] * sR[1];
#pragma unroll
          for (int i = 2; i < 5; ++i) {
            R d = (Wt[0][0] - Wt[i - 1][0]) * col[0] + (Wt[0][1] - Wt[i - 1][1]) * col[1] +
                  (Wt[0][3] - Wt[i - 1][3]) * col[3];
            if (i == 4) d += (Wt[1][4] - Wt[3][4]) * col[4];
            acc[i + 10] += s1 * d;
          }
          // block 33, entries (3,4), (4,4), (5,4): row 2 of Atau3 for m = 1, 2, 3;
          // entry (5,3): Atau3(3,5) for Atau3(5,5)
          wat(2, Wt);
#pragma unroll
          for (int m = 0; m < 5; ++m) {
            R r[5];
            arow(2, m, r);
            col[m] = r[3];
          }
          const R s2 = sR[2] * sR[2];
#pragma unroll
          for (int i = 2; i < 5; ++i)
            acc[i + 15] += s2 * ((Wt[0][0] - Wt[i - 1][0]) * col[0] +
                                 (Wt[0][1] - Wt[i - 1][1]) * col[1] +
                                 (Wt[0][2] - Wt[i - 1][2]) * col[2]);
          arow(2, 4, r4);
          acc[4 + 10] += s2 * ((Wt[1][4] - Wt[3][4]) * r4[2]);
        }
      } else {
        shakib_pq<R>(uq, rho, drdp, drdT, e1p, e3p, e4p, (R)o[kPtT0], (R)o[kPtT1],
                     (R)o[kPtT2], cfacR, sR, rR, Na, Nb, acc);
      }
    }
    // viscous part sum_q WdetJ K_ab
    visc_add<R>(W, sR, rR, ub, muR, lamR, kapR, acc);
    if constexpr (kDC) {
      R sab = (R)1;
      if constexpr (!kMT) {
        // giju of the element from dxi/dx = 2 (shape gradients of nodes 1..3)
        R ix[3][3], G[6], trG, GG, gu[6];
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
          for (int c = 0; c < 3; ++c) ix[k][c] = (R)2 * (R)g16[3 * k + c];
        tet_metric_tensor(ix, G, trG, GG);
        metric_inverse(G, gu);
        sab = (R)0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) sab += sR[i] * gu[gu_idx(i, j)] * rR[j];
      }
      dc_block<R>(W * sab, dA, acc);
    }
  }
  // the node's residual: element residuals of the diagonal blocks' elements
  if (diag) {
    for (int k = 0; k < cnt; ++k) {
      const int code = A.blocks[L.x + k];
      const double* ro = A.resid + ((size_t)(code >> 4) * 4 + ((code >> 2) & 3)) * 5;
#pragma unroll
      for (int f = 0; f < 5; ++f) racc[f] += ro[f];
    }
  }
  // combine the chunks of one block in lane order: the leader adds the
  // partials of the next nch - 1 lanes (the entry's chunks are consecutive
  // lanes of the wavefront) through cross-lane reads, so the pass holds no
  // LDS and its occupancy is set by its registers
  const bool leader = (L.w & 1) && l < A.nlanes;
  const int nch = (L.w >> 8) & 0xff;
  const int total = L.w >> 16;
  int mx = nch;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = max(mx, __shfl_xor(mx, o, 64));
  // the chunk sums stay in the tangent's type (a few float additions more
  // on the mixed path), which keeps the pass under 64 VGPRs in float
  R csum[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) csum[i] = acc[i];
  for (int c = 1; c < mx; ++c) {
#pragma unroll
    for (int i = 0; i < 25; ++i) {
      const R v = __shfl_down(acc[i], c, 64);
      if (c < nch) csum[i] += v;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      const double v = __shfl_down(racc[f], c, 64);
      if (c < nch) racc[f] += v;
    }
  }
  if (!leader) return;
  double sum[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) sum[i] = (double)csum[i];
  const int code0 = A.blocks[L.x];
  const long long e0 = code0 >> 4;
  const int a0 = (code0 >> 2) & 3, b0 = code0 & 3;
  const long long nel = A.nel;
  const int v = A.ien[a0 * nel + e0], wcol = A.ien[b0 * nel + e0];
  if (diag) {
#pragma unroll
    for (int f = 0; f < 5; ++f) A.res[5 * (size_t)v + f] = racc[f];
    // block diagonal before the Dirichlet rows/columns (asigmr.f:92-103)
    if (A.prec_on) {
#pragma unroll
      for (int i = 0; i < 25; ++i) A.bdiag[25 * (size_t)v + i] = sum[i];
    }
  }
  if (!A.lhs_on) return;
  // bc3LHS on every element block: constrained rows of node a and columns of
  // node b zeroed, a unit diagonal per element block of the node
  const uint8_t ma = A.bcm[v], mb = A.bcm[wcol];
  if (ma | mb) {
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
      for (int i = 0; i < 5; ++i) {
        if (((ma >> i) & 1) || ((mb >> j) & 1)) sum[i + 5 * j] = 0.0;
        if (diag && i == j && ((ma >> i) & 1)) sum[i + 5 * j] = (double)total;
      }
  }
  T* dst = reinterpret_cast<T*>(A.lhs) + (size_t)aos_stride<T>() * L.y;
  if constexpr (sizeof(T) == 4) {
    float4* v = reinterpret_cast<float4*>(dst);
#pragma unroll
    for (int k = 0; k < 6; ++k)
      v[k] = make_float4((float)sum[4 * k], (float)sum[4 * k + 1], (float)sum[4 * k + 2],
                         (float)sum[4 * k + 3]);
    v[6] = make_float4((float)sum[24], 0.f, 0.f, 0.f);
  } else {
    double2* v = reinterpret_cast<double2*>(dst);
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = make_double2(sum[2 * k], sum[2 * k + 1]);
    v[12] = make_double2(sum[24], 0.0);
  }
}

// The block pass with lhsK: the entries (A,B) and (B,A) come from the same
// element blocks with a and b exchanged, so one lane forms both from one
// read of the point records (the pass is bound by those per-lane reads).
// Diagonal entries run in wavefronts of their own, so every lane of a
// wavefront forms the same number of blocks.  Same arithmetic per block as
// k_asm_blocks; with DC (kDC) the point weights and s_a.giju s_b are
// symmetric in a and b and serve both blocks; the matrix tau (kMT) through
// mtau_pq.  Used for every flag set in float, for Shakib's or Franca's tau
// without DC in double (the other double variants exceed the registers).
template <typename T, typename S, bool kDC, bool kMT>
__device__ __forceinline__ void asm_pairs(const BlkArgs& A) {
  using R = T;
  const long long per = (gridDim.x + 7) / 8;
  const long long wg = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const long long l = wg * 64 + threadIdx.x;
  int4 L = make_int4(0, 0, 0, 0);
  int k2 = -1;
  if (l < A.nlanes) {
    L = A.lanes[l];
    k2 = A.lanes_k2[l];
  }
  const bool diag = (L.w >> 1) & 1;
  const bool pair = k2 >= 0;
  const int cnt = L.z;
  R acc[25], acc2[25];
#pragma unroll
  for (int i = 0; i < 25; ++i) acc[i] = acc2[i] = (R)0;
  double racc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  const S* qps = reinterpret_cast<const S*>(A.qps);
  const S* geo = reinterpret_cast<const S*>(A.geo);
  const R muR = (R)A.mu, lamR = (R)A.lam, kapR = (R)A.kap, cfacR = (R)A.cfac;
  for (int k = 0; k < cnt; ++k) {
    const int code = A.blocks[L.x + k];
    const long long e = code >> 4;
    const int a = (code >> 2) & 3, b = code & 3;
    S g16[16];
    load_rec<S>(geo + (size_t)e * rec_stride<S>(), g16);
    R sR[3], rR[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sR[c] = (R)sel4t(g16[c], g16[3 + c], g16[6 + c], g16[9 + c], a);
      rR[c] = (R)sel4t(g16[c], g16[3 + c], g16[6 + c], g16[9 + c], b);
    }
    const R W = (R)g16[12];
    R ub[3] = {(R)0, (R)0, (R)0};
    R dA[14];
    if constexpr (kDC) {
#pragma unroll
      for (int i = 0; i < 14; ++i) dA[i] = (R)0;
    }
    const S* qrec = qps + (size_t)e * 4 * pt_stride<S>();
    // float: two points' record loads in flight (136 VGPRs, three
    // wavefronts per SIMD: 8.23 ms at 128^3 against 8.51 not unrolled);
    // double and the DC / matrix-tau variants: not unrolled (registers)
    constexpr int kPairUnroll = (sizeof(S) == 4 && !kDC && !kMT) ? 2 : 1;
#pragma unroll kPairUnroll
    for (int q = 0; q < 4; ++q) {
      S o[12];
      load_pt<S>(qrec + q * pt_stride<S>(), o);
      const R uq[3] = {(R)o[0], (R)o[1], (R)o[2]};
      ub[0] += uq[0];
      ub[1] += uq[1];
      ub[2] += uq[2];
      const R rho = (R)o[kPtRho], drdp = (R)o[kPtDrdp], drdT = (R)o[kPtDrdT];
      const R e1p = (R)o[kPtE1p], e3p = (R)o[kPtE3p], e4p = (R)o[kPtE4p];
      const R t0 = (R)o[kPtT0], wt1 = (R)o[kPtT1], wt2 = (R)o[kPtT2];
      const R Nqa = (R)sel4(cQ.N[q][0], cQ.N[q][1], cQ.N[q][2], cQ.N[q][3], a);
      const R Nqb = (R)sel4(cQ.N[q][0], cQ.N[q][1], cQ.N[q][2], cQ.N[q][3], b);
      if constexpr (kDC) {
        R dq;
        if constexpr (kMT) {
          // s_a . (DC giju_q) s_b of the point's rotated giju
          const double* d = A.dcw + ((size_t)e * 4 + q) * 8;
          dq = (R)0;
#pragma unroll
          for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) dq += sR[i] * (R)d[gu_idx(i, j)] * rR[j];
        } else {
          dq = (R)A.dcw[(size_t)e * 4 + q];
        }
        dc_point<R>(dq, drdp, drdT, rho, uq, e1p, e4p, dA);
      }
      if constexpr (kMT) {
        R WPT[25];
        const S* mp = reinterpret_cast<const S*>(A.mtau) + ((size_t)e * 4 + q) * mt_stride<S>();
#pragma unroll
        for (int k = 0; k < 25; ++k) WPT[k] = (R)mp[k];
        mtau_pq<R>(uq, rho, drdp, drdT, e1p, e3p, e4p, WPT, cfacR, sR, rR, W * Nqa, Nqb, acc);
        if (pair)
          mtau_pq<R>(uq, rho, drdp, drdT, e1p, e3p, e4p, WPT, cfacR, rR, sR, W * Nqb, Nqa,
                     acc2);
      } else {
        shakib_pq<R>(uq, rho, drdp, drdT, e1p, e3p, e4p, t0, wt1, wt2, cfacR, sR, rR, W * Nqa,
                     Nqb, acc);
        if (pair)
          shakib_pq<R>(uq, rho, drdp, drdT, e1p, e3p, e4p, t0, wt1, wt2, cfacR, rR, sR,
                       W * Nqb, Nqa, acc2);
      }
    }
    visc_add<R>(W, sR, rR, ub, muR, lamR, kapR, acc);
    if (pair) visc_add<R>(W, rR, sR, ub, muR, lamR, kapR, acc2);
    if constexpr (kDC) {
      R sab = (R)1;
      if constexpr (!kMT) {
        // giju of the element from dxi/dx = 2 (shape gradients of nodes 1..3)
        R ix[3][3], G[6], trG, GG, gu[6];
#pragma unroll
        for (int kk = 0; kk < 3; ++kk)
#pragma unroll
          for (int c = 0; c < 3; ++c) ix[kk][c] = (R)2 * (R)g16[3 * kk + c];
        tet_metric_tensor(ix, G, trG, GG);
        metric_inverse(G, gu);
        sab = (R)0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) sab += sR[i] * gu[gu_idx(i, j)] * rR[j];
      }
      dc_block<R>(W * sab, dA, acc);
      if (pair) dc_block<R>(W * sab, dA, acc2);
    }
  }
  if (diag) {
    for (int k = 0; k < cnt; ++k) {
      const int code = A.blocks[L.x + k];
      const double* ro = A.resid + ((size_t)(code >> 4) * 4 + ((code >> 2) & 3)) * 5;
#pragma unroll
      for (int f = 0; f < 5; ++f) racc[f] += ro[f];
    }
  }
  // chunk combine as in k_asm_blocks, one block at a time (the second one's
  // partials move only in wavefronts of pairs, wavefront-uniform: padding
  // lanes hold zeros), each stored by the leader before the next
  const bool leader = (L.w & 1) && l < A.nlanes;
  const int nch = (L.w >> 8) & 0xff;
  const int total = L.w >> 16;
  int mx = nch, wp = pair ? 1 : 0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    mx = max(mx, __shfl_xor(mx, o, 64));
    wp = max(wp, __shfl_xor(wp, o, 64));
  }
  R csum[25];
  auto combine = [&](const R* ac) {
#pragma unroll
    for (int i = 0; i < 25; ++i) csum[i] = ac[i];
    for (int c = 1; c < mx; ++c) {
#pragma unroll
      for (int i = 0; i < 25; ++i) {
        const R v = __shfl_down(ac[i], c, 64);
        if (c < nch) csum[i] += v;
      }
    }
  };
  if (!wp) {
    for (int c = 1; c < mx; ++c) {
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        const double v = __shfl_down(racc[f], c, 64);
        if (c < nch) racc[f] += v;
      }
    }
  }
  int v = 0, wcol = 0;
  uint8_t ma = 0, mb = 0;
  if (leader) {
    const int code0 = A.blocks[L.x];
    const long long e0 = code0 >> 4;
    const int a0 = (code0 >> 2) & 3, b0 = code0 & 3;
    const long long nel = A.nel;
    v = A.ien[a0 * nel + e0];
    wcol = A.ien[b0 * nel + e0];
    ma = A.bcm[v];
    mb = A.bcm[wcol];
  }
  // bc3LHS and the store of one block: rows masked by mr, columns by mc
  auto store = [&](long long entry, uint8_t mr, uint8_t mc, bool dg) {
    double sum[25];
#pragma unroll
    for (int i = 0; i < 25; ++i) sum[i] = (double)csum[i];
    if (dg && A.prec_on) {
#pragma unroll
      for (int i = 0; i < 25; ++i) A.bdiag[25 * (size_t)v + i] = sum[i];
    }
    if (mr | mc) {
#pragma unroll
      for (int j = 0; j < 5; ++j)
#pragma unroll
        for (int i = 0; i < 5; ++i) {
          if (((mr >> i) & 1) || ((mc >> j) & 1)) sum[i + 5 * j] = 0.0;
          if (dg && i == j && ((mr >> i) & 1)) sum[i + 5 * j] = (double)total;
        }
    }
    T* dst = reinterpret_cast<T*>(A.lhs) + (size_t)aos_stride<T>() * entry;
    if constexpr (sizeof(T) == 4) {
      float4* d4 = reinterpret_cast<float4*>(dst);
#pragma unroll
      for (int k = 0; k < 6; ++k)
        d4[k] = make_float4((float)sum[4 * k], (float)sum[4 * k + 1], (float)sum[4 * k + 2],
                            (float)sum[4 * k + 3]);
      d4[6] = make_float4((float)sum[24], 0.f, 0.f, 0.f);
    } else {
      double2* d2 = reinterpret_cast<double2*>(dst);
#pragma unroll
      for (int k = 0; k < 12; ++k) d2[k] = make_double2(sum[2 * k], sum[2 * k + 1]);
      d2[12] = make_double2(sum[24], 0.0);
    }
  };
  combine(acc);
  if (leader) {
    if (diag) {
#pragma unroll
      for (int f = 0; f < 5; ++f) A.res[5 * (size_t)v + f] = racc[f];
    }
    store(L.y, ma, mb, diag);
  }
  if (wp) {
    combine(acc2);
    if (leader && pair) store(k2, mb, ma, false);
  }
}

// Shakib's or Franca's tau without DC, float or double
template <typename T, typename S>
__global__ void __launch_bounds__(64) k_asm_pairs(BlkArgs A) {
  asm_pairs<T, S, false, false>(A);
}

// DC and / or the matrix tau, mixed precision
template <bool kDC, bool kMT>
__global__ void __launch_bounds__(64) k_asm_pairs_dcmt(BlkArgs A) {
  asm_pairs<float, float, kDC, kMT>(A);
}

// Wall faces: floating convective, pressure, viscous and heat fluxes
// (iBCB = 0), residual only (iLHScond = 0).  Nodes 1..3 lie on the face,
// ordered so (x2-x1) x (x3-x1) points outward.
struct WallArgs {
  const int* ien;   // 4 planes of nelb, colour order
  const double* node;
  double* res;      // 5 per node
  long long nshg, nelb;
  int p0, p1;
  PhysParams ph;
};

__global__ void __launch_bounds__(64) k_wall(WallArgs A) {
  const int p = A.p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.p1) return;
  const long long nshg = A.nshg, nelb = A.nelb;
  const PhysParams& ph = A.ph;
  int nd[4];
  double yl[4][5], xl[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    nd[a] = A.ien[a * nelb + p];
    const double* v = A.node + 16 * (size_t)nd[a];
#pragma unroll
    for (int f = 0; f < 5; ++f) yl[a][f] = v[f];
#pragma unroll
    for (int c = 0; c < 3; ++c) xl[a][c] = v[10 + c];
  }
  // volume metrics for the gradients
  double d[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) d[i][j] = 0.5 * (xl[j][i] - xl[3][i]);
  double ix[3][3];
  ix[0][0] = d[1][1] * d[2][2] - d[2][1] * d[1][2];
  ix[0][1] = d[2][1] * d[0][2] - d[0][1] * d[2][2];
  ix[0][2] = d[0][1] * d[1][2] - d[0][2] * d[1][1];
  const double idet = 1.0 / (ix[0][0] * d[0][0] + ix[0][1] * d[1][0] + ix[0][2] * d[2][0]);
  ix[0][0] *= idet;
  ix[0][1] *= idet;
  ix[0][2] *= idet;
  ix[1][0] = (d[1][2] * d[2][0] - d[1][0] * d[2][2]) * idet;
  ix[1][1] = (d[0][0] * d[2][2] - d[2][0] * d[0][2]) * idet;
  ix[1][2] = (d[1][0] * d[0][2] - d[0][0] * d[1][2]) * idet;
  ix[2][0] = (d[1][0] * d[2][1] - d[1][1] * d[2][0]) * idet;
  ix[2][1] = (d[2][0] * d[0][1] - d[0][0] * d[2][1]) * idet;
  ix[2][2] = (d[0][0] * d[1][1] - d[0][1] * d[1][0]) * idet;
  double shg[4][3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    shg[0][k] = 0.5 * ix[0][k];
    shg[1][k] = 0.5 * ix[1][k];
    shg[2][k] = 0.5 * ix[2][k];
    shg[3][k] = -0.5 * (ix[0][k] + ix[1][k] + ix[2][k]);
  }
  double g[3][5];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int f = 0; f < 5; ++f)
      g[i][f] = shg[0][i] * yl[0][f] + shg[1][i] * yl[1][f] + shg[2][i] * yl[2][f] +
                shg[3][i] * yl[3][f];
  // outward unit normal and weighted area
  double v1[3], v2[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    v1[c] = xl[1][c] - xl[0][c];
    v2[c] = xl[2][c] - xl[0][c];
  }
  double nrm[3] = {v1[1] * v2[2] - v2[1] * v1[2], v2[0] * v1[2] - v1[0] * v2[2],
                   v1[0] * v2[1] - v2[0] * v1[1]};
  const double area2 = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
  const double inv = 1.0 / area2;
#pragma unroll
  for (int c = 0; c < 3; ++c) nrm[c] *= inv;
  const double WdetJb = cQ.Qwtb * area2 * 0.25;

  const double cp = ph.Rgas * ph.gamma / ph.gamma1;
  const double mu = ph.mu, lam = -(2.0 / 3.0) * mu, kap = mu * cp / ph.pr;
  const double divu = g[0][1] + g[1][2] + g[2][3];
  // wall traction and heat flux use the volume gradients (constant)
  double tn[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
      acc += nrm[i] * (mu * (g[i][1 + m] + g[m][1 + i]) + (m == i ? lam * divu : 0.0));
    tn[m] = acc;
  }
  const double heat = -kap * (nrm[0] * g[0][4] + nrm[1] * g[1][4] + nrm[2] * g[2][4]);

  double rl[3][5];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int f = 0; f < 5; ++f) rl[a][f] = 0.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double* Nq = cQ.Nb[q];
    double Y[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) Y[f] = Nq[0] * yl[0][f] + Nq[1] * yl[1][f] + Nq[2] * yl[2][f];
    const double pres = Y[0], u1 = Y[1], u2 = Y[2], u3 = Y[3], Tq = Y[4];
    const double rk = 0.5 * (u1 * u1 + u2 * u2 + u3 * u3);
    const double rho = pres / (ph.Rgas * Tq);
    const double ei = Tq * (ph.Rgas / ph.gamma1);
    const double un = nrm[0] * u1 + nrm[1] * u2 + nrm[2] * u3;
    const double rou = rho * un;
    double F[5];
    F[0] = rou;
    F[1] = rou * u1 + nrm[0] * pres - tn[0];
    F[2] = rou * u2 + nrm[1] * pres - tn[1];
    F[3] = rou * u3 + nrm[2] * pres - tn[2];
    F[4] = rou * (ei + rk) + un * pres - (u1 * tn[0] + u2 * tn[1] + u3 * tn[2]) + heat;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int f = 0; f < 5; ++f) rl[a][f] += WdetJb * Nq[a] * F[f];
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int f = 0; f < 5; ++f) A.res[5 * (size_t)nd[a] + f] += rl[a][f];
}

// Accumulation layout -> reference planes, with the Dirichlet rows of res
// and BDiag (bc3res.f:25-153, bc3bdg.f:25-330)
// The assembly leaves res (5) and BDiag (25) per node contiguous; a block of
// 128 nodes stages its records through LDS (coalesced reads) and writes the
// Dirichlet-masked values as planes.
__global__ void __launch_bounds__(128) k_bc_nodes(const double* __restrict__ ares,
                                                  const double* __restrict__ abd,
                                                  double* __restrict__ res,
                                                  double* __restrict__ bdiag,
                                                  const uint8_t* __restrict__ bcm, long long nshg,
                                                  int prec) {
  __shared__ double sb[128 * 25];
  const long long n0 = blockIdx.x * 128LL;
  const int cnt = (int)(nshg - n0 < 128 ? nshg - n0 : 128);
  const int t = threadIdx.x;
  const long long n = n0 + t;
  for (int k = t; k < 5 * cnt; k += 128) sb[k] = ares[5 * n0 + k];
  __syncthreads();
  const uint8_t m = t < cnt ? bcm[n] : 0;
  if (t < cnt) {
#pragma unroll
    for (int f = 0; f < 5; ++f) res[n + f * nshg] = ((m >> f) & 1) ? 0.0 : sb[5 * t + f];
  }
  if (!prec) return;
  __syncthreads();
  for (int k = t; k < 25 * cnt; k += 128) sb[k] = abd[25 * n0 + k];
  __syncthreads();
  if (t >= cnt) return;
#pragma unroll
  for (int j = 0; j < 5; ++j)
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      double v = sb[25 * t + i + 5 * j];
      if (((m >> i) & 1) || ((m >> j) & 1)) v = (i == j) ? 1.0 : 0.0;
      bdiag[n + (i + 5 * j) * nshg] = v;
    }
}

// y, ac (global order, planes) and x -> one 16-double line per node
__global__ void k_pack_nodes(const double* y, const double* ac, const double* x, double* pack,
                             long long nshg) {
  long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= nshg) return;
  double v[16];
  gather_local(y, nshg, (int)n, v);
  gather_local(ac, nshg, (int)n, v + 5);
  v[10] = x[n];
  v[11] = x[n + nshg];
  v[12] = x[n + 2 * nshg];
  v[13] = v[14] = v[15] = 0.0;
  double2* d = reinterpret_cast<double2*>(pack + 16 * n);
#pragma unroll
  for (int k = 0; k < 8; ++k) d[k] = make_double2(v[2 * k], v[2 * k + 1]);
}

// Periodic fold: res(master) += res(slaves) in ascending slave order, slave
// rows zeroed; BDiag(master) += BDiag(slaves), then slaves copy the master
// (bc3res.f:157-163, bc3bdg.f:338-352).
__global__ void k_fold_res_bdiag(double* res, double* bdiag, const int* master,
                                 const int* ptr, const int* slave, int nfold,
                                 long long nshg, int prec) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfold) return;
  const long long m = master[i];
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const long long s = slave[k];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      res[m + f * nshg] += res[s + f * nshg];
      res[s + f * nshg] = 0.0;
    }
  }
  if (!prec) return;
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const long long s = slave[k];
#pragma unroll
    for (int e = 0; e < 25; ++e) bdiag[m + e * nshg] += bdiag[s + e * nshg];
  }
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const long long s = slave[k];
#pragma unroll
    for (int e = 0; e < 25; ++e) bdiag[s + e * nshg] = bdiag[m + e * nshg];
  }
}

// ---------------- global reconstruction of the diffusive flux ----------------
// (AsIq/e3q/e3qvar + qpbc, elmgmr.f:361-391, idiff 1 or 3.)  For linear tets
// the velocity and temperature gradients are constant on an element, so the
// point flux qdi (e3q.f:109-163) is, per direction i, the viscous stress row
// tau_i (momentum) and u(x_q) . tau_i + k T,i (energy).  Its nodal residual
// ql(a) = sum_q N_a(x_q) WdetJ qdi(x_q) is therefore sN_a W tau_i and
// (Md_a u_a + Mo_a sum_b u_b) . W tau_i + sN_a W k T,i: k_q_elem stores per
// element W tau (6), W k grad T (3), (sum_b u_b) . W tau_i (3) and W, and
// k_q_nodes sums over the node's elements in element order with its own
// velocity u_a.  k_q_fold and k_q_scale apply qpbc (periodic fold, copy
// back, q / lumped mass).
constexpr int kQrec = 16;  // element record (13 values + pad): one 128-byte line

__global__ void __launch_bounds__(256) k_q_elem(const int* __restrict__ ien,
                                                const double* __restrict__ node,
                                                double* __restrict__ qrec, long long nel, double mu,
                                                double lam, double kap) {
  // contiguous element ranges per XCD (workgroups go round robin over 8;
  // the grid is a multiple of 8), so neighbouring elements share an L2
  const long long per = gridDim.x / 8;
  const long long wg = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const long long e = wg * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  double xl[4][3], vl[4][4];  // coordinates; u, v, w, T
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double2* v = reinterpret_cast<const double2*>(node + 16 * (size_t)ien[k * nel + e]);
    const double2 pu = v[0], vw = v[1], tp = v[2], x01 = v[5], x2 = v[6];
    vl[k][0] = pu.y;
    vl[k][1] = vw.x;
    vl[k][2] = vw.y;
    vl[k][3] = tp.x;
    xl[k][0] = x01.x;
    xl[k][1] = x01.y;
    xl[k][2] = x2.x;
  }
  double d[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) d[i][j] = 0.5 * (xl[j][i] - xl[3][i]);
  double ix[3][3];
  ix[0][0] = d[1][1] * d[2][2] - d[2][1] * d[1][2];
  ix[0][1] = d[2][1] * d[0][2] - d[0][1] * d[2][2];
  ix[0][2] = d[0][1] * d[1][2] - d[0][2] * d[1][1];
  const double det = ix[0][0] * d[0][0] + ix[0][1] * d[1][0] + ix[0][2] * d[2][0];
  const double idet = 1.0 / det;
  ix[0][0] *= idet;
  ix[0][1] *= idet;
  ix[0][2] *= idet;
  ix[1][0] = (d[1][2] * d[2][0] - d[1][0] * d[2][2]) * idet;
  ix[1][1] = (d[0][0] * d[2][2] - d[2][0] * d[0][2]) * idet;
  ix[1][2] = (d[1][0] * d[0][2] - d[0][0] * d[1][2]) * idet;
  ix[2][0] = (d[1][0] * d[2][1] - d[1][1] * d[2][0]) * idet;
  ix[2][1] = (d[2][0] * d[0][1] - d[0][0] * d[2][1]) * idet;
  ix[2][2] = (d[0][0] * d[1][1] - d[0][1] * d[1][0]) * idet;
  const double W = cQ.Qwt * det;
  // g[i][f]: d/dx_i of (u, v, w, T)
  double g[3][4];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double s0 = 0.5 * ix[0][i], s1 = 0.5 * ix[1][i], s2 = 0.5 * ix[2][i];
    const double s3 = -(s0 + s1 + s2);
#pragma unroll
    for (int f = 0; f < 4; ++f) g[i][f] = s0 * vl[0][f] + s1 * vl[1][f] + s2 * vl[2][f] + s3 * vl[3][f];
  }
  const double l2m = lam + 2.0 * mu;
  double r[kQrec];
  r[0] = W * (l2m * g[0][0] + lam * g[1][1] + lam * g[2][2]);  // tau_11
  r[1] = W * (mu * (g[0][1] + g[1][0]));                       // tau_12
  r[2] = W * (mu * (g[0][2] + g[2][0]));                       // tau_13
  r[3] = W * (lam * g[0][0] + l2m * g[1][1] + lam * g[2][2]);  // tau_22
  r[4] = W * (mu * (g[1][2] + g[2][1]));                       // tau_23
  r[5] = W * (lam * g[0][0] + lam * g[1][1] + l2m * g[2][2]);  // tau_33
  r[6] = W * (kap * g[0][3]);
  r[7] = W * (kap * g[1][3]);
  r[8] = W * (kap * g[2][3]);
  const double su[3] = {vl[0][0] + vl[1][0] + vl[2][0] + vl[3][0],
                        vl[0][1] + vl[1][1] + vl[2][1] + vl[3][1],
                        vl[0][2] + vl[1][2] + vl[2][2] + vl[3][2]};
  r[9] = su[0] * r[0] + su[1] * r[1] + su[2] * r[2];
  r[10] = su[0] * r[1] + su[1] * r[3] + su[2] * r[4];
  r[11] = su[0] * r[2] + su[1] * r[4] + su[2] * r[5];
  r[12] = W;
  r[13] = r[14] = r[15] = 0.0;
  double2* o = reinterpret_cast<double2*>(qrec + (size_t)kQrec * e);
#pragma unroll
  for (int k = 0; k < kQrec / 2; ++k) o[k] = make_double2(r[2 * k], r[2 * k + 1]);
}

// one lane per node, 64-node groups: the node's elements in element order,
// stored sliced (entry j of lane l of group g at inc[goff[g] + 64 j + l],
// -1 past the node's count) so that the lanes' reads coalesce
__global__ void __launch_bounds__(256) k_q_nodes(const long long* __restrict__ goff,
                                                 const int* __restrict__ inc,
                                                 const double* __restrict__ node,
                                                 const double* __restrict__ qrec,
                                                 double* __restrict__ qnode, long long nshg,
                                                 int idiff) {
  // contiguous node ranges per XCD (grid a multiple of 8): the records of an
  // element are read by its four nodes through one L2
  const long long per = gridDim.x / 8;
  const long long wg = (blockIdx.x % 8) * per + blockIdx.x / 8;
  const long long n = wg * blockDim.x + threadIdx.x;
  if (n >= nshg) return;
  const long long g = n >> 6;
  const int lane = (int)(n & 63);
  const long long o0 = goff[g];
  const int deg = (int)((goff[g + 1] - o0) >> 6);
  const double2* nv = reinterpret_cast<const double2*>(node + 16 * (size_t)n);
  const double2 pu = nv[0], vw = nv[1];
  const double un[3] = {pu.y, vw.x, vw.y};
  double acc[13];
#pragma unroll
  for (int k = 0; k < 13; ++k) acc[k] = 0.0;
  for (int j = 0; j < deg; ++j) {
    const int code = inc[o0 + 64LL * j + lane];
    if (code < 0) continue;
    const long long e = code >> 2;
    const int a = code & 3;
    const double2* rv = reinterpret_cast<const double2*>(qrec + (size_t)kQrec * e);
    double r[14];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const double2 t2 = rv[k];
      r[2 * k] = t2.x;
      r[2 * k + 1] = t2.y;
    }
    const double sN = sel4(cQ.sN[0], cQ.sN[1], cQ.sN[2], cQ.sN[3], a);
    const double md = sel4(cQ.Md[0], cQ.Md[1], cQ.Md[2], cQ.Md[3], a);
    const double mo = sel4(cQ.Mo[0], cQ.Mo[1], cQ.Mo[2], cQ.Mo[3], a);
    // stress rows: direction 1 (11, 12, 13), 2 (12, 22, 23), 3 (13, 23, 33)
    const double t1[3] = {r[0], r[1], r[2]}, t2[3] = {r[1], r[3], r[4]}, t3[3] = {r[2], r[4], r[5]};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      acc[k] += sN * t1[k];
      acc[4 + k] += sN * t2[k];
      acc[8 + k] += sN * t3[k];
    }
    acc[3] += md * (un[0] * t1[0] + un[1] * t1[1] + un[2] * t1[2]) + mo * r[9] + sN * r[6];
    acc[7] += md * (un[0] * t2[0] + un[1] * t2[1] + un[2] * t2[2]) + mo * r[10] + sN * r[7];
    acc[11] += md * (un[0] * t3[0] + un[1] * t3[1] + un[2] * t3[2]) + mo * r[11] + sN * r[8];
    const double lump = idiff == 3 ? sel4(cQ.lump3[0], cQ.lump3[1], cQ.lump3[2], cQ.lump3[3], a) : sN;
    acc[12] += lump * r[12];
  }
  double2* o = reinterpret_cast<double2*>(qnode + 16 * (size_t)n);
#pragma unroll
  for (int k = 0; k < 6; ++k) o[k] = make_double2(acc[2 * k], acc[2 * k + 1]);
  o[6] = make_double2(acc[12], 0.0);
  o[7] = make_double2(0.0, 0.0);
}

// qpbc.f: master += slaves (ascending), then the slaves copy the master
__global__ void k_q_fold(double* qnode, const int* master, const int* ptr, const int* slave,
                         int nfold) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nfold) return;
  double* qm = qnode + 16 * (size_t)master[i];
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    const double* qs = qnode + 16 * (size_t)slave[k];
#pragma unroll
    for (int f = 0; f < 13; ++f) qm[f] += qs[f];
  }
  for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
    double* qs = qnode + 16 * (size_t)slave[k];
#pragma unroll
    for (int f = 0; f < 13; ++f) qs[f] = qm[f];
  }
}

// qpbc.f: rmass = 1/rmass, q = rmass q
__global__ void k_q_scale(double* qnode, long long nshg) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= nshg) return;
  double* q = qnode + 16 * (size_t)n;
  const double r = 1.0 / q[12];
#pragma unroll
  for (int f = 0; f < 12; ++f) q[f] = r * q[f];
  q[12] = r;
}

template <typename T>
__global__ void k_fill(T* p, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = T(0);
}

bool g_quad_ready = false;

// ---------------- matrix-free GMRES (SolMFG) ----------------
// The modified residual of the matrix-free solver (e3 with ires = 2, and the
// rmes of ElmMFG with ires = 3; e3.f:95-300) of the variables yp at the
// state y, per element and node.  Lane q = quadrature point q.  The state
// (pressure, velocity, temperature and everything derived from them: the
// A matrices, tau, viscosity) comes from y (e3ivar.f:150-155); the point
// values and gradients from yp (e3ivar.f:96-101,259-334).  With
//     L  = A_i yp_,i                        (e3conv.f:100-184, non-integrated)
//     M  = L + c U(yp),  c = almi/(gami alfi) Dtgl, U the conservative
//          variables of yp at the point    (e3ivar.f:108-129, e3ls.f:103,
//                                            e3massr.f:71-80)
// the point contributes WdetJ [ s_a,i (F^v_i(yp) + A_i tau M) + N_a M ] to
// node a (e3visc.f:278-343, e3tau.f:263-271, e3ls.f:231-348,
// e3wmlt.f:97-122): no Euler flux and no acceleration.  The four lanes of
// an element sum their contributions as the residual pass does.  iabs
// takes |.| of each element-node value (asires.f:90, itrFDI's estimate of
// the cancellation error).  Everything that comes from the state is read
// from the records the point pass stored at that state (the point values,
// thermodynamics and W tau, the shape gradients and WdetJ): the solver's
// products all run at one state, so only yp's values and gradients are
// formed here.  With the matrix tau (itau >= 10) tau M is PTau M
// (e3tau.f:723-732), PTau the point's matrix tau at the state, from the
// W PTau records of the same point pass (it depends on the state alone:
// e3ls.f:198-214 forms it from the state's point values, acceleration and
// metrics).
struct MresArgs {
  const int* ien;       // 4 planes of nel
  const double* node;   // 16 per node: y (local order), ac, x
  const double* yp;     // variables in the local order, yp_stride doubles per node
                        // (the node lines themselves for yp = y)
  double* resid;        // 5 per (e, a) at slot eslot[4 e + a]
  const int* eslot;
  long long nel;
  PhysParams ph;
  double fct1;          // almi/gami/alfi*Dtgl
  int yp_stride;        // 16: yp is y itself (ElmMFG), 8: a packed perturbed state
  int iabs;
  const double* mtau;   // itau = 10: nel*4 x mt_stride<double>() W PTau; else null
  const double* qps;    // the point pass's FP64 point records at the state
  const double* geo;    // and its geometry records
};

__global__ void __launch_bounds__(256) k_mres(MresArgs A) {
  const long long e_raw = blockIdx.x * 64LL + threadIdx.x / 4;
  const int q = threadIdx.x & 3;
  const bool valid = e_raw < A.nel;
  const long long e = valid ? e_raw : A.nel - 1;  // idle lanes load a valid element
  const long long nel = A.nel;
  const PhysParams& ph = A.ph;
  int nd[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) nd[k] = A.ien[k * nel + e];
  // the state's point values, thermodynamics and W tau, and the element's
  // shape gradients and WdetJ, as the point pass stored them at this state
  double rec[12], g16[16];
  load_pt<double>(A.qps + ((size_t)e * 4 + q) * pt_stride<double>(), rec);
  load_rec<double>(A.geo + (size_t)e * rec_stride<double>(), g16);
  const double WdetJ = g16[12];
  double shg[4][3];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) shg[k][c] = g16[3 * k + c];
  Thermo t;
  t.u[0] = rec[0];
  t.u[1] = rec[1];
  t.u[2] = rec[2];
  t.rho = rec[kPtRho];
  t.drdp = rec[kPtDrdp];
  t.drdT = rec[kPtDrdT];
  t.e1p = rec[kPtE1p];
  t.e2p = t.e1p + 1.0;
  t.e3p = rec[kPtE3p];
  t.e4p = rec[kPtE4p];
  t.tau[0] = rec[kPtT0] / WdetJ;
  t.tau[1] = rec[kPtT1] / WdetJ;
  t.tau[2] = rec[kPtT2] / WdetJ;
  double Nq[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) Nq[k] = sel4(cQ.N[0][k], cQ.N[1][k], cQ.N[2][k], cQ.N[3][k], q);
  double Yp[5], g[3][5];
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    Yp[f] = 0.0;
    g[0][f] = g[1][f] = g[2][f] = 0.0;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double2* vp = reinterpret_cast<const double2*>(A.yp + (size_t)A.yp_stride * nd[k]);
    double p6[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double2 b = vp[c];
      p6[2 * c] = b.x;
      p6[2 * c + 1] = b.y;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      Yp[f] += Nq[k] * p6[f];
#pragma unroll
      for (int i = 0; i < 3; ++i) g[i][f] += shg[k][i] * p6[f];
    }
  }
  const double cp = ph.Rgas * ph.gamma / ph.gamma1;
  const double mu = ph.mu, lam = -(2.0 / 3.0) * mu, kap = mu * cp / ph.pr;
  // L = A_i yp_,i and M = L + c U(yp)
  double M[5], tmp[5];
  Ai_times(t, 0, g[0], M);
#pragma unroll
  for (int i = 1; i < 3; ++i) {
    Ai_times(t, i, g[i], tmp);
#pragma unroll
    for (int f = 0; f < 5; ++f) M[f] += tmp[f];
  }
  {
    const double rkp = 0.5 * (Yp[1] * Yp[1] + Yp[2] * Yp[2] + Yp[3] * Yp[3]);
    const double rhop = Yp[0] / (ph.Rgas * Yp[4]);
    const double eip = Yp[4] * (ph.Rgas / ph.gamma1);
    M[0] += A.fct1 * rhop;
    M[1] += A.fct1 * (rhop * Yp[1]);
    M[2] += A.fct1 * (rhop * Yp[2]);
    M[3] += A.fct1 * (rhop * Yp[3]);
    M[4] += A.fct1 * (rhop * (eip + rkp));
  }
  double TM[5];
  if (A.mtau) {
    const double* wp = A.mtau + ((size_t)e * 4 + q) * mt_stride<double>();
    const double iW = 1.0 / WdetJ;
#pragma unroll
    for (int i = 0; i < 5; ++i)
      TM[i] = (wp[5 * i] * M[0] + wp[5 * i + 1] * M[1] + wp[5 * i + 2] * M[2] +
               wp[5 * i + 3] * M[3] + wp[5 * i + 4] * M[4]) * iW;
  } else {
    TM[0] = t.tau[0] * M[0];
    TM[1] = t.tau[1] * M[1];
    TM[2] = t.tau[1] * M[2];
    TM[3] = t.tau[1] * M[3];
    TM[4] = t.tau[2] * M[4];
  }
  const double divu = g[0][1] + g[1][2] + g[2][3];
  double F[3][5];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Ai_times(t, i, TM, F[i]);
    double vf[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      vf[m] = mu * (g[i][1 + m] + g[m][1 + i]) + (m == i ? lam * divu : 0.0);
      F[i][1 + m] += vf[m];
    }
    F[i][4] += t.u[0] * vf[0] + t.u[1] * vf[1] + t.u[2] * vf[2] + kap * g[i][4];
  }
  double mine[5];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double r5[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
      r5[f] = WdetJ * (shg[a][0] * F[0][f] + shg[a][1] * F[1][f] + shg[a][2] * F[2][f] +
                       Nq[a] * M[f]);
      r5[f] += __shfl_xor(r5[f], 1, 64);
      r5[f] += __shfl_xor(r5[f], 2, 64);
      if (a == q) mine[f] = r5[f];
    }
  }
  if (valid) {
    double* ro = A.resid + (size_t)A.eslot[e * 4 + q] * 5;
#pragma unroll
    for (int f = 0; f < 5; ++f) ro[f] = A.iabs ? fabs(mine[f]) : mine[f];
  }
}

// out (+)= the sum of the node's element contributions in element order
// (the scatter of asires.f:91 / asimfg.f:96; read from the incidence slots
// k_mres writes), then bc3Res's Dirichlet rows
// (bc3res.f:25-153); the periodic fold follows (k_fold_res_bdiag)
__global__ void __launch_bounds__(256) k_mres_nodes(const long long* __restrict__ goff,
                                                    const int* __restrict__ inc,
                                                    const double* __restrict__ resid,
                                                    const uint8_t* __restrict__ bcm,
                                                    double* __restrict__ out, long long nshg,
                                                    int accumulate) {
  const long long n = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (n >= nshg) return;
  const long long g = n >> 6;
  const int lane = (int)(n & 63);
  const long long o0 = goff[g];
  const int deg = (int)((goff[g + 1] - o0) >> 6);
  double acc[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) acc[f] = accumulate ? out[n + f * nshg] : 0.0;
  for (int j = 0; j < deg; ++j) {
    const long long slot = o0 + 64LL * j + lane;
    if (inc[slot] < 0) continue;
    const double* r = resid + slot * 5;
#pragma unroll
    for (int f = 0; f < 5; ++f) acc[f] += r[f];
  }
  const uint8_t m = bcm[n];
#pragma unroll
  for (int f = 0; f < 5; ++f) out[n + f * nshg] = ((m >> f) & 1) ? 0.0 : acc[f];
}

// ypack(k) = the local-order variables U^-1 (ypre + s d) of the
// preconditioned point (au1mfg.f:64-68, au2mfg.f:67-69, itrfdi.f:61-65;
// the local order is yshuffle's old order, so no shuffle is needed).  With
// bc, itrBC (itrbc.f:60-139, ires = 2: y only): the Dirichlet values of the
// node's periodic master, whose values every slave copies.
__global__ void k_mfg_point(const double* __restrict__ ypre, const double* __restrict__ d,
                            double s, const double* __restrict__ D,  // nullptr: no U^-1
                            const uint8_t* __restrict__ bcm, const double* __restrict__ BC,
                            const int* __restrict__ iper, double* __restrict__ ypack,
                            long long n, int bc) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  const long long m = bc ? (long long)iper[k] : k;
  double v[5];
#pragma unroll
  for (int f = 0; f < 5; ++f) v[f] = d ? ypre[f * n + m] + s * d[f * n + m] : ypre[f * n + m];
  // i3LU 'backward' (i3lu.f:119-139) with the node's factors
  if (D) {
#pragma unroll
    for (int i = 4; i >= 0; --i) {
      double acc = v[i];
#pragma unroll
      for (int c = 4; c > i; --c) acc -= v[c] * D[(i + 5 * c) * n + m];
      v[i] = D[(i + 5 * i) * n + m] * acc;
    }
  }
  if (bc) {
    const uint8_t msk = bcm[m];
    if (msk & 16) v[4] = BC[m + n];
    if (msk & 2) {
      v[1] = BC[m + 2 * n];
      v[2] = BC[m + 3 * n];
      v[3] = BC[m + 4 * n];
    }
    if (msk & 1) v[0] = BC[m];
  }
  double2* o = reinterpret_cast<double2*>(ypack + 8 * k);
  o[0] = make_double2(v[0], v[1]);
  o[1] = make_double2(v[2], v[3]);
  o[2] = make_double2(v[4], 0.0);
  o[3] = make_double2(0.0, 0.0);
}

// w = (L^-1 t - rmes) / e (au1mfg.f:76-83: i3LU 'forward', then the one-sided
// difference quotient)
__global__ void k_mfg_fwd_diff(const double* __restrict__ D, const double* __restrict__ t,
                               const double* __restrict__ rmes, double e,
                               double* __restrict__ w, long long n) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) v[i] = t[i * n + k];
#pragma unroll
  for (int i = 1; i < 5; ++i) {
    double acc = v[i];
#pragma unroll
    for (int c = 0; c < i; ++c) acc -= D[(i + 5 * c) * n + k] * v[c];
    v[i] = acc;
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) w[i * n + k] = (v[i] - rmes[i * n + k]) / e;
}

// u = res - (u1 - u2) / (2 eps) (au2mfg.f:112), or with mode 1 itrFDI's
// second difference d = (u1 - 2 rmes) / epsM (itrfdi.f:134), stored in u
__global__ void k_mfg_combine(const double* __restrict__ a, const double* __restrict__ u1,
                              const double* __restrict__ u2, double h, double* __restrict__ u,
                              long long n5, int mode) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n5) return;
  if (mode == 0)
    u[k] = a[k] - (u1[k] - u2[k]) / h;
  else
    u[k] = (u1[k] - 2.0 * a[k]) / h;
}

int upload_quadrature() {
  if (g_quad_ready) return 0;
  Quad h{};
  const double a4 = 0.5854101966249685, b4 = 0.1381966011250150;
  const double r4[4][3] = {{a4, b4, b4}, {b4, a4, b4}, {b4, b4, a4}, {b4, b4, b4}};
  for (int q = 0; q < 4; ++q) {
    h.N[q][0] = r4[q][0];
    h.N[q][1] = r4[q][1];
    h.N[q][2] = r4[q][2];
    h.N[q][3] = 1.0 - r4[q][0] - r4[q][1] - r4[q][2];
  }
  const double A = 0.3333333333333330, B = 0.666666666666667, C = 0.166666666666667;
  const double r3[3][3] = {{B, C, C}, {C, B, C}, {C, C, B}};
  for (int q = 0; q < 3; ++q) {
    h.Nb[q][0] = r3[q][0];
    h.Nb[q][1] = r3[q][1];
    h.Nb[q][2] = r3[q][2];
    h.Nb[q][3] = 1.0 - r3[q][0] - r3[q][1] - r3[q][2];
  }
  h.Qwt = (4.0 / 3.0) * 0.25;
  h.Qwtb = 2.0 * A;
  double ssq = 0.0;
  for (int a = 0; a < 4; ++a) {
    double M[4];
    for (int b = 0; b < 4; ++b) {
      M[b] = 0.0;
      for (int q = 0; q < 4; ++q) M[b] += h.N[q][a] * h.N[q][b];
    }
    h.Mo[a] = 0.0;
    for (int b = 0; b < 4; ++b)
      if (b != a) h.Mo[a] += M[b] / 3.0;
    h.Md[a] = M[a] - h.Mo[a];
    h.sN[a] = 0.0;
    h.lump3[a] = 0.0;
    for (int q = 0; q < 4; ++q) {
      h.sN[a] += h.N[q][a];
      h.lump3[a] += h.N[q][a] * h.N[q][a];
    }
    ssq += h.lump3[a];
  }
  for (int a = 0; a < 4; ++a) h.lump3[a] *= 4.0 / ssq;
  PHG_HIP(hipMemcpyToSymbol(HIP_SYMBOL(cQ), &h, sizeof(Quad)));
  g_quad_ready = true;
  return 0;
}

PhysParams phys_of(const phg_params& p) {
  return PhysParams{p.Rgas,   p.gamma,  p.gamma1, p.pr,   p.mu,   p.dtsfct, p.taucfct,
                    p.temper, p.almi,   p.alfi,   p.gami, p.Dtgl, p.itau};
}

}  // namespace

// point pass for the flag set: diffusive-flux correction, DC, matrix tau
template <typename S>
void launch_qp(int g1, hipStream_t st, const QpArgs& Q, bool diff, bool dc, bool mt) {
  if (mt && diff && dc)
    k_asm_qp<S, true, true, true><<<g1, 256, 0, st>>>(Q);
  else if (mt && dc)
    k_asm_qp<S, false, true, true><<<g1, 256, 0, st>>>(Q);
  else if (mt && diff)
    k_asm_qp<S, true, false, true><<<g1, 256, 0, st>>>(Q);
  else if (mt)
    k_asm_qp<S, false, false, true><<<g1, 256, 0, st>>>(Q);
  else if (diff && dc)
    k_asm_qp<S, true, true, false><<<g1, 256, 0, st>>>(Q);
  else if (diff)
    k_asm_qp<S, true, false, false><<<g1, 256, 0, st>>>(Q);
  else if (dc)
    k_asm_qp<S, false, true, false><<<g1, 256, 0, st>>>(Q);
  else
    k_asm_qp<S, false, false, false><<<g1, 256, 0, st>>>(Q);
}

int launch_fill(phg_ctx* c, void* p, size_t bytes) {
  size_t n = bytes / 4;
  int blocks = (int)std::min<size_t>((n + 255) / 256, 8192);
  if (blocks < 1) return 0;
  k_fill<float><<<blocks, 256, 0, c->stream>>>(reinterpret_cast<float*>(p), n);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_pack_nodes(phg_ctx* c) {
  const long long n = c->nshg;
  k_pack_nodes<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(c->y, c->ac, c->x, c->nodepack, n);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_diffusive_flux(phg_ctx* c) {
  if (!c->qnode || !c->qrec) {
    phg_set_error("diffusive flux buffers not allocated");
    return PHG_EINVAL;
  }
  PHG_TRY(upload_quadrature());
  const PhysParams ph = phys_of(c->prm);
  const double cp = ph.Rgas * ph.gamma / ph.gamma1;
  const long long nel = c->nel, n = c->nshg;
  auto xcd_grid = [](long long items) { return (int)(((items + 255) / 256 + 7) / 8 * 8); };
  k_q_elem<<<xcd_grid(nel), 256, 0, c->stream>>>(c->ienc, c->nodepack, c->qrec, nel, ph.mu,
                                                 -(2.0 / 3.0) * ph.mu, ph.mu * cp / ph.pr);
  PHG_HIP(hipGetLastError());
  k_q_nodes<<<xcd_grid(n), 256, 0, c->stream>>>(c->inc_goff, c->inc, c->nodepack, c->qrec,
                                                c->qnode, n, c->prm.idiff);
  PHG_HIP(hipGetLastError());
  if (c->nfold > 0) {
    k_q_fold<<<(c->nfold + 255) / 256, 256, 0, c->stream>>>(c->qnode, c->fold_master, c->fold_ptr,
                                                           c->fold_slave, c->nfold);
    PHG_HIP(hipGetLastError());
  }
  k_q_scale<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(c->qnode, n);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_assembly(phg_ctx* c, bool lhs, bool prec) {
  if (lhs) PHG_TRY(require_matrix(c, "ElmGMRs with lhs = 1"));
  PHG_TRY(upload_quadrature());
  PHG_TRY(launch_pack_nodes(c));
  const bool mixed = c->precision == PHG_MIXED;
  const bool diff = c->prm.idiff != 0;
  if (diff) PHG_TRY(launch_diffusive_flux(c));
  const PhysParams ph = phys_of(c->prm);
  // pass 1: point states and element residuals
  QpArgs Q{};
  Q.ien = c->ienc;
  Q.node = c->nodepack;
  Q.qnode = c->qnode;
  Q.qps = mixed ? (void*)c->qps32 : (void*)c->qps;
  Q.geo = mixed ? (void*)c->geo32 : (void*)c->geo;
  Q.resid = c->resid;
  Q.nel = c->nel;
  Q.ph = ph;
  const bool dc = c->prm.idc != 0;
  if (dc && !c->dcw) {
    phg_set_error("iDC %d: the DC factor buffer is not allocated", c->prm.idc);
    return PHG_EINVAL;
  }
  const bool mt = c->prm.itau >= 10;
  if (mt && !c->mtau) {
    phg_set_error("itau %d: the matrix tau buffer is not allocated", c->prm.itau);
    return PHG_EINVAL;
  }
  Q.dcw = c->dcw;
  Q.mtau = c->mtau;
  Q.epsM = c->prm.epsM;
  Q.idc = c->prm.idc;
  const int g1 = (int)((c->nel + 63) / 64);
  PHG_HIP(hipEventRecord(c->aev[0], c->stream));
  if (mixed)
    launch_qp<float>(g1, c->stream, Q, diff, dc, mt);
  else
    launch_qp<double>(g1, c->stream, Q, diff, dc, mt);
  PHG_HIP(hipGetLastError());
  PHG_HIP(hipEventRecord(c->aev[1], c->stream));
  // pass 2: block owners (res, BDiag, lhsK written once)
  BlkArgs B{};
  // without lhsK only the diagonal entries' owners have work (the node's
  // residual and BDiag): run their lanes alone; with lhsK the paired lanes
  // of k_asm_pairs / k_asm_pairs_dcmt (in double for Shakib's or Franca's
  // tau without DC)
  const bool paired = lhs && (mixed || (!dc && !mt));
  B.lanes = paired ? c->lanes_pair : lhs ? c->lanes : c->lanes_diag;
  B.lanes_k2 = c->lanes_k2;
  B.blocks = c->blocks;
  B.ien = c->ienc;
  B.bcm = c->bcm;
  B.qps = Q.qps;
  B.geo = Q.geo;
  B.resid = c->resid;
  B.dcw = c->dcw;
  B.mtau = c->mtau;
  B.res = c->acc_res;
  B.bdiag = c->acc_bd;
  B.lhs = mixed ? (void*)c->lhs_aos32 : (void*)c->lhs_aos;
  B.nel = c->nel;
  B.nlanes = paired ? c->nlanes_pair : lhs ? c->nlanes : c->nlanes_diag;
  B.lhs_on = lhs;
  B.prec_on = prec;
  const double cp = ph.Rgas * ph.gamma / ph.gamma1;
  B.mu = ph.mu;
  B.lam = -(2.0 / 3.0) * ph.mu;
  B.kap = ph.mu * cp / ph.pr;
  B.cfac = ph.almi / ph.gami / ph.alfi * ph.Dtgl;
  const int g2 = (int)(((B.nlanes + 63) / 64 + 7) / 8 * 8);  // whole XCD rounds
  if (paired && mixed && dc && mt)
    k_asm_pairs_dcmt<true, true><<<g2, 64, 0, c->stream>>>(B);
  else if (paired && mixed && dc)
    k_asm_pairs_dcmt<true, false><<<g2, 64, 0, c->stream>>>(B);
  else if (paired && mixed && mt)
    k_asm_pairs_dcmt<false, true><<<g2, 64, 0, c->stream>>>(B);
  else if (paired && mixed)
    k_asm_pairs<float, float><<<g2, 64, 0, c->stream>>>(B);
  else if (paired)
    k_asm_pairs<double, double><<<g2, 64, 0, c->stream>>>(B);
  else if (mixed && mt && dc)
    k_asm_blocks<float, float, true, true><<<g2, 64, 0, c->stream>>>(B);
  else if (mixed && mt)
    k_asm_blocks<float, float, false, true><<<g2, 64, 0, c->stream>>>(B);
  else if (mixed && dc)
    k_asm_blocks<float, float, true, false><<<g2, 64, 0, c->stream>>>(B);
  else if (mixed)
    k_asm_blocks<float, float, false, false><<<g2, 64, 0, c->stream>>>(B);
  else if (mt && dc)
    k_asm_blocks<double, double, true, true><<<g2, 64, 0, c->stream>>>(B);
  else if (mt && !lhs && prec)  // the matrix-free solver's e3bdg_nd
    k_asm_blocks<double, double, false, true, true><<<g2, 64, 0, c->stream>>>(B);
  else if (mt)
    k_asm_blocks<double, double, false, true><<<g2, 64, 0, c->stream>>>(B);
  else if (dc)
    k_asm_blocks<double, double, true, false><<<g2, 64, 0, c->stream>>>(B);
  else
    k_asm_blocks<double, double, false, false><<<g2, 64, 0, c->stream>>>(B);
  PHG_HIP(hipGetLastError());
  PHG_HIP(hipEventRecord(c->aev[2], c->stream));
  // wall faces: residual only, added colour by colour
  WallArgs W{};
  W.ien = c->ienbc;
  W.node = c->nodepack;
  W.res = c->acc_res;
  W.nshg = c->nshg;
  W.nelb = c->nelb;
  W.ph = ph;
  for (size_t k = 0; k + 1 < c->colb_off.size(); ++k) {
    W.p0 = c->colb_off[k];
    W.p1 = c->colb_off[k + 1];
    int n = W.p1 - W.p0;
    if (n <= 0) continue;
    k_wall<<<(n + 63) / 64, 64, 0, c->stream>>>(W);
  }
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_jv(phg_ctx* c, const double* v, double* w) {
  PHG_TRY(upload_quadrature());
  const PhysParams ph = phys_of(c->prm);
  JvArgs J{};
  J.ien = c->ienc;
  J.node = c->nodepack;
  J.v = v;
  J.iper = c->iper;
  J.bcm = c->bcm;
  J.out = c->resid;
  J.eslot = c->eslot;
  J.dcw = c->dcw;
  J.epsM = c->prm.epsM;
  J.nel = c->nel;
  J.nshg = c->nshg;
  J.ph = ph;
  J.cfac = ph.almi / ph.gami / ph.alfi * ph.Dtgl;
  const int gj = (int)((c->nel + 63) / 64);
  if (c->prm.idc != 0) {
    if (!c->dcw) {
      phg_set_error("tangent product with iDC %d: no assembly has formed the DC factors",
                    c->prm.idc);
      return PHG_EINVAL;
    }
    if (c->prm.itau >= 10)
      k_jv<true, true><<<gj, 256, 0, c->stream>>>(J);
    else
      k_jv<true, false><<<gj, 256, 0, c->stream>>>(J);
  } else if (c->prm.itau >= 10) {
    k_jv<false, true><<<gj, 256, 0, c->stream>>>(J);
  } else {
    k_jv<false, false><<<gj, 256, 0, c->stream>>>(J);
  }
  PHG_HIP(hipGetLastError());
  const long long n = c->nshg;
  k_jv_nodes<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(c->inc_goff, c->inc, c->resid, v,
                                                            c->iper, c->bcm, w, n);
  PHG_HIP(hipGetLastError());
  return launch_bc3per(c, w, 5);
}

int launch_bc_res_bdiag(phg_ctx* c, bool prec) {
  long long n = c->nshg;
  k_bc_nodes<<<(int)((n + 127) / 128), 128, 0, c->stream>>>(c->acc_res, c->acc_bd, c->res, c->bdiag,
                                                             c->bcm, n, prec);
  PHG_HIP(hipGetLastError());
  if (c->nfold > 0) {
    k_fold_res_bdiag<<<(c->nfold + 255) / 256, 256, 0, c->stream>>>(
        c->res, c->bdiag, c->fold_master, c->fold_ptr, c->fold_slave, c->nfold, n, prec);
    PHG_HIP(hipGetLastError());
  }
  return 0;
}

// The modified residual of the variables yp (stride 8, packed by
// launch_mfg_point; nullptr: y itself, ElmMFG) at the resident state,
// summed into out (accumulate: added to out's values), Dirichlet rows
// zeroed and the periodic slaves folded (bc3res.f)
int launch_mres(phg_ctx* c, const double* yp, double* out, bool accumulate, bool iabs) {
  PHG_TRY(upload_quadrature());
  const PhysParams ph = phys_of(c->prm);
  MresArgs M{};
  M.ien = c->ienc;
  M.node = c->nodepack;
  M.yp = yp ? yp : c->nodepack;
  M.yp_stride = yp ? 8 : 16;
  M.resid = c->resid;
  M.eslot = c->eslot;
  M.nel = c->nel;
  M.ph = ph;
  M.fct1 = ph.almi / ph.gami / ph.alfi * ph.Dtgl;
  M.iabs = iabs ? 1 : 0;
  M.mtau = c->prm.itau >= 10 ? reinterpret_cast<const double*>(c->mtau) : nullptr;
  if (!c->qps || !c->geo) {
    phg_set_error("modified residual: the FP64 point records are not allocated");
    return PHG_EINVAL;
  }
  M.qps = c->qps;
  M.geo = c->geo;
  if (c->prm.itau >= 10 && (!c->mtau || c->precision != PHG_FP64)) {
    phg_set_error("modified residual: the matrix tau needs the FP64 W PTau records");
    return PHG_EINVAL;
  }
  k_mres<<<(int)((c->nel + 63) / 64), 256, 0, c->stream>>>(M);
  PHG_HIP(hipGetLastError());
  const long long n = c->nshg;
  k_mres_nodes<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(c->inc_goff, c->inc, c->resid,
                                                              c->bcm, out, n, accumulate ? 1 : 0);
  PHG_HIP(hipGetLastError());
  if (c->nfold > 0) {
    k_fold_res_bdiag<<<(c->nfold + 255) / 256, 256, 0, c->stream>>>(
        out, nullptr, c->fold_master, c->fold_ptr, c->fold_slave, c->nfold, n, 0);
    PHG_HIP(hipGetLastError());
  }
  return 0;
}

int launch_mfg_point(phg_ctx* c, const double* ypre, const double* d, double s, double* ypack,
                     bool bc, bool lu) {
  const long long n = c->nshg;
  k_mfg_point<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(
      ypre, d, s, lu ? c->bdiag : nullptr, c->bcm, c->BCv, c->iper, ypack, n, bc ? 1 : 0);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_mfg_fwd_diff(phg_ctx* c, const double* t, const double* rmes, double e, double* w) {
  const long long n = c->nshg;
  k_mfg_fwd_diff<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(c->bdiag, t, rmes, e, w, n);
  PHG_HIP(hipGetLastError());
  return 0;
}

int launch_mfg_combine(phg_ctx* c, const double* a, const double* u1, const double* u2, double h,
                       double* u, int mode) {
  const long long n5 = 5 * c->nshg;
  k_mfg_combine<<<(int)((n5 + 255) / 256), 256, 0, c->stream>>>(a, u1, u2, h, u, n5, mode);
  PHG_HIP(hipGetLastError());
  return 0;
}
