/*
 * phoracle.c -- CPU oracle (test infrastructure, not the product).
 *
 * Plain-C FP64 restatement of the reference compressible hot path, operation
 * for operation, so that its results can be compared with the GPU library
 * and, where oracle/_ref builds, with the reference's own compiled routines.
 * Sums are written in the reference's order; terms the reference leaves out
 * because the A-matrix entry is structurally zero are kept as exact zeros
 * (x + 0 == x), which leaves every sum bit-identical.  Build with
 * -ffp-contract=off so no multiply-add is fused.
 *
 * Reference files followed (paths relative to the reference root):
 *   phSolver/compressible/elmgmr.f:280-612   ElmGMRs
 *   phSolver/compressible/asigmr.f:1-119     AsIGMR
 *   phSolver/common/localy.f, local.f        gather/scatter order
 *   phSolver/compressible/e3.f:1-306         e3 driver
 *   phSolver/compressible/e3ivar.f:1-489     e3ivar
 *   phSolver/compressible/getthm.f:106-172   getthm (ipress=0)
 *   phSolver/compressible/getdiff.f:1-269    getDiff (constant mu)
 *   phSolver/common/e3metric.f:8-80          e3metric
 *   phSolver/compressible/e3mtrx.f:1-231     e3mtrx
 *   phSolver/compressible/e3conv.f:1-352     e3conv
 *   phSolver/compressible/e3visc.f:1-357     e3visc
 *   phSolver/compressible/e3ls.f:1-771       e3LS
 *   phSolver/compressible/e3tau.f:1-283,1397-1507  e3tau (itau=0), e3gijd
 *   phSolver/compressible/e3massr.f, e3massl.f, e3wmlt.f
 *   phSolver/compressible/asiq.f, e3q.f, e3qvar.f, phSolver/common/qpbc.f
 *                                            diffusive flux (idiff 1, 3)
 *   phSolver/compressible/bc3lhs.f, bc3res.f, bc3bdg.f, bc3per.f
 *   phSolver/common/fillsparse.f:68-129,234-270  fillsparseC, sparseloc
 *   phSolver/compressible/asbmfg.f, e3b.f, e3bvar.f  wall-face fluxes
 *   phSolver/compressible/solgmr.f:368-744   SolGMRs
 *   phSolver/compressible/i3lu.f, spsi3pre.f, sparseap.f
 *   phSolver/common/mpitools.f:104-138       sumgat (single rank)
 *   phSolver/compressible/rstat.f, rstatCheck.f, tnanq.f
 *   phSolver/compressible/itrPC.f:56-148, itrbc.f:1-198
 *   phSolver/common/symtet.c, symtri.c, genint.f:1-76, genshp.f, genshpb.f,
 *   shapeFunction/src/uniformP.c:8-34         quadrature and shapes
 */
#include "phoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PT125 0.1250000000000000000000000000000
#define PT5 0.5000000000000000000000000000000
#define PT66 0.6666666666666666666666666666667
#define ONE 1.0
#define TWO 2.0

/* ------------------------------------------------------------------ */
/* quadrature: symtet.c:26-27,76-78 (4-point rule), genint.f:74 (4/3),
 * symtri.c:22-24,55-56 (3-point rule), genint.f:75 (x2), shape functions
 * uniformP.c:16-33, gradients halved genshp.f:34-35 / genshpb.f:33-34 */
void orc_quadrature(double shp[4][4], double shgl[4][3][4], double Qwt[4],
                    double shpb[3][4], double shglb[3][3][4], double Qwtb[3]) {
  const double a4 = 0.5854101966249685, b4 = 0.1381966011250150;
  const double rstw4[4][4] = {{a4, b4, b4, b4}, {b4, a4, b4, b4},
                              {b4, b4, a4, b4}, {b4, b4, b4, a4}};
  const double A = 0.3333333333333330000000000000000000;
  const double B = 0.666666666666667, C = 0.166666666666667;
  const double rst3[3][3] = {{B, C, C}, {C, B, C}, {C, C, B}};
  double dN[4][3];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) dN[i][j] = (i == 3) ? -1.0e0 : (i == j ? 1.0e0 : 0.0e0);
  for (int q = 0; q < 4; ++q) {
    double L[4];
    L[0] = rstw4[q][0];
    L[1] = rstw4[q][1];
    L[2] = rstw4[q][2];
    L[3] = 1.0e0 - rstw4[q][0] - rstw4[q][1] - rstw4[q][2];
    for (int a = 0; a < 4; ++a) {
      if (shp) shp[q][a] = L[a];
      for (int j = 0; j < 3; ++j)
        if (shgl) shgl[q][j][a] = dN[a][j] / TWO;
    }
    if (Qwt) Qwt[q] = (4.0 / 3.0) * 0.25;
  }
  for (int q = 0; q < 3; ++q) {
    double L[4];
    L[0] = rst3[q][0];
    L[1] = rst3[q][1];
    L[2] = rst3[q][2];
    L[3] = 1.0e0 - rst3[q][0] - rst3[q][1] - rst3[q][2];
    for (int a = 0; a < 4; ++a) {
      if (shpb) shpb[q][a] = L[a];
      for (int j = 0; j < 3; ++j)
        if (shglb) shglb[q][j][a] = dN[a][j] / TWO;
    }
    if (Qwtb) Qwtb[q] = TWO * A;
  }
}

/* e3metric.f:22-77 */
static void e3metric(const double xl[4][3], const double sg[3][4], double Qw,
                     double dxidx[3][3], double shg[4][3], double* WdetJ) {
  double d[3][3] = {{0}};
  for (int n = 0; n < 4; ++n)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) d[i][j] = d[i][j] + xl[n][i] * sg[j][n];
  dxidx[0][0] = d[1][1] * d[2][2] - d[2][1] * d[1][2];
  dxidx[0][1] = d[2][1] * d[0][2] - d[0][1] * d[2][2];
  dxidx[0][2] = d[0][1] * d[1][2] - d[0][2] * d[1][1];
  double tmp = ONE / (dxidx[0][0] * d[0][0] + dxidx[0][1] * d[1][0] +
                      dxidx[0][2] * d[2][0]);
  dxidx[0][0] = dxidx[0][0] * tmp;
  dxidx[0][1] = dxidx[0][1] * tmp;
  dxidx[0][2] = dxidx[0][2] * tmp;
  dxidx[1][0] = (d[1][2] * d[2][0] - d[1][0] * d[2][2]) * tmp;
  dxidx[1][1] = (d[0][0] * d[2][2] - d[2][0] * d[0][2]) * tmp;
  dxidx[1][2] = (d[1][0] * d[0][2] - d[0][0] * d[1][2]) * tmp;
  dxidx[2][0] = (d[1][0] * d[2][1] - d[1][1] * d[2][0]) * tmp;
  dxidx[2][1] = (d[2][0] * d[0][1] - d[0][0] * d[2][1]) * tmp;
  dxidx[2][2] = (d[0][0] * d[1][1] - d[0][1] * d[1][0]) * tmp;
  *WdetJ = Qw / tmp;
  for (int n = 0; n < 4; ++n)
    for (int k = 0; k < 3; ++k)
      shg[n][k] = sg[0][n] * dxidx[0][k] + sg[1][n] * dxidx[1][k] + sg[2][n] * dxidx[2][k];
}

/* e3tau.f:1397-1507, tet branch (lcsyst == 1) */
static void e3gijd(const double dx[3][3], double g[6]) {
  const double c1 = 1.259921049894873e+00, c2 = 6.299605249474365e-01;
  double t1, t2, t3;
  t1 = c1 * dx[0][0] + c2 * (dx[1][0] + dx[2][0]);
  t2 = c1 * dx[1][0] + c2 * (dx[0][0] + dx[2][0]);
  t3 = c1 * dx[2][0] + c2 * (dx[0][0] + dx[1][0]);
  g[0] = dx[0][0] * t1 + dx[1][0] * t2 + dx[2][0] * t3;
  t1 = c1 * dx[0][1] + c2 * (dx[1][1] + dx[2][1]);
  t2 = c1 * dx[1][1] + c2 * (dx[0][1] + dx[2][1]);
  t3 = c1 * dx[2][1] + c2 * (dx[0][1] + dx[1][1]);
  g[1] = dx[0][0] * t1 + dx[1][0] * t2 + dx[2][0] * t3;
  g[2] = dx[0][1] * t1 + dx[1][1] * t2 + dx[2][1] * t3;
  t1 = c1 * dx[0][2] + c2 * (dx[1][2] + dx[2][2]);
  t2 = c1 * dx[1][2] + c2 * (dx[0][2] + dx[2][2]);
  t3 = c1 * dx[2][2] + c2 * (dx[0][2] + dx[1][2]);
  g[3] = dx[0][0] * t1 + dx[1][0] * t2 + dx[2][0] * t3;
  g[4] = dx[0][1] * t1 + dx[1][1] * t2 + dx[2][1] * t3;
  g[5] = dx[0][2] * t1 + dx[1][2] * t2 + dx[2][2] * t3;
}

typedef struct {
  double shp[4][4], shgl[4][3][4], Qwt[4];
  double shpb[3][4], shglb[3][3][4], Qwtb[3];
} quad_t;

static const quad_t* quad(void) {
  static quad_t q;
  static int init = 0;
  if (!init) {
    orc_quadrature(q.shp, q.shgl, q.Qwt, q.shpb, q.shglb, q.Qwtb);
    init = 1;
  }
  return &q;
}

/* dot of row r of M with vector v, columns in ascending order */
static inline double rowdot(const double M[5][5], int r, const double v[5]) {
  return M[r][0] * v[0] + M[r][1] * v[1] + M[r][2] * v[2] + M[r][3] * v[3] +
         M[r][4] * v[4];
}

void orc_e3(const orc_params* P, const double yl[4][5],
            const double acl[4][5], const double xl[4][3], double rl[4][5],
            double EG[20][20], int want_lhs) {
  orc_e3_q(P, yl, acl, xl, NULL, rl, EG, want_lhs);
}

void orc_e3_q(const orc_params* P, const double yl[4][5],
              const double acl[4][5], const double xl[4][3],
              const double ql[4][12], double rl[4][5], double EG[20][20],
              int want_lhs) {
  const quad_t* Q = quad();
  const int lhs = want_lhs;
  memset(rl, 0, sizeof(double) * 20);
  if (lhs) memset(EG, 0, sizeof(double) * 400);
  const double Rgas = P->Rgas, gamma = P->gamma, gamma1 = P->gamma1;
  for (int q = 0; q < 4; ++q) {
    if (Q->Qwt[q] == 0.0) continue; /* e3.f:112 */
    /* getshp: hierarchic.f:30-57 (ipord = 1) */
    double shape[4], sg[3][4];
    for (int a = 0; a < 4; ++a) {
      shape[a] = Q->shp[q][a];
      for (int j = 0; j < 3; ++j) sg[j][a] = Q->shgl[q][j][a];
    }
    double ri[20], rmi[20], stiff[15][15];
    memset(ri, 0, sizeof ri);
    memset(rmi, 0, sizeof rmi);
    if (lhs) memset(stiff, 0, sizeof stiff);

    /* ---- e3ivar.f:142-156,186-194 primitive variables, acceleration ---- */
    double pres = 0, u1 = 0, u2 = 0, u3 = 0, T = 0, aci[5] = {0, 0, 0, 0, 0};
    for (int n = 0; n < 4; ++n) {
      pres = pres + shape[n] * yl[n][0];
      u1 = u1 + shape[n] * yl[n][1];
      u2 = u2 + shape[n] * yl[n][2];
      u3 = u3 + shape[n] * yl[n][3];
      T = T + shape[n] * yl[n][4];
    }
    for (int n = 0; n < 4; ++n)
      for (int f = 0; f < 5; ++f) aci[f] = aci[f] + shape[n] * acl[n][f];
    /* e3ivar.f:203 */
    double rk = PT5 * (u1 * u1 + u2 * u2 + u3 * u3);
    /* getthm.f:130-160, ithm = 7 */
    double rho = pres / (Rgas * T);
    double ei = T * (Rgas / gamma1);
    double h = T * (Rgas * gamma / gamma1);
    double cv = Rgas / gamma1;
    double cp = Rgas * gamma / gamma1;
    double alfap = ONE / T;
    double betaT = ONE / pres;
    /* getdiff.f:126-128,163-171,263 (constant mu, Stokes lambda, DNS) */
    double rmu = P->mu;
    double rlm = -PT66 * rmu;
    double con = rmu * cp / P->pr;
    double xmut = 0.0;
    rlm = rlm - PT66 * xmut;
    rmu = rmu + xmut;
    double rlm2mu = rlm + TWO * rmu;
    con = con + xmut * cp / P->pr;
    /* e3metric */
    double dxidx[3][3], shg[4][3], WdetJ;
    e3metric(xl, sg, Q->Qwt[q], dxidx, shg, &WdetJ);
    /* e3ivar.f:259-334 global gradients (nshl == 4 branch) */
    double g1[5], g2[5], g3[5];
    for (int f = 0; f < 5; ++f) {
      g1[f] = 0.0;
      g2[f] = 0.0;
      g3[f] = 0.0;
      g1[f] = g1[f] + shg[0][0] * yl[0][f] + shg[1][0] * yl[1][f] +
              shg[2][0] * yl[2][f] + shg[3][0] * yl[3][f];
      g2[f] = g2[f] + shg[0][1] * yl[0][f] + shg[1][1] * yl[1][f] +
              shg[2][1] * yl[2][f] + shg[3][1] * yl[3][f];
      g3[f] = g3[f] + shg[0][2] * yl[0][f] + shg[1][2] * yl[1][f] +
              shg[2][2] * yl[2][f] + shg[3][2] * yl[3][f];
    }

    /* ---- e3mtrx.f:87-224 ---- */
    double A0[5][5], A1[5][5], A2[5][5], A3[5][5];
    memset(A0, 0, sizeof A0);
    memset(A1, 0, sizeof A1);
    memset(A2, 0, sizeof A2);
    memset(A3, 0, sizeof A3);
    double drdp = rho * betaT;
    double drdT = -rho * alfap;
    A0[4][0] = drdp * (h + rk) - alfap * T;
    double e2p = A0[4][0] + ONE;
    double e3p = rho * (h + rk);
    double e4p = drdT * (h + rk) + rho * cp;
    A0[0][0] = drdp;
    A0[0][4] = drdT;
    A0[1][0] = drdp * u1;
    A0[1][1] = rho;
    A0[1][4] = drdT * u1;
    A0[2][0] = drdp * u2;
    A0[2][2] = rho;
    A0[2][4] = drdT * u2;
    A0[3][0] = drdp * u3;
    A0[3][3] = rho;
    A0[3][4] = drdT * u3;
    A0[4][1] = rho * u1;
    A0[4][2] = rho * u2;
    A0[4][3] = rho * u3;
    A0[4][4] = e4p;

    A1[0][0] = drdp * u1;
    A1[0][1] = rho;
    A1[0][4] = drdT * u1;
    A1[1][0] = drdp * u1 * u1 + 1;
    A1[1][1] = TWO * rho * u1;
    A1[1][4] = drdT * u1 * u1;
    A1[2][0] = drdp * u1 * u2;
    A1[2][1] = rho * u2;
    A1[2][2] = rho * u1;
    A1[2][4] = drdT * u1 * u2;
    A1[3][0] = drdp * u1 * u3;
    A1[3][1] = rho * u3;
    A1[3][3] = rho * u1;
    A1[3][4] = drdT * u1 * u3;
    A1[4][0] = u1 * e2p;
    A1[4][1] = e3p + rho * u1 * u1;
    A1[4][2] = rho * u1 * u2;
    A1[4][3] = rho * u1 * u3;
    A1[4][4] = u1 * e4p;

    A2[0][0] = drdp * u2;
    A2[0][2] = rho;
    A2[0][4] = drdT * u2;
    A2[1][0] = drdp * u1 * u2;
    A2[1][1] = rho * u2;
    A2[1][2] = rho * u1;
    A2[1][4] = drdT * u1 * u2;
    A2[2][0] = drdp * u2 * u2 + 1;
    A2[2][2] = TWO * rho * u2;
    A2[2][4] = drdT * u2 * u2;
    A2[3][0] = drdp * u2 * u3;
    A2[3][2] = rho * u3;
    A2[3][3] = rho * u2;
    A2[3][4] = drdT * u2 * u3;
    A2[4][0] = u2 * e2p;
    A2[4][1] = rho * u1 * u2;
    A2[4][2] = e3p + rho * u2 * u2;
    A2[4][3] = rho * u2 * u3;
    A2[4][4] = u2 * e4p;

    A3[0][0] = drdp * u3;
    A3[0][3] = rho;
    A3[0][4] = drdT * u3;
    A3[1][0] = drdp * u1 * u3;
    A3[1][1] = rho * u3;
    A3[1][3] = rho * u1;
    A3[1][4] = drdT * u1 * u3;
    A3[2][0] = drdp * u3 * u2;
    A3[2][2] = rho * u3;
    A3[2][3] = rho * u2;
    A3[2][4] = drdT * u3 * u2;
    A3[3][0] = drdp * u3 * u3 + 1;
    A3[3][3] = TWO * rho * u3;
    A3[3][4] = drdT * u3 * u3;
    A3[4][0] = u3 * e2p;
    A3[4][1] = rho * u1 * u3;
    A3[4][2] = rho * u2 * u3;
    A3[4][3] = e3p + rho * u3 * u3;
    A3[4][4] = u3 * e4p;

    /* ---- e3conv.f:72-96 Euler flux (integrated by parts) ---- */
    ri[0] = (-u1) * rho;
    ri[1] = (-u1) * rho * u1 - pres;
    ri[2] = (-u1) * rho * u2;
    ri[3] = (-u1) * rho * u3;
    ri[4] = (-u1) * rho * (ei + rk) - u1 * pres;
    ri[5] = (-u2) * rho;
    ri[6] = (-u2) * rho * u1;
    ri[7] = (-u2) * rho * u2 - pres;
    ri[8] = (-u2) * rho * u3;
    ri[9] = (-u2) * rho * (ei + rk) - u2 * pres;
    ri[10] = (-u3) * rho;
    ri[11] = (-u3) * rho * u1;
    ri[12] = (-u3) * rho * u2;
    ri[13] = (-u3) * rho * u3 - pres;
    ri[14] = (-u3) * rho * (ei + rk) - u3 * pres;
    /* e3conv.f:100-179 rLyi = A_i Y,i */
    double rLyi[5];
    for (int i = 0; i < 5; ++i)
      rLyi[i] = A1[i][0] * g1[0] + A1[i][1] * g1[1] + A1[i][2] * g1[2] +
                A1[i][3] * g1[3] + A1[i][4] * g1[4] + A2[i][0] * g2[0] +
                A2[i][1] * g2[1] + A2[i][2] * g2[2] + A2[i][3] * g2[3] +
                A2[i][4] * g2[4] + A3[i][0] * g3[0] + A3[i][1] * g3[1] +
                A3[i][2] * g3[2] + A3[i][3] * g3[3] + A3[i][4] * g3[4];
    /* e3conv.f:188-345 LHS: EGmass(a,b) += N_a (WdetJ N_b,i A_i) */
    if (lhs) {
      for (int j = 0; j < 4; ++j) {
        int j0 = 5 * j;
        double f1 = WdetJ * shg[j][0], f2 = WdetJ * shg[j][1], f3 = WdetJ * shg[j][2];
        double AiNbi[5][5];
        for (int r = 0; r < 5; ++r)
          for (int c = 0; c < 5; ++c)
            AiNbi[r][c] = f1 * A1[r][c] + f2 * A2[r][c] + f3 * A3[r][c];
        for (int i = 0; i < 4; ++i) {
          int i0 = 5 * i;
          for (int jd = 0; jd < 5; ++jd)
            for (int r = 0; r < 5; ++r)
              EG[i0 + r][j0 + jd] = EG[i0 + r][j0 + jd] + shape[i] * AiNbi[r][jd];
        }
      }
    }

    /* ---- e3visc.f:63-141 diffusivity K_ij (stiff) ---- */
    if (lhs) {
      stiff[1][1] = rlm2mu;
      stiff[2][2] = rmu;
      stiff[3][3] = rmu;
      stiff[4][1] = rlm2mu * u1;
      stiff[4][2] = rmu * u2;
      stiff[4][3] = rmu * u3;
      stiff[4][4] = con;
      stiff[1][7] = rlm;
      stiff[2][6] = rmu;
      stiff[4][6] = rmu * u2;
      stiff[4][7] = rlm * u1;
      stiff[1][13] = rlm;
      stiff[3][11] = rmu;
      stiff[4][11] = rmu * u3;
      stiff[4][13] = rlm * u1;
      stiff[6][2] = rmu;
      stiff[7][1] = rlm;
      stiff[9][1] = rlm * u2;
      stiff[9][2] = rmu * u1;
      stiff[6][6] = rmu;
      stiff[7][7] = rlm2mu;
      stiff[8][8] = rmu;
      stiff[9][6] = rmu * u1;
      stiff[9][7] = rlm2mu * u2;
      stiff[9][8] = rmu * u3;
      stiff[9][9] = con;
      stiff[7][13] = rlm;
      stiff[8][12] = rmu;
      stiff[9][12] = rmu * u3;
      stiff[9][13] = rlm * u2;
      stiff[11][3] = rmu;
      stiff[13][1] = rlm;
      stiff[14][1] = rlm * u3;
      stiff[14][3] = rmu * u1;
      stiff[12][8] = rmu;
      stiff[13][7] = rlm;
      stiff[14][7] = rlm * u3;
      stiff[14][8] = rmu * u2;
      stiff[11][11] = rmu;
      stiff[12][12] = rmu;
      stiff[13][13] = rlm2mu;
      stiff[14][11] = rmu * u1;
      stiff[14][12] = rmu * u2;
      stiff[14][13] = rlm2mu * u3;
      stiff[14][14] = con;
    }
    /* e3visc.f:278-343 viscous and heat fluxes (rlsli = 0) */
    const double rls = 0.0;
    rmi[1] = rlm2mu * g1[1] + rlm * g2[2] + rlm * g3[3] - rls;
    rmi[2] = rmu * g1[2] + rmu * g2[1] - rls;
    rmi[3] = rmu * g1[3] + rmu * g3[1] - rls;
    rmi[4] = rlm2mu * u1 * g1[1] + rmu * u2 * g1[2] + rmu * u3 * g1[3] +
             rmu * u2 * g2[1] + rlm * u1 * g2[2] + rmu * u3 * g3[1] +
             rlm * u1 * g3[3] + con * g1[4];
    for (int k = 1; k <= 4; ++k) ri[k] = ri[k] + rmi[k];
    rmi[6] = rmu * g1[2] + rmu * g2[1] - rls;
    rmi[7] = rlm * g1[1] + rlm2mu * g2[2] + rlm * g3[3] - rls;
    rmi[8] = rmu * g2[3] + rmu * g3[2] - rls;
    rmi[9] = rlm * u2 * g1[1] + rmu * u1 * g1[2] + rmu * u1 * g2[1] +
             rlm2mu * u2 * g2[2] + rmu * u3 * g2[3] + rmu * u3 * g3[2] +
             rlm * u2 * g3[3] + con * g2[4];
    for (int k = 6; k <= 9; ++k) ri[k] = ri[k] + rmi[k];
    rmi[11] = rmu * g1[3] + rmu * g3[1] - rls;
    rmi[12] = rmu * g2[3] + rmu * g3[2] - rls;
    rmi[13] = rlm * g1[1] + rlm * g2[2] + rlm2mu * g3[3] - rls;
    rmi[14] = rlm * u3 * g1[1] + rmu * u1 * g1[3] + rlm * u3 * g2[2] +
              rmu * u2 * g2[3] + rmu * u1 * g3[1] + rmu * u2 * g3[2] +
              rlm2mu * u3 * g3[3] + con * g3[4];
    for (int k = 11; k <= 14; ++k) ri[k] = ri[k] + rmi[k];

    /* ---- e3ls.f:105-143 add the time term to the LS residual ---- */
    for (int i = 0; i < 5; ++i)
      rLyi[i] = rLyi[i] + A0[i][0] * aci[0] + A0[i][1] * aci[1] +
                A0[i][2] * aci[2] + A0[i][3] * aci[3] + A0[i][4] * aci[4];
    /* ---- e3ivar.f:361-395 divergence of the reconstructed diffusive flux,
     * e3ls.f:148-155 subtracted from the LS residual (isurf = 0) ---- */
    if (ql) {
      double divqi[4] = {0.0, 0.0, 0.0, 0.0};
      for (int n = 0; n < 4; ++n)
        for (int k = 0; k < 4; ++k)
          divqi[k] = divqi[k] + shg[n][0] * ql[n][k] + shg[n][1] * ql[n][4 + k] +
                     shg[n][2] * ql[n][8 + k];
      rLyi[1] = rLyi[1] - divqi[0];
      rLyi[2] = rLyi[2] - divqi[1];
      rLyi[3] = rLyi[3] - divqi[2];
      rLyi[4] = rLyi[4] - divqi[3];
    }
    /* ---- e3tau.f:1-283 (itau = 0, ipord = 1) ---- */
    double gijd[6];
    e3gijd(dxidx, gijd);
    const double fff = 36.0;
    double dts = P->dtsfct * P->Dtgl;
    double tau[3];
    double t2 = rho * rho *
                    ((TWO * dts) * (TWO * dts) +
                     (u1 * (u1 * gijd[0] + TWO * (u2 * gijd[1] + u3 * gijd[3])) +
                      u2 * (u2 * gijd[2] + TWO * u3 * gijd[4]) + u3 * u3 * gijd[5])) +
                fff * (rmu * rmu) *
                    (gijd[0] * gijd[0] + gijd[2] * gijd[2] + gijd[5] * gijd[5] +
                     TWO * (gijd[1] * gijd[1] + gijd[3] * gijd[3] + gijd[4] * gijd[4]));
    double fct = sqrt(t2);
    tau[0] = PT125 * fct / (rho * (gijd[0] + gijd[2] + gijd[5])) * P->taucfct;
    tau[1] = ONE / fct;
    tau[2] = tau[1] / cv * P->temper;
    rLyi[0] = tau[0] * rLyi[0];
    rLyi[1] = tau[1] * rLyi[1];
    rLyi[2] = tau[1] * rLyi[2];
    rLyi[3] = tau[1] * rLyi[3];
    rLyi[4] = tau[2] * rLyi[4];
    /* e3ls.f:351-458 RHS ri += A_i (tau L) */
    for (int i = 0; i < 5; ++i) ri[i] = rowdot(A1, i, rLyi) + ri[i];
    for (int i = 0; i < 5; ++i) ri[5 + i] = rowdot(A2, i, rLyi) + ri[5 + i];
    for (int i = 0; i < 5; ++i) ri[10 + i] = rowdot(A3, i, rLyi) + ri[10 + i];
    /* e3ls.f:464-766 LHS */
    if (lhs) {
      const double* Ai[3] = {&A1[0][0], &A2[0][0], &A3[0][0]};
      double AitauA0[3][5][5];
      for (int k = 0; k < 3; ++k) {
        const double (*Ak)[5] = (const double (*)[5])Ai[k];
        double Atau[5][5];
        for (int i = 0; i < 5; ++i) {
          Atau[i][0] = Ak[i][0] * tau[0];
          Atau[i][1] = Ak[i][1] * tau[1];
          Atau[i][2] = Ak[i][2] * tau[1];
          Atau[i][3] = Ak[i][3] * tau[1];
          Atau[i][4] = Ak[i][4] * tau[2];
        }
        for (int j = 0; j < 5; ++j)
          for (int i = 0; i < 5; ++i)
            AitauA0[k][i][j] = Atau[i][0] * A0[0][j] + Atau[i][1] * A0[1][j] +
                               Atau[i][2] * A0[2][j] + Atau[i][3] * A0[3][j] +
                               Atau[i][4] * A0[4][j];
        for (int l = 0; l < 3; ++l) {
          const double (*Al)[5] = (const double (*)[5])Ai[l];
          for (int j = 0; j < 5; ++j)
            for (int i = 0; i < 5; ++i)
              stiff[i + 5 * k][j + 5 * l] =
                  stiff[i + 5 * k][j + 5 * l] +
                  (Atau[i][0] * Al[0][j] + Atau[i][1] * Al[1][j] + Atau[i][2] * Al[2][j] +
                   Atau[i][3] * Al[3][j] + Atau[i][4] * Al[4][j]);
        }
      }
      /* e3ls.f:713-760 least-squares time term */
      for (int i = 0; i < 4; ++i) {
        int i0 = 5 * i;
        double At[5][5];
        for (int id = 0; id < 5; ++id)
          for (int jd = 0; jd < 5; ++jd)
            At[id][jd] = shg[i][0] * AitauA0[0][id][jd] + shg[i][1] * AitauA0[1][id][jd] +
                         shg[i][2] * AitauA0[2][id][jd];
        for (int j = 0; j < 4; ++j) {
          int j0 = 5 * j;
          double fa = shape[j] * WdetJ * P->almi / P->gami / P->alfi * P->Dtgl;
          for (int id = 0; id < 5; ++id)
            for (int jd = 0; jd < 5; ++jd)
              EG[i0 + id][j0 + jd] = EG[i0 + id][j0 + jd] + fa * At[id][jd];
        }
      }
    }
    /* ---- e3massr.f:34-72 ---- */
    for (int i = 0; i < 5; ++i)
      ri[15 + i] = ri[15 + i] + A0[i][0] * aci[0] + A0[i][1] * aci[1] +
                   A0[i][2] * aci[2] + A0[i][3] * aci[3] + A0[i][4] * aci[4];
    /* ---- e3massl.f:29-71 (bcool = 0) ---- */
    if (lhs) {
      double temp = WdetJ * P->almi / P->gami / P->alfi * P->Dtgl;
      double bcool = WdetJ * 0.0;
      const double sponge = 0.0;
      for (int j = 0; j < 4; ++j) {
        int j0 = 5 * j;
        for (int i = 0; i < 4; ++i) {
          int i0 = 5 * i;
          double shpij = shape[i] * shape[j];
          double fa = shpij * (temp + bcool * sponge);
          for (int jd = 0; jd < 5; ++jd)
            for (int f = 0; f < 5; ++f)
              EG[i0 + f][j0 + jd] = EG[i0 + f][j0 + jd] + fa * A0[f][jd];
        }
      }
    }
    /* ---- e3wmlt.f:72-143 RHS ---- */
    for (int i = 0; i < 4; ++i)
      for (int f = 0; f < 5; ++f)
        rl[i][f] = rl[i][f] + WdetJ * (shg[i][0] * ri[f] + shg[i][1] * ri[5 + f] +
                                       shg[i][2] * ri[10 + f]);
    for (int i = 0; i < 4; ++i)
      for (int f = 0; f < 5; ++f)
        rl[i][f] = rl[i][f] + shape[i] * WdetJ * ri[15 + f];
    /* ---- e3wmlt.f:150-227 LHS stiffness ---- */
    if (lhs) {
      for (int j = 0; j < 4; ++j) {
        int j0 = 5 * j;
        double s1 = WdetJ * shg[j][0], s2 = WdetJ * shg[j][1], s3 = WdetJ * shg[j][2];
        double st1[5][5], st2[5][5], st3[5][5];
        for (int jd = 0; jd < 5; ++jd)
          for (int id = 0; id < 5; ++id) {
            st1[id][jd] = s1 * stiff[id][jd] + s2 * stiff[id][jd + 5] + s3 * stiff[id][jd + 10];
            st2[id][jd] = s1 * stiff[id + 5][jd] + s2 * stiff[id + 5][jd + 5] +
                          s3 * stiff[id + 5][jd + 10];
            st3[id][jd] = s1 * stiff[id + 10][jd] + s2 * stiff[id + 10][jd + 5] +
                          s3 * stiff[id + 10][jd + 10];
          }
        for (int i = 0; i < 4; ++i) {
          int i0 = 5 * i;
          for (int jd = 0; jd < 5; ++jd)
            for (int r = 0; r < 5; ++r)
              EG[i0 + r][j0 + jd] = EG[i0 + r][j0 + jd] + shg[i][0] * st1[r][jd] +
                                    shg[i][1] * st2[r][jd] + shg[i][2] * st3[r][jd];
        }
      }
    }
  }
}

/* e3b.f:111-300 with e3bvar.f:86-375, iBCB = 0 (floating fluxes) */
/* e3q.f:1-246 with e3qvar.f (idiff = 1 or 3, isurf = 0): one element's
 * diffusive flux residual ql[a][k] and lumped mass rmassl[a] */
static void e3q(const orc_params* P, const double yl[4][5], const double xl[4][3],
                double ql[4][12], double rmassl[4]) {
  const quad_t* Q = quad();
  const double Rgas = P->Rgas, gamma = P->gamma, gamma1 = P->gamma1;
  memset(ql, 0, sizeof(double) * 48);
  memset(rmassl, 0, sizeof(double) * 4);
  double alph1 = 0.0, alph2 = 0.0;
  for (int q = 0; q < 4; ++q) {
    if (Q->Qwt[q] == 0.0) continue;
    double shape[4], sg[3][4];
    for (int a = 0; a < 4; ++a) {
      shape[a] = Q->shp[q][a];
      for (int j = 0; j < 3; ++j) sg[j][a] = Q->shgl[q][j][a];
    }
    /* e3qvar.f: primitive variables, getthm (ithm = 7), metrics, gradients */
    double pres = 0, u1 = 0, u2 = 0, u3 = 0, T = 0;
    for (int n = 0; n < 4; ++n) {
      pres = pres + shape[n] * yl[n][0];
      u1 = u1 + shape[n] * yl[n][1];
      u2 = u2 + shape[n] * yl[n][2];
      u3 = u3 + shape[n] * yl[n][3];
      T = T + shape[n] * yl[n][4];
    }
    double cp = Rgas * gamma / gamma1;
    double dxidx[3][3], shg[4][3], WdetJ;
    e3metric(xl, sg, Q->Qwt[q], dxidx, shg, &WdetJ);
    double g1[5], g2[5], g3[5];
    for (int f = 0; f < 5; ++f) {
      g1[f] = 0.0;
      g2[f] = 0.0;
      g3[f] = 0.0;
    }
    for (int n = 0; n < 4; ++n)
      for (int f = 0; f < 5; ++f) {
        g1[f] = g1[f] + shg[n][0] * yl[n][f];
        g2[f] = g2[f] + shg[n][1] * yl[n][f];
        g3[f] = g3[f] + shg[n][2] * yl[n][f];
      }
    /* getdiff.f (constant mu, Stokes lambda, DNS) */
    double rmu = P->mu;
    double rlm = -PT66 * rmu;
    double con = rmu * cp / P->pr;
    double xmut = 0.0;
    rlm = rlm - PT66 * xmut;
    rmu = rmu + xmut;
    double rlm2mu = rlm + TWO * rmu;
    con = con + xmut * cp / P->pr;
    /* e3q.f:109-163 diffusive flux at the point */
    double qdi[12];
    qdi[0] = rlm2mu * g1[1] + rlm * g2[2] + rlm * g3[3];
    qdi[1] = rmu * g1[2] + rmu * g2[1];
    qdi[2] = rmu * g1[3] + rmu * g3[1];
    qdi[3] = rlm2mu * u1 * g1[1] + rmu * u2 * g1[2] + rmu * u3 * g1[3] + rmu * u2 * g2[1] +
             rlm * u1 * g2[2] + rmu * u3 * g3[1] + rlm * u1 * g3[3] + con * g1[4];
    qdi[4] = rmu * g1[2] + rmu * g2[1];
    qdi[5] = rlm * g1[1] + rlm2mu * g2[2] + rlm * g3[3];
    qdi[6] = rmu * g2[3] + rmu * g3[2];
    qdi[7] = rlm * u2 * g1[1] + rmu * u1 * g1[2] + rmu * u1 * g2[1] + rlm2mu * u2 * g2[2] +
             rmu * u3 * g2[3] + rmu * u3 * g3[2] + rlm * u2 * g3[3] + con * g2[4];
    qdi[8] = rmu * g1[3] + rmu * g3[1];
    qdi[9] = rmu * g2[3] + rmu * g3[2];
    qdi[10] = rlm * g1[1] + rlm * g2[2] + rlm2mu * g3[3];
    qdi[11] = rlm * u3 * g1[1] + rmu * u1 * g1[3] + rlm * u3 * g2[2] + rmu * u2 * g2[3] +
              rmu * u1 * g3[1] + rmu * u2 * g3[2] + rlm2mu * u3 * g3[3] + con * g3[4];
    /* e3q.f:166-180 */
    for (int i = 0; i < 4; ++i)
      for (int k = 0; k < 12; ++k) ql[i][k] = ql[i][k] + shape[i] * WdetJ * qdi[k];
    /* e3q.f:185-205 row-sum (idiff 1) or special lumping (idiff 3) */
    if (P->idiff == 1)
      for (int i = 0; i < 4; ++i) rmassl[i] = rmassl[i] + shape[i] * WdetJ;
    if (P->idiff == 3) {
      double shpsum = 0.0;
      for (int i = 0; i < 4; ++i) {
        shpsum = shpsum + shape[i] * shape[i];
        rmassl[i] = rmassl[i] + shape[i] * shape[i] * WdetJ;
      }
      alph1 = alph1 + WdetJ;
      alph2 = alph2 + shpsum * WdetJ;
    }
  }
  /* e3q.f:236-240 */
  if (P->idiff == 3)
    for (int i = 0; i < 4; ++i) rmassl[i] = rmassl[i] * alph1 / alph2;
}

int orc_asiq(const orc_mesh* m, const orc_params* P, const double* y,
             double* qres, double* rmass) {
  if (P->idiff != 1 && P->idiff != 3) return 1;
  const long long nshg = m->nshg, nel = m->nel;
  memset(qres, 0, sizeof(double) * 12 * nshg);
  memset(rmass, 0, sizeof(double) * nshg);
  const int bk = m->ibksiz > 0 ? m->ibksiz : 64;
  double(*ql)[4][12] = malloc(sizeof(double) * 48 * bk);
  double(*rml)[4] = malloc(sizeof(double) * 4 * bk);
  int(*nod)[4] = malloc(sizeof(int) * 4 * bk);
  if (!ql || !rml || !nod) return 4;
  for (long long e0 = 0; e0 < nel; e0 += bk) {
    int npro = (int)((nel - e0) < bk ? (nel - e0) : bk);
    for (int e = 0; e < npro; ++e) {
      double yl[4][5], xl[4][3];
      for (int a = 0; a < 4; ++a) {
        long long g = m->ien[(e0 + e) + nel * a] - 1;
        nod[e][a] = (int)g;
        yl[a][1] = y[g];
        yl[a][2] = y[g + nshg];
        yl[a][3] = y[g + 2 * nshg];
        yl[a][0] = y[g + 3 * nshg];
        yl[a][4] = y[g + 4 * nshg];
        for (int c = 0; c < 3; ++c) xl[a][c] = m->x[g + nshg * c];
      }
      e3q(P, (const double(*)[5])yl, (const double(*)[3])xl, ql[e], rml[e]);
    }
    /* asiq.f:62-63, local.f:67-74 scatter order */
    for (int k = 0; k < 12; ++k)
      for (int i = 0; i < 4; ++i)
        for (int e = 0; e < npro; ++e)
          qres[nod[e][i] + nshg * k] = qres[nod[e][i] + nshg * k] + ql[e][i][k];
    for (int i = 0; i < 4; ++i)
      for (int e = 0; e < npro; ++e) rmass[nod[e][i]] = rmass[nod[e][i]] + rml[e][i];
  }
  free(ql);
  free(rml);
  free(nod);
  /* qpbc.f: periodic fold, copy back, invert the mass, scale */
  for (long long j = 0; j < nshg; ++j)
    if (m->iBC[j] & 1024) {
      long long i = m->iper[j] - 1;
      rmass[i] = rmass[i] + rmass[j];
      for (int k = 0; k < 12; ++k) qres[i + nshg * k] = qres[i + nshg * k] + qres[j + nshg * k];
    }
  for (long long j = 0; j < nshg; ++j)
    if (m->iBC[j] & 1024) {
      long long i = m->iper[j] - 1;
      rmass[j] = rmass[i];
      for (int k = 0; k < 12; ++k) qres[j + nshg * k] = qres[i + nshg * k];
    }
  for (long long n = 0; n < nshg; ++n) rmass[n] = ONE / rmass[n];
  for (int k = 0; k < 12; ++k)
    for (long long n = 0; n < nshg; ++n) qres[n + nshg * k] = rmass[n] * qres[n + nshg * k];
  return 0;
}

void orc_e3b(const orc_params* P, const double yl[4][5], const double xl[4][3],
             double rl[4][5]) {
  const quad_t* Q = quad();
  const double Rgas = P->Rgas, gamma = P->gamma, gamma1 = P->gamma1;
  memset(rl, 0, sizeof(double) * 20);
  for (int q = 0; q < 3; ++q) {
    if (Q->Qwtb[q] == 0.0) continue;
    double shape[4], sg[3][4];
    for (int a = 0; a < 4; ++a) {
      shape[a] = Q->shpb[q][a];
      for (int j = 0; j < 3; ++j) sg[j][a] = Q->shglb[q][j][a];
    }
    /* e3bvar.f:94-112 (nshlb = 3 face nodes) */
    double pres = 0, u1 = 0, u2 = 0, u3 = 0, T = 0;
    for (int n = 0; n < 3; ++n) {
      pres = pres + shape[n] * yl[n][0];
      u1 = u1 + shape[n] * yl[n][1];
      u2 = u2 + shape[n] * yl[n][2];
      u3 = u3 + shape[n] * yl[n][3];
      T = T + shape[n] * yl[n][4];
    }
    double rk = PT5 * (u1 * u1 + u2 * u2 + u3 * u3);
    double rho = pres / (Rgas * T);
    double ei = T * (Rgas / gamma1);
    double cp = Rgas * gamma / gamma1;
    /* e3bvar.f:131-146 deformation gradient */
    double d[3][3] = {{0}};
    for (int n = 0; n < 4; ++n)
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) d[i][j] = d[i][j] + xl[n][i] * sg[j][n];
    /* e3bvar.f:150-193 normal and weighted area (lcsyst = 1) */
    double v1[3], v2[3];
    for (int c = 0; c < 3; ++c) {
      v1[c] = xl[1][c] - xl[0][c];
      v2[c] = xl[2][c] - xl[0][c];
    }
    double t1 = v1[1] * v2[2] - v2[1] * v1[2];
    double t2 = v2[0] * v1[2] - v1[0] * v2[2];
    double t3 = v1[0] * v2[1] - v2[0] * v1[1];
    double tmp = ONE / sqrt(t1 * t1 + t2 * t2 + t3 * t3);
    double bn[3] = {t1 * tmp, t2 * tmp, t3 * tmp};
    double WdetJb = Q->Qwtb[q] / (4.0 * tmp);
    /* e3bvar.f:199-226 inverse deformation gradient */
    double dx[3][3];
    dx[0][0] = d[1][1] * d[2][2] - d[2][1] * d[1][2];
    dx[0][1] = d[2][1] * d[0][2] - d[0][1] * d[2][2];
    dx[0][2] = d[0][1] * d[1][2] - d[0][2] * d[1][1];
    double tt = ONE / (dx[0][0] * d[0][0] + dx[0][1] * d[1][0] + dx[0][2] * d[2][0]);
    dx[0][0] = dx[0][0] * tt;
    dx[0][1] = dx[0][1] * tt;
    dx[0][2] = dx[0][2] * tt;
    dx[1][0] = (d[1][2] * d[2][0] - d[1][0] * d[2][2]) * tt;
    dx[1][1] = (d[0][0] * d[2][2] - d[2][0] * d[0][2]) * tt;
    dx[1][2] = (d[1][0] * d[0][2] - d[0][0] * d[1][2]) * tt;
    dx[2][0] = (d[1][0] * d[2][1] - d[1][1] * d[2][0]) * tt;
    dx[2][1] = (d[2][0] * d[0][1] - d[0][0] * d[2][1]) * tt;
    dx[2][2] = (d[0][0] * d[1][1] - d[0][1] * d[1][0]) * tt;
    /* e3bvar.f:231-300 local then global gradients, dofs 2..5 */
    double gl[3][5] = {{0}};
    for (int n = 0; n < 4; ++n)
      for (int k = 0; k < 3; ++k)
        for (int f = 0; f < 5; ++f) gl[k][f] = gl[k][f] + sg[k][n] * yl[n][f];
    double g1[5] = {0}, g2[5] = {0}, g3[5] = {0};
    for (int f = 1; f < 5; ++f) {
      g1[f] = dx[0][0] * gl[0][f] + dx[1][0] * gl[1][f] + dx[2][0] * gl[2][f];
      g2[f] = dx[0][1] * gl[0][f] + dx[1][1] * gl[1][f] + dx[2][1] * gl[2][f];
      g3[f] = dx[0][2] * gl[0][f] + dx[1][2] * gl[1][f] + dx[2][2] * gl[2][f];
    }
    /* e3b.f:146-171 convective and pressure flux */
    double un = bn[0] * u1 + bn[1] * u2 + bn[2] * u3;
    double rou = rho * (un);
    double p = pres;
    double F1 = rou;
    double F2 = rou * u1 + bn[0] * p;
    double F3 = rou * u2 + bn[1] * p;
    double F4 = rou * u3 + bn[2] * p;
    double F5 = rou * (ei + rk) + un * p;
    /* getDiff for the wall point */
    double rmu = P->mu;
    double rlm = -PT66 * rmu;
    double con = rmu * cp / P->pr;
    rlm = rlm - PT66 * 0.0;
    rmu = rmu + 0.0;
    double rlm2mu = rlm + TWO * rmu;
    con = con + 0.0 * cp / P->pr;
    /* e3b.f:198-228 floating viscous traction and heat flux */
    double tau1n = bn[0] * (rlm2mu * g1[1] + rlm * g2[2] + rlm * g3[3]) +
                   bn[1] * (rmu * (g2[1] + g1[2])) + bn[2] * (rmu * (g3[1] + g1[3]));
    double tau2n = bn[0] * (rmu * (g2[1] + g1[2])) +
                   bn[1] * (rlm * g1[1] + rlm2mu * g2[2] + rlm * g3[3]) +
                   bn[2] * (rmu * (g3[2] + g2[3]));
    double tau3n = bn[0] * (rmu * (g3[1] + g1[3])) + bn[1] * (rmu * (g3[2] + g2[3])) +
                   bn[2] * (rlm * g1[1] + rlm * g2[2] + rlm2mu * g3[3]);
    double Fv2 = tau1n, Fv3 = tau2n, Fv4 = tau3n;
    double Fv5 = u1 * Fv2 + u2 * Fv3 + u3 * Fv4;
    double heat = -con * (bn[0] * g1[4] + bn[1] * g2[4] + bn[2] * g3[4]);
    double Fh5 = heat;
    F2 = F2 - Fv2;
    F3 = F3 - Fv3;
    F4 = F4 - Fv4;
    F5 = F5 - Fv5 + Fh5;
    /* e3b.f:266-290 */
    for (int n = 0; n < 3; ++n) {
      rl[n][0] = rl[n][0] + WdetJb * shape[n] * F1;
      rl[n][1] = rl[n][1] + WdetJb * shape[n] * F2;
      rl[n][2] = rl[n][2] + WdetJb * shape[n] * F3;
      rl[n][3] = rl[n][3] + WdetJb * shape[n] * F4;
      rl[n][4] = rl[n][4] + WdetJb * shape[n] * F5;
    }
  }
}

/* ---- boundary-condition helpers (bc3lhs.f, bc3res.f, bc3bdg.f) ---- */
static int vel_code(int ibc) { return (ibc >> 3) & 7; }
static int bc_supported(int ibc) {
  int v = vel_code(ibc);
  if (ibc & 1) return 0;           /* density: not in this path */
  if (ibc & (1 << 11)) return 0;   /* scaled plane extraction   */
  return v == 0 || v == 7;
}

/* bc3lhs.f:39-290 for one element */
static void bc3lhs_elem(const int* iBC, const int nodes[4], double EG[20][20]) {
  for (int a = 0; a < 4; ++a) {
    int ibc = iBC[nodes[a]];
    if (ibc == 0) continue;
    int off = 5 * a;
    int cons[5] = {0, 0, 0, 0, 0};
    if (ibc & 4) cons[0] = 1;
    if (vel_code(ibc) == 7) cons[1] = cons[2] = cons[3] = 1;
    if (ibc & 2) cons[4] = 1;
    for (int d = 0; d < 5; ++d) {
      if (!cons[d]) continue;
      int c = off + d;
      for (int i = 0; i < 20; ++i) {
        EG[i][c] = 0.0;
        EG[c][i] = 0.0;
      }
      EG[c][c] = ONE;
    }
  }
}

/* fillsparse.f:243-270 */
static long long sparseloc(const int* list, int n, int target) {
  int rowvl = 1, rowvh = n + 1;
  while (rowvh - rowvl > 1) {
    int rowv = (rowvh + rowvl) / 2;
    if (list[rowv - 1] > target)
      rowvh = rowv;
    else
      rowvl = rowv;
  }
  return rowvl;
}

int orc_elmgmrs(const orc_mesh* m, const orc_params* P, const double* y,
                const double* ac, double* res, double* rmes, double* BDiag,
                double* lhsK) {
  const long long nshg = m->nshg, nel = m->nel, nelb = m->nelb;
  for (long long n = 0; n < nshg; ++n)
    if (!bc_supported(m->iBC[n])) return 3;
  if (P->idiff != 0 && P->idiff != 1 && P->idiff != 3) return 5;
  /* elmgmr.f:361-391 global reconstruction of the diffusive flux */
  double* qres = NULL;
  if (P->idiff != 0) {
    qres = malloc(sizeof(double) * 13 * nshg);
    if (!qres) return 4;
    int rc = orc_asiq(m, P, y, qres, qres + 12 * nshg);
    if (rc) {
      free(qres);
      return rc;
    }
  }
  /* elmgmr.f:399-403 */
  memset(res, 0, sizeof(double) * nshg * 5);
  if (rmes) memset(rmes, 0, sizeof(double) * nshg * 5);
  if (P->lhs == 1) memset(lhsK, 0, sizeof(double) * 25 * m->nnz_tot);
  if (P->iprec != 0) memset(BDiag, 0, sizeof(double) * 25 * nshg);
  const int bk = m->ibksiz > 0 ? m->ibksiz : 64;
  /* Threads (OpenMP) evaluate e3 for kSuper element blocks at a time into
   * per-element buffers; the scatter then runs in the reference's order
   * (block by block, local.f:67-74 and fillsparse.f:101-126), each thread
   * applying only the contributions to its own range of rows, so every
   * entry of res, BDiag and lhsK receives its terms in the same sequence as
   * a serial run: the result is bitwise independent of the thread count. */
  enum { kSuper = 64 };
  const long long cap = (long long)kSuper * bk;
  double(*rl)[4][5] = malloc(sizeof(double) * 20 * cap);
  double(*EGb)[20][20] = malloc(sizeof(double) * 400 * cap);
  double(*BDl)[4][25] = malloc(sizeof(double) * 100 * cap);
  int(*nod)[4] = malloc(sizeof(int) * 4 * cap);
  long long* kloc = P->lhs == 1 ? malloc(sizeof(long long) * 16 * cap) : NULL;
  if (!rl || !EGb || !BDl || !nod || (P->lhs == 1 && !kloc)) return 4;
  for (long long s0 = 0; s0 < nel; s0 += cap) {
    const long long ns = (nel - s0) < cap ? (nel - s0) : cap;
#pragma omp parallel for schedule(dynamic, 1)
    for (long long e = 0; e < ns; ++e) {
      double yl[4][5], acl[4][5], xl[4][3];
      for (int a = 0; a < 4; ++a) {
        long long g = m->ien[(s0 + e) + nel * a] - 1;
        nod[e][a] = (int)g;
        /* localy.f:53-73: (u,v,w,p,T) -> (p,u,v,w,T) */
        yl[a][1] = y[g];
        yl[a][2] = y[g + nshg];
        yl[a][3] = y[g + 2 * nshg];
        yl[a][0] = y[g + 3 * nshg];
        yl[a][4] = y[g + 4 * nshg];
        acl[a][1] = ac[g];
        acl[a][2] = ac[g + nshg];
        acl[a][3] = ac[g + 2 * nshg];
        acl[a][0] = ac[g + 3 * nshg];
        acl[a][4] = ac[g + 4 * nshg];
        for (int c = 0; c < 3; ++c) xl[a][c] = m->x[g + nshg * c];
      }
      double qll[4][12];
      if (qres) /* asigmr.f:55 */
        for (int a = 0; a < 4; ++a)
          for (int k = 0; k < 12; ++k) qll[a][k] = qres[nod[e][a] + nshg * k];
      orc_e3_q(P, (const double(*)[5])yl, (const double(*)[5])acl,
               (const double(*)[3])xl, qres ? (const double(*)[12])qll : NULL, rl[e], EGb[e],
               P->lhs == 1);
      /* asigmr.f:92-103 BDiagl(:,i,j,k) = EGmass(:,5(i-1)+j,5(i-1)+k) */
      if (P->iprec != 0)
        for (int i = 0; i < 4; ++i)
          for (int k = 0; k < 5; ++k)
            for (int j = 0; j < 5; ++j) BDl[e][i][j + 5 * k] = EGb[e][5 * i + j][5 * i + k];
      if (P->lhs == 1) {
        bc3lhs_elem(m->iBC, nod[e], EGb[e]);
        /* sparseloc (fillsparse.f:243-270) of the element's 16 blocks */
        for (int aa = 0; aa < 4; ++aa) {
          const int i = nod[e][aa];
          const int c = m->colm[i];
          const int n = m->colm[i + 1] - c;
          for (int b = 0; b < 4; ++b)
            kloc[16 * e + 4 * aa + b] = sparseloc(m->rowp + (c - 1), n, nod[e][b] + 1) + c - 1;
        }
      }
    }
#pragma omp parallel
    {
      int nt = 1, t = 0;
#ifdef _OPENMP
      nt = omp_get_num_threads();
      t = omp_get_thread_num();
#endif
      const long long r0 = nshg * t / nt, r1 = nshg * (t + 1) / nt;
#define MINE(node) ((node) >= r0 && (node) < r1)
      for (long long e0 = 0; e0 < ns; e0 += bk) {
        const int npro = (int)((ns - e0) < bk ? (ns - e0) : bk);
        /* local.f:67-74 scatter order: dof, local node, element */
        for (int f = 0; f < 5; ++f)
          for (int i = 0; i < 4; ++i)
            for (int e = (int)e0; e < e0 + npro; ++e)
              if (MINE(nod[e][i]))
                res[nod[e][i] + nshg * f] = res[nod[e][i] + nshg * f] + rl[e][i][f];
        if (P->iprec != 0)
          for (int jk = 0; jk < 25; ++jk)
            for (int i = 0; i < 4; ++i)
              for (int e = (int)e0; e < e0 + npro; ++e)
                if (MINE(nod[e][i]))
                  BDiag[nod[e][i] + nshg * jk] = BDiag[nod[e][i] + nshg * jk] + BDl[e][i][jk];
        if (P->lhs == 1)
          /* fillsparse.f:101-126 */
          for (int e = (int)e0; e < e0 + npro; ++e)
            for (int aa = 0; aa < 4; ++aa) {
              if (!MINE(nod[e][aa])) continue;
              for (int b = 0; b < 4; ++b) {
                double* blk = lhsK + 25 * (kloc[16 * e + 4 * aa + b] - 1);
                for (int g = 0; g < 5; ++g)
                  for (int f = 0; f < 5; ++f)
                    blk[5 * g + f] = blk[5 * g + f] + EGb[e][5 * aa + f][5 * b + g];
              }
            }
      }
#undef MINE
    }
  }
  free(kloc);
  /* boundary elements: asbmfg.f:39-56, blocks of ibksiz (genbkb.f) */
  for (long long e0 = 0; e0 < nelb; e0 += bk) {
    int npro = (int)((nelb - e0) < bk ? (nelb - e0) : bk);
    for (int e = 0; e < npro; ++e) {
      double yl[4][5], xl[4][3];
      for (int a = 0; a < 4; ++a) {
        long long g = m->ienb[(e0 + e) + nelb * a] - 1;
        nod[e][a] = (int)g;
        yl[a][1] = y[g];
        yl[a][2] = y[g + nshg];
        yl[a][3] = y[g + 2 * nshg];
        yl[a][0] = y[g + 3 * nshg];
        yl[a][4] = y[g + 4 * nshg];
        for (int c = 0; c < 3; ++c) xl[a][c] = m->x[g + nshg * c];
      }
      orc_e3b(P, (const double(*)[5])yl, (const double(*)[3])xl, rl[e]);
    }
    for (int f = 0; f < 5; ++f)
      for (int i = 0; i < 4; ++i)
        for (int e = 0; e < npro; ++e)
          res[nod[e][i] + nshg * f] = res[nod[e][i] + nshg * f] + rl[e][i][f];
  }
  free(rl);
  free(EGb);
  free(BDl);
  free(nod);
  free(qres);

  /* bc3res.f:25-163 */
  for (long long n = 0; n < nshg; ++n) {
    int ibc = m->iBC[n];
    if (ibc & 4) res[n] = 0.0;
    if (vel_code(ibc) == 7) {
      res[n + nshg] = 0.0;
      res[n + 2 * nshg] = 0.0;
      res[n + 3 * nshg] = 0.0;
    }
    if (ibc & 2) res[n + 4 * nshg] = 0.0;
  }
  for (long long j = 0; j < nshg; ++j)
    if (m->iBC[j] & 1024) {
      long long i = m->iper[j] - 1;
      for (int f = 0; f < 5; ++f) {
        res[i + nshg * f] = res[i + nshg * f] + res[j + nshg * f];
        res[j + nshg * f] = 0.0;
      }
    }
  /* bc3bdg.f:25-352 */
  if (P->iprec != 0) {
    for (long long n = 0; n < nshg; ++n) {
      int ibc = m->iBC[n];
      int cons[5] = {(ibc & 4) != 0, vel_code(ibc) == 7, vel_code(ibc) == 7,
                     vel_code(ibc) == 7, (ibc & 2) != 0};
      for (int d = 0; d < 5; ++d) {
        if (!cons[d]) continue;
        for (int o = 0; o < 5; ++o) {
          BDiag[n + nshg * (d + 5 * o)] = 0.0;
          BDiag[n + nshg * (o + 5 * d)] = 0.0;
        }
        BDiag[n + nshg * (d + 5 * d)] = ONE;
      }
    }
    for (long long j = 0; j < nshg; ++j)
      if (m->iBC[j] & 1024) {
        long long i = m->iper[j] - 1;
        for (int jk = 0; jk < 25; ++jk)
          BDiag[i + nshg * jk] = BDiag[i + nshg * jk] + BDiag[j + nshg * jk];
      }
    for (long long j = 0; j < nshg; ++j)
      if (m->iBC[j] & 1024) {
        long long i = m->iper[j] - 1;
        for (int jk = 0; jk < 25; ++jk) BDiag[j + nshg * jk] = BDiag[i + nshg * jk];
      }
  }
  return 0;
}

/* i3lu.f:33-171 */
#define DG(a, b) Diag[n + nshg * ((a - 1) + 5 * (b - 1))]
#define RR(a) r[n + nshg * (a - 1)]
int orc_i3lu(long long nshg, double* Diag, double* r, int code) {
  if (code == 0) {
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < nshg; ++n) {
      DG(1, 1) = ONE / DG(1, 1);
      DG(2, 1) = DG(1, 1) * DG(2, 1);
      DG(3, 1) = DG(1, 1) * DG(3, 1);
      DG(4, 1) = DG(1, 1) * DG(4, 1);
      DG(5, 1) = DG(1, 1) * DG(5, 1);
      DG(2, 2) = DG(2, 2) - DG(2, 1) * DG(1, 2);
      DG(2, 3) = DG(2, 3) - DG(2, 1) * DG(1, 3);
      DG(2, 4) = DG(2, 4) - DG(2, 1) * DG(1, 4);
      DG(2, 5) = DG(2, 5) - DG(2, 1) * DG(1, 5);
      DG(2, 2) = ONE / DG(2, 2);
      DG(3, 2) = DG(2, 2) * (DG(3, 2) - DG(3, 1) * DG(1, 2));
      DG(4, 2) = DG(2, 2) * (DG(4, 2) - DG(4, 1) * DG(1, 2));
      DG(5, 2) = DG(2, 2) * (DG(5, 2) - DG(5, 1) * DG(1, 2));
      DG(3, 3) = DG(3, 3) - DG(3, 1) * DG(1, 3) - DG(3, 2) * DG(2, 3);
      DG(3, 4) = DG(3, 4) - DG(3, 1) * DG(1, 4) - DG(3, 2) * DG(2, 4);
      DG(3, 5) = DG(3, 5) - DG(3, 1) * DG(1, 5) - DG(3, 2) * DG(2, 5);
      DG(3, 3) = ONE / DG(3, 3);
      DG(4, 3) = DG(3, 3) * (DG(4, 3) - DG(4, 1) * DG(1, 3) - DG(4, 2) * DG(2, 3));
      DG(5, 3) = DG(3, 3) * (DG(5, 3) - DG(5, 1) * DG(1, 3) - DG(5, 2) * DG(2, 3));
      DG(4, 4) = DG(4, 4) - DG(4, 1) * DG(1, 4) - DG(4, 2) * DG(2, 4) - DG(4, 3) * DG(3, 4);
      DG(4, 4) = ONE / DG(4, 4);
      DG(5, 4) = DG(4, 4) * (DG(5, 4) - DG(5, 1) * DG(1, 4) - DG(5, 2) * DG(2, 4) -
                             DG(5, 3) * DG(3, 4));
      DG(5, 5) = DG(5, 5) - DG(5, 1) * DG(1, 5) - DG(5, 2) * DG(2, 5) -
                 DG(5, 3) * DG(3, 5) - DG(5, 4) * DG(4, 5);
      DG(5, 5) = ONE / DG(5, 5);
    }
    return 0;
  }
  if (code == 1) {
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < nshg; ++n) {
      RR(2) = RR(2) - DG(2, 1) * RR(1);
      RR(3) = RR(3) - DG(3, 1) * RR(1) - DG(3, 2) * RR(2);
      RR(4) = RR(4) - DG(4, 1) * RR(1) - DG(4, 2) * RR(2) - DG(4, 3) * RR(3);
      RR(5) = RR(5) - DG(5, 1) * RR(1) - DG(5, 2) * RR(2) - DG(5, 3) * RR(3) - DG(5, 4) * RR(4);
    }
    return 0;
  }
  if (code == 2) {
#pragma omp parallel for schedule(static)
    for (long long n = 0; n < nshg; ++n) {
      RR(5) = DG(5, 5) * RR(5);
      RR(4) = DG(4, 4) * (RR(4) - RR(5) * DG(4, 5));
      RR(3) = DG(3, 3) * (RR(3) - RR(5) * DG(3, 5) - RR(4) * DG(3, 4));
      RR(2) = DG(2, 2) * (RR(2) - RR(5) * DG(2, 5) - RR(4) * DG(2, 4) - RR(3) * DG(2, 3));
      RR(1) = DG(1, 1) * (RR(1) - RR(5) * DG(1, 5) - RR(4) * DG(1, 4) - RR(3) * DG(1, 3) -
                          RR(2) * DG(1, 2));
    }
    return 0;
  }
  if (code == 3) {
    for (long long n = 0; n < nshg; ++n) {
      RR(1) = RR(1) / DG(1, 1) + RR(2) * DG(1, 2) + RR(3) * DG(1, 3) + RR(4) * DG(1, 4) +
              RR(5) * DG(1, 5);
      RR(2) = RR(2) / DG(2, 2) + RR(3) * DG(2, 3) + RR(4) * DG(2, 4) + RR(5) * DG(2, 5);
      RR(3) = RR(3) / DG(3, 3) + RR(4) * DG(3, 4) + RR(5) * DG(3, 5);
      RR(4) = RR(4) / DG(4, 4) + RR(5) * DG(4, 5);
      RR(5) = RR(5) / DG(5, 5);
    }
    return 0;
  }
  return 1;
}
#undef RR

/* spsi3pre.f:46-217; BDiag is LU-factored by i3LU */
#define BD(i, a, b) BDiag[(i) + nshg * ((a - 1) + 5 * (b - 1))]
#define LK(e) L[(e)-1]
void orc_spsi3pre(long long nshg, const double* BDiag, double* lhsK,
                  const int* colm, const int* rowp) {
  /* rows are independent: threads over rows leave every value bitwise
   * equal to a serial run */
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < nshg; ++i)
    for (long long k = colm[i] - 1; k < colm[i + 1] - 1; ++k) {
      double* L = lhsK + 25 * k;
      for (int c = 0; c < 5; ++c) {
        int o = 5 * c; /* column c of the block: entries o+1..o+5 */
        LK(o + 2) = LK(o + 2) - BD(i, 2, 1) * LK(o + 1);
        LK(o + 3) = LK(o + 3) - BD(i, 3, 1) * LK(o + 1) - BD(i, 3, 2) * LK(o + 2);
        LK(o + 4) = LK(o + 4) - BD(i, 4, 1) * LK(o + 1) - BD(i, 4, 2) * LK(o + 2) -
                    BD(i, 4, 3) * LK(o + 3);
        LK(o + 5) = LK(o + 5) - BD(i, 5, 1) * LK(o + 1) - BD(i, 5, 2) * LK(o + 2) -
                    BD(i, 5, 3) * LK(o + 3) - BD(i, 5, 4) * LK(o + 4);
      }
    }
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < nshg; ++i)
    for (long long k = colm[i] - 1; k < colm[i + 1] - 1; ++k) {
      long long j = rowp[k] - 1;
      double* L = lhsK + 25 * k;
      for (int f = 1; f <= 5; ++f) LK(f) = BD(j, 1, 1) * LK(f);
      for (int f = 1; f <= 5; ++f) LK(5 + f) = BD(j, 2, 2) * (LK(5 + f) - BD(j, 1, 2) * LK(f));
      for (int f = 1; f <= 5; ++f)
        LK(10 + f) = BD(j, 3, 3) * (LK(10 + f) - BD(j, 1, 3) * LK(f) - BD(j, 2, 3) * LK(5 + f));
      for (int f = 1; f <= 5; ++f)
        LK(15 + f) = BD(j, 4, 4) * (LK(15 + f) - BD(j, 1, 4) * LK(f) - BD(j, 2, 4) * LK(5 + f) -
                                    BD(j, 3, 4) * LK(10 + f));
      for (int f = 1; f <= 5; ++f)
        LK(20 + f) = BD(j, 5, 5) * (LK(20 + f) - BD(j, 1, 5) * LK(f) - BD(j, 2, 5) * LK(5 + f) -
                                    BD(j, 3, 5) * LK(10 + f) - BD(j, 4, 5) * LK(15 + f));
    }
}
#undef LK
#undef BD

/* sparseap.f:31-102 (single rank, no axisymmetric BC) */
void orc_sparse_ap(long long nshg, const int* iper, const int* colm,
                   const int* rowp, const double* lhsK, double* p) {
  double* pp = malloc(sizeof(double) * 5 * nshg);
  double* q = malloc(sizeof(double) * 5 * nshg);
  for (int f = 0; f < 5; ++f)
    for (long long n = 0; n < nshg; ++n) pp[n + nshg * f] = p[(iper[n] - 1) + nshg * f];
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < nshg; ++i) {
    double t1 = 0., t2 = 0., t3 = 0., t4 = 0., t5 = 0.;
    for (long long k = colm[i] - 1; k < colm[i + 1] - 1; ++k) {
      long long j = rowp[k] - 1;
      const double* L = lhsK + 25 * k;
      double p1 = pp[j], p2 = pp[j + nshg], p3 = pp[j + 2 * nshg], p4 = pp[j + 3 * nshg],
             p5 = pp[j + 4 * nshg];
      t1 = t1 + L[0] * p1 + L[5] * p2 + L[10] * p3 + L[15] * p4 + L[20] * p5;
      t2 = t2 + L[1] * p1 + L[6] * p2 + L[11] * p3 + L[16] * p4 + L[21] * p5;
      t3 = t3 + L[2] * p1 + L[7] * p2 + L[12] * p3 + L[17] * p4 + L[22] * p5;
      t4 = t4 + L[3] * p1 + L[8] * p2 + L[13] * p3 + L[18] * p4 + L[23] * p5;
      t5 = t5 + L[4] * p1 + L[9] * p2 + L[14] * p3 + L[19] * p4 + L[24] * p5;
    }
    q[i] = 0.0 + t1;
    q[i + nshg] = 0.0 + t2;
    q[i + 2 * nshg] = 0.0 + t3;
    q[i + 3 * nshg] = 0.0 + t4;
    q[i + 4 * nshg] = 0.0 + t5;
  }
  memcpy(p, q, sizeof(double) * 5 * nshg);
  free(pp);
  free(q);
}

/* bc3per.f:25-34 */
void orc_bc3per(long long nshg, const int* iBC, const int* iper, double* res,
                int nQs) {
  for (long long j = 0; j < nshg; ++j)
    if (iBC[j] & 1024) {
      long long i = iper[j] - 1;
      for (int f = 0; f < nQs; ++f) {
        res[i + nshg * f] = res[i + nshg * f] + res[j + nshg * f];
        res[j + nshg * f] = 0.0;
      }
    }
}

/* mpitools.f:113 sum(u) over (nshg,n), column-major */
static double sumsq_prod(long long len, const double* a, const double* b) {
  double s = 0.0;
  for (long long i = 0; i < len; ++i) s = s + a[i] * b[i];
  return s;
}

/* tnanq.f:10-28 */
static void tnanq(long long len, double* u, long long* nnan, long long* nlarge) {
  for (long long i = 0; i < len; ++i) {
    if (fabs(u[i]) > 1.0e10) (*nlarge)++;
    if (u[i] != u[i]) {
      (*nnan)++;
      u[i] = 9.876543e21;
    }
  }
}

int orc_solgmrs(const orc_mesh* m, const orc_params* P, const double* y,
                const double* ac, double* res, double* rmes, double* BDiag,
                double* lhsK, double* Dy, orc_stats* st) {
  const long long nshg = m->nshg, V = 5 * nshg;
  const int K = P->Kspace;
  int rc = orc_elmgmrs(m, P, y, ac, res, NULL, BDiag, lhsK);
  if (rc) return rc;
  if (rmes) memcpy(rmes, res, sizeof(double) * V); /* solgmr.f:456 */
  st->nnan = st->nlarge = 0;
  tnanq(V, res, &st->nnan, &st->nlarge);
  tnanq(25 * nshg, BDiag, &st->nnan, &st->nlarge);
  if (st->nnan + st->nlarge > 0) return 10;
  if (P->iprec != 0) orc_i3lu(nshg, BDiag, NULL, 0);
  orc_i3lu(nshg, BDiag, res, 1);
  /* rstatCheck.f:36-79 */
  {
    double rtot = 0.0;
    for (long long n = 0; n < nshg; ++n) {
      double r2 = 0.0;
      for (int f = 0; f < 5; ++f) r2 = r2 + res[n + nshg * f] * res[n + nshg * f];
      rtot = rtot + r2;
    }
    double totres = sqrt(rtot / (double)(float)nshg);
    /* the divergence test skips the first solve of a time step (iter is
     * reset at itrdrv.f:409 and incremented before each solve, :434) */
    if (P->iter > 1 && totres > 10000.0 * st->reslast) return 11;
    st->reslast = totres;
  }
  memset(Dy, 0, sizeof(double) * V);
  if (P->lhs == 1) orc_spsi3pre(nshg, BDiag, lhsK, m->colm, m->rowp);
  double* u = malloc(sizeof(double) * V * (K + 1));
  double* H = calloc((size_t)(K + 1) * K, sizeof(double));
  double* e = calloc(K + 1, sizeof(double));
  double* yB = calloc(K + 1, sizeof(double));
  double* Rc = calloc(K + 1, sizeof(double));
  double* Rs = calloc(K + 1, sizeof(double));
#define HB(a, b) H[((a)-1) + (K + 1) * ((b)-1)]
#define UB(k) (u + V * ((k)-1))
  memcpy(UB(1), res, sizeof(double) * V);
  double unorm = sqrt(sumsq_prod(V, res, res));
  int iKs = 0;
  double epsnrm = P->etol * unorm, echeck = 0.0;
  st->unorm0 = unorm;
  st->epsnrm = epsnrm;
  if (!(unorm < 100. * P->epsM * P->epsM)) {
    for (int mG = 1; mG <= P->nGMRES; ++mG) {
      st->lGMRES = mG - 1;
      /* restart branch (solgmr.f:526-554) never runs: it tests lGMRES,
       * which this routine does not update; nGMRES = 1 is assumed. */
      memset(e, 0, sizeof(double) * (K + 1));
      e[0] = unorm;
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < V; ++i) UB(1)[i] = UB(1)[i] / unorm;
      for (int iK = 1; iK <= K; ++iK) {
        iKs = iK;
        double* w = UB(iK + 1);
        memcpy(w, UB(iK), sizeof(double) * V);
        orc_sparse_ap(nshg, m->iper, m->colm, m->rowp, lhsK, w);
        orc_bc3per(nshg, m->iBC, m->iper, w, 5);
        double beta = 0.0;
        for (int jK = 1; jK <= iK + 1; ++jK) {
          if (jK == 1) {
            beta = sumsq_prod(V, w, UB(1));
          } else {
            const double* um = UB(jK - 1);
#pragma omp parallel for schedule(static)
            for (long long i = 0; i < V; ++i) w[i] = w[i] - beta * um[i];
            beta = sumsq_prod(V, w, UB(jK));
          }
          HB(jK, iK) = beta;
        }
        unorm = sqrt(beta);
        HB(iK + 1, iK) = unorm;
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < V; ++i) w[i] = w[i] / unorm;
        for (int jK = 1; jK <= iK - 1; ++jK) {
          double tmp = Rc[jK] * HB(jK, iK) + Rs[jK] * HB(jK + 1, iK);
          HB(jK + 1, iK) = -Rs[jK] * HB(jK, iK) + Rc[jK] * HB(jK + 1, iK);
          HB(jK, iK) = tmp;
        }
        double tmp = sqrt(HB(iK, iK) * HB(iK, iK) + HB(iK + 1, iK) * HB(iK + 1, iK));
        Rc[iK] = HB(iK, iK) / tmp;
        Rs[iK] = HB(iK + 1, iK) / tmp;
        HB(iK, iK) = tmp;
        HB(iK + 1, iK) = 0.0;
        tmp = Rc[iK] * e[iK - 1] + Rs[iK] * e[iK];
        e[iK] = -Rs[iK] * e[iK - 1] + Rc[iK] * e[iK];
        e[iK - 1] = tmp;
        st->ntotGM++;
        echeck = fabs(e[iK]);
        if (echeck <= epsnrm && iK >= P->minIters) break;
      }
      for (int jK = iKs; jK >= 1; --jK) {
        yB[jK] = e[jK - 1] / HB(jK, jK);
        for (int lK = 1; lK <= jK - 1; ++lK) e[lK - 1] = e[lK - 1] - yB[jK] * HB(lK, jK);
      }
      for (int jK = 1; jK <= iKs; ++jK) {
        const double* uj = UB(jK);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < V; ++i) Dy[i] = Dy[i] + yB[jK] * uj[i];
      }
      echeck = fabs(e[iKs]);
      if (echeck <= epsnrm) break;
    }
  }
#undef HB
#undef UB
  orc_i3lu(nshg, BDiag, Dy, 2);
  st->iKs = iKs;
  st->echeck = echeck;
  st->converged = echeck <= epsnrm;
  /* rstat.f:95-116 */
  {
    double s1 = 0.0, s2 = 0.0;
    for (long long n = 0; n < nshg; ++n) {
      double a = 0.0, b = 0.0;
      for (int f = 0; f < 5; ++f) {
        a = a + res[n + nshg * f] * res[n + nshg * f];
        if (rmes) b = b + rmes[n + nshg * f] * rmes[n + nshg * f];
      }
      s1 = s1 + a;
      s2 = s2 + b;
    }
    st->totres[0] = sqrt(s1) / (double)(float)nshg;
    st->totres[1] = sqrt(s2) / (double)(float)nshg;
  }
  free(u);
  free(H);
  free(e);
  free(yB);
  free(Rc);
  free(Rs);
  return 0;
}

/* itrPC.f:56-73, the same-velocity predictor.  ipred 2-4 (itrPC.f:75-119)
 * pass undeclared, implicitly typed scalars for iBC, BC, iper and ilwork to
 * itrBC, so the reference defines no result for them: rejected (1). */
int orc_itr_predict(long long nshg, const orc_params* P, const double* yold,
                    const double* acold, double* y, double* ac) {
  if (P->ipred != 1) return 1;
  for (long long i = 0; i < 5 * nshg; ++i) {
    y[i] = yold[i];
    ac[i] = acold[i] * (ONE - P->almi / P->gami);
  }
  return 0;
}

/* itrPC.f:128-148 */
void orc_itr_correct(long long nshg, const orc_params* P, double* y,
                     double* ac, const double* yold, const double* acold,
                     const double* Dy) {
  for (long long n = 0; n < nshg; ++n) {
    y[n] = y[n] - Dy[n + nshg];
    y[n + nshg] = y[n + nshg] - Dy[n + 2 * nshg];
    y[n + 2 * nshg] = y[n + 2 * nshg] - Dy[n + 3 * nshg];
    y[n + 3 * nshg] = y[n + 3 * nshg] - Dy[n];
    y[n + 4 * nshg] = y[n + 4 * nshg] - Dy[n + 4 * nshg];
  }
  double fct1 = ONE - P->almi / P->gami;
  double fct2 = P->almi * P->Dtgl / P->gami / P->alfi;
  for (long long i = 0; i < 5 * nshg; ++i) ac[i] = acold[i] * fct1 + (y[i] - yold[i]) * fct2;
}

/* itrPC.f:176-208, branch without level-set redistancing (iLset != 2) */
void orc_itr_update(long long nshg, const orc_params* P, double* yold,
                    double* acold, const double* y, const double* ac) {
  double fct2 = ONE / P->almi;
  double fct3 = ONE / P->alfi;
  for (long long i = 0; i < 5 * nshg; ++i) acold[i] = acold[i] + (ac[i] - acold[i]) * fct2;
  for (long long i = 0; i < 5 * nshg; ++i) yold[i] = yold[i] + (y[i] - yold[i]) * fct3;
}

/* itrbc.f:60-178 (codes used by this path) */
int orc_itr_bc(const orc_mesh* m, double* y, double* ac) {
  const long long nshg = m->nshg;
  const double* BC = m->BC;
  for (long long n = 0; n < nshg; ++n) {
    int ibc = m->iBC[n];
    if (!bc_supported(ibc)) return 3;
    if (ibc & 2) y[n + 4 * nshg] = BC[n + nshg * 1];
    if (vel_code(ibc) == 7) {
      y[n] = BC[n + nshg * 2];
      y[n + nshg] = BC[n + nshg * 3];
      y[n + 2 * nshg] = BC[n + nshg * 4];
    }
    if (ibc & 4) y[n + 3 * nshg] = BC[n];
  }
  for (int f = 0; f < 5; ++f)
    for (long long n = 0; n < nshg; ++n) {
      long long i = m->iper[n] - 1;
      y[n + nshg * f] = y[i + nshg * f];
      if (ac) ac[n + nshg * f] = ac[i + nshg * f];
    }
  return 0;
}
