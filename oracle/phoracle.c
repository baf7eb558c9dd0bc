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
 *   phSolver/compressible/e3tau.f:1-283,1397-1507  e3tau (itau 0 and 1), e3gijd
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

#include "phoracle_e3.c"


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
  if (P->idc < 0 || P->idc > 3) return 5; /* input_fform.cc:683-686 */
  /* Shakib's or Franca's diagonal tau, or e3tau_nd; itau 2 and 3 have no
   * branch in the compressible e3tau (tau is left unset) */
  if (P->itau != 0 && P->itau != 1 && P->itau != 10 && P->itau != 11) return 5;
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

/* ---------------- matrix-free GMRES: SolMFG and its helpers ---------------- */

/* bc3res.f:25-163 on one residual-type array */
static void mfg_bc3res(const orc_mesh* m, double* res) {
  const long long nshg = m->nshg;
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
}

/* bc3bdg.f:25-352 */
static void mfg_bc3bdg(const orc_mesh* m, double* BDiag) {
  const long long nshg = m->nshg;
  for (long long n = 0; n < nshg; ++n) {
    int ibc = m->iBC[n];
    int cons[5] = {(ibc & 4) != 0, vel_code(ibc) == 7, vel_code(ibc) == 7, vel_code(ibc) == 7,
                   (ibc & 2) != 0};
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
      for (int jk = 0; jk < 25; ++jk) BDiag[i + nshg * jk] = BDiag[i + nshg * jk] + BDiag[j + nshg * jk];
    }
  for (long long j = 0; j < nshg; ++j)
    if (m->iBC[j] & 1024) {
      long long i = m->iper[j] - 1;
      for (int jk = 0; jk < 25; ++jk) BDiag[j + nshg * jk] = BDiag[i + nshg * jk];
    }
}

/* localy.f:53-73 gather of node g in the local order (p,u,v,w,T) */
static void gather_local5(const double* y, long long nshg, long long g, double v[5]) {
  v[1] = y[g];
  v[2] = y[g + nshg];
  v[3] = y[g + 2 * nshg];
  v[0] = y[g + 3 * nshg];
  v[4] = y[g + 4 * nshg];
}

/* AsIMFG (asimfg.f:43-110, ires = 3) or AsIRes (asires.f:29-94, ires = 2)
 * over the interior elements, ibksiz at a time, scattered in the
 * reference's order (local.f:67-74).  yp: e3's yl, y: e3's ycl (the same
 * array for AsIMFG).  res / BDiag only for ires = 3, rmes always; iabres:
 * |rml| per element before the scatter (asires.f:90, itrFDI's cancellation
 * estimate). */
static int mfg_interior(const orc_mesh* m, const orc_params* P, const double* yp,
                        const double* y, const double* ac, const double* qres, int ires,
                        int iabres, double* res, double* rmes, double* BDiag) {
  const long long nshg = m->nshg, nel = m->nel;
  const int bk = m->ibksiz > 0 ? m->ibksiz : 64;
  const int bdg = ires == 3 && P->iprec == 1 && P->lhs != 1; /* e3.f:261 */
  enum { kSuper = 64 };
  const long long cap = (long long)kSuper * bk;
  double(*rl)[4][5] = malloc(sizeof(double) * 20 * cap);
  double(*rml)[4][5] = malloc(sizeof(double) * 20 * cap);
  double(*BDl)[4][25] = bdg ? malloc(sizeof(double) * 100 * cap) : NULL;
  int(*nod)[4] = malloc(sizeof(int) * 4 * cap);
  if (!rl || !rml || (bdg && !BDl) || !nod) {
    free(rl);
    free(rml);
    free(BDl);
    free(nod);
    return 4;
  }
  for (long long s0 = 0; s0 < nel; s0 += cap) {
    const long long ns = (nel - s0) < cap ? (nel - s0) : cap;
#pragma omp parallel for schedule(dynamic, 1)
    for (long long e = 0; e < ns; ++e) {
      double ylp[4][5], yl[4][5], acl[4][5], xl[4][3], qll[4][12];
      for (int a = 0; a < 4; ++a) {
        long long g = m->ien[(s0 + e) + nel * a] - 1;
        nod[e][a] = (int)g;
        gather_local5(yp, nshg, g, ylp[a]);
        gather_local5(y, nshg, g, yl[a]);
        gather_local5(ac, nshg, g, acl[a]);
        for (int c = 0; c < 3; ++c) xl[a][c] = m->x[g + nshg * c];
        if (qres)
          for (int k = 0; k < 12; ++k) qll[a][k] = qres[g + nshg * k];
      }
      orc_e3_mfg(P, (const double(*)[5])ylp, (const double(*)[5])yl, (const double(*)[5])acl,
                 (const double(*)[3])xl, qres ? (const double(*)[12])qll : NULL, ires, rl[e],
                 rml[e], bdg ? BDl[e] : NULL);
      if (iabres)
        for (int a = 0; a < 4; ++a)
          for (int f = 0; f < 5; ++f) rml[e][a][f] = fabs(rml[e][a][f]);
    }
    for (long long e0 = 0; e0 < ns; e0 += bk) {
      const int npro = (int)((ns - e0) < bk ? (ns - e0) : bk);
      if (ires == 3)
        for (int f = 0; f < 5; ++f)
          for (int i = 0; i < 4; ++i)
            for (int e = (int)e0; e < e0 + npro; ++e)
              res[nod[e][i] + nshg * f] = res[nod[e][i] + nshg * f] + rl[e][i][f];
      for (int f = 0; f < 5; ++f)
        for (int i = 0; i < 4; ++i)
          for (int e = (int)e0; e < e0 + npro; ++e)
            rmes[nod[e][i] + nshg * f] = rmes[nod[e][i] + nshg * f] + rml[e][i][f];
      if (bdg)
        for (int jk = 0; jk < 25; ++jk)
          for (int i = 0; i < 4; ++i)
            for (int e = (int)e0; e < e0 + npro; ++e)
              BDiag[nod[e][i] + nshg * jk] = BDiag[nod[e][i] + nshg * jk] + BDl[e][i][jk];
    }
  }
  free(rl);
  free(rml);
  free(BDl);
  free(nod);
  return 0;
}

static int mfg_flags(const orc_mesh* m, const orc_params* P) {
  for (long long n = 0; n < m->nshg; ++n)
    if (!bc_supported(m->iBC[n])) return 3;
  if (P->idiff != 0 && P->idiff != 1 && P->idiff != 3) return 5;
  /* e3dc's ires = 2 branch reads giju, rTLS and raLS, which e3tau forms only
   * for ires 1 and 3 (e3tau.f:185-262) */
  if (P->idc != 0) return 5;
  if (P->itau != 0 && P->itau != 1 && P->itau != 10 && P->itau != 11) return 5;
  return 0;
}

int orc_elmmfg(const orc_mesh* m, const orc_params* P, const double* y, const double* ac,
               double* res, double* rmes, double* BDiag) {
  const long long nshg = m->nshg, nelb = m->nelb;
  int rc = mfg_flags(m, P);
  if (rc) return rc;
  double* qres = NULL;
  if (P->idiff != 0) { /* elmmfg.f:68-124 */
    qres = malloc(sizeof(double) * 13 * nshg);
    if (!qres) return 4;
    rc = orc_asiq(m, P, y, qres, qres + 12 * nshg);
    if (rc) {
      free(qres);
      return rc;
    }
  }
  /* elmmfg.f:128-132 */
  memset(res, 0, sizeof(double) * nshg * 5);
  memset(rmes, 0, sizeof(double) * nshg * 5);
  if (P->iprec != 0) memset(BDiag, 0, sizeof(double) * 25 * nshg);
  rc = mfg_interior(m, P, y, y, ac, qres, 3, 0, res, rmes, BDiag);
  free(qres);
  if (rc) return rc;
  /* AsBMFG (elmmfg.f:178-226): the wall flux enters res only, e3b adds to
   * rmes only with Jactyp = 1 (e3b.f:296-297) and itrSetup sets Jactyp = 0
   * (itrPC.f:29) */
  const int bk = m->ibksiz > 0 ? m->ibksiz : 64;
  double(*rl)[4][5] = malloc(sizeof(double) * 20 * bk);
  int(*nod)[4] = malloc(sizeof(int) * 4 * bk);
  if (!rl || !nod) {
    free(rl);
    free(nod);
    return 4;
  }
  for (long long e0 = 0; e0 < nelb; e0 += bk) {
    int npro = (int)((nelb - e0) < bk ? (nelb - e0) : bk);
    for (int e = 0; e < npro; ++e) {
      double yl[4][5], xl[4][3];
      for (int a = 0; a < 4; ++a) {
        long long g = m->ienb[(e0 + e) + nelb * a] - 1;
        nod[e][a] = (int)g;
        gather_local5(y, nshg, g, yl[a]);
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
  free(nod);
  /* elmmfg.f:247-256 */
  mfg_bc3res(m, res);
  mfg_bc3res(m, rmes);
  if (P->iprec != 0) mfg_bc3bdg(m, BDiag);
  return 0;
}

int orc_itrres(const orc_mesh* m, const orc_params* P, const double* yp, const double* yc,
               const double* ac, double* rmes, int iabres) {
  int rc = mfg_flags(m, P);
  if (rc) return rc;
  /* itrres.f:40-81: interior elements only (the boundary loop needs
   * Jactyp != 0, itrres.f:86), accumulated into rmes, then bc3Res */
  rc = mfg_interior(m, P, yp, yc, ac, NULL, 2, iabres, NULL, rmes, NULL);
  if (rc) return rc;
  mfg_bc3res(m, rmes);
  return 0;
}

/* shuffle.f:1-30: (p,u,v,w,T) <-> (u,v,w,p,T) of an (nshg,5) array */
void orc_yshuffle(long long nshg, double* r, int new2old) {
  for (long long n = 0; n < nshg; ++n) {
    double t[5];
    for (int f = 0; f < 5; ++f) t[f] = r[n + nshg * f];
    if (new2old) {
      r[n] = t[3];
      r[n + nshg] = t[0];
      r[n + 2 * nshg] = t[1];
      r[n + 3 * nshg] = t[2];
    } else {
      r[n] = t[1];
      r[n + nshg] = t[2];
      r[n + 2 * nshg] = t[3];
      r[n + 3 * nshg] = t[0];
    }
  }
}

/* y-variables of the direction: old2new(U^-1 (ypre + s d)) (au1mfg.f:64-68,
 * itrfdi.f:61-65) */
static void mfg_point(long long nshg, const double* BDiag, const double* ypre, double s,
                      const double* d, double* out) {
  const long long V = 5 * nshg;
  if (d)
    for (long long i = 0; i < V; ++i) out[i] = ypre[i] + s * d[i];
  else
    memcpy(out, ypre, sizeof(double) * V);
  orc_i3lu(nshg, (double*)BDiag, out, 2);
  orc_yshuffle(nshg, out, 0);
}

int orc_itrfdi(const orc_mesh* m, const orc_params* P, const double* ypre, const double* y,
               const double* ac, const double* rmes, const double* u, const double* BDiag,
               double* eGMRES) {
  const long long nshg = m->nshg, V = 5 * nshg;
  double* yt = malloc(sizeof(double) * V);
  double* rt = calloc(V, sizeof(double));
  if (!yt || !rt) {
    free(yt);
    free(rt);
    return 4;
  }
  int rc;
  /* itrfdi.f:56-82 cancellation error epsA */
  mfg_point(nshg, BDiag, ypre, 0.0, NULL, yt);
  if ((rc = orc_itrres(m, P, yt, y, ac, rt, 1))) goto out;
  orc_i3lu(nshg, (double*)BDiag, rt, 1);
  const double epsA = (P->epsM * P->epsM) * sqrt(sumsq_prod(V, rt, rt));
  /* itrfdi.f:86-130 second difference at epsSD = sqrt(epsM) along u */
  const double epsSD = sqrt(P->epsM);
  memset(rt, 0, sizeof(double) * V);
  mfg_point(nshg, BDiag, ypre, epsSD, u, yt);
  if ((rc = orc_itrres(m, P, yt, y, ac, rt, 0))) goto out;
  mfg_point(nshg, BDiag, ypre, -epsSD, u, yt);
  if ((rc = orc_itrres(m, P, yt, y, ac, rt, 0))) goto out;
  orc_i3lu(nshg, (double*)BDiag, rt, 1);
  for (long long i = 0; i < V; ++i) {
    double d = (rt[i] - TWO * rmes[i]) / P->epsM;
    rt[i] = d * d;
  }
  {
    double SDnrm = 0.0;
    for (long long i = 0; i < V; ++i) SDnrm = SDnrm + rt[i];
    SDnrm = sqrt(SDnrm);
    *eGMRES = TWO * sqrt(epsA / SDnrm); /* itrfdi.f:139 */
  }
out:
  free(yt);
  free(rt);
  return rc;
}

int orc_au1mfg(const orc_mesh* m, const orc_params* P, const double* ypre, const double* y,
               const double* ac, const double* rmes, const double* BDiag, double eGMRES,
               double* u) {
  const long long nshg = m->nshg, V = 5 * nshg;
  double* ub = calloc(V, sizeof(double));
  if (!ub) return 4;
  /* au1mfg.f:62-83 */
  for (long long i = 0; i < V; ++i) u[i] = ypre[i] + eGMRES * u[i];
  orc_i3lu(nshg, (double*)BDiag, u, 2);
  orc_yshuffle(nshg, u, 0);
  int rc = orc_itr_bc(m, u, NULL); /* ires = 2: y only (itrbc.f:137-139) */
  if (!rc) rc = orc_itrres(m, P, u, y, ac, ub, 0);
  if (!rc) {
    orc_i3lu(nshg, (double*)BDiag, ub, 1);
    for (long long i = 0; i < V; ++i) u[i] = (ub[i] - rmes[i]) / eGMRES;
  }
  free(ub);
  return rc;
}

int orc_au2mfg(const orc_mesh* m, const orc_params* P, const double* ypre, const double* y,
               const double* ac, const double* res, const double* Dy, const double* BDiag,
               double* u) {
  const long long nshg = m->nshg, V = 5 * nshg;
  double* u1 = calloc(V, sizeof(double));
  double* u2 = calloc(V, sizeof(double));
  double* w = malloc(sizeof(double) * V);
  int rc = 4;
  if (!u1 || !u2 || !w) goto out;
  /* au2mfg.f:59-108 */
  const double eps = pow(P->epsM, PT66) / sqrt(sumsq_prod(V, Dy, Dy));
  for (long long i = 0; i < V; ++i) w[i] = ypre[i] + eps * Dy[i];
  orc_i3lu(nshg, (double*)BDiag, w, 2);
  orc_yshuffle(nshg, w, 0);
  if ((rc = orc_itr_bc(m, w, NULL))) goto out;
  if ((rc = orc_itrres(m, P, w, y, ac, u1, 0))) goto out;
  orc_i3lu(nshg, (double*)BDiag, u1, 1);
  for (long long i = 0; i < V; ++i) w[i] = ypre[i] - eps * Dy[i];
  orc_i3lu(nshg, (double*)BDiag, w, 2);
  orc_yshuffle(nshg, w, 0);
  if ((rc = orc_itr_bc(m, w, NULL))) goto out;
  if ((rc = orc_itrres(m, P, w, y, ac, u2, 0))) goto out;
  orc_i3lu(nshg, (double*)BDiag, u2, 1);
  for (long long i = 0; i < V; ++i) u[i] = res[i] - (u1[i] - u2[i]) / (TWO * eps);
out:
  free(u1);
  free(u2);
  free(w);
  return rc;
}

int orc_solmfg(const orc_mesh* m, const orc_params* P, const double* y, const double* ac,
               double* res, double* rmes, double* BDiag, double* Dy, double* eGMRES, int fdi,
               orc_stats* st) {
  const long long nshg = m->nshg, V = 5 * nshg;
  const int K = P->Kspace;
  int rc = orc_elmmfg(m, P, y, ac, res, rmes, BDiag);
  if (rc) return rc;
  /* solmfg.f:101-118 */
  if (P->iprec != 0) orc_i3lu(nshg, BDiag, NULL, 0);
  orc_i3lu(nshg, BDiag, res, 1);
  orc_i3lu(nshg, BDiag, rmes, 1);
  double* u = malloc(sizeof(double) * V * (K + 1));
  double* ypre = malloc(sizeof(double) * V);
  double* H = calloc((size_t)(K + 1) * K, sizeof(double));
  double* e = calloc(K + 1, sizeof(double));
  double* yB = calloc(K + 1, sizeof(double));
  double* Rc = calloc(K + 1, sizeof(double));
  double* Rs = calloc(K + 1, sizeof(double));
  if (!u || !ypre || !H || !e || !yB || !Rc || !Rs) {
    rc = 4;
    goto out;
  }
#define HB(a, b) H[((a)-1) + (K + 1) * ((b)-1)]
#define UB(k) (u + V * ((k)-1))
  memcpy(UB(1), res, sizeof(double) * V);
  memset(Dy, 0, sizeof(double) * V);
  /* solmfg.f:124-133 ypre = U new2old(y) */
  memcpy(ypre, y, sizeof(double) * V);
  orc_yshuffle(nshg, ypre, 1);
  orc_i3lu(nshg, BDiag, ypre, 3);
  double unorm = sqrt(sumsq_prod(V, res, res));
  int iKs = 0;
  double epsnrm = 0.0, echeck = 0.0;
  st->lGMRES = 0;
  st->unorm0 = unorm;
  if (!(unorm < 100. * P->epsM * P->epsM)) {
    epsnrm = P->etol * unorm;
    /* solmfg.f:163-170 (the caller's iter = 1 and mod(istep, 20) = 0) */
    if (fdi && (rc = orc_itrfdi(m, P, ypre, y, ac, rmes, res, BDiag, eGMRES))) goto out;
    for (int mG = 1; mG <= P->nGMRES; ++mG) {
      st->lGMRES = mG - 1;
      if (mG > 1) { /* solmfg.f:181-202 R - A Dy by the two-sided difference */
        if ((rc = orc_au2mfg(m, P, ypre, y, ac, res, Dy, BDiag, UB(1)))) goto out;
        unorm = sqrt(sumsq_prod(V, UB(1), UB(1)));
      }
      memset(e, 0, sizeof(double) * (K + 1));
      e[0] = unorm;
      for (long long i = 0; i < V; ++i) UB(1)[i] = UB(1)[i] / unorm;
      for (int iK = 1; iK <= K; ++iK) {
        iKs = iK;
        double* w = UB(iK + 1);
        memcpy(w, UB(iK), sizeof(double) * V);
        if ((rc = orc_au1mfg(m, P, ypre, y, ac, rmes, BDiag, *eGMRES, w))) goto out;
        double beta = 0.0;
        for (int jK = 1; jK <= iK + 1; ++jK) {
          if (jK == 1) {
            beta = sumsq_prod(V, w, UB(1));
          } else {
            const double* um = UB(jK - 1);
            for (long long i = 0; i < V; ++i) w[i] = w[i] - beta * um[i];
            beta = sumsq_prod(V, w, UB(jK));
          }
          HB(jK, iK) = beta;
        }
        unorm = sqrt(beta);
        HB(iK + 1, iK) = unorm;
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
        for (long long i = 0; i < V; ++i) Dy[i] = Dy[i] + yB[jK] * uj[i];
      }
      echeck = fabs(e[iKs]);
      if (echeck <= epsnrm) break;
    }
  }
#undef HB
#undef UB
  /* solmfg.f:364 */
  orc_i3lu(nshg, BDiag, Dy, 2);
  st->iKs = iKs;
  st->epsnrm = epsnrm;
  st->echeck = echeck;
  st->converged = echeck <= epsnrm;
  {
    /* rstat.f:95-116 on res; solmfg.f:368 passes no second array, so the
     * second norm here is the preconditioned modified residual's */
    double s1 = 0.0, s2 = 0.0;
    for (long long n = 0; n < nshg; ++n) {
      double a = 0.0, b = 0.0;
      for (int f = 0; f < 5; ++f) {
        a = a + res[n + nshg * f] * res[n + nshg * f];
        b = b + rmes[n + nshg * f] * rmes[n + nshg * f];
      }
      s1 = s1 + a;
      s2 = s2 + b;
    }
    st->totres[0] = sqrt(s1) / (double)(float)nshg;
    st->totres[1] = sqrt(s2) / (double)(float)nshg;
  }
out:
  free(u);
  free(ypre);
  free(H);
  free(e);
  free(yB);
  free(Rc);
  free(Rs);
  return rc;
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

int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}
