/*
 * phoracle_e3.c -- element routines of the CPU oracle (test infrastructure,
 * not the product): quadrature, e3metric, e3gijd and the interior element
 * residual and tangent e3.  Included by phoracle.c; kept in a file of its own
 * so a flop-counting build can compile it with an instrumented number type.
 */

#define PT125 0.1250000000000000000000000000000
#define PT5 0.5000000000000000000000000000000
#define PT66 0.6666666666666666666666666666667
#define ONE 1.0
#define TWO 2.0


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

/* sign(one, x) of the reference (IEEE sign bit of x) */
static inline double fsign1(double x) { return copysign(ONE, x); }

/* e3eig1.f:1-221 (lcsyst = 1): eigenvectors Q of the Euler Jacobian along
 * the streamline, in entropy variables; rotates dxidx into the streamline
 * frame in place (as the reference does, so later uses of dxidx in the same
 * point see the rotated rows) and returns in a1..a3 the squared column norms
 * of the rotated dxidx with the tet correction */
static void e3eig1(const orc_params* P, double rho, double T, double cp, double gamb,
                   double c, double u1, double u2, double u3, double* a1, double* a2,
                   double* a3, double eb1, double dxidx[3][3], double u, double Q[5][5]) {
  const double pt39 = 0.3968502629920498686879264098181;
  const double sqt2 = 1.4142135623730950488016887242097;
  double R[3][3], fact;
  if (u != 0.0) {
    fact = ONE / u;
    R[0][0] = u1 * fact;
    R[0][1] = u2 * fact;
    R[0][2] = u3 * fact;
  } else {
    R[0][0] = ONE;
    R[0][1] = 0.0;
    R[0][2] = 0.0;
  }
  fact = *a1 * R[0][0] + *a2 * R[0][1] + *a3 * R[0][2];
  *a1 = *a1 - fact * R[0][0];
  *a2 = *a2 - fact * R[0][1];
  *a3 = *a3 - fact * R[0][2];
  fact = *a1 * *a1 + *a2 * *a2 + *a3 * *a3;
  if (fact > P->epsM) {
    R[1][0] = *a1;
    R[1][1] = *a2;
    R[1][2] = *a3;
  } else {
    const double sg = fsign1(R[0][1] * R[0][2]);
    R[1][0] = R[0][1] + sg * R[0][2];
    R[1][1] = -R[0][0] - sg * R[0][2];
    R[1][2] = sg * (R[0][1] - R[0][0]);
    fact = R[1][0] * R[1][0] + R[1][1] * R[1][1] + R[1][2] * R[1][2];
  }
  fact = ONE / sqrt(fact);
  R[1][0] = R[1][0] * fact;
  R[1][1] = R[1][1] * fact;
  R[1][2] = R[1][2] * fact;
  R[2][0] = R[0][1] * R[1][2] - R[0][2] * R[1][1];
  R[2][1] = R[0][2] * R[1][0] - R[0][0] * R[1][2];
  R[2][2] = R[0][0] * R[1][1] - R[0][1] * R[1][0];
  for (int k = 0; k < 3; ++k) {
    const double b1 = R[0][0] * dxidx[k][0] + R[0][1] * dxidx[k][1] + R[0][2] * dxidx[k][2];
    const double b2 = R[1][0] * dxidx[k][0] + R[1][1] * dxidx[k][1] + R[1][2] * dxidx[k][2];
    const double b3 = R[2][0] * dxidx[k][0] + R[2][1] * dxidx[k][1] + R[2][2] * dxidx[k][2];
    dxidx[k][0] = b1;
    dxidx[k][1] = b2;
    dxidx[k][2] = b3;
  }
  double c1 = dxidx[0][0] * dxidx[0][0] + dxidx[1][0] * dxidx[1][0] + dxidx[2][0] * dxidx[2][0];
  double c2 = dxidx[0][1] * dxidx[0][1] + dxidx[1][1] * dxidx[1][1] + dxidx[2][1] * dxidx[2][1];
  double c3 = dxidx[0][2] * dxidx[0][2] + dxidx[1][2] * dxidx[1][2] + dxidx[2][2] * dxidx[2][2];
  double s1 = dxidx[0][0] + dxidx[1][0] + dxidx[2][0];
  double s2 = dxidx[0][1] + dxidx[1][1] + dxidx[2][1];
  double s3 = dxidx[0][2] + dxidx[1][2] + dxidx[2][2];
  *a1 = (c1 + s1 * s1) * pt39;
  *a2 = (c2 + s2 * s2) * pt39;
  *a3 = (c3 + s3 * s3) * pt39;
  fact = (ONE / sqrt(TWO * rho * T)) / c;
  Q[0][4] = fact * ((c + u) * c - eb1 * gamb);
  Q[1][4] = -fact * (c + u * gamb) * R[0][0];
  Q[2][4] = -fact * (c + u * gamb) * R[0][1];
  Q[3][4] = -fact * (c + u * gamb) * R[0][2];
  Q[4][4] = fact * gamb;
  Q[0][0] = fact * ((c - u) * c - eb1 * gamb);
  Q[1][0] = fact * (c - u * gamb) * R[0][0];
  Q[2][0] = fact * (c - u * gamb) * R[0][1];
  Q[3][0] = fact * (c - u * gamb) * R[0][2];
  Q[4][0] = fact * gamb;
  Q[0][2] = 0.0;
  Q[1][2] = -fact * c * sqt2 * R[1][0];
  Q[2][2] = -fact * c * sqt2 * R[1][1];
  Q[3][2] = -fact * c * sqt2 * R[1][2];
  Q[4][2] = 0.0;
  Q[0][1] = 0.0;
  Q[1][1] = fact * c * sqt2 * R[2][0];
  Q[2][1] = fact * c * sqt2 * R[2][1];
  Q[3][1] = fact * c * sqt2 * R[2][2];
  Q[4][1] = 0.0;
  fact = (ONE / sqrt(cp * rho)) / T;
  Q[0][3] = fact * eb1;
  Q[1][3] = fact * u * R[0][0];
  Q[2][3] = fact * u * R[0][1];
  Q[3][3] = fact * u * R[0][2];
  Q[4][3] = -fact;
}

/* e3eig1.f:224-342 (e3eig2): eigenvalues of the rotated metric-weighted
 * Jacobian, the (1,5) pair diagonalised by one Jacobi rotation of Q; with
 * the time term (iremoveStabTimeTerm = 0) */
static void e3eig2(const orc_params* P, double u, double c, double AR1, double AR2, double AR3,
                   double rlam[5], double Q[5][5], double* eigmax) {
  const double offd = (AR2 + AR3) * PT5 * (c * c);
  rlam[0] = AR1 * ((u + c) * (u + c)) + offd;
  rlam[1] = AR1 * (u * u) + AR3 * (c * c);
  rlam[2] = AR1 * (u * u) + AR2 * (c * c);
  rlam[3] = AR1 * (u * u);
  rlam[4] = AR1 * ((u - c) * (u - c)) + offd;
  const double tmp = P->dtsfct * 4.0 * (P->Dtgl * P->Dtgl);
  for (int i = 0; i < 5; ++i) rlam[i] = rlam[i] + tmp;
  double t = PT5 * (rlam[0] - rlam[4]) / offd;
  t = fsign1(t) / (fabs(t) + sqrt(ONE + t * t));
  const double Rc = ONE / sqrt(ONE + t * t);
  const double Rs = t * Rc;
  rlam[0] = rlam[0] - t * offd;
  rlam[4] = rlam[4] + t * offd;
  for (int r = 0; r < 5; ++r) {
    const double tt = Rc * Q[r][0] - Rs * Q[r][4];
    Q[r][4] = Rs * Q[r][0] + Rc * Q[r][4];
    Q[r][0] = tt;
  }
  double m = rlam[0];
  for (int i = 1; i < 5; ++i) m = fmax(m, rlam[i]);
  *eigmax = m;
}

/* e3visc.f:143-268 (itau >= 10, lcsyst = 1): the compact diffusivity in
 * local parent coordinates */
static void e3compk(const double dxidx[3][3], double T, double rmu, double rlm,
                    double rlm2mu, double con, double u1, double u2, double u3,
                    double compK[10]) {
  const double pt39 = 0.3968502629920498686879264098181;
  const double (*d)[3] = dxidx;
  double f1 = d[0][0] * d[0][0] + d[1][0] * d[1][0] + d[2][0] * d[2][0];
  double f2 = d[0][0] * d[0][1] + d[1][0] * d[1][1] + d[2][0] * d[2][1];
  double f3 = d[0][1] * d[0][1] + d[1][1] * d[1][1] + d[2][1] * d[2][1];
  double f4 = d[0][0] * d[0][2] + d[1][0] * d[1][2] + d[2][0] * d[2][2];
  double f5 = d[0][1] * d[0][2] + d[1][1] * d[1][2] + d[2][1] * d[2][2];
  double f6 = d[0][2] * d[0][2] + d[1][2] * d[1][2] + d[2][2] * d[2][2];
  const double s1 = d[0][0] + d[1][0] + d[2][0];
  const double s2 = d[0][1] + d[1][1] + d[2][1];
  const double s3 = d[0][2] + d[1][2] + d[2][2];
  f1 = (f1 + s1 * s1) * pt39;
  f2 = (f2 + s1 * s2) * pt39;
  f3 = (f3 + s2 * s2) * pt39;
  f4 = (f4 + s1 * s3) * pt39;
  f5 = (f5 + s2 * s3) * pt39;
  f6 = (f6 + s3 * s3) * pt39;
  compK[0] = f1 * T * rlm2mu + f3 * T * rmu + f6 * T * rmu;
  compK[1] = f2 * T * (rlm + rmu);
  compK[2] = f1 * T * rmu + f3 * T * rlm2mu + f6 * T * rmu;
  compK[3] = f4 * T * (rlm + rmu);
  compK[4] = f5 * T * (rlm + rmu);
  compK[5] = f1 * T * rmu + f3 * T * rmu + f6 * T * rlm2mu;
  compK[6] = f1 * T * rlm2mu * u1 + f2 * T * (rlm + rmu) * u2 + f3 * T * rmu * u1 +
             f4 * T * (rlm + rmu) * u3 + f6 * T * rmu * u1;
  compK[7] = f1 * T * rmu * u2 + f2 * T * (rlm + rmu) * u1 + f3 * T * rlm2mu * u2 +
             f5 * T * (rlm + rmu) * u3 + f6 * T * rmu * u2;
  compK[8] = f1 * T * rmu * u3 + f3 * T * rmu * u3 + f4 * T * (rlm + rmu) * u1 +
             f5 * T * (rlm + rmu) * u2 + f6 * T * rlm2mu * u3;
  const double rk = PT5 * (u1 * u1 + u2 * u2 + u3 * u3);
  compK[9] = f1 * T * (con * T + TWO * rmu * rk + (rlm + rmu) * (u1 * u1)) +
             f2 * T * (rlm + rmu) * TWO * u1 * u2 +
             f3 * T * (con * T + TWO * rmu * rk + (rlm + rmu) * (u2 * u2)) +
             f4 * T * (rlm + rmu) * TWO * u1 * u3 + f5 * T * (rlm + rmu) * TWO * u2 * u3 +
             f6 * T * (con * T + TWO * rmu * rk + (rlm + rmu) * (u3 * u3));
}

/* e3tau.f:284-518 (e3tau_nd, ipord = 1, Navier = 1): the matrix tau in
 * primitive variables, PTau = dY/dV Q diag(lambda) Q^T; dxidx is rotated in
 * place by e3eig1 */
static void e3tau_nd(const orc_params* P, double rho, double cp, double T, double u1,
                     double u2, double u3, double a1, double a2, double a3,
                     double dxidx[3][3], const double compK[10], double PTau[5][5]) {
  const double pt33 = 0.3333333333333333333333333333333;
  const double gamb = P->gamma1;
  double c = sqrt((P->gamma * P->Rgas) * T);
  const double u = sqrt(u1 * u1 + u2 * u2 + u3 * u3);
  const double eb1 = cp * T - PT5 * (u1 * u1 + u2 * u2 + u3 * u3);
  double Q[5][5], rlam[5], eigmax;
  e3eig1(P, rho, T, cp, gamb, c, u1, u2, u3, &a1, &a2, &a3, eb1, dxidx, u, Q);
  e3eig2(P, u, c, a1, a2, a3, rlam, Q, &eigmax);
  for (int i = 0; i < 5; ++i) {
    if (rlam[i] > (P->epsM * P->epsM) * eigmax)
      rlam[i] = ONE / sqrt(rlam[i]);
    else
      rlam[i] = 0.0;
  }
  for (int i = 0; i < 5; ++i) {
    const double q2 = Q[1][i], q3 = Q[2][i], q4 = Q[3][i], q5 = Q[4][i];
    c = pt33 * (q2 * (compK[0] * q2 + compK[1] * q3 + compK[3] * q4 + compK[6] * q5) +
                q3 * (compK[1] * q2 + compK[2] * q3 + compK[4] * q4 + compK[7] * q5) +
                q4 * (compK[3] * q2 + compK[4] * q3 + compK[5] * q4 + compK[8] * q5) +
                q5 * (compK[6] * q2 + compK[7] * q3 + compK[8] * q4 + compK[9] * q5));
    const double Pe = ONE / (c * rlam[i]);
    rlam[i] = rlam[i] * fmin(pt33 * Pe, ONE);
  }
  /* STau, symmetric, entropy variables: 11 12 22 13 23 33 14 24 34 44 15 25 35 45 55 */
  static const int sr[15] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3, 0, 1, 2, 3, 4};
  static const int sc[15] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 4};
  double S[15];
  for (int k = 0; k < 15; ++k) {
    const int r = sr[k], cc = sc[k];
    S[k] = rlam[0] * Q[r][0] * Q[cc][0] + rlam[1] * Q[r][1] * Q[cc][1] +
           rlam[2] * Q[r][2] * Q[cc][2] + rlam[3] * Q[r][3] * Q[cc][3] +
           rlam[4] * Q[r][4] * Q[cc][4];
  }
  /* PTau = dY/dV STau (G. Hauke's dY/dV) */
  double b = cp * T + PT5 * (u1 * u1 + u2 * u2 + u3 * u3);
  PTau[0][0] = rho * T * (S[0] + u1 * S[1] + u2 * S[3] + u3 * S[6] + b * S[10]);
  PTau[0][1] = rho * T * (S[1] + u1 * S[2] + u2 * S[4] + u3 * S[7] + b * S[11]);
  PTau[0][2] = rho * T * (S[3] + u1 * S[4] + u2 * S[5] + u3 * S[8] + b * S[12]);
  PTau[0][3] = rho * T * (S[6] + u1 * S[7] + u2 * S[8] + u3 * S[9] + b * S[13]);
  PTau[0][4] = rho * T * (S[10] + u1 * S[11] + u2 * S[12] + u3 * S[13] + b * S[14]);
  const double uu[3] = {u1, u2, u3};
  static const int col[3][5] = {{1, 2, 4, 7, 11}, {3, 4, 5, 8, 12}, {6, 7, 8, 9, 13}};
  static const int last[5] = {10, 11, 12, 13, 14};
  for (int m = 0; m < 3; ++m)
    for (int j = 0; j < 5; ++j) PTau[1 + m][j] = T * (S[col[m][j]] + uu[m] * S[last[j]]);
  b = T * T;
  for (int j = 0; j < 5; ++j) PTau[4][j] = b * S[last[j]];
}

/* e3bdg.f:1-2596 (itau < 10, ngauss = 4, Weak Form SUPG: ivart = 2, the
 * input default input.config:226), the block diagonal of the matrix-free
 * solver at one point, added to BDl[a][i + 5 k] = BDiagl(a, i, k):
 *   N_a A_i N_a,i                                     (e3bdg.f:34-189)
 *   N_a^2 WdetJ Dtgl almi/(gami alfi) A0               (e3bdg.f:216-246)
 *   N_a WdetJ almi/(gami alfi) Dtgl N_a,i A_i tau A0   (e3bdg.f:251-342)
 *   N_a,k N_a,l WdetJ (A_k tau A_l + K_kl)             (e3bdg.f:344-2595)
 * The reference writes each product out entry by entry and leaves out the
 * terms in which an A-matrix entry is zero by construction (e3mtrx.f); the
 * full sums here add those terms as exact zeros, in the same grouping and
 * order (tau class by tau class, A_k-first terms before A_l-first terms),
 * so the result is the same to the bit.  The stiffness blocks come in the
 * reference's order 11, 22, 33, 12+21, 13+31, 23+32, each with its
 * viscous K (e3bdg.f:602-610, 862-870, 1123-1131, 1609-1614, 2092-2097,
 * 2575-2580). */
/* e3bdg.f:34-246 = e3bdg_nd.f:34-247: the convective and mass terms */
static void e3bdg_head(const orc_params* P, const double shape[4], const double shg[4][3],
                       double WdetJ, const double A1[5][5], const double A2[5][5],
                       const double A3[5][5], const double A0[5][5], double BDl[4][25]) {
  for (int j = 0; j < 4; ++j) {
    const double tmp = shape[j] * WdetJ;
    for (int k = 0; k < 5; ++k)
      for (int i = 0; i < 5; ++i)
        BDl[j][i + 5 * k] =
            BDl[j][i + 5 * k] +
            tmp * (shg[j][0] * A1[i][k] + shg[j][1] * A2[i][k] + shg[j][2] * A3[i][k]);
  }
  const double ff = P->almi / P->gami / P->alfi;
  const double tmp = WdetJ * (P->Dtgl * ff + 0.0);
  for (int j = 0; j < 4; ++j) {
    const double tmp2 = (shape[j] * shape[j]) * tmp;
    for (int k = 0; k < 5; ++k)
      for (int i = 0; i < 5; ++i) BDl[j][i + 5 * k] = BDl[j][i + 5 * k] + tmp2 * A0[i][k];
  }
}

/* the viscous K of stiffness block (k,l) (e3bdg.f:602-610, 1609-1614 and
 * the like; e3bdg_nd.f:565-573, 1341-1346 and the like) */
static void e3bdg_visc(int k, int l, double u1, double u2, double u3, double rmu,
                       double rlm2mu, double con, double M[5][5]) {
  const double uu[3] = {u1, u2, u3};
  if (k == l) {
    for (int m = 0; m < 3; ++m) M[1 + m][1 + m] = M[1 + m][1 + m] + (m == k ? rlm2mu : rmu);
    for (int m = 0; m < 3; ++m) M[4][1 + m] = M[4][1 + m] + (m == k ? rlm2mu : rmu) * uu[m];
    M[4][4] = M[4][4] + con;
  } else {
    M[1 + k][1 + l] = M[1 + k][1 + l] + rlm2mu - rmu;
    M[1 + l][1 + k] = M[1 + l][1 + k] + rlm2mu - rmu;
    M[4][1 + k] = M[4][1 + k] + (rlm2mu - rmu) * uu[l];
    M[4][1 + l] = M[4][1 + l] + (rlm2mu - rmu) * uu[k];
  }
}

static void e3bdg(const orc_params* P, const double shape[4], const double shg[4][3],
                  double WdetJ, const double A1[5][5], const double A2[5][5],
                  const double A3[5][5], const double A0[5][5], const double tau[3],
                  double u1, double u2, double u3, double rmu, double rlm2mu, double con,
                  double BDl[4][25]) {
  const double(*Am[3])[5] = {A1, A2, A3};
  e3bdg_head(P, shape, shg, WdetJ, A1, A2, A3, A0, BDl);
  /* A_k tau A0 with the diagonal tau (tau1 continuity, tau2 momentum, tau3
   * energy columns) */
  const double tc[5] = {tau[0], tau[1], tau[1], tau[1], tau[2]};
  double AtA0[3][5][5];
  for (int d = 0; d < 3; ++d) {
    double At[5][5];
    for (int i = 0; i < 5; ++i)
      for (int m = 0; m < 5; ++m) At[i][m] = Am[d][i][m] * tc[m];
    for (int c = 0; c < 5; ++c)
      for (int i = 0; i < 5; ++i)
        AtA0[d][i][c] = At[i][0] * A0[0][c] + At[i][1] * A0[1][c] + At[i][2] * A0[2][c] +
                        At[i][3] * A0[3][c] + At[i][4] * A0[4][c];
  }
  for (int a = 0; a < 4; ++a) {
    const double tmp1 = shape[a] * WdetJ * P->almi / P->gami / P->alfi * P->Dtgl;
    for (int i = 0; i < 5; ++i)
      for (int c = 0; c < 5; ++c)
        BDl[a][i + 5 * c] = BDl[a][i + 5 * c] +
                            tmp1 * (shg[a][0] * AtA0[0][i][c] + shg[a][1] * AtA0[1][i][c] +
                                    shg[a][2] * AtA0[2][i][c]);
  }
  /* stiffness blocks (k,l): diagonal ones tau1 A(i,1)A(1,j) + tau2 (...) +
   * tau3 A(i,5)A(5,j); off-diagonal ones the sum of both orders per class */
  static const int kl[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  for (int b = 0; b < 6; ++b) {
    const int k = kl[b][0], l = kl[b][1];
    const double(*Ak)[5] = Am[k];
    const double(*Al)[5] = Am[l];
    double M[5][5];
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) {
        if (k == l) {
          M[i][j] = tau[0] * Ak[i][0] * Ak[0][j] +
                    tau[1] * (Ak[i][1] * Ak[1][j] + Ak[i][2] * Ak[2][j] + Ak[i][3] * Ak[3][j]) +
                    tau[2] * Ak[i][4] * Ak[4][j];
        } else {
          M[i][j] = tau[0] * (Ak[i][0] * Al[0][j] + Al[i][0] * Ak[0][j]) +
                    tau[1] * (Ak[i][1] * Al[1][j] + Ak[i][2] * Al[2][j] + Ak[i][3] * Al[3][j] +
                              Al[i][1] * Ak[1][j] + Al[i][2] * Ak[2][j] + Al[i][3] * Ak[3][j]) +
                    tau[2] * (Ak[i][4] * Al[4][j] + Al[i][4] * Ak[4][j]);
        }
      }
    e3bdg_visc(k, l, u1, u2, u3, rmu, rlm2mu, con, M);
    for (int a = 0; a < 4; ++a) {
      const double tmp = WdetJ * shg[a][k] * shg[a][l];
      for (int c = 0; c < 5; ++c)
        for (int i = 0; i < 5; ++i) BDl[a][i + 5 * c] = BDl[a][i + 5 * c] + tmp * M[i][c];
    }
  }
}

/* e3bdg_nd.f:1-2048 (itau >= 10, ivart = 2): e3bdg with the matrix tau.
 * Atau_k = A_k PTau (e3bdg_nd.f:252-314, accumulated from zero over the
 * inner index), A_k tau A0 and the least-squares time term as in e3bdg, and
 * the stiffness blocks (e3bdg_nd.f:354-2044) in the order 11, 22, 33, 12,
 * 13, 23: entry (i,j) of block (k,l) is sum_m Atau_k(i,m) A_l(m,j), followed
 * for k != l by sum_m Atau_l(i,m) A_k(m,j), left to right; the reference
 * leaves out the terms whose A entry is zero by construction, which the
 * full sums add as exact zeros.  Eight of the reference's statements read
 * another row of Atau_k than the entry's, and are restated as written
 * (1-based rows/columns):
 *   block 11, entry (5,3): Atau1(3,5) for Atau1(5,5)        (e3bdg_nd.f:539)
 *   block 22, entries (3,3), (4,3), (5,3): Atau2(2,m) for m = 1, 2, 4
 *     (e3bdg_nd.f:678, 716, 754), and in (5,3) Atau2(3,5) for Atau2(5,5)
 *   block 33, entries (3,4), (4,4), (5,4): Atau3(2,m) for m = 1, 2, 3
 *     (e3bdg_nd.f:902, 940, 978); entry (5,3): Atau3(3,5) for Atau3(5,5)
 *     (e3bdg_nd.f:971)
 * so this block diagonal is the tangent's only outside those entries. */
static void e3bdg_nd(const orc_params* P, const double shape[4], const double shg[4][3],
                     double WdetJ, const double A1[5][5], const double A2[5][5],
                     const double A3[5][5], const double A0[5][5], const double PTau[5][5],
                     double u1, double u2, double u3, double rmu, double rlm2mu, double con,
                     double BDl[4][25]) {
  const double(*Am[3])[5] = {A1, A2, A3};
  e3bdg_head(P, shape, shg, WdetJ, A1, A2, A3, A0, BDl);
  double Atau[3][5][5], AtA0[3][5][5];
  for (int d = 0; d < 3; ++d) {
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) {
        double acc = 0.0;
        for (int m = 0; m < 5; ++m) acc = acc + Am[d][i][m] * PTau[m][j];
        Atau[d][i][j] = acc;
      }
    for (int c = 0; c < 5; ++c)
      for (int i = 0; i < 5; ++i)
        AtA0[d][i][c] = Atau[d][i][0] * A0[0][c] + Atau[d][i][1] * A0[1][c] +
                        Atau[d][i][2] * A0[2][c] + Atau[d][i][3] * A0[3][c] +
                        Atau[d][i][4] * A0[4][c];
  }
  for (int a = 0; a < 4; ++a) {
    const double tmp1 = shape[a] * WdetJ * P->almi / P->gami / P->alfi * P->Dtgl;
    for (int i = 0; i < 5; ++i)
      for (int c = 0; c < 5; ++c)
        BDl[a][i + 5 * c] = BDl[a][i + 5 * c] +
                            tmp1 * (shg[a][0] * AtA0[0][i][c] + shg[a][1] * AtA0[1][i][c] +
                                    shg[a][2] * AtA0[2][i][c]);
  }
  static const int kl[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  for (int b = 0; b < 6; ++b) {
    const int k = kl[b][0], l = kl[b][1];
    double M[5][5];
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) {
        /* row of Atau_k read for each m (the entry's own row but for the
         * eight statements above) */
        double X[5];
        for (int m = 0; m < 5; ++m) X[m] = Atau[k][i][m];
        if (k == l) {
          if (k == 0 && i == 4 && j == 2) X[4] = Atau[0][2][4];
          if (k == 1 && j == 2 && i >= 2) {
            X[0] = Atau[1][1][0];
            X[1] = Atau[1][1][1];
            X[3] = Atau[1][1][3];
            if (i == 4) X[4] = Atau[1][2][4];
          }
          if (k == 2 && j == 3 && i >= 2) {
            X[0] = Atau[2][1][0];
            X[1] = Atau[2][1][1];
            X[2] = Atau[2][1][2];
          }
          if (k == 2 && i == 4 && j == 2) X[4] = Atau[2][2][4];
        }
        const double(*Al)[5] = Am[l];
        double acc = X[0] * Al[0][j];
        for (int m = 1; m < 5; ++m) acc = acc + X[m] * Al[m][j];
        if (k != l) {
          const double(*Ak)[5] = Am[k];
          for (int m = 0; m < 5; ++m) acc = acc + Atau[l][i][m] * Ak[m][j];
        }
        M[i][j] = acc;
      }
    e3bdg_visc(k, l, u1, u2, u3, rmu, rlm2mu, con, M);
    for (int a = 0; a < 4; ++a) {
      const double tmp = WdetJ * shg[a][k] * shg[a][l];
      for (int c = 0; c < 5; ++c)
        for (int i = 0; i < 5; ++i) BDl[a][i + 5 * c] = BDl[a][i + 5 * c] + tmp * M[i][c];
    }
  }
}

/* e3 of one element (e3.f:95-300) in the reference's three residual modes
 * (common /genpar/ ires): 1 the residual rl, 2 the modified residual rml of
 * the matrix-free solver, 3 both.  ylp is e3's yl: the variables whose point
 * values dui and gradients enter (e3ivar.f:96-101,259-334); yl is e3's ycl:
 * the state (pressure, velocity, temperature) at the points
 * (e3ivar.f:150-155).  ElmGMRs and ElmMFG pass the same array for both;
 * ItrRes/AsIRes pass the perturbed variables as ylp (asires.f:56-59,83-86).
 * BDl (non-NULL) receives e3bdg's block diagonal (e3.f:261-272, with
 * iprec = 1 and lhs != 1). */
static void e3_body(const orc_params* P, const double ylp[4][5], const double yl[4][5],
                    const double acl[4][5], const double xl[4][3], const double ql[4][12],
                    int ires, double rl[4][5], double rml[4][5], double EG[20][20],
                    int want_lhs, double BDl[4][25]) {
  const quad_t* Q = quad();
  const int lhs = want_lhs;
  const int want_r = ires == 1 || ires == 3, want_m = ires != 1;
  if (want_r) memset(rl, 0, sizeof(double) * 20);
  if (want_m) memset(rml, 0, sizeof(double) * 20);
  if (BDl) memset(BDl, 0, sizeof(double) * 100);
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
    /* e3ivar.f:94-129 dui: the conservative variables (rho, rho u, rho E)
     * of e3's yl at the point (getthm ithm = 6; used by the modified
     * residual only) */
    double dui[5] = {0, 0, 0, 0, 0};
    if (want_m) {
      for (int n = 0; n < 4; ++n)
        for (int f = 0; f < 5; ++f) dui[f] = dui[f] + shape[n] * ylp[n][f];
      const double rkp = PT5 * (dui[1] * dui[1] + dui[2] * dui[2] + dui[3] * dui[3]);
      const double rhop = dui[0] / (Rgas * dui[4]);
      const double eip = dui[4] * (Rgas / gamma1);
      dui[0] = rhop;
      dui[1] = rhop * dui[1];
      dui[2] = rhop * dui[2];
      dui[3] = rhop * dui[3];
      dui[4] = rhop * (eip + rkp);
    }
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
      g1[f] = g1[f] + shg[0][0] * ylp[0][f] + shg[1][0] * ylp[1][f] +
              shg[2][0] * ylp[2][f] + shg[3][0] * ylp[3][f];
      g2[f] = g2[f] + shg[0][1] * ylp[0][f] + shg[1][1] * ylp[1][f] +
              shg[2][1] * ylp[2][f] + shg[3][1] * ylp[3][f];
      g3[f] = g3[f] + shg[0][2] * ylp[0][f] + shg[1][2] * ylp[1][f] +
              shg[2][2] * ylp[2][f] + shg[3][2] * ylp[3][f];
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
    /* e3mtrx.f:232-298 (idc != 0): A0DC (the DC norm's metric), A0inv
     * (symmetric storage, 15) and dV/dY (entropy variables, 15); cv is
     * recomputed locally there from cp, alfap, T, rho, betaT */
    double A0DC[4] = {0, 0, 0, 0}, A0inv[15] = {0}, dVdY[15] = {0};
    if (P->idc != 0) {
      double s1 = ONE / (rho * rho * betaT * T);
      double cvl = cp - (alfap * alfap * T / rho / betaT);
      A0DC[0] = (rho * betaT) * (rho * betaT) * s1;
      A0DC[1] = -rho * alfap * rho * betaT * s1;
      A0DC[2] = rho / T;
      A0DC[3] = (-rho * alfap) * (-rho * alfap) * s1 + (rho * cvl / (T * T));
      double fact1 = ONE / (rho * cvl * (T * T));
      double d = alfap * T / rho / betaT;
      double e1bar = h - rk;
      double e2bar = e1bar - d;
      double e3bar = e2bar - cvl * T;
      double e5bar = e1bar * e1bar - 2.0 * e1bar * d + 2.0 * rk * cvl * T + cp * T / rho / betaT;
      double c1bar = u1 * u1 + cvl * T;
      double c2bar = u2 * u2 + cvl * T;
      double c3bar = u3 * u3 + cvl * T;
      double u12 = u1 * u2, u31 = u3 * u1, u23 = u2 * u3;
      A0inv[0] = e5bar * fact1;
      A0inv[1] = c1bar * fact1;
      A0inv[2] = c2bar * fact1;
      A0inv[3] = c3bar * fact1;
      A0inv[4] = 1.0 * fact1;
      A0inv[5] = u1 * e3bar * fact1;
      A0inv[6] = u2 * e3bar * fact1;
      A0inv[7] = u3 * e3bar * fact1;
      A0inv[8] = -e2bar * fact1;
      A0inv[9] = u12 * fact1;
      A0inv[10] = u31 * fact1;
      A0inv[11] = -u1 * fact1;
      A0inv[12] = u23 * fact1;
      A0inv[13] = -u2 * fact1;
      A0inv[14] = -u3 * fact1;
      double f1 = 1.0 / T;
      double f2 = f1 / T;
      dVdY[0] = f1 / rho;
      dVdY[1] = -f1 * u1;
      dVdY[2] = f1;
      dVdY[3] = -f1 * u2;
      dVdY[4] = 0.0;
      dVdY[5] = f1;
      dVdY[6] = -f1 * u3;
      dVdY[7] = 0.0;
      dVdY[8] = 0.0;
      dVdY[9] = f1;
      dVdY[10] = -(h - rk) * f2;
      dVdY[11] = -f2 * u1;
      dVdY[12] = -f2 * u2;
      dVdY[13] = -f2 * u3;
      dVdY[14] = f2;
    }

    /* ---- e3conv.f:72-96 Euler flux (integrated by parts), ires 1 or 3 ---- */
    if (want_r) {
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
    }
    /* e3conv.f:100-179 rLyi = A_i Y,i */
    double rLyi[5];
    for (int i = 0; i < 5; ++i)
      rLyi[i] = A1[i][0] * g1[0] + A1[i][1] * g1[1] + A1[i][2] * g1[2] +
                A1[i][3] * g1[3] + A1[i][4] * g1[4] + A2[i][0] * g2[0] +
                A2[i][1] * g2[1] + A2[i][2] * g2[2] + A2[i][3] * g2[3] +
                A2[i][4] * g2[4] + A3[i][0] * g3[0] + A3[i][1] * g3[1] +
                A3[i][2] * g3[2] + A3[i][3] * g3[3] + A3[i][4] * g3[4];
    /* e3conv.f:183-184: the modified residual takes the convective term in
     * non-integrated form */
    if (want_m)
      for (int i = 0; i < 5; ++i) rmi[15 + i] = rLyi[i];
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

    /* ---- e3ls.f:98-103 the modified LS residual: A_i Y,i + fct1 Y ---- */
    double rLymi[5] = {0, 0, 0, 0, 0};
    if (want_m) {
      const double fct1 = P->almi / P->gami / P->alfi * P->Dtgl;
      for (int i = 0; i < 5; ++i) rLymi[i] = rLyi[i] + fct1 * dui[i];
    }
    /* ---- e3ls.f:105-143 add the time term to the LS residual ---- */
    if (want_r)
    for (int i = 0; i < 5; ++i)
      rLyi[i] = rLyi[i] + A0[i][0] * aci[0] + A0[i][1] * aci[1] +
                A0[i][2] * aci[2] + A0[i][3] * aci[3] + A0[i][4] * aci[4];
    /* ---- e3ivar.f:361-395 divergence of the reconstructed diffusive flux,
     * e3ls.f:148-155 subtracted from the LS residual (isurf = 0) ---- */
    if (ql && want_r) {
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
    /* ---- e3tau.f:1-283 (itau = 0 or 1, ipord = 1); e3tau_nd (itau 10 and
     * 11: e3ls.f:188-214 calls it for every itau >= 10, and it reads no
     * itau) ---- */
    double gijd[6];
    double PTau[5][5];
    if (P->itau >= 10) {
      /* compK is formed in e3visc from the unrotated dxidx (e3visc.f:143);
       * e3tau_nd then rotates dxidx, and its DC giju uses the rotated rows */
      double compK[10];
      e3compk((const double(*)[3])dxidx, T, rmu, rlm, rlm2mu, con, u1, u2, u3, compK);
      e3tau_nd(P, rho, cp, T, u1, u2, u3, aci[1], aci[2], aci[3], dxidx, compK, PTau);
    }
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
    if (P->itau == 1) {
      /* e3tau.f:54-139 (Diagonal-Franca; ipord = 1: celt = pt66; the
       * reference's 1.0e-15 is a default-kind real literal) */
      const double g1 = gijd[0], g2 = gijd[1], g3 = gijd[2], g4 = gijd[3], g5 = gijd[4],
                   g6 = gijd[5];
      double fa = g1 * g3 * g6 - g1 * g5 * g5 - g3 * g4 * g4 - g6 * g2 * g2 + g2 * g4 * g5 * TWO;
      double uh1 = u1 * u1 * (g3 * g6 - g5 * g5) + u2 * u2 * (g1 * g6 - g4 * g4) +
                   u3 * u3 * (g1 * g3 - g2 * g2) +
                   TWO * (u1 * u2 * (g4 * g5 - g2 * g6) + u1 * u3 * (g2 * g5 - g4 * g3) +
                          u2 * u3 * (g4 * g2 - g1 * g5));
      uh1 = TWO * sqrt(uh1 / fa);
      double h2o2u = u1 * u1 * g1 + u2 * u2 * g3 + u3 * u3 * g6 +
                     (u1 * u2 * g2 + u1 * u3 * g4 + u2 * u3 * g5) * TWO + (double)1.0e-15f;
      h2o2u = ONE / sqrt(h2o2u);
      fa = PT5 * rho * uh1 / rmu;
      tau[0] = PT5 * uh1 * fmin(ONE, fa) * P->taucfct;
      /* iremoveStabTimeTerm = 0 */
      const double dtsn = ONE / (P->Dtgl * P->dtsfct);
      tau[1] = fmin(dtsn, h2o2u);
      tau[1] = tau[1] / rho;
      tau[2] = tau[1] / cv;
      fa = h2o2u * h2o2u * rk * PT66 / rmu;
      tau[1] = fmin(tau[1], fa);
      tau[2] = fmin(tau[2], fa * rmu / con) * P->temper;
    }
    double Lt[5]; /* e3tau.f:188: rLyitemp, the LS residual before tau */
    for (int f = 0; f < 5; ++f) Lt[f] = rLyi[f];
    if (!want_r) {
      /* e3tau.f:185 (and :638): rLyi keeps A_i Y,i for ires = 2 */
    } else if (P->itau >= 10) {
      /* e3tau.f:633-643 rLyi = Tau rLyi */
      for (int i = 0; i < 5; ++i) {
        double acc = 0.0;
        for (int j = 0; j < 5; ++j) acc = acc + PTau[i][j] * Lt[j];
        rLyi[i] = acc;
      }
    } else {
      rLyi[0] = tau[0] * rLyi[0];
      rLyi[1] = tau[1] * rLyi[1];
      rLyi[2] = tau[1] * rLyi[2];
      rLyi[3] = tau[1] * rLyi[3];
      rLyi[4] = tau[2] * rLyi[4];
    }
    /* e3tau.f:263-271 and :723-732: tau times the modified LS residual */
    if (want_m) {
      if (P->itau >= 10) {
        double em[5];
        for (int j = 0; j < 5; ++j) em[j] = rLymi[j];
        for (int i = 0; i < 5; ++i) {
          double acc = 0.0;
          for (int j = 0; j < 5; ++j) acc = acc + PTau[i][j] * em[j];
          rLymi[i] = acc;
        }
      } else {
        rLymi[0] = tau[0] * rLymi[0];
        rLymi[1] = tau[1] * rLymi[1];
        rLymi[2] = tau[1] * rLymi[2];
        rLymi[3] = tau[1] * rLymi[3];
        rLymi[4] = tau[2] * rLymi[4];
      }
    }
    /* e3tau.f:198-268 (iDC != 0): rTLS = L.(dV/dY tau L), raLS = L.A0inv L
     * and giju, the inverse of the metric tensor */
    double rTLS = 0.0, raLS = 0.0, giju[6] = {0, 0, 0, 0, 0, 0};
    if (P->idc != 0 && want_r) {
      const double* L = Lt;
      const double* V = dVdY;
      const double* r = rLyi;
      rTLS = L[0] * (r[0] * V[0] + V[1] * r[1] + V[3] * r[2] + r[3] * V[6] + V[10] * r[4]) +
             L[1] * (r[1] * V[2] + r[2] * V[4] + V[7] * r[3] + r[4] * V[11]) +
             L[2] * (r[2] * V[5] + V[8] * r[3] + V[12] * r[4]) +
             L[3] * (r[3] * V[9] + V[13] * r[4]) + L[4] * (V[14] * r[4]);
      const double* Ai = A0inv;
      raLS = TWO * L[3] * L[4] * Ai[14] + TWO * L[2] * L[4] * Ai[13] +
             TWO * L[0] * L[1] * Ai[5] + TWO * L[1] * L[2] * Ai[9] +
             TWO * L[1] * L[3] * Ai[10] + TWO * L[0] * L[2] * Ai[6] +
             TWO * L[2] * L[3] * Ai[12] + TWO * L[1] * L[4] * Ai[11] +
             TWO * L[0] * L[3] * Ai[7] + TWO * L[0] * L[4] * Ai[8] + L[0] * L[0] * Ai[0] +
             L[1] * L[1] * Ai[1] + L[2] * L[2] * Ai[2] + L[3] * L[3] * Ai[3] +
             L[4] * L[4] * Ai[4];
      /* gijdu: 11, 22, 33, 12, 13, 23 */
      const double m1 = gijd[0], m2 = gijd[2], m3 = gijd[5], m4 = gijd[1], m5 = gijd[3],
                   m6 = gijd[4];
      double detI = ONE / (m1 * m2 * m3 - m1 * m6 * m6 - m4 * m4 * m3 + m4 * m5 * m6 * TWO -
                           m5 * m5 * m2);
      giju[0] = detI * (m2 * m3 - m6 * m6);
      giju[1] = detI * (m1 * m3 - m5 * m5);
      giju[2] = detI * (m1 * m2 - m4 * m4);
      giju[3] = detI * (m5 * m6 - m4 * m3);
      giju[4] = detI * (m4 * m6 - m5 * m2);
      giju[5] = detI * (m4 * m5 - m1 * m6);
    }
    /* e3ls.f:231-348 modified residual rmi += A_i (tau Lm) */
    if (want_m) {
      for (int i = 0; i < 5; ++i) rmi[i] = rowdot(A1, i, rLymi) + rmi[i];
      for (int i = 0; i < 5; ++i) rmi[5 + i] = rowdot(A2, i, rLymi) + rmi[5 + i];
      for (int i = 0; i < 5; ++i) rmi[10 + i] = rowdot(A3, i, rLymi) + rmi[10 + i];
    }
    /* e3ls.f:351-458 RHS ri += A_i (tau L) */
    if (want_r) {
    for (int i = 0; i < 5; ++i) ri[i] = rowdot(A1, i, rLyi) + ri[i];
    for (int i = 0; i < 5; ++i) ri[5 + i] = rowdot(A2, i, rLyi) + ri[5 + i];
    for (int i = 0; i < 5; ++i) ri[10 + i] = rowdot(A3, i, rLyi) + ri[10 + i];
    }
    /* e3ls.f:464-766 LHS */
    if (lhs) {
      const double* Ai[3] = {&A1[0][0], &A2[0][0], &A3[0][0]};
      double AitauA0[3][5][5];
      for (int k = 0; k < 3; ++k) {
        const double (*Ak)[5] = (const double (*)[5])Ai[k];
        double Atau[5][5];
        if (P->itau >= 10) {
          /* e3ls.f:477-486 Atau = A_k PTau */
          for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) {
              double acc = 0.0;
              for (int m = 0; m < 5; ++m) acc = acc + Ak[i][m] * PTau[m][j];
              Atau[i][j] = acc;
            }
        } else {
          for (int i = 0; i < 5; ++i) {
            Atau[i][0] = Ak[i][0] * tau[0];
            Atau[i][1] = Ak[i][1] * tau[1];
            Atau[i][2] = Ak[i][2] * tau[1];
            Atau[i][3] = Ak[i][3] * tau[1];
            Atau[i][4] = Ak[i][4] * tau[2];
          }
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
    /* ---- e3dc.f:44-451 discontinuity capturing (e3.f:217-222); the
     * ires = 2 branch (e3dc.f:380-409) is not restated: the drivers of the
     * modified residual refuse iDC != 0 ---- */
    if (P->idc != 0 && want_r) {
      const double* gy[3] = {g1, g2, g3};
      double A0g[15];
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 5; ++i)
          A0g[5 * j + i] = A0[i][0] * gy[j][0] + A0[i][1] * gy[j][1] + A0[i][2] * gy[j][2] +
                           A0[i][3] * gy[j][3] + A0[i][4] * gy[j][4];
      /* giju rows: (1 4 5), (4 2 6), (5 6 3) */
      static const int gr[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
      double gA[15];
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 5; ++i)
          gA[5 * j + i] = giju[gr[j][0]] * A0g[i] + giju[gr[j][1]] * A0g[5 + i] +
                          giju[gr[j][2]] * A0g[10 + i];
      double yAy[6];
      for (int j = 0; j < 3; ++j) {
        const double* g = gy[j];
        yAy[j] = A0DC[0] * (g[0] * g[0]) + TWO * g[0] * A0DC[1] * g[4] +
                 A0DC[2] * (g[1] * g[1]) + A0DC[2] * (g[2] * g[2]) + A0DC[2] * (g[3] * g[3]) +
                 A0DC[3] * (g[4] * g[4]);
      }
      static const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
      for (int k = 0; k < 3; ++k) {
        const double* a = gy[pr[k][0]];
        const double* b = gy[pr[k][1]];
        yAy[3 + k] = a[0] * A0DC[0] * b[0] + a[0] * A0DC[1] * b[4] + a[1] * A0DC[2] * b[1] +
                     a[2] * A0DC[2] * b[2] + a[3] * A0DC[2] * b[3] + a[4] * A0DC[1] * b[0] +
                     a[4] * A0DC[3] * b[4];
      }
      double gnorm = ONE / (giju[0] * yAy[0] + TWO * giju[3] * yAy[3] + TWO * giju[4] * yAy[4] +
                            giju[1] * yAy[1] + TWO * giju[5] * yAy[5] + giju[2] * yAy[2] + P->epsM);
      double DC = 0.0;
      if (P->idc == 1) {
        const double fact = ONE; /* ipord = 1 */
        DC = fmax(0.0, (fact * sqrt(raLS * gnorm)) - (rTLS * gnorm));
      } else if (P->idc == 2) {
        DC = TWO * rTLS * gnorm;
      } else if (P->idc == 3) {
        const double fact = ONE;
        DC = fmin(fmax(0.0, fact * sqrt(raLS * gnorm) - rTLS * gnorm), TWO * rTLS * gnorm);
      }
      /* ires = 1 */
      for (int k = 0; k < 15; ++k) ri[k] = ri[k] + DC * gA[k];
      /* e3dc.f:410-446 (iprec = 1): stiff[I,J] += DC A0 giju(IJ) */
      if (lhs && P->iprec == 1) {
        for (int j = 0; j < 5; ++j)
          for (int i = 0; i < 5; ++i) {
            double dtmp = A0[i][j] * DC;
            stiff[i][j] = stiff[i][j] + dtmp * giju[0];
            stiff[i][j + 5] = stiff[i][j + 5] + dtmp * giju[3];
            stiff[i][j + 10] = stiff[i][j + 10] + dtmp * giju[4];
            stiff[i + 5][j] = stiff[i + 5][j] + dtmp * giju[3];
            stiff[i + 5][j + 5] = stiff[i + 5][j + 5] + dtmp * giju[1];
            stiff[i + 5][j + 10] = stiff[i + 5][j + 10] + dtmp * giju[5];
            stiff[i + 10][j] = stiff[i + 10][j] + dtmp * giju[4];
            stiff[i + 10][j + 5] = stiff[i + 10][j + 5] + dtmp * giju[5];
            stiff[i + 10][j + 10] = stiff[i + 10][j + 10] + dtmp * giju[2];
          }
      }
    }
    /* ---- e3massr.f:34-72 ---- */
    if (want_r)
    for (int i = 0; i < 5; ++i)
      ri[15 + i] = ri[15 + i] + A0[i][0] * aci[0] + A0[i][1] * aci[1] +
                   A0[i][2] * aci[2] + A0[i][3] * aci[3] + A0[i][4] * aci[4];
    if (want_m) {
      const double fct1 = P->almi / P->gami / P->alfi * P->Dtgl;
      for (int i = 0; i < 5; ++i) rmi[15 + i] = rmi[15 + i] + fct1 * dui[i];
    }
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
    /* ---- e3.f:261-272 the block diagonal of the matrix-free solver ---- */
    if (BDl && P->itau >= 10)
      e3bdg_nd(P, shape, (const double(*)[3])shg, WdetJ, A1, A2, A3, A0,
               (const double(*)[5])PTau, u1, u2, u3, rmu, rlm2mu, con, BDl);
    else if (BDl)
      e3bdg(P, shape, (const double(*)[3])shg, WdetJ, A1, A2, A3, A0, tau, u1, u2, u3, rmu,
            rlm2mu, con, BDl);
    /* ---- e3wmlt.f:72-143 RHS ---- */
    if (want_r) {
    for (int i = 0; i < 4; ++i)
      for (int f = 0; f < 5; ++f)
        rl[i][f] = rl[i][f] + WdetJ * (shg[i][0] * ri[f] + shg[i][1] * ri[5 + f] +
                                       shg[i][2] * ri[10 + f]);
    }
    if (want_m)
      for (int i = 0; i < 4; ++i)
        for (int f = 0; f < 5; ++f)
          rml[i][f] = rml[i][f] + WdetJ * (shg[i][0] * rmi[f] + shg[i][1] * rmi[5 + f] +
                                           shg[i][2] * rmi[10 + f] + shape[i] * rmi[15 + f]);
    /* e3wmlt.f:124-143 runs for every ires; with ires = 2 ri(16:20) is zero */
    if (want_r)
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

void orc_e3_q(const orc_params* P, const double yl[4][5],
              const double acl[4][5], const double xl[4][3],
              const double ql[4][12], double rl[4][5], double EG[20][20],
              int want_lhs) {
  e3_body(P, yl, yl, acl, xl, ql, 1, rl, NULL, EG, want_lhs, NULL);
}

void orc_e3_mfg(const orc_params* P, const double ylp[4][5], const double yl[4][5],
                const double acl[4][5], const double xl[4][3], const double ql[4][12],
                int ires, double rl[4][5], double rml[4][5], double BDl[4][25]) {
  e3_body(P, ylp, yl, acl, xl, ql, ires, rl, rml, NULL, 0, BDl);
}

void orc_e3(const orc_params* P, const double yl[4][5],
            const double acl[4][5], const double xl[4][3], double rl[4][5],
            double EG[20][20], int want_lhs) {
  orc_e3_q(P, yl, acl, xl, NULL, rl, EG, want_lhs);
}
