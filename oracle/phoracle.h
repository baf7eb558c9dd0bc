/*
 * phoracle.h -- CPU oracle for the compressible hot path (test
 * infrastructure only: imported by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py; never by the product path).
 *
 * A plain-C FP64 restatement of PHASTA's ElmGMRs -> AsIGMR -> e3 assembly
 * and SolGMRs block-diagonal GMRES, following the reference operation order
 * (file:line citations in phoracle.c).  Arrays use the reference's Fortran
 * layouts: column-major, node index fastest, 1-based connectivity values.
 */
#ifndef PHORACLE_H
#define PHORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh {
  long long nshg;     /* nodes incl. periodic slaves (numnp == nshg, p=1) */
  long long nel;      /* interior tets                                    */
  long long nelb;     /* boundary (wall) tets                             */
  long long nnz_tot;  /* 5x5 blocks in lhsK                               */
  int ndofBC;         /* second dimension of BC                           */
  int ibksiz;         /* elements per block (input.config:255, 64)        */
  const double* x;    /* (nshg,3)                                         */
  const int* ien;     /* (nel,4)  1-based                                 */
  const int* ienb;    /* (nelb,4) 1-based, face nodes first               */
  const int* iBC;     /* (nshg)                                           */
  const double* BC;   /* (nshg,ndofBC)                                    */
  const int* iper;    /* (nshg) 1-based                                   */
  const int* colm;    /* (nshg+1) 1-based row pointers                    */
  const int* rowp;    /* (nnz_tot) 1-based sorted column indices          */
} orc_mesh;

typedef struct orc_params {
  double Rgas, gamma, gamma1, pr, mu;   /* /mmatpar/, datmat(1,2,1)        */
  double dtsfct, taucfct, temper;       /* /genpar/, /outpar/              */
  double almi, alfi, gami, Dtgl;        /* /timdat/                        */
  double etol, epsM;                    /* epstol(1), /precis/             */
  int Kspace, nGMRES, minIters;         /* /solpar/, /incomp/              */
  int lhs, iprec;                       /* per-call flags (itrdrv.f:456)   */
  int ipred;                            /* predictor 1-4 (itrPC.f:67-119)  */
  int idiff;                            /* diffusive flux reconstruction   */
  int iter;                             /* /timdat/ Newton iteration of the
                                           time step (itrdrv.f:409,434)    */
  int idc;                              /* /solpar/ iDC 0-3 (e3dc.f)       */
  int itau;                             /* /genpar/ itau: 0 diagonal (Shakib),
                                           10 matrix tau (e3tau_nd)         */
} orc_params;

typedef struct orc_stats {
  int iKs;          /* Krylov vectors used in the last cycle              */
  int lGMRES;       /* cycles - 1                                         */
  int ntotGM;       /* running total of Krylov iterations                 */
  int converged;    /* |e_{iKs+1}| <= epsnrm                              */
  double unorm0;    /* ||L^-1 r|| at entry                                */
  double epsnrm;    /* etol * unorm0                                      */
  double echeck;    /* |e_{iKs+1}| at exit                                */
  double totres[2]; /* rstat: sqrt(sum res^2)/nshg for res and rmes       */
  double reslast;   /* rstatCheck memory (input/output)                   */
  long long nnan;   /* tnanq counts over res and BDiag                    */
  long long nlarge;
} orc_stats;

/* quadrature/shape tables for linear tets (symtet.c, symtri.c, genint.f,
 * genshp.f, genshpb.f, shptet.c).  shp[q][a], shgl[q][j][a], Qwt[q];
 * boundary: shpb[q][a], shglb[q][j][a], Qwtb[q] (3 points). */
void orc_quadrature(double shp[4][4], double shgl[4][3][4], double Qwt[4],
                    double shpb[3][4], double shglb[3][3][4], double Qwtb[3]);

/* e3 for one element; yl/acl local order (p,u,v,w,T).  rl[a][f] and
 * EGmass[5a+f][5b+g] are overwritten (not accumulated). */
void orc_e3(const orc_params* pr, const double yl[4][5],
            const double acl[4][5], const double xl[4][3], double rl[4][5],
            double EGmass[20][20], int want_lhs);

/* e3 with the element's nodal reconstructed diffusive flux ql[a][0..11]
 * (qres gathered, asigmr.f:55): its divergence is subtracted from the
 * least-squares residual (e3ivar.f:361-395, e3ls.f:148-155).  ql = NULL
 * is orc_e3 (idiff = 0). */
void orc_e3_q(const orc_params* pr, const double yl[4][5],
              const double acl[4][5], const double xl[4][3],
              const double ql[4][12], double rl[4][5], double EGmass[20][20],
              int want_lhs);

/* e3 in its residual modes ires (1 res, 2 modified residual, 3 both;
 * e3.f:95-300).  ylp: e3's yl (point values and gradients), yl: e3's ycl
 * (the state); the same array except in ItrRes.  rl / rml / BDl (e3bdg's
 * BDiagl(a,i,k) at [a][i+5k], or NULL) are overwritten. */
void orc_e3_mfg(const orc_params* pr, const double ylp[4][5], const double yl[4][5],
                const double acl[4][5], const double xl[4][3], const double ql[4][12],
                int ires, double rl[4][5], double rml[4][5], double BDl[4][25]);

/* AsIq over all interior elements + qpbc (elmgmr.f:361-391, asiq.f,
 * e3q.f, e3qvar.f, qpbc.f): qres (nshg,12) the reconstructed diffusive
 * flux, rmass (nshg) the inverted lumped mass.  idiff 1 (row-sum lumping)
 * or 3 (special lumping); returns 1 for other idiff. */
int orc_asiq(const orc_mesh* m, const orc_params* pr, const double* y,
             double* qres, double* rmass);

/* e3b for one wall element (iBCB = 0: floating fluxes). rl overwritten. */
void orc_e3b(const orc_params* pr, const double yl[4][5],
             const double xl[4][3], double rl[4][5]);

/* ElmGMRs (elmgmr.f:280-612).  res,rmes (nshg,5); BDiag (nshg,5,5);
 * lhsK (25,nnz_tot).  lhsK/BDiag touched only when pr->lhs / pr->iprec.
 * idiff 1 or 3 runs orc_asiq first; idiff 2 (qres never formed in the
 * reference) and other values return 5. */
int orc_elmgmrs(const orc_mesh* m, const orc_params* pr, const double* y,
                const double* ac, double* res, double* rmes, double* BDiag,
                double* lhsK);

/* i3LU (i3lu.f): code 0 = 'LU_Fact ', 1 = 'forward ', 2 = 'backward',
 * 3 = 'product '. */
int orc_i3lu(long long nshg, double* Diag, double* r, int code);
void orc_spsi3pre(long long nshg, const double* BDiag, double* lhsK,
                  const int* colm, const int* rowp);
void orc_sparse_ap(long long nshg, const int* iper, const int* colm,
                   const int* rowp, const double* lhsK, double* p);
void orc_bc3per(long long nshg, const int* iBC, const int* iper, double* res,
                int nQs);

/* SolGMRs (solgmr.f:368-744) including ElmGMRs.  Dy (nshg,5) output;
 * res returns the preconditioned residual, rmes the raw one. */
int orc_solgmrs(const orc_mesh* m, const orc_params* pr, const double* y,
                const double* ac, double* res, double* rmes, double* BDiag,
                double* lhsK, double* Dy, orc_stats* st);

/* Matrix-free GMRES (impl = 20, itrdrv.f:489-506).  The frozen flags plus
 * idiff 0, 1, 3; iDC != 0 and itau != 0 return 5.
 * ElmMFG (elmmfg.f): res, the modified residual rmes and (iprec) e3bdg's
 * BDiag, after bc3Res / bc3BDg. */
int orc_elmmfg(const orc_mesh* m, const orc_params* pr, const double* y,
               const double* ac, double* res, double* rmes, double* BDiag);
/* ItrRes (itrres.f): rmes += the modified residual of the variables yp
 * (global order) at the state yc, interior elements only, then bc3Res;
 * iabres: element |rml| (itrFDI's cancellation estimate). */
int orc_itrres(const orc_mesh* m, const orc_params* pr, const double* yp,
               const double* yc, const double* ac, double* rmes, int iabres);
/* yshuffle (shuffle.f): new2old (u,v,w,p,T) -> (p,u,v,w,T), else back */
void orc_yshuffle(long long nshg, double* r, int new2old);
/* itrFDI (itrfdi.f:1-141): the finite-difference interval eGMRES from the
 * preconditioned base point ypre, rmes (forward-reduced) and direction u */
int orc_itrfdi(const orc_mesh* m, const orc_params* pr, const double* ypre,
               const double* y, const double* ac, const double* rmes,
               const double* u, const double* BDiag, double* eGMRES);
/* Au1MFG (au1mfg.f): u <- (L^-1 Rmod(y(ypre + eGMRES u)) - rmes) / eGMRES */
int orc_au1mfg(const orc_mesh* m, const orc_params* pr, const double* ypre,
               const double* y, const double* ac, const double* rmes,
               const double* BDiag, double eGMRES, double* u);
/* Au2MFG (au2mfg.f): u <- res - the two-sided difference along Dy */
int orc_au2mfg(const orc_mesh* m, const orc_params* pr, const double* ypre,
               const double* y, const double* ac, const double* res,
               const double* Dy, const double* BDiag, double* u);
/* SolMFG (solmfg.f:1-381) including ElmMFG.  eGMRES: the /itrpar/ interval
 * (in/out); fdi: recompute it by itrFDI first (the reference does so when
 * iter = 1 and mod(istep, 20) = 0, solmfg.f:163).  res returns
 * preconditioned, rmes forward-reduced, BDiag LU-factored (iprec), Dy the
 * increment consumed by itrCorrect. */
int orc_solmfg(const orc_mesh* m, const orc_params* pr, const double* y,
               const double* ac, double* res, double* rmes, double* BDiag,
               double* Dy, double* eGMRES, int fdi, orc_stats* st);

/* itrPredict / itrCorrect / itrBC (itrPC.f:56-148, itrbc.f:1-198);
 * y, ac, yold, acold are (nshg,5) in the global order (u,v,w,p,T).
 * orc_itr_predict returns 1 for ipred != 1 (no defined reference result). */
int orc_itr_predict(long long nshg, const orc_params* pr, const double* yold,
                    const double* acold, double* y, double* ac);
void orc_itr_correct(long long nshg, const orc_params* pr, double* y,
                     double* ac, const double* yold, const double* acold,
                     const double* Dy);
void orc_itr_update(long long nshg, const orc_params* pr, double* yold,
                    double* acold, const double* y, const double* ac);
int orc_itr_bc(const orc_mesh* m, double* y, double* ac);
/* OpenMP threads of the parallel loops (1: the serial oracle); returns the
 * count in effect, 1 when built without OpenMP. */
int orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
