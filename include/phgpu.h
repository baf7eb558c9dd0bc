/*
 * phgpu.h -- C-ABI of the MI355X (gfx950) implementation of PHASTA's
 * compressible implicit hot path: ElmGMRs/AsIGMR/e3 assembly and the
 * block-diagonal-preconditioned sparse GMRES SolGMRs.
 *
 * Every array argument keeps the reference's Fortran layout (column-major,
 * node index fastest, 1-based connectivity values) so that the Fortran host
 * can pass its own arrays through iso_c_binding unchanged:
 *   y, ac, yold, acold (nshg,ndof)  global dof order (u,v,w,p,T)
 *   res, rmes, Dy      (nshg,5)     equation order (continuity, mom x/y/z,
 *                                   energy) = local dof order (p,u,v,w,T)
 *   BDiag              (nshg,5,5)
 *   lhsK               (25,nnz_tot) 5x5 block, column-major inside a block
 *   colm (nshg+1), rowp (nnz_tot)   CSR (colm = row pointers, rowp = sorted
 *                                   column indices; genadj.f:56-65)
 * All entry points return 0 on success and a nonzero PHG_E* code otherwise
 * (the reference aborts through error(); the Fortran shim converts a nonzero
 * status into `call error('SolGMRs ','hip     ',code)`).  The library is not
 * reentrant: one context per process and GPU, one host thread, one stream.
 */
#ifndef PHGPU_H
#define PHGPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PHG_OK = 0,
  PHG_EINVAL = 1,     /* bad argument / unsupported flag or BC code        */
  PHG_EHIP = 2,       /* HIP runtime error                                  */
  PHG_ENOMEM = 3,     /* device or host allocation failed                   */
  PHG_ENAN = 10,      /* tnanq: NaN or |x| > 1e10 in res or BDiag           */
  PHG_EDIVERGE = 11,  /* rstatCheck: residual grew by more than 1e4         */
  PHG_ECOLOR = 12     /* element colouring needs more than 64 colours      */
};

/* PHG_FP64_MATRIX_FREE: an FP64 context for SolMFG alone.  As itrdrv.f:163-185
 * allocates lhsK, rowp and colm only for the sparse solver (impl 1x) and
 * BDiag alone for the matrix-free one (impl 2x), it holds no lhsK, no matvec
 * layout and no owner lanes of the off-diagonal entries (about 13 GB less at
 * 12.6 M elements); the entry points that need lhsK return PHG_EINVAL. */
enum { PHG_FP64 = 0, PHG_MIXED = 1, PHG_FP64_MATRIX_FREE = 2 };

/* COMMON-block scalars read on the path (common.h:121-259); same field
 * order as the CPU oracle's parameter block. */
typedef struct phg_params {
  double Rgas, gamma, gamma1, pr, mu;   /* /mmatpar/, datmat(1,2,1)         */
  double dtsfct, taucfct, temper;       /* /genpar/, /outpar/               */
  double almi, alfi, gami, Dtgl;        /* /timdat/                         */
  double etol, epsM;                    /* epstol(1), /precis/              */
  int Kspace, nGMRES, minIters;         /* /solpar/, /incomp/               */
  int lhs, iprec;                       /* itrdrv.f:456-457                 */
  int ipred;                            /* /genpar/ predictor; 1 only      */
  int idiff;                            /* /genpar/ 0, 1 or 3 (elmgmr.f:361) */
  int iter;                             /* /timdat/ Newton iteration of the
                                           time step (itrdrv.f:409,434); the
                                           rstatCheck divergence abort applies
                                           when iter > 1 (rstatCheck.f:73).
                                           phg_predict sets 0 and
                                           phg_newton_iteration adds 1 before
                                           its solve; phg_solgmrs takes the
                                           caller's value.                  */
  int idc;                              /* /solpar/ iDC: 0 off, 1 DC-mallet,
                                           2 DC-quadratic, 3 DC-minimum
                                           (input_fform.cc:683-686, e3dc.f) */
  int itau;                             /* /genpar/ itau (input_fform.cc:
                                           663-674): 0 Shakib's diagonal tau
                                           (e3tau.f:140-176), 1 Franca's
                                           (e3tau.f:54-139), 10 and 11 the
                                           matrix tau of e3tau_nd
                                           (e3tau.f:284-744; it reads no
                                           itau); 2 and 3, which the
                                           compressible e3tau leaves unset,
                                           are refused                      */
} phg_params;

/* Mesh arrays as the Fortran host holds them (readnblk.f, genadj.f);
 * copied to the device by phg_init and not retained. */
typedef struct phg_mesh {
  long long nshg, nel, nelb, nnz_tot;
  int ndofBC, reserved;
  const double* x;     /* (nshg,3)                                         */
  const int* ien;      /* (nel,4)  interior tets, mien blocks concatenated  */
  const int* ienb;     /* (nelb,4) wall tets, face nodes 1..3 (mienb)       */
  const int* iBC;      /* (nshg)                                           */
  const double* BC;    /* (nshg,ndofBC)                                    */
  const int* iper;     /* (nshg)                                           */
  const int* colm;     /* (nshg+1)                                         */
  const int* rowp;     /* (nnz_tot)                                        */
} phg_mesh;

/* Solver statistics (the /itrpar/ outputs iKs, lGMRES, ntotGM read by
 * rstat.f:95-139, plus the rstat/rstatCheck norms). */
typedef struct phg_stats {
  int iKs, lGMRES, ntotGM, converged;
  double unorm0, epsnrm, echeck;
  double totres[2];      /* rstat: ||res||/nshg, ||rmes||/nshg             */
  double reslast;        /* rstatCheck memory across calls                 */
  long long nnan, nlarge;
  double ms_assembly, ms_solve, ms_total;
  double ms_spmv, ms_mgs; /* summed HIP-event times of the matvec and MGS  */
  int n_spmv, n_mgs;      /* launches behind ms_spmv / ms_mgs              */
  int refine_cycles;      /* GMRES cycles run (1 on the plain FP64 path)   */
  int pc_fallback;        /* learned PC: 0 its cycle converged (or unused),
                             1 discarded for block-diagonal GMRES, 2 finished
                             by block-diagonal cycles, 3 its first direction
                             left a larger one-step residual than the
                             block-diagonal direction and the cycle ran
                             without the network, 4 set aside because it
                             lost that test earlier in the same time step
                             (resident stepping: phg_predict ... phg_update) */
  double orth_loss;       /* max over the solve's cycles of
                             max_ij |u_i . u_j - delta_ij| of the Krylov basis
                             (phg_diagnostics flag PHG_DIAG_ORTH; else 0)   */
  double ms_point;        /* HIP-event time of the assembly's point pass     */
  double ms_blocks;       /* HIP-event time of its block (tangent) pass      */
} phg_stats;

typedef struct phg_ctx phg_ctx;

/* Set-up: copies mesh and BC data to the GPU, colours the elements, builds
 * the device block-CSR layout and the element-to-block map that replaces
 * sparseloc (fillsparse.f:243-270).  precision: PHG_FP64, PHG_MIXED or
 * PHG_FP64_MATRIX_FREE. */
int phg_init(const phg_mesh* mesh, const phg_params* prm, int precision,
             phg_ctx** out);
int phg_finalize(phg_ctx* ctx);
int phg_set_params(phg_ctx* ctx, const phg_params* prm);
const char* phg_last_error(void);
/* Informational: number of interior / boundary colours, device bytes. */
int phg_info(phg_ctx* ctx, int* ncolors, int* ncolors_b, long long* dev_bytes);

/* ElmGMRs(y,ac,x,shp,shgl,iBC,BC,shpb,shglb,res,rmes,BDiag,iper,ilwork,
 *         lhsK,col,row,rerr)                      elmgmr.f:280-285
 * Host arrays; lhsK may be NULL (it stays resident on the device). */
int phg_elmgmrs(phg_ctx* ctx, const double* y, const double* ac, double* res,
                double* rmes, double* BDiag, double* lhsK);

/* SolGMRs(y,ac,yold,acold,x,iBC,BC,col,row,lhsk,res,BDiag,HBrg,eBrg,yBrg,
 *         Rcos,Rsin,iper,ilwork,shp,shgl,shpb,shglb,Dy,rerr) solgmr.f:368
 * Host arrays.  BDiag returns LU-factored, lhsK (if non-NULL) scaled, res
 * preconditioned, rmes raw, Dy the increment consumed by itrCorrect. */
int phg_solgmrs(phg_ctx* ctx, const double* y, const double* ac,
                double* res, double* rmes, double* BDiag, double* lhsK,
                double* Dy, phg_stats* st);

/* Device-resident time stepping around the same call (itrdrv.f:398-599):
 * the state stays in HBM between calls. */
int phg_state_upload(phg_ctx* ctx, const double* yold, const double* acold);
int phg_state_download(phg_ctx* ctx, double* y, double* ac, double* Dy);
int phg_state_download_old(phg_ctx* ctx, double* yold, double* acold);
/* Current iterate (y, ac) from the host, e.g. to resume inside a step. */
int phg_state_upload_cur(phg_ctx* ctx, const double* y, const double* ac);
/* itrPredict + itrBC (itrdrv.f:398-399)        itrPC.f:56-121, itrbc.f
 * Any generalized-alpha almi, alfi, gami (itrPC.f:16-25) with the
 * same-velocity predictor ipred = 1.  PHG_EINVAL for ipred 2-4: the
 * reference's itrPredict passes undeclared (implicitly typed, unset)
 * scalars for iBC, BC, iper and ilwork to itrBC there (itrPC.f:88,101,
 * 114), so those predictors have no defined result to reproduce. */
int phg_predict(phg_ctx* ctx);
/* SolGMRs + itrCorrect + itrBC on the resident state  itrdrv.f:477-599 */
int phg_newton_iteration(phg_ctx* ctx, phg_stats* st);

/* itrUpdate + itrBC on (yold, acold): end of a time step
 *                                   itrPC.f:176-208, itrdrv.f:651-652 */
int phg_update(phg_ctx* ctx);

/* Matrix-free GMRES (Solver Type: GMRES Matrix Free, impl = 20; selected at
 * itrdrv.f:489-506 with lhs = 0, iprec from LHSupd).  PHG_FP64 and
 * PHG_FP64_MATRIX_FREE contexts,
 * lhs = 0, iDC = 0 (PHG_EINVAL otherwise); idiff 0, 1, 3; itau 0, 1 (BDiag
 * from e3bdg.f), 10, 11 (from e3bdg_nd.f).
 *
 * ElmMFG(y,ac,x,shp,shgl,iBC,BC,shpb,shglb,res,rmes,BDiag,iper,ilwork,rerr)
 *                                                     elmmfg.f:1
 * res as ElmGMRs forms it, rmes the modified residual (e3 with ires = 3),
 * BDiag (iprec != 0) the block diagonal of e3bdg.f (itau >= 10: of
 * e3bdg_nd.f, whose eight mis-indexed statements it reproduces); no lhsK is
 * formed. */
int phg_elmmfg(phg_ctx* ctx, const double* y, const double* ac, double* res,
               double* rmes, double* BDiag);
/* ItrRes(yp,yc,x,shp,shgl,iBC,BC,shpb,shglb,rmes,iper,ilwork,ac)
 *                                                     itrres.f:1
 * rmes (in/out, accumulated as the reference does) += the modified residual
 * of the variables yp (nshg,5, global order) at the state yc, interior
 * elements only, then bc3Res; iabres = 1 takes |.| per element node
 * (asires.f:90, the /precis/ flag itrFDI sets). */
int phg_itrres(phg_ctx* ctx, const double* yp, const double* yc,
               const double* ac, double* rmes, int iabres);
/* Au1MFG(ypre,y,ac,x,rmes,res,uBrg,BDiag,iBC,BC,iper,ilwork,shp,shgl,shpb,
 *        shglb)                                       au1mfg.f:1
 * u <- (L^-1 Rmod(itrBC(old2new(U^-1 (ypre + eGMRES u)))) - rmes) / eGMRES:
 * the one-sided finite-difference product of SolMFG's Krylov loop, at the
 * state y.  ypre (nshg,5) = U new2old(y) and u in the equation order, rmes
 * forward-reduced, BDiag LU-factored (as SolMFG holds them). */
int phg_au1mfg(phg_ctx* ctx, const double* ypre, const double* y,
               const double* ac, const double* rmes, const double* BDiag,
               double eGMRES, double* u);
/* SolMFG(y,ac,yold,acold,x,iBC,BC,res,BDiag,HBrg,eBrg,yBrg,Rcos,Rsin,iper,
 *        ilwork,shp,shgl,shpb,shglb,Dy,rerr)          solmfg.f:1
 * ElmMFG, then GMRES on finite-difference products of the modified residual
 * (Au1MFG, au1mfg.f; Au2MFG's two-sided difference for the restarts,
 * au2mfg.f).  eGMRES: the /itrpar/ finite-difference interval, in/out; with
 * fdi = 1 itrFDI recomputes it first (itrfdi.f; the reference does so when
 * iter = 1 and mod(istep, 20) = 0, solmfg.f:163).  Returns res
 * preconditioned, rmes forward-reduced, BDiag LU-factored, Dy the increment
 * consumed by itrCorrect; st->iKs is the last cycle's Krylov count. */
int phg_solmfg(phg_ctx* ctx, const double* y, const double* ac, double* res,
               double* rmes, double* BDiag, double* Dy, double* eGMRES, int fdi,
               phg_stats* st);
/* SolMFG + itrCorrect + itrBC on the resident state (itrdrv.f:489-599);
 * eGMRES as phg_solmfg (NULL: the context's interval). */
int phg_newton_iteration_mfg(phg_ctx* ctx, double* eGMRES, int fdi,
                             phg_stats* st);

/* Fine-grained entry points (unit parity; host arrays in/out). */
/* AsIq over the interior elements + qpbc, the global reconstruction of the
 * diffusive flux for idiff 1 or 3          elmgmr.f:361-391, asiq.f:1,
 * qpbc.f:1.  qres (nshg,12): the flux, direction-major (x1: momentum x,
 * y, z, energy; then x2, x3); rmass (nshg): the inverted lumped mass. */
int phg_asiq(phg_ctx* ctx, const double* y, double* qres, double* rmass);
/* SparseAp(iper,ilwork,iBC,col,row,lhsK,p)          sparseap.f:1     */
int phg_sparse_ap(phg_ctx* ctx, const double* lhsK, double* p);
/* p <- bc3per(A p(iper)) with A the FP64 assembled tangent of state y
 * (EGmass of every element after bc3LHS, elmgmr.f:441-459), applied element
 * by element without forming or storing lhsK: the unscaled operator of
 * SparseAp + bc3per.  The mixed-precision path forms its refinement
 * residual with it.  Matrix-free SolMFG (solmfg.f, au1mfg.f:56-78) instead
 * differences the residual at y + eps u; this is the exact product.  With
 * iDC != 0 the DC term uses the DC factors of the context's last assembly
 * (phg_elmgmrs / phg_solgmrs / phg_newton_iteration at the same state),
 * as the tangent that assembly formed does; PHG_EINVAL before any. */
int phg_tangent_ap(phg_ctx* ctx, const double* y, double* p);
/* Transpose of SparseAp's operator, p <- K^T p with K p = bc3per(A p(iper))
 * (no counterpart in the reference: the learned preconditioner's training
 * differentiates through the matvec). */
int phg_sparse_ap_t(phg_ctx* ctx, const double* lhsK, double* p);
/* i3LU(Diag,r,code), code 'LU_Fact ','forward ','backward','product '
 *                                                   i3lu.f:1         */
int phg_i3lu(phg_ctx* ctx, double* Diag, double* r, const char* code);
/* Spsi3pre(BDiag,lhsK,col,row)                      spsi3pre.f:1     */
int phg_spsi3pre(phg_ctx* ctx, const double* BDiag, double* lhsK);
/* phg_spsi3pre without allocation (no counterpart in the reference): the
 * same result as phg_spsi3pre, for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  lhsK is staged through a static 64 MiB device array
 * instead of a buffer allocated on first use, so nnz_tot is limited to
 * 335,544 blocks (25 doubles each); larger contexts get PHG_EINVAL.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (two
 * host threads, two processes' contexts are separate) race.  FP64
 * contexts only, as phg_spsi3pre. */
int phg_spsi3pre_noalloc(phg_ctx* ctx, const double* BDiag, double* lhsK);
/* bc3per(iBC,res,iper,ilwork,nQs)                   bc3per.f:1       */
int phg_bc3per(phg_ctx* ctx, double* res, int nQs);

/* Kernel timing hook for the benchmark: run `reps` SpMV launches on the
 * resident lhsK and return the mean launch time (HIP events on the
 * library stream). */
int phg_bench_spmv(phg_ctx* ctx, int reps, double* ms_per_launch,
                   double* bytes_per_launch);
int phg_bench_assembly(phg_ctx* ctx, int reps, double* ms_per_launch);
/* Plain device copy inside the Krylov space (no counterpart in the
 * reference): a kernel of known traffic that calibrates the memory
 * counters of a profiler run; bytes = read + written bytes per launch. */
int phg_bench_copy(phg_ctx* ctx, int reps, double* ms_per_launch, double* bytes);
/* ---- Learned preconditioner (no counterpart in the reference) ----
 * A 3-layer message-passing network over the periodic node graph (hidden
 * width 64, bf16 MFMA with float accumulation) that preconditions flexible
 * GMRES on the block-diagonally scaled system, with block-diagonal GMRES as
 * the fallback whenever its cycle does not reduce the true residual.
 * Parameters: phg_mpnn_param_count() floats, W0 (32x64), b0, W1 (320x64),
 * b1, W2, b2, W3, b3, Wout (64x5), bout, matrices row-major [in][out].    */
int phg_mpnn_param_count(void);
/* Builds the node graph from the same mesh arrays as phg_init; dscale: the
 * 25 divisors of the BDiag inputs. */
int phg_mpnn_init(phg_ctx* ctx, const phg_mesh* mesh, const float* dscale);
int phg_mpnn_set_weights(phg_ctx* ctx, const float* w, int count);
int phg_mpnn_get_weights(phg_ctx* ctx, float* w, int count);
/* use_pc > 0: flexible GMRES inside phg_solgmrs/phg_newton_iteration, the
 * network preconditioning the first use_pc directions of each cycle (the
 * block-diagonal preconditioner the others); 0: block-diagonal GMRES.
 * keep_pair: keep (BDiag before factorisation, b, z) of each solve. */
int phg_mpnn_enable(phg_ctx* ctx, int use_pc, int keep_pair);
/* z = network(v) for host arrays (bdiag_raw (nshg,5,5), v, z (nshg,5)). */
int phg_mpnn_apply(phg_ctx* ctx, const double* bdiag_raw, const double* v, double* z);
/* One training step on a pair (v, t): forward, loss, backward; an Adam
 * step when lr > 0.  grad (nullable) receives the gradient. */
int phg_mpnn_train_step(phg_ctx* ctx, const double* bdiag_raw, const double* v,
                        const double* t, float lr, double* loss, float* grad);
/* Pair of the last solve: BDiag before factorisation, the scaled right-hand
 * side b = L^-1 res and the scaled solution z (Dy = U^-1 z). */
int phg_mpnn_last_pair(phg_ctx* ctx, double* bdiag_raw, double* b, double* zhat);

/* Training over several meshes: dst uses src's parameters and optimiser
 * state from now on (same device). */
int phg_mpnn_share(phg_ctx* dst, phg_ctx* src);
/* Device-resident training set of `capacity` pairs (35 floats per node). */
int phg_mpnn_dataset(phg_ctx* ctx, int capacity);
/* Stores the pair of the last solve (keep_pair enabled) in a slot. */
int phg_mpnn_store_pair(phg_ctx* ctx, int slot);
/* One training step (Adam when lr > 0) on a stored pair; loss nullable. */
int phg_mpnn_train_pair(phg_ctx* ctx, int slot, float lr, double* loss);

/* Training on the residual of preconditioned directions: one Adam step per
 * Krylov vector v_1..v_nvec of the last solve (keep_pair enabled), loss
 * ||K z(v) - v||^2 / ||v||^2 with K = bc3per o SparseAp the scaled operator
 * of that solve (gradient through K^T, as phg_sparse_ap_t).  mean_loss and
 * grad (the gradient of the last vector's loss, phg_mpnn_param_count()
 * floats) are nullable.  The vectors are the context's Krylov basis as it
 * stands: host-array entry points that run in between (phg_sparse_ap,
 * phg_sparse_ap_t, phg_tangent_ap) stage their arguments in its
 * first vectors and the matrix they are given replaces K. */
int phg_mpnn_train_krylov(phg_ctx* ctx, int nvec, float lr, double* mean_loss, float* grad);

/* Diagnostics (no counterpart in the reference): PHG_DIAG_ORTH measures
 * the orthogonality of every GMRES cycle's Krylov basis into
 * phg_stats.orth_loss, at the cost of (iKs+1)(iKs+2)/2 dot products per
 * cycle; 0 turns diagnostics off. */
enum { PHG_DIAG_ORTH = 1 };
int phg_diagnostics(phg_ctx* ctx, int flags);

/* Wait for all work queued on the library stream. */
int phg_sync(phg_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
