"""Frozen flag set and synthetic case definitions for the hot path.

Every value that the reference reads from ``solver.inp``/``input.config`` on
this path is pinned here once (reference defaults unless noted):

* calorically perfect gas, ``ipress = 0`` (getthm.f:106-172); gamma = 1.4;
  Rgas = 1/(0.79/Rs(N2) + 0.21/Rs(O2)) with Rs = Rh/M (input.f:165-182,
  common.f:100-112): 288.30... J/(kg K)
* constant viscosity mu = 1.8e-5 Pa s (matflg(2)=0, getdiff.f:126-128),
  Stokes bulk viscosity (matflg(3)=0), Prandtl 0.72 (input.config:113)
* Navier = 1, ipord = 1 linear tets, 4-point interior / 3-point wall rule
* itau = 0 (Shakib diagonal tau), dtsfct = taucfct = temper = 1
  (input.config:233-266), iDC = 0, idiff = 0 ("Include Viscous Correction
  in Stabilization: False"), iLHScond = 0
* backward Euler ("First Order": almi = alfi = gami = 1, ipred = 1,
  itrPC.f:16-20) on the workloads; generalized alpha (rho_inf in [0, 1])
  and the four predictors through time_integration; LHSupd = 1 so
  lhs = iprec = 1 each Newton iteration
* nGMRES = 1, Kspace 10 (16^3 case) or 20, epstol 1e-3, minIters 0
  (reference default 10; set explicitly and identically in every arm)
* wall thermal condition: floating flux (iBCB = 0, e3b.f:149-228)
* time step: dt = CFL * h_min / (c0 (1 + M)), h_min the smallest grid-line
  spacing: 1/N on the uniform boxes with CFL = 1 (10-11 block-diagonal
  GMRES iterations per solve on the 16^3 box, the range of the reference's
  legacy compressible case); on the y-graded boxes h_min is the first wall
  cell and CFL = 3 (128^3, ratio 1.1: h_min = 1.12e-4, 10-19 Krylov
  iterations per Newton iteration; with h = 1/N the wall-normal acoustic
  CFL is about 50, block-diagonal GMRES(20) stalls and the state breaks
  down in the second Newton iteration)
"""
import math

RH = 8.31441000000000e+0          # common.f:104
MSH = (2.8e-2, 3.2e-2)            # common.f:108 (N2, O2)
XN2, XO2 = 0.79, 0.21             # common.f:106-107
RS = (RH / MSH[0], RH / MSH[1])
RGAS = 1.0 / (XN2 / RS[0] + XO2 / RS[1])   # input.f:182
GAMMA = 1.4
GAMMA1 = 0.4
MU = 1.8e-5
PRANDTL = 0.72
T0 = 300.0
C0 = math.sqrt(GAMMA * RGAS * T0)
EPSM = math.sqrt(2.220446049250313e-16)   # input.f:30 sqrt(epsilon(one))
CFL = 1.0
IBKSIZ = 64

MESH_SEED = 20261202
FIELD_SEED = 7

# Synthetic workloads (unit box, Kuhn 6-tet split, periodic x/z, walls y).
CASES = {
    "box16": dict(N=16, mach=0.3, jitter=0.1, grade=1.0, kspace=10,
                  newton=1, steps=1, precision="fp64", cfl=1.0),
    "box64": dict(N=64, mach=0.5, jitter=0.1, grade=1.0, kspace=20,
                  newton=3, steps=1, precision="fp64", cfl=1.0),
    "box128": dict(N=128, mach=0.3, jitter=0.1, grade=1.1, kspace=20,
                   newton=3, steps=1, precision="mixed", cfl=3.0),
    # boundary layer: first wall cell h/100 (aspect ratio 100), growing by
    # 1.08 per cell up to the uniform core (67 graded cells per wall, core
    # cells 1.75 h tall)
    "box256": dict(N=256, mach=0.6, jitter=0.1, grade=1.08, first=0.01, kspace=20,
                   newton=3, steps=5, precision="mixed", cfl=3.0),
}


def min_spacing(N, grade=1.0, first=0.0):
    """Smallest grid-line spacing of the box: 1/N; with geometric
    y-grading of ratio `grade` from both walls the first wall cell
    h0 = (1/2) (grade - 1) / (grade^(N/2) - 1) (the mesh generator's ygrid);
    with a capped boundary layer (first > 0) the first cell first/N."""
    if grade == 1.0:
        return 1.0 / N
    if first > 0.0:
        return first / N
    return 0.5 * (grade - 1.0) / (grade ** (N // 2) - 1.0)


def time_step(N, mach, grade=1.0, cfl=CFL, first=0.0):
    return cfl * min_spacing(N, grade, first) / (C0 * (1.0 + mach))


# "Predictor at Start of Step" values (input_fform.cc:636-643)
PREDICTORS = {"Same Velocity": 1, "Zero Acceleration": 2, "Same Acceleration": 3,
              "Same Delta": 4}


def time_integration(rhoinf=-1.0, ipred=1):
    """itrSetup (itrPC.f:6-37): the generalized-alpha factors of the
    Hulbert family for a spectral radius rho_inf in [0, 1] (second order),
    backward Euler with the same-velocity predictor otherwise ("First
    Order" sets rho_inf = -1, input_fform.cc:633-635); the same-delta
    predictor falls back to same acceleration at rho_inf = 0.  The
    start-up shift y -= (1 - alfi) Delt CFLfl acold of ipred = 4
    (itrPC.f:31-32) is zero in the reference, which never assigns CFLfl,
    so the first step starts from y = yold (itrdrv.f:148)."""
    if rhoinf < 0.0 or rhoinf > 1.0:
        return dict(almi=1.0, alfi=1.0, gami=1.0, ipred=1)
    almi = (3.0 - rhoinf) / (1.0 + rhoinf) / 2.0
    alfi = 1.0 / (1.0 + rhoinf)
    gami = 0.5 + almi - alfi
    if ipred == 4 and rhoinf == 0.0:
        ipred = 3
    return dict(almi=almi, alfi=alfi, gami=gami, ipred=ipred)


def solver_params(N, mach, kspace, etol=1e-3, min_iters=0, grade=1.0, cfl=CFL,
                  rhoinf=-1.0, ipred=1, idiff=0, first=0.0, idc=0,
                  itau=0):
    """The COMMON-block scalars the path reads, as a plain dict.  idiff:
    0, 1 (what "Include Viscous Correction in Stabilization: True" sets for
    linear elements, input_fform.cc:706-711) or 3: global reconstruction of
    the diffusive flux with row-sum or special lumping (elmgmr.f:361-391,
    e3q.f:185-205).  idc: discontinuity capturing, 0 off, 1 DC-mallet,
    2 DC-quadratic, 3 DC-minimum (input_fform.cc:683-686, e3dc.f).  itau:
    0 Shakib's diagonal tau, 10 the matrix tau of e3tau_nd ("Tau Matrix:
    Matrix-Mallet", input_fform.cc)."""
    dt = time_step(N, mach, grade, cfl, first)
    return dict(Rgas=RGAS, gamma=GAMMA, gamma1=GAMMA1, pr=PRANDTL, mu=MU,
                dtsfct=1.0, taucfct=1.0, temper=1.0, Dtgl=1.0 / dt,
                etol=etol, epsM=EPSM, Kspace=kspace, nGMRES=1,
                minIters=min_iters, lhs=1, iprec=1, idiff=idiff, idc=idc, itau=itau,
                **time_integration(rhoinf, ipred))
