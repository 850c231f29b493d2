/*
 * pic_oracle.c -- CPU oracle for the PIC particle stage (test infrastructure).
 *
 * Plain C99, FP64, compiled with -ffp-contract=off so that every product and
 * sum is rounded exactly as written.  Each routine restates one routine of
 * the reference (Fortran, PICLas v4.1.0) and cites it as file:line, paths
 * relative to the reference's src/ directory.  The tests compare the HIP path
 * against it, and it is the CPU baseline timed beside it; the product never
 * calls it.
 *
 * Scope: 3D hexahedral meshes without mortars; TriaTracking (periodic and
 * open particle boundaries); cell_volweight_mean and 3D shape_function
 * deposition; the Boris-Leapfrog push of the 508 time stepper.
 */
#include "pic_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 16

/* ---------------------------------------------------------------- errors */
static char g_err[512];
static void set_err(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
const char *om_last_error(void) { return g_err; }

/* --------------------------------------------------------------- constants
 * globals_vars.f90:48-90 */
static const double PI_ = 3.14159265358979323846264338327950288;
static const double EPS_MACH = DBL_EPSILON; /* EPSILON(0.0) with 8-byte reals */
static double c2_inv(void) {
  const double c = 299792458.0;
  const double c2 = c * c;
  return 1.0 / c2;
}

/* ================================================================ basis.f90 */
/* ALMOSTEQUAL_UNITY, basis.f90:1011-1022 */
static int almost_equal_unity(double x, double y) {
  if (x == 0.0 || y == 0.0) return fabs(x - y) <= 2.0 * DBL_EPSILON;
  return (fabs(x - y) <= DBL_EPSILON * fabs(x)) && (fabs(x - y) <= DBL_EPSILON * fabs(y));
}

/* LagrangeInterpolationPolys, basis.f90:1223-1264 (barycentric form 2) */
void om_lagrange_polys(double x, int32_t N, const double *xGP, const double *wb, double *L) {
  int eq = 0;
  for (int i = 0; i <= N; i++) {
    L[i] = 0.0;
    if (almost_equal_unity(x, xGP[i])) { L[i] = 1.0; eq = 1; }
  }
  if (eq) return;
  double sum = 0.0;
  for (int i = 0; i <= N; i++) { L[i] = wb[i] / (x - xGP[i]); sum = sum + L[i]; }
  for (int i = 0; i <= N; i++) L[i] = L[i] / sum;
}

/* LegendrePolynomialAndDerivative, basis.f90:643-688 (normalised) */
static void legendre_poly_and_deriv(int N, double x, double *L, double *Lder) {
  double Lv = 0.0, Ld = 0.0;
  if (N == 0) { Lv = 1.0; Ld = 0.0; }
  else if (N == 1) { Lv = x; Ld = 1.0; }
  else {
    double Lm2 = 1.0, Lm1 = x, Ldm2 = 0.0, Ldm1 = 1.0;
    for (int i = 2; i <= N; i++) {
      Lv = ((double)(2 * i - 1) * x * Lm1 - (double)(i - 1) * Lm2) / (double)i;
      Ld = Ldm2 + (double)(2 * i - 1) * Lm1;
      Lm2 = Lm1; Lm1 = Lv; Ldm2 = Ldm1; Ldm1 = Ld;
    }
  }
  *L = Lv * sqrt((double)N + 0.5);
  *Lder = Ld * sqrt((double)N + 0.5);
}

/* LegendreGaussNodesAndWeights, basis.f90:758-835 */
void om_legendre_gauss_nodes(int32_t N, double *x, double *w) {
  double L, Ld;
  if (N == 0) { x[0] = 0.0; if (w) w[0] = 2.0; return; }
  if (N == 1) {
    x[0] = -sqrt(1.0 / 3.0); x[1] = -x[0];
    if (w) { w[0] = 1.0; w[1] = 1.0; }
    return;
  }
  const double tol = 1.e-15;
  const int nIter = 10;
  const double cheb = 2.0 * atan(1.0) / (double)(N + 1);
  for (int i = 0; i <= (N + 1) / 2 - 1; i++) {
    x[i] = -cos(cheb * (double)(2 * i + 1));
    for (int it = 0; it <= nIter; it++) {
      legendre_poly_and_deriv(N + 1, x[i], &L, &Ld);
      double dx = -L / Ld;
      x[i] = x[i] + dx;
      if (fabs(dx) < tol * fabs(x[i])) break;
    }
    legendre_poly_and_deriv(N + 1, x[i], &L, &Ld);
    x[N - i] = -x[i];
    if (w) {
      w[i] = (2.0 * N + 3) / ((1.0 - x[i] * x[i]) * Ld * Ld);
      w[N - i] = w[i];
    }
  }
  if (N % 2 == 0) {
    x[N / 2] = 0.0;
    legendre_poly_and_deriv(N + 1, x[N / 2], &L, &Ld);
    if (w) w[N / 2] = (2.0 * N + 3) / (Ld * Ld);
  }
}

/* ChebyGaussLobNodesAndWeights, basis.f90:719-756 */
void om_cheby_gauss_lobatto_nodes(int32_t N, double *x) {
  if (N == 0) { x[0] = 0.0; return; }
  for (int i = 0; i <= N; i++) x[i] = -cos((double)i / (double)N * acos(-1.0));
}

/* BarycentricWeights, basis.f90:952-976 */
void om_barycentric_weights(int32_t N, const double *x, double *wb) {
  for (int i = 0; i <= N; i++) wb[i] = 1.0;
  for (int i = 1; i <= N; i++)
    for (int j = 0; j <= i - 1; j++) {
      wb[j] = wb[j] * (x[j] - x[i]);
      wb[i] = wb[i] * (x[i] - x[j]);
    }
  for (int i = 0; i <= N; i++) wb[i] = 1.0 / wb[i];
}

/* PolynomialDerivativeMatrix, basis.f90 (D(iGP,iLagrange)) */
static void derivative_matrix(int N, const double *x, double *D /* [N+1][N+1], D[i*(N+1)+l] */) {
  double wb[MAXN];
  om_barycentric_weights(N, x, wb);
  for (int i = 0; i < (N + 1) * (N + 1); i++) D[i] = 0.0;
  for (int l = 0; l <= N; l++)
    for (int i = 0; i <= N; i++)
      if (l != i) {
        D[i * (N + 1) + l] = wb[l] / (wb[i] * (x[i] - x[l]));
        D[i * (N + 1) + i] = D[i * (N + 1) + i] - D[i * (N + 1) + l];
      }
}

/* ============================================================ state */
struct om_state {
  int nElems, nSides, nNodes, NGeo, N, nDof, nGP;  /* nGP = (NGeo+1)^3 */
  int32_t *ElemInfo, *SideInfo;
  double *NodeCoords;
  int32_t *NodeInfo; /* [nNodes] 0-based unique id */
  int nUnique;
  int nBCs;
  int32_t *bcAlpha, *partBC;
  int nPV;
  double PV[3][3];
  /* derived side info (particle_mesh_readin.f90:385-418) */
  int32_t *sideElem, *sideLocal, *sideNbSide;
  /* geometry */
  int32_t *ElemSideNodeID; /* [e][6][4] 0-based node index */
  uint8_t *Concave;        /* [e][6] */
  int32_t *ElemNodeID;     /* [e][8] 0-based node index, CGNS corners */
  double *XCL;             /* [e][nGP][3] */
  double *dXCL;            /* [e][nGP][3 (x)][3 (d)] = dX_x/dxi_d */
  double *bary, *radius;   /* [e][3], [e]: BuildElementRadiusTria */
  double *xiBasis, *slen;  /* [e][6][3], [e][6] */
  double *xGPe, *sJ;       /* [e][nDof][3], [e][nDof] */
  /* reference mapping (RefMapping tracking, NGeo > 1) */
  uint8_t *curved;         /* [e] ElemCurved                                  */
  double *dXCL1;           /* [e][8][9] dXCL_NGeo at the 8 corner CL points   */
  double *epsOneCell;      /* [e] ElemEpsOneCell                              */
  uint8_t *openElem;       /* [e] element has an open particle boundary side  */
  /* boundary tracking with RefMapping (pic_oracle_bc.h) */
  int cartPeriodic;        /* every boundary periodic: PeriodicMovement (CartesianPeriodic = T) */
  int nBC;                 /* boundary sides (HOPR side rows with a BC)        */
  int32_t *bcSide, *sideToBC;
  double *bz;              /* [nBC][NGeo+1][NGeo+1][3] BezierControlPoints3D   */
  uint8_t *sideType, *bbEmpty;
  double *sideNorm, *sideDist, *slabN, *slabI, *baseV, *bcOrigin, *bcRadius;
  int32_t *e2bOff, *e2bSide; /* ElemToBCSides (CSR) and SideBCMetrics side    */
  double *e2bDist;           /* SideBCMetrics(BCSIDE_DISTANCE)                */
  double bcHalo;
  double gmin[3], gmax[3]; /* GEO%x/y/zminglob, maxglob                       */
  /* self-consistent field through a fixed linear operator (adjoint workload) */
  double *Gop;             /* [3*nUnique][nElems]: unique-node E from element-mean rho */
  double *rhoBar, *Enode;  /* [nElems], [nUnique][3]                           */
  int dirPV[3];            /* GEO%DirPeriodicVectors (0-based axis)           */
  int32_t *perOff, *perNodes; /* twins of each unique node (CSR) */
  double *NodeVolume;      /* [nUnique] */
  /* 1D bases */
  double xiCL[MAXN], wBaryCL[MAXN];
  double xGP[MAXN], wGP[MAXN], wBaryGP[MAXN];
  /* background bins (FIBGM stand-in built from element bounding boxes) */
  double bmin[3], bmax[3], bdelta[3];
  int nb[3];
  int32_t *binOff, *binElems;
  /* params */
  om_params par;
  double r_sf, r2_sf, r2_sf_inv, w_sf;
  int nCases;
  int caseMat[27][3];
  /* fields */
  double *E, *PS, *NodeSource;
  /* particles */
  int n;
  double *PartState, *LastPartPos, *Field, *PartPosRef;
  int32_t *elem, *lastElem, *species; /* 0-based elem; -1 unknown */
  uint8_t *inside, *isNew;
  int64_t lost;
  int64_t multiCross; /* walk iterations that took the multi-crossing branch */
  double *tbLog;      /* tie-break events (om_set_tiebreak_log), OM_TB_REC doubles each */
  int64_t tbCap, tbN;
  int plant;          /* planted mistakes (om_plant_mistake), 0 normally */
  int nthreads;
  /* threaded shape-function deposit: per-thread PartSource copies, kept
   * between calls, and the wall time spent zeroing and summing them */
  double **priv;
  int nPriv;
  double tPrivReduce;
};

#define ELEMINFO(e, c) (s->ElemInfo[(size_t)(e) * 6 + (c)])
#define SIDEINFO(sd, c) (s->SideInfo[(size_t)(sd) * 5 + (c)])
/* HOPR column indices, src/piclas.h:146-167 (0-based here) */
enum { ELEM_TYPE = 0, ELEM_ZONE, ELEM_FIRSTSIDEIND, ELEM_LASTSIDEIND, ELEM_FIRSTNODEIND, ELEM_LASTNODEIND };
enum { SIDE_TYPE = 0, SIDE_ID, SIDE_NBELEMID, SIDE_FLIP, SIDE_BCID };

static void *xcalloc(size_t n, size_t sz) {
  void *p = calloc(n ? n : 1, sz);
  if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  return p;
}

/* GetCornerNodeMapCGNS, mesh_tools.f90:740-769 (1-based node numbers within an element) */
static void corner_node_map(int NGeo, int CNS[8], int NodeMap[6][4]) {
  const int n1 = NGeo + 1;
  CNS[0] = 1;
  CNS[1] = n1;
  CNS[2] = n1 * n1;
  CNS[3] = n1 * NGeo + 1;
  CNS[4] = n1 * n1 * NGeo + 1;
  CNS[5] = n1 * n1 * NGeo + n1;
  CNS[6] = n1 * n1 * n1;
  CNS[7] = n1 * n1 * NGeo + n1 * NGeo + 1;
  const int m[6][4] = {{1, 4, 3, 2}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 4, 8, 7}, {1, 5, 8, 4}, {5, 6, 7, 8}};
  for (int sd = 0; sd < 6; sd++)
    for (int k = 0; k < 4; k++) NodeMap[sd][k] = CNS[m[sd][k] - 1];
}

/* ChangeBasis3D, changeBasis.f90:54-111, for Dim1 = 3 */
static void change_basis_3d(int Nin, int Nout, const double *Vdm /* [Nout+1][Nin+1] */,
                            const double *in /* [k][j][i][3] */, double *out) {
  const int a = Nin + 1, b = Nout + 1;
  double *buf1 = xcalloc((size_t)b * a * a * 3, sizeof(double));
  double *buf2 = xcalloc((size_t)b * b * a * 3, sizeof(double));
  for (int k = 0; k < a; k++) for (int j = 0; j < a; j++) for (int i = 0; i < a; i++)
    for (int io = 0; io < b; io++) for (int c = 0; c < 3; c++)
      buf1[((k * a + j) * b + io) * 3 + c] += Vdm[io * a + i] * in[((k * a + j) * a + i) * 3 + c];
  for (int k = 0; k < a; k++) for (int j = 0; j < a; j++) for (int jo = 0; jo < b; jo++)
    for (int io = 0; io < b; io++) for (int c = 0; c < 3; c++)
      buf2[((k * b + jo) * b + io) * 3 + c] += Vdm[jo * a + j] * buf1[((k * a + j) * b + io) * 3 + c];
  for (int i = 0; i < b * b * b * 3; i++) out[i] = 0.0;
  for (int k = 0; k < a; k++) for (int ko = 0; ko < b; ko++) for (int jo = 0; jo < b; jo++)
    for (int io = 0; io < b; io++) for (int c = 0; c < 3; c++)
      out[((ko * b + jo) * b + io) * 3 + c] += Vdm[ko * a + k] * buf2[((k * b + jo) * b + io) * 3 + c];
  free(buf1);
  free(buf2);
}

/* InitializeVandermonde, basis.f90:615-639: Vdm(iOut,:) = L(xOut(iOut)) */
static void vandermonde(int Nin, const double *xin, int Nout, const double *xout, double *Vdm) {
  double wb[MAXN];
  om_barycentric_weights(Nin, xin, wb);
  for (int io = 0; io <= Nout; io++) om_lagrange_polys(xout[io], Nin, xin, wb, Vdm + io * (Nin + 1));
}

/* ------------------------------------------------------------ mesh init */
#include "pic_oracle_bc.h"

static int build_mesh(om_state *s, const om_mesh_desc *m) {
  s->nElems = m->nElems; s->nSides = m->nSides; s->nNodes = m->nNodes;
  s->NGeo = m->NGeo; s->N = m->N;
  s->nDof = (m->N + 1) * (m->N + 1) * (m->N + 1);
  s->nGP = (m->NGeo + 1) * (m->NGeo + 1) * (m->NGeo + 1);
  if (m->N < 1 || m->N > 12 || m->NGeo < 1 || m->NGeo > 8) { set_err("unsupported N/NGeo"); return -1; }
  const int ne = s->nElems;
  s->ElemInfo = xcalloc((size_t)ne * 6, sizeof(int32_t));
  memcpy(s->ElemInfo, m->ElemInfo, (size_t)ne * 6 * sizeof(int32_t));
  s->SideInfo = xcalloc((size_t)s->nSides * 5, sizeof(int32_t));
  memcpy(s->SideInfo, m->SideInfo, (size_t)s->nSides * 5 * sizeof(int32_t));
  s->NodeCoords = xcalloc((size_t)s->nNodes * 3, sizeof(double));
  memcpy(s->NodeCoords, m->NodeCoords, (size_t)s->nNodes * 3 * sizeof(double));
  s->nBCs = m->nBCs;
  s->bcAlpha = xcalloc(m->nBCs, sizeof(int32_t));
  s->partBC = xcalloc(m->nBCs, sizeof(int32_t));
  for (int b = 0; b < m->nBCs; b++) {
    s->bcAlpha[b] = m->BCType[b * 4 + 3];    /* BoundaryType(:,BC_ALPHA) */
    s->partBC[b] = m->PartBoundCond[b];
  }
  s->nPV = m->nPeriodicVectors;
  for (int v = 0; v < s->nPV && v < 3; v++)
    for (int c = 0; c < 3; c++) s->PV[v][c] = m->PeriodicVectors[v * 3 + c];

  /* unique node ids (NodeInfo_Shared) */
  s->NodeInfo = xcalloc(s->nNodes, sizeof(int32_t));
  int maxid = 0;
  for (int i = 0; i < s->nNodes; i++) {
    s->NodeInfo[i] = m->GlobalNodeIDs[i] - 1;
    if (m->GlobalNodeIDs[i] > maxid) maxid = m->GlobalNodeIDs[i];
  }
  s->nUnique = maxid;

  /* SIDE_ELEMID / SIDE_LOCALID / SIDE_NBSIDEID, particle_mesh_readin.f90:385-418 */
  s->sideElem = xcalloc(s->nSides, sizeof(int32_t));
  s->sideLocal = xcalloc(s->nSides, sizeof(int32_t));
  s->sideNbSide = xcalloc(s->nSides, sizeof(int32_t));
  for (int e = 0; e < ne; e++) {
    int first = ELEMINFO(e, ELEM_FIRSTSIDEIND), last = ELEMINFO(e, ELEM_LASTSIDEIND);
    if (last - first != 6) { set_err("element %d has %d sides (mortars unsupported)", e + 1, last - first); return -1; }
    int cnt = 0;
    for (int l = 1; l <= last - first; l++) {
      int sd = first + l - 1;
      s->sideElem[sd] = e;
      if (SIDEINFO(sd, SIDE_TYPE) <= 100) s->sideLocal[sd] = ++cnt;
      else { set_err("mortar side unsupported"); return -1; }
      int nb = SIDEINFO(sd, SIDE_NBELEMID);
      if (nb == 0) s->sideNbSide[sd] = -1;
      else if (nb < 0) { set_err("mortar neighbour unsupported"); return -1; }
      else {
        int nf = ELEMINFO(nb - 1, ELEM_FIRSTSIDEIND), nl = ELEMINFO(nb - 1, ELEM_LASTSIDEIND);
        s->sideNbSide[sd] = -1;
        for (int jl = 1; jl <= nl - nf; jl++) {
          int nsd = nf + jl - 1;
          if (abs(SIDEINFO(sd, SIDE_ID)) == abs(SIDEINFO(nsd, SIDE_ID))) { s->sideNbSide[sd] = nsd; break; }
        }
      }
    }
  }

  /* ElemSideNodeID and ConcaveElemSide, particle_mesh_tools.f90:1800-1930 */
  int CNS[8], NodeMap[6][4];
  corner_node_map(s->NGeo, CNS, NodeMap);
  s->ElemSideNodeID = xcalloc((size_t)ne * 24, sizeof(int32_t));
  s->Concave = xcalloc((size_t)ne * 6, sizeof(uint8_t));
  s->ElemNodeID = xcalloc((size_t)ne * 8, sizeof(int32_t));
  for (int e = 0; e < ne; e++) {
    int first = ELEMINFO(e, ELEM_FIRSTSIDEIND), fn = ELEMINFO(e, ELEM_FIRSTNODEIND);
    for (int c = 0; c < 8; c++) s->ElemNodeID[e * 8 + c] = fn + CNS[c] - 1; /* InitElemNodeIDs, mesh_tools.f90:634-669 */
    for (int l = 1; l <= 6; l++) {
      int sd = first + l - 1, ls = s->sideLocal[sd];
      int nStart = SIDEINFO(sd, SIDE_ID) > 0 ? 0 : (SIDEINFO(sd, SIDE_FLIP) % 10 - 1 > 0 ? SIDEINFO(sd, SIDE_FLIP) % 10 - 1 : 0);
      for (int k = 0; k < 4; k++)
        s->ElemSideNodeID[(e * 6 + ls - 1) * 4 + k] = fn + NodeMap[ls - 1][(nStart + k) % 4] - 1;
    }
    for (int l = 1; l <= 6; l++) {
      int sd = first + l - 1, ls = s->sideLocal[sd];
      double A[3][3]; /* A(:,NodeNum) */
      const int32_t *nid = &s->ElemSideNodeID[(e * 6 + ls - 1) * 4];
      for (int nn = 0; nn < 3; nn++)
        for (int c = 0; c < 3; c++) A[c][nn] = s->NodeCoords[nid[nn] * 3 + c] - s->NodeCoords[nid[3] * 3 + c];
      double detcon = ((A[1][0] * A[2][1] - A[2][0] * A[1][1]) * A[0][2] +
                       (A[2][0] * A[0][1] - A[0][0] * A[2][1]) * A[1][2] +
                       (A[0][0] * A[1][1] - A[1][0] * A[0][1]) * A[2][2]);
      if (detcon < 0) s->Concave[e * 6 + ls - 1] = 1;
      else if (detcon == 0.0) {
        /* GlobalSideID.LT.SideInfo_Shared(SIDE_NBSIDEID,...) with 1-based ids */
        int nbs = s->sideNbSide[sd];
        if (sd + 1 < (nbs >= 0 ? nbs + 1 : 0)) s->Concave[e * 6 + ls - 1] = 1;
      }
    }
  }

  /* 1D bases: CL nodes of NGeo (metrics.f90:288-290), Gauss nodes of N */
  om_cheby_gauss_lobatto_nodes(s->NGeo, s->xiCL);
  om_barycentric_weights(s->NGeo, s->xiCL, s->wBaryCL);
  om_legendre_gauss_nodes(s->N, s->xGP, s->wGP);
  om_barycentric_weights(s->N, s->xGP, s->wBaryGP);

  /* XCL_NGeo and dXCL_NGeo, metrics.f90:293-338 */
  const int ng = s->NGeo, g1 = ng + 1, nGP = s->nGP;
  double xEQ[MAXN], Vdm[MAXN * MAXN], D[MAXN * MAXN];
  for (int i = 0; i <= ng; i++) xEQ[i] = 2.0 * (double)i / (double)ng - 1.0;
  vandermonde(ng, xEQ, ng, s->xiCL, Vdm);
  derivative_matrix(ng, s->xiCL, D);
  s->XCL = xcalloc((size_t)ne * nGP * 3, sizeof(double));
  s->dXCL = xcalloc((size_t)ne * nGP * 9, sizeof(double));
  double *nc = xcalloc((size_t)nGP * 3, sizeof(double));
  for (int e = 0; e < ne; e++) {
    int fn = ELEMINFO(e, ELEM_FIRSTNODEIND);
    if (ELEMINFO(e, ELEM_LASTNODEIND) - fn != nGP) { set_err("element %d: node count != (NGeo+1)^3", e + 1); free(nc); return -1; }
    memcpy(nc, &s->NodeCoords[(size_t)fn * 3], (size_t)nGP * 3 * sizeof(double));
    double *X = &s->XCL[(size_t)e * nGP * 3];
    change_basis_3d(ng, ng, Vdm, nc, X);
    double *dX = &s->dXCL[(size_t)e * nGP * 9];
    for (int k = 0; k < g1; k++) for (int j = 0; j < g1; j++) for (int i = 0; i < g1; i++) {
      int p = (k * g1 + j) * g1 + i;
      for (int q = 0; q < 9; q++) dX[p * 9 + q] = 0.0;
      for (int ll = 0; ll < g1; ll++)
        for (int x = 0; x < 3; x++) {
          dX[p * 9 + x * 3 + 0] += D[i * g1 + ll] * X[((k * g1 + j) * g1 + ll) * 3 + x];
          dX[p * 9 + x * 3 + 1] += D[j * g1 + ll] * X[((k * g1 + ll) * g1 + i) * 3 + x];
          dX[p * 9 + x * 3 + 2] += D[k * g1 + ll] * X[((ll * g1 + j) * g1 + i) * 3 + x];
        }
    }
  }
  free(nc);

  /* ElemBaryNGeo / ElemRadius(2)NGeo for TriaTracking, particle_mesh_build.f90:217-309 */
  s->bary = xcalloc((size_t)ne * 3, sizeof(double));
  s->radius = xcalloc(ne, sizeof(double));
  for (int e = 0; e < ne; e++) {
    int fn = ELEMINFO(e, ELEM_FIRSTNODEIND);
    double xp[3] = {0, 0, 0};
    for (int nn = 1; nn <= 8; nn++)
      for (int c = 0; c < 3; c++) xp[c] = xp[c] + s->NodeCoords[(size_t)(fn + nn - 1) * 3 + c];
    for (int c = 0; c < 3; c++) s->bary[e * 3 + c] = xp[c] / 8.0;
    double r = 0.0;
    for (int nn = 1; nn <= 8; nn++) {
      double d[3];
      for (int c = 0; c < 3; c++) d[c] = s->NodeCoords[(size_t)(fn + nn - 1) * 3 + c] - s->bary[e * 3 + c];
      double v = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (v > r) r = v;
    }
    s->radius[e] = r;
  }

  /* XiEtaZetaBasis, slenXiEtaZetaBasis, particle_mesh_build.f90:386-444 */
  s->xiBasis = xcalloc((size_t)ne * 18, sizeof(double));
  s->slen = xcalloc((size_t)ne * 6, sizeof(double));
  {
    const double Xi[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
    for (int e = 0; e < ne; e++) {
      const double *X = &s->XCL[(size_t)e * nGP * 3];
      for (int d = 0; d < 6; d++) {
        double L1[MAXN], L2[MAXN], L3[MAXN];
        om_lagrange_polys(Xi[d][0], ng, s->xiCL, s->wBaryCL, L1);
        om_lagrange_polys(Xi[d][1], ng, s->xiCL, s->wBaryCL, L2);
        om_lagrange_polys(Xi[d][2], ng, s->xiCL, s->wBaryCL, L3);
        double xp[3] = {0, 0, 0};
        for (int k = 0; k < g1; k++) for (int j = 0; j < g1; j++) for (int i = 0; i < g1; i++)
          for (int c = 0; c < 3; c++) xp[c] = xp[c] + X[((k * g1 + j) * g1 + i) * 3 + c] * L1[i] * L2[j] * L3[k];
        double *b = &s->xiBasis[(e * 6 + d) * 3];
        for (int c = 0; c < 3; c++) b[c] = xp[c] - s->bary[e * 3 + c];
        s->slen[e * 6 + d] = 1.0 / (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      }
    }
  }

  /* Elem_xGP and sJ at the Gauss points of N.  The reference evaluates them in
   * CalcMetrics (metrics.f90:293-356) by change of basis XCL_NGeo -> Gauss
   * and detJ on NGeoRef -> Gauss; for the polynomial maps here that is the
   * mapping and its Jacobian determinant evaluated at the Gauss points. */
  const int n1 = s->N + 1, nDof = s->nDof;
  s->xGPe = xcalloc((size_t)ne * nDof * 3, sizeof(double));
  s->sJ = xcalloc((size_t)ne * nDof, sizeof(double));
  {
    double Vg[MAXN * MAXN];
    vandermonde(ng, s->xiCL, s->N, s->xGP, Vg);
    double *tmp = xcalloc((size_t)nDof * 9, sizeof(double));
    double *dxin = xcalloc((size_t)nGP * 3, sizeof(double));
    double *dxout = xcalloc((size_t)nDof * 3, sizeof(double));
    for (int e = 0; e < ne; e++) {
      const double *X = &s->XCL[(size_t)e * nGP * 3];
      change_basis_3d(ng, s->N, Vg, X, &s->xGPe[(size_t)e * nDof * 3]);
      /* Jacobian columns d: interpolate dX(:,d) to the Gauss points */
      for (int d = 0; d < 3; d++) {
        for (int p = 0; p < nGP; p++)
          for (int x = 0; x < 3; x++) dxin[p * 3 + x] = s->dXCL[((size_t)e * nGP + p) * 9 + x * 3 + d];
        change_basis_3d(ng, s->N, Vg, dxin, dxout);
        for (int p = 0; p < nDof; p++)
          for (int x = 0; x < 3; x++) tmp[p * 9 + x * 3 + d] = dxout[p * 3 + x];
      }
      for (int p = 0; p < nDof; p++) {
        const double *J = &tmp[p * 9]; /* J[x*3+d] */
        /* DetJac_Ref formula, metrics.f90:343-347, with dX(d,x) = J[x*3+d] */
#define DX(d, x) J[(x) * 3 + (d)]
        double det = DX(0, 0) * (DX(1, 1) * DX(2, 2) - DX(2, 1) * DX(1, 2)) +
                     DX(1, 0) * (DX(2, 1) * DX(0, 2) - DX(0, 1) * DX(2, 2)) +
                     DX(2, 0) * (DX(0, 1) * DX(1, 2) - DX(1, 1) * DX(0, 2));
#undef DX
        s->sJ[(size_t)e * nDof + p] = 1.0 / det;
      }
    }
    free(tmp); free(dxin); free(dxout);
  }
  (void)n1;

  /* ElemCurved (particle_mesh_tools.f90:976-996): the trilinear map of the 8
   * corner CL points differs from XCL_NGeo by more than 1e-14 somewhere;
   * dXCL_NGeo at the corners feeds the trilinear Newton of the others. */
  s->curved = xcalloc(ne, 1);
  s->dXCL1 = xcalloc((size_t)ne * 72, sizeof(double));
  {
    double xi1[2] = {-1.0, 1.0}, V1[MAXN * 2];
    vandermonde(1, xi1, ng, s->xiCL, V1);
    double X1[8 * 3], Xn[MAXN * MAXN * MAXN * 3];
    for (int e = 0; e < ne; e++) {
      const double *X = &s->XCL[(size_t)e * nGP * 3];
      for (int t = 0; t < 8; t++) {
        int i = (t & 1) * ng, j = ((t >> 1) & 1) * ng, k = ((t >> 2) & 1) * ng;
        int p = (k * g1 + j) * g1 + i;
        for (int c = 0; c < 3; c++) X1[t * 3 + c] = X[p * 3 + c];
        for (int q = 0; q < 9; q++) s->dXCL1[(size_t)e * 72 + t * 9 + q] = s->dXCL[((size_t)e * nGP + p) * 9 + q];
      }
      if (ng == 1) continue;
      change_basis_3d(1, ng, V1, X1, Xn);
      for (int p = 0; p < nGP && !s->curved[e]; p++)
        for (int c = 0; c < 3; c++)
          if (fabs(Xn[p * 3 + c] - X[p * 3 + c]) > 1e-14) { s->curved[e] = 1; break; }
    }
  }

  /* elements with an open particle boundary side */
  s->openElem = xcalloc(ne, 1);
  for (int sd = 0; sd < s->nSides; sd++) {
    int bc = SIDEINFO(sd, SIDE_BCID);
    if (bc > 0 && s->partBC[bc - 1] == 1) s->openElem[s->sideElem[sd]] = 1;
  }
  /* GEO%x/y/zminglob and maxglob over the mesh nodes (TriaTracking,
   * particle_mesh_tools.f90:775-783; RefMapping replaces them below);
   * DirPeriodicVectors */
  for (int c = 0; c < 3; c++) { s->gmin[c] = 1e300; s->gmax[c] = -1e300; }
  for (int i = 0; i < s->nNodes; i++)
    for (int c = 0; c < 3; c++) {
      if (s->NodeCoords[(size_t)i * 3 + c] < s->gmin[c]) s->gmin[c] = s->NodeCoords[(size_t)i * 3 + c];
      if (s->NodeCoords[(size_t)i * 3 + c] > s->gmax[c]) s->gmax[c] = s->NodeCoords[(size_t)i * 3 + c];
    }
  for (int v = 0; v < s->nPV; v++) {
    s->dirPV[v] = -1;
    for (int c = 0; c < 3; c++)
      if (s->PV[v][c] != 0.0) {
        if (s->dirPV[v] >= 0) { set_err("periodic vector %d is not aligned with an axis", v + 1); return -1; }
        s->dirPV[v] = c;
      }
  }

  if (s->par.tracking == OM_TRACK_REFMAPPING) {
    /* ElemBaryNGeo = X(0,0,0) (BuildElementOriginShared, particle_mesh_build.f90:1385-1475) */
    double L0[MAXN];
    om_lagrange_polys(0.0, ng, s->xiCL, s->wBaryCL, L0);
    for (int e = 0; e < ne; e++) {
      const double *X = &s->XCL[(size_t)e * nGP * 3];
      double xp[3] = {0, 0, 0};
      for (int k = 0; k < g1; k++) for (int j = 0; j < g1; j++) {
        double buf = L0[j] * L0[k];
        for (int i = 0; i < g1; i++)
          for (int c = 0; c < 3; c++) xp[c] = xp[c] + X[((k * g1 + j) * g1 + i) * 3 + c] * L0[i] * buf;
      }
      for (int c = 0; c < 3; c++) s->bary[e * 3 + c] = xp[c];
    }
    /* XiEtaZetaBasis from that origin and ElemRadiusNGeo = the largest distance
     * of a Bezier control point of the six faces (BuildElementBasisAndRadius,
     * particle_mesh_build.f90:1546-1685).  Face control points: the face's
     * CL points mapped to the Bernstein basis. */
    const double Xi[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
    double Vb[MAXN * MAXN], sVb[MAXN * MAXN];
    for (int a = 0; a < g1; a++)
      for (int i = 0; i < g1; i++) {
        double binom = 1.0;
        for (int q = 1; q <= i; q++) binom = binom * (double)(ng - q + 1) / (double)q;
        Vb[a * g1 + i] = binom * pow(0.5 * (1.0 + s->xiCL[a]), i) * pow(0.5 * (1.0 - s->xiCL[a]), ng - i);
      }
    { /* Gauss-Jordan inverse of the small Bernstein Vandermonde matrix */
      double A[MAXN][2 * MAXN];
      for (int r = 0; r < g1; r++)
        for (int c = 0; c < 2 * g1; c++) A[r][c] = c < g1 ? Vb[r * g1 + c] : (c - g1 == r ? 1.0 : 0.0);
      for (int c = 0; c < g1; c++) {
        int piv = c;
        for (int r = c + 1; r < g1; r++) if (fabs(A[r][c]) > fabs(A[piv][c])) piv = r;
        for (int q = 0; q < 2 * g1; q++) { double t = A[c][q]; A[c][q] = A[piv][q]; A[piv][q] = t; }
        double d = A[c][c];
        for (int q = 0; q < 2 * g1; q++) A[c][q] /= d;
        for (int r = 0; r < g1; r++)
          if (r != c) { double f = A[r][c]; for (int q = 0; q < 2 * g1; q++) A[r][q] -= f * A[c][q]; }
      }
      for (int r = 0; r < g1; r++) for (int c = 0; c < g1; c++) sVb[r * g1 + c] = A[r][g1 + c];
    }
    double bzmin[3] = {1e300, 1e300, 1e300}, bzmax[3] = {-1e300, -1e300, -1e300};
    for (int e = 0; e < ne; e++) {
      const double *X = &s->XCL[(size_t)e * nGP * 3];
      for (int d = 0; d < 6; d++) {
        double L1[MAXN], L2[MAXN], L3[MAXN];
        om_lagrange_polys(Xi[d][0], ng, s->xiCL, s->wBaryCL, L1);
        om_lagrange_polys(Xi[d][1], ng, s->xiCL, s->wBaryCL, L2);
        om_lagrange_polys(Xi[d][2], ng, s->xiCL, s->wBaryCL, L3);
        double xp[3] = {0, 0, 0};
        for (int k = 0; k < g1; k++) for (int j = 0; j < g1; j++) for (int i = 0; i < g1; i++)
          for (int c = 0; c < 3; c++) xp[c] = xp[c] + X[((k * g1 + j) * g1 + i) * 3 + c] * L1[i] * L2[j] * L3[k];
        double *b = &s->xiBasis[(e * 6 + d) * 3];
        for (int c = 0; c < 3; c++) b[c] = xp[c] - s->bary[e * 3 + c];
        s->slen[e * 6 + d] = 1.0 / (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      }
      double r = 0.0;
      for (int f = 0; f < 6; f++) {
        int ax = f % 3, fixed = f < 3 ? ng : 0; /* faces xi=+-1, eta=+-1, zeta=+-1 */
        double P[MAXN][MAXN][3], Bz[MAXN][MAXN][3], T[MAXN][MAXN][3];
        for (int a = 0; a < g1; a++) for (int bq = 0; bq < g1; bq++) {
          int ijk[3];
          ijk[ax] = fixed;
          ijk[(ax + 1) % 3] = a;
          ijk[(ax + 2) % 3] = bq;
          for (int c = 0; c < 3; c++) P[a][bq][c] = X[((ijk[2] * g1 + ijk[1]) * g1 + ijk[0]) * 3 + c];
        }
        for (int a = 0; a < g1; a++) for (int bq = 0; bq < g1; bq++) for (int c = 0; c < 3; c++) {
          double acc = 0.0;
          for (int q = 0; q < g1; q++) acc += sVb[a * g1 + q] * P[q][bq][c];
          T[a][bq][c] = acc;
        }
        for (int a = 0; a < g1; a++) for (int bq = 0; bq < g1; bq++) for (int c = 0; c < 3; c++) {
          double acc = 0.0;
          for (int q = 0; q < g1; q++) acc += sVb[bq * g1 + q] * T[a][q][c];
          Bz[a][bq][c] = acc;
        }
        for (int a = 0; a < g1; a++) for (int bq = 0; bq < g1; bq++) {
          double d0 = Bz[a][bq][0] - s->bary[e * 3], d1 = Bz[a][bq][1] - s->bary[e * 3 + 1], d2 = Bz[a][bq][2] - s->bary[e * 3 + 2];
          double v = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
          if (v > r) r = v;
          for (int c = 0; c < 3; c++) {
            if (Bz[a][bq][c] < bzmin[c]) bzmin[c] = Bz[a][bq][c];
            if (Bz[a][bq][c] > bzmax[c]) bzmax[c] = Bz[a][bq][c];
          }
        }
      }
      s->radius[e] = r;
    }
    /* with RefMapping GEO%x/y/zminglob, maxglob bound the sides' Bezier
     * control points (GetMeshMinMax, particle_mesh_tools.f90:735-747) */
    for (int c = 0; c < 3; c++) { s->gmin[c] = bzmin[c]; s->gmax[c] = bzmax[c]; }
    /* ElemEpsOneCell = 1 + sqrt(3 scaleJ RefMappingEps), scaleJ = max sJ / min sJ
     * (BuildEpsOneCell, particle_mesh_build.f90:447-642) */
    s->epsOneCell = xcalloc(ne, sizeof(double));
    for (int e = 0; e < ne; e++) {
      double mx = -1e300, mn = 1e300;
      for (int p = 0; p < nDof; p++) {
        double v = s->sJ[(size_t)e * nDof + p];
        if (v > mx) mx = v;
        if (v < mn) mn = v;
      }
      s->epsOneCell[e] = 1.0 + sqrt(3.0 * (mx / mn) * s->par.refmapping_eps);
    }
    if (build_bc_tables(s) != 0) return -1;
  }

  /* Periodic node twins (InitializePeriodicNodes, pic_depo.f90:1218-): every
   * unique node that coincides with this node under a combination of the
   * periodic vectors with coefficients in {-1,0,1}.  The reference finds the
   * same closure by walking periodic sides; here the nodes are matched by
   * position. */
  {
    const int nu = s->nUnique;
    double *ux = xcalloc((size_t)nu * 3, sizeof(double));
    for (int i = 0; i < s->nNodes; i++)
      for (int c = 0; c < 3; c++) ux[(size_t)s->NodeInfo[i] * 3 + c] = s->NodeCoords[(size_t)i * 3 + c];
    s->perOff = xcalloc((size_t)nu + 1, sizeof(int32_t));
    int cap = 16, cnt = 0;
    s->perNodes = xcalloc(cap, sizeof(int32_t));
    /* spatial hash: sort unique nodes by rounded x for matching */
    double scale = 0.0;
    for (int e = 0; e < ne; e++) if (s->radius[e] > scale) scale = s->radius[e];
    const double tol = 1e-9 * (scale > 0 ? scale : 1.0);
    int ncomb = 1;
    for (int v = 0; v < s->nPV; v++) ncomb *= 3;
    /* bucket grid over unique nodes */
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int u = 0; u < nu; u++) for (int c = 0; c < 3; c++) {
      if (ux[u * 3 + c] < lo[c]) lo[c] = ux[u * 3 + c];
      if (ux[u * 3 + c] > hi[c]) hi[c] = ux[u * 3 + c];
    }
    int gb[3]; double gd[3];
    for (int c = 0; c < 3; c++) {
      double ext = hi[c] - lo[c];
      gb[c] = (int)cbrt((double)nu) + 1;
      gd[c] = ext > 0 ? ext / gb[c] : 1.0;
    }
    int nbk = gb[0] * gb[1] * gb[2];
    int32_t *bcount = xcalloc((size_t)nbk + 1, sizeof(int32_t)), *blist = xcalloc(nu, sizeof(int32_t));
#define BIDX(p, c) ((int)fmin(fmax(floor(((p) - lo[c]) / gd[c]), 0.0), gb[c] - 1))
    for (int u = 0; u < nu; u++) {
      int b = (BIDX(ux[u * 3 + 2], 2) * gb[1] + BIDX(ux[u * 3 + 1], 1)) * gb[0] + BIDX(ux[u * 3], 0);
      bcount[b + 1]++;
    }
    for (int b = 0; b < nbk; b++) bcount[b + 1] += bcount[b];
    int32_t *bfill = xcalloc(nbk, sizeof(int32_t));
    for (int u = 0; u < nu; u++) {
      int b = (BIDX(ux[u * 3 + 2], 2) * gb[1] + BIDX(ux[u * 3 + 1], 1)) * gb[0] + BIDX(ux[u * 3], 0);
      blist[bcount[b] + bfill[b]++] = u;
    }
    for (int u = 0; u < nu; u++) {
      s->perOff[u] = cnt;
      for (int comb = 0; comb < ncomb; comb++) {
        int cc = comb, coef[3] = {0, 0, 0};
        for (int v = 0; v < s->nPV; v++) { coef[v] = cc % 3 - 1; cc /= 3; }
        if (coef[0] == 0 && coef[1] == 0 && coef[2] == 0) continue;
        double p[3];
        for (int c = 0; c < 3; c++) {
          p[c] = ux[u * 3 + c];
          for (int v = 0; v < s->nPV; v++) p[c] += coef[v] * s->PV[v][c];
        }
        int inside = 1;
        for (int c = 0; c < 3; c++) if (p[c] < lo[c] - tol || p[c] > hi[c] + tol) inside = 0;
        if (!inside) continue;
        int bi[3];
        for (int c = 0; c < 3; c++) bi[c] = BIDX(p[c], c);
        for (int dz = -1; dz <= 1; dz++) for (int dy = -1; dy <= 1; dy++) for (int dx = -1; dx <= 1; dx++) {
          int z = bi[2] + dz, y = bi[1] + dy, x = bi[0] + dx;
          if (x < 0 || y < 0 || z < 0 || x >= gb[0] || y >= gb[1] || z >= gb[2]) continue;
          int b = (z * gb[1] + y) * gb[0] + x;
          for (int q = bcount[b]; q < bcount[b + 1]; q++) {
            int w = blist[q];
            if (w == u) continue;
            if (fabs(ux[w * 3] - p[0]) <= tol && fabs(ux[w * 3 + 1] - p[1]) <= tol && fabs(ux[w * 3 + 2] - p[2]) <= tol) {
              int dup = 0;
              for (int t = s->perOff[u]; t < cnt; t++) if (s->perNodes[t] == w) dup = 1;
              if (!dup) {
                if (cnt == cap) { cap *= 2; s->perNodes = realloc(s->perNodes, cap * sizeof(int32_t)); }
                s->perNodes[cnt++] = w;
              }
            }
          }
        }
      }
    }
    s->perOff[nu] = cnt;
#undef BIDX
    free(bcount); free(blist); free(bfill); free(ux);
  }

  /* NodeVolume, CalcCellLocNodeVolumes, pic_depo_tools.f90:224-355 */
  {
    const int nu = s->nUnique;
    s->NodeVolume = xcalloc(nu, sizeof(double));
    const double *x = s->xGP, *w = s->wGP;
    for (int e = 0; e < ne; e++) {
      int NodeID[8];
      for (int c = 0; c < 8; c++) NodeID[c] = s->NodeInfo[s->ElemNodeID[e * 8 + c]];
      for (int i = 0; i <= s->N; i++) for (int j = 0; j <= s->N; j++) for (int k = 0; k <= s->N; k++) {
        int r = (k * (s->N + 1) + j) * (s->N + 1) + i;
        double sJ = s->sJ[(size_t)e * nDof + r];
        double *NV = s->NodeVolume;
        NV[NodeID[0]] = NV[NodeID[0]] + 1. / sJ * ((1. - x[i]) * (1. - x[j]) * (1. - x[k]) * w[i] * w[j] * w[k] / 8.);
        NV[NodeID[1]] = NV[NodeID[1]] + 1. / sJ * ((1. + x[i]) * (1. - x[j]) * (1. - x[k]) * w[i] * w[j] * w[k] / 8.);
        NV[NodeID[2]] = NV[NodeID[2]] + 1. / sJ * ((1. + x[i]) * (1. + x[j]) * (1. - x[k]) * w[i] * w[j] * w[k] / 8.);
        NV[NodeID[3]] = NV[NodeID[3]] + 1. / sJ * ((1. - x[i]) * (1. + x[j]) * (1. - x[k]) * w[i] * w[j] * w[k] / 8.);
        NV[NodeID[4]] = NV[NodeID[4]] + 1. / sJ * ((1. - x[i]) * (1. - x[j]) * (1. + x[k]) * w[i] * w[j] * w[k] / 8.);
        NV[NodeID[5]] = NV[NodeID[5]] + 1. / sJ * ((1. + x[i]) * (1. - x[j]) * (1. + x[k]) * w[i] * w[j] * w[k] / 8.);
        NV[NodeID[6]] = NV[NodeID[6]] + 1. / sJ * ((1. + x[i]) * (1. + x[j]) * (1. + x[k]) * w[i] * w[j] * w[k] / 8.);
        NV[NodeID[7]] = NV[NodeID[7]] + 1. / sJ * ((1. - x[i]) * (1. + x[j]) * (1. + x[k]) * w[i] * w[j] * w[k] / 8.);
      }
    }
    if (s->nPV > 0) {
      double *loc = xcalloc(nu, sizeof(double));
      for (int u = 0; u < nu; u++)
        for (int t = s->perOff[u]; t < s->perOff[u + 1]; t++) loc[u] = loc[u] + s->NodeVolume[s->perNodes[t]];
      for (int u = 0; u < nu; u++) s->NodeVolume[u] = s->NodeVolume[u] + loc[u];
      free(loc);
    }
  }

  /* Background bins over element bounding boxes (the role of FIBGM,
   * particle_bgm.f90:79): cell size = the largest element extent. */
  {
    double ext = 0.0;
    for (int c = 0; c < 3; c++) { s->bmin[c] = 1e300; s->bmax[c] = -1e300; }
    double *emin = xcalloc((size_t)ne * 3, sizeof(double)), *emax = xcalloc((size_t)ne * 3, sizeof(double));
    for (int e = 0; e < ne; e++) {
      for (int c = 0; c < 3; c++) { emin[e * 3 + c] = 1e300; emax[e * 3 + c] = -1e300; }
      int fn = ELEMINFO(e, ELEM_FIRSTNODEIND), ln = ELEMINFO(e, ELEM_LASTNODEIND);
      if (s->par.tracking == OM_TRACK_REFMAPPING) {
        /* the bounding sphere (ElemBaryNGeo, ElemRadiusNGeo) holds the curved element */
        for (int c = 0; c < 3; c++) {
          emin[e * 3 + c] = s->bary[e * 3 + c] - s->radius[e];
          emax[e * 3 + c] = s->bary[e * 3 + c] + s->radius[e];
        }
      } else
        for (int nn = fn; nn < ln; nn++) for (int c = 0; c < 3; c++) {
          double v = s->NodeCoords[(size_t)nn * 3 + c];
          if (v < emin[e * 3 + c]) emin[e * 3 + c] = v;
          if (v > emax[e * 3 + c]) emax[e * 3 + c] = v;
        }
      for (int c = 0; c < 3; c++) {
        if (emin[e * 3 + c] < s->bmin[c]) s->bmin[c] = emin[e * 3 + c];
        if (emax[e * 3 + c] > s->bmax[c]) s->bmax[c] = emax[e * 3 + c];
        if (emax[e * 3 + c] - emin[e * 3 + c] > ext) ext = emax[e * 3 + c] - emin[e * 3 + c];
      }
    }
    for (int c = 0; c < 3; c++) {
      double L = s->bmax[c] - s->bmin[c];
      s->nb[c] = (int)ceil(L / ext);
      if (s->nb[c] < 1) s->nb[c] = 1;
      s->bdelta[c] = L / s->nb[c];
      if (s->bdelta[c] <= 0) s->bdelta[c] = 1.0;
    }
    size_t nbins = (size_t)s->nb[0] * s->nb[1] * s->nb[2];
    s->binOff = xcalloc(nbins + 1, sizeof(int32_t));
    for (int pass = 0; pass < 2; pass++) {
      int32_t *fill = pass ? xcalloc(nbins, sizeof(int32_t)) : NULL;
      for (int e = 0; e < ne; e++) {
        int lo3[3], hi3[3];
        for (int c = 0; c < 3; c++) {
          lo3[c] = (int)floor((emin[e * 3 + c] - s->bmin[c]) / s->bdelta[c]);
          hi3[c] = (int)floor((emax[e * 3 + c] - s->bmin[c]) / s->bdelta[c]);
          if (lo3[c] < 0) lo3[c] = 0;
          if (hi3[c] > s->nb[c] - 1) hi3[c] = s->nb[c] - 1;
        }
        for (int z = lo3[2]; z <= hi3[2]; z++) for (int y = lo3[1]; y <= hi3[1]; y++) for (int x = lo3[0]; x <= hi3[0]; x++) {
          size_t b = ((size_t)z * s->nb[1] + y) * s->nb[0] + x;
          if (!pass) s->binOff[b + 1]++;
          else s->binElems[s->binOff[b] + fill[b]++] = e;
        }
      }
      if (!pass) {
        for (size_t b = 0; b < nbins; b++) s->binOff[b + 1] += s->binOff[b];
        s->binElems = xcalloc(s->binOff[nbins], sizeof(int32_t));
      } else free(fill);
    }
    free(emin); free(emax);
  }
  return 0;
}

/* ================================================================ eval_xyz.f90 */
/* getDet / getInv, eval_xyz.f90:448-495 (M[r][c] = Mat(r+1,c+1)) */
static double get_det(const double M[3][3]) {
  return (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * M[2][2] +
         (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * M[2][0] +
         (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * M[2][1];
}
static void get_inv(const double M[3][3], double sdet, double I[3][3]) {
  I[0][0] = (M[1][1] * M[2][2] - M[1][2] * M[2][1]) * sdet;
  I[0][1] = (M[0][2] * M[2][1] - M[0][1] * M[2][2]) * sdet;
  I[0][2] = (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * sdet;
  I[1][0] = (M[1][2] * M[2][0] - M[1][0] * M[2][2]) * sdet;
  I[1][1] = (M[0][0] * M[2][2] - M[0][2] * M[2][0]) * sdet;
  I[1][2] = (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * sdet;
  I[2][0] = (M[1][0] * M[2][1] - M[1][1] * M[2][0]) * sdet;
  I[2][1] = (M[0][1] * M[2][0] - M[0][0] * M[2][1]) * sdet;
  I[2][2] = (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * sdet;
}

/* RefElemNewton, eval_xyz.f90:298-445.  mode 1 = ForceMode (aborts without
 * isSuccessful), mode 2 = default.  Returns isSuccessful (1) or 0; sets
 * *aborted when the reference would abort. */
static int ref_elem_newton_basis(const om_state *s, double xi[3], const double x_in[3], int ng,
                                 const double *xiCL, const double *wBaryCL, const double *X, const double *dX,
                                 int mode, int want_success, int *aborted) {
  const int g1 = ng + 1;
  double Lag[3][MAXN], F[3], xi_old[3];
  int ok = 1;
  *aborted = 0;
  for (int d = 0; d < 3; d++) om_lagrange_polys(xi[d], ng, xiCL, wBaryCL, Lag[d]);
  for (int c = 0; c < 3; c++) F[c] = -x_in[c];
  for (int k = 0; k < g1; k++) for (int j = 0; j < g1; j++) {
    double buff = Lag[1][j] * Lag[2][k];
    for (int i = 0; i < g1; i++)
      for (int c = 0; c < 3; c++) F[c] = F[c] + X[((k * g1 + j) * g1 + i) * 3 + c] * Lag[0][i] * buff;
  }
  double deltaXi2 = (fabs(F[0]) < EPS_MACH && fabs(F[1]) < EPS_MACH && fabs(F[2]) < EPS_MACH) ? 0.0 : 1.0;
  double NormF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
  double NormF_old = NormF;
  int it = 0;
  while (deltaXi2 > s->par.refmapping_eps && it < 100) {
    it++;
    double Jac[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < g1; k++) for (int j = 0; j < g1; j++) {
      double buff = Lag[1][j] * Lag[2][k];
      for (int i = 0; i < g1; i++) {
        double buff2 = Lag[0][i] * buff;
        const double *dp = &dX[((k * g1 + j) * g1 + i) * 9];
        /* Jac(a,1:3) += dXCL(1:3,a)*buff2 ; dp[x*3+d] = dXCL(d,x) */
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) Jac[a][b] = Jac[a][b] + dp[a * 3 + b] * buff2;
      }
    }
    double sdet = get_det(Jac);
    if (sdet > 0.) sdet = 1. / sdet;
    double sJac[3][3];
    get_inv(Jac, sdet, sJac);
    double dxi[3];
    for (int r = 0; r < 3; r++) dxi[r] = sJac[r][0] * F[0] + sJac[r][1] * F[1] + sJac[r][2] * F[2];
    deltaXi2 = dxi[0] * dxi[0] + dxi[1] * dxi[1] + dxi[2] * dxi[2];
    for (int d = 0; d < 3; d++) xi_old[d] = xi[d];
    NormF_old = NormF;
    NormF = NormF * 2.;
    double lambda = 1.0;
    int iArmijo = 1;
    while (NormF > NormF_old * (1. - 0.0001 * lambda) && iArmijo <= 8) {
      for (int d = 0; d < 3; d++) xi[d] = xi_old[d] - lambda * dxi[d];
      for (int d = 0; d < 3; d++) om_lagrange_polys(xi[d], ng, xiCL, wBaryCL, Lag[d]);
      for (int c = 0; c < 3; c++) F[c] = -x_in[c];
      for (int k = 0; k < g1; k++) for (int j = 0; j < g1; j++) {
        double buff = Lag[1][j] * Lag[2][k];
        for (int i = 0; i < g1; i++) {
          double buff2 = Lag[0][i] * buff;
          for (int c = 0; c < 3; c++) F[c] = F[c] + X[((k * g1 + j) * g1 + i) * 3 + c] * buff2;
        }
      }
      lambda = 0.2 * lambda;
      iArmijo = iArmijo + 1;
      NormF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
    }
    if (fabs(xi[0]) > 1.5 || fabs(xi[1]) > 1.5 || fabs(xi[2]) > 1.5) {
      if (want_success) { ok = 0; break; }
      else if (mode == 1) { *aborted = 1; break; }
      else break;
    }
  }
  return ok;
}

/* GetPositionInRefElem's choice of map (eval_xyz.f90:86-117): the full NGeo
 * map for curved elements and NGeo = 1 meshes, else the trilinear map of the
 * 8 corner CL points (XCL_NGeo1, dXCL_NGeo1). */
static int ref_elem_newton(const om_state *s, double xi[3], const double x_in[3], int e,
                           int mode, int want_success, int *aborted) {
  if (s->NGeo == 1 || (s->curved && s->curved[e]))
    return ref_elem_newton_basis(s, xi, x_in, s->NGeo, s->xiCL, s->wBaryCL, &s->XCL[(size_t)e * s->nGP * 3],
                                 &s->dXCL[(size_t)e * s->nGP * 9], mode, want_success, aborted);
  static const double xi1[2] = {-1.0, 1.0}, wb1[2] = {-0.5, 0.5};
  const int g = s->NGeo, g1 = g + 1;
  double X1[8][3];
  for (int t = 0; t < 8; t++) {
    int i = (t & 1) * g, j = ((t >> 1) & 1) * g, k = ((t >> 2) & 1) * g;
    for (int c = 0; c < 3; c++) X1[t][c] = s->XCL[((size_t)e * s->nGP + (k * g1 + j) * g1 + i) * 3 + c];
  }
  return ref_elem_newton_basis(s, xi, x_in, 1, xi1, wb1, &X1[0][0], &s->dXCL1[(size_t)e * 72], mode, want_success,
                               aborted);
}

/* GetRefNewtonStartValue, eval_xyz.f90:498-612 */
static void ref_newton_start(const om_state *s, const double x[3], double xi[3], int e) {
  const double epsOne = 1.0 + s->par.refmapping_eps;
  switch (s->par.refmapping_guess) {
  case 1: {
    double P[3], XL[6];
    for (int c = 0; c < 3; c++) P[c] = x[c] - s->bary[e * 3 + c];
    for (int d = 0; d < 6; d++) {
      const double *b = &s->xiBasis[(e * 6 + d) * 3];
      XL[d] = (P[0] * b[0] + P[1] * b[1] + P[2] * b[2]) * s->slen[e * 6 + d];
    }
    for (int d = 0; d < 3; d++) xi[d] = 0.5 * (XL[d] - XL[d + 3]);
    double mx = fmax(fabs(xi[0]), fmax(fabs(xi[1]), fabs(xi[2])));
    if (mx > epsOne)
      for (int d = 0; d < 3; d++) xi[d] = fmax(fmin(1.0, xi[d]), -1.0);
    break;
  }
  case 2: {
    const int n1 = s->N + 1;
    const double *xg = &s->xGPe[(size_t)e * s->nDof * 3];
    double d0[3] = {x[0] - xg[0], x[1] - xg[1], x[2] - xg[2]};
    double win = sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
    xi[0] = xi[1] = xi[2] = s->xGP[0];
    for (int i = 0; i < n1; i++) for (int j = 0; j < n1; j++) for (int k = 0; k < n1; k++) {
      int r = (k * n1 + j) * n1 + i;
      double dX = fabs(x[0] - xg[r * 3]);
      if (dX > win) continue;
      double dY = fabs(x[1] - xg[r * 3 + 1]);
      if (dY > win) continue;
      double dZ = fabs(x[2] - xg[r * 3 + 2]);
      if (dZ > win) continue;
      double dist = sqrt(dX * dX + dY * dY + dZ * dZ);
      if (dist < win) { win = dist; xi[0] = s->xGP[i]; xi[1] = s->xGP[j]; xi[2] = s->xGP[k]; }
    }
    break;
  }
  case 3: {
    const int g1 = s->NGeo + 1;
    const double *X = &s->XCL[(size_t)e * s->nGP * 3];
    double d0[3] = {x[0] - X[0], x[1] - X[1], x[2] - X[2]};
    double win = sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
    xi[0] = xi[1] = xi[2] = s->xiCL[0];
    for (int i = 0; i < g1; i++) for (int j = 0; j < g1; j++) for (int k = 0; k < g1; k++) {
      int r = (k * g1 + j) * g1 + i;
      double dX = fabs(x[0] - X[r * 3]);
      if (dX > win) continue;
      double dY = fabs(x[1] - X[r * 3 + 1]);
      if (dY > win) continue;
      double dZ = fabs(x[2] - X[r * 3 + 2]);
      if (dZ > win) continue;
      double dist = sqrt(dX * dX + dY * dY + dZ * dZ);
      if (dist < win) { win = dist; xi[0] = s->xiCL[i]; xi[1] = s->xiCL[j]; xi[2] = s->xiCL[k]; }
    }
    break;
  }
  default:
    xi[0] = xi[1] = xi[2] = 0.0;
  }
}

/* GetPositionInRefElem, eval_xyz.f90:35-123 (curved/NGeo=1 branch; with
 * TriaTracking ElemCurved is .FALSE. and NGeo=1 meshes take the full map) */
static int position_in_ref_elem(const om_state *s, const double x[3], double xi[3], int e,
                                int force_mode, int want_success, int *aborted) {
  ref_newton_start(s, x, xi, e);
  return ref_elem_newton(s, xi, x, e, force_mode ? 1 : 2, want_success, aborted);
}

/* EvaluateFieldAtRefPos, eval_xyz.f90:167-295 (NVar_in = 3, no BG field) */
static void eval_field_at_ref_pos(const om_state *s, const double xi[3], int e, double U[3]) {
  const int N = s->N, n1 = N + 1;
  double L[3][MAXN];
  for (int d = 0; d < 3; d++) om_lagrange_polys(xi[d], N, s->xGP, s->wBaryGP, L[d]);
  const double *Ue = &s->E[(size_t)e * s->nDof * 3];
  U[0] = U[1] = U[2] = 0.0;
  for (int k = 0; k < n1; k++) for (int j = 0; j < n1; j++) {
    double Lez = L[1][j] * L[2][k];
    for (int i = 0; i < n1; i++) {
      const double *u = &Ue[((k * n1 + j) * n1 + i) * 3];
      for (int c = 0; c < 3; c++) U[c] = U[c] + u[c] * L[0][i] * Lez;
    }
  }
}

/* GetEMFieldDW, pic_interpolation_tools.f90:458-572 (E only).  The triple
 * loop is written DO k; DO l; DO m, so the EXIT taken when the particle sits
 * on a Gauss point leaves only the m-loop: the weights of the points visited
 * before are zeroed, that point gets 1, and the loop goes on with the next l. */
static void em_field_dw(const om_state *s, int e, const double x[3], double U[3]) {
  const int n1 = s->N + 1, nd = s->nDof;
  double dist[MAXN * MAXN * MAXN];
  double sum = 0.0;
  for (int q = 0; q < nd; q++) dist[q] = 0.0;
  const double *xg = &s->xGPe[(size_t)e * nd * 3];
  for (int k = 0; k < n1; k++) for (int l = 0; l < n1; l++) for (int m = 0; m < n1; m++) {
    int r = (m * n1 + l) * n1 + k;
    double d[3] = {xg[r * 3] - x[0], xg[r * 3 + 1] - x[1], xg[r * 3 + 2] - x[2]};
    double norm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (norm > 0.) { dist[r] = 1. / norm; }
    else {
      for (int q = 0; q < nd; q++) dist[q] = 0.0;
      dist[r] = 1.0; sum = 1.0;
      break;
    }
    sum = sum + dist[r];
  }
  U[0] = U[1] = U[2] = 0.0;
  const double *Ue = &s->E[(size_t)e * nd * 3];
  for (int k = 0; k < n1; k++) for (int l = 0; l < n1; l++) for (int m = 0; m < n1; m++) {
    int r = (m * n1 + l) * n1 + k;
    for (int c = 0; c < 3; c++) U[c] = U[c] + dist[r] / sum * Ue[r * 3 + c];
  }
}

/* ================================================================ API basics */
om_state *om_create(const om_mesh_desc *mesh, const om_params *par) {
  if (par->tracking != OM_TRACK_TRIA && par->tracking != OM_TRACK_REFMAPPING) { set_err("tracking method"); return NULL; }
  if (par->tracking == OM_TRACK_REFMAPPING && par->deposition == OM_DEPO_CVWM) {
    /* pic_depo_method.f90:127-129 */
    set_err("cell_volweight_mean needs TriaTracking");
    return NULL;
  }
  om_state *s = xcalloc(1, sizeof(om_state));
  s->par = *par;
  s->nthreads = 1;
  if (build_mesh(s, mesh) != 0) { om_destroy(s); return NULL; }
  s->E = xcalloc((size_t)s->nElems * s->nDof * 3, sizeof(double));
  s->PS = xcalloc((size_t)s->nElems * s->nDof * 4, sizeof(double));
  s->NodeSource = xcalloc((size_t)s->nUnique * 4, sizeof(double));
  /* shape function constants, InitShapeFunctionDimensionalty 3D branch,
   * pic_depo_shapefunction_tools.f90:1131-1294, beta() pic_depo_tools.f90:491-509 */
  if (par->deposition == OM_DEPO_SF) {
    s->r_sf = par->sf_radius;
    s->r2_sf = s->r_sf * s->r_sf;
    s->r2_sf_inv = 1. / s->r2_sf;
    double a = (double)par->sf_alpha;
    double BetaFac = tgamma(1.5) * tgamma(a + 1.) / tgamma(1.5 + (a + 1.));
    s->w_sf = 1. / (2. * BetaFac * a + 2 * BetaFac) * (a + 1.) / (PI_ * (s->r_sf * s->r_sf * s->r_sf));
    /* InitPeriodicSFCaseMatrix, pic_depo.f90:828-910 (dim_sf = 3) */
    if (s->nPV <= 0) s->nCases = 0;
    else if (s->nPV < 3) {
      /* the reference's 3D table (below) adds PeriodicVectors(:,2:3), which
       * do not exist with fewer vectors; use the 3^nPV combinations of the
       * vectors that do, in the same +1 / 0 / -1 order */
      s->nCases = s->nPV == 1 ? 3 : 9;
      memset(s->caseMat, 0, sizeof s->caseMat);
      for (int c = 0; c < s->nCases; c++) {
        int cc = c;
        for (int v = s->nPV - 1; v >= 0; v--) { s->caseMat[c][v] = 1 - cc % 3; cc /= 3; }
      }
    } else {
      s->nCases = 27;
      memset(s->caseMat, 0, sizeof s->caseMat);
      for (int c = 0; c < 9; c++) s->caseMat[c][0] = 1;
      for (int c = 18; c < 27; c++) s->caseMat[c][0] = -1;
      for (int I = 1; I <= 3; I++) {
        for (int c = I * 9 - 8; c <= I * 9 - 6; c++) s->caseMat[c - 1][1] = 1;
        for (int c = I * 9 - 2; c <= I * 9; c++) s->caseMat[c - 1][1] = -1;
        for (int J = 1; J <= 3; J++) {
          s->caseMat[(J * 3 - 2) + (I - 1) * 9 - 1][2] = 1;
          s->caseMat[(J * 3) + (I - 1) * 9 - 1][2] = -1;
        }
      }
    }
  }
  return s;
}

void om_destroy(om_state *s) {
  if (!s) return;
  void *ptrs[] = {s->ElemInfo, s->SideInfo, s->NodeCoords, s->NodeInfo, s->bcAlpha, s->partBC, s->sideElem,
                  s->sideLocal, s->sideNbSide, s->ElemSideNodeID, s->Concave, s->ElemNodeID, s->XCL, s->dXCL,
                  s->bary, s->radius, s->xiBasis, s->slen, s->xGPe, s->sJ, s->perOff, s->perNodes,
                  s->NodeVolume, s->binOff, s->binElems, s->E, s->PS, s->NodeSource, s->PartState,
                  s->LastPartPos, s->Field, s->elem, s->lastElem, s->species, s->inside, s->isNew,
                  s->curved, s->dXCL1, s->epsOneCell, s->openElem, s->PartPosRef, s->Gop, s->rhoBar, s->Enode,
                  s->bcSide, s->sideToBC, s->bz, s->sideType, s->bbEmpty, s->sideNorm, s->sideDist, s->slabN,
                  s->slabI, s->baseV, s->bcOrigin, s->bcRadius, s->e2bOff, s->e2bSide, s->e2bDist};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++) free(ptrs[i]);
  for (int t = 0; t < s->nPriv; t++) free(s->priv[t]);
  free(s->priv);
  free(s->tbLog);
  free(s);
}

void om_set_threads(om_state *s, int nthreads) { s->nthreads = nthreads > 0 ? nthreads : 1; }

int om_set_field(om_state *s, const double *E) {
  memcpy(s->E, E, (size_t)s->nElems * s->nDof * 3 * sizeof(double));
  return 0;
}

int om_set_particles(om_state *s, int32_t n, const double *PartState, const int32_t *elem,
                     const int32_t *species, const uint8_t *is_new) {
  free(s->PartState); free(s->LastPartPos); free(s->Field); free(s->elem); free(s->lastElem);
  free(s->species); free(s->inside); free(s->isNew); free(s->PartPosRef);
  s->n = n;
  s->PartPosRef = xcalloc((size_t)n * 3, sizeof(double));
  s->PartState = xcalloc((size_t)n * 6, sizeof(double));
  s->LastPartPos = xcalloc((size_t)n * 3, sizeof(double));
  s->Field = xcalloc((size_t)n * 6, sizeof(double));
  s->elem = xcalloc(n, sizeof(int32_t));
  s->lastElem = xcalloc(n, sizeof(int32_t));
  s->species = xcalloc(n, sizeof(int32_t));
  s->inside = xcalloc(n, sizeof(uint8_t));
  s->isNew = xcalloc(n, sizeof(uint8_t));
  memcpy(s->PartState, PartState, (size_t)n * 6 * sizeof(double));
  for (int i = 0; i < n; i++) {
    s->elem[i] = elem ? elem[i] - 1 : -1;
    s->species[i] = species ? species[i] - 1 : 0;
    if (s->species[i] < 0 || s->species[i] >= s->par.nSpecies) { set_err("bad species"); return -1; }
    s->inside[i] = 1;
    s->isNew[i] = is_new ? is_new[i] : 0;
    if (s->elem[i] >= s->nElems) { set_err("bad element id"); return -1; }
    if (s->par.tracking == OM_TRACK_REFMAPPING && s->elem[i] >= 0) {
      /* LocateParticleInElement leaves PartPosRef of the owner (particle_localization.f90:47-74) */
      int ab;
      position_in_ref_elem(s, &s->PartState[(size_t)i * 6], &s->PartPosRef[(size_t)i * 3], s->elem[i], 0, 0, &ab);
    }
  }
  s->lost = 0;
  return 0;
}

int32_t om_get_nparticles(const om_state *s) { return s->n; }
double om_get_private_reduce_seconds(const om_state *s) { return s->tPrivReduce; }
int64_t om_get_lost_count(const om_state *s) { return s->lost; }
int64_t om_get_multi_crossings(const om_state *s) { return s->multiCross; }
int om_set_tiebreak_log(om_state *s, int64_t cap) {
  free(s->tbLog);
  s->tbLog = NULL;
  s->tbN = 0;
  s->tbCap = cap > 0 ? cap : 0;
  if (s->tbCap) s->tbLog = xcalloc((size_t)s->tbCap * OM_TB_REC, sizeof(double));
  return 0;
}
int64_t om_get_tiebreak_log(const om_state *s, double *out) {
  const int64_t n = s->tbN < s->tbCap ? s->tbN : s->tbCap;
  if (out && n) memcpy(out, s->tbLog, (size_t)n * OM_TB_REC * sizeof(double));
  return n;
}
void om_plant_mistake(om_state *s, int which) { s->plant = which; }

int om_get_particles(const om_state *s, double *PartState, int32_t *elem, uint8_t *inside) {
  if (PartState) memcpy(PartState, s->PartState, (size_t)s->n * 6 * sizeof(double));
  for (int i = 0; i < s->n; i++) {
    if (elem) elem[i] = s->elem[i] + 1;
    if (inside) inside[i] = s->inside[i];
  }
  return 0;
}
int om_get_last_part_pos(const om_state *s, double *L) { memcpy(L, s->LastPartPos, (size_t)s->n * 3 * sizeof(double)); return 0; }
int om_get_field_at_particle(const om_state *s, double *F) { memcpy(F, s->Field, (size_t)s->n * 6 * sizeof(double)); return 0; }
int om_get_part_source(const om_state *s, double *PS) { memcpy(PS, s->PS, (size_t)s->nElems * s->nDof * 4 * sizeof(double)); return 0; }
int om_set_part_source(om_state *s, const double *PS) { memcpy(s->PS, PS, (size_t)s->nElems * s->nDof * 4 * sizeof(double)); return 0; }
int om_get_field(const om_state *s, double *E) { memcpy(E, s->E, (size_t)s->nElems * s->nDof * 3 * sizeof(double)); return 0; }
int om_get_elem_xgp(const om_state *s, double *x) { memcpy(x, s->xGPe, (size_t)s->nElems * s->nDof * 3 * sizeof(double)); return 0; }
int om_get_sj(const om_state *s, double *sJ) { memcpy(sJ, s->sJ, (size_t)s->nElems * s->nDof * sizeof(double)); return 0; }
int om_get_elem_side_node_id(const om_state *s, int32_t *id) { memcpy(id, s->ElemSideNodeID, (size_t)s->nElems * 24 * sizeof(int32_t)); return 0; }
int om_get_concave(const om_state *s, uint8_t *c) { memcpy(c, s->Concave, (size_t)s->nElems * 6); return 0; }
int om_get_node_volume(const om_state *s, double *nv) { memcpy(nv, s->NodeVolume, (size_t)s->nUnique * sizeof(double)); return 0; }
int32_t om_get_n_unique_nodes(const om_state *s) { return s->nUnique; }
int om_get_elem_bary(const om_state *s, double *b, double *r) {
  if (b) memcpy(b, s->bary, (size_t)s->nElems * 3 * sizeof(double));
  if (r) memcpy(r, s->radius, (size_t)s->nElems * sizeof(double));
  return 0;
}
int om_get_part_pos_ref(const om_state *s, double *ref) {
  memcpy(ref, s->PartPosRef, (size_t)s->n * 3 * sizeof(double));
  return 0;
}
int om_get_refmapping_tables(const om_state *s, uint8_t *curved, double *eps) {
  if (curved) memcpy(curved, s->curved, (size_t)s->nElems);
  if (eps) for (int e = 0; e < s->nElems; e++) eps[e] = s->epsOneCell ? s->epsOneCell[e] : 0.0;
  return 0;
}
int om_get_xieta_basis(const om_state *s, double *b, double *sl) {
  if (b) memcpy(b, s->xiBasis, (size_t)s->nElems * 18 * sizeof(double));
  if (sl) memcpy(sl, s->slen, (size_t)s->nElems * 6 * sizeof(double));
  return 0;
}

int om_position_in_ref_elem(om_state *s, int32_t n, const double *x, const int32_t *elem, double *xi, uint8_t *ok) {
  for (int i = 0; i < n; i++) {
    int ab;
    ok[i] = (uint8_t)position_in_ref_elem(s, &x[i * 3], &xi[i * 3], elem[i] - 1, 0, 1, &ab);
  }
  return 0;
}
int om_eval_field_at_ref_pos(om_state *s, int32_t n, const double *xi, const int32_t *elem, double *U) {
  for (int i = 0; i < n; i++) eval_field_at_ref_pos(s, &xi[i * 3], elem[i] - 1, &U[i * 3]);
  return 0;
}
int32_t om_get_n_bc_sides(const om_state *s) { return s->nBC; }
int om_get_bc_sides(const om_state *s, int32_t *side, uint8_t *type, double *bez, double *origin, double *radius) {
  const int g1 = s->NGeo + 1;
  for (int b = 0; b < s->nBC; b++) {
    side[b] = s->bcSide[b];
    type[b] = s->sideType[b];
    radius[b] = s->bcRadius[b];
    for (int c = 0; c < 3; c++) origin[b * 3 + c] = s->bcOrigin[b * 3 + c];
  }
  memcpy(bez, s->bz, (size_t)s->nBC * g1 * g1 * 3 * sizeof(double));
  return 0;
}
int32_t om_get_bc_list_size(const om_state *s) { return s->e2bOff ? s->e2bOff[s->nElems] : 0; }
int om_get_bc_elem_lists(const om_state *s, int32_t *off, int32_t *bcside, double *dist) {
  if (!s->e2bOff) return -1;
  memcpy(off, s->e2bOff, (size_t)(s->nElems + 1) * sizeof(int32_t));
  memcpy(bcside, s->e2bSide, (size_t)s->e2bOff[s->nElems] * sizeof(int32_t));
  memcpy(dist, s->e2bDist, (size_t)s->e2bOff[s->nElems] * sizeof(double));
  return 0;
}
int om_bc_side_intersection(om_state *s, int32_t b, const double *last, const double *x, double *out) {
  if (b < 0 || b >= s->nBC) { set_err("boundary side out of range"); return -1; }
  double traj[3] = {x[0] - last[0], x[1] - last[1], x[2] - last[2]};
  const double len = sqrt(dot3(traj, traj));
  for (int c = 0; c < 3; c++) traj[c] = traj[c] / len;
  const int sd = s->bcSide[b];
  const int flip = SIDEINFO(sd, SIDE_ID) > 0 ? 0 : SIDEINFO(sd, SIDE_FLIP) % 10;
  return side_intersection(s, b, flip, last, x, traj, len, &out[0], &out[1], &out[2]);
}
int om_em_field_dw(om_state *s, int32_t n, const double *x, const int32_t *elem, double *U) {
  for (int i = 0; i < n; i++) em_field_dw(s, elem[i] - 1, &x[i * 3], &U[i * 3]);
  return 0;
}

/* ================================================================ tracking */
/* ParticleInsideQuad3D, particle_mesh_tools.f90:78-222 (regular sides) */
static int inside_quad_3d(const om_state *s, const double x[3], int e, double Det[6][2]) {
  int in = 1;
  for (int l = 1; l <= 6; l++) {
    const int32_t *nid = &s->ElemSideNodeID[(e * 6 + l - 1) * 4];
    double A[3][4];
    for (int nn = 0; nn < 4; nn++)
      for (int c = 0; c < 3; c++) A[c][nn] = s->NodeCoords[(size_t)nid[nn] * 3 + c] - x[c];
    double cp[3];
    cp[0] = A[1][0] * A[2][2] - A[2][0] * A[1][2];
    cp[1] = A[2][0] * A[0][2] - A[0][0] * A[2][2];
    cp[2] = A[0][0] * A[1][2] - A[1][0] * A[0][2];
    Det[l - 1][0] = cp[0] * A[0][1] + cp[1] * A[1][1] + cp[2] * A[2][1];
    Det[l - 1][0] = -Det[l - 1][0];
    Det[l - 1][1] = cp[0] * A[0][3] + cp[1] * A[1][3] + cp[2] * A[2][3];
    int neg = 0, pos = 0;
    if (Det[l - 1][0] < 0) neg = 1; else pos = 1;
    if (Det[l - 1][1] < 0) neg = 1; else pos = 1;
    if (s->Concave[e * 6 + l - 1]) { if (!pos) in = 0; }
    else { if (neg) in = 0; }
  }
  return in;
}

typedef struct { double traj[3], len, alpha; } track_info;

/* ParticleThroughSideCheck3DFast, particle_intersection.f90:167-280 */
static int through_side_fast(const om_state *s, const double last[3], const track_info *ti, int l, int e, int TriNum) {
  const int32_t *nid = &s->ElemSideNodeID[(e * 6 + l - 1) * 4];
  const double Px = last[0], Py = last[1], Pz = last[2];
  const double Vx = ti->traj[0], Vy = ti->traj[1], Vz = ti->traj[2];
  double Ax[3], Ay[3], Az[3];
  Ax[0] = s->NodeCoords[(size_t)nid[0] * 3] - Px;
  Ay[0] = s->NodeCoords[(size_t)nid[0] * 3 + 1] - Py;
  Az[0] = s->NodeCoords[(size_t)nid[0] * 3 + 2] - Pz;
  for (int n = 2; n <= 3; n++) {
    int NodeID = n + TriNum - 1; /* 1-based side node */
    const double *p = &s->NodeCoords[(size_t)nid[NodeID - 1] * 3];
    Ax[n - 1] = p[0] - Px; Ay[n - 1] = p[1] - Py; Az[n - 1] = p[2] - Pz;
  }
  double dc[3][3];
  dc[0][0] = (Ay[0] * Vz - Az[0] * Vy) * Ax[2];
  dc[0][1] = (Az[0] * Vx - Ax[0] * Vz) * Ay[2];
  dc[0][2] = (Ax[0] * Vy - Ay[0] * Vx) * Az[2];
  dc[1][0] = (Ay[1] * Vz - Az[1] * Vy) * Ax[0];
  dc[1][1] = (Az[1] * Vx - Ax[1] * Vz) * Ay[0];
  dc[1][2] = (Ax[1] * Vy - Ay[1] * Vx) * Az[0];
  dc[2][0] = (Ay[2] * Vz - Az[2] * Vy) * Ax[1];
  dc[2][1] = (Az[2] * Vx - Ax[2] * Vz) * Ay[1];
  dc[2][2] = (Ax[2] * Vy - Ay[2] * Vx) * Az[1];
  double det[3], minc[3];
  for (int r = 0; r < 3; r++) {
    det[r] = dc[r][0] + dc[r][1] + dc[r][2];
    /* -epsMach*MINVAL(ABS(.),MASK=ABS(.)>0); MINVAL of an empty set is HUGE */
    double mn = DBL_MAX;
    for (int c = 0; c < 3; c++) if (fabs(dc[r][c]) > 0.0 && fabs(dc[r][c]) < mn) mn = fabs(dc[r][c]);
    minc[r] = -EPS_MACH * mn;
  }
  return (det[0] >= minc[0]) && (det[1] >= minc[1]) && (det[2] >= minc[2]);
}

/* ParticleThroughSideLastPosCheck, particle_intersection.f90:430-512 (regular side) */
static int through_side_last_pos(const om_state *s, const double last[3], int l, int e, int TriNum, double *detOut) {
  const int32_t *nid = &s->ElemSideNodeID[(e * 6 + l - 1) * 4];
  double NC[3][3]; /* NodeCoord(ind,iNode) -> NC[ind][iNode] */
  for (int c = 0; c < 3; c++) NC[c][0] = s->NodeCoords[(size_t)nid[0] * 3 + c];
  for (int iNode = 2; iNode <= 3; iNode++) {
    int NodeNum = iNode + TriNum - 1;
    for (int c = 0; c < 3; c++) NC[c][iNode - 1] = s->NodeCoords[(size_t)nid[NodeNum - 1] * 3 + c];
  }
  double Ax[3], Ay[3], Az[3];
  for (int ind = 0; ind < 3; ind++) {
    Ax[ind] = NC[0][ind] - last[0];
    Ay[ind] = NC[1][ind] - last[1];
    Az[ind] = NC[2][ind] - last[2];
  }
  double det = ((Ay[0] * Az[1] - Az[0] * Ay[1]) * Ax[2] +
                (Az[0] * Ax[1] - Ax[0] * Az[1]) * Ay[2] +
                (Ax[0] * Ay[1] - Ay[0] * Ax[1]) * Az[2]);
  *detOut = det;
  return !((det < 0) || (det != det));
}

/* IntersectionWithWall, particle_intersection.f90:79-164 */
static void intersection_with_wall(const om_state *s, const double last[3], track_info *ti, int l, int e, int TriNum) {
  const int32_t *nid = &s->ElemSideNodeID[(e * 6 + l - 1) * 4];
  const double *n0 = &s->NodeCoords[(size_t)nid[0] * 3];
  double xN = n0[0], yN = n0[1], zN = n0[2];
  int Node1 = TriNum + 1, Node2 = TriNum + 2;
  const double *p1 = &s->NodeCoords[(size_t)nid[Node1 - 1] * 3], *p2 = &s->NodeCoords[(size_t)nid[Node2 - 1] * 3];
  double V1[3] = {p1[0] - xN, p1[1] - yN, p1[2] - zN};
  double V2[3] = {p2[0] - xN, p2[1] - yN, p2[2] - zN};
  double nx = V1[1] * V2[2] - V1[2] * V2[1];
  double ny = V1[2] * V2[0] - V1[0] * V2[2];
  double nz = V1[0] * V2[1] - V1[1] * V2[0];
  double nVal = sqrt(nx * nx + ny * ny + nz * nz);
  nx = nx / nVal; ny = ny / nVal; nz = nz / nVal;
  double bx = last[0] - xN, by = last[1] - yN, bz = last[2] - zN;
  double ax = bx - nx * (bx * nx + by * ny + bz * nz);
  double ay = by - ny * (bx * nx + by * ny + bz * nz);
  double az = bz - nz * (bx * nx + by * ny + bz * nz);
  double dist = sqrt(((ay * bz - az * by) * (ay * bz - az * by) +
                      (az * bx - ax * bz) * (az * bx - ax * bz) +
                      (ax * by - ay * bx) * (ax * by - ay * bx)) /
                     (ax * ax + ay * ay + az * az));
  if (dist != dist) dist = sqrt(bx * bx + by * by + bz * bz);
  ti->alpha = ti->traj[0] * nx + ti->traj[1] * ny + ti->traj[2] * nz;
  if (fabs(ti->alpha) > 0.) ti->alpha = dist / ti->alpha;
}

/* SingleParticleTriaTracking3D, particle_triatracking.f90:137-484 (no mortars,
 * open and periodic particle boundaries). Returns 0, or -1 when the reference
 * would abort. */
static int tria_track_one(om_state *s, int i) {
  double *x = &s->PartState[(size_t)i * 6];
  double *last = &s->LastPartPos[(size_t)i * 3];
  int ElemID = s->lastElem[i];
  int done = 0;
  int DoneLastElem[6][4];
  memset(DoneLastElem, 0, sizeof DoneLastElem);
  int SideID = -1, LocalSide = 0, TriNum = 0;
  track_info ti = {{0, 0, 0}, 0, 0};
  int guard = 0;
  while (!done) {
    if (++guard > 10000) { set_err("tracking did not terminate"); return -1; }
    double det[6][2];
    if (inside_quad_3d(s, x, ElemID, det)) {
      s->elem[i] = ElemID;
      done = 1;
      break;
    }
    int nThrough = 0, LocSidesTemp[6] = {0}, TriNumTemp[6] = {0}, GlobSideTemp[6] = {0};
    for (int c = 0; c < 3; c++) ti.traj[c] = x[c] - last[c];
    ti.len = sqrt(ti.traj[0] * ti.traj[0] + ti.traj[1] * ti.traj[1] + ti.traj[2] * ti.traj[2]);
    if (fabs(ti.len) > 0.)
      for (int c = 0; c < 3; c++) ti.traj[c] = ti.traj[c] / ti.len;
    int first = ELEMINFO(ElemID, ELEM_FIRSTSIDEIND);
    for (int l = 1; l <= 6; l++) {
      int sd = first + l - 1;
      int ls = s->sideLocal[sd];
      if (ls <= 0) continue;
      for (int tn = 1; tn <= 2; tn++) {
        if (det[ls - 1][tn - 1] <= -0.0) {
          if (through_side_fast(s, last, &ti, ls, ElemID, tn)) {
            nThrough++;
            LocSidesTemp[nThrough - 1] = ls;
            TriNumTemp[nThrough - 1] = tn;
            GlobSideTemp[nThrough - 1] = sd + 1;
            SideID = sd;
            LocalSide = ls;
          }
        }
      }
    }
    TriNum = TriNumTemp[0];
    if (nThrough != 1) {
      if (nThrough == 0) {
#pragma omp atomic
        s->lost++;
        s->inside[i] = 0;
        done = 1;
        break;
      } else {
        int second = 0;
        double minRatio = 0;
#pragma omp atomic
        s->multiCross++;
        int tbFlag[6] = {0};
        for (int i2 = 0; i2 < nThrough; i2++) {
          int doCheck = 1;
          for (int is = 2; is <= 6; is++)
            if (DoneLastElem[is - 1][0] == ElemID + 1 && DoneLastElem[is - 1][3] == GlobSideTemp[i2] &&
                DoneLastElem[is - 1][2] == TriNumTemp[i2])
              doCheck = 0;
          if (!doCheck) continue;
          tbFlag[i2] = 1;
          double detM;
          if (through_side_last_pos(s, last, LocSidesTemp[i2], ElemID, TriNumTemp[i2], &detM)) {
            tbFlag[i2] = 3;
            double dPart = det[LocSidesTemp[i2] - 1][TriNumTemp[i2] - 1];
            if (detM == 0 && dPart == 0) continue;
            if (detM == 0 && minRatio == 0) {
              second++;
              SideID = GlobSideTemp[i2] - 1;
              LocalSide = LocSidesTemp[i2];
              TriNum = TriNumTemp[i2];
            } else {
              if (detM == 0) continue;
              double ratio = dPart / detM;
              /* planted mistake 1 (tests only): take the side crossed last */
              const int take = (s->plant == 1) ? (second == 0 || ratio > minRatio) : (ratio < minRatio);
              if (take) {
                minRatio = ratio;
                second++;
                SideID = GlobSideTemp[i2] - 1;
                LocalSide = LocSidesTemp[i2];
                TriNum = TriNumTemp[i2];
              }
            }
          }
        }
        if (s->tbCap) {
          int64_t slot;
#pragma omp atomic capture
          slot = s->tbN++;
          if (slot < s->tbCap) {
            double *r = s->tbLog + slot * OM_TB_REC;
            r[0] = i; r[1] = ElemID + 1; r[2] = nThrough; r[3] = second ? LocalSide : 0; r[4] = second ? TriNum : 0;
            for (int c = 0; c < 3; c++) { r[5 + c] = last[c]; r[8 + c] = x[c]; }
            for (int i2 = 0; i2 < nThrough; i2++) {
              r[11 + 3 * i2] = LocSidesTemp[i2]; r[12 + 3 * i2] = TriNumTemp[i2]; r[13 + 3 * i2] = tbFlag[i2];
            }
          }
        }
        if (second == 0) {
#pragma omp atomic
          s->lost++;
          s->inside[i] = 0;
          done = 1;
          break;
        }
      }
    }
    int bcid = SIDEINFO(SideID, SIDE_BCID);
    if (bcid > 0) {
      int OldElemID = ElemID;
      int BCType = s->partBC[bcid - 1];
      if (BCType != 1) intersection_with_wall(s, last, &ti, LocalSide, ElemID, TriNum);
      /* GetBoundaryInteraction, particle_boundary_condition.f90:35-221 */
      if (BCType == 1) {                 /* open: RemoveParticle */
        s->inside[i] = 0;
      } else if (BCType == 3) {          /* PeriodicBoundary, :224-284 */
        int PVID = s->bcAlpha[bcid - 1];
        int v = abs(PVID) - 1;
        for (int c = 0; c < 3; c++) last[c] = last[c] + ti.traj[c] * ti.alpha;
        for (int c = 0; c < 3; c++) last[c] = last[c] + copysign(fabs(s->PV[v][c]), (double)PVID);
        for (int c = 0; c < 3; c++) x[c] = last[c] + (ti.len - ti.alpha) * ti.traj[c];
        ti.len = ti.len - ti.alpha;
        ElemID = SIDEINFO(SideID, SIDE_NBELEMID) - 1;
      } else {
        set_err("particle boundary condition %d not supported by the oracle", BCType);
        return -1;
      }
      if (!s->inside[i]) done = 1;
      if (BCType == 2 || BCType == 10) memset(DoneLastElem, 0, sizeof DoneLastElem);
      else {
        for (int k = 4; k >= 0; k--) memcpy(DoneLastElem[k + 1], DoneLastElem[k], sizeof DoneLastElem[0]);
        DoneLastElem[0][0] = OldElemID + 1;
        DoneLastElem[0][1] = LocalSide;
        DoneLastElem[0][2] = TriNum;
        DoneLastElem[0][3] = SideID + 1;
      }
    } else {
      for (int k = 4; k >= 0; k--) memcpy(DoneLastElem[k + 1], DoneLastElem[k], sizeof DoneLastElem[0]);
      DoneLastElem[0][0] = ElemID + 1;
      DoneLastElem[0][1] = LocalSide;
      DoneLastElem[0][2] = TriNum;
      DoneLastElem[0][3] = SideID + 1;
      ElemID = SIDEINFO(SideID, SIDE_NBELEMID) - 1;
    }
    if (ElemID < 0 || ElemID >= s->nElems) { set_err("element not defined after tracking"); return -1; }
  }
  return 0;
}

static double max_abs3(const double v[3]) { return fmax(fabs(v[0]), fmax(fabs(v[1]), fabs(v[2]))); }

/* PeriodicMovement with FastPeriodic = F (particle_reftracking.f90:697-983):
 * a particle beyond the global bounds along the axis of a periodic vector is
 * moved back by that vector, LastPartPos with it.  Returns 1 if it moved. */
static int periodic_movement(const om_state *s, double *x, double *last) {
  int moved = 0;
  for (int dir = 0; dir < 3; dir++) {
    int v = -1;
    for (int q = 0; q < s->nPV; q++)
      if (s->dirPV[q] == dir) { v = q; break; }
    if (v < 0) continue;
    if (x[dir] > s->gmax[dir]) {
      const double sg = s->PV[v][dir] > 0 ? -1.0 : 1.0;
      for (int c = 0; c < 3; c++) { x[c] = x[c] + sg * s->PV[v][c]; last[c] = last[c] + sg * s->PV[v][c]; }
      moved = 1;
    }
    if (x[dir] < s->gmin[dir]) {
      const double sg = s->PV[v][dir] > 0 ? 1.0 : -1.0;
      for (int c = 0; c < 3; c++) { x[c] = x[c] + sg * s->PV[v][c]; last[c] = last[c] + sg * s->PV[v][c]; }
      moved = 1;
    }
  }
  return moved;
}

/* FIBGM cell of a point as ParticleRefTracking computes it (:180-186) */
static size_t bgm_cell(const om_state *s, const double x[3]) {
  int b3[3];
  for (int c = 0; c < 3; c++) {
    int v = (int)floor((x[c] - s->bmin[c]) / s->bdelta[c]);
    if (v < 0) v = 0;
    v = v + 1;
    if (v > s->nb[c]) v = s->nb[c];
    b3[c] = v - 1;
  }
  return ((size_t)b3[2] * s->nb[1] + b3[1]) * s->nb[0] + b3[0];
}

static int single_point_to_element(const om_state *s, const double x[3]);
static int is_bc_elem(const om_state *s, int e) { return s->e2bOff && s->e2bOff[e + 1] > s->e2bOff[e]; }

/* ParticleRefTracking, particle_reftracking.f90:30-414.  When every boundary
 * is periodic (CartesianPeriodic = T) periodic boundaries are crossed by
 * PeriodicMovement; otherwise particles of elements with boundary sides in
 * range go through ParticleBCTracking (open walls and periodic sides by
 * intersecting the trajectory with the boundary faces). */
static int ref_track_one(om_state *s, int i) {
  double *x = &s->PartState[(size_t)i * 6], *last = &s->LastPartPos[(size_t)i * 3];
  double *ref = &s->PartPosRef[(size_t)i * 3];
  const int LastElemID = s->lastElem[i];
  int ab;
  if (is_bc_elem(s, LastElemID)) {
    int ElemID = LastElemID, done = 0, moved = 0;
    double len0 = 0.;
    if (bc_tracking(s, i, &len0, &ElemID, 1, &done, &moved) != 0) return -1;
    if (done) return 0;
    position_in_ref_elem(s, x, ref, ElemID, 0, 0, &ab);
    if (max_abs3(ref) < 1.0) { s->elem[i] = ElemID; return 0; }
  } else {
    if (s->cartPeriodic && s->nPV > 0)
      for (int g = 0; g < 4; g++)
        if (!periodic_movement(s, x, last)) break;
    position_in_ref_elem(s, x, ref, LastElemID, 0, 0, &ab);
    if (max_abs3(ref) < 1.0) { s->elem[i] = LastElemID; return 0; }
  }
  const int oldElemID = LastElemID;
  const size_t b = bgm_cell(s, x);
  const int n = s->binOff[b + 1] - s->binOff[b];
  double dist[512];
  int list[512];
  if (n > 512) { set_err("too many elements in a background cell"); return -1; }
  if (n > 1) {
    for (int q = 0; q < n; q++) {
      int e = s->binElems[s->binOff[b] + q];
      list[q] = e;
      if (e == oldElemID) dist[q] = -HUGE_VAL;
      else {
        double d0 = x[0] - s->bary[e * 3], d1 = x[1] - s->bary[e * 3 + 1], d2 = x[2] - s->bary[e * 3 + 2];
        dist[q] = d0 * d0 + d1 * d1 + d2 * d2;
        if (dist[q] > s->radius[e] * s->radius[e]) dist[q] = -HUGE_VAL;
      }
    }
    for (int a = 1; a < n; a++) { /* InsertionSort (utils), stable */
      double dv = dist[a]; int lv = list[a]; int j = a - 1;
      while (j >= 0 && dist[j] > dv) { dist[j + 1] = dist[j]; list[j + 1] = list[j]; j--; }
      dist[j + 1] = dv; list[j + 1] = lv;
    }
  } else if (n == 1) {
    dist[0] = 0.0;
    list[0] = s->binElems[s->binOff[b]];
  }
  double oldXi[3] = {ref[0], ref[1], ref[2]}, newXi[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL};
  int newElemID = -1;
  for (int q = 0; q < n; q++) {
    if (dist[q] == -HUGE_VAL) continue;
    const int e = list[q];
    position_in_ref_elem(s, x, ref, e, 0, 0, &ab);
    if (max_abs3(ref) < 1.0) { s->elem[i] = e; return 0; }
    if (max_abs3(ref) < max_abs3(newXi)) {
      for (int c = 0; c < 3; c++) newXi[c] = ref[c];
      newElemID = e;
    }
  }
  int TestElem;
  if (max_abs3(oldXi) < max_abs3(newXi)) {
    for (int c = 0; c < 3; c++) ref[c] = oldXi[c];
    TestElem = oldElemID;
  } else {
    for (int c = 0; c < 3; c++) ref[c] = newXi[c];
    TestElem = newElemID;
  }
  s->elem[i] = TestElem;
  if (TestElem >= 0 && max_abs3(ref) <= s->epsOneCell[TestElem]) return 0;
  if (TestElem < 0 || !is_bc_elem(s, TestElem)) {   /* Tolerance Issue with internal element: abort */
    set_err("RefMapping: particle %d not inside of element %d", i + 1, TestElem + 1);
    return -1;
  }
  /* boundary element (:325-405): FallBackFaceIntersection (NGeo > 1), then
   * ParticleBCTracking again, then the element or LocateParticleInElement */
  if (s->NGeo > 1) {
    const int rc = fallback_face_intersection(s, i, TestElem);
    if (rc < 0) return -1;
    if (rc == 1) {   /* no face hit: LocateParticleInElement; not found aborts */
      const int e = single_point_to_element(s, x);
      if (e < 0) { s->inside[i] = 0; s->lost++; return 0; }
      s->elem[i] = e;
      position_in_ref_elem(s, x, ref, e, 0, 0, &ab);
      return 0;
    }
  }
  int done = 0, moved = 0;
  double len0 = 0.;
  if (bc_tracking(s, i, &len0, &TestElem, 1, &done, &moved) != 0) return -1;
  if (done) return 0;
  position_in_ref_elem(s, x, ref, TestElem, 0, 0, &ab);
  if (max_abs3(ref) > s->epsOneCell[TestElem]) {   /* Tolerance Issue with BC element, relocating */
    const int e = single_point_to_element(s, x);
    if (e < 0) { s->inside[i] = 0; s->lost++; return 0; }
    s->elem[i] = e;
    position_in_ref_elem(s, x, ref, e, 0, 0, &ab);
  } else {
    s->elem[i] = TestElem;
  }
  return 0;
}

static int track_one(om_state *s, int i) {
  return s->par.tracking == OM_TRACK_REFMAPPING ? ref_track_one(s, i) : tria_track_one(s, i);
}

int om_perform_tracking(om_state *s) {
  int err = 0;
  /* ParticleTriaTracking (particle_triatracking.f90:63-134) / ParticleRefTracking */
  for (int i = 0; i < s->n; i++)
    if (s->inside[i] && s->lastElem[i] >= 0)
      if (track_one(s, i) != 0) err = -1;
  return err;
}

/* SinglePointToElement with TriaTracking, particle_localization.f90:81-203 */
static int single_point_to_element(const om_state *s, const double x[3]) {
  int b3[3];
  for (int c = 0; c < 3; c++) {
    int v = (int)ceil((x[c] - s->bmin[c]) / s->bdelta[c]);
    if (v > s->nb[c]) v = s->nb[c];
    if (v < 1) v = 1;
    b3[c] = v - 1;
  }
  size_t b = ((size_t)b3[2] * s->nb[1] + b3[1]) * s->nb[0] + b3[0];
  int n = s->binOff[b + 1] - s->binOff[b];
  if (n <= 0) return -1;
  double *dist = xcalloc(n, sizeof(double));
  int *list = xcalloc(n, sizeof(int));
  int any = 0;
  for (int q = 0; q < n; q++) {
    int e = s->binElems[s->binOff[b] + q];
    double d[3] = {x[0] - s->bary[e * 3], x[1] - s->bary[e * 3 + 1], x[2] - s->bary[e * 3 + 2]};
    double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    dist[q] = d2 <= s->radius[e] * s->radius[e] ? d2 : -1.0;
    if (dist[q] != -1.0) any = 1;
    list[q] = e;
  }
  int found = -1;
  if (any) {
    for (int a = 1; a < n; a++) { /* InsertionSort (utils) by distance */
      double dv = dist[a]; int lv = list[a]; int j = a - 1;
      while (j >= 0 && dist[j] > dv) { dist[j + 1] = dist[j]; list[j + 1] = list[j]; j--; }
      dist[j + 1] = dv; list[j + 1] = lv;
    }
    for (int q = 0; q < n; q++) {
      if (dist[q] == -1.0) continue;
      if (s->par.tracking == OM_TRACK_REFMAPPING) {
        double xi[3];
        int ab;
        position_in_ref_elem(s, x, xi, list[q], 0, 0, &ab);
        if (max_abs3(xi) <= s->epsOneCell[list[q]]) { found = list[q]; break; }
      } else {
        double det[6][2];
        if (inside_quad_3d(s, x, list[q], det)) { found = list[q]; break; }
      }
    }
  }
  free(dist); free(list);
  return found;
}

int om_locate_particles(om_state *s) {
  int missing = 0;
  for (int i = 0; i < s->n; i++)
    if (s->elem[i] < 0) {
      s->elem[i] = single_point_to_element(s, &s->PartState[(size_t)i * 6]);
      if (s->elem[i] < 0) { s->inside[i] = 0; missing++; }
      else if (s->par.tracking == OM_TRACK_REFMAPPING) {
        int ab;
        position_in_ref_elem(s, &s->PartState[(size_t)i * 6], &s->PartPosRef[(size_t)i * 3], s->elem[i], 0, 0, &ab);
      }
    }
  return missing;
}

/* ================================================================ interpolation and push */
static void interpolate_one(om_state *s, int i) {
  double *F = &s->Field[(size_t)i * 6];
  const double *x = &s->PartState[(size_t)i * 6];
  int e = s->elem[i];
  /* GetExternalFieldAtParticle, pic_interpolation_tools.f90:54-86 */
  for (int c = 0; c < 6; c++) F[c] = s->par.external_field[c];
  /* GetInterpolatedFieldPartPos, pic_interpolation_tools.f90:198-254 */
  double xi[3];
  int ab, ok = 1;
  if (s->par.tracking == OM_TRACK_REFMAPPING) /* PartPosRef from the tracking (:242-244) */
    for (int c = 0; c < 3; c++) xi[c] = s->PartPosRef[(size_t)i * 3 + c];
  else
    ok = position_in_ref_elem(s, x, xi, e, 0, 1, &ab);
  double U[6] = {0, 0, 0, 0, 0, 0};
  if (!ok && s->par.deposition == OM_DEPO_CVWM) em_field_dw(s, e, x, U);
  else eval_field_at_ref_pos(s, xi, e, U); /* GetEMField, :399-455 */
  for (int c = 0; c < 6; c++) F[c] = F[c] + U[c];
}

int om_interpolate_field_to_particle(om_state *s) {
  /* InterpolateFieldToParticle, pic_interpolation.f90:248-322 (particle loop) */
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int i = 0; i < s->n; i++) {
    if (!s->inside[i]) continue;
    if (fabs(s->par.charge[s->species[i]]) > 0.0) interpolate_one(s, i);
  }
  return 0;
}

/* Boris-Leapfrog push, timedisc_TimeStepPoissonByBorisLeapfrog.f90:128-198 */
static void push_one(om_state *s, int i, double dt) {
  const double c2i = c2_inv();
  double *st = &s->PartState[(size_t)i * 6];
  const double *F = &s->Field[(size_t)i * 6];
  const int sp = s->species[i];
  const double q = s->par.charge[sp], m = s->par.mass[sp];
  const int push = fabs(q) > 0.0; /* isPushParticle, particle_tools.f90:950-976 */
  double dtFrac = dt;
  if (s->isNew[i]) {
    if (push) {
      /* CalcPartRHSSingleParticle -> PartRHS_NR, particle_rhs.f90:203-253 */
      double qmt = q / m;
      double Pt[3] = {F[0] * qmt, F[1] * qmt, F[2] * qmt};
      for (int c = 0; c < 3; c++) st[3 + c] = st[3 + c] - Pt[c] * dt * 0.5;
    }
    s->isNew[i] = 0;
  }
  if (push) {
    double c_1 = (q * dtFrac) / (m * 2.);
    double *v = &st[3];
    double gamma = 1. / sqrt(1 - ((v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * c2i));
    double vmo[3] = {v[0] * gamma, v[1] * gamma, v[2] * gamma};
    double vm[3] = {vmo[0] + c_1 * F[0], vmo[1] + c_1 * F[1], vmo[2] + c_1 * F[2]};
    double gm = sqrt(1 + (vm[0] * vm[0] + vm[1] * vm[1] + vm[2] * vm[2]) * c2i);
    const double *B = &F[3];
    double Bn = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]); /* VECNORM3D */
    double ub[3];                                               /* UNITVECTOR, globals.f90:922-934 */
    double invL = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    if (fabs(invL) > 0.0) { invL = 1. / invL; ub[0] = B[0] * invL; ub[1] = B[1] * invL; ub[2] = B[2] * invL; }
    else { ub[0] = ub[1] = ub[2] = 0.0; }
    double tf = tan(c_1 / gm * Bn);
    double t[3] = {tf * ub[0], tf * ub[1], tf * ub[2]};
    double vp[3] = {vm[0] + (vm[1] * t[2] - vm[2] * t[1]), vm[1] + (vm[2] * t[0] - vm[0] * t[2]),
                    vm[2] + (vm[0] * t[1] - vm[1] * t[0])};
    double f2 = 2.0 / (1. + (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
    double cr[3] = {vp[1] * t[2] - vp[2] * t[1], vp[2] * t[0] - vp[0] * t[2], vp[0] * t[1] - vp[1] * t[0]};
    double vpl[3] = {vm[0] + f2 * cr[0], vm[1] + f2 * cr[1], vm[2] + f2 * cr[2]};
    double vn1[3] = {vpl[0] + c_1 * F[0], vpl[1] + c_1 * F[1], vpl[2] + c_1 * F[2]};
    double den = sqrt(1 + (vn1[0] * vn1[0] + vn1[1] * vn1[1] + vn1[2] * vn1[2]) * c2i);
    v[0] = vn1[0] / den; v[1] = vn1[1] / den; v[2] = vn1[2] / den;
  }
  st[0] = st[0] + st[3] * dtFrac;
  st[1] = st[1] + st[4] * dtFrac;
  st[2] = st[2] + st[5] * dtFrac;
}

/* LastPartPos <- PartState and LastGlobalElemID <- GlobalElemID (stepper
 * lines 109-110), then the push loop (lines 128-198). */
int om_push_boris_leapfrog(om_state *s, double dt) {
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (int i = 0; i < s->n; i++) {
    s->LastPartPos[i * 3 + 0] = s->PartState[i * 6 + 0];
    s->LastPartPos[i * 3 + 1] = s->PartState[i * 6 + 1];
    s->LastPartPos[i * 3 + 2] = s->PartState[i * 6 + 2];
    s->lastElem[i] = s->elem[i];
    if (s->inside[i]) push_one(s, i, dt);
  }
  return 0;
}

/* ================================================================ deposition */
/* DepositionMethod_CVWM, pic_depo_method.f90:373-744 */
static int deposit_cvwm(om_state *s) {
  const int nu = s->nUnique, N = s->N, n1 = N + 1, nd = s->nDof;
  double *NS = s->NodeSource;
  for (size_t q = 0; q < (size_t)nu * 4; q++) NS[q] = 0.0;
  for (int i = 0; i < s->n; i++) {
    if (!s->inside[i]) continue;
    const int sp = s->species[i];
    if (!(fabs(s->par.charge[sp]) > 0.0)) continue;
    const double Charge = s->par.charge[sp] * s->par.mpf[sp];
    const double *st = &s->PartState[(size_t)i * 6];
    const int e = s->elem[i];
    double xi[3];
    int ab;
    int ok = position_in_ref_elem(s, st, xi, e, 1, 1, &ab);
    double TS[4] = {st[3] * Charge, st[4] * Charge, st[5] * Charge, Charge};
    double w[8];
    int NodeID[8];
    if (ok) {
      double a1 = 0.5 * (xi[0] + 1.0), a2 = 0.5 * (xi[1] + 1.0), a3 = 0.5 * (xi[2] + 1.0);
      w[0] = (1 - a1) * (1 - a2) * (1 - a3);
      w[1] = (a1) * (1 - a2) * (1 - a3);
      w[2] = (a1) * (a2) * (1 - a3);
      w[3] = (1 - a1) * (a2) * (1 - a3);
      w[4] = (1 - a1) * (1 - a2) * (a3);
      w[5] = (a1) * (1 - a2) * (a3);
      w[6] = (a1) * (a2) * (a3);
      w[7] = (1 - a1) * (a2) * (a3);
      for (int c = 0; c < 8; c++) NodeID[c] = s->NodeInfo[s->ElemNodeID[e * 8 + c]];
      for (int c = 0; c < 8; c++) {
        for (int k = 0; k < 4; k++) NS[NodeID[c] * 4 + k] = NS[NodeID[c] * 4 + k] + (TS[k] * w[c]);
        for (int t = s->perOff[NodeID[c]]; t < s->perOff[NodeID[c] + 1]; t++)
          for (int k = 0; k < 4; k++) NS[s->perNodes[t] * 4 + k] = NS[s->perNodes[t] * 4 + k] + (TS[k] * w[c]);
      }
    } else {
      for (int c = 0; c < 8; c++) w[c] = 0.0;
      int nid[8];
      for (int c = 0; c < 8; c++) nid[c] = s->ElemNodeID[e * 8 + c];
      for (int c = 0; c < 8; c++) {
        const double *p = &s->NodeCoords[(size_t)nid[c] * 3];
        double d[3] = {p[0] - st[0], p[1] - st[1], p[2] - st[2]};
        double norm = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (norm > 0.) w[c] = 1. / norm;
        else { for (int q = 0; q < 8; q++) w[q] = 0.; w[c] = 1.0; break; }
      }
      double sum = 0.0;
      for (int c = 0; c < 8; c++) sum = sum + w[c];
      for (int c = 0; c < 8; c++) {
        int u = s->NodeInfo[nid[c]];
        for (int k = 0; k < 4; k++) NS[u * 4 + k] = NS[u * 4 + k] + w[c] / sum * TS[k];
        for (int t = s->perOff[u]; t < s->perOff[u + 1]; t++)
          for (int k = 0; k < 4; k++) NS[s->perNodes[t] * 4 + k] = NS[s->perNodes[t] * 4 + k] + w[c] / sum * TS[k];
      }
    }
  }
  for (int u = 0; u < nu; u++)
    if (s->NodeVolume[u] > 0.)
      for (int k = 0; k < 4; k++) NS[u * 4 + k] = NS[u * 4 + k] / s->NodeVolume[u];
  /* node -> DOF, :704-733 with CellVolWeight%Fac = (xGP+1)/2 (pic_depo.f90:271-276) */
  double Fac[MAXN];
  for (int q = 0; q < n1; q++) Fac[q] = (s->xGP[q] + 1.0) / 2.0;
  for (int e = 0; e < s->nElems; e++) {
    int NodeID[8];
    for (int c = 0; c < 8; c++) NodeID[c] = s->NodeInfo[s->ElemNodeID[e * 8 + c]];
    for (int kk = 0; kk < n1; kk++) for (int ll = 0; ll < n1; ll++) for (int mm = 0; mm < n1; mm++) {
      double a1 = Fac[kk], a2 = Fac[ll], a3 = Fac[mm];
      double *ps = &s->PS[((size_t)e * nd + (mm * n1 + ll) * n1 + kk) * 4];
      for (int k = 0; k < 4; k++)
        ps[k] = NS[NodeID[0] * 4 + k] * (1 - a1) * (1 - a2) * (1 - a3) +
                NS[NodeID[1] * 4 + k] * (a1) * (1 - a2) * (1 - a3) +
                NS[NodeID[2] * 4 + k] * (a1) * (a2) * (1 - a3) +
                NS[NodeID[3] * 4 + k] * (1 - a1) * (a2) * (1 - a3) +
                NS[NodeID[4] * 4 + k] * (1 - a1) * (1 - a2) * (a3) +
                NS[NodeID[5] * 4 + k] * (a1) * (1 - a2) * (a3) +
                NS[NodeID[6] * 4 + k] * (a1) * (a2) * (a3) +
                NS[NodeID[7] * 4 + k] * (1 - a1) * (a2) * (a3);
    }
  }
  return 0;
}

/* depoChargeOnDOFsSF, pic_depo_shapefunction_tools.f90:292-422, accumulating into PS */
static void depo_charge_on_dofs_sf(const om_state *s, const double pos[3], const double Fac[4],
                                   double *PS, uint8_t *done) {
  int kmin[3], kmax[3];
  for (int c = 0; c < 3; c++) {
    kmax[c] = (int)ceil((pos[c] + s->r_sf - s->bmin[c]) / s->bdelta[c]);
    kmin[c] = (int)floor((pos[c] - s->r_sf - s->bmin[c]) / s->bdelta[c] + 1);
    if (kmax[c] > s->nb[c]) kmax[c] = s->nb[c];
    if (kmin[c] < 1) kmin[c] = 1;
  }
  const int n1 = s->N + 1, nd = s->nDof;
  int nUsed = 0;
  int usedList[4096];
  for (int kk = kmin[0]; kk <= kmax[0]; kk++)
    for (int ll = kmin[1]; ll <= kmax[1]; ll++)
      for (int mm = kmin[2]; mm <= kmax[2]; mm++) {
        size_t b = ((size_t)(mm - 1) * s->nb[1] + (ll - 1)) * s->nb[0] + (kk - 1);
        for (int q = s->binOff[b]; q < s->binOff[b + 1]; q++) {
          int e = s->binElems[q];
          if (done[e]) continue;
          double d[3] = {pos[0] - s->bary[e * 3], pos[1] - s->bary[e * 3 + 1], pos[2] - s->bary[e * 3 + 2]};
          if (sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) > (s->r_sf + s->radius[e])) continue;
          const double *xg = &s->xGPe[(size_t)e * nd * 3];
          for (int m = 0; m < n1; m++) for (int l = 0; l < n1; l++) for (int k = 0; k < n1; k++) {
            int r = (m * n1 + l) * n1 + k;
            double v0 = pos[0] - xg[r * 3], v1 = pos[1] - xg[r * 3 + 1], v2 = pos[2] - xg[r * 3 + 2];
            double radius2 = v0 * v0 + v1 * v1 + v2 * v2; /* SFRadius2, :1029-1058 */
            if (radius2 <= s->r2_sf) {
              double S = 1. - s->r2_sf_inv * radius2;
              double S1 = S * S;
              for (int ex = 3; ex <= s->par.sf_alpha; ex++) S1 = S * S1;
              double *ps = &PS[((size_t)e * nd + r) * 4];
              for (int c = 0; c < 4; c++) ps[c] = ps[c] + S1 * Fac[c]; /* UpdatePartSource, :911-967 */
            }
          }
          done[e] = 1;
          if (nUsed < 4096) usedList[nUsed++] = e;
        }
      }
  for (int u = 0; u < nUsed; u++) done[usedList[u]] = 0;
  if (nUsed >= 4096) memset(done, 0, s->nElems);
}

/* DepositionMethod_SF -> calcSfSource (shape_function, SourceSize 4),
 * pic_depo_method.f90:851-1003, pic_depo_shapefunction_tools.f90:30-164 */
static void sf_source_one(const om_state *s, int i, double *PS, uint8_t *done) {
  const int sp = s->species[i];
  const double Charge = s->par.charge[sp] * s->par.mpf[sp];
  const double *st = &s->PartState[(size_t)i * 6];
  double Fac[4] = {st[3] * Charge, st[4] * Charge, st[5] * Charge, Charge};
  if (s->nCases > 1) {
    for (int c = 0; c < 4; c++) Fac[c] = Fac[c] * s->w_sf;
    for (int ic = 0; ic < s->nCases; ic++) {
      double p[3]; /* GetPartPosShifted, :1061-1124 (3D) */
      for (int c = 0; c < 3; c++)
        p[c] = st[c] + s->caseMat[ic][0] * s->PV[0][c] + s->caseMat[ic][1] * s->PV[1][c] + s->caseMat[ic][2] * s->PV[2][c];
      depo_charge_on_dofs_sf(s, p, Fac, PS, done);
    }
  } else {
    double F2[4];
    for (int c = 0; c < 4; c++) F2[c] = Fac[c] * s->w_sf;
    depo_charge_on_dofs_sf(s, st, F2, PS, done);
  }
}

static int deposit_sf(om_state *s) {
  const size_t nps = (size_t)s->nElems * s->nDof * 4;
  if (s->nthreads <= 1) {
    uint8_t *done = xcalloc(s->nElems, 1);
    for (int i = 0; i < s->n; i++) {
      if (!s->inside[i] || !(fabs(s->par.charge[s->species[i]]) > 0.0)) continue;
      sf_source_one(s, i, s->PS, done);
    }
    free(done);
    return 0;
  }
#ifdef _OPENMP
  /* timing build: a private PS per thread (kept between calls), summed
   * afterwards; zeroing and summing the copies is timed on its own
   * (tPrivReduce), since a domain-decomposed code has no such copies */
  int nt = s->nthreads;
  if (s->nPriv != nt) {
    for (int t = 0; t < s->nPriv; t++) free(s->priv[t]);
    free(s->priv);
    s->priv = xcalloc(nt, sizeof(double *));
    s->nPriv = nt;
  }
  double **priv = s->priv;
  double tz = 0.0, tr = 0.0;
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    double t0 = omp_get_wtime();
    if (!priv[t]) priv[t] = xcalloc(nps, sizeof(double));
    else memset(priv[t], 0, nps * sizeof(double));
#pragma omp barrier
    if (t == 0) tz = omp_get_wtime() - t0;
    uint8_t *done = xcalloc(s->nElems, 1);
#pragma omp for schedule(static)
    for (int i = 0; i < s->n; i++) {
      if (!s->inside[i] || !(fabs(s->par.charge[s->species[i]]) > 0.0)) continue;
      sf_source_one(s, i, priv[t], done);
    }
    free(done);
    double t1 = omp_get_wtime();
#pragma omp for schedule(static)
    for (size_t q = 0; q < nps; q++) {
      double acc = 0.0;
      for (int u = 0; u < nt; u++) acc += priv[u][q];
      s->PS[q] = acc;
    }
    if (t == 0) tr = omp_get_wtime() - t1;
  }
  s->tPrivReduce += tz + tr;
#endif
  return 0;
}

/* Deposition, pic_depo.f90:944-1018 */
int om_deposition(om_state *s) {
  memset(s->PS, 0, (size_t)s->nElems * s->nDof * 4 * sizeof(double));
  if (s->par.deposition == OM_DEPO_CVWM) return deposit_cvwm(s);
  if (s->par.deposition == OM_DEPO_SF) return deposit_sf(s);
  return 0;
}

/* ================================================================ 508 step */
int om_time_step(om_state *s, double dt, int32_t nsteps) {
  for (int it = 0; it < nsteps; it++) {
    if (om_deposition(s) != 0) return -1;
    /* (HDG field solve: outside the stage, E frozen) */
    om_interpolate_field_to_particle(s);
    om_push_boris_leapfrog(s, dt);
    int err = 0;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) reduction(| : err)
    for (int i = 0; i < s->n; i++)
      if (s->inside[i] && s->lastElem[i] >= 0)
        if (track_one(s, i) != 0) err |= 1;
    if (err) return -1;
  }
  return 0;
}

/* ================================================================ self-consistent field
 * A fixed linear field operator closes the loop for the adjoint workload:
 *   rhoBar_e = sum_ijk w_i w_j w_k rho_e,ijk / 8       (reference-space mean)
 *   E_node   = G rhoBar                                  (unique corner nodes)
 *   E_dof    = trilinear interpolation of the 8 corner nodes at the Gauss points
 * G is supplied by the caller (e.g. a spectral periodic Poisson solve). */
int om_sc_set_operator(om_state *s, const double *G) {
  const size_t rows = (size_t)3 * s->nUnique, cols = (size_t)s->nElems;
  free(s->Gop); free(s->rhoBar); free(s->Enode);
  s->Gop = xcalloc(rows * cols, sizeof(double));
  memcpy(s->Gop, G, rows * cols * sizeof(double));
  s->rhoBar = xcalloc(cols, sizeof(double));
  s->Enode = xcalloc(rows, sizeof(double));
  return 0;
}

int om_sc_field_from_source(om_state *s) {
  if (!s->Gop) { set_err("no field operator set"); return -1; }
  const int ne = s->nElems, n1 = s->N + 1, nd = s->nDof;
  for (int e = 0; e < ne; e++) {
    double acc = 0.0;
    for (int k = 0; k < n1; k++) for (int j = 0; j < n1; j++) for (int i = 0; i < n1; i++) {
      int r = (k * n1 + j) * n1 + i;
      acc = acc + s->wGP[i] * s->wGP[j] * s->wGP[k] * s->PS[((size_t)e * nd + r) * 4 + 3];
    }
    s->rhoBar[e] = acc / 8.0;
  }
  const size_t rows = (size_t)3 * s->nUnique;
#pragma omp parallel for num_threads(s->nthreads) schedule(static)
  for (size_t r = 0; r < rows; r++) {
    const double *g = &s->Gop[r * ne];
    double acc = 0.0;
    for (int e = 0; e < ne; e++) acc = acc + g[e] * s->rhoBar[e];
    s->Enode[r] = acc;
  }
  for (int e = 0; e < ne; e++) {
    int uid[8];
    for (int c = 0; c < 8; c++) uid[c] = s->NodeInfo[s->ElemNodeID[e * 8 + c]];
    for (int k = 0; k < n1; k++) for (int j = 0; j < n1; j++) for (int i = 0; i < n1; i++) {
      int r = (k * n1 + j) * n1 + i;
      double a1 = (s->xGP[i] + 1.0) / 2.0, a2 = (s->xGP[j] + 1.0) / 2.0, a3 = (s->xGP[k] + 1.0) / 2.0;
      double w[8] = {(1 - a1) * (1 - a2) * (1 - a3), a1 * (1 - a2) * (1 - a3), a1 * a2 * (1 - a3), (1 - a1) * a2 * (1 - a3),
                     (1 - a1) * (1 - a2) * a3, a1 * (1 - a2) * a3, a1 * a2 * a3, (1 - a1) * a2 * a3};
      for (int c = 0; c < 3; c++) {
        double v = 0.0;
        for (int q = 0; q < 8; q++) v = v + w[q] * s->Enode[(size_t)uid[q] * 3 + c];
        s->E[((size_t)e * nd + r) * 3 + c] = v;
      }
    }
  }
  return 0;
}

/* K self-consistent steps (deposit, field, interpolate, push, track), then the
 * loss: field energy of the field of the final deposition. */
int om_sc_forward(om_state *s, double dt, int32_t K, double *loss) {
  for (int it = 0; it < K; it++) {
    if (om_deposition(s) != 0) return -1;
    if (om_sc_field_from_source(s) != 0) return -1;
    om_interpolate_field_to_particle(s);
    om_push_boris_leapfrog(s, dt);
    int err = 0;
#pragma omp parallel for num_threads(s->nthreads) schedule(static) reduction(| : err)
    for (int i = 0; i < s->n; i++)
      if (s->inside[i] && s->lastElem[i] >= 0)
        if (track_one(s, i) != 0) err |= 1;
    if (err) return -1;
  }
  if (om_deposition(s) != 0) return -1;
  if (om_sc_field_from_source(s) != 0) return -1;
  *loss = om_field_energy(s);
  return 0;
}

/* ================================================================ analysis */
/* CalcDepositedCharge, pic_analyze.f90:136-210 */
double om_deposited_charge(const om_state *s) {
  const int n1 = s->N + 1;
  double ch = 0.0;
  for (int e = 0; e < s->nElems; e++)
    for (int k = 0; k < n1; k++) for (int j = 0; j < n1; j++) for (int i = 0; i < n1; i++) {
      int r = (k * n1 + j) * n1 + i;
      double wijk = s->wGP[i] * s->wGP[j] * s->wGP[k];
      double J = 1. / s->sJ[(size_t)e * s->nDof + r];
      ch = ch + wijk * s->PS[((size_t)e * s->nDof + r) * 4 + 3] * J;
    }
  return ch;
}
double om_particle_charge(const om_state *s) {
  double ch = 0.0;
  for (int i = 0; i < s->n; i++)
    if (s->inside[i]) ch = ch + s->par.charge[s->species[i]] * s->par.mpf[s->species[i]];
  return ch;
}
/* CalcPotentialEnergy, analyzefield_tools.f90:99-228 (WEl) */
double om_field_energy(const om_state *s) {
  const double eps0 = 8.8541878176e-12;
  const int n1 = s->N + 1;
  double W = 0.0;
  for (int e = 0; e < s->nElems; e++) {
    double We = 0.0;
    for (int k = 0; k < n1; k++) for (int j = 0; j < n1; j++) for (int i = 0; i < n1; i++) {
      int r = (k * n1 + j) * n1 + i;
      const double *E = &s->E[((size_t)e * s->nDof + r) * 3];
      double Eabs = E[0] * E[0] + E[1] * E[1] + E[2] * E[2];
      We = We + s->wGP[i] * s->wGP[j] * s->wGP[k] / s->sJ[(size_t)e * s->nDof + r] * Eabs;
    }
    W = W + We;
  }
  return W * eps0 * 0.5;
}
