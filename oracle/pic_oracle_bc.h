/* RefMapping boundary tracking of the oracle (test infrastructure: the CPU
 * checker of the GPU stage, never the product).  Included once by
 * pic_oracle.c, after the mesh state and the Newton routines.
 *
 * Restates ParticleBCTracking (src/particles/tracking/particle_reftracking.f90:417-694),
 * GetBoundaryInteraction's RefMapping branch and PeriodicBoundary
 * (src/particles/boundary/particle_boundary_condition.f90:35-284), the face
 * intersections it calls (src/particles/tracking/particle_intersection.f90:
 * ComputePlanarRectIntersection 515-684, ComputePlanarCurvedIntersection
 * 687-852, ComputeCurvedIntersection 1243-1632, BezierClipRecursive
 * 1635-1747, BezierNewton 1750-1988, InsideBoundingBox 1991-2038,
 * BoundingBoxIntersection 2040-2109, FlatBoundingBoxIntersection 2112-2200,
 * ComputeBezierIntersectionPoint 2276-2456, CheckXiClip / CheckEtaClip
 * 2459-3242, calcLineNormVec3 3285-3363, CalcSminSmax2 3366-3451),
 * FallBackFaceIntersection (particle_reftracking.f90:986-1099), and the side
 * tables they read: BezierControlPoints3D (CalcBezierControlPoints,
 * particle_mesh_tools.f90:1650-1760), side slabs
 * (GetSideSlabNormalsAndIntervals, particle_surfaces.f90:619-816), SideType /
 * SideNormVec / SideDistance (IdentifyElemAndSideType,
 * particle_mesh_tools.f90:848-1160), BaseVectors and BCSideMetrics
 * (particle_mesh_build.f90:1760-1930), ElemToBCSides and SideBCMetrics
 * (BuildBCElemDistance, particle_mesh_build.f90:700-1080).
 *
 * Used when not every boundary is periodic (CartesianPeriodic = F): open
 * walls and periodic sides are then met by intersecting the trajectory with
 * the element's boundary faces.  Bilinear sides (deformed linear elements
 * with open walls) are refused at set-up. */

enum { ST_PLANAR_RECT = 0, ST_PLANAR_NONRECT = 1, ST_BILINEAR = 2, ST_PLANAR_CURVED = 3, ST_CURVED = 4 };
enum { LS_ZETA_MINUS = 1, LS_ETA_MINUS = 2, LS_XI_PLUS = 3, LS_ETA_PLUS = 4, LS_XI_MINUS = 5, LS_ZETA_PLUS = 6 };

/* defaults of the intersection parameters, particle_surfaces.f90:133-158 */
#define BZ_CLIP_TOL 1e-8
#define BZ_NEWTON_TOL2 (1e-4 * 1e-4)
#define BZ_NEWTON_MAXITER 100
#define BZ_SPLIT_LIMIT (2.0 * 0.6)
#define BZ_CLIP_MAXITER 100
#define BZ_CLIP_HIT (1.0 + 100.0 * BZ_CLIP_TOL)
#define BZ_EPSTOL (100.0 * DBL_EPSILON)
#define BZ_NEWTON_ANGLE 1.570796326
#define BZ_G1MAX 3            /* NGeo <= 2 */
#define BZ_MAXINTER 6         /* BezierClipMaxIntersec = 2 (NGeo + 1) */

static int almost_zero(double x) { return fabs(x) <= 2.22e-16; } /* ALMOSTZERO, piclas.h:83 */

static void cross3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static void unit3(double v[3]) { /* UNITVECTOR */
  const double l = sqrt(dot3(v, v));
  v[0] = v[0] / l; v[1] = v[1] / l; v[2] = v[2] / l;
}
static void crossnorm3(const double a[3], const double b[3], double c[3]) { /* CROSSNORM */
  cross3(a, b, c);
  unit3(c);
}

/* Bezier control point (p,q) component c of BC side b */
#define BZP(b, p, q, c) (s->bz[((((size_t)(b) * g1 + (p)) * g1) + (q)) * 3 + (c)])

static double choose_d(int n, int k) {
  double r = 1.0;
  for (int q = 1; q <= k; q++) r = r * (double)(n - q + 1) / (double)q;
  return r;
}
static double ipow(double x, int k) {
  double r = 1.0;
  for (int q = 0; q < k; q++) r = r * x;
  return r;
}
/* FacNchooseK(p,l) = C(p,l) / 2^p (BuildBezierVdm, basis.f90:312-324) */
static double fac_nck(int p, int l) { return (1.0 / (double)(1 << p)) * choose_d(p, l); }

/* sVdm_Bezier = INVERSE(Vdm_Bezier), Vdm_Bezier(i,j) = B_j(xi_i) on the CL
 * nodes (BuildBezierVdm, basis.f90:275-358; INVERSE, mathtools.f90:44-102) */
static void bezier_svdm(int ng, const double *xi, double sV[BZ_G1MAX][BZ_G1MAX]) {
  const int n = ng + 1;
  double A[BZ_G1MAX][BZ_G1MAX];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      A[i][j] = (1.0 / (double)(1 << ng)) * choose_d(ng, j) * ipow(xi[i] + 1.0, j) * ipow(1.0 - xi[i], ng - j);
  if (n == 2) {
    const double sdet = 1.0 / (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
    sV[0][0] = (A[1][1]) * sdet; sV[0][1] = (-A[0][1]) * sdet;
    sV[1][0] = (-A[1][0]) * sdet; sV[1][1] = (A[0][0]) * sdet;
  } else {
    const double sdet = 1.0 / ((A[0][0] * A[1][1] - A[0][1] * A[1][0]) * A[2][2] +
                               (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * A[2][0] +
                               (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * A[2][1]);
    sV[0][0] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) * sdet;
    sV[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * sdet;
    sV[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * sdet;
    sV[1][0] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) * sdet;
    sV[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * sdet;
    sV[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * sdet;
    sV[2][0] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) * sdet;
    sV[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * sdet;
    sV[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * sdet;
  }
}

/* CL-point slice of local side ls of an element, tmp[a][b] (the two free
 * volume indices in order), CalcBezierControlPoints :1718-1731 */
static void side_slice(const double *X, int g1, int ls, double tmp[BZ_G1MAX][BZ_G1MAX][3]) {
  const int ng = g1 - 1;
  for (int a = 0; a < g1; a++)
    for (int b = 0; b < g1; b++) {
      int i = 0, j = 0, k = 0;
      switch (ls) {
        case LS_XI_MINUS: i = 0; j = a; k = b; break;
        case LS_XI_PLUS: i = ng; j = a; k = b; break;
        case LS_ETA_MINUS: i = a; j = 0; k = b; break;
        case LS_ETA_PLUS: i = a; j = ng; k = b; break;
        case LS_ZETA_MINUS: i = a; j = b; k = 0; break;
        default: i = a; j = b; k = ng; break;
      }
      for (int c = 0; c < 3; c++) tmp[a][b][c] = X[((k * g1 + j) * g1 + i) * 3 + c];
    }
}

/* CGNS_SideToVol2 (mesh/mapping.f90:320-356) */
static void cgns_side_to_vol2(int ng, int p, int q, int ls, int pq[2]) {
  switch (ls) {
    case LS_XI_MINUS: pq[0] = q; pq[1] = p; break;
    case LS_ETA_PLUS: pq[0] = ng - p; pq[1] = q; break;
    case LS_ZETA_MINUS: pq[0] = q; pq[1] = p; break;
    default: pq[0] = p; pq[1] = q; break;
  }
}

/* Bezier control points of local side ls of an element from its CL points:
 * ChangeBasis2D with sVdm_Bezier, then CGNS side orientation */
static void side_bezier(const double *X, int ng, int ls, const double sV[BZ_G1MAX][BZ_G1MAX],
                        double out[BZ_G1MAX][BZ_G1MAX][3]) {
  const int g1 = ng + 1;
  double tmp[BZ_G1MAX][BZ_G1MAX][3], buf[BZ_G1MAX][BZ_G1MAX][3], tmp2[BZ_G1MAX][BZ_G1MAX][3];
  side_slice(X, g1, ls, tmp);
  /* ChangeBasis2D (interpolation/changebasis.f90): first index, then second */
  for (int jn = 0; jn < g1; jn++)
    for (int io = 0; io < g1; io++) {
      double v[3] = {0, 0, 0};
      for (int in = 0; in < g1; in++)
        for (int c = 0; c < 3; c++) v[c] = v[c] + sV[io][in] * tmp[in][jn][c];
      for (int c = 0; c < 3; c++) buf[io][jn][c] = v[c];
    }
  for (int jo = 0; jo < g1; jo++)
    for (int io = 0; io < g1; io++) {
      double v[3] = {0, 0, 0};
      for (int jn = 0; jn < g1; jn++)
        for (int c = 0; c < 3; c++) v[c] = v[c] + sV[jo][jn] * buf[io][jn][c];
      for (int c = 0; c < 3; c++) tmp2[io][jo][c] = v[c];
    }
  for (int q = 0; q < g1; q++)
    for (int p = 0; p < g1; p++) {
      int pq[2];
      cgns_side_to_vol2(ng, p, q, ls, pq);
      for (int c = 0; c < 3; c++) out[p][q][c] = tmp2[pq[0]][pq[1]][c];
    }
}

/* GetSideSlabNormalsAndIntervals (particle_surfaces.f90:619-816).  Returns 0,
 * or -1 for a degenerated side (the reference aborts). */
static int side_slab(int ng, double B[BZ_G1MAX][BZ_G1MAX][3], double N[3][3], double I[6], uint8_t *empty) {
  double n2[3];
  for (int c = 0; c < 3; c++) N[0][c] = B[ng][0][c] - B[0][0][c] + B[ng][ng][c] - B[0][ng][c];
  if (N[0][0] == 0 && N[0][1] == 0 && N[0][2] == 0) return -1;
  const double l0 = sqrt(dot3(N[0], N[0]));
  for (int c = 0; c < 3; c++) N[0][c] = N[0][c] / l0;
  for (int c = 0; c < 3; c++) n2[c] = B[0][ng][c] - B[0][0][c] + B[ng][ng][c] - B[ng][0][c];
  crossnorm3(N[0], n2, N[1]);
  crossnorm3(N[1], N[0], N[2]);
  for (int c = 0; c < 6; c++) I[c] = 0.0;
  for (int q = 0; q <= ng; q++)
    for (int p = 0; p <= ng; p++) {
      if (p == 0 && q == 0) continue;
      double d[3] = {B[p][q][0] - B[0][0][0], B[p][q][1] - B[0][0][1], B[p][q][2] - B[0][0][2]};
      for (int a = 0; a < 3; a++) {
        const double sp = dot3(d, N[a]);
        if (sp < 0.) I[2 * a] = fmin(I[2 * a], sp);
        else if (sp > 0.) I[2 * a + 1] = fmax(I[2 * a + 1], sp);
      }
    }
  const double dx = fabs(I[1] - I[0]), dy = fabs(I[3] - I[2]), dz = fabs(I[5] - I[4]);
  double h = dy;
  const double w = fmin(dx, dz), l = fmax(dx, dz);
  const double dMax = fmax(fmax(h, l), w);
  if (l / dMax < 1.0e-6 || w / dMax < 1.0e-6) return -1;
  if (h / dMax < 1.0e-6) { I[2] = 0.; I[3] = 0.; h = 0.; }
  *empty = almost_zero(h / sqrt(l * l + w * w + h * h));
  return 0;
}

/* DeCasteljauInterpolation (basis.f90:429-476) at (xi, eta) in [-1,1]^2 */
static void de_casteljau(int ng, const double *bz /* [g1][g1][3] */, const double xin[2], double out[3]) {
  const int g1 = ng + 1;
  double R[BZ_G1MAX][BZ_G1MAX][3];
  for (int p = 0; p < g1; p++) for (int q = 0; q < g1; q++) for (int c = 0; c < 3; c++) R[p][q][c] = bz[(p * g1 + q) * 3 + c];
  const double Xi = 0.5 * (xin[0] + 1.), Eta = 0.5 * (xin[1] + 1.);
  const double MXi = 1.0 - Xi, MEta = 1.0 - Eta;
  int l = ng - 1;
  for (int it = 1; it <= ng; it++) {
    for (int q = 0; q <= l; q++)
      for (int p = 0; p <= l; p++)
        for (int c = 0; c < 3; c++)
          R[p][q][c] = MXi * R[p][q][c] * MEta + MXi * R[p][q + 1][c] * Eta + Xi * R[p + 1][q][c] * MEta +
                       Xi * R[p + 1][q + 1][c] * Eta;
    l = l - 1;
  }
  for (int c = 0; c < 3; c++) out[c] = R[0][0][c];
}

/* Side tables of every boundary side and the element -> boundary side lists.
 * BC_halo_eps (BuildBCElemDistance, MPI branch :748-817) is taken as twice the
 * largest ElemRadiusNGeo: every side a particle can reach within one step of
 * at most that length is listed (a longer step is an error). */
static int build_bc_tables(om_state *s) {
  const int ne = s->nElems, ng = s->NGeo, g1 = ng + 1, nGP = s->nGP;
  s->cartPeriodic = 1;
  int nb = 0;
  for (int sd = 0; sd < s->nSides; sd++) {
    const int bc = SIDEINFO(sd, SIDE_BCID);
    if (bc <= 0) continue;
    nb++;
    if (s->partBC[bc - 1] != 3) s->cartPeriodic = 0;
  }
  if (s->nPV == 0) s->cartPeriodic = 0;
  if (s->cartPeriodic || nb == 0) return 0;
  if (ng > 2) { set_err("boundary tracking: NGeo > 2 unsupported"); return -1; }
  s->nBC = nb;
  s->bcSide = xcalloc(nb, sizeof(int32_t));
  s->sideToBC = xcalloc(s->nSides, sizeof(int32_t));
  s->bz = xcalloc((size_t)nb * g1 * g1 * 3, sizeof(double));
  s->sideType = xcalloc(nb, 1);
  s->sideNorm = xcalloc((size_t)nb * 3, sizeof(double));
  s->sideDist = xcalloc(nb, sizeof(double));
  s->slabN = xcalloc((size_t)nb * 9, sizeof(double));
  s->slabI = xcalloc((size_t)nb * 6, sizeof(double));
  s->bbEmpty = xcalloc(nb, 1);
  s->baseV = xcalloc((size_t)nb * 9, sizeof(double));
  s->bcOrigin = xcalloc((size_t)nb * 3, sizeof(double));
  s->bcRadius = xcalloc(nb, sizeof(double));
  double sV[BZ_G1MAX][BZ_G1MAX];
  bezier_svdm(ng, s->xiCL, sV);
  double xi1[2] = {-1.0, 1.0}, V1[MAXN * 2];
  vandermonde(1, xi1, ng, s->xiCL, V1);
  int b = 0;
  for (int sd = 0; sd < s->nSides; sd++) {
    s->sideToBC[sd] = -1;
    if (SIDEINFO(sd, SIDE_BCID) <= 0) continue;
    s->sideToBC[sd] = b;
    s->bcSide[b] = sd;
    const int e = s->sideElem[sd], ls = s->sideLocal[sd];
    const double *X = &s->XCL[(size_t)e * nGP * 3];
    double B[BZ_G1MAX][BZ_G1MAX][3];
    side_bezier(X, ng, ls, sV, B);
    for (int p = 0; p < g1; p++) for (int q = 0; q < g1; q++) for (int c = 0; c < 3; c++) BZP(b, p, q, c) = B[p][q][c];
    double N[3][3], I[6];
    uint8_t empty = 0;
    if (side_slab(ng, B, N, I, &empty) != 0) { set_err("boundary side %d: degenerated side slab", sd + 1); return -1; }
    for (int a = 0; a < 3; a++) for (int c = 0; c < 3; c++) s->slabN[(size_t)b * 9 + a * 3 + c] = N[a][c];
    for (int c = 0; c < 6; c++) s->slabI[(size_t)b * 6 + c] = I[c];
    s->bbEmpty[b] = empty;
    /* IdentifyElemAndSideType (particle_mesh_tools.f90:976-1155) */
    const int flip = SIDEINFO(sd, SIDE_ID) > 0 ? 0 : SIDEINFO(sd, SIDE_FLIP) % 10;
    double *nv = &s->sideNorm[b * 3];
    double v1[3], v2[3], v3[3];
    for (int c = 0; c < 3; c++) {
      v1[c] = -B[0][0][c] + B[ng][0][c] - B[0][ng][c] + B[ng][ng][c];
      v2[c] = -B[0][0][c] - B[ng][0][c] + B[0][ng][c] + B[ng][ng][c];
    }
    int type;
    int planar = 0, rectTolSigned = 0;
    if (!s->curved[e]) {
      if (empty) planar = 1;
      else type = ST_BILINEAR;
    } else {
      /* curved side: its CL points differ from the trilinear map's */
      double X1[8 * 3], Xn[MAXN * MAXN * MAXN * 3];
      for (int t = 0; t < 8; t++) {
        int i = (t & 1) * ng, j = ((t >> 1) & 1) * ng, k = ((t >> 2) & 1) * ng;
        for (int c = 0; c < 3; c++) X1[t * 3 + c] = X[((k * g1 + j) * g1 + i) * 3 + c];
      }
      change_basis_3d(1, ng, V1, X1, Xn);
      double So[BZ_G1MAX][BZ_G1MAX][3], Sn[BZ_G1MAX][BZ_G1MAX][3];
      side_slice(X, g1, ls, So);
      side_slice(Xn, g1, ls, Sn);
      int curvedSide = 0;
      for (int a = 0; a < g1 && !curvedSide; a++)
        for (int bb = 0; bb < g1 && !curvedSide; bb++)
          if (fabs(Sn[a][bb][0] - So[a][bb][0]) > 1e-14 || fabs(Sn[a][bb][1] - So[a][bb][1]) > 1e-14 ||
              fabs(Sn[a][bb][2] - So[a][bb][2]) > 1e-14)
            curvedSide = 1;
      if (curvedSide) {
        if (empty) {
          type = ST_PLANAR_CURVED;
          crossnorm3(v1, v2, nv);
          double c4[3];
          for (int c = 0; c < 3; c++) c4[c] = 0.25 * (B[0][0][c] + B[ng][0][c] + B[0][ng][c] + B[ng][ng][c]);
          double d[3] = {c4[0] - s->bary[e * 3], c4[1] - s->bary[e * 3 + 1], c4[2] - s->bary[e * 3 + 2]};
          if (flip == 0) { if (dot3(d, nv) < 0) for (int c = 0; c < 3; c++) nv[c] = -nv[c]; }
          else if (dot3(d, nv) > 0) for (int c = 0; c < 3; c++) nv[c] = -nv[c];
          s->sideDist[b] = dot3(c4, nv);
        } else {
          type = ST_CURVED;
        }
      } else if (empty) {
        planar = 1;
        rectTolSigned = 1;
      } else {
        type = ST_BILINEAR;
      }
    }
    if (planar) {
      crossnorm3(v1, v2, nv);
      double c4[3];
      for (int c = 0; c < 3; c++) c4[c] = 0.25 * (B[0][0][c] + B[ng][0][c] + B[0][ng][c] + B[ng][ng][c]);
      double d[3] = {c4[0] - s->bary[e * 3], c4[1] - s->bary[e * 3 + 1], c4[2] - s->bary[e * 3 + 2]};
      if (flip == 0) { if (dot3(d, nv) < 0) for (int c = 0; c < 3; c++) nv[c] = -nv[c]; }
      else if (dot3(d, nv) > 0) for (int c = 0; c < 3; c++) nv[c] = -nv[c];
      s->sideDist[b] = dot3(c4, nv);
      int rect = 1;
      double w1[3], w2[3], w3[3];
      for (int c = 0; c < 3; c++) {
        w1[c] = B[0][ng][c] - B[0][0][c];
        w2[c] = B[ng][0][c] - B[0][0][c];
        w3[c] = B[ng][ng][c] - B[0][ng][c];
      }
      unit3(w1); unit3(w2); unit3(w3);
      /* flat sides of curved elements test the signed product against 1e-14 */
#define NOT_ORTHO(a, b) (rectTolSigned ? (dot3(a, b) > 1e-14) : !almost_zero(dot3(a, b)))
      if (NOT_ORTHO(w1, w2)) rect = 0;
      if (NOT_ORTHO(w1, w3)) rect = 0;
      if (rect) {
        for (int c = 0; c < 3; c++) w1[c] = B[ng][ng][c] - B[ng][0][c];
        unit3(w1);
        if (NOT_ORTHO(w1, w2)) rect = 0;
        if (NOT_ORTHO(w1, w3)) rect = 0;
      }
#undef NOT_ORTHO
      type = rect ? ST_PLANAR_RECT : ST_PLANAR_NONRECT;
    }
    if (type == ST_BILINEAR || type == ST_PLANAR_NONRECT) {
      set_err("boundary tracking: bilinear or non-rectangular planar boundary side %d unsupported", sd + 1);
      return -1;
    }
    (void)v3;
    s->sideType[b] = (uint8_t)type;
    /* BaseVectors0..2 (particle_mesh_build.f90:1917-1924) */
    double *bv = &s->baseV[(size_t)b * 9];
    for (int c = 0; c < 3; c++) {
      bv[c] = (+B[0][0][c] + B[ng][0][c] + B[0][ng][c] + B[ng][ng][c]);
      bv[3 + c] = (-B[0][0][c] + B[ng][0][c] - B[0][ng][c] + B[ng][ng][c]);
      bv[6 + c] = (-B[0][0][c] - B[ng][0][c] + B[0][ng][c] + B[ng][ng][c]);
    }
    /* BCSideMetrics: origin at xi = eta = 0 and the largest distance of a control point */
    const double x0[2] = {0.0, 0.0};
    de_casteljau(ng, &BZP(b, 0, 0, 0), x0, &s->bcOrigin[b * 3]);
    double rmax = 0.;
    for (int q = 0; q <= ng; q++)
      for (int p = 0; p <= ng; p++) {
        double d[3] = {B[p][q][0] - s->bcOrigin[b * 3], B[p][q][1] - s->bcOrigin[b * 3 + 1],
                       B[p][q][2] - s->bcOrigin[b * 3 + 2]};
        rmax = fmax(rmax, dot3(d, d));
      }
    s->bcRadius[b] = sqrt(rmax);
    b++;
  }
  /* ElemToBCSides / SideBCMetrics (non-fullMesh branch :876-905, :996-1024) */
  double rmaxE = 0.0;
  for (int e = 0; e < ne; e++) rmaxE = fmax(rmaxE, s->radius[e]);
  s->bcHalo = 2.0 * rmaxE;
  s->e2bOff = xcalloc(ne + 1, sizeof(int32_t));
  int32_t *tmpList = xcalloc(nb + 6, sizeof(int32_t));
  size_t cap = 1024, cnt = 0;
  s->e2bSide = xcalloc(cap, sizeof(int32_t));
  s->e2bDist = xcalloc(cap, sizeof(double));
  for (int e = 0; e < ne; e++) {
    int nl = 0;
    const int first = ELEMINFO(e, ELEM_FIRSTSIDEIND);
    for (int l = 0; l < 6; l++) {
      const int sd = first + l;
      if (SIDEINFO(sd, SIDE_BCID) > 0) tmpList[nl++] = s->sideToBC[sd];
    }
    for (int q = 0; q < nb; q++) {
      const int be = s->sideElem[s->bcSide[q]];
      if (be == e) continue;
      double d[3] = {s->bary[e * 3] - s->bary[be * 3], s->bary[e * 3 + 1] - s->bary[be * 3 + 1],
                     s->bary[e * 3 + 2] - s->bary[be * 3 + 2]};
      if (sqrt(dot3(d, d)) > s->bcHalo + s->radius[e] + s->radius[be]) continue;
      double d2[3] = {s->bary[e * 3] - s->bcOrigin[q * 3], s->bary[e * 3 + 1] - s->bcOrigin[q * 3 + 1],
                      s->bary[e * 3 + 2] - s->bcOrigin[q * 3 + 2]};
      if (sqrt(dot3(d2, d2)) <= s->bcHalo + s->radius[e] + s->bcRadius[q]) tmpList[nl++] = q;
    }
    if (cnt + nl > cap) {
      while (cnt + nl > cap) cap *= 2;
      s->e2bSide = realloc(s->e2bSide, cap * sizeof(int32_t));
      s->e2bDist = realloc(s->e2bDist, cap * sizeof(double));
    }
    s->e2bOff[e] = (int32_t)cnt;
    for (int q = 0; q < nl; q++) {
      const int bs = tmpList[q];
      double d[3] = {s->bary[e * 3] - s->bcOrigin[bs * 3], s->bary[e * 3 + 1] - s->bcOrigin[bs * 3 + 1],
                     s->bary[e * 3 + 2] - s->bcOrigin[bs * 3 + 2]};
      s->e2bSide[cnt + q] = bs;
      s->e2bDist[cnt + q] = sqrt(dot3(d, d)) - s->radius[e] - s->bcRadius[bs];
    }
    /* InsertionSort by distance (stable) */
    for (int a = 1; a < nl; a++) {
      const double dv = s->e2bDist[cnt + a];
      const int32_t sv = s->e2bSide[cnt + a];
      int j = a - 1;
      while (j >= 0 && s->e2bDist[cnt + j] > dv) {
        s->e2bDist[cnt + j + 1] = s->e2bDist[cnt + j];
        s->e2bSide[cnt + j + 1] = s->e2bSide[cnt + j];
        j--;
      }
      s->e2bDist[cnt + j + 1] = dv;
      s->e2bSide[cnt + j + 1] = sv;
    }
    cnt += nl;
  }
  s->e2bOff[ne] = (int32_t)cnt;
  free(tmpList);
  return 0;
}

/* ---------------------------------------------------------------- intersections */
typedef struct {
  double traj[3], len, alpha, xi, eta;  /* TrackInfo */
} bc_track_info;

typedef struct {         /* module work arrays of the Bezier clipping */
  double locAlpha[BZ_MAXINTER], locXi[BZ_MAXINTER], locEta[BZ_MAXINTER];
  double XiArray[BZ_CLIP_MAXITER + 1][2], EtaArray[BZ_CLIP_MAXITER + 1][2];
  double localTol;       /* BezierClipLocalTol */
  int err;
} bz_work;

/* EvaluateBezierPolynomialAndGradient (particle_surfaces.f90:474-616), no
 * precomputed derivative: point and/or gradient of an iSize-dimensional patch */
static void bezier_eval(int ng, int iSize, const double *B /* [p][q][iSize] */, const double X[2], double *P,
                        double G[2][3]) {
  double xiup[BZ_G1MAX], etaup[BZ_G1MAX], xidown[BZ_G1MAX], etadown[BZ_G1MAX], xit[BZ_G1MAX], etat[BZ_G1MAX];
  const int g1 = ng + 1, M = ng - 1;
  const double MinusXi = 1.0 - X[0], PlusXi = 1.0 + X[0], MinusEta = 1.0 - X[1], PlusEta = 1.0 + X[1];
  xiup[0] = 1.0; etaup[0] = 1.0; xidown[ng] = 1.0; etadown[ng] = 1.0;
  for (int p = 1; p <= ng; p++) {
    xiup[p] = xiup[p - 1] * PlusXi;
    etaup[p] = etaup[p - 1] * PlusEta;
    xidown[ng - p] = xidown[ng - p + 1] * MinusXi;
    etadown[ng - p] = etadown[ng - p + 1] * MinusEta;
  }
  for (int p = 0; p <= ng; p++) { xit[p] = xiup[p] * xidown[p]; etat[p] = etaup[p] * etadown[p]; }
  if (P) {
    for (int c = 0; c < iSize; c++) P[c] = 0.;
    for (int q = 0; q <= ng; q++)
      for (int p = 0; p <= ng; p++)
        for (int c = 0; c < iSize; c++)
          P[c] = P[c] + B[(p * g1 + q) * iSize + c] * fac_nck(ng, p) * xit[p] * fac_nck(ng, q) * etat[q];
  }
  if (G) {
    for (int d = 0; d < 2; d++) for (int c = 0; c < iSize; c++) G[d][c] = 0.;
    for (int q = 0; q <= ng; q++)
      for (int p = 0; p <= M; p++)
        for (int c = 0; c < iSize; c++) {
          G[0][c] = G[0][c] + (B[((p + 1) * g1 + q) * iSize + c] - B[(p * g1 + q) * iSize + c]) * fac_nck(M, p) *
                                  xiup[p] * xidown[p + 1] * fac_nck(ng, q) * etat[q];
          G[1][c] = G[1][c] + (B[(q * g1 + p + 1) * iSize + c] - B[(q * g1 + p) * iSize + c]) * fac_nck(ng, q) *
                                  xit[q] * fac_nck(M, p) * etaup[p] * etadown[p + 1];
        }
    for (int d = 0; d < 2; d++) for (int c = 0; c < iSize; c++) G[d][c] = 0.5 * (double)ng * G[d][c];
  }
}

/* CalcNormAndTangBezier (particle_surfaces.f90:443-471) */
static void bezier_normal(const om_state *s, int b, double xi, double eta, double n[3]) {
  const int g1 = s->NGeo + 1;
  double G[2][3];
  const double X[2] = {xi, eta};
  bezier_eval(s->NGeo, 3, &s->bz[(size_t)b * g1 * g1 * 3], X, NULL, G);
  crossnorm3(G[0], G[1], n);
}

static int inside_bbox(const om_state *s, int b, const double pos[3]) {
  const double *N = &s->slabN[(size_t)b * 9], *I = &s->slabI[(size_t)b * 6];
  const int g1 = s->NGeo + 1;
  double P[3];
  for (int c = 0; c < 3; c++) P[c] = pos[c] - s->bz[(size_t)b * g1 * g1 * 3 + c];
  const double eps = 100. * DBL_EPSILON;
  const double y = dot3(P, &N[3]);
  if (y < I[2] - eps || y > I[3] + eps) return 0;
  const double x = dot3(P, &N[0]);
  if (x < I[0] - eps || x > I[1] + eps) return 0;
  const double z = dot3(P, &N[6]);
  if (z < I[4] - eps || z > I[5] + eps) return 0;
  return 1;
}

static void slab_alpha(const om_state *s, int b, int i, const double last[3], const double traj[3], double dnk,
                       double *amin, double *amax) {
  const double *N = &s->slabN[(size_t)b * 9 + i * 3], *I = &s->slabI[(size_t)b * 6];
  const int g1 = s->NGeo + 1;
  const double *B0 = &s->bz[(size_t)b * g1 * g1 * 3];
  double d[3] = {B0[0] - last[0], B0[1] - last[1], B0[2] - last[2]};
  const double base = dot3(d, N);
  (void)traj;
  if (dnk < 0.) {
    *amin = (base + I[2 * i + 1]) / dnk;
    *amax = (base + I[2 * i]) / dnk;
  } else {
    *amin = (base + I[2 * i]) / dnk;
    *amax = (base + I[2 * i + 1]) / dnk;
  }
}

/* BoundingBoxIntersection (particle_intersection.f90:2040-2109) */
static int bbox_intersection(const om_state *s, int b, const double last[3], const double traj[3], double len) {
  const double epsM = DBL_EPSILON;
  double lo[3], hi[3];
  for (int i = 0; i < 3; i++) {
    double dnk = dot3(traj, &s->slabN[(size_t)b * 9 + i * 3]);
    if (fabs(dnk) < 100. * epsM) dnk = 100. * epsM;
    slab_alpha(s, b, i, last, traj, dnk, &lo[i], &hi[i]);
  }
  const double maxv = fmax(fmax(lo[0], lo[1]), lo[2]);
  const double minv = fmin(fmin(hi[0], hi[1]), hi[2]);
  if (maxv <= minv) return (maxv < len + 100 * epsM) && (maxv + 100 * epsM > 0.);
  return 0;
}

/* FlatBoundingBoxIntersection (particle_intersection.f90:2112-2200) */
static int flat_bbox_intersection(const om_state *s, int b, const double last[3], const double traj[3], double len) {
  const double epsM = DBL_EPSILON;
  double lo[3], hi[3];
  for (int i = 0; i < 3; i += 2) {
    const double dnk = dot3(traj, &s->slabN[(size_t)b * 9 + i * 3]);
    if (fabs(dnk) < 100. * epsM) { lo[i] = -HUGE_VAL; hi[i] = HUGE_VAL; }
    else slab_alpha(s, b, i, last, traj, dnk, &lo[i], &hi[i]);
  }
  const double maxv = fmax(lo[0], lo[2]);
  const double minv = fmin(hi[0], hi[2]);
  if (maxv <= minv) {
    if (maxv < 0 && minv > 0) return 1;
    return (maxv < len + 100. * epsM) && (maxv + 100. * epsM > 0.);
  }
  return 0;
}

/* ComputeBezierIntersectionPoint (particle_intersection.f90:2276-2456) */
static void bezier_intersection_point(const om_state *s, int b, bz_work *W, int nXiClip, int nEtaClip,
                                      int *nInter, const double last[3], const double traj[3], double len) {
  const int ng = s->NGeo, g1 = ng + 1;
  double Xi, Eta;
  if (nXiClip == 0) Xi = 0.;
  else {
    Xi = 0.5 * (W->XiArray[nXiClip][0] + W->XiArray[nXiClip][1]);
    for (int k = nXiClip - 1; k >= 1; k--) Xi = W->XiArray[k][0] + 0.5 * (Xi + 1) * (W->XiArray[k][1] - W->XiArray[k][0]);
  }
  if (nEtaClip == 0) Eta = 0.;
  else {
    Eta = 0.5 * (W->EtaArray[nEtaClip][0] + W->EtaArray[nEtaClip][1]);
    for (int k = nEtaClip - 1; k >= 1; k--) Eta = W->EtaArray[k][0] + 0.5 * (Eta + 1) * (W->EtaArray[k][1] - W->EtaArray[k][0]);
  }
  const double tmpXi = Xi, tmpEta = Eta;
  Xi = 0.5 * (Xi + 1);
  Eta = 0.5 * (Eta + 1);
  if (fabs(Eta) > BZ_CLIP_HIT || fabs(Xi) > BZ_CLIP_HIT) return;
  const double MXi = 1.0 - Xi, MEta = 1.0 - Eta;
  double R[BZ_G1MAX][BZ_G1MAX][3];
  for (int p = 0; p < g1; p++) for (int q = 0; q < g1; q++) for (int c = 0; c < 3; c++) R[p][q][c] = BZP(b, p, q, c);
  int l = ng - 1;
  for (int it = 1; it <= ng; it++) {
    for (int q = 0; q <= l; q++)
      for (int p = 0; p <= l; p++)
        for (int c = 0; c < 3; c++)
          R[p][q][c] = MXi * R[p][q][c] * MEta + MXi * R[p][q + 1][c] * Eta + Xi * R[p + 1][q][c] * MEta +
                       Xi * R[p + 1][q + 1][c] * Eta;
    l = l - 1;
  }
  double iv[3] = {R[0][0][0] - last[0], R[0][0][1] - last[1], R[0][0][2] - last[2]};
  const double alpha = dot3(iv, traj);
  const double alphaNorm = alpha / len;
  if (alphaNorm <= 1.0 && alphaNorm > -BZ_EPSTOL) {
    if (*nInter >= BZ_MAXINTER) { W->err = 1; return; }   /* the reference STOPs */
    int isNew = 1;
    for (int k = 0; k < *nInter && *nInter > 0; k++) {
      if (fabs(W->locXi[k] - tmpXi) > 0.01) continue;
      if (fabs(W->locEta[k] - tmpEta) > 0.01) continue;
      isNew = 0;
      break;
    }
    if (*nInter == 0 || isNew) {
      W->locAlpha[*nInter] = alpha;
      W->locXi[*nInter] = tmpXi;
      W->locEta[*nInter] = tmpEta;
      (*nInter)++;
    }
  }
}

/* calcLineNormVec3 (particle_intersection.f90:3285-3363) */
static void line_norm_vec(int ng, double B[BZ_G1MAX][BZ_G1MAX][2], double L[2][2], bz_work *W) {
  double LXi[2], LEta[2];
  for (int c = 0; c < 2; c++) {
    LXi[c] = (B[0][ng][c] - B[0][0][c]) + (B[ng][ng][c] - B[ng][0][c]);
    LEta[c] = (B[ng][0][c] - B[0][0][c]) + (B[ng][ng][c] - B[0][ng][c]);
  }
  double len = sqrt(LXi[0] * LXi[0] + LXi[1] * LXi[1]);
  if (len == 0.) { W->err = 2; return; }   /* the reference STOPs */
  LXi[0] = LXi[0] / len; LXi[1] = LXi[1] / len;
  len = sqrt(LEta[0] * LEta[0] + LEta[1] * LEta[1]);
  if (len == 0.) { W->err = 2; return; }
  LEta[0] = LEta[0] / len; LEta[1] = LEta[1] / len;
  /* LineNormVec(:,1) = (-LXi(2), LXi(1)), LineNormVec(:,2) = (-Leta(2), Leta(1)) */
  L[0][0] = -LXi[1]; L[1][0] = LXi[0];
  L[0][1] = -LEta[1]; L[1][1] = LEta[0];
}

/* CalcSminSmax2 (particle_intersection.f90:3366-3451), Xi_NGeo equidistant */
static void calc_smin_smax(int ng, double mm[2][BZ_G1MAX], double *Smin, double *Smax, int iter) {
  double X[BZ_G1MAX];
  for (int i = 0; i <= ng; i++) X[i] = 2.0 * (double)i / (double)ng - 1.0;
  *Smin = 1.5;
  *Smax = -1.5;
  for (int p = 0; p <= ng; p++)
    for (int q = 0; q <= ng; q++) {
      if (p == q) continue;
      if (mm[1][p] * mm[1][q] <= 0.) {
        const double m = (mm[1][q] - mm[1][p]) / (X[q] - X[p]);
        if (m < 0.) *Smax = fmax(*Smax, X[q] - mm[1][q] / m);
        if (m > 0.) *Smin = fmin(*Smin, X[q] - mm[1][q] / m);
      }
      if (mm[0][p] * mm[0][q] <= 0.) {
        const double m = (mm[0][q] - mm[0][p]) / (X[q] - X[p]);
        if (m > 0.) *Smax = fmax(*Smax, X[q] - mm[0][q] / m);
        if (m < 0.) *Smin = fmin(*Smin, X[q] - mm[0][q] / m);
      }
    }
  if (mm[0][0] * mm[1][0] < 0.) *Smin = -1.;
  if (mm[0][ng] * mm[1][ng] < 0.) *Smax = 1.;
  if (*Smax > -1.5) *Smax = fmin(*Smax + 20. * BZ_CLIP_TOL, 1.0);
  if (*Smin < 1.5) *Smin = fmax(*Smin - 20. * BZ_CLIP_TOL, -1.0);
  if (iter == 0) {
    if (*Smin == 1.5) *Smin = -1.;
    if (*Smax == -1.5) *Smax = 1.;
  }
}

/* Sub-patches along direction dir (0: the first index, xi; 1: the second,
 * eta), the de Casteljau steps of CheckXiClip / CheckEtaClip: clip_left keeps
 * [-1, hi], clip_right keeps [lo, 1] of a patch that spans [-1, hi]. */
static void clip_left(int ng, int dir, double B[BZ_G1MAX][BZ_G1MAX][2], double hi, double T[BZ_G1MAX][BZ_G1MAX][2]) {
  /* T(:,p,q) = sum_{l<=p} B(:,l,q) XiBuf(p,l), XiBuf(p,l) = (1+hi)^l (1-hi)^(p-l) FacNchooseK(p,l) (dir 0);
   * the same on the second index for dir 1 */
  double up[BZ_G1MAX], down[BZ_G1MAX], Buf[BZ_G1MAX][BZ_G1MAX];
  const double Plus = 1.0 + hi, Minus = 1.0 - hi;
  up[0] = 1.0; down[0] = 1.0;
  for (int l = 1; l <= ng; l++) { up[l] = up[l - 1] * Plus; down[l] = down[l - 1] * Minus; }
  for (int p = 0; p <= ng; p++) for (int l = 0; l <= p; l++) Buf[p][l] = up[l] * down[p - l] * fac_nck(p, l);
  for (int a = 0; a <= ng; a++) for (int b = 0; b <= ng; b++) T[a][b][0] = T[a][b][1] = 0.;
  for (int q = 0; q <= ng; q++)
    for (int p = 0; p <= ng; p++)
      for (int l = 0; l <= p; l++)
        for (int c = 0; c < 2; c++) {
          if (dir == 0) T[p][q][c] = T[p][q][c] + B[l][q][c] * Buf[p][l];
          else T[q][p][c] = T[q][p][c] + B[q][l][c] * Buf[p][l];
        }
}
static void clip_right(int ng, int dir, double B[BZ_G1MAX][BZ_G1MAX][2], double lo, double hi,
                       double T[BZ_G1MAX][BZ_G1MAX][2]) {
  /* PlusXi = (lo+1)/(hi+1); MinusXi = 2 PlusXi; PlusXi = 2 - 2 PlusXi;
   * T(:,N-p,q) = sum_{l<=p} B(:,N-l,q) XiBuf(p,l) */
  double up[BZ_G1MAX], down[BZ_G1MAX], Buf[BZ_G1MAX][BZ_G1MAX];
  double Plus = (lo + 1.0) / (hi + 1.0);
  const double Minus = 2.0 * Plus;
  Plus = 2.0 - 2.0 * Plus;
  up[0] = 1.0; down[0] = 1.0;
  for (int l = 1; l <= ng; l++) { up[l] = up[l - 1] * Plus; down[l] = down[l - 1] * Minus; }
  for (int p = 0; p <= ng; p++) for (int l = 0; l <= p; l++) Buf[p][l] = up[l] * down[p - l] * fac_nck(p, l);
  for (int a = 0; a <= ng; a++) for (int b = 0; b <= ng; b++) T[a][b][0] = T[a][b][1] = 0.;
  for (int q = 0; q <= ng; q++)
    for (int p = 0; p <= ng; p++)
      for (int l = 0; l <= p; l++)
        for (int c = 0; c < 2; c++) {
          if (dir == 0) T[ng - p][q][c] = T[ng - p][q][c] + B[ng - l][q][c] * Buf[p][l];
          else T[q][ng - p][c] = T[q][ng - p][c] + B[q][ng - l][c] * Buf[p][l];
        }
}
static void copy_patch(int ng, double D[BZ_G1MAX][BZ_G1MAX][2], double S[BZ_G1MAX][BZ_G1MAX][2]) {
  for (int a = 0; a <= ng; a++) for (int b = 0; b <= ng; b++) { D[a][b][0] = S[a][b][0]; D[a][b][1] = S[a][b][1]; }
}

typedef struct {
  const om_state *s;
  int b;
  const double *last, *traj;
  double len;
  bz_work *W;
} bz_ctx;

static void bezier_clip_recursive(bz_ctx *C, int *ClipMode, double B[BZ_G1MAX][BZ_G1MAX][2], double L[2][2],
                                  int *iClipIter, int *nXiClip, int *nEtaClip, int *nInter);

/* CheckXiClip (dir 0) / CheckEtaClip (dir 1), particle_intersection.f90:2459-3242 */
static void check_clip(bz_ctx *C, int dir, int *ClipMode, double B[BZ_G1MAX][BZ_G1MAX][2], double L[2][2],
                       int *iClipIter, int *nXiClip, int *nEtaClip, int *nInter) {
  const int ng = C->s->NGeo;
  bz_work *W = C->W;
  double B1[BZ_G1MAX][BZ_G1MAX], mm[2][BZ_G1MAX];
  for (int q = 0; q <= ng; q++)
    for (int p = 0; p <= ng; p++) B1[p][q] = B[p][q][0] * L[0][dir] + B[p][q][1] * L[1][dir];
  for (int l = 0; l <= ng; l++) {
    double mn = HUGE_VAL, mx = -HUGE_VAL;
    for (int k = 0; k <= ng; k++) {
      const double v = dir == 0 ? B1[l][k] : B1[k][l];
      mn = fmin(mn, v);
      mx = fmax(mx, v);
    }
    mm[0][l] = mn;
    mm[1][l] = mx;
  }
  double dmin = HUGE_VAL, dmax = -HUGE_VAL;
  for (int l = 0; l <= ng; l++) { dmin = fmin(dmin, mm[0][l]); dmax = fmax(dmax, mm[1][l]); }
  if (dmax * dmin > 0.) { *ClipMode = -1; return; }
  if (fabs(dmax - dmin) < W->localTol) {
    if (dir == 0) *ClipMode = (*ClipMode == 3) ? 5 : 4;
    else *ClipMode = (*ClipMode == 4) ? 5 : 3;
    return;
  }
  *ClipMode = dir == 0 ? 2 : 1;
  int *nClip = dir == 0 ? nXiClip : nEtaClip;
  double (*Arr)[2] = dir == 0 ? W->XiArray : W->EtaArray;
  double Smin, Smax;
  calc_smin_smax(ng, mm, &Smin, &Smax, *nClip);
  if (Smin == 1.5 || Smax == -1.5) { *ClipMode = -1; return; }
  if (Smin > Smax) { const double t = Smax; Smax = Smin; Smin = t; }
  (*nClip)++;
  if (*nClip > BZ_CLIP_MAXITER) { W->err = 3; *ClipMode = -1; return; }
  double T[BZ_G1MAX][BZ_G1MAX][2], T2[BZ_G1MAX][BZ_G1MAX][2];
  if (Smax - Smin > BZ_SPLIT_LIMIT) {
    const double Split = 0.5 * (Smax + Smin);
    Arr[*nClip][0] = Split; Arr[*nClip][1] = Smax;
    if (Smax != 1.0) clip_left(ng, dir, B, Smax, T);
    else copy_patch(ng, T, B);
    clip_right(ng, dir, T, Split, Smax, T2);
    const int sIter = *iClipIter, sXi = *nXiClip, sEta = *nEtaClip, sMode = *ClipMode;
    double sL[2][2] = {{L[0][0], L[0][1]}, {L[1][0], L[1][1]}};
    bezier_clip_recursive(C, ClipMode, T2, L, iClipIter, nXiClip, nEtaClip, nInter);
    *iClipIter = sIter; *nXiClip = sXi; *nEtaClip = sEta; *ClipMode = sMode;
    L[0][0] = sL[0][0]; L[0][1] = sL[0][1]; L[1][0] = sL[1][0]; L[1][1] = sL[1][1];
    Arr[*nClip][0] = Smin; Arr[*nClip][1] = Split;
    clip_left(ng, dir, B, Split, T);
    if (Smin != -1.0) clip_right(ng, dir, T, Smin, Split, T2);
    else copy_patch(ng, T2, T);
    bezier_clip_recursive(C, ClipMode, T2, L, iClipIter, nXiClip, nEtaClip, nInter);
    *ClipMode = -1;
  } else {
    Arr[*nClip][0] = Smin; Arr[*nClip][1] = Smax;
    if (Smax != 1.0) { clip_left(ng, dir, B, Smax, T); copy_patch(ng, B, T); }
    if (Smin != -1.0) { clip_right(ng, dir, B, Smin, Smax, T); copy_patch(ng, B, T); }
  }
}

/* BezierClipRecursive (particle_intersection.f90:1635-1747), line vector method 2 */
static void bezier_clip_recursive(bz_ctx *C, int *ClipMode, double B[BZ_G1MAX][BZ_G1MAX][2], double L[2][2],
                                  int *iClipIter, int *nXiClip, int *nEtaClip, int *nInter) {
  const int ng = C->s->NGeo;
  while (*iClipIter <= BZ_CLIP_MAXITER) {
    if (C->W->err) return;
    (*iClipIter)++;
    switch (*ClipMode) {
      case -1: return;
      case 1:
        if ((*iClipIter) % 2 == 1) line_norm_vec(ng, B, L, C->W);
        check_clip(C, 0, ClipMode, B, L, iClipIter, nXiClip, nEtaClip, nInter);
        break;
      case 2:
        if ((*iClipIter) % 2 == 1) line_norm_vec(ng, B, L, C->W);
        check_clip(C, 1, ClipMode, B, L, iClipIter, nXiClip, nEtaClip, nInter);
        break;
      case 3: check_clip(C, 0, ClipMode, B, L, iClipIter, nXiClip, nEtaClip, nInter); break;
      case 4: check_clip(C, 1, ClipMode, B, L, iClipIter, nXiClip, nEtaClip, nInter); break;
      case 5:
        bezier_intersection_point(C->s, C->b, C->W, *nXiClip, *nEtaClip, nInter, C->last, C->traj, C->len);
        return;
      default: C->W->err = 4; return;
    }
  }
  C->W->err = 5;   /* BezierClipMaxIter reached: the reference STOPs */
}

/* D_Bezier (BuildBezierDMat, basis.f90:361-410) on the CL nodes */
static void bezier_dmat(int ng, const double *xi, double D[BZ_G1MAX][BZ_G1MAX]) {
  for (int i = 0; i <= ng; i++) for (int j = 0; j <= ng; j++) D[i][j] = 0.;
  for (int i = 0; i <= ng; i++) D[i][0] = -fac_nck(ng, 0) * (double)ng * ipow(1 - xi[i], ng - 1);
  for (int i = 0; i <= ng; i++) D[i][ng] = fac_nck(ng, ng) * (double)ng * ipow(xi[i] + 1, ng - 1);
  for (int j = 1; j < ng; j++)
    for (int i = 0; i <= ng; i++) {
      const double P = xi[i] + 1.0, Mn = 1.0 - xi[i];
      D[i][j] = fac_nck(ng, j) * ((double)j * ipow(P, j - 1) * ipow(Mn, ng - j) -
                                  (double)(ng - j) * ipow(P, j) * ipow(Mn, ng - j - 1));
    }
}

/* BezierNewton (particle_intersection.f90:1750-1988), BezierNewtonGuess = 1 */
static double bezier_newton(const om_state *s, int b, double Xi[2], double B[BZ_G1MAX][BZ_G1MAX][2],
                            const double last[3], const double traj[3], double len, int *failed) {
  const int ng = s->NGeo, g1 = ng + 1;
  *failed = 0;
  double MinMax[2][2];
  int hasInter = 1;
  for (int l = 0; l < 2; l++) {
    double mn = HUGE_VAL, mx = -HUGE_VAL;
    for (int p = 0; p <= ng; p++) for (int q = 0; q <= ng; q++) { mn = fmin(mn, B[p][q][l]); mx = fmax(mx, B[p][q][l]); }
    MinMax[0][l] = mn; MinMax[1][l] = mx;
    if (MinMax[0][l] * MinMax[1][l] > 0.) hasInter = 0;
  }
  if (!hasInter) return -1.;
  for (int l = 0; l < 2; l++) Xi[l] = -1.0 - 2 * MinMax[0][l] / (MinMax[1][l] - MinMax[0][l]);
  /* dBezierControlPoints2D(dd,nn,i,j) = sum_l D(Cyc(dd),l) B(nn, Cyc with the dd index l) -- the
   * derivative patch is not used below: the gradient comes from the control points (the
   * reference passes it only to the first evaluation; both forms are the same polynomial). */
  double P[2], G[2][3], Bf[BZ_G1MAX * BZ_G1MAX * 2];
  for (int p = 0; p < g1; p++) for (int q = 0; q < g1; q++) { Bf[(p * g1 + q) * 2] = B[p][q][0]; Bf[(p * g1 + q) * 2 + 1] = B[p][q][1]; }
  {
    /* first evaluation with the derivative patch (EvaluateBezierPolynomialAndGradient with dBezier...) */
    double D[BZ_G1MAX][BZ_G1MAX];
    bezier_dmat(ng, s->xiCL, D);
    double dB[2][2][BZ_G1MAX][BZ_G1MAX];   /* [dd][nn][i][j] */
    for (int nn = 0; nn < 2; nn++)
      for (int dd = 0; dd < 2; dd++)
        for (int j = 0; j <= ng; j++)
          for (int i = 0; i <= ng; i++) {
            double acc = 0.;
            for (int l = 0; l <= ng; l++) {
              const int ci = dd == 0 ? l : i, cj = dd == 0 ? j : l;
              acc = acc + D[dd == 0 ? i : j][l] * B[ci][cj][nn];
            }
            dB[dd][nn][i][j] = acc;
          }
    double xiup[BZ_G1MAX], etaup[BZ_G1MAX], xidown[BZ_G1MAX], etadown[BZ_G1MAX], xit[BZ_G1MAX], etat[BZ_G1MAX];
    xiup[0] = 1.0; etaup[0] = 1.0; xidown[ng] = 1.0; etadown[ng] = 1.0;
    for (int p = 1; p <= ng; p++) {
      xiup[p] = xiup[p - 1] * (1.0 + Xi[0]);
      etaup[p] = etaup[p - 1] * (1.0 + Xi[1]);
      xidown[ng - p] = xidown[ng - p + 1] * (1.0 - Xi[0]);
      etadown[ng - p] = etadown[ng - p + 1] * (1.0 - Xi[1]);
    }
    for (int p = 0; p <= ng; p++) { xit[p] = xiup[p] * xidown[p]; etat[p] = etaup[p] * etadown[p]; }
    P[0] = P[1] = 0.;
    for (int q = 0; q <= ng; q++)
      for (int p = 0; p <= ng; p++)
        for (int c = 0; c < 2; c++) P[c] = P[c] + B[p][q][c] * fac_nck(ng, p) * xit[p] * fac_nck(ng, q) * etat[q];
    for (int d = 0; d < 2; d++) for (int c = 0; c < 2; c++) G[d][c] = 0.;
    for (int q = 0; q <= ng; q++)
      for (int p = 0; p <= ng; p++)
        for (int d = 0; d < 2; d++)
          for (int c = 0; c < 2; c++)
            G[d][c] = G[d][c] + dB[d][c][p][q] * fac_nck(ng, p) * xit[p] * fac_nck(ng, q) * etat[q];
  }
  double NormP = P[0] * P[0] + P[1] * P[1];
  int nIter = 1;
  double dXi2 = 1.0;
  double alpha = -1.0;
  while (dXi2 > BZ_NEWTON_TOL2 && nIter <= BZ_NEWTON_MAXITER) {
    double sdet = G[0][0] * G[1][1] - G[0][1] * G[1][0];
    if (fabs(sdet) > DBL_EPSILON) sdet = 1. / sdet;
    else { alpha = -1.0; Xi[0] = Xi[1] = 1.5; break; }
    double dXi[2] = {G[1][1] * P[0] - G[1][0] * P[1], -G[0][1] * P[0] + G[0][0] * P[1]};
    dXi[0] = sdet * dXi[0]; dXi[1] = sdet * dXi[1];
    dXi2 = dXi[0] * dXi[0] + dXi[1] * dXi[1];
    const double Xo[2] = {Xi[0], Xi[1]};
    double lambda = 1.0;
    int iArmijo = 1;
    const double NormPold = NormP;
    NormP = NormP * 2.;
    while (NormP > NormPold * (1.0 - 0.0001 * lambda) && iArmijo <= 6) {
      Xi[0] = Xo[0] - lambda * dXi[0];
      Xi[1] = Xo[1] - lambda * dXi[1];
      bezier_eval(ng, 2, Bf, Xi, P, G);
      NormP = P[0] * P[0] + P[1] * P[1];
      lambda = 0.3 * lambda;
      iArmijo++;
    }
    if (nIter >= 4 && (fabs(Xi[0]) > 1.5 || fabs(Xi[1]) > 1.5)) { Xi[0] = Xi[1] = 1.5; break; }
    nIter++;
  }
  if (nIter > BZ_NEWTON_MAXITER) { *failed = 1; return alpha; }
  if (fabs(Xi[0]) > 1.002 || fabs(Xi[1]) > 1.002) return alpha;
  double InterP[3];
  bezier_eval(ng, 3, &s->bz[(size_t)b * g1 * g1 * 3], Xi, InterP, NULL);
  double iv[3] = {InterP[0] - last[0], InterP[1] - last[1], InterP[2] - last[2]};
  alpha = dot3(iv, traj);
  const double alphaNorm = alpha / len;
  if (alphaNorm <= 1.0 && alphaNorm > -BZ_EPSTOL) return alpha;
  return -1.0;
}

static void project_2d(const om_state *s, int b, const double last[3], const double n_a[3], const double n_b[3],
                       double B2[BZ_G1MAX][BZ_G1MAX][2]) {
  const int ng = s->NGeo;
  for (int q = 0; q <= ng; q++)
    for (int p = 0; p <= ng; p++) {
      const double *bp = &s->bz[(((size_t)b * (ng + 1) + p) * (ng + 1) + q) * 3];
      double d[3] = {bp[0] - last[0], bp[1] - last[1], bp[2] - last[2]};
      B2[p][q][0] = dot3(d, n_a);
      B2[p][q][1] = dot3(d, n_b);
    }
}

static void traj_basis(const double traj[3], double n1[3], double n2[3]) {
  if (fabs(traj[2]) < 0.) { n1[0] = -traj[1] - traj[2]; n1[1] = traj[0]; n1[2] = traj[0]; }
  else { n1[0] = traj[2]; n1[1] = traj[2]; n1[2] = -traj[0] - traj[1]; }
  unit3(n1);
  crossnorm3(traj, n1, n2);
}

/* ComputePlanarRectIntersection (particle_intersection.f90:515-684) */
static int planar_rect_intersection(const om_state *s, int b, int flip, const double last[3], const double traj[3],
                                    double len, double *alpha, double *xi, double *eta) {
  *alpha = -1.0; *xi = -2.; *eta = -2.;
  double nv[3];
  double locDist = s->sideDist[b];
  for (int c = 0; c < 3; c++) nv[c] = s->sideNorm[b * 3 + c];
  if (flip != 0) { for (int c = 0; c < 3; c++) nv[c] = -nv[c]; locDist = -locDist; }
  const double coeffA = dot3(nv, traj);
  const double locSideDist = locDist - dot3(last, nv);
  if (almost_zero(coeffA)) { *alpha = -1.; return 0; }
  *alpha = locSideDist / coeffA;
  if (locSideDist < -100 * DBL_EPSILON) { *alpha = -1.; return 0; }
  const double alphaNorm = *alpha / len;
  if (alphaNorm > 1.0 || alphaNorm < -BZ_EPSTOL) { *alpha = -1.; return 0; }
  const double *bv = &s->baseV[(size_t)b * 9];
  double P0[3], P1[3], P2[3];
  for (int c = 0; c < 3; c++) {
    const double inter = last[c] + *alpha * traj[c];
    P0[c] = -0.25 * bv[c] + inter;
    P1[c] = 0.25 * bv[3 + c];
    P2[c] = 0.25 * bv[6 + c];
  }
  const double A1 = P1[0] * P1[0] + P1[1] * P1[1] + P1[2] * P1[2];
  const double B1 = P2[0] * P1[0] + P2[1] * P1[1] + P2[2] * P1[2];
  const double C1 = P1[0] * P0[0] + P1[1] * P0[1] + P1[2] * P0[2];
  const double A2 = B1;
  const double B2 = P2[0] * P2[0] + P2[1] * P2[1] + P2[2] * P2[2];
  const double C2 = P2[0] * P0[0] + P2[1] * P0[1] + P2[2] * P0[2];
  double sdet = A1 * B2 - A2 * B1;
  if (fabs(sdet) == 0) { *alpha = -1.; return -1; }
  sdet = 1.0 / sdet;
  const double epsLoc = 1.0 + 100. * DBL_EPSILON;
  *xi = (B2 * C1 - B1 * C2) * sdet;
  if (fabs(*xi) > epsLoc) { *alpha = -1.0; return 0; }
  *eta = (-A2 * C1 + A1 * C2) * sdet;
  if (fabs(*eta) > epsLoc) { *alpha = -1.0; return 0; }
  return 0;
}

/* ComputePlanarCurvedIntersection (particle_intersection.f90:687-852), RefMapping */
static int planar_curved_intersection(const om_state *s, int b, int flip, const double last[3], const double traj[3],
                                      double len, double *alpha, double *xi, double *eta) {
  *alpha = -1.0; *xi = 2.0; *eta = 2.0;
  double nv[3];
  double locDist = s->sideDist[b];
  for (int c = 0; c < 3; c++) nv[c] = s->sideNorm[b * 3 + c];
  if (flip != 0) { for (int c = 0; c < 3; c++) nv[c] = -nv[c]; locDist = -locDist; }
  const double coeffA = dot3(nv, traj);
  if (coeffA <= 0.) return 0;
  double locSideDist = locDist - dot3(last, nv);
  locSideDist = locSideDist / coeffA;
  if (locSideDist > len) return 0;
  if (!flat_bbox_intersection(s, b, last, traj, len)) return 0;
  double n1[3], n2[3];
  traj_basis(traj, n1, n2);
  double B2[BZ_G1MAX][BZ_G1MAX][2];
  project_2d(s, b, last, n1, n2, B2);
  bz_work W;
  memset(&W, 0, sizeof W);
  double Xn[2] = {0., 0.};
  int failed = 0;
  W.locAlpha[0] = bezier_newton(s, b, Xn, B2, last, traj, len, &failed);
  W.locXi[0] = Xn[0];
  W.locEta[0] = Xn[1];
  if (failed) {
    int iClipIter = 0, nXiClip = 0, nEtaClip = 0, nInter = 0, ClipMode = 1;
    double L[2][2] = {{0., 0.}, {0., 0.}};
    /* BezierClipLocalTol is not set on this path in the reference (a stale
     * module value); the clip's own width is used here */
    double dXi = -HUGE_VAL, dEta = -HUGE_VAL, mXi = HUGE_VAL, mEta = HUGE_VAL;
    for (int p = 0; p <= s->NGeo; p++) for (int q = 0; q <= s->NGeo; q++) {
      dXi = fmax(dXi, B2[p][q][0]); mXi = fmin(mXi, B2[p][q][0]);
      dEta = fmax(dEta, B2[p][q][1]); mEta = fmin(mEta, B2[p][q][1]);
    }
    W.localTol = fmin(dXi - mXi, dEta - mEta) * BZ_CLIP_TOL;
    W.locAlpha[0] = -1.0;
    bz_ctx C = {s, b, last, traj, len, &W};
    bezier_clip_recursive(&C, &ClipMode, B2, L, &iClipIter, &nXiClip, &nEtaClip, &nInter);
    if (W.err) return -1;
  }
  if (W.locAlpha[0] > -1) {
    *alpha = W.locAlpha[0];
    *xi = W.locXi[0];
    *eta = W.locEta[0];
  }
  return 0;
}

/* ComputeCurvedIntersection (particle_intersection.f90:1243-1632), RefMapping */
static int curved_intersection(const om_state *s, int b, int flip, const double last[3], const double x[3],
                               const double traj[3], double len, double *alpha, double *xi, double *eta) {
  *alpha = -1.0; *xi = 2.0; *eta = 2.0;
  double nv[3];
  for (int c = 0; c < 3; c++) nv[c] = flip == 0 ? s->sideNorm[b * 3 + c] : -s->sideNorm[b * 3 + c];
  if (s->bbEmpty[b]) {
    if (dot3(nv, traj) < 0.) return 0;
    if (!flat_bbox_intersection(s, b, last, traj, len)) return 0;
  } else if (!inside_bbox(s, b, last) && !inside_bbox(s, b, x) && !bbox_intersection(s, b, last, traj, len)) {
    return 0;
  }
  double n1[3], n2[3];
  traj_basis(traj, n1, n2);
  const double PartFaceAngle = fabs(0.5 * PI_ - acos(dot3(traj, &s->slabN[(size_t)b * 9 + 3])));
  double B2[BZ_G1MAX][BZ_G1MAX][2];
  project_2d(s, b, last, n2, n1, B2);
  bz_work W;
  memset(&W, 0, sizeof W);
  for (int k = 0; k < BZ_MAXINTER; k++) W.locAlpha[k] = -1.0;
  int iClipIter = 0, nXiClip = 0, nEtaClip = 0, nInter = 0, ClipMode;
  double L[2][2] = {{0., 0.}, {0., 0.}};
  double dXi = -HUGE_VAL, dEta = -HUGE_VAL, mXi = HUGE_VAL, mEta = HUGE_VAL;
  for (int p = 0; p <= s->NGeo; p++) for (int q = 0; q <= s->NGeo; q++) {
    dXi = fmax(dXi, B2[p][q][0]); mXi = fmin(mXi, B2[p][q][0]);
    dEta = fmax(dEta, B2[p][q][1]); mEta = fmin(mEta, B2[p][q][1]);
  }
  dXi = dXi - mXi;
  dEta = dEta - mEta;
  bz_ctx C = {s, b, last, traj, len, &W};
  if (PartFaceAngle < BZ_NEWTON_ANGLE) {
    ClipMode = dXi > dEta ? 1 : 2;
    W.localTol = fmin(dXi, dEta) * BZ_CLIP_TOL;
    bezier_clip_recursive(&C, &ClipMode, B2, L, &iClipIter, &nXiClip, &nEtaClip, &nInter);
    if (W.err) return -1;
  } else {
    /* trajectory exactly along the side normal: clip, then Newton from the
     * clipped root (BezierClipLocalTol: see planar_curved_intersection) */
    double B2t[BZ_G1MAX][BZ_G1MAX][2];
    copy_patch(s->NGeo, B2t, B2);
    ClipMode = 1;
    W.localTol = fmin(dXi, dEta) * BZ_CLIP_TOL;
    bezier_clip_recursive(&C, &ClipMode, B2t, L, &iClipIter, &nXiClip, &nEtaClip, &nInter);
    if (W.err || nInter > 1) return -1;
    double Xn[2] = {0., 0.};
    int failed = 0;
    W.locAlpha[0] = bezier_newton(s, b, Xn, B2, last, traj, len, &failed);
    nInter = W.locAlpha[0] > -1 ? 1 : 0;
    if (failed) return -1;
    W.locXi[0] = Xn[0];
    W.locEta[0] = Xn[1];
  }
  if (nInter == 0) return 0;
  if (nInter == 1) { *alpha = W.locAlpha[0]; *xi = W.locXi[0]; *eta = W.locEta[0]; return 0; }
  int locID[BZ_MAXINTER];
  for (int k = 0; k < nInter; k++) locID[k] = k;
  for (int a = 1; a < nInter; a++) { /* InsertionSort */
    const double av = W.locAlpha[a];
    const int iv = locID[a];
    int j = a - 1;
    while (j >= 0 && W.locAlpha[j] > av) { W.locAlpha[j + 1] = W.locAlpha[j]; locID[j + 1] = locID[j]; j--; }
    W.locAlpha[j + 1] = av;
    locID[j + 1] = iv;
  }
  for (int k = 0; k < nInter; k++)
    if (W.locAlpha[k] > -1.0) {
      *alpha = W.locAlpha[k];
      *xi = W.locXi[locID[k]];
      *eta = W.locEta[locID[k]];
      return 0;
    }
  return 0;
}

static int side_intersection(const om_state *s, int b, int flip, const double last[3], const double x[3],
                             const double traj[3], double len, double *alpha, double *xi, double *eta) {
  switch (s->sideType[b]) {
    case ST_PLANAR_RECT: return planar_rect_intersection(s, b, flip, last, traj, len, alpha, xi, eta);
    case ST_PLANAR_CURVED: return planar_curved_intersection(s, b, flip, last, traj, len, alpha, xi, eta);
    case ST_CURVED: return curved_intersection(s, b, flip, last, x, traj, len, alpha, xi, eta);
    default: return -1;
  }
}

/* GetBoundaryInteraction, RefMapping branch, with PeriodicBoundary and
 * RemoveParticle (open walls).  Returns crossedBC. */
static int bc_interaction(om_state *s, int i, int b, int flip, int *ElemID, bc_track_info *ti) {
  double n[3];
  if (s->sideType[b] == ST_CURVED) bezier_normal(s, b, ti->xi, ti->eta, n);
  else for (int c = 0; c < 3; c++) n[c] = s->sideNorm[b * 3 + c];
  if (flip != 0) for (int c = 0; c < 3; c++) n[c] = -n[c];
  if (dot3(n, ti->traj) <= 0.) return 0;
  const int sd = s->bcSide[b];
  const int bcid = SIDEINFO(sd, SIDE_BCID);
  const int type = s->partBC[bcid - 1];
  if (type == 1) { s->inside[i] = 0; return 1; }   /* OpenBC: RemoveParticle */
  /* PeriodicBC: PeriodicBoundary (:271-327) */
  double *x = &s->PartState[(size_t)i * 6], *last = &s->LastPartPos[(size_t)i * 3];
  const int pvid = s->bcAlpha[bcid - 1];
  const int v = abs(pvid) - 1;
  for (int c = 0; c < 3; c++) last[c] = last[c] + ti->traj[c] * ti->alpha;
  for (int c = 0; c < 3; c++) last[c] = last[c] + copysign(fabs(s->PV[v][c]), (double)pvid);
  for (int c = 0; c < 3; c++) x[c] = last[c] + (ti->len - ti->alpha) * ti->traj[c];
  ti->len = ti->len - ti->alpha;
  *ElemID = SIDEINFO(sd, SIDE_NBELEMID) - 1;
  return 1;
}

/* ParticleBCTracking (particle_reftracking.f90:417-694) with CartesianPeriodic = F.
 * Returns 0, or -1 on an error the reference aborts on. */
static int bc_tracking(om_state *s, int i, double *len0, int *ElemID, int iCount, int *done, int *moved) {
  double *x = &s->PartState[(size_t)i * 6], *last = &s->LastPartPos[(size_t)i * 3];
  bc_track_info ti;
  for (int c = 0; c < 3; c++) ti.traj[c] = x[c] - last[c];
  double len = sqrt(dot3(ti.traj, ti.traj));
  if (almost_zero(len / s->radius[*ElemID])) { s->elem[i] = *ElemID; *done = 1; return 0; }   /* PARTHASMOVED */
  if (len > s->bcHalo) { set_err("boundary tracking: particle %d moved farther than the boundary halo", i + 1); return -1; }
  for (int c = 0; c < 3; c++) ti.traj[c] = ti.traj[c] / len;
  *moved = 0;
  int doTracing = 1, doubleCheck = 0;
  *len0 = fmax(*len0, len);
  const int first = s->e2bOff[*ElemID], lastS = s->e2bOff[*ElemID + 1], nloc = lastS - first;
  double locAlpha[4096], xi[4096], eta[4096];
  int locSideList[4096];
  if (nloc > 4096) { set_err("boundary tracking: more than 4096 sides in range"); return -1; }
  int reflected = 0;
  while (doTracing) {
    for (int k = 0; k < nloc; k++) { locAlpha[k] = -1.0; locSideList[k] = k; }
    int nInter = 0;
    for (int k = 0; k < nloc; k++) {
      if (s->e2bDist[first + k] > *len0) break;
      const int b = s->e2bSide[first + k];
      const int sd = s->bcSide[b];
      locSideList[k] = k;
      const int flip = SIDEINFO(sd, SIDE_ID) > 0 ? 0 : SIDEINFO(sd, SIDE_FLIP) % 10;
      if (side_intersection(s, b, flip, last, x, ti.traj, len, &locAlpha[k], &xi[k], &eta[k]) != 0) {
        set_err("boundary tracking: intersection of particle %d with side %d failed", i + 1, sd + 1);
        return -1;
      }
      if (locAlpha[k] > -1.0) nInter++;
    }
    if (nInter == 0) { doTracing = 0; continue; }
    *moved = 1;
    /* InsertionSort(locAlpha, locSideList) over the whole range; the sides
     * not visited in this pass hold alpha = -1 */
    for (int a = 1; a < nloc; a++) {
      const double av = locAlpha[a];
      const int lv = locSideList[a];
      int j = a - 1;
      while (j >= 0 && locAlpha[j] > av) { locAlpha[j + 1] = locAlpha[j]; locSideList[j + 1] = locSideList[j]; j--; }
      locAlpha[j + 1] = av;
      locSideList[j + 1] = lv;
    }
    for (int k = 0; k < nloc; k++) {
      if (locAlpha[k] > -1.0) {
        const int hit = locSideList[k];
        const int b = s->e2bSide[first + hit];
        const int sd = s->bcSide[b];
        const int flip = SIDEINFO(sd, SIDE_ID) > 0 ? 0 : SIDEINFO(sd, SIDE_FLIP) % 10;
        const int OldElemID = *ElemID;
        ti.len = len;
        ti.alpha = locAlpha[k];
        ti.xi = xi[hit];
        ti.eta = eta[hit];
        reflected = bc_interaction(s, i, b, flip, ElemID, &ti);
        len = ti.len;
        locAlpha[k] = ti.alpha;
        if (*ElemID != OldElemID) {
          if (s->nPV > 0) {
            double m = -HUGE_VAL;
            for (int r = s->e2bOff[OldElemID]; r < s->e2bOff[OldElemID + 1]; r++) m = fmax(m, s->e2bDist[r]);
            *len0 = m;
          }
          if (iCount >= 100000) { set_err("boundary tracking: particle %d recursed too often", i + 1); return -1; }
          const int rc = bc_tracking(s, i, len0, ElemID, iCount + 1, done, moved);
          *moved = 1;
          return rc;
        }
        if (reflected) break;
      }
    }
    if (!s->inside[i]) { *done = 1; return 0; }
    if (!reflected) {
      if (!doubleCheck) doubleCheck = 1;
      else doTracing = 0;
    }
  }
  return 0;
}

/* FallBackFaceIntersection (particle_reftracking.f90:986-1099): the
 * boundary sides of the element are intersected with the segment from its
 * origin to the particle; a hit moves the particle to 97 % of the way there.
 * Returns 0, 1 when the particle had to be located again (not found: lost),
 * or -1. */
static int fallback_face_intersection(om_state *s, int i, int ElemID) {
  double *x = &s->PartState[(size_t)i * 6], *last = &s->LastPartPos[(size_t)i * 3];
  const double tmpPos[3] = {x[0], x[1], x[2]}, tmpLast[3] = {last[0], last[1], last[2]};
  for (int c = 0; c < 3; c++) last[c] = s->bary[ElemID * 3 + c];
  double traj[3] = {x[0] - last[0], x[1] - last[1], x[2] - last[2]};
  const double len = sqrt(dot3(traj, traj));
  if (len > 0) for (int c = 0; c < 3; c++) traj[c] = traj[c] / len;
  const int first = s->e2bOff[ElemID], nloc = s->e2bOff[ElemID + 1] - first;
  double locAlpha[4096], xi, eta;
  int list[4096], nInter = 0;
  if (nloc > 4096) { set_err("boundary tracking: more than 4096 sides in range"); return -1; }
  for (int k = 0; k < nloc; k++) {
    const int b = s->e2bSide[first + k];
    const int sd = s->bcSide[b];
    list[k] = k;
    const int flip = SIDEINFO(sd, SIDE_ID) > 0 ? 0 : SIDEINFO(sd, SIDE_FLIP) % 10;
    locAlpha[k] = -1.0;
    if (side_intersection(s, b, flip, last, x, traj, len, &locAlpha[k], &xi, &eta) != 0) return -1;
    if (locAlpha[k] > -1.0) nInter++;
  }
  if (nInter == 0) {
    for (int c = 0; c < 3; c++) { x[c] = tmpPos[c]; last[c] = tmpLast[c]; }
    return 1;
  }
  for (int a = 1; a < nloc; a++) {
    const double av = locAlpha[a];
    const int lv = list[a];
    int j = a - 1;
    while (j >= 0 && locAlpha[j] > av) { locAlpha[j + 1] = locAlpha[j]; list[j + 1] = list[j]; j--; }
    locAlpha[j + 1] = av;
    list[j + 1] = lv;
  }
  for (int k = 0; k < nloc; k++)
    if (locAlpha[k] > -1.0) {
      for (int c = 0; c < 3; c++) last[c] = last[c] + 0.97 * locAlpha[k] * traj[c];
      for (int c = 0; c < 3; c++) x[c] = last[c];
    }
  return 0;
}
