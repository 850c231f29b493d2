// Host-side producers of the particle-mesh tables the stage kernels read.
//
// Each table restates the reference's particle-mesh initialisation (paths
// relative to the reference's src/):
//   ElemSideNodeID, ConcaveElemSide   particles/particle_mesh/particle_mesh_tools.f90:1800-1930
//   XCL_NGeo, dXCL_NGeo               mesh/metrics.f90:293-338
//   ElemBaryNGeo, ElemRadiusNGeo      particles/particle_mesh/particle_mesh_build.f90:217-309
//   XiEtaZetaBasis, slen...           particles/particle_mesh/particle_mesh_build.f90:386-444
//   Elem_xGP, sJ                      mesh/metrics.f90:293-356 (evaluated at the Gauss points)
//   NodeVolume + periodic twins       particles/pic/deposition/pic_depo_tools.f90:224-355,
//                                     particles/pic/deposition/pic_depo.f90:1218-
// plus the shape-function source lists, which have no counterpart in the
// reference (it searches its FIBGM per particle, pic_depo_shapefunction_tools.f90:292-422;
// the stage instead lists, per target element, every element whose particles
// can reach it, so deposition is a gather).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

#include "stage_internal.h"

namespace pic {
namespace {

// basis.f90:1011-1022
bool almost_equal_unity(double x, double y) {
  const double eps = 2.220446049250313e-16;
  if (x == 0.0 || y == 0.0) return std::fabs(x - y) <= 2.0 * eps;
  return std::fabs(x - y) <= eps * std::fabs(x) && std::fabs(x - y) <= eps * std::fabs(y);
}

// basis.f90:1223-1264
void lagrange(double x, int N, const double *xg, const double *wb, double *L) {
  bool eq = false;
  for (int i = 0; i <= N; ++i) {
    L[i] = 0.0;
    if (almost_equal_unity(x, xg[i])) { L[i] = 1.0; eq = true; }
  }
  if (eq) return;
  double sum = 0.0;
  for (int i = 0; i <= N; ++i) { L[i] = wb[i] / (x - xg[i]); sum += L[i]; }
  for (int i = 0; i <= N; ++i) L[i] = L[i] / sum;
}

// basis.f90:643-688, normalised Legendre polynomial and derivative
void legendre(int N, double x, double &L, double &Ld) {
  double l = 0, d = 0;
  if (N == 0) { l = 1; d = 0; }
  else if (N == 1) { l = x; d = 1; }
  else {
    double l2 = 1, l1 = x, d2 = 0, d1 = 1;
    for (int i = 2; i <= N; ++i) {
      l = (double(2 * i - 1) * x * l1 - double(i - 1) * l2) / double(i);
      d = d2 + double(2 * i - 1) * l1;
      l2 = l1; l1 = l; d2 = d1; d1 = d;
    }
  }
  L = l * std::sqrt(double(N) + 0.5);
  Ld = d * std::sqrt(double(N) + 0.5);
}

// basis.f90:758-835 Gauss-Legendre nodes by Newton from Chebyshev guesses
void gauss_nodes(int N, double *x, double *w) {
  double L, Ld;
  if (N == 0) { x[0] = 0; w[0] = 2; return; }
  if (N == 1) { x[0] = -std::sqrt(1.0 / 3.0); x[1] = -x[0]; w[0] = w[1] = 1.0; return; }
  const double cheb = 2.0 * std::atan(1.0) / double(N + 1);
  for (int i = 0; i <= (N + 1) / 2 - 1; ++i) {
    x[i] = -std::cos(cheb * double(2 * i + 1));
    for (int it = 0; it <= 10; ++it) {
      legendre(N + 1, x[i], L, Ld);
      double dx = -L / Ld;
      x[i] += dx;
      if (std::fabs(dx) < 1e-15 * std::fabs(x[i])) break;
    }
    legendre(N + 1, x[i], L, Ld);
    x[N - i] = -x[i];
    w[i] = (2.0 * N + 3) / ((1.0 - x[i] * x[i]) * Ld * Ld);
    w[N - i] = w[i];
  }
  if (N % 2 == 0) {
    x[N / 2] = 0.0;
    legendre(N + 1, 0.0, L, Ld);
    w[N / 2] = (2.0 * N + 3) / (Ld * Ld);
  }
}

// basis.f90:952-976
void bary_weights(int N, const double *x, double *wb) {
  for (int i = 0; i <= N; ++i) wb[i] = 1.0;
  for (int i = 1; i <= N; ++i)
    for (int j = 0; j < i; ++j) { wb[j] *= (x[j] - x[i]); wb[i] *= (x[i] - x[j]); }
  for (int i = 0; i <= N; ++i) wb[i] = 1.0 / wb[i];
}

// Sum-factorised tensor-product change of basis of a 3-vector field
// (mesh/changeBasis.f90:54-111).
void change_basis(int Nin, int Nout, const double *V, const double *in, double *out) {
  const int a = Nin + 1, b = Nout + 1;
  std::vector<double> b1(size_t(b) * a * a * 3, 0.0), b2(size_t(b) * b * a * 3, 0.0);
  for (int k = 0; k < a; ++k) for (int j = 0; j < a; ++j) for (int i = 0; i < a; ++i)
    for (int io = 0; io < b; ++io) for (int c = 0; c < 3; ++c)
      b1[((k * a + j) * b + io) * 3 + c] += V[io * a + i] * in[((k * a + j) * a + i) * 3 + c];
  for (int k = 0; k < a; ++k) for (int j = 0; j < a; ++j) for (int jo = 0; jo < b; ++jo)
    for (int io = 0; io < b; ++io) for (int c = 0; c < 3; ++c)
      b2[((k * b + jo) * b + io) * 3 + c] += V[jo * a + j] * b1[((k * a + j) * b + io) * 3 + c];
  std::fill(out, out + size_t(b) * b * b * 3, 0.0);
  for (int k = 0; k < a; ++k) for (int ko = 0; ko < b; ++ko) for (int jo = 0; jo < b; ++jo)
    for (int io = 0; io < b; ++io) for (int c = 0; c < 3; ++c)
      out[((ko * b + jo) * b + io) * 3 + c] += V[ko * a + k] * b2[((k * b + jo) * b + io) * 3 + c];
}

// ---- RefMapping boundary tracking tables (ParticleBCTracking) -------------
// Local side ids of the reference (piclas.h:204-209).
enum { kZetaMinus = 1, kEtaMinus = 2, kXiPlus = 3, kEtaPlus = 4, kXiMinus = 5, kZetaPlus = 6 };

double choose_d(int n, int k) {
  double r = 1.0;
  for (int q = 1; q <= k; ++q) r = r * double(n - q + 1) / double(q);
  return r;
}
double ipow(double x, int k) {
  double r = 1.0;
  for (int q = 0; q < k; ++q) r = r * x;
  return r;
}

// sVdm_Bezier = INVERSE(Vdm_Bezier) on the CL nodes (BuildBezierVdm,
// basis.f90:275-358; INVERSE's closed forms for 2x2 and 3x3, mathtools.f90:44-102)
void bezier_svdm(int ng, const double *xi, double sV[3][3]) {
  double A[3][3];
  for (int i = 0; i <= ng; ++i)
    for (int j = 0; j <= ng; ++j)
      A[i][j] = (1.0 / double(1 << ng)) * choose_d(ng, j) * ipow(xi[i] + 1.0, j) * ipow(1.0 - xi[i], ng - j);
  if (ng == 1) {
    const double sdet = 1.0 / (A[0][0] * A[1][1] - A[0][1] * A[1][0]);
    sV[0][0] = (A[1][1]) * sdet; sV[0][1] = (-A[0][1]) * sdet;
    sV[1][0] = (-A[1][0]) * sdet; sV[1][1] = (A[0][0]) * sdet;
    return;
  }
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

// the CL points of local side ls, tmp[a][b] over the two free volume indices
void side_slice(const double *X, int g1, int ls, double tmp[3][3][3]) {
  const int ng = g1 - 1;
  for (int a = 0; a < g1; ++a)
    for (int b = 0; b < g1; ++b) {
      int i, j, k;
      switch (ls) {
        case kXiMinus: i = 0; j = a; k = b; break;
        case kXiPlus: i = ng; j = a; k = b; break;
        case kEtaMinus: i = a; j = 0; k = b; break;
        case kEtaPlus: i = a; j = ng; k = b; break;
        case kZetaMinus: i = a; j = b; k = 0; break;
        default: i = a; j = b; k = ng; break;
      }
      for (int c = 0; c < 3; ++c) tmp[a][b][c] = X[((k * g1 + j) * g1 + i) * 3 + c];
    }
}

// BezierControlPoints3D of local side ls (CalcBezierControlPoints,
// particle_mesh_tools.f90:1718-1741): the slice in the Bernstein basis
// (ChangeBasis2D, first index then second), in the CGNS side orientation
// (CGNS_SideToVol2, mesh/mapping.f90:320-356)
void side_bezier(const double *X, int ng, int ls, const double sV[3][3], double out[3][3][3]) {
  const int g1 = ng + 1;
  double tmp[3][3][3], buf[3][3][3], t2[3][3][3];
  side_slice(X, g1, ls, tmp);
  for (int jn = 0; jn < g1; ++jn)
    for (int io = 0; io < g1; ++io) {
      double v[3] = {0, 0, 0};
      for (int in = 0; in < g1; ++in)
        for (int c = 0; c < 3; ++c) v[c] = v[c] + sV[io][in] * tmp[in][jn][c];
      for (int c = 0; c < 3; ++c) buf[io][jn][c] = v[c];
    }
  for (int jo = 0; jo < g1; ++jo)
    for (int io = 0; io < g1; ++io) {
      double v[3] = {0, 0, 0};
      for (int jn = 0; jn < g1; ++jn)
        for (int c = 0; c < 3; ++c) v[c] = v[c] + sV[jo][jn] * buf[io][jn][c];
      for (int c = 0; c < 3; ++c) t2[io][jo][c] = v[c];
    }
  for (int q = 0; q < g1; ++q)
    for (int p = 0; p < g1; ++p) {
      int a = p, b = q;
      if (ls == kXiMinus || ls == kZetaMinus) { a = q; b = p; }
      else if (ls == kEtaPlus) { a = ng - p; b = q; }
      for (int c = 0; c < 3; ++c) out[p][q][c] = t2[a][b][c];
    }
}

double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
void unit3(double v[3]) {
  const double l = std::sqrt(dot3(v, v));
  v[0] = v[0] / l; v[1] = v[1] / l; v[2] = v[2] / l;
}
void crossnorm3(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  unit3(c);
}

// GetSideSlabNormalsAndIntervals (particle_surfaces.f90:619-816); false for a
// degenerated side (the reference aborts)
bool side_slab(int ng, double B[3][3][3], double N[3][3], double I[6], uint8_t &empty) {
  double n2[3];
  for (int c = 0; c < 3; ++c) N[0][c] = B[ng][0][c] - B[0][0][c] + B[ng][ng][c] - B[0][ng][c];
  if (N[0][0] == 0 && N[0][1] == 0 && N[0][2] == 0) return false;
  const double l0 = std::sqrt(dot3(N[0], N[0]));
  for (int c = 0; c < 3; ++c) N[0][c] = N[0][c] / l0;
  for (int c = 0; c < 3; ++c) n2[c] = B[0][ng][c] - B[0][0][c] + B[ng][ng][c] - B[ng][0][c];
  crossnorm3(N[0], n2, N[1]);
  crossnorm3(N[1], N[0], N[2]);
  for (int c = 0; c < 6; ++c) I[c] = 0.0;
  for (int q = 0; q <= ng; ++q)
    for (int p = 0; p <= ng; ++p) {
      if (p == 0 && q == 0) continue;
      const double d[3] = {B[p][q][0] - B[0][0][0], B[p][q][1] - B[0][0][1], B[p][q][2] - B[0][0][2]};
      for (int a = 0; a < 3; ++a) {
        const double sp = dot3(d, N[a]);
        if (sp < 0.) I[2 * a] = std::fmin(I[2 * a], sp);
        else if (sp > 0.) I[2 * a + 1] = std::fmax(I[2 * a + 1], sp);
      }
    }
  const double dx = std::fabs(I[1] - I[0]), dy = std::fabs(I[3] - I[2]), dz = std::fabs(I[5] - I[4]);
  double h = dy;
  const double w = std::fmin(dx, dz), l = std::fmax(dx, dz);
  const double dMax = std::fmax(std::fmax(h, l), w);
  if (l / dMax < 1.0e-6 || w / dMax < 1.0e-6) return false;
  if (h / dMax < 1.0e-6) { I[2] = 0.; I[3] = 0.; h = 0.; }
  empty = std::fabs(h / std::sqrt(l * l + w * w + h * h)) <= 2.22e-16;
  return true;
}

// Boundary side tables and the element -> boundary side lists for
// ParticleBCTracking, built when not every boundary is periodic
// (CartesianPeriodic = F): SideType / SideNormVec / SideDistance
// (IdentifyElemAndSideType, particle_mesh_tools.f90:976-1155), BaseVectors
// and BCSideMetrics (particle_mesh_build.f90:1760-1930), ElemToBCSides and
// SideBCMetrics sorted by distance (BuildBCElemDistance, non-fullMesh
// branch, particle_mesh_build.f90:876-1080) with BC_halo_eps = twice the
// largest ElemRadiusNGeo.
std::string build_bc_tables(const pic_mesh_desc &m, HostMesh &hm) {
  const int ne = hm.nElems, ng = hm.NGeo, g1 = ng + 1, nGP = hm.nGP;
  auto SI = [&](int s, int c) { return m.SideInfo[size_t(s) * 5 + c]; };
  hm.cartPeriodic = hm.nPV > 0;
  int nb = 0;
  for (int sd = 0; sd < m.nSides; ++sd) {
    const int bc = SI(sd, 4);
    if (bc <= 0) continue;
    ++nb;
    if (m.PartBoundCond[bc - 1] != PIC_BC_PERIODIC) hm.cartPeriodic = false;
  }
  if (hm.cartPeriodic || nb == 0) return "";
  hm.nBC = nb;
  hm.bcElemNb.assign(nb, -1);
  hm.bcKind.assign(nb, 0);
  hm.bcPV.assign(nb, 0);
  hm.bcFlip.assign(nb, 0);
  hm.bz.assign(size_t(nb) * g1 * g1 * 3, 0.0);
  hm.sideType.assign(nb, 0);
  hm.bbEmpty.assign(nb, 0);
  hm.sideNorm.assign(size_t(nb) * 3, 0.0);
  hm.sideDist.assign(nb, 0.0);
  hm.slabN.assign(size_t(nb) * 9, 0.0);
  hm.slabI.assign(size_t(nb) * 6, 0.0);
  hm.baseV.assign(size_t(nb) * 9, 0.0);
  std::vector<double> origin(size_t(nb) * 3), bradius(nb);
  std::vector<int32_t> bcElem(nb), sideToBC(m.nSides, -1), sideElem(m.nSides, 0), sideLocal(m.nSides, 0);
  for (int e = 0; e < ne; ++e)
    for (int l = 0; l < 6; ++l) {
      const int sd = m.ElemInfo[size_t(e) * 6 + 2] + l;
      sideElem[sd] = e;
      sideLocal[sd] = l + 1;
    }
  double sV[3][3];
  bezier_svdm(ng, hm.xiCL, sV);
  double V1[8 * 2];
  {
    const double x1[2] = {-1.0, 1.0};
    double w1[2];
    bary_weights(1, x1, w1);
    for (int io = 0; io <= ng; ++io) lagrange(hm.xiCL[io], 1, x1, w1, &V1[io * 2]);
  }
  std::vector<double> X1(24), Xn(size_t(nGP) * 3);
  int b = 0;
  for (int sd = 0; sd < m.nSides; ++sd) {
    const int bcid = SI(sd, 4);
    if (bcid <= 0) continue;
    sideToBC[sd] = b;
    const int e = sideElem[sd], ls = sideLocal[sd];
    bcElem[b] = e;
    hm.bcElemNb[b] = SI(sd, 2) - 1;
    hm.bcKind[b] = uint8_t(m.PartBoundCond[bcid - 1]);
    hm.bcPV[b] = int8_t(m.BCType[(bcid - 1) * 4 + 3]);
    const int flip = SI(sd, 1) > 0 ? 0 : SI(sd, 3) % 10;
    hm.bcFlip[b] = uint8_t(flip);
    const double *X = &hm.XCL[size_t(e) * nGP * 3];
    double B[3][3][3];
    side_bezier(X, ng, ls, sV, B);
    for (int p = 0; p < g1; ++p) for (int q = 0; q < g1; ++q) for (int c = 0; c < 3; ++c)
      hm.bz[((size_t(b) * g1 + p) * g1 + q) * 3 + c] = B[p][q][c];
    double N[3][3], I[6];
    uint8_t empty = 0;
    if (!side_slab(ng, B, N, I, empty)) return "boundary side " + std::to_string(sd + 1) + ": degenerated side slab";
    for (int a = 0; a < 3; ++a) for (int c = 0; c < 3; ++c) hm.slabN[size_t(b) * 9 + a * 3 + c] = N[a][c];
    for (int c = 0; c < 6; ++c) hm.slabI[size_t(b) * 6 + c] = I[c];
    hm.bbEmpty[b] = empty;
    double *nv = &hm.sideNorm[size_t(b) * 3];
    double v1[3], v2[3];
    for (int c = 0; c < 3; ++c) {
      v1[c] = -B[0][0][c] + B[ng][0][c] - B[0][ng][c] + B[ng][ng][c];
      v2[c] = -B[0][0][c] - B[ng][0][c] + B[0][ng][c] + B[ng][ng][c];
    }
    auto orient = [&]() {   // outward for flip = 0, inward otherwise; SideDistance
      crossnorm3(v1, v2, nv);
      double c4[3];
      for (int c = 0; c < 3; ++c) c4[c] = 0.25 * (B[0][0][c] + B[ng][0][c] + B[0][ng][c] + B[ng][ng][c]);
      const double d[3] = {c4[0] - hm.bary[e * 3], c4[1] - hm.bary[e * 3 + 1], c4[2] - hm.bary[e * 3 + 2]};
      if (flip == 0) { if (dot3(d, nv) < 0) for (int c = 0; c < 3; ++c) nv[c] = -nv[c]; }
      else if (dot3(d, nv) > 0) for (int c = 0; c < 3; ++c) nv[c] = -nv[c];
      hm.sideDist[b] = dot3(c4, nv);
    };
    int type = -1;
    bool planar = false, signedTol = false;
    if (!hm.curved[e]) {
      if (empty) planar = true;
      else type = kSideBilinear;
    } else {
      for (int t = 0; t < 8; ++t) {
        const int i = (t & 1) * ng, j = ((t >> 1) & 1) * ng, k = ((t >> 2) & 1) * ng;
        for (int c = 0; c < 3; ++c) X1[t * 3 + c] = X[((k * g1 + j) * g1 + i) * 3 + c];
      }
      change_basis(1, ng, V1, X1.data(), Xn.data());
      double So[3][3][3], Sn[3][3][3];
      side_slice(X, g1, ls, So);
      side_slice(Xn.data(), g1, ls, Sn);
      bool curvedSide = false;
      for (int a = 0; a < g1 && !curvedSide; ++a)
        for (int bb = 0; bb < g1 && !curvedSide; ++bb)
          if (std::fabs(Sn[a][bb][0] - So[a][bb][0]) > 1e-14 || std::fabs(Sn[a][bb][1] - So[a][bb][1]) > 1e-14 ||
              std::fabs(Sn[a][bb][2] - So[a][bb][2]) > 1e-14)
            curvedSide = true;
      if (curvedSide) {
        if (empty) { type = kSidePlanarCurved; orient(); }
        else type = kSideCurved;
      } else if (empty) {
        planar = true;
        signedTol = true;
      } else {
        type = kSideBilinear;
      }
    }
    if (planar) {
      orient();
      double w1[3], w2[3], w3[3];
      for (int c = 0; c < 3; ++c) {
        w1[c] = B[0][ng][c] - B[0][0][c];
        w2[c] = B[ng][0][c] - B[0][0][c];
        w3[c] = B[ng][ng][c] - B[0][ng][c];
      }
      unit3(w1); unit3(w2); unit3(w3);
      // flat sides of curved elements test the signed product against 1e-14
      auto notOrtho = [&](const double *a, const double *c) {
        return signedTol ? (dot3(a, c) > 1e-14) : !(std::fabs(dot3(a, c)) <= 2.22e-16);
      };
      bool rect = !notOrtho(w1, w2) && !notOrtho(w1, w3);
      if (rect) {
        for (int c = 0; c < 3; ++c) w1[c] = B[ng][ng][c] - B[ng][0][c];
        unit3(w1);
        rect = !notOrtho(w1, w2) && !notOrtho(w1, w3);
      }
      type = rect ? kSidePlanarRect : kSidePlanarNonRect;
    }
    if (type == kSideBilinear || type == kSidePlanarNonRect)
      return "RefMapping with non-periodic boundaries: bilinear or non-rectangular planar boundary side " +
             std::to_string(sd + 1) + " is not supported";
    hm.sideType[b] = uint8_t(type);
    double *bv = &hm.baseV[size_t(b) * 9];
    for (int c = 0; c < 3; ++c) {
      bv[c] = (+B[0][0][c] + B[ng][0][c] + B[0][ng][c] + B[ng][ng][c]);
      bv[3 + c] = (-B[0][0][c] + B[ng][0][c] - B[0][ng][c] + B[ng][ng][c]);
      bv[6 + c] = (-B[0][0][c] - B[ng][0][c] + B[0][ng][c] + B[ng][ng][c]);
    }
    // BCSideMetrics: DeCasteljauInterpolation at xi = eta = 0 (basis.f90:429-476), radius
    double R[3][3][3];
    std::copy(&B[0][0][0], &B[0][0][0] + 27, &R[0][0][0]);
    const double Xi = 0.5, Eta = 0.5, MXi = 1.0 - Xi, MEta = 1.0 - Eta;
    int l = ng - 1;
    for (int it = 1; it <= ng; ++it) {
      for (int q = 0; q <= l; ++q)
        for (int p = 0; p <= l; ++p)
          for (int c = 0; c < 3; ++c)
            R[p][q][c] = MXi * R[p][q][c] * MEta + MXi * R[p][q + 1][c] * Eta + Xi * R[p + 1][q][c] * MEta +
                         Xi * R[p + 1][q + 1][c] * Eta;
      l = l - 1;
    }
    for (int c = 0; c < 3; ++c) origin[b * 3 + c] = R[0][0][c];
    double rmax = 0.;
    for (int q = 0; q <= ng; ++q)
      for (int p = 0; p <= ng; ++p) {
        const double d[3] = {B[p][q][0] - origin[b * 3], B[p][q][1] - origin[b * 3 + 1], B[p][q][2] - origin[b * 3 + 2]};
        rmax = std::fmax(rmax, dot3(d, d));
      }
    bradius[b] = std::sqrt(rmax);
    ++b;
  }
  double rmaxE = 0.0;
  for (int e = 0; e < ne; ++e) rmaxE = std::fmax(rmaxE, hm.radius[e]);
  hm.bcHalo = 2.0 * rmaxE;
  hm.e2bOff.assign(ne + 1, 0);
  hm.e2bSide.clear();
  hm.e2bDist.clear();
  std::vector<int32_t> lst;
  for (int e = 0; e < ne; ++e) {
    lst.clear();
    for (int l = 0; l < 6; ++l) {
      const int sd = m.ElemInfo[size_t(e) * 6 + 2] + l;
      if (SI(sd, 4) > 0) lst.push_back(sideToBC[sd]);
    }
    for (int q = 0; q < nb; ++q) {
      const int be = bcElem[q];
      if (be == e) continue;
      const double d[3] = {hm.bary[e * 3] - hm.bary[be * 3], hm.bary[e * 3 + 1] - hm.bary[be * 3 + 1],
                           hm.bary[e * 3 + 2] - hm.bary[be * 3 + 2]};
      if (std::sqrt(dot3(d, d)) > hm.bcHalo + hm.radius[e] + hm.radius[be]) continue;
      const double d2[3] = {hm.bary[e * 3] - origin[q * 3], hm.bary[e * 3 + 1] - origin[q * 3 + 1],
                            hm.bary[e * 3 + 2] - origin[q * 3 + 2]};
      if (std::sqrt(dot3(d2, d2)) <= hm.bcHalo + hm.radius[e] + bradius[q]) lst.push_back(q);
    }
    const size_t base = hm.e2bSide.size();
    hm.e2bOff[e] = int32_t(base);
    for (int32_t bs : lst) {
      const double d[3] = {hm.bary[e * 3] - origin[bs * 3], hm.bary[e * 3 + 1] - origin[bs * 3 + 1],
                           hm.bary[e * 3 + 2] - origin[bs * 3 + 2]};
      hm.e2bSide.push_back(bs);
      hm.e2bDist.push_back(std::sqrt(dot3(d, d)) - hm.radius[e] - bradius[bs]);
    }
    for (size_t a = 1; a < lst.size(); ++a) {   // InsertionSort by distance (stable)
      const double dv = hm.e2bDist[base + a];
      const int32_t sv = hm.e2bSide[base + a];
      long j = long(a) - 1;
      while (j >= 0 && hm.e2bDist[base + j] > dv) {
        hm.e2bDist[base + j + 1] = hm.e2bDist[base + j];
        hm.e2bSide[base + j + 1] = hm.e2bSide[base + j];
        --j;
      }
      hm.e2bDist[base + j + 1] = dv;
      hm.e2bSide[base + j + 1] = sv;
    }
  }
  hm.e2bOff[ne] = int32_t(hm.e2bSide.size());
  // the normal slab of each listed side: unit normal n (SideSlabNormals(:,2))
  // and the interval of n . X that holds the patch (its control points'
  // interval, SideSlabIntervals 3:4, widened by 1e-3 of the side's extent):
  // the device's exact no-hit test for a step
  hm.e2bSlab.resize(hm.e2bSide.size() * 6);
  for (size_t k = 0; k < hm.e2bSide.size(); ++k) {
    const int bs = hm.e2bSide[k];
    const double *N1 = &hm.slabN[size_t(bs) * 9 + 3], *I = &hm.slabI[size_t(bs) * 6];
    const double *B0 = &hm.bz[size_t(bs) * g1 * g1 * 3];
    const double c = dot3(N1, B0), mrg = 1e-3 * std::fmax(I[1] - I[0], I[5] - I[4]);
    double *r = &hm.e2bSlab[k * 6];
    r[0] = N1[0]; r[1] = N1[1]; r[2] = N1[2];
    r[3] = c + I[2] - mrg;
    r[4] = c + I[3] + mrg;
    r[5] = 0.0;
  }
  return "";
}

struct UnionFind {
  std::vector<int32_t> p;
  explicit UnionFind(int n) : p(n) { std::iota(p.begin(), p.end(), 0); }
  int find(int x) { while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; } return x; }
  void unite(int a, int b) { a = find(a); b = find(b); if (a != b) p[std::max(a, b)] = std::min(a, b); }
};

}  // namespace

std::string build_host_mesh(const pic_mesh_desc &m, const pic_params &par, HostMesh &hm) {
  const int ne = m.nElems;
  if (ne <= 0 || m.nSides != 6 * ne) return "mesh: expected 6 sides per element (no mortars)";
  if (m.NGeo < 1 || m.NGeo > kMaxNGeo) return "mesh: NGeo must be 1 or 2";
  const bool refmap = par.tracking == PIC_TRACK_REFMAPPING;
  if (!refmap && m.NGeo != 1) return "mesh: TriaTracking needs linear elements (NGeo = 1); use RefMapping for NGeo = 2";
  if (m.N < 1 || m.N > kMaxN) return "mesh: N out of range 1..7";
  hm.nElems = ne; hm.nSides = m.nSides; hm.nNodes = m.nNodes; hm.NGeo = m.NGeo; hm.N = m.N;
  hm.nDof = (m.N + 1) * (m.N + 1) * (m.N + 1);
  hm.nGP = (m.NGeo + 1) * (m.NGeo + 1) * (m.NGeo + 1);
  const int g1 = m.NGeo + 1, nGP = hm.nGP, nDof = hm.nDof;
  auto EI = [&](int e, int c) { return m.ElemInfo[size_t(e) * 6 + c]; };
  auto SI = [&](int s, int c) { return m.SideInfo[size_t(s) * 5 + c]; };

  hm.nPV = m.nPeriodicVectors;
  if (hm.nPV > 3) return "mesh: more than 3 periodic vectors";
  for (int v = 0; v < hm.nPV; ++v)
    for (int c = 0; c < 3; ++c) hm.PV[v][c] = m.PeriodicVectors[v * 3 + c];
  for (int b = 0; b < m.nBCs; ++b) {
    int t = m.PartBoundCond[b];
    if (t != PIC_BC_OPEN && t != PIC_BC_PERIODIC)
      return "mesh: particle boundary condition " + std::to_string(t) + " is not supported (open, periodic)";
  }

  // CGNS corners (mesh_tools.f90:740-769) as tensor corner indices t=i+2j+4k
  const int cgns_t[8] = {0, 1, 3, 2, 4, 5, 7, 6};
  const int side_cgns[6][4] = {{1, 4, 3, 2}, {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 4, 8, 7}, {1, 5, 8, 4}, {5, 6, 7, 8}};
  // tensor corner t -> element-local node number (0-based) for NGeo
  auto corner_node = [&](int t) {
    int i = (t & 1) * m.NGeo, j = ((t >> 1) & 1) * m.NGeo, k = ((t >> 2) & 1) * m.NGeo;
    return (k * g1 + j) * g1 + i;
  };

  hm.corner.assign(size_t(ne) * 24, 0.0);
  hm.sideNb.assign(size_t(ne) * 6, -1);
  hm.sideGID.assign(size_t(ne) * 6, 0);
  hm.sideNode.assign(size_t(ne) * 24, 0);
  hm.sideBC.assign(size_t(ne) * 6, 0);
  hm.sidePV.assign(size_t(ne) * 6, 0);
  hm.concave.assign(ne, 0);
  for (int c = 0; c < 3; ++c) { hm.lo[c] = 1e300; hm.hi[c] = -1e300; }

  for (int e = 0; e < ne; ++e) {
    const int fs = EI(e, 2), ls_ = EI(e, 3), fn = EI(e, 4), ln = EI(e, 5);
    if (ls_ - fs != 6) return "mesh: element with != 6 sides";
    if (ln - fn != nGP) return "mesh: element node count != (NGeo+1)^3";
    for (int t = 0; t < 8; ++t)
      for (int c = 0; c < 3; ++c) {
        double v = m.NodeCoords[size_t(fn + corner_node(t)) * 3 + c];
        hm.corner[size_t(e) * 24 + t * 3 + c] = v;
        hm.lo[c] = std::min(hm.lo[c], v);
        hm.hi[c] = std::max(hm.hi[c], v);
      }
    for (int l = 0; l < 6; ++l) {
      const int sd = fs + l;  // 0-based side index; SIDE_LOCALID = l+1 for non-mortar hexahedra
      if (SI(sd, 0) > 100) return "mesh: mortar sides are not supported";
      const int nb = SI(sd, 2);
      if (nb < 0) return "mesh: mortar neighbours are not supported";
      hm.sideNb[e * 6 + l] = nb - 1;
      hm.sideGID[e * 6 + l] = sd + 1;
      const int bcid = SI(sd, 4);
      if (bcid > 0) {
        hm.sideBC[e * 6 + l] = uint8_t(m.PartBoundCond[bcid - 1]);
        hm.sidePV[e * 6 + l] = int8_t(m.BCType[(bcid - 1) * 4 + 3]);
        if (m.PartBoundCond[bcid - 1] == PIC_BC_PERIODIC && (hm.sidePV[e * 6 + l] == 0 || std::abs(hm.sidePV[e * 6 + l]) > hm.nPV))
          return "mesh: periodic side without a valid periodic vector";
      } else if (nb == 0) {
        return "mesh: side without neighbour and without boundary condition";
      }
      // ElemSideNodeID with the flip rotation (particle_mesh_tools.f90:1874-1898)
      const int flip = SI(sd, 3) % 10;
      const int nStart = SI(sd, 1) > 0 ? 0 : std::max(0, flip - 1);
      for (int k = 0; k < 4; ++k)
        hm.sideNode[(e * 6 + l) * 4 + k] = uint8_t(cgns_t[side_cgns[l][(nStart + k) % 4] - 1]);
    }
    // ConcaveElemSide (particle_mesh_tools.f90:1904-1930)
    for (int l = 0; l < 6; ++l) {
      const uint8_t *sn = &hm.sideNode[(e * 6 + l) * 4];
      const double *C = &hm.corner[size_t(e) * 24];
      double A[3][3];
      for (int q = 0; q < 3; ++q)
        for (int c = 0; c < 3; ++c) A[c][q] = C[sn[q] * 3 + c] - C[sn[3] * 3 + c];
      double det = ((A[1][0] * A[2][1] - A[2][0] * A[1][1]) * A[0][2] +
                    (A[2][0] * A[0][1] - A[0][0] * A[2][1]) * A[1][2] +
                    (A[0][0] * A[1][1] - A[1][0] * A[0][1]) * A[2][2]);
      bool cc = false;
      if (det < 0) cc = true;
      else if (det == 0.0) {
        // GlobalSideID < SIDE_NBSIDEID (1-based; 0 when there is no neighbour)
        const int sd = fs + l, nb = SI(sd, 2);
        int nbside = 0;
        if (nb > 0) {
          const int nfs = EI(nb - 1, 2);
          for (int q = 0; q < 6; ++q)
            if (std::abs(SI(nfs + q, 1)) == std::abs(SI(sd, 1))) { nbside = nfs + q + 1; break; }
        }
        cc = (sd + 1) < nbside;
      }
      if (cc) hm.concave[e] |= uint8_t(1u << l);
    }
  }

  // 1D bases
  for (int i = 0; i <= m.NGeo; ++i) hm.xiCL[i] = -std::cos(double(i) / double(m.NGeo) * std::acos(-1.0));
  bary_weights(m.NGeo, hm.xiCL, hm.wBaryCL);
  gauss_nodes(m.N, hm.xGP, hm.wGP);
  bary_weights(m.N, hm.xGP, hm.wBaryGP);

  // XCL_NGeo (equidistant -> CL change of basis) and dXCL_NGeo = D x XCL
  double xEQ[8], wEQ[8], Vdm[64], D[64];
  for (int i = 0; i <= m.NGeo; ++i) xEQ[i] = 2.0 * double(i) / double(m.NGeo) - 1.0;
  bary_weights(m.NGeo, xEQ, wEQ);
  for (int io = 0; io <= m.NGeo; ++io) lagrange(hm.xiCL[io], m.NGeo, xEQ, wEQ, &Vdm[io * g1]);
  // PolynomialDerivativeMatrix on the CL nodes
  std::fill(D, D + 64, 0.0);
  for (int l = 0; l < g1; ++l)
    for (int i = 0; i < g1; ++i)
      if (l != i) {
        D[i * g1 + l] = hm.wBaryCL[l] / (hm.wBaryCL[i] * (hm.xiCL[i] - hm.xiCL[l]));
        D[i * g1 + i] -= D[i * g1 + l];
      }
  hm.XCL.assign(size_t(ne) * nGP * 3, 0.0);
  hm.dXCL.assign(size_t(ne) * nGP * 9, 0.0);
  for (int e = 0; e < ne; ++e) {
    const double *nc = &m.NodeCoords[size_t(EI(e, 4)) * 3];
    double *X = &hm.XCL[size_t(e) * nGP * 3];
    change_basis(m.NGeo, m.NGeo, Vdm, nc, X);
    double *dX = &hm.dXCL[size_t(e) * nGP * 9];
    for (int k = 0; k < g1; ++k) for (int j = 0; j < g1; ++j) for (int i = 0; i < g1; ++i) {
      const int p = (k * g1 + j) * g1 + i;
      for (int ll = 0; ll < g1; ++ll)
        for (int x = 0; x < 3; ++x) {
          dX[p * 9 + x * 3 + 0] += D[i * g1 + ll] * X[((k * g1 + j) * g1 + ll) * 3 + x];
          dX[p * 9 + x * 3 + 1] += D[j * g1 + ll] * X[((k * g1 + ll) * g1 + i) * 3 + x];
          dX[p * 9 + x * 3 + 2] += D[k * g1 + ll] * X[((ll * g1 + j) * g1 + i) * 3 + x];
        }
    }
  }

  // ElemBaryNGeo / ElemRadiusNGeo (TriaTracking flavour: corner mean, max corner distance)
  hm.bary.assign(size_t(ne) * 3, 0.0);
  hm.radius.assign(ne, 0.0);
  for (int e = 0; e < ne; ++e) {
    const double *nc = &m.NodeCoords[size_t(EI(e, 4)) * 3];
    double xp[3] = {0, 0, 0};
    for (int q = 0; q < 8; ++q)
      for (int c = 0; c < 3; ++c) xp[c] += nc[q * 3 + c];
    for (int c = 0; c < 3; ++c) hm.bary[e * 3 + c] = xp[c] / 8.0;
    double r = 0;
    for (int q = 0; q < 8; ++q) {
      double d0 = nc[q * 3] - hm.bary[e * 3], d1 = nc[q * 3 + 1] - hm.bary[e * 3 + 1], d2 = nc[q * 3 + 2] - hm.bary[e * 3 + 2];
      r = std::max(r, std::sqrt(d0 * d0 + d1 * d1 + d2 * d2));
    }
    hm.radius[e] = r;
  }

  // XiEtaZetaBasis / slenXiEtaZetaBasis
  hm.xiBasis.assign(size_t(ne) * 18, 0.0);
  hm.slen.assign(size_t(ne) * 6, 0.0);
  {
    const double Xi[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
    double L[3][8];
    for (int e = 0; e < ne; ++e) {
      const double *X = &hm.XCL[size_t(e) * nGP * 3];
      for (int d = 0; d < 6; ++d) {
        for (int q = 0; q < 3; ++q) lagrange(Xi[d][q], m.NGeo, hm.xiCL, hm.wBaryCL, L[q]);
        double xp[3] = {0, 0, 0};
        for (int k = 0; k < g1; ++k) for (int j = 0; j < g1; ++j) for (int i = 0; i < g1; ++i)
          for (int c = 0; c < 3; ++c) xp[c] = xp[c] + X[((k * g1 + j) * g1 + i) * 3 + c] * L[0][i] * L[1][j] * L[2][k];
        double *b = &hm.xiBasis[(e * 6 + d) * 3];
        for (int c = 0; c < 3; ++c) b[c] = xp[c] - hm.bary[e * 3 + c];
        hm.slen[e * 6 + d] = 1.0 / (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      }
    }
  }

  // Elem_xGP and sJ at the Gauss points (mapping and Jacobian of XCL_NGeo)
  hm.xGPe.assign(size_t(ne) * nDof * 3, 0.0);
  hm.sJ.assign(size_t(ne) * nDof, 0.0);
  {
    double Vg[8 * (kMaxN + 1)];
    for (int io = 0; io <= m.N; ++io) lagrange(hm.xGP[io], m.NGeo, hm.xiCL, hm.wBaryCL, &Vg[io * g1]);
    std::vector<double> din(size_t(nGP) * 3), dout(size_t(nDof) * 3), J(size_t(nDof) * 9);
    for (int e = 0; e < ne; ++e) {
      change_basis(m.NGeo, m.N, Vg, &hm.XCL[size_t(e) * nGP * 3], &hm.xGPe[size_t(e) * nDof * 3]);
      for (int d = 0; d < 3; ++d) {
        for (int p = 0; p < nGP; ++p)
          for (int x = 0; x < 3; ++x) din[p * 3 + x] = hm.dXCL[(size_t(e) * nGP + p) * 9 + x * 3 + d];
        change_basis(m.NGeo, m.N, Vg, din.data(), dout.data());
        for (int p = 0; p < nDof; ++p)
          for (int x = 0; x < 3; ++x) J[p * 9 + x * 3 + d] = dout[p * 3 + x];
      }
      for (int p = 0; p < nDof; ++p) {
        const double *Jp = &J[p * 9];
        auto dx = [&](int d, int x) { return Jp[x * 3 + d]; };
        double det = dx(0, 0) * (dx(1, 1) * dx(2, 2) - dx(2, 1) * dx(1, 2)) +
                     dx(1, 0) * (dx(2, 1) * dx(0, 2) - dx(0, 1) * dx(2, 2)) +
                     dx(2, 0) * (dx(0, 1) * dx(1, 2) - dx(1, 1) * dx(0, 2));
        if (!(det > 0)) return "mesh: non-positive Jacobian at a Gauss point of element " + std::to_string(e + 1);
        hm.sJ[size_t(e) * nDof + p] = 1.0 / det;
      }
    }
  }

  // ---- reference-mapping tables ---------------------------------------------
  // ElemCurved: the trilinear map of the 8 corner CL points differs from
  // XCL_NGeo by more than 1e-14 (particle_mesh_tools.f90:976-996); dXCL_NGeo
  // at the corners drives the trilinear Newton of the other elements.
  hm.curved.assign(ne, 0);
  hm.dXCL1.assign(size_t(ne) * 72, 0.0);
  {
    double V1[8 * 2];
    const double x1[2] = {-1.0, 1.0};
    double w1[2];
    bary_weights(1, x1, w1);
    for (int io = 0; io <= m.NGeo; ++io) lagrange(hm.xiCL[io], 1, x1, w1, &V1[io * 2]);
    std::vector<double> X1(24), Xn(size_t(nGP) * 3);
    for (int e = 0; e < ne; ++e) {
      const double *X = &hm.XCL[size_t(e) * nGP * 3];
      for (int t = 0; t < 8; ++t) {
        const int p = corner_node(t);
        for (int c = 0; c < 3; ++c) X1[t * 3 + c] = X[p * 3 + c];
        for (int q = 0; q < 9; ++q) hm.dXCL1[size_t(e) * 72 + t * 9 + q] = hm.dXCL[(size_t(e) * nGP + p) * 9 + q];
      }
      if (m.NGeo == 1) continue;
      change_basis(1, m.NGeo, V1, X1.data(), Xn.data());
      for (int p = 0; p < nGP && !hm.curved[e]; ++p)
        for (int c = 0; c < 3; ++c)
          if (std::fabs(Xn[p * 3 + c] - X[p * 3 + c]) > 1e-14) { hm.curved[e] = 1; break; }
    }
  }
  hm.openElem.assign(ne, 0);
  for (int e = 0; e < ne; ++e)
    for (int l = 0; l < 6; ++l)
      if (hm.sideBC[e * 6 + l] == PIC_BC_OPEN) hm.openElem[e] = 1;
  // GEO%x/y/zminglob, maxglob over the mesh nodes (TriaTracking,
  // particle_mesh_tools.f90:775-783; RefMapping replaces them below)
  for (int c = 0; c < 3; ++c) { hm.gmin[c] = 1e300; hm.gmax[c] = -1e300; }
  for (int i = 0; i < m.nNodes; ++i)
    for (int c = 0; c < 3; ++c) {
      hm.gmin[c] = std::min(hm.gmin[c], m.NodeCoords[size_t(i) * 3 + c]);
      hm.gmax[c] = std::max(hm.gmax[c], m.NodeCoords[size_t(i) * 3 + c]);
    }
  for (int v = 0; v < hm.nPV; ++v) {
    hm.dirPV[v] = -1;
    for (int c = 0; c < 3; ++c)
      if (hm.PV[v][c] != 0.0) {
        if (hm.dirPV[v] >= 0) return "mesh: periodic vectors must be aligned with an axis";
        hm.dirPV[v] = c;
      }
  }
  if (refmap) {
    // ElemBaryNGeo = X(0,0,0) (BuildElementOriginShared, particle_mesh_build.f90:1385-1475)
    double L0[8];
    lagrange(0.0, m.NGeo, hm.xiCL, hm.wBaryCL, L0);
    for (int e = 0; e < ne; ++e) {
      const double *X = &hm.XCL[size_t(e) * nGP * 3];
      double xp[3] = {0, 0, 0};
      for (int k = 0; k < g1; ++k) for (int j = 0; j < g1; ++j) {
        const double buf = L0[j] * L0[k];
        for (int i = 0; i < g1; ++i)
          for (int c = 0; c < 3; ++c) xp[c] = xp[c] + X[((k * g1 + j) * g1 + i) * 3 + c] * L0[i] * buf;
      }
      for (int c = 0; c < 3; ++c) hm.bary[e * 3 + c] = xp[c];
    }
    // XiEtaZetaBasis from that origin, ElemRadiusNGeo = largest distance of a
    // Bezier control point of the six faces (BuildElementBasisAndRadius,
    // particle_mesh_build.f90:1546-1685)
    const double Xi[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
    double sVb[9];
    {
      double A[3][6];
      for (int a = 0; a < g1; ++a)
        for (int i = 0; i < g1; ++i) {
          double binom = 1.0;
          for (int q = 1; q <= i; ++q) binom = binom * double(m.NGeo - q + 1) / double(q);
          A[a][i] = binom * std::pow(0.5 * (1.0 + hm.xiCL[a]), i) * std::pow(0.5 * (1.0 - hm.xiCL[a]), m.NGeo - i);
          A[a][g1 + i] = a == i ? 1.0 : 0.0;
        }
      for (int c = 0; c < g1; ++c) {  // Gauss-Jordan with partial pivoting
        int piv = c;
        for (int r = c + 1; r < g1; ++r) if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
        for (int q = 0; q < 2 * g1; ++q) std::swap(A[c][q], A[piv][q]);
        const double d = A[c][c];
        for (int q = 0; q < 2 * g1; ++q) A[c][q] /= d;
        for (int r = 0; r < g1; ++r)
          if (r != c) { const double f = A[r][c]; for (int q = 0; q < 2 * g1; ++q) A[r][q] -= f * A[c][q]; }
      }
      for (int r = 0; r < g1; ++r) for (int c = 0; c < g1; ++c) sVb[r * g1 + c] = A[r][g1 + c];
    }
    double bzmin[3] = {1e300, 1e300, 1e300}, bzmax[3] = {-1e300, -1e300, -1e300};
    for (int e = 0; e < ne; ++e) {
      const double *X = &hm.XCL[size_t(e) * nGP * 3];
      double L[3][8];
      for (int d = 0; d < 6; ++d) {
        for (int q = 0; q < 3; ++q) lagrange(Xi[d][q], m.NGeo, hm.xiCL, hm.wBaryCL, L[q]);
        double xp[3] = {0, 0, 0};
        for (int k = 0; k < g1; ++k) for (int j = 0; j < g1; ++j) for (int i = 0; i < g1; ++i)
          for (int c = 0; c < 3; ++c) xp[c] = xp[c] + X[((k * g1 + j) * g1 + i) * 3 + c] * L[0][i] * L[1][j] * L[2][k];
        double *b = &hm.xiBasis[(e * 6 + d) * 3];
        for (int c = 0; c < 3; ++c) b[c] = xp[c] - hm.bary[e * 3 + c];
        hm.slen[e * 6 + d] = 1.0 / (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      }
      double r = 0.0;
      for (int f = 0; f < 6; ++f) {
        const int ax = f % 3, fixed = f < 3 ? m.NGeo : 0;
        double P[3][3][3], T[3][3][3];
        for (int a = 0; a < g1; ++a) for (int bq = 0; bq < g1; ++bq) {
          int ijk[3];
          ijk[ax] = fixed; ijk[(ax + 1) % 3] = a; ijk[(ax + 2) % 3] = bq;
          for (int c = 0; c < 3; ++c) P[a][bq][c] = X[((ijk[2] * g1 + ijk[1]) * g1 + ijk[0]) * 3 + c];
        }
        for (int a = 0; a < g1; ++a) for (int bq = 0; bq < g1; ++bq) for (int c = 0; c < 3; ++c) {
          double acc = 0.0;
          for (int q = 0; q < g1; ++q) acc += sVb[a * g1 + q] * P[q][bq][c];
          T[a][bq][c] = acc;
        }
        for (int a = 0; a < g1; ++a) for (int bq = 0; bq < g1; ++bq) {
          double B[3];
          for (int c = 0; c < 3; ++c) {
            double acc = 0.0;
            for (int q = 0; q < g1; ++q) acc += sVb[bq * g1 + q] * T[a][q][c];
            B[c] = acc;
          }
          const double d0 = B[0] - hm.bary[e * 3], d1 = B[1] - hm.bary[e * 3 + 1], d2 = B[2] - hm.bary[e * 3 + 2];
          r = std::max(r, std::sqrt(d0 * d0 + d1 * d1 + d2 * d2));
          for (int c = 0; c < 3; ++c) { bzmin[c] = std::min(bzmin[c], B[c]); bzmax[c] = std::max(bzmax[c], B[c]); }
        }
      }
      hm.radius[e] = r;
    }
    // with RefMapping GEO%x/y/zminglob, maxglob bound the sides' Bezier
    // control points (GetMeshMinMax, particle_mesh_tools.f90:735-747)
    for (int c = 0; c < 3; ++c) { hm.gmin[c] = bzmin[c]; hm.gmax[c] = bzmax[c]; }
    // ElemEpsOneCell = 1 + sqrt(3 scaleJ RefMappingEps), scaleJ = max sJ / min sJ (BuildEpsOneCell)
    hm.epsOneCell.assign(ne, 0.0);
    for (int e = 0; e < ne; ++e) {
      double mx = -1e300, mn = 1e300;
      for (int p = 0; p < nDof; ++p) {
        mx = std::max(mx, hm.sJ[size_t(e) * nDof + p]);
        mn = std::min(mn, hm.sJ[size_t(e) * nDof + p]);
      }
      hm.epsOneCell[e] = 1.0 + std::sqrt(3.0 * (mx / mn) * par.refmapping_eps);
    }
    // background cells over the bounding spheres; cell size = the largest extent
    double ext = 0.0;
    for (int c = 0; c < 3; ++c) { hm.bmin[c] = 1e300; }
    double bmax[3] = {-1e300, -1e300, -1e300};
    for (int e = 0; e < ne; ++e)
      for (int c = 0; c < 3; ++c) {
        hm.bmin[c] = std::min(hm.bmin[c], hm.bary[e * 3 + c] - hm.radius[e]);
        bmax[c] = std::max(bmax[c], hm.bary[e * 3 + c] + hm.radius[e]);
        ext = std::max(ext, 2.0 * hm.radius[e]);
      }
    for (int c = 0; c < 3; ++c) {
      const double L = bmax[c] - hm.bmin[c];
      hm.nb[c] = std::max(1, int(std::ceil(L / ext)));
      hm.bdelta[c] = L / hm.nb[c];
      if (!(hm.bdelta[c] > 0)) hm.bdelta[c] = 1.0;
    }
    const size_t nbins = size_t(hm.nb[0]) * hm.nb[1] * hm.nb[2];
    hm.binOff.assign(nbins + 1, 0);
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<int32_t> fill(pass ? nbins : 0, 0);
      for (int e = 0; e < ne; ++e) {
        int lo3[3], hi3[3];
        for (int c = 0; c < 3; ++c) {
          lo3[c] = std::max(0, int(std::floor((hm.bary[e * 3 + c] - hm.radius[e] - hm.bmin[c]) / hm.bdelta[c])));
          hi3[c] = std::min(hm.nb[c] - 1, int(std::floor((hm.bary[e * 3 + c] + hm.radius[e] - hm.bmin[c]) / hm.bdelta[c])));
        }
        for (int z = lo3[2]; z <= hi3[2]; ++z) for (int y = lo3[1]; y <= hi3[1]; ++y) for (int x = lo3[0]; x <= hi3[0]; ++x) {
          const size_t b = (size_t(z) * hm.nb[1] + y) * hm.nb[0] + x;
          if (!pass) hm.binOff[b + 1]++;
          else hm.binElems[hm.binOff[b] + fill[b]++] = e;
        }
      }
      if (!pass) {
        for (size_t q = 0; q < nbins; ++q) hm.binOff[q + 1] += hm.binOff[q];
        hm.binElems.resize(hm.binOff[nbins]);
      }
    }
    for (int e = 0; e < ne; ++e)  // the domain box holds the curved elements
      for (int c = 0; c < 3; ++c) {
        hm.lo[c] = std::min(hm.lo[c], hm.bary[e * 3 + c] - hm.radius[e]);
        hm.hi[c] = std::max(hm.hi[c], hm.bary[e * 3 + c] + hm.radius[e]);
      }
    for (size_t q = 0; q < nbins; ++q)
      if (hm.binOff[q + 1] - hm.binOff[q] > kMaxBinElems)
        return "RefMapping: a background cell holds more than " + std::to_string(kMaxBinElems) + " elements";
    std::string err = build_bc_tables(m, hm);
    if (!err.empty()) return err;
  }

  if (!refmap) {
    // background cells for SinglePointToElement with TriaTracking: elements
    // binned by the bounding box of their nodes; cell size = largest extent
    std::vector<double> emin(size_t(ne) * 3, 1e300), emax(size_t(ne) * 3, -1e300);
    double ext = 0.0, bmax[3] = {-1e300, -1e300, -1e300};
    for (int c = 0; c < 3; ++c) hm.bmin[c] = 1e300;
    for (int e = 0; e < ne; ++e) {
      const int fn = EI(e, 4), ln = EI(e, 5);
      for (int q = fn; q < ln; ++q)
        for (int c = 0; c < 3; ++c) {
          emin[e * 3 + c] = std::min(emin[e * 3 + c], m.NodeCoords[size_t(q) * 3 + c]);
          emax[e * 3 + c] = std::max(emax[e * 3 + c], m.NodeCoords[size_t(q) * 3 + c]);
        }
      for (int c = 0; c < 3; ++c) {
        hm.bmin[c] = std::min(hm.bmin[c], emin[e * 3 + c]);
        bmax[c] = std::max(bmax[c], emax[e * 3 + c]);
        ext = std::max(ext, emax[e * 3 + c] - emin[e * 3 + c]);
      }
    }
    for (int c = 0; c < 3; ++c) {
      const double L = bmax[c] - hm.bmin[c];
      hm.nb[c] = std::max(1, int(std::ceil(L / ext)));
      hm.bdelta[c] = L / hm.nb[c];
      if (!(hm.bdelta[c] > 0)) hm.bdelta[c] = 1.0;
    }
    const size_t nbins = size_t(hm.nb[0]) * hm.nb[1] * hm.nb[2];
    hm.binOff.assign(nbins + 1, 0);
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<int32_t> fill(pass ? nbins : 0, 0);
      for (int e = 0; e < ne; ++e) {
        int lo3[3], hi3[3];
        for (int c = 0; c < 3; ++c) {
          lo3[c] = std::max(0, int(std::floor((emin[e * 3 + c] - hm.bmin[c]) / hm.bdelta[c])));
          hi3[c] = std::min(hm.nb[c] - 1, int(std::floor((emax[e * 3 + c] - hm.bmin[c]) / hm.bdelta[c])));
        }
        for (int z = lo3[2]; z <= hi3[2]; ++z) for (int y = lo3[1]; y <= hi3[1]; ++y) for (int x = lo3[0]; x <= hi3[0]; ++x) {
          const size_t b = (size_t(z) * hm.nb[1] + y) * hm.nb[0] + x;
          if (!pass) hm.binOff[b + 1]++;
          else hm.binElems[hm.binOff[b] + fill[b]++] = e;
        }
      }
      if (!pass) {
        for (size_t q = 0; q < nbins; ++q) hm.binOff[q + 1] += hm.binOff[q];
        hm.binElems.resize(hm.binOff[nbins]);
      }
    }
  }

  // ---- cell_volweight_mean node classes ------------------------------------
  int maxid = 0;
  for (int i = 0; i < m.nNodes; ++i) maxid = std::max(maxid, m.GlobalNodeIDs[i]);
  hm.nUnique = maxid;
  std::vector<double> ux(size_t(maxid) * 3, 0.0);
  for (int i = 0; i < m.nNodes; ++i)
    for (int c = 0; c < 3; ++c) ux[size_t(m.GlobalNodeIDs[i] - 1) * 3 + c] = m.NodeCoords[size_t(i) * 3 + c];
  UnionFind uf(maxid);
  if (hm.nPV > 0) {
    double scale = 0;
    for (int e = 0; e < ne; ++e) scale = std::max(scale, hm.radius[e]);
    const double tol = 1e-9 * (scale > 0 ? scale : 1.0), q = 1e-6 * (scale > 0 ? scale : 1.0);
    auto key = [&](double x, double y, double z) {
      int64_t a = int64_t(std::floor(x / q)), b = int64_t(std::floor(y / q)), c = int64_t(std::floor(z / q));
      return (uint64_t(a) * 73856093ull) ^ (uint64_t(b) * 19349663ull) ^ (uint64_t(c) * 83492791ull);
    };
    std::unordered_multimap<uint64_t, int32_t> grid;
    grid.reserve(size_t(maxid) * 2);
    for (int u = 0; u < maxid; ++u) grid.emplace(key(ux[u * 3], ux[u * 3 + 1], ux[u * 3 + 2]), u);
    int ncomb = 1;
    for (int v = 0; v < hm.nPV; ++v) ncomb *= 3;
    for (int u = 0; u < maxid; ++u)
      for (int comb = 0; comb < ncomb; ++comb) {
        int cc = comb, co[3] = {0, 0, 0};
        for (int v = 0; v < hm.nPV; ++v) { co[v] = cc % 3 - 1; cc /= 3; }
        double p[3];
        for (int c = 0; c < 3; ++c) {
          p[c] = ux[u * 3 + c];
          for (int v = 0; v < hm.nPV; ++v) p[c] += co[v] * hm.PV[v][c];
        }
        for (int dz = -1; dz <= 1; ++dz) for (int dy = -1; dy <= 1; ++dy) for (int dx = -1; dx <= 1; ++dx) {
          auto rg = grid.equal_range(key(p[0] + dx * q, p[1] + dy * q, p[2] + dz * q));
          for (auto it = rg.first; it != rg.second; ++it) {
            int w = it->second;
            if (std::fabs(ux[w * 3] - p[0]) <= tol && std::fabs(ux[w * 3 + 1] - p[1]) <= tol &&
                std::fabs(ux[w * 3 + 2] - p[2]) <= tol)
              uf.unite(u, w);
          }
        }
      }
  }
  std::vector<int32_t> cls(maxid, -1);
  int ncl = 0;
  for (int u = 0; u < maxid; ++u) {
    int r = uf.find(u);
    if (cls[r] < 0) cls[r] = ncl++;
    cls[u] = cls[r];
  }
  hm.nClasses = ncl;
  hm.elemClass.assign(size_t(ne) * 8, 0);
  hm.elemNode.assign(size_t(ne) * 8, 0);
  std::vector<double> nodeVol(maxid, 0.0);
  {
    const int n1 = m.N + 1;
    const double *x = hm.xGP, *w = hm.wGP;
    for (int e = 0; e < ne; ++e) {
      const int fn = EI(e, 4);
      int uid[8];
      for (int c = 0; c < 8; ++c) {
        uid[c] = m.GlobalNodeIDs[fn + corner_node(cgns_t[c])] - 1;
        hm.elemClass[size_t(e) * 8 + c] = cls[uid[c]];
        hm.elemNode[size_t(e) * 8 + c] = uid[c];
      }
      // CalcCellLocNodeVolumes (pic_depo_tools.f90:291-303)
      for (int i = 0; i < n1; ++i) for (int j = 0; j < n1; ++j) for (int k = 0; k < n1; ++k) {
        const double s = hm.sJ[size_t(e) * nDof + (k * n1 + j) * n1 + i];
        const double wt = w[i] * w[j] * w[k] / 8.;
        nodeVol[uid[0]] += 1. / s * ((1. - x[i]) * (1. - x[j]) * (1. - x[k]) * wt);
        nodeVol[uid[1]] += 1. / s * ((1. + x[i]) * (1. - x[j]) * (1. - x[k]) * wt);
        nodeVol[uid[2]] += 1. / s * ((1. + x[i]) * (1. + x[j]) * (1. - x[k]) * wt);
        nodeVol[uid[3]] += 1. / s * ((1. - x[i]) * (1. + x[j]) * (1. - x[k]) * wt);
        nodeVol[uid[4]] += 1. / s * ((1. - x[i]) * (1. - x[j]) * (1. + x[k]) * wt);
        nodeVol[uid[5]] += 1. / s * ((1. + x[i]) * (1. - x[j]) * (1. + x[k]) * wt);
        nodeVol[uid[6]] += 1. / s * ((1. + x[i]) * (1. + x[j]) * (1. + x[k]) * wt);
        nodeVol[uid[7]] += 1. / s * ((1. - x[i]) * (1. + x[j]) * (1. + x[k]) * wt);
      }
    }
  }
  // class volume = own volume + twins' volumes (periodic merge, pic_depo_tools.f90:319-338)
  hm.classVolume.assign(ncl, 0.0);
  for (int u = 0; u < maxid; ++u) hm.classVolume[cls[u]] += nodeVol[u];
  // incidence CSR, class -> (elem*8 + corner), in element order
  hm.classOff.assign(ncl + 1, 0);
  for (size_t q = 0; q < hm.elemClass.size(); ++q) hm.classOff[hm.elemClass[q] + 1]++;
  for (int c = 0; c < ncl; ++c) hm.classOff[c + 1] += hm.classOff[c];
  hm.classInc.assign(hm.elemClass.size(), 0);
  {
    std::vector<int32_t> fill(ncl, 0);
    for (size_t q = 0; q < hm.elemClass.size(); ++q) {
      int c = hm.elemClass[q];
      hm.classInc[hm.classOff[c] + fill[c]++] = int32_t(q);
    }
  }

  // unique node -> (element, corner) incidences, in element order
  hm.nodeOff.assign(size_t(maxid) + 1, 0);
  for (size_t q = 0; q < hm.elemNode.size(); ++q) hm.nodeOff[hm.elemNode[q] + 1]++;
  for (int u = 0; u < maxid; ++u) hm.nodeOff[u + 1] += hm.nodeOff[u];
  hm.nodeInc.assign(hm.elemNode.size(), 0);
  {
    std::vector<int32_t> fill(maxid, 0);
    for (size_t q = 0; q < hm.elemNode.size(); ++q) {
      const int u = hm.elemNode[q];
      hm.nodeInc[hm.nodeOff[u] + fill[u]++] = int32_t(q);
    }
  }

  // ---- shape-function source lists -----------------------------------------
  if (par.deposition == PIC_DEPO_SF) {
    const double R = par.sf_radius;
    if (!(R > 0)) return "shape_function: radius must be positive";
    // shift table: the periodic cases (pic_depo.f90:858-888): every combination
    // of -1/0/+1 times each periodic vector (27 for three vectors, 3^nPV for
    // fewer, where the reference's fixed 3D table would use absent vectors)
    int nsh = 1;
    for (int v = 0; v < hm.nPV; ++v) nsh *= 3;
    hm.nShift = nsh;
    for (int s = 0; s < 27; ++s) for (int c = 0; c < 3; ++c) hm.shift[s][c] = 0.0;
    for (int s = 0; s < nsh && hm.nPV > 0; ++s) {
      int cc = s, co[3] = {0, 0, 0};
      for (int v = hm.nPV - 1; v >= 0; --v) { co[v] = cc % 3 - 1; cc /= 3; }
      for (int c = 0; c < 3; ++c) hm.shift[s][c] = co[0] * hm.PV[0][c] + co[1] * hm.PV[1][c] + co[2] * hm.PV[2][c];
    }
    // element bounding boxes (sources: all nodes; targets: Gauss points)
    std::vector<double> sbb(size_t(ne) * 6), tbb(size_t(ne) * 6);
    double ext = 0;
    for (int e = 0; e < ne; ++e) {
      const double *nc = &m.NodeCoords[size_t(EI(e, 4)) * 3];
      for (int c = 0; c < 3; ++c) { sbb[e * 6 + c] = 1e300; sbb[e * 6 + 3 + c] = -1e300; tbb[e * 6 + c] = 1e300; tbb[e * 6 + 3 + c] = -1e300; }
      if (m.NGeo == 1)
        for (int q = 0; q < nGP; ++q)
          for (int c = 0; c < 3; ++c) {
            sbb[e * 6 + c] = std::min(sbb[e * 6 + c], nc[q * 3 + c]);
            sbb[e * 6 + 3 + c] = std::max(sbb[e * 6 + 3 + c], nc[q * 3 + c]);
          }
      else  // a curved element lies inside its bounding sphere
        for (int c = 0; c < 3; ++c) {
          sbb[e * 6 + c] = hm.bary[e * 3 + c] - hm.radius[e];
          sbb[e * 6 + 3 + c] = hm.bary[e * 3 + c] + hm.radius[e];
        }
      for (int q = 0; q < nDof; ++q)
        for (int c = 0; c < 3; ++c) {
          double v = hm.xGPe[(size_t(e) * nDof + q) * 3 + c];
          tbb[e * 6 + c] = std::min(tbb[e * 6 + c], v);
          tbb[e * 6 + 3 + c] = std::max(tbb[e * 6 + 3 + c], v);
        }
      for (int c = 0; c < 3; ++c) ext = std::max(ext, sbb[e * 6 + 3 + c] - sbb[e * 6 + c]);
    }
    // bins over source boxes
    int nb[3];
    double bd[3];
    for (int c = 0; c < 3; ++c) {
      double L = hm.hi[c] - hm.lo[c];
      nb[c] = std::max(1, int(std::ceil(L / ext)));
      bd[c] = L > 0 ? L / nb[c] : 1.0;
    }
    const size_t nbins = size_t(nb[0]) * nb[1] * nb[2];
    std::vector<int32_t> boff(nbins + 1, 0), belem;
    auto binr = [&](double lo_, double hi_, int c, int &a, int &b) {
      a = int(std::floor((lo_ - hm.lo[c]) / bd[c]));
      b = int(std::floor((hi_ - hm.lo[c]) / bd[c]));
      a = std::max(a, 0); b = std::min(b, nb[c] - 1);
    };
    for (int pass = 0; pass < 2; ++pass) {
      std::vector<int32_t> fill(pass ? nbins : 0, 0);
      for (int e = 0; e < ne; ++e) {
        int a[3], b[3];
        for (int c = 0; c < 3; ++c) binr(sbb[e * 6 + c], sbb[e * 6 + 3 + c], c, a[c], b[c]);
        for (int z = a[2]; z <= b[2]; ++z) for (int y = a[1]; y <= b[1]; ++y) for (int x = a[0]; x <= b[0]; ++x) {
          size_t bi = (size_t(z) * nb[1] + y) * nb[0] + x;
          if (!pass) boff[bi + 1]++;
          else belem[boff[bi] + fill[bi]++] = e;
        }
      }
      if (!pass) {
        for (size_t q = 0; q < nbins; ++q) boff[q + 1] += boff[q];
        belem.resize(boff[nbins]);
      }
    }
    const double pad = 1e-9 * ext + 0.0;
    hm.sfOff.assign(ne + 1, 0);
    hm.sfSrc.clear();
    hm.sfShift.clear();
    hm.sfRunOff.assign(ne + 1, 0);
    hm.sfRunFirst.clear();
    hm.sfRunEnd.clear();
    hm.sfRunShift.clear();
    std::vector<int64_t> cand;
    for (int t = 0; t < ne; ++t) {
      cand.clear();
      for (int s = 0; s < nsh; ++s) {
        // query box: target Gauss-point box minus shift, dilated by R
        double qlo[3], qhi[3];
        bool any = true;
        for (int c = 0; c < 3; ++c) {
          qlo[c] = tbb[t * 6 + c] - hm.shift[s][c] - R - pad;
          qhi[c] = tbb[t * 6 + 3 + c] - hm.shift[s][c] + R + pad;
          if (qhi[c] < hm.lo[c] || qlo[c] > hm.hi[c]) any = false;
        }
        if (!any) continue;
        int a[3], b[3];
        for (int c = 0; c < 3; ++c) binr(qlo[c], qhi[c], c, a[c], b[c]);
        for (int z = a[2]; z <= b[2]; ++z) for (int y = a[1]; y <= b[1]; ++y) for (int x = a[0]; x <= b[0]; ++x) {
          size_t bi = (size_t(z) * nb[1] + y) * nb[0] + x;
          for (int q = boff[bi]; q < boff[bi + 1]; ++q) {
            int se = belem[q];
            // box-to-box distance between shifted source box and target Gauss-point box
            double d2 = 0;
            for (int c = 0; c < 3; ++c) {
              double slo = sbb[se * 6 + c] + hm.shift[s][c], shi = sbb[se * 6 + 3 + c] + hm.shift[s][c];
              double g = std::max(0.0, std::max(slo - tbb[t * 6 + 3 + c], tbb[t * 6 + c] - shi));
              d2 += g * g;
            }
            if (d2 <= (R + pad) * (R + pad)) cand.push_back(int64_t(se) * 32 + s);
          }
        }
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (int64_t v : cand) { hm.sfSrc.push_back(int32_t(v / 32)); hm.sfShift.push_back(uint8_t(v % 32)); }
      hm.sfOff[t + 1] = int32_t(hm.sfSrc.size());
      // runs: by shift, then element; consecutive elements of one shift merge
      for (int64_t &v : cand) v = (v % 32) * (int64_t(1) << 32) + v / 32;
      std::sort(cand.begin(), cand.end());
      for (size_t a = 0; a < cand.size();) {
        const int64_t sh = cand[a] >> 32;
        const int32_t first = int32_t(cand[a] & 0xffffffff);
        size_t b = a + 1;
        while (b < cand.size() && (cand[b] >> 32) == sh && int32_t(cand[b] & 0xffffffff) == first + int32_t(b - a)) ++b;
        hm.sfRunFirst.push_back(first);
        hm.sfRunEnd.push_back(first + int32_t(b - a));
        hm.sfRunShift.push_back(uint8_t(sh));
        a = b;
      }
      hm.sfRunOff[t + 1] = int32_t(hm.sfRunFirst.size());
    }
    // runs of the union of the lists of four consecutive targets (the
    // workgroup-cooperative gather stages each of their records once for
    // the four), ordered as every target's own runs: by shift, then element
    const int ng4 = (ne + 3) / 4;
    hm.sfGrpRunOff.assign(ng4 + 1, 0);
    hm.sfGrpRunFirst.clear();
    hm.sfGrpRunEnd.clear();
    hm.sfGrpRunShift.clear();
    for (int g = 0; g < ng4; ++g) {
      cand.clear();
      for (int t = 4 * g; t < std::min(ne, 4 * g + 4); ++t)
        for (int q = hm.sfOff[t]; q < hm.sfOff[t + 1]; ++q)
          cand.push_back(int64_t(hm.sfShift[q]) * (int64_t(1) << 32) + hm.sfSrc[q]);
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (size_t a = 0; a < cand.size();) {
        const int64_t sh = cand[a] >> 32;
        const int32_t first = int32_t(cand[a] & 0xffffffff);
        size_t b = a + 1;
        while (b < cand.size() && (cand[b] >> 32) == sh && int32_t(cand[b] & 0xffffffff) == first + int32_t(b - a)) ++b;
        hm.sfGrpRunFirst.push_back(first);
        hm.sfGrpRunEnd.push_back(first + int32_t(b - a));
        hm.sfGrpRunShift.push_back(uint8_t(sh));
        a = b;
      }
      hm.sfGrpRunOff[g + 1] = int32_t(hm.sfGrpRunFirst.size());
    }
    // inverse lists: source element -> (target, shift), targets ascending
    hm.sfSrcOff.assign(ne + 1, 0);
    for (int32_t sv : hm.sfSrc) hm.sfSrcOff[sv + 1]++;
    for (int e = 0; e < ne; ++e) hm.sfSrcOff[e + 1] += hm.sfSrcOff[e];
    hm.sfTgt.assign(hm.sfSrc.size(), 0);
    hm.sfTgtShift.assign(hm.sfSrc.size(), 0);
    std::vector<int32_t> fill(ne, 0);
    for (int t = 0; t < ne; ++t)
      for (int q = hm.sfOff[t]; q < hm.sfOff[t + 1]; ++q) {
        const int sv = hm.sfSrc[q], slot = hm.sfSrcOff[sv] + fill[sv]++;
        hm.sfTgt[slot] = t;
        hm.sfTgtShift[slot] = hm.sfShift[q];
      }
  }
  return "";
}

}  // namespace pic
