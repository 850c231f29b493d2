// RefMapping boundary tracking on the device: ParticleBCTracking
// (particles/tracking/particle_reftracking.f90:417-694) with the face
// intersections it calls (particles/tracking/particle_intersection.f90:
// ComputePlanarRectIntersection 515-684, ComputePlanarCurvedIntersection
// 687-852, ComputeCurvedIntersection 1243-1632, BezierClipRecursive 1635-1747,
// BezierNewton 1750-1988, InsideBoundingBox 1991-2038,
// BoundingBoxIntersection 2040-2109, FlatBoundingBoxIntersection 2112-2200,
// ComputeBezierIntersectionPoint 2276-2456, CheckXiClip / CheckEtaClip
// 2459-3242, calcLineNormVec3 3285-3363, CalcSminSmax2 3366-3451),
// GetBoundaryInteraction's RefMapping branch with PeriodicBoundary and the
// removal at open walls (particles/boundary/particle_boundary_condition.f90:35-327),
// and FallBackFaceIntersection (particle_reftracking.f90:986-1099).
//
// Used on meshes whose boundaries are not all periodic (CartesianPeriodic =
// F), for the few particles of boundary elements whose step reaches a
// boundary side; the side tables come from pic_mesh.cpp.  The reference's
// recursions (Bezier clipping after a split, ParticleBCTracking after a
// periodic crossing) run as loops here, the clipping with an explicit stack
// of the second halves of split patches, in the same depth-first order.
#pragma once

#include "stage_internal.h"

namespace pic {

// Curved and planar-curved boundary sides exist only on curved elements, and
// RefMapping runs on NGeo = 1 or 2: the Bezier routines below work on
// patches of degree 2, fixed at compile time so that the patch loops unroll
// and the patches stay in registers.
constexpr int kBzNG = 2;
constexpr int kBzMaxInter = 6;     // BezierClipMaxIntersec = 2 (NGeo + 1)
constexpr int kBzClipMax = 100;    // BezierClipMaxIter
constexpr int kBzArr = 24;         // clip levels kept per direction (beyond: error)
constexpr int kBzStack = 8;        // pending split halves (beyond: error)
constexpr int kBcHits = 16;        // sides hit in one element pass (beyond: error)
constexpr double kBzClipTol = 1e-8;                 // BezierClipTolerance
constexpr double kBzNewtonTol2 = 1e-4 * 1e-4;       // BezierNewtonTolerance^2
constexpr int kBzNewtonMaxIter = 100;
constexpr double kBzSplitLimit = 2.0 * 0.6;         // 2 BezierSplitLimit
constexpr double kBzClipHit = 1.0 + 100.0 * 1e-8;   // 1 + 100 BezierClipTolerance
constexpr double kBzEpsTol = 100.0 * kEpsMach;      // epsilontol
constexpr double kBzNewtonAngle = 1.570796326;
enum BcError { kBcOk = 0, kBcClipOverflow = 1, kBcClipMaxIter = 2, kBcTooManySides = 3, kBcHalo = 4,
               kBcNewtonFailed = 5, kBcUnsupported = 6 };

__device__ inline bool bz_almost_zero(double x) { return fabs(x) <= 2.22e-16; }
__device__ inline double bz_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
__device__ inline void bz_unit(double v[3]) {
  const double l = sqrt(bz_dot(v, v));
  v[0] = v[0] / l; v[1] = v[1] / l; v[2] = v[2] / l;
}
__device__ inline void bz_crossnorm(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  bz_unit(c);
}
__device__ inline double bz_choose(int n, int k) {
  double r = 1.0;
  for (int q = 1; q <= k; ++q) r = r * double(n - q + 1) / double(q);
  return r;
}
__device__ inline double bz_fac(int p, int l) { return (1.0 / double(1 << p)) * bz_choose(p, l); }   // FacNchooseK
__device__ inline double bz_ipow(double x, int k) {
  double r = 1.0;
  for (int q = 0; q < k; ++q) r = r * x;
  return r;
}
__device__ inline bool bc_elem(const DevMesh &M, int e) { return M.e2bOff && M.e2bOff[e + 1] > M.e2bOff[e]; }

// control point (p, q) of boundary side b
__device__ inline const double *bz_pt(const DevMesh &M, int b, int p, int q) {
  const int g1 = M.NGeo + 1;
  return M.bz + ((size_t(b) * g1 + p) * g1 + q) * 3;
}

struct BzPatch { double v[3][3][2]; };

struct BzWork {
  double locAlpha[kBzMaxInter], locXi[kBzMaxInter], locEta[kBzMaxInter];
  double arr[2][kBzArr + 1][2];   // XiArray, EtaArray
  double localTol;                // BezierClipLocalTol
  int err;
};

// EvaluateBezierPolynomialAndGradient without precomputed derivative
// (particle_surfaces.f90:474-616): point and/or gradient of a patch with
// iSize components; B indexed [(p*g1+q)*iSize+c]
__device__ inline void bz_eval(int ng, int iSize, const double *B, const double X[2], double *P, double G[2][3]) {
  double xiup[3], etaup[3], xidown[3], etadown[3], xit[3], etat[3];
  const int g1 = ng + 1, Mm = ng - 1;
  const double MinusXi = 1.0 - X[0], PlusXi = 1.0 + X[0], MinusEta = 1.0 - X[1], PlusEta = 1.0 + X[1];
  xiup[0] = 1.0; etaup[0] = 1.0; xidown[ng] = 1.0; etadown[ng] = 1.0;
  for (int p = 1; p <= ng; ++p) {
    xiup[p] = xiup[p - 1] * PlusXi;
    etaup[p] = etaup[p - 1] * PlusEta;
    xidown[ng - p] = xidown[ng - p + 1] * MinusXi;
    etadown[ng - p] = etadown[ng - p + 1] * MinusEta;
  }
  for (int p = 0; p <= ng; ++p) { xit[p] = xiup[p] * xidown[p]; etat[p] = etaup[p] * etadown[p]; }
  if (P) {
    for (int c = 0; c < iSize; ++c) P[c] = 0.;
    for (int q = 0; q <= ng; ++q)
      for (int p = 0; p <= ng; ++p)
        for (int c = 0; c < iSize; ++c)
          P[c] = P[c] + B[(p * g1 + q) * iSize + c] * bz_fac(ng, p) * xit[p] * bz_fac(ng, q) * etat[q];
  }
  if (G) {
    for (int d = 0; d < 2; ++d) for (int c = 0; c < iSize; ++c) G[d][c] = 0.;
    for (int q = 0; q <= ng; ++q)
      for (int p = 0; p <= Mm; ++p)
        for (int c = 0; c < iSize; ++c) {
          G[0][c] = G[0][c] + (B[((p + 1) * g1 + q) * iSize + c] - B[(p * g1 + q) * iSize + c]) * bz_fac(Mm, p) *
                                  xiup[p] * xidown[p + 1] * bz_fac(ng, q) * etat[q];
          G[1][c] = G[1][c] + (B[(q * g1 + p + 1) * iSize + c] - B[(q * g1 + p) * iSize + c]) * bz_fac(ng, q) *
                                  xit[q] * bz_fac(Mm, p) * etaup[p] * etadown[p + 1];
        }
    for (int d = 0; d < 2; ++d) for (int c = 0; c < iSize; ++c) G[d][c] = 0.5 * double(ng) * G[d][c];
  }
}

__device__ inline bool bz_inside_bbox(const DevMesh &M, int b, const double pos[3]) {
  const double *N = M.slabN + size_t(b) * 9, *I = M.slabI + size_t(b) * 6, *B0 = bz_pt(M, b, 0, 0);
  const double P[3] = {pos[0] - B0[0], pos[1] - B0[1], pos[2] - B0[2]};
  const double eps = 100. * kEpsMach;
  const double y = bz_dot(P, N + 3);
  if (y < I[2] - eps || y > I[3] + eps) return false;
  const double xx = bz_dot(P, N);
  if (xx < I[0] - eps || xx > I[1] + eps) return false;
  const double z = bz_dot(P, N + 6);
  if (z < I[4] - eps || z > I[5] + eps) return false;
  return true;
}

__device__ inline void bz_slab_alpha(const DevMesh &M, int b, int i, const double last[3], double dnk, double &amin,
                                     double &amax) {
  const double *N = M.slabN + size_t(b) * 9 + i * 3, *I = M.slabI + size_t(b) * 6, *B0 = bz_pt(M, b, 0, 0);
  const double d[3] = {B0[0] - last[0], B0[1] - last[1], B0[2] - last[2]};
  const double base = bz_dot(d, N);
  if (dnk < 0.) { amin = (base + I[2 * i + 1]) / dnk; amax = (base + I[2 * i]) / dnk; }
  else { amin = (base + I[2 * i]) / dnk; amax = (base + I[2 * i + 1]) / dnk; }
}

__device__ inline bool bz_bbox_intersection(const DevMesh &M, int b, const double last[3], const double traj[3],
                                            double len) {
  double lo[3], hi[3];
  for (int i = 0; i < 3; ++i) {
    double dnk = bz_dot(traj, M.slabN + size_t(b) * 9 + i * 3);
    if (fabs(dnk) < 100. * kEpsMach) dnk = 100. * kEpsMach;
    bz_slab_alpha(M, b, i, last, dnk, lo[i], hi[i]);
  }
  const double maxv = fmax(fmax(lo[0], lo[1]), lo[2]), minv = fmin(fmin(hi[0], hi[1]), hi[2]);
  if (maxv <= minv) return (maxv < len + 100 * kEpsMach) && (maxv + 100 * kEpsMach > 0.);
  return false;
}

__device__ inline bool bz_flat_bbox_intersection(const DevMesh &M, int b, const double last[3], const double traj[3],
                                                 double len) {
  double lo[3], hi[3];
  for (int i = 0; i < 3; i += 2) {
    const double dnk = bz_dot(traj, M.slabN + size_t(b) * 9 + i * 3);
    if (fabs(dnk) < 100. * kEpsMach) { lo[i] = -HUGE_VAL; hi[i] = HUGE_VAL; }
    else bz_slab_alpha(M, b, i, last, dnk, lo[i], hi[i]);
  }
  const double maxv = fmax(lo[0], lo[2]), minv = fmin(hi[0], hi[2]);
  if (maxv <= minv) {
    if (maxv < 0 && minv > 0) return true;
    return (maxv < len + 100. * kEpsMach) && (maxv + 100. * kEpsMach > 0.);
  }
  return false;
}

struct BzCtx {
  const DevMesh *M;
  int b;
  const double *last, *traj;
  double len;
  BzWork *W;
};

// ComputeBezierIntersectionPoint
__device__ inline void bz_intersection_point(const BzCtx &C, int nXi, int nEta, int &nInter) {
  const DevMesh &M = *C.M;
  BzWork &W = *C.W;
  constexpr int ng = kBzNG, g1 = ng + 1;
  double Xi, Eta;
  if (nXi == 0) Xi = 0.;
  else {
    Xi = 0.5 * (W.arr[0][nXi][0] + W.arr[0][nXi][1]);
    for (int k = nXi - 1; k >= 1; --k) Xi = W.arr[0][k][0] + 0.5 * (Xi + 1) * (W.arr[0][k][1] - W.arr[0][k][0]);
  }
  if (nEta == 0) Eta = 0.;
  else {
    Eta = 0.5 * (W.arr[1][nEta][0] + W.arr[1][nEta][1]);
    for (int k = nEta - 1; k >= 1; --k) Eta = W.arr[1][k][0] + 0.5 * (Eta + 1) * (W.arr[1][k][1] - W.arr[1][k][0]);
  }
  const double tXi = Xi, tEta = Eta;
  Xi = 0.5 * (Xi + 1);
  Eta = 0.5 * (Eta + 1);
  if (fabs(Eta) > kBzClipHit || fabs(Xi) > kBzClipHit) return;
  const double MXi = 1.0 - Xi, MEta = 1.0 - Eta;
  double R[3][3][3];
  for (int p = 0; p < g1; ++p) for (int q = 0; q < g1; ++q) for (int c = 0; c < 3; ++c) R[p][q][c] = bz_pt(M, C.b, p, q)[c];
  int l = ng - 1;
  for (int it = 1; it <= ng; ++it) {
    for (int q = 0; q <= l; ++q)
      for (int p = 0; p <= l; ++p)
        for (int c = 0; c < 3; ++c)
          R[p][q][c] = MXi * R[p][q][c] * MEta + MXi * R[p][q + 1][c] * Eta + Xi * R[p + 1][q][c] * MEta +
                       Xi * R[p + 1][q + 1][c] * Eta;
    l = l - 1;
  }
  const double iv[3] = {R[0][0][0] - C.last[0], R[0][0][1] - C.last[1], R[0][0][2] - C.last[2]};
  const double alpha = bz_dot(iv, C.traj);
  const double alphaNorm = alpha / C.len;
  if (alphaNorm <= 1.0 && alphaNorm > -kBzEpsTol) {
    if (nInter >= kBzMaxInter) { W.err = kBcClipOverflow; return; }   // the reference STOPs
    bool isNew = true;
    for (int k = 0; k < nInter; ++k) {
      if (fabs(W.locXi[k] - tXi) > 0.01) continue;
      if (fabs(W.locEta[k] - tEta) > 0.01) continue;
      isNew = false;
      break;
    }
    if (nInter == 0 || isNew) {
      W.locAlpha[nInter] = alpha;
      W.locXi[nInter] = tXi;
      W.locEta[nInter] = tEta;
      ++nInter;
    }
  }
}

// calcLineNormVec3
__device__ inline void bz_line_norm_vec(int ng, const BzPatch &B, double L[2][2], BzWork &W) {
  double LXi[2], LEta[2];
  for (int c = 0; c < 2; ++c) {
    LXi[c] = (B.v[0][ng][c] - B.v[0][0][c]) + (B.v[ng][ng][c] - B.v[ng][0][c]);
    LEta[c] = (B.v[ng][0][c] - B.v[0][0][c]) + (B.v[ng][ng][c] - B.v[0][ng][c]);
  }
  double l = sqrt(LXi[0] * LXi[0] + LXi[1] * LXi[1]);
  if (l == 0.) { W.err = kBcClipMaxIter; return; }
  LXi[0] = LXi[0] / l; LXi[1] = LXi[1] / l;
  l = sqrt(LEta[0] * LEta[0] + LEta[1] * LEta[1]);
  if (l == 0.) { W.err = kBcClipMaxIter; return; }
  LEta[0] = LEta[0] / l; LEta[1] = LEta[1] / l;
  L[0][0] = -LXi[1]; L[1][0] = LXi[0];
  L[0][1] = -LEta[1]; L[1][1] = LEta[0];
}

// CalcSminSmax2 (Xi_NGeo equidistant)
__device__ inline void bz_smin_smax(int ng, const double mm[2][3], double &Smin, double &Smax, int iter) {
  double X[3];
  for (int i = 0; i <= ng; ++i) X[i] = 2.0 * double(i) / double(ng) - 1.0;
  Smin = 1.5;
  Smax = -1.5;
  for (int p = 0; p <= ng; ++p)
    for (int q = 0; q <= ng; ++q) {
      if (p == q) continue;
      if (mm[1][p] * mm[1][q] <= 0.) {
        const double m = (mm[1][q] - mm[1][p]) / (X[q] - X[p]);
        if (m < 0.) Smax = fmax(Smax, X[q] - mm[1][q] / m);
        if (m > 0.) Smin = fmin(Smin, X[q] - mm[1][q] / m);
      }
      if (mm[0][p] * mm[0][q] <= 0.) {
        const double m = (mm[0][q] - mm[0][p]) / (X[q] - X[p]);
        if (m > 0.) Smax = fmax(Smax, X[q] - mm[0][q] / m);
        if (m < 0.) Smin = fmin(Smin, X[q] - mm[0][q] / m);
      }
    }
  if (mm[0][0] * mm[1][0] < 0.) Smin = -1.;
  if (mm[0][ng] * mm[1][ng] < 0.) Smax = 1.;
  if (Smax > -1.5) Smax = fmin(Smax + 20. * kBzClipTol, 1.0);
  if (Smin < 1.5) Smin = fmax(Smin - 20. * kBzClipTol, -1.0);
  if (iter == 0) {
    if (Smin == 1.5) Smin = -1.;
    if (Smax == -1.5) Smax = 1.;
  }
}

// sub-patches along dir (0: xi, first index; 1: eta, second): clip_left keeps
// [-1, hi]; clip_right keeps [lo, 1] of a patch spanning [-1, hi]
__device__ inline void bz_clip_left(int ng, int dir, const BzPatch &B, double hi, BzPatch &T) {
  double up[3], down[3], Buf[3][3];
  const double Plus = 1.0 + hi, Minus = 1.0 - hi;
  up[0] = 1.0; down[0] = 1.0;
  for (int l = 1; l <= ng; ++l) { up[l] = up[l - 1] * Plus; down[l] = down[l - 1] * Minus; }
  for (int p = 0; p <= ng; ++p) for (int l = 0; l <= p; ++l) Buf[p][l] = up[l] * down[p - l] * bz_fac(p, l);
  for (int a = 0; a <= ng; ++a) for (int b = 0; b <= ng; ++b) T.v[a][b][0] = T.v[a][b][1] = 0.;
  for (int q = 0; q <= ng; ++q)
    for (int p = 0; p <= ng; ++p)
      for (int l = 0; l <= p; ++l)
        for (int c = 0; c < 2; ++c) {
          if (dir == 0) T.v[p][q][c] = T.v[p][q][c] + B.v[l][q][c] * Buf[p][l];
          else T.v[q][p][c] = T.v[q][p][c] + B.v[q][l][c] * Buf[p][l];
        }
}
__device__ inline void bz_clip_right(int ng, int dir, const BzPatch &B, double lo, double hi, BzPatch &T) {
  double up[3], down[3], Buf[3][3];
  double Plus = (lo + 1.0) / (hi + 1.0);
  const double Minus = 2.0 * Plus;
  Plus = 2.0 - 2.0 * Plus;
  up[0] = 1.0; down[0] = 1.0;
  for (int l = 1; l <= ng; ++l) { up[l] = up[l - 1] * Plus; down[l] = down[l - 1] * Minus; }
  for (int p = 0; p <= ng; ++p) for (int l = 0; l <= p; ++l) Buf[p][l] = up[l] * down[p - l] * bz_fac(p, l);
  for (int a = 0; a <= ng; ++a) for (int b = 0; b <= ng; ++b) T.v[a][b][0] = T.v[a][b][1] = 0.;
  for (int q = 0; q <= ng; ++q)
    for (int p = 0; p <= ng; ++p)
      for (int l = 0; l <= p; ++l)
        for (int c = 0; c < 2; ++c) {
          if (dir == 0) T.v[ng - p][q][c] = T.v[ng - p][q][c] + B.v[ng - l][q][c] * Buf[p][l];
          else T.v[q][ng - p][c] = T.v[q][ng - p][c] + B.v[q][ng - l][c] * Buf[p][l];
        }
}

// state of one BezierClipRecursive activation
struct BzState {
  BzPatch B;
  double L[2][2];
  int clipMode, iClipIter, nXi, nEta;
};
// the second half of a split patch, run once the first half's subtree is done
struct BzPending {
  BzState st;
  int dir, level;
  double lo, hi;   // XiArray / EtaArray entry at level
};

// BezierClipRecursive + CheckXiClip / CheckEtaClip, depth first, with the
// second halves of split patches on a stack
__device__ __attribute__((noinline)) void bz_clip(const BzCtx &C, BzState st, int &nInter) {
  constexpr int ng = kBzNG;
  BzWork &W = *C.W;
  BzPending stack[kBzStack];
  int sp = 0;
  for (;;) {
    bool finished = false;   // this activation returned
    while (!finished) {
      if (W.err) return;
      if (st.iClipIter > kBzClipMax) { W.err = kBcClipMaxIter; return; }
      ++st.iClipIter;
      int dir;
      switch (st.clipMode) {
        case -1: finished = true; continue;
        case 1:
          if (st.iClipIter % 2 == 1) bz_line_norm_vec(ng, st.B, st.L, W);
          dir = 0;
          break;
        case 2:
          if (st.iClipIter % 2 == 1) bz_line_norm_vec(ng, st.B, st.L, W);
          dir = 1;
          break;
        case 3: dir = 0; break;
        case 4: dir = 1; break;
        case 5:
          bz_intersection_point(C, st.nXi, st.nEta, nInter);
          finished = true;
          continue;
        default: W.err = kBcClipMaxIter; return;
      }
      // CheckXiClip (dir 0) / CheckEtaClip (dir 1)
      double mm[2][3];
      for (int l = 0; l <= ng; ++l) {
        double mn = HUGE_VAL, mx = -HUGE_VAL;
        for (int k = 0; k <= ng; ++k) {
          const int p = dir == 0 ? l : k, q = dir == 0 ? k : l;
          const double v = st.B.v[p][q][0] * st.L[0][dir] + st.B.v[p][q][1] * st.L[1][dir];
          mn = fmin(mn, v);
          mx = fmax(mx, v);
        }
        mm[0][l] = mn;
        mm[1][l] = mx;
      }
      double dmin = HUGE_VAL, dmax = -HUGE_VAL;
      for (int l = 0; l <= ng; ++l) { dmin = fmin(dmin, mm[0][l]); dmax = fmax(dmax, mm[1][l]); }
      if (dmax * dmin > 0.) { st.clipMode = -1; continue; }
      if (fabs(dmax - dmin) < W.localTol) {
        if (dir == 0) st.clipMode = (st.clipMode == 3) ? 5 : 4;
        else st.clipMode = (st.clipMode == 4) ? 5 : 3;
        continue;
      }
      st.clipMode = dir == 0 ? 2 : 1;
      int &nClip = dir == 0 ? st.nXi : st.nEta;
      double Smin, Smax;
      bz_smin_smax(ng, mm, Smin, Smax, nClip);
      if (Smin == 1.5 || Smax == -1.5) { st.clipMode = -1; continue; }
      if (Smin > Smax) { const double t = Smax; Smax = Smin; Smin = t; }
      ++nClip;
      if (nClip > kBzArr) { W.err = kBcClipOverflow; return; }
      BzPatch T, T2;
      if (Smax - Smin > kBzSplitLimit) {
        const double Split = 0.5 * (Smax + Smin);
        W.arr[dir][nClip][0] = Split;
        W.arr[dir][nClip][1] = Smax;
        if (Smax != 1.0) bz_clip_left(ng, dir, st.B, Smax, T);
        else T = st.B;
        bz_clip_right(ng, dir, T, Split, Smax, T2);
        // second half: the state as it is now, patch [Smin, Split]
        if (sp >= kBzStack) { W.err = kBcClipOverflow; return; }
        BzPending &pd = stack[sp++];
        pd.st = st;
        bz_clip_left(ng, dir, st.B, Split, T);
        if (Smin != -1.0) bz_clip_right(ng, dir, T, Smin, Split, pd.st.B);
        else pd.st.B = T;
        pd.dir = dir;
        pd.level = nClip;
        pd.lo = Smin;
        pd.hi = Split;
        // first half: a child activation in the current state; after both
        // halves this activation ends (ClipMode = -1)
        st.B = T2;
      } else {
        W.arr[dir][nClip][0] = Smin;
        W.arr[dir][nClip][1] = Smax;
        if (Smax != 1.0) { bz_clip_left(ng, dir, st.B, Smax, T); st.B = T; }
        if (Smin != -1.0) { bz_clip_right(ng, dir, st.B, Smin, Smax, T); st.B = T; }
      }
    }
    if (sp == 0) return;
    BzPending &pd = stack[--sp];
    st = pd.st;
    W.arr[pd.dir][pd.level][0] = pd.lo;
    W.arr[pd.dir][pd.level][1] = pd.hi;
  }
}

// BezierNewton with BezierNewtonGuess = 1; the first gradient from the
// derivative patch D_Bezier (BuildBezierDMat, basis.f90:361-410) on the CL nodes
__device__ inline double bz_newton(const DevMesh &M, int b, double Xi[2], const BzPatch &B, const double last[3],
                                   const double traj[3], double len, bool &failed) {
  constexpr int ng = kBzNG, g1 = ng + 1;
  failed = false;
  double MinMax[2][2];
  bool hasInter = true;
  for (int l = 0; l < 2; ++l) {
    double mn = HUGE_VAL, mx = -HUGE_VAL;
    for (int p = 0; p <= ng; ++p) for (int q = 0; q <= ng; ++q) { mn = fmin(mn, B.v[p][q][l]); mx = fmax(mx, B.v[p][q][l]); }
    MinMax[0][l] = mn; MinMax[1][l] = mx;
    if (MinMax[0][l] * MinMax[1][l] > 0.) hasInter = false;
  }
  if (!hasInter) return -1.;
  for (int l = 0; l < 2; ++l) Xi[l] = -1.0 - 2 * MinMax[0][l] / (MinMax[1][l] - MinMax[0][l]);
  double D[3][3];
  for (int i = 0; i <= ng; ++i) for (int j = 0; j <= ng; ++j) D[i][j] = 0.;
  for (int i = 0; i <= ng; ++i) D[i][0] = -bz_fac(ng, 0) * double(ng) * bz_ipow(1 - M.xiCLg[i], ng - 1);
  for (int i = 0; i <= ng; ++i) D[i][ng] = bz_fac(ng, ng) * double(ng) * bz_ipow(M.xiCLg[i] + 1, ng - 1);
  for (int j = 1; j < ng; ++j)
    for (int i = 0; i <= ng; ++i) {
      const double Pp = M.xiCLg[i] + 1.0, Mn = 1.0 - M.xiCLg[i];
      D[i][j] = bz_fac(ng, j) * (double(j) * bz_ipow(Pp, j - 1) * bz_ipow(Mn, ng - j) -
                                 double(ng - j) * bz_ipow(Pp, j) * bz_ipow(Mn, ng - j - 1));
    }
  double P[2], G[2][3], Bf[18];
  for (int p = 0; p < g1; ++p) for (int q = 0; q < g1; ++q) { Bf[(p * g1 + q) * 2] = B.v[p][q][0]; Bf[(p * g1 + q) * 2 + 1] = B.v[p][q][1]; }
  {
    double xiup[3], etaup[3], xidown[3], etadown[3], xit[3], etat[3];
    xiup[0] = 1.0; etaup[0] = 1.0; xidown[ng] = 1.0; etadown[ng] = 1.0;
    for (int p = 1; p <= ng; ++p) {
      xiup[p] = xiup[p - 1] * (1.0 + Xi[0]);
      etaup[p] = etaup[p - 1] * (1.0 + Xi[1]);
      xidown[ng - p] = xidown[ng - p + 1] * (1.0 - Xi[0]);
      etadown[ng - p] = etadown[ng - p + 1] * (1.0 - Xi[1]);
    }
    for (int p = 0; p <= ng; ++p) { xit[p] = xiup[p] * xidown[p]; etat[p] = etaup[p] * etadown[p]; }
    P[0] = P[1] = 0.;
    for (int q = 0; q <= ng; ++q)
      for (int p = 0; p <= ng; ++p)
        for (int c = 0; c < 2; ++c) P[c] = P[c] + B.v[p][q][c] * bz_fac(ng, p) * xit[p] * bz_fac(ng, q) * etat[q];
    for (int d = 0; d < 2; ++d) for (int c = 0; c < 2; ++c) G[d][c] = 0.;
    for (int q = 0; q <= ng; ++q)
      for (int p = 0; p <= ng; ++p)
        for (int d = 0; d < 2; ++d)
          for (int c = 0; c < 2; ++c) {
            double dB = 0.;
            for (int l = 0; l <= ng; ++l) {
              const int ci = d == 0 ? l : p, cj = d == 0 ? q : l;
              dB = dB + D[d == 0 ? p : q][l] * B.v[ci][cj][c];
            }
            G[d][c] = G[d][c] + dB * bz_fac(ng, p) * xit[p] * bz_fac(ng, q) * etat[q];
          }
  }
  double NormP = P[0] * P[0] + P[1] * P[1];
  int nIter = 1;
  double dXi2 = 1.0, alpha = -1.0;
  while (dXi2 > kBzNewtonTol2 && nIter <= kBzNewtonMaxIter) {
    double sdet = G[0][0] * G[1][1] - G[0][1] * G[1][0];
    if (fabs(sdet) > kEpsMach) sdet = 1. / sdet;
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
      bz_eval(ng, 2, Bf, Xi, P, G);
      NormP = P[0] * P[0] + P[1] * P[1];
      lambda = 0.3 * lambda;
      ++iArmijo;
    }
    if (nIter >= 4 && (fabs(Xi[0]) > 1.5 || fabs(Xi[1]) > 1.5)) { Xi[0] = Xi[1] = 1.5; break; }
    ++nIter;
  }
  if (nIter > kBzNewtonMaxIter) { failed = true; return alpha; }
  if (fabs(Xi[0]) > 1.002 || fabs(Xi[1]) > 1.002) return alpha;
  double InterP[3];
  bz_eval(ng, 3, M.bz + size_t(b) * g1 * g1 * 3, Xi, InterP, nullptr);
  const double iv[3] = {InterP[0] - last[0], InterP[1] - last[1], InterP[2] - last[2]};
  alpha = bz_dot(iv, traj);
  const double alphaNorm = alpha / len;
  if (alphaNorm <= 1.0 && alphaNorm > -kBzEpsTol) return alpha;
  return -1.0;
}

__device__ inline void bz_project(const DevMesh &M, int b, const double last[3], const double na[3], const double nb[3],
                                  BzPatch &B) {
  constexpr int ng = kBzNG;
  for (int q = 0; q <= ng; ++q)
    for (int p = 0; p <= ng; ++p) {
      const double *bp = bz_pt(M, b, p, q);
      const double d[3] = {bp[0] - last[0], bp[1] - last[1], bp[2] - last[2]};
      B.v[p][q][0] = bz_dot(d, na);
      B.v[p][q][1] = bz_dot(d, nb);
    }
}

__device__ inline void bz_traj_basis(const double traj[3], double n1[3], double n2[3]) {
  n1[0] = traj[2]; n1[1] = traj[2]; n1[2] = -traj[0] - traj[1];   // ABS(traj(3)) < 0 never holds
  bz_unit(n1);
  bz_crossnorm(traj, n1, n2);
}

// ComputePlanarRectIntersection
__device__ inline int bc_planar_rect(const DevMesh &M, int b, int flip, const double last[3], const double traj[3],
                                     double len, double &alpha, double &xi, double &eta) {
  alpha = -1.0; xi = -2.; eta = -2.;
  const double *sn = M.sideNorm + size_t(b) * 3;
  double nv[3] = {sn[0], sn[1], sn[2]}, locDist = M.sideDist[b];
  if (flip != 0) { nv[0] = -nv[0]; nv[1] = -nv[1]; nv[2] = -nv[2]; locDist = -locDist; }
  const double coeffA = bz_dot(nv, traj);
  const double locSideDist = locDist - bz_dot(last, nv);
  if (bz_almost_zero(coeffA)) { alpha = -1.; return 0; }
  alpha = locSideDist / coeffA;
  if (locSideDist < -100 * kEpsMach) { alpha = -1.; return 0; }
  const double alphaNorm = alpha / len;
  if (alphaNorm > 1.0 || alphaNorm < -kBzEpsTol) { alpha = -1.; return 0; }
  const double *bv = M.baseV + size_t(b) * 9;
  double P0[3], P1[3], P2[3];
  for (int c = 0; c < 3; ++c) {
    const double inter = last[c] + alpha * traj[c];
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
  if (fabs(sdet) == 0) { alpha = -1.; return kBcUnsupported; }
  sdet = 1.0 / sdet;
  const double epsLoc = 1.0 + 100. * kEpsMach;
  xi = (B2 * C1 - B1 * C2) * sdet;
  if (fabs(xi) > epsLoc) { alpha = -1.0; return 0; }
  eta = (-A2 * C1 + A1 * C2) * sdet;
  if (fabs(eta) > epsLoc) { alpha = -1.0; return 0; }
  return 0;
}

__device__ inline double bz_patch_width(int ng, const BzPatch &B, int c) {
  double mx = -HUGE_VAL, mn = HUGE_VAL;
  for (int p = 0; p <= ng; ++p) for (int q = 0; q <= ng; ++q) { mx = fmax(mx, B.v[p][q][c]); mn = fmin(mn, B.v[p][q][c]); }
  return mx - mn;
}

// ComputePlanarCurvedIntersection (RefMapping)
__device__ inline int bc_planar_curved(const DevMesh &M, int b, int flip, const double last[3], const double traj[3],
                                       double len, double &alpha, double &xi, double &eta) {
  alpha = -1.0; xi = 2.0; eta = 2.0;
  const double *sn = M.sideNorm + size_t(b) * 3;
  double nv[3] = {sn[0], sn[1], sn[2]}, locDist = M.sideDist[b];
  if (flip != 0) { nv[0] = -nv[0]; nv[1] = -nv[1]; nv[2] = -nv[2]; locDist = -locDist; }
  const double coeffA = bz_dot(nv, traj);
  if (coeffA <= 0.) return 0;
  double locSideDist = locDist - bz_dot(last, nv);
  locSideDist = locSideDist / coeffA;
  if (locSideDist > len) return 0;
  if (!bz_flat_bbox_intersection(M, b, last, traj, len)) return 0;
  double n1[3], n2[3];
  bz_traj_basis(traj, n1, n2);
  BzPatch B;
  bz_project(M, b, last, n1, n2, B);
  BzWork W;
  W.err = 0;
  double Xn[2] = {0., 0.};
  bool failed = false;
  W.locAlpha[0] = bz_newton(M, b, Xn, B, last, traj, len, failed);
  W.locXi[0] = Xn[0];
  W.locEta[0] = Xn[1];
  if (failed) {
    // BezierClipLocalTol is a stale module value on this path in the
    // reference; the patch's own width is used
    W.localTol = fmin(bz_patch_width(kBzNG, B, 0), bz_patch_width(kBzNG, B, 1)) * kBzClipTol;
    W.locAlpha[0] = -1.0;
    BzState st;
    st.B = B;
    st.L[0][0] = st.L[0][1] = st.L[1][0] = st.L[1][1] = 0.;
    st.clipMode = 1; st.iClipIter = 0; st.nXi = 0; st.nEta = 0;
    int nInter = 0;
    BzCtx C = {&M, b, last, traj, len, &W};
    bz_clip(C, st, nInter);
    if (W.err) return W.err;
  }
  if (W.locAlpha[0] > -1) { alpha = W.locAlpha[0]; xi = W.locXi[0]; eta = W.locEta[0]; }
  return 0;
}

// ComputeCurvedIntersection (RefMapping)
__device__ inline int bc_curved(const DevMesh &M, int b, int flip, const double last[3], const double x[3],
                                const double traj[3], double len, double &alpha, double &xi, double &eta) {
  alpha = -1.0; xi = 2.0; eta = 2.0;
  const double *sn = M.sideNorm + size_t(b) * 3;
  const double nv[3] = {flip == 0 ? sn[0] : -sn[0], flip == 0 ? sn[1] : -sn[1], flip == 0 ? sn[2] : -sn[2]};
  if (M.bbEmpty[b]) {
    if (bz_dot(nv, traj) < 0.) return 0;
    if (!bz_flat_bbox_intersection(M, b, last, traj, len)) return 0;
  } else if (!bz_inside_bbox(M, b, last) && !bz_inside_bbox(M, b, x) && !bz_bbox_intersection(M, b, last, traj, len)) {
    return 0;
  }
  double n1[3], n2[3];
  bz_traj_basis(traj, n1, n2);
  const double PartFaceAngle = fabs(0.5 * M_PI - acos(bz_dot(traj, M.slabN + size_t(b) * 9 + 3)));
  BzPatch B;
  bz_project(M, b, last, n2, n1, B);
  BzWork W;
  W.err = 0;
  for (int k = 0; k < kBzMaxInter; ++k) W.locAlpha[k] = -1.0;
  const double dXi = bz_patch_width(kBzNG, B, 0), dEta = bz_patch_width(kBzNG, B, 1);
  W.localTol = fmin(dXi, dEta) * kBzClipTol;
  BzState st;
  st.B = B;
  st.L[0][0] = st.L[0][1] = st.L[1][0] = st.L[1][1] = 0.;
  st.iClipIter = 0; st.nXi = 0; st.nEta = 0;
  int nInter = 0;
  BzCtx C = {&M, b, last, traj, len, &W};
  if (PartFaceAngle < kBzNewtonAngle) {
    st.clipMode = dXi > dEta ? 1 : 2;
    bz_clip(C, st, nInter);
    if (W.err) return W.err;
  } else {
    // trajectory along the side normal: clip, then Newton from the clipped root
    st.clipMode = 1;
    bz_clip(C, st, nInter);
    if (W.err) return W.err;
    if (nInter > 1) return kBcNewtonFailed;
    double Xn[2] = {0., 0.};
    bool failed = false;
    W.locAlpha[0] = bz_newton(M, b, Xn, B, last, traj, len, failed);
    nInter = W.locAlpha[0] > -1 ? 1 : 0;
    if (failed) return kBcNewtonFailed;
    W.locXi[0] = Xn[0];
    W.locEta[0] = Xn[1];
  }
  if (nInter == 0) return 0;
  if (nInter == 1) { alpha = W.locAlpha[0]; xi = W.locXi[0]; eta = W.locEta[0]; return 0; }
  int locID[kBzMaxInter];
  for (int k = 0; k < nInter; ++k) locID[k] = k;
  for (int a = 1; a < nInter; ++a) {   // InsertionSort
    const double av = W.locAlpha[a];
    const int iv = locID[a];
    int j = a - 1;
    while (j >= 0 && W.locAlpha[j] > av) { W.locAlpha[j + 1] = W.locAlpha[j]; locID[j + 1] = locID[j]; --j; }
    W.locAlpha[j + 1] = av;
    locID[j + 1] = iv;
  }
  for (int k = 0; k < nInter; ++k)
    if (W.locAlpha[k] > -1.0) { alpha = W.locAlpha[k]; xi = W.locXi[locID[k]]; eta = W.locEta[locID[k]]; return 0; }
  return 0;
}

__device__ inline int bc_side_intersection(const DevMesh &M, int b, const double last[3], const double x[3],
                                           const double traj[3], double len, double &alpha, double &xi, double &eta) {
  const int flip = M.bcFlip[b];
  switch (M.sideType[b]) {
    case kSidePlanarRect: return bc_planar_rect(M, b, flip, last, traj, len, alpha, xi, eta);
    case kSidePlanarCurved:
      if (M.NGeo != kBzNG) { alpha = -1.0; return kBcUnsupported; }
      return bc_planar_curved(M, b, flip, last, traj, len, alpha, xi, eta);
    case kSideCurved:
      if (M.NGeo != kBzNG) { alpha = -1.0; return kBcUnsupported; }
      return bc_curved(M, b, flip, last, x, traj, len, alpha, xi, eta);
    default: alpha = -1.0; return kBcUnsupported;
  }
}

// Whether ParticleBCTracking's first pass can find an intersection for the
// step last -> x (len > 0) of a particle of boundary element e: the early
// returns of the intersection routines for every side in range, in the same
// arithmetic (planar rectangular sides: the whole test).  When it is false
// the pass finds none and the particle is mapped in e as any other.
__device__ inline bool bc_may_hit(const DevMesh &M, int e, const double x[3], const double last[3], double len) {
  double traj[3] = {x[0] - last[0], x[1] - last[1], x[2] - last[2]};
  traj[0] = traj[0] / len; traj[1] = traj[1] / len; traj[2] = traj[2] / len;
  const int first = M.e2bOff[e], nloc = M.e2bOff[e + 1] - first;
  for (int k = 0; k < nloc; ++k) {
    if (M.e2bDist[first + k] > len) break;
    const int b = M.e2bSide[first + k];
    const int flip = M.bcFlip[b];
    const double *sn = M.sideNorm + size_t(b) * 3;
    switch (M.sideType[b]) {
      case kSidePlanarRect: {
        double a, xi, eta;
        if (bc_planar_rect(M, b, flip, last, traj, len, a, xi, eta) != kBcOk || a > -1.0) return true;
        break;
      }
      case kSidePlanarCurved: {
        double nv[3] = {sn[0], sn[1], sn[2]}, locDist = M.sideDist[b];
        if (flip != 0) { nv[0] = -nv[0]; nv[1] = -nv[1]; nv[2] = -nv[2]; locDist = -locDist; }
        const double coeffA = bz_dot(nv, traj);
        if (coeffA <= 0.) break;
        if ((locDist - bz_dot(last, nv)) / coeffA > len) break;
        if (!bz_flat_bbox_intersection(M, b, last, traj, len)) break;
        return true;
      }
      case kSideCurved: {
        if (M.bbEmpty[b]) {
          const double nv[3] = {flip == 0 ? sn[0] : -sn[0], flip == 0 ? sn[1] : -sn[1], flip == 0 ? sn[2] : -sn[2]};
          if (bz_dot(nv, traj) < 0.) break;
          if (!bz_flat_bbox_intersection(M, b, last, traj, len)) break;
        } else if (!bz_inside_bbox(M, b, last) && !bz_inside_bbox(M, b, x) &&
                   !bz_bbox_intersection(M, b, last, traj, len)) {
          break;
        }
        return true;
      }
      default: return true;
    }
  }
  return false;
}

// GetBoundaryInteraction, RefMapping branch: returns crossedBC; open walls
// remove the particle, periodic sides move it (PeriodicBoundary)
__device__ inline bool bc_interaction(const DevMesh &M, int b, int &ElemID, double x[3], double last[3],
                                      const double traj[3], double &len, double alpha, double xi, double eta,
                                      bool &removed) {
  double n[3];
  if (M.sideType[b] == kSideCurved) {
    constexpr int g1 = kBzNG + 1;
    double G[2][3];
    const double X[2] = {xi, eta};
    bz_eval(kBzNG, 3, M.bz + size_t(b) * g1 * g1 * 3, X, nullptr, G);
    bz_crossnorm(G[0], G[1], n);
  } else {
    const double *sn = M.sideNorm + size_t(b) * 3;
    n[0] = sn[0]; n[1] = sn[1]; n[2] = sn[2];
  }
  if (M.bcFlip[b] != 0) { n[0] = -n[0]; n[1] = -n[1]; n[2] = -n[2]; }
  if (bz_dot(n, traj) <= 0.) return false;
  if (M.bcKind[b] == PIC_BC_OPEN) { removed = true; return true; }
  const int pvid = M.bcPV[b];
  const int v = abs(pvid) - 1;
  for (int c = 0; c < 3; ++c) last[c] = last[c] + traj[c] * alpha;
  for (int c = 0; c < 3; ++c) last[c] = last[c] + copysign(fabs(M.PV[v][c]), double(pvid));
  for (int c = 0; c < 3; ++c) x[c] = last[c] + (len - alpha) * traj[c];
  len = len - alpha;
  ElemID = M.bcElemNb[b];
  return true;
}

// ParticleBCTracking (CartesianPeriodic = F).  On return: done (particle did
// not move, or left through an open wall: removed), else ElemID is the
// element to map the particle in.  Returns kBcOk or an error where the
// reference aborts.
__device__ __attribute__((noinline)) int bc_tracking(const DevMesh &M, int &ElemID, double x[3], double last[3],
                                                     bool &done, bool &removed) {
  done = false;
  removed = false;
  double len0 = 0.0;
  for (int pass = 0; pass < 1000; ++pass) {   // one pass per element; the reference recurses
    double traj[3] = {x[0] - last[0], x[1] - last[1], x[2] - last[2]};
    double len = sqrt(bz_dot(traj, traj));
    if (bz_almost_zero(len / M.radius[ElemID])) { done = true; return kBcOk; }   // PARTHASMOVED
    if (len > M.bcHalo) return kBcHalo;
    traj[0] = traj[0] / len; traj[1] = traj[1] / len; traj[2] = traj[2] / len;
    len0 = fmax(len0, len);
    const int first = M.e2bOff[ElemID], nloc = M.e2bOff[ElemID + 1] - first;
    bool doTracing = true, doubleCheck = false, crossed = false, newElem = false;
    while (doTracing) {
      // the sides hit, in list order (the others keep alpha = -1 and are
      // passed over after the sort)
      double locAlpha[kBcHits], xs[kBcHits], es[kBcHits];
      int list[kBcHits];
      int nv = 0;
      for (int k = 0; k < nloc; ++k) {
        if (M.e2bDist[first + k] > len0) break;
        double a, xk, ek;
        const int rc = bc_side_intersection(M, M.e2bSide[first + k], last, x, traj, len, a, xk, ek);
        if (rc != kBcOk) return rc;
        if (a > -1.0) {
          if (nv >= kBcHits) return kBcTooManySides;
          locAlpha[nv] = a; xs[nv] = xk; es[nv] = ek; list[nv] = k;
          ++nv;
        }
      }
      if (nv == 0) { doTracing = false; continue; }
      for (int a = 1; a < nv; ++a) {   // InsertionSort(locAlpha, locSideList): stable
        const double av = locAlpha[a], xa = xs[a], ea = es[a];
        const int la = list[a];
        int j = a - 1;
        while (j >= 0 && locAlpha[j] > av) {
          locAlpha[j + 1] = locAlpha[j]; xs[j + 1] = xs[j]; es[j + 1] = es[j]; list[j + 1] = list[j];
          --j;
        }
        locAlpha[j + 1] = av; xs[j + 1] = xa; es[j + 1] = ea; list[j + 1] = la;
      }
      for (int k = 0; k < nv; ++k) {
        if (!(locAlpha[k] > -1.0)) continue;
        const int b = M.e2bSide[first + list[k]];
        const int OldElemID = ElemID;
        crossed = bc_interaction(M, b, ElemID, x, last, traj, len, locAlpha[k], xs[k], es[k], removed);
        if (ElemID != OldElemID) {
          if (M.nPV > 0) {   // the largest distance of the old element's list
            double mx = -HUGE_VAL;
            for (int r = first; r < first + nloc; ++r) mx = fmax(mx, M.e2bDist[r]);
            len0 = mx;
          }
          newElem = true;
          break;
        }
        if (crossed) break;
      }
      if (newElem) break;
      if (removed) { done = true; return kBcOk; }
      if (!crossed) {
        if (!doubleCheck) doubleCheck = true;
        else doTracing = false;
      }
    }
    if (!newElem) return kBcOk;
  }
  return kBcHalo;
}

// FallBackFaceIntersection: returns 0 (face hit: the particle moved to 97 %
// of the way from the element origin), 1 (no hit: positions restored) or an
// error code
__device__ __attribute__((noinline)) int bc_fallback(const DevMesh &M, int ElemID, double x[3], double last[3]) {
  const double tmpPos[3] = {x[0], x[1], x[2]}, tmpLast[3] = {last[0], last[1], last[2]};
  for (int c = 0; c < 3; ++c) last[c] = M.bary[ElemID * 3 + c];
  double traj[3] = {x[0] - last[0], x[1] - last[1], x[2] - last[2]};
  const double len = sqrt(bz_dot(traj, traj));
  if (len > 0) { traj[0] = traj[0] / len; traj[1] = traj[1] / len; traj[2] = traj[2] / len; }
  const int first = M.e2bOff[ElemID], nloc = M.e2bOff[ElemID + 1] - first;
  double locAlpha[kBcHits];
  int nInter = 0;
  for (int k = 0; k < nloc; ++k) {
    double a, xi, eta;
    const int rc = bc_side_intersection(M, M.e2bSide[first + k], last, x, traj, len, a, xi, eta);
    if (rc != kBcOk) return rc;
    if (a > -1.0) {
      if (nInter >= kBcHits) return kBcTooManySides;
      locAlpha[nInter++] = a;
    }
  }
  if (nInter == 0) {
    for (int c = 0; c < 3; ++c) { x[c] = tmpPos[c]; last[c] = tmpLast[c]; }
    return 1;
  }
  for (int a = 1; a < nInter; ++a) {   // InsertionSort
    const double av = locAlpha[a];
    int j = a - 1;
    while (j >= 0 && locAlpha[j] > av) { locAlpha[j + 1] = locAlpha[j]; --j; }
    locAlpha[j + 1] = av;
  }
  for (int k = 0; k < nInter; ++k) {   // every hit, as the reference's loop
    for (int c = 0; c < 3; ++c) last[c] = last[c] + 0.97 * locAlpha[k] * traj[c];
    for (int c = 0; c < 3; ++c) x[c] = last[c];
  }
  return 0;
}

}  // namespace pic
