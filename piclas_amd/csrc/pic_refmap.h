// Device routines of reference-mapping tracking (RefMapping) on linear and
// curved (NGeo = 2) hexahedra.  Every routine restates one routine of the
// reference (paths relative to its src/) with the same operation order, but
// for the Lagrange basis of the map, formed in product form (lagrange_prod:
// the same polynomials, no division per evaluation; round-off apart).
#pragma once

#include "pic_device.h"
#include "pic_bctrack.h"

namespace pic {

// RefElemNewton on a map of G points per direction, mode 2 without
// isSuccessful (interpolation/eval_xyz.f90:298-445): Newton with Armijo
// back-tracking until |dxi|^2 <= RefMappingEps; leaves the loop when a
// component of xi exceeds 1.5.
template <int G>
__device__ inline void newton_map(const double *X, const double *dX, const double *xiCL, const double *cf,
                                  double refEps, const double xin[3], double xi[3]) {
  double L[3][G], F[3], xo[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) lagrange_prod<G>(xi[d], xiCL, cf, L[d]);
#pragma unroll
  for (int c = 0; c < 3; ++c) F[c] = -xin[c];
#pragma unroll
  for (int k = 0; k < G; ++k)
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const double buff = L[1][j] * L[2][k];
#pragma unroll
      for (int i = 0; i < G; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) F[c] = F[c] + X[((k * G + j) * G + i) * 3 + c] * L[0][i] * buff;
    }
  double deltaXi2 = (fabs(F[0]) < kEpsMach && fabs(F[1]) < kEpsMach && fabs(F[2]) < kEpsMach) ? 0.0 : 1.0;
  double NormF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
  double NormF_old;
  int it = 0;
  while (deltaXi2 > refEps && it < 100) {
    ++it;
    asm volatile("" ::: "memory");
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll 1
    for (int k = 0; k < G; ++k)
#pragma unroll
      for (int j = 0; j < G; ++j) {
        const double buff = L[1][j] * L[2][k];
#pragma unroll
        for (int i = 0; i < G; ++i) {
          const double b2 = L[0][i] * buff;
          const double *dp = dX + ((k * G + j) * G + i) * 9;
#pragma unroll
          for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) J[a][b] = J[a][b] + dp[a * 3 + b] * b2;
        }
      }
    double sdet = get_det(J);
    if (sdet > 0.) sdet = 1. / sdet;
    double S[3][3];
    S[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * sdet;
    S[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * sdet;
    S[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * sdet;
    S[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * sdet;
    S[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * sdet;
    S[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * sdet;
    S[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * sdet;
    S[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * sdet;
    S[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * sdet;
    double dxi[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) dxi[r] = S[r][0] * F[0] + S[r][1] * F[1] + S[r][2] * F[2];
    deltaXi2 = dxi[0] * dxi[0] + dxi[1] * dxi[1] + dxi[2] * dxi[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) xo[d] = xi[d];
    NormF_old = NormF;
    NormF = NormF * 2.;
    double lambda = 1.0;
    int iArmijo = 1;
    while (NormF > NormF_old * (1. - 0.0001 * lambda) && iArmijo <= 8) {
      asm volatile("" ::: "memory");
#pragma unroll
      for (int d = 0; d < 3; ++d) xi[d] = xo[d] - lambda * dxi[d];
#pragma unroll
      for (int d = 0; d < 3; ++d) lagrange_prod<G>(xi[d], xiCL, cf, L[d]);
#pragma unroll
      for (int c = 0; c < 3; ++c) F[c] = -xin[c];
#pragma unroll 1
      for (int k = 0; k < G; ++k)
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const double buff = L[1][j] * L[2][k];
#pragma unroll
          for (int i = 0; i < G; ++i) {
            const double b2 = L[0][i] * buff;
#pragma unroll
            for (int c = 0; c < 3; ++c) F[c] = F[c] + X[((k * G + j) * G + i) * 3 + c] * b2;
          }
        }
      lambda = 0.2 * lambda;
      iArmijo = iArmijo + 1;
      NormF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
    }
    if (fabs(xi[0]) > 1.5 || fabs(xi[1]) > 1.5 || fabs(xi[2]) > 1.5) break;
  }
}

// ---- wave-uniform element: the map read by scalar loads (sload_chunk) ----
// F = -xin + sum_r X_r (L_i L_j L_k), the association of the first evaluation
// in newton_map ((X L_i) buff) or of the Armijo loop (X (L_i buff))
template <int G, bool FIRST>
__device__ inline void map_s(const double *X, const double L[3][G], const double xin[3], double F[3]) {
  constexpr int NV = G * G * G * 3;
#pragma unroll
  for (int c = 0; c < 3; ++c) F[c] = -xin[c];
  double buff = 0.0, li = 0.0;
#pragma unroll
  for (int ch = 0; ch < (NV + 15) / 16; ++ch) {
    double buf[16];
    sload_chunk<NV>(X, ch, buf);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int v = 16 * ch + q;
      if (v < NV) {
        const int r = v / 3, cc = v % 3, i = r % G, j = (r / G) % G, k = r / (G * G);
        if (cc == 0) {
          buff = L[1][j] * L[2][k];
          li = FIRST ? L[0][i] : L[0][i] * buff;
        }
        if (FIRST) F[cc] = F[cc] + buf[q] * li * buff;
        else F[cc] = F[cc] + buf[q] * li;
      }
    }
  }
}

template <int G>
__device__ inline void newton_map_s(const double *X, const double *dX, const double *xiCL, const double *cf,
                                    double refEps, const double xin[3], double xi[3]) {
  constexpr int NJ = G * G * G * 9;
  double L[3][G], F[3], xo[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) lagrange_prod<G>(xi[d], xiCL, cf, L[d]);
  map_s<G, true>(X, L, xin, F);
  double deltaXi2 = (fabs(F[0]) < kEpsMach && fabs(F[1]) < kEpsMach && fabs(F[2]) < kEpsMach) ? 0.0 : 1.0;
  double NormF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
  double NormF_old;
  int it = 0;
  while (deltaXi2 > refEps && it < 100) {
    ++it;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double b2 = 0.0;
#pragma unroll
    for (int ch = 0; ch < (NJ + 15) / 16; ++ch) {
      double buf[16];
      sload_chunk<NJ>(dX, ch, buf);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int v = 16 * ch + q;
        if (v < NJ) {
          const int r = v / 9, ab = v % 9, i = r % G, j = (r / G) % G, k = r / (G * G);
          if (ab == 0) {
            const double buff = L[1][j] * L[2][k];
            b2 = L[0][i] * buff;
          }
          J[ab / 3][ab % 3] = J[ab / 3][ab % 3] + buf[q] * b2;
        }
      }
    }
    double sdet = get_det(J);
    if (sdet > 0.) sdet = 1. / sdet;
    double S[3][3];
    S[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * sdet;
    S[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * sdet;
    S[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * sdet;
    S[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * sdet;
    S[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * sdet;
    S[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * sdet;
    S[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * sdet;
    S[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * sdet;
    S[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * sdet;
    double dxi[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) dxi[r] = S[r][0] * F[0] + S[r][1] * F[1] + S[r][2] * F[2];
    deltaXi2 = dxi[0] * dxi[0] + dxi[1] * dxi[1] + dxi[2] * dxi[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) xo[d] = xi[d];
    NormF_old = NormF;
    NormF = NormF * 2.;
    double lambda = 1.0;
    int iArmijo = 1;
    while (NormF > NormF_old * (1. - 0.0001 * lambda) && iArmijo <= 8) {
#pragma unroll
      for (int d = 0; d < 3; ++d) xi[d] = xo[d] - lambda * dxi[d];
#pragma unroll
      for (int d = 0; d < 3; ++d) lagrange_prod<G>(xi[d], xiCL, cf, L[d]);
      map_s<G, false>(X, L, xin, F);
      lambda = 0.2 * lambda;
      iArmijo = iArmijo + 1;
      NormF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
    }
    if (fabs(xi[0]) > 1.5 || fabs(xi[1]) > 1.5 || fabs(xi[2]) > 1.5) break;
  }
}

// GetRefNewtonStartValue guess 3: the closest CL point of XCL_NGeo
// (interpolation/eval_xyz.f90:580-598); guesses 1, 2 and 4 as ref_newton_start
template <int N, int G, bool UNI = false>
__device__ inline void rm_start(const DevMesh &M, int e, const double x[3], double xi[3]) {
  if (M.refGuess != 3) {
    ref_newton_start<N, UNI>(M, e, x, xi);
    return;
  }
  const double *X = M.XCL + (size_t)e * G * G * G * 3;
  const double d0[3] = {x[0] - X[0], x[1] - X[1], x[2] - X[2]};
  double win = sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
  xi[0] = xi[1] = xi[2] = M.xiCLg[0];
#pragma unroll 1
  for (int i = 0; i < G; ++i)
#pragma unroll 1
    for (int j = 0; j < G; ++j)
#pragma unroll 1
      for (int k = 0; k < G; ++k) {
        const int r = (k * G + j) * G + i;
        const double dX = fabs(x[0] - X[r * 3]);
        if (dX > win) continue;
        const double dY = fabs(x[1] - X[r * 3 + 1]);
        if (dY > win) continue;
        const double dZ = fabs(x[2] - X[r * 3 + 2]);
        if (dZ > win) continue;
        const double dist = sqrt(dX * dX + dY * dY + dZ * dZ);
        if (dist < win) { win = dist; xi[0] = M.xiCLg[i]; xi[1] = M.xiCLg[j]; xi[2] = M.xiCLg[k]; }
      }
}

// GetPositionInRefElem (interpolation/eval_xyz.f90:35-123): start value, then
// Newton on the full NGeo map for curved elements and NGeo = 1 meshes, else on
// the trilinear map of the 8 corner CL points.
template <int N, int G>
__device__ inline void rm_position(const DevMesh &M, int e, const double x[3], double xi[3]) {
  rm_start<N, G>(M, e, x, xi);
  if (G == 2 || M.curved[e])
    newton_map<G>(M.XCL + (size_t)e * G * G * G * 3, M.dXCLg + (size_t)e * G * G * G * 9, M.xiCLg, M.cCLg,
                  M.refEps, x, xi);
  else
    newton_map<2>(M.corner + (size_t)e * 24, M.dXCL + (size_t)e * 72, M.xiCL, M.cCL, M.refEps, x, xi);
}

__device__ inline double max_abs3(const double v[3]) { return fmax(fabs(v[0]), fmax(fabs(v[1]), fabs(v[2]))); }

// PeriodicMovement with FastPeriodic = F (particles/tracking/particle_reftracking.f90:697-983)
__device__ inline bool periodic_movement(const DevMesh &M, double x[3], double last[3]) {
  bool moved = false;
#pragma unroll 1
  for (int dir = 0; dir < 3; ++dir) {
    int v = -1;
    for (int q = 0; q < M.nPV; ++q)
      if (M.dirPV[q] == dir) { v = q; break; }
    if (v < 0) continue;
    if (x[dir] > M.gmax[dir]) {
      const double sg = M.PV[v][dir] > 0 ? -1.0 : 1.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) { x[c] = x[c] + sg * M.PV[v][c]; last[c] = last[c] + sg * M.PV[v][c]; }
      moved = true;
    }
    if (x[dir] < M.gmin[dir]) {
      const double sg = M.PV[v][dir] > 0 ? 1.0 : -1.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) { x[c] = x[c] + sg * M.PV[v][c]; last[c] = last[c] + sg * M.PV[v][c]; }
      moved = true;
    }
  }
  return moved;
}

// Candidate order of ParticleRefTracking / SinglePointToElement: a stable
// ascending sort of the background cell's elements by |x - bary|^2, without
// the excluded ones.  Instead of sorting, each call returns the next entry
// after (dPrev, qPrev) in that order (or -1).
__device__ inline int next_candidate(const DevMesh &M, int b0, int n, const double x[3], int exclude, bool single,
                                     double &dPrev, int &qPrev) {
  double best = 0.0;
  int bq = -1;
  for (int q = 0; q < n; ++q) {
    const int e = M.binElems[b0 + q];
    double d2 = 0.0;
    if (!single) {
      if (e == exclude) continue;
      const double *b = M.bary + (size_t)e * 3;
      const double v0 = x[0] - b[0], v1 = x[1] - b[1], v2 = x[2] - b[2];
      d2 = v0 * v0 + v1 * v1 + v2 * v2;
      if (d2 > M.radius[e] * M.radius[e]) continue;
    }
    if (d2 < dPrev || (d2 == dPrev && q <= qPrev)) continue;
    if (bq < 0 || d2 < best) { best = d2; bq = q; }
  }
  if (bq >= 0) { dPrev = best; qPrev = bq; }
  return bq;
}

enum RmStatus { kRmInside = 0, kRmRemoved = 1, kRmNotFound = 2, kRmDeferred = 3 };
// first part, in the last element e0 (ParticleRefTracking :87-161): whether
// the particle is still inside e0 (xi: PartPosRef there, else the mapping in
// e0 for the candidate search), and whether it is a boundary candidate: a
// particle of a boundary element whose step reaches the nearest boundary
// side in range, for which ParticleBCTracking can find an intersection
// (decided by bc_may_hit before rm_bc_track runs)
struct RmHead {
  bool inside, cand;
};

// rm_position for a wave whose lanes all map in element e (e wave-uniform):
// the Newton reads the map by scalar loads
template <int N, int G>
__device__ inline void rm_position_s(const DevMesh &M, int e, const double x[3], double xi[3]) {
  rm_start<N, G, true>(M, e, x, xi);
  if (G == 2 || M.curved[e])
    newton_map_s<G>(M.XCL + (size_t)e * G * G * G * 3, M.dXCLg + (size_t)e * G * G * G * 9, M.xiCLg, M.cCLg,
                    M.refEps, x, xi);
  else
    newton_map_s<2>(M.corner + (size_t)e * 24, M.dXCL + (size_t)e * 72, M.xiCL, M.cCL, M.refEps, x, xi);
}

template <int N, int G>
__device__ inline int rm_locate(const DevMesh &M, const double x[3], double xi[3]);
// PeriodicMovement (CartesianPeriodic meshes, which have no boundary
// tables), or for a particle of a boundary element the reference's early
// return when it did not move (PartPosRef kept, done) and the candidate
// test; then the mapping in e0.  Nothing of the step but x and xi is live
// across the Newton iteration.
template <int N, int G, bool UNI = false>
__device__ inline RmHead rm_track_head(const DevMesh &M, int e0, double x[3], double last[3], double xi[3]) {
  RmHead h = {false, false};
  if (M.cartPeriodic && M.nPV > 0) {
    for (int g = 0; g < 4; ++g)
      if (!periodic_movement(M, x, last)) break;
  } else if (bc_elem(M, e0)) {
    const double d0 = x[0] - last[0], d1 = x[1] - last[1], d2 = x[2] - last[2];
    const double len = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    if (bz_almost_zero(len / M.radius[e0])) {   // PARTHASMOVED: PartPosRef kept
      h.inside = true;
      return h;
    }
    // Of the sides in range (SideBCMetrics distance <= the step), the
    // intersection routines can report a hit only where the step reaches
    // the side's normal slab: the patch lies in the hull of its control
    // points, and the clip and Newton tolerances accept points at most
    // ~1e-6 of the side's extent outside it, far inside the slab's 1e-3
    // margin.  A step that reaches no slab hits nothing, so the boundary
    // tracking would leave it as it is; only the others are candidates.
    const int first = M.e2bOff[e0], nloc = M.e2bOff[e0 + 1] - first;
    for (int k = 0; k < nloc; ++k) {
      if (M.e2bDist[first + k] > len) break;
      const double *sl = M.e2bSlab + size_t(first + k) * 6;
      const double a = sl[0] * last[0] + sl[1] * last[1] + sl[2] * last[2];
      const double c = sl[0] * x[0] + sl[1] * x[1] + sl[2] * x[2];
      if (fmax(a, c) >= sl[3] && fmin(a, c) <= sl[4]) {
        h.cand = true;
        break;
      }
    }
  }
  // (e0 may sit in a vector register; the scalar loads need it wave-uniform)
  if (UNI) rm_position_s<N, G>(M, __builtin_amdgcn_readfirstlane(e0), x, xi);
  else rm_position<N, G>(M, e0, x, xi);
  h.inside = max_abs3(xi) < 1.0;
  return h;
}

// second part, for particles that left: the candidate search (:167-324)
// around FIBGM cell of x, excluding LastElemID e0; xi holds oldXi on entry.
// Not found within ElemEpsOneCell in a boundary element: the fallback
// (:325-405) with FallBackFaceIntersection, ParticleBCTracking again and
// LocateParticleInElement.  NOBC: a particle that needs that fallback is
// left as it was (status kRmDeferred, x and last untouched) for a kernel
// that holds the boundary code.
template <int N, int G, bool NOBC = false>
__device__ inline int rm_search(const DevMesh &M, int e0, double x[3], double last[3], double xi[3], int &status) {
  status = kRmInside;
  // FIBGM cell (:180-186)
  int b3[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    int v = int(floor((x[c] - M.bmin[c]) / M.bdelta[c]));
    v = v < 0 ? 0 : v;
    v = v + 1;
    v = v > M.nb[c] ? M.nb[c] : v;
    b3[c] = v - 1;
  }
  const int b = (b3[2] * M.nb[1] + b3[1]) * M.nb[0] + b3[0];
  const int b0 = M.binOff[b], n = M.binOff[b + 1] - b0;
  const double oldXi[3] = {xi[0], xi[1], xi[2]};
  double newXi[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL};
  int newElem = -1;
  double dPrev = -HUGE_VAL;
  int qPrev = -1;
  int q;
  while ((q = next_candidate(M, b0, n, x, e0, n == 1, dPrev, qPrev)) >= 0) {
    const int e = M.binElems[b0 + q];
    double t[3];
    rm_position<N, G>(M, e, x, t);
    if (max_abs3(t) < 1.0) {
      xi[0] = t[0]; xi[1] = t[1]; xi[2] = t[2];
      return e;
    }
    if (max_abs3(t) < max_abs3(newXi)) {
      newXi[0] = t[0]; newXi[1] = t[1]; newXi[2] = t[2];
      newElem = e;
    }
  }
  int test;
  if (max_abs3(oldXi) < max_abs3(newXi)) {
    xi[0] = oldXi[0]; xi[1] = oldXi[1]; xi[2] = oldXi[2];
    test = e0;
  } else {
    xi[0] = newXi[0]; xi[1] = newXi[1]; xi[2] = newXi[2];
    test = newElem;
  }
  if (test >= 0 && max_abs3(xi) <= M.epsOneCell[test]) return test;
  if (test < 0 || !bc_elem(M, test)) {   // Tolerance Issue with internal element: the reference aborts
    status = kRmNotFound;
    return -1;
  }
  if (NOBC) {
    status = kRmDeferred;
    return -1;
  }
  if (M.NGeo > 1) {
    const int rc = bc_fallback(M, test, x, last);
    if (rc == 1) {   // no face hit: LocateParticleInElement; not found aborts
      const int e = rm_locate<N, G>(M, x, xi);
      if (e < 0) status = kRmNotFound;
      return e;
    }
    if (rc != kBcOk) { status = kRmNotFound; return -1; }
  }
  bool done, removed;
  if (bc_tracking(M, test, x, last, done, removed) != kBcOk) { status = kRmNotFound; return -1; }
  if (removed) { status = kRmRemoved; return -1; }
  if (done) return test;
  rm_position<N, G>(M, test, x, xi);
  if (max_abs3(xi) > M.epsOneCell[test]) {   // Tolerance Issue with BC element, relocating
    const int e = rm_locate<N, G>(M, x, xi);
    if (e < 0) status = kRmNotFound;
    return e;
  }
  return test;
}

// ParticleBCTracking, the mapping in the element it ends in and, if that
// fails, the candidate search, for a particle of boundary element e0
template <int N, int G>
__device__ inline int rm_bc_track(const DevMesh &M, int e0, double x[3], double last[3], double xi[3], int &status) {
  status = kRmInside;
  int e = e0;
  bool done, removed;
  if (bc_tracking(M, e, x, last, done, removed) != kBcOk) { status = kRmNotFound; return -1; }
  if (removed) { status = kRmRemoved; return -1; }
  if (done) return e;
  rm_position<N, G>(M, e, x, xi);
  if (max_abs3(xi) < 1.0) return e;
  return rm_search<N, G>(M, e0, x, last, xi, status);
}

// the rest of ParticleRefTracking after the head (xi: the head's mapping in
// e0); FILTERED: the candidate bit already holds the outcome of bc_may_hit
template <int N, int G, bool FILTERED = false>
__device__ inline int rm_track_tail(const DevMesh &M, int e0, RmHead h, double x[3], double last[3], double xi[3],
                                    int &status) {
  status = kRmInside;
  if (h.cand) {
    if (FILTERED) return rm_bc_track<N, G>(M, e0, x, last, xi, status);
    const double d0 = x[0] - last[0], d1 = x[1] - last[1], d2 = x[2] - last[2];
    if (bc_may_hit(M, e0, x, last, sqrt(d0 * d0 + d1 * d1 + d2 * d2))) return rm_bc_track<N, G>(M, e0, x, last, xi, status);
  }
  if (h.inside) return e0;
  return rm_search<N, G>(M, e0, x, last, xi, status);
}

template <int N, int G>
__device__ inline int rm_track(const DevMesh &M, int e0, double x[3], double last[3], double xi[3], int &status) {
  const RmHead h = rm_track_head<N, G>(M, e0, x, last, xi);
  return rm_track_tail<N, G>(M, e0, h, x, last, xi, status);
}

// SinglePointToElement for RefMapping (particles/tracking/particle_localization.f90:81-203)
template <int N, int G>
__device__ inline int rm_locate(const DevMesh &M, const double x[3], double xi[3]) {
  int b3[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    int v = int(ceil((x[c] - M.bmin[c]) / M.bdelta[c]));
    v = v > M.nb[c] ? M.nb[c] : v;
    v = v < 1 ? 1 : v;
    b3[c] = v - 1;
  }
  const int b = (b3[2] * M.nb[1] + b3[1]) * M.nb[0] + b3[0];
  const int b0 = M.binOff[b], n = M.binOff[b + 1] - b0;
  double dPrev = -HUGE_VAL;
  int qPrev = -1, q;
  while ((q = next_candidate(M, b0, n, x, -1, false, dPrev, qPrev)) >= 0) {
    const int e = M.binElems[b0 + q];
    rm_position<N, G>(M, e, x, xi);
    if (max_abs3(xi) <= M.epsOneCell[e]) return e;
  }
  return -1;
}

// SinglePointToElement with TriaTracking (particles/tracking/particle_localization.f90:81-203):
// candidates within ElemRadius, nearest first, ParticleInsideQuad3D
__device__ inline int tria_locate(const DevMesh &M, const double x[3]) {
  int b3[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    int v = int(ceil((x[c] - M.bmin[c]) / M.bdelta[c]));
    v = v > M.nb[c] ? M.nb[c] : v;
    v = v < 1 ? 1 : v;
    b3[c] = v - 1;
  }
  const int b = (b3[2] * M.nb[1] + b3[1]) * M.nb[0] + b3[0];
  const int b0 = M.binOff[b], n = M.binOff[b + 1] - b0;
  double dPrev = -HUGE_VAL;
  int qPrev = -1, q;
  while ((q = next_candidate(M, b0, n, x, -1, false, dPrev, qPrev)) >= 0) {
    const int e = M.binElems[b0 + q];
    double det[6][2];
    if (inside_quad(M, e, x, det)) return e;
  }
  return -1;
}

}  // namespace pic
