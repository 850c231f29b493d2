// Device-side views and per-particle routines of the particle stage
// (gfx950).  Every routine restates one routine of the reference (paths
// relative to the reference's src/) with the same operation order; the
// whole translation unit is compiled without floating-point contraction so
// that the geometric predicates of the tracking round exactly like the
// reference's expressions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pic {

struct DevMesh {
  int nElems, N, nDof, deposition, refGuess, nSpecies;
  double refEps;
  // tracking
  const double *corner;     // [e][8][3] tensor corner order
  const int32_t *sideNb;    // [e][6]
  const int32_t *sideGID;   // [e][6]
  const uint8_t *sideNode;  // [e][6][4]
  const double *sideXYZ;    // [e][6][4][3] corner coordinates of every side in ElemSideNodeID order
  const uint8_t *sideBC;    // [e][6]
  const int8_t *sidePV;     // [e][6]
  const uint8_t *concave;   // [e]
  // reference mapping
  const double *dXCL;       // [e][8][3][3]
  const double *bary;       // [e][3]
  const double *radius;     // [e]
  const double *xiBasis;    // [e][6][3]
  const double *slen;       // [e][6]
  const double *xGPe;       // [e][nDof][3]
  const double *E;          // [e][nDof][3]
  double xiCL[2], wBaryCL[2];   // linear CL basis (corner map)
  double cCL[2];                // its product-form Lagrange factors
  double xGP[8], wBaryGP[8];
  double cGP[8];            // 1 / prod_{m != i} (xGP_i - xGP_m): product-form Lagrange factors
  double PV[3][3];
  double ext[6];
  double charge[8], mass[8], mpf[8];
  // reference-mapping tracking (curved NGeo = 2 or linear elements)
  int NGeo, nGP, tracking, nPV;
  const double *XCL;        // [e][nGP][3] XCL_NGeo
  const double *dXCLg;      // [e][nGP][3][3] dXCL_NGeo
  const uint8_t *curved;    // [e] ElemCurved
  const double *epsOneCell; // [e]
  const uint8_t *openElem;  // [e] element with an open boundary side
  const int32_t *binOff, *binElems;
  double bmin[3], bdelta[3];
  int nb[3];
  double gmin[3], gmax[3];
  int dirPV[3];
  double xiCLg[3], wBaryCLg[3];  // CL basis of NGeo
  double cCLg[3];                // its product-form Lagrange factors
  // affine elements (parallelepipeds): corner-map centre and inverse
  // Jacobian, [e][12] = c(3), A^-1 (3x3 row-major); null unless every
  // element is affine
  const double *affine;
  // mixed precision: E and the Gauss basis in FP32
  const float *E32;         // [e][nDof][3]
  float xGP32[8], wBaryGP32[8];
  // RefMapping boundary tracking (ParticleBCTracking); null unless some
  // boundary is not periodic (then cartPeriodic = 0)
  int cartPeriodic;         // every boundary periodic: PeriodicMovement
  double bcHalo;            // BC_halo_eps of the element lists
  const int32_t *bcElemNb;  // [nBC] element across (periodic), -1
  const uint8_t *bcKind;    // [nBC] PIC_BC_*
  const int8_t *bcPV;       // [nBC] periodic vector id
  const uint8_t *bcFlip;    // [nBC]
  const double *bz;         // [nBC][NGeo+1][NGeo+1][3] Bezier control points
  const uint8_t *sideType;  // [nBC]
  const uint8_t *bbEmpty;   // [nBC]
  const double *sideNorm, *sideDist, *slabN, *slabI, *baseV;
  const int32_t *e2bOff;    // [nElems+1] boundary sides in range of each element
  const int32_t *e2bSide;   //   boundary side index
  const double *e2bDist;    //   distance (sorted ascending per element)
  const double *e2bSlab;    //   [6] the side's normal slab (n, lo, hi of n . X, 0)
};

struct DevParts {
  double *x, *y, *z, *vx, *vy, *vz;
  int32_t *elem;    // owner element (0-based), -1 = lost
  int32_t *id;      // caller's particle index
  uint8_t *spec;    // species (0-based)
  uint8_t *isNew;   // PDM%IsNewPart
  double *lx, *ly, *lz;   // LastPartPos
  int32_t *lastElem;      // PEM%LastGlobalElemID
  double *F;              // FieldAtParticle [n][6]
  double *rx, *ry, *rz;   // PartPosRef (RefMapping)
};

constexpr double kEpsMach = 2.220446049250313e-16;

// c^2 and 1/c^2 of globals_vars.f90:87-89
__device__ inline double c2_inv() {
  const double c = 299792458.0;
  return 1.0 / (c * c);
}

// ALMOSTEQUAL_UNITY, interpolation/basis.f90:1011-1022
__device__ inline bool almost_equal_unity(double x, double y) {
  if (x == 0.0 || y == 0.0) return fabs(x - y) <= 2.0 * kEpsMach;
  return (fabs(x - y) <= kEpsMach * fabs(x)) && (fabs(x - y) <= kEpsMach * fabs(y));
}

// LagrangeInterpolationPolys, interpolation/basis.f90:1223-1264
template <int NP>
__device__ inline void lagrange(double x, const double *xg, const double *wb, double *L) {
  bool eq = false;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    L[i] = 0.0;
    if (almost_equal_unity(x, xg[i])) { L[i] = 1.0; eq = true; }
  }
  if (eq) return;
  double sum = 0.0;
#pragma unroll
  for (int i = 0; i < NP; ++i) { L[i] = wb[i] / (x - xg[i]); sum = sum + L[i]; }
#pragma unroll
  for (int i = 0; i < NP; ++i) L[i] = L[i] / sum;
}

// The Lagrange basis at the Gauss points for the field evaluation, in
// product form L_i(x) = c_i prod_{m != i} (x - x_m) with c_i =
// 1 / prod_{m != i} (x_i - x_m): the polynomial of LagrangeInterpolationPolys'
// barycentric form (interpolation/basis.f90:1223-1264; the two agree to
// round-off) with no division per particle; a node hit within ALMOSTEQUAL
// gives the unit vector as there.
template <int NP>
__device__ inline void lagrange_prod(double x, const double *xg, const double *c, double *L) {
  bool eq = false;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    L[i] = 0.0;
    if (almost_equal_unity(x, xg[i])) { L[i] = 1.0; eq = true; }
  }
  if (eq) return;
  double d[NP];
#pragma unroll
  for (int i = 0; i < NP; ++i) d[i] = x - xg[i];
  double pre = 1.0;
#pragma unroll
  for (int i = 0; i < NP; ++i) { L[i] = pre; pre = pre * d[i]; }
  double suf = 1.0;
#pragma unroll
  for (int i = NP - 1; i >= 0; --i) { L[i] = L[i] * suf * c[i]; suf = suf * d[i]; }
}

// getDet / getInv, interpolation/eval_xyz.f90:448-495
__device__ inline double get_det(const double M[3][3]) {
  return (M[0][0] * M[1][1] - M[0][1] * M[1][0]) * M[2][2] + (M[0][1] * M[1][2] - M[0][2] * M[1][1]) * M[2][0] +
         (M[0][2] * M[1][0] - M[0][0] * M[1][2]) * M[2][1];
}

// RefElemNewton for NGeo=1, interpolation/eval_xyz.f90:298-445.
// mode 1 = ForceMode; with wantSuccess the routine reports failure instead
// of aborting.  Returns isSuccessful.
__device__ inline bool ref_elem_newton(const DevMesh &M, int e, const double xin[3], double xi[3], bool wantSuccess) {
  const double *X = M.corner + (size_t)e * 24;
  const double *dX = M.dXCL + (size_t)e * 72;
  double L[3][2], F[3], xo[3];
  bool ok = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) lagrange_prod<2>(xi[d], M.xiCL, M.cCL, L[d]);
#pragma unroll
  for (int c = 0; c < 3; ++c) F[c] = -xin[c];
#pragma unroll
  for (int k = 0; k < 2; ++k)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const double buff = L[1][j] * L[2][k];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int c = 0; c < 3; ++c) F[c] = F[c] + X[(k * 4 + j * 2 + i) * 3 + c] * L[0][i] * buff;
    }
  double deltaXi2 = (fabs(F[0]) < kEpsMach && fabs(F[1]) < kEpsMach && fabs(F[2]) < kEpsMach) ? 0.0 : 1.0;
  double NormF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
  double NormF_old;
  int it = 0;
  while (deltaXi2 > M.refEps && it < 100) {
    ++it;
    // keep the element tables in memory (cache) instead of pinning 96
    // loop-invariant doubles in registers across the iteration
    asm volatile("" ::: "memory");
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll 1
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const double buff = L[1][j] * L[2][k];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const double b2 = L[0][i] * buff;
          const double *dp = dX + (k * 4 + j * 2 + i) * 9;
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
      for (int d = 0; d < 3; ++d) lagrange_prod<2>(xi[d], M.xiCL, M.cCL, L[d]);
#pragma unroll
      for (int c = 0; c < 3; ++c) F[c] = -xin[c];
#pragma unroll
      for (int k = 0; k < 2; ++k)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const double buff = L[1][j] * L[2][k];
#pragma unroll
          for (int i = 0; i < 2; ++i) {
            const double b2 = L[0][i] * buff;
#pragma unroll
            for (int c = 0; c < 3; ++c) F[c] = F[c] + X[(k * 4 + j * 2 + i) * 3 + c] * b2;
          }
        }
      lambda = 0.2 * lambda;
      iArmijo = iArmijo + 1;
      NormF = F[0] * F[0] + F[1] * F[1] + F[2] * F[2];
    }
    if (fabs(xi[0]) > 1.5 || fabs(xi[1]) > 1.5 || fabs(xi[2]) > 1.5) {
      if (wantSuccess) ok = false;
      break;
    }
  }
  return ok;
}

// A wave-uniform pointer in scalar registers: the scalar loads below take
// their address from SGPRs, and a uniform value that the compiler keeps in
// vector registers (across divergent control flow) is read from the first
// active lane (no instruction when it is already scalar).
__device__ inline const double *uniform_ptr(const double *p) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(a)), hi = __builtin_amdgcn_readfirstlane(uint32_t(a >> 32));
  return reinterpret_cast<const double *>((uint64_t(hi) << 32) | uint64_t(lo));
}

// GetRefNewtonStartValue, interpolation/eval_xyz.f90:498-612 (guesses 1, 2, 4)
// the 3 coordinates of one point by scalar loads (p wave-uniform; exactly 24 bytes read)
__device__ inline void sload_xyz(const double *p, double out[3]) {
  p = uniform_ptr(p);
  typedef uint32_t dw4 __attribute__((ext_vector_type(4)));
  typedef uint32_t dw2 __attribute__((ext_vector_type(2)));
  dw4 a;
  dw2 b;
  asm volatile("s_load_dwordx4 %0, %2, 0x0\n\ts_load_dwordx2 %1, %2, 0x10\n\ts_waitcnt lgkmcnt(0)"
               : "=&s"(a), "=&s"(b) : "s"(p));
  out[0] = __builtin_bit_cast(double, (uint64_t(a[1]) << 32) | uint64_t(a[0]));
  out[1] = __builtin_bit_cast(double, (uint64_t(a[3]) << 32) | uint64_t(a[2]));
  out[2] = __builtin_bit_cast(double, (uint64_t(b[1]) << 32) | uint64_t(b[0]));
}

// UNI: e is wave-uniform; guess 2 then reads the Gauss points by scalar loads
template <int N, bool UNI = false>
__device__ inline void ref_newton_start(const DevMesh &M, int e, const double x[3], double xi[3]) {
  if (M.refGuess == 1) {
    const double *b0 = M.bary + (size_t)e * 3;
    const double P[3] = {x[0] - b0[0], x[1] - b0[1], x[2] - b0[2]};
    double XL[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      const double *b = M.xiBasis + ((size_t)e * 6 + d) * 3;
      XL[d] = (P[0] * b[0] + P[1] * b[1] + P[2] * b[2]) * M.slen[(size_t)e * 6 + d];
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) xi[d] = 0.5 * (XL[d] - XL[d + 3]);
    const double mx = fmax(fabs(xi[0]), fmax(fabs(xi[1]), fabs(xi[2])));
    if (mx > 1.0 + M.refEps)
#pragma unroll
      for (int d = 0; d < 3; ++d) xi[d] = fmax(fmin(1.0, xi[d]), -1.0);
  } else if (M.refGuess == 2) {
    constexpr int n1 = N + 1;
    const double *xg = M.xGPe + (size_t)e * n1 * n1 * n1 * 3;
    double g[3];
    if (UNI) sload_xyz(xg, g);
    else { g[0] = xg[0]; g[1] = xg[1]; g[2] = xg[2]; }
    const double d0[3] = {x[0] - g[0], x[1] - g[1], x[2] - g[2]};
    double win = sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
    xi[0] = xi[1] = xi[2] = M.xGP[0];
#pragma unroll 1
    for (int i = 0; i < n1; ++i)
#pragma unroll 1
      for (int j = 0; j < n1; ++j)
#pragma unroll 1
        for (int k = 0; k < n1; ++k) {
          const int r = (k * n1 + j) * n1 + i;
          double dX, dY, dZ;
          if (UNI) {
            sload_xyz(xg + r * 3, g);
            dX = fabs(x[0] - g[0]);
            if (dX > win) continue;
            dY = fabs(x[1] - g[1]);
            if (dY > win) continue;
            dZ = fabs(x[2] - g[2]);
            if (dZ > win) continue;
          } else {
            dX = fabs(x[0] - xg[r * 3]);
            if (dX > win) continue;
            dY = fabs(x[1] - xg[r * 3 + 1]);
            if (dY > win) continue;
            dZ = fabs(x[2] - xg[r * 3 + 2]);
            if (dZ > win) continue;
          }
          const double dist = sqrt(dX * dX + dY * dY + dZ * dZ);
          if (dist < win) { win = dist; xi[0] = M.xGP[i]; xi[1] = M.xGP[j]; xi[2] = M.xGP[k]; }
        }
  } else {
    xi[0] = xi[1] = xi[2] = 0.0;
  }
}

// EvaluateFieldAtRefPos with NVar=3 (interpolation/eval_xyz.f90:167-295),
// summed by factors: U = sum_k L_k sum_j L_j sum_i L_i U_ijk (the reference
// sums sum_ijk U_ijk L_i (L_j L_k); the two agree to round-off).
// (Ue: the element's field values)
template <int N>
__device__ inline void eval_field_p(const DevMesh &M, const double *Ue, const double xi[3], double U[3]) {
#pragma clang fp contract(fast)
  constexpr int n1 = N + 1;
  double L[3][n1];
#pragma unroll
  for (int d = 0; d < 3; ++d) lagrange_prod<n1>(xi[d], M.xGP, M.cGP, L[d]);
  U[0] = U[1] = U[2] = 0.0;
#pragma unroll 1
  for (int k = 0; k < n1; ++k) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int j = 0; j < n1; ++j) {
      double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
      for (int i = 0; i < n1; ++i) {
        const double *u = Ue + ((k * n1 + j) * n1 + i) * 3;
        t0 += u[0] * L[0][i];
        t1 += u[1] * L[0][i];
        t2 += u[2] * L[0][i];
      }
      s0 += t0 * L[1][j];
      s1 += t1 * L[1][j];
      s2 += t2 * L[1][j];
    }
    U[0] += s0 * L[2][k];
    U[1] += s1 * L[2][k];
    U[2] += s2 * L[2][k];
  }
}
template <int N>
__device__ inline void eval_field(const DevMesh &M, int e, const double xi[3], double U[3]) {
  eval_field_p<N>(M, M.E + (size_t)e * (N + 1) * (N + 1) * (N + 1) * 3, xi, U);
}

// ---- wave-uniform element: table reads by scalar loads ---------------------
// When every lane of a wave maps in the same element, the element's tables
// are read once per wave into scalar registers (s_load, loads only) instead
// of once per lane through the vector L1, where the replicated reads of the
// 27-point map bound the Newton.  The arithmetic is that of newton_map.
typedef uint32_t sdw16 __attribute__((ext_vector_type(16)));
typedef uint32_t sdw8 __attribute__((ext_vector_type(8)));
typedef uint32_t sdw4 __attribute__((ext_vector_type(4)));
typedef uint32_t sdw2 __attribute__((ext_vector_type(2)));

__device__ inline double dw_to_d(uint32_t lo, uint32_t hi) {
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | uint64_t(lo));
}

// values [16c, min(16c + 16, NV)) of the table at p (p wave-uniform; exact
// size, nothing past the table is read)
template <int NV>
__device__ inline void sload_chunk(const double *p, int c, double buf[16]) {
  const double *q = uniform_ptr(p + 16 * c);
  const int m = NV - 16 * c;
  if (m >= 16) {
    // two loads in flight, one wait
    sdw16 r0, r1;
    asm volatile("s_load_dwordx16 %0, %2, 0x0\n\ts_load_dwordx16 %1, %2, 0x40\n\ts_waitcnt lgkmcnt(0)"
                 : "=&s"(r0), "=&s"(r1) : "s"(q));
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      buf[k] = dw_to_d(r0[2 * k], r0[2 * k + 1]);
      buf[8 + k] = dw_to_d(r1[2 * k], r1[2 * k + 1]);
    }
    return;
  }
  int k = 0;
  if (m & 8) {
    sdw16 r;
    asm volatile("s_load_dwordx16 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=s"(r) : "s"(q));
#pragma unroll
    for (int t = 0; t < 8; ++t) buf[t] = dw_to_d(r[2 * t], r[2 * t + 1]);
    k = 8;
  }
  if (m & 4) {
    sdw8 r;
    asm volatile("s_load_dwordx8 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=s"(r) : "s"(q + k));
#pragma unroll
    for (int t = 0; t < 4; ++t) buf[k + t] = dw_to_d(r[2 * t], r[2 * t + 1]);
    k += 4;
  }
  if (m & 2) {
    sdw4 r;
    asm volatile("s_load_dwordx4 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=s"(r) : "s"(q + k));
    buf[k] = dw_to_d(r[0], r[1]);
    buf[k + 1] = dw_to_d(r[2], r[3]);
    k += 2;
  }
  if (m & 1) {
    sdw2 r;
    asm volatile("s_load_dwordx2 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=s"(r) : "s"(q + k));
    buf[k] = dw_to_d(r[0], r[1]);
  }
}

// eval_field for a wave whose lanes are all in element e (e wave-uniform):
// the element's E is read one k-plane at a time by scalar loads; same
// arithmetic as eval_field.
template <int N>
__device__ inline void eval_field_s(const DevMesh &M, int e, const double xi[3], double U[3]) {
#pragma clang fp contract(fast)
  constexpr int n1 = N + 1, NP = n1 * n1 * 3;
  double L[3][n1];
#pragma unroll
  for (int d = 0; d < 3; ++d) lagrange_prod<n1>(xi[d], M.xGP, M.cGP, L[d]);
  const double *Ue = M.E + (size_t)e * n1 * n1 * n1 * 3;
  U[0] = U[1] = U[2] = 0.0;
#pragma unroll 1
  for (int k = 0; k < n1; ++k) {
    double s[3] = {0.0, 0.0, 0.0}, t[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int ch = 0; ch < (NP + 15) / 16; ++ch) {
      double buf[16];
      sload_chunk<NP>(Ue + k * NP, ch, buf);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int v = 16 * ch + q;
        if (v < NP) {
          const int c = v % 3, i = (v / 3) % n1, j = v / (3 * n1);
          if (i == 0) t[c] = 0.0;
          t[c] += buf[q] * L[0][i];
          if (i == n1 - 1) s[c] += t[c] * L[1][j];
        }
      }
    }
    U[0] += s[0] * L[2][k];
    U[1] += s[1] * L[2][k];
    U[2] += s[2] * L[2][k];
  }
}

// GetEMFieldDW (E only), particles/pic/interpolation/pic_interpolation_tools.f90:458-572.
// The reference's loop nest is DO k; DO l; DO m, and the EXIT taken when the
// particle sits on a Gauss point leaves only the m-loop: weights of points
// visited before are zeroed, that point weighs 1, the rest of its m-row is
// skipped and the loop continues with the next l.  Two passes restate that
// without a per-point weight array.
// Kept out of line and unrolled once: it runs only where Newton failed, and
// inlined into the push kernels it raised their register count (affine push
// 118 -> 188 VGPRs).
template <int N>
__device__ __attribute__((noinline)) void em_field_dw(const DevMesh &M, int e, const double x[3], double U[3]) {
  constexpr int n1 = N + 1, nd = n1 * n1 * n1;
  const double *xg = M.xGPe + (size_t)e * nd * 3;
  const double *Ue = M.E + (size_t)e * nd * 3;
  auto norm_of = [&](int r) {
    const double d0 = xg[r * 3] - x[0], d1 = xg[r * 3 + 1] - x[1], d2 = xg[r * 3 + 2] - x[2];
    return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  };
  // pass 1: DistSum and the (last) Gauss point hit, in loop order
  double sum = 0.0;
  int hitK = -1, hitL = -1, hitM = -1;
#pragma unroll 1
  for (int k = 0; k < n1; ++k)
#pragma unroll 1
    for (int l = 0; l < n1; ++l)
#pragma unroll 1
      for (int m = 0; m < n1; ++m) {
        const double nrm = norm_of((m * n1 + l) * n1 + k);
        if (nrm > 0.) sum = sum + 1. / nrm;
        else { sum = 1.0; hitK = k; hitL = l; hitM = m; break; }
      }
  // pass 2: weights; before the hit they are zero, in the hit's m-row after it too
  U[0] = U[1] = U[2] = 0.0;
#pragma unroll 1
  for (int k = 0; k < n1; ++k)
#pragma unroll 1
    for (int l = 0; l < n1; ++l)
#pragma unroll 1
      for (int m = 0; m < n1; ++m) {
        const int r = (m * n1 + l) * n1 + k;
        double w;
        if (hitK < 0) w = 1. / norm_of(r);
        else {
          const int ord = (k * n1 + l) * n1 + m, hord = (hitK * n1 + hitL) * n1 + hitM;
          if (ord < hord) w = 0.0;
          else if (ord == hord) w = 1.0;
          else if (k == hitK && l == hitL) w = 0.0;
          else { const double nrm = norm_of(r); w = nrm > 0. ? 1. / nrm : 1.0; }
        }
        U[0] = U[0] + w / sum * Ue[r * 3 + 0];
        U[1] = U[1] + w / sum * Ue[r * 3 + 1];
        U[2] = U[2] + w / sum * Ue[r * 3 + 2];
      }
}

// Mixed precision: LagrangeInterpolationPolys and EvaluateFieldAtRefPos in
// FP32 (nodes equal to the FP32 machine epsilon count as hit).  The FP32
// basis keeps the barycentric form w_i / (x - x_i) / sum_j w_j / (x - x_j):
// its weights' FP32 rounding cancels in the quotient, where the product
// form's factors c_i rounded to FP32 bias every particle's field the same
// way (kinetic energies of a 1000-step run 2.9e-8 apart from the FP64 run
// against 1.6e-11 with this form).
template <int NP>
__device__ inline void lagrange_f32(float x, const float *xg, const float *wb, float *L) {
  bool eq = false;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    L[i] = 0.f;
    const float d = fabsf(x - xg[i]);
    if ((x == 0.f || xg[i] == 0.f) ? d <= 2.f * 1.1920929e-7f
                                   : (d <= 1.1920929e-7f * fabsf(x) && d <= 1.1920929e-7f * fabsf(xg[i]))) {
      L[i] = 1.f;
      eq = true;
    }
  }
  if (eq) return;
  float sum = 0.f;
#pragma unroll
  for (int i = 0; i < NP; ++i) { L[i] = wb[i] / (x - xg[i]); sum = sum + L[i]; }
#pragma unroll
  for (int i = 0; i < NP; ++i) L[i] = L[i] / sum;
}

template <int N>
__device__ inline void eval_field_f32_p(const DevMesh &M, const float *Ue, const double xi[3], double U[3]) {
  constexpr int n1 = N + 1;
  float L[3][n1];
#pragma unroll
  for (int d = 0; d < 3; ++d) lagrange_f32<n1>(float(xi[d]), M.xGP32, M.wBaryGP32, L[d]);
  float u0 = 0.f, u1 = 0.f, u2 = 0.f;
#pragma unroll 1
  for (int k = 0; k < n1; ++k)
#pragma unroll 1
    for (int j = 0; j < n1; ++j) {
      const float Lez = L[1][j] * L[2][k];
#pragma unroll
      for (int i = 0; i < n1; ++i) {
        const float *u = Ue + ((k * n1 + j) * n1 + i) * 3;
        const float w = L[0][i] * Lez;
        u0 = u0 + u[0] * w;
        u1 = u1 + u[1] * w;
        u2 = u2 + u[2] * w;
      }
    }
  U[0] = u0; U[1] = u1; U[2] = u2;
}
template <int N>
__device__ inline void eval_field_f32(const DevMesh &M, int e, const double xi[3], double U[3]) {
  eval_field_f32_p<N>(M, M.E32 + (size_t)e * (N + 1) * (N + 1) * (N + 1) * 3, xi, U);
}

// eval_field_f32 for a wave whose lanes are all in element e (e wave-uniform):
// the FP32 field read one k-plane at a time by scalar loads; same arithmetic
template <int N>
__device__ inline void eval_field_f32_s(const DevMesh &M, int e, const double xi[3], double U[3]) {
  constexpr int n1 = N + 1, NP = n1 * n1 * 3;
  float L[3][n1];
#pragma unroll
  for (int d = 0; d < 3; ++d) lagrange_f32<n1>(float(xi[d]), M.xGP32, M.wBaryGP32, L[d]);
  if constexpr (NP % 2 != 0) {   // planes not 8-byte aligned: the vector loads (mixed runs at N = 3)
    eval_field_f32<N>(M, e, xi, U);
    return;
  }
  const float *Ue = M.E32 + (size_t)e * n1 * n1 * n1 * 3;
  float u[3] = {0.f, 0.f, 0.f};
#pragma unroll 1
  for (int k = 0; k < n1; ++k) {
    float Lez = 0.f, w = 0.f;
    // the plane's NP floats as NP/2 8-byte words (plane offsets are 8-byte aligned)
    const double *Pk = reinterpret_cast<const double *>(Ue + k * NP);
#pragma unroll
    for (int ch = 0; ch < (NP / 2 + 15) / 16; ++ch) {
      double dbuf[16];
      sload_chunk<NP / 2>(Pk, ch, dbuf);
      float buf[32];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const uint64_t bits = __builtin_bit_cast(uint64_t, dbuf[t]);
        buf[2 * t] = __builtin_bit_cast(float, uint32_t(bits));
        buf[2 * t + 1] = __builtin_bit_cast(float, uint32_t(bits >> 32));
      }
#pragma unroll
      for (int q = 0; q < 32; ++q) {
        const int v = 32 * ch + q;
        if (v < NP) {
          const int c = v % 3, i = (v / 3) % n1, j = v / (3 * n1);
          if (c == 0) {
            if (i == 0) Lez = L[1][j] * L[2][k];
            w = L[0][i] * Lez;
          }
          u[c] = u[c] + buf[q] * w;
        }
      }
    }
  }
  U[0] = u[0]; U[1] = u[1]; U[2] = u[2];
}

// InterpolateFieldToSingleParticle for the 508 stepper
// (particles/pic/interpolation/pic_interpolation.f90:325-369):
// external field + E(xi) with xi from GetPositionInRefElem.
// On affine elements (parallelepipeds) the corner map is X(xi) = c + A xi and
// RefElemNewton converges to xi = A^-1 (x - c), which is formed directly
// (the two agree to round-off).  Used when every element is affine.
template <bool UNI = false>
__device__ inline void affine_xi(const DevMesh &M, int e, const double x[3], double xi[3]) {
  double buf[16];
  const double *a = M.affine + (size_t)e * 12;
  if (UNI) {
    sload_chunk<12>(a, 0, buf);
    a = buf;
  }
  const double d0 = x[0] - a[0], d1 = x[1] - a[1], d2 = x[2] - a[2];
  xi[0] = a[3] * d0 + a[4] * d1 + a[5] * d2;
  xi[1] = a[6] * d0 + a[7] * d1 + a[8] * d2;
  xi[2] = a[9] * d0 + a[10] * d1 + a[11] * d2;
}

// UNI: e is wave-uniform (every lane of the wave in element e), E by scalar loads
template <int N, bool MIXED = false, bool AFFINE = false, bool UNI = false>
__device__ inline void field_at_particle(const DevMesh &M, int e, const double x[3], double F[6]) {
#pragma unroll
  for (int c = 0; c < 6; ++c) F[c] = M.ext[c];
  double xi[3];
  bool ok = true;
  if (AFFINE) {
    affine_xi<UNI>(M, e, x, xi);
    // RefElemNewton's first step lands on this xi and gives up when a
    // component exceeds 1.5: the particle is then far
    // outside its element and isSuccessful is .FALSE. (eval_xyz.f90:424-441)
    ok = !(fabs(xi[0]) > 1.5 || fabs(xi[1]) > 1.5 || fabs(xi[2]) > 1.5);
  } else {
    ref_newton_start<N>(M, e, x, xi);
    ok = ref_elem_newton(M, e, x, xi, true);
  }
  double U[3];
  if (!ok && M.deposition == 1) em_field_dw<N>(M, e, x, U);
  else if (MIXED && UNI) eval_field_f32_s<N>(M, e, xi, U);
  else if (MIXED) eval_field_f32<N>(M, e, xi, U);
  else if (UNI) eval_field_s<N>(M, e, xi, U);
  else eval_field<N>(M, e, xi, U);
  F[0] = F[0] + U[0];
  F[1] = F[1] + U[1];
  F[2] = F[2] + U[2];
}

// Boris-Leapfrog push of one particle,
// timedisc/timedisc_TimeStepPoissonByBorisLeapfrog.f90:128-198 (with the
// IsNewPart half step through PartRHS_NR, particles/particle_rhs.f90:203-253)
__device__ inline void boris_push(double q, double m, bool isNew, const double F[6], double dt, double x[3], double v[3]) {
  const bool push = fabs(q) > 0.0;
  if (isNew && push) {
    const double qmt = q / m;
    const double Pt[3] = {F[0] * qmt, F[1] * qmt, F[2] * qmt};
    v[0] = v[0] - Pt[0] * dt * 0.5;
    v[1] = v[1] - Pt[1] * dt * 0.5;
    v[2] = v[2] - Pt[2] * dt * 0.5;
  }
  if (push) {
    const double c2i = c2_inv();
    const double c_1 = (q * dt) / (m * 2.);
    const double gamma = 1. / sqrt(1 - ((v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * c2i));
    const double vmo[3] = {v[0] * gamma, v[1] * gamma, v[2] * gamma};
    const double vm[3] = {vmo[0] + c_1 * F[0], vmo[1] + c_1 * F[1], vmo[2] + c_1 * F[2]};
    const double gm = sqrt(1 + (vm[0] * vm[0] + vm[1] * vm[1] + vm[2] * vm[2]) * c2i);
    const double Bn = sqrt(F[3] * F[3] + F[4] * F[4] + F[5] * F[5]);
    double ub[3] = {0.0, 0.0, 0.0};
    double invL = sqrt(F[3] * F[3] + F[4] * F[4] + F[5] * F[5]);
    if (fabs(invL) > 0.0) { invL = 1. / invL; ub[0] = F[3] * invL; ub[1] = F[4] * invL; ub[2] = F[5] * invL; }
    const double tf = tan(c_1 / gm * Bn);
    const double t[3] = {tf * ub[0], tf * ub[1], tf * ub[2]};
    const double vp[3] = {vm[0] + (vm[1] * t[2] - vm[2] * t[1]), vm[1] + (vm[2] * t[0] - vm[0] * t[2]),
                          vm[2] + (vm[0] * t[1] - vm[1] * t[0])};
    const double f2 = 2.0 / (1. + (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
    const double cr[3] = {vp[1] * t[2] - vp[2] * t[1], vp[2] * t[0] - vp[0] * t[2], vp[0] * t[1] - vp[1] * t[0]};
    const double vpl[3] = {vm[0] + f2 * cr[0], vm[1] + f2 * cr[1], vm[2] + f2 * cr[2]};
    const double vn1[3] = {vpl[0] + c_1 * F[0], vpl[1] + c_1 * F[1], vpl[2] + c_1 * F[2]};
    const double den = sqrt(1 + (vn1[0] * vn1[0] + vn1[1] * vn1[1] + vn1[2] * vn1[2]) * c2i);
    // v_n+1 / gamma_n+1 as one reciprocal and three products (round-off apart)
    const double rden = 1.0 / den;
    v[0] = vn1[0] * rden;
    v[1] = vn1[1] * rden;
    v[2] = vn1[2] * rden;
  }
  x[0] = x[0] + v[0] * dt;
  x[1] = x[1] + v[1] * dt;
  x[2] = x[2] + v[2] * dt;
}

// ParticleInsideQuad3D, particles/particle_mesh/particle_mesh_tools.f90:78-222
// UNI: e is wave-uniform, the side corners come by scalar loads
template <bool UNI = false>
__device__ inline bool inside_quad(const DevMesh &M, int e, const double x[3], double det[6][2]) {
  const uint8_t cc = M.concave[e];
  bool in = true;
#pragma unroll 1
  for (int l = 0; l < 6; ++l) {
    const double *sx = M.sideXYZ + ((size_t)e * 6 + l) * 12;
    double A[3][4];
    if (UNI) {
      double buf[16];
      sload_chunk<12>(sx, 0, buf);
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) A[c][q] = buf[q * 3 + c] - x[c];
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) A[c][q] = sx[q * 3 + c] - x[c];
    }
    const double cp0 = A[1][0] * A[2][2] - A[2][0] * A[1][2];
    const double cp1 = A[2][0] * A[0][2] - A[0][0] * A[2][2];
    const double cp2 = A[0][0] * A[1][2] - A[1][0] * A[0][2];
    double d1 = cp0 * A[0][1] + cp1 * A[1][1] + cp2 * A[2][1];
    d1 = -d1;
    const double d2 = cp0 * A[0][3] + cp1 * A[1][3] + cp2 * A[2][3];
    det[l][0] = d1;
    det[l][1] = d2;
    const bool neg = (d1 < 0) || (d2 < 0);
    const bool pos = !(d1 < 0) || !(d2 < 0);
    if ((cc >> l) & 1) { if (!pos) in = false; }
    else { if (neg) in = false; }
  }
  return in;
}

// ParticleThroughSideCheck3DFast, particles/tracking/particle_intersection.f90:167-280
__device__ inline bool through_side_fast(const DevMesh &M, int e, int l, int tri, const double P[3], const double V[3]) {
  const double *sx = M.sideXYZ + ((size_t)e * 6 + l) * 12;
  double Ax[3], Ay[3], Az[3];
  Ax[0] = sx[0] - P[0];
  Ay[0] = sx[1] - P[1];
  Az[0] = sx[2] - P[2];
#pragma unroll
  for (int n = 1; n < 3; ++n) {
    const int node = n + tri - 1;  // 0-based side node for n = 1,2
    Ax[n] = sx[node * 3 + 0] - P[0];
    Ay[n] = sx[node * 3 + 1] - P[1];
    Az[n] = sx[node * 3 + 2] - P[2];
  }
  double dc[3][3];
  dc[0][0] = (Ay[0] * V[2] - Az[0] * V[1]) * Ax[2];
  dc[0][1] = (Az[0] * V[0] - Ax[0] * V[2]) * Ay[2];
  dc[0][2] = (Ax[0] * V[1] - Ay[0] * V[0]) * Az[2];
  dc[1][0] = (Ay[1] * V[2] - Az[1] * V[1]) * Ax[0];
  dc[1][1] = (Az[1] * V[0] - Ax[1] * V[2]) * Ay[0];
  dc[1][2] = (Ax[1] * V[1] - Ay[1] * V[0]) * Az[0];
  dc[2][0] = (Ay[2] * V[2] - Az[2] * V[1]) * Ax[1];
  dc[2][1] = (Az[2] * V[0] - Ax[2] * V[2]) * Ay[1];
  dc[2][2] = (Ax[2] * V[1] - Ay[2] * V[0]) * Az[1];
  bool through = true;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const double det = dc[r][0] + dc[r][1] + dc[r][2];
    double mn = 1.7976931348623157e308;  // MINVAL over an empty mask is HUGE
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double a = fabs(dc[r][c]);
      if (a > 0.0 && a < mn) mn = a;
    }
    if (!(det >= -kEpsMach * mn)) through = false;
  }
  return through;
}

// ParticleThroughSideLastPosCheck (regular side), particle_intersection.f90:430-512
__device__ inline bool through_side_last_pos(const DevMesh &M, int e, int l, int tri, const double P[3], double &det) {
  const double *sx = M.sideXYZ + ((size_t)e * 6 + l) * 12;
  const int nodes[3] = {0, tri, tri + 1};
  double Ax[3], Ay[3], Az[3];
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    Ax[q] = sx[nodes[q] * 3 + 0] - P[0];
    Ay[q] = sx[nodes[q] * 3 + 1] - P[1];
    Az[q] = sx[nodes[q] * 3 + 2] - P[2];
  }
  det = ((Ay[0] * Az[1] - Az[0] * Ay[1]) * Ax[2] + (Az[0] * Ax[1] - Ax[0] * Az[1]) * Ay[2] +
         (Ax[0] * Ay[1] - Ay[0] * Ax[1]) * Az[2]);
  return !((det < 0) || (det != det));
}

// IntersectionWithWall, particle_intersection.f90:79-164
__device__ inline double intersection_with_wall(const DevMesh &M, int e, int l, int tri, const double P[3], const double traj[3]) {
  const double *sx = M.sideXYZ + ((size_t)e * 6 + l) * 12;
  const double xN = sx[0], yN = sx[1], zN = sx[2];
  const int n1 = tri, n2 = tri + 1;
  const double V1[3] = {sx[n1 * 3] - xN, sx[n1 * 3 + 1] - yN, sx[n1 * 3 + 2] - zN};
  const double V2[3] = {sx[n2 * 3] - xN, sx[n2 * 3 + 1] - yN, sx[n2 * 3 + 2] - zN};
  double nx = V1[1] * V2[2] - V1[2] * V2[1];
  double ny = V1[2] * V2[0] - V1[0] * V2[2];
  double nz = V1[0] * V2[1] - V1[1] * V2[0];
  const double nVal = sqrt(nx * nx + ny * ny + nz * nz);
  nx = nx / nVal; ny = ny / nVal; nz = nz / nVal;
  const double bx = P[0] - xN, by = P[1] - yN, bz = P[2] - zN;
  const double ax = bx - nx * (bx * nx + by * ny + bz * nz);
  const double ay = by - ny * (bx * nx + by * ny + bz * nz);
  const double az = bz - nz * (bx * nx + by * ny + bz * nz);
  double dist = sqrt(((ay * bz - az * by) * (ay * bz - az * by) + (az * bx - ax * bz) * (az * bx - ax * bz) +
                      (ax * by - ay * bx) * (ax * by - ay * bx)) /
                     (ax * ax + ay * ay + az * az));
  if (dist != dist) dist = sqrt(bx * bx + by * by + bz * bz);
  double alpha = traj[0] * nx + traj[1] * ny + traj[2] * nz;
  if (fabs(alpha) > 0.) alpha = dist / alpha;
  return alpha;
}

// SingleParticleTriaTracking3D, particles/tracking/particle_triatracking.f90:137-484
// (regular sides; open and periodic particle boundaries, particle_boundary_condition.f90:35-284).
// x and last are updated in place by periodic shifts.  Returns the new owner
// element, or -1 when the particle is lost or removed; *removed_open tells
// the two apart.
__device__ inline int tria_track(const DevMesh &M, int elem, double x[3], double last[3], bool &lost, int &steps) {
  int DoneLastElem[6][4];
#pragma unroll
  for (int a = 0; a < 6; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) DoneLastElem[a][b] = 0;
  double traj[3] = {0, 0, 0}, len = 0.0;
  lost = false;
  steps = 0;
  for (int guard = 0; guard < 4096; ++guard) {
    ++steps;
    double det[6][2];
    if (inside_quad(M, elem, x, det)) return elem;
    traj[0] = x[0] - last[0];
    traj[1] = x[1] - last[1];
    traj[2] = x[2] - last[2];
    len = sqrt(traj[0] * traj[0] + traj[1] * traj[1] + traj[2] * traj[2]);
    if (fabs(len) > 0.) { traj[0] = traj[0] / len; traj[1] = traj[1] / len; traj[2] = traj[2] / len; }
    int nThrough = 0, locS[12], triS[12];
    int side = -1, tri = 0;
    for (int l = 0; l < 6; ++l)
      for (int t = 1; t <= 2; ++t)
        if (det[l][t - 1] <= -0.0 && through_side_fast(M, elem, l, t, last, traj)) {
          locS[nThrough] = l;
          triS[nThrough] = t;
          ++nThrough;
          side = l;
        }
    tri = nThrough > 0 ? triS[0] : 0;
    if (nThrough != 1) {
      if (nThrough == 0) { lost = true; return -1; }
      int second = 0;
      double minRatio = 0;
      for (int q = 0; q < nThrough; ++q) {
        const int gid = M.sideGID[(size_t)elem * 6 + locS[q]];
        bool doCheck = true;
        for (int is = 1; is < 6; ++is)
          if (DoneLastElem[is][0] == elem + 1 && DoneLastElem[is][3] == gid && DoneLastElem[is][2] == triS[q]) doCheck = false;
        if (!doCheck) continue;
        double detM;
        if (through_side_last_pos(M, elem, locS[q], triS[q], last, detM)) {
          const double dP = det[locS[q]][triS[q] - 1];
          if (detM == 0 && dP == 0) continue;
          if (detM == 0 && minRatio == 0) {
            ++second; side = locS[q]; tri = triS[q];
          } else {
            if (detM == 0) continue;
            const double ratio = dP / detM;
            if (ratio < minRatio) { minRatio = ratio; ++second; side = locS[q]; tri = triS[q]; }
          }
        }
      }
      if (second == 0) { lost = true; return -1; }
    }
    const size_t si = (size_t)elem * 6 + side;
    const int bc = M.sideBC[si];
    const int gid = M.sideGID[si];
    if (bc != 0) {
      const int oldElem = elem;
      if (bc == 1) {  // open boundary: RemoveParticle
        lost = false;
        return -1;
      }
      // periodic: IntersectionWithWall then PeriodicBoundary
      const double alpha = intersection_with_wall(M, elem, side, tri, last, traj);
      const int pvid = M.sidePV[si];
      const int v = (pvid < 0 ? -pvid : pvid) - 1;
      for (int c = 0; c < 3; ++c) last[c] = last[c] + traj[c] * alpha;
      for (int c = 0; c < 3; ++c) last[c] = last[c] + copysign(fabs(M.PV[v][c]), (double)pvid);
      for (int c = 0; c < 3; ++c) x[c] = last[c] + (len - alpha) * traj[c];
      elem = M.sideNb[si];
      for (int k = 4; k >= 0; --k)
        for (int b = 0; b < 4; ++b) DoneLastElem[k + 1][b] = DoneLastElem[k][b];
      DoneLastElem[0][0] = oldElem + 1;
      DoneLastElem[0][1] = side + 1;
      DoneLastElem[0][2] = tri;
      DoneLastElem[0][3] = gid;
    } else {
      for (int k = 4; k >= 0; --k)
        for (int b = 0; b < 4; ++b) DoneLastElem[k + 1][b] = DoneLastElem[k][b];
      DoneLastElem[0][0] = elem + 1;
      DoneLastElem[0][1] = side + 1;
      DoneLastElem[0][2] = tri;
      DoneLastElem[0][3] = gid;
      elem = M.sideNb[si];
    }
    if (elem < 0 || elem >= M.nElems) { lost = true; return -1; }
  }
  lost = true;
  return -1;
}

}  // namespace pic
