// Self-consistent step through a fixed linear field operator, and the
// hand-written adjoint of interpolate + push + deposit over K such steps.
// Included once, at the end of pic_stage.hip (it uses the handle and the
// launch helpers defined there).
//
// Forward step n (state x^n, v^n sorted by element):
//   rho^n   = Dep(x^n)                      shape-function deposition (rho component)
//   rbar^n  = A rho^n                       rbar_e = sum_ijk w_i w_j w_k rho_e,ijk / 8
//   En^n    = G rbar^n                      E at the unique corner nodes (G given)
//   E^n     = P En^n                        trilinear corner interpolation at the Gauss points
//   F^n     = ext + E^n(x^n)                InterpolateFieldToParticle
//   v^{n+1} = Boris(v^n, F^n),  x^{n+1} = x^n + v^{n+1} dt   (then tracking and re-sort)
// Loss after K steps: W = eps0/2 sum_e sum_ijk w_i w_j w_k / sJ |E^K|^2 with
// E^K = P G A Dep(x^K) (CalcPotentialEnergy of the final deposition).
//
// Reverse sweep, adjoints lx, lv held per particle id:
//   lE^K = eps0 w / sJ E^K ;  lrbar = G^T P^T lE ;  lx += Dep_x^T (A^T lrbar)
//   step n, backwards:  lv' = lv + dt lx ;  (lv, lF) = (dB/dv, dB/dF)^T lv' ;
//                        lx += (dE/dx)^T lF ;  lE^n = sum_p L(xi_p) lF_p ;
//                        lrbar = G^T P^T lE^n ;  lx += Dep_x^T (A^T lrbar)
// The states of every step are recomputed from checkpoints taken every
// `every` steps, segment by segment.

namespace {

// ---- field operator kernels ---------------------------------------------------
template <int N>
__global__ __launch_bounds__(kBlock) void k_elem_mean(int nElems, const double *PS, QuadArgs Q, double *rbar) {
  constexpr int n1 = N + 1, nd = n1 * n1 * n1;
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= nElems) return;
  double acc = 0.0;
  for (int k = 0; k < n1; ++k)
    for (int j = 0; j < n1; ++j)
#pragma unroll
      for (int i = 0; i < n1; ++i) {
        const int r = (k * n1 + j) * n1 + i;
        acc = acc + Q.wGP[i] * Q.wGP[j] * Q.wGP[k] * PS[((size_t)e * nd + r) * 4 + 3];
      }
  rbar[e] = acc / 8.0;
}

// y = M x for a dense row-major M[rows][cols]: one wave per row, lane-strided
// partial sums reduced by a butterfly (a fixed order)
__global__ __launch_bounds__(kBlock) void k_matvec(int64_t rows, int64_t cols, const double *Mx, const double *x, double *y) {
  const int64_t r = (int64_t(blockIdx.x) * kBlock + threadIdx.x) / kWave;
  const int lane = lane_id();
  if (r >= rows) return;
  const double *m = Mx + r * cols;
  double acc = 0.0;
  for (int64_t c = lane; c < cols; c += kWave) acc += m[c] * x[c];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc = acc + __shfl_xor(acc, off, kWave);
  if (lane == 0) y[r] = acc;
}

// trilinear weight of CGNS corner q at the Gauss point (a1, a2, a3 in [0,1])
__device__ inline double corner_weight(int q, double a1, double a2, double a3) {
  switch (q) {
    case 0: return (1 - a1) * (1 - a2) * (1 - a3);
    case 1: return a1 * (1 - a2) * (1 - a3);
    case 2: return a1 * a2 * (1 - a3);
    case 3: return (1 - a1) * a2 * (1 - a3);
    case 4: return (1 - a1) * (1 - a2) * a3;
    case 5: return a1 * (1 - a2) * a3;
    case 6: return a1 * a2 * a3;
    default: return (1 - a1) * a2 * a3;
  }
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_node_to_dof(DevMesh M, const int32_t *elemNode, const double *En, double *E) {
  constexpr int n1 = N + 1, nd = n1 * n1 * n1;
  const int64_t g = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (g >= int64_t(M.nElems) * nd) return;
  const int e = int(g / nd), r = int(g % nd);
  const int i = r % n1, j = (r / n1) % n1, k = r / (n1 * n1);
  const double a1 = (M.xGP[i] + 1.0) / 2.0, a2 = (M.xGP[j] + 1.0) / 2.0, a3 = (M.xGP[k] + 1.0) / 2.0;
  const int32_t *u = elemNode + (size_t)e * 8;
  for (int c = 0; c < 3; ++c) {
    double v = 0.0;
    for (int q = 0; q < 8; ++q) v = v + corner_weight(q, a1, a2, a3) * En[(size_t)u[q] * 3 + c];
    E[g * 3 + c] = v;
  }
}

// ---- adjoint kernels --------------------------------------------------------------
// lE = eps0 w_i w_j w_k / sJ E   (gradient of the field energy)
template <int N>
__global__ __launch_bounds__(kBlock) void k_adj_energy(int nElems, const double *E, const double *sJ, QuadArgs Q, double *lE) {
  constexpr int n1 = N + 1, nd = n1 * n1 * n1;
  const int64_t g = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (g >= int64_t(nElems) * nd) return;
  const int r = int(g % nd);
  const int i = r % n1, j = (r / n1) % n1, k = r / (n1 * n1);
  const double f = 8.8541878176e-12 * (Q.wGP[i] * Q.wGP[j] * Q.wGP[k] / sJ[g]);
  for (int c = 0; c < 3; ++c) lE[g * 3 + c] = f * E[g * 3 + c];
}

// P^T: per unique node, its incidences (element, corner) in element order
template <int N>
__global__ __launch_bounds__(kBlock) void k_adj_node(DevMesh M, int nNodes, const int32_t *off, const int32_t *inc,
                                                     const double *lE, double *lEn) {
  constexpr int n1 = N + 1, nd = n1 * n1 * n1;
  const int u = blockIdx.x * kBlock + threadIdx.x;
  if (u >= nNodes) return;
  double a[3] = {0.0, 0.0, 0.0};
  for (int q = off[u]; q < off[u + 1]; ++q) {
    const int e = inc[q] / 8, cq = inc[q] % 8;
    for (int r = 0; r < nd; ++r) {
      const int i = r % n1, j = (r / n1) % n1, k = r / (n1 * n1);
      const double w = corner_weight(cq, (M.xGP[i] + 1.0) / 2.0, (M.xGP[j] + 1.0) / 2.0, (M.xGP[k] + 1.0) / 2.0);
      for (int c = 0; c < 3; ++c) a[c] = a[c] + w * lE[((size_t)e * nd + r) * 3 + c];
    }
  }
  for (int c = 0; c < 3; ++c) lEn[(size_t)u * 3 + c] = a[c];
}

// Dep_x^T: for every particle, the derivative of sum_dof lrho_dof rho_dof
// with respect to its position; lrho_dof = lrbar_e w_i w_j w_k / 8 (A^T).
// rho_dof = sum_p f_p S^alpha, S = 1 - |x_p + shift - x_dof|^2 / R^2 (>= 0).
__global__ __launch_bounds__(kBlock) void k_adj_sf(DevMesh M, DevParts P, int64_t n, SFArgs A, const int32_t *srcOff,
                                                   const int32_t *tgt, const uint8_t *tgtShift, const double *lrbar,
                                                   QuadArgs Q, double *lx) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int e = P.elem[i];
  if (e < 0) return;
  const int sp = P.spec[i];
  const double q0 = M.charge[sp];
  if (!(fabs(q0) > 0.0)) return;
  const double f = q0 * M.mpf[sp] * A.w_sf;
  const int n1 = M.N + 1, nd = M.nDof;
  const double x0[3] = {P.x[i], P.y[i], P.z[i]};
  double g[3] = {0.0, 0.0, 0.0};
  for (int q = srcOff[e]; q < srcOff[e + 1]; ++q) {
    const int t = tgt[q];
    const int si = tgtShift[q];
    const double p[3] = {x0[0] + A.sh[si][0], x0[1] + A.sh[si][1], x0[2] + A.sh[si][2]};
    const double lr = lrbar[t] / 8.0;
    const double *xg = M.xGPe + (size_t)t * nd * 3;
    for (int r = 0; r < nd; ++r) {
      const double d0 = p[0] - xg[r * 3], d1 = p[1] - xg[r * 3 + 1], d2 = p[2] - xg[r * 3 + 2];
      const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
      if (r2 > A.r2_sf) continue;
      const double S = 1. - A.r2_sf_inv * r2;
      double Sa1 = 1.0;   // S^(alpha-1)
      for (int ex = 2; ex <= A.alpha; ++ex) Sa1 = Sa1 * S;
      const int ii = r % n1, jj = (r / n1) % n1, kk = r / (n1 * n1);
      const double c = lr * Q.wGP[ii] * Q.wGP[jj] * Q.wGP[kk] * f * A.alpha * Sa1 * (-2.0 * A.r2_sf_inv);
      g[0] += c * d0; g[1] += c * d1; g[2] += c * d2;
    }
  }
  const int id = P.id[i];
  lx[(size_t)id * 3 + 0] += g[0];
  lx[(size_t)id * 3 + 1] += g[1];
  lx[(size_t)id * 3 + 2] += g[2];
}

// forward-mode dual number for the Boris Jacobians
struct Dual {
  double v, d;
};
__device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ inline Dual dconst(double c) { return {c, 0.0}; }
__device__ inline Dual dsqrt(Dual a) {
  const double s = sqrt(a.v);
  return {s, a.d / (2.0 * s)};
}
__device__ inline Dual dtan(Dual a) {
  const double t = tan(a.v);
  return {t, a.d * (1.0 + t * t)};
}

// velocity update of boris_push with dual numbers (E and v may carry tangents;
// the external B does not)
__device__ inline void boris_vel_dual(double q, double m, const Dual E[3], const double B[3], double dt, Dual v[3]) {
  const double c2i = c2_inv();
  const Dual c_1 = dconst((q * dt) / (m * 2.));
  const Dual gamma = dconst(1.) / dsqrt(dconst(1) - ((v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) * dconst(c2i)));
  const Dual vmo[3] = {v[0] * gamma, v[1] * gamma, v[2] * gamma};
  const Dual vm[3] = {vmo[0] + c_1 * E[0], vmo[1] + c_1 * E[1], vmo[2] + c_1 * E[2]};
  const Dual gm = dsqrt(dconst(1) + (vm[0] * vm[0] + vm[1] * vm[1] + vm[2] * vm[2]) * dconst(c2i));
  const double Bn = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
  double ub[3] = {0.0, 0.0, 0.0};
  double invL = Bn;
  if (fabs(invL) > 0.0) { invL = 1. / invL; ub[0] = B[0] * invL; ub[1] = B[1] * invL; ub[2] = B[2] * invL; }
  const Dual tf = dtan(c_1 / gm * dconst(Bn));
  const Dual t[3] = {tf * dconst(ub[0]), tf * dconst(ub[1]), tf * dconst(ub[2])};
  const Dual vp[3] = {vm[0] + (vm[1] * t[2] - vm[2] * t[1]), vm[1] + (vm[2] * t[0] - vm[0] * t[2]),
                      vm[2] + (vm[0] * t[1] - vm[1] * t[0])};
  const Dual f2 = dconst(2.0) / (dconst(1.) + (t[0] * t[0] + t[1] * t[1] + t[2] * t[2]));
  const Dual cr[3] = {vp[1] * t[2] - vp[2] * t[1], vp[2] * t[0] - vp[0] * t[2], vp[0] * t[1] - vp[1] * t[0]};
  const Dual vpl[3] = {vm[0] + f2 * cr[0], vm[1] + f2 * cr[1], vm[2] + f2 * cr[2]};
  const Dual vn1[3] = {vpl[0] + c_1 * E[0], vpl[1] + c_1 * E[1], vpl[2] + c_1 * E[2]};
  const Dual den = dsqrt(dconst(1) + (vn1[0] * vn1[0] + vn1[1] * vn1[1] + vn1[2] * vn1[2]) * dconst(c2i));
  v[0] = vn1[0] / den;
  v[1] = vn1[1] / den;
  v[2] = vn1[2] / den;
}

// Lagrange basis on the Gauss nodes and its derivative (product form)
template <int NP>
__device__ inline void lagrange_d(double x, const double *xg, double *L, double *dL) {
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    double l = 1.0, dl = 0.0;
#pragma unroll
    for (int m = 0; m < NP; ++m) {
      if (m == i) continue;
      const double inv = 1.0 / (xg[i] - xg[m]);
      dl = dl * ((x - xg[m]) * inv) + l * inv;
      l = l * ((x - xg[m]) * inv);
    }
    L[i] = l;
    dL[i] = dl;
  }
}

// Reverse of one step for every particle of the stored step-n state:
// lv' = lv + dt lx ;  (lv, lF) = J_B^T lv' ;  lx += (dE/dx)^T lF.
// Writes lF and xi (sorted order) for the field scatter.
template <int N>
__global__ __launch_bounds__(kBlock) void k_adj_push(DevMesh M, DevParts P, int64_t n, double dt, double *lx, double *lv,
                                                     double *lF, double *xiOut) {
  constexpr int n1 = N + 1, nd = n1 * n1 * n1;
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int e = P.elem[i];
  const int id = P.id[i];
  const int sp = P.spec[i];
  const double q = M.charge[sp], m = M.mass[sp];
  const double x[3] = {P.x[i], P.y[i], P.z[i]};
  double lvp[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) lvp[c] = lv[(size_t)id * 3 + c] + dt * lx[(size_t)id * 3 + c];
  double xi[3] = {0, 0, 0};
  double lFo[3] = {0, 0, 0};
  double lvo[3] = {lvp[0], lvp[1], lvp[2]};
  if (fabs(q) > 0.0) {
    // the field of the forward step at this particle
    ref_newton_start<N>(M, e, x, xi);
    ref_elem_newton(M, e, x, xi, true);
    double L[3][n1], dL[3][n1];
#pragma unroll
    for (int d = 0; d < 3; ++d) lagrange_d<n1>(xi[d], M.xGP, L[d], dL[d]);
    const double *Ue = M.E + (size_t)e * nd * 3;
    double U[3] = {0, 0, 0}, dU[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // dU[c][d] = dE_c/dxi_d
    for (int k = 0; k < n1; ++k)
      for (int j = 0; j < n1; ++j)
        for (int ii = 0; ii < n1; ++ii) {
          const double *u = Ue + ((k * n1 + j) * n1 + ii) * 3;
          const double w = L[0][ii] * L[1][j] * L[2][k];
          const double w0 = dL[0][ii] * L[1][j] * L[2][k], w1 = L[0][ii] * dL[1][j] * L[2][k],
                       w2 = L[0][ii] * L[1][j] * dL[2][k];
          for (int c = 0; c < 3; ++c) {
            U[c] += u[c] * w;
            dU[c][0] += u[c] * w0; dU[c][1] += u[c] * w1; dU[c][2] += u[c] * w2;
          }
        }
    const double Bx[3] = {M.ext[3], M.ext[4], M.ext[5]};
    const double F[3] = {M.ext[0] + U[0], M.ext[1] + U[1], M.ext[2] + U[2]};
    const double v0[3] = {P.vx[i], P.vy[i], P.vz[i]};
    // columns of the Jacobian of v' with respect to v (0..2) and E (3..5)
    for (int c = 0; c < 3; ++c) lvo[c] = 0.0;
    for (int s = 0; s < 6; ++s) {
      Dual E[3], v[3];
      for (int c = 0; c < 3; ++c) {
        E[c] = {F[c], s == 3 + c ? 1.0 : 0.0};
        v[c] = {v0[c], s == c ? 1.0 : 0.0};
      }
      boris_vel_dual(q, m, E, Bx, dt, v);
      const double dot = v[0].d * lvp[0] + v[1].d * lvp[1] + v[2].d * lvp[2];
      if (s < 3) lvo[s] = dot;
      else lFo[s - 3] = dot;
    }
    // dxi/dx = J^{-1}, J[a][b] = dx_a/dxi_b of the element map at xi
    double Lc[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d) lagrange<2>(xi[d], M.xiCL, M.wBaryCL, Lc[d]);
    const double *dX = M.dXCL + (size_t)e * 72;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int k = 0; k < 2; ++k)
      for (int j = 0; j < 2; ++j)
        for (int ii = 0; ii < 2; ++ii) {
          const double b2 = Lc[0][ii] * Lc[1][j] * Lc[2][k];
          const double *dp = dX + ((k * 2 + j) * 2 + ii) * 9;
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) J[a][b] += dp[a * 3 + b] * b2;
        }
    const double det = get_det(J), sd = 1.0 / det;
    double Ji[3][3];
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * sd;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * sd;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * sd;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * sd;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * sd;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * sd;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * sd;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * sd;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * sd;
    // lx_a += sum_c lF_c dE_c/dx_a = sum_c lF_c sum_d dU[c][d] Ji[d][a]
    for (int a = 0; a < 3; ++a) {
      double acc = 0.0;
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) acc += lFo[c] * dU[c][d] * Ji[d][a];
      lx[(size_t)id * 3 + a] += acc;
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    lv[(size_t)id * 3 + c] = lvo[c];
    lF[i * 3 + c] = lFo[c];
    xiOut[i * 3 + c] = xi[c];
  }
}

// lE^n_dof = sum over the element's particles of L_ijk(xi_p) lF_p: one wave
// per element, lanes over the DOFs, particles in segment order
template <int N>
__global__ __launch_bounds__(kBlock) void k_adj_scatter(DevMesh M, const int32_t *start, const double *lF,
                                                        const double *xi, double *lE) {
  constexpr int n1 = N + 1, nd = n1 * n1 * n1;
  const int w = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = lane_id();
  if (w >= M.nElems) return;
  for (int r = lane; r < nd; r += kWave) {
    const int i = r % n1, j = (r / n1) % n1, k = r / (n1 * n1);
    double a[3] = {0, 0, 0};
    for (int p = start[w]; p < start[w + 1]; ++p) {
      double L0[n1], L1[n1], L2[n1];
      lagrange<n1>(xi[p * 3 + 0], M.xGP, M.wBaryGP, L0);
      lagrange<n1>(xi[p * 3 + 1], M.xGP, M.wBaryGP, L1);
      lagrange<n1>(xi[p * 3 + 2], M.xGP, M.wBaryGP, L2);
      const double wgt = L0[i] * L1[j] * L2[k];
      for (int c = 0; c < 3; ++c) a[c] += wgt * lF[p * 3 + c];
    }
    for (int c = 0; c < 3; ++c) lE[((size_t)w * nd + r) * 3 + c] = a[c];
  }
}

__global__ void k_copy_f64(const double *a, double *b, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// ---- host side ------------------------------------------------------------------
int alloc_state(pic_handle *h, StepState &S, bool withE) {
  int rc = 0;
  rc |= dalloc(h, &S.x, size_t(h->cap) * 6);
  rc |= dalloc(h, &S.elem, h->cap);
  rc |= dalloc(h, &S.id, h->cap);
  rc |= dalloc(h, &S.spec, h->cap);
  rc |= dalloc(h, &S.start, h->hm.nElems + 1);
  if (withE) rc |= dalloc(h, &S.E, size_t(h->hm.nElems) * h->hm.nDof * 3);
  return rc;
}

int save_state(pic_handle *h, StepState &S) {
  const size_t n = size_t(h->n), cap = size_t(h->cap);
  double *src[6] = {h->P.x, h->P.y, h->P.z, h->P.vx, h->P.vy, h->P.vz};
  for (int c = 0; c < 6; ++c)
    HIPCHK(hipMemcpyAsync(S.x + c * cap, src[c], n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(S.elem, h->P.elem, n * sizeof(int32_t), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(S.id, h->P.id, n * sizeof(int32_t), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(S.spec, h->P.spec, n, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(S.start, h->start, (h->hm.nElems + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, h->stream));
  S.n = h->n;
  return 0;
}

int restore_state(pic_handle *h, const StepState &S) {
  const size_t n = size_t(S.n), cap = size_t(h->cap);
  double *dst[6] = {h->P.x, h->P.y, h->P.z, h->P.vx, h->P.vy, h->P.vz};
  for (int c = 0; c < 6; ++c)
    HIPCHK(hipMemcpyAsync(dst[c], S.x + c * cap, n * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->P.elem, S.elem, n * sizeof(int32_t), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->P.id, S.id, n * sizeof(int32_t), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->P.spec, S.spec, n, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->start, S.start, (h->hm.nElems + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, h->stream));
  h->recFresh = false;
  hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(int64_t(n))), dim3(kBlock), 0, h->stream, h->P.isNew, int64_t(n), uint8_t(0));
  h->n = S.n;
  return 0;
}

// a DevParts view of a stored state
DevParts view_of(pic_handle *h, const StepState &S) {
  DevParts v{};
  const size_t cap = size_t(h->cap);
  v.x = S.x; v.y = S.x + cap; v.z = S.x + 2 * cap;
  v.vx = S.x + 3 * cap; v.vy = S.x + 4 * cap; v.vz = S.x + 5 * cap;
  v.elem = S.elem; v.id = S.id; v.spec = S.spec;
  return v;
}

QuadArgs quad_of(pic_handle *h) {
  QuadArgs Q{};
  for (int i = 0; i <= h->hm.N; ++i) Q.wGP[i] = h->hm.wGP[i];
  return Q;
}

// E = P G A rho of the current deposition (h->PS) into h->E
int field_from_source(pic_handle *h) {
  const int ne = h->hm.nElems, N = h->hm.N;
  QuadArgs Q = quad_of(h);
  LAUNCH_N(k_elem_mean, N, dim3(grid_for(ne)), ne, h->PS, Q, h->adj.rbar);
  const int64_t rows = int64_t(h->hm.nUnique) * 3;
  hipLaunchKernelGGL(k_matvec, dim3(grid_for(rows * kWave)), dim3(kBlock), 0, h->stream, rows, int64_t(ne), h->adj.G,
                     h->adj.rbar, h->adj.En);
  LAUNCH_N(k_node_to_dof, N, dim3(grid_for(int64_t(ne) * h->hm.nDof)), h->M, h->adj.elemNode, h->adj.En, h->E);
  HIPCHK(hipGetLastError());
  return 0;
}

// one self-consistent step: deposit, field, push/track, re-sort
int sc_step(pic_handle *h, double dt) {
  if (h->par.tracking != PIC_TRACK_TRIA) return fail(PIC_EINVAL, "self-consistent steps: TriaTracking only");
  if (int rc = run_deposition(h, false)) return rc;
  if (int rc = field_from_source(h)) return rc;
  const int64_t n = h->n;
  const int32_t *nDev = h->start + h->hm.nElems;
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kBlock), 0, h->stream, h->mv.count, int64_t(1), 0);
  LAUNCH_N(k_push_inside, h->hm.N, dim3(grid_for(std::max<int64_t>(n, 1))), h->M, h->P, n, dt, h->newElem, h->mv, nDev);
  h->recFresh = false;
  hipLaunchKernelGGL(k_track_movers, dim3(1024), dim3(kBlock), 0, h->stream, h->M, h->P, h->newElem, h->mv,
                     h->lostCount, h->insideById, h->stateById);
  HIPCHK(hipGetLastError());
  if (int rc = enqueue_sort(h, n, false)) return rc;
  const int64_t n0 = h->n;
  if (int rc = read_count(h)) return rc;
  if (h->n != n0) return fail(PIC_ETRACK, "self-consistent steps: particles left the domain (periodic meshes only)");
  return 0;
}

// loss of the current state: field energy of the field of its deposition
int sc_loss(pic_handle *h, double *loss) {
  if (int rc = run_deposition(h, false)) return rc;
  if (int rc = field_from_source(h)) return rc;
  double w = 0.0;
  if (int rc = elem_integral(h, 1, &w)) return rc;
  *loss = w * 8.8541878176e-12 * 0.5;
  return 0;
}

// lrbar = G^T P^T lE, then lx += Dep_x^T(A^T lrbar) at the state in h->P / view
int adj_field_to_x(pic_handle *h, const DevParts &V, int64_t n) {
  const int N = h->hm.N;
  LAUNCH_N(k_adj_node, N, dim3(grid_for(h->hm.nUnique)), h->M, h->hm.nUnique, h->adj.nodeOff, h->adj.nodeInc,
           h->adj.lE, h->adj.lEn);
  const int64_t rows = h->hm.nElems, cols = int64_t(h->hm.nUnique) * 3;
  hipLaunchKernelGGL(k_matvec, dim3(grid_for(rows * kWave)), dim3(kBlock), 0, h->stream, rows, cols, h->adj.GT,
                     h->adj.lEn, h->adj.lrbar);
  QuadArgs Q = quad_of(h);
  hipLaunchKernelGGL(k_adj_sf, dim3(grid_for(std::max<int64_t>(n, 1))), dim3(kBlock), 0, h->stream, h->M, V, n, h->sf,
                     h->adj.srcOff, h->adj.tgt, h->adj.tgtShift, h->adj.lrbar, Q, h->adj.lx);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // namespace

extern "C" {

int pic_set_field_operator(pic_handle *h, const double *G) {
  if (!h || !G) return fail(PIC_EINVAL, "pic_set_field_operator: null argument");
  if (h->par.deposition != PIC_DEPO_SF) return fail(PIC_EINVAL, "pic_set_field_operator: shape_function deposition only");
  if (h->par.tracking != PIC_TRACK_TRIA) return fail(PIC_EINVAL, "pic_set_field_operator: TriaTracking only");
  if (h->par.precision != PIC_PREC_FP64) return fail(PIC_EINVAL, "pic_set_field_operator: FP64 only");
  if (set_device(h)) return PIC_EHIP;
  const int ne = h->hm.nElems, nu = h->hm.nUnique;
  const size_t rows = size_t(nu) * 3;
  Adjoint &a = h->adj;
  if (!a.G) {
    int rc = 0;
    rc |= dalloc(h, &a.G, rows * ne);
    rc |= dalloc(h, &a.GT, rows * ne);
    rc |= dalloc(h, &a.rbar, ne);
    rc |= dalloc(h, &a.lrbar, ne);
    rc |= dalloc(h, &a.En, rows);
    rc |= dalloc(h, &a.lEn, rows);
    rc |= dalloc(h, &a.lE, size_t(ne) * h->hm.nDof * 3);
    rc |= upload(h, &a.elemNode, h->hm.elemNode);
    rc |= upload(h, &a.nodeOff, h->hm.nodeOff);
    rc |= upload(h, &a.nodeInc, h->hm.nodeInc);
    rc |= upload(h, &a.srcOff, h->hm.sfSrcOff);
    rc |= upload(h, &a.tgt, h->hm.sfTgt);
    rc |= upload(h, &a.tgtShift, h->hm.sfTgtShift);
    if (rc) return PIC_EHIP;
  }
  HIPCHK(hipMemcpy(a.G, G, rows * ne * sizeof(double), hipMemcpyHostToDevice));
  std::vector<double> T(rows * ne);
  for (size_t r = 0; r < rows; ++r)
    for (int e = 0; e < ne; ++e) T[size_t(e) * rows + r] = G[r * ne + e];
  HIPCHK(hipMemcpy(a.GT, T.data(), rows * ne * sizeof(double), hipMemcpyHostToDevice));
  return PIC_OK;
}

int pic_sc_time_step(pic_handle *h, double dt, int32_t nsteps, double *loss) {
  if (!h) return fail(PIC_EINVAL, "pic_sc_time_step: null handle");
  if (!h->adj.G) return fail(PIC_ESTATE, "pic_sc_time_step: call pic_set_field_operator first");
  if (int rc = require_particles(h)) return rc;
  if (set_device(h)) return PIC_EHIP;
  for (int it = 0; it < nsteps; ++it)
    if (int rc = sc_step(h, dt)) return rc;
  if (loss)
    if (int rc = sc_loss(h, loss)) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  h->stagedValid = false;
  return PIC_OK;
}

int pic_adjoint(pic_handle *h, double dt, int32_t K, int32_t every, double *loss, double *dLdx0, double *dLdv0) {
  if (!h || !loss) return fail(PIC_EINVAL, "pic_adjoint: null argument");
  if (!h->adj.G) return fail(PIC_ESTATE, "pic_adjoint: call pic_set_field_operator first");
  if (K < 1 || every < 1) return fail(PIC_EINVAL, "pic_adjoint: K and the checkpoint interval must be >= 1");
  if (int rc = require_particles(h)) return rc;
  if (set_device(h)) return PIC_EHIP;
  Adjoint &a = h->adj;
  const int nck = (K + every - 1) / every;
  // buffers: checkpoints, one segment of step states, adjoint vectors
  if (int(a.ck.size()) < nck) {
    a.ck.resize(nck);
    for (int c = 0; c < nck; ++c)
      if (!a.ck[c].x && alloc_state(h, a.ck[c], false)) return PIC_EHIP;
  }
  if (int(a.seg.size()) < every) {
    a.seg.resize(every);
    for (int c = 0; c < every; ++c)
      if (!a.seg[c].x && alloc_state(h, a.seg[c], true)) return PIC_EHIP;
  }
  if (!a.lx) {
    int rc = 0;
    rc |= dalloc(h, &a.lx, size_t(h->cap) * 3);
    rc |= dalloc(h, &a.lv, size_t(h->cap) * 3);
    rc |= dalloc(h, &a.lF, size_t(h->cap) * 3);
    rc |= dalloc(h, &a.xi, size_t(h->cap) * 3);
    if (rc) return PIC_EHIP;
  }
  const int64_t n = h->n;
  // the steps start from running particles (no IsNewPart half step)
  hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(std::max<int64_t>(n, 1))), dim3(kBlock), 0, h->stream, h->P.isNew, n,
                     uint8_t(0));
  // forward: checkpoints every `every` steps
  for (int it = 0; it < K; ++it) {
    if (it % every == 0)
      if (int rc = save_state(h, a.ck[it / every])) return rc;
    if (int rc = sc_step(h, dt)) return rc;
  }
  if (int rc = sc_loss(h, loss)) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  if (!dLdx0 && !dLdv0) return PIC_OK;
  // reverse: gradient of the field energy at step K
  const int N = h->hm.N;
  const int ne = h->hm.nElems;
  QuadArgs Q = quad_of(h);
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(int64_t(h->cap) * 3)), dim3(kBlock), 0, h->stream, a.lx, int64_t(h->cap) * 3);
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(int64_t(h->cap) * 3)), dim3(kBlock), 0, h->stream, a.lv, int64_t(h->cap) * 3);
  LAUNCH_N(k_adj_energy, N, dim3(grid_for(int64_t(ne) * h->hm.nDof)), ne, h->E, h->sJ, Q, a.lE);
  if (int rc = adj_field_to_x(h, h->P, n)) return rc;
  // segments, last first: recompute the step states from the checkpoint, then reverse them
  for (int c = nck - 1; c >= 0; --c) {
    const int s0 = c * every, s1 = std::min(K, s0 + every);
    if (int rc = restore_state(h, a.ck[c])) return rc;
    for (int it = s0; it < s1; ++it) {
      StepState &S = a.seg[it - s0];
      if (int rc = save_state(h, S)) return rc;
      if (int rc = sc_step(h, dt)) return rc;
      // sc_step computed the field of step `it` into h->E before pushing
      hipLaunchKernelGGL(k_copy_f64, dim3(grid_for(int64_t(ne) * h->hm.nDof * 3)), dim3(kBlock), 0, h->stream, h->E, S.E,
                         int64_t(ne) * h->hm.nDof * 3);
    }
    for (int it = s1 - 1; it >= s0; --it) {
      StepState &S = a.seg[it - s0];
      DevParts V = view_of(h, S);
      // the forward step read E^it: point the mesh view at the stored field
      DevMesh Mstep = h->M;
      Mstep.E = S.E;
      LAUNCH_N(k_adj_push, N, dim3(grid_for(S.n)), Mstep, V, S.n, dt, a.lx, a.lv, a.lF, a.xi);
      LAUNCH_N(k_adj_scatter, N, dim3(grid_for(int64_t(ne) * kWave)), Mstep, S.start, a.lF, a.xi, a.lE);
      HIPCHK(hipGetLastError());
      if (int rc = adj_field_to_x(h, V, S.n)) return rc;
    }
  }
  // the stage ends in the state after K steps
  if (int rc = restore_state(h, a.ck[nck - 1])) return rc;
  for (int it = (nck - 1) * every; it < K; ++it)
    if (int rc = sc_step(h, dt)) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  const int64_t nt = h->nTotal;
  if (dLdx0) HIPCHK(hipMemcpy(dLdx0, a.lx, nt * 3 * sizeof(double), hipMemcpyDeviceToHost));
  if (dLdv0) HIPCHK(hipMemcpy(dLdv0, a.lv, nt * 3 * sizeof(double), hipMemcpyDeviceToHost));
  h->stagedValid = false;
  return PIC_OK;
}

}  // extern "C"
