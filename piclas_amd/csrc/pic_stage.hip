// MI355X (gfx950) particle stage: kernels and the C-ABI of include/pic_stage.h.
//
// Step structure (TimeStepPoissonByBorisLeapfrog without the field solve,
// reference src/timedisc/timedisc_TimeStepPoissonByBorisLeapfrog.f90:32-286):
//   1. deposition from the element-sorted particles (gather, no atomics)
//   2. one fused kernel: LastPartPos <- x, interpolate E, Boris push, TriaTracking
//   3. element re-sort: particles that stay keep their order, arrivals are
//      appended in caller-id order, so every run is bit-for-bit repeatable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "pic_device.h"
#include "pic_refmap.h"
#include "stage_internal.h"

#pragma clang fp contract(off)

using namespace pic;

namespace {

thread_local std::string g_err;
int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
#define HIPCHK(x)                                                                                        \
  do {                                                                                                   \
    hipError_t e_ = (x);                                                                                 \
    if (e_ != hipSuccess) return fail(PIC_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));         \
  } while (0)

// hipMemcpy that treats an empty copy as done (no particles is a valid state)
inline hipError_t copy_bytes(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  return bytes == 0 ? hipSuccess : hipMemcpy(dst, src, bytes, kind);
}

constexpr int kBlock = 256;
constexpr int kWave = 64;

inline unsigned grid_for(int64_t n, int block = kBlock) { return unsigned((n + block - 1) / block); }

// Workgroups are handed to the 8 XCDs round-robin; this maps hardware block
// b to a logical block so each XCD works on a contiguous range (its own L2
// then holds the neighbourhood of the elements it deposits).  The grid is a
// multiple of 8 blocks, which makes the map a permutation of [0, grid).
__device__ inline int xcd_block(int b, int grid) {
  const int per = grid / 8;
  return (b % 8) * per + b / 8;
}
inline unsigned xcd_grid(int64_t nblk) { return unsigned((nblk + 7) / 8 * 8); }

__device__ inline int lane_id() { return __lane_id(); }

// ------------------------------------------------------------------ kernels

// Interpolation only (InterpolateFieldToParticle): F[i] = ext + E(x_i)
template <int N, bool MIXED = false, bool AFFINE = false>
__global__ __launch_bounds__(kBlock) void k_interp(DevMesh M, DevParts P, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int e = P.elem[i];
  double F[6] = {0, 0, 0, 0, 0, 0};
  const int sp = P.spec[i];
  if (e >= 0 && fabs(M.charge[sp]) > 0.0) {
    const double x[3] = {P.x[i], P.y[i], P.z[i]};
    field_at_particle<N, MIXED, AFFINE>(M, e, x, F);
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) P.F[i * 6 + c] = F[c];
}

// Push only: LastPartPos <- PartState, Boris push with the stored field
__global__ __launch_bounds__(kBlock) void k_push(DevMesh M, DevParts P, int64_t n, double dt) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int e = P.elem[i];
  double x[3] = {P.x[i], P.y[i], P.z[i]};
  P.lx[i] = x[0]; P.ly[i] = x[1]; P.lz[i] = x[2];
  P.lastElem[i] = e;
  if (e < 0) return;
  double v[3] = {P.vx[i], P.vy[i], P.vz[i]};
  double F[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) F[c] = P.F[i * 6 + c];
  const int sp = P.spec[i];
  boris_push(M.charge[sp], M.mass[sp], P.isNew[i] != 0, F, dt, x, v);
  P.isNew[i] = 0;
  P.x[i] = x[0]; P.y[i] = x[1]; P.z[i] = x[2];
  P.vx[i] = v[0]; P.vy[i] = v[1]; P.vz[i] = v[2];
}

// Tracking only (PerformTracking): from lastElem / LastPartPos to the new owner
__global__ __launch_bounds__(kBlock) void k_track(DevMesh M, DevParts P, int64_t n, int32_t *newElem,
                                                  unsigned long long *lostCount, uint8_t *insideById, double *stateById) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int e0 = P.lastElem[i];
  if (e0 < 0) { newElem[i] = -1; return; }
  double x[3] = {P.x[i], P.y[i], P.z[i]};
  double last[3] = {P.lx[i], P.ly[i], P.lz[i]};
  bool lost;
  int steps;
  const int e = tria_track(M, e0, x, last, lost, steps);
  P.x[i] = x[0]; P.y[i] = x[1]; P.z[i] = x[2];
  P.lx[i] = last[0]; P.ly[i] = last[1]; P.lz[i] = last[2];
  newElem[i] = e;
  P.elem[i] = e;
  if (e < 0) {
    const int id = P.id[i];
    insideById[id] = 0;
    double *st = stateById + (size_t)id * 6;
    st[0] = x[0]; st[1] = x[1]; st[2] = x[2]; st[3] = P.vx[i]; st[4] = P.vy[i]; st[5] = P.vz[i];
    if (lost) atomicAdd(lostCount, 1ull);
  }
}

// Fused step body: interpolate + push + track for one particle
template <int N>
__global__ __launch_bounds__(kBlock) void k_push_track(DevMesh M, DevParts P, int64_t n, double dt, int32_t *newElem,
                                                       unsigned long long *lostCount, uint8_t *insideById,
                                                       double *stateById) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int e0 = P.elem[i];
  if (e0 < 0) { newElem[i] = -1; return; }
  double x[3] = {P.x[i], P.y[i], P.z[i]};
  double v[3] = {P.vx[i], P.vy[i], P.vz[i]};
  const double last0[3] = {x[0], x[1], x[2]};
  const int sp = P.spec[i];
  double F[6] = {0, 0, 0, 0, 0, 0};
  if (fabs(M.charge[sp]) > 0.0) field_at_particle<N>(M, e0, x, F);
  boris_push(M.charge[sp], M.mass[sp], P.isNew[i] != 0, F, dt, x, v);
  double last[3] = {last0[0], last0[1], last0[2]};
  bool lost;
  int steps;
  const int e = tria_track(M, e0, x, last, lost, steps);
  P.x[i] = x[0]; P.y[i] = x[1]; P.z[i] = x[2];
  P.vx[i] = v[0]; P.vy[i] = v[1]; P.vz[i] = v[2];
  P.isNew[i] = 0;
  newElem[i] = e;
  if (e < 0) {
    const int id = P.id[i];
    insideById[id] = 0;
    double *st = stateById + (size_t)id * 6;
    st[0] = x[0]; st[1] = x[1]; st[2] = x[2]; st[3] = v[0]; st[4] = v[1]; st[5] = v[2];
    if (lost) atomicAdd(lostCount, 1ull);
  }
}

// Split form of the fused step for throughput.  Kernel A does, per particle,
// interpolate + push and the first iteration of SingleParticleTriaTracking3D
// (ParticleInsideQuad3D in the last element); particles still inside are
// done.  The others (about one in a hundred) are appended to a list and
// kernel B runs the full face walk for them from the same state, so the
// result is the one of the single loop.  Splitting keeps kernel A's register
// footprint low enough for several waves per SIMD.
struct Movers {
  int32_t *count;   // [1]
  int32_t *idx;     // [cap]
  double *last;     // [cap][3] LastPartPos
  uint8_t *flag;    // [cap] RefMapping: bit 0 boundary candidate, bit 1 still inside its element
  int32_t *slot2;   // [cap] RefMapping: the slots that k_rm_bc_filter passes on (count at count[1])
};

template <int N, bool MIXED = false, bool AFFINE = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE_PUSH))) void k_push_inside(DevMesh M, DevParts P, int64_t n, double dt, int32_t *newElem,
                                                        Movers mv, const int32_t *nDev) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (nDev) n = *nDev;   // particle count on the device (graph replay)
  bool mover = false;
  double last[3] = {0, 0, 0};
  // whether all particles of the wave sit in one element (sorted order: most
  // waves); its field is then read by scalar loads
  const int eAll = i < n ? P.elem[i] : -1;
  const int eu = __builtin_amdgcn_readfirstlane(eAll);
  const bool uni = eu >= 0 && __ballot(i < n && eAll != eu) == 0;
  if (i < n) {
    const int e0 = eAll;
    if (e0 < 0) {
      newElem[i] = -1;
    } else {
      double x[3] = {P.x[i], P.y[i], P.z[i]};
      double v[3] = {P.vx[i], P.vy[i], P.vz[i]};
      last[0] = x[0]; last[1] = x[1]; last[2] = x[2];
      const int sp = P.spec[i];
      double F[6] = {0, 0, 0, 0, 0, 0};
      if (fabs(M.charge[sp]) > 0.0) {
        if (uni) field_at_particle<N, MIXED, AFFINE, true>(M, eu, x, F);
        else field_at_particle<N, MIXED, AFFINE>(M, e0, x, F);
      }
      boris_push(M.charge[sp], M.mass[sp], P.isNew[i] != 0, F, dt, x, v);
      P.x[i] = x[0]; P.y[i] = x[1]; P.z[i] = x[2];
      P.vx[i] = v[0]; P.vy[i] = v[1]; P.vz[i] = v[2];
      P.isNew[i] = 0;
      double det[6][2];
      bool in;
      if (AFFINE) {
        // On a parallelepiped (corner map affine to 1e-14) ParticleInsideQuad3D
        // decides as |xi| <= 1 does, xi = A^-1 (x - c), except within rounding
        // of a face: the reference's side test runs only for particles with
        // max |xi| within 1e-10 of 1, the others are decided by xi.
        double xi[3];
        if (uni) affine_xi<true>(M, eu, x, xi);
        else affine_xi(M, e0, x, xi);
        const double mx = max_abs3(xi);
        if (mx < 1.0 - 1e-10) in = true;
        else if (mx > 1.0 + 1e-10) in = false;
        else in = uni ? inside_quad<true>(M, eu, x, det) : inside_quad(M, e0, x, det);
      } else {
        in = uni ? inside_quad<true>(M, eu, x, det) : inside_quad(M, e0, x, det);
      }
      if (in) newElem[i] = e0;
      else mover = true;
    }
  }
  // one atomic per wave for the mover slots
  const unsigned long long m = __ballot(mover);
  if (m == 0) return;
  const int lane = lane_id();
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(mv.count, __popcll(m));
  base = __shfl(base, leader, kWave);
  if (mover) {
    const int slot = base + __popcll(m & ((1ull << lane) - 1ull));
    mv.idx[slot] = int32_t(i);
    mv.last[(size_t)slot * 3 + 0] = last[0];
    mv.last[(size_t)slot * 3 + 1] = last[1];
    mv.last[(size_t)slot * 3 + 2] = last[2];
  }
}

__global__ __launch_bounds__(kBlock) void k_track_movers(DevMesh M, DevParts P, int32_t *newElem, Movers mv,
                                                         unsigned long long *lostCount, uint8_t *insideById,
                                                         double *stateById) {
  const int cnt = *mv.count;
  for (int k = blockIdx.x * kBlock + threadIdx.x; k < cnt; k += gridDim.x * kBlock) {
    const int i = mv.idx[k];
    double x[3] = {P.x[i], P.y[i], P.z[i]};
    double last[3] = {mv.last[(size_t)k * 3], mv.last[(size_t)k * 3 + 1], mv.last[(size_t)k * 3 + 2]};
    bool lost;
    int steps;
    const int e = tria_track(M, P.elem[i], x, last, lost, steps);
    P.x[i] = x[0]; P.y[i] = x[1]; P.z[i] = x[2];
    newElem[i] = e;
    if (e < 0) {
      const int id = P.id[i];
      insideById[id] = 0;
      double *st = stateById + (size_t)id * 6;
      st[0] = x[0]; st[1] = x[1]; st[2] = x[2]; st[3] = P.vx[i]; st[4] = P.vy[i]; st[5] = P.vz[i];
      if (lost) atomicAdd(lostCount, 1ull);
    }
  }
}

// ---- RefMapping (ParticleRefTracking) kernels -------------------------------
// G = NGeo + 1 geometry points per direction.  The field is evaluated at the
// stored PartPosRef (GetInterpolatedFieldPartPos, pic_interpolation_tools.f90:242-244).
template <int N, bool MIXED = false, bool UNI = false>
__device__ inline void field_at_ref(const DevMesh &M, int e, const double xi[3], double F[6]) {
#pragma unroll
  for (int c = 0; c < 6; ++c) F[c] = M.ext[c];
  double U[3];
  if (MIXED && UNI) eval_field_f32_s<N>(M, e, xi, U);
  else if (MIXED) eval_field_f32<N>(M, e, xi, U);
  else if (UNI) eval_field_s<N>(M, e, xi, U);
  else eval_field<N>(M, e, xi, U);
  F[0] = F[0] + U[0];
  F[1] = F[1] + U[1];
  F[2] = F[2] + U[2];
}

__device__ inline void rm_removed(DevParts &P, int64_t i, const double x[3], const double v[3], int status,
                                  unsigned long long *lostCount, uint8_t *insideById, double *stateById) {
  const int id = P.id[i];
  insideById[id] = 0;
  double *st = stateById + (size_t)id * 6;
  st[0] = x[0]; st[1] = x[1]; st[2] = x[2]; st[3] = v[0]; st[4] = v[1]; st[5] = v[2];
  if (status == kRmNotFound) atomicAdd(lostCount, 1ull);
}

// Interpolation, push and ParticleRefTracking in two kernels (as for TriaTracking): kernel A
// interpolates, pushes and maps the particle in its last element; the few
// that left it go to a list that kernel B searches, from the same state.
template <int N, int G, bool MIXED = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE_RM))) void k_rm_push_inside(DevMesh M, DevParts P, int64_t n, double dt, int32_t *newElem,
                                                           Movers mv, const int32_t *nDev) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (nDev) n = *nDev;
  bool mover = false;
  // whether all particles of the wave sit in one element (sorted order: most
  // waves); its map is then read by scalar loads
  const int eAll = i < n ? P.elem[i] : -1;
  const int eu = __builtin_amdgcn_readfirstlane(eAll);
  const bool uni = eu >= 0 && __ballot(i < n && eAll != eu) == 0;
  RmHead head = {true, false};
  double x[3] = {0, 0, 0};
  if (i < n) {
    const int e0 = eAll;
    if (e0 < 0) {
      newElem[i] = -1;
    } else {
      x[0] = P.x[i]; x[1] = P.y[i]; x[2] = P.z[i];
      double v[3] = {P.vx[i], P.vy[i], P.vz[i]};
      double last[3] = {x[0], x[1], x[2]};
      double xi[3] = {P.rx[i], P.ry[i], P.rz[i]};
      const int sp = P.spec[i];
      double F[6] = {0, 0, 0, 0, 0, 0};
      if (fabs(M.charge[sp]) > 0.0) {
        if (uni) field_at_ref<N, MIXED, true>(M, eu, xi, F);
        else field_at_ref<N, MIXED>(M, e0, xi, F);
      }
      boris_push(M.charge[sp], M.mass[sp], P.isNew[i] != 0, F, dt, x, v);
      P.vx[i] = v[0]; P.vy[i] = v[1]; P.vz[i] = v[2];
      P.isNew[i] = 0;
      head = uni ? rm_track_head<N, G, true>(M, eu, x, last, xi) : rm_track_head<N, G>(M, e0, x, last, xi);
      P.rx[i] = xi[0]; P.ry[i] = xi[1]; P.rz[i] = xi[2];
      if (head.inside && !head.cand) {
        newElem[i] = e0;
        P.x[i] = x[0]; P.y[i] = x[1]; P.z[i] = x[2];
      } else {
        mover = true;   // x is stored after LastPartPos is read back below
      }
    }
  }
  const unsigned long long m = __ballot(mover);
  if (m == 0) return;
  const int lane = lane_id();
  const int leader = __ffsll((long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(mv.count, __popcll(m));
  base = __shfl(base, leader, kWave);
  if (mover) {
    const int slot = base + __popcll(m & ((1ull << lane) - 1ull));
    mv.idx[slot] = int32_t(i);
    mv.flag[slot] = uint8_t((head.cand ? 1 : 0) | (head.inside ? 2 : 0));
    // LastPartPos: the position before the push, read back from the arrays
    // (not held in registers across the Newton iteration; after a
    // PeriodicMovement it is not the shifted one, which no later stage reads
    // on CartesianPeriodic meshes)
    mv.last[(size_t)slot * 3 + 0] = P.x[i];
    mv.last[(size_t)slot * 3 + 1] = P.y[i];
    mv.last[(size_t)slot * 3 + 2] = P.z[i];
    P.x[i] = x[0]; P.y[i] = x[1]; P.z[i] = x[2];
  }
}

// Kernel B for RefMapping: the particles that left their element, and those
// of boundary elements whose step reaches a boundary side's range.  The
// boundary test (bc_may_hit: the early exits of the face intersections of
// every boundary side in range) and the candidate search run here; a
// particle whose step may hit a boundary side, or that needs the fallback
// of the search, is left untouched for k_rm_bc_movers, which holds the
// intersection code (its registers and stack would slow the many movers
// for the few).
template <int N, int G>
__global__ __launch_bounds__(kBlock) void k_rm_movers(DevMesh M, DevParts P, int32_t *newElem, Movers mv,
                                                      unsigned long long *lostCount, uint8_t *insideById,
                                                      double *stateById) {
  const int cnt = *mv.count;
  const int lane = lane_id();
  const int stride = gridDim.x * kBlock;
  // wave-uniform trip count: every lane takes part in the ballots
  for (int k0 = blockIdx.x * kBlock + (threadIdx.x & ~(kWave - 1)); k0 < cnt; k0 += stride) {
    const int k = k0 + lane;
    bool defer = false;
    if (k < cnt) {
      const int i = mv.idx[k];
      uint8_t f = mv.flag[k];
      const int e0 = P.elem[i];
      double x[3] = {P.x[i], P.y[i], P.z[i]};
      double last[3] = {mv.last[(size_t)k * 3], mv.last[(size_t)k * 3 + 1], mv.last[(size_t)k * 3 + 2]};
      bool done = false;
      if (f & 1) {
        const double d0 = x[0] - last[0], d1 = x[1] - last[1], d2 = x[2] - last[2];
        if (bc_may_hit(M, e0, x, last, sqrt(d0 * d0 + d1 * d1 + d2 * d2))) {
          defer = true;
        } else if (f & 2) {
          newElem[i] = e0;
          done = true;
        } else {
          f = uint8_t(f & ~1);
          mv.flag[k] = f;
        }
      } else if (f & 2) {   // (not a mover: kept for completeness)
        newElem[i] = e0;
        done = true;
      }
      if (!defer && !done) {
        double xi[3] = {P.rx[i], P.ry[i], P.rz[i]};
        int status;
        const int e = rm_search<N, G, true>(M, e0, x, last, xi, status);
        if (status == kRmDeferred) {
          defer = true;
        } else {
          P.rx[i] = xi[0]; P.ry[i] = xi[1]; P.rz[i] = xi[2];
          newElem[i] = e;
          if (e < 0) {
            const double v[3] = {P.vx[i], P.vy[i], P.vz[i]};
            rm_removed(P, i, x, v, status, lostCount, insideById, stateById);
          }
        }
      }
    }
    const unsigned long long m = __ballot(defer);
    if (m == 0) continue;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(mv.count + 1, __popcll(m));
    base = __shfl(base, leader, kWave);
    if (defer) mv.slot2[base + __popcll(m & ((1ull << lane) - 1ull))] = k;
  }
}

// Kernel C for RefMapping: ParticleBCTracking for the movers whose step may
// hit a boundary side, and the search's fallback, from the state kernel A
// left (the candidate bit holds the outcome of bc_may_hit)
template <int N, int G>
__global__ __launch_bounds__(kBlock) void k_rm_bc_movers(DevMesh M, DevParts P, int32_t *newElem, Movers mv,
                                                         unsigned long long *lostCount, uint8_t *insideById,
                                                         double *stateById) {
  const int cnt = mv.count[1];
  for (int j = blockIdx.x * kBlock + threadIdx.x; j < cnt; j += gridDim.x * kBlock) {
    const int k = mv.slot2[j];
    const int i = mv.idx[k];
    const uint8_t f = mv.flag[k];
    double x[3] = {P.x[i], P.y[i], P.z[i]};
    double last[3] = {mv.last[(size_t)k * 3], mv.last[(size_t)k * 3 + 1], mv.last[(size_t)k * 3 + 2]};
    double xi[3] = {P.rx[i], P.ry[i], P.rz[i]};
    int status;
    const RmHead h = {(f & 2) != 0, (f & 1) != 0};
    const int e = rm_track_tail<N, G, true>(M, P.elem[i], h, x, last, xi, status);
    P.x[i] = x[0]; P.y[i] = x[1]; P.z[i] = x[2];
    P.rx[i] = xi[0]; P.ry[i] = xi[1]; P.rz[i] = xi[2];
    newElem[i] = e;
    if (e < 0) {
      const double v[3] = {P.vx[i], P.vy[i], P.vz[i]};
      rm_removed(P, i, x, v, status, lostCount, insideById, stateById);
    }
  }
}

// staged InterpolateFieldToParticle with RefMapping
template <int N, bool MIXED = false>
__global__ __launch_bounds__(kBlock) void k_rm_interp(DevMesh M, DevParts P, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int e = P.elem[i];
  double F[6] = {0, 0, 0, 0, 0, 0};
  const int sp = P.spec[i];
  if (e >= 0 && fabs(M.charge[sp]) > 0.0) {
    const double xi[3] = {P.rx[i], P.ry[i], P.rz[i]};
    field_at_ref<N, MIXED>(M, e, xi, F);
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) P.F[i * 6 + c] = F[c];
}

// staged PerformTracking with RefMapping
template <int N, int G>
__global__ __launch_bounds__(kBlock) void k_rm_track(DevMesh M, DevParts P, int64_t n, int32_t *newElem,
                                                     unsigned long long *lostCount, uint8_t *insideById,
                                                     double *stateById) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int e0 = P.lastElem[i];
  if (e0 < 0) { newElem[i] = -1; return; }
  double x[3] = {P.x[i], P.y[i], P.z[i]};
  double last[3] = {P.lx[i], P.ly[i], P.lz[i]};
  double xi[3] = {P.rx[i], P.ry[i], P.rz[i]};
  int status;
  const int e = rm_track<N, G>(M, e0, x, last, xi, status);
  P.x[i] = x[0]; P.y[i] = x[1]; P.z[i] = x[2];
  P.lx[i] = last[0]; P.ly[i] = last[1]; P.lz[i] = last[2];
  P.rx[i] = xi[0]; P.ry[i] = xi[1]; P.rz[i] = xi[2];
  newElem[i] = e;
  P.elem[i] = e;
  if (e < 0) {
    const double v[3] = {P.vx[i], P.vy[i], P.vz[i]};
    rm_removed(P, i, x, v, status, lostCount, insideById, stateById);
  }
}

// PartPosRef of every particle in its element (LocateParticleInElement's
// GetPositionInRefElem, particle_localization.f90:47-74)
template <int N, int G>
__global__ __launch_bounds__(kBlock) void k_rm_init(DevMesh M, DevParts P, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const double x[3] = {P.x[i], P.y[i], P.z[i]};
  double xi[3];
  rm_position<N, G>(M, P.elem[i], x, xi);
  P.rx[i] = xi[0]; P.ry[i] = xi[1]; P.rz[i] = xi[2];
}

// SinglePointToElement for particles handed over without an element
template <int N, int G>
__global__ __launch_bounds__(kBlock) void k_rm_locate(DevMesh M, int64_t n, const double *xyz, int32_t *elem) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  if (elem[i] >= 0) return;
  const double x[3] = {xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2]};
  double xi[3];
  elem[i] = rm_locate<N, G>(M, x, xi);
}
__global__ __launch_bounds__(kBlock) void k_tria_locate(DevMesh M, int64_t n, const double *xyz, int32_t *elem) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  if (elem[i] >= 0) return;
  const double x[3] = {xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2]};
  elem[i] = tria_locate(M, x);
}

// ---- sort by element --------------------------------------------------------
// One wave per old element segment: the particles that stay get their rank
// among the element's stayers, ordered by 8 bins of their z across the
// element's Gauss points and within a bin by their old order; arrivals take
// a slot of their new element by an atomic (put in caller-id order by
// k_fix_arrivals).  Ordered by z, the groups of 4 records the shape-function
// gathers stage together reach the same z-planes of a target's DOFs more
// often, and more of their plane products are skipped.
constexpr int kZBins = 8;
__global__ __launch_bounds__(kBlock) void k_seg_rank(int nElems, const int32_t *oldStart, const int32_t *newElem,
                                                     const double *pz, const double *xGPe, int nDof,
                                                     int32_t *stayCnt, int32_t *arrCnt, int32_t *rank) {
  const int w = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = lane_id();
  if (w >= nElems) return;
  const int s0 = oldStart[w], s1 = oldStart[w + 1];
  const double za = xGPe[(size_t)w * nDof * 3 + 2], zb = xGPe[((size_t)w * nDof + nDof - 1) * 3 + 2];
  const double zlo = fmin(za, zb), zinv = zb != za ? double(kZBins) / fabs(zb - za) : 0.0;
  auto bin_of = [&](int i) { return int(fmin(fmax((pz[i] - zlo) * zinv, 0.0), double(kZBins - 1))); };
  const unsigned long long lt = (1ull << lane) - 1ull;
  // pass 1: stayers per bin
  int off[kZBins];
#pragma unroll
  for (int b = 0; b < kZBins; ++b) off[b] = 0;
  for (int base = s0; base < s1; base += kWave) {
    const int i = base + lane;
    const bool st = i < s1 && newElem[i] == w;
    const int bin = st ? bin_of(i) : -1;
#pragma unroll
    for (int b = 0; b < kZBins; ++b) off[b] += __popcll(__ballot(bin == b));
  }
  int stay = 0;
#pragma unroll
  for (int b = 0; b < kZBins; ++b) {
    const int c = off[b];
    off[b] = stay;
    stay += c;
  }
  // pass 2: ranks
  for (int base = s0; base < s1; base += kWave) {
    const int i = base + lane;
    const bool valid = i < s1;
    const int ne = valid ? newElem[i] : -2;
    const bool st = valid && ne == w;
    const int bin = st ? bin_of(i) : -1;
#pragma unroll
    for (int b = 0; b < kZBins; ++b) {
      const unsigned long long m = __ballot(bin == b);
      if (bin == b) rank[i] = off[b] + __popcll(m & lt);
      off[b] += __popcll(m);
    }
    if (valid && !st && ne >= 0) rank[i] = atomicAdd(&arrCnt[ne], 1);
  }
  if (lane == 0) stayCnt[w] = stay;
}

// exclusive scan of stayCnt+arrCnt into newStart[0..nElems], single
// workgroup: tiles of 4096 counts, 4 consecutive per thread (one 16-byte
// load each), wave scans, wave totals through LDS, the next tile's loads in
// flight during the current tile's scan.
__global__ __launch_bounds__(1024) void k_scan(int nElems, const int32_t *stayCnt, const int32_t *arrCnt, int32_t *newStart) {
  __shared__ int32_t wsum[1024 / kWave];
  const int t = threadIdx.x, lane = lane_id(), w = t / kWave;
  auto load4 = [&](int base, int4 &a, int4 &b) {
    const int e = base + 4 * t;
    if (e + 3 < nElems) {
      a = *reinterpret_cast<const int4 *>(stayCnt + e);
      b = *reinterpret_cast<const int4 *>(arrCnt + e);
    } else {
      a = make_int4(e < nElems ? stayCnt[e] : 0, e + 1 < nElems ? stayCnt[e + 1] : 0,
                    e + 2 < nElems ? stayCnt[e + 2] : 0, 0);
      b = make_int4(e < nElems ? arrCnt[e] : 0, e + 1 < nElems ? arrCnt[e + 1] : 0,
                    e + 2 < nElems ? arrCnt[e + 2] : 0, 0);
    }
  };
  int carry = 0;
  int4 sa, sb;
  load4(0, sa, sb);
  for (int base = 0; base < nElems; base += 4096) {
    const int v0 = sa.x + sb.x, v1 = sa.y + sb.y, v2 = sa.z + sb.z, v3 = sa.w + sb.w;
    if (base + 4096 < nElems) load4(base + 4096, sa, sb);
    const int s = v0 + v1 + v2 + v3;
    int x = s;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const int y = __shfl_up(x, off, kWave);
      if (lane >= off) x += y;
    }
    if (lane == kWave - 1) wsum[w] = x;
    __syncthreads();
    int before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 1024 / kWave; ++k) {
      const int ws = wsum[k];
      before += k < w ? ws : 0;
      tot += ws;
    }
    __syncthreads();
    const int e = base + 4 * t;
    int run = carry + before + x - s;
    if (e < nElems) newStart[e] = run;
    run += v0;
    if (e + 1 < nElems) newStart[e + 1] = run;
    run += v1;
    if (e + 2 < nElems) newStart[e + 2] = run;
    run += v2;
    if (e + 3 < nElems) newStart[e + 3] = run;
    carry += tot;
  }
  if (t == 0) newStart[nElems] = carry;
}

__global__ __launch_bounds__(kBlock) void k_dest(int64_t n, const int32_t *oldElemOf, const int32_t *newElem,
                                                 const int32_t *rank, const int32_t *newStart, const int32_t *stayCnt,
                                                 int32_t *perm, const int32_t *nDev) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (nDev) n = *nDev;
  if (i >= n) return;
  const int ne = newElem[i];
  if (ne < 0) return;
  const int oe = oldElemOf[i];
  int64_t d;
  if (ne == oe) d = newStart[ne] + rank[i];
  else d = int64_t(newStart[ne]) + stayCnt[ne] + rank[i];
  perm[d] = int32_t(i);
}

// per element: order the arrival tail by caller id (insertion sort, tails are short)
__global__ __launch_bounds__(kBlock) void k_fix_arrivals(int nElems, const int32_t *newStart, const int32_t *stayCnt,
                                                         const int32_t *arrCnt, const int32_t *id, int32_t *perm) {
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= nElems) return;
  const int na = arrCnt[e];
  if (na < 2) return;
  int32_t *p = perm + newStart[e] + stayCnt[e];
  for (int a = 1; a < na; ++a) {
    const int32_t v = p[a];
    const int32_t key = id[v];
    int b = a - 1;
    while (b >= 0 && id[p[b]] > key) { p[b + 1] = p[b]; --b; }
    p[b + 1] = v;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(const int32_t *total, const int32_t *perm, const T *__restrict__ src,
                                                   T *__restrict__ dst, int width) {
  const int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (j >= int64_t(*total) * width) return;
  const int64_t r = j / width, c = j % width;
  dst[j] = src[int64_t(perm[r]) * width + c];
}

// six FP64 state arrays + id/spec/isNew/elem in one pass
// (and, when rec is given, the shape-function deposition records of the new
// order; fsp holds ChargeIC * MacroParticleFactor * w_sf per species)
struct SpeciesFactor {
  double fsp[8];
};
// Bounding box of the records of one block of 64 (the wave's records j =
// 64 b .. 64 b + 63; lanes past the end or with f = 0, which no gather
// keeps, count as empty): blk[2 b] = (min x, min y, min z, 0), blk[2 b + 1] =
// (max x, max y, max z, 0), +inf / -inf for a block without records.  The
// gathers on general elements skip the blocks whose box misses a target.
// Called by every lane of the wave.
__device__ inline void record_block_box(double4 *blk, int64_t j, bool has, double x, double y, double z) {
  const double inf = __builtin_inf();
  double lo[3] = {has ? x : inf, has ? y : inf, has ? z : inf};
  double hi[3] = {has ? x : -inf, has ? y : -inf, has ? z : -inf};
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      lo[c] = fmin(lo[c], __shfl_xor(lo[c], off, kWave));
      hi[c] = fmax(hi[c], __shfl_xor(hi[c], off, kWave));
    }
  if ((j & (kWave - 1)) == 0) {
    blk[2 * (j >> 6)] = make_double4(lo[0], lo[1], lo[2], 0.0);
    blk[2 * (j >> 6) + 1] = make_double4(hi[0], hi[1], hi[2], 0.0);
  }
}

__global__ __launch_bounds__(kBlock) void k_gather_state(const int32_t *total, const int32_t *perm, DevParts s, DevParts d,
                                                         const int32_t *newElem, SpeciesFactor sf, double4 *rec,
                                                         double4 *recB, double4 *blk) {
  const int64_t j = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t n = *total;
  if ((j & ~int64_t(kWave - 1)) >= n) return;   // the whole wave is past the end
  const bool valid = j < n;
  double x = 0.0, y = 0.0, z = 0.0, f = 0.0;
  if (valid) {
    const int32_t i = perm[j];
    x = s.x[i]; y = s.y[i]; z = s.z[i];
    const double vx = s.vx[i], vy = s.vy[i], vz = s.vz[i];
    const uint8_t sp = s.spec[i];
    d.x[j] = x; d.y[j] = y; d.z[j] = z;
    d.vx[j] = vx; d.vy[j] = vy; d.vz[j] = vz;
    d.id[j] = s.id[i]; d.spec[j] = sp; d.isNew[j] = s.isNew[i];
    d.elem[j] = newElem[i];
    if (s.rx) { d.rx[j] = s.rx[i]; d.ry[j] = s.ry[i]; d.rz[j] = s.rz[i]; }
    if (rec) {
      f = sf.fsp[sp];
      rec[j] = make_double4(x, y, z, f);
      recB[j] = make_double4(vx * f, vy * f, vz * f, f);
    }
  }
  if (blk) record_block_box(blk, j, valid && f != 0.0, x, y, z);
}

// ---- cell_volweight_mean ----------------------------------------------------
// One wave per element segment: trilinear weights of every particle (after
// the ForceMode Newton of GetPositionInRefElem) summed per CGNS corner in a
// fixed lane/butterfly order (DepositionMethod_CVWM, pic_depo_method.f90:465-544).
template <int N>
__global__ __launch_bounds__(kBlock) void k_cvwm_elem(DevMesh M, DevParts P, const int32_t *start, double *elemCorner) {
  const int w = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = lane_id();
  if (w >= M.nElems) return;
  double acc[8][4];
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[c][k] = 0.0;
  const int s0 = start[w], s1 = start[w + 1];
  const double *C = M.corner + (size_t)w * 24;
  for (int i = s0 + lane; i < s1; i += kWave) {
    const int sp = P.spec[i];
    const double q = M.charge[sp];
    if (!(fabs(q) > 0.0)) continue;
    const double Charge = q * M.mpf[sp];
    const double x[3] = {P.x[i], P.y[i], P.z[i]};
    double xi[3];
    ref_newton_start<N>(M, w, x, xi);
    const bool ok = ref_elem_newton(M, w, x, xi, true);
    const double TS[4] = {P.vx[i] * Charge, P.vy[i] * Charge, P.vz[i] * Charge, Charge};
    double wt[8];
    if (ok) {
      const double a1 = 0.5 * (xi[0] + 1.0), a2 = 0.5 * (xi[1] + 1.0), a3 = 0.5 * (xi[2] + 1.0);
      wt[0] = (1 - a1) * (1 - a2) * (1 - a3);
      wt[1] = (a1) * (1 - a2) * (1 - a3);
      wt[2] = (a1) * (a2) * (1 - a3);
      wt[3] = (1 - a1) * (a2) * (1 - a3);
      wt[4] = (1 - a1) * (1 - a2) * (a3);
      wt[5] = (a1) * (1 - a2) * (a3);
      wt[6] = (a1) * (a2) * (a3);
      wt[7] = (1 - a1) * (a2) * (a3);
#pragma unroll
      for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[c][k] = acc[c][k] + (TS[k] * wt[c]);
    } else {
      // inverse-distance weights to the 8 corners (pic_depo_method.f90:512-539)
      const int cgns_t[8] = {0, 1, 3, 2, 4, 5, 7, 6};
      int hit = -1;
      for (int c = 0; c < 8; ++c) {
        const double *p = C + cgns_t[c] * 3;
        const double d0 = p[0] - x[0], d1 = p[1] - x[1], d2 = p[2] - x[2];
        const double nrm = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        if (nrm > 0.) wt[c] = 1. / nrm;
        else { hit = c; break; }
      }
      if (hit >= 0)
        for (int c = 0; c < 8; ++c) wt[c] = c == hit ? 1.0 : 0.0;
      double sum = 0.0;
      for (int c = 0; c < 8; ++c) sum = sum + wt[c];
#pragma unroll
      for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int k = 0; k < 4; ++k) acc[c][k] = acc[c][k] + wt[c] / sum * TS[k];
    }
  }
  // butterfly reduction over the wave
#pragma unroll
  for (int c = 0; c < 8; ++c)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double v = acc[c][k];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, kWave);
      acc[c][k] = v;
    }
  if (lane < 32) {
    double out = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (c * 4 + k == lane) out = acc[c][k];
    elemCorner[(size_t)w * 32 + lane] = out;
  }
}

// node classes (a node and its periodic twins): sum incidences, divide by
// NodeVolume (pic_depo_method.f90:688-695)
__global__ __launch_bounds__(kBlock) void k_cvwm_class(int nClasses, const int32_t *off, const int32_t *inc,
                                                       const double *vol, const double *elemCorner, double *classSrc) {
  const int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= nClasses) return;
  double s[4] = {0, 0, 0, 0};
  for (int q = off[c]; q < off[c + 1]; ++q) {
    const double *ec = elemCorner + (size_t)inc[q] * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) s[k] = s[k] + ec[k];
  }
  const double v = vol[c];
#pragma unroll
  for (int k = 0; k < 4; ++k) classSrc[(size_t)c * 4 + k] = v > 0. ? s[k] / v : s[k];
}

// node -> DOF trilinear interpolation (pic_depo_method.f90:704-733)
template <int N>
__global__ __launch_bounds__(kBlock) void k_cvwm_dof(DevMesh M, const int32_t *elemClass, const double *classSrc, double *PS) {
  constexpr int n1 = N + 1, nd = n1 * n1 * n1;
  const int64_t g = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (g >= int64_t(M.nElems) * nd) return;
  const int e = int(g / nd), r = int(g % nd);
  const int kk = r % n1, ll = (r / n1) % n1, mm = r / (n1 * n1);
  const double a1 = (M.xGP[kk] + 1.0) / 2.0, a2 = (M.xGP[ll] + 1.0) / 2.0, a3 = (M.xGP[mm] + 1.0) / 2.0;
  const int32_t *cl = elemClass + (size_t)e * 8;
  double *ps = PS + g * 4;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double *S0 = classSrc + (size_t)cl[0] * 4, *S1 = classSrc + (size_t)cl[1] * 4;
    const double *S2 = classSrc + (size_t)cl[2] * 4, *S3 = classSrc + (size_t)cl[3] * 4;
    const double *S4 = classSrc + (size_t)cl[4] * 4, *S5 = classSrc + (size_t)cl[5] * 4;
    const double *S6 = classSrc + (size_t)cl[6] * 4, *S7 = classSrc + (size_t)cl[7] * 4;
    ps[k] = S0[k] * (1 - a1) * (1 - a2) * (1 - a3) + S1[k] * (a1) * (1 - a2) * (1 - a3) +
            S2[k] * (a1) * (a2) * (1 - a3) + S3[k] * (1 - a1) * (a2) * (1 - a3) +
            S4[k] * (1 - a1) * (1 - a2) * (a3) + S5[k] * (a1) * (1 - a2) * (a3) +
            S6[k] * (a1) * (a2) * (a3) + S7[k] * (1 - a1) * (a2) * (a3);
  }
}

// ---- shape_function ---------------------------------------------------------
// S = 1 - r^2/R^2 <= 1, so max(S, 0) = min(max(S, 0), 1): written this way the
// clamp becomes the clamp bit of the instruction that forms S (one operation
// less per particle-DOF pair).
__device__ inline double clamp01(double x) { return fmin(fmax(x, 0.0), 1.0); }
__device__ inline float clamp01f(float x) { return fminf(fmaxf(x, 0.f), 1.f); }

// Gather form of depoChargeOnDOFsSF (pic_depo_shapefunction_tools.f90:292-422):
// one wave per target element, one lane per DOF.  For every (source element,
// periodic shift) of the target's list the wave streams the source segment,
// keeps the particles the reference's element cull admits
// (|x - bary| <= r_sf + ElemRadius), stages them in LDS and every lane adds
// S^alpha * w_sf * q * (v,1) of the particles within r_sf of its DOF.
struct SFArgs {
  const int32_t *off, *src;
  const uint8_t *shift;
  const int32_t *roff, *rfirst, *rend;   // the lists as runs of consecutive source elements
  const uint8_t *rshift;
  double sh[27][3];
  double r_sf, r2_sf, r2_sf_inv, w_sf;
  double fsp[8];   // per species: ChargeIC * MacroParticleFactor * w_sf
  // per sorted particle, in two planes: rec[i] = (x, y, z, f) (what the culls
  // read) and recB[i] = (f vx, f vy, f vz, f) (read for the staged records)
  const double4 *rec;
  const double4 *recB;
  // per block of 64 records: the box of its records (record_block_box)
  const double4 *blk;
  double r_inv;         // 1 / r_sf
  int nspec;
  int cartesian;        // every element's Gauss points form an axis-aligned tensor grid
  int alpha;
  int linear;   // NGeo == 1: the barycentre cull is implied by the DOF-box test
  int32_t *work;   // [8] target counters of the persistent gathers (zeroed before each launch)
  const int32_t *groff, *grfirst, *grend;   // runs of the union of the lists of 4 consecutive targets
  const uint8_t *grshift;
};

// record c of a run whose plane starts at the wave-uniform base b: a 32-bit
// offset from a scalar base (the saddr + voffset form of the load)
__device__ inline double4 rec_at(const double4 *b, int c) {
  return *reinterpret_cast<const double4 *>(reinterpret_cast<const char *>(b) + unsigned(c) * 32u);
}

template <int N, int ALPHA>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE))) void k_sf_gather(DevMesh M, DevParts P, const int32_t *start, SFArgs A, double *PS) {
  // Results need only agree with the reference to round-off (the sums run in
  // another order anyway), so this kernel allows fused multiply-add.
#pragma clang fp contract(fast)
  constexpr int n1 = N + 1, nd = n1 * n1 * n1, nch = (nd + kWave - 1) / kWave;
  __shared__ double lds[kBlock / kWave][kWave][8];
  const int wib = threadIdx.x / kWave;
  const int lane = lane_id();
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int t = lb * (kBlock / kWave) + wib;
  if (t >= M.nElems) return;
  // DOFs of the target in coordinates centred on its barycentre:
  //   r^2 = |d|^2 + |p|^2 - 2 d.p   with d, p relative to the barycentre
  const double bx = M.bary[t * 3], by = M.bary[t * 3 + 1], bz = M.bary[t * 3 + 2];
  //   S = 1 - r^2/R^2 = cd - (|p|^2/R^2) + (2/R^2) d.p   (one add, three FMA)
  // lanes without a DOF get cd = -1e300 so that max(S,0) = 0.
  const double R2i0 = A.r2_sf_inv;
  double mx[nch], my[nch], mz[nch], cd[nch], acc[nch][4];
  bool has[nch];
  double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
#pragma unroll
  for (int c = 0; c < nch; ++c) {
    const int d = c * kWave + lane;
    has[c] = d < nd;
    const double *xg = M.xGPe + ((size_t)t * nd + (has[c] ? d : 0)) * 3;
    const double dx = xg[0] - bx, dy = xg[1] - by, dz = xg[2] - bz;
    mx[c] = 2.0 * R2i0 * dx; my[c] = 2.0 * R2i0 * dy; mz[c] = 2.0 * R2i0 * dz;
    cd[c] = has[c] ? 1.0 - R2i0 * (dx * dx + dy * dy + dz * dz) : -1e300;
    if (has[c]) {
      blo[0] = fmin(blo[0], dx); blo[1] = fmin(blo[1], dy); blo[2] = fmin(blo[2], dz);
      bhi[0] = fmax(bhi[0], dx); bhi[1] = fmax(bhi[1], dy); bhi[2] = fmax(bhi[2], dz);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) acc[c][k] = 0.0;
  }
  // bounding box of the target DOFs (wave min/max)
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      blo[a] = fmin(blo[a], __shfl_xor(blo[a], off, kWave));
      bhi[a] = fmax(bhi[a], __shfl_xor(bhi[a], off, kWave));
    }
  // the reference's element cull |x - bary| > r_sf + ElemRadius, compared squared
  const double thr = A.r_sf + M.radius[t], thr2 = thr * thr;
  const double R2 = A.r2_sf;
  const double R2box = R2 * (1.0 + 1e-12);
  double (*L)[8] = lds[wib];
  for (int q = A.off[t]; q < A.off[t + 1]; ++q) {
    const int s = A.src[q];
    const int si = A.shift[q];
    const double shx = A.sh[si][0], shy = A.sh[si][1], shz = A.sh[si][2];
    const int s0 = start[s], s1 = start[s + 1];
    for (int base = s0; base < s1; base += kWave) {
      const int i = base + lane;
      bool pass = false;
      double px = 0, py = 0, pz = 0, f = 0, vx = 0, vy = 0, vz = 0;
      if (i < s1) {
        const int sp = P.spec[i];
        const double q0 = M.charge[sp];
        // shifted position (GetPartPosShifted) relative to the target barycentre
        px = (P.x[i] + shx) - bx; py = (P.y[i] + shy) - by; pz = (P.z[i] + shz) - bz;
        // the reference's element cull |x - bary| > r_sf + ElemRadius, and the
        // exact test that some DOF of the target lies within r_sf
        const double gx = fmax(0.0, fmax(blo[0] - px, px - bhi[0]));
        const double gy = fmax(0.0, fmax(blo[1] - py, py - bhi[1]));
        const double gz = fmax(0.0, fmax(blo[2] - pz, pz - bhi[2]));
        // On linear elements the cull is implied by the box test (every DOF
        // lies within ElemRadius of the barycentre), so only the box test runs.
        pass = (fabs(q0) > 0.0) && (gx * gx + gy * gy + gz * gz <= R2box) &&
               (A.linear || !(px * px + py * py + pz * pz > thr2));
        f = q0 * M.mpf[sp] * A.w_sf;
        vx = P.vx[i]; vy = P.vy[i]; vz = P.vz[i];
      }
      const unsigned long long m = __ballot(pass);
      const int cnt = __popcll(m);
      if (cnt == 0) continue;
      if (pass) {
        const int slot = __popcll(m & ((1ull << lane) - 1ull));
        L[slot][0] = px; L[slot][1] = py; L[slot][2] = pz;
        L[slot][3] = R2i0 * (px * px + py * py + pz * pz);
        L[slot][4] = vx * f; L[slot][5] = vy * f; L[slot][6] = vz * f; L[slot][7] = f;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      auto deposit_one = [&](int j) {
        const double qx = L[j][0], qy = L[j][1], qz = L[j][2], qq = L[j][3];
        const double g0 = L[j][4], g1 = L[j][5], g2 = L[j][6], g3 = L[j][7];
#pragma unroll
        for (int c = 0; c < nch; ++c) {
          // radius2 <= r2_sf  <=>  S >= 0 (up to round-off, where S^alpha ~ 0)
          const double S = clamp01(mx[c] * qx + (my[c] * qy + (mz[c] * qz + (cd[c] - qq))));
          double S1 = S * S;
          if (ALPHA == 0) {
            for (int ex = 3; ex <= A.alpha; ++ex) S1 = S * S1;
          } else {
#pragma unroll
            for (int ex = 3; ex <= ALPHA; ++ex) S1 = S * S1;
          }
          acc[c][0] += S1 * g0;
          acc[c][1] += S1 * g1;
          acc[c][2] += S1 * g2;
          acc[c][3] += S1 * g3;
        }
      };
      // two particles per trip so the LDS reads of the second overlap the first's arithmetic
      int j = 0;
      for (; j + 1 < cnt; j += 2) {
        deposit_one(j);
        deposit_one(j + 1);
      }
      if (j < cnt) deposit_one(j);
      __builtin_amdgcn_wave_barrier();
    }
  }
#pragma unroll
  for (int c = 0; c < nch; ++c)
    if (has[c]) {
      double *ps = PS + ((size_t)t * nd + c * kWave + lane) * 4;
      ps[0] = acc[c][0]; ps[1] = acc[c][1]; ps[2] = acc[c][2]; ps[3] = acc[c][3];
    }
}

// Shape-function gather for N = 3 (64 DOFs) with the accumulation on the
// matrix cores.  PartSource(dof, c) += sum_p S(dof, p)^alpha * f_p (v_p, 1)_c
// is, per group of 4 staged particles, a 64x4x4 product; v_mfma_f64_4x4x4
// (4 blocks of 4x4x4) does a 16-DOF quarter of it, and runs beside the VALU
// that forms S.  Lane layout of that instruction (A: lane 16k+4b+i holds
// A_b[i][k]; B: lane 16k+4b+j holds B_b[k][j]; D: lane 16i+4b+j holds
// D_b[i][j]) puts particle k = lane/16 and DOF 16Q + lane%16 on each lane for
// the operands, and DOF 16Q + 4((lane/4)%4) + lane/16, component lane%4 on
// the accumulator.  The wave is one target element; the segment of every
// (source element, shift) of its list is walked in chunks of 64 particle
// records.
// Deposition records of the sorted particles for the shape-function
// gathers: (x, y, z, 0), (f vx, f vy, f vz, f) with f = ChargeIC *
// MacroParticleFactor * w_sf of the particle's species, so that a gather
// culls on one 32-byte record per particle and target and reads the second
// only for the particles it keeps.
__global__ __launch_bounds__(kBlock) void k_sf_records(DevParts P, const int32_t *count, SFArgs A, double4 *rec,
                                                       double4 *recB, double4 *blk) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t n = *count;
  if ((i & ~int64_t(kWave - 1)) >= n) return;   // the whole wave is past the end
  const bool valid = i < n;
  double x = 0.0, y = 0.0, z = 0.0, f = 0.0;
  if (valid) {
    f = A.fsp[P.spec[i]];
    x = P.x[i]; y = P.y[i]; z = P.z[i];
    rec[i] = make_double4(x, y, z, f);
    recB[i] = make_double4(P.vx[i] * f, P.vy[i] * f, P.vz[i] * f, f);
  }
  if (blk) record_block_box(blk, i, valid && f != 0.0, x, y, z);
}

// The next target of a persistent gather wave: targets are taken in
// order from 8 contiguous ranges, one per XCD (the XCD that runs a workgroup
// is its block index mod 8), each by an atomic counter; a wave whose range
// is exhausted helps in the next.  Every target is computed by one wave in
// the same way, so the results do not depend on which wave takes it.
__device__ inline int sf_next_target(const SFArgs &A, int nE, int &range, int lane) {
  for (int tried = 0; tried < 8; ++tried) {
    int v = 0;
    if (lane == 0) v = atomicAdd(&A.work[range], 1);
    v = __builtin_amdgcn_readfirstlane(v);
    const int lo = int(int64_t(nE) * range / 8) + v, hi = int(int64_t(nE) * (range + 1) / 8);
    if (lo < hi) return lo;
    range = (range + 1) & 7;
  }
  return nE;
}

// Per wave a ring of 128 staged records (x, y, z relative to the target
// barycentre, |x|^2/R^2, f vx, f vy, f vz, f): particles that pass the cull
// are appended, and groups of 4 are consumed once 64 are pending, so groups
// span source chunks and only the very last group is padded.
constexpr int kRingMfma = 128;
template <int ALPHA, bool LINEAR>
__device__ __attribute__((always_inline)) inline void sf_gather_mfma_target(const DevMesh &M, const int32_t *start,
                                                                            const SFArgs &A, double *PS, int t,
                                                                            double (*L)[8], int lane) {
#pragma clang fp contract(fast)
  constexpr int nd = 64;
  constexpr int kRing = kRingMfma;
  const double bx = M.bary[t * 3], by = M.bary[t * 3 + 1], bz = M.bary[t * 3 + 2];
  const double R2i0 = A.r2_sf_inv;
  // operand-side DOF constants, one DOF per quarter:  S = cd - qq + m.q
  double mx[4], my[4], mz[4], cd[4], acc[4];
  double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
#pragma unroll
  for (int Q = 0; Q < 4; ++Q) {
    const int d = 16 * Q + (lane & 15);
    const double *xg = M.xGPe + ((size_t)t * nd + d) * 3;
    const double dx = xg[0] - bx, dy = xg[1] - by, dz = xg[2] - bz;
    mx[Q] = 2.0 * R2i0 * dx; my[Q] = 2.0 * R2i0 * dy; mz[Q] = 2.0 * R2i0 * dz;
    cd[Q] = 1.0 - R2i0 * (dx * dx + dy * dy + dz * dz);
    blo[0] = fmin(blo[0], dx); blo[1] = fmin(blo[1], dy); blo[2] = fmin(blo[2], dz);
    bhi[0] = fmax(bhi[0], dx); bhi[1] = fmax(bhi[1], dy); bhi[2] = fmax(bhi[2], dz);
    acc[Q] = 0.0;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      blo[a] = fmin(blo[a], __shfl_xor(blo[a], off, kWave));
      bhi[a] = fmax(bhi[a], __shfl_xor(bhi[a], off, kWave));
    }
  // the reference's element cull |x - bary| > r_sf + ElemRadius, compared squared
  const double thr = A.r_sf + M.radius[t], thr2 = thr * thr;
  const double R2box = A.r2_sf * (1.0 + 1e-12);
  const double R2skip = R2box * (1.0 + 1e-9), thr2skip = thr2 * (1.0 + 1e-9);
  const int kp = lane >> 4, jc = lane & 3;
  // groups read 64 consecutive ring records from a multiple of 64 (no wrap)
  auto group = [&](const double (*Lh)[8], int g) {
    const double *rec = Lh[4 * g + kp];
    const double qx = rec[0], qy = rec[1], qz = rec[2], qq = rec[3];
    const double bval = rec[4 + jc];
#pragma unroll
    for (int Q = 0; Q < 4; ++Q) {
      const double S = clamp01(mx[Q] * qx + (my[Q] * qy + (mz[Q] * qz + (cd[Q] - qq))));
      double S1 = S * S;
      if (ALPHA == 0) {
        for (int ex = 3; ex <= A.alpha; ++ex) S1 = S * S1;
      } else {
#pragma unroll
        for (int ex = 3; ex <= ALPHA; ++ex) S1 = S * S1;
      }
      acc[Q] = __builtin_amdgcn_mfma_f64_4x4x4f64(S1, bval, acc[Q], 0, 0, 0);
    }
  };
  int head = 0, tail = 0;   // ring positions in records (head multiple of 64)
  // The target's list, as runs of consecutive source elements with one
  // shift (contiguous particle segments), is read 64 runs at a time, one per
  // lane (segment bounds and shift); the walk takes them with readlane.
  const int q0 = A.roff[t], qend = A.roff[t + 1];
  for (int qb = q0; qb < qend; qb += kWave) {
    int ms0 = 0, ms1 = 0, msi = 0;
    if (qb + lane < qend) {
      ms0 = start[A.rfirst[qb + lane]]; ms1 = start[A.rend[qb + lane]]; msi = A.rshift[qb + lane];
    }
    const int jn = min(kWave, qend - qb);
    for (int j = 0; j < jn; ++j) {
      const int s0 = __builtin_amdgcn_readlane(ms0, j), s1 = __builtin_amdgcn_readlane(ms1, j);
      const int si = __builtin_amdgcn_readlane(msi, j);
      // shifted position (GetPartPosShifted) relative to the target barycentre
      const double ox = A.sh[si][0] - bx, oy = A.sh[si][1] - by, oz = A.sh[si][2] - bz;
      // the run's records from a wave-uniform base
      const double4 *ra_s = A.rec + s0, *rb_s = A.recB + s0;
      const int last = s1 - 1 - s0;
      // the exact test that some DOF of the target lies within r_sf, and on
      // curved elements the reference's element cull |x - bary| > r_sf + ElemRadius
      auto cull = [&](const double4 &ra, bool valid, double &px, double &py, double &pz) {
        px = ra.x + ox; py = ra.y + oy; pz = ra.z + oz;
        const double gx = fmax(0.0, fmax(blo[0] - px, px - bhi[0]));
        const double gy = fmax(0.0, fmax(blo[1] - py, py - bhi[1]));
        const double gz = fmax(0.0, fmax(blo[2] - pz, pz - bhi[2]));
        bool pass = valid && (ra.w != 0.0) && (gx * gx + gy * gy + gz * gz <= R2box);
        // (on linear elements every DOF lies within ElemRadius of the
        // barycentre, so the box test implies the cull)
        if (!LINEAR) pass = pass && !(px * px + py * py + pz * pz > thr2);
        return pass;
      };
      // stage the kept records of one chunk and consume 64 once pending
      auto stage = [&](bool pass, double px, double py, double pz, const double4 &rb) {
        const unsigned long long m = __ballot(pass);
        if (m == 0) return;
        if (pass) {
          double *r = L[(tail + __popcll(m & ((1ull << lane) - 1ull))) & (kRing - 1)];
          r[0] = px; r[1] = py; r[2] = pz;
          r[3] = R2i0 * (px * px + py * py + pz * pz);
          r[4] = rb.x; r[5] = rb.y; r[6] = rb.z; r[7] = rb.w;
        }
        tail += __popcll(m);
        if (tail - head < kWave) return;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int g = 0; g < kWave / 4; ++g) group(L + (head & (kRing - 1)), g);
        head += kWave;
        __builtin_amdgcn_wave_barrier();
      };
      // The run is walked in the aligned blocks of 64 records that it
      // overlaps.  A block whose box (record_block_box) lies farther than
      // r_sf from the DOF box, or on curved elements farther than
      // r_sf + ElemRadius from the barycentre, holds no record the cull
      // keeps (each record's distances are at least the box's, and the
      // skip thresholds carry a 1e-9 margin over the cull's for round-off):
      // it is not read.  The blocks of up to 64 at a time are tested one
      // per lane; the kept records are staged in record order, as without
      // the skip.
      const int bfirst = s0 >> 6, blast = s1 > s0 ? (s1 - 1) >> 6 : bfirst - 1;
      for (int wb = bfirst; wb <= blast; wb += kWave) {
        bool reach = false;
        if (wb + lane <= blast) {
          const double4 lo = A.blk[2 * (wb + lane)], hi = A.blk[2 * (wb + lane) + 1];
          const double lx = lo.x + ox, ly = lo.y + oy, lz = lo.z + oz;
          const double hx = hi.x + ox, hy = hi.y + oy, hz = hi.z + oz;
          const double gx = fmax(0.0, fmax(blo[0] - hx, lx - bhi[0]));
          const double gy = fmax(0.0, fmax(blo[1] - hy, ly - bhi[1]));
          const double gz = fmax(0.0, fmax(blo[2] - hz, lz - bhi[2]));
          reach = !(gx * gx + gy * gy + gz * gz > R2skip);
          if (!LINEAR) {
            const double nx = fmax(0.0, fmax(lx, -hx)), ny = fmax(0.0, fmax(ly, -hy)), nz = fmax(0.0, fmax(lz, -hz));
            reach = reach && !(nx * nx + ny * ny + nz * nz > thr2skip);
          }
        }
        unsigned long long km = __ballot(reach);
        // two kept blocks per pass: both position records loaded before
        // either is culled, then the current records of the kept particles
        while (km) {
          const int ia = __builtin_ctzll(km);
          km &= km - 1;
          const bool two = km != 0;
          const int ib = two ? __builtin_ctzll(km) : ia;
          if (two) km &= km - 1;
          const int k0 = ((wb + ia) << 6) + lane - s0, k1 = ((wb + ib) << 6) + lane - s0;
          const bool v0 = k0 >= 0 && k0 <= last, v1 = two && k1 >= 0 && k1 <= last;
          const double4 ra0 = rec_at(ra_s, min(max(k0, 0), last));
          const double4 ra1 = rec_at(ra_s, min(max(k1, 0), last));
          double px0, py0, pz0, px1, py1, pz1;
          const bool p0 = cull(ra0, v0, px0, py0, pz0);
          const bool p1 = cull(ra1, v1, px1, py1, pz1);
          double4 rb0 = make_double4(0.0, 0.0, 0.0, 0.0), rb1 = rb0;
          if (p0) rb0 = rec_at(rb_s, k0);
          if (p1) rb1 = rec_at(rb_s, k1);
          stage(p0, px0, py0, pz0, rb0);
          if (two) stage(p1, px1, py1, pz1, rb1);
        }
      }
    }
  }
  // the rest, with the last group padded by weightless records
  const int rest = tail - head;
  if (rest > 0) {
    const int ng = (rest + 3) >> 2;
    if (lane >= rest && lane < 4 * ng) {
      double *r = L[(head + lane) & (kRing - 1)];
#pragma unroll
      for (int c = 0; c < 8; ++c) r[c] = 0.0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int g = 0; g < ng; ++g) group(L + (head & (kRing - 1)), g);
  }
  const int dofLo = 4 * ((lane >> 2) & 3) + (lane >> 4);
#pragma unroll
  for (int Q = 0; Q < 4; ++Q) PS[((size_t)t * nd + 16 * Q + dofLo) * 4 + jc] = acc[Q];
}

// One wave per target element (PERSIST false: target from the block index,
// XCD-contiguous), or persistent waves that take targets from the counters
// of sf_next_target (PERSIST true: launched at the resident capacity).
template <int ALPHA, bool LINEAR, bool PERSIST>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE_MX))) void k_sf_gather_mfma(DevMesh M, const int32_t *start, SFArgs A, double *PS) {
  __shared__ double lds[kBlock / kWave][kRingMfma][8];
  const int wib = threadIdx.x / kWave;
  const int lane = lane_id();
  if (!PERSIST) {
    const int t = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * (kBlock / kWave) + wib);
    if (t < M.nElems) sf_gather_mfma_target<ALPHA, LINEAR>(M, start, A, PS, t, lds[wib], lane);
    return;
  }
  int range = blockIdx.x & 7;
  for (int t = sf_next_target(A, M.nElems, range, lane); t < M.nElems; t = sf_next_target(A, M.nElems, range, lane)) {
    sf_gather_mfma_target<ALPHA, LINEAR>(M, start, A, PS, t, lds[wib], lane);
    __builtin_amdgcn_wave_barrier();
  }
}

// The general gather with the loads shared by a workgroup: its four waves
// take four consecutive target elements (neighbours that share most of their
// sources) and walk the union of their lists in chunks of 128 records, staged
// once in LDS for the four; each wave culls every staged record against its
// own target.  A target's own records pass its cull in the order of its own
// runs (a record of an element that is not in its list lies farther than
// r_sf from its DOF box), so each target's ring, groups and sums are those of
// k_sf_gather_mfma, bit for bit; the records are read from memory once per
// workgroup instead of once per wave, for ~1.3x the cull work per wave.
constexpr int kStageCoop = 128;
template <int ALPHA, bool LINEAR>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE_MX))) void k_sf_gather_mfma_coop(DevMesh M, const int32_t *start, SFArgs A, double *PS) {
#pragma clang fp contract(fast)
  constexpr int nd = 64;
  constexpr int kRing = kRingMfma;
  __shared__ double lds[kBlock / kWave][kRingMfma][8];
  __shared__ double4 stg[kStageCoop][2];
  const int wib = threadIdx.x / kWave;
  const int lane = lane_id();
  const int grp = xcd_block(blockIdx.x, gridDim.x);               // block-uniform
  if (grp * (kBlock / kWave) >= M.nElems) return;                 // blocks of the XCD padding
  const int t = __builtin_amdgcn_readfirstlane(grp * (kBlock / kWave) + wib);
  const bool live = t < M.nElems;                                 // wave-uniform
  const int te = live ? t : 0;
  double (*L)[8] = lds[wib];
  const double bx = M.bary[te * 3], by = M.bary[te * 3 + 1], bz = M.bary[te * 3 + 2];
  const double R2i0 = A.r2_sf_inv;
  double mx[4], my[4], mz[4], cd[4], acc[4];
  double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
#pragma unroll
  for (int Q = 0; Q < 4; ++Q) {
    const int d = 16 * Q + (lane & 15);
    const double *xg = M.xGPe + ((size_t)te * nd + d) * 3;
    const double dx = xg[0] - bx, dy = xg[1] - by, dz = xg[2] - bz;
    mx[Q] = 2.0 * R2i0 * dx; my[Q] = 2.0 * R2i0 * dy; mz[Q] = 2.0 * R2i0 * dz;
    cd[Q] = 1.0 - R2i0 * (dx * dx + dy * dy + dz * dz);
    blo[0] = fmin(blo[0], dx); blo[1] = fmin(blo[1], dy); blo[2] = fmin(blo[2], dz);
    bhi[0] = fmax(bhi[0], dx); bhi[1] = fmax(bhi[1], dy); bhi[2] = fmax(bhi[2], dz);
    acc[Q] = 0.0;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      blo[a] = fmin(blo[a], __shfl_xor(blo[a], off, kWave));
      bhi[a] = fmax(bhi[a], __shfl_xor(bhi[a], off, kWave));
    }
  const double thr = A.r_sf + M.radius[te], thr2 = thr * thr;
  const double R2box = A.r2_sf * (1.0 + 1e-12);
  const int kp = lane >> 4, jc = lane & 3;
  auto group = [&](const double (*Lh)[8], int g) {
    const double *rec = Lh[4 * g + kp];
    const double qx = rec[0], qy = rec[1], qz = rec[2], qq = rec[3];
    const double bval = rec[4 + jc];
#pragma unroll
    for (int Q = 0; Q < 4; ++Q) {
      const double S = clamp01(mx[Q] * qx + (my[Q] * qy + (mz[Q] * qz + (cd[Q] - qq))));
      double S1 = S * S;
      if (ALPHA == 0) {
        for (int ex = 3; ex <= A.alpha; ++ex) S1 = S * S1;
      } else {
#pragma unroll
        for (int ex = 3; ex <= ALPHA; ++ex) S1 = S * S1;
      }
      acc[Q] = __builtin_amdgcn_mfma_f64_4x4x4f64(S1, bval, acc[Q], 0, 0, 0);
    }
  };
  int head = 0, tail = 0;
  const int q0 = A.groff[grp], qend = A.groff[grp + 1];           // block-uniform
  for (int qb = q0; qb < qend; qb += kWave) {
    int ms0 = 0, ms1 = 0, msi = 0;
    if (qb + lane < qend) {
      ms0 = start[A.grfirst[qb + lane]]; ms1 = start[A.grend[qb + lane]]; msi = A.grshift[qb + lane];
    }
    const int jn = min(kWave, qend - qb);
    for (int j = 0; j < jn; ++j) {
      const int s0 = __builtin_amdgcn_readlane(ms0, j), s1 = __builtin_amdgcn_readlane(ms1, j);
      const int si = __builtin_amdgcn_readlane(msi, j);
      const double ox = A.sh[si][0] - bx, oy = A.sh[si][1] - by, oz = A.sh[si][2] - bz;
      const int cnt = s1 - s0;
      for (int base = 0; base < cnt; base += kStageCoop) {
        // stage the chunk once for the workgroup (the previous chunk's reads
        // ended at the barrier below)
        if (threadIdx.x < kStageCoop && base + int(threadIdx.x) < cnt) {
          const size_t r = size_t(s0 + base + int(threadIdx.x));
          stg[threadIdx.x][0] = A.rec[r];
          stg[threadIdx.x][1] = A.recB[r];
        }
        __syncthreads();
        if (live) {
          for (int sub = 0; sub < kStageCoop && base + sub < cnt; sub += kWave) {
            const int k = sub + lane;
            const bool valid = base + k < cnt;
            const double4 ra = stg[valid ? k : 0][0], rb = stg[valid ? k : 0][1];
            const double px = ra.x + ox, py = ra.y + oy, pz = ra.z + oz;
            const double gx = fmax(0.0, fmax(blo[0] - px, px - bhi[0]));
            const double gy = fmax(0.0, fmax(blo[1] - py, py - bhi[1]));
            const double gz = fmax(0.0, fmax(blo[2] - pz, pz - bhi[2]));
            bool pass = valid && (ra.w != 0.0) && (gx * gx + gy * gy + gz * gz <= R2box);
            if (!LINEAR) pass = pass && !(px * px + py * py + pz * pz > thr2);
            const unsigned long long m = __ballot(pass);
            if (m == 0) continue;
            if (pass) {
              double *r = L[(tail + __popcll(m & ((1ull << lane) - 1ull))) & (kRing - 1)];
              r[0] = px; r[1] = py; r[2] = pz;
              r[3] = R2i0 * (px * px + py * py + pz * pz);
              r[4] = rb.x; r[5] = rb.y; r[6] = rb.z; r[7] = rb.w;
            }
            tail += __popcll(m);
            if (tail - head < kWave) continue;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int g = 0; g < kWave / 4; ++g) group(L + (head & (kRing - 1)), g);
            head += kWave;
            __builtin_amdgcn_wave_barrier();
          }
        }
        __syncthreads();
      }
    }
  }
  if (!live) return;
  const int rest = tail - head;
  if (rest > 0) {
    const int ng = (rest + 3) >> 2;
    if (lane >= rest && lane < 4 * ng) {
      double *r = L[(head + lane) & (kRing - 1)];
#pragma unroll
      for (int c = 0; c < 8; ++c) r[c] = 0.0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int g = 0; g < ng; ++g) group(L + (head & (kRing - 1)), g);
  }
  const int dofLo = 4 * ((lane >> 2) & 3) + (lane >> 4);
#pragma unroll
  for (int Q = 0; Q < 4; ++Q) PS[((size_t)t * nd + 16 * Q + dofLo) * 4 + jc] = acc[Q];
}

// The same gather for targets whose Gauss points form an axis-aligned tensor
// grid (Cartesian elements): with u = x/r_sf, S = 1 - (u_i - pu)^2 -
// (v_j - pv)^2 - (w_k - pw)^2 = [K_ij + 2 u_i pu + 2 v_j pv] - C_k with the
// lane constant K_ij = 1 - u_i^2 - v_j^2 and the per-particle
// C_k = (w_k - pw)^2 + pu^2 + pv^2, formed once at staging.  The bracket costs
// two fused multiply-adds per lane and group of 4 particles; a pair then costs
// a subtraction, the clamp and the square.  Lane and accumulator layout as
// k_sf_gather_mfma, with quarter Q = k and lane%16 = i + 4j.
// ring of 128 staged records: pu, pv, C_k for k = 0..3, f vx, f vy, f vz, f
// (staging writes pw in the place of C_0; a pass with one record per lane
// forms the four C_k before each batch of 64 is consumed)
constexpr int kRingCart = 128, kRecCart = 10;
template <int ALPHA>
__device__ __attribute__((always_inline)) inline void sf_gather_cart_target(const DevMesh &M, const int32_t *start,
                                                                            const SFArgs &A, double *PS, int t,
                                                                            double (*L)[kRecCart], int lane) {
#pragma clang fp contract(fast)
  constexpr int nd = 64;
  constexpr int kRing = kRingCart, kRec = kRecCart;
  const double bx = M.bary[t * 3], by = M.bary[t * 3 + 1], bz = M.bary[t * 3 + 2];
  const double Ri = A.r_inv;
  const double *xg = M.xGPe + (size_t)t * nd * 3;
  // scaled DOF coordinates relative to the barycentre: u_i = x(i,0,0), v_j = y(0,j,0), w_k = z(0,0,k)
  double u[4], v[4], w[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    u[a] = (xg[a * 3 + 0] - bx) * Ri;
    v[a] = (xg[4 * a * 3 + 1] - by) * Ri;
    w[a] = (xg[16 * a * 3 + 2] - bz) * Ri;
  }
  const double ulo = fmin(fmin(u[0], u[1]), fmin(u[2], u[3])), uhi = fmax(fmax(u[0], u[1]), fmax(u[2], u[3]));
  const double vlo = fmin(fmin(v[0], v[1]), fmin(v[2], v[3])), vhi = fmax(fmax(v[0], v[1]), fmax(v[2], v[3]));
  const double wlo = fmin(fmin(w[0], w[1]), fmin(w[2], w[3])), whi = fmax(fmax(w[0], w[1]), fmax(w[2], w[3]));
  // Staged coordinates are taken relative to the centre of the DOF box, so
  // that the cull is |p| - half width per axis (the absolute value is a free
  // operand modifier); the DOF constants move to the same frame.
  const double uc = 0.5 * (ulo + uhi), vc = 0.5 * (vlo + vhi), wc = 0.5 * (wlo + whi);
  const double uh = 0.5 * (uhi - ulo), vh = 0.5 * (vhi - vlo), wh = 0.5 * (whi - wlo);
#pragma unroll
  for (int a = 0; a < 4; ++a) w[a] -= wc;
  const double ui = (xg[(lane & 3) * 3 + 0] - bx) * Ri - uc;
  const double vj = (xg[4 * ((lane >> 2) & 3) * 3 + 1] - by) * Ri - vc;
  const double Kij = 1.0 - ui * ui - vj * vj, tu = 2.0 * ui, tv = 2.0 * vj;
  const double box2 = 1.0 + 1e-12;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int kp = lane >> 4, jc = lane & 3;
  // live[Q] bit r: staged record r of the batch reaches some DOF of the
  // plane k = Q; a group of 4 records with no live bit in a plane has S = 0
  // at all 16 of its DOFs (the clamp), so that plane's products are skipped
  auto group = [&](const double (*Lh)[kRec], int g, const unsigned long long *live) {
    const double *rec = Lh[4 * g + kp];
    const double W = fma(tv, rec[1], fma(tu, rec[0], Kij));
    double bval = rec[6 + jc];
    // the four C_k and f(v,1) are loaded together (the empty asm keeps the
    // loads ahead of the plane tests); planes no record of the group reaches
    // are skipped
    double C[4] = {rec[2], rec[3], rec[4], rec[5]};
    asm volatile("" : "+v"(C[0]), "+v"(C[1]), "+v"(C[2]), "+v"(C[3]), "+v"(bval));
#pragma unroll
    for (int Q = 0; Q < 4; ++Q) {
      if (((live[Q] >> (4 * g)) & 0xFull) == 0) continue;
      const double S = clamp01(W - C[Q]);
      double S1 = S * S;
      if (ALPHA == 0) {
        for (int ex = 3; ex <= A.alpha; ++ex) S1 = S * S1;
      } else {
#pragma unroll
        for (int ex = 3; ex <= ALPHA; ++ex) S1 = S * S1;
      }
      acc[Q] = __builtin_amdgcn_mfma_f64_4x4x4f64(S1, bval, acc[Q], 0, 0, 0);
    }
  };
  // C_k = (w_k - pw)^2 + pu^2 + pv^2 of the 64 records from Lh, one per lane
  // (rows past the staged ones are overwritten harmlessly and never read),
  // and the planes each record reaches: plane k is out of reach when the
  // distance from (pu, pv) to the box of the (u_i, v_j), squared, plus
  // (w_k - pw)^2 is not below 1 (a margin of 1e-9 covers the rounding of S)
  auto form_ck = [&](double (*Lh)[kRec], unsigned long long *live) {
    double *r = Lh[lane];
    const double pu = r[0], pv = r[1], pw = r[2];
    const double c0 = w[0] - pw, c1 = w[1] - pw, c2 = w[2] - pw, c3 = w[3] - pw;
    const double P = fma(pv, pv, pu * pu);
    r[2] = fma(c0, c0, P); r[3] = fma(c1, c1, P); r[4] = fma(c2, c2, P); r[5] = fma(c3, c3, P);
    const double gu = fmax(fabs(pu) - uh, 0.0), gv = fmax(fabs(pv) - vh, 0.0);
    const double dxy = fma(gv, gv, gu * gu);
    constexpr double reach = 1.0 + 1e-9;
    live[0] = __ballot(fma(c0, c0, dxy) < reach);
    live[1] = __ballot(fma(c1, c1, dxy) < reach);
    live[2] = __ballot(fma(c2, c2, dxy) < reach);
    live[3] = __ballot(fma(c3, c3, dxy) < reach);
  };
  int head = 0, tail = 0;   // ring positions in records (head multiple of 64)
  const int q0 = A.roff[t], qend = A.roff[t + 1];
  for (int qb = q0; qb < qend; qb += kWave) {
    int ms0 = 0, ms1 = 0, msi = 0;
    if (qb + lane < qend) {
      ms0 = start[A.rfirst[qb + lane]]; ms1 = start[A.rend[qb + lane]]; msi = A.rshift[qb + lane];
    }
    const int jn = min(kWave, qend - qb);
    for (int j = 0; j < jn; ++j) {
      const int s0 = __builtin_amdgcn_readlane(ms0, j), s1 = __builtin_amdgcn_readlane(ms1, j);
      const int si = __builtin_amdgcn_readlane(msi, j);
      const double ox = A.sh[si][0] - bx, oy = A.sh[si][1] - by, oz = A.sh[si][2] - bz;
      // the run's records from a wave-uniform base
      const double4 *ra_s = A.rec + s0, *rb_s = A.recB + s0;
      const int last = s1 - 1 - s0;
      // cull, stage and (64 pending) consume one chunk of 64 records
      auto chunk = [&](const double4 &ra, const double4 &rb, bool valid) {
        // shifted position (GetPartPosShifted) relative to the DOF-box centre, scaled
        const double pu = fma(ra.x + ox, Ri, -uc), pv = fma(ra.y + oy, Ri, -vc), pw = fma(ra.z + oz, Ri, -wc);
        const double gx = fmax(fabs(pu) - uh, 0.0);
        const double gy = fmax(fabs(pv) - vh, 0.0);
        const double gz = fmax(fabs(pw) - wh, 0.0);
        const bool pass = valid && (ra.w != 0.0) && (gx * gx + gy * gy + gz * gz <= box2);
        const unsigned long long m = __ballot(pass);
        if (m == 0) return;
        if (pass) {
          double *r = L[(tail + __popcll(m & ((1ull << lane) - 1ull))) & (kRing - 1)];
          r[0] = pu; r[1] = pv; r[2] = pw;
          r[6] = rb.x; r[7] = rb.y; r[8] = rb.z; r[9] = rb.w;
        }
        tail += __popcll(m);
        if (tail - head < kWave) return;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        unsigned long long live[4];
        form_ck(L + (head & (kRing - 1)), live);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int g = 0; g < kWave / 4; ++g) group(L + (head & (kRing - 1)), g, live);
        head += kWave;
        __builtin_amdgcn_wave_barrier();
      };
      // two chunks per pass, both loads issued before either is culled (the
      // second at a clamped index when the run ends within the first)
      for (int base = 0; base <= last; base += 2 * kWave) {
        const int k0 = base + lane, k1 = base + kWave + lane;
        const int c0 = min(k0, last), c1 = min(k1, last);
        const double4 ra0 = rec_at(ra_s, c0), rb0 = rec_at(rb_s, c0);
        const double4 ra1 = rec_at(ra_s, c1), rb1 = rec_at(rb_s, c1);
        chunk(ra0, rb0, k0 <= last);
        if (base + kWave <= last) chunk(ra1, rb1, k1 <= last);
      }
    }
  }
  // the rest, with the last group padded by weightless records far outside
  const int rest = tail - head;
  if (rest > 0) {
    const int ng = (rest + 3) >> 2;
    if (lane >= rest && lane < 4 * ng) {
      double *r = L[(head + lane) & (kRing - 1)];
      r[0] = 0.0; r[1] = 0.0; r[2] = 1e3;
#pragma unroll
      for (int c = 6; c < kRec; ++c) r[c] = 0.0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    unsigned long long live[4];
    form_ck(L + (head & (kRing - 1)), live);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int g = 0; g < ng; ++g) group(L + (head & (kRing - 1)), g, live);
  }
  const int dofLo = 4 * ((lane >> 2) & 3) + (lane >> 4);
#pragma unroll
  for (int Q = 0; Q < 4; ++Q) PS[((size_t)t * nd + 16 * Q + dofLo) * 4 + jc] = acc[Q];
}

template <int ALPHA, bool PERSIST>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE_MX))) void k_sf_gather_cart(DevMesh M, const int32_t *start, SFArgs A, double *PS) {
  __shared__ double lds[kBlock / kWave][kRingCart][kRecCart];
  const int wib = threadIdx.x / kWave;
  const int lane = lane_id();
  if (!PERSIST) {
    const int t = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * (kBlock / kWave) + wib);
    if (t < M.nElems) sf_gather_cart_target<ALPHA>(M, start, A, PS, t, lds[wib], lane);
    return;
  }
  int range = blockIdx.x & 7;
  for (int t = sf_next_target(A, M.nElems, range, lane); t < M.nElems; t = sf_next_target(A, M.nElems, range, lane)) {
    sf_gather_cart_target<ALPHA>(M, start, A, PS, t, lds[wib], lane);
    __builtin_amdgcn_wave_barrier();
  }
}

// Mixed-precision shape-function gather for N = 3 on general elements: S and
// the products in FP32, summed in FP32 on the matrix cores, then added into
// FP64 after every batch of 64 staged particles.  The products run 16
// particles per sequence of v_mfma_f32_4x4x1 (16 blocks of 4x4): operand A
// is f (v,1) of particle p, taken for every block from the lanes 4p..4p+3
// (broadcast by CBSZ = 4, ABID = p), operand B is S^alpha of DOF lane, and
// lane l gathers the four source components of DOF l.  Positions relative to
// the target barycentre and the cull are formed in FP64 from the same
// records and runs as the FP64 gathers.  Records in LDS (FP32): px, py, pz,
// |p|^2/R^2, f (v,1).
typedef float v4f __attribute__((ext_vector_type(4)));
#define SF_MIXED_GSTEP(P)                                                        \
  {                                                                              \
    const float *r_ = Lh[g + (P)];                                               \
    const float S_ = clamp01f(mx * r_[0] + (my * r_[1] + (mz * r_[2] + (cd - r_[3])))); \
    float S1_ = S_ * S_;                                                         \
    if (ALPHA == 0) {                                                            \
      for (int ex = 3; ex <= A.alpha; ++ex) S1_ = S_ * S1_;                      \
    } else {                                                                     \
      _Pragma("unroll") for (int ex = 3; ex <= ALPHA; ++ex) S1_ = S_ * S1_;      \
    }                                                                            \
    c = __builtin_amdgcn_mfma_f32_4x4x1f32(fa, S1_, c, 4, (P), 0);               \
  }
template <int ALPHA>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE_MX))) void k_sf_gather_mixed(DevMesh M, const int32_t *start, SFArgs A, double *PS) {
#pragma clang fp contract(fast)
  constexpr int nd = 64, kRing = 128;
  constexpr int kRec = 8;
  __shared__ float lds[kBlock / kWave][kRing][kRec];
  const int wib = threadIdx.x / kWave;
  const int lane = lane_id();
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int t = __builtin_amdgcn_readfirstlane(lb * (kBlock / kWave) + wib);
  if (t >= M.nElems) return;
  const double bx = M.bary[t * 3], by = M.bary[t * 3 + 1], bz = M.bary[t * 3 + 2];
  const double R2i0 = A.r2_sf_inv;
  const double *xg = M.xGPe + (size_t)t * nd * 3;
  // S = cd - qq + m.q for DOF lane
  const double dx = xg[lane * 3] - bx, dy = xg[lane * 3 + 1] - by, dz = xg[lane * 3 + 2] - bz;
  const float mx = float(2.0 * R2i0 * dx), my = float(2.0 * R2i0 * dy), mz = float(2.0 * R2i0 * dz);
  const float cd = float(1.0 - R2i0 * (dx * dx + dy * dy + dz * dz));
  double blo[3] = {dx, dy, dz}, bhi[3] = {dx, dy, dz};
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      blo[c] = fmin(blo[c], __shfl_xor(blo[c], off, kWave));
      bhi[c] = fmax(bhi[c], __shfl_xor(bhi[c], off, kWave));
    }
  // the reference's element cull |x - bary| > r_sf + ElemRadius, compared squared
  const double thr = A.r_sf + M.radius[t], thr2 = thr * thr;
  const double box2 = A.r2_sf * (1.0 + 1e-12);
  float (*L)[kRec] = lds[wib];
  // ring rows past the last record of a batch are read with zero weight: keep them finite
#pragma unroll
  for (int q = 0; q < 2 * kRec; ++q) L[2 * lane + q / kRec][q % kRec] = 0.f;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // the batch of n staged records from ring position h0 (no wrap)
  auto batch = [&](int h0, int n) {
    const float (*Lh)[kRec] = L + (h0 & (kRing - 1));
    v4f c = {0.f, 0.f, 0.f, 0.f};
    for (int g = 0; g < n; g += 16) {
      const int pf = g + (lane >> 2);
      const float fa = pf < n ? Lh[pf][4 + (lane & 3)] : 0.f;
      SF_MIXED_GSTEP(0) SF_MIXED_GSTEP(1) SF_MIXED_GSTEP(2) SF_MIXED_GSTEP(3)
      SF_MIXED_GSTEP(4) SF_MIXED_GSTEP(5) SF_MIXED_GSTEP(6) SF_MIXED_GSTEP(7)
      SF_MIXED_GSTEP(8) SF_MIXED_GSTEP(9) SF_MIXED_GSTEP(10) SF_MIXED_GSTEP(11)
      SF_MIXED_GSTEP(12) SF_MIXED_GSTEP(13) SF_MIXED_GSTEP(14) SF_MIXED_GSTEP(15)
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += double(c[q]);
  };
  int head = 0, tail = 0;   // ring positions in records (head multiple of 64)
  const int q0 = A.roff[t], qend = A.roff[t + 1];
  for (int qb = q0; qb < qend; qb += kWave) {
    int ms0 = 0, ms1 = 0, msi = 0;
    if (qb + lane < qend) {
      ms0 = start[A.rfirst[qb + lane]]; ms1 = start[A.rend[qb + lane]]; msi = A.rshift[qb + lane];
    }
    const int jn = min(kWave, qend - qb);
    for (int j = 0; j < jn; ++j) {
      const int s0 = __builtin_amdgcn_readlane(ms0, j), s1 = __builtin_amdgcn_readlane(ms1, j);
      const int si = __builtin_amdgcn_readlane(msi, j);
      const double ox = A.sh[si][0] - bx, oy = A.sh[si][1] - by, oz = A.sh[si][2] - bz;
      // the run's records from a wave-uniform base
      const double4 *ra_s = A.rec + s0, *rb_s = A.recB + s0;
      const int last = s1 - 1 - s0;
      for (int base = 0; base <= last; base += kWave) {
        const int k = base + lane;
        const bool valid = k <= last;
        const int c = valid ? k : last;
        const double4 ra = rec_at(ra_s, c), rb = rec_at(rb_s, c);
        const double px = ra.x + ox, py = ra.y + oy, pz = ra.z + oz;
        const double gx = fmax(0.0, fmax(blo[0] - px, px - bhi[0]));
        const double gy = fmax(0.0, fmax(blo[1] - py, py - bhi[1]));
        const double gz = fmax(0.0, fmax(blo[2] - pz, pz - bhi[2]));
        bool pass = valid && (ra.w != 0.0) && (gx * gx + gy * gy + gz * gz <= box2);
        if (!A.linear) pass = pass && !(px * px + py * py + pz * pz > thr2);
        const unsigned long long m = __ballot(pass);
        if (m == 0) continue;
        if (pass) {
          float *r = L[(tail + __popcll(m & ((1ull << lane) - 1ull))) & (kRing - 1)];
          r[0] = float(px); r[1] = float(py); r[2] = float(pz);
          r[3] = float(R2i0 * (px * px + py * py + pz * pz));
          r[4] = float(rb.x); r[5] = float(rb.y); r[6] = float(rb.z); r[7] = float(rb.w);
        }
        tail += __popcll(m);
        if (tail - head < kWave) continue;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        batch(head, kWave);
        head += kWave;
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  if (tail > head) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    batch(head, tail - head);
  }
  double *out = PS + ((size_t)t * nd + lane) * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = acc[q];
}
#undef SF_MIXED_GSTEP

// Mixed-precision gather on axis-aligned tensor grids (N = 3).  The cull and
// the ring follow k_sf_gather_cart; the ring keeps each survivor's position
// (FP64, relative to the DOF-box centre, scaled by 1/r_sf) and f (v,1) in
// FP32.  Once 64 are pending, one pass with a survivor per lane forms the 16
// (i,j) parts 1 - (u_i - pu)^2 - (v_j - pv)^2 and the 4 k parts
// (w_k - pw)^2 of S in FP32.  The products then run 16 particles per
// sequence of v_mfma_f32_4x4x1 (16 blocks): operand A is f (v,1) of the
// particle, taken for every block from the lanes 4p..4p+3 that hold
// particle p (broadcast by the instruction's CBSZ = 4, ABID = p), operand B
// is S^alpha of DOF lane, and lane l gathers the four source components of
// DOF l in its four result registers.  A pair costs two LDS reads, the
// subtraction with the clamp and the square; the FP32 sums of each batch of
// 64 are added into FP64.
#define SF_MIXED_STEP(P)                                                         \
  {                                                                              \
    const float *o_ = Pr[g + (P)];                                               \
    const float S_ = clamp01f(o_[wq] - o_[16 + zk]);                             \
    float S1_ = S_ * S_;                                                         \
    if (ALPHA == 0) {                                                            \
      for (int ex = 3; ex <= A.alpha; ++ex) S1_ = S_ * S1_;                      \
    } else {                                                                     \
      _Pragma("unroll") for (int ex = 3; ex <= ALPHA; ++ex) S1_ = S_ * S1_;      \
    }                                                                            \
    c = __builtin_amdgcn_mfma_f32_4x4x1f32(fa, S1_, c, 4, (P), 0);               \
  }
template <int ALPHA>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WPE_MX))) void k_sf_gather_mixed_cart(DevMesh M, const int32_t *start, SFArgs A, double *PS) {
#pragma clang fp contract(fast)
  constexpr int nd = 64, kRing = 128;
  __shared__ double ringP[kBlock / kWave][kRing][3];
  __shared__ float ringF[kBlock / kWave][kRing][4];
  __shared__ float prep[kBlock / kWave][kWave][20];
  const int wib = threadIdx.x / kWave;
  const int lane = lane_id();
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int t = __builtin_amdgcn_readfirstlane(lb * (kBlock / kWave) + wib);
  if (t >= M.nElems) return;
  const double bx = M.bary[t * 3], by = M.bary[t * 3 + 1], bz = M.bary[t * 3 + 2];
  const double Ri = A.r_inv;
  const double *xg = M.xGPe + (size_t)t * nd * 3;
  double u[4], v[4], w[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    u[a] = (xg[a * 3 + 0] - bx) * Ri;
    v[a] = (xg[4 * a * 3 + 1] - by) * Ri;
    w[a] = (xg[16 * a * 3 + 2] - bz) * Ri;
  }
  const double ulo = fmin(fmin(u[0], u[1]), fmin(u[2], u[3])), uhi = fmax(fmax(u[0], u[1]), fmax(u[2], u[3]));
  const double vlo = fmin(fmin(v[0], v[1]), fmin(v[2], v[3])), vhi = fmax(fmax(v[0], v[1]), fmax(v[2], v[3]));
  const double wlo = fmin(fmin(w[0], w[1]), fmin(w[2], w[3])), whi = fmax(fmax(w[0], w[1]), fmax(w[2], w[3]));
  const double uc = 0.5 * (ulo + uhi), vc = 0.5 * (vlo + vhi), wc = 0.5 * (wlo + whi);
  const double uh = 0.5 * (uhi - ulo), vh = 0.5 * (vhi - vlo), wh = 0.5 * (whi - wlo);
#pragma unroll
  for (int a = 0; a < 4; ++a) { u[a] -= uc; v[a] -= vc; w[a] -= wc; }
  const double box2 = 1.0 + 1e-12;
  double (*RP)[3] = ringP[wib];
  float (*RF)[4] = ringF[wib];
  float (*Pr)[20] = prep[wib];
  // rows past the last record of a batch are read with zero weight: keep them finite
#pragma unroll
  for (int q = 0; q < 20; ++q) Pr[lane][q] = 0.f;
  const int wq = lane & 15, zk = lane >> 4;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // the batch of n <= 64 ring records from position h0 (no wrap)
  auto batch = [&](int h0, int n) {
    if (lane < n) {
      const double *p = RP[(h0 + lane) & (kRing - 1)];
      float *o = Pr[lane];
      float X[4], Y[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float a = float(u[q] - p[0]), b = float(v[q] - p[1]), cz = float(w[q] - p[2]);
        X[q] = 1.f - a * a;
        Y[q] = b * b;
        o[16 + q] = cz * cz;
      }
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) o[ii + 4 * jj] = X[ii] - Y[jj];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    v4f c = {0.f, 0.f, 0.f, 0.f};
    for (int g = 0; g < n; g += 16) {
      const int pf = g + (lane >> 2);
      const float fa = pf < n ? RF[(h0 + pf) & (kRing - 1)][lane & 3] : 0.f;
      SF_MIXED_STEP(0) SF_MIXED_STEP(1) SF_MIXED_STEP(2) SF_MIXED_STEP(3)
      SF_MIXED_STEP(4) SF_MIXED_STEP(5) SF_MIXED_STEP(6) SF_MIXED_STEP(7)
      SF_MIXED_STEP(8) SF_MIXED_STEP(9) SF_MIXED_STEP(10) SF_MIXED_STEP(11)
      SF_MIXED_STEP(12) SF_MIXED_STEP(13) SF_MIXED_STEP(14) SF_MIXED_STEP(15)
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += double(c[q]);
    __builtin_amdgcn_wave_barrier();
  };
  int head = 0, tail = 0;   // ring positions in records (head multiple of 64)
  const int q0 = A.roff[t], qend = A.roff[t + 1];
  for (int qb = q0; qb < qend; qb += kWave) {
    int ms0 = 0, ms1 = 0, msi = 0;
    if (qb + lane < qend) {
      ms0 = start[A.rfirst[qb + lane]]; ms1 = start[A.rend[qb + lane]]; msi = A.rshift[qb + lane];
    }
    const int jn = min(kWave, qend - qb);
    for (int j = 0; j < jn; ++j) {
      const int s0 = __builtin_amdgcn_readlane(ms0, j), s1 = __builtin_amdgcn_readlane(ms1, j);
      const int si = __builtin_amdgcn_readlane(msi, j);
      const double ox = A.sh[si][0] - bx, oy = A.sh[si][1] - by, oz = A.sh[si][2] - bz;
      const double4 *ra_s = A.rec + s0, *rb_s = A.recB + s0;
      const int last = s1 - 1 - s0;
      for (int base = 0; base <= last; base += kWave) {
        const int k = base + lane;
        const bool valid = k <= last;
        const int c = valid ? k : last;
        const double4 ra = rec_at(ra_s, c), rb = rec_at(rb_s, c);
        const double pu = fma(ra.x + ox, Ri, -uc), pv = fma(ra.y + oy, Ri, -vc), pw = fma(ra.z + oz, Ri, -wc);
        const double gx = fmax(fabs(pu) - uh, 0.0);
        const double gy = fmax(fabs(pv) - vh, 0.0);
        const double gz = fmax(fabs(pw) - wh, 0.0);
        const bool pass = valid && (ra.w != 0.0) && (gx * gx + gy * gy + gz * gz <= box2);
        const unsigned long long m = __ballot(pass);
        if (m == 0) continue;
        if (pass) {
          const int slot = (tail + __popcll(m & ((1ull << lane) - 1ull))) & (kRing - 1);
          RP[slot][0] = pu; RP[slot][1] = pv; RP[slot][2] = pw;
          RF[slot][0] = float(rb.x); RF[slot][1] = float(rb.y); RF[slot][2] = float(rb.z); RF[slot][3] = float(rb.w);
        }
        tail += __popcll(m);
        if (tail - head < kWave) continue;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        batch(head, kWave);
        head += kWave;
      }
    }
  }
  if (tail > head) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    batch(head, tail - head);
  }
  double *out = PS + ((size_t)t * nd + lane) * 4;
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = acc[q];
}
#undef SF_MIXED_STEP

// ---- analysis integrals ------------------------------------------------------
// Per element, in the reference's i-fastest DOF order:
//   mode 0  sum_ijk (w_i w_j w_k * rho) * (1/sJ)       CalcDepositedCharge, pic_analyze.f90:172-179
//   mode 1  sum_ijk  w_i w_j w_k / sJ * |E|^2           CalcPotentialEnergy, analyzefield_tools.f90:191,209
struct QuadArgs {
  double wGP[kMaxN + 1];
};
template <int N>
__global__ __launch_bounds__(kBlock) void k_elem_integral(int nElems, int mode, const double *sJ, const double *src,
                                                          QuadArgs Q, double *elemPart) {
  constexpr int n1 = N + 1, nd = n1 * n1 * n1;
  const int e = blockIdx.x * kBlock + threadIdx.x;
  if (e >= nElems) return;
  double acc = 0.0;
  for (int k = 0; k < n1; ++k)
    for (int j = 0; j < n1; ++j)
#pragma unroll
      for (int i = 0; i < n1; ++i) {
        const int r = (k * n1 + j) * n1 + i;
        const size_t g = size_t(e) * nd + r;
        const double wijk = Q.wGP[i] * Q.wGP[j] * Q.wGP[k];
        if (mode == 0) {
          acc = acc + wijk * src[g * 4 + 3] * (1. / sJ[g]);
        } else {
          const double *E = src + g * 3;
          const double Eabs = E[0] * E[0] + E[1] * E[1] + E[2] * E[2];
          acc = acc + wijk / sJ[g] * Eabs;
        }
      }
  elemPart[e] = acc;
}

// Sum of n values in a fixed order (strided per-thread sums, then a tree in
// LDS): one block of 1024 threads, so the result does not depend on timing.
__global__ __launch_bounds__(1024) void k_sum_fixed(int64_t n, const double *v, double *out) {
  __shared__ double red[1024];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int64_t i = t; i < n; i += 1024) s = s + v[i];
  red[t] = s;
  __syncthreads();
  for (int w = 512; w >= 1; w >>= 1) {
    if (t < w) red[t] = red[t] + red[t + w];
    __syncthreads();
  }
  if (t == 0) *out = red[0];
}

// Particles per species among the n element-sorted particles (all inside).
__global__ __launch_bounds__(kBlock) void k_species_count(int64_t n, const uint8_t *spec,
                                                          unsigned long long *cnt) {
  __shared__ unsigned int c[kMaxSpecies];
  if (threadIdx.x < kMaxSpecies) c[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBlock)
    atomicAdd(&c[spec[i]], 1u);
  __syncthreads();
  if (threadIdx.x < kMaxSpecies && c[threadIdx.x]) atomicAdd(&cnt[threadIdx.x], (unsigned long long)c[threadIdx.x]);
}

__global__ void k_to_f32(const double *a, float *b, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) b[i] = float(a[i]);
}
__global__ void k_fill_u8(uint8_t *p, int64_t n, uint8_t v) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double *p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = 0.0;
}
__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}
// element id of each sorted slot, from the segment offsets
__global__ void k_elem_of(int nElems, const int32_t *start, int32_t *elemOf) {
  const int w = (blockIdx.x * kBlock + threadIdx.x) / kWave;
  const int lane = lane_id();
  if (w >= nElems) return;
  for (int i = start[w] + lane; i < start[w + 1]; i += kWave) elemOf[i] = w;
}

}  // namespace

namespace {
struct StepState {       // a stored particle state, sorted by element
  double *x = nullptr;   // [6][cap]: x y z vx vy vz
  int32_t *elem = nullptr, *id = nullptr, *start = nullptr;
  uint8_t *spec = nullptr;
  double *E = nullptr;   // field of the step [nElems][nDof][3]
  int64_t n = 0;
};
// self-consistent field operator and adjoint buffers
struct Adjoint {
  double *G = nullptr, *GT = nullptr, *rbar = nullptr, *lrbar = nullptr, *En = nullptr, *lEn = nullptr, *lE = nullptr;
  double *lx = nullptr, *lv = nullptr, *lF = nullptr, *xi = nullptr;
  int32_t *elemNode = nullptr, *nodeOff = nullptr, *nodeInc = nullptr, *srcOff = nullptr, *tgt = nullptr;
  uint8_t *tgtShift = nullptr;
  std::vector<StepState> ck, seg;
};
}  // namespace

// ======================================================================= handle
struct pic_handle {
  pic_mesh_desc meshDesc{};
  pic_params par{};
  HostMesh hm;
  DevMesh M{};
  hipStream_t stream = nullptr;
  int device = 0;
  int64_t cap = 0;      // particle capacity
  int64_t n = 0;        // particles in the sorted arrays
  int64_t nTotal = 0;   // caller's particle count (ids 0..nTotal-1)
  bool particlesSet = false;   // pic_set_particles (or an emission / restart) was called, possibly with none
  // device buffers
  std::vector<void *> allocs;
  DevParts P{}, Q{};    // current and scratch particle arrays
  int32_t *start = nullptr, *startNew = nullptr, *newElem = nullptr, *rank = nullptr, *perm = nullptr;
  int32_t *stayCnt = nullptr, *arrCnt = nullptr, *elemOf = nullptr, *dTotal = nullptr;
  uint8_t *insideById = nullptr;
  double *stateById = nullptr;
  unsigned long long *lostCount = nullptr;
  double *E = nullptr, *PS = nullptr, *elemCorner = nullptr, *classSrc = nullptr;
  float *E32 = nullptr;   // mixed precision copy of E
  int32_t *elemClass = nullptr, *classOff = nullptr, *classInc = nullptr;
  double *classVol = nullptr;
  double *sJ = nullptr, *elemPart = nullptr, *sumOut = nullptr;
  unsigned long long *specCnt = nullptr;
  SFArgs sf{};
  Adjoint adj;
  bool sfMatrix = true;   // N = 3 shape function: accumulate on the matrix cores
  bool recFresh = false;  // sf.rec holds the records of the current particle arrays
  bool useGraph = false;  // replay captured two-step graphs in pic_time_step
  bool sfPersist = false; // N = 3 matrix-core gathers: persistent waves that take targets from counters
  bool sfCoop = false;    // general matrix-core gather: a workgroup's four targets share their record loads
  int nCU = 0;            // compute units of the device
  std::vector<std::pair<const void *, unsigned>> persistGrid;   // resident blocks per persistent kernel
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  double graphDt = 0;
  int64_t graphN = 0;
  const double *graphP = nullptr;
  Movers mv{};
  bool stagedValid = false;   // LastPartPos/FieldAtParticle refer to the current order
  std::vector<std::pair<int64_t, std::vector<double>>> unplaced;  // particles no element holds at set-up
  // statistics
  hipEvent_t ev[11] = {};   // ev[9], ev[10]: around the push kernel (kernel A)
  double msDeposit = 0, msPushTrack = 0, msSort = 0, msDepositMain = 0, msPushMain = 0;
  int nDepositLaunches = 0;
  int64_t nSteps = 0;
};

namespace {

// Blocks of a persistent kernel that the device holds at once (its
// occupancy per CU times the CUs, a multiple of 8 so that every XCD range of
// the target counters starts with its own blocks), found once per kernel.
unsigned persist_grid(pic_handle *h, const void *kernel) {
  for (auto &pg : h->persistGrid)
    if (pg.first == kernel) return pg.second;
  int perCU = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, kernel, kBlock, 0) != hipSuccess || perCU < 1) perCU = 1;
  const unsigned g = std::max(8u, unsigned(perCU * h->nCU) / 8u * 8u);
  h->persistGrid.emplace_back(kernel, g);
  return g;
}

// the record block boxes, written with the records where the general-element
// gather reads them (null on tensor grids, whose gather does not)
inline double4 *block_boxes(pic_handle *h) {
  return h->sf.cartesian ? nullptr : const_cast<double4 *>(h->sf.blk);
}

template <typename T>
int dalloc(pic_handle *h, T **p, size_t count) {
  void *q = nullptr;
  HIPCHK(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
  h->allocs.push_back(q);
  *p = static_cast<T *>(q);
  return 0;
}
template <typename T>
int upload(pic_handle *h, T **p, const std::vector<T> &v) {
  if (int rc = dalloc(h, p, v.size())) return rc;
  if (!v.empty()) HIPCHK(copy_bytes(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

int alloc_parts(pic_handle *h, DevParts &P, int64_t cap) {
  int rc = 0;
  rc |= dalloc(h, &P.x, cap); rc |= dalloc(h, &P.y, cap); rc |= dalloc(h, &P.z, cap);
  rc |= dalloc(h, &P.vx, cap); rc |= dalloc(h, &P.vy, cap); rc |= dalloc(h, &P.vz, cap);
  rc |= dalloc(h, &P.elem, cap); rc |= dalloc(h, &P.id, cap);
  rc |= dalloc(h, &P.spec, cap); rc |= dalloc(h, &P.isNew, cap);
  return rc;
}

#define LAUNCH_N(KERNEL, N, GRID, ...)                                                   \
  switch (N) {                                                                           \
    case 1: hipLaunchKernelGGL(KERNEL<1>, GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 2: hipLaunchKernelGGL(KERNEL<2>, GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 3: hipLaunchKernelGGL(KERNEL<3>, GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 4: hipLaunchKernelGGL(KERNEL<4>, GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 5: hipLaunchKernelGGL(KERNEL<5>, GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 6: hipLaunchKernelGGL(KERNEL<6>, GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 7: hipLaunchKernelGGL(KERNEL<7>, GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    default: break;                                                                      \
  }

#define LAUNCH_N2(KERNEL, N, A2, GRID, ...)                                                        \
  switch (N) {                                                                                      \
    case 1: hipLaunchKernelGGL((KERNEL<1, A2>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 2: hipLaunchKernelGGL((KERNEL<2, A2>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 3: hipLaunchKernelGGL((KERNEL<3, A2>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 4: hipLaunchKernelGGL((KERNEL<4, A2>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 5: hipLaunchKernelGGL((KERNEL<5, A2>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 6: hipLaunchKernelGGL((KERNEL<6, A2>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    case 7: hipLaunchKernelGGL((KERNEL<7, A2>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break; \
    default: break;                                                                                 \
  }

#define LAUNCH_NG_N(KERNEL, N, G, GRID, ...)                                                                  \
  switch (N) {                                                                                                 \
    case 1: hipLaunchKernelGGL((KERNEL<1, G>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break;           \
    case 2: hipLaunchKernelGGL((KERNEL<2, G>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break;           \
    case 3: hipLaunchKernelGGL((KERNEL<3, G>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break;           \
    case 4: hipLaunchKernelGGL((KERNEL<4, G>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break;           \
    case 5: hipLaunchKernelGGL((KERNEL<5, G>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break;           \
    case 6: hipLaunchKernelGGL((KERNEL<6, G>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break;           \
    case 7: hipLaunchKernelGGL((KERNEL<7, G>), GRID, dim3(kBlock), 0, h->stream, __VA_ARGS__); break;           \
    default: break;                                                                                            \
  }
// kernels templated on the DG degree and the geometry points per direction
#define LAUNCH_NG(KERNEL, N, NGEO, GRID, ...)                   \
  do {                                                          \
    if ((NGEO) == 1) { LAUNCH_NG_N(KERNEL, N, 2, GRID, __VA_ARGS__) } \
    else { LAUNCH_NG_N(KERNEL, N, 3, GRID, __VA_ARGS__) }       \
  } while (0)

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, a, b);
  return ms;
}

int set_device(pic_handle *h) {
  HIPCHK(hipSetDevice(h->device));
  return 0;
}

void drop_graph(pic_handle *h);
int emit_scratch_init(int device);
int set_particles_impl(pic_handle *h, int64_t n, const double *PartState, const int32_t *elem, const int32_t *species,
                       const uint8_t *is_new, bool outsideMarked);

// the dominant kernel run_deposition launches (PIC_KERNEL_*)
int deposit_kernel(const pic_handle *h) {
  if (h->par.deposition == PIC_DEPO_NONE) return PIC_KERNEL_NONE;
  if (h->par.deposition == PIC_DEPO_CVWM) return PIC_KERNEL_CVWM;
  if (h->hm.N == 3 && h->par.precision == PIC_PREC_MIXED) return PIC_KERNEL_SF_MIXED;
  if (h->hm.N == 3 && h->sfMatrix) return h->sf.cartesian ? PIC_KERNEL_SF_CARTESIAN : PIC_KERNEL_SF_MATRIX;
  return PIC_KERNEL_SF_VALU;
}

// Deposition() on the current sorted particles
int run_deposition(pic_handle *h, bool timed) {
  const int ne = h->hm.nElems, N = h->hm.N;
  const size_t nps = size_t(ne) * h->hm.nDof * 4;
  if (timed) HIPCHK(hipEventRecord(h->ev[0], h->stream));
  // ev[8] marks the start of the dominant kernel (after the record pass of the
  // matrix-core gather)
  const bool mixed = h->par.precision == PIC_PREC_MIXED;
  const bool records = h->par.deposition == PIC_DEPO_SF && N == 3 && (h->sfMatrix || mixed);
  if (timed && !records) HIPCHK(hipEventRecord(h->ev[8], h->stream));
  if (h->par.deposition == PIC_DEPO_NONE) {
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(int64_t(nps))), dim3(kBlock), 0, h->stream, h->PS, int64_t(nps));
  } else if (h->par.deposition == PIC_DEPO_CVWM) {
    LAUNCH_N(k_cvwm_elem, N, dim3(grid_for(int64_t(ne) * kWave)), h->M, h->P, h->start, h->elemCorner);
    if (timed) HIPCHK(hipEventRecord(h->ev[1], h->stream));
    hipLaunchKernelGGL(k_cvwm_class, dim3(grid_for(h->hm.nClasses)), dim3(kBlock), 0, h->stream, h->hm.nClasses,
                       h->classOff, h->classInc, h->classVol, h->elemCorner, h->classSrc);
    LAUNCH_N(k_cvwm_dof, N, dim3(grid_for(int64_t(ne) * h->hm.nDof)), h->M, h->elemClass, h->classSrc, h->PS);
  } else {
    const int waves = kBlock / kWave;
    const unsigned nblk = xcd_grid((ne + waves - 1) / waves);
    if (records) {
      // the element sort writes the records; after other changes of the
      // particle arrays they are formed here
      if (!h->recFresh)
        hipLaunchKernelGGL(k_sf_records, dim3(grid_for(std::max<int64_t>(h->n, 1))), dim3(kBlock), 0, h->stream, h->P,
                           h->start + ne, h->sf, const_cast<double4 *>(h->sf.rec), const_cast<double4 *>(h->sf.recB),
                           block_boxes(h));
      h->recFresh = true;
      if (timed) HIPCHK(hipEventRecord(h->ev[8], h->stream));
    }
    if (N == 3 && mixed) {
      const bool cart = h->sf.cartesian != 0;
      if (h->sf.alpha == 2 && cart)
        hipLaunchKernelGGL((k_sf_gather_mixed_cart<2>), dim3(nblk), dim3(kBlock), 0, h->stream, h->M, h->start, h->sf, h->PS);
      else if (h->sf.alpha == 2)
        hipLaunchKernelGGL((k_sf_gather_mixed<2>), dim3(nblk), dim3(kBlock), 0, h->stream, h->M, h->start, h->sf, h->PS);
      else if (cart)
        hipLaunchKernelGGL((k_sf_gather_mixed_cart<0>), dim3(nblk), dim3(kBlock), 0, h->stream, h->M, h->start, h->sf, h->PS);
      else
        hipLaunchKernelGGL((k_sf_gather_mixed<0>), dim3(nblk), dim3(kBlock), 0, h->stream, h->M, h->start, h->sf, h->PS);
    } else if (N == 3 && h->sfMatrix) {
      const bool lin = h->sf.linear != 0;
      // persistent waves: the counters start at zero (a fill kernel, which a
      // captured graph replays), the grid is what the device holds at once
      if (h->sfPersist)
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kBlock), 0, h->stream, h->sf.work, int64_t(8), 0);
#define SF_GATHER(K, ...)                                                                                    \
  do {                                                                                                     \
    if (h->sfPersist) {                                                                                    \
      auto kp = K<__VA_ARGS__, true>;                                                                      \
      hipLaunchKernelGGL(kp, dim3(persist_grid(h, (const void *)kp)), dim3(kBlock), 0, h->stream, h->M,    \
                         h->start, h->sf, h->PS);                                                          \
    } else {                                                                                               \
      auto kp = K<__VA_ARGS__, false>;                                                                     \
      hipLaunchKernelGGL(kp, dim3(nblk), dim3(kBlock), 0, h->stream, h->M, h->start, h->sf, h->PS);         \
    }                                                                                                      \
  } while (0)
      if (h->sfCoop && !h->sf.cartesian) {
        if (h->sf.alpha == 2 && lin)
          hipLaunchKernelGGL((k_sf_gather_mfma_coop<2, true>), dim3(nblk), dim3(kBlock), 0, h->stream, h->M, h->start, h->sf, h->PS);
        else if (h->sf.alpha == 2)
          hipLaunchKernelGGL((k_sf_gather_mfma_coop<2, false>), dim3(nblk), dim3(kBlock), 0, h->stream, h->M, h->start, h->sf, h->PS);
        else if (lin)
          hipLaunchKernelGGL((k_sf_gather_mfma_coop<0, true>), dim3(nblk), dim3(kBlock), 0, h->stream, h->M, h->start, h->sf, h->PS);
        else
          hipLaunchKernelGGL((k_sf_gather_mfma_coop<0, false>), dim3(nblk), dim3(kBlock), 0, h->stream, h->M, h->start, h->sf, h->PS);
      } else if (h->sf.cartesian && h->sf.alpha == 2)
        SF_GATHER(k_sf_gather_cart, 2);
      else if (h->sf.cartesian)
        SF_GATHER(k_sf_gather_cart, 0);
      else if (h->sf.alpha == 2 && lin)
        SF_GATHER(k_sf_gather_mfma, 2, true);
      else if (h->sf.alpha == 2)
        SF_GATHER(k_sf_gather_mfma, 2, false);
      else if (lin)
        SF_GATHER(k_sf_gather_mfma, 0, true);
      else
        SF_GATHER(k_sf_gather_mfma, 0, false);
#undef SF_GATHER
    } else if (h->sf.alpha == 2) {
      LAUNCH_N2(k_sf_gather, N, 2, dim3(nblk), h->M, h->P, h->start, h->sf, h->PS);
    } else {
      LAUNCH_N2(k_sf_gather, N, 0, dim3(nblk), h->M, h->P, h->start, h->sf, h->PS);
    }
    if (timed) HIPCHK(hipEventRecord(h->ev[1], h->stream));
  }
  HIPCHK(hipGetLastError());
  if (timed) HIPCHK(hipEventRecord(h->ev[2], h->stream));
  return 0;
}

// element re-sort after tracking; newElem[i] holds the owner (-1 = gone)
// Enqueues the re-sort without waiting on the host: grids cover nGrid
// particles (>= the count), kernels bound themselves by the count that the
// old segment table holds on the device.  Pointer swaps happen on the host.
int enqueue_sort(pic_handle *h, int64_t nGrid, bool withStaged) {
  const int ne = h->hm.nElems;
  const int64_t n = std::max<int64_t>(nGrid, 1);   // the kernels bound themselves by the device count
  const int32_t *nOld = h->start + ne;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(ne)), dim3(kBlock), 0, h->stream, h->arrCnt, int64_t(ne), 0);
  hipLaunchKernelGGL(k_elem_of, dim3(grid_for(int64_t(ne) * kWave)), dim3(kBlock), 0, h->stream, ne, h->start, h->elemOf);
  hipLaunchKernelGGL(k_seg_rank, dim3(grid_for(int64_t(ne) * kWave)), dim3(kBlock), 0, h->stream, ne, h->start,
                     h->newElem, h->P.z, h->M.xGPe, h->hm.nDof, h->stayCnt, h->arrCnt, h->rank);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, h->stream, ne, h->stayCnt, h->arrCnt, h->startNew);
  hipLaunchKernelGGL(k_dest, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, n, h->elemOf, h->newElem, h->rank,
                     h->startNew, h->stayCnt, h->perm, nOld);
  hipLaunchKernelGGL(k_fix_arrivals, dim3(grid_for(ne)), dim3(kBlock), 0, h->stream, ne, h->startNew, h->stayCnt,
                     h->arrCnt, h->P.id, h->perm);
  const int32_t *total = h->startNew + ne;
  // records for the shape-function gathers that read them
  const bool useRec = h->sf.rec && (h->sfMatrix || h->par.precision == PIC_PREC_MIXED);
  SpeciesFactor sfac{};
  for (int k = 0; k < 8; ++k) sfac.fsp[k] = h->sf.fsp[k];
  hipLaunchKernelGGL(k_gather_state, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, total, h->perm, h->P, h->Q, h->newElem,
                     sfac, useRec ? const_cast<double4 *>(h->sf.rec) : nullptr,
                     useRec ? const_cast<double4 *>(h->sf.recB) : nullptr,
                     useRec ? block_boxes(h) : nullptr);
  if (withStaged) {
    // keep LastPartPos / LastElem / FieldAtParticle in the new order
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, total, h->perm, h->P.lx, h->Q.lx, 1);
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, total, h->perm, h->P.ly, h->Q.ly, 1);
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, total, h->perm, h->P.lz, h->Q.lz, 1);
    hipLaunchKernelGGL(k_gather<int32_t>, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, total, h->perm, h->P.lastElem,
                       h->Q.lastElem, 1);
    hipLaunchKernelGGL(k_gather<double>, dim3(grid_for(n * 6)), dim3(kBlock), 0, h->stream, total, h->perm, h->P.F, h->Q.F, 6);
    std::swap(h->P.lx, h->Q.lx); std::swap(h->P.ly, h->Q.ly); std::swap(h->P.lz, h->Q.lz);
    std::swap(h->P.lastElem, h->Q.lastElem); std::swap(h->P.F, h->Q.F);
  }
  HIPCHK(hipGetLastError());
  std::swap(h->P.x, h->Q.x); std::swap(h->P.y, h->Q.y); std::swap(h->P.z, h->Q.z);
  std::swap(h->P.vx, h->Q.vx); std::swap(h->P.vy, h->Q.vy); std::swap(h->P.vz, h->Q.vz);
  std::swap(h->P.elem, h->Q.elem); std::swap(h->P.id, h->Q.id);
  std::swap(h->P.spec, h->Q.spec); std::swap(h->P.isNew, h->Q.isNew);
  std::swap(h->P.rx, h->Q.rx); std::swap(h->P.ry, h->Q.ry); std::swap(h->P.rz, h->Q.rz);
  std::swap(h->start, h->startNew);
  h->recFresh = useRec;   // the gather wrote the records of the new order
  return 0;
}

// particle count after the last enqueued sort (synchronises the stream)
int read_count(pic_handle *h) {
  int32_t total_h = 0;
  HIPCHK(hipMemcpyAsync(&total_h, h->start + h->hm.nElems, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->n = total_h;
  return 0;
}

int run_sort(pic_handle *h, bool withStaged) {
  if (int rc = enqueue_sort(h, h->n, withStaged)) return rc;
  return read_count(h);
}

int require_particles(pic_handle *h) {
  if (!h->particlesSet) return fail(PIC_ESTATE, "no particles set");
  return 0;
}

}  // namespace

// ========================================================================= C ABI
extern "C" {

const char *pic_last_error(void) { return g_err.c_str(); }

int pic_create(const pic_mesh_desc *mesh, const pic_params *par, pic_handle **out) {
  if (!mesh || !par || !out) return fail(PIC_EINVAL, "pic_create: null argument");
  *out = nullptr;
  if (par->tracking != PIC_TRACK_TRIA && par->tracking != PIC_TRACK_REFMAPPING)
    return fail(PIC_EINVAL, "pic_create: tracking must be TriaTracking or RefMapping");
  if (par->tracking == PIC_TRACK_REFMAPPING && par->deposition == PIC_DEPO_CVWM)  // pic_depo_method.f90:127-129
    return fail(PIC_EINVAL, "pic_create: cell_volweight_mean needs TriaTracking");
  if (par->deposition < PIC_DEPO_NONE || par->deposition > PIC_DEPO_SF) return fail(PIC_EINVAL, "pic_create: deposition type");
  if (par->nSpecies < 1 || par->nSpecies > kMaxSpecies) return fail(PIC_EINVAL, "pic_create: nSpecies must be 1..8");
  if (par->refmapping_guess < 1 || par->refmapping_guess > 4 ||
      (par->tracking == PIC_TRACK_TRIA && par->refmapping_guess == 3))
    return fail(PIC_EINVAL, "pic_create: RefMappingGuess must be 1..4 (1, 2 or 4 with TriaTracking)");
  if (par->deposition == PIC_DEPO_SF && par->sf_alpha < 2) return fail(PIC_EINVAL, "pic_create: PIC-shapefunction-alpha must be >= 2");
  if (par->max_particles < 1 || par->max_particles > (int64_t(1) << 31) - 1)
    return fail(PIC_EINVAL, "pic_create: max_particles must be in 1..2^31-1");
  if (par->precision != PIC_PREC_FP64 && par->precision != PIC_PREC_MIXED)
    return fail(PIC_EINVAL, "pic_create: precision must be PIC_PREC_FP64 or PIC_PREC_MIXED");
  if (par->precision == PIC_PREC_MIXED && !(par->deposition == PIC_DEPO_SF && mesh->N == 3))
    return fail(PIC_EINVAL, "pic_create: mixed precision is implemented for shape_function deposition with N = 3");
  pic_handle *h = new pic_handle();
  h->meshDesc = *mesh;
  h->par = *par;
  h->device = par->device;
  std::string err = build_host_mesh(*mesh, *par, h->hm);
  if (!err.empty()) { delete h; return fail(PIC_EINVAL, err); }
  if (set_device(h)) { delete h; return PIC_EHIP; }
  if (int rc = emit_scratch_init(h->device)) { delete h; return rc; }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) { delete h; return fail(PIC_EHIP, "stream"); }
  for (auto &e : h->ev) (void)hipEventCreate(&e);
  const HostMesh &hm = h->hm;
  DevMesh &M = h->M;
  M.nElems = hm.nElems; M.N = hm.N; M.nDof = hm.nDof; M.deposition = par->deposition;
  M.refGuess = par->refmapping_guess; M.refEps = par->refmapping_eps; M.nSpecies = par->nSpecies;
  int rc = 0;
  rc |= upload(h, const_cast<double **>(&M.corner), hm.corner);
  rc |= upload(h, const_cast<int32_t **>(&M.sideNb), hm.sideNb);
  rc |= upload(h, const_cast<int32_t **>(&M.sideGID), hm.sideGID);
  rc |= upload(h, const_cast<uint8_t **>(&M.sideNode), hm.sideNode);
  {
    std::vector<double> sxyz(size_t(hm.nElems) * 72);
    for (int e = 0; e < hm.nElems; ++e)
      for (int l = 0; l < 6; ++l)
        for (int q = 0; q < 4; ++q)
          for (int c = 0; c < 3; ++c)
            sxyz[((size_t(e) * 6 + l) * 4 + q) * 3 + c] =
                hm.corner[(size_t(e) * 8 + hm.sideNode[(size_t(e) * 6 + l) * 4 + q]) * 3 + c];
    rc |= upload(h, const_cast<double **>(&M.sideXYZ), sxyz);
  }
  rc |= upload(h, const_cast<uint8_t **>(&M.sideBC), hm.sideBC);
  rc |= upload(h, const_cast<int8_t **>(&M.sidePV), hm.sidePV);
  rc |= upload(h, const_cast<uint8_t **>(&M.concave), hm.concave);
  rc |= upload(h, const_cast<double **>(&M.dXCL), hm.dXCL1);
  rc |= upload(h, const_cast<double **>(&M.bary), hm.bary);
  rc |= upload(h, const_cast<double **>(&M.radius), hm.radius);
  rc |= upload(h, const_cast<double **>(&M.xiBasis), hm.xiBasis);
  rc |= upload(h, const_cast<double **>(&M.slen), hm.slen);
  rc |= upload(h, const_cast<double **>(&M.xGPe), hm.xGPe);
  {
    // affine elements: the trilinear corner map has no xi.eta, xi.zeta,
    // eta.zeta or xi.eta.zeta terms (to round-off)
    std::vector<double> aff(size_t(hm.nElems) * 12, 0.0);
    bool all = hm.NGeo == 1 && hm.nElems > 0;
    for (int e = 0; e < hm.nElems && all; ++e) {
      all = false;
      const double *X = &hm.corner[size_t(e) * 24];
      double co[8][3] = {};   // coefficients of 1, xi, eta, xi.eta, zeta, xi.zeta, eta.zeta, xi.eta.zeta
      for (int t = 0; t < 8; ++t)
        for (int m = 0; m < 8; ++m) {
          double sg = 1.0;
          for (int d = 0; d < 3; ++d)
            if ((m >> d) & 1) sg *= ((t >> d) & 1) ? 1.0 : -1.0;
          for (int c = 0; c < 3; ++c) co[m][c] += 0.125 * sg * X[t * 3 + c];
        }
      double h1 = 0, hn = 0;
      for (int c = 0; c < 3; ++c) {
        h1 = std::max({h1, std::fabs(co[1][c]), std::fabs(co[2][c]), std::fabs(co[4][c])});
        hn = std::max({hn, std::fabs(co[3][c]), std::fabs(co[5][c]), std::fabs(co[6][c]), std::fabs(co[7][c])});
      }
      if (!(hn <= 1e-14 * h1)) continue;
      const double A[3][3] = {{co[1][0], co[2][0], co[4][0]}, {co[1][1], co[2][1], co[4][1]}, {co[1][2], co[2][2], co[4][2]}};
      const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
      if (!(std::fabs(det) > 0)) continue;
      double *a = &aff[size_t(e) * 12];
      a[0] = co[0][0]; a[1] = co[0][1]; a[2] = co[0][2];
      a[3] = (A[1][1] * A[2][2] - A[1][2] * A[2][1]) / det;
      a[4] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) / det;
      a[5] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) / det;
      a[6] = (A[1][2] * A[2][0] - A[1][0] * A[2][2]) / det;
      a[7] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) / det;
      a[8] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) / det;
      a[9] = (A[1][0] * A[2][1] - A[1][1] * A[2][0]) / det;
      a[10] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) / det;
      a[11] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) / det;
      all = true;
    }
    if (const char *k = std::getenv("PIC_AFFINE")) all = all && std::strcmp(k, "0") != 0;
    M.affine = nullptr;
    if (all) rc |= upload(h, const_cast<double **>(&M.affine), aff);
  }
  rc |= dalloc(h, &h->E, size_t(hm.nElems) * hm.nDof * 3);
  M.E = h->E;
  M.xiCL[0] = -1.0; M.xiCL[1] = 1.0;      // linear CL basis of the corner map
  M.wBaryCL[0] = -0.5; M.wBaryCL[1] = 0.5;
  M.cCL[0] = -0.5; M.cCL[1] = 0.5;        // 1 / (xi_0 - xi_1), 1 / (xi_1 - xi_0)
  M.NGeo = hm.NGeo; M.nGP = hm.nGP; M.tracking = par->tracking; M.nPV = hm.nPV;
  for (int i = 0; i <= hm.NGeo; ++i) {
    M.xiCLg[i] = hm.xiCL[i];
    M.wBaryCLg[i] = hm.wBaryCL[i];
    double p = 1.0;
    for (int m = 0; m <= hm.NGeo; ++m)
      if (m != i) p *= hm.xiCL[i] - hm.xiCL[m];
    M.cCLg[i] = 1.0 / p;
  }
  for (int c = 0; c < 3; ++c) {
    M.gmin[c] = hm.gmin[c]; M.gmax[c] = hm.gmax[c]; M.dirPV[c] = hm.dirPV[c];
    M.bmin[c] = hm.bmin[c]; M.bdelta[c] = hm.bdelta[c]; M.nb[c] = hm.nb[c];
  }
  if (par->tracking == PIC_TRACK_REFMAPPING) {
    if (hm.NGeo == 1) M.XCL = M.corner;
    else rc |= upload(h, const_cast<double **>(&M.XCL), hm.XCL);
    rc |= upload(h, const_cast<double **>(&M.dXCLg), hm.dXCL);
    rc |= upload(h, const_cast<uint8_t **>(&M.curved), hm.curved);
    rc |= upload(h, const_cast<double **>(&M.epsOneCell), hm.epsOneCell);
    rc |= upload(h, const_cast<uint8_t **>(&M.openElem), hm.openElem);
    M.cartPeriodic = hm.cartPeriodic ? 1 : 0;
    if (hm.nBC > 0) {   // ParticleBCTracking tables
      M.bcHalo = hm.bcHalo;
      rc |= upload(h, const_cast<int32_t **>(&M.bcElemNb), hm.bcElemNb);
      rc |= upload(h, const_cast<uint8_t **>(&M.bcKind), hm.bcKind);
      rc |= upload(h, const_cast<int8_t **>(&M.bcPV), hm.bcPV);
      rc |= upload(h, const_cast<uint8_t **>(&M.bcFlip), hm.bcFlip);
      rc |= upload(h, const_cast<double **>(&M.bz), hm.bz);
      rc |= upload(h, const_cast<uint8_t **>(&M.sideType), hm.sideType);
      rc |= upload(h, const_cast<uint8_t **>(&M.bbEmpty), hm.bbEmpty);
      rc |= upload(h, const_cast<double **>(&M.sideNorm), hm.sideNorm);
      rc |= upload(h, const_cast<double **>(&M.sideDist), hm.sideDist);
      rc |= upload(h, const_cast<double **>(&M.slabN), hm.slabN);
      rc |= upload(h, const_cast<double **>(&M.slabI), hm.slabI);
      rc |= upload(h, const_cast<double **>(&M.baseV), hm.baseV);
      rc |= upload(h, const_cast<int32_t **>(&M.e2bOff), hm.e2bOff);
      rc |= upload(h, const_cast<int32_t **>(&M.e2bSide), hm.e2bSide);
      rc |= upload(h, const_cast<double **>(&M.e2bDist), hm.e2bDist);
      rc |= upload(h, const_cast<double **>(&M.e2bSlab), hm.e2bSlab);
    }
  }
  rc |= upload(h, const_cast<int32_t **>(&M.binOff), hm.binOff);
  rc |= upload(h, const_cast<int32_t **>(&M.binElems), hm.binElems);
  for (int i = 0; i <= hm.N; ++i) {
    M.xGP[i] = hm.xGP[i];
    M.wBaryGP[i] = hm.wBaryGP[i];
    double p = 1.0;
    for (int m = 0; m <= hm.N; ++m)
      if (m != i) p *= hm.xGP[i] - hm.xGP[m];
    M.cGP[i] = 1.0 / p;
  }
  for (int v = 0; v < 3; ++v)
    for (int c = 0; c < 3; ++c) M.PV[v][c] = hm.PV[v][c];
  for (int c = 0; c < 6; ++c) M.ext[c] = par->external_field[c];
  for (int s = 0; s < kMaxSpecies; ++s) { M.charge[s] = par->charge[s]; M.mass[s] = par->mass[s]; M.mpf[s] = par->mpf[s]; }
  rc |= dalloc(h, &h->PS, size_t(hm.nElems) * hm.nDof * 4);
  if (par->precision == PIC_PREC_MIXED) {
    rc |= dalloc(h, &h->E32, size_t(hm.nElems) * hm.nDof * 3);
    M.E32 = h->E32;
    for (int i = 0; i <= hm.N; ++i) {
      M.xGP32[i] = float(hm.xGP[i]);
      M.wBaryGP32[i] = float(hm.wBaryGP[i]);
    }
  }
  rc |= upload(h, &h->sJ, hm.sJ);
  rc |= dalloc(h, &h->elemPart, hm.nElems);
  rc |= dalloc(h, &h->sumOut, 1);
  rc |= dalloc(h, &h->specCnt, kMaxSpecies);
  if (par->deposition == PIC_DEPO_CVWM) {
    rc |= dalloc(h, &h->elemCorner, size_t(hm.nElems) * 32);
    rc |= dalloc(h, &h->classSrc, size_t(hm.nClasses) * 4);
    rc |= upload(h, &h->elemClass, hm.elemClass);
    rc |= upload(h, &h->classOff, hm.classOff);
    rc |= upload(h, &h->classInc, hm.classInc);
    rc |= upload(h, &h->classVol, hm.classVolume);
  }
  if (par->deposition == PIC_DEPO_SF) {
    rc |= upload(h, const_cast<int32_t **>(&h->sf.off), hm.sfOff);
    rc |= upload(h, const_cast<int32_t **>(&h->sf.src), hm.sfSrc);
    rc |= upload(h, const_cast<uint8_t **>(&h->sf.shift), hm.sfShift);
    rc |= upload(h, const_cast<int32_t **>(&h->sf.roff), hm.sfRunOff);
    rc |= upload(h, const_cast<int32_t **>(&h->sf.rfirst), hm.sfRunFirst);
    rc |= upload(h, const_cast<int32_t **>(&h->sf.rend), hm.sfRunEnd);
    rc |= upload(h, const_cast<uint8_t **>(&h->sf.rshift), hm.sfRunShift);
    rc |= upload(h, const_cast<int32_t **>(&h->sf.groff), hm.sfGrpRunOff);
    rc |= upload(h, const_cast<int32_t **>(&h->sf.grfirst), hm.sfGrpRunFirst);
    rc |= upload(h, const_cast<int32_t **>(&h->sf.grend), hm.sfGrpRunEnd);
    rc |= upload(h, const_cast<uint8_t **>(&h->sf.grshift), hm.sfGrpRunShift);
    for (int s = 0; s < 27; ++s)
      for (int c = 0; c < 3; ++c) h->sf.sh[s][c] = hm.shift[s][c];
    const double a = par->sf_alpha, r = par->sf_radius;
    h->sf.r_sf = r;
    h->sf.r2_sf = r * r;
    h->sf.r2_sf_inv = 1. / h->sf.r2_sf;
    // InitShapeFunctionDimensionalty 3D branch (pic_depo_shapefunction_tools.f90:1243-1246)
    const double BetaFac = std::tgamma(1.5) * std::tgamma(a + 1.) / std::tgamma(1.5 + (a + 1.));
    h->sf.w_sf = 1. / (2. * BetaFac * a + 2 * BetaFac) * (a + 1.) / (M_PI * (r * r * r));
    h->sf.alpha = par->sf_alpha;
    h->sf.linear = mesh->NGeo == 1;
    h->sf.nspec = par->nSpecies;
    h->sf.r_inv = 1. / r;
    // axis-aligned tensor grids of Gauss points: x depends on i only, y on j, z on k
    if (hm.N == 3) {
      bool cart = true;
      for (int e = 0; e < hm.nElems && cart; ++e) {
        const double *xg = &hm.xGPe[size_t(e) * 64 * 3];
        const double tol = 1e-12 * hm.radius[e];
        for (int k = 0; k < 4 && cart; ++k)
          for (int j = 0; j < 4 && cart; ++j)
            for (int i = 0; i < 4 && cart; ++i) {
              const double *p = xg + (i + 4 * j + 16 * k) * 3;
              cart = std::fabs(p[0] - xg[i * 3]) <= tol && std::fabs(p[1] - xg[4 * j * 3 + 1]) <= tol &&
                     std::fabs(p[2] - xg[16 * k * 3 + 2]) <= tol;
            }
      }
      h->sf.cartesian = cart;
      if (const char *k = std::getenv("PIC_SF_CARTESIAN")) h->sf.cartesian = cart && std::strcmp(k, "0") != 0;
    }
    for (int k = 0; k < 8; ++k) h->sf.fsp[k] = k < par->nSpecies ? par->charge[k] * par->mpf[k] * h->sf.w_sf : 0.0;
    // PIC_SF_KERNEL=valu selects the all-VALU gather (for comparisons)
    if (const char *k = std::getenv("PIC_SF_KERNEL")) h->sfMatrix = std::strcmp(k, "valu") != 0;
  }
  // particles
  const int64_t cap = par->max_particles;
  h->cap = cap;
  rc |= alloc_parts(h, h->P, cap);
  rc |= alloc_parts(h, h->Q, cap);
  rc |= dalloc(h, &h->P.lx, cap); rc |= dalloc(h, &h->P.ly, cap); rc |= dalloc(h, &h->P.lz, cap);
  rc |= dalloc(h, &h->Q.lx, cap); rc |= dalloc(h, &h->Q.ly, cap); rc |= dalloc(h, &h->Q.lz, cap);
  rc |= dalloc(h, &h->P.lastElem, cap); rc |= dalloc(h, &h->Q.lastElem, cap);
  rc |= dalloc(h, &h->P.F, cap * 6); rc |= dalloc(h, &h->Q.F, cap * 6);
  if (par->tracking == PIC_TRACK_REFMAPPING) {
    rc |= dalloc(h, &h->P.rx, cap); rc |= dalloc(h, &h->P.ry, cap); rc |= dalloc(h, &h->P.rz, cap);
    rc |= dalloc(h, &h->Q.rx, cap); rc |= dalloc(h, &h->Q.ry, cap); rc |= dalloc(h, &h->Q.rz, cap);
  }
  if (par->deposition == PIC_DEPO_SF && hm.N == 3) {
    double4 *rec = nullptr;
    rc |= dalloc(h, &rec, size_t(cap) * 2);
    h->sf.rec = rec;
    h->sf.recB = rec + cap;
    double4 *blk = nullptr;
    rc |= dalloc(h, &blk, size_t((cap + kWave - 1) / kWave) * 2);
    h->sf.blk = blk;
    rc |= dalloc(h, &h->sf.work, 8);
    if (hipDeviceGetAttribute(&h->nCU, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess || h->nCU < 1)
      h->nCU = 1;
  }
  rc |= dalloc(h, &h->start, hm.nElems + 1);
  rc |= dalloc(h, &h->startNew, hm.nElems + 1);
  rc |= dalloc(h, &h->newElem, cap);
  rc |= dalloc(h, &h->rank, cap);
  rc |= dalloc(h, &h->perm, cap);
  rc |= dalloc(h, &h->elemOf, cap);
  rc |= dalloc(h, &h->stayCnt, hm.nElems);
  rc |= dalloc(h, &h->arrCnt, hm.nElems);
  rc |= dalloc(h, &h->insideById, cap);
  rc |= dalloc(h, &h->stateById, cap * 6);
  rc |= dalloc(h, &h->mv.count, 2);
  rc |= dalloc(h, &h->mv.idx, cap);
  rc |= dalloc(h, &h->mv.last, cap * 3);
  if (par->tracking == PIC_TRACK_REFMAPPING) {
    rc |= dalloc(h, &h->mv.flag, cap);
    rc |= dalloc(h, &h->mv.slot2, cap);
  }
  rc |= dalloc(h, &h->lostCount, 1);
  if (rc) { pic_destroy(h); return PIC_EHIP; }
  HIPCHK(hipMemset(h->lostCount, 0, sizeof(unsigned long long)));
  HIPCHK(hipMemset(h->E, 0, size_t(hm.nElems) * hm.nDof * 3 * sizeof(double)));
  HIPCHK(hipMemset(h->PS, 0, size_t(hm.nElems) * hm.nDof * 4 * sizeof(double)));
  *out = h;
  return PIC_OK;
}

int pic_destroy(pic_handle *h) {
  if (!h) return PIC_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph(h);
  for (void *p : h->allocs) (void)hipFree(p);
  for (auto &e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PIC_OK;
}

namespace {
// mixed precision: the FP32 copy of E read by the interpolation
int field_to_f32(pic_handle *h) {
  if (!h->E32) return 0;
  const int64_t n = int64_t(h->hm.nElems) * h->hm.nDof * 3;
  hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, h->E, h->E32, n);
  HIPCHK(hipGetLastError());
  return 0;
}
}  // namespace

int pic_set_field(pic_handle *h, const double *E) {
  if (!h || !E) return fail(PIC_EINVAL, "pic_set_field: null argument");
  if (set_device(h)) return PIC_EHIP;
  HIPCHK(hipMemcpyAsync(h->E, E, size_t(h->hm.nElems) * h->hm.nDof * 3 * sizeof(double), hipMemcpyHostToDevice, h->stream));
  if (int rc = field_to_f32(h)) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return PIC_OK;
}

int pic_set_field_device(pic_handle *h, const double *E_dev) {
  if (!h || !E_dev) return fail(PIC_EINVAL, "pic_set_field_device: null argument");
  if (set_device(h)) return PIC_EHIP;
  HIPCHK(hipMemcpyAsync(h->E, E_dev, size_t(h->hm.nElems) * h->hm.nDof * 3 * sizeof(double), hipMemcpyDeviceToDevice,
                        h->stream));
  if (int rc = field_to_f32(h)) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  return PIC_OK;
}

int pic_set_particles(pic_handle *h, int64_t n, const double *PartState, const int32_t *elem, const int32_t *species,
                      const uint8_t *is_new) {
  return set_particles_impl(h, n, PartState, elem, species, is_new, false);
}

}  // extern "C"

namespace {
// pic_set_particles; with outsideMarked an element id of -1 marks a particle
// that no element holds (set up as not inside)
int set_particles_impl(pic_handle *h, int64_t n, const double *PartState, const int32_t *elem, const int32_t *species,
                       const uint8_t *is_new, bool outsideMarked) {
  if (h) h->recFresh = false;
  if (!h || (n > 0 && !PartState)) return fail(PIC_EINVAL, "pic_set_particles: null argument");
  if (n < 0 || n > h->cap) return fail(PIC_ECAPACITY, "pic_set_particles: more particles than max_particles");
  if (set_device(h)) return PIC_EHIP;
  const int ne = h->hm.nElems;
  std::vector<int32_t> el(n), sp(n, 0);
  std::vector<uint8_t> ins(n, 1);
  const bool refmap = h->par.tracking == PIC_TRACK_REFMAPPING;
  std::vector<int64_t> pending;
  for (int64_t i = 0; i < n; ++i) {
    int e = elem ? elem[i] : 0;
    if (outsideMarked && e == -1) {
      el[i] = -1;
    } else {
      if (e < 0 || e > ne) return fail(PIC_EINVAL, "pic_set_particles: element id out of range");
      if (e == 0) pending.push_back(i);
      el[i] = e - 1;
    }
    sp[i] = species ? species[i] - 1 : 0;
    if (sp[i] < 0 || sp[i] >= h->par.nSpecies) return fail(PIC_EINVAL, "pic_set_particles: species out of range");
  }
  if (!pending.empty()) {
    // SinglePointToElement on the device for the particles without an element
    const int64_t np = int64_t(pending.size());
    std::vector<double> xyz(np * 3);
    std::vector<int32_t> pe(np, -1);
    for (int64_t k = 0; k < np; ++k)
      for (int c = 0; c < 3; ++c) xyz[k * 3 + c] = PartState[pending[k] * 6 + c];
    HIPCHK(copy_bytes(h->mv.last, xyz.data(), np * 3 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(copy_bytes(h->newElem, pe.data(), np * sizeof(int32_t), hipMemcpyHostToDevice));
    if (refmap) { LAUNCH_NG(k_rm_locate, h->hm.N, h->hm.NGeo, dim3(grid_for(np)), h->M, np, h->mv.last, h->newElem); }
    else hipLaunchKernelGGL(k_tria_locate, dim3(grid_for(np)), dim3(kBlock), 0, h->stream, h->M, np, h->mv.last, h->newElem);
    HIPCHK(hipGetLastError());
    // the stage's stream does not synchronise with the null stream's copies
    HIPCHK(hipStreamSynchronize(h->stream));
    HIPCHK(copy_bytes(pe.data(), h->newElem, np * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < np; ++k) el[pending[k]] = pe[k];
  }
  for (int64_t i = 0; i < n; ++i)
    if (el[i] < 0) ins[i] = 0;
  // counting sort by element, stable in caller order
  std::vector<int64_t> cnt(ne + 1, 0);
  for (int64_t i = 0; i < n; ++i)
    if (el[i] >= 0) cnt[el[i] + 1]++;
  for (int e = 0; e < ne; ++e) cnt[e + 1] += cnt[e];
  const int64_t nin = cnt[ne];
  std::vector<int32_t> start(ne + 1);
  for (int e = 0; e <= ne; ++e) start[e] = int32_t(cnt[e]);
  std::vector<int64_t> order(nin);
  {
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < n; ++i)
      if (el[i] >= 0) order[fill[el[i]]++] = i;
  }
  std::vector<double> buf(nin);
  double *dst[6] = {h->P.x, h->P.y, h->P.z, h->P.vx, h->P.vy, h->P.vz};
  for (int c = 0; c < 6; ++c) {
    for (int64_t j = 0; j < nin; ++j) buf[j] = PartState[order[j] * 6 + c];
    HIPCHK(copy_bytes(dst[c], buf.data(), nin * sizeof(double), hipMemcpyHostToDevice));
  }
  std::vector<int32_t> ib(nin);
  std::vector<uint8_t> ub(nin);
  for (int64_t j = 0; j < nin; ++j) ib[j] = el[order[j]];
  HIPCHK(copy_bytes(h->P.elem, ib.data(), nin * sizeof(int32_t), hipMemcpyHostToDevice));
  for (int64_t j = 0; j < nin; ++j) ib[j] = int32_t(order[j]);
  HIPCHK(copy_bytes(h->P.id, ib.data(), nin * sizeof(int32_t), hipMemcpyHostToDevice));
  for (int64_t j = 0; j < nin; ++j) ub[j] = uint8_t(sp[order[j]]);
  HIPCHK(copy_bytes(h->P.spec, ub.data(), nin, hipMemcpyHostToDevice));
  for (int64_t j = 0; j < nin; ++j) ub[j] = is_new ? is_new[order[j]] : 0;
  HIPCHK(copy_bytes(h->P.isNew, ub.data(), nin, hipMemcpyHostToDevice));
  HIPCHK(copy_bytes(h->start, start.data(), (ne + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
  HIPCHK(copy_bytes(h->insideById, ins.data(), n, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(h->lostCount, 0, sizeof(unsigned long long)));
  // particles no element holds keep their state for pic_get_particles
  h->unplaced.clear();
  for (int64_t i = 0; i < n; ++i)
    if (!ins[i]) h->unplaced.emplace_back(i, std::vector<double>(PartState + i * 6, PartState + i * 6 + 6));
  h->n = nin;
  h->nTotal = n;
  h->particlesSet = true;
  h->stagedValid = false;
  drop_graph(h);   // recorded for the previous particle arrays
  if (refmap && nin > 0) {
    LAUNCH_NG(k_rm_init, h->hm.N, h->hm.NGeo, dim3(grid_for(nin)), h->M, h->P, nin);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
  }
  return PIC_OK;
}

}  // namespace

extern "C" {

int pic_deposition(pic_handle *h) {
  if (!h) return fail(PIC_EINVAL, "pic_deposition: null handle");
  if (set_device(h)) return PIC_EHIP;
  if (int rc = run_deposition(h, true)) return rc;
  HIPCHK(hipStreamSynchronize(h->stream));
  h->msDeposit += elapsed(h->ev[0], h->ev[2]);
  if (h->par.deposition != PIC_DEPO_NONE) { h->msDepositMain += elapsed(h->ev[8], h->ev[1]); h->nDepositLaunches++; }
  return PIC_OK;
}

int pic_interpolate_field_to_particle(pic_handle *h) {
  if (!h) return fail(PIC_EINVAL, "pic_interpolate_field_to_particle: null handle");
  if (int rc = require_particles(h)) return rc;
  if (set_device(h)) return PIC_EHIP;
  if (h->n > 0) {
    const bool mixed = h->par.precision == PIC_PREC_MIXED;   // N = 3
    if (h->par.tracking == PIC_TRACK_REFMAPPING) {
      if (mixed) hipLaunchKernelGGL((k_rm_interp<3, true>), dim3(grid_for(h->n)), dim3(kBlock), 0, h->stream, h->M, h->P, h->n);
      else { LAUNCH_N(k_rm_interp, h->hm.N, dim3(grid_for(h->n)), h->M, h->P, h->n); }
    } else {
      const bool aff = h->M.affine != nullptr && h->hm.N == 3;
      if (mixed && aff)
        hipLaunchKernelGGL((k_interp<3, true, true>), dim3(grid_for(h->n)), dim3(kBlock), 0, h->stream, h->M, h->P, h->n);
      else if (mixed)
        hipLaunchKernelGGL((k_interp<3, true>), dim3(grid_for(h->n)), dim3(kBlock), 0, h->stream, h->M, h->P, h->n);
      else if (aff)
        hipLaunchKernelGGL((k_interp<3, false, true>), dim3(grid_for(h->n)), dim3(kBlock), 0, h->stream, h->M, h->P, h->n);
      else { LAUNCH_N(k_interp, h->hm.N, dim3(grid_for(h->n)), h->M, h->P, h->n); }
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  h->stagedValid = true;
  return PIC_OK;
}

int pic_push_boris_leapfrog(pic_handle *h, double dt) {
  if (!h) return fail(PIC_EINVAL, "pic_push_boris_leapfrog: null handle");
  if (int rc = require_particles(h)) return rc;
  if (!h->stagedValid) return fail(PIC_ESTATE, "pic_push_boris_leapfrog: call pic_interpolate_field_to_particle first");
  if (set_device(h)) return PIC_EHIP;
  if (h->n > 0) hipLaunchKernelGGL(k_push, dim3(grid_for(h->n)), dim3(kBlock), 0, h->stream, h->M, h->P, h->n, dt);
  h->recFresh = false;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  return PIC_OK;
}

int pic_perform_tracking(pic_handle *h) {
  if (!h) return fail(PIC_EINVAL, "pic_perform_tracking: null handle");
  if (int rc = require_particles(h)) return rc;
  if (!h->stagedValid) return fail(PIC_ESTATE, "pic_perform_tracking: call pic_push_boris_leapfrog first");
  if (set_device(h)) return PIC_EHIP;
  if (h->n > 0) {
    if (h->par.tracking == PIC_TRACK_REFMAPPING) {
      LAUNCH_NG(k_rm_track, h->hm.N, h->hm.NGeo, dim3(grid_for(h->n)), h->M, h->P, h->n, h->newElem, h->lostCount,
                h->insideById, h->stateById);
    } else {
      hipLaunchKernelGGL(k_track, dim3(grid_for(h->n)), dim3(kBlock), 0, h->stream, h->M, h->P, h->n, h->newElem,
                         h->lostCount, h->insideById, h->stateById);
    }
    HIPCHK(hipGetLastError());
    if (int rc = run_sort(h, true)) return rc;
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return PIC_OK;
}

}  // extern "C"

namespace {
// One 508 step without the field solve, enqueued on the stage's stream with
// no host synchronisation: Deposition; LastPartPos <- x, interpolate, push,
// track; re-sort.  nGrid bounds the particle count (the kernels read the
// count itself from the segment table), so the same launches stay valid
// while particles are removed; this makes the step capturable in a graph.
int enqueue_step(pic_handle *h, double dt, int64_t nGrid, bool timed) {
  if (int rc = run_deposition(h, timed)) return rc;
  h->recFresh = false;   // the push below moves the particles; the re-sort writes new records
  if (timed) HIPCHK(hipEventRecord(h->ev[3], h->stream));
  const int ne = h->hm.nElems;
  const int32_t *nDev = h->start + ne;
  const bool mixed = h->par.precision == PIC_PREC_MIXED;   // N = 3
  const dim3 grid(grid_for(std::max<int64_t>(nGrid, 1)));
  if (h->par.tracking == PIC_TRACK_REFMAPPING) {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kBlock), 0, h->stream, h->mv.count, int64_t(2), 0);
    if (timed) HIPCHK(hipEventRecord(h->ev[9], h->stream));
    if (mixed && h->hm.NGeo == 1)
      hipLaunchKernelGGL((k_rm_push_inside<3, 2, true>), grid, dim3(kBlock), 0, h->stream, h->M, h->P, nGrid, dt,
                         h->newElem, h->mv, nDev);
    else if (mixed)
      hipLaunchKernelGGL((k_rm_push_inside<3, 3, true>), grid, dim3(kBlock), 0, h->stream, h->M, h->P, nGrid, dt,
                         h->newElem, h->mv, nDev);
    else
      LAUNCH_NG(k_rm_push_inside, h->hm.N, h->hm.NGeo, grid, h->M, h->P, nGrid, dt, h->newElem, h->mv, nDev);
    if (timed) HIPCHK(hipEventRecord(h->ev[10], h->stream));
    LAUNCH_NG(k_rm_movers, h->hm.N, h->hm.NGeo, dim3(std::min<unsigned>(grid.x, 4096u)), h->M, h->P, h->newElem, h->mv,
              h->lostCount, h->insideById, h->stateById);
    if (h->hm.nBC > 0)
      LAUNCH_NG(k_rm_bc_movers, h->hm.N, h->hm.NGeo, dim3(1024), h->M, h->P, h->newElem, h->mv, h->lostCount,
                h->insideById, h->stateById);
  } else {
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kBlock), 0, h->stream, h->mv.count, int64_t(1), 0);
    if (timed) HIPCHK(hipEventRecord(h->ev[9], h->stream));
    // every element affine (N = 3): xi without Newton
    const bool aff = h->M.affine != nullptr && h->hm.N == 3;
    if (mixed && aff)
      hipLaunchKernelGGL((k_push_inside<3, true, true>), grid, dim3(kBlock), 0, h->stream, h->M, h->P, nGrid, dt,
                         h->newElem, h->mv, nDev);
    else if (mixed)
      hipLaunchKernelGGL((k_push_inside<3, true>), grid, dim3(kBlock), 0, h->stream, h->M, h->P, nGrid, dt, h->newElem,
                         h->mv, nDev);
    else if (aff)
      hipLaunchKernelGGL((k_push_inside<3, false, true>), grid, dim3(kBlock), 0, h->stream, h->M, h->P, nGrid, dt,
                         h->newElem, h->mv, nDev);
    else { LAUNCH_N(k_push_inside, h->hm.N, grid, h->M, h->P, nGrid, dt, h->newElem, h->mv, nDev); }
    if (timed) HIPCHK(hipEventRecord(h->ev[10], h->stream));
    hipLaunchKernelGGL(k_track_movers, dim3(1024), dim3(kBlock), 0, h->stream, h->M, h->P, h->newElem, h->mv,
                       h->lostCount, h->insideById, h->stateById);
  }
  HIPCHK(hipGetLastError());
  if (timed) HIPCHK(hipEventRecord(h->ev[4], h->stream));
  return enqueue_sort(h, nGrid, false);
}

void drop_graph(pic_handle *h) {
  if (h->graphExec) (void)hipGraphExecDestroy(h->graphExec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->graphExec = nullptr;
  h->graph = nullptr;
}

// Two steps captured in one graph: the particle buffers swap every step, so
// after two the pointers the graph was recorded with are current again.
// deposition records of the current arrays, if the gathers read them and
// the last sort did not write them
int refresh_records(pic_handle *h) {
  const bool records = h->par.deposition == PIC_DEPO_SF && h->hm.N == 3 &&
                       (h->sfMatrix || h->par.precision == PIC_PREC_MIXED) && h->sf.rec;
  if (!records || h->recFresh) return 0;
  hipLaunchKernelGGL(k_sf_records, dim3(grid_for(std::max<int64_t>(h->n, 1))), dim3(kBlock), 0, h->stream, h->P,
                     h->start + h->hm.nElems, h->sf, const_cast<double4 *>(h->sf.rec),
                     const_cast<double4 *>(h->sf.recB), block_boxes(h));
  HIPCHK(hipGetLastError());
  h->recFresh = true;
  return 0;
}

int ensure_graph(pic_handle *h, double dt) {
  if (h->graphExec && h->graphDt == dt && h->graphN >= h->n && h->graphP == h->P.x) return 0;
  drop_graph(h);
  const int64_t nGrid = h->n;
  // the captured steps read the records the previous step's sort wrote: form
  // them now if the arrays changed since, and before every replay
  if (int rc = refresh_records(h)) return rc;
  HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_step(h, dt, nGrid, false);
  if (!rc) rc = enqueue_step(h, dt, nGrid, false);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->stream, &g);
  if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return fail(PIC_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  h->graph = g;
  HIPCHK(hipGraphInstantiate(&h->graphExec, g, nullptr, nullptr, 0));
  h->graphDt = dt;
  h->graphN = nGrid;
  h->graphP = h->P.x;
  return 0;
}
}  // namespace

extern "C" {

int pic_time_step(pic_handle *h, double dt, int32_t nsteps) {
  if (!h) return fail(PIC_EINVAL, "pic_time_step: null handle");
  if (int rc = require_particles(h)) return rc;
  if (nsteps < 0) return fail(PIC_EINVAL, "pic_time_step: nsteps < 0");
  if (set_device(h)) return PIC_EHIP;
  int it = 0;
  if (h->useGraph && nsteps >= 2 && h->n > 0) {
    // replay the two-step graph; the device time of the replays is booked
    // as push/track time (the stages are not timed one by one)
    if (int rc = ensure_graph(h, dt)) return rc;
    if (int rc = refresh_records(h)) return rc;
    HIPCHK(hipEventRecord(h->ev[6], h->stream));
    for (; it + 1 < nsteps; it += 2) HIPCHK(hipGraphLaunch(h->graphExec, h->stream));
    HIPCHK(hipEventRecord(h->ev[7], h->stream));
    if (int rc = read_count(h)) return rc;
    h->msPushTrack += elapsed(h->ev[6], h->ev[7]);
    h->nSteps += it;
  }
  for (; it < nsteps; ++it) {
    if (int rc = enqueue_step(h, dt, h->n, true)) return rc;
    if (int rc = read_count(h)) return rc;
    HIPCHK(hipEventRecord(h->ev[5], h->stream));
    HIPCHK(hipEventSynchronize(h->ev[5]));
    h->msDeposit += elapsed(h->ev[0], h->ev[2]);
    if (h->par.deposition != PIC_DEPO_NONE) { h->msDepositMain += elapsed(h->ev[8], h->ev[1]); h->nDepositLaunches++; }
    h->msPushTrack += elapsed(h->ev[3], h->ev[4]);
    h->msPushMain += elapsed(h->ev[9], h->ev[10]);
    h->msSort += elapsed(h->ev[4], h->ev[5]);
    h->nSteps++;
  }
  h->stagedValid = false;
  return PIC_OK;
}

int pic_set_tuning(pic_handle *h, const pic_tuning *t) {
  if (!h || !t) return fail(PIC_EINVAL, "pic_set_tuning: null argument");
  if (t->sf_kernel < 0 || t->sf_kernel > 2) return fail(PIC_EINVAL, "pic_set_tuning: sf_kernel must be 0, 1 or 2");
  if (t->sf_persistent < 0 || t->sf_persistent > 1) return fail(PIC_EINVAL, "pic_set_tuning: sf_persistent must be 0 or 1");
  if (t->sf_shared_loads < 0 || t->sf_shared_loads > 1)
    return fail(PIC_EINVAL, "pic_set_tuning: sf_shared_loads must be 0 or 1");
  if (set_device(h)) return PIC_EHIP;
  // a captured graph holds the kernel choice it was recorded with, and the
  // record refresh before a replay follows the current choice: any change
  // of the tuning records the graph again at the next replayed step
  HIPCHK(hipStreamSynchronize(h->stream));
  drop_graph(h);
  h->useGraph = t->use_graph != 0;
  if (t->sf_kernel != 0) h->sfMatrix = t->sf_kernel == 2;
  h->sfPersist = t->sf_persistent != 0 && h->sf.work != nullptr;
  h->sfCoop = t->sf_shared_loads != 0 && h->sf.groff != nullptr;
  return PIC_OK;
}

int pic_get_particles(pic_handle *h, double *PartState, int32_t *elem, uint8_t *inside) {
  if (!h) return fail(PIC_EINVAL, "pic_get_particles: null handle");
  if (set_device(h)) return PIC_EHIP;
  HIPCHK(hipStreamSynchronize(h->stream));
  const int64_t n = h->n, nt = h->nTotal;
  std::vector<double> a(n);
  std::vector<int32_t> id(n), el(n);
  std::vector<uint8_t> ins(nt);
  HIPCHK(copy_bytes(id.data(), h->P.id, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIPCHK(copy_bytes(el.data(), h->P.elem, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIPCHK(copy_bytes(ins.data(), h->insideById, nt, hipMemcpyDeviceToHost));
  if (PartState) {
    double *src[6] = {h->P.x, h->P.y, h->P.z, h->P.vx, h->P.vy, h->P.vz};
    for (int c = 0; c < 6; ++c) {
      HIPCHK(copy_bytes(a.data(), src[c], n * sizeof(double), hipMemcpyDeviceToHost));
      for (int64_t j = 0; j < n; ++j) PartState[int64_t(id[j]) * 6 + c] = a[j];
    }
    int64_t nlost = 0;
    for (int64_t i = 0; i < nt; ++i) nlost += ins[i] ? 0 : 1;
    if (nlost > 0) {
      std::vector<double> st(size_t(nt) * 6);
      HIPCHK(copy_bytes(st.data(), h->stateById, size_t(nt) * 6 * sizeof(double), hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < nt; ++i)
        if (!ins[i])
          for (int c = 0; c < 6; ++c) PartState[i * 6 + c] = st[i * 6 + c];
      for (auto &u : h->unplaced)
        for (int c = 0; c < 6; ++c) PartState[u.first * 6 + c] = u.second[c];
    }
  }
  if (elem) {
    for (int64_t i = 0; i < nt; ++i) elem[i] = 0;
    for (int64_t j = 0; j < n; ++j) elem[id[j]] = el[j] + 1;
  }
  if (inside)
    for (int64_t i = 0; i < nt; ++i) inside[i] = ins[i];
  return PIC_OK;
}

int pic_get_last_part_pos(pic_handle *h, double *L) {
  if (!h || !L) return fail(PIC_EINVAL, "pic_get_last_part_pos: null argument");
  if (!h->stagedValid) return fail(PIC_ESTATE, "pic_get_last_part_pos: no staged push since the last fused step");
  if (set_device(h)) return PIC_EHIP;
  const int64_t n = h->n;
  std::vector<int32_t> id(n);
  std::vector<double> a(n);
  HIPCHK(copy_bytes(id.data(), h->P.id, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  double *src[3] = {h->P.lx, h->P.ly, h->P.lz};
  for (int c = 0; c < 3; ++c) {
    HIPCHK(copy_bytes(a.data(), src[c], n * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < n; ++j) L[int64_t(id[j]) * 3 + c] = a[j];
  }
  return PIC_OK;
}

int pic_get_part_pos_ref(pic_handle *h, double *ref) {
  if (!h || !ref) return fail(PIC_EINVAL, "pic_get_part_pos_ref: null argument");
  if (h->par.tracking != PIC_TRACK_REFMAPPING) return fail(PIC_ESTATE, "pic_get_part_pos_ref: RefMapping only");
  if (set_device(h)) return PIC_EHIP;
  HIPCHK(hipStreamSynchronize(h->stream));
  const int64_t n = h->n;
  std::vector<int32_t> id(n);
  std::vector<double> a(n);
  HIPCHK(copy_bytes(id.data(), h->P.id, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  double *src[3] = {h->P.rx, h->P.ry, h->P.rz};
  for (int c = 0; c < 3; ++c) {
    HIPCHK(copy_bytes(a.data(), src[c], n * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < n; ++j) ref[int64_t(id[j]) * 3 + c] = a[j];
  }
  return PIC_OK;
}

int pic_get_field_at_particle(pic_handle *h, double *F) {
  if (!h || !F) return fail(PIC_EINVAL, "pic_get_field_at_particle: null argument");
  if (!h->stagedValid) return fail(PIC_ESTATE, "pic_get_field_at_particle: call pic_interpolate_field_to_particle first");
  if (set_device(h)) return PIC_EHIP;
  const int64_t n = h->n;
  std::vector<int32_t> id(n);
  std::vector<double> a(size_t(n) * 6);
  HIPCHK(copy_bytes(id.data(), h->P.id, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIPCHK(copy_bytes(a.data(), h->P.F, size_t(n) * 6 * sizeof(double), hipMemcpyDeviceToHost));
  for (int64_t j = 0; j < n; ++j)
    for (int c = 0; c < 6; ++c) F[int64_t(id[j]) * 6 + c] = a[j * 6 + c];
  return PIC_OK;
}

int pic_get_part_source(pic_handle *h, double *PS) {
  if (!h || !PS) return fail(PIC_EINVAL, "pic_get_part_source: null argument");
  if (set_device(h)) return PIC_EHIP;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(copy_bytes(PS, h->PS, size_t(h->hm.nElems) * h->hm.nDof * 4 * sizeof(double), hipMemcpyDeviceToHost));
  return PIC_OK;
}

int pic_part_source_device(pic_handle *h, const double **PS_dev) {
  if (!h || !PS_dev) return fail(PIC_EINVAL, "pic_part_source_device: null argument");
  *PS_dev = h->PS;
  return PIC_OK;
}

namespace {
// element integral of the current PS (mode 0) or E (mode 1), summed on the device
int elem_integral(pic_handle *h, int mode, double *out) {
  const int ne = h->hm.nElems;
  QuadArgs Q{};
  for (int i = 0; i <= h->hm.N; ++i) Q.wGP[i] = h->hm.wGP[i];
  const double *src = mode == 0 ? h->PS : h->E;
  LAUNCH_N(k_elem_integral, h->hm.N, dim3(grid_for(ne)), ne, mode, h->sJ, src, Q, h->elemPart);
  hipLaunchKernelGGL(k_sum_fixed, dim3(1), dim3(1024), 0, h->stream, int64_t(ne), h->elemPart, h->sumOut);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, h->sumOut, sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}
}  // namespace

int pic_deposited_charge(pic_handle *h, double *part_charge) {
  if (!h || !part_charge) return fail(PIC_EINVAL, "pic_deposited_charge: null argument");
  if (set_device(h)) return PIC_EHIP;
  double dep = 0.0;
  if (int rc = elem_integral(h, 0, &dep)) return rc;
  // charge of the particles inside: per-species counts (exact) times ChargeIC * MacroParticleFactor
  unsigned long long cnt[kMaxSpecies] = {0};
  HIPCHK(hipMemsetAsync(h->specCnt, 0, sizeof(cnt), h->stream));
  if (h->n > 0)
    hipLaunchKernelGGL(k_species_count, dim3(std::min<int64_t>(grid_for(h->n), 2048)), dim3(kBlock), 0, h->stream,
                       h->n, h->P.spec, h->specCnt);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(cnt, h->specCnt, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  double part = 0.0;
  for (int s = 0; s < h->par.nSpecies; ++s) part = part + double(cnt[s]) * (h->par.charge[s] * h->par.mpf[s]);
  part_charge[0] = dep;
  part_charge[1] = std::fabs(part - dep);
  // ALMOSTZERO (piclas.h:83): no relative error for a zero total
  part_charge[2] = std::fabs(part) <= 2.22e-16 ? 0.0 : std::fabs(part - dep) / part;
  return PIC_OK;
}

int pic_field_energy(pic_handle *h, double *w_el) {
  if (!h || !w_el) return fail(PIC_EINVAL, "pic_field_energy: null argument");
  if (set_device(h)) return PIC_EHIP;
  double w = 0.0;
  if (int rc = elem_integral(h, 1, &w)) return rc;
  const double eps0 = 8.8541878176e-12;   // globals_vars.f90:84
  *w_el = w * eps0 * 0.5;
  return PIC_OK;
}

int pic_get_stats(pic_handle *h, pic_stats *st) {
  if (!h || !st) return fail(PIC_EINVAL, "pic_get_stats: null argument");
  if (set_device(h)) return PIC_EHIP;
  unsigned long long lost = 0;
  HIPCHK(copy_bytes(&lost, h->lostCount, sizeof(lost), hipMemcpyDeviceToHost));
  st->n_particles = h->n;
  st->n_lost = int64_t(lost);
  st->n_steps = h->nSteps;
  st->ms_deposit = h->msDeposit;
  st->ms_push_track = h->msPushTrack;
  st->ms_sort = h->msSort;
  st->n_deposit_launches = h->nDepositLaunches;
  st->ms_deposit_main = h->msDepositMain;
  st->ms_push_main = h->msPushMain;
  st->deposit_kernel = deposit_kernel(h);
  return PIC_OK;
}

int pic_reset_stats(pic_handle *h) {
  if (!h) return fail(PIC_EINVAL, "pic_reset_stats: null handle");
  h->msDeposit = h->msPushTrack = h->msSort = h->msDepositMain = h->msPushMain = 0;
  h->nDepositLaunches = 0;
  h->nSteps = 0;
  return PIC_OK;
}

int pic_synchronize(pic_handle *h) {
  if (!h) return fail(PIC_EINVAL, "pic_synchronize: null handle");
  if (set_device(h)) return PIC_EHIP;
  HIPCHK(hipStreamSynchronize(h->stream));
  return PIC_OK;
}

}  // extern "C"

#include "pic_adjoint_impl.h"
#include "pic_emission.h"
#include "pic_state_io.h"
