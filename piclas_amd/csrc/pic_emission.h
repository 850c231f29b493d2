// Particle emission on the device: the initial insertion of
// particles/emission/particle_position_and_velocity.f90 (SetParticlePosition,
// SetParticleVelocity) for the space types 'cuboid' and 'sin_deviation' and
// the velocity distributions 'constant' and 'maxwell_lpn'.  Positions and
// velocities are formed on the device, the particles are located there
// (SinglePointToElement) and then set as the stage's particles.
//
// Random numbers come from a counter-based generator, Philox4x32-10
// (Salmon, Moraes, Dror and Shaw, "Parallel random numbers: as easy as
// 1, 2, 3", SC'11): key = the 64-bit seed, counter = (particle id, draw
// index), so that every particle's numbers are fixed by (seed, id) whatever
// the launch shape.  The reference's RANDOM_NUMBER stream is not reproduced.
//
// Included at the end of pic_stage.hip (uses the handle and its helpers).
#pragma once

namespace {

struct Philox4 {
  uint32_t v[4];
};

__device__ inline Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
  constexpr uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;   // round multipliers
  constexpr uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;   // Weyl key increments
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = uint64_t(kM0) * c0, p1 = uint64_t(kM1) * c2;
    const uint32_t hi0 = uint32_t(p0 >> 32), lo0 = uint32_t(p0), hi1 = uint32_t(p1 >> 32), lo1 = uint32_t(p1);
    const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += kW0; k1 += kW1;
  }
  return {{c0, c1, c2, c3}};
}

// two uniform doubles in [0,1) (53 random bits each) for (particle, draw)
__device__ inline void uniform2(uint64_t seed, int64_t id, uint32_t draw, double &u0, double &u1) {
  const Philox4 r = philox4x32_10(uint32_t(uint64_t(id)), uint32_t(uint64_t(id) >> 32), draw, 0u, uint32_t(seed),
                                  uint32_t(seed >> 32));
  const double s = 1.0 / 9007199254740992.0;   // 2^-53
  u0 = double(((uint64_t(r.v[0]) << 32) | r.v[1]) >> 11) * s;
  u1 = double(((uint64_t(r.v[2]) << 32) | r.v[3]) >> 11) * s;
}

struct EmitInit {
  int64_t first;             // first particle id of this init
  int64_t count;
  int space, velocity, species;
  double base[3], v1[3], v2[3], line[3];   // cuboid: BasePointIC, BaseVectorIC 1/2, unit normal * CuboidHeightIC
  int nxyz[3];
  double amplitude, waveNumber;
  double veloIC, veloVec[3], temperature, mass;
};
struct EmitArgs {
  int nInits;
  EmitInit init[8];
  double gmin[3], gmax[3];   // GEO%x/y/z min/max glob
  uint64_t seed;
};

__global__ __launch_bounds__(kBlock) void k_emit(EmitArgs A, int64_t n, double *state, double *xyz, int32_t *elem,
                                                 uint8_t *spec) {
  const int64_t p = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (p >= n) return;
  int a = 0;
  while (a + 1 < A.nInits && p >= A.init[a + 1].first) ++a;
  const EmitInit &I = A.init[a];
  const int64_t q = p - I.first;
  uint32_t draw = 0;
  double x[3];
  if (I.space == PIC_SPACE_CUBOID) {
    // SetParticlePositionCuboidCylinder 'cuboid' (particle_emission_tools.f90:1157-1161)
    double r0, r1, r2, r3;
    uniform2(A.seed, p, draw++, r0, r1);
    uniform2(A.seed, p, draw++, r2, r3);
    for (int c = 0; c < 3; ++c) {
      double v = I.base[c] + I.v1[c] * r0;
      v = v + I.v2[c] * r1;
      x[c] = v + I.line[c] * r2;
    }
  } else {
    // SetParticlePositionSinDeviation (particle_emission_tools.f90:1267-1294):
    // the lattice loops i (x), j (y), k (z), k innermost, from 1
    const int nz = I.nxyz[2], ny = I.nxyz[1];
    const int64_t k = q % nz, j = (q / nz) % ny, i = q / (int64_t(nz) * ny);
    const double xlen = fabs(A.gmax[0] - A.gmin[0]), ylen = fabs(A.gmax[1] - A.gmin[1]);
    const double zlen = fabs(A.gmax[2] - A.gmin[2]);
    const double pilen = 2.0 * M_PI / xlen;
    const double xs = xlen / I.nxyz[0], ys = ylen / ny, zs = zlen / nz;
    const double xp = (double(i + 1) * xs - xs * 0.5);
    x[0] = A.gmin[0] + xp + I.amplitude * sin(I.waveNumber * pilen * xp);
    x[1] = A.gmin[1] + double(j + 1) * ys - ys * 0.5;
    x[2] = A.gmin[2] + double(k + 1) * zs - zs * 0.5;
  }
  double v[3];
  if (I.velocity == PIC_VELO_CONSTANT) {
    for (int c = 0; c < 3; ++c) v[c] = I.veloVec[c] * I.veloIC;
  } else {
    // CalcVelocity_maxwell_lpn (particle_emission_tools.f90:338-359): Marsaglia's
    // polar method, x and y from one accepted pair, z from a second one
    const double kB = 1.380648813e-23;   // BoltzmannConst, globals_vars.f90:61
    double w[4];
    for (int pair = 0; pair < 2; ++pair) {
      double a0 = 0.0, a1 = 0.0, s = 2.0;
      for (int tries = 0; tries < 64 && (s >= 1.0 || s == 0.0); ++tries) {
        double r0, r1;
        uniform2(A.seed, p, draw++, r0, r1);
        a0 = 2. * r0 - 1.;
        a1 = 2. * r1 - 1.;
        s = a0 * a0 + a1 * a1;
      }
      const double f = sqrt(-2 * kB * I.temperature / I.mass * log(s) / s);
      w[2 * pair] = a0 * f;
      w[2 * pair + 1] = a1 * f;
    }
    v[0] = w[0]; v[1] = w[1]; v[2] = w[2];
    for (int c = 0; c < 3; ++c) v[c] = v[c] + I.veloIC * I.veloVec[c];
  }
  for (int c = 0; c < 3; ++c) {
    state[p * 6 + c] = x[c];
    state[p * 6 + 3 + c] = v[c];
    xyz[p * 3 + c] = x[c];
  }
  elem[p] = -1;
  spec[p] = uint8_t(I.species);
}

// Staging of the allocation-free emission (pic_emit_particles_noalloc): a
// file-scope device array, one copy per GPU, for at most kEmitScratchMax
// particles, carved as PartState [max][6] doubles, positions [max][3]
// doubles, element [max] int32, species [max] uint8 (every region starts at
// a multiple of 256 bytes).  Host code never passes `scratch` itself to a
// kernel or copy: it uses the device address that hipGetSymbolAddress
// returned for the current GPU (emit_scratch_ptr).  Every call on one GPU
// shares the array; the emission kernels write every entry they later read,
// so nothing needs clearing between calls.  Calls that overlap on one GPU
// (two handles on one device driven from two threads) race.
constexpr int64_t kEmitScratchMax = int64_t(1) << 19;
constexpr size_t kEmitScratchBytes = size_t(kEmitScratchMax) * (6 * 8 + 3 * 8 + 4 + 1);
static __device__ __align__(256) unsigned char scratch[kEmitScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void *g_scratchPtr[kMaxDevices] = {};

// once per device, with that device current (called by pic_create)
int emit_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(PIC_EINVAL, "pic_create: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if (g_scratchPtr[device]) return 0;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(PIC_EHIP, std::string("pic_create: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratchPtr[device] = ptr;
  return 0;
}

int emit_impl(pic_handle *h, int32_t n_inits, const pic_emission *inits, uint64_t seed, int64_t *n_emitted,
              bool noalloc) {
  const char *who = noalloc ? "pic_emit_particles_noalloc" : "pic_emit_particles";
  if (!h || (n_inits > 0 && !inits)) return fail(PIC_EINVAL, std::string(who) + ": null argument");
  if (n_inits < 1 || n_inits > 8) return fail(PIC_EINVAL, std::string(who) + ": 1 to 8 emission inits");
  if (set_device(h)) return PIC_EHIP;
  EmitArgs A{};
  A.nInits = n_inits;
  A.seed = seed;
  // GEO%x/y/z min/max glob (GetMeshMinMax): all mesh nodes with
  // TriaTracking, the sides' Bezier control points with RefMapping
  for (int c = 0; c < 3; ++c) { A.gmin[c] = h->hm.gmin[c]; A.gmax[c] = h->hm.gmax[c]; }
  int64_t n = 0;
  for (int a = 0; a < n_inits; ++a) {
    const pic_emission &E = inits[a];
    EmitInit &I = A.init[a];
    if (E.species < 1 || E.species > h->par.nSpecies) return fail(PIC_EINVAL, std::string(who) + ": species out of range");
    I.species = E.species - 1;
    I.space = E.space;
    I.velocity = E.velocity;
    I.first = n;
    if (E.space == PIC_SPACE_CUBOID) {
      if (E.particle_number < 0) return fail(PIC_EINVAL, std::string(who) + ": negative ParticleNumber");
      I.count = E.particle_number;
      // unit normal of the base vectors (particle_emission_tools.f90:1146-1152)
      const double *a1 = E.base_vector1, *a2 = E.base_vector2;
      double l[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]};
      const double nl = std::sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
      if (!(nl > 0.0)) return fail(PIC_EINVAL, std::string(who) + ": BaseVectors are parallel");
      for (int c = 0; c < 3; ++c) {
        I.base[c] = E.base_point[c];
        I.v1[c] = a1[c];
        I.v2[c] = a2[c];
        I.line[c] = l[c] / nl * E.cuboid_height;
      }
    } else if (E.space == PIC_SPACE_SIN_DEVIATION) {
      if (E.max_particle_number[0] < 1 || E.max_particle_number[1] < 1 || E.max_particle_number[2] < 1)
        return fail(PIC_EINVAL, std::string(who) + ": maxParticleNumber must be positive in x, y and z");
      I.count = int64_t(E.max_particle_number[0]) * E.max_particle_number[1] * E.max_particle_number[2];
      for (int c = 0; c < 3; ++c) I.nxyz[c] = E.max_particle_number[c];
      I.amplitude = E.amplitude;
      I.waveNumber = E.wave_number;
    } else {
      return fail(PIC_EINVAL, std::string(who) + ": unknown space type");
    }
    if (E.velocity != PIC_VELO_CONSTANT && E.velocity != PIC_VELO_MAXWELL_LPN)
      return fail(PIC_EINVAL, std::string(who) + ": unknown velocity distribution");
    I.veloIC = E.velo_ic;
    for (int c = 0; c < 3; ++c) I.veloVec[c] = E.velo_vec[c];
    I.temperature = E.temperature;
    I.mass = h->par.mass[E.species - 1];
    n += I.count;
  }
  if (n > h->cap) return fail(PIC_ECAPACITY, std::string(who) + ": more particles than max_particles");
  // staging on the device: PartState [n][6], positions [n][3], element, species
  double *state = nullptr, *xyz = nullptr;
  int32_t *elem = nullptr;
  uint8_t *spec = nullptr;
  const size_t nn = size_t(std::max<int64_t>(n, 1));
  auto release = [&]() {
    if (noalloc) return;
    if (state) (void)hipFree(state);
    if (xyz) (void)hipFree(xyz);
    if (elem) (void)hipFree(elem);
    if (spec) (void)hipFree(spec);
  };
  if (noalloc) {
    if (n > kEmitScratchMax)
      return fail(PIC_ECAPACITY, std::string(who) + ": " + std::to_string(n) + " particles exceed the staging limit of " +
                                     std::to_string(kEmitScratchMax) + " (use pic_emit_particles)");
    unsigned char *base = (h->device >= 0 && h->device < kMaxDevices)
                              ? static_cast<unsigned char *>(g_scratchPtr[h->device]) : nullptr;
    if (!base) return fail(PIC_ESTATE, std::string(who) + ": staging of this device not initialised");
    state = reinterpret_cast<double *>(base);
    xyz = reinterpret_cast<double *>(base + size_t(kEmitScratchMax) * 48);
    elem = reinterpret_cast<int32_t *>(base + size_t(kEmitScratchMax) * 72);
    spec = base + size_t(kEmitScratchMax) * 76;
  } else if (hipMalloc(&state, nn * 6 * sizeof(double)) != hipSuccess ||
             hipMalloc(&xyz, nn * 3 * sizeof(double)) != hipSuccess ||
             hipMalloc(&elem, nn * sizeof(int32_t)) != hipSuccess || hipMalloc(&spec, nn) != hipSuccess) {
    release();
    return fail(PIC_EHIP, "pic_emit_particles: device staging allocation");
  }
  std::vector<double> hs(size_t(n) * 6);
  std::vector<int32_t> he(n);
  std::vector<uint8_t> hsp(n);
  if (n > 0) {
    hipLaunchKernelGGL(k_emit, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, A, n, state, xyz, elem, spec);
    // SinglePointToElement on the device
    if (h->par.tracking == PIC_TRACK_REFMAPPING) {
      LAUNCH_NG(k_rm_locate, h->hm.N, h->hm.NGeo, dim3(grid_for(n)), h->M, n, xyz, elem);
    } else {
      hipLaunchKernelGGL(k_tria_locate, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, h->M, n, xyz, elem);
    }
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess ||
        hipMemcpy(hs.data(), state, size_t(n) * 6 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(he.data(), elem, size_t(n) * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hsp.data(), spec, size_t(n), hipMemcpyDeviceToHost) != hipSuccess) {
      release();
      return fail(PIC_EHIP, std::string(who) + ": emission kernels");
    }
  }
  release();
  // set them as the stage's particles: 1-based elements (particles outside
  // the mesh are kept as not inside), 1-based species, IsNewPart
  std::vector<int32_t> el(n), sp(n);
  std::vector<uint8_t> isNew(n, 1);
  int64_t inside = 0;
  for (int64_t i = 0; i < n; ++i) {
    el[i] = he[i] >= 0 ? he[i] + 1 : -1;
    sp[i] = int32_t(hsp[i]) + 1;
    inside += he[i] >= 0;
  }
  // set_particles locates element 0; an element of -1 marks the emitted
  // particles no element holds
  if (int rc = set_particles_impl(h, n, hs.data(), el.data(), sp.data(), isNew.data(), true)) return rc;
  if (n_emitted) *n_emitted = inside;
  return PIC_OK;
}

}  // namespace

extern "C" {

int pic_emit_particles(pic_handle *h, int32_t n_inits, const pic_emission *inits, uint64_t seed, int64_t *n_emitted) {
  return emit_impl(h, n_inits, inits, seed, n_emitted, false);
}

int pic_emit_particles_noalloc(pic_handle *h, int32_t n_inits, const pic_emission *inits, uint64_t seed,
                               int64_t *n_emitted) {
  return emit_impl(h, n_inits, inits, seed, n_emitted, true);
}

}  // extern "C"
