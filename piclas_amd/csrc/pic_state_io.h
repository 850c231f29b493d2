// Particle state in the reference's state-file layout: the element-sorted
// PartData / PartInt arrays that io_hdf5/hdf5_output_particle.f90 writes
// (FillParticleData, :1514-1620) and particles/restart reads
// (particle_readin.f90:503-637, particle_restart.f90:134-440).
//
//   PartInt [nElems][2]  first and last particle offset of every element
//                        (particles first+1 .. last in the reference's
//                        1-based numbering, i.e. [first, last) 0-based)
//   PartData [n][7]      x y z vx vy vz species (species as a real)
//
// The HDF5 file itself is read and written by the host (Python, the image's
// HDF5 library); these entry points move the arrays in and out of the stage.
//
// Included at the end of pic_stage.hip (uses the handle and its helpers).
#pragma once

namespace {

constexpr int kPartDataSize = PIC_PART_DATA_SIZE;

// ParticleRestart's in-element check of the stored element
// (particle_restart.f90:285-330, TriaTracking): ParticleInsideQuad; a particle
// that fails it is marked (-1) for SinglePointToElement.
__global__ __launch_bounds__(kBlock) void k_tria_restart_check(DevMesh M, int64_t n, const double *xyz, int32_t *elem) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int e = elem[i];
  if (e < 0) return;
  const double x[3] = {xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2]};
  double det[6][2];
  if (!inside_quad(M, e, x, det)) elem[i] = -1;
}

// The same check with RefMapping (particle_restart.f90:389-440):
// GetPositionInRefElem in the stored element, inside when every |xi| is at
// most ElemEpsOneCell.
template <int N, int G>
__global__ __launch_bounds__(kBlock) void k_rm_restart_check(DevMesh M, int64_t n, const double *xyz, int32_t *elem) {
  const int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const int e = elem[i];
  if (e < 0) return;
  const double x[3] = {xyz[i * 3], xyz[i * 3 + 1], xyz[i * 3 + 2]};
  double xi[3];
  rm_position<N, G>(M, e, x, xi);
  if (!(max_abs3(xi) <= M.epsOneCell[e])) elem[i] = -1;
}

}  // namespace

extern "C" {

int pic_restart_particles(pic_handle *h, int64_t n, const double *PartData, const int64_t *PartInt,
                          int64_t *n_missing, int64_t *n_lost) {
  if (!h || (n > 0 && !PartData) || !PartInt) return fail(PIC_EINVAL, "pic_restart_particles: null argument");
  if (n < 0 || n > h->cap) return fail(PIC_ECAPACITY, "pic_restart_particles: more particles than max_particles");
  const int ne = h->hm.nElems;
  // offsetnPart = PartInt(1, first element); the elements' ranges must tile [0, n)
  const int64_t off = PartInt[0];
  int64_t prev = off;
  for (int e = 0; e < ne; ++e) {
    const int64_t a = PartInt[2 * e] - off, b = PartInt[2 * e + 1] - off;
    if (a != prev - off || b < a || b > n)
      return fail(PIC_EINVAL, "pic_restart_particles: PartInt ranges do not tile the particle data");
    prev = PartInt[2 * e + 1];
  }
  if (prev - off != n) return fail(PIC_EINVAL, "pic_restart_particles: PartInt does not cover the particle data");
  // PartState, species and the stored element (particle_restart.f90:143-258)
  std::vector<double> state(size_t(n) * 6);
  std::vector<int32_t> el(n), sp(n);
  for (int e = 0; e < ne; ++e)
    for (int64_t p = PartInt[2 * e] - off; p < PartInt[2 * e + 1] - off; ++p) {
      const double s = PartData[p * kPartDataSize + 6];
      const int spec = int(s);
      if (spec <= 0) return fail(PIC_EINVAL, "pic_restart_particles: species ID zero (corrupted restart data)");
      if (spec > h->par.nSpecies) return fail(PIC_EINVAL, "pic_restart_particles: species out of range");
      for (int c = 0; c < 6; ++c) state[p * 6 + c] = PartData[p * kPartDataSize + c];
      sp[p] = spec;
      el[p] = e;
    }
  if (set_device(h)) return PIC_EHIP;
  int64_t missing = 0, lost = 0;
  if (n > 0) {
    // in-element check, then SinglePointToElement for the particles that fail it
    std::vector<double> xyz(size_t(n) * 3);
    for (int64_t p = 0; p < n; ++p)
      for (int c = 0; c < 3; ++c) xyz[p * 3 + c] = state[p * 6 + c];
    HIPCHK(copy_bytes(h->mv.last, xyz.data(), size_t(n) * 3 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(copy_bytes(h->newElem, el.data(), size_t(n) * sizeof(int32_t), hipMemcpyHostToDevice));
    const bool refmap = h->par.tracking == PIC_TRACK_REFMAPPING;
    if (refmap) { LAUNCH_NG(k_rm_restart_check, h->hm.N, h->hm.NGeo, dim3(grid_for(n)), h->M, n, h->mv.last, h->newElem); }
    else hipLaunchKernelGGL(k_tria_restart_check, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, h->M, n, h->mv.last,
                            h->newElem);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<int32_t> chk(n);
    HIPCHK(copy_bytes(chk.data(), h->newElem, size_t(n) * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t p = 0; p < n; ++p) missing += chk[p] < 0;
    if (missing > 0) {
      if (refmap) { LAUNCH_NG(k_rm_locate, h->hm.N, h->hm.NGeo, dim3(grid_for(n)), h->M, n, h->mv.last, h->newElem); }
      else hipLaunchKernelGGL(k_tria_locate, dim3(grid_for(n)), dim3(kBlock), 0, h->stream, h->M, n, h->mv.last,
                              h->newElem);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(h->stream));
      HIPCHK(copy_bytes(chk.data(), h->newElem, size_t(n) * sizeof(int32_t), hipMemcpyDeviceToHost));
    }
    // 1-based elements for set_particles; -1 marks a lost particle (not inside)
    for (int64_t p = 0; p < n; ++p) {
      el[p] = chk[p] >= 0 ? chk[p] + 1 : -1;
      lost += chk[p] < 0;
    }
  }
  if (int rc = set_particles_impl(h, n, state.data(), el.data(), sp.data(), nullptr, true)) return rc;
  if (n_missing) *n_missing = missing;
  if (n_lost) *n_lost = lost;
  return PIC_OK;
}

int pic_get_particle_data(pic_handle *h, int64_t cap, double *PartData, int64_t *PartInt, int64_t *n_out) {
  if (!h || !PartInt || !n_out) return fail(PIC_EINVAL, "pic_get_particle_data: null argument");
  if (set_device(h)) return PIC_EHIP;
  HIPCHK(hipStreamSynchronize(h->stream));
  const int64_t n = h->n, nt = h->nTotal;
  const int ne = h->hm.nElems;
  std::vector<int32_t> id(n), el(n);
  std::vector<uint8_t> spec(n);
  HIPCHK(copy_bytes(id.data(), h->P.id, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIPCHK(copy_bytes(el.data(), h->P.elem, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  HIPCHK(copy_bytes(spec.data(), h->P.spec, n, hipMemcpyDeviceToHost));
  // particles inside, by element; within an element in caller order (the
  // reference's PEM%pNext chain runs in particle index order,
  // particles/particle_tools.f90:99-133)
  std::vector<int64_t> slotById(nt, -1);
  std::vector<int64_t> cnt(ne + 1, 0);
  for (int64_t j = 0; j < n; ++j)
    if (el[j] >= 0) {
      slotById[id[j]] = j;
      cnt[el[j] + 1]++;
    }
  for (int e = 0; e < ne; ++e) cnt[e + 1] += cnt[e];
  const int64_t nin = cnt[ne];
  *n_out = nin;
  for (int e = 0; e < ne; ++e) {
    PartInt[2 * e] = cnt[e];
    PartInt[2 * e + 1] = cnt[e + 1];
  }
  if (nin > cap) return fail(PIC_ECAPACITY, "pic_get_particle_data: PartData holds fewer rows than particles inside");
  if (nin > 0 && !PartData) return fail(PIC_EINVAL, "pic_get_particle_data: null PartData");
  std::vector<int64_t> row(n, -1);
  {
    std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
    for (int64_t i = 0; i < nt; ++i) {
      const int64_t j = slotById[i];
      if (j >= 0) row[j] = fill[el[j]]++;
    }
  }
  std::vector<double> a(n);
  double *src[6] = {h->P.x, h->P.y, h->P.z, h->P.vx, h->P.vy, h->P.vz};
  for (int c = 0; c < 6; ++c) {
    HIPCHK(copy_bytes(a.data(), src[c], n * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t j = 0; j < n; ++j)
      if (row[j] >= 0) PartData[row[j] * kPartDataSize + c] = a[j];
  }
  for (int64_t j = 0; j < n; ++j)
    if (row[j] >= 0) PartData[row[j] * kPartDataSize + 6] = double(int(spec[j]) + 1);
  return PIC_OK;
}

}  // extern "C"
