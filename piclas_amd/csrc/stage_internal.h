// Internal declarations of the MI355X particle stage: the host-side mesh
// tables (built once from the HOPR arrays) and the device views the kernels
// read.  Memory layout in HBM:
//   particles   structure of arrays, sorted by element (segment offsets
//               elemStart[nElems+1]); x,y,z,vx,vy,vz FP64, elem/id int32
//   per element corner coordinates, side topology, Newton tables, E field,
//               Gauss-point positions: element-major, read by the particles
//               of one element segment back to back.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "pic_stage.h"

namespace pic {

constexpr int kMaxN = 7;          // DG degree supported by the kernels
constexpr int kMaxNGeo = 2;       // geometry degree (TriaTracking: 1; RefMapping: 1 or 2)
constexpr int kMaxSpecies = 8;
constexpr int kMaxBinElems = 256;  // elements per RefMapping background cell

// Host-side mesh tables: the reference's particle-mesh arrays, derived from
// the HOPR arrays exactly as its init routines derive them.
struct HostMesh {
  int nElems = 0, nSides = 0, nNodes = 0, NGeo = 1, N = 1, nDof = 8, nGP = 8;
  int nUnique = 0, nClasses = 0;
  // tracking (particle_mesh_tools.f90:1800-1930)
  std::vector<double> corner;      // [e][8][3], tensor corner order t=i+2j+4k
  std::vector<int32_t> sideNb;     // [e][6] 0-based neighbour element or -1
  std::vector<int32_t> sideGID;    // [e][6] 1-based global side index
  std::vector<uint8_t> sideNode;   // [e][6][4] tensor corner index of ElemSideNodeID
  std::vector<uint8_t> sideBC;     // [e][6] 0 interior, PIC_BC_* on boundary sides
  std::vector<int8_t> sidePV;      // [e][6] periodic vector id (BC_ALPHA)
  std::vector<uint8_t> concave;    // [e] bit l set: ConcaveElemSide(l+1)
  // reference mapping (metrics.f90, particle_mesh_build.f90)
  std::vector<double> XCL;         // [e][nGP][3]
  std::vector<double> dXCL;        // [e][nGP][3(x)][3(d)]
  std::vector<double> bary;        // [e][3]  ElemBaryNGeo (corner mean)
  std::vector<double> radius;      // [e]     ElemRadiusNGeo
  std::vector<double> xiBasis;     // [e][6][3]
  std::vector<double> slen;        // [e][6]
  std::vector<double> xGPe;        // [e][nDof][3] Elem_xGP
  std::vector<double> sJ;          // [e][nDof]
  // RefMapping (particle_mesh_build.f90, particle_mesh_tools.f90:976-996)
  std::vector<uint8_t> curved;     // [e] ElemCurved
  std::vector<double> dXCL1;       // [e][8][3][3] dXCL_NGeo at the corner CL points
  std::vector<double> epsOneCell;  // [e] ElemEpsOneCell
  std::vector<uint8_t> openElem;   // [e] element has an open particle boundary side
  double gmin[3] = {0}, gmax[3] = {0};  // GEO%x/y/zminglob, maxglob
  int dirPV[3] = {-1, -1, -1};     // axis of each periodic vector
  // background cells for the RefMapping candidate search: elements whose
  // bounding sphere (bary, radius) overlaps the cell
  double bmin[3] = {0}, bdelta[3] = {1, 1, 1};
  int nb[3] = {1, 1, 1};
  std::vector<int32_t> binOff, binElems;
  // 1D bases
  double xiCL[8] = {0}, wBaryCL[8] = {0};
  double xGP[kMaxN + 1] = {0}, wGP[kMaxN + 1] = {0}, wBaryGP[kMaxN + 1] = {0};
  // cell_volweight_mean: node classes (a unique node and its periodic twins)
  std::vector<int32_t> elemClass;  // [e][8] class of each CGNS corner
  std::vector<int32_t> elemNode;   // [e][8] unique node (GlobalNodeIDs - 1) of each CGNS corner
  std::vector<int32_t> nodeOff;    // [nUnique+1] CSR over (e*8+corner) incidences of each unique node
  std::vector<int32_t> nodeInc;
  std::vector<int32_t> classOff;   // [nClasses+1] CSR over (e*8+corner) incidences
  std::vector<int32_t> classInc;
  std::vector<double> classVolume; // NodeVolume of the class
  int nShift = 1;                  // shape-function periodic cases (1, 3, 9 or 27)
  // shape function: target element -> (source element, periodic shift) list
  std::vector<int32_t> sfOff;      // [nElems+1]
  std::vector<int32_t> sfSrc;      // source element
  std::vector<uint8_t> sfShift;    // index into shift table (27 combos)
  // the same lists as runs of consecutive source elements with one shift,
  // whose particle segments are contiguous in the element-sorted arrays
  std::vector<int32_t> sfRunOff;   // [nElems+1]
  std::vector<int32_t> sfRunFirst, sfRunEnd;   // elements [first, end)
  std::vector<uint8_t> sfRunShift;
  std::vector<int32_t> sfGrpRunOff, sfGrpRunFirst, sfGrpRunEnd;   // runs of the union of 4 consecutive targets' lists
  std::vector<uint8_t> sfGrpRunShift;
  // the same pairs by source element: source -> (target element, shift)
  std::vector<int32_t> sfSrcOff, sfTgt;
  std::vector<uint8_t> sfTgtShift;
  double shift[27][3] = {{0}};
  int nPV = 0;
  double PV[3][3] = {{0}};
  // domain bounding box
  double lo[3] = {0}, hi[3] = {0};
  // RefMapping boundary tracking (ParticleBCTracking), built when not every
  // boundary is periodic: per boundary side (a HOPR side row with a BC)
  bool cartPeriodic = false;       // every boundary periodic: PeriodicMovement
  int nBC = 0;
  std::vector<int32_t> bcElemNb;   // [nBC] 0-based neighbour across the side, -1
  std::vector<uint8_t> bcKind;     // [nBC] PIC_BC_*
  std::vector<int8_t> bcPV;        // [nBC] BC_ALPHA (periodic vector id)
  std::vector<uint8_t> bcFlip;     // [nBC] flip as the tracking reads it
  std::vector<double> bz;          // [nBC][NGeo+1][NGeo+1][3] BezierControlPoints3D
  std::vector<uint8_t> sideType;   // [nBC] kSide*
  std::vector<uint8_t> bbEmpty;    // [nBC] BoundingBoxIsEmpty
  std::vector<double> sideNorm;    // [nBC][3] SideNormVec
  std::vector<double> sideDist;    // [nBC] SideDistance
  std::vector<double> slabN;       // [nBC][3][3] SideSlabNormals (slab, xyz)
  std::vector<double> slabI;       // [nBC][6] SideSlabIntervals
  std::vector<double> baseV;       // [nBC][3][3] BaseVectors0..2
  std::vector<int32_t> e2bOff;     // [nElems+1] ElemToBCSides
  std::vector<int32_t> e2bSide;    // SideBCMetrics(BCSIDE_SIDEID) as boundary side index
  std::vector<double> e2bDist;     // SideBCMetrics(BCSIDE_DISTANCE)
  std::vector<double> e2bSlab;     // [list entry][6] the side's normal slab: n, lo, hi, 0
  double bcHalo = 0;               // BC_halo_eps
};

// SideType values (piclas.h:262-266)
enum { kSidePlanarRect = 0, kSidePlanarNonRect = 1, kSideBilinear = 2, kSidePlanarCurved = 3, kSideCurved = 4 };

// Builds every table above.  Returns empty string on success, else a message.
std::string build_host_mesh(const pic_mesh_desc &m, const pic_params &p, HostMesh &hm);

}  // namespace pic
