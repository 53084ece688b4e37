// Host-side case construction for the BCM solver: cube lattice, cell flags,
// immersed-boundary (IB) tables, interface tables, domain-boundary cube lists
// and the initial state, all in the reference's Array.h layouts.
#pragma once

#include <stdint.h>

#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "bcm_layout.h"

namespace bcm {

struct CaseParams {
  int lattice[3];            // cubes along x, y, z
  double cube_size;          // edge length of one cube
  double origin[3];          // lower corner of the box
  double sphere_center[3];
  double sphere_diameter;
  int ico_level;             // icosphere subdivision level (20 * 4^L triangles)
  double mach, reynolds;
  uint64_t seed;             // initial density noise seed
  double noise_amp;          // relative density noise amplitude (1e-3)
  // Optional refinement: the lattice cubes with index in [refine_lo, refine_hi)
  // along every axis are each split into 8 cubes of half the edge (a two-level
  // grid); no refinement when the box is empty on some axis.
  int refine_lo[3], refine_hi[3];
  // Optional second refinement inside the first: the half-edge cubes whose
  // index on the doubled lattice lies in [refine2_lo, refine2_hi) are split
  // again (a three-level grid).  Neighbouring cubes may differ by one level.
  int refine2_lo[3], refine2_hi[3];
};

// Immersed-boundary table of one list (minus or plus) in the 1-based flat
// format of BCM_Reading_IBM.cpp:48-75: GCindex[4n+1], IPsur[4n+1], weight[8n+1].
struct IBTable {
  int n = 0;
  std::vector<int> gcindex;   // size 4n+1
  std::vector<int> ipsur;     // size 4n+1
  std::vector<double> weight; // size 8n+1
};

// The 15 interface lists of BCM_Interface_table.cpp:20-308, 1-based.
enum IfaceList {
  L_X_EQ, L_Y_EQ, L_Z_EQ,
  L_X_BS_PLUS, L_X_SB_PLUS, L_X_BS_MINUS, L_X_SB_MINUS,
  L_Y_BS_PLUS, L_Y_SB_PLUS, L_Y_BS_MINUS, L_Y_SB_MINUS,
  L_Z_BS_PLUS, L_Z_SB_PLUS, L_Z_BS_MINUS, L_Z_SB_MINUS,
  L_COUNT
};

struct Triangle {
  double n[3];      // facet normal as written to the STL
  double v[3][3];   // vertices
};

// Allocator that leaves new elements uninitialised: the state arrays are
// written in full by the setup's parallel passes, so the first touch of
// their pages happens on the threads that fill them.
template <typename T>
struct NoInitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <typename U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  template <typename U>
  void construct(U* p) noexcept {
    ::new (static_cast<void*>(p)) U;
  }
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
  }
};
using StateVec = std::vector<double, NoInitAlloc<double>>;

struct Case {
  CaseParams p;
  Physics phys;
  int ncube = 0;          // number of cubes + 1 (index 0 unused)
  int n_wallcube = 0;     // number of wall cubes + 1
  std::vector<double> cube_size, Xcube, Ycube, Zcube;   // [ncube]
  std::vector<double> Xcnt, Ycnt, Zcnt;                  // [ncube][20]
  std::vector<int> csl;                                  // [ncube]
  std::vector<int> adj_number;                           // [ncube][5][7]
  std::vector<int> wallcube;                             // [n_wallcube]
  std::vector<int> fws;                                  // [ncube][20][20][20]
  std::vector<int> fws_grid;   // flags as read from the grid file (before setup passes)
  std::vector<int> nadj;       // [L_COUNT] list lengths
  std::vector<std::vector<int>> adj;   // [L_COUNT][ncube], 1-based lists
  std::vector<int> bc_lists[6];        // Xbc_l, Xbc_u, Ybc_l, Ybc_u, Zbc_l, Zbc_u; 1-based
  std::vector<Triangle> stl;           // all triangles in file order
  IBTable ib_minus, ib_plus;
  int n_ghost_candidates = 0;          // ghost cells examined by the IB setup
  // State arrays, AoS [ncube][20][20][20][5].
  StateVec U1_, U1, U1q;
};

// A surface triangle with the quantities of the ray test that depend on the
// triangle alone (BCM_Immersed_boundary.cpp:722-750, 1437-1495), computed on
// the host in the reference's operation order.
struct SurfaceTri {
  double v0[3], v1[3], v2[3];
  double e1[3], e2[3];  // v1 - v0, v2 - v0
  double nu[3];         // e1 x e2 (unnormalised normal)
  double dir[3];        // -nu / |nu|: the ray direction
  double pv[3];         // dir x e2
  double det, inv_det;  // e1 . pv and 1 / det
  double dd;            // dir . dir
};
void surface_tri(const Triangle& T, SurfaceTri& s);

// Device for the two quadratic passes of the IB setup.  Both must return the
// decisions of the host loops exactly:
// * filter: keep[t] = 1 when a vertex of stl[t] lies in some cube's box
//   padded by 1.5 cells (BCM_Immersed_boundary.cpp:380-470);
// * search: per ghost-cell centre xyz[3g..3g+2], hit_any[g] = 1 when some
//   triangle passes the ray test, hit_tri[g] = the first triangle of smallest
//   |t| (dir . dir) below 1e8 (or -1); for cells with no hit, near_tri[g] = the
//   last triangle of smallest squared distance to its closest point at or
//   below 1e8 (or -1).
struct IbDevice {
  virtual ~IbDevice() {}
  virtual int filter(const Case& cs, std::vector<uint8_t>& keep, std::string& err) = 0;
  virtual int search(const std::vector<SurfaceTri>& tri, const std::vector<double>& xyz,
                     std::vector<int>& hit_any, std::vector<int>& hit_tri,
                     std::vector<int>& near_tri, std::string& err) = 0;
};

// One face ghost region filled across a change of cube size, the on-rank
// two-level part of BCM_Interface (x: Interface.cpp:1790-2600, y: :2651-3470,
// z: :3518-4360).  The big cube's face meets four small cubes, one per slot:
// slot 2 covers the upper half of the first tangential axis, slot 3 the upper
// half of the second (tangential axes in x, y, z order), slot 4 both.
// * inject (restrict_ == 0): dst is the small cube; its two ghost layers on
//   side `hi` take the big cube's cell at the face (index 17 for the small
//   cube's lower side, 2 for its upper side) under each pair of ghost cells;
// * restrict (restrict_ == 1): dst is the big cube; the 2 x 8 x 8 ghost cells
//   of its quadrant on side `hi` take the mean of the 2 x 2 x 2 small cells
//   beneath, summed in the reference's order.
struct MlTask {
  int dst, src;
  int8_t axis;       // 0 x, 1 y, 2 z
  int8_t hi;         // 1: dst's upper ghost layers (18, 19); 0: lower (0, 1)
  int8_t restrict_;  // 0 inject into the small cube, 1 restrict into the big one
  int8_t slot;       // 1..4
};
// Tasks of the two-level faces in reference order, from the interface lists
// (lists[L_*_BS_PLUS], lists[L_*_BS_MINUS]) and adj_number.
std::vector<MlTask> ml_tasks(const int* const* lists, const int* nadj, const int* adj_number);
// Applies one task to an AoS state array.
void ml_apply_host(const MlTask& t, double* U);

// Builds the whole case.  Returns 0 on success, non-zero with a message in err.
// With dev, the triangle filter and the surface search of the IB setup run
// there; the case is the same either way.
int build_case(const CaseParams& p, Case& cs, std::string& err, IbDevice* dev = nullptr);
// build_case and read_case with the IB search on a GPU (bcm_ib.hip).
int build_case_on_device(const CaseParams& p, int device, Case& cs, std::string& err);
int read_case_on_device(const std::string& dir, const CaseParams& p, int device, Case& cs,
                        std::string& err);

// Cell centres and size levels from cube sizes and corners.
void grid_geometry(Case& cs);
// The setup passes that follow the grid and the surface: boundary lists,
// interface tables, flags, initial state, ghost detection, IB tables.
int finish_setup(Case& cs, std::string& err, IbDevice* dev = nullptr);

// Individual passes, exposed for tests.
void compute_physics(const CaseParams& p, Physics& ph);
void make_icosphere(int level, const double center[3], double radius,
                    std::vector<Triangle>& tris);
double noise_unit(uint64_t seed, int cube, int i, int j, int k);
void interface_tables(int ncube, const int* csl, const int* adj_number,
                      int* nadj, int* const* lists);
void host_interface_copy(int ncube, const int* const* lists, const int* nadj,
                         const int* adj_number, double* U);

// Writers for the on-disk inputs the reference reads.
int write_bcmgrid(const Case& cs, const std::string& path);
int write_stl(const Case& cs, const std::string& path);

// Case I/O (bcm_io.cpp).  Readers return 0 or non-zero with a message in err.
int read_bcmgrid(const std::string& path, Case& cs, std::string& err);
int read_stl(const std::string& path, std::vector<Triangle>& tris, std::string& err);
int read_ibm(const std::string& path, IBTable& tb, std::string& err);
// IBM_minus00000.dat and IBM_plus00000.dat in dir.
int write_ibm(const Case& cs, const std::string& dir);
// Rank-decomposition files (MPI_communication_table_in_turn.dat,
// Communication_irank.00000.dat, rank_map.dat) of a one-rank run of
// mpi_ncube cubes; the check accepts absent files and refuses a decomposition
// over several ranks.
int write_rank_tables(int mpi_ncube, const std::string& dir);
int check_rank_tables(const std::string& dir, int mpi_ncube, std::string& err);
// A case from BCMgrid00000.dat and BCM_STL_fine.stl in dir; p supplies the
// flow parameters (Mach, Reynolds number, reference length, noise).
int read_case(const std::string& dir, const CaseParams& p, Case& cs, std::string& err,
              IbDevice* dev = nullptr);
// PLOT3D grid BCMdata00001.x and solution qBCMdata<step>.q in dir for cubes
// 1..ncube-1 with cell centres Xcnt/Ycnt/Zcnt [ncube][20] and the AoS state U
// [ncube][20][20][20][5]; the headers announce mpi_ncube blocks.
int write_plot3d(int ncube, const double* Xcnt, const double* Ycnt, const double* Zcnt,
                 int mpi_ncube, const double* U, int step, const std::string& dir,
                 std::string& err);
// Restart snapshot: step, norm scales E1..E5 and U1_, U1, U1q (AoS, ncube
// slots each).  E5 may be null on write (zeros) and on read (ignored).
int write_restart(const std::string& path, int ncube, long long step, const double* E5,
                  const double* U1_, const double* U1, const double* U1q, std::string& err);
int read_restart(const std::string& path, int ncube, long long* step, double* E5, double* U1_,
                 double* U1, double* U1q, std::string& err);

}  // namespace bcm
