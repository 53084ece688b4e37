// Device data layout and kernel launch interface of the BCM engine.
//
// State arrays live in HBM as structure-of-arrays per cube:
//   [cube][var][i][j][k], 20^3 doubles per (cube, var) plane, k fastest.
// A wavefront of 64 lanes covers four 16-cell k-rows of one i-plane, so every
// stencil load is a coalesced row segment.  Flags are int8 [cube][i][j][k].
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bcm {

constexpr int PLANE = 8000;               // 20^3 cells of one variable of one cube
constexpr int CUBE_STRIDE = 5 * PLANE;    // doubles per cube in a state array

struct DevPhys {
  double rho0, U0, Uin, V0, W0, P0, T0, mu_L;
};

// Freestream conserved state of a case: (rho0, rho0 Uin, rho0 V0, rho0 W0,
// P0/(K-1) + rho0 |V|^2 / 2).  The mixed-precision path stores the inner
// iterate as its FP32 deviation from this state.
__host__ __device__ inline void ref_state(const DevPhys& ph, double (&u)[5]) {
  u[0] = ph.rho0;
  u[1] = ph.rho0 * ph.Uin;
  u[2] = ph.rho0 * ph.V0;
  u[3] = ph.rho0 * ph.W0;
  u[4] = ph.P0 / (1.4 - 1) + 0.5 * ph.rho0 * (ph.Uin * ph.Uin + ph.V0 * ph.V0 + ph.W0 * ph.W0);
}

// Edge ghost cells (layer-1/18 lines where two ghost layers meet).  The edge
// passes of an exchange (Interface.cpp:5640-5732) set each edge cell a flux
// kernel reads to the mean of two face ghost cells of the same cube: the
// corners (j, k in {1, 18}) of interior planes i = 2..17 (i-edges), and in
// the ghost planes i = 1, 18 the border cells with one of j, k in {1, 18}
// (k-edges along j = 1, 18, j-edges along k = 1, 18).  For a cell (i, j, k)
// with i, j, k in 1..18 this returns true and the two source cells (flat
// i*400 + j*20 + k) when the cell is one of those; cube corners, which no
// flux stencil reads, and all other cells return false.
__host__ __device__ inline bool edge_sources(int i, int j, int k, int& a, int& b) {
  const bool ej = j == 1 || j == 18, ek = k == 1 || k == 18;
  const int in = i == 1 ? 2 : 17, jn = j == 1 ? 2 : 17, kn = k == 1 ? 2 : 17;
  if (i != 1 && i != 18) {
    if (!(ej && ek)) return false;
    a = i * 400 + j * 20 + kn;
    b = i * 400 + jn * 20 + k;
  } else if (ej && !ek) {
    a = i * 400 + jn * 20 + k;
    b = in * 400 + j * 20 + k;
  } else if (ek && !ej) {
    a = i * 400 + j * 20 + kn;
    b = in * 400 + j * 20 + k;
  } else {
    return false;
  }
  return true;
}

// A cell where two or three ghost layers meet (edge or corner of the padded
// cube).
__host__ __device__ inline bool edge_cell(int i, int j, int k) {
  const int g = (i < 2 || i > 17) + (j < 2 || j > 17) + (k < 2 || k > 17);
  return g >= 2;
}

// Physics of every cube of a pool that holds several independent cases:
// phys[cube_case[c]] is the case constants of cube c; cube_case[c] < 0 marks
// a cube that belongs to no case (kernels skip it).
struct PhysTab {
  const DevPhys* phys;
  const int* cube_case;
};

// One resident case in the per-case norm reduction: cubes [first, first+n)
// of the pool, norm scaling cube count mpi_ncube, norm slot and physics index.
struct CaseDesc {
  int first, n, mpi_ncube, slot;
};

// One IB ghost-fill entry in level order (see the ghost schedule in the engine).
struct GhostEntry {
  int cube;
  int16_t gi, gj, gk, pi, pj, pk;   // ghost cell and interpolation base corner
  int16_t lc[8], lg;                // LDS slots read for the corners and written
  double w[8];
};

// IB fill of one cube in LDS: value slots for the cells its entries touch
// plus one slot per entry; a cube needing more slots works on memory.
constexpr int GHOST_LDS_SLOTS = 1536;

// Device tables of the IB fill: entries grouped per cube and level, and per
// cube the list of cells its entries touch (written ghost cells first).
// In LDS (slots <= GHOST_LDS_SLOTS) every entry writes a slot of its own and
// reads the slots of the values the sequential fill would read, so levels
// follow read-after-write dependencies only; touched_final names the slot
// holding each written cell's last value.  On memory the fill is in place
// and levels also order writes after earlier reads and writes of a cell.
struct GhostTables {
  const GhostEntry* ent = nullptr;
  const int* cube_level_off = nullptr;   // per cube: first level (+1 sentinel)
  const int* level_off = nullptr;        // per level: first entry (+1 sentinel)
  const int16_t* touched = nullptr;      // cells i*400 + j*20 + k
  const int16_t* touched_final = nullptr;// slot of the last value of each written cell
  const int* touched_off = nullptr;      // per cube: first touched cell (+1 sentinel)
  const int* touched_nw = nullptr;       // per cube: written cells at the head of its list
  int ncubes = 0;
  int max_slots = 0;
};

// Source cubes for the six face halos of a cube (0 = no equal-size neighbour).
struct HaloSrc {
  int src[6];   // -X, +X, -Y, +Y, -Z, +Z
};

// One ghost region of a face across a change of cube size (two-level grid):
// dst's two ghost layers on side hi (1 upper, 0 lower) of axis `axis` from
// the neighbour src.  restrict_ 0: dst is the small cube, each ghost cell
// copies the big cube's face cell above it (BCM_Interface.cpp:1790-2021 for
// x); restrict_ 1: dst is the big cube, the 2 x 8 x 8 ghost cells of quadrant
// `slot` (1..4) take 0.125 times the sum of the 8 small cells beneath, in the
// reference's order (:2026-2600 for x; y and z alike).
struct MlTaskDev {
  int dst, src;
  int8_t axis, hi, restrict_, slot;
};

// Boundary-cube mask bit (shifted left by 1 for z-max) marking a z-BC face
// whose BC the pushing flux kernels apply to the interior planes i = 2..17
// themselves (BCM_Z_boundary_condition.cpp:47-88 reads only the plane's cell
// k = 2, or 17, which nothing changes between a stage's BCs and its flux on
// such cubes); the stage opening then does only the ghost planes.
constexpr int ZBC_IN_FLUX = 64;

// The z-min BC's w-momentum, -x, in storage form: FP64 stores the state, FP32
// its deviation from the case's reference state (r3 = rho0 W0).
__device__ inline double zbc_neg(double x, double) { return -x; }
__device__ inline float zbc_neg(float x, double r3) { return (float)(-((double)x + r3) - r3); }

// Copies, with the threads of one workgroup, the ghost regions the domain BCs
// of cube c write (mask bits 1 x-min, 2 x-max, 4 y-min, 8 y-max, 16 z-min,
// 32 z-max) from Uin to Uout: x faces over j,k in [2,17], y faces over all i
// and k in [2,17], z faces over all i,j.  The regions are disjoint and no
// neighbour writes them.
template <typename T>
__device__ inline void copy_bc_regions(const T* __restrict__ Uin, T* __restrict__ Uout, int c,
                                       int m, double r3) {
  const size_t base = (size_t)c * CUBE_STRIDE;
  auto cp = [&](int cell) {
#pragma unroll
    for (int v = 0; v < 5; ++v) Uout[base + (size_t)v * PLANE + cell] = Uin[base + (size_t)v * PLANE + cell];
  };
  for (int side = 0; side < 2; ++side) {
    const int off = side ? 18 : 0;
    if (m & (1 << side))
      for (int n = threadIdx.x; n < 512; n += blockDim.x)
        cp((off + n / 256) * 400 + ((n % 256) / 16 + 2) * 20 + (n % 16 + 2));
    if (m & (4 << side))
      for (int n = threadIdx.x; n < 640; n += blockDim.x)
        cp(((n % 320) / 16) * 400 + (off + n / 320) * 20 + (n % 16 + 2));
    if (m & (16 << side)) {
      // Planes whose z-BC this stage's flux kernel applied (ZBC_IN_FLUX):
      // the BC values, from the plane's cell k = 2 (or 17).
      const bool influx = m & (ZBC_IN_FLUX << side);
      for (int n = threadIdx.x; n < 800; n += blockDim.x) {
        const int i = (n % 400) / 20, cell = i * 400 + ((n % 400) % 20) * 20 + off + n / 400;
        if (influx && i >= 2 && i <= 17) {
          const int src = cell - (off + n / 400) + (side ? 17 : 2);
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            const T x = Uin[base + (size_t)v * PLANE + src];
            Uout[base + (size_t)v * PLANE + cell] = (v == 3 && !side) ? zbc_neg(x, r3) : x;
          }
        } else {
          cp(cell);
        }
      }
    }
  }
}

// Launches cover pool cubes 1..ncube-1; "ncube" is one past the last cube in use.
// The BC, IB fill and halo kernels exist for an FP64 inner iterate and for the
// FP32 one of the mixed-precision path, which holds the deviation from the
// case's freestream state (ref_state); arithmetic is FP64 on full states
// either way, copies move stored values as they are.
void launch_bc(hipStream_t s, double* U, const int* cubes, const uint8_t* masks, int n,
               PhysTab pt);
void launch_bc(hipStream_t s, float* U, const int* cubes, const uint8_t* masks, int n,
               PhysTab pt);
void launch_ghost(hipStream_t s, double* U, const GhostTables& g, PhysTab pt);
void launch_ghost(hipStream_t s, float* U, const GhostTables& g, PhysTab pt);
void launch_halo_edge(hipStream_t s, double* U, const HaloSrc* src, int ncube, const int* cube_case);
void launch_halo_edge(hipStream_t s, float* U, const HaloSrc* src, int ncube, const int* cube_case);
// Pre-flux refresh: copy (source, destination) cell pairs (cube*8000 + cell).
void launch_halo_pairs(hipStream_t s, double* U, const int2* pairs, int npairs);
void launch_halo_pairs(hipStream_t s, float* U, const int2* pairs, int npairs);
// Edge passes of the listed cubes.
void launch_edge_list(hipStream_t s, double* U, const int* list, int n);
void launch_edge_list(hipStream_t s, float* U, const int* list, int n);
// Edge passes of the n listed cubes (entry c; entry -c: only the k-edge cells
// the z-BC reads), then the domain BCs of those that are boundary cubes
// (cube_mask[c]: the k_bc side bits of pool cube c, 0 for none).  Used after
// a flux kernel that wrote the face halos itself.
// Opening of a pushed stage in one launch: per item (x: edge/BC entry as
// above or 0, y: IB ghost-table index or -1) a cube's edge passes and BCs,
// then its IB fill.
void launch_open(hipStream_t s, double* U, const int2* items, int n, const uint8_t* cube_mask,
                 const GhostTables& g, PhysTab pt);
void launch_open(hipStream_t s, float* U, const int2* items, int n, const uint8_t* cube_mask,
                 const GhostTables& g, PhysTab pt);
// Edge passes of every cube.
void launch_edge(hipStream_t s, double* U, int ncube, const int* cube_case);
void launch_edge(hipStream_t s, float* U, int ncube, const int* cube_case);
// Two-level faces: one workgroup per task.
void launch_halo_ml(hipStream_t s, double* U, const MlTaskDev* tasks, int ntask, PhysTab pt);
void launch_halo_ml(hipStream_t s, float* U, const MlTaskDev* tasks, int ntask, PhysTab pt);
void launch_flux(hipStream_t s, const double* U, const double* U1, const double* U1q,
                 const int8_t* fws, const double* cube_size, double* Rku, double* Res,
                 int ncube, PhysTab pt, double deltaT, double deltaTau, double e);
// Marching variant of the flux kernel: one workgroup per cube segment
// (nseg in {1, 2, 4, 8, 16} segments of the cube's 16 i-planes).  With
// fuse != 0 and nseg == 1 it also applies the RK update in place and, at
// RK 3, writes one row of 8 norm partials per cube (as k_update does per block).
void launch_flux_march(hipStream_t s, double* U, const double* U1, const double* U1q,
                       const int8_t* fws, const double* cube_size, double* Rku, double* Res,
                       int ncube, PhysTab pt, double deltaT, double deltaTau, double e, int nseg,
                       int RK, int fuse, double* U1p1, double* U1p2, double* partials);
// Marching kernel with the update written to Uout (one workgroup per cube):
// every interior cell, and the two layers next to each face with an
// equal-size neighbour (halo[c]) into that neighbour's ghost layers in Uout.
// With aux != 0 it writes Rku, Res, U1p1 (RK1) and U1p2 (RK2) bitwise as the
// in-place fused kernel does (nothing reads them back during stages); at RK 3
// one row of norm partials per cube either way.
// Ghost regions the domain BCs wrote (cube_mask) are copied from Uin to Uout.
void launch_flux_push(hipStream_t s, const double* Uin, double* Uout, const HaloSrc* halo,
                      const uint8_t* cube_mask, const double* U1, const double* U1q,
                      const int8_t* fws, const double* cube_size, double* Rku, double* Res,
                      int ncube, PhysTab pt, double deltaT, double deltaTau, double e, int RK,
                      double* U1p1, double* U1p2, double* partials, int aux);
void launch_copy_d(hipStream_t s, const double* src, double* dst, size_t n);
// Runs n threads of the shared-denominator division against a / b (four
// quotients each); *mismatches receives the count of differing bit patterns.
// Returns nonzero on a HIP error.
int div_selftest(unsigned long long seed, int n, unsigned long long* mismatches);
void launch_update(hipStream_t s, int RK, double* U, double* U1p1, double* U1p2,
                   const double* Rku, const double* Res, const int8_t* fws, int ncube,
                   const int* cube_case, double* partials);
// Norms er[0..7] of each case in cases[0..ncases): rows_per_cube partial rows
// per cube (1 after the fused FP64 flux kernels, 16 after k_update and the
// mixed flux kernel; several rows are first summed in row order into each
// cube's first row); case d writes er[d.slot*8 .. d.slot*8+7].
void launch_reduce(hipStream_t s, double* partials, int rows_per_cube, const CaseDesc* cases,
                   int ncases, const DevPhys* phys, double* er);
void launch_shift(hipStream_t s, const double* U, double* U1, double* U1q, int ncube,
                  const int* cube_case);
// AoS <-> SoA transposes of ncube whole cubes (cube 0 of both arrays included).
void launch_aos_to_soa(hipStream_t s, const double* aos, double* soa, int ncube);
void launch_soa_to_aos(hipStream_t s, const double* soa, double* aos, int ncube);
int update_blocks(int ncube);

// ---- Mixed-precision path (bcm_kernels_mixed.hip) ----
// Inner iterate U in FP32 SoA as the deviation from ref_state of the cube's
// case; U1, U1q (physical time levels) in FP64 as full states; Bsrc, the
// dual-time source constant 3 Uref - 4 U1 + U1q of every interior cell (FP64),
// which the stages read in place of U1 and U1q.
// One workgroup per cube: flux, residual, preconditioned increment and the RK
// update of stage RK, reading Uin and writing every interior cell to Uout
// (updated fluid cells, other cells unchanged).  Cells of the two layers
// next to a face with an equal-size neighbour (halo[c]) are also written into
// that neighbour's ghost layers in Uout: the face part of the post-update
// exchange.  nseg (1, 2 or 4) runs of 16/nseg planes per cube, one workgroup
// each; results are bitwise the same for every nseg.  At RK 3,
// MIXED_ROWS_PER_CUBE rows of 8 norm partials per cube.
constexpr int MIXED_ROWS_PER_CUBE = 16;
void launch_flux_mixed(hipStream_t s, const float* Uin, float* Uout, const HaloSrc* halo,
                       const uint8_t* cube_mask, const double* Bsrc, const int8_t* fws,
                       const double* cube_size, int ncube, PhysTab pt, double deltaT,
                       double deltaTau, double e, int RK, double* partials, int nseg);
// dst[0..n) <- src[0..n).
void launch_copy_f(hipStream_t s, const float* src, float* dst, size_t n);
// U1q <- U1, U1 <- ref_state + U and Bsrc <- 3 ref_state - 4 U1 + U1q on
// interior cells.
void launch_shift_mixed(hipStream_t s, const float* U, double* U1, double* U1q, double* Bsrc,
                        int ncube, PhysTab pt);
// Bsrc <- 3 ref_state - 4 U1 + U1q on the interior cells of cubes
// first..first+n-1 (after U1 or U1q were set from outside).
void launch_bsrc_mixed(hipStream_t s, const double* U1, const double* U1q, double* Bsrc, int first,
                       int n, PhysTab pt);
// FP64 AoS full states <-> FP32 SoA deviations from the reference state ref
// (5 doubles) of ncube whole cubes of one case.
void launch_aos_to_soa_f(hipStream_t s, const double* aos, float* soa, int ncube,
                         const double* ref5);
void launch_soa_to_aos_f(hipStream_t s, const float* soa, double* aos, int ncube,
                         const double* ref5);

}  // namespace bcm
