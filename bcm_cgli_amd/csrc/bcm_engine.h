// Device-resident solver engine: one HIP stream and a pool of cubes in HBM
// that holds one or more independent flow cases, each in a contiguous range
// of pool cubes, and the stage sequence of the reference main loop
// (main.cpp:1157-1245) as kernel launches over every cube of every case.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <vector>

#include "bcm_kernels.h"
#include "bcm_setup.h"

namespace bcm {

// Sets the message bcm_last_error() returns on the calling thread.
void set_last_error(const std::string& m);

struct GhostSchedule {
  std::vector<GhostEntry> entries;   // grouped per cube, level-major within the cube
  std::vector<int> cube_level_off;   // per ghost cube: first level index (+1 sentinel)
  std::vector<int> level_off;        // per level: first entry index (+1 sentinel)
  std::vector<int16_t> touched;      // per ghost cube: written ghost cells, then other corners
  std::vector<int16_t> touched_final;// slot of each written cell's last value
  std::vector<int> touched_off;      // per ghost cube: first touched cell (+1 sentinel)
  std::vector<int> touched_nw;       // per ghost cube: written cells
  int max_levels = 0;
  int max_slots = 0;
};

// Device copy of a schedule (owns its arrays).
struct GhostDevice {
  GhostTables t;
  std::vector<void*> owned;
};
int ghost_upload(const GhostSchedule& gs, GhostDevice& d, std::string& err);
void ghost_free(GhostDevice& d);

// Builds the level schedule of the sequential in-place IB fill for entries
// given in reference order (minus list, then plus list): per cube, levels
// over renamed LDS slots when the cube needs at most slot_limit slots
// (GHOST_LDS_SLOTS normally; lower only to exercise the other schedule),
// else in place on memory.
GhostSchedule build_ghost_schedule(const std::vector<GhostEntry>& ordered,
                                   int slot_limit = GHOST_LDS_SLOTS);

// Device tables of one case in its own cube numbering (cubes 1..n).
struct CaseTables {
  int n = 0;            // cubes
  int mpi_ncube = 0;    // cube count of the norm scaling (MPI_Ncube)
  DevPhys ph{};
  std::vector<int> bc_cubes;
  std::vector<uint8_t> bc_masks;
  GhostSchedule ghost;
  std::vector<HaloSrc> halo;       // [n + 1]
  std::vector<int2> halo_pairs;    // (source, destination) flat cells, cube * 8000 + cell
  std::vector<MlTaskDev> ml;       // two-level faces (empty on one-size grids)
};
int build_case_tables(const Case& cs, CaseTables& tb, std::string& err,
                      int ghost_slot_limit = GHOST_LDS_SLOTS);

struct CaseSlot {
  bool active = false;
  int first = 0;        // pool cube of the case's cube 1
  CaseTables tb;
  std::vector<double> centres;   // Xcnt, Ycnt, Zcnt of cubes 0..n, [3][n + 1][20]
};

struct OutputQueue;

struct Engine {
  int device = 0;
  hipStream_t stream = nullptr;
  int cap = 0;         // pool cubes allocated (slot 0 included)
  int ncube = 0;       // one past the last pool cube in use; launches cover 1..ncube-1
  double deltaT = 0.0001, deltaTau = 0.0001 / 200., e = 0.1;

  // Resident cases.
  int max_cases = 1;
  std::vector<CaseSlot> cases;
  std::vector<int> cube_case_host;    // [cap], -1 = free
  int n_active = 0;
  DevPhys* phys_dev = nullptr;        // [max_cases]
  int* cube_case = nullptr;           // [cap]
  CaseDesc* case_desc = nullptr;      // [n_active], ordered by first cube
  PhysTab pt() const { return PhysTab{phys_dev, cube_case}; }

  // Arithmetic of the pool: 0 plain FP64 (bitwise the reference), 1 mixed
  // (FP32 inner iterate and stage arithmetic, FP64 time levels, source and norms).
  int prec = 0;
  // Inner iterate of the mixed path: the current copy and the one the next
  // flux kernel writes (the two are swapped after every flux kernel).
  float* Uf = nullptr;
  float* Uf2 = nullptr;
  uint8_t* bc_mask_cube = nullptr;   // [cap] BC side bits per pool cube

  // Device arrays (SoA per cube, cap cubes each).  The mixed path keeps U1,
  // U1q, staging, flags and sizes; U, U1p1, U1p2, Rku and Res stay null.
  double *U = nullptr, *U1 = nullptr, *U1q = nullptr, *U1p1 = nullptr, *U1p2 = nullptr;
  double *Rku = nullptr, *Res = nullptr, *staging = nullptr;
  // Mixed path: the dual-time source constant 3 Uref - 4 U1 + U1q of every
  // interior cell (FP64), formed when U1 and U1q change (the physical-time
  // shift, uploads) instead of in every stage.
  double* Bsrc = nullptr;
  double* U2 = nullptr;         // second copy of the FP64 iterate (flux with halo push)
  int8_t* fws = nullptr;
  double* cube_size = nullptr;
  // Tables of all resident cases in pool numbering.
  int* bc_cubes = nullptr;
  uint8_t* bc_masks = nullptr;
  int n_bc = 0;
  GhostDevice ghost;
  int n_ghost_cubes = 0;
  int n_ghost_entries = 0;
  int ghost_levels = 0;
  HaloSrc* halo = nullptr;
  // Pre-flux halo refresh: (source, destination) cells of the IB ghost cells
  // that lie in a neighbour's gather source layers.
  int2* halo_pairs = nullptr;
  int n_halo_pairs = 0;
  // Two-level faces of all resident cases, refreshed after the equal-size
  // faces and before the edge passes wherever an exchange runs.
  MlTaskDev* ml = nullptr;
  int n_ml = 0;
  // Cubes whose face ghosts can change between a stage's opening edge pass
  // and its pre-flux exchange (BC, IB-refresh and two-level destinations):
  // the pre-flux edge pass covers only these.
  int* edge_list = nullptr;
  int n_edge_list = 0;
  // Opening of a stage after a flux kernel that pushed its face halos, one
  // item per cube that needs one: (edge entry, ghost-table index or -1), IB
  // cubes first.  Edge entry c: all edge passes, then BCs (the cube's IB fill
  // touches an edge cell); -c: the k-edge cells the z-BC copies, then BCs
  // (other boundary cubes); 0: none.  Elsewhere only the pushing flux kernels
  // read edge cells, and they form them themselves.
  int2* open_items = nullptr;
  int n_open_items = 0;
  // Set when an inner iterate was replaced from outside (state upload, new
  // case): the next pre-flux exchange must then be the full one.
  bool halo_full_next = false;
  double* partials = nullptr;
  double* er_dev = nullptr;     // [er_slots][max_cases][8]
  int er_slots = 0;
  double* er_host = nullptr;    // pinned mirror of er_dev

  // Flux kernel choice: variant 0 = one workgroup per cell plane (k_flux),
  // 1 = marching workgroups over nseg segments of each cube (k_flux_march).
  int flux_variant = 1;
  int flux_nseg = 1;
  bool flux_fuse = true;     // marching kernel with nseg 1 also applies the RK update
  // Stages run the marching kernel that writes the update and the face halos
  // into the other copy of the iterate (engine stages only; the stage pieces
  // and drop-ins keep the in-place kernels).
  bool flux_push = true;
  // The pushing FP64 kernel stores Rku1, U1p1, U1p2 and Residual1 (which no
  // stage reads back) only when set: captured steps clear it for all but
  // their last inner iteration.
  bool push_aux = true;
  // Mixed path: runs per cube of the flux kernel (1, 2 or 4); the autotuner
  // picks it per size class of the resident pool (log2 of its cubes) and
  // keeps each class's pick.
  int mixed_nseg = 1;
  int tuned_class = -1;
  std::map<int, int> nseg_by_class;
  bool flux_tuned = false;
  bool flux_pinned = false;    // chosen with set_flux: no re-tuning
  // Cubes whose IB fill needs more LDS slots than this take the in-place
  // schedule on memory (GHOST_LDS_SLOTS unless lowered to test that path).
  int ghost_slot_limit = GHOST_LDS_SLOTS;

  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  OutputQueue* out = nullptr;   // asynchronous output (created on first use)
  bool graph_valid = false;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_inner = 0;
};

// Bytes of device memory per pool cube for a precision (0 FP64, 1 mixed).
size_t engine_bytes_per_cube(int prec);

int engine_create(int device, Engine** out, std::string& err);
// Allocates a pool of cap cubes (slot 0 included) for up to max_cases cases
// in precision prec (0 FP64, 1 mixed).
int engine_alloc(Engine* e, int cap, int max_cases, std::string& err, int prec = 0);
int ensure_er(Engine* e, int slots, std::string& err);
void engine_destroy(Engine* e);
// Single-case use: a pool of exactly cs.ncube cubes holding cs in slot 0.
int engine_load_case(Engine* e, const Case& cs, std::string& err);
// Adds a case to the pool in the first free range of its size; the slot
// index it gets is returned through slot.
int engine_add_case(Engine* e, const Case& cs, int* slot, std::string& err);
int engine_remove_case(Engine* e, int slot, std::string& err);
// Largest free contiguous range of pool cubes.
int engine_free_range(const Engine* e);
// Registers cubes 1..ncube-1 as one case with no tables (drop-in binding).
int engine_bind_single(Engine* e, int mpi_ncube, const DevPhys& ph, std::string& err);
// Whole-pool state transfers, AoS layout of cubes 0..ncube-1.
int engine_upload_state(Engine* e, int which, const double* aos);
int engine_download_state(Engine* e, int which, double* aos);
int engine_upload_flags(Engine* e, const int* fws_aos);
// One case's state, AoS layout of its cubes 0..n (cube 0 is zeroed on download
// and ignored on upload).
int engine_case_state(Engine* e, int slot, int which, double* aos, bool upload,
                      std::string& err);

// Stage pieces (asynchronous on the engine stream).
void engine_bc(Engine* e);
void engine_ghost(Engine* e);
void engine_interface(Engine* e);
void engine_halo_refresh(Engine* e);
void engine_flux(Engine* e, int RK);
void engine_update(Engine* e, int RK, int er_slot);
// Flux and RK update of one stage, fused into one kernel when the flux
// choice allows it (marching kernel, one workgroup per cube).
void engine_flux_update(Engine* e, int RK, int er_slot);
bool engine_fused(const Engine* e);
// Flux and update of one stage written to the other copy of the iterate with
// the face halos (the form stages use when flux_push is set); swaps copies.
void engine_flux_push(Engine* e, int RK, int er_slot);
void engine_shift(Engine* e);
void engine_stage(Engine* e, int RK, int er_slot, bool full_pre_halo = false);
// n_steps physical steps of n_inner inner iterations for every resident case.
// er_hist (may be null) receives er[1..7] as [step][inner][max_cases][7].
int engine_run(Engine* e, int n_steps, int n_inner, double* er_hist, bool use_graph,
               std::string& err);
// n_steps physical steps under the reference's inner-iteration control
// (main.cpp:1249-1320): per case, after each inner iteration E[m] =
// max(E[m], er[m]) and the step's inner loop ends when every er[m] / E[m]
// (m = 1..5) is below tol or after max_inner iterations; then the case's
// time levels shift.  E: [max_cases][5] running maxima, read and updated.
// iters: [n_steps][max_cases] inner iterations run (0 for empty slots).
// er_hist (may be null): er[1..7] as [step][inner][max_cases][7], rows after
// a case's break untouched.
int engine_run_converge(Engine* e, int n_steps, int max_inner, double tol, double* E, int* iters,
                        double* er_hist, std::string& err);
int engine_sync(Engine* e, std::string& err);
// Times one RK stage for every flux choice on the resident cases (state
// restored afterwards) and keeps the fastest.  Returns the chosen
// (variant, nseg) through the pointers.
int engine_autotune(Engine* e, int* variant, int* nseg, std::string& err);
// Tunes only when no choice was made yet, or (mixed path, choice not pinned
// with set_flux) when the pool's size class has changed since the last tuning.
int engine_maybe_tune(Engine* e, std::string& err);

// Asynchronous PLOT3D output (bcm_output.cpp): queues the current inner
// iterate U1_ of a case for writing as BCMdata00001.x and qBCMdata<step>.q in
// dir.  The solver stream only pays a layout conversion; the copy to the host
// and the file writes overlap later work.  Errors of the writes are reported
// by engine_output_wait, which also waits for every queued output.
int engine_output(Engine* e, int slot, int step, const std::string& dir, std::string& err);
int engine_output_wait(Engine* e, std::string& err);
void engine_output_free(Engine* e);

}  // namespace bcm
