// Device-resident solver engine: one HIP stream, device mirrors of the state
// and table arrays of one flow case, and the stage sequence of the reference
// main loop (main.cpp:1157-1245) as kernel launches.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "bcm_kernels.h"
#include "bcm_setup.h"

namespace bcm {

struct GhostSchedule {
  std::vector<GhostEntry> entries;   // grouped per cube, level-major within the cube
  std::vector<int> cube_level_off;   // per ghost cube: first level index (+1 sentinel)
  std::vector<int> level_off;        // per level: first entry index (+1 sentinel)
  int max_levels = 0;
};

// Builds the level schedule of the sequential in-place IB fill for entries
// given in reference order (minus list, then plus list).
GhostSchedule build_ghost_schedule(const std::vector<GhostEntry>& ordered);

struct Engine {
  int device = 0;
  hipStream_t stream = nullptr;
  int ncube = 0;       // cubes + 1 (slot 0 unused)
  int mpi_ncube = 0;   // cube count used by the norm scaling
  DevPhys ph{};
  double deltaT = 0.0001, deltaTau = 0.0001 / 200., e = 0.1;

  // Device arrays.
  double *U = nullptr, *U1 = nullptr, *U1q = nullptr, *U1p1 = nullptr, *U1p2 = nullptr;
  double *Rku = nullptr, *Res = nullptr, *staging = nullptr;
  int8_t* fws = nullptr;
  double* cube_size = nullptr;
  int* bc_cubes = nullptr;
  uint8_t* bc_masks = nullptr;
  int n_bc = 0;
  GhostEntry* ghost = nullptr;
  int* cube_level_off = nullptr;
  int* level_off = nullptr;
  int n_ghost_cubes = 0;
  int n_ghost_entries = 0;
  int ghost_levels = 0;
  HaloSrc* halo = nullptr;
  // Pre-flux halo refresh: (source, destination) cells of the IB ghost cells
  // that lie in a neighbour's gather source layers.
  int2* halo_pairs = nullptr;
  int n_halo_pairs = 0;
  // Set when the inner iterate was replaced from outside: the next pre-flux
  // exchange must then be the full one.
  bool halo_full_next = false;
  double* partials = nullptr;
  double* er_dev = nullptr;     // 8 doubles per inner iteration slot
  int er_slots = 0;
  double* er_host = nullptr;    // pinned mirror of er_dev
  size_t state_bytes = 0;

  // Flux kernel choice: variant 0 = one workgroup per cell plane (k_flux),
  // 1 = marching workgroups over nseg segments of each cube (k_flux_march).
  int flux_variant = 1;
  int flux_nseg = 1;
  bool flux_fuse = true;     // marching kernel with nseg 1 also applies the RK update
  bool flux_tuned = false;

  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool graph_valid = false;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_inner = 0;
};

int engine_create(int device, Engine** out, std::string& err);
// Allocates zeroed device arrays for ncube (cubes + 1) cubes.
int engine_alloc(Engine* e, int ncube, std::string& err);
int ensure_er(Engine* e, int slots, std::string& err);
void engine_destroy(Engine* e);
// Uploads a case built on the host (tables and state).
int engine_load_case(Engine* e, const Case& cs, std::string& err);
// Uploads tables only (flags, cube sizes, lists) for drop-in use.
int engine_upload_state(Engine* e, int which, const double* aos);
int engine_download_state(Engine* e, int which, double* aos);
int engine_upload_flags(Engine* e, const int* fws_aos);

// Stage pieces (asynchronous on the engine stream).
void engine_bc(Engine* e);
void engine_ghost(Engine* e);
void engine_interface(Engine* e);
void engine_flux(Engine* e, int RK);
void engine_update(Engine* e, int RK, int er_slot);
// Flux and RK update of one stage, fused into one kernel when the flux
// choice allows it (marching kernel, one workgroup per cube).
void engine_flux_update(Engine* e, int RK, int er_slot);
bool engine_fused(const Engine* e);
void engine_shift(Engine* e);
void engine_stage(Engine* e, int RK, int er_slot, bool full_pre_halo = false);
// n_steps physical steps of n_inner inner iterations; er_hist (may be null)
// receives er[1..7] per inner iteration.
int engine_run(Engine* e, int n_steps, int n_inner, double* er_hist, bool use_graph,
               std::string& err);
int engine_sync(Engine* e, std::string& err);
// Times every flux variant on the resident case and keeps the fastest.
// Returns the chosen (variant, nseg) through the pointers.
int engine_autotune(Engine* e, int* variant, int* nseg, std::string& err);

}  // namespace bcm
