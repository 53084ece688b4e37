// Device-resident BCM engine: memory, tables, and the reference stage order.
#include "bcm_engine.h"

#include <string.h>

#include <algorithm>
#include <map>

namespace bcm {

namespace {

#define HIP_TRY(x)                                                             \
  do {                                                                         \
    hipError_t _e = (x);                                                       \
    if (_e != hipSuccess) {                                                    \
      err = std::string(#x) + ": " + hipGetErrorString(_e);                    \
      return 1;                                                                \
    }                                                                          \
  } while (0)

template <typename T>
int dev_alloc(T** p, size_t n, std::string& err) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIP_TRY(hipMalloc((void**)p, n * sizeof(T)));
  return 0;
}

template <typename T>
int dev_upload(T** p, const std::vector<T>& v, std::string& err) {
  if (dev_alloc(p, v.size(), err)) return 1;
  if (!v.empty()) HIP_TRY(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

double* state_ptr(Engine* e, int which) {
  switch (which) {
    case 0: return e->U;
    case 1: return e->U1;
    case 2: return e->U1q;
    case 3: return e->U1p1;
    case 4: return e->U1p2;
    case 5: return e->Rku;
    case 6: return e->Res;
  }
  return nullptr;
}

void free_graph(Engine* e) {
  if (e->graph_exec) hipGraphExecDestroy(e->graph_exec);
  if (e->graph) hipGraphDestroy(e->graph);
  e->graph_exec = nullptr;
  e->graph = nullptr;
  e->graph_valid = false;
}

void free_case(Engine* e) {
  free_graph(e);
  void* ptrs[] = {e->U, e->U1, e->U1q, e->U1p1, e->U1p2, e->Rku, e->Res, e->staging,
                  e->fws, e->cube_size, e->bc_cubes, e->bc_masks, e->ghost,
                  e->cube_level_off, e->level_off, e->halo, e->partials, e->er_dev,
                  e->halo_pairs};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (e->er_host) hipHostFree(e->er_host);
  e->U = e->U1 = e->U1q = e->U1p1 = e->U1p2 = e->Rku = e->Res = e->staging = nullptr;
  e->fws = nullptr;
  e->cube_size = nullptr;
  e->bc_cubes = nullptr;
  e->bc_masks = nullptr;
  e->ghost = nullptr;
  e->cube_level_off = e->level_off = nullptr;
  e->halo = nullptr;
  e->halo_pairs = nullptr;
  e->n_halo_pairs = 0;
  e->partials = e->er_dev = e->er_host = nullptr;
  e->er_slots = 0;
  e->ncube = 0;
}

}  // namespace

int ensure_er(Engine* e, int slots, std::string& err);

// ---------------------------------------------------------------------------
GhostSchedule build_ghost_schedule(const std::vector<GhostEntry>& ordered) {
  GhostSchedule gs;
  // Group by cube, keeping reference order inside each cube.
  std::map<int, std::vector<int>> by_cube;
  for (int n = 0; n < (int)ordered.size(); ++n) by_cube[ordered[n].cube].push_back(n);
  std::vector<int> wlev(8000), rlev(8000);
  auto cell = [](int i, int j, int k) { return i * 400 + j * 20 + k; };
  for (auto& kv : by_cube) {
    std::fill(wlev.begin(), wlev.end(), 0);
    std::fill(rlev.begin(), rlev.end(), 0);
    const std::vector<int>& idx = kv.second;
    std::vector<int> level(idx.size());
    int nlev = 0;
    for (size_t a = 0; a < idx.size(); ++a) {
      const GhostEntry& E = ordered[idx[a]];
      int lv = 1;
      for (int m = 0; m < 8; ++m) {
        int ci = cell(E.pi + (m == 1 || m == 4 || m == 5 || m == 7),
                      E.pj + (m == 2 || m == 4 || m == 6 || m == 7),
                      E.pk + (m == 3 || m == 5 || m == 6 || m == 7));
        lv = std::max(lv, wlev[ci] + 1);   // read after an earlier write
      }
      int g = cell(E.gi, E.gj, E.gk);
      lv = std::max(lv, rlev[g] + 1);      // write after an earlier read
      lv = std::max(lv, wlev[g] + 1);      // write after an earlier write
      level[a] = lv;
      wlev[g] = lv;
      for (int m = 0; m < 8; ++m) {
        int ci = cell(E.pi + (m == 1 || m == 4 || m == 5 || m == 7),
                      E.pj + (m == 2 || m == 4 || m == 6 || m == 7),
                      E.pk + (m == 3 || m == 5 || m == 6 || m == 7));
        rlev[ci] = std::max(rlev[ci], lv);
      }
      nlev = std::max(nlev, lv);
    }
    gs.cube_level_off.push_back((int)gs.level_off.size());
    for (int lv = 1; lv <= nlev; ++lv) {
      gs.level_off.push_back((int)gs.entries.size());
      for (size_t a = 0; a < idx.size(); ++a)
        if (level[a] == lv) gs.entries.push_back(ordered[idx[a]]);
    }
    gs.max_levels = std::max(gs.max_levels, nlev);
  }
  gs.cube_level_off.push_back((int)gs.level_off.size());
  gs.level_off.push_back((int)gs.entries.size());
  return gs;
}

int engine_create(int device, Engine** out, std::string& err) {
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    err = "no such HIP device";
    return 1;
  }
  HIP_TRY(hipSetDevice(device));
  Engine* e = new Engine();
  e->device = device;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess) {
    err = "stream/event creation failed";
    delete e;
    return 1;
  }
  *out = e;
  return 0;
}

void engine_destroy(Engine* e) {
  if (!e) return;
  hipSetDevice(e->device);
  if (e->stream) hipStreamSynchronize(e->stream);
  free_case(e);
  if (e->ev0) hipEventDestroy(e->ev0);
  if (e->ev1) hipEventDestroy(e->ev1);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
}

int engine_alloc(Engine* e, int ncube, std::string& err) {
  HIP_TRY(hipSetDevice(e->device));
  free_case(e);
  if (ncube < 2) {
    err = "a case needs at least one cube";
    return 1;
  }
  e->ncube = ncube;
  e->mpi_ncube = ncube - 1;
  const size_t ncells = (size_t)ncube * PLANE;
  e->state_bytes = ncells * 5 * sizeof(double);
  for (double** p : {&e->U, &e->U1, &e->U1q, &e->U1p1, &e->U1p2, &e->Rku, &e->Res, &e->staging}) {
    if (dev_alloc(p, ncells * 5, err)) return 1;
    HIP_TRY(hipMemsetAsync(*p, 0, e->state_bytes, e->stream));
  }
  if (dev_alloc(&e->fws, ncells, err)) return 1;
  HIP_TRY(hipMemsetAsync(e->fws, 0, ncells, e->stream));
  if (dev_alloc(&e->cube_size, (size_t)ncube, err)) return 1;
  HIP_TRY(hipMemsetAsync(e->cube_size, 0, ncube * sizeof(double), e->stream));
  if (dev_alloc(&e->partials, (size_t)update_blocks(ncube) * 8, err)) return 1;
  if (ensure_er(e, 1, err)) return 1;
  HIP_TRY(hipStreamSynchronize(e->stream));
  return 0;
}

int engine_load_case(Engine* e, const Case& cs, std::string& err) {
  if (engine_alloc(e, cs.ncube, err)) return 1;
  e->flux_tuned = false;
  e->ph = DevPhys{cs.phys.rho0, cs.phys.U0, cs.phys.Uin, cs.phys.V0,
                  cs.phys.W0,   cs.phys.P0, cs.phys.T0,  cs.phys.mu_L};
  if (engine_upload_flags(e, cs.fws.data())) {
    err = "flag upload failed";
    return 1;
  }
  HIP_TRY(hipMemcpy(e->cube_size, cs.cube_size.data(), cs.ncube * sizeof(double),
                    hipMemcpyHostToDevice));
  // Boundary cubes with their side masks.
  {
    std::vector<uint8_t> mask(cs.ncube, 0);
    for (int s = 0; s < 6; ++s)
      for (size_t n = 1; n < cs.bc_lists[s].size(); ++n) mask[cs.bc_lists[s][n]] |= (uint8_t)(1 << s);
    std::vector<int> cubes;
    std::vector<uint8_t> m;
    for (int c = 1; c < cs.ncube; ++c)
      if (mask[c]) {
        cubes.push_back(c);
        m.push_back(mask[c]);
      }
    e->n_bc = (int)cubes.size();
    if (dev_upload(&e->bc_cubes, cubes, err) || dev_upload(&e->bc_masks, m, err)) return 1;
  }
  // IB ghost schedule: minus list, then plus list, each in list order.
  {
    std::vector<GhostEntry> ordered;
    for (const IBTable* tb : {&cs.ib_minus, &cs.ib_plus})
      for (int n = 1; n <= tb->n; ++n) {
        GhostEntry E;
        const int* g = &tb->gcindex[(n - 1) * 4 + 1];
        const int* q = &tb->ipsur[(n - 1) * 4 + 1];
        if (g[0] != q[0]) {
          err = "IB entry interpolates from another cube";
          return 1;
        }
        if (q[1] < 0 || q[1] > 18 || q[2] < 0 || q[2] > 18 || q[3] < 0 || q[3] > 18 ||
            g[1] < 2 || g[1] > 17 || g[2] < 2 || g[2] > 17 || g[3] < 2 || g[3] > 17 || g[0] < 1 ||
            g[0] >= cs.ncube) {
          err = "IB entry out of range";
          return 1;
        }
        E.cube = g[0];
        E.gi = (int16_t)g[1]; E.gj = (int16_t)g[2]; E.gk = (int16_t)g[3];
        E.pi = (int16_t)q[1]; E.pj = (int16_t)q[2]; E.pk = (int16_t)q[3];
        for (int m = 0; m < 8; ++m) E.w[m] = tb->weight[(n - 1) * 8 + 1 + m];
        ordered.push_back(E);
      }
    GhostSchedule gs = build_ghost_schedule(ordered);
    e->n_ghost_cubes = (int)gs.cube_level_off.size() - 1;
    e->n_ghost_entries = (int)gs.entries.size();
    e->ghost_levels = gs.max_levels;
    if (dev_upload(&e->ghost, gs.entries, err) || dev_upload(&e->cube_level_off, gs.cube_level_off, err) ||
        dev_upload(&e->level_off, gs.level_off, err))
      return 1;
  }
  // Face halo sources from the equal-size interface lists.
  {
    std::vector<HaloSrc> hs(cs.ncube);
    for (auto& h : hs) memset(&h, 0, sizeof h);
    const int dirs[3] = {2, 4, 6};
    for (int d = 0; d < 3; ++d)
      for (int a = 1; a <= cs.nadj[d]; ++a) {
        int c0 = cs.adj[d][a];
        int c2 = cs.adj_number[adj_index(c0, 1, dirs[d])];
        if (c2 <= 0 || c2 >= cs.ncube) {
          err = "interface list names a missing neighbour";
          return 1;
        }
        hs[c2].src[2 * d] = c0;      // c2's minus face from c0's top layers
        hs[c0].src[2 * d + 1] = c2;  // c0's plus face from c2's bottom layers
      }
    if (dev_upload(&e->halo, hs, err)) return 1;
    // Cells of the IB lists that some neighbour gathers: the only interior
    // cells that change between a post-update exchange and the next stage's
    // pre-flux exchange.
    std::vector<char> is_ghost((size_t)cs.ncube * PLANE, 0);
    for (const IBTable* tb : {&cs.ib_minus, &cs.ib_plus})
      for (int n = 1; n <= tb->n; ++n) {
        const int* g = &tb->gcindex[(n - 1) * 4 + 1];
        is_ghost[(size_t)g[0] * PLANE + g[1] * 400 + g[2] * 20 + g[3]] = 1;
      }
    std::vector<int2> pairs;
    for (int n = 1; n < cs.ncube; ++n)
      for (int f = 0; f < 6; ++f) {
        const int sc = hs[n].src[f];
        if (sc <= 0) continue;
        const int ax = f >> 1;
        const bool plus = f & 1;
        for (int l = 0; l < 2; ++l)
          for (int p = ILO; p <= IHI; ++p)
            for (int r = ILO; r <= IHI; ++r) {
              const int dl = plus ? 18 + l : l, sl = plus ? 2 + l : 16 + l;
              const int si = ax == 0 ? sl : p, sj = ax == 1 ? sl : (ax == 0 ? p : r),
                        sk = ax == 2 ? sl : r;
              const int di = ax == 0 ? dl : p, dj = ax == 1 ? dl : (ax == 0 ? p : r),
                        dk = ax == 2 ? dl : r;
              const int scell = si * 400 + sj * 20 + sk;
              if (!is_ghost[(size_t)sc * PLANE + scell]) continue;
              pairs.push_back(make_int2(sc * PLANE + scell, n * PLANE + di * 400 + dj * 20 + dk));
            }
      }
    e->n_halo_pairs = (int)pairs.size();
    if (dev_upload(&e->halo_pairs, pairs, err)) return 1;
    e->halo_full_next = false;
  }
  // State upload.
  const std::vector<double>* st[3] = {&cs.U1_, &cs.U1, &cs.U1q};
  for (int w = 0; w < 3; ++w)
    if (engine_upload_state(e, w, st[w]->data())) {
      err = "state upload failed";
      return 1;
    }
  HIP_TRY(hipStreamSynchronize(e->stream));
  return 0;
}

int engine_upload_state(Engine* e, int which, const double* aos) {
  double* d = state_ptr(e, which);
  if (!d) return 1;
  if (which == 0) e->halo_full_next = true;
  if (hipMemcpyAsync(e->staging, aos, e->state_bytes, hipMemcpyHostToDevice, e->stream) != hipSuccess)
    return 1;
  launch_aos_to_soa(e->stream, e->staging, d, e->ncube);
  return hipStreamSynchronize(e->stream) == hipSuccess ? 0 : 1;
}

int engine_download_state(Engine* e, int which, double* aos) {
  double* d = state_ptr(e, which);
  if (!d) return 1;
  launch_soa_to_aos(e->stream, d, e->staging, e->ncube);
  if (hipMemcpyAsync(aos, e->staging, e->state_bytes, hipMemcpyDeviceToHost, e->stream) != hipSuccess)
    return 1;
  return hipStreamSynchronize(e->stream) == hipSuccess ? 0 : 1;
}

int engine_upload_flags(Engine* e, const int* fws_aos) {
  std::vector<int8_t> f((size_t)e->ncube * PLANE);
  for (size_t n = 0; n < f.size(); ++n) f[n] = (int8_t)fws_aos[n];
  if (hipMemcpyAsync(e->fws, f.data(), f.size(), hipMemcpyHostToDevice, e->stream) != hipSuccess)
    return 1;
  return hipStreamSynchronize(e->stream) == hipSuccess ? 0 : 1;
}

void engine_bc(Engine* e) { launch_bc(e->stream, e->U, e->bc_cubes, e->bc_masks, e->n_bc, e->ph); }

void engine_ghost(Engine* e) {
  launch_ghost(e->stream, e->U, e->ghost, e->cube_level_off, e->level_off, nullptr,
               e->n_ghost_cubes, e->ph);
}

void engine_interface(Engine* e) { launch_halo_edge(e->stream, e->U, e->halo, e->ncube); }

void engine_flux(Engine* e, int RK) {
  if (e->flux_variant == 0)
    launch_flux(e->stream, e->U, e->U1, e->U1q, e->fws, e->cube_size, e->Rku, e->Res, e->ncube,
                e->ph, e->deltaT, e->deltaTau, e->e);
  else
    launch_flux_march(e->stream, e->U, e->U1, e->U1q, e->fws, e->cube_size, e->Rku, e->Res,
                      e->ncube, e->ph, e->deltaT, e->deltaTau, e->e, e->flux_nseg, RK, 0,
                      e->U1p1, e->U1p2, e->partials);
}

bool engine_fused(const Engine* e) {
  return e->flux_variant == 1 && e->flux_nseg == 1 && e->flux_fuse;
}

void engine_flux_update(Engine* e, int RK, int er_slot) {
  if (!engine_fused(e)) {
    engine_flux(e, RK);
    engine_update(e, RK, er_slot);
    return;
  }
  launch_flux_march(e->stream, e->U, e->U1, e->U1q, e->fws, e->cube_size, e->Rku, e->Res,
                    e->ncube, e->ph, e->deltaT, e->deltaTau, e->e, 1, RK, 1, e->U1p1, e->U1p2,
                    e->partials);
  if (RK == 3)
    launch_reduce(e->stream, e->partials, e->ncube - 1, e->er_dev + 8 * er_slot, e->mpi_ncube,
                  e->ph);
}

int engine_autotune(Engine* e, int* variant, int* nseg, std::string& err) {
  HIP_TRY(hipSetDevice(e->device));
  if (e->ncube <= 1) {
    err = "no case loaded";
    return 1;
  }
  // Candidates (variant, nseg, fuse): the plane kernel, the marching kernel
  // with 1..8 segments, and the marching kernel with the RK update fused.
  // Each is timed as flux + update of one RK stage on the resident case; the
  // arrays the update writes are saved first and restored afterwards.
  const int cand[6][3] = {{0, 1, 0}, {1, 1, 1}, {1, 1, 0}, {1, 2, 0}, {1, 4, 0}, {1, 8, 0}};
  double* saved[3] = {nullptr, nullptr, nullptr};
  double* live[3] = {e->U, e->U1p1, e->U1p2};
  for (int a = 0; a < 3; ++a) {
    HIP_TRY(hipMalloc((void**)&saved[a], e->state_bytes));
    HIP_TRY(hipMemcpyAsync(saved[a], live[a], e->state_bytes, hipMemcpyDeviceToDevice, e->stream));
  }
  float best = 1e30f;
  int bv = e->flux_variant, bs = e->flux_nseg;
  bool bf = e->flux_fuse;
  for (auto& cv : cand) {
    e->flux_variant = cv[0];
    e->flux_nseg = cv[1];
    e->flux_fuse = cv[2] != 0;
    engine_flux_update(e, 2, 0);  // warm up
    HIP_TRY(hipEventRecord(e->ev0, e->stream));
    for (int r = 0; r < 3; ++r) engine_flux_update(e, 2, 0);
    HIP_TRY(hipEventRecord(e->ev1, e->stream));
    HIP_TRY(hipEventSynchronize(e->ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
    if (ms < best) {
      best = ms;
      bv = cv[0];
      bs = cv[1];
      bf = cv[2] != 0;
    }
  }
  for (int a = 0; a < 3; ++a) {
    HIP_TRY(hipMemcpyAsync(live[a], saved[a], e->state_bytes, hipMemcpyDeviceToDevice, e->stream));
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  for (int a = 0; a < 3; ++a) hipFree(saved[a]);
  e->flux_variant = bv;
  e->flux_nseg = bs;
  e->flux_fuse = bf;
  e->flux_tuned = true;
  e->graph_valid = false;
  if (variant) *variant = bv;
  if (nseg) *nseg = bs;
  return 0;
}

void engine_update(Engine* e, int RK, int er_slot) {
  launch_update(e->stream, RK, e->U, e->U1p1, e->U1p2, e->Rku, e->Res, e->fws, e->ncube,
                e->partials);
  if (RK == 3)
    launch_reduce(e->stream, e->partials, update_blocks(e->ncube), e->er_dev + 8 * er_slot,
                  e->mpi_ncube, e->ph);
}

void engine_shift(Engine* e) { launch_shift(e->stream, e->U, e->U1, e->U1q, e->ncube); }

void engine_stage(Engine* e, int RK, int er_slot, bool full_pre_halo) {
  engine_bc(e);
  engine_ghost(e);
  if (full_pre_halo || !e->halo_pairs)
    engine_interface(e);
  else
    launch_halo_sparse(e->stream, e->U, e->halo_pairs, e->n_halo_pairs, e->ncube);
  engine_flux_update(e, RK, er_slot);
  engine_interface(e);
}

int ensure_er(Engine* e, int slots, std::string& err) {
  if (e->er_slots >= slots) return 0;
  if (e->er_dev) hipFree(e->er_dev);
  if (e->er_host) hipHostFree(e->er_host);
  e->er_dev = nullptr;
  e->er_host = nullptr;
  free_graph(e);
  HIP_TRY(hipMalloc((void**)&e->er_dev, (size_t)slots * 8 * sizeof(double)));
  HIP_TRY(hipHostMalloc((void**)&e->er_host, (size_t)slots * 8 * sizeof(double)));
  e->er_slots = slots;
  return 0;
}

namespace {

void enqueue_step(Engine* e, int n_inner, bool full_first) {
  for (int it = 0; it < n_inner; ++it)
    for (int RK = 1; RK <= 3; ++RK) engine_stage(e, RK, it, full_first && it == 0 && RK == 1);
  engine_shift(e);
}

}  // namespace

int engine_run(Engine* e, int n_steps, int n_inner, double* er_hist, bool use_graph,
               std::string& err) {
  HIP_TRY(hipSetDevice(e->device));
  if (e->ncube <= 1) {
    err = "no case loaded";
    return 1;
  }
  if (ensure_er(e, n_inner, err)) return 1;
  if (!e->flux_tuned) {
    int v, n;
    if (engine_autotune(e, &v, &n, err)) return 1;
  }
  if (use_graph && (!e->graph_valid || e->graph_inner != n_inner)) {
    free_graph(e);
    HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(e, n_inner, false);
    HIP_TRY(hipStreamEndCapture(e->stream, &e->graph));
    HIP_TRY(hipGraphInstantiate(&e->graph_exec, e->graph, nullptr, nullptr, 0));
    e->graph_valid = true;
    e->graph_inner = n_inner;
  }
  for (int s = 0; s < n_steps; ++s) {
    if (e->halo_full_next) {
      // The inner iterate was set from outside: this step starts with the
      // full pre-flux exchange.
      enqueue_step(e, n_inner, true);
      e->halo_full_next = false;
    } else if (use_graph) {
      HIP_TRY(hipGraphLaunch(e->graph_exec, e->stream));
    } else {
      enqueue_step(e, n_inner, false);
    }
    if (er_hist) {
      HIP_TRY(hipMemcpyAsync(e->er_host, e->er_dev, (size_t)n_inner * 8 * sizeof(double),
                             hipMemcpyDeviceToHost, e->stream));
      HIP_TRY(hipStreamSynchronize(e->stream));
      for (int it = 0; it < n_inner; ++it)
        for (int m = 0; m < 7; ++m)
          er_hist[((size_t)s * n_inner + it) * 7 + m] = e->er_host[it * 8 + 1 + m];
    }
  }
  HIP_TRY(hipGetLastError());
  return 0;
}

int engine_sync(Engine* e, std::string& err) {
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace bcm
