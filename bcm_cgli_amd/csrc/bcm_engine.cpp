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
  if (e->graph_exec) (void)hipGraphExecDestroy(e->graph_exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->graph_exec = nullptr;
  e->graph = nullptr;
  e->graph_valid = false;
}

// Tables of the resident cases (rebuilt whenever a case comes or goes).
void free_tables(Engine* e) {
  void* ptrs[] = {e->bc_cubes,     e->bc_masks, e->halo,     e->halo_pairs,
                  e->case_desc,    e->bc_mask_cube, e->ml, e->edge_list, e->open_items};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  ghost_free(e->ghost);
  e->bc_cubes = nullptr;
  e->bc_masks = nullptr;
  e->bc_mask_cube = nullptr;
  e->halo = nullptr;
  e->halo_pairs = nullptr;
  e->case_desc = nullptr;
  e->ml = nullptr;
  e->edge_list = nullptr;
  e->open_items = nullptr;
  e->n_halo_pairs = e->n_bc = e->n_ghost_cubes = e->n_ghost_entries = e->n_ml = 0;
  e->n_edge_list = e->n_open_items = 0;
}

void free_case(Engine* e) {
  free_graph(e);
  free_tables(e);
  void* ptrs[] = {e->U,   e->U1,        e->U1q,      e->U1p1,     e->U1p2,   e->Rku,
                  e->Res, e->staging,   e->fws,      e->cube_size, e->partials, e->er_dev,
                  e->phys_dev, e->cube_case, e->U2, e->Bsrc};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (e->Uf) (void)hipFree(e->Uf);
  if (e->Uf2) (void)hipFree(e->Uf2);
  e->Uf = e->Uf2 = nullptr;
  if (e->er_host) (void)hipHostFree(e->er_host);
  e->U = e->U1 = e->U1q = e->U1p1 = e->U1p2 = e->Rku = e->Res = e->staging = e->U2 = nullptr;
  e->Bsrc = nullptr;
  e->fws = nullptr;
  e->cube_size = nullptr;
  e->partials = e->er_dev = e->er_host = nullptr;
  e->phys_dev = nullptr;
  e->cube_case = nullptr;
  e->er_slots = 0;
  e->ncube = 0;
  e->cap = 0;
  e->n_active = 0;
  e->cases.clear();
  e->cube_case_host.clear();
}

}  // namespace

int ensure_er(Engine* e, int slots, std::string& err);
int rebuild_tables(Engine* e, std::string& err);

// ---------------------------------------------------------------------------
GhostSchedule build_ghost_schedule(const std::vector<GhostEntry>& ordered, int slot_limit) {
  GhostSchedule gs;
  // Group by cube, keeping reference order inside each cube.
  std::map<int, std::vector<int>> by_cube;
  for (int n = 0; n < (int)ordered.size(); ++n) by_cube[ordered[n].cube].push_back(n);
  auto cell = [](int i, int j, int k) { return i * 400 + j * 20 + k; };
  auto corner = [&](const GhostEntry& E, int m) {
    return cell(E.pi + (m == 1 || m == 4 || m == 5 || m == 7),
                E.pj + (m == 2 || m == 4 || m == 6 || m == 7),
                E.pk + (m == 3 || m == 5 || m == 6 || m == 7));
  };
  std::vector<int> pos(8000), cur(8000), wlev(8000), rlev(8000);
  for (auto& kv : by_cube) {
    const std::vector<int>& idx = kv.second;
    const int ne = (int)idx.size();
    // Touched cells: written ghost cells first, then the other corners.
    std::fill(pos.begin(), pos.end(), -1);
    std::vector<int16_t> list;
    for (int a = 0; a < ne; ++a) {
      const GhostEntry& E = ordered[idx[a]];
      const int g = cell(E.gi, E.gj, E.gk);
      if (pos[g] < 0) {
        pos[g] = (int)list.size();
        list.push_back((int16_t)g);
      }
    }
    const int nw = (int)list.size();
    for (int a = 0; a < ne; ++a)
      for (int m = 0; m < 8; ++m) {
        const int ci = corner(ordered[idx[a]], m);
        if (pos[ci] < 0) {
          pos[ci] = (int)list.size();
          list.push_back((int16_t)ci);
        }
      }
    const int nt = (int)list.size();
    std::vector<GhostEntry> ents(ne);
    std::vector<int> level(ne);
    std::vector<int16_t> fin(nt);
    int nlev = 0;
    if (nt + ne <= std::min(slot_limit, GHOST_LDS_SLOTS)) {
      // Renamed: entry a writes slot nt + a and reads the slot holding the
      // value the sequential fill would read; only reads of earlier writes
      // order the levels.
      std::vector<int> slot_level(nt + ne, 0);
      for (int n = 0; n < nt; ++n) cur[list[n]] = n;
      for (int a = 0; a < ne; ++a) {
        GhostEntry E = ordered[idx[a]];
        int lv = 1;
        for (int m = 0; m < 8; ++m) {
          const int sl = cur[corner(E, m)];
          E.lc[m] = (int16_t)sl;
          lv = std::max(lv, slot_level[sl] + 1);
        }
        E.lg = (int16_t)(nt + a);
        slot_level[nt + a] = lv;
        cur[cell(E.gi, E.gj, E.gk)] = nt + a;
        ents[a] = E;
        level[a] = lv;
        nlev = std::max(nlev, lv);
      }
      for (int n = 0; n < nt; ++n) fin[n] = (int16_t)cur[list[n]];
      gs.max_slots = std::max(gs.max_slots, nt + ne);
    } else {
      // In place: an entry also waits for earlier reads and writes of its
      // ghost cell.
      std::fill(wlev.begin(), wlev.end(), 0);
      std::fill(rlev.begin(), rlev.end(), 0);
      for (int a = 0; a < ne; ++a) {
        GhostEntry E = ordered[idx[a]];
        int lv = 1;
        for (int m = 0; m < 8; ++m) lv = std::max(lv, wlev[corner(E, m)] + 1);
        const int g = cell(E.gi, E.gj, E.gk);
        lv = std::max(lv, rlev[g] + 1);
        lv = std::max(lv, wlev[g] + 1);
        wlev[g] = lv;
        for (int m = 0; m < 8; ++m) rlev[corner(E, m)] = std::max(rlev[corner(E, m)], lv);
        for (int m = 0; m < 8; ++m) E.lc[m] = (int16_t)pos[corner(E, m)];
        E.lg = (int16_t)pos[g];
        ents[a] = E;
        level[a] = lv;
        nlev = std::max(nlev, lv);
      }
      for (int n = 0; n < nt; ++n) fin[n] = (int16_t)n;
    }
    gs.touched_off.push_back((int)gs.touched.size());
    gs.touched_nw.push_back(nw);
    gs.touched.insert(gs.touched.end(), list.begin(), list.end());
    gs.touched_final.insert(gs.touched_final.end(), fin.begin(), fin.end());
    gs.cube_level_off.push_back((int)gs.level_off.size());
    for (int lv = 1; lv <= nlev; ++lv) {
      gs.level_off.push_back((int)gs.entries.size());
      for (int a = 0; a < ne; ++a)
        if (level[a] == lv) gs.entries.push_back(ents[a]);
    }
    gs.max_levels = std::max(gs.max_levels, nlev);
  }
  gs.cube_level_off.push_back((int)gs.level_off.size());
  gs.level_off.push_back((int)gs.entries.size());
  gs.touched_off.push_back((int)gs.touched.size());
  return gs;
}

int ghost_upload(const GhostSchedule& gs, GhostDevice& d, std::string& err) {
  ghost_free(d);
  GhostEntry* ent;
  int *clo, *lo, *toff, *tnw;
  int16_t *tc, *tf;
  if (dev_upload(&ent, gs.entries, err)) return 1;
  d.owned.push_back(ent);
  if (dev_upload(&clo, gs.cube_level_off, err)) return 1;
  d.owned.push_back(clo);
  if (dev_upload(&lo, gs.level_off, err)) return 1;
  d.owned.push_back(lo);
  if (dev_upload(&tc, gs.touched, err)) return 1;
  d.owned.push_back(tc);
  if (dev_upload(&tf, gs.touched_final, err)) return 1;
  d.owned.push_back(tf);
  if (dev_upload(&toff, gs.touched_off, err)) return 1;
  d.owned.push_back(toff);
  if (dev_upload(&tnw, gs.touched_nw, err)) return 1;
  d.owned.push_back(tnw);
  d.t.ent = ent;
  d.t.cube_level_off = clo;
  d.t.level_off = lo;
  d.t.touched = tc;
  d.t.touched_final = tf;
  d.t.touched_off = toff;
  d.t.touched_nw = tnw;
  d.t.ncubes = (int)gs.cube_level_off.size() - 1;
  d.t.max_slots = gs.max_slots;
  return 0;
}

void ghost_free(GhostDevice& d) {
  for (void* p : d.owned)
    if (p) (void)hipFree(p);
  d.owned.clear();
  d.t = GhostTables();
}

size_t engine_bytes_per_cube(int prec) {
  const size_t state = prec == 1 ? (size_t)CUBE_STRIDE * (2 * sizeof(float) + 4 * sizeof(double))
                                 : (size_t)9 * CUBE_STRIDE * sizeof(double);
  return state + PLANE + sizeof(double) + sizeof(int) + (size_t)16 * 8 * sizeof(double);
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
  (void)hipSetDevice(e->device);
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  engine_output_free(e);
  free_case(e);
  if (e->ev0) (void)hipEventDestroy(e->ev0);
  if (e->ev1) (void)hipEventDestroy(e->ev1);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

int engine_alloc(Engine* e, int cap, int max_cases, std::string& err, int prec) {
  HIP_TRY(hipSetDevice(e->device));
  if (engine_output_wait(e, err)) return 1;
  free_case(e);
  if (cap < 2) {
    err = "a pool needs at least one cube";
    return 1;
  }
  if (max_cases < 1) {
    err = "a pool needs at least one case slot";
    return 1;
  }
  if (prec != 0 && prec != 1) {
    err = "precision must be 0 (FP64) or 1 (mixed)";
    return 1;
  }
  e->prec = prec;
  e->cap = cap;
  e->ncube = 1;
  e->max_cases = max_cases;
  e->cases.assign(max_cases, CaseSlot());
  e->cube_case_host.assign(cap, -1);
  e->n_active = 0;
  const size_t ncells = (size_t)cap * PLANE;
  const size_t bytes = ncells * 5 * sizeof(double);
  if (prec == 1) {
    if (dev_alloc(&e->Uf, ncells * 5, err) || dev_alloc(&e->Uf2, ncells * 5, err)) return 1;
    HIP_TRY(hipMemsetAsync(e->Uf, 0, ncells * 5 * sizeof(float), e->stream));
    HIP_TRY(hipMemsetAsync(e->Uf2, 0, ncells * 5 * sizeof(float), e->stream));
    for (double** p : {&e->U1, &e->U1q, &e->staging, &e->Bsrc}) {
      if (dev_alloc(p, ncells * 5, err)) return 1;
      HIP_TRY(hipMemsetAsync(*p, 0, bytes, e->stream));
    }
  } else {
    for (double** p : {&e->U, &e->U1, &e->U1q, &e->U1p1, &e->U1p2, &e->Rku, &e->Res, &e->staging,
                       &e->U2}) {
      if (dev_alloc(p, ncells * 5, err)) return 1;
      HIP_TRY(hipMemsetAsync(*p, 0, bytes, e->stream));
    }
  }
  if (dev_alloc(&e->fws, ncells, err)) return 1;
  HIP_TRY(hipMemsetAsync(e->fws, 0, ncells, e->stream));
  if (dev_alloc(&e->cube_size, (size_t)cap, err)) return 1;
  HIP_TRY(hipMemsetAsync(e->cube_size, 0, cap * sizeof(double), e->stream));
  if (dev_alloc(&e->partials, (size_t)update_blocks(cap) * 8, err)) return 1;
  if (dev_alloc(&e->phys_dev, (size_t)max_cases, err)) return 1;
  if (dev_alloc(&e->cube_case, (size_t)cap, err)) return 1;
  if (ensure_er(e, 1, err)) return 1;
  if (rebuild_tables(e, err)) return 1;
  HIP_TRY(hipStreamSynchronize(e->stream));
  return 0;
}

int build_case_tables(const Case& cs, CaseTables& tb, std::string& err, int ghost_slot_limit) {
  tb = CaseTables();
  tb.n = cs.ncube - 1;
  tb.mpi_ncube = cs.ncube - 1;
  tb.ph = DevPhys{cs.phys.rho0, cs.phys.U0, cs.phys.Uin, cs.phys.V0,
                  cs.phys.W0,   cs.phys.P0, cs.phys.T0,  cs.phys.mu_L};
  // Boundary cubes with their side masks.
  {
    std::vector<uint8_t> mask(cs.ncube, 0);
    for (int s = 0; s < 6; ++s)
      for (size_t n = 1; n < cs.bc_lists[s].size(); ++n) {
        const int c = cs.bc_lists[s][n];
        if (c < 1 || c >= cs.ncube) {
          err = "boundary list names a missing cube";
          return 1;
        }
        mask[c] |= (uint8_t)(1 << s);
      }
    for (int c = 1; c < cs.ncube; ++c)
      if (mask[c]) {
        tb.bc_cubes.push_back(c);
        tb.bc_masks.push_back(mask[c]);
      }
  }
  // IB ghost schedule: minus list, then plus list, each in list order.
  {
    std::vector<GhostEntry> ordered;
    for (const IBTable* t : {&cs.ib_minus, &cs.ib_plus})
      for (int n = 1; n <= t->n; ++n) {
        GhostEntry E;
        const int* g = &t->gcindex[(n - 1) * 4 + 1];
        const int* q = &t->ipsur[(n - 1) * 4 + 1];
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
        for (int m = 0; m < 8; ++m) E.w[m] = t->weight[(n - 1) * 8 + 1 + m];
        ordered.push_back(E);
      }
    tb.ghost = build_ghost_schedule(ordered, ghost_slot_limit);
  }
  // Face halo sources from the equal-size interface lists.
  tb.halo.resize(cs.ncube);
  for (auto& h : tb.halo) memset(&h, 0, sizeof h);
  const int dirs[3] = {2, 4, 6};
  for (int d = 0; d < 3; ++d)
    for (int a = 1; a <= cs.nadj[d]; ++a) {
      int c0 = cs.adj[d][a];
      int c2 = c0 >= 1 && c0 < cs.ncube ? cs.adj_number[adj_index(c0, 1, dirs[d])] : 0;
      if (c0 < 1 || c0 >= cs.ncube || c2 <= 0 || c2 >= cs.ncube) {
        err = "interface list names a missing neighbour";
        return 1;
      }
      tb.halo[c2].src[2 * d] = c0;      // c2's minus face from c0's top layers
      tb.halo[c0].src[2 * d + 1] = c2;  // c0's plus face from c2's bottom layers
    }
  // Two-level faces.
  {
    const int* lp[L_COUNT];
    for (int l = 0; l < L_COUNT; ++l) lp[l] = cs.adj[l].data();
    for (const MlTask& t : ml_tasks(lp, cs.nadj.data(), cs.adj_number.data())) {
      if (t.dst < 1 || t.dst >= cs.ncube || t.src < 1 || t.src >= cs.ncube || t.axis < 0 ||
          t.axis > 2 || t.slot < 1 || t.slot > 4) {
        err = "two-level interface names a missing cube";
        return 1;
      }
      tb.ml.push_back(MlTaskDev{t.dst, t.src, t.axis, t.hi, t.restrict_, t.slot});
    }
  }
  // Cells of the IB lists that some neighbour gathers: the only interior
  // cells that change between a post-update exchange and the next stage's
  // pre-flux exchange.
  std::vector<char> is_ghost((size_t)cs.ncube * PLANE, 0);
  for (const IBTable* t : {&cs.ib_minus, &cs.ib_plus})
    for (int n = 1; n <= t->n; ++n) {
      const int* g = &t->gcindex[(n - 1) * 4 + 1];
      is_ghost[(size_t)g[0] * PLANE + g[1] * 400 + g[2] * 20 + g[3]] = 1;
    }
  for (int n = 1; n < cs.ncube; ++n)
    for (int f = 0; f < 6; ++f) {
      const int sc = tb.halo[n].src[f];
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
            tb.halo_pairs.push_back(
                make_int2(sc * PLANE + scell, n * PLANE + di * 400 + dj * 20 + dk));
          }
    }
  return 0;
}

// Concatenates the tables of all active cases in pool numbering and uploads
// them with the cube -> case map and the per-case physics and norm records.
int rebuild_tables(Engine* e, std::string& err) {
  free_tables(e);
  std::vector<int> order;
  for (int s = 0; s < (int)e->cases.size(); ++s)
    if (e->cases[s].active) order.push_back(s);
  std::sort(order.begin(), order.end(),
            [&](int a, int b) { return e->cases[a].first < e->cases[b].first; });
  std::vector<int> bc;
  std::vector<uint8_t> bm;
  GhostSchedule gs;
  std::vector<int2> pairs;
  std::vector<MlTaskDev> ml;
  std::vector<CaseDesc> desc;
  std::vector<DevPhys> phys(e->max_cases);
  for (int s = 0; s < e->max_cases; ++s) phys[s] = e->cases[s].tb.ph;
  int top = 1;
  for (int s : order) {
    const CaseSlot& cs = e->cases[s];
    const CaseTables& tb = cs.tb;
    const int off = cs.first - 1;
    top = std::max(top, cs.first + tb.n);
    for (size_t a = 0; a < tb.bc_cubes.size(); ++a) {
      bc.push_back(tb.bc_cubes[a] + off);
      bm.push_back(tb.bc_masks[a]);
    }
    const int e0 = (int)gs.entries.size(), l0 = (int)gs.level_off.size();
    const int t0 = (int)gs.touched.size();
    for (GhostEntry E : tb.ghost.entries) {
      E.cube += off;
      gs.entries.push_back(E);
    }
    for (size_t a = 0; a + 1 < tb.ghost.level_off.size(); ++a)
      gs.level_off.push_back(tb.ghost.level_off[a] + e0);
    for (size_t a = 0; a + 1 < tb.ghost.cube_level_off.size(); ++a) {
      gs.cube_level_off.push_back(tb.ghost.cube_level_off[a] + l0);
      gs.touched_off.push_back(tb.ghost.touched_off[a] + t0);
      gs.touched_nw.push_back(tb.ghost.touched_nw[a]);
    }
    gs.touched.insert(gs.touched.end(), tb.ghost.touched.begin(), tb.ghost.touched.end());
    gs.touched_final.insert(gs.touched_final.end(), tb.ghost.touched_final.begin(),
                            tb.ghost.touched_final.end());
    gs.max_levels = std::max(gs.max_levels, tb.ghost.max_levels);
    gs.max_slots = std::max(gs.max_slots, tb.ghost.max_slots);
    for (int2 pr : tb.halo_pairs) pairs.push_back(make_int2(pr.x + off * PLANE, pr.y + off * PLANE));
    for (MlTaskDev t : tb.ml) {
      t.dst += off;
      t.src += off;
      ml.push_back(t);
    }
    desc.push_back(CaseDesc{cs.first, tb.n, tb.mpi_ncube, s});
  }
  gs.cube_level_off.push_back((int)gs.level_off.size());
  gs.level_off.push_back((int)gs.entries.size());
  gs.touched_off.push_back((int)gs.touched.size());
  std::vector<uint8_t> cube_mask(e->cap, 0);
  for (size_t a = 0; a < bc.size(); ++a) cube_mask[bc[a]] = bm[a];
  {
    // A z-BC reads the plane k = 2 (or 17) of its own cube.  Where nothing
    // writes the cube between the stage's BCs and its flux (no IB fill, no
    // IB-refresh copy, no two-level face), the pushing flux kernels apply the
    // z-BC to the interior planes themselves.
    std::vector<char> busy(top, 0);
    for (const GhostEntry& E : gs.entries) busy[E.cube] = 1;
    for (const int2& pr : pairs) busy[pr.y / PLANE] = 1;
    for (const MlTaskDev& t : ml) busy[t.dst] = 1;
    for (int c : bc) {
      if (busy[c]) continue;
      if (cube_mask[c] & 16) cube_mask[c] |= ZBC_IN_FLUX;
      if (cube_mask[c] & 32) cube_mask[c] |= ZBC_IN_FLUX << 1;
    }
  }
  std::vector<HaloSrc> hs(top);
  for (auto& h : hs) memset(&h, 0, sizeof h);
  for (int s : order) {
    const CaseSlot& cs = e->cases[s];
    const int off = cs.first - 1;
    for (int c = 1; c <= cs.tb.n && c < (int)cs.tb.halo.size(); ++c)
      for (int f = 0; f < 6; ++f) {
        const int src = cs.tb.halo[c].src[f];
        hs[c + off].src[f] = src > 0 ? src + off : 0;
      }
  }
  e->ncube = top;
  e->n_active = (int)order.size();
  e->n_bc = (int)bc.size();
  e->n_ghost_cubes = (int)gs.cube_level_off.size() - 1;
  e->n_ghost_entries = (int)gs.entries.size();
  e->ghost_levels = gs.max_levels;
  e->n_halo_pairs = (int)pairs.size();
  e->n_ml = (int)ml.size();
  if (dev_upload(&e->ml, ml, err)) return 1;
  {
    std::vector<char> mark(top, 0);
    for (int c : bc) mark[c] = 1;
    for (const int2& pr : pairs) mark[pr.y / PLANE] = 1;
    for (const MlTaskDev& t : ml) mark[t.dst] = 1;
    std::vector<int> list;
    for (int c = 1; c < top; ++c)
      if (mark[c]) list.push_back(c);
    e->n_edge_list = (int)list.size();
    if (dev_upload(&e->edge_list, list, err)) return 1;
  }
  {
    // 2: the IB fill touches an edge cell (all edge passes), 1: boundary cube
    // only (the edge cells its BCs read).
    std::vector<char> mark(top, 0);
    for (int c : bc) mark[c] = 1;
    for (const GhostEntry& E : gs.entries) {
      bool touch = edge_cell(E.gi, E.gj, E.gk);
      for (int d = 0; d < 8; ++d)
        touch = touch || edge_cell(E.pi + (d & 1), E.pj + ((d >> 1) & 1), E.pk + ((d >> 2) & 1));
      if (touch) mark[E.cube] = 2;
    }
    // The refresh copies land on face ghost cells, which the flux kernels
    // read from memory; an edge destination would be skipped by them.
    for (const int2& pr : pairs) {
      const int cell = pr.y % PLANE;
      if (edge_cell(cell / 400, cell / 20 % 20, cell % 20)) {
        err = "an IB refresh copy lands on an edge ghost cell";
        return 1;
      }
    }
    // One item per cube: IB cubes first (their fills are the longest chains),
    // then the other listed cubes.
    std::vector<int> gidx(top, -1);
    const int ng = (int)gs.cube_level_off.size() - 1;
    for (int g = 0; g < ng; ++g) {
      if (gs.cube_level_off[g] == gs.cube_level_off[g + 1]) continue;
      const int c = gs.entries[gs.level_off[gs.cube_level_off[g]]].cube;
      if (gidx[c] >= 0) {
        err = "a cube appears twice in the IB fill schedule";
        return 1;
      }
      gidx[c] = g;
    }
    std::vector<int2> items;
    for (int c = 1; c < top; ++c)
      if (gidx[c] >= 0) items.push_back(make_int2(mark[c] ? (mark[c] == 2 ? c : -c) : 0, gidx[c]));
    for (int c = 1; c < top; ++c)
      if (gidx[c] < 0 && mark[c]) items.push_back(make_int2(mark[c] == 2 ? c : -c, -1));
    e->n_open_items = (int)items.size();
    if (dev_upload(&e->open_items, items, err)) return 1;
  }
  if (dev_upload(&e->bc_cubes, bc, err) || dev_upload(&e->bc_masks, bm, err) ||
      ghost_upload(gs, e->ghost, err) || dev_upload(&e->halo, hs, err) ||
      dev_upload(&e->halo_pairs, pairs, err) || dev_upload(&e->case_desc, desc, err) ||
      dev_upload(&e->bc_mask_cube, cube_mask, err))
    return 1;
  HIP_TRY(hipMemcpy(e->phys_dev, phys.data(), phys.size() * sizeof(DevPhys),
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(e->cube_case, e->cube_case_host.data(), e->cap * sizeof(int),
                    hipMemcpyHostToDevice));
  free_graph(e);
  return 0;
}

int engine_free_range(const Engine* e) {
  int best = 0, run = 0;
  for (int c = 1; c < e->cap; ++c) {
    run = e->cube_case_host[c] < 0 ? run + 1 : 0;
    best = std::max(best, run);
  }
  return best;
}

namespace {
// First pool cube of the smallest free run of at least n cubes (the lowest
// such run on ties), or 0.  Best fit keeps large runs whole for large cases.
int find_range(const Engine* e, int n) {
  int best = 0, best_len = 0, run = 0;
  for (int c = 1; c <= e->cap; ++c) {
    if (c < e->cap && e->cube_case_host[c] < 0) {
      ++run;
      continue;
    }
    if (run >= n && (best_len == 0 || run < best_len)) {
      best = c - run;
      best_len = run;
    }
    run = 0;
  }
  return best;
}
}  // namespace

namespace {
int place_case(Engine* e, const Case& cs, int slot, std::string& err);
}  // namespace

int engine_add_case(Engine* e, const Case& cs, int* slot_out, std::string& err) {
  HIP_TRY(hipSetDevice(e->device));
  if (slot_out) *slot_out = -1;
  int slot = -1;
  for (int s = 0; s < e->max_cases; ++s)
    if (!e->cases[s].active) {
      slot = s;
      break;
    }
  if (slot < 0) {
    err = "all case slots are in use";
    return 1;
  }
  CaseTables tb;
  if (build_case_tables(cs, tb, err, e->ghost_slot_limit)) return 1;
  const int first = find_range(e, tb.n);
  if (first == 0) {
    err = "no free range of " + std::to_string(tb.n) + " cubes in the pool";
    return 1;
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  CaseSlot& S = e->cases[slot];
  S.first = first;
  S.tb = std::move(tb);
  S.active = true;
  const int n = S.tb.n;
  if (place_case(e, cs, slot, err)) {
    // Roll back: the slot and its cubes are free again and the tables of the
    // other cases are rebuilt without it.
    for (int c = first; c < first + n; ++c) e->cube_case_host[c] = -1;
    S = CaseSlot();
    std::string err2;
    (void)rebuild_tables(e, err2);
    e->halo_full_next = true;
    return 1;
  }
  if (slot_out) *slot_out = slot;
  return 0;
}

namespace {
// Uploads the flags, sizes and state of a case just given slot `slot` and
// its range, and rebuilds the pool tables.
int place_case(Engine* e, const Case& cs, int slot, std::string& err) {
  CaseSlot& S = e->cases[slot];
  const int first = S.first, n = S.tb.n;
  S.centres.assign((size_t)3 * (n + 1) * NP, 0.0);
  for (int a = 0; a < 3; ++a) {
    const std::vector<double>& src = a == 0 ? cs.Xcnt : a == 1 ? cs.Ycnt : cs.Zcnt;
    if (src.size() >= (size_t)(n + 1) * NP)
      std::copy(src.begin(), src.begin() + (size_t)(n + 1) * NP,
                S.centres.begin() + (size_t)a * (n + 1) * NP);
  }
  // Flags, cube sizes and state of the range; RK scratch, increments and
  // force densities start at zero as in the reference's fresh arrays.
  {
    std::vector<int8_t> f((size_t)n * PLANE);
    for (size_t a = 0; a < f.size(); ++a) f[a] = (int8_t)cs.fws[PLANE + a];
    HIP_TRY(hipMemcpy(e->fws + (size_t)first * PLANE, f.data(), f.size(), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(e->cube_size + first, cs.cube_size.data() + 1, n * sizeof(double),
                    hipMemcpyHostToDevice));
  for (double* p : {e->U1p1, e->U1p2, e->Rku, e->Res})
    if (p)
      HIP_TRY(hipMemsetAsync(p + (size_t)first * CUBE_STRIDE, 0,
                             (size_t)n * CUBE_STRIDE * sizeof(double), e->stream));
  for (int c = first; c < first + n; ++c) e->cube_case_host[c] = slot;
  if (rebuild_tables(e, err)) return 1;
  const double* st[3] = {cs.U1_.data(), cs.U1.data(), cs.U1q.data()};
  for (int w = 0; w < 3; ++w)
    if (engine_case_state(e, slot, w, const_cast<double*>(st[w]), true, err)) return 1;
  e->halo_full_next = true;
  return 0;
}
}  // namespace

int engine_remove_case(Engine* e, int slot, std::string& err) {
  if (slot < 0 || slot >= e->max_cases || !e->cases[slot].active) {
    err = "no such case slot";
    return 1;
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  CaseSlot& S = e->cases[slot];
  for (int c = S.first; c < S.first + S.tb.n; ++c) e->cube_case_host[c] = -1;
  S = CaseSlot();
  return rebuild_tables(e, err);
}

int engine_bind_single(Engine* e, int mpi_ncube, const DevPhys& ph, std::string& err) {
  for (int s = 0; s < e->max_cases; ++s) e->cases[s] = CaseSlot();
  CaseSlot& S = e->cases[0];
  S.active = true;
  S.first = 1;
  S.tb.n = e->cap - 1;
  S.tb.mpi_ncube = mpi_ncube;
  S.tb.ph = ph;
  for (int c = 0; c < e->cap; ++c) e->cube_case_host[c] = c == 0 ? -1 : 0;
  return rebuild_tables(e, err);
}

int engine_load_case(Engine* e, const Case& cs, std::string& err) {
  if (engine_alloc(e, cs.ncube, 1, err)) return 1;
  e->flux_tuned = false;
  e->flux_pinned = false;
  int slot;
  if (engine_add_case(e, cs, &slot, err)) return 1;
  e->halo_full_next = false;
  HIP_TRY(hipStreamSynchronize(e->stream));
  return 0;
}

int engine_case_state(Engine* e, int slot, int which, double* aos, bool upload,
                      std::string& err) {
  if (slot < 0 || slot >= e->max_cases || !e->cases[slot].active) {
    err = "no such case slot";
    return 1;
  }
  const bool f32 = e->prec == 1 && which == 0;
  double* d = f32 ? nullptr : state_ptr(e, which);
  if (!d && !f32) {
    err = e->prec == 1 ? "the mixed path keeps only U1_, U1 and U1q" : "unknown state array";
    return 1;
  }
  const CaseSlot& S = e->cases[slot];
  const size_t cube_bytes = (size_t)CUBE_STRIDE * sizeof(double);
  const size_t n = S.tb.n;
  double* dst = f32 ? nullptr : d + (size_t)S.first * CUBE_STRIDE;
  float* dstf = f32 ? e->Uf + (size_t)S.first * CUBE_STRIDE : nullptr;
  if (upload) {
    if (which == 0) e->halo_full_next = true;
    HIP_TRY(hipMemcpyAsync(e->staging, aos + CUBE_STRIDE, n * cube_bytes, hipMemcpyHostToDevice,
                           e->stream));
    double ref[5];
    ref_state(S.tb.ph, ref);
    if (f32)
      launch_aos_to_soa_f(e->stream, e->staging, dstf, (int)n, ref);
    else
      launch_aos_to_soa(e->stream, e->staging, dst, (int)n);
    // The second copy of the iterate starts equal (cells no kernel rewrites
    // keep their values in both).
    if (which == 0 && f32)
      launch_copy_f(e->stream, dstf, e->Uf2 + (size_t)S.first * CUBE_STRIDE, n * CUBE_STRIDE);
    if (which == 0 && !f32 && e->U2)
      launch_copy_d(e->stream, dst, e->U2 + (size_t)S.first * CUBE_STRIDE, n * CUBE_STRIDE);
    // The mixed stages read the dual-time source constant formed from U1, U1q.
    if (e->prec == 1 && which != 0)
      launch_bsrc_mixed(e->stream, e->U1, e->U1q, e->Bsrc, S.first, (int)n, e->pt());
  } else {
    double ref[5];
    ref_state(S.tb.ph, ref);
    if (f32)
      launch_soa_to_aos_f(e->stream, dstf, e->staging, (int)n, ref);
    else
      launch_soa_to_aos(e->stream, dst, e->staging, (int)n);
    HIP_TRY(hipMemcpyAsync(aos + CUBE_STRIDE, e->staging, n * cube_bytes, hipMemcpyDeviceToHost,
                           e->stream));
    memset(aos, 0, cube_bytes);
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  return 0;
}

namespace {
size_t pool_bytes(const Engine* e) { return (size_t)e->ncube * CUBE_STRIDE * sizeof(double); }
}  // namespace

int engine_upload_state(Engine* e, int which, const double* aos) {
  if (e->prec != 0) return 1;
  double* d = state_ptr(e, which);
  if (!d) return 1;
  if (which == 0) e->halo_full_next = true;
  if (hipMemcpyAsync(e->staging, aos, pool_bytes(e), hipMemcpyHostToDevice, e->stream) != hipSuccess)
    return 1;
  launch_aos_to_soa(e->stream, e->staging, d, e->ncube);
  if (which == 0 && e->U2) launch_copy_d(e->stream, d, e->U2, (size_t)e->ncube * CUBE_STRIDE);
  return hipStreamSynchronize(e->stream) == hipSuccess ? 0 : 1;
}

int engine_download_state(Engine* e, int which, double* aos) {
  if (e->prec != 0) return 1;
  double* d = state_ptr(e, which);
  if (!d) return 1;
  launch_soa_to_aos(e->stream, d, e->staging, e->ncube);
  if (hipMemcpyAsync(aos, e->staging, pool_bytes(e), hipMemcpyDeviceToHost, e->stream) != hipSuccess)
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

void engine_bc(Engine* e) {
  if (e->prec == 1)
    launch_bc(e->stream, e->Uf, e->bc_cubes, e->bc_masks, e->n_bc, e->pt());
  else
    launch_bc(e->stream, e->U, e->bc_cubes, e->bc_masks, e->n_bc, e->pt());
}

void engine_ghost(Engine* e) {
  if (e->prec == 1)
    launch_ghost(e->stream, e->Uf, e->ghost.t, e->pt());
  else
    launch_ghost(e->stream, e->U, e->ghost.t, e->pt());
}

namespace {
// Two-level faces, then the edge passes of every cube (the equal-size faces
// are already in place).
template <typename T>
void ml_then_edges(Engine* e, T* U) {
  launch_halo_ml(e->stream, U, e->ml, e->n_ml, e->pt());
  launch_edge(e->stream, U, e->ncube, e->cube_case);
}
}  // namespace

void engine_interface(Engine* e) {
  if (e->prec == 1) {
    launch_halo_edge(e->stream, e->Uf, e->halo, e->ncube, e->cube_case);
    if (e->n_ml) ml_then_edges(e, e->Uf);
  } else {
    launch_halo_edge(e->stream, e->U, e->halo, e->ncube, e->cube_case);
    if (e->n_ml) ml_then_edges(e, e->U);
  }
}

namespace {
// IB ghost cells in equal-size gather layers, then (two-level grids) the
// two-level faces, whose restrictions read small cubes' cells near the face,
// then the edge passes of the cubes whose face ghosts changed since the
// stage's opening edge pass (elsewhere they would give the same values).
// Before a halo-pushing flux kernel (which forms the edge cells it reads)
// the edge passes are left out.
template <typename T>
void halo_refresh(Engine* e, T* U, bool edges) {
  launch_halo_pairs(e->stream, U, e->halo_pairs, e->n_halo_pairs);
  launch_halo_ml(e->stream, U, e->ml, e->n_ml, e->pt());
  if (edges) launch_edge_list(e->stream, U, e->edge_list, e->n_edge_list);
}
}  // namespace

void engine_halo_refresh(Engine* e) {
  if (e->prec == 1)
    halo_refresh(e, e->Uf, true);
  else
    halo_refresh(e, e->U, true);
}

namespace {
// Opening of a stage after a halo-pushing flux kernel: two-level faces, then
// the edge passes and BCs of the cubes that need their edge cells in memory,
// then the IB fill and the pre-flux exchange (full, or the refresh without
// edge passes).
template <typename T>
void push_stage_open(Engine* e, T* U, bool full_pre_halo) {
  launch_halo_ml(e->stream, U, e->ml, e->n_ml, e->pt());
  launch_open(e->stream, U, e->open_items, e->n_open_items, e->bc_mask_cube, e->ghost.t, e->pt());
  if (full_pre_halo)
    engine_interface(e);
  else
    halo_refresh(e, U, false);
}
}  // namespace

namespace {
double* er_row(Engine* e, int er_slot) { return e->er_dev + (size_t)er_slot * e->max_cases * 8; }
}  // namespace

void engine_flux(Engine* e, int RK) {
  if (e->flux_variant == 0)
    launch_flux(e->stream, e->U, e->U1, e->U1q, e->fws, e->cube_size, e->Rku, e->Res, e->ncube,
                e->pt(), e->deltaT, e->deltaTau, e->e);
  else
    launch_flux_march(e->stream, e->U, e->U1, e->U1q, e->fws, e->cube_size, e->Rku, e->Res,
                      e->ncube, e->pt(), e->deltaT, e->deltaTau, e->e, e->flux_nseg, RK, 0,
                      e->U1p1, e->U1p2, e->partials);
}

bool engine_fused(const Engine* e) {
  return e->prec == 1 || (e->flux_variant == 1 && e->flux_nseg == 1 && e->flux_fuse);
}

void engine_flux_update(Engine* e, int RK, int er_slot) {
  if (e->prec == 1) {
    launch_flux_mixed(e->stream, e->Uf, e->Uf2, e->halo, e->bc_mask_cube, e->Bsrc, e->fws,
                      e->cube_size, e->ncube, e->pt(), e->deltaT, e->deltaTau, e->e, RK,
                      e->partials, e->mixed_nseg);
    std::swap(e->Uf, e->Uf2);
    if (RK == 3)
      launch_reduce(e->stream, e->partials, MIXED_ROWS_PER_CUBE, e->case_desc, e->n_active,
                    e->phys_dev, er_row(e, er_slot));
    return;
  }
  if (!engine_fused(e)) {
    engine_flux(e, RK);
    engine_update(e, RK, er_slot);
    return;
  }
  launch_flux_march(e->stream, e->U, e->U1, e->U1q, e->fws, e->cube_size, e->Rku, e->Res,
                    e->ncube, e->pt(), e->deltaT, e->deltaTau, e->e, 1, RK, 1, e->U1p1, e->U1p2,
                    e->partials);
  if (RK == 3)
    launch_reduce(e->stream, e->partials, 1, e->case_desc, e->n_active, e->phys_dev,
                  er_row(e, er_slot));
}

namespace {
// Size class of the resident pool: floor(log2(resident cubes)).
int size_class(const Engine* e) {
  long n = 0;
  for (const CaseSlot& S : e->cases)
    if (S.active) n += S.tb.n;
  int c = 0;
  while (n > 1) {
    n >>= 1;
    ++c;
  }
  return c;
}
}  // namespace

// Tunes the flux kernel choice when it has not been chosen yet, or (mixed
// path, choice not pinned) when the pool's size class changed since.
int engine_maybe_tune(Engine* e, std::string& err) {
  if (e->flux_tuned && (e->prec == 0 || e->flux_pinned || size_class(e) == e->tuned_class))
    return 0;
  int v, n;
  return engine_autotune(e, &v, &n, err);
}

int engine_autotune(Engine* e, int* variant, int* nseg, std::string& err) {
  HIP_TRY(hipSetDevice(e->device));
  if (e->n_active == 0) {
    err = "no case loaded";
    return 1;
  }
  if (e->prec == 1) {
    // The mixed flux kernel in 1, 2 or 4 runs per cube (bitwise the same
    // results), each timed as one whole RK stage on the resident cases, once
    // per size class of the pool; the iterate the stages write is restored.
    // The candidates are timed in three interleaved rounds and each keeps its
    // fastest round, so a clock ramp at the start favours none of them.
    const int cls = size_class(e);
    auto known = e->nseg_by_class.find(cls);
    if (known == e->nseg_by_class.end()) {
      const size_t bytes = (size_t)e->ncube * CUBE_STRIDE * sizeof(float);
      struct Saved {
        float* p[2] = {nullptr, nullptr};
        ~Saved() {
          for (float* q : p)
            if (q) (void)hipFree(q);
        }
      } saved;
      float* const live[2] = {e->Uf, e->Uf2};
      for (int a = 0; a < 2; ++a) {
        HIP_TRY(hipMalloc((void**)&saved.p[a], bytes));
        HIP_TRY(hipMemcpyAsync(saved.p[a], live[a], bytes, hipMemcpyDeviceToDevice, e->stream));
      }
      const bool full = e->halo_full_next;
      const int cands[3] = {1, 2, 4};
      float fastest[3] = {1e30f, 1e30f, 1e30f};
      for (int round = 0; round < 3; ++round)
        for (int c = 0; c < 3; ++c) {
          e->mixed_nseg = cands[c];
          engine_stage(e, 2, 0, full);   // warm up
          HIP_TRY(hipEventRecord(e->ev0, e->stream));
          for (int r = 0; r < 3; ++r) engine_stage(e, 2, 0, full);
          HIP_TRY(hipEventRecord(e->ev1, e->stream));
          HIP_TRY(hipEventSynchronize(e->ev1));
          float ms = 0;
          HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
          e->Uf = live[0];
          e->Uf2 = live[1];
          for (int a = 0; a < 2; ++a)
            HIP_TRY(
                hipMemcpyAsync(live[a], saved.p[a], bytes, hipMemcpyDeviceToDevice, e->stream));
          fastest[c] = std::min(fastest[c], ms);
        }
      HIP_TRY(hipStreamSynchronize(e->stream));
      int bn = cands[0];
      float best = fastest[0];
      for (int c = 1; c < 3; ++c)
        if (fastest[c] < best) {
          best = fastest[c];
          bn = cands[c];
        }
      known = e->nseg_by_class.emplace(cls, bn).first;
    }
    if (e->mixed_nseg != known->second) e->graph_valid = false;
    e->mixed_nseg = known->second;
    e->tuned_class = cls;
    e->flux_tuned = true;
    if (variant) *variant = 2;
    if (nseg) *nseg = e->mixed_nseg;
    return 0;
  }
  // Candidates (variant, nseg, fuse): the marching kernel with the RK update
  // fused in place, and the marching kernel writing the update and face halos
  // into the other copy of the iterate (fuse 2).  Both reduce the norms over
  // one partial row per cube, so the norms do not depend on which one wins.
  // (The plane kernel and the segmented marching kernel, selectable with
  // set_flux, reduce over per-plane rows and are not candidates.)  Each is
  // timed as one whole RK stage on the resident cases, as the inner
  // iterations of a step run it, in three interleaved rounds of which each
  // keeps its fastest; every array a stage writes is saved first and
  // restored afterwards.
  const int cand[2][3] = {{1, 1, 1}, {1, 1, 2}};
  double* const live[6] = {e->U, e->U2, e->U1p1, e->U1p2, e->Rku, e->Res};
  struct Saved {
    double* p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    ~Saved() {
      for (double* q : p)
        if (q) (void)hipFree(q);
    }
  } saved;
  for (int a = 0; a < 6; ++a) {
    HIP_TRY(hipMalloc((void**)&saved.p[a], pool_bytes(e)));
    HIP_TRY(hipMemcpyAsync(saved.p[a], live[a], pool_bytes(e), hipMemcpyDeviceToDevice,
                           e->stream));
  }
  const bool full = e->halo_full_next;
  float fastest[2] = {1e30f, 1e30f};
  for (int round = 0; round < 3; ++round)
    for (int c = 0; c < 2; ++c) {
      const int* cv = cand[c];
      e->flux_variant = cv[0];
      e->flux_nseg = cv[1];
      e->flux_fuse = cv[2] != 0;
      e->flux_push = cv[2] == 2;
      e->push_aux = false;
      engine_stage(e, 2, 0, full);  // warm up
      HIP_TRY(hipEventRecord(e->ev0, e->stream));
      for (int r = 0; r < 3; ++r) engine_stage(e, 2, 0, full);
      HIP_TRY(hipEventRecord(e->ev1, e->stream));
      e->push_aux = true;
      HIP_TRY(hipEventSynchronize(e->ev1));
      float ms = 0;
      HIP_TRY(hipEventElapsedTime(&ms, e->ev0, e->ev1));
      e->U = live[0];
      e->U2 = live[1];
      for (int a = 0; a < 6; ++a)
        HIP_TRY(hipMemcpyAsync(live[a], saved.p[a], pool_bytes(e), hipMemcpyDeviceToDevice,
                               e->stream));
      fastest[c] = std::min(fastest[c], ms);
    }
  HIP_TRY(hipStreamSynchronize(e->stream));
  const int* win = cand[fastest[1] < fastest[0] ? 1 : 0];
  const int bv = win[0], bs = win[1];
  const bool bf = win[2] != 0, bp = win[2] == 2;
  e->flux_push = bp;
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
                e->cube_case, e->partials);
  if (RK == 3)
    launch_reduce(e->stream, e->partials, 16, e->case_desc, e->n_active, e->phys_dev,
                  er_row(e, er_slot));
}

void engine_shift(Engine* e) {
  if (e->prec == 1)
    launch_shift_mixed(e->stream, e->Uf, e->U1, e->U1q, e->Bsrc, e->ncube, e->pt());
  else
    launch_shift(e->stream, e->U, e->U1, e->U1q, e->ncube, e->cube_case);
}

void engine_flux_push(Engine* e, int RK, int er_slot) {
  if (e->prec == 1) {
    engine_flux_update(e, RK, er_slot);
    return;
  }
  launch_flux_push(e->stream, e->U, e->U2, e->halo, e->bc_mask_cube, e->U1, e->U1q, e->fws,
                   e->cube_size, e->Rku, e->Res, e->ncube, e->pt(), e->deltaT, e->deltaTau, e->e,
                   RK, e->U1p1, e->U1p2, e->partials, e->push_aux ? 1 : 0);
  std::swap(e->U, e->U2);
  if (RK == 3)
    launch_reduce(e->stream, e->partials, 1, e->case_desc, e->n_active, e->phys_dev,
                  er_row(e, er_slot));
}

void engine_stage(Engine* e, int RK, int er_slot, bool full_pre_halo) {
  if (e->prec == 1) {
    // The previous flux kernel wrote the equal-size face halos; the
    // two-level faces and, where needed, the edge passes of that exchange and
    // this stage's BCs open the stage (all idempotent when no interior cell
    // changed, as at the start of a run).
    push_stage_open(e, e->Uf, full_pre_halo);
    engine_flux_update(e, RK, er_slot);
    return;
  }
  if (e->flux_push) {
    push_stage_open(e, e->U, full_pre_halo);
    engine_flux_push(e, RK, er_slot);
    return;
  }
  engine_bc(e);
  engine_ghost(e);
  if (full_pre_halo)
    engine_interface(e);
  else
    engine_halo_refresh(e);
  engine_flux_update(e, RK, er_slot);
  engine_interface(e);
}

int ensure_er(Engine* e, int slots, std::string& err) {
  if (e->er_slots >= slots) return 0;
  if (e->er_dev) (void)hipFree(e->er_dev);
  if (e->er_host) (void)hipHostFree(e->er_host);
  e->er_dev = nullptr;
  e->er_host = nullptr;
  free_graph(e);
  const size_t n = (size_t)slots * e->max_cases * 8;
  HIP_TRY(hipMalloc((void**)&e->er_dev, n * sizeof(double)));
  HIP_TRY(hipMemset(e->er_dev, 0, n * sizeof(double)));
  HIP_TRY(hipHostMalloc((void**)&e->er_host, n * sizeof(double)));
  e->er_slots = slots;
  return 0;
}

namespace {

void enqueue_step(Engine* e, int n_inner, bool full_first) {
  float* startf = e->Uf;
  double* startd = e->U;
  // Only the last inner iteration's stages store Rku1, U1p1, U1p2 and
  // Residual1, which no stage reads back: the step ends with the arrays the
  // reference leaves.
  for (int it = 0; it < n_inner; ++it) {
    e->push_aux = it == n_inner - 1;
    for (int RK = 1; RK <= 3; ++RK) engine_stage(e, RK, it, full_first && it == 0 && RK == 1);
  }
  e->push_aux = true;
  engine_shift(e);
  // An odd number of flux kernels that swap copies: bring the iterate back to
  // the copy the step started from, so that every replay of a captured step
  // is alike.
  if (e->prec == 1 && e->Uf != startf) {
    launch_copy_f(e->stream, e->Uf, startf, (size_t)e->ncube * CUBE_STRIDE);
    std::swap(e->Uf, e->Uf2);
  }
  if (e->prec == 0 && e->U != startd) {
    launch_copy_d(e->stream, e->U, startd, (size_t)e->ncube * CUBE_STRIDE);
    std::swap(e->U, e->U2);
  }
}

}  // namespace

int engine_run(Engine* e, int n_steps, int n_inner, double* er_hist, bool use_graph,
               std::string& err) {
  HIP_TRY(hipSetDevice(e->device));
  if (e->n_active == 0) {
    err = "no case loaded";
    return 1;
  }
  if (ensure_er(e, n_inner, err)) return 1;
  if (engine_maybe_tune(e, err)) return 1;
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
      const size_t row = (size_t)e->max_cases * 8;
      HIP_TRY(hipMemcpyAsync(e->er_host, e->er_dev, n_inner * row * sizeof(double),
                             hipMemcpyDeviceToHost, e->stream));
      HIP_TRY(hipStreamSynchronize(e->stream));
      for (int it = 0; it < n_inner; ++it)
        for (int c = 0; c < e->max_cases; ++c)
          for (int m = 0; m < 7; ++m)
            er_hist[(((size_t)s * n_inner + it) * e->max_cases + c) * 7 + m] =
                e->er_host[it * row + c * 8 + 1 + m];
    }
  }
  // The last flux kernel wrote the equal-size face halos; the two-level
  // faces and the edge passes complete that exchange (as the reference's
  // closing BCM_Interface does).
  if (e->prec == 1 && n_steps > 0) ml_then_edges(e, e->Uf);
  if (e->prec == 0 && e->flux_push && n_steps > 0) ml_then_edges(e, e->U);
  HIP_TRY(hipGetLastError());
  return 0;
}

namespace {

// Completes the exchange the last flux kernel began (two-level faces and edge
// passes), so that the iterate holds every ghost value a closing
// BCM_Interface would give.
void complete_exchange(Engine* e) {
  if (e->prec == 1)
    ml_then_edges(e, e->Uf);
  else if (e->flux_push)
    ml_then_edges(e, e->U);
}

// Physical-time shift of one case's cubes (pool cubes first..first+n-1): the
// shift kernels see the case's range as cubes 1..n of a smaller pool.
void shift_case(Engine* e, const CaseSlot& S) {
  const size_t off = (size_t)(S.first - 1) * CUBE_STRIDE;
  int* cc = e->cube_case + (S.first - 1);
  if (e->prec == 1)
    launch_shift_mixed(e->stream, e->Uf + off, e->U1 + off, e->U1q + off, e->Bsrc + off, S.tb.n + 1,
                       PhysTab{e->phys_dev, cc});
  else
    launch_shift(e->stream, e->U + off, e->U1 + off, e->U1q + off, S.tb.n + 1, cc);
}

// Copies one case's cubes of the current iterate to (save) or from the
// staging buffer, which no run uses otherwise.
void stash_case(Engine* e, const CaseSlot& S, bool save) {
  const size_t off = (size_t)S.first * CUBE_STRIDE, n = (size_t)S.tb.n * CUBE_STRIDE;
  if (e->prec == 1) {
    float* st = reinterpret_cast<float*>(e->staging) + off;
    if (save)
      launch_copy_f(e->stream, e->Uf + off, st, n);
    else
      launch_copy_f(e->stream, st, e->Uf + off, n);
  } else {
    double* st = e->staging + off;
    if (save)
      launch_copy_d(e->stream, e->U + off, st, n);
    else
      launch_copy_d(e->stream, st, e->U + off, n);
  }
}

}  // namespace

int engine_run_converge(Engine* e, int n_steps, int max_inner, double tol, double* E, int* iters,
                        double* er_hist, std::string& err) {
  HIP_TRY(hipSetDevice(e->device));
  if (e->n_active == 0) {
    err = "no case loaded";
    return 1;
  }
  if (n_steps < 0 || max_inner < 1) {
    err = "invalid step counts";
    return 1;
  }
  if (ensure_er(e, 1, err)) return 1;
  if (engine_maybe_tune(e, err)) return 1;
  const int mc = e->max_cases;
  const size_t row = (size_t)mc * 8;
  std::vector<int> done(mc);
  std::vector<int> stashed;
  for (int s = 0; s < n_steps; ++s) {
    float* startf = e->Uf;
    double* startd = e->U;
    int running = 0;
    for (int c = 0; c < mc; ++c) {
      done[c] = e->cases[c].active ? 0 : -1;
      running += e->cases[c].active;
    }
    stashed.clear();
    const bool full = e->halo_full_next;
    e->halo_full_next = false;
    for (int it = 1; running > 0; ++it) {
      for (int RK = 1; RK <= 3; ++RK) engine_stage(e, RK, 0, full && it == 1 && RK == 1);
      HIP_TRY(hipMemcpyAsync(e->er_host, e->er_dev, row * sizeof(double), hipMemcpyDeviceToHost,
                             e->stream));
      HIP_TRY(hipStreamSynchronize(e->stream));
      std::vector<int> ended;
      for (int c = 0; c < mc; ++c) {
        if (done[c] != 0) continue;
        double* er = e->er_host + c * 8;
        if (er_hist)
          for (int m = 0; m < 7; ++m)
            er_hist[(((size_t)s * max_inner + it - 1) * mc + c) * 7 + m] = er[1 + m];
        // main.cpp:1249-1271: running maxima (the max macro of main.cpp:14),
        // normalised norms, and the break test.
        bool below = true;
        for (int m = 1; m <= 5; ++m) {
          double& Em = E[(size_t)c * 5 + m - 1];
          Em = Em > er[m] ? Em : er[m];
          below = below && er[m] / Em < tol;
        }
        if (below || it == max_inner) ended.push_back(c);
      }
      for (int c : ended) done[c] = it;
      running -= (int)ended.size();
      if (running > 0 && !ended.empty()) {
        // Cases that end their inner loop while others go on: shift them now
        // and keep their iterate aside until the step ends; their cubes keep
        // being computed with the rest, and the results are discarded.
        complete_exchange(e);
        for (int c : ended) {
          shift_case(e, e->cases[c]);
          stash_case(e, e->cases[c], true);
          stashed.push_back(c);
        }
      }
    }
    if (stashed.empty()) {
      engine_shift(e);
    } else {
      for (int c = 0; c < mc; ++c)
        if (done[c] > 0 && std::find(stashed.begin(), stashed.end(), c) == stashed.end())
          shift_case(e, e->cases[c]);
      for (int c : stashed) stash_case(e, e->cases[c], false);
      // Restored iterates: the next pre-flux exchange is the full one.
      e->halo_full_next = true;
    }
    if (e->prec == 1 && e->Uf != startf) {
      launch_copy_f(e->stream, e->Uf, startf, (size_t)e->ncube * CUBE_STRIDE);
      std::swap(e->Uf, e->Uf2);
    }
    if (e->prec == 0 && e->U != startd) {
      launch_copy_d(e->stream, e->U, startd, (size_t)e->ncube * CUBE_STRIDE);
      std::swap(e->U, e->U2);
    }
    for (int c = 0; c < mc; ++c) iters[(size_t)s * mc + c] = done[c] > 0 ? done[c] : 0;
  }
  if (n_steps > 0) complete_exchange(e);
  HIP_TRY(hipGetLastError());
  return 0;
}

int engine_sync(Engine* e, std::string& err) {
  HIP_TRY(hipStreamSynchronize(e->stream));
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace bcm
