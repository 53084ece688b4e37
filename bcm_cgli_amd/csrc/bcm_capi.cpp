// C ABI of libbcm_cgli.so: host case construction, the device engine and the
// drop-in replacements of the reference's BCM_* routines.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/bcm_cgli.h"
#include "bcm_engine.h"
#include "bcm_setup.h"

using namespace bcm;

namespace {
thread_local std::string g_err;

int fail(const std::string& m) {
  g_err = m;
  return 1;
}

[[noreturn]] void fatal(const char* where, const std::string& m) {
  fprintf(stderr, "bcm_cgli: %s: %s\n", where, m.c_str());
  fflush(stderr);
  abort();
}
}  // namespace

void bcm::set_last_error(const std::string& m) { g_err = m; }

struct bcm_case {
  Case cs;
  std::vector<int> ib_minus_gc, ib_minus_ip, ib_plus_gc, ib_plus_ip;
};

// Per-engine drop-in binding: host arrays of the caller and device copies of
// the tables it passes per call.
struct DropIn {
  int ncube = 0;
  double *U1_ = nullptr, *U1 = nullptr, *U1q = nullptr, *U1p1 = nullptr, *U1p2 = nullptr;
  double *Rku1 = nullptr, *Residual1 = nullptr, *cube_size = nullptr;
  int* FWS = nullptr;
  struct Bc {
    int* cubes = nullptr;
    uint8_t* masks = nullptr;
    int n = 0;
  };
  std::map<std::tuple<int, const int*, int, const int*, int>, Bc> bc;
  std::map<std::tuple<const int*, const int*, const double*, int>, GhostDevice> gh;
  std::map<std::vector<const void*>, HaloSrc*> halo;
  std::map<std::vector<const void*>, std::pair<MlTaskDev*, int>> ml;
  std::vector<void*> owned;
};

struct bcm_engine {
  Engine* e = nullptr;
  std::unique_ptr<DropIn> di;
};

namespace {
bcm_engine* g_current = nullptr;

template <typename T>
T* upload_vec(DropIn& d, const std::vector<T>& v) {
  T* p = nullptr;
  size_t n = v.empty() ? 1 : v.size();
  if (hipMalloc((void**)&p, n * sizeof(T)) != hipSuccess) return nullptr;
  if (!v.empty() && hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return nullptr;
  d.owned.push_back(p);
  return p;
}

bcm_engine* current(const char* where, double* U1_) {
  if (!g_current || !g_current->di) fatal(where, "no engine bound (call bcm_bind first)");
  if (U1_ && U1_ != g_current->di->U1_) fatal(where, "U1_ is not the bound state array");
  return g_current;
}

void check(const char* where, hipError_t st) {
  if (st != hipSuccess) fatal(where, hipGetErrorString(st));
}

void run_bc_axis(const char* where, int axis, int nl, int nu, int* l, int* u, double* U1_) {
  bcm_engine* h = current(where, U1_);
  DropIn& d = *h->di;
  if (nl < 0 || nu < 0) fatal(where, "negative list length");
  auto key = std::make_tuple(axis, (const int*)l, nl, (const int*)u, nu);
  auto it = d.bc.find(key);
  if (it == d.bc.end()) {
    std::vector<uint8_t> mask(d.ncube, 0);
    for (int n = 1; n <= nl; ++n) {
      if (l[n] < 1 || l[n] >= d.ncube) fatal(where, "cube index out of range");
      mask[l[n]] |= (uint8_t)(1 << (2 * axis));
    }
    for (int n = 1; n <= nu; ++n) {
      if (u[n] < 1 || u[n] >= d.ncube) fatal(where, "cube index out of range");
      mask[u[n]] |= (uint8_t)(1 << (2 * axis + 1));
    }
    std::vector<int> cubes;
    std::vector<uint8_t> m;
    for (int c = 1; c < d.ncube; ++c)
      if (mask[c]) {
        cubes.push_back(c);
        m.push_back(mask[c]);
      }
    DropIn::Bc b;
    b.n = (int)cubes.size();
    b.cubes = upload_vec(d, cubes);
    b.masks = upload_vec(d, m);
    if (!b.cubes || !b.masks) fatal(where, "device allocation failed");
    it = d.bc.emplace(key, b).first;
  }
  launch_bc(h->e->stream, h->e->U, it->second.cubes, it->second.masks, it->second.n, h->e->pt());
  check(where, hipGetLastError());
}

void run_ghost(const char* where, int* NNBC, double* weight, int* GCindex, int* IPsur,
               double* U1_) {
  bcm_engine* h = current(where, U1_);
  DropIn& d = *h->di;
  const int nbc = *NNBC;
  if (nbc < 0) fatal(where, "negative entry count");
  auto key = std::make_tuple((const int*)GCindex, (const int*)IPsur, (const double*)weight, nbc);
  auto it = d.gh.find(key);
  if (it == d.gh.end()) {
    std::vector<GhostEntry> ordered;
    for (int n = 1; n <= nbc; ++n) {
      const int* g = &GCindex[(n - 1) * 4 + 1];
      const int* q = &IPsur[(n - 1) * 4 + 1];
      if (g[0] < 1 || g[0] >= d.ncube || g[0] != q[0] || g[1] < 0 || g[1] > 19 || g[2] < 0 ||
          g[2] > 19 || g[3] < 0 || g[3] > 19 || q[1] < 0 || q[1] > 18 || q[2] < 0 ||
          q[2] > 18 || q[3] < 0 || q[3] > 18)
        fatal(where, "IB entry out of range or interpolating from another cube");
      GhostEntry E;
      E.cube = g[0];
      E.gi = (int16_t)g[1]; E.gj = (int16_t)g[2]; E.gk = (int16_t)g[3];
      E.pi = (int16_t)q[1]; E.pj = (int16_t)q[2]; E.pk = (int16_t)q[3];
      for (int m = 0; m < 8; ++m) E.w[m] = weight[(n - 1) * 8 + 1 + m];
      ordered.push_back(E);
    }
    GhostSchedule gs = build_ghost_schedule(ordered);
    GhostDevice g;
    std::string err;
    if (ghost_upload(gs, g, err)) fatal(where, err);
    d.owned.insert(d.owned.end(), g.owned.begin(), g.owned.end());
    g.owned.clear();
    it = d.gh.emplace(key, g).first;
  }
  launch_ghost(h->e->stream, h->e->U, it->second.t, h->e->pt());
  check(where, hipGetLastError());
}
}  // namespace

extern "C" {

const char* bcm_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Host cases.
namespace {
CaseParams to_params(const bcm_case_params* p) {
  CaseParams cp;
  for (int d = 0; d < 3; ++d) {
    cp.lattice[d] = p->lattice[d];
    cp.origin[d] = p->origin[d];
    cp.sphere_center[d] = p->sphere_center[d];
  }
  cp.cube_size = p->cube_size;
  cp.sphere_diameter = p->sphere_diameter;
  cp.ico_level = p->ico_level;
  cp.mach = p->mach;
  cp.reynolds = p->reynolds;
  cp.seed = p->seed;
  cp.noise_amp = p->noise_amp;
  for (int d = 0; d < 3; ++d) {
    cp.refine_lo[d] = p->refine_lo[d];
    cp.refine_hi[d] = p->refine_hi[d];
    cp.refine2_lo[d] = p->refine2_lo[d];
    cp.refine2_hi[d] = p->refine2_hi[d];
  }
  return cp;
}

// IB tables also exposed with their 1-based index arrays.
void finish_handle(bcm_case* c) {
  c->ib_minus_gc = c->cs.ib_minus.gcindex;
  c->ib_plus_gc = c->cs.ib_plus.gcindex;
  c->ib_minus_ip = c->cs.ib_minus.ipsur;
  c->ib_plus_ip = c->cs.ib_plus.ipsur;
}
}  // namespace

int bcm_case_build(const bcm_case_params* p, bcm_case** out) {
  if (!p || !out) return fail("null argument");
  *out = nullptr;
  std::unique_ptr<bcm_case> c(new bcm_case());
  std::string err;
  if (build_case(to_params(p), c->cs, err)) return fail(err);
  finish_handle(c.get());
  *out = c.release();
  return 0;
}

int bcm_case_build_on(const bcm_case_params* p, int device, bcm_case** out) {
  if (!p || !out) return fail("null argument");
  *out = nullptr;
  std::unique_ptr<bcm_case> c(new bcm_case());
  std::string err;
  if (build_case_on_device(to_params(p), device, c->cs, err)) return fail(err);
  finish_handle(c.get());
  *out = c.release();
  return 0;
}

int bcm_case_read_on(const char* dir, const bcm_case_params* p, int device, bcm_case** out) {
  if (!dir || !p || !out) return fail("null argument");
  *out = nullptr;
  std::unique_ptr<bcm_case> c(new bcm_case());
  std::string err;
  if (read_case_on_device(dir, to_params(p), device, c->cs, err)) return fail(err);
  finish_handle(c.get());
  *out = c.release();
  return 0;
}

int bcm_case_read(const char* dir, const bcm_case_params* p, bcm_case** out) {
  if (!dir || !p || !out) return fail("null argument");
  *out = nullptr;
  std::unique_ptr<bcm_case> c(new bcm_case());
  std::string err;
  if (read_case(dir, to_params(p), c->cs, err)) return fail(err);
  finish_handle(c.get());
  *out = c.release();
  return 0;
}

void bcm_case_free(bcm_case* c) { delete c; }

int bcm_case_info(const bcm_case* c, int* info, double* phys) {
  if (!c) return fail("null case");
  const Case& cs = c->cs;
  if (info) {
    int v[16] = {cs.ncube,
                 cs.n_wallcube,
                 cs.ib_minus.n,
                 cs.ib_plus.n,
                 (int)cs.bc_lists[0].size() - 1,
                 (int)cs.bc_lists[1].size() - 1,
                 (int)cs.bc_lists[2].size() - 1,
                 (int)cs.bc_lists[3].size() - 1,
                 (int)cs.bc_lists[4].size() - 1,
                 (int)cs.bc_lists[5].size() - 1,
                 cs.nadj[L_X_EQ],
                 cs.nadj[L_Y_EQ],
                 cs.nadj[L_Z_EQ],
                 (int)cs.stl.size(),
                 cs.n_ghost_candidates,
                 0};
    memcpy(info, v, sizeof v);
  }
  if (phys) {
    const Physics& p = cs.phys;
    double v[8] = {p.rho0, p.U0, p.Uin, p.V0, p.W0, p.P0, p.T0, p.mu_L};
    memcpy(phys, v, sizeof v);
  }
  return 0;
}

void* bcm_case_array(bcm_case* c, const char* name, long long* count) {
  if (!c || !name) {
    fail("null argument");
    return nullptr;
  }
  Case& cs = c->cs;
  auto ret = [&](auto& v) -> void* {
    if (count) *count = (long long)v.size();
    return (void*)v.data();
  };
  std::string n(name);
  if (n == "U1_") return ret(cs.U1_);
  if (n == "Xcnt") return ret(cs.Xcnt);
  if (n == "Ycnt") return ret(cs.Ycnt);
  if (n == "Zcnt") return ret(cs.Zcnt);
  if (n == "U1") return ret(cs.U1);
  if (n == "U1q") return ret(cs.U1q);
  if (n == "cube_size") return ret(cs.cube_size);
  if (n == "Xcube") return ret(cs.Xcube);
  if (n == "Ycube") return ret(cs.Ycube);
  if (n == "Zcube") return ret(cs.Zcube);
  if (n == "FWS") return ret(cs.fws);
  if (n == "FWS_grid") return ret(cs.fws_grid);
  if (n == "csl") return ret(cs.csl);
  if (n == "adj_number") return ret(cs.adj_number);
  if (n == "ib_minus_gc") return ret(c->ib_minus_gc);
  if (n == "ib_minus_ip") return ret(c->ib_minus_ip);
  if (n == "ib_plus_gc") return ret(c->ib_plus_gc);
  if (n == "ib_plus_ip") return ret(c->ib_plus_ip);
  if (n == "ib_minus_w") return ret(cs.ib_minus.weight);
  if (n == "ib_plus_w") return ret(cs.ib_plus.weight);
  const char* bcn[6] = {"Xbc_l", "Xbc_u", "Ybc_l", "Ybc_u", "Zbc_l", "Zbc_u"};
  for (int s = 0; s < 6; ++s)
    if (n == bcn[s]) return ret(cs.bc_lists[s]);
  static const char* lists[L_COUNT] = {
      "adjX_eq",       "adjY_eq",       "adjZ_eq",        "adjX_bs_plus",  "adjX_sb_plus",
      "adjX_bs_minus", "adjX_sb_minus", "adjY_bs_plus",   "adjY_sb_plus",  "adjY_bs_minus",
      "adjY_sb_minus", "adjZ_bs_plus",  "adjZ_sb_plus",   "adjZ_bs_minus", "adjZ_sb_minus"};
  for (int l = 0; l < L_COUNT; ++l)
    if (n == lists[l]) return ret(cs.adj[l]);
  if (n == "nadj") return ret(cs.nadj);
  fail("unknown case array " + n);
  return nullptr;
}

int bcm_case_write_inputs(const bcm_case* c, const char* dir) {
  if (!c || !dir) return fail("null argument");
  std::string d(dir);
  if (write_bcmgrid(c->cs, d + "/BCMgrid00000.dat")) return fail("cannot write grid file");
  if (write_stl(c->cs, d + "/BCM_STL_fine.stl")) return fail("cannot write STL file");
  if (write_rank_tables(c->cs.ncube - 1, d)) return fail("cannot write the rank tables");
  return 0;
}

int bcm_case_write_ibm(const bcm_case* c, const char* dir) {
  if (!c || !dir) return fail("null argument");
  if (write_ibm(c->cs, dir)) return fail(std::string("cannot write the IB tables in ") + dir);
  return 0;
}

int bcm_case_write_plot3d(const bcm_case* c, const double* aos, int step, const char* dir) {
  if (!c || !dir) return fail("null argument");
  if (step < 0 || step > 99999) return fail("output step must lie in 0..99999");
  const Case& cs = c->cs;
  std::string err;
  if (write_plot3d(cs.ncube, cs.Xcnt.data(), cs.Ycnt.data(), cs.Zcnt.data(), cs.ncube - 1,
                   aos ? aos : cs.U1_.data(), step, dir, err))
    return fail(err);
  return 0;
}

int bcm_write_restart(const char* path, int ncube, long long step, const double* E5,
                      const double* U1_, const double* U1, const double* U1q) {
  if (!path || !U1_ || !U1 || !U1q) return fail("null argument");
  if (ncube < 2) return fail("a restart needs at least one cube");
  std::string err;
  if (write_restart(path, ncube, step, E5, U1_, U1, U1q, err)) return fail(err);
  return 0;
}

int bcm_read_restart(const char* path, int ncube, long long* step, double* E5, double* U1_,
                     double* U1, double* U1q) {
  if (!path || !U1_ || !U1 || !U1q) return fail("null argument");
  std::string err;
  if (read_restart(path, ncube, step, E5, U1_, U1, U1q, err)) return fail(err);
  return 0;
}

// ---------------------------------------------------------------------------
// Engine.
int bcm_engine_create(int device, bcm_engine** out) {
  if (!out) return fail("null argument");
  *out = nullptr;
  std::string err;
  Engine* e = nullptr;
  if (engine_create(device, &e, err)) return fail(err);
  bcm_engine* h = new bcm_engine();
  h->e = e;
  *out = h;
  return 0;
}

void bcm_engine_destroy(bcm_engine* h) {
  if (!h) return;
  if (g_current == h) g_current = nullptr;
  if (h->di)
    for (void* p : h->di->owned) (void)hipFree(p);
  engine_destroy(h->e);
  delete h;
}

int bcm_engine_load_case(bcm_engine* h, const bcm_case* c) {
  if (!h || !c) return fail("null argument");
  std::string err;
  if (engine_load_case(h->e, c->cs, err)) return fail(err);
  return 0;
}

int bcm_engine_alloc_pool(bcm_engine* h, int cap_cubes, int max_cases, int precision) {
  if (!h) return fail("null engine");
  if (g_current == h) g_current = nullptr;
  if (h->di) {
    for (void* p : h->di->owned) (void)hipFree(p);
    h->di.reset();
  }
  std::string err;
  if (engine_alloc(h->e, cap_cubes + 1, max_cases, err, precision)) return fail(err);
  h->e->flux_tuned = false;
  h->e->flux_pinned = false;
  return 0;
}

int bcm_engine_add_case(bcm_engine* h, const bcm_case* c, int* slot) {
  if (!h || !c || !slot) return fail("null argument");
  if (h->di) return fail("engine is bound to host arrays");
  if (h->e->cap == 0) return fail("no pool allocated");
  std::string err;
  if (engine_add_case(h->e, c->cs, slot, err)) return fail(err);
  return 0;
}

int bcm_engine_remove_case(bcm_engine* h, int slot) {
  if (!h) return fail("null engine");
  std::string err;
  if (engine_remove_case(h->e, slot, err)) return fail(err);
  return 0;
}

int bcm_engine_case_state(bcm_engine* h, int slot, int which, double* aos, int upload) {
  if (!h || !aos) return fail("null argument");
  if (which < 0 || which > 6) return fail("unknown state array");
  std::string err;
  if (engine_case_state(h->e, slot, which, aos, upload != 0, err)) return fail(err);
  return 0;
}

int bcm_engine_output(bcm_engine* h, int slot, int step, const char* dir) {
  if (!h || !dir) return fail("null argument");
  std::string err;
  if (engine_output(h->e, slot, step, dir, err)) return fail(err);
  return 0;
}

int bcm_engine_output_wait(bcm_engine* h) {
  if (!h) return fail("null engine");
  std::string err;
  if (engine_output_wait(h->e, err)) return fail(err);
  return 0;
}

int bcm_engine_save_restart(bcm_engine* h, int slot, const char* path, long long step,
                            const double* E5) {
  if (!h || !path) return fail("null argument");
  Engine* e = h->e;
  if (slot < 0 || slot >= e->max_cases || !e->cases[slot].active) return fail("no such case slot");
  const int ncube = e->cases[slot].tb.n + 1;
  std::vector<double> st[3];
  std::string err;
  for (int w = 0; w < 3; ++w) {
    st[w].resize((size_t)ncube * CUBE_STRIDE);
    if (engine_case_state(e, slot, w, st[w].data(), false, err)) return fail(err);
  }
  if (write_restart(path, ncube, step, E5, st[0].data(), st[1].data(), st[2].data(), err))
    return fail(err);
  return 0;
}

int bcm_engine_load_restart(bcm_engine* h, int slot, const char* path, long long* step,
                            double* E5) {
  if (!h || !path) return fail("null argument");
  Engine* e = h->e;
  if (slot < 0 || slot >= e->max_cases || !e->cases[slot].active) return fail("no such case slot");
  const int ncube = e->cases[slot].tb.n + 1;
  std::vector<double> st[3];
  for (auto& v : st) v.resize((size_t)ncube * CUBE_STRIDE);
  std::string err;
  if (read_restart(path, ncube, step, E5, st[0].data(), st[1].data(), st[2].data(), err))
    return fail(err);
  for (int w = 0; w < 3; ++w)
    if (engine_case_state(e, slot, w, st[w].data(), true, err)) return fail(err);
  return 0;
}

int bcm_engine_case_er(bcm_engine* h, int slot, int inner, double* er8) {
  if (!h || !er8) return fail("null argument");
  Engine* e = h->e;
  if (slot < 0 || slot >= e->max_cases) return fail("no such case slot");
  if (inner < 0 || inner >= e->er_slots) return fail("no such inner iteration record");
  if (hipMemcpy(er8, e->er_dev + ((size_t)inner * e->max_cases + slot) * 8, 8 * sizeof(double),
                hipMemcpyDeviceToHost) != hipSuccess)
    return fail("er copy failed");
  return 0;
}

int bcm_engine_pool_info(bcm_engine* h, int* info6) {
  if (!h || !info6) return fail("null argument");
  const Engine* e = h->e;
  info6[0] = e->cap > 0 ? e->cap - 1 : 0;
  info6[1] = e->ncube > 0 ? e->ncube - 1 : 0;
  info6[2] = e->n_active;
  info6[3] = e->max_cases;
  info6[4] = e->cap > 0 ? engine_free_range(e) : 0;
  int used = 0;
  for (const CaseSlot& s : e->cases)
    if (s.active) used += s.tb.n;
  info6[5] = used;
  return 0;
}

long long bcm_bytes_per_cube(int precision) {
  return (long long)engine_bytes_per_cube(precision == 1 ? 1 : 0);
}

int bcm_engine_precision(bcm_engine* h) { return h ? h->e->prec : -1; }

int bcm_engine_set_time(bcm_engine* h, double deltaT, double deltaTau, double e_floor) {
  if (!h) return fail("null engine");
  if (!(deltaT > 0) || !(deltaTau > 0)) return fail("time steps must be positive");
  h->e->deltaT = deltaT;
  h->e->deltaTau = deltaTau;
  h->e->e = e_floor;
  h->e->graph_valid = false;
  return 0;
}

int bcm_engine_run(bcm_engine* h, int n_steps, int n_inner, double* er_hist, int use_graph) {
  if (!h) return fail("null engine");
  if (n_steps < 0 || n_inner < 1) return fail("invalid step counts");
  std::string err;
  if (engine_run(h->e, n_steps, n_inner, er_hist, use_graph != 0, err)) return fail(err);
  if (engine_sync(h->e, err)) return fail(err);
  return 0;
}

int bcm_engine_run_converge(bcm_engine* h, int n_steps, int max_inner, double tol, double* E,
                            int* iters, double* er_hist) {
  if (!h) return fail("null engine");
  if (n_steps < 0 || max_inner < 1) return fail("invalid step counts");
  if (!E || !iters) return fail("null norm maxima or iteration counts");
  std::string err;
  if (engine_run_converge(h->e, n_steps, max_inner, tol, E, iters, er_hist, err)) return fail(err);
  if (engine_sync(h->e, err)) return fail(err);
  return 0;
}

int bcm_engine_part(bcm_engine* h, int part, int RK) {
  if (!h) return fail("null engine");
  Engine* e = h->e;
  if (e->n_active == 0) return fail("no case loaded");
  if (RK < 1 || RK > 3) return fail("RK must be 1, 2 or 3");
  if (e->prec == 1 && (part == 3 || part == 4))
    return fail("the mixed path fuses flux and update (use part 7)");
  switch (part) {
    case 0: engine_bc(e); break;
    case 1: engine_ghost(e); break;
    case 2: engine_interface(e); break;
    case 3: engine_flux(e, RK); break;
    case 4: engine_update(e, RK, 0); break;
    case 5: engine_shift(e); break;
    case 6: engine_stage(e, RK, 0, e->halo_full_next); e->halo_full_next = false; break;
    case 7: engine_flux_update(e, RK, 0); break;
    case 8: engine_halo_refresh(e); break;
    case 9: engine_flux_push(e, RK, 0); break;
    default: return fail("unknown part");
  }
  std::string err;
  if (engine_sync(e, err)) return fail(err);
  return 0;
}

int bcm_engine_get_state(bcm_engine* h, int which, double* aos) {
  if (!h || !aos) return fail("null argument");
  if (which < 0 || which > 6) return fail("unknown state array");
  if (engine_download_state(h->e, which, aos)) return fail("download failed");
  return 0;
}

int bcm_engine_set_state(bcm_engine* h, int which, const double* aos) {
  if (!h || !aos) return fail("null argument");
  if (which < 0 || which > 6) return fail("unknown state array");
  h->e->graph_valid = false;
  if (engine_upload_state(h->e, which, aos)) return fail("upload failed");
  return 0;
}

int bcm_engine_get_er(bcm_engine* h, double* er8) {
  if (!h || !er8) return fail("null argument");
  if (hipMemcpy(er8, h->e->er_dev, 8 * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return fail("er copy failed");
  return 0;
}

int bcm_engine_sync(bcm_engine* h) {
  if (!h) return fail("null engine");
  std::string err;
  if (engine_sync(h->e, err)) return fail(err);
  return 0;
}

int bcm_engine_time_part(bcm_engine* h, int part, int RK, int reps, double* ms) {
  if (!h || !ms) return fail("null argument");
  Engine* e = h->e;
  if (e->n_active == 0) return fail("no case loaded");
  if (reps < 1) return fail("reps must be positive");
  if (e->prec == 1 && (part == 3 || part == 4))
    return fail("the mixed path fuses flux and update (use part 7)");
  // The pushing stage kernel is timed as the inner iterations of a step run
  // it, without the stores of Rku1, U1p1, U1p2 and Residual1 that only a
  // step's last inner iteration makes.
  e->push_aux = part != 9;
  if (hipEventRecord(e->ev0, e->stream) != hipSuccess) return fail("event record failed");
  for (int r = 0; r < reps; ++r) {
    switch (part) {
      case 0: engine_bc(e); break;
      case 1: engine_ghost(e); break;
      case 2: engine_interface(e); break;
      case 3: engine_flux(e, RK); break;
      case 4: engine_update(e, RK, 0); break;
      case 5: engine_shift(e); break;
      case 6: engine_stage(e, RK, 0); break;
      case 7: engine_flux_update(e, RK, 0); break;
      case 8: engine_halo_refresh(e); break;
      case 9: engine_flux_push(e, RK, 0); break;
      default: e->push_aux = true; return fail("unknown part");
    }
  }
  e->push_aux = true;
  if (hipEventRecord(e->ev1, e->stream) != hipSuccess) return fail("event record failed");
  if (hipEventSynchronize(e->ev1) != hipSuccess) return fail("event wait failed");
  float t = 0;
  if (hipEventElapsedTime(&t, e->ev0, e->ev1) != hipSuccess) return fail("event timing failed");
  *ms = (double)t / reps;
  std::string err;
  if (engine_sync(e, err)) return fail(err);
  return 0;
}

int bcm_engine_set_flux(bcm_engine* h, int variant, int nseg, int fuse) {
  if (!h) return fail("null engine");
  if (h->e->prec == 1) {
    // One kernel; the choice is its runs per cube (variant and fuse ignored).
    if (nseg != 1 && nseg != 2 && nseg != 4) return fail("mixed path: nseg must be 1, 2 or 4");
    h->e->mixed_nseg = nseg;
    h->e->flux_tuned = true;
    h->e->flux_pinned = true;
    h->e->graph_valid = false;
    return 0;
  }
  if (variant != 0 && variant != 1) return fail("flux variant must be 0 or 1");
  if (nseg != 1 && nseg != 2 && nseg != 4 && nseg != 8 && nseg != 16)
    return fail("nseg must divide 16");
  if (fuse < 0 || fuse > 2) return fail("fuse must be 0, 1 or 2");
  if (fuse == 2 && (variant != 1 || nseg != 1)) return fail("fuse 2 needs variant 1, nseg 1");
  h->e->flux_fuse = fuse != 0;
  h->e->flux_push = fuse == 2;
  h->e->flux_variant = variant;
  h->e->flux_nseg = nseg;
  h->e->flux_tuned = true;
  h->e->flux_pinned = true;
  h->e->graph_valid = false;
  return 0;
}

int bcm_engine_autotune(bcm_engine* h, int* out3) {
  if (!h) return fail("null engine");
  std::string err;
  int v = 0, n = 1;
  h->e->flux_pinned = false;
  if (engine_autotune(h->e, &v, &n, err)) return fail(err);
  if (out3) {
    out3[0] = v;
    out3[1] = n;
    out3[2] = h->e->flux_push ? 2 : (h->e->flux_fuse ? 1 : 0);
  }
  return 0;
}

int bcm_engine_ghost_stats(bcm_engine* h, int* stats3) {
  if (!h || !stats3) return fail("null argument");
  stats3[0] = h->e->n_ghost_entries;
  stats3[1] = h->e->n_ghost_cubes;
  stats3[2] = h->e->ghost_levels;
  return 0;
}

int bcm_selftest_division(unsigned long long seed, int n, unsigned long long* mismatches) {
  if (!mismatches || n < 0) return fail("bad argument");
  if (div_selftest(seed, n, mismatches)) return fail("HIP error in the division self-test");
  return 0;
}

int bcm_engine_set_ghost_lds_limit(bcm_engine* h, int slots) {
  if (!h) return fail("null engine");
  if (slots < 0 || slots > GHOST_LDS_SLOTS)
    return fail("slot limit must be in 0.." + std::to_string(GHOST_LDS_SLOTS));
  h->e->ghost_slot_limit = slots;
  return 0;
}

// ---------------------------------------------------------------------------
// Drop-ins.
int bcm_bind(bcm_engine* h, int ncube, int n_cubes_norm, double* U1_, double* U1, double* U1q,
             double* U1p1, double* U1p2, double* Rku1, double* Residual1, int* FWS,
             double* cube_size, const double* phys7) {
  if (!h || !U1_ || !U1 || !U1q || !U1p1 || !U1p2 || !Rku1 || !Residual1 || !FWS || !cube_size ||
      !phys7)
    return fail("null argument");
  std::string err;
  if (h->di)
    for (void* p : h->di->owned) (void)hipFree(p);
  h->di.reset(new DropIn());
  if (engine_alloc(h->e, ncube, 1, err)) return fail(err);
  Engine* e = h->e;
  const double rho0 = phys7[0], P0 = phys7[5];
  const DevPhys ph{rho0, phys7[1], phys7[2], phys7[3], phys7[4], P0, P0 / rho0 / 287.0, phys7[6]};
  if (engine_bind_single(e, n_cubes_norm, ph, err)) return fail(err);
  DropIn& d = *h->di;
  d.ncube = ncube;
  d.U1_ = U1_; d.U1 = U1; d.U1q = U1q; d.U1p1 = U1p1; d.U1p2 = U1p2;
  d.Rku1 = Rku1; d.Residual1 = Residual1; d.FWS = FWS; d.cube_size = cube_size;
  // Empty halo table until BCM_Interface is called.
  g_current = h;
  return bcm_sync_to_device(h);
}

int bcm_sync_to_device(bcm_engine* h) {
  if (!h || !h->di) return fail("engine not bound");
  DropIn& d = *h->di;
  Engine* e = h->e;
  double* arr[7] = {d.U1_, d.U1, d.U1q, d.U1p1, d.U1p2, d.Rku1, d.Residual1};
  for (int w = 0; w < 7; ++w)
    if (engine_upload_state(e, w, arr[w])) return fail("state upload failed");
  if (engine_upload_flags(e, d.FWS)) return fail("flag upload failed");
  if (hipMemcpy(e->cube_size, d.cube_size, d.ncube * sizeof(double), hipMemcpyHostToDevice) !=
      hipSuccess)
    return fail("cube size upload failed");
  return 0;
}

int bcm_sync_to_host(bcm_engine* h) {
  if (!h || !h->di) return fail("engine not bound");
  DropIn& d = *h->di;
  double* arr[7] = {d.U1_, d.U1, d.U1q, d.U1p1, d.U1p2, d.Rku1, d.Residual1};
  for (int w = 0; w < 7; ++w)
    if (engine_download_state(h->e, w, arr[w])) return fail("state download failed");
  return 0;
}

void BCM_X_boundary_condition(int myid, int nXbc_l, int nXbc_u, int* Xbc_l, int* Xbc_u,
                              double* U1_) {
  (void)myid;
  run_bc_axis("BCM_X_boundary_condition", 0, nXbc_l, nXbc_u, Xbc_l, Xbc_u, U1_);
}

void BCM_Y_boundary_condition(int nYbc_l, int nYbc_u, int* Ybc_l, int* Ybc_u, double* U1_) {
  run_bc_axis("BCM_Y_boundary_condition", 1, nYbc_l, nYbc_u, Ybc_l, Ybc_u, U1_);
}

void BCM_Z_boundary_condition(int nZbc_l, int nZbc_u, int* Zbc_l, int* Zbc_u, double* U1_) {
  run_bc_axis("BCM_Z_boundary_condition", 2, nZbc_l, nZbc_u, Zbc_l, Zbc_u, U1_);
}

void BCM_Ghostcell_minus(int myid, int* NNBC, double* weight, int* GCindex, int* IPsur,
                         int* FWS, double* U1_) {
  (void)myid;
  (void)FWS;
  run_ghost("BCM_Ghostcell_minus", NNBC, weight, GCindex, IPsur, U1_);
}

void BCM_Ghostcell_plus(int myid, int* NNBC, double* weight, int* GCindex, int* IPsur,
                        int* FWS, double* U1_) {
  (void)myid;
  (void)FWS;
  run_ghost("BCM_Ghostcell_plus", NNBC, weight, GCindex, IPsur, U1_);
}

void BCM_Interface_table(int myid, int ncube, int n_wallcube, int* nadjX_eq, int* nadjY_eq,
                         int* nadjZ_eq, int* nadjX_bs_plus, int* nadjX_sb_plus,
                         int* nadjX_bs_minus, int* nadjX_sb_minus, int* nadjY_bs_plus,
                         int* nadjY_sb_plus, int* nadjY_bs_minus, int* nadjY_sb_minus,
                         int* nadjZ_bs_plus, int* nadjZ_sb_plus, int* nadjZ_bs_minus,
                         int* nadjZ_sb_minus, int* csl, int* adj_number, int* adjX_eq,
                         int* adjY_eq, int* adjZ_eq, int* adjX_bs_plus, int* adjX_sb_plus,
                         int* adjX_bs_minus, int* adjX_sb_minus, int* adjY_bs_plus,
                         int* adjY_sb_plus, int* adjY_bs_minus, int* adjY_sb_minus,
                         int* adjZ_bs_plus, int* adjZ_sb_plus, int* adjZ_bs_minus,
                         int* adjZ_sb_minus) {
  (void)myid;
  (void)n_wallcube;
  int nadj[L_COUNT];
  int* lists[L_COUNT] = {adjX_eq,       adjY_eq,       adjZ_eq,       adjX_bs_plus,
                         adjX_sb_plus,  adjX_bs_minus, adjX_sb_minus, adjY_bs_plus,
                         adjY_sb_plus,  adjY_bs_minus, adjY_sb_minus, adjZ_bs_plus,
                         adjZ_sb_plus,  adjZ_bs_minus, adjZ_sb_minus};
  interface_tables(ncube, csl, adj_number, nadj, lists);
  int* outs[L_COUNT] = {nadjX_eq,       nadjY_eq,       nadjZ_eq,       nadjX_bs_plus,
                        nadjX_sb_plus,  nadjX_bs_minus, nadjX_sb_minus, nadjY_bs_plus,
                        nadjY_sb_plus,  nadjY_bs_minus, nadjY_sb_minus, nadjZ_bs_plus,
                        nadjZ_sb_plus,  nadjZ_bs_minus, nadjZ_sb_minus};
  for (int l = 0; l < L_COUNT; ++l) *outs[l] = nadj[l];
}

void BCM_Interface(int myid, int ncube, int mPI_Nadj, int ncpu_bs, int ncpu_eq, int ncpu_sb,
                   int max_nei_bs, int max_nei_eq, int max_nei_sb, int nadjX_eq, int nadjY_eq,
                   int nadjZ_eq, int nadjX_bs_plus, int nadjX_sb_plus, int nadjX_bs_minus,
                   int nadjX_sb_minus, int nadjY_bs_plus, int nadjY_sb_plus, int nadjY_bs_minus,
                   int nadjY_sb_minus, int nadjZ_bs_plus, int nadjZ_sb_plus, int nadjZ_bs_minus,
                   int nadjZ_sb_minus, int* rank_map, int* MPI_cpu, int* MPI_cube,
                   int* MPI_cpu_adj, int* MPI_cube_adj, int* MPI_direction, int* MPI_interface,
                   int* neighbor_cpu_eq, int* Ncube_Ncpu_eq, int* neighbor_cpu_sb,
                   int* Ncube_Ncpu_sb, int* neighbor_cpu_bs, int* Ncube_Ncpu_bs,
                   int* Scube_Ncpu_eq, int* Rcube_Ncpu_eq, double* send_data_curr_eq,
                   double* recv_data_curr_eq, double* send_data_neig_eq,
                   double* recv_data_neig_eq, int* Sdir_eq, int* Rdir_eq, int* Scube_Ncpu_sb,
                   int* Rcube_Ncpu_sb, double* send_data_curr_sb, double* recv_data_curr_sb,
                   double* send_data_neig_sb, double* recv_data_neig_sb, int* Sdir_sb,
                   int* Rdir_sb, int* Scube_Ncpu_bs, int* Rcube_Ncpu_bs,
                   double* send_data_curr_bs, double* recv_data_curr_bs,
                   double* send_data_neig_bs, double* recv_data_neig_bs, int* Sdir_bs,
                   int* Rdir_bs, int* ist_eq, int* ist_sb, int* ist_bs, int* adjN_sb,
                   int* RadjN_bs, int* SadjN_bs, int* csl, int* adj_number, int* adjX_eq,
                   int* adjY_eq, int* adjZ_eq, int* adjX_bs_plus, int* adjX_sb_plus,
                   int* adjX_bs_minus, int* adjX_sb_minus, int* adjY_bs_plus,
                   int* adjY_sb_plus, int* adjY_bs_minus, int* adjY_sb_minus,
                   int* adjZ_bs_plus, int* adjZ_sb_plus, int* adjZ_bs_minus,
                   int* adjZ_sb_minus, double* U1_) {
  const char* where = "BCM_Interface";
  bcm_engine* h = current(where, U1_);
  DropIn& d = *h->di;
  (void)myid; (void)mPI_Nadj; (void)max_nei_bs; (void)max_nei_eq; (void)max_nei_sb;
  (void)rank_map; (void)MPI_cpu; (void)MPI_cube; (void)MPI_cpu_adj; (void)MPI_cube_adj;
  (void)MPI_direction; (void)MPI_interface; (void)neighbor_cpu_eq; (void)Ncube_Ncpu_eq;
  (void)neighbor_cpu_sb; (void)Ncube_Ncpu_sb; (void)neighbor_cpu_bs; (void)Ncube_Ncpu_bs;
  (void)Scube_Ncpu_eq; (void)Rcube_Ncpu_eq; (void)send_data_curr_eq; (void)recv_data_curr_eq;
  (void)send_data_neig_eq; (void)recv_data_neig_eq; (void)Sdir_eq; (void)Rdir_eq;
  (void)Scube_Ncpu_sb; (void)Rcube_Ncpu_sb; (void)send_data_curr_sb; (void)recv_data_curr_sb;
  (void)send_data_neig_sb; (void)recv_data_neig_sb; (void)Sdir_sb; (void)Rdir_sb;
  (void)Scube_Ncpu_bs; (void)Rcube_Ncpu_bs; (void)send_data_curr_bs; (void)recv_data_curr_bs;
  (void)send_data_neig_bs; (void)recv_data_neig_bs; (void)Sdir_bs; (void)Rdir_bs;
  (void)ist_eq; (void)ist_sb; (void)ist_bs; (void)adjN_sb; (void)RadjN_bs; (void)SadjN_bs;
  (void)csl; (void)adjX_sb_plus; (void)adjX_sb_minus; (void)adjY_sb_plus; (void)adjY_sb_minus;
  (void)adjZ_sb_plus; (void)adjZ_sb_minus;
  (void)nadjX_sb_plus; (void)nadjX_sb_minus; (void)nadjY_sb_plus; (void)nadjY_sb_minus;
  (void)nadjZ_sb_plus; (void)nadjZ_sb_minus;
  if (ncpu_bs != 0 || ncpu_eq != 0 || ncpu_sb != 0)
    fatal(where, "neighbour ranks are not supported on a single device");
  if (ncube != d.ncube) fatal(where, "ncube differs from the bound arrays");
  std::vector<const void*> key = {adjX_eq, adjY_eq, adjZ_eq, adj_number,
                                  (const void*)(intptr_t)nadjX_eq, (const void*)(intptr_t)nadjY_eq,
                                  (const void*)(intptr_t)nadjZ_eq};
  auto it = d.halo.find(key);
  if (it == d.halo.end()) {
    std::vector<HaloSrc> hs(ncube);
    for (auto& x : hs) memset(&x, 0, sizeof x);
    const int n[3] = {nadjX_eq, nadjY_eq, nadjZ_eq};
    const int* l[3] = {adjX_eq, adjY_eq, adjZ_eq};
    for (int a = 0; a < 3; ++a)
      for (int m = 1; m <= n[a]; ++m) {
        int c0 = l[a][m];
        if (c0 < 1 || c0 >= ncube) fatal(where, "interface list entry out of range");
        int c2 = adj_number[adj_index(c0, 1, 2 * a + 2)];
        if (c2 < 1 || c2 >= ncube) fatal(where, "interface neighbour out of range");
        hs[c2].src[2 * a] = c0;
        hs[c0].src[2 * a + 1] = c2;
      }
    HaloSrc* dp = upload_vec(d, hs);
    if (!dp) fatal(where, "device allocation failed");
    it = d.halo.emplace(key, dp).first;
  }
  launch_halo_edge(h->e->stream, h->e->U, it->second, ncube, h->e->cube_case);
  check(where, hipGetLastError());
  // Two-level faces (the bs lists; the sb lists serve only the reference's
  // exchanges with other ranks), then the edge passes again.
  const int nbs[6] = {nadjX_bs_plus, nadjX_bs_minus, nadjY_bs_plus,
                      nadjY_bs_minus, nadjZ_bs_plus, nadjZ_bs_minus};
  if (nbs[0] || nbs[1] || nbs[2] || nbs[3] || nbs[4] || nbs[5]) {
    const int* bsl[6] = {adjX_bs_plus, adjX_bs_minus, adjY_bs_plus,
                         adjY_bs_minus, adjZ_bs_plus, adjZ_bs_minus};
    std::vector<const void*> mkey = {adj_number};
    for (int q = 0; q < 6; ++q) {
      mkey.push_back(bsl[q]);
      mkey.push_back((const void*)(intptr_t)nbs[q]);
    }
    auto mt = d.ml.find(mkey);
    if (mt == d.ml.end()) {
      int nadj[L_COUNT] = {0};
      const int* lists[L_COUNT] = {nullptr};
      const int li[6] = {L_X_BS_PLUS, L_X_BS_MINUS, L_Y_BS_PLUS,
                         L_Y_BS_MINUS, L_Z_BS_PLUS, L_Z_BS_MINUS};
      for (int q = 0; q < 6; ++q) {
        nadj[li[q]] = nbs[q];
        lists[li[q]] = bsl[q];
        for (int m = 1; m <= nbs[q]; ++m)
          if (bsl[q][m] < 1 || bsl[q][m] >= ncube) fatal(where, "interface list entry out of range");
      }
      std::vector<MlTaskDev> tasks;
      for (const MlTask& t : ml_tasks(lists, nadj, adj_number)) {
        if (t.dst < 1 || t.dst >= ncube || t.src < 1 || t.src >= ncube)
          fatal(where, "two-level neighbour out of range");
        tasks.push_back(MlTaskDev{t.dst, t.src, t.axis, t.hi, t.restrict_, t.slot});
      }
      MlTaskDev* dp = upload_vec(d, tasks);
      if (!dp) fatal(where, "device allocation failed");
      mt = d.ml.emplace(mkey, std::make_pair(dp, (int)tasks.size())).first;
    }
    launch_halo_ml(h->e->stream, h->e->U, mt->second.first, mt->second.second, h->e->pt());
    launch_edge(h->e->stream, h->e->U, ncube, h->e->cube_case);
    check(where, hipGetLastError());
  }
}

void BCM_Flux_XYZ_Viscous_Runge_kutta(int myid, int ncube, int RK, double deltaT,
                                      double deltaTau, double e, int* FWS, int* csl,
                                      double* cube_size, double* U1_, double* U1, double* U1q,
                                      double* U1p1, double* U1p2, double* Rku1,
                                      double* Residual1, double* er) {
  const char* where = "BCM_Flux_XYZ_Viscous_Runge_kutta";
  bcm_engine* h = current(where, U1_);
  DropIn& d = *h->di;
  (void)myid;
  (void)csl;
  if (ncube != d.ncube || U1 != d.U1 || U1q != d.U1q || U1p1 != d.U1p1 || U1p2 != d.U1p2 ||
      Rku1 != d.Rku1 || Residual1 != d.Residual1 || FWS != d.FWS || cube_size != d.cube_size)
    fatal(where, "arguments differ from the bound arrays");
  if (RK < 1 || RK > 3) fatal(where, "RK must be 1, 2 or 3");
  Engine* en = h->e;
  en->deltaT = deltaT;
  en->deltaTau = deltaTau;
  en->e = e;
  engine_flux_update(en, RK, 0);
  check(where, hipGetLastError());
  if (RK == 3) {
    double tmp[8];
    check(where, hipMemcpyAsync(tmp, en->er_dev, sizeof tmp, hipMemcpyDeviceToHost, en->stream));
    check(where, hipStreamSynchronize(en->stream));
    for (int m = 1; m <= 7; ++m) er[m] = tmp[m];
  }
}

void BCM_Grid(int myid, int ncube, int n_wallcube, double* cube_size, int* csl, double* Xcube,
              double* Ycube, double* Zcube, double* Xcnt, double* Ycnt, double* Zcnt, int* FWS,
              int* adj_number, int* wallcube) {
  const char* where = "BCM_Grid";
  char name[64];
  snprintf(name, sizeof name, "BCMgrid%05d.dat", myid);
  Case cs;
  std::string err;
  if (read_bcmgrid(name, cs, err)) fatal(where, err);
  if (cs.ncube != ncube || cs.n_wallcube != n_wallcube)
    fatal(where, "cube or wall-cube count differs from the grid file");
  grid_geometry(cs);
  // Entries the reference's reader sets: cubes 1..ncube-1, adjacency slots
  // 1..4 of directions 1..6, wall cubes 1..n_wallcube-1.
  for (int c = 1; c < ncube; ++c) {
    cube_size[c] = cs.cube_size[c];
    Xcube[c] = cs.Xcube[c];
    Ycube[c] = cs.Ycube[c];
    Zcube[c] = cs.Zcube[c];
    csl[c] = cs.csl[c];
    memcpy(Xcnt + (size_t)c * NP, &cs.Xcnt[(size_t)c * NP], NP * sizeof(double));
    memcpy(Ycnt + (size_t)c * NP, &cs.Ycnt[(size_t)c * NP], NP * sizeof(double));
    memcpy(Zcnt + (size_t)c * NP, &cs.Zcnt[(size_t)c * NP], NP * sizeof(double));
    memcpy(FWS + (size_t)c * CELLS_PADDED, &cs.fws[(size_t)c * CELLS_PADDED],
           CELLS_PADDED * sizeof(int));
    for (int slot = 1; slot <= 4; ++slot)
      for (int d = 1; d <= 6; ++d)
        adj_number[adj_index(c, slot, d)] = cs.adj_number[adj_index(c, slot, d)];
  }
  for (int w = 1; w < n_wallcube; ++w) wallcube[w] = cs.wallcube[w];
}

void BCM_Reading_IBM(int myid, int* NNBC, int mp_switch, double* weight, int* GCindex,
                     int* IPsur) {
  const char* where = "BCM_Reading_IBM";
  if (!NNBC || (mp_switch != 1 && mp_switch != -1)) fatal(where, "mp_switch must be 1 or -1");
  char name[64];
  snprintf(name, sizeof name, "IBM_%s%05d.dat", mp_switch == 1 ? "plus" : "minus", myid);
  IBTable tb;
  std::string err;
  if (read_ibm(name, tb, err)) fatal(where, err);
  const int n = *NNBC;
  if (n < 0 || n > tb.n) fatal(where, "the file holds fewer entries than requested");
  memcpy(GCindex + 1, tb.gcindex.data() + 1, (size_t)4 * n * sizeof(int));
  memcpy(IPsur + 1, tb.ipsur.data() + 1, (size_t)4 * n * sizeof(int));
  memcpy(weight + 1, tb.weight.data() + 1, (size_t)8 * n * sizeof(double));
}

void BCM_Output(int myid, int ncube, int step, int* rank_map, double* U1, double* U1q,
                double* cube_size, double* Xcnt, double* Ycnt, double* Zcnt) {
  const char* where = "BCM_Output";
  (void)myid;
  (void)rank_map;
  (void)U1q;
  (void)cube_size;
  if (step < 0 || step > 99999) fatal(where, "step must lie in 0..99999");
  // The state of a bound engine lives on the device: bring U1_ home first.
  if (g_current && g_current->di && U1 == g_current->di->U1_) {
    if (g_current->di->ncube != ncube) fatal(where, "ncube differs from the bound arrays");
    if (engine_download_state(g_current->e, 0, U1)) fatal(where, "state download failed");
  }
  std::string err;
  if (write_plot3d(ncube, Xcnt, Ycnt, Zcnt, ncube - 1, U1, step, ".", err)) fatal(where, err);
}

void BCM_Physical_time_shift(double* U1_, double* U1, double* U1q) {
  const char* where = "BCM_Physical_time_shift";
  bcm_engine* h = current(where, U1_);
  if (U1 != h->di->U1 || U1q != h->di->U1q) fatal(where, "arguments differ from the bound arrays");
  engine_shift(h->e);
  check(where, hipGetLastError());
}

}  // extern "C"
