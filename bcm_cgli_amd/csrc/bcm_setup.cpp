// Host-side case construction.  Every pass restates a setup step of the
// reference solver so that the flags, IB tables and initial state handed to
// the device are the ones the reference computes for the same grid and
// surface.  Citations are file:line in the reference tree.
#include "bcm_setup.h"
#include "bcm_geom.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <thread>
#include <map>

namespace bcm {

namespace {

// Runs fn(i) for i in [0, n) on up to 16 host threads (fewer when
// OMP_NUM_THREADS says so).  Every use writes disjoint data per i, so the
// result does not depend on the schedule.
template <typename F>
void parallel_for(int n, F fn) {
  int nth = (int)std::thread::hardware_concurrency();
  if (const char* env = getenv("OMP_NUM_THREADS")) nth = std::min(nth, atoi(env));
  nth = std::max(1, std::min(nth, 16));
  if (n <= 1 || nth == 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  nth = std::min(nth, n);
  std::atomic<int> next(0);
  std::vector<std::thread> pool;
  for (int t = 0; t < nth; ++t)
    pool.emplace_back([&]() {
      for (int i; (i = next.fetch_add(1)) < n;) fn(i);
    });
  for (auto& th : pool) th.join();
}

inline double dmin(double a, double b) { return ((a) < (b)) ? (a) : (b); }
inline double dmax(double a, double b) { return ((a) > (b)) ? (a) : (b); }

}  // namespace

// ---------------------------------------------------------------------------
// Physical constants.  prm.h:25-37 fixes rho0, P0, V0, W0 and T0 = P0/rho0/R;
// the flow case chooses U0 = Uin from the Mach number and mu_L from the
// Reynolds number on the sphere diameter.
void compute_physics(const CaseParams& p, Physics& ph) {
  ph.rho0 = 1.1842;
  ph.P0 = 101300.0;
  ph.V0 = 0.0;
  ph.W0 = 0.0;
  const double c0 = sqrt(KGAS * ph.P0 / ph.rho0);
  ph.U0 = p.mach * c0;
  ph.Uin = ph.U0;
  ph.mu_L = ph.rho0 * ph.U0 * p.sphere_diameter / p.reynolds;
  ph.T0 = ph.P0 / ph.rho0 / RGAS;
}

// ---------------------------------------------------------------------------
// Counter-based noise: splitmix64 of (seed, cube, i, j, k) mapped to [-1, 1).
double noise_unit(uint64_t seed, int cube, int i, int j, int k) {
  uint64_t key = (uint64_t)cube * (uint64_t)CELLS_PADDED +
                 (uint64_t)((i * NP + j) * NP + k);
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + key + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

// ---------------------------------------------------------------------------
// Icosphere: a regular icosahedron with counter-clockwise (outward) faces,
// each triangle split into four per level, new vertices pushed to the sphere.
void make_icosphere(int level, const double center[3], double radius,
                    std::vector<Triangle>& tris) {
  const double t = (1.0 + sqrt(5.0)) / 2.0;
  std::vector<std::array<double, 3>> v = {
      {-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0},
      {0, -1, t}, {0, 1, t}, {0, -1, -t}, {0, 1, -t},
      {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto& a : v) {
    double n = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    a[0] /= n; a[1] /= n; a[2] /= n;
  }
  std::vector<std::array<int, 3>> f = {
      {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
      {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
      {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
      {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  for (int l = 0; l < level; ++l) {
    std::map<std::pair<int, int>, int> mid;
    auto midpoint = [&](int a, int b) {
      std::pair<int, int> key(std::min(a, b), std::max(a, b));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      std::array<double, 3> m = {(v[a][0] + v[b][0]) * 0.5, (v[a][1] + v[b][1]) * 0.5,
                                 (v[a][2] + v[b][2]) * 0.5};
      double n = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      m[0] /= n; m[1] /= n; m[2] /= n;
      v.push_back(m);
      int id = (int)v.size() - 1;
      mid[key] = id;
      return id;
    };
    std::vector<std::array<int, 3>> g;
    g.reserve(f.size() * 4);
    for (auto& tri : f) {
      int a = tri[0], b = tri[1], c = tri[2];
      int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      g.push_back({a, ab, ca});
      g.push_back({b, bc, ab});
      g.push_back({c, ca, bc});
      g.push_back({ab, bc, ca});
    }
    f.swap(g);
  }
  tris.clear();
  tris.reserve(f.size());
  for (auto& tri : f) {
    Triangle T;
    for (int m = 0; m < 3; ++m)
      for (int d = 0; d < 3; ++d) T.v[m][d] = center[d] + radius * v[tri[m]][d];
    double e1[3], e2[3];
    for (int d = 0; d < 3; ++d) {
      e1[d] = T.v[1][d] - T.v[0][d];
      e2[d] = T.v[2][d] - T.v[0][d];
    }
    double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                   e1[0] * e2[1] - e1[1] * e2[0]};
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int d = 0; d < 3; ++d) T.n[d] = n[d] / nn;
    tris.push_back(T);
  }
}

// ---------------------------------------------------------------------------
// Interface lists, BCM_Interface_table.cpp:96-305.  An equal-size pair is
// recorded once, from its + side; the - side only records size changes.
void interface_tables(int ncube, const int* csl, const int* adj, int* nadj,
                      int* const* lists) {
  for (int l = 0; l < L_COUNT; ++l) nadj[l] = 0;
  auto A = [&](int c, int s, int d) { return adj[adj_index(c, s, d)]; };
  auto push = [&](int l, int c) { nadj[l] += 1; lists[l][nadj[l]] = c; };
  struct Dir { int dir, eq, sb, bs; bool plus; };
  // Order of the six tests per cube follows the reference: +X, +Y, +Z, -X, -Y, -Z.
  const Dir dirs[6] = {{2, L_X_EQ, L_X_SB_PLUS, L_X_BS_PLUS, true},
                       {4, L_Y_EQ, L_Y_SB_PLUS, L_Y_BS_PLUS, true},
                       {6, L_Z_EQ, L_Z_SB_PLUS, L_Z_BS_PLUS, true},
                       {1, -1, L_X_SB_MINUS, L_X_BS_MINUS, false},
                       {3, -1, L_Y_SB_MINUS, L_Y_BS_MINUS, false},
                       {5, -1, L_Z_SB_MINUS, L_Z_BS_MINUS, false}};
  for (int c = 1; c < ncube; ++c) {
    for (const Dir& d : dirs) {
      int a1 = A(c, 1, d.dir), a2 = A(c, 2, d.dir), a3 = A(c, 3, d.dir), a4 = A(c, 4, d.dir);
      if (a1 > 0 || a2 > 0 || a3 > 0 || a4 > 0) {
        int acube = std::max(std::max(std::max(a1, a2), a3), a4);
        if (d.plus) {
          if (csl[c] == csl[acube]) push(d.eq, c);
          else if (csl[c] < csl[acube]) push(d.sb, c);
          else if (csl[c] > csl[acube]) push(d.bs, c);
        } else {
          if (csl[c] < csl[acube]) push(d.sb, c);
          if (csl[c] > csl[acube]) push(d.bs, c);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
std::vector<MlTask> ml_tasks(const int* const* lists, const int* nadj, const int* adj) {
  std::vector<MlTask> out;
  const int bs_plus[3] = {L_X_BS_PLUS, L_Y_BS_PLUS, L_Z_BS_PLUS};
  const int bs_minus[3] = {L_X_BS_MINUS, L_Y_BS_MINUS, L_Z_BS_MINUS};
  for (int A = 0; A < 3; ++A) {
    const int dplus = 2 * A + 2, dminus = 2 * A + 1;
    // Small cubes on the big cube's upper side: their lower ghosts.
    for (int a = 1; a <= nadj[bs_plus[A]]; ++a) {
      const int big = lists[bs_plus[A]][a];
      for (int s = 1; s <= 4; ++s) {
        const int small = adj[adj_index(big, s, dplus)];
        if (small > 0) out.push_back({small, big, (int8_t)A, 0, 0, (int8_t)s});
      }
    }
    // Small cubes on the big cube's lower side: their upper ghosts.
    for (int a = 1; a <= nadj[bs_minus[A]]; ++a) {
      const int big = lists[bs_minus[A]][a];
      for (int s = 1; s <= 4; ++s) {
        const int small = adj[adj_index(big, s, dminus)];
        if (small > 0) out.push_back({small, big, (int8_t)A, 1, 0, (int8_t)s});
      }
    }
    // The big cube's upper ghosts from the small cubes above it.
    for (int a = 1; a <= nadj[bs_plus[A]]; ++a) {
      const int big = lists[bs_plus[A]][a];
      for (int s = 1; s <= 4; ++s) {
        const int small = adj[adj_index(big, s, dplus)];
        if (small > 0) out.push_back({big, small, (int8_t)A, 1, 1, (int8_t)s});
      }
    }
    // The big cube's lower ghosts from the small cubes below it.
    for (int a = 1; a <= nadj[bs_minus[A]]; ++a) {
      const int big = lists[bs_minus[A]][a];
      for (int s = 1; s <= 4; ++s) {
        const int small = adj[adj_index(big, s, dminus)];
        if (small > 0) out.push_back({big, small, (int8_t)A, 0, 1, (int8_t)s});
      }
    }
  }
  return out;
}

void ml_apply_host(const MlTask& t, double* U) {
  const int A = t.axis;
  const int T0 = A == 0 ? 1 : 0, T1 = A == 2 ? 1 : 2;   // tangential axes
  const int o0 = (t.slot == 2 || t.slot == 4) ? NC / 2 : 0;
  const int o1 = (t.slot == 3 || t.slot == 4) ? NC / 2 : 0;
  int d[3], q[3];
  if (!t.restrict_) {
    for (int l = 0; l < NB; ++l)
      for (int p = ILO; p <= IHI; ++p)
        for (int r = ILO; r <= IHI; ++r) {
          d[A] = t.hi ? IHI + 1 + l : l;
          d[T0] = p;
          d[T1] = r;
          q[A] = t.hi ? NB + l / NB : IHI + l / NB;
          q[T0] = p / NB + 1 + o0;
          q[T1] = r / NB + 1 + o1;
          for (int v = 0; v < NV; ++v)
            U[aos_index(t.dst, d[0], d[1], d[2], v)] = U[aos_index(t.src, q[0], q[1], q[2], v)];
        }
    return;
  }
  for (int l = 0; l < NB; ++l)
    for (int p = ILO; p <= (IHI - NB + 1) / 2 + 1; ++p)
      for (int r = ILO; r <= (IHI - NB + 1) / 2 + 1; ++r) {
        d[A] = t.hi ? NB + NC + l : l;
        d[T0] = p + o0;
        d[T1] = r + o1;
        q[A] = t.hi ? NB * (l + 1) : IHI - 2 * NB + NB * (l + 1) - 1;
        q[T0] = (p - NB + 1) * 2;
        q[T1] = (r - NB + 1) * 2;
        const int i2 = q[0], j2 = q[1], k2 = q[2];
        for (int v = 0; v < NV; ++v) {
          auto S = [&](int a, int b, int c) { return U[aos_index(t.src, a, b, c, v)]; };
          U[aos_index(t.dst, d[0], d[1], d[2], v)] =
              0.125 * (S(i2, j2, k2) + S(i2 + 1, j2, k2) + S(i2, j2 + 1, k2) + S(i2, j2, k2 + 1) +
                       S(i2 + 1, j2 + 1, k2) + S(i2 + 1, j2, k2 + 1) + S(i2, j2 + 1, k2 + 1) +
                       S(i2 + 1, j2 + 1, k2 + 1));
        }
      }
}

// ---------------------------------------------------------------------------
// Equal-size halo copies and the edge-ghost averaging of BCM_Interface on a
// host state array (Interface.cpp:1745-1782, 2604-2641, 3474-3511, 5640-5732).
void host_interface_copy(int ncube, const int* const* lists, const int* nadj,
                         const int* adj, double* U) {
  auto at = [&](int c, int i, int j, int k, int v) -> double& {
    return U[aos_index(c, i, j, k, v)];
  };
  parallel_for(nadj[L_X_EQ], [&](int a) {
    int c0 = lists[L_X_EQ][a + 1], c2 = adj[adj_index(c0, 1, 2)];
    for (int i = 0; i < NB; ++i)
      for (int j = ILO; j <= IHI; ++j)
        for (int k = ILO; k <= IHI; ++k)
          for (int v = 0; v < NV; ++v) {
            at(c2, i, j, k, v) = at(c0, NC + i, j, k, v);
            at(c0, NB + NC + i, j, k, v) = at(c2, NB + i, j, k, v);
          }
  });
  parallel_for(nadj[L_Y_EQ], [&](int a) {
    int c0 = lists[L_Y_EQ][a + 1], c2 = adj[adj_index(c0, 1, 4)];
    for (int i = ILO; i <= IHI; ++i)
      for (int j = 0; j < NB; ++j)
        for (int k = ILO; k <= IHI; ++k)
          for (int v = 0; v < NV; ++v) {
            at(c2, i, j, k, v) = at(c0, i, NC + j, k, v);
            at(c0, i, NB + NC + j, k, v) = at(c2, i, NB + j, k, v);
          }
  });
  parallel_for(nadj[L_Z_EQ], [&](int a) {
    int c0 = lists[L_Z_EQ][a + 1], c2 = adj[adj_index(c0, 1, 6)];
    for (int i = ILO; i <= IHI; ++i)
      for (int j = ILO; j <= IHI; ++j)
        for (int k = 0; k < NB; ++k)
          for (int v = 0; v < NV; ++v) {
            at(c2, i, j, k, v) = at(c0, i, j, NC + k, v);
            at(c0, i, j, NB + NC + k, v) = at(c2, i, j, NB + k, v);
          }
  });
  // Two-level faces: each task writes its own ghost region from interior
  // cells, so the tasks are independent.
  const std::vector<MlTask> ml = ml_tasks(lists, nadj, adj);
  parallel_for((int)ml.size(), [&](int n) { ml_apply_host(ml[n], U); });
  const int L1 = NB - 1, H1 = IHI + 1;  // layers 1 and 18
  parallel_for(ncube - 1, [&](int c0) {
    const int c = c0 + 1;
    for (int i = ILO; i <= IHI; ++i)
      for (int v = 0; v < NV; ++v) {
        at(c, i, L1, L1, v) = 0.5 * (at(c, i, L1, ILO, v) + at(c, i, ILO, L1, v));
        at(c, i, H1, L1, v) = 0.5 * (at(c, i, H1, ILO, v) + at(c, i, IHI, L1, v));
        at(c, i, L1, H1, v) = 0.5 * (at(c, i, L1, IHI, v) + at(c, i, ILO, H1, v));
        at(c, i, H1, H1, v) = 0.5 * (at(c, i, H1, IHI, v) + at(c, i, IHI, H1, v));
      }
    for (int j = L1; j <= H1; ++j)
      for (int v = 0; v < NV; ++v) {
        at(c, L1, j, L1, v) = 0.5 * (at(c, L1, j, ILO, v) + at(c, ILO, j, L1, v));
        at(c, L1, j, H1, v) = 0.5 * (at(c, ILO, j, H1, v) + at(c, L1, j, IHI, v));
        at(c, H1, j, L1, v) = 0.5 * (at(c, IHI, j, L1, v) + at(c, H1, j, ILO, v));
        at(c, H1, j, H1, v) = 0.5 * (at(c, IHI, j, H1, v) + at(c, H1, j, IHI, v));
      }
    for (int k = L1; k <= H1; ++k)
      for (int v = 0; v < NV; ++v) {
        at(c, L1, L1, k, v) = 0.5 * (at(c, L1, ILO, k, v) + at(c, ILO, L1, k, v));
        at(c, L1, H1, k, v) = 0.5 * (at(c, ILO, H1, k, v) + at(c, L1, IHI, k, v));
        at(c, H1, L1, k, v) = 0.5 * (at(c, IHI, L1, k, v) + at(c, H1, ILO, k, v));
        at(c, H1, H1, k, v) = 0.5 * (at(c, IHI, H1, k, v) + at(c, H1, IHI, k, v));
      }
  });
}

namespace {

// Flag halo exchange plus hole treatment, BCM_FWS_Interface.cpp:114-783 at a
// single rank: equal-size face copies through a double copy of the flags,
// rounding back to {-1,0,1}, then cells with two solid opposite neighbours
// become solid until a sweep changes nothing (the sweep count in the
// reference grows with every change, so it always ends on a quiet sweep;
// the rule only ever adds solid cells, so the result is that fixed point).
void fws_interface(Case& cs) {
  const int ncube = cs.ncube;
  std::vector<double> tmp((size_t)ncube * CELLS_PADDED, 0.0);
  parallel_for(ncube - 1, [&](int c0) {
    const size_t c = c0 + 1;
    if (cs.csl[c] == 0)
      for (int n = 0; n < CELLS_PADDED; ++n) tmp[c * CELLS_PADDED + n] = cs.fws[c * CELLS_PADDED + n];
  });
  auto T = [&](int c, int i, int j, int k) -> double& { return tmp[flag_index(c, i, j, k)]; };
  const int* adj = cs.adj_number.data();
  parallel_for(cs.nadj[L_X_EQ], [&](int a) {
    int c0 = cs.adj[L_X_EQ][a + 1], c2 = adj[adj_index(c0, 1, 2)];
    for (int i = 0; i < NB; ++i)
      for (int j = ILO; j <= IHI; ++j)
        for (int k = ILO; k <= IHI; ++k) {
          T(c2, i, j, k) = T(c0, NC + i, j, k);
          T(c0, NB + NC + i, j, k) = T(c2, NB + i, j, k);
        }
  });
  parallel_for(cs.nadj[L_Y_EQ], [&](int a) {
    int c0 = cs.adj[L_Y_EQ][a + 1], c2 = adj[adj_index(c0, 1, 4)];
    for (int i = ILO; i <= IHI; ++i)
      for (int j = 0; j < NB; ++j)
        for (int k = ILO; k <= IHI; ++k) {
          T(c2, i, j, k) = T(c0, i, NC + j, k);
          T(c0, i, NB + NC + j, k) = T(c2, i, NB + j, k);
        }
  });
  parallel_for(cs.nadj[L_Z_EQ], [&](int a) {
    int c0 = cs.adj[L_Z_EQ][a + 1], c2 = adj[adj_index(c0, 1, 6)];
    for (int i = ILO; i <= IHI; ++i)
      for (int j = ILO; j <= IHI; ++j)
        for (int k = 0; k < NB; ++k) {
          T(c2, i, j, k) = T(c0, i, j, NC + k);
          T(c0, i, j, NB + NC + k) = T(c2, i, j, NB + k);
        }
  });
  parallel_for(ncube - 1, [&](int c0) {
    const int c = c0 + 1;
    if (cs.csl[c] != 0) return;
    for (int n = 0; n < CELLS_PADDED; ++n) {
      double t = tmp[(size_t)c * CELLS_PADDED + n];
      int& f = cs.fws[(size_t)c * CELLS_PADDED + n];
      if (t > 0.5) f = 1;
      else if (t < -0.5) f = -1;
      else f = 0;
    }
    // Hole treatment of the cube (reads and writes only this cube).
    auto F = [&](int i, int j, int k) -> int& { return cs.fws[flag_index(c, i, j, k)]; };
    bool changed = true;
    while (changed) {
      changed = false;
      for (int i = NB - 1; i <= IHI + 1; ++i)
        for (int j = NB - 1; j <= IHI + 1; ++j)
          for (int k = NB - 1; k <= IHI + 1; ++k) {
            if (F(i, j, k) >= IGHOST &&
                ((F(i + 1, j, k) + F(i - 1, j, k) == 2 * ISOLID) ||
                 (F(i, j + 1, k) + F(i, j - 1, k) == 2 * ISOLID) ||
                 (F(i, j, k + 1) + F(i, j, k - 1) == 2 * ISOLID))) {
              F(i, j, k) = ISOLID;
              changed = true;
            }
          }
    }
  });
}

// Gauss-Jordan inverse with partial pivoting in the operation order of
// BCM_Immersed_boundary.cpp:1498-1579.
void inverse_matrix(const double* in, double* out, int n) {
  double a[8][8], inv[8][8];
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) { a[j][i] = in[i + n * j]; inv[j][i] = (i == j) ? 1.0 : 0.0; }
  for (int k = 0; k < n; ++k) {
    int mx = k;
    for (int j = k + 1; j < n; ++j)
      if (fabs(a[j][k]) > fabs(a[mx][k])) mx = j;
    if (mx != k)
      for (int i = 0; i < n; ++i) {
        std::swap(a[mx][i], a[k][i]);
        std::swap(inv[mx][i], inv[k][i]);
      }
    double tmp = a[k][k];
    for (int i = 0; i < n; ++i) { a[k][i] /= tmp; inv[k][i] /= tmp; }
    for (int j = 0; j < n; ++j) {
      if (j == k) continue;
      tmp = a[j][k] / a[k][k];
      for (int i = 0; i < n; ++i) {
        a[j][i] = a[j][i] - a[k][i] * tmp;
        inv[j][i] = inv[j][i] - inv[k][i] * tmp;
      }
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) out[i + n * j] = inv[j][i];
}

// Weights are serialised with "%f" and read back with "%lf" in the
// reference (BCM_Immersed_boundary.cpp:1017,1032; BCM_Reading_IBM.cpp:63).
double round_f6(double w) {
  char buf[64];
  snprintf(buf, sizeof buf, "%f", w);
  return strtod(buf, nullptr);
}

struct GhostHit {
  bool hit;
  double tmin, dir[3], dotp, bip[3];
};

}  // namespace

void surface_tri(const Triangle& T, SurfaceTri& s) {
  for (int d = 0; d < 3; ++d) {
    s.v0[d] = T.v[0][d];
    s.v1[d] = T.v[1][d];
    s.v2[d] = T.v[2][d];
  }
  double v12[3], v13[3];
  for (int d = 0; d < 3; ++d) { v12[d] = T.v[1][d] - T.v[0][d]; v13[d] = T.v[2][d] - T.v[0][d]; }
  s.nu[0] = v12[1] * v13[2] - v12[2] * v13[1];
  s.nu[1] = v12[2] * v13[0] - v12[0] * v13[2];
  s.nu[2] = v12[0] * v13[1] - v12[1] * v13[0];
  const double dis = sqrt(s.nu[0] * s.nu[0] + s.nu[1] * s.nu[1] + s.nu[2] * s.nu[2]);
  s.dir[0] = -s.nu[0] / dis;
  s.dir[1] = -s.nu[1] / dis;
  s.dir[2] = -s.nu[2] / dis;
  // The edges and the triangle-only part of intersect_triangle.
  for (int d = 0; d < 3; ++d) { s.e1[d] = T.v[1][d] - T.v[0][d]; s.e2[d] = T.v[2][d] - T.v[0][d]; }
  s.pv[0] = s.dir[1] * s.e2[2] - s.dir[2] * s.e2[1];
  s.pv[1] = s.dir[2] * s.e2[0] - s.dir[0] * s.e2[2];
  s.pv[2] = s.dir[0] * s.e2[1] - s.dir[1] * s.e2[0];
  s.det = s.e1[0] * s.pv[0] + s.e1[1] * s.pv[1] + s.e1[2] * s.pv[2];
  s.inv_det = 1.0 / s.det;
  s.dd = s.dir[0] * s.dir[0] + s.dir[1] * s.dir[1] + s.dir[2] * s.dir[2];
}

namespace {

// Triangles with any vertex inside some cube box padded by 1.5 cells
// (BCM_Immersed_boundary.cpp:380-470), on the host.
void filter_host(const Case& cs, std::vector<uint8_t>& keep) {
  keep.assign(cs.stl.size(), 0);
  for (size_t t = 0; t < cs.stl.size(); ++t) {
    const Triangle& T = cs.stl[t];
    bool k = false;
    for (int m = 0; m < 3 && !k; ++m) {
      const double* pv = T.v[m];
      for (int c = 1; c < cs.ncube; ++c) {
        double dx = cs.cube_size[c] / NC, dy = dx, dz = dx;
        if (pv[0] >= cs.Xcube[c] - 1.5 * dx && pv[0] <= cs.Xcube[c] + cs.cube_size[c] + 1.5 * dx &&
            pv[1] >= cs.Ycube[c] - 1.5 * dy && pv[1] <= cs.Ycube[c] + cs.cube_size[c] + 1.5 * dy &&
            pv[2] >= cs.Zcube[c] - 1.5 * dz && pv[2] <= cs.Zcube[c] + cs.cube_size[c] + 1.5 * dz) {
          k = true;
          break;
        }
      }
    }
    keep[t] = k;
  }
}

// Surface search of one ghost cell over all kept triangles
// (BCM_Immersed_boundary.cpp:722-860): the ray along each triangle's inward
// normal with the smallest |t|, else the closest point on any triangle.
void search_host(const std::vector<SurfaceTri>& tri, const double orig[3], GhostHit& h) {
  h.hit = false;
  h.tmin = BIGNUM;
  h.dotp = 0.0;
  const int Ntri = (int)tri.size();
  double ndis_min = BIGNUM;
  for (int it = 0; it < Ntri; ++it) {
    const SurfaceTri& T = tri[it];
    double t = 0, u, v;
    int f = intersect_triangle(orig, T.dir, T.v0, T.v1, T.v2, &t, &u, &v);
    double ndis = fabs(t) * T.dd;
    if (f == 1) {
      h.hit = true;
      if (ndis < ndis_min) {
        ndis_min = ndis;
        h.tmin = fabs(t);
        h.dir[0] = T.dir[0]; h.dir[1] = T.dir[1]; h.dir[2] = T.dir[2];
        h.dotp = T.nu[0] * (T.v0[0] - orig[0]) + T.nu[1] * (T.v0[1] - orig[1]) +
                 T.nu[2] * (T.v0[2] - orig[2]);
      }
    }
  }
  if (!h.hit) {
    ndis_min = BIGNUM;
    for (int it = 0; it < Ntri; ++it) {
      double p[3];
      closest_point(tri[it].v0, tri[it].v1, tri[it].v2, orig, p);
      double nd = (p[0] - orig[0]) * (p[0] - orig[0]) + (p[1] - orig[1]) * (p[1] - orig[1]) +
                  (p[2] - orig[2]) * (p[2] - orig[2]);
      if (nd <= ndis_min) {
        h.bip[0] = p[0]; h.bip[1] = p[1]; h.bip[2] = p[2];
        ndis_min = nd;
      }
    }
  }
}

// Image-point construction and trilinear weights for every ghost cell,
// BCM_Immersed_boundary.cpp:127-1081.
int immersed_boundary(Case& cs, IbDevice* dev, std::string& err) {
  const int ncube = cs.ncube;
  std::vector<uint8_t> keep;
  if (dev) {
    if (dev->filter(cs, keep, err)) return 1;
  } else {
    filter_host(cs, keep);
  }
  std::vector<SurfaceTri> tri;
  for (size_t t = 0; t < cs.stl.size(); ++t)
    if (keep[t]) {
      SurfaceTri s;
      surface_tri(cs.stl[t], s);
      tri.push_back(s);
    }
  const int Ntri = (int)tri.size();
  // Ghost cells of finest cubes, interior only, cube-i-j-k order (:500-545).
  struct GC { int c, i, j, k; double x, y, z; };
  std::vector<GC> gcs;
  for (int c = 1; c < ncube; ++c) {
    if (cs.csl[c] != 0) continue;
    for (int i = ILO; i <= IHI; ++i)
      for (int j = ILO; j <= IHI; ++j)
        for (int k = ILO; k <= IHI; ++k)
          if (cs.fws[flag_index(c, i, j, k)] == IGHOST)
            gcs.push_back({c, i, j, k, cs.Xcnt[c * NP + i], cs.Ycnt[c * NP + j], cs.Zcnt[c * NP + k]});
  }
  cs.n_ghost_candidates = (int)gcs.size();
  const int Ngc = (int)gcs.size();
  std::vector<GhostHit> hits(Ngc);
  if (dev && Ntri > 0 && Ngc > 0) {
    // The device names the winning triangles; their values are recomputed
    // here exactly as the host search computes them.
    std::vector<double> xyz((size_t)3 * Ngc);
    for (int g = 0; g < Ngc; ++g) {
      xyz[3 * g] = gcs[g].x;
      xyz[3 * g + 1] = gcs[g].y;
      xyz[3 * g + 2] = gcs[g].z;
    }
    std::vector<int> hit_any, hit_tri, near_tri;
    if (dev->search(tri, xyz, hit_any, hit_tri, near_tri, err)) return 1;
    for (int g = 0; g < Ngc; ++g) {
      GhostHit& h = hits[g];
      h.hit = hit_any[g] != 0;
      h.tmin = BIGNUM;
      h.dotp = 0.0;
      const double orig[3] = {gcs[g].x, gcs[g].y, gcs[g].z};
      if (h.hit && hit_tri[g] >= 0) {
        const SurfaceTri& T = tri[hit_tri[g]];
        double t = 0, u, v;
        intersect_triangle(orig, T.dir, T.v0, T.v1, T.v2, &t, &u, &v);
        h.tmin = fabs(t);
        h.dir[0] = T.dir[0]; h.dir[1] = T.dir[1]; h.dir[2] = T.dir[2];
        h.dotp = T.nu[0] * (T.v0[0] - orig[0]) + T.nu[1] * (T.v0[1] - orig[1]) +
                 T.nu[2] * (T.v0[2] - orig[2]);
      } else if (!h.hit && near_tri[g] >= 0) {
        const SurfaceTri& T = tri[near_tri[g]];
        closest_point(T.v0, T.v1, T.v2, orig, h.bip);
      }
    }
  } else if (Ntri > 0) {
    // Independent across ghost cells, so the cells are split over host threads.
    int nth = (int)std::thread::hardware_concurrency();
    if (const char* env = getenv("OMP_NUM_THREADS")) nth = std::min(nth, atoi(env));
    if (nth < 1) nth = 1;
    if (nth > 16) nth = 16;
    std::atomic<int> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < nth; ++t)
      pool.emplace_back([&]() {
        for (int g; (g = next.fetch_add(16)) < Ngc;)
          for (int q = g; q < std::min(g + 16, Ngc); ++q) {
            const double orig[3] = {gcs[q].x, gcs[q].y, gcs[q].z};
            search_host(tri, orig, hits[q]);
          }
      });
    for (auto& th : pool) th.join();
  }
  // Sequential pass: a ghost cell without a ray hit reuses the orientation
  // (dotp) of the last ghost cell that had one, as the reference's dotp
  // variable carries over between iterations of its ghost loop.
  double dotp = 0.0;
  std::vector<int> ltmp;
  cs.ib_minus = IBTable();
  cs.ib_plus = IBTable();
  for (IBTable* tb : {&cs.ib_minus, &cs.ib_plus}) {
    tb->gcindex.assign(1, 0);
    tb->ipsur.assign(1, 0);
    tb->weight.assign(1, 0.0);
  }
  for (int g = 0; g < Ngc; ++g) {
    if (Ntri == 0) continue;
    const GhostHit& h = hits[g];
    if (h.hit) dotp = h.dotp;
    const double orig[3] = {gcs[g].x, gcs[g].y, gcs[g].z};
    const int c = gcs[g].c;
    double BI[3], IP[3];
    if (dotp == 0) {
      for (int d = 0; d < 3; ++d) BI[d] = IP[d] = orig[d];
    } else if (dotp > 0) {
      if (h.hit) {
        for (int d = 0; d < 3; ++d) {
          BI[d] = orig[d] - h.tmin * h.dir[d];
          IP[d] = orig[d] + h.tmin * h.dir[d];
        }
      } else {
        for (int d = 0; d < 3; ++d) { BI[d] = h.bip[d]; IP[d] = 2 * orig[d] - BI[d]; }
      }
    } else {
      if (h.hit) {
        for (int d = 0; d < 3; ++d) {
          BI[d] = orig[d] + h.tmin * h.dir[d];
          IP[d] = orig[d] - h.tmin * h.dir[d];
        }
      } else {
        for (int d = 0; d < 3; ++d) { BI[d] = h.bip[d]; IP[d] = 2 * orig[d] - BI[d]; }
      }
    }
    const double dx = cs.cube_size[c] / NC, dy = dx, dz = dx;
    double dis = sqrt((BI[0] - orig[0]) * (BI[0] - orig[0]) + (BI[1] - orig[1]) * (BI[1] - orig[1]) +
                      (BI[2] - orig[2]) * (BI[2] - orig[2]));
    if (dis > sqrt(2.0) * dx) {
      cs.fws[flag_index(c, gcs[g].i, gcs[g].j, gcs[g].k)] = IFLUID;
      continue;
    }
    int icount = 0, ii, jj, kk;
    const double* Xc = &cs.Xcnt[c * NP];
    const double* Yc = &cs.Ycnt[c * NP];
    const double* Zc = &cs.Zcnt[c * NP];
    for (ii = 0; ii <= IHI + 1; ++ii)
      if (IP[0] >= Xc[ii] && IP[0] <= Xc[ii] + dx) { icount++; break; }
    for (jj = 0; jj <= IHI + 1; ++jj)
      if (IP[1] >= Yc[jj] && IP[1] <= Yc[jj] + dy) { icount++; break; }
    for (kk = 0; kk <= IHI + 1; ++kk)
      if (IP[2] >= Zc[kk] && IP[2] <= Zc[kk] + dz) { icount++; break; }
    if (icount < 3) continue;
    double s1 = IP[0] - Xc[ii], s2 = IP[1] - Yc[jj], s3 = IP[2] - Zc[kk];
    double hx = Xc[ii + 1] - Xc[ii], hy = Yc[jj + 1] - Yc[jj], hz = Zc[kk + 1] - Zc[kk];
    const double sx[8] = {0, hx, 0, 0, hx, hx, 0, hx};
    const double sy[8] = {0, 0, hy, 0, hy, 0, hy, hy};
    const double sz[8] = {0, 0, 0, hz, 0, hz, hz, hz};
    double V[64];
    for (int vi = 0; vi < 8; ++vi) {
      V[vi * 8 + 7] = 1;
      V[vi * 8 + 6] = sz[vi];
      V[vi * 8 + 5] = sy[vi];
      V[vi * 8 + 4] = sx[vi];
      V[vi * 8 + 3] = sy[vi] * sz[vi];
      V[vi * 8 + 2] = sx[vi] * sz[vi];
      V[vi * 8 + 1] = sx[vi] * sy[vi];
      V[vi * 8 + 0] = sx[vi] * sy[vi] * sz[vi];
    }
    inverse_matrix(V, V, 8);
    double w[8];
    for (int vi = 0; vi < 8; ++vi)
      w[vi] = V[vi] * s1 * s2 * s3 + V[vi + 8] * s1 * s2 + V[vi + 16] * s1 * s3 +
              V[vi + 24] * s2 * s3 + V[vi + 32] * s1 + V[vi + 40] * s2 + V[vi + 48] * s3 +
              V[vi + 56];
    bool ok = true;
    for (int vi = 0; vi < 8; ++vi)
      ok = ok && (w[vi] >= (-SMALLNUM) && w[vi] <= (1 + SMALLNUM));
    if (!ok) continue;
    IBTable& tb = (dotp >= 0) ? cs.ib_minus : cs.ib_plus;
    tb.n += 1;
    const int gidx[4] = {c, gcs[g].i, gcs[g].j, gcs[g].k};
    const int pidx[4] = {c, ii, jj, kk};
    for (int m = 0; m < 4; ++m) { tb.gcindex.push_back(gidx[m]); tb.ipsur.push_back(pidx[m]); }
    for (int vi = 0; vi < 8; ++vi) tb.weight.push_back(round_f6(w[vi]));
  }
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
namespace {

// nx x ny x nz cubes of one size, numbered x fastest, one neighbour per face.
void uniform_lattice(const CaseParams& p, Case& cs) {
  const int nxc = p.lattice[0], nyc = p.lattice[1], nzc = p.lattice[2];
  const int ncube = nxc * nyc * nzc + 1;
  cs.ncube = ncube;
  cs.cube_size.assign(ncube, 0.0);
  cs.Xcube.assign(ncube, 0.0);
  cs.Ycube.assign(ncube, 0.0);
  cs.Zcube.assign(ncube, 0.0);
  cs.adj_number.assign((size_t)ncube * 35, 0);
  auto cid = [&](int ix, int iy, int iz) { return 1 + ix + nxc * (iy + nyc * iz); };
  for (int iz = 0; iz < nzc; ++iz)
    for (int iy = 0; iy < nyc; ++iy)
      for (int ix = 0; ix < nxc; ++ix) {
        int c = cid(ix, iy, iz);
        cs.cube_size[c] = p.cube_size;
        cs.Xcube[c] = p.origin[0] + ix * p.cube_size;
        cs.Ycube[c] = p.origin[1] + iy * p.cube_size;
        cs.Zcube[c] = p.origin[2] + iz * p.cube_size;
        int* a = &cs.adj_number[(size_t)c * 35];
        a[1 * 7 + 1] = ix > 0 ? cid(ix - 1, iy, iz) : 0;
        a[1 * 7 + 2] = ix < nxc - 1 ? cid(ix + 1, iy, iz) : 0;
        a[1 * 7 + 3] = iy > 0 ? cid(ix, iy - 1, iz) : 0;
        a[1 * 7 + 4] = iy < nyc - 1 ? cid(ix, iy + 1, iz) : 0;
        a[1 * 7 + 5] = iz > 0 ? cid(ix, iy, iz - 1) : 0;
        a[1 * 7 + 6] = iz < nzc - 1 ? cid(ix, iy, iz + 1) : 0;
      }
}

// Two-level grid: lattice cubes of edge cube_size, those inside the
// refinement box split into 8 cubes of half the edge (x fastest within the
// parent).  Cubes are numbered in lattice order, a split cube's 8 children in
// place of it.  Adjacency follows the slot convention of BCM_Interface.cpp:
// a face with one neighbour (same size or the bigger cube) uses slot 1; a
// face with four smaller neighbours lists them in slots 1..4 with slot 2 on
// the upper half of the first tangential axis and slot 3 on the upper half of
// the second (tangential axes in x, y, z order), e.g. :1825-1883 for +x.
// With a second refinement box (on the doubled lattice) the split cubes'
// children inside it are split again, recursively in place: a three-level
// grid.  Positions count in units of the smallest cube.
int two_level_lattice(const CaseParams& p, Case& cs) {
  bool second = true;
  for (int d = 0; d < 3; ++d)
    if (p.refine2_hi[d] <= p.refine2_lo[d]) second = false;
  const int S = second ? 4 : 2;                       // lattice cube edge in smallest cubes
  const int L[3] = {p.lattice[0], p.lattice[1], p.lattice[2]};
  const int F[3] = {S * L[0], S * L[1], S * L[2]};   // finest positions per axis
  const double h = p.cube_size / S;
  std::vector<int> owner((size_t)F[0] * F[1] * F[2], 0);
  auto own = [&](int x, int y, int z) -> int& { return owner[((size_t)z * F[1] + y) * F[0] + x]; };
  struct Cube { int x0[3]; int s; };
  std::vector<Cube> cubes(1);
  auto in_box = [](const int* ii, const int* lo, const int* hi) {
    bool in = true;
    for (int d = 0; d < 3; ++d) in = in && ii[d] >= lo[d] && ii[d] < hi[d];
    return in;
  };
  for (int iz = 0; iz < L[2]; ++iz)
    for (int iy = 0; iy < L[1]; ++iy)
      for (int ix = 0; ix < L[0]; ++ix) {
        const int ii[3] = {ix, iy, iz};
        if (!in_box(ii, p.refine_lo, p.refine_hi)) {
          cubes.push_back({{S * ix, S * iy, S * iz}, S});
          continue;
        }
        for (int dz = 0; dz < 2; ++dz)
          for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
              const int jj[3] = {2 * ix + dx, 2 * iy + dy, 2 * iz + dz};   // doubled lattice
              const int s = S / 2;
              if (!second || !in_box(jj, p.refine2_lo, p.refine2_hi)) {
                cubes.push_back({{s * jj[0], s * jj[1], s * jj[2]}, s});
                continue;
              }
              for (int ez = 0; ez < 2; ++ez)
                for (int ey = 0; ey < 2; ++ey)
                  for (int ex = 0; ex < 2; ++ex)
                    cubes.push_back({{s * jj[0] + ex, s * jj[1] + ey, s * jj[2] + ez}, 1});
            }
      }
  const int ncube = (int)cubes.size();
  for (int c = 1; c < ncube; ++c) {
    const Cube& q = cubes[c];
    for (int a = 0; a < q.s; ++a)
      for (int b = 0; b < q.s; ++b)
        for (int e = 0; e < q.s; ++e) own(q.x0[0] + e, q.x0[1] + b, q.x0[2] + a) = c;
  }
  cs.ncube = ncube;
  cs.cube_size.assign(ncube, 0.0);
  cs.Xcube.assign(ncube, 0.0);
  cs.Ycube.assign(ncube, 0.0);
  cs.Zcube.assign(ncube, 0.0);
  cs.adj_number.assign((size_t)ncube * 35, 0);
  for (int c = 1; c < ncube; ++c) {
    const Cube& q = cubes[c];
    cs.cube_size[c] = q.s * h;
    cs.Xcube[c] = p.origin[0] + q.x0[0] * h;
    cs.Ycube[c] = p.origin[1] + q.x0[1] * h;
    cs.Zcube[c] = p.origin[2] + q.x0[2] * h;
    for (int dir = 1; dir <= 6; ++dir) {
      const int A = (dir - 1) / 2;
      const bool plus = dir % 2 == 0;
      const int t0 = A == 0 ? 1 : 0, t1 = A == 2 ? 1 : 2;
      const int pa = plus ? q.x0[A] + q.s : q.x0[A] - 1;
      if (pa < 0 || pa >= F[A]) continue;
      // The neighbour at each quadrant of the face (u0 on the first
      // tangential axis fastest): one cube of the same or twice the size
      // covers them all, or four cubes of half the size cover one each.
      int ids[4], n = 0;
      const int hs = q.s > 1 ? q.s / 2 : 1;
      for (int u1 = 0; u1 < q.s; u1 += hs)
        for (int u0 = 0; u0 < q.s; u0 += hs) {
          int pos[3];
          pos[A] = pa;
          pos[t0] = q.x0[t0] + u0;
          pos[t1] = q.x0[t1] + u1;
          ids[n++] = own(pos[0], pos[1], pos[2]);
        }
      int* a = &cs.adj_number[(size_t)c * 35];
      const int s0 = cubes[ids[0]].s;
      if (n == 1 || (ids[0] == ids[1] && ids[0] == ids[2] && ids[0] == ids[3])) {
        if (s0 != q.s && s0 != 2 * q.s) return 1;   // more than one level apart
        a[1 * 7 + dir] = ids[0];
      } else {
        for (int m = 0; m < 4; ++m) {
          if (cubes[ids[m]].s != hs) return 1;   // neither one neighbour nor four half-size ones
          a[(1 + m) * 7 + dir] = ids[m];
        }
      }
    }
  }
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
int build_case(const CaseParams& p, Case& cs, std::string& err, IbDevice* dev) {
  cs = Case();
  cs.p = p;
  if (p.lattice[0] < 1 || p.lattice[1] < 1 || p.lattice[2] < 1) {
    err = "lattice dimensions must be positive";
    return 1;
  }
  if (!(p.cube_size > 0) || !(p.sphere_diameter >= 0) || p.ico_level < 0 || p.ico_level > 7 ||
      !(p.mach > 0) || !(p.reynolds > 0)) {
    err = "invalid case parameters";
    return 1;
  }
  compute_physics(p, cs.phys);
  bool refined = true;
  for (int d = 0; d < 3; ++d) {
    if (p.refine_hi[d] <= p.refine_lo[d]) refined = false;
    if (p.refine_lo[d] < 0 || p.refine_hi[d] > p.lattice[d]) {
      err = "refinement box outside the lattice";
      return 1;
    }
  }
  bool refined2 = true;
  for (int d = 0; d < 3; ++d) {
    if (p.refine2_hi[d] <= p.refine2_lo[d]) refined2 = false;
    if (p.refine2_lo[d] < 0 || p.refine2_hi[d] > 2 * p.lattice[d]) {
      err = "second refinement box outside the doubled lattice";
      return 1;
    }
  }
  if (refined2 && !refined) {
    err = "a second refinement box needs a first one";
    return 1;
  }
  if (refined) {
    if (two_level_lattice(p, cs)) {
      err = "refined lattice could not be built (neighbouring cubes more than one level apart)";
      return 1;
    }
  } else {
    uniform_lattice(p, cs);
  }
  const int ncube = cs.ncube;
  grid_geometry(cs);

  // Grid flags: all cells fluid, interior cells whose centre lies inside the
  // sphere solid.  This is the content of the grid file the reference reads
  // (BCM_Grid.cpp:172-278): ghost layers stay fluid.
  cs.fws.assign((size_t)ncube * CELLS_PADDED, IFLUID);
  for (int n = 0; n < CELLS_PADDED; ++n) cs.fws[n] = 0;  // slot 0 is unused
  const double r = 0.5 * p.sphere_diameter;
  std::vector<uint8_t> is_wall(ncube, 0);
  parallel_for(ncube - 1, [&](int c0) {
    const int c = c0 + 1;
    bool wall = false;
    for (int i = ILO; i <= IHI; ++i)
      for (int j = ILO; j <= IHI; ++j)
        for (int k = ILO; k <= IHI; ++k) {
          double ddx = cs.Xcnt[c * NP + i] - p.sphere_center[0];
          double ddy = cs.Ycnt[c * NP + j] - p.sphere_center[1];
          double ddz = cs.Zcnt[c * NP + k] - p.sphere_center[2];
          if (ddx * ddx + ddy * ddy + ddz * ddz < r * r) {
            cs.fws[flag_index(c, i, j, k)] = ISOLID;
            wall = true;
          }
        }
    is_wall[c] = wall;
  });
  cs.wallcube.assign(1, 0);
  for (int c = 1; c < ncube; ++c)
    if (is_wall[c]) cs.wallcube.push_back(c);
  cs.n_wallcube = (int)cs.wallcube.size();
  cs.fws_grid = cs.fws;
  make_icosphere(p.ico_level, p.sphere_center, r, cs.stl);
  return finish_setup(cs, err, dev);
}

// Grid geometry derived from cube sizes and corners: cell centres
// (BCM_Grid.cpp:123-150) and cube size levels (:155-168).
void grid_geometry(Case& cs) {
  const int ncube = cs.ncube;
  cs.Xcnt.assign((size_t)ncube * NP, 0.0);
  cs.Ycnt.assign((size_t)ncube * NP, 0.0);
  cs.Zcnt.assign((size_t)ncube * NP, 0.0);
  for (int c = 1; c < ncube; ++c) {
    double dx = cs.cube_size[c] / NC;
    for (int i = 0; i < NP; ++i) {
      cs.Xcnt[c * NP + i] = cs.Xcube[c] + (i - 1.5) * dx;
      cs.Ycnt[c * NP + i] = cs.Ycube[c] + (i - 1.5) * dx;
      cs.Zcnt[c * NP + i] = cs.Zcube[c] + (i - 1.5) * dx;
    }
  }
  cs.csl.assign(ncube, 0);
  double size_min = BIGNUM;
  for (int c = 1; c < ncube; ++c) size_min = dmin(size_min, cs.cube_size[c]);
  for (int c = 1; c < ncube; ++c)
    cs.csl[c] = static_cast<int>(0.5 + log(cs.cube_size[c] / size_min) / log(2.0));
}

// Everything after the grid (cube geometry, adjacency, grid flags, wall
// cubes) and the surface triangles are known: main.cpp:108-1139.
int finish_setup(Case& cs, std::string& err, IbDevice* dev) {
  const CaseParams& p = cs.p;
  const int ncube = cs.ncube;

  // Domain-boundary cube lists, main.cpp:108-189.
  double Xmax = -BIGNUM, Xmin = BIGNUM, Ymax = -BIGNUM, Ymin = BIGNUM, Zmax = -BIGNUM, Zmin = BIGNUM;
  for (int c = 1; c < ncube; ++c) {
    Xmin = dmin(Xmin, cs.Xcube[c]);
    Xmax = dmax(Xmax, cs.Xcube[c] + cs.cube_size[c]);
    Ymin = dmin(Ymin, cs.Ycube[c]);
    Ymax = dmax(Ymax, cs.Ycube[c] + cs.cube_size[c]);
    Zmin = dmin(Zmin, cs.Zcube[c]);
    Zmax = dmax(Zmax, cs.Zcube[c] + cs.cube_size[c]);
  }
  for (auto& l : cs.bc_lists) l.assign(1, 0);
  for (int c = 1; c < ncube; ++c) {
    if (fabs(cs.Xcube[c] + cs.cube_size[c] - Xmax) < SMALLNUM) cs.bc_lists[1].push_back(c);
    if (fabs(cs.Ycube[c] + cs.cube_size[c] - Ymax) < SMALLNUM) cs.bc_lists[3].push_back(c);
    if (fabs(cs.Zcube[c] + cs.cube_size[c] - Zmax) < SMALLNUM) cs.bc_lists[5].push_back(c);
    if (fabs(cs.Xcube[c] - Xmin) < SMALLNUM) cs.bc_lists[0].push_back(c);
    if (fabs(cs.Ycube[c] - Ymin) < SMALLNUM) cs.bc_lists[2].push_back(c);
    if (fabs(cs.Zcube[c] - Zmin) < SMALLNUM) cs.bc_lists[4].push_back(c);
  }

  // Interface lists.
  cs.nadj.assign(L_COUNT, 0);
  cs.adj.assign(L_COUNT, std::vector<int>(ncube, 0));
  {
    int* lp[L_COUNT];
    for (int l = 0; l < L_COUNT; ++l) lp[l] = cs.adj[l].data();
    interface_tables(ncube, cs.csl.data(), cs.adj_number.data(), cs.nadj.data(), lp);
  }
  // Size changes of more than one level are not two-level faces.
  for (int c = 1; c < ncube; ++c)
    for (int dir = 1; dir <= 6; ++dir)
      for (int sl = 1; sl <= 4; ++sl) {
        const int n = cs.adj_number[adj_index(c, sl, dir)];
        if (n > 0 && abs(cs.csl[n] - cs.csl[c]) > 1) {
          err = "neighbouring cubes differ by more than one size level";
          return 2;
        }
      }

  // First flag halo exchange and hole treatment (main.cpp:634-647).
  fws_interface(cs);

  // Initial state, BCM_Initial_condition.cpp:46-86, then seeded density noise
  // on fluid cells, then U1_ = U1q = U1 (main.cpp:661-685).
  const Physics& ph = cs.phys;
  const size_t nstate = (size_t)ncube * CELLS_PADDED * NV;
  cs.U1.resize(nstate);
  cs.U1_.resize(nstate);
  cs.U1q.resize(nstate);
  for (StateVec* v : {&cs.U1, &cs.U1_, &cs.U1q})
    std::fill(v->begin(), v->begin() + (size_t)CELLS_PADDED * NV, 0.0);  // cube 0
  parallel_for(ncube - 1, [&](int c0) {
    const int c = c0 + 1;
    for (int i = 0; i < NP; ++i)
      for (int j = 0; j < NP; ++j)
        for (int k = 0; k < NP; ++k) {
          double u[NV];
          int f = cs.fws[flag_index(c, i, j, k)];
          if (f > ISOLID) {
            double rho = ph.rho0, U = ph.U0, V = ph.V0, W = ph.W0;
            double VV = U * U + V * V + W * W;
            double P = ph.P0;
            u[0] = rho; u[1] = rho * U; u[2] = rho * V; u[3] = rho * W;
            u[4] = P / (KGAS - 1) + 0.5 * rho * VV;
          } else {
            u[0] = ph.rho0; u[1] = 0; u[2] = 0; u[3] = 0;
            u[4] = ph.P0 / (KGAS - 1);
          }
          if (f == IFLUID && p.noise_amp != 0.0) {
            double fac = 1.0 + p.noise_amp * noise_unit(p.seed, c, i, j, k);
            u[0] = u[0] * fac;
          }
          const size_t o = aos_index(c, i, j, k, 0);
          for (int v = 0; v < NV; ++v) cs.U1[o + v] = cs.U1_[o + v] = cs.U1q[o + v] = u[v];
        }
  });

  // Ghost detection on finest cubes, main.cpp:694-803.
  {
    std::vector<int> tmp = cs.fws;
    parallel_for(ncube - 1, [&](int c0) {
      const int c = c0 + 1;
      if (cs.csl[c] != 0) return;
      auto F = [&](int i, int j, int k) { return cs.fws[flag_index(c, i, j, k)]; };
      for (int i = 1; i <= IHI + 1; ++i)
        for (int j = 1; j <= IHI + 1; ++j)
          for (int k = 1; k <= IHI + 1; ++k)
            if (F(i, j, k) == ISOLID &&
                (F(i + 1, j, k) == IFLUID || F(i - 1, j, k) == IFLUID || F(i, j - 1, k) == IFLUID ||
                 F(i, j + 1, k) == IFLUID || F(i, j, k - 1) == IFLUID || F(i, j, k + 1) == IFLUID))
              tmp[flag_index(c, i, j, k)] = -2;
    });
    parallel_for(ncube - 1, [&](int c0) {
      const int c = c0 + 1;
      if (cs.csl[c] != 0) return;
      for (int n = 0; n < CELLS_PADDED; ++n)
        if (tmp[(size_t)c * CELLS_PADDED + n] == -2) cs.fws[(size_t)c * CELLS_PADDED + n] = IGHOST;
    });
  }

  // IB tables (main.cpp:811).
  if (immersed_boundary(cs, dev, err)) return 1;

  int* lp[L_COUNT];
  for (int l = 0; l < L_COUNT; ++l) lp[l] = cs.adj[l].data();
  host_interface_copy(ncube, lp, cs.nadj.data(), cs.adj_number.data(), cs.U1_.data());
  for (const IBTable* tb : {&cs.ib_minus, &cs.ib_plus})
    for (int n = 1; n <= tb->n; ++n) {
      const int* g = &tb->gcindex[(n - 1) * 4 + 1];
      cs.fws[flag_index(g[0], g[1], g[2], g[3])] = IGHOST;
    }
  host_interface_copy(ncube, lp, cs.nadj.data(), cs.adj_number.data(), cs.U1_.data());
  fws_interface(cs);
  // Listed ghost cells start at rest (main.cpp:1087-1139).
  for (const IBTable* tb : {&cs.ib_minus, &cs.ib_plus})
    for (int n = 1; n <= tb->n; ++n) {
      const int* g = &tb->gcindex[(n - 1) * 4 + 1];
      for (double* U : {cs.U1.data(), cs.U1_.data()}) {
        double* u = &U[aos_index(g[0], g[1], g[2], g[3], 0)];
        u[0] = ph.rho0; u[1] = 0; u[2] = 0; u[3] = 0; u[4] = ph.P0 / (KGAS - 1);
      }
    }
  return 0;
}

// ---------------------------------------------------------------------------
// Grid file in the layout BCM_Grid.cpp:63-261 and main.cpp:84-97 read.
int write_bcmgrid(const Case& cs, const std::string& path) {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return 1;
  const int n = cs.ncube - 1;
  fprintf(f, "BCM grid generated for a %dx%dx%d cube lattice\n", cs.p.lattice[0],
          cs.p.lattice[1], cs.p.lattice[2]);
  fprintf(f, "#_of_cubes in the computational region >>\n%d\n", n);
  fprintf(f, "cube_size\tXcube\tYcube\tZcube\n");
  for (int c = 1; c < cs.ncube; ++c)
    fprintf(f, "%.17g\t%.17g\t%.17g\t%.17g\n", cs.cube_size[c], cs.Xcube[c], cs.Ycube[c],
            cs.Zcube[c]);
  fprintf(f, "adjacent cubes (4 slots) for -X +X -Y +Y -Z +Z\n");
  for (int c = 1; c < cs.ncube; ++c)
    for (int d = 1; d <= 6; ++d)
      fprintf(f, "%d\t%d\t%d\t%d\n", cs.adj_number[adj_index(c, 1, d)],
              cs.adj_number[adj_index(c, 2, d)], cs.adj_number[adj_index(c, 3, d)],
              cs.adj_number[adj_index(c, 4, d)]);
  fprintf(f, "#_of_cells in a cube >>\n%d\n", NC * NC * NC);
  fprintf(f, "#_of_wall_cubes >>\n%d\n", cs.n_wallcube - 1);
  fprintf(f, "wall cubes\n");
  for (int w = 1; w < cs.n_wallcube; ++w) fprintf(f, "%d\n", cs.wallcube[w]);
  fprintf(f, "run-length encoded flags of the wall cubes\n");
  std::vector<int> seq(NC * NC * NC + 1);
  for (int w = 1; w < cs.n_wallcube; ++w) {
    int c = cs.wallcube[w];
    for (int k = 1; k <= NC; ++k)
      for (int j = 1; j <= NC; ++j)
        for (int i = 1; i <= NC; ++i) {
          int pp = j % 2, qq = k % 2;
          int i1 = NC * (j - 1) + pp * i + (1 - pp) * (NC + 1 - i);
          int i2 = NC * (NC - j) + pp * (NC + 1 - i) + (1 - pp) * i;
          int ii = qq * i1 + (1 - qq) * i2 + NC * NC * (k - 1);
          seq[ii] = cs.fws_grid[flag_index(c, i + NB - 1, j + NB - 1, k + NB - 1)] == IFLUID ? 1 : 0;
        }
    std::vector<int> ch;
    for (int ii = 2; ii <= NC * NC * NC; ++ii)
      if (seq[ii] != seq[ii - 1]) ch.push_back(ii);
    fprintf(f, "%d\n%d\n", (int)ch.size(), seq[1]);
    for (int x : ch) fprintf(f, "%d\n", x);
  }
  fclose(f);
  return 0;
}

// ASCII STL as read by BCM_Immersed_boundary.cpp:198-275.
int write_stl(const Case& cs, const std::string& path) {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return 1;
  fprintf(f, "solid icosphere\n");
  for (const Triangle& T : cs.stl) {
    fprintf(f, "facet normal %.17g %.17g %.17g\n", T.n[0], T.n[1], T.n[2]);
    fprintf(f, "outer loop\n");
    for (int m = 0; m < 3; ++m) fprintf(f, "vertex %.17g %.17g %.17g\n", T.v[m][0], T.v[m][1], T.v[m][2]);
    fprintf(f, "endloop\n");
    fprintf(f, "endfacet\n");
  }
  fprintf(f, "endsolid icosphere\n");
  fclose(f);
  return 0;
}

}  // namespace bcm
