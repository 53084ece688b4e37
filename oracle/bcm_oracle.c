/*
 * bcm_oracle.c -- CPU restatement, in FP64 C, of the per-iteration hot path of
 * the BCM_CGLI reference solver.
 *
 * TEST INFRASTRUCTURE.  This file is the parity checker for the HIP kernels:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it, and never as the thing measured or shipped.  The product path does not
 * link or call it.
 *
 * Every function follows the reference routine it names, in the same
 * operation order (build with -ffp-contract=off so no multiply-add is fused),
 * so its results are bitwise those of the reference compiled the same way.
 * Arrays use the reference layouts of Array.h:11-36: state [cube][20][20][20][5]
 * (k fastest, variable innermost), flags [cube][20][20][20], cubes 1-based.
 *
 * Pinning (see tests/test_oracle_golden.py): every function here is checked
 * bitwise against outputs of the reference's own sources compiled in the build
 * container (fixtures in tests/golden/), and by closed forms in
 * tests/test_oracle_properties.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NP 20
#define NV 5
#define ILO 2
#define IHI 17
#define K_ 1.4
#define R_ 287.0
#define CV_ 717.5
#define PR_ 0.72
#define EP_ 1.

typedef struct {
  double rho0, U0, Uin, V0, W0, P0, T0, mu_L;
} orc_phys;

static inline int64_t ix(int c, int i, int j, int k) {
  return (((int64_t)c * NP + i) * NP + j) * NP + k;
}
#define S(arr, c, i, j, k, v) (arr)[ix(c, i, j, k) * NV + (v)]
#define F(arr, c, i, j, k) (arr)[ix(c, i, j, k)]
#define MAXD(a, b) (((a) > (b)) ? (a) : (b))

/* ------------------------------------------------------------------------
 * Domain boundary conditions.
 * X: BCM_X_boundary_condition.cpp:43-118 (inflow at x-min from cell i=2's
 *    density and pressure with velocity (Uin,V0,W0); copy of i=17 at x-max).
 * Y: BCM_Y_boundary_condition.cpp:36-80 (copy of j=2 / j=17, i over 0..19).
 * Z: BCM_Z_boundary_condition.cpp:42-92 (symmetry at z-min, w negated;
 *    copy of k=17 at z-max; i, j over 0..19).
 */
void orc_bc_x(int nl, int nu, const int* l, const int* u, double* U, const orc_phys* ph) {
  for (int n = 1; n <= nl; ++n) {
    int c = l[n];
    for (int j = ILO; j <= IHI; ++j)
      for (int k = ILO; k <= IHI; ++k) {
        double rho = S(U, c, 2, j, k, 0);
        double Uu = S(U, c, 2, j, k, 1) / rho;
        double V = S(U, c, 2, j, k, 2) / rho;
        double W = S(U, c, 2, j, k, 3) / rho;
        double VV = Uu * Uu + V * V + W * W;
        double P = (S(U, c, 2, j, k, 4) - 0.5 * rho * VV) * (K_ - 1);
        Uu = ph->Uin;
        V = ph->V0;
        W = ph->W0;
        VV = Uu * Uu + V * V + W * W;
        for (int i = 1; i >= 0; --i) {
          S(U, c, i, j, k, 0) = rho;
          S(U, c, i, j, k, 1) = rho * Uu;
          S(U, c, i, j, k, 2) = rho * V;
          S(U, c, i, j, k, 3) = rho * W;
          S(U, c, i, j, k, 4) = P / (K_ - 1) + 0.5 * rho * VV;
        }
      }
  }
  for (int n = 1; n <= nu; ++n) {
    int c = u[n];
    for (int j = ILO; j <= IHI; ++j)
      for (int k = ILO; k <= IHI; ++k)
        for (int v = 0; v < NV; ++v) {
          S(U, c, 18, j, k, v) = S(U, c, 17, j, k, v);
          S(U, c, 19, j, k, v) = S(U, c, 17, j, k, v);
        }
  }
}

void orc_bc_y(int nl, int nu, const int* l, const int* u, double* U) {
  for (int n = 1; n <= nl; ++n) {
    int c = l[n];
    for (int i = 0; i < NP; ++i)
      for (int k = ILO; k <= IHI; ++k)
        for (int v = 0; v < NV; ++v) {
          S(U, c, i, 1, k, v) = S(U, c, i, 2, k, v);
          S(U, c, i, 0, k, v) = S(U, c, i, 2, k, v);
        }
  }
  for (int n = 1; n <= nu; ++n) {
    int c = u[n];
    for (int i = 0; i < NP; ++i)
      for (int k = ILO; k <= IHI; ++k)
        for (int v = 0; v < NV; ++v) {
          S(U, c, i, 18, k, v) = S(U, c, i, 17, k, v);
          S(U, c, i, 19, k, v) = S(U, c, i, 17, k, v);
        }
  }
}

void orc_bc_z(int nl, int nu, const int* l, const int* u, double* U) {
  for (int n = 1; n <= nl; ++n) {
    int c = l[n];
    for (int i = 0; i < NP; ++i)
      for (int j = 0; j < NP; ++j)
        for (int v = 0; v < NV; ++v) {
          double s = S(U, c, i, j, 2, v);
          S(U, c, i, j, 1, v) = (v == 3) ? -s : s;
          S(U, c, i, j, 0, v) = (v == 3) ? -s : s;
        }
  }
  for (int n = 1; n <= nu; ++n) {
    int c = u[n];
    for (int i = 0; i < NP; ++i)
      for (int j = 0; j < NP; ++j)
        for (int v = 0; v < NV; ++v) {
          S(U, c, i, j, 18, v) = S(U, c, i, j, 17, v);
          S(U, c, i, j, 19, v) = S(U, c, i, j, 17, v);
        }
  }
}

/* ------------------------------------------------------------------------
 * Immersed-boundary ghost fill, BCM_Ghostcell_minus.cpp:72-489 (the _plus file
 * is the same routine).  Entries are processed in list order and write U in
 * place, so a later entry sees earlier writes among its interpolation corners.
 */
void orc_ghostcell(int nbc, const double* weight, const int* gcindex, const int* ipsur,
                   double* U, const orc_phys* ph) {
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  const double rho0 = ph->rho0, P0 = ph->P0;
  for (int n = 1; n <= nbc; ++n) {
    const int* q = &ipsur[(n - 1) * 4 + 1];
    const int* g = &gcindex[(n - 1) * 4 + 1];
    const double* w = &weight[(n - 1) * 8 + 1];
    double rc[8], uc[8], vc[8], wc[8], pc[8];
    for (int m = 0; m < 8; ++m) {
      const double* s = &S(U, q[0], q[1] + off[m][0], q[2] + off[m][1], q[3] + off[m][2], 0);
      rc[m] = s[0];
      uc[m] = s[1] / rc[m];
      vc[m] = s[2] / rc[m];
      wc[m] = s[3] / rc[m];
      double VV = uc[m] * uc[m] + vc[m] * vc[m] + wc[m] * wc[m];
      pc[m] = (s[4] - 0.5 * rc[m] * VV) * (K_ - 1);
    }
    double rho = 0, P = 0, Uu = 0, V = 0, W = 0;
    /* rho = w1*(rho000-rho0)+...+w8*(rho111-rho0)+rho0 with left-to-right sums */
    rho = w[0] * (rc[0] - rho0);
    P = w[0] * (pc[0] - P0);
    Uu = w[0] * uc[0];
    V = w[0] * vc[0];
    W = w[0] * wc[0];
    for (int m = 1; m < 8; ++m) {
      rho = rho + w[m] * (rc[m] - rho0);
      P = P + w[m] * (pc[m] - P0);
      Uu = Uu + w[m] * uc[m];
      V = V + w[m] * vc[m];
      W = W + w[m] * wc[m];
    }
    rho = rho + rho0;
    P = P + P0;
    double VV = Uu * Uu + V * V + W * W;
    double* t = &S(U, g[0], g[1], g[2], g[3], 0);
    t[0] = rho;
    t[1] = 0.5 * rho * Uu;
    t[2] = 0.5 * rho * V;
    t[3] = 0.5 * rho * W;
    t[4] = P / (K_ - 1) + 0.5 * rho * 0.25 * VV;
    for (int s = 0; s < 8; ++s) {
      if (!(g[1] == q[1] + off[s][0] && g[2] == q[2] + off[s][1] && g[3] == q[3] + off[s][2]))
        continue;
      /* The ghost cell is corner s: interpolate from the other seven. */
      if (w[s] > 0.9999) {
        rho = rc[s];
        P = pc[s];
      } else {
        double sr = 0, sp = 0;
        int first = 1;
        for (int m = 0; m < 8; ++m) {
          if (m == s) continue;
          if (first) {
            sr = w[m] * (rc[m] - rho0);
            sp = w[m] * (pc[m] - P0);
            first = 0;
          } else {
            sr = sr + w[m] * (rc[m] - rho0);
            sp = sp + w[m] * (pc[m] - P0);
          }
        }
        rho = sr / (1 - w[s]) + rho0;
        P = sp / (1 - w[s]) + P0;
      }
      double su = 0, sv = 0, sw = 0;
      int first = 1;
      for (int m = 0; m < 8; ++m) {
        if (m == s) continue;
        if (first) {
          su = w[m] * uc[m];
          sv = w[m] * vc[m];
          sw = w[m] * wc[m];
          first = 0;
        } else {
          su = su + w[m] * uc[m];
          sv = sv + w[m] * vc[m];
          sw = sw + w[m] * wc[m];
        }
      }
      Uu = su / (2 - w[s]);
      V = sv / (2 - w[s]);
      W = sw / (2 - w[s]);
      VV = Uu * Uu + V * V + W * W;
      t[0] = rho;
      t[1] = rho * Uu;
      t[2] = rho * V;
      t[3] = rho * W;
      t[4] = P / (K_ - 1) + 0.5 * rho * VV;
    }
  }
}

/* ------------------------------------------------------------------------
 * On-rank part of BCM_Interface:
 * equal-size face copies X (Interface.cpp:1745-1782), Y (:2604-2641),
 * Z (:3474-3511); two-level faces X (:1790-2600), Y (:2651-3470),
 * Z (:3518-4360); then the per-cube edge-ghost averaging (:5640-5732).
 * eq[d] lists cubes whose +d neighbour (adj_number[c][1][2d+2]) has the same
 * size; bs[2d] (bs_plus) and bs[2d+1] (bs_minus) list big cubes with four
 * small neighbours on their + and - side of axis d.
 */
#define ADJ(c, s, d) adj_number[((c) * 5 + (s)) * 7 + (d)]

/* Position of a point given its axis-d index a and tangential indices (p, r)
 * in x, y, z order of the other two axes. */
static inline void place(int d, int a, int p, int r, int* x) {
  x[d] = a;
  x[d == 0 ? 1 : 0] = p;
  x[d == 2 ? 1 : 2] = r;
}

/* Big cube -> small cube: the small cube's two ghost layers copy the big
 * cube's face cell over each 2 x 2 x 1 group (:1795-1898 and :1908-2021 for x). */
static void inject(int d, int big, int small, int slot, int small_upper, double* U) {
  int op = (slot == 2 || slot == 4) ? 8 : 0, orr = (slot == 3 || slot == 4) ? 8 : 0;
  for (int l = 0; l < 2; ++l)
    for (int p = ILO; p <= IHI; ++p)
      for (int r = ILO; r <= IHI; ++r) {
        int dst[3], src[3];
        place(d, small_upper ? IHI + l + 1 : l, p, r, dst);
        place(d, small_upper ? 2 + l / 2 : IHI + l / 2, p / 2 + 1 + op, r / 2 + 1 + orr, src);
        for (int v = 0; v < NV; ++v)
          S(U, small, dst[0], dst[1], dst[2], v) = S(U, big, src[0], src[1], src[2], v);
      }
}

/* Small cube -> big cube: the big cube's ghost cells of one face quadrant
 * take 0.125 times the sum of the 8 small cells beneath, in the order
 * (0,0,0) (1,0,0) (0,1,0) (0,0,1) (1,1,0) (1,0,1) (0,1,1) (1,1,1) of (i,j,k)
 * offsets (:2026-2310 and :2314-2600 for x). */
static void restrict8(int d, int big, int small, int slot, int big_upper, double* U) {
  int op = (slot == 2 || slot == 4) ? 8 : 0, orr = (slot == 3 || slot == 4) ? 8 : 0;
  for (int l = 0; l < 2; ++l)
    for (int p = ILO; p <= 9; ++p)
      for (int r = ILO; r <= 9; ++r) {
        int dst[3], b[3];
        place(d, big_upper ? 18 + l : l, p + op, r + orr, dst);
        place(d, big_upper ? 2 * (l + 1) : 17 - 4 + 2 * (l + 1) - 1, (p - 1) * 2, (r - 1) * 2, b);
        for (int v = 0; v < NV; ++v) {
          double s = S(U, small, b[0], b[1], b[2], v);
          s += S(U, small, b[0] + 1, b[1], b[2], v);
          s += S(U, small, b[0], b[1] + 1, b[2], v);
          s += S(U, small, b[0], b[1], b[2] + 1, v);
          s += S(U, small, b[0] + 1, b[1] + 1, b[2], v);
          s += S(U, small, b[0] + 1, b[1], b[2] + 1, v);
          s += S(U, small, b[0], b[1] + 1, b[2] + 1, v);
          s += S(U, small, b[0] + 1, b[1] + 1, b[2] + 1, v);
          S(U, big, dst[0], dst[1], dst[2], v) = 0.125 * s;
        }
      }
}

void orc_interface(int ncube, const int* neq, const int* const* eq, const int* nbs,
                   const int* const* bs, const int* adj_number, double* U) {
  for (int d = 0; d < 3; ++d) {
    for (int a = 1; a <= neq[d]; ++a) {
      int c0 = eq[d][a];
      int c2 = adj_number[(c0 * 5 + 1) * 7 + 2 * d + 2];
      for (int l = 0; l < 2; ++l)
        for (int p = ILO; p <= IHI; ++p)
          for (int r = ILO; r <= IHI; ++r) {
            int a0[3], a2[3], b0[3], b2[3];
            /* (p, r) are the two tangential indices in increasing axis order */
            int t0 = d == 0 ? 1 : 0, t1 = d == 2 ? 1 : 2;
            a0[d] = 16 + l; a2[d] = l; b0[d] = 18 + l; b2[d] = 2 + l;
            a0[t0] = a2[t0] = b0[t0] = b2[t0] = p;
            a0[t1] = a2[t1] = b0[t1] = b2[t1] = r;
            for (int v = 0; v < NV; ++v) {
              S(U, c2, a2[0], a2[1], a2[2], v) = S(U, c0, a0[0], a0[1], a0[2], v);
              S(U, c0, b0[0], b0[1], b0[2], v) = S(U, c2, b2[0], b2[1], b2[2], v);
            }
          }
    }
    for (int a = 1; a <= nbs[2 * d]; ++a)
      for (int s = 1; s <= 4; ++s) {
        int big = bs[2 * d][a], small = ADJ(big, s, 2 * d + 2);
        if (small > 0) inject(d, big, small, s, 0, U);
      }
    for (int a = 1; a <= nbs[2 * d + 1]; ++a)
      for (int s = 1; s <= 4; ++s) {
        int big = bs[2 * d + 1][a], small = ADJ(big, s, 2 * d + 1);
        if (small > 0) inject(d, big, small, s, 1, U);
      }
    for (int a = 1; a <= nbs[2 * d]; ++a)
      for (int s = 1; s <= 4; ++s) {
        int big = bs[2 * d][a], small = ADJ(big, s, 2 * d + 2);
        if (small > 0) restrict8(d, big, small, s, 1, U);
      }
    for (int a = 1; a <= nbs[2 * d + 1]; ++a)
      for (int s = 1; s <= 4; ++s) {
        int big = bs[2 * d + 1][a], small = ADJ(big, s, 2 * d + 1);
        if (small > 0) restrict8(d, big, small, s, 0, U);
      }
  }
  for (int c = 1; c < ncube; ++c) {
    for (int i = ILO; i <= IHI; ++i)
      for (int v = 0; v < NV; ++v) {
        S(U, c, i, 1, 1, v) = 0.5 * (S(U, c, i, 1, 2, v) + S(U, c, i, 2, 1, v));
        S(U, c, i, 18, 1, v) = 0.5 * (S(U, c, i, 18, 2, v) + S(U, c, i, 17, 1, v));
        S(U, c, i, 1, 18, v) = 0.5 * (S(U, c, i, 1, 17, v) + S(U, c, i, 2, 18, v));
        S(U, c, i, 18, 18, v) = 0.5 * (S(U, c, i, 18, 17, v) + S(U, c, i, 17, 18, v));
      }
    for (int j = 1; j <= 18; ++j)
      for (int v = 0; v < NV; ++v) {
        S(U, c, 1, j, 1, v) = 0.5 * (S(U, c, 1, j, 2, v) + S(U, c, 2, j, 1, v));
        S(U, c, 1, j, 18, v) = 0.5 * (S(U, c, 2, j, 18, v) + S(U, c, 1, j, 17, v));
        S(U, c, 18, j, 1, v) = 0.5 * (S(U, c, 17, j, 1, v) + S(U, c, 18, j, 2, v));
        S(U, c, 18, j, 18, v) = 0.5 * (S(U, c, 17, j, 18, v) + S(U, c, 18, j, 17, v));
      }
    for (int k = 1; k <= 18; ++k)
      for (int v = 0; v < NV; ++v) {
        S(U, c, 1, 1, k, v) = 0.5 * (S(U, c, 1, 2, k, v) + S(U, c, 2, 1, k, v));
        S(U, c, 1, 18, k, v) = 0.5 * (S(U, c, 2, 18, k, v) + S(U, c, 1, 17, k, v));
        S(U, c, 18, 1, k, v) = 0.5 * (S(U, c, 17, 1, k, v) + S(U, c, 18, 2, k, v));
        S(U, c, 18, 18, k, v) = 0.5 * (S(U, c, 17, 18, k, v) + S(U, c, 18, 17, k, v));
      }
  }
}

/* ------------------------------------------------------------------------
 * Flux / residual / preconditioned point-implicit RK stage,
 * BCM_Flux_XYZ_Viscous_Runge_kutta.cpp:326-2655.
 */

/* One face state pair -> preconditioned Roe-type flux (Flux.cpp:611-797 for Z;
 * X :1061-1247 and Y are the same with the normal velocity swapped).
 * n is the normal velocity component 1..3 (u, v, w). */
static void roe_face(const double ML[5], const double MR[5], int n, double e, double inF[5],
                     double* dP) {
  double dU[5];
  for (int m = 0; m < 5; ++m) dU[m] = MR[m] - ML[m];
  double _rho = ML[0];
  double _q[3] = {ML[1] / _rho, ML[2] / _rho, ML[3] / _rho};
  double _VV = _q[0] * _q[0] + _q[1] * _q[1] + _q[2] * _q[2];
  double _P = (ML[4] - 0.5 * _rho * _VV) * (K_ - 1);
  double _C = K_ * _P / _rho;
  double _H = 0.5 * _VV + _C / (K_ - 1);
  double rho_ = MR[0];
  double q_[3] = {MR[1] / rho_, MR[2] / rho_, MR[3] / rho_};
  double VV_ = q_[0] * q_[0] + q_[1] * q_[1] + q_[2] * q_[2];
  double P_ = (MR[4] - 0.5 * rho_ * VV_) * (K_ - 1);
  double C_ = K_ * P_ / rho_;
  double H_ = 0.5 * VV_ + C_ / (K_ - 1);

  double temp5 = sqrt(_rho);
  double temp6 = sqrt(rho_);
  double temp4 = temp5 + temp6;
  double rho = sqrt(_rho * rho_);
  double q[3];
  for (int d = 0; d < 3; ++d) q[d] = (temp5 * _q[d] + temp6 * q_[d]) / temp4;
  double VV = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  double H = (temp5 * _H + temp6 * H_) / temp4;
  double C = (H - 0.5 * VV) * (K_ - 1);
  double beta = MAXD(VV / C, e);
  double Un = q[n - 1], _Un = _q[n - 1], Un_ = q_[n - 1];
  double temp = 0.5 * (1 + beta) * Un;
  double Sx = 0.5 * sqrt(4 * beta * C + Un * Un * (1 - beta) * (1 - beta));
  double temp1 = (P_ - _P) / rho / beta / C;
  double temp2 = temp / Sx * (Un_ - _Un);
  double temp3 = 0.5 * (1 - beta) * Un * temp / Sx;
  double deltaU = (Sx - temp3 - beta * fabs(Un)) * temp1 + temp2;
  double deltaP = temp / Sx * (P_ - _P) + (Sx - fabs(Un) + temp3) * rho * (Un_ - _Un);
  double Fav[5];
  Fav[0] = fabs(Un) * dU[0] + deltaU * rho;
  for (int m = 1; m <= 3; ++m) {
    Fav[m] = fabs(Un) * dU[m] + deltaU * rho * q[m - 1];
    if (m == n) Fav[m] = Fav[m] + deltaP;
  }
  Fav[4] = fabs(Un) * dU[4] + deltaU * rho * H + deltaP * Un;
  inF[0] = 0.5 * ((_rho * _Un + rho_ * Un_) - EP_ * Fav[0]);
  for (int m = 1; m <= 3; ++m) {
    if (m == n)
      inF[m] = 0.5 * ((_rho * _q[m - 1] * _q[m - 1] + _P + rho_ * q_[m - 1] * q_[m - 1] + P_) -
                      EP_ * Fav[m]);
    else
      inF[m] = 0.5 * ((_rho * _q[m - 1] * _Un + rho_ * q_[m - 1] * Un_) - EP_ * Fav[m]);
  }
  inF[4] = 0.5 * ((_Un * (3.5 * _P + 0.5 * _rho * _VV) + Un_ * (3.5 * P_ + 0.5 * rho_ * VV_)) -
                  EP_ * Fav[4]);
  *dP = 0.5 * (_P + P_);
}

/* Face states at the two faces of a cell along one axis (Flux.cpp:497-596).
 * s[0..6] are the conserved states at offsets -3..+3 along the axis, f[0..6]
 * the flags; pos is the cell index along the axis. */
static void reconstruct(const double s[7][5], const int f[7], int pos, double ML[5],
                        double MR[5], double MLi[5], double MRi[5]) {
  const double c6 = 1. / 6, c60 = 1. / 60;
  if (f[4] == 1 && f[2] == 1) {
    if (pos == 2 | f[1] != 1) {
      for (int v = 0; v < 5; ++v) {
        ML[v] = c6 * (-1 * s[1][v] + 5 * s[2][v] + 2 * s[3][v]);
        MR[v] = c6 * (2 * s[2][v] + 5 * s[3][v] - 1 * s[4][v]);
      }
    } else {
      for (int v = 0; v < 5; ++v) {
        ML[v] = c60 * (2 * s[0][v] - 13 * s[1][v] + 47 * s[2][v] + 27 * s[3][v] - 3 * s[4][v]);
        MR[v] = c60 * (-3 * s[1][v] + 27 * s[2][v] + 47 * s[3][v] - 13 * s[4][v] + 2 * s[5][v]);
      }
    }
    if (pos == 17 | f[5] != 1) {
      for (int v = 0; v < 5; ++v) {
        MLi[v] = c6 * (-1 * s[2][v] + 5 * s[3][v] + 2 * s[4][v]);
        MRi[v] = c6 * (2 * s[3][v] + 5 * s[4][v] - 1 * s[5][v]);
      }
    } else {
      for (int v = 0; v < 5; ++v) {
        MLi[v] = c60 * (2 * s[1][v] - 13 * s[2][v] + 47 * s[3][v] + 27 * s[4][v] - 3 * s[5][v]);
        MRi[v] = c60 * (-3 * s[2][v] + 27 * s[3][v] + 47 * s[4][v] - 13 * s[5][v] + 2 * s[6][v]);
      }
    }
  } else {
    for (int v = 0; v < 5; ++v) {
      ML[v] = s[2][v];
      MR[v] = s[3][v];
      MLi[v] = s[3][v];
      MRi[v] = s[4][v];
    }
  }
}

typedef struct {
  double U, V, W, T;
} prim_t;

/* Primitive velocity and temperature of a neighbour cell as Flux.cpp:1724-1893
 * computes them (T = P/(rho*R)). */
static inline prim_t nb_prim(const double* u) {
  prim_t p;
  double rho = u[0];
  p.U = u[1] / rho;
  p.V = u[2] / rho;
  p.W = u[3] / rho;
  double P = (u[4] - 0.5 * rho * (p.U * p.U + p.V * p.V + p.W * p.W)) * (K_ - 1);
  p.T = P / (rho * R_);
  return p;
}

/* Viscous face flux (Flux.cpp:1946-2272).  d = face normal axis (0,1,2),
 * side = -1 (backward) or +1 (forward).  Q(o) returns the neighbour primitive at
 * offset o; self holds the cell's own values (T = P/rho/R). */
static void viscous_face(int d, int side, prim_t self, const prim_t nb[3][3][3],
                         double T0, double mu_L, double inv[3], double L[5]) {
#define Q(a, b, c) nb[(a) + 1][(b) + 1][(c) + 1]
  int o[3] = {0, 0, 0};
  o[d] = side;
  prim_t f = Q(o[0], o[1], o[2]);
  double Tf = 0.5 * (self.T + f.T);
  double temp = Tf / T0;
  double mu_E = mu_L * temp * sqrt(temp) * (T0 + 110.) / (Tf + 110.);
  double g[3][4]; /* g[axis][u,v,w,T] */
  for (int t = 0; t < 3; ++t) {
    if (t == d) {
      if (side < 0) {
        g[t][0] = (self.U - f.U) * inv[t];
        g[t][1] = (self.V - f.V) * inv[t];
        g[t][2] = (self.W - f.W) * inv[t];
        g[t][3] = (self.T - f.T) * inv[t];
      } else {
        g[t][0] = (f.U - self.U) * inv[t];
        g[t][1] = (f.V - self.V) * inv[t];
        g[t][2] = (f.W - self.W) * inv[t];
        g[t][3] = (f.T - self.T) * inv[t];
      }
    } else {
      int p1[3] = {0, 0, 0}, p2[3], m1[3] = {0, 0, 0}, m2[3];
      p1[t] = 1;
      m1[t] = -1;
      for (int a = 0; a < 3; ++a) { p2[a] = p1[a] + o[a]; m2[a] = m1[a] + o[a]; }
      prim_t A = Q(p1[0], p1[1], p1[2]), B = Q(p2[0], p2[1], p2[2]);
      prim_t Cc = Q(m1[0], m1[1], m1[2]), D = Q(m2[0], m2[1], m2[2]);
      g[t][0] = (A.U + B.U - Cc.U - D.U) / 4 * inv[t];
      g[t][1] = (A.V + B.V - Cc.V - D.V) / 4 * inv[t];
      g[t][2] = (A.W + B.W - Cc.W - D.W) / 4 * inv[t];
      g[t][3] = (A.T + B.T - Cc.T - D.T) / 4 * inv[t];
    }
  }
#undef Q
  double Uf = 0.5 * (self.U + f.U), Vf = 0.5 * (self.V + f.V), Wf = 0.5 * (self.W + f.W);
  double div = g[0][0] + g[1][1] + g[2][2];
  double sg[3];
  /* stress row d: sigma_{d,m} for m = x,y,z */
  for (int m = 0; m < 3; ++m) {
    if (m == d)
      sg[m] = mu_E * (2 * g[d][d] - 2.0 / 3.0 * div);
    else
      sg[m] = mu_E * (g[m][d] + g[d][m]);
  }
  L[1] = sg[0];
  L[2] = sg[1];
  L[3] = sg[2];
  L[4] = Uf * sg[0] + Vf * sg[1] + Wf * sg[2] + mu_E * CV_ * K_ * g[d][3] / (PR_);
}

static void flux_cell(int c, int i, int j, int k, double deltaT, double deltaTau, double e,
                      const int* FWS, double dx, const double* U1_, const double* U1,
                      const double* U1q, double* Rku1, double* Residual1, const orc_phys* ph) {
  const int IF = F(FWS, c, i, j, k);
  const double* u = &S(U1_, c, i, j, k, 0);
  const double* uu = &S(U1, c, i, j, k, 0);
  const double* uq = &S(U1q, c, i, j, k, 0);
  double Rp[5];
  for (int v = 0; v < 5; ++v) Rp[v] = -(3 * u[v] - 4 * uu[v] + uq[v]) / (2 * deltaT);

  double Rfd[3][5], dPb[3], dPf[3];
  const int pos[3] = {i, j, k};
  for (int d = 0; d < 3; ++d) {
    double s[7][5];
    int f[7];
    for (int o = -3; o <= 3; ++o) {
      int a[3] = {i, j, k};
      a[d] += o;
      if (a[d] < 0 || a[d] >= NP) { /* never used: guarded like k>2 / k<nz */
        f[o + 3] = 0;
        for (int v = 0; v < 5; ++v) s[o + 3][v] = 0;
        continue;
      }
      f[o + 3] = F(FWS, c, a[0], a[1], a[2]);
      for (int v = 0; v < 5; ++v) s[o + 3][v] = S(U1_, c, a[0], a[1], a[2], v);
    }
    double ML[5], MR[5], MLi[5], MRi[5], inF[5], inFi[5];
    reconstruct(s, f, pos[d], ML, MR, MLi, MRi);
    roe_face(ML, MR, d + 1, e, inF, &dPb[d]);
    roe_face(MLi, MRi, d + 1, e, inFi, &dPf[d]);
    for (int v = 0; v < 5; ++v) Rfd[d][v] = inFi[v] - inF[v];
  }
  const double dy = dx, dz = dx;
  double Rf[5];
  for (int v = 0; v < 5; ++v) Rf[v] = -(Rfd[0][v] / dx + Rfd[1][v] / dy + Rfd[2][v] / dz);

  /* viscous terms */
  prim_t nb[3][3][3];
  for (int a = -1; a <= 1; ++a)
    for (int b = -1; b <= 1; ++b)
      for (int cc = -1; cc <= 1; ++cc) {
        int nz = (a != 0) + (b != 0) + (cc != 0);
        if (nz == 1 || nz == 2) nb[a + 1][b + 1][cc + 1] = nb_prim(&S(U1_, c, i + a, j + b, k + cc, 0));
      }
  prim_t self;
  double rho = u[0];
  self.U = u[1] / u[0];
  self.V = u[2] / u[0];
  self.W = u[3] / u[0];
  {
    double VV = self.U * self.U + self.V * self.V + self.W * self.W;
    double P = (u[4] - 0.5 * rho * VV) * (K_ - 1);
    self.T = P / rho / R_;
  }
  double inv[3] = {1. / dx, 1. / dx, 1. / dx};
  double Lb[3][5], Lf[3][5];
  for (int d = 0; d < 3; ++d) {
    viscous_face(d, -1, self, nb, ph->T0, ph->mu_L, inv, Lb[d]);
    viscous_face(d, +1, self, nb, ph->T0, ph->mu_L, inv, Lf[d]);
  }
  double vF[5];
  for (int v = 1; v < 5; ++v)
    vF[v] = (Lf[0][v] - Lb[0][v]) * inv[0] + (Lf[1][v] - Lb[1][v]) * inv[1] +
            (Lf[2][v] - Lb[2][v]) * inv[2];
  double Rk[5];
  Rk[0] = Rp[0] + Rf[0];
  for (int v = 1; v < 5; ++v) Rk[v] = Rp[v] + Rf[v] + vF[v];

  /* preconditioner D = M (Gamma + 3 dtau/(2 dT) M)^-1, Flux.cpp:2295-2343 */
  double U = u[1] / u[0], V = u[2] / u[0], W = u[3] / u[0];
  double VV = U * U + V * V + W * W;
  double P = (u[4] - 0.5 * rho * VV) * (K_ - 1);
  double C = K_ * P / rho;
  double T = P / rho;
  double H = 0.5 * VV + C / (K_ - 1);
  double beta = MAXD(VV / C, e);
  double temp = K_ * T * (2 * deltaT + 3 * deltaTau) * (2 * deltaT + 3 * beta * deltaTau);
  double temp2 = (K_ - 1) * (beta - 1) * deltaT * deltaT;
  double temp3 = H - H * K_ - VV + K_ * (T + VV);
  double d[5][5];
  d[0][0] = (2 * deltaT * (-2 * (-(K_ - 1) * (H - VV) - temp3 * beta) * deltaT + 3 * K_ * T * beta * deltaTau)) / temp;
  d[0][1] = (-4 * U * temp2) / temp;
  d[0][2] = (-4 * V * temp2) / temp;
  d[0][3] = (-4 * W * temp2) / temp;
  d[0][4] = (4 * temp2) / temp;
  d[1][0] = (4 * U * temp3 * (beta - 1) * deltaT * deltaT) / temp;
  d[1][1] = (2 * deltaT * (2 * K_ * (T - U * U * (beta - 1)) * deltaT + 2 * U * U * (beta - 1) * deltaT + 3 * K_ * T * beta * deltaTau)) / temp;
  d[1][2] = (-4 * U * V * temp2) / temp;
  d[1][3] = (-4 * U * W * temp2) / temp;
  d[1][4] = (4 * U * temp2) / temp;
  d[2][0] = (4 * V * temp3 * (beta - 1) * deltaT * deltaT) / temp;
  d[2][1] = (-4 * U * V * temp2) / temp;
  d[2][2] = (2 * deltaT * (2 * K_ * (T - V * V * (beta - 1)) * deltaT + 2 * V * V * (beta - 1) * deltaT + 3 * K_ * T * beta * deltaTau)) / temp;
  d[2][3] = (-4 * V * W * temp2) / temp;
  d[2][4] = (4 * V * temp2) / temp;
  d[3][0] = (4 * W * temp3 * (beta - 1) * deltaT * deltaT) / temp;
  d[3][1] = (-4 * U * W * temp2) / temp;
  d[3][2] = (-4 * V * W * temp2) / temp;
  d[3][3] = (2 * deltaT * (2 * K_ * (T - W * W * (beta - 1)) * deltaT + 2 * W * W * (beta - 1) * deltaT + 3 * K_ * T * beta * deltaTau)) / temp;
  d[3][4] = (4 * W * temp2) / temp;
  d[4][0] = (4 * H * temp3 * (beta - 1) * deltaT * deltaT) / temp;
  d[4][1] = (-4 * H * U * temp2) / temp;
  d[4][2] = (-4 * H * V * temp2) / temp;
  d[4][3] = (-4 * H * W * temp2) / temp;
  d[4][4] = (2 * deltaT * (2 * H * (K_ - 1) * (beta - 1) * deltaT + K_ * T * (2 * deltaT + 3 * beta * deltaTau))) / temp;
  double* out = &S(Rku1, c, i, j, k, 0);
  for (int r = 0; r < 5; ++r) {
    if (IF == 1)
      out[r] = deltaTau * (d[r][0] * Rk[0] + d[r][1] * Rk[1] + d[r][2] * Rk[2] + d[r][3] * Rk[3] +
                           d[r][4] * Rk[4]);
    else
      out[r] = 0;
  }
  if (IF < 1) {
    double* res = &S(Residual1, c, i, j, k, 0);
    res[0] = -(-(dPf[0] - dPb[0]) / dx + vF[1]) * dx * dy * dz;
    res[1] = -(-(dPf[2] - dPb[2]) / dz + vF[3]) * dx * dy * dz;
  }
}

void orc_flux_rk(int ncube, int mpi_ncube, int RK, double deltaT, double deltaTau, double e,
                 const int* FWS, const double* cube_size, double* U1_, const double* U1,
                 const double* U1q, double* U1p1, double* U1p2, double* Rku1, double* Residual1,
                 double* er, const orc_phys* ph, int nthreads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int c = 1; c < ncube; ++c) {
    double dx = cube_size[c] / 16;
    for (int i = ILO; i <= IHI; ++i)
      for (int j = ILO; j <= IHI; ++j)
        for (int k = ILO; k <= IHI; ++k)
          flux_cell(c, i, j, k, deltaT, deltaTau, e, FWS, dx, U1_, U1, U1q, Rku1, Residual1, ph);
  }
  if (RK == 1) {
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1)
    for (int c = 1; c < ncube; ++c)
      for (int i = ILO; i <= IHI; ++i)
        for (int j = ILO; j <= IHI; ++j)
          for (int k = ILO; k <= IHI; ++k)
            if (F(FWS, c, i, j, k) == 1)
              for (int v = 0; v < 5; ++v)
                S(U1p1, c, i, j, k, v) = S(U1_, c, i, j, k, v) =
                    S(U1_, c, i, j, k, v) + S(Rku1, c, i, j, k, v);
  }
  if (RK == 2) {
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1)
    for (int c = 1; c < ncube; ++c)
      for (int i = ILO; i <= IHI; ++i)
        for (int j = ILO; j <= IHI; ++j)
          for (int k = ILO; k <= IHI; ++k)
            if (F(FWS, c, i, j, k) == 1)
              for (int v = 0; v < 5; ++v)
                S(U1_, c, i, j, k, v) = S(U1p2, c, i, j, k, v) =
                    0.75 * S(U1_, c, i, j, k, v) + 0.25 * S(U1p1, c, i, j, k, v) +
                    0.25 * S(Rku1, c, i, j, k, v);
  }
  if (RK == 3) {
    double e1 = 0, e2 = 0, e3 = 0, e4 = 0, e5 = 0, e6 = 0, e7 = 0;
    /* Update and norms in the reference's sequential cube-i-j-k order. */
    for (int c = 1; c < ncube; ++c)
      for (int i = ILO; i <= IHI; ++i)
        for (int j = ILO; j <= IHI; ++j)
          for (int k = ILO; k <= IHI; ++k) {
            double* u = &S(U1_, c, i, j, k, 0);
            double rhoold = u[0];
            double Uold = u[1] / rhoold, Vold = u[2] / rhoold, Wold = u[3] / rhoold;
            double VVold = Uold * Uold + Vold * Vold + Wold * Wold;
            double Pold = (u[4] - 0.5 * rhoold * VVold) * (K_ - 1);
            double Told = Pold / rhoold;
            if (F(FWS, c, i, j, k) == 1) {
              const double* p2 = &S(U1p2, c, i, j, k, 0);
              const double* rk = &S(Rku1, c, i, j, k, 0);
              for (int v = 0; v < 5; ++v)
                u[v] = 1.0 / 3.0 * u[v] + 2.0 / 3.0 * p2[v] + 2.0 / 3.0 * rk[v];
              double rho = u[0];
              double U = u[1] / rho, V = u[2] / rho, W = u[3] / rho;
              double VV = U * U + V * V + W * W;
              double P = (u[4] - 0.5 * rho * VV) * (K_ - 1);
              double T = P / rho;
              e1 = e1 + (P - Pold) * (P - Pold);
              e2 = e2 + (U - Uold) * (U - Uold);
              e3 = e3 + (V - Vold) * (V - Vold);
              e4 = e4 + (W - Wold) * (W - Wold);
              e5 = e5 + (T - Told) * (T - Told);
            }
          }
    for (int c = 1; c < ncube; ++c)
      for (int i = ILO; i <= IHI; ++i)
        for (int j = ILO; j <= IHI; ++j)
          for (int k = ILO; k <= IHI; ++k) {
            e6 = e6 + S(Residual1, c, i, j, k, 0);
            e7 = e7 + S(Residual1, c, i, j, k, 1);
          }
    double DN = 1. / (mpi_ncube * 16 * 16 * 16);
    er[1] = sqrt(e1) * DN;
    er[2] = sqrt(e2) * DN;
    er[3] = sqrt(e3) * DN;
    er[4] = sqrt(e4) * DN;
    er[5] = sqrt(e5) * DN;
    er[6] = e6 / (0.5 * ph->rho0 * ph->U0 * ph->U0 * 2.602);
    er[7] = e7 / (0.5 * ph->rho0 * ph->U0 * ph->U0 * 2.602);
  }
}

/* Physical-time shift, main.cpp:1292-1315: U1q <- U1, U1 <- U1_ on interior cells. */
void orc_shift(int ncube, double* U1, double* U1q, const double* U1_) {
  for (int c = 1; c < ncube; ++c)
    for (int i = ILO; i <= IHI; ++i)
      for (int j = ILO; j <= IHI; ++j)
        for (int k = ILO; k <= IHI; ++k)
          for (int v = 0; v < 5; ++v) {
            S(U1q, c, i, j, k, v) = S(U1, c, i, j, k, v);
            S(U1, c, i, j, k, v) = S(U1_, c, i, j, k, v);
          }
}

/* Case tables the stage driver needs. */
typedef struct {
  int ncube, mpi_ncube;
  const int* FWS;
  const double* cube_size;
  const int* adj_number;
  int neq[3];
  const int* eq[3];
  int nbs[6];       /* X bs_plus, X bs_minus, Y bs_plus, Y bs_minus, Z bs_plus, Z bs_minus */
  const int* bs[6];
  int nbc[6];
  const int* bc[6]; /* Xbc_l, Xbc_u, Ybc_l, Ybc_u, Zbc_l, Zbc_u (1-based) */
  int nib[2];
  const double* ib_weight[2]; /* minus, plus */
  const int* ib_gc[2];
  const int* ib_ip[2];
  orc_phys phys;
} orc_case;

/* One RK stage in the order of main.cpp:1161-1241. */
void orc_stage(const orc_case* cs, int RK, double deltaT, double deltaTau, double e,
               double* U1_, double* U1, double* U1q, double* U1p1, double* U1p2, double* Rku1,
               double* Residual1, double* er, int nthreads) {
  orc_bc_x(cs->nbc[0], cs->nbc[1], cs->bc[0], cs->bc[1], U1_, &cs->phys);
  orc_bc_y(cs->nbc[2], cs->nbc[3], cs->bc[2], cs->bc[3], U1_);
  orc_bc_z(cs->nbc[4], cs->nbc[5], cs->bc[4], cs->bc[5], U1_);
  for (int s = 0; s < 2; ++s)
    orc_ghostcell(cs->nib[s], cs->ib_weight[s], cs->ib_gc[s], cs->ib_ip[s], U1_, &cs->phys);
  orc_interface(cs->ncube, cs->neq, cs->eq, cs->nbs, cs->bs, cs->adj_number, U1_);
  orc_flux_rk(cs->ncube, cs->mpi_ncube, RK, deltaT, deltaTau, e, cs->FWS, cs->cube_size, U1_,
              U1, U1q, U1p1, U1p2, Rku1, Residual1, er, &cs->phys, nthreads);
  orc_interface(cs->ncube, cs->neq, cs->eq, cs->nbs, cs->bs, cs->adj_number, U1_);
}

/* n_steps physical steps of n_inner inner iterations each (fixed counts, no
 * convergence break), then the physical-time shift.  er_hist receives
 * er[1..7] after every inner iteration (7 doubles each) if non-null. */
void orc_run(const orc_case* cs, int n_steps, int n_inner, double deltaT, double deltaTau,
             double e, double* U1_, double* U1, double* U1q, double* U1p1, double* U1p2,
             double* Rku1, double* Residual1, double* er_hist, int nthreads) {
  double er[8] = {0};
  int it = 0;
  for (int step = 0; step < n_steps; ++step) {
    for (int inner = 0; inner < n_inner; ++inner) {
      for (int RK = 1; RK <= 3; ++RK)
        orc_stage(cs, RK, deltaT, deltaTau, e, U1_, U1, U1q, U1p1, U1p2, Rku1, Residual1, er,
                  nthreads);
      if (er_hist)
        for (int m = 0; m < 7; ++m) er_hist[it * 7 + m] = er[m + 1];
      ++it;
    }
    orc_shift(cs->ncube, U1, U1q, U1_);
  }
}

/* n_steps physical steps under the reference's inner-iteration control
 * (main.cpp:1249-1320): after each inner iteration E[m] = max(E[m], er[m])
 * and er[m] /= E[m] for m = 1..5 (the max macro of main.cpp:14); the inner
 * loop ends when all five are below tol or at iteration max_inner (the
 * reference's iteration_end_step), then the time levels shift.  E holds the
 * five running maxima E1..E5 (zero at the start of a run, main.cpp:66-70) and
 * is updated.  iters[step] receives each step's inner-iteration count; rows of
 * er_hist (max_inner per step, 7 doubles each, raw er[1..7]) after the break
 * are left as they were. */
void orc_run_converge(const orc_case* cs, int n_steps, int max_inner, double tol, double* E,
                      int* iters, double deltaT, double deltaTau, double e, double* U1_,
                      double* U1, double* U1q, double* U1p1, double* U1p2, double* Rku1,
                      double* Residual1, double* er_hist, int nthreads) {
  double er[8] = {0};
  for (int step = 0; step < n_steps; ++step) {
    for (int it = 1;; ++it) {
      for (int RK = 1; RK <= 3; ++RK)
        orc_stage(cs, RK, deltaT, deltaTau, e, U1_, U1, U1q, U1p1, U1p2, Rku1, Residual1, er,
                  nthreads);
      if (er_hist)
        for (int m = 0; m < 7; ++m) er_hist[((size_t)step * max_inner + it - 1) * 7 + m] = er[m + 1];
      int below = 1;
      for (int m = 1; m <= 5; ++m) {
        E[m - 1] = E[m - 1] > er[m] ? E[m - 1] : er[m];
        er[m] = er[m] / E[m - 1];
        below = below && er[m] < tol;
      }
      if (below || it == max_inner) {
        iters[step] = it;
        break;
      }
    }
    orc_shift(cs->ncube, U1, U1q, U1_);
  }
}
