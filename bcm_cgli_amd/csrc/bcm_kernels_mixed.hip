// Mixed-precision form of the per-iteration hot path for gfx950.
//
// FP32: the inner iterate U1_ (storage and arithmetic), the inviscid and
// viscous fluxes, the residual, the preconditioner and the RK stages.  The
// iterate is stored as its deviation d = U1_ - Uref from the case's freestream
// state Uref, and fluxes are formed as F(U) - F(Uref) from deviations only, so
// the FP32 rounding applies to the deviations (small away from the body)
// rather than to the full state, whose pressure and energy carry a large
// constant offset at these Mach numbers.
// FP64: the physical time levels U1 and U1q, the dual-time source
// -(3u - 4U1 + U1q)/(2 dT) (a difference of nearly equal states), the norms
// and force sums, and the IB interpolation (bcm_kernels.hip, FP64 arithmetic on
// the FP32 iterate).  The algorithm is the reference's
// (BCM_Flux_XYZ_Viscous_Runge_kutta.cpp:326-2655, cited as Flux.cpp); the
// operation order is free here, so divisions become reciprocal multiplies and
// square roots the hardware approximations.
//
// Kernel shape: one workgroup of 256 lanes per cube, lane (j,k), marching
// along i.  Each lane keeps its own i-column of conserved values (planes
// i-3..i+3) and flags (i-2..i+2) in registers for the x-stencils; the plane
// being processed is staged in LDS for the y- and z-stencils; primitives of
// three planes sit in an LDS ring for the viscous terms.  The kernel reads
// one copy of the iterate and writes the other, so each cell's RK update is
// stored as soon as it is computed: no stage increment array, no RK scratch
// arrays; the two layers next to a face with a neighbour are written into
// the neighbour's ghost layers as well, which replaces the face part of the
// post-update halo exchange.  The reference's RK2/RK3 read U1p1
// and U1p2, which equal U1_ at every fluid cell when those stages run (only
// ghost and non-fluid cells change in between), so the stages use U1_.
#include <hip/hip_runtime.h>
#include <math.h>

#include "bcm_kernels.h"

namespace bcm {
namespace {

constexpr int WG = 256;
constexpr float KGf = 1.4f;
constexpr float KM1f = 0.4f;
constexpr float IKM1f = 2.5f;          // 1/(K-1)
constexpr float IRGf = 1.0f / 287.0f;  // 1/R
constexpr float HEATf = 717.5f * 1.4f / 0.72f;   // Cv K / Pr
constexpr double KM1d = 1.4 - 1;

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fmaxr(float a, float b) { return (a > b) ? a : b; }

// Reference (freestream) state of a case in FP32 and the quantities derived
// from it that the perturbation forms below need.
struct RefF {
  float rho, u[3], m[3], E, P, PoR, T;   // PoR = P0/rho0, T = P0/(rho0 R)
};

__device__ __forceinline__ RefF make_ref(const DevPhys& ph) {
  double r[5];
  ref_state(ph, r);
  RefF f;
  f.rho = (float)r[0];
  f.u[0] = (float)ph.Uin;
  f.u[1] = (float)ph.V0;
  f.u[2] = (float)ph.W0;
  for (int k = 0; k < 3; ++k) f.m[k] = (float)r[k + 1];
  f.E = (float)r[4];
  f.P = (float)ph.P0;
  f.PoR = (float)(ph.P0 / ph.rho0);
  f.T = (float)(ph.P0 / (ph.rho0 * 287.0));
  return f;
}

// One state d = U - Uref in perturbation form: density and its reciprocal,
// velocity u = u0 + du, pressure P = P0 + dP, total enthalpy H.  du and dP are
// formed from the deviations only (no difference of nearly equal numbers):
//   du = (dm - u0 drho) / rho,
//   dP = (K-1) (dE - (drho |u|^2 + rho0 sum du (u + u0)) / 2).
struct Side {
  float r, ir, du[3], u[3], dP, P, H;
};

__device__ __forceinline__ void side(const float (&d)[5], const RefF& rf, Side& s) {
  s.r = rf.rho + d[0];
  s.ir = frcp(s.r);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    s.du[k] = (d[k + 1] - rf.u[k] * d[0]) * s.ir;
    s.u[k] = rf.u[k] + s.du[k];
  }
  const float VV = s.u[0] * s.u[0] + s.u[1] * s.u[1] + s.u[2] * s.u[2];
  const float dke = 0.5f * (d[0] * VV + rf.rho * (s.du[0] * (s.u[0] + rf.u[0]) +
                                                  s.du[1] * (s.u[1] + rf.u[1]) +
                                                  s.du[2] * (s.u[2] + rf.u[2])));
  s.dP = KM1f * (d[4] - dke);
  s.P = rf.P + s.dP;
  s.H = 0.5f * VV + KGf * s.P * s.ir * IKM1f;
}

// Preconditioned Roe-type flux at a face with normal axis d (Flux.cpp:611-797)
// in perturbation form: F returns F(U) - F(Uref), which differs from the flux
// by a constant that cancels in every face difference of the residual.
//   mass      dm_d
//   momentum  dm_m u_d + m0_m du_d + dP [m = d]
//   energy    du_d (E + P) + u0_d (dE + dP)
// The dissipation is built from differences R - L as in the reference.  dP
// returns the deviation of the face pressure average.
__device__ __forceinline__ void roe_f(const float (&L)[5], const float (&R)[5], int d, float e,
                                      const RefF& rf, float (&F)[5], float& dP) {
  Side l, r;
  side(L, rf, l);
  side(R, rf, r);
  const float t5 = fsqrt(l.r), t6 = fsqrt(r.r), it4 = frcp(t5 + t6);
  const float rho = t5 * t6;
  float q[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) q[m] = (t5 * l.u[m] + t6 * r.u[m]) * it4;
  const float VV = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
  const float H = (t5 * l.H + t6 * r.H) * it4;
  const float C = (H - 0.5f * VV) * KM1f;
  const float beta = fmaxr(VV * frcp(C), e);
  const float un = d == 0 ? q[0] : (d == 1 ? q[1] : q[2]);
  const float lun = d == 0 ? l.u[0] : (d == 1 ? l.u[1] : l.u[2]);
  const float run = d == 0 ? r.u[0] : (d == 1 ? r.u[1] : r.u[2]);
  const float ldu = d == 0 ? l.du[0] : (d == 1 ? l.du[1] : l.du[2]);
  const float rdu = d == 0 ? r.du[0] : (d == 1 ? r.du[1] : r.du[2]);
  const float u0n = d == 0 ? rf.u[0] : (d == 1 ? rf.u[1] : rf.u[2]);
  const float temp = 0.5f * (1 + beta) * un;
  const float S = 0.5f * fsqrt(4 * beta * C + un * un * (1 - beta) * (1 - beta));
  const float iS = frcp(S);
  const float dPr = r.dP - l.dP, dun = rdu - ldu;
  const float temp1 = dPr * frcp(rho * beta * C);
  const float tS = temp * iS;
  const float temp3 = 0.5f * (1 - beta) * un * tS;
  const float an = fabsf(un);
  const float deltaU = (S - temp3 - beta * an) * temp1 + tS * dun;
  const float deltaP = tS * dPr + (S - an + temp3) * rho * dun;
  const float dUr = deltaU * rho;
  F[0] = 0.5f * ((L[d + 1] + R[d + 1]) - (an * (R[0] - L[0]) + dUr));
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    const float fav = an * (R[m + 1] - L[m + 1]) + dUr * q[m] + (m == d ? deltaP : 0.f);
    const float cen = (L[m + 1] * lun + rf.m[m] * ldu) + (R[m + 1] * run + rf.m[m] * rdu) +
                      (m == d ? l.dP + r.dP : 0.f);
    F[m + 1] = 0.5f * (cen - fav);
  }
  const float EPl = rf.E + rf.P + L[4] + l.dP, EPr = rf.E + rf.P + R[4] + r.dP;
  const float cenE = (ldu * EPl + u0n * (L[4] + l.dP)) + (rdu * EPr + u0n * (R[4] + r.dP));
  const float fav4 = an * (R[4] - L[4]) + dUr * H + deltaP * un;
  F[4] = 0.5f * (cenE - fav4);
  dP = 0.5f * (l.dP + r.dP);
}

// Face states of order 5, 3 or 1 (Flux.cpp:499-596) from six cells s0..s5 =
// p-2..p+3 of the face between p and p+1.
__device__ __forceinline__ void recon(int order, float s0, float s1, float s2, float s3, float s4,
                                      float s5, float& L, float& R) {
  if (order == 5) {
    L = (1.f / 60) * (2 * s0 - 13 * s1 + 47 * s2 + 27 * s3 - 3 * s4);
    R = (1.f / 60) * (-3 * s1 + 27 * s2 + 47 * s3 - 13 * s4 + 2 * s5);
  } else if (order == 3) {
    L = (1.f / 6) * (-s1 + 5 * s2 + 2 * s3);
    R = (1.f / 6) * (2 * s2 + 5 * s3 - s4);
  } else {
    L = s2;
    R = s3;
  }
}

__device__ __forceinline__ int order_fwd(int p, int fm1, int fp1, int fp2) {
  if (fp1 == 1 && fm1 == 1) return (p == 17 || fp2 != 1) ? 3 : 5;
  return 1;
}
__device__ __forceinline__ int order_bwd(int q, int fm2, int fm1, int fp1) {
  if (fp1 == 1 && fm1 == 1) return (q == 2 || fm2 != 1) ? 3 : 5;
  return 1;
}

constexpr int PR = 18 * 18;       // primitive plane, j,k in [1,18]
constexpr int NF = 17 * 16;       // faces of one axis in one plane

// Flux of the face between p and p+1 along axis d (1 or 2) of the staged
// plane: base is the plane index of cell p, s the index stride of the axis.
__device__ __forceinline__ void plane_face(const float (*cp)[400], int base, int s, int d,
                                           int order, float e, const RefF& rf, float (&F)[5],
                                           float& dP) {
  float L[5], R[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const float* q = cp[v] + base;
    const float s1 = q[-s], s2 = q[0], s3 = q[s], s4 = q[2 * s];
    const float s0 = order == 5 ? q[-2 * s] : 0.f;
    const float s5 = order == 5 ? q[3 * s] : 0.f;
    recon(order, s0, s1, s2, s3, s4, s5, L[v], R[v]);
  }
  roe_f(L, R, d, e, rf, F, dP);
}

// Viscous flux through the face between cell P = (i+a0, j0, k0) and P + e_D
// (Flux.cpp:1946-2272, forward-face formulas of P): one evaluation serves
// both adjacent cells.  ring holds deviations of u, v, w, T of planes i-1..i+1
// (slot = plane & 3, j,k in [1,18]).  L receives the x, y, z stress and the
// energy flux.
template <int D>
__device__ __forceinline__ void visc_face(const float (*ring)[4][18 * 18], int i, int a0, int j0,
                                          int k0, const RefF& rf, float inv, float iT0,
                                          float muL, float sut, float (&L)[4]) {
  constexpr int da = D == 0, dj = D == 1, dk = D == 2;
  auto P = [&](int a, int jj, int kk, int m) -> float {
    return ring[(i + a) & 3][m][(jj - 1) * 18 + (kk - 1)];
  };
  float me[4], fv[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    me[m] = P(a0, j0, k0, m);
    fv[m] = P(a0 + da, j0 + dj, k0 + dk, m);
  }
  const float Tf = rf.T + 0.5f * (me[3] + fv[3]);
  const float temp = Tf * iT0;
  const float mu = muL * temp * fsqrt(temp) * sut * frcp(Tf + 110.f);
  float g[3][4];
#pragma unroll
  for (int tt = 0; tt < 3; ++tt) {
    if (tt == D) {
#pragma unroll
      for (int m = 0; m < 4; ++m) g[tt][m] = (fv[m] - me[m]) * inv;
    } else {
      const int ta = tt == 0, tj = tt == 1, tk = tt == 2;
#pragma unroll
      for (int m = 0; m < 4; ++m)
        g[tt][m] = (P(a0 + ta, j0 + tj, k0 + tk, m) + P(a0 + ta + da, j0 + tj + dj, k0 + tk + dk, m) -
                    P(a0 - ta, j0 - tj, k0 - tk, m) - P(a0 - ta + da, j0 - tj + dj, k0 - tk + dk, m)) *
                   (0.25f * inv);
    }
  }
  const float uf = rf.u[0] + 0.5f * (me[0] + fv[0]), vf = rf.u[1] + 0.5f * (me[1] + fv[1]),
              wf = rf.u[2] + 0.5f * (me[2] + fv[2]);
  const float div = g[0][0] + g[1][1] + g[2][2];
#pragma unroll
  for (int m = 0; m < 3; ++m)
    L[m] = (m == D) ? mu * (2 * g[D][D] - (2.0f / 3.0f) * div) : mu * (g[m][D] + g[D][m]);
  L[3] = uf * L[0] + vf * L[1] + wf * L[2] + mu * HEATf * g[D][3];
}

// Border cells of a 20x20 plane (j or k in {0,1,18,19}): 144 of them.
__device__ __forceinline__ int border20(int t) {
  const int edge[4] = {0, 1, 18, 19};
  if (t < 80) return edge[t / 20] * 20 + t % 20;
  const int u = t - 80;
  return (2 + u / 4) * 20 + edge[u % 4];
}
// Border cells of the 18x18 primitive plane (j or k in {1,18}): 68 of them,
// as (j,k) in padded-plane numbering.
__device__ __forceinline__ int border18(int t) {
  if (t < 18) return 1 * 20 + (1 + t);
  if (t < 36) return 18 * 20 + (1 + t - 18);
  if (t < 52) return (2 + t - 36) * 20 + 1;
  return (2 + t - 52) * 20 + 18;
}
// The border lanes holding the two face ghost cells each corner of the plane
// averages, in corner order (1,1), (1,18), (18,1), (18,18): (1,2) and (2,1),
// (1,17) and (2,18), (18,2) and (17,1), (18,17) and (17,18).  Returns the
// lane's slot 0..7, or -1.
__device__ __forceinline__ int corner_source_slot(int t) {
  switch (t) {
    case 1: return 0;
    case 36: return 1;
    case 16: return 2;
    case 52: return 3;
    case 19: return 4;
    case 51: return 5;
    case 34: return 6;
    case 67: return 7;
    default: return -1;
  }
}

__global__ __launch_bounds__(WG, 2) void k_flux_mixed(
    const float* __restrict__ U, float* __restrict__ Uout, const HaloSrc* __restrict__ halo,
    const uint8_t* __restrict__ cube_mask, const double* __restrict__ U1,
    const double* __restrict__ U1q,
    const int8_t* __restrict__ fws, const double* __restrict__ cube_size, PhysTab pt,
    double deltaT, double deltaTau, double e_d, int RK, double* __restrict__ partials) {
  __shared__ float cp[2][5][400];     // conserved values of plane i (double-buffered)
  __shared__ int8_t fl[2][400];       // flags of plane i
  __shared__ float pr[4][4][PR];      // u, v, w, T of planes (ring of 4)
  __shared__ float fy[5][NF];
  __shared__ float fz[6][NF];         // flux + face pressure average
  __shared__ float vy[4][NF];         // viscous fluxes of the y- and z-faces
  __shared__ float vz[4][NF];
  __shared__ double red[(WG / 64) * 7];
  __shared__ float esrc[2][8][5];     // corner sources of interior planes (by parity)

  const int c = blockIdx.x + 1;
  const int cc = pt.cube_case[c];
  if (cc < 0) return;
  const DevPhys ph = pt.phys[cc];
  const int t = threadIdx.x;
  const int j = (t >> 4) + 2, k = (t & 15) + 2;
  const int col = j * 20 + k;
  const double dxd = cube_size[c] / 16;
  const float inv = (float)(1.0 / dxd), vol = (float)(dxd * dxd * dxd);
  const float e = (float)e_d;
  const RefF rf = make_ref(ph);
  double uref[5];
  ref_state(ph, uref);
  const float iT0 = (float)(1.0 / ph.T0), muL = (float)ph.mu_L;
  const float sut = (float)(ph.T0 + 110.);
  const double i2dT = 1.0 / (2 * deltaT);
  const float dT = (float)deltaT, dtau = (float)deltaTau;
  const float* __restrict__ Uc = U + (size_t)c * CUBE_STRIDE;
  float* __restrict__ Oc = Uout + (size_t)c * CUBE_STRIDE;
  const HaloSrc hs = halo[c];
  const int8_t* __restrict__ Fc = fws + (size_t)c * PLANE;
  // ZBC_IN_FLUX faces: the z-ghost cells (k = 0, 1 or 18, 19) of the interior
  // planes are read as the z-BC sets them, from the cell k = 2 (w negated) or
  // 17 of the same row.  Prefetched border values take the redirected address
  // at once and the negation when they are used.
  const int zm = cube_mask[c];
  const bool zlo = (zm & ZBC_IN_FLUX) != 0, zhi = (zm & (ZBC_IN_FLUX << 1)) != 0;
  auto zsrc = [&](int cell, bool& neg) {
    neg = false;
    if (!(zlo || zhi)) return cell;
    const int kk = cell % 20;
    if (zlo && kk < 2 && cell >= 800 && cell < 7200) {
      neg = true;
      return cell - kk + 2;
    }
    if (zhi && kk > 17 && cell >= 800 && cell < 7200) return cell - kk + 17;
    return cell;
  };
  const double r3 = uref[3];
  const size_t gcol = (size_t)c * CUBE_STRIDE + col;
  // Own column: w[m] = plane i-3+m, fw[m] = flag of plane i-2+m.
  float w[7][5];
  int fw[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) w[0][v] = 1.f;
#pragma unroll
  for (int m = 1; m < 7; ++m)
#pragma unroll
    for (int v = 0; v < 5; ++v) w[m][v] = Uc[v * PLANE + (m - 1) * 400 + col];
#pragma unroll
  for (int m = 0; m < 5; ++m) fw[m] = Fc[m * 400 + col];

  // Deviations of the viscous block's primitives u, v, w and T = P/(rho R)
  // from the reference state: dT = (dP - (P0/rho0) drho) / (rho R).
  auto prims = [&](const float (&q)[5], float& u, float& v, float& ww, float& T) {
    Side sd;
    side(q, rf, sd);
    u = sd.du[0];
    v = sd.du[1];
    ww = sd.du[2];
    T = (sd.dP - rf.PoR * q[0]) * sd.ir * IRGf;
  };
  // Primitive plane p into ring slot p & 3: the own cell from the column
  // window (own holds plane p), the 68 border cells (lanes t < 68) from bq,
  // their conserved values loaded beforehand.
  auto fill_prims = [&](int p, const float (&own)[5], const float (&bq)[5]) {
    float* dst = &pr[p & 3][0][0];
    float u, v, ww, T;
    prims(own, u, v, ww, T);
    const int o = (j - 1) * 18 + (k - 1);
    dst[o] = u;
    dst[PR + o] = v;
    dst[2 * PR + o] = ww;
    dst[3 * PR + o] = T;
    if (t < 68) {
      const int cell = border18(t);
      float q[5];
#pragma unroll
      for (int m = 0; m < 5; ++m) q[m] = bq[m];
      int ea, eb;
      if (p == 1 || p == 18) {
        // Ghost planes: the border cells other than the cube corners are
        // edge ghost cells, the mean of two face ghost cells.
        if (edge_sources(p, cell / 20, cell % 20, ea, eb)) {
          bool na, nb;
          ea = zsrc(ea, na);
          eb = zsrc(eb, nb);
#pragma unroll
          for (int m = 0; m < 5; ++m) {
            const float x = Uc[m * PLANE + ea], y = Uc[m * PLANE + eb];
            q[m] = 0.5f * ((m == 3 && na ? zbc_neg(x, r3) : x) + (m == 3 && nb ? zbc_neg(y, r3) : y));
          }
        }
      } else {
        if (zlo && cell % 20 == 1) q[3] = zbc_neg(q[3], r3);
        // Interior planes: the sources of the plane's corners, edge ghost
        // cells formed by corner_prims.
        const int es = corner_source_slot(t);
        if (es >= 0) {
#pragma unroll
          for (int m = 0; m < 5; ++m) esrc[p & 1][es][m] = q[m];
        }
      }
      prims(q, u, v, ww, T);
      const int ob = (cell / 20 - 1) * 18 + (cell % 20 - 1);
      dst[ob] = u;
      dst[PR + ob] = v;
      dst[2 * PR + ob] = ww;
      dst[3 * PR + ob] = T;
    }
  };
  // The four corners of interior plane p are edge ghost cells: the mean of
  // two face ghost cells each, as the pre-flux exchange's edge passes would
  // leave them (no edge pass runs between the stage's face updates and this
  // kernel).  Only the plane's own face pass reads them; four lanes with one
  // face fewer to evaluate form them from the sources fill_prims kept.
  auto corner_prims = [&](int p) {
    const int q4 = t - (WG - 4);
    if (q4 < 0) return;
    float q[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) q[m] = 0.5f * (esrc[p & 1][2 * q4][m] + esrc[p & 1][2 * q4 + 1][m]);
    float u, v, ww, T;
    prims(q, u, v, ww, T);
    const int o = (q4 & 2 ? 17 * 18 : 0) + (q4 & 1 ? 17 : 0);
    float* dst = &pr[p & 3][0][0];
    dst[o] = u;
    dst[PR + o] = v;
    dst[2 * PR + o] = ww;
    dst[3 * PR + o] = T;
  };
  // Border values are loaded one iteration ahead of their use: bcp/bfl the
  // 144 border cells of the next plane to stage, bpr the 68 primitive-ring
  // border cells of the plane after it.
  auto load_b18 = [&](int p, float (&q)[5]) {
    if (t < 68 && p <= 19) {
      const int cell = border18(t);
      bool neg;
      const int src = zsrc(p * 400 + cell, neg);
#pragma unroll
      for (int v = 0; v < 5; ++v) q[v] = Uc[v * PLANE + src];
    }
  };
  auto load_b20 = [&](int p, float (&q)[5], int& f) {
    if (t < 144 && p <= 17) {
      const int cell = border20(t);
      bool neg;
      const int src = zsrc(p * 400 + cell, neg);
#pragma unroll
      for (int v = 0; v < 5; ++v) q[v] = Uc[v * PLANE + src];
      f = Fc[p * 400 + cell];
    }
  };
  float bcp[5] = {0, 0, 0, 0, 0}, bpr[5] = {1, 0, 0, 0, 0};
  int bfl = 0;
  load_b18(1, bpr);
  fill_prims(1, w[2], bpr);
  load_b18(2, bpr);
  fill_prims(2, w[3], bpr);
  __syncthreads();
  corner_prims(2);
  load_b20(2, bcp, bfl);
  load_b18(3, bpr);

  float cF[5] = {0, 0, 0, 0, 0}, cdP = 0.f;
  float cL[4] = {0, 0, 0, 0};   // viscous x-face flux carried from the previous plane
  int corder = -1;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};

#pragma unroll 1
  for (int i = 2; i <= 17; ++i) {
    const int b = i & 1;
    // ---- stage plane i (conserved, flags) and primitives of plane i+1 ----
#pragma unroll
    for (int v = 0; v < 5; ++v) cp[b][v][col] = w[3][v];
    fl[b][col] = (int8_t)fw[2];
    if (t < 144) {
      const int cell = border20(t);
#pragma unroll
      for (int v = 0; v < 5; ++v)
        cp[b][v][cell] = v == 3 && zlo && cell % 20 < 2 ? zbc_neg(bcp[v], r3) : bcp[v];
      fl[b][cell] = (int8_t)bfl;
    }
    fill_prims(i + 1, w[4], bpr);
    // Loads for later: the next iteration's borders, this cell's time levels
    // folded into the FP64 source constant Bv = 3 Uref - 4 U1 + U1q.
    load_b20(i + 1, bcp, bfl);
    load_b18(i + 2, bpr);
    double Bv[5];
    {
      const size_t gb = gcol + (size_t)i * 400;
#pragma unroll
      for (int v = 0; v < 5; ++v)
        Bv[v] = 3.0 * uref[v] - 4.0 * U1[gb + (size_t)v * PLANE] + U1q[gb + (size_t)v * PLANE];
    }
    __syncthreads();

    // ---- y- and z-faces of plane i: 272 + 272 faces over 256 lanes ----
    if (i + 1 <= 17) corner_prims(i + 1);
    for (int f = t; f < 2 * NF; f += WG) {
      const bool yface = f < NF;
      const int g = yface ? f : f - NF;
      const int p = yface ? g / 16 + 1 : g % 17 + 1;
      const int base = yface ? p * 20 + (g % 16 + 2) : (g / 17 + 2) * 20 + p;
      const int s = yface ? 20 : 1;
      const int8_t* fb = fl[b] + base;
      const int o = p >= 2 ? order_fwd(p, fb[-s], fb[s], fb[2 * s])
                           : order_bwd(2, fb[-s], fb[0], fb[2 * s]);
      float F[5], dP, Lv[4];
      plane_face(cp[b], base, s, yface ? 1 : 2, o, e, rf, F, dP);
      if (yface) {
        visc_face<1>(pr, i, 0, p, g % 16 + 2, rf, inv, iT0, muL, sut, Lv);
#pragma unroll
        for (int v = 0; v < 5; ++v) fy[v][g] = F[v];
#pragma unroll
        for (int v = 0; v < 4; ++v) vy[v][g] = Lv[v];
      } else {
        visc_face<2>(pr, i, 0, g / 17 + 2, p, rf, inv, iT0, muL, sut, Lv);
#pragma unroll
        for (int v = 0; v < 5; ++v) fz[v][g] = F[v];
        fz[5][g] = dP;
#pragma unroll
        for (int v = 0; v < 4; ++v) vz[v][g] = Lv[v];
      }
    }
    __syncthreads();

    const int IF = fw[2];
    // ---- inviscid x: forward face from the column, backward carried ----
    float Rx[5], Ry[5], Rz[5], dPxb, dPxf, dPzb, dPzf;
    {
      const int of = order_fwd(i, fw[1], fw[3], fw[4]);
      float L[5], R[5], Ff[5];
#pragma unroll
      for (int v = 0; v < 5; ++v)
        recon(of, w[1][v], w[2][v], w[3][v], w[4][v], w[5][v], w[6][v], L[v], R[v]);
      roe_f(L, R, 0, e, rf, Ff, dPxf);
      const int ob = order_bwd(i, fw[0], fw[1], fw[3]);
      if (i != 2 && ob == corder) {
        dPxb = cdP;
#pragma unroll
        for (int v = 0; v < 5; ++v) Rx[v] = Ff[v] - cF[v];
      } else {
        float Fb[5];
#pragma unroll
        for (int v = 0; v < 5; ++v)
          recon(ob, w[0][v], w[1][v], w[2][v], w[3][v], w[4][v], w[5][v], L[v], R[v]);
        roe_f(L, R, 0, e, rf, Fb, dPxb);
#pragma unroll
        for (int v = 0; v < 5; ++v) Rx[v] = Ff[v] - Fb[v];
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) cF[v] = Ff[v];
      cdP = dPxf;
      corder = of;
    }
    const int8_t* fc = fl[b] + col;
    // ---- inviscid y ----
    {
      const int fm2 = fc[-40], fm1 = fc[-20], fp1 = fc[20];
      const int ff = (j - 1) * 16 + (k - 2);
      const int ob = order_bwd(j, fm2, fm1, fp1);
      if (j == 2 || ob == order_fwd(j - 1, fm2, fc[0], fp1)) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Ry[v] = fy[v][ff] - fy[v][ff - 16];
      } else {
        float Fb[5], dPu;
        plane_face(cp[b], col - 20, 20, 1, ob, e, rf, Fb, dPu);
#pragma unroll
        for (int v = 0; v < 5; ++v) Ry[v] = fy[v][ff] - Fb[v];
      }
    }
    // ---- inviscid z ----
    {
      const int fm2 = fc[-2], fm1 = fc[-1], fp1 = fc[1];
      const int ff = (j - 2) * 17 + (k - 1);
      const int ob = order_bwd(k, fm2, fm1, fp1);
      dPzf = fz[5][ff];
      if (k == 2 || ob == order_fwd(k - 1, fm2, fc[0], fp1)) {
        dPzb = fz[5][ff - 1];
#pragma unroll
        for (int v = 0; v < 5; ++v) Rz[v] = fz[v][ff] - fz[v][ff - 1];
      } else {
        float Fb[5];
        plane_face(cp[b], col - 1, 1, 2, ob, e, rf, Fb, dPzb);
#pragma unroll
        for (int v = 0; v < 5; ++v) Rz[v] = fz[v][ff] - Fb[v];
      }
    }

    // ---- viscous (Flux.cpp:1722-2279): x-face carried, y/z faces shared ----
    float vF[5];
    {
      float Lf[4];
      visc_face<0>(pr, i, 0, j, k, rf, inv, iT0, muL, sut, Lf);
      if (i == 2) visc_face<0>(pr, i, -1, j, k, rf, inv, iT0, muL, sut, cL);
      const int fyf = (j - 1) * 16 + (k - 2), fzf = (j - 2) * 17 + (k - 1);
      vF[0] = 0.f;
#pragma unroll
      for (int v = 1; v < 5; ++v)
        vF[v] = ((Lf[v - 1] - cL[v - 1]) + (vy[v - 1][fyf] - vy[v - 1][fyf - 16]) +
                 (vz[v - 1][fzf] - vz[v - 1][fzf - 1])) * inv;
#pragma unroll
      for (int v = 0; v < 4; ++v) cL[v] = Lf[v];
    }

    // ---- residual with the FP64 dual-time source ----
    float u[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v] = w[3][v];
    float Rk[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double Rp = -(3.0 * (double)u[v] + Bv[v]) * i2dT;
      const float Rf = -(Rx[v] + Ry[v] + Rz[v]) * inv + (v == 0 ? 0.f : vF[v]);
      Rk[v] = (float)(Rp + (double)Rf);
    }

    // ---- preconditioned increment (Flux.cpp:2295-2375) and RK update ----
    float nv[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) nv[v] = u[v];
    if (IF == 1) {
      Side me;
      side(w[3], rf, me);
      const float U_ = me.u[0], V_ = me.u[1], W_ = me.u[2];
      const float irho = me.ir;
      const float VV = U_ * U_ + V_ * V_ + W_ * W_;
      const float P = me.P;
      const float C = KGf * P * irho;
      const float T = P * irho;
      const float H = 0.5f * VV + C * IKM1f;
      const float beta = fmaxr(VV * frcp(C), e);
      const float bm1 = beta - 1;
      const float itemp = frcp(KGf * T * (2 * dT + 3 * dtau) * (2 * dT + 3 * beta * dtau));
      const float temp2 = KM1f * bm1 * dT * dT;
      const float temp3 = H - H * KGf - VV + KGf * (T + VV);
      const float kTbt = 3 * KGf * T * beta * dtau;
      const float a4 = 4 * temp2 * itemp;              // 4 temp2 / temp
      const float b4 = 4 * temp3 * bm1 * dT * dT * itemp;
      const float d11 = 2 * dT * (-2 * (-KM1f * (H - VV) - temp3 * beta) * dT + kTbt) * itemp;
      const float dmm[3] = {
          2 * dT * (2 * KGf * (T - U_ * U_ * bm1) * dT + 2 * U_ * U_ * bm1 * dT + kTbt) * itemp,
          2 * dT * (2 * KGf * (T - V_ * V_ * bm1) * dT + 2 * V_ * V_ * bm1 * dT + kTbt) * itemp,
          2 * dT * (2 * KGf * (T - W_ * W_ * bm1) * dT + 2 * W_ * W_ * bm1 * dT + kTbt) * itemp};
      const float d55 = 2 * dT * (2 * H * KM1f * bm1 * dT + KGf * T * (2 * dT + 3 * beta * dtau)) * itemp;
      const float vel[3] = {U_, V_, W_};
      // Row 0: d11, -a4 U, -a4 V, -a4 W, a4.  Rows 1..3 (velocity m): b4 vm,
      // -a4 vm vn off-diagonal, dmm on the diagonal, a4 vm.  Row 4: b4 H,
      // -a4 H U, -a4 H V, -a4 H W, d55.
      const float sv = U_ * Rk[1] + V_ * Rk[2] + W_ * Rk[3];
      float out[5];
      out[0] = dtau * (d11 * Rk[0] - a4 * sv + a4 * Rk[4]);
#pragma unroll
      for (int m = 0; m < 3; ++m)
        out[m + 1] = dtau * (b4 * vel[m] * Rk[0] - a4 * vel[m] * (sv - vel[m] * Rk[m + 1]) +
                             dmm[m] * Rk[m + 1] + a4 * vel[m] * Rk[4]);
      out[4] = dtau * (b4 * H * Rk[0] - a4 * H * sv + d55 * Rk[4]);
      float un[5];
      if (RK == 1) {
#pragma unroll
        for (int v = 0; v < 5; ++v) un[v] = u[v] + out[v];
      } else if (RK == 2) {
#pragma unroll
        for (int v = 0; v < 5; ++v) un[v] = 0.75f * u[v] + 0.25f * u[v] + 0.25f * out[v];
      } else {
#pragma unroll
        for (int v = 0; v < 5; ++v)
          un[v] = (1.0f / 3.0f) * u[v] + (2.0f / 3.0f) * u[v] + (2.0f / 3.0f) * out[v];
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) nv[v] = un[v];
      if (RK == 3) {
        // Change norms in FP64 (Flux.cpp:2530-2566).
        double fo[5], fn[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          fo[v] = (double)u[v] + uref[v];
          fn[v] = (double)un[v] + uref[v];
        }
        const double ro = fo[0], iro = 1.0 / ro;
        const double uo = fo[1] * iro, vo = fo[2] * iro, wo = fo[3] * iro;
        const double Po = (fo[4] - 0.5 * ro * (uo * uo + vo * vo + wo * wo)) * KM1d;
        const double r = fn[0], ir = 1.0 / r;
        const double uu = fn[1] * ir, vv = fn[2] * ir, ww = fn[3] * ir;
        const double Pn = (fn[4] - 0.5 * r * (uu * uu + vv * vv + ww * ww)) * KM1d;
        const double dTn = Pn * ir - Po * iro;
        acc[0] += (Pn - Po) * (Pn - Po);
        acc[1] += (uu - uo) * (uu - uo);
        acc[2] += (vv - vo) * (vv - vo);
        acc[3] += (ww - wo) * (ww - wo);
        acc[4] += dTn * dTn;
      }
    } else if (IF < 1 && RK == 3) {
      // Force density of non-fluid cells (Flux.cpp:2395-2398), summed at RK 3.
      acc[5] += (double)(-(-(dPxf - dPxb) * inv + vF[1]) * vol);
      acc[6] += (double)(-(-(dPzf - dPzb) * inv + vF[3]) * vol);
    }
    // Every interior cell goes to Uout; cells next to a face with a
    // neighbour also go to that neighbour's ghost layers.
    {
      const int o = i * 400 + col;
      const int nx = i <= 3 ? hs.src[0] : (i >= 16 ? hs.src[1] : 0);
      const int ny = j <= 3 ? hs.src[2] : (j >= 16 ? hs.src[3] : 0);
      const int nz = k <= 3 ? hs.src[4] : (k >= 16 ? hs.src[5] : 0);
      const int ox = o + (i <= 3 ? 16 * 400 : -16 * 400);
      const int oy = o + (j <= 3 ? 16 * 20 : -16 * 20);
      const int oz = o + (k <= 3 ? 16 : -16);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        Oc[v * PLANE + o] = nv[v];
        if (nx > 0) Uout[(size_t)nx * CUBE_STRIDE + v * PLANE + ox] = nv[v];
        if (ny > 0) Uout[(size_t)ny * CUBE_STRIDE + v * PLANE + oy] = nv[v];
        if (nz > 0) Uout[(size_t)nz * CUBE_STRIDE + v * PLANE + oz] = nv[v];
      }
    }

    // ---- advance the column window ----
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
      for (int v = 0; v < 5; ++v) w[m][v] = w[m + 1][v];
#pragma unroll
    for (int m = 0; m < 4; ++m) fw[m] = fw[m + 1];
    if (i + 4 <= 19) {
#pragma unroll
      for (int v = 0; v < 5; ++v) w[6][v] = Uc[v * PLANE + (i + 4) * 400 + col];
    }
    if (i + 3 <= 19) fw[4] = Fc[(i + 3) * 400 + col];
  }

  // Ghost regions the domain BCs wrote this stage go to Uout as they are (no
  // neighbour writes them), so Uout holds them as an in-place update would.
  {
    const int bm = cube_mask[c];
    if (bm) copy_bc_regions(U, Uout, c, bm, uref[3]);
  }
  if (RK == 3) {
#pragma unroll
    for (int m = 0; m < 7; ++m)
      for (int off = 32; off > 0; off >>= 1) acc[m] += __shfl_down(acc[m], off, 64);
    const int lane = t & 63, wave = t >> 6;
    if (lane == 0)
#pragma unroll
      for (int m = 0; m < 7; ++m) red[wave * 7 + m] = acc[m];
    __syncthreads();
    if (t == 0)
#pragma unroll
      for (int m = 0; m < 7; ++m) {
        double s = red[m];
        for (int wv = 1; wv < WG / 64; ++wv) s += red[wv * 7 + m];
        partials[(size_t)blockIdx.x * 8 + m] = s;
      }
  }
}

__global__ __launch_bounds__(WG) void k_shift_mixed(const float* __restrict__ U,
                                                    double* __restrict__ U1,
                                                    double* __restrict__ U1q, PhysTab pt) {
  const int c = blockIdx.y + 1;
  const int cc = pt.cube_case[c];
  if (cc < 0) return;
  double ref[5];
  ref_state(pt.phys[cc], ref);
  const int i = blockIdx.x + 2;
  const int j = (threadIdx.x >> 4) + 2;
  const int k = (threadIdx.x & 15) + 2;
  const size_t b = (size_t)c * CUBE_STRIDE + i * 400 + j * 20 + k;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const size_t o = b + (size_t)v * PLANE;
    U1q[o] = U1[o];
    U1[o] = (double)U[o] + ref[v];
  }
}

struct Ref5 {
  double r[5];
};

__global__ __launch_bounds__(WG) void k_copy_f(const float* __restrict__ src,
                                               float* __restrict__ dst, size_t n) {
  for (size_t g = (size_t)blockIdx.x * WG + threadIdx.x; g < n; g += (size_t)gridDim.x * WG)
    dst[g] = src[g];
}

__global__ __launch_bounds__(WG) void k_aos_to_soa_f(const double* __restrict__ aos,
                                                     float* __restrict__ soa, int ncells,
                                                     Ref5 ref) {
  const int g = blockIdx.x * WG + threadIdx.x;
  if (g >= ncells) return;
  const int c = g / PLANE, cell = g % PLANE;
#pragma unroll
  for (int v = 0; v < 5; ++v)
    soa[(size_t)c * CUBE_STRIDE + (size_t)v * PLANE + cell] =
        (float)(aos[(size_t)g * 5 + v] - ref.r[v]);
}

__global__ __launch_bounds__(WG) void k_soa_to_aos_f(const float* __restrict__ soa,
                                                     double* __restrict__ aos, int ncells,
                                                     Ref5 ref) {
  const int g = blockIdx.x * WG + threadIdx.x;
  if (g >= ncells) return;
  const int c = g / PLANE, cell = g % PLANE;
#pragma unroll
  for (int v = 0; v < 5; ++v)
    aos[(size_t)g * 5 + v] =
        (double)soa[(size_t)c * CUBE_STRIDE + (size_t)v * PLANE + cell] + ref.r[v];
}

}  // namespace

void launch_flux_mixed(hipStream_t s, const float* Uin, float* Uout, const HaloSrc* halo,
                       const uint8_t* cube_mask, const double* U1, const double* U1q,
                       const int8_t* fws, const double* cube_size, int ncube, PhysTab pt,
                       double deltaT, double deltaTau, double e, int RK, double* partials) {
  if (ncube > 1)
    hipLaunchKernelGGL(k_flux_mixed, dim3(ncube - 1), dim3(WG), 0, s, Uin, Uout, halo, cube_mask,
                       U1, U1q, fws, cube_size, pt, deltaT, deltaTau, e, RK, partials);
}

void launch_copy_f(hipStream_t s, const float* src, float* dst, size_t n) {
  if (n == 0) return;
  const size_t blocks = (n + WG - 1) / WG;
  hipLaunchKernelGGL(k_copy_f, dim3(blocks < 65536 ? (unsigned)blocks : 65536u), dim3(WG), 0, s,
                     src, dst, n);
}

void launch_shift_mixed(hipStream_t s, const float* U, double* U1, double* U1q, int ncube,
                        PhysTab pt) {
  if (ncube > 1)
    hipLaunchKernelGGL(k_shift_mixed, dim3(16, ncube - 1), dim3(WG), 0, s, U, U1, U1q, pt);
}

void launch_aos_to_soa_f(hipStream_t s, const double* aos, float* soa, int ncube,
                         const double* ref5) {
  const int n = ncube * PLANE;
  Ref5 r;
  for (int v = 0; v < 5; ++v) r.r[v] = ref5[v];
  hipLaunchKernelGGL(k_aos_to_soa_f, dim3((n + WG - 1) / WG), dim3(WG), 0, s, aos, soa, n, r);
}

void launch_soa_to_aos_f(hipStream_t s, const float* soa, double* aos, int ncube,
                         const double* ref5) {
  const int n = ncube * PLANE;
  Ref5 r;
  for (int v = 0; v < 5; ++v) r.r[v] = ref5[v];
  hipLaunchKernelGGL(k_soa_to_aos_f, dim3((n + WG - 1) / WG), dim3(WG), 0, s, soa, aos, n, r);
}

}  // namespace bcm
