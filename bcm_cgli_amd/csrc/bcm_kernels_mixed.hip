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
// -(3u - 4U1 + U1q)/(2 dT) (a difference of nearly equal states; its constant
// part 3 Uref - 4 U1 + U1q is formed once per physical step, at the shift), the
// norms and force sums, and the IB interpolation (bcm_kernels.hip, FP64
// arithmetic on the FP32 iterate).  The algorithm is the reference's
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
// First lanes of the 16 that evaluate the y- and z-faces next to the ghost
// layers (waves 1 and 2: wave 0 stages the plane's border cells and forms
// their primitives, wave 3 forms the corner primitives).
#ifndef BCM_MIXED_BND_Y
#define BCM_MIXED_BND_Y 64
#endif
#ifndef BCM_MIXED_BND_Z
#define BCM_MIXED_BND_Z 128
#endif
constexpr int BND_Y = BCM_MIXED_BND_Y, BND_Z = BCM_MIXED_BND_Z;
constexpr float KGf = 1.4f;
constexpr float KM1f = 0.4f;
constexpr float IKM1f = 2.5f;          // 1/(K-1)
constexpr float IRGf = 1.0f / 287.0f;  // 1/R
constexpr float HEATf = 717.5f * 1.4f / 0.72f;   // Cv K / Pr
constexpr double KM1d = 1.4 - 1;

// x of the lane CTRL's data-parallel move selects (0 where it selects none),
// both halves of a double moved alike.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROW_MASK, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROW_MASK, 0xf, false);
  return __hiloint2double(hi, lo);
}
// Sum of x over the 64 lanes of the wave, complete in lane 63, in a fixed
// order: prefix sums along each row of 16 lanes (row_shr 1, 2, 4, 8), then
// row 0's total into row 1 and row 2's into row 3 (row_bcast 15), then the
// first half's total into the second half (row_bcast 31).
__device__ __forceinline__ double wave_sum(double x) {
  x += dpp_d<0x111, 0xf>(x);
  x += dpp_d<0x112, 0xf>(x);
  x += dpp_d<0x114, 0xf>(x);
  x += dpp_d<0x118, 0xf>(x);
  x += dpp_d<0x142, 0xa>(x);
  x += dpp_d<0x143, 0xc>(x);
  return x;
}

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
// p-2..p+3 of the face between p and p+1, as weighted sums with the order's
// weights (zero for the cells an order does not use), so that lanes of one
// wave with different orders run the same instructions:
//   5: L = (2 s0 - 13 s1 + 47 s2 + 27 s3 - 3 s4) / 60,
//      R = (-3 s1 + 27 s2 + 47 s3 - 13 s4 + 2 s5) / 60
//   3: L = (-s1 + 5 s2 + 2 s3) / 6,  R = (2 s2 + 5 s3 - s4) / 6
//   1: L = s2, R = s3.
// Cells outside an order's stencil must hold finite values (callers pass
// an in-range cell in their place).
struct ReconW {
  float a0, a1, a2, a3, a4, b1, b2, b3, b4, b5;
};
__device__ __forceinline__ ReconW recon_weights(int order) {
  const bool o5 = order == 5, o3 = order == 3;
  ReconW w;
  w.a0 = o5 ? 2.f / 60 : 0.f;
  w.a1 = o5 ? -13.f / 60 : (o3 ? -1.f / 6 : 0.f);
  w.a2 = o5 ? 47.f / 60 : (o3 ? 5.f / 6 : 1.f);
  w.a3 = o5 ? 27.f / 60 : (o3 ? 2.f / 6 : 0.f);
  w.a4 = o5 ? -3.f / 60 : 0.f;
  w.b1 = o5 ? -3.f / 60 : 0.f;
  w.b2 = o5 ? 27.f / 60 : (o3 ? 2.f / 6 : 0.f);
  w.b3 = o5 ? 47.f / 60 : (o3 ? 5.f / 6 : 1.f);
  w.b4 = o5 ? -13.f / 60 : (o3 ? -1.f / 6 : 0.f);
  w.b5 = o5 ? 2.f / 60 : 0.f;
  return w;
}
__device__ __forceinline__ void recon(const ReconW& w, float s0, float s1, float s2, float s3,
                                      float s4, float s5, float& L, float& R) {
  L = w.a0 * s0 + w.a1 * s1 + w.a2 * s2 + w.a3 * s3 + w.a4 * s4;
  R = w.b1 * s1 + w.b2 * s2 + w.b3 * s3 + w.b4 * s4 + w.b5 * s5;
}

__device__ __forceinline__ int order_fwd(int p, int fm1, int fp1, int fp2) {
  if (fp1 == 1 && fm1 == 1) return (p == 17 || fp2 != 1) ? 3 : 5;
  return 1;
}
__device__ __forceinline__ int order_bwd(int q, int fm2, int fm1, int fp1) {
  if (fp1 == 1 && fm1 == 1) return (q == 2 || fm2 != 1) ? 3 : 5;
  return 1;
}

// LDS layouts.  Lane t of a workgroup owns cell (j, k) with k = t % 16 + 2
// and j = 2 + w + 4 q for wave w = t / 64 and q = (t / 16) % 4: the four
// 16-cell rows a wave reads at once lie 4 rows apart, so row strides of 20
// (staged plane, primitive ring) put them on distinct LDS banks.
constexpr int PRW = 20;           // row stride of the primitive ring
constexpr int PR = 18 * PRW;      // primitive plane, j,k in [1,18]
constexpr int NF = 17 * 16;       // faces of one axis in one plane
constexpr int NFY = 4 * 80;       // y-face slots (17 rows of 16 in 4 blocks of 5 rows)
__device__ __forceinline__ int ring_idx(int jj, int kk) { return (jj - 1) * PRW + (kk - 1); }
// y-face between rows p and p+1 at k = kk (p = 1..17): block (p-1) % 4, row
// (p-1) / 4 in it, so a wave's rows are consecutive slots.
__device__ __forceinline__ int fy_idx(int p, int kk) {
  return ((p - 1) & 3) * 80 + ((p - 1) >> 2) * 16 + (kk - 2);
}
// z-face between columns p and p+1 in row jj (jj = 2..17, p = 1..17).
__device__ __forceinline__ int fz_idx(int jj, int p) {
  return ((jj - 2) & 3) * 68 + ((jj - 2) >> 2) * 17 + (p - 1);
}

// Flux of the face between p and p+1 along axis D (1 or 2) of the staged
// plane: base is the plane index of cell p.  Orders below 5 read cell p in
// place of the cells p-2 and p+3 they do not use (which may lie beyond the
// plane).
template <int D>
__device__ __forceinline__ void plane_face(const float (*cp)[400], int base, int order, float e,
                                           const RefF& rf, float (&F)[5], float& dP) {
  constexpr int s = D == 1 ? 20 : 1;
  const ReconW rw = recon_weights(order);
  const int o0 = order == 5 ? -2 * s : 0, o5 = order == 5 ? 3 * s : 0;
  float L[5], R[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const float* q = cp[v] + base;
    recon(rw, q[o0], q[-s], q[0], q[s], q[2 * s], q[o5], L[v], R[v]);
  }
  roe_f(L, R, D, e, rf, F, dP);
}

// Viscous flux through the face between cell P = (i+a0, j0, k0) and P + e_D
// (Flux.cpp:1946-2272, forward-face formulas of P): one evaluation serves
// both adjacent cells.  ring holds deviations of u, v, w, T of planes i-1..i+1
// (slot = plane & 3, j,k in [1,18]).  L receives the x, y, z stress and the
// energy flux.
template <int D>
__device__ __forceinline__ void visc_face(const float (*ring)[4][PR], int i, int a0, int j0,
                                          int k0, const RefF& rf, float inv, float iT0,
                                          float muL, float sut, float (&L)[4]) {
  constexpr int da = D == 0, dj = D == 1, dk = D == 2;
  auto P = [&](int a, int jj, int kk, int m) -> float {
    return ring[(i + a) & 3][m][ring_idx(jj, kk)];
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

// y-face (D = 1) between (p, kk) and (p + 1, kk), or z-face (D = 2) between
// (jj, p) and (jj, p + 1), of staged plane i: inviscid flux into fy / fz (face
// slot fy_idx(p, kk), or fz_idx(jj, p); fz also keeps the face pressure
// average) and viscous flux into vy / vz.  Face p = 1 lies between the first
// ghost layer and the first interior cell.
template <int D>
__device__ __forceinline__ void yz_face(const float (*cp)[400], const int8_t* fl,
                                        const float (*ring)[4][PR], int i, int a, int p,
                                        float e, const RefF& rf, float inv, float iT0, float muL,
                                        float sut, float* fo, float* vo) {
  constexpr int W = D == 1 ? NFY : NF;   // row length of the face arrays
  constexpr int s = D == 1 ? 20 : 1;
  const int base = D == 1 ? p * 20 + a : a * 20 + p;
  const int8_t* fb = fl + base;
  const int o = p >= 2 ? order_fwd(p, fb[-s], fb[s], fb[2 * s])
                       : order_bwd(2, fb[-s], fb[0], fb[2 * s]);
  float F[5], dP, Lv[4];
  plane_face<D>(cp, base, o, e, rf, F, dP);
  if (D == 1)
    visc_face<1>(ring, i, 0, p, a, rf, inv, iT0, muL, sut, Lv);
  else
    visc_face<2>(ring, i, 0, a, p, rf, inv, iT0, muL, sut, Lv);
  const int g = D == 1 ? fy_idx(p, a) : fz_idx(a, p);
#pragma unroll
  for (int v = 0; v < 5; ++v) fo[v * W + g] = F[v];
  if (D == 2) fo[5 * W + g] = dP;
#pragma unroll
  for (int v = 0; v < 4; ++v) vo[v * W + g] = Lv[v];
}

// Border cells of a 20x20 plane (j or k in {0,1,18,19}): 144 of them.  (The
// edge rows are formed arithmetically: a table indexed by lane would live in
// memory.)
__device__ __forceinline__ int edge4(int e) { return e < 2 ? e : e + 16; }
__device__ __forceinline__ int border20(int t) {
  if (t < 80) return edge4(t / 20) * 20 + t % 20;
  const int u = t - 80;
  return (2 + u / 4) * 20 + edge4(u % 4);
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

// Segmented march: with nseg > 1 each cube's interior planes 2..17 are split
// into nseg runs of 16/nseg planes, one workgroup each (more workgroups when
// few cubes are resident).  A run starting at i0 > 2 first executes one
// prologue iteration at i0 - 1 that only forms the x-face i0 - 1/2 (inviscid
// and viscous) the run carries into its first plane: the same loop body, and
// so the same instructions and bits, as the unsegmented march computes it
// with.  Norm partials are rows per (group of 4 planes, wave), 16 per cube,
// summed in a fixed order, so every result is bitwise the same for every
// nseg.

__global__ __launch_bounds__(WG, 2) void k_flux_mixed(
    const float* __restrict__ U, float* __restrict__ Uout, const HaloSrc* __restrict__ halo,
    const uint8_t* __restrict__ cube_mask, const double* __restrict__ Bsrc,
    const int8_t* __restrict__ fws, const double* __restrict__ cube_size, PhysTab pt,
    double deltaT, double deltaTau, double e_d, int RK, double* __restrict__ partials,
    int nseg) {
  __shared__ float cp[2][5][400];     // conserved values of plane i (double-buffered)
  __shared__ int8_t fl[2][400];       // flags of plane i
  __shared__ float pr[4][4][PR];      // u, v, w, T of planes (ring of 4)
  __shared__ float fy[5][NFY];
  __shared__ float fz[6][NF];         // flux + face pressure average
  __shared__ float vy[4][NFY];        // viscous fluxes of the y- and z-faces
  __shared__ float vz[4][NF];
  __shared__ float esrc[2][8][5];     // corner sources of interior planes (by parity)

  const int seg = blockIdx.x % nseg;
  const int c = blockIdx.x / nseg + 1;
  const int cc = pt.cube_case[c];
  if (cc < 0) return;
  const DevPhys ph = pt.phys[cc];
  // Planes i0..i1 of this run; the march starts at `start` (a prologue
  // iteration at i0 - 1 for runs after the first).
  const int len = 16 / nseg;
  const int i0 = 2 + seg * len, i1 = i0 + len - 1;
  const int start = i0 == 2 ? 2 : i0 - 1;
  const int t = threadIdx.x;
  const int j = 2 + (t >> 6) + 4 * ((t >> 4) & 3), k = (t & 15) + 2;
  const int col = j * 20 + k;
  // Border work goes to waves 0 and 3, whose face pass has no boundary
  // round: of a staged plane's 144 border cells, s20a (wave 0: 0..63, wave 3:
  // 64..127) and s20b (wave 3: 128..143); of a primitive plane's 68 border
  // cells, s18 (wave 0: 0..63, wave 3: 64..67); -1 for none.  c20a, c20b and
  // c18 are the cells.
  const int s20a = t < 64 ? t : (t >= 192 ? t - 128 : -1);
  const int s20b = t >= 196 && t < 212 ? t - 68 : -1;
  const int s18 = t < 64 ? t : (t >= 192 && t < 196 ? t - 128 : -1);
  const int c20a = border20(s20a < 0 ? 0 : s20a), c20b = border20(s20b < 0 ? 0 : s20b);
  const int c18 = border18(s18 < 0 ? 0 : s18);
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
  // Own column: w[m] = plane i-3+m, fw[m] = flag of plane i-2+m (plane -1,
  // before the first ghost plane, is a placeholder no stencil uses).
  // The leading plane of each (w[6], fw[4]) is loaded at the top of the
  // iteration; its first use, the x-face, comes after the face pass.
  float w[7][5];
  int fw[5];
#pragma unroll
  for (int m = 0; m < 6; ++m)
#pragma unroll
    for (int v = 0; v < 5; ++v)
      w[m][v] = start - 3 + m >= 0 ? Uc[v * PLANE + (start - 3 + m) * 400 + col] : 1.f;
#pragma unroll
  for (int m = 0; m < 4; ++m) fw[m] = Fc[(start - 2 + m) * 400 + col];

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
  // window (own holds plane p).
  auto fill_own = [&](int p, const float (&own)[5]) {
    float* dst = &pr[p & 3][0][0];
    float u, v, ww, T;
    prims(own, u, v, ww, T);
    const int o = ring_idx(j, k);
    dst[o] = u;
    dst[PR + o] = v;
    dst[2 * PR + o] = ww;
    dst[3 * PR + o] = T;
  };
  // The border cells of primitive plane p from bq, their conserved values
  // loaded beforehand.
  auto fill_border = [&](int p, const float (&bq)[5]) {
    float* dst = &pr[p & 3][0][0];
    float u, v, ww, T;
    if (s18 >= 0) {
      const int cell = c18;
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
        const int es = corner_source_slot(s18);
        if (es >= 0) {
#pragma unroll
          for (int m = 0; m < 5; ++m) esrc[p & 1][es][m] = q[m];
        }
      }
      prims(q, u, v, ww, T);
      const int ob = ring_idx(cell / 20, cell % 20);
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
  // face fewer to evaluate form them from the sources fill_border kept.
  auto corner_prims = [&](int p) {
    const int q4 = t - (WG - 4);
    if (q4 < 0) return;
    float q[5];
#pragma unroll
    for (int m = 0; m < 5; ++m) q[m] = 0.5f * (esrc[p & 1][2 * q4][m] + esrc[p & 1][2 * q4 + 1][m]);
    float u, v, ww, T;
    prims(q, u, v, ww, T);
    const int o = ring_idx(q4 & 2 ? 18 : 1, q4 & 1 ? 18 : 1);
    float* dst = &pr[p & 3][0][0];
    dst[o] = u;
    dst[PR + o] = v;
    dst[2 * PR + o] = ww;
    dst[3 * PR + o] = T;
  };
  // Border values are loaded one iteration ahead of their use: bcp/bfl (and
  // bcq/bfq) the border cells of a plane to stage, bpr those of a primitive
  // plane.
  auto load_b18 = [&](int p, float (&q)[5]) {
    if (s18 >= 0 && p <= 19) {
      bool neg;
      const int src = zsrc(p * 400 + c18, neg);
#pragma unroll
      for (int v = 0; v < 5; ++v) q[v] = Uc[v * PLANE + src];
    }
  };
  auto load_b20 = [&](int p, float (&qa)[5], int& fa, float (&qb)[5], int& fb) {
    if (p > 17) return;
    bool neg;
    if (s20a >= 0) {
      const int src = zsrc(p * 400 + c20a, neg);
#pragma unroll
      for (int v = 0; v < 5; ++v) qa[v] = Uc[v * PLANE + src];
      fa = Fc[p * 400 + c20a];
    }
    if (s20b >= 0) {
      const int src = zsrc(p * 400 + c20b, neg);
#pragma unroll
      for (int v = 0; v < 5; ++v) qb[v] = Uc[v * PLANE + src];
      fb = Fc[p * 400 + c20b];
    }
  };
  // The border cells of plane p into its staging buffer.
  auto stage_border = [&](int p, const float (&qa)[5], int fa, const float (&qb)[5], int fb) {
    if (s20a >= 0) {
#pragma unroll
      for (int v = 0; v < 5; ++v)
        cp[p & 1][v][c20a] = v == 3 && zlo && c20a % 20 < 2 ? zbc_neg(qa[v], r3) : qa[v];
      fl[p & 1][c20a] = (int8_t)fa;
    }
    if (s20b >= 0) {
#pragma unroll
      for (int v = 0; v < 5; ++v)
        cp[p & 1][v][c20b] = v == 3 && zlo && c20b % 20 < 2 ? zbc_neg(qb[v], r3) : qb[v];
      fl[p & 1][c20b] = (int8_t)fb;
    }
  };
  // Before the march: primitive planes start-1, start and the border of
  // start+1; the border of the first staged plane.  In the march, iteration
  // i stages plane i's own cells and forms plane i+1's own primitives before
  // its face pass; during the face pass waves 0 and 3 stage the border of
  // plane i+1 and form the border primitives of plane i+2.
  float bcp[5] = {0, 0, 0, 0, 0}, bcq[5] = {0, 0, 0, 0, 0}, bpr[5] = {1, 0, 0, 0, 0};
  int bfl = 0, bfq = 0;
  load_b18(start - 1, bpr);
  fill_own(start - 1, w[2]);
  fill_border(start - 1, bpr);
  load_b18(start, bpr);
  fill_own(start, w[3]);
  fill_border(start, bpr);
  load_b18(start + 1, bpr);
  fill_border(start + 1, bpr);
  if (start >= i0) {
    load_b20(start, bcp, bfl, bcq, bfq);
    stage_border(start, bcp, bfl, bcq, bfq);
  }
  __syncthreads();
  corner_prims(start);
  load_b20(start + 1, bcp, bfl, bcq, bfq);
  load_b18(start + 2, bpr);

  float cF[5] = {0, 0, 0, 0, 0}, cdP = 0.f;
  float cL[4] = {0, 0, 0, 0};   // viscous x-face flux carried from the previous plane
  int corder = -1;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};

#pragma unroll 1
  for (int i = start; i <= i1; ++i) {
    const int b = i & 1;
    const bool pro = i < i0;    // prologue: only the x-face i + 1/2
    // Planes beyond the last ghost plane repeat it (a finite placeholder no
    // stencil uses).
#pragma unroll
    for (int v = 0; v < 5; ++v) w[6][v] = Uc[v * PLANE + min(i + 3, 19) * 400 + col];
    fw[4] = Fc[(i + 2) * 400 + col];
    // ---- stage plane i's own cells, primitives of plane i+1's own cells ----
    if (!pro) {
#pragma unroll
      for (int v = 0; v < 5; ++v) cp[b][v][col] = w[3][v];
      fl[b][col] = (int8_t)fw[2];
    }
    fill_own(i + 1, w[4]);
    // This cell's FP64 source constant Bv = 3 Uref - 4 U1 + U1q, first used
    // in the residual after the face pass.
    double Bv[5] = {0, 0, 0, 0, 0};
    if (!pro) {
      const size_t gb = gcol + (size_t)i * 400;
#pragma unroll
      for (int v = 0; v < 5; ++v) Bv[v] = Bsrc[gb + (size_t)v * PLANE];
    }
    __syncthreads();

    // ---- y- and z-faces of plane i: 272 + 272 faces over 256 lanes ----
    if (i + 1 <= 17) corner_prims(i + 1);
    // Border work for later planes (waves 0 and 3), and its loads for the
    // next iteration.  Plane i+1's staging buffer last held plane i-1, whose
    // readers all passed this iteration's first barrier; ring slot i+2 last
    // held plane i-2.
    if (i + 1 <= i1) stage_border(i + 1, bcp, bfl, bcq, bfq);
    if (i + 2 <= i1 + 1) fill_border(i + 2, bpr);
    if (i + 2 <= i1) load_b20(i + 2, bcp, bfl, bcq, bfq);
    if (i + 3 <= i1 + 1) load_b18(i + 3, bpr);
    // Round 0: each lane the y-face above its cell; round 1: the z-face above
    // it (every wave evaluates one face kind at a time); rounds 2 and 3: the
    // 16 y-faces and the 16 z-faces between the first ghost layer and the
    // first interior cells, on lanes of two different waves.  One face per
    // loop iteration keeps the faces' working sets apart.
#pragma unroll 1
    for (int r = pro ? 4 : 0; r < 4; ++r) {
      const bool bnd = r >= 2;
      const int l0 = r == 2 ? BND_Y : BND_Z;
      if (bnd && (t < l0 || t >= l0 + 16)) continue;
      if ((r & 1) == 0)
        yz_face<1>(cp[b], fl[b], pr, i, bnd ? t - l0 + 2 : k, bnd ? 1 : j, e, rf, inv, iT0, muL,
                   sut, &fy[0][0], &vy[0][0]);
      else
        yz_face<2>(cp[b], fl[b], pr, i, bnd ? t - l0 + 2 : j, bnd ? 1 : k, e, rf, inv, iT0, muL,
                   sut, &fz[0][0], &vz[0][0]);
    }
    __syncthreads();

    const int IF = fw[2];
    // ---- inviscid x: forward face from the column, backward carried ----
    float Rx[5], Ry[5], Rz[5], dPxb, dPxf, dPzb, dPzf;
    {
      // Column cells an order does not use hold finite values (placeholder
      // planes beyond the cube repeat a loaded one).
      const int of = order_fwd(i, fw[1], fw[3], fw[4]);
      float L[5], R[5], Ff[5];
      const ReconW rwf = recon_weights(of);
#pragma unroll
      for (int v = 0; v < 5; ++v)
        recon(rwf, w[1][v], w[2][v], w[3][v], w[4][v], w[5][v], w[6][v], L[v], R[v]);
      roe_f(L, R, 0, e, rf, Ff, dPxf);
      const int ob = order_bwd(i, fw[0], fw[1], fw[3]);
      if (pro) {
      } else if (i != 2 && ob == corder) {
        dPxb = cdP;
#pragma unroll
        for (int v = 0; v < 5; ++v) Rx[v] = Ff[v] - cF[v];
      } else {
        float Fb[5];
        const ReconW rwb = recon_weights(ob);
#pragma unroll
        for (int v = 0; v < 5; ++v)
          recon(rwb, w[0][v], w[1][v], w[2][v], w[3][v], w[4][v], w[5][v], L[v], R[v]);
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
    if (!pro) {
      const int fm2 = fc[-40], fm1 = fc[-20], fp1 = fc[20];
      const int ff = fy_idx(j, k), fb = fy_idx(j - 1, k);
      const int ob = order_bwd(j, fm2, fm1, fp1);
      if (j == 2 || ob == order_fwd(j - 1, fm2, fc[0], fp1)) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Ry[v] = fy[v][ff] - fy[v][fb];
      } else {
        float Fb[5], dPu;
        plane_face<1>(cp[b], col - 20, ob, e, rf, Fb, dPu);
#pragma unroll
        for (int v = 0; v < 5; ++v) Ry[v] = fy[v][ff] - Fb[v];
      }
    }
    // ---- inviscid z ----
    if (!pro) {
      const int fm2 = fc[-2], fm1 = fc[-1], fp1 = fc[1];
      const int ff = fz_idx(j, k);   // the backward face is ff - 1
      const int ob = order_bwd(k, fm2, fm1, fp1);
      dPzf = fz[5][ff];
      if (k == 2 || ob == order_fwd(k - 1, fm2, fc[0], fp1)) {
        dPzb = fz[5][ff - 1];
#pragma unroll
        for (int v = 0; v < 5; ++v) Rz[v] = fz[v][ff] - fz[v][ff - 1];
      } else {
        float Fb[5];
        plane_face<2>(cp[b], col - 1, ob, e, rf, Fb, dPzb);
#pragma unroll
        for (int v = 0; v < 5; ++v) Rz[v] = fz[v][ff] - Fb[v];
      }
    }

    // ---- viscous (Flux.cpp:1722-2279): x-face carried, y/z faces shared ----
    float vF[5] = {0, 0, 0, 0, 0};
    {
      float Lf[4];
      visc_face<0>(pr, i, 0, j, k, rf, inv, iT0, muL, sut, Lf);
      if (i == 2) visc_face<0>(pr, i, -1, j, k, rf, inv, iT0, muL, sut, cL);
      if (!pro) {
        const int fyf = fy_idx(j, k), fyb = fy_idx(j - 1, k), fzf = fz_idx(j, k);
#pragma unroll
        for (int v = 1; v < 5; ++v)
          vF[v] = ((Lf[v - 1] - cL[v - 1]) + (vy[v - 1][fyf] - vy[v - 1][fyb]) +
                   (vz[v - 1][fzf] - vz[v - 1][fzf - 1])) * inv;
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) cL[v] = Lf[v];
    }
    if (!pro) {

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
    // Norm partials at the end of each group of 4 planes: one row per wave.
    if (RK == 3 && ((i - 2) & 3) == 3) {
#pragma unroll
      for (int m = 0; m < 7; ++m) acc[m] = wave_sum(acc[m]);
      if ((t & 63) == 63) {
        double* row =
            partials + ((size_t)(c - 1) * MIXED_ROWS_PER_CUBE + ((i - 2) >> 2) * 4 + (t >> 6)) * 8;
#pragma unroll
        for (int m = 0; m < 7; ++m) row[m] = acc[m];
      }
#pragma unroll
      for (int m = 0; m < 7; ++m) acc[m] = 0.0;
    }
    }  // !pro

    // ---- advance the column window ----
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
      for (int v = 0; v < 5; ++v) w[m][v] = w[m + 1][v];
#pragma unroll
    for (int m = 0; m < 4; ++m) fw[m] = fw[m + 1];
  }

  // Ghost regions the domain BCs wrote this stage go to Uout as they are (no
  // neighbour writes them), so Uout holds them as an in-place update would.
  if (seg == 0) {
    const int bm = cube_mask[c];
    if (bm) copy_bc_regions(U, Uout, c, bm, uref[3]);
  }
}

__global__ __launch_bounds__(WG) void k_shift_mixed(const float* __restrict__ U,
                                                    double* __restrict__ U1,
                                                    double* __restrict__ U1q,
                                                    double* __restrict__ Bsrc, PhysTab pt) {
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
    const double q = U1[o], n = (double)U[o] + ref[v];
    U1q[o] = q;
    U1[o] = n;
    Bsrc[o] = 3.0 * ref[v] - 4.0 * n + q;
  }
}

// Bsrc <- 3 Uref - 4 U1 + U1q on the interior cells of cubes first + blockIdx.y.
__global__ __launch_bounds__(WG) void k_bsrc_mixed(const double* __restrict__ U1,
                                                   const double* __restrict__ U1q,
                                                   double* __restrict__ Bsrc, int first,
                                                   PhysTab pt) {
  const int c = first + blockIdx.y;
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
    Bsrc[o] = 3.0 * ref[v] - 4.0 * U1[o] + U1q[o];
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
                       const uint8_t* cube_mask, const double* Bsrc, const int8_t* fws,
                       const double* cube_size, int ncube, PhysTab pt, double deltaT,
                       double deltaTau, double e, int RK, double* partials, int nseg) {
  if (ncube > 1)
    hipLaunchKernelGGL(k_flux_mixed, dim3((ncube - 1) * nseg), dim3(WG), 0, s, Uin, Uout, halo,
                       cube_mask, Bsrc, fws, cube_size, pt, deltaT, deltaTau, e, RK, partials,
                       nseg);
}

void launch_copy_f(hipStream_t s, const float* src, float* dst, size_t n) {
  if (n == 0) return;
  const size_t blocks = (n + WG - 1) / WG;
  hipLaunchKernelGGL(k_copy_f, dim3(blocks < 65536 ? (unsigned)blocks : 65536u), dim3(WG), 0, s,
                     src, dst, n);
}

void launch_shift_mixed(hipStream_t s, const float* U, double* U1, double* U1q, double* Bsrc,
                        int ncube, PhysTab pt) {
  if (ncube > 1)
    hipLaunchKernelGGL(k_shift_mixed, dim3(16, ncube - 1), dim3(WG), 0, s, U, U1, U1q, Bsrc, pt);
}

void launch_bsrc_mixed(hipStream_t s, const double* U1, const double* U1q, double* Bsrc, int first,
                       int n, PhysTab pt) {
  if (n > 0)
    hipLaunchKernelGGL(k_bsrc_mixed, dim3(16, n), dim3(WG), 0, s, U1, U1q, Bsrc, first, pt);
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
