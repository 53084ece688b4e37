// CDNA4 (gfx950) kernels of the BCM per-iteration hot path, FP64.
//
// Every kernel evaluates the reference expressions in the reference order and
// the library is built with -ffp-contract=off, so results are bitwise those
// of the reference's FP64 CPU path (checked against the oracle in tests/).
// Citations are file:line in the reference tree; "Flux.cpp" is
// BCM_Flux_XYZ_Viscous_Runge_kutta.cpp.
#include <hip/hip_runtime.h>
#include <math.h>
#include <type_traits>

#include "bcm_kernels.h"

namespace bcm {
namespace {

constexpr double KG = 1.4;          // K
constexpr double KM1 = 1.4 - 1;     // (K-1), folded exactly as the reference folds it
constexpr double RG = 287.0;        // R
constexpr double CVG = 717.5;       // Cv
constexpr double PRG = 0.72;        // Pr_L
constexpr int WG = 256;

__device__ __forceinline__ size_t sidx(int c, int v, int i, int j, int k) {
  return (size_t)c * CUBE_STRIDE + (size_t)v * PLANE + i * 400 + j * 20 + k;
}
__device__ __forceinline__ int fidx(int c, int i, int j, int k) {
  return c * PLANE + i * 400 + j * 20 + k;
}
// max(a,b) of the reference: (((a)>(b))?(a):(b)), NaN falls to b.
__device__ __forceinline__ double rmax(double a, double b) { return (a > b) ? a : b; }

// Several correctly rounded divisions by one denominator.  a / b compiles for
// gfx950 to: the denominator scaled (div_scale), its reciprocal refined by
// two Newton steps, then per quotient the numerator scaled, one product, one
// correction and the special-case fixup (div_fmas, div_fixup).  DivBy keeps
// the denominator's part once, scaled as for the numerator 1 (a zero
// numerator there would give NaN; the quotient's fixup handles zeros);
// div_sh runs the quotient's part.  The bits are those of a / b whenever the
// denominator's scaling does not depend on the numerator: for finite
// operands, unless an exponent difference reaches the scaling thresholds
// (768) or an operand is subnormal, far outside the solver's values; the
// self-test below compares against a / b.
struct DivBy {
  double b, ds, r;   // denominator, as scaled, refined reciprocal of ds
};
__device__ __forceinline__ DivBy div_by(double b) {
  bool f;
  DivBy d;
  d.b = b;
  d.ds = __builtin_amdgcn_div_scale(1.0, b, false, &f);
  const double r0 = __builtin_amdgcn_rcp(d.ds);
  const double e0 = __builtin_fma(-d.ds, r0, 1.0);
  const double r1 = __builtin_fma(r0, e0, r0);
  const double e1 = __builtin_fma(-d.ds, r1, 1.0);
  d.r = __builtin_fma(r1, e1, r1);
  return d;
}
__device__ __forceinline__ double div_sh(const DivBy& d, double a) {
  bool f;
  const double as = __builtin_amdgcn_div_scale(a, d.b, true, &f);
  const double q = as * d.r;
  const double rem = __builtin_fma(-d.ds, q, as);
  return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(rem, d.r, q, f), d.b, a);
}

// Full-state access to an inner iterate.  FP64 storage holds the state
// itself; FP32 storage holds its deviation from the case's reference state.
struct NoOff {
  static constexpr bool exact = true;   // the reference's operations, bitwise
  __device__ explicit NoOff(const DevPhys&) {}
  __device__ double ld(const double* U, size_t o, int) const { return U[o]; }
  __device__ void st(double* U, size_t o, int, double x) const { U[o] = x; }
  __device__ double rt(double x, int) const { return x; }
};
struct RefOff {
  static constexpr bool exact = false;  // mixed path: operation order free
  double r[5];
  __device__ explicit RefOff(const DevPhys& ph) { ref_state(ph, r); }
  __device__ double ld(const float* U, size_t o, int v) const { return (double)U[o] + r[v]; }
  __device__ void st(float* U, size_t o, int v, double x) const { U[o] = (float)(x - r[v]); }
  // x as storage would return it.
  __device__ double rt(double x, int v) const { return (double)(float)(x - r[v]) + r[v]; }
};
template <typename T> struct OffOf;
template <> struct OffOf<double> { using type = NoOff; };
template <> struct OffOf<float> { using type = RefOff; };

// ---------------------------------------------------------------------------
// Preconditioned Roe-type flux at one face (Flux.cpp:611-797; X :1061-1247,
// Y :1520-1695).  D is the face-normal axis.
template <int D>
__device__ __forceinline__ void roe_flux(const double (&L)[5], const double (&R)[5], double e,
                                         double (&F)[5], double& dP) {
  // Divisions sharing a denominator go through DivBy (the same bits).
  const double lr = L[0];
  const DivBy dl = div_by(lr);
  const double lq0 = div_sh(dl, L[1]), lq1 = div_sh(dl, L[2]), lq2 = div_sh(dl, L[3]);
  const double lVV = lq0 * lq0 + lq1 * lq1 + lq2 * lq2;
  const double lP = (L[4] - 0.5 * lr * lVV) * KM1;
  const double lC = div_sh(dl, KG * lP);
  const DivBy dk = div_by(KM1);
  const double lH = 0.5 * lVV + div_sh(dk, lC);
  const double rr = R[0];
  const DivBy dr = div_by(rr);
  const double rq0 = div_sh(dr, R[1]), rq1 = div_sh(dr, R[2]), rq2 = div_sh(dr, R[3]);
  const double rVV = rq0 * rq0 + rq1 * rq1 + rq2 * rq2;
  const double rP = (R[4] - 0.5 * rr * rVV) * KM1;
  const double rC = div_sh(dr, KG * rP);
  const double rH = 0.5 * rVV + div_sh(dk, rC);

  const double t5 = sqrt(lr), t6 = sqrt(rr), t4 = t5 + t6;
  const double rho = sqrt(lr * rr);
  const DivBy d4 = div_by(t4);
  const double q0 = div_sh(d4, t5 * lq0 + t6 * rq0);
  const double q1 = div_sh(d4, t5 * lq1 + t6 * rq1);
  const double q2 = div_sh(d4, t5 * lq2 + t6 * rq2);
  const double VV = q0 * q0 + q1 * q1 + q2 * q2;
  const double H = div_sh(d4, t5 * lH + t6 * rH);
  const double C = (H - 0.5 * VV) * KM1;
  const DivBy dC = div_by(C);
  const double beta = rmax(div_sh(dC, VV), e);
  const double un = D == 0 ? q0 : (D == 1 ? q1 : q2);
  const double lun = D == 0 ? lq0 : (D == 1 ? lq1 : lq2);
  const double run = D == 0 ? rq0 : (D == 1 ? rq1 : rq2);
  const double temp = 0.5 * (1 + beta) * un;
  const double S = 0.5 * sqrt(4 * beta * C + un * un * (1 - beta) * (1 - beta));
  const double temp1 = div_sh(dC, (rP - lP) / rho / beta);
  const DivBy dS = div_by(S);
  const double tS = div_sh(dS, temp);
  const double temp2 = tS * (run - lun);
  const double temp3 = div_sh(dS, 0.5 * (1 - beta) * un * temp);
  const double an = fabs(un);
  const double deltaU = (S - temp3 - beta * an) * temp1 + temp2;
  const double deltaP = tS * (rP - lP) + (S - an + temp3) * rho * (run - lun);

  // Dissipation and the averaged physical flux; Ep = 1 multiplies exactly.
  const double fav0 = an * (R[0] - L[0]) + deltaU * rho;
  double fav1 = an * (R[1] - L[1]) + deltaU * rho * q0;
  double fav2 = an * (R[2] - L[2]) + deltaU * rho * q1;
  double fav3 = an * (R[3] - L[3]) + deltaU * rho * q2;
  if (D == 0) fav1 = fav1 + deltaP;
  if (D == 1) fav2 = fav2 + deltaP;
  if (D == 2) fav3 = fav3 + deltaP;
  const double fav4 = an * (R[4] - L[4]) + deltaU * rho * H + deltaP * un;
  F[0] = 0.5 * ((lr * lun + rr * run) - fav0);
  F[1] = D == 0 ? 0.5 * ((lr * lq0 * lq0 + lP + rr * rq0 * rq0 + rP) - fav1)
                : 0.5 * ((lr * lq0 * lun + rr * rq0 * run) - fav1);
  F[2] = D == 1 ? 0.5 * ((lr * lq1 * lq1 + lP + rr * rq1 * rq1 + rP) - fav2)
                : 0.5 * ((lr * lq1 * lun + rr * rq1 * run) - fav2);
  F[3] = D == 2 ? 0.5 * ((lr * lq2 * lq2 + lP + rr * rq2 * rq2 + rP) - fav3)
                : 0.5 * ((lr * lq2 * lun + rr * rq2 * run) - fav3);
  F[4] = 0.5 * ((lun * (3.5 * lP + 0.5 * lr * lVV) + run * (3.5 * rP + 0.5 * rr * rVV)) - fav4);
  dP = 0.5 * (lP + rP);
}

// Inviscid flux difference of one cell along axis D (Flux.cpp:367-806 for Z):
// upwind-biased face states of 5th, 3rd or 1st order depending on the cell
// position and the neighbour flags, then the two face fluxes.
template <int D>
__device__ __forceinline__ void axis_flux(const double* __restrict__ U,
                                          const int8_t* __restrict__ fws, int c, int i, int j,
                                          int k, double e, double (&Rf)[5], double& dPb,
                                          double& dPf) {
  constexpr int st = D == 0 ? 400 : (D == 1 ? 20 : 1);
  const int pos = D == 0 ? i : (D == 1 ? j : k);
  const size_t b = sidx(c, 0, i, j, k);
  const int8_t* fb = fws + fidx(c, i, j, k);
  const int fm2 = fb[-2 * st], fm1 = fb[-st], fp1 = fb[st], fp2 = fb[2 * st];
  double s[7][5];
#pragma unroll
  for (int o = -2; o <= 2; ++o)
#pragma unroll
    for (int v = 0; v < 5; ++v) s[o + 3][v] = U[b + (size_t)v * PLANE + o * st];
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    s[0][v] = pos > 2 ? U[b + (size_t)v * PLANE - 3 * st] : 0.0;
    s[6][v] = pos < 17 ? U[b + (size_t)v * PLANE + 3 * st] : 0.0;
  }
  const double c6 = 1. / 6, c60 = 1. / 60;
  double ML[5], MR[5], MLi[5], MRi[5];
  if (fp1 == 1 && fm1 == 1) {
    if (pos == 2 || fm2 != 1) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        ML[v] = c6 * (-s[1][v] + 5 * s[2][v] + 2 * s[3][v]);
        MR[v] = c6 * (2 * s[2][v] + 5 * s[3][v] - s[4][v]);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        ML[v] = c60 * (2 * s[0][v] - 13 * s[1][v] + 47 * s[2][v] + 27 * s[3][v] - 3 * s[4][v]);
        MR[v] = c60 * (-3 * s[1][v] + 27 * s[2][v] + 47 * s[3][v] - 13 * s[4][v] + 2 * s[5][v]);
      }
    }
    if (pos == 17 || fp2 != 1) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        MLi[v] = c6 * (-s[2][v] + 5 * s[3][v] + 2 * s[4][v]);
        MRi[v] = c6 * (2 * s[3][v] + 5 * s[4][v] - s[5][v]);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        MLi[v] = c60 * (2 * s[1][v] - 13 * s[2][v] + 47 * s[3][v] + 27 * s[4][v] - 3 * s[5][v]);
        MRi[v] = c60 * (-3 * s[2][v] + 27 * s[3][v] + 47 * s[4][v] - 13 * s[5][v] + 2 * s[6][v]);
      }
    }
  } else {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      ML[v] = s[2][v];
      MR[v] = s[3][v];
      MLi[v] = s[3][v];
      MRi[v] = s[4][v];
    }
  }
  double Fb[5], Ff[5];
  roe_flux<D>(ML, MR, e, Fb, dPb);
  roe_flux<D>(MLi, MRi, e, Ff, dPf);
#pragma unroll
  for (int v = 0; v < 5; ++v) Rf[v] = Ff[v] - Fb[v];
}

struct Prim {
  double u, v, w, T;
};

// Neighbour primitives as the viscous block computes them (T = P/(rho*R)).
__device__ __forceinline__ Prim nb_prim(const double* __restrict__ U, size_t b) {
  const double r = U[b];
  Prim p;
  p.u = U[b + PLANE] / r;
  p.v = U[b + 2 * PLANE] / r;
  p.w = U[b + 3 * PLANE] / r;
  const double P = (U[b + 4 * PLANE] - 0.5 * r * (p.u * p.u + p.v * p.v + p.w * p.w)) * KM1;
  p.T = P / (r * RG);
  return p;
}

// Slot of a neighbour offset (a,b,c) in {-1,0,1}^3 in a 27-entry table.
__device__ __forceinline__ constexpr int nslot(int a, int b, int c) {
  return (a + 1) * 9 + (b + 1) * 3 + (c + 1);
}

// Viscous flux through the face of axis D on side SIDE (-1 backward,
// +1 forward), Flux.cpp:1946-2272.
template <int D, int SIDE>
__device__ __forceinline__ void visc_face(const Prim (&q)[27], const Prim& me, double T0,
                                          double mu_L, double inv, double (&L)[5]) {
  constexpr int o0 = D == 0 ? SIDE : 0, o1 = D == 1 ? SIDE : 0, o2 = D == 2 ? SIDE : 0;
  const Prim f = q[nslot(o0, o1, o2)];
  const double Tf = 0.5 * (me.T + f.T);
  const double temp = Tf / T0;
  const double mu = mu_L * temp * sqrt(temp) * (T0 + 110.) / (Tf + 110.);
  double g[3][4];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    if (t == D) {
      if (SIDE < 0) {
        g[t][0] = (me.u - f.u) * inv;
        g[t][1] = (me.v - f.v) * inv;
        g[t][2] = (me.w - f.w) * inv;
        g[t][3] = (me.T - f.T) * inv;
      } else {
        g[t][0] = (f.u - me.u) * inv;
        g[t][1] = (f.v - me.v) * inv;
        g[t][2] = (f.w - me.w) * inv;
        g[t][3] = (f.T - me.T) * inv;
      }
    } else {
      const int p0 = t == 0 ? 1 : 0, p1 = t == 1 ? 1 : 0, p2 = t == 2 ? 1 : 0;
      const Prim& A = q[nslot(p0, p1, p2)];
      const Prim& B = q[nslot(p0 + o0, p1 + o1, p2 + o2)];
      const Prim& Cm = q[nslot(-p0, -p1, -p2)];
      const Prim& Dm = q[nslot(-p0 + o0, -p1 + o1, -p2 + o2)];
      g[t][0] = (A.u + B.u - Cm.u - Dm.u) / 4 * inv;
      g[t][1] = (A.v + B.v - Cm.v - Dm.v) / 4 * inv;
      g[t][2] = (A.w + B.w - Cm.w - Dm.w) / 4 * inv;
      g[t][3] = (A.T + B.T - Cm.T - Dm.T) / 4 * inv;
    }
  }
  const double uf = 0.5 * (me.u + f.u), vf = 0.5 * (me.v + f.v), wf = 0.5 * (me.w + f.w);
  const double div = g[0][0] + g[1][1] + g[2][2];
  double sg[3];
#pragma unroll
  for (int m = 0; m < 3; ++m)
    sg[m] = (m == D) ? mu * (2 * g[D][D] - 2.0 / 3.0 * div) : mu * (g[m][D] + g[D][m]);
  L[1] = sg[0];
  L[2] = sg[1];
  L[3] = sg[2];
  L[4] = uf * sg[0] + vf * sg[1] + wf * sg[2] + mu * CVG * KG * g[D][3] / (PRG);
}

// ---------------------------------------------------------------------------
// Flux, residual and preconditioned stage increment of one interior cell
// (Flux.cpp:331-2398).  Grid: (16 i-planes, cubes), 256 lanes = 16 j x 16 k.
__global__ __launch_bounds__(WG) void k_flux(const double* __restrict__ U,
                                             const double* __restrict__ U1,
                                             const double* __restrict__ U1q,
                                             const int8_t* __restrict__ fws,
                                             const double* __restrict__ cube_size,
                                             double* __restrict__ Rku, double* __restrict__ Res,
                                             PhysTab pt, double deltaT, double deltaTau,
                                             double e) {
  const int c = blockIdx.y + 1;
  const int cc = pt.cube_case[c];
  if (cc < 0) return;
  const DevPhys ph = pt.phys[cc];
  const int i = blockIdx.x + 2;
  const int j = (threadIdx.x >> 4) + 2;
  const int k = (threadIdx.x & 15) + 2;
  const double dx = cube_size[c] / 16;
  const double dy = dx, dz = dx;
  const double inv = 1. / dx;
  const size_t b = sidx(c, 0, i, j, k);
  const int IF = fws[fidx(c, i, j, k)];

  double u[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) u[v] = U[b + (size_t)v * PLANE];

  // Inviscid part.
  double Rx[5], Ry[5], Rz[5], dPxb, dPxf, dPyb, dPyf, dPzb, dPzf;
  axis_flux<2>(U, fws, c, i, j, k, e, Rz, dPzb, dPzf);
  axis_flux<0>(U, fws, c, i, j, k, e, Rx, dPxb, dPxf);
  axis_flux<1>(U, fws, c, i, j, k, e, Ry, dPyb, dPyf);
  double Rf[5];
#pragma unroll
  for (int v = 0; v < 5; ++v) Rf[v] = -(Rx[v] / dx + Ry[v] / dy + Rz[v] / dz);

  // Viscous part: primitives of the 6 face and 12 edge neighbours.
  Prim q[27];
#pragma unroll
  for (int a = -1; a <= 1; ++a)
#pragma unroll
    for (int bb = -1; bb <= 1; ++bb)
#pragma unroll
      for (int cc = -1; cc <= 1; ++cc) {
        const int nz = (a != 0) + (bb != 0) + (cc != 0);
        if (nz == 1 || nz == 2) q[nslot(a, bb, cc)] = nb_prim(U, b + a * 400 + bb * 20 + cc);
      }
  Prim me;
  const double rho = u[0];
  me.u = u[1] / u[0];
  me.v = u[2] / u[0];
  me.w = u[3] / u[0];
  {
    const double VV = me.u * me.u + me.v * me.v + me.w * me.w;
    const double P = (u[4] - 0.5 * rho * VV) * KM1;
    me.T = P / rho / RG;
  }
  double Lxb[5], Lxf[5], Lyb[5], Lyf[5], Lzb[5], Lzf[5];
  visc_face<0, -1>(q, me, ph.T0, ph.mu_L, inv, Lxb);
  visc_face<0, +1>(q, me, ph.T0, ph.mu_L, inv, Lxf);
  visc_face<1, -1>(q, me, ph.T0, ph.mu_L, inv, Lyb);
  visc_face<1, +1>(q, me, ph.T0, ph.mu_L, inv, Lyf);
  visc_face<2, -1>(q, me, ph.T0, ph.mu_L, inv, Lzb);
  visc_face<2, +1>(q, me, ph.T0, ph.mu_L, inv, Lzf);
  double vF[5];
#pragma unroll
  for (int v = 1; v < 5; ++v)
    vF[v] = (Lxf[v] - Lxb[v]) * inv + (Lyf[v] - Lyb[v]) * inv + (Lzf[v] - Lzb[v]) * inv;

  // Residual with the dual-time source.
  double Rk[5];
  const DivBy d2t = div_by(2 * deltaT);
  {
    const double uu0 = U1[b], uq0 = U1q[b];
    Rk[0] = div_sh(d2t, -(3 * u[0] - 4 * uu0 + uq0)) + Rf[0];
  }
#pragma unroll
  for (int v = 1; v < 5; ++v) {
    const double uu = U1[b + (size_t)v * PLANE], uq = U1q[b + (size_t)v * PLANE];
    const double Rp = div_sh(d2t, -(3 * u[v] - 4 * uu + uq));
    Rk[v] = Rp + Rf[v] + vF[v];
  }

  // Closed-form preconditioner D = M (Gamma + 3 dtau/(2 dT) M)^-1 (Flux.cpp:2295-2343).
  const double U_ = u[1] / u[0], V_ = u[2] / u[0], W_ = u[3] / u[0];
  const double VV = U_ * U_ + V_ * V_ + W_ * W_;
  const double P = (u[4] - 0.5 * rho * VV) * KM1;
  const double C = KG * P / rho;
  const double T = P / rho;
  const double H = 0.5 * VV + C / KM1;
  const double beta = rmax(VV / C, e);
  const double temp = KG * T * (2 * deltaT + 3 * deltaTau) * (2 * deltaT + 3 * beta * deltaTau);
  const double temp2 = KM1 * (beta - 1) * deltaT * deltaT;
  const double temp3 = H - H * KG - VV + KG * (T + VV);
  double out[5];
  if (IF == 1) {
    const DivBy dT = div_by(temp);
    const double d11 = div_sh(dT, (2 * deltaT * (-2 * (-KM1 * (H - VV) - temp3 * beta) * deltaT + 3 * KG * T * beta * deltaTau)));
    const double d12 = div_sh(dT, (-4 * U_ * temp2));
    const double d13 = div_sh(dT, (-4 * V_ * temp2));
    const double d14 = div_sh(dT, (-4 * W_ * temp2));
    const double d15 = div_sh(dT, (4 * temp2));
    out[0] = deltaTau * (d11 * Rk[0] + d12 * Rk[1] + d13 * Rk[2] + d14 * Rk[3] + d15 * Rk[4]);
    const double d21 = div_sh(dT, (4 * U_ * temp3 * (beta - 1) * deltaT * deltaT));
    const double d22 = div_sh(dT, (2 * deltaT * (2 * KG * (T - U_ * U_ * (beta - 1)) * deltaT + 2 * U_ * U_ * (beta - 1) * deltaT + 3 * KG * T * beta * deltaTau)));
    const double d23 = div_sh(dT, (-4 * U_ * V_ * temp2));
    const double d24 = div_sh(dT, (-4 * U_ * W_ * temp2));
    const double d25 = div_sh(dT, (4 * U_ * temp2));
    out[1] = deltaTau * (d21 * Rk[0] + d22 * Rk[1] + d23 * Rk[2] + d24 * Rk[3] + d25 * Rk[4]);
    const double d31 = div_sh(dT, (4 * V_ * temp3 * (beta - 1) * deltaT * deltaT));
    const double d33 = div_sh(dT, (2 * deltaT * (2 * KG * (T - V_ * V_ * (beta - 1)) * deltaT + 2 * V_ * V_ * (beta - 1) * deltaT + 3 * KG * T * beta * deltaTau)));
    const double d34 = div_sh(dT, (-4 * V_ * W_ * temp2));
    const double d35 = div_sh(dT, (4 * V_ * temp2));
    out[2] = deltaTau * (d31 * Rk[0] + d23 * Rk[1] + d33 * Rk[2] + d34 * Rk[3] + d35 * Rk[4]);
    const double d41 = div_sh(dT, (4 * W_ * temp3 * (beta - 1) * deltaT * deltaT));
    const double d44 = div_sh(dT, (2 * deltaT * (2 * KG * (T - W_ * W_ * (beta - 1)) * deltaT + 2 * W_ * W_ * (beta - 1) * deltaT + 3 * KG * T * beta * deltaTau)));
    const double d45 = div_sh(dT, (4 * W_ * temp2));
    out[3] = deltaTau * (d41 * Rk[0] + d24 * Rk[1] + d34 * Rk[2] + d44 * Rk[3] + d45 * Rk[4]);
    const double d51 = div_sh(dT, (4 * H * temp3 * (beta - 1) * deltaT * deltaT));
    const double d52 = div_sh(dT, (-4 * H * U_ * temp2));
    const double d53 = div_sh(dT, (-4 * H * V_ * temp2));
    const double d54 = div_sh(dT, (-4 * H * W_ * temp2));
    const double d55 = div_sh(dT, (2 * deltaT * (2 * H * KM1 * (beta - 1) * deltaT + KG * T * (2 * deltaT + 3 * beta * deltaTau))));
    out[4] = deltaTau * (d51 * Rk[0] + d52 * Rk[1] + d53 * Rk[2] + d54 * Rk[3] + d55 * Rk[4]);
  } else {
#pragma unroll
    for (int v = 0; v < 5; ++v) out[v] = 0;
  }
#pragma unroll
  for (int v = 0; v < 5; ++v) Rku[b + (size_t)v * PLANE] = out[v];
  if (IF < 1) {
    Res[b] = -(-(dPxf - dPxb) / dx + vF[1]) * dx * dy * dz;
    Res[b + PLANE] = -(-(dPzf - dPzb) / dz + vF[3]) * dx * dy * dz;
  }
}

// Block sum of NS doubles per lane in a fixed tree order; lane 0 returns the totals.
template <int NS>
__device__ __forceinline__ void block_sum(double (&x)[NS], double* sh) {
#pragma unroll
  for (int m = 0; m < NS; ++m)
    for (int off = 32; off > 0; off >>= 1) x[m] += __shfl_down(x[m], off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int m = 0; m < NS; ++m) sh[wave * NS + m] = x[m];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
      double s = sh[m];
      for (int w = 1; w < nw; ++w) s += sh[w * NS + m];
      x[m] = s;
    }
  }
}

// ---------------------------------------------------------------------------
// Face-based formulation of the same flux (used by k_flux_march).
//
// The states at the face between cells p and p+1 depend only on the
// reconstruction order and on the six cells p-2..p+3: the reference's
// backward-face formulas of cell p+1 and forward-face formulas of cell p are
// the same expressions (Flux.cpp:499-562).  Only the order choice differs
// between the two views, so one evaluation serves both cells whenever their
// orders agree, with bitwise the same result.

// Order of face p+1/2 as cell p sees it (forward face, Flux.cpp:497,533).
__device__ __forceinline__ int order_fwd(int p, int fm1, int fp1, int fp2) {
  if (fp1 == 1 && fm1 == 1) return (p == 17 || fp2 != 1) ? 3 : 5;
  return 1;
}
// Order of face q-1/2 as cell q sees it (backward face, Flux.cpp:497,499).
__device__ __forceinline__ int order_bwd(int q, int fm2, int fm1, int fp1) {
  if (fp1 == 1 && fm1 == 1) return (q == 2 || fm2 != 1) ? 3 : 5;
  return 1;
}

// Flux through the face between p and p+1 along axis D.  b indexes cell p
// (variable 0) in an array whose variables are VS apart (PLANE in memory, one
// 20x20 plane in LDS); only the cells the order needs are loaded.
template <int D, int VS = PLANE>
__device__ __forceinline__ void face_flux(const double* __restrict__ U, size_t b, int order,
                                          double e, double (&F)[5], double& dP) {
  constexpr int st = D == 0 ? 400 : (D == 1 ? 20 : 1);
  const double c6 = 1. / 6, c60 = 1. / 60;
  double L[5], R[5];
  if (order == 5) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double* q = U + b + (size_t)v * VS;
      const double s0 = q[-2 * st], s1 = q[-st], s2 = q[0], s3 = q[st], s4 = q[2 * st], s5 = q[3 * st];
      L[v] = c60 * (2 * s0 - 13 * s1 + 47 * s2 + 27 * s3 - 3 * s4);
      R[v] = c60 * (-3 * s1 + 27 * s2 + 47 * s3 - 13 * s4 + 2 * s5);
    }
  } else if (order == 3) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double* q = U + b + (size_t)v * VS;
      const double s1 = q[-st], s2 = q[0], s3 = q[st], s4 = q[2 * st];
      L[v] = c6 * (-s1 + 5 * s2 + 2 * s3);
      R[v] = c6 * (2 * s2 + 5 * s3 - s4);
    }
  } else {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double* q = U + b + (size_t)v * VS;
      L[v] = q[0];
      R[v] = q[st];
    }
  }
  roe_flux<D>(L, R, e, F, dP);
}

// face_flux<D> on cells already loaded: s[m][v] is variable v of cell
// p-2+m.  The same expressions, so the same bits.
template <int D>
__device__ __forceinline__ void face_flux_v(const double (&s)[6][5], int order, double e,
                                            double (&F)[5], double& dP) {
  const double c6 = 1. / 6, c60 = 1. / 60;
  double L[5], R[5];
  if (order == 5) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      L[v] = c60 * (2 * s[0][v] - 13 * s[1][v] + 47 * s[2][v] + 27 * s[3][v] - 3 * s[4][v]);
      R[v] = c60 * (-3 * s[1][v] + 27 * s[2][v] + 47 * s[3][v] - 13 * s[4][v] + 2 * s[5][v]);
    }
  } else if (order == 3) {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      L[v] = c6 * (-s[1][v] + 5 * s[2][v] + 2 * s[3][v]);
      R[v] = c6 * (2 * s[2][v] + 5 * s[3][v] - s[4][v]);
    }
  } else {
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      L[v] = s[2][v];
      R[v] = s[3][v];
    }
  }
  roe_flux<D>(L, R, e, F, dP);
}

// LDS layout of k_flux_march.
constexpr int PR_PLANE = 18 * 18;   // primitives of j,k in [1,18]
constexpr int NFACE = 17 * 16;      // faces of one axis in one i-plane

// Flux, residual and stage increment with the work of one cube (or one
// segment of its i-range) per workgroup, marching along i.  Per plane:
// primitives of the next plane go into a 3-plane LDS ring (each computed once
// instead of once per neighbour); the 272 y-faces and 272 z-faces of the plane
// are evaluated once each into LDS; the x-face i+1/2 is carried in registers to
// the next plane.  A cell whose backward-face order differs from its
// neighbour's forward-face order evaluates its own backward face.  All
// arithmetic is the reference's, in its order: results are bitwise those of
// k_flux.
//
// PUSH: the update is not done in place but written to Uout, with every
// interior cell (non-fluid cells unchanged) and, for the two layers next to a
// face with an equal-size neighbour, into that neighbour's ghost layers: the
// face part of the post-update exchange.  It uses U for U1p1 (RK2) and U1p2
// (RK3): those equal U1_ bitwise at fluid cells during a stage sequence,
// since only ghost and non-fluid cells change between stages; both are still
// written as the reference writes them.
template <bool PUSH>
__global__ __launch_bounds__(WG, 2) void k_flux_march(
    double* __restrict__ U, const double* __restrict__ U1, const double* __restrict__ U1q,
    const int8_t* __restrict__ fws, const double* __restrict__ cube_size,
    double* __restrict__ Rku, double* __restrict__ Res, PhysTab pt, double deltaT,
    double deltaTau, double e, int nseg, int RK, int fuse, double* __restrict__ U1p1,
    double* __restrict__ U1p2, double* __restrict__ partials, double* __restrict__ Uout,
    const HaloSrc* __restrict__ halo, const uint8_t* __restrict__ cube_mask, int aux) {
  __shared__ double pr[4][3 * PR_PLANE];   // u, v, w, T = P/(rho R) of 3 planes
  __shared__ double fy[5][NFACE];
  __shared__ double fz[6][NFACE];          // flux + face pressure average
  __shared__ __align__(16) double cpl[5 * 400];   // conserved values of plane i
  __shared__ int8_t fpl[400];              // flags of plane i
  const int seg = blockIdx.x % nseg;
  const int c = blockIdx.x / nseg + 1;
  const int cc = pt.cube_case[c];
  if (cc < 0) return;
  const DevPhys ph = pt.phys[cc];
  const int len = 16 / nseg;
  const int i0 = 2 + seg * len, i1 = i0 + len - 1;
  // t, j and k are re-formed at the top of every plane iteration (below).
  int t = threadIdx.x;
  int j = (t >> 4) + 2, k = (t & 15) + 2;
  const double dx = cube_size[c] / 16;
  const double dy = dx, dz = dx;
  const double inv = 1. / dx;
  // x / 2^n equals x * 2^-n exactly: use the multiply when dx is a power of two.
  int ex;
  const bool pow2 = frexp(dx, &ex) == 0.5;

  // PUSH: the edge ghost cells the viscous stencil reads are not read from
  // memory (no edge pass runs between the stage's face updates and this
  // kernel) but formed as the pre-flux exchange's edge passes would leave
  // them, the mean of two face ghost cells.  In the ghost planes 1 and 18
  // they are the border cells other than the cube corners, formed here (every
  // other cell is the mean of itself and itself, which is exact).  In the
  // other planes they are the four corners, read only by the viscous stencil
  // of their own plane: they are formed from the staged plane during its face
  // pass (corner_prims below).
  // ZBC_IN_FLUX faces: the z-ghost cells (k = 0, 1 or 18, 19) of the
  // interior planes are read as the z-BC sets them, from the cell k = 2 (w
  // negated) or 17 of the same row.
  const int zm = PUSH ? cube_mask[c] : 0;
  const bool zlo = zm & ZBC_IN_FLUX, zhi = zm & (ZBC_IN_FLUX << 1);
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
  // Primitive cell n (j,k = n / 18 + 1, n % 18 + 1) of plane ip: its
  // conserved values (fill_load), then u, v, w and T into the ring slot
  // (fill_store).
  auto fill_load = [&](int ip, int n, double (&q)[5]) {
    const bool gplane = ip == 1 || ip == 18;
    const int jj = n / 18 + 1, kk = n % 18 + 1;
    if (PUSH && gplane) {
      int ea = 0, eb = 0;
      if (!edge_sources(ip, jj, kk, ea, eb)) ea = eb = ip * 400 + jj * 20 + kk;
      bool na, nb;
      ea = zsrc(ea, na);
      eb = zsrc(eb, nb);
      const size_t cb = (size_t)c * CUBE_STRIDE;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double x = U[cb + (size_t)v * PLANE + ea], y = U[cb + (size_t)v * PLANE + eb];
        q[v] = 0.5 * ((v == 3 && na ? -x : x) + (v == 3 && nb ? -y : y));
      }
    } else if (PUSH) {
      bool na;
      const int ea = zsrc(ip * 400 + jj * 20 + kk, na);
      const size_t cb = (size_t)c * CUBE_STRIDE;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double x = U[cb + (size_t)v * PLANE + ea];
        q[v] = v == 3 && na ? -x : x;
      }
    } else {
      const size_t o = sidx(c, 0, ip, jj, kk);
#pragma unroll
      for (int v = 0; v < 5; ++v) q[v] = U[o + (size_t)v * PLANE];
    }
  };
  auto fill_store = [&](int ip, int n, const double (&q)[5]) {
    const int slot = ip % 3;
    const double r = q[0];
    const DivBy dq = div_by(r);
    const double u = div_sh(dq, q[1]), v = div_sh(dq, q[2]), w = div_sh(dq, q[3]);
    const double P = (q[4] - 0.5 * r * (u * u + v * v + w * w)) * KM1;
    pr[0][slot * PR_PLANE + n] = u;
    pr[1][slot * PR_PLANE + n] = v;
    pr[2][slot * PR_PLANE + n] = w;
    pr[3][slot * PR_PLANE + n] = P / (r * RG);
  };
  // The whole 18x18 plane: the 68 cells beyond one per lane go to lanes
  // 188..255, most of which stage no part of the plane.
  auto fill_prims = [&](int ip) {
#pragma unroll 1
    for (int rd = 0; rd < 2; ++rd) {
      const int n = rd == 0 ? t : WG + t - 188;
      if (rd == 1 && t < 188) continue;
      double q[5];
      fill_load(ip, n, q);
      fill_store(ip, n, q);
    }
  };
  // PUSH: the four corner cells of interior plane ip, formed from the staged
  // plane (cpl) and written over the primitive ring's corners by lanes that
  // have one face fewer to evaluate in the face pass.
  auto corner_prims = [&](int ip) {
    const int q4 = t - (WG - 4);
    if (q4 < 0) return;
    const int jj = q4 & 2 ? 18 : 1, kk = q4 & 1 ? 18 : 1;
    int ea, eb;
    edge_sources(ip, jj, kk, ea, eb);
    ea %= 400;
    eb %= 400;
    double q[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) q[v] = 0.5 * (cpl[v * 400 + ea] + cpl[v * 400 + eb]);
    const double r = q[0];
    const DivBy dq = div_by(r);
    const double u = div_sh(dq, q[1]), v = div_sh(dq, q[2]), w = div_sh(dq, q[3]);
    const double P = (q[4] - 0.5 * r * (u * u + v * v + w * w)) * KM1;
    const int n = (jj - 1) * 18 + (kk - 1), slot = ip % 3;
    pr[0][slot * PR_PLANE + n] = u;
    pr[1][slot * PR_PLANE + n] = v;
    pr[2][slot * PR_PLANE + n] = w;
    pr[3][slot * PR_PLANE + n] = P / (r * RG);
  };
  fill_prims(i0 - 1);
  fill_prims(i0);

  double cF[5], cdP = 0;  // x-face flux carried from the previous plane
  int corder = -1;
  // With one workgroup per cube (fuse), the RK update of a plane runs three
  // planes behind the flux: no later plane's stencil reaches back that far,
  // and no other workgroup reads this cube.  RK3 norms accumulate per lane.
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  auto update_plane = [&](int ip) {
    const size_t bu = sidx(c, 0, ip, j, k);
    if (fws[fidx(c, ip, j, k)] != 1) return;
    if (RK == 1) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const size_t o = bu + (size_t)v * PLANE;
        const double x = U[o] + Rku[o];
        U[o] = x;
        U1p1[o] = x;
      }
    } else if (RK == 2) {
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const size_t o = bu + (size_t)v * PLANE;
        const double x = 0.75 * U[o] + 0.25 * U1p1[o] + 0.25 * Rku[o];
        U[o] = x;
        U1p2[o] = x;
      }
    } else {
      double uo[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) uo[v] = U[bu + (size_t)v * PLANE];
      const double ro = uo[0];
      const double uo1 = uo[1] / ro, vo1 = uo[2] / ro, wo1 = uo[3] / ro;
      const double VVo = uo1 * uo1 + vo1 * vo1 + wo1 * wo1;
      const double Po = (uo[4] - 0.5 * ro * VVo) * KM1;
      const double To = Po / ro;
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const size_t o = bu + (size_t)v * PLANE;
        uo[v] = 1.0 / 3.0 * uo[v] + 2.0 / 3.0 * U1p2[o] + 2.0 / 3.0 * Rku[o];
        U[o] = uo[v];
      }
      const double r = uo[0];
      const double un = uo[1] / r, vn = uo[2] / r, wn = uo[3] / r;
      const double VV = un * un + vn * vn + wn * wn;
      const double P = (uo[4] - 0.5 * r * VV) * KM1;
      const double T = P / r;
      acc[0] += (P - Po) * (P - Po);
      acc[1] += (un - uo1) * (un - uo1);
      acc[2] += (vn - vo1) * (vn - vo1);
      acc[3] += (wn - wo1) * (wn - wo1);
      acc[4] += (T - To) * (T - To);
    }
  };
#pragma unroll 1
  for (int i = i0; i <= i1; ++i) {
    // The lane index passes through an empty asm statement, so every
    // per-lane quantity derived from it (LDS addresses, face and flag
    // indices) is computed inside the iteration that uses it rather than
    // hoisted before the loop and held across it: held, they exceed the
    // register budget of two waves per SIMD and spill to scratch memory,
    // whose reloads sit on the plane's critical path.
    asm volatile("" : "+v"(t));
    j = (t >> 4) + 2;
    k = (t & 15) + 2;
    // Plane i-4 was last read during iteration i-1 (backward-face stencil of
    // plane i-1 reaches i-4); the barrier that ended that iteration makes its
    // update safe now.
    if (!PUSH && fuse && i - 4 >= i0) update_plane(i - 4);
    {
      // Plane i's conserved values go from memory straight into the staging
      // buffer (global_load_lds: lane l of a wave writes the 16 bytes at the
      // wave's base + 16 l), issued before the primitive fill so that the
      // fill's loads and arithmetic run while they land; a wait on the
      // vector memory counter before the barrier completes them.
      if (t < 200) {
        const size_t o = sidx(c, 0, i, 0, 0) + 2 * t;
#pragma unroll
        for (int v = 0; v < 5; ++v)
          __builtin_amdgcn_global_load_lds(
              (const __attribute__((address_space(1))) void*)(U + o + (size_t)v * PLANE),
              (__attribute__((address_space(3))) void*)(cpl + v * 400 + 2 * (t & ~63)), 16, 0, 0);
      }
      // Primitives of plane i+1.  PUSH: the plane's four corners are never
      // read from the fill (ghost planes are never the current plane, and an
      // interior plane's corners come from corner_prims), so its 320 other
      // cells fill the 256 lanes here and wave 3 in the face pass.
#pragma unroll 1
      for (int rd = 0; rd < (PUSH ? 1 : 2); ++rd) {
        const int m = rd == 0 ? t : WG + t - 188;
        if (rd == 1 && t < 188) continue;
        const int n = PUSH ? m + 1 + (m >= 16) + (m >= 304) : m;
        double q[5];
        fill_load(i + 1, n, q);
        fill_store(i + 1, n, q);
      }
      // ZBC_IN_FLUX faces: the staged plane's z-ghost cells (k = 0, 1 from
      // k = 2 with w negated; k = 18, 19 from k = 17), 4 per row of 20, are
      // loaded here and written over the copied plane once every wave's copy
      // has landed (a barrier of their own).
      double zq[5];
      const int zrow = t >> 2, zk = t & 3;
      const bool zfix = t < 80 && (zk < 2 ? zlo : zhi);
      const int zdst = zrow * 20 + (zk < 2 ? zk : zk + 16), zs = zrow * 20 + (zk < 2 ? 2 : 17);
      if (zfix) {
        const size_t o = sidx(c, 0, i, 0, 0) + zs;
#pragma unroll
        for (int v = 0; v < 5; ++v) zq[v] = U[o + (size_t)v * PLANE];
      }
      if (t < 200) {
        const int8_t* fr = fws + fidx(c, i, 0, 0) + 2 * t;
        fpl[2 * t] = fr[0];
        fpl[2 * t + 1] = fr[1];
      }
      if (zlo || zhi) {
        __builtin_amdgcn_s_waitcnt(0xF70);   // vmcnt(0)
        __syncthreads();
        if (zfix) {
#pragma unroll
          for (int v = 0; v < 5; ++v) cpl[v * 400 + zdst] = v == 3 && zk < 2 ? -zq[v] : zq[v];
        }
      }
      __builtin_amdgcn_s_waitcnt(0xF70);   // vmcnt(0)
    }
    __syncthreads();
    // y-faces (face f = (p-1)*16 + kk-2 joins j = p and p+1 at k = kk) and
    // z-faces (g = (jj-2)*17 + p-1 joins k = p and p+1 at j = jj) of plane i.
    // Round 0: each lane the y-face above its cell; round 1: the z-face
    // above it (every wave evaluates one face kind at a time); rounds 2 and
    // 3: the 16 y-faces and 16 z-faces next to the first ghost layers
    // (p = 1), on lanes of waves 0 and 1.
    if (PUSH) corner_prims(i);
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
      const bool bnd = r >= 2;
      const int l0 = r == 2 ? 0 : 64;
      if (bnd && (t < l0 || t >= l0 + 16)) continue;
      if ((r & 1) == 0) {
        const int p = bnd ? 1 : j, kk = bnd ? t - l0 + 2 : k;
        const int f = (p - 1) * 16 + (kk - 2);
        const int8_t* fb = fpl + p * 20 + kk;
        const int o = p >= 2 ? order_fwd(p, fb[-20], fb[20], fb[40]) : order_bwd(2, fb[-20], fb[0], fb[40]);
        double F[5], dP;
        face_flux<1, 400>(cpl, p * 20 + kk, o, e, F, dP);
#pragma unroll
        for (int v = 0; v < 5; ++v) fy[v][f] = F[v];
      } else {
        const int jj = bnd ? t - l0 + 2 : j, p = bnd ? 1 : k;
        const int g = (jj - 2) * 17 + (p - 1);
        const int8_t* fb = fpl + jj * 20 + p;
        const int o = p >= 2 ? order_fwd(p, fb[-1], fb[1], fb[2]) : order_bwd(2, fb[-1], fb[0], fb[2]);
        double F[5], dP;
        face_flux<2, 400>(cpl, jj * 20 + p, o, e, F, dP);
#pragma unroll
        for (int v = 0; v < 5; ++v) fz[v][g] = F[v];
        fz[5][g] = dP;
      }
    }
    // PUSH: the fill's second round (wave 3, which has no boundary faces)
    // runs in the face pass; plane i+1's primitives are first read after
    // the next barrier, and its ring slot was last read before the previous
    // one.
    if (PUSH && t >= 192) {
      const int m = WG + t - 192;
      const int n = m + 1 + (m >= 16) + (m >= 304);
      double q[5];
      fill_load(i + 1, n, q);
      fill_store(i + 1, n, q);
    }
    // The x-stencil of the cell's forward face, planes i-2..i+3, and the
    // flags it needs, loaded before the barrier so that they arrive while the
    // workgroup gathers.  (Plane i+3 = 20 would lie beyond the cube; the last
    // interior plane's forward face has order 3 or 1, which does not read
    // it, so plane 19 stands in.)
    const int8_t* fc = fws + fidx(c, i, j, k);
    const int IF = fc[0];
    const int fm2 = fc[-800], fm1 = fc[-400], fp1 = fc[400], fp2 = fc[800];
    double xs[6][5];
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
      for (int v = 0; v < 5; ++v) xs[m][v] = U[sidx(c, v, min(i - 2 + m, 19), j, k)];
    __syncthreads();

    const size_t b = sidx(c, 0, i, j, k);
    // ---- inviscid, x: forward face computed, backward face carried ----
    double Rx[5], Ry[5], Rz[5], dPxb, dPxf, dPzb, dPzf;
    {
      const int of = order_fwd(i, fm1, fp1, fp2);
      double Ff[5];
      face_flux_v<0>(xs, of, e, Ff, dPxf);
      const int ob = order_bwd(i, fm2, fm1, fp1);
      if (i != i0 && ob == corder) {
        dPxb = cdP;
#pragma unroll
        for (int v = 0; v < 5; ++v) Rx[v] = Ff[v] - cF[v];
      } else {
        double Fb[5];
        face_flux<0>(U, b - 400, ob, e, Fb, dPxb);
#pragma unroll
        for (int v = 0; v < 5; ++v) Rx[v] = Ff[v] - Fb[v];
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) cF[v] = Ff[v];
      cdP = dPxf;
      corder = of;
    }
    // The cell's time levels, loaded here so that the y-, z- and viscous
    // terms run while they arrive.
    double tu1[5], tuq[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      tu1[v] = U1[b + (size_t)v * PLANE];
      tuq[v] = U1q[b + (size_t)v * PLANE];
    }
    // ---- inviscid, y ----
    {
      const int8_t* fl = fpl + j * 20 + k;
      const int fm2 = fl[-40], fm1 = fl[-20], fp1 = fl[20];
      const int ff = (j - 1) * 16 + (k - 2), fbk = ff - 16;
      const int ob = order_bwd(j, fm2, fm1, fp1);
      const bool share = j == 2 || ob == order_fwd(j - 1, fm2, fl[0], fp1);
      if (share) {
#pragma unroll
        for (int v = 0; v < 5; ++v) Ry[v] = fy[v][ff] - fy[v][fbk];
      } else {
        double Fb[5], dPu;
        face_flux<1, 400>(cpl, j * 20 + k - 20, ob, e, Fb, dPu);
#pragma unroll
        for (int v = 0; v < 5; ++v) Ry[v] = fy[v][ff] - Fb[v];
      }
    }
    // ---- inviscid, z ----
    {
      const int8_t* fl = fpl + j * 20 + k;
      const int fm2 = fl[-2], fm1 = fl[-1], fp1 = fl[1];
      const int ff = (j - 2) * 17 + (k - 1), fbk = ff - 1;
      const int ob = order_bwd(k, fm2, fm1, fp1);
      const bool share = k == 2 || ob == order_fwd(k - 1, fm2, fl[0], fp1);
      dPzf = fz[5][ff];
      if (share) {
        dPzb = fz[5][fbk];
#pragma unroll
        for (int v = 0; v < 5; ++v) Rz[v] = fz[v][ff] - fz[v][fbk];
      } else {
        double Fb[5];
        face_flux<2, 400>(cpl, j * 20 + k - 1, ob, e, Fb, dPzb);
#pragma unroll
        for (int v = 0; v < 5; ++v) Rz[v] = fz[v][ff] - Fb[v];
      }
    }
    double Rf[5];
    if (pow2) {
#pragma unroll
      for (int v = 0; v < 5; ++v) Rf[v] = -(Rx[v] * inv + Ry[v] * inv + Rz[v] * inv);
    } else {
#pragma unroll
      for (int v = 0; v < 5; ++v) Rf[v] = -(Rx[v] / dx + Ry[v] / dy + Rz[v] / dz);
    }

    // ---- viscous (Flux.cpp:1722-2279) from the primitive ring ----
    // The cell's conserved values from the staged plane (interior cells are
    // staged as they are).
    double u[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v] = cpl[v * 400 + j * 20 + k];
    const double rho = u[0];
    auto Q = [&](int a, int bb, int cc, int m) -> double {
      return pr[m][((i + a) % 3) * PR_PLANE + (j + bb - 1) * 18 + (k + cc - 1)];
    };
    Prim me;
    me.u = Q(0, 0, 0, 0);
    me.v = Q(0, 0, 0, 1);
    me.w = Q(0, 0, 0, 2);
    {
      const double VV = me.u * me.u + me.v * me.v + me.w * me.w;
      const double P = (u[4] - 0.5 * rho * VV) * KM1;
      me.T = P / rho / RG;
    }
    double vF[5] = {0, 0, 0, 0, 0};
    double Lb[5], Lf[5];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
#pragma unroll
      for (int side = -1; side <= 1; side += 2) {
        const int o0 = d == 0 ? side : 0, o1 = d == 1 ? side : 0, o2 = d == 2 ? side : 0;
        Prim f;
        f.u = Q(o0, o1, o2, 0);
        f.v = Q(o0, o1, o2, 1);
        f.w = Q(o0, o1, o2, 2);
        f.T = Q(o0, o1, o2, 3);
        const double Tf = 0.5 * (me.T + f.T);
        const double temp = Tf / ph.T0;
        const double mu = ph.mu_L * temp * sqrt(temp) * (ph.T0 + 110.) / (Tf + 110.);
        double g[3][4];
#pragma unroll
        for (int tt = 0; tt < 3; ++tt) {
          if (tt == d) {
            if (side < 0) {
              g[tt][0] = (me.u - f.u) * inv;
              g[tt][1] = (me.v - f.v) * inv;
              g[tt][2] = (me.w - f.w) * inv;
              g[tt][3] = (me.T - f.T) * inv;
            } else {
              g[tt][0] = (f.u - me.u) * inv;
              g[tt][1] = (f.v - me.v) * inv;
              g[tt][2] = (f.w - me.w) * inv;
              g[tt][3] = (f.T - me.T) * inv;
            }
          } else {
            const int p0 = tt == 0, p1 = tt == 1, p2 = tt == 2;
#pragma unroll
            for (int m = 0; m < 4; ++m)
              g[tt][m] = (Q(p0, p1, p2, m) + Q(p0 + o0, p1 + o1, p2 + o2, m) - Q(-p0, -p1, -p2, m) -
                          Q(-p0 + o0, -p1 + o1, -p2 + o2, m)) / 4 * inv;
          }
        }
        const double uf = 0.5 * (me.u + f.u), vf = 0.5 * (me.v + f.v), wf = 0.5 * (me.w + f.w);
        const double div = g[0][0] + g[1][1] + g[2][2];
        double sg[3];
#pragma unroll
        for (int m = 0; m < 3; ++m)
          sg[m] = (m == d) ? mu * (2 * g[d][d] - 2.0 / 3.0 * div) : mu * (g[m][d] + g[d][m]);
        double* L = side < 0 ? Lb : Lf;
        L[1] = sg[0];
        L[2] = sg[1];
        L[3] = sg[2];
        L[4] = uf * sg[0] + vf * sg[1] + wf * sg[2] + mu * CVG * KG * g[d][3] / (PRG);
      }
#pragma unroll
      for (int v = 1; v < 5; ++v) {
        const double term = (Lf[v] - Lb[v]) * inv;
        vF[v] = d == 0 ? term : vF[v] + term;
      }
    }

    // ---- residual with the dual-time source, preconditioner, increment ----
    double Rk[5];
    const DivBy d2t = div_by(2 * deltaT);
    {
      const double uu0 = tu1[0], uq0 = tuq[0];
      Rk[0] = div_sh(d2t, -(3 * u[0] - 4 * uu0 + uq0)) + Rf[0];
    }
#pragma unroll
    for (int v = 1; v < 5; ++v) {
      const double uu = tu1[v], uq = tuq[v];
      const double Rp = div_sh(d2t, -(3 * u[v] - 4 * uu + uq));
      Rk[v] = Rp + Rf[v] + vF[v];
    }
    double out[5] = {0, 0, 0, 0, 0};
    if (IF == 1) {
      const double U_ = me.u, V_ = me.v, W_ = me.w;
      const double VV = U_ * U_ + V_ * V_ + W_ * W_;
      const double P = (u[4] - 0.5 * rho * VV) * KM1;
      const double C = KG * P / rho;
      const double T = P / rho;
      const double H = 0.5 * VV + C / KM1;
      const double beta = rmax(VV / C, e);
      const double temp = KG * T * (2 * deltaT + 3 * deltaTau) * (2 * deltaT + 3 * beta * deltaTau);
      const double temp2 = KM1 * (beta - 1) * deltaT * deltaT;
      const double temp3 = H - H * KG - VV + KG * (T + VV);
      const DivBy dT = div_by(temp);
      const double d11 = div_sh(dT, (2 * deltaT * (-2 * (-KM1 * (H - VV) - temp3 * beta) * deltaT + 3 * KG * T * beta * deltaTau)));
      const double d12 = div_sh(dT, (-4 * U_ * temp2));
      const double d13 = div_sh(dT, (-4 * V_ * temp2));
      const double d14 = div_sh(dT, (-4 * W_ * temp2));
      const double d15 = div_sh(dT, (4 * temp2));
      out[0] = deltaTau * (d11 * Rk[0] + d12 * Rk[1] + d13 * Rk[2] + d14 * Rk[3] + d15 * Rk[4]);
      const double d21 = div_sh(dT, (4 * U_ * temp3 * (beta - 1) * deltaT * deltaT));
      const double d22 = div_sh(dT, (2 * deltaT * (2 * KG * (T - U_ * U_ * (beta - 1)) * deltaT + 2 * U_ * U_ * (beta - 1) * deltaT + 3 * KG * T * beta * deltaTau)));
      const double d23 = div_sh(dT, (-4 * U_ * V_ * temp2));
      const double d24 = div_sh(dT, (-4 * U_ * W_ * temp2));
      const double d25 = div_sh(dT, (4 * U_ * temp2));
      out[1] = deltaTau * (d21 * Rk[0] + d22 * Rk[1] + d23 * Rk[2] + d24 * Rk[3] + d25 * Rk[4]);
      const double d31 = div_sh(dT, (4 * V_ * temp3 * (beta - 1) * deltaT * deltaT));
      const double d33 = div_sh(dT, (2 * deltaT * (2 * KG * (T - V_ * V_ * (beta - 1)) * deltaT + 2 * V_ * V_ * (beta - 1) * deltaT + 3 * KG * T * beta * deltaTau)));
      const double d34 = div_sh(dT, (-4 * V_ * W_ * temp2));
      const double d35 = div_sh(dT, (4 * V_ * temp2));
      out[2] = deltaTau * (d31 * Rk[0] + d23 * Rk[1] + d33 * Rk[2] + d34 * Rk[3] + d35 * Rk[4]);
      const double d41 = div_sh(dT, (4 * W_ * temp3 * (beta - 1) * deltaT * deltaT));
      const double d44 = div_sh(dT, (2 * deltaT * (2 * KG * (T - W_ * W_ * (beta - 1)) * deltaT + 2 * W_ * W_ * (beta - 1) * deltaT + 3 * KG * T * beta * deltaTau)));
      const double d45 = div_sh(dT, (4 * W_ * temp2));
      out[3] = deltaTau * (d41 * Rk[0] + d24 * Rk[1] + d34 * Rk[2] + d44 * Rk[3] + d45 * Rk[4]);
      const double d51 = div_sh(dT, (4 * H * temp3 * (beta - 1) * deltaT * deltaT));
      const double d52 = div_sh(dT, (-4 * H * U_ * temp2));
      const double d53 = div_sh(dT, (-4 * H * V_ * temp2));
      const double d54 = div_sh(dT, (-4 * H * W_ * temp2));
      const double d55 = div_sh(dT, (2 * deltaT * (2 * H * KM1 * (beta - 1) * deltaT + KG * T * (2 * deltaT + 3 * beta * deltaTau))));
      out[4] = deltaTau * (d51 * Rk[0] + d52 * Rk[1] + d53 * Rk[2] + d54 * Rk[3] + d55 * Rk[4]);
    }
    // The pushing kernel reads none of Rku1, U1p1, U1p2 and Residual1 back;
    // it stores them only when asked (aux), which the engine does in the
    // last inner iteration of a step, so the arrays end as the reference's.
    if (!PUSH || aux) {
#pragma unroll
      for (int v = 0; v < 5; ++v) Rku[b + (size_t)v * PLANE] = out[v];
    }
    if (IF < 1) {
      const double px = pow2 ? -(dPxf - dPxb) * inv : -(dPxf - dPxb) / dx;
      const double pz = pow2 ? -(dPzf - dPzb) * inv : -(dPzf - dPzb) / dz;
      const double r0 = -(px + vF[1]) * dx * dy * dz;
      const double r1 = -(pz + vF[3]) * dx * dy * dz;
      if (!PUSH || aux) {
        Res[b] = r0;
        Res[b + PLANE] = r1;
      }
      if ((fuse || PUSH) && RK == 3) {
        acc[5] += r0;
        acc[6] += r1;
      }
    }
    if constexpr (PUSH) {
      // The RK update of update_plane with U1p1 = U1p2 = U (see above).
      double nv[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) nv[v] = u[v];
      if (IF == 1) {
        if (RK == 1) {
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            const double x = u[v] + out[v];
            nv[v] = x;
            if (aux) U1p1[b + (size_t)v * PLANE] = x;
          }
        } else if (RK == 2) {
#pragma unroll
          for (int v = 0; v < 5; ++v) {
            const double x = 0.75 * u[v] + 0.25 * u[v] + 0.25 * out[v];
            nv[v] = x;
            if (aux) U1p2[b + (size_t)v * PLANE] = x;
          }
        } else {
          const double ro = u[0];
          const double uo1 = u[1] / ro, vo1 = u[2] / ro, wo1 = u[3] / ro;
          const double VVo = uo1 * uo1 + vo1 * vo1 + wo1 * wo1;
          const double Po = (u[4] - 0.5 * ro * VVo) * KM1;
          const double To = Po / ro;
#pragma unroll
          for (int v = 0; v < 5; ++v) nv[v] = 1.0 / 3.0 * u[v] + 2.0 / 3.0 * u[v] + 2.0 / 3.0 * out[v];
          const double r = nv[0];
          const double un = nv[1] / r, vn = nv[2] / r, wn = nv[3] / r;
          const double VV = un * un + vn * vn + wn * wn;
          const double P = (nv[4] - 0.5 * r * VV) * KM1;
          const double T = P / r;
          acc[0] += (P - Po) * (P - Po);
          acc[1] += (un - uo1) * (un - uo1);
          acc[2] += (vn - vo1) * (vn - vo1);
          acc[3] += (wn - wo1) * (wn - wo1);
          acc[4] += (T - To) * (T - To);
        }
      }
      const HaloSrc hs = halo[c];
      const int o = i * 400 + j * 20 + k;
      const int nx = i <= 3 ? hs.src[0] : (i >= 16 ? hs.src[1] : 0);
      const int ny = j <= 3 ? hs.src[2] : (j >= 16 ? hs.src[3] : 0);
      const int nz = k <= 3 ? hs.src[4] : (k >= 16 ? hs.src[5] : 0);
      const int ox = o + (i <= 3 ? 16 * 400 : -16 * 400);
      const int oy = o + (j <= 3 ? 16 * 20 : -16 * 20);
      const int oz = o + (k <= 3 ? 16 : -16);
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        Uout[(size_t)c * CUBE_STRIDE + (size_t)v * PLANE + o] = nv[v];
        if (nx > 0) Uout[(size_t)nx * CUBE_STRIDE + (size_t)v * PLANE + ox] = nv[v];
        if (ny > 0) Uout[(size_t)ny * CUBE_STRIDE + (size_t)v * PLANE + oy] = nv[v];
        if (nz > 0) Uout[(size_t)nz * CUBE_STRIDE + (size_t)v * PLANE + oz] = nv[v];
      }
    }
    __syncthreads();
  }
  if constexpr (PUSH) {
    const int bm = cube_mask[c];
    if (bm) copy_bc_regions(U, Uout, c, bm, 0.0);
  }
  if (fuse || PUSH) {
    if (!PUSH)
      for (int ip = i1 - 3 < i0 ? i0 : i1 - 3; ip <= i1; ++ip) update_plane(ip);
    if (RK == 3) {
      __shared__ double sh[(WG / 64) * 7];
      block_sum<7>(acc, sh);
      if (threadIdx.x == 0)
#pragma unroll
        for (int m = 0; m < 7; ++m) partials[(size_t)blockIdx.x * 8 + m] = acc[m];
    }
  }
}

// RK stage update of fluid cells (Flux.cpp:2417-2599).  At RK3 each block
// also writes its partial sums of the five change norms (:2530-2566) and of
// the two force-density components (:2603-2625).
__global__ __launch_bounds__(WG) void k_update(int RK, double* __restrict__ U,
                                               double* __restrict__ U1p1,
                                               double* __restrict__ U1p2,
                                               const double* __restrict__ Rku,
                                               const double* __restrict__ Res,
                                               const int8_t* __restrict__ fws,
                                               const int* __restrict__ cube_case,
                                               double* __restrict__ partials) {
  const int c = blockIdx.y + 1;
  if (cube_case[c] < 0) return;
  const int i = blockIdx.x + 2;
  const int j = (threadIdx.x >> 4) + 2;
  const int k = (threadIdx.x & 15) + 2;
  const size_t b = sidx(c, 0, i, j, k);
  const bool fluid = fws[fidx(c, i, j, k)] == 1;
  if (RK == 1) {
    if (fluid)
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const double x = U[b + (size_t)v * PLANE] + Rku[b + (size_t)v * PLANE];
        U[b + (size_t)v * PLANE] = x;
        U1p1[b + (size_t)v * PLANE] = x;
      }
    return;
  }
  if (RK == 2) {
    if (fluid)
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        const size_t o = b + (size_t)v * PLANE;
        const double x = 0.75 * U[o] + 0.25 * U1p1[o] + 0.25 * Rku[o];
        U[o] = x;
        U1p2[o] = x;
      }
    return;
  }
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  if (fluid) {
    double u[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) u[v] = U[b + (size_t)v * PLANE];
    const double ro = u[0];
    const double uo = u[1] / ro, vo = u[2] / ro, wo = u[3] / ro;
    const double VVo = uo * uo + vo * vo + wo * wo;
    const double Po = (u[4] - 0.5 * ro * VVo) * KM1;
    const double To = Po / ro;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const size_t o = b + (size_t)v * PLANE;
      u[v] = 1.0 / 3.0 * u[v] + 2.0 / 3.0 * U1p2[o] + 2.0 / 3.0 * Rku[o];
      U[o] = u[v];
    }
    const double r = u[0];
    const double un = u[1] / r, vn = u[2] / r, wn = u[3] / r;
    const double VV = un * un + vn * vn + wn * wn;
    const double P = (u[4] - 0.5 * r * VV) * KM1;
    const double T = P / r;
    acc[0] = (P - Po) * (P - Po);
    acc[1] = (un - uo) * (un - uo);
    acc[2] = (vn - vo) * (vn - vo);
    acc[3] = (wn - wo) * (wn - wo);
    acc[4] = (T - To) * (T - To);
  }
  acc[5] = Res[b];
  acc[6] = Res[b + PLANE];
  __shared__ double sh[(WG / 64) * 7];
  block_sum<7>(acc, sh);
  if (threadIdx.x == 0) {
    const size_t blk = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
    for (int m = 0; m < 7; ++m) partials[blk * 8 + m] = acc[m];
  }
}

// With several partial rows per cube, each cube's rows are first summed in
// row order into its first row, one thread per cube (blockIdx.y: case).
__global__ __launch_bounds__(WG) void k_fold_rows(double* __restrict__ partials,
                                                  int rows_per_cube,
                                                  const CaseDesc* __restrict__ cases) {
  const CaseDesc cd = cases[blockIdx.y];
  for (int q = blockIdx.x * WG + threadIdx.x; q < cd.n; q += gridDim.x * WG) {
    double* r = partials + (size_t)(cd.first - 1 + q) * rows_per_cube * 8;
    double acc[7];
#pragma unroll
    for (int m = 0; m < 7; ++m) acc[m] = r[m];
    for (int b = 1; b < rows_per_cube; ++b)
#pragma unroll
      for (int m = 0; m < 7; ++m) acc[m] += r[b * 8 + m];
#pragma unroll
    for (int m = 0; m < 7; ++m) r[m] = acc[m];
  }
}

// Final norm reduction (Flux.cpp:2627-2652 at one rank): er[1..7] of one
// case per block from one partial row per cube (every stride-th row).  The
// rows of a case are summed in an order fixed relative to the case's first
// cube, so a case gives the same bits alone and next to other cases.
__global__ __launch_bounds__(WG) void k_reduce(const double* __restrict__ partials,
                                               int stride,
                                               const CaseDesc* __restrict__ cases,
                                               const DevPhys* __restrict__ phys,
                                               double* __restrict__ er_all) {
  const CaseDesc cd = cases[blockIdx.x];
  const DevPhys ph = phys[cd.slot];
  const double* rows = partials + (size_t)(cd.first - 1) * stride * 8;
  const int nrows = cd.n;
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int bi = threadIdx.x; bi < nrows; bi += WG)
#pragma unroll
    for (int m = 0; m < 7; ++m) acc[m] += rows[(size_t)bi * stride * 8 + m];
  __shared__ double sh[(WG / 64) * 7];
  block_sum<7>(acc, sh);
  if (threadIdx.x == 0) {
    double* er = er_all + (size_t)cd.slot * 8;
    const double DN = 1. / (cd.mpi_ncube * 16 * 16 * 16);
    er[0] = 0;
    for (int m = 0; m < 5; ++m) er[m + 1] = sqrt(acc[m]) * DN;
    er[6] = acc[5] / (0.5 * ph.rho0 * ph.U0 * ph.U0 * 2.602);
    er[7] = acc[6] / (0.5 * ph.rho0 * ph.U0 * ph.U0 * 2.602);
  }
}

// ---------------------------------------------------------------------------
// Domain boundary conditions of one cube per block: X (BCM_X_boundary_condition
// .cpp:43-118), then Y (BCM_Y_boundary_condition.cpp:36-80), then Z
// (BCM_Z_boundary_condition.cpp:42-92).  mask bits: 1 x-min, 2 x-max, 4 y-min,
// 8 y-max, 16 z-min, 32 z-max.
// In each of the three phases the cells read (the layers next to the ghost
// layers) and the cells written (the ghost layers) are disjoint, so every
// lane issues all its loads of a phase before its first store: one memory
// round trip per phase instead of one per value.
template <typename T>
__device__ __forceinline__ void bc_cube(T* __restrict__ U, int c, int m, PhysTab pt) {
  const DevPhys ph = pt.phys[pt.cube_case[c]];
  const typename OffOf<T>::type F(ph);
  const int t = threadIdx.x;
  {
    const int j = (t >> 4) + 2, k = (t & 15) + 2;
    T lo[5], hi[5];
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      if (m & 1) lo[v] = U[sidx(c, v, 2, j, k)];
      if (m & 2) hi[v] = U[sidx(c, v, 17, j, k)];
    }
    if (m & 1) {
      const double rho = F.ld(lo, 0, 0);
      const double uu = F.ld(lo, 1, 1) / rho;
      const double vv = F.ld(lo, 2, 2) / rho;
      const double ww = F.ld(lo, 3, 3) / rho;
      double VV = uu * uu + vv * vv + ww * ww;
      const double P = (F.ld(lo, 4, 4) - 0.5 * rho * VV) * KM1;
      VV = ph.Uin * ph.Uin + ph.V0 * ph.V0 + ph.W0 * ph.W0;
      const double E = P / KM1 + 0.5 * rho * VV;
      for (int i = 0; i < 2; ++i) {
        U[sidx(c, 0, i, j, k)] = lo[0];
        F.st(U, sidx(c, 1, i, j, k), 1, rho * ph.Uin);
        F.st(U, sidx(c, 2, i, j, k), 2, rho * ph.V0);
        F.st(U, sidx(c, 3, i, j, k), 3, rho * ph.W0);
        F.st(U, sidx(c, 4, i, j, k), 4, E);
      }
    }
    if (m & 2)
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        U[sidx(c, v, 18, j, k)] = hi[v];
        U[sidx(c, v, 19, j, k)] = hi[v];
      }
  }
  __syncthreads();
  {
    // Rows (i, k) of the y-faces, i = 0..19: w = t and t + 256 (w < 320).
    T lo[2][5], hi[2][5];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int w = t + h * WG, i = w >> 4, k = (w & 15) + 2;
      if (w < 20 * 16)
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          if (m & 4) lo[h][v] = U[sidx(c, v, i, 2, k)];
          if (m & 8) hi[h][v] = U[sidx(c, v, i, 17, k)];
        }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int w = t + h * WG, i = w >> 4, k = (w & 15) + 2;
      if (w < 20 * 16)
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          if (m & 4) {
            U[sidx(c, v, i, 1, k)] = lo[h][v];
            U[sidx(c, v, i, 0, k)] = lo[h][v];
          }
          if (m & 8) {
            U[sidx(c, v, i, 18, k)] = hi[h][v];
            U[sidx(c, v, i, 19, k)] = hi[h][v];
          }
        }
    }
  }
  __syncthreads();
  {
    // Rows (i, j) of the z-faces: w = t and t + 256 (w < 400).  Interior
    // planes of a face marked ZBC_IN_FLUX get their z-BC from the flux
    // kernel.
    T lo[2][5], hi[2][5];
    bool dlo[2], dhi[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int w = t + h * WG, i = w / 20, j = w % 20;
      const bool inner = i >= 2 && i <= 17;
      dlo[h] = w < 400 && (m & 16) && !(inner && (m & ZBC_IN_FLUX));
      dhi[h] = w < 400 && (m & 32) && !(inner && (m & (ZBC_IN_FLUX << 1)));
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        if (dlo[h]) lo[h][v] = U[sidx(c, v, i, j, 2)];
        if (dhi[h]) hi[h][v] = U[sidx(c, v, i, j, 17)];
      }
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int w = t + h * WG, i = w / 20, j = w % 20;
      if (dlo[h])
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          if (v == 3) {
            const double y = -F.ld(lo[h], 3, 3);
            F.st(U, sidx(c, v, i, j, 1), 3, y);
            F.st(U, sidx(c, v, i, j, 0), 3, y);
          } else {
            U[sidx(c, v, i, j, 1)] = lo[h][v];
            U[sidx(c, v, i, j, 0)] = lo[h][v];
          }
        }
      if (dhi[h])
#pragma unroll
        for (int v = 0; v < 5; ++v) {
          U[sidx(c, v, i, j, 18)] = hi[h][v];
          U[sidx(c, v, i, j, 19)] = hi[h][v];
        }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(WG) void k_bc(T* __restrict__ U, const int* __restrict__ cubes,
                                           const uint8_t* __restrict__ masks, PhysTab pt) {
  bc_cube(U, cubes[blockIdx.x], masks[blockIdx.x], pt);
}

// ---------------------------------------------------------------------------
// Immersed-boundary ghost fill (BCM_Ghostcell_minus.cpp:72-489).  The
// reference walks the minus list then the plus list in order, writing in
// place; entries here come grouped per cube into dependency levels built on
// the host (an entry's level is above every earlier entry it reads from or
// whose corners it overwrites), so all entries of one level run in parallel
// and levels are separated by workgroup barriers.  Interpolation corners
// always lie in the ghost cell's own padded cube.  The cells a cube's entries
// touch are staged in LDS as FP64 conserved values (full states), every level
// runs out of LDS with one value slot per entry (see GhostTables), and the
// last value of each written ghost cell goes back to memory once; a cube
// needing more than lds_cap slots works on memory in place.  A value written
// to LDS is rounded as storage would round it, so both ways give the same
// results.

// IB fill of ghost-table cube g by the calling workgroup; gsm holds
// [5][lds_cap] doubles.
// The fields of a GhostEntry the fill reads, loaded one by one so that a copy
// held across a level stays in registers.
struct GhostEntryR {
  int16_t gi, gj, gk, pi, pj, pk, lg;
  int16_t lc[8];
  double w[8];
};
__device__ __forceinline__ GhostEntryR load_entry(const GhostEntry* p) {
  GhostEntryR r;
  r.gi = p->gi;
  r.gj = p->gj;
  r.gk = p->gk;
  r.pi = p->pi;
  r.pj = p->pj;
  r.pk = p->pk;
  r.lg = p->lg;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    r.lc[m] = p->lc[m];
    r.w[m] = p->w[m];
  }
  return r;
}

template <typename T>
__device__ __forceinline__ void ghost_cube(T* __restrict__ U, const GhostTables& gt, int g,
                                           int lds_cap, PhysTab pt, double* gsm) {
  const int l0 = gt.cube_level_off[g], l1 = gt.cube_level_off[g + 1];
  if (l0 == l1) return;
  const int c = gt.ent[gt.level_off[l0]].cube;
  const DevPhys ph = pt.phys[pt.cube_case[c]];
  const typename OffOf<T>::type F(ph);
  const double rho0 = ph.rho0, P0 = ph.P0;
  const int t0 = gt.touched_off[g];
  const int nt = gt.touched_off[g + 1] - t0;
  const int ne = gt.level_off[l1] - gt.level_off[l0];
  const bool lds = nt + ne <= lds_cap;
  if (lds) {
    // Up to six cells per lane: the loads of two cells are issued together.
    for (int n = threadIdx.x; n < nt; n += 2 * WG) {
      const int n2 = n + WG;
      const size_t o = (size_t)c * CUBE_STRIDE + gt.touched[t0 + n];
      const size_t o2 = (size_t)c * CUBE_STRIDE + gt.touched[t0 + (n2 < nt ? n2 : n)];
      double a[5], a2[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) {
        a[v] = F.ld(U, o + (size_t)v * PLANE, v);
        a2[v] = F.ld(U, o2 + (size_t)v * PLANE, v);
      }
#pragma unroll
      for (int v = 0; v < 5; ++v) gsm[v * lds_cap + n] = a[v];
      if (n2 < nt) {
#pragma unroll
        for (int v = 0; v < 5; ++v) gsm[v * lds_cap + n2] = a2[v];
      }
    }
    __syncthreads();
  }
  // Density, velocity and pressure of one interpolation corner.
  auto corner_prim = [&](const double (&q)[5], double& r, double& u, double& v, double& w,
                         double& p) {
    r = q[0];
    if constexpr (OffOf<T>::type::exact) {
      u = q[1] / r;
      v = q[2] / r;
      w = q[3] / r;
    } else {
      // One reciprocal per corner instead of three divisions.
      const double ir = 1.0 / r;
      u = q[1] * ir;
      v = q[2] * ir;
      w = q[3] * ir;
    }
    const double VV = u * u + v * v + w * w;
    p = (q[4] - 0.5 * r * VV) * KM1;
  };
  // One entry from its corners' primitives: into its slot (LDS) or its ghost
  // cell (memory).
  auto combine = [&](const GhostEntryR& E, const double (&rc)[8], const double (&uc)[8],
                     const double (&vc)[8], const double (&wc)[8], const double (&pc)[8]) {
    double rho = E.w[0] * (rc[0] - rho0), P = E.w[0] * (pc[0] - P0);
    double uu = E.w[0] * uc[0], vv = E.w[0] * vc[0], ww = E.w[0] * wc[0];
#pragma unroll
    for (int m = 1; m < 8; ++m) {
      rho = rho + E.w[m] * (rc[m] - rho0);
      P = P + E.w[m] * (pc[m] - P0);
      uu = uu + E.w[m] * uc[m];
      vv = vv + E.w[m] * vc[m];
      ww = ww + E.w[m] * wc[m];
    }
    rho = rho + rho0;
    P = P + P0;
    double VV = uu * uu + vv * vv + ww * ww;
    double o0 = rho, o1 = 0.5 * rho * uu, o2 = 0.5 * rho * vv, o3 = 0.5 * rho * ww;
    double o4 = P / KM1 + 0.5 * rho * 0.25 * VV;
    // The ghost cell may itself be an interpolation corner (:192-449).
    const int di = E.gi - E.pi, dj = E.gj - E.pj, dk = E.gk - E.pk;
    if (di >= 0 && di <= 1 && dj >= 0 && dj <= 1 && dk >= 0 && dk <= 1) {
      const int s = di && dj && dk ? 7
                    : (dj && dk ? 6 : (di && dk ? 5 : (di && dj ? 4 : (dk ? 3 : (dj ? 2 : (di ? 1 : 0))))));
      // Corner s's weight and values, selected without indexing by s (which
      // would put the arrays in memory).
      double ws = 0, rs = 0, ps = 0;
#pragma unroll
      for (int m = 0; m < 8; ++m)
        if (m == s) {
          ws = E.w[m];
          rs = rc[m];
          ps = pc[m];
        }
      if (ws > 0.9999) {
        rho = rs;
        P = ps;
      } else {
        double sr = 0, sp = 0;
        bool first = true;
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          if (m == s) continue;
          if (first) {
            sr = E.w[m] * (rc[m] - rho0);
            sp = E.w[m] * (pc[m] - P0);
            first = false;
          } else {
            sr = sr + E.w[m] * (rc[m] - rho0);
            sp = sp + E.w[m] * (pc[m] - P0);
          }
        }
        rho = sr / (1 - ws) + rho0;
        P = sp / (1 - ws) + P0;
      }
      double su = 0, sv = 0, sw = 0;
      bool first = true;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        if (m == s) continue;
        if (first) {
          su = E.w[m] * uc[m];
          sv = E.w[m] * vc[m];
          sw = E.w[m] * wc[m];
          first = false;
        } else {
          su = su + E.w[m] * uc[m];
          sv = sv + E.w[m] * vc[m];
          sw = sw + E.w[m] * wc[m];
        }
      }
      uu = su / (2 - ws);
      vv = sv / (2 - ws);
      ww = sw / (2 - ws);
      VV = uu * uu + vv * vv + ww * ww;
      o0 = rho;
      o1 = rho * uu;
      o2 = rho * vv;
      o3 = rho * ww;
      o4 = P / KM1 + 0.5 * rho * VV;
    }
    const double out[5] = {o0, o1, o2, o3, o4};
    if (lds) {
#pragma unroll
      for (int v = 0; v < 5; ++v) gsm[v * lds_cap + E.lg] = F.rt(out[v], v);
    } else {
      const size_t go = sidx(c, 0, E.gi, E.gj, E.gk);
#pragma unroll
      for (int v = 0; v < 5; ++v) F.st(U, go + (size_t)v * PLANE, v, out[v]);
    }
  };
  // One entry by one lane: interpolation from the eight corners.
  auto fill_entry = [&](const GhostEntryR& E) {
    double rc[8], uc[8], vc[8], wc[8], pc[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const int a = (m == 1 || m == 4 || m == 5 || m == 7);
      const int bq = (m == 2 || m == 4 || m == 6 || m == 7);
      const int cq = (m == 3 || m == 5 || m == 6 || m == 7);
      double q[5];
      if (lds) {
#pragma unroll
        for (int v = 0; v < 5; ++v) q[v] = gsm[v * lds_cap + E.lc[m]];
      } else {
        const size_t o = sidx(c, 0, E.pi + a, E.pj + bq, E.pk + cq);
#pragma unroll
        for (int v = 0; v < 5; ++v) q[v] = F.ld(U, o + (size_t)v * PLANE, v);
      }
      corner_prim(q, rc[m], uc[m], vc[m], wc[m], pc[m]);
    }
    combine(E, rc, uc, vc, wc, pc);
  };
  // Levels run one after the other, each taking one entry's latency.  A level
  // of at most 64 (32) entries spreads each entry over 4 (8) lanes: they form
  // its corners' primitives into LDS, and the entry's first lane combines them
  // as one lane would.  The lane's entry (and corner slots) of the next level
  // are loaded while the current one is computed; entries beyond the first
  // 256 of a level load in place.  (A lane without an entry in a level loads
  // entry 0 of the schedule, which it does not use.)
  __shared__ double cpr[64 * 8 * 5];   // corner primitives [entry][corner][5]
  const int t = threadIdx.x;
  auto lanes_of = [&](int nl) { return !lds || nl > 64 ? 1 : (nl > 32 ? 4 : 8); };
  // Entry of lane t in a level starting at e with nl entries.
  auto entry_of = [&](int e, int nl) {
    const int L = lanes_of(nl);
    const int q = t / L;
    return q < nl ? e + q : 0;
  };
  // Slots of the corners lane t forms in a level starting at e with nl
  // entries (spread levels only; corners r*K and r*K+1 of entry q).
  auto slots_of = [&](int e, int nl, int& s0, int& s1) {
    const int L = lanes_of(nl);
    s0 = s1 = 0;
    if (L == 1) return;
    const int q = t / L, r = t % L, K = 8 / L;
    if (q >= nl) return;
    s0 = gt.ent[e + q].lc[r * K];
    if (K == 2) s1 = gt.ent[e + q].lc[r * K + 1];
  };
  int nl = gt.level_off[l0 + 1] - gt.level_off[l0];
  GhostEntryR Enext = load_entry(gt.ent + entry_of(gt.level_off[l0], nl));
  int sn0, sn1;
  slots_of(gt.level_off[l0], nl, sn0, sn1);
  for (int lv = l0; lv < l1; ++lv) {
    const int e0 = gt.level_off[lv], e1 = gt.level_off[lv + 1];
    nl = e1 - e0;
    const int L = lanes_of(nl);
    const GhostEntryR Ecur = Enext;
    const int sc0 = sn0, sc1 = sn1;
    if (lv + 1 < l1) {
      const int nn = gt.level_off[lv + 2] - e1;
      Enext = load_entry(gt.ent + entry_of(e1, nn));
      slots_of(e1, nn, sn0, sn1);
    }
    if (L == 1) {
      const int n0 = e0 + t;
      if (n0 < e1) fill_entry(Ecur);
      for (int n = n0 + WG; n < e1; n += WG) fill_entry(load_entry(gt.ent + n));
    } else {
      // Lane r of entry q forms corners r*K .. r*K+K-1 (the corner index only
      // addresses memory, so the lanes of a wave run one instruction stream).
      const int q = t / L, r = t % L, K = 8 / L;
      if (q < nl) {
        for (int kk = 0; kk < K; ++kk) {
          const int m = r * K + kk, sl = kk == 0 ? sc0 : sc1;
          double qv[5], pv[5];
#pragma unroll
          for (int v = 0; v < 5; ++v) qv[v] = gsm[v * lds_cap + sl];
          corner_prim(qv, pv[0], pv[1], pv[2], pv[3], pv[4]);
#pragma unroll
          for (int v = 0; v < 5; ++v) cpr[(q * 8 + m) * 5 + v] = pv[v];
        }
      }
      __syncthreads();
      if (r == 0 && q < nl) {
        double rc[8], uc[8], vc[8], wc[8], pc[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
          const double* a = cpr + (q * 8 + m) * 5;
          rc[m] = a[0];
          uc[m] = a[1];
          vc[m] = a[2];
          wc[m] = a[3];
          pc[m] = a[4];
        }
        combine(Ecur, rc, uc, vc, wc, pc);
      }
    }
    __syncthreads();
  }
  if (lds) {
    const int nw = gt.touched_nw[g];
    for (int n = threadIdx.x; n < nw; n += WG) {
      const size_t o = (size_t)c * CUBE_STRIDE + gt.touched[t0 + n];
      const int sl = gt.touched_final[t0 + n];
#pragma unroll
      for (int v = 0; v < 5; ++v) F.st(U, o + (size_t)v * PLANE, v, gsm[v * lds_cap + sl]);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(WG) void k_ghost(T* __restrict__ U, GhostTables gt, int lds_cap,
                                              PhysTab pt) {
  extern __shared__ double gsm[];   // [5][lds_cap]
  ghost_cube(U, gt, blockIdx.x, lds_cap, pt, gsm);
}

// Edge-ghost averaging of one cube in the reference's three sequential passes
// (Interface.cpp:5640-5732): i-edges, then j-edges, then k-edges.  Called by a
// whole workgroup; the face ghosts it reads must be complete.  A pass reads
// face ghost cells and writes edge cells only, so each lane loads both of its
// items (w = t and t + 256) before storing either.
template <typename T>
__device__ __forceinline__ void edge_fill(T* __restrict__ U, int c) {
  const int t = threadIdx.x;
  // Pass a (0: i-edges at (j,k) in {1,18}^2, i = 2..17; 1: j-edges at (i,k),
  // j = 1..18; 2: k-edges at (i,j), k = 1..18).  Item w: edge e = w & 3,
  // variable v = (w >> 2) % 5, position along the edge (w >> 2) / 5.
  auto pass = [&](auto axis) {
    constexpr int A = decltype(axis)::value;
    constexpr int n = (A == 0 ? 16 : 18) * 4 * 5, p0 = A == 0 ? 2 : 1;
    T xa[2], xb[2];
    size_t dst[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int w = t + h * WG;
      if (w >= n) continue;
      const int q = (w >> 2) / 5 + p0, e = w & 3, v = (w >> 2) % 5;
      size_t a, b;
      if (A == 0) {
        if (e == 0) a = sidx(c, v, q, 1, 2), b = sidx(c, v, q, 2, 1);
        else if (e == 1) a = sidx(c, v, q, 18, 2), b = sidx(c, v, q, 17, 1);
        else if (e == 2) a = sidx(c, v, q, 1, 17), b = sidx(c, v, q, 2, 18);
        else a = sidx(c, v, q, 18, 17), b = sidx(c, v, q, 17, 18);
        dst[h] = sidx(c, v, q, (e == 1 || e == 3) ? 18 : 1, (e >= 2) ? 18 : 1);
      } else if (A == 1) {
        if (e == 0) a = sidx(c, v, 1, q, 2), b = sidx(c, v, 2, q, 1);
        else if (e == 1) a = sidx(c, v, 2, q, 18), b = sidx(c, v, 1, q, 17);
        else if (e == 2) a = sidx(c, v, 17, q, 1), b = sidx(c, v, 18, q, 2);
        else a = sidx(c, v, 17, q, 18), b = sidx(c, v, 18, q, 17);
        dst[h] = sidx(c, v, (e >= 2) ? 18 : 1, q, (e == 1 || e == 3) ? 18 : 1);
      } else {
        if (e == 0) a = sidx(c, v, 1, 2, q), b = sidx(c, v, 2, 1, q);
        else if (e == 1) a = sidx(c, v, 2, 18, q), b = sidx(c, v, 1, 17, q);
        else if (e == 2) a = sidx(c, v, 17, 1, q), b = sidx(c, v, 18, 2, q);
        else a = sidx(c, v, 17, 18, q), b = sidx(c, v, 18, 17, q);
        dst[h] = sidx(c, v, (e >= 2) ? 18 : 1, (e == 1 || e == 3) ? 18 : 1, q);
      }
      xa[h] = U[a];
      xb[h] = U[b];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
      if (t + h * WG < n) U[dst[h]] = 0.5 * (xa[h] + xb[h]);
  };
  pass(std::integral_constant<int, 0>());
  __syncthreads();
  pass(std::integral_constant<int, 1>());
  __syncthreads();
  pass(std::integral_constant<int, 2>());
}

// ---------------------------------------------------------------------------
// Cube-to-cube halo of equal-size neighbours as a gather into each cube's own
// face layers (Interface.cpp:1745-1782, 2604-2641, 3474-3511), then the
// cube's edge-ghost averaging in the reference's three sequential passes
// (:5640-5732).  Sources are interior layers, which nothing in this kernel
// writes, so cubes proceed independently.
template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

template <typename T>
__global__ __launch_bounds__(WG) void k_halo_edge(T* __restrict__ U,
                                                  const HaloSrc* __restrict__ srcs,
                                                  const int* __restrict__ cube_case) {
  using T2 = typename Vec2<T>::type;
  const int c = blockIdx.x + 1;
  if (cube_case[c] < 0) return;
  const HaloSrc hs = srcs[c];
  const int t = threadIdx.x;
  const int p = (t >> 4) + 2, r = (t & 15) + 2;
  // All loads first, then all stores: one memory round trip per lane.  The
  // two layers of a z-face are adjacent in k and move as one 2-element vector
  // (rows start 16-byte aligned for FP64 and 8-byte aligned for FP32).
  // Faces without a neighbour read their own cube (discarded) so that every
  // load is unconditional.
  const int sx0 = hs.src[0], sx1 = hs.src[1], sy0 = hs.src[2], sy1 = hs.src[3];
  const int sz0 = hs.src[4], sz1 = hs.src[5];
  const int ax0 = sx0 > 0 ? sx0 : c, ax1 = sx1 > 0 ? sx1 : c, ay0 = sy0 > 0 ? sy0 : c,
            ay1 = sy1 > 0 ? sy1 : c, az0 = sz0 > 0 ? sz0 : c, az1 = sz1 > 0 ? sz1 : c;
  T x0[2][5], x1[2][5], y0[2][5], y1[2][5], z0[2][5], z1[2][5];
#pragma unroll
  for (int l = 0; l < 2; ++l)
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      x0[l][v] = U[sidx(ax0, v, 16 + l, p, r)];
      x1[l][v] = U[sidx(ax1, v, 2 + l, p, r)];
      y0[l][v] = U[sidx(ay0, v, p, 16 + l, r)];
      y1[l][v] = U[sidx(ay1, v, p, 2 + l, r)];
    }
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const T2 a = *reinterpret_cast<const T2*>(U + sidx(az0, v, p, r, 16));
    const T2 b = *reinterpret_cast<const T2*>(U + sidx(az1, v, p, r, 2));
    z0[0][v] = a.x;
    z0[1][v] = a.y;
    z1[0][v] = b.x;
    z1[1][v] = b.y;
  }
#pragma unroll
  for (int l = 0; l < 2; ++l)
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      if (sx0 > 0) U[sidx(c, v, l, p, r)] = x0[l][v];
      if (sx1 > 0) U[sidx(c, v, 18 + l, p, r)] = x1[l][v];
      if (sy0 > 0) U[sidx(c, v, p, l, r)] = y0[l][v];
      if (sy1 > 0) U[sidx(c, v, p, 18 + l, r)] = y1[l][v];
    }
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    T2 a, b;
    a.x = z0[0][v];
    a.y = z0[1][v];
    b.x = z1[0][v];
    b.y = z1[1][v];
    if (sz0 > 0) *reinterpret_cast<T2*>(U + sidx(c, v, p, r, 0)) = a;
    if (sz1 > 0) *reinterpret_cast<T2*>(U + sidx(c, v, p, r, 18)) = b;
  }
  __syncthreads();
  edge_fill(U, c);
}

// Pre-flux halo refresh.  Between the post-update exchange and the next
// stage's pre-flux exchange the only interior cells that change are the IB
// ghost cells written by the ghost fill, and the domain BCs only write ghost
// layers that no gather writes.  The full gather therefore reduces to copying
// those ghost cells that lie in a neighbour's source layers, followed by the
// edge passes, which read face ghosts the BCs may have changed.
template <typename T>
__global__ __launch_bounds__(WG) void k_halo_sparse(T* __restrict__ U,
                                                    const int2* __restrict__ pairs, int n) {
  const int g = blockIdx.x * WG + threadIdx.x;
  if (g >= n) return;
  const int2 pr = pairs[g];  // (source flat cell, destination flat cell), cube*8000 + cell
  const size_t so = (size_t)(pr.x / PLANE) * CUBE_STRIDE + pr.x % PLANE;
  const size_t dd = (size_t)(pr.y / PLANE) * CUBE_STRIDE + pr.y % PLANE;
#pragma unroll
  for (int v = 0; v < 5; ++v) U[dd + (size_t)v * PLANE] = U[so + (size_t)v * PLANE];
}

template <typename T>
__global__ __launch_bounds__(WG) void k_edge(T* __restrict__ U,
                                             const int* __restrict__ cube_case) {
  if (cube_case[blockIdx.x + 1] < 0) return;
  edge_fill(U, blockIdx.x + 1);
}

// Edge passes of the listed cubes only.
template <typename T>
__global__ __launch_bounds__(WG) void k_edge_list(T* __restrict__ U, const int* __restrict__ list) {
  edge_fill(U, list[blockIdx.x]);
}

// Stage start after a flux kernel that wrote its face halos into the
// neighbours (flux_push), for one cube (entry c): the edge passes of the
// post-update exchange (Interface.cpp:5640-5732), then the domain BCs of the
// next stage.  Entry -c asks only for the edge cells the BCs read: the z-BC
// (BCM_Z_boundary_condition.cpp:47-88) copies the plane k = 2 (or 17) whole,
// four of whose cells are k-edges, each the mean of two face ghost cells
// taken before the X- and Y-BCs overwrite them.
template <typename T>
__device__ __forceinline__ void edge_bc_cube(T* __restrict__ U, int entry,
                                             const uint8_t* __restrict__ cube_mask, PhysTab pt) {
  const int c = entry > 0 ? entry : -entry;
  if (pt.cube_case[c] < 0) return;
  const int m = cube_mask[c];
  if (entry > 0) {
    edge_fill(U, c);
  } else if (threadIdx.x < 40) {
    const int side = threadIdx.x / 20, cell = (threadIdx.x / 5) % 4, v = threadIdx.x % 5;
    if (m & (side ? 32 : 16)) {
      const int ii = cell & 1 ? 18 : 1, jj = cell & 2 ? 18 : 1, kk = side ? 17 : 2;
      int a, b;
      edge_sources(ii, jj, kk, a, b);
      const size_t o = (size_t)c * CUBE_STRIDE + (size_t)v * PLANE;
      U[o + ii * 400 + jj * 20 + kk] = 0.5 * (U[o + a] + U[o + b]);
    }
  }
  if (m) {
    __syncthreads();
    bc_cube(U, c, m, pt);
  }
}

// Opening of a pushed stage in one launch: per item, a cube's edge passes and
// BCs (x: entry as for edge_bc_cube, 0 for none), then its IB fill (y: ghost
// table index, -1 for none).  Both touch only the cube's own cells, so cubes
// proceed independently, each in the reference's order.
template <typename T>
__global__ __launch_bounds__(WG) void k_open(T* __restrict__ U, const int2* __restrict__ items,
                                             const uint8_t* __restrict__ cube_mask,
                                             GhostTables gt, int lds_cap, PhysTab pt) {
  extern __shared__ double gsm[];   // [5][lds_cap]
  const int2 it = items[blockIdx.x];
  if (it.x != 0) edge_bc_cube(U, it.x, cube_mask, pt);
  if (it.y >= 0) {
    __syncthreads();
    ghost_cube(U, gt, it.y, lds_cap, pt, gsm);
  }
}

// Two-level faces (BCM_Interface.cpp:1790-2600 for x, :2651-3470 for y,
// :3518-4360 for z): one task per workgroup.  Injection copies storage values
// (the FP32 iterate holds deviations from one reference state per case, and
// both cubes of a face belong to the same case); restriction sums the 8 full
// states in the reference's order, (0,0,0) (1,0,0) (0,1,0) (0,0,1) (1,1,0)
// (1,0,1) (0,1,1) (1,1,1) of (i,j,k) offsets, and scales by 0.125.
template <typename T>
__global__ __launch_bounds__(WG) void k_halo_ml(T* __restrict__ U,
                                                const MlTaskDev* __restrict__ tasks,
                                                PhysTab pt) {
  const MlTaskDev t = tasks[blockIdx.x];
  const int A = t.axis, T0 = A == 0 ? 1 : 0, T1 = A == 2 ? 1 : 2;
  const int o0 = (t.slot == 2 || t.slot == 4) ? 8 : 0;
  const int o1 = (t.slot == 3 || t.slot == 4) ? 8 : 0;
  int d[3], q[3];
  if (!t.restrict_) {
    for (int n = threadIdx.x; n < 2 * 256; n += WG) {
      const int l = n >> 8, p = ((n >> 4) & 15) + 2, r = (n & 15) + 2;
      d[A] = t.hi ? 18 + l : l;
      d[T0] = p;
      d[T1] = r;
      q[A] = t.hi ? 2 : 17;
      q[T0] = p / 2 + 1 + o0;
      q[T1] = r / 2 + 1 + o1;
      const size_t dd = sidx(t.dst, 0, d[0], d[1], d[2]);
      const size_t so = sidx(t.src, 0, q[0], q[1], q[2]);
#pragma unroll
      for (int v = 0; v < 5; ++v) U[dd + (size_t)v * PLANE] = U[so + (size_t)v * PLANE];
    }
    return;
  }
  const typename OffOf<T>::type F(pt.phys[pt.cube_case[t.dst]]);
  for (int n = threadIdx.x; n < 2 * 64; n += WG) {
    const int l = n >> 6, p = ((n >> 3) & 7) + 2, r = (n & 7) + 2;
    d[A] = t.hi ? 18 + l : l;
    d[T0] = p + o0;
    d[T1] = r + o1;
    q[A] = t.hi ? 2 * (l + 1) : 14 + 2 * l;
    q[T0] = (p - 1) * 2;
    q[T1] = (r - 1) * 2;
    const size_t dd = sidx(t.dst, 0, d[0], d[1], d[2]);
    const size_t b = sidx(t.src, 0, q[0], q[1], q[2]);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const size_t o = b + (size_t)v * PLANE;
      double s = F.ld(U, o, v);
      s = s + F.ld(U, o + 400, v);
      s = s + F.ld(U, o + 20, v);
      s = s + F.ld(U, o + 1, v);
      s = s + F.ld(U, o + 420, v);
      s = s + F.ld(U, o + 401, v);
      s = s + F.ld(U, o + 21, v);
      s = s + F.ld(U, o + 421, v);
      F.st(U, dd + (size_t)v * PLANE, v, 0.125 * s);
    }
  }
}

// Physical-time shift of interior cells (main.cpp:1292-1315).
__global__ __launch_bounds__(WG) void k_shift(const double* __restrict__ U,
                                              double* __restrict__ U1,
                                              double* __restrict__ U1q,
                                              const int* __restrict__ cube_case) {
  const int c = blockIdx.y + 1;
  if (cube_case[c] < 0) return;
  const int i = blockIdx.x + 2;
  const int j = (threadIdx.x >> 4) + 2;
  const int k = (threadIdx.x & 15) + 2;
  const size_t b = sidx(c, 0, i, j, k);
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const size_t o = b + (size_t)v * PLANE;
    U1q[o] = U1[o];
    U1[o] = U[o];
  }
}

__global__ __launch_bounds__(WG) void k_aos_to_soa(const double* __restrict__ aos,
                                                   double* __restrict__ soa, int ncells) {
  const int g = blockIdx.x * WG + threadIdx.x;
  if (g >= ncells) return;
  const int c = g / PLANE, cell = g % PLANE;
#pragma unroll
  for (int v = 0; v < 5; ++v) soa[(size_t)c * CUBE_STRIDE + (size_t)v * PLANE + cell] = aos[(size_t)g * 5 + v];
}

__global__ __launch_bounds__(WG) void k_soa_to_aos(const double* __restrict__ soa,
                                                   double* __restrict__ aos, int ncells) {
  const int g = blockIdx.x * WG + threadIdx.x;
  if (g >= ncells) return;
  const int c = g / PLANE, cell = g % PLANE;
#pragma unroll
  for (int v = 0; v < 5; ++v) aos[(size_t)g * 5 + v] = soa[(size_t)c * CUBE_STRIDE + (size_t)v * PLANE + cell];
}

}  // namespace

// ---------------------------------------------------------------------------
int update_blocks(int ncube) { return 16 * (ncube - 1); }

template <typename T>
void launch_bc_t(hipStream_t s, T* U, const int* cubes, const uint8_t* masks, int n, PhysTab pt) {
  if (n > 0) hipLaunchKernelGGL(k_bc<T>, dim3(n), dim3(WG), 0, s, U, cubes, masks, pt);
}
void launch_bc(hipStream_t s, double* U, const int* cubes, const uint8_t* masks, int n,
               PhysTab pt) {
  launch_bc_t(s, U, cubes, masks, n, pt);
}
void launch_bc(hipStream_t s, float* U, const int* cubes, const uint8_t* masks, int n,
               PhysTab pt) {
  launch_bc_t(s, U, cubes, masks, n, pt);
}

template <typename T>
void launch_ghost_t(hipStream_t s, T* U, const GhostTables& g, PhysTab pt) {
  if (g.ncubes <= 0) return;
  const int cap = g.max_slots < GHOST_LDS_SLOTS ? g.max_slots : GHOST_LDS_SLOTS;
  const size_t shmem = (size_t)(cap > 0 ? cap : 1) * 5 * sizeof(double);
  hipLaunchKernelGGL(k_ghost<T>, dim3(g.ncubes), dim3(WG), shmem, s, U, g, cap, pt);
}
void launch_ghost(hipStream_t s, double* U, const GhostTables& g, PhysTab pt) {
  launch_ghost_t(s, U, g, pt);
}
void launch_ghost(hipStream_t s, float* U, const GhostTables& g, PhysTab pt) {
  launch_ghost_t(s, U, g, pt);
}

template <typename T>
void launch_open_t(hipStream_t s, T* U, const int2* items, int n, const uint8_t* cube_mask,
                   const GhostTables& g, PhysTab pt) {
  if (n <= 0) return;
  const int cap = g.max_slots < GHOST_LDS_SLOTS ? g.max_slots : GHOST_LDS_SLOTS;
  const size_t shmem = (size_t)(cap > 0 ? cap : 1) * 5 * sizeof(double);
  hipLaunchKernelGGL(k_open<T>, dim3(n), dim3(WG), shmem, s, U, items, cube_mask, g, cap, pt);
}
void launch_open(hipStream_t s, double* U, const int2* items, int n, const uint8_t* cube_mask,
                 const GhostTables& g, PhysTab pt) {
  launch_open_t(s, U, items, n, cube_mask, g, pt);
}
void launch_open(hipStream_t s, float* U, const int2* items, int n, const uint8_t* cube_mask,
                 const GhostTables& g, PhysTab pt) {
  launch_open_t(s, U, items, n, cube_mask, g, pt);
}

template <typename T>
void launch_edge_t(hipStream_t s, T* U, int ncube, const int* cube_case) {
  if (ncube > 1) hipLaunchKernelGGL(k_edge<T>, dim3(ncube - 1), dim3(WG), 0, s, U, cube_case);
}
template <typename T>
void launch_halo_ml_t(hipStream_t s, T* U, const MlTaskDev* tasks, int ntask, PhysTab pt) {
  if (ntask > 0) hipLaunchKernelGGL(k_halo_ml<T>, dim3(ntask), dim3(WG), 0, s, U, tasks, pt);
}
void launch_halo_ml(hipStream_t s, double* U, const MlTaskDev* tasks, int ntask, PhysTab pt) {
  launch_halo_ml_t(s, U, tasks, ntask, pt);
}
void launch_halo_ml(hipStream_t s, float* U, const MlTaskDev* tasks, int ntask, PhysTab pt) {
  launch_halo_ml_t(s, U, tasks, ntask, pt);
}
void launch_edge(hipStream_t s, double* U, int ncube, const int* cube_case) {
  launch_edge_t(s, U, ncube, cube_case);
}
void launch_edge(hipStream_t s, float* U, int ncube, const int* cube_case) {
  launch_edge_t(s, U, ncube, cube_case);
}

template <typename T>
void launch_halo_edge_t(hipStream_t s, T* U, const HaloSrc* src, int ncube, const int* cube_case) {
  if (ncube > 1)
    hipLaunchKernelGGL(k_halo_edge<T>, dim3(ncube - 1), dim3(WG), 0, s, U, src, cube_case);
}
void launch_halo_edge(hipStream_t s, double* U, const HaloSrc* src, int ncube,
                      const int* cube_case) {
  launch_halo_edge_t(s, U, src, ncube, cube_case);
}
void launch_halo_edge(hipStream_t s, float* U, const HaloSrc* src, int ncube,
                      const int* cube_case) {
  launch_halo_edge_t(s, U, src, ncube, cube_case);
}

template <typename T>
void launch_halo_pairs_t(hipStream_t s, T* U, const int2* pairs, int npairs) {
  if (npairs > 0)
    hipLaunchKernelGGL(k_halo_sparse<T>, dim3((npairs + WG - 1) / WG), dim3(WG), 0, s, U, pairs,
                       npairs);
}
void launch_halo_pairs(hipStream_t s, double* U, const int2* pairs, int npairs) {
  launch_halo_pairs_t(s, U, pairs, npairs);
}
void launch_halo_pairs(hipStream_t s, float* U, const int2* pairs, int npairs) {
  launch_halo_pairs_t(s, U, pairs, npairs);
}

template <typename T>
void launch_edge_list_t(hipStream_t s, T* U, const int* list, int n) {
  if (n > 0) hipLaunchKernelGGL(k_edge_list<T>, dim3(n), dim3(WG), 0, s, U, list);
}
void launch_edge_list(hipStream_t s, double* U, const int* list, int n) {
  launch_edge_list_t(s, U, list, n);
}
void launch_edge_list(hipStream_t s, float* U, const int* list, int n) {
  launch_edge_list_t(s, U, list, n);
}

void launch_flux(hipStream_t s, const double* U, const double* U1, const double* U1q,
                 const int8_t* fws, const double* cube_size, double* Rku, double* Res,
                 int ncube, PhysTab pt, double deltaT, double deltaTau, double e) {
  if (ncube > 1)
    hipLaunchKernelGGL(k_flux, dim3(16, ncube - 1), dim3(WG), 0, s, U, U1, U1q, fws, cube_size,
                       Rku, Res, pt, deltaT, deltaTau, e);
}

void launch_flux_march(hipStream_t s, double* U, const double* U1, const double* U1q,
                       const int8_t* fws, const double* cube_size, double* Rku, double* Res,
                       int ncube, PhysTab pt, double deltaT, double deltaTau, double e, int nseg,
                       int RK, int fuse, double* U1p1, double* U1p2, double* partials) {
  if (ncube > 1)
    hipLaunchKernelGGL(k_flux_march<false>, dim3((ncube - 1) * nseg), dim3(WG), 0, s, U, U1, U1q,
                       fws, cube_size, Rku, Res, pt, deltaT, deltaTau, e, nseg, RK,
                       fuse && nseg == 1, U1p1, U1p2, partials, nullptr, nullptr, nullptr, 1);
}

void launch_flux_push(hipStream_t s, const double* Uin, double* Uout, const HaloSrc* halo,
                      const uint8_t* cube_mask, const double* U1, const double* U1q,
                      const int8_t* fws, const double* cube_size, double* Rku, double* Res,
                      int ncube, PhysTab pt, double deltaT, double deltaTau, double e, int RK,
                      double* U1p1, double* U1p2, double* partials, int aux) {
  if (ncube > 1)
    hipLaunchKernelGGL(k_flux_march<true>, dim3(ncube - 1), dim3(WG), 0, s,
                       const_cast<double*>(Uin), U1, U1q, fws, cube_size, Rku, Res, pt, deltaT,
                       deltaTau, e, 1, RK, 0, U1p1, U1p2, partials, Uout, halo, cube_mask, aux);
}

__global__ __launch_bounds__(WG) void k_copy_d(const double* __restrict__ src,
                                               double* __restrict__ dst, size_t n) {
  for (size_t g = (size_t)blockIdx.x * WG + threadIdx.x; g < n; g += (size_t)gridDim.x * WG)
    dst[g] = src[g];
}

namespace {
// DivBy/div_sh against a / b: per thread one denominator and four numerators
// from a seeded hash (random signs, significands and exponents in
// [-126, 126]; about one numerator in 16 is a signed zero).
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double test_value(unsigned long long r) {
  const double m = 1.0 + (double)(r >> 12) * 0x1p-52;
  const double v = ldexp(m, (int)((r >> 1) % 253) - 126);
  return (r & 1) ? -v : v;
}
__global__ __launch_bounds__(WG) void k_div_selftest(unsigned long long seed, int n,
                                                     unsigned long long* __restrict__ bad) {
  const int g = blockIdx.x * WG + threadIdx.x;
  if (g >= n) return;
  const unsigned long long base = mix64(seed ^ ((unsigned long long)g << 20));
  const double b = test_value(mix64(base));
  double a[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    a[k] = test_value(mix64(base + 1 + k));
    if ((mix64(base + 10 + k) & 15) == 0) a[k] = (k & 1) ? -0.0 : 0.0;
  }
  const DivBy d = div_by(b);
  unsigned long long nb = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k)
    nb += __double_as_longlong(div_sh(d, a[k])) != __double_as_longlong(a[k] / b);
  if (nb) atomicAdd(bad, nb);
}
}  // namespace

int div_selftest(unsigned long long seed, int n, unsigned long long* mismatches) {
  unsigned long long* d = nullptr;
  if (n < 0 || hipMalloc((void**)&d, sizeof(unsigned long long)) != hipSuccess) return 1;
  int rc = hipMemset(d, 0, sizeof(unsigned long long)) != hipSuccess;
  if (!rc && n > 0) hipLaunchKernelGGL(k_div_selftest, dim3((n + WG - 1) / WG), dim3(WG), 0, 0, seed, n, d);
  rc = rc || hipGetLastError() != hipSuccess ||
       hipMemcpy(mismatches, d, sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess;
  (void)hipFree(d);
  return rc;
}

void launch_copy_d(hipStream_t s, const double* src, double* dst, size_t n) {
  if (n == 0) return;
  const size_t blocks = (n + WG - 1) / WG;
  hipLaunchKernelGGL(k_copy_d, dim3(blocks < 65536 ? (unsigned)blocks : 65536u), dim3(WG), 0, s,
                     src, dst, n);
}

void launch_update(hipStream_t s, int RK, double* U, double* U1p1, double* U1p2,
                   const double* Rku, const double* Res, const int8_t* fws, int ncube,
                   const int* cube_case, double* partials) {
  if (ncube > 1)
    hipLaunchKernelGGL(k_update, dim3(16, ncube - 1), dim3(WG), 0, s, RK, U, U1p1, U1p2, Rku,
                       Res, fws, cube_case, partials);
}

void launch_reduce(hipStream_t s, double* partials, int rows_per_cube, const CaseDesc* cases,
                   int ncases, const DevPhys* phys, double* er) {
  if (ncases <= 0) return;
  if (rows_per_cube > 1)
    hipLaunchKernelGGL(k_fold_rows, dim3(8, ncases), dim3(WG), 0, s, partials, rows_per_cube,
                       cases);
  hipLaunchKernelGGL(k_reduce, dim3(ncases), dim3(WG), 0, s, partials, rows_per_cube, cases, phys,
                     er);
}

void launch_shift(hipStream_t s, const double* U, double* U1, double* U1q, int ncube,
                  const int* cube_case) {
  if (ncube > 1)
    hipLaunchKernelGGL(k_shift, dim3(16, ncube - 1), dim3(WG), 0, s, U, U1, U1q, cube_case);
}

void launch_aos_to_soa(hipStream_t s, const double* aos, double* soa, int ncube) {
  const int n = ncube * PLANE;
  hipLaunchKernelGGL(k_aos_to_soa, dim3((n + WG - 1) / WG), dim3(WG), 0, s, aos, soa, n);
}

void launch_soa_to_aos(hipStream_t s, const double* soa, double* aos, int ncube) {
  const int n = ncube * PLANE;
  hipLaunchKernelGGL(k_soa_to_aos, dim3((n + WG - 1) / WG), dim3(WG), 0, s, soa, aos, n);
}

}  // namespace bcm
