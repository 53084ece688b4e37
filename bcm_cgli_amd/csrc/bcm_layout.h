// Cube geometry, flag values and physical constants shared by the host setup
// and the HIP engine.
//
// A cube holds NC^3 interior cells padded by NB ghost layers on every side
// (Resolution.h:9-22): storage extent NP = NC + 2*NB = 20 per axis, interior
// index range [NB, NB+NC-1] = [2, 17].  Cubes are numbered from 1; slot 0 of
// every per-cube array is unused, exactly as in Array.h:11-36.
#pragma once

#include <stdint.h>

namespace bcm {

constexpr int NC = 16;            // interior cells per axis (NcubeX/Y/Z)
constexpr int NB = 2;             // ghost layers (n_buffer)
constexpr int NP = NC + 2 * NB;   // padded extent (X_size) = 20
constexpr int NV = 5;             // conserved variables (Ndim)
constexpr int ILO = NB;           // first interior index = 2
constexpr int IHI = NB + NC - 1;  // last interior index (nx) = 17
constexpr int CELLS_PADDED = NP * NP * NP;   // 8000
constexpr int CELLS_INTERIOR = NC * NC * NC; // 4096

// Cell flags (BCM.h:6-8).
constexpr int IFLUID = 1;
constexpr int IGHOST = 0;
constexpr int ISOLID = -1;

// Gas constants (prm.h:5-13).  R is an integer literal in the reference.
constexpr double KGAS = 1.4;
constexpr double RGAS = 287.0;
constexpr double CV = 717.5;
constexpr double PR_L = 0.72;
constexpr double EPFLUX = 1.0;      // Ep
constexpr double BIGNUM = 100000000.0;   // MAX
constexpr double SMALLNUM = 0.0000001;   // minimum

// Reference-layout flat indices.  State arrays are AoS per cell:
// [cube][i][j][k][v] with k fastest, v innermost (Array.h:11-23).
inline int64_t aos_index(int c, int i, int j, int k, int v) {
  return ((((int64_t)c * NP + i) * NP + j) * NP + k) * NV + v;
}
inline int64_t flag_index(int c, int i, int j, int k) {
  return (((int64_t)c * NP + i) * NP + j) * NP + k;
}
// adj_number[cube][slot 0..4][dir 0..6] (Array.h:39).
inline int64_t adj_index(int c, int slot, int dir) {
  return ((int64_t)c * 5 + slot) * 7 + dir;
}

// Physical constants of one flow case.  The reference keeps them as
// per-function locals from prm.h; a case here fixes M and Re by choosing
// U0 = Uin = M * sqrt(K * P0 / rho0) and mu_L = rho0 * U0 * D / Re.
struct Physics {
  double rho0, U0, Uin, V0, W0, P0, T0, mu_L;
};

}  // namespace bcm
