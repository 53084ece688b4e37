// bcm_cgli.hpp -- C++ overloads with the reference's own parameter types
// (Array.h array pointers, main.h:170-513), forwarding to the C ABI of
// bcm_cgli.h, so a caller written against main.h compiles unchanged.
#ifndef BCM_CGLI_HPP
#define BCM_CGLI_HPP

#include "bcm_cgli.h"

namespace bcm_cgli_types {
typedef double (*State)[20][20][20][5];
typedef int (*Flags)[20][20][20];
typedef int (*Adj)[5][7];
}  // namespace bcm_cgli_types

inline void BCM_X_boundary_condition(int myid, int nl, int nu, int* l, int* u,
                                     bcm_cgli_types::State U1_) {
  BCM_X_boundary_condition(myid, nl, nu, l, u, &U1_[0][0][0][0][0]);
}
inline void BCM_Y_boundary_condition(int nl, int nu, int* l, int* u, bcm_cgli_types::State U1_) {
  BCM_Y_boundary_condition(nl, nu, l, u, &U1_[0][0][0][0][0]);
}
inline void BCM_Z_boundary_condition(int nl, int nu, int* l, int* u, bcm_cgli_types::State U1_) {
  BCM_Z_boundary_condition(nl, nu, l, u, &U1_[0][0][0][0][0]);
}
inline void BCM_Ghostcell_minus(int myid, int* NNBC, double* weight, int* GCindex, int* IPsur,
                                bcm_cgli_types::Flags FWS, bcm_cgli_types::State U1_) {
  BCM_Ghostcell_minus(myid, NNBC, weight, GCindex, IPsur, &FWS[0][0][0][0], &U1_[0][0][0][0][0]);
}
inline void BCM_Ghostcell_plus(int myid, int* NNBC, double* weight, int* GCindex, int* IPsur,
                               bcm_cgli_types::Flags FWS, bcm_cgli_types::State U1_) {
  BCM_Ghostcell_plus(myid, NNBC, weight, GCindex, IPsur, &FWS[0][0][0][0], &U1_[0][0][0][0][0]);
}
inline void BCM_Flux_XYZ_Viscous_Runge_kutta(int myid, int ncube, int RK, double deltaT,
                                             double deltaTau, double e,
                                             bcm_cgli_types::Flags FWS, int* csl,
                                             double* cube_size, bcm_cgli_types::State U1_,
                                             bcm_cgli_types::State U1, bcm_cgli_types::State U1q,
                                             bcm_cgli_types::State U1p1,
                                             bcm_cgli_types::State U1p2,
                                             bcm_cgli_types::State Rku1,
                                             bcm_cgli_types::State Residual1, double* er) {
  BCM_Flux_XYZ_Viscous_Runge_kutta(myid, ncube, RK, deltaT, deltaTau, e, &FWS[0][0][0][0], csl,
                                   cube_size, &U1_[0][0][0][0][0], &U1[0][0][0][0][0],
                                   &U1q[0][0][0][0][0], &U1p1[0][0][0][0][0],
                                   &U1p2[0][0][0][0][0], &Rku1[0][0][0][0][0],
                                   &Residual1[0][0][0][0][0], er);
}
inline void BCM_Physical_time_shift(bcm_cgli_types::State U1_, bcm_cgli_types::State U1,
                                    bcm_cgli_types::State U1q) {
  BCM_Physical_time_shift(&U1_[0][0][0][0][0], &U1[0][0][0][0][0], &U1q[0][0][0][0][0]);
}

#endif  // BCM_CGLI_HPP
