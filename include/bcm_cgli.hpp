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

// BCM_Grid (main.h:25-43) with the Array.h pointer types.
inline void BCM_Grid(int myid, int ncube, int n_wallcube, double* cube_size, int* csl,
                     double* Xcube, double* Ycube, double* Zcube, double (*Xcnt)[20],
                     double (*Ycnt)[20], double (*Zcnt)[20], bcm_cgli_types::Flags FWS,
                     int (*adj_number)[5][7], int* wallcube) {
  BCM_Grid(myid, ncube, n_wallcube, cube_size, csl, Xcube, Ycube, Zcube, &Xcnt[0][0], &Ycnt[0][0],
           &Zcnt[0][0], &FWS[0][0][0][0], &adj_number[0][0][0], wallcube);
}
// BCM_Output (main.h:515-533) for one rank; rank_map is int[2][MPI_Ncube].
inline void BCM_Output(int myid, int ncube, int step, int* rank_map, bcm_cgli_types::State U1,
                       bcm_cgli_types::State U1q, double* cube_size, double (*Xcnt)[20],
                       double (*Ycnt)[20], double (*Zcnt)[20]) {
  BCM_Output(myid, ncube, step, rank_map, &U1[0][0][0][0][0], &U1q[0][0][0][0][0], cube_size,
             &Xcnt[0][0], &Ycnt[0][0], &Zcnt[0][0]);
}

#endif  // BCM_CGLI_HPP
