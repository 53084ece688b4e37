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

// BCM_Interface_table (main.h:170-219): adj_number is int (*)[5][7].
inline void BCM_Interface_table(int myid, int ncube, int n_wallcube, int* nadjX_eq, int* nadjY_eq,
                                int* nadjZ_eq, int* nadjX_bs_plus, int* nadjX_sb_plus,
                                int* nadjX_bs_minus, int* nadjX_sb_minus, int* nadjY_bs_plus,
                                int* nadjY_sb_plus, int* nadjY_bs_minus, int* nadjY_sb_minus,
                                int* nadjZ_bs_plus, int* nadjZ_sb_plus, int* nadjZ_bs_minus,
                                int* nadjZ_sb_minus, int* csl, bcm_cgli_types::Adj adj_number,
                                int* adjX_eq, int* adjY_eq, int* adjZ_eq, int* adjX_bs_plus,
                                int* adjX_sb_plus, int* adjX_bs_minus, int* adjX_sb_minus,
                                int* adjY_bs_plus, int* adjY_sb_plus, int* adjY_bs_minus,
                                int* adjY_sb_minus, int* adjZ_bs_plus, int* adjZ_sb_plus,
                                int* adjZ_bs_minus, int* adjZ_sb_minus) {
  BCM_Interface_table(myid, ncube, n_wallcube, nadjX_eq, nadjY_eq, nadjZ_eq, nadjX_bs_plus,
                      nadjX_sb_plus, nadjX_bs_minus, nadjX_sb_minus, nadjY_bs_plus, nadjY_sb_plus,
                      nadjY_bs_minus, nadjY_sb_minus, nadjZ_bs_plus, nadjZ_sb_plus, nadjZ_bs_minus,
                      nadjZ_sb_minus, csl, &adj_number[0][0][0], adjX_eq, adjY_eq, adjZ_eq,
                      adjX_bs_plus, adjX_sb_plus, adjX_bs_minus, adjX_sb_minus, adjY_bs_plus,
                      adjY_sb_plus, adjY_bs_minus, adjY_sb_minus, adjZ_bs_plus, adjZ_sb_plus,
                      adjZ_bs_minus, adjZ_sb_minus);
}

// BCM_Interface (main.h:224-354): rank_map is int (*)[MPI_Ncube], whose
// extent the caller's Resolution.h fixes at compile time, hence the template;
// adj_number is int (*)[5][7] and U1_ the Array.h state pointer.
template <int MPI_NCUBE>
inline void BCM_Interface(
    int myid, int ncube, int mPI_Nadj, int ncpu_bs, int ncpu_eq, int ncpu_sb, int max_nei_bs,
    int max_nei_eq, int max_nei_sb, int nadjX_eq, int nadjY_eq, int nadjZ_eq, int nadjX_bs_plus,
    int nadjX_sb_plus, int nadjX_bs_minus, int nadjX_sb_minus, int nadjY_bs_plus,
    int nadjY_sb_plus, int nadjY_bs_minus, int nadjY_sb_minus, int nadjZ_bs_plus,
    int nadjZ_sb_plus, int nadjZ_bs_minus, int nadjZ_sb_minus, int (*rank_map)[MPI_NCUBE],
    int* MPI_cpu, int* MPI_cube, int* MPI_cpu_adj, int* MPI_cube_adj, int* MPI_direction,
    int* MPI_interface, int* neighbor_cpu_eq, int* Ncube_Ncpu_eq, int* neighbor_cpu_sb,
    int* Ncube_Ncpu_sb, int* neighbor_cpu_bs, int* Ncube_Ncpu_bs, int* Scube_Ncpu_eq,
    int* Rcube_Ncpu_eq, double* send_data_curr_eq, double* recv_data_curr_eq,
    double* send_data_neig_eq, double* recv_data_neig_eq, int* Sdir_eq, int* Rdir_eq,
    int* Scube_Ncpu_sb, int* Rcube_Ncpu_sb, double* send_data_curr_sb, double* recv_data_curr_sb,
    double* send_data_neig_sb, double* recv_data_neig_sb, int* Sdir_sb, int* Rdir_sb,
    int* Scube_Ncpu_bs, int* Rcube_Ncpu_bs, double* send_data_curr_bs, double* recv_data_curr_bs,
    double* send_data_neig_bs, double* recv_data_neig_bs, int* Sdir_bs, int* Rdir_bs,
    int* ist_eq, int* ist_sb, int* ist_bs, int* adjN_sb, int* RadjN_bs, int* SadjN_bs, int* csl,
    bcm_cgli_types::Adj adj_number, int* adjX_eq, int* adjY_eq, int* adjZ_eq, int* adjX_bs_plus,
    int* adjX_sb_plus, int* adjX_bs_minus, int* adjX_sb_minus, int* adjY_bs_plus,
    int* adjY_sb_plus, int* adjY_bs_minus, int* adjY_sb_minus, int* adjZ_bs_plus,
    int* adjZ_sb_plus, int* adjZ_bs_minus, int* adjZ_sb_minus, bcm_cgli_types::State U1_) {
  BCM_Interface(myid, ncube, mPI_Nadj, ncpu_bs, ncpu_eq, ncpu_sb, max_nei_bs, max_nei_eq,
                max_nei_sb, nadjX_eq, nadjY_eq, nadjZ_eq, nadjX_bs_plus, nadjX_sb_plus,
                nadjX_bs_minus, nadjX_sb_minus, nadjY_bs_plus, nadjY_sb_plus, nadjY_bs_minus,
                nadjY_sb_minus, nadjZ_bs_plus, nadjZ_sb_plus, nadjZ_bs_minus, nadjZ_sb_minus,
                &rank_map[0][0], MPI_cpu, MPI_cube, MPI_cpu_adj, MPI_cube_adj, MPI_direction,
                MPI_interface, neighbor_cpu_eq, Ncube_Ncpu_eq, neighbor_cpu_sb, Ncube_Ncpu_sb,
                neighbor_cpu_bs, Ncube_Ncpu_bs, Scube_Ncpu_eq, Rcube_Ncpu_eq, send_data_curr_eq,
                recv_data_curr_eq, send_data_neig_eq, recv_data_neig_eq, Sdir_eq, Rdir_eq,
                Scube_Ncpu_sb, Rcube_Ncpu_sb, send_data_curr_sb, recv_data_curr_sb,
                send_data_neig_sb, recv_data_neig_sb, Sdir_sb, Rdir_sb, Scube_Ncpu_bs,
                Rcube_Ncpu_bs, send_data_curr_bs, recv_data_curr_bs, send_data_neig_bs,
                recv_data_neig_bs, Sdir_bs, Rdir_bs, ist_eq, ist_sb, ist_bs, adjN_sb, RadjN_bs,
                SadjN_bs, csl, &adj_number[0][0][0], adjX_eq, adjY_eq, adjZ_eq, adjX_bs_plus,
                adjX_sb_plus, adjX_bs_minus, adjX_sb_minus, adjY_bs_plus, adjY_sb_plus,
                adjY_bs_minus, adjY_sb_minus, adjZ_bs_plus, adjZ_sb_plus, adjZ_bs_minus,
                adjZ_sb_minus, &U1_[0][0][0][0][0]);
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
