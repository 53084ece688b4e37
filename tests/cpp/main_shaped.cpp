// A caller shaped like the reference's main program: its arrays are declared
// with the Array.h types (pointers to [20][20][20][5] doubles, [20][20][20]
// ints, [5][7] adjacency rows, rank_map as int (*)[MPI_Ncube]) and the drop-in
// routines are called with the argument lists of the reference's call sites
// (main.cpp:562-572 for the interface tables, main.cpp:1157-1245 for one RK
// stage sequence, main.cpp:1346-1351 for the output).  It has to compile
// against include/bcm_cgli.hpp and link against libbcm_cgli.so unchanged;
// it is not run (no device here).
#include "bcm_cgli.hpp"

#define MPI_Ncube 64
#define Ncube (MPI_Ncube + 1)
#define N_wallcube 9
#define X_size 20
#define Y_size 20
#define Z_size 20
#define Ndim 5

double (*U1_)[X_size][Y_size][Z_size][Ndim] = new double[Ncube][X_size][Y_size][Z_size][Ndim];
double (*U1)[X_size][Y_size][Z_size][Ndim] = new double[Ncube][X_size][Y_size][Z_size][Ndim];
double (*U1q)[X_size][Y_size][Z_size][Ndim] = new double[Ncube][X_size][Y_size][Z_size][Ndim];
double (*U1p1)[X_size][Y_size][Z_size][Ndim] = new double[Ncube][X_size][Y_size][Z_size][Ndim];
double (*U1p2)[X_size][Y_size][Z_size][Ndim] = new double[Ncube][X_size][Y_size][Z_size][Ndim];
double (*Residual1)[X_size][Y_size][Z_size][Ndim] = new double[Ncube][X_size][Y_size][Z_size][Ndim];
double (*Rku1)[X_size][Y_size][Z_size][Ndim] = new double[Ncube][X_size][Y_size][Z_size][Ndim];
double* cube_size = new double[Ncube];
int* csl = new int[Ncube];
double* Xcube = new double[Ncube];
double* Ycube = new double[Ncube];
double* Zcube = new double[Ncube];
double (*Xcnt)[X_size] = new double[Ncube][X_size];
double (*Ycnt)[Y_size] = new double[Ncube][Y_size];
double (*Zcnt)[Z_size] = new double[Ncube][Z_size];
int (*FWS)[X_size][Y_size][Z_size] = new int[Ncube][X_size][Y_size][Z_size];
int (*adj_number)[5][7] = new int[Ncube][5][7];
int* wallcube = new int[N_wallcube];
int *adjX_eq = new int[Ncube], *adjY_eq = new int[Ncube], *adjZ_eq = new int[Ncube];
int *adjX_bs_plus = new int[Ncube], *adjX_sb_plus = new int[Ncube];
int *adjX_bs_minus = new int[Ncube], *adjX_sb_minus = new int[Ncube];
int *adjY_bs_plus = new int[Ncube], *adjY_sb_plus = new int[Ncube];
int *adjY_bs_minus = new int[Ncube], *adjY_sb_minus = new int[Ncube];
int *adjZ_bs_plus = new int[Ncube], *adjZ_sb_plus = new int[Ncube];
int *adjZ_bs_minus = new int[Ncube], *adjZ_sb_minus = new int[Ncube];

int main() {
  int myid = 0;
  const double deltaT = 0.0001, deltaTau = deltaT / 200., e = 0.1;
  double* er = new double[8];
  int nadjX_eq, nadjY_eq, nadjZ_eq;
  int nadjX_bs_plus, nadjX_sb_plus, nadjX_bs_minus, nadjX_sb_minus;
  int nadjY_bs_plus, nadjY_sb_plus, nadjY_bs_minus, nadjY_sb_minus;
  int nadjZ_bs_plus, nadjZ_sb_plus, nadjZ_bs_minus, nadjZ_sb_minus;
  // One rank: no neighbour ranks, no MPI adjacencies.
  int MPI_Nadj = 1, Ncpu_bs = 0, Ncpu_eq = 0, Ncpu_sb = 0;
  int Max_nei_bs = 0, Max_nei_eq = 0, Max_nei_sb = 0;
  int(*rank_map)[MPI_Ncube] = new int[2][MPI_Ncube];
  int *MPI_cpu = new int[MPI_Nadj], *MPI_cube = new int[MPI_Nadj];
  int *MPI_cpu_adj = new int[MPI_Nadj], *MPI_cube_adj = new int[MPI_Nadj];
  int *MPI_direction = new int[MPI_Nadj], *MPI_interface = new int[MPI_Nadj];
  int *neighbor_cpu_eq = new int[1], *Ncube_Ncpu_eq = new int[1];
  int *neighbor_cpu_sb = new int[1], *Ncube_Ncpu_sb = new int[1];
  int *neighbor_cpu_bs = new int[1], *Ncube_Ncpu_bs = new int[1];
  int *Scube_Ncpu_eq = new int[Max_nei_eq + 1], *Rcube_Ncpu_eq = new int[Max_nei_eq + 1];
  int *Scube_Ncpu_sb = new int[Max_nei_sb + 1], *Rcube_Ncpu_sb = new int[Max_nei_sb + 1];
  int *Scube_Ncpu_bs = new int[Max_nei_bs + 1], *Rcube_Ncpu_bs = new int[Max_nei_bs + 1];
  double *send_data_curr_eq = new double[1], *recv_data_curr_eq = new double[1];
  double *send_data_neig_eq = new double[1], *recv_data_neig_eq = new double[1];
  double *send_data_curr_sb = new double[1], *recv_data_curr_sb = new double[1];
  double *send_data_neig_sb = new double[1], *recv_data_neig_sb = new double[1];
  double *send_data_curr_bs = new double[1], *recv_data_curr_bs = new double[1];
  double *send_data_neig_bs = new double[1], *recv_data_neig_bs = new double[1];
  int *Sdir_eq = new int[1], *Rdir_eq = new int[1], *Sdir_sb = new int[1], *Rdir_sb = new int[1];
  int *Sdir_bs = new int[1], *Rdir_bs = new int[1];
  int *ist_eq = new int[1], *ist_sb = new int[1], *ist_bs = new int[1];
  int *adjN_sb = new int[1], *RadjN_bs = new int[1], *SadjN_bs = new int[1];
  int NXbc_l = 0, NXbc_u = 0, NYbc_l = 0, NYbc_u = 0, NZbc_l = 0, NZbc_u = 0;
  int *Xbc_l = new int[1], *Xbc_u = new int[1], *Ybc_l = new int[1], *Ybc_u = new int[1];
  int *Zbc_l = new int[1], *Zbc_u = new int[1];
  int NBC_minus = 0, NBC_plus = 0;
  double *weight_minus = new double[1], *weight_plus = new double[1];
  int *GCindex_minus = new int[1], *GCindex_plus = new int[1];
  int *IPsur_minus = new int[1], *IPsur_plus = new int[1];

  BCM_Grid(myid, Ncube, N_wallcube, cube_size, csl, Xcube, Ycube, Zcube, Xcnt, Ycnt, Zcnt, FWS,
           adj_number, wallcube);

  BCM_Interface_table(myid, Ncube, N_wallcube,
                      &nadjX_eq, &nadjY_eq, &nadjZ_eq,
                      &nadjX_bs_plus, &nadjX_sb_plus, &nadjX_bs_minus, &nadjX_sb_minus,
                      &nadjY_bs_plus, &nadjY_sb_plus, &nadjY_bs_minus, &nadjY_sb_minus,
                      &nadjZ_bs_plus, &nadjZ_sb_plus, &nadjZ_bs_minus, &nadjZ_sb_minus,
                      csl,
                      adj_number,
                      adjX_eq, adjY_eq, adjZ_eq,
                      adjX_bs_plus, adjX_sb_plus, adjX_bs_minus, adjX_sb_minus,
                      adjY_bs_plus, adjY_sb_plus, adjY_bs_minus, adjY_sb_minus,
                      adjZ_bs_plus, adjZ_sb_plus, adjZ_bs_minus, adjZ_sb_minus);

  BCM_Reading_IBM(myid, &NBC_minus, -1, weight_minus, GCindex_minus, IPsur_minus);
  BCM_Reading_IBM(myid, &NBC_plus, 1, weight_plus, GCindex_plus, IPsur_plus);

  bcm_engine* eng = nullptr;
  bcm_engine_create(0, &eng);
  const double phys7[7] = {1.1842, 70.0, 70.0, 0.0, 0.0, 101300.0, 1e-3};
  bcm_bind(eng, Ncube, MPI_Ncube, &U1_[0][0][0][0][0], &U1[0][0][0][0][0], &U1q[0][0][0][0][0],
           &U1p1[0][0][0][0][0], &U1p2[0][0][0][0][0], &Rku1[0][0][0][0][0],
           &Residual1[0][0][0][0][0], &FWS[0][0][0][0], cube_size, phys7);

  for (int RK = 1; RK < 4; ++RK) {
    BCM_X_boundary_condition(myid, NXbc_l, NXbc_u, Xbc_l, Xbc_u, U1_);
    BCM_Y_boundary_condition(NYbc_l, NYbc_u, Ybc_l, Ybc_u, U1_);
    BCM_Z_boundary_condition(NZbc_l, NZbc_u, Zbc_l, Zbc_u, U1_);
    BCM_Ghostcell_minus(myid, &NBC_minus, weight_minus, GCindex_minus, IPsur_minus, FWS, U1_);
    BCM_Ghostcell_plus(myid, &NBC_plus, weight_plus, GCindex_plus, IPsur_plus, FWS, U1_);
    for (int pass = 0; pass < 2; ++pass) {
      // pass 0: before the flux call; pass 1: after it (main.cpp:1177, 1214).
      if (pass == 1)
        BCM_Flux_XYZ_Viscous_Runge_kutta(myid, Ncube, RK, deltaT, deltaTau, e, FWS, csl, cube_size,
                                         U1_, U1, U1q, U1p1, U1p2,
                                         Rku1, Residual1,
                                         er);
      BCM_Interface(myid, Ncube,
                    MPI_Nadj,
                    Ncpu_bs, Ncpu_eq, Ncpu_sb,
                    Max_nei_bs, Max_nei_eq, Max_nei_bs,
                    nadjX_eq, nadjY_eq, nadjZ_eq,
                    nadjX_bs_plus, nadjX_sb_plus, nadjX_bs_minus, nadjX_sb_minus,
                    nadjY_bs_plus, nadjY_sb_plus, nadjY_bs_minus, nadjY_sb_minus,
                    nadjZ_bs_plus, nadjZ_sb_plus, nadjZ_bs_minus, nadjZ_sb_minus,
                    rank_map,
                    MPI_cpu, MPI_cube, MPI_cpu_adj, MPI_cube_adj, MPI_direction, MPI_interface,
                    neighbor_cpu_eq, Ncube_Ncpu_eq, neighbor_cpu_sb, Ncube_Ncpu_sb,
                    neighbor_cpu_bs, Ncube_Ncpu_bs,
                    Scube_Ncpu_eq, Rcube_Ncpu_eq, send_data_curr_eq, recv_data_curr_eq,
                    send_data_neig_eq, recv_data_neig_eq, Sdir_eq, Rdir_eq,
                    Scube_Ncpu_sb, Rcube_Ncpu_sb, send_data_curr_sb, recv_data_curr_sb,
                    send_data_neig_sb, recv_data_neig_sb, Sdir_sb, Rdir_sb,
                    Scube_Ncpu_bs, Rcube_Ncpu_bs, send_data_curr_bs, recv_data_curr_bs,
                    send_data_neig_bs, recv_data_neig_bs, Sdir_bs, Rdir_bs,
                    ist_eq, ist_sb, ist_bs, adjN_sb, RadjN_bs, SadjN_bs,
                    csl,
                    adj_number,
                    adjX_eq, adjY_eq, adjZ_eq,
                    adjX_bs_plus, adjX_sb_plus, adjX_bs_minus, adjX_sb_minus,
                    adjY_bs_plus, adjY_sb_plus, adjY_bs_minus, adjY_sb_minus,
                    adjZ_bs_plus, adjZ_sb_plus, adjZ_bs_minus, adjZ_sb_minus,
                    U1_);
    }
  }
  BCM_Physical_time_shift(U1_, U1, U1q);
  BCM_Output(myid, Ncube, 1, &rank_map[0][0], U1_, U1q, cube_size, Xcnt, Ycnt, Zcnt);
  bcm_sync_to_host(eng);
  bcm_engine_destroy(eng);
  return 0;
}
