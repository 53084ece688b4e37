/*
 * bcm_cgli.h -- C ABI of the MI355X BCM solver library (libbcm_cgli.so).
 *
 * Three groups of entry points:
 *   1. Drop-ins with the names and parameter order of the reference's BCM_*
 *      routines (main.h prototypes).  They run on the device mirror of host
 *      arrays registered with bcm_bind(); tables passed per call are copied to
 *      the device on first use and reused while their host pointer and length
 *      stay the same.  Like the reference (MPI_ERRORS_ARE_FATAL) they return
 *      nothing and abort the process with a message on any error.
 *   2. Host case construction (lattice, flags, IB tables, initial state).
 *   3. A device engine that keeps one case resident and runs the reference
 *      stage order, optionally as a captured HIP graph.
 * Every function of groups 2 and 3 returns 0 on success and non-zero on
 * failure, with the message in bcm_last_error().  Array arguments use the
 * reference layouts of Array.h:11-36: state [ncube][20][20][20][5] doubles
 * (k fastest, variable innermost), flags [ncube][20][20][20] ints, cubes
 * numbered from 1 (slot 0 unused).  One host thread per engine.
 */
#ifndef BCM_CGLI_H
#define BCM_CGLI_H

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the most recent failure on the calling thread. */
const char* bcm_last_error(void);

/* ------------------------------------------------------------------------ */
/* 1. Drop-in entry points.                                                   */

typedef struct bcm_engine bcm_engine;

/* Registers the caller's state arrays (Array.h:11-23), flags, cube sizes and
 * flow constants with engine e, uploads them, and makes e the engine the
 * drop-ins below run on.  The constants replace prm.h:25-37 (rho0, U0, Uin,
 * V0, W0, P0, mu_L; T0 = P0/rho0/R) and Resolution.h:7 (n_cubes = MPI_Ncube).
 * No counterpart in the reference, whose arrays live in host memory only. */
int bcm_bind(bcm_engine* e, int ncube, int n_cubes_norm, double* U1_, double* U1, double* U1q,
             double* U1p1, double* U1p2, double* Rku1, double* Residual1, int* FWS,
             double* cube_size, const double* phys7);
/* Copies the device mirrors of the bound arrays back to / from the host. */
int bcm_sync_to_host(bcm_engine* e);
int bcm_sync_to_device(bcm_engine* e);

/* Replaces BCM_X_boundary_condition.cpp:19-122 (main.h:418-436). */
void BCM_X_boundary_condition(int myid, int nXbc_l, int nXbc_u, int* Xbc_l, int* Xbc_u,
                              double* U1_);
/* Replaces BCM_Y_boundary_condition.cpp:17-88 (main.h:438-453). */
void BCM_Y_boundary_condition(int nYbc_l, int nYbc_u, int* Ybc_l, int* Ybc_u, double* U1_);
/* Replaces BCM_Z_boundary_condition.cpp:20-96 (main.h:455-470). */
void BCM_Z_boundary_condition(int nZbc_l, int nZbc_u, int* Zbc_l, int* Zbc_u, double* U1_);
/* Replace BCM_Ghostcell_minus.cpp:17-495 and BCM_Ghostcell_plus.cpp:16-488
 * (main.h:378-413).  GCindex/IPsur are 1-based [4*NNBC+1], weight [8*NNBC+1]. */
void BCM_Ghostcell_minus(int myid, int* NNBC, double* weight, int* GCindex, int* IPsur,
                         int* FWS, double* U1_);
void BCM_Ghostcell_plus(int myid, int* NNBC, double* weight, int* GCindex, int* IPsur,
                        int* FWS, double* U1_);
/* Replaces BCM_Interface_table.cpp:20-308 (main.h:170-219); host-side table
 * construction.  adj_number is [ncube][5][7]; lists are 1-based [ncube]. */
void BCM_Interface_table(int myid, int ncube, int n_wallcube, int* nadjX_eq, int* nadjY_eq,
                         int* nadjZ_eq, int* nadjX_bs_plus, int* nadjX_sb_plus,
                         int* nadjX_bs_minus, int* nadjX_sb_minus, int* nadjY_bs_plus,
                         int* nadjY_sb_plus, int* nadjY_bs_minus, int* nadjY_sb_minus,
                         int* nadjZ_bs_plus, int* nadjZ_sb_plus, int* nadjZ_bs_minus,
                         int* nadjZ_sb_minus, int* csl, int* adj_number, int* adjX_eq,
                         int* adjY_eq, int* adjZ_eq, int* adjX_bs_plus, int* adjX_sb_plus,
                         int* adjX_bs_minus, int* adjX_sb_minus, int* adjY_bs_plus,
                         int* adjY_sb_plus, int* adjY_bs_minus, int* adjY_sb_minus,
                         int* adjZ_bs_plus, int* adjZ_sb_plus, int* adjZ_bs_minus,
                         int* adjZ_sb_minus);
/* Replaces BCM_Interface.cpp:27-5735 (main.h:224-354) on one device: the
 * equal-size halo copies, the two-level faces of the bs lists (injection into
 * the small cubes' ghost layers, 8-cell restriction into the big cube's;
 * :1790-2600 for x, y and z alike) and the edge-ghost averaging.  The MPI
 * arguments must describe a single rank (neighbour-rank counts ncpu_* zero,
 * else the call aborts); the sb lists serve only the reference's exchanges
 * with other ranks and are ignored. */
void BCM_Interface(int myid, int ncube, int mPI_Nadj, int ncpu_bs, int ncpu_eq, int ncpu_sb,
                   int max_nei_bs, int max_nei_eq, int max_nei_sb, int nadjX_eq, int nadjY_eq,
                   int nadjZ_eq, int nadjX_bs_plus, int nadjX_sb_plus, int nadjX_bs_minus,
                   int nadjX_sb_minus, int nadjY_bs_plus, int nadjY_sb_plus, int nadjY_bs_minus,
                   int nadjY_sb_minus, int nadjZ_bs_plus, int nadjZ_sb_plus, int nadjZ_bs_minus,
                   int nadjZ_sb_minus, int* rank_map, int* MPI_cpu, int* MPI_cube,
                   int* MPI_cpu_adj, int* MPI_cube_adj, int* MPI_direction, int* MPI_interface,
                   int* neighbor_cpu_eq, int* Ncube_Ncpu_eq, int* neighbor_cpu_sb,
                   int* Ncube_Ncpu_sb, int* neighbor_cpu_bs, int* Ncube_Ncpu_bs,
                   int* Scube_Ncpu_eq, int* Rcube_Ncpu_eq, double* send_data_curr_eq,
                   double* recv_data_curr_eq, double* send_data_neig_eq,
                   double* recv_data_neig_eq, int* Sdir_eq, int* Rdir_eq, int* Scube_Ncpu_sb,
                   int* Rcube_Ncpu_sb, double* send_data_curr_sb, double* recv_data_curr_sb,
                   double* send_data_neig_sb, double* recv_data_neig_sb, int* Sdir_sb,
                   int* Rdir_sb, int* Scube_Ncpu_bs, int* Rcube_Ncpu_bs,
                   double* send_data_curr_bs, double* recv_data_curr_bs,
                   double* send_data_neig_bs, double* recv_data_neig_bs, int* Sdir_bs,
                   int* Rdir_bs, int* ist_eq, int* ist_sb, int* ist_bs, int* adjN_sb,
                   int* RadjN_bs, int* SadjN_bs, int* csl, int* adj_number, int* adjX_eq,
                   int* adjY_eq, int* adjZ_eq, int* adjX_bs_plus, int* adjX_sb_plus,
                   int* adjX_bs_minus, int* adjX_sb_minus, int* adjY_bs_plus,
                   int* adjY_sb_plus, int* adjY_bs_minus, int* adjY_sb_minus,
                   int* adjZ_bs_plus, int* adjZ_sb_plus, int* adjZ_bs_minus,
                   int* adjZ_sb_minus, double* U1_);
/* Replaces BCM_Flux_XYZ_Viscous_Runge_kutta.cpp:18-2658 (main.h:473-513):
 * flux, residual and preconditioned stage update of every cube; at RK == 3
 * also the change norms and force sums, returned in er[1..7]. */
void BCM_Flux_XYZ_Viscous_Runge_kutta(int myid, int ncube, int RK, double deltaT,
                                      double deltaTau, double e, int* FWS, int* csl,
                                      double* cube_size, double* U1_, double* U1, double* U1q,
                                      double* U1p1, double* U1p2, double* Rku1,
                                      double* Residual1, double* er);
/* Physical-time shift U1q <- U1, U1 <- U1_ on interior cells, which the
 * reference does inline in its main loop (main.cpp:1292-1315). */
void BCM_Physical_time_shift(double* U1_, double* U1, double* U1q);
/* Replaces BCM_Grid.cpp:18-347 (main.h:25-43): reads BCMgrid<myid>.dat from
 * the working directory into the caller's arrays (cubes 1..ncube-1, slots
 * 1..4 of adjacency directions 1..6, wall cubes 1..n_wallcube-1); cell
 * centres and size levels follow from the cube sizes and corners. */
void BCM_Grid(int myid, int ncube, int n_wallcube, double* cube_size, int* csl, double* Xcube,
              double* Ycube, double* Zcube, double* Xcnt, double* Ycnt, double* Zcnt, int* FWS,
              int* adj_number, int* wallcube);
/* Replaces BCM_Reading_IBM.cpp:15-77 (main.h:359-373): reads *NNBC entries
 * of IBM_plus<myid>.dat (mp_switch 1) or IBM_minus<myid>.dat (mp_switch -1)
 * into the 1-based GCindex[4n+1], IPsur[4n+1] and weight[8n+1]. */
void BCM_Reading_IBM(int myid, int* NNBC, int mp_switch, double* weight, int* GCindex,
                     int* IPsur);
/* Replaces BCM_Output.cpp:17-303 (main.h:515-533) for one rank: writes
 * BCMdata00001.x and qBCMdata<step>.q of U1 (the caller passes U1_, as
 * main.cpp:1348 does) into the working directory.  When U1 is the bound
 * engine's U1_, the device state is copied into it first.  rank_map, U1q
 * and cube_size are accepted and unused, as in the reference. */
void BCM_Output(int myid, int ncube, int step, int* rank_map, double* U1, double* U1q,
                double* cube_size, double* Xcnt, double* Ycnt, double* Zcnt);

/* ------------------------------------------------------------------------ */
/* 2. Host case construction.                                                 */

typedef struct bcm_case_params {
  int lattice[3];             /* cubes along x, y, z */
  double cube_size;           /* cube edge length */
  double origin[3];           /* lower box corner */
  double sphere_center[3];
  double sphere_diameter;
  int ico_level;              /* icosphere level: 20*4^L triangles */
  double mach, reynolds;
  unsigned long long seed;    /* density-noise seed */
  double noise_amp;           /* relative density noise amplitude */
  int refine_lo[3];           /* lattice cubes with index in [refine_lo, refine_hi) */
  int refine_hi[3];           /* on every axis are split into 8 (a two-level grid) */
  int refine2_lo[3];          /* their children with index on the doubled lattice in */
  int refine2_hi[3];          /* [refine2_lo, refine2_hi) are split again (three levels) */
} bcm_case_params;

typedef struct bcm_case bcm_case;

int bcm_case_build(const bcm_case_params* p, bcm_case** out);
/* bcm_case_build with the two quadratic passes of the IB setup (triangle
 * filter, BCM_Immersed_boundary.cpp:380-470; surface search per ghost cell,
 * :722-860) as kernels on GPU `device`; the case is bitwise the one
 * bcm_case_build gives. */
int bcm_case_build_on(const bcm_case_params* p, int device, bcm_case** out);
void bcm_case_free(bcm_case* c);
/* info[16]: ncube, n_wallcube, nbc_minus, nbc_plus, nXbc_l, nXbc_u, nYbc_l,
 * nYbc_u, nZbc_l, nZbc_u, nadjX_eq, nadjY_eq, nadjZ_eq, n_triangles,
 * n_ghost_candidates, 0.  phys[8]: rho0, U0, Uin, V0, W0, P0, T0, mu_L. */
int bcm_case_info(const bcm_case* c, int* info, double* phys);
/* Pointer to a case array and its element count.  Names: "U1_", "U1", "U1q",
 * "cube_size", "Xcube", "Ycube", "Zcube", "Xcnt", "Ycnt", "Zcnt" (double); "FWS",
 * "FWS_grid" (the flags as the grid file holds them), "csl",
 * "adj_number", "ib_minus_gc", "ib_minus_ip", "ib_plus_gc", "ib_plus_ip",
 * "Xbc_l", "Xbc_u", "Ybc_l", "Ybc_u", "Zbc_l", "Zbc_u", the 15 interface lists of
 * BCM_Interface_table.cpp ("adjX_eq", "adjY_eq", "adjZ_eq", "adjX_bs_plus",
 * "adjX_sb_plus", "adjX_bs_minus", "adjX_sb_minus", then the same for Y and
 * Z), "nadj" (their 15 lengths) (int); "ib_minus_w", "ib_plus_w" (double). */
void* bcm_case_array(bcm_case* c, const char* name, long long* count);
/* Writes BCMgrid00000.dat and BCM_STL_fine.stl, the reference's input files,
 * and the rank-decomposition files it reads at start-up (main.cpp:213-221,
 * 262-554, 584-596) for a decomposition into one rank:
 * MPI_communication_table_in_turn.dat (0 cross-rank connections),
 * Communication_irank.00000.dat (no neighbour ranks) and rank_map.dat (every
 * cube on rank 0, local number = global number + 1). */
int bcm_case_write_inputs(const bcm_case* c, const char* dir);
/* Builds a case from the reference's input files in dir: the grid file
 * BCMgrid00000.dat as BCM_Grid.cpp:63-261 reads it (cube sizes and corners,
 * adjacency, wall cubes, run-length coded flags) and the ASCII STL surface
 * BCM_STL_fine.stl as BCM_Immersed_boundary.cpp:198-275 reads it; then the
 * setup of main.cpp:108-1139.  p supplies the flow parameters the reference
 * keeps in prm.h: mach, reynolds, sphere_diameter (the reference length of
 * mu_L), seed and noise_amp; its lattice, cube_size, origin, sphere_center
 * and ico_level are not used.  Rank-decomposition files present in dir are
 * checked: a decomposition over several ranks (cross-rank connections,
 * neighbour ranks, cubes on another rank or a rank map of another cube
 * count) is refused with an error, since one GPU runs a one-rank case. */
int bcm_case_read(const char* dir, const bcm_case_params* p, bcm_case** out);
/* bcm_case_read with the IB search on GPU `device` (as bcm_case_build_on). */
int bcm_case_read_on(const char* dir, const bcm_case_params* p, int device, bcm_case** out);
/* Writes the IB tables as IBM_minus00000.dat and IBM_plus00000.dat in the
 * byte format of BCM_Immersed_boundary.cpp:1009-1065 (read back by
 * BCM_Reading_IBM.cpp:48-75). */
int bcm_case_write_ibm(const bcm_case* c, const char* dir);
/* Writes the PLOT3D grid BCMdata00001.x and solution qBCMdata<step>.q of the
 * state aos ((ncube) cubes of 20^3 x 5 doubles; null: the case's U1_) in the
 * byte layout BCM_Output.cpp:86-301 writes on one rank (replaces BCM_Output,
 * main.h, called at main.cpp:1346-1351).  step in 0..99999. */
int bcm_case_write_plot3d(const bcm_case* c, const double* aos, int step, const char* dir);
/* Restart snapshot (the reference has none: its reader,
 * BCM_Initial_condition.cpp:91-227, is disabled, and its output omits U1q):
 * step, the norm scales E1..E5 (E5 may be null: zeros on write, ignored on
 * read) and U1_, U1, U1q, each ncube cubes of 20^3 x 5 doubles.  A read
 * fails unless the file holds exactly ncube cubes. */
int bcm_write_restart(const char* path, int ncube, long long step, const double* E5,
                      const double* U1_, const double* U1, const double* U1q);
int bcm_read_restart(const char* path, int ncube, long long* step, double* E5, double* U1_,
                     double* U1, double* U1q);

/* ------------------------------------------------------------------------ */
/* 3. Device engine.                                                          */

int bcm_engine_create(int device, bcm_engine** out);
void bcm_engine_destroy(bcm_engine* e);
/* Uploads a host case: state, flags and every table. */
int bcm_engine_load_case(bcm_engine* e, const bcm_case* c);
/* Physical and pseudo time steps and the preconditioning floor e
 * (main.cpp:57-59 defaults 1e-4, 1e-4/200, 0.1). */
int bcm_engine_set_time(bcm_engine* e, double deltaT, double deltaTau, double e_floor);
/* n_steps physical steps of n_inner inner iterations (3 RK stages each) of
 * every resident case, each step closed by the physical-time shift.
 * er_hist, if not null, receives er[1..7] of every case slot after every
 * inner iteration: [n_steps][n_inner][max_cases][7] doubles (max_cases is 1
 * after bcm_engine_load_case). */
int bcm_engine_run(bcm_engine* e, int n_steps, int n_inner, double* er_hist, int use_graph);
/* n_steps physical steps under the reference's inner-iteration control, the
 * loop of main.cpp:1146-1320: per case, after every inner iteration
 * E[m] = max(E[m], er[m]) and er[m] /= E[m] (m = 1..5); the inner loop ends
 * when all five are below tol or at iteration max_inner, and the case's time
 * levels shift (the reference's tol is 0.02, max_inner its
 * iteration_end_step).  Each case decides for itself; a case that ends before
 * others is shifted at once and its iterate is kept aside until the step ends.
 * E: [max_cases][5] running maxima E1..E5, read and updated (zeros at the
 * start of a run, as main.cpp:66-70).  iters: [n_steps][max_cases] inner
 * iterations of each step (0 for empty slots).  er_hist (may be null):
 * [n_steps][max_inner][max_cases][7] doubles, er[1..7] of the iterations each
 * case ran; rows after its break are left untouched. */
int bcm_engine_run_converge(bcm_engine* e, int n_steps, int max_inner, double tol, double* E,
                            int* iters, double* er_hist);
/* One piece of a stage: 0 domain BCs, 1 IB ghost fill, 2 halo + edges,
 * 3 flux, 4 RK update (with norms at RK 3), 5 time shift, 6 whole stage,
 * 7 flux and RK update as a stage runs them (fused when the choice allows),
 * 8 the pre-flux halo refresh a stage uses (IB ghost copies + edge passes),
 * 9 flux and RK update written to the other copy of the iterate together
 * with the face halos (the stage form of flux choice fuse 2). */
int bcm_engine_part(bcm_engine* e, int part, int RK);
/* which: 0 U1_, 1 U1, 2 U1q, 3 U1p1, 4 U1p2, 5 Rku1, 6 Residual1. */
int bcm_engine_get_state(bcm_engine* e, int which, double* aos);
int bcm_engine_set_state(bcm_engine* e, int which, const double* aos);
/* er[0..7] as last reduced (er[1..7] meaningful). */
int bcm_engine_get_er(bcm_engine* e, double* er8);
int bcm_engine_sync(bcm_engine* e);
/* Average device time in ms over reps launches of one part (events on the
 * engine stream). */
int bcm_engine_time_part(bcm_engine* e, int part, int RK, int reps, double* ms);
/* Flux kernel choice: variant 0 (one workgroup per cell plane) or 1 (one
 * workgroup per cube segment marching along i, nseg in {1,2,4,8,16}); fuse 1
 * lets variant 1 with nseg 1 apply the RK update in the same kernel, in
 * place; fuse 2 (variant 1, nseg 1) writes the update and the face halos into
 * a second copy of the iterate in engine stages, which then open with the
 * edge passes and BCs instead of closing with the halo exchange.  All choices
 * give bitwise the same results.  Setting a choice disables tuning. */
int bcm_engine_set_flux(bcm_engine* e, int variant, int nseg, int fuse);
/* Times the flux choices (one whole RK stage each) on the resident cases,
 * restoring the state afterwards, and keeps the fastest; runs automatically
 * before the first bcm_engine_run otherwise.  out3 (may be null) receives
 * {variant, nseg, fuse}. */
int bcm_engine_autotune(bcm_engine* e, int* out3);
/* Ghost schedule statistics: entries, cubes with entries, maximum levels. */
int bcm_engine_ghost_stats(bcm_engine* e, int* stats3);
/* The IB fill of a cube runs out of LDS when its value slots fit the limit
 * (at most 1536, the default), else in place on memory; a lower limit,
 * applied to cases loaded or added afterwards, puts more cubes on the memory
 * schedule (bitwise the same results).  Not a reference interface: a test
 * and diagnostic hook. */
int bcm_engine_set_ghost_lds_limit(bcm_engine* e, int slots);
/* The FP64 kernels divide several numerators by one denominator with the
 * denominator's part of the correctly rounded division formed once.  This
 * runs n threads of four such quotients on random operands against plain
 * division on the current device; *mismatches receives the number of quotients
 * whose bits differ (0 expected).  Not a reference interface: a self-test. */
int bcm_selftest_division(unsigned long long seed, int n, unsigned long long* mismatches);

/* ------------------------------------------------------------------------ */
/* 4. Case pool: several independent cases resident on one device.           */
/* Not a reference interface: the reference runs one case per process.  The
 * cases of a pool run together, every kernel launch covering all their
 * cubes; each case keeps its own physics, tables and norms, and its results
 * are bitwise the ones it gives alone. */

/* Device memory per pool cube, bytes, for a precision (0 FP64, 1 mixed). */
long long bcm_bytes_per_cube(int precision);
/* Replaces whatever the engine holds by an empty pool of cap_cubes cubes
 * with max_cases case slots.  precision 0: plain FP64, bitwise the
 * reference.  precision 1: mixed; the inner iterate U1_, the fluxes, the
 * residual, the preconditioner and the RK stages in FP32; the time levels U1
 * and U1q, the dual-time source, the norms and the IB interpolation in FP64.
 * A mixed pool keeps only U1_, U1 and U1q (state arrays 0..2) and runs flux
 * and update as one kernel (part 7). */
int bcm_engine_alloc_pool(bcm_engine* e, int cap_cubes, int max_cases, int precision);
/* 0 FP64, 1 mixed. */
int bcm_engine_precision(bcm_engine* e);
/* Adds a case in the first free range of its cube count; *slot receives its
 * slot index.  Fails when no slot or no range is free. */
int bcm_engine_add_case(bcm_engine* e, const bcm_case* c, int* slot);
int bcm_engine_remove_case(bcm_engine* e, int slot);
/* One case's state array in the reference layout, (ncube+1) cubes of
 * 20^3 x 5 doubles (cube 0 zeroed on download, ignored on upload). */
int bcm_engine_case_state(bcm_engine* e, int slot, int which, double* aos, int upload);
/* er[0..7] of one case after inner iteration `inner` of the last step. */
int bcm_engine_case_er(bcm_engine* e, int slot, int inner, double* er8);
/* info6: pool cubes, cubes spanned, active cases, case slots, largest free
 * range, cubes in use. */
int bcm_engine_pool_info(bcm_engine* e, int* info6);

/* ------------------------------------------------------------------------ */
/* 5. Output and restart of resident cases.                                   */

/* Queues the PLOT3D output of a resident case's inner iterate U1_ (the array
 * main.cpp:1348 hands to BCM_Output) as BCMdata00001.x and qBCMdata<step>.q in
 * dir, in the bytes bcm_case_write_plot3d writes.  The engine stream pays one
 * layout conversion; the copy to the host runs on a second stream and the
 * files are written by a host thread, both overlapping later engine work.
 * At most two outputs are in flight; a third waits for the oldest. */
int bcm_engine_output(bcm_engine* e, int slot, int step, const char* dir);
/* Waits for every queued output; returns the first write error since the
 * last wait. */
int bcm_engine_output_wait(bcm_engine* e);
/* Restart snapshot of a resident case (U1_, U1, U1q, with the caller's step
 * and norm scales) in the bcm_write_restart format, and its inverse.  A
 * mixed pool saves U1_ as the FP64 sum of the freestream state and the FP32
 * deviation; loading re-rounds the deviation to FP32. */
int bcm_engine_save_restart(bcm_engine* e, int slot, const char* path, long long step,
                            const double* E5);
int bcm_engine_load_restart(bcm_engine* e, int slot, const char* path, long long* step,
                            double* E5);

/* ------------------------------------------------------------------------ */
/* 6. Flux/RK stage that allocates nothing.                                   */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the work of BCM_Flux_XYZ_Viscous_Runge_kutta (same
 * kernels, bitwise the same results) on arrays the caller owns on the
 * current device, in the library's device layout: state arrays [ncube][5][20]
 * [20][20] doubles (variable-major per cube, k fastest), flags FWS [ncube][20]
 * [20][20] int8, cube_size [ncube]; cube 0 unused.  U1_ is updated in place
 * and U1p1 (RK 1), U1p2 (RK 2), Rku1 and Residual1 written; at RK 3, er[0..7]
 * (device, 8 doubles) receives 0 and the norms and force sums er[1..7].
 * phys7 (host) holds rho0, U0, Uin, V0, W0, P0, mu_L as for bcm_bind, and
 * n_cubes_norm the norm scaling's cube count (MPI_Ncube).  Work is queued on
 * `stream` (a hipStream_t; NULL for the default stream) and not waited for.
 * The norm partials and case tables live in a fixed device array of
 * BCM_NOALLOC_SCRATCH_BYTES per device, held until the process ends, so at
 * most BCM_NOALLOC_MAX_CUBES cubes (ncube - 1) per call; larger calls fail.
 * Calls that overlap on one device share that array and race. */
#define BCM_NOALLOC_MAX_CUBES 16384
#define BCM_NOALLOC_SCRATCH_BYTES 1114624
int bcm_flux_rk_noalloc(int ncube, int n_cubes_norm, int RK, double deltaT, double deltaTau,
                        double e, const signed char* FWS, const double* cube_size, double* U1_,
                        const double* U1, const double* U1q, double* U1p1, double* U1p2,
                        double* Rku1, double* Residual1, double* er, const double* phys7,
                        void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BCM_CGLI_H */
