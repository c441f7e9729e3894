/*
 * bhah_hip.h -- C-ABI of the MI355X (gfx950) evolution backend for NRPy's
 * BlackHoles@Home (BHaH) single-patch infrastructure.
 *
 * Every entry point is the counterpart of a C function that NRPy generates
 * for a BHaH project (file:line of the reference emitter or its frozen
 * output are cited per function).  The reference's functions take
 * project-generated structs (commondata_struct, params_struct, rfm_struct,
 * bc_struct, griddata_struct) whose layouts differ per project; this ABI
 * replaces them with one opaque patch handle built from a plain descriptor.
 * A generated BHaH project binds these through a thin adapter (see the
 * integration notes shipped with the repository).
 *
 * Conventions
 *  - All REAL arrays are FP64 device pointers (hipMalloc'd or torch tensors),
 *    gf-major SoA with IDX4(gf,i0,i1,i2) = i0 + Nx0*(i1 + Nx1*(i2 + Nx2*gf)),
 *    Nxi = Nxxi + 2*NGHOSTS, xx0 fastest -- the BHaH layout
 *    (nrpy/infrastructures/BHaH/tests/BHaH_defines_h_BHaH_defines.c:149-162).
 *  - Evolved gridfunction order (24): aDD00..aDD22, alpha, betU0..2, cf,
 *    hDD00..hDD22, lambdaU0..2, trK, vetU0..2 (nrpy/grid.py:248-251);
 *    auxevol: RbarDD00..RbarDD22 (6).
 *  - Work is enqueued on the patch's stream; calls return without
 *    synchronising unless stated.
 *  - Every function returns 0 on success, a negative BHAH_ERR_* code
 *    otherwise; bhah_hip_last_error() describes the last failure.  (The
 *    reference functions are void and abort on error.)
 */
#ifndef BHAH_HIP_H
#define BHAH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BHAH_HIP_ABI_VERSION 2

enum {
  BHAH_OK = 0,
  BHAH_ERR_ARG = -1,
  BHAH_ERR_HIP = -2,
  BHAH_ERR_UNSUPPORTED = -3,
  BHAH_ERR_BCSTRUCT = -4,
};

enum { BHAH_CARTESIAN = 0, BHAH_SPHERICAL = 1, BHAH_SINHSPHERICAL = 2 };
enum { BHAH_SYSTEM_BSSN = 0, BHAH_SYSTEM_WAVE = 1 };
enum { BHAH_OUTER_BC_RADIATION = 0, BHAH_OUTER_BC_EXTRAPOLATION = 1 };
enum { BHAH_KERNEL_RHS = 0, BHAH_KERNEL_RICCI = 1 };

#define BHAH_NUM_EVOL_GFS_BSSN 24
#define BHAH_NUM_AUXEVOL_GFS_BSSN 6
#define BHAH_NUM_EVOL_GFS_WAVE 2
/* Diagnostic gridfunctions written by the constraint kernels, in this order: H, MU0, MU1,
 * MU2, MSQUARED, LAMBDA_CONSTRAINT. */
#define BHAH_NUM_DIAG_GFS 6
enum { BHAH_DIAG_H = 0, BHAH_DIAG_MU0 = 1, BHAH_DIAG_MSQUARED = 4, BHAH_DIAG_LAMBDA_CONSTRAINT = 5 };

/* Plain descriptor replacing commondata_struct/params_struct for one patch. */
typedef struct bhah_patch_desc {
  int system;                /* BHAH_SYSTEM_BSSN or BHAH_SYSTEM_WAVE */
  int coord;                 /* BHAH_CARTESIAN, BHAH_SPHERICAL, BHAH_SINHSPHERICAL */
  int Nxx[3];                /* interior points per direction */
  int fd_order;              /* 4 or 6 */
  int symmetry_axes;         /* bit d set: derivatives along xx_d vanish (e.g. 4 = phi) */
  int enable_KO;             /* Kreiss-Oliger dissipation on (BSSN) */
  int outer_bc;              /* BHAH_OUTER_BC_* (BSSN) */
  int radiation_bc_fd_order; /* FD order of d/dr in the radiation BC (reference default 4) */
  double grid_physical_size; /* RMAX (Spherical), AMPL (SinhSpherical), half-width (Cartesian) */
  double SINHW;              /* SinhSpherical width */
  double CFL_FACTOR;         /* dt = CFL_FACTOR * min ds */
  double eta;                /* Gamma-driver damping */
  double KO_strength_gauge, KO_strength_nongauge;
  double wavespeed;          /* scalar wave speed */
  int setup_on_device;       /* 0: set-up on the host (the reference's loops); 1: on the device */
} bhah_patch_desc;

typedef struct bhah_patch bhah_patch;

/* MoL gridfunction arrays (MoL_gridfunctions_struct,
 * nrpy/infrastructures/BHaH/MoLtimestepping/MoL_malloc_intermediate_stage_gfs.py:20). */
typedef struct bhah_MoL_gridfunctions {
  double *y_n_gfs;
  double *y_nplus1_running_total_gfs;
  double *k_odd_gfs;
  double *k_even_gfs;
  double *auxevol_gfs;
} bhah_MoL_gridfunctions;

/* The time members of commondata_struct that MoL_step_forward_in_time updates
 * (nrpy/infrastructures/BHaH/MoLtimestepping/MoL_step_forward_in_time.py:473-481;
 * t_0 and nn_0 are reset after reading a checkpoint, read_checkpoint__openmp.c:183-185). */
typedef struct bhah_clock {
  double time, t_0, dt;
  int nn, nn_0;
} bhah_clock;

/* Volume integrals of the reference's two diagnostics recipes
 * (nrpy/infrastructures/BHaH/general_relativity/diagnostics_volume_integration.py:177-207):
 * recipe 0 "whole_domain", recipe 1 "sphere_R_gt_8" (points with |x| > 8 about the origin).
 * Integrand k is f_k^2 for f = H, MSQUARED, LAMBDA_CONSTRAINT (is_squared = 1), with the
 * proper volume element dV = sqrt(det gammahat) |dxx0 dxx1 dxx2| at interior points
 * (diagnostics_volume_integration_helpers.h:396-476, sqrt_detgammahat_d3xx_volume_element.py:33-130).
 * L2 = sqrt(integral), RMS = sqrt(integral / volume), 0 when the volume is 0
 * (diagnostics_volume_integration_helpers.h:332-345). */
#define BHAH_DIAG_NUM_RECIPES 2
#define BHAH_DIAG_NUM_INTEGRANDS 3
typedef struct bhah_volume_integrals {
  double volume[BHAH_DIAG_NUM_RECIPES];
  double integral[BHAH_DIAG_NUM_RECIPES][BHAH_DIAG_NUM_INTEGRANDS];
  double L2[BHAH_DIAG_NUM_RECIPES][BHAH_DIAG_NUM_INTEGRANDS];
  double RMS[BHAH_DIAG_NUM_RECIPES][BHAH_DIAG_NUM_INTEGRANDS];
} bhah_volume_integrals;

/* Grid geometry of a created patch (numerical_grid_params_Nxx_dxx_xx). */
typedef struct bhah_patch_info {
  int Nxx[3], nghosts, Nxx_plus_2NGHOSTS[3]; /* lower-case: a project #defines NGHOSTS */
  int num_evol_gfs, num_auxevol_gfs;
  int64_t Ntot;       /* points per gridfunction, ghost zones included */
  double xxmin[3], xxmax[3], dxx[3], invdxx[3];
  double dt;          /* CFL-limited timestep */
  int num_inner_boundary_points;
  int num_pure_outer_boundary_points;
} bhah_patch_info;

const char *bhah_hip_last_error(void);
int bhah_hip_abi_version(void);

/* Set up grid, reference-metric precompute arrays, boundary structs and CFL
 * timestep on the current HIP device.  Replaces numerical_grids_and_timestep
 * (nrpy/infrastructures/BHaH/numerical_grids_and_timestep.py:31-760),
 * rfm_precompute (nrpy/infrastructures/BHaH/rfm_precompute.py:84-400) and
 * bcstruct_set_up (nrpy/infrastructures/BHaH/CurviBoundaryConditions/bcstruct_set_up.py:38-792).
 * With desc->setup_on_device the per-point work runs on the device: the reference-metric
 * arrays, the eigen-coordinate classification of every ghost point with its parity signs,
 * order-preserving compaction into the inner and pure-outer lists (the same lists, in the
 * same order, as the host loops), and the CFL min-reduction; only the O(surface) lists and
 * the 1D arrays come back to the host. */
int bhah_hip_patch_create(const bhah_patch_desc *desc, bhah_patch **out);
int bhah_hip_patch_destroy(bhah_patch *patch);
int bhah_hip_patch_get_info(const bhah_patch *patch, bhah_patch_info *info);
/* Use an existing hipStream_t (e.g. torch.cuda.current_stream().cuda_stream). */
int bhah_hip_patch_set_stream(bhah_patch *patch, void *hip_stream);
int bhah_hip_synchronize(bhah_patch *patch);
/* Host copies of coordinates (3 arrays of Nxx_plus_2NGHOSTS[d]) and, for
 * tests, of the boundary structs. */
int bhah_hip_patch_get_xx(const bhah_patch *patch, double *xx0, double *xx1, double *xx2);
int bhah_hip_patch_get_inner_bc(const bhah_patch *patch, int64_t *dstpt, int64_t *srcpt, int8_t *parity10);
int bhah_hip_patch_get_outer_bc_count(const bhah_patch *patch, int which_gz, int dirn);
int bhah_hip_patch_get_outer_bc(const bhah_patch *patch, int which_gz, int dirn, int16_t *i0i1i2_face6);
/* Host copies of the reference-metric precompute arrays (rfm_precompute.py:196-340):
 * rfm0[k * Nx0 + i0] = d^k f0/dxx0^k (k = 0..3, f0 = r(xx0)), rfm1[k * Nx1 + i1] = d^k f1/dxx1^k
 * (k = 0..2, f1 = sin(xx1)). */
int bhah_hip_patch_get_rfm(const bhah_patch *patch, double *rfm0, double *rfm1);

/* rhs_eval (nrpy/infrastructures/BHaH/general_relativity/rhs_eval.py:97-120; call
 * nrpy/examples/two_blackholes_collide.py:294): BSSN right-hand sides at interior points. */
int bhah_hip_rhs_eval(bhah_patch *patch, const double *auxevol_gfs, const double *in_gfs, double *rhs_gfs);

/* Ricci_eval (nrpy/infrastructures/BHaH/general_relativity/Ricci_eval.py:81-93; call
 * two_blackholes_collide.py:292): RbarDD into auxevol_gfs at interior points. */
int bhah_hip_Ricci_eval(bhah_patch *patch, const double *in_gfs, double *auxevol_gfs);

/* enforce_detgammabar_equals_detgammahat (…/enforce_detgammabar_equals_detgammahat.py:63-83;
 * call two_blackholes_collide.py:307): all points, in place. */
int bhah_hip_enforce_detgammabar_equals_detgammahat(bhah_patch *patch, double *in_gfs);

/* apply_bcs_outerradiation_and_inner (CurviBoundaryConditions/tests/
 * apply_bcs_outerradiation_and_inner__rfm__openmp__HoleySinhSpherical.c:211-237):
 * Sommerfeld RHS at pure outer points, then inner parity on rhs_gfs. */
int bhah_hip_apply_bcs_outerradiation_and_inner(bhah_patch *patch, const double *gfs, double *rhs_gfs);

/* apply_bcs_outerextrap_and_inner (…/apply_bcs_outerextrap_and_inner__openmp.c:80-91). */
int bhah_hip_apply_bcs_outerextrap_and_inner(bhah_patch *patch, double *gfs);

/* apply_bcs_inner_only (…/apply_bcs_inner_only_apply_bcs_inner_only__openmp.c:58-70). */
int bhah_hip_apply_bcs_inner_only(bhah_patch *patch, double *gfs);

/* MoL_malloc_intermediate_stage_gfs (MoLtimestepping/MoL_malloc_intermediate_stage_gfs.py:20):
 * allocates y_nplus1_running_total, k_odd and k_even on the device (zeroed); y_n_gfs and
 * auxevol_gfs are allocated too when they are NULL. */
int bhah_hip_MoL_malloc_intermediate_stage_gfs(bhah_patch *patch, bhah_MoL_gridfunctions *gf);
int bhah_hip_MoL_free(bhah_patch *patch, bhah_MoL_gridfunctions *gf);

/* MoL_step_forward_in_time, RK4 (MoLtimestepping/MoL_step_forward_in_time.py:39-507; frozen
 * MoL_step_forward_in_time_CUDA__MoL_step_forward_in_time__RK4.cu:182-287): one full step of
 * size clock->dt, then the reference's bookkeeping
 * time = t_0 + (nn - nn_0 + 1) * dt; nn++ (MoL_step_forward_in_time.py:473-481). */
int bhah_hip_MoL_step_forward_in_time(bhah_patch *patch, bhah_MoL_gridfunctions *gf, bhah_clock *clock);
/* nsteps steps, each followed by the same clock update; with use_graph != 0 the four
 * substeps are captured once in a hipGraph and replayed (no HIP API work per step beyond the
 * graph launch). */
int bhah_hip_MoL_step_n(bhah_patch *patch, bhah_MoL_gridfunctions *gf, bhah_clock *clock, int nsteps, int use_graph);

/* Kernel schedules.  rhs_eval and Ricci_eval each come in several generated
 * schedules (partitions of the outputs into kernels x register budgets); all
 * compute the same values.  The autotuner (nrpy_amd/autotune.py) times them on
 * the device and stores the fastest per workload.  No reference counterpart:
 * the reference emits one kernel per function
 * (nrpy/helpers/parallelization/utilities.py:108-196). */
int bhah_hip_patch_num_schedules(const bhah_patch *patch, int kind);
const char *bhah_hip_patch_schedule_name(const bhah_patch *patch, int kind, int index);
int bhah_hip_patch_select_schedule(bhah_patch *patch, int kind, int index);
int bhah_hip_patch_get_schedule(const bhah_patch *patch, int kind);
/* Interior thread-block shape (tile_x * tile_y = 256, tile_x along xx0); replaces the
 * reference's BHAH_THREADS_IN_{X,Y,Z}_DIR_* macros (nrpy/helpers/parallelization/gpu_kernel.py:131-158). */
int bhah_hip_patch_set_tile(bhah_patch *patch, int tile_x, int tile_y);

/* Wave equation: rhs_eval (nrpy/examples/wave_equation_cartesian.py:318-357) and quadratic
 * extrapolation apply_bcs (wave_equation_cartesian.py:374-431). */
int bhah_hip_wave_rhs_eval(bhah_patch *patch, const double *in_gfs, double *rhs_gfs);
int bhah_hip_wave_apply_bcs(bhah_patch *patch, double *gfs);

/* Constraint diagnostics: constraints_eval (nrpy/infrastructures/BHaH/general_relativity/
 * constraints_eval.py; math nrpy/equations/general_relativity/BSSN_constraints.py:82-203) at
 * interior points into diag (device, BHAH_NUM_DIAG_GFS * Ntot): H, MU^i, M^2 =
 * gammabar_ij MU^i MU^j e^{4 phi} and |Lambdabar^i - Delta^i| = sqrt(gammabar_ij dL^i dL^j)
 * (the reference's DIAG_HAMILTONIAN, DIAG_MSQUARED and DIAG_LAMBDA_CONSTRAINT,
 * diagnostic_gfs_set.py:75-85).  Needs RbarDD in auxevol_gfs (bhah_hip_Ricci_eval on in_gfs
 * first).  If integrals is not NULL it receives the two recipes' volume integrals
 * (diagnostics_volume_integration.py:177-207), reduced on the device in a fixed order, so
 * repeated calls give bitwise equal results; this synchronises the stream. */
int bhah_hip_diagnostics_constraints(bhah_patch *patch, const double *auxevol_gfs, const double *in_gfs, double *diag,
                                     bhah_volume_integrals *integrals);

/* The same constraint diagnostics and volume integrals, allocating nothing and never
 * synchronising: no counterpart in the reference.  It is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * reduction's partial sums go to one 64 KiB device array per GPU that the library holds
 * until the process ends (whatever the grid size, at most 1024 x 8 partial sums), so calls
 * that overlap on one GPU (other streams, other threads) race and must be serialised by the
 * caller.  integral_sums (device, 2 * 4 doubles) receives per recipe the proper volume and
 * the integrals of H^2, MSQUARED^2 and LAMBDA_CONSTRAINT^2 (the volume and integral members of
 * bhah_volume_integrals, bitwise equal to bhah_hip_diagnostics_constraints'; L2 and RMS are
 * their square roots).  diag as in bhah_hip_diagnostics_constraints. */
int bhah_hip_diagnostics_constraints_noalloc(bhah_patch *patch, const double *auxevol_gfs, const double *in_gfs,
                                             double *diag, double *integral_sums);

/* diagnostics (nrpy/infrastructures/BHaH/diagnostics/diagnostics.py:46-270) for one grid: on
 * output steps, fabs(round(time / every) * every - time) < dt / 2 (diagnostics.py:195), it
 * evaluates Ricci_eval on y_n into auxevol_gfs, the constraint diagnostics into diag and the
 * volume integrals into *integrals, and sets *output_step = 1; otherwise it does nothing and
 * sets *output_step = 0.  File output stays with the caller (the reference's
 * diags_integration_write_row columns are volume, integral, L2, RMS). */
int bhah_hip_diagnostics(bhah_patch *patch, const bhah_clock *clock, double diagnostics_output_every,
                         const double *y_n_gfs, double *auxevol_gfs, double *diag, bhah_volume_integrals *integrals,
                         int *output_step);
/* Device buffer for the diagnostic gridfunctions (BHAH_NUM_DIAG_GFS * Ntot doubles); the
 * reference's driver allocates its diagnostic_gfs the same size (diagnostics.py:204-209, 249). */
int bhah_hip_diagnostic_gfs_malloc(const bhah_patch *patch, double **diag);
int bhah_hip_diagnostic_gfs_free(double *diag);

/* Checkpoints (write_checkpoint / read_checkpoint, BHaH/tests/write_checkpoint_write_checkpoint__openmp.c:84-117,
 * read_checkpoint_read_checkpoint__openmp.c:64-160).  The reference's file is
 * commondata_struct, then per grid params_struct followed by this payload:
 * int count; int indices[count] (strictly increasing); double data[count * NUM_EVOL_GFS],
 * point-major (data[e * NUM_EVOL_GFS + gf] = y_n[gf * Ntot + indices[e]]).  The structs are
 * project-generated, so the caller writes them; these calls produce and consume the
 * payload.  Pack lists every point; unpack validates like the reference reader, zeroes
 * all of y_n (device) and scatters the listed points into it
 * (read_checkpoint_read_checkpoint__openmp.c:163-171).  Both synchronise the patch stream.
 * After unpacking, the caller sets clock.t_0 = clock.time and clock.nn_0 = clock.nn, as the
 * reference does (read_checkpoint__openmp.c:183-185). */
int64_t bhah_hip_checkpoint_payload_bytes(const bhah_patch *patch);
int bhah_hip_checkpoint_pack(bhah_patch *patch, const double *y_n, void *out);
int bhah_hip_checkpoint_unpack(bhah_patch *patch, const void *payload, int64_t bytes, double *y_n);

/* Brill-Lindquist initial data on the device at every point, ghost zones included
 * (InitialData_Cartesian BrillLindquist with the ADM-to-BSSN conversion,
 * ADM_Initial_Data_Reader__BSSN_Converter.py:268-580): the data are conformally flat, so
 * gammabar = gammahat, h_ij = a_ij = lambda^i = 0 and K = beta = B = 0; cf = alpha = psi^-2 with
 * psi = 1 + m1 / (2 |x - x1|) + m2 / (2 |x - x2|).  bh = {m1, x1, y1, z1, m2, x2, y2, z2}. */
int bhah_hip_initial_data_brill_lindquist(bhah_patch *patch, const double bh[8], double *y_n);

/* General initial data: initial_data_reader__convert_ADM_Cartesian_to_BSSN
 * (nrpy/infrastructures/BHaH/general_relativity/ADM_Initial_Data_Reader__BSSN_Converter.py:268-580,
 * 1019-1129) followed by the outer extrapolation initial_data() applies (initial_data.py:252).
 * adm_cartesian (device, 19 * Ntot doubles) holds ADM data in the Cartesian basis at every
 * point, ghost zones included: gammaDD00, 01, 02, 11, 12, 22, KDD00, ..., KDD22, alpha,
 * betaU0-2, BU0-2.  Per point: BSSN in the Cartesian basis (ADM_to_BSSN.py:80-160), the basis
 * change to the patch's coordinates with the Jacobian of its map, rescaling; lambda^i = 0.
 * Then inner boundary conditions, lambda^i = Lambdabar^i / ReU^i from finite differences of
 * hDD at interior points, and apply_bcs_outerextrap_and_inner on y_n. */
int bhah_hip_initial_data_convert_ADM_Cartesian_to_BSSN(bhah_patch *patch, const double *adm_cartesian, double *y_n);

/* Reverse mode of MoL_step_forward_in_time (one RK4 step; no reference counterpart --
 * the reference has no adjoint).  Given y_n, the state at the start of the step, and
 * lam, the adjoint (gradient of a scalar loss) with respect to the state after the
 * step, overwrite lam with the adjoint with respect to y_n.  The step's stages are
 * recomputed from y_n into a patch-owned workspace (allocated on the first call).
 * Builds with adjoint kernels: Spherical, symmetry_axes 4, no KO, fd_order 4, radiation
 * outer BCs.  Derivatives are those of the discrete step; the upwind stencil choice is
 * piecewise constant in beta^i and contributes none. */
int bhah_hip_MoL_step_adjoint(bhah_patch *patch, const double *y_n, double dt, double *lam);

/* The same reverse step split around a tape (no reference counterpart either), for callers
 * that keep the forward run's stage data instead of recomputing it: HBM holds many steps'
 * tapes.  A step's tape is bhah_hip_MoL_step_tape_doubles(patch) doubles of device memory:
 * the state y_n, the stage inputs Y1..Y3, the right-hand sides K1..K4 after the boundary
 * conditions and the RbarDD of every stage.
 * bhah_hip_MoL_step_forward_taped advances y_n by one RK4 step in place, with the kernels and
 * order of bhah_hip_MoL_step_forward_in_time (the same values bit for bit), and fills the
 * tape; bhah_hip_MoL_step_adjoint_taped then does what bhah_hip_MoL_step_adjoint does for
 * that step, reading the tape instead of recomputing the stages.  Both are asynchronous on
 * the patch's stream and use the patch's adjoint workspace. */
int64_t bhah_hip_MoL_step_tape_doubles(const bhah_patch *patch);
int bhah_hip_MoL_step_forward_taped(bhah_patch *patch, double *y_n, double dt, double *tape);
int bhah_hip_MoL_step_adjoint_taped(bhah_patch *patch, const double *tape, double dt, double *lam);

/* Batched ensembles: many independent BSSN patches of one kernel build, with uneven
 * grids, advanced together.  No single reference counterpart: it batches
 * MoL_step_forward_in_time (MoLtimestepping/MoL_step_forward_in_time.py:39-507) over
 * independent evolutions.  Each stage of a substep is one persistent launch over the
 * tile list of all members; bhah_hip_ensemble_step_n captures a whole step into a
 * hipGraph.  Members must share coordinate system, FD order, symmetry axes and KO
 * (spherical-like coordinates, radiation outer BCs). */
typedef struct bhah_ensemble bhah_ensemble;
int bhah_hip_ensemble_create(const bhah_patch_desc *descs, int n, bhah_ensemble **out);
int bhah_hip_ensemble_destroy(bhah_ensemble *ens);
int bhah_hip_ensemble_size(const bhah_ensemble *ens);
/* The member's patch (grid, boundary structs, dt); owned by the ensemble. */
int bhah_hip_ensemble_member(bhah_ensemble *ens, int m, bhah_patch **patch);
/* Bind each member's five MoL arrays (n entries, all allocated, auxevol included) and
 * timestep (dt == NULL: each member's CFL timestep). */
int bhah_hip_ensemble_bind(bhah_ensemble *ens, const bhah_MoL_gridfunctions *gfs, const double *dt);
int bhah_hip_ensemble_set_stream(bhah_ensemble *ens, void *hip_stream);
int bhah_hip_ensemble_select_schedule(bhah_ensemble *ens, int kind, int index);
/* Interior tile (tile_x * tile_y = 256) and the persistent grid size
 * (max_workgroups; 0 = one workgroup per tile). */
int bhah_hip_ensemble_set_launch(bhah_ensemble *ens, int tile_x, int tile_y, int max_workgroups);
/* nsteps RK4 steps of every member, each with its own dt. */
int bhah_hip_ensemble_step_n(bhah_ensemble *ens, int nsteps, int use_graph);
int bhah_hip_ensemble_synchronize(bhah_ensemble *ens);

#ifdef __cplusplus
}
#endif
#endif /* BHAH_HIP_H */
