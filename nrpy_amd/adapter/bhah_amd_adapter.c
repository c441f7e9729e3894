/* bhah_amd_adapter.c -- binds the C functions of a generated BHaH project to libbhah_hip.
 *
 * Compile it inside the project, against the project's own BHaH_defines.h (struct layouts
 * are generated per project, BHaH/BHaH_defines_h.py:270-275), and link libbhah_hip.so:
 *
 *   gcc -O2 -c bhah_amd_adapter.c -I<project> -I<repo>/include
 *   hipcc -o project ... bhah_amd_adapter.o -L<repo>/nrpy_amd/lib -lbhah_hip
 *
 * It replaces the project's definitions of the functions below; their signatures are those
 * of the reference emitters: rhs_eval.py:97-120, Ricci_eval.py:81-93,
 * enforce_detgammabar_equals_detgammahat.py:63-83,
 * apply_bcs_outerradiation_and_inner__rfm__openmp__HoleySinhSpherical.c:211-215,
 * apply_bcs_outerextrap_and_inner.py:252, apply_bcs_inner_only.py:52-111,
 * MoL_step_forward_in_time.py:39-507, diagnostics/diagnostics.py:46-270.  All REAL arrays
 * must be device memory, as in the reference's CUDA builds.
 *
 * Members read: commondata->{time, t_0, dt, nn, nn_0, NUMGRIDS, CFL_FACTOR, eta,
 * KreissOliger_strength_gauge, KreissOliger_strength_nongauge, diagnostics_output_every,
 * convergence_factor}; params->{CoordSystem_hash, Nxx0, Nxx1, Nxx2, grid_physical_size,
 * SINHW, grid_idx}; griddata[g].gridfuncs.{y_n_gfs, y_nplus1_running_total_gfs, k_odd_gfs,
 * k_even_gfs, auxevol_gfs}.  The generated project defines SPHERICAL and SINHSPHERICAL as
 * the coordinate-system hashes (rfm_wrapper_functions.py:17-33, 130). */
#include "BHaH_defines.h"
#include "BHaH_function_prototypes.h"
#include "bhah_hip.h"

#include <stdio.h>
#include <stdlib.h>

#ifndef BHAH_AMD_SYMMETRY_AXES
#define BHAH_AMD_SYMMETRY_AXES 0 /* 4 when the project was generated with symmetry_axes="2" */
#endif
#ifndef BHAH_AMD_ENABLE_KO
#define BHAH_AMD_ENABLE_KO 1 /* as the project's rhs_eval was generated */
#endif

static bhah_patch *patch_of_grid[MAXNUMGRIDS];
static REAL *diag_of_grid[MAXNUMGRIDS];

static void check(int rc) {
  if (rc != BHAH_OK) {
    fprintf(stderr, "libbhah_hip: %s\n", bhah_hip_last_error());
    exit(1);
  }
}

static bhah_clock clock_of(const commondata_struct *cd) {
  bhah_clock c;
  c.time = cd->time;
  c.t_0 = cd->t_0;
  c.dt = cd->dt;
  c.nn = cd->nn;
  c.nn_0 = cd->nn_0;
  return c;
}

/* Call once per grid after numerical_grids_and_timestep(). */
void bhah_amd_attach(const commondata_struct *cd, const params_struct *p, int grid) {
  bhah_patch_desc d = {0};
  d.system = BHAH_SYSTEM_BSSN;
  d.coord = (p->CoordSystem_hash == SINHSPHERICAL) ? BHAH_SINHSPHERICAL : BHAH_SPHERICAL;
  d.Nxx[0] = p->Nxx0;
  d.Nxx[1] = p->Nxx1;
  d.Nxx[2] = p->Nxx2;
  d.fd_order = 2 * (NGHOSTS - 1);
  d.symmetry_axes = BHAH_AMD_SYMMETRY_AXES;
  d.enable_KO = BHAH_AMD_ENABLE_KO;
  d.outer_bc = BHAH_OUTER_BC_RADIATION;
  d.radiation_bc_fd_order = 4;
  d.grid_physical_size = p->grid_physical_size;
  d.SINHW = p->SINHW;
  d.CFL_FACTOR = cd->CFL_FACTOR;
  d.eta = cd->eta;
  d.KO_strength_gauge = cd->KreissOliger_strength_gauge;
  d.KO_strength_nongauge = cd->KreissOliger_strength_nongauge;
  check(bhah_hip_patch_create(&d, &patch_of_grid[grid]));
}

void bhah_amd_detach(int grid) {
  if (diag_of_grid[grid]) check(bhah_hip_diagnostic_gfs_free(diag_of_grid[grid]));
  diag_of_grid[grid] = NULL;
  check(bhah_hip_patch_destroy(patch_of_grid[grid]));
  patch_of_grid[grid] = NULL;
}

static bhah_patch *P(const params_struct *p) { return patch_of_grid[p->grid_idx]; }

/* The library patch of a grid (for initial data, checkpoints and the CFL timestep). */
bhah_patch *bhah_amd_patch(int grid) { return patch_of_grid[grid]; }

/* MoL_malloc_intermediate_stage_gfs (MoLtimestepping/MoL_malloc_intermediate_stage_gfs.py:80-88):
 * y_nplus1_running_total, k_odd and k_even on the device; y_n and auxevol as well when the
 * project has not allocated them. */
void MoL_malloc_intermediate_stage_gfs(const commondata_struct *cd, const params_struct *p,
                                       MoL_gridfunctions_struct *gf) {
  (void)cd;
  bhah_MoL_gridfunctions h = {gf->y_n_gfs, NULL, NULL, NULL, gf->auxevol_gfs};
  check(bhah_hip_MoL_malloc_intermediate_stage_gfs(P(p), &h));
  gf->y_n_gfs = h.y_n_gfs;
  gf->y_nplus1_running_total_gfs = h.y_nplus1_running_total_gfs;
  gf->k_odd_gfs = h.k_odd_gfs;
  gf->k_even_gfs = h.k_even_gfs;
  gf->auxevol_gfs = h.auxevol_gfs;
}

void rhs_eval(const commondata_struct *cd, const params_struct *p, const rfm_struct *rfm, const REAL *auxevol_gfs,
              const REAL *in_gfs, REAL *rhs_gfs) {
  (void)cd;
  (void)rfm;
  check(bhah_hip_rhs_eval(P(p), auxevol_gfs, in_gfs, rhs_gfs));
}

void Ricci_eval(const params_struct *p, const rfm_struct *rfm, const REAL *in_gfs, REAL *auxevol_gfs) {
  (void)rfm;
  check(bhah_hip_Ricci_eval(P(p), in_gfs, auxevol_gfs));
}

void enforce_detgammabar_equals_detgammahat(const params_struct *p, const rfm_struct *rfm, REAL *in_gfs,
                                            const REAL *auxevol_gfs) {
  (void)rfm;
  (void)auxevol_gfs;
  check(bhah_hip_enforce_detgammabar_equals_detgammahat(P(p), in_gfs));
}

/* The patch holds the boundary structs, wavespeeds and f_infinity; the project's own
 * arguments for them are not read. */
void apply_bcs_outerradiation_and_inner(const commondata_struct *cd, const params_struct *p, const bc_struct *bc,
                                        REAL *restrict xx[3], const REAL wavespeed[NUM_EVOL_GFS],
                                        const REAL f_infinity[NUM_EVOL_GFS], REAL *gfs, REAL *rhs_gfs) {
  (void)cd;
  (void)bc;
  (void)xx;
  (void)wavespeed;
  (void)f_infinity;
  check(bhah_hip_apply_bcs_outerradiation_and_inner(P(p), gfs, rhs_gfs));
}

void apply_bcs_outerextrap_and_inner(const commondata_struct *cd, const params_struct *p, const bc_struct *bc,
                                     REAL *gfs) {
  (void)cd;
  (void)bc;
  check(bhah_hip_apply_bcs_outerextrap_and_inner(P(p), gfs));
}

void apply_bcs_inner_only(const commondata_struct *cd, const params_struct *p, const bc_struct *bc, REAL *gfs) {
  (void)cd;
  (void)bc;
  check(bhah_hip_apply_bcs_inner_only(P(p), gfs));
}

/* One RK4 step of every grid, then the reference's bookkeeping
 * time = t_0 + (nn - nn_0 + 1) dt, nn++ (MoL_step_forward_in_time.py:473-481). */
void MoL_step_forward_in_time(commondata_struct *cd, griddata_struct *gd) {
  bhah_clock c = clock_of(cd);
  for (int g = 0; g < cd->NUMGRIDS; g++) {
    MoL_gridfunctions_struct *m = &gd[g].gridfuncs;
    bhah_MoL_gridfunctions h = {m->y_n_gfs, m->y_nplus1_running_total_gfs, m->k_odd_gfs, m->k_even_gfs,
                                m->auxevol_gfs};
    c = clock_of(cd);
    check(bhah_hip_MoL_step_forward_in_time(patch_of_grid[g], &h, &c));
  }
  cd->time = c.time;
  cd->nn = c.nn;
}

/* diagnostics(): on output steps, the constraint volume integrals of every grid, written as
 * the reference's rows (time, volume, then integral, L2 and RMS per integrand) to
 * out3d-integrals-conv_factor<cf>-<recipe>.txt (diagnostics_volume_integration_helpers.h:
 * 332-345, 739-741).  The diagnostic gridfunctions stay on the device. */
void diagnostics(commondata_struct *cd, griddata_struct *gd) {
  static const char *recipe[BHAH_DIAG_NUM_RECIPES] = {"whole_domain", "sphere_R_gt_8"};
  static const char *integrand[BHAH_DIAG_NUM_INTEGRANDS] = {"DIAG_HAMILTONIAN", "DIAG_MSQUARED",
                                                            "DIAG_LAMBDA_CONSTRAINT"};
  const bhah_clock c = clock_of(cd);
  for (int g = 0; g < cd->NUMGRIDS; g++) {
    /* the reference allocates diagnostic_gfs on each output step (diagnostics.py:204-209, 249);
     * here one device buffer per grid is kept for the run */
    if (!diag_of_grid[g]) check(bhah_hip_diagnostic_gfs_malloc(patch_of_grid[g], &diag_of_grid[g]));
    bhah_volume_integrals vi;
    int output_step = 0;
    check(bhah_hip_diagnostics(patch_of_grid[g], &c, cd->diagnostics_output_every, gd[g].gridfuncs.y_n_gfs,
                               gd[g].gridfuncs.auxevol_gfs, diag_of_grid[g], &vi, &output_step));
    if (!output_step) continue;
    for (int q = 0; q < BHAH_DIAG_NUM_RECIPES; q++) {
      char fname[256];
      snprintf(fname, sizeof(fname), "out3d-integrals-conv_factor%.2f-%s.txt", cd->convergence_factor, recipe[q]);
      FILE *f = fopen(fname, cd->nn == 0 ? "w" : "a");
      if (!f) {
        fprintf(stderr, "diagnostics: could not open %s\n", fname);
        exit(1);
      }
      if (cd->nn == 0) {
        fprintf(f, "# time remaining_volume");
        for (int k = 0; k < BHAH_DIAG_NUM_INTEGRANDS; k++)
          fprintf(f, " INT[%s^2] L2[%s] RMS[%s]", integrand[k], integrand[k], integrand[k]);
        fprintf(f, "\n");
      }
      fprintf(f, "%.15e %.15e", c.time, vi.volume[q]);
      for (int k = 0; k < BHAH_DIAG_NUM_INTEGRANDS; k++)
        fprintf(f, " %.15e %.15e %.15e", vi.integral[q][k], vi.L2[q][k], vi.RMS[q][k]);
      fprintf(f, "\n");
      fclose(f);
    }
  }
}
