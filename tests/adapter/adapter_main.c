/* TEST INFRASTRUCTURE ONLY: a BHaH project main() (the shape of BHaH/tests/main_c__openmp.c:
 * set-up, initial data, MoL_malloc_intermediate_stage_gfs, then
 * while (time < t_final) { diagnostics(); MoL_step_forward_in_time(); }) running on
 * libbhah_hip through nrpy_amd/adapter/bhah_amd_adapter.c.
 *
 *   adapter_main Nxx0 Nxx1 Nxx2 t_final payload_out
 *
 * Spherical, RMAX 7.5, fd4 (NGHOSTS 3), symmetry axis phi, head-on Brill-Lindquist data
 * (m = 0.5 at z = +-0.5), eta 2, CFL 0.5, diagnostics every 0.25.  Writes the final y_n as a
 * checkpoint payload and prints "time nn dt" on stdout. */
#include "BHaH_defines.h"
#include "BHaH_function_prototypes.h"
#include "bhah_hip.h"

bhah_patch *bhah_amd_patch(int grid);

static void check(int rc) {
  if (rc != BHAH_OK) {
    fprintf(stderr, "libbhah_hip: %s\n", bhah_hip_last_error());
    exit(1);
  }
}

int main(int argc, char **argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: %s Nxx0 Nxx1 Nxx2 t_final payload_out\n", argv[0]);
    return 2;
  }
  commondata_struct cd = {0};
  cd.CFL_FACTOR = 0.5;
  cd.KreissOliger_strength_gauge = 0.3;
  cd.KreissOliger_strength_nongauge = 0.3;
  cd.NUMGRIDS = 1;
  cd.convergence_factor = 1.0;
  cd.diagnostics_output_every = 0.25;
  cd.eta = 2.0;
  cd.t_final = atof(argv[4]);
  griddata_struct gd[1] = {0};
  params_struct *p = &gd[0].params;
  p->CoordSystem_hash = SPHERICAL;
  p->Nxx0 = atoi(argv[1]);
  p->Nxx1 = atoi(argv[2]);
  p->Nxx2 = atoi(argv[3]);
  p->Nxx_plus_2NGHOSTS0 = p->Nxx0 + 2 * NGHOSTS;
  p->Nxx_plus_2NGHOSTS1 = p->Nxx1 + 2 * NGHOSTS;
  p->Nxx_plus_2NGHOSTS2 = p->Nxx2 + 2 * NGHOSTS;
  p->SINHW = 0.2;
  p->grid_idx = 0;
  p->grid_physical_size = 7.5;
  /* grid, boundary structs and CFL timestep (numerical_grids_and_timestep + bcstruct_set_up) */
  bhah_amd_attach(&cd, p, 0);
  bhah_patch_info info;
  check(bhah_hip_patch_get_info(bhah_amd_patch(0), &info));
  cd.dt = info.dt;
  /* storage and initial data, then the post-ID boundary conditions */
  MoL_malloc_intermediate_stage_gfs(&cd, p, &gd[0].gridfuncs);
  const double bh[8] = {0.5, 0.0, 0.0, 0.5, 0.5, 0.0, 0.0, -0.5};
  check(bhah_hip_initial_data_brill_lindquist(bhah_amd_patch(0), bh, gd[0].gridfuncs.y_n_gfs));
  apply_bcs_outerextrap_and_inner(&cd, p, &gd[0].bcstruct, gd[0].gridfuncs.y_n_gfs);
  /* main loop (main_c__openmp.c:76-89) */
  while (cd.time < cd.t_final) {
    diagnostics(&cd, gd);
    MoL_step_forward_in_time(&cd, gd);
  }
  const int64_t n = bhah_hip_checkpoint_payload_bytes(bhah_amd_patch(0));
  if (n < 0) check((int)n);
  char *buf = (char *)malloc((size_t)n);
  check(bhah_hip_checkpoint_pack(bhah_amd_patch(0), gd[0].gridfuncs.y_n_gfs, buf));
  FILE *f = fopen(argv[5], "wb");
  if (!f || fwrite(buf, 1, (size_t)n, f) != (size_t)n) {
    fprintf(stderr, "cannot write %s\n", argv[5]);
    return 1;
  }
  fclose(f);
  free(buf);
  printf("%.17g %d %.17g\n", cd.time, cd.nn, cd.dt);
  bhah_amd_detach(0);
  return 0;
}
