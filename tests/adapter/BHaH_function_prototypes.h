/* TEST INFRASTRUCTURE ONLY: the generated prototypes the adapter replaces, with the
 * reference emitters' signatures (see nrpy_amd/adapter/bhah_amd_adapter.c). */
#include "BHaH_defines.h"
void rhs_eval(const commondata_struct *restrict commondata, const params_struct *restrict params,
              const rfm_struct *restrict rfmstruct, const REAL *restrict auxevol_gfs, const REAL *restrict in_gfs,
              REAL *restrict rhs_gfs);
void Ricci_eval(const params_struct *restrict params, const rfm_struct *restrict rfmstruct,
                const REAL *restrict in_gfs, REAL *restrict auxevol_gfs);
void enforce_detgammabar_equals_detgammahat(const params_struct *restrict params,
                                            const rfm_struct *restrict rfmstruct, REAL *restrict in_gfs,
                                            const REAL *restrict auxevol_gfs);
void apply_bcs_outerradiation_and_inner(const commondata_struct *restrict commondata,
                                        const params_struct *restrict params, const bc_struct *restrict bcstruct,
                                        REAL *restrict xx[3], const REAL custom_wavespeed[NUM_EVOL_GFS],
                                        const REAL custom_f_infinity[NUM_EVOL_GFS], REAL *restrict gfs,
                                        REAL *restrict rhs_gfs);
void apply_bcs_outerextrap_and_inner(const commondata_struct *restrict commondata,
                                     const params_struct *restrict params, const bc_struct *restrict bcstruct,
                                     REAL *restrict gfs);
void apply_bcs_inner_only(const commondata_struct *restrict commondata, const params_struct *restrict params,
                          const bc_struct *restrict bcstruct, REAL *restrict gfs);
void MoL_malloc_intermediate_stage_gfs(const commondata_struct *restrict commondata,
                                       const params_struct *restrict params,
                                       MoL_gridfunctions_struct *restrict gridfuncs);
void MoL_step_forward_in_time(commondata_struct *restrict commondata, griddata_struct *restrict griddata);
void diagnostics(commondata_struct *restrict commondata, griddata_struct *restrict griddata);
/* adapter set-up */
void bhah_amd_attach(const commondata_struct *cd, const params_struct *p, int grid);
void bhah_amd_detach(int grid);
