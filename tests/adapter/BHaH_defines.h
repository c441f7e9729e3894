/* TEST INFRASTRUCTURE ONLY: a stand-in for the BHaH_defines.h that NRPy generates for a
 * BSSN project, with the reference's type and member names (structs and macros as emitted
 * by BHaH/BHaH_defines_h.py, MoLtimestepping/BHaH_defines.py:32-40,
 * CurviBoundaryConditions/BHaH_defines.py:158-190, rfm_wrapper_functions.py:130; frozen
 * example BHaH/tests/BHaH_defines_h_BHaH_defines.c:121-218).  Member sets are cut down to
 * what the adapter and the test program use; a generated header has more. */
#ifndef __BHAH_DEFINES_H__
#define __BHAH_DEFINES_H__
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#define REAL double
#define NGHOSTS 3
#define NUM_EVOL_GFS 24
#define NUM_AUXEVOL_GFS 6
#define MAXNUMGRIDS 15
#define SINHSPHERICAL 579707936
#define SPHERICAL 1307621703

typedef struct __commondata_struct__ {
  REAL CFL_FACTOR;
  REAL KreissOliger_strength_gauge;
  REAL KreissOliger_strength_nongauge;
  int NUMGRIDS;
  REAL convergence_factor;
  REAL diagnostics_output_every;
  REAL dt;
  REAL eta;
  int nn;
  int nn_0;
  REAL t_0;
  REAL t_final;
  REAL time;
} commondata_struct;

typedef struct __params_struct__ {
  int CoordSystem_hash;
  int Nxx0, Nxx1, Nxx2;
  int Nxx_plus_2NGHOSTS0, Nxx_plus_2NGHOSTS1, Nxx_plus_2NGHOSTS2;
  REAL SINHW;
  int grid_idx;
  REAL grid_physical_size;
} params_struct;

typedef struct __MoL_gridfunctions_struct__ {
  REAL *y_n_gfs;
  REAL *y_nplus1_running_total_gfs;
  REAL *k_odd_gfs;
  REAL *k_even_gfs;
  REAL *auxevol_gfs;
} MoL_gridfunctions_struct;

typedef struct __rfmstruct__ {
  REAL *f0_of_xx0;
} rfm_struct;

typedef struct __innerpt_bc_struct__ {
  int dstpt, srcpt;
  int8_t parity[10];
} innerpt_bc_struct;
typedef struct __outerpt_bc_struct__ {
  short i0, i1, i2;
  int8_t FACEX0, FACEX1, FACEX2;
} outerpt_bc_struct;
typedef struct __bc_info_struct__ {
  int num_inner_boundary_points;
  int num_pure_outer_boundary_points[NGHOSTS][3];
  int bc_loop_bounds[NGHOSTS][6][6];
} bc_info_struct;
typedef struct __bc_struct__ {
  innerpt_bc_struct *inner_bc_array;
  outerpt_bc_struct *pure_outer_bc_array[NGHOSTS * 3];
  bc_info_struct bc_info;
} bc_struct;

typedef struct __griddata__ {
  REAL *xx[3];
  MoL_gridfunctions_struct gridfuncs;
  params_struct params;
  rfm_struct *rfmstruct;
  bc_struct bcstruct;
} griddata_struct;
#endif
