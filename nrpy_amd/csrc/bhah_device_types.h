// Parameter blocks shared by the generated kernels (gen/*.hip) and the host
// runtime (bhah_hip.hip).
#pragma once

// Grid and physics parameters of one patch, passed by value to every kernel.
struct bhah_kparams {
  int Nx0, Nx1, Nx2, NG;
  long Ntot, s1, s2;
  double invdxx0, invdxx1, invdxx2;
  double eta, KO_gauge, KO_nongauge, wavespeed;
  double dt;
  int substep;
};

// One member of a batched ensemble as the batched kernels see it: its
// parameters, reference-metric arrays and its five method-of-lines arrays
// (arr[0] y_n, arr[1] y_nplus1_running_total, arr[2] k_odd, arr[3] k_even,
// arr[4] auxevol).  Kernels pick arrays by index so one table serves every
// substep.
struct bhah_member {
  bhah_kparams kp;
  const double *rfm0, *rfm1, *xx0;
  double *arr[5];
  int coord;
  double AMPL, SINHW;
};

// One unit of batched work: member index and tile coordinates (tile along
// xx0, tile along xx1, xx2 plane) of the member's launch grid.
struct bhah_item {
  int m, bx, by, bz;
};
