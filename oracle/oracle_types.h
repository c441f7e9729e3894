/* ORACLE / TEST INFRASTRUCTURE ONLY -- never linked into the product library.
 * Shared types of the CPU restatement in bhah_oracle.c. */
#pragma once

enum { OG_CENTRE = 0, OG_AUX = 1, OG_DERIV = 2, OG_RFM = 3, OG_PARAM = 4 };
enum { OG_dD = 0, OG_dDD = 1, OG_dupD = 2, OG_dKOD = 3, OG_ddnD = 4 };

typedef struct {
  int kind; /* OG_CENTRE ... */
  int a;    /* gf index / rfm slot / param slot */
  int b;    /* derivative op */
  int c, d; /* derivative directions */
} og_slot;

/* One grid: sizes, ghost width, coordinate bounds and spacings, coordinate parameters. */
typedef struct {
  int coord;
  int Nxx[3], NG, Nx[3];
  double xxmin[3], xxmax[3], dxx[3], invdxx[3];
  double p0, p1; /* RMAX | AMPL,SINHW | grid_physical_size */
  int fd_order;
} og_grid;

/* Unrolled-stencil form of a pointwise kernel over the interior (native-flags build only):
 * (grid, params, rfm0, rfm1, aux, in, out). */
typedef void (*og_unrolled_fn)(const og_grid *restrict, const double *restrict, const double *restrict,
                               const double *restrict, const double *restrict, const double *restrict,
                               double *restrict);

/* 1D stencil as exact rationals num/den at integer offsets. */
typedef struct {
  int n;
  int off[12];
  long long num[12];
  long long den[12];
} og_stencil;
