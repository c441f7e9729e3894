/* ORACLE / TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C CPU restatement of the reference's BHaH single-patch evolution
 * path, used by the tests and by bench.py's CPU-baseline leg as the checker.
 * It is never linked into, loaded by, or called from the product library.
 *
 * Restated from the reference (nrpy/nrpy), file:line cited per function.
 * Layout: gf-major SoA doubles, IDX4(gf,i0,i1,i2) = i0 + Nx0*(i1 + Nx1*(i2 + Nx2*gf))
 * (nrpy/infrastructures/BHaH/tests/BHaH_defines_h_BHaH_defines.c:149-162).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gen/oracle_pointwise.h"
#include "oracle_types.h"

#define NRPYMIN(a, b) ((a) < (b) ? (a) : (b))

enum { OG_CARTESIAN = 0, OG_SPHERICAL = 1, OG_SINHSPHERICAL = 2 };

/* numerical_grid_params_Nxx_dxx_xx: nrpy/infrastructures/BHaH/tests/
 * numerical_grids_and_timestep_numerical_grid_params_Nxx_dxx_xx__openmp__HoleySinhSpherical.c:60-104
 * and nrpy/reference_metric.py:1201-1274 (xxmin/xxmax of Spherical, SinhSpherical);
 * Cartesian as nrpy/examples/wave_equation_cartesian.py:93-122. */
void og_grid_init(og_grid *g, int coord, int N0, int N1, int N2, int NG, int fd_order, double p0, double p1) {
  memset(g, 0, sizeof(*g));
  g->coord = coord;
  g->Nxx[0] = N0;
  g->Nxx[1] = N1;
  g->Nxx[2] = N2;
  g->NG = NG;
  g->fd_order = fd_order;
  g->p0 = p0;
  g->p1 = p1;
  if (coord == OG_CARTESIAN) {
    for (int d = 0; d < 3; d++) {
      g->xxmin[d] = -p0;
      g->xxmax[d] = p0;
    }
  } else {
    g->xxmin[0] = 0.0;
    g->xxmax[0] = (coord == OG_SPHERICAL) ? p0 : 1.0;
    g->xxmin[1] = 0.0;
    g->xxmax[1] = M_PI;
    g->xxmin[2] = -M_PI;
    g->xxmax[2] = M_PI;
  }
  for (int d = 0; d < 3; d++) {
    g->Nx[d] = g->Nxx[d] + 2 * NG;
    g->dxx[d] = (g->xxmax[d] - g->xxmin[d]) / ((double)g->Nxx[d]);
    g->invdxx[d] = ((double)g->Nxx[d]) / (g->xxmax[d] - g->xxmin[d]);
  }
}

void og_grid_get(const og_grid *g, int *ints /*10*/, double *dbls /*14*/) {
  ints[0] = g->coord;
  for (int d = 0; d < 3; d++) {
    ints[1 + d] = g->Nxx[d];
    ints[5 + d] = g->Nx[d];
    dbls[d] = g->xxmin[d];
    dbls[3 + d] = g->xxmax[d];
    dbls[6 + d] = g->dxx[d];
    dbls[9 + d] = g->invdxx[d];
  }
  ints[4] = g->NG;
  ints[8] = g->fd_order;
  dbls[12] = g->p0;
  dbls[13] = g->p1;
}

/* Cell-centred coordinates: xx[j] = xxmin + ((j - NGHOSTS) + 1/2) dxx
 * (numerical_grid_params_Nxx_dxx_xx__openmp__HoleySinhSpherical.c:4-8). */
void og_xx(const og_grid *g, double *x0, double *x1, double *x2) {
  double *xx[3] = {x0, x1, x2};
  for (int d = 0; d < 3; d++)
    for (int j = 0; j < g->Nx[d]; j++)
      xx[d][j] = g->xxmin[d] + ((double)(j - g->NG) + (1.0 / 2.0)) * g->dxx[d];
}

/* rfm_precompute 1D arrays (nrpy/infrastructures/BHaH/rfm_precompute.py:196-340) for the
 * spherical-like closed forms f0 = r(xx0), f1 = sin(xx1) (nrpy/reference_metric.py:1338-1357).
 * rfm0[k*Nx0 + i0]: f0, f0', f0'', f0'''; rfm1[k*Nx1 + i1]: f1, f1', f1''. */
void og_rfm(const og_grid *g, const double *x0, const double *x1, double *rfm0, double *rfm1) {
  const int N0 = g->Nx[0], N1 = g->Nx[1];
  for (int i = 0; i < N0; i++) {
    const double x = x0[i];
    double f = 1, d1 = 0, d2 = 0, d3 = 0;
    if (g->coord == OG_SPHERICAL) {
      f = x;
      d1 = 1.0;
    } else if (g->coord == OG_SINHSPHERICAL) {
      const double A = g->p0, w = g->p1;
      const double den = exp(1.0 / w) - exp(-1.0 / w);
      const double ep = exp(x / w), em = exp(-x / w);
      f = A * (ep - em) / den;
      d1 = A * (ep + em) / (w * den);
      d2 = A * (ep - em) / (w * w * den);
      d3 = A * (ep + em) / (w * w * w * den);
    }
    rfm0[0 * N0 + i] = f;
    rfm0[1 * N0 + i] = d1;
    rfm0[2 * N0 + i] = d2;
    rfm0[3 * N0 + i] = d3;
  }
  for (int j = 0; j < N1; j++) {
    const double y = x1[j];
    const int cart = (g->coord == OG_CARTESIAN);
    rfm1[0 * N1 + j] = cart ? 1.0 : sin(y);
    rfm1[1 * N1 + j] = cart ? 0.0 : cos(y);
    rfm1[2 * N1 + j] = cart ? 0.0 : -sin(y);
  }
}

/* ------------------------------------------------------------------------ */
/* Finite differences (nrpy/finite_difference.py:202-330, 1043-1068).        */
static double st_w(const og_stencil *s, int k) { return (double)s->num[k] / (double)s->den[k]; }

static double fd1(const og_grid *g, const double *gf, long idx, int op, int dir) {
  const og_stencil *s = &OG_STENCIL[g->fd_order / 2 - 1][op];
  const long stride = dir == 0 ? 1 : (dir == 1 ? g->Nx[0] : (long)g->Nx[0] * g->Nx[1]);
  double acc = 0.0;
  for (int k = 0; k < s->n; k++)
    acc += st_w(s, k) * gf[idx + s->off[k] * stride];
  return acc * g->invdxx[dir];
}

static double fd_dDD(const og_grid *g, const double *gf, long idx, int d0, int d1) {
  const long str[3] = {1, g->Nx[0], (long)g->Nx[0] * g->Nx[1]};
  if (d0 == d1) {
    const og_stencil *s = &OG_STENCIL[g->fd_order / 2 - 1][OG_dDD];
    double acc = 0.0;
    for (int k = 0; k < s->n; k++)
      acc += st_w(s, k) * gf[idx + s->off[k] * str[d0]];
    return acc * g->invdxx[d0] * g->invdxx[d0];
  }
  /* mixed: products of centred first-derivative weights, exact rationals */
  const og_stencil *s = &OG_STENCIL[g->fd_order / 2 - 1][OG_dD];
  double acc = 0.0;
  for (int a = 0; a < s->n; a++)
    for (int b = 0; b < s->n; b++) {
      const double w = (double)(s->num[a] * s->num[b]) / (double)(s->den[a] * s->den[b]);
      acc += w * gf[idx + s->off[a] * str[d0] + s->off[b] * str[d1]];
    }
  return acc * g->invdxx[d0] * g->invdxx[d1];
}

/* Fill the slot array v[] for one point.  Upwinded slots (dupD) follow
 * nrpy/c_codegen.py:1085-1150: U = (betaU[d] > 0), v = U*(dupD - ddnD) + ddnD. */
static void fill_slots(const og_grid *g, const og_slot *slots, int nslots, const double *in, const double *aux,
                       const double *rfm0, const double *rfm1, const double *params, int i0, int i1, int i2, double *v,
                       void (*upwind)(const double *, double *)) {
  const long Ntot = (long)g->Nx[0] * g->Nx[1] * g->Nx[2];
  const long idx = i0 + (long)g->Nx[0] * (i1 + (long)g->Nx[1] * i2);
  for (int k = 0; k < nslots; k++) {
    const og_slot *s = &slots[k];
    switch (s->kind) {
    case OG_CENTRE:
      v[k] = in[s->a * Ntot + idx];
      break;
    case OG_AUX:
      v[k] = aux[s->a * Ntot + idx];
      break;
    case OG_RFM:
      v[k] = (s->a < 4) ? rfm0[s->a * g->Nx[0] + i0] : rfm1[(s->a - 4) * g->Nx[1] + i1];
      break;
    case OG_PARAM:
      v[k] = params[s->a];
      break;
    case OG_DERIV: {
      const double *gf = &in[s->a * Ntot];
      if (s->b == OG_dD || s->b == OG_dKOD)
        v[k] = fd1(g, gf, idx, s->b, s->c);
      else if (s->b == OG_dDD)
        v[k] = fd_dDD(g, gf, idx, s->c, s->d);
      break;
    }
    }
  }
  if (upwind) {
    double betaU[3];
    upwind(v, betaU);
    for (int k = 0; k < nslots; k++) {
      const og_slot *s = &slots[k];
      if (s->kind == OG_DERIV && s->b == OG_dupD) {
        const double *gf = &in[s->a * Ntot];
        const double up = fd1(g, gf, idx, OG_dupD, s->c);
        const double dn = fd1(g, gf, idx, OG_ddnD, s->c);
        const double U = betaU[s->c] > 0.0 ? 1.0 : 0.0;
        v[k] = U * (up - dn) + dn;
      }
    }
  }
}

typedef struct {
  const og_slot *slots;
  int nslots, nout;
  const int *out_gf;
  void (*pointwise)(const double *, double *);
  void (*upwind)(const double *, double *);
} og_kernel;

#define KERNEL(name, up) {name##_slots, sizeof(name##_slots) / sizeof(og_slot), sizeof(name##_out_gf) / sizeof(int), name##_out_gf, name##_pointwise, up}

/* variant 0: Spherical, symmetry axis 2, no KO; variant 1: SinhSpherical, full 3D, KO */
static const og_kernel K_RHS[2] = {KERNEL(bssn_rhs_Spherical_sym2_KO0, bssn_rhs_Spherical_sym2_KO0_upwind),
                                   KERNEL(bssn_rhs_SinhSpherical_symnone_KO1, bssn_rhs_SinhSpherical_symnone_KO1_upwind)};
static const og_kernel K_RICCI[2] = {KERNEL(ricci_Spherical_sym2_KO0, NULL), KERNEL(ricci_SinhSpherical_symnone_KO1, NULL)};
static const og_kernel K_DET[2] = {KERNEL(detgamma_Spherical_sym2_KO0, NULL), KERNEL(detgamma_SinhSpherical_symnone_KO1, NULL)};
static const og_kernel K_CONS[2] = {KERNEL(constraints_Spherical_sym2_KO0, NULL),
                                    KERNEL(constraints_SinhSpherical_symnone_KO1, NULL)};
static const og_kernel K_LAM[2] = {KERNEL(initial_lambda_Spherical_sym2_KO0, NULL),
                                   KERNEL(initial_lambda_SinhSpherical_symnone_KO1, NULL)};
static const og_kernel K_WAVE = KERNEL(wave_rhs, NULL);

/* Interior ("loop_region=interior") or all-points loop applying a pointwise kernel. */
static void run_kernel(const og_kernel *K, const og_grid *g, const double *rfm0, const double *rfm1, const double *params,
                       const double *aux, const double *in, double *out, int all_points) {
  const long Ntot = (long)g->Nx[0] * g->Nx[1] * g->Nx[2];
  const int lo = all_points ? 0 : g->NG;
#pragma omp parallel for schedule(static)
  for (int i2 = lo; i2 < g->Nx[2] - lo; i2++) {
    double v[512], o[32];
    for (int i1 = lo; i1 < g->Nx[1] - lo; i1++)
      for (int i0 = lo; i0 < g->Nx[0] - lo; i0++) {
        fill_slots(g, K->slots, K->nslots, in, aux, rfm0, rfm1, params, i0, i1, i2, v, K->upwind);
        K->pointwise(v, o);
        const long idx = i0 + (long)g->Nx[0] * (i1 + (long)g->Nx[1] * i2);
        for (int k = 0; k < K->nout; k++)
          out[K->out_gf[k] * Ntot + idx] = o[k];
      }
  }
}

/* rhs_eval (nrpy/infrastructures/BHaH/general_relativity/rhs_eval.py:336-379): interior only.
 * params[]: eta, KO_gauge, KO_nongauge, wavespeed. */
#ifdef OG_UNROLLED
/* The native-flags build (liboracle_bhah_native.so) runs the unrolled-stencil forms where one
 * exists for the variant and finite-difference order (OG_UNROLLED_TABLE: kind 0 = rhs_eval,
 * 1 = Ricci_eval); everything else is shared with the checker build. */
#include "gen/oracle_unrolled.h"

static og_unrolled_fn unrolled_find(int kind, int variant, int fd_order) {
  for (size_t k = 0; k < sizeof(OG_UNROLLED_TABLE) / sizeof(OG_UNROLLED_TABLE[0]); k++) {
    const og_unrolled_entry *e = &OG_UNROLLED_TABLE[k];
    if (e->kind == kind && e->variant == variant && e->fd_order == fd_order) return e->fn;
  }
  return NULL;
}
#define unrolled_rhs(v, o) unrolled_find(0, (v), (o))
#define unrolled_ricci(v, o) unrolled_find(1, (v), (o))

int og_unrolled_available(int variant, int fd_order) {
  return unrolled_rhs(variant, fd_order) != NULL && unrolled_ricci(variant, fd_order) != NULL;
}
#else
int og_unrolled_available(int variant, int fd_order) {
  (void)variant;
  (void)fd_order;
  return 0;
}
#endif

void og_bssn_rhs(const og_grid *g, int variant, const double *params, const double *rfm0, const double *rfm1,
                 const double *auxevol, const double *in, double *rhs) {
#ifdef OG_UNROLLED
  const og_unrolled_fn f = unrolled_rhs(variant, g->fd_order);
  if (f) {
    f(g, params, rfm0, rfm1, auxevol, in, rhs);
    return;
  }
#endif
  run_kernel(&K_RHS[variant], g, rfm0, rfm1, params, auxevol, in, rhs, 0);
}

/* Ricci_eval (nrpy/infrastructures/BHaH/general_relativity/Ricci_eval.py:95-142): interior only;
 * writes RbarDD into the 6 auxevol gridfunctions. */
void og_ricci(const og_grid *g, int variant, const double *rfm0, const double *rfm1, const double *in, double *auxevol) {
#ifdef OG_UNROLLED
  const og_unrolled_fn f = unrolled_ricci(variant, g->fd_order);
  if (f) {
    f(g, NULL, rfm0, rfm1, NULL, in, auxevol);
    return;
  }
#endif
  run_kernel(&K_RICCI[variant], g, rfm0, rfm1, NULL, NULL, in, auxevol, 0);
}

/* enforce_detgammabar_equals_detgammahat (…/enforce_detgammabar_equals_detgammahat.py:84-128):
 * all points, in place on hDD. */
void og_enforce_detgamma(const og_grid *g, int variant, const double *rfm0, const double *rfm1, double *gfs) {
  const og_kernel *K = &K_DET[variant];
  /* In place: the kernel reads only its own point's values (no stencil slots) and writes hDD
   * there after every slot of that point has been read.  A stencil slot would make the update
   * order-dependent, so evaluate from a copy in that case. */
  int stencil = 0;
  for (int k = 0; k < K->nslots; k++) stencil |= K->slots[k].kind == OG_DERIV;
  if (!stencil) {
    run_kernel(K, g, rfm0, rfm1, NULL, NULL, gfs, gfs, 1);
    return;
  }
  const long Ntot = (long)g->Nx[0] * g->Nx[1] * g->Nx[2];
  double *copy = malloc(sizeof(double) * 24 * Ntot);
  memcpy(copy, gfs, sizeof(double) * 24 * Ntot);
  run_kernel(K, g, rfm0, rfm1, NULL, NULL, copy, gfs, 1);
  free(copy);
}

/* lambda^i of converted initial data from finite differences of hDD at interior points, written
 * into the lambdaU gridfunctions of gfs in place (initial_data_lambdaU_grid_interior,
 * ADM_Initial_Data_Reader__BSSN_Converter.py:462-580; it reads only hDD and the rfm arrays). */
void og_initial_lambda(const og_grid *g, int variant, const double *rfm0, const double *rfm1, double *gfs) {
  run_kernel(&K_LAM[variant], g, rfm0, rfm1, NULL, NULL, gfs, gfs, 0);
}

/* Constraints H, MU^i, M^2 and |Lambda constraint| (BSSN_constraints.py:82-203) at interior
 * points: out[6*Ntot]. */
void og_constraints(const og_grid *g, int variant, const double *rfm0, const double *rfm1, const double *auxevol,
                    const double *in, double *out) {
  run_kernel(&K_CONS[variant], g, rfm0, rfm1, NULL, auxevol, in, out, 0);
}

/* Scalar-wave RHS (nrpy/examples/wave_equation_cartesian.py:318-357): interior only. */
void og_wave_rhs(const og_grid *g, double wavespeed, const double *in, double *rhs) {
  const double params[4] = {0, 0, 0, wavespeed};
  run_kernel(&K_WAVE, g, NULL, NULL, params, NULL, in, rhs, 0);
}

/* ------------------------------------------------------------------------ */
/* Boundary-condition struct (CurviBoundaryConditions/tests/
 * bcstruct_set_up_bcstruct_set_up__rfm__openmp__HoleySinhSpherical.c:43-520), restated for
 * spherical-like coordinates whose eigen-coordinate is Spherical.                */
typedef struct {
  int num_inner;
  long *inner_dst, *inner_src;
  int8_t *inner_parity; /* [num_inner*10] */
  int num_outer[8][3];
  short *outer[8][3]; /* 6 shorts per point: i0,i1,i2,FACEX0,FACEX1,FACEX2 */
} og_bcstruct;

static void eigen_inbounds(const og_grid *g, const double *xx[3], int i0, int i1, int i2, int out[3], double xin[3]) {
  /* Spherical eigen-coordinate double map xx -> Cart -> xx' */
  const double r = xx[0][i0], th = xx[1][i1], ph = xx[2][i2];
  const double X = r * sin(th) * cos(ph), Y = r * sin(th) * sin(ph), Z = r * cos(th);
  const double rr = sqrt(X * X + Y * Y + Z * Z);
  const double c[3] = {rr, acos(Z / rr), atan2(Y, X)};
  for (int d = 0; d < 3; d++) {
    const double xxmin = xx[d][0] - ((double)(0 - g->NG) + (1.0 / 2.0)) * g->dxx[d];
    out[d] = (int)((c[d] - xxmin - (1.0 / 2.0) * g->dxx[d] + ((double)g->NG) * g->dxx[d]) / g->dxx[d] + 0.5);
  }
  for (int d = 0; d < 3; d++)
    xin[d] = xx[d][out[d]];
}

static void parity_dots(double xx1, double xx2, const double xin[3], int8_t par[10]) {
  const double t0 = cos(xx1) * cos(xin[1]);
  const double t1 = sin(xx1) * sin(xin[1]);
  const double t2 = sin(xx2) * sin(xin[2]);
  const double t3 = cos(xx2) * cos(xin[2]);
  const double p1 = t0 + t1 * t2 + t1 * t3; /* rhat . rhat' */
  const double p2 = t0 * t2 + t0 * t3 + t1; /* thhat . thhat' */
  const double p3 = t2 + t3;                /* phhat . phhat' */
  const double d[10] = {1, p1, p2, p3, p1 * p1, p1 * p2, p1 * p3, p2 * p2, p2 * p3, p3 * p3};
  for (int k = 0; k < 10; k++) {
    if (fabs(d[k]) < 1 - 1e-8 || fabs(d[k]) > 1 + 1e-8) {
      fprintf(stderr, "oracle: parity dot product %e not +-1\n", d[k]);
      exit(1);
    }
    par[k] = d[k] < 0 ? -1 : 1;
  }
}

og_bcstruct *og_bcstruct_new(const og_grid *g, const double *x0, const double *x1, const double *x2) {
  og_bcstruct *bc = calloc(1, sizeof(og_bcstruct));
  const double *xx[3] = {x0, x1, x2};
  const int NG = g->NG, *Nx = g->Nx;
  /* inner points, in i2-i1-i0 loop order */
  long cap = 1024;
  bc->inner_dst = malloc(sizeof(long) * cap);
  bc->inner_src = malloc(sizeof(long) * cap);
  bc->inner_parity = malloc(10 * cap);
  for (int i2 = 0; i2 < Nx[2]; i2++)
    for (int i1 = 0; i1 < Nx[1]; i1++)
      for (int i0 = 0; i0 < Nx[0]; i0++) {
        const int interior = i0 >= NG && i0 < Nx[0] - NG && i1 >= NG && i1 < Nx[1] - NG && i2 >= NG && i2 < Nx[2] - NG;
        if (interior)
          continue;
        int ib[3];
        double xin[3];
        eigen_inbounds(g, xx, i0, i1, i2, ib, xin);
        if (ib[0] == i0 && ib[1] == i1 && ib[2] == i2)
          continue;
        if (bc->num_inner == cap) {
          cap *= 2;
          bc->inner_dst = realloc(bc->inner_dst, sizeof(long) * cap);
          bc->inner_src = realloc(bc->inner_src, sizeof(long) * cap);
          bc->inner_parity = realloc(bc->inner_parity, 10 * cap);
        }
        const int n = bc->num_inner++;
        bc->inner_dst[n] = i0 + (long)Nx[0] * (i1 + (long)Nx[1] * i2);
        bc->inner_src[n] = ib[0] + (long)Nx[0] * (ib[1] + (long)Nx[1] * ib[2]);
        parity_dots(xx[1][i1], xx[2][i2], xin, &bc->inner_parity[10 * n]);
      }
  /* outer points, layer by layer, faces x0min,x0max,x1min,... */
  int imin[3] = {NG, NG, NG};
  int imax[3] = {Nx[0] - NG, Nx[1] - NG, Nx[2] - NG};
  /* face ranges {i0lo,i0hi,i1lo,i1hi,i2lo,i2hi}: each face uses the current bounds, then
   * widens its own direction (bcstruct_set_up…HoleySinhSpherical.c:382-475) */
  int range[8][6][6];
  for (int gz = 0; gz < NG; gz++)
    for (int d = 0; d < 3; d++)
      for (int side = 0; side < 2; side++) {
        int *r = range[gz][2 * d + side];
        for (int e = 0; e < 3; e++) {
          r[2 * e] = imin[e];
          r[2 * e + 1] = imax[e];
        }
        if (side == 0) {
          r[2 * d] = imin[d] - 1;
          r[2 * d + 1] = imin[d];
          imin[d]--;
        } else {
          r[2 * d] = imax[d];
          r[2 * d + 1] = imax[d] + 1;
          imax[d]++;
        }
      }
  for (int gz = 0; gz < NG; gz++)
    for (int d = 0; d < 3; d++) {
      long n = 0;
      for (int f = 2 * d; f <= 2 * d + 1; f++) {
        const int *r = range[gz][f];
        n += (long)(r[1] - r[0]) * (r[3] - r[2]) * (r[5] - r[4]);
      }
      short *arr = malloc(sizeof(short) * 6 * (n + 1));
      int cnt = 0;
      for (int f = 2 * d; f <= 2 * d + 1; f++) {
        const int *r = range[gz][f];
        const int F0 = (f == 0) - (f == 1), F1 = (f == 2) - (f == 3), F2 = (f == 4) - (f == 5);
        for (int i2 = r[4]; i2 < r[5]; i2++)
          for (int i1 = r[2]; i1 < r[3]; i1++)
            for (int i0 = r[0]; i0 < r[1]; i0++) {
              int ib[3];
              double xin[3];
              eigen_inbounds(g, xx, i0, i1, i2, ib, xin);
              if (ib[0] == i0 && ib[1] == i1 && ib[2] == i2) {
                short *p = &arr[6 * cnt++];
                p[0] = i0;
                p[1] = i1;
                p[2] = i2;
                p[3] = F0;
                p[4] = F1;
                p[5] = F2;
              }
            }
      }
      bc->num_outer[gz][d] = cnt;
      bc->outer[gz][d] = arr;
    }
  return bc;
}

int og_bc_num_inner(const og_bcstruct *bc) { return bc->num_inner; }
void og_bc_inner(const og_bcstruct *bc, long *dst, long *src, int8_t *parity) {
  memcpy(dst, bc->inner_dst, sizeof(long) * bc->num_inner);
  memcpy(src, bc->inner_src, sizeof(long) * bc->num_inner);
  memcpy(parity, bc->inner_parity, 10 * (size_t)bc->num_inner);
}
int og_bc_num_outer(const og_bcstruct *bc, int gz, int d) { return bc->num_outer[gz][d]; }
void og_bc_outer(const og_bcstruct *bc, int gz, int d, short *out) {
  memcpy(out, bc->outer[gz][d], sizeof(short) * 6 * bc->num_outer[gz][d]);
}
void og_bcstruct_free(og_bcstruct *bc, int NG) {
  free(bc->inner_dst);
  free(bc->inner_src);
  free(bc->inner_parity);
  for (int gz = 0; gz < NG; gz++)
    for (int d = 0; d < 3; d++)
      free(bc->outer[gz][d]);
  free(bc);
}

/* apply_bcs_inner_only (CurviBoundaryConditions/tests/apply_bcs_inner_only_apply_bcs_inner_only__openmp.c:27-70):
 * gf[dst] = parity * gf[src], sign applied as a sign-bit flip. */
void og_apply_bcs_inner_only(const og_grid *g, const og_bcstruct *bc, int ngfs, const int *gf_parity, double *gfs) {
  const long Ntot = (long)g->Nx[0] * g->Nx[1] * g->Nx[2];
  for (int gf = 0; gf < ngfs; gf++) {
    double *f = &gfs[gf * Ntot];
    for (int p = 0; p < bc->num_inner; p++) {
      const double v = f[bc->inner_src[p]];
      f[bc->inner_dst[p]] = bc->inner_parity[10 * p + gf_parity[gf]] < 0 ? -v : v;
    }
  }
}

/* apply_bcs_outerextrap_and_inner (…/apply_bcs_outerextrap_and_inner__openmp.c:8-91):
 * f0 = 3 f1 - 3 f2 + f3 along the face normal, innermost layer first, then inner BCs. */
void og_apply_bcs_outerextrap_and_inner(const og_grid *g, const og_bcstruct *bc, int ngfs, const int *gf_parity, double *gfs) {
  const long Ntot = (long)g->Nx[0] * g->Nx[1] * g->Nx[2];
  const long N0 = g->Nx[0], N01 = (long)g->Nx[0] * g->Nx[1];
  for (int gz = 0; gz < g->NG; gz++)
    for (int d = 0; d < 3; d++)
      for (int p = 0; p < bc->num_outer[gz][d]; p++) {
        const short *q = &bc->outer[gz][d][6 * p];
        const long i0 = q[0] + N0 * q[1] + N01 * q[2];
        const long s = q[3] + N0 * q[4] + N01 * q[5];
        for (int gf = 0; gf < ngfs; gf++) {
          double *f = &gfs[gf * Ntot];
          f[i0] = +3.0 * f[i0 + s] - 3.0 * f[i0 + 2 * s] + 1.0 * f[i0 + 3 * s];
        }
      }
  og_apply_bcs_inner_only(g, bc, ngfs, gf_parity, gfs);
}

/* r(xx0) and dxx0/dr for spherical-like coordinates
 * (…/apply_bcs_outerradiation_and_inner__rfm__openmp__HoleySinhSpherical.c:42-56). */
static void r_and_dx0dr(const og_grid *g, double xx0, double *r, double *dx0dr) {
  if (g->coord == OG_SPHERICAL) {
    *r = xx0;
    *dx0dr = 1.0;
  } else {
    const double tmp0 = (1.0 / (g->p1));
    const double tmp1 = exp(tmp0) - exp(-tmp0);
    const double tmp3 = exp(tmp0 * xx0);
    const double tmp4 = exp(-tmp0 * xx0);
    *r = g->p0 * (tmp3 - tmp4) / tmp1;
    *dx0dr = tmp1 / (g->p0 * (tmp0 * tmp3 + tmp0 * tmp4));
  }
}

/* partial_r f with a maximally-centred, bounds-clamped first-derivative stencil along xx0
 * (…__HoleySinhSpherical.c:60-101; FD order = radiation_BC_fd_order). */
static double partial_r_f(const og_grid *g, const double *f, int i0, int i1, int i2, int FACE0, double dx0dr, int rad_order) {
  const int R = rad_order / 2;
  int off = FACE0;
  if (i0 < R)
    off = R - i0;
  else if (i0 > g->Nx[0] - R - 1)
    off = (g->Nx[0] - R - 1) - i0;
  const og_stencil *s = &OG_RAD_STENCIL[R - 1][off + 3];
  const long idx = i0 + (long)g->Nx[0] * (i1 + (long)g->Nx[1] * i2);
  double acc = 0.0;
  for (int k = 0; k < s->n; k++)
    acc += st_w(s, k) * f[idx + s->off[k]];
  return dx0dr * (acc * g->invdxx[0]);
}

/* apply_bcs_outerradiation_and_inner (…__HoleySinhSpherical.c:104-237): Sommerfeld
 * condition on the RHS at pure outer points, innermost layer first, then inner BCs on the RHS. */
void og_apply_bcs_outerradiation_and_inner(const og_grid *g, const og_bcstruct *bc, const double *x0, int ngfs,
                                           const int *gf_parity, const double *wavespeed, const double *f_inf,
                                           int rad_order, const double *gfs, double *rhs) {
  const long Ntot = (long)g->Nx[0] * g->Nx[1] * g->Nx[2];
  const long N0 = g->Nx[0], N01 = (long)g->Nx[0] * g->Nx[1];
  for (int gz = 0; gz < g->NG; gz++)
    for (int d = 0; d < 3; d++)
      for (int p = 0; p < bc->num_outer[gz][d]; p++) {
        const short *q = &bc->outer[gz][d][6 * p];
        const int i0 = q[0], i1 = q[1], i2 = q[2];
        const int j0 = i0 + q[3], j1 = i1 + q[4], j2 = i2 + q[5];
        double r, dxr, r_int, dxr_int;
        r_and_dx0dr(g, x0[i0], &r, &dxr);
        r_and_dx0dr(g, x0[j0], &r_int, &dxr_int);
        const long idx = i0 + N0 * i1 + N01 * i2;
        const long idx_int = j0 + N0 * j1 + N01 * j2;
        for (int gf = 0; gf < ngfs; gf++) {
          const double *f = &gfs[gf * Ntot];
          const double dfr = partial_r_f(g, f, i0, i1, i2, q[3], dxr, rad_order);
          const double dfr_int = partial_r_f(g, f, j0, j1, j2, q[3], dxr_int, rad_order);
          const double c = wavespeed[gf], finf = f_inf[gf];
          const double dtf_int = rhs[gf * Ntot + idx_int];
          const double dtf_int_out = -c * (dfr_int + (f[idx_int] - finf) / r_int);
          const double k = r_int * r_int * r_int * (dtf_int - dtf_int_out);
          const double rinv = 1.0 / r;
          const double dtf_out = -c * (dfr + (f[idx] - finf) * rinv);
          rhs[gf * Ntot + idx] = dtf_out + k * rinv * rinv * rinv;
        }
      }
  og_apply_bcs_inner_only(g, bc, ngfs, gf_parity, rhs);
}

/* Quadratic extrapolation on all six faces of a Cartesian box
 * (nrpy/examples/wave_equation_cartesian.py:374-431). */
void og_wave_apply_bcs(const og_grid *g, double *gfs) {
  const long Ntot = (long)g->Nx[0] * g->Nx[1] * g->Nx[2];
  const int N0 = g->Nx[0], N1 = g->Nx[1];
#define IX(a, b, c) ((a) + (long)N0 * ((b) + (long)N1 * (c)))
  for (int gf = 0; gf < 2; gf++) {
    double *f = &gfs[gf * Ntot];
    int imin[3] = {g->NG, g->NG, g->NG};
    int imax[3] = {g->Nx[0] - g->NG, g->Nx[1] - g->NG, g->Nx[2] - g->NG};
    for (int gz = 0; gz < g->NG; gz++) {
      for (int d = 0; d < 3; d++) {
        for (int side = 0; side < 2; side++) {
          int lo[3] = {imin[0], imin[1], imin[2]}, hi[3] = {imax[0], imax[1], imax[2]};
          int F[3] = {0, 0, 0};
          if (side == 0) {
            lo[d] = imin[d] - 1;
            hi[d] = imin[d];
            F[d] = 1;
          } else {
            lo[d] = imax[d];
            hi[d] = imax[d] + 1;
            F[d] = -1;
          }
          for (int i2 = lo[2]; i2 < hi[2]; i2++)
            for (int i1 = lo[1]; i1 < hi[1]; i1++)
              for (int i0 = lo[0]; i0 < hi[0]; i0++)
                f[IX(i0, i1, i2)] = +3.0 * f[IX(i0 + F[0], i1 + F[1], i2 + F[2])] - 3.0 * f[IX(i0 + 2 * F[0], i1 + 2 * F[1], i2 + 2 * F[2])] +
                                    1.0 * f[IX(i0 + 3 * F[0], i1 + 3 * F[1], i2 + 3 * F[2])];
          if (side == 0)
            imin[d]--;
          else
            imax[d]++;
        }
      }
    }
  }
#undef IX
}

/* ------------------------------------------------------------------------ */
/* RK4 substeps over all points (MoLtimestepping/tests/register_all_MoL_step_forward_in_time.c:4-103). */
void og_rk_substep(int s, long n, double dt, double *k, double *y_n, double *y_run) {
  const double R16 = 1.0 / 6.0, R12 = 1.0 / 2.0, R13 = 1.0 / 3.0;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) {
    const double kv = k[i], yn = y_n[i];
    if (s == 1) {
      y_run[i] = R16 * dt * kv;
      k[i] = R12 * dt * kv + yn;
    } else if (s == 2) {
      const double yr = y_run[i];
      y_run[i] = R13 * dt * kv + yr;
      k[i] = R12 * dt * kv + yn;
    } else if (s == 3) {
      const double yr = y_run[i];
      y_run[i] = R13 * dt * kv + yr;
      k[i] = dt * kv + yn;
    } else {
      const double yr = y_run[i];
      y_n[i] = R16 * dt * kv + yn + yr;
    }
  }
}

/* MoL_step_forward_in_time for BSSN, RK4 (…__RK4.cu:182-287 with the rhs/post-rhs strings of
 * nrpy/examples/two_blackholes_collide.py:290-313): per substep Ricci_eval -> rhs_eval ->
 * radiation+inner BCs on the RHS -> RK update -> det(gammabar) enforcement.
 * Arrays: y_n, y_run, k_odd, k_even: 24*Ntot; aux: 6*Ntot. */
void og_bssn_mol_step(const og_grid *g, int variant, const double *params, const double *rfm0, const double *rfm1,
                      const og_bcstruct *bc, const double *x0, int rad_order, double dt, double *y_n, double *y_run,
                      double *k_odd, double *k_even, double *aux) {
  const long Ntot = (long)g->Nx[0] * g->Nx[1] * g->Nx[2];
  static const int parity[24] = {4, 5, 6, 7, 8, 9, 0, 1, 2, 3, 0, 4, 5, 6, 7, 8, 9, 1, 2, 3, 0, 1, 2, 3};
  double ws[24], fi[24];
  for (int k = 0; k < 24; k++) {
    ws[k] = 1.0;
    fi[k] = 0.0;
  }
  ws[6] = sqrt(2.0); /* alpha */
  fi[6] = 1.0;       /* alpha */
  fi[10] = 1.0;      /* cf */
  double *in_s[4] = {y_n, k_odd, k_even, k_odd};
  double *out_s[4] = {k_odd, k_even, k_odd, k_even};
  for (int s = 1; s <= 4; s++) {
    double *in = in_s[s - 1], *out = out_s[s - 1];
    og_ricci(g, variant, rfm0, rfm1, in, aux);
    og_bssn_rhs(g, variant, params, rfm0, rfm1, aux, in, out);
    og_apply_bcs_outerradiation_and_inner(g, bc, x0, 24, parity, ws, fi, rad_order, in, out);
    og_rk_substep(s, 24 * Ntot, dt, out, y_n, y_run);
    og_enforce_detgamma(g, variant, rfm0, rfm1, s == 4 ? y_n : out);
  }
}

/* Wave-equation RK4 step (nrpy/examples/wave_equation_cartesian.py:440-444): rhs_eval, RK update,
 * then quadratic-extrapolation BCs on the new stage. Arrays are 2*Ntot. */
void og_wave_mol_step(const og_grid *g, double wavespeed, double dt, double *y_n, double *y_run, double *k_odd, double *k_even) {
  const long Ntot = (long)g->Nx[0] * g->Nx[1] * g->Nx[2];
  double *in_s[4] = {y_n, k_odd, k_even, k_odd};
  double *out_s[4] = {k_odd, k_even, k_odd, k_even};
  for (int s = 1; s <= 4; s++) {
    og_wave_rhs(g, wavespeed, in_s[s - 1], out_s[s - 1]);
    og_rk_substep(s, 2 * Ntot, dt, out_s[s - 1], y_n, y_run);
    og_wave_apply_bcs(g, s == 4 ? y_n : out_s[s - 1]);
  }
}

/* dt = CFL * min over all points of min(ds0, ds1, ds2)
 * (BHaH/tests/numerical_grids_and_timestep_cfl_limited_timestep__openmp.c and
 *  …ds_min_single_pt__openmp__HoleySinhSpherical.c). */
double og_cfl_dt(const og_grid *g, const double *x0, const double *x1, const double *x2, double CFL) {
  (void)x2;
  double ds_min = 1e38;
  if (g->coord == OG_CARTESIAN)
    return CFL * NRPYMIN(g->dxx[0], NRPYMIN(g->dxx[1], g->dxx[2]));
  for (int i2 = 0; i2 < g->Nx[2]; i2++)
    for (int i1 = 0; i1 < g->Nx[1]; i1++)
      for (int i0 = 0; i0 < g->Nx[0]; i0++) {
        double ds0, ds1, ds2;
        if (g->coord == OG_SPHERICAL) {
          ds0 = fabs(g->dxx[0]);
          ds1 = fabs(x0[i0] * g->dxx[1]);
          ds2 = fabs(x0[i0] * g->dxx[2] * sin(x1[i1]));
        } else {
          const double tmp0 = (1.0 / (g->p1));
          const double tmp4 = g->p0 / (exp(tmp0) - exp(-tmp0));
          const double tmp2 = exp(tmp0 * x0[i0]);
          const double tmp3 = exp(-tmp0 * x0[i0]);
          const double tmp5 = tmp4 * (tmp2 - tmp3);
          ds0 = fabs(g->dxx[0] * tmp4 * (tmp0 * tmp2 + tmp0 * tmp3));
          ds1 = fabs(g->dxx[1] * tmp5);
          ds2 = fabs(g->dxx[2] * tmp5 * sin(x1[i1]));
        }
        ds_min = NRPYMIN(ds_min, NRPYMIN(ds0, NRPYMIN(ds1, ds2)));
      }
  return ds_min * CFL;
}

/* ---- Checkpoint payload (after params_struct), as write_checkpoint writes it
 * (BHaH/tests/write_checkpoint_write_checkpoint__openmp.c:84-117): int count, int indices[count]
 * (every point, increasing), then REAL compact[count * ngf] with compact[el * ngf + gf] =
 * y[ntot * gf + indices[el]].  Returns bytes written. */
long og_checkpoint_pack(long ntot, int ngf, const double *y, void *out) {
  char *o = (char *)out;
  const int count = (int)ntot;
  memcpy(o, &count, sizeof(int));
  int *ind = (int *)(o + sizeof(int));
  double *compact = (double *)(o + sizeof(int) * (1 + ntot));
  long which_el = 0;
  for (long i = 0; i < ntot; i++) {
    ind[which_el] = (int)i;
    for (int gf = 0; gf < ngf; gf++) compact[which_el * ngf + gf] = y[ntot * gf + i];
    which_el++;
  }
  return (long)sizeof(int) * (1 + ntot) + (long)sizeof(double) * ntot * ngf;
}

/* Inverse, as read_checkpoint fills y_n (read_checkpoint_read_checkpoint__openmp.c:64-160):
 * returns 0, or -1 on an invalid count or index (the reference exits). */
int og_checkpoint_unpack(long ntot, int ngf, const void *in, double *y) {
  const char *c = (const char *)in;
  int count;
  memcpy(&count, c, sizeof(int));
  if (count < 0 || count > ntot) return -1;
  const int *ind = (const int *)(c + sizeof(int));
  const double *compact = (const double *)(c + sizeof(int) * (1 + (long)count));
  /* validate every index first (read_checkpoint__openmp.c:130-145) */
  for (int i = 0; i < count; i++)
    if (ind[i] < 0 || ind[i] >= ntot || (i > 0 && ind[i] <= ind[i - 1])) return -1;
  /* zero all of y_n, then scatter the listed points (read_checkpoint__openmp.c:163-171) */
  for (long i = 0; i < ntot * ngf; i++) y[i] = 0.0;
  for (int i = 0; i < count; i++)
    for (int gf = 0; gf < ngf; gf++) y[ntot * gf + ind[i]] = compact[(long)i * ngf + gf];
  return 0;
}
