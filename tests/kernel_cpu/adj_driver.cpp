/* TEST INFRASTRUCTURE ONLY.  Runs the generated adjoint kernels of one BSSN build
 * on the CPU.  Built with -DKERNEL_FILE="<gen adjoint .hip>" -DTAG=<build tag>.
 *
 * Input: int32 mode(0: rhs+Ricci+gather, 1: det via adj_stage, 2: det via adj_final) Nx0 Nx1 Nx2 NG;
 * float64 invdxx0..2
 * eta KO_gauge KO_nongauge wavespeed; in[24*Ntot]; aux[6*Ntot]; lam[24*Ntot];
 * rfm0[4*Nx0]; rfm1[3*Nx1].  Output: lam_y[24*Ntot] (mode 0: transposed
 * rhs_eval o Ricci_eval applied to lam at interior points; modes 1, 2: transposed
 * det(gammabar) enforcement at u = in). */
#include KERNEL_FILE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

static double *rd(FILE *f, size_t n) {
  double *p = (double *)malloc((n ? n : 1) * sizeof(double));
  if (n && fread(p, sizeof(double), n, f) != n) exit(2);
  return p;
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = fopen(argv[1], "rb");
  int h[5];
  double par[7];
  if (!f || fread(h, sizeof(int), 5, f) != 5 || fread(par, sizeof(double), 7, f) != 7) return 2;
  bhah_kparams p;
  memset(&p, 0, sizeof(p));
  p.Nx0 = h[1]; p.Nx1 = h[2]; p.Nx2 = h[3]; p.NG = h[4];
  p.s1 = p.Nx0;
  p.s2 = (long)p.Nx0 * p.Nx1;
  p.Ntot = p.s2 * p.Nx2;
  p.invdxx0 = par[0]; p.invdxx1 = par[1]; p.invdxx2 = par[2];
  p.eta = par[3]; p.KO_gauge = par[4]; p.KO_nongauge = par[5]; p.wavespeed = par[6];
  const size_t N = (size_t)p.Ntot;
  double *in = rd(f, 24 * N), *aux = rd(f, 6 * N), *lam = rd(f, 24 * N);
  double *rfm0 = rd(f, 4 * (size_t)p.Nx0), *rfm1 = rd(f, 3 * (size_t)p.Nx1);
  fclose(f);
  double *out = (double *)calloc(24 * N, sizeof(double));
  if (h[0] == 0) {
    const int ns = CAT(adj_nslots_, TAG)();
    double *S = (double *)calloc((size_t)ns * N, sizeof(double));
    dim3 b(64, 4, 1), g((p.Nx0 - 2 * p.NG + 63) / 64, (p.Nx1 - 2 * p.NG + 3) / 4, p.Nx2 - 2 * p.NG);
    CAT(launch_adj_rhs_, TAG)(g, b, 0, in, aux, lam, S, rfm0, rfm1, p);
    CAT(launch_adj_ricci_, TAG)(g, b, 0, in, S, rfm0, rfm1, p);
    dim3 b1(256), g1((unsigned)((24 * N + 255) / 256));
    CAT(launch_adj_gather_, TAG)(g1, b1, 0, S, out, p);
    free(S);
  } else {
    /* the RK/det adjoints with zero combination weights: u = in, so out = J_D(in)^T lam
     * (mode 1 through adj_stage: lamYN += ..., mode 2 through adj_final: lamYN = ...) */
    double *K = (double *)calloc(4 * 24 * N, sizeof(double)), *lamK = (double *)calloc(4 * 24 * N, sizeof(double));
    dim3 b(64, 4, 1), g((p.Nx0 + 63) / 64, (p.Nx1 + 3) / 4, p.Nx2);
    if (h[0] == 1)
      CAT(launch_adj_stage_, TAG)(g, b, 0, in, K, 0.0, lam, out, lamK, rfm0, rfm1, p);
    else
      CAT(launch_adj_final_, TAG)(g, b, 0, in, K, 0.0, 0.0, 0.0, 0.0, lam, out, lamK, rfm0, rfm1, p);
    free(K);
    free(lamK);
  }
  FILE *o = fopen(argv[2], "wb");
  fwrite(out, sizeof(double), 24 * N, o);
  fclose(o);
  free(in); free(aux); free(lam); free(rfm0); free(rfm1); free(out);
  return 0;
}
