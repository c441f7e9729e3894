/* TEST INFRASTRUCTURE ONLY.  Runs one generated rhs/Ricci schedule launcher on
 * the CPU.  Built per schedule with -DKERNEL_FILE="<gen .hip>" -DLAUNCH=<launcher>
 * against the shim hip/hip_runtime.h; every array is a separate exact-size heap
 * block so AddressSanitizer reports any out-of-bounds access.
 *
 * Input file: int32 Nx0 Nx1 Nx2 NG tile_x tile_y ngf_in naux nout; float64
 * invdxx0..2 eta KO_gauge KO_nongauge wavespeed; then in[ngf_in*Ntot],
 * aux[naux*Ntot], rfm0[4*Nx0], rfm1[3*Nx1].  Output file: out[nout*Ntot]. */
#include KERNEL_FILE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double *read_block(FILE *f, size_t n) {
  double *p = (double *)malloc((n ? n : 1) * sizeof(double));
  if (n && fread(p, sizeof(double), n, f) != n) {
    fprintf(stderr, "short input\n");
    exit(2);
  }
  return p;
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  int hdr[9];
  double par[7];
  if (fread(hdr, sizeof(int), 9, f) != 9 || fread(par, sizeof(double), 7, f) != 7) return 2;
  bhah_kparams p;
  memset(&p, 0, sizeof(p));
  p.Nx0 = hdr[0]; p.Nx1 = hdr[1]; p.Nx2 = hdr[2]; p.NG = hdr[3];
  p.s1 = p.Nx0;
  p.s2 = (long)p.Nx0 * p.Nx1;
  p.Ntot = p.s2 * p.Nx2;
  p.invdxx0 = par[0]; p.invdxx1 = par[1]; p.invdxx2 = par[2];
  p.eta = par[3]; p.KO_gauge = par[4]; p.KO_nongauge = par[5]; p.wavespeed = par[6];
  const size_t N = (size_t)p.Ntot;
  double *in = read_block(f, hdr[6] * N);
  double *aux = hdr[7] ? read_block(f, hdr[7] * N) : NULL;
  double *rfm0 = read_block(f, 4 * (size_t)p.Nx0);
  double *rfm1 = read_block(f, 3 * (size_t)p.Nx1);
  fclose(f);
  double *out = (double *)calloc(hdr[8] * N, sizeof(double));
  const int tx = hdr[4], ty = hdr[5], NG = p.NG;
  dim3 b(tx, ty, 1);
  dim3 g((p.Nx0 - 2 * NG + tx - 1) / tx, (p.Nx1 - 2 * NG + ty - 1) / ty, p.Nx2 - 2 * NG);
  LAUNCH(g, b, (hipStream_t)0, in, aux, out, rfm0, rfm1, p);
  FILE *o = fopen(argv[2], "wb");
  fwrite(out, sizeof(double), hdr[8] * N, o);
  fclose(o);
  free(in); free(aux); free(rfm0); free(rfm1); free(out);
  return 0;
}
