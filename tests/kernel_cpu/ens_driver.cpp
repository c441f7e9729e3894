/* TEST INFRASTRUCTURE ONLY.  Runs one generated batched-ensemble launcher on the
 * CPU over several members of different grid sizes.  Built with
 * -DKERNEL_FILE="<gen .hip>" -DLAUNCH=<ensemble launcher> -DKIND_RHS=0|1.
 *
 * Input: int32 nmem tile_x tile_y max_wg tile_z (interior tiles of tile_x x tile_y x tile_z
 * points: tile_z > 1 for LDS-tiled schedules); per member: int32 Nx0 Nx1 Nx2 NG,
 * float64 invdxx0..2 eta KO_gauge KO_nongauge wavespeed, in[24*Ntot],
 * aux[6*Ntot] (rhs only), rfm0[4*Nx0], rfm1[3*Nx1].  Output: per member
 * out[nout*Ntot].  Every array is its own exact-size heap block. */
#include KERNEL_FILE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>

static double *rd(FILE *f, size_t n) {
  double *p = (double *)malloc((n ? n : 1) * sizeof(double));
  if (n && fread(p, sizeof(double), n, f) != n) exit(2);
  return p;
}

int main(int argc, char **argv) {
  if (argc != 3) return 2;
  FILE *f = fopen(argv[1], "rb");
  int h[5];
  if (!f || fread(h, sizeof(int), 5, f) != 5) return 2;
  const int n = h[0], tx = h[1], ty = h[2], maxwg = h[3], tz = h[4];
  std::vector<bhah_member> mem(n);
  std::vector<bhah_item> items;
  std::vector<size_t> nout(n);
  for (int m = 0; m < n; m++) {
    int g[4];
    double par[7];
    if (fread(g, sizeof(int), 4, f) != 4 || fread(par, sizeof(double), 7, f) != 7) return 2;
    bhah_kparams &p = mem[m].kp;
    memset(&p, 0, sizeof(p));
    p.Nx0 = g[0]; p.Nx1 = g[1]; p.Nx2 = g[2]; p.NG = g[3];
    p.s1 = p.Nx0;
    p.s2 = (long)p.Nx0 * p.Nx1;
    p.Ntot = p.s2 * p.Nx2;
    p.invdxx0 = par[0]; p.invdxx1 = par[1]; p.invdxx2 = par[2];
    p.eta = par[3]; p.KO_gauge = par[4]; p.KO_nongauge = par[5]; p.wavespeed = par[6];
    const size_t N = (size_t)p.Ntot;
    for (int a = 0; a < 5; a++) mem[m].arr[a] = NULL;
    mem[m].arr[0] = rd(f, 24 * N);
    if (KIND_RHS) mem[m].arr[4] = rd(f, 6 * N);
    mem[m].rfm0 = rd(f, 4 * (size_t)p.Nx0);
    mem[m].rfm1 = rd(f, 3 * (size_t)p.Nx1);
    nout[m] = (KIND_RHS ? 24 : 6) * N;
    mem[m].arr[2] = (double *)calloc(nout[m], sizeof(double));
    if (!KIND_RHS) mem[m].arr[4] = mem[m].arr[2];
    const int n0 = p.Nx0 - 2 * p.NG, n1 = p.Nx1 - 2 * p.NG, n2 = p.Nx2 - 2 * p.NG;
    for (int bz = 0; bz < (n2 + tz - 1) / tz; bz++)
      for (int by = 0; by < (n1 + ty - 1) / ty; by++)
        for (int bx = 0; bx < (n0 + tx - 1) / tx; bx++) items.push_back({m, bx, by, bz});
  }
  fclose(f);
  const int ni = (int)items.size();
  bhah_item *it = (bhah_item *)malloc(sizeof(bhah_item) * ni);
  memcpy(it, items.data(), sizeof(bhah_item) * ni);
  bhah_member *dm = (bhah_member *)malloc(sizeof(bhah_member) * n);
  memcpy(dm, mem.data(), sizeof(bhah_member) * n);
  const int grid = (maxwg > 0 && maxwg < ni) ? maxwg : ni;
  // rhs: in = arr[0], aux = arr[4], out = arr[2]; Ricci: in = arr[0], out = arr[2]
  LAUNCH(dim3(grid), dim3(tx, ty, 1), (hipStream_t)0, dm, it, ni, 0, 4, 2);
  FILE *o = fopen(argv[2], "wb");
  for (int m = 0; m < n; m++) fwrite(mem[m].arr[2], sizeof(double), nout[m], o);
  fclose(o);
  return 0;
}
