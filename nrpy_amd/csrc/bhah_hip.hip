// Host runtime and boundary-condition kernels of the gfx950 BHaH evolution backend.
//
// The generated per-variant translation units (gen/*.hip) hold the interior
// stencil kernels; this file owns patch setup (grid, reference-metric
// precompute arrays, boundary structs, CFL timestep), the boundary-condition
// kernels, the RK4 method-of-lines driver and the C-ABI declared in
// include/bhah_hip.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bhah_hip.h"

#include "bhah_device_types.h"

// ---------------------------------------------------------------------------
// Generated kernels: per (coordinate, symmetry, KO, fd order) build, a set of
// candidate schedules for rhs_eval and Ricci_eval (output partitions x register
// budgets), plus the RK4+det pass, det enforcement and constraints.
typedef hipError_t (*bhah_cand_fn)(dim3, dim3, hipStream_t, const double *, const double *, double *,
                                   const double *, const double *, bhah_kparams);
typedef hipError_t (*bhah_ens_fn)(dim3, dim3, hipStream_t, const bhah_member *, const bhah_item *, int, int, int, int);
struct BssnCandidate {
  const char *name;
  bhah_cand_fn launch;
  bhah_ens_fn ens;  // batched-ensemble form, where the build has one
  int lds[3];       // workgroup tile (points along xx0, xx1, xx2) of LDS-tiled schedules; 0 otherwise
};
struct BssnKernels {
  int coord, sym_axes, KO, fd_order;
  const char *tag;
  const BssnCandidate *rhs;
  int n_rhs;
  const BssnCandidate *ricci;
  int n_ricci;
  hipError_t (*rk_det)(dim3, dim3, hipStream_t, const double *, const double *, double *, double *, const double *,
                       const double *, bhah_kparams);
  // the same pass taking the inner-parity boundary condition of k with it (pairs of points only)
  hipError_t (*rk_det_inner)(hipStream_t, const double *, const double *, double *, double *, const double *,
                             const double *, bhah_kparams, const int *, const int *);
  hipError_t (*detgamma)(dim3, dim3, hipStream_t, double *, const double *, const double *, bhah_kparams);
  hipError_t (*constraints)(dim3, dim3, hipStream_t, const double *, const double *, double *, const double *,
                            const double *, bhah_kparams);
  hipError_t (*initial_lambda)(dim3, dim3, hipStream_t, double *, const double *, const double *, bhah_kparams);
  hipError_t (*ens_rk_det)(dim3, dim3, hipStream_t, const bhah_member *, const bhah_item *, int, int, int, int);
  // reverse-mode kernels of one RK4 step (builds with adjoints only)
  int (*adj_nslots)(void);
  hipError_t (*adj_rhs)(dim3, dim3, hipStream_t, const double *, const double *, const double *, double *,
                        const double *, const double *, bhah_kparams);
  hipError_t (*adj_ricci)(dim3, dim3, hipStream_t, const double *, double *, const double *, const double *,
                          bhah_kparams);
  hipError_t (*adj_final)(dim3, dim3, hipStream_t, const double *, const double *, double, double, double, double,
                          const double *, double *, double *, const double *, const double *, bhah_kparams);
  hipError_t (*adj_stage)(dim3, dim3, hipStream_t, const double *, const double *, double, const double *, double *,
                          double *, const double *, const double *, bhah_kparams);
  hipError_t (*adj_gather)(dim3, dim3, hipStream_t, const double *, double *, bhah_kparams);
};
#include "gen/kernel_table.inc"

extern "C" hipError_t launch_wave_rhs_fd4(dim3, dim3, hipStream_t, const double *, double *, bhah_kparams);
extern "C" hipError_t launch_wave_rhs_rk_fd4(dim3, dim3, hipStream_t, const double *, const double *, double *,
                                             double *, bhah_kparams);

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(call)                                                                                                 \
  do {                                                                                                                \
    hipError_t e_ = (call);                                                                                           \
    if (e_ != hipSuccess) return fail(BHAH_ERR_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                                      __LINE__);                                                                      \
  } while (0)

// Parity type per evolved gridfunction (nrpy/grid.py:116-126, 270-321).
const int kBssnParity[24] = {4, 5, 6, 7, 8, 9, 0, 1, 2, 3, 0, 4, 5, 6, 7, 8, 9, 1, 2, 3, 0, 1, 2, 3};

}  // namespace

struct bhah_patch {
  bhah_patch_desc d;
  const BssnKernels *k = nullptr;
  int NG = 0, ngfs = 0, naux = 0;
  int Nx[3] = {0, 0, 0};
  long Ntot = 0;
  double xxmin[3], xxmax[3], dxx[3], invdxx[3];
  double dt = 0;
  std::vector<double> xx[3];
  // host boundary structs
  std::vector<int64_t> inner_dst, inner_src;
  std::vector<int8_t> inner_parity;  // 10 per point
  std::vector<int16_t> outer[8][3];  // 6 shorts per point
  // device arrays
  double *d_rfm0 = nullptr, *d_rfm1 = nullptr, *d_xx0 = nullptr, *d_xx1 = nullptr, *d_xx2 = nullptr;
  int *d_inner_dst = nullptr, *d_inner_src = nullptr;
  unsigned short *d_inner_mask = nullptr;  // bit t set: parity type t flips sign
  // per pair of points (2q, 2q+1): the 16-byte-aligned pair that holds the sources of both
  // points' right-hand sides (the pair itself away from inner ghost points) and a code: bit 0 /
  // bit 1 set when point 0 / point 1 takes that pair's second value, bits 4-15 / 16-27 the parity
  // masks of points 0 / 1.  Null when some pair's sources are not in one aligned pair.
  int *d_inner_pairs = nullptr;
  // the MoL step takes the inner boundary condition of k inside the RK pass (environment
  // BHAH_HIP_SEPARATE_INNER_BC=1 at patch creation keeps the separate copy, for comparison)
  bool fold_inner = true;
  // inner points grouped by source: d_inner_by_src[d_inner_src_start[u] .. [u+1]) lists the
  // points (indices into the inner arrays) copied from source d_inner_src_list[u]
  int *d_inner_by_src = nullptr, *d_inner_src_start = nullptr, *d_inner_src_list = nullptr;
  int n_inner_src = 0;
  int16_t *d_outer[8][3] = {};
  int *d_rad_lines = nullptr;  // (i1, i2) per radial line when the outer points form lines
  int n_rad_lines = 0;
  double *d_wavespeed = nullptr, *d_finf = nullptr;
  double *d_red = nullptr;  // volume-integral partial sums (VI_MAX_BLOCKS x VI_SUMS) and totals
  int *d_gf_parity = nullptr;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // selected schedules (indices into k->rhs / k->ricci) and interior tile shape
  int rhs_choice = 0, ricci_choice = 0;
  // the selected schedules stage through 16-byte global-to-LDS loads (input must be 16-byte aligned)
  bool rhs_dma = false, ricci_dma = false;
  int tile_x = 64, tile_y = 4;
  // adjoint workspace (allocated on the first MoL_step_adjoint call)
  double *d_adj = nullptr;
  // cached graph of one reverse step (bhah_hip_MoL_step_adjoint) and the arguments it was captured with
  hipGraphExec_t adj_exec = nullptr;
  const double *adj_y = nullptr;
  double *adj_lam = nullptr;
  double adj_dt = 0;
  // cached graph for MoL_step_n
  hipGraphExec_t graph_exec = nullptr;
  bhah_MoL_gridfunctions graph_gf{};
  double graph_dt = 0;

  bhah_kparams kp(double dt_, int substep) const {
    bhah_kparams p;
    p.Nx0 = Nx[0];
    p.Nx1 = Nx[1];
    p.Nx2 = Nx[2];
    p.NG = NG;
    p.Ntot = Ntot;
    p.s1 = Nx[0];
    p.s2 = (long)Nx[0] * Nx[1];
    p.invdxx0 = invdxx[0];
    p.invdxx1 = invdxx[1];
    p.invdxx2 = invdxx[2];
    p.eta = d.eta;
    p.KO_gauge = d.KO_strength_gauge;
    p.KO_nongauge = d.KO_strength_nongauge;
    p.wavespeed = d.wavespeed;
    p.dt = dt_;
    p.substep = substep;
    return p;
  }
  // Interior launch: tile_x x tile_y (= 256) threads over (i0, i1), one block row per i2.
  void interior_dims(dim3 &g, dim3 &b) const {
    b = dim3(tile_x, tile_y, 1);
    g = dim3((Nx[0] - 2 * NG + tile_x - 1) / tile_x, (Nx[1] - 2 * NG + tile_y - 1) / tile_y, Nx[2] - 2 * NG);
  }
  void invalidate_graph() {
    if (graph_exec) {
      (void)hipGraphExecDestroy(graph_exec);
      graph_exec = nullptr;
    }
    if (adj_exec) {
      (void)hipGraphExecDestroy(adj_exec);
      adj_exec = nullptr;
    }
  }
  void all_dims(dim3 &g, dim3 &b) const {
    b = dim3(64, 4, 1);
    g = dim3((Nx[0] + 63) / 64, (Nx[1] + 3) / 4, Nx[2]);
  }
};

// ---------------------------------------------------------------------------
// Boundary-condition kernels.

// Sommerfeld outgoing-radiation RHS at pure outer points of one (layer, direction)
// list, spherical-like coordinates (CurviBoundaryConditions/tests/
// apply_bcs_outerradiation_and_inner__rfm__openmp__HoleySinhSpherical.c:42-198).
// One thread per (point, gridfunction).
struct RadStencil {
  int n[7];
  int off[7][7];
  double w[7][7];
};

__device__ __forceinline__ void r_and_dx0dr(int coord, double AMPL, double SINHW, double xx0, double &r, double &dx0dr) {
#pragma clang fp contract(off)
  if (coord == BHAH_SPHERICAL) {
    r = xx0;
    dx0dr = 1.0;
  } else {
    const double tmp0 = (1.0 / (SINHW));
    const double tmp1 = exp(tmp0) - exp(-tmp0);
    const double tmp3 = exp(tmp0 * xx0);
    const double tmp4 = exp(-tmp0 * xx0);
    r = AMPL * (tmp3 - tmp4) / tmp1;
    dx0dr = tmp1 / (AMPL * (tmp0 * tmp3 + tmp0 * tmp4));
  }
}

__device__ __forceinline__ double partial_r_f(const RadStencil &st, int R, const double *__restrict__ f, int Nx0,
                                              long idx, int i0, int FACE0, double dx0dr, double invdxx0) {
#pragma clang fp contract(off)
  int off = FACE0;
  if (i0 < R)
    off = R - i0;
  else if (i0 > Nx0 - R - 1)
    off = (Nx0 - R - 1) - i0;
  const int k = off + R;
  double acc = 0.0;
  for (int j = 0; j < st.n[k]; j++) acc += st.w[k][j] * f[idx + st.off[k][j]];
  return dx0dr * (acc * invdxx0);
}

__device__ __forceinline__ void rad_point(const int16_t *q, int gf, const double *__restrict__ gfs,
                                          double *__restrict__ rhs, const double *__restrict__ xx0, double c,
                                          double f_inf, const RadStencil &st, int R, int coord, double AMPL,
                                          double SINHW, int Nx0, long s1, long s2, long Ntot, double invdxx0) {
#pragma clang fp contract(off)
  const int i0 = q[0], i1 = q[1], i2 = q[2];
  const int j0 = i0 + q[3], j1 = i1 + q[4], j2 = i2 + q[5];
  double r, dxr, r_int, dxr_int;
  r_and_dx0dr(coord, AMPL, SINHW, xx0[i0], r, dxr);
  r_and_dx0dr(coord, AMPL, SINHW, xx0[j0], r_int, dxr_int);
  const long idx = i0 + s1 * i1 + s2 * i2;
  const long idx_int = j0 + s1 * j1 + s2 * j2;
  const double *f = &gfs[gf * Ntot];
  const double dfr = partial_r_f(st, R, f, Nx0, idx, i0, q[3], dxr, invdxx0);
  const double dfr_int = partial_r_f(st, R, f, Nx0, idx_int, j0, q[3], dxr_int, invdxx0);
  const double dtf_int = rhs[gf * Ntot + idx_int];
  const double dtf_int_out = -c * (dfr_int + (f[idx_int] - f_inf) / r_int);
  const double k = r_int * r_int * r_int * (dtf_int - dtf_int_out);
  const double rinv = 1.0 / r;
  const double dtf_out = -c * (dfr + (f[idx] - f_inf) * rinv);
  rhs[gf * Ntot + idx] = dtf_out + k * rinv * rinv * rinv;
}

__global__ void __launch_bounds__(256)
bc_radiation_kernel(const int16_t *__restrict__ pts, int npts, int ngfs, const double *__restrict__ gfs,
                    double *__restrict__ rhs, const double *__restrict__ xx0, const double *__restrict__ wavespeed,
                    const double *__restrict__ finf, RadStencil st, int R, int coord, double AMPL, double SINHW,
                    int Nx0, long s1, long s2, long Ntot, double invdxx0) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)npts * ngfs) return;
  const int p = (int)(t % npts), gf = (int)(t / npts);
  rad_point(&pts[6 * p], gf, gfs, rhs, xx0, wavespeed[gf], finf[gf], st, R, coord, AMPL, SINHW, Nx0, s1, s2, Ntot,
            invdxx0);
}

// All ghost layers of the radiation condition in one launch, for grids whose pure outer
// points are radial lines (i1, i2 fixed, i0 = Nx0-NG .. Nx0-1; spherical-like coordinates):
// one thread per (line, gridfunction) walks the layers innermost first; the right-hand side
// at a layer's interior neighbour is the value the thread formed one layer earlier (the
// first layer's is loaded), kept in a register.  Same arithmetic as rad_point.
__global__ void __launch_bounds__(256)
bc_radiation_lines_kernel(const int *__restrict__ lines, int nlines, int ngfs, int NG, const double *__restrict__ gfs,
                          double *__restrict__ rhs, const double *__restrict__ xx0,
                          const double *__restrict__ wavespeed, const double *__restrict__ finf, RadStencil st,
                          int R, int coord, double AMPL, double SINHW, int Nx0, long s1, long s2, long Ntot,
                          double invdxx0) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nlines * ngfs) return;
  const int l = (int)(t % nlines), gf = (int)(t / nlines);
  const long base = s1 * lines[2 * l] + s2 * lines[2 * l + 1];
  const double *f = &gfs[gf * Ntot];
  double *rr = &rhs[gf * Ntot];
  const double c = wavespeed[gf], f_inf = finf[gf];
  double dtf_int = rr[base + Nx0 - NG - 1];
  double r_int, dxr_int;
  r_and_dx0dr(coord, AMPL, SINHW, xx0[Nx0 - NG - 1], r_int, dxr_int);
  for (int gz = 0; gz < NG; gz++) {
    const int i0 = Nx0 - NG + gz, j0 = i0 - 1;
    const long idx = base + i0, idx_int = base + j0;
    double r, dxr;
    r_and_dx0dr(coord, AMPL, SINHW, xx0[i0], r, dxr);
    const double dfr = partial_r_f(st, R, f, Nx0, idx, i0, -1, dxr, invdxx0);
    const double dfr_int = partial_r_f(st, R, f, Nx0, idx_int, j0, -1, dxr_int, invdxx0);
    const double dtf_int_out = -c * (dfr_int + (f[idx_int] - f_inf) / r_int);
    const double k = r_int * r_int * r_int * (dtf_int - dtf_int_out);
    const double rinv = 1.0 / r;
    const double dtf_out = -c * (dfr + (f[idx] - f_inf) * rinv);
    const double v = dtf_out + k * rinv * rinv * rinv;
    rr[idx] = v;
    dtf_int = v;
    r_int = r;
    dxr_int = dxr;
  }
}

// The same walk with each line's last RAD_SEG values of the gridfunction staged in LDS by
// coalesced loads (lanes of a wave read consecutive values of four lines), instead of every
// lane reading its own line with a separate cache line per load; the layers' results are
// written back the same way.  Needs Nx0 >= RAD_SEG and every stencil point of the walk
// within the segment (NG + R + 2 <= RAD_SEG, checked by the host).  Same arithmetic as
// bc_radiation_lines_kernel.
constexpr int RAD_SEG = 16, RAD_PITCH = RAD_SEG + 1;
constexpr long RAD_STAGED_MIN_ITEMS = 8192;

__device__ __forceinline__ double partial_r_seg(const RadStencil &st, int R, const double *seg, int Nx0, int i0,
                                                double dx0dr, double invdxx0) {
#pragma clang fp contract(off)
  // one-sided towards the interior at the outer face (FACE0 = -1), as partial_r_f
  const int off = (i0 > Nx0 - R - 1) ? (Nx0 - R - 1) - i0 : -1;
  const int k = off + R;
  const int c = i0 - (Nx0 - RAD_SEG);
  double acc = 0.0;
  for (int j = 0; j < st.n[k]; j++) acc += st.w[k][j] * seg[c + st.off[k][j]];
  return dx0dr * (acc * invdxx0);
}

__global__ void __launch_bounds__(256)
bc_radiation_lines_lds_kernel(const int *__restrict__ lines, int nlines, int ngfs, int NG,
                              const double *__restrict__ gfs, double *__restrict__ rhs, const double *__restrict__ xx0,
                              const double *__restrict__ wavespeed, const double *__restrict__ finf, RadStencil st,
                              int R, int coord, double AMPL, double SINHW, int Nx0, long s1, long s2, long Ntot,
                              double invdxx0) {
#pragma clang fp contract(off)
  __shared__ double seg[256 * RAD_PITCH];
  __shared__ long row[256];  // gf * Ntot + line base of each item of the block
  const int tid = (int)threadIdx.x;
  const long nitems = (long)nlines * ngfs, t0 = (long)blockIdx.x * 256, t = t0 + tid;
  const bool active = t < nitems;
  if (active) {
    const int l = (int)(t % nlines), gf = (int)(t / nlines);
    row[tid] = gf * Ntot + s1 * lines[2 * l] + s2 * lines[2 * l + 1];
  }
  __syncthreads();
  for (int j = 0; j < RAD_SEG; j++) {
    const int e = tid + 256 * j, it = e / RAD_SEG, x = e % RAD_SEG;
    if (t0 + it < nitems) seg[it * RAD_PITCH + x] = gfs[row[it] + (Nx0 - RAD_SEG) + x];
  }
  __syncthreads();
  if (active) {
    const double *f = &seg[tid * RAD_PITCH];
    const long base = row[tid];
    const int gf = (int)(t / nlines);
    const double c = wavespeed[gf], f_inf = finf[gf];
    double dtf_int = rhs[base + Nx0 - NG - 1];
    double r_int, dxr_int;
    r_and_dx0dr(coord, AMPL, SINHW, xx0[Nx0 - NG - 1], r_int, dxr_int);
    double v[8];
    for (int gz = 0; gz < NG; gz++) {
      const int i0 = Nx0 - NG + gz, j0 = i0 - 1;
      double r, dxr;
      r_and_dx0dr(coord, AMPL, SINHW, xx0[i0], r, dxr);
      const double dfr = partial_r_seg(st, R, f, Nx0, i0, dxr, invdxx0);
      const double dfr_int = partial_r_seg(st, R, f, Nx0, j0, dxr_int, invdxx0);
      const double dtf_int_out = -c * (dfr_int + (f[j0 - (Nx0 - RAD_SEG)] - f_inf) / r_int);
      const double k = r_int * r_int * r_int * (dtf_int - dtf_int_out);
      const double rinv = 1.0 / r;
      const double dtf_out = -c * (dfr + (f[i0 - (Nx0 - RAD_SEG)] - f_inf) * rinv);
      v[gz] = dtf_out + k * rinv * rinv * rinv;
      dtf_int = v[gz];
      r_int = r;
      dxr_int = dxr;
    }
    // this thread alone read its segment: its results take the segment's first NG slots
    for (int gz = 0; gz < NG; gz++) seg[tid * RAD_PITCH + gz] = v[gz];
  }
  __syncthreads();
  for (int j = 0; j < NG; j++) {
    const int e = tid + 256 * j, it = e / NG, gz = e % NG;
    if (t0 + it < nitems) rhs[row[it] + (Nx0 - NG) + gz] = seg[it * RAD_PITCH + gz];
  }
}

constexpr int BC_INNER_BATCH = 12;  // gridfunctions in flight per thread in the inner-parity copy

// Batched forms over an ensemble: one record per (member, boundary point), one
// thread per (record, gridfunction); member arrays are picked by index.
struct ens_outer_pt {
  int m;
  int16_t q[6];  // i0, i1, i2, face normal
};
struct ens_inner_pt {
  int m, dst, src;
  unsigned mask;
};

__global__ void __launch_bounds__(256)
ens_bc_radiation_kernel(const bhah_member *__restrict__ mem, const ens_outer_pt *__restrict__ pts, int npts,
                        int ngfs, int in_sel, int rhs_sel, const double *__restrict__ wavespeed,
                        const double *__restrict__ finf, RadStencil st, int R) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)npts * ngfs) return;
  const int p = (int)(t % npts), gf = (int)(t / npts);
  const ens_outer_pt e = pts[p];
  const bhah_member *M = &mem[e.m];
  rad_point(e.q, gf, M->arr[in_sel], M->arr[rhs_sel], M->xx0, wavespeed[gf], finf[gf], st, R, M->coord, M->AMPL,
            M->SINHW, M->kp.Nx0, M->kp.s1, M->kp.s2, M->kp.Ntot, M->kp.invdxx0);
}

__global__ void __launch_bounds__(256)
ens_bc_inner_kernel(const bhah_member *__restrict__ mem, const ens_inner_pt *__restrict__ pts, int npts, int ngfs,
                    const int *__restrict__ gf_parity, int sel) {
  // one thread per boundary point, all gridfunctions: the record is read once
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= npts) return;
  const ens_inner_pt e = pts[p];
  const bhah_member *M = &mem[e.m];
  double *f = M->arr[sel];
  const long Ntot = M->kp.Ntot;
  for (int g0 = 0; g0 < ngfs; g0 += BC_INNER_BATCH) {  // loads of a batch before its stores
    double v[BC_INNER_BATCH];
#pragma unroll
    for (int j = 0; j < BC_INNER_BATCH; j++)
      if (g0 + j < ngfs) v[j] = f[(g0 + j) * Ntot + e.src];
#pragma unroll
    for (int j = 0; j < BC_INNER_BATCH; j++)
      if (g0 + j < ngfs) f[(g0 + j) * Ntot + e.dst] = ((e.mask >> gf_parity[g0 + j]) & 1u) ? -v[j] : v[j];
  }
}

// Quadratic extrapolation f0 = 3 f1 - 3 f2 + f3 along the face normal
// (…/apply_bcs_outerextrap_and_inner__openmp.c:8-34).
__global__ void __launch_bounds__(256)
bc_extrap_kernel(const int16_t *__restrict__ pts, int npts, int ngfs, double *__restrict__ gfs, long s1, long s2,
                 long Ntot) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)npts * ngfs) return;
  const int p = (int)(t % npts), gf = (int)(t / npts);
  const int16_t *q = &pts[6 * p];
  const long i = q[0] + s1 * q[1] + s2 * q[2];
  const long s = q[3] + s1 * q[4] + s2 * q[5];
  double *f = &gfs[gf * Ntot];
  f[i] = +3.0 * f[i + s] - 3.0 * f[i + 2 * s] + 1.0 * f[i + 3 * s];
}

// gf[dst] = +-gf[src] by a sign-bit flip (…/apply_bcs_inner_only__openmp.c:3-57).
__global__ void __launch_bounds__(256)
bc_inner_kernel(const int *__restrict__ dst, const int *__restrict__ src, const unsigned short *__restrict__ mask,
                int npts, int ngfs, const int *__restrict__ gf_parity, double *__restrict__ gfs, long Ntot) {
  // one thread per boundary point, all gridfunctions: the point's record is read once
  const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (p >= npts) return;
  const long d = dst[p], sr = src[p];
  const unsigned m = mask[p];
  // Sources are interior points, never another boundary point's destination, so the loads
  // of a batch of gridfunctions may all be issued before its stores (the compiler cannot
  // know that the stores do not alias later loads of the same array).
  for (int g0 = 0; g0 < ngfs; g0 += BC_INNER_BATCH) {
    double v[BC_INNER_BATCH];
#pragma unroll
    for (int j = 0; j < BC_INNER_BATCH; j++)
      if (g0 + j < ngfs) v[j] = gfs[(g0 + j) * Ntot + sr];
#pragma unroll
    for (int j = 0; j < BC_INNER_BATCH; j++)
      if (g0 + j < ngfs) gfs[(g0 + j) * Ntot + d] = ((m >> gf_parity[g0 + j]) & 1u) ? -v[j] : v[j];
  }
}

// ---- Adjoints of the boundary conditions (reverse mode; linear maps).

// Weights and flat indices of the radial derivative at one point, as partial_r_f forms it.
__device__ __forceinline__ int partial_r_taps(const RadStencil &st, int R, int Nx0, long idx, int i0, int FACE0,
                                             long *at, double *w) {
  int off = FACE0;
  if (i0 < R)
    off = R - i0;
  else if (i0 > Nx0 - R - 1)
    off = (Nx0 - R - 1) - i0;
  const int k = off + R;
  for (int j = 0; j < st.n[k]; j++) {
    at[j] = idx + st.off[k][j];
    w[j] = st.w[k][j];
  }
  return st.n[k];
}

// Reverse of one radiation-BC layer: the forward set rhs[p] from rhs[p_int] and the
// stage input f around p and p_int; push rhs[p]'s adjoint to rhs[p_int] (distinct per
// point of a layer) and, atomically, to f, then clear it.
__global__ void __launch_bounds__(256)
bc_radiation_adjoint_kernel(const int16_t *__restrict__ pts, int npts, int ngfs, double *__restrict__ lam_rhs,
                            double *__restrict__ lam_f, const double *__restrict__ xx0,
                            const double *__restrict__ wavespeed, RadStencil st, int R, int coord, double AMPL,
                            double SINHW, int Nx0, long s1, long s2, long Ntot, double invdxx0) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)npts * ngfs) return;
  const int p = (int)(t % npts), gf = (int)(t / npts);
  const int16_t *q = &pts[6 * p];
  const int i0 = q[0], i1 = q[1], i2 = q[2];
  const int j0 = i0 + q[3], j1 = i1 + q[4], j2 = i2 + q[5];
  const long idx = i0 + s1 * i1 + s2 * i2;
  const long idx_int = j0 + s1 * j1 + s2 * j2;
  double *lr = &lam_rhs[gf * Ntot];
  double *lf = &lam_f[gf * Ntot];
  const double lam = lr[idx];
  lr[idx] = 0.0;
  if (lam == 0.0) return;
  double r, dxr, r_int, dxr_int;
  r_and_dx0dr(coord, AMPL, SINHW, xx0[i0], r, dxr);
  r_and_dx0dr(coord, AMPL, SINHW, xx0[j0], r_int, dxr_int);
  const double c = wavespeed[gf];
  const double ratio = (r_int * r_int * r_int) / (r * r * r);
  lr[idx_int] += ratio * lam;
  long at[8];
  double w[8];
  int n = partial_r_taps(st, R, Nx0, idx, i0, q[3], at, w);
  for (int j = 0; j < n; j++) atomicAdd(&lf[at[j]], -c * dxr * invdxx0 * w[j] * lam);
  atomicAdd(&lf[idx], -c / r * lam);
  n = partial_r_taps(st, R, Nx0, idx_int, j0, q[3], at, w);
  for (int j = 0; j < n; j++) atomicAdd(&lf[at[j]], ratio * c * dxr_int * invdxx0 * w[j] * lam);
  atomicAdd(&lf[idx_int], ratio * c / r_int * lam);
}

// The same reverse for outer points that form radial lines (the layout of
// bc_radiation_lines_kernel): one thread per (line, gridfunction) walks the layers from the
// outermost in, with the operations of bc_radiation_adjoint_kernel in the same order.  Only
// this thread touches its line's values, so the additions need no atomics and the result is
// the per-layer kernels' bit for bit.
__global__ void __launch_bounds__(256)
bc_radiation_adjoint_lines_kernel(const int *__restrict__ lines, int nlines, int ngfs, int NG,
                                  double *__restrict__ lam_rhs, double *__restrict__ lam_f,
                                  const double *__restrict__ xx0, const double *__restrict__ wavespeed,
                                  RadStencil st, int R, int coord, double AMPL, double SINHW, int Nx0, long s1,
                                  long s2, long Ntot, double invdxx0) {
#pragma clang fp contract(off)
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nlines * ngfs) return;
  const int l = (int)(t % nlines), gf = (int)(t / nlines);
  const long base = s1 * lines[2 * l] + s2 * lines[2 * l + 1];
  double *lr = &lam_rhs[gf * Ntot];
  double *lf = &lam_f[gf * Ntot];
  const double c = wavespeed[gf];
  for (int gz = NG - 1; gz >= 0; gz--) {
    const int i0 = Nx0 - NG + gz, j0 = i0 - 1;
    const long idx = base + i0, idx_int = base + j0;
    const double lam = lr[idx];
    lr[idx] = 0.0;
    if (lam == 0.0) continue;
    double r, dxr, r_int, dxr_int;
    r_and_dx0dr(coord, AMPL, SINHW, xx0[i0], r, dxr);
    r_and_dx0dr(coord, AMPL, SINHW, xx0[j0], r_int, dxr_int);
    const double ratio = (r_int * r_int * r_int) / (r * r * r);
    lr[idx_int] += ratio * lam;
    long at[8];
    double w[8];
    int n = partial_r_taps(st, R, Nx0, idx, i0, -1, at, w);
    for (int j = 0; j < n; j++) lf[at[j]] += -c * dxr * invdxx0 * w[j] * lam;
    lf[idx] += -c / r * lam;
    n = partial_r_taps(st, R, Nx0, idx_int, j0, -1, at, w);
    for (int j = 0; j < n; j++) lf[at[j]] += ratio * c * dxr_int * invdxx0 * w[j] * lam;
    lf[idx_int] += ratio * c / r_int * lam;
  }
}

// Reverse of gf[dst] = +-gf[src]: lam[src] += sum of +-lam[dst] over the points copied from
// src, lam[dst] = 0.  One thread per (source, gridfunction), summing its points in a fixed
// order, so the result does not depend on scheduling.  Sources are never destinations, and
// every destination belongs to one source, so no two threads touch the same value.
__global__ void __launch_bounds__(256)
bc_inner_adjoint_kernel(const int *__restrict__ dst, const int *__restrict__ by_src,
                        const int *__restrict__ src_start, const int *__restrict__ src_list, int nsrc,
                        const unsigned short *__restrict__ mask, int ngfs, const int *__restrict__ gf_parity,
                        double *__restrict__ lam, long Ntot) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nsrc * ngfs) return;
  const int u = (int)(t % nsrc), gf = (int)(t / nsrc);
  double *l = &lam[gf * Ntot];
  const int par = gf_parity[gf];
  double acc = 0.0;
  for (int j = src_start[u]; j < src_start[u + 1]; j++) {
    const int p = by_src[j];
    const double v = l[dst[p]];
    l[dst[p]] = 0.0;
    acc += ((mask[p] >> par) & 1u) ? -v : v;
  }
  l[src_list[u]] += acc;
}

// dst[0..n) = v
__global__ void __launch_bounds__(256)
fill_kernel(double *__restrict__ dst, double v, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = v;
}

// out = a + b
__global__ void __launch_bounds__(256)
add_kernel(double *__restrict__ out, const double *__restrict__ a, const double *__restrict__ b, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = a[t] + b[t];
}

// ---- Checkpoint payload (write_checkpoint__openmp.c:95-117): point-major copy of y_n.
__global__ void __launch_bounds__(256)
checkpoint_pack_kernel(const double *__restrict__ y, double *__restrict__ compact, long ntot, int ngf) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntot * ngf) return;
  const long p = t / ngf;
  const int g = (int)(t - p * ngf);
  compact[t] = y[g * ntot + p];
}
__global__ void __launch_bounds__(256)
checkpoint_unpack_kernel(const int *__restrict__ idx, const double *__restrict__ compact, long count, int ngf,
                         long ntot, double *__restrict__ y) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count * ngf) return;
  const long e = t / ngf;
  const int g = (int)(t - e * ngf);
  y[g * ntot + idx[e]] = compact[t];
}

// ---- initial_data_reader__convert_ADM_Cartesian_to_BSSN on the device
// (ADM_Initial_Data_Reader__BSSN_Converter.py:268-460, math ADM_to_BSSN.py:80-160): ADM data in
// the Cartesian basis at every point (gammaDD00,01,02,11,12,22, KDD.., alpha, betaU, BU) ->
// BSSN in the Cartesian basis (det ghat = 1 there) -> the patch's spherical-like basis with
// the Jacobian dx_Cart/dxx of the map -> rescaled variables; lambda^i = 0 here (it is formed
// by finite differences afterwards).
__global__ void __launch_bounds__(256)
adm_cartesian_to_bssn_kernel(const double *__restrict__ adm, double *__restrict__ y, const double *__restrict__ rfm0,
                             const double *__restrict__ xx1, const double *__restrict__ xx2, int Nx0, int Nx1,
                             long Ntot) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= Ntot) return;
  const int i0 = (int)(p % Nx0), i1 = (int)((p / Nx0) % Nx1), i2 = (int)(p / ((long)Nx0 * Nx1));
  const int sym[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  double g[3][3], K[3][3];
  for (int k = 0; k < 6; k++) {
    const int i = sym[k][0], j = sym[k][1];
    g[i][j] = g[j][i] = adm[k * Ntot + p];
    K[i][j] = K[j][i] = adm[(6 + k) * Ntot + p];
  }
  auto det3 = [](const double m[3][3]) {
    return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
           m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  };
  const double detg = det3(g);
  double gUU[3][3];
  gUU[0][0] = (g[1][1] * g[2][2] - g[1][2] * g[2][1]) / detg;
  gUU[0][1] = gUU[1][0] = (g[0][2] * g[2][1] - g[0][1] * g[2][2]) / detg;
  gUU[0][2] = gUU[2][0] = (g[0][1] * g[1][2] - g[0][2] * g[1][1]) / detg;
  gUU[1][1] = (g[0][0] * g[2][2] - g[0][2] * g[2][0]) / detg;
  gUU[1][2] = gUU[2][1] = (g[0][2] * g[1][0] - g[0][0] * g[1][2]) / detg;
  gUU[2][2] = (g[0][0] * g[1][1] - g[0][1] * g[1][0]) / detg;
  // gammabar_ij = (det ghat / det gamma)^(1/3) gamma_ij, W = (det gamma / det gammabar)^(-1/6),
  // K = gamma^ij K_ij, Abar_ij = (det ghat / det gamma)^(1/3) (K_ij - gamma_ij K / 3)
  const double fac = cbrt(1.0 / detg);
  double gb[3][3], Ab[3][3];
  double trK = 0.0;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) trK += gUU[i][j] * K[i][j];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      gb[i][j] = fac * g[i][j];
      Ab[i][j] = fac * (K[i][j] - (1.0 / 3.0) * g[i][j] * trK);
    }
  const double cf = pow(detg / det3(gb), -1.0 / 6.0);
  // Jacobian of x = f0(xx0) (sin xx1 cos xx2, sin xx1 sin xx2, cos xx1); its columns are
  // orthogonal with lengths s = (f0', f0, f0 sin xx1) = 1 / ReU
  const double f0 = rfm0[i0], f0p = rfm0[Nx0 + i0];
  const double st = sin(xx1[i1]), ct = cos(xx1[i1]), sp = sin(xx2[i2]), cp = cos(xx2[i2]);
  const double J[3][3] = {{f0p * st * cp, f0 * ct * cp, -f0 * st * sp},
                          {f0p * st * sp, f0 * ct * sp, f0 * st * cp},
                          {f0p * ct, -f0 * st, 0.0}};
  const double s[3] = {f0p, f0, f0 * st};
  double gbr[3][3], Abr[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double a = 0.0, b = 0.0;
      for (int m = 0; m < 3; m++)
        for (int n = 0; n < 3; n++) {
          a += J[m][i] * gb[m][n] * J[n][j];
          b += J[m][i] * Ab[m][n] * J[n][j];
        }
      gbr[i][j] = a;
      Abr[i][j] = b;
    }
  // evolved order: aDD00..22 (0-5), alpha 6, betU 7-9, cf 10, hDD 11-16, lambdaU 17-19, trK 20, vetU 21-23
  for (int k = 0; k < 6; k++) {
    const int i = sym[k][0], j = sym[k][1];
    const double ReDD = s[i] * s[j];
    y[(11 + k) * Ntot + p] = (gbr[i][j] - (i == j ? s[i] * s[i] : 0.0)) / ReDD;
    y[k * Ntot + p] = Abr[i][j] / ReDD;
  }
  for (int i = 0; i < 3; i++) {
    // beta_rfm^i = (J^-1)^i_a beta^a = J[a][i] beta^a / s_i^2; vet^i = beta_rfm^i / ReU^i
    double vb = 0.0, vB = 0.0;
    for (int a = 0; a < 3; a++) {
      vb += J[a][i] * adm[(13 + a) * Ntot + p];
      vB += J[a][i] * adm[(16 + a) * Ntot + p];
    }
    y[(21 + i) * Ntot + p] = vb / s[i];
    y[(7 + i) * Ntot + p] = vB / s[i];
    y[(17 + i) * Ntot + p] = 0.0;
  }
  y[6 * Ntot + p] = adm[12 * Ntot + p];
  y[10 * Ntot + p] = cf;
  y[20 * Ntot + p] = trK;
}

// ---- Brill-Lindquist initial data on every point (ghost zones included): conformal
// factor W = psi^-2 and lapse psi^-2, all other BSSN variables zero
// (InitialData_Cartesian BrillLindquist; conformally flat, so gammabar = gammahat and the
// rescaled h_ij, lambda^i vanish).
__global__ void __launch_bounds__(256)
initial_data_bl_kernel(double *__restrict__ y, int coord, double AMPL, double SINHW, int Nx0, int Nx1, int Nx2, int NG,
                       double xmin0, double xmin1, double xmin2, double dx0, double dx1, double dx2, double m1,
                       double x1, double y1, double z1, double m2, double x2, double y2, double z2, long ntot) {
#pragma clang fp contract(off)
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const int i0 = (int)(p % Nx0), i1 = (int)((p / Nx0) % Nx1), i2 = (int)(p / ((long)Nx0 * Nx1));
  const double q0 = xmin0 + ((double)(i0 - NG) + (1.0 / 2.0)) * dx0;
  const double q1 = xmin1 + ((double)(i1 - NG) + (1.0 / 2.0)) * dx1;
  const double q2 = xmin2 + ((double)(i2 - NG) + (1.0 / 2.0)) * dx2;
  double X, Y, Z;
  if (coord == BHAH_CARTESIAN) {
    X = q0;
    Y = q1;
    Z = q2;
  } else {
    double r = q0;
    if (coord == BHAH_SINHSPHERICAL)  // reference_metric.py:1224-1235, exponential form
      r = AMPL * (exp(q0 / SINHW) - exp(-q0 / SINHW)) / (exp(1.0 / SINHW) - exp(-1.0 / SINHW));
    X = r * sin(q1) * cos(q2);
    Y = r * sin(q1) * sin(q2);
    Z = r * cos(q1);
  }
  const double r1 = sqrt((X - x1) * (X - x1) + (Y - y1) * (Y - y1) + (Z - z1) * (Z - z1));
  const double r2 = sqrt((X - x2) * (X - x2) + (Y - y2) * (Y - y2) + (Z - z2) * (Z - z2));
  const double psi = 1.0 + m1 / (2.0 * r1) + m2 / (2.0 * r2);
  const double w = 1.0 / (psi * psi);
  for (int g = 0; g < BHAH_NUM_EVOL_GFS_BSSN; g++) y[g * ntot + p] = (g == 6 || g == 10) ? w : 0.0;
}

// ---- Volume integrals of the diagnostics recipes (whole domain; |x| > 8), reduced in a
// fixed order so that every call gives bitwise the same sums: each workgroup of the first
// pass walks interior points in array order with a grid stride and writes its eight partial
// sums (per recipe: volume, integral of H^2, of MSQUARED^2, of LAMBDA_CONSTRAINT^2); the
// second pass adds the partials of all workgroups in workgroup order.
constexpr int VI_SUMS = 4 * BHAH_DIAG_NUM_RECIPES;
constexpr int VI_MAX_BLOCKS = 1024;
constexpr double VI_SPHERE_RADIUS = 8.0;  // diagnostics_volume_integration.py:193

__global__ void __launch_bounds__(256)
volume_integrals_partial_kernel(const double *__restrict__ diag, const double *__restrict__ rfm0,
                                const double *__restrict__ rfm1, const double *__restrict__ xx1,
                                const double *__restrict__ xx2, bhah_kparams p, double dxx012,
                                double *__restrict__ partial) {
  __shared__ double sh[VI_SUMS][256];
  const int n0 = p.Nx0 - 2 * p.NG, n1 = p.Nx1 - 2 * p.NG, n2 = p.Nx2 - 2 * p.NG;
  const long nint = (long)n0 * n1 * n2;
  double acc[VI_SUMS];
#pragma unroll
  for (int k = 0; k < VI_SUMS; k++) acc[k] = 0.0;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < nint; q += (long)gridDim.x * blockDim.x) {
    const int i0 = p.NG + (int)(q % n0), i1 = p.NG + (int)((q / n0) % n1), i2 = p.NG + (int)(q / ((long)n0 * n1));
    const long idx = i0 + p.s1 * i1 + p.s2 * i2;
    // sqrt(det gammahat) |dxx0 dxx1 dxx2| of spherical-like coordinates: |f0' f0^2 sin(xx1)| d^3xx
    const double f0 = rfm0[i0], f0p = rfm0[p.Nx0 + i0];
    const double dV = fabs(f0p * f0 * f0 * rfm1[i1]) * dxx012;
    if (dV == 0.0) continue;
    const double H = diag[idx], M2 = diag[BHAH_DIAG_MSQUARED * p.Ntot + idx],
                 L = diag[BHAH_DIAG_LAMBDA_CONSTRAINT * p.Ntot + idx];
    const double v[4] = {dV, H * H * dV, M2 * M2 * dV, L * L * dV};
#pragma unroll
    for (int k = 0; k < 4; k++) acc[k] += v[k];
    // recipe 1 keeps points outside the sphere |x| <= 8 (exclude_inside = 1)
    const double X = f0 * sin(xx1[i1]) * cos(xx2[i2]), Y = f0 * sin(xx1[i1]) * sin(xx2[i2]), Z = f0 * cos(xx1[i1]);
    if (!(sqrt(X * X + Y * Y + Z * Z) <= VI_SPHERE_RADIUS)) {
#pragma unroll
      for (int k = 0; k < 4; k++) acc[4 + k] += v[k];
    }
  }
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < VI_SUMS; k++) sh[k][tid] = acc[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) {
#pragma unroll
      for (int k = 0; k < VI_SUMS; k++) sh[k][tid] += sh[k][tid + s];
    }
    __syncthreads();
  }
  if (tid < VI_SUMS) partial[(long)blockIdx.x * VI_SUMS + tid] = sh[tid][0];
}

// Partial sums of bhah_hip_diagnostics_constraints_noalloc, which allocates nothing: one
// array per GPU for the life of the process (64 KiB), shared by every call on that GPU.
// Host code never uses this name as an address; it uses the per-device pointer from
// hipGetSymbolAddress (noalloc_scratch).
constexpr int NOALLOC_SCRATCH_DOUBLES = VI_MAX_BLOCKS * VI_SUMS;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

__global__ void __launch_bounds__(64)
volume_integrals_final_kernel(const double *__restrict__ partial, int nblocks, double *__restrict__ out) {
  const int k = threadIdx.x;
  if (k >= VI_SUMS) return;
  double s = 0.0;
  for (int b = 0; b < nblocks; b++) s += partial[(long)b * VI_SUMS + k];
  out[k] = s;
}

// ---- Device-side set-up (bhah_hip_patch_create_on_device): reference-metric arrays, the
// classification of ghost points into inner and pure outer boundary points, order-preserving
// compaction of both lists, and the CFL reduction, with the host set-up's arithmetic
// (contraction into fused multiply-adds is off, so coordinates and index maps agree bit for
// bit; exp, sin, cos, acos and atan2 come from the device math library).
struct SetupGeom {
  int coord, NG, Nx0, Nx1, Nx2;
  long Ntot;
  double xxmin[3], dxx[3], AMPL, SINHW;
};
struct SetupBox {
  int lo[3], hi[3];  // points lo <= i < hi, enumerated with i0 fastest (the host loops' order)
  int face[3];       // outward face normal recorded with pure outer points
};
constexpr int SETUP_BLOCK = 256;

__device__ __forceinline__ double setup_xx(const SetupGeom &g, int d, int j) {
#pragma clang fp contract(off)
  return g.xxmin[d] + ((double)(j - g.NG) + (1.0 / 2.0)) * g.dxx[d];
}

__global__ void __launch_bounds__(256) setup_rfm_kernel(SetupGeom g, double *__restrict__ rfm0,
                                                        double *__restrict__ rfm1) {
#pragma clang fp contract(off)
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t < g.Nx0) {
    const double x = setup_xx(g, 0, t);
    double f = 1, d1 = 0, d2 = 0, d3 = 0;
    if (g.coord == BHAH_SPHERICAL) {
      f = x;
      d1 = 1.0;
    } else if (g.coord == BHAH_SINHSPHERICAL) {
      const double A = g.AMPL, w = g.SINHW;
      const double den = exp(1.0 / w) - exp(-1.0 / w);
      const double ep = exp(x / w), em = exp(-x / w);
      f = A * (ep - em) / den;
      d1 = A * (ep + em) / (w * den);
      d2 = A * (ep - em) / (w * w * den);
      d3 = A * (ep + em) / (w * w * w * den);
    }
    rfm0[t] = f;
    rfm0[g.Nx0 + t] = d1;
    rfm0[2 * g.Nx0 + t] = d2;
    rfm0[3 * g.Nx0 + t] = d3;
  }
  if (t < g.Nx1) {
    const double y = setup_xx(g, 1, t);
    const bool cart = g.coord == BHAH_CARTESIAN;
    rfm1[t] = cart ? 1.0 : sin(y);
    rfm1[g.Nx1 + t] = cart ? 0.0 : cos(y);
    rfm1[2 * g.Nx1 + t] = cart ? 0.0 : -sin(y);
  }
}

// kind[p]: 0 interior, 1 inner boundary point (src[p], mask[p] set), 2 pure outer point.
// err bit 1: an inner point maps outside the grid; bit 2: a parity dot product is not +-1.
__global__ void __launch_bounds__(256) setup_classify_kernel(SetupGeom g, unsigned char *__restrict__ kind,
                                                             int *__restrict__ src, unsigned short *__restrict__ mask,
                                                             int *__restrict__ err) {
#pragma clang fp contract(off)
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.Ntot) return;
  const int i0 = (int)(p % g.Nx0), i1 = (int)((p / g.Nx0) % g.Nx1), i2 = (int)(p / ((long)g.Nx0 * g.Nx1));
  const int NG = g.NG;
  if (i0 >= NG && i0 < g.Nx0 - NG && i1 >= NG && i1 < g.Nx1 - NG && i2 >= NG && i2 < g.Nx2 - NG) {
    kind[p] = 0;
    return;
  }
  if (g.coord == BHAH_CARTESIAN) {
    kind[p] = 2;
    return;
  }
  // spherical eigen-coordinate double map xx -> Cart -> xx' (as eigen_inbounds on the host)
  const int Nx[3] = {g.Nx0, g.Nx1, g.Nx2};
  const double r = setup_xx(g, 0, i0), th = setup_xx(g, 1, i1), ph = setup_xx(g, 2, i2);
  const double X = r * sin(th) * cos(ph), Y = r * sin(th) * sin(ph), Z = r * cos(th);
  const double rr = sqrt(X * X + Y * Y + Z * Z);
  const double c[3] = {rr, acos(Z / rr), atan2(Y, X)};
  int ib[3];
  for (int d = 0; d < 3; d++) {
    const double xxmin = setup_xx(g, d, 0) - ((double)(0 - NG) + (1.0 / 2.0)) * g.dxx[d];
    ib[d] = (int)((c[d] - xxmin - (1.0 / 2.0) * g.dxx[d] + ((double)NG) * g.dxx[d]) / g.dxx[d] + 0.5);
  }
  if (ib[0] == i0 && ib[1] == i1 && ib[2] == i2) {
    kind[p] = 2;
    return;
  }
  for (int d = 0; d < 3; d++)
    if (ib[d] < 0 || ib[d] >= Nx[d]) {
      atomicOr(err, 1);
      kind[p] = 3;
      return;
    }
  const double xin1 = setup_xx(g, 1, ib[1]), xin2 = setup_xx(g, 2, ib[2]);
  const double t0 = cos(th) * cos(xin1);
  const double t1 = sin(th) * sin(xin1);
  const double t2 = sin(ph) * sin(xin2);
  const double t3 = cos(ph) * cos(xin2);
  const double p1 = t0 + t1 * t2 + t1 * t3;
  const double p2 = t0 * t2 + t0 * t3 + t1;
  const double p3 = t2 + t3;
  const double dots[10] = {1, p1, p2, p3, p1 * p1, p1 * p2, p1 * p3, p2 * p2, p2 * p3, p3 * p3};
  unsigned m = 0;
  for (int k = 0; k < 10; k++) {
    if (fabs(dots[k]) < 1 - 1e-8 || fabs(dots[k]) > 1 + 1e-8) atomicOr(err, 2);
    if (dots[k] < 0) m |= 1u << k;
  }
  kind[p] = 1;
  src[p] = ib[0] + Nx[0] * (ib[1] + Nx[1] * ib[2]);
  mask[p] = (unsigned short)m;
}

__device__ __forceinline__ long setup_box_point(const SetupGeom &g, const SetupBox &b, long q, int &i0, int &i1,
                                                int &i2) {
  const int n0 = b.hi[0] - b.lo[0], n1 = b.hi[1] - b.lo[1];
  i0 = b.lo[0] + (int)(q % n0);
  i1 = b.lo[1] + (int)((q / n0) % n1);
  i2 = b.lo[2] + (int)(q / ((long)n0 * n1));
  return i0 + (long)g.Nx0 * (i1 + (long)g.Nx1 * i2);
}

// Points of kind `want` per block of SETUP_BLOCK consecutive box points.
__global__ void __launch_bounds__(SETUP_BLOCK) setup_count_kernel(SetupGeom g, SetupBox b, long n,
                                                                  const unsigned char *__restrict__ kind, int want,
                                                                  int *__restrict__ counts) {
  __shared__ int wc[SETUP_BLOCK / 64];
  const long q = (long)blockIdx.x * SETUP_BLOCK + threadIdx.x;
  int i0, i1, i2;
  const bool f = q < n && kind[setup_box_point(g, b, q < n ? q : 0, i0, i1, i2)] == want;
  const unsigned long long m = __ballot(f);
  if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int s = 0;
    for (int w = 0; w < SETUP_BLOCK / 64; w++) s += wc[w];
    counts[blockIdx.x] = s;
  }
}

// Write the selected points in box order from each block's offset: inner points as
// (dst, src, mask), pure outer points as (i0, i1, i2, face normal).
__global__ void __launch_bounds__(SETUP_BLOCK) setup_scatter_kernel(
    SetupGeom g, SetupBox b, long n, const unsigned char *__restrict__ kind, int want, const int *__restrict__ offsets,
    const int *__restrict__ src, const unsigned short *__restrict__ mask, int *__restrict__ o_dst,
    int *__restrict__ o_src, unsigned short *__restrict__ o_mask, int16_t *__restrict__ o_outer) {
  __shared__ int wc[SETUP_BLOCK / 64];
  const long q = (long)blockIdx.x * SETUP_BLOCK + threadIdx.x;
  int i0 = 0, i1 = 0, i2 = 0;
  const long p = setup_box_point(g, b, q < n ? q : 0, i0, i1, i2);
  const bool f = q < n && kind[p] == want;
  const unsigned long long m = __ballot(f);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wc[wave] = __popcll(m);
  __syncthreads();
  if (!f) return;
  int pos = offsets[blockIdx.x];
  for (int w = 0; w < wave; w++) pos += wc[w];
  pos += __popcll(m & ((1ull << lane) - 1ull));
  if (want == 1) {
    o_dst[pos] = (int)p;
    o_src[pos] = src[p];
    o_mask[pos] = mask[p];
  } else {
    int16_t *o = &o_outer[6 * (long)pos];
    o[0] = (int16_t)i0;
    o[1] = (int16_t)i1;
    o[2] = (int16_t)i2;
    o[3] = (int16_t)b.face[0];
    o[4] = (int16_t)b.face[1];
    o[5] = (int16_t)b.face[2];
  }
}

// Per-block minimum over all points of min(ds0, ds1, ds2) (the CFL condition).
__global__ void __launch_bounds__(256) setup_cfl_kernel(SetupGeom g, double *__restrict__ partial) {
#pragma clang fp contract(off)
  __shared__ double sh[256];
  double mn = 1e38;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < g.Ntot; p += (long)gridDim.x * blockDim.x) {
    const int i0 = (int)(p % g.Nx0), i1 = (int)((p / g.Nx0) % g.Nx1);
    const double x0 = setup_xx(g, 0, i0), x1 = setup_xx(g, 1, i1);
    double ds0, ds1, ds2;
    if (g.coord == BHAH_SPHERICAL) {
      ds0 = fabs(g.dxx[0]);
      ds1 = fabs(x0 * g.dxx[1]);
      ds2 = fabs(x0 * g.dxx[2] * sin(x1));
    } else {
      const double tmp0 = (1.0 / (g.SINHW));
      const double tmp4 = g.AMPL / (exp(tmp0) - exp(-tmp0));
      const double tmp2 = exp(tmp0 * x0);
      const double tmp3 = exp(-tmp0 * x0);
      const double tmp5 = tmp4 * (tmp2 - tmp3);
      ds0 = fabs(g.dxx[0] * tmp4 * (tmp0 * tmp2 + tmp0 * tmp3));
      ds1 = fabs(g.dxx[1] * tmp5);
      ds2 = fabs(g.dxx[2] * tmp5 * sin(x1));
    }
    const double m = fmin(ds0, fmin(ds1, ds2));
    mn = m < mn ? m : mn;
  }
  sh[threadIdx.x] = mn;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// ---------------------------------------------------------------------------
// Host-side setup.
namespace {

// First-derivative weights on 2R+1 points shifted by o (exact rationals, then rounded once):
// Vandermonde solve, as nrpy/finite_difference.py:93-131 with the radiation BC's
// arbitrary-offset stencils (CurviBoundaryConditions/apply_bcs_outerradiation_and_inner.py:126-160).
void rad_stencil(int R, RadStencil &st) {
  memset(&st, 0, sizeof(st));
  for (int o = -R; o <= R; o++) {
    const int n = 2 * R + 1;
    // solve sum_j w_j x_j^i = delta_{i1} with x_j = j - R + o, by Lagrange basis derivative at 0
    int cnt = 0;
    for (int j = 0; j < n; j++) {
      const int xj = j - R + o;
      // w_j = L_j'(0) = sum_{m != j} [1/(xj - xm)] prod_{l != j, m} (0 - xl)/(xj - xl)
      long double w = 0.0L;
      for (int m = 0; m < n; m++) {
        if (m == j) continue;
        const int xm = m - R + o;
        long double term = 1.0L / (long double)(xj - xm);
        for (int l = 0; l < n; l++) {
          if (l == j || l == m) continue;
          const int xl = l - R + o;
          term *= (long double)(0 - xl) / (long double)(xj - xl);
        }
        w += term;
      }
      if (w != 0.0L) {
        st.off[o + R][cnt] = xj;
        st.w[o + R][cnt] = (double)w;
        cnt++;
      }
    }
    st.n[o + R] = cnt;
  }
}

// Spherical eigen-coordinate double map xx -> Cart -> xx' (bcstruct_set_up…HoleySinhSpherical.c:43-185).
void eigen_inbounds(const bhah_patch &P, int i0, int i1, int i2, int ib[3], double xin[3]) {
  if (P.d.coord == BHAH_CARTESIAN) {
    ib[0] = i0;
    ib[1] = i1;
    ib[2] = i2;
    for (int d = 0; d < 3; d++) xin[d] = P.xx[d][ib[d]];
    return;
  }
  const double r = P.xx[0][i0], th = P.xx[1][i1], ph = P.xx[2][i2];
  const double X = r * sin(th) * cos(ph), Y = r * sin(th) * sin(ph), Z = r * cos(th);
  const double rr = sqrt(X * X + Y * Y + Z * Z);
  const double c[3] = {rr, acos(Z / rr), atan2(Y, X)};
  for (int d = 0; d < 3; d++) {
    const double xxmin = P.xx[d][0] - ((double)(0 - P.NG) + (1.0 / 2.0)) * P.dxx[d];
    ib[d] = (int)((c[d] - xxmin - (1.0 / 2.0) * P.dxx[d] + ((double)P.NG) * P.dxx[d]) / P.dxx[d] + 0.5);
  }
  for (int d = 0; d < 3; d++) xin[d] = P.xx[d][ib[d]];
}

int parity_signs(double xx1, double xx2, const double xin[3], int8_t par[10]) {
  const double t0 = cos(xx1) * cos(xin[1]);
  const double t1 = sin(xx1) * sin(xin[1]);
  const double t2 = sin(xx2) * sin(xin[2]);
  const double t3 = cos(xx2) * cos(xin[2]);
  const double p1 = t0 + t1 * t2 + t1 * t3;
  const double p2 = t0 * t2 + t0 * t3 + t1;
  const double p3 = t2 + t3;
  const double dots[10] = {1, p1, p2, p3, p1 * p1, p1 * p2, p1 * p3, p2 * p2, p2 * p3, p3 * p3};
  for (int k = 0; k < 10; k++) {
    if (fabs(dots[k]) < 1 - 1e-8 || fabs(dots[k]) > 1 + 1e-8) return -1;
    par[k] = dots[k] < 0 ? -1 : 1;
  }
  return 0;
}

int build_bcstruct(bhah_patch &P) {
  const int NG = P.NG;
  const int *Nx = P.Nx;
  for (int i2 = 0; i2 < Nx[2]; i2++)
    for (int i1 = 0; i1 < Nx[1]; i1++)
      for (int i0 = 0; i0 < Nx[0]; i0++) {
        const bool interior =
            i0 >= NG && i0 < Nx[0] - NG && i1 >= NG && i1 < Nx[1] - NG && i2 >= NG && i2 < Nx[2] - NG;
        if (interior) continue;
        int ib[3];
        double xin[3];
        eigen_inbounds(P, i0, i1, i2, ib, xin);
        if (ib[0] == i0 && ib[1] == i1 && ib[2] == i2) continue;
        for (int d = 0; d < 3; d++)
          if (ib[d] < 0 || ib[d] >= Nx[d])
            return fail(BHAH_ERR_BCSTRUCT, "inner boundary point (%d,%d,%d) maps outside the grid", i0, i1, i2);
        int8_t par[10];
        if (parity_signs(P.xx[1][i1], P.xx[2][i2], xin, par))
          return fail(BHAH_ERR_BCSTRUCT, "parity dot product not +-1 at (%d,%d,%d)", i0, i1, i2);
        P.inner_dst.push_back(i0 + (int64_t)Nx[0] * (i1 + (int64_t)Nx[1] * i2));
        P.inner_src.push_back(ib[0] + (int64_t)Nx[0] * (ib[1] + (int64_t)Nx[1] * ib[2]));
        for (int k = 0; k < 10; k++) P.inner_parity.push_back(par[k]);
      }
  // outer boundary lists, innermost layer first; faces x0min,x0max,x1min,... each using the
  // current bounds and then widening its own direction (bcstruct_set_up…HoleySinhSpherical.c:382-520)
  int imin[3] = {NG, NG, NG};
  int imax[3] = {Nx[0] - NG, Nx[1] - NG, Nx[2] - NG};
  for (int gz = 0; gz < NG; gz++)
    for (int d = 0; d < 3; d++)
      for (int side = 0; side < 2; side++) {
        int r[6];
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
        const int face = 2 * d + side;
        const int F0 = (face == 0) - (face == 1), F1 = (face == 2) - (face == 3), F2 = (face == 4) - (face == 5);
        for (int i2 = r[4]; i2 < r[5]; i2++)
          for (int i1 = r[2]; i1 < r[3]; i1++)
            for (int i0 = r[0]; i0 < r[1]; i0++) {
              int ib[3];
              double xin[3];
              eigen_inbounds(P, i0, i1, i2, ib, xin);
              if (ib[0] == i0 && ib[1] == i1 && ib[2] == i2) {
                const int16_t v[6] = {(int16_t)i0, (int16_t)i1, (int16_t)i2, (int16_t)F0, (int16_t)F1, (int16_t)F2};
                P.outer[gz][d].insert(P.outer[gz][d].end(), v, v + 6);
              }
            }
      }
  return BHAH_OK;
}

template <typename T>
int upload(const std::vector<T> &h, T **dptr) {
  *dptr = nullptr;
  if (h.empty()) return BHAH_OK;
  HIP_TRY(hipMalloc((void **)dptr, sizeof(T) * h.size()));
  HIP_TRY(hipMemcpy(*dptr, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return BHAH_OK;
}

int launch_1d(long n, dim3 &g, dim3 &b) {
  b = dim3(256);
  g = dim3((unsigned)((n + 255) / 256));
  return BHAH_OK;
}

int check_ptr(const void *p, const char *name) {
  if (!p) return fail(BHAH_ERR_ARG, "%s is NULL", name);
  return BHAH_OK;
}

// Two-pass volume integrals of the diagnostic gridfunctions diag: the first pass writes
// VI_SUMS partial sums per workgroup into partial (VI_MAX_BLOCKS * VI_SUMS doubles), the
// second their totals into out (VI_SUMS doubles).  Every partial a call reads is written by
// the same call, so the buffers need no clearing.
int enqueue_volume_integrals(bhah_patch *P, const double *diag, double *partial, double *out) {
  const long nint = (long)P->d.Nxx[0] * P->d.Nxx[1] * P->d.Nxx[2];
  const int nb = (int)std::min<long>(VI_MAX_BLOCKS, std::max<long>(1, (nint + 255) / 256));
  hipLaunchKernelGGL(volume_integrals_partial_kernel, dim3(nb), dim3(256), 0, P->stream, diag, P->d_rfm0, P->d_rfm1,
                     P->d_xx1, P->d_xx2, P->kp(0, 0), fabs(P->dxx[0] * P->dxx[1] * P->dxx[2]), partial);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(volume_integrals_final_kernel, dim3(1), dim3(64), 0, P->stream, partial, nb, out);
  HIP_TRY(hipGetLastError());
  return BHAH_OK;
}

// Device address of `scratch` on the current GPU, looked up once per GPU with that GPU current
// (each GPU holds its own copy of the array at its own address).
constexpr int NOALLOC_MAX_DEVICES = 64;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

int noalloc_scratch(double **out) {
  *out = nullptr;
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(BHAH_ERR_UNSUPPORTED, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(BHAH_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (double *)p;
  }
  *out = g_scratch_ptr[dev];
  return BHAH_OK;
}

// L2 and RMS columns of diags_integration_write_row (diagnostics_volume_integration_helpers.h:332-345).
void finish_volume_integrals(const double sums[VI_SUMS], bhah_volume_integrals *r) {
  for (int q = 0; q < BHAH_DIAG_NUM_RECIPES; q++) {
    const double vol = sums[4 * q];
    r->volume[q] = vol;
    for (int k = 0; k < BHAH_DIAG_NUM_INTEGRANDS; k++) {
      const double I = sums[4 * q + 1 + k];
      r->integral[q][k] = I;
      r->L2[q][k] = sqrt(fmax(I, 0.0));
      r->RMS[q][k] = vol > 0.0 ? sqrt(fmax(I, 0.0) / vol) : 0.0;
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI.
extern "C" {

const char *bhah_hip_last_error(void) { return g_last_error.c_str(); }
int bhah_hip_abi_version(void) { return BHAH_HIP_ABI_VERSION; }

static int setup_on_host(bhah_patch &P, std::vector<double> &rfm0, std::vector<double> &rfm1);
static int setup_on_device(bhah_patch &P, std::vector<double> &rfm0, std::vector<double> &rfm1);
static int patch_upload(bhah_patch *P, const std::vector<double> &rfm0, const std::vector<double> &rfm1,
                        bhah_patch **out);
static int patch_create_impl(const bhah_patch_desc *desc, bhah_patch **out, bool on_device);

int bhah_hip_patch_create(const bhah_patch_desc *desc, bhah_patch **out) {
  return patch_create_impl(desc, out, desc && desc->setup_on_device);
}

static int patch_create_impl(const bhah_patch_desc *desc, bhah_patch **out, bool on_device) {
  if (!desc || !out) return fail(BHAH_ERR_ARG, "NULL argument");
  *out = nullptr;
  const bhah_patch_desc &d = *desc;
  if (d.fd_order != 2 && d.fd_order != 4 && d.fd_order != 6 && d.fd_order != 8)
    return fail(BHAH_ERR_ARG, "fd_order %d unsupported", d.fd_order);
  for (int i = 0; i < 3; i++)
    if (d.Nxx[i] < 1) return fail(BHAH_ERR_ARG, "Nxx[%d] = %d", i, d.Nxx[i]);
  bhah_patch *P = new bhah_patch();
  P->d = d;
  if (d.system == BHAH_SYSTEM_BSSN) {
    for (const auto &k : kBssnKernels)
      if (k.coord == d.coord && k.sym_axes == d.symmetry_axes && k.KO == (d.enable_KO != 0) && k.fd_order == d.fd_order)
        P->k = &k;
    if (!P->k) {
      delete P;
      return fail(BHAH_ERR_UNSUPPORTED, "no BSSN kernel build for coord=%d symmetry_axes=%d KO=%d fd_order=%d", d.coord,
                  d.symmetry_axes, d.enable_KO, d.fd_order);
    }
    P->NG = d.fd_order / 2 + 1;  // upwinding / KO need one extra ghost zone (BHaH_defines_h.py:278-292)
    P->ngfs = BHAH_NUM_EVOL_GFS_BSSN;
    P->naux = BHAH_NUM_AUXEVOL_GFS_BSSN;
  } else if (d.system == BHAH_SYSTEM_WAVE) {
    if (d.coord != BHAH_CARTESIAN || d.fd_order != 4) {
      delete P;
      return fail(BHAH_ERR_UNSUPPORTED, "wave system supports Cartesian, fd_order 4 only");
    }
    P->NG = d.fd_order / 2;
    P->ngfs = BHAH_NUM_EVOL_GFS_WAVE;
  } else {
    delete P;
    return fail(BHAH_ERR_ARG, "unknown system %d", d.system);
  }
  // Grid (numerical_grid_params_Nxx_dxx_xx…HoleySinhSpherical.c:60-104; reference_metric.py:1201-1274).
  if (d.coord == BHAH_CARTESIAN) {
    for (int i = 0; i < 3; i++) {
      P->xxmin[i] = -d.grid_physical_size;
      P->xxmax[i] = d.grid_physical_size;
    }
  } else {
    P->xxmin[0] = 0.0;
    P->xxmax[0] = (d.coord == BHAH_SPHERICAL) ? d.grid_physical_size : 1.0;
    P->xxmin[1] = 0.0;
    P->xxmax[1] = M_PI;
    P->xxmin[2] = -M_PI;
    P->xxmax[2] = M_PI;
  }
  for (int i = 0; i < 3; i++) {
    P->Nx[i] = d.Nxx[i] + 2 * P->NG;
    P->dxx[i] = (P->xxmax[i] - P->xxmin[i]) / ((double)d.Nxx[i]);
    P->invdxx[i] = ((double)d.Nxx[i]) / (P->xxmax[i] - P->xxmin[i]);
    P->xx[i].resize(P->Nx[i]);
    for (int j = 0; j < P->Nx[i]; j++) P->xx[i][j] = P->xxmin[i] + ((double)(j - P->NG) + (1.0 / 2.0)) * P->dxx[i];
  }
  P->Ntot = (long)P->Nx[0] * P->Nx[1] * P->Nx[2];
  if (P->Nx[0] > 32767 || P->Nx[1] > 32767 || P->Nx[2] > 32767 || P->Ntot * (long)P->ngfs > (1L << 31) * 64) {
    delete P;
    return fail(BHAH_ERR_ARG, "grid too large");
  }
  std::vector<double> rfm0(4 * P->Nx[0]), rfm1(3 * P->Nx[1]);
  int rc = on_device ? setup_on_device(*P, rfm0, rfm1) : setup_on_host(*P, rfm0, rfm1);
  double *scr = nullptr;
  if (!rc) rc = noalloc_scratch(&scr);  // set-up of the no-allocation entry point on this GPU
  if (rc) {
    delete P;
    return rc;
  }
  return patch_upload(P, rfm0, rfm1, out);
}

// Reference-metric arrays, boundary structs and CFL timestep on the host.
static int setup_on_host(bhah_patch &PP, std::vector<double> &rfm0, std::vector<double> &rfm1) {
  bhah_patch *P = &PP;
  const bhah_patch_desc &d = P->d;
  // rfm precompute arrays: f0, f0', f0'', f0''' over xx0 and f1, f1', f1'' over xx1
  // (rfm_precompute.py:196-340 with the closed forms of reference_metric.py:1338-1357).
  for (int i = 0; i < P->Nx[0]; i++) {
    const double x = P->xx[0][i];
    double f = 1, d1 = 0, d2 = 0, d3 = 0;
    if (d.coord == BHAH_SPHERICAL) {
      f = x;
      d1 = 1.0;
    } else if (d.coord == BHAH_SINHSPHERICAL) {
      const double A = d.grid_physical_size, w = d.SINHW;
      const double den = exp(1.0 / w) - exp(-1.0 / w);
      const double ep = exp(x / w), em = exp(-x / w);
      f = A * (ep - em) / den;
      d1 = A * (ep + em) / (w * den);
      d2 = A * (ep - em) / (w * w * den);
      d3 = A * (ep + em) / (w * w * w * den);
    }
    rfm0[i] = f;
    rfm0[P->Nx[0] + i] = d1;
    rfm0[2 * P->Nx[0] + i] = d2;
    rfm0[3 * P->Nx[0] + i] = d3;
  }
  for (int j = 0; j < P->Nx[1]; j++) {
    const double y = P->xx[1][j];
    const bool cart = d.coord == BHAH_CARTESIAN;
    rfm1[j] = cart ? 1.0 : sin(y);
    rfm1[P->Nx[1] + j] = cart ? 0.0 : cos(y);
    rfm1[2 * P->Nx[1] + j] = cart ? 0.0 : -sin(y);
  }
  int rc = build_bcstruct(*P);
  if (rc) return rc;
  // CFL timestep: dt = CFL * min over all points of min_i ds_i
  // (numerical_grids_and_timestep_cfl_limited_timestep__openmp.c; ds_min_single_pt…HoleySinhSpherical.c).
  double ds_min = 1e38;
  if (d.coord == BHAH_CARTESIAN) {
    ds_min = fmin(P->dxx[0], fmin(P->dxx[1], P->dxx[2]));
  } else {
    for (int i2 = 0; i2 < P->Nx[2]; i2++)
      for (int i1 = 0; i1 < P->Nx[1]; i1++)
        for (int i0 = 0; i0 < P->Nx[0]; i0++) {
          const double x0 = P->xx[0][i0], x1 = P->xx[1][i1];
          double ds0, ds1, ds2;
          if (d.coord == BHAH_SPHERICAL) {
            ds0 = fabs(P->dxx[0]);
            ds1 = fabs(x0 * P->dxx[1]);
            ds2 = fabs(x0 * P->dxx[2] * sin(x1));
          } else {
            const double tmp0 = (1.0 / (d.SINHW));
            const double tmp4 = d.grid_physical_size / (exp(tmp0) - exp(-tmp0));
            const double tmp2 = exp(tmp0 * x0);
            const double tmp3 = exp(-tmp0 * x0);
            const double tmp5 = tmp4 * (tmp2 - tmp3);
            ds0 = fabs(P->dxx[0] * tmp4 * (tmp0 * tmp2 + tmp0 * tmp3));
            ds1 = fabs(P->dxx[1] * tmp5);
            ds2 = fabs(P->dxx[2] * tmp5 * sin(x1));
          }
          const double m = fmin(ds0, fmin(ds1, ds2));
          ds_min = m < ds_min ? m : ds_min;
        }
  }
  P->dt = d.CFL_FACTOR * ds_min;
  return BHAH_OK;
}

// Device buffers of the set-up, freed when it returns.
struct SetupTmp {
  std::vector<void *> ptrs;
  ~SetupTmp() {
    for (void *q : ptrs) (void)hipFree(q);
  }
  int alloc(void **out, size_t bytes) {
    *out = nullptr;
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess)
      return fail(BHAH_ERR_HIP, "hipMalloc of the set-up workspace failed");
    ptrs.push_back(*out);
    return BHAH_OK;
  }
};
#define SETUP_ALLOC(t, ptr, n) (t).alloc((void **)&(ptr), sizeof(*(ptr)) * (size_t)(n))

// The points of kind `want` in box b, in box order, appended to the host lists: inner points
// to (dst, src, mask), pure outer points to outer (6 shorts each).
static int setup_select(const SetupGeom &g, const SetupBox &b, const unsigned char *kind, int want, const int *src,
                        const unsigned short *mask, std::vector<int> &dst_h, std::vector<int> &src_h,
                        std::vector<unsigned short> &mask_h, std::vector<int16_t> &outer_h) {
  const long n = (long)(b.hi[0] - b.lo[0]) * (b.hi[1] - b.lo[1]) * (b.hi[2] - b.lo[2]);
  if (n <= 0) return BHAH_OK;
  const int nb = (int)((n + SETUP_BLOCK - 1) / SETUP_BLOCK);
  SetupTmp t;
  int rc, *counts, *offs;
  if ((rc = SETUP_ALLOC(t, counts, nb)) || (rc = SETUP_ALLOC(t, offs, nb))) return rc;
  hipLaunchKernelGGL(setup_count_kernel, dim3(nb), dim3(SETUP_BLOCK), 0, nullptr, g, b, n, kind, want, counts);
  HIP_TRY(hipGetLastError());
  std::vector<int> h(nb);
  HIP_TRY(hipMemcpy(h.data(), counts, sizeof(int) * nb, hipMemcpyDeviceToHost));
  long total = 0;
  for (int i = 0; i < nb; i++) {
    const int c = h[i];
    h[i] = (int)total;
    total += c;
  }
  if (!total) return BHAH_OK;
  HIP_TRY(hipMemcpy(offs, h.data(), sizeof(int) * nb, hipMemcpyHostToDevice));
  int *o_dst = nullptr, *o_src = nullptr;
  unsigned short *o_mask = nullptr;
  int16_t *o_outer = nullptr;
  if (want == 1) {
    if ((rc = SETUP_ALLOC(t, o_dst, total)) || (rc = SETUP_ALLOC(t, o_src, total)) || (rc = SETUP_ALLOC(t, o_mask, total))) return rc;
  } else if ((rc = SETUP_ALLOC(t, o_outer, 6 * total))) {
    return rc;
  }
  hipLaunchKernelGGL(setup_scatter_kernel, dim3(nb), dim3(SETUP_BLOCK), 0, nullptr, g, b, n, kind, want, offs, src,
                     mask, o_dst, o_src, o_mask, o_outer);
  HIP_TRY(hipGetLastError());
  if (want == 1) {
    const size_t s = dst_h.size();
    dst_h.resize(s + total);
    src_h.resize(s + total);
    mask_h.resize(s + total);
    HIP_TRY(hipMemcpy(dst_h.data() + s, o_dst, sizeof(int) * total, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(src_h.data() + s, o_src, sizeof(int) * total, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(mask_h.data() + s, o_mask, sizeof(unsigned short) * total, hipMemcpyDeviceToHost));
  } else {
    const size_t s = outer_h.size();
    outer_h.resize(s + 6 * total);
    HIP_TRY(hipMemcpy(outer_h.data() + s, o_outer, sizeof(int16_t) * 6 * total, hipMemcpyDeviceToHost));
  }
  return BHAH_OK;
}

// Reference-metric arrays, boundary structs and CFL timestep computed on the device; the
// boundary lists (O(surface)) and the 1D arrays come back to the host copies the patch keeps.
static int setup_on_device(bhah_patch &P, std::vector<double> &rfm0, std::vector<double> &rfm1) {
  const bhah_patch_desc &d = P.d;
  SetupGeom g;
  g.coord = d.coord;
  g.NG = P.NG;
  g.Nx0 = P.Nx[0];
  g.Nx1 = P.Nx[1];
  g.Nx2 = P.Nx[2];
  g.Ntot = P.Ntot;
  for (int i = 0; i < 3; i++) {
    g.xxmin[i] = P.xxmin[i];
    g.dxx[i] = P.dxx[i];
  }
  g.AMPL = d.grid_physical_size;
  g.SINHW = d.SINHW;
  SetupTmp t;
  int rc;
  double *d_r0, *d_r1, *partial;
  unsigned char *kind;
  int *src, *err;
  unsigned short *mask;
  if ((rc = SETUP_ALLOC(t, d_r0, rfm0.size())) || (rc = SETUP_ALLOC(t, d_r1, rfm1.size())) || (rc = SETUP_ALLOC(t, kind, P.Ntot)) ||
      (rc = SETUP_ALLOC(t, src, P.Ntot)) || (rc = SETUP_ALLOC(t, mask, P.Ntot)) || (rc = SETUP_ALLOC(t, err, 1)) ||
      (rc = SETUP_ALLOC(t, partial, VI_MAX_BLOCKS)))
    return rc;
  const int n1d = std::max(P.Nx[0], P.Nx[1]);
  hipLaunchKernelGGL(setup_rfm_kernel, dim3((n1d + 255) / 256), dim3(256), 0, nullptr, g, d_r0, d_r1);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(rfm0.data(), d_r0, sizeof(double) * rfm0.size(), hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(rfm1.data(), d_r1, sizeof(double) * rfm1.size(), hipMemcpyDeviceToHost));
  // classification of every ghost point, then the inner list in index order
  HIP_TRY(hipMemset(err, 0, sizeof(int)));
  hipLaunchKernelGGL(setup_classify_kernel, dim3((unsigned)((P.Ntot + 255) / 256)), dim3(256), 0, nullptr, g, kind,
                     src, mask, err);
  HIP_TRY(hipGetLastError());
  int herr = 0;
  HIP_TRY(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
  if (herr & 1) return fail(BHAH_ERR_BCSTRUCT, "an inner boundary point maps outside the grid");
  if (herr & 2) return fail(BHAH_ERR_BCSTRUCT, "a parity dot product is not +-1");
  std::vector<int> dst_h, src_h;
  std::vector<unsigned short> mask_h;
  std::vector<int16_t> none;
  const SetupBox all = {{0, 0, 0}, {P.Nx[0], P.Nx[1], P.Nx[2]}, {0, 0, 0}};
  if ((rc = setup_select(g, all, kind, 1, src, mask, dst_h, src_h, mask_h, none))) return rc;
  for (size_t q = 0; q < dst_h.size(); q++) {
    P.inner_dst.push_back(dst_h[q]);
    P.inner_src.push_back(src_h[q]);
    for (int k = 0; k < 10; k++) P.inner_parity.push_back(((mask_h[q] >> k) & 1u) ? -1 : 1);
  }
  // pure outer points, face by face in the order of build_bcstruct (innermost layer first;
  // each face spans the current bounds, then widens its own direction)
  const int NG = P.NG;
  int imin[3] = {NG, NG, NG}, imax[3] = {P.Nx[0] - NG, P.Nx[1] - NG, P.Nx[2] - NG};
  for (int gz = 0; gz < NG; gz++)
    for (int dd = 0; dd < 3; dd++)
      for (int side = 0; side < 2; side++) {
        SetupBox b;
        for (int e = 0; e < 3; e++) {
          b.lo[e] = imin[e];
          b.hi[e] = imax[e];
        }
        if (side == 0) {
          b.lo[dd] = imin[dd] - 1;
          b.hi[dd] = imin[dd];
          imin[dd]--;
        } else {
          b.lo[dd] = imax[dd];
          b.hi[dd] = imax[dd] + 1;
          imax[dd]++;
        }
        const int face = 2 * dd + side;
        b.face[0] = (face == 0) - (face == 1);
        b.face[1] = (face == 2) - (face == 3);
        b.face[2] = (face == 4) - (face == 5);
        std::vector<int> u0, u1;
        std::vector<unsigned short> u2;
        if ((rc = setup_select(g, b, kind, 2, src, mask, u0, u1, u2, P.outer[gz][dd]))) return rc;
      }
  // CFL timestep: dt = CFL * min over all points of min_i ds_i
  double ds_min = 1e38;
  if (d.coord == BHAH_CARTESIAN) {
    ds_min = fmin(P.dxx[0], fmin(P.dxx[1], P.dxx[2]));
  } else {
    const int nb = (int)std::min<long>(VI_MAX_BLOCKS, (P.Ntot + 255) / 256);
    hipLaunchKernelGGL(setup_cfl_kernel, dim3(nb), dim3(256), 0, nullptr, g, partial);
    HIP_TRY(hipGetLastError());
    std::vector<double> h(nb);
    HIP_TRY(hipMemcpy(h.data(), partial, sizeof(double) * nb, hipMemcpyDeviceToHost));
    for (double v : h) ds_min = v < ds_min ? v : ds_min;
  }
  P.dt = d.CFL_FACTOR * ds_min;
  return BHAH_OK;
}

// Upload the set-up to the device and finish the patch (both set-up paths).
static int patch_upload(bhah_patch *P, const std::vector<double> &rfm0, const std::vector<double> &rfm1,
                        bhah_patch **out) {
  const bhah_patch_desc &d = P->d;
  int rc;
  std::vector<unsigned short> mask(P->inner_dst.size());
  std::vector<int> dst(P->inner_dst.size()), src(P->inner_dst.size());
  for (size_t p = 0; p < mask.size(); p++) {
    unsigned m = 0;
    for (int t = 0; t < 10; t++)
      if (P->inner_parity[10 * p + t] < 0) m |= 1u << t;
    mask[p] = (unsigned short)m;
    dst[p] = (int)P->inner_dst[p];
    src[p] = (int)P->inner_src[p];
  }
  std::vector<double> ws(P->ngfs, 1.0), fi(P->ngfs, 0.0);
  if (d.system == BHAH_SYSTEM_BSSN) {
    ws[6] = sqrt(2.0);  // alpha (BSSN_quantities.py:62-86)
    fi[6] = 1.0;        // alpha
    fi[10] = 1.0;       // cf
  }
  std::vector<int> par(kBssnParity, kBssnParity + 24);
  // The inner points grouped by source (sources ascending, each group's points in list order):
  // the adjoint of the copies gathers per source in this fixed order, so it is deterministic.
  std::vector<int> order(dst.size()), by_src_start, by_src_src;
  for (size_t p = 0; p < order.size(); p++) order[p] = (int)p;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return src[a] < src[b]; });
  for (size_t j = 0; j < order.size(); j++)
    if (j == 0 || src[order[j]] != src[order[j - 1]]) {
      by_src_start.push_back((int)j);
      by_src_src.push_back(src[order[j]]);
    }
  by_src_start.push_back((int)order.size());
  P->n_inner_src = (int)by_src_src.size();
  if ((rc = upload(rfm0, &P->d_rfm0)) || (rc = upload(rfm1, &P->d_rfm1)) || (rc = upload(P->xx[0], &P->d_xx0)) ||
      (rc = upload(P->xx[1], &P->d_xx1)) || (rc = upload(P->xx[2], &P->d_xx2)) ||
      (rc = upload(dst, &P->d_inner_dst)) || (rc = upload(src, &P->d_inner_src)) ||
      (rc = upload(mask, &P->d_inner_mask)) || (rc = upload(ws, &P->d_wavespeed)) || (rc = upload(fi, &P->d_finf)) ||
      (rc = upload(par, &P->d_gf_parity)) || (rc = upload(order, &P->d_inner_by_src)) ||
      (rc = upload(by_src_start, &P->d_inner_src_start)) || (rc = upload(by_src_src, &P->d_inner_src_list))) {
    bhah_hip_patch_destroy(P);
    return rc;
  }
  if (d.system == BHAH_SYSTEM_BSSN && !(P->Ntot & 1) && !dst.empty()) {
    std::vector<int> imap((size_t)P->Ntot, -1);
    for (size_t p = 0; p < dst.size(); p++) imap[(size_t)dst[p]] = (int)p;
    std::vector<int> pairs((size_t)P->Ntot);  // (source pair, code) per pair
    bool ok = true;
    for (long q = 0; q < P->Ntot / 2 && ok; q++) {
      long sr[2];
      unsigned m[2];
      for (int e = 0; e < 2; e++) {
        const int ip = imap[(size_t)(2 * q + e)];
        sr[e] = ip >= 0 ? (long)src[(size_t)ip] : 2 * q + e;
        m[e] = ip >= 0 ? mask[(size_t)ip] : 0u;
        if (m[e] >> 12) ok = false;
      }
      if ((sr[0] >> 1) != (sr[1] >> 1)) ok = false;
      pairs[(size_t)(2 * q)] = (int)(sr[0] >> 1);
      pairs[(size_t)(2 * q + 1)] = (int)((sr[0] & 1) | ((sr[1] & 1) << 1) | (m[0] << 4) | (m[1] << 16));
    }
    if (ok && (rc = upload(pairs, &P->d_inner_pairs))) {
      bhah_hip_patch_destroy(P);
      return rc;
    }
    const char *sep = getenv("BHAH_HIP_SEPARATE_INNER_BC");
    P->fold_inner = !(sep && sep[0] == '1');
  }
  for (int gz = 0; gz < P->NG; gz++)
    for (int dd = 0; dd < 3; dd++)
      if ((rc = upload(P->outer[gz][dd], &P->d_outer[gz][dd]))) {
        bhah_hip_patch_destroy(P);
        return rc;
      }
  // radial-line form of the outer points (every layer lists the same (i1, i2) on the x0max face)
  if (d.coord != BHAH_CARTESIAN && P->NG > 0) {
    bool lines = true;
    const auto &L0 = P->outer[0][0];
    for (int gz = 0; gz < P->NG && lines; gz++) {
      if (!P->outer[gz][1].empty() || !P->outer[gz][2].empty() || P->outer[gz][0].size() != L0.size()) lines = false;
      for (size_t q = 0; lines && q < L0.size(); q += 6) {
        const int16_t *a = &P->outer[gz][0][q], *b = &L0[q];
        if (a[0] != P->Nx[0] - P->NG + gz || a[1] != b[1] || a[2] != b[2] || a[3] != -1 || a[4] || a[5]) lines = false;
      }
    }
    if (lines && !L0.empty()) {
      std::vector<int> ln;
      for (size_t q = 0; q < L0.size(); q += 6) {
        ln.push_back(L0[q + 1]);
        ln.push_back(L0[q + 2]);
      }
      if ((rc = upload(ln, &P->d_rad_lines))) {
        bhah_hip_patch_destroy(P);
        return rc;
      }
      P->n_rad_lines = (int)(L0.size() / 6);
    }
  }
  if (hipMalloc((void **)&P->d_red, (VI_MAX_BLOCKS + 1) * VI_SUMS * sizeof(double)) != hipSuccess) {
    bhah_hip_patch_destroy(P);
    return fail(BHAH_ERR_HIP, "hipMalloc failed");
  }
  P->stream = nullptr;  // default (null) stream unless set
  *out = P;
  return BHAH_OK;
}

int bhah_hip_patch_destroy(bhah_patch *P) {
  if (!P) return BHAH_OK;
  if (P->graph_exec) hipGraphExecDestroy(P->graph_exec);
  if (P->adj_exec) hipGraphExecDestroy(P->adj_exec);
  hipFree(P->d_rfm0);
  hipFree(P->d_rfm1);
  hipFree(P->d_xx0);
  hipFree(P->d_xx1);
  hipFree(P->d_xx2);
  hipFree(P->d_inner_dst);
  hipFree(P->d_inner_by_src);
  hipFree(P->d_inner_src_start);
  hipFree(P->d_inner_src_list);
  hipFree(P->d_inner_src);
  hipFree(P->d_inner_mask);
  hipFree(P->d_inner_pairs);
  hipFree(P->d_wavespeed);
  hipFree(P->d_finf);
  hipFree(P->d_red);
  hipFree(P->d_gf_parity);
  hipFree(P->d_adj);
  hipFree(P->d_rad_lines);
  for (int gz = 0; gz < 8; gz++)
    for (int d = 0; d < 3; d++) hipFree(P->d_outer[gz][d]);
  if (P->own_stream) hipStreamDestroy(P->stream);
  delete P;
  return BHAH_OK;
}

int bhah_hip_patch_get_info(const bhah_patch *P, bhah_patch_info *info) {
  if (!P || !info) return fail(BHAH_ERR_ARG, "NULL argument");
  memset(info, 0, sizeof(*info));
  for (int i = 0; i < 3; i++) {
    info->Nxx[i] = P->d.Nxx[i];
    info->Nxx_plus_2NGHOSTS[i] = P->Nx[i];
    info->xxmin[i] = P->xxmin[i];
    info->xxmax[i] = P->xxmax[i];
    info->dxx[i] = P->dxx[i];
    info->invdxx[i] = P->invdxx[i];
  }
  info->nghosts = P->NG;
  info->num_evol_gfs = P->ngfs;
  info->num_auxevol_gfs = P->naux;
  info->Ntot = P->Ntot;
  info->dt = P->dt;
  info->num_inner_boundary_points = (int)P->inner_dst.size();
  int n = 0;
  for (int gz = 0; gz < P->NG; gz++)
    for (int d = 0; d < 3; d++) n += (int)(P->outer[gz][d].size() / 6);
  info->num_pure_outer_boundary_points = n;
  return BHAH_OK;
}

int bhah_hip_patch_set_stream(bhah_patch *P, void *s) {
  if (!P) return fail(BHAH_ERR_ARG, "NULL patch");
  P->stream = (hipStream_t)s;
  P->invalidate_graph();
  return BHAH_OK;
}

int bhah_hip_synchronize(bhah_patch *P) {
  if (!P) return fail(BHAH_ERR_ARG, "NULL patch");
  HIP_TRY(hipStreamSynchronize(P->stream));
  return BHAH_OK;
}

int bhah_hip_patch_get_xx(const bhah_patch *P, double *x0, double *x1, double *x2) {
  if (!P || !x0 || !x1 || !x2) return fail(BHAH_ERR_ARG, "NULL argument");
  double *o[3] = {x0, x1, x2};
  for (int d = 0; d < 3; d++) memcpy(o[d], P->xx[d].data(), sizeof(double) * P->Nx[d]);
  return BHAH_OK;
}

int bhah_hip_patch_get_inner_bc(const bhah_patch *P, int64_t *dst, int64_t *src, int8_t *par) {
  if (!P || !dst || !src || !par) return fail(BHAH_ERR_ARG, "NULL argument");
  memcpy(dst, P->inner_dst.data(), sizeof(int64_t) * P->inner_dst.size());
  memcpy(src, P->inner_src.data(), sizeof(int64_t) * P->inner_src.size());
  memcpy(par, P->inner_parity.data(), P->inner_parity.size());
  return BHAH_OK;
}

int bhah_hip_patch_get_rfm(const bhah_patch *P, double *rfm0, double *rfm1) {
  if (!P || !rfm0 || !rfm1) return fail(BHAH_ERR_ARG, "NULL argument");
  HIP_TRY(hipMemcpy(rfm0, P->d_rfm0, sizeof(double) * 4 * P->Nx[0], hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(rfm1, P->d_rfm1, sizeof(double) * 3 * P->Nx[1], hipMemcpyDeviceToHost));
  return BHAH_OK;
}

int bhah_hip_patch_get_outer_bc_count(const bhah_patch *P, int gz, int d) {
  if (!P || gz < 0 || gz >= P->NG || d < 0 || d > 2) return fail(BHAH_ERR_ARG, "bad arguments");
  return (int)(P->outer[gz][d].size() / 6);
}

int bhah_hip_patch_get_outer_bc(const bhah_patch *P, int gz, int d, int16_t *out) {
  if (!P || !out || gz < 0 || gz >= P->NG || d < 0 || d > 2) return fail(BHAH_ERR_ARG, "bad arguments");
  memcpy(out, P->outer[gz][d].data(), sizeof(int16_t) * P->outer[gz][d].size());
  return BHAH_OK;
}

// ---- interior kernels
int bhah_hip_rhs_eval(bhah_patch *P, const double *aux, const double *in, double *rhs) {
  if (!P || !P->k) return fail(BHAH_ERR_ARG, "patch is not a BSSN patch");
  int rc;
  if ((rc = check_ptr(aux, "auxevol_gfs")) || (rc = check_ptr(in, "in_gfs")) || (rc = check_ptr(rhs, "rhs_gfs"))) return rc;
  dim3 g, b;
  P->interior_dims(g, b);
  // a schedule with 16-byte global-to-LDS loads needs a 16-byte aligned input; any other
  // pointer runs the first (plain) schedule for this call
  const int c = (P->rhs_dma && ((uintptr_t)in & 15)) ? 0 : P->rhs_choice;
  HIP_TRY(P->k->rhs[c].launch(g, b, P->stream, in, aux, rhs, P->d_rfm0, P->d_rfm1, P->kp(0, 0)));
  return BHAH_OK;
}

int bhah_hip_Ricci_eval(bhah_patch *P, const double *in, double *aux) {
  if (!P || !P->k) return fail(BHAH_ERR_ARG, "patch is not a BSSN patch");
  int rc;
  if ((rc = check_ptr(in, "in_gfs")) || (rc = check_ptr(aux, "auxevol_gfs"))) return rc;
  dim3 g, b;
  P->interior_dims(g, b);
  const int c = (P->ricci_dma && ((uintptr_t)in & 15)) ? 0 : P->ricci_choice;
  HIP_TRY(P->k->ricci[c].launch(g, b, P->stream, in, nullptr, aux, P->d_rfm0, P->d_rfm1, P->kp(0, 0)));
  return BHAH_OK;
}

int bhah_hip_enforce_detgammabar_equals_detgammahat(bhah_patch *P, double *gfs) {
  if (!P || !P->k) return fail(BHAH_ERR_ARG, "patch is not a BSSN patch");
  if (!gfs) return fail(BHAH_ERR_ARG, "in_gfs is NULL");
  dim3 g, b;
  P->all_dims(g, b);
  HIP_TRY(P->k->detgamma(g, b, P->stream, gfs, P->d_rfm0, P->d_rfm1, P->kp(0, 0)));
  return BHAH_OK;
}

int bhah_hip_apply_bcs_inner_only(bhah_patch *P, double *gfs) {
  if (!P || !gfs) return fail(BHAH_ERR_ARG, "NULL argument");
  const int n = (int)P->inner_dst.size();
  if (n == 0) return BHAH_OK;
  dim3 g, b;
  launch_1d((long)n, g, b);
  hipLaunchKernelGGL(bc_inner_kernel, g, b, 0, P->stream, P->d_inner_dst, P->d_inner_src, P->d_inner_mask, n, P->ngfs,
                     P->d_gf_parity, gfs, P->Ntot);
  HIP_TRY(hipGetLastError());
  return BHAH_OK;
}

int bhah_hip_apply_bcs_outerextrap_and_inner(bhah_patch *P, double *gfs) {
  if (!P || !gfs) return fail(BHAH_ERR_ARG, "NULL argument");
  for (int gz = 0; gz < P->NG; gz++)
    for (int d = 0; d < 3; d++) {
      const int n = (int)(P->outer[gz][d].size() / 6);
      if (!n) continue;
      dim3 g, b;
      launch_1d((long)n * P->ngfs, g, b);
      hipLaunchKernelGGL(bc_extrap_kernel, g, b, 0, P->stream, P->d_outer[gz][d], n, P->ngfs, gfs, (long)P->Nx[0],
                         (long)P->Nx[0] * P->Nx[1], P->Ntot);
      HIP_TRY(hipGetLastError());
    }
  return bhah_hip_apply_bcs_inner_only(P, gfs);
}

// The radiation condition on the outer points of rhs alone; the caller applies the inner one.
static int apply_bcs_outerradiation(bhah_patch *P, const double *gfs, double *rhs) {
  if (P->d.coord == BHAH_CARTESIAN) return fail(BHAH_ERR_UNSUPPORTED, "radiation BCs implemented for spherical-like coordinates");
  const int R = (P->d.radiation_bc_fd_order > 0 ? P->d.radiation_bc_fd_order : 4) / 2;
  if (R < 1 || R > 3) return fail(BHAH_ERR_ARG, "radiation_bc_fd_order must be 2, 4 or 6");
  RadStencil st;
  rad_stencil(R, st);
  if (P->n_rad_lines) {
    dim3 g, b;
    launch_1d((long)P->n_rad_lines * P->ngfs, g, b);
    // Staged form from a few thousand (line, gridfunction) items up (259,584 items at
    // 192x96x96: 71 -> 26 us per launch). Small patches keep the one-pass form, whose launch
    // has no block-wide steps: at 3,456 items (72x12x2) the staged one took ~3 us longer.
    const long items = (long)P->n_rad_lines * P->ngfs;
    if (items >= RAD_STAGED_MIN_ITEMS && P->Nx[0] >= RAD_SEG && P->NG + R + 2 <= RAD_SEG && P->NG <= 8)
      hipLaunchKernelGGL(bc_radiation_lines_lds_kernel, dim3((unsigned)(((long)P->n_rad_lines * P->ngfs + 255) / 256)),
                         dim3(256), 0, P->stream, P->d_rad_lines, P->n_rad_lines, P->ngfs, P->NG, gfs, rhs, P->d_xx0,
                         P->d_wavespeed, P->d_finf, st, R, P->d.coord, P->d.grid_physical_size, P->d.SINHW, P->Nx[0],
                         (long)P->Nx[0], (long)P->Nx[0] * P->Nx[1], P->Ntot, P->invdxx[0]);
    else
      hipLaunchKernelGGL(bc_radiation_lines_kernel, g, b, 0, P->stream, P->d_rad_lines, P->n_rad_lines, P->ngfs,
                         P->NG, gfs, rhs, P->d_xx0, P->d_wavespeed, P->d_finf, st, R, P->d.coord,
                         P->d.grid_physical_size, P->d.SINHW, P->Nx[0], (long)P->Nx[0], (long)P->Nx[0] * P->Nx[1],
                         P->Ntot, P->invdxx[0]);
    HIP_TRY(hipGetLastError());
    return BHAH_OK;
  }
  for (int gz = 0; gz < P->NG; gz++)
    for (int d = 0; d < 3; d++) {
      const int n = (int)(P->outer[gz][d].size() / 6);
      if (!n) continue;
      dim3 g, b;
      launch_1d((long)n * P->ngfs, g, b);
      hipLaunchKernelGGL(bc_radiation_kernel, g, b, 0, P->stream, P->d_outer[gz][d], n, P->ngfs, gfs, rhs, P->d_xx0,
                         P->d_wavespeed, P->d_finf, st, R, P->d.coord, P->d.grid_physical_size, P->d.SINHW, P->Nx[0],
                         (long)P->Nx[0], (long)P->Nx[0] * P->Nx[1], P->Ntot, P->invdxx[0]);
      HIP_TRY(hipGetLastError());
    }
  return BHAH_OK;
}

int bhah_hip_apply_bcs_outerradiation_and_inner(bhah_patch *P, const double *gfs, double *rhs) {
  if (!P || !gfs || !rhs) return fail(BHAH_ERR_ARG, "NULL argument");
  int rc;
  if ((rc = apply_bcs_outerradiation(P, gfs, rhs))) return rc;
  return bhah_hip_apply_bcs_inner_only(P, rhs);
}

int bhah_hip_MoL_malloc_intermediate_stage_gfs(bhah_patch *P, bhah_MoL_gridfunctions *gf) {
  if (!P || !gf) return fail(BHAH_ERR_ARG, "NULL argument");
  const size_t bytes = sizeof(double) * (size_t)P->Ntot * P->ngfs;
  if (!gf->y_n_gfs) HIP_TRY(hipMalloc((void **)&gf->y_n_gfs, bytes));
  HIP_TRY(hipMalloc((void **)&gf->y_nplus1_running_total_gfs, bytes));
  HIP_TRY(hipMalloc((void **)&gf->k_odd_gfs, bytes));
  HIP_TRY(hipMalloc((void **)&gf->k_even_gfs, bytes));
  const bool new_aux = P->naux && !gf->auxevol_gfs;
  if (new_aux) HIP_TRY(hipMalloc((void **)&gf->auxevol_gfs, sizeof(double) * (size_t)P->Ntot * P->naux));
  const size_t all = bytes;
  HIP_TRY(hipMemsetAsync(gf->y_nplus1_running_total_gfs, 0, all, P->stream));
  HIP_TRY(hipMemsetAsync(gf->k_odd_gfs, 0, all, P->stream));
  HIP_TRY(hipMemsetAsync(gf->k_even_gfs, 0, all, P->stream));
  if (new_aux) HIP_TRY(hipMemsetAsync(gf->auxevol_gfs, 0, sizeof(double) * (size_t)P->Ntot * P->naux, P->stream));
  return BHAH_OK;
}

int bhah_hip_MoL_free(bhah_patch *P, bhah_MoL_gridfunctions *gf) {
  if (!P || !gf) return fail(BHAH_ERR_ARG, "NULL argument");
  hipFree(gf->y_nplus1_running_total_gfs);
  hipFree(gf->k_odd_gfs);
  hipFree(gf->k_even_gfs);
  hipFree(gf->auxevol_gfs);
  gf->y_nplus1_running_total_gfs = gf->k_odd_gfs = gf->k_even_gfs = gf->auxevol_gfs = nullptr;
  return BHAH_OK;
}

// Enqueue the four RK4 substeps (no host synchronisation).
static int enqueue_step(bhah_patch *P, bhah_MoL_gridfunctions *gf, double dt) {
  double *in_s[4] = {gf->y_n_gfs, gf->k_odd_gfs, gf->k_even_gfs, gf->k_odd_gfs};
  double *out_s[4] = {gf->k_odd_gfs, gf->k_even_gfs, gf->k_odd_gfs, gf->k_even_gfs};
  int rc;
  if (P->d.system == BHAH_SYSTEM_WAVE) {
    // rhs_eval + RK update fused at interior points, then extrapolation BCs on the new stage
    // (nrpy/examples/wave_equation_cartesian.py:440-444).
    for (int s = 1; s <= 4; s++) {
      dim3 g, b;
      P->interior_dims(g, b);
      double *next = (s == 4) ? gf->y_n_gfs : out_s[s - 1];
      HIP_TRY(launch_wave_rhs_rk_fd4(g, b, P->stream, in_s[s - 1], gf->y_n_gfs, gf->y_nplus1_running_total_gfs, next,
                                     P->kp(dt, s)));
      if ((rc = bhah_hip_wave_apply_bcs(P, next))) return rc;
    }
    return BHAH_OK;
  }
  // BSSN: Ricci -> RHS -> radiation/extrapolation + inner BCs on RHS -> RK update + det(gammabar)
  // enforcement in one pass (two_blackholes_collide.py:290-313).
  const uintptr_t al = reinterpret_cast<uintptr_t>(gf->y_n_gfs) | reinterpret_cast<uintptr_t>(gf->k_odd_gfs) |
                       reinterpret_cast<uintptr_t>(gf->k_even_gfs) |
                       reinterpret_cast<uintptr_t>(gf->y_nplus1_running_total_gfs);
  if (P->d.outer_bc == BHAH_OUTER_BC_RADIATION && P->fold_inner && P->k->rk_det_inner && P->d_inner_pairs &&
      !(P->Ntot & 1) && !(al & 15)) {
    // The inner copy of k rides in the RK pass, which gathers k at each inner ghost point from
    // its source: k stays in k_even and every stage goes to k_odd, so the pass never writes
    // the array it gathers from.  The same values reach every array as with the separate copy.
    for (int s = 1; s <= 4; s++) {
      double *in = (s == 1) ? gf->y_n_gfs : gf->k_odd_gfs, *k = gf->k_even_gfs;
      double *next = (s == 4) ? gf->y_n_gfs : gf->k_odd_gfs;
      if ((rc = bhah_hip_Ricci_eval(P, in, gf->auxevol_gfs))) return rc;
      if ((rc = bhah_hip_rhs_eval(P, gf->auxevol_gfs, in, k))) return rc;
      if ((rc = apply_bcs_outerradiation(P, in, k))) return rc;
      HIP_TRY(P->k->rk_det_inner(P->stream, k, gf->y_n_gfs, gf->y_nplus1_running_total_gfs, next, P->d_rfm0,
                                 P->d_rfm1, P->kp(dt, s), P->d_inner_pairs, P->d_gf_parity));
    }
    return BHAH_OK;
  }
  for (int s = 1; s <= 4; s++) {
    double *in = in_s[s - 1], *k = out_s[s - 1];
    if ((rc = bhah_hip_Ricci_eval(P, in, gf->auxevol_gfs))) return rc;
    if ((rc = bhah_hip_rhs_eval(P, gf->auxevol_gfs, in, k))) return rc;
    if (P->d.outer_bc == BHAH_OUTER_BC_RADIATION) {
      if ((rc = bhah_hip_apply_bcs_outerradiation_and_inner(P, in, k))) return rc;
    }
    dim3 g, b;
    P->all_dims(g, b);
    double *next = (s == 4) ? gf->y_n_gfs : k;
    HIP_TRY(P->k->rk_det(g, b, P->stream, k, gf->y_n_gfs, gf->y_nplus1_running_total_gfs, next, P->d_rfm0, P->d_rfm1,
                         P->kp(dt, s)));
    if (P->d.outer_bc == BHAH_OUTER_BC_EXTRAPOLATION) {
      // post-RHS extrapolation BCs then det enforcement (two_blackholes_collide.py:304-307)
      if ((rc = bhah_hip_apply_bcs_outerextrap_and_inner(P, next))) return rc;
      if ((rc = bhah_hip_enforce_detgammabar_equals_detgammahat(P, next))) return rc;
    }
  }
  return BHAH_OK;
}

// The reference's time bookkeeping after a step: time from the iteration count, so that
// rounding does not accumulate (MoL_step_forward_in_time.py:473-481).
static void advance_clock(bhah_clock *c) {
  c->time = c->t_0 + (double)(c->nn - c->nn_0 + 1) * c->dt;
  c->nn++;
}

int bhah_hip_MoL_step_forward_in_time(bhah_patch *P, bhah_MoL_gridfunctions *gf, bhah_clock *clk) {
  if (!P || !gf || !clk || !gf->y_n_gfs || !gf->k_odd_gfs || !gf->k_even_gfs || !gf->y_nplus1_running_total_gfs ||
      (P->naux && !gf->auxevol_gfs))
    return fail(BHAH_ERR_ARG, "MoL gridfunctions or clock not set");
  int rc = enqueue_step(P, gf, clk->dt);
  if (rc) return rc;
  advance_clock(clk);
  return BHAH_OK;
}

int bhah_hip_MoL_step_n(bhah_patch *P, bhah_MoL_gridfunctions *gf, bhah_clock *clk, int nsteps, int use_graph) {
  if (!P || !gf || !clk || nsteps < 0) return fail(BHAH_ERR_ARG, "bad arguments");
  if (!use_graph) {
    for (int n = 0; n < nsteps; n++) {
      int rc = bhah_hip_MoL_step_forward_in_time(P, gf, clk);
      if (rc) return rc;
    }
    return BHAH_OK;
  }
  if (!gf->y_n_gfs || !gf->k_odd_gfs || !gf->k_even_gfs || !gf->y_nplus1_running_total_gfs ||
      (P->naux && !gf->auxevol_gfs))
    return fail(BHAH_ERR_ARG, "MoL gridfunctions not allocated");
  const double dt = clk->dt;
  const bool reuse = P->graph_exec && memcmp(&P->graph_gf, gf, sizeof(*gf)) == 0 && P->graph_dt == dt;
  if (!reuse) {
    if (P->graph_exec) {
      hipGraphExecDestroy(P->graph_exec);
      P->graph_exec = nullptr;
    }
    hipStream_t cap = P->stream;
    bool tmp_stream = false;
    if (!cap) {  // the null stream cannot be captured
      HIP_TRY(hipStreamCreate(&cap));
      tmp_stream = true;
    }
    hipStream_t saved = P->stream;
    P->stream = cap;
    HIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_step(P, gf, dt);
    hipGraph_t graph;
    hipError_t e = hipStreamEndCapture(cap, &graph);
    P->stream = saved;
    if (tmp_stream) hipStreamDestroy(cap);
    if (rc) {
      if (e == hipSuccess) hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(BHAH_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&P->graph_exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(BHAH_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    P->graph_gf = *gf;
    P->graph_dt = dt;
  }
  for (int n = 0; n < nsteps; n++) {
    HIP_TRY(hipGraphLaunch(P->graph_exec, P->stream));
    advance_clock(clk);
  }
  return BHAH_OK;
}

int bhah_hip_patch_num_schedules(const bhah_patch *P, int kind) {
  if (!P || !P->k) return fail(BHAH_ERR_ARG, "patch is not a BSSN patch");
  if (kind == BHAH_KERNEL_RHS) return P->k->n_rhs;
  if (kind == BHAH_KERNEL_RICCI) return P->k->n_ricci;
  return fail(BHAH_ERR_ARG, "unknown kernel kind %d", kind);
}

const char *bhah_hip_patch_schedule_name(const bhah_patch *P, int kind, int index) {
  if (!P || !P->k) return nullptr;
  if (kind == BHAH_KERNEL_RHS && index >= 0 && index < P->k->n_rhs) return P->k->rhs[index].name;
  if (kind == BHAH_KERNEL_RICCI && index >= 0 && index < P->k->n_ricci) return P->k->ricci[index].name;
  return nullptr;
}

int bhah_hip_patch_select_schedule(bhah_patch *P, int kind, int index) {
  const int n = bhah_hip_patch_num_schedules(P, kind);
  if (n < 0) return n;
  if (index < 0 || index >= n) return fail(BHAH_ERR_ARG, "schedule index %d out of range [0,%d)", index, n);
  const char *nm = (kind == BHAH_KERNEL_RHS ? P->k->rhs : P->k->ricci)[index].name;
  if (strstr(nm, "_u32") && P->Ntot * 24 * 8 >= (1L << 31))
    return fail(BHAH_ERR_UNSUPPORTED, "schedule %s addresses arrays below 2 GiB only", nm);
  const bool dma = strstr(nm, "_dma") != nullptr;
  if (dma && (P->Nx[0] & 1))
    return fail(BHAH_ERR_UNSUPPORTED, "schedule %s loads xx0 pairs and needs an even Nx0 (have %d)", nm, P->Nx[0]);
  if (dma && P->Ntot * 8 >= (1L << 32))
    return fail(BHAH_ERR_UNSUPPORTED, "schedule %s addresses a gridfunction below 4 GiB only", nm);
  if (kind == BHAH_KERNEL_RHS) {
    P->rhs_choice = index;
    P->rhs_dma = dma;
  } else {
    P->ricci_choice = index;
    P->ricci_dma = dma;
  }
  P->invalidate_graph();
  return BHAH_OK;
}

int bhah_hip_patch_get_schedule(const bhah_patch *P, int kind) {
  if (!P || !P->k) return fail(BHAH_ERR_ARG, "patch is not a BSSN patch");
  return kind == BHAH_KERNEL_RHS ? P->rhs_choice : P->ricci_choice;
}

int bhah_hip_patch_set_tile(bhah_patch *P, int tile_x, int tile_y) {
  if (!P) return fail(BHAH_ERR_ARG, "NULL patch");
  if (tile_x * tile_y != 256 || tile_x < 16 || tile_x % 16) return fail(BHAH_ERR_ARG, "tile must be 256 threads, tile_x a multiple of 16");
  P->tile_x = tile_x;
  P->tile_y = tile_y;
  P->invalidate_graph();
  return BHAH_OK;
}

int bhah_hip_wave_rhs_eval(bhah_patch *P, const double *in, double *rhs) {
  if (!P || P->d.system != BHAH_SYSTEM_WAVE) return fail(BHAH_ERR_ARG, "patch is not a wave patch");
  if (!in || !rhs) return fail(BHAH_ERR_ARG, "NULL argument");
  dim3 g, b;
  P->interior_dims(g, b);
  HIP_TRY(launch_wave_rhs_fd4(g, b, P->stream, in, rhs, P->kp(0, 0)));
  return BHAH_OK;
}

int bhah_hip_wave_apply_bcs(bhah_patch *P, double *gfs) {
  if (!P || !gfs) return fail(BHAH_ERR_ARG, "NULL argument");
  // Cartesian: every ghost point is a pure outer point; the outer lists follow the same
  // face order as the example's FACE_UPDATE sweep.
  return bhah_hip_apply_bcs_outerextrap_and_inner(P, gfs);
}

int bhah_hip_diagnostics_constraints(bhah_patch *P, const double *aux, const double *in, double *diag,
                                     bhah_volume_integrals *integrals) {
  if (!P || !P->k) return fail(BHAH_ERR_ARG, "patch is not a BSSN patch");
  if (!aux || !in || !diag) return fail(BHAH_ERR_ARG, "NULL argument");
  dim3 g, b;
  P->interior_dims(g, b);
  HIP_TRY(P->k->constraints(g, b, P->stream, in, aux, diag, P->d_rfm0, P->d_rfm1, P->kp(0, 0)));
  if (!integrals) return BHAH_OK;
  int rc = enqueue_volume_integrals(P, diag, P->d_red, P->d_red + VI_MAX_BLOCKS * VI_SUMS);
  if (rc) return rc;
  double sums[VI_SUMS];
  HIP_TRY(hipMemcpyAsync(sums, P->d_red + VI_MAX_BLOCKS * VI_SUMS, sizeof(sums), hipMemcpyDeviceToHost, P->stream));
  HIP_TRY(hipStreamSynchronize(P->stream));
  finish_volume_integrals(sums, integrals);
  return BHAH_OK;
}

int bhah_hip_diagnostics_constraints_noalloc(bhah_patch *P, const double *aux, const double *in, double *diag,
                                             double *integral_sums) {
  if (!P || !P->k) return fail(BHAH_ERR_ARG, "patch is not a BSSN patch");
  if (!aux || !in || !diag || !integral_sums) return fail(BHAH_ERR_ARG, "NULL argument");
  // the reduction writes at most VI_MAX_BLOCKS x VI_SUMS partial sums whatever the grid size
  const long nint = (long)P->d.Nxx[0] * P->d.Nxx[1] * P->d.Nxx[2];
  const long nb = std::min<long>(VI_MAX_BLOCKS, std::max<long>(1, (nint + 255) / 256));
  if (nb * VI_SUMS > NOALLOC_SCRATCH_DOUBLES)
    return fail(BHAH_ERR_UNSUPPORTED, "%ld partial sums exceed the %d-value scratch", nb * VI_SUMS,
                NOALLOC_SCRATCH_DOUBLES);
  double *scr = nullptr;
  int rc = noalloc_scratch(&scr);
  if (rc) return rc;
  dim3 g, b;
  P->interior_dims(g, b);
  HIP_TRY(P->k->constraints(g, b, P->stream, in, aux, diag, P->d_rfm0, P->d_rfm1, P->kp(0, 0)));
  return enqueue_volume_integrals(P, diag, scr, integral_sums);
}

int bhah_hip_diagnostics(bhah_patch *P, const bhah_clock *clk, double every, const double *y_n, double *aux,
                         double *diag, bhah_volume_integrals *integrals, int *output_step) {
  if (!P || !P->k) return fail(BHAH_ERR_ARG, "patch is not a BSSN patch");
  if (!clk || !y_n || !aux || !diag || !integrals || !output_step) return fail(BHAH_ERR_ARG, "NULL argument");
  if (!(every > 0.0)) return fail(BHAH_ERR_ARG, "diagnostics_output_every must be > 0 (have %g)", every);
  *output_step = 0;
  // output step rule of the reference driver (diagnostics.py:195)
  if (!(fabs(round(clk->time / every) * every - clk->time) < 0.5 * clk->dt)) return BHAH_OK;
  int rc;
  if ((rc = bhah_hip_Ricci_eval(P, y_n, aux))) return rc;
  if ((rc = bhah_hip_diagnostics_constraints(P, aux, y_n, diag, integrals))) return rc;
  *output_step = 1;
  return BHAH_OK;
}

int bhah_hip_diagnostic_gfs_malloc(const bhah_patch *P, double **diag) {
  if (!P || !diag) return fail(BHAH_ERR_ARG, "NULL argument");
  *diag = nullptr;
  HIP_TRY(hipMalloc((void **)diag, sizeof(double) * BHAH_NUM_DIAG_GFS * (size_t)P->Ntot));
  return BHAH_OK;
}

int bhah_hip_diagnostic_gfs_free(double *diag) {
  HIP_TRY(hipFree(diag));
  return BHAH_OK;
}

// ---------------------------------------------------------------------------
// Checkpoints in the reference's per-grid payload layout, and initial data on the device.

extern "C" int64_t bhah_hip_checkpoint_payload_bytes(const bhah_patch *P) {
  if (!P) return fail(BHAH_ERR_ARG, "NULL patch");
  return (int64_t)sizeof(int) * (1 + P->Ntot) + (int64_t)sizeof(double) * P->Ntot * P->ngfs;
}

extern "C" int bhah_hip_checkpoint_pack(bhah_patch *P, const double *y_n, void *out) {
  if (!P || !y_n || !out) return fail(BHAH_ERR_ARG, "NULL argument");
  if (P->Ntot > 0x7fffffffL) return fail(BHAH_ERR_ARG, "grid has more points than an int index holds");
  const long n = P->Ntot, nv = n * P->ngfs;
  char *o = (char *)out;
  const int count = (int)n;
  memcpy(o, &count, sizeof(int));
  int *ind = (int *)(o + sizeof(int));
  for (long i = 0; i < n; i++) ind[i] = (int)i;
  double *dtmp = nullptr;
  HIP_TRY(hipMallocAsync((void **)&dtmp, nv * sizeof(double), P->stream));
  dim3 g, b;
  launch_1d(nv, g, b);
  hipLaunchKernelGGL(checkpoint_pack_kernel, g, b, 0, P->stream, y_n, dtmp, n, P->ngfs);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(o + sizeof(int) * (1 + n), dtmp, nv * sizeof(double), hipMemcpyDeviceToHost,
                                          P->stream);
  hipFreeAsync(dtmp, P->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
  if (e != hipSuccess) return fail(BHAH_ERR_HIP, "checkpoint pack failed: %s", hipGetErrorString(e));
  return BHAH_OK;
}

extern "C" int bhah_hip_checkpoint_unpack(bhah_patch *P, const void *payload, int64_t bytes, double *y_n) {
  if (!P || !payload || !y_n) return fail(BHAH_ERR_ARG, "NULL argument");
  const char *c = (const char *)payload;
  if (bytes < (int64_t)sizeof(int)) return fail(BHAH_ERR_ARG, "checkpoint payload truncated");
  int count;
  memcpy(&count, c, sizeof(int));
  if (count < 0 || count > P->Ntot) return fail(BHAH_ERR_ARG, "invalid count=%d (ntot=%ld)", count, P->Ntot);
  const int64_t need = (int64_t)sizeof(int) * (1 + count) + (int64_t)sizeof(double) * count * P->ngfs;
  if (bytes < need) return fail(BHAH_ERR_ARG, "checkpoint payload truncated (%lld < %lld bytes)", (long long)bytes,
                                (long long)need);
  const int *ind = (const int *)(c + sizeof(int));
  for (int i = 0; i < count; i++) {
    if (ind[i] < 0 || ind[i] >= P->Ntot) return fail(BHAH_ERR_ARG, "invalid gridpoint index %d at element %d", ind[i], i);
    if (i > 0 && ind[i] <= ind[i - 1]) return fail(BHAH_ERR_ARG, "non-increasing gridpoint index at element %d", i);
  }
  // the reference reader zeroes every point of y_n, then scatters the listed ones
  // (read_checkpoint__openmp.c:163-171)
  {
    dim3 g, b;
    launch_1d(P->Ntot * P->ngfs, g, b);
    hipLaunchKernelGGL(fill_kernel, g, b, 0, P->stream, y_n, 0.0, P->Ntot * P->ngfs);
    HIP_TRY(hipGetLastError());
  }
  if (!count) {
    HIP_TRY(hipStreamSynchronize(P->stream));
    return BHAH_OK;
  }
  const long nv = (long)count * P->ngfs;
  int *dind = nullptr;
  double *ddat = nullptr;
  HIP_TRY(hipMallocAsync((void **)&dind, count * sizeof(int), P->stream));
  HIP_TRY(hipMallocAsync((void **)&ddat, nv * sizeof(double), P->stream));
  hipError_t e = hipMemcpyAsync(dind, ind, count * sizeof(int), hipMemcpyHostToDevice, P->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(ddat, c + sizeof(int) * (1 + (long)count), nv * sizeof(double), hipMemcpyHostToDevice,
                       P->stream);
  if (e == hipSuccess) {
    dim3 g, b;
    launch_1d(nv, g, b);
    hipLaunchKernelGGL(checkpoint_unpack_kernel, g, b, 0, P->stream, dind, ddat, (long)count, P->ngfs, P->Ntot, y_n);
    e = hipGetLastError();
  }
  hipFreeAsync(dind, P->stream);
  hipFreeAsync(ddat, P->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(P->stream);
  if (e != hipSuccess) return fail(BHAH_ERR_HIP, "checkpoint unpack failed: %s", hipGetErrorString(e));
  return BHAH_OK;
}

extern "C" int bhah_hip_initial_data_brill_lindquist(bhah_patch *P, const double bh[8], double *y_n) {
  if (!P || !bh || !y_n) return fail(BHAH_ERR_ARG, "NULL argument");
  if (P->d.system != BHAH_SYSTEM_BSSN) return fail(BHAH_ERR_ARG, "patch is not a BSSN patch");
  dim3 g, b;
  launch_1d(P->Ntot, g, b);
  hipLaunchKernelGGL(initial_data_bl_kernel, g, b, 0, P->stream, y_n, P->d.coord, P->d.grid_physical_size, P->d.SINHW,
                     P->Nx[0], P->Nx[1], P->Nx[2], P->NG, P->xxmin[0], P->xxmin[1], P->xxmin[2], P->dxx[0], P->dxx[1],
                     P->dxx[2], bh[0], bh[1], bh[2], bh[3], bh[4], bh[5], bh[6], bh[7], P->Ntot);
  HIP_TRY(hipGetLastError());
  return BHAH_OK;
}

extern "C" int bhah_hip_initial_data_convert_ADM_Cartesian_to_BSSN(bhah_patch *P, const double *adm, double *y_n) {
  if (!P || !P->k) return fail(BHAH_ERR_ARG, "patch is not a BSSN patch");
  if (!adm || !y_n) return fail(BHAH_ERR_ARG, "NULL argument");
  if (P->d.coord == BHAH_CARTESIAN) return fail(BHAH_ERR_UNSUPPORTED, "BSSN patches are spherical-like");
  dim3 g, b;
  launch_1d(P->Ntot, g, b);
  hipLaunchKernelGGL(adm_cartesian_to_bssn_kernel, g, b, 0, P->stream, adm, y_n, P->d_rfm0, P->d_xx1, P->d_xx2,
                     P->Nx[0], P->Nx[1], P->Ntot);
  HIP_TRY(hipGetLastError());
  int rc;
  // inner boundary points hold parity copies before lambda^i is differenced
  // (ADM_Initial_Data_Reader__BSSN_Converter.py:986-1010)
  if ((rc = bhah_hip_apply_bcs_inner_only(P, y_n))) return rc;
  P->interior_dims(g, b);
  HIP_TRY(P->k->initial_lambda(g, b, P->stream, y_n, P->d_rfm0, P->d_rfm1, P->kp(0, 0)));
  // then initial_data() applies the outer extrapolation and inner BCs to everything (initial_data.py:252)
  return bhah_hip_apply_bcs_outerextrap_and_inner(P, y_n);
}

// ---------------------------------------------------------------------------
// Reverse mode of one RK4 step (BSSN, radiation outer BCs).  The step's stages are
// recomputed from y_n into a patch-owned workspace, then the adjoint runs backwards:
// det-enforcement Jacobians, RK combinations, boundary conditions, rhs_eval and
// Ricci_eval transposes and the transposed stencils.

static int adjoint_rhs_bcs(bhah_patch *P, double *lam_k, double *lam_f) {
  // forward order: radiation layers gz = 0..NG-1 (directions 0..2), then inner parity
  if (P->n_inner_src) {
    dim3 g, b;
    launch_1d((long)P->n_inner_src * P->ngfs, g, b);
    hipLaunchKernelGGL(bc_inner_adjoint_kernel, g, b, 0, P->stream, P->d_inner_dst, P->d_inner_by_src,
                       P->d_inner_src_start, P->d_inner_src_list, P->n_inner_src, P->d_inner_mask, P->ngfs,
                       P->d_gf_parity, lam_k, P->Ntot);
    HIP_TRY(hipGetLastError());
  }
  const int R = (P->d.radiation_bc_fd_order > 0 ? P->d.radiation_bc_fd_order : 4) / 2;
  RadStencil st;
  rad_stencil(R, st);
  if (P->n_rad_lines) {  // outer points on radial lines: every layer in one launch
    dim3 g, b;
    launch_1d((long)P->n_rad_lines * P->ngfs, g, b);
    hipLaunchKernelGGL(bc_radiation_adjoint_lines_kernel, g, b, 0, P->stream, P->d_rad_lines, P->n_rad_lines,
                       P->ngfs, P->NG, lam_k, lam_f, P->d_xx0, P->d_wavespeed, st, R, P->d.coord,
                       P->d.grid_physical_size, P->d.SINHW, P->Nx[0], (long)P->Nx[0], (long)P->Nx[0] * P->Nx[1],
                       P->Ntot, P->invdxx[0]);
    HIP_TRY(hipGetLastError());
    return BHAH_OK;
  }
  for (int gz = P->NG - 1; gz >= 0; gz--)
    for (int d = 2; d >= 0; d--) {
      const int n = (int)(P->outer[gz][d].size() / 6);
      if (!n) continue;
      dim3 g, b;
      launch_1d((long)n * P->ngfs, g, b);
      hipLaunchKernelGGL(bc_radiation_adjoint_kernel, g, b, 0, P->stream, P->d_outer[gz][d], n, P->ngfs, lam_k, lam_f,
                         P->d_xx0, P->d_wavespeed, st, R, P->d.coord, P->d.grid_physical_size, P->d.SINHW, P->Nx[0],
                         (long)P->Nx[0], (long)P->Nx[0] * P->Nx[1], P->Ntot, P->invdxx[0]);
      HIP_TRY(hipGetLastError());
    }
  return BHAH_OK;
}

// A step's tape: the stage data the reverse needs, in one block of 8 * ngfs * Ntot +
// 4 * 6 * Ntot doubles: Y0 (the state y_n), Y1..Y3 (stage inputs), K1..K4 (right-hand sides
// after the BCs), AUX1..AUX4 (RbarDD of each stage).
struct StepTape {
  double *Y[4], *K, *AUX;
};

static StepTape tape_view(const bhah_patch *P, double *t) {
  const long N = P->Ntot, nv = N * P->ngfs;
  StepTape T;
  for (int s = 0; s < 4; s++) T.Y[s] = t + s * nv;
  T.K = t + 4 * nv;
  T.AUX = t + 8 * nv;
  return T;
}

static long tape_doubles(const bhah_patch *P) { return (8L * P->ngfs + 4L * 6) * P->Ntot; }

// Enqueue the stages of one RK4 step from T.Y[0] into the tape, the kernels and their order
// being those of the MoL step.  With y_next, the step's last update (rk_det of stage 4) also
// writes y_{n+1} there, as the MoL step writes it into y_n.
static int enqueue_stages(bhah_patch *P, const StepTape &T, double *yrun, double dt, double *y_next) {
  const long N = P->Ntot, nv = N * P->ngfs;
  dim3 ga, ba;
  P->all_dims(ga, ba);
  int rc;
  for (int s = 1; s <= 4; s++) {
    double *in = T.Y[s - 1], *k = T.K + (s - 1) * nv, *aux = T.AUX + (s - 1) * 6 * N;
    if ((rc = bhah_hip_Ricci_eval(P, in, aux))) return rc;
    if ((rc = bhah_hip_rhs_eval(P, aux, in, k))) return rc;
    if ((rc = bhah_hip_apply_bcs_outerradiation_and_inner(P, in, k))) return rc;
    if (s < 4)
      HIP_TRY(P->k->rk_det(ga, ba, P->stream, k, T.Y[0], yrun, T.Y[s], P->d_rfm0, P->d_rfm1, P->kp(dt, s)));
    else if (y_next)
      HIP_TRY(P->k->rk_det(ga, ba, P->stream, k, T.Y[0], yrun, y_next, P->d_rfm0, P->d_rfm1, P->kp(dt, s)));
  }
  return BHAH_OK;
}

// Workspace of the reverse sweep, after the recomputed tape and y_run in P->d_adj:
// lamK1..4, lamY, lamYN (6 nv) and the adjoint slots S (nslots * Ntot, zero at every
// non-interior point: filled once when the workspace is made, then written at interior points
// only, every slot at every point, by adj_rhs and adj_ricci).
struct AdjWork {
  double *lamK, *lamY, *lamYN, *S;
};

static AdjWork adj_work(const bhah_patch *P) {
  const long nv = P->Ntot * P->ngfs;
  double *w = P->d_adj + tape_doubles(P) + nv;
  return {w, w + 4 * nv, w + 5 * nv, w + 6 * nv};
}

static int ensure_adj_workspace(bhah_patch *P) {
  if (P->d_adj) return BHAH_OK;
  const long N = P->Ntot, nv = N * P->ngfs;
  const int nslots = P->k->adj_nslots();
  const size_t total = (size_t)tape_doubles(P) + (size_t)nv + 6 * (size_t)nv + (size_t)nslots * N;
  HIP_TRY(hipMalloc((void **)&P->d_adj, total * sizeof(double)));
  dim3 g, b;
  launch_1d((long)nslots * N, g, b);
  hipLaunchKernelGGL(fill_kernel, g, b, 0, P->stream, adj_work(P).S, 0.0, (long)nslots * N);
  HIP_TRY(hipGetLastError());
  return BHAH_OK;
}

// Enqueue the reverse of one RK4 step from its tape on P->stream (no allocation, no memset, so
// the sequence can be captured in a graph).
static int enqueue_reverse(bhah_patch *P, const StepTape &T, double dt, double *lam) {
  const long N = P->Ntot, nv = N * P->ngfs;
  const AdjWork W = adj_work(P);
  const double *y_n = T.Y[0];
  const double *K = T.K;
  hipStream_t st = P->stream;
  dim3 ga, ba, gi, bi, g1, b1;
  P->all_dims(ga, ba);
  P->interior_dims(gi, bi);
  launch_1d(nv, g1, b1);
  int rc;
  const double a_stage[4] = {0.5 * dt, 0.5 * dt, dt, 0.0};  // Y[s] = D(y_n + a_stage[s-1] K_s), s = 1..3
  const double b_final[4] = {dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0};
  // y_{n+1} = D(y_n + sum b K): lamYN = J_D^T lam, then lamK_j = b_j lamYN
  HIP_TRY(P->k->adj_final(ga, ba, st, y_n, K, b_final[0], b_final[1], b_final[2], b_final[3], lam, W.lamYN, W.lamK,
                          P->d_rfm0, P->d_rfm1, P->kp(dt, 0)));
  for (int s = 4; s >= 1; s--) {
    const double *in = T.Y[s - 1], *aux = T.AUX + (s - 1) * 6 * N;
    double *lk = W.lamK + (s - 1) * nv;
    hipLaunchKernelGGL(fill_kernel, g1, b1, 0, st, W.lamY, 0.0, nv);
    HIP_TRY(hipGetLastError());
    if ((rc = adjoint_rhs_bcs(P, lk, W.lamY))) return rc;
    HIP_TRY(P->k->adj_rhs(gi, bi, st, in, aux, lk, W.S, P->d_rfm0, P->d_rfm1, P->kp(dt, 0)));
    HIP_TRY(P->k->adj_ricci(gi, bi, st, in, W.S, P->d_rfm0, P->d_rfm1, P->kp(dt, 0)));
    HIP_TRY(P->k->adj_gather(g1, b1, st, W.S, W.lamY, P->kp(dt, 0)));
    if (s == 1) {
      // the adjoint of y_n: lamYN + lamY, written straight to the caller's array
      hipLaunchKernelGGL(add_kernel, g1, b1, 0, st, lam, W.lamYN, W.lamY, nv);
      HIP_TRY(hipGetLastError());
      break;
    }
    // Y[s-1] = D(y_n + a K_{s-1}): lamYN += J_D^T lamY, lamK_{s-1} += a J_D^T lamY
    HIP_TRY(P->k->adj_stage(ga, ba, st, y_n, K + (s - 2) * nv, a_stage[s - 2], W.lamY, W.lamYN,
                            W.lamK + (s - 2) * nv, P->d_rfm0, P->d_rfm1, P->kp(dt, 0)));
  }
  return BHAH_OK;
}

// The reverse of one step from y_n alone: recompute the tape into the workspace, then reverse.
static int enqueue_adjoint(bhah_patch *P, const double *y_n, double dt, double *lam) {
  StepTape T = tape_view(P, P->d_adj);
  T.Y[0] = const_cast<double *>(y_n);  // read only: the stages are written to Y1..Y3, K, AUX
  int rc;
  if ((rc = enqueue_stages(P, T, P->d_adj + tape_doubles(P), dt, nullptr))) return rc;
  return enqueue_reverse(P, T, dt, lam);
}

static int check_adjoint_patch(const bhah_patch *P) {
  if (!P->k->adj_rhs) return fail(BHAH_ERR_UNSUPPORTED, "kernel build %s has no adjoint kernels", P->k->tag);
  if (P->d.outer_bc != BHAH_OUTER_BC_RADIATION) return fail(BHAH_ERR_UNSUPPORTED, "adjoint needs radiation outer BCs");
  return BHAH_OK;
}

extern "C" int64_t bhah_hip_MoL_step_tape_doubles(const bhah_patch *P) {
  if (!P || !P->k) return -1;
  return tape_doubles(P);
}

extern "C" int bhah_hip_MoL_step_forward_taped(bhah_patch *P, double *y_n, double dt, double *tape) {
  if (!P || !P->k || !y_n || !tape) return fail(BHAH_ERR_ARG, "bad arguments");
  int rc;
  if ((rc = check_adjoint_patch(P)) || (rc = ensure_adj_workspace(P))) return rc;
  const long nv = P->Ntot * P->ngfs;
  StepTape T = tape_view(P, tape);
  HIP_TRY(hipMemcpyAsync(T.Y[0], y_n, nv * sizeof(double), hipMemcpyDeviceToDevice, P->stream));
  return enqueue_stages(P, T, P->d_adj + tape_doubles(P), dt, y_n);
}

extern "C" int bhah_hip_MoL_step_adjoint_taped(bhah_patch *P, const double *tape, double dt, double *lam) {
  if (!P || !P->k || !tape || !lam) return fail(BHAH_ERR_ARG, "bad arguments");
  int rc;
  if ((rc = check_adjoint_patch(P)) || (rc = ensure_adj_workspace(P))) return rc;
  return enqueue_reverse(P, tape_view(P, const_cast<double *>(tape)), dt, lam);
}

extern "C" int bhah_hip_MoL_step_adjoint(bhah_patch *P, const double *y_n, double dt, double *lam) {
  if (!P || !P->k || !y_n || !lam) return fail(BHAH_ERR_ARG, "bad arguments");
  int rc0;
  if ((rc0 = check_adjoint_patch(P)) || (rc0 = ensure_adj_workspace(P))) return rc0;
  // The whole reverse step is one hipGraph, replayed while y_n, lam and dt stay the same (a
  // caller walking back over stored states copies each into one buffer).
  if (!(P->adj_exec && P->adj_y == y_n && P->adj_lam == lam && P->adj_dt == dt)) {
    if (P->adj_exec) {
      (void)hipGraphExecDestroy(P->adj_exec);
      P->adj_exec = nullptr;
    }
    hipStream_t cap = P->stream;
    bool tmp_stream = false;
    if (!cap) {  // the null stream cannot be captured
      HIP_TRY(hipStreamCreate(&cap));
      tmp_stream = true;
    }
    hipStream_t saved = P->stream;
    P->stream = cap;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    const int rc = e == hipSuccess ? enqueue_adjoint(P, y_n, dt, lam) : BHAH_OK;
    hipGraph_t graph = nullptr;
    if (e == hipSuccess) e = hipStreamEndCapture(cap, &graph);
    P->stream = saved;
    if (tmp_stream) (void)hipStreamDestroy(cap);
    if (rc || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc ? rc : fail(BHAH_ERR_HIP, "adjoint graph capture failed: %s", hipGetErrorString(e));
    }
    e = hipGraphInstantiate(&P->adj_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(BHAH_ERR_HIP, "adjoint graph instantiate failed: %s", hipGetErrorString(e));
    P->adj_y = y_n;
    P->adj_lam = lam;
    P->adj_dt = dt;
  }
  HIP_TRY(hipGraphLaunch(P->adj_exec, P->stream));
  return BHAH_OK;
}

// ---------------------------------------------------------------------------
// Batched ensembles: many independent BSSN patches of one kernel build (uneven
// grids allowed) advanced together.  Every stage of an RK4 substep is one
// persistent launch over the tile list of all members (a fixed grid of
// workgroups walks the list), so a whole step is ~5 launches per substep
// whatever the member count, and is captured once into a hipGraph.

int bhah_hip_ensemble_destroy(bhah_ensemble *E);

}  // extern "C"

struct bhah_ensemble {
  std::vector<bhah_patch *> P;
  const BssnKernels *k = nullptr;
  int NG = 0;
  std::vector<bhah_member> hmem;
  bhah_member *d_mem = nullptr;
  bhah_item *d_int = nullptr, *d_all = nullptr;
  int n_int = 0, n_all = 0;
  // tile lists of the selected LDS-tiled schedules, in their own tile units ([0] rhs, [1] Ricci)
  bhah_item *d_lds[2] = {};
  int n_lds[2] = {};
  bool aligned16 = true;  // every bound stage array 16-byte aligned (schedules with pair loads need it)
  ens_outer_pt *d_outer[8] = {};
  int n_outer[8] = {};
  ens_inner_pt *d_inner = nullptr;
  int n_inner = 0;
  double *d_wavespeed = nullptr, *d_finf = nullptr;
  int *d_gf_parity = nullptr;
  RadStencil st;
  int R = 2;
  int rhs_choice = 0, ricci_choice = 0;
  int tile_x = 64, tile_y = 4, max_wg = 0;
  bool bound = false;
  hipStream_t stream = nullptr;
  hipGraphExec_t graph_exec = nullptr;

  void invalidate_graph() {
    if (graph_exec) {
      hipGraphExecDestroy(graph_exec);
      graph_exec = nullptr;
    }
  }
  dim3 grid_for(int nitems) const {
    const int g = (max_wg > 0 && max_wg < nitems) ? max_wg : nitems;
    return dim3(g > 0 ? g : 1);
  }
};

namespace {

// Tile lists: interior tiles (tile_x x tile_y threads over (i0, i1), one xx2 plane each)
// and, for the RK pass, chunks of 256 consecutive points of each member's arrays, all members
// concatenated.
int ensemble_build_items(bhah_ensemble *E) {
  std::vector<bhah_item> in, all;
  for (int m = 0; m < (int)E->P.size(); m++) {
    const bhah_patch *P = E->P[m];
    const int n0 = P->d.Nxx[0], n1 = P->d.Nxx[1], n2 = P->d.Nxx[2];
    for (int bz = 0; bz < n2; bz++)
      for (int by = 0; by < (n1 + E->tile_y - 1) / E->tile_y; by++)
        for (int bx = 0; bx < (n0 + E->tile_x - 1) / E->tile_x; bx++) in.push_back({m, bx, by, bz});
    for (long c = 0; c < (P->Ntot + 255) / 256; c++) all.push_back({m, (int)c, 0, 0});
  }
  hipFree(E->d_int);
  hipFree(E->d_all);
  E->d_int = E->d_all = nullptr;
  int rc;
  if ((rc = upload(in, &E->d_int)) || (rc = upload(all, &E->d_all))) return rc;
  E->n_int = (int)in.size();
  E->n_all = (int)all.size();
  return BHAH_OK;
}

// Tile list of a selected LDS-tiled schedule (kind 0: rhs, 1: Ricci): every member's interior
// in the schedule's own TX x TY x TZ tiles.
int ensemble_build_lds_items(bhah_ensemble *E, int kind) {
  const BssnCandidate &c = kind == 0 ? E->k->rhs[E->rhs_choice] : E->k->ricci[E->ricci_choice];
  hipFree(E->d_lds[kind]);
  E->d_lds[kind] = nullptr;
  E->n_lds[kind] = 0;
  if (!c.lds[0]) return BHAH_OK;
  std::vector<bhah_item> v;
  for (int m = 0; m < (int)E->P.size(); m++) {
    const bhah_patch *P = E->P[m];
    const int t0 = (P->d.Nxx[0] + c.lds[0] - 1) / c.lds[0], t1 = (P->d.Nxx[1] + c.lds[1] - 1) / c.lds[1],
              t2 = (P->d.Nxx[2] + c.lds[2] - 1) / c.lds[2];
    for (int bz = 0; bz < t2; bz++)
      for (int by = 0; by < t1; by++)
        for (int bx = 0; bx < t0; bx++) v.push_back({m, bx, by, bz});
  }
  int rc;
  if ((rc = upload(v, &E->d_lds[kind]))) return rc;
  E->n_lds[kind] = (int)v.size();
  return BHAH_OK;
}

int ensemble_enqueue_step(bhah_ensemble *E) {
  const int in_s[4] = {0, 2, 3, 2}, k_s[4] = {2, 3, 2, 3};
  const BssnCandidate &ric = E->k->ricci[E->ricci_choice], &rhs = E->k->rhs[E->rhs_choice];
  const dim3 tb(E->tile_x, E->tile_y, 1);
  if (!E->aligned16 && (strstr(ric.name, "_dma") || strstr(rhs.name, "_dma")))
    return fail(BHAH_ERR_UNSUPPORTED, "schedules with 16-byte pair loads need 16-byte aligned member arrays");
  // LDS-tiled schedules walk their own tile lists (the launcher sets the workgroup size)
  const bhah_item *ric_items = ric.lds[0] ? E->d_lds[1] : E->d_int, *rhs_items = rhs.lds[0] ? E->d_lds[0] : E->d_int;
  const int ric_n = ric.lds[0] ? E->n_lds[1] : E->n_int, rhs_n = rhs.lds[0] ? E->n_lds[0] : E->n_int;
  for (int s = 1; s <= 4; s++) {
    const int in = in_s[s - 1], k = k_s[s - 1], out = (s == 4) ? 0 : k;
    HIP_TRY(ric.ens(E->grid_for(ric_n), tb, E->stream, E->d_mem, ric_items, ric_n, in, 4, 4));
    HIP_TRY(rhs.ens(E->grid_for(rhs_n), tb, E->stream, E->d_mem, rhs_items, rhs_n, in, 4, k));
    for (int gz = 0; gz < E->NG; gz++) {
      if (!E->n_outer[gz]) continue;
      dim3 g, b;
      launch_1d((long)E->n_outer[gz] * BHAH_NUM_EVOL_GFS_BSSN, g, b);
      hipLaunchKernelGGL(ens_bc_radiation_kernel, g, b, 0, E->stream, E->d_mem, E->d_outer[gz], E->n_outer[gz],
                         BHAH_NUM_EVOL_GFS_BSSN, in, k, E->d_wavespeed, E->d_finf, E->st, E->R);
      HIP_TRY(hipGetLastError());
    }
    if (E->n_inner) {
      dim3 g, b;
      launch_1d((long)E->n_inner, g, b);
      hipLaunchKernelGGL(ens_bc_inner_kernel, g, b, 0, E->stream, E->d_mem, E->d_inner, E->n_inner,
                         BHAH_NUM_EVOL_GFS_BSSN, E->d_gf_parity, k);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(E->k->ens_rk_det(E->grid_for(E->n_all), dim3(256, 1, 1), E->stream, E->d_mem, E->d_all, E->n_all, k, out, s));
  }
  return BHAH_OK;
}

}  // namespace

extern "C" {

int bhah_hip_ensemble_create(const bhah_patch_desc *descs, int n, bhah_ensemble **out) {
  if (!descs || !out || n < 1) return fail(BHAH_ERR_ARG, "bad arguments");
  *out = nullptr;
  bhah_ensemble *E = new bhah_ensemble();
  int rc;
  for (int m = 0; m < n; m++) {
    const bhah_patch_desc &d = descs[m];
    if (d.system != BHAH_SYSTEM_BSSN || d.outer_bc != BHAH_OUTER_BC_RADIATION || d.coord == BHAH_CARTESIAN) {
      bhah_hip_ensemble_destroy(E);
      return fail(BHAH_ERR_UNSUPPORTED, "ensembles take BSSN patches in spherical-like coordinates with radiation BCs");
    }
    bhah_patch *P = nullptr;
    if ((rc = bhah_hip_patch_create(&d, &P))) {
      bhah_hip_ensemble_destroy(E);
      return rc;
    }
    E->P.push_back(P);
    if (m == 0) {
      E->k = P->k;
      E->NG = P->NG;
    } else if (P->k != E->k || d.radiation_bc_fd_order != descs[0].radiation_bc_fd_order) {
      bhah_hip_ensemble_destroy(E);
      return fail(BHAH_ERR_ARG, "member %d uses a different kernel build than member 0", m);
    }
  }
  if (!E->k->ens_rk_det || !E->k->rhs[0].ens) {
    bhah_hip_ensemble_destroy(E);
    return fail(BHAH_ERR_UNSUPPORTED, "kernel build %s has no batched-ensemble kernels", E->k->tag);
  }
  E->R = (descs[0].radiation_bc_fd_order > 0 ? descs[0].radiation_bc_fd_order : 4) / 2;
  if (E->R < 1 || E->R > 3) {
    bhah_hip_ensemble_destroy(E);
    return fail(BHAH_ERR_ARG, "radiation_bc_fd_order must be 2, 4 or 6");
  }
  rad_stencil(E->R, E->st);
  // concatenated boundary records
  std::vector<ens_inner_pt> inner;
  std::vector<ens_outer_pt> outer[8];
  for (int m = 0; m < n; m++) {
    const bhah_patch *P = E->P[m];
    for (size_t p = 0; p < P->inner_dst.size(); p++) {
      unsigned mk = 0;
      for (int t = 0; t < 10; t++)
        if (P->inner_parity[10 * p + t] < 0) mk |= 1u << t;
      inner.push_back({m, (int)P->inner_dst[p], (int)P->inner_src[p], mk});
    }
    for (int gz = 0; gz < E->NG; gz++)
      for (int d = 0; d < 3; d++)
        for (size_t q = 0; q < P->outer[gz][d].size(); q += 6) {
          ens_outer_pt e;
          e.m = m;
          for (int j = 0; j < 6; j++) e.q[j] = P->outer[gz][d][q + j];
          outer[gz].push_back(e);
        }
  }
  std::vector<double> ws(24, 1.0), fi(24, 0.0);
  ws[6] = sqrt(2.0);
  fi[6] = 1.0;
  fi[10] = 1.0;
  std::vector<int> par(kBssnParity, kBssnParity + 24);
  if ((rc = upload(inner, &E->d_inner)) || (rc = upload(ws, &E->d_wavespeed)) || (rc = upload(fi, &E->d_finf)) ||
      (rc = upload(par, &E->d_gf_parity))) {
    bhah_hip_ensemble_destroy(E);
    return rc;
  }
  E->n_inner = (int)inner.size();
  for (int gz = 0; gz < E->NG; gz++) {
    if ((rc = upload(outer[gz], &E->d_outer[gz]))) {
      bhah_hip_ensemble_destroy(E);
      return rc;
    }
    E->n_outer[gz] = (int)outer[gz].size();
  }
  if ((rc = ensemble_build_items(E))) {
    bhah_hip_ensemble_destroy(E);
    return rc;
  }
  if (hipMalloc((void **)&E->d_mem, sizeof(bhah_member) * n) != hipSuccess) {
    bhah_hip_ensemble_destroy(E);
    return fail(BHAH_ERR_HIP, "hipMalloc failed");
  }
  E->hmem.resize(n);
  *out = E;
  return BHAH_OK;
}

int bhah_hip_ensemble_destroy(bhah_ensemble *E) {
  if (!E) return BHAH_OK;
  if (E->graph_exec) hipGraphExecDestroy(E->graph_exec);
  for (bhah_patch *P : E->P) bhah_hip_patch_destroy(P);
  hipFree(E->d_mem);
  hipFree(E->d_int);
  hipFree(E->d_all);
  hipFree(E->d_lds[0]);
  hipFree(E->d_lds[1]);
  for (int gz = 0; gz < 8; gz++) hipFree(E->d_outer[gz]);
  hipFree(E->d_inner);
  hipFree(E->d_wavespeed);
  hipFree(E->d_finf);
  hipFree(E->d_gf_parity);
  delete E;
  return BHAH_OK;
}

int bhah_hip_ensemble_size(const bhah_ensemble *E) {
  if (!E) return fail(BHAH_ERR_ARG, "NULL ensemble");
  return (int)E->P.size();
}

int bhah_hip_ensemble_member(bhah_ensemble *E, int m, bhah_patch **patch) {
  if (!E || !patch || m < 0 || m >= (int)E->P.size()) return fail(BHAH_ERR_ARG, "bad arguments");
  *patch = E->P[m];
  return BHAH_OK;
}

int bhah_hip_ensemble_bind(bhah_ensemble *E, const bhah_MoL_gridfunctions *gfs, const double *dt) {
  if (!E || !gfs) return fail(BHAH_ERR_ARG, "NULL argument");
  E->aligned16 = true;
  for (size_t m = 0; m < E->P.size(); m++) {
    const bhah_patch *P = E->P[m];
    const bhah_MoL_gridfunctions &g = gfs[m];
    if (!g.y_n_gfs || !g.y_nplus1_running_total_gfs || !g.k_odd_gfs || !g.k_even_gfs || !g.auxevol_gfs)
      return fail(BHAH_ERR_ARG, "member %d: MoL gridfunctions not allocated", (int)m);
    bhah_member &M = E->hmem[m];
    M.kp = P->kp(dt ? dt[m] : P->dt, 0);
    M.rfm0 = P->d_rfm0;
    M.rfm1 = P->d_rfm1;
    M.xx0 = P->d_xx0;
    M.arr[0] = g.y_n_gfs;
    M.arr[1] = g.y_nplus1_running_total_gfs;
    M.arr[2] = g.k_odd_gfs;
    M.arr[3] = g.k_even_gfs;
    M.arr[4] = g.auxevol_gfs;
    for (int a = 0; a < 4; a++)
      if ((uintptr_t)M.arr[a] & 15) E->aligned16 = false;
    M.coord = P->d.coord;
    M.AMPL = P->d.grid_physical_size;
    M.SINHW = P->d.SINHW;
  }
  HIP_TRY(hipMemcpy(E->d_mem, E->hmem.data(), sizeof(bhah_member) * E->hmem.size(), hipMemcpyHostToDevice));
  E->bound = true;
  E->invalidate_graph();
  return BHAH_OK;
}

int bhah_hip_ensemble_set_stream(bhah_ensemble *E, void *s) {
  if (!E) return fail(BHAH_ERR_ARG, "NULL ensemble");
  E->stream = (hipStream_t)s;
  E->invalidate_graph();
  return BHAH_OK;
}

int bhah_hip_ensemble_select_schedule(bhah_ensemble *E, int kind, int index) {
  if (!E) return fail(BHAH_ERR_ARG, "NULL ensemble");
  const int n = kind == BHAH_KERNEL_RHS ? E->k->n_rhs : kind == BHAH_KERNEL_RICCI ? E->k->n_ricci : -1;
  if (n < 0) return fail(BHAH_ERR_ARG, "unknown kernel kind %d", kind);
  if (index < 0 || index >= n) return fail(BHAH_ERR_ARG, "schedule index %d out of range [0,%d)", index, n);
  const BssnCandidate &c = (kind == BHAH_KERNEL_RHS ? E->k->rhs : E->k->ricci)[index];
  if (!c.ens) return fail(BHAH_ERR_UNSUPPORTED, "schedule %s has no batched form", c.name);
  for (const bhah_patch *P : E->P) {
    if (strstr(c.name, "_u32") && P->Ntot * 24 * 8 >= (1L << 31))
      return fail(BHAH_ERR_UNSUPPORTED, "schedule %s addresses arrays below 2 GiB only", c.name);
    if (strstr(c.name, "_dma") && (P->Nx[0] & 1))
      return fail(BHAH_ERR_UNSUPPORTED, "schedule %s loads xx0 pairs and needs an even Nx0 in every member", c.name);
    if (strstr(c.name, "_dma") && P->Ntot * 8 >= (1L << 32))
      return fail(BHAH_ERR_UNSUPPORTED, "schedule %s addresses a gridfunction below 4 GiB only", c.name);
  }
  if (kind == BHAH_KERNEL_RHS) E->rhs_choice = index;
  else E->ricci_choice = index;
  E->invalidate_graph();
  return ensemble_build_lds_items(E, kind == BHAH_KERNEL_RHS ? 0 : 1);
}

int bhah_hip_ensemble_set_launch(bhah_ensemble *E, int tile_x, int tile_y, int max_workgroups) {
  if (!E) return fail(BHAH_ERR_ARG, "NULL ensemble");
  if (tile_x * tile_y != 256 || tile_x < 16 || tile_x % 16) return fail(BHAH_ERR_ARG, "tile must be 256 threads, tile_x a multiple of 16");
  if (max_workgroups < 0) return fail(BHAH_ERR_ARG, "max_workgroups must be >= 0 (0: one per tile)");
  E->tile_x = tile_x;
  E->tile_y = tile_y;
  E->max_wg = max_workgroups;
  E->invalidate_graph();
  return ensemble_build_items(E);
}

int bhah_hip_ensemble_step_n(bhah_ensemble *E, int nsteps, int use_graph) {
  if (!E || nsteps < 0) return fail(BHAH_ERR_ARG, "bad arguments");
  if (!E->bound) return fail(BHAH_ERR_ARG, "ensemble arrays not bound (bhah_hip_ensemble_bind)");
  if (!use_graph) {
    for (int n = 0; n < nsteps; n++) {
      int rc = ensemble_enqueue_step(E);
      if (rc) return rc;
    }
    return BHAH_OK;
  }
  if (!E->graph_exec) {
    hipStream_t cap = E->stream;
    bool tmp = false;
    if (!cap) {
      HIP_TRY(hipStreamCreate(&cap));
      tmp = true;
    }
    hipStream_t saved = E->stream;
    E->stream = cap;
    HIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    int rc = ensemble_enqueue_step(E);
    hipGraph_t graph;
    hipError_t e = hipStreamEndCapture(cap, &graph);
    E->stream = saved;
    if (tmp) hipStreamDestroy(cap);
    if (rc) return rc;
    if (e != hipSuccess) return fail(BHAH_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&E->graph_exec, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(BHAH_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
  }
  for (int n = 0; n < nsteps; n++) HIP_TRY(hipGraphLaunch(E->graph_exec, E->stream));
  return BHAH_OK;
}

int bhah_hip_ensemble_synchronize(bhah_ensemble *E) {
  if (!E) return fail(BHAH_ERR_ARG, "NULL ensemble");
  HIP_TRY(hipStreamSynchronize(E->stream));
  return BHAH_OK;
}

}  // extern "C"
