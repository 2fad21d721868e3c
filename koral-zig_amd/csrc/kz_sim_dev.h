// kz_sim_dev.h — device-side run state of one grid (field planes, grid
// geometry, step scalars) and the indexing and load helpers shared by the
// step kernels of kz_step.hip and kz_options.hip.
#ifndef KZ_SIM_DEV_H
#define KZ_SIM_DEV_H

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/kz_step.h"
#include "../../include/kz_implicit.h"
#include "kz_dev.h"
#include "kz_dev_solve.h"

using namespace kzd;

// geometry planes: the upper triangle of g (symmetric bit for bit, checked
// at create), all 16 entries of G (its mathematically zero entries carry
// either sign, so it is kept whole), then gdet, gttpert, alpha
#define GEO_NC 29
#define KR_NC 40  // Christoffels Gamma^l_{nu k}: symmetric in (nu, k)
#define BCX_REC 66

struct GridD {
  int nx, ny, nz, ng, ngx, ngy, ngz, sx, sy, sz;
  long ncell;
  double minx, miny, minz, dx, dy, dz;
  int fsx[3], fsy[3];
  long nface[3];
};

struct Dev {
  double *p, *u, *ut0, *ut1, *ut2, *dut1, *drt1, *drt2, *uforget;
  double *flb[3];
  double *scal;  // planes: ahd_m[3], arad_m[3], tstepden
  int32_t *flags;
  double *emf[3];
  double *gc, *gf[3], *kris;
  double *bcx;
  unsigned long long *cnt;  // [0] n_fail [1] n_iters [2] n_solves [3] n_guard
  unsigned long long *tsd;  // [0] max bits [1] min bits
  int *err;
  double *ss;  // step scalars (see k_begin_step)
  double *stage;  // staging for AoS <-> SoA
  // split implicit solve: per domain cell (index t = ix + nx*(iy + ny*iz))
  double *imp_uu00;  // [13][nd] conserved state of the incoming primitives
  double *imp_pp0;   // [13][nd] Newton start (after the bisection warm start)
  double *imp_aux;   // [IMP_NAUX][nd] reference-state scalars, see ImpAux
  int32_t *imp_slow;           // cells that need rungs past the first
  unsigned long long *imp_nslow;
  int32_t *imp_list;             // [2][nd] cells by starting block (MHD, RAD)
  unsigned long long *imp_ncnt;  // [2] lengths of the two lists
  int32_t *imp_rl;                // [2][nd] cells the persistent first rung left
  unsigned long long *imp_rc;     // [2] their lengths
  unsigned long long *imp_qpos;   // [3] queue claim positions
  // the u2p fixup list
  unsigned *sync;     // [SY_NFIX] u2p fixup list length
  int32_t *fix_list;  // domain cells flagged for a u2p fixup
  // radiative viscosity and dynamo (allocated when enabled)
  double *bl;       // [BL_NC][ncell] Boyer-Lindquist sidecar, see BlPlane
  double *rijvisc;  // [16][ncell] viscous R^i_j (domain + one-cell ring, corners 0)
  double *dyn_a;    // [ncell] Delta A_phi of the last mimicDynamo
  double *scaleth;  // [nx] density-weighted theta scale height per column
};
// BL sidecar planes (MetricCache.blGeom + jacMy2Bl, precompute.zig:379,457-470)
enum { BL_G = 0, BL_GC = 16, BL_ALPHA = 32, BL_GDET = 33, BL_R = 34, BL_TH = 35, BL_J = 36, BL_NC = 52 };
enum { SY_NFIX = 0, SY_N = 4 };

// imp_aux planes
enum { IA_EHAT = 0, IA_UINT, IA_SGAS, IA_KES, IA_KGA, IA_KRA, IA_KRR, IA_SW, IMP_NAUX };

enum { SC_AHD = 0, SC_ARAD = 3, SC_TSTEPDEN = 6, SC_N = 7 };
enum { SS_DT = 0, SS_DTG, SS_INV_DTG, SS_NINV_DTG, SS_INV_DT, SS_NINV_DT, SS_C12, SS_HALF, SS_T, SS_OWN_DT, SS_USER, SS_N };

static const double GAM_IMEX = 0x1.2bec333018867p-2;  // 1 - 1/sqrt(2), correctly rounded
static constexpr double KZ_PI = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// indexing helpers
__device__ __forceinline__ long cidx(const GridD &g, long ix, long iy, long iz) {
  return (ix + g.ngx) + (iy + g.ngy) * (long)g.sx + (iz + g.ngz) * (long)g.sy * g.sx;
}
__device__ __forceinline__ long fidx(const GridD &g, int d, long ix, long iy, long iz) {
  long jx = ix + g.ngx, jy = iy + g.ngy, jz = iz + g.ngz;
  return (jz * g.fsy[d] + jy) * (long)g.fsx[d] + jx;
}
__device__ __forceinline__ double xl_(const GridD &g, long i) { return g.minx + (double)i * g.dx; }
__device__ __forceinline__ double yl_(const GridD &g, long i) { return g.miny + (double)i * g.dy; }
__device__ __forceinline__ double zl_(const GridD &g, long i) { return g.minz + (double)i * g.dz; }
__device__ __forceinline__ double xc_(const GridD &g, long i) { return 0.5 * (xl_(g, i) + xl_(g, i + 1)); }
__device__ __forceinline__ double yc_(const GridD &g, long i) { return 0.5 * (yl_(g, i) + yl_(g, i + 1)); }
__device__ __forceinline__ double zc_(const GridD &g, long i) { return 0.5 * (zl_(g, i) + zl_(g, i + 1)); }
__device__ __forceinline__ double csize(const GridD &g, long i, int d) {
  return d == 0 ? xl_(g, i + 1) - xl_(g, i) : (d == 1 ? yl_(g, i + 1) - yl_(g, i) : zl_(g, i + 1) - zl_(g, i));
}
// index of the symmetric pair (a, b), a <= b, in a 10-entry upper triangle
__host__ __device__ constexpr int tri(int a, int b) { return a * 4 - a * (a - 1) / 2 + (b - a); }
__device__ __forceinline__ void load_geo(const double *__restrict__ t, long stride, long i, double x1, Geo &gm) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = a; b < 4; ++b) {
      double v = t[tri(a, b) * stride + i];
      gm.g[a][b] = v;
      gm.g[b][a] = v;
    }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) gm.G[a][b] = t[(10 + a * 4 + b) * stride + i];
  gm.gdet = t[26 * stride + i];
  gm.gttpert = t[27 * stride + i];
  gm.alpha = t[28 * stride + i];
  gm.x1 = x1;
}
__device__ __forceinline__ void geo_c(const GridD &g, const Dev &d, long ix, long iy, long iz, Geo &gm) {
  load_geo(d.gc, g.ncell, cidx(g, ix, iy, iz), xc_(g, ix), gm);
}
__device__ __forceinline__ void geo_f(const GridD &g, const Dev &d, int dim, long ix, long iy, long iz, Geo &gm) {
  load_geo(d.gf[dim], g.nface[dim], fidx(g, dim, ix, iy, iz), dim == 0 ? xl_(g, ix) : xc_(g, ix), gm);
}
__device__ __forceinline__ void ld(const double *__restrict__ f, long n, long c, double *pp) {
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) pp[iv] = f[iv * n + c];
}
__device__ __forceinline__ void st(double *__restrict__ f, long n, long c, const double *pp) {
#pragma unroll
  for (int iv = 0; iv < NV; ++iv) f[iv * n + c] = pp[iv];
}
__device__ __forceinline__ void set_err(int *err, int code) {
  if (code) atomicCAS(err, 0, code);
}
__device__ __forceinline__ bool polar_row(const kz_physics &ph, const GridD &g, long iy) {
  if (!ph.correct_polaraxis) return false;
  long nc = ph.nccorrectpolar;
  return iy < nc || iy > g.ny - nc - 1;
}

#define TID (blockIdx.x * (long)blockDim.x + threadIdx.x)
static inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Minimum waves per SIMD each heavy kernel is compiled for (1 = no bound;
// n bounds the kernel to 512/n vector registers).  Chosen by a measured
// sweep at 384x360: 2 waves help the memory-bound update and u2p passes and
// the implicit preparation (their register need is just above 256); the face
// flux and the first Newton rung lose more to spills than they gain.
#define KZ_WPE(n) __attribute__((amdgpu_waves_per_eu(n)))
#ifndef KZ_WPE_FACE
#define KZ_WPE_FACE 2
#endif
#ifndef KZ_WPE_UPDATE
#define KZ_WPE_UPDATE 2
#endif
#ifndef KZ_WPE_U2P
#define KZ_WPE_U2P 2
#endif
#ifndef KZ_WPE_WS
#define KZ_WPE_WS 1
#endif
#ifndef KZ_WPE_PREP
#define KZ_WPE_PREP 2
#endif
#ifndef KZ_WPE_RUNG0
#define KZ_WPE_RUNG0 1
#endif
#ifndef KZ_WPE_BC
#define KZ_WPE_BC 1
#endif
// Width of the persistent implicit launches in blocks of one wave per CU
// (4: one wave per SIMD, which is what the first-rung registers allow to be
// resident at once); chosen by the tuner on the independent-cell batch.
#ifndef KZ_PERSIST_PER_CU
#define KZ_PERSIST_PER_CU 4
#endif

#endif
