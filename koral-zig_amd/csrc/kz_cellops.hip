// kz_cellops.hip — batched per-point entry points over the device physics.
//
// Each entry point runs one per-point function of the hot path over n
// records on the GPU (one thread per record): the same device code the
// step kernels inline, exposed so that the reference's per-function golden
// vectors (koral/testing/*_golden_tests.zig) can be replayed against it.
// Buffers are host memory; each call stages them through device memory on
// the null stream and returns when the results are back.
#include <hip/hip_runtime.h>

#include <cstdio>

#include "../../include/kz_cellops.h"
#include "kz_dev.h"
#include "kz_dev_solve.h"

using namespace kzd;

#define TID (blockIdx.x * (long)blockDim.x + threadIdx.x)

__device__ __forceinline__ void geo_from_kz(const kz_geom &k, Geo &o) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      o.g[i][j] = k.gg[i][j];
      o.G[i][j] = k.GG[i][j];
    }
  o.gdet = k.gdet;
  o.alpha = k.alpha;
  o.gttpert = k.gttpert;
  o.x1 = k.xx[1];
}

__global__ void b_p2u(long n, const double *pp, const kz_geom *g, double gam, double *uu) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  double p[NV], u[NV];
  for (int k = 0; k < NV; ++k) p[k] = pp[i * NV + k];
  p2u(p, u, gm, gam);
  for (int k = 0; k < NV; ++k) uu[i * NV + k] = u[k];
}

__global__ void b_u2p_mhd(long n, const double *uu, double *pp, const kz_geom *g, double gam, kz_floor_params f,
                          kz_u2p_result *r) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  double p[NV], u[NV];
  for (int k = 0; k < NV; ++k) {
    p[k] = pp[i * NV + k];
    u[k] = uu[i * NV + k];
  }
  int ent = 0;
  int ret = u2p_mhd(u, p, gm, gam, f, &ent);
  for (int k = 0; k < NV; ++k) pp[i * NV + k] = p[k];
  kz_u2p_result o;
  o.ret = ret;
  o.entropy_used = ent;
  o.fixup = ret < -1;
  o._pad = 0;
  r[i] = o;
}

template <bool ZAMO>
__global__ void b_check_floors_mhd(long n, double *pp, const kz_geom *g, double gam, kz_floor_params f, int32_t *ret,
                                   unsigned long long *nguard) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  double p[NV];
  for (int k = 0; k < NV; ++k) p[k] = pp[i * NV + k];
  int guard = 0;
  ret[i] = check_floors_mhd<ZAMO>(p, gm, gam, f, &guard);
  for (int k = 0; k < NV; ++k) pp[i * NV + k] = p[k];
  if (guard) atomicAdd(nguard, (unsigned long long)guard);
}

__global__ void b_u2p_rad(long n, const double *uu, double *pp, const kz_geom *g, kz_rad_params rp,
                          kz_u2prad_result *r) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  double p[NV], u[NV];
  for (int k = 0; k < NV; ++k) {
    p[k] = pp[i * NV + k];
    u[k] = uu[i * NV + k];
  }
  int corr = 0;
  int ret = u2p_rad(u, p, gm, rp, &corr);
  for (int k = 0; k < NV; ++k) pp[i * NV + k] = p[k];
  kz_u2prad_result o;
  o.ret = ret;
  o.corrected = corr;
  r[i] = o;
}

__global__ void b_check_floors_rad(long n, double *pp, const kz_geom *g, kz_rad_params rp, int32_t *ret) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  double p[NV];
  for (int k = 0; k < NV; ++k) p[k] = pp[i * NV + k];
  ret[i] = check_floors_rad(p, gm, rp);
  for (int k = 0; k < NV; ++k) pp[i * NV + k] = p[k];
}

__global__ void __launch_bounds__(64) b_implicit(long n, double *uu, double *pp, const kz_geom *g, const double *dt,
                                                 double gam, kz_rad_params rp, kz_opac_params op,
                                                 kz_implicit_params ip, kz_implicit_result *r) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  double p[NV], u[NV];
  for (int k = 0; k < NV; ++k) {
    p[k] = pp[i * NV + k];
    u[k] = uu[i * NV + k];
  }
  r[i] = solve_implicit_lab(u, p, gm, dt[i], gam, rp, op, ip);
  for (int k = 0; k < NV; ++k) {
    pp[i * NV + k] = p[k];
    uu[i * NV + k] = u[k];
  }
}

__global__ void b_reconstruct(long n, const int32_t *order, const double *theta, const double *u5, const double *dx5,
                              double *l, double *r) {
  long i = TID;
  if (i >= n) return;
  double u[5], d[5];
  for (int k = 0; k < 5; ++k) {
    u[k] = u5[i * 5 + k];
    d[k] = dx5[i * 5 + k];
  }
  reconstruct(order[i], theta[i], u, d, l + i, r + i);
}

__global__ void b_u2p_solver_w(long n, const double *uu, double *pp, const kz_geom *g, int etype, double gam,
                               int32_t *ret) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  double p[NV], u[NV];
  for (int k = 0; k < NV; ++k) {
    p[k] = pp[i * NV + k];
    u[k] = uu[i * NV + k];
  }
  ret[i] = u2p_solver_w(u, p, gm, etype, gam);
  for (int k = 0; k < NV; ++k) pp[i * NV + k] = p[k];
}

__global__ void b_flux_prime(long n, const double *pp, const int32_t *idim, const kz_geom *g, double gam, double *ff,
                             int32_t *err) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  double p[NV], f[NV];
  for (int k = 0; k < NV; ++k) p[k] = pp[i * NV + k];
  err[i] = flux_prime(p, idim[i], gm, gam, f);
  for (int k = 0; k < NV; ++k) ff[i * NV + k] = f[k];
}

__global__ void b_gas_ws(long n, const double *pp, const kz_geom *g, double gam, int a0, int a1, int a2, double *a6) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  int act[3] = {a0, a1, a2};
  double a[6];
  gas_wavespeeds(pp + i * NV, gm, gam, act, a);
  for (int k = 0; k < 6; ++k) a6[i * 6 + k] = a[k];
}

__global__ void b_rad_ws(long n, const double *pp, const kz_geom *g, const double *tau3, int a0, int a1, int a2,
                         double *a12) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  int act[3] = {a0, a1, a2};
  double t[3] = {tau3[i * 3], tau3[i * 3 + 1], tau3[i * 3 + 2]};
  double a[12];
  rad_wavespeeds(pp + i * NV, gm, t, act, a);
  for (int k = 0; k < 12; ++k) a12[i * 12 + k] = a[k];
}

__global__ void b_rij(long n, const double *pp, const kz_geom *g, double *out21) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  double rij[4][4], ucon[4], ucov[4];
  calc_rij(pp + i * NV, gm, rij);
  double rtt = calc_ff_rtt(pp + i * NV, gm);
  ucon_ucov(pp + i * NV, gm, ucon, ucov);
  double *o = out21 + i * 21;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) o[a * 4 + b] = rij[a][b];
  o[16] = rtt;
  for (int a = 0; a < 4; ++a) o[17 + a] = ucon[a];
}

__global__ void b_radstate(long n, const double *pp, const kz_geom *g, double gam, kz_opac_params op, int full,
                           double *out10) {
  long i = TID;
  if (i >= n) return;
  Geo gm;
  geo_from_kz(g[i], gm);
  RadState st;
  fill_radstate(pp + i * NV, gm, gam, op, full != 0, st);
  double *o = out10 + i * 10;
  o[0] = st.tgas;
  o[1] = st.te;
  o[2] = st.ehat;
  o[3] = st.trad;
  o[4] = st.kappaes;
  o[5] = st.op.gas_abs;
  o[6] = st.op.rad_abs;
  o[7] = st.op.rad_ross;
  o[8] = st.sgas;
  o[9] = calc_chi_slim(pp + i * NV, gm, gam, op);
}

__global__ void b_math(long n, int fn, const double *x, const double *y, double *out) {
  long i = TID;
  if (i >= n) return;
  double a = x[i], b = y ? y[i] : 0.0, r;
  switch (fn) {
    case 0: r = kzm_exp(a); break;
    case 1: r = kzm_log(a); break;
    case 2: r = kzm_pow(a, b); break;
    case 3: r = kzm_cbrt(a); break;
    case 4: r = kzm_log1p(a); break;
    case 5: r = kzm_sqrt(a); break;
    default: r = a / b; break;
  }
  out[i] = r;
}

// ---------------------------------------------------------------------------
// host staging

namespace {

struct Stage {
  void *ptrs[12];
  int n = 0;
  bool bad = false;
  ~Stage() {
    for (int i = 0; i < n; ++i) (void)hipFree(ptrs[i]);
  }
  template <class T>
  T *in(const T *h, size_t count) {
    T *d = alloc<T>(count);
    if (d && h && hipMemcpy(d, h, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) bad = true;
    return d;
  }
  template <class T>
  T *alloc(size_t count) {
    void *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
      bad = true;
      return nullptr;
    }
    (void)hipMemset(p, 0, count * sizeof(T));
    ptrs[n++] = p;
    return (T *)p;
  }
  template <class T>
  void out(T *h, const T *d, size_t count) {
    if (!bad && h && hipMemcpy(h, d, count * sizeof(T), hipMemcpyDeviceToHost) != hipSuccess) bad = true;
  }
  int finish_launch() {
    if (bad) return KZ_E_HIP;
    if (hipGetLastError() != hipSuccess) return KZ_E_HIP;
    if (hipDeviceSynchronize() != hipSuccess) return KZ_E_HIP;
    return 0;
  }
};

inline unsigned nb(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

extern "C" {

int kz_p2u_batch(int64_t n, const double *pp, const kz_geom *g, double gam, double *uu) {
  if (n <= 0) return 0;
  Stage s;
  double *dp = s.in(pp, n * NV), *du = s.alloc<double>(n * NV);
  kz_geom *dg = s.in(g, n);
  if (s.bad) return KZ_E_HIP;
  b_p2u<<<nb(n, 128), 128>>>(n, dp, dg, gam, du);
  if (int e = s.finish_launch()) return e;
  s.out(uu, du, n * NV);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_u2p_mhd_batch(int64_t n, const double *uu, double *pp, const kz_geom *g, double gam, const kz_floor_params *f,
                     kz_u2p_result *r) {
  if (n <= 0) return 0;
  Stage s;
  double *du = s.in(uu, n * NV), *dp = s.in(pp, n * NV);
  kz_geom *dg = s.in(g, n);
  kz_u2p_result *dr = s.alloc<kz_u2p_result>(n);
  if (s.bad) return KZ_E_HIP;
  b_u2p_mhd<<<nb(n, 64), 64>>>(n, du, dp, dg, gam, *f, dr);
  if (int e = s.finish_launch()) return e;
  s.out(pp, dp, n * NV);
  s.out(r, dr, n);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_check_floors_mhd_batch(int64_t n, double *pp, const kz_geom *g, double gam, const kz_floor_params *f,
                              int32_t *ret, uint64_t *nguard) {
  if (n <= 0) return 0;
  Stage s;
  double *dp = s.in(pp, n * NV);
  kz_geom *dg = s.in(g, n);
  int32_t *dr = s.alloc<int32_t>(n);
  unsigned long long *dn = s.alloc<unsigned long long>(1);
  if (s.bad) return KZ_E_HIP;
  if (f->b2rhofloorframe < 0 || f->b2rhofloorframe > 1) return KZ_E_INVALID_CONFIG;
  if (f->b2rhofloorframe == 1) b_check_floors_mhd<true><<<nb(n, 64), 64>>>(n, dp, dg, gam, *f, dr, dn);
  else b_check_floors_mhd<false><<<nb(n, 64), 64>>>(n, dp, dg, gam, *f, dr, dn);
  if (int e = s.finish_launch()) return e;
  s.out(pp, dp, n * NV);
  s.out(ret, dr, n);
  unsigned long long hn = 0;
  s.out(&hn, dn, 1);
  if (nguard) *nguard = hn;
  return s.bad ? KZ_E_HIP : 0;
}

int kz_u2p_rad_batch(int64_t n, const double *uu, double *pp, const kz_geom *g, const kz_rad_params *rp,
                     kz_u2prad_result *r) {
  if (n <= 0) return 0;
  Stage s;
  double *du = s.in(uu, n * NV), *dp = s.in(pp, n * NV);
  kz_geom *dg = s.in(g, n);
  kz_u2prad_result *dr = s.alloc<kz_u2prad_result>(n);
  if (s.bad) return KZ_E_HIP;
  b_u2p_rad<<<nb(n, 64), 64>>>(n, du, dp, dg, *rp, dr);
  if (int e = s.finish_launch()) return e;
  s.out(pp, dp, n * NV);
  s.out(r, dr, n);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_check_floors_rad_batch(int64_t n, double *pp, const kz_geom *g, const kz_rad_params *rp, int32_t *ret) {
  if (n <= 0) return 0;
  Stage s;
  double *dp = s.in(pp, n * NV);
  kz_geom *dg = s.in(g, n);
  int32_t *dr = s.alloc<int32_t>(n);
  if (s.bad) return KZ_E_HIP;
  b_check_floors_rad<<<nb(n, 64), 64>>>(n, dp, dg, *rp, dr);
  if (int e = s.finish_launch()) return e;
  s.out(pp, dp, n * NV);
  s.out(ret, dr, n);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_solve_implicit_lab_batch(int64_t n, double *uu, double *pp, const kz_geom *g, const double *dt, double gam,
                                const kz_rad_params *rp, const kz_opac_params *op, const kz_implicit_params *ip,
                                kz_implicit_result *r) {
  if (n <= 0) return 0;
  Stage s;
  double *du = s.in(uu, n * NV), *dp = s.in(pp, n * NV), *ddt = s.in(dt, n);
  kz_geom *dg = s.in(g, n);
  kz_implicit_result *dr = s.alloc<kz_implicit_result>(n);
  if (s.bad) return KZ_E_HIP;
  b_implicit<<<nb(n, 64), 64>>>(n, du, dp, dg, ddt, gam, *rp, *op, *ip, dr);
  if (int e = s.finish_launch()) return e;
  s.out(uu, du, n * NV);
  s.out(pp, dp, n * NV);
  s.out(r, dr, n);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_reconstruct_batch(int64_t n, const int32_t *order, const double *theta, const double *u5, const double *dx5,
                         double *l, double *r) {
  if (n <= 0) return 0;
  Stage s;
  int32_t *dor = s.in(order, n);
  double *dth = s.in(theta, n), *du = s.in(u5, n * 5), *dd = s.in(dx5, n * 5);
  double *dl = s.alloc<double>(n), *dr = s.alloc<double>(n);
  if (s.bad) return KZ_E_HIP;
  b_reconstruct<<<nb(n, 256), 256>>>(n, dor, dth, du, dd, dl, dr);
  if (int e = s.finish_launch()) return e;
  s.out(l, dl, n);
  s.out(r, dr, n);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_u2p_solver_w_batch(int64_t n, const double *uu, double *pp, const kz_geom *g, int32_t etype, double gam,
                          int32_t *ret) {
  if (n <= 0) return 0;
  if (etype != 0 && etype != 1) return KZ_E_INVALID_CONFIG;
  Stage s;
  double *du = s.in(uu, n * NV), *dp = s.in(pp, n * NV);
  kz_geom *dg = s.in(g, n);
  int32_t *dr = s.alloc<int32_t>(n);
  if (s.bad) return KZ_E_HIP;
  b_u2p_solver_w<<<nb(n, 64), 64>>>(n, du, dp, dg, etype, gam, dr);
  if (int e = s.finish_launch()) return e;
  s.out(pp, dp, n * NV);
  s.out(ret, dr, n);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_flux_prime_batch(int64_t n, const double *pp, const int32_t *idim, const kz_geom *g, double gam, double *ff,
                        int32_t *err) {
  if (n <= 0) return 0;
  for (int64_t i = 0; i < n; ++i)
    if (idim[i] < 0 || idim[i] > 2) return KZ_E_INVALID_CONFIG;
  Stage s;
  double *dp = s.in(pp, n * NV), *df = s.alloc<double>(n * NV);
  int32_t *di = s.in(idim, n), *de = s.alloc<int32_t>(n);
  kz_geom *dg = s.in(g, n);
  if (s.bad) return KZ_E_HIP;
  b_flux_prime<<<nb(n, 64), 64>>>(n, dp, di, dg, gam, df, de);
  if (int e = s.finish_launch()) return e;
  s.out(ff, df, n * NV);
  s.out(err, de, n);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_gas_wavespeeds_batch(int64_t n, const double *pp, const kz_geom *g, double gam, const int32_t *act,
                            double *a6) {
  if (n <= 0) return 0;
  Stage s;
  double *dp = s.in(pp, n * NV), *da = s.alloc<double>(n * 6);
  kz_geom *dg = s.in(g, n);
  if (s.bad) return KZ_E_HIP;
  b_gas_ws<<<nb(n, 64), 64>>>(n, dp, dg, gam, act[0], act[1], act[2], da);
  if (int e = s.finish_launch()) return e;
  s.out(a6, da, n * 6);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_rad_wavespeeds_batch(int64_t n, const double *pp, const kz_geom *g, const double *tau3, const int32_t *act,
                            double *a12) {
  if (n <= 0) return 0;
  Stage s;
  double *dp = s.in(pp, n * NV), *dt = s.in(tau3, n * 3), *da = s.alloc<double>(n * 12);
  kz_geom *dg = s.in(g, n);
  if (s.bad) return KZ_E_HIP;
  b_rad_ws<<<nb(n, 64), 64>>>(n, dp, dg, dt, act[0], act[1], act[2], da);
  if (int e = s.finish_launch()) return e;
  s.out(a12, da, n * 12);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_rij_batch(int64_t n, const double *pp, const kz_geom *g, double *out21) {
  if (n <= 0) return 0;
  Stage s;
  double *dp = s.in(pp, n * NV), *dout = s.alloc<double>(n * 21);
  kz_geom *dg = s.in(g, n);
  if (s.bad) return KZ_E_HIP;
  b_rij<<<nb(n, 64), 64>>>(n, dp, dg, dout);
  if (int e = s.finish_launch()) return e;
  s.out(out21, dout, n * 21);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_radstate_batch(int64_t n, const double *pp, const kz_geom *g, double gam, const kz_opac_params *op,
                      int32_t full, double *out10) {
  if (n <= 0) return 0;
  Stage s;
  double *dp = s.in(pp, n * NV), *dout = s.alloc<double>(n * 10);
  kz_geom *dg = s.in(g, n);
  if (s.bad) return KZ_E_HIP;
  b_radstate<<<nb(n, 64), 64>>>(n, dp, dg, gam, *op, full, dout);
  if (int e = s.finish_launch()) return e;
  s.out(out10, dout, n * 10);
  return s.bad ? KZ_E_HIP : 0;
}

int kz_math_batch(int64_t n, int32_t fn, const double *x, const double *y, double *out) {
  if (n <= 0) return 0;
  if (fn < 0 || fn > 6) return KZ_E_INVALID_CONFIG;
  Stage s;
  double *dx = s.in(x, n), *dy = y ? s.in(y, n) : nullptr, *dout = s.alloc<double>(n);
  if (s.bad) return KZ_E_HIP;
  b_math<<<nb(n, 256), 256>>>(n, fn, dx, dy, dout);
  if (int e = s.finish_launch()) return e;
  s.out(out, dout, n);
  return s.bad ? KZ_E_HIP : 0;
}

}  // extern "C"
