// moloch_core.hip -- MOLOCH non-hydrostatic timestep on MI355X (gfx950).
//
// Drop-in device implementation of moloch(istep, lprint)
// (nompi/mod_moloch.F90:272-421) behind the C ABI of include/moloch_hip.h.
//
// Device layout.  Every field is a 3-D (or 2-D) array with j fastest, one
// (j,i) plane per level, and a halo of two points on each horizontal side,
// the union of the reference's ga/gb halos (mod_atmosphere.F90:104-162).
// Element (j,i,k) lives at (k-1)*plane + (i+1)*pj + (j+JOFF); JOFF puts j=1
// on a 256-byte boundary (FP64) so a wavefront's 64 lanes read aligned lines.
//
// Periodic halos.  The nompi tree fills halos by copies at fixed points of
// the step (nompi/mod_moloch.F90:514-529, 716-723, 1075-1078, ...).  Here the
// kernel that produces a field writes its boundary values into the halo
// slots as it stores them, so no separate exchange pass reads the field
// again; the copies are taken at the same point of the step as in the
// reference, which reproduces its one-substep-stale Exner halo at the seam
// (MLC params halo_mpi = 0) or the MPI tree's fresh one (halo_mpi = 1).
//
// Precision.  Kernels are templated on the storage type T of the state
// (double for MLC_FP64, float for MLC_FP32).  The Exner function is always
// stored in FP64, and the terms in which hydrostatic balance cancels -- the
// explicit w of the column solve and the pressure-gradient force -- as well
// as the whole tridiagonal elimination are evaluated in FP64.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <string>
#include <vector>

#include "mlc_constants.h"
#include "mlc_internal.h"
#include "moloch_lac.h"
#include "moloch_output.h"

// ------------------------------------------------------------------ errors
static thread_local char g_errbuf[512] = "no error";

int mlc_fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_errbuf, sizeof(g_errbuf), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char *mlc_last_error(void) { return g_errbuf; }

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return mlc_fail(MLC_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                      __LINE__);                                                            \
  } while (0)

namespace mlc {

constexpr int JOFF = 31;
// segmented column solve: level segments per column, levels per segment,
// columns per block
constexpr int SEG_NS = 8, SEG_L = 8, SEG_CW = 32;

struct Geo {
  int nj, ni, nk;
  long pj, plane;
  __host__ __device__ __forceinline__ long at(int j, int i, int k) const {
    return (long)(k - 1) * plane + (long)(i + 1) * pj + (long)(j + JOFF);
  }
  __host__ __device__ __forceinline__ long at2(int j, int i) const {
    return (long)(i + 1) * pj + (long)(j + JOFF);
  }
};

// Element i of an array: with an unsigned index the byte offset is formed in
// 32 bits, so the load or store can take the SGPR-base plus 32-bit-offset
// form (one VGPR per address); the caller keeps every array below 4 GiB.
template <typename X>
__device__ __forceinline__ X ldx(const X *b, unsigned i) {
  return *reinterpret_cast<const X *>(reinterpret_cast<const char *>(b) + i * (unsigned)sizeof(X));
}
template <typename X>
__device__ __forceinline__ X ldx(const X *b, long i) {
  return b[i];
}
template <typename X>
__device__ __forceinline__ void stx(X *b, unsigned i, X v) {
  *reinterpret_cast<X *>(reinterpret_cast<char *>(b) + i * (unsigned)sizeof(X)) = v;
}
template <typename X>
__device__ __forceinline__ void stx(X *b, long i, X v) {
  b[i] = v;
}

// Per-run scalar parameters handed to every kernel.
struct Par {
  double dx, dzita, anu2;
  int rotllr, halo_mpi, fulleq, divfilter;
};

// All device pointers of one run.  T: storage type of the state.
template <typename T>
struct Dev {
  Geo g;
  // state
  T *u, *v, *un, *vn, *w, *qv, *qc, *qi, *qr, *qs;
  T *tvirt, *p, *rho, *qsat, *ux, *vx, *ps;
  // the Exner function, theta and the temperature stay FP64 in both
  // precisions: the hydrostatic balance of the w equation cancels theta
  // times the Exner difference against g dts, which amplifies a rounding of
  // theta by g dts / |w|
  double *pai, *t, *tetav;
  // geometry
  T *fmz, *fmzf, *zeta, *hx, *hy, *mu, *mv, *mx, *rmu, *rmv, *mx2, *coru, *corv;
  const double *gzitak, *gzitakh, *xknu, *ffilt;
  // scratch
  T *s, *deltaw, *zdh, *zds, *zdf, *wx, *qwltot, *qwitot;
  double *wz, *p0;        // advection scratch, sized for an FP64 field
  double *pf, *smd, *tf, *smt;  // filter scratch
  double *thg, *thc;      // tridiagonal elimination (FP64)
  double *diag;           // reduction partials
  // increments of the substep being run (its result minus its start) of u,
  // v, w (level k holds w(k+1)) and pi, written by the sound kernels when
  // not null: the learned corrector's inputs without phase-start copies
  float *iu, *iv, *iw, *ipai;
};

// ------------------------------------------------------------ device math
// pfesat / pfwsat (nompi/mod_moloch.F90:119-181), evaluated in FP64.
__device__ __forceinline__ double d_pfwsat(double t, double p) {
  double td = t - kTzero;
  if (td > 100.0) td = 100.0;
  if (td < -75.0) td = -75.0;
  double esat;
  if (td >= 0.0) {
    esat = 0.611213476e+03 +
           td * (0.444007856e+02 +
                 td * (0.143064234e+01 +
                       td * (0.264461437e-01 +
                             td * (0.305903558e-03 +
                                   td * (0.196237241e-05 +
                                         td * (0.892344772e-08 + td * (-0.373208410e-10 + td * 0.209339997e-13)))))));
  } else {
    esat = 0.611123516e+03 +
           td * (0.503109514e+02 +
                 td * (0.188369801e+01 +
                       td * (0.420547422e-01 +
                             td * (0.614396778e-03 +
                                   td * (0.602780717e-05 +
                                         td * (0.387940929e-07 + td * (0.149436277e-09 + td * 0.262655803e-12)))))));
  }
  double cap = 0.15 * p;
  double es = esat < cap ? esat : cap;
  return kEp2 * (es / (p - es));
}

// superbee limiter in the form of nompi/mod_moloch.F90:993
template <typename R>
__device__ __forceinline__ R waf_b(R r) {
  R m = fmin(R(2) * r, R(1));
  m = fmax(r, m);
  m = fmin(R(2), m);
  return fmax(R(0), m);
}
// Quotient of the advection's limiter ratios and metric coefficients: IEEE
// division in FP64 (the FP64 core and the FP64 riders keep the oracle's
// rounding); in FP32 the numerator times the hardware reciprocal (1 ulp), a
// quotient within 2 ulp of the rounded one in two instructions instead of
// the eleven of the correctly rounded sequence (scale, reciprocal, four
// refinements, fixup).  Denominators here are at least kMinden (1e-30) in
// magnitude, far from the reciprocal's denormal range.
template <typename R>
__device__ __forceinline__ R wdiv(R a, R b) {
  return a / b;
}
template <>
__device__ __forceinline__ float wdiv<float>(float a, float b) {
  return a * __builtin_amdgcn_rcpf(b);
}
template <typename R>
__device__ __forceinline__ R sgn_floor(R z) {
  return copysign(fmax(fabs(z), R(kMinden)), z);
}

// Store a value at (j,i,k) and into the 1-wide face halo slot it feeds
// (faces only, no corners: the pattern of nompi/mod_moloch.F90:716-723).
template <typename F>
__device__ __forceinline__ void store_faces(F *f, const Geo &g, int j, int i, int k, F val) {
  if (j == 1) f[g.at(g.nj + 1, i, k)] = val;
  if (j == g.nj) f[g.at(0, i, k)] = val;
  if (i == 1) f[g.at(j, g.ni + 1, k)] = val;
  if (i == g.ni) f[g.at(j, 0, k)] = val;
}

// ============================================================ halo kernels
// mode 0: faces 1-wide (rows and columns of the interior)
// mode 1: faces 1-wide plus the four corners (:514-529)
// mode 2: 2-wide in j for rows i=1..ni (:1395-1398, :1133-1136)
// mode 3: 2-wide in i for columns j=1..nj (:1399-1402, :1075-1078)
template <typename F>
__global__ void __launch_bounds__(256) k_halo(F *f, Geo g, int mode) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int k = blockIdx.y + 1;
  int nj = g.nj, ni = g.ni;
  if (mode == 0 || mode == 1) {
    if (t >= 1 && t <= ni) {
      f[g.at(0, t, k)] = f[g.at(nj, t, k)];
      f[g.at(nj + 1, t, k)] = f[g.at(1, t, k)];
    }
    if (t >= 1 && t <= nj) {
      f[g.at(t, 0, k)] = f[g.at(t, ni, k)];
      f[g.at(t, ni + 1, k)] = f[g.at(t, 1, k)];
    }
    if (mode == 1 && t == 0) {
      f[g.at(nj + 1, ni + 1, k)] = f[g.at(1, 1, k)];
      f[g.at(0, ni + 1, k)] = f[g.at(nj, 1, k)];
      f[g.at(0, 0, k)] = f[g.at(nj, ni, k)];
      f[g.at(nj + 1, 0, k)] = f[g.at(1, ni, k)];
    }
  } else if (mode == 2) {
    if (t >= 1 && t <= ni) {
      f[g.at(0, t, k)] = f[g.at(nj, t, k)];
      f[g.at(nj + 1, t, k)] = f[g.at(1, t, k)];
      f[g.at(-1, t, k)] = f[g.at(nj - 1, t, k)];
      f[g.at(nj + 2, t, k)] = f[g.at(2, t, k)];
    }
  } else {
    if (t >= 1 && t <= nj) {
      f[g.at(t, 0, k)] = f[g.at(t, ni, k)];
      f[g.at(t, ni + 1, k)] = f[g.at(t, 1, k)];
      f[g.at(t, -1, k)] = f[g.at(t, ni - 1, k)];
      f[g.at(t, ni + 2, k)] = f[g.at(t, 2, k)];
    }
  }
}

// Column kernels use a (64,4) block over (j,i).  (Numbering the tiles so
// that each XCD sweeps a contiguous run of them measured slower: 213 against
// 208 ms per 2048 x 2048 x 60 FP32 step.)
#define COL_IDX                                      \
  const int j = blockIdx.x * 64 + threadIdx.x + 1;   \
  const int i = blockIdx.y * 4 + threadIdx.y + 1;    \
  if (j > D.g.nj || i > D.g.ni) return;              \
  const Geo &g = D.g;                                \
  (void)g;
// COL_IDX for (64, BY) blocks
#define COL_IDX_B(BY)                                \
  const int j = blockIdx.x * 64 + threadIdx.x + 1;   \
  const int i = blockIdx.y * (BY) + threadIdx.y + 1; \
  if (j > D.g.nj || i > D.g.ni) return;              \
  const Geo &g = D.g;                                \
  (void)g;

// ================================================================ prologue
// :291-323 -- p and qsat, tvirt, qwltot/qwitot, tetav; filter copies.
template <typename T>
__global__ void __launch_bounds__(256) k_prologue(Dev<T> D, Par P, int do_pf, int do_tf) {
  COL_IDX
  for (int k = 1; k <= g.nk; k++) {
    long x = g.at(j, i, k);
    double pai = D.pai[x];
    double t = (double)D.t[x];
    double p = pow(pai, kCpovr) * kP00;
    D.qsat[x] = (T)d_pfwsat(t, p);
    T qv = D.qv[x], qc = D.qc[x], qi = D.qi[x], qr = D.qr[x], qs = D.qs[x];
    double tv = t * (1.0 + kEp1 * (double)qv - (double)qc - (double)qi - (double)qr - (double)qs);
    if (P.fulleq) {
      D.qwltot[x] = qc + qr;
      D.qwitot[x] = qi + qs;
    }
    double tet = tv / pai;
    D.tetav[x] = tet;
    if (do_pf) D.pf[x] = pai;
    if (do_tf) D.tf[x] = tet;
  }
}

// ======================================================== sound, phase A
// Surface w (:533-545), s (:547-570) and the horizontal divergence
// (:575-624) from the substep-start winds; zdiv2's face halos are stored
// alongside, as divdamp fills them before first use (:859-862).
template <typename T>
__global__ void __launch_bounds__(256) k_sound_a(Dev<T> D, Par P, double dts) {
  COL_IDX
  const int nk = g.nk;
  const T dtrdx = (T)(dts / P.dx), dtrdy = (T)(dts / P.dx);
  const long c2 = g.at2(j, i);
  const T hx0 = D.hx[c2], hx1 = D.hx[g.at2(j + 1, i)];
  const T hy0 = D.hy[c2], hy1 = D.hy[g.at2(j, i + 1)];
  const T rmv0 = D.rmv[c2], rmv1 = D.rmv[g.at2(j, i + 1)];
  const T rmu0 = D.rmu[c2], rmu1 = D.rmu[g.at2(j + 1, i)];
  const T mx = D.mx[c2], mx2 = D.mx2[c2];
  T u0m = 0, u1m = 0, v0m = 0, v1m = 0;
  for (int k = 1; k <= nk; k++) {
    long x = g.at(j, i, k);
    T u0 = D.u[x], u1 = D.u[x + 1], v0 = D.v[x], v1 = D.v[x + g.pj];
    if (k >= 2) {
      T zuh = (u0 + u0m) * hx0 + (u1 + u1m) * hx1;
      T zvh = (v0 + v0m) * hy0 + (v1 + v1m) * hy1;
      D.s[x] = T(-0.25) * (zuh + zvh) * (T)D.gzitak[k - 1];
    }
    if (k == nk) {
      T zuh = u0 * hx0 + u1 * hx1;
      T zvh = v0 * hy0 + v1 * hy1;
      T wb = T(0.5) * (zuh + zvh);
      if (D.iw) D.iw[g.at(j, i, nk)] = (float)(wb - D.w[g.at(j, i, nk + 1)]);
      D.w[g.at(j, i, nk + 1)] = wb;
      D.s[g.at(j, i, nk + 1)] = -wb;
    }
    T f0 = D.fmz[x];
    T zrfmzum = T(2) / (f0 + D.fmz[x - 1]);
    T zrfmzvm = T(2) / (f0 + D.fmz[x - g.pj]);
    T zrfmzup = T(2) / (f0 + D.fmz[x + 1]);
    T zrfmzvp = T(2) / (f0 + D.fmz[x + g.pj]);
    T zdv;
    if (P.rotllr) {
      T zum = dtrdx * u0 * zrfmzum;
      T zup = dtrdx * u1 * zrfmzup;
      T zvm = dtrdy * v0 * zrfmzvm * rmv0;
      T zvp = dtrdy * v1 * zrfmzvp * rmv1;
      zdv = f0 * mx * ((zup - zum) + (zvp - zvm));
    } else {
      T zum = dtrdx * u0 * rmu0 * zrfmzum;
      T zup = dtrdx * u1 * rmu1 * zrfmzup;
      T zvm = dtrdy * v0 * rmv0 * zrfmzvm;
      T zvp = dtrdy * v1 * rmv1 * zrfmzvp;
      zdv = mx2 * f0 * ((zup - zum) + (zvp - zvm));
    }
    D.zdh[x] = zdv;
    store_faces(D.zdh, g, j, i, k, zdv);
    u0m = u0;
    u1m = u1;
    v0m = v0;
    v1m = v1;
  }
}

// ======================================================== sound, phase B
// divdamp (:853-895), vertical divergence (:627-630), the implicit w solve
// (:634-686), the new Exner function and theta (:690-728), one column per
// lane: the reference's loop order, kept for params.unfused and for kz
// beyond what k_sound_bs holds in registers.  Forward elimination runs bottom-up storing the elimination
// coefficients in FP64 scratch; back substitution runs top-down and
// finishes each level as soon as w above and below it are known.
// split != 0 (divergence filter on): stop after theta, leaving zdiv2 in zds.
template <typename T>
__global__ void __launch_bounds__(256) k_sound_b(Dev<T> D, Par P, double dts, int ns, int split) {
  COL_IDX
  const int nk = g.nk;
  const double dtrdz = dts / P.dzita;
  const double zcs2 = dtrdz * dtrdz * kRdrcv;
  const T ddamp = (T)(0.125 * (P.dx / dts));
  const T ddamp2 = (T)(kXdamp * 0.015625 / dts);
  const T dtrdz_t = (T)dtrdz;
  const long c2 = g.at2(j, i);
  const T mu = D.mu[c2], mv = D.mv[c2];

  // divdamp (u, v) and the smoothed + vertical divergence of level k
  auto level_a = [&](int k) -> T {
    long x = g.at(j, i, k);
    T zc = D.zdh[x];
    T zjm = D.zdh[x - 1], zjp = D.zdh[x + 1], zim = D.zdh[x - g.pj], zip = D.zdh[x + g.pj];
    T xk = (T)D.xknu[k - 1];
    D.u[x] = D.u[x] + xk * ddamp * mu * (zc - zjm);
    if (P.rotllr)
      D.v[x] = D.v[x] + xk * ddamp * (zc - zim);
    else
      D.v[x] = D.v[x] + xk * ddamp * mv * (zc - zim);
    T p2d = T(0.125) * (zjm + zjp + zim + zip) - T(0.5) * zc;
    T z = zc + ddamp2 * p2d;
    z = z + D.fmz[x] * dtrdz_t * (D.s[x] - D.s[x + g.plane]);
    D.zds[x] = z;
    return z;
  };

  // ---- forward elimination, k = nk .. 2
  double gk1 = (double)D.w[g.at(j, i, nk + 1)];  // w(kz+1) from phase A
  double ck1 = 0.0;                                // wwkw(kzp1) = 0
  long xk = g.at(j, i, nk);
  double zk = (double)level_a(nk);
  double tk = (double)D.tetav[xk], pk = D.pai[xk], fzk = (double)D.fmz[xk];
  double qsk = (double)D.qsat[xk];
  for (int k = nk; k >= 2; k--) {
    long xm = xk - g.plane;  // level k-1
    double zkm = (double)level_a(k - 1);
    double tkm = (double)D.tetav[xm], pkm = D.pai[xm], fzkm = (double)D.fmz[xm];
    double qskm = (double)D.qsat[xm];
    double wk = (double)D.w[xk];
    double ffk = (double)D.fmzf[xk];
    double zrom1w = 0.5 * kCpd * ffk * (tkm + tk);
    zrom1w = zrom1w - kCpd * wk * ffk * ffk * (double)ns * dtrdz * (tkm - tk);
    if ((double)D.qv[xk] > 0.96 * qsk && wk > 0.1) {
      double zqs = 0.5 * (qsk + qskm);
      double tm = (double)D.t[xm];
      double zdth = kEgrav * wk * (double)(ns - 1) * dts * kWlhv * kWlhv * zqs / (kCpd * pkm * kRwat * tm * tm);
      zrom1w = zrom1w + zdth * ffk;
    }
    double zwexpl = wk - zrom1w * dtrdz * (pkm - pk) - kEgrav * dts;
    zwexpl = zwexpl + kRdrcv * zrom1w * dtrdz * (pkm * zkm - pk * zk);
    double ff = D.ffilt[k - 1];
    double zu = zcs2 * fzkm * zrom1w * pkm + ff;
    double zd = zcs2 * fzk * zrom1w * pk + ff;
    double zrapp = 1.0 / (1.0 + zd + zu - zd * ck1);
    double gk = zrapp * (zwexpl + zd * gk1);
    double ck = zrapp * zu;
    D.thg[xk] = gk;
    D.thc[xk] = ck;
    gk1 = gk;
    ck1 = ck;
    xk = xm;
    zk = zkm;
    tk = tkm;
    pk = pkm;
    fzk = fzkm;
    qsk = qskm;
  }

  // ---- back substitution, k = 2 .. nk, finishing level k-1 at step k
  auto finish = [&](int k, double wtop, double wbot) {
    long x = g.at(j, i, k);
    T zf = D.zds[x] + dtrdz_t * D.fmz[x] * ((T)wtop - (T)wbot);
    T qv = D.qv[x], qc = D.qc[x];
    if (P.fulleq) {
      zf = zf * (T(1) + T(0.86) * qv + T(3.2) * qc) / (T(1) + T(0.96) * qv + T(4.8) * qc);
      // the increment in the storage type, 1 + increment and the product in FP64
      double tet = D.tetav[x] *
                   (1.0 + (double)((T)kRdrcv * zf * (T(0.25) * qv + T(4.2) * D.qwltot[x] + T(2.1) * D.qwitot[x])));
      D.tetav[x] = tet;
      store_faces(D.tetav, g, j, i, k, tet);
    }
    if (split) {
      D.zds[x] = zf;
      return;
    }
    double po = D.pai[x];
    if (!P.halo_mpi) store_faces(D.pai, g, j, i, k, po);
    double pn = po * (1.0 - kRdrcv * (double)zf);
    D.pai[x] = pn;
    if (P.halo_mpi) store_faces(D.pai, g, j, i, k, pn);
  };

  double wprev = (double)D.w[g.at(j, i, 1)];  // w(1) = 0
  T wprev_t = D.w[g.at(j, i, 1)];
  for (int k = 2; k <= nk; k++) {
    long x = g.at(j, i, k);
    double wn = D.thg[x] + D.thc[x] * wprev;
    T wo = D.w[x];
    T wn_t = (T)wn;
    D.w[x] = wn_t;
    T dw = -wo + wn_t;
    D.deltaw[x] = dw;
    store_faces(D.deltaw, g, j, i, k, dw);
    finish(k - 1, (double)wprev_t, (double)wn_t);
    wprev = wn;
    wprev_t = wn_t;
  }
  finish(nk, (double)wprev_t, (double)D.w[g.at(j, i, nk + 1)]);
}

// ============================================ sound, phase B (segmented)
// The arithmetic of k_sound_b without divdamp's wind update, which
// k_sound_c<T, true> applies as it reads the winds.  Each column's unknown
// levels 2..kz are cut into NS contiguous segments of at most L levels
// (segment 0, the top one, also takes level 1); a block holds CW columns of
// one row and one lane per (column, segment), 64 / CW segments per wave.
// Each lane keeps what the column solve and the theta/pi update need for its
// levels on chip: in registers the explicit right-hand side and the two
// coupling coefficients (FP64), the divergence and pi; in LDS theta, the old
// w, qv and fmz.  The forward elimination (bottom segment first) and the back
// substitution (top segment first) pass through the segments in turn,
// handing the (g, c) pair and then w across segment boundaries through LDS,
// one block barrier per segment: the operations and their order are those of
// the one-lane sweep of k_sound_b, so the results are bit-identical, while
// neither the elimination coefficients nor the divergence go through HBM.
// Needs 2 <= kz <= NS*L + 1; nbj = number of CW-column tiles per row.
template <typename T, int NS, int L, int CW>
__global__ void __launch_bounds__(NS * CW) k_sound_bs(Dev<T> D, Par P, double dts, int ns, int split, int nbj) {
  static_assert(64 % CW == 0 && (NS * CW) % 64 == 0, "whole waves of whole columns");
  __shared__ double hg[NS][CW], hc[NS][CW], hw[NS][CW], hw1[NS][CW];
  // theta, old w, qv and fmz of each lane's levels, kept from the first
  // phase to the last (in registers they would cost occupancy, read again
  // from memory they would cost traffic)
  __shared__ T stash[3][L][NS * CW];  // old w, qv, fmz
  __shared__ double sth[L][NS * CW];
  const int tl = threadIdx.y * 64 + threadIdx.x;
  const Geo &g = D.g;
  const int lane = threadIdx.x % CW;                           // column within the block
  const int wv = threadIdx.y * (64 / CW) + threadIdx.x / CW;  // segment
  // Workgroups are dispatched to the 8 XCDs round robin; number the tiles so
  // that each XCD sweeps a contiguous band of rows (the i +- 1 divergence
  // reads then meet rows its own L2 has seen).
  const int nblk = gridDim.x, lin = blockIdx.x;
  const int per = nblk >> 3, rem = nblk & 7, xcd = lin & 7;
  const int tile = (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + (lin >> 3);
  const int i = tile / nbj + 1;
  const int jr = (tile % nbj) * CW + lane + 1;
  const bool act = jr <= g.nj;
  const int j = act ? jr : g.nj;  // idle lanes read a valid column and store nothing
  const int nk = g.nk;
  const int nu = nk - 1, base = nu / NS, rm = nu % NS;
  const int cnt = base + (wv < rm ? 1 : 0);            // empty segments only at the bottom
  const int a = 2 + wv * base + (wv < rm ? wv : rm);  // levels a .. a+cnt-1
  const int bl = a + cnt - 1;
  const double dtrdz = dts / P.dzita;
  const double zcs2 = dtrdz * dtrdz * kRdrcv;
  const T ddamp2 = (T)(kXdamp * 0.015625 / dts);
  const T dtrdz_t = (T)dtrdz;
  const long surf = g.at(j, i, nk + 1);

  // smoothed + vertical divergence of the level at x (level_a of k_sound_b)
  auto level_z = [&](long x) -> T {
    T zc = D.zdh[x];
    T zjm = D.zdh[x - 1], zjp = D.zdh[x + 1], zim = D.zdh[x - g.pj], zip = D.zdh[x + g.pj];
    T p2d = T(0.125) * (zjm + zjp + zim + zip) - T(0.5) * zc;
    T z = zc + ddamp2 * p2d;
    z = z + D.fmz[x] * dtrdz_t * (D.s[x] - D.s[x + g.plane]);
    return z;
  };

  double rz[L], ru[L], rd[L];  // zwexpl, zu, zd; then g, c; then the new w
  T zz[L];
  double pp[L];
  // level a-1 (the segment's upper neighbour; level 1 for wave 0)
  T zt = 0, fzt = 0;
  double ppt = 0, tht = 0;
  if (cnt > 0) {
    {
      const long x = g.at(j, i, a - 1);
      zt = level_z(x);
      tht = D.tetav[x];
      ppt = D.pai[x];
      fzt = D.fmz[x];
    }
    double zkm = (double)zt, tkm = (double)tht, pkm = ppt, fzkm = (double)fzt;
    double qskm = (double)D.qsat[g.at(j, i, a - 1)];
#pragma unroll
    for (int l = 0; l < L; l++) {
      const int k = a + l <= bl ? a + l : bl;  // past the segment: repeat its last level
      const long xk = g.at(j, i, k), xm = xk - g.plane;
      const T zkt = level_z(xk);
      zz[l] = zkt;
      const double tkt = D.tetav[xk];
      const T wkt = D.w[xk], qvt = D.qv[xk], fzt_k = D.fmz[xk];
      sth[l][tl] = tkt;
      stash[0][l][tl] = wkt;
      stash[1][l][tl] = qvt;
      stash[2][l][tl] = fzt_k;
      pp[l] = D.pai[xk];
      const double zk = (double)zkt, tk = (double)tkt, pk = pp[l], fzk = (double)fzt_k;
      const double qsk = (double)D.qsat[xk];
      const double wk = (double)wkt;
      const double ffk = (double)D.fmzf[xk];
      double zrom1w = 0.5 * kCpd * ffk * (tkm + tk);
      zrom1w = zrom1w - kCpd * wk * ffk * ffk * (double)ns * dtrdz * (tkm - tk);
      if ((double)qvt > 0.96 * qsk && wk > 0.1) {
        double zqs = 0.5 * (qsk + qskm);
        double tm = (double)D.t[xm];
        double zdth = kEgrav * wk * (double)(ns - 1) * dts * kWlhv * kWlhv * zqs / (kCpd * pkm * kRwat * tm * tm);
        zrom1w = zrom1w + zdth * ffk;
      }
      double zwexpl = wk - zrom1w * dtrdz * (pkm - pk) - kEgrav * dts;
      zwexpl = zwexpl + kRdrcv * zrom1w * dtrdz * (pkm * zkm - pk * zk);
      double ff = D.ffilt[k - 1];
      rz[l] = zwexpl;
      ru[l] = zcs2 * fzkm * zrom1w * pkm + ff;
      rd[l] = zcs2 * fzk * zrom1w * pk + ff;
      zkm = zk;
      tkm = tk;
      pkm = pk;
      fzkm = fzk;
      qskm = qsk;
    }
  }

  // forward elimination, bottom segment first: g_k, c_k from g_{k+1}, c_{k+1}
  for (int st = NS - 1; st >= 0; st--) {
    if (wv == st) {
      double g1, c1;
      if (st == NS - 1) {
        g1 = (double)D.w[surf];  // w(kz+1) from phase A
        c1 = 0.0;                 // wwkw(kzp1) = 0
      } else {
        g1 = hg[st + 1][lane];
        c1 = hc[st + 1][lane];
      }
#pragma unroll
      for (int l = L - 1; l >= 0; l--) {
        if (l < cnt) {
          double zrapp = 1.0 / (1.0 + rd[l] + ru[l] - rd[l] * c1);
          double gk = zrapp * (rz[l] + rd[l] * g1);
          double ck = zrapp * ru[l];
          rz[l] = gk;
          ru[l] = ck;
          g1 = gk;
          c1 = ck;
        }
      }
      hg[st][lane] = g1;
      hc[st][lane] = c1;
    }
    __syncthreads();
  }
  // back substitution, top segment first
  const T w1t = D.w[g.at(j, i, 1)];  // w(1) = 0
  for (int st = 0; st < NS; st++) {
    if (wv == st) {
      double wp = st == 0 ? (double)w1t : hw[st - 1][lane];
#pragma unroll
      for (int l = 0; l < L; l++) {
        if (l < cnt) {
          double wn = rz[l] + ru[l] * wp;
          rz[l] = wn;
          wp = wn;
        }
      }
      hw[st][lane] = wp;
      hw1[st][lane] = cnt > 0 ? rz[0] : wp;
    }
    __syncthreads();
  }
  if (cnt == 0) return;

  // w, deltaw; theta and pi of each level once w above and below are known
  auto finish = [&](long x, int k, T zk, T fzk, double thk, T qv, double po, T wtop, T wbot) {
    T zf = zk + dtrdz_t * fzk * (wtop - wbot);
    T qc = D.qc[x];
    if (P.fulleq) {
      zf = zf * (T(1) + T(0.86) * qv + T(3.2) * qc) / (T(1) + T(0.96) * qv + T(4.8) * qc);
      // the increment in the storage type, 1 + increment and the product in FP64
      double tet =
          thk * (1.0 + (double)((T)kRdrcv * zf * (T(0.25) * qv + T(4.2) * D.qwltot[x] + T(2.1) * D.qwitot[x])));
      D.tetav[x] = tet;
      store_faces(D.tetav, g, j, i, k, tet);
    }
    if (split) {
      D.zds[x] = zf;
      return;
    }
    if (!P.halo_mpi) store_faces(D.pai, g, j, i, k, po);
    double pn = po * (1.0 - kRdrcv * (double)zf);
    D.pai[x] = pn;
    if (P.halo_mpi) store_faces(D.pai, g, j, i, k, pn);
    if (D.ipai) D.ipai[x] = (float)(pn - po);
  };
  if (!act) return;
  const T wnext = bl == nk ? D.w[surf] : (T)hw1[wv + 1 < NS ? wv + 1 : wv][lane];
  if (wv == 0) finish(g.at(j, i, 1), 1, zt, fzt, tht, D.qv[g.at(j, i, 1)], ppt, w1t, (T)rz[0]);
#pragma unroll
  for (int l = 0; l < L; l++) {
    if (l < cnt) {
      const int k = a + l;
      const long x = g.at(j, i, k);
      const T wn_t = (T)rz[l];
      T dw = -stash[0][l][tl] + wn_t;
      D.w[x] = wn_t;
      D.deltaw[x] = dw;
      if (D.iw) D.iw[x - g.plane] = (float)dw;
      store_faces(D.deltaw, g, j, i, k, dw);
      const T wb = l + 1 < cnt ? (T)rz[l + 1 < L ? l + 1 : l] : wnext;
      finish(x, k, zz[l], stash[2][l][tl], sth[l][tl], stash[1][l][tl], pp[l], wn_t, wb);
    }
  }
}

// Exner update after the divergence filter (split path): :716-728
template <typename T>
__global__ void __launch_bounds__(256) k_sound_b2(Dev<T> D, Par P) {
  COL_IDX
  for (int k = 1; k <= g.nk; k++) {
    long x = g.at(j, i, k);
    double po = D.pai[x];
    if (!P.halo_mpi) store_faces(D.pai, g, j, i, k, po);
    double pn = po * (1.0 - kRdrcv * (double)D.zdf[x]);
    D.pai[x] = pn;
    if (P.halo_mpi) store_faces(D.pai, g, j, i, k, pn);
  }
}

// ======================================================== sound, phase C
// Horizontal momentum, Equations 17 and 18 (:730-838); results go to un/vn,
// evaluated in FP64.  The winds read are the ud/vd snapshots of :730-735:
// the post-divdamp interior and the substep-start halos.  DAMP = false: u
// and v already hold them (k_sound_b applied divdamp).  DAMP = true: u and v
// hold the substep-start winds and each interior value read gets divdamp's
// increment (:865-872) here, from the horizontal divergence zdiv2 of phase
// A, in the operand order of k_sound_b, so the winds agree bit for bit.
// BY: rows per (64, BY) block (16 for the FP32 state: the rows above and
// below a block are read once per 16 rows instead of once per 4, 2.46
// against 2.73 ms per launch at 2048 x 2048 x 60; the FP64 kernel needs more
// than the 128 VGPRs of a 1024-lane block and keeps 4).
template <typename T, bool DAMP, typename I, int BY>
__global__ void __launch_bounds__(64 * BY) k_sound_c(Dev<T> D, Par P, double dts) {
  COL_IDX_B(BY)
  const double dtrdx = dts / P.dx, dtrdy = dts / P.dx;
  // element indices in I (unsigned: 32-bit byte offsets from the SGPR bases)
  const I c2 = (I)g.at2(j, i);
  const double zcx = dtrdx * (double)ldx(D.mu, c2);
  const double zcy = P.rotllr ? dtrdy : dtrdy * (double)ldx(D.mv, c2);
  const double hx = (double)ldx(D.hx, c2), hy = (double)ldx(D.hy, c2);
  const double coru = (double)ldx(D.coru, c2), corv = (double)ldx(D.corv, c2);
  const I pj = (I)g.pj, plane = (I)g.plane;
  // divdamp of the winds read (DAMP): the points are u(j,i), u(j,i-1),
  // u(j+1,i), u(j+1,i-1) and v(j,i), v(j-1,i), v(j-1,i+1), v(j,i+1); each
  // gets xknu*ddamp*m*(zdiv2 difference) if it is an interior point.  Every
  // load below lies within the two-point halo, so all are unconditional and
  // the interior tests (column constants) only select.
  const T ddamp = (T)(0.125 * (P.dx / dts));
  const bool jlo = j >= 2, jhi = j <= g.nj - 1, ilo = i >= 2, ihi = i <= g.ni - 1;
  T mu_c = 0, mu_im = 0, mu_jp = 0, mu_jpim = 0, mv_c = 0, mv_jm = 0, mv_jmip = 0, mv_ip = 0;
  if (DAMP) {
    mu_c = ldx(D.mu, c2);
    mu_im = ldx(D.mu, (I)(c2 - pj));
    mu_jp = ldx(D.mu, (I)(c2 + 1));
    mu_jpim = ldx(D.mu, (I)(c2 + 1 - pj));
    mv_c = ldx(D.mv, c2);
    mv_jm = ldx(D.mv, (I)(c2 - 1));
    mv_jmip = ldx(D.mv, (I)(c2 - 1 + pj));
    mv_ip = ldx(D.mv, (I)(c2 + pj));
  }
  I x = (I)g.at(j, i, 1);
  double dwc = (double)ldx(D.deltaw, x), dwjm = (double)ldx(D.deltaw, (I)(x - 1)), dwim = (double)ldx(D.deltaw, (I)(x - pj));
  for (int k = 1; k <= g.nk; k++, x += plane) {
    I xp = x + plane;
    double dwc1 = (double)ldx(D.deltaw, xp), dwjm1 = (double)ldx(D.deltaw, (I)(xp - 1)), dwim1 = (double)ldx(D.deltaw, (I)(xp - pj));
    double tc = (double)ldx(D.tetav, x);
    double pc = ldx(D.pai, x);
    double gzh = D.gzitakh[k - 1];
    T u_c = ldx(D.u, x), u_im = ldx(D.u, (I)(x - pj)), u_jp = ldx(D.u, (I)(x + 1)), u_jpim = ldx(D.u, (I)(x + 1 - pj));
    T v_c = ldx(D.v, x), v_jm = ldx(D.v, (I)(x - 1)), v_jmip = ldx(D.v, (I)(x - 1 + pj)), v_ip = ldx(D.v, (I)(x + pj));
    const T u_start = u_c, v_start = v_c;
    if (DAMP) {
      const T xk = (T)D.xknu[k - 1];
      const T z00 = ldx(D.zdh, x), zm0 = ldx(D.zdh, (I)(x - 1)), zp0 = ldx(D.zdh, (I)(x + 1));
      const T z0m = ldx(D.zdh, (I)(x - pj)), zmm = ldx(D.zdh, (I)(x - 1 - pj)), zpm = ldx(D.zdh, (I)(x + 1 - pj));
      const T z0p = ldx(D.zdh, (I)(x + pj)), zmp = ldx(D.zdh, (I)(x - 1 + pj));
      // the operand order of k_sound_b's level update
      const T du_c = u_c + xk * ddamp * mu_c * (z00 - zm0);
      const T du_im = u_im + xk * ddamp * mu_im * (z0m - zmm);
      const T du_jp = u_jp + xk * ddamp * mu_jp * (zp0 - z00);
      const T du_jpim = u_jpim + xk * ddamp * mu_jpim * (zpm - z0m);
      T dv_c, dv_jm, dv_jmip, dv_ip;
      if (P.rotllr) {
        dv_c = v_c + xk * ddamp * (z00 - z0m);
        dv_jm = v_jm + xk * ddamp * (zm0 - zmm);
        dv_jmip = v_jmip + xk * ddamp * (zmp - zm0);
        dv_ip = v_ip + xk * ddamp * (z0p - z00);
      } else {
        dv_c = v_c + xk * ddamp * mv_c * (z00 - z0m);
        dv_jm = v_jm + xk * ddamp * mv_jm * (zm0 - zmm);
        dv_jmip = v_jmip + xk * ddamp * mv_jmip * (zmp - zm0);
        dv_ip = v_ip + xk * ddamp * mv_ip * (z0p - z00);
      }
      u_c = du_c;
      u_im = ilo ? du_im : u_im;
      u_jp = jhi ? du_jp : u_jp;
      u_jpim = (jhi && ilo) ? du_jpim : u_jpim;
      v_c = dv_c;
      v_jm = jlo ? dv_jm : v_jm;
      v_jmip = (jlo && ihi) ? dv_jmip : v_jmip;
      v_ip = ihi ? dv_ip : v_ip;
    }
    // Equation 17
    {
      double zfz = kEgrav * dts + 0.25 * (dwjm + dwjm1 + dwc + dwc1);
      double zrom1u = 0.5 * kCpd * ((double)ldx(D.tetav, (I)(x - 1)) + tc);
      double zcor1u = coru * dts * 0.25 * ((double)v_c + (double)v_jm + (double)v_jmip + (double)v_ip);
      double un = (double)u_c + zcor1u - zfz * hx * gzh - zcx * zrom1u * (pc - ldx(D.pai, (I)(x - 1)));
      stx(D.un, x, (T)un);
      if (DAMP && D.iu) stx(D.iu, x, (float)((T)un - u_start));
    }
    // Equation 18
    {
      double zfz = kEgrav * dts + 0.25 * (dwim + dwim1 + dwc + dwc1);
      double zrom1v = 0.5 * kCpd * ((double)ldx(D.tetav, (I)(x - pj)) + tc);
      double zcor1v = corv * dts * 0.25 * ((double)u_c + (double)u_im + (double)u_jp + (double)u_jpim);
      double vn = (double)v_c - zcor1v - zfz * hy * gzh - zcy * zrom1v * (pc - ldx(D.pai, (I)(x - pj)));
      stx(D.vn, x, (T)vn);
      if (DAMP && D.iv) stx(D.iv, x, (float)((T)vn - v_start));
    }
    dwc = dwc1;
    dwjm = dwjm1;
    dwim = dwim1;
  }
}

// =========================================================== 5-point smooth
// dst = src + coef_k * (1/8*sum of 4 neighbours - 1/2*src), interior only:
// divergence_filter (:423-438), filtpai (:440-460), filttheta (:462-481).
template <typename F>
__global__ void __launch_bounds__(256) k_smooth5(const F *src, F *dst, Geo g, const double *xknu, double coef) {
  const int j = blockIdx.x * 64 + threadIdx.x + 1;
  const int i = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (j > g.nj || i > g.ni) return;
  long x = g.at(j, i, k);
  F c = src[x];
  F p2d = F(0.125) * (src[x - 1] + src[x + 1] + src[x - g.pj] + src[x + g.pj]) - F(0.5) * c;
  F cf = xknu ? (F)(coef * xknu[k - 1]) : (F)coef;
  dst[x] = c + cf * p2d;
}

// a = a - b (interior) / a = s + b (interior); filter bracket of :333-353
template <typename F, typename G>
__global__ void __launch_bounds__(256) k_sub(F *a, const G *b, Geo g) {
  const int j = blockIdx.x * 64 + threadIdx.x + 1;
  const int i = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (j > g.nj || i > g.ni) return;
  long x = g.at(j, i, k);
  a[x] = a[x] - (F)b[x];
}
template <typename F, typename G>
__global__ void __launch_bounds__(256) k_addto(F *a, const F *s, const G *b, Geo g) {
  const int j = blockIdx.x * 64 + threadIdx.x + 1;
  const int i = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (j > g.nj || i > g.ni) return;
  long x = g.at(j, i, k);
  a[x] = s[x] + (F)b[x];
}

// ======================================================= advection set-up
// Completion of Equation 10 (:843-849), wstagtox (:1327-1341) and the
// cross-point winds of uvstagtox (:1404-1416).  u and v carry 2-wide halos.
template <typename T>
__global__ void __launch_bounds__(256) k_adv_in(Dev<T> D) {
  COL_IDX
  const int nk = g.nk;
  const long pl = g.plane;
  // one pass down the column, w in a window of levels k-1 .. k+2
  const long x1 = g.at(j, i, 1);
  T wm = D.w[x1], w0 = wm, w1 = D.w[x1 + pl], w2 = nk >= 2 ? D.w[x1 + 2 * pl] : w1;
  for (int k = 1; k <= nk; k++) {
    const long x = x1 + (long)(k - 1) * pl;
    if (k >= 2) D.s[x] = (w0 + D.s[x]) * D.fmzf[x];
    if (k == 1)
      D.wx[x] = T(0.5) * (w1 + w0);
    else if (k == nk)
      D.wx[x] = T(0.5) * (w1 + w0);
    else
      D.wx[x] = T(0.5625) * (w1 + w0) - T(0.0625) * (w2 + wm);
    D.ux[x] = T(0.5625) * (D.u[x + 1] + D.u[x]) - T(0.0625) * (D.u[x + 2] + D.u[x - 1]);
    D.vx[x] = T(0.5625) * (D.v[x + g.pj] + D.v[x]) - T(0.0625) * (D.v[x + 2 * g.pj] + D.v[x - g.pj]);
    wm = w0;
    w0 = w1;
    w1 = w2;
    if (k + 3 <= nk + 1) w2 = D.w[x + 3 * pl];
  }
  D.s[x1] = T(0);
  D.s[g.at(j, i, nk + 1)] = T(0);
}

// xtouvstag (:1374-1386) and xtowstag (:1343-1357); ux/vx carry halos.
// kz >= 3 (mlc_create).
template <typename T>
__global__ void __launch_bounds__(256) k_adv_out(Dev<T> D) {
  COL_IDX
  const int nk = g.nk;
  const long pl = g.plane;
  // one pass down the column
  for (int k = 1; k <= nk; k++) {
    const long x = g.at(j, i, k);
    D.u[x] = T(0.5625) * (D.ux[x] + D.ux[x - 1]) - T(0.0625) * (D.ux[x + 1] + D.ux[x - 2]);
    D.v[x] = T(0.5625) * (D.vx[x] + D.vx[x - g.pj]) - T(0.0625) * (D.vx[x + g.pj] + D.vx[x - 2 * g.pj]);
    if (k == 2) D.w[x] = T(0.5) * (D.wx[x] + D.wx[x - pl]);
    if (k >= 3 && k <= nk - 1) D.w[x] = T(0.5625) * (D.wx[x] + D.wx[x - pl]) - T(0.0625) * (D.wx[x + pl] + D.wx[x - 2 * pl]);
    if (k == nk && k >= 3) D.w[x] = T(0.5) * (D.wx[x] + D.wx[x - pl]);
  }
}

// ================================================================= wafone
// Vertical WAF passes (:951-1073), one column per lane, a sliding window of
// four levels so the second pass can update its input in place.  Q is the
// storage type of the advected field (FP64 for the Exner function).
template <typename T, typename Q>
__device__ __forceinline__ void waf_vpass(const Dev<T> &D, const Geo &g, int j, int i, const Q *src, Q *dst,
                                          Q dtrdz, bool halo) {
  const int nk = g.nk;
  const long c0 = g.at(j, i, 1);
  auto L = [&](int k) -> Q { return src[c0 + (long)(k - 1) * g.plane]; };
  Q qm1 = L(1), q0 = L(1), q1 = L(2), q2 = nk >= 3 ? L(3) : L(2);
  Q wfw_k = Q(0);  // wfw(1) = 0
  for (int k = 1; k <= nk; k++) {
    long x = c0 + (long)(k - 1) * g.plane;
    Q sk = (Q)D.s[x], sk1 = (Q)D.s[x + g.plane];
    Q wfw_k1 = Q(0);
    if (k <= nk - 1) {
      Q zamu = sk1 * dtrdz, is, a, b;
      if (zamu >= Q(0)) {
        is = Q(1);
        a = q1;                  // level k+1
        b = (k + 2 > nk) ? q1 : q2;  // level min(k+2, nk)
      } else {
        is = Q(-1);
        a = (k - 1 < 1) ? q0 : qm1;  // level max(k-1, 1)
        b = q0;                      // level k
      }
      Q zzden = sgn_floor(q0 - q1);
      Q r = wdiv(a - b, zzden);
      Q bb = waf_b(r);
      Q zphi = is + zamu * bb - is * bb;
      wfw_k1 = Q(0.5) * sk1 * ((Q(1) + zphi) * q1 + (Q(1) - zphi) * q0);
    }
    Q f0 = (Q)D.fmz[x];
    Q zrfmu = wdiv(dtrdz * f0, (Q)D.fmzf[x]);
    Q zrfmd = wdiv(dtrdz * f0, (Q)D.fmzf[x + g.plane]);
    Q zdv = (sk * zrfmu - sk1 * zrfmd) * q0;
    Q out = q0 - wfw_k * zrfmu + wfw_k1 * zrfmd + zdv;
    // advance the window before the in-place store of level k
    Q q3 = (k + 3 <= nk) ? L(k + 3) : q2;
    dst[x] = out;
    if (halo) {
      int ni = g.ni;
      if (i == 1) dst[g.at(j, ni + 1, k)] = out;
      if (i == 2) dst[g.at(j, ni + 2, k)] = out;
      if (i == ni) dst[g.at(j, 0, k)] = out;
      if (i == ni - 1) dst[g.at(j, -1, k)] = out;
    }
    qm1 = q0;
    q0 = q1;
    q1 = q2;
    q2 = q3;
    wfw_k = wfw_k1;
  }
}

template <typename T, typename Q>
__global__ void __launch_bounds__(256) k_waf_v(Dev<T> D, Par P, const Q *pp, Q *wz, double dta, int twice) {
  COL_IDX
  Q dtrdz = (Q)(dta / P.dzita);
  if (twice) dtrdz = Q(0.5) * dtrdz;
  waf_vpass<T, Q>(D, g, j, i, pp, wz, dtrdz, !twice);
  if (twice) waf_vpass<T, Q>(D, g, j, i, wz, wz, dtrdz, true);
}

// Meridional sweep (:1082-1131 ROTLLR, :1190-1240 otherwise) -> p0 with its
// 2-wide j halo (:1133-1136).  Both fluxes of a cell are evaluated by its
// lane.
template <typename T, typename Q>
__global__ void __launch_bounds__(256) k_waf_m(Dev<T> D, Par P, const Q *pp, const Q *wz, Q *p0, double dta) {
  const int j = blockIdx.x * 64 + threadIdx.x + 1;
  const int i = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  const Geo &g = D.g;
  if (j > g.nj || i > g.ni) return;
  const Q dtrdy = (Q)(dta / P.dx);
  const int imin = -1, imax = g.ni + 2;  // :264-267
  auto flux = [&](int ii) -> Q {
    long xx = g.at(j, ii, k);
    Q vv = (Q)D.v[xx];
    Q zamu = P.rotllr ? vv * dtrdy : vv * (Q)D.rmv[g.at2(j, ii)] * dtrdy;
    Q is;
    int ih;
    if (zamu > Q(0)) {
      is = Q(1);
      ih = ii - 1;
    } else {
      is = Q(-1);
      ih = ii + 1 < imax ? ii + 1 : imax;
    }
    int ihm1 = ih - 1 > imin ? ih - 1 : imin;
    Q wc = wz[xx], wm = wz[xx - g.pj];
    Q zzden = sgn_floor(wc - wm);
    Q r = wdiv(wz[g.at(j, ih, k)] - wz[g.at(j, ihm1, k)], zzden);
    Q b = waf_b(r);
    Q zphi = is + zamu * b - is * b;
    if (P.rotllr) return Q(0.5) * vv * ((Q(1) + zphi) * wm + (Q(1) - zphi) * wc);
    return Q(0.5) * vv * (Q)D.rmv[g.at2(j, ii)] * ((Q(1) + zphi) * wm + (Q(1) - zphi) * wc);
  };
  long x = g.at(j, i, k);
  Q zs = flux(i), zn = flux(i + 1);
  Q f0 = (Q)D.fmz[x];
  Q out;
  if (P.rotllr) {
    Q mx = (Q)D.mx[g.at2(j, i)];
    Q zhxvtn = dtrdy * (Q)D.rmv[g.at2(j, i + 1)] * mx;
    Q zhxvts = dtrdy * (Q)D.rmv[g.at2(j, i)] * mx;
    Q zrfmn = wdiv(zhxvtn * Q(2) * f0, f0 + (Q)D.fmz[x + g.pj]);
    Q zrfms = wdiv(zhxvts * Q(2) * f0, f0 + (Q)D.fmz[x - g.pj]);
    Q zdv = ((Q)D.v[x + g.pj] * zrfmn - (Q)D.v[x] * zrfms) * pp[x];
    out = wz[x] + zs * zrfms - zn * zrfmn + zdv;
  } else {
    Q zrfmn = wdiv(dtrdy * Q(2) * f0, f0 + (Q)D.fmz[x + g.pj]);
    Q zrfms = wdiv(dtrdy * Q(2) * f0, f0 + (Q)D.fmz[x - g.pj]);
    Q zdv = ((Q)D.v[x + g.pj] * (Q)D.rmv[g.at2(j, i + 1)] * zrfmn - (Q)D.v[x] * (Q)D.rmv[g.at2(j, i)] * zrfms) * pp[x];
    out = wz[x] + (Q)D.mx2[g.at2(j, i)] * (zs * zrfms - zn * zrfmn + zdv);
  }
  p0[x] = out;
  const int nj = g.nj;
  if (j == 1) p0[g.at(nj + 1, i, k)] = out;
  if (j == 2) p0[g.at(nj + 2, i, k)] = out;
  if (j == nj) p0[g.at(0, i, k)] = out;
  if (j == nj - 1) p0[g.at(-1, i, k)] = out;
}

// Zonal sweep (:1138-1187 ROTLLR, :1247-1296 otherwise) -> pp (interior).
template <typename T, typename Q>
__global__ void __launch_bounds__(256) k_waf_z(Dev<T> D, Par P, Q *pp, const Q *p0, double dta) {
  const int j = blockIdx.x * 64 + threadIdx.x + 1;
  const int i = blockIdx.y * 4 + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  const Geo &g = D.g;
  if (j > g.nj || i > g.ni) return;
  const Q dtrdx = (Q)(dta / P.dx);
  const int jmin = -1, jmax = g.nj + 2;  // :264-267
  auto flux = [&](int jj) -> Q {
    long xx = g.at(jj, i, k);
    Q uu = (Q)D.u[xx];
    Q zamu = P.rotllr ? uu * (Q)D.mu[g.at2(jj, i)] * dtrdx : uu * (Q)D.rmu[g.at2(jj, i)] * dtrdx;
    Q is;
    int jh;
    if (zamu > Q(0)) {
      is = Q(1);
      jh = jj - 1;
    } else {
      is = Q(-1);
      jh = jj + 1 < jmax ? jj + 1 : jmax;
    }
    int jhm1 = jh - 1 > jmin ? jh - 1 : jmin;
    Q pc = p0[xx], pm = p0[xx - 1];
    Q zzden = sgn_floor(pc - pm);
    Q r = wdiv(p0[g.at(jh, i, k)] - p0[g.at(jhm1, i, k)], zzden);
    Q b = waf_b(r);
    Q zphi = is + zamu * b - is * b;
    if (P.rotllr) return Q(0.5) * uu * ((Q(1) + zphi) * pm + (Q(1) - zphi) * pc);
    return Q(0.5) * uu * (Q)D.rmu[g.at2(jj, i)] * ((Q(1) + zphi) * pm + (Q(1) - zphi) * pc);
  };
  long x = g.at(j, i, k);
  Q zw = flux(j), ze = flux(j + 1);
  Q f0 = (Q)D.fmz[x];
  Q out;
  if (P.rotllr) {
    Q zcostx = dtrdx * (Q)D.mu[g.at2(j, i)];
    Q zrfme = wdiv(zcostx * Q(2) * f0, f0 + (Q)D.fmz[x + 1]);
    Q zrfmw = wdiv(zcostx * Q(2) * f0, f0 + (Q)D.fmz[x - 1]);
    Q zdv = ((Q)D.u[x + 1] * zrfme - (Q)D.u[x] * zrfmw) * pp[x];
    out = p0[x] + zw * zrfmw - ze * zrfme + zdv;
  } else {
    Q zrfme = wdiv(dtrdx * Q(2) * f0, f0 + (Q)D.fmz[x + 1]);
    Q zrfmw = wdiv(dtrdx * Q(2) * f0, f0 + (Q)D.fmz[x - 1]);
    Q zdv = ((Q)D.u[x + 1] * (Q)D.rmu[g.at2(j + 1, i)] * zrfme - (Q)D.u[x] * (Q)D.rmu[g.at2(j, i)] * zrfmw) * pp[x];
    out = p0[x] + (Q)D.mx2[g.at2(j, i)] * (zw * zrfmw - ze * zrfme + zdv);
  }
  pp[x] = out;
}

// WAF flux through interface m (between levels m-1 above and m below),
// :966-1001 written for a window: qm2..q1 are levels m-2..m+1, clamped.
template <typename Q>
__device__ __forceinline__ Q waf_vflux(Q s, Q dtrdz, Q qm2, Q qm1, Q q0, Q q1) {
  Q zamu = s * dtrdz, is, a, b;
  if (zamu >= Q(0)) {
    is = Q(1);
    a = q0;
    b = q1;
  } else {
    is = Q(-1);
    a = qm2;
    b = qm1;
  }
  Q zzden = sgn_floor(qm1 - q0);
  Q r = wdiv(a - b, zzden);
  Q bb = waf_b(r);
  Q zphi = is + zamu * bb - is * bb;
  return Q(0.5) * s * ((Q(1) + zphi) * q0 + (Q(1) - zphi) * qm1);
}

// ============================================= multi-field wafone, 2 kernels
// The advected set (:911-920) goes through wafone in groups of fields of one
// storage type (at most waf_gmax<Q>() per group); in the FP32 core the FP64
// θ and Exner function ride along with the last group in the horizontal
// kernel (FieldD).  k_wafv runs both vertical half passes for every field of
// the group, one column per lane, sharing the loads of s, fmz, fmzf and the
// zrfmu/zrfmd divisions between the fields.  k_wafh then does the meridional
// and zonal sweeps of every field of the group for one strip of one level
// per wave, sharing the loads of u, v, fmz and the map factors.  Same
// arithmetic as the single-field kernels.
// Group sizes: all eight FP32 fields in one group (the shared loads once per
// cycle: 2 ms per 2048 x 2048 x 60 step faster than groups of five and three,
// at 166 VGPRs for the horizontal kernel with its FP64 riders); FP64 fields
// in groups of five (eight need 212-254 VGPRs).
constexpr int MAXF = 8;
template <typename Q>
constexpr int waf_gmax() {
  return sizeof(Q) == sizeof(double) ? 5 : 8;
}

template <typename Q>
struct FieldSet {
  const Q *in[MAXF];  // fields before advection (also read by the zdv terms)
  Q *wz[MAXF];        // after the vertical passes
  Q *out[MAXF];       // after the horizontal sweeps
};

// An FP64 field riding along with a group of FP32 fields (the Exner function
// of the FP32 core): advected with every coefficient evaluated in FP64 from
// the same loads, i.e. exactly as in a group of its own.
// FP64 riders of a group of FP32 fields (theta and the Exner function).
constexpr int MAXD = 2;
struct FieldD {
  const double *in[MAXD];
  double *wz[MAXD];
  double *out[MAXD];
};

// Field-independent window of the vertical passes in precision R: s at
// interfaces m-4..m, zrfmu / zrfmd (:1010-1011) of levels m-4..m.
template <typename R>
struct VCoef {
  R S[5], ZU[5], ZD[5], ffz, dtrdz;
  __device__ __forceinline__ void init(R ffz0, double dta, double dzita) {
    dtrdz = R(0.5) * (R)(dta / dzita);
    ffz = ffz0;
#pragma unroll
    for (int q = 0; q < 5; q++) S[q] = R(0), ZU[q] = R(0), ZD[q] = R(0);
  }
  __device__ __forceinline__ void advance(int m, int nk, R sm, R fzm, R ff1) {
#pragma unroll
    for (int q = 0; q < 4; q++) S[q] = S[q + 1], ZU[q] = ZU[q + 1], ZD[q] = ZD[q + 1];
    S[4] = sm;
    if (m <= nk) {
      ZU[4] = wdiv(dtrdz * fzm, ffz);  // zrfmu (:1010)
      ZD[4] = wdiv(dtrdz * fzm, ff1);  // zrfmd (:1011)
      ffz = ff1;
    }
  }
};

// One field's window of the two vertical half passes: levels m-3..m of the
// field (Pw), first-pass levels m-5..m-2 (Ww), the last fluxes.  step(m)
// takes the field at level min(m, kz) and stores the result of level m-4.
template <typename R>
struct VField {
  R Pw[4], Ww[4], F1, F2;
  __device__ __forceinline__ void init(R p1) {
#pragma unroll
    for (int q = 0; q < 4; q++) Pw[q] = p1;
    F1 = R(0);
    F2 = R(0);
  }
  template <typename I>
  __device__ __forceinline__ void step(const VCoef<R> &c, int m, int nk, R pm, R *wz, I c0, I pl) {
    const int kb = m - 2, kc = m - 4;
#pragma unroll
    for (int q = 0; q < 3; q++) Pw[q] = Pw[q + 1];
    Pw[3] = pm;
    // first pass, level kb from interfaces kb (F1) and kb+1 = m-1
    if (kb >= 1 && kb <= nk) {
      R f1 = R(0);
      if (m - 1 >= 2 && m - 1 <= nk) f1 = waf_vflux(c.S[3], c.dtrdz, Pw[0], Pw[1], Pw[2], Pw[3]);
      R q = Pw[1];
      R zdv = (c.S[2] * c.ZU[2] - c.S[3] * c.ZD[2]) * q;
      R w1 = q - F1 * c.ZU[2] + f1 * c.ZD[2] + zdv;
      F1 = f1;
      if (kb == 1) {
        Ww[0] = Ww[1] = Ww[2] = Ww[3] = w1;
      } else {
        Ww[0] = Ww[1];
        Ww[1] = Ww[2];
        Ww[2] = Ww[3];
        Ww[3] = w1;
      }
    } else if (kb > nk) {
      Ww[0] = Ww[1];
      Ww[1] = Ww[2];
      Ww[2] = Ww[3];
    }
    // second pass, level kc from interfaces kc (F2) and kc+1 = m-3
    if (kc >= 1 && kc <= nk) {
      R f2 = R(0);
      if (m - 3 >= 2 && m - 3 <= nk) f2 = waf_vflux(c.S[1], c.dtrdz, Ww[0], Ww[1], Ww[2], Ww[3]);
      R q = Ww[1];
      R zdv = (c.S[0] * c.ZU[0] - c.S[1] * c.ZD[0]) * q;
      stx(wz, (I)(c0 + (I)(kc - 1) * pl), q - F2 * c.ZU[0] + f2 * c.ZD[0] + zdv);
      F2 = f2;
    }
  }
};

// Both vertical half passes (:951-1073) of NF fields of type Q, one column
// per lane, k swept as a sliding window; the loads of s, fmz, fmzf and the
// zrfmu/zrfmd divisions are shared by the fields.  (An FP64 rider of the
// horizontal kernel gets a launch of its own here: in one kernel its FP64
// window and coefficients cost half the occupancy.)
template <typename T, typename Q, int NF, typename I>
__global__ void __launch_bounds__(256) k_wafv(Dev<T> D, Par P, FieldSet<Q> F, double dta) {
  COL_IDX
  const int nk = g.nk;
  // element indices in I (unsigned: 32-bit byte offsets, as in k_wafh)
  const I c0 = (I)g.at(j, i, 1);
  const I pl = (I)g.plane;
  VCoef<Q> cq;
  VField<Q> vf[NF];
  // the loads of iteration m (s at interface m, fmz and the fields at level
  // min(m, kz), fmzf below it), issued one iteration ahead; past the column
  // they repeat the last level and are not used
  T n_s, n_fz, n_ff;
  Q n_p[NF];
  auto fetch = [&](int m) {
    const int mm = m <= nk ? m : nk, ms = m <= nk + 1 ? m : nk + 1;
    const I xm = (I)(c0 + (I)(mm - 1) * pl);
    n_s = ldx(D.s, (I)(c0 + (I)(ms - 1) * pl));
    n_fz = ldx(D.fmz, xm);
    n_ff = ldx(D.fmzf, (I)(xm + pl));
#pragma unroll
    for (int f = 0; f < NF; f++) n_p[f] = ldx(F.in[f], xm);
  };
  fetch(1);
  cq.init((Q)ldx(D.fmzf, c0), dta, P.dzita);
  for (int m = 1; m <= nk + 4; m++) {
    const T sm = m <= nk + 1 ? n_s : T(0);
    const T fzm = n_fz, ff1 = n_ff;
    Q pin[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) pin[f] = n_p[f];
    fetch(m + 1);
    if (m == 1) {
#pragma unroll
      for (int f = 0; f < NF; f++) vf[f].init(pin[f]);
    }
    cq.advance(m, nk, (Q)sm, (Q)fzm, (Q)ff1);
#pragma unroll
    for (int f = 0; f < NF; f++) vf[f].step(cq, m, nk, pin[f], F.wz[f], c0, pl);
  }
}

// Horizontal sweeps of one level, one wave per strip of HJ columns x HR rows:
// the wave's 64 lanes are the strip's columns plus the two-column reach of
// the zonal sweep on either side.
constexpr int HJ = 60, HR = 32;

// Block index -> work item, grouping consecutive items on one XCD: the
// hardware hands consecutive workgroups to the eight XCDs in turn, and
// neighbouring strips share halo lines through that XCD's L2.
__device__ __forceinline__ long xcd_remap(long b, long total) {
  const long nx = 8;
  long chunk = (total + nx - 1) / nx;
  return (b % nx) * chunk + b / nx;
}

// Periodic index in 1..n of any integer.
__device__ __forceinline__ int wrap_idx(int x, int n) { return ((x - 1) % n + n) % n + 1; }

// For each field of the group the rows of the strip are swept south to
// north with a window of five rows of wz (rows r-2 .. r+2) in registers:
// the meridional flux at the north face of row r (:1082-1100) comes from the
// window, p0 of row r (:1104-1131) from that flux and the one carried from
// row r-1, the zonal face fluxes (:1138-1160) and the result (:1162-1187)
// from p0 of the neighbouring lanes.  No block barrier; the loads of row
// r+1 are issued before row r is evaluated.  Everything that does not depend
// on the advected field (Courant numbers, fmz ratios, the divergence-
// correction coefficients) is evaluated once per row for all fields.
// meridional flux through the south face of a row (Courant number mz,
// amplitude ma) from wz of rows c-2, c-1, c, c+1 (:1082-1131)
template <typename R>
__device__ __forceinline__ R waf_mflux(R mz, R ma, R wmm, R wm, R wc, R wp) {
  R is, a, b;
  if (mz > R(0)) {
    is = R(1);
    a = wm;
    b = wmm;
  } else {
    is = R(-1);
    a = wp;
    b = wc;
  }
  const R zzden = sgn_floor(wc - wm);
  const R rq = wdiv(a - b, zzden);
  const R bb = waf_b(rq);
  const R zphi = is + mz * bb - is * bb;
  return ma * ((R(1) + zphi) * wm + (R(1) - zphi) * wc);
}

// field-independent loads of one row of a strip (storage type T), the
// coefficients made of them in precision R
template <typename T>
struct HRow {
  T vn, fn, fe, fw, uc, up, rvn, m0, m1, m2;
};
template <typename R>
struct HCoef {
  R mzn, man, pfs, pfn, pcd, pm2, zz, za, zfw, zfe, zcd, zm2;
};
template <typename R, bool ROT, typename T>
__device__ __forceinline__ HCoef<R> waf_hcoef(const HRow<T> &ci, T v_c, T rv_c, T f_s, T f_c, R dtrdx, R dtrdy) {
  HCoef<R> c;
  const R vn = (R)ci.vn, rvn = (R)ci.rvn;
  if constexpr (ROT) {
    c.mzn = vn * dtrdy;
    c.man = R(0.5) * vn;
  } else {
    c.mzn = vn * rvn * dtrdy;
    c.man = R(0.5) * vn * rvn;
  }
  {
    const R f0 = (R)f_c, fn = (R)ci.fn, fs = (R)f_s;
    const R vc = (R)v_c, vp = vn;
    if constexpr (ROT) {
      const R mx = (R)ci.m0;
      const R zhxvtn = dtrdy * rvn * mx;
      const R zhxvts = dtrdy * (R)rv_c * mx;
      c.pfn = wdiv(zhxvtn * R(2) * f0, f0 + fn);
      c.pfs = wdiv(zhxvts * R(2) * f0, f0 + fs);
      c.pcd = vp * c.pfn - vc * c.pfs;
      c.pm2 = R(1);
    } else {
      c.pfn = wdiv(dtrdy * R(2) * f0, f0 + fn);
      c.pfs = wdiv(dtrdy * R(2) * f0, f0 + fs);
      c.pcd = vp * rvn * c.pfn - vc * (R)rv_c * c.pfs;
      c.pm2 = (R)ci.m0;
    }
  }
  {
    const R uc = (R)ci.uc, up = (R)ci.up;
    const R f0 = (R)f_c, fe = (R)ci.fe, fw = (R)ci.fw;
    if constexpr (ROT) {
      const R mu = (R)ci.m1;
      c.zz = uc * mu * dtrdx;
      c.za = R(0.5) * uc;
      const R zcostx = dtrdx * mu;
      c.zfe = wdiv(zcostx * R(2) * f0, f0 + fe);
      c.zfw = wdiv(zcostx * R(2) * f0, f0 + fw);
      c.zcd = up * c.zfe - uc * c.zfw;
      c.zm2 = R(1);
    } else {
      const R rm = (R)ci.m1;
      c.zz = uc * rm * dtrdx;
      c.za = R(0.5) * uc * rm;
      c.zfe = wdiv(dtrdx * R(2) * f0, f0 + fe);
      c.zfw = wdiv(dtrdx * R(2) * f0, f0 + fw);
      c.zcd = up * (R)ci.m2 * c.zfe - uc * rm * c.zfw;
      c.zm2 = (R)ci.m0;
    }
  }
  return c;
}

// One field, one row: the meridional flux at the north face and p0 of the
// row (:1082-1131), then the zonal flux at each lane's west face and the
// result (:1138-1187); the window of wz rows slides north.
template <typename R, bool ROT>
__device__ __forceinline__ R waf_hrow(R (&w)[5], R &zs, R pp, const HCoef<R> &c, int lm1, int lm2, int lp1) {
  const R zn = waf_mflux(c.mzn, c.man, w[1], w[2], w[3], w[4]);
  R p0;
  const R zdv = c.pcd * pp;
  if constexpr (ROT)
    p0 = w[2] + zs * c.pfs - zn * c.pfn + zdv;
  else
    p0 = w[2] + c.pm2 * (zs * c.pfs - zn * c.pfn + zdv);
  const R pm = __shfl(p0, lm1, 64), pmm = __shfl(p0, lm2, 64), ppl = __shfl(p0, lp1, 64);
  R is, a, b;
  if (c.zz > R(0)) {
    is = R(1);
    a = pm;
    b = pmm;
  } else {
    is = R(-1);
    a = ppl;
    b = p0;
  }
  const R zzden = sgn_floor(p0 - pm);
  const R rq = wdiv(a - b, zzden);
  const R bb = waf_b(rq);
  const R zphi = is + c.zz * bb - is * bb;
  const R zw = c.za * ((R(1) + zphi) * pm + (R(1) - zphi) * p0);
  const R ze = __shfl(zw, lp1, 64);
  const R zdu = c.zcd * pp;
  R o;
  if constexpr (ROT)
    o = p0 + zw * c.zfw - ze * c.zfe + zdu;
  else
    o = p0 + c.zm2 * (zw * c.zfw - ze * c.zfe + zdu);
  w[0] = w[1];
  w[1] = w[2];
  w[2] = w[3];
  w[3] = w[4];
  zs = zn;
  return o;
}

// Both horizontal sweeps of NF fields of type Q and ND (0 or 1) FP64 fields
// for one strip of one level per wave; the field-independent coefficients
// are evaluated once per row and precision.
template <typename T, typename Q, int NF, int ND, bool ROT, typename I>
__global__ void __launch_bounds__(256) k_wafh(Dev<T> D, Par P, FieldSet<Q> F, FieldD FD, double dta, int ntj,
                                              int nti) {
  static_assert(ND >= 0 && ND <= MAXD, "at most MAXD FP64 riders");
  const Geo &g = D.g;
  const int nj = g.nj, ni = g.ni;
  const long total = (long)ntj * nti * g.nk;  // strips
  const long nblk = (total + 3) / 4;
  const long lb = xcd_remap((long)blockIdx.x, nblk);
  if (lb >= nblk) return;
  const long item = lb * 4 + (threadIdx.x >> 6);
  if (item >= total) return;  // the whole wave
  const int lane = threadIdx.x & 63;
  const int tj = (int)(item % ntj), ti = (int)((item / ntj) % nti), k = (int)(item / ((long)ntj * nti)) + 1;
  const int j0 = tj * HJ + 1, r0 = ti * HR + 1;
  const int nrows = ni - r0 + 1 < HR ? ni - r0 + 1 : HR;
  const int ja = j0 - 2 + lane;  // the lane's column (unwrapped)
  const bool store = lane >= 2 && lane < 2 + HJ && ja <= nj;
  const int jw = wrap_idx(ja, nj);
  const int jp = jw == nj ? 1 : jw + 1, jm = jw == 1 ? nj : jw - 1;
  const Q dtrdx = (Q)(dta / P.dx), dtrdy = (Q)(dta / P.dx);
  const double dtrdx_d = dta / P.dx, dtrdy_d = dta / P.dx;
  // element indices in I: unsigned (32-bit byte offsets from the SGPR
  // bases, one VGPR per address) when every array is below 4 GiB, else long
  const I pl = (I)((long)(k - 1) * g.plane), pj = (I)g.pj;
  auto at3 = [&](int jj, int ii) { return (I)(pl + (I)(ii + 1) * pj + (I)(jj + JOFF)); };
  auto at2 = [&](int jj, int ii) { return (I)((I)(ii + 1) * pj + (I)(jj + JOFF)); };
  auto nxt = [&](int ii) { return ii == ni ? 1 : ii + 1; };
  const int lm1 = lane > 0 ? lane - 1 : 0, lm2 = lane > 1 ? lane - 2 : 0, lp1 = lane < 63 ? lane + 1 : 63;

  // rows r-2 .. r+2 of the first row r = r0
  int i_m2 = wrap_idx(r0 - 2, ni);
  int i_m1 = nxt(i_m2), i_c = nxt(i_m1), i_p1 = nxt(i_c), i_p2 = nxt(i_p1);
  Q w[NF][5], zs[NF];
  double wd[ND > 0 ? ND : 1][5], zsd[ND > 0 ? ND : 1];
  {
    // Courant number and amplitude of the south face of the first row
    const T vv = ldx(D.v, at3(jw, i_c));
    const T rm = ROT ? T(0) : ldx(D.rmv, at2(jw, i_c));
    Q mz, ma;
    double mzd, mad;
    if constexpr (ROT) {
      mz = (Q)vv * dtrdy;
      ma = Q(0.5) * (Q)vv;
      mzd = (double)vv * dtrdy_d;
      mad = 0.5 * (double)vv;
    } else {
      mz = (Q)vv * (Q)rm * dtrdy;
      ma = Q(0.5) * (Q)vv * (Q)rm;
      mzd = (double)vv * (double)rm * dtrdy_d;
      mad = 0.5 * (double)vv * (double)rm;
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
      w[f][0] = ldx(F.wz[f], at3(jw, i_m2));
      w[f][1] = ldx(F.wz[f], at3(jw, i_m1));
      w[f][2] = ldx(F.wz[f], at3(jw, i_c));
      w[f][3] = ldx(F.wz[f], at3(jw, i_p1));
      zs[f] = waf_mflux(mz, ma, w[f][0], w[f][1], w[f][2], w[f][3]);
    }
#pragma unroll
    for (int d = 0; d < ND; d++) {
      wd[d][0] = ldx(FD.wz[d], at3(jw, i_m2));
      wd[d][1] = ldx(FD.wz[d], at3(jw, i_m1));
      wd[d][2] = ldx(FD.wz[d], at3(jw, i_c));
      wd[d][3] = ldx(FD.wz[d], at3(jw, i_p1));
      zsd[d] = waf_mflux(mzd, mad, wd[d][0], wd[d][1], wd[d][2], wd[d][3]);
    }
  }
  // field-independent loads of the row centred at ic, issued a row ahead:
  // v, fmz and rmv at its north neighbour (the south ones are carried from
  // the previous row), fmz east and west, u here and east, the map factors
  auto row_in = [&](int ic, int in) {
    HRow<T> o;
    o.vn = ldx(D.v, at3(jw, in));
    o.fn = ldx(D.fmz, at3(jw, in));
    o.fe = ldx(D.fmz, at3(jp, ic));
    o.fw = ldx(D.fmz, at3(jm, ic));
    o.uc = ldx(D.u, at3(jw, ic));
    o.up = ldx(D.u, at3(jp, ic));
    o.rvn = ldx(D.rmv, at2(jw, in));
    if constexpr (ROT) {
      o.m0 = ldx(D.mx, at2(jw, ic));
      o.m1 = ldx(D.mu, at2(jw, ic));
      o.m2 = T(0);
    } else {
      o.m0 = ldx(D.mx2, at2(jw, ic));
      o.m1 = ldx(D.rmu, at2(jw, ic));
      o.m2 = ldx(D.rmu, at2(jp, ic));
    }
    return o;
  };
  // carried south values of the current row
  T v_c = ldx(D.v, at3(jw, i_c)), rv_c = ldx(D.rmv, at2(jw, i_c));
  T f_s = ldx(D.fmz, at3(jw, i_m1)), f_c = ldx(D.fmz, at3(jw, i_c));
  HRow<T> nx = row_in(i_c, i_p1);
  // loads of the first row
  Q nw[NF], np[NF];
  double nwd[ND > 0 ? ND : 1], npd[ND > 0 ? ND : 1];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    nw[f] = ldx(F.wz[f], at3(jw, i_p2));
    np[f] = ldx(F.in[f], at3(jw, i_c));
  }
#pragma unroll
  for (int d = 0; d < ND; d++) {
    nwd[d] = ldx(FD.wz[d], at3(jw, i_p2));
    npd[d] = ldx(FD.in[d], at3(jw, i_c));
  }
  for (int r = 0; r < nrows; r++) {
    Q pp[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) {
      w[f][4] = nw[f];
      pp[f] = np[f];
    }
    double ppd[ND > 0 ? ND : 1];
#pragma unroll
    for (int d = 0; d < ND; d++) {
      ppd[d] = npd[d];
      wd[d][4] = nwd[d];
    }
    const HRow<T> ci = nx;
    if (r + 1 < nrows) {  // uniform over the wave
      const int i_n = nxt(i_p2);
#pragma unroll
      for (int f = 0; f < NF; f++) {
        nw[f] = ldx(F.wz[f], at3(jw, i_n));
        np[f] = ldx(F.in[f], at3(jw, i_p1));
      }
#pragma unroll
      for (int d = 0; d < ND; d++) {
        nwd[d] = ldx(FD.wz[d], at3(jw, i_n));
        npd[d] = ldx(FD.in[d], at3(jw, i_p1));
      }
      nx = row_in(i_p1, i_p2);
    }
    const HCoef<Q> cq = waf_hcoef<Q, ROT>(ci, v_c, rv_c, f_s, f_c, dtrdx, dtrdy);
    const I xo = at3(ja <= nj ? ja : jw, i_c);
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const Q o = waf_hrow<Q, ROT>(w[f], zs[f], pp[f], cq, lm1, lm2, lp1);
      if (store) stx(F.out[f], xo, o);
    }
    if (ND) {
      const HCoef<double> cd = waf_hcoef<double, ROT>(ci, v_c, rv_c, f_s, f_c, dtrdx_d, dtrdy_d);
#pragma unroll
      for (int d = 0; d < ND; d++) {
        const double o = waf_hrow<double, ROT>(wd[d], zsd[d], ppd[d], cd, lm1, lm2, lp1);
        if (store) stx(FD.out[d], xo, o);
      }
    }
    // the north values become the next row's south ones
    v_c = ci.vn;
    rv_c = ci.rvn;
    f_s = f_c;
    f_c = ci.fn;
    i_m2 = i_m1;
    i_m1 = i_c;
    i_c = i_p1;
    i_p1 = i_p2;
    i_p2 = nxt(i_p2);
  }
}

// ================================================================ epilogue
// :355-391 -- tvirt, t, p, rho, ps (level-kz extrapolation), qsat.
template <typename T>
__global__ void __launch_bounds__(256) k_epilogue(Dev<T> D) {
  // one point per thread (k in the grid); the lowest level's thread also
  // extrapolates ps, with tvirt of the level above formed again
  COL_IDX
  const int nk = g.nk;
  double tvk = 0, tvkm = 0, pk = 0;
  {
    const int k = blockIdx.z + 1;
    long x = g.at(j, i, k);
    double pai = D.pai[x];
    double tv = D.tetav[x] * pai;
    T tv_t = (T)tv;
    D.tvirt[x] = tv_t;
    // the temperature from the FP64 virtual temperature (tv_t == tv in FP64)
    double t = tv / (1.0 + kEp1 * (double)D.qv[x] - (double)D.qc[x] - (double)D.qi[x] - (double)D.qr[x] -
                     (double)D.qs[x]);
    D.t[x] = t;
    double p = pow(pai, kCpovr) * kP00;
    T p_t = (T)p;
    D.p[x] = p_t;
    D.rho[x] = (T)((double)p_t / (kRgas * t));
    D.qsat[x] = (T)d_pfwsat(t, (double)p_t);
    if (k != nk) return;
    tvkm = (double)(T)(D.tetav[x - g.plane] * D.pai[x - g.plane]);
    tvk = (double)tv_t;
    pk = (double)p_t;
  }
  double zk = (double)D.zeta[g.at(j, i, nk)], zkm = (double)D.zeta[g.at(j, i, nk - 1)];
  double zdgz = zk * kEgrav;
  double lrt = (tvkm - tvk) / (zkm - zk);
  lrt = 0.65 * lrt - 0.35 * kLrate;
  double tv = tvk - 0.5 * zk * lrt;
  D.ps[g.at2(j, i)] = (T)(pk * exp(zdgz / (kRgas * tv)));
}

// =================================================== learned acoustic corrector
// A network corrects one long acoustic substep of (nsound-1)*dts toward
// nsound-1 ordinary substeps; the last ordinary substep then runs as usual.
// Inputs of a column (j,i), taken after the long substep: the increments it
// made to u, v, w (levels 2..kz+1) and the Exner function, and theta and qv,
// over kz levels at the 3 x 3 columns around it.  "fl" = field * kz +
// level - 1 is padded to nflp = roundup(6 kz, 16) with zero features.
// Feature o packs 8 consecutive fl at one stencil position p, (dj, di) =
// (p%3-1, p/3-1): chunk c = o/8 -> pair of fl groups dg = c/18, half
// h = (c%18)/9, p = c%9.  Outputs: increments of u, v, w (levels 2..kz+1)
// and Exner, o = f*kz + k-1.  Rows of the whole grid are ordered by 64 x 4
// tiles of columns: r -> tile r/256 = (tj, ti) in j-fastest order, then
// j-run (r%256)/64 and j-offset r%64.
constexpr int LAC_IN_FIELDS = 6, LAC_OUT_FIELDS = 4;

__host__ __device__ __forceinline__ void lac_feature(int o, int &fl, int &p) {
  const int c = o >> 3, cc = c % 18;
  p = cc % 9;
  fl = (c / 18) * 16 + (cc / 9) * 8 + (o & 7);
}

struct LacRows {
  int ntj;  // tiles along j
  __device__ __forceinline__ void colof(long r, const Geo &g, int &j, int &i) const {
    const long t = r >> 8;
    const int w = (int)(r & 255);
    j = (int)(t % ntj) * 64 + (w & 63) + 1;
    i = (int)(t / ntj) * 4 + (w >> 6) + 1;
  }
};

// The increments that are network inputs: from phase-start copies of u, v,
// w, pi (inc = 0), or as the sound kernels wrote them (inc = 1: iu, iv, iw,
// ipai, the same float values).
template <typename T>
struct LacSnap {
  const T *u, *v, *w;
  const double *pai;
  const float *iu, *iv, *iw, *ipai;
  int inc;
};

template <typename T>
__device__ __forceinline__ float lac_value(const Dev<T> &D, const LacSnap<T> &S, int fl, int j, int i) {
  const Geo &g = D.g;
  const int nk = g.nk, f = fl / nk, k = fl % nk + 1;
  if (f >= LAC_IN_FIELDS) return 0.f;
  j = j < 1 ? j + g.nj : (j > g.nj ? j - g.nj : j);
  i = i < 1 ? i + g.ni : (i > g.ni ? i - g.ni : i);
  const long x = g.at(j, i, k);
  if (S.inc && f < 4) {
    const float *d = f == 0 ? S.iu : f == 1 ? S.iv : f == 2 ? S.iw : S.ipai;
    return d[x];
  }
  switch (f) {
    case 0: return (float)(D.u[x] - S.u[x]);
    case 1: return (float)(D.v[x] - S.v[x]);
    case 2: return (float)(D.w[x + g.plane] - S.w[x + g.plane]);
    case 3: return (float)(D.pai[x] - S.pai[x]);
    case 4: return (float)D.tetav[x];
    default: return (float)D.qv[x];
  }
}

// Features of 64 rows per block into row-major X (row stride nin): rows
// cols[r] = (i-1)*nj + (j-1) when cols is given (raw fp32 into Xf), else
// rows row0 .. row0+nrows-1 of the tile order (normalised bf16
// (x - mean[o]) * rstd[o] into Xb; rows outside the grid get zeros).
template <typename T>
__global__ void __launch_bounds__(256) k_lac_gather(Dev<T> D, LacSnap<T> S, LacRows R, const int *cols, long row0,
                                                    int nrows, int nin, float *Xf, __bf16 *Xb, const float *mean,
                                                    const float *rstd) {
  __shared__ float tile[64][65];
  const Geo &g = D.g;
  const int r0 = blockIdx.x * 64;
  const int tr = threadIdx.x & 63, tq = threadIdx.x >> 6;
  const int row = r0 + tr;
  int j = 0, i = 0;
  bool ok = row < nrows;
  if (ok) {
    if (cols) {
      const int c = cols[row];
      j = c % g.nj + 1;
      i = c / g.nj + 1;
    } else {
      R.colof(row0 + row, g, j, i);
      ok = j <= g.nj && i <= g.ni;
    }
  }
  for (int o0 = 0; o0 < nin; o0 += 64) {
    for (int q = tq; q < 64; q += 4) {
      const int o = o0 + q;
      float val = 0.f;
      if (ok && o < nin) {
        int fl, p;
        lac_feature(o, fl, p);
        val = lac_value(D, S, fl, j + p % 3 - 1, i + p / 3 - 1);
        if (Xb) val = (val - mean[o]) * rstd[o];
      }
      tile[tr][q] = val;
    }
    __syncthreads();
    for (int q = tq; q < 64; q += 4) {
      const int rr = r0 + q, o = o0 + tr;
      if (rr < nrows && o < nin) {
        const float v = tile[q][tr];
        if (Xb)
          Xb[(long)rr * nin + o] = (__bf16)v;
        else
          Xf[(long)rr * nin + o] = v;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------ fused network inference
// Inputs of the whole grid are first written as normalised bf16 "feature
// planes" F[pair dg][column][16 fl] (a pair of fl groups per plane, 32
// bytes per column), over the columns of the row tiles plus a periodic
// one-column ring: extended column (jj, ii), jj in [0, fpj), ii in [0, fpi),
// is grid column (wrap(jj), wrap(ii)).  The values are the training
// gather's, bit for bit.
struct FeatPlanes {
  __bf16 *F;
  int fpj, fpi;  // extended columns along j and i
  __host__ __device__ long plane() const { return (long)fpj * fpi * 16; }
};

template <typename T>
__global__ void __launch_bounds__(256) k_lac_features(Dev<T> D, LacSnap<T> S, FeatPlanes P, int ndg,
                                                      const float *fmean, const float *frstd) {
  typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf8;
  const int jj = blockIdx.x * 64 + threadIdx.x % 64, ii = blockIdx.y * 4 + threadIdx.x / 64;
  if (jj >= P.fpj || ii >= P.fpi) return;
  const long col = (long)ii * P.fpj + jj;
  const int j = ((jj - 1) % D.g.nj + D.g.nj) % D.g.nj + 1, i = ((ii - 1) % D.g.ni + D.g.ni) % D.g.ni + 1;
  for (int dg = 0; dg < ndg; dg++) {
    bf8 h[2];
#pragma unroll
    for (int fi = 0; fi < 16; fi++) {
      const int fl = dg * 16 + fi;
      const float v = lac_value(D, S, fl, j, i);
      h[fi >> 3][fi & 7] = (__bf16)((v - fmean[fl]) * frstd[fl]);
    }
    bf8 *dst = reinterpret_cast<bf8 *>(P.F + dg * P.plane() + col * 16);
    dst[0] = h[0];
    dst[1] = h[1];
  }
}

// Layer 1: block tile of 256 rows x 256 outputs on
// v_mfma_f32_32x32x16_bf16: eight waves (two per SIMD, so one wave's MFMAs
// cover the other's memory waits) of 128 rows x 64 outputs each.  B comes
// from weight fragments packed in the MFMA lane order, loaded PD k-steps
// ahead into registers and staged through LDS one k-step (16) at a time.  A
// is the feature planes: the block's 256 rows are one 64 x 4 tile of
// columns; for each pair of fl groups (16 fl, 9 stencil positions = 9
// k-steps) the 66 x 6 halo of the tile is copied into LDS as [column][16 fl]
// and every stencil position reads it at its offset; the next pair is loaded
// when the current one starts.  Epilogue: bias, then ReLU into bf16 H (or
// plain fp32 into Y).
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 lac_bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float lac_f32x16;

constexpr int FM = 256, FN = 256, FT = 512;  // block tile, threads
constexpr int HALO_J = 66, HALO_I = 6, HALO_C = HALO_J * HALO_I;
constexpr int ACOL = 24;  // bf16 per LDS column / row (16 used): 48-byte stride, conflict-free 16-byte reads
constexpr int DB = 4;     // k-steps of register prefetch
static_assert(HALO_C <= FT, "one halo column per thread");

struct FusedArgs {
  const lac_bf16x8 *Bp;  // packed fragments
  int bnt;               // n-tiles of 32 in the packed layout
  int nks;               // k-steps of 16
  int N;                 // valid outputs
  const float *bias;
  const __bf16 *A;       // row-major activations (rows of this launch), or null
  long lda;
  __bf16 *H;             // bf16 output with ReLU (ldh), or null
  long ldh;
  float *Y;              // fp32 output (ldy) when H is null
  long ldy;
  long row0;             // tile-order row of the launch's first row
  int nrt;               // row tiles (of 256) in this launch
  int nN;                // output tiles (of 256)
  FeatPlanes fp;         // layer 1 inputs
};

__global__ void __launch_bounds__(FT) k_mlp_fused(LacRows R, FusedArgs a) {
  // The k loop is unrolled in groups of KG = 9 steps (one pair of fl
  // groups) with prefetch depth PD dividing KG, so register slots are
  // compile-time and the loop body is free of branches: prefetches past the
  // end are clamped to the last stage and land in buffers nobody reads any
  // more.  Threads without a halo column copy into spare LDS columns.
  constexpr int KG = 9, PD = 3;
  __shared__ __attribute__((aligned(16))) __bf16 sA[2][FT * ACOL];
  __shared__ __attribute__((aligned(16))) lac_bf16x8 sB[2][FN / 32][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;  // rows wm*128, outputs wn*64
  // XCD-aware order: consecutive work items on the same XCD (round-robin
  // dispatch), so the output tiles of one row tile and neighbouring row
  // tiles share that XCD's L2.
  const int nb = gridDim.x, b = blockIdx.x;
  const int lin = (b & 7) * (nb >> 3) + (b >> 3);
  const int rt = lin / a.nN, ntile = lin % a.nN;
  if (rt >= a.nrt) return;
  const long rowb = a.row0 + (long)rt * FM;
  const int n0 = ntile * FN;
  const int nks = a.nks;  // a multiple of KG

  lac_f32x16 acc[4][2];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[x][y][r] = 0.f;

  // B of k-step ks: 8 fragments x 64 lanes, one per thread
  const lac_bf16x8 *bsrc = a.Bp + ((long)(n0 >> 5) + (tid >> 6)) * 64 + lane;
  lac_bf16x8 breg[PD];
  auto loadB = [&](int ks, int slot) {
    ks = ks < nks ? ks : nks - 1;
    breg[slot] = bsrc[(long)ks * a.bnt * 64];
  };
  auto commitB = [&](int slot, int buf) { sB[buf][tid >> 6][lane] = breg[slot]; };
  // feature-plane halo: column tid of the tile's 66 x 6 halo (threads past
  // the halo re-read its last column into their spare LDS column)
  const __bf16 *hsrc;
  {
    const int t = (int)(rowb >> 8), c = tid < HALO_C ? tid : HALO_C - 1;
    const int tj = t % R.ntj, ti = t / R.ntj;
    hsrc = a.fp.F + ((long)(ti * 4 + c / HALO_J) * a.fp.fpj + tj * 64 + c % HALO_J) * 16;
  }
  const int ndg = nks / 9;
  lac_bf16x8 hreg[2];
  auto loadH = [&](int dg) {
    dg = dg < ndg ? dg : ndg - 1;
    const lac_bf16x8 *src = reinterpret_cast<const lac_bf16x8 *>(hsrc + dg * a.fp.plane());
    hreg[0] = src[0];
    hreg[1] = src[1];
  };
  auto commitH = [&](int buf) {
    lac_bf16x8 *dst = reinterpret_cast<lac_bf16x8 *>(&sA[buf][tid * ACOL]);
    dst[0] = hreg[0];
    dst[1] = hreg[1];
  };

  auto mfma_step = [&](int abuf, int bbuf, int t) {
    lac_bf16x8 av[4], bv[2];
#pragma unroll
    for (int x = 0; x < 4; x++) {
      const int m = wm * 128 + 32 * x + (lane & 31);
      const int cc = 2 * t + (lane >> 5), h = cc / 9, p = cc % 9;
      const int col = ((m >> 6) + p / 3) * HALO_J + (m & 63) + p % 3;
      av[x] = *reinterpret_cast<const lac_bf16x8 *>(&sA[abuf][col * ACOL + h * 8]);
    }
#pragma unroll
    for (int y = 0; y < 2; y++) bv[y] = sB[bbuf][wn * 2 + y][lane];
#pragma unroll
    for (int x = 0; x < 4; x++)
#pragma unroll
      for (int y = 0; y < 2; y++) acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[x], bv[y], acc[x][y], 0, 0, 0);
  };

  // prologue: stage 0 into LDS; stages 1 .. PD into register slots s % PD
  loadH(0);
  commitH(0);
  loadB(0, 0);
  commitB(0, 0);
#pragma unroll
  for (int u = 1; u <= PD; u++) loadB(u, u % PD);
  __syncthreads();
  for (int k0 = 0; k0 < nks; k0 += KG) {
    const int dg = k0 / 9;  // the pair of fl groups of this group of k-steps
    loadH(dg + 1);
#pragma unroll
    for (int u = 0; u < KG; u++) {
      const int ks = k0 + u;
      mfma_step(dg & 1, ks & 1, u);
      // stage ks+1 from its register slot into the free LDS buffer, then
      // refill that slot with stage ks+1+PD (stage s lives in slot s % PD)
      commitB((u + 1) % PD, (ks + 1) & 1);
      if (u == KG - 1) commitH((dg + 1) & 1);
      loadB(ks + 1 + PD, (u + 1) % PD);
      __syncthreads();
    }
  }

  // epilogue: D[m][n] of tile (x, y): n = lane%32, m = 8*(r/4) + 4*(lane/32) + r%4
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) {
      const int n = n0 + wn * 64 + 32 * y + (lane & 31);
      if (n >= a.N) continue;
      const float bn = a.bias[n];
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int m = wm * 128 + 32 * x + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const long row = rowb - a.row0 + m;
        const float v = acc[x][y][r] + bn;
        if (a.H)
          a.H[row * a.ldh + n] = (__bf16)(v > 0.f ? v : 0.f);
        else
          a.Y[row * a.ldy + n] = v;
      }
    }
}

typedef __attribute__((address_space(3))) void lds_void;

// One 16-byte-per-lane global -> LDS copy: lane l's 16 bytes at gptr land at
// LDS byte address lds + 16 l (lds wave-uniform).  Issued from inline asm so
// that the compiler's wait insertion, which cannot tell the buffers apart,
// adds no wait for it; the kernels wait with wait_vm before reading the
// buffer, and before they end.  The LDS address goes through m0, which the
// compiler reserves: the asm saves m0 in a scalar temporary and restores it,
// so the compiler's own m0 value survives.  No vector output registers.
__device__ __forceinline__ void dma16(const void *gptr, const void *lds_ptr) {
  const unsigned lds = __builtin_amdgcn_readfirstlane((unsigned)(unsigned long)(const lds_void *)lds_ptr);
  unsigned save;
  __asm__ volatile(
      "s_mov_b32 %0, m0\n\ts_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off\n\ts_nop 0\n\ts_mov_b32 m0, %0"
      : "=&s"(save)
      : "v"(gptr), "s"(lds)
      : "memory");
}

// wait until at most n vector-memory loads of this wave are outstanding
template <int N>
__device__ __forceinline__ void wait_vm() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Layers 2 and 3: row-major bf16 activations A (rows of this launch) times
// packed weight fragments, the block tile and wave layout of k_mlp_fused.
// Stages of two k-steps (32 k) go straight from global memory into one of
// four LDS buffers (global_load_lds, 16 bytes per lane, lane l of a wave
// instruction landing at base + 16 l), three stages ahead of the MFMAs, so
// no registers hold prefetched data and a block barrier comes every two
// k-steps.  An A row of a stage is 64 bytes, its four 16-byte chunks stored
// at chunk ^ ((row >> 1) & 3) so the 16-byte reads of eight consecutive rows
// hit distinct banks.  The MFMAs are issued in k_mlp_fused's order, so the
// results are bit-identical.
constexpr int DKS = 2, DNB = 4;  // k-steps per stage, LDS buffers
__global__ void __launch_bounds__(FT) k_mlp_dense(FusedArgs a) {
  // one LDS object per buffer and the stage loop unrolled over the buffers,
  // so each DMA and each read names its buffer at compile time
  __shared__ __attribute__((aligned(16))) __bf16 sA0[FM * DKS * 16], sA1[FM * DKS * 16], sA2[FM * DKS * 16],
      sA3[FM * DKS * 16];
  __shared__ __attribute__((aligned(16))) lac_bf16x8 sB0[DKS][FN / 32][64], sB1[DKS][FN / 32][64],
      sB2[DKS][FN / 32][64], sB3[DKS][FN / 32][64];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;  // rows wm*128, outputs wn*64
  const int nb = gridDim.x, b = blockIdx.x;
  const int lin = (b & 7) * (nb >> 3) + (b >> 3);  // consecutive items on one XCD
  const int rt = lin / a.nN, ntile = lin % a.nN;
  if (rt >= a.nrt) return;  // the whole block
  const long rowb = (long)rt * FM;  // within the launch
  const int n0 = ntile * FN;
  const int nst = a.nks / DKS;  // nks: a multiple of DKS (K a multiple of 32)

  lac_f32x16 acc[4][2];
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[x][y][r] = 0.f;

  // this lane's share of a stage: A rows g*16 + lane/4 of row groups g =
  // 2*wave + q, chunk slot lane%4; B fragment lane of (k-step, n-tile) pairs
  // p = 2*wave + q
  const __bf16 *asrc[2];
  const lac_bf16x8 *bsrc[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int row = (2 * wave + q) * 16 + (lane >> 2);
    const int c = (lane & 3) ^ ((row >> 1) & 3);
    asrc[q] = a.A + (rowb + row) * a.lda + c * 8;
    const int pr = 2 * wave + q, t = pr >> 3, fr = pr & 7;
    bsrc[q] = a.Bp + ((long)t * a.bnt + (n0 >> 5) + fr) * 64 + lane;
  }
  auto issue = [&](int st, __bf16 *dA, lac_bf16x8 (*dB)[FN / 32][64]) {
#pragma unroll
    for (int q = 0; q < 2; q++) {
      dma16(asrc[q] + st * (DKS * 16), &dA[(2 * wave + q) * 16 * (DKS * 16)]);
      const int pr = 2 * wave + q;
      dma16(bsrc[q] + (long)st * DKS * a.bnt * 64, &dB[pr >> 3][pr & 7][0]);
    }
  };
  auto compute = [&](const __bf16 *cA, const lac_bf16x8 (*cB)[FN / 32][64]) {
#pragma unroll
    for (int t = 0; t < DKS; t++) {
      lac_bf16x8 av[4], bv[2];
#pragma unroll
      for (int x = 0; x < 4; x++) {
        const int m = wm * 128 + 32 * x + (lane & 31);
        const int c = (2 * t + (lane >> 5)) ^ ((m >> 1) & 3);
        av[x] = *reinterpret_cast<const lac_bf16x8 *>(&cA[m * (DKS * 16) + c * 8]);
      }
#pragma unroll
      for (int y = 0; y < 2; y++) bv[y] = cB[t][wn * 2 + y][lane];
#pragma unroll
      for (int x = 0; x < 4; x++)
#pragma unroll
        for (int y = 0; y < 2; y++)
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
  };
  // stage st from its buffer; its loads have landed when at most the later
  // stages' (four per stage) are outstanding; stage st+3 then goes to the
  // buffer of stage st-1, which every wave finished reading before the
  // barrier
  auto stage = [&](int st, const __bf16 *cA, const lac_bf16x8 (*cB)[FN / 32][64], __bf16 *nA,
                   lac_bf16x8 (*nB)[FN / 32][64]) {
    if (st + 2 < nst)
      wait_vm<8>();
    else if (st + 1 < nst)
      wait_vm<4>();
    else
      wait_vm<0>();
    // a bare barrier: the explicit wait above already covers this wave's
    // DMA, and a workgroup fence would wait for the later stages' too
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
    compute(cA, cB);
    if (st + 3 < nst) issue(st + 3, nA, nB);
  };

  // only stages below nst are loaded or computed (K may be one or a few stages)
  issue(0, sA0, sB0);
  if (nst > 1) issue(1, sA1, sB1);
  if (nst > 2) issue(2, sA2, sB2);
  for (int st = 0; st < nst; st += DNB) {
    stage(st, sA0, sB0, sA3, sB3);
    if (st + 1 < nst) stage(st + 1, sA1, sB1, sA0, sB0);
    if (st + 2 < nst) stage(st + 2, sA2, sB2, sA1, sB1);
    if (st + 3 < nst) stage(st + 3, sA3, sB3, sA2, sB2);
  }

  // epilogue: D[m][n] of tile (x, y): n = lane%32, m = 8*(r/4) + 4*(lane/32) + r%4
#pragma unroll
  for (int x = 0; x < 4; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) {
      const int n = n0 + wn * 64 + 32 * y + (lane & 31);
      if (n >= a.N) continue;
      const float bn = a.bias[n];
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int m = wm * 128 + 32 * x + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const long row = rowb + m;
        const float v = acc[x][y][r] + bn;
        if (a.H)
          a.H[row * a.ldh + n] = (__bf16)(v > 0.f ? v : 0.f);
        else
          a.Y[row * a.ldy + n] = v;
      }
    }
}

// Layers 2 and 3 in one kernel when the hidden width is TH = 512 and layer 3
// has at most 256 outputs.  A block owns TM = 128 rows.  Layer 2
// (h2 = relu(h1 W2 + b2), 128 x 512) streams its operands through LDS as
// k_mlp_dense does: stages of two k-steps by global_load_lds, three buffers,
// two stages ahead, one block barrier per stage; eight waves of 64 rows x
// 128 outputs.  Its bf16 results stay in LDS (the stage buffers lie inside
// that tile, which is written only after the last stage has been read), and
// layer 3 (y = h2 W3 + b3) takes its A operand from there and its weight
// fragments from global memory, four k-steps ahead; eight waves of 64 rows x
// 64 outputs.  Every accumulator sees its MFMAs in k_mlp_dense's k order, so
// the results are bit-identical to the two launches; h1 is read once and h2
// never reaches HBM.
constexpr int TM = 128, TH = 512, TKS = TH / 16;     // rows, hidden width, k-steps
constexpr int TSA = TM * 32;                          // bf16 of one A stage
constexpr int TSB = 2 * (TH / 32) * 64;               // 16-byte fragment lanes of one B stage
constexpr int TBUF = TSA + TSB * 8;                   // bf16 of one stage buffer
static_assert(3 * TBUF <= TM * TH, "stage buffers inside the h2 tile");

struct TailArgs {
  const lac_bf16x8 *Bp2, *Bp3;  // packed fragments: layer 2 (16 n-tiles), layer 3 (8 n-tiles)
  const float *b2, *b3;
  const __bf16 *A;              // h1 of the launch's rows, row-major, TH wide
  float *Y;                     // layer 3 output (ldy)
  long ldy;
  int N3;                       // outputs of layer 3 (<= 256)
  int nrt;                      // row tiles of TM
};

// Chunk swizzle of row m of k_mlp_tail's h2 tile: bits 0 and 2 of m go to
// bits 2 and 3, bits 1 and 3 to bits 0 and 1.
__device__ __forceinline__ int h2_swz(int m) {
  return ((m & 1) << 2) | (((m >> 2) & 1) << 3) | ((m >> 1) & 1) | (((m >> 3) & 1) << 1);
}

__global__ void __launch_bounds__(FT) k_mlp_tail(TailArgs a) {
  // h2 tile [TM][TH] bf16; the 16-byte chunk c of row m is stored at
  // c ^ h2_swz(m), a permutation of the bits of m & 15: the layer-3
  // A-fragment reads of 16 consecutive rows hit 16 distinct chunks, and the
  // four rows m, m+1, m+4, m+5 (m even) that one h2 store instruction
  // writes land in four different groups of four chunks, i.e. distinct banks
  __shared__ __attribute__((aligned(16))) __bf16 sH[TM * TH];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nb = gridDim.x, b = blockIdx.x;
  const int rt = (b & 7) * (nb >> 3) + (b >> 3);  // consecutive tiles on one XCD
  if (rt >= a.nrt) return;                        // the whole block
  const long rowb = (long)rt * TM;
  const int wm = wave >> 2, wn = wave & 3;

  // ---- layer 2: rows wm*64 + 32x, outputs wn*128 + 32y
  lac_f32x16 acc[2][4];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[x][y][r] = 0.f;
  // this lane's share of a stage: A row wave*16 + lane/4, chunk slot lane%4
  // (holding chunk slot ^ ((row >> 1) & 3) of the row's 32 k); B fragments
  // p = 4*wave + q (k-step p/16 of the stage, n-tile p%16)
  const int arow = wave * 16 + (lane >> 2);
  const __bf16 *asrc = a.A + (rowb + arow) * TH + ((lane & 3) ^ ((arow >> 1) & 3)) * 8;
  auto issue = [&](int st, int s) {
    __bf16 *base = sH + s * TBUF;
    dma16(asrc + st * 32, base + wave * 16 * 32);
    lac_bf16x8 *bb = reinterpret_cast<lac_bf16x8 *>(base + TSA);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int p = 4 * wave + q;
      dma16(a.Bp2 + ((long)(2 * st + (p >> 4)) * (TH / 32) + (p & 15)) * 64 + lane, bb + p * 64);
    }
  };
  auto compute = [&](int s) {
    const __bf16 *cA = sH + s * TBUF;
    const lac_bf16x8 *cB = reinterpret_cast<const lac_bf16x8 *>(cA + TSA);
#pragma unroll
    for (int t = 0; t < 2; t++) {
      lac_bf16x8 av[2], bv[4];
#pragma unroll
      for (int x = 0; x < 2; x++) {
        const int m = wm * 64 + 32 * x + (lane & 31);
        const int c = (2 * t + (lane >> 5)) ^ ((m >> 1) & 3);
        av[x] = *reinterpret_cast<const lac_bf16x8 *>(&cA[m * 32 + c * 8]);
      }
#pragma unroll
      for (int y = 0; y < 4; y++) bv[y] = cB[(t * 16 + wn * 4 + y) * 64 + lane];
#pragma unroll
      for (int x = 0; x < 2; x++)
#pragma unroll
        for (int y = 0; y < 4; y++)
          acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
  };
  constexpr int NST = TKS / 2;  // stages
  issue(0, 0);
  issue(1, 1);
#pragma unroll
  for (int st = 0; st < NST; st++) {
    // stage st has landed when at most stage st+1's five copies are pending
    if (st + 1 < NST)
      wait_vm<5>();
    else
      wait_vm<0>();
    __asm__ volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
    compute(st % 3);
    // into the buffer of stage st-1, which every wave finished before the barrier
    if (st + 2 < NST) issue(st + 2, (st + 2) % 3);
  }

  // ---- layer 3 weights: n-tiles wn*2 + y, the first four k-steps now
  const int nt0 = wn * 2;
  const bool live0 = nt0 * 32 < a.N3, live1 = (nt0 + 1) * 32 < a.N3;  // uniform per wave
  lac_bf16x8 bq[4][2];
  auto bload = [&](int t, int slot) {
    if (live0) bq[slot][0] = a.Bp3[((long)t * 8 + nt0) * 64 + lane];
    if (live1) bq[slot][1] = a.Bp3[((long)t * 8 + nt0 + 1) * 64 + lane];
  };
#pragma unroll
  for (int t = 0; t < 4; t++) bload(t, t);

  // ---- h2 = relu(acc + b2) as bf16 into the tile, once every wave is done
  // with the stage buffers (its reads of them have fed its MFMAs; a bare
  // barrier, so the weight loads above stay in flight).  Lanes 2q and 2q+1
  // hold outputs n and n+1 of the same rows; they trade one value per pair
  // of rows (a DPP swap within the lane pair), so that the even lane stores
  // both outputs of the even row and the odd lane both of the odd row, each
  // as one 32-bit word (two lanes writing the halves of one word cost LDS
  // bank conflicts).
  __asm__ volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  __asm__ volatile("" ::: "memory");
  const int odd = lane & 1;
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 4; y++) {
      const int n = wn * 128 + 32 * y + (lane & 31);
      const int ne = n & ~1;  // the pair's even output
      const float bn = a.b2[n];
#pragma unroll
      for (int r = 0; r < 16; r += 2) {
        float v0 = acc[x][y][r] + bn, v1 = acc[x][y][r + 1] + bn;
        v0 = v0 > 0.f ? v0 : 0.f;
        v1 = v1 > 0.f ? v1 : 0.f;
        // even lane sends its row r+1 value, odd lane its row r value
        const float send = odd ? v0 : v1;
        const float got = __int_as_float(
            __builtin_amdgcn_update_dpp(0, __float_as_int(send), 0xB1 /* quad_perm 1,0,3,2 */, 0xF, 0xF, false));
        const float lo = odd ? got : v0, hi = odd ? v1 : got;
        const int rr = r + odd;
        const int m = wm * 64 + 32 * x + 8 * (rr >> 2) + 4 * (lane >> 5) + (rr & 3);
        const __bf16 blo = (__bf16)lo, bhi = (__bf16)hi;
        const unsigned word = (unsigned)__builtin_bit_cast(unsigned short, blo) |
                              ((unsigned)__builtin_bit_cast(unsigned short, bhi) << 16);
        *reinterpret_cast<unsigned *>(&sH[m * TH + (((ne >> 3) ^ h2_swz(m)) << 3) + (ne & 7)]) = word;
      }
    }
  __syncthreads();

  // ---- layer 3: rows wm*64 + 32x, n-tiles wn*2 + y
  lac_f32x16 acc3[2][2];
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc3[x][y][r] = 0.f;
#pragma unroll
  for (int t = 0; t < TKS; t++) {
    lac_bf16x8 cur0 = bq[t & 3][0], cur1 = bq[t & 3][1];
    if (t + 4 < TKS) bload(t + 4, t & 3);
    lac_bf16x8 av[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const int m = wm * 64 + 32 * x + (lane & 31);
      const int c = (2 * t + (lane >> 5)) ^ h2_swz(m);
      av[x] = *reinterpret_cast<const lac_bf16x8 *>(&sH[m * TH + c * 8]);
    }
#pragma unroll
    for (int x = 0; x < 2; x++) {
      if (live0) acc3[x][0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[x], cur0, acc3[x][0], 0, 0, 0);
      if (live1) acc3[x][1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[x], cur1, acc3[x][1], 0, 0, 0);
    }
  }
#pragma unroll
  for (int x = 0; x < 2; x++)
#pragma unroll
    for (int y = 0; y < 2; y++) {
      const int n = (nt0 + y) * 32 + (lane & 31);
      if (n >= a.N3) continue;
      const float bn = a.b3[n];
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const int m = wm * 64 + 32 * x + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        a.Y[(rowb + m) * a.ldy + n] = acc3[x][y][r] + bn;
      }
    }
}

// Training targets at columns cols[r]: ordinary-substep result (the state)
// minus the long-substep result (fu, fv, fw, fpai).
template <typename T>
__global__ void __launch_bounds__(256) k_lac_target(Dev<T> D, const int *cols, int nrows, const T *fu, const T *fv,
                                                    const T *fw, const double *fpai, float *Tt) {
  const Geo &g = D.g;
  const int nk = g.nk, nout = LAC_OUT_FIELDS * nk;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nrows * nout) return;
  const int r = (int)(t / nout), o = (int)(t % nout);
  const int k = o % nk + 1, f = o / nk;
  const int c = cols[r];
  const long x = g.at(c % g.nj + 1, c / g.nj + 1, k);
  double d;
  switch (f) {
    case 0: d = (double)D.u[x] - (double)fu[x]; break;
    case 1: d = (double)D.v[x] - (double)fv[x]; break;
    case 2: d = (double)D.w[x + g.plane] - (double)fw[x + g.plane]; break;
    default: d = D.pai[x] - fpai[x]; break;
  }
  Tt[t] = (float)d;
}

// Add the de-normalised network outputs Y (tile-ordered rows, row stride
// nout) to u, v, w and the Exner function; 64 rows (one j-run) per block,
// transposed through LDS so both the Y reads and the field writes coalesce.
template <typename T>
__global__ void __launch_bounds__(256) k_lac_apply(Dev<T> D, LacRows R, const float *Y, const float *omean,
                                                   const float *ostd) {
  __shared__ float tile[64][65];
  const Geo &g = D.g;
  const int nk = g.nk, nout = LAC_OUT_FIELDS * nk;
  const long r0 = (long)blockIdx.x * 64;
  const int tr = threadIdx.x & 63, tq = threadIdx.x >> 6;
  int j, i;
  R.colof(r0 + tr, g, j, i);
  const bool ok = j <= g.nj && i <= g.ni;
  for (int o0 = 0; o0 < nout; o0 += 64) {
    for (int q = tq; q < 64; q += 4) {
      const int o = o0 + tr;
      tile[q][tr] = o < nout ? Y[(r0 + q) * nout + o] * ostd[o] + omean[o] : 0.f;
    }
    __syncthreads();
    if (ok)
      for (int q = tq; q < 64; q += 4) {
        const int o = o0 + q;
        if (o >= nout) break;
        const float d = tile[tr][q];
        const int f = o / nk, k = o % nk + 1;
        const long x = g.at(j, i, k);
        switch (f) {
          case 0: D.u[x] = D.u[x] + (T)d; break;
          case 1: D.v[x] = D.v[x] + (T)d; break;
          case 2: D.w[x + g.plane] = D.w[x + g.plane] + (T)d; break;
          default: D.pai[x] = D.pai[x] + (double)d; break;
        }
      }
    __syncthreads();
  }
}

// Per-column mean and standard deviation of a row-major fp32 matrix.
__global__ void __launch_bounds__(256) k_colstats(const float *X, long rows, int cols, double *mean, double *sd) {
  const int o = blockIdx.x;
  double s = 0.0, s2 = 0.0;
  for (long r = threadIdx.x; r < rows; r += 256) {
    const double v = X[r * cols + o];
    s += v;
    s2 += v * v;
  }
  __shared__ double a[256], b[256];
  a[threadIdx.x] = s;
  b[threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) a[threadIdx.x] += a[threadIdx.x + w], b[threadIdx.x] += b[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double m = a[0] / (double)rows, var = b[0] / (double)rows - m * m;
    mean[o] = m;
    sd[o] = sqrt(var > 0.0 ? var : 0.0);
  }
}

// Training batch: rows idx[r] of the raw set, normalised (bf16 inputs, fp32
// targets).
__global__ void __launch_bounds__(256) k_lac_batch(const float *Xraw, const float *Traw, const int *idx, int nin,
                                                   int nout, const float *imean, const float *irstd,
                                                   const float *omean, const float *ostd, __bf16 *Xb, float *Tn) {
  const long r = blockIdx.y;
  const long s = idx[r];
  for (int o = blockIdx.x * 256 + threadIdx.x; o < nin; o += gridDim.x * 256)
    Xb[r * nin + o] = (__bf16)((Xraw[s * nin + o] - imean[o]) * irstd[o]);
  for (int o = blockIdx.x * 256 + threadIdx.x; o < nout; o += gridDim.x * 256)
    Tn[r * nout + o] = (Traw[s * nout + o] - omean[o]) / ostd[o];
}

// Accumulate sum (a - b)^2 and sum b^2 over the interior, levels k0..k1.
template <typename F>
__global__ void __launch_bounds__(256) k_diffnorm(const F *a, const F *b, Geo g, int k0, int k1, double *out) {
  __shared__ double s1[256], s2[256];
  const int j = blockIdx.x * 64 + threadIdx.x % 64 + 1;
  const int i = blockIdx.y * 4 + threadIdx.x / 64 + 1;
  double d2 = 0.0, b2 = 0.0;
  if (j <= g.nj && i <= g.ni)
    for (int k = k0; k <= k1; k++) {
      const long x = g.at(j, i, k);
      const double d = (double)a[x] - (double)b[x];
      d2 += d * d;
      b2 += (double)b[x] * (double)b[x];
    }
  s1[threadIdx.x] = d2;
  s2[threadIdx.x] = b2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) s1[threadIdx.x] += s1[threadIdx.x + w], s2[threadIdx.x] += s2[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], s1[0]);
    atomicAdd(&out[1], s2[0]);
  }
}

// ============================================================= diagnostics
// :395-419 -- extrema of ps, u, v over the interior.  Extrema skip NaN
// entries and are NaN only when every entry is NaN.
struct Ext {
  double mx[3], mn[3];
  unsigned long long cnt[3];
};
__device__ __forceinline__ void ext_add(Ext &e, int f, double x) {
  if (x == x) {
    e.mx[f] = fmax(e.mx[f], x);
    e.mn[f] = fmin(e.mn[f], x);
    e.cnt[f]++;
  }
}
template <typename T>
__global__ void __launch_bounds__(256) k_diag_partial(Dev<T> D) {
  __shared__ double smx[3][256], smn[3][256];
  __shared__ unsigned long long scnt[3][256];
  const Geo &g = D.g;
  int tid = threadIdx.y * 64 + threadIdx.x;
  int j = blockIdx.x * 64 + threadIdx.x + 1;
  int i = blockIdx.y * 4 + threadIdx.y + 1;
  Ext e;
  for (int f = 0; f < 3; f++) {
    e.mx[f] = -INFINITY;
    e.mn[f] = INFINITY;
    e.cnt[f] = 0;
  }
  if (j <= g.nj && i <= g.ni) {
    ext_add(e, 0, (double)D.ps[g.at2(j, i)]);
    for (int k = 1; k <= g.nk; k++) {
      long x = g.at(j, i, k);
      ext_add(e, 1, (double)D.u[x]);
      ext_add(e, 2, (double)D.v[x]);
    }
  }
  for (int f = 0; f < 3; f++) {
    smx[f][tid] = e.mx[f];
    smn[f][tid] = e.mn[f];
    scnt[f][tid] = e.cnt[f];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s)
      for (int f = 0; f < 3; f++) {
        smx[f][tid] = fmax(smx[f][tid], smx[f][tid + s]);
        smn[f][tid] = fmin(smn[f][tid], smn[f][tid + s]);
        scnt[f][tid] += scnt[f][tid + s];
      }
    __syncthreads();
  }
  if (tid == 0) {
    long b = blockIdx.y * gridDim.x + blockIdx.x;
    double *o = D.diag + b * 9;
    for (int f = 0; f < 3; f++) {
      o[f] = smx[f][0];
      o[3 + f] = smn[f][0];
      o[6 + f] = (double)scnt[f][0];
    }
  }
}
__global__ void __launch_bounds__(256) k_diag_final(double *partial, long nblk, double *out) {
  __shared__ double smx[3][256], smn[3][256], scn[3][256];
  int tid = threadIdx.x;
  double mx[3] = {-INFINITY, -INFINITY, -INFINITY}, mn[3] = {INFINITY, INFINITY, INFINITY}, cn[3] = {0, 0, 0};
  for (long b = tid; b < nblk; b += 256)
    for (int f = 0; f < 3; f++) {
      mx[f] = fmax(mx[f], partial[b * 9 + f]);
      mn[f] = fmin(mn[f], partial[b * 9 + 3 + f]);
      cn[f] += partial[b * 9 + 6 + f];
    }
  for (int f = 0; f < 3; f++) {
    smx[f][tid] = mx[f];
    smn[f][tid] = mn[f];
    scn[f][tid] = cn[f];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s)
      for (int f = 0; f < 3; f++) {
        smx[f][tid] = fmax(smx[f][tid], smx[f][tid + s]);
        smn[f][tid] = fmin(smn[f][tid], smn[f][tid + s]);
        scn[f][tid] += scn[f][tid + s];
      }
    __syncthreads();
  }
  if (tid == 0) {
    for (int f = 0; f < 3; f++) {
      out[2 * f] = scn[f][0] > 0 ? smx[f][0] : NAN;
      out[2 * f + 1] = scn[f][0] > 0 ? smn[f][0] : NAN;
    }
    out[6] = scn[0][0];  // ps values that are not NaN
  }
}

// ============================================================ case builder
// The 3-D part of the case builder on the device (the host keeps the levels,
// the 2-D geometry and ps): vertical coordinate (mod_zita.F90:87-179),
// temperature, humidity and wind, the discrete hydrostatic Exner recursion
// (mod_atmosphere.F90:210-234) and the perturbations, one column per lane,
// all in FP64 as on the host; the splitmix64 draws of the theta noise are
// taken by their index in the host's loop order.
struct CaseDev {
  int kind, noise_levels;
  double ztop, zh, a0, tsfc, lapse, ttrop, psfc, rh, u0, v0, jet_amp, jet_z, jet_width;
  double noise_amp, bubble_amp, bubble_rcells, bubble_z, bubble_rz, dzita;
  unsigned long long rng;
  const double *zita, *zitah, *ht, *ps;  // ht, ps in the device 2-D layout
};

__device__ __forceinline__ double cd_gzita(double z, const CaseDev &c) {
  const double r = z / c.ztop;
  return ((0.0 - 1.0 * c.a0) * r - (3.0 - 2.0 * c.a0) * (r * r) + (2.0 - 1.0 * c.a0) * (r * r * r)) + 1.0;
}
__device__ __forceinline__ double cd_zeta(double z, double geopot, const CaseDev &c) {
  const double zfz = c.ztop / (exp(c.ztop / c.zh) - 1.0);
  return geopot * kRegrav * (cd_gzita(z, c) - 1.0) + zfz * (exp(z / c.zh) - 1.0);
}
__device__ __forceinline__ double cd_fmz(double z, double geopot, const CaseDev &c) {
  const double r = z / c.ztop, zfz = c.ztop / (exp(c.ztop / c.zh) - 1.0);
  const double gp = ((0.0 - 1.0 * c.a0) * 1.0 - (6.0 - 4.0 * c.a0) * r + (6.0 - 3.0 * c.a0) * (r * r)) / c.ztop;
  return 1.0 / (gp * (geopot * kRegrav) + zfz * exp(z / c.zh) / c.zh);
}
__device__ __forceinline__ double cd_T(double z, const CaseDev &c) { return fmax(c.tsfc - c.lapse * z, c.ttrop); }
__device__ __forceinline__ double cd_p(double z, const CaseDev &c) {
  const double expo = kEgrav / (kRgas * c.lapse), zt = (c.tsfc - c.ttrop) / c.lapse;
  if (z <= zt) return c.psfc * pow(cd_T(z, c) / c.tsfc, expo);
  return c.psfc * pow(c.ttrop / c.tsfc, expo) * exp(-kEgrav * (z - zt) / (kRgas * c.ttrop));
}
// temperature and humidity before the perturbations that need pai
__device__ __forceinline__ void cd_tq(const CaseDev &c, const Geo &g, int j, int i, int k, double geopot, double &t,
                                      double &q) {
  const double zeta = cd_zeta(c.zitah[k - 1], geopot, c);
  if (c.kind == MLC_CASE_REFERENCE) {
    // fill_atmosphere (mod_atmosphere.F90:171-209)
    const double ts = 300.0, qs0 = 0.01865, tropoh = 15000.0;
    const double tvs = ts * (1.0 + kEp1 * qs0), tvt = tvs - kLrate * tropoh;
    if (zeta < tropoh) {
      q = qs0 * exp(-zeta / 4000.0) * exp((-zeta / 7500.0) * (-zeta / 7500.0));
      t = (tvs - kLrate * zeta) / (1.0 + q);
    } else {
      q = 1.0e-8;
      t = tvt;
    }
    if (k == g.nk) {  // the warm disc at the lowest level (:196-203), REAL() arithmetic
      const float fjx = (float)g.nj, fiy = (float)g.ni;
      const double aa = (double)fjx / 5.0;
      const float bb = fiy / 5.0f, b2 = bb * bb;
      const double mr = sqrt(aa * aa + (double)b2) / 2.0;
      const float dj = (float)(j - g.nj / 2), di = (float)(i - g.ni / 2);
      const double md = (double)sqrtf(dj * dj + di * di);
      if (md < mr) t += (mr - md) / mr;
    }
  } else {
    const double z = zeta + geopot * kRegrav;
    t = cd_T(z, c);
    q = c.rh * d_pfwsat(t, cd_p(z, c));
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_case_build(Dev<T> D, CaseDev c) {
  COL_IDX
  const int nk = g.nk;
  const long c2 = g.at2(j, i);
  const double geo = c.ht[c2];
  // vertical coordinate, humidity, wind
  for (int k = 1; k <= nk; k++) {
    const long x = g.at(j, i, k);
    D.zeta[x] = (T)cd_zeta(c.zitah[k - 1], geo, c);
    D.fmz[x] = (T)cd_fmz(c.zitah[k - 1], geo, c);
    double t, q;
    cd_tq(c, g, j, i, k, geo, t, q);
    D.qv[x] = (T)q;
    double uu = 1.0, vv = 1.0;
    if (c.kind != MLC_CASE_REFERENCE) {
      // u on the west face of cell j: heights averaged over j-1 and j
      const double gw = c.ht[g.at2(j - 1, i)];
      const double zf = 0.5 * ((cd_zeta(c.zitah[k - 1], gw, c) + gw * kRegrav) +
                               (cd_zeta(c.zitah[k - 1], geo, c) + geo * kRegrav));
      uu = c.u0;
      if (c.jet_amp != 0.0) {
        const double e = (zf - c.jet_z) / c.jet_width;
        uu += c.jet_amp * exp(-e * e);
      }
      vv = c.v0;
    }
    D.u[x] = (T)uu;
    D.v[x] = (T)vv;
  }
  for (int k = 1; k <= nk + 1; k++) D.fmzf[g.at(j, i, k)] = (T)cd_fmz(c.zita[k - 1], geo, c);
  // discrete hydrostatic Exner function, bottom up
  double tk, qk;
  cd_tq(c, g, j, i, nk, geo, tk, qk);
  double pk;
  {
    double tm, qm;
    cd_tq(c, g, j, i, nk - 1, geo, tm, qm);
    const double zk = cd_zeta(c.zitah[nk - 1], geo, c), zm = cd_zeta(c.zitah[nk - 2], geo, c);
    const double zdelta = zk * kEgrav;
    const double tv1 = tk * (1.0 + kEp1 * qk), tv2 = tm * (1.0 + kEp1 * qm);
    double lrt = (tv2 - tv1) / (zm - zk);
    lrt = 0.65 * lrt - 0.35 * kLrate;
    const double tv = tv1 - 0.5 * zk * lrt;
    const double zz = 1.0 / (kRgas * tv);
    const double p = c.ps[c2] * exp(-zdelta * zz);
    pk = pow(p / kP00, kRovcp);
    D.pai[g.at(j, i, nk)] = pk;
  }
  for (int k = nk - 1; k >= 1; k--) {
    double t, q;
    cd_tq(c, g, j, i, k, geo, t, q);
    const double tv1 = t * (1.0 + kEp1 * q), tv2 = tk * (1.0 + kEp1 * qk);
    const double ff = cd_fmz(c.zita[k], geo, c);  // fmzf at level k+1
    const double zb = 2.0 * kEgrav * c.dzita / (ff * kCpd) + tv1 - tv2;
    const double zdelta = sqrt(zb * zb + 4.0 * tv2 * tv1);
    pk = -pk / (2.0 * tv2) * (zb - zdelta);
    D.pai[g.at(j, i, k)] = pk;
    tk = t;
    qk = q;
  }
  // temperature with the perturbations applied as T' = theta' * pai
  const int k0 = nk - c.noise_levels + 1 < 1 ? 1 : nk - c.noise_levels + 1;
  const double jc = 0.5 * (double)(g.nj + 1), ic = 0.5 * (double)(g.ni + 1);
  for (int k = 1; k <= nk; k++) {
    const long x = g.at(j, i, k);
    double t, q;
    cd_tq(c, g, j, i, k, geo, t, q);
    if (c.kind != MLC_CASE_REFERENCE) {
      const double pk2 = D.pai[x];
      if (c.noise_amp > 0.0 && k >= k0) {
        const unsigned long long n = ((unsigned long long)(k - k0) * g.ni + (i - 1)) * g.nj + (j - 1);
        unsigned long long z = c.rng + (n + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        const double uni = (double)(z >> 11) * (1.0 / 9007199254740992.0);
        t += (-c.noise_amp + (c.noise_amp - -c.noise_amp) * uni) * pk2;
      }
      if (c.bubble_amp != 0.0) {
        const double dh = sqrt(((double)j - jc) * ((double)j - jc) + ((double)i - ic) * ((double)i - ic)) /
                          c.bubble_rcells;
        const double dz = (cd_zeta(c.zitah[k - 1], geo, c) - c.bubble_z) / c.bubble_rz;
        const double r = sqrt(dh * dh + dz * dz);
        if (r < 1.0) {
          const double cc = cos(0.5 * kMathpi * r);
          t += c.bubble_amp * cc * cc * pk2;
        }
      }
    }
    D.t[x] = t;
  }
}

// Every halo cell of the two-point ring from its periodic partner.
template <typename F>
__global__ void __launch_bounds__(256) k_periodic(F *f, Geo g) {
  const int e = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y + 1;
  const int wj = g.nj + 4, wi = g.ni + 4;
  if (e >= wj * wi) return;
  const int j = e % wj - 1, i = e / wj - 1;
  if (j >= 1 && j <= g.nj && i >= 1 && i <= g.ni) return;
  const int jj = ((j - 1) % g.nj + g.nj) % g.nj + 1, ii = ((i - 1) % g.ni + g.ni) % g.ni + 1;
  f[g.at(j, i, k)] = f[g.at(jj, ii, k)];
}

// ============================================================ output packing
// Interior points of a field, j fastest, as doubles: the layout of the
// reference's output records (nompi/mod_mppparam.F90:197,310).
// Debug: a quiet NaN into every cell of the two-point halo ring of level k
// (rows i = -1, 0, ni+1, ni+2 and columns j = -1, 0, nj+1, nj+2).
template <typename F>
__global__ void __launch_bounds__(256) k_poison(F *f, Geo g) {
  const int t = blockIdx.x * 256 + threadIdx.x, k = blockIdx.y + 1;
  const int ncol = 4 * (g.ni + 4), nrow = 4 * g.nj;
  int j, i;
  if (t < ncol) {
    const int c = t % 4;
    i = t / 4 - 1;
    j = c < 2 ? c - 1 : g.nj + c - 1;
  } else if (t < ncol + nrow) {
    const int r = (t - ncol) % 4;
    j = (t - ncol) / 4 + 1;
    i = r < 2 ? r - 1 : g.ni + r - 1;
  } else {
    return;
  }
  f[g.at(j, i, k)] = (F)__builtin_nan("");
}

template <typename F>
__global__ void __launch_bounds__(256) k_pack_interior(const F *src, Geo g, double *dst) {
  const int j = blockIdx.x * 64 + threadIdx.x % 64 + 1;
  const int i = blockIdx.y * 4 + threadIdx.x / 64 + 1;
  const int k = blockIdx.z + 1;
  if (j > g.nj || i > g.ni) return;
  dst[((long)(k - 1) * g.ni + (i - 1)) * g.nj + (j - 1)] = (double)src[g.at(j, i, k)];
}

// ============================================================ host driver
// Kernel kinds for the per-launch event timing of mlc_profile_*.
enum KernelKind {
  KK_HALO = 0, KK_PROLOGUE, KK_SOUND_A, KK_SOUND_B, KK_SOUND_B2, KK_SOUND_C, KK_SMOOTH5, KK_FILTER_BRACKET,
  KK_ADV_IN, KK_WAF_V, KK_WAF_M, KK_WAF_Z, KK_WAF_H, KK_ADV_OUT, KK_EPILOGUE, KK_DIAG,
  KK_LAC_GATHER, KK_LAC_MLP, KK_LAC_APPLY, KK_LAC_COPY, KK_COUNT
};
static const char *kKernelNames[KK_COUNT] = {
    "halo", "prologue", "sound_a", "sound_b", "sound_b2", "sound_c", "smooth5", "filter_bracket",
    "adv_in", "waf_v", "waf_m", "waf_z", "waf_h", "adv_out", "epilogue", "diag",
    "lac_gather", "lac_mlp", "lac_apply", "lac_copy"};

struct CoreBase {
  virtual ~CoreBase() {}
  virtual int set_field(int f, const double *host, const int *lb, const int *ub) = 0;
  virtual int get_field(int f, double *host, const int *lb, const int *ub) = 0;
  virtual int init() = 0;
  virtual int step(int istep, int lprint, mlc_diag *diag) = 0;
  virtual int sync() = 0;
  virtual int poison_halos(int enable) = 0;
  virtual int profile(int enable) = 0;
  virtual int profile_get(int kind, double *total_ms, long long *launches, double *bytes) = 0;
  virtual int lac_create(const mlc_lac_config &c) = 0;
  virtual int lac_collect(int nsteps) = 0;
  virtual int lac_train(int iters, int batch, double lr, double *loss) = 0;
  virtual int lac_mode(int mode) = 0;
  virtual int lac_stats(mlc_lac_stats *s) = 0;
  virtual mlc_mlp *lac_network() = 0;
  virtual int lac_dataset(long first, int n, float *X, float *T, int *cols) = 0;
  // ps then t (interior, doubles) into dst on the handle's stream; done is
  // recorded after them; wait_event makes the stream wait for an event
  virtual int pack_output(double *dst, hipEvent_t done) = 0;
  virtual int wait_event(hipEvent_t e) = 0;
  virtual int build_case(const mlc_case &c, const CaseHost &H) = 0;
};

static long round_up(long a, long b) { return (a + b - 1) / b * b; }

template <typename T>
struct Core : CoreBase {
  mlc_dims dims;
  mlc_params prm;
  Geo g;
  hipStream_t stream = nullptr;
  Dev<T> D{};
  std::vector<void *> allocs;
  std::vector<double> zita, zitah, ffilt;                   // host levels
  std::vector<double> h_msfu, h_msfv, h_msfx, h_ulat, h_vlat;  // host geometry
  bool initialised = false;
  bool unfused_advection = false;  // params.unfused: the three-sweep wafone kernels
  T *gwz[MAXF] = {}, *gout[MAXF] = {};  // scratch of the grouped advection
  double *dgout[2] = {};                // results of its FP64 group (theta, pi) in the FP32 core
  long nblk_diag = 0;
  double *d_diag_out = nullptr;

  // ---------------------------------------------- per-kernel event timing
  struct ProfRec {
    int kind;
    double bytes;
    hipEvent_t a, b;
  };
  bool prof = false;
  std::vector<hipEvent_t> evpool;
  size_t evnext = 0;
  std::vector<ProfRec> pending;
  double acc_ms[KK_COUNT] = {0};
  long long acc_n[KK_COUNT] = {0};
  double acc_bytes[KK_COUNT] = {0};

  // Algorithmic bytes of one launch: nT passes over kz-level fields of the
  // storage type, nD passes over FP64 fields (the Exner function).  A pass is
  // a field that must cross a grid-wide synchronisation of the step (read or
  // written once per phase); each kernel is charged the passes of its share
  // of the step: prologue 11, per acoustic substep 25 (sound_a: u, v, fmz;
  // sound_b: the other 15 of the divergence / column solve / theta-pi phase;
  // sound_c: the 7 of the momentum phase), per advection cycle the s write
  // and the u, v, w reads (adv_in), the advected fields read once with s,
  // fmz, fmzf (waf_v) and written once (waf_h), the u, v, w writes
  // (adv_out), epilogue 12.  Intermediate arrays of this implementation
  // (zdiv2, the elimination coefficients, wz, p0) are not charged.
  double fbytes(double nT, double nD) const {
    return (double)g.nj * g.ni * g.nk * (nT * sizeof(T) + nD * sizeof(double));
  }

  int resolve() {
    if (pending.empty()) return MLC_OK;
    HIP_TRY(hipEventSynchronize(pending.back().b));
    for (auto &r : pending) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
      acc_ms[r.kind] += ms;
      acc_n[r.kind] += 1;
      acc_bytes[r.kind] += r.bytes;
    }
    pending.clear();
    evnext = 0;
    return MLC_OK;
  }

  template <typename Kern, typename... Args>
  void launch(int kind, double bytes, Kern kern, dim3 grid, dim3 block, Args... args) {
    timed(kind, bytes, [&]() { hipLaunchKernelGGL(kern, grid, block, 0, stream, args...); });
  }

  // Run fn (work queued on the stream) between two timing events.
  template <typename Fn>
  void timed(int kind, double bytes, Fn fn) {
    ProfRec r{kind, bytes, nullptr, nullptr};
    if (prof) {
      if (evnext + 2 > evpool.size()) {
        if (evpool.size() < 16384) {
          for (int n = 0; n < 256; n++) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) break;
            evpool.push_back(e);
          }
        }
        if (evnext + 2 > evpool.size()) (void)resolve();
      }
      if (evnext + 2 <= evpool.size()) {
        r.a = evpool[evnext++];
        r.b = evpool[evnext++];
        (void)hipEventRecord(r.a, stream);
      }
    }
    fn();
    if (prof && r.a) {
      (void)hipEventRecord(r.b, stream);
      pending.push_back(r);
    }
  }

  int profile(int enable) override {
    int rc = resolve();
    if (rc) return rc;
    prof = enable != 0;
    for (int k = 0; k < KK_COUNT; k++) acc_ms[k] = 0, acc_n[k] = 0, acc_bytes[k] = 0;
    return MLC_OK;
  }

  int profile_get(int kind, double *total_ms, long long *launches, double *bytes) override {
    if (kind < 0 || kind >= KK_COUNT) return mlc_fail(MLC_EINVAL, "mlc_profile_get: kind out of range");
    int rc = resolve();
    if (rc) return rc;
    if (total_ms) *total_ms = acc_ms[kind];
    if (launches) *launches = acc_n[kind];
    if (bytes) *bytes = acc_bytes[kind];
    return MLC_OK;
  }

  int alloc_field(T **p, int nlev) { return alloc_bytes((void **)p, (size_t)g.plane * nlev * sizeof(T)); }
  int alloc_d(double **p, long n) { return alloc_bytes((void **)p, (size_t)n * sizeof(double)); }
  int alloc_bytes(void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return mlc_fail(MLC_ENOMEM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    allocs.push_back(*p);
    e = hipMemsetAsync(*p, 0, bytes, stream);
    if (e != hipSuccess) return mlc_fail(MLC_EHIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
    return MLC_OK;
  }

  ~Core() override {
    if (stream) (void)hipStreamSynchronize(stream);
    if (L.net) mlp_free(L.net);
    for (hipEvent_t e : evpool) (void)hipEventDestroy(e);
    for (void *p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int create(const mlc_dims &d, const mlc_params &p) {
    dims = d;
    prm = p;
    unfused_advection = p.unfused != 0;
    HIP_TRY(hipSetDevice(p.device));
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    g.nj = d.jx;
    g.ni = d.iy;
    g.nk = d.kz;
    g.pj = round_up(JOFF + d.jx + 3, 32);
    g.plane = g.pj * (long)(d.iy + 4);
    const int K = d.kz, K1 = d.kz + 1;
    int rc;
#define AF(ptr, nl) \
  if ((rc = alloc_field(&D.ptr, nl)) != MLC_OK) return rc;
    AF(u, K) AF(v, K) AF(un, K) AF(vn, K) AF(w, K1);
    AF(qv, K) AF(qc, K) AF(qi, K) AF(qr, K) AF(qs, K);
    AF(tvirt, K) AF(p, K) AF(rho, K) AF(qsat, K) AF(ux, K) AF(vx, K) AF(ps, 1);
    AF(fmz, K) AF(fmzf, K1) AF(zeta, K) AF(hx, 1) AF(hy, 1) AF(mu, 1) AF(mv, 1) AF(mx, 1);
    AF(rmu, 1) AF(rmv, 1) AF(mx2, 1) AF(coru, 1) AF(corv, 1);
    AF(s, K1) AF(deltaw, K1) AF(zdh, K) AF(wx, K) AF(qwltot, K) AF(qwitot, K);
    if (!segmented_solve() || p.divfilter) AF(zds, K);
    if (p.divfilter) AF(zdf, K);
    for (int f = 0; f < waf_gmax<T>(); f++) {
      if ((rc = alloc_field(&gwz[f], K)) != MLC_OK) return rc;
      if ((rc = alloc_field(&gout[f], K)) != MLC_OK) return rc;
    }
#undef AF
    if ((rc = alloc_d(&D.pai, g.plane * K)) != MLC_OK) return rc;
    if ((rc = alloc_d(&D.wz, g.plane * K)) != MLC_OK) return rc;
    if ((rc = alloc_d(&D.p0, g.plane * K)) != MLC_OK) return rc;
    if ((rc = alloc_d(&D.t, g.plane * K)) != MLC_OK) return rc;
    if ((rc = alloc_d(&D.tetav, g.plane * K)) != MLC_OK) return rc;
    // scratch of the optional paths, only where they run
    if (p.filterpai) {
      if ((rc = alloc_d(&D.pf, g.plane * K)) != MLC_OK) return rc;
      if ((rc = alloc_d(&D.smd, g.plane * K)) != MLC_OK) return rc;
    }
    if (p.fulleq && p.filtertheta) {
      if ((rc = alloc_d(&D.tf, g.plane * K)) != MLC_OK) return rc;
      if ((rc = alloc_d(&D.smt, g.plane * K)) != MLC_OK) return rc;
    }
    if constexpr (sizeof(T) != sizeof(double)) {
      // results of the FP64 advection group (theta, pi)
      for (int f = 0; f < 2; f++)
        if ((rc = alloc_d(&dgout[f], g.plane * K)) != MLC_OK) return rc;
    }
    if (!segmented_solve()) {  // the one-lane column solve keeps its coefficients in HBM
      if ((rc = alloc_d(&D.thg, g.plane * K1)) != MLC_OK) return rc;
      if ((rc = alloc_d(&D.thc, g.plane * K1)) != MLC_OK) return rc;
    }
    double *lv;
    if ((rc = alloc_d(&lv, 4L * K1)) != MLC_OK) return rc;
    D.gzitak = lv;
    D.gzitakh = lv + K1;
    D.xknu = lv + 2 * K1;
    D.ffilt = lv + 3 * K1;
    nblk_diag = (long)((d.jx + 63) / 64) * ((d.iy + 3) / 4);
    if ((rc = alloc_d(&D.diag, nblk_diag * 9)) != MLC_OK) return rc;
    if ((rc = alloc_d(&d_diag_out, 8)) != MLC_OK) return rc;
    D.g = g;
    zita.assign(K1, 0.0);
    zitah.assign(K, 0.0);
    ffilt.assign(K, 0.0);
    long n2 = (long)(d.jx + 4) * (d.iy + 4);
    h_msfu.assign(n2, 1.0);
    h_msfv.assign(n2, 1.0);
    h_msfx.assign(n2, 1.0);
    h_ulat.assign(n2, 0.0);
    h_vlat.assign(n2, 0.0);
    HIP_TRY(hipStreamSynchronize(stream));
    return MLC_OK;
  }

  // ---------------------------------------------------------- transfers
  struct Slot {
    void *dev;
    bool is_double;
    int nlev;
    std::vector<double> *host2d;  // geometry kept on the host as well
  };
  int slot(int f, Slot &s) {
    const int K = g.nk;
    s.host2d = nullptr;
    s.is_double = false;
    switch (f) {
      case MLC_U: s.dev = D.u, s.nlev = K; break;
      case MLC_V: s.dev = D.v, s.nlev = K; break;
      case MLC_W: s.dev = D.w, s.nlev = K + 1; break;
      case MLC_PAI: s.dev = D.pai, s.nlev = K, s.is_double = true; break;
      case MLC_T: s.dev = D.t, s.nlev = K, s.is_double = true; break;
      case MLC_TETAV: s.dev = D.tetav, s.nlev = K, s.is_double = true; break;
      case MLC_QV: s.dev = D.qv, s.nlev = K; break;
      case MLC_QC: s.dev = D.qc, s.nlev = K; break;
      case MLC_QI: s.dev = D.qi, s.nlev = K; break;
      case MLC_QR: s.dev = D.qr, s.nlev = K; break;
      case MLC_QS: s.dev = D.qs, s.nlev = K; break;
      case MLC_TVIRT: s.dev = D.tvirt, s.nlev = K; break;
      case MLC_P: s.dev = D.p, s.nlev = K; break;
      case MLC_RHO: s.dev = D.rho, s.nlev = K; break;
      case MLC_QSAT: s.dev = D.qsat, s.nlev = K; break;
      case MLC_UX: s.dev = D.ux, s.nlev = K; break;
      case MLC_VX: s.dev = D.vx, s.nlev = K; break;
      case MLC_PS: s.dev = D.ps, s.nlev = 1; break;
      case MLC_FMZ: s.dev = D.fmz, s.nlev = K; break;
      case MLC_FMZF: s.dev = D.fmzf, s.nlev = K + 1; break;
      case MLC_ZETA: s.dev = D.zeta, s.nlev = K; break;
      case MLC_HX: s.dev = D.hx, s.nlev = 1; break;
      case MLC_HY: s.dev = D.hy, s.nlev = 1; break;
      case MLC_MSFU: s.dev = D.mu, s.nlev = 1, s.host2d = &h_msfu; break;
      case MLC_MSFV: s.dev = D.mv, s.nlev = 1, s.host2d = &h_msfv; break;
      case MLC_MSFX: s.dev = D.mx, s.nlev = 1, s.host2d = &h_msfx; break;
      case MLC_ULAT: s.dev = nullptr, s.nlev = 1, s.host2d = &h_ulat; break;
      case MLC_VLAT: s.dev = nullptr, s.nlev = 1, s.host2d = &h_vlat; break;
      default: return mlc_fail(MLC_EINVAL, "unknown field id %d", f);
    }
    return MLC_OK;
  }

  // Visit the intersection of a Fortran host array (lb..ub) with the device
  // extent (-1:nj+2, -1:ni+2, 1:nlev); fn(host_index, device_index).
  template <typename Fn>
  int visit(const Slot &s, const int *lb, const int *ub, Fn fn) {
    int rank = s.nlev == 1 ? 2 : 3;
    int l[3] = {lb[0], lb[1], rank == 3 ? lb[2] : 1};
    int u[3] = {ub[0], ub[1], rank == 3 ? ub[2] : 1};
    for (int r = 0; r < rank; r++)
      if (u[r] < l[r]) return mlc_fail(MLC_EINVAL, "empty bounds in dimension %d", r + 1);
    long n0 = (long)u[0] - l[0] + 1, n1 = (long)u[1] - l[1] + 1;
    int j0 = l[0] > -1 ? l[0] : -1, j1 = u[0] < g.nj + 2 ? u[0] : g.nj + 2;
    int i0 = l[1] > -1 ? l[1] : -1, i1 = u[1] < g.ni + 2 ? u[1] : g.ni + 2;
    int k0 = l[2] > 1 ? l[2] : 1, k1 = u[2] < s.nlev ? u[2] : s.nlev;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = k0; k <= k1; k++)
      for (int i = i0; i <= i1; i++)
        for (int j = j0; j <= j1; j++) {
          long hidx = ((long)(k - l[2]) * n1 + (i - l[1])) * n0 + (j - l[0]);
          fn(hidx, g.at(j, i, k));
        }
    return MLC_OK;
  }

  int set_field(int f, const double *host, const int *lb, const int *ub) override {
    Slot s;
    int rc = slot(f, s);
    if (rc) return rc;
    if (!host || !lb || !ub) return mlc_fail(MLC_EINVAL, "mlc_set_field: NULL argument");
    if (s.host2d) {
      long pj = g.nj + 4;
      rc = visit(s, lb, ub, [&](long h, long d) {
        long jj = d % g.pj - JOFF, ii = d / g.pj - 1;
        (*s.host2d)[(ii + 1) * pj + (jj + 1)] = host[h];
      });
      if (rc) return rc;
      if (!s.dev) return MLC_OK;
    }
    size_t n = (size_t)g.plane * s.nlev;
    // a host array covering the whole device extent needs no read-back
    const int rank = s.nlev == 1 ? 2 : 3;
    const bool covers = lb[0] <= -1 && lb[1] <= -1 && ub[0] >= g.nj + 2 && ub[1] >= g.ni + 2 &&
                        (rank == 2 || (lb[2] <= 1 && ub[2] >= s.nlev));
    if (s.is_double) {
      std::vector<double> buf(n);
      if (!covers) {
        HIP_TRY(hipMemcpyAsync(buf.data(), s.dev, n * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
      }
      rc = visit(s, lb, ub, [&](long h, long d) { buf[d] = host[h]; });
      if (rc) return rc;
      HIP_TRY(hipMemcpyAsync(s.dev, buf.data(), n * sizeof(double), hipMemcpyHostToDevice, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    } else {
      std::vector<T> buf(n);
      if (!covers) {
        HIP_TRY(hipMemcpyAsync(buf.data(), s.dev, n * sizeof(T), hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
      }
      rc = visit(s, lb, ub, [&](long h, long d) { buf[d] = (T)host[h]; });
      if (rc) return rc;
      HIP_TRY(hipMemcpyAsync(s.dev, buf.data(), n * sizeof(T), hipMemcpyHostToDevice, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    }
    return MLC_OK;
  }

  int get_field(int f, double *host, const int *lb, const int *ub) override {
    Slot s;
    int rc = slot(f, s);
    if (rc) return rc;
    if (!host || !lb || !ub) return mlc_fail(MLC_EINVAL, "mlc_get_field: NULL argument");
    if (!s.dev) {
      long pj = g.nj + 4;
      return visit(s, lb, ub, [&](long h, long d) {
        long jj = d % g.pj - JOFF, ii = d / g.pj - 1;
        host[h] = (*s.host2d)[(ii + 1) * pj + (jj + 1)];
      });
    }
    size_t n = (size_t)g.plane * s.nlev;
    if (s.is_double) {
      std::vector<double> buf(n);
      HIP_TRY(hipMemcpyAsync(buf.data(), s.dev, n * sizeof(double), hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
      return visit(s, lb, ub, [&](long h, long d) { host[h] = buf[d]; });
    }
    std::vector<T> buf(n);
    HIP_TRY(hipMemcpyAsync(buf.data(), s.dev, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return visit(s, lb, ub, [&](long h, long d) { host[h] = (double)buf[d]; });
  }

  // ------------------------------------------------------------- init
  // init_moloch (nompi/mod_moloch.F90:219-268) and xknu of allocate_moloch
  // (:206-210), evaluated on the host in FP64.
  int init() override {
    const int K = g.nk, nj = g.nj, ni = g.ni;
    std::vector<double> lv(4L * (K + 1), 0.0);
    for (int k = 1; k <= K + 1; k++) lv[k - 1] = gzita_host(zita[k - 1], prm.ztop, prm.a0);
    for (int k = 1; k <= K; k++) lv[(K + 1) + k - 1] = gzita_host(zitah[k - 1], prm.ztop, prm.a0);
    for (int k = 1; k <= K; k++)
      lv[2 * (K + 1) + k - 1] =
          kXdamp + (1.0 - kXdamp) * std::sin(0.5 * kMathpi * (1.0 - (double)(k - 1) / (double)(K - 1)));
    for (int k = 1; k <= K; k++) lv[3 * (K + 1) + k - 1] = ffilt[k - 1];
    HIP_TRY(hipMemcpyAsync((void *)D.gzitak, lv.data(), lv.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    long pjh = nj + 4;
    auto H = [&](const std::vector<double> &a, int j, int i) { return a[(long)(i + 1) * pjh + (j + 1)]; };
    std::vector<T> coru(g.plane, T(0)), corv(g.plane, T(0)), mx2(g.plane, T(0)), rmu(g.plane, T(0)),
        rmv(g.plane, T(0));
    for (int i = 1; i <= ni; i++)
      for (int j = 1; j <= nj; j++) {
        coru[g.at2(j, i)] = (T)(kEomeg2 * std::sin(H(h_ulat, j, i) * kDegrad));
        corv[g.at2(j, i)] = (T)(kEomeg2 * std::sin(H(h_vlat, j, i) * kDegrad));
        mx2[g.at2(j, i)] = (T)(H(h_msfx, j, i) * H(h_msfx, j, i));
      }
    for (int i = -1; i <= ni + 2; i++)
      for (int j = -1; j <= nj + 2; j++) {
        double a = H(h_msfu, j, i), b = H(h_msfv, j, i);
        rmu[g.at2(j, i)] = (T)(a != 0.0 ? 1.0 / a : 0.0);
        rmv[g.at2(j, i)] = (T)(b != 0.0 ? 1.0 / b : 0.0);
      }
    HIP_TRY(hipMemcpyAsync(D.coru, coru.data(), g.plane * sizeof(T), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(D.corv, corv.data(), g.plane * sizeof(T), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(D.mx2, mx2.data(), g.plane * sizeof(T), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(D.rmu, rmu.data(), g.plane * sizeof(T), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipMemcpyAsync(D.rmv, rmv.data(), g.plane * sizeof(T), hipMemcpyHostToDevice, stream));
    // w(:,:,1) = 0 (:260-262)
    HIP_TRY(hipMemsetAsync(D.w, 0, g.plane * sizeof(T), stream));
    HIP_TRY(hipStreamSynchronize(stream));
    initialised = true;
    return MLC_OK;
  }

  // --------------------------------------------------------- launchers
  dim3 col_grid() const { return dim3((g.nj + 63) / 64, (g.ni + 3) / 4, 1); }
  dim3 pt_grid() const { return dim3((g.nj + 63) / 64, (g.ni + 3) / 4, g.nk); }
  dim3 blk() const { return dim3(64, 4, 1); }
  dim3 col_grid(int by) const { return dim3((g.nj + 63) / 64, (g.ni + by - 1) / by, 1); }
  Par par() const {
    Par P;
    P.dx = prm.dx;
    P.dzita = prm.dzita;
    P.anu2 = prm.anu2;
    P.rotllr = prm.rotllr;
    P.halo_mpi = prm.halo_mpi;
    P.fulleq = prm.fulleq;
    P.divfilter = prm.divfilter;
    return P;
  }
  template <typename F>
  void halo(F *f, int nlev, int mode) {
    int n = (g.nj > g.ni ? g.nj : g.ni) + 1;
    launch(KK_HALO, 0.0, k_halo<F>, dim3((n + 255) / 256, nlev), dim3(256), f, g, mode);
  }

  // The segmented column solve (k_sound_bs, divdamp folded into k_sound_c)
  // unless params.unfused asks for the one-lane sweep, or kz exceeds what
  // SEG_NS segments of SEG_L levels hold.
  bool segmented_solve() const { return !prm.unfused && g.nk >= 2 && g.nk <= SEG_NS * SEG_L + 1; }

  // One acoustic substep of length dts (the body of the loop at :511-839);
  // ns is the substep's ordinal, which enters the w solve (:641-657).
  void substep(double dts, int ns) {
    Par P = par();
    halo(D.u, g.nk, 1);
    halo(D.v, g.nk, 1);
    launch(KK_SOUND_A, fbytes(3, 0), k_sound_a<T>, col_grid(), blk(), D, P, dts);
    int split = prm.divfilter ? 1 : 0;
    const bool seg = segmented_solve();
    if (seg) {
      const int nbj = (g.nj + SEG_CW - 1) / SEG_CW;
      launch(KK_SOUND_B, fbytes(11, 4), k_sound_bs<T, SEG_NS, SEG_L, SEG_CW>, dim3(nbj * g.ni),
             dim3(64, SEG_NS * SEG_CW / 64), D, P, dts, ns, split, nbj);
    } else {
      launch(KK_SOUND_B, fbytes(11, 4), k_sound_b<T>, col_grid(), blk(), D, P, dts, ns, split);
    }
    if (split) {
      // divergence_filter (:423-438) then the Exner update
      halo(D.zds, g.nk, 0);
      launch(KK_SMOOTH5, fbytes(2, 0), k_smooth5<T>, pt_grid(), blk(), (const T *)D.zds, D.zdf, g,
             (const double *)D.xknu, prm.anu2);
      launch(KK_SOUND_B2, fbytes(1, 2), k_sound_b2<T>, col_grid(), blk(), D, P);
    }
    // 32-bit offsets for the FP32 state only: the FP64 kernel measured
    // slower with them (79 against 63 ms per 2048 x 2048 x 60 step)
    const bool so = sizeof(T) == sizeof(float) && small_offsets();
    if (seg && so)
      launch(KK_SOUND_C, fbytes(5, 2), k_sound_c<T, true, unsigned, 16>, col_grid(16), dim3(64, 16), D, P, dts);
    else if (seg)
      launch(KK_SOUND_C, fbytes(5, 2), k_sound_c<T, true, long, 4>, col_grid(), blk(), D, P, dts);
    else if (so)
      launch(KK_SOUND_C, fbytes(5, 2), k_sound_c<T, false, unsigned, 16>, col_grid(16), dim3(64, 16), D, P, dts);
    else
      launch(KK_SOUND_C, fbytes(5, 2), k_sound_c<T, false, long, 4>, col_grid(), blk(), D, P, dts);
    std::swap(D.u, D.un);
    std::swap(D.v, D.vn);
  }

  // sound (:483-851): nsound substeps of dts.
  void sound(double dts) {
    if (!prm.fulleq) halo(D.tetav, g.nk, 0);
    for (int ns = 1; ns <= prm.nsound; ns++) substep(dts, ns);
  }

  // wafone (:929-1311) on one field with the three sweep kernels (the
  // single vertical pass, or params.unfused).
  template <typename Q>
  void wafone(Q *pp, double dta) {
    Par P = par();
    const double q = sizeof(Q) == sizeof(double) ? 1.0 : 0.0;  // advected field in FP64?
    Q *wz = (Q *)D.wz, *p0 = (Q *)D.p0;
    // vertical: pp, s, fmz, fmzf -> wz;  meridional: wz, v, fmz, pp -> p0;
    // zonal: p0, u, fmz, pp -> pp
    launch(KK_WAF_V, fbytes(3 + (1 - q), q), k_waf_v<T, Q>, col_grid(), blk(), D, P, (const Q *)pp, wz, dta,
           prm.vadvtwice ? 1 : 0);
    launch(KK_WAF_M, 0.0, k_waf_m<T, Q>, pt_grid(), blk(), D, P, (const Q *)pp,
           (const Q *)wz, p0, dta);
    launch(KK_WAF_Z, fbytes(1 - q, q), k_waf_z<T, Q>, pt_grid(), blk(), D, P, pp, (const Q *)p0, dta);
  }

  // every 3-D array (kz + 1 levels, FP64 at most) below 4 GiB: kernels may
  // address them by 32-bit byte offsets
  bool small_offsets() const { return (double)g.plane * (g.nk + 1) * sizeof(double) < 4294967296.0; }

  // wafone on a group of fields of one storage type: k_wafv then k_wafh; each
  // field's result buffer then takes the field's place.
  // One group: NF fields of type Q (in *fld[f], scratch wzs[f], results into
  // outs[f], then swapped into place) and ND FP64 riders (*dfld[d], scratch
  // dwzs[d], results into douts[d], swapped likewise).
  // Charged bytes: s, fmz and fmzf once per advection cycle (sfm_charged),
  // and one read and one write of each advected field except the first
  // nfree of the group (ux, vx, wx, which stand in for the u, v, w passes of
  // k_adv_in and k_adv_out).
  bool sfm_charged = false;
  template <typename Q, int NF, int ND>
  void waf_group(Q **fld[], Q **outs, Q **wzs, double **dfld[], double **douts, double **dwzs, double dta,
                 int nfree = 0) {
    Par P = par();
    FieldSet<Q> F;
    for (int f = 0; f < MAXF; f++) F.in[f] = nullptr, F.wz[f] = nullptr, F.out[f] = nullptr;
    for (int f = 0; f < NF; f++) {
      F.in[f] = *fld[f];
      F.wz[f] = wzs[f];
      F.out[f] = outs[f];
    }
    FieldD FD;
    for (int d = 0; d < MAXD; d++) FD.in[d] = nullptr, FD.wz[d] = nullptr, FD.out[d] = nullptr;
    for (int d = 0; d < ND; d++) {
      FD.in[d] = *dfld[d];
      FD.wz[d] = dwzs[d];
      FD.out[d] = douts[d];
    }
    const double q = sizeof(Q) == sizeof(double) ? 1.0 : 0.0;
    const int nc = NF - nfree;
    // vertical: s, fmz, fmzf + the fields -> wz (the riders on their own)
    const double vb = (sfm_charged ? 0.0 : fbytes(3, 0)) + (q ? fbytes(0, nc) : fbytes(nc, 0));
    sfm_charged = true;
    if (small_offsets())
      launch(KK_WAF_V, vb, k_wafv<T, Q, NF, unsigned>, col_grid(), blk(), D, P, F, dta);
    else
      launch(KK_WAF_V, vb, k_wafv<T, Q, NF, long>, col_grid(), blk(), D, P, F, dta);
    if constexpr (ND > 0) {
      FieldSet<double> Fd;
      for (int f = 0; f < MAXF; f++) Fd.in[f] = nullptr, Fd.wz[f] = nullptr, Fd.out[f] = nullptr;
      for (int d = 0; d < ND; d++) {
        Fd.in[d] = FD.in[d];
        Fd.wz[d] = FD.wz[d];
      }
      if (small_offsets())
        launch(KK_WAF_V, fbytes(0, ND), k_wafv<T, double, ND, unsigned>, col_grid(), blk(), D, P, Fd, dta);
      else
        launch(KK_WAF_V, fbytes(0, ND), k_wafv<T, double, ND, long>, col_grid(), blk(), D, P, Fd, dta);
    }
    // horizontal: u, v, fmz + the fields' wz and starting values -> results
    const int ntj = (g.nj + HJ - 1) / HJ, nti = (g.ni + HR - 1) / HR;
    const long total = ((long)ntj * nti * g.nk + 3) / 4;  // blocks of four strips
    const double hb = (q ? fbytes(0, nc) : fbytes(nc, 0)) + fbytes(0, ND);
    const dim3 hg((unsigned)round_up(total, 8));
    if (small_offsets()) {
      if (prm.rotllr)
        launch(KK_WAF_H, hb, k_wafh<T, Q, NF, ND, true, unsigned>, hg, dim3(256), D, P, F, FD, dta, ntj, nti);
      else
        launch(KK_WAF_H, hb, k_wafh<T, Q, NF, ND, false, unsigned>, hg, dim3(256), D, P, F, FD, dta, ntj, nti);
    } else {
      if (prm.rotllr)
        launch(KK_WAF_H, hb, k_wafh<T, Q, NF, ND, true, long>, hg, dim3(256), D, P, F, FD, dta, ntj, nti);
      else
        launch(KK_WAF_H, hb, k_wafh<T, Q, NF, ND, false, long>, hg, dim3(256), D, P, F, FD, dta, ntj, nti);
    }
    for (int f = 0; f < NF; f++) std::swap(*fld[f], outs[f]);
    for (int d = 0; d < ND; d++) std::swap(*dfld[d], douts[d]);
  }
  // The n fields in groups of at most waf_gmax<Q>(); the nd FP64 riders, if
  // any, go with the last group.
  // The first nfree fields are not charged (see waf_group).
  template <typename Q>
  void waf_groups(Q **fld[], int n, Q **outs, Q **wzs, double **dfld[], int nd, double **douts, double **dwzs,
                  double dta, int nfree = 0) {
    constexpr int GM = waf_gmax<Q>();
    for (int b = 0; b < n; b += GM) {
      const int m = n - b < GM ? n - b : GM;
      const bool rider = nd > 0 && b + m == n;
      const int nf = nfree - b < 0 ? 0 : (nfree - b > m ? m : nfree - b);
#define MLC_WAF_GROUP(NF)                                                   \
  if constexpr (sizeof(Q) != sizeof(double)) {                              \
    if (rider) {                                                            \
      waf_group<Q, NF, MAXD>(fld + b, outs, wzs, dfld, douts, dwzs, dta, nf); \
      break;                                                                \
    }                                                                       \
  }                                                                         \
  waf_group<Q, NF, 0>(fld + b, outs, wzs, nullptr, nullptr, nullptr, dta, nf);
      switch (m) {
        case 1: MLC_WAF_GROUP(1) break;
        case 2: MLC_WAF_GROUP(2) break;
        case 3: MLC_WAF_GROUP(3) break;
        case 4: MLC_WAF_GROUP(4) break;
        case 5: MLC_WAF_GROUP(5) break;
        default:
          if constexpr (GM == 8) {
            switch (m) {
              case 6: MLC_WAF_GROUP(6) break;
              case 7: MLC_WAF_GROUP(7) break;
              default: MLC_WAF_GROUP(8) break;
            }
          }
          break;
      }
#undef MLC_WAF_GROUP
    }
  }


  void advection(double dta) {
    halo(D.u, g.nk, 2);
    halo(D.v, g.nk, 3);
    launch(KK_ADV_IN, fbytes(4, 0), k_adv_in<T>, col_grid(), blk(), D);
    if (prm.vadvtwice && !unfused_advection) {
      T **qm[5] = {&D.qv, &D.qc, &D.qi, &D.qr, &D.qs};
      T **list[10];
      int n = 0;
      list[n++] = &D.ux;
      list[n++] = &D.vx;
      list[n++] = &D.wx;
      if constexpr (sizeof(T) == sizeof(double)) {
        list[n++] = reinterpret_cast<T **>(&D.tetav);
        list[n++] = reinterpret_cast<T **>(&D.pai);
      }
      list[n++] = &D.qv;
      for (int s = prm.iqfrst; s <= dims.nqx; s++) list[n++] = qm[s - 1];
      sfm_charged = false;
      if constexpr (sizeof(T) != sizeof(double)) {
        // theta and the Exner function (FP64) ride with the last group
        static_assert(MAXD == 2, "two FP64 riders");
        double **dl[MAXD] = {&D.tetav, &D.pai};
        double *dwz[MAXD] = {D.p0, D.wz};
        waf_groups<T>(list, n, gout, gwz, dl, MAXD, dgout, dwz, dta, 3);
      } else {
        waf_groups<T>(list, n, gout, gwz, nullptr, 0, nullptr, nullptr, dta, 3);
      }
      halo(D.ux, g.nk, 2);
      halo(D.vx, g.nk, 3);
      launch(KK_ADV_OUT, fbytes(3, 0), k_adv_out<T>, col_grid(), blk(), D);
      return;
    }
    wafone<double>(D.tetav, dta);
    wafone<double>(D.pai, dta);
    wafone<T>(D.ux, dta);
    wafone<T>(D.vx, dta);
    wafone<T>(D.wx, dta);
    wafone<T>(D.qv, dta);
    T *q[5] = {D.qv, D.qc, D.qi, D.qr, D.qs};
    for (int n = prm.iqfrst; n <= dims.nqx; n++) wafone<T>(q[n - 1], dta);
    halo(D.ux, g.nk, 2);
    halo(D.vx, g.nk, 3);
    launch(KK_ADV_OUT, fbytes(3, 0), k_adv_out<T>, col_grid(), blk(), D);
  }

  // ------------------------------------------- learned acoustic corrector
  struct LacState {
    bool exists = false;
    mlc_lac_config cfg{};
    mlc_mlp *net = nullptr;
    int nflp = 0, nin = 0, nout = 0, chunk = 0;
    LacRows rows{};
    long rtot = 0;
    T *su = nullptr, *sv = nullptr, *sw = nullptr;  // phase-start snapshot
    double *stet = nullptr;
    double *spai = nullptr;
    T *ru = nullptr, *rv = nullptr, *rw = nullptr;  // long substep without increment
    double *rpai = nullptr;
    T *cu = nullptr, *cv = nullptr, *cw = nullptr;  // long substep with increment
    double *cpai = nullptr;
    float *Y = nullptr;        // network outputs of every row [rtot x nout]
    FeatPlanes fp{};           // normalised inputs of every column (fused engine)
    __bf16 *Xc = nullptr;      // features of one inference batch
    float *Xraw = nullptr, *Traw = nullptr;  // training set (ring)
    long count = 0, next = 0;
    int *cols = nullptr, *Craw = nullptr;  // sampled columns; their ring copy
    std::vector<int> hcols;
    __bf16 *Xb = nullptr;
    float *Tn = nullptr;
    int *bidx = nullptr;
    long bcap = 0, icap = 0;
    float *stat = nullptr;  // imean, irstd [nin]; omean, ostd [nout]; fmean, frstd [nflp]
    bool normalised = false;
    double *norms = nullptr;  // check accumulators
    int mode = MLC_LAC_OFF;
    long long learned_steps = 0;
    int retrain_left = 0;  // collection steps left before an automatic retraining
    int consecutive = 0;   // retrainings since the last passed check
    mlc_lac_stats st{};
    uint64_t rng = 0;
    uint64_t next_u64() {
      uint64_t z = (rng += 0x9E3779B97F4A7C15ull);
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    }
  } L;

  int lac_create(const mlc_lac_config &c) override {
    if (!initialised) return mlc_fail(MLC_ESTATE, "mlc_lac_create before mlc_init");
    if (L.exists) return mlc_fail(MLC_ESTATE, "mlc_lac_create: the handle already has a corrector");
    if (prm.nsound < 2) return mlc_fail(MLC_EINVAL, "mlc_lac_create: needs nsound >= 2");
    if (c.nhid < 16 || c.samples_per_phase < 1 || c.capacity < 1 || c.chunk_rows < 256 || c.chunk_rows % 256 ||
        c.check_every < 0 || !(c.check_tol >= 0.0) || (c.engine != 0 && c.engine != 1))
      return mlc_fail(MLC_EINVAL, "mlc_lac_create: bad configuration");
    if (!(c.rel_tol >= 0.0) || c.retrain_steps < 0 || c.max_retrains < 0 ||
        (c.retrain_steps > 0 && (c.retrain_iters < 1 || c.retrain_batch < 1 || !(c.retrain_lr > 0.0))))
      return mlc_fail(MLC_EINVAL, "mlc_lac_create: bad drift-check or retraining settings");
    if (c.engine == 1 && c.nhid % 64)
      return mlc_fail(MLC_EINVAL, "mlc_lac_create: the fused engine needs nhid a multiple of 64");
    L.cfg = c;
    const int K = g.nk;
    L.nflp = (int)round_up(LAC_IN_FIELDS * K, 16);
    L.nin = 9 * L.nflp;
    L.nout = LAC_OUT_FIELDS * K;
    L.rows.ntj = (g.nj + 63) / 64;
    L.rtot = (long)L.rows.ntj * ((g.ni + 3) / 4) * 256;
    L.chunk = (int)(c.chunk_rows < L.rtot ? c.chunk_rows : L.rtot);
    int rc = mlp_new(L.nin, c.nhid, L.nout, c.seed, prm.device, stream, &L.net);
    if (rc) return rc;
#define AL(ptr, bytes) \
  if ((rc = alloc_bytes((void **)&ptr, (size_t)(bytes))) != MLC_OK) return rc;
    const size_t fT = (size_t)g.plane * K * sizeof(T), fT1 = (size_t)g.plane * (K + 1) * sizeof(T);
    const size_t fD = (size_t)g.plane * K * sizeof(double);
    AL(L.su, fT) AL(L.sv, fT) AL(L.sw, fT1) AL(L.stet, fD) AL(L.spai, fD);
    AL(L.ru, fT) AL(L.rv, fT) AL(L.rw, fT1) AL(L.rpai, fD);
    AL(L.cu, fT) AL(L.cv, fT) AL(L.cw, fT1) AL(L.cpai, fD);
    AL(L.Y, (size_t)L.rtot * L.nout * sizeof(float));
    if (c.engine == 0) AL(L.Xc, (size_t)L.chunk * L.nin * sizeof(__bf16));
    if (c.engine == 1) {
      L.fp.fpj = L.rows.ntj * 64 + 2;
      L.fp.fpi = ((g.ni + 3) / 4) * 4 + 2;
      AL(L.fp.F, (size_t)(L.nflp / 16) * L.fp.plane() * sizeof(__bf16));
    }
    AL(L.Xraw, (size_t)c.capacity * L.nin * sizeof(float));
    AL(L.Traw, (size_t)c.capacity * L.nout * sizeof(float));
    AL(L.cols, (size_t)c.samples_per_phase * sizeof(int));
    AL(L.Craw, (size_t)c.capacity * sizeof(int));
    AL(L.stat, (size_t)(2 * L.nin + 2 * L.nout + 2 * L.nflp) * sizeof(float));
    AL(L.norms, 16 * sizeof(double));
#undef AL
    L.hcols.assign(c.samples_per_phase, 0);
    L.rng = c.seed ^ 0xA5A5A5A55A5A5A5Aull;
    L.exists = true;
    HIP_TRY(hipStreamSynchronize(stream));
    return MLC_OK;
  }

  template <typename F>
  void copy_field(F *dst, const F *src, int nlev) {
    const size_t n = (size_t)g.plane * nlev * sizeof(F);
    timed(KK_LAC_COPY, 2.0 * g.nj * g.ni * nlev * sizeof(F),
          [&]() { (void)hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, stream); });
  }
  // theta is only needed to restore the phase start (checks, collection).
  void lac_snapshot(bool theta) {
    copy_field(L.su, D.u, g.nk);
    copy_field(L.sv, D.v, g.nk);
    copy_field(L.sw, D.w, g.nk + 1);
    if (theta) copy_field(L.stet, D.tetav, g.nk);
    copy_field(L.spai, D.pai, g.nk);
  }
  void lac_restore() {
    copy_field(D.u, L.su, g.nk);
    copy_field(D.v, L.sv, g.nk);
    copy_field(D.w, L.sw, g.nk + 1);
    copy_field(D.tetav, L.stet, g.nk);
    copy_field(D.pai, L.spai, g.nk);
  }
  void lac_save(T *u, T *v, T *w, double *pai) {
    copy_field(u, D.u, g.nk);
    copy_field(v, D.v, g.nk);
    copy_field(w, D.w, g.nk + 1);
    copy_field(pai, D.pai, g.nk);
  }

  LacSnap<T> lac_snap(bool inc = false) const {
    return LacSnap<T>{L.su, L.sv, L.sw, L.spai, (const float *)L.su, (const float *)L.sv, (const float *)L.sw,
                      (const float *)L.spai, inc ? 1 : 0};
  }
  // Increment mode: the long substep's kernels write the increments of u, v,
  // w, pi into the snapshot buffers (as floats), so a learned phase without a
  // check needs no phase-start copies.  Only on the segmented path (the
  // one-lane sweep updates the winds in place) and without the divergence
  // filter (which updates pi in a kernel of its own).
  bool lac_increments_ok() const { return segmented_solve() && !prm.divfilter && L.cfg.engine == 1; }
  void lac_increments(bool on) {
    D.iu = on ? (float *)L.su : nullptr;
    D.iv = on ? (float *)L.sv : nullptr;
    D.iw = on ? (float *)L.sw : nullptr;
    D.ipai = on ? (float *)L.spai : nullptr;
  }

  // Network outputs for every row from the current state and the snapshot.
  int lac_infer(bool inc) {
    const float *st = L.stat;
    const double nh = L.cfg.nhid;
    int rc = mlp_reserve_rows(L.net, L.chunk);
    if (rc) return rc;
    if (L.cfg.engine == 1 && (rc = mlp_pack_dev(L.net))) return rc;
    if (L.cfg.engine == 1) {
      const float *fs = L.stat + 2 * L.nin + 2 * L.nout;
      // inputs: the float increments of u, v, w, pi kept by the substep
      // kernels (inc) or u, v, w, pi and their phase-start copies, then
      // theta and qv; the feature planes written once
      const double in_b = inc ? (double)g.nj * g.ni * g.nk * (4.0 * sizeof(float) + 2.0 * sizeof(T))
                              : fbytes(8, 2);
      launch(KK_LAC_GATHER, in_b + (double)(L.nflp / 16) * L.fp.plane() * sizeof(__bf16),
             k_lac_features<T>, dim3((L.fp.fpj + 63) / 64, (L.fp.fpi + 3) / 4), dim3(256), D, lac_snap(inc), L.fp,
             L.nflp / 16, fs, fs + L.nflp);
    }
    for (long r0 = 0; r0 < L.rtot; r0 += L.chunk) {
      const int n = (int)(L.rtot - r0 < L.chunk ? L.rtot - r0 : L.chunk);
      // algorithmic flops: the 9 x 6 kz real inputs, not the zero padding to nin
      const double nreal = 9.0 * LAC_IN_FIELDS * g.nk;
      const double flops = 2.0 * n * (nreal * nh + nh * nh + nh * L.nout);
      if (L.cfg.engine == 1) {
        timed(KK_LAC_MLP, flops, [&]() { fused_forward(r0, n); });
        continue;
      }
      timed(KK_LAC_GATHER, (double)n * L.nin * sizeof(__bf16) + fbytes(6, 0) * n / L.rtot, [&]() {
        hipLaunchKernelGGL(k_lac_gather<T>, dim3(n / 64), dim3(256), 0, stream, D, lac_snap(inc), L.rows,
                           (const int *)nullptr, r0, n, L.nin, (float *)nullptr, L.Xc, st, st + L.nin);
      });
      timed(KK_LAC_MLP, flops, [&]() { rc = mlp_forward_dev(L.net, L.Xc, L.nin, n, L.Y + r0 * L.nout, L.nout); });
      if (rc) return rc;
    }
    return MLC_OK;
  }

  // The three layers of rows r0 .. r0+n-1 (n a multiple of 256) on the fused
  // kernels: layer 1 gathers its inputs from the fields.
  void fused_forward(long r0, int n) {
    mlc_mlp *p = L.net;
    const int nrt = n / FM;
    auto grid = [&](int nN) { return dim3((unsigned)round_up((long)nrt * nN, 8)); };
    FusedArgs a{};
    a.row0 = r0;
    a.nrt = nrt;
    a.fp = L.fp;
    // layer 1
    a.Bp = reinterpret_cast<const lac_bf16x8 *>(p->Wp[0]);
    a.bnt = p->wp_nt[0];
    a.nks = p->wp_ks[0];
    a.N = p->nhid;
    a.nN = (p->nhid + FN - 1) / FN;
    a.bias = p->bvec[0];
    a.H = p->H1;
    a.ldh = p->nhid;
    hipLaunchKernelGGL(k_mlp_fused, grid(a.nN), dim3(FT), 0, stream, L.rows, a);
    if (p->nhid == TH && p->nout <= 256 && p->wp_nt[1] == TH / 32 && p->wp_ks[1] == TKS && p->wp_nt[2] == 8 &&
        p->wp_ks[2] == TKS && n % TM == 0) {
      // layers 2 and 3 in one kernel
      TailArgs t{};
      t.Bp2 = reinterpret_cast<const lac_bf16x8 *>(p->Wp[1]);
      t.Bp3 = reinterpret_cast<const lac_bf16x8 *>(p->Wp[2]);
      t.b2 = p->bvec[1];
      t.b3 = p->bvec[2];
      t.A = p->H1;
      t.Y = L.Y + r0 * L.nout;
      t.ldy = L.nout;
      t.N3 = p->nout;
      t.nrt = n / TM;
      hipLaunchKernelGGL(k_mlp_tail, dim3((unsigned)round_up((long)t.nrt, 8)), dim3(FT), 0, stream, t);
      return;
    }
    // layer 2
    a.Bp = reinterpret_cast<const lac_bf16x8 *>(p->Wp[1]);
    a.bnt = p->wp_nt[1];
    a.nks = p->wp_ks[1];
    a.bias = p->bvec[1];
    a.A = p->H1;
    a.lda = p->nhid;
    a.H = p->H2;
    hipLaunchKernelGGL(k_mlp_dense, grid(a.nN), dim3(FT), 0, stream, a);
    // layer 3
    a.Bp = reinterpret_cast<const lac_bf16x8 *>(p->Wp[2]);
    a.bnt = p->wp_nt[2];
    a.nks = p->wp_ks[2];
    a.N = p->nout;
    a.nN = (p->nout + FN - 1) / FN;
    a.bias = p->bvec[2];
    a.A = p->H2;
    a.H = nullptr;
    a.Y = L.Y + r0 * L.nout;
    a.ldy = L.nout;
    hipLaunchKernelGGL(k_mlp_dense, grid(a.nN), dim3(FT), 0, stream, a);
  }

  void lac_apply() {
    const float *st = L.stat + 2 * L.nin;
    launch(KK_LAC_APPLY, (double)L.rtot * L.nout * sizeof(float) + fbytes(6, 2), k_lac_apply<T>,
           dim3((unsigned)(L.rtot / 64)), dim3(256), D, L.rows, (const float *)L.Y, st, st + L.nout);
  }

  // Squared norms of (a - b) and b for the four corrected fields into
  // L.norms[4*f + 2*which ...]: which 0 = with the increment, 1 = without.
  void lac_norms(int which, const T *u, const T *v, const T *w, const double *pai) {
    double *o = L.norms + 2 * which;
    launch(KK_DIAG, 0.0, k_diffnorm<T>, col_grid(), dim3(256), u, (const T *)D.u, g, 1, g.nk, o);
    launch(KK_DIAG, 0.0, k_diffnorm<T>, col_grid(), dim3(256), v, (const T *)D.v, g, 1, g.nk, o + 4);
    launch(KK_DIAG, 0.0, k_diffnorm<T>, col_grid(), dim3(256), w, (const T *)D.w, g, 2, g.nk, o + 8);
    launch(KK_DIAG, 0.0, k_diffnorm<double>, col_grid(), dim3(256), pai, (const double *)D.pai, g, 1, g.nk, o + 12);
  }

  // One acoustic phase with the long substep.  correct: add the network's
  // increment; check: also run the ordinary substeps from the same start,
  // compare, and continue from the ordinary result.
  int sound_lac(double dts, bool correct, bool check, double *worst, double *worst_rel) {
    if (!prm.fulleq) halo(D.tetav, g.nk, 0);
    const int nb = prm.nsound - 1;
    int rc;
    const bool inc = correct && !check && lac_increments_ok();
    if (inc)
      lac_increments(true);
    else
      lac_snapshot(check);
    substep(dts * nb, 1);
    lac_increments(false);
    if (correct && (rc = lac_infer(inc))) return rc;
    if (check) lac_save(L.ru, L.rv, L.rw, L.rpai);
    if (correct) lac_apply();
    if (check) {
      lac_save(L.cu, L.cv, L.cw, L.cpai);
      lac_restore();
      for (int ns = 1; ns <= nb; ns++) substep(dts, ns);
      HIP_TRY(hipMemsetAsync(L.norms, 0, 16 * sizeof(double), stream));
      lac_norms(0, L.cu, L.cv, L.cw, L.cpai);
      lac_norms(1, L.ru, L.rv, L.rw, L.rpai);
      double nr[16];
      HIP_TRY(hipMemcpyAsync(nr, L.norms, sizeof(nr), hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
      for (int f = 0; f < 4; f++) {
        const double ec = nr[4 * f], eu = nr[4 * f + 2], ref = nr[4 * f + 1];
        const double ratio = eu > 0.0 ? std::sqrt(ec / eu) : (ec > 0.0 ? INFINITY : 0.0);
        L.st.ratio[f] = ratio;
        L.st.rel[f] = ref > 0.0 ? std::sqrt(ec / ref) : std::sqrt(ec);
        // a NaN ratio (a network that puts out NaN) stays the worst
        if (!std::isnan(*worst) && !(ratio <= *worst)) *worst = ratio;
        if (!std::isnan(*worst_rel) && !(L.st.rel[f] <= *worst_rel)) *worst_rel = L.st.rel[f];
      }
      L.st.checks++;
    }
    substep(dts, prm.nsound);
    if (check)
      L.st.plain_phases++;
    else if (correct)
      L.st.learned_phases++;
    return MLC_OK;
  }

  // One ordinary acoustic phase that also records training pairs at
  // sampled columns; the state evolves exactly as in sound().
  int sound_collect(double dts) {
    if (!prm.fulleq) halo(D.tetav, g.nk, 0);
    const int nb = prm.nsound - 1, ns = L.cfg.samples_per_phase;
    const long ncol = (long)g.nj * g.ni;
    for (int r = 0; r < ns; r++) L.hcols[r] = (int)(L.next_u64() % (uint64_t)ncol);
    HIP_TRY(hipMemcpyAsync(L.cols, L.hcols.data(), ns * sizeof(int), hipMemcpyHostToDevice, stream));
    if (L.next + ns > L.cfg.capacity) L.next = 0;
    const long base = L.next;
    const long rows_fit = L.cfg.capacity - base;
    const int n = (int)(ns < rows_fit ? ns : rows_fit);
    lac_snapshot(true);
    substep(dts * nb, 1);
    launch(KK_LAC_GATHER, 0.0, k_lac_gather<T>, dim3((n + 63) / 64), dim3(256), D, lac_snap(), L.rows,
           (const int *)L.cols, 0L, n, L.nin, L.Xraw + base * L.nin, (__bf16 *)nullptr, (const float *)nullptr,
           (const float *)nullptr);
    lac_save(L.ru, L.rv, L.rw, L.rpai);
    lac_restore();
    for (int s = 1; s <= nb; s++) substep(dts, s);
    launch(KK_LAC_GATHER, 0.0, k_lac_target<T>, dim3((unsigned)(((long)n * L.nout + 255) / 256)), dim3(256), D,
           (const int *)L.cols, n, (const T *)L.ru, (const T *)L.rv, (const T *)L.rw, (const double *)L.rpai,
           L.Traw + base * L.nout);
    substep(dts, prm.nsound);
    HIP_TRY(hipMemcpyAsync(L.Craw + base, L.cols, n * sizeof(int), hipMemcpyDeviceToDevice, stream));
    L.next = base + n;
    if (L.next > L.count) L.count = L.next;
    L.st.samples = L.count;
    L.st.plain_phases++;
    HIP_TRY(hipStreamSynchronize(stream));  // hcols is reused by the next phase
    return MLC_OK;
  }

  // Input normalisation per (field, level), pooled over the stencil
  // positions; target normalisation per output.
  int lac_normalise() {
    const int nin = L.nin, nout = L.nout;
    double *dm = nullptr;
    HIP_TRY(hipMalloc((void **)&dm, (size_t)2 * (nin + nout) * sizeof(double)));
    hipLaunchKernelGGL(k_colstats, dim3(nin), dim3(256), 0, stream, (const float *)L.Xraw, L.count, nin, dm,
                       dm + nin + nout);
    hipLaunchKernelGGL(k_colstats, dim3(nout), dim3(256), 0, stream, (const float *)L.Traw, L.count, nout, dm + nin,
                       dm + 2 * nin + nout);
    std::vector<double> h(2 * (nin + nout));
    hipError_t e = hipMemcpyAsync(h.data(), dm, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(dm);
    if (e != hipSuccess) return mlc_fail(MLC_EHIP, "lac normalisation: %s", hipGetErrorString(e));
    const double *mean = h.data(), *sd = h.data() + nin + nout;
    std::vector<double> fm(L.nflp, 0.0), fq(L.nflp, 0.0);
    for (int o = 0; o < nin; o++) {
      int fl, p;
      lac_feature(o, fl, p);
      fm[fl] += mean[o] / 9.0;
      fq[fl] += (sd[o] * sd[o] + mean[o] * mean[o]) / 9.0;
    }
    std::vector<float> st(2 * nin + 2 * nout + 2 * L.nflp);
    float *fmean = st.data() + 2 * nin + 2 * nout, *frstd = fmean + L.nflp;
    for (int fl = 0; fl < L.nflp; fl++) {
      const double var = fq[fl] - fm[fl] * fm[fl];
      const double s = var > 0.0 ? std::sqrt(var) : 0.0;
      fmean[fl] = (float)fm[fl];
      frstd[fl] = s > 1e-30 * (1.0 + std::fabs(fm[fl])) ? (float)(1.0 / s) : 0.f;
    }
    for (int o = 0; o < nin; o++) {
      int fl, p;
      lac_feature(o, fl, p);
      st[o] = fmean[fl];
      st[nin + o] = frstd[fl];
    }
    // targets: mean per output, spread per field (root mean square of the
    // levels' standard deviations), so that the loss weights levels by their
    // size as the fields' L2 errors do
    for (int f = 0; f < LAC_OUT_FIELDS; f++) {
      double q = 0.0;
      for (int k = 0; k < g.nk; k++) q += sd[nin + f * g.nk + k] * sd[nin + f * g.nk + k];
      const double s = std::sqrt(q / g.nk);
      for (int k = 0; k < g.nk; k++) {
        const int o = f * g.nk + k;
        st[2 * nin + o] = (float)mean[nin + o];
        st[2 * nin + nout + o] = s > 0.0 ? (float)s : 1.f;
      }
    }
    HIP_TRY(hipMemcpyAsync(L.stat, st.data(), st.size() * sizeof(float), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    L.normalised = true;
    return MLC_OK;
  }

  int lac_train(int iters, int batch, double lr, double *loss) override {
    if (!L.exists) return mlc_fail(MLC_ESTATE, "mlc_lac_train: no corrector (mlc_lac_create)");
    if (iters < 0 || batch < 1 || !(lr > 0.0)) return mlc_fail(MLC_EINVAL, "mlc_lac_train: bad arguments");
    if (L.count < 1) return mlc_fail(MLC_ESTATE, "mlc_lac_train: the training set is empty (mlc_lac_collect)");
    int rc;
    if (!L.normalised && (rc = lac_normalise())) return rc;
    if (batch > L.bcap) {
      if ((rc = alloc_bytes((void **)&L.Xb, (size_t)batch * L.nin * sizeof(__bf16)))) return rc;
      if ((rc = alloc_bytes((void **)&L.Tn, (size_t)batch * L.nout * sizeof(float)))) return rc;
      L.bcap = batch;
    }
    const long nidx = (long)iters * batch;
    if (nidx > L.icap) {
      if ((rc = alloc_bytes((void **)&L.bidx, (size_t)nidx * sizeof(int)))) return rc;
      L.icap = nidx;
    }
    std::vector<int> idx(nidx);
    for (long r = 0; r < nidx; r++) idx[r] = (int)(L.next_u64() % (uint64_t)L.count);
    HIP_TRY(hipMemcpyAsync(L.bidx, idx.data(), nidx * sizeof(int), hipMemcpyHostToDevice, stream));
    const float *st = L.stat;
    double l = NAN;
    const unsigned gx = (unsigned)((L.nin + 255) / 256);
    for (int it = 0; it < iters; it++) {
      hipLaunchKernelGGL(k_lac_batch, dim3(gx, batch), dim3(256), 0, stream, (const float *)L.Xraw,
                         (const float *)L.Traw, (const int *)(L.bidx + (long)it * batch), L.nin, L.nout, st,
                         st + L.nin, st + 2 * L.nin, st + 2 * L.nin + L.nout, L.Xb, L.Tn);
      if ((rc = mlp_backward_dev(L.net, L.Xb, L.nin, L.Tn, batch, it == iters - 1 ? &l : nullptr))) return rc;
      if ((rc = mlp_adam_dev(L.net, lr))) return rc;
    }
    HIP_TRY(hipStreamSynchronize(stream));
    HIP_TRY(hipGetLastError());
    if (iters > 0) {
      L.st.last_loss = l;
      L.st.fallback = 0;
      L.retrain_left = 0;
    }
    if (loss) *loss = l;
    return MLC_OK;
  }

  int lac_collect(int nsteps) override {
    if (!L.exists) return mlc_fail(MLC_ESTATE, "mlc_lac_collect: no corrector (mlc_lac_create)");
    if (nsteps < 0) return mlc_fail(MLC_EINVAL, "mlc_lac_collect: nsteps < 0");
    for (int s = 0; s < nsteps; s++) {
      int rc = run(0, nullptr, true);
      if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(stream));
    HIP_TRY(hipGetLastError());
    return MLC_OK;
  }

  int lac_mode(int mode) override {
    if (!L.exists) return mlc_fail(MLC_ESTATE, "mlc_lac_mode: no corrector (mlc_lac_create)");
    if (mode != MLC_LAC_OFF && mode != MLC_LAC_ON && mode != MLC_LAC_REDUCED)
      return mlc_fail(MLC_EINVAL, "mlc_lac_mode: unknown mode %d", mode);
    if (mode == MLC_LAC_ON && !L.normalised) return mlc_fail(MLC_ESTATE, "mlc_lac_mode: the corrector is not trained");
    L.mode = mode;
    L.learned_steps = 0;
    L.retrain_left = 0;
    L.consecutive = 0;
    return MLC_OK;
  }

  int lac_stats(mlc_lac_stats *s) override {
    if (!L.exists) return mlc_fail(MLC_ESTATE, "mlc_lac_stats: no corrector (mlc_lac_create)");
    *s = L.st;
    return MLC_OK;
  }

  mlc_mlp *lac_network() override { return L.exists ? L.net : nullptr; }

  int pack_output(double *dst, hipEvent_t done) override {
    if (!initialised) return mlc_fail(MLC_ESTATE, "mlc_output_write before mlc_init");
    const long n2 = (long)g.nj * g.ni;
    hipLaunchKernelGGL(k_pack_interior<T>, dim3((g.nj + 63) / 64, (g.ni + 3) / 4, 1), dim3(256), 0, stream,
                       (const T *)D.ps, g, dst);
    hipLaunchKernelGGL(k_pack_interior<double>, dim3((g.nj + 63) / 64, (g.ni + 3) / 4, g.nk), dim3(256), 0, stream,
                       (const double *)D.t, g, dst + n2);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(done, stream));
    return MLC_OK;
  }

  int wait_event(hipEvent_t e) override {
    HIP_TRY(hipStreamWaitEvent(stream, e, 0));
    return MLC_OK;
  }

  // 3-D fields of a case on the device; the 2-D fields and the levels have
  // been set from H.
  int build_case(const mlc_case &cs, const CaseHost &H) override {
    const int K = g.nk;
    const long n2h = (long)(g.nj + 4) * (g.ni + 4);
    double *d2 = nullptr;  // ht, ps in the device 2-D layout, zita, zitah
    const size_t n2 = (size_t)g.plane;
    HIP_TRY(hipMalloc((void **)&d2, (2 * n2 + 2 * (K + 1)) * sizeof(double)));
    std::vector<double> h2(2 * n2 + 2 * (K + 1), 0.0);
    for (int i = -1; i <= g.ni + 2; i++)
      for (int j = -1; j <= g.nj + 2; j++) {
        const long hh = (long)(i + 1) * (g.nj + 4) + (j + 1);
        h2[g.at2(j, i)] = H.ht[hh];
        h2[n2 + g.at2(j, i)] = H.ps[hh];
      }
    (void)n2h;
    for (int k = 0; k <= K; k++) h2[2 * n2 + k] = H.zita[k];
    for (int k = 0; k < K; k++) h2[2 * n2 + K + 1 + k] = H.zitah[k];
    hipError_t e = hipMemcpyAsync(d2, h2.data(), h2.size() * sizeof(double), hipMemcpyHostToDevice, stream);
    CaseDev c{};
    c.kind = cs.kind;
    c.noise_levels = cs.noise_levels;
    c.ztop = cs.ztop, c.zh = cs.zh, c.a0 = cs.a0, c.tsfc = cs.tsfc, c.lapse = cs.lapse, c.ttrop = cs.ttrop;
    c.psfc = cs.psfc, c.rh = cs.rh, c.u0 = cs.u0, c.v0 = cs.v0, c.jet_amp = cs.jet_amp, c.jet_z = cs.jet_z;
    c.jet_width = cs.jet_width, c.noise_amp = cs.noise_amp, c.bubble_amp = cs.bubble_amp;
    c.bubble_rcells = cs.bubble_rcells, c.bubble_z = cs.bubble_z, c.bubble_rz = cs.bubble_rz;
    c.dzita = H.zita[K - 1];
    c.rng = H.rng;
    c.ht = d2;
    c.ps = d2 + n2;
    c.zita = d2 + 2 * n2;
    c.zitah = d2 + 2 * n2 + K + 1;
    // fields the case leaves at zero
    T *zf[] = {D.w, D.qc, D.qi, D.qr, D.qs, D.tvirt, D.p, D.rho, D.qsat, D.ux, D.vx};
    const int zl[] = {K + 1, K, K, K, K, K, K, K, K, K, K};
    for (int f = 0; f < 11 && e == hipSuccess; f++) e = hipMemsetAsync(zf[f], 0, g.plane * zl[f] * sizeof(T), stream);
    if (e == hipSuccess) e = hipMemsetAsync(D.tetav, 0, g.plane * K * sizeof(double), stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_case_build<T>, col_grid(), blk(), 0, stream, D, c);
      const dim3 pg((unsigned)(((g.nj + 4) * (g.ni + 4) + 255) / 256), 1);
      T *pf[] = {D.u, D.v, D.qv, D.zeta, D.fmz};
      for (T *f : pf) hipLaunchKernelGGL(k_periodic<T>, dim3(pg.x, K), dim3(256), 0, stream, f, g);
      hipLaunchKernelGGL(k_periodic<T>, dim3(pg.x, K + 1), dim3(256), 0, stream, D.fmzf, g);
      hipLaunchKernelGGL(k_periodic<double>, dim3(pg.x, K), dim3(256), 0, stream, D.pai, g);
      hipLaunchKernelGGL(k_periodic<double>, dim3(pg.x, K), dim3(256), 0, stream, D.t, g);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    (void)hipFree(d2);
    if (e != hipSuccess) return mlc_fail(MLC_EHIP, "mlc_case_load_device: %s", hipGetErrorString(e));
    return MLC_OK;
  }

  int lac_dataset(long first, int n, float *X, float *Tt, int *cols) override {
    if (!L.exists) return mlc_fail(MLC_ESTATE, "mlc_lac_dataset: no corrector (mlc_lac_create)");
    if (first < 0 || n < 0 || first + n > L.count) return mlc_fail(MLC_EINVAL, "mlc_lac_dataset: rows out of range");
    if (X) HIP_TRY(hipMemcpyAsync(X, L.Xraw + first * L.nin, (size_t)n * L.nin * 4, hipMemcpyDeviceToHost, stream));
    if (Tt)
      HIP_TRY(hipMemcpyAsync(Tt, L.Traw + first * L.nout, (size_t)n * L.nout * 4, hipMemcpyDeviceToHost, stream));
    if (cols) HIP_TRY(hipMemcpyAsync(cols, L.Craw + first, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return MLC_OK;
  }

  int step(int istep, int lprint, mlc_diag *diag) override {
    (void)istep;
    return run(lprint, diag, false);
  }

  // Debug mode: before every step, NaN into the halo ring of every state and
  // scratch field (not the geometry), so a read of a halo cell the step has
  // not filled first shows up as a NaN in the interior.
  bool poison = false;
  int poison_halos(int enable) override {
    poison = enable != 0;
    return MLC_OK;
  }
  void poison_now() {
    const int K = g.nk;
    const dim3 pg((unsigned)((4 * (g.ni + 4) + 4 * g.nj + 255) / 256), 1);
    T *tf[] = {D.u, D.v, D.qv, D.qc, D.qi, D.qr, D.qs, D.ux, D.vx, D.wx, D.zdh, D.tvirt, D.p, D.rho, D.qsat};
    for (T *f : tf) hipLaunchKernelGGL(k_poison<T>, dim3(pg.x, K), dim3(256), 0, stream, f, g);
    T *tf1[] = {D.w, D.s};
    for (T *f : tf1) hipLaunchKernelGGL(k_poison<T>, dim3(pg.x, K + 1), dim3(256), 0, stream, f, g);
    // deltaw at levels 2..kz only: levels 1 and kz+1 (halos included) are the
    // zeros reset_tendencies leaves there for good (:1313-1325), which the
    // momentum equations read at the bottom level (:748-750, :773-775)
    if (K > 1) hipLaunchKernelGGL(k_poison<T>, dim3(pg.x, K - 1), dim3(256), 0, stream, D.deltaw + g.plane, g);
    double *df[] = {D.pai, D.tetav, D.t};
    for (double *f : df) hipLaunchKernelGGL(k_poison<double>, dim3(pg.x, K), dim3(256), 0, stream, f, g);
  }

  // collect: record training pairs in every acoustic phase (ordinary loop).
  int run(int lprint, mlc_diag *diag, bool collect) {
    if (!initialised) return mlc_fail(MLC_ESTATE, "mlc_step before mlc_init");
    if (lprint && !diag) return mlc_fail(MLC_EINVAL, "mlc_step: lprint set but diag is NULL");
    const double dtstepa = prm.dt / (double)prm.nadv;
    const double dtsound = dtstepa / (double)prm.nsound;
    // acoustic loop of this step
    const bool learned = L.exists && L.mode == MLC_LAC_ON && !L.st.fallback && !collect;
    // after a failed check: ordinary steps that collect fresh training pairs
    const bool recollect = L.exists && L.mode == MLC_LAC_ON && L.st.fallback && L.retrain_left > 0 && !collect;
    const bool reduced = L.exists && L.mode == MLC_LAC_REDUCED && !collect;
    if (learned && !L.normalised) return mlc_fail(MLC_ESTATE, "mlc_step: the corrector is not trained");
    const bool check = learned && L.cfg.check_every > 0 && L.learned_steps % L.cfg.check_every == 0;
    Par P = par();
    if (poison) poison_now();
    launch(KK_PROLOGUE, fbytes(8, 3), k_prologue<T>, col_grid(), blk(), D, P, prm.filterpai ? 1 : 0,
           (prm.fulleq && prm.filtertheta) ? 1 : 0);
    double worst = 0.0, worst_rel = 0.0;
    for (int n = 1; n <= prm.nadv; n++) {
      int rc = MLC_OK;
      if (collect || recollect) {
        rc = sound_collect(dtsound);
        if (recollect) L.st.collect_phases++;
      } else if (learned || reduced)
        rc = sound_lac(dtsound, learned, check, &worst, &worst_rel);
      else {
        sound(dtsound);
        if (L.exists) L.st.plain_phases++;
      }
      if (rc) return rc;
      advection(dtstepa);
    }
    if (learned) {
      L.learned_steps++;
      if (check) {
        L.st.worst_ratio = worst;
        L.st.worst_rel = worst_rel;
        // NaN included: the comparisons are false
        const bool ok = worst <= L.cfg.check_tol && (L.cfg.rel_tol <= 0.0 || worst_rel <= L.cfg.rel_tol);
        if (ok) {
          L.consecutive = 0;
        } else {
          L.st.fallback = 1;
          L.st.failed_checks++;
          if (L.cfg.retrain_steps > 0 && L.consecutive < L.cfg.max_retrains) {
            L.retrain_left = L.cfg.retrain_steps;
            L.count = 0;  // the training set restarts with the current flow
            L.next = 0;
          }
        }
      }
    }
    const bool retrain_now = recollect && --L.retrain_left == 0;
    if (prm.filterpai) {  // :333-341
      launch(KK_FILTER_BRACKET, fbytes(0, 3), k_sub<double, double>, pt_grid(), blk(), D.pai, (const double *)D.pf, g);
      halo(D.pai, g.nk, 0);
      launch(KK_SMOOTH5, fbytes(0, 2), k_smooth5<double>, pt_grid(), blk(), (const double *)D.pai, D.smd, g,
             (const double *)nullptr, kNupaitq);
      launch(KK_FILTER_BRACKET, fbytes(0, 3), k_addto<double, double>, pt_grid(), blk(), D.pai,
             (const double *)D.smd, (const double *)D.pf, g);
    }
    if (prm.fulleq && prm.filtertheta) {  // :343-353
      launch(KK_FILTER_BRACKET, fbytes(0, 3), k_sub<double, double>, pt_grid(), blk(), D.tetav,
             (const double *)D.tf, g);
      halo(D.tetav, g.nk, 0);
      launch(KK_SMOOTH5, fbytes(0, 2), k_smooth5<double>, pt_grid(), blk(), (const double *)D.tetav, D.smt, g,
             (const double *)nullptr, kNupaitq);
      launch(KK_FILTER_BRACKET, fbytes(0, 3), k_addto<double, double>, pt_grid(), blk(), D.tetav,
             (const double *)D.smt, (const double *)D.tf, g);
    }
    launch(KK_EPILOGUE, fbytes(9, 3), k_epilogue<T>, pt_grid(), blk(), D);
    HIP_TRY(hipGetLastError());
    if (retrain_now) {
      // retrain on the fresh pairs, normalisation recomputed, then resume
      // with a check at the next step
      L.consecutive++;
      L.st.retrains++;
      L.normalised = false;
      int rc = lac_train(L.cfg.retrain_iters, L.cfg.retrain_batch, L.cfg.retrain_lr, nullptr);
      if (rc) return rc;
      L.learned_steps = 0;
    }
    if (lprint) {
      launch(KK_DIAG, fbytes(2, 0), k_diag_partial<T>, col_grid(), blk(), D);
      launch(KK_DIAG, 0.0, k_diag_final, dim3(1), dim3(256), D.diag, nblk_diag, d_diag_out);
      double out[7];
      HIP_TRY(hipMemcpyAsync(out, d_diag_out, sizeof(out), hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
      diag->ps_max = out[0];
      diag->ps_min = out[1];
      diag->u_max = out[2];
      diag->u_min = out[3];
      diag->v_max = out[4];
      diag->v_min = out[5];
      // any NaN in ps, or NaN/Inf extrema (the extrema skip NaN values)
      diag->cfl_violation = out[6] < (double)g.nj * g.ni || std::isnan(out[0]) || std::isnan(out[1]) ||
                            std::isinf(out[0]) || std::isinf(out[1]);
    }
    return MLC_OK;
  }

  int sync() override {
    HIP_TRY(hipStreamSynchronize(stream));
    HIP_TRY(hipGetLastError());
    return MLC_OK;
  }
};

}  // namespace mlc

// ================================================================= C ABI
struct mlc_handle {
  mlc::CoreBase *core;
  mlc_dims dims;
  mlc_params prm;
  std::vector<double> zita, zitah, ffilt;
};

using namespace mlc;

extern "C" int mlc_default_params(mlc_params *p) {
  if (!p) return mlc_fail(MLC_EINVAL, "mlc_default_params: NULL");
  std::memset(p, 0, sizeof(*p));
  p->dt = 30.0;
  p->dx = 3000.0;
  p->ztop = 36000.0;
  p->zh = 12000.0;
  p->a0 = 0.0;
  p->anu2 = 0.6;
  p->nadv = 3;
  p->nsound = 5;
  p->iqfrst = 2;
  p->rotllr = 1;
  p->divfilter = 0;
  p->halo_mpi = 0;
  p->fulleq = 1;
  p->divdamp = 1;
  p->vadvtwice = 1;
  p->filterpai = 0;
  p->filtertheta = 0;
  p->precision = MLC_FP64;
  p->device = 0;
  p->unfused = 0;
  return MLC_OK;
}

extern "C" int mlc_create(const mlc_dims *dims, const mlc_params *params, mlc_handle **out) {
  if (!dims || !params || !out) return mlc_fail(MLC_EINVAL, "mlc_create: NULL argument");
  *out = nullptr;
  if (dims->jx < 3 || dims->iy < 3 || dims->kz < 3 || dims->kz > 1024)
    return mlc_fail(MLC_EINVAL, "mlc_create: need jx, iy >= 3 and 3 <= kz <= 1024 (got %d, %d, %d)", dims->jx,
                    dims->iy, dims->kz);
  if (dims->nqx != 5) return mlc_fail(MLC_EINVAL, "mlc_create: nqx must be 5 (qv, qc, qi, qr, qs)");
  if (params->nadv < 1 || params->nsound < 1 || !(params->dt > 0) || !(params->dx > 0) || !(params->dzita > 0))
    return mlc_fail(MLC_EINVAL, "mlc_create: nadv, nsound >= 1 and dt, dx, dzita > 0 required");
  if (params->iqfrst < 2 || params->iqfrst > 6)
    return mlc_fail(MLC_EINVAL, "mlc_create: iqfrst must be in 2..6");
  if (!params->divdamp)
    return mlc_fail(MLC_EINVAL, "mlc_create: divdamp = 0 is not supported (the reference fixes do_divdamp = .true.)");
  if (params->precision != MLC_FP64 && params->precision != MLC_FP32)
    return mlc_fail(MLC_EINVAL, "mlc_create: precision must be MLC_FP64 or MLC_FP32");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return mlc_fail(MLC_EHIP, "mlc_create: no HIP device");
  if (params->device < 0 || params->device >= ndev) return mlc_fail(MLC_EINVAL, "mlc_create: bad device ordinal");
  mlc_handle *h = new (std::nothrow) mlc_handle();
  if (!h) return mlc_fail(MLC_ENOMEM, "mlc_create: out of host memory");
  h->dims = *dims;
  h->prm = *params;
  int rc;
  if (params->precision == MLC_FP64) {
    auto *c = new (std::nothrow) Core<double>();
    h->core = c;
    rc = c ? c->create(*dims, *params) : mlc_fail(MLC_ENOMEM, "out of host memory");
  } else {
    auto *c = new (std::nothrow) Core<float>();
    h->core = c;
    rc = c ? c->create(*dims, *params) : mlc_fail(MLC_ENOMEM, "out of host memory");
  }
  if (rc != MLC_OK) {
    delete h->core;
    delete h;
    return rc;
  }
  *out = h;
  return MLC_OK;
}

extern "C" int mlc_set_field(mlc_handle *h, int field, const double *host, const int *lb, const int *ub) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_set_field: NULL handle");
  return h->core->set_field(field, host, lb, ub);
}

extern "C" int mlc_get_field(mlc_handle *h, int field, double *host, const int *lb, const int *ub) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_get_field: NULL handle");
  return h->core->get_field(field, host, lb, ub);
}

extern "C" int mlc_set_levels(mlc_handle *h, int which, const double *values, int n) {
  if (!h || !values) return mlc_fail(MLC_EINVAL, "mlc_set_levels: NULL argument");
  int K = h->dims.kz;
  int want = which == MLC_ZITA ? K + 1 : K;
  if (which < 0 || which >= MLC_NLEVELS) return mlc_fail(MLC_EINVAL, "mlc_set_levels: bad level id %d", which);
  if (n != want) return mlc_fail(MLC_EINVAL, "mlc_set_levels: expected %d values, got %d", want, n);
  std::vector<double> v(values, values + n);
  if (h->prm.precision == MLC_FP64) {
    auto *c = static_cast<Core<double> *>(h->core);
    (which == MLC_ZITA ? c->zita : which == MLC_ZITAH ? c->zitah : c->ffilt) = v;
  } else {
    auto *c = static_cast<Core<float> *>(h->core);
    (which == MLC_ZITA ? c->zita : which == MLC_ZITAH ? c->zitah : c->ffilt) = v;
  }
  return MLC_OK;
}

extern "C" int mlc_init(mlc_handle *h) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_init: NULL handle");
  return h->core->init();
}

extern "C" int mlc_step(mlc_handle *h, int istep, int lprint, mlc_diag *diag) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_step: NULL handle");
  return h->core->step(istep, lprint, diag);
}

extern "C" int mlc_sync(mlc_handle *h) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_sync: NULL handle");
  return h->core->sync();
}

extern "C" int mlc_debug_poison_halos(mlc_handle *h, int enable) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_debug_poison_halos: NULL handle");
  return h->core->poison_halos(enable);
}

extern "C" int mlc_profile(mlc_handle *h, int enable) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_profile: NULL handle");
  return h->core->profile(enable);
}

extern "C" int mlc_profile_get(mlc_handle *h, int kind, const char **name, double *total_ms, long long *launches,
                               double *bytes) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_profile_get: NULL handle");
  if (kind < 0 || kind >= KK_COUNT) return mlc_fail(MLC_EINVAL, "mlc_profile_get: kind out of range");
  if (name) *name = kKernelNames[kind];
  return h->core->profile_get(kind, total_ms, launches, bytes);
}

extern "C" int mlc_profile_kinds(void) { return KK_COUNT; }

extern "C" int mlc_destroy(mlc_handle *h) {
  if (!h) return MLC_OK;
  delete h->core;
  delete h;
  return MLC_OK;
}

extern "C" int mlc_case_load(mlc_handle *h, const mlc_case *c) {
  if (!h || !c) return mlc_fail(MLC_EINVAL, "mlc_case_load: NULL argument");
  if (c->jx != h->dims.jx || c->iy != h->dims.iy || c->kz != h->dims.kz)
    return mlc_fail(MLC_EINVAL, "mlc_case_load: case grid differs from the handle's");
  const int jx = c->jx, iy = c->iy, kz = c->kz;
  const long plane = (long)(jx + 4) * (iy + 4);
  std::vector<std::vector<double>> buf(MLC_NFIELDS);
  std::vector<double *> fp(MLC_NFIELDS);
  auto nlev = [&](int f) {
    switch (f) {
      case MLC_W:
      case MLC_FMZF: return kz + 1;
      case MLC_PS: case MLC_HX: case MLC_HY: case MLC_MSFU: case MLC_MSFV: case MLC_MSFX:
      case MLC_ULAT: case MLC_VLAT: return 1;
      default: return kz;
    }
  };
  for (int f = 0; f < MLC_NFIELDS; f++) {
    buf[f].assign(plane * nlev(f), 0.0);
    fp[f] = buf[f].data();
  }
  std::vector<double> zita(kz + 1), zitah(kz), ffilt(kz);
  double *lv[3] = {zita.data(), zitah.data(), ffilt.data()};
  int rc = mlc_case_build(c, fp.data(), lv);
  if (rc) return rc;
  for (int f = 0; f < MLC_NFIELDS; f++) {
    int lb[3] = {-1, -1, 1}, ub[3] = {jx + 2, iy + 2, nlev(f)};
    if ((rc = mlc_set_field(h, f, fp[f], lb, ub)) != MLC_OK) return rc;
  }
  if ((rc = mlc_set_levels(h, MLC_ZITA, zita.data(), kz + 1))) return rc;
  if ((rc = mlc_set_levels(h, MLC_ZITAH, zitah.data(), kz))) return rc;
  if ((rc = mlc_set_levels(h, MLC_FFILT, ffilt.data(), kz))) return rc;
  return mlc_init(h);
}

// ============================================== learned acoustic corrector
extern "C" int mlc_lac_default_config(mlc_lac_config *c) {
  if (!c) return mlc_fail(MLC_EINVAL, "mlc_lac_default_config: NULL");
  c->nhid = 512;
  c->samples_per_phase = 4096;
  c->capacity = 131072;
  c->chunk_rows = 65536;
  c->check_every = 50;
  c->check_tol = 1.0;
  c->seed = 1;
  c->engine = 1;
  c->rel_tol = 1e-3;
  c->retrain_steps = 2;
  c->retrain_iters = 1500;
  c->retrain_batch = 1024;
  c->retrain_lr = 3e-4;
  c->max_retrains = 3;
  return MLC_OK;
}

extern "C" int mlc_lac_create(mlc_handle *h, const mlc_lac_config *c) {
  if (!h || !c) return mlc_fail(MLC_EINVAL, "mlc_lac_create: NULL argument");
  return h->core->lac_create(*c);
}

extern "C" int mlc_lac_collect(mlc_handle *h, int nsteps) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_lac_collect: NULL handle");
  return h->core->lac_collect(nsteps);
}

extern "C" int mlc_lac_train(mlc_handle *h, int iters, int batch, double lr, double *loss) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_lac_train: NULL handle");
  return h->core->lac_train(iters, batch, lr, loss);
}

extern "C" int mlc_lac_mode(mlc_handle *h, int mode) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_lac_mode: NULL handle");
  return h->core->lac_mode(mode);
}

extern "C" int mlc_lac_get_stats(mlc_handle *h, mlc_lac_stats *s) {
  if (!h || !s) return mlc_fail(MLC_EINVAL, "mlc_lac_get_stats: NULL argument");
  return h->core->lac_stats(s);
}

extern "C" mlc_mlp *mlc_lac_network(mlc_handle *h) { return h ? h->core->lac_network() : nullptr; }

extern "C" int mlc_lac_dataset(mlc_handle *h, long first, int n, float *X, float *T, int *cols) {
  if (!h) return mlc_fail(MLC_EINVAL, "mlc_lac_dataset: NULL handle");
  return h->core->lac_dataset(first, n, X, T, cols);
}

// ================================================================= output
struct mlc_output {
  mlc_handle *h = nullptr;
  mlc_ncfile *fps = nullptr, *ft = nullptr;
  int device = 0;
  long n2 = 0, n3 = 0;  // values of ps and of t per record
  hipStream_t ost = nullptr;
  hipEvent_t ready[2] = {}, copied[2] = {};
  double *dstage[2] = {}, *hbuf[2] = {};
  bool busy[2] = {false, false};
  int next = 0;
  std::deque<int> queue;
  std::mutex m;
  std::condition_variable cv;
  bool stop = false;
  std::string err;
  std::thread writer;

  void run() {
    (void)hipSetDevice(device);
    for (;;) {
      int s;
      {
        std::unique_lock<std::mutex> lk(m);
        cv.wait(lk, [&] { return stop || !queue.empty(); });
        if (queue.empty()) return;
        s = queue.front();
        queue.pop_front();
      }
      std::string e;
      if (hipEventSynchronize(copied[s]) != hipSuccess)
        e = "device-to-host copy of an output record failed";
      else if (mlc_nc_append(fps, hbuf[s]) != MLC_OK || mlc_nc_append(ft, hbuf[s] + n2) != MLC_OK)
        e = mlc_last_error();
      std::lock_guard<std::mutex> lk(m);
      if (!e.empty() && err.empty()) err = e;
      busy[s] = false;
      cv.notify_all();
    }
  }
};

static void output_release(mlc_output *o) {
  if (o->writer.joinable()) {
    {
      std::lock_guard<std::mutex> lk(o->m);
      o->stop = true;
    }
    o->cv.notify_all();
    o->writer.join();
  }
  for (int s = 0; s < 2; s++) {
    if (o->hbuf[s]) (void)hipHostFree(o->hbuf[s]);
    if (o->dstage[s]) (void)hipFree(o->dstage[s]);
    if (o->ready[s]) (void)hipEventDestroy(o->ready[s]);
    if (o->copied[s]) (void)hipEventDestroy(o->copied[s]);
  }
  if (o->ost) (void)hipStreamDestroy(o->ost);
  (void)mlc_nc_close(o->fps);
  (void)mlc_nc_close(o->ft);
  delete o;
}

extern "C" int mlc_output_open(mlc_handle *h, const char *dir, mlc_output **out) {
  if (!h || !dir || !out) return mlc_fail(MLC_EINVAL, "mlc_output_open: NULL argument");
  *out = nullptr;
  auto *o = new (std::nothrow) mlc_output();
  if (!o) return mlc_fail(MLC_ENOMEM, "mlc_output_open: out of host memory");
  o->h = h;
  o->device = h->prm.device;
  o->n2 = (long)h->dims.jx * h->dims.iy;
  o->n3 = o->n2 * h->dims.kz;
  const std::string d(dir);
  int rc = mlc_nc_create((d + "/ps.nc").c_str(), "ps", h->dims.jx, h->dims.iy, 0, &o->fps);
  if (!rc) rc = mlc_nc_create((d + "/t.nc").c_str(), "t", h->dims.jx, h->dims.iy, h->dims.kz, &o->ft);
  const size_t bytes = (size_t)(o->n2 + o->n3) * sizeof(double);
  auto hip_ok = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && !rc) rc = mlc_fail(MLC_EHIP, "mlc_output_open: %s: %s", what, hipGetErrorString(e));
  };
  if (!rc) hip_ok(hipSetDevice(o->device), "hipSetDevice");
  if (!rc) hip_ok(hipStreamCreateWithFlags(&o->ost, hipStreamNonBlocking), "stream");
  for (int s = 0; s < 2 && !rc; s++) {
    hip_ok(hipEventCreateWithFlags(&o->ready[s], hipEventDisableTiming), "event");
    if (!rc) hip_ok(hipEventCreateWithFlags(&o->copied[s], hipEventDisableTiming), "event");
    if (!rc) hip_ok(hipMalloc((void **)&o->dstage[s], bytes), "device staging");
    if (!rc) hip_ok(hipHostMalloc((void **)&o->hbuf[s], bytes, hipHostMallocDefault), "pinned host buffer");
  }
  if (rc) {
    output_release(o);
    return rc;
  }
  o->writer = std::thread([o] { o->run(); });
  *out = o;
  return MLC_OK;
}

extern "C" int mlc_output_write(mlc_output *o) {
  if (!o) return mlc_fail(MLC_EINVAL, "mlc_output_write: NULL");
  const int s = o->next;
  {
    std::unique_lock<std::mutex> lk(o->m);
    o->cv.wait(lk, [&] { return !o->busy[s] || !o->err.empty(); });
    if (!o->err.empty()) return mlc_fail(MLC_EINVAL, "mlc_output_write: %s", o->err.c_str());
    o->busy[s] = true;
  }
  o->next ^= 1;
  // the slot's previous copy has finished (its file write has), so packing
  // over its staging buffer on the handle's stream is safe
  int rc = o->h->core->pack_output(o->dstage[s], o->ready[s]);
  hipError_t e = hipSuccess;
  if (!rc) e = hipStreamWaitEvent(o->ost, o->ready[s], 0);
  if (!rc && e == hipSuccess)
    e = hipMemcpyAsync(o->hbuf[s], o->dstage[s], (size_t)(o->n2 + o->n3) * sizeof(double), hipMemcpyDeviceToHost,
                       o->ost);
  if (!rc && e == hipSuccess) e = hipEventRecord(o->copied[s], o->ost);
  if (!rc && e != hipSuccess) rc = mlc_fail(MLC_EHIP, "mlc_output_write: %s", hipGetErrorString(e));
  std::lock_guard<std::mutex> lk(o->m);
  if (rc) {
    o->busy[s] = false;
    return rc;
  }
  o->queue.push_back(s);
  o->cv.notify_all();
  return MLC_OK;
}

extern "C" int mlc_output_flush(mlc_output *o) {
  if (!o) return mlc_fail(MLC_EINVAL, "mlc_output_flush: NULL");
  std::unique_lock<std::mutex> lk(o->m);
  o->cv.wait(lk, [&] { return (!o->busy[0] && !o->busy[1]) || !o->err.empty(); });
  if (!o->err.empty()) return mlc_fail(MLC_EINVAL, "mlc_output: %s", o->err.c_str());
  return MLC_OK;
}

extern "C" int mlc_output_close(mlc_output *o) {
  if (!o) return MLC_OK;
  int rc = mlc_output_flush(o);
  output_release(o);
  return rc;
}

extern "C" int mlc_case_load_device(mlc_handle *h, const mlc_case *c) {
  if (!h || !c) return mlc_fail(MLC_EINVAL, "mlc_case_load_device: NULL argument");
  if (c->jx != h->dims.jx || c->iy != h->dims.iy || c->kz != h->dims.kz)
    return mlc_fail(MLC_EINVAL, "mlc_case_load_device: case grid differs from the handle's");
  CaseHost H;
  int rc = case_host_part(c, H);
  if (rc) return rc;
  const int jx = c->jx, iy = c->iy, kz = c->kz;
  int lb[3] = {-1, -1, 1}, ub[3] = {jx + 2, iy + 2, 1};
  const std::pair<int, const std::vector<double> *> two[] = {
      {MLC_PS, &H.ps}, {MLC_HX, &H.hx}, {MLC_HY, &H.hy}, {MLC_MSFU, &H.msf}, {MLC_MSFV, &H.msf},
      {MLC_MSFX, &H.msf}, {MLC_ULAT, &H.ulat}, {MLC_VLAT, &H.vlat}};
  for (auto &f : two)
    if ((rc = mlc_set_field(h, f.first, f.second->data(), lb, ub))) return rc;
  std::vector<double> ffilt(kz, 0.0);
  if ((rc = mlc_set_levels(h, MLC_ZITA, H.zita.data(), kz + 1))) return rc;
  if ((rc = mlc_set_levels(h, MLC_ZITAH, H.zitah.data(), kz))) return rc;
  if ((rc = mlc_set_levels(h, MLC_FFILT, ffilt.data(), kz))) return rc;
  if ((rc = h->core->build_case(*c, H))) return rc;
  return mlc_init(h);
}
