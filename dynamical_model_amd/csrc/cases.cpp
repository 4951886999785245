// cases.cpp -- host builder of the geometry and initial state of a run.
//
// Geometry follows `param` (nompi/regcm_dynamical_core.F90:93-194) and the
// vertical coordinate of mod_zita.F90; the Exner profile uses the discrete
// hydrostatic recursion of fill_atmosphere (mod_atmosphere.F90:210-234).
// MLC_CASE_REFERENCE reproduces fill_atmosphere's state (:171-235) exactly.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "mlc_constants.h"
#include "mlc_internal.h"
#include "moloch_hip.h"

using namespace mlc;

namespace {

// ---------------------------------------------------------------- mod_zita
double zfz(double ztop, double zh) { return ztop / (std::exp(ztop / zh) - 1.0); }  // :87-92
double bzita(double z, double ztop, double zh) { return zfz(ztop, zh) * (std::exp(z / zh) - 1.0); }  // :94-99
double bzitap(double z, double ztop, double zh) { return zfz(ztop, zh) * std::exp(z / zh) / zh; }  // :101-106
double gzitap(double z, double ztop, double a0) {  // :123-133
  double ratio = z / ztop;
  return ((0.0 - 1.0 * a0) * 1.0 - (6.0 - 4.0 * a0) * ratio + (6.0 - 3.0 * a0) * (ratio * ratio)) / ztop;
}
// md_zeta / md_fmz with geopotential orography (:136-179)
double md_fmz(double z, double geopot, double ztop, double zh, double a0) {
  double orog = geopot * kRegrav;
  return 1.0 / (gzitap(z, ztop, a0) * orog + bzitap(z, ztop, zh));
}
double md_zeta(double z, double geopot, double ztop, double zh, double a0) {
  double orog = geopot * kRegrav;
  return orog * (gzita_host(z, ztop, a0) - 1.0) + bzita(z, ztop, zh);
}

// splitmix64 -> uniform [0,1)
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

// Canonical host layout (-1:jx+2, -1:iy+2, 1:nlev)
struct HostLayout {
  int jx, iy;
  long pj, plane;
  HostLayout(int jx_, int iy_) : jx(jx_), iy(iy_), pj(jx_ + 4), plane((long)(jx_ + 4) * (iy_ + 4)) {}
  long at(int j, int i, int k = 1) const { return (long)(k - 1) * plane + (long)(i + 1) * pj + (j + 1); }
};

// Fill every halo cell of the halo-2 extent periodically from the interior.
void periodic_fill(const HostLayout &L, double *f, int nlev) {
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= nlev; k++)
    for (int i = -1; i <= L.iy + 2; i++)
      for (int j = -1; j <= L.jx + 2; j++) {
        if (j >= 1 && j <= L.jx && i >= 1 && i <= L.iy) continue;
        int jj = ((j - 1) % L.jx + L.jx) % L.jx + 1;
        int ii = ((i - 1) % L.iy + L.iy) % L.iy + 1;
        f[L.at(j, i, k)] = f[L.at(jj, ii, k)];
      }
}

// Standard atmosphere: temperature and dry pressure at height z (m).
struct StdAtm {
  double tsfc, lapse, ttrop, psfc;
  double ztrop() const { return (tsfc - ttrop) / lapse; }
  double T(double z) const { return std::fmax(tsfc - lapse * z, ttrop); }
  double p(double z) const {
    double expo = kEgrav / (kRgas * lapse);
    double zt = ztrop();
    if (z <= zt) return psfc * std::pow(T(z) / tsfc, expo);
    double pt = psfc * std::pow(ttrop / tsfc, expo);
    return pt * std::exp(-kEgrav * (z - zt) / (kRgas * ttrop));
  }
};

}  // namespace

// gzita (mod_zita.F90:110-120), shared with the device set-up.
double gzita_host(double zita, double ztop, double a0) {
  double ratio = zita / ztop;
  return ((0.0 - 1.0 * a0) * ratio - (3.0 - 2.0 * a0) * (ratio * ratio) +
          (2.0 - 1.0 * a0) * (ratio * ratio * ratio)) +
         1.0;
}

// pfesat / pfwsat (nompi/mod_moloch.F90:119-181), host copy for the builder.
double pfwsat_host(double t, double p) {
  static const double a[9] = {0.611213476e+03, 0.444007856e+02, 0.143064234e+01,
                              0.264461437e-01, 0.305903558e-03, 0.196237241e-05,
                              0.892344772e-08, -0.373208410e-10, 0.209339997e-13};
  static const double c[9] = {0.611123516e+03, 0.503109514e+02, 0.188369801e+01,
                              0.420547422e-01, 0.614396778e-03, 0.602780717e-05,
                              0.387940929e-07, 0.149436277e-09, 0.262655803e-12};
  double td = t - kTzero;
  if (td > 100.0) td = 100.0;
  if (td < -75.0) td = -75.0;
  const double *q = td >= 0.0 ? a : c;
  double esat = q[0] + td * (q[1] + td * (q[2] + td * (q[3] + td * (q[4] + td * (q[5] + td * (q[6] + td * (q[7] + td * q[8])))))));
  double es = std::fmin(esat, 0.15 * p);
  return kEp2 * (es / (p - es));
}

extern "C" int mlc_case_config(int config, mlc_case *c) {
  if (!c) return mlc_fail(MLC_EINVAL, "mlc_case_config: NULL case");
  std::memset(c, 0, sizeof(*c));
  c->kind = MLC_CASE_STANDARD;
  c->ztop = 36000.0;
  c->zh = 12000.0;
  c->a0 = 0.0;
  c->nhills = 1;
  c->hill_h = 500.0;
  c->hill_hw = 10.0;
  c->tsfc = 288.0;
  c->lapse = 0.0065;
  c->ttrop = 216.65;
  c->psfc = 100000.0;
  c->rh = 0.5;
  c->u0 = 10.0;
  c->v0 = 0.0;
  c->jet_amp = 0.0;
  c->jet_z = 10000.0;
  c->jet_width = 3000.0;
  c->noise_amp = 0.1;
  c->noise_levels = 5;
  c->bubble_amp = 0.0;
  c->bubble_rcells = 20.0;
  c->bubble_z = 1500.0;
  c->bubble_rz = 1500.0;
  c->zero_coriolis = 0;
  switch (config) {
    case 0:  // 100x100x30, one 500 m hill of half-width 10 cells
      c->jx = c->iy = 100;
      c->kz = 30;
      c->seed = 1;
      break;
    case 1:  // 512x512x40, hill half-width 40, noise 0.5 K, 2 K bubble
      c->jx = c->iy = 512;
      c->kz = 40;
      c->hill_hw = 40.0;
      c->noise_amp = 0.5;
      c->bubble_amp = 2.0;
      c->seed = 2;
      break;
    case 2:  // 1536x1536x60, 20 km top, 20 random hills, jet, RH 60 %
    case 3:  // training set-up: configuration 2 at 512x512x40
    case 4:  // 2048x2048x60, configuration 2 with a new seed
      c->jx = c->iy = (config == 2) ? 1536 : (config == 3 ? 512 : 2048);
      c->kz = (config == 3) ? 40 : 60;
      c->ztop = 20000.0;
      c->zh = 20000.0 / 3.0;
      c->nhills = 20;
      c->hill_hmin = 200.0;
      c->hill_hmax = 2000.0;
      c->hill_hwmin = 8.0;
      c->hill_hwmax = 40.0;
      c->jet_amp = 30.0;
      c->rh = 0.6;
      c->noise_amp = 0.2;
      c->seed = (config == 4) ? 4 : 3;
      break;
    default:
      return mlc_fail(MLC_EINVAL, "mlc_case_config: configuration must be 0..4");
  }
  return MLC_OK;
}

extern "C" int mlc_case_params(const mlc_case *c, mlc_params *p) {
  if (!c || !p) return mlc_fail(MLC_EINVAL, "mlc_case_params: NULL argument");
  p->ztop = c->ztop;
  p->zh = c->zh;
  p->a0 = c->a0;
  p->dzita = c->ztop / (double)c->kz;  // zita(kz) of model_zitaf (mod_zita.F90:57-70)
  return MLC_OK;
}

// Levels, 2-D geometry and surface pressure of a case (canonical layout
// (-1:jx+2, -1:iy+2)); *rng receives the generator state after the terrain
// draws, where the theta noise continues.
int case_host_part(const mlc_case *c, CaseHost &o) {
  if (!c) return mlc_fail(MLC_EINVAL, "mlc_case_build: NULL case");
  const int jx = c->jx, iy = c->iy, kz = c->kz;
  if (jx < 3 || iy < 3 || kz < 3 || kz > 1024)
    return mlc_fail(MLC_EINVAL, "mlc_case_build: need jx, iy >= 3 and 3 <= kz <= 1024");
  HostLayout L(jx, iy);
  const long n2 = L.plane;
  o.zita.assign(kz + 1, 0.0);
  o.zitah.assign(kz, 0.0);
  // model_zitaf / model_zitah (mod_zita.F90:57-85)
  {
    double dz = c->ztop / (double)kz;
    o.zita[kz] = 0.0;
    for (int k = kz; k >= 1; k--) o.zita[k - 1] = o.zita[k] + dz;
    o.zitah[kz - 1] = dz * 0.5;
    for (int k = kz - 1; k >= 1; k--) o.zitah[k - 1] = o.zitah[k] + dz;
  }
  o.ht.assign(n2, 0.0);
  o.hx.assign(n2, 0.0);
  o.hy.assign(n2, 0.0);
  o.msf.assign(n2, 1.0);
  o.ulat.assign(n2, 0.0);
  o.vlat.assign(n2, 0.0);
  o.ps.assign(n2, 0.0);
  std::vector<double> &ht = o.ht, &hx = o.hx, &hy = o.hy, &msf = o.msf, &ulat = o.ulat, &vlat = o.vlat;
  Rng rng(c->seed);

  // terrain (geopotential, as `param` stores mddom%ht)
  if (c->nhills == 1 && c->hill_hw > 0.0) {
    for (int i = 1; i <= iy; i++)
      for (int j = 1; j <= jx; j++) {
        double rr = std::sqrt(((double)j - 0.5 * (double)(jx + 1)) * ((double)j - 0.5 * (double)(jx + 1)) +
                              ((double)i - 0.5 * (double)(iy + 1)) * ((double)i - 0.5 * (double)(iy + 1)));
        ht[L.at(j, i)] = kEgrav * c->hill_h * std::exp(-std::log(2.0) * ((rr / c->hill_hw) * (rr / c->hill_hw)));
      }
  } else if (c->nhills > 1) {
    std::vector<double> h(n2, 0.0);
    for (int n = 0; n < c->nhills; n++) {
      double xc = rng.uniform(0.0, (double)jx), yc = rng.uniform(0.0, (double)iy);
      double hh = rng.uniform(c->hill_hmin, c->hill_hmax);
      double hw = rng.uniform(c->hill_hwmin, c->hill_hwmax);
#pragma omp parallel for schedule(static)
      for (int i = 1; i <= iy; i++)
        for (int j = 1; j <= jx; j++) {
          double dxp = std::fabs((double)j - xc), dyp = std::fabs((double)i - yc);
          dxp = std::fmin(dxp, (double)jx - dxp);
          dyp = std::fmin(dyp, (double)iy - dyp);
          double r2 = (dxp * dxp + dyp * dyp) / (hw * hw);
          h[L.at(j, i)] += hh * std::exp(-std::log(2.0) * r2);
        }
    }
    for (long x = 0; x < n2; x++) ht[x] = kEgrav * h[x];
  }
  periodic_fill(L, ht.data(), 1);

  // latitudes as in param (:139-147), map factors 1 (:129-131)
  {
    double dl = kRaddeg * 3000.0 / kEarthrad;
    for (int i = -1; i <= iy + 2; i++)
      for (int j = -1; j <= jx + 2; j++) {
        if (!c->zero_coriolis) {
          ulat[L.at(j, i)] = dl * ((double)i - (double)iy * 0.5 - 0.5);
          vlat[L.at(j, i)] = dl * ((double)i - (double)iy * 0.5);
        }
      }
  }
  // terrain slopes (:148-162), then their periodic halos (:163-166)
  {
    double rdx = 1.0 / 3000.0;
    for (int i = 1; i <= iy; i++)
      for (int j = 1; j <= jx; j++) {
        hx[L.at(j, i)] = (ht[L.at(j, i)] - ht[L.at(j - 1, i)]) * msf[L.at(j, i)] * rdx * kRegrav;
        hy[L.at(j, i)] = (ht[L.at(j, i)] - ht[L.at(j, i - 1)]) * rdx * kRegrav;
      }
    periodic_fill(L, hx.data(), 1);
    periodic_fill(L, hy.data(), 1);
  }
  // surface pressure: fill_atmosphere's constant (mod_atmosphere.F90:184) or
  // the standard atmosphere at the ground
  if (c->kind == MLC_CASE_REFERENCE) {
    for (long x = 0; x < n2; x++) o.ps[x] = 101480.0;
  } else {
    StdAtm A{c->tsfc, c->lapse, c->ttrop, c->psfc};
    for (int i = 1; i <= iy; i++)
      for (int j = 1; j <= jx; j++) o.ps[L.at(j, i)] = A.p(ht[L.at(j, i)] * kRegrav);
    periodic_fill(L, o.ps.data(), 1);
  }
  o.rng = rng.s;
  return MLC_OK;
}

extern "C" int mlc_case_build(const mlc_case *c, double *const *fields, double *const *levels) {
  CaseHost H;
  int rc = case_host_part(c, H);
  if (rc) return rc;
  const int jx = c->jx, iy = c->iy, kz = c->kz;
  HostLayout L(jx, iy);
  const long n2 = L.plane, n3 = L.plane * kz, n3p = L.plane * (kz + 1);
  std::vector<double> &zita = H.zita, &zitah = H.zitah;
  const double dzita = zita[kz - 1];
  std::vector<double> &ht = H.ht, &hx = H.hx, &hy = H.hy, &msf = H.msf, &ulat = H.ulat, &vlat = H.vlat;
  std::vector<double> zeta(n3, 0.0), fmz(n3, 0.0), fmzf(n3p, 0.0), zabs(n3, 0.0);
  std::vector<double> u(n3, 0.0), v(n3, 0.0), w(n3p, 0.0), pai(n3, 0.0), t(n3, 0.0), qv(n3, 0.0);
  std::vector<double> &ps = H.ps;
  Rng rng(H.rng);
  (void)n2;

  // zeta, fmz (:167-172), fmzf (:185-190)
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= kz; k++)
    for (int i = 1; i <= iy; i++)
      for (int j = 1; j <= jx; j++) {
        double g = ht[L.at(j, i)];
        zeta[L.at(j, i, k)] = md_zeta(zitah[k - 1], g, c->ztop, c->zh, c->a0);
        fmz[L.at(j, i, k)] = md_fmz(zitah[k - 1], g, c->ztop, c->zh, c->a0);
        zabs[L.at(j, i, k)] = zeta[L.at(j, i, k)] + g * kRegrav;
      }
#pragma omp parallel for schedule(static)
  for (int k = 1; k <= kz + 1; k++)
    for (int i = 1; i <= iy; i++)
      for (int j = 1; j <= jx; j++) fmzf[L.at(j, i, k)] = md_fmz(zita[k - 1], ht[L.at(j, i)], c->ztop, c->zh, c->a0);
  periodic_fill(L, zeta.data(), kz);
  periodic_fill(L, fmz.data(), kz);
  periodic_fill(L, fmzf.data(), kz + 1);
  periodic_fill(L, zabs.data(), kz);

  if (c->kind == MLC_CASE_REFERENCE) {
    // fill_atmosphere (mod_atmosphere.F90:171-209)
    const double ts = 300.0, qs0 = 0.01865, ps0 = 101480.0, tropoh = 15000.0;
    double tvs = ts * (1.0 + kEp1 * qs0);
    double tvt = tvs - kLrate * tropoh;
    (void)ps0;
    for (int i = 1; i <= iy; i++)
      for (int j = 1; j <= jx; j++)
        for (int k = 1; k <= kz; k++) {
          long x = L.at(j, i, k);
          if (zeta[x] < tropoh) {
            double zi = zeta[x];
            double qi = qs0 * std::exp(-zi / 4000.0) * std::exp((-zi / 7500.0) * (-zi / 7500.0));
            qv[x] = qi;
            t[x] = (tvs - kLrate * zi) / (1.0 + qi);
          } else {
            qv[x] = 1.0e-8;
            t[x] = tvt;
          }
        }
    int mxc = jx / 2, myc = iy / 2;
    // mr = sqrt((real(jx)/5.0_rkx)**2+(real(iy)/5_rkx)**2)/2.0_rkx (:196):
    // REAL() is default (single) precision; the first quotient is promoted
    // to double, the second divides by the integer 5 and stays single.
    float fjx = (float)jx, fiy = (float)iy;
    double a = (double)fjx / 5.0;
    float b = fiy / 5.0f;
    float b2 = b * b;
    double mr = std::sqrt(a * a + (double)b2) / 2.0;
    for (int i = 1; i <= iy; i++)
      for (int j = 1; j <= jx; j++) {
        // md = sqrt(real(j-mx)**2+real(i-my)**2) in single precision (:199)
        float dj = (float)(j - mxc), di = (float)(i - myc);
        double md = (double)std::sqrt(dj * dj + di * di);
        if (md < mr) t[L.at(j, i, kz)] += (mr - md) / mr;
      }
    for (long x = 0; x < n3; x++) {
      u[x] = 1.0;
      v[x] = 1.0;
    }
  } else {
    StdAtm A{c->tsfc, c->lapse, c->ttrop, c->psfc};
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= kz; k++)
      for (int i = 1; i <= iy; i++)
        for (int j = 1; j <= jx; j++) {
          long x = L.at(j, i, k);
          double z = zabs[x];
          t[x] = A.T(z);
          qv[x] = c->rh * pfwsat_host(t[x], A.p(z));
        }
    // wind: u on the west face of cell j, heights averaged over j-1 and j
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= kz; k++)
      for (int i = 1; i <= iy; i++)
        for (int j = 1; j <= jx; j++) {
          long x = L.at(j, i, k);
          double zf = 0.5 * (zabs[L.at(j - 1, i, k)] + zabs[x]);
          double uj = c->u0;
          if (c->jet_amp != 0.0) {
            double e = (zf - c->jet_z) / c->jet_width;
            uj += c->jet_amp * std::exp(-e * e);
          }
          u[x] = uj;
          v[x] = c->v0;
        }
  }

  // discrete hydrostatic Exner function (mod_atmosphere.F90:210-234)
#pragma omp parallel for schedule(static)
  for (int i = 1; i <= iy; i++)
    for (int j = 1; j <= jx; j++) {
      long xk = L.at(j, i, kz), xk1 = L.at(j, i, kz - 1);
      double zdelta = zeta[xk] * kEgrav;
      double tv1 = t[xk] * (1.0 + kEp1 * qv[xk]);
      double tv2 = t[xk1] * (1.0 + kEp1 * qv[xk1]);
      double lrt = (tv2 - tv1) / (zeta[xk1] - zeta[xk]);
      lrt = 0.65 * lrt - 0.35 * kLrate;
      double tv = tv1 - 0.5 * zeta[xk] * lrt;
      double zz = 1.0 / (kRgas * tv);
      double p = ps[L.at(j, i)] * std::exp(-zdelta * zz);
      pai[xk] = std::pow(p / kP00, kRovcp);
    }
  for (int k = kz - 1; k >= 1; k--)
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= iy; i++)
      for (int j = 1; j <= jx; j++) {
        long x = L.at(j, i, k), xp = L.at(j, i, k + 1);
        double tv1 = t[x] * (1.0 + kEp1 * qv[x]);
        double tv2 = t[xp] * (1.0 + kEp1 * qv[xp]);
        double zb = 2.0 * kEgrav * dzita / (fmzf[xp] * kCpd) + tv1 - tv2;
        double zdelta = std::sqrt(zb * zb + 4.0 * tv2 * tv1);
        pai[x] = -pai[xp] / (2.0 * tv2) * (zb - zdelta);
      }

  if (c->kind == MLC_CASE_STANDARD) {
    // potential-temperature perturbations, applied as T' = theta' * pai
    if (c->noise_amp > 0.0) {
      int k0 = kz - c->noise_levels + 1;
      if (k0 < 1) k0 = 1;
      for (int k = k0; k <= kz; k++)
        for (int i = 1; i <= iy; i++)
          for (int j = 1; j <= jx; j++) {
            long x = L.at(j, i, k);
            t[x] += rng.uniform(-c->noise_amp, c->noise_amp) * pai[x];
          }
    }
    if (c->bubble_amp != 0.0) {
      double jc = 0.5 * (double)(jx + 1), ic = 0.5 * (double)(iy + 1);
#pragma omp parallel for schedule(static)
      for (int k = 1; k <= kz; k++)
        for (int i = 1; i <= iy; i++)
          for (int j = 1; j <= jx; j++) {
            long x = L.at(j, i, k);
            double dh = std::sqrt(((double)j - jc) * ((double)j - jc) + ((double)i - ic) * ((double)i - ic)) / c->bubble_rcells;
            double dz = (zeta[x] - c->bubble_z) / c->bubble_rz;
            double r = std::sqrt(dh * dh + dz * dz);
            if (r < 1.0) {
              double cc = std::cos(0.5 * kMathpi * r);
              t[x] += c->bubble_amp * cc * cc * pai[x];
            }
          }
    }
  }

  periodic_fill(L, u.data(), kz);
  periodic_fill(L, v.data(), kz);
  periodic_fill(L, pai.data(), kz);
  periodic_fill(L, t.data(), kz);
  periodic_fill(L, qv.data(), kz);

  auto put = [&](int f, const std::vector<double> &src) {
    if (fields && fields[f]) std::memcpy(fields[f], src.data(), src.size() * sizeof(double));
  };
  auto zero = [&](int f, long n) {
    if (fields && fields[f]) std::memset(fields[f], 0, n * sizeof(double));
  };
  put(MLC_U, u);
  put(MLC_V, v);
  put(MLC_W, w);
  put(MLC_PAI, pai);
  put(MLC_T, t);
  put(MLC_QV, qv);
  zero(MLC_TETAV, n3);
  zero(MLC_QC, n3);
  zero(MLC_QI, n3);
  zero(MLC_QR, n3);
  zero(MLC_QS, n3);
  zero(MLC_TVIRT, n3);
  zero(MLC_P, n3);
  zero(MLC_RHO, n3);
  zero(MLC_QSAT, n3);
  zero(MLC_UX, n3);
  zero(MLC_VX, n3);
  put(MLC_PS, ps);
  put(MLC_FMZ, fmz);
  put(MLC_FMZF, fmzf);
  put(MLC_ZETA, zeta);
  put(MLC_HX, hx);
  put(MLC_HY, hy);
  put(MLC_MSFU, msf);
  put(MLC_MSFV, msf);
  put(MLC_MSFX, msf);
  put(MLC_ULAT, ulat);
  put(MLC_VLAT, vlat);
  if (levels) {
    if (levels[MLC_ZITA]) std::memcpy(levels[MLC_ZITA], zita.data(), (kz + 1) * sizeof(double));
    if (levels[MLC_ZITAH]) std::memcpy(levels[MLC_ZITAH], zitah.data(), kz * sizeof(double));
    if (levels[MLC_FFILT]) std::memset(levels[MLC_FFILT], 0, kz * sizeof(double));
  }
  return MLC_OK;
}
