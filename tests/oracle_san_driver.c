/*
 * oracle_san_driver.c -- runs the CPU oracle (oracle/moloch_oracle.c) as a
 * standalone program, so that it can be built with AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_oracle_sanitizers.py).  Test
 * infrastructure only.
 *
 * Every array is allocated with exactly the size the Python wrapper
 * (tests/oracle_lib.py) gives it, so an access outside the cells the oracle
 * owns is reported.
 *
 * Input file (native endianness):
 *   int32 nj, ni, nk, nsteps
 *   double dt, dx, dzita, ztop, zh, a0, anu2
 *   int32 nadv, nsound, nqx, iqfrst, rotllr, divfilter, halo_mpi,
 *         do_fulleq, do_divdamp, do_vadvtwice, do_filterpai, do_filtertheta
 *   then every array of the state in the order of NAMES below, whole.
 * Output file: the same arrays after init_moloch and nsteps steps (lprint on
 * the last), then the six diagnostics and the CFL flag (as a double).
 */
#include <stdio.h>
#include <stdlib.h>
#include <stddef.h>
#include <string.h>

#include "../oracle/moloch_oracle.h"

enum { K1D, K2D, K3D, K3D1, KQX, KP2D };

static const struct {
  const char *name;
  int kind;
  size_t off;
} NAMES[] = {
#define A(n, k) {#n, k, offsetof(ora_state, n)}
    A(zita, K1D),    A(zitah, K1D),  A(ffilt, K1D), A(gzitak, K1D), A(gzitakh, K1D), A(xknu, K1D),
    A(hx, K2D),      A(hy, K2D),     A(mu, K2D),    A(mv, K2D),     A(mx, K2D),      A(ulat, K2D),
    A(vlat, K2D),    A(ps, K2D),     A(rmu, K2D),   A(rmv, K2D),    A(mx2, K2D),     A(coru, K2D),
    A(corv, K2D),    A(fmz, K3D),    A(fmzf, K3D1), A(zeta, K3D),   A(u, K3D),       A(v, K3D),
    A(w, K3D1),      A(pai, K3D),    A(t, K3D),     A(tetav, K3D),  A(tvirt, K3D),   A(p, K3D),
    A(rho, K3D),     A(qsat, K3D),   A(ux, K3D),    A(vx, K3D),     A(qx, KQX),      A(s, K3D1),
    A(deltaw, K3D1), A(zdiv2, K3D),  A(wx, K3D),    A(wz, K3D),     A(p0, K3D),      A(ud, K3D),
    A(vd, K3D),      A(qwltot, K3D), A(qwitot, K3D), A(pf, K3D),    A(tf, K3D),      A(p2d, KP2D),
    A(wwkw, K3D1),   A(wfw, K3D1),   A(zpby, K3D),  A(zpbw, K3D),
#undef A
};
#define NARR (sizeof(NAMES) / sizeof(NAMES[0]))

static size_t count(const ora_state *S, int kind) {
  const size_t plane = (size_t)(S->nj + 4) * (size_t)(S->ni + 4);
  switch (kind) {
    case K1D: return (size_t)S->nk + 1;
    case K2D: case KP2D: return plane;
    case K3D: return plane * (size_t)S->nk;
    case K3D1: return plane * (size_t)(S->nk + 1);
    default: return plane * (size_t)S->nk * 5;
  }
}

static int rd(FILE *f, void *p, size_t n) { return fread(p, 1, n, f) == n ? 0 : -1; }

int main(int argc, char **argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s input output\n", argv[0]);
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f) return 2;
  ora_state S;
  memset(&S, 0, sizeof(S));
  int hdr[4], sw[12];
  double par[7];
  if (rd(f, hdr, sizeof(hdr)) || rd(f, par, sizeof(par)) || rd(f, sw, sizeof(sw))) return 2;
  S.nj = hdr[0], S.ni = hdr[1], S.nk = hdr[2];
  const int nsteps = hdr[3];
  S.dt = par[0], S.dx = par[1], S.dzita = par[2], S.ztop = par[3], S.zh = par[4], S.a0 = par[5], S.anu2 = par[6];
  S.nadv = sw[0], S.nsound = sw[1], S.nqx = sw[2], S.iqfrst = sw[3], S.rotllr = sw[4], S.divfilter = sw[5];
  S.halo_mpi = sw[6], S.do_fulleq = sw[7], S.do_divdamp = sw[8], S.do_vadvtwice = sw[9];
  S.do_filterpai = sw[10], S.do_filtertheta = sw[11];
  for (size_t a = 0; a < NARR; a++) {
    const size_t n = count(&S, NAMES[a].kind);
    double *p = (double *)malloc(n * sizeof(double));
    if (!p || rd(f, p, n * sizeof(double))) return 2;
    memcpy((char *)&S + NAMES[a].off, &p, sizeof(p));
  }
  fclose(f);

  ora_init_moloch(&S);
  for (int n = 1; n <= nsteps; n++) ora_moloch(&S, n, n == nsteps);

  FILE *o = fopen(argv[2], "wb");
  if (!o) return 2;
  for (size_t a = 0; a < NARR; a++) {
    double *p;
    memcpy(&p, (char *)&S + NAMES[a].off, sizeof(p));
    if (fwrite(p, sizeof(double), count(&S, NAMES[a].kind), o) != count(&S, NAMES[a].kind)) return 2;
    free(p);
  }
  double tail[7];
  memcpy(tail, S.diag, sizeof(S.diag));
  tail[6] = (double)S.cfl_violation;
  if (fwrite(tail, sizeof(double), 7, o) != 7) return 2;
  fclose(o);
  return 0;
}
