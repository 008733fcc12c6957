/* kmm_cpu.c -- CPU restatement in C (FP64, OpenMP) of the reference's KMM
 * channel time step for the channel + passive scalar + Goldstein
 * immersed-boundary subset (FFTW_DNS/code/src/dns/main_dns.f90:198-811 and
 * the routines it calls).  TEST INFRASTRUCTURE AND CPU BASELINE ONLY: it is
 * checked against the numpy oracle (oracle/kmm_oracle.py, whose functions it
 * follows pass by pass) and timed as the benchmark's CPU leg; the product
 * path never links or calls it.
 *
 * Layout (as the numpy oracle): spectral fields complex (nxh, nz, nyp) with
 * the Chebyshev index fastest; physical fields and IB integrals (mx, mz, nyp).
 * Transforms: own mixed-radix (4, 2, 3, 5) Stockham FFTs; the Chebyshev
 * transform (DCT-I) through a length-2 ny FFT of the even extension; x c2r /
 * r2c with two real rows packed in one complex transform.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef double complex cplx;

typedef struct {
  int nx, ny, nz;
  double xl, yl, zl, re, dt, theta, dPdx, gain, ugain;
  int flow_select, perturbtime;
  int scalar, scl_flag;
  double diff, deltaT;
  double atu, btu, gtu, abu, bbu, gbu, atw, btw, gtw, abw, bbw, gbw;
  int threads;
} kc_params;

typedef struct {
  cplx *v, *fn, *fnm1, *omy, *gn, *gnm1, *sc, *scn, *scnm1;
  double *u0, *w0, *h1n, *h1nm1, *h3n, *h3nm1;
  double *fxi, *fyi, *fzi;          /* (mx, mz, nyp) */
  const int *imat;                  /* (mx, mz, nyp) or NULL */
  const cplx *gf1, *gf2;            /* Green's functions (nxh, nz, nyp) */
  const cplx *dnv1b, *dnv1t, *dnv2b, *dnv2t;  /* (nxh, nz) */
} kc_state;

/* ------------------------------------------------------------------ FFT -- */
typedef struct {
  int n, nst, r[32];
  cplx *tw; /* exp(-2 pi i k / n) */
} fftplan;

static void plan_init(fftplan *P, int n) {
  P->n = n;
  P->nst = 0;
  int m = n;
  while (m % 4 == 0) { P->r[P->nst++] = 4; m /= 4; }
  while (m % 2 == 0) { P->r[P->nst++] = 2; m /= 2; }
  while (m % 3 == 0) { P->r[P->nst++] = 3; m /= 3; }
  while (m % 5 == 0) { P->r[P->nst++] = 5; m /= 5; }
  P->tw = malloc(sizeof(cplx) * n);
  for (int k = 0; k < n; ++k) P->tw[k] = cexp(-2.0 * M_PI * I * (double)k / (double)n);
}
static void plan_free(fftplan *P) { free(P->tw); }

/* unnormalised transform of x (scratch y of the same length); dir -1:
 * exp(-i ..) (forward), +1: exp(+i ..) (backward).  Returns the buffer with
 * the result. */
static cplx *fft_exec(const fftplan *P, cplx *x, cplx *y, int dir) {
  const int n = P->n;
  int Ns = 1;
  for (int s = 0; s < P->nst; ++s) {
    const int R = P->r[s], nb = n / R, ts = n / (Ns * R), ng = nb / Ns;
    /* butterfly j = g Ns + k reads x[j + m nb], writes y[g Ns R + k + m Ns] */
    for (int k = 0; k < Ns; ++k) {
      cplx tw[5];
      for (int m = 1; m < R; ++m) {
        tw[m] = P->tw[m * k * ts];
        if (dir > 0) tw[m] = conj(tw[m]);
      }
      for (int g = 0; g < ng; ++g) {
        const int j = g * Ns + k, od = g * Ns * R + k;
        cplx v[5];
        v[0] = x[j];
        for (int m = 1; m < R; ++m) v[m] = (k ? x[j + m * nb] * tw[m] : x[j + m * nb]);
        if (R == 2) {
          const cplx a = v[0] + v[1], b = v[0] - v[1];
          v[0] = a; v[1] = b;
        } else if (R == 4) {
          const cplx t0 = v[0] + v[2], t1 = v[0] - v[2], t2 = v[1] + v[3], t3 = v[1] - v[3];
          const cplx it3 = (dir < 0) ? -I * t3 : I * t3;
          v[0] = t0 + t2; v[2] = t0 - t2; v[1] = t1 + it3; v[3] = t1 - it3;
        } else if (R == 3) {
          const double s3 = 0.86602540378443864676 * (dir < 0 ? -1.0 : 1.0);
          const cplx t1 = v[1] + v[2], t2 = v[0] - 0.5 * t1, t3 = I * s3 * (v[1] - v[2]);
          v[0] = v[0] + t1; v[1] = t2 + t3; v[2] = t2 - t3;
        } else {  /* 5: from the definition */
          cplx o[5];
          for (int q = 0; q < 5; ++q) {
            cplx acc = 0.0;
            for (int m = 0; m < 5; ++m) {
              cplx w = P->tw[((q * m) % 5) * (n / 5)];
              if (dir > 0) w = conj(w);
              acc += v[m] * w;
            }
            o[q] = acc;
          }
          for (int q = 0; q < 5; ++q) v[q] = o[q];
        }
        for (int m = 0; m < R; ++m) y[od + m * Ns] = v[m];
      }
    }
    cplx *t = x; x = y; y = t;
    Ns *= R;
  }
  return x;
}

/* ------------------------------------------------------------ context -- */
typedef struct {
  const kc_params *p;
  int nxh, nz, nzh, nyp, ny, nyh, mx, mz, P;
  double alpha, beta, dyde, g, gs;
  double *wavx, *wavz, *ycoord, *seght, *smz, *smy;
  fftplan fx, fz, fc;   /* mx, mz, 2 ny */
  cplx *T[9];           /* [y][kx][kz] physical-y transfer fields */
  cplx *NL[4];          /* [y][kx][kz] plane-stage outputs */
  cplx *u, *w, *omx, *omz, *sx, *sy, *sz;  /* spectral work fields */
  int kmax;
  double fs[10][10];
} kc_ctx;

#define SIDX(c, kx, kz) (((size_t)(kx) * (c)->nz + (kz)) * (c)->nyp)

/* ---------------------------------------------------- pencil routines -- */
/* derivatives.f90:258-314 */
static void cderiv(const cplx *f, cplx *df, int ny, double dyde) {
  df[ny] = 0.0;
  df[ny - 1] = 2.0 * ny * f[ny] * dyde;
  for (int n = ny - 2; n >= 0; --n) df[n] = (n + 2 <= ny ? df[n + 2] : 0.0) + 2.0 * (n + 1) * f[n + 1] * dyde;
  df[0] *= 0.5;
}

/* solvers.f90:324-383: 1-based coefficient arrays (size nyh + 2) */
static void evnc(double wn, double dyde, int nyh, double *a1, double *a2, double *a3) {
  for (int i = 0; i < nyh + 2; ++i) a1[i] = a2[i] = a3[i] = 0.0;
  for (int i = 1; i <= nyh; ++i) {
    const double j = 2.0 * i;
    a1[i] = wn * (1.0 / (4 * j * (j - 1)));
    a2[i] = -dyde * dyde - wn * (1.0 / (2 * (j * j - 1)));
    a3[i] = wn * (1.0 / (4 * j * (j + 1)));
  }
  a1[1] *= 2.0;
  a3[nyh - 1] = 0.0;
  a2[nyh] = -dyde * dyde;
  a3[nyh] = 0.0;
}
static void oddc(double wn, double dyde, int nyh, double *a1, double *a2, double *a3) {
  const int nyhm = nyh - 1;
  for (int i = 0; i < nyh + 2; ++i) a1[i] = a2[i] = a3[i] = 0.0;
  for (int i = 1; i <= nyhm; ++i) {
    const double j = 2.0 * i + 1;
    a1[i] = wn * (1.0 / (4 * j * (j - 1)));
    a2[i] = -dyde * dyde - wn * (1.0 / (2 * (j * j - 1)));
    a3[i] = wn * (1.0 / (4 * j * (j + 1)));
  }
  a3[nyhm - 1] = 0.0;
  a2[nyhm] = -dyde * dyde;
  a3[nyhm] = 0.0;
}
/* solvers.f90:387-496 */
static void evns(const cplx *s, int ny, cplx *w) {
  const int nyh = ny / 2, nyhm = nyh - 1;
  for (int i = 0; i < nyh + 2; ++i) w[i] = 0.0;
  for (int i = 2; i <= nyhm; ++i) {
    const int j = 2 * i - 2;
    const double cm = (j - 2 == 0) ? 2.0 : 1.0;
    w[i] = -cm * s[j - 2] * (1.0 / (4.0 * j * (j - 1))) + s[j] * (1.0 / (2.0 * (j * j - 1))) -
           s[j + 2] * (1.0 / (4.0 * j * (j + 1)));
  }
  int j = ny - 2;
  w[nyh] = -s[j - 2] / (4.0 * j * (j - 1)) + s[j] / (2.0 * (j * j - 1));
  j = ny;
  w[nyh + 1] = -s[j - 2] / (4.0 * j * (j - 1));
}
static void odds(const cplx *s, int ny, cplx *w) {
  const int nyh = ny / 2, nyhm = nyh - 1;
  for (int i = 0; i < nyh + 2; ++i) w[i] = 0.0;
  for (int i = 2; i < nyhm; ++i) {
    const int j = 2 * i - 1;
    w[i] = -s[j - 2] * (1.0 / (4.0 * j * (j - 1))) + s[j] * (1.0 / (2.0 * (j * j - 1))) -
           s[j + 2] * (1.0 / (4.0 * j * (j + 1)));
  }
  int j = ny - 3;
  w[nyhm] = -s[j - 2] / (4.0 * j * (j - 1)) + s[j] / (2.0 * (j * j - 1));
  j = ny - 1;
  w[nyh] = -s[j - 2] / (4.0 * j * (j - 1));
}
/* solvers.f90:500-553 */
static void eqsolve(const double *a1, const double *a2, const double *a3, const double *t, cplx *w, int nrank,
                    double *d) {
  const int m = nrank - 1;
  d[m] = a2[m];
  for (int i = m - 1; i >= 1; --i) {
    d[i] = a2[i] - a3[i] * a1[i + 1] / d[i + 1];
    w[i + 1] = w[i + 1] - a3[i] * w[i + 2] / d[i + 1];
  }
  double f = t[m] - t[nrank] * a1[m] / d[m];
  w[1] = w[1] - t[nrank] * w[nrank] / d[m];
  for (int i = m - 1; i >= 1; --i) {
    w[1] = w[1] - f * w[i + 1] / d[i];
    f = t[i] - f * a1[i] / d[i];
  }
  w[1] = w[1] / f;
  for (int i = 2; i <= nrank; ++i) w[i] = (w[i] - a1[i - 1] * w[i - 1]) / d[i - 1];
}

#define KC_MAXH 200  /* nyh + 2 <= KC_MAXH */

/* solvers.f90:174-283 (one pencil) */
static void tau_solve(const cplx *f, double wn, double dyde, cplx bct, cplx bcb, int ib, int ny, cplx *u) {
  const int nyh = ny / 2, nyhp = nyh + 1;
  double a1[KC_MAXH], a2[KC_MAXH], a3[KC_MAXH], d[KC_MAXH], t[KC_MAXH];
  cplx w[KC_MAXH], we[KC_MAXH];
  int kzero = 0;
  if (ib != 0 && wn <= 1.0e-20) { kzero = 1; wn = 1.0; }
  evnc(wn, dyde, nyh, a1, a2, a3);
  evns(f, ny, w);
  w[1] = (ib == 0) ? 0.5 * (bct + bcb) : 0.5 * (bct - bcb);
  t[0] = 0.0;
  for (int q = 1; q <= nyhp; ++q) t[q] = (ib == 0) ? 1.0 : dyde * (double)(2 * (q - 1)) * (2 * (q - 1));
  eqsolve(a1, a2, a3, t, w, nyhp, d);
  for (int q = 0; q < nyh + 2; ++q) we[q] = w[q];
  oddc(wn, dyde, nyh, a1, a2, a3);
  odds(f, ny, w);
  w[1] = (ib == 0) ? 0.5 * (bct - bcb) : 0.5 * (bct + bcb);
  for (int q = 1; q <= nyh; ++q) t[q] = (ib == 0) ? 1.0 : dyde * (double)(2 * (q - 1) + 1) * (2 * (q - 1) + 1);
  eqsolve(a1, a2, a3, t, w, nyh, d);
  for (int q = 0; q <= nyh; ++q) u[2 * q] = we[q + 1];
  for (int q = 0; q < nyh; ++q) u[2 * q + 1] = w[q + 1];
  if (kzero)
    for (int n = 0; n <= ny; ++n) u[n] = 0.0;
}

/* pntsolve pair (solvers.f90:906-1017) with Robin rows tt, tb (nyp each) */
static void pntsolve_pair(double wn, double dyde, const cplx *f, const double *tt, const double *tb, cplx gtop,
                          cplx gbot, int ny, cplx *u) {
  const int nyh = ny / 2;
  double A1[2][KC_MAXH], A2[2][KC_MAXH], A3[2][KC_MAXH], ds[2][KC_MAXH], fj[2][3];
  cplx gsv[2][KC_MAXH];
  cplx s1 = gtop, s2 = gbot;
  evnc(wn, dyde, nyh, A1[0], A2[0], A3[0]);
  oddc(wn, dyde, nyh, A1[1], A2[1], A3[1]);
  evns(f, ny, gsv[0]);
  odds(f, ny, gsv[1]);
  gsv[0][1] = 0.0;
  gsv[1][1] = 0.0;
  for (int ip = 0; ip < 2; ++ip) {
    const int n = ip == 0 ? nyh + 1 : nyh, m = n - 1;
    const double *a1 = A1[ip], *a2 = A2[ip], *a3 = A3[ip];
    double *dd = ds[ip];
    cplx *gs = gsv[ip];
    dd[m] = a2[m];
    for (int i = m - 1; i >= 1; --i) {
      dd[i] = a2[i] - a3[i] * a1[i + 1] / dd[i + 1];
      gs[i + 1] = gs[i + 1] - a3[i] * gs[i + 2] / dd[i + 1];
    }
    for (int js = 2; js >= 1; --js) {
      const double *tv = js == 2 ? tb : tt;
      double t[KC_MAXH];
      t[0] = 0.0;
      for (int q = 1; q <= n; ++q) t[q] = ip == 0 ? tv[2 * (q - 1)] : tv[2 * q - 1];
      double fv = t[m] - t[n] * a1[m] / dd[m];
      cplx *sj = js == 1 ? &s1 : &s2;
      *sj = *sj - t[n] * gs[n] / dd[m];
      for (int i = m - 1; i >= 1; --i) {
        *sj = *sj - fv * gs[i + 1] / dd[i];
        fv = t[i] - fv * a1[i] / dd[i];
      }
      fj[ip][js] = fv;
    }
  }
  const double det = fj[0][1] * fj[1][2] - fj[0][2] * fj[1][1];
  u[0] = (s1 * fj[1][2] - s2 * fj[1][1]) / det;
  u[1] = (s2 * fj[0][1] - s1 * fj[0][2]) / det;
  for (int i = 2; i <= nyh + 1; ++i) u[2 * i - 2] = (gsv[0][i] - A1[0][i - 1] * u[2 * i - 4]) / ds[0][i - 1];
  for (int i = 2; i <= nyh; ++i) u[2 * i - 1] = (gsv[1][i] - A1[1][i - 1] * u[2 * i - 3]) / ds[1][i - 1];
}

/* mixedbc / mixedbc1d rows (solvers.f90:777-803, 1249-1269) */
static void mixed_rows(double at, double ab, double bt, double bb, double dyde, int nyp, double bsign, double *tt,
                       double *tb) {
  for (int n = 0; n < nyp; ++n) {
    tt[n] = at + dyde * bt * n * n;
    tb[n] = ab + bsign * dyde * bb * n * n;
    if (n & 1) tb[n] = -tb[n];
  }
}

/* penta (solvers.f90:680-773), ib = 0 */
static void tau_penta(const cplx *f, double wn, double dyde, double at, double bt, cplx gtop, double ab, double bb,
                      cplx gbot, int ny, cplx *u) {
  double tt[2 * KC_MAXH], tb[2 * KC_MAXH];
  mixed_rows(at, ab, bt, bb, dyde, ny + 1, 1.0, tt, tb);
  pntsolve_pair(wn, dyde, f, tt, tb, gtop, gbot, ny, u);
}

/* solvers.f90:555-676 (theta <= 0.99) */
static void phirhs(const cplx *v, const cplx *fn, const cplx *fnm1, double var, double k2, const kc_ctx *c,
                   cplx *out, cplx *w1, cplx *w2) {
  const kc_params *p = c->p;
  const double re = p->re, dt = p->dt, th = p->theta;
  cderiv(v, w1, c->ny, c->dyde);
  cderiv(w1, w2, c->ny, c->dyde);
  for (int n = 0; n < c->nyp; ++n) {
    cplx wr = -re * var / (dt * th) * v[n] - re * var / (2.0 * th) * (3.0 * fn[n] - fnm1[n]);
    wr = wr + (1.0 - th) * k2 * v[n] / th;
    out[n] = wr - (1.0 - th) * w2[n] / th;
  }
}

/* uzero (solvers.f90:1065-1176) through psolve1d's mixedbc1d rows */
static void uzero(double *u0, const double *h1, const double *h1l, int uflag, const kc_ctx *c) {
  const kc_params *p = c->p;
  const double re = p->re, dt = p->dt, th = p->theta;
  const int nyp = c->nyp;
  cplx f[2 * KC_MAXH], d1[2 * KC_MAXH], d2[2 * KC_MAXH], r[2 * KC_MAXH], u[2 * KC_MAXH];
  for (int n = 0; n < nyp; ++n) f[n] = u0[n];
  cderiv(f, d1, c->ny, c->dyde);
  cderiv(d1, d2, c->ny, c->dyde);
  for (int n = 0; n < nyp; ++n)
    r[n] = -re * u0[n] / (dt * th) - re * (3.0 * h1[n] - h1l[n]) / (2.0 * th) - (1.0 - th) * d2[n] / th;
  double tt[2 * KC_MAXH], tb[2 * KC_MAXH];
  if (uflag == 1) {
    mixed_rows(p->atu, p->abu, p->btu, p->bbu, c->dyde, nyp, -1.0, tt, tb);
    pntsolve_pair(re / (th * dt), c->dyde, r, tt, tb, p->gtu, p->gbu, c->ny, u);
  } else {  /* btu for w, as the reference passes it (solvers.f90:1173) */
    mixed_rows(p->atw, p->abw, p->btu, p->bbw, c->dyde, nyp, -1.0, tt, tb);
    pntsolve_pair(re / (th * dt), c->dyde, r, tt, tb, p->gtw, p->gbw, c->ny, u);
  }
  for (int n = 0; n < nyp; ++n) u0[n] = creal(u[n]);
}

/* helpers.f90:91-158 on the kx = 0 plane and the z-Nyquist row */
static void norm_field(cplx *a, const kc_ctx *c) {
  const int nz = c->nz, nyp = c->nyp;
  for (int n = 0; n < nyp; ++n) a[SIDX(c, 0, 0) + n] = creal(a[SIDX(c, 0, 0) + n]);
  for (int j = 1; j < nz / 2; ++j) {
    cplx *x = a + SIDX(c, 0, j), *y = a + SIDX(c, 0, nz - j);
    for (int n = 0; n < nyp; ++n) {
      x[n] = 0.5 * (x[n] + conj(y[n]));
      y[n] = conj(x[n]);
    }
  }
  for (int kx = 0; kx < c->nxh; ++kx) memset(a + SIDX(c, kx, nz / 2), 0, sizeof(cplx) * nyp);
}

/* Chebyshev values at y_k of sum a_n T_n (c_n/2 scaling, DCT-I) */
static void cheb_inverse(const cplx *a, cplx *out, const kc_ctx *c, cplx *b1, cplx *b2) {
  const int ny = c->ny, M = 2 * ny;
  for (int n = 0; n <= ny; ++n) b1[n] = (n == 0 || n == ny) ? a[n] : 0.5 * a[n];
  for (int n = 1; n < ny; ++n) b1[M - n] = b1[n];
  const cplx *r = fft_exec(&c->fc, b1, b2, -1);
  for (int n = 0; n <= ny; ++n) out[n] = r[n];
}
/* DCT-I / ny, ends halved */
static void cheb_forward(const cplx *v, cplx *out, const kc_ctx *c, cplx *b1, cplx *b2) {
  const int ny = c->ny, M = 2 * ny;
  for (int n = 0; n <= ny; ++n) b1[n] = v[n];
  for (int n = 1; n < ny; ++n) b1[M - n] = b1[n];
  const cplx *r = fft_exec(&c->fc, b1, b2, -1);
  for (int n = 0; n <= ny; ++n) out[n] = r[n] / (double)ny;
  out[0] *= 0.5;
  out[ny] *= 0.5;
}

/* --------------------------------------------------------------- step -- */
static double k2_of(const kc_ctx *c, int kx, int kz) { return c->wavx[kx] * c->wavx[kx] + c->wavz[kz] * c->wavz[kz]; }

/* solves of the step for one pencil (main_dns.f90:324-484) */
static void pencil_solves(kc_ctx *c, const kc_state *s, int kx, int kz, cplx *wk) {
  const kc_params *p = c->p;
  const int nyp = c->nyp, ny = c->ny;
  const size_t o = SIDX(c, kx, kz);
  const double k2 = k2_of(c, kx, kz);
  cplx *a = wk, *b = wk + nyp, *r = wk + 2 * nyp, *w1 = wk + 3 * nyp, *w2 = wk + 4 * nyp, *vp = wk + 5 * nyp;
  if (p->scalar) {
    phirhs(s->sc + o, s->scn + o, s->scnm1 + o, p->diff, k2, c, r, w1, w2);
    memcpy(s->scnm1 + o, s->scn + o, sizeof(cplx) * nyp);
    tau_penta(r, k2 + c->gs, c->dyde, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0, ny, s->sc + o);
  }
  /* phi = D^2 v - k^2 v */
  cderiv(s->v + o, a, ny, c->dyde);
  cderiv(a, b, ny, c->dyde);
  for (int n = 0; n < nyp; ++n) b[n] -= k2 * s->v[o + n];
  phirhs(b, s->fn + o, s->fnm1 + o, 1.0, k2, c, r, w1, w2);
  memcpy(s->fnm1 + o, s->fn + o, sizeof(cplx) * nyp);
  tau_solve(r, k2 + c->g, c->dyde, 0.0, 0.0, 0, ny, a);
  tau_solve(a, k2, c->dyde, 0.0, 0.0, 0, ny, vp);
  phirhs(s->omy + o, s->gn + o, s->gnm1 + o, 1.0, k2, c, r, w1, w2);
  memcpy(s->gnm1 + o, s->gn + o, sizeof(cplx) * nyp);
  cplx vb = 0.0, vt = 0.0;
  for (int n = 0; n < nyp; ++n) {
    const double nn = (double)n * n * c->dyde;
    vt += vp[n] * nn;
    vb += vp[n] * ((n & 1) ? nn : -nn);
  }
  const size_t q = (size_t)kx * c->nz + kz;
  const cplx dnm = s->dnv1t[q] * s->dnv2b[q] - s->dnv2t[q] * s->dnv1b[q];
  const cplx c1 = (vb * s->dnv2t[q] - vt * s->dnv2b[q]) / dnm;
  const cplx c2 = (vt * s->dnv1b[q] - vb * s->dnv1t[q]) / dnm;
  for (int n = 0; n < nyp; ++n) s->v[o + n] = vp[n] + c1 * s->gf1[o + n] + c2 * s->gf2[o + n];
  tau_penta(r, k2 + c->g, c->dyde, p->atu, p->btu, 0.0, p->abu, p->bbu, 0.0, ny, s->omy + o);
}

/* veloc (solvers.f90:1576-1659) for one pencil */
static void pencil_veloc(kc_ctx *c, kc_state *s, int kx, int kz, cplx *wk) {
  const int nyp = c->nyp;
  const size_t o = SIDX(c, kx, kz);
  if (kx == 0 && kz == 0) {
    for (int n = 0; n < nyp; ++n) {
      c->u[o + n] = s->u0[n];
      c->w[o + n] = s->w0[n];
      s->v[o + n] = 0.0;
    }
    return;
  }
  cderiv(s->v + o, wk, c->ny, c->dyde);
  const double kxw = c->wavx[kx], kzw = c->wavz[kz], w2 = kxw * kxw + kzw * kzw;
  for (int n = 0; n < nyp; ++n) {
    c->u[o + n] = -(-I * kxw * wk[n] + I * kzw * s->omy[o + n]) / w2;
    c->w[o + n] = -(-I * kzw * wk[n] - I * kxw * s->omy[o + n]) / w2;
  }
}
/* vort (solvers.f90:1663-1722) */
static void pencil_vort(kc_ctx *c, kc_state *s, int kx, int kz, cplx *wk) {
  const int nyp = c->nyp;
  const size_t o = SIDX(c, kx, kz);
  const double kxw = c->wavx[kx], kzw = c->wavz[kz];
  cderiv(c->u + o, wk, c->ny, c->dyde);
  cderiv(c->w + o, wk + nyp, c->ny, c->dyde);
  for (int n = 0; n < nyp; ++n) {
    c->omz[o + n] = I * kxw * s->v[o + n] - wk[n];
    c->omx[o + n] = wk[nyp + n] - I * kzw * s->v[o + n];
  }
}
/* gradscl (derivatives.f90:90-144) */
static void pencil_gradscl(kc_ctx *c, kc_state *s, int kx, int kz) {
  const int nyp = c->nyp;
  const size_t o = SIDX(c, kx, kz);
  cderiv(s->sc + o, c->sy + o, c->ny, c->dyde);
  for (int n = 0; n < nyp; ++n) {
    c->sx[o + n] = I * c->wavx[kx] * s->sc[o + n];
    c->sz[o + n] = I * c->wavz[kz] * s->sc[o + n];
  }
}
/* inverse Chebyshev transforms into the [y][kx][kz] transfer fields */
static void pencil_to_transfer(kc_ctx *c, kc_state *s, int kx, int kz, cplx *wk) {
  const int nyp = c->nyp, nf = c->p->scalar ? 9 : 6;
  const size_t o = SIDX(c, kx, kz), P = (size_t)c->P, pp = (size_t)kx * c->nz + kz;
  const cplx *src[9] = {c->u, s->v, c->w, c->omx, s->omy, c->omz, c->sx, c->sy, c->sz};
  cplx *out = wk, *b1 = wk + nyp, *b2 = wk + nyp + 2 * c->ny;
  for (int f = 0; f < nf; ++f) {
    cheb_inverse(src[f] + o, out, c, b1, b2);
    for (int n = 0; n < nyp; ++n) c->T[f][(size_t)n * P + pp] = out[n];
  }
}

/* the plane stage for y-plane y: z inverse, x c2r, products, CFL, dPdx and
 * IB, x r2c, z forward (solvers.f90:2550-3275); returns the plane's CFL max */
static double plane(kc_ctx *c, kc_state *s, int y, int it, cplx *zb, double *ph, cplx *l1, cplx *l2) {
  const kc_params *p = c->p;
  const int nxh = c->nxh, nz = c->nz, nzh = c->nzh, mx = c->mx, mz = c->mz;
  const int nf = p->scalar ? 9 : 6, nout = p->scalar ? 4 : 3;
  const size_t P = (size_t)c->P, pl = (size_t)mx * mz;
  /* ph[f][x][z] physical */
  for (int f = 0; f < nf; ++f) {
    const cplx *t = c->T[f] + (size_t)y * P;
    /* z inverse per kx: zb[kx][z] */
    for (int kx = 0; kx < nxh; ++kx) {
      for (int z = 0; z < mz; ++z) l1[z] = 0.0;
      for (int j = 0; j < nz; ++j) l1[j < nzh ? j : mz - nz + j] = t[(size_t)kx * nz + j];
      const cplx *r = fft_exec(&c->fz, l1, l2, +1);
      for (int z = 0; z < mz; ++z) zb[(size_t)kx * mz + z] = r[z];
    }
    /* x c2r, rows z and z + 1 packed */
    double *out = ph + (size_t)f * pl;
    for (int z = 0; z < mz; z += 2) {
      for (int x = 0; x < mx; ++x) l1[x] = 0.0;
      const cplx a0 = zb[z], b0 = zb[z + 1];
      l1[0] = creal(a0) + I * creal(b0);
      for (int k = 1; k < nxh; ++k) {
        const cplx a = zb[(size_t)k * mz + z], b = zb[(size_t)k * mz + z + 1];
        l1[k] = a + I * b;
        l1[mx - k] = conj(a) + I * conj(b);
      }
      const cplx *r = fft_exec(&c->fx, l1, l2, +1);
      for (int x = 0; x < mx; ++x) {
        out[(size_t)x * mz + z] = creal(r[x]);
        out[(size_t)x * mz + z + 1] = cimag(r[x]);
      }
    }
  }
  /* products (solvers.f90:2891-2898), CFL (2985-2986), scalar (3059-3063) */
  double *U = ph, *V = ph + pl, *W = ph + 2 * pl, *OX = ph + 3 * pl, *OY = ph + 4 * pl, *OZ = ph + 5 * pl;
  double *vw = ph + 9 * pl;  /* vwx, vwy, vwz, vc */
  const double delxm = p->xl / mx, delzm = p->zl / mz;
  double cmax = 0.0;
  for (size_t e = 0; e < pl; ++e) {
    const double u = U[e], v = V[e], w = W[e];
    vw[e] = v * OZ[e] - w * OY[e];
    vw[pl + e] = -u * OZ[e] + w * OX[e];
    vw[2 * pl + e] = u * OY[e] - v * OX[e];
    const double cf = (fabs(u) / delxm + fabs(v) / c->seght[y] + fabs(w) / delzm) * p->dt;
    if (cf > cmax) cmax = cf;
    if (p->scalar) {
      double vc = -(u * ph[6 * pl + e] + v * ph[7 * pl + e] + w * ph[8 * pl + e]);
      if (p->scl_flag == 6) vc += p->deltaT;
      vw[3 * pl + e] = vc;
    }
  }
  const int *im = s->imat;
  if (p->flow_select == 1)
    for (int x = 0; x < mx; ++x)
      for (int z = 0; z < mz; ++z)
        if (!im || im[((size_t)x * mz + z) * c->nyp + y] == 0) vw[(size_t)x * mz + z] -= p->dPdx;
  if (im) {
    /* Goldstein feedback (solvers.f90:2996-3003) and the order-dependent
     * spreading (3064-3108) for cell types 1 and 4 */
    double *drag = ph + 13 * pl;  /* dragx, dragy, dragz */
    int any = 0;
    for (int x = 0; x < mx; ++x)
      for (int z = 0; z < mz; ++z) {
        const size_t e = (size_t)x * mz + z, g3 = e * c->nyp + y;
        drag[e] = drag[pl + e] = drag[2 * pl + e] = 0.0;
        const int t = im[g3];
        if (t == 1 || (t == 4 && it <= p->perturbtime)) {
          s->fxi[g3] += U[e];
          s->fyi[g3] += V[e];
          s->fzi[g3] += W[e];
          drag[e] = -p->ugain * U[e] - p->gain * s->fxi[g3];
          drag[pl + e] = -p->ugain * V[e] - p->gain * s->fyi[g3];
          drag[2 * pl + e] = -p->ugain * W[e] - p->gain * s->fzi[g3];
          any = 1;
        } else if (t == 4) {
          any = 1;
        }
      }
    if (any && (y + 1) <= c->kmax) {
      for (int x = 0; x < mx; ++x)
        for (int z = 0; z < mz; ++z) {
          const size_t e = (size_t)x * mz + z;
          const int t = im[e * c->nyp + y];
          if (!(t == 1 || t == 4)) continue;
          const long ls = (long)x * mz + z;
          for (int ii = -2; ii <= 2; ++ii) {
            int ip = x + ii;
            if (ip < 0) ip += mx;
            if (ip >= mx) ip -= mx;
            for (int jj = -2; jj <= 2; ++jj) {
              const double wgt = c->fs[abs(jj)][abs(ii)];
              const long lt = (long)ip * mz + (z + jj);
              if (lt < 0 || lt >= (long)pl) return -1.0;  /* spreading leaves the plane */
              if (lt <= ls) {
                vw[lt] += wgt * drag[e];
                vw[pl + lt] += wgt * drag[pl + e];
                vw[2 * pl + lt] += wgt * drag[2 * pl + e];
              }
            }
          }
        }
    }
  }
  /* forward: x r2c (rows packed), z forward, keep nz modes, 1/(mx mz) */
  const double sc = 1.0 / ((double)mx * (double)mz);
  for (int o = 0; o < nout; ++o) {
    const double *in = vw + (size_t)o * pl;
    for (int z = 0; z < mz; z += 2) {
      for (int x = 0; x < mx; ++x) l1[x] = in[(size_t)x * mz + z] + I * in[(size_t)x * mz + z + 1];
      const cplx *r = fft_exec(&c->fx, l1, l2, -1);
      for (int k = 0; k < nxh; ++k) {
        const cplx zk = r[k], zm = conj(r[(mx - k) % mx]);
        zb[(size_t)k * mz + z] = 0.5 * (zk + zm);
        zb[(size_t)k * mz + z + 1] = -0.5 * I * (zk - zm);
      }
    }
    cplx *dst = c->NL[o] + (size_t)y * P;
    for (int kx = 0; kx < nxh; ++kx) {
      for (int z = 0; z < mz; ++z) l1[z] = zb[(size_t)kx * mz + z];
      const cplx *r = fft_exec(&c->fz, l1, l2, -1);
      for (int j = 0; j < nz; ++j) dst[(size_t)kx * nz + j] = r[j < nzh ? j : mz - nz + j] * sc;
    }
  }
  return cmax;
}

static int kc_init(kc_ctx *c, const kc_params *p, const kc_state *s) {
  memset(c, 0, sizeof(*c));
  c->p = p;
  c->nxh = p->nx / 2; c->nz = p->nz; c->nzh = p->nz / 2; c->ny = p->ny; c->nyp = p->ny + 1; c->nyh = p->ny / 2;
  c->mx = 3 * p->nx / 2; c->mz = 3 * p->nz / 2; c->P = c->nxh * c->nz;
  if (c->nyh + 2 > KC_MAXH || p->theta > 0.99) return 2;
  const double pi = 2.0 * acos(0.0);
  c->alpha = 2.0 * pi / p->xl; c->beta = 2.0 * pi / p->zl; c->dyde = 2.0 / p->yl;
  c->g = p->re / (p->dt * p->theta);
  c->gs = p->re * p->diff / (p->dt * p->theta);
  c->wavx = malloc(sizeof(double) * c->nxh);
  c->wavz = malloc(sizeof(double) * c->nz);
  for (int k = 0; k < c->nxh; ++k) c->wavx[k] = k * c->alpha;
  for (int k = 0; k < c->nz; ++k) c->wavz[k] = (k < c->nzh ? k : k - c->nz) * c->beta;
  c->ycoord = malloc(sizeof(double) * c->nyp);
  c->seght = malloc(sizeof(double) * c->nyp);
  for (int n = 0; n < c->nyp; ++n) c->ycoord[n] = cos(n * pi / c->ny) * (p->yl / 2.0);
  c->ycoord[0] = p->yl / 2.0;
  for (int n = 1; n < c->nyp; ++n) c->seght[n] = fabs(c->ycoord[n] - c->ycoord[n - 1]);
  c->seght[0] = c->seght[c->ny];
  c->smz = malloc(sizeof(double) * c->nz);
  c->smy = malloc(sizeof(double) * c->nyp);
  for (int j = 0; j < c->nz; ++j) {
    const int jj = j < c->nzh ? j : c->nz - j;
    c->smz[j] = exp(-3.0 * pow((double)jj / c->nzh, 10));
  }
  for (int n = 0; n < c->nyp; ++n) c->smy[n] = exp(-3.0 * pow((double)n / c->ny, 10));
  plan_init(&c->fx, c->mx);
  plan_init(&c->fz, c->mz);
  plan_init(&c->fc, 2 * c->ny);
  const size_t S = (size_t)c->P * c->nyp;
  for (int f = 0; f < 9; ++f) c->T[f] = calloc(S, sizeof(cplx));
  for (int f = 0; f < 4; ++f) c->NL[f] = calloc(S, sizeof(cplx));
  c->u = calloc(S, sizeof(cplx)); c->w = calloc(S, sizeof(cplx));
  c->omx = calloc(S, sizeof(cplx)); c->omz = calloc(S, sizeof(cplx));
  c->sx = calloc(S, sizeof(cplx)); c->sy = calloc(S, sizeof(cplx)); c->sz = calloc(S, sizeof(cplx));
  /* main_dns.f90:1526-1558: kmaxsurf and fspread */
  c->kmax = 0;
  if (s->imat)
    for (int x = 0; x < c->mx; ++x)
      for (size_t e = 0; e < (size_t)c->mz * c->nyp; ++e) {
        const int t = s->imat[(size_t)x * c->mz * c->nyp + e];
        if (t == 1 || t == 3 || t == 4) { if (x + 1 > c->kmax) c->kmax = x + 1; }
        else if (t != 0) return 2;   /* cell types beyond the solid ones are not restated here */
      }
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 10; ++j) {
      const double dsq = (double)(i * i + j * j);
      c->fs[j][i] = dsq < 13 ? exp(-2.0 * dsq) : 0.0;
    }
  return 0;
}

static void kc_free(kc_ctx *c) {
  free(c->wavx); free(c->wavz); free(c->ycoord); free(c->seght); free(c->smz); free(c->smy);
  plan_free(&c->fx); plan_free(&c->fz); plan_free(&c->fc);
  for (int f = 0; f < 9; ++f) free(c->T[f]);
  for (int f = 0; f < 4; ++f) free(c->NL[f]);
  free(c->u); free(c->w); free(c->omx); free(c->omz); free(c->sx); free(c->sy); free(c->sz);
}

/* one time step (Solver.step of the oracle, main_dns.f90:198-811) */
static int kc_step(kc_ctx *c, kc_state *s, int it, double *cflmax) {
  const kc_params *p = c->p;
  const int nxh = c->nxh, nz = c->nz, nyp = c->nyp, P = c->P;
  const int wkn = 16 * nyp + 8 * c->ny;
  /* solves (every pencil) */
#pragma omp parallel
  {
    cplx *wk = malloc(sizeof(cplx) * wkn);
#pragma omp for schedule(dynamic, 16)
    for (int q = 0; q < P; ++q) pencil_solves(c, s, q / nz, q % nz, wk);
    free(wk);
  }
  uzero(s->u0, s->h1n, s->h1nm1, 1, c);
  uzero(s->w0, s->h3n, s->h3nm1, 3, c);
  memset(s->omy + SIDX(c, 0, 0), 0, sizeof(cplx) * nyp);
  memcpy(s->h1nm1, s->h1n, sizeof(double) * nyp);
  memcpy(s->h3nm1, s->h3n, sizeof(double) * nyp);
  norm_field(s->v, c);
  norm_field(s->omy, c);
  /* veloc, vort, gradscl and the inverse transforms: the kx > 0 pencils in
   * one pass, the kx = 0 plane op by op with its norm projections between */
#pragma omp parallel
  {
    cplx *wk = malloc(sizeof(cplx) * wkn);
#pragma omp for schedule(dynamic, 16)
    for (int q = nz; q < P; ++q) {
      const int kx = q / nz, kz = q % nz;
      if (kz == c->nzh) {  /* norm zeroes the Nyquist row of every derived field */
        pencil_veloc(c, s, kx, kz, wk);
        const size_t o = SIDX(c, kx, kz);
        memset(c->u + o, 0, sizeof(cplx) * nyp); memset(c->w + o, 0, sizeof(cplx) * nyp);
        memset(c->omx + o, 0, sizeof(cplx) * nyp); memset(c->omz + o, 0, sizeof(cplx) * nyp);
        if (p->scalar) {
          memset(s->sc + o, 0, sizeof(cplx) * nyp);
          memset(c->sx + o, 0, sizeof(cplx) * nyp); memset(c->sy + o, 0, sizeof(cplx) * nyp);
          memset(c->sz + o, 0, sizeof(cplx) * nyp);
        }
      } else {
        pencil_veloc(c, s, kx, kz, wk);
        pencil_vort(c, s, kx, kz, wk);
        if (p->scalar) pencil_gradscl(c, s, kx, kz);
      }
      pencil_to_transfer(c, s, kx, kz, wk);
    }
    free(wk);
  }
  {
    cplx *wk = malloc(sizeof(cplx) * wkn);
    for (int kz = 0; kz < nz; ++kz) pencil_veloc(c, s, 0, kz, wk);
    norm_field(c->u, c);
    norm_field(c->w, c);
    for (int kz = 0; kz < nz; ++kz) pencil_vort(c, s, 0, kz, wk);
    norm_field(c->omx, c);
    norm_field(c->omz, c);
    if (p->scalar) {
      norm_field(s->sc, c);
      for (int kz = 0; kz < nz; ++kz) pencil_gradscl(c, s, 0, kz);
      norm_field(c->sx, c);
      norm_field(c->sy, c);
      norm_field(c->sz, c);
    }
    for (int kz = 0; kz < nz; ++kz) pencil_to_transfer(c, s, 0, kz, wk);
    free(wk);
  }
  /* plane stage */
  double cm = 0.0;
  int bad = 0;
#pragma omp parallel
  {
    const size_t pl = (size_t)c->mx * c->mz;
    cplx *zb = malloc(sizeof(cplx) * (size_t)nxh * c->mz);
    double *ph = malloc(sizeof(double) * 16 * pl);
    const int L = c->mx > c->mz ? c->mx : c->mz;
    cplx *l1 = malloc(sizeof(cplx) * L), *l2 = malloc(sizeof(cplx) * L);
    double lm = 0.0;
#pragma omp for schedule(dynamic, 1)
    for (int y = 0; y < nyp; ++y) {
      const double r = plane(c, s, y, it, zb, ph, l1, l2);
      if (r < 0.0) {
#pragma omp atomic write
        bad = 1;
      }
      if (r > lm) lm = r;
    }
#pragma omp critical
    if (lm > cm) cm = lm;
    free(zb); free(ph); free(l1); free(l2);
  }
  if (bad) return 3;
  *cflmax = cm;
  if (cm > 1.0) return 1;
  /* y forward transforms, then norm, smoothing, h1n/h3n, norm, nonlin */
  const int nout = p->scalar ? 4 : 3;
  cplx *outs[4] = {s->gn, s->fn, c->omz, s->scn};   /* vwx -> gn, vwy -> fn, vwz -> omz, vc -> scn */
#pragma omp parallel
  {
    cplx *wk = malloc(sizeof(cplx) * wkn);
    cplx *col = wk, *b1 = wk + nyp, *b2 = wk + nyp + 2 * c->ny;
#pragma omp for schedule(dynamic, 16)
    for (int q = 0; q < P; ++q) {
      const size_t o = (size_t)q * nyp;
      for (int f = 0; f < nout; ++f) {
        for (int n = 0; n < nyp; ++n) col[n] = c->NL[f][(size_t)n * P + q];
        cheb_forward(col, outs[f] + o, c, b1, b2);
      }
    }
    free(wk);
  }
  for (int f = 0; f < nout; ++f) norm_field(outs[f], c);
  for (int q = 0; q < P; ++q) {
    const size_t o = (size_t)q * nyp;
    const double sz = c->smz[q % nz];
    for (int f = 0; f < nout; ++f)
      for (int n = 0; n < nyp; ++n) outs[f][o + n] *= sz * c->smy[n];
  }
  for (int n = 0; n < nyp; ++n) {
    s->h1n[n] = creal(s->gn[n]);
    s->h3n[n] = creal(c->omz[n]);
  }
  for (int f = 0; f < 3; ++f) norm_field(outs[f], c);
  /* nonlin (solvers.f90:1497-1572): fn = -k2 fn - D(i (kz omz + kx gn)), gn = i (kz gn - kx omz) */
#pragma omp parallel
  {
    cplx *wk = malloc(sizeof(cplx) * wkn);
#pragma omp for schedule(dynamic, 16)
    for (int q = 0; q < P; ++q) {
      const int kx = q / nz, kz = q % nz;
      const size_t o = (size_t)q * nyp;
      const double kxw = c->wavx[kx], kzw = c->wavz[kz], k2 = kxw * kxw + kzw * kzw;
      for (int n = 0; n < nyp; ++n) wk[n] = I * (kzw * c->omz[o + n] + kxw * s->gn[o + n]);
      cderiv(wk, wk + nyp, c->ny, c->dyde);
      for (int n = 0; n < nyp; ++n) {
        s->fn[o + n] = -k2 * s->fn[o + n] - wk[nyp + n];
        s->gn[o + n] = I * (kzw * s->gn[o + n] - kxw * c->omz[o + n]);
      }
    }
    free(wk);
  }
  norm_field(s->fn, c);
  norm_field(s->gn, c);
  if (p->scalar) norm_field(s->scn, c);
  return 0;
}

/* nsteps steps from step index it0 + 1; 0 ok, 1 CFL > 1 (cflmax set, the
 * step's results kept as far as the reference's stop), 2 unsupported
 * configuration, 3 IB spreading leaves the plane */
int kc_steps(const kc_params *p, kc_state *s, int it0, int nsteps, double *cflmax) {
  if (p->threads > 0) omp_set_num_threads(p->threads);
  kc_ctx c;
  int r = kc_init(&c, p, s);
  if (r) { kc_free(&c); return r; }
  double cm = 0.0;
  for (int k = 0; k < nsteps && !r; ++k) r = kc_step(&c, s, it0 + 1 + k, &cm);
  if (cflmax) *cflmax = cm;
  kc_free(&c);
  return r;
}

int kc_max_threads(void) { return omp_get_max_threads(); }
