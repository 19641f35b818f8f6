/*
 * gems_oracle.c -- CPU FP64 restatement of the GEMS force-evaluation hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; see gems_oracle.h).
 * Compiled with -ffp-contract=off so that every product and sum rounds the
 * way the reference's scalar Fortran does on x86-64 (no FMA contraction).
 *
 * Parity status of each part (how it is pinned) is listed per function.
 */
#include "gems_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* constants: src/Constants.F90:114,124-127,173-175                     */
/* ------------------------------------------------------------------ */
static double c_ui_ev(void) {
  const double axps_mxs = 1.0e2, uma_kg = 1.6605402e-27, qe_si = 1.60219e-19;
  const double joule_ev = 1.0 / qe_si;
  return axps_mxs * axps_mxs * uma_kg * joule_ev;
}

void or_constants(double out[5]) {
  double ui_ev = c_ui_ev();
  double ev_ui = 1.0 / ui_ev;
  double kB_ev = 8.617385e-05;
  /* eV_kcm is a default-real literal in the reference (Constants.F90:150) */
  double ev_kcm = (double)23.06035f;
  out[0] = ev_ui;
  out[1] = ui_ev;
  out[2] = kB_ev * ev_ui;
  out[3] = ui_ev * ev_kcm;
  out[4] = kB_ev;
}

double or_element_mass(int z) {
  /* src/Elements.F90: masses are default-real literals stored in real(dp) */
  switch (z) {
  case 2: return (double)4.0032f;
  case 10: return (double)20.183f;
  case 18: return (double)39.948f;
  case 28: return (double)58.69f;
  case 29: return (double)63.55f;
  case 46: return (double)106.4f;
  case 47: return (double)107.87f;
  case 79: return (double)196.97f;
  default: return -1.0;
  }
}

/* ------------------------------------------------------------------ */
/* RNG: src/Random.F90:145-172 (lcg_ranu, lcg_init_ran), 213-250        */
/* ------------------------------------------------------------------ */
void or_rng_init(or_rng *r, int seed) {
  /* lcg_init_ran: idum=-abs(lseed); ix, iy keep their module values but
   * lcg_ranu re-mixes them because idum<=0 on the next call. */
  r->idum = -abs(seed);
  r->ix = -1;
  r->iy = -1;
  r->stored = 0;
  r->g = 0.0;
}

double or_ranu(or_rng *r) {
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  /* am = nearest(1.0,-1.0)/im evaluated in default real (single):
   * im converts to 2^31 in single precision, so am = (1-2^-24)*2^-31. */
  const double am = (double)(nextafterf(1.0f, 0.0f) / (float)im);
  int32_t k;
  if (r->idum <= 0 || r->iy < 0) {
    r->iy = (888889999 ^ abs(r->idum)) | 1;
    r->ix = 777755555 ^ abs(r->idum);
    r->idum = abs(r->idum) + 1;
  }
  uint32_t ux = (uint32_t)r->ix;
  ux ^= ux << 13;
  ux ^= ux >> 17;
  ux ^= ux << 5;
  r->ix = (int32_t)ux;
  k = r->iy / iq;
  r->iy = ia * (r->iy - k * iq) - ir * k;
  if (r->iy < 0) r->iy += im;
  return am * (double)((im & (r->ix ^ r->iy)) | 1);
}

double or_gasdev(or_rng *r) {
  double v1, v2, rsq;
  if (r->stored) {
    r->stored = 0;
    return r->g;
  }
  for (;;) {
    v1 = 2.0 * or_ranu(r) - 1.0;
    v2 = 2.0 * or_ranu(r) - 1.0;
    rsq = v1 * v1 + v2 * v2;
    if (rsq > 0.0 && rsq < 1.0) break;
  }
  rsq = sqrt(-2.0 * log(rsq) / rsq);
  r->g = v2 * rsq;
  r->stored = 1;
  return v1 * rsq;
}

void or_rang(or_rng *r, double *r1, double *r2) {
  double aux;
  *r1 = or_gasdev(r);
  aux = or_gasdev(r);
  if (r2) *r2 = aux;
}

/* ------------------------------------------------------------------ */
/* box + vdistance: Program_Types.F90:126-141, Inq_Properties.F90:137-157 */
/* ------------------------------------------------------------------ */
void or_box_set(or_box *b, int boxed, const double L[3], const int pbc[3]) {
  b->boxed = boxed;
  b->mic = 1;
  for (int d = 0; d < 3; ++d) {
    b->L[d] = boxed ? L[d] : 1.0e6;
    b->one_box[d] = boxed ? 1.0 / L[d] : 1.0e-6;
    b->pbc[d] = pbc ? pbc[d] : 0;
  }
}

static inline double idnint(double v) { return round(v); } /* half away */

void or_vdistance(const or_box *b, const double *xa, const double *xb,
                  double vd[3]) {
  for (int l = 0; l < 3; ++l) vd[l] = xa[l] - xb[l];
  if (!b->mic) return;
  for (int l = 0; l < 3; ++l)
    if (b->pbc[l]) vd[l] = vd[l] - b->L[l] * idnint(vd[l] * b->one_box[l]);
}

static inline double dot3(const double v[3]) {
  return v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
}

/* ------------------------------------------------------------------ */
/* neighbour lists: Neighbour.F90:587-664 (setcells), 676-759           */
/* (verlet_self), 927-989 (sort_in_cells), 991-1087 (linkedcell_self),  */
/* 1350-1376 (cell_pbc), 1413-1432 (icell2vcell), map at :75-83          */
/* ------------------------------------------------------------------ */
static const int nb_map[27][3] = {
    {0, 0, 0},  {1, 0, 0},   {1, 1, 0},  {0, 1, 0},   {-1, 1, 0},
    {1, 0, -1}, {1, 1, -1},  {0, 1, -1}, {-1, 1, -1}, {1, 0, 1},
    {1, 1, 1},  {0, 1, 1},   {-1, 1, 1}, {0, 0, 1},   {-1, 0, 0},
    {-1, -1, 0}, {0, -1, 0}, {1, -1, 0}, {-1, 0, 1},  {-1, -1, 1},
    {0, -1, 1}, {1, -1, 1},  {-1, 0, -1}, {-1, -1, -1}, {0, -1, -1},
    {1, -1, -1}, {0, 0, -1}};

int or_nb_brute(const double *x, int n, const or_box *b, double rlist,
                int strict, int maxnb, int *nn, int *list) {
  double rc2 = rlist * rlist;
  int need = 0;
  for (int i = 0; i < n; ++i) {
    int m = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double vd[3];
      or_vdistance(b, x + 3 * j, x + 3 * i, vd);
      double rd = dot3(vd);
      int keep = strict ? (rd < rc2) : (rd <= rc2);
      if (!keep) continue;
      if (m < maxnb) list[(size_t)i * maxnb + m] = j;
      ++m;
    }
    nn[i] = m < maxnb ? m : maxnb;
    if (m > need) need = m;
  }
  return need > maxnb ? need : 0;
}

int or_nb_build(const double *x, int n, const or_box *b, double rlist,
                int maxnb, int *nn, int *list, int *mode_out,
                int ncells_out[3], int nthreads) {
  double rcut2 = rlist * rlist;
  int nc[3] = {1, 1, 1};
  int linked = 0;
  if (b->boxed) {
    for (int d = 0; d < 3; ++d) nc[d] = (int)(b->L[d] / rlist);
    if (!(nc[0] < 4 && nc[1] < 4 && nc[2] < 4)) linked = 1;
  }
  if (mode_out) *mode_out = linked;
  if (ncells_out)
    for (int d = 0; d < 3; ++d) ncells_out[d] = linked ? nc[d] : 0;
  (void)nthreads;

  if (!linked) {
    /* verlet_self: vd = vdistance(la, lb) = x_i - x_j; rd>rcut -> cycle */
    int need = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(max : need)
    for (int i = 0; i < n; ++i) {
      int m = 0;
      for (int j = 0; j < n; ++j) {
        if (i == j) continue;
        double vd[3];
        or_vdistance(b, x + 3 * i, x + 3 * j, vd);
        double rd = dot3(vd);
        if (rd > rcut2) continue;
        if (m < maxnb) list[(size_t)i * maxnb + m] = j;
        ++m;
      }
      nn[i] = m < maxnb ? m : maxnb;
      if (m > need) need = m;
    }
    return need > maxnb ? need : 0;
  }

  /* sort_in_cells: head/chain, the chain runs from the last inserted atom */
  int ncell = nc[0] * nc[1] * nc[2];
  double cell[3], one_cell[3];
  for (int d = 0; d < 3; ++d) {
    cell[d] = b->L[d] / nc[d];
    one_cell[d] = 1.0 / cell[d];
  }
  int *head = (int *)malloc(sizeof(int) * ncell);
  int *clist = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *cellof = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int bad = 0;
  for (int c = 0; c < ncell; ++c) head[c] = -1;
  for (int i = 0; i < n; ++i) {
    int a[3];
    for (int d = 0; d < 3; ++d) {
      /* Fortran int() truncates toward zero; the reference does no wrap or
       * clamp here (Neighbour.F90:950-957).  Atoms outside the grid are
       * clamped to the edge cell and reported. */
      double t = x[3 * i + d] * one_cell[d];
      int ai = (int)t;
      if (ai < 0 || t < 0.0) { ai = 0; if (t <= -1.0) bad = 1; }
      if (ai >= nc[d]) { if (ai > nc[d]) bad = 1; ai = nc[d] - 1; }
      a[d] = ai;
    }
    int c = a[0] + nc[0] * (a[1] + nc[1] * a[2]);
    cellof[i] = c;
    clist[i] = head[c];
    head[c] = i;
  }
  (void)bad;

  int need = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads) reduction(max : need)
  for (int ic = 0; ic < ncell; ++ic) {
    int rc[3];
    rc[2] = ic / (nc[0] * nc[1]);
    rc[1] = (ic % (nc[0] * nc[1])) / nc[0];
    rc[0] = ic % nc[0];
    for (int i = head[ic]; i >= 0; i = clist[i]) nn[i] = 0;
    for (int nabor = 0; nabor < 27; ++nabor) {
      int ncv[3];
      int ok = 1;
      for (int d = 0; d < 3; ++d) {
        ncv[d] = nb_map[nabor][d] + rc[d];
        if (b->mic) {
          ncv[d] = ((ncv[d] % nc[d]) + nc[d]) % nc[d];
        } else if (ncv[d] < 0 || ncv[d] >= nc[d]) {
          ok = 0;
        }
      }
      if (!ok) continue;
      int jc = ncv[0] + nc[0] * (ncv[1] + nc[1] * ncv[2]);
      for (int i = head[ic]; i >= 0; i = clist[i]) {
        for (int j = head[jc]; j >= 0; j = clist[j]) {
          if (i == j) continue;
          double vd[3];
          or_vdistance(b, x + 3 * j, x + 3 * i, vd);
          double rd = dot3(vd);
          if (rd < rcut2) {
            int m = nn[i];
            if (m < maxnb) list[(size_t)i * maxnb + m] = j;
            nn[i] = m + 1;
          }
        }
      }
    }
    for (int i = head[ic]; i >= 0; i = clist[i]) {
      if (nn[i] > need) need = nn[i];
      if (nn[i] > maxnb) nn[i] = maxnb;
    }
  }
  free(head);
  free(clist);
  free(cellof);
  return need > maxnb ? need : 0;
}

/* ------------------------------------------------------------------ */
/* SMA-TB: TB.F90:106-170 (smatb_set), 367-455 (pass 1), 457-586 (pass 2) */
/* ------------------------------------------------------------------ */
void or_smatb_init(or_smatb *s, int nt) {
  memset(s, 0, sizeof(*s));
  s->nt = nt;
}

void or_smatb_set(or_smatb *s, int i, int j, double a, double eps, double p,
                  double q, double r0, double rci, double rce) {
  int ij = i * OR_MAXT + j, ji = j * OR_MAXT + i;
  double dcut, ab, bb, cb, aux;
  s->prm[ij] = 1;
  s->ca[ij] = a;
  s->eps[ij] = eps;
  s->p[ij] = p;
  s->q[ij] = q;
  s->r0[ij] = r0;
  s->rci[ij] = rci;
  s->rce[ij] = rce;
  /* quintic tail, Logsdail et al. (TB.F90:134-149); dcut**3 etc. are
   * integer powers, evaluated as repeated products */
  dcut = rce - rci;
  ab = -1.0 / (dcut * dcut * dcut);
  bb = -(q / r0) / (dcut * dcut);
  cb = -((q / r0) * (q / r0)) / dcut;
  aux = eps * exp(-q * (rci / r0 - 1.0));
  s->b1[ij] = aux * (12.0 * ab - 6.0 * bb + cb) / (2.0 * (dcut * dcut));
  s->b2[ij] = aux * (15.0 * ab - 7.0 * bb + cb) / dcut;
  s->b3[ij] = aux * (20.0 * ab - 8.0 * bb + cb) / 2.0;
  aux = a * exp(-p * (rci / r0 - 1.0));
  bb = -(p / r0) / (dcut * dcut);
  cb = -((p / r0) * (p / r0)) / dcut;
  s->r1[ij] = aux * (12.0 * ab - 6.0 * bb + cb) / (2.0 * (dcut * dcut));
  s->r2[ij] = aux * (15.0 * ab - 7.0 * bb + cb) / dcut;
  s->r3[ij] = aux * (20.0 * ab - 8.0 * bb + cb) / 2.0;
  s->prm[ji] = 1;
  s->ca[ji] = s->ca[ij];
  s->eps[ji] = s->eps[ij];
  s->p[ji] = s->p[ij];
  s->q[ji] = s->q[ij];
  s->r0[ji] = s->r0[ij];
  s->rci[ji] = s->rci[ij];
  s->rce[ji] = s->rce[ij];
  s->b1[ji] = s->b1[ij];
  s->b2[ji] = s->b2[ij];
  s->b3[ji] = s->b3[ij];
  s->r1[ji] = s->r1[ij];
  s->r2[ji] = s->r2[ij];
  s->r3[ji] = s->r3[ij];
}

double or_smatb_rcut(const or_smatb *s) {
  double m = 0.0;
  for (int i = 0; i < s->nt; ++i)
    for (int j = 0; j < s->nt; ++j)
      if (s->rce[i * OR_MAXT + j] > m) m = s->rce[i * OR_MAXT + j];
  return m;
}

void or_smatb_force(const double *x, const int *type, int n, const or_box *b,
                    const or_smatb *s, const int *nn, const int *list,
                    int maxnb, double *f, double *e, double *rho_out,
                    double *eband_out, double *epot, int nthreads) {
  double ev_ui = 1.0 / c_ui_ev();
  double *band = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  double *eband = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  (void)nthreads;

  /* pass 1: smatb_preinteraction (TB.F90:390-452) */
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int i = 0; i < n; ++i) {
    int z1 = type[i];
    double acc = 0.0;
    for (int l = 0; l < nn[i]; ++l) {
      int j = list[(size_t)i * maxnb + l];
      int zz = z1 * OR_MAXT + type[j];
      double vd[3];
      or_vdistance(b, x + 3 * j, x + 3 * i, vd);
      double dr = sqrt(dot3(vd));
      if (dr > s->rce[zz]) continue;
      double banda;
      if (dr < s->rci[zz]) {
        double r0 = s->r0[zz];
        banda = s->eps[zz] * s->eps[zz] * exp(-2.0 * s->q[zz] * (dr / r0 - 1.0));
      } else {
        double drm = dr - s->rce[zz];
        double drm2 = drm * drm, drm3 = drm2 * drm, drm4 = drm3 * drm,
               drm5 = drm4 * drm;
        banda = s->b1[zz] * drm5 + s->b2[zz] * drm4 + s->b3[zz] * drm3;
        banda = banda * banda;
      }
      acc = acc + banda;
    }
    band[i] = acc;
  }
  for (int i = 0; i < n; ++i) eband[i] = 1.0 / sqrt(band[i]);

  /* pass 2: smatb_interact (TB.F90:479-580), newton off */
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int i = 0; i < n; ++i) {
    int z1 = type[i];
    double fi[3] = {0.0, 0.0, 0.0}, ei = 0.0;
    for (int l = 0; l < nn[i]; ++l) {
      int j = list[(size_t)i * maxnb + l];
      int zz = z1 * OR_MAXT + type[j];
      double vd[3];
      or_vdistance(b, x + 3 * j, x + 3 * i, vd);
      double dr = sqrt(dot3(vd));
      double factor = 0.0, repul;
      if (dr > s->rce[zz]) continue;
      if (dr <= s->rci[zz]) {
        double r0 = s->r0[zz];
        repul = s->ca[zz] * exp(-s->p[zz] * (dr / r0 - 1.0));
        factor = 2.0 * repul * (s->p[zz] / r0) -
                 s->eps[zz] * s->eps[zz] * 1.0 / 2.0 *
                     exp(-2.0 * s->q[zz] * (dr / r0 - 1.0)) *
                     (2.0 * s->q[zz] / r0) * (eband[i] + eband[j]);
      } else {
        double drm = dr - s->rce[zz];
        double drm2 = drm * drm, drm3 = drm2 * drm, drm4 = drm3 * drm,
               drm5 = drm4 * drm;
        factor = (s->b1[zz] * drm5 + s->b2[zz] * drm4 + s->b3[zz] * drm3) *
                 (5.0 * s->b1[zz] * drm4 + 4.0 * s->b2[zz] * drm3 +
                  3.0 * s->b3[zz] * drm2) *
                 (eband[i] + eband[j]);
        factor = factor - 2.0 * (5.0 * s->r1[zz] * drm4 +
                                 4.0 * s->r2[zz] * drm3 +
                                 3.0 * s->r3[zz] * drm2);
        repul = s->r1[zz] * drm5 + s->r2[zz] * drm4 + s->r3[zz] * drm3;
      }
      for (int d = 0; d < 3; ++d) {
        double factor2 = factor * ev_ui * vd[d] / dr;
        fi[d] = fi[d] - factor2;
      }
      ei = ei + repul * ev_ui;
    }
    double aux2 = sqrt(band[i]) * ev_ui;
    ei = ei - aux2;
    for (int d = 0; d < 3; ++d) f[3 * i + d] = fi[d];
    e[i] = ei;
  }
  double acc = 0.0;
  for (int i = 0; i < n; ++i) acc += e[i];
  if (epot) *epot = acc;
  if (rho_out) memcpy(rho_out, band, sizeof(double) * n);
  if (eband_out) memcpy(eband_out, eband, sizeof(double) * n);
  free(band);
  free(eband);
}

/* ------------------------------------------------------------------ */
/* Lennard-Jones: Pairs.F90:198-279, newton off                          */
/* ------------------------------------------------------------------ */
void or_lj_force(const double *x, int n, const or_box *b, double eps,
                 double sigma, double rc, const int *nn, const int *list,
                 int maxnb, double *f, double *e, double *epot,
                 int nthreads) {
  double ev_ui = 1.0 / c_ui_ev();
  double rcut2 = rc * rc;
  (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int i = 0; i < n; ++i) {
    double fi[3] = {0.0, 0.0, 0.0}, ei = 0.0;
    for (int l = 0; l < nn[i]; ++l) {
      int j = list[(size_t)i * maxnb + l];
      double vd[3];
      or_vdistance(b, x + 3 * j, x + 3 * i, vd);
      double dr = dot3(vd);
      if (dr > rcut2) continue;
      dr = sqrt(dr);
      double t = sigma / dr;
      double t2 = t * t, t4 = t2 * t2;
      double s_dr = t4 * t2; /* (s/dr)**6 */
      double p = 4.0 * eps * (s_dr * s_dr - s_dr);
      double fr = eps * (48.0 * s_dr - 24.0) * s_dr / dr;
      p = p * ev_ui;
      fr = fr * ev_ui;
      ei = ei + p * 0.5;
      for (int d = 0; d < 3; ++d) fi[d] = fi[d] - fr * vd[d] / dr;
    }
    for (int d = 0; d < 3; ++d) f[3 * i + d] = fi[d];
    e[i] = ei;
  }
  double acc = 0.0;
  for (int i = 0; i < n; ++i) acc += e[i];
  if (epot) *epot = acc;
}

/* ------------------------------------------------------------------ */
/* shifted Lennard-Jones: Pairs.F90:282-370 (plj_set, plj_interact)      */
/* ------------------------------------------------------------------ */
void or_plj_init(or_plj *p, int nt, double rc) {
  memset(p, 0, sizeof(*p));
  p->nt = nt;
  p->rc = rc;
}

void or_plj_set(or_plj *p, int i, int j, double e, double s) {
  double rc = p->rc;
  double prc = 4.0 * e * (pow(s / rc, 12) - pow(s / rc, 6));
  double frc = 24.0 * e * (2.0 * (pow(s, 12) / pow(rc, 13)) - (pow(s, 6) / pow(rc, 7)));
  int k[2] = {i * OR_MAXT + j, j * OR_MAXT + i};
  for (int u = 0; u < 2; ++u) {
    p->e[k[u]] = e;
    p->s[k[u]] = s;
    p->prc[k[u]] = prc;
    p->frc[k[u]] = frc;
    p->rcp[k[u]] = rc;
  }
}

static double pow6(double x) {  /* x**6 as repeated multiplication */
  double x2 = x * x, x4 = x2 * x2;
  return x4 * x2;
}

void or_slj_set(or_plj *p, int i, int j, double e, double s, double rs) {
  double s_dr = pow6(s / rs);
  double urs = 4.0 * e * (s_dr * s_dr - s_dr);
  double dudrs = -e * (48.0 * s_dr - 24.0) * s_dr / rs;
  double rc = rs - 32.0 * urs / (11.0 * dudrs);
  double d = rc - rs, d2 = d * d, d4 = d2 * d2, d8 = d4 * d4;
  double sa4 = (8.0 * urs + d * dudrs) / (4 * d4);
  double sa8 = -(4.0 * urs + d * dudrs) / (4 * d8);
  int k[2] = {i * OR_MAXT + j, j * OR_MAXT + i};
  for (int u = 0; u < 2; ++u) {
    p->e[k[u]] = e;
    p->s[k[u]] = s;
    p->rs[k[u]] = rs;
    p->sa4[k[u]] = sa4;
    p->sa8[k[u]] = sa8;
    p->rcp[k[u]] = rc;
  }
  if (rc > p->rc) p->rc = rc;
}

void or_wca_set(or_plj *p, int i, int j, double e, double s) {
  /* 2._dp**(1./6.): the exponent is a default-real (FP32) literal */
  double rc = pow(2.0, (double)(1.0f / 6.0f)) * s;
  int k[2] = {i * OR_MAXT + j, j * OR_MAXT + i};
  for (int u = 0; u < 2; ++u) {
    p->e[k[u]] = e;
    p->s[k[u]] = s;
    p->rcp[k[u]] = rc;
  }
  if (rc > p->rc) p->rc = rc;
}

void or_sm1_set(or_plj *p, double r1, double r2, double e) {
  p->kind = 3;
  for (int k = 0; k < OR_MAXT * OR_MAXT; ++k) {
    p->e[k] = e;
    p->rs[k] = r1;
    p->rcp[k] = r2;
  }
  p->rc = r2;
}

void or_plj_force(const double *x, const int *type, int n, const or_box *b, const or_plj *pl,
                  const int *nn, const int *list, int maxnb, double *f, double *e,
                  double *epot, int nthreads) {
  double ev_ui = 1.0 / c_ui_ev();
  double rcut = pl->rc;
  (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int i = 0; i < n; ++i) {
    double fi[3] = {0.0, 0.0, 0.0}, ei = 0.0;
    for (int l = 0; l < nn[i]; ++l) {
      int j = list[(size_t)i * maxnb + l];
      int k = type[i] * OR_MAXT + type[j];
      double vd[3];
      or_vdistance(b, x + 3 * j, x + 3 * i, vd);
      double dr = dot3(vd);
      double rck = pl->rcp[k];
      if (dr > rck * rck) continue;
      dr = sqrt(dr);
      double p, fr;
      if (pl->kind == 3) {
        /* sm1_interact (Pairs.F90:925-990): the reference adds f vd / dr to
         * atom i, so fr = -f here; pi = 4 atan(1), pio2 = pi / 2 */
        const double pi = 4.0 * atan(1.0), pio2 = pi * 0.5;
        const double r1 = pl->rs[k], r2 = rck;
        if (dr > r1) {
          double aux = pi * (dr - (r1 + r2) * 0.5) / (r2 - r1);
          p = 0.5 + 0.5 * sin(aux);
          fr = cos(aux) * pio2 * dr / (r2 - r1);
          p = p * pl->e[k] - pl->e[k];
          fr = -(fr * pl->e[k]);
        } else {
          p = -pl->e[k];
          fr = 0.0;
        }
      } else if (pl->kind == 1 && !(dr < pl->rs[k])) {
        double d = dr - rck, d2 = d * d;
        double s4 = d2 * d2; /* (dr-rcut)**4 */
        p = pl->sa4[k] * s4 + pl->sa8[k] * s4 * s4;
        fr = -(4.0 * pl->sa4[k] * s4 + 8.0 * pl->sa8[k] * s4 * s4) / (dr - rck);
      } else {
        double s_dr = pow6(pl->s[k] / dr); /* (s/dr)**6 */
        if (pl->kind == 0)
          p = 4.0 * pl->e[k] * (s_dr * s_dr - s_dr) - pl->prc[k] - pl->frc[k] * (dr - rcut);
        else if (pl->kind == 1)
          p = 4.0 * pl->e[k] * (s_dr - 1.0) * s_dr;
        else
          p = 4.0 * pl->e[k] * (s_dr * s_dr - s_dr) + pl->e[k];
        fr = pl->e[k] * (48.0 * s_dr - 24.0) * s_dr / dr + (pl->kind == 0 ? pl->frc[k] : 0.0);
      }
      p = p * ev_ui;
      fr = fr * ev_ui;
      ei = ei + p * 0.5;
      for (int d = 0; d < 3; ++d) fi[d] = fi[d] - fr * vd[d] / dr;
    }
    for (int d = 0; d < 3; ++d) f[3 * i + d] = fi[d];
    e[i] = ei;
  }
  double acc = 0.0;
  for (int i = 0; i < n; ++i) acc += e[i];
  if (epot) *epot = acc;
}

/* ------------------------------------------------------------------ */
/* integrators: Integration.F90:909-1031 (Ermak), 1150-1186 (VV)         */
/* ------------------------------------------------------------------ */
void or_vv_a(int n, double dt, double *x, double *v, double *a,
             const double *f, const double *one_mass) {
  double acel_dv = dt * 0.5;
  double acel_dx = acel_dv * dt;
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      int k = 3 * i + d;
      a[k] = f[k] * one_mass[i];
      x[k] = x[k] + dt * v[k] + acel_dx * a[k];
      v[k] = v[k] + acel_dv * a[k];
    }
}

void or_vv_b(int n, double dt, double *v, double *a, const double *f,
             const double *one_mass) {
  double acel_dv = dt * 0.5;
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      int k = 3 * i + d;
      a[k] = f[k] * one_mass[i];
      v[k] = v[k] + acel_dv * a[k];
    }
}

void or_ermak_set(or_ermak *k, double dt, double gama, double T) {
  double ev_ui = 1.0 / c_ui_ev();
  double kB_ui = 8.617385e-05 * ev_ui;
  k->c0 = exp(-dt * gama);
  k->c1 = (1.0 - k->c0) / gama;
  k->c2 = (1.0 - k->c1 / dt) / gama;
  k->sdr = sqrt(dt / gama *
                (2.0 - (3.0 - 4.0 * k->c0 + k->c0 * k->c0) / (dt * gama)));
  k->sdv = sqrt(1.0 - k->c0 * k->c0);
  k->crv1 = (1.0 - k->c0) * (1.0 - k->c0) / (gama * k->sdr * k->sdv);
  k->crv2 = sqrt(1.0 - (k->crv1 * k->crv1));
  k->skt = sqrt(kB_ui * T);
  k->dt = dt;
  k->gamma = gama;
}

void or_ermak_a(const or_ermak *k, or_rng *r, int n, double dt, double *x,
                const double *v, const double *a,
                const double *one_sqrt_mass, double *pos_v, double *vel_v) {
  for (int i = 0; i < n; ++i) {
    double f_lan = k->skt * one_sqrt_mass[i];
    for (int d = 0; d < 3; ++d) {
      double r1, r2;
      or_rang(r, &r1, &r2);
      pos_v[3 * i + d] = f_lan * k->sdr * r1;
      vel_v[3 * i + d] = f_lan * k->sdv * (k->crv1 * r1 + k->crv2 * r2);
    }
    for (int d = 0; d < 3; ++d) {
      int q = 3 * i + d;
      x[q] = x[q] + k->c1 * v[q] + k->c2 * dt * a[q] + pos_v[q];
    }
  }
}

void or_ermak_b(const or_ermak *k, int n, double *v, double *a,
                const double *f, const double *one_mass,
                const double *vel_v) {
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      int q = 3 * i + d;
      v[q] = k->c0 * v[q] + (k->c1 - k->c2) * a[q] +
             k->c2 * f[q] * one_mass[i] + vel_v[q];
      a[q] = f[q] * one_mass[i];
    }
}

void or_do_pbc(int n, const or_box *b, double *x, int *boxcr) {
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      if (!b->pbc[d]) continue;
      int q = 3 * i + d;
      if (x[q] >= b->L[d]) {
        x[q] = x[q] - b->L[d];
        if (boxcr) boxcr[q] += 1;
      } else if (x[q] < 0.0) {
        x[q] = x[q] + b->L[d];
        if (boxcr) boxcr[q] -= 1;
      }
    }
}

/* ------------------------------------------------------------------ */
/* observables: Inq_Properties.F90:399-434, 579-595, 669-697;            */
/* Set_Properties.f90:285-377                                             */
/* ------------------------------------------------------------------ */
static void cm_vel(int n, const double *v, const double *mass, double cm[3]) {
  double M = 0.0;
  for (int i = 0; i < n; ++i) M = M + mass[i];
  if (n == 1) {
    for (int d = 0; d < 3; ++d) cm[d] = v[d];
    return;
  }
  cm[0] = cm[1] = cm[2] = 0.0;
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) cm[d] = cm[d] + mass[i] * v[3 * i + d];
  for (int d = 0; d < 3; ++d) cm[d] = cm[d] / M;
}

double or_ekin(int n, const double *v, const double *mass) {
  double cm[3], ek = 0.0;
  cm_vel(n, v, mass, cm);
  for (int i = 0; i < n; ++i) {
    double w[3];
    for (int d = 0; d < 3; ++d) w[d] = v[3 * i + d] - cm[d];
    ek = ek + dot3(w) * mass[i];
  }
  return ek * 0.5;
}

void or_set_gdist(or_rng *r, int n, double T, double *v, const double *mass,
                  const double *one_mass) {
  double ev_ui = 1.0 / c_ui_ev();
  double kB_ui = 8.617385e-05 * ev_ui;
  double cm[3];
  cm_vel(n, v, mass, cm);
  double factor = kB_ui * T;
  for (int i = 0; i < n; ++i) {
    double vel_med = sqrt(factor * one_mass[i]);
    for (int d = 0; d < 3; ++d) {
      double r1;
      or_rang(r, &r1, NULL);
      v[3 * i + d] = vel_med * r1 + cm[d];
    }
  }
  /* set_scal_vel */
  double temp = 2.0 * or_ekin(n, v, mass) / (kB_ui * (3 * n));
  if (temp != 0.0) {
    double fac = sqrt(T / temp);
    for (int i = 0; i < 3 * n; ++i) v[i] = v[i] * fac;
  }
  /* remove the CM drift */
  cm_vel(n, v, mass, cm);
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) v[3 * i + d] = v[3 * i + d] + (-cm[d]);
}

/* ------------------------------------------------------------------ */
/* MD loop: Programs.f90:114-167, Interaction.f90:135-241,               */
/* Neighbour.F90:1242-1346 (update/test_update)                          */
/* ------------------------------------------------------------------ */
static double md_rcut(const or_md *md) {
  return md->pot == 0 ? or_smatb_rcut(md->tb) : md->pot == 2 ? md->plj->rc : md->lj_rc;
}

static int md_update(or_md *md) {
  md->nupd += 1;
  memcpy(md->pos_old, md->x, sizeof(double) * 3 * md->n);
  int mode, nc[3];
  return or_nb_build(md->x, md->n, &md->box, md_rcut(md) + md->skin,
                     md->maxnb, md->nn, md->list, &mode, nc, md->nthreads);
}

int or_md_interact(or_md *md) {
  int rc = 0;
  if (md->first_call) {
    md->first_call = 0;
    rc = md_update(md);
  } else {
    /* test_update: two largest displacements since the last build;
     * pos_old is not wrapped by do_pbc (Set_Properties.f90:444-470) */
    double d1 = 1.0e-16, d2 = 1.0e-16;
    for (int i = 0; i < md->n; ++i) {
      double vd[3];
      for (int d = 0; d < 3; ++d)
        vd[d] = md->x[3 * i + d] - md->pos_old[3 * i + d];
      double rd = dot3(vd);
      if (rd > d1) {
        d2 = d1;
        d1 = rd;
      } else if (rd > d2) {
        d2 = rd;
      }
    }
    if (sqrt(d1) + sqrt(d2) > md->skin) rc = md_update(md);
  }
  if (rc != 0) return rc;
  if (md->pot == 0)
    or_smatb_force(md->x, md->type, md->n, &md->box, md->tb, md->nn, md->list,
                   md->maxnb, md->f, md->e, NULL, NULL, &md->epot,
                   md->nthreads);
  else if (md->pot == 2)
    or_plj_force(md->x, md->type, md->n, &md->box, md->plj, md->nn, md->list, md->maxnb,
                 md->f, md->e, &md->epot, md->nthreads);
  else
    or_lj_force(md->x, md->n, &md->box, md->lj_eps, md->lj_sigma, md->lj_rc,
                md->nn, md->list, md->maxnb, md->f, md->e, &md->epot,
                md->nthreads);
  return 0;
}

int or_md_run(or_md *md) {
  int n = md->n, rc;
  double cst[5];
  or_constants(cst);
  double ui_ev = cst[1];
  double *one_mass = (double *)malloc(sizeof(double) * n);
  double *one_sqrt_mass = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) {
    one_mass[i] = 1.0 / md->mass[i];
    one_sqrt_mass[i] = sqrt(one_mass[i]);
  }
  or_ermak ek;
  if (md->integ == 1) or_ermak_set(&ek, md->dt, md->gamma, md->T);
  rc = or_md_interact(md);
  int iout = 0;
  for (int ns = 1; ns <= md->nsteps && rc == 0; ++ns) {
    if (md->integ == 0)
      or_vv_a(n, md->dt, md->x, md->v, md->a, md->f, one_mass);
    else
      or_ermak_a(&ek, md->rng, n, md->dt, md->x, md->v, md->a, one_sqrt_mass,
                 md->pos_v, md->vel_v);
    or_do_pbc(n, &md->box, md->x, md->boxcr);
    rc = or_md_interact(md);
    if (rc) break;
    if (md->integ == 0)
      or_vv_b(n, md->dt, md->v, md->a, md->f, one_mass);
    else
      or_ermak_b(&ek, n, md->v, md->a, md->f, one_mass, md->vel_v);
    or_do_pbc(n, &md->box, md->x, md->boxcr);
    if (md->each > 0 && ns % md->each == 0 && md->energy) {
      double ep = 0.0;
      for (int i = 0; i < n; ++i) ep = ep + md->e[i];
      double f1 = ep * ui_ev;
      double f2 = or_ekin(n, md->v, md->mass) * ui_ev;
      md->energy[3 * iout + 0] = f1;
      md->energy[3 * iout + 1] = f2;
      md->energy[3 * iout + 2] = f1 + f2;
      ++iout;
    }
  }
  free(one_mass);
  free(one_sqrt_mass);
  return rc;
}

/* ------------------------------------------------------------------ */
/* FIRE (Bitzek et al. 2006), see gems_oracle.h                          */
/* ------------------------------------------------------------------ */
static double fire_forces(const double *x, const int *type, int n, const or_box *b,
                          const or_smatb *s, double rl, int *maxnb, int **nn, int **list,
                          double *f, double *e) {
  for (;;) {
    int need = or_nb_build(x, n, b, rl, *maxnb, *nn, *list, NULL, NULL, 1);
    if (need == 0) break;
    *maxnb = need + 16;
    free(*list);
    *list = (int *)malloc(sizeof(int) * (size_t)n * *maxnb);
  }
  double ep = 0.0;
  or_smatb_force(x, type, n, b, s, *nn, *list, *maxnb, f, e, NULL, NULL, &ep, 1);
  return ep;
}

int or_fire(double *x, const int *type, int n, const or_smatb *s, const double *mass,
            double skin, int max_iter, double dt_max, double alpha0, double ftol,
            double *epot, double *fmax, int *converged) {
  double ui_ev = c_ui_ev();
  or_box b;
  or_box_set(&b, 0, NULL, NULL);
  double rl = or_smatb_rcut(s) + skin;
  int maxnb = 128;
  int *nn = (int *)malloc(sizeof(int) * n);
  int *list = (int *)malloc(sizeof(int) * (size_t)n * maxnb);
  double *f = (double *)malloc(sizeof(double) * 3 * n);
  double *e = (double *)malloc(sizeof(double) * n);
  double *v = (double *)calloc(3 * (size_t)n, sizeof(double));
  double ep = fire_forces(x, type, n, &b, s, rl, &maxnb, &nn, &list, f, e);
  double dt = dt_max * 0.1, alpha = alpha0, fm = 0.0;
  int npos = 0, it, conv = 0;
  for (it = 0; it < max_iter; ++it) {
    fm = 0.0;
    for (int i = 0; i < n; ++i) {
      double q = f[3 * i] * f[3 * i] + f[3 * i + 1] * f[3 * i + 1] + f[3 * i + 2] * f[3 * i + 2];
      if (q > fm) fm = q;
    }
    fm = sqrt(fm) * ui_ev;
    if (fm < ftol) {
      conv = 1;
      break;
    }
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        v[3 * i + d] += 0.5 * dt * f[3 * i + d] / mass[i];
        x[3 * i + d] += dt * v[3 * i + d];
      }
    ep = fire_forces(x, type, n, &b, s, rl, &maxnb, &nn, &list, f, e);
    double P = 0.0, vv = 0.0, ff = 0.0;
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) {
        v[3 * i + d] += 0.5 * dt * f[3 * i + d] / mass[i];
        P += f[3 * i + d] * v[3 * i + d];
        vv += v[3 * i + d] * v[3 * i + d];
        ff += f[3 * i + d] * f[3 * i + d];
      }
    double mixv = ff > 0.0 ? alpha * sqrt(vv) / sqrt(ff) : 0.0;
    int up = P > 0.0;
    for (int k = 0; k < 3 * n; ++k) v[k] = up ? (1.0 - alpha) * v[k] + mixv * f[k] : 0.0;
    if (up) {
      ++npos;
      if (npos > 5) {
        dt = fmin(dt * 1.1, dt_max);
        alpha *= 0.99;
      }
    } else {
      npos = 0;
      dt *= 0.5;
      alpha = alpha0;
    }
  }
  *epot = ep * ui_ev;
  *fmax = fm;
  *converged = conv;
  free(nn);
  free(list);
  free(f);
  free(e);
  free(v);
  return it;
}
