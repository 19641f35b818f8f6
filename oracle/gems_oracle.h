/*
 * gems_oracle.h -- CPU FP64 restatement of the GEMS force-evaluation hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, the smoke check and the CPU-baseline leg of bench.py load it.
 * The product library (gems_amd/csrc) never links, calls or falls back to it.
 *
 * Every function cites the reference routine it restates (paths relative to
 * the GEMS source tree).  Layout follows the reference's per-atom pos(3):
 * x[3*i+d], i = 0..n-1 in the reference's atom order.
 *
 * Units are GEMS internal units: Angstrom, amu, ps, energy "ui" = amu A^2/ps^2.
 */
#ifndef GEMS_ORACLE_H
#define GEMS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXT 8

/* ---- constants (src/Constants.F90:101-175, src/Elements.F90:55-246) ---- */
/* out[0]=ev_ui out[1]=ui_ev out[2]=kB_ui out[3]=ui_kcm out[4]=kB_ev        */
void or_constants(double out[5]);
/* element mass as the reference stores it: a default-real literal widened to
 * real(dp) (src/Elements.F90:92,110,142).  z = atomic number. */
double or_element_mass(int z);

/* ---- random numbers (src/Random.F90:145-250) ---- */
typedef struct {
  int32_t idum, ix, iy;
  int stored;
  double g;
} or_rng;
void or_rng_init(or_rng *r, int seed);          /* lcg_init_ran(lseed)   */
double or_ranu(or_rng *r);                      /* lcg_ranu              */
double or_gasdev(or_rng *r);                    /* gasdev                */
void or_rang(or_rng *r, double *r1, double *r2); /* rang (r2 may be NULL) */

/* ---- box and minimum image (src/Program_Types.F90:126-141,
 *      src/Inq_Properties.F90:137-157) ---- */
typedef struct {
  int boxed;      /* a box was defined (box size)                          */
  int mic;        /* minimum image convention (default .true.)             */
  int pbc[3];     /* per-dimension periodicity (set pbc, uniform per atom)  */
  double L[3];    /* box lengths (1e6 when unboxed)                        */
  double one_box[3];
} or_box;
void or_box_set(or_box *b, int boxed, const double L[3], const int pbc[3]);
/* vd = x_a - x_b with MIC */
void or_vdistance(const or_box *b, const double *xa, const double *xb, double vd[3]);

/* ---- neighbour lists (src/Neighbour.F90) ---- */
/* Full (half=.false.) self lists as the OpenMP build makes them.
 * mode_out: 0 = verlet_self O(N^2) (Neighbour.F90:676-759, rd>rc2 -> cycle),
 *           1 = linkedcell_self (Neighbour.F90:991-1087, rd<rc2 keeps).
 * list is row-major [n][maxnb].  Returns 0, or the needed maxnb (>0) on
 * overflow, or -1 on an atom outside the cell grid. */
int or_nb_build(const double *x, int n, const or_box *b, double rlist,
                int maxnb, int *nn, int *list, int *mode_out, int ncells_out[3],
                int nthreads);
/* Brute-force O(N^2) enumeration: keeps j != i with rd < rlist^2 (strict=1)
 * or rd <= rlist^2 (strict=0), j ascending. */
int or_nb_brute(const double *x, int n, const or_box *b, double rlist,
                int strict, int maxnb, int *nn, int *list);

/* ---- SMA-TB / Gupta (src/TB.F90) ---- */
typedef struct {
  int nt;
  int prm[OR_MAXT * OR_MAXT];
  double ca[OR_MAXT * OR_MAXT], eps[OR_MAXT * OR_MAXT], p[OR_MAXT * OR_MAXT],
      q[OR_MAXT * OR_MAXT], r0[OR_MAXT * OR_MAXT], rci[OR_MAXT * OR_MAXT],
      rce[OR_MAXT * OR_MAXT], b1[OR_MAXT * OR_MAXT], b2[OR_MAXT * OR_MAXT],
      b3[OR_MAXT * OR_MAXT], r1[OR_MAXT * OR_MAXT], r2[OR_MAXT * OR_MAXT],
      r3[OR_MAXT * OR_MAXT];
} or_smatb;
void or_smatb_init(or_smatb *s, int nt);
/* smatb_set (TB.F90:106-170); i,j are 0-based internal types */
void or_smatb_set(or_smatb *s, int i, int j, double a, double eps, double p,
                  double q, double r0, double rci, double rce);
double or_smatb_rcut(const or_smatb *s); /* maxval(rce) (TB.F90:283) */
/* smatb_preinteraction + smatb_interact (TB.F90:367-586), full list, newton
 * off.  f,e are OVERWRITTEN (the caller zeroes them in the reference,
 * Interaction.f90:180-188).  rho (band) and eband are optional outputs. */
void or_smatb_force(const double *x, const int *type, int n, const or_box *b,
                    const or_smatb *s, const int *nn, const int *list,
                    int maxnb, double *f, double *e, double *rho,
                    double *eband, double *epot, int nthreads);

/* ---- Lennard-Jones (src/Pairs.F90:198-279) ---- */
void or_lj_force(const double *x, int n, const or_box *b, double eps,
                 double sigma, double rc, const int *nn, const int *list,
                 int maxnb, double *f, double *e, double *epot, int nthreads);

/* ---- shifted Lennard-Jones "plj" (src/Pairs.F90:282-370) ----
 * One table over all type pairs (the reference has one plj intergroup per
 * group pair); one cut rc for all (the script form fixes 10 A). */
/* The same table serves the splined (slj, Pairs.F90:374-493) and WCA
 * (wca, Pairs.F90:495-560) forms: kind 0 = plj, 1 = slj, 2 = wca, 3 = sm1; rcp holds
 * each pair's cut (plj: rc; slj: from slj_set; wca: 2^(1/6) s) and rc their
 * maximum. */
typedef struct {
  int nt, kind;
  double rc;
  double e[OR_MAXT * OR_MAXT], s[OR_MAXT * OR_MAXT], prc[OR_MAXT * OR_MAXT],
      frc[OR_MAXT * OR_MAXT], rs[OR_MAXT * OR_MAXT], sa4[OR_MAXT * OR_MAXT],
      sa8[OR_MAXT * OR_MAXT], rcp[OR_MAXT * OR_MAXT];
} or_plj;
void or_plj_init(or_plj *p, int nt, double rc);
/* plj_set (Pairs.F90:282-291) for types i, j (0-based), both orders */
void or_plj_set(or_plj *p, int i, int j, double e, double s);
/* slj_set (Pairs.F90:374-401) / wca_set (Pairs.F90:495-506); the table's
 * kind must be 1 / 2 (or_plj_init then set kind) */
void or_slj_set(or_plj *p, int i, int j, double e, double s, double rs);
void or_wca_set(or_plj *p, int i, int j, double e, double s);
/* sm1_set (Pairs.F90:916-922): one (r1, r2, e) for every type pair; kind 3
 * (rs holds r1, rcp the cut r2) */
void or_sm1_set(or_plj *p, double r1, double r2, double e);
/* plj_interact (Pairs.F90:293-370), slj_interact (:403-493) or
 * wca_interact (:508-560) by the table's kind; full list, newton off; f, e
 * overwritten */
void or_plj_force(const double *x, const int *type, int n, const or_box *b, const or_plj *p,
                  const int *nn, const int *list, int maxnb, double *f, double *e,
                  double *epot, int nthreads);

/* ---- integrators (src/Integration.F90) ---- */
void or_vv_a(int n, double dt, double *x, double *v, double *a,
             const double *f, const double *one_mass);      /* :1150-1169 */
void or_vv_b(int n, double dt, double *v, double *a, const double *f,
             const double *one_mass);                        /* :1171-1186 */
typedef struct {
  double c0, c1, c2, sdr, sdv, crv1, crv2, skt, dt, gamma;
} or_ermak;
void or_ermak_set(or_ermak *k, double dt, double gamma, double T); /* :909-961 */
void or_ermak_a(const or_ermak *k, or_rng *r, int n, double dt, double *x,
                const double *v, const double *a,
                const double *one_sqrt_mass, double *pos_v,
                double *vel_v);                               /* :963-1011 */
void or_ermak_b(const or_ermak *k, int n, double *v, double *a,
                const double *f, const double *one_mass,
                const double *vel_v);                         /* :1013-1031 */
/* do_pbc_1 (src/Set_Properties.f90:444-470); boxcr may be NULL */
void or_do_pbc(int n, const or_box *b, double *x, int *boxcr);

/* ---- observables ---- */
/* inq_kin_energy (src/Inq_Properties.F90:399-422): KE about the CM, ui */
double or_ekin(int n, const double *v, const double *mass);
/* set_gdist + set_scal_vel (src/Set_Properties.f90:301-377) */
void or_set_gdist(or_rng *r, int n, double T, double *v, const double *mass,
                  const double *one_mass);

/* ---- the MD loop (src/Programs.f90:114-167 + Interaction.f90:135-241) ----*/
typedef struct {
  /* system */
  int n;
  double *x, *v, *a, *f, *e;   /* [3n],[3n],[3n],[3n],[n] */
  const int *type;             /* [n] internal types (SMA-TB), 0-based */
  const double *mass;          /* [n] */
  or_box box;
  /* potential: 0 = smatb, 1 = lj, 2 = plj */
  int pot;
  const or_smatb *tb;
  double lj_eps, lj_sigma, lj_rc;
  const or_plj *plj;
  double skin;                 /* nb_dcut */
  /* integrator: 0 = velocity Verlet, 1 = Ermak */
  int integ;
  double dt, T, gamma;
  or_rng *rng;
  /* run */
  int nsteps, each;
  double *energy;              /* [nsteps/each][3] eV: epot, ekin, total */
  int nthreads;
  /* state kept between calls (first call: set first_call=1) */
  int first_call;
  double *pos_old;             /* [3n] */
  int maxnb;
  int *nn, *list;              /* [n], [n*maxnb] */
  int *boxcr;                  /* [3n] */
  double *pos_v, *vel_v;       /* [3n] Ermak scratch */
  /* outputs */
  int nupd;                    /* neighbour rebuilds (nupd_vlist) */
  double epot;                 /* last ig%epot, ui */
} or_md;
/* Returns 0, or a positive maxnb needed (list overflow), or -1. */
int or_md_run(or_md *md);
/* One interact() call: test_update/update + forces.  Returns like above. */
int or_md_interact(or_md *md);

/* ---- FIRE minimisation (Bitzek et al., PRL 97, 170201 (2006)) ----
 * Not in the reference build (src/FIRE.f90 does not compile); restated from
 * the paper for the batched relaxations: velocity-Verlet step, then
 * P = F.v, v <- (1-a) v + a |v| F/|F|; P > 0 for more than 5 steps:
 * dt <- min(1.1 dt, dt_max), a <- 0.99 a; P <= 0: dt <- dt/2, v <- 0,
 * a <- alpha0.  Start: v = 0, dt = dt_max/10.  Free system, SMA-TB, FP64.
 * Returns iterations; x is updated; *fmax in eV/A, *epot in eV. */
int or_fire(double *x, const int *type, int n, const or_smatb *s, const double *mass,
            double skin, int max_iter, double dt_max, double alpha0, double ftol,
            double *epot, double *fmax, int *converged);

#ifdef __cplusplus
}
#endif
#endif
