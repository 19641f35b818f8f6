/*
 * gems_hip.h -- C ABI of the MI355X (gfx950) force-evaluation library for the
 * GEMS molecular simulator.
 *
 * The library replaces GEMS's hot path: neighbour build (src/Neighbour.F90),
 * the Lennard-Jones pair force (src/Pairs.F90), the SMA-TB/Gupta two-pass
 * force (src/TB.F90) and the velocity-Verlet / Ermak updates
 * (src/Integration.F90).  Every entry point names the reference routine it
 * stands in for.
 *
 * Conventions
 *  - Plain C types only; no exceptions cross the boundary.
 *  - Return 0 on success, a negative GH_E* code on failure; gh_last_error()
 *    then holds a message.  (GEMS's own errors STOP via werr, Errors.f90:43-59;
 *    a Fortran caller maps a non-zero return to werr(gh_last_error(ctx)).)
 *  - Units are GEMS internal units: Angstrom, amu, ps, energy "ui"
 *    (amu*A^2/ps^2).  Per-atom arrays are in the caller's atom order with the
 *    reference's per-atom layout x[3*i+d].  The library copies in and out and
 *    never keeps host pointers.
 *  - One context = one system on one device, driven by one host thread, with
 *    one HIP stream of its own (default priority).
 */
#ifndef GEMS_HIP_H
#define GEMS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gh_ctx gh_ctx;

/* error codes */
#define GH_OK 0
#define GH_EINVAL -1    /* bad argument                                   */
#define GH_EHIP -2      /* HIP runtime error                              */
#define GH_ESTATE -3    /* call out of order (e.g. no atoms / potential)  */
#define GH_EOVERFLOW -4 /* neighbour list capacity exceeded               */
#define GH_ERANGE -5    /* an atom left the cell grid of a non-periodic box */

/* precision of the pair arithmetic */
#define GH_FP64 0   /* plain FP64 everywhere                                */
#define GH_MIXED 1  /* FP32 pair terms; FP64 positions, sums, integrator    */

/* potentials */
#define GH_POT_NONE 0
#define GH_POT_SMATB 1
#define GH_POT_LJ 2
#define GH_POT_PLJ 3
#define GH_POT_SLJ 4
#define GH_POT_WCA 5
#define GH_POT_SM1 6

/* integrators */
#define GH_INT_VV 0     /* velocity Verlet   (Integration.F90:1150-1186) */
#define GH_INT_ERMAK 1  /* Ermak / Langevin  (Integration.F90:909-1031)  */

/* random streams for Ermak */
#define GH_RNG_LCG 0    /* the reference "prng lcg" stream, drawn on the host
                           in atom order (Random.F90:145-250)              */
#define GH_RNG_PHILOX 1 /* counter-based Philox4x32-10 on the device       */

/* Context lifetime.  No reference counterpart: GEMS keeps global state
 * (src/Program_Types.F90:37-102); a context holds that state for one system.
 * device: HIP ordinal.  precision: GH_FP64 or GH_MIXED. */
int gh_create(gh_ctx **ctx, int device, int precision);
void gh_destroy(gh_ctx *ctx);
const char *gh_last_error(const gh_ctx *ctx);
int gh_set_precision(gh_ctx *ctx, int precision);

/* Box.  Replaces `box size` + `set pbc` (box_setvars, Program_Types.F90:126-141;
 * set_pbc; do_pbc_1, Set_Properties.f90:444-470).  L == NULL: no box (free
 * system; the reference then uses the O(N^2) verlet_self list). */
int gh_set_box(gh_ctx *ctx, const double *L, const int *pbc);

/* Periodic images.  Replaces `box mic T|F` (CLInterpreter.F90:1723-1726):
 * ghost = 0 (the default, `mic T`) pairs each atom with the nearest image of
 * another (vdistance); ghost = 1 (`mic F`) with every first-shell image of
 * every atom, itself included, as the reference's ghost atoms do (pbcghost /
 * pbcfullghost, Neighbour.F90:1520-1794).  The two agree once each periodic
 * side is at least twice the cut plus skin; below that one atom can see
 * several images of another and only the ghost form counts them all. */
int gh_set_images(gh_ctx *ctx, int ghost);

/* One-sided interactions.  Replaces `interact A feels B` (also `A <- B`,
 * `B -> A`; Interaction.f90:69-77, a cross group with half and newton off,
 * :128-131) and lets several interactions of one pair potential share a
 * context: atom i pairs with atom j only when feel[i] & felt[j] is nonzero.
 * Give interaction k its bit: `under G` sets it in feel and felt of G's
 * atoms, `A feels B` in feel of A's and felt of B's.  i takes the whole pair
 * force and half the pair energy, as in the reference; j gets no reaction
 * unless it feels i too.  n ints each, caller order; NULL, NULL restores
 * every pair (the default).  Call after gh_set_atoms (which clears them).
 * Pair potentials only; a masked context uses the O(N^2) list. */
int gh_set_pair_masks(gh_ctx *ctx, const int *feel, const int *felt);

/* Atoms.  Replaces the atom creation commands and the intergroup atom table
 * (intergroup%at, Neighbour.F90:114; adda, :359-392).  x: 3n positions (A);
 * type: n internal types 0..nt-1 (TB.F90:194-215 numbering). */
int gh_set_atoms(gh_ctx *ctx, int n, const double *x, const int *type);
/* Per-type masses in amu (atom%mass, Atoms.F90:44; set mass). */
int gh_set_masses(gh_ctx *ctx, int nt, const double *mass);
/* Velocities, 3n (A/ps); NULL sets them to zero. */
int gh_set_velocities(gh_ctx *ctx, const double *v);
/* Accelerations F/m, 3n (A/ps^2 in ui): atom%acel (Atoms.F90:70), which
 * ermak_a (Integration.F90:963-1011) reads from the previous step's ermak_b;
 * NULL sets them to zero (as gh_set_atoms does).  Velocity Verlet recomputes
 * them from the forces. */
int gh_set_accelerations(gh_ctx *ctx, const double *a);

/* SMA-TB parameters for every type pair, row-major nt*nt arrays (eV, A).
 * Replaces smatb_set / smatb_readprm (TB.F90:106-170, 291-365); the quintic
 * tail coefficients are derived the same way.  Sets rcut = max(rce)
 * (TB.F90:283). */
int gh_set_smatb(gh_ctx *ctx, int nt, const double *A, const double *xi,
                 const double *p, const double *q, const double *r0,
                 const double *rci, const double *rce);

/* Lennard-Jones 12-6, unshifted: `interact ... under lj e s rc`
 * (lj_set, Pairs.F90:198-205).  eps in eV, sigma and rc in A. */
int gh_set_lj(gh_ctx *ctx, double eps, double sigma, double rc);

/* Lennard-Jones with its value and slope shifted to zero at the cut:
 * `interact A with B plj e s` (plj_set / plj_interact, Pairs.F90:282-370;
 * the script form fixes rc = 10 A, Pairs.F90:148-151).  One context holds
 * every pair of types: eps (eV) and sigma (A) are nt x nt, row-major and
 * symmetric (entry [ti][tj] is the reference's plj group for types ti, tj),
 * rc (A) is the cut of all of them; types come from gh_set_atoms.  Each
 * directed pair adds half its energy to atom i, as the full-list build
 * does. */
int gh_set_plj(gh_ctx *ctx, int nt, const double *eps, const double *sigma, double rc);

/* Lennard-Jones with a splined tail (Gronbech-Jensen & Farago 2014):
 * `interact ... slj e s rs` (slj_set / slj_interact, Pairs.F90:374-493).
 * 12-6 below rs, sa4 (r-rc)^4 + sa8 (r-rc)^8 from rs to the cut rc, which
 * slj_set derives from (e, s, rs).  eps, sigma, rs are nt x nt symmetric
 * (eV, A, A). */
int gh_set_slj(gh_ctx *ctx, int nt, const double *eps, const double *sigma, const double *rs);

/* Weeks-Chandler-Andersen: 12-6 + e, cut at 2^(1/6) s (wca_set /
 * wca_interact, Pairs.F90:495-560; the exponent 1/6 is a default-real
 * literal there and here).  eps, sigma nt x nt symmetric. */
int gh_set_wca(gh_ctx *ctx, int nt, const double *eps, const double *sigma);

/* sm1, a smooth attractive step (sm1_set / sm1_interact,
 * Pairs.F90:916-990): p = -e for r <= r1; from r1 to the cut r2
 * p = e (1/2 + 1/2 sin a) - e with a = pi (r - (r1 + r2)/2) / (r2 - r1), and
 * the force along the pair e cos(a) (pi/2) r / (r2 - r1) toward the other
 * atom (the reference's factor r reproduced as written).  e in eV, r1 < r2
 * in A; one set for every type pair. */
int gh_set_sm1(gh_ctx *ctx, double r1, double r2, double e);

/* Verlet skin `nb_dcut` (Neighbour.F90:237), A. */
int gh_set_skin(gh_ctx *ctx, double skin);

/* Cell-tile shape of the tiled passes.  The reference fixes only the cell
 * grid (intergroup_setcells, Neighbour.F90:587-664); the device groups cells
 * into bricks of bd[0] x bd[1] x bd[2] cells (1..4 each, 0 = automatic), one
 * workgroup per brick (or per share of a brick's atoms where gh_autotune
 * chose several; setting the shape resets that to one).  No counterpart in
 * the reference. */
int gh_set_tile(gh_ctx *ctx, const int bd[3]);
/* Autotuner: times `reps` force evaluations for each brick shape whose halo
 * fits the LDS budget and keeps the fastest; where the bricks are too few to
 * fill the chip it also tries two and four workgroups per brick, each
 * staging the brick's halo and taking a share of its atoms (results do not
 * depend on it; GEMS_HIP_CHUNKS=1|2|3|4 at context creation fixes the
 * count).  With velocities and a time step
 * set, each shape's time is that of a force evaluation plus its rebuild
 * chain times the expected rebuilds per step (2 v_max dt / skin, at most 1).
 * Writes the chosen shape and its time per force evaluation (us); bd_out =
 * {0,0,0} when the O(N^2) list is in use (nothing to tune).  Positions and
 * velocities are unchanged.  No counterpart in the reference. */
int gh_autotune(gh_ctx *ctx, int reps, int bd_out[3], double *us_out);

/* New positions (3n, A, caller order) for the next gh_compute when the
 * host integrates: what an interaction plugin reads from ig%at(i)%o%pos
 * (Neighbour.F90:114) at each interact() (Interaction.f90:135-241).  The
 * neighbour lists and cell order are kept; each coordinate is taken at the
 * periodic image nearest the atom's current one (so the host's do_pbc wrap,
 * Set_Properties.f90:444-470, is no displacement) and the skin test of the
 * next gh_compute sees the displacement since the last build. */
int gh_set_positions(gh_ctx *ctx, const double *x);

/* One force evaluation == interact() (Interaction.f90:135-241): skin test
 * and neighbour rebuild (test_update/update, Neighbour.F90:1242-1346), then
 * smatb_interact (TB.F90:457-586) or lj_interact (Pairs.F90:207-279).
 * F: 3n forces (ui/A), e_atom: n per-atom energies (ui), e_tot: sum (ui).
 * Any output pointer may be NULL. */
int gh_compute(gh_ctx *ctx, double *F, double *e_atom, double *e_tot);

/* The current full neighbour list (ig%nn, ig%list; Neighbour.F90:120-121)
 * at radius rcut+skin, in the caller's atom indices, each row sorted
 * ascending (with ghost images, an atom once per image in range, itself
 * included).  nn: n counts; idx: n*cap entries, row i at idx[i*cap].
 * Returns GH_EOVERFLOW if a row is longer than cap. */
int gh_neighbors(gh_ctx *ctx, int *nn, int *idx, int cap);

/* Random stream: `prng lcg` + `prng seed s` (Random.F90:137-172). */
int gh_prng_lcg(gh_ctx *ctx, int seed);

/* The `prng lcg` stream's state as a checkpoint stores it: (idum, iy, ix),
 * the record lcg_write_chpseed writes and lcg_read_chpseed reads
 * (Random.F90:178-190).  Setting it resumes the reference's stream where a
 * checkpoint left it (draws come in pairs, so no cached normal is pending). */
int gh_get_lcg_state(gh_ctx *ctx, int state[3]);
int gh_set_lcg_state(gh_ctx *ctx, const int state[3]);
/* `set tempgdist T` on all atoms (set_gdist + set_scal_vel,
 * Set_Properties.f90:301-377), drawing from the lcg stream. */
int gh_set_gdist(gh_ctx *ctx, double T);

/* Integrators: `evolve v_verlet` and `evolve ermak T gamma`
 * (integrate_cli, Integration.F90:118-230; set_ermak :909-961).
 * dt in ps.  rng: GH_RNG_LCG (reproduces GEMS's stream) or GH_RNG_PHILOX
 * (device counter-based; philox_seed keys it). */
int gh_integrator_vv(gh_ctx *ctx, double dt);
int gh_integrator_ermak(gh_ctx *ctx, double dt, double T, double gamma,
                        int rng, uint64_t philox_seed);

/* `dinamica nsteps` (dinamic, Programs.f90:114-167).  Every `each` steps
 * (each > 0) one row [epot, ekin, epot+ekin] in eV is written to energy
 * (write_energy, Output.F90:824-831; ekin about the centre of mass,
 * Inq_Properties.F90:399-422).  energy may be NULL when each <= 0. */
int gh_run(gh_ctx *ctx, int nsteps, int each, double *energy);

/* Energies of the current state in eV (write_energy). */
int gh_energies(gh_ctx *ctx, double *epot, double *ekin);

/* `lbfgs` (lbfgs_minimizator, Quasi_Newton.f90:160-232): limited-memory
 * BFGS with m correction pairs (the reference: 5) on all atoms, forces from
 * the device path; stops when |grad| / max(1, |x|) < eps (the reference's
 * test, eps = 1e-8 there, gradient in eV/A) or after max_iter iterations.
 * The minimum stays in the context; iterations and the energy (eV) out. */
int gh_lbfgs(gh_ctx *ctx, int m, double eps, int max_iter, int *iters, double *epot_ev);

/* Copy the state back: 3n positions, 3n velocities, 3n forces (ui/A), n
 * per-atom energies (ui).  Any pointer may be NULL. */
int gh_get_state(gh_ctx *ctx, double *x, double *v, double *F, double *e);

/* Counters: nupd = neighbour rebuilds (nupd_vlist, Neighbour.F90:238),
 * maxnn = longest list row at the last build, cap = row capacity, steps =
 * MD steps taken, launches = kernels enqueued so far. */
typedef struct {
  int64_t nupd;
  int32_t maxnn;
  int32_t cap;
  int32_t ncells[3];
  int32_t list_mode; /* 0 = O(N^2) (verlet_self), 1 = cells (linkedcell_self) */
  int64_t steps;
  int64_t launches;  /* kernels the context has enqueued (no counterpart) */
  int32_t list_build; /* kernel of the last enqueued list build: 0 none, 1 one
                         thread per atom, 2 cooperative (no counterpart) */
  int32_t chunks;     /* workgroups per brick of the force passes, each taking
                         a share of its atoms (cell mode; no counterpart) */
} gh_stats;
int gh_get_stats(gh_ctx *ctx, gh_stats *s);

/* Kernel timing with HIP events on the context's stream.  While enabled,
 * the library records events around the force kernels; gh_get_timing
 * returns accumulated milliseconds and launch counts for: [0] density pass
 * (SMA-TB pass 1), [1] force pass (SMA-TB pass 2 or LJ), [2] integrator,
 * [3] neighbour rebuild. */
int gh_set_timing(gh_ctx *ctx, int enable);
int gh_get_timing(gh_ctx *ctx, double ms[4], int64_t count[4]);

/* ------------------------------------------------------------------------
 * Batched free clusters.  GEMS runs one system per process
 * (src/Program_Types.F90:89-102) and replica exchange as one replica per MPI
 * rank (src/ReplicaExchange.f90:27-151); a batch holds many independent
 * unboxed systems on one device and advances all of them in one launch, one
 * workgroup per system (largest first).  Neighbour lists are the O(N^2)
 * verlet_self lists of each system (Neighbour.F90:676-759); forces are the
 * SMA-TB passes of TB.F90:367-586; Langevin is ermak_a/ermak_b
 * (Integration.F90:963-1031) with the Philox stream.
 * ------------------------------------------------------------------------ */
typedef struct gh_batch gh_batch;
int gh_batch_create(gh_batch **b, int device, int precision);
void gh_batch_destroy(gh_batch *b);
const char *gh_batch_last_error(const gh_batch *b);
int gh_batch_set_precision(gh_batch *b, int precision);
/* nsys systems; atoms of system s are off[s] .. off[s+1]-1 of x (3 per atom,
 * A) and type (internal types 0..nt-1).  Velocities are set to zero. */
int gh_batch_set_systems(gh_batch *b, int nsys, const int *off, const double *x,
                         const int *type);
int gh_batch_set_smatb(gh_batch *b, int nt, const double *A, const double *xi,
                       const double *p, const double *q, const double *r0,
                       const double *rci, const double *rce);
int gh_batch_set_masses(gh_batch *b, int nt, const double *mass);
int gh_batch_set_skin(gh_batch *b, double skin);
int gh_batch_set_velocities(gh_batch *b, const double *v);
/* Per-system temperatures (K) used by gh_batch_langevin. */
int gh_batch_set_temperatures(gh_batch *b, const double *T);
/* Forces (ui/A, 3 per atom) and per-system potential energies (eV) of the
 * current positions: one interact() per system.  Either pointer may be NULL. */
int gh_batch_compute(gh_batch *b, double *F, double *epot);
/* nsteps Ermak steps (dt ps, gamma 1/ps) of every system in one launch.
 * epot, ekin (eV, per system, after the last step; ekin about the CM) may be
 * NULL. */
int gh_batch_langevin(gh_batch *b, int nsteps, double dt, double gamma, uint64_t seed,
                      double *epot, double *ekin);
/* One replica-exchange sweep over ladders of systems (the pairing and
 * Metropolis rule of src/ReplicaExchange.f90:63-138): rung r of ladder l is
 * system sys[l*R + r] at temperature T[sys]; pairs (r, r+1) with
 * r % 2 == parity exchange temperatures with probability min(1, exp(-du)),
 * du = (E_upper - E_lower)(beta_lower - beta_upper), energies from the last
 * gh_batch_langevin, and the velocities of both are rescaled by
 * sqrt(T_new / T_old).  The rung order is updated in sys.  Returns the number
 * of accepted swaps in *naccept. */
int gh_batch_exchange(gh_batch *b, int nladder, int R, int *sys, int parity, uint64_t seed,
                      int *naccept);
/* FIRE relaxation (Bitzek et al., PRL 97, 170201 (2006)) of every system, one
 * workgroup per system: dt starts at dt_max/10, N_min = 5, f_inc = 1.1,
 * f_dec = 0.5, f_alpha = 0.99.  With precision GH_MIXED the pair terms are
 * FP32 until max|F| < fswitch (eV/A), then FP64 to max|F| < ftol.  Outputs
 * per system: iterations, converged flag, final epot (eV), final max|F|. */
int gh_batch_fire(gh_batch *b, int max_iter, double dt_max, double alpha0, double fswitch,
                  double ftol, int *iters, int *converged, double *epot, double *fmax);
/* Copy back positions and velocities (3 per atom). */
int gh_batch_get_state(gh_batch *b, double *x, double *v);

/* ------------------------------------------------------------------------
 * Forces without allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  The SMA-TB forces of one free cluster (no box), computed
 * as gh_compute computes them on a context without a box (the O(N^2)
 * verlet_self list, both passes, `precision` GH_FP64 or GH_MIXED), from the
 * caller's DEVICE arrays: x (3n, AoS, A), type (n, values 0..nt-1) in, F
 * (3n, AoS, ui/A), e_atom (n, ui), epot (1, ui) and status (1: 0, or
 * GH_EOVERFLOW when a list row exceeded its capacity) out.  The parameter
 * matrices (nt x nt, row-major, as gh_set_smatb) are host arrays.  Every
 * intermediate lives in one static device array of the library (about
 * 1.6 MB per device, held until the process ends); nothing is allocated.
 * Limits: n <= 2048, nt <= 2 (GH_EINVAL beyond).  Work is enqueued on
 * `stream` (a hipStream_t; NULL: the null stream) and nothing waits for it.
 * Calls that overlap on one device share the static array and race.
 * ------------------------------------------------------------------------ */
int gh_cluster_forces_static(int device, int precision, int n, int nt, const double *x,
                             const int *type, const double *A, const double *xi,
                             const double *p, const double *q, const double *r0,
                             const double *rci, const double *rce, double skin, double *F,
                             double *e_atom, double *epot, int *status, void *stream);

#ifdef __cplusplus
}
#endif
#endif
