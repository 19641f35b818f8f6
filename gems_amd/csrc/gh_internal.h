// Internal declarations of the gfx950 GEMS hot-path library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/gems_hip.h"

#define GH_MAXT 8
#define GH_JMASK_HOST 0x07FFFFFF

// ---------------------------------------------------------------------------
// GEMS units and constants.  The reference evaluates them in real(dp) from
// the literals of src/Constants.F90:114,124-127,173-175; the element masses
// are default-real literals widened to real(dp) (src/Elements.F90).
// ---------------------------------------------------------------------------
namespace ghc {
inline double ui_ev() {
  const double axps_mxs = 1.0e2, uma_kg = 1.6605402e-27, qe_si = 1.60219e-19;
  const double joule_ev = 1.0 / qe_si;
  return axps_mxs * axps_mxs * uma_kg * joule_ev;
}
inline double ev_ui() { return 1.0 / ui_ev(); }
inline double kB_ui() { return 8.617385e-05 * ev_ui(); }
}  // namespace ghc

// ---------------------------------------------------------------------------
// Host side of the `prng lcg` stream (src/Random.F90:145-250).
// ---------------------------------------------------------------------------
struct GhLcg {
  int32_t idum = -123456789, ix = -1, iy = -1;
  bool stored = false;
  double g = 0.0;
  void seed(int s);
  double ranu();
  double gasdev();
  void rang(double *r1, double *r2);
};

// SMA-TB coefficients for one ordered type pair (TB.F90:106-170).
struct GhTbPair {
  double A, xi, p, q, r0, rci, rce;
  double b1, b2, b3, r1, r2, r3;
};
void gh_tb_coeffs(GhTbPair &c);

// FP32 form of one type pair's SMA-TB terms (the mixed-precision passes), in
// a length unit u = 2^-k Angstrom (k = 0: Angstrom; the fixed-point tiles
// use the unit of their integer positions).  With s = r - r0f (r0f the float
// nearest r0; s is formed from r^2 in FP64 so that it carries no rounding of
// r itself; the kernels carry s2 = 2 s) and E2' = 2^{bq s} (E2 = e^{-2q(r/r0
// - 1)} = E2' 2^{bq (r0f - r0)}, the last factor folded into the amplitudes):
//   density   xi^2 E2                               = rho_a E2'
//   force     2A(p/r0) e^{-p(r/r0-1)} - xi^2 (q/r0) E2 (eb_i + eb_j)
//           = E2' [a em1(g s) + c0 - B0 (d_i + d_j)]
//   energy    A e^{-p(r/r0-1)} = Ar E2' (1 + em1(g s))
// where em1(x) = e^x - 1, g s = -(p - 2q)(r - r0f)/r0, eb = e0 + d with e0 a
// context constant near every atom's 1/sqrt(rho) (d stays small, so its FP32
// rounding is far below eb's), a and Ar carry e^{-(p-2q)(r0f-r0)/r0} and c0 =
// a - 2 B0 e0 is rounded once from FP64.  The two large terms of the force
// (about 1.46 eV/A each for an Au nearest-neighbour pair, 0.03 apart) thereby
// cancel before any FP32 rounding.  Tail (rci < r <= rce): the quintic
// polynomials of TB.F90:134-149 in Angstrom, d = (r - rce).
struct TbF {
  double rce2q;            // rce^2 in u^2
  float r0q2, rciq;        // 2 r0f and rci in u
  float sc, rce;           // u and rce in Angstrom (tail: d = r_u sc - rce)
  float rho_a, bqh;        // density xi^2 E2 = rho_a 2^{bqh s2}, s2 = 2 s in u
  float a, c0, B0, Ar;     // force bracket and repulsion amplitude (E2 = 2^{bqh s2})
  float gh;                // -(p - 2q)/(2 r0) per u: g s = gh s2
  float e02;               // 2 e0 (tail terms)
  float b1f, b2f, b3f, b1x5, b2x4, b3x3;          // tail density polynomial and 5b1, 4b2, 3b3
  float r1f, r2f, r3f, r1m10, r2m8, r3m6;          // tail repulsion and -10r1, -8r2, -6r3
};

// Kernel-side parameter block (one per type pair): the FP64 form (the
// reference's arithmetic) and the FP32 form in Angstrom.
struct TbDev {
  double ca, eps2, p, q, r0, one_r0, rci, rce, rce2, b1, b2, b3, r1, r2, r3;
  double p_r0, q2_r0;  // p/r0, 2q/r0
  TbF f;               // FP32 form, u = 1 A
};
// Fills the FP32 form of pair t in the unit 2^-k A with the context constant e0.
void gh_tbf_fill(const GhTbPair &t, int k, double e0, TbF &o);
// e0 of a parameter set: 1/sqrt(rho) of a type-0 atom with twelve neighbours at r0.
double gh_tb_e0(const GhTbPair *tb);
// Tabulated Lennard-Jones forms per type pair (plj, slj, wca; the context's
// pot says which): prc and frc are plj's value and slope terms at the cut
// (plj_set, Pairs.F90:282-291); rs, sa4, sa8 slj's spline (slj_set,
// Pairs.F90:374-401); rc this pair's cut.  FP64 and FP32 forms.
// per type pair; sm1 keeps r1 in rs, r2 in rc and (r1 + r2)/2 in sa8
struct PljDev {
  double e, s, prc, frc, rs, sa4, sa8, rc, rc2;
  float ef, s2f, prcf, frcf, rsf, sa4f, sa8f, rcf;
};
inline bool gh_is_table_pot(int pot) { return pot >= 3 && pot <= 6; }

// Fixed-point positions of the mixed-precision SMA-TB tiles: 2^fxk steps per
// Angstrom relative to each brick's centre, with fxk (at most GH_FXK_MAX)
// chosen at grid set-up so that every staged atom lies within 2^30 steps of
// it.  The batched clusters stage positions relative to each system's first
// atom at 2^GH_BFXK steps per Angstrom (systems within 64 A of that atom).
#define GH_FXK_MAX 25
#define GH_BFXK 24
#define GH_BFXS 16777216.0  // 2^GH_BFXK

// Cell grid and its grouping into bricks.  Atoms are sorted by key =
// brick * cpb + cell-within-brick, so each brick's atoms are contiguous and
// one workgroup stages the brick's halo (the bricks' cells grown by one cell)
// into LDS.  Grid axes: nc interior cells; non-periodic axes add one guard
// cell on each side (gofs = 1, gn = nc + 2).
struct TileGeom {
  int nc[3], gn[3], gofs[3], pbc[3];
  int bd[3];      // brick edge in cells
  int nbr[3];     // bricks per axis (ceil(gn / bd))
  int cpb;        // cells per brick
  int nbrick;     // bricks
  int nkey;       // nbrick * cpb
  int slot_cap;   // LDS slots per workgroup (halo atoms)
  int tpb;        // threads per workgroup of the force passes (a multiple of 64)
  int tpb_max, tpb_avg;  // its two candidates: the fullest brick, the mean brick
  int lanes = 1;  // lanes per atom in the force passes: 4 below 131072 atoms (fixed by n)
  int chunks = 1; // workgroups per brick in the force passes (each a share of its atoms)
  int fxk;        // fixed-point steps per Angstrom: 2^fxk (mixed SMA-TB)
  double fxs;     // 2^fxk
  double L[3];
  double one_cell[3];
};
#define GH_MAXHALO 216  // (4 + 2)^3 halo cells at most (brick edge <= 4)
// Dynamic LDS of the list build: two workgroups (with their static Halo) fit
// one CU's 160 KB.
#define GH_BUILD_LDS 77824
// Quad rows every neighbour-list allocation carries past its capacity: the
// list traversals prefetch two quads ahead without a bounds clamp.
#define GH_LIST_PAD_QUADS 2

struct GhStatus {                     // device-resident counters
  unsigned long long maxdisp2_bits;   // max |dx|^2 since the last build
  int rebuild;                        // rebuild at this step
  int overflow;                       // a row exceeded the capacity (sticky)
  int maxnn;                          // longest row at the last build
  int outside;                        // an atom left a non-periodic grid
  long long nupd;                     // rebuilds so far
  int force_rebuild;                  // host request
  int tile_overflow;                  // a brick's halo exceeded slot_cap (sticky)
  int max_halo;                       // largest halo (atoms) at the last build
  int max_brick;                      // largest brick (atoms) at the last build
  int ticket;                         // binning workgroups done (last one scans)
  int halt;                           // GH_HALT_*: every state-changing kernel returns
  long long halt_step;                // step at which halt was raised
};

// Why the device stopped a run: the step's rebuild was due but not enqueued
// (the host re-runs the step with the rebuild chain), or the step cannot be
// computed (a row or a halo outgrew its capacity, an atom left the grid).
#define GH_HALT_REBUILD 1
#define GH_HALT_OVERFLOW 2
#define GH_HALT_TILE 3
#define GH_HALT_RANGE 4

// The step the device has started and the stop as it stood then, published
// by the step's first force pass into pinned host memory (the host throttles
// on `step` and reads halt without a stream operation).
struct GhProgress {
  long long step;
  long long halt_step;
  unsigned long long maxd2;  // displacement maximum (bits of |dx|^2) at the step's start
  int halt;
  int pad;
};

struct gh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int precision = GH_FP64;
  std::string err;

  // box
  int boxed = 0;
  double L[3] = {1e6, 1e6, 1e6};
  int pbc[3] = {0, 0, 0};
  int ghost = 0;  // periodic images as ghost atoms (box mic F) instead of minimum image

  // atoms
  int n = 0, ncap = 0;
  int nt = 1;
  double mass_t[GH_MAXT] = {1, 1, 1, 1, 1, 1, 1, 1};
  bool have_atoms = false;

  // potential
  int pot = GH_POT_NONE;
  GhTbPair tb[GH_MAXT * GH_MAXT];
  double lj_eps = 0, lj_sigma = 0, lj_rc = 0;
  double rcut = 0;      // potential cut radius
  double skin = 1.0;    // nb_dcut default (Neighbour.F90:237)

  // integrator
  int integ = GH_INT_VV;
  double dt = 0.0;
  double T = 0.0, gamma = 1.0;
  int rng_mode = GH_RNG_LCG;
  uint64_t philox_seed = 0;
  GhLcg lcg;
  bool first_force = true;   // forces not yet valid (interact not called)
  long long step = 0;
  long long launches = 0;  // kernels enqueued on the context's stream
  // a run stopped at step `step` + 1 after its stepa: the forces and stepb of
  // that step are still due (the next gh_run applies them first)
  bool mid_step = false;

  // neighbour grid
  int list_mode = 1;          // 0 = O(N^2), 1 = cells
  int nc[3] = {1, 1, 1};      // cells per dim (interior)
  int gofs[3] = {0, 0, 0};    // 1 for a non-periodic dim (one guard cell)
  int gn[3] = {1, 1, 1};      // cells per dim including guard cells
  double cell[3] = {1, 1, 1}; // cell edge
  int ncell = 1;
  TileGeom tg;                // brick grouping (cell mode)
  int bd_req[3] = {0, 0, 0};  // requested brick edge (0 = automatic)
  int tpb_mean = 0;           // 1: force-pass workgroups sized to the mean brick (autotuned)
  int chunks_req = 0;         // force-pass workgroups per brick (autotuned; 0 = 1)
  int chunks_env = 0;         // GEMS_HIP_CHUNKS: forced workgroups per brick (0 = free)
  int max_brick = 0;          // atoms of the fullest brick at the last sizing
  int list_rows = 0;          // GEMS_HIP_LIST_BUILD: 1 = rows (one thread per atom), 2 = coop, 0 auto
  int list_build = 0;         // kernel of the last enqueued build: 1 per-thread, 2 cooperative
  int cap = 0;                // list row capacity
  int cap_req = 0;

  // device state (cell-sorted order)
  double4 *pos = nullptr, *pos_t = nullptr;   // x,y,z,(type as double)
  double4 *pos_next = nullptr;                // positions written by a fused force pass
  double *vel = nullptr, *vel_t = nullptr;    // [3][ncap] SoA
  double *acc = nullptr, *acc_t = nullptr;    // [3][ncap]
  double *frc = nullptr;                      // [3][ncap]
  double *epa = nullptr;                      // [ncap] per-atom energy (ui)
  double *rho = nullptr, *eband = nullptr;    // [ncap]
  float *eband32 = nullptr;                   // [ncap]
  int *type = nullptr, *type_t = nullptr;     // [ncap]
  int *orig = nullptr, *orig_t = nullptr;     // caller index of sorted atom
  double *xold = nullptr;                     // [3][ncap] positions at build
  double *posv = nullptr;                     // Ermak random displacement [3][ncap]
  double *xin = nullptr;                      // gh_set_positions upload [3n] (caller order)
  double *velv = nullptr;                     // Ermak random velocity [3][ncap]
  double *gauss = nullptr;                    // host-stream gaussians [6][ncap] (caller order)
  double *gauss_h = nullptr;                  // pinned host staging
  int *cellof = nullptr;                      // [ncap] cell of sorted atom
  int *cellid = nullptr;                      // [ncap] cell of unsorted atom (rebuild scratch)
  int *cell_count = nullptr, *cell_start = nullptr;
  int *perm = nullptr;                        // [ncap]
  int *nn = nullptr;                          // [ncap]
  int *list = nullptr;                        // [cap][ncap] ELL (O(N^2) mode)
  unsigned short *list16 = nullptr;           // [cap][ncap] ELL of LDS slots (cell mode)
  void *chain_dev = nullptr;                  // rebuild-chain arguments (device copy)
  void *chain_pin = nullptr;                  // their pinned staging copy
  std::vector<unsigned char> chain_last;      // bytes last uploaded
  TbDev *tbd = nullptr;                       // [MAXT*MAXT]
  TbF *tbf = nullptr;                         // [MAXT*MAXT]
  TbF tbf0;                                   // host copy of tbf[0] (one-type kernel argument)
  int tbf_k = -1;                             // fixed-point unit exponent tbf was filled for
  double tb_e0 = 0.0;                         // eband reference of the FP32 forms
  PljDev *pljd = nullptr;                     // [MAXT*MAXT] shifted LJ per type pair
  // one-sided interactions: i feels j iff mfeel[i] & mfelt[j] (caller order)
  int *mfeel = nullptr, *mfelt = nullptr;
  bool masked = false;
  GhStatus *status = nullptr;
  GhProgress *prog_h = nullptr;               // pinned: the latest step's outcome
  GhProgress *prog_d = nullptr;               // its device address
  double *red = nullptr;                      // reduction scratch
  int nred = 0;
  double *red_h = nullptr;                    // pinned result
  int cell_alloc = 0;

  // timing: a pool of events, 8 per step, read back when the pool fills
  bool timing = false;
  std::vector<hipEvent_t> evpool;
  int evused = 0;  // steps recorded in the pool
  double t_ms[4] = {0, 0, 0, 0};
  int64_t t_cnt[4] = {0, 0, 0, 0};

  // host copies in the caller's order
  std::vector<int> type_h;
  bool ready = false;       // grid and list sized for the current setup
  bool pos_dirty = true;    // host changed positions since the last build
};

// error helpers
int gh_fail(gh_ctx *c, int code, const std::string &msg);
#define GH_HIP(c, call)                                                    \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return gh_fail(c, GH_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define GH_TRY(x)             \
  do {                        \
    int rc_ = (x);            \
    if (rc_ != 0) return rc_; \
  } while (0)

// Velocity Verlet fused into the force pass (stepb of this step, stepa of the
// next) for the atoms whose forces the pass has just summed.
struct FuseVV {
  int on;
  double dt, acel_dv, acel_dx;
  double om[GH_MAXT];
  double *vel;
  const double *xold;
  double4 *pos_next;
  unsigned long long *maxd2;  // GhStatus::maxdisp2_bits
  int ncap;
};

// Device address of the displacement maximum the integrators fold into.
inline unsigned long long *gh_maxd2(gh_ctx *c) { return &c->status->maxdisp2_bits; }

// launchers (gh_kernels.hip)
int gh_launch_rebuild(gh_ctx *c, bool force);
int gh_launch_set_positions(gh_ctx *c, const double *x_dev);
int gh_launch_forces(gh_ctx *c, hipEvent_t mid, bool fuse_vv = false);
bool gh_can_fuse_vv(const gh_ctx *c);
int gh_launch_vv_a(gh_ctx *c);
int gh_launch_vv_b(gh_ctx *c);
int gh_launch_vv_ba(gh_ctx *c);
int gh_launch_ermak_a(gh_ctx *c);
int gh_launch_ermak_b(gh_ctx *c);
int gh_launch_energies(gh_ctx *c, double *epot_ui, double *ekin_ui);
int gh_launch_rebuild_cells(gh_ctx *c);
int gh_launch_forces_tiled(gh_ctx *c, hipEvent_t mid, bool fuse_vv);
int gh_fill_int(gh_ctx *c, int *p, int v, int n);
int gh_launch_wrap(gh_ctx *c);
int gh_tile_occupancy(gh_ctx *c, int *max_halo, int *max_brick);
int gh_request_rebuild(gh_ctx *c);  // the next rebuild chain rebuilds whatever the displacement
int gh_tiled_to_global(gh_ctx *c, int *dst);
int gh_upload_tb(gh_ctx *c);
int gh_upload_tbf(gh_ctx *c, int k);
