// Host side of the C ABI declared in include/gems_hip.h.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstdlib>
#include <new>
#include <deque>
#include <thread>
#include <vector>

#include "gh_internal.h"

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
int gh_fail(gh_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

static inline int roundup(int v, int m) { return (v + m - 1) / m * m; }

template <typename T>
static int dalloc(gh_ctx *c, T **p, size_t count) {
  if (*p) {
    GH_HIP(c, hipFree(*p));
    *p = nullptr;
  }
  if (count == 0) return 0;
  GH_HIP(c, hipMalloc((void **)p, sizeof(T) * count));
  return 0;
}


static int read_status(gh_ctx *c, GhStatus *h) {
  GH_HIP(c, hipMemcpyAsync(h, c->status, sizeof(GhStatus), hipMemcpyDeviceToHost, c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}
static int write_status(gh_ctx *c, const GhStatus *h) {
  GH_HIP(c, hipMemcpyAsync(c->status, h, sizeof(GhStatus), hipMemcpyHostToDevice, c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// `prng lcg` (Random.F90:145-250): xorshift(13,-17,5) on ix combined with a
// Park-Miller (Schrage) generator iy; Marsaglia polar normals in pairs.
// ---------------------------------------------------------------------------
void GhLcg::seed(int s) {
  idum = -std::abs(s);
  stored = false;
}

double GhLcg::ranu() {
  const int32_t ia = 16807, im = 2147483647, iq = 127773, ir = 2836;
  // am is evaluated in default (single) real by the reference: the integer
  // im rounds to 2^31 and am = (1 - 2^-24) * 2^-31 exactly.
  const double am = (double)(nextafterf(1.0f, 0.0f) / (float)im);
  if (idum <= 0 || iy < 0) {
    iy = (888889999 ^ std::abs(idum)) | 1;
    ix = 777755555 ^ std::abs(idum);
    idum = std::abs(idum) + 1;
  }
  uint32_t u = (uint32_t)ix;
  u ^= u << 13;
  u ^= u >> 17;
  u ^= u << 5;
  ix = (int32_t)u;
  const int32_t k = iy / iq;
  iy = ia * (iy - k * iq) - ir * k;
  if (iy < 0) iy += im;
  return am * (double)((im & (ix ^ iy)) | 1);
}

double GhLcg::gasdev() {
  if (stored) {
    stored = false;
    return g;
  }
  double v1, v2, rsq;
  do {
    v1 = 2.0 * ranu() - 1.0;
    v2 = 2.0 * ranu() - 1.0;
    rsq = v1 * v1 + v2 * v2;
  } while (!(rsq > 0.0 && rsq < 1.0));
  rsq = sqrt(-2.0 * log(rsq) / rsq);
  g = v2 * rsq;
  stored = true;
  return v1 * rsq;
}

void GhLcg::rang(double *r1, double *r2) {
  *r1 = gasdev();
  const double aux = gasdev();
  if (r2) *r2 = aux;
}

// Quintic tail of Logsdail et al. as smatb_set derives it (TB.F90:134-149).
void gh_tb_coeffs(GhTbPair &c) {
  const double dcut = c.rce - c.rci;
  const double ab = -1.0 / (dcut * dcut * dcut);
  double bb = -(c.q / c.r0) / (dcut * dcut);
  double cb = -((c.q / c.r0) * (c.q / c.r0)) / dcut;
  double aux = c.xi * exp(-c.q * (c.rci / c.r0 - 1.0));
  c.b1 = aux * (12.0 * ab - 6.0 * bb + cb) / (2.0 * (dcut * dcut));
  c.b2 = aux * (15.0 * ab - 7.0 * bb + cb) / dcut;
  c.b3 = aux * (20.0 * ab - 8.0 * bb + cb) / 2.0;
  aux = c.A * exp(-c.p * (c.rci / c.r0 - 1.0));
  bb = -(c.p / c.r0) / (dcut * dcut);
  cb = -((c.p / c.r0) * (c.p / c.r0)) / dcut;
  c.r1 = aux * (12.0 * ab - 6.0 * bb + cb) / (2.0 * (dcut * dcut));
  c.r2 = aux * (15.0 * ab - 7.0 * bb + cb) / dcut;
  c.r3 = aux * (20.0 * ab - 8.0 * bb + cb) / 2.0;
}

double gh_tb_e0(const GhTbPair *tb) { return 1.0 / (tb[0].xi * sqrt(12.0)); }

void gh_tbf_fill(const GhTbPair &t, int k, double e0, TbF &o) {
  memset(&o, 0, sizeof(o));
  const double log2e = 1.4426950408889634;
  const double fxs = ldexp(1.0, k);
  const float fxsf = (float)fxs, uf = (float)ldexp(1.0, -k);  // powers of two: exact scalings
  const float r0f = (float)t.r0;
  const double dr0 = (double)r0f - t.r0;  // r - r0 = s + dr0 with s = r - r0f
  const double bq = -2.0 * t.q / t.r0 * log2e;
  const double gam = (t.p - 2.0 * t.q) / t.r0;
  const double ebq = exp2(bq * dr0);   // E2 = E2' ebq
  const double kap = exp(-gam * dr0);  // e^{-p(r/r0-1)} = E2 e^{-gam s} kap
  o.rce2q = t.rce * t.rce * fxs * fxs;
  o.r0q2 = 2.0f * r0f * fxsf;
  o.rciq = (float)t.rci * fxsf;
  o.sc = uf;
  o.rce = (float)t.rce;
  o.bqh = (float)bq * uf * 0.5f;
  o.rho_a = (float)(t.xi * t.xi * ebq);
  const double a = 2.0 * t.A * t.p / t.r0 * kap, B0 = t.xi * t.xi * t.q / t.r0;
  o.a = (float)(a * ebq);
  o.B0 = (float)(B0 * ebq);
  o.c0 = (float)((a - 2.0 * B0 * e0) * ebq);
  o.Ar = (float)(t.A * kap * ebq);
  o.gh = (float)(-gam) * uf * 0.5f;
  o.e02 = (float)(2.0 * e0);
  o.b1f = (float)t.b1;
  o.b2f = (float)t.b2;
  o.b3f = (float)t.b3;
  o.b1x5 = (float)(5.0 * t.b1);
  o.b2x4 = (float)(4.0 * t.b2);
  o.b3x3 = (float)(3.0 * t.b3);
  o.r1f = (float)t.r1;
  o.r2f = (float)t.r2;
  o.r3f = (float)t.r3;
  o.r1m10 = (float)(-10.0 * t.r1);
  o.r2m8 = (float)(-8.0 * t.r2);
  o.r3m6 = (float)(-6.0 * t.r3);
}

// FP32 forms of every type pair in the unit 2^-k A (the tiled passes' fixed
// point), uploaded on the context's stream.
int gh_upload_tbf(gh_ctx *c, int k) {
  TbF f[GH_MAXT * GH_MAXT];
  memset(f, 0, sizeof(f));
  for (int a = 0; a < c->nt; ++a)
    for (int b = 0; b < c->nt; ++b) gh_tbf_fill(c->tb[a * GH_MAXT + b], k, c->tb_e0, f[a * GH_MAXT + b]);
  if (!c->tbf) GH_TRY(dalloc(c, &c->tbf, GH_MAXT * GH_MAXT));
  GH_HIP(c, hipMemcpyAsync(c->tbf, f, sizeof(f), hipMemcpyHostToDevice, c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));  // f lives until the copy is done
  c->tbf0 = f[0];
  c->tbf_k = k;
  return 0;
}

int gh_upload_tb(gh_ctx *c) {
  TbDev h[GH_MAXT * GH_MAXT];
  memset(h, 0, sizeof(h));
  c->tb_e0 = gh_tb_e0(c->tb);
  for (int a = 0; a < c->nt; ++a)
    for (int b = 0; b < c->nt; ++b) {
      const GhTbPair &t = c->tb[a * GH_MAXT + b];
      TbDev &d = h[a * GH_MAXT + b];
      d.ca = t.A;
      d.eps2 = t.xi * t.xi;
      d.p = t.p;
      d.q = t.q;
      d.r0 = t.r0;
      d.one_r0 = 1.0 / t.r0;
      d.rci = t.rci;
      d.rce = t.rce;
      d.rce2 = t.rce * t.rce;
      d.b1 = t.b1;
      d.b2 = t.b2;
      d.b3 = t.b3;
      d.r1 = t.r1;
      d.r2 = t.r2;
      d.r3 = t.r3;
      d.p_r0 = t.p / t.r0;
      d.q2_r0 = 2.0 * t.q / t.r0;
      gh_tbf_fill(t, 0, c->tb_e0, d.f);
    }
  GH_TRY(dalloc(c, &c->tbd, GH_MAXT * GH_MAXT));
  // ordered on the context's stream (a non-blocking stream does not wait for
  // null-stream copies); h lives until the stream has drained
  GH_HIP(c, hipMemcpyAsync(c->tbd, h, sizeof(h), hipMemcpyHostToDevice, c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  return gh_upload_tbf(c, c->tbf_k < 0 ? GH_FXK_MAX : c->tbf_k);
}

// ---------------------------------------------------------------------------
// grid / list sizing
// ---------------------------------------------------------------------------
// Cell grid as intergroup_setcells chooses it (Neighbour.F90:587-664): cells
// of edge >= rcut+skin; no box, or fewer than 4 cells on every axis, selects
// the O(N^2) list.  Periodic axes need >= 3 cells for the 27-cell stencil to
// visit distinct cells; otherwise the O(N^2) list is used.  Non-periodic axes
// get one guard cell on each side so atoms that drift out of the box within
// a cell length are still binned exactly.
static const int GH_LDS_BYTES = 65536;

// Halo slots a workgroup of the tiled passes can stage: the LDS left after
// the static tables, over the larger per-slot need of the two passes.
size_t gh_tile_slot_bytes(int pass, int precision, int nt);
size_t gh_tile_static_bytes(int precision, int nt);
size_t gh_pair_slot_bytes(int pot, int precision);
static int max_slots(const gh_ctx *c) {
  const bool pair = c->pot == GH_POT_LJ || gh_is_table_pot(c->pot);
  const size_t per = pair ? gh_pair_slot_bytes(c->pot, c->precision)
                          : std::max(gh_tile_slot_bytes(0, c->precision, c->nt),
                                     gh_tile_slot_bytes(1, c->precision, c->nt));
  const size_t stat = pair ? gh_tile_static_bytes(GH_FP64, 1) + sizeof(PljDev) * GH_MAXT * GH_MAXT
                           : gh_tile_static_bytes(c->precision, c->nt);
  return (int)std::min<size_t>((GH_LDS_BYTES - stat) / per - 1, 65534);  // - the sentinel
}

// Threads per workgroup of the force passes, two candidates: one per atom of
// the fullest brick at sizing (plus 1/16 for the atoms' motion), so that no
// brick's atom loop runs a second round (spare wavefronts of emptier bricks
// end at once); or one per atom of the mean brick (fewer wave slots held,
// a few bricks run a short second round).  Rounded up to whole wavefronts,
// at least one thread per halo cell (halo_setup), at most 1024.  The
// autotuner times both; the fullest-brick rule is the default.
#ifndef GH_STAGE_ROUNDS
#define GH_STAGE_ROUNDS 3
#endif
static int force_pass_threads(const TileGeom &t, int atoms) {
  const int nh = (t.bd[0] + 2) * (t.bd[1] + 2) * (t.bd[2] + 2);
  const int need = std::max(nh, atoms);
  return std::min(1024, (need + 63) / 64 * 64);
}
// Below 131,072 atoms (fewer than two wavefronts per SIMD at one lane per
// atom) four lanes share each atom's row, so a pass has four times the
// wavefronts to hide its latencies; the lane count depends on the atom count
// alone, which fixes the order of every FP64 sum whatever the brick shape.
// With t.chunks workgroups per brick, each takes a contiguous share of the
// brick's atoms and stages the whole halo: it gets at least a third as many
// threads as halo slots, so that the staging takes at most three rounds of
// loads (the atom loop's spare threads return at once).
static void set_force_pass_threads(gh_ctx *c) {
  TileGeom &t = c->tg;
  t.lanes = c->n < 131072 ? 4 : 1;
  t.chunks = std::max(1, c->chunks_req);
  const int mb = (c->max_brick + t.chunks - 1) / t.chunks;
  const long long units = (long long)t.nbrick * t.chunks;
  const int stage = t.chunks > 1 ? (t.slot_cap + GH_STAGE_ROUNDS - 1) / GH_STAGE_ROUNDS : 0;
  t.tpb_max = force_pass_threads(t, std::max(stage, t.lanes * (mb + mb / 16 + 1)));
  t.tpb_avg = force_pass_threads(t, std::max(stage, t.lanes * (int)((c->n + units - 1) / units)));
  t.tpb = c->tpb_mean ? t.tpb_avg : t.tpb_max;
}

// Brick edges (cells) for the tiled passes: the largest brick whose expected
// halo fits the LDS slot budget with a 30% margin; ties go to the smaller
// halo.  bd_req (gh_set_tile) overrides.
static void choose_bricks(gh_ctx *c, int limit) {
  TileGeom &t = c->tg;
  double cells = 1.0;
  for (int d = 0; d < 3; ++d) cells *= t.nc[d];
  const double apc = c->n / cells;  // atoms per interior cell
  int best[3] = {1, 1, 1};
  long bestc = 0, besth = 1L << 40;
  for (int bz = 1; bz <= 4; ++bz)
    for (int by = 1; by <= 4; ++by)
      for (int bx = 1; bx <= 4; ++bx) {
        const int b[3] = {bx, by, bz};
        bool ok = true;
        for (int d = 0; d < 3; ++d)
          if (b[d] > t.gn[d]) ok = false;
        if (!ok) continue;
        const long halo = (long)(bx + 2) * (by + 2) * (bz + 2);
        if (apc * halo * 1.3 > limit && !(bx == 1 && by == 1 && bz == 1)) continue;
        const long nc = (long)bx * by * bz;
        if (nc > bestc || (nc == bestc && halo < besth)) {
          bestc = nc;
          besth = halo;
          best[0] = bx;
          best[1] = by;
          best[2] = bz;
        }
      }
  for (int d = 0; d < 3; ++d) {
    t.bd[d] = c->bd_req[d] > 0 ? std::min(c->bd_req[d], std::min(4, t.gn[d])) : best[d];
  }
}

static void finish_tiles(gh_ctx *c) {
  TileGeom &t = c->tg;
  t.cpb = t.bd[0] * t.bd[1] * t.bd[2];
  t.nbrick = 1;
  for (int d = 0; d < 3; ++d) {
    t.nbr[d] = (t.gn[d] + t.bd[d] - 1) / t.bd[d];
    t.nbrick *= t.nbr[d];
  }
  t.nkey = t.nbrick * t.cpb;
  c->ncell = t.nkey;
  // fixed-point unit of the mixed tiles (SMA-TB, LJ): a staged atom lies within half
  // a brick plus its halo cell of the brick centre, plus half the skin (the
  // largest displacement between rebuilds); 1 A of margin
  double ext = 0.0;
  for (int d = 0; d < 3; ++d) ext = std::max(ext, (0.5 * t.bd[d] + 1.0) * c->cell[d]);
  ext += 0.5 * c->skin + 1.0;
  t.fxk = GH_FXK_MAX;
  while (t.fxk > 0 && ext * ldexp(1.0, t.fxk) >= 1073741824.0) --t.fxk;
  t.fxs = ldexp(1.0, t.fxk);
}

static void setup_grid(gh_ctx *c) {
  const double rl = c->rcut + c->skin;
  c->list_mode = 0;
  for (int d = 0; d < 3; ++d) {
    c->nc[d] = 1;
    c->gn[d] = 1;
    c->gofs[d] = 0;
    c->cell[d] = c->L[d];
  }
  c->ncell = 1;
  if (!c->boxed || c->masked) return;
  int nc[3];
  bool all_small = true, pbc_small = false;
  for (int d = 0; d < 3; ++d) {
    nc[d] = (int)(c->L[d] / rl);
    if (nc[d] >= 4) all_small = false;
    if (c->pbc[d] && nc[d] < 3) pbc_small = true;
    if (nc[d] < 1) nc[d] = 1;
  }
  if (all_small || pbc_small) return;
  c->list_mode = 1;
  TileGeom &t = c->tg;
  for (int d = 0; d < 3; ++d) {
    c->nc[d] = nc[d];
    c->cell[d] = c->L[d] / nc[d];
    c->gofs[d] = c->pbc[d] ? 0 : 1;
    c->gn[d] = nc[d] + 2 * c->gofs[d];
    t.nc[d] = c->nc[d];
    t.gn[d] = c->gn[d];
    t.gofs[d] = c->gofs[d];
    t.pbc[d] = c->pbc[d];
    t.L[d] = c->L[d];
    t.one_cell[d] = 1.0 / c->cell[d];
  }
  choose_bricks(c, max_slots(c));
  finish_tiles(c);
}

static int alloc_cells(gh_ctx *c) {
  if (c->ncell > c->cell_alloc) {
    GH_TRY(dalloc(c, &c->cell_count, (size_t)c->ncell));
    GH_TRY(dalloc(c, &c->cell_start, (size_t)c->ncell + 1));
    c->cell_alloc = c->ncell;
  }
  return 0;
}

static int ensure_ready(gh_ctx *c) {
  if (!c->have_atoms) return gh_fail(c, GH_ESTATE, "no atoms: call gh_set_atoms first");
  if (c->pot == GH_POT_NONE) return gh_fail(c, GH_ESTATE, "no potential set");
  if (c->masked && c->pot == GH_POT_SMATB)
    return gh_fail(c, GH_EINVAL, "one-sided interactions are served for pair potentials only");
  if (c->ready) return 0;
  setup_grid(c);
  if (c->boxed) GH_TRY(gh_launch_wrap(c));
  GhStatus h;
  GH_TRY(read_status(c, &h));
  const long long nupd0 = h.nupd;
  int maxnn = 0;
  if (c->list_mode == 1) {
    // sizing: bin, count rows, measure the largest halo; shrink the bricks
    // until the halo fits the LDS budget
    for (int attempt = 0;; ++attempt) {
      GH_TRY(alloc_cells(c));
      GH_TRY(gh_fill_int(c, c->cell_count, 0, c->tg.nkey));
      c->tg.slot_cap = 65535;  // count pass: any halo (the build reads L2 past its LDS)
      c->cap = 0;
      GH_TRY(gh_launch_rebuild(c, true));
      int max_halo = 0, max_brick = 0;
      GH_TRY(gh_tile_occupancy(c, &max_halo, &max_brick));
      GH_TRY(read_status(c, &h));
      int slot_cap = max_halo + std::max(32, max_halo / 8);
      slot_cap = roundup(slot_cap, 32);
      if (slot_cap <= max_slots(c) && !h.tile_overflow) {
        c->tg.slot_cap = slot_cap;
        c->max_brick = max_brick;
        set_force_pass_threads(c);
        maxnn = h.maxnn;
        break;
      }
      int *bd = c->tg.bd;
      int dmax = 0;
      for (int d = 1; d < 3; ++d)
        if (bd[d] > bd[dmax]) dmax = d;
      if (bd[dmax] == 1 || attempt > 12)
        return gh_fail(c, GH_EINVAL,
                       "atom density too high for the LDS halo of one cell (halo of " +
                           std::to_string(max_halo) + " atoms)");
      bd[dmax] -= 1;
      c->bd_req[0] = c->bd_req[1] = c->bd_req[2] = 0;
      finish_tiles(c);
      h.tile_overflow = 0;
      h.halt = 0;
      GH_TRY(write_status(c, &h));
    }
  } else {
    c->cap = 0;
    GH_TRY(gh_launch_rebuild(c, true));
    GH_TRY(read_status(c, &h));
    maxnn = h.maxnn;
  }
  int cap = maxnn + std::max(16, maxnn / 4);
  if (c->cap_req > cap) cap = c->cap_req;
  cap = roundup(std::max(cap, 8), 8);
  if ((size_t)cap * c->ncap > (size_t)1 << 31)
    return gh_fail(c, GH_EINVAL, "neighbour list too large");
  if (c->list_mode == 1) {
    if (c->list) {
      GH_HIP(c, hipFree(c->list));
      c->list = nullptr;
    }
    GH_TRY(dalloc(c, &c->list16, (size_t)(cap + 4 * GH_LIST_PAD_QUADS) * c->ncap));
  } else {
    if (c->list16) {
      GH_HIP(c, hipFree(c->list16));
      c->list16 = nullptr;
    }
    GH_TRY(dalloc(c, &c->list, (size_t)cap * c->ncap));
  }
  c->cap = cap;
  h.overflow = 0;
  h.tile_overflow = 0;
  h.halt = 0;
  h.halt_step = 0;
  h.nupd = nupd0;
  GH_TRY(write_status(c, &h));
  GH_TRY(gh_launch_rebuild(c, true));
  GH_TRY(read_status(c, &h));
  if (h.overflow || h.tile_overflow) {
    const bool rows = h.overflow != 0;
    h.halt = 0;
    h.halt_step = 0;
    h.overflow = 0;
    h.tile_overflow = 0;
    GH_TRY(write_status(c, &h));
    return gh_fail(c, GH_EOVERFLOW,
                   rows ? "neighbour list overflow at sizing" : "LDS halo overflow at sizing");
  }
  c->ready = true;
  c->pos_dirty = false;
  return 0;
}

// Reports a stop the device raised (GhStatus::halt) and re-arms the status:
// the halt, the outside flag and the overflow flags are cleared on the
// device; an overflow also makes the next call re-size the list or tiles.
static int report_halt(gh_ctx *c, GhStatus &h) {
  const int code = h.halt;
  const long long at = h.halt_step;
  const int maxnn = h.maxnn;
  h.halt = 0;
  h.halt_step = 0;
  h.outside = 0;
  h.overflow = 0;
  h.tile_overflow = 0;
  GH_TRY(write_status(c, &h));
  const std::string where = " at step " + std::to_string(at);
  switch (code) {
    case GH_HALT_OVERFLOW:
      c->ready = false;
      c->cap_req = maxnn + std::max(16, maxnn / 4);
      return gh_fail(c, GH_EOVERFLOW,
                     "neighbour list row capacity exceeded (" + std::to_string(maxnn) + " > " +
                         std::to_string(c->cap) + ")" + where +
                         "; the state is that step's start; the next call re-sizes the list");
    case GH_HALT_TILE:
      c->ready = false;
      return gh_fail(c, GH_EOVERFLOW,
                     "a brick's halo outgrew the LDS slots sized at setup" + where +
                         "; the state is that step's start; the next call re-sizes the tiles");
    case GH_HALT_RANGE:
      return gh_fail(c, GH_ERANGE,
                     "an atom left the cell grid of a non-periodic box axis" + where);
    default:
      return gh_fail(c, GH_ESTATE, "a rebuild was due but not enqueued" + where);
  }
}

static int check_status(gh_ctx *c) {
  GhStatus h;
  GH_TRY(read_status(c, &h));
  if (h.halt) return report_halt(c, h);
  if (h.overflow) {
    c->ready = false;
    c->cap_req = h.maxnn + std::max(16, h.maxnn / 4);
    return gh_fail(c, GH_EOVERFLOW,
                   "neighbour list row capacity exceeded (" + std::to_string(h.maxnn) + " > " +
                       std::to_string(c->cap) + "); results of this call are invalid; the "
                       "next call re-sizes the list");
  }
  if (h.tile_overflow) {
    c->ready = false;
    return gh_fail(c, GH_EOVERFLOW,
                   "a brick's halo outgrew the LDS slots sized at setup; results of this call "
                   "are invalid; the next call re-sizes the tiles");
  }
  if (h.outside) {
    h.outside = 0;
    GH_TRY(write_status(c, &h));
    return gh_fail(c, GH_ERANGE, "an atom left the cell grid of a non-periodic box axis");
  }
  return 0;
}

// ---------------------------------------------------------------------------
// timing pool: 8 events per step
//   0 start, 1 after stepa, 2 after rebuild, 3 after pass 1, 4 after pass 2,
//   5 after stepb
// ---------------------------------------------------------------------------
static const int EV_PER_STEP = 6;
static const int EV_STEPS = 128;

static int timing_flush(gh_ctx *c) {
  if (!c->timing || c->evused == 0) return 0;
  GH_HIP(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->evused; ++k) {
    hipEvent_t *e = &c->evpool[(size_t)k * EV_PER_STEP];
    float t;
    GH_HIP(c, hipEventElapsedTime(&t, e[2], e[3]));
    c->t_ms[0] += t;
    c->t_cnt[0] += 1;
    GH_HIP(c, hipEventElapsedTime(&t, e[3], e[4]));
    c->t_ms[1] += t;
    c->t_cnt[1] += 1;
    float ta, tb;
    GH_HIP(c, hipEventElapsedTime(&ta, e[0], e[1]));
    GH_HIP(c, hipEventElapsedTime(&tb, e[4], e[5]));
    c->t_ms[2] += ta + tb;
    c->t_cnt[2] += 1;
    GH_HIP(c, hipEventElapsedTime(&t, e[1], e[2]));
    c->t_ms[3] += t;
    c->t_cnt[3] += 1;
  }
  c->evused = 0;
  return 0;
}

static hipEvent_t ev_at(gh_ctx *c, int k) {
  if (!c->timing) return nullptr;
  return c->evpool[(size_t)c->evused * EV_PER_STEP + k];
}

static int rec(gh_ctx *c, int k) {
  if (!c->timing) return 0;
  GH_HIP(c, hipEventRecord(ev_at(c, k), c->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
// The nt x nt table of a tabulated Lennard-Jones form: validates symmetry,
// fills each pair with fill(entry, a, b), uploads, and selects the form.
template <class FILL>
static int set_pair_table(gh_ctx *c, int pot, int nt, const double *eps, const double *sigma,
                          const char *who, FILL &&fill) {
  if (!c || nt <= 0 || nt > GH_MAXT || !eps || !sigma)
    return gh_fail(c, GH_EINVAL, std::string(who) + ": bad arguments");
  PljDev h[GH_MAXT * GH_MAXT];
  memset(h, 0, sizeof(h));
  double rmax = 0.0;
  for (int a = 0; a < nt; ++a)
    for (int b = 0; b < nt; ++b) {
      const double e = eps[a * nt + b], sg = sigma[a * nt + b];
      if (!(sg > 0) || e != eps[b * nt + a] || sg != sigma[b * nt + a])
        return gh_fail(c, GH_EINVAL,
                       std::string(who) + ": eps and sigma must be symmetric, sigma > 0");
      PljDev &d = h[a * GH_MAXT + b];
      d.e = e;
      d.s = sg;
      if (!fill(d, a, b)) return gh_fail(c, GH_EINVAL, std::string(who) + ": bad arguments");
      d.rc2 = d.rc * d.rc;
      d.ef = (float)e;
      d.s2f = (float)(sg * sg);
      d.prcf = (float)d.prc;
      d.frcf = (float)d.frc;
      d.rsf = (float)d.rs;
      d.sa4f = (float)d.sa4;
      d.sa8f = (float)d.sa8;
      d.rcf = (float)d.rc;
      rmax = std::max(rmax, d.rc);
    }
  GH_HIP(c, hipSetDevice(c->device));
  GH_TRY(dalloc(c, &c->pljd, GH_MAXT * GH_MAXT));
  GH_HIP(c, hipMemcpyAsync(c->pljd, h, sizeof(h), hipMemcpyHostToDevice, c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  c->nt = nt;
  c->lj_eps = 0.0;
  c->lj_sigma = 0.0;
  c->lj_rc = rmax;
  c->pot = pot;
  c->rcut = rmax;
  c->ready = false;
  return 0;
}

static double pow6(double x) {  // x**6 as the reference's compiled integer power
  const double x2 = x * x, x4 = x2 * x2;
  return x4 * x2;
}

extern "C" {

int gh_create(gh_ctx **out, int device, int precision) {
  if (!out) return GH_EINVAL;
  *out = nullptr;
  if (precision != GH_FP64 && precision != GH_MIXED) return GH_EINVAL;
  gh_ctx *c = new (std::nothrow) gh_ctx();
  if (!c) return GH_EINVAL;
  c->device = device;
  c->precision = precision;
  // list-build kernel (both write the same rows): "rows" one thread per
  // atom, "coop" a wavefront per atom wherever its LDS fits; unset, the
  // choice of coop_args
  const char *lb = getenv("GEMS_HIP_LIST_BUILD");
  c->list_rows = !lb ? 0 : strcmp(lb, "rows") == 0 ? 1 : strcmp(lb, "coop") == 0 ? 2 : 0;
  // force-pass workgroups per brick, 1..4 (the autotuner then keeps it)
  if (const char *ck = getenv("GEMS_HIP_CHUNKS")) c->chunks_env = std::min(4, std::max(0, atoi(ck)));
  c->chunks_req = c->chunks_env;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
    delete c;
    return GH_EHIP;
  }
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void **)&c->status, sizeof(GhStatus)) != hipSuccess ||
      hipHostMalloc((void **)&c->red_h, sizeof(double) * 16, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&c->prog_h, sizeof(GhProgress), hipHostMallocCoherent) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&c->prog_d, c->prog_h, 0) != hipSuccess) {
    gh_destroy(c);
    return GH_EHIP;
  }
  memset(c->prog_h, 0, sizeof(GhProgress));
  GhStatus h;
  memset(&h, 0, sizeof(h));
  if (hipMemcpyAsync(c->status, &h, sizeof(h), hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    gh_destroy(c);
    return GH_EHIP;
  }
  c->lcg.seed(123456789);
  *out = c;
  return 0;
}

void gh_destroy(gh_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void *dev[] = {c->pos, c->pos_t, c->pos_next, c->vel, c->vel_t, c->acc, c->acc_t, c->frc, c->epa,
                 c->rho, c->eband, c->eband32, c->type, c->type_t, c->orig, c->orig_t,
                 c->xold, c->posv, c->velv, c->gauss, c->cellof, c->cellid, c->cell_count,
                 c->cell_start, c->perm, c->nn, c->list, c->tbd, c->tbf, c->status,
                 c->red, c->list16, c->xin, c->chain_dev, c->pljd, c->mfeel, c->mfelt};
  for (void *p : dev)
    if (p) (void)hipFree(p);
  if (c->chain_pin) (void)hipHostFree(c->chain_pin);
  if (c->gauss_h) (void)hipHostFree(c->gauss_h);
  if (c->red_h) (void)hipHostFree(c->red_h);
  if (c->prog_h) (void)hipHostFree(c->prog_h);
  for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char *gh_last_error(const gh_ctx *c) { return c ? c->err.c_str() : "null context"; }

int gh_set_tile(gh_ctx *c, const int *bd) {
  if (!c) return GH_EINVAL;
  for (int d = 0; d < 3; ++d)
    if (bd && (bd[d] < 0 || bd[d] > 4)) return gh_fail(c, GH_EINVAL, "brick edges must be 0..4");
  for (int d = 0; d < 3; ++d) c->bd_req[d] = bd ? bd[d] : 0;
  c->tpb_mean = 0;
  c->chunks_req = c->chunks_env;
  c->ready = false;
  return 0;
}

int gh_autotune(gh_ctx *c, int reps, int *bd_out, double *us_out) {
  if (!c || reps <= 0) return GH_EINVAL;
  GH_HIP(c, hipSetDevice(c->device));
  GH_TRY(ensure_ready(c));
  if (c->list_mode != 1) {
    if (bd_out) bd_out[0] = bd_out[1] = bd_out[2] = 0;
    if (us_out) *us_out = 0.0;
    return 0;
  }
  GhStatus h0;
  GH_TRY(read_status(c, &h0));  // the trial rebuilds are not counted in nupd
  double cells = 1.0;
  for (int d = 0; d < 3; ++d) cells *= c->nc[d];
  const double apc = c->n / cells;
  // Expected rebuilds per step: one falls due when an atom has moved half the
  // skin, so at most 2 v_max dt / skin (ballistic motion of the fastest
  // atom).  Each shape is scored by its force evaluation plus that share of
  // its rebuild chain; before velocities and a time step are set the rate is
  // 0 and the force evaluation alone decides.
  double rate = 0.0;
  if (c->dt > 0.0 && c->skin > 0.0) {
    std::vector<double> v(3 * (size_t)c->ncap);
    GH_HIP(c, hipMemcpyAsync(v.data(), c->vel, sizeof(double) * v.size(), hipMemcpyDeviceToHost,
                             c->stream));
    GH_HIP(c, hipStreamSynchronize(c->stream));
    double v2 = 0.0;
    for (int i = 0; i < c->n; ++i) {
      const double a = v[i], b = v[c->ncap + i], d = v[2 * (size_t)c->ncap + i];
      v2 = std::max(v2, a * a + b * b + d * d);
    }
    rate = std::min(1.0, 2.0 * sqrt(v2) * c->dt / c->skin);
  }
  hipEvent_t e0, e1;
  GH_HIP(c, hipEventCreate(&e0));
  GH_HIP(c, hipEventCreate(&e1));
  int best[3] = {0, 0, 0}, best_mean = 0, best_chunks = 1;
  double best_us = 1e30, best_score = 1e30;
  for (int bz = 1; bz <= 4; ++bz)
    for (int by = bz; by <= 4; ++by)
      for (int bx = by; bx <= 4; ++bx) {
        const int b[3] = {bx, by, bz};
        bool ok = true;
        for (int d = 0; d < 3; ++d) ok = ok && b[d] <= c->gn[d];
        const double halo = apc * (bx + 2) * (by + 2) * (bz + 2);
        // bricks under 4 cells stage >8 halo atoms per own atom: never faster
        if (!ok || halo * 1.15 > max_slots(c) || bx * by * bz < 4) continue;
        for (int d = 0; d < 3; ++d) c->bd_req[d] = b[d];
        c->ready = false;
        if (ensure_ready(c) != 0) continue;
        bool same = true;  // the sizing may have shrunk the request
        for (int d = 0; d < 3; ++d) same = same && c->tg.bd[d] == b[d];
        if (!same) continue;
        double chain_us = 0.0;
        if (rate > 0.0) {
          GH_TRY(gh_request_rebuild(c));
          GH_HIP(c, hipEventRecord(e0, c->stream));
          GH_TRY(gh_launch_rebuild(c, false));
          GH_HIP(c, hipEventRecord(e1, c->stream));
          GH_HIP(c, hipEventSynchronize(e1));
          float ms = 0.f;
          GH_HIP(c, hipEventElapsedTime(&ms, e0, e1));
          if (check_status(c) != 0) continue;
          chain_us = 1e3 * ms;
        }
        // workgroups per brick: 1, and 2 and 4 where the bricks are too few
        // to fill the chip (the grid stays within two workgroups per CU);
        // GEMS_HIP_CHUNKS fixes it
        for (int ki = 0; ki < 3; ++ki) {
          const int k = c->chunks_env ? c->chunks_env : 1 << ki;
          if (c->chunks_env && ki > 0) break;
          if (k > 1 && !c->chunks_env && c->tg.nbrick * k > 512) continue;
          c->chunks_req = k;
          set_force_pass_threads(c);
          const int cand[2] = {c->tg.tpb_max, c->tg.tpb_avg};
          for (int m = 0; m < 2; ++m) {
            if (m == 1 && cand[1] == cand[0]) break;
            c->tg.tpb = cand[m];
            GH_TRY(gh_launch_forces(c, nullptr));
            GH_HIP(c, hipEventRecord(e0, c->stream));
            for (int r = 0; r < reps; ++r) GH_TRY(gh_launch_forces(c, nullptr));
            GH_HIP(c, hipEventRecord(e1, c->stream));
            GH_HIP(c, hipEventSynchronize(e1));
            float ms = 0.f;
            GH_HIP(c, hipEventElapsedTime(&ms, e0, e1));
            if (check_status(c) != 0) break;
            const double us = 1e3 * ms / reps;
            if (us + rate * chain_us < best_score) {
              best_score = us + rate * chain_us;
              best_us = us;
              best_mean = m;
              best_chunks = k;
              for (int d = 0; d < 3; ++d) best[d] = b[d];
            }
          }
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (int d = 0; d < 3; ++d) c->bd_req[d] = best[d];
  c->tpb_mean = best_mean;
  c->chunks_req = best_chunks;
  c->ready = false;
  c->err.clear();
  GH_TRY(ensure_ready(c));
  GH_TRY(gh_launch_forces(c, nullptr));  // forces of the kept shape
  GH_TRY(check_status(c));
  GhStatus h1;
  GH_TRY(read_status(c, &h1));
  h1.nupd = h0.nupd;
  GH_TRY(write_status(c, &h1));
  if (bd_out)
    for (int d = 0; d < 3; ++d) bd_out[d] = c->tg.bd[d];
  if (us_out) *us_out = best_us;
  return 0;
}

int gh_set_precision(gh_ctx *c, int precision) {
  if (!c) return GH_EINVAL;
  if (precision != GH_FP64 && precision != GH_MIXED)
    return gh_fail(c, GH_EINVAL, "precision must be GH_FP64 or GH_MIXED");
  if (precision != c->precision) c->ready = false;  // LDS budgets depend on it
  c->precision = precision;
  return 0;
}

int gh_set_box(gh_ctx *c, const double *L, const int *pbc) {
  if (!c) return GH_EINVAL;
  if (L) {
    for (int d = 0; d < 3; ++d)
      if (!(L[d] > 0.0)) return gh_fail(c, GH_EINVAL, "box lengths must be positive");
    c->boxed = 1;
    for (int d = 0; d < 3; ++d) {
      c->L[d] = L[d];
      c->pbc[d] = pbc ? (pbc[d] != 0) : 0;
    }
  } else {
    c->boxed = 0;
    for (int d = 0; d < 3; ++d) {
      c->L[d] = 1.0e6;  // the reference's unboxed box (Program_Types.F90:69)
      c->pbc[d] = 0;
    }
  }
  c->ready = false;
  c->pos_dirty = true;
  return 0;
}

int gh_set_pair_masks(gh_ctx *c, const int *feel, const int *felt) {
  if (!c) return GH_EINVAL;
  if ((feel == nullptr) != (felt == nullptr))
    return gh_fail(c, GH_EINVAL, "gh_set_pair_masks: give both masks or neither");
  if (!c->have_atoms) return gh_fail(c, GH_ESTATE, "gh_set_pair_masks: call gh_set_atoms first");
  GH_HIP(c, hipSetDevice(c->device));
  c->ready = false;
  if (!feel) {
    c->masked = false;
    return 0;
  }
  if (!c->mfeel) {
    GH_TRY(dalloc(c, &c->mfeel, c->ncap));
    GH_TRY(dalloc(c, &c->mfelt, c->ncap));
  }
  GH_HIP(c, hipMemcpyAsync(c->mfeel, feel, sizeof(int) * c->n, hipMemcpyHostToDevice, c->stream));
  GH_HIP(c, hipMemcpyAsync(c->mfelt, felt, sizeof(int) * c->n, hipMemcpyHostToDevice, c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  c->masked = true;
  return 0;
}

int gh_set_images(gh_ctx *c, int ghost) {
  if (!c || (ghost != 0 && ghost != 1))
    return gh_fail(c, GH_EINVAL, "gh_set_images: ghost must be 0 or 1");
  c->ghost = ghost;
  c->ready = false;
  return 0;
}

int gh_set_atoms(gh_ctx *c, int n, const double *x, const int *type) {
  if (!c || n <= 0 || !x) return gh_fail(c, GH_EINVAL, "gh_set_atoms: bad arguments");
  if (n > (int)GH_JMASK_HOST) return gh_fail(c, GH_EINVAL, "too many atoms (max 134217727)");
  GH_HIP(c, hipSetDevice(c->device));
  const int ncap = roundup(n, 256);
  c->masked = false;  // new atoms: masks, if any, are set again after this call
  if (ncap != c->ncap) {
    c->ncap = ncap;
    GH_TRY(dalloc(c, &c->mfeel, 0));
    GH_TRY(dalloc(c, &c->mfelt, 0));
    GH_TRY(dalloc(c, &c->pos, ncap));
    GH_TRY(dalloc(c, &c->pos_t, ncap));
    GH_TRY(dalloc(c, &c->pos_next, ncap));
    GH_TRY(dalloc(c, &c->vel, 3 * (size_t)ncap));
    GH_TRY(dalloc(c, &c->vel_t, 3 * (size_t)ncap));
    GH_TRY(dalloc(c, &c->acc, 3 * (size_t)ncap));
    GH_TRY(dalloc(c, &c->acc_t, 3 * (size_t)ncap));
    GH_TRY(dalloc(c, &c->frc, 3 * (size_t)ncap));
    GH_TRY(dalloc(c, &c->epa, ncap));
    GH_TRY(dalloc(c, &c->rho, ncap));
    GH_TRY(dalloc(c, &c->eband, ncap));
    GH_TRY(dalloc(c, &c->eband32, ncap));
    GH_TRY(dalloc(c, &c->type, ncap));
    GH_TRY(dalloc(c, &c->type_t, ncap));
    GH_TRY(dalloc(c, &c->orig, ncap));
    GH_TRY(dalloc(c, &c->orig_t, ncap));
    GH_TRY(dalloc(c, &c->xold, 3 * (size_t)ncap));
    GH_TRY(dalloc(c, &c->cellof, ncap));
    GH_TRY(dalloc(c, &c->cellid, ncap));
    GH_TRY(dalloc(c, &c->perm, ncap));
    GH_TRY(dalloc(c, &c->nn, ncap));
    c->nred = std::min(1024, std::max(1, (ncap + 255) / 256));
    GH_TRY(dalloc(c, &c->red, (size_t)c->nred * 10 + 16));
    if (c->gauss) {
      GH_HIP(c, hipFree(c->gauss));
      c->gauss = nullptr;
    }
    if (c->xin) {
      GH_HIP(c, hipFree(c->xin));
      c->xin = nullptr;
    }
    if (c->gauss_h) {
      GH_HIP(c, hipHostFree(c->gauss_h));
      c->gauss_h = nullptr;
    }
    if (c->list) {
      GH_HIP(c, hipFree(c->list));
      c->list = nullptr;
    }
    c->cap = 0;
  }
  c->n = n;
  std::vector<double4> hp(ncap);
  c->type_h.assign(n, 0);
  std::vector<int> ho(ncap), ht(ncap, 0);
  for (int i = 0; i < ncap; ++i) {
    if (i < n) {
      double v[3];
      for (int d = 0; d < 3; ++d) {
        v[d] = x[3 * i + d];
        // do_pbc_1 (Set_Properties.f90:444-470): periodic axes start wrapped
        if (c->boxed && c->pbc[d]) {
          if (v[d] >= c->L[d]) v[d] -= c->L[d];
          else if (v[d] < 0.0) v[d] += c->L[d];
          if (v[d] < 0.0 || v[d] >= c->L[d])
            return gh_fail(c, GH_EINVAL, "atom more than one box length outside a periodic box");
        }
      }
      hp[i] = make_double4(v[0], v[1], v[2], 0.0);
      int t = type ? type[i] : 0;
      if (t < 0 || t >= GH_MAXT) return gh_fail(c, GH_EINVAL, "atom type out of range 0..7");
      ht[i] = t;
      c->type_h[i] = t;
    } else {
      hp[i] = make_double4(0, 0, 0, 0);
    }
    ho[i] = i;
  }
  // uploads and clears ordered on the context's stream, then drained (the host
  // vectors die on return)
  hipStream_t st = c->stream;
  GH_HIP(c, hipMemcpyAsync(c->pos, hp.data(), sizeof(double4) * ncap, hipMemcpyHostToDevice, st));
  GH_HIP(c, hipMemcpyAsync(c->type, ht.data(), sizeof(int) * ncap, hipMemcpyHostToDevice, st));
  GH_HIP(c, hipMemcpyAsync(c->orig, ho.data(), sizeof(int) * ncap, hipMemcpyHostToDevice, st));
  GH_HIP(c, hipMemsetAsync(c->vel, 0, sizeof(double) * 3 * ncap, st));
  GH_HIP(c, hipMemsetAsync(c->acc, 0, sizeof(double) * 3 * ncap, st));
  GH_HIP(c, hipMemsetAsync(c->frc, 0, sizeof(double) * 3 * ncap, st));
  GH_HIP(c, hipMemsetAsync(c->xold, 0, sizeof(double) * 3 * ncap, st));
  GH_HIP(c, hipMemsetAsync(c->epa, 0, sizeof(double) * ncap, st));
  GH_HIP(c, hipStreamSynchronize(st));
  c->have_atoms = true;
  c->ready = false;
  c->pos_dirty = true;
  c->first_force = true;
  return 0;
}

int gh_set_masses(gh_ctx *c, int nt, const double *mass) {
  if (!c || nt <= 0 || nt > GH_MAXT || !mass) return gh_fail(c, GH_EINVAL, "gh_set_masses");
  for (int t = 0; t < nt; ++t) {
    if (!(mass[t] > 0.0)) return gh_fail(c, GH_EINVAL, "masses must be positive");
    c->mass_t[t] = mass[t];
  }
  return 0;
}

// Host copies in caller order <-> device arrays in sorted order.
static int fetch_orig(gh_ctx *c, std::vector<int> &orig) {
  orig.resize(c->n);
  GH_HIP(c, hipMemcpyAsync(orig.data(), c->orig, sizeof(int) * c->n, hipMemcpyDeviceToHost,
                           c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

static int fetch_soa3(gh_ctx *c, const double *d, const std::vector<int> &orig, double *out) {
  std::vector<double> h(3 * (size_t)c->ncap);
  GH_HIP(c, hipMemcpyAsync(h.data(), d, sizeof(double) * 3 * c->ncap, hipMemcpyDeviceToHost,
                           c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  for (int s = 0; s < c->n; ++s)
    for (int k = 0; k < 3; ++k) out[3 * (size_t)orig[s] + k] = h[(size_t)k * c->ncap + s];
  return 0;
}

static int put_soa3(gh_ctx *c, double *d, const std::vector<int> &orig, const double *in) {
  std::vector<double> h(3 * (size_t)c->ncap, 0.0);
  for (int s = 0; s < c->n; ++s)
    for (int k = 0; k < 3; ++k) h[(size_t)k * c->ncap + s] = in ? in[3 * (size_t)orig[s] + k] : 0.0;
  GH_HIP(c, hipMemcpyAsync(d, h.data(), sizeof(double) * 3 * c->ncap, hipMemcpyHostToDevice,
                           c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gh_set_positions(gh_ctx *c, const double *x) {
  if (!c || !x) return gh_fail(c, GH_EINVAL, "gh_set_positions: bad arguments");
  if (!c->have_atoms) return gh_fail(c, GH_ESTATE, "no atoms: call gh_set_atoms first");
  GH_HIP(c, hipSetDevice(c->device));
  if (!c->xin) GH_TRY(dalloc(c, &c->xin, 3 * (size_t)c->ncap));
  GH_HIP(c, hipMemcpyAsync(c->xin, x, sizeof(double) * 3 * (size_t)c->n, hipMemcpyHostToDevice,
                           c->stream));
  return gh_launch_set_positions(c, c->xin);
}

int gh_set_velocities(gh_ctx *c, const double *v) {
  if (!c || !c->have_atoms) return gh_fail(c, GH_ESTATE, "no atoms");
  std::vector<int> orig;
  GH_TRY(fetch_orig(c, orig));
  return put_soa3(c, c->vel, orig, v);
}

int gh_set_accelerations(gh_ctx *c, const double *a) {
  if (!c || !c->have_atoms) return gh_fail(c, GH_ESTATE, "no atoms");
  std::vector<int> orig;
  GH_TRY(fetch_orig(c, orig));
  return put_soa3(c, c->acc, orig, a);
}

int gh_set_smatb(gh_ctx *c, int nt, const double *A, const double *xi, const double *p,
                 const double *q, const double *r0, const double *rci, const double *rce) {
  if (!c || nt <= 0 || nt > GH_MAXT || !A || !xi || !p || !q || !r0 || !rci || !rce)
    return gh_fail(c, GH_EINVAL, "gh_set_smatb: bad arguments");
  double rmax = 0.0;
  for (int a = 0; a < nt; ++a)
    for (int b = 0; b < nt; ++b) {
      const int k = a * nt + b;
      if (!(rce[k] > rci[k]) || !(r0[k] > 0) || !(A[k] > 0) || !(xi[k] > 0))
        return gh_fail(c, GH_EINVAL, "SMA-TB parameters need A, xi, r0 > 0 and rce > rci");
      GhTbPair &t = c->tb[a * GH_MAXT + b];
      t.A = A[k];
      t.xi = xi[k];
      t.p = p[k];
      t.q = q[k];
      t.r0 = r0[k];
      t.rci = rci[k];
      t.rce = rce[k];
      gh_tb_coeffs(t);
      rmax = std::max(rmax, rce[k]);
    }
  c->nt = nt;
  c->pot = GH_POT_SMATB;
  c->rcut = rmax;
  GH_TRY(gh_upload_tb(c));
  c->ready = false;
  return 0;
}

int gh_set_lj(gh_ctx *c, double eps, double sigma, double rc) {
  if (!c || !(sigma > 0) || !(rc > 0)) return gh_fail(c, GH_EINVAL, "gh_set_lj: bad arguments");
  c->lj_eps = eps;
  c->lj_sigma = sigma;
  c->lj_rc = rc;
  c->pot = GH_POT_LJ;
  c->rcut = rc;
  c->ready = false;
  return 0;
}

int gh_set_plj(gh_ctx *c, int nt, const double *eps, const double *sigma, double rc) {
  if (!(rc > 0)) return gh_fail(c, GH_EINVAL, "gh_set_plj: bad arguments");
  return set_pair_table(c, GH_POT_PLJ, nt, eps, sigma, "gh_set_plj", [&](PljDev &d, int, int) {
    // plj_set (Pairs.F90:282-291)
    d.prc = 4.0 * d.e * (pow(d.s / rc, 12) - pow(d.s / rc, 6));
    d.frc = 24.0 * d.e * (2.0 * (pow(d.s, 12) / pow(rc, 13)) - (pow(d.s, 6) / pow(rc, 7)));
    d.rc = rc;
    return true;
  });
}

int gh_set_slj(gh_ctx *c, int nt, const double *eps, const double *sigma, const double *rs) {
  if (!rs) return gh_fail(c, GH_EINVAL, "gh_set_slj: bad arguments");
  return set_pair_table(c, GH_POT_SLJ, nt, eps, sigma, "gh_set_slj", [&](PljDev &d, int a, int b) {
    // slj_set (Pairs.F90:374-401)
    const double r = rs[a * nt + b];
    if (!(r > 0) || r != rs[b * nt + a]) return false;
    const double s_dr = pow6(d.s / r);
    const double urs = 4.0 * d.e * (s_dr * s_dr - s_dr);
    const double dudrs = -d.e * (48.0 * s_dr - 24.0) * s_dr / r;
    d.rs = r;
    d.rc = r - 32.0 * urs / (11.0 * dudrs);
    const double dd = d.rc - r, d2 = dd * dd, d4 = d2 * d2, d8 = d4 * d4;
    d.sa4 = (8.0 * urs + dd * dudrs) / (4 * d4);
    d.sa8 = -(4.0 * urs + dd * dudrs) / (4 * d8);
    return d.rc > r;
  });
}

int gh_set_wca(gh_ctx *c, int nt, const double *eps, const double *sigma) {
  return set_pair_table(c, GH_POT_WCA, nt, eps, sigma, "gh_set_wca", [&](PljDev &d, int, int) {
    // wca_set (Pairs.F90:495-506): 2._dp**(1./6.), the exponent a default-real literal
    d.rc = pow(2.0, (double)(1.0f / 6.0f)) * d.s;
    return true;
  });
}

int gh_set_sm1(gh_ctx *c, double r1, double r2, double e) {
  if (!c) return GH_EINVAL;
  if (!(r1 >= 0.0) || !(r2 > r1)) return gh_fail(c, GH_EINVAL, "gh_set_sm1: need 0 <= r1 < r2");
  double eps[GH_MAXT * GH_MAXT], one[GH_MAXT * GH_MAXT];
  for (int k = 0; k < GH_MAXT * GH_MAXT; ++k) {
    eps[k] = e;
    one[k] = 1.0;
  }
  // the same entry for every pair of the eight types, whatever types the atoms carry
  return set_pair_table(c, GH_POT_SM1, GH_MAXT, eps, one, "gh_set_sm1", [&](PljDev &d, int, int) {
    d.rs = r1;
    d.rc = r2;
    d.sa8 = (r1 + r2) * 0.5;
    return true;
  });
}

int gh_set_skin(gh_ctx *c, double skin) {
  if (!c || !(skin >= 0.0)) return gh_fail(c, GH_EINVAL, "skin must be >= 0");
  c->skin = skin;
  c->ready = false;
  return 0;
}

static int interact(gh_ctx *c) {
  GH_TRY(ensure_ready(c));
  if (c->pos_dirty) {
    GH_TRY(gh_launch_rebuild(c, true));
    c->pos_dirty = false;
  } else {
    GH_TRY(gh_launch_rebuild(c, false));
  }
  GH_TRY(gh_launch_forces(c, nullptr));
  c->first_force = false;
  return 0;
}

int gh_compute(gh_ctx *c, double *F, double *e_atom, double *e_tot) {
  if (!c) return GH_EINVAL;
  GH_HIP(c, hipSetDevice(c->device));
  GH_TRY(interact(c));
  GH_TRY(check_status(c));
  std::vector<int> orig;
  GH_TRY(fetch_orig(c, orig));
  if (F) GH_TRY(fetch_soa3(c, c->frc, orig, F));
  if (e_atom || e_tot) {
    std::vector<double> h(c->n);
    GH_HIP(c, hipMemcpyAsync(h.data(), c->epa, sizeof(double) * c->n, hipMemcpyDeviceToHost,
                             c->stream));
    GH_HIP(c, hipStreamSynchronize(c->stream));
    if (e_atom)
      for (int s = 0; s < c->n; ++s) e_atom[orig[s]] = h[s];
    if (e_tot) {
      double ep, ek;
      GH_TRY(gh_launch_energies(c, &ep, &ek));
      *e_tot = ep;
    }
  }
  return 0;
}

int gh_neighbors(gh_ctx *c, int *nn_out, int *idx, int cap) {
  if (!c || !nn_out || !idx || cap <= 0) return gh_fail(c, GH_EINVAL, "gh_neighbors: bad arguments");
  if (!c->ready) GH_TRY(interact(c));
  std::vector<int> orig;
  GH_TRY(fetch_orig(c, orig));
  std::vector<int> nn(c->n);
  std::vector<unsigned> lst((size_t)c->cap * c->ncap);
  if (c->list_mode == 1) {
    int *tmp = nullptr;
    GH_HIP(c, hipMalloc((void **)&tmp, sizeof(int) * (size_t)c->cap * c->ncap));
    int rc2 = gh_tiled_to_global(c, tmp);
    if (rc2 == 0) {
      GH_HIP(c, hipMemcpyAsync(lst.data(), tmp, sizeof(int) * lst.size(), hipMemcpyDeviceToHost,
                               c->stream));
      GH_HIP(c, hipStreamSynchronize(c->stream));
    }
    GH_HIP(c, hipFree(tmp));
    if (rc2) return rc2;
  } else {
    GH_HIP(c, hipMemcpyAsync(lst.data(), c->list, sizeof(unsigned) * lst.size(),
                             hipMemcpyDeviceToHost, c->stream));
  }
  GH_HIP(c, hipMemcpyAsync(nn.data(), c->nn, sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  int rc = 0;
  std::vector<int> row;
  for (int s = 0; s < c->n; ++s) {
    row.clear();
    for (int k = 0; k < nn[s]; ++k) row.push_back(orig[lst[(size_t)k * c->ncap + s] & 0x07FFFFFFu]);
    std::sort(row.begin(), row.end());
    const int i = orig[s];
    const int m = std::min((int)row.size(), cap);
    if ((int)row.size() > cap) rc = GH_EOVERFLOW;
    nn_out[i] = m;
    for (int k = 0; k < m; ++k) idx[(size_t)i * cap + k] = row[k];
  }
  if (rc) return gh_fail(c, rc, "gh_neighbors: cap too small");
  return 0;
}

int gh_prng_lcg(gh_ctx *c, int seed) {
  if (!c) return GH_EINVAL;
  c->lcg.seed(seed);
  return 0;
}

int gh_get_lcg_state(gh_ctx *c, int state[3]) {
  if (!c || !state) return GH_EINVAL;
  state[0] = c->lcg.idum;
  state[1] = c->lcg.iy;
  state[2] = c->lcg.ix;
  return 0;
}

int gh_set_lcg_state(gh_ctx *c, const int state[3]) {
  if (!c || !state) return GH_EINVAL;
  c->lcg.idum = state[0];
  c->lcg.iy = state[1];
  c->lcg.ix = state[2];
  c->lcg.stored = false;
  return 0;
}

// set_gdist + set_scal_vel + CM removal (Set_Properties.f90:301-377).
int gh_set_gdist(gh_ctx *c, double T) {
  if (!c || !c->have_atoms) return gh_fail(c, GH_ESTATE, "no atoms");
  const int n = c->n;
  std::vector<int> orig;
  GH_TRY(fetch_orig(c, orig));
  std::vector<double> v(3 * (size_t)n);
  GH_TRY(fetch_soa3(c, c->vel, orig, v.data()));
  std::vector<double> m(n);
  for (int i = 0; i < n; ++i) m[i] = c->mass_t[c->type_h[i]];
  auto cmv = [&](double cm[3]) {
    double M = 0.0;
    for (int i = 0; i < n; ++i) M = M + m[i];
    if (n == 1) {
      for (int d = 0; d < 3; ++d) cm[d] = v[d];
      return;
    }
    cm[0] = cm[1] = cm[2] = 0.0;
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) cm[d] = cm[d] + m[i] * v[3 * (size_t)i + d];
    for (int d = 0; d < 3; ++d) cm[d] = cm[d] / M;
  };
  const double kB_ui = ghc::kB_ui();
  double cm[3];
  cmv(cm);
  const double factor = kB_ui * T;
  for (int i = 0; i < n; ++i) {
    const double vel_med = sqrt(factor * (1.0 / m[i]));
    for (int d = 0; d < 3; ++d) {
      double r1;
      c->lcg.rang(&r1, nullptr);
      v[3 * (size_t)i + d] = vel_med * r1 + cm[d];
    }
  }
  cmv(cm);
  double ek = 0.0;
  for (int i = 0; i < n; ++i) {
    double w[3];
    for (int d = 0; d < 3; ++d) w[d] = v[3 * (size_t)i + d] - cm[d];
    ek = ek + (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) * m[i];
  }
  ek = ek * 0.5;
  const double temp = 2.0 * ek / (kB_ui * (3 * n));
  if (temp != 0.0) {
    const double f = sqrt(T / temp);
    for (auto &x : v) x = x * f;
  }
  cmv(cm);
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) v[3 * (size_t)i + d] = v[3 * (size_t)i + d] + (-cm[d]);
  return put_soa3(c, c->vel, orig, v.data());
}

int gh_integrator_vv(gh_ctx *c, double dt) {
  if (!c || !(dt != 0.0)) return gh_fail(c, GH_EINVAL, "dt must be non-zero");
  c->integ = GH_INT_VV;
  c->dt = dt;
  return 0;
}

int gh_integrator_ermak(gh_ctx *c, double dt, double T, double gamma, int rng,
                        uint64_t philox_seed) {
  if (!c || !(dt > 0.0) || !(gamma > 0.0) || !(T >= 0.0))
    return gh_fail(c, GH_EINVAL, "ermak needs dt > 0, gamma > 0, T >= 0");
  if (rng != GH_RNG_LCG && rng != GH_RNG_PHILOX) return gh_fail(c, GH_EINVAL, "bad rng");
  c->integ = GH_INT_ERMAK;
  c->dt = dt;
  c->T = T;
  c->gamma = gamma;
  c->rng_mode = rng;
  c->philox_seed = philox_seed;
  return 0;
}

// Draw one step of the reference stream: per atom in caller order, per axis,
// rang(r1, r2) (ermak_a, Integration.F90:990-996).
static int lcg_step_normals(gh_ctx *c) {
  const int n = c->n, ncap = c->ncap;
  if (!c->gauss) {
    GH_TRY(dalloc(c, &c->gauss, 6 * (size_t)ncap));
    GH_HIP(c, hipHostMalloc((void **)&c->gauss_h, sizeof(double) * 6 * ncap, hipHostMallocDefault));
  }
  GH_HIP(c, hipStreamSynchronize(c->stream));  // the staging buffer is free again
  for (int i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      double r1, r2;
      c->lcg.rang(&r1, &r2);
      c->gauss_h[(size_t)(2 * d) * ncap + i] = r1;
      c->gauss_h[(size_t)(2 * d + 1) * ncap + i] = r2;
    }
  GH_HIP(c, hipMemcpyAsync(c->gauss, c->gauss_h, sizeof(double) * 6 * ncap,
                           hipMemcpyHostToDevice, c->stream));
  return 0;
}

// dinamic (Programs.f90:114-167): nsteps steps of the selected integrator,
// [epot, ekin, sum] in eV every `each` steps.
//
// Each step enqueues stepa (unless the previous step's force pass applied
// it), the rebuild chain, the force passes and stepb.  On the tiled velocity
// Verlet path the rebuild chain is enqueued only where the host expects a
// rebuild: every step the first force pass tests the device's rebuild rule
// itself, and if a rebuild is due without its chain it stops the run at that
// step (GhStatus::halt; every state-changing kernel then returns at once, so
// the state stays the one the step started from).  The last pass of each
// step publishes the step and any stop into pinned host memory; the host
// keeps at most kLead steps enqueued beyond the last one published and, on
// a stop, drains the stream, restores the positions buffer of that step,
// clears the stop and re-enqueues from it with the chain.  Rows are read
// after a stream synchronisation, with the status, so a row never holds a
// stopped step.  A stop for an overflow or an atom out of the grid ends the
// run with the error and the step; the state is then the start of that
// step's force evaluation (after its stepa) and the next gh_run completes
// that step first.
namespace {
struct StepRec {
  double4 *pos, *pos_next;
};
}  // namespace

static int finish_mid_step(gh_ctx *c) {
  // the stopped step's forces (the list re-sized if needed) and its stepb
  GH_TRY(interact(c));
  GH_TRY(check_status(c));
  GH_TRY(c->integ == GH_INT_ERMAK ? gh_launch_ermak_b(c) : gh_launch_vv_b(c));
  c->step += 1;
  c->mid_step = false;
  return 0;
}

#ifndef GH_CHAIN_AHEAD
#define GH_CHAIN_AHEAD 2
#endif
int gh_run(gh_ctx *c, int nsteps, int each, double *energy) {
  if (!c || nsteps < 0) return GH_EINVAL;
  if (each > 0 && !energy) return gh_fail(c, GH_EINVAL, "energy buffer required when each > 0");
  if (!(c->dt != 0.0)) return gh_fail(c, GH_ESTATE, "no integrator: call gh_integrator_*");
  GH_HIP(c, hipSetDevice(c->device));
  const double ui_ev = ghc::ui_ev();
  if (c->timing && c->evpool.empty()) {
    c->evpool.resize((size_t)EV_STEPS * EV_PER_STEP);
    for (auto &e : c->evpool) GH_HIP(c, hipEventCreate(&e));
  }
  const bool ermak = c->integ == GH_INT_ERMAK;
  const bool lcg = ermak && c->rng_mode == GH_RNG_LCG;
  int ns = 1;     // step of this call being enqueued (1-based)
  int iout = 0;   // energy rows written
  if (c->mid_step) {
    // the stopped step of the previous call is this call's first step
    if (nsteps == 0) return 0;
    GH_TRY(finish_mid_step(c));
    if (each == 1) {
      double ep, ek;
      GH_TRY(gh_launch_energies(c, &ep, &ek));
      energy[0] = ep * ui_ev;
      energy[1] = ek * ui_ev;
      energy[2] = (ep + ek) * ui_ev;
      iout = 1;
    }
    ns = 2;
  } else {
    // dinamic: one force evaluation before the loop (Programs.f90:129)
    GH_TRY(interact(c));
    GH_TRY(check_status(c));
  }
  // the chain only where expected: tiled velocity Verlet without per-step events
  const bool lazy = !ermak && !c->timing && gh_can_fuse_vv(c);
  // the tiled passes publish each step: the host then stays within kLead steps
  const bool publishes = c->list_mode == 1;
  const int kLead = 3;
  const int kRing = 64;
  std::vector<StepRec> ring(kRing);
  const long long s0 = c->step - (ns - 1);  // step s0 + k is this call's k-th
  volatile GhProgress *pg = c->prog_h;
  pg->halt = 0;
  pg->step = c->step;
  // velocity Verlet: stepb of step n and stepa of step n+1 run as one kernel
  // unless the state after stepb is observed (energy output, end of the run).
  // In the tiled path the force pass itself applies stepb of step n and stepa
  // of step n+1 (fused); elsewhere one kernel does both (k_vv_ba).
  bool pending_b = false, stepa_done = false, chain = false;
  // The host enqueues the rebuild chain where it expects the device's rule to
  // fire: from the published displacement maximum (the running maximum since
  // the last build, at the start of the latest published step) and its
  // fastest growth per step since that build, it chains every step whose
  // displacement could pass half the skin with half a step's growth to
  // spare.  A rebuild it misses still stops the run at its step (resume).
  long long seen_step = -1;
  double seen_r = 0.0, rate = 0.0;
  // Chained steps the device has not yet started (published): the host,
  // up to kLead steps ahead, would otherwise chain again the steps after a
  // rebuild from the displacement it last saw; at most GH_CHAIN_AHEAD
  // unseen chains (a rebuild that none of them catches stops the run at its
  // step, and the step is re-enqueued with the chain).
  std::deque<long long> unseen;
  auto rebuild_expected = [&](long long step) {
    const long long p = pg->step;
    const double r = sqrt(__builtin_bit_cast(double, (unsigned long long)pg->maxd2));
    if (p != seen_step) {
      if (seen_step < 0 || p < seen_step || r < seen_r) rate = 0.0;  // a build since
      else if (p > seen_step) rate = std::max(rate, (r - seen_r) / (double)(p - seen_step));
      seen_step = p;
      seen_r = r;
    }
    return rate > 0.0 && 2.0 * (r + rate * ((double)(step - p) + 0.5)) > c->skin;
  };
  // resumes at step index k (1-based in this run) after a stop there
  auto resume = [&](GhStatus &h) -> int {
    GH_HIP(c, hipStreamSynchronize(c->stream));
    GH_TRY(read_status(c, &h));
    const long long k = h.halt_step - s0;
    if (k < 1 || k > ns) return gh_fail(c, GH_ESTATE, "stop outside the run");
    if (k >= ns - kRing + 1) {
      const StepRec &r = ring[k % kRing];
      c->pos = r.pos;
      c->pos_next = r.pos_next;
    } else if (publishes) {
      return gh_fail(c, GH_ESTATE, "stop older than the step record");
    }  // else no step swaps the positions buffers
    c->step = h.halt_step - 1;
    iout = each > 0 ? (int)((k - 1) / each) : 0;
    if (h.halt != GH_HALT_REBUILD) {
      c->mid_step = true;
      return report_halt(c, h);
    }
    h.halt = 0;
    h.halt_step = 0;
    GH_TRY(write_status(c, &h));
    pg->halt = 0;
    pg->step = c->step;
    ns = (int)k;
    stepa_done = true;  // the stop came after the step's stepa
    pending_b = false;
    chain = true;
    unseen.clear();
    return 0;
  };
  for (;;) {
    while (ns <= nsteps) {
      if (publishes) {
        // at most kLead steps beyond the last one published
        const long long need = s0 + ns - 1 - kLead;
        for (int spin = 0; pg->step < need && !pg->halt; ++spin) {
          std::this_thread::yield();
          if ((spin & 1023) == 1023) {  // an idle stream has published every step it ran
            const hipError_t q = hipStreamQuery(c->stream);
            if (q == hipSuccess && pg->step < need && !pg->halt)
              return gh_fail(c, GH_ESTATE, "a step was not published");
            if (q != hipSuccess && q != hipErrorNotReady) GH_HIP(c, q);
          }
        }
        if (pg->halt) {
          GhStatus h;
          GH_TRY(resume(h));
          continue;
        }
      }
      if (lazy && !chain) {
        while (!unseen.empty() && unseen.front() <= pg->step) unseen.pop_front();
        chain = (int)unseen.size() < GH_CHAIN_AHEAD && rebuild_expected(s0 + ns);
      }
      c->step = s0 + ns;
      ring[ns % kRing] = StepRec{c->pos, c->pos_next};
      if (lcg) GH_TRY(lcg_step_normals(c));
      GH_TRY(rec(c, 0));
      if (pending_b) GH_TRY(gh_launch_vv_ba(c));
      else if (!stepa_done) GH_TRY(ermak ? gh_launch_ermak_a(c) : gh_launch_vv_a(c));
      pending_b = stepa_done = false;
      GH_TRY(rec(c, 1));
      if (!lazy || chain) {
        GH_TRY(gh_launch_rebuild(c, false));
        if (lazy) unseen.push_back(c->step);
      }
      chain = false;
      GH_TRY(rec(c, 2));
      const bool observed = ns == nsteps || (each > 0 && ns % each == 0);
      const bool fuse = !ermak && !observed && gh_can_fuse_vv(c);
      GH_TRY(gh_launch_forces(c, ev_at(c, 3), fuse));
      if (fuse) {
        std::swap(c->pos, c->pos_next);
        stepa_done = true;
      }
      GH_TRY(rec(c, 4));
      if (ermak || observed) GH_TRY(ermak ? gh_launch_ermak_b(c) : gh_launch_vv_b(c));
      else if (!fuse) pending_b = true;
      GH_TRY(rec(c, 5));
      if (c->timing) {
        c->evused += 1;
        if (c->evused == EV_STEPS) GH_TRY(timing_flush(c));
      }
      if (each > 0 && ns % each == 0) {
        double ep, ek;
        GH_TRY(gh_launch_energies(c, &ep, &ek));  // synchronises the stream
        GhStatus h;
        GH_TRY(read_status(c, &h));
        if (h.halt) {
          GH_TRY(resume(h));
          continue;
        }
        const double f1 = ep * ui_ev, f2 = ek * ui_ev;
        energy[3 * iout + 0] = f1;
        energy[3 * iout + 1] = f2;
        energy[3 * iout + 2] = f1 + f2;
        ++iout;
      }
      ++ns;
    }
    // the last steps: no stop may remain unseen
    GH_HIP(c, hipStreamSynchronize(c->stream));
    GhStatus h;
    GH_TRY(read_status(c, &h));
    if (!h.halt) break;
    GH_TRY(resume(h));
  }
  GH_TRY(timing_flush(c));
  GH_TRY(check_status(c));
  return 0;
}

int gh_energies(gh_ctx *c, double *epot, double *ekin) {
  if (!c) return GH_EINVAL;
  if (c->first_force) GH_TRY(interact(c));
  double ep, ek;
  GH_TRY(gh_launch_energies(c, &ep, &ek));
  if (epot) *epot = ep * ghc::ui_ev();
  if (ekin) *ekin = ek * ghc::ui_ev();
  return 0;
}

// lbfgs_minimizator (Quasi_Newton.f90:160-232): limited-memory BFGS on all
// atoms with the device force path.  The reference hands GEMS's energy
// (eV) and gradient to Nocedal's lbfgs (m = 5 pairs, eps = 1e-8, stop when
// |g| / max(1, |x|) < eps); here the same quantities, the two-loop recursion
// with the newest pair's scaling as the initial Hessian, and a backtracking
// line search to the Armijo condition (c1 = 1e-4), first step 1/|g|.
int gh_lbfgs(gh_ctx *c, int m, double eps, int max_iter, int *iters, double *epot_ev) {
  if (!c || m < 1 || !(eps > 0) || max_iter < 0) return gh_fail(c, GH_EINVAL, "gh_lbfgs: bad arguments");
  if (!c->have_atoms) return gh_fail(c, GH_ESTATE, "no atoms");
  const int n3 = 3 * c->n;
  const double ui_ev = ghc::ui_ev();
  std::vector<double> x(n3), g(n3), F(n3);
  GH_TRY(gh_get_state(c, x.data(), nullptr, nullptr, nullptr));
  auto eval = [&](const std::vector<double> &xx, double &f, std::vector<double> &gg) -> int {
    GH_TRY(gh_set_positions(c, xx.data()));
    double et = 0.0;
    GH_TRY(gh_compute(c, F.data(), nullptr, &et));
    f = et * ui_ev;
    for (int k = 0; k < n3; ++k) gg[k] = -F[k] * ui_ev;
    return 0;
  };
  auto dot = [&](const std::vector<double> &a, const std::vector<double> &b) {
    double s = 0.0;
    for (int k = 0; k < n3; ++k) s += a[k] * b[k];
    return s;
  };
  double f = 0.0;
  GH_TRY(eval(x, f, g));
  std::vector<std::vector<double>> S, Y;
  std::vector<double> rho;
  std::vector<double> d(n3), xn(n3), gn(n3), q(n3);
  int it = 0;
  for (; it < max_iter; ++it) {
    if (std::sqrt(dot(g, g)) / std::max(1.0, std::sqrt(dot(x, x))) < eps) break;
    // search direction: the two-loop recursion
    q = g;
    std::vector<double> alpha(S.size());
    for (int i = (int)S.size() - 1; i >= 0; --i) {
      alpha[i] = rho[i] * dot(S[i], q);
      for (int k = 0; k < n3; ++k) q[k] -= alpha[i] * Y[i][k];
    }
    double h0 = 1.0;
    if (!S.empty()) h0 = dot(S.back(), Y.back()) / dot(Y.back(), Y.back());
    for (int k = 0; k < n3; ++k) q[k] *= h0;
    for (size_t i = 0; i < S.size(); ++i) {
      const double b = rho[i] * dot(Y[i], q);
      for (int k = 0; k < n3; ++k) q[k] += S[i][k] * (alpha[i] - b);
    }
    for (int k = 0; k < n3; ++k) d[k] = -q[k];
    double gd = dot(g, d);
    if (!(gd < 0.0)) {  // not a descent direction: restart from steepest descent
      S.clear();
      Y.clear();
      rho.clear();
      for (int k = 0; k < n3; ++k) d[k] = -g[k];
      gd = dot(g, d);
    }
    double stp = S.empty() ? 1.0 / std::sqrt(dot(g, g)) : 1.0;
    double fn = 0.0;
    int ls = 0;
    for (;; ++ls) {
      for (int k = 0; k < n3; ++k) xn[k] = x[k] + stp * d[k];
      GH_TRY(eval(xn, fn, gn));
      if (fn <= f + 1e-4 * stp * gd || ls >= 40) break;
      stp *= 0.5;
    }
    if (ls >= 40 && fn > f) break;  // no decrease along d: stop where we are
    std::vector<double> sv(n3), yv(n3);
    for (int k = 0; k < n3; ++k) {
      sv[k] = xn[k] - x[k];
      yv[k] = gn[k] - g[k];
    }
    const double sy = dot(sv, yv);
    if (sy > 1e-300) {
      if ((int)S.size() == m) {
        S.erase(S.begin());
        Y.erase(Y.begin());
        rho.erase(rho.begin());
      }
      S.push_back(sv);
      Y.push_back(yv);
      rho.push_back(1.0 / sy);
    }
    x.swap(xn);
    g.swap(gn);
    f = fn;
  }
  GH_TRY(gh_set_positions(c, x.data()));
  double et = 0.0;
  GH_TRY(gh_compute(c, nullptr, nullptr, &et));
  if (iters) *iters = it;
  if (epot_ev) *epot_ev = et * ui_ev;
  return 0;
}

int gh_get_state(gh_ctx *c, double *x, double *v, double *F, double *e) {
  if (!c || !c->have_atoms) return gh_fail(c, GH_ESTATE, "no atoms");
  std::vector<int> orig;
  GH_TRY(fetch_orig(c, orig));
  if (x) {
    std::vector<double4> h(c->n);
    GH_HIP(c, hipMemcpyAsync(h.data(), c->pos, sizeof(double4) * c->n, hipMemcpyDeviceToHost,
                             c->stream));
    GH_HIP(c, hipStreamSynchronize(c->stream));
    for (int s = 0; s < c->n; ++s) {
      double v[3] = {h[s].x, h[s].y, h[s].z};
      // positions run continuously between rebuilds (at most half a skin
      // outside the box); report them wrapped as do_pbc_1 does
      for (int d = 0; d < 3; ++d)
        if (c->boxed && c->pbc[d]) {
          if (v[d] >= c->L[d]) v[d] -= c->L[d];
          else if (v[d] < 0.0) v[d] += c->L[d];
        }
      for (int d = 0; d < 3; ++d) x[3 * (size_t)orig[s] + d] = v[d];
    }
  }
  if (v) GH_TRY(fetch_soa3(c, c->vel, orig, v));
  if (F) GH_TRY(fetch_soa3(c, c->frc, orig, F));
  if (e) {
    std::vector<double> h(c->n);
    GH_HIP(c, hipMemcpyAsync(h.data(), c->epa, sizeof(double) * c->n, hipMemcpyDeviceToHost,
                             c->stream));
    GH_HIP(c, hipStreamSynchronize(c->stream));
    for (int s = 0; s < c->n; ++s) e[orig[s]] = h[s];
  }
  return 0;
}

int gh_get_stats(gh_ctx *c, gh_stats *s) {
  if (!c || !s) return GH_EINVAL;
  GhStatus h;
  GH_TRY(read_status(c, &h));
  s->nupd = h.nupd;
  s->maxnn = h.maxnn;
  s->cap = c->cap;
  for (int d = 0; d < 3; ++d) s->ncells[d] = c->list_mode ? c->nc[d] : 0;
  s->list_mode = c->list_mode;
  s->steps = c->step;
  s->launches = c->launches;
  s->list_build = c->list_build;
  s->chunks = c->list_mode == 1 ? c->tg.chunks : 0;
  return 0;
}

int gh_set_timing(gh_ctx *c, int enable) {
  if (!c) return GH_EINVAL;
  GH_TRY(timing_flush(c));
  c->timing = enable != 0;
  for (int k = 0; k < 4; ++k) {
    c->t_ms[k] = 0;
    c->t_cnt[k] = 0;
  }
  return 0;
}

int gh_get_timing(gh_ctx *c, double ms[4], int64_t count[4]) {
  if (!c || !ms || !count) return GH_EINVAL;
  GH_TRY(timing_flush(c));
  for (int k = 0; k < 4; ++k) {
    ms[k] = c->t_ms[k];
    count[k] = c->t_cnt[k];
  }
  return 0;
}

}  // extern "C"
