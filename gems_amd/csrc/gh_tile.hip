// Brick-tiled neighbour list and force passes (cell mode).
//
// Atoms are sorted by key = brick * cpb + cell-within-brick (see TileGeom).
// One workgroup owns one brick: it stages the positions of the brick's halo
// (every cell of the brick grown by one cell, periodic images shifted) into
// LDS, and the neighbour list of each atom holds 16-bit LDS slot indices into
// that halo.  The force passes then read neighbour positions from LDS only;
// HBM sees each position a few times per pass (once per halo it belongs to)
// instead of once per neighbour pair, and list rows cost 2 B per entry.
//
// LDS slot layout per workgroup: the halo cells in order (hx fastest), each
// cell's atoms in sorted order.  hbase[h] is the first slot of halo cell h and
// the atom in slot s of cell h is s + delta[h].
//
// List layout ("quad ELL"): entry k of atom a is the 16-bit word
// ((k / 4) * ld + a) * 4 + k % 4, so one 8-byte load per lane fetches four
// consecutive entries and a wavefront's load covers 512 contiguous bytes.
//
// Mixed precision stages positions as fixed point relative to the brick
// centre, 2^fxk steps per Angstrom (fxk <= 25, chosen so that every staged
// atom lies within 2^30 steps): pair vectors are exact integer differences,
// converted exactly to FP64, so the only rounding of a pair's geometry is the
// 2^-(fxk+1) A quantisation of each position (1.5e-8 A at fxk = 25).
#include <math.h>
#include <cstring>

#include "gh_device.h"

#define GH_TBLOCK 256   // list build, occupancy and layout kernels
#define GH_TBMAX 1024   // force passes: TileGeom::tpb threads, at most this

namespace {
using namespace ghd;

struct Halo {
  int nh, total, a0, a1, key0, hd0, hd1, pad;
  double org[3];                  // brick centre: origin of the fixed-point frame
  int delta[GH_MAXHALO];          // atom in slot s of halo cell h: s + delta[h]
  int hbase[GH_MAXHALO + 1];      // first slot of halo cell h
  int wsum[GH_TBMAX / 64];        // halo_setup's per-wavefront count totals
  signed char shc[3][GH_MAXHALO]; // periodic image shift of halo cell h (box lengths)
};

__device__ __forceinline__ int grid_key(const TileGeom &t, int cx, int cy, int cz) {
  const int bx = cx / t.bd[0], by = cy / t.bd[1], bz = cz / t.bd[2];
  const int lx = cx - bx * t.bd[0], ly = cy - by * t.bd[1], lz = cz - bz * t.bd[2];
  return ((bz * t.nbr[1] + by) * t.nbr[0] + bx) * t.cpb + (lz * t.bd[1] + ly) * t.bd[0] + lx;
}

__device__ __forceinline__ double shift(const TileGeom &t, const Halo &H, int d, int h) {
  return (double)H.shc[d][h] * t.L[d];  // exact: 0 or +-L
}

// Halo of brick b: atom ranges, image shifts and LDS slot bases of the
// (bd+2)^3 halo cells.  Returns false, uniformly over the workgroup, if the
// brick holds no atoms or (check) its halo exceeds slot_cap, which stops the
// run at `step`.
__device__ bool halo_setup(const TileGeom &t, int b, const int *__restrict__ cell_start, Halo &H,
                           GhStatus *st, bool check, long long step = -1) {
  const int bx = b % t.nbr[0], by = (b / t.nbr[0]) % t.nbr[1], bz = b / (t.nbr[0] * t.nbr[1]);
  const int hd0 = t.bd[0] + 2, hd1 = t.bd[1] + 2, hd2 = t.bd[2] + 2;
  const int nh = hd0 * hd1 * hd2;
  const int key0 = b * t.cpb;
  const int h = threadIdx.x;
  int cnt = 0, start = 0;
  if (h < nh) {
    const int hx = h % hd0, hy = (h / hd0) % hd1, hz = h / (hd0 * hd1);
    int g[3] = {bx * t.bd[0] + hx - 1, by * t.bd[1] + hy - 1, bz * t.bd[2] + hz - 1};
    const int bc[3] = {bx, by, bz};
    int sc[3] = {0, 0, 0};
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int lo = bc[d] * t.bd[d];
      const int hi = min(lo + t.bd[d], t.gn[d]);  // one past the brick's last cell
      if (g[d] > hi) ok = false;                  // not adjacent to any cell of the brick
      if (t.pbc[d]) {
        if (g[d] < 0) {
          g[d] += t.nc[d];
          sc[d] = -1;
        } else if (g[d] >= t.nc[d]) {
          g[d] -= t.nc[d];
          sc[d] = 1;
        }
      } else if (g[d] < 0 || g[d] >= t.gn[d]) {
        ok = false;
      }
    }
    if (ok) {
      const int key = grid_key(t, g[0], g[1], g[2]);
      start = cell_start[key];
      cnt = cell_start[key + 1] - start;
    }
    H.shc[0][h] = (signed char)sc[0];
    H.shc[1][h] = (signed char)sc[1];
    H.shc[2][h] = (signed char)sc[2];
  }
  // exclusive scan of the halo cell counts: within each wavefront by
  // shuffles, then across the four wavefronts
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int inc = cnt;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) H.wsum[wv] = inc;
  __syncthreads();
  int woff = 0, total = 0;
  const int nw = (int)(blockDim.x >> 6);
  for (int w = 0; w < nw; ++w) {
    const int x = H.wsum[w];
    if (w < wv) woff += x;
    total += x;
  }
  if (h < nh) {
    const int base = woff + inc - cnt;
    H.hbase[h] = base;
    H.delta[h] = start - base;
  }
  if (h == 0) {
    H.hbase[nh] = total;
    H.total = total;
    H.nh = nh;
    H.hd0 = hd0;
    H.hd1 = hd1;
    H.key0 = key0;
    H.a0 = cell_start[key0];
    H.a1 = cell_start[key0 + t.cpb];
    const int bc[3] = {bx, by, bz};
#pragma unroll
    for (int d = 0; d < 3; ++d)
      H.org[d] = ((double)(bc[d] * t.bd[d] - t.gofs[d]) + 0.5 * t.bd[d]) * (t.L[d] / t.nc[d]);
  }
  __syncthreads();
  if (H.a1 <= H.a0) return false;
  if (check && H.total > t.slot_cap) {
    if (threadIdx.x == 0) {
      st->tile_overflow = 1;
      raise_halt(st, GH_HALT_TILE, step);
    }
    return false;
  }
  return true;
}

// Slot -> halo cell map in LDS (halo cells number < 256).
__device__ __forceinline__ void slot_map(const Halo &H, unsigned char *scell) {
  const int h = threadIdx.x;
  if (h < H.nh)
    for (int s = H.hbase[h]; s < H.hbase[h + 1]; ++s) scell[s] = (unsigned char)h;
  __syncthreads();
}

// Halo cell holding LDS slot s (hbase non-decreasing; empty cells repeat).
__device__ __forceinline__ int halo_of_slot(const Halo &H, int s) {
  int lo = 0, hi = H.nh;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (H.hbase[mid] <= s) lo = mid;
    else hi = mid;
  }
  return lo;
}

// stepb of this step and stepa of the next for atom a (velocity_verlet_b,
// Integration.F90:1171-1186, then velocity_verlet_a, :1150-1169): the same
// operations in the same order as k_vv_ba, on the force just summed.
// New positions go to pos_next (other workgroups still stage pos); returns
// |x_new - x_old|^2 for the skin test.
// The update's loads of atom a (velocity, build position), issued together
// before its arithmetic; the mixed SMA-TB and the LJ passes over small
// systems issue them before the pair loop (those passes wait on each memory
// round trip), the others after it.
struct VvLoad {
  double v[3], xo[3];
};
__device__ __forceinline__ VvLoad vv_load(const FuseVV &f, int a) {
  VvLoad r;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r.v[d] = f.vel[d * f.ncap + a];
    r.xo[d] = f.xold[d * f.ncap + a];
  }
  return r;
}
__device__ __forceinline__ double fused_vv(const FuseVV &f, int a, const double4 &p, int t,
                                           double Fx, double Fy, double Fz, const VvLoad &ld) {
  const double om = f.om[t];
  const double F[3] = {Fx, Fy, Fz};
  double x[3] = {p.x, p.y, p.z};
  double d2 = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double acc = F[d] * om;
    const double v = ld.v[d] + f.acel_dv * acc;  // stepb
    x[d] = x[d] + f.dt * v + f.acel_dx * acc;    // stepa
    f.vel[d * f.ncap + a] = v + f.acel_dv * acc;
    const double dd = x[d] - ld.xo[d];
    d2 += dd * dd;
  }
  f.pos_next[a] = make_double4(x[0], x[1], x[2], p.w);
  return d2;
}
__device__ __forceinline__ double fused_vv(const FuseVV &f, int a, const double4 &p, int t,
                                           double Fx, double Fy, double Fz) {
  return fused_vv(f, a, p, t, Fx, Fy, Fz, vv_load(f, a));
}

// Workgroup maximum of d2, stored at dst[b] (all threads call it).
// Workgroup maximum of |x - x_old|^2 folded into the device-wide maximum.
__device__ __forceinline__ void block_max_atomic(double d2, unsigned long long *maxd2) {
  __shared__ double part[GH_TBMAX / 64];
  d2 = wave_max(d2);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x >> 6);
    for (int w = 1; w < nw; ++w) d2 = fmax(d2, part[w]);
    atomicMax(maxd2, (unsigned long long)__double_as_longlong(d2));
  }
}

// Lanes of the wavefront below this one set in mask.
__device__ __forceinline__ int lane_rank(unsigned long long mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

__device__ __forceinline__ size_t qidx(int k, int a, int ld) {
  return (((size_t)(k >> 2) * ld + a) << 2) + (k & 3);
}


// ---------------------------------------------------------------------------
// Neighbour list (linkedcell_self, Neighbour.F90:991-1087): for each atom of
// the brick, the slots of its 27 neighbour cells with rd < rl^2, the atom
// itself excluded, where rd is rounded exactly as the reference's vdistance +
// dot_product of vd = (x_j - x_i) + image shift, so the sets are bit-identical.
//
// One sweep.  With the halo staged (STAGE), a candidate is first tested in
// FP32 on brick-local coordinates; only candidates within `margin` of rl^2
// (FP32 error bound x 10) are decided by the exact FP64 rd from L2.
//
// Row order.  Each row lists its pairs by class of rd at build time, the
// candidate order kept within a class: 0 below (rin - m)^2, 1 below
// (rin + m)^2, 2 below (rout + m)^2, 3 the rest, where rin is the bound of
// the pair function's inner branch (min rci for SMA-TB, rc for LJ), rout its
// cutoff (max rce, rc) and m a quarter of the skin.  Entry k of every row of
// a wavefront then mostly falls in one class, so the lanes take one branch of
// the pair function, and the trailing class-3 quads (pairs that stay beyond
// the cutoff until they move m) are skipped by the whole wavefront.  The
// class is decided in FP32 unless r^2 lies within `margin` of a bound, where
// the exact FP64 rd decides: the order, and with it the FP64 sums, do not
// depend on the brick shape.  The last quad is padded with `pad`, a sentinel
// LDS slot far from every atom.
// ---------------------------------------------------------------------------
#define GH_FARMAX 48

struct BuildArgs {
  double rl2;
  float lo2f, hi2f;  // FP32 prefilter band around rl^2
  double cls2[3];    // row-order class bounds on r^2 (non-decreasing)
  float clsf[3];     // the same in FP32
  float clsm;        // within clsm of a bound the FP64 r^2 decides the class
  int cap, ld, ncap, lds_slots, lds_rows;
  // cooperative build (k_build_coop): chunks per brick, staged slots,
  // candidate-list entries and row-buffer entries per wavefront in its LDS
  int chunks, nsl, ncand, rstride;
  int *nn;
  unsigned short *list16;
  double *xold;
};

__device__ __forceinline__ int row_class(double rd, const BuildArgs &A) {
  return (rd >= A.cls2[0]) + (rd >= A.cls2[1]) + (rd >= A.cls2[2]);
}

// Class from the FP32 r^2, or -1 when it lies within clsm of a bound.
__device__ __forceinline__ int row_class_f(float r2f, const BuildArgs &A) {
  const float lo = r2f - A.clsm, hi = r2f + A.clsm;
  const int cl = (lo >= A.clsf[0]) + (lo >= A.clsf[1]) + (lo >= A.clsf[2]);
  const int ch = (hi >= A.clsf[0]) + (hi >= A.clsf[1]) + (hi >= A.clsf[2]);
  return cl == ch ? cl : -1;
}

// rd of atom a (position pa) and slot s of halo cell h, rounded as the
// reference's vdistance + dot_product.
__device__ __forceinline__ double pair_rd(const Halo &H, const TileGeom &t,
                                          const double4 *__restrict__ pos, const double4 &pa,
                                          int s, int h) {
  const double4 pj = pos[s + H.delta[h]];
  const double vx = (pj.x - pa.x) + shift(t, H, 0, h);
  const double vy = (pj.y - pa.y) + shift(t, H, 1, h);
  const double vz = (pj.z - pa.z) + shift(t, H, 2, h);
  return dist2_rn(vx, vy, vz);
}

// Rows through a sequential writer.  Staged: class-0 pairs go out at once and
// the others wait in an LDS buffer, tagged with their class (GH_FARMAX per
// thread; any excess joins the first run, which only costs order), to follow
// class by class.  Unstaged (slot numbers may need all 16 bits): classes 0-2
// go out in candidate order and class 3 waits.
template <bool STAGE>
__device__ __forceinline__ void build_rows(const Halo &H, const TileGeom &t,
                                           const double4 *__restrict__ pos,
                                           const int *__restrict__ cellof, const float *xf,
                                           const float *yf, const float *zf,
                                           unsigned short *far, const BuildArgs &A, int pad,
                                           int &mmax) {
  const int cap = A.cap;
  for (int a = H.a0 + threadIdx.x; a < H.a1; a += blockDim.x) {
    const int lc = cellof[a] - H.key0;
    const int lx = lc % t.bd[0], ly = (lc / t.bd[0]) % t.bd[1], lz = lc / (t.bd[0] * t.bd[1]);
    const int ha = (lx + 1) + H.hd0 * ((ly + 1) + H.hd1 * (lz + 1));
    const int own = a - H.delta[ha];
    const double4 pa = pos[a];
    float xi = 0.f, yi = 0.f, zi = 0.f;
    if (STAGE) {
      xi = xf[own];
      yi = yf[own];
      zi = zf[own];
    }
    RowWriter row{reinterpret_cast<unsigned long long *>(A.list16), A.ld, a, cap, 0, 0ull};
    int nfar = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int h = ha + dx + H.hd0 * (dy + H.hd1 * dz);
          const int se = H.hbase[h + 1];
          for (int s = H.hbase[h]; s < se; ++s) {
            if (s == own) continue;
            int c;
            if (STAGE) {
              const float ux = xf[s] - xi, uy = yf[s] - yi, uz = zf[s] - zi;
              const float r2f = fmaf(ux, ux, fmaf(uy, uy, uz * uz));
              if (r2f > A.hi2f) continue;
              double rd = -1.0;
              if (r2f >= A.lo2f) {  // near the cut: the reference's rounding decides
                rd = pair_rd(H, t, pos, pa, s, h);
                if (!(rd < A.rl2)) continue;
              }
              c = row_class_f(r2f, A);
              if (c < 0) c = row_class(rd >= 0.0 ? rd : pair_rd(H, t, pos, pa, s, h), A);
            } else {
              const double rd = pair_rd(H, t, pos, pa, s, h);
              if (!(rd < A.rl2)) continue;
              c = row_class(rd, A);
            }
            if (c < (STAGE ? 1 : 3) || nfar == GH_FARMAX) row.put(s);
            else far[(nfar++) * GH_TBLOCK + threadIdx.x] =
                (unsigned short)(STAGE ? s | (c << 14) : s);
          }
        }
    if (STAGE) {
      for (int c = 1; c < 4; ++c)
        for (int k = 0; k < nfar; ++k) {
          const int e = far[k * GH_TBLOCK + threadIdx.x];
          if ((e >> 14) == c) row.put(e & 0x3FFF);
        }
    } else {
      for (int k = 0; k < nfar; ++k) row.put(far[k * GH_TBLOCK + threadIdx.x]);
    }
    row.finish(pad);
    A.nn[a] = min(row.m, cap);
    A.xold[a] = pa.x;
    A.xold[A.ncap + a] = pa.y;
    A.xold[2 * A.ncap + a] = pa.z;
    mmax = max(mmax, row.m);
  }
}

// Staged build with the rows assembled in LDS (rb: entry k of this thread's
// row at rb[k * GH_TBLOCK], k = cap is a dump slot): every candidate is
// written through a computed index, so the candidate loop has no branch.
// A second sweep makes the exact tests, tags each entry with its class and
// counts the classes; a third writes every entry to its place in the list.
__device__ __forceinline__ void build_rows_lds(const Halo &H, const TileGeom &t,
                                               const double4 *__restrict__ pos,
                                               const int *__restrict__ cellof, const float *xf,
                                               const float *yf, const float *zf,
                                               const unsigned char *scell,
                                               unsigned short *rowbuf, const BuildArgs &A,
                                               int pad, int &mmax) {
  unsigned short *rb = rowbuf + threadIdx.x;
  const int cap = A.cap;
  for (int a = H.a0 + threadIdx.x; a < H.a1; a += blockDim.x) {
    const int lc = cellof[a] - H.key0;
    const int lx = lc % t.bd[0], ly = (lc / t.bd[0]) % t.bd[1], lz = lc / (t.bd[0] * t.bd[1]);
    const int ha = (lx + 1) + H.hd0 * ((ly + 1) + H.hd1 * (lz + 1));
    const int own = a - H.delta[ha];
    const double4 pa = pos[a];
    const float xi = xf[own], yi = yf[own], zi = zf[own];
    int m = 0;
    // The cells dx = -1, 0, 1 of a stencil row are consecutive halo cells,
    // so their slots form one run: nine runs, four candidates per step.  The
    // candidate loop has no branch: a candidate inside the FP32 band keeps a
    // flag (bit 15) for the exact test of the next sweep.
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy) {
        const int h = ha + H.hd0 * (dy + H.hd1 * dz);
        const int sb = H.hbase[h - 1], se = H.hbase[h + 2];
        for (int s0 = sb; s0 < se; s0 += 4) {
          float r2[4];
          int sl[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            sl[u] = min(s0 + u, se - 1);
            const float ux = xf[sl[u]] - xi, uy = yf[sl[u]] - yi, uz = zf[sl[u]] - zi;
            r2[u] = fmaf(ux, ux, fmaf(uy, uy, uz * uz));
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const bool keep = (s0 + u < se) & (sl[u] != own) & (r2[u] <= A.hi2f);
            const int e = sl[u] | (r2[u] >= A.lo2f ? 0x8000 : 0);
            rb[(keep ? min(m, cap) : cap) * GH_TBLOCK] = (unsigned short)e;
            m += keep ? 1 : 0;
          }
        }
      }
    // Second sweep: exact test of the flagged entries (the reference's
    // rounding decides; failures are dropped, compacting in place) and the
    // class of each kept pair in bits 14-15 (staged slots number below 16384).
    int mw = 0, n0 = 0, n1 = 0, n2 = 0;
    for (int k = 0, ke = min(m, cap); k < ke; ++k) {
      const int e = rb[k * GH_TBLOCK];
      const int s = e & 0x3FFF;
      int c;
      if (e & 0x8000) {
        const double rd = pair_rd(H, t, pos, pa, s, scell[s]);
        if (!(rd < A.rl2)) continue;
        c = row_class(rd, A);
      } else {
        const float ux = xf[s] - xi, uy = yf[s] - yi, uz = zf[s] - zi;
        c = row_class_f(fmaf(ux, ux, fmaf(uy, uy, uz * uz)), A);
        if (c < 0) c = row_class(pair_rd(H, t, pos, pa, s, scell[s]), A);
      }
      rb[(mw++) * GH_TBLOCK] = (unsigned short)(s | (c << 14));
      n0 += c == 0 ? 1 : 0;
      n1 += c == 1 ? 1 : 0;
      n2 += c == 2 ? 1 : 0;
    }
    // One sweep writes each entry straight to its place in class order.
    unsigned short *row = A.list16 + ((size_t)a << 2);
    const size_t qstride = (size_t)A.ld << 2;
    int d0 = 0, d1 = n0, d2 = n0 + n1, d3 = n0 + n1 + n2;
    for (int k = 0; k < mw; ++k) {
      const int e = rb[k * GH_TBLOCK];
      const int c = e >> 14;
      const int d = c == 0 ? d0 : c == 1 ? d1 : c == 2 ? d2 : d3;
      d0 += c == 0 ? 1 : 0;
      d1 += c == 1 ? 1 : 0;
      d2 += c == 2 ? 1 : 0;
      d3 += c == 3 ? 1 : 0;
      row[(d >> 2) * qstride + (d & 3)] = (unsigned short)(e & 0x3FFF);
    }
    for (int d = mw; d & 3; ++d) row[(d >> 2) * qstride + (d & 3)] = (unsigned short)pad;
    A.nn[a] = mw;
    A.xold[a] = pa.x;
    A.xold[A.ncap + a] = pa.y;
    A.xold[2 * A.ncap + a] = pa.z;
    mmax = max(mmax, m);  // before the exact drops: an overflow may have lost entries
  }
}

// ---------------------------------------------------------------------------
// Binning (sort_in_cells, Neighbour.F90:927-989, as a counting sort by key =
// brick * cpb + cell within the brick): wrap, copy the per-atom state aside
// and count per key; exclusive scan; scatter; order each cell by caller index
// (deterministic layout) and gather the state back in the new order.
// ---------------------------------------------------------------------------
struct CellArgs {
  GhStatus *st;
  TileGeom t;
  double skin;
  int n, ncap, nkey;
  double4 *pos, *pos_t;
  double *vel, *vel_t, *acc, *acc_t;
  int *type, *type_t, *orig, *orig_t;
  int *cellid, *cellof, *count, *start, *perm;
};

__device__ __forceinline__ int tile_coord(double x, const TileGeom &t, int d, int *outside) {
  int c = (int)floor(x * t.one_cell[d]);
  if (t.pbc[d]) {
    c %= t.nc[d];
    if (c < 0) c += t.nc[d];
  } else {
    c += t.gofs[d];
    if (c < 0) {
      c = 0;
      *outside = 1;
    }
    if (c >= t.gn[d]) {
      c = t.gn[d] - 1;
      *outside = 1;
    }
  }
  return c;
}

// Wraps periodic coordinates into the box (do_pbc_1, Set_Properties.f90:
// 444-470), copies the atom's state to the _t arrays (the gather reads it
// there and writes the new order in place), computes the atom's key and
// counts atoms per key.
__device__ void cell_bin(const CellArgs &R, const double4 *__restrict__ pos, int i0, int stride,
                         long long step) {
  const TileGeom &t = R.t;
  const int ncap = R.ncap;
  for (int i = i0; i < R.n; i += stride) {
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (t.pbc[d]) {
        x[d] = x[d] - t.L[d] * floor(x[d] / t.L[d]);
        if (x[d] >= t.L[d]) x[d] -= t.L[d];
      }
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
    R.pos_t[i] = p;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      R.vel_t[d * ncap + i] = R.vel[d * ncap + i];
      R.acc_t[d * ncap + i] = R.acc[d * ncap + i];
    }
    R.type_t[i] = R.type[i];
    R.orig_t[i] = R.orig[i];
    int out = 0;
    const int cx = tile_coord(x[0], t, 0, &out);
    const int cy = tile_coord(x[1], t, 1, &out);
    const int cz = tile_coord(x[2], t, 2, &out);
    if (out) {
      R.st->outside = 1;
      raise_halt(R.st, GH_HALT_RANGE, step);
    }
    const int key = grid_key(t, cx, cy, cz);
    R.cellid[i] = key;
    // one atomic per distinct key of the wavefront (atoms sorted at the last
    // build share cells: a wavefront holds a few keys)
    for (unsigned long long todo = __ballot(1); todo;) {
      const int leader = __ffsll((long long)todo) - 1;
      const int lk = __shfl(key, leader, 64);
      const unsigned long long grp = __ballot(key == lk) & todo;
      if ((int)(threadIdx.x & 63) == leader) atomicAdd(&R.count[lk], __popcll(grp));
      todo &= ~grp;
    }
  }
}

// Workgroup (BS threads) exclusive scan of one value per thread in thread
// order; returns this thread's offset and the workgroup total.
template <int BS>
__device__ __forceinline__ int block_scan(int v, int &total) {
  __shared__ int wsum[BS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int inc = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  int woff = 0;
  total = 0;
  for (int w = 0; w < BS / 64; ++w) {
    const int x = wsum[w];
    if (w < wv) woff += x;
    total += x;
  }
  __syncthreads();
  return woff + inc - v;
}

// Exclusive scan of the cell counts by one workgroup of BS threads.  Each
// pass gives every thread GH_SCANRUN consecutive keys, loaded and stored 16 B
// at a time, so a pass costs one memory round trip for the loads and one for
// the stores; a carry links the passes.  start = the prefix; start[nkey] =
// the total.
#define GH_SCANRUN 32
template <int BS>
__device__ void cell_scan(const CellArgs &R) {
  int carry = 0;
  for (int base = 0; base < R.nkey; base += GH_SCANRUN * BS) {
    const int k0 = base + GH_SCANRUN * threadIdx.x;
    int v[GH_SCANRUN];
#pragma unroll
    for (int u = 0; u < GH_SCANRUN; u += 4) {
      const int k = k0 + u;
      if (k + 3 < R.nkey) {
        const int4 q = *reinterpret_cast<const int4 *>(R.count + k);
        v[u] = q.x;
        v[u + 1] = q.y;
        v[u + 2] = q.z;
        v[u + 3] = q.w;
      } else {
#pragma unroll
        for (int w = 0; w < 4; ++w) v[u + w] = k + w < R.nkey ? R.count[k + w] : 0;
      }
    }
    int tot = 0;
#pragma unroll
    for (int u = 0; u < GH_SCANRUN; ++u) tot += v[u];
    int total;
    int run = carry + block_scan<BS>(tot, total);
#pragma unroll
    for (int u = 0; u < GH_SCANRUN; u += 4) {
      const int k = k0 + u;
      int4 q;
      q.x = run;
      q.y = q.x + v[u];
      q.z = q.y + v[u + 1];
      q.w = q.z + v[u + 2];
      run = q.w + v[u + 3];
      if (k + 3 < R.nkey) {
        *reinterpret_cast<int4 *>(R.start + k) = q;
      } else {
        if (k < R.nkey) R.start[k] = q.x;
        if (k + 1 < R.nkey) R.start[k + 1] = q.y;
        if (k + 2 < R.nkey) R.start[k + 2] = q.z;
      }
    }
    carry += total;
  }
  if (threadIdx.x == 0) R.start[R.nkey] = carry;
}

// Places each atom in its cell's range, counting the cell's count down to
// zero (ready for the next binning); the order within a cell is the atomics'
// and is fixed by the rank gather.
__device__ void cell_scatter(const CellArgs &R, int i0, int stride) {
  const int lane = threadIdx.x & 63;
  for (int i = i0; i < R.n; i += stride) {
    const int key = R.cellid[i];
    // one atomic per distinct key of the wavefront: the group's lanes take
    // the slots below the count the atomic returns, one each
    int slot = 0;
    for (unsigned long long todo = __ballot(1); todo;) {
      const int leader = __ffsll((long long)todo) - 1;
      const int lk = __shfl(key, leader, 64);
      const unsigned long long grp = __ballot(key == lk) & todo;
      int old = 0;
      if (lane == leader) old = atomicSub(&R.count[lk], __popcll(grp));
      old = __shfl(old, leader, 64);
      if (key == lk) slot = R.start[lk] + old - 1 - lane_rank(grp);
      todo &= ~grp;
    }
    R.perm[slot] = i;
  }
}

// One thread per slot: the atom placed there by the scatter takes the rank
// of its caller index among its cell's atoms (the layout is then independent
// of the atomics' order), and its state moves from the _t copies to that
// slot.  Every thread reads only the scatter's output and the copies, so
// no ordering between threads is needed.
__device__ void cell_rank_gather(const CellArgs &R, double4 *__restrict__ pos, int s0,
                                 int stride) {
  const int ncap = R.ncap;
  for (int s = s0; s < R.n; s += stride) {
    const int v = R.perm[s];
    const int c = R.cellid[v];
    const int b = R.start[c], e = R.start[c + 1];
    const int key = R.orig_t[v];
    // eight cell members per round, their loads issued before any is used
    // (a cell of the 4,000-atom LJ box holds ~32 atoms: 8 dependent round
    // trips instead of 64)
    int rank = 0;
    for (int k0 = b; k0 < e; k0 += 8) {
      int m[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) m[u] = k0 + u < e ? R.perm[k0 + u] : v;
#pragma unroll
      for (int u = 0; u < 8; ++u) rank += R.orig_t[m[u]] < key ? 1 : 0;
    }
    const int d = b + rank;
    pos[d] = R.pos_t[v];
#pragma unroll
    for (int dd = 0; dd < 3; ++dd) {
      R.vel[dd * ncap + d] = R.vel_t[dd * ncap + v];
      R.acc[dd * ncap + d] = R.acc_t[dd * ncap + v];
    }
    R.type[d] = R.type_t[v];
    R.orig[d] = key;
    R.cellof[d] = c;
  }
}

// Arguments of the rebuild chain, resident in device memory: the kernels take
// a pointer (small kernel arguments launch faster, and most launches only
// read the rebuild flag and return).  The positions buffer alternates from
// step to step, so it comes as a kernel argument (R.pos is unused).
struct ChainArgs {
  CellArgs R;
  BuildArgs A;
};

// The binning phases; every launch returns at once when no rebuild is due.
// k_cell_bin decides (each workgroup evaluates the same test on the
// unchanged status; workgroup 0 records it), bins, and its last workgroup to
// finish scans the counts; k_cell_scatter books the decision (every kernel
// that tests rebuild_due has finished by then) and scatters; the later
// kernels read st->rebuild.
#define GH_BINBLOCK 1024  // binning workgroup: its last workgroup scans the counts
__global__ __launch_bounds__(GH_BINBLOCK) void k_cell_bin(GhStatus *st,
                                                        const ChainArgs *__restrict__ C,
                                                        double4 *pos, double skin,
                                                        long long step) {
  const bool due = !st->halt && rebuild_due(st, skin);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->rebuild = due ? 1 : 0;
  if (!due) return;
  const CellArgs &R = C->R;
  cell_bin(R, pos, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, step);
  __shared__ int last;
  __syncthreads();  // this workgroup's counts are issued
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(&st->ticket, 1) == (int)gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();  // every workgroup's counts are visible
  cell_scan<GH_BINBLOCK>(R);
  if (threadIdx.x == 0) st->ticket = 0;
}
__global__ __launch_bounds__(GH_TBLOCK) void k_cell_scatter(GhStatus *st,
                                                            const ChainArgs *__restrict__ C) {
  if (blockIdx.x == 0 && threadIdx.x == 0) rebuild_book(st);
  if (!st->rebuild) return;
  cell_scatter(C->R, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}
__global__ __launch_bounds__(GH_TBLOCK) void k_cell_rank_gather(GhStatus *st,
                                                                const ChainArgs *__restrict__ C,
                                                                double4 *pos) {
  if (!st->rebuild) return;
  cell_rank_gather(C->R, pos, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// Brick-local FP32 positions of the halo slots as three arrays.
__device__ __forceinline__ void stage_local(const Halo &H, const TileGeom &t,
                                            const double4 *__restrict__ pos,
                                             const unsigned char *scell, float *xf, float *yf,
                                             float *zf) {
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = scell[s];
    const double4 p = pos[s + H.delta[h]];
    xf[s] = (float)(p.x - (H.org[0] - shift(t, H, 0, h)));
    yf[s] = (float)(p.y - (H.org[1] - shift(t, H, 1, h)));
    zf[s] = (float)(p.z - (H.org[2] - shift(t, H, 2, h)));
  }
  __syncthreads();
}

// The workgroup's longest row folded into st->maxnn, one atomic per
// workgroup (one per wavefront queued thousands of atomics on one address at
// the end of a build); a row past the capacity flags the overflow and, when
// halt, stops the run.  All threads call it.
__device__ __forceinline__ void rows_max_atomic(int m, GhStatus *st, int cap, bool halt,
                                                long long step) {
  __shared__ int part[GH_TBMAX / 64];
  m = wave_imax(m);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (int)(blockDim.x >> 6);
    for (int w = 1; w < nw; ++w) m = max(m, part[w]);
    if (m > 0) {
      atomicMax(&st->maxnn, m);
      if (m > cap) {
        st->overflow = 1;
        if (halt) raise_halt(st, GH_HALT_OVERFLOW, step);
      }
    }
  }
}

// Rows of the atoms of brick b (all threads of the workgroup call it; LDS
// is free again on return).
__device__ void build_brick(GhStatus *st, const TileGeom &t, const double4 *__restrict__ pos,
                            const int *__restrict__ cellof, const int *__restrict__ cell_start,
                            const BuildArgs &A, float *dynf, Halo &H, int b, long long step) {
  if (!halo_setup(t, b, cell_start, H, st, true, step)) return;
  int mmax = 0;
  const int pad = t.slot_cap;  // sentinel slot of the force passes
  const int cap = A.cap, lds_slots = A.lds_slots;
  if (A.lds_rows) {
    // LDS: rows [(cap + 1) * GH_TBLOCK] u16, then float4 positions, slot map
    unsigned short *rowbuf = reinterpret_cast<unsigned short *>(dynf);
    float *xf = dynf + (cap + 1) * GH_TBLOCK / 2;
    if (H.total <= lds_slots) {
      float *yf = xf + lds_slots, *zf = yf + lds_slots;
      unsigned char *scell = reinterpret_cast<unsigned char *>(zf + lds_slots);
      slot_map(H, scell);
      stage_local(H, t, pos, scell, xf, yf, zf);
      build_rows_lds(H, t, pos, cellof, xf, yf, zf, scell, rowbuf, A, pad, mmax);
    } else {
      // halo too large for the slots left beside the rows: unstaged path
      unsigned short *far = rowbuf;  // GH_FARMAX * GH_TBLOCK fits in the row buffer
      build_rows<false>(H, t, pos, cellof, nullptr, nullptr, nullptr, far, A, pad, mmax);
    }
  } else {
    unsigned short *far = reinterpret_cast<unsigned short *>(dynf);
    float *xf = dynf + GH_FARMAX * GH_TBLOCK / 2;
    if (H.total <= lds_slots) {
      float *yf = xf + lds_slots, *zf = yf + lds_slots;
      unsigned char *scell = reinterpret_cast<unsigned char *>(zf + lds_slots);
      slot_map(H, scell);
      stage_local(H, t, pos, scell, xf, yf, zf);
      build_rows<true>(H, t, pos, cellof, xf, yf, zf, far, A, pad, mmax);
    } else {
      build_rows<false>(H, t, pos, cellof, nullptr, nullptr, nullptr, far, A, pad, mmax);
    }
  }
  rows_max_atomic(mmax, st, cap, cap > 0, step);  // cap 0: the sizing count
  __syncthreads();
}

__global__ __launch_bounds__(GH_TBLOCK) void k_build_tiled(GhStatus *st,
                                                           const ChainArgs *__restrict__ C,
                                                           const double4 *__restrict__ pos,
                                                           long long step) {
  if (!st->rebuild) return;
  extern __shared__ float dynf[];
  __shared__ Halo H;
  const CellArgs &R = C->R;
  const int b = xcd_block();
  if (b >= R.t.nbrick) return;
  build_brick(st, R.t, pos, R.cellof, R.start, C->A, dynf, H, b, step);
}

// ---------------------------------------------------------------------------
// Cooperative list build: a wavefront builds one atom's row at a time, its 64
// lanes testing 64 consecutive candidates per step.  The candidates of a
// brick cell (the slots of its nine runs, in the (dz, dy) order of the
// stencil, slots ascending within a run) are laid out once per workgroup as
// one LDS list, so a step reads 64 consecutive list entries and the halo
// positions they name (float4, brick-local FP32) without bank conflicts.
// Kept candidates go to the wavefront's row buffer in candidate order (ballot
// and lane prefix count); a second sweep makes the exact tests and tags each
// entry with its class, a third writes every entry to its place in class
// order.  The rows are the ones build_rows_lds writes: same prefilter, same
// exact tests, same order.  A workgroup takes a range of the brick's cells
// (several chunks per brick where bricks are fewer than the CUs), in batches
// whose candidate lists fit its LDS (a single cell always fits: its lists
// hold at most the halo's slots).
// ---------------------------------------------------------------------------
#define GH_CBLOCK 1024  // workgroup of the cooperative build: 16 wavefronts


// Halo cell of brick cell X (0 <= X < cpb).
__device__ __forceinline__ int halo_cell(const TileGeom &t, const Halo &H, int X) {
  const int lx = X % t.bd[0], ly = (X / t.bd[0]) % t.bd[1], lz = X / (t.bd[0] * t.bd[1]);
  return (lx + 1) + H.hd0 * ((ly + 1) + H.hd1 * (lz + 1));
}

__global__ __launch_bounds__(GH_CBLOCK) void k_build_coop(GhStatus *st,
                                                          const ChainArgs *__restrict__ C,
                                                          const double4 *__restrict__ pos,
                                                          long long step) {
  if (!st->rebuild) return;
  extern __shared__ float4 cl4[];
  __shared__ Halo H;
  __shared__ int coff[65];  // candidate-list offsets of the batch's cells (cpb <= 64)
  __shared__ int cst[65];   // their first atoms (cst[nfit]: one past the last)
  __shared__ int bend;      // the batch's end cell
  const CellArgs &R = C->R;
  const BuildArgs &A = C->A;
  const TileGeom &t = R.t;
  const int K = A.chunks;
  const int lin = xcd_block();
  const int b = lin / K, kc = lin - b * K;
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, R.start, H, st, true, step)) return;
  float4 *xyz = cl4;
  unsigned short *cand = reinterpret_cast<unsigned short *>(cl4 + A.nsl);
  unsigned short *rows = cand + A.ncand;
  unsigned char *scell = reinterpret_cast<unsigned char *>(rows + (GH_CBLOCK / 64) * A.rstride);
  slot_map(H, scell);
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = scell[s];
    const double4 p = pos[s + H.delta[h]];
    xyz[s] = make_float4((float)(p.x - (H.org[0] - shift(t, H, 0, h))),
                         (float)(p.y - (H.org[1] - shift(t, H, 1, h))),
                         (float)(p.z - (H.org[2] - shift(t, H, 2, h))), __int_as_float(s));
  }
  // the padding slot of the candidate lists: beyond every cut, never kept
  const int sent = A.nsl - 1;
  if (threadIdx.x == 0) xyz[sent] = make_float4(1e30f, 1e30f, 1e30f, __int_as_float(-1));
  // wavefront-uniform values go through readfirstlane, so that the loops
  // over a wavefront's atoms and candidates run on scalar registers
  const int lane = threadIdx.x & 63, nw = (int)(blockDim.x >> 6);
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  unsigned short *rw = rows + wv * A.rstride;
  const int cap = A.cap, pad = t.slot_cap;
  const float hi2f = A.hi2f;
  const size_t qstride = (size_t)A.ld << 2;
  const int X1 = (kc + 1) * t.cpb / K;
  int mmax = 0;
  for (int xb = kc * t.cpb / K; xb < X1;) {
    __syncthreads();  // the staging, or the previous batch, is complete
    if (wv == 0) {
      // candidate counts of cells xb.. and the cells whose lists fit
      const int X = xb + lane;
      int T = 0;
      if (X < X1) {
        const int ha = halo_cell(t, H, X);
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy) {
            const int h = ha + H.hd0 * (dy + H.hd1 * dz);
            T += H.hbase[h + 2] - H.hbase[h - 1];
          }
      }
      T = (T + 127) & ~127;  // padded to whole iterations of the candidate loop
      int inc = T;
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o, 64);
        if (lane >= o) inc += y;
      }
      const int nfit = max(__popcll(__ballot(X < X1 && inc <= A.ncand)), 1);
      // offsets of the batch's lists, and its cells' first atoms up to the
      // one past its last cell (cst[nfit], which lane 0 writes: nfit may be 64)
      if (lane < nfit) {
        coff[lane + 1] = inc;
        cst[lane] = R.start[H.key0 + xb + lane];
      }
      if (lane == 0) {
        coff[0] = 0;
        cst[nfit] = R.start[H.key0 + xb + nfit];
        bend = xb + nfit;
      }
    }
    __syncthreads();
    const int xe = __builtin_amdgcn_readfirstlane(bend);
    for (int i = wv; i < xe - xb; i += nw) {  // one wavefront per cell
      const int ha = halo_cell(t, H, xb + i);
      int off = coff[i];
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
          const int h = ha + H.hd0 * (dy + H.hd1 * dz);
          const int s0 = H.hbase[h - 1], len = H.hbase[h + 2] - s0;
          for (int j = lane; j < len; j += 64)
            if (off + j < A.ncand) cand[off + j] = (unsigned short)(s0 + j);
          off += len;
        }
      for (int j = off + lane; j < coff[i + 1]; j += 64)
        if (j < A.ncand) cand[j] = (unsigned short)sent;
    }
    __syncthreads();
    // the batch's atoms (contiguous: sorted by cell) split among the wavefronts
    const int aa0 = __builtin_amdgcn_readfirstlane(cst[0]);
    const int aa1 = __builtin_amdgcn_readfirstlane(cst[xe - xb]);
    const int per = (aa1 - aa0 + nw - 1) / nw;
    const int ab = aa0 + wv * per, ae = min(aa1, ab + per);
    int i = 0, dl = 0, cb = 0, tot = 0;  // the cell of atom a (batch index) and its lists
    int cnext = ab;                      // first atom past cell i
    for (int a = ab; a < ae; ++a) {
      if (a >= cnext) {
        while (__builtin_amdgcn_readfirstlane(cst[i + 1]) <= a) ++i;
        cnext = __builtin_amdgcn_readfirstlane(cst[i + 1]);
        dl = __builtin_amdgcn_readfirstlane(H.delta[halo_cell(t, H, xb + i)]);
        cb = __builtin_amdgcn_readfirstlane(coff[i]);
        tot = __builtin_amdgcn_readfirstlane(coff[i + 1]) - cb;
      }
      const int own = a - dl;
      const float4 pi = xyz[own];
      // sweep 1: FP32 prefilter of every candidate, kept ones in order;
      // two steps of 64 per iteration, so their LDS reads overlap
      // (lists padded with the far slot: no bounds test; the wave masks of
      // the compares are the ballots)
      int m = 0;
      const unsigned short *cl = cand + cb + lane;
      for (int c0 = 0; c0 < tot; c0 += 128) {
        const int sa = cl[c0], sc = cl[c0 + 64];
        const float4 pa4 = xyz[sa], pc4 = xyz[sc];
        float ux = pa4.x - pi.x, uy = pa4.y - pi.y, uz = pa4.z - pi.z;
        const float r2a = fmaf(ux, ux, fmaf(uy, uy, uz * uz));
        ux = pc4.x - pi.x, uy = pc4.y - pi.y, uz = pc4.z - pi.z;
        const float r2c = fmaf(ux, ux, fmaf(uy, uy, uz * uz));
        // (w holds the slot: the own test reads it, so each position is one
        // 16-byte read, conflict-free over consecutive slots)
        const unsigned long long ma =
            __builtin_amdgcn_fcmpf(r2a, hi2f, 5 /* ordered <= */) &
            __builtin_amdgcn_uicmp(__float_as_uint(pa4.w), (unsigned)own, 33 /* != */);
        const unsigned long long mc = __builtin_amdgcn_fcmpf(r2c, hi2f, 5) &
                                      __builtin_amdgcn_uicmp(__float_as_uint(pc4.w), (unsigned)own, 33);
        const int da = m + lane_rank(ma);
        if (__builtin_amdgcn_inverse_ballot_w64(ma) && da < cap) rw[da] = (unsigned short)sa;
        m += __popcll(ma);
        const int dc = m + lane_rank(mc);
        if (__builtin_amdgcn_inverse_ballot_w64(mc) && dc < cap) rw[dc] = (unsigned short)sc;
        m += __popcll(mc);
      }
      __builtin_amdgcn_wave_barrier();
      // entry k after the exact test near the cut (the reference's rounding
      // decides; a failure gives 0xFFFF), tagged with its class in bits 14-15
      auto classify = [&](int k) -> int {
        const int s = rw[k];
        const float4 pj = xyz[s];
        const float vx = pj.x - pi.x, vy = pj.y - pi.y, vz = pj.z - pi.z;
        const float r2f = fmaf(vx, vx, fmaf(vy, vy, vz * vz));
        int cls;
        if (r2f >= A.lo2f) {
          const double rd = pair_rd(H, t, pos, pos[a], s, scell[s]);
          if (!(rd < A.rl2)) return 0xFFFF;
          cls = row_class(rd, A);
        } else {
          cls = row_class_f(r2f, A);
          if (cls < 0) cls = row_class(pair_rd(H, t, pos, pos[a], s, scell[s]), A);
        }
        return s | (cls << 14);
      };
      unsigned short *row = A.list16 + ((size_t)a << 2);
      auto put = [&](int e, int d) {
        if (e != 0xFFFF) row[(size_t)(d >> 2) * qstride + (d & 3)] = (unsigned short)(e & 0x3FFF);
      };
      const int mk = min(m, cap);
      int mw;
      if (mk <= 64) {
        // one chunk: classify, count by class (wave masks of the class
        // bits), place
        const int e = lane < mk ? classify(lane) : 0xFFFF;
        const unsigned long long ok = __builtin_amdgcn_uicmp((unsigned)e, 0xFFFFu, 33);
        const unsigned long long hb = ok & __builtin_amdgcn_uicmp((unsigned)e & 0x8000u, 0u, 33);
        const unsigned long long lb = ok & __builtin_amdgcn_uicmp((unsigned)e & 0x4000u, 0u, 33);
        const unsigned long long q0 = ok & ~(hb | lb), q1 = lb & ~hb, q2 = hb & ~lb, q3 = hb & lb;
        const int n0 = __popcll(q0), n1 = __popcll(q1), n2 = __popcll(q2);
        mw = __popcll(ok);
        const int cls = e >> 14;
        const int d = cls == 0   ? lane_rank(q0)
                      : cls == 1 ? n0 + lane_rank(q1)
                      : cls == 2 ? n0 + n1 + lane_rank(q2)
                                 : n0 + n1 + n2 + lane_rank(q3);
        put(e, d);
      } else if (mk <= 128) {
        // two chunks held in registers: classify, count, place
        const int e0 = classify(lane);
        const int e1 = lane + 64 < mk ? classify(lane + 64) : 0xFFFF;
        const int c0 = e0 >> 14, c1 = e1 >> 14;
        const bool o0 = e0 != 0xFFFF, o1 = e1 != 0xFFFF;
        unsigned long long b0[4], b1[4];
        b0[0] = __ballot(o0 && c0 == 0);
        b0[1] = __ballot(o0 && c0 == 1);
        b0[2] = __ballot(o0 && c0 == 2);
        b0[3] = __ballot(o0 && c0 == 3);
        b1[0] = __ballot(o1 && c1 == 0);
        b1[1] = __ballot(o1 && c1 == 1);
        b1[2] = __ballot(o1 && c1 == 2);
        b1[3] = __ballot(o1 && c1 == 3);
        int base[4], run = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          base[q] = run;
          run += __popcll(b0[q]) + __popcll(b1[q]);
        }
        mw = run;
        int d0 = 0, d1 = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          if (c0 == q) d0 = base[q] + lane_rank(b0[q]);
          if (c1 == q) d1 = base[q] + __popcll(b0[q]) + lane_rank(b1[q]);
        }
        put(e0, d0);
        put(e1, d1);
      } else {
        // longer rows: the entries go through the row buffer twice
        int n0 = 0, n1 = 0, n2 = 0, n3 = 0;
        for (int k0 = 0; k0 < mk; k0 += 64) {
          const int k = k0 + lane;
          const int e = k < mk ? classify(k) : 0xFFFF;
          if (k < mk) rw[k] = (unsigned short)e;
          const bool ok = e != 0xFFFF;
          n0 += __popcll(__ballot(ok && (e >> 14) == 0));
          n1 += __popcll(__ballot(ok && (e >> 14) == 1));
          n2 += __popcll(__ballot(ok && (e >> 14) == 2));
          n3 += __popcll(__ballot(ok && (e >> 14) == 3));
        }
        __builtin_amdgcn_wave_barrier();
        int d0 = 0, d1 = n0, d2 = n0 + n1, d3 = n0 + n1 + n2;
        mw = d3 + n3;
        for (int k0 = 0; k0 < mk; k0 += 64) {
          const int k = k0 + lane;
          const int e = k < mk ? (int)rw[k] : 0xFFFF;
          const bool ok = e != 0xFFFF;
          const int cls = e >> 14;
          const unsigned long long q0 = __ballot(ok && cls == 0), q1 = __ballot(ok && cls == 1),
                                   q2 = __ballot(ok && cls == 2), q3 = __ballot(ok && cls == 3);
          const int d = cls == 0   ? d0 + lane_rank(q0)
                        : cls == 1 ? d1 + lane_rank(q1)
                        : cls == 2 ? d2 + lane_rank(q2)
                                   : d3 + lane_rank(q3);
          put(e, d);
          d0 += __popcll(q0);
          d1 += __popcll(q1);
          d2 += __popcll(q2);
          d3 += __popcll(q3);
        }
      }
      if (lane < ((4 - (mw & 3)) & 3)) {
        const int d = mw + lane;
        row[(size_t)(d >> 2) * qstride + (d & 3)] = (unsigned short)pad;
      }
      if (lane == 0) A.nn[a] = mw;
      mmax = max(mmax, m);  // before the exact drops: an overflow may have lost entries
      __builtin_amdgcn_wave_barrier();
    }
    // the build positions of the batch's atoms for the skin test
    for (int a = aa0 + (int)threadIdx.x; a < aa1; a += blockDim.x) {
      const double4 p = pos[a];
      A.xold[a] = p.x;
      A.xold[A.ncap + a] = p.y;
      A.xold[2 * A.ncap + a] = p.z;
    }
    xb = xe;
  }
  rows_max_atomic(mmax, st, cap, true, step);
}

// Largest halo and brick (atoms) over all bricks; sizes slot_cap.
__global__ __launch_bounds__(GH_TBLOCK) void k_tile_occupancy(TileGeom t,
                                                              const int *__restrict__ cell_start,
                                                              GhStatus *st) {
  __shared__ Halo H;
  const int b = blockIdx.x;
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, false)) return;
  if (threadIdx.x == 0) {
    atomicMax(&st->max_halo, H.total);
    atomicMax(&st->max_brick, H.a1 - H.a0);
  }
}

// Slot -> sorted atom index, for reading the list back on the host
// (dst is a plain ELL, entry k of atom a at k * ld + a).
__global__ __launch_bounds__(GH_TBLOCK) void k_tiled_to_global(
    TileGeom t, const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, int *__restrict__ dst, GhStatus *st) {
  __shared__ Halo H;
  const int b = blockIdx.x;
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true)) return;
  for (int a = H.a0 + threadIdx.x; a < H.a1; a += blockDim.x)
    for (int k = 0; k < nn[a]; ++k) {
      const int s = list16[qidx(k, a, ld)];
      dst[(size_t)k * ld + a] = s + H.delta[halo_of_slot(H, s)];
    }
}

// Brick of a force-pass workgroup: t.chunks consecutive workgroups (one
// XCD's) share a brick, each staging its halo and taking the contiguous share
// kc of its atoms (chunk_range).
__device__ __forceinline__ int pass_brick(const TileGeom &t, int &kc) {
  const int lin = xcd_block();
  const int b = lin / t.chunks;
  kc = lin - b * t.chunks;
  return b;
}
__device__ __forceinline__ void chunk_range(const TileGeom &t, const Halo &H, int kc, int &a0,
                                            int &a1) {
  const int na = H.a1 - H.a0;
  a0 = H.a0 + na * kc / t.chunks;
  a1 = H.a0 + na * (kc + 1) / t.chunks;
}

// Per-step arguments of the force passes.  The first pass of a step runs
// only if the run is not halted and the step's rebuild is not due (when the
// host has not enqueued the rebuild chain for the step, a due rebuild stops
// the run there; the host re-runs the step with the chain); its workgroup 0
// publishes the step (step_publish).
struct StepArgs {
  long long step;
  double skin;
  GhProgress *prog;  // null: nothing published
};
__device__ __forceinline__ bool step_gate(GhStatus *st, const StepArgs &sa) {
  const int h = halted(st);
  const bool stop = h || rebuild_due(st, sa.skin);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (stop && !h) raise_halt(st, GH_HALT_REBUILD, sa.step);
    step_publish(sa.prog, sa.step, h ? h : (stop ? GH_HALT_REBUILD : 0),
                 h ? st->halt_step : sa.step, st->maxdisp2_bits);
  }
  return stop;
}

// ---------------------------------------------------------------------------
// SMA-TB, mixed precision (fixed-point LDS, 2^t.fxk steps per Angstrom about
// the brick centre; the FP32 pair forms of TbF in that unit).  Pass 1:
// smatb_preinteraction (TB.F90:367-455); LDS per slot: qx, qy, qz in three
// arrays, then the slot map (and the type).  Stores rho (when asked) and
// d = 1/sqrt(rho) - e0 in FP32.
// S > 0: the slot arrays are static LDS of S slots (S >= slot_cap + 1), so a
// neighbour's three coordinates are one address (the slot shifted) and
// compile-time offsets; S = 0: dynamic LDS sized to the halo.
// ---------------------------------------------------------------------------
// A copy of x held in a vector register: the empty asm makes the value opaque,
// so the compiler keeps one VGPR copy instead of re-materialising it from its
// scalar register before every instruction that takes two scalar operands.
__device__ __forceinline__ float in_vgpr(float x) {
  float v;
  asm("" : "=v"(v) : "0"(x));
  return v;
}
// The one-type parameters with the addends of the FP32 pair forms in VGPRs.
__device__ __forceinline__ TbF tbf_vgpr(const TbF &p) {
  TbF v = p;
  v.rce = in_vgpr(p.rce);
  v.b2f = in_vgpr(p.b2f);
  v.b2x4 = in_vgpr(p.b2x4);
  v.r2m8 = in_vgpr(p.r2m8);
  v.c0 = in_vgpr(p.c0);
  return v;
}

template <bool SINGLE, int S, int G>
__global__ __launch_bounds__(GH_TBMAX) void k_tb_rho_fx(
    GhStatus *st, TileGeom t, const double4 *__restrict__ pos, const int *__restrict__ type,
    const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, const TbF *__restrict__ tbf,
    double *__restrict__ rho, float *__restrict__ ebd32, double e0, TbF p1, StepArgs sa) {
  if (step_gate(st, sa)) return;
  extern __shared__ int qa_dyn[];
  __shared__ int qa_st[S > 0 ? 3 * S + (S + 3) / 4 * (SINGLE ? 1 : 2) : 1];
  int *qa = S > 0 ? qa_st : qa_dyn;
  __shared__ Halo H;
  __shared__ TbF sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];  // one type: the parameters come in p1
  if (!SINGLE)
    for (int k = threadIdx.x; k < GH_MAXT * GH_MAXT; k += blockDim.x) sp[k] = tbf[k];
  int kc;  // this workgroup's share of the brick's atoms
  const int b = pass_brick(t, kc);
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true, sa.step)) return;
  const int ns = t.slot_cap + 1;  // slots + the sentinel
  const int st3 = S ? S : ns;     // stride of the three coordinate arrays
  int *qxa = qa, *qya = qa + st3, *qza = qa + 2 * st3;
  unsigned char *scell = reinterpret_cast<unsigned char *>(qa + 3 * st3);
  unsigned char *ty = scell + ns;
  slot_map(H, scell);
  const double fxs = t.fxs;
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = scell[s];
    const int src = s + H.delta[h];
    const double4 p = pos[src];
    qxa[s] = fxq(p.x - (H.org[0] - shift(t, H, 0, h)), fxs);
    qya[s] = fxq(p.y - (H.org[1] - shift(t, H, 1, h)), fxs);
    qza[s] = fxq(p.z - (H.org[2] - shift(t, H, 2, h)), fxs);
    if (!SINGLE) ty[s] = (unsigned char)type[src];
  }
  if (threadIdx.x == 0) {
    qxa[t.slot_cap] = qya[t.slot_cap] = qza[t.slot_cap] = GH_FXFAR;
    if (!SINGLE) ty[t.slot_cap] = 0;
  }
  __syncthreads();
  const TbF pv = tbf_vgpr(p1);
  const int sub = (int)(threadIdx.x & (G - 1));  // G lanes per atom (small systems)
  int ca0, ca1;
  chunk_range(t, H, kc, ca0, ca1);
  for (int a = ca0 + (int)(threadIdx.x / G); a < ca1; a += (int)(blockDim.x / G)) {
    const double4 pa = pos[a];
    const int qx = fxq(pa.x - H.org[0], fxs), qy = fxq(pa.y - H.org[1], fxs),
              qz = fxq(pa.z - H.org[2], fxs);
    const int ti = SINGLE ? 0 : type[a];
    double acc = 0.0;
    auto pair = [&](int jx, int jy, int jz, int s) {
      const double dx = fxd(jx, qx), dy = fxd(jy, qy), dz = fxd(jz, qz);
      float band;
      if (tb_rho_fx(SINGLE ? pv : sp[ti * GH_MAXT + ty[s]], fma(dx, dx, fma(dy, dy, dz * dz)),
                    band))
        acc += (double)band;
    };
    auto quad = [&](unsigned long long w) {
      const int s0 = GH_QSLOT(w, 0), s1 = GH_QSLOT(w, 1), s2 = GH_QSLOT(w, 2), s3 = GH_QSLOT(w, 3);
      const int x0 = qxa[s0], y0 = qya[s0], z0 = qza[s0];
      const int x1 = qxa[s1], y1 = qya[s1], z1 = qza[s1];
      const int x2 = qxa[s2], y2 = qya[s2], z2 = qza[s2];
      const int x3 = qxa[s3], y3 = qya[s3], z3 = qza[s3];
      pair(x0, y0, z0, s0);
      pair(x1, y1, z1, s1);
      pair(x2, y2, z2, s2);
      pair(x3, y3, z3, s3);
    };
    if (G == 1) for_each_quad(list16, ld, a, nn[a], quad);
    else for_each_quad_g<G>(list16, ld, a, nn[a], sub, quad);
    if (G > 1) acc = group_sum<G>(acc);
    if (sub == 0) {
      if (rho) rho[a] = acc;  // null on fused steps: only energy output reads it
      double yr = __builtin_amdgcn_rsq(acc);
      const double hr = 0.5 * acc;
      yr = yr * fma(-hr * yr, yr, 1.5);
      yr = yr * fma(-hr * yr, yr, 1.5);
      ebd32[a] = (float)(yr - e0);
    }
  }
}

// Pass 2: smatb_interact (TB.F90:457-586); LDS slot = {qx, qy, qz, d_j}
// (one 16-B read per neighbour: here fewer instructions outweigh the bank
// conflicts that make the density pass prefer separate arrays), in static
// LDS of S slots as in pass 1 (S = 0: dynamic).  FUSED: the
// velocity Verlet update of the atoms (no forces or energies stored, no
// repulsion summed); otherwise forces and per-atom energies are stored.
template <bool SINGLE, bool FUSED, int S, int G>
__device__ __forceinline__ void tb_force_fx_body(
    GhStatus *st, const TileGeom &t, const double4 *__restrict__ pos, const int *__restrict__ type,
    const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, int ncap, const TbF *__restrict__ tbf,
    const double *__restrict__ rho, const float *__restrict__ ebd32, double ev_ui,
    double *__restrict__ frc, double *__restrict__ epa, const FuseVV &fv, const TbF &p1,
    long long step) {
  extern __shared__ int4 q4_dyn[];
  __shared__ int4 q4_st[S > 0 ? S + (S + 15) / 16 * (SINGLE ? 1 : 2) : 1];
  int4 *q4 = S > 0 ? q4_st : q4_dyn;
  __shared__ Halo H;
  __shared__ TbF sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];  // one type: the parameters come in p1
  if (!SINGLE)
    for (int k = threadIdx.x; k < GH_MAXT * GH_MAXT; k += blockDim.x) sp[k] = tbf[k];
  int kc;  // this workgroup's share of the brick's atoms
  const int b = pass_brick(t, kc);
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true, step)) return;
  const int ns = t.slot_cap + 1;  // slots + the sentinel
  unsigned char *scell = reinterpret_cast<unsigned char *>(q4 + (S > 0 ? S : ns));
  unsigned char *ty = scell + (S > 0 ? S : ns);
  slot_map(H, scell);
  const double fxs = t.fxs;
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = scell[s];
    const int src = s + H.delta[h];
    const double4 p = pos[src];
    q4[s] = make_int4(fxq(p.x - (H.org[0] - shift(t, H, 0, h)), fxs),
                      fxq(p.y - (H.org[1] - shift(t, H, 1, h)), fxs),
                      fxq(p.z - (H.org[2] - shift(t, H, 2, h)), fxs), __float_as_int(ebd32[src]));
    if (!SINGLE) ty[s] = (unsigned char)type[src];
  }
  if (threadIdx.x == 0) {
    q4[t.slot_cap] = make_int4(GH_FXFAR, GH_FXFAR, GH_FXFAR, 0);
    if (!SINGLE) ty[t.slot_cap] = 0;
  }
  __syncthreads();
  // tb_force_fx returns factor / r per unit of the pair vector, so the sums
  // over integer pair vectors come out in eV/A
  double dmax2 = 0.0;
  const TbF pv = tbf_vgpr(p1);
  const int sub = (int)(threadIdx.x & (G - 1));  // G lanes per atom (small systems)
  int ca0, ca1;
  chunk_range(t, H, kc, ca0, ca1);
  for (int a = ca0 + (int)(threadIdx.x / G); a < ca1; a += (int)(blockDim.x / G)) {
    const double4 pa = pos[a];
    const int qx = fxq(pa.x - H.org[0], fxs), qy = fxq(pa.y - H.org[1], fxs),
              qz = fxq(pa.z - H.org[2], fxs);
    const int ti = SINGLE ? 0 : type[a];
    const float di = ebd32[a];
    VvLoad vl{};
    if (FUSED && G > 1 && sub == 0) vl = vv_load(fv, a);
    double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
    auto pair = [&](const int4 &q, int s) {
      const double dx = fxd(q.x, qx), dy = fxd(q.y, qy), dz = fxd(q.z, qz);
      float sf, repul;
      if (tb_force_fx<!FUSED>(SINGLE ? pv : sp[ti * GH_MAXT + ty[s]],
                              fma(dx, dx, fma(dy, dy, dz * dz)), di + __int_as_float(q.w), sf,
                              repul)) {
        const double sfd = (double)sf;
        fx = fma(-sfd, dx, fx);
        fy = fma(-sfd, dy, fy);
        fz = fma(-sfd, dz, fz);
        if (!FUSED) ei += (double)repul;
      }
    };
    auto quad = [&](unsigned long long w) {
      const int s0 = GH_QSLOT(w, 0), s1 = GH_QSLOT(w, 1), s2 = GH_QSLOT(w, 2), s3 = GH_QSLOT(w, 3);
      const int4 q0 = q4[s0], q1 = q4[s1], q2 = q4[s2], q3 = q4[s3];
      pair(q0, s0);
      pair(q1, s1);
      pair(q2, s2);
      pair(q3, s3);
    };
    if (G == 1) for_each_quad(list16, ld, a, nn[a], quad);
    else for_each_quad_g<G>(list16, ld, a, nn[a], sub, quad);
    if (G > 1) {
      fx = group_sum<G>(fx);
      fy = group_sum<G>(fy);
      fz = group_sum<G>(fz);
      if (!FUSED) ei = group_sum<G>(ei);
    }
    if (sub != 0) continue;
    const double Fx = fx * ev_ui, Fy = fy * ev_ui, Fz = fz * ev_ui;
    if (FUSED) {
      dmax2 = fmax(dmax2, G > 1 ? fused_vv(fv, a, pa, ti, Fx, Fy, Fz, vl)
                                : fused_vv(fv, a, pa, ti, Fx, Fy, Fz));
    } else {
      frc[a] = Fx;
      frc[ncap + a] = Fy;
      frc[2 * ncap + a] = Fz;
      epa[a] = (ei - sqrt(rho[a])) * ev_ui;
    }
  }
  if (FUSED) block_max_atomic(dmax2, fv.maxd2);
}
template <bool SINGLE, bool FUSED, int S, int G>
__global__ __launch_bounds__(GH_TBMAX) void k_tb_force_fx(
    GhStatus *st, TileGeom t, const double4 *__restrict__ pos, const int *__restrict__ type,
    const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, int ncap, const TbF *__restrict__ tbf,
    const double *__restrict__ rho, const float *__restrict__ ebd32, double ev_ui,
    double *__restrict__ frc, double *__restrict__ epa, FuseVV fv, TbF p1, StepArgs sa) {
  if (!halted(st))
    tb_force_fx_body<SINGLE, FUSED, S, G>(st, t, pos, type, cell_start, nn, list16, ld, ncap, tbf, rho,
                                    ebd32, ev_ui, frc, epa, fv, p1, sa.step);
}

// ---------------------------------------------------------------------------
// SMA-TB, plain FP64 (the reference's arithmetic; LDS holds FP64 positions
// with the image shift applied).
// ---------------------------------------------------------------------------
template <bool SINGLE, int G>
__global__ __launch_bounds__(GH_TBMAX) void k_tb_rho64_tiled(
    GhStatus *st, TileGeom t, const double4 *__restrict__ pos, const int *__restrict__ type,
    const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, const TbDev *__restrict__ tbd, int nt,
    double *__restrict__ rho, double *__restrict__ eband, float *__restrict__ eband32,
    StepArgs sa) {
  if (step_gate(st, sa)) return;
  extern __shared__ double dyn[];
  __shared__ Halo H;
  __shared__ TbDev sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];
  if (SINGLE) {
    if (threadIdx.x == 0) sp[0] = tbd[0];
  } else {
    load_params(sp, tbd, nt);
  }
  int kc;  // this workgroup's share of the brick's atoms
  const int b = pass_brick(t, kc);
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true, sa.step)) return;
  const int cap = t.slot_cap + 1;  // slots + the sentinel
  double *px = dyn, *py = dyn + cap, *pz = dyn + 2 * cap;
  int *ty = reinterpret_cast<int *>(dyn + 3 * cap);
  unsigned char *scell = reinterpret_cast<unsigned char *>(ty + (SINGLE ? 0 : cap));
  slot_map(H, scell);
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = scell[s];
    const int src = s + H.delta[h];
    const double4 p = pos[src];
    px[s] = p.x + shift(t, H, 0, h);
    py[s] = p.y + shift(t, H, 1, h);
    pz[s] = p.z + shift(t, H, 2, h);
    if (!SINGLE) ty[s] = type[src];
  }
  if (threadIdx.x == 0) {
    px[t.slot_cap] = py[t.slot_cap] = pz[t.slot_cap] = GH_FAR;
    if (!SINGLE) ty[t.slot_cap] = 0;
  }
  __syncthreads();
  const int sub = (int)(threadIdx.x & (G - 1));  // G lanes per atom (small systems)
  int ca0, ca1;
  chunk_range(t, H, kc, ca0, ca1);
  for (int a = ca0 + (int)(threadIdx.x / G); a < ca1; a += (int)(blockDim.x / G)) {
    const double4 pi = pos[a];
    const int ti = SINGLE ? 0 : type[a];
    double acc = 0.0;
    auto slot = [&](int s) {
      const double vx = px[s] - pi.x, vy = py[s] - pi.y, vz = pz[s] - pi.z;
      const TbDev &P = SINGLE ? sp[0] : sp[ti * GH_MAXT + ty[s]];
      const double dr = sqrt(dist2_rn(vx, vy, vz));
      if (dr <= P.rce) acc += tb_band64(P, dr);
    };
    if (G == 1) for_each_slot(list16, ld, a, nn[a], slot);
    else for_each_slot_g<G>(list16, ld, a, nn[a], sub, slot);
    if (G > 1) acc = group_sum<G>(acc);
    if (sub != 0) continue;
    if (rho) rho[a] = acc;  // null on fused steps: only energy output reads it
    const double eb = 1.0 / sqrt(acc);
    eband[a] = eb;
    eband32[a] = (float)eb;
  }
}

template <bool SINGLE, int G>
__device__ __forceinline__ void tb_force64_body(
    GhStatus *st, const TileGeom &t, const double4 *__restrict__ pos, const int *__restrict__ type,
    const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, int ncap, const TbDev *__restrict__ tbd,
    int nt, const double *__restrict__ rho, const double *__restrict__ eband, double ev_ui,
    double *__restrict__ frc, double *__restrict__ epa, const FuseVV &fv, long long step) {
  extern __shared__ double dyn[];
  __shared__ Halo H;
  __shared__ TbDev sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];
  if (SINGLE) {
    if (threadIdx.x == 0) sp[0] = tbd[0];
  } else {
    load_params(sp, tbd, nt);
  }
  int kc;  // this workgroup's share of the brick's atoms
  const int b = pass_brick(t, kc);
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true, step)) return;
  const int cap = t.slot_cap + 1;  // slots + the sentinel
  double *px = dyn, *py = dyn + cap, *pz = dyn + 2 * cap, *eb = dyn + 3 * cap;
  int *ty = reinterpret_cast<int *>(dyn + 4 * cap);
  unsigned char *scell = reinterpret_cast<unsigned char *>(ty + (SINGLE ? 0 : cap));
  slot_map(H, scell);
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = scell[s];
    const int src = s + H.delta[h];
    const double4 p = pos[src];
    px[s] = p.x + shift(t, H, 0, h);
    py[s] = p.y + shift(t, H, 1, h);
    pz[s] = p.z + shift(t, H, 2, h);
    eb[s] = eband[src];
    if (!SINGLE) ty[s] = type[src];
  }
  if (threadIdx.x == 0) {
    px[t.slot_cap] = py[t.slot_cap] = pz[t.slot_cap] = GH_FAR;
    eb[t.slot_cap] = 0.0;
    if (!SINGLE) ty[t.slot_cap] = 0;
  }
  __syncthreads();
  double dmax2 = 0.0;
  const int sub = (int)(threadIdx.x & (G - 1));  // G lanes per atom (small systems)
  int ca0, ca1;
  chunk_range(t, H, kc, ca0, ca1);
  for (int a = ca0 + (int)(threadIdx.x / G); a < ca1; a += (int)(blockDim.x / G)) {
    const double4 pi = pos[a];
    const int ti = SINGLE ? 0 : type[a];
    const double ebi = eband[a];
    double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
    auto slot = [&](int s) {
      const double vx = px[s] - pi.x, vy = py[s] - pi.y, vz = pz[s] - pi.z;
      const TbDev &P = SINGLE ? sp[0] : sp[ti * GH_MAXT + ty[s]];
      const double dr = sqrt(dist2_rn(vx, vy, vz));
      if (dr <= P.rce) {
        double factor, repul;
        tb_pair64(P, dr, ebi + eb[s], factor, repul);
        const double sf = factor / dr;
        fx -= sf * vx;
        fy -= sf * vy;
        fz -= sf * vz;
        ei += repul;
      }
    };
    if (G == 1) for_each_slot(list16, ld, a, nn[a], slot);
    else for_each_slot_g<G>(list16, ld, a, nn[a], sub, slot);
    if (G > 1) {
      fx = group_sum<G>(fx);
      fy = group_sum<G>(fy);
      fz = group_sum<G>(fz);
      ei = group_sum<G>(ei);
    }
    if (sub != 0) continue;
    const double Fx = fx * ev_ui, Fy = fy * ev_ui, Fz = fz * ev_ui;
    if (fv.on) {
      dmax2 = fmax(dmax2, fused_vv(fv, a, pi, ti, Fx, Fy, Fz));
    } else {
      frc[a] = Fx;
      frc[ncap + a] = Fy;
      frc[2 * ncap + a] = Fz;
      epa[a] = (ei - sqrt(rho[a])) * ev_ui;
    }
  }
  if (fv.on) block_max_atomic(dmax2, fv.maxd2);
}
template <bool SINGLE, int G>
__global__ __launch_bounds__(GH_TBMAX) void k_tb_force64_tiled(
    GhStatus *st, TileGeom t, const double4 *__restrict__ pos, const int *__restrict__ type,
    const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, int ncap, const TbDev *__restrict__ tbd,
    int nt, const double *__restrict__ rho, const double *__restrict__ eband, double ev_ui,
    double *__restrict__ frc, double *__restrict__ epa, FuseVV fv, StepArgs sa) {
  if (!halted(st))
    tb_force64_body<SINGLE, G>(st, t, pos, type, cell_start, nn, list16, ld, ncap, tbd, nt, rho,
                            eband, ev_ui, frc, epa, fv, sa.step);
}

// Lennard-Jones (lj_interact, Pairs.F90:207-279), tiled, FP64 LDS positions.
// POT: GH_POT_LJ (eps, sigma as arguments) or a tabulated form (plj, slj,
// wca: per type pair, types staged in LDS).
template <bool MIXED, int POT, bool FUSED, int G>
__device__ __forceinline__ void lj_tiled_body(
    GhStatus *st, const TileGeom &t, const double4 *__restrict__ pos,
    const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, int ncap, double eps, double sigma,
    double rc2, double ev_ui, double *__restrict__ frc, double *__restrict__ epa,
    const int *__restrict__ type, const FuseVV &fv, const PljDev *__restrict__ pljd,
    long long step) {
  extern __shared__ double dyn[];
  __shared__ Halo H;
  constexpr bool TAB = POT != GH_POT_LJ;
  __shared__ PljDev pp[TAB ? GH_MAXT * GH_MAXT : 1];
  if (TAB)
    for (int k = threadIdx.x; k < GH_MAXT * GH_MAXT; k += blockDim.x) pp[k] = pljd[k];
  int kc;  // this workgroup's share of the brick's atoms
  const int b = pass_brick(t, kc);
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true, step)) return;
  const int cap = t.slot_cap + 1;  // slots + the sentinel
  double *px = dyn, *py = dyn + cap, *pz = dyn + 2 * cap;
  unsigned char *scell = reinterpret_cast<unsigned char *>(dyn + 3 * cap);
  unsigned char *ty = scell + cap;  // tabulated forms: the slots' types
  slot_map(H, scell);
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = scell[s];
    const int src = s + H.delta[h];
    const double4 p = pos[src];
    px[s] = p.x + shift(t, H, 0, h);
    py[s] = p.y + shift(t, H, 1, h);
    pz[s] = p.z + shift(t, H, 2, h);
    if (TAB) ty[s] = (unsigned char)type[src];
  }
  if (threadIdx.x == 0) {
    px[t.slot_cap] = py[t.slot_cap] = pz[t.slot_cap] = GH_FAR;
    if (TAB) ty[t.slot_cap] = 0;
  }
  __syncthreads();
  const float epsf = (float)eps, sig2f = (float)(sigma * sigma);
  double dmax2 = 0.0;
  const int sub = (int)(threadIdx.x & (G - 1));  // G lanes per atom (small systems)
  int ca0, ca1;
  chunk_range(t, H, kc, ca0, ca1);
  for (int a = ca0 + (int)(threadIdx.x / G); a < ca1; a += (int)(blockDim.x / G)) {
    const double4 pi = pos[a];
    const int ti = type[a];
    VvLoad vl{};
    if (FUSED && G > 1 && sub == 0) vl = vv_load(fv, a);
    double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
    auto slot = [&](int s) {
      const double vx = px[s] - pi.x, vy = py[s] - pi.y, vz = pz[s] - pi.z;
      if (MIXED) {
        const double r2 = fma(vx, vx, fma(vy, vy, vz * vz));
        if (r2 <= rc2) {
          double sf;
          if (TAB) {
            const PljDev &P = pp[ti * GH_MAXT + ty[s]];
            if (r2 > P.rc2) return;  // this pair's cut
            float p, f_r;
            table_pair32<POT>(P, (float)r2, p, f_r);
            if (!FUSED) ei += (double)p;
            sf = (double)f_r;
          } else {
            const float ir2 = __builtin_amdgcn_rcpf((float)r2);
            const float s2 = sig2f * ir2;
            const float s6 = s2 * s2 * s2;
            if (!FUSED) ei += (double)(4.0f * epsf * (s6 * s6 - s6));
            sf = (double)(epsf * fmaf(48.0f, s6, -24.0f) * s6 * ir2);
          }
          fx = fma(-sf, vx, fx);
          fy = fma(-sf, vy, fy);
          fz = fma(-sf, vz, fz);
        }
      } else {
        const double r2 = dist2_rn(vx, vy, vz);
        if (r2 <= rc2) {
          double sf;
          if (TAB) {
            const PljDev &P = pp[ti * GH_MAXT + ty[s]];
            if (r2 > P.rc2) return;  // this pair's cut
            double p;
            table_pair64<POT>(P, r2, p, sf);
            if (!FUSED) ei += p;
          } else {
            const double dr = sqrt(r2);
            const double tt = sigma / dr;
            const double t2 = tt * tt, t4 = t2 * t2, s6 = t4 * t2;
            if (!FUSED) ei += 4.0 * eps * (s6 * s6 - s6);
            sf = eps * (48.0 * s6 - 24.0) * s6 / dr / dr;
          }
          fx -= sf * vx;
          fy -= sf * vy;
          fz -= sf * vz;
        }
      }
    };
    if (G == 1) for_each_slot(list16, ld, a, nn[a], slot);
    else for_each_slot_g<G>(list16, ld, a, nn[a], sub, slot);
    if (G > 1) {
      fx = group_sum<G>(fx);
      fy = group_sum<G>(fy);
      fz = group_sum<G>(fz);
      if (!FUSED) ei = group_sum<G>(ei);
    }
    if (sub != 0) continue;
    const double Fx = fx * ev_ui, Fy = fy * ev_ui, Fz = fz * ev_ui;
    if (FUSED) {
      dmax2 = fmax(dmax2, G > 1 ? fused_vv(fv, a, pi, ti, Fx, Fy, Fz, vl)
                                : fused_vv(fv, a, pi, ti, Fx, Fy, Fz));
    } else {
      frc[a] = Fx;
      frc[ncap + a] = Fy;
      frc[2 * ncap + a] = Fz;
      epa[a] = ei * 0.5 * ev_ui;
    }
  }
  if (FUSED) block_max_atomic(dmax2, fv.maxd2);
}
// One pass per step: gated (and publishing) as a first pass.
template <bool MIXED, int POT, bool FUSED, int G>
__global__ __launch_bounds__(GH_TBMAX) void k_lj_tiled(
    GhStatus *st, TileGeom t, const double4 *__restrict__ pos, const int *__restrict__ cell_start,
    const int *__restrict__ nn, const unsigned short *__restrict__ list16, int ld, int ncap,
    double eps, double sigma, double rc2, double ev_ui, double *__restrict__ frc,
    double *__restrict__ epa, const int *__restrict__ type, FuseVV fv,
    const PljDev *__restrict__ pljd, StepArgs sa) {
  if (!step_gate(st, sa))
    lj_tiled_body<MIXED, POT, FUSED, G>(st, t, pos, cell_start, nn, list16, ld, ncap, eps, sigma, rc2,
                              ev_ui, frc, epa, type, fv, pljd, sa.step);
}

// Lennard-Jones, mixed precision (lj_interact, Pairs.F90:207-279), on the
// fixed-point tiles of the SMA-TB passes: LDS slot = {qx, qy, qz, -} in steps
// of 2^-t.fxk A about the brick centre, one 16-byte read per neighbour; the
// pair vector is the exact integer difference and r^2 is exact in FP64, the
// pair terms are FP32 in the same unit (sig2 = sigma^2 and rc2 = rc^2 in
// steps^2), so f/r per step^2 summed over integer vectors, times 2^t.fxk,
// is the force in eV/A.  FUSED: the velocity Verlet update, no energies.
struct LjF {
  double rc2;      // rc^2 in steps^2
  float sig2, e4;  // sigma^2 in steps^2, 4 eps
  float e24, e48;  // 24 eps, 48 eps
};
template <bool FUSED, int G>
__global__ __launch_bounds__(GH_TBMAX) void k_lj_fx(
    GhStatus *st, TileGeom t, const double4 *__restrict__ pos, const int *__restrict__ cell_start,
    const int *__restrict__ nn, const unsigned short *__restrict__ list16, int ld, int ncap,
    LjF p, double ev_ui, double *__restrict__ frc, double *__restrict__ epa,
    const int *__restrict__ type, FuseVV fv, StepArgs sa) {
  if (step_gate(st, sa)) return;
  extern __shared__ int4 lq4[];
  __shared__ Halo H;
  int kc;  // this workgroup's share of the brick's atoms
  const int b = pass_brick(t, kc);
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true, sa.step)) return;
  const int ns = t.slot_cap + 1;  // slots + the sentinel
  unsigned char *scell = reinterpret_cast<unsigned char *>(lq4 + ns);
  slot_map(H, scell);
  const double fxs = t.fxs;
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = scell[s];
    const double4 q = pos[s + H.delta[h]];
    lq4[s] = make_int4(fxq(q.x - (H.org[0] - shift(t, H, 0, h)), fxs),
                       fxq(q.y - (H.org[1] - shift(t, H, 1, h)), fxs),
                       fxq(q.z - (H.org[2] - shift(t, H, 2, h)), fxs), 0);
  }
  if (threadIdx.x == 0) lq4[t.slot_cap] = make_int4(GH_FXFAR, GH_FXFAR, GH_FXFAR, 0);
  __syncthreads();
  const float sig2 = in_vgpr(p.sig2), e24 = p.e24, e48 = p.e48, e4 = p.e4;
  const double scale = fxs * ev_ui;
  double dmax2 = 0.0;
  const int sub = (int)(threadIdx.x & (G - 1));  // G lanes per atom (small systems)
  int ca0, ca1;
  chunk_range(t, H, kc, ca0, ca1);
  for (int a = ca0 + (int)(threadIdx.x / G); a < ca1; a += (int)(blockDim.x / G)) {
    const double4 pa = pos[a];
    const int qx = fxq(pa.x - H.org[0], fxs), qy = fxq(pa.y - H.org[1], fxs),
              qz = fxq(pa.z - H.org[2], fxs);
    VvLoad vl{};
    if (FUSED && G > 1 && sub == 0) vl = vv_load(fv, a);
    double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
    auto pair = [&](const int4 &q) {
      const double dx = fxd(q.x, qx), dy = fxd(q.y, qy), dz = fxd(q.z, qz);
      const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
      if (r2 <= p.rc2) {
        // 1/r^2: the hardware estimate and one Newton step on the exact FP64
        // r^2 (half an ulp instead of one plus the rounding of r^2 to FP32:
        // max|dF|/rms 1.2e-5 -> 9e-6 and |dE|/E 7e-9 -> 4e-11 at 131k atoms)
        float ir2 = __builtin_amdgcn_rcpf((float)r2);
        ir2 = fmaf(ir2, (float)fma(-r2, (double)ir2, 1.0), ir2);
        const float s2 = sig2 * ir2;
        const float s6 = s2 * s2 * s2;
        if (!FUSED) ei += (double)(e4 * (s6 * s6 - s6));
        const double sf = (double)(fmaf(e48, s6, -e24) * s6 * ir2);
        fx = fma(-sf, dx, fx);
        fy = fma(-sf, dy, fy);
        fz = fma(-sf, dz, fz);
      }
    };
    auto quad = [&](unsigned long long w) {
      const int4 q0 = lq4[GH_QSLOT(w, 0)], q1 = lq4[GH_QSLOT(w, 1)], q2 = lq4[GH_QSLOT(w, 2)],
                 q3 = lq4[GH_QSLOT(w, 3)];
      pair(q0);
      pair(q1);
      pair(q2);
      pair(q3);
    };
    if (G == 1) for_each_quad(list16, ld, a, nn[a], quad);
    else for_each_quad_g<G>(list16, ld, a, nn[a], sub, quad);
    if (G > 1) {
      fx = group_sum<G>(fx);
      fy = group_sum<G>(fy);
      fz = group_sum<G>(fz);
      if (!FUSED) ei = group_sum<G>(ei);
    }
    if (sub != 0) continue;
    const double Fx = fx * scale, Fy = fy * scale, Fz = fz * scale;
    if (FUSED) {
      dmax2 = fmax(dmax2, G > 1 ? fused_vv(fv, a, pa, type[a], Fx, Fy, Fz, vl)
                                : fused_vv(fv, a, pa, type[a], Fx, Fy, Fz));
    } else {
      frc[a] = Fx;
      frc[ncap + a] = Fy;
      frc[2 * ncap + a] = Fz;
      epa[a] = ei * 0.5 * ev_ui;
    }
  }
  if (FUSED) block_max_atomic(dmax2, fv.maxd2);
}

}  // namespace

// ---------------------------------------------------------------------------
// LDS budgets (bytes per halo slot, and static bytes) of the tiled passes.
// ---------------------------------------------------------------------------
size_t gh_tile_slot_bytes(int pass, int precision, int nt) {
  const bool multi = nt > 1;
  if (precision == GH_MIXED) {
    if (pass == 0) return 12 + 1 + (multi ? 1 : 0);  // rho: qx, qy, qz + map (+ type)
    return 16 + 1 + (multi ? 1 : 0);                 // force: int4 + map (+ type)
  }
  if (pass == 0) return 24 + 1 + (multi ? 4 : 0);  // rho: xyz + map (+ type)
  return 32 + 1 + (multi ? 4 : 0);                 // force: xyz + eband + map (+ type)
}
// Pair potentials: FP64 xyz + map per slot (+ the type for the tabulated
// forms); mixed-precision LJ: fixed-point int4 + map.
size_t gh_pair_slot_bytes(int pot, int precision) {
  if (pot == GH_POT_LJ && precision == GH_MIXED) return 16 + 1;
  return 25 + (gh_is_table_pot(pot) ? 1 : 0);
}
size_t gh_tile_static_bytes(int precision, int nt) {
  const size_t params = nt > 1 ? (precision == GH_MIXED ? sizeof(TbF) : sizeof(TbDev)) *
                                     GH_MAXT * GH_MAXT
                               : sizeof(TbDev);
  return sizeof(Halo) + params;
}

// LDS of the build: the far-entry buffer, then FP32 local coordinates (three
// arrays: 12 B; one float4 array reads with bank conflicts) and the slot map
// (1 B) per staged slot.
static size_t build_far_bytes() { return (size_t)GH_FARMAX * GH_TBLOCK * 2; }
static size_t build_slots() { return (GH_BUILD_LDS - build_far_bytes()) / 13; }

// Bounds of the pair function's inner branch and of its cutoff over all type
// pairs: rin = min rci, rout = max rce (SMA-TB); rin = rout = rc (LJ).
static void inner_outer(const gh_ctx *c, double *rin, double *rout) {
  if (c->pot == GH_POT_LJ || gh_is_table_pot(c->pot)) {
    *rin = *rout = c->lj_rc;
    return;
  }
  *rin = 1e300;
  *rout = 0.0;
  for (int a = 0; a < c->nt; ++a)
    for (int b = 0; b < c->nt; ++b) {
      *rin = std::min(*rin, c->tb[a * GH_MAXT + b].rci);
      *rout = std::max(*rout, c->tb[a * GH_MAXT + b].rce);
    }
}

static_assert(GH_BUILD_LDS / 13 < 16384, "staged slots must leave two tag bits in a row entry");

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// Build arguments and dynamic LDS of the tiled list build.
static BuildArgs build_args(const gh_ctx *c, int cap, size_t *lds_out) {
  BuildArgs A;
  int lds_slots = (int)std::min<size_t>(build_slots(), (size_t)c->tg.slot_cap);
  size_t lds = build_far_bytes() + 13 * (size_t)lds_slots;
  // rows in LDS when they and the staged halo fit (the branch-free variant;
  // 13 B per staged slot: three floats and the slot map)
  const size_t rows = ((size_t)(cap + 1) * GH_TBLOCK * 2 + 3) / 4 * 4;
  const size_t slots_left = rows < GH_BUILD_LDS ? (GH_BUILD_LDS - rows) / 13 : 0;
  int lds_rows = 0;
  if (slots_left >= (size_t)c->tg.slot_cap || slots_left >= 2048) {
    lds_rows = 1;
    lds_slots = (int)std::min<size_t>(slots_left, (size_t)c->tg.slot_cap);
    lds = rows + 13 * (size_t)lds_slots;
  }
  const double rl = c->rcut + c->skin;
  A.rl2 = rl * rl;
  // FP32 prefilter: |r^2_f32 - r^2| stays below 3e-4 A^2 for local
  // coordinates within 40 A of the brick centre; the margin is ten times that
  const double margin = std::max(3e-3, 1e-4 * A.rl2);
  A.lo2f = (float)(A.rl2 - margin);
  A.hi2f = (float)(A.rl2 + margin);
  // row-order classes: bounds at rin -+ m and rout + m, m = skin / 4
  double rin, rout;
  inner_outer(c, &rin, &rout);
  const double m = 0.25 * c->skin;
  A.cls2[0] = std::max(rin - m, 0.0) * std::max(rin - m, 0.0);
  A.cls2[1] = (rin + m) * (rin + m);
  A.cls2[2] = std::max((rout + m) * (rout + m), A.cls2[1]);
  for (int k = 0; k < 3; ++k) A.clsf[k] = (float)A.cls2[k];
  A.clsm = (float)margin;
  A.cap = cap;
  A.ld = c->ncap;
  A.ncap = c->ncap;
  A.lds_slots = lds_slots;
  A.lds_rows = lds_rows;
  A.nn = c->nn;
  A.list16 = c->list16;
  A.xold = c->xold;
  A.chunks = A.nsl = A.ncand = A.rstride = 0;
  *lds_out = lds;
  return A;
}

// LDS of the cooperative build: the staged halo and a far slot (float4 + slot
// map, 17 B per slot), the candidate lists, one row buffer of cap entries per
// wavefront.
// Sized like the per-thread build's, so that two workgroups share a CU.
#define GH_COOP_LDS GH_BUILD_LDS

// Sets the cooperative build's arguments; false where it does not apply: the
// sizing count (cap 0, any halo), slot numbers beyond the 14 bits a tagged
// row entry leaves, or LDS too small for one cell's candidate lists beside
// the staged halo.  Automatic choice: the per-thread build with its rows in
// LDS where bricks fill the chip (1024 and more: 1M-atom SMA-TB, 526 us
// against 563 us per rebuild), the cooperative build elsewhere (few bricks,
// or rows too long for LDS: the 1M-atom LJ liquid 824 us against 1241 us,
// 4000 atoms 37 us against 235 us).
static bool coop_args(const gh_ctx *c, BuildArgs &A, size_t *lds_out) {
  const int slot_cap = c->tg.slot_cap;
  if (c->list_rows == 1 || A.cap <= 0 || slot_cap >= 16000) return false;
  if (c->list_rows == 0 && A.lds_rows && c->tg.nbrick >= 1024) return false;
  const int nsl = (slot_cap + 4) / 4 * 4;  // + the far slot that pads the lists
  const int rstride = (A.cap + 1) / 2 * 2;
  const size_t fixed = (size_t)nsl * 17 + (size_t)(GH_CBLOCK / 64) * rstride * 2;
  // one cell's padded lists: at most the halo's slots + 127
  if (fixed + 2 * ((size_t)nsl + 128) > GH_COOP_LDS) return false;
  A.nsl = nsl;
  A.rstride = rstride;
  A.ncand = (int)((GH_COOP_LDS - fixed) / 2);
  // chunks of cells per brick: about two workgroups per CU when bricks are few
  const int nb = c->tg.nbrick;
  A.chunks = std::min(c->tg.cpb, std::max(1, (512 + nb - 1) / nb));
  *lds_out = fixed + 2 * (size_t)A.ncand;
  return true;
}


int gh_launch_rebuild_cells(gh_ctx *c) {
  ChainArgs h;
  std::memset(&h, 0, sizeof(h));  // padding too: the bytes are compared below
  CellArgs &R = h.R;
  R.st = c->status;
  R.t = c->tg;
  R.skin = c->skin;
  R.n = c->n;
  R.ncap = c->ncap;
  R.nkey = c->tg.nkey;
  R.pos = nullptr;  // alternates: passed to the kernels
  R.pos_t = c->pos_t;
  R.vel = c->vel;
  R.vel_t = c->vel_t;
  R.acc = c->acc;
  R.acc_t = c->acc_t;
  R.type = c->type;
  R.type_t = c->type_t;
  R.orig = c->orig;
  R.orig_t = c->orig_t;
  R.cellid = c->cellid;
  R.cellof = c->cellof;
  R.count = c->cell_count;
  R.start = c->cell_start;
  R.perm = c->perm;
  size_t lds = 0, lds_coop = 0;
  h.A = build_args(c, c->cap, &lds);
  const bool coop = coop_args(c, h.A, &lds_coop);
  hipStream_t s = c->stream;
  // upload on change only, ordered on the stream; the pinned staging copy is
  // rewritten only after the stream has drained
  if (c->chain_last.size() != sizeof(h) || std::memcmp(c->chain_last.data(), &h, sizeof(h)) != 0) {
    if (!c->chain_dev) {
      GH_HIP(c, hipMalloc(&c->chain_dev, sizeof(h)));
      GH_HIP(c, hipHostMalloc(&c->chain_pin, sizeof(h), hipHostMallocDefault));
    }
    GH_HIP(c, hipStreamSynchronize(s));
    std::memcpy(c->chain_pin, &h, sizeof(h));
    GH_HIP(c, hipMemcpyAsync(c->chain_dev, c->chain_pin, sizeof(h), hipMemcpyHostToDevice, s));
    c->chain_last.assign(reinterpret_cast<const unsigned char *>(&h),
                         reinterpret_cast<const unsigned char *>(&h) + sizeof(h));
  }
  const ChainArgs *C = static_cast<const ChainArgs *>(c->chain_dev);
  // grid-stride launches of at most 2048 workgroups: cheap when no rebuild is due
  const int ga = std::min((R.n + GH_TBLOCK - 1) / GH_TBLOCK, 2048);
  const int gb = std::min((R.n + GH_BINBLOCK - 1) / GH_BINBLOCK, 1024);
  { k_cell_bin<<<gb, GH_BINBLOCK, 0, s>>>(c->status, C, c->pos, c->skin, c->step); ++c->launches; }
  { k_cell_scatter<<<ga, GH_TBLOCK, 0, s>>>(c->status, C); ++c->launches; }
  { k_cell_rank_gather<<<ga, GH_TBLOCK, 0, s>>>(c->status, C, c->pos); ++c->launches; }
  c->list_build = coop ? 2 : 1;
  if (coop) {
    // 4 to 16 wavefronts: about 24 atoms each of a workgroup's mean share
    // where the workgroups fill the CUs, about two where they do not (each
    // CU then holds one workgroup, whose rows wait on their latencies)
    const long long units = (long long)c->tg.nbrick * h.A.chunks;
    const long long share = (long long)c->n / units + 1;
    const int apw = units < 256 ? 2 : 24;
    const int waves = (int)std::min<long long>(GH_CBLOCK / 64,
                                               std::max<long long>(4, (share + apw - 1) / apw));
    const int grid = xgrid(c->tg.nbrick * h.A.chunks);
    k_build_coop<<<grid, 64 * waves, lds_coop, s>>>(c->status, C, c->pos, c->step);
  } else {
    k_build_tiled<<<xgrid(c->tg.nbrick), GH_TBLOCK, lds, s>>>(c->status, C, c->pos, c->step);
  }
  ++c->launches;
  GH_HIP(c, hipGetLastError());
  return 0;
}

int gh_tile_occupancy(gh_ctx *c, int *max_halo, int *max_brick) {
  GhStatus h;
  GH_HIP(c, hipMemcpyAsync(&h, c->status, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  h.max_halo = 0;
  h.max_brick = 0;
  GH_HIP(c, hipMemcpyAsync(c->status, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
  { k_tile_occupancy<<<c->tg.nbrick, GH_TBLOCK, 0, c->stream>>>(c->tg, c->cell_start, c->status); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  GH_HIP(c, hipMemcpyAsync(&h, c->status, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  *max_halo = h.max_halo;
  *max_brick = h.max_brick;
  return 0;
}

int gh_tiled_to_global(gh_ctx *c, int *dst) {
  { k_tiled_to_global<<<c->tg.nbrick, GH_TBLOCK, 0, c->stream>>>(c->tg, c->cell_start, c->nn,
                                                               c->list16, c->ncap, dst,
                                                               c->status); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  return 0;
}

int gh_launch_forces_tiled(gh_ctx *c, hipEvent_t mid, bool fuse_vv) {
  hipStream_t s = c->stream;
  FuseVV fv;
  fv.on = fuse_vv ? 1 : 0;
  fv.dt = c->dt;
  fv.acel_dv = c->dt * 0.5;
  fv.acel_dx = fv.acel_dv * c->dt;
  for (int t = 0; t < GH_MAXT; ++t) fv.om[t] = 1.0 / c->mass_t[t];
  fv.vel = c->vel;
  fv.xold = c->xold;
  fv.pos_next = c->pos_next;
  fv.maxd2 = gh_maxd2(c);
  fv.ncap = c->ncap;
  const int grid = xgrid(c->tg.nbrick * c->tg.chunks);
  const int ld = c->ncap;
  const double ev_ui = ghc::ev_ui();
  const bool single = c->nt == 1;
  const bool mixed = c->precision == GH_MIXED;
  const size_t slots = (size_t)c->tg.slot_cap + 1;  // + the sentinel slot
  // a fused step stores no forces, per-atom energies or densities: the
  // steps whose state is observed run unfused and store them
  double *rho1 = fuse_vv ? nullptr : c->rho;
  StepArgs sa;
  sa.step = c->step;
  sa.skin = c->skin;
  sa.prog = c->prog_d;
  if (c->pot == GH_POT_SMATB) {
    const size_t lds1 = gh_tile_slot_bytes(0, c->precision, c->nt) * slots;
    const size_t lds2 = gh_tile_slot_bytes(1, c->precision, c->nt) * slots;
    if (mixed) {
      if (c->tbf_k != c->tg.fxk) GH_TRY(gh_upload_tbf(c, c->tg.fxk));
      const double e0 = c->tb_e0;
#define RHO_FX(SG, SS, GG)                                                                  \
  do {                                                                                      \
    const size_t lds = SS ? 0 : lds1;                                                        \
    k_tb_rho_fx<SG, SS, GG><<<grid, c->tg.tpb, lds, s>>>(c->status, c->tg, c->pos, c->type,  \
                                                     c->cell_start, c->nn, c->list16, ld,    \
                                                     c->tbf, rho1, c->eband32, e0, c->tbf0,  \
                                                     sa);                                    \
    ++c->launches;                                                                          \
  } while (0)
#define RHO_S(SG)                                                                           \
  do {                                                                                      \
    const int need = c->tg.slot_cap + 1;                                                    \
    if (c->tg.lanes == 4) RHO_FX(SG, 0, 4);                                                 \
    else if (need <= 768) RHO_FX(SG, 768, 1);                                               \
    else if (need <= 1024) RHO_FX(SG, 1024, 1);                                             \
    else if (need <= 1280) RHO_FX(SG, 1280, 1);                                             \
    else if (need <= 1536) RHO_FX(SG, 1536, 1);                                             \
    else if (need <= 1792) RHO_FX(SG, 1792, 1);                                             \
    else if (need <= 2048) RHO_FX(SG, 2048, 1);                                             \
    else if (need <= 2560) RHO_FX(SG, 2560, 1);                                             \
    else if (need <= 3072) RHO_FX(SG, 3072, 1);                                             \
    else RHO_FX(SG, 0, 1);                                                                  \
  } while (0)
      if (single) RHO_S(true);
      else RHO_S(false);
#undef RHO_S
#undef RHO_FX
      if (mid) GH_HIP(c, hipEventRecord(mid, s));
#define TB_FX3(SG, FU, SS, GG)                                                              \
  do {                                                                                      \
    k_tb_force_fx<SG, FU, SS, GG><<<grid, c->tg.tpb, SS ? 0 : lds2, s>>>(                    \
        c->status, c->tg, c->pos, c->type, c->cell_start, c->nn, c->list16, ld, c->ncap,    \
        c->tbf, c->rho, c->eband32, ev_ui, c->frc, c->epa, fv, c->tbf0, sa);                \
    ++c->launches;                                                                          \
  } while (0)
#define TB_FX(SG, FU)                                                                       \
  do {                                                                                      \
    const int need = c->tg.slot_cap + 1;                                                    \
    if (c->tg.lanes == 4) TB_FX3(SG, FU, 0, 4);                                             \
    else if (need <= 768) TB_FX3(SG, FU, 768, 1);                                           \
    else if (need <= 1024) TB_FX3(SG, FU, 1024, 1);                                         \
    else if (need <= 1280) TB_FX3(SG, FU, 1280, 1);                                         \
    else if (need <= 1536) TB_FX3(SG, FU, 1536, 1);                                         \
    else if (need <= 1792) TB_FX3(SG, FU, 1792, 1);                                         \
    else if (need <= 2048) TB_FX3(SG, FU, 2048, 1);                                         \
    else if (need <= 2560) TB_FX3(SG, FU, 2560, 1);                                         \
    else if (need <= 3072) TB_FX3(SG, FU, 3072, 1);                                         \
    else TB_FX3(SG, FU, 0, 1);                                                              \
  } while (0)
      if (single) {
        if (fuse_vv) TB_FX(true, true);
        else TB_FX(true, false);
      } else {
        if (fuse_vv) TB_FX(false, true);
        else TB_FX(false, false);
      }
#undef TB_FX
#undef TB_FX3
    } else {
#define RHO64(SG, GG)                                                                       \
  k_tb_rho64_tiled<SG, GG><<<grid, c->tg.tpb, lds1, s>>>(                                    \
      c->status, c->tg, c->pos, c->type, c->cell_start, c->nn, c->list16, ld, c->tbd, c->nt, \
      rho1, c->eband, c->eband32, sa)
#define FORCE64(SG, GG)                                                                     \
  k_tb_force64_tiled<SG, GG><<<grid, c->tg.tpb, lds2, s>>>(                                  \
      c->status, c->tg, c->pos, c->type, c->cell_start, c->nn, c->list16, ld, c->ncap,       \
      c->tbd, c->nt, c->rho, c->eband, ev_ui, c->frc, c->epa, fv, sa)
      const bool g4 = c->tg.lanes == 4;
      if (single) {
        if (g4) RHO64(true, 4);
        else RHO64(true, 1);
      } else {
        if (g4) RHO64(false, 4);
        else RHO64(false, 1);
      }
      ++c->launches;
      if (mid) GH_HIP(c, hipEventRecord(mid, s));
      if (single) {
        if (g4) FORCE64(true, 4);
        else FORCE64(true, 1);
      } else {
        if (g4) FORCE64(false, 4);
        else FORCE64(false, 1);
      }
      ++c->launches;
#undef RHO64
#undef FORCE64
    }
  } else if (c->pot == GH_POT_LJ || gh_is_table_pot(c->pot)) {
    if (mid) GH_HIP(c, hipEventRecord(mid, s));
    const double rc2 = c->lj_rc * c->lj_rc;
    const size_t lds = gh_pair_slot_bytes(c->pot, c->precision) * slots;
#define LJ_LAUNCH4(MX, PT, FU, GG)                                                          \
  k_lj_tiled<MX, PT, FU, GG><<<grid, c->tg.tpb, lds, s>>>(                                   \
      c->status, c->tg, c->pos, c->cell_start, c->nn, c->list16, ld, c->ncap, c->lj_eps,     \
      c->lj_sigma, rc2, ev_ui, c->frc, c->epa, c->type, fv, c->pljd, sa)
#define LJ_LAUNCH(MX, PT)                                                                   \
  do {                                                                                      \
    if (c->tg.lanes == 4) {                                                                 \
      if (fuse_vv) LJ_LAUNCH4(MX, PT, true, 4);                                             \
      else LJ_LAUNCH4(MX, PT, false, 4);                                                    \
    } else {                                                                                \
      if (fuse_vv) LJ_LAUNCH4(MX, PT, true, 1);                                             \
      else LJ_LAUNCH4(MX, PT, false, 1);                                                    \
    }                                                                                       \
    ++c->launches;                                                                          \
  } while (0)
#define LJ_POTS(MX)                                                                         \
  switch (c->pot) {                                                                         \
    case GH_POT_PLJ: LJ_LAUNCH(MX, GH_POT_PLJ); break;                                      \
    case GH_POT_SLJ: LJ_LAUNCH(MX, GH_POT_SLJ); break;                                      \
    case GH_POT_WCA: LJ_LAUNCH(MX, GH_POT_WCA); break;                                      \
    case GH_POT_SM1: LJ_LAUNCH(MX, GH_POT_SM1); break;                                      \
    default: LJ_LAUNCH(MX, GH_POT_LJ);                                                      \
  }
    if (mixed && c->pot == GH_POT_LJ) {
      LjF p;
      const double f2 = c->tg.fxs * c->tg.fxs;
      p.rc2 = rc2 * f2;
      p.sig2 = (float)(c->lj_sigma * c->lj_sigma * f2);
      p.e4 = (float)(4.0 * c->lj_eps);
      p.e24 = (float)(24.0 * c->lj_eps);
      p.e48 = (float)(48.0 * c->lj_eps);
#define LJFX(FU, GG)                                                                        \
  k_lj_fx<FU, GG><<<grid, c->tg.tpb, lds, s>>>(c->status, c->tg, c->pos, c->cell_start, c->nn, \
                                             c->list16, ld, c->ncap, p, ev_ui, c->frc,       \
                                             c->epa, c->type, fv, sa)
      if (c->tg.lanes == 4) {
        if (fuse_vv) LJFX(true, 4);
        else LJFX(false, 4);
      } else {
        if (fuse_vv) LJFX(true, 1);
        else LJFX(false, 1);
      }
#undef LJFX
      ++c->launches;
    } else if (mixed) {
      LJ_POTS(true)
    } else {
      LJ_POTS(false)
    }
#undef LJ_POTS
#undef LJ_LAUNCH
#undef LJ_LAUNCH4
  } else {
    return gh_fail(c, GH_ESTATE, "no potential set");
  }
  GH_HIP(c, hipGetLastError());
  return 0;
}
