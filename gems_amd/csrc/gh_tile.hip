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
// cell's atoms in sorted order.  hbase[h] is the first slot of halo cell h.
#include <math.h>

#include "gh_device.h"

#define GH_TBLOCK 256

namespace {
using namespace ghd;

struct Halo {
  int nh, total, a0, a1, key0, hd0, hd1;
  int hstart[GH_MAXHALO];
  int hbase[GH_MAXHALO + 1];
  int scan[GH_TBLOCK];
  double shx[GH_MAXHALO], shy[GH_MAXHALO], shz[GH_MAXHALO];
};

__device__ __forceinline__ int grid_key(const TileGeom &t, int cx, int cy, int cz) {
  const int bx = cx / t.bd[0], by = cy / t.bd[1], bz = cz / t.bd[2];
  const int lx = cx - bx * t.bd[0], ly = cy - by * t.bd[1], lz = cz - bz * t.bd[2];
  return ((bz * t.nbr[1] + by) * t.nbr[0] + bx) * t.cpb + (lz * t.bd[1] + ly) * t.bd[0] + lx;
}

// Halo of brick b: atom ranges, image shifts and LDS slot bases of the
// (bd+2)^3 halo cells.  Returns false, uniformly over the workgroup, if the
// brick holds no atoms or (check) its halo exceeds slot_cap.
__device__ bool halo_setup(const TileGeom &t, int b, const int *__restrict__ cell_start, Halo &H,
                           GhStatus *st, bool check) {
  const int bx = b % t.nbr[0], by = (b / t.nbr[0]) % t.nbr[1], bz = b / (t.nbr[0] * t.nbr[1]);
  const int hd0 = t.bd[0] + 2, hd1 = t.bd[1] + 2, hd2 = t.bd[2] + 2;
  const int nh = hd0 * hd1 * hd2;
  const int key0 = b * t.cpb;
  const int h = threadIdx.x;
  int cnt = 0;
  if (h < nh) {
    const int hx = h % hd0, hy = (h / hd0) % hd1, hz = h / (hd0 * hd1);
    int g[3] = {bx * t.bd[0] + hx - 1, by * t.bd[1] + hy - 1, bz * t.bd[2] + hz - 1};
    const int bc[3] = {bx, by, bz};
    double sh[3] = {0.0, 0.0, 0.0};
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int lo = bc[d] * t.bd[d];
      const int hi = min(lo + t.bd[d], t.gn[d]);  // one past the brick's last cell
      if (g[d] > hi) ok = false;                  // not adjacent to any cell of the brick
      if (t.pbc[d]) {
        if (g[d] < 0) {
          g[d] += t.nc[d];
          sh[d] = -t.L[d];
        } else if (g[d] >= t.nc[d]) {
          g[d] -= t.nc[d];
          sh[d] = t.L[d];
        }
      } else if (g[d] < 0 || g[d] >= t.gn[d]) {
        ok = false;
      }
    }
    int start = 0;
    if (ok) {
      const int key = grid_key(t, g[0], g[1], g[2]);
      start = cell_start[key];
      cnt = cell_start[key + 1] - start;
    }
    H.hstart[h] = start;
    H.shx[h] = sh[0];
    H.shy[h] = sh[1];
    H.shz[h] = sh[2];
  }
  // exclusive scan of the halo cell counts (Hillis-Steele over the block)
  H.scan[threadIdx.x] = cnt;
  __syncthreads();
  for (int o = 1; o < GH_TBLOCK; o <<= 1) {
    const int v = threadIdx.x >= o ? H.scan[threadIdx.x - o] : 0;
    __syncthreads();
    H.scan[threadIdx.x] += v;
    __syncthreads();
  }
  if (h < nh) H.hbase[h] = H.scan[h] - cnt;
  if (h == 0) {
    H.hbase[nh] = H.scan[nh - 1];
    H.total = H.scan[nh - 1];
    H.nh = nh;
    H.hd0 = hd0;
    H.hd1 = hd1;
    H.key0 = key0;
    H.a0 = cell_start[key0];
    H.a1 = cell_start[key0 + t.cpb];
  }
  __syncthreads();
  if (H.a1 <= H.a0) return false;
  if (check && H.total > t.slot_cap) {
    if (threadIdx.x == 0) st->tile_overflow = 1;
    return false;
  }
  return true;
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

__device__ __forceinline__ unsigned short ld_slot(const unsigned short *p) {
  return __builtin_nontemporal_load(p);
}

// ---------------------------------------------------------------------------
// Neighbour list (linkedcell_self, Neighbour.F90:991-1087): for each atom of
// the brick, the slots of its 27 neighbour cells with rd < rl^2, the atom
// itself excluded.  vd = (x_j - x_i) + image shift, rounded exactly as the
// reference's vdistance + dot_product, so the sets are bit-identical.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GH_TBLOCK) void k_build_tiled(
    GhStatus *st, TileGeom t, const double4 *__restrict__ pos, const int *__restrict__ cellof,
    const int *__restrict__ cell_start, double rl2, int cap, int ld, int ncap,
    int *__restrict__ nn, unsigned short *__restrict__ list16, double *__restrict__ xold) {
  if (!st->rebuild) return;
  extern __shared__ double dyn[];
  __shared__ Halo H;
  const int b = xcd_block();
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true)) return;
  double *px = dyn, *py = dyn + t.slot_cap, *pz = dyn + 2 * t.slot_cap;
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = halo_of_slot(H, s);
    const double4 p = pos[H.hstart[h] + (s - H.hbase[h])];
    px[s] = p.x;
    py[s] = p.y;
    pz[s] = p.z;
  }
  __syncthreads();
  int mmax = 0;
  for (int a = H.a0 + threadIdx.x; a < H.a1; a += blockDim.x) {
    const int lc = cellof[a] - H.key0;
    const int lx = lc % t.bd[0], ly = (lc / t.bd[0]) % t.bd[1], lz = lc / (t.bd[0] * t.bd[1]);
    const int ha = (lx + 1) + H.hd0 * ((ly + 1) + H.hd1 * (lz + 1));
    const int own = H.hbase[ha] + (a - H.hstart[ha]);
    const double xi = px[own], yi = py[own], zi = pz[own];
    int m = 0;
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int h = ha + dx + H.hd0 * (dy + H.hd1 * dz);
          const double sx = H.shx[h], sy = H.shy[h], sz = H.shz[h];
          const int se = H.hbase[h + 1];
          for (int s = H.hbase[h]; s < se; ++s) {
            if (s == own) continue;
            const double vx = (px[s] - xi) + sx;
            const double vy = (py[s] - yi) + sy;
            const double vz = (pz[s] - zi) + sz;
            if (dist2_rn(vx, vy, vz) < rl2) {
              if (m < cap) list16[(size_t)m * ld + a] = (unsigned short)s;
              ++m;
            }
          }
        }
    nn[a] = min(m, cap);
    xold[a] = xi;
    xold[ncap + a] = yi;
    xold[2 * ncap + a] = zi;
    mmax = max(mmax, m);
  }
  mmax = wave_imax(mmax);
  if ((threadIdx.x & 63) == 0 && mmax > 0) {
    atomicMax(&st->maxnn, mmax);
    if (mmax > cap) st->overflow = 1;
  }
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

// Slot -> sorted atom index, for reading the list back on the host.
__global__ __launch_bounds__(GH_TBLOCK) void k_tiled_to_global(
    TileGeom t, const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, int *__restrict__ dst, GhStatus *st) {
  __shared__ Halo H;
  const int b = blockIdx.x;
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true)) return;
  for (int a = H.a0 + threadIdx.x; a < H.a1; a += blockDim.x)
    for (int k = 0; k < nn[a]; ++k) {
      const int s = list16[(size_t)k * ld + a];
      const int h = halo_of_slot(H, s);
      dst[(size_t)k * ld + a] = H.hstart[h] + (s - H.hbase[h]);
    }
}

// Stages the halo's image-shifted positions (and optionally types) into LDS.
template <bool TYPES>
__device__ __forceinline__ void stage_pos(const Halo &H, const double4 *__restrict__ pos,
                                          const int *__restrict__ type, double *px, double *py,
                                          double *pz, int *ty) {
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = halo_of_slot(H, s);
    const int src = H.hstart[h] + (s - H.hbase[h]);
    const double4 p = pos[src];
    px[s] = p.x + H.shx[h];
    py[s] = p.y + H.shy[h];
    pz[s] = p.z + H.shz[h];
    if (TYPES) ty[s] = type[src];
  }
}

// ---------------------------------------------------------------------------
// SMA-TB pass 1 (smatb_preinteraction, TB.F90:367-455), tiled.
// MIXED: FP32 distance and exponential, FP64 pair vector and accumulation.
// ---------------------------------------------------------------------------
template <bool MIXED, bool SINGLE>
__global__ __launch_bounds__(GH_TBLOCK) void k_tb_rho_tiled(
    GhStatus *st, TileGeom t, const double4 *__restrict__ pos, const int *__restrict__ type,
    const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, const TbDev *__restrict__ tbd, int nt,
    double *__restrict__ rho, double *__restrict__ eband, float *__restrict__ eband32) {
  extern __shared__ double dyn[];
  __shared__ Halo H;
  __shared__ TbDev sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];
  if (SINGLE) {
    if (threadIdx.x == 0) sp[0] = tbd[0];
  } else {
    load_params(sp, tbd, nt);
  }
  const int b = xcd_block();
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true)) return;
  double *px = dyn, *py = dyn + t.slot_cap, *pz = dyn + 2 * t.slot_cap;
  int *ty = (int *)(dyn + 3 * t.slot_cap);
  stage_pos<!SINGLE>(H, pos, type, px, py, pz, ty);
  __syncthreads();
  for (int a = H.a0 + threadIdx.x; a < H.a1; a += blockDim.x) {
    const double4 pi = pos[a];
    const int ti = SINGLE ? 0 : type[a];
    const int m = nn[a];
    double acc = 0.0;
    for (int k = 0; k < m; ++k) {
      const int s = ld_slot(&list16[(size_t)k * ld + a]);
      const double vx = px[s] - pi.x, vy = py[s] - pi.y, vz = pz[s] - pi.z;
      const TbDev &P = SINGLE ? sp[0] : sp[ti * GH_MAXT + ty[s]];
      if (MIXED) {
        const double r2 = fma(vx, vx, fma(vy, vy, vz * vz));
        if (r2 > P.rce2) continue;
        const float r2f = (float)r2;
        const float r = r2f * __builtin_amdgcn_rsqf(r2f);
        float phi;
        if (r < P.rci_f) {
          phi = P.rho_a * __builtin_amdgcn_exp2f(P.rho_b * (r - P.r0f));
        } else {
          const float d = r - P.rce_f;
          const float pp = d * d * d * fmaf(d, fmaf(d, P.b1f, P.b2f), P.b3f);
          phi = pp * pp;
        }
        acc += (double)phi;
      } else {
        const double dr = sqrt(dist2_rn(vx, vy, vz));
        if (dr > P.rce) continue;
        double banda;
        if (dr < P.rci) {
          banda = P.eps2 * exp(-2.0 * P.q * (dr * P.one_r0 - 1.0));
        } else {
          const double d = dr - P.rce;
          const double d2 = d * d, d3 = d2 * d, d4 = d3 * d, d5 = d4 * d;
          banda = P.b1 * d5 + P.b2 * d4 + P.b3 * d3;
          banda = banda * banda;
        }
        acc += banda;
      }
    }
    rho[a] = acc;
    const double eb = 1.0 / sqrt(acc);
    eband[a] = eb;
    eband32[a] = (float)eb;
  }
}

// SMA-TB pass 2 (smatb_interact, TB.F90:457-586), tiled.
template <bool MIXED, bool SINGLE>
__global__ __launch_bounds__(GH_TBLOCK) void k_tb_force_tiled(
    GhStatus *st, TileGeom t, const double4 *__restrict__ pos, const int *__restrict__ type,
    const int *__restrict__ cell_start, const int *__restrict__ nn,
    const unsigned short *__restrict__ list16, int ld, int ncap, const TbDev *__restrict__ tbd,
    int nt, const double *__restrict__ rho, const double *__restrict__ eband,
    const float *__restrict__ eband32, double ev_ui, double *__restrict__ frc,
    double *__restrict__ epa) {
  extern __shared__ double dyn[];
  __shared__ Halo H;
  __shared__ TbDev sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];
  if (SINGLE) {
    if (threadIdx.x == 0) sp[0] = tbd[0];
  } else {
    load_params(sp, tbd, nt);
  }
  const int b = xcd_block();
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true)) return;
  const int cap = t.slot_cap;
  double *px = dyn, *py = dyn + cap, *pz = dyn + 2 * cap;
  double *eb64 = dyn + 3 * cap;                 // FP64 form
  float *eb32 = (float *)(dyn + 3 * cap);       // mixed form
  int *ty = MIXED ? (int *)(eb32 + cap) : (int *)(eb64 + cap);
  for (int s = threadIdx.x; s < H.total; s += blockDim.x) {
    const int h = halo_of_slot(H, s);
    const int src = H.hstart[h] + (s - H.hbase[h]);
    const double4 p = pos[src];
    px[s] = p.x + H.shx[h];
    py[s] = p.y + H.shy[h];
    pz[s] = p.z + H.shz[h];
    if (MIXED) eb32[s] = eband32[src];
    else eb64[s] = eband[src];
    if (!SINGLE) ty[s] = type[src];
  }
  __syncthreads();
  for (int a = H.a0 + threadIdx.x; a < H.a1; a += blockDim.x) {
    const double4 pi = pos[a];
    const int ti = SINGLE ? 0 : type[a];
    const int m = nn[a];
    double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
    if (MIXED) {
      const float ebi = eband32[a];
      for (int k = 0; k < m; ++k) {
        const int s = ld_slot(&list16[(size_t)k * ld + a]);
        const double vx = px[s] - pi.x, vy = py[s] - pi.y, vz = pz[s] - pi.z;
        const TbDev &P = SINGLE ? sp[0] : sp[ti * GH_MAXT + ty[s]];
        const double r2 = fma(vx, vx, fma(vy, vy, vz * vz));
        if (r2 > P.rce2) continue;
        const float r2f = (float)r2;
        const float rinv = __builtin_amdgcn_rsqf(r2f);
        const float r = r2f * rinv;
        const float ebs = ebi + eb32[s];
        float factor, repul;
        if (r <= P.rci_f) {
          const float s0 = r - P.r0f;
          repul = P.rep_a * __builtin_amdgcn_exp2f(P.rep_b * s0);
          const float band = P.rho_a * __builtin_amdgcn_exp2f(P.rho_b * s0);
          factor = fmaf(repul, P.frep, -band * P.fband * ebs);
        } else {
          const float d = r - P.rce_f;
          const float d2 = d * d, d3 = d2 * d;
          const float pp = d3 * fmaf(d, fmaf(d, P.b1f, P.b2f), P.b3f);
          const float dp = d2 * fmaf(d, fmaf(d, 5.0f * P.b1f, 4.0f * P.b2f), 3.0f * P.b3f);
          const float dr_ = d2 * fmaf(d, fmaf(d, 5.0f * P.r1f, 4.0f * P.r2f), 3.0f * P.r3f);
          factor = fmaf(pp * dp, ebs, -2.0f * dr_);
          repul = d3 * fmaf(d, fmaf(d, P.r1f, P.r2f), P.r3f);
        }
        const double sf = (double)(factor * rinv);
        fx = fma(-sf, vx, fx);
        fy = fma(-sf, vy, fy);
        fz = fma(-sf, vz, fz);
        ei += (double)repul;
      }
    } else {
      const double ebi = eband[a];
      for (int k = 0; k < m; ++k) {
        const int s = ld_slot(&list16[(size_t)k * ld + a]);
        const double vx = px[s] - pi.x, vy = py[s] - pi.y, vz = pz[s] - pi.z;
        const TbDev &P = SINGLE ? sp[0] : sp[ti * GH_MAXT + ty[s]];
        const double dr = sqrt(dist2_rn(vx, vy, vz));
        if (dr > P.rce) continue;
        double factor, repul;
        if (dr <= P.rci) {
          const double tt = dr * P.one_r0 - 1.0;
          repul = P.ca * exp(-P.p * tt);
          factor = 2.0 * repul * P.p_r0 -
                   P.eps2 * 0.5 * exp(-2.0 * P.q * tt) * P.q2_r0 * (ebi + eb64[s]);
        } else {
          const double d = dr - P.rce;
          const double d2 = d * d, d3 = d2 * d, d4 = d3 * d, d5 = d4 * d;
          factor = (P.b1 * d5 + P.b2 * d4 + P.b3 * d3) *
                   (5.0 * P.b1 * d4 + 4.0 * P.b2 * d3 + 3.0 * P.b3 * d2) * (ebi + eb64[s]);
          factor = factor - 2.0 * (5.0 * P.r1 * d4 + 4.0 * P.r2 * d3 + 3.0 * P.r3 * d2);
          repul = P.r1 * d5 + P.r2 * d4 + P.r3 * d3;
        }
        const double sf = factor / dr;
        fx -= sf * vx;
        fy -= sf * vy;
        fz -= sf * vz;
        ei += repul;
      }
    }
    frc[a] = fx * ev_ui;
    frc[ncap + a] = fy * ev_ui;
    frc[2 * ncap + a] = fz * ev_ui;
    epa[a] = (ei - sqrt(rho[a])) * ev_ui;
  }
}

// Lennard-Jones (lj_interact, Pairs.F90:207-279), tiled.
template <bool MIXED>
__global__ __launch_bounds__(GH_TBLOCK) void k_lj_tiled(
    GhStatus *st, TileGeom t, const double4 *__restrict__ pos, const int *__restrict__ cell_start,
    const int *__restrict__ nn, const unsigned short *__restrict__ list16, int ld, int ncap,
    double eps, double sigma, double rc2, double ev_ui, double *__restrict__ frc,
    double *__restrict__ epa) {
  extern __shared__ double dyn[];
  __shared__ Halo H;
  const int b = xcd_block();
  if (b >= t.nbrick) return;
  if (!halo_setup(t, b, cell_start, H, st, true)) return;
  double *px = dyn, *py = dyn + t.slot_cap, *pz = dyn + 2 * t.slot_cap;
  stage_pos<false>(H, pos, nullptr, px, py, pz, nullptr);
  __syncthreads();
  const float epsf = (float)eps, sig2f = (float)(sigma * sigma);
  for (int a = H.a0 + threadIdx.x; a < H.a1; a += blockDim.x) {
    const double4 pi = pos[a];
    const int m = nn[a];
    double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
    for (int k = 0; k < m; ++k) {
      const int s = ld_slot(&list16[(size_t)k * ld + a]);
      const double vx = px[s] - pi.x, vy = py[s] - pi.y, vz = pz[s] - pi.z;
      if (MIXED) {
        const double r2 = fma(vx, vx, fma(vy, vy, vz * vz));
        if (r2 > rc2) continue;
        const float ir2 = __builtin_amdgcn_rcpf((float)r2);
        const float s2 = sig2f * ir2;
        const float s6 = s2 * s2 * s2;
        ei += (double)(4.0f * epsf * (s6 * s6 - s6));
        const double sf = (double)(epsf * fmaf(48.0f, s6, -24.0f) * s6 * ir2);
        fx = fma(-sf, vx, fx);
        fy = fma(-sf, vy, fy);
        fz = fma(-sf, vz, fz);
      } else {
        const double r2 = dist2_rn(vx, vy, vz);
        if (r2 > rc2) continue;
        const double dr = sqrt(r2);
        const double tt = sigma / dr;
        const double t2 = tt * tt, t4 = t2 * t2, s6 = t4 * t2;
        ei += 4.0 * eps * (s6 * s6 - s6);
        const double sf = eps * (48.0 * s6 - 24.0) * s6 / dr / dr;
        fx -= sf * vx;
        fy -= sf * vy;
        fz -= sf * vz;
      }
    }
    frc[a] = fx * ev_ui;
    frc[ncap + a] = fy * ev_ui;
    frc[2 * ncap + a] = fz * ev_ui;
    epa[a] = ei * 0.5 * ev_ui;
  }
}

inline size_t lds_bytes_force(const gh_ctx *c) {
  const bool multi = c->nt > 1;
  const size_t per = 24 + (c->precision == GH_MIXED ? 4 : 8) + (multi ? 4 : 0);
  return per * (size_t)c->tg.slot_cap;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int gh_launch_build_tiled(gh_ctx *c, int cap, int count_only) {
  const size_t lds = 24 * (size_t)c->tg.slot_cap;
  k_build_tiled<<<xgrid(c->tg.nbrick), GH_TBLOCK, lds, c->stream>>>(
      c->status, c->tg, c->pos, c->cellof, c->cell_start, (c->rcut + c->skin) * (c->rcut + c->skin),
      count_only ? 0 : cap, c->ncap, c->ncap, c->nn, c->list16, c->xold);
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
  k_tile_occupancy<<<c->tg.nbrick, GH_TBLOCK, 0, c->stream>>>(c->tg, c->cell_start, c->status);
  GH_HIP(c, hipGetLastError());
  GH_HIP(c, hipMemcpyAsync(&h, c->status, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  GH_HIP(c, hipStreamSynchronize(c->stream));
  *max_halo = h.max_halo;
  *max_brick = h.max_brick;
  return 0;
}

int gh_tiled_to_global(gh_ctx *c, int *dst) {
  k_tiled_to_global<<<c->tg.nbrick, GH_TBLOCK, 0, c->stream>>>(c->tg, c->cell_start, c->nn,
                                                               c->list16, c->ncap, dst,
                                                               c->status);
  GH_HIP(c, hipGetLastError());
  return 0;
}

int gh_launch_forces_tiled(gh_ctx *c, hipEvent_t mid) {
  hipStream_t s = c->stream;
  const int grid = xgrid(c->tg.nbrick);
  const int ld = c->ncap;
  const double ev_ui = ghc::ev_ui();
  const bool single = c->nt == 1;
  const bool mixed = c->precision == GH_MIXED;
  const size_t lds_pos = (24 + (single ? 0 : 4)) * (size_t)c->tg.slot_cap;
  if (c->pot == GH_POT_SMATB) {
#define RHO(MX, SG)                                                                          \
  k_tb_rho_tiled<MX, SG><<<grid, GH_TBLOCK, lds_pos, s>>>(c->status, c->tg, c->pos, c->type, \
                                                          c->cell_start, c->nn, c->list16,   \
                                                          ld, c->tbd, c->nt, c->rho,         \
                                                          c->eband, c->eband32)
    if (mixed) {
      if (single) RHO(true, true); else RHO(true, false);
    } else {
      if (single) RHO(false, true); else RHO(false, false);
    }
#undef RHO
    if (mid) GH_HIP(c, hipEventRecord(mid, s));
    const size_t lds = lds_bytes_force(c);
#define FRC(MX, SG)                                                                     \
  k_tb_force_tiled<MX, SG><<<grid, GH_TBLOCK, lds, s>>>(                                \
      c->status, c->tg, c->pos, c->type, c->cell_start, c->nn, c->list16, ld, c->ncap,  \
      c->tbd, c->nt, c->rho, c->eband, c->eband32, ev_ui, c->frc, c->epa)
    if (mixed) {
      if (single) FRC(true, true); else FRC(true, false);
    } else {
      if (single) FRC(false, true); else FRC(false, false);
    }
#undef FRC
  } else if (c->pot == GH_POT_LJ) {
    if (mid) GH_HIP(c, hipEventRecord(mid, s));
    const double rc2 = c->lj_rc * c->lj_rc;
    const size_t lds = 24 * (size_t)c->tg.slot_cap;
    if (mixed)
      k_lj_tiled<true><<<grid, GH_TBLOCK, lds, s>>>(c->status, c->tg, c->pos, c->cell_start,
                                                    c->nn, c->list16, ld, c->ncap, c->lj_eps,
                                                    c->lj_sigma, rc2, ev_ui, c->frc, c->epa);
    else
      k_lj_tiled<false><<<grid, GH_TBLOCK, lds, s>>>(c->status, c->tg, c->pos, c->cell_start,
                                                     c->nn, c->list16, ld, c->ncap, c->lj_eps,
                                                     c->lj_sigma, rc2, ev_ui, c->frc, c->epa);
  } else {
    return gh_fail(c, GH_ESTATE, "no potential set");
  }
  GH_HIP(c, hipGetLastError());
  return 0;
}
