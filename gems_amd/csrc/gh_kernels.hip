// gfx950 kernels of the GEMS force-evaluation hot path.
//
// Data layout (HBM, cell-sorted order after each rebuild):
//   pos    double4[ncap]   x, y, z, w unused             32 B/atom
//   vel, acc, frc          3 x double[ncap] (SoA)         24 B/atom each
//   epa, rho, eband        double[ncap]
//   eband32                float[ncap] (mixed precision: 1/sqrt(rho) - e0)
//   type, orig, cellof     int[ncap]
//   list   uint32[cap][ld] ELL: row k of all atoms is contiguous, so a
//          wavefront reading entry k of 64 consecutive atoms is coalesced.
//          entry = j | (image code << 27); image code w = (wx+1)+3(wy+1)+9(wz+1)
//          with the neighbour's periodic image at x_j + wd*L_d.
// Wavefronts are 64 lanes; per-atom kernels use 256-thread workgroups.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "gh_device.h"


namespace {
using namespace ghd;


__global__ void k_fill_int(int *__restrict__ p, int v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Rebuild decision (test_update, Neighbour.F90:1305-1346).  The reference
// rebuilds when the two largest displacements since the last build sum past
// the skin.  Here the displacement is the true (minimum-image) one, and the
// test uses 2*max, which bounds d1+d2 from above.  Integrators fold each
// workgroup's max |x - x_old|^2 into GhStatus::maxdisp2_bits with one
// atomicMax (non-negative doubles order as their bit patterns); it runs over
// the steps since the last build and the first kernel of a rebuild sequence
// reads it (rebuild_due, rebuild_book).  State-changing kernels return at
// once while GhStatus::halt is set (raise_halt).
// ---------------------------------------------------------------------------
// Decision of the O(N^2) list mode (one thread).
__global__ void k_decide(GhStatus *st, double skin) {
  st->rebuild = !st->halt && rebuild_due(st, skin);
  rebuild_book(st);
}

struct GridArgs {
  int nc[3], gn[3], gofs[3], pbc[3];
  double L[3], one_cell[3];
  int ncell;
  int ghost;
};



// O(N^2) build (verlet_self, Neighbour.F90:676-759): rd > rl^2 -> skip, so
// equality is kept.  Minimum image by rounding as vdistance does; or, with
// ghost images (box mic F), every first-shell image x_j + w L, w in {-1,0,1}
// on each periodic axis, of every atom j including i itself (not w = 0), as
// the reference's ghost atoms are (pbcghost / pbcfullghost,
// Neighbour.F90:1520-1794).  Positions are wrapped into [0, L) just before a
// ghost-image build, as do_pbc does before the reference creates its ghosts,
// so the window is the reference's: it holds every image within the cut when
// each periodic side is at least the cut plus skin.  With pair masks (a
// one-sided `feels` interaction: a cross group with half and newton off,
// Interaction.f90:69-77, 128-131) atom i lists j only when
// mfeel[i] & mfelt[j] is nonzero; i then takes the whole pair force and half
// the pair energy, j no reaction.
__global__ __launch_bounds__(GH_BLOCK) void k_build_n2(
    GhStatus *st, const double4 *__restrict__ pos, int n, GridArgs g, double rl2,
    int cap, int ld, int count_only, int *__restrict__ nn, unsigned *__restrict__ list,
    double *__restrict__ xold, int ncap, const int *__restrict__ orig,
    const int *__restrict__ mfeel, const int *__restrict__ mfelt, long long step) {
  if (!st->rebuild) return;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  int m = 0;
  if (s < n) {
    const double4 pi = pos[s];
    // one-sided interactions: j enters i's row only if i feels what j is felt by
    const int fi = mfeel ? mfeel[orig[s]] : 0;
    const double ob0 = 1.0 / g.L[0], ob1 = 1.0 / g.L[1], ob2 = 1.0 / g.L[2];
    for (int j = 0; j < n && g.ghost; ++j) {
      if (mfeel && !(fi & mfelt[orig[j]])) continue;
      const double4 pj = pos[j];
      const double ux = pj.x - pi.x, uy = pj.y - pi.y, uz = pj.z - pi.z;
      for (int w = 0; w < 27; ++w) {
        const int wx = w % 3 - 1, wy = (w / 3) % 3 - 1, wz = w / 9 - 1;
        if ((wx && !g.pbc[0]) || (wy && !g.pbc[1]) || (wz && !g.pbc[2])) continue;
        if (j == s && w == 13) continue;
        const double rd = dist2_rn(ux + wx * g.L[0], uy + wy * g.L[1], uz + wz * g.L[2]);
        if (rd <= rl2) {
          if (!count_only && m < cap) list[(size_t)m * ld + s] = (unsigned)j | ((unsigned)w << 27);
          ++m;
        }
      }
    }
    for (int j = 0; j < n && !g.ghost; ++j) {
      if (j == s) continue;
      if (mfeel && !(fi & mfelt[orig[j]])) continue;
      const double4 pj = pos[j];
      double vx = pj.x - pi.x, vy = pj.y - pi.y, vz = pj.z - pi.z;
      int wx = 0, wy = 0, wz = 0;
      if (g.pbc[0]) { wx = -(int)round(vx * ob0); vx = vx + wx * g.L[0]; }
      if (g.pbc[1]) { wy = -(int)round(vy * ob1); vy = vy + wy * g.L[1]; }
      if (g.pbc[2]) { wz = -(int)round(vz * ob2); vz = vz + wz * g.L[2]; }
      const double rd = dist2_rn(vx, vy, vz);
      if (rd <= rl2) {
        if (!count_only && m < cap) {
          const unsigned code = (unsigned)((wx + 1) + 3 * (wy + 1) + 9 * (wz + 1)) << 27;
          list[(size_t)m * ld + s] = (unsigned)j | code;
        }
        ++m;
      }
    }
    if (!count_only) {
      nn[s] = min(m, cap);
      xold[s] = pi.x;
      xold[ncap + s] = pi.y;
      xold[2 * ncap + s] = pi.z;
    }
  }
  int wm = wave_imax(m);
  if ((threadIdx.x & 63) == 0 && wm > 0) {
    atomicMax(&st->maxnn, wm);
    if (wm > cap) {
      st->overflow = 1;
      if (cap > 0) raise_halt(st, GH_HALT_OVERFLOW, step);  // cap 0: the sizing count
    }
  }
}

// image shift table in LDS: shift[w] = (wx*L0, wy*L1, wz*L2)
struct ShiftLds {
  double x[27], y[27], z[27];
};
__device__ __forceinline__ void load_shift(ShiftLds &sh, double L0, double L1, double L2) {
  if (threadIdx.x < 27) {
    int w = threadIdx.x;
    sh.x[w] = (w % 3 - 1) * L0;
    sh.y[w] = ((w / 3) % 3 - 1) * L1;
    sh.z[w] = (w / 9 - 1) * L2;
  }
}

// ---------------------------------------------------------------------------
// SMA-TB pass 1 (smatb_preinteraction, TB.F90:367-455):
//   rho_i = sum_j [ r < rci ? xi^2 exp(-2q(r/r0-1)) : (b1 x^5 + b2 x^4 + b3 x^3)^2 ]
//   with x = r - rce and r > rce skipped; eband_i = 1/sqrt(rho_i).
// FP64 form.
// ---------------------------------------------------------------------------
template <bool SINGLE>
__global__ __launch_bounds__(GH_BLOCK) void k_tb_rho64(
    const double4 *__restrict__ pos, const int *__restrict__ type, const int *__restrict__ nn,
    const unsigned *__restrict__ list, int ld, int n, const TbDev *__restrict__ tbd, int nt,
    double L0, double L1, double L2, double *__restrict__ rho, double *__restrict__ eband,
    float *__restrict__ eband32) {
  __shared__ ShiftLds sh;
  __shared__ TbDev sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];
  load_shift(sh, L0, L1, L2);
  if (SINGLE) {
    if (threadIdx.x == 0) sp[0] = tbd[0];
  } else {
    load_params(sp, tbd, nt);
  }
  __syncthreads();
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 pi = pos[i];
  const int ti = SINGLE ? 0 : type[i];
  const int m = nn[i];
  double acc = 0.0;
  for (int k = 0; k < m; ++k) {
    const unsigned e = ld_stream(&list[(size_t)k * ld + i]);
    const int j = (int)(e & GH_JMASK);
    const int w = (int)(e >> 27);
    const double4 pj = pos[j];
    const double vx = (pj.x - pi.x) + sh.x[w];
    const double vy = (pj.y - pi.y) + sh.y[w];
    const double vz = (pj.z - pi.z) + sh.z[w];
    const TbDev &P = SINGLE ? sp[0] : sp[ti * GH_MAXT + type[j]];
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
  rho[i] = acc;
  const double eb = 1.0 / sqrt(acc);
  eband[i] = eb;
  eband32[i] = (float)eb;
}

// SMA-TB pass 2 (smatb_interact, TB.F90:457-586), full list, newton off:
//   r <= rci: repul = A e^{-p(r/r0-1)};
//             factor = 2 repul p/r0 - xi^2/2 e^{-2q(r/r0-1)} (2q/r0)(eb_i+eb_j)
//   rci < r <= rce: factor = P P' (eb_i+eb_j) - 2 R';  repul = R
//   F_i -= factor * vd / r  (vd = x_j - x_i), e_i = sum repul - sqrt(rho_i).
template <bool SINGLE>
__global__ __launch_bounds__(GH_BLOCK) void k_tb_force64(
    const double4 *__restrict__ pos, const int *__restrict__ type, const int *__restrict__ nn,
    const unsigned *__restrict__ list, int ld, int n, int ncap, const TbDev *__restrict__ tbd,
    int nt, double L0, double L1, double L2, const double *__restrict__ rho,
    const double *__restrict__ eband, double ev_ui, double *__restrict__ frc,
    double *__restrict__ epa) {
  __shared__ ShiftLds sh;
  __shared__ TbDev sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];
  load_shift(sh, L0, L1, L2);
  if (SINGLE) {
    if (threadIdx.x == 0) sp[0] = tbd[0];
  } else {
    load_params(sp, tbd, nt);
  }
  __syncthreads();
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 pi = pos[i];
  const int ti = SINGLE ? 0 : type[i];
  const double ebi = eband[i];
  const int m = nn[i];
  double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
  for (int k = 0; k < m; ++k) {
    const unsigned e = ld_stream(&list[(size_t)k * ld + i]);
    const int j = (int)(e & GH_JMASK);
    const int w = (int)(e >> 27);
    const double4 pj = pos[j];
    const double vx = (pj.x - pi.x) + sh.x[w];
    const double vy = (pj.y - pi.y) + sh.y[w];
    const double vz = (pj.z - pi.z) + sh.z[w];
    const TbDev &P = SINGLE ? sp[0] : sp[ti * GH_MAXT + type[j]];
    const double dr = sqrt(dist2_rn(vx, vy, vz));
    if (dr > P.rce) continue;
    double factor, repul;
    if (dr <= P.rci) {
      const double t = dr * P.one_r0 - 1.0;
      repul = P.ca * exp(-P.p * t);
      factor = 2.0 * repul * P.p_r0 -
               P.eps2 * 0.5 * exp(-2.0 * P.q * t) * P.q2_r0 * (ebi + eband[j]);
    } else {
      const double d = dr - P.rce;
      const double d2 = d * d, d3 = d2 * d, d4 = d3 * d, d5 = d4 * d;
      factor = (P.b1 * d5 + P.b2 * d4 + P.b3 * d3) *
               (5.0 * P.b1 * d4 + 4.0 * P.b2 * d3 + 3.0 * P.b3 * d2) * (ebi + eband[j]);
      factor = factor - 2.0 * (5.0 * P.r1 * d4 + 4.0 * P.r2 * d3 + 3.0 * P.r3 * d2);
      repul = P.r1 * d5 + P.r2 * d4 + P.r3 * d3;
    }
    const double s = factor / dr;
    fx -= s * vx;
    fy -= s * vy;
    fz -= s * vz;
    ei += repul;
  }
  frc[i] = fx * ev_ui;
  frc[ncap + i] = fy * ev_ui;
  frc[2 * ncap + i] = fz * ev_ui;
  epa[i] = (ei - sqrt(rho[i])) * ev_ui;
}

// ---------------------------------------------------------------------------
// Mixed precision SMA-TB: positions and the pair vector stay FP64 (the
// difference x_j - x_i is exact to FP64 rounding); the pair terms are the
// FP32 forms of TbDev::f (Angstrom units; r^2 in FP64); rho, energies and
// forces accumulate in FP64.  Pass 1 stores d = 1/sqrt(rho) - e0 in FP32.
// ---------------------------------------------------------------------------
template <bool SINGLE>
__global__ __launch_bounds__(GH_BLOCK) void k_tb_rho_mx(
    const double4 *__restrict__ pos, const int *__restrict__ type, const int *__restrict__ nn,
    const unsigned *__restrict__ list, int ld, int n, const TbDev *__restrict__ tbd, int nt,
    double L0, double L1, double L2, double *__restrict__ rho, double *__restrict__ eband,
    float *__restrict__ ebd32, double e0) {
  __shared__ ShiftLds sh;
  __shared__ TbDev sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];
  load_shift(sh, L0, L1, L2);
  if (SINGLE) {
    if (threadIdx.x == 0) sp[0] = tbd[0];
  } else {
    load_params(sp, tbd, nt);
  }
  __syncthreads();
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 pi = pos[i];
  const int ti = SINGLE ? 0 : type[i];
  const int m = nn[i];
  double acc = 0.0;
  for (int k = 0; k < m; ++k) {
    const unsigned e = ld_stream(&list[(size_t)k * ld + i]);
    const int j = (int)(e & GH_JMASK);
    const int w = (int)(e >> 27);
    const double4 pj = pos[j];
    const double vx = (pj.x - pi.x) + sh.x[w];
    const double vy = (pj.y - pi.y) + sh.y[w];
    const double vz = (pj.z - pi.z) + sh.z[w];
    const TbDev &P = SINGLE ? sp[0] : sp[ti * GH_MAXT + type[j]];
    float band;
    if (tb_rho_fx(P.f, fma(vx, vx, fma(vy, vy, vz * vz)), band)) acc += (double)band;
  }
  rho[i] = acc;
  const double eb = 1.0 / sqrt(acc);
  eband[i] = eb;
  ebd32[i] = (float)(eb - e0);
}

template <bool SINGLE>
__global__ __launch_bounds__(GH_BLOCK) void k_tb_force_mx(
    const double4 *__restrict__ pos, const int *__restrict__ type, const int *__restrict__ nn,
    const unsigned *__restrict__ list, int ld, int n, int ncap, const TbDev *__restrict__ tbd,
    int nt, double L0, double L1, double L2, const double *__restrict__ rho,
    const float *__restrict__ ebd32, double ev_ui, double *__restrict__ frc,
    double *__restrict__ epa) {
  __shared__ ShiftLds sh;
  __shared__ TbDev sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];
  load_shift(sh, L0, L1, L2);
  if (SINGLE) {
    if (threadIdx.x == 0) sp[0] = tbd[0];
  } else {
    load_params(sp, tbd, nt);
  }
  __syncthreads();
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 pi = pos[i];
  const int ti = SINGLE ? 0 : type[i];
  const float di = ebd32[i];
  const int m = nn[i];
  double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
  for (int k = 0; k < m; ++k) {
    const unsigned e = ld_stream(&list[(size_t)k * ld + i]);
    const int j = (int)(e & GH_JMASK);
    const int w = (int)(e >> 27);
    const double4 pj = pos[j];
    const double vx = (pj.x - pi.x) + sh.x[w];
    const double vy = (pj.y - pi.y) + sh.y[w];
    const double vz = (pj.z - pi.z) + sh.z[w];
    const TbDev &P = SINGLE ? sp[0] : sp[ti * GH_MAXT + type[j]];
    float sf, repul;
    if (!tb_force_fx<true>(P.f, fma(vx, vx, fma(vy, vy, vz * vz)), di + ebd32[j], sf, repul))
      continue;
    const double s = (double)sf;
    fx = fma(-s, vx, fx);
    fy = fma(-s, vy, fy);
    fz = fma(-s, vz, fz);
    ei += (double)repul;
  }
  frc[i] = fx * ev_ui;
  frc[ncap + i] = fy * ev_ui;
  frc[2 * ncap + i] = fz * ev_ui;
  epa[i] = (ei - sqrt(rho[i])) * ev_ui;
}

// ---------------------------------------------------------------------------
// Lennard-Jones (lj_interact, Pairs.F90:207-279), full list, newton off:
//   r^2 > rc^2 skipped; p = 4e(s6^2 - s6); f = e(48 s6 - 24) s6 / r;
//   e_i += p/2; F_i -= f vd / r.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GH_BLOCK) void k_lj64(
    const double4 *__restrict__ pos, const int *__restrict__ nn, const unsigned *__restrict__ list,
    int ld, int n, int ncap, double eps, double sigma, double rc2, double L0, double L1,
    double L2, double ev_ui, double *__restrict__ frc, double *__restrict__ epa) {
  __shared__ ShiftLds sh;
  load_shift(sh, L0, L1, L2);
  __syncthreads();
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 pi = pos[i];
  const int m = nn[i];
  double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
  for (int k = 0; k < m; ++k) {
    const unsigned e = ld_stream(&list[(size_t)k * ld + i]);
    const int j = (int)(e & GH_JMASK);
    const int w = (int)(e >> 27);
    const double4 pj = pos[j];
    const double vx = (pj.x - pi.x) + sh.x[w];
    const double vy = (pj.y - pi.y) + sh.y[w];
    const double vz = (pj.z - pi.z) + sh.z[w];
    const double r2 = dist2_rn(vx, vy, vz);
    if (r2 > rc2) continue;
    const double dr = sqrt(r2);
    const double t = sigma / dr;
    const double t2 = t * t, t4 = t2 * t2, s6 = t4 * t2;
    const double p = 4.0 * eps * (s6 * s6 - s6);
    const double f = eps * (48.0 * s6 - 24.0) * s6 / dr;
    ei += p;
    const double s = f / dr;
    fx -= s * vx;
    fy -= s * vy;
    fz -= s * vz;
  }
  frc[i] = fx * ev_ui;
  frc[ncap + i] = fy * ev_ui;
  frc[2 * ncap + i] = fz * ev_ui;
  epa[i] = ei * 0.5 * ev_ui;
}

// Tabulated Lennard-Jones forms (plj_interact, Pairs.F90:293-370;
// slj_interact, :403-493; wca_interact, :508-560) on the O(N^2)-mode list,
// newton off, per type pair: e_i += p/2; F_i -= f vd / r.
template <bool MIXED, int POT>
__global__ __launch_bounds__(GH_BLOCK) void k_plj(
    const double4 *__restrict__ pos, const int *__restrict__ type, const int *__restrict__ nn,
    const unsigned *__restrict__ list, int ld, int n, int ncap, const PljDev *__restrict__ pljd,
    double rc2, double L0, double L1, double L2, double ev_ui, double *__restrict__ frc,
    double *__restrict__ epa) {
  __shared__ ShiftLds sh;
  __shared__ PljDev pp[GH_MAXT * GH_MAXT];
  load_shift(sh, L0, L1, L2);
  for (int k = threadIdx.x; k < GH_MAXT * GH_MAXT; k += blockDim.x) pp[k] = pljd[k];
  __syncthreads();
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 pi = pos[i];
  const int ti = type[i], m = nn[i];
  double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
  for (int k = 0; k < m; ++k) {
    const unsigned e = ld_stream(&list[(size_t)k * ld + i]);
    const int j = (int)(e & GH_JMASK);
    const int w = (int)(e >> 27);
    const double4 pj = pos[j];
    const double vx = (pj.x - pi.x) + sh.x[w];
    const double vy = (pj.y - pi.y) + sh.y[w];
    const double vz = (pj.z - pi.z) + sh.z[w];
    const double r2 = MIXED ? fma(vx, vx, fma(vy, vy, vz * vz)) : dist2_rn(vx, vy, vz);
    if (r2 > rc2) continue;
    const PljDev &P = pp[ti * GH_MAXT + type[j]];
    if (r2 > P.rc2) continue;  // this pair's cut
    double p, sf;
    if (MIXED) {
      float pf, fr;
      table_pair32<POT>(P, (float)r2, pf, fr);
      p = (double)pf;
      sf = (double)fr;
    } else {
      table_pair64<POT>(P, r2, p, sf);
    }
    ei += p;
    fx -= sf * vx;
    fy -= sf * vy;
    fz -= sf * vz;
  }
  frc[i] = fx * ev_ui;
  frc[ncap + i] = fy * ev_ui;
  frc[2 * ncap + i] = fz * ev_ui;
  epa[i] = ei * 0.5 * ev_ui;
}

__global__ __launch_bounds__(GH_BLOCK) void k_lj_mx(
    const double4 *__restrict__ pos, const int *__restrict__ nn, const unsigned *__restrict__ list,
    int ld, int n, int ncap, float eps, float sigma2, double rc2, double L0, double L1,
    double L2, double ev_ui, double *__restrict__ frc, double *__restrict__ epa) {
  __shared__ ShiftLds sh;
  load_shift(sh, L0, L1, L2);
  __syncthreads();
  const int i = xcd_block() * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double4 pi = pos[i];
  const int m = nn[i];
  double fx = 0.0, fy = 0.0, fz = 0.0, ei = 0.0;
  for (int k = 0; k < m; ++k) {
    const unsigned e = ld_stream(&list[(size_t)k * ld + i]);
    const int j = (int)(e & GH_JMASK);
    const int w = (int)(e >> 27);
    const double4 pj = pos[j];
    const double vx = (pj.x - pi.x) + sh.x[w];
    const double vy = (pj.y - pi.y) + sh.y[w];
    const double vz = (pj.z - pi.z) + sh.z[w];
    const double r2 = fma(vx, vx, fma(vy, vy, vz * vz));
    if (r2 > rc2) continue;
    const float ir2 = __builtin_amdgcn_rcpf((float)r2);
    const float s2 = sigma2 * ir2;
    const float s6 = s2 * s2 * s2;
    const float p = 4.0f * eps * (s6 * s6 - s6);
    const float fr2 = eps * fmaf(48.0f, s6, -24.0f) * s6 * ir2;  // f / r
    ei += (double)p;
    const double s = (double)fr2;
    fx = fma(-s, vx, fx);
    fy = fma(-s, vy, fy);
    fz = fma(-s, vz, fz);
  }
  frc[i] = fx * ev_ui;
  frc[ncap + i] = fy * ev_ui;
  frc[2 * ncap + i] = fz * ev_ui;
  epa[i] = ei * 0.5 * ev_ui;
}

// ---------------------------------------------------------------------------
// Integrators.  Per-type masses come in the kernel argument block.
// ---------------------------------------------------------------------------
struct MassArgs {
  double om[GH_MAXT];   // 1/m
  double osm[GH_MAXT];  // sqrt(1/m)
};

struct WrapArgs {
  int pbc[3];
  double L[3];
};

// Positions are not wrapped between neighbour rebuilds: the list stores each
// pair's periodic image, which stays valid while atoms move continuously.
// They are wrapped into the box at every rebuild (k_wrap_rebuild) and when
// copied out, which reproduces do_pbc_1 (Set_Properties.f90:444-470).
// Workgroup maximum of |x - x_old|^2 folded into the device-wide maximum with
// one atomicMax (no fence: the next kernel reads it).
__device__ __forceinline__ void disp_reduce(double d2, unsigned long long *__restrict__ maxd2) {
  __shared__ double part[GH_BLOCK / 64];
  d2 = wave_max(d2);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = d2;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < GH_BLOCK / 64; ++w) d2 = fmax(d2, part[w]);
    atomicMax(maxd2, (unsigned long long)__double_as_longlong(d2));
  }
}

// velocity_verlet_a (Integration.F90:1150-1169) + do_pbc + displacement.
__global__ __launch_bounds__(GH_BLOCK) void k_vv_a(
    int n, int ncap, double dt, MassArgs ma,
    const int *__restrict__ type, double4 *__restrict__ pos, double *__restrict__ vel,
    double *__restrict__ acc, const double *__restrict__ frc, const double *__restrict__ xold,
    GhStatus *st) {
  if (halted(st)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (i < n) {
    const double acel_dv = dt * 0.5, acel_dx = acel_dv * dt;
    const double om = ma.om[type[i]];
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double a = frc[d * ncap + i] * om;
      const double v = vel[d * ncap + i];
      x[d] = x[d] + dt * v + acel_dx * a;
      vel[d * ncap + i] = v + acel_dv * a;
      acc[d * ncap + i] = a;
      const double dd = x[d] - xold[d * ncap + i];
      d2 += dd * dd;
    }
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
    pos[i] = p;
  }
  disp_reduce(d2, &st->maxdisp2_bits);
}

// velocity_verlet_b of one step fused with velocity_verlet_a of the next
// (Integration.F90:1171-1186 then 1150-1169): the same operations in the same
// order as the two kernels, with a = F/m formed once.  The acceleration is
// not stored (stepa recomputes it; the last stepb of a run stores it).
__global__ __launch_bounds__(GH_BLOCK) void k_vv_ba(
    int n, int ncap, double dt, MassArgs ma,
    const int *__restrict__ type, double4 *__restrict__ pos, double *__restrict__ vel,
    const double *__restrict__ frc, const double *__restrict__ xold, GhStatus *st) {
  if (halted(st)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (i < n) {
    const double acel_dv = dt * 0.5, acel_dx = acel_dv * dt;
    const double om = ma.om[type[i]];
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double a = frc[d * ncap + i] * om;
      const double v = vel[d * ncap + i] + acel_dv * a;  // stepb
      x[d] = x[d] + dt * v + acel_dx * a;                // stepa
      vel[d * ncap + i] = v + acel_dv * a;
      const double dd = x[d] - xold[d * ncap + i];
      d2 += dd * dd;
    }
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
    pos[i] = p;
  }
  disp_reduce(d2, &st->maxdisp2_bits);
}

// velocity_verlet_b (Integration.F90:1171-1186).
__global__ __launch_bounds__(GH_BLOCK) void k_vv_b(int n, int ncap, double dt, MassArgs ma,
                                                   const int *__restrict__ type,
                                                   double *__restrict__ vel,
                                                   double *__restrict__ acc,
                                                   const double *__restrict__ frc,
                                                   const GhStatus *st) {
  if (halted(st)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double acel_dv = dt * 0.5;
  const double om = ma.om[type[i]];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double a = frc[d * ncap + i] * om;
    acc[d * ncap + i] = a;
    vel[d * ncap + i] = vel[d * ncap + i] + acel_dv * a;
  }
}

struct ErmakArgs {
  double c0, c1, c2, sdr, sdv, crv1, crv2, skt, dt;
  int philox;
  unsigned long long seed;
  long long step;
};

__device__ __forceinline__ void ermak_gauss(const ErmakArgs &k, int id,
                                            const double *__restrict__ gauss, int ncap,
                                            double g[6]) {
  if (k.philox) {
    philox_normals(k.seed, k.step, id, g);
  } else {
#pragma unroll
    for (int q = 0; q < 6; ++q) g[q] = gauss[(size_t)q * ncap + id];
  }
}

// ermak_a (Integration.F90:963-1011) + do_pbc + displacement.
__global__ __launch_bounds__(GH_BLOCK) void k_ermak_a(
    int n, int ncap, ErmakArgs k, MassArgs ma,
    const int *__restrict__ type,
    const int *__restrict__ orig, const double *__restrict__ gauss, double4 *__restrict__ pos,
    const double *__restrict__ vel, const double *__restrict__ acc,
    const double *__restrict__ xold, GhStatus *st) {
  if (halted(st)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (i < n) {
    double g[6];
    ermak_gauss(k, orig[i], gauss, ncap, g);
    const double f_lan = k.skt * ma.osm[type[i]];
    double4 p = pos[i];
    double x[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double pos_v = f_lan * k.sdr * g[2 * d];
      x[d] = x[d] + k.c1 * vel[d * ncap + i] + k.c2 * k.dt * acc[d * ncap + i] + pos_v;
      const double dd = x[d] - xold[d * ncap + i];
      d2 += dd * dd;
    }
    p.x = x[0];
    p.y = x[1];
    p.z = x[2];
    pos[i] = p;
  }
  disp_reduce(d2, &st->maxdisp2_bits);
}

// ermak_b (Integration.F90:1013-1031); vel_v is re-derived from the same
// normals that ermak_a used (same stream entry), not stored.
__global__ __launch_bounds__(GH_BLOCK) void k_ermak_b(
    int n, int ncap, ErmakArgs k, MassArgs ma, const int *__restrict__ type,
    const int *__restrict__ orig, const double *__restrict__ gauss, double *__restrict__ vel,
    double *__restrict__ acc, const double *__restrict__ frc, const GhStatus *st) {
  if (halted(st)) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double g[6];
  ermak_gauss(k, orig[i], gauss, ncap, g);
  const int t = type[i];
  const double f_lan = k.skt * ma.osm[t];
  const double om = ma.om[t];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double vel_v = f_lan * k.sdv * (k.crv1 * g[2 * d] + k.crv2 * g[2 * d + 1]);
    const double f = frc[d * ncap + i];
    vel[d * ncap + i] = k.c0 * vel[d * ncap + i] + (k.c1 - k.c2) * acc[d * ncap + i] +
                        k.c2 * f * om + vel_v;
    acc[d * ncap + i] = f * om;
  }
}

// ---------------------------------------------------------------------------
// Energies (write_energy, Output.F90:824-831): deterministic two-stage sums,
// compensated: every running sum carries its rounding error beside it
// (TwoSum, Knuth), partial sums travel as (sum, error) pairs and the result
// is sum + error, so the total does not depend on the grid's order of
// accumulation beyond the last bit.  (The reference sums per atom with an
// OpenMP reduction, TB.F90:472-477.)
// stage A: [epot, M, Px, Py, Pz] partials; stage B: sum m |v - vcm|^2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void dd_add(double &s, double &c, double x) {
  const double t = s + x;
  const double bp = t - s;
  c += (s - (t - bp)) + (x - bp);
  s = t;
}

// Block-wide compensated sums of NV values; deterministic order.  Result
// (sum, error) valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum_dd(double (&s)[NV], double (&c)[NV]) {
  __shared__ double part[GH_BLOCK / 64][NV][2];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k)
    for (int o = 32; o > 0; o >>= 1) {
      const double s2 = __shfl_xor(s[k], o, 64), c2 = __shfl_xor(c[k], o, 64);
      dd_add(s[k], c[k], s2);
      c[k] += c2;
    }
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      part[wid][k][0] = s[k];
      part[wid][k][1] = c[k];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double ss = 0.0, cc = 0.0;
      for (int w = 0; w < GH_BLOCK / 64; ++w) {
        dd_add(ss, cc, part[w][k][0]);
        cc += part[w][k][1];
      }
      s[k] = ss;
      c[k] = cc;
    }
  }
}

__global__ __launch_bounds__(GH_BLOCK) void k_red_a(int n, int ncap, MassArgs ma,
                                                    const int *__restrict__ type,
                                                    const double *__restrict__ epa,
                                                    const double *__restrict__ vel,
                                                    double *__restrict__ part) {
  double v[5] = {0, 0, 0, 0, 0}, e[5] = {0, 0, 0, 0, 0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double m = 1.0 / ma.om[type[i]];
    dd_add(v[0], e[0], epa[i]);
    dd_add(v[1], e[1], m);
    dd_add(v[2], e[2], m * vel[i]);
    dd_add(v[3], e[3], m * vel[ncap + i]);
    dd_add(v[4], e[4], m * vel[2 * ncap + i]);
  }
  block_sum_dd<5>(v, e);
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; ++k) {
      part[(blockIdx.x * 5 + k) * 2] = v[k];
      part[(blockIdx.x * 5 + k) * 2 + 1] = e[k];
    }
}

// One workgroup: out[k] = the compensated total of the nblk (sum, error)
// partials of value k (nv <= 5).
__global__ __launch_bounds__(GH_BLOCK) void k_red_final(int nblk, int nv,
                                                        const double *__restrict__ part,
                                                        double *__restrict__ out) {
  double v[5] = {0, 0, 0, 0, 0}, e[5] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblk; b += blockDim.x)
    for (int k = 0; k < nv; ++k) {
      dd_add(v[k], e[k], part[(b * nv + k) * 2]);
      e[k] += part[(b * nv + k) * 2 + 1];
    }
  block_sum_dd<5>(v, e);
  if (threadIdx.x == 0)
    for (int k = 0; k < nv; ++k) out[k] = v[k] + e[k];
}

__global__ __launch_bounds__(GH_BLOCK) void k_red_b(int n, int ncap, MassArgs ma,
                                                    const int *__restrict__ type,
                                                    const double *__restrict__ vel,
                                                    const double *__restrict__ sums,
                                                    double *__restrict__ part) {
  const double M = sums[1];
  const double cx = sums[2] / M, cy = sums[3] / M, cz = sums[4] / M;
  double v[1] = {0}, e[1] = {0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double m = 1.0 / ma.om[type[i]];
    const double wx = vel[i] - cx, wy = vel[ncap + i] - cy, wz = vel[2 * ncap + i] - cz;
    dd_add(v[0], e[0], (wx * wx + wy * wy + wz * wz) * m);
  }
  block_sum_dd<1>(v, e);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = v[0];
    part[2 * blockIdx.x + 1] = e[0];
  }
}

// Wraps positions into [0, L) on periodic axes (used when the box or the
// atoms change; the timestep kernels apply do_pbc_1's single shift).
__global__ void k_wrap(int n, WrapArgs wa, double4 *__restrict__ pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 p = pos[i];
  double x[3] = {p.x, p.y, p.z};
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (wa.pbc[d]) {
      x[d] = x[d] - wa.L[d] * floor(x[d] / wa.L[d]);
      if (x[d] >= wa.L[d]) x[d] -= wa.L[d];
    }
  p.x = x[0];
  p.y = x[1];
  p.z = x[2];
  pos[i] = p;
}

// New positions from the host, in caller order (x: 3n, AoS): each atom's
// coordinate moves to the periodic image nearest its current one (a host-side
// wrap is no displacement), and its displacement since the last build joins
// the rebuild test.
__global__ __launch_bounds__(GH_BLOCK) void k_set_positions(
    int n, int ncap, WrapArgs wa, const int *__restrict__ orig, const double *__restrict__ x,
    double4 *__restrict__ pos, const double *__restrict__ xold,
    unsigned long long *__restrict__ maxd2) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0;
  if (s < n) {
    const int o = orig[s];
    double4 p = pos[s];
    double v[3] = {x[3 * (size_t)o], x[3 * (size_t)o + 1], x[3 * (size_t)o + 2]};
    const double old[3] = {p.x, p.y, p.z};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (wa.pbc[d]) v[d] += wa.L[d] * rint((old[d] - v[d]) / wa.L[d]);
      const double dd = v[d] - xold[d * ncap + s];
      d2 += dd * dd;
    }
    p.x = v[0];
    p.y = v[1];
    p.z = v[2];
    pos[s] = p;
  }
  disp_reduce(d2, maxd2);
}

__global__ void k_wrap_rebuild(const GhStatus *st, int n, WrapArgs wa,
                               double4 *__restrict__ pos) {
  if (!st->rebuild) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 p = pos[i];
  double x[3] = {p.x, p.y, p.z};
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (wa.pbc[d]) {
      x[d] = x[d] - wa.L[d] * floor(x[d] / wa.L[d]);
      if (x[d] >= wa.L[d]) x[d] -= wa.L[d];
    }
  p.x = x[0];
  p.y = x[1];
  p.z = x[2];
  pos[i] = p;
}

inline int nblocks(int n) { return (n + GH_BLOCK - 1) / GH_BLOCK; }

GridArgs make_grid(const gh_ctx *c) {
  GridArgs g;
  for (int d = 0; d < 3; ++d) {
    g.nc[d] = c->nc[d];
    g.gn[d] = c->gn[d];
    g.gofs[d] = c->gofs[d];
    g.pbc[d] = c->pbc[d];
    g.L[d] = c->L[d];
    g.one_cell[d] = 1.0 / c->cell[d];
  }
  g.ncell = c->ncell;
  g.ghost = c->ghost;
  return g;
}

MassArgs make_mass(const gh_ctx *c) {
  MassArgs m;
  for (int t = 0; t < GH_MAXT; ++t) {
    m.om[t] = 1.0 / c->mass_t[t];
    m.osm[t] = sqrt(m.om[t]);
  }
  return m;
}

WrapArgs make_wrap(const gh_ctx *c) {
  WrapArgs w;
  for (int d = 0; d < 3; ++d) {
    w.pbc[d] = c->boxed ? c->pbc[d] : 0;
    w.L[d] = c->L[d];
  }
  return w;
}

ErmakArgs make_ermak(const gh_ctx *c) {
  ErmakArgs k;
  const double dt = c->dt, gama = c->gamma;
  k.c0 = exp(-dt * gama);
  k.c1 = (1.0 - k.c0) / gama;
  k.c2 = (1.0 - k.c1 / dt) / gama;
  k.sdr = sqrt(dt / gama * (2.0 - (3.0 - 4.0 * k.c0 + k.c0 * k.c0) / (dt * gama)));
  k.sdv = sqrt(1.0 - k.c0 * k.c0);
  k.crv1 = (1.0 - k.c0) * (1.0 - k.c0) / (gama * k.sdr * k.sdv);
  k.crv2 = sqrt(1.0 - (k.crv1 * k.crv1));
  k.skt = sqrt(ghc::kB_ui() * c->T);
  k.dt = dt;
  k.philox = c->rng_mode == GH_RNG_PHILOX;
  k.seed = c->philox_seed;
  k.step = c->step;
  return k;
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// Neighbour rebuild.  force: request a rebuild regardless of displacement.
// Every kernel of the rebuild reads the device flag set by k_decide, so the
// sequence can be enqueued every step without a host round trip.
using namespace ghd;

int gh_request_rebuild(gh_ctx *c) {
  hipStream_t s = c->stream;
  GhStatus h;
  GH_HIP(c, hipMemcpyAsync(&h, c->status, sizeof(GhStatus), hipMemcpyDeviceToHost, s));
  GH_HIP(c, hipStreamSynchronize(s));
  h.force_rebuild = 1;
  GH_HIP(c, hipMemcpyAsync(c->status, &h, sizeof(GhStatus), hipMemcpyHostToDevice, s));
  GH_HIP(c, hipStreamSynchronize(s));
  return 0;
}

int gh_launch_rebuild(gh_ctx *c, bool force) {
  hipStream_t s = c->stream;
  if (force) GH_TRY(gh_request_rebuild(c));
  const int n = c->n, nb = nblocks(n);
  if (c->list_mode == 1) return gh_launch_rebuild_cells(c);  // decides in its first kernels
  { k_decide<<<1, 1, 0, s>>>(c->status, c->skin); ++c->launches; }
  if (c->list_mode == 0 && c->boxed && (c->pbc[0] || c->pbc[1] || c->pbc[2]))
    { k_wrap_rebuild<<<nb, GH_BLOCK, 0, s>>>(c->status, n, make_wrap(c), c->pos); ++c->launches; }
  const GridArgs g = make_grid(c);
  const double rl = c->rcut + c->skin;
  const double rl2 = rl * rl;
  const int ld = c->ncap;
  { k_build_n2<<<nb, GH_BLOCK, 0, s>>>(c->status, c->pos, n, g, rl2, c->cap, ld, 0, c->nn,
                                     (unsigned *)c->list, c->xold, c->ncap, c->orig,
                                     c->masked ? c->mfeel : nullptr,
                                     c->masked ? c->mfelt : nullptr, c->step); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  return 0;
}

bool gh_can_fuse_vv(const gh_ctx *c) {
  return c->list_mode == 1 && c->integ == GH_INT_VV &&
         (c->pot == GH_POT_SMATB || c->pot == GH_POT_LJ || gh_is_table_pot(c->pot));
}

int gh_launch_forces(gh_ctx *c, hipEvent_t mid, bool fuse_vv) {
  if (c->list_mode == 1) return gh_launch_forces_tiled(c, mid, fuse_vv);
  if (fuse_vv) return gh_fail(c, GH_ESTATE, "fused integration needs the tiled path");
  hipStream_t s = c->stream;
  const int n = c->n, nb = nblocks(n);
  const int ld = c->ncap;
  const double ev_ui = ghc::ev_ui();
  const double L0 = c->L[0], L1 = c->L[1], L2 = c->L[2];
  const unsigned *list = (const unsigned *)c->list;
  const bool single = c->nt == 1;
  const bool mixed = c->precision == GH_MIXED;
  if (c->pot == GH_POT_SMATB) {
    if (mixed) {
      if (single)
        { k_tb_rho_mx<true><<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->type, c->nn, list, ld, n, c->tbd,
                                                  c->nt, L0, L1, L2, c->rho, c->eband,
                                                  c->eband32, c->tb_e0); ++c->launches; }
      else
        { k_tb_rho_mx<false><<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->type, c->nn, list, ld, n, c->tbd,
                                                   c->nt, L0, L1, L2, c->rho, c->eband,
                                                   c->eband32, c->tb_e0); ++c->launches; }
    } else {
      if (single)
        { k_tb_rho64<true><<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->type, c->nn, list, ld, n, c->tbd,
                                                 c->nt, L0, L1, L2, c->rho, c->eband,
                                                 c->eband32); ++c->launches; }
      else
        { k_tb_rho64<false><<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->type, c->nn, list, ld, n, c->tbd,
                                                  c->nt, L0, L1, L2, c->rho, c->eband,
                                                  c->eband32); ++c->launches; }
    }
    if (mid) GH_HIP(c, hipEventRecord(mid, s));
    if (mixed) {
      if (single)
        { k_tb_force_mx<true><<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->type, c->nn, list, ld, n,
                                                    c->ncap, c->tbd, c->nt, L0, L1, L2, c->rho,
                                                    c->eband32, ev_ui, c->frc, c->epa); ++c->launches; }
      else
        { k_tb_force_mx<false><<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->type, c->nn, list, ld, n,
                                                     c->ncap, c->tbd, c->nt, L0, L1, L2,
                                                     c->rho, c->eband32, ev_ui, c->frc,
                                                     c->epa); ++c->launches; }
    } else {
      if (single)
        { k_tb_force64<true><<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->type, c->nn, list, ld, n,
                                                   c->ncap, c->tbd, c->nt, L0, L1, L2, c->rho,
                                                   c->eband, ev_ui, c->frc, c->epa); ++c->launches; }
      else
        { k_tb_force64<false><<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->type, c->nn, list, ld, n,
                                                    c->ncap, c->tbd, c->nt, L0, L1, L2, c->rho,
                                                    c->eband, ev_ui, c->frc, c->epa); ++c->launches; }
    }
  } else if (c->pot == GH_POT_LJ) {
    if (mid) GH_HIP(c, hipEventRecord(mid, s));
    const double rc2 = c->lj_rc * c->lj_rc;
    if (mixed)
      { k_lj_mx<<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->nn, list, ld, n, c->ncap, (float)c->lj_eps,
                                      (float)(c->lj_sigma * c->lj_sigma), rc2, L0, L1, L2,
                                      ev_ui, c->frc, c->epa); ++c->launches; }
    else
      { k_lj64<<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->nn, list, ld, n, c->ncap, c->lj_eps,
                                     c->lj_sigma, rc2, L0, L1, L2, ev_ui, c->frc, c->epa); ++c->launches; }
  } else if (gh_is_table_pot(c->pot)) {
    if (mid) GH_HIP(c, hipEventRecord(mid, s));
    const double rc2 = c->lj_rc * c->lj_rc;
#define PLJ_LAUNCH(MX, PT)                                                                  \
  do {                                                                                      \
    k_plj<MX, PT><<<xgrid(nb), GH_BLOCK, 0, s>>>(c->pos, c->type, c->nn, list, ld, n,        \
                                                 c->ncap, c->pljd, rc2, L0, L1, L2, ev_ui,  \
                                                 c->frc, c->epa);                           \
    ++c->launches;                                                                          \
  } while (0)
#define PLJ_POTS(MX)                                                                        \
  switch (c->pot) {                                                                         \
    case GH_POT_SLJ: PLJ_LAUNCH(MX, GH_POT_SLJ); break;                                     \
    case GH_POT_WCA: PLJ_LAUNCH(MX, GH_POT_WCA); break;                                     \
    case GH_POT_SM1: PLJ_LAUNCH(MX, GH_POT_SM1); break;                                     \
    default: PLJ_LAUNCH(MX, GH_POT_PLJ);                                                    \
  }
    if (mixed) {
      PLJ_POTS(true)
    } else {
      PLJ_POTS(false)
    }
#undef PLJ_POTS
#undef PLJ_LAUNCH
  } else {
    return gh_fail(c, GH_ESTATE, "no potential set");
  }
  GH_HIP(c, hipGetLastError());
  return 0;
}

int gh_launch_vv_a(gh_ctx *c) {
  { k_vv_a<<<nblocks(c->n), GH_BLOCK, 0, c->stream>>>(c->n, c->ncap, c->dt,
                                                    make_mass(c), c->type, c->pos, c->vel,
                                                    c->acc, c->frc, c->xold, c->status); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  return 0;
}

int gh_launch_vv_ba(gh_ctx *c) {
  { k_vv_ba<<<nblocks(c->n), GH_BLOCK, 0, c->stream>>>(c->n, c->ncap, c->dt,
                                                     make_mass(c), c->type, c->pos, c->vel,
                                                     c->frc, c->xold, c->status); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  return 0;
}

int gh_launch_vv_b(gh_ctx *c) {
  { k_vv_b<<<nblocks(c->n), GH_BLOCK, 0, c->stream>>>(c->n, c->ncap, c->dt, make_mass(c), c->type,
                                                    c->vel, c->acc, c->frc, c->status); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  return 0;
}

int gh_launch_ermak_a(gh_ctx *c) {
  { k_ermak_a<<<nblocks(c->n), GH_BLOCK, 0, c->stream>>>(
      c->n, c->ncap, make_ermak(c), make_mass(c), c->type, c->orig,
      c->gauss, c->pos, c->vel, c->acc, c->xold, c->status); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  return 0;
}

int gh_launch_ermak_b(gh_ctx *c) {
  { k_ermak_b<<<nblocks(c->n), GH_BLOCK, 0, c->stream>>>(c->n, c->ncap, make_ermak(c),
                                                       make_mass(c), c->type, c->orig,
                                                       c->gauss, c->vel, c->acc, c->frc,
                                                       c->status); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  return 0;
}

// epot and ekin (ui) of the current state; host-synchronous.
int gh_launch_energies(gh_ctx *c, double *epot_ui, double *ekin_ui) {
  hipStream_t s = c->stream;
  const int nblk = std::min(c->nred, std::max(1, nblocks(c->n)));
  const MassArgs ma = make_mass(c);
  double *part = c->red;                 // [nblk*5] (sum, error) pairs
  double *sums = c->red + 10 * c->nred;  // [5]
  double *kin = sums + 8;                // [1]
  { k_red_a<<<nblk, GH_BLOCK, 0, s>>>(c->n, c->ncap, ma, c->type, c->epa, c->vel, part); ++c->launches; }
  { k_red_final<<<1, GH_BLOCK, 0, s>>>(nblk, 5, part, sums); ++c->launches; }
  { k_red_b<<<nblk, GH_BLOCK, 0, s>>>(c->n, c->ncap, ma, c->type, c->vel, sums, part); ++c->launches; }
  { k_red_final<<<1, GH_BLOCK, 0, s>>>(nblk, 1, part, kin); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  GH_HIP(c, hipMemcpyAsync(c->red_h, sums, sizeof(double) * 9, hipMemcpyDeviceToHost, s));
  GH_HIP(c, hipStreamSynchronize(s));
  *epot_ui = c->red_h[0];
  *ekin_ui = c->red_h[8] * 0.5;
  return 0;
}

int gh_launch_set_positions(gh_ctx *c, const double *x_dev) {
  { k_set_positions<<<nblocks(c->n), GH_BLOCK, 0, c->stream>>>(
      c->n, c->ncap, make_wrap(c), c->orig, x_dev, c->pos, c->xold, gh_maxd2(c)); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  return 0;
}

int gh_launch_wrap(gh_ctx *c) {
  { k_wrap<<<nblocks(c->n), GH_BLOCK, 0, c->stream>>>(c->n, make_wrap(c), c->pos); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  return 0;
}

// Fill helper for host code (no hipMemset in code that may be captured).
int gh_fill_int(gh_ctx *c, int *p, int v, int n) {
  if (n <= 0) return 0;
  { k_fill_int<<<nblocks(n), GH_BLOCK, 0, c->stream>>>(p, v, n); ++c->launches; }
  GH_HIP(c, hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// gh_cluster_forces_static: gh_compute of one free cluster (no box: the
// O(N^2) verlet_self list and the SMA-TB passes above) from the caller's
// device arrays, every intermediate in one file-scope device array.  No
// device memory is allocated; the array lives as long as the process, one
// copy per device.  Calls that overlap on one device share it and race.
// ---------------------------------------------------------------------------
#define GH_STATIC_MAXN 2048  // atoms per call
#define GH_STATIC_CAP 160    // list row capacity (a free cluster's rows hold < 90)
#define GH_STATIC_MAXT 2     // atom types per call
namespace {
constexpr size_t al256(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kSOffPos = al256(sizeof(GhStatus));
constexpr size_t kSOffList = kSOffPos + al256(sizeof(double4) * GH_STATIC_MAXN);
constexpr size_t kSOffNn = kSOffList + al256(sizeof(unsigned) * GH_STATIC_CAP * GH_STATIC_MAXN);
constexpr size_t kSOffXold = kSOffNn + al256(sizeof(int) * GH_STATIC_MAXN);
constexpr size_t kSOffRho = kSOffXold + al256(sizeof(double) * 3 * GH_STATIC_MAXN);
constexpr size_t kSOffEband = kSOffRho + al256(sizeof(double) * GH_STATIC_MAXN);
constexpr size_t kSOffEb32 = kSOffEband + al256(sizeof(double) * GH_STATIC_MAXN);
constexpr size_t kSOffFrc = kSOffEb32 + al256(sizeof(float) * GH_STATIC_MAXN);
constexpr size_t kSOffTb = kSOffFrc + al256(sizeof(double) * 3 * GH_STATIC_MAXN);
constexpr size_t kScratchBytes = kSOffTb + al256(sizeof(TbDev) * GH_MAXT * GH_MAXT);

// the parameter blocks of up to GH_STATIC_MAXT types, by value
struct StaticTb {
  TbDev t[GH_STATIC_MAXT * GH_STATIC_MAXT];
};

// Re-arms the status for one build, stages positions (3n AoS -> double4)
// and writes the parameter table.  Every field the later kernels read is
// written here: nothing depends on an earlier call's contents.
__global__ __launch_bounds__(GH_BLOCK) void k_static_init(unsigned char *scr, int n, int nt,
                                                          const double *__restrict__ x,
                                                          StaticTb tb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  GhStatus *st = reinterpret_cast<GhStatus *>(scr);
  if (i == 0) {
    st->maxdisp2_bits = 0ull;
    st->rebuild = 1;
    st->overflow = 0;
    st->maxnn = 0;
    st->outside = 0;
    st->halt = 0;
    st->halt_step = 0;
  }
  TbDev *tbd = reinterpret_cast<TbDev *>(scr + kSOffTb);
  if (i < nt * nt) tbd[(i / nt) * GH_MAXT + i % nt] = tb.t[i];
  if (i < n) {
    double4 *pos = reinterpret_cast<double4 *>(scr + kSOffPos);
    pos[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], 0.0);
  }
}

// Forces to the caller's AoS layout; one workgroup then sums the energies
// (compensated) and reports the build's row overflow.
__global__ __launch_bounds__(GH_BLOCK) void k_static_out(const unsigned char *scr, int n,
                                                         double *__restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *frc = reinterpret_cast<const double *>(scr + kSOffFrc);
#pragma unroll
  for (int d = 0; d < 3; ++d) F[3 * i + d] = frc[d * GH_STATIC_MAXN + i];
}
__global__ __launch_bounds__(GH_BLOCK) void k_static_sum(const unsigned char *scr, int n,
                                                         const double *__restrict__ e,
                                                         double *__restrict__ epot,
                                                         int *__restrict__ status) {
  double v[1] = {0.0}, c[1] = {0.0};
  for (int i = threadIdx.x; i < n; i += blockDim.x) dd_add(v[0], c[0], e[i]);
  block_sum_dd<1>(v, c);
  if (threadIdx.x == 0) {
    *epot = v[0] + c[0];
    const GhStatus *st = reinterpret_cast<const GhStatus *>(scr);
    *status = st->overflow ? GH_EOVERFLOW : 0;
  }
}
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device, taken once per device with
// that device current (the host-side name is only a stand-in).
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

extern "C" int gh_cluster_forces_static(int device, int precision, int n, int nt,
                                        const double *x, const int *type, const double *A,
                                        const double *xi, const double *p, const double *q,
                                        const double *r0, const double *rci, const double *rce,
                                        double skin, double *F, double *e_atom, double *epot,
                                        int *status, void *stream) {
  if (device < 0 || device >= 64 || (precision != GH_FP64 && precision != GH_MIXED))
    return GH_EINVAL;
  if (n < 1 || n > GH_STATIC_MAXN || nt < 1 || nt > GH_STATIC_MAXT || !(skin >= 0.0))
    return GH_EINVAL;  // n <= GH_STATIC_MAXN atoms, nt <= GH_STATIC_MAXT types
  if (!x || !type || !A || !xi || !p || !q || !r0 || !rci || !rce || !F || !e_atom || !epot ||
      !status)
    return GH_EINVAL;
  int prev = 0;
  if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) return GH_EHIP;
  int rc = 0;
  unsigned char *scr = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
      void *ptr = nullptr;
      if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) rc = GH_EHIP;
      else g_scratch_ptr[device] = ptr;
    }
    scr = static_cast<unsigned char *>(g_scratch_ptr[device]);
  }
  // parameter blocks (TB.F90:106-170) as gh_set_smatb and gh_upload_tb build them
  StaticTb tb;
  memset(&tb, 0, sizeof(tb));
  double rcut = 0.0;
  const GhTbPair t00{A[0], xi[0], p[0], q[0], r0[0], rci[0], rce[0], 0, 0, 0, 0, 0, 0};
  const double e0 = gh_tb_e0(&t00);
  for (int a = 0; a < nt && rc == 0; ++a)
    for (int b = 0; b < nt; ++b) {
      const int k = a * nt + b;
      if (!(r0[k] > 0) || !(rci[k] > 0) || !(rce[k] > rci[k])) rc = GH_EINVAL;
      GhTbPair t{A[k], xi[k], p[k], q[k], r0[k], rci[k], rce[k], 0, 0, 0, 0, 0, 0};
      gh_tb_coeffs(t);
      TbDev &d = tb.t[k];
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
      gh_tbf_fill(t, 0, e0, d.f);
      rcut = std::max(rcut, t.rce);
    }
  if (rc == 0) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    GhStatus *st = reinterpret_cast<GhStatus *>(scr);
    double4 *pos = reinterpret_cast<double4 *>(scr + kSOffPos);
    unsigned *list = reinterpret_cast<unsigned *>(scr + kSOffList);
    int *nn = reinterpret_cast<int *>(scr + kSOffNn);
    double *xold = reinterpret_cast<double *>(scr + kSOffXold);
    double *rho = reinterpret_cast<double *>(scr + kSOffRho);
    double *eband = reinterpret_cast<double *>(scr + kSOffEband);
    float *eb32 = reinterpret_cast<float *>(scr + kSOffEb32);
    double *frc = reinterpret_cast<double *>(scr + kSOffFrc);
    const TbDev *tbd = reinterpret_cast<const TbDev *>(scr + kSOffTb);
    const int nb = nblocks(n), ld = GH_STATIC_MAXN;
    const double L = 1e6, rl = rcut + skin;
    GridArgs g;
    memset(&g, 0, sizeof(g));
    for (int d = 0; d < 3; ++d) {
      g.nc[d] = g.gn[d] = 1;
      g.L[d] = L;
      g.one_cell[d] = 1.0 / L;
    }
    g.ncell = 1;
    const double ev_ui = ghc::ev_ui();
    const bool single = nt == 1;
    k_static_init<<<std::max(nb, 1), GH_BLOCK, 0, s>>>(scr, n, nt, x, tb);
    k_build_n2<<<nb, GH_BLOCK, 0, s>>>(st, pos, n, g, rl * rl, GH_STATIC_CAP, ld, 0, nn, list,
                                        xold, ld, nullptr, nullptr, nullptr, 0);
    if (precision == GH_MIXED) {
      if (single) {
        k_tb_rho_mx<true><<<xgrid(nb), GH_BLOCK, 0, s>>>(pos, type, nn, list, ld, n, tbd, nt, L, L,
                                                         L, rho, eband, eb32, e0);
        k_tb_force_mx<true><<<xgrid(nb), GH_BLOCK, 0, s>>>(pos, type, nn, list, ld, n, ld, tbd,
                                                           nt, L, L, L, rho, eb32, ev_ui, frc,
                                                           e_atom);
      } else {
        k_tb_rho_mx<false><<<xgrid(nb), GH_BLOCK, 0, s>>>(pos, type, nn, list, ld, n, tbd, nt, L,
                                                          L, L, rho, eband, eb32, e0);
        k_tb_force_mx<false><<<xgrid(nb), GH_BLOCK, 0, s>>>(pos, type, nn, list, ld, n, ld, tbd,
                                                            nt, L, L, L, rho, eb32, ev_ui, frc,
                                                            e_atom);
      }
    } else {
      if (single) {
        k_tb_rho64<true><<<xgrid(nb), GH_BLOCK, 0, s>>>(pos, type, nn, list, ld, n, tbd, nt, L, L,
                                                        L, rho, eband, eb32);
        k_tb_force64<true><<<xgrid(nb), GH_BLOCK, 0, s>>>(pos, type, nn, list, ld, n, ld, tbd, nt,
                                                          L, L, L, rho, eband, ev_ui, frc,
                                                          e_atom);
      } else {
        k_tb_rho64<false><<<xgrid(nb), GH_BLOCK, 0, s>>>(pos, type, nn, list, ld, n, tbd, nt, L,
                                                         L, L, rho, eband, eb32);
        k_tb_force64<false><<<xgrid(nb), GH_BLOCK, 0, s>>>(pos, type, nn, list, ld, n, ld, tbd,
                                                           nt, L, L, L, rho, eband, ev_ui, frc,
                                                           e_atom);
      }
    }
    k_static_out<<<nb, GH_BLOCK, 0, s>>>(scr, n, F);
    k_static_sum<<<1, GH_BLOCK, 0, s>>>(scr, n, e_atom, epot, status);
    if (hipGetLastError() != hipSuccess) rc = GH_EHIP;
  }
  (void)hipSetDevice(prev);
  return rc;
}
