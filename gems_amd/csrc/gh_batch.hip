// Batched free clusters: many independent unboxed systems advanced in one
// launch, one workgroup per system, largest system first.
//
// GEMS has no batching (one system per process, replica exchange over MPI
// ranks, src/ReplicaExchange.f90:27-151); per system the work is the same hot
// path: the O(N^2) verlet_self list (Neighbour.F90:676-759, rd <= rl^2 kept),
// the two SMA-TB passes (TB.F90:367-586) and the Ermak update
// (Integration.F90:963-1031).  A system's atoms are contiguous; all state is
// FP64 in HBM (SoA with stride ld), neighbour lists are ELL rows of 16-bit
// indices local to the system.  Within a workgroup the steps of a system are
// separated by workgroup barriers only, so nsteps need one launch.
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "gh_device.h"

#define GH_BBLOCK 512  // largest workgroup of the batch kernels for the ordinary systems
#define GH_BIGSYS 1024  // systems above this many atoms get 1024-thread workgroups

struct gh_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int precision = GH_FP64;
  std::string err;
  int nsys = 0, ntot = 0, ld = 0;
  int nt = 1;
  double mass_t[GH_MAXT] = {1, 1, 1, 1, 1, 1, 1, 1};
  GhTbPair tb[GH_MAXT * GH_MAXT];
  bool have_pot = false;
  double rcut = 0.0, skin = 1.0;
  int cap = 0;
  int cap_req = 0;  // row capacity asked for by the last overflow
  long long step = 0;
  bool lists_valid = false;
  bool forces_valid = false;
  std::vector<int> off_h, order_h;
  std::vector<double> T_h, epot_h;
  // device
  int *off = nullptr, *order = nullptr, *type = nullptr, *nn = nullptr, *flags = nullptr;
  double4 *pos = nullptr;
  double *vel = nullptr, *acc = nullptr, *frc = nullptr, *e = nullptr, *rho = nullptr,
         *eband = nullptr, *xold = nullptr, *T = nullptr, *out = nullptr, *scale = nullptr;
  float *eband32 = nullptr;
  float *gn = nullptr;  // [6][ld] this step's normals (stepa -> stepb)
  unsigned short *list = nullptr;
  TbDev *tbd = nullptr;
  TbF *tbf = nullptr;
  double tb_e0 = 0.0;
  int nst = 0;   // LDS staging slots per workgroup (largest system + sentinel)
  int nmax = 0;  // atoms of the largest system
  // systems above GH_BIGSYS atoms (the first nbig of the largest-first order)
  // run on a second stream beside the rest; nmax_rest = the largest of the rest
  int nbig = 0, nmax_rest = 0;
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

namespace {
using namespace ghd;

struct BatchArgs {
  const int *off, *order, *type;
  int first;  // this launch's workgroup b works on system order[first + b]
  double4 *pos;
  double *vel, *acc, *frc, *e, *rho, *eband, *xold;
  float *eband32;
  float *gn;
  int *nn;
  unsigned short *list;
  int cap, ld;
  const TbDev *tbd;
  const TbF *tbf;
  int nt;
  double om[GH_MAXT], osm[GH_MAXT];
  double rl2;
  float near2f, lo2f, hi2f;  // build: inner-branch bound, FP32 prefilter band
  int nst;                   // LDS staging slots (largest system + sentinel); 0 = none
  double ev_ui;
  const double *T;
  double *out;  // per system: [0] epot ui, [1] ekin ui, [2] rebuilds, [3] fmax eV/A,
                // [4] iterations, [5] converged
  int *flags;   // [0] list overflow, [1] longest row asked for past the capacity
  double e0;    // eband reference of the FP32 pair forms
};

#define OUTW 6

template <int BS>
struct BlockRed {
  double v[BS / 64][8];  // up to 8 values per reduction
};

// Block-wide sums of NV values; the result is returned to every thread.
template <int BS, int NV>
__device__ __forceinline__ void bsum(BlockRed<BS> &R, double (&v)[NV]) {
  static_assert(NV <= 8, "BlockRed holds 8 values per wave");
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) R.v[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
#pragma unroll 4  // fully unrolled, 16 wavefronts x 8 values would all sit in registers
    for (int w = 0; w < BS / 64; ++w) s += R.v[w][k];
    v[k] = s;
  }
  __syncthreads();
}

template <int BS>
__device__ __forceinline__ double bmax(BlockRed<BS> &R, double v) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) R.v[threadIdx.x >> 6][0] = v;
  __syncthreads();
  double m = R.v[0][0];
#pragma unroll 4
  for (int w = 1; w < BS / 64; ++w) m = fmax(m, R.v[w][0]);
  __syncthreads();
  return m;
}

// Per-workgroup state of one system: its atom range, the LDS staging of its
// positions (fixed point around its first atom) and the parameter tables.
struct Sys {
  int a0, n;
  bool staged;          // positions are in q (all within 64 A of the origin, 2^24 steps per A)
  int4 *q;              // [n + 1]: positions, payload; q[n] = sentinel
  unsigned char *ty;    // [n + 1] types
  const TbDev *sp;      // FP64 parameters
  const TbF *sf;        // FP32 parameters
};

// Stages positions (and types) of the system; `staged` is uniform.
__device__ void sys_stage(const BatchArgs &B, Sys &S, int *flag, bool multi) {
  if (B.nst == 0) {
    S.staged = false;
    return;
  }
  const double4 o = B.pos[S.a0];
  if (threadIdx.x == 0) *flag = 0;
  __syncthreads();
  for (int k = threadIdx.x; k < S.n; k += blockDim.x) {
    const double4 p = B.pos[S.a0 + k];
    const double dx = (p.x - o.x) * GH_BFXS, dy = (p.y - o.y) * GH_BFXS,
                 dz = (p.z - o.z) * GH_BFXS;
    if (fabs(dx) >= GH_FXLIM || fabs(dy) >= GH_FXLIM || fabs(dz) >= GH_FXLIM) *flag = 1;
    const int t = B.type[S.a0 + k];
    S.q[k] = make_int4(__double2int_rn(dx), __double2int_rn(dy), __double2int_rn(dz), t);
    if (multi) S.ty[k] = (unsigned char)t;
  }
  if (threadIdx.x == 0) {
    S.q[S.n] = make_int4(GH_FXFAR, GH_FXFAR, GH_FXFAR, 0);
    if (multi) S.ty[S.n] = 0;
  }
  __syncthreads();
  S.staged = *flag == 0;
  __syncthreads();
}

// verlet_self list of one system (Neighbour.F90:676-759: all pairs with
// rd <= rl^2, i excluded; rd rounded as the reference's dot_product).  Rows
// hold system-local indices, the pairs with r^2 < near2 first; the sentinel
// index is n.  Staged: an FP32 prefilter on the exact fixed-point differences
// decides all pairs but those within the band [lo2f, hi2f] around rl^2, which
// the FP64 rd decides.  Ends with a barrier.
__device__ void sys_build(const BatchArgs &B, const Sys &S) {
  const int a0 = S.a0, n = S.n;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int i = a0 + k;
    const double4 pi = B.pos[i];
    RowWriter row{reinterpret_cast<unsigned long long *>(B.list), B.ld, i, B.cap, 0, 0ull};
    if (S.staged) {
      const int4 qi = S.q[k];
      for (int sweep = 0; sweep < 2; ++sweep)
        for (int j = 0; j < n; ++j) {
          if (j == k) continue;
          const int4 qj = S.q[j];
          const float ux = (float)(qj.x - qi.x), uy = (float)(qj.y - qi.y),
                      uz = (float)(qj.z - qi.z);
          const float r2f = fmaf(ux, ux, fmaf(uy, uy, uz * uz)) * (float)(1.0 / (GH_BFXS * GH_BFXS));
          if (r2f > B.hi2f || ((r2f < B.near2f) != (sweep == 0))) continue;
          if (r2f >= B.lo2f) {
            const double4 pj = B.pos[a0 + j];
            if (dist2_rn(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z) > B.rl2) continue;
          }
          row.put(j);
        }
    } else {
      for (int j = 0; j < n; ++j) {
        if (j == k) continue;
        const double4 pj = B.pos[a0 + j];
        if (dist2_rn(pj.x - pi.x, pj.y - pi.y, pj.z - pi.z) <= B.rl2) row.put(j);
      }
    }
    row.finish(n);
    if (row.m > B.cap) {
      B.flags[0] = 1;
      atomicMax(&B.flags[1], row.m);
    }
    B.nn[i] = min(row.m, B.cap);
    B.xold[i] = pi.x;
    B.xold[B.ld + i] = pi.y;
    B.xold[2 * B.ld + i] = pi.z;
  }
  __syncthreads();
}

// Both SMA-TB passes of one system (TB.F90:367-586); ends with a barrier.
// MIXED and staged: LDS fixed-point pair vectors and the FP32 pair forms;
// otherwise FP64 positions from global memory (L1/L2 resident per system).
template <bool MIXED, bool SINGLE>
__device__ void sys_forces(const BatchArgs &B, const Sys &S) {
  const int a0 = S.a0, n = S.n;
  const bool fx = MIXED && S.staged;
  // pass 1: density
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int i = a0 + k;
    const int ti = SINGLE ? 0 : B.type[i];
    double acc = 0.0;
    if (fx) {
      const int4 qi = S.q[k];
      auto pair = [&](const int4 &q) {
        const double dx = fxd(q.x, qi.x), dy = fxd(q.y, qi.y), dz = fxd(q.z, qi.z);
        float band;
        if (tb_rho_fx(SINGLE ? S.sf[0] : S.sf[ti * GH_MAXT + q.w],
                      fma(dx, dx, fma(dy, dy, dz * dz)), band))
          acc += (double)band;
      };
      for_each_quad(B.list, B.ld, i, B.nn[i], [&](unsigned long long w) {
        const int4 q0 = S.q[GH_QSLOT(w, 0)], q1 = S.q[GH_QSLOT(w, 1)], q2 = S.q[GH_QSLOT(w, 2)],
                   q3 = S.q[GH_QSLOT(w, 3)];
        pair(q0);
        pair(q1);
        pair(q2);
        pair(q3);
      });
    } else {
      const double4 pi = B.pos[i];
      for_each_slot(B.list, B.ld, i, B.nn[i], [&](int s) {
        if (s >= n) return;  // sentinel
        const int j = a0 + s;
        const double4 pj = B.pos[j];
        const TbDev &P = SINGLE ? S.sp[0] : S.sp[ti * GH_MAXT + B.type[j]];
        const double vx = pj.x - pi.x, vy = pj.y - pi.y, vz = pj.z - pi.z;
        if (MIXED) {
          float band;
          if (tb_rho_fx(P.f, fma(vx, vx, fma(vy, vy, vz * vz)), band)) acc += (double)band;
        } else {
          const double dr = sqrt(dist2_rn(vx, vy, vz));
          if (dr <= P.rce) acc += tb_band64(P, dr);
        }
      });
    }
    B.rho[i] = acc;
    const double eb = 1.0 / sqrt(acc);
    B.eband[i] = eb;
    B.eband32[i] = (float)(eb - B.e0);  // the FP32 forms take 1/sqrt(rho) - e0
  }
  __syncthreads();
  if (fx) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) S.q[k].w = __float_as_int(B.eband32[a0 + k]);
    if (threadIdx.x == 0) S.q[n].w = 0;
    __syncthreads();
  }
  // pass 2: forces and per-atom energies
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    const int i = a0 + k;
    const int ti = SINGLE ? 0 : B.type[i];
    double fxs = 0.0, fys = 0.0, fzs = 0.0, ei = 0.0, scale = B.ev_ui;
    if (fx) {
      const int4 qi = S.q[k];
      const float di = __int_as_float(qi.w);
      auto pair = [&](const int4 &q, int s) {
        const double dx = fxd(q.x, qi.x), dy = fxd(q.y, qi.y), dz = fxd(q.z, qi.z);
        float sf, repul;
        if (tb_force_fx<true>(SINGLE ? S.sf[0] : S.sf[ti * GH_MAXT + S.ty[s]],
                              fma(dx, dx, fma(dy, dy, dz * dz)), di + __int_as_float(q.w), sf,
                              repul)) {
          const double sfd = (double)sf;
          fxs = fma(-sfd, dx, fxs);
          fys = fma(-sfd, dy, fys);
          fzs = fma(-sfd, dz, fzs);
          ei += (double)repul;
        }
      };
      for_each_quad(B.list, B.ld, i, B.nn[i], [&](unsigned long long w) {
        const int s0 = GH_QSLOT(w, 0), s1 = GH_QSLOT(w, 1), s2 = GH_QSLOT(w, 2),
                  s3 = GH_QSLOT(w, 3);
        const int4 q0 = S.q[s0], q1 = S.q[s1], q2 = S.q[s2], q3 = S.q[s3];
        pair(q0, s0);
        pair(q1, s1);
        pair(q2, s2);
        pair(q3, s3);
      });
      scale = B.ev_ui;  // tb_force_fx returns factor / r per step: forces come out in ui/A
    } else {
      const double4 pi = B.pos[i];
      const double ebi = B.eband[i];
      const float di = B.eband32[i];
      for_each_slot(B.list, B.ld, i, B.nn[i], [&](int s) {
        if (s >= n) return;  // sentinel
        const int j = a0 + s;
        const double4 pj = B.pos[j];
        const TbDev &P = SINGLE ? S.sp[0] : S.sp[ti * GH_MAXT + B.type[j]];
        const double vx = pj.x - pi.x, vy = pj.y - pi.y, vz = pj.z - pi.z;
        if (MIXED) {
          float sf32, repul;
          if (tb_force_fx<true>(P.f, fma(vx, vx, fma(vy, vy, vz * vz)), di + B.eband32[j],
                                sf32, repul)) {
            const double sf = (double)sf32;
            fxs = fma(-sf, vx, fxs);
            fys = fma(-sf, vy, fys);
            fzs = fma(-sf, vz, fzs);
            ei += (double)repul;
          }
        } else {
          const double dr = sqrt(dist2_rn(vx, vy, vz));
          if (dr <= P.rce) {
            double factor, repul;
            tb_pair64(P, dr, ebi + B.eband[j], factor, repul);
            const double sf = factor / dr;
            fxs -= sf * vx;
            fys -= sf * vy;
            fzs -= sf * vz;
            ei += repul;
          }
        }
      });
    }
    B.frc[i] = fxs * scale;
    B.frc[B.ld + i] = fys * scale;
    B.frc[2 * B.ld + i] = fzs * scale;
    B.e[i] = (ei - sqrt(B.rho[i])) * B.ev_ui;
  }
  __syncthreads();
}

// Shared tables and the system of this workgroup.
#define SYS_SETUP(SINGLE)                                                                  \
  extern __shared__ int4 dyn_q[];                                                          \
  __shared__ TbDev sp[SINGLE ? 1 : GH_MAXT * GH_MAXT];                                     \
  __shared__ TbF sf[SINGLE ? 1 : GH_MAXT * GH_MAXT];                                       \
  __shared__ BlockRed<BS> R;                                                               \
  __shared__ int sflag;                                                                    \
  for (int k_ = threadIdx.x; k_ < (SINGLE ? 1 : GH_MAXT * GH_MAXT); k_ += blockDim.x) {    \
    sp[k_] = B.tbd[k_];                                                                    \
    sf[k_] = B.tbf[k_];                                                                    \
  }                                                                                        \
  const int s = B.order[B.first + blockIdx.x];                                             \
  Sys S;                                                                                   \
  S.a0 = B.off[s];                                                                         \
  S.n = B.off[s + 1] - S.a0;                                                               \
  S.q = dyn_q;                                                                             \
  S.ty = reinterpret_cast<unsigned char *>(dyn_q + B.nst);                                 \
  S.sp = sp;                                                                               \
  S.sf = sf;                                                                               \
  S.staged = false;                                                                        \
  __syncthreads();

// epot (ui) and ekin about the centre of mass (ui) of one system.
template <int BS>
__device__ void sys_energies(const BatchArgs &B, BlockRed<BS> &R, int a0, int a1, double &epot,
                             double &ekin) {
  double s[5] = {0, 0, 0, 0, 0};
  for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double m = 1.0 / B.om[B.type[i]];
    s[0] += B.e[i];
    s[1] += m;
    s[2] += m * B.vel[i];
    s[3] += m * B.vel[B.ld + i];
    s[4] += m * B.vel[2 * B.ld + i];
  }
  bsum<BS, 5>(R, s);
  const double cx = s[2] / s[1], cy = s[3] / s[1], cz = s[4] / s[1];
  double k[1] = {0.0};
  for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double m = 1.0 / B.om[B.type[i]];
    const double wx = B.vel[i] - cx, wy = B.vel[B.ld + i] - cy, wz = B.vel[2 * B.ld + i] - cz;
    k[0] += (wx * wx + wy * wy + wz * wz) * m;
  }
  bsum<BS, 1>(R, k);
  epot = s[0];
  ekin = 0.5 * k[0];
}

template <bool MIXED, bool SINGLE, int BS>
__global__ __launch_bounds__(BS) void k_batch_compute(BatchArgs B) {
  SYS_SETUP(SINGLE)
  sys_stage(B, S, &sflag, !SINGLE);
  sys_build(B, S);
  sys_forces<MIXED, SINGLE>(B, S);
  double ep, ek;
  sys_energies(B, R, S.a0, S.a0 + S.n, ep, ek);
  if (threadIdx.x == 0) {
    B.out[s * OUTW + 0] = ep;
    B.out[s * OUTW + 1] = ek;
  }
}

struct ErmakC {
  double c0, c1, c2, sdr, sdv, crv1, crv2, dt, kB_ui;
};

// nsteps Ermak steps of every system (ermak_a / ermak_b with the Philox
// stream keyed by (seed, step, atom)); lists rebuilt per system when
// 2 max|x - x_old| exceeds the skin.
template <bool MIXED, bool SINGLE, int BS>
__global__ __launch_bounds__(BS) void k_batch_langevin(BatchArgs B, int nsteps,
                                                              long long step0, ErmakC k,
                                                              unsigned long long seed,
                                                              double skin) {
  // The step constants live in LDS and are read where they are used: held in
  // scalar registers for the whole kernel they were spilled, and the ROCm 7.2
  // compiler has mis-allocated such spills here (a later kernel-argument load
  // overwrote the registers before they were saved).
  __shared__ ErmakC kc;
  __shared__ unsigned long long kseed;
  __shared__ double kskin;
  if (threadIdx.x == 0) {
    kc = k;
    kseed = seed;
    kskin = skin;
  }
  SYS_SETUP(SINGLE)
  const int a0 = S.a0, a1 = S.a0 + S.n;
  const double skt = sqrt(kc.kB_ui * B.T[s]);
  int nreb = 0;
  for (int st = 0; st < nsteps; ++st) {
    const long long step = step0 + st;
    double d2 = 0.0;
    for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
      double g[6];
      philox_normals(kseed, step, i, g);
#pragma unroll
      for (int q = 0; q < 6; ++q) B.gn[q * B.ld + i] = (float)g[q];  // exact: FP32 normals
      const double f_lan = skt * B.osm[B.type[i]];
      double4 p = B.pos[i];
      double x[3] = {p.x, p.y, p.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double pos_v = f_lan * kc.sdr * g[2 * d];
        x[d] = x[d] + kc.c1 * B.vel[d * B.ld + i] + kc.c2 * kc.dt * B.acc[d * B.ld + i] + pos_v;
        const double dd = x[d] - B.xold[d * B.ld + i];
        d2 += dd * dd;
      }
      p.x = x[0];
      p.y = x[1];
      p.z = x[2];
      B.pos[i] = p;
    }
    const double dmax2 = bmax(R, d2);
    if (MIXED) sys_stage(B, S, &sflag, !SINGLE);
    if (2.0 * sqrt(dmax2) > kskin) {
      if (!MIXED) sys_stage(B, S, &sflag, !SINGLE);
      sys_build(B, S);
      ++nreb;
    }
    sys_forces<MIXED, SINGLE>(B, S);
    for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
      double g[6];
#pragma unroll
      for (int q = 0; q < 6; ++q) g[q] = (double)B.gn[q * B.ld + i];  // this step's stepa draw
      const int t = B.type[i];
      const double f_lan = skt * B.osm[t];
      const double om = B.om[t];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double vel_v = f_lan * kc.sdv * (kc.crv1 * g[2 * d] + kc.crv2 * g[2 * d + 1]);
        const double f = B.frc[d * B.ld + i];
        B.vel[d * B.ld + i] = kc.c0 * B.vel[d * B.ld + i] + (kc.c1 - kc.c2) * B.acc[d * B.ld + i] +
                              kc.c2 * f * om + vel_v;
        B.acc[d * B.ld + i] = f * om;
      }
    }
    __syncthreads();
  }
  double ep, ek;
  sys_energies(B, R, a0, a1, ep, ek);
  if (threadIdx.x == 0) {
    B.out[s * OUTW + 0] = ep;
    B.out[s * OUTW + 1] = ek;
    B.out[s * OUTW + 2] += nreb;
  }
}

struct FireC {
  double dt_max, alpha0, fswitch, ftol, ui_ev;
  int max_iter, mixed;
};

template <bool SINGLE>
__device__ __forceinline__ void fire_forces(const BatchArgs &B, Sys &S, int *flag, bool mixed) {
  if (mixed) {
    sys_stage(B, S, flag, !SINGLE);
    sys_forces<true, SINGLE>(B, S);
  } else {
    sys_forces<false, SINGLE>(B, S);
  }
}

template <int BS>
__device__ double sys_fmax(const BatchArgs &B, BlockRed<BS> &R, int a0, int a1) {
  double m = 0.0;
  for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
    const double fx = B.frc[i], fy = B.frc[B.ld + i], fz = B.frc[2 * B.ld + i];
    m = fmax(m, fx * fx + fy * fy + fz * fz);
  }
  return sqrt(bmax(R, m));
}

// FIRE (Bitzek et al. 2006) per system: velocity-Verlet MD step, then
// P = F.v; v <- (1-a) v + a |v| F/|F|; P > 0 for more than N_min steps grows
// dt (x1.1, capped at dt_max) and shrinks a (x0.99); P <= 0 halves dt, zeroes
// v and resets a.  Mixed precision switches to FP64 pair terms once
// max|F| < fswitch; convergence is max|F| < ftol in FP64.
template <bool SINGLE, int BS>
__global__ __launch_bounds__(BS) void k_batch_fire(BatchArgs B, FireC fc_arg, double skin_arg) {
  // constants in LDS, read where used (see k_batch_langevin)
  __shared__ FireC fc;
  __shared__ double skin;
  if (threadIdx.x == 0) {
    fc = fc_arg;
    skin = skin_arg;
  }
  SYS_SETUP(SINGLE)
  const int a0 = S.a0, a1 = S.a0 + S.n;
  bool mixed = fc.mixed != 0;
  for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x)
    for (int d = 0; d < 3; ++d) B.vel[d * B.ld + i] = 0.0;
  sys_stage(B, S, &sflag, !SINGLE);
  sys_build(B, S);
  fire_forces<SINGLE>(B, S, &sflag, mixed);
  double dt = fc.dt_max * 0.1, alpha = fc.alpha0;
  int npos = 0, it = 0, conv = 0;
  double fmax = 0.0;
  for (it = 0; it < fc.max_iter; ++it) {
    fmax = sys_fmax(B, R, a0, a1) * fc.ui_ev;  // eV/A
    if (mixed && fmax < fc.fswitch) {
      mixed = false;
      fire_forces<SINGLE>(B, S, &sflag, mixed);
      fmax = sys_fmax(B, R, a0, a1) * fc.ui_ev;
    }
    if (!mixed && fmax < fc.ftol) {
      conv = 1;
      break;
    }
    // velocity Verlet step
    double d2 = 0.0;
    for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
      const double om = B.om[B.type[i]];
      double4 p = B.pos[i];
      double x[3] = {p.x, p.y, p.z};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double v = B.vel[d * B.ld + i] + 0.5 * dt * B.frc[d * B.ld + i] * om;
        B.vel[d * B.ld + i] = v;
        x[d] += dt * v;
        const double dd = x[d] - B.xold[d * B.ld + i];
        d2 += dd * dd;
      }
      p.x = x[0];
      p.y = x[1];
      p.z = x[2];
      B.pos[i] = p;
    }
    const double dmax2 = bmax(R, d2);
    if (2.0 * sqrt(dmax2) > skin) {
      sys_stage(B, S, &sflag, !SINGLE);
      sys_build(B, S);
    }
    fire_forces<SINGLE>(B, S, &sflag, mixed);
    double pv[3] = {0.0, 0.0, 0.0};  // F.v, |v|^2, |F|^2
    for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x) {
      const double om = B.om[B.type[i]];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double f = B.frc[d * B.ld + i];
        const double v = B.vel[d * B.ld + i] + 0.5 * dt * f * om;
        B.vel[d * B.ld + i] = v;
        pv[0] += f * v;
        pv[1] += v * v;
        pv[2] += f * f;
      }
    }
    bsum<BS, 3>(R, pv);
    const double mixv = pv[2] > 0.0 ? alpha * sqrt(pv[1]) / sqrt(pv[2]) : 0.0;
    const bool up = pv[0] > 0.0;
    for (int i = a0 + threadIdx.x; i < a1; i += blockDim.x)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const double v = B.vel[d * B.ld + i];
        B.vel[d * B.ld + i] = up ? (1.0 - alpha) * v + mixv * B.frc[d * B.ld + i] : 0.0;
      }
    if (up) {
      ++npos;
      if (npos > 5) {
        dt = fmin(dt * 1.1, fc.dt_max);
        alpha *= 0.99;
      }
    } else {
      npos = 0;
      dt *= 0.5;
      alpha = fc.alpha0;
    }
    __syncthreads();
  }
  double ep, ek;
  sys_energies(B, R, a0, a1, ep, ek);
  if (threadIdx.x == 0) {
    B.out[s * OUTW + 0] = ep;
    B.out[s * OUTW + 3] = fmax;
    B.out[s * OUTW + 4] = it;
    B.out[s * OUTW + 5] = conv;
  }
}

__global__ void k_batch_scale(int nsys, const int *__restrict__ off,
                              const double *__restrict__ scale, double *__restrict__ vel,
                              int ld) {
  const int s = blockIdx.x;
  if (s >= nsys) return;
  const double f = scale[s];
  for (int i = off[s] + threadIdx.x; i < off[s + 1]; i += blockDim.x)
    for (int d = 0; d < 3; ++d) vel[d * ld + i] *= f;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int bfail(gh_batch *b, int code, const std::string &m) {
  if (b) b->err = m;
  return code;
}
#define BH(b, call)                                                                      \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return bfail(b, GH_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));        \
  } while (0)
#define BT(x)                 \
  do {                        \
    int rc_ = (x);            \
    if (rc_ != 0) return rc_; \
  } while (0)

template <typename T>
static int balloc(gh_batch *b, T **p, size_t n) {
  if (*p) {
    BH(b, hipFree(*p));
    *p = nullptr;
  }
  if (n) BH(b, hipMalloc((void **)p, sizeof(T) * n));
  return 0;
}

static BatchArgs bargs(const gh_batch *b) {
  BatchArgs A;
  A.off = b->off;
  A.order = b->order;
  A.first = 0;
  A.type = b->type;
  A.pos = b->pos;
  A.vel = b->vel;
  A.acc = b->acc;
  A.frc = b->frc;
  A.e = b->e;
  A.rho = b->rho;
  A.eband = b->eband;
  A.xold = b->xold;
  A.eband32 = b->eband32;
  A.gn = b->gn;
  A.nn = b->nn;
  A.list = b->list;
  A.cap = b->cap;
  A.ld = b->ld;
  A.tbd = b->tbd;
  A.tbf = b->tbf;
  A.e0 = b->tb_e0;
  A.nt = b->nt;
  A.nst = b->nst;
  for (int t = 0; t < GH_MAXT; ++t) {
    A.om[t] = 1.0 / b->mass_t[t];
    A.osm[t] = sqrt(A.om[t]);
  }
  const double rl = b->rcut + b->skin;
  A.rl2 = rl * rl;
  // FP32 prefilter of the build: its error on r^2 stays below 1e-4 A^2 for
  // staged clusters (< 64 A); the band is thirty times that
  const double margin = std::max(3e-3, 1e-4 * A.rl2);
  A.lo2f = (float)(A.rl2 - margin);
  A.hi2f = (float)(A.rl2 + margin);
  double rci = 1e300;
  for (int i = 0; i < b->nt; ++i)
    for (int j = 0; j < b->nt; ++j) rci = std::min(rci, b->tb[i * GH_MAXT + j].rci);
  A.near2f = (float)(rci * rci);
  A.ev_ui = ghc::ev_ui();
  A.T = b->T;
  A.out = b->out;
  A.flags = b->flags;
  return A;
}

// Reports a row overflow of the last call and re-arms the flags; the next
// call re-sizes the rows to the longest one asked for (plus a quarter).
static int check_flags(gh_batch *b) {
  int f[2] = {0, 0};
  BH(b, hipMemcpyAsync(f, b->flags, sizeof(f), hipMemcpyDeviceToHost, b->stream));
  BH(b, hipStreamSynchronize(b->stream));
  if (f[0]) {
    const int zero[2] = {0, 0};
    BH(b, hipMemcpyAsync(b->flags, zero, sizeof(zero), hipMemcpyHostToDevice, b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    b->lists_valid = false;
    b->forces_valid = false;
    b->cap_req = f[1] + std::max(16, f[1] / 4);
    const int cap = b->cap;
    b->cap = 0;
    return bfail(b, GH_EOVERFLOW,
                 "a cluster's neighbour row exceeded the capacity (" + std::to_string(f[1]) +
                     " > " + std::to_string(cap) + "); results of this call are invalid; "
                     "the next call re-sizes the rows");
  }
  return 0;
}

static int fetch_out(gh_batch *b, std::vector<double> &o) {
  o.resize((size_t)b->nsys * OUTW);
  BH(b, hipMemcpyAsync(o.data(), b->out, sizeof(double) * o.size(), hipMemcpyDeviceToHost,
                       b->stream));
  BH(b, hipStreamSynchronize(b->stream));
  return 0;
}

// capacity of a list row: every other atom of the largest system, capped by
// the densest packing within rcut+skin (fcc metals: < 90 neighbours)
static int batch_ready(gh_batch *b) {
  if (!b->nsys) return bfail(b, GH_ESTATE, "no systems");
  if (!b->have_pot) return bfail(b, GH_ESTATE, "no potential");
  if (b->cap == 0) {
    int nmax = 0;
    for (int s = 0; s < b->nsys; ++s) nmax = std::max(nmax, b->off_h[s + 1] - b->off_h[s]);
    const double rl = b->rcut + b->skin;
    // close packing density of the smallest pair distance (0.8 r0 at most)
    double rmin = 1e30;
    for (int i = 0; i < b->nt; ++i)
      for (int j = 0; j < b->nt; ++j) rmin = std::min(rmin, 0.8 * b->tb[i * GH_MAXT + j].r0);
    const double dens = 1.4142135623730951 / (rmin * rmin * rmin);
    int est = (int)(4.18879 * rl * rl * rl * dens) + 16;
    b->cap = std::max(8, std::min(nmax, std::max(est, b->cap_req)));
    b->cap = (b->cap + 7) / 8 * 8;
    BT(balloc(b, &b->list, (size_t)(b->cap + 4 * GH_LIST_PAD_QUADS) * b->ld));
    // LDS staging of the largest system (int4 + type byte per atom, and the
    // sentinel) beside the static tables; larger systems use global memory
    const size_t stat = (b->nt > 1 ? GH_MAXT * GH_MAXT : 1) * (sizeof(TbDev) + sizeof(TbF)) +
                        sizeof(double) * 8 * (GH_BBLOCK / 64) + 64;
    const size_t need = 17 * (size_t)(nmax + 1);
    b->nmax = nmax;
    b->nst = stat + need <= 65536 ? nmax + 1 : 0;
  }
  return 0;
}

static size_t batch_lds(const gh_batch *b) { return 17 * (size_t)b->nst; }

// One batch kernel over all systems.  The systems above GH_BIGSYS atoms bound
// the run (one workgroup each, largest first), so they get 1024-thread
// workgroups (four wavefronts per SIMD) on a second stream while the rest
// run on the batch stream; `big(first, count, stream)` and `rest(first,
// count, nmax, stream)` launch the two parts.  Variants whose 1024-thread
// form would spill registers pass big_ok = false and keep 512 threads.
template <class BIG, class REST>
static int split_launch(gh_batch *b, bool big_ok, BIG &&big, REST &&rest) {
  const int nbig = big_ok ? b->nbig : 0;
  if (nbig > 0) {
    BH(b, hipEventRecord(b->ev_fork, b->stream));
    BH(b, hipStreamWaitEvent(b->stream2, b->ev_fork, 0));
    big(0, nbig, b->stream2);
    BH(b, hipGetLastError());
  }
  if (b->nsys > nbig) {
    rest(nbig, b->nsys - nbig, nbig > 0 ? b->nmax_rest : b->nmax, b->stream);
    BH(b, hipGetLastError());
  }
  if (nbig > 0) {
    BH(b, hipEventRecord(b->ev_join, b->stream2));
    BH(b, hipStreamWaitEvent(b->stream, b->ev_join, 0));
  }
  return 0;
}

extern "C" {

int gh_batch_create(gh_batch **out, int device, int precision) {
  if (!out) return GH_EINVAL;
  *out = nullptr;
  if (precision != GH_FP64 && precision != GH_MIXED) return GH_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return GH_EHIP;
  gh_batch *b = new gh_batch();
  b->device = device;
  b->precision = precision;
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void **)&b->flags, sizeof(int) * 4) != hipSuccess ||
      hipMemsetAsync(b->flags, 0, sizeof(int) * 4, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess) {
    delete b;
    return GH_EHIP;
  }
  *out = b;
  return 0;
}

void gh_batch_destroy(gh_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->stream2) (void)hipStreamSynchronize(b->stream2);
  void *p[] = {b->off, b->order, b->type, b->nn, b->flags, b->pos, b->vel, b->acc, b->frc,
               b->e, b->rho, b->eband, b->xold, b->T, b->out, b->scale, b->eband32, b->list,
               b->tbd, b->tbf, b->gn};
  for (void *q : p)
    if (q) (void)hipFree(q);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->stream2) (void)hipStreamDestroy(b->stream2);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  delete b;
}

const char *gh_batch_last_error(const gh_batch *b) { return b ? b->err.c_str() : "null batch"; }

int gh_batch_set_precision(gh_batch *b, int precision) {
  if (!b) return GH_EINVAL;
  if (precision != GH_FP64 && precision != GH_MIXED) return bfail(b, GH_EINVAL, "precision");
  b->precision = precision;
  b->forces_valid = false;
  return 0;
}

int gh_batch_set_systems(gh_batch *b, int nsys, const int *off, const double *x,
                         const int *type) {
  if (!b || nsys <= 0 || !off || !x) return bfail(b, GH_EINVAL, "gh_batch_set_systems");
  if (off[0] != 0) return bfail(b, GH_EINVAL, "off[0] must be 0");
  for (int s = 0; s < nsys; ++s) {
    const int n = off[s + 1] - off[s];
    if (n < 1 || n > 65535) return bfail(b, GH_EINVAL, "systems need 1..65535 atoms");
  }
  BH(b, hipSetDevice(b->device));
  b->nsys = nsys;
  b->ntot = off[nsys];
  b->ld = (b->ntot + 255) / 256 * 256;
  const int ld = b->ld;
  b->off_h.assign(off, off + nsys + 1);
  b->order_h.resize(nsys);
  std::iota(b->order_h.begin(), b->order_h.end(), 0);
  std::stable_sort(b->order_h.begin(), b->order_h.end(), [&](int a, int c) {
    return off[a + 1] - off[a] > off[c + 1] - off[c];
  });
  b->nbig = 0;
  while (b->nbig < nsys && off[b->order_h[b->nbig] + 1] - off[b->order_h[b->nbig]] > GH_BIGSYS)
    ++b->nbig;
  b->nmax_rest = b->nbig < nsys ? off[b->order_h[b->nbig] + 1] - off[b->order_h[b->nbig]] : 0;
  b->T_h.assign(nsys, 0.0);
  BT(balloc(b, &b->off, nsys + 1));
  BT(balloc(b, &b->order, nsys));
  BT(balloc(b, &b->type, ld));
  BT(balloc(b, &b->nn, ld));
  BT(balloc(b, &b->pos, ld));
  BT(balloc(b, &b->vel, 3 * (size_t)ld));
  BT(balloc(b, &b->acc, 3 * (size_t)ld));
  BT(balloc(b, &b->frc, 3 * (size_t)ld));
  BT(balloc(b, &b->e, ld));
  BT(balloc(b, &b->rho, ld));
  BT(balloc(b, &b->eband, ld));
  BT(balloc(b, &b->eband32, ld));
  BT(balloc(b, &b->gn, 6 * (size_t)ld));
  BT(balloc(b, &b->xold, 3 * (size_t)ld));
  BT(balloc(b, &b->T, nsys));
  BT(balloc(b, &b->out, (size_t)nsys * OUTW));
  BT(balloc(b, &b->scale, nsys));
  b->cap = 0;
  b->cap_req = 0;
  if (b->list) {
    BH(b, hipFree(b->list));
    b->list = nullptr;
  }
  std::vector<double4> hp(ld, make_double4(0, 0, 0, 0));
  std::vector<int> ht(ld, 0);
  for (int i = 0; i < b->ntot; ++i) {
    hp[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], 0.0);
    ht[i] = type ? type[i] : 0;
    if (ht[i] < 0 || ht[i] >= GH_MAXT) return bfail(b, GH_EINVAL, "type out of range");
  }
  // every upload and clear is ordered on the batch stream (a non-blocking
  // stream does not wait for work on the null stream)
  hipStream_t st = b->stream;
  BH(b, hipMemcpyAsync(b->pos, hp.data(), sizeof(double4) * ld, hipMemcpyHostToDevice, st));
  BH(b, hipMemcpyAsync(b->type, ht.data(), sizeof(int) * ld, hipMemcpyHostToDevice, st));
  BH(b, hipMemcpyAsync(b->off, off, sizeof(int) * (nsys + 1), hipMemcpyHostToDevice, st));
  BH(b, hipMemcpyAsync(b->order, b->order_h.data(), sizeof(int) * nsys, hipMemcpyHostToDevice,
                       st));
  BH(b, hipMemsetAsync(b->vel, 0, sizeof(double) * 3 * ld, st));
  BH(b, hipMemsetAsync(b->acc, 0, sizeof(double) * 3 * ld, st));
  BH(b, hipMemsetAsync(b->frc, 0, sizeof(double) * 3 * ld, st));
  BH(b, hipMemsetAsync(b->out, 0, sizeof(double) * nsys * OUTW, st));
  BH(b, hipMemsetAsync(b->T, 0, sizeof(double) * nsys, st));
  BH(b, hipStreamSynchronize(st));
  b->lists_valid = false;
  b->forces_valid = false;
  return 0;
}

int gh_batch_set_smatb(gh_batch *b, int nt, const double *A, const double *xi, const double *p,
                       const double *q, const double *r0, const double *rci, const double *rce) {
  if (!b || nt <= 0 || nt > GH_MAXT) return bfail(b, GH_EINVAL, "gh_batch_set_smatb");
  // reuse the single-system parameter upload through a scratch context
  gh_ctx tmp;
  tmp.device = b->device;
  tmp.tbf_k = GH_BFXK;  // the FP32 forms in the unit of the staged positions
  int rc = gh_set_smatb(&tmp, nt, A, xi, p, q, r0, rci, rce);
  if (rc) return bfail(b, rc, tmp.err);
  for (int k = 0; k < GH_MAXT * GH_MAXT; ++k) b->tb[k] = tmp.tb[k];
  if (b->tbd) BH(b, hipFree(b->tbd));
  b->tbd = tmp.tbd;  // ownership moves to the batch
  tmp.tbd = nullptr;
  if (b->tbf) BH(b, hipFree(b->tbf));
  b->tbf = tmp.tbf;
  tmp.tbf = nullptr;
  b->tb_e0 = tmp.tb_e0;
  b->nt = nt;
  b->rcut = tmp.rcut;
  b->have_pot = true;
  b->cap = 0;
  b->lists_valid = false;
  b->forces_valid = false;
  return 0;
}

int gh_batch_set_masses(gh_batch *b, int nt, const double *mass) {
  if (!b || nt <= 0 || nt > GH_MAXT || !mass) return bfail(b, GH_EINVAL, "gh_batch_set_masses");
  for (int t = 0; t < nt; ++t) {
    if (!(mass[t] > 0)) return bfail(b, GH_EINVAL, "masses must be positive");
    b->mass_t[t] = mass[t];
  }
  return 0;
}

int gh_batch_set_skin(gh_batch *b, double skin) {
  if (!b || !(skin >= 0)) return bfail(b, GH_EINVAL, "skin must be >= 0");
  b->skin = skin;
  b->cap = 0;
  b->lists_valid = false;
  return 0;
}

int gh_batch_set_velocities(gh_batch *b, const double *v) {
  if (!b || !b->nsys) return bfail(b, GH_ESTATE, "no systems");
  std::vector<double> h(3 * (size_t)b->ld, 0.0);
  if (v)
    for (int i = 0; i < b->ntot; ++i)
      for (int d = 0; d < 3; ++d) h[(size_t)d * b->ld + i] = v[3 * (size_t)i + d];
  BH(b, hipMemcpyAsync(b->vel, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice,
                       b->stream));
  BH(b, hipStreamSynchronize(b->stream));
  return 0;
}

int gh_batch_set_temperatures(gh_batch *b, const double *T) {
  if (!b || !b->nsys || !T) return bfail(b, GH_ESTATE, "no systems");
  b->T_h.assign(T, T + b->nsys);
  BH(b, hipMemcpyAsync(b->T, T, sizeof(double) * b->nsys, hipMemcpyHostToDevice, b->stream));
  BH(b, hipStreamSynchronize(b->stream));
  return 0;
}

int gh_batch_compute(gh_batch *b, double *F, double *epot) {
  if (!b) return GH_EINVAL;
  BH(b, hipSetDevice(b->device));
  BT(batch_ready(b));
  const BatchArgs A = bargs(b);
  const bool mixed = b->precision == GH_MIXED, single = b->nt == 1;
  const size_t lds = batch_lds(b);
#define LAUNCH(MX, SG)                                                                     \
  BT(split_launch(                                                                         \
      b, true,                                                                             \
      [&](int first, int count, hipStream_t st) {                                          \
        BatchArgs A1 = A;                                                                  \
        A1.first = first;                                                                  \
        k_batch_compute<MX, SG, 1024><<<count, 1024, lds, st>>>(A1);                       \
      },                                                                                   \
      [&](int first, int count, int nmax, hipStream_t st) {                                \
        BatchArgs A1 = A;                                                                  \
        A1.first = first;                                                                  \
        if (nmax > 512)                                                                    \
          k_batch_compute<MX, SG, 512><<<count, 512, lds, st>>>(A1);                       \
        else                                                                               \
          k_batch_compute<MX, SG, 256><<<count, 256, lds, st>>>(A1);                       \
      }))
  if (mixed) {
    if (single) LAUNCH(true, true); else LAUNCH(true, false);
  } else {
    if (single) LAUNCH(false, true); else LAUNCH(false, false);
  }
#undef LAUNCH
  BT(check_flags(b));
  b->lists_valid = true;
  b->forces_valid = true;
  std::vector<double> o;
  BT(fetch_out(b, o));
  const double ui_ev = ghc::ui_ev();
  b->epot_h.resize(b->nsys);
  for (int s = 0; s < b->nsys; ++s) b->epot_h[s] = o[s * OUTW] * ui_ev;
  if (epot)
    for (int s = 0; s < b->nsys; ++s) epot[s] = b->epot_h[s];
  if (F) {
    std::vector<double> h(3 * (size_t)b->ld);
    BH(b, hipMemcpyAsync(h.data(), b->frc, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                         b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    for (int i = 0; i < b->ntot; ++i)
      for (int d = 0; d < 3; ++d) F[3 * (size_t)i + d] = h[(size_t)d * b->ld + i];
  }
  return 0;
}

int gh_batch_langevin(gh_batch *b, int nsteps, double dt, double gamma, uint64_t seed,
                      double *epot, double *ekin) {
  if (!b || nsteps < 0 || !(dt > 0) || !(gamma > 0))
    return bfail(b, GH_EINVAL, "langevin needs nsteps >= 0, dt > 0, gamma > 0");
  if (!b->forces_valid) BT(gh_batch_compute(b, nullptr, nullptr));
  ErmakC k;
  k.c0 = exp(-dt * gamma);
  k.c1 = (1.0 - k.c0) / gamma;
  k.c2 = (1.0 - k.c1 / dt) / gamma;
  k.sdr = sqrt(dt / gamma * (2.0 - (3.0 - 4.0 * k.c0 + k.c0 * k.c0) / (dt * gamma)));
  k.sdv = sqrt(1.0 - k.c0 * k.c0);
  k.crv1 = (1.0 - k.c0) * (1.0 - k.c0) / (gamma * k.sdr * k.sdv);
  k.crv2 = sqrt(1.0 - (k.crv1 * k.crv1));
  k.dt = dt;
  k.kB_ui = ghc::kB_ui();
  const BatchArgs A = bargs(b);
  const bool mixed = b->precision == GH_MIXED, single = b->nt == 1;
  const size_t lds = batch_lds(b);
  const long long step0 = b->step + 1;
  // one-type variants need more than the 128 VGPRs of a 1024-thread workgroup
#define LAUNCH(MX, SG)                                                                     \
  BT(split_launch(                                                                         \
      b, !SG,                                                                              \
      [&](int first, int count, hipStream_t st) {                                          \
        if constexpr (!SG) {                                                               \
          BatchArgs A1 = A;                                                                \
          A1.first = first;                                                                \
          k_batch_langevin<MX, SG, 1024><<<count, 1024, lds, st>>>(A1, nsteps, step0, k,   \
                                                                   seed, b->skin);         \
        }                                                                                  \
      },                                                                                   \
      [&](int first, int count, int nmax, hipStream_t st) {                                \
        BatchArgs A1 = A;                                                                  \
        A1.first = first;                                                                  \
        if (nmax > 512)                                                                    \
          k_batch_langevin<MX, SG, 512><<<count, 512, lds, st>>>(A1, nsteps, step0, k,     \
                                                                 seed, b->skin);           \
        else                                                                               \
          k_batch_langevin<MX, SG, 256><<<count, 256, lds, st>>>(A1, nsteps, step0, k,     \
                                                                 seed, b->skin);           \
      }))
  if (mixed) {
    if (single) LAUNCH(true, true); else LAUNCH(true, false);
  } else {
    if (single) LAUNCH(false, true); else LAUNCH(false, false);
  }
#undef LAUNCH
  b->step += nsteps;
  BT(check_flags(b));
  std::vector<double> o;
  BT(fetch_out(b, o));
  const double ui_ev = ghc::ui_ev();
  b->epot_h.resize(b->nsys);
  for (int s = 0; s < b->nsys; ++s) {
    b->epot_h[s] = o[s * OUTW] * ui_ev;
    if (epot) epot[s] = b->epot_h[s];
    if (ekin) ekin[s] = o[s * OUTW + 1] * ui_ev;
  }
  return 0;
}

// splitmix64: host stream for the Metropolis draws of the exchanges
static double smix_u01(uint64_t &st) {
  uint64_t z = (st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

int gh_batch_exchange(gh_batch *b, int nladder, int R, int *sys, int parity, uint64_t seed,
                      int *naccept) {
  if (!b || nladder <= 0 || R < 2 || !sys) return bfail(b, GH_EINVAL, "gh_batch_exchange");
  if ((int)b->epot_h.size() != b->nsys) return bfail(b, GH_ESTATE, "no energies yet");
  const double kB = ghc::kB_ui(), ev_ui = ghc::ev_ui();
  std::vector<double> scale(b->nsys, 1.0);
  uint64_t st = seed;
  int acc = 0;
  for (int l = 0; l < nladder; ++l)
    for (int r = parity & 1; r + 1 < R; r += 2) {
      const int lo = sys[l * R + r], hi = sys[l * R + r + 1];
      // ReplicaExchange.f90:86-110: deltab = beta - beta_partner,
      // deltau = (E_partner - E) deltab; swap if deltau <= 0 or u < e^{-deltau}
      const double beta = 1.0 / (kB * b->T_h[lo]), betap = 1.0 / (kB * b->T_h[hi]);
      const double du = (b->epot_h[hi] - b->epot_h[lo]) * ev_ui * (beta - betap);
      bool swap = true;
      if (du > 0) swap = smix_u01(st) < exp(-du);
      if (!swap) continue;
      ++acc;
      // the configurations trade temperatures: velocities scale by sqrt(T_new/T_old)
      std::swap(b->T_h[lo], b->T_h[hi]);
      scale[lo] = sqrt(b->T_h[lo] / b->T_h[hi]);
      scale[hi] = sqrt(b->T_h[hi] / b->T_h[lo]);
      std::swap(sys[l * R + r], sys[l * R + r + 1]);
    }
  BH(b, hipMemcpyAsync(b->T, b->T_h.data(), sizeof(double) * b->nsys, hipMemcpyHostToDevice,
                       b->stream));
  BH(b, hipMemcpyAsync(b->scale, scale.data(), sizeof(double) * b->nsys, hipMemcpyHostToDevice,
                       b->stream));
  k_batch_scale<<<b->nsys, 256, 0, b->stream>>>(b->nsys, b->off, b->scale, b->vel, b->ld);
  BH(b, hipGetLastError());
  BH(b, hipStreamSynchronize(b->stream));
  if (naccept) *naccept = acc;
  return 0;
}

int gh_batch_fire(gh_batch *b, int max_iter, double dt_max, double alpha0, double fswitch,
                  double ftol, int *iters, int *converged, double *epot, double *fmax) {
  if (!b || max_iter < 0 || !(dt_max > 0)) return bfail(b, GH_EINVAL, "gh_batch_fire");
  BH(b, hipSetDevice(b->device));
  BT(batch_ready(b));
  FireC fc;
  fc.dt_max = dt_max;
  fc.alpha0 = alpha0;
  fc.fswitch = fswitch;
  fc.ftol = ftol;
  fc.ui_ev = ghc::ui_ev();
  fc.max_iter = max_iter;
  fc.mixed = b->precision == GH_MIXED;
  const BatchArgs A = bargs(b);
  const size_t lds = batch_lds(b);
#define LAUNCH(SG)                                                                         \
  BT(split_launch(                                                                         \
      b, !SG,                                                                              \
      [&](int first, int count, hipStream_t st) {                                          \
        if constexpr (!SG) {                                                               \
          BatchArgs A1 = A;                                                                \
          A1.first = first;                                                                \
          k_batch_fire<SG, 1024><<<count, 1024, lds, st>>>(A1, fc, b->skin);               \
        }                                                                                  \
      },                                                                                   \
      [&](int first, int count, int nmax, hipStream_t st) {                                \
        BatchArgs A1 = A;                                                                  \
        A1.first = first;                                                                  \
        if (nmax > 512)                                                                    \
          k_batch_fire<SG, 512><<<count, 512, lds, st>>>(A1, fc, b->skin);                 \
        else                                                                               \
          k_batch_fire<SG, 256><<<count, 256, lds, st>>>(A1, fc, b->skin);                 \
      }))
  if (b->nt == 1) {
    LAUNCH(true);
  } else {
    LAUNCH(false);
  }
#undef LAUNCH
  BT(check_flags(b));
  b->lists_valid = true;
  b->forces_valid = false;  // the last forces may be FP64 while precision is mixed
  std::vector<double> o;
  BT(fetch_out(b, o));
  const double ui_ev = ghc::ui_ev();
  b->epot_h.resize(b->nsys);
  for (int s = 0; s < b->nsys; ++s) {
    b->epot_h[s] = o[s * OUTW] * ui_ev;
    if (epot) epot[s] = b->epot_h[s];
    if (fmax) fmax[s] = o[s * OUTW + 3];
    if (iters) iters[s] = (int)o[s * OUTW + 4];
    if (converged) converged[s] = (int)o[s * OUTW + 5];
  }
  return 0;
}

int gh_batch_get_state(gh_batch *b, double *x, double *v) {
  if (!b || !b->nsys) return bfail(b, GH_ESTATE, "no systems");
  if (x) {
    std::vector<double4> h(b->ld);
    BH(b, hipMemcpyAsync(h.data(), b->pos, sizeof(double4) * b->ld, hipMemcpyDeviceToHost,
                         b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    for (int i = 0; i < b->ntot; ++i) {
      x[3 * (size_t)i] = h[i].x;
      x[3 * (size_t)i + 1] = h[i].y;
      x[3 * (size_t)i + 2] = h[i].z;
    }
  }
  if (v) {
    std::vector<double> h(3 * (size_t)b->ld);
    BH(b, hipMemcpyAsync(h.data(), b->vel, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                         b->stream));
    BH(b, hipStreamSynchronize(b->stream));
    for (int i = 0; i < b->ntot; ++i)
      for (int d = 0; d < 3; ++d) v[3 * (size_t)i + d] = h[(size_t)d * b->ld + i];
  }
  return 0;
}

}  // extern "C"
