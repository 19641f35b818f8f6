// Device helpers shared by the kernel files.
#pragma once

#include "gh_internal.h"

#define GH_BLOCK 256
#define GH_JMASK 0x07FFFFFFu
#define GH_NXCD 8

namespace ghd {

// Rebuild test on the displacement maximum the integrators fold into
// GhStatus::maxdisp2_bits (one atomicMax per workgroup).
__device__ __forceinline__ bool rebuild_due(const GhStatus *st, double skin) {
  const double d2 = __longlong_as_double((long long)st->maxdisp2_bits);
  return st->force_rebuild || 2.0 * sqrt(d2) > skin;
}

// Books the decision in st->rebuild once every kernel that tests
// rebuild_due has run: counts the rebuild, clears the host request and the
// row maximum, and restarts the displacement maximum.  Between rebuilds the
// maximum runs over all steps since the build (an upper bound of the
// current displacement, so a test on it never rebuilds late).
__device__ __forceinline__ void rebuild_book(GhStatus *st) {
  if (st->rebuild) {
    st->nupd += 1;
    st->force_rebuild = 0;
    st->maxnn = 0;
    st->maxdisp2_bits = 0ull;
  }
}

// Stops the run at `step` (the first raiser wins and records the step):
// every kernel that changes the atoms' state returns at once while
// st->halt is set, so the state stays the one the step started from.
__device__ __forceinline__ void raise_halt(GhStatus *st, int code, long long step) {
  if (atomicCAS(&st->halt, 0, code) == 0) st->halt_step = step;
}
// A kernel reads the stop as it stood when the kernel started (a plain,
// wave-uniform load: every earlier kernel's writes are visible then).
__device__ __forceinline__ int halted(const GhStatus *st) { return st->halt; }

// The step's first pass publishes, from its workgroup 0 as it starts, the
// step and the stop as they stand into pinned host memory (three stores;
// the host throttles on the step and reads the stop without a stream
// operation; a stop raised later in the step is seen at the next step's
// start, and always by the host's status read after a synchronisation).
__device__ __forceinline__ void step_publish(GhProgress *prog, long long step, int halt,
                                             long long halt_step, unsigned long long maxd2) {
  if (!prog) return;
  __hip_atomic_store(&prog->maxd2, maxd2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&prog->halt, halt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&prog->halt_step, halt_step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(&prog->step, step, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Workgroups are dealt round-robin to the 8 XCDs, each with its own L2.
// Remapping the workgroup id makes XCD k run a contiguous range of
// workgroups, i.e. one slab of cell-sorted atoms whose neighbours (and their
// positions) stay in that XCD's L2.  The grid is launched as a multiple of 8.
__device__ __forceinline__ int xcd_block() {
  const int per = gridDim.x / GH_NXCD;
  return (blockIdx.x % GH_NXCD) * per + blockIdx.x / GH_NXCD;
}
inline int xgrid(int nb) { return (nb + GH_NXCD - 1) / GH_NXCD * GH_NXCD; }

__device__ __forceinline__ unsigned ld_stream(const unsigned *p) {
  return __builtin_nontemporal_load(p);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ int wave_imax(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// |vd|^2 rounded exactly as the reference's dot_product (no contraction).
__device__ __forceinline__ double dist2_rn(double x, double y, double z) {
  return __dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z));
}

__device__ __forceinline__ void load_params(TbDev *sp, const TbDev *__restrict__ tbd, int nt) {
  for (int k = threadIdx.x; k < GH_MAXT * GH_MAXT; k += blockDim.x) {
    int a = k / GH_MAXT, b = k % GH_MAXT;
    if (a < nt && b < nt) sp[k] = tbd[k];
  }
}

// ---------------------------------------------------------------------------
// SMA-TB pair terms (TB.F90:405-423 density, 493-531 force), for r <= rce.
// FP64 form: the reference's arithmetic.  FP32 form: see TbDev.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double tb_band64(const TbDev &P, double dr) {
  if (dr < P.rci) return P.eps2 * exp(-2.0 * P.q * (dr * P.one_r0 - 1.0));
  const double d = dr - P.rce;
  const double d2 = d * d, d3 = d2 * d, d4 = d3 * d, d5 = d4 * d;
  const double b = P.b1 * d5 + P.b2 * d4 + P.b3 * d3;
  return b * b;
}

// factor = -dE/dr of the pair; repul = the pair's repulsion booked on atom i.
__device__ __forceinline__ void tb_pair64(const TbDev &P, double dr, double ebs, double &factor,
                                          double &repul) {
  if (dr <= P.rci) {
    const double t = dr * P.one_r0 - 1.0;
    repul = P.ca * exp(-P.p * t);
    factor = 2.0 * repul * P.p_r0 - P.eps2 * 0.5 * exp(-2.0 * P.q * t) * P.q2_r0 * ebs;
  } else {
    const double d = dr - P.rce;
    const double d2 = d * d, d3 = d2 * d, d4 = d3 * d, d5 = d4 * d;
    factor = (P.b1 * d5 + P.b2 * d4 + P.b3 * d3) *
             (5.0 * P.b1 * d4 + 4.0 * P.b2 * d3 + 3.0 * P.b3 * d2) * ebs;
    factor = factor - 2.0 * (5.0 * P.r1 * d4 + 4.0 * P.r2 * d3 + 3.0 * P.r3 * d2);
    repul = P.r1 * d5 + P.r2 * d4 + P.r3 * d3;
  }
}

// ---------------------------------------------------------------------------
// Neighbour rows ("quad ELL"): entry k of atom a is the 16-bit word
// ((k / 4) * ld + a) * 4 + k % 4, so one 8-byte load per lane fetches four
// consecutive entries.  Rows end with a padded quad whose spare entries name
// a sentinel slot far from every atom.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long ld_quad(const unsigned long long *p) {
  return __builtin_nontemporal_load(p);
}

// Fixed-point positions: staged atoms lie within 2^30 steps of their frame's
// origin and the sentinel slot holds INT_MIN in every coordinate, so the
// difference of two atoms never overflows 32 bits and the wrapped difference
// between the sentinel and any atom exceeds 2^30 steps in magnitude on every
// axis (far outside every cut).  Differences are taken in unsigned
// arithmetic (wrapping) and read back as int.
__device__ __forceinline__ int fxq(double d, double fxs) { return __double2int_rn(d * fxs); }
__device__ __forceinline__ double fxd(int a, int b) {
  return (double)(int)((unsigned)a - (unsigned)b);
}
#define GH_FXFAR ((int)0x80000000)
#define GH_FXLIM 1073741824.0  // 2^30
#define GH_FAR 1.0e30         // sentinel slot, FP64

struct RowWriter {  // quad-packed row of one atom
  unsigned long long *lq;
  int ld, a, cap, m;
  unsigned long long w;
  __device__ void put(int s) {
    w |= (unsigned long long)s << (16 * (m & 3));
    if ((m & 3) == 3) {
      if (m < cap) lq[(size_t)(m >> 2) * ld + a] = w;
      w = 0;
    }
    ++m;
  }
  __device__ void finish(int pad) {
    if ((m & 3) && m < cap) {
      for (int k = m & 3; k < 4; ++k) w |= (unsigned long long)pad << (16 * k);
      lq[(size_t)(m >> 2) * ld + a] = w;
    }
  }
};


// Runs f(slot) over an atom's list, one 8-byte load per four entries; the
// row's last quad is padded with the sentinel slot, so every entry of every
// quad is processed without a bounds test.  The loads run two quads ahead of
// use, unconditionally and through a running pointer: the list allocation
// holds GH_LIST_PAD_QUADS rows past its capacity, so the last prefetches
// stay in bounds (their values are never used).  No wait on the list stream
// sits between a quad's load and its use, and no 64-bit multiply or clamp
// sits in the loop.
template <class F>
__device__ __forceinline__ void for_each_slot(const unsigned short *__restrict__ list16, int ld,
                                              int a, int m, F &&f) {
  const int ng = (m + 3) >> 2;
  if (ng == 0) return;
  const unsigned long long *lp = reinterpret_cast<const unsigned long long *>(list16) + a;
  const size_t sld = (size_t)ld;
  unsigned long long w0 = ld_quad(lp);
  unsigned long long w1 = ld_quad(lp + sld);
  const unsigned long long *pn = lp + 2 * sld;
  for (int g = 0; g < ng; g += 2) {
    const unsigned long long n0 = ld_quad(pn);
    const unsigned long long n1 = ld_quad(pn + sld);
    pn += 2 * sld;
#pragma unroll
    for (int u = 0; u < 4; ++u) f((int)((w0 >> (16 * u)) & 0xFFFFull));
    if (g + 1 < ng) {
#pragma unroll
      for (int u = 0; u < 4; ++u) f((int)((w1 >> (16 * u)) & 0xFFFFull));
    }
    w0 = n0;
    w1 = n1;
  }
}


// As for_each_slot, one call per quad with its four packed slots (the
// caller issues the four LDS reads together).
template <class F>
__device__ __forceinline__ void for_each_quad(const unsigned short *__restrict__ list16, int ld,
                                              int a, int m, F &&f) {
  const int ng = (m + 3) >> 2;
  if (ng == 0) return;
  const unsigned long long *lp = reinterpret_cast<const unsigned long long *>(list16) + a;
  const size_t sld = (size_t)ld;
  unsigned long long w0 = ld_quad(lp);
  unsigned long long w1 = ld_quad(lp + sld);
  const unsigned long long *pn = lp + 2 * sld;
  for (int g = 0; g < ng; g += 2) {
    const unsigned long long n0 = ld_quad(pn);
    const unsigned long long n1 = ld_quad(pn + sld);
    pn += 2 * sld;
    f(w0);
    if (g + 1 < ng) f(w1);
    w0 = n0;
    w1 = n1;
  }
}
#define GH_QSLOT(w, u) ((int)(((w) >> (16 * (u))) & 0xFFFFull))

// Small systems: G consecutive lanes share one atom's row, lane `sub` taking
// quads sub, sub + G, ... (loaded four at a time before use, so a lane's list
// loads are in flight together); the partial sums meet in group_sum.  The
// order of the FP64 sums is fixed by G, which the host derives from the atom
// count alone.
template <int G, class F>
__device__ __forceinline__ void for_each_quad_g(const unsigned short *__restrict__ list16, int ld,
                                                int a, int m, int sub, F &&f) {
  const int ng = (m + 3) >> 2;
  const unsigned long long *lp = reinterpret_cast<const unsigned long long *>(list16) + a;
  const size_t sld = (size_t)ld;
  for (int g0 = sub; g0 < ng; g0 += 4 * G) {
    unsigned long long w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int g = g0 + u * G;
      w[u] = g < ng ? ld_quad(lp + (size_t)g * sld) : 0ull;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (g0 + u * G < ng) f(w[u]);
  }
}
template <int G, class F>
__device__ __forceinline__ void for_each_slot_g(const unsigned short *__restrict__ list16, int ld,
                                                int a, int m, int sub, F &&f) {
  for_each_quad_g<G>(list16, ld, a, m, sub, [&](unsigned long long w) {
#pragma unroll
    for (int u = 0; u < 4; ++u) f(GH_QSLOT(w, u));
  });
}
// Sum over the G lanes of an atom group (aligned groups of G lanes of one
// wavefront, all active together); every lane gets the total.
template <int G>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// SMA-TB FP32 pair forms (the TbF terms; r2 in the TbF unit, FP64).  r is
// the FP32 r2 times its hardware reciprocal square root y; for the inner
// branch s2 = 2 (r - r0f) is recovered to FP32 precision of s itself from
// the residual r2 - r_h^2, exact in FP64 (r_h^2 of a float is exact), so the
// rounding of r (a few 1e-7 A) never enters the exponentials; 2 r_h - 2 r0f
// is exact for r_h within a factor two of r0f.
__device__ __forceinline__ float tb_s2(const TbF &P, double r2, float y, float rh) {
  const float e = (float)fma(-(double)rh, (double)rh, r2);
  return fmaf(e, y, fmaf(rh, 2.0f, -P.r0q2));
}
// e^x - 1 to about one FP32 ulp of the result: x + x^2 (1/2 + x Q(x)) for
// |x| < 1/2 (Q a least-squares fit of (e^x - 1 - x - x^2/2)/x^3 there), else
// from the hardware exponential, evaluated only when a lane of the wave
// needs it (rows ordered by distance class keep the nearest shell inside).
__device__ __forceinline__ float tb_em1(float x) {
  float q = fmaf(1.9983829406555742e-04f, x, 1.3986654812470078e-03f);
  q = fmaf(q, x, 8.3331037312746048e-03f);
  q = fmaf(q, x, 4.1665747761726379e-02f);
  q = fmaf(q, x, 1.6666667163372040e-01f);
  float em1 = fmaf(x * x, fmaf(q, x, 0.5f), x);
  const bool far = !(fabsf(x) < 0.5f);
  if (__builtin_expect(__any(far), 0)) {
    const float t = __builtin_amdgcn_exp2f(x * 1.4426950408889634f) - 1.0f;
    if (far) em1 = t;
  }
  return em1;
}
// Density term of a pair within rce (smatb_preinteraction, TB.F90:405-423).
__device__ __forceinline__ bool tb_rho_fx(const TbF &P, double r2, float &band) {
  if (r2 > P.rce2q) return false;
  const float r2f = (float)r2;
  const float y = __builtin_amdgcn_rsqf(r2f);
  const float rh = r2f * y;
  if (rh < P.rciq) {
    band = P.rho_a * __builtin_amdgcn_exp2f(P.bqh * tb_s2(P, r2, y, rh));
  } else {
    const float d = fmaf(rh, P.sc, -P.rce);
    const float pp = d * d * d * fmaf(d, fmaf(d, P.b1f, P.b2f), P.b3f);
    band = pp * pp;
  }
  return true;
}
// Force of a pair within rce (smatb_interact, TB.F90:493-531): sf = factor/r
// per unit of the pair vector, repul the pair's repulsion (ENERGY only);
// dl = d_i + d_j (eband offsets from e0).
template <bool ENERGY>
__device__ __forceinline__ bool tb_force_fx(const TbF &P, double r2, float dl, float &sf,
                                            float &repul) {
  if (r2 > P.rce2q) return false;
  const float r2f = (float)r2;
  const float y = __builtin_amdgcn_rsqf(r2f);
  const float rh = r2f * y;
  float factor;
  if (rh <= P.rciq) {
    const float s2 = tb_s2(P, r2, y, rh);
    const float e2 = __builtin_amdgcn_exp2f(P.bqh * s2);
    const float em1 = tb_em1(P.gh * s2);
    factor = e2 * fmaf(P.a, em1, fmaf(-P.B0, dl, P.c0));
    if (ENERGY) {
      const float ar = P.Ar * e2;
      repul = fmaf(ar, em1, ar);
    }
  } else {
    const float d = fmaf(rh, P.sc, -P.rce);
    const float d2 = d * d, d3 = d2 * d;
    const float pp = d3 * fmaf(d, fmaf(d, P.b1f, P.b2f), P.b3f);
    const float dp = d2 * fmaf(d, fmaf(d, P.b1x5, P.b2x4), P.b3x3);
    const float drm = d2 * fmaf(d, fmaf(d, P.r1m10, P.r2m8), P.r3m6);  // -2 dR/dd
    factor = fmaf(pp * dp, P.e02 + dl, drm);
    if (ENERGY) repul = d3 * fmaf(d, fmaf(d, P.r1f, P.r2f), P.r3f);
  }
  sf = factor * y;
  return true;
}

// Tabulated Lennard-Jones pairs: energy p and f/r with f = -dp/dr, for r^2
// within the pair's cut.  POT: GH_POT_PLJ (plj_interact, Pairs.F90:293-370),
// GH_POT_SLJ (slj_interact, :403-493), GH_POT_WCA (wca_interact, :508-560).
// FP64 follows the reference's arithmetic; FP32 uses the hardware
// reciprocal square root.
template <int POT>
__device__ __forceinline__ void table_pair64(const PljDev &P, double r2, double &p,
                                             double &f_r) {
  const double dr = sqrt(r2);
  double f;
  if (POT == 6) {  // sm1: f here is minus the reference's attractive factor
    if (dr > P.rs) {
      const double w = P.rc - P.rs;
      const double a = M_PI * (dr - P.sa8) / w;
      p = (0.5 + 0.5 * sin(a)) * P.e - P.e;
      f = -(cos(a) * M_PI_2 * dr / w) * P.e;
    } else {
      p = -P.e;
      f = 0.0;
    }
  } else if (POT == 4 && !(dr < P.rs)) {
    const double d = dr - P.rc, d2 = d * d, s4 = d2 * d2;  // (dr - rcut)**4
    p = P.sa4 * s4 + P.sa8 * s4 * s4;
    f = -(4.0 * P.sa4 * s4 + 8.0 * P.sa8 * s4 * s4) / (dr - P.rc);
  } else {
    const double t = P.s / dr;
    const double t2 = t * t, t4 = t2 * t2, s6 = t4 * t2;  // (s/dr)**6
    if (POT == 3) p = 4.0 * P.e * (s6 * s6 - s6) - P.prc - P.frc * (dr - P.rc);
    else if (POT == 4) p = 4.0 * P.e * (s6 - 1.0) * s6;
    else p = 4.0 * P.e * (s6 * s6 - s6) + P.e;
    f = P.e * (48.0 * s6 - 24.0) * s6 / dr + (POT == 3 ? P.frc : 0.0);
  }
  f_r = POT == 6 && f == 0.0 ? 0.0 : f / dr;
}
template <int POT>
__device__ __forceinline__ void table_pair32(const PljDev &P, float r2f, float &p, float &f_r) {
  const float ir = __builtin_amdgcn_rsqf(r2f);
  const float dr = r2f * ir, ir2 = ir * ir;
  if (POT == 6) {
    if (dr > P.rsf) {
      const float w = P.rcf - P.rsf;
      const float a = (float)M_PI * (dr - P.sa8f) / w;
      p = (0.5f + 0.5f * sinf(a)) * P.ef - P.ef;
      f_r = -(cosf(a) * (float)M_PI_2 / w) * P.ef;
    } else {
      p = -P.ef;
      f_r = 0.0f;
    }
  } else if (POT == 4 && !(dr < P.rsf)) {
    const float d = dr - P.rcf, d2 = d * d, s4 = d2 * d2;
    p = fmaf(P.sa8f * s4, s4, P.sa4f * s4);
    f_r = -fmaf(8.0f * P.sa8f * s4, s4, 4.0f * P.sa4f * s4) / d * ir;
  } else {
    const float s2 = P.s2f * ir2;
    const float s6 = s2 * s2 * s2;
    if (POT == 3) p = 4.0f * P.ef * (s6 * s6 - s6) - P.prcf - P.frcf * (dr - P.rcf);
    else if (POT == 4) p = 4.0f * P.ef * (s6 - 1.0f) * s6;
    else p = 4.0f * P.ef * (s6 * s6 - s6) + P.ef;
    f_r = fmaf(P.ef * fmaf(48.0f, s6, -24.0f), s6, POT == 3 ? P.frcf * dr : 0.0f) * ir2;
  }
}

// Philox4x32-10 (Salmon et al., SC'11): counter-based stream.
__device__ __forceinline__ uint4 philox(uint4 c, uint2 k) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c.x;
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c.z;
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
    k.x += 0x9E3779B9u;
    k.y += 0xBB67AE85u;
  }
  return c;
}

// Six standard normals for atom `id` at step `step`: two Philox4x32-10 draws
// (counter (id, b, step), key = seed) give 24-bit uniforms, turned into three
// Box-Muller pairs in FP32 with the fast hardware log and sin/cos (the noise
// needs no more precision than FP32; its tail reaches 5.8 sigma).
__device__ __forceinline__ void philox_normals(unsigned long long seed, long long step, int id,
                                               double g[6]) {
  const uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  const unsigned s0 = (unsigned)step, s1 = (unsigned)((unsigned long long)step >> 32);
  const uint4 c0 = philox(make_uint4((unsigned)id, 0u, s0, s1), key);
  const uint4 c1 = philox(make_uint4((unsigned)id, 1u, s0, s1), key);
  const unsigned w[6] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y};
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const float u1 = ((float)(w[2 * b] >> 8) + 0.5f) * 5.9604644775390625e-08f;  // (0,1)
    const float a = (float)(w[2 * b + 1] >> 8) * 3.7450702829239286e-07f;        // 2 pi u
    const float r = __builtin_sqrtf(-2.0f * __logf(u1));
    g[2 * b] = (double)(r * __cosf(a));
    g[2 * b + 1] = (double)(r * __sinf(a));
  }
}

}  // namespace ghd
