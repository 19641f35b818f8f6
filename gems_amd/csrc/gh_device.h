// Device helpers shared by the kernel files.
#pragma once

#include "gh_internal.h"

#define GH_BLOCK 256
#define GH_JMASK 0x07FFFFFFu
#define GH_NXCD 8

namespace ghd {

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

// FP32 forms take TbDev or TbF (same FP32 field names).
template <class TP>
__device__ __forceinline__ float tb_band32(const TP &P, float r) {
  if (r < P.rci_f) return P.rho_a * __builtin_amdgcn_exp2f(P.rho_b * (r - P.r0f));
  const float d = r - P.rce_f;
  const float pp = d * d * d * fmaf(d, fmaf(d, P.b1f, P.b2f), P.b3f);
  return pp * pp;
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

template <class TP>
__device__ __forceinline__ void tb_pair32(const TP &P, float r, float ebs, float &factor,
                                          float &repul) {
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

__device__ __forceinline__ double u01(unsigned a, unsigned b) {
  // 53-bit uniform in (0,1)
  const unsigned long long v = ((unsigned long long)a << 21) ^ (unsigned long long)b;
  return ((double)(v & ((1ull << 53) - 1)) + 0.5) * (1.0 / 9007199254740992.0);
}

// Six standard normals for atom `id` at step `step` (Box-Muller).
__device__ __forceinline__ void philox_normals(unsigned long long seed, long long step, int id,
                                               double g[6]) {
  const uint2 key = make_uint2((unsigned)seed, (unsigned)(seed >> 32));
  double u[6];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const uint4 c = philox(make_uint4((unsigned)id, (unsigned)b, (unsigned)step,
                                      (unsigned)((unsigned long long)step >> 32)), key);
    u[2 * b] = u01(c.x, c.y);
    u[2 * b + 1] = u01(c.z, c.w);
  }
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    const double r = sqrt(-2.0 * log(u[2 * b]));
    double sn, cs;
    sincospi(2.0 * u[2 * b + 1], &sn, &cs);
    g[2 * b] = r * cs;
    g[2 * b + 1] = r * sn;
  }
}

}  // namespace ghd
