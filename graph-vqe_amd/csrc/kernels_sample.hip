// Perfect sampling and the finite-shot Z-string estimator (gfx950).
//
//   k_sample  one wavefront per (instance, block of 16 shots).  The 16 shot
//             vectors are the columns of V^T (16 x 16, bond x shot) and stay in
//             MFMA registers for the whole site sweep:
//               W_i^T = A_i^T V^T          (ket amplitudes of bit i)
//               Z_i^T = Rt_{k+1} W_i^T      (right environment applied)
//               p_i[s] = Re sum_b Z_i[s][b] conj(W_i[s][b])
//             bit 0 iff u < p0/(p0+p1), u from the SplitMix64 stream of
//             (seed, site, shot); V <- W_bit / sqrt(p_bit).  This restates
//             k_sampleW / k_sampleProbabilities / k_chooseSample
//             (sharded_mps.cuh:103-188) with all shots of a block in one wave.
//             `bits` (one byte per site) may be null: then only the packed
//             words are written.
//   k_zsums   one thread per (instance, term): sum over shots of
//             (-1)^popcount(bits & mask) (k_zSampleSums, sharded_mps.cuh:193-214),
//             on bit-packed shot words instead of one byte per site.
#include "device_common.h"

using namespace gv;

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

extern "C" __global__ void __launch_bounds__(256)
k_sample(const InstDev *__restrict__ insts, int ninst, const double2 *__restrict__ arena,
         const double2 *__restrict__ envs, const unsigned *__restrict__ seeds, int shots, int blocks_per_inst, int stride,
         unsigned char *__restrict__ bits, unsigned long long *__restrict__ packed) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + wave_id();
  if (w >= ninst * blocks_per_inst) return;
  const int ii = w / blocks_per_inst, blk = w % blocks_per_inst;
  const InstDev inst = uload(insts + ii);
  const int n = inst.n;
  const double2 *sites = arena + inst.site_off;
  const double2 *Rt = envs + inst.env_off + (size_t)(n + 1) * GV_ENV;
  const unsigned seed = uload(seeds + ii);
  const int s_loc = lane & 15, g = lane >> 4;
  const int shot = blk * 16 + s_loc;
  const bool live = shot < shots;

  // V^T: lane (s + 16g), register t holds V[s][kdx(g,t)]; start from [1].
  double vr[4], vi[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    vr[t] = (kdx(g, t) == 0) ? 1.0 : 0.0;
    vi[t] = 0.0;
  }
  unsigned long long word0 = 0, word1 = 0;
  Frag f;
  Env r;
  for (int k = 0; k < n; ++k) {
    frag_load(sites + (size_t)k * GV_SLOT, lane, f, false);  // as A operand: A_i^T
    env_load(Rt + (size_t)(k + 1) * GV_ENV, lane, r);
    d4 wr[2], wi[2];
    double p[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      d4 ar = {0.0, 0.0, 0.0, 0.0}, ai = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        ar = mfma(f.re[i][t], vr[t], ar);
        ar = mfma(-f.im[i][t], vi[t], ar);
        ai = mfma(f.re[i][t], vi[t], ai);
        ai = mfma(f.im[i][t], vr[t], ai);
      }
      d4 zr = {0.0, 0.0, 0.0, 0.0}, zi = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        zr = mfma(r.re[t], ar[t], zr);
        zr = mfma(-r.im[t], ai[t], zr);
        zi = mfma(r.re[t], ai[t], zi);
        zi = mfma(r.im[t], ar[t], zi);
      }
      double acc = 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) acc += zr[v] * ar[v] + zi[v] * ai[v];  // Re(z conj(w))
      acc += __shfl_xor(acc, 16, 64);
      acc += __shfl_xor(acc, 32, 64);
      p[i] = acc > 0.0 ? acc : 0.0;
      wr[i] = ar;
      wi[i] = ai;
    }
    double p0 = p[0], p1 = p[1], total = p0 + p1;
    if (total <= 0.0) {
      p0 = p1 = 0.5;
      total = 1.0;
    }
    const unsigned long long raw =
        splitmix64(((unsigned long long)seed << 32) ^ ((unsigned long long)k << 20) ^ (unsigned long long)shot);
    const double u = (double)(raw >> 11) * 0x1.0p-53;
    const int sel = u < p0 / total ? 0 : 1;
    const double ps = sel ? p1 : p0;
    const double inv = ps > 0.0 ? 1.0 / sqrt(ps) : 1.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      vr[t] = (sel ? wr[1][t] : wr[0][t]) * inv;
      vi[t] = (sel ? wi[1][t] : wi[0][t]) * inv;
    }
    if (live && g == 0) {
      if (bits) bits[((size_t)ii * shots + shot) * stride + k] = (unsigned char)sel;
      if (sel) {
        if (k < 64) word0 |= 1ULL << k;
        else word1 |= 1ULL << (k - 64);
      }
    }
  }
  if (live && g == 0) {
    packed[((size_t)ii * shots + shot) * 2 + 0] = word0;
    packed[((size_t)ii * shots + shot) * 2 + 1] = word1;
    // rows of smaller instances in a mixed batch: zero the padding columns
    if (bits)
      for (int k = n; k < stride; ++k) bits[((size_t)ii * shots + shot) * stride + k] = 0;
  }
}

// Real circuits (every gate matrix real) leave a real MPS, so the right
// environments and the shot vectors are real: the same sweep in real
// arithmetic, one MFMA product per complex product (16 MFMAs per site
// instead of 64), on the packed real arena (A_i, then A_i^T, 512 doubles per
// site slot).  The complex kernels add only exact zeros to the same real
// accumulations, so both give the same probabilities and bits.
namespace {

// Real 16x16 environment in E layout, row-major in memory (2 KiB of a slot).
__device__ __forceinline__ void renv_ld(const double *p, int lane, double (&e)[4]) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) e[t] = p[a * 16 + kdx(g, t)];
}

__device__ __forceinline__ void renv_st(double *p, int lane, const double (&e)[4]) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) p[a * 16 + kdx(g, t)] = e[t];
}

// lane (j + 16 g), register t: M_i[kdx(g, t)][j] of a real site slot.
__device__ __forceinline__ void rfrag_ld(const double *slot, int lane, double (&f)[2][4]) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int t = 0; t < 4; ++t) f[i][t] = slot[(kdx(g, t) * 2 + i) * 16 + j];
}

}  // namespace

// Transposed right environments Rt_s = sum_i (A_i^T)^T Rt_{s+1} A_i^T of
// every instance, one wavefront each, as real row-major 16 x 16 blocks at
// the start of each 4 KiB environment slot (k_envs' Rt slots).
extern "C" __global__ void __launch_bounds__(256)
k_renvs_right(const InstDev *__restrict__ insts, int ninst, const double *__restrict__ rarenaT,
              double2 *__restrict__ envs) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + wave_id();
  if (w >= ninst) return;
  const InstDev inst = uload(insts + w);
  const int n = inst.n;
  const double *sites = rarenaT + inst.site_off;
  double2 *Rt = envs + inst.env_off + (size_t)(n + 1) * GV_ENV;
  double e[4], f[2][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) e[t] = (lane == 0 && t == 0) ? 1.0 : 0.0;
  renv_st((double *)(Rt + (size_t)n * GV_ENV), lane, e);
  for (int s = n - 1; s >= 0; --s) {
    rfrag_ld(sites + (size_t)s * GV_SLOT, lane, f);
    d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) x = mfma(e[t], f[i][t], x);   // X_i = E M_i
      d4 yi = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) yi = mfma(x[t], f[i][t], yi);  // M_i^T X_i (E layout)
      y = i == 0 ? yi : y + yi;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) e[t] = y[t];
    renv_st((double *)(Rt + (size_t)s * GV_ENV), lane, e);
  }
}

extern "C" __global__ void __launch_bounds__(256)
k_sample_real(const InstDev *__restrict__ insts, int ninst, const double *__restrict__ rarena,
              const double2 *__restrict__ envs, const unsigned *__restrict__ seeds, int shots, int blocks_per_inst,
              int stride, unsigned char *__restrict__ bits, unsigned long long *__restrict__ packed) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + wave_id();
  if (w >= ninst * blocks_per_inst) return;
  const int ii = w / blocks_per_inst, blk = w % blocks_per_inst;
  const InstDev inst = uload(insts + ii);
  const int n = inst.n;
  const double *sites = rarena + inst.site_off;
  const double2 *Rt = envs + inst.env_off + (size_t)(n + 1) * GV_ENV;
  const unsigned seed = uload(seeds + ii);
  const int s_loc = lane & 15, g = lane >> 4;
  const int shot = blk * 16 + s_loc;
  const bool live = shot < shots;
  double vr[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) vr[t] = (kdx(g, t) == 0) ? 1.0 : 0.0;
  unsigned long long word0 = 0, word1 = 0;
  double f[2][4], r[4];
  for (int k = 0; k < n; ++k) {
    rfrag_ld(sites + (size_t)k * GV_SLOT, lane, f);   // as A operand: A_i^T
    renv_ld((const double *)(Rt + (size_t)(k + 1) * GV_ENV), lane, r);
    d4 wr[2];
    double p[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      d4 ar = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) ar = mfma(f[i][t], vr[t], ar);
      d4 zr = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) zr = mfma(r[t], ar[t], zr);
      double acc = 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) acc += zr[v] * ar[v];
      acc += __shfl_xor(acc, 16, 64);
      acc += __shfl_xor(acc, 32, 64);
      p[i] = acc > 0.0 ? acc : 0.0;
      wr[i] = ar;
    }
    double p0 = p[0], p1 = p[1], total = p0 + p1;
    if (total <= 0.0) {
      p0 = p1 = 0.5;
      total = 1.0;
    }
    const unsigned long long raw =
        splitmix64(((unsigned long long)seed << 32) ^ ((unsigned long long)k << 20) ^ (unsigned long long)shot);
    const double u = (double)(raw >> 11) * 0x1.0p-53;
    const int sel = u < p0 / total ? 0 : 1;
    const double ps = sel ? p1 : p0;
    const double inv = ps > 0.0 ? 1.0 / sqrt(ps) : 1.0;
#pragma unroll
    for (int t = 0; t < 4; ++t) vr[t] = (sel ? wr[1][t] : wr[0][t]) * inv;
    if (live && g == 0) {
      if (bits) bits[((size_t)ii * shots + shot) * stride + k] = (unsigned char)sel;
      if (sel) {
        if (k < 64) word0 |= 1ULL << k;
        else word1 |= 1ULL << (k - 64);
      }
    }
  }
  if (live && g == 0) {
    packed[((size_t)ii * shots + shot) * 2 + 0] = word0;
    packed[((size_t)ii * shots + shot) * 2 + 1] = word1;
    if (bits)
      for (int k = n; k < stride; ++k) bits[((size_t)ii * shots + shot) * stride + k] = 0;
  }
}

// Term masks in site order, two 64-bit words per term (n <= 128).
extern "C" __global__ void __launch_bounds__(256)
k_zsums(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
        const unsigned long long *__restrict__ const *masks, const unsigned long long *__restrict__ packed,
        int shots, double2 *__restrict__ terms) {
  const int ii = blockIdx.y;
  const InstDev inst = uload(insts + ii);
  const ProbDev P = uload(probs + inst.prob);
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= P.nterms) return;
  const unsigned long long m0 = masks[inst.prob][2 * k], m1 = masks[inst.prob][2 * k + 1];
  const unsigned long long *pk = packed + (size_t)ii * shots * 2;
  int sum = 0;
  for (int s = 0; s < shots; ++s) {
    const int par = (__popcll(pk[2 * s] & m0) + __popcll(pk[2 * s + 1] & m1)) & 1;
    sum += par ? -1 : 1;
  }
  terms[inst.term_off + k] = make_double2((double)sum / shots, 0.0);
}

namespace gv {

hipError_t kl_sample(int ninst, int shots, int stride, hipStream_t s, const InstDev *insts, const double2 *arena,
                     const double2 *envs, const unsigned *seeds, unsigned char *bits, unsigned long long *packed) {
  const int bpi = (shots + 15) / 16;
  const int waves = ninst * bpi;
  hipLaunchKernelGGL(k_sample, dim3((waves + 3) / 4), dim3(256), 0, s, insts, ninst, arena, envs, seeds, shots, bpi, stride,
                     bits, packed);
  return hipGetLastError();
}

hipError_t kl_renvs_right(int ninst, hipStream_t s, const InstDev *insts, const double *rarena, long long nsite,
                          double2 *envs) {
  hipLaunchKernelGGL(k_renvs_right, dim3((ninst + 3) / 4), dim3(256), 0, s, insts, ninst, rarena + nsite, envs);
  return hipGetLastError();
}

hipError_t kl_sample_real(int ninst, int shots, int stride, hipStream_t s, const InstDev *insts, const double *rarena,
                          const double2 *envs, const unsigned *seeds, unsigned char *bits, unsigned long long *packed) {
  const int bpi = (shots + 15) / 16;
  const int waves = ninst * bpi;
  hipLaunchKernelGGL(k_sample_real, dim3((waves + 3) / 4), dim3(256), 0, s, insts, ninst, rarena, envs, seeds, shots,
                     bpi, stride, bits, packed);
  return hipGetLastError();
}

hipError_t kl_zsums(int ninst, int max_terms, int shots, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                    const unsigned long long *const *masks, const unsigned long long *packed, double2 *terms) {
  hipLaunchKernelGGL(k_zsums, dim3((max_terms + 255) / 256, ninst), dim3(256), 0, s, insts, probs, masks, packed,
                     shots, terms);
  return hipGetLastError();
}

}  // namespace gv
