// Real copies and zero structure of prepared sites (gfx950): the per-slot
// and per-chain pieces of k_pack_real, k_site_masks and k_chain_flags
// (kernels_real.hip).
#pragma once
#include "device_common.h"

namespace gv {

// Zero structure of a site, per (instance, site, copy):
//   bits 0-3 / 4-7   row groups t (rows 4t..4t+3) of M_0 / M_1 with a nonzero
//   bit 8            the column supports of M_0 and M_1 are disjoint
//   bits 16-31       the columns of M_1
// Both products of a step contract over the rows of M_i in k-steps of four
// rows, so a k-step over an all-zero row group only adds exact zeros and can
// be skipped; the chain steps do so for the row split of bulk sites (M_0 in
// rows 0-7, M_1 in rows 8-15: groups 0-1 and 2-3), in straight-line code (a
// branch per k-step would break the MFMA issue).  With disjoint columns, X = E (M_0 + M_1) holds X_0 = E M_0 in
// the columns of M_0 and X_1 in those of M_1, and Y = (M_0 + M_1)^T X holds
// M_a^T X_b in the block (columns of M_a) x (columns of M_b): one product
// pair gives Y00, Y11, Y01 and Y10, each entry from the same products in the
// same order as the separate products (the rest being exact zeros).
#define GV_SM_DISJOINT 0x100u
// Bit 9: the rows of M_0 are even and those of M_1 odd.  The bond-doubling
// state of the reverse gradient has it in the transposed copy of every site
// whose program ends with a CNOT's control half (the doubled bond's new bit
// is its lowest, so A_i sits in the columns of parity i); the gradient's
// forward chains then contract over the rows in the order 0, 2, .., 14,
// 1, 3, .., 15, where the two components are a row split.
#define GV_SM_PARITY 0x200u

// Merged site fragment F = M_0 + M_1, row-major 16 x 16 per slot and copy.
#define GV_FSLOT (GV_CHI * GV_CHI)

// A site step has structure when its components are column-disjoint or row
// split (the two cases the merged fragments serve).
__device__ __forceinline__ bool site_structured(unsigned mw) {
  return (mw & GV_SM_DISJOINT) != 0u || (mw & 0xffu) == 0xc3u;
}

// Element e of the real copies of a site arena: the direct copy holds A
// (element (l, i, r) at (l*2 + i)*16 + r), the transposed copy A_i^T in the
// same layout, i.e. its element (row, i, j) is A[j][i][row].
__device__ __forceinline__ long long pack_transposed_src(long long e) {
  const long long base = e - (e % GV_SLOT);
  const int k = (int)(e % GV_SLOT), row = k >> 5, i = (k >> 4) & 1, j = k & 15;
  return base + (j * 2 + i) * 16 + row;
}

// Zero structure of one real site slot A (one wavefront, A's values as
// site_slot_load leaves them): for the direct
// copy (M_i = A_i) and the transposed one (M_i = A_i^T), the row groups of
// M_i with a nonzero, whether M_0 and M_1 have disjoint column supports, and
// the columns of M_1 (bit layout at GV_SM_DISJOINT); and the merged
// fragments F = A_0 + A_1 (direct copy, into fd) and its transpose
// (transposed copy, into ft), meaningful where the supports are disjoint.
// Lane j + 16 g holds a[i][t] = A[g + 4t][i][j].
__device__ __forceinline__ void site_slot_load(const double *A, int lane, double (&a)[2][4]) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int t = 0; t < 4; ++t) a[i][t] = A[((g + 4 * t) * 2 + i) * 16 + j];
}

__device__ __forceinline__ void site_mask_slot(const double (&av)[2][4], double *fd, double *ft, unsigned *sm2,
                                               int lane) {
  const int j = lane & 15, g = lane >> 4;
  unsigned rows[2] = {0u, 0u}, cols[2] = {0u, 0u};  // supports of A_i: rows l, columns r
  double f[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const double a = av[i][t];
      f[t] += a;
      const unsigned long long b = __ballot(a != 0.0);
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if ((b >> (16 * q)) & 0xffffull) rows[i] |= 1u << (q + 4 * t);
      cols[i] |= (unsigned)((b | (b >> 16) | (b >> 32) | (b >> 48)) & 0xffffull);
    }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    fd[(g + 4 * t) * 16 + j] = f[t];
    ft[j * 16 + g + 4 * t] = f[t];
  }
  auto groups = [](unsigned supp) {
    unsigned m = 0;
#pragma unroll
    for (int t = 0; t < 4; ++t)
      if ((supp >> (4 * t)) & 15u) m |= 1u << t;
    return m;
  };
  // direct copy: rows of M_i are rows of A_i, columns are columns; transposed: swapped
  auto parity = [](unsigned s0, unsigned s1) {
    return ((s0 & 0xaaaau) == 0u && (s1 & 0x5555u) == 0u) ? GV_SM_PARITY : 0u;
  };
  const unsigned direct = groups(rows[0]) | (groups(rows[1]) << 4) |
                          ((cols[0] & cols[1]) == 0u ? GV_SM_DISJOINT : 0u) | parity(rows[0], rows[1]) |
                          (cols[1] << 16);
  const unsigned transposed = groups(cols[0]) | (groups(cols[1]) << 4) |
                              ((rows[0] & rows[1]) == 0u ? GV_SM_DISJOINT : 0u) | parity(cols[0], cols[1]) |
                              (rows[1] << 16);
  if (lane == 0) {
    sm2[0] = direct;
    sm2[1] = transposed;
  }
}

// Structure the gradient's forward chains use: disjoint columns in the direct
// copy (prefix chains, merged products), the parity split in the transposed
// one (suffix chains, row-split products over the permuted rows).
__device__ __forceinline__ bool site_structured_grad(unsigned mw, int dir) {
  return (mw & (dir != 0 ? GV_SM_PARITY : GV_SM_DISJOINT)) != 0u;
}

// Chain flag of chain c of an instance: 1 when at most a quarter of its
// steps are on sites without structure (the chain then runs over the merged
// fragments; grad: the structure site_structured_grad asks for).  sm: the
// instance's mask words (two per site: direct, transposed).  Eight steps per
// round: their descriptor loads, then their mask loads, are issued together
// instead of one dependent pair per step.
__device__ __forceinline__ int chain_flag(const ProbDev &P, int c, const unsigned *sm, bool grad) {
  const ChainDev ch = P.chains[c];
  const unsigned *smc = sm + (ch.dir != 0 ? 1 : 0);
  int plain = 0;
  for (int s0 = 0; s0 < ch.nsteps; s0 += 8) {
    int site[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) site[q] = s0 + q < ch.nsteps ? P.steps[ch.step0 + s0 + q].site : -1;
    unsigned mw[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) mw[q] = site[q] >= 0 ? smc[2 * site[q]] : GV_SM_DISJOINT | GV_SM_PARITY;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      plain += (grad ? site_structured_grad(mw[q], ch.dir) : site_structured(mw[q])) ? 0 : 1;
  }
  return 4 * plain <= ch.nsteps ? 1 : 0;
}

}  // namespace gv
