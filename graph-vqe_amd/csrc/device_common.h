// Device helpers shared by the gfx950 kernels: complex128 arithmetic and the
// FP64 MFMA 16x16x4 fragment machinery used for environment transfers.
//
// MFMA operand layout (v_mfma_f64_16x16x4_f64, wave64):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, register v holds D[(l/16) + 4v][l % 16]
// A 16x16x16 product is 4 k-steps; in k-step t lane group g = l/16 supplies
// contraction index k = g + 4t.  With that k order the D registers of one
// product are directly the A operand of the next (see Env below), so a
// chain of transfers runs in registers with no lane shuffles.
#pragma once
#include <hip/hip_runtime.h>

#include "gvmps_internal.h"

typedef double d4 __attribute__((ext_vector_type(4)));

// Contraction / row index carried by lane group g (= lane / 16) and register
// or k-step t.  It equals the D-register row of the MFMA output (measured on
// gfx950: D[i][j] sits in lane j + 16*(i % 4), register i / 4), so one
// product's D registers are the next product's A operand.
__device__ __forceinline__ constexpr int kdx(int g, int t) { return g + 4 * t; }

// Chain kernels keep an environment, two site fragments and the products in
// registers (~240 VGPRs); capping them at 256 registers gives two waves per
// SIMD, so one wave's memory latency hides behind the other's MFMAs.
#define GV_TWO_WAVES __attribute__((amdgpu_waves_per_eu(2, 2)))
// The real-arithmetic chains need about 120 registers: four waves per SIMD.
#define GV_FOUR_WAVES __attribute__((amdgpu_waves_per_eu(4, 4)))
// Real chains of I/Z-only forests keep no site fragment or X_i product past
// the Y products, and contract one inline term end at a time: they fit six
// waves per SIMD (80 registers; measured 2% faster than five, 4% than four).
#define GV_IZ_WAVES __attribute__((amdgpu_waves_per_eu(6, 6)))
// The same over merged site fragments (4 doubles per lane instead of 8).
#ifndef GV_IZM_WAVES
#define GV_IZM_WAVES __attribute__((amdgpu_waves_per_eu(6, 6)))
#endif
// The per-term form keeps more live state: five waves.
// Complex chains of I/Z-only forests (k_chains*_iz).
#ifndef GV_CIZ_WAVES
#define GV_CIZ_WAVES __attribute__((amdgpu_waves_per_eu(2, 2)))
#endif
#define GV_FIVE_WAVES __attribute__((amdgpu_waves_per_eu(5, 5)))
// The I/Z forms of the reverse-gradient chains (forward and backward).
#define GV_FWD_IZ_WAVES __attribute__((amdgpu_waves_per_eu(5, 5)))
#define GV_BWD_IZ_WAVES __attribute__((amdgpu_waves_per_eu(5, 5)))
#define GV_IZ_TERM_UNROLL 1

// Index of this wavefront in its workgroup as a wave-uniform value.  Without
// readfirstlane the compiler treats threadIdx.x >> 6 as per-lane, so every
// work-item descriptor derived from it is fetched with vector loads into
// VGPRs and all control flow on it runs under exec masks.
__device__ __forceinline__ int wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// This wave's work item under the XCD-aware block order: block b runs on XCD
// b % 8, and XCD x takes a contiguous range of logical blocks, so
// consecutive items (one instance's chains or leaves) share an XCD and its L2.
__device__ __forceinline__ int xcd_item() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = b & 7;
  const int logical = xcd * q8 + min(xcd, r8) + (b >> 3);
  return logical * (int)(blockDim.x >> 6) + wave_id();
}

// Wave-uniform load of read-only problem data (plans, work lists, tables the
// host uploads before the launch) through the constant address space: scalar
// loads into SGPRs.  A plain load of the same data becomes a per-lane vector
// load whenever the kernel also stores through other pointers, because the
// compiler cannot prove the table is not overwritten.  The words are read as
// ints: a struct copy would go through a generic reference and lose the
// address space.
template <typename T>
__device__ __forceinline__ T uload(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "uload reads whole dwords");
  typedef __attribute__((address_space(4))) const int cint;
  const cint *q = (const cint *)p;
  int w[sizeof(T) / 4];
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) w[i] = q[i];
  T v;
  __builtin_memcpy(&v, w, sizeof(T));
  return v;
}

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // a * conj(b)
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// Environment E (16x16 complex) in "E layout": lane (a + 16 g), register t
// holds E[a][kdx(g, t)].  Memory form is row-major E[a*16 + k].
struct Env {
  double re[4], im[4];
};

// Site tensor fragments: lane (j + 16 g) holds M_i[kdx(g, t)][j] for physical
// index i, where M_i is A_i (left sweeps / chains) or A_i^T (right sweeps).
struct Frag {
  double re[2][4], im[2][4];
};

__device__ __forceinline__ void env_load(const double2 *p, int lane, Env &e) {
  const int a = lane & 15, g = lane >> 4;
  const double2 *q = p + a * 16;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    double2 v = q[kdx(g, t)];
    e.re[t] = v.x;
    e.im[t] = v.y;
  }
}

__device__ __forceinline__ void env_store(double2 *p, int lane, const Env &e) {
  const int a = lane & 15, g = lane >> 4;
  double2 *q = p + a * 16;
#pragma unroll
  for (int t = 0; t < 4; ++t) q[kdx(g, t)] = make_double2(e.re[t], e.im[t]);
}

// Column-major slots (element (a, k) at k * 16 + a): each store or load
// instruction covers four whole rows of 16 elements, where the row-major form
// touches sixteen.  Used for slots that only the chain kernels write and read
// (chain hand-off spills, suffix stores and leaves of the complex energy
// forests, the gradient's checkpointed step environments).
__device__ __forceinline__ void env_load_cm(const double2 *p, int lane, Env &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const double2 v = p[kdx(g, t) * 16 + a];
    e.re[t] = v.x;
    e.im[t] = v.y;
  }
}

__device__ __forceinline__ void env_store_cm(double2 *p, int lane, const Env &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) p[kdx(g, t) * 16 + a] = make_double2(e.re[t], e.im[t]);
}

// Non-temporal forms, for slots written once and read once much later (the
// reverse gradient's checkpointed step environments).
__device__ __forceinline__ void nt_store2(double2 *p, double re, double im) {
  __builtin_nontemporal_store(re, &p->x);
  __builtin_nontemporal_store(im, &p->y);
}

__device__ __forceinline__ void env_load_cm_nt(const double2 *p, int lane, Env &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const double2 *q = p + kdx(g, t) * 16 + a;
    e.re[t] = __builtin_nontemporal_load(&q->x);
    e.im[t] = __builtin_nontemporal_load(&q->y);
  }
}

__device__ __forceinline__ void env_store_cm_nt(double2 *p, int lane, const Env &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) nt_store2(p + kdx(g, t) * 16 + a, e.re[t], e.im[t]);
}

// Forest environments of Hermitian Pauli words are Hermitian, and so are
// their adjoints.  The reverse-gradient kernels' hand-off slots keep the upper
// triangle: element (a, k), a <= k, at a (31 - a) / 2 + k (136 entries, 2176
// of 4096 bytes); loads take E[a][k] = conj(E[k][a]) below the diagonal.  The
// complex energy chains keep full slots: they are at their register cap, and
// the index arithmetic of packed slots makes them spill.
__device__ __forceinline__ int herm_idx(int a, int k) { return a <= k ? a * (31 - a) / 2 + k : k * (31 - k) / 2 + a; }

__device__ __forceinline__ void env_store_herm(double2 *p, int lane, const Env &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int k = kdx(g, t);
    if (k >= a) p[herm_idx(a, k)] = make_double2(e.re[t], e.im[t]);
  }
}

__device__ __forceinline__ void env_load_herm(const double2 *p, int lane, Env &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int k = kdx(g, t);
    const double2 v = p[herm_idx(a, k)];
    e.re[t] = v.x;
    e.im[t] = k < a ? -v.y : v.y;
  }
}

__device__ __forceinline__ void env_unit(int lane, Env &e) {  // [1] in the (0,0) corner
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    e.re[t] = (lane == 0 && t == 0) ? 1.0 : 0.0;
    e.im[t] = 0.0;
  }
}

// Site slot layout: element (l, i, r) at (l*2 + i)*16 + r.
__device__ __forceinline__ void frag_load(const double2 *slot, int lane, Frag &f, bool transposed) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int row = kdx(g, t);
      const int idx = transposed ? ((j * 2 + i) * 16 + row) : ((row * 2 + i) * 16 + j);
      double2 v = slot[idx];
      f.re[i][t] = v.x;
      f.im[i][t] = v.y;
    }
}

// Complex 16x16x16 products take three real MFMA products (Gauss):
//   (a + ib)(c + id) = (ac - bd) + i((a + b)(c + d) - ac - bd),
// 12 MFMAs instead of 16; the operand sums are four VALU adds per side.  The
// method is normwise stable (error <= a few eps ||A|| ||B|| per entry).

// X_i = E * M_i  (result in D layout: lane (j+16g), reg v = X_i[kdx(g,v)][j]).
__device__ __forceinline__ void prod_x(const Env &e, const Frag &f, int i, d4 &xr, d4 &xi) {
  d4 t1 = {0.0, 0.0, 0.0, 0.0}, t2 = {0.0, 0.0, 0.0, 0.0}, t3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    t1 = mfma(e.re[t], f.re[i][t], t1);
    t2 = mfma(e.im[t], f.im[i][t], t2);
    t3 = mfma(e.re[t] + e.im[t], f.re[i][t] + f.im[i][t], t3);
  }
  xr = t1 - t2;
  xi = t3 - t1 - t2;
}

// Y = M_j^H X, returned in E layout (computed as Y^T = X^T conj(M_j)):
//   Re = Xr Mr + Xi Mi,  Im = Xi Mr - Xr Mi = (Xr + Xi)(Mr - Mi) - Xr Mr + Xi Mi.
__device__ __forceinline__ void prod_y(const d4 &xr, const d4 &xi, const Frag &f, int j, d4 &yr, d4 &yi) {
  d4 t1 = {0.0, 0.0, 0.0, 0.0}, t2 = {0.0, 0.0, 0.0, 0.0}, t3 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    t1 = mfma(xr[t], f.re[j][t], t1);
    t2 = mfma(xi[t], f.im[j][t], t2);
    t3 = mfma(xr[t] + xi[t], f.re[j][t] - f.im[j][t], t3);
  }
  yr = t1 + t2;
  yi = t3 - t1 + t2;
}

// Identity transfer T_I(E) = sum_i M_i^H E M_i.
__device__ __forceinline__ void transfer_identity(const Env &e, const Frag &f, Env &out) {
  d4 xr, xi, y0r, y0i, y1r, y1i;
  prod_x(e, f, 0, xr, xi);
  prod_y(xr, xi, f, 0, y0r, y0i);
  prod_x(e, f, 1, xr, xi);
  prod_y(xr, xi, f, 1, y1r, y1i);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    out.re[v] = y0r[v] + y1r[v];
    out.im[v] = y0i[v] + y1i[v];
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ------------------------------------------- site adjoints of the backward
// A step's site adjoint arrives in the D layout: lane a16 + 16 g4, register
// v holds Mbar_i[kdx(g4, v)][a16] (rows = left bond, columns = right bond) in
// m0 (i = 0) and m1 (i = 1).  Before it is stored it is pulled back through
// the CNOT halves that end its site's local program (ProbDev::site_tail:
// count | last kind << 4 | kind before it << 8, at most one per bond side),
// the sums the local backward would otherwise make per site:
//   control, right:  Abar[l,i,r] = A'bar[l,i,2r+i]
//   target,  right:  Abar[l,i,r] = A'bar[l,i,2r] + A'bar[l,i^1,2r+1]
//   control, left:   Abar[l,i,r] = A'bar[2l+i,i,r]
//   target,  left:   Abar[l,i,r] = A'bar[2l,i,r] + A'bar[2l+1,i^1,r]
// (a CNOT half appends the new bond bit as the index's lowest bit).  Row
// parity is g4 & 1 and column parity a16 & 1, so the first pull-back leaves
// each lane one component i (its column parity after a right half, its row
// parity after a left half), pairing lanes 1 or 16 apart; a second one, on
// the other bond, pairs lanes 17 apart and keeps half of the lanes.  The
// result is stored compactly, ((l * 2 + i) * rc + r) over the pulled-back
// dimensions: 16x2x16 (no pull-back), 16x2x8, 8x2x16 or 8x2x8.

__device__ __forceinline__ bool tail_right_first(int tail) {
  const int k1 = (tail >> 4) & 15;
  return k1 == gv::LO_CTRL_R || k1 == gv::LO_TARG_R;
}
// Halved dimensions after the pull-back: rows (left bond), columns (right bond).
__device__ __forceinline__ bool tail_rows_half(int tail) {
  const int nt = tail & 3;
  return nt == 2 || (nt == 1 && !tail_right_first(tail));
}
__device__ __forceinline__ bool tail_cols_half(int tail) {
  const int nt = tail & 3;
  return nt == 2 || (nt == 1 && tail_right_first(tail));
}

// Where this lane's pulled-back values go: register v lands at
// ((lh ? row >> 1 : row) * 2 + c) * rc + col, row = kdx(g4, v), if keep.
struct TailMap {
  bool keep, lh;
  int c, rc, col;
};
__device__ __forceinline__ int tail_index(const TailMap &m, int g4, int v) {
  const int row = kdx(g4, v);
  return ((m.lh ? row >> 1 : row) * 2 + m.c) * m.rc + m.col;
}

// The pulled-back values of this lane per register, and where they go
// (tail count >= 1).  y may alias neither m0 nor m1.
__device__ __forceinline__ TailMap tail_pull(int tail, int a16, int g4, const d4 &m0, const d4 &m1, double y[4]) {
  const int nt = tail & 3, k1 = (tail >> 4) & 15;
  const bool r1 = k1 == gv::LO_CTRL_R || k1 == gv::LO_TARG_R;
  TailMap m;
  m.c = r1 ? (a16 & 1) : (g4 & 1);
  if (k1 == gv::LO_TARG_R || k1 == gv::LO_TARG_L) {
    const int x = r1 ? 1 : 16;
#pragma unroll
    for (int v = 0; v < 4; ++v) y[v] = m0[v] + __shfl_xor(m1[v], x);
  } else {
#pragma unroll
    for (int v = 0; v < 4; ++v) y[v] = m.c ? m1[v] : m0[v];
  }
  m.keep = true;
  if (nt == 2) {
    const int k2 = (tail >> 8) & 15;
    if (k2 == gv::LO_TARG_R || k2 == gv::LO_TARG_L) {
#pragma unroll
      for (int v = 0; v < 4; ++v) y[v] += __shfl_xor(y[v], 17);
      m.keep = ((r1 ? g4 : a16) & 1) == 0;
    } else {
      m.keep = (a16 & 1) == (g4 & 1);
    }
  }
  m.lh = !r1 || nt == 2;
  const bool rh = r1 || nt == 2;
  m.rc = rh ? 8 : 16;
  m.col = rh ? a16 >> 1 : a16;
  return m;
}

// Compact index of site-slot element e = (l*2 + i)*16 + r, or -1 when the
// element lies beyond the pulled-back dimensions (its adjoint is zero there).
__device__ __forceinline__ int tail_compact_index(int tail, int e) {
  const int l = e >> 5, i = (e >> 4) & 1, r = e & 15;
  const bool lh = tail_rows_half(tail), rh = tail_cols_half(tail);
  if ((lh && l >= 8) || (rh && r >= 8)) return -1;
  return (l * 2 + i) * (rh ? 8 : 16) + r;
}

// ------------------------------------------------ energy reduction
// One workgroup (256 threads) per instance: identity words' coefficients and
// the chain (and leaf) partial energies, Neumaier-compensated, reduced in a
// fixed order.  ssum / scomp: 256 doubles of LDS each.
__device__ __forceinline__ void reduce_energy_block(const gv::InstDev &inst, const gv::ProbDev &P,
                                                   const double *__restrict__ parts, double *__restrict__ energies,
                                                   int with_leaves, double *ssum, double *scomp) {
  const int tid = threadIdx.x;
  double sum = 0.0, comp = 0.0;
  auto add = [&](double term) {
    const double t = sum + term;
    comp += (fabs(sum) >= fabs(term)) ? (sum - t) + term : (term - t) + sum;
    sum = t;
  };
  for (int k = tid; k < P.nterms; k += 256)
    if (!P.active[k]) add(P.coeffs[2 * k]);
  const int nparts = P.nchains + (with_leaves ? P.nleaf : 0);
  for (int c = tid; c < nparts; c += 256) add(parts[inst.part_off + c]);
  ssum[tid] = sum;
  scomp[tid] = comp;
  __syncthreads();
  for (int off = 128; off >= 1; off >>= 1) {
    if (tid < off) {
      const double a = ssum[tid], b = ssum[tid + off];
      const double t = a + b;
      const double e = (fabs(a) >= fabs(b)) ? (a - t) + b : (b - t) + a;
      ssum[tid] = t;
      scomp[tid] += scomp[tid + off] + e;
    }
    __syncthreads();
  }
  if (tid == 0) energies[inst.out] = ssum[0] + scomp[0];
  __syncthreads();  // ssum / scomp are reused by the workgroup's next instance
}


