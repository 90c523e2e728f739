// Real-arithmetic energy path (gfx950).
//
// A circuit whose gate matrices are all real (RY, CX, H, X, Z, SWAP, ...)
// applied to |0...0> keeps a real MPS: the Theta products, Householder
// reflectors (e = x_j / |x_j| = +-1), Jacobi rotations (gamma real) and
// rank rule of the preparation leave every imaginary part exactly zero.  For
// Pauli words without Y every forest environment is then real as well
// (I, X and Z are real matrices), and the chains can run in real arithmetic:
//
//   complex path (Gauss form)   t1 = Er Mr, t2 = Ei Mi = 0, t3 = (Er+Ei)(Mr+Mi) = t1
//   real path                   X = E M   (one MFMA product per complex product)
//
// 16 MFMAs per I/Z step instead of 48, 4 KiB site slots instead of 8 KiB, and
// symmetric environments kept as packed upper triangles (1088 bytes) in the
// 4 KiB spill, store and leaf slots.  The packed real arena holds every site
// twice: as A_i and as A_i^T (suffix-forest chains read the transposed copy),
// so every fragment load is a row load with 128-byte lines per 16 lanes.  The
// chain/forest schedule, work items, term ends and partial-energy layout are
// those of kernels_energy.hip, so k_reduce_energy / k_reduce combine the
// results unchanged.
#include "device_common.h"
#include "site_pack.h"

using namespace gv;

namespace {

// Real environment, E layout: lane (a + 16 g), register t holds E[a][kdx(g, t)].
struct REnv {
  double v[4];
};

// Real site fragments: lane (j + 16 g) holds M_i[kdx(g, t)][j].
struct RFrag {
  double v[2][4];
};

// Slot k of a complex 4 KiB slot array, used as 256 doubles (row-major E[a][k]).
__device__ __forceinline__ double *rslot(double2 *base, long long k) { return (double *)(base + k * GV_ENV); }
__device__ __forceinline__ const double *rslot(const double2 *base, long long k) {
  return (const double *)(base + k * GV_ENV);
}

__device__ __forceinline__ void renv_load(const double *p, int lane, REnv &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) e.v[t] = p[a * 16 + kdx(g, t)];
}

__device__ __forceinline__ void renv_store(double *p, int lane, const REnv &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) p[a * 16 + kdx(g, t)] = e.v[t];
}

__device__ __forceinline__ void renv_unit(int lane, REnv &e) {
#pragma unroll
  for (int t = 0; t < 4; ++t) e.v[t] = (lane == 0 && t == 0) ? 1.0 : 0.0;
}

// Environments of I/X/Z words over a real MPS are symmetric.  The energy
// kernels' slots (chain hand-off spills, suffix-forest stores, deferred prefix
// leaves) keep the upper triangle: element (a, k), a <= k, at a (31 - a) / 2 + k, 136 doubles
// (9 of the 16 cache lines of a full 2 KiB environment).  Loads mirror it.
// (Column by column, k (k + 1) / 2 + a, makes the stores contiguous but its
// index arithmetic spills the I/Z kernels: 2.03 -> 2.13 ms.)
__device__ __forceinline__ int sym_idx(int a, int k) { return a <= k ? a * (31 - a) / 2 + k : k * (31 - k) / 2 + a; }

__device__ __forceinline__ void renv_store_sym(double *p, int lane, const REnv &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int k = kdx(g, t);
    if (k >= a) p[sym_idx(a, k)] = e.v[t];
  }
}

__device__ __forceinline__ void renv_load_sym(const double *p, int lane, REnv &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) e.v[t] = p[sym_idx(a, kdx(g, t))];
}

// The gradient's checkpointed step inputs are packed triangles too, written
// once by the forward pass and read once by the backward pass much later, so
// both go non-temporal.  The forward pass stores them contiguously: lane l
// writes packed elements 2l and 2l+1, lanes 0-3 also 128 + 2l and 129 + 2l,
// read from the wave's 16 x 17 LDS tile at the positions pack_tile_offsets
// gives (in doubles); the backward pass gathers them with renv_load_sym_nt.
struct PackTile {
  int o[4];
};
typedef double gv_d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ PackTile pack_tile_offsets(int lane) {
  PackTile pt;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int p = q < 2 ? 2 * lane + q : min(128 + 2 * lane + (q - 2), 135);
    int a = 0;  // row of packed element p: rows start at r (33 - r) / 2
#pragma unroll
    for (int r = 1; r < 16; ++r) a += p >= r * (33 - r) / 2 ? 1 : 0;
    pt.o[q] = a * 17 + (p - a * (31 - a) / 2);
  }
  return pt;
}

// Symmetrised environment: each lane keeps its entries on and above the
// diagonal and takes the mirror of those below it, through the wave's tile,
// so E[a][k] and E[k][a] are the same bits.  The forward pass continues every
// chain from the symmetrised input it checkpoints (STORE: into ck, as a
// packed triangle, half the bytes of the square), and the backward pass
// symmetrises the inputs it recomputes, so every checkpoint stride sees the
// same environments.
// Row order of the parity-split products: 0, 2, .., 14, 1, 3, .., 15 (row x of
// the permuted matrices is row par_perm(x) of the natural ones).
__device__ __forceinline__ constexpr int par_perm(int x) { return x < 8 ? 2 * x : 2 * x - 15; }

// perm: e leaves as P E P^T (both indices in par_perm order) instead, the
// form the parity-split products contract; the checkpoint stays natural.
template <bool STORE>
__device__ __forceinline__ void renv_symmetrize(REnv &e, int lane, double (*tl)[17], const int *ptab, double *ck,
                                                bool perm = false) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) tl[a][kdx(g, t)] = e.v[t];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  if (perm) {
    const int pa = par_perm(a);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int pk = par_perm(kdx(g, t));
      e.v[t] = tl[min(pa, pk)][max(pa, pk)];
    }
  } else {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = kdx(g, t);
      const double m = tl[k][a];
      if (k < a) e.v[t] = m;
    }
  }
  if (STORE && ck) {  // the upper triangle, packed, non-temporal
    const int4 o = *reinterpret_cast<const int4 *>(ptab + 4 * lane);
    const double *tf = &tl[0][0];
    gv_d2 *q = reinterpret_cast<gv_d2 *>(ck);
    __builtin_nontemporal_store(gv_d2{tf[o.x], tf[o.y]}, q + lane);
    if (lane < 4) __builtin_nontemporal_store(gv_d2{tf[o.z], tf[o.w]}, q + 64 + lane);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void renv_load_sym_nt(const double *p, int lane, REnv &e) {
  const int a = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) e.v[t] = __builtin_nontemporal_load(p + sym_idx(a, kdx(g, t)));
}

// Real site slot layout: element (l, i, r) at (l*2 + i)*16 + r (512 doubles);
// in the transposed copy the slot holds A_i^T in the same layout.
__device__ __forceinline__ void rfrag_load(const double *slot, int lane, RFrag &f) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int t = 0; t < 4; ++t) f.v[i][t] = slot[(kdx(g, t) * 2 + i) * 16 + j];
}

// rfrag_load with the rows in par_perm order.
__device__ __forceinline__ void rfrag_load_perm(const double *slot, int lane, RFrag &f) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int t = 0; t < 4; ++t) f.v[i][t] = slot[(par_perm(kdx(g, t)) * 2 + i) * 16 + j];
}

// X_i = E M_i (D layout), the Gauss form's t1 of prod_x.
__device__ __forceinline__ d4 rprod_x(const REnv &e, const RFrag &f, int i) {
  d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 4; ++t) x = mfma(e.v[t], f.v[i][t], x);
  return x;
}

// Y = M_j^T X in E layout (computed as Y^T = X^T M_j), the t1 of prod_y.
__device__ __forceinline__ d4 rprod_y(const d4 &x, const RFrag &f, int j) {
  d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 4; ++t) y = mfma(x[t], f.v[j][t], y);
  return y;
}

// Zero structure of the sites (site_pack.h): the mask words select, per
// step, the merged and row-split products below.
// Products over the k-steps in the compile-time set G (row groups of M_i).
template <unsigned G>
__device__ __forceinline__ d4 rprod_x_g(const REnv &e, const RFrag &f, int i) {
  d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 4; ++t)
    if (G & (1u << t)) x = mfma(e.v[t], f.v[i][t], x);
  return x;
}

template <unsigned G>
__device__ __forceinline__ d4 rprod_y_g(const d4 &x, const RFrag &f, int j) {
  d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 4; ++t)
    if (G & (1u << t)) y = mfma(x[t], f.v[j][t], y);
  return y;
}

// Y = F^T (E F) for F = M_0 + M_1 (disjoint columns), E layout.
__device__ __forceinline__ d4 rprod_merged(const REnv &e, const RFrag &f) {
  double F[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) F[t] = f.v[0][t] + f.v[1][t];
  d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 4; ++t) x = mfma(e.v[t], F[t], x);
  d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 4; ++t) y = mfma(x[t], F[t], y);
  return y;
}

// Y_ab of a merged product ym is its block (columns of M_a) x (columns of
// M_b), the columns of M_1 being c1 = mw >> 16: pass 0 takes (0,0) and (1,1)
// (the I/Z children), pass 1 (0,1) and (1,0) (the X children).
__device__ __forceinline__ void rsplit(const d4 &ym, unsigned mw, int pass, int a16, int g4, d4 &ya, d4 &yb) {
  const unsigned c1 = mw >> 16;
  const unsigned ra = (c1 >> a16) & 1u;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const unsigned cb = ((c1 >> kdx(g4, v)) & 1u) ^ (unsigned)pass;
    ya[v] = (ra == 0u && cb == 0u) ? ym[v] : 0.0;
    yb[v] = (ra == 1u && cb == 1u) ? ym[v] : 0.0;
  }
}

// The children of one step pass from (ya, yb) = (Y00, Y11) (pass 0) or
// (Y01, Y10) (pass 1): I/X children take ya + yb, Z children ya - yb; each is
// stored, deferred, contracted with its inline term ends, or continues the
// chain (cont).
template <bool TERMS, bool IZ>
__device__ __forceinline__ void rstep_children(const ProbDev &P, const InstDev &inst, const StepDev &sd, int pass,
                                               const d4 &ya, const d4 &yb, int lane, double2 *__restrict__ spill,
                                               double2 *__restrict__ leafs, double2 *__restrict__ tv, double &acc,
                                               REnv &cont) {
  for (int c = 0; c < sd.nchild; ++c) {
    const ChildDev cd = uload(P.children + sd.child0 + c);
    if (!IZ && (cd.kind == CK_X) != (pass == 1)) continue;
    // ya +- yb as one FMA with a wave-uniform sign: the product is exact, so
    // each entry is the sum or difference rounded once, as an add would be
    const double sgn = cd.kind == CK_Z ? -1.0 : 1.0;
    REnv ce;
#pragma unroll
    for (int v = 0; v < 4; ++v) ce.v[v] = __builtin_fma(sgn, yb[v], ya[v]);
    if (cd.store >= 0) renv_store_sym(rslot(spill, cd.store), lane, ce);
    if (cd.defer >= 0) renv_store_sym(rslot(leafs, inst.leaf_off + cd.defer), lane, ce);
    const int nterm_inline = cd.defer >= 0 ? 0 : cd.nterm;
#pragma unroll(IZ ? GV_IZ_TERM_UNROLL : 2)
    for (int t = 0; t < nterm_inline; ++t) {
      const TermEnd te = uload(P.term_list + cd.term0 + t);
      REnv r;
      if (te.rslot >= 0) renv_load_sym(rslot(spill, te.rslot), lane, r);
      else renv_unit(lane, r);  // term ends at site n-1
      double sr = 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) sr += ce.v[v] * r.v[v];
      if (TERMS) {
        sr = wave_sum(sr);
        if (lane == 0) tv[te.term] = make_double2(sr, 0.0);
      } else {
        acc += te.cre * sr;  // Re(c <P>), <P> real
      }
    }
    if (cd.action == -2) cont = ce;
    else if (cd.action >= 0) renv_store_sym(rslot(spill, cd.action), lane, ce);
  }
}

// A child environment's packed upper triangle through the wave's 16 x 17 LDS
// tile: lane l gets packed elements 2l, 2l+1 (lo) and, lanes 0-3, 128 + 2l,
// 129 + 2l (hi), at the tile positions ptab holds (4 per lane, from
// pack_tile_offsets).  rpacked_store writes them with two contiguous stores.
__device__ __forceinline__ void renv_pack_tile(const REnv &e, int lane, double (*tl)[17], const int *ptab, gv_d2 &lo,
                                               gv_d2 &hi) {
  const int a = lane & 15, g = lane >> 4;
  const int4 o = *reinterpret_cast<const int4 *>(ptab + 4 * lane);
#pragma unroll
  for (int t = 0; t < 4; ++t) tl[a][kdx(g, t)] = e.v[t];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const double *tf = &tl[0][0];
  lo = gv_d2{tf[o.x], tf[o.y]};
  hi = gv_d2{tf[o.z], tf[o.w]};
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The table of tile positions, filled by the workgroup's first wave before
// any wave returns (a workgroup barrier).
__device__ __forceinline__ void pack_table(int *ptab) {
  if (threadIdx.x < 64) {
    const PackTile pt = pack_tile_offsets(threadIdx.x);
#pragma unroll
    for (int q = 0; q < 4; ++q) ptab[4 * threadIdx.x + q] = pt.o[q];
  }
  __syncthreads();
}

__device__ __forceinline__ void rpacked_store(double *p, int lane, const gv_d2 &lo, const gv_d2 &hi) {
  gv_d2 *q = reinterpret_cast<gv_d2 *>(p);
  q[lane] = lo;
  if (lane < 4) q[64 + lane] = hi;
}

// The children of an I/Z step (at most two: I and Z) from records the chain
// loop fetched one step ahead, so no descriptor load waits here.
template <bool TERMS>
__device__ __forceinline__ void rstep_children_iz(const ProbDev &P, const InstDev &inst, int nchild,
                                                  const ChildDev (&cds)[2], const d4 &ya, const d4 &yb, int lane,
                                                  double2 *__restrict__ spill, double2 *__restrict__ leafs,
                                                  double2 *__restrict__ tv, double &acc, REnv &cont,
                                                  double (*tl)[17], const int *ptab) {
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    if (c >= nchild) break;
    const ChildDev &cd = cds[c];
    const double sgn = cd.kind == CK_Z ? -1.0 : 1.0;
    REnv ce;
#pragma unroll
    for (int v = 0; v < 4; ++v) ce.v[v] = __builtin_fma(sgn, yb[v], ya[v]);
    if (cd.store >= 0 || cd.defer >= 0 || cd.action >= 0) {
      // the packed triangle through the wave's tile, so each target takes
      // two contiguous stores instead of four masked gathers of the square
      gv_d2 lo, hi;
      renv_pack_tile(ce, lane, tl, ptab, lo, hi);
      if (cd.store >= 0) rpacked_store(rslot(spill, cd.store), lane, lo, hi);
      if (cd.defer >= 0) rpacked_store(rslot(leafs, inst.leaf_off + cd.defer), lane, lo, hi);
      if (cd.action >= 0) rpacked_store(rslot(spill, cd.action), lane, lo, hi);
    }
    const int nterm_inline = cd.defer >= 0 ? 0 : cd.nterm;
#pragma unroll(GV_IZ_TERM_UNROLL)
    for (int t = 0; t < nterm_inline; ++t) {
      const TermEnd te = uload(P.term_list + cd.term0 + t);
      REnv r;
      if (te.rslot >= 0) renv_load_sym(rslot(spill, te.rslot), lane, r);
      else renv_unit(lane, r);  // term ends at site n-1
      double sr = 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) sr += ce.v[v] * r.v[v];
      if (TERMS) {
        sr = wave_sum(sr);
        if (lane == 0) tv[te.term] = make_double2(sr, 0.0);
      } else {
        acc += te.cre * sr;
      }
    }
    if (cd.action == -2) cont = ce;
  }
}

// One generation of prefix/suffix chains in real arithmetic; the same work
// items and slot numbering as chains_body in kernels_energy.hip.  IZ: every
// step of the forest has I/Z children only (Hamiltonians without X words), so
// the site fragment and the X_i products die once Y00 and Y11 are formed;
// with fewer live registers the kernel runs more waves per SIMD.  The
// arithmetic per instance is the same in both forms.
template <bool TERMS, bool IZ>
__device__ __forceinline__ void rchain_item(const int2 it, const int lane, const InstDev *__restrict__ insts,
                                            const ProbDev *__restrict__ probs, const double *__restrict__ rarena,
                                            const double *__restrict__ rarenaT, const unsigned *__restrict__ smask,
                                            double2 *__restrict__ leafs, double2 *__restrict__ spills,
                                            double2 *__restrict__ terms, double *__restrict__ parts) {
  const InstDev inst = uload(insts + it.x);
  const ProbDev P = uload(probs + inst.prob);
  const ChainDev ch = uload(P.chains + it.y);
  const double *sites = (ch.dir != 0 ? rarenaT : rarena) + inst.site_off;
  // this instance's zero-structure words for the copy the chain reads
  const unsigned *sm = smask && !IZ ? smask + (inst.site_off / GV_SLOT) * 2 + (ch.dir != 0 ? 1 : 0) : nullptr;
  const int a16 = lane & 15, g4 = lane >> 4;
  double2 *spill = spills + inst.spill_off;
  double2 *tv = terms + inst.term_off;

  REnv e;
  if (ch.src >= 0) renv_load_sym(rslot(spill, ch.src), lane, e);
  else renv_unit(lane, e);
  double acc = 0.0;

  StepDev sd0 = uload(P.steps + ch.step0);
  RFrag fa, fb;
  rfrag_load(sites + (size_t)sd0.site * GV_SLOT, lane, fa);
  // one chain step on fragment f; prefetches the next step's fragment into nf.
  // The prefetch is unconditional (the last step reloads its own site): with
  // the same number of loads in flight on every path, the wait for f leaves
  // the prefetch outstanding instead of draining it (vmcnt counts in order).
  auto run_step = [&](int st, const StepDev &sd, StepDev &nsd, const RFrag &f, RFrag &nf) {
    nsd = uload(P.steps + ch.step0 + min(st + 1, ch.nsteps - 1));
    rfrag_load(sites + (size_t)nsd.site * GV_SLOT, lane, nf);
    const unsigned mw = sm ? uload(sm + 2 * sd.site) : 0xffu;
    const bool merged = (mw & GV_SM_DISJOINT) != 0;
    const bool rowsplit = !merged && (mw & 0xffu) == 0xc3u;  // M_0 rows 0-7, M_1 rows 8-15
    auto split = [&](const d4 &ym, int pass, d4 &ya, d4 &yb) { rsplit(ym, mw, pass, a16, g4, ya, yb); };
    d4 x0, x1, ym;
    if (IZ) {
      // the I/Z form keeps only (ya, yb) live past this point (the masked
      // I/Z steps run in rchain_item_merged)
      const d4 p0 = rprod_x(e, f, 0), p1 = rprod_x(e, f, 1);
      x0 = rprod_y(p0, f, 0);
      x1 = rprod_y(p1, f, 1);
    } else if (merged) {
      ym = rprod_merged(e, f);
    } else if (rowsplit) {
      x0 = rprod_x_g<3u>(e, f, 0);
      x1 = rprod_x_g<12u>(e, f, 1);
    } else {
      x0 = rprod_x(e, f, 0);
      x1 = rprod_x(e, f, 1);
    }
    REnv cont;
    // pass 0: I/Z children from Y00, Y11; pass 1: X children from Y01, Y10
#pragma nounroll
    for (int pass = 0; pass < (IZ ? 1 : 2); ++pass) {
      if (!IZ && !(sd.need & (1 << pass))) continue;
      d4 ya, yb;
      if (IZ) {  // formed above: x0 = Y00, x1 = Y11
        ya = x0;
        yb = x1;
      } else if (merged) {
        split(ym, pass, ya, yb);
      } else if (rowsplit) {
        ya = rprod_y_g<3u>(pass ? x1 : x0, f, 0);
        yb = rprod_y_g<12u>(pass ? x0 : x1, f, 1);
      } else {
        ya = rprod_y(pass ? x1 : x0, f, 0);
        yb = rprod_y(pass ? x0 : x1, f, 1);
      }
      rstep_children<TERMS, IZ>(P, inst, sd, pass, ya, yb, lane, spill, leafs, tv, acc, cont);
    }
    e = cont;
  };
  for (int st = 0; st < ch.nsteps; st += 2) {
    StepDev sd1;
    run_step(st, sd0, sd1, fa, fb);
    if (st + 1 >= ch.nsteps) break;
    run_step(st + 1, sd1, sd0, fb, fa);
  }
  if (!TERMS) {
    acc = wave_sum(acc);
    if (lane == 0) parts[inst.part_off + it.y] = acc;
  }
}

template <bool TERMS, bool IZ>
__device__ __forceinline__ void rchains_body(const int2 *__restrict__ items, int nitems,
                                             const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
                                             const double *__restrict__ rarena, const double *__restrict__ rarenaT,
                                             const unsigned *__restrict__ smask, double2 *__restrict__ leafs,
                                             double2 *__restrict__ spills, double2 *__restrict__ terms,
                                             double *__restrict__ parts) {
  const int w = xcd_item();
  if (w >= nitems) return;
  rchain_item<TERMS, IZ>(uload(items + w), threadIdx.x & 63, insts, probs, rarena, rarenaT, smask, leafs, spills,
                         terms, parts);
}

// Merged site fragments.  Where the two components of a site have disjoint
// supports (both zero-structure cases above), F = M_0 + M_1 holds them both:
// the merged products read F, and so does the row split (rows 0-7 of F are
// those of M_0, rows 8-15 those of M_1).  k_site_masks writes F for every
// slot and copy, row-major 16x16 (GV_FSLOT doubles), so an I/Z chain step
// loads 2 KiB instead of 4 and keeps 4 doubles per lane per fragment instead
// of 8; the few steps whose site has no such structure (bonds below 16 at
// the chain ends) load the full slot when they run.
__device__ __forceinline__ void ffrag_load(const double *slot, int lane, double (&f)[4]) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) f[t] = slot[kdx(g, t) * 16 + j];
}

__device__ __forceinline__ void ffrag_load_perm(const double *slot, int lane, double (&f)[4]) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int t = 0; t < 4; ++t) f[t] = slot[par_perm(kdx(g, t)) * 16 + j];
}

template <bool TERMS>
__device__ __forceinline__ void rchain_item_merged(const int2 it, const int lane, const InstDev *__restrict__ insts,
                                                   const ProbDev *__restrict__ probs,
                                                   const double *__restrict__ rarena, const double *__restrict__ rarenaT,
                                                   const double *__restrict__ farena,
                                                   const double *__restrict__ farenaT,
                                                   const unsigned *__restrict__ smask, double2 *__restrict__ leafs,
                                                   double2 *__restrict__ spills, double2 *__restrict__ terms,
                                                   double *__restrict__ parts, double (*tl)[17], const int *ptab) {
  const InstDev inst = uload(insts + it.x);
  const ProbDev P = uload(probs + inst.prob);
  const ChainDev ch = uload(P.chains + it.y);
  const long long slot0 = inst.site_off / GV_SLOT;
  const double *sites = (ch.dir != 0 ? rarenaT : rarena) + inst.site_off;
  const double *fsites = (ch.dir != 0 ? farenaT : farena) + slot0 * GV_FSLOT;
  const unsigned *sm = smask + slot0 * 2 + (ch.dir != 0 ? 1 : 0);
  const int a16 = lane & 15, g4 = lane >> 4;
  double2 *spill = spills + inst.spill_off;
  double2 *tv = terms + inst.term_off;

  REnv e;
  if (ch.src >= 0) renv_load_sym(rslot(spill, ch.src), lane, e);
  else renv_unit(lane, e);
  double acc = 0.0;

  // Scalar pipeline: the descriptors a step needs (its StepDev, mask word and
  // child records) are fetched one step ahead, beside the fragment prefetch,
  // so the step's only scalar waits are for loads issued a whole step before.
  auto children_of = [&](const StepDev &d, ChildDev (&c)[2]) {
    // unconditional (the batch pads the child table with one record), so no
    // branch or copy waits for them before they are used
    c[0] = uload(P.children + d.child0);
    c[1] = uload(P.children + d.child0 + 1);
  };
  StepDev sd = uload(P.steps + ch.step0);
  StepDev nsd = uload(P.steps + ch.step0 + min(1, ch.nsteps - 1));
  unsigned mw = uload(sm + 2 * sd.site);
  ChildDev cds[2];
  children_of(sd, cds);
  double fa[4], fb[4];
  ffrag_load(fsites + (size_t)sd.site * GV_FSLOT, lane, fa);
  auto run_step = [&](int st, const double(&f)[4], double(&nf)[4]) {
    // next step: fragment (unconditional, as in rchain_item), mask, children;
    // the step after: its descriptor
    ffrag_load(fsites + (size_t)nsd.site * GV_FSLOT, lane, nf);
    const unsigned nmw = uload(sm + 2 * nsd.site);
    ChildDev ncds[2];
    children_of(nsd, ncds);
    const StepDev nnsd = uload(P.steps + ch.step0 + min(st + 2, ch.nsteps - 1));
    d4 y0, y1;
    if (mw & GV_SM_DISJOINT) {
      d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) x = mfma(e.v[t], f[t], x);
      d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) y = mfma(x[t], f[t], y);
      if ((mw >> 16) == 0xff00u) {
        // the bond-16 bulk: M_0 in columns 0-7, M_1 in 8-15, so Y00 is
        // lanes a < 8 in registers 0-1 and Y11 lanes a >= 8 in registers
        // 2-3 (the same selects as rsplit, with the bit tests folded)
        const bool lo = a16 < 8;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          y0[v] = (v < 2 && lo) ? y[v] : 0.0;
          y1[v] = (v >= 2 && !lo) ? y[v] : 0.0;
        }
      } else {
        rsplit(y, mw, 0, a16, g4, y0, y1);
      }
    } else if ((mw & 0xffu) == 0xc3u) {  // M_0 rows 0-7 (f[0], f[1]), M_1 rows 8-15 (f[2], f[3])
      d4 p0 = {0.0, 0.0, 0.0, 0.0}, p1 = {0.0, 0.0, 0.0, 0.0};
      p0 = mfma(e.v[0], f[0], p0);
      p1 = mfma(e.v[2], f[2], p1);
      p0 = mfma(e.v[1], f[1], p0);
      p1 = mfma(e.v[3], f[3], p1);
      d4 q0 = {0.0, 0.0, 0.0, 0.0}, q1 = {0.0, 0.0, 0.0, 0.0};
      q0 = mfma(p0[0], f[0], q0);
      q1 = mfma(p1[2], f[2], q1);
      q0 = mfma(p0[1], f[1], q0);
      q1 = mfma(p1[3], f[3], q1);
      y0 = q0;
      y1 = q1;
    } else {  // no structure: the components from the full slot, one at a time
      const double *slot = sites + (size_t)sd.site * GV_SLOT;
      double m[4];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int t = 0; t < 4; ++t) m[t] = slot[(kdx(g4, t) * 2 + i) * 16 + a16];
        d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 4; ++t) x = mfma(e.v[t], m[t], x);
        d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 4; ++t) y = mfma(x[t], m[t], y);
        if (i == 0) y0 = y;
        else y1 = y;
      }
    }
    REnv cont;
    rstep_children_iz<TERMS>(P, inst, sd.nchild, cds, y0, y1, lane, spill, leafs, tv, acc, cont, tl, ptab);
    e = cont;
    sd = nsd;
    nsd = nnsd;
    mw = nmw;
    cds[0] = ncds[0];
    cds[1] = ncds[1];
  };
  for (int st = 0; st < ch.nsteps; st += 2) {
    run_step(st, fa, fb);
    if (st + 1 >= ch.nsteps) break;
    run_step(st + 1, fb, fa);
  }
  if (!TERMS) {
    acc = wave_sum(acc);
    if (lane == 0) parts[inst.part_off + it.y] = acc;
  }
}


// Chains with at most a quarter of their steps on sites without structure run
// over the merged fragments (those steps load their full slots when they run);
// the rest, e.g. every chain of a state prepared with a decomposition at each
// gate, run the full-slot loop with its prefetch.  The flag sits at the
// chain's partial-energy index.
template <bool TERMS>
__device__ __forceinline__ void rchains_merged_body(const int2 *__restrict__ items, int nitems,
                                                    const InstDev *__restrict__ insts,
                                                    const ProbDev *__restrict__ probs,
                                                    const double *__restrict__ rarena,
                                                    const double *__restrict__ rarenaT,
                                                    const double *__restrict__ farena,
                                                    const double *__restrict__ farenaT,
                                                    const unsigned *__restrict__ smask,
                                                    const int *__restrict__ cflags, double2 *__restrict__ leafs,
                                                    double2 *__restrict__ spills, double2 *__restrict__ terms,
                                                    double *__restrict__ parts) {
  // per-wave tiles for the packed child stores and the lanes' tile positions
  __shared__ double tiles[4][16][17];
  __shared__ int ptab[64 * 4];
  pack_table(ptab);
  const int w = xcd_item();
  if (w >= nitems) return;
  const int2 it = uload(items + w);
  const InstDev inst = uload(insts + it.x);
  if (uload(cflags + inst.part_off + it.y))
    rchain_item_merged<TERMS>(it, threadIdx.x & 63, insts, probs, rarena, rarenaT, farena, farenaT, smask, leafs,
                              spills, terms, parts, tiles[wave_id()], ptab);
  else
    rchain_item<TERMS, true>(it, threadIdx.x & 63, insts, probs, rarena, rarenaT, nullptr, leafs, spills, terms,
                             parts);
}

// Deferred term ends contract two symmetric environments, both stored as
// packed upper triangles (136 doubles): <L, R> = sum over the triangle of
// w L R with w = 1 on the diagonal and 2 off it.  Lane l reads packed
// elements 2l, 2l+1 (one 16-byte load) and lanes 0-3 also 128 + 2l, 129 + 2l,
// so each environment is two contiguous loads per lane instead of four
// gathers of the mirrored square.  The weights are folded into L once.
__device__ __forceinline__ bool sym_diag(int p) {
  // diagonal positions of the packed triangle: a (33 - a) / 2, a = 0..15
  return p == 0 || p == 16 || p == 31 || p == 45 || p == 58 || p == 70 || p == 81 || p == 91 || p == 100 ||
         p == 108 || p == 115 || p == 121 || p == 126 || p == 130 || p == 133 || p == 135;
}

struct RPacked {
  double2 lo, hi;  // packed elements 2l, 2l+1 and (lanes 0-3) 128 + 2l, 129 + 2l
};

__device__ __forceinline__ void rpacked_load(const double *p, int lane, RPacked &e) {
  const double2 *q = reinterpret_cast<const double2 *>(p);
  e.lo = q[lane];
  e.hi = lane < 4 ? q[64 + lane] : make_double2(0.0, 0.0);
}

__device__ __forceinline__ void rpacked_unit(int lane, RPacked &e) {
  e.lo = make_double2(lane == 0 ? 1.0 : 0.0, 0.0);
  e.hi = make_double2(0.0, 0.0);
}

__device__ __forceinline__ void rpacked_weight(int lane, RPacked &e) {
  if (!sym_diag(2 * lane)) e.lo.x *= 2.0;
  if (!sym_diag(2 * lane + 1)) e.lo.y *= 2.0;
  if (!sym_diag(128 + 2 * lane)) e.hi.x *= 2.0;
  if (!sym_diag(129 + 2 * lane)) e.hi.y *= 2.0;
}

template <bool TERMS>
__device__ __forceinline__ void rleaf_item(const LeafItem &li, const int lane, const double2 *__restrict__ leafs,
                                           const double2 *__restrict__ spills, double2 *__restrict__ terms,
                                           double *__restrict__ parts) {
  const double2 *spill = spills + li.spill;
  RPacked e;
  rpacked_load((const double *)(leafs + li.env), lane, e);
  rpacked_weight(lane, e);
  double acc = 0.0;
#pragma unroll 4
  for (int t = 0; t < li.nterm; ++t) {
    const TermEnd te = uload(li.tl + t);
    RPacked r;
    if (te.rslot >= 0) rpacked_load(rslot(spill, te.rslot), lane, r);
    else rpacked_unit(lane, r);
    double sr = e.lo.x * r.lo.x;
    sr = __builtin_fma(e.lo.y, r.lo.y, sr);
    sr = __builtin_fma(e.hi.x, r.hi.x, sr);
    sr = __builtin_fma(e.hi.y, r.hi.y, sr);
    if (TERMS) {
      sr = wave_sum(sr);
      if (lane == 0) terms[li.tv + te.term] = make_double2(sr, 0.0);
    } else {
      acc = __builtin_fma(te.cre, sr, acc);
    }
  }
  if (!TERMS) {
    acc = wave_sum(acc);
    if (lane == 0) parts[li.part] = acc;
  }
}

template <bool TERMS>
__device__ __forceinline__ void rleaf_body(const LeafItem *__restrict__ items, int nitems,
                                           const double2 *__restrict__ leafs, const double2 *__restrict__ spills,
                                           double2 *__restrict__ terms, double *__restrict__ parts) {
  const int w = xcd_item();
  if (w >= nitems) return;
  rleaf_item<TERMS>(uload(items + w), threadIdx.x & 63, leafs, spills, terms, parts);
}

__device__ __forceinline__ int sidx(int l, int i, int r) { return (l * 2 + i) * 16 + r; }

// Input environment of the step after `step` in its chain (the continuing
// child), with the products and combination of k_rchains_gfwd: bit for bit
// the environment the forward pass had.  smc: the mask words of the chain's
// copy when the forward ran it over the zero structure (a suffix chain's
// parity-split steps contracted P E P^T with the rows of M_i in par_perm
// order: here the full products in that order, which add the same nonzero
// terms in the same order), else null; tl: the wave's LDS tile.
__device__ __forceinline__ void radvance(const ProbDev &P, const double *sites, int step, int lane, REnv &e,
                                         const unsigned *smc, double (*tl)[17]) {
  const StepDev sd = uload(P.steps + step);
  RFrag f;
  if (smc && (uload(smc + 2 * sd.site) & GV_SM_PARITY)) {
    rfrag_load_perm(sites + (size_t)sd.site * GV_SLOT, lane, f);
    renv_symmetrize<false>(e, lane, tl, nullptr, nullptr, true);  // e is symmetric: P E P^T
  } else {
    rfrag_load(sites + (size_t)sd.site * GV_SLOT, lane, f);
  }
  int kind = CK_I;
  for (int c = 0; c < sd.nchild; ++c) {
    const ChildDev cd = uload(P.children + sd.child0 + c);
    if (cd.action == -2) kind = cd.kind;
  }
  const d4 x0 = rprod_x(e, f, 0), x1 = rprod_x(e, f, 1);
  const bool xk = kind == CK_X;
  const d4 ya = rprod_y(xk ? x1 : x0, f, 0);
  const d4 yb = rprod_y(xk ? x0 : x1, f, 1);
#pragma unroll
  for (int v = 0; v < 4; ++v) e.v[v] = __builtin_fma(kind == CK_Z ? -1.0 : 1.0, yb[v], ya[v]);  // ya +- yb
}

// Stores one step's site adjoint (m0, m1: Mbar_0, Mbar_1 in the D layout),
// pulled back through the CNOT halves that end its site's program, in the
// compact layout of tail_pull (device_common.h).  Non-temporal: the site
// reduction reads it once, after the whole backward pass.
__device__ __forceinline__ void radj_store(double *slot, int tail, int a16, int g4, const d4 &m0, const d4 &m1) {
  if ((tail & 3) == 0) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      __builtin_nontemporal_store(m0[v], slot + sidx(kdx(g4, v), 0, a16));
      __builtin_nontemporal_store(m1[v], slot + sidx(kdx(g4, v), 1, a16));
    }
    return;
  }
  double y[4];
  const TailMap tm = tail_pull(tail, a16, g4, m0, m1, y);
  if (tm.keep) {
#pragma unroll
    for (int v = 0; v < 4; ++v) __builtin_nontemporal_store(y[v], slot + tail_index(tm, g4, v));
  }
}

}  // namespace

// ------------------------------------------------ reverse gradient, real
// The bond-doubling state of a real tape (RY, CX, ...) is real, and so are
// the forest environments of I/X/Z words and their adjoints (the adjoint of a
// real symmetric environment is real symmetric: conj and transpose drop out
// of the complex formulas).  Same slots as kernels_grad.hip; a step's site
// adjoint, pulled back through the CNOT halves that end its site's program
// (radj_store), uses the first 1 to 4 KiB of its 8 KiB step slot.  Hand-off spills, suffix stores,
// prefix leaves and adjoint spills are packed symmetric; the checkpointed
// step environments stay full, so a recomputed step sees exactly the
// environment the forward pass used and every checkpoint stride gives the
// same bits.

// Forward: energy chains keeping the checkpointed step environments and the
// prefix leaves the backward pass reads (k_chains_gfwd in real arithmetic).
// It runs the energy launches (shallow prefix depths beside the suffix
// forest): deferred term ends go to leaf slots for k_rleaf_energy, exactly as
// in the energy pass.  The backward pass reads only what this pass stores, so
// it keeps its own generation order.
// IZ: every step has I/Z children only (one pass, as in rchains_body).
template <bool IZ>
__device__ __forceinline__ void rgfwd_body(const int2 *__restrict__ items, int nitems,
                                           const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
                                           const double *__restrict__ rgarena, const double *__restrict__ rgarenaT,
                                           double2 *__restrict__ spills, double *__restrict__ parts,
                                           double2 *__restrict__ estore, double2 *__restrict__ gleafs,
                                           double2 *__restrict__ leafs, double (*tr)[16][17], const int *ptab) {
  const int lane = threadIdx.x & 63;
  double (*tl)[17] = tr[wave_id()];
  const int nb = gridDim.x, b = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = b & 7;
  const int logical = xcd * q8 + min(xcd, r8) + (b >> 3);
  const int w = logical * (int)(blockDim.x >> 6) + wave_id();
  if (w >= nitems) return;
  const int2 it = uload(items + w);
  const InstDev inst = uload(insts + it.x);
  const ProbDev P = uload(probs + inst.prob);
  const ChainDev ch = uload(P.chains + it.y);
  const double *sites = (ch.dir != 0 ? rgarenaT : rgarena) + inst.site_off;
  double2 *spill = spills + inst.spill_off;
  double2 *es = estore + inst.ck_off * GV_ENV;
  double2 *gl = gleafs + inst.gleaf_off * GV_ENV;
  REnv e;
  if (ch.src >= 0) renv_load_sym(rslot(spill, ch.src), lane, e);
  else renv_unit(lane, e);
  double acc = 0.0;
  StepDev sd0 = uload(P.steps + ch.step0);
  RFrag fa, fb;
  rfrag_load(sites + (size_t)sd0.site * GV_SLOT, lane, fa);
  auto run_step = [&](int st, const StepDev &sd, StepDev &nsd, const RFrag &f, RFrag &nf) {
    // the step's input, symmetrised, is what it contracts and checkpoints
    // (store first: the prefetch below is then the youngest memory operation
    // when this step waits for f)
    const int cs = uload(P.ck_slot + ch.step0 + st);
    renv_symmetrize<true>(e, lane, tl, ptab, cs >= 0 ? rslot(es, cs) : nullptr);
    nsd = uload(P.steps + ch.step0 + min(st + 1, ch.nsteps - 1));  // unconditional prefetch, as above
    rfrag_load(sites + (size_t)nsd.site * GV_SLOT, lane, nf);
    const d4 x0 = rprod_x(e, f, 0), x1 = rprod_x(e, f, 1);
    REnv cont;
#pragma nounroll
    for (int pass = 0; pass < (IZ ? 1 : 2); ++pass) {
      if (!IZ && !(sd.need & (1 << pass))) continue;
      const d4 ya = rprod_y(pass ? x1 : x0, f, 0);
      const d4 yb = rprod_y(pass ? x0 : x1, f, 1);
      for (int c = 0; c < sd.nchild; ++c) {
        const ChildDev cd = uload(P.children + sd.child0 + c);
        if (!IZ && (cd.kind == CK_X) != (pass == 1)) continue;
        REnv ce;
#pragma unroll
        for (int v = 0; v < 4; ++v) ce.v[v] = __builtin_fma(cd.kind == CK_Z ? -1.0 : 1.0, yb[v], ya[v]);  // ya +- yb
        if (cd.store >= 0 || cd.gleaf >= 0 || cd.defer >= 0 || cd.action >= 0) {
          // the packed triangle through the wave's tile: two contiguous
          // stores per target (as in rstep_children_iz)
          gv_d2 lo, hi;
          renv_pack_tile(ce, lane, tl, ptab, lo, hi);
          if (cd.store >= 0) rpacked_store(rslot(spill, cd.store), lane, lo, hi);
          if (cd.gleaf >= 0) rpacked_store(rslot(gl, cd.gleaf), lane, lo, hi);
          if (cd.defer >= 0) rpacked_store(rslot(leafs, inst.leaf_off + cd.defer), lane, lo, hi);
          if (cd.action >= 0) rpacked_store(rslot(spill, cd.action), lane, lo, hi);
        }
        const int nterm_inline = cd.defer >= 0 ? 0 : cd.nterm;
        for (int t = 0; t < nterm_inline; ++t) {
          const TermEnd te = uload(P.term_list + cd.term0 + t);
          REnv r;
          if (te.rslot >= 0) renv_load_sym(rslot(spill, te.rslot), lane, r);
          else renv_unit(lane, r);
          double sr = 0.0;
#pragma unroll
          for (int v = 0; v < 4; ++v) sr += ce.v[v] * r.v[v];
          acc += te.cre * sr;  // Re(c_k) <P_k>
        }
        if (cd.action == -2) cont = ce;
      }
    }
    e = cont;
  };
  for (int st = 0; st < ch.nsteps; st += 2) {
    StepDev sd1;
    run_step(st, sd0, sd1, fa, fb);
    if (st + 1 >= ch.nsteps) break;
    run_step(st + 1, sd1, sd0, fb, fa);
  }
  acc = wave_sum(acc);
  if (lane == 0) parts[inst.part_off + it.y] = acc;
}

// The forward of an I/Z chain over the zero structure of the bond-doubling
// state (its chain flag set by k_chain_flags in gradient mode): a prefix
// chain's steps on sites with disjoint columns take the merged products over
// F = M_0 + M_1, a suffix chain's steps on parity-split sites the row-split
// products over P F and P E P^T (the rows in par_perm order: M_0 in rows 0-7,
// M_1 in rows 8-15), each in half the MFMAs; other steps take the full
// products of rgfwd_body.  The merged products give the bits of the full
// ones; the permuted ones those of the full products in the permuted order,
// which radvance repeats.  Checkpoints, stores and term ends as rgfwd_body.
__device__ __forceinline__ void rgfwd_struct_item(const int2 it, const int lane, const InstDev *__restrict__ insts,
                                                  const ProbDev *__restrict__ probs,
                                                  const double *__restrict__ rgarena,
                                                  const double *__restrict__ rgarenaT,
                                                  const double *__restrict__ farena,
                                                  const double *__restrict__ farenaT,
                                                  const unsigned *__restrict__ smask, double2 *__restrict__ spills,
                                                  double *__restrict__ parts, double2 *__restrict__ estore,
                                                  double2 *__restrict__ gleafs, double2 *__restrict__ leafs,
                                                  double (*tl)[17], const int *ptab) {
  const InstDev inst = uload(insts + it.x);
  const ProbDev P = uload(probs + inst.prob);
  const ChainDev ch = uload(P.chains + it.y);
  const bool dir1 = ch.dir != 0;
  const long long slot0 = inst.site_off / GV_SLOT;
  const double *sites = (dir1 ? rgarenaT : rgarena) + inst.site_off;
  const double *fsites = (dir1 ? farenaT : farena) + slot0 * GV_FSLOT;
  const unsigned *sm = smask + slot0 * 2 + (dir1 ? 1 : 0);
  const int a16 = lane & 15, g4 = lane >> 4;
  double2 *spill = spills + inst.spill_off;
  double2 *es = estore + inst.ck_off * GV_ENV;
  double2 *gl = gleafs + inst.gleaf_off * GV_ENV;
  REnv e;
  if (ch.src >= 0) renv_load_sym(rslot(spill, ch.src), lane, e);
  else renv_unit(lane, e);
  double acc = 0.0;
  // a suffix chain's fragments are loaded in par_perm row order
  auto fload = [&](int site, double(&f)[4]) {
    if (dir1) ffrag_load_perm(fsites + (size_t)site * GV_FSLOT, lane, f);
    else ffrag_load(fsites + (size_t)site * GV_FSLOT, lane, f);
  };
  StepDev sd0 = uload(P.steps + ch.step0);
  double fa[4], fb[4];
  fload(sd0.site, fa);
  auto run_step = [&](int st, const StepDev &sd, StepDev &nsd, const double(&f)[4], double(&nf)[4]) {
    const unsigned mw = uload(sm + 2 * sd.site);
    const bool merged = !dir1 && (mw & GV_SM_DISJOINT) != 0u;
    const bool perm = dir1 && (mw & GV_SM_PARITY) != 0u;
    const int cs = uload(P.ck_slot + ch.step0 + st);
    renv_symmetrize<true>(e, lane, tl, ptab, cs >= 0 ? rslot(es, cs) : nullptr, perm);
    nsd = uload(P.steps + ch.step0 + min(st + 1, ch.nsteps - 1));  // unconditional prefetch
    fload(nsd.site, nf);
    d4 y0, y1;
    if (merged) {
      d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) x = mfma(e.v[t], f[t], x);
      d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) y = mfma(x[t], f[t], y);
      rsplit(y, mw, 0, a16, g4, y0, y1);
    } else if (perm) {  // P M_0 in rows 0-7 (f[0], f[1]), P M_1 in rows 8-15 (f[2], f[3])
      d4 p0 = {0.0, 0.0, 0.0, 0.0}, p1 = {0.0, 0.0, 0.0, 0.0};
      p0 = mfma(e.v[0], f[0], p0);
      p1 = mfma(e.v[2], f[2], p1);
      p0 = mfma(e.v[1], f[1], p0);
      p1 = mfma(e.v[3], f[3], p1);
      d4 q0 = {0.0, 0.0, 0.0, 0.0}, q1 = {0.0, 0.0, 0.0, 0.0};
      q0 = mfma(p0[0], f[0], q0);
      q1 = mfma(p1[2], f[2], q1);
      q0 = mfma(p0[1], f[1], q0);
      q1 = mfma(p1[3], f[3], q1);
      y0 = q0;
      y1 = q1;
    } else {  // no structure: the components from the full slot, one at a time
      const double *slot = sites + (size_t)sd.site * GV_SLOT;
      double m[4];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int t = 0; t < 4; ++t) m[t] = slot[(kdx(g4, t) * 2 + i) * 16 + a16];
        d4 x = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 4; ++t) x = mfma(e.v[t], m[t], x);
        d4 y = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < 4; ++t) y = mfma(x[t], m[t], y);
        if (i == 0) y0 = y;
        else y1 = y;
      }
    }
    REnv cont;
    for (int c = 0; c < sd.nchild; ++c) {
      const ChildDev cd = uload(P.children + sd.child0 + c);
      REnv ce;
#pragma unroll
      for (int v = 0; v < 4; ++v) ce.v[v] = __builtin_fma(cd.kind == CK_Z ? -1.0 : 1.0, y1[v], y0[v]);  // y0 +- y1
      if (cd.store >= 0 || cd.gleaf >= 0 || cd.defer >= 0 || cd.action >= 0) {
        gv_d2 lo, hi;
        renv_pack_tile(ce, lane, tl, ptab, lo, hi);
        if (cd.store >= 0) rpacked_store(rslot(spill, cd.store), lane, lo, hi);
        if (cd.gleaf >= 0) rpacked_store(rslot(gl, cd.gleaf), lane, lo, hi);
        if (cd.defer >= 0) rpacked_store(rslot(leafs, inst.leaf_off + cd.defer), lane, lo, hi);
        if (cd.action >= 0) rpacked_store(rslot(spill, cd.action), lane, lo, hi);
      }
      const int nterm_inline = cd.defer >= 0 ? 0 : cd.nterm;
      for (int t = 0; t < nterm_inline; ++t) {
        const TermEnd te = uload(P.term_list + cd.term0 + t);
        REnv r;
        if (te.rslot >= 0) renv_load_sym(rslot(spill, te.rslot), lane, r);
        else renv_unit(lane, r);
        double sr = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) sr += ce.v[v] * r.v[v];
        acc += te.cre * sr;  // Re(c_k) <P_k>
      }
      if (cd.action == -2) cont = ce;
    }
    e = cont;
  };
  for (int st = 0; st < ch.nsteps; st += 2) {
    StepDev sd1;
    run_step(st, sd0, sd1, fa, fb);
    if (st + 1 >= ch.nsteps) break;
    run_step(st + 1, sd1, sd0, fb, fa);
  }
  acc = wave_sum(acc);
  if (lane == 0) parts[inst.part_off + it.y] = acc;
}

extern "C" __global__ void __launch_bounds__(256) GV_FOUR_WAVES
k_rchains_gfwd(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
               const ProbDev *__restrict__ probs, const double *__restrict__ rgarena,
               const double *__restrict__ rgarenaT, double2 *__restrict__ spills, double *__restrict__ parts,
               double2 *__restrict__ estore, double2 *__restrict__ gleafs, double2 *__restrict__ leafs) {
  __shared__ double tr[4][16][17];
  __shared__ int ptab[64 * 4];
  pack_table(ptab);
  rgfwd_body<false>(items, nitems, insts, probs, rgarena, rgarenaT, spills, parts, estore, gleafs, leafs, tr, ptab);
}

extern "C" __global__ void __launch_bounds__(256) GV_FWD_IZ_WAVES
k_rchains_gfwd_iz(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
                  const ProbDev *__restrict__ probs, const double *__restrict__ rgarena,
                  const double *__restrict__ rgarenaT, double2 *__restrict__ spills, double *__restrict__ parts,
                  double2 *__restrict__ estore, double2 *__restrict__ gleafs, double2 *__restrict__ leafs) {
  __shared__ double tr[4][16][17];
  __shared__ int ptab[64 * 4];
  pack_table(ptab);
  rgfwd_body<true>(items, nitems, insts, probs, rgarena, rgarenaT, spills, parts, estore, gleafs, leafs, tr, ptab);
}

// I/Z forests with the zero structure: flagged chains run rgfwd_struct_item.
extern "C" __global__ void __launch_bounds__(256) GV_FWD_IZ_WAVES
k_rchains_gfwd_izs(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
                   const ProbDev *__restrict__ probs, const double *__restrict__ rgarena,
                   const double *__restrict__ rgarenaT, const double *__restrict__ farena,
                   const double *__restrict__ farenaT, const unsigned *__restrict__ smask,
                   const int *__restrict__ cflags, double2 *__restrict__ spills, double *__restrict__ parts,
                   double2 *__restrict__ estore, double2 *__restrict__ gleafs, double2 *__restrict__ leafs) {
  __shared__ double tr[4][16][17];
  __shared__ int ptab[64 * 4];
  pack_table(ptab);
  const int w = xcd_item();
  if (w >= nitems) return;
  const int2 it = uload(items + w);
  const InstDev inst = uload(insts + it.x);
  if (uload(cflags + inst.part_off + it.y))
    rgfwd_struct_item(it, threadIdx.x & 63, insts, probs, rgarena, rgarenaT, farena, farenaT, smask, spills, parts,
                      estore, gleafs, leafs, tr[wave_id()], ptab);
  else
    rgfwd_body<true>(items, nitems, insts, probs, rgarena, rgarenaT, spills, parts, estore, gleafs, leafs, tr, ptab);
}

// Backward (k_chains_bwd in real arithmetic).  For a step with input E and
// site matrices M_i: Ybar00 = I+Z, Ybar11 = I-Z, Ybar01 = Ybar10 = X (child
// adjoints), G_i = sum_j M_j Ybar_ji, Ebar = sum_i G_i M_i^T, Mbar_i = 2 E G_i.
// IZ: I/Z children only, so Ybar01 = Ybar10 = 0 and G_i = Ybar_ii M_i.
template <bool IZ>
__device__ __forceinline__ void rbwd_body(const int2 *__restrict__ items, int nitems,
                                          const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
                                          const double *__restrict__ rgarena, const double *__restrict__ rgarenaT,
                                          const double2 *__restrict__ spills, double2 *__restrict__ adjspill,
                                          const double2 *__restrict__ estore, const double2 *__restrict__ gleafs,
                                          double2 *__restrict__ aslot, const unsigned *__restrict__ smask,
                                          const int *__restrict__ cflags, double (*tr)[16][17]) {
  const int lane = threadIdx.x & 63;
  const int wv = wave_id();
  const int nb = gridDim.x, b = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = b & 7;
  const int logical = xcd * q8 + min(xcd, r8) + (b >> 3);
  const int w = logical * (int)(blockDim.x >> 6) + wv;
  if (w >= nitems) return;
  const int2 it = uload(items + w);
  const InstDev inst = uload(insts + it.x);
  const ProbDev P = uload(probs + inst.prob);
  const ChainDev ch = uload(P.chains + it.y);
  // M_i as the forward pass read it (radvance) and in the A-operand form (its transpose)
  const double *sites = (ch.dir != 0 ? rgarenaT : rgarena) + inst.site_off;
  const double *sites_a = (ch.dir != 0 ? rgarena : rgarenaT) + inst.site_off;
  const double2 *spill = spills + inst.spill_off;
  double2 *aspill = adjspill + inst.spill_off;
  const double2 *es = estore + inst.ck_off * GV_ENV;
  const double2 *gl = gleafs + inst.gleaf_off * GV_ENV;
  double2 *as = aslot + inst.step_off * GV_SLOT;
  const bool transposed = ch.dir != 0;
  const int a16 = lane & 15, g4 = lane >> 4;
  // mask words of a suffix chain the forward ran over the zero structure
  // (k_rchains_gfwd_izs), for the recomputed steps
  const unsigned *smc = smask && transposed && uload(cflags + inst.part_off + it.y)
                            ? smask + (inst.site_off / GV_SLOT) * 2 + 1
                            : nullptr;

  REnv eb;  // adjoint of the environment continuing the chain
#pragma unroll
  for (int t = 0; t < 4; ++t) eb.v[t] = 0.0;
  for (int st = ch.nsteps - 1; st >= 0; --st) {
    const int step = ch.step0 + st;
    const StepDev sd = uload(P.steps + step);
    REnv E;
    const int cs = uload(P.ck_slot + step);
    if (cs >= 0) {
      renv_load_sym_nt(rslot(es, cs), lane, E);
    } else {  // recompute from the nearest checkpoint (step 0 of a chain is always kept)
      int s0 = st;
      while (uload(P.ck_slot + ch.step0 + s0) < 0) --s0;
      renv_load_sym_nt(rslot(es, uload(P.ck_slot + ch.step0 + s0)), lane, E);
      for (int q = s0; q < st; ++q) {
        radvance(P, sites, ch.step0 + q, lane, E, smc, tr[wv]);
        renv_symmetrize<false>(E, lane, tr[wv], nullptr, nullptr);  // as the forward pass did before step q + 1
      }
    }
    RFrag fa;  // M_i in the A-operand form
    rfrag_load(sites_a + (size_t)sd.site * GV_SLOT, lane, fa);
    REnv y00, y11, y01;
#pragma unroll
    for (int t = 0; t < 4; ++t) y00.v[t] = y11.v[t] = y01.v[t] = 0.0;  // y01 stays 0 when IZ
    for (int c = 0; c < sd.nchild; ++c) {
      const ChildDev cd = uload(P.children + sd.child0 + c);
      REnv cb;
      if (cd.action == -2) cb = eb;
      else if (cd.action >= 0) renv_load_sym(rslot(aspill, cd.action), lane, cb);
      else {
#pragma unroll
        for (int t = 0; t < 4; ++t) cb.v[t] = 0.0;
      }
      // term ends (prefix forest): Cbar += Re(c_k) Rt_k
      for (int t = 0; t < cd.nterm; ++t) {
        const TermEnd te = uload(P.term_list + cd.term0 + t);
        const double ck = te.cre;
        if (te.rslot >= 0) {
          REnv r;
          renv_load_sym(rslot(spill, te.rslot), lane, r);
#pragma unroll
          for (int v = 0; v < 4; ++v) cb.v[v] += ck * r.v[v];
        } else if (lane == 0) {
          cb.v[0] += ck;  // unit right environment
        }
      }
      // suffix-forest slot read by prefix term ends: Cbar += Re(c_k) E_left(k)
      if (cd.store >= 0) {
        const int q1 = uload(P.seed_begin + cd.store + 1);
        for (int q = uload(P.seed_begin + cd.store); q < q1; ++q) {
          const SeedTerm sdt = uload(P.seed_terms + q);
          REnv le;
          renv_load_sym(rslot(gl, sdt.gleaf), lane, le);
          const double ck = uload(P.coeffs + 2 * sdt.term);
#pragma unroll
          for (int v = 0; v < 4; ++v) cb.v[v] += ck * le.v[v];
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        if (cd.kind == CK_I) { y00.v[t] += cb.v[t]; y11.v[t] += cb.v[t]; }
        else if (IZ || cd.kind == CK_Z) { y00.v[t] += cb.v[t]; y11.v[t] -= cb.v[t]; }
        else y01.v[t] += cb.v[t];  // X: Ybar01 = Ybar10
      }
    }
    // G_i (as the next products' A operand): G_0 = Ybar00 M_0 + Ybar01 M_1, G_1 = Ybar11 M_1 + Ybar10 M_0
    d4 g0 = {0.0, 0.0, 0.0, 0.0}, g1 = {0.0, 0.0, 0.0, 0.0};
    if (IZ || (sd.need & 1)) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        g0 = mfma(y00.v[t], fa.v[0][t], g0);
        g1 = mfma(y11.v[t], fa.v[1][t], g1);
      }
    }
    if (!IZ && (sd.need & 2)) {
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        g0 = mfma(y01.v[t], fa.v[1][t], g0);
        g1 = mfma(y01.v[t], fa.v[0][t], g1);
      }
    }
    // Ebar = sum_i G_i M_i^T
    d4 er = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      er = mfma(g0[t], fa.v[0][t], er);
      er = mfma(g1[t], fa.v[1][t], er);
    }
    // Mbar_i = 2 E G_i: G_i to the B-operand layout through LDS
    d4 mb[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const d4 &g = i ? g1 : g0;
#pragma unroll
      for (int t = 0; t < 4; ++t) tr[wv][a16][kdx(g4, t)] = g[t];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double d[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) d[v] = 2.0 * tr[wv][kdx(g4, v)][a16];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // d is 2 G_i as the B operand and 2 G_i^T as the A operand, and E in
      // its layout is E^T as the B operand: Mbar_i = E (2 G_i), and for suffix
      // chains (whose slot holds the adjoint of A_i^T) Mbar_i^T = (2 G_i^T) E^T.
      // D layout either way, so lanes with consecutive columns store
      // consecutive slot elements.
      d4 m = {0.0, 0.0, 0.0, 0.0};
      if (transposed) {
#pragma unroll
        for (int t = 0; t < 4; ++t) m = mfma(d[t], E.v[t], m);
      } else {
#pragma unroll
        for (int t = 0; t < 4; ++t) m = mfma(E.v[t], d[t], m);
      }
      mb[i] = m;
    }
    radj_store((double *)(as + (size_t)step * GV_SLOT), uload(P.site_tail + sd.site), a16, g4, mb[0], mb[1]);
#pragma unroll
    for (int t = 0; t < 4; ++t) eb.v[t] = er[t];
  }
  if (ch.src >= 0) renv_store_sym(rslot(aspill, ch.src), lane, eb);
}

extern "C" __global__ void __launch_bounds__(256) GV_FOUR_WAVES
k_rchains_bwd(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
              const ProbDev *__restrict__ probs, const double *__restrict__ rgarena,
              const double *__restrict__ rgarenaT, const double2 *__restrict__ spills, double2 *__restrict__ adjspill,
              const double2 *__restrict__ estore, const double2 *__restrict__ gleafs, double2 *__restrict__ aslot,
              const unsigned *__restrict__ smask, const int *__restrict__ cflags) {
  __shared__ double tr[4][16][17];
  rbwd_body<false>(items, nitems, insts, probs, rgarena, rgarenaT, spills, adjspill, estore, gleafs, aslot, smask,
                   cflags, tr);
}

extern "C" __global__ void __launch_bounds__(256) GV_BWD_IZ_WAVES
k_rchains_bwd_iz(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
                 const ProbDev *__restrict__ probs, const double *__restrict__ rgarena,
                 const double *__restrict__ rgarenaT, const double2 *__restrict__ spills,
                 double2 *__restrict__ adjspill, const double2 *__restrict__ estore,
                 const double2 *__restrict__ gleafs, double2 *__restrict__ aslot,
                 const unsigned *__restrict__ smask, const int *__restrict__ cflags) {
  __shared__ double tr[4][16][17];
  rbwd_body<true>(items, nitems, insts, probs, rgarena, rgarenaT, spills, adjspill, estore, gleafs, aslot, smask,
                  cflags, tr);
}

// Per (instance, site): sum of its steps' real adjoints (radj_store's compact
// layout, 512 >> tail-count entries) in step order, written as a complex slot
// (imaginary parts zero, zeros beyond the pulled-back dimensions) for
// k_site_bwd, which resumes the local program before the pulled-back halves.
extern "C" __global__ void __launch_bounds__(256)
k_rsite_reduce(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
               const double2 *__restrict__ aslot, double2 *__restrict__ abar) {
  __shared__ double red[512];
  const InstDev inst = uload(insts + blockIdx.y);
  const int s = blockIdx.x;
  if (s >= inst.n) return;
  const ProbDev P = uload(probs + inst.prob);
  const int t = threadIdx.x;
  const int tail = uload(P.site_tail + s), nt = tail & 3;
  const double2 *as = aslot + inst.step_off * GV_SLOT;
  double s0 = 0.0, s1 = 0.0;
  if (2 * t < (512 >> nt)) {
    const int q1 = uload(P.site_steps_begin + s + 1);
    int q = uload(P.site_steps_begin + s);
    for (; q + 8 <= q1; q += 8) {
      double2 a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = as[(size_t)uload(P.site_steps + q + u) * GV_SLOT + t];  // doubles 2t, 2t+1
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        s0 += a[u].x;
        s1 += a[u].y;
      }
    }
    for (; q < q1; ++q) {
      const double2 a = as[(size_t)uload(P.site_steps + q) * GV_SLOT + t];
      s0 += a.x;
      s1 += a.y;
    }
  }
  red[2 * t] = s0;
  red[2 * t + 1] = s1;
  __syncthreads();
  double2 *dst = abar + inst.site_off + (size_t)s * GV_SLOT;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int k = tail_compact_index(tail, 2 * t + u);
    dst[2 * t + u] = make_double2(k >= 0 ? red[k] : 0.0, 0.0);
  }
}

// Real parts of the prepared site arena (imaginary parts are exactly zero on
// this path), as A_i into rarena and as A_i^T into rarenaT.
extern "C" __global__ void __launch_bounds__(256)
k_pack_real(const double2 *__restrict__ arena, double *__restrict__ rarena, double *__restrict__ rarenaT,
            long long n) {
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    rarena[e] = arena[e].x;
    rarenaT[e] = arena[pack_transposed_src(e)].x;
  }
}

extern "C" __global__ void __launch_bounds__(256) GV_FOUR_WAVES
k_rchains(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
          const ProbDev *__restrict__ probs, const double *__restrict__ rarena, long long nsite,
                    const unsigned *__restrict__ smask,
          double2 *__restrict__ leafs, double2 *__restrict__ spills, double2 *__restrict__ terms) {
  rchains_body<true, false>(items, nitems, insts, probs, rarena, rarena + nsite, smask, leafs, spills, terms, nullptr);
}

extern "C" __global__ void __launch_bounds__(256) GV_FOUR_WAVES
k_rchains_energy(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
                 const ProbDev *__restrict__ probs, const double *__restrict__ rarena, long long nsite,
                    const unsigned *__restrict__ smask,
                 double2 *__restrict__ leafs, double2 *__restrict__ spills, double *__restrict__ parts) {
  rchains_body<false, false>(items, nitems, insts, probs, rarena, rarena + nsite, smask, leafs, spills, nullptr, parts);
}

// Forests whose steps have I/Z children only.
extern "C" __global__ void __launch_bounds__(256) GV_FIVE_WAVES
k_rchains_iz(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
             const ProbDev *__restrict__ probs, const double *__restrict__ rarena, long long nsite,
                    const unsigned *__restrict__ smask,
             double2 *__restrict__ leafs, double2 *__restrict__ spills, double2 *__restrict__ terms) {
  rchains_body<true, true>(items, nitems, insts, probs, rarena, rarena + nsite, smask, leafs, spills, terms, nullptr);
}

extern "C" __global__ void __launch_bounds__(256) GV_IZ_WAVES
k_rchains_energy_iz(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
                    const ProbDev *__restrict__ probs, const double *__restrict__ rarena, long long nsite,
                    const unsigned *__restrict__ smask,
                    double2 *__restrict__ leafs, double2 *__restrict__ spills, double *__restrict__ parts) {
  rchains_body<false, true>(items, nitems, insts, probs, rarena, rarena + nsite, smask, leafs, spills, nullptr, parts);
}

// I/Z forests over the merged fragments (zero-structure masks on).
extern "C" __global__ void __launch_bounds__(256) GV_FIVE_WAVES
k_rchains_izm(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
              const ProbDev *__restrict__ probs, const double *__restrict__ rarena, long long nsite,
              const double *__restrict__ farena, long long nslots, const unsigned *__restrict__ smask,
              const int *__restrict__ cflags, double2 *__restrict__ leafs, double2 *__restrict__ spills,
              double2 *__restrict__ terms) {
  rchains_merged_body<true>(items, nitems, insts, probs, rarena, rarena + nsite, farena, farena + nslots * GV_FSLOT,
                            smask, cflags, leafs, spills, terms, nullptr);
}

extern "C" __global__ void __launch_bounds__(256) GV_IZM_WAVES
k_rchains_energy_izm(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
                     const ProbDev *__restrict__ probs, const double *__restrict__ rarena, long long nsite,
                     const double *__restrict__ farena, long long nslots, const unsigned *__restrict__ smask,
                     const int *__restrict__ cflags, double2 *__restrict__ leafs, double2 *__restrict__ spills,
                     double *__restrict__ parts) {
  rchains_merged_body<false>(items, nitems, insts, probs, rarena, rarena + nsite, farena, farena + nslots * GV_FSLOT,
                             smask, cflags, leafs, spills, nullptr, parts);
}

extern "C" __global__ void __launch_bounds__(256)
k_rleaf_terms(const LeafItem *__restrict__ items, int nitems, const double2 *__restrict__ leafs,
              const double2 *__restrict__ spills, double2 *__restrict__ terms) {
  rleaf_body<true>(items, nitems, leafs, spills, terms, nullptr);
}

extern "C" __global__ void __launch_bounds__(256)
k_rleaf_energy(const LeafItem *__restrict__ items, int nitems, const double2 *__restrict__ leafs,
               const double2 *__restrict__ spills, double *__restrict__ parts) {
  rleaf_body<false>(items, nitems, leafs, spills, nullptr, parts);
}

namespace gv {

hipError_t kl_pack_real(long long n, hipStream_t s, const double2 *arena, double *rarena) {
  const long long blocks = std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_pack_real, dim3((unsigned)blocks), dim3(256), 0, s, arena, rarena, rarena + n, n);
  return hipGetLastError();
}

hipError_t kl_rchains(int nitems, int wpb, bool iz, hipStream_t s, const int2 *items, const InstDev *insts,
                      const ProbDev *probs, const double *rarena, long long nsite, const unsigned *smask,
                      const double *farena, const int *cflags, double2 *leafs, double2 *spills, double2 *terms,
                      double *parts) {
  const dim3 grid((nitems + wpb - 1) / wpb), block(64 * wpb);
  const long long nslots = nsite / GV_SLOT;
  if (iz && smask) {  // I/Z forests over the merged fragments
    if (terms)
      hipLaunchKernelGGL(k_rchains_izm, grid, block, 0, s, items, nitems, insts, probs, rarena, nsite, farena, nslots,
                         smask, cflags, leafs, spills, terms);
    else
      hipLaunchKernelGGL(k_rchains_energy_izm, grid, block, 0, s, items, nitems, insts, probs, rarena, nsite, farena,
                         nslots, smask, cflags, leafs, spills, parts);
    return hipGetLastError();
  }
  if (terms)
    hipLaunchKernelGGL(iz ? k_rchains_iz : k_rchains, grid, block, 0, s, items, nitems, insts, probs, rarena, nsite,
                       smask, leafs, spills, terms);
  else
    hipLaunchKernelGGL(iz ? k_rchains_energy_iz : k_rchains_energy, grid, block, 0, s, items, nitems, insts, probs,
                       rarena, nsite, smask, leafs, spills, parts);
  return hipGetLastError();
}

// Zero structure of every site (one wave per site slot) from the real arena
// A (element (l, i, r) at (l*2 + i)*16 + r): for the direct copy (M_i = A_i)
// and the transposed one (M_i = A_i^T), the row groups of M_i with a nonzero,
// whether M_0 and M_1 have disjoint column supports, and the columns of M_1
// (bit layout at GV_SM_DISJOINT).
extern "C" __global__ void __launch_bounds__(256)
k_site_masks(long long nslots, const double *__restrict__ rarena, unsigned *__restrict__ smask,
             double *__restrict__ farena) {
  const long long w = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= nslots) return;
  const int lane = threadIdx.x & 63;
  double a[2][4];
  site_slot_load(rarena + w * GV_SLOT, lane, a);
  site_mask_slot(a, farena + w * GV_FSLOT, farena + (nslots + w) * GV_FSLOT, smask + 2 * w, lane);
}

hipError_t kl_site_masks(long long nslots, hipStream_t s, const double *rarena, unsigned *smask, double *farena) {
  if (nslots <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_site_masks, dim3((unsigned)((nslots + 3) / 4)), dim3(256), 0, s, nslots, rarena, smask,
                     farena);
  return hipGetLastError();
}

// Per (instance, chain): 1 when at most a quarter of the chain's steps are on
// sites without structure (rchains_merged_body).
extern "C" __global__ void __launch_bounds__(256)
k_chain_flags(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
              const unsigned *__restrict__ smask, int *__restrict__ cflags, int grad) {
  const InstDev inst = uload(insts + blockIdx.x);
  const ProbDev P = uload(probs + inst.prob);
  const int c = blockIdx.y * blockDim.x + threadIdx.x;
  if (c >= P.nchains) return;
  cflags[inst.part_off + c] = chain_flag(P, c, smask + (inst.site_off / GV_SLOT) * 2, grad != 0);
}

hipError_t kl_chain_flags(int ninst, int max_chains, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                          const unsigned *smask, int *cflags, int grad) {
  if (ninst <= 0 || max_chains <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_chain_flags, dim3(ninst, (max_chains + 255) / 256), dim3(256), 0, s, insts, probs, smask,
                     cflags, grad);
  return hipGetLastError();
}

hipError_t kl_rchains_gfwd(int nitems, int wpb, bool iz, hipStream_t s, const int2 *items, const InstDev *insts,
                           const ProbDev *probs, const double *rgarena, long long nsite, double2 *spills, double *parts,
                           double2 *estore, double2 *gleafs, double2 *leafs, const double *farena,
                           const unsigned *smask, const int *cflags) {
  if (iz && smask) {
    const long long nslots = nsite / GV_SLOT;
    hipLaunchKernelGGL(k_rchains_gfwd_izs, dim3((nitems + wpb - 1) / wpb), dim3(64 * wpb), 0, s, items, nitems, insts,
                       probs, rgarena, rgarena + nsite, farena, farena + nslots * GV_FSLOT, smask, cflags, spills,
                       parts, estore, gleafs, leafs);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(iz ? k_rchains_gfwd_iz : k_rchains_gfwd, dim3((nitems + wpb - 1) / wpb), dim3(64 * wpb), 0, s, items, nitems, insts, probs,
                     rgarena, rgarena + nsite, spills, parts, estore, gleafs, leafs);
  return hipGetLastError();
}

hipError_t kl_rchains_bwd(int nitems, int wpb, bool iz, hipStream_t s, const int2 *items, const InstDev *insts,
                          const ProbDev *probs, const double *rgarena, long long nsite, const double2 *spills,
                          double2 *adjspill, const double2 *estore, const double2 *gleafs, double2 *aslot,
                          const unsigned *smask, const int *cflags) {
  hipLaunchKernelGGL(iz ? k_rchains_bwd_iz : k_rchains_bwd, dim3((nitems + wpb - 1) / wpb), dim3(64 * wpb), 0, s, items, nitems, insts, probs,
                     rgarena, rgarena + nsite, spills, adjspill, estore, gleafs, aslot, smask, cflags);
  return hipGetLastError();
}

hipError_t kl_rsite_reduce(int ninst, int max_n, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                           const double2 *aslot, double2 *abar) {
  hipLaunchKernelGGL(k_rsite_reduce, dim3(max_n, ninst), dim3(256), 0, s, insts, probs, aslot, abar);
  return hipGetLastError();
}

hipError_t kl_rleaf_terms(int nitems, hipStream_t s, const LeafItem *items, const double2 *leafs,
                          const double2 *spills, double2 *terms, double *parts) {
  if (nitems == 0) return hipSuccess;
  if (terms)
    hipLaunchKernelGGL(k_rleaf_terms, dim3((nitems + 3) / 4), dim3(256), 0, s, items, nitems, leafs, spills, terms);
  else
    hipLaunchKernelGGL(k_rleaf_energy, dim3((nitems + 3) / 4), dim3(256), 0, s, items, nitems, leafs, spills, parts);
  return hipGetLastError();
}

}  // namespace gv
