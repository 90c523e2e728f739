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
// environments of 2 KiB (stored in the first half of the complex 4 KiB spill
// and leaf slots).  The chain/forest schedule, work items, term ends and
// partial-energy layout are those of kernels_energy.hip, so k_reduce_energy /
// k_reduce combine the results unchanged.
#include "device_common.h"

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

// Real site slot layout: element (l, i, r) at (l*2 + i)*16 + r (512 doubles).
__device__ __forceinline__ void rfrag_load(const double *slot, int lane, RFrag &f, bool transposed) {
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int row = kdx(g, t);
      f.v[i][t] = slot[transposed ? ((j * 2 + i) * 16 + row) : ((row * 2 + i) * 16 + j)];
    }
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

// One generation of prefix/suffix chains in real arithmetic; the same work
// items and slot numbering as chains_body in kernels_energy.hip.
template <bool TERMS>
__device__ __forceinline__ void rchains_body(const int2 *__restrict__ items, int nitems,
                                             const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
                                             const double *__restrict__ rarena, double2 *__restrict__ leafs,
                                             double2 *__restrict__ spills, double2 *__restrict__ terms,
                                             double *__restrict__ parts) {
  const int lane = threadIdx.x & 63;
  const int nb = gridDim.x, b = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = b & 7;
  const int logical = xcd * q8 + min(xcd, r8) + (b >> 3);
  const int w = logical * (int)(blockDim.x >> 6) + wave_id();
  if (w >= nitems) return;
  const int2 it = uload(items + w);
  const InstDev inst = uload(insts + it.x);
  const ProbDev P = uload(probs + inst.prob);
  const ChainDev ch = uload(P.chains + it.y);
  const double *sites = rarena + inst.site_off;
  double2 *spill = spills + inst.spill_off;
  double2 *tv = terms + inst.term_off;

  REnv e;
  if (ch.src >= 0) renv_load(rslot(spill, ch.src), lane, e);
  else renv_unit(lane, e);
  const bool transposed = ch.dir != 0;
  double acc = 0.0;

  StepDev sd0 = uload(P.steps + ch.step0);
  RFrag fa, fb;
  rfrag_load(sites + (size_t)sd0.site * GV_SLOT, lane, fa, transposed);
  // one chain step on fragment f; prefetches the next step's fragment into nf
  auto run_step = [&](int st, const StepDev &sd, StepDev &nsd, const RFrag &f, RFrag &nf) {
    if (st + 1 < ch.nsteps) {
      nsd = uload(P.steps + ch.step0 + st + 1);
      rfrag_load(sites + (size_t)nsd.site * GV_SLOT, lane, nf, transposed);
    }
    const d4 x0 = rprod_x(e, f, 0), x1 = rprod_x(e, f, 1);
    REnv cont;
    // pass 0: I/Z children from Y00, Y11; pass 1: X children from Y01, Y10
#pragma nounroll
    for (int pass = 0; pass < 2; ++pass) {
      if (!(sd.need & (1 << pass))) continue;
      const d4 ya = rprod_y(pass ? x1 : x0, f, 0);
      const d4 yb = rprod_y(pass ? x0 : x1, f, 1);
      for (int c = 0; c < sd.nchild; ++c) {
        const ChildDev cd = uload(P.children + sd.child0 + c);
        if ((cd.kind == CK_X) != (pass == 1)) continue;
        REnv ce;
#pragma unroll
        for (int v = 0; v < 4; ++v) ce.v[v] = cd.kind == CK_Z ? ya[v] - yb[v] : ya[v] + yb[v];
        if (cd.store >= 0) renv_store(rslot(spill, cd.store), lane, ce);
        if (cd.defer >= 0) renv_store(rslot(leafs, inst.leaf_off + cd.defer), lane, ce);
        const int nterm_inline = cd.defer >= 0 ? 0 : cd.nterm;
#pragma unroll 2
        for (int t = 0; t < nterm_inline; ++t) {
          const TermEnd te = uload(P.term_list + cd.term0 + t);
          REnv r;
          if (te.rslot >= 0) renv_load(rslot(spill, te.rslot), lane, r);
          else renv_unit(lane, r);  // term ends at site n-1
          double sr = 0.0;
#pragma unroll
          for (int v = 0; v < 4; ++v) sr += ce.v[v] * r.v[v];
          if (TERMS) {
            sr = wave_sum(sr);
            if (lane == 0) tv[te.term] = make_double2(sr, 0.0);
          } else {
            acc += uload(P.coeffs + 2 * te.term) * sr;  // Re(c <P>), <P> real
          }
        }
        if (cd.action == -2) cont = ce;
        else if (cd.action >= 0) renv_store(rslot(spill, cd.action), lane, ce);
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
  if (!TERMS) {
    acc = wave_sum(acc);
    if (lane == 0) parts[inst.part_off + it.y] = acc;
  }
}

template <bool TERMS>
__device__ __forceinline__ void rleaf_body(const int2 *__restrict__ items, int nitems,
                                           const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
                                           const double2 *__restrict__ leafs, const double2 *__restrict__ spills,
                                           double2 *__restrict__ terms, double *__restrict__ parts) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * (int)(blockDim.x >> 6) + wave_id();
  if (w >= nitems) return;
  const int2 it = uload(items + w);
  const InstDev inst = uload(insts + it.x);
  const ProbDev P = uload(probs + inst.prob);
  const LeafDev lf = uload(P.leaves + it.y);
  const double2 *spill = spills + inst.spill_off;
  REnv e;
  renv_load(rslot(leafs, inst.leaf_off + it.y), lane, e);
  double acc = 0.0;
#pragma unroll 4
  for (int t = 0; t < lf.nterm; ++t) {
    const TermEnd te = uload(P.term_list + lf.term0 + t);
    REnv r;
    if (te.rslot >= 0) renv_load(rslot(spill, te.rslot), lane, r);
    else renv_unit(lane, r);
    double sr = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) sr += e.v[v] * r.v[v];
    if (TERMS) {
      sr = wave_sum(sr);
      if (lane == 0) terms[inst.term_off + te.term] = make_double2(sr, 0.0);
    } else {
      acc += uload(P.coeffs + 2 * te.term) * sr;
    }
  }
  if (!TERMS) {
    acc = wave_sum(acc);
    if (lane == 0) parts[inst.part_off + P.nchains + it.y] = acc;
  }
}

}  // namespace

// Real parts of the prepared site arena (imaginary parts are exactly zero on
// this path).
extern "C" __global__ void __launch_bounds__(256)
k_pack_real(const double2 *__restrict__ arena, double *__restrict__ rarena, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    rarena[i] = arena[i].x;
}

extern "C" __global__ void __launch_bounds__(256) GV_FOUR_WAVES
k_rchains(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
          const ProbDev *__restrict__ probs, const double *__restrict__ rarena, double2 *__restrict__ leafs,
          double2 *__restrict__ spills, double2 *__restrict__ terms) {
  rchains_body<true>(items, nitems, insts, probs, rarena, leafs, spills, terms, nullptr);
}

extern "C" __global__ void __launch_bounds__(256) GV_FOUR_WAVES
k_rchains_energy(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
                 const ProbDev *__restrict__ probs, const double *__restrict__ rarena, double2 *__restrict__ leafs,
                 double2 *__restrict__ spills, double *__restrict__ parts) {
  rchains_body<false>(items, nitems, insts, probs, rarena, leafs, spills, nullptr, parts);
}

extern "C" __global__ void __launch_bounds__(256)
k_rleaf_terms(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
              const ProbDev *__restrict__ probs, const double2 *__restrict__ leafs, const double2 *__restrict__ spills,
              double2 *__restrict__ terms) {
  rleaf_body<true>(items, nitems, insts, probs, leafs, spills, terms, nullptr);
}

extern "C" __global__ void __launch_bounds__(256)
k_rleaf_energy(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
               const ProbDev *__restrict__ probs, const double2 *__restrict__ leafs, const double2 *__restrict__ spills,
               double *__restrict__ parts) {
  rleaf_body<false>(items, nitems, insts, probs, leafs, spills, nullptr, parts);
}

namespace gv {

hipError_t kl_pack_real(long long n, hipStream_t s, const double2 *arena, double *rarena) {
  const long long blocks = std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_pack_real, dim3((unsigned)blocks), dim3(256), 0, s, arena, rarena, n);
  return hipGetLastError();
}

hipError_t kl_rchains(int nitems, int wpb, hipStream_t s, const int2 *items, const InstDev *insts,
                      const ProbDev *probs, const double *rarena, double2 *leafs, double2 *spills, double2 *terms,
                      double *parts) {
  if (terms)
    hipLaunchKernelGGL(k_rchains, dim3((nitems + wpb - 1) / wpb), dim3(64 * wpb), 0, s, items, nitems, insts, probs,
                       rarena, leafs, spills, terms);
  else
    hipLaunchKernelGGL(k_rchains_energy, dim3((nitems + wpb - 1) / wpb), dim3(64 * wpb), 0, s, items, nitems, insts,
                       probs, rarena, leafs, spills, parts);
  return hipGetLastError();
}

hipError_t kl_rleaf_terms(int nitems, hipStream_t s, const int2 *items, const InstDev *insts, const ProbDev *probs,
                          const double2 *leafs, const double2 *spills, double2 *terms, double *parts) {
  if (nitems == 0) return hipSuccess;
  if (terms)
    hipLaunchKernelGGL(k_rleaf_terms, dim3((nitems + 3) / 4), dim3(256), 0, s, items, nitems, insts, probs, leafs,
                       spills, terms);
  else
    hipLaunchKernelGGL(k_rleaf_energy, dim3((nitems + 3) / 4), dim3(256), 0, s, items, nitems, insts, probs, leafs,
                       spills, parts);
  return hipGetLastError();
}

}  // namespace gv
