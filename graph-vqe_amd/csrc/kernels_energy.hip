// Analytic-energy kernels (gfx950); state preparation is in kernels_prep.hip.
//
//   k_envs    one wavefront per (instance, direction): identity left
//             environments L_s and transposed right environments Rt_s.
//   k_chains  one wavefront per (instance, chain) of the prefix and suffix
//             forests: the environment stays in MFMA registers along the
//             chain; children that start other chains are spilled to HBM.
//   k_reduce  one workgroup per instance: E = Re sum_k c_k <P_k>, compensated
//             and in a fixed order.
#include "device_common.h"

using namespace gv;

// Identity environments.  Wave w: instance w/2, direction w&1 (0 = L, 1 = Rt).
//   L_0 = [1],  L_{s+1} = sum_i A_i^H L_s A_i
//   Rt_n = [1], Rt_s = sum_i (A_i^T)^H Rt_{s+1} A_i^T  (Rt = R^T, R as in
//   prepareSamplingRightEnvironments, sharded_mps.cuh:61-90, 814-836)
extern "C" __global__ void __launch_bounds__(256)
k_envs(const InstDev *__restrict__ insts, int ninst, const double2 *__restrict__ arena, double2 *__restrict__ envs) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + wave_id();
  if (w >= 2 * ninst) return;
  const InstDev inst = uload(insts + (w >> 1));
  const int dir = w & 1;
  const int n = inst.n;
  const double2 *sites = arena + inst.site_off;
  double2 *L = envs + inst.env_off;
  double2 *Rt = L + (size_t)(n + 1) * GV_ENV;
  Env e, nx;
  Frag f;
  env_unit(lane, e);
  if (dir == 0) {
    env_store(L, lane, e);
    for (int s = 0; s < n; ++s) {
      frag_load(sites + (size_t)s * GV_SLOT, lane, f, false);
      transfer_identity(e, f, nx);
      e = nx;
      env_store(L + (size_t)(s + 1) * GV_ENV, lane, e);
    }
  } else {
    env_store(Rt + (size_t)n * GV_ENV, lane, e);
    for (int s = n - 1; s >= 0; --s) {
      frag_load(sites + (size_t)s * GV_SLOT, lane, f, true);
      transfer_identity(e, f, nx);
      e = nx;
      env_store(Rt + (size_t)s * GV_ENV, lane, e);
    }
  }
}

// One generation of prefix/suffix chains.  Work item (instance, chain) per
// wave.  TERMS: write every <P_k> (observe's per-term data); otherwise each
// lane accumulates its share of sum_k c_k <P_k> and the chain writes one
// partial energy (no per-term wave reductions).
template <bool TERMS, bool IZ>
__device__ __forceinline__ void chains_body(const int2 *__restrict__ items, int nitems,
                                            const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
                                            const double2 *__restrict__ arena, long long nsite,
                                            double2 *__restrict__ leafs, double2 *__restrict__ spills,
                                            double2 *__restrict__ terms, double *__restrict__ parts) {
  const int lane = threadIdx.x & 63;
  // XCD-aware block order: hardware block b runs on XCD b % 8; give each XCD
  // a contiguous range of work items so one instance's chains share an L2.
  const int nb = gridDim.x, b = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = b & 7;
  const int logical = xcd * q8 + min(xcd, r8) + (b >> 3);
  const int w = logical * (int)(blockDim.x >> 6) + wave_id();
  if (w >= nitems) return;
  const int2 it = uload(items + w);
  const InstDev inst = uload(insts + it.x);
  const ProbDev P = uload(probs + inst.prob);
  const ChainDev ch = uload(P.chains + it.y);
  const int n = inst.n;
  // suffix-forest chains read the transposed copy (second half of the arena)
  const double2 *sites = arena + (ch.dir != 0 ? nsite : 0) + inst.site_off;
  double2 *spill = spills + inst.spill_off;
  double2 *tv = terms + inst.term_off;

  // each forest has one root chain starting from the unit environment [1]
  // (left of site 0 / right of site n-1); every other chain starts from a
  // spilled node
  Env e;
  if (ch.src >= 0) env_load_cm(spill + (size_t)ch.src * GV_ENV, lane, e);
  else env_unit(lane, e);
  double acc = 0.0;  // energy mode: this lane's share of sum_k c_k <P_k>
  StepDev sd = uload(P.steps + ch.step0);
  Frag f;
  frag_load(sites + (size_t)sd.site * GV_SLOT, lane, f, false);
  for (int st = 0; st < ch.nsteps; ++st) {
    // prefetch the next step's site tensor while this step computes.  The
    // load is unconditional (the last step reloads its own site): with the
    // same number of loads in flight on every path, the wait for this step's
    // operands leaves the prefetch outstanding (vmcnt counts in order).
    const StepDev nsd = uload(P.steps + ch.step0 + min(st + 1, ch.nsteps - 1));
    Frag nf;
    frag_load(sites + (size_t)nsd.site * GV_SLOT, lane, nf, false);
    d4 x0r, x0i, x1r, x1i;
    prod_x(e, f, 0, x0r, x0i);
    prod_x(e, f, 1, x1r, x1i);
    Env cont;
    // pass 0: I/Z children from Y00 = A_0^H E A_0 and Y11 = A_1^H E A_1;
    // pass 1: X/Y children from Y01 = A_0^H E A_1 and Y10 = A_1^H E A_0
    // (one pair of products live at a time)
#pragma nounroll
    for (int pass = 0; pass < (IZ ? 1 : 2); ++pass) {
    if (!IZ && !(sd.need & (1 << pass))) continue;
    d4 yar, yai, ybr, ybi;
    if (pass == 0) {
      prod_y(x0r, x0i, f, 0, yar, yai);
      prod_y(x1r, x1i, f, 1, ybr, ybi);
    } else {
      prod_y(x1r, x1i, f, 0, yar, yai);
      prod_y(x0r, x0i, f, 1, ybr, ybi);
    }
    for (int c = 0; c < sd.nchild; ++c) {
      const ChildDev cd = uload(P.children + sd.child0 + c);
      if (!IZ && (cd.kind == CK_X || cd.kind == CK_Y) != (pass == 1)) continue;
      Env ce;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // analyticLeft phases, analytic_batch.cuh:11-14
        if (cd.kind == CK_I || cd.kind == CK_X) { ce.re[v] = yar[v] + ybr[v]; ce.im[v] = yai[v] + ybi[v]; }
        else if (cd.kind == CK_Z) { ce.re[v] = yar[v] - ybr[v]; ce.im[v] = yai[v] - ybi[v]; }
        else { ce.re[v] = yai[v] - ybi[v]; ce.im[v] = -yar[v] + ybr[v]; }  // -i Y01 + i Y10
      }
      if (cd.store >= 0) env_store_cm(spill + (size_t)cd.store * GV_ENV, lane, ce);
      // <P> = Tr(E_left R_right) = sum E[b][b'] Rt[b][b'] with the term's
      // suffix-forest environment (or the identity right environment);
      // nodes with many term ends leave them to k_leaf_terms
      if (cd.defer >= 0) env_store_cm(leafs + (inst.leaf_off + cd.defer) * GV_ENV, lane, ce);
      const int nterm_inline = cd.defer >= 0 ? 0 : cd.nterm;
#pragma unroll 2
      for (int t = 0; t < nterm_inline; ++t) {
        const TermEnd te = uload(P.term_list + cd.term0 + t);
        Env r;
        if (te.rslot >= 0) env_load_cm(spill + (size_t)te.rslot * GV_ENV, lane, r);
        else env_unit(lane, r);  // term ends at site n-1
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          sr += ce.re[v] * r.re[v] - ce.im[v] * r.im[v];
          si += ce.re[v] * r.im[v] + ce.im[v] * r.re[v];
        }
        if (TERMS) {
          sr = wave_sum(sr);
          si = wave_sum(si);
          if (lane == 0) tv[te.term] = make_double2(sr, si);
        } else {
          const double2 cf = uload((const double2 *)P.coeffs + te.term);
          acc += cf.x * sr - cf.y * si;  // Re(c <P>)
        }
      }
      if (cd.action == -2) cont = ce;
      else if (cd.action >= 0) env_store_cm(spill + (size_t)cd.action * GV_ENV, lane, ce);
    }
    }
    e = cont;
    sd = nsd;
    f = nf;
  }
  if (!TERMS) {
    acc = wave_sum(acc);
    if (lane == 0) parts[inst.part_off + it.y] = acc;
  }
}

extern "C" __global__ void __launch_bounds__(256) GV_TWO_WAVES
k_chains(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
         const ProbDev *__restrict__ probs, const double2 *__restrict__ arena, long long nsite,
         double2 *__restrict__ leafs, double2 *__restrict__ spills, double2 *__restrict__ terms) {
  chains_body<true, false>(items, nitems, insts, probs, arena, nsite, leafs, spills, terms, nullptr);
}

extern "C" __global__ void __launch_bounds__(256) GV_TWO_WAVES
k_chains_energy(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
                const ProbDev *__restrict__ probs, const double2 *__restrict__ arena, long long nsite,
                double2 *__restrict__ leafs, double2 *__restrict__ spills, double *__restrict__ parts) {
  chains_body<false, false>(items, nitems, insts, probs, arena, nsite, leafs, spills, nullptr, parts);
}

// Forests whose steps have I/Z children only (no X or Y words): one product
// pass, and the site fragment and X_i products die once Y00 and Y11 are
// formed.  Same arithmetic per instance.
extern "C" __global__ void __launch_bounds__(256) GV_CIZ_WAVES
k_chains_iz(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
            const ProbDev *__restrict__ probs, const double2 *__restrict__ arena, long long nsite,
            double2 *__restrict__ leafs, double2 *__restrict__ spills, double2 *__restrict__ terms) {
  chains_body<true, true>(items, nitems, insts, probs, arena, nsite, leafs, spills, terms, nullptr);
}

extern "C" __global__ void __launch_bounds__(256) GV_CIZ_WAVES
k_chains_energy_iz(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
                   const ProbDev *__restrict__ probs, const double2 *__restrict__ arena, long long nsite,
                   double2 *__restrict__ leafs, double2 *__restrict__ spills, double *__restrict__ parts) {
  chains_body<false, true>(items, nitems, insts, probs, arena, nsite, leafs, spills, nullptr, parts);
}

// Transposed copy of the site arena: slot s of the second half holds A_i^T
// (element (l, i, r) = A[r][i][l]), so suffix-forest chains load row fragments.
extern "C" __global__ void __launch_bounds__(256)
k_transpose_sites(double2 *__restrict__ arena, long long n) {
  for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long long)gridDim.x * 256) {
    const long long base = e - (e % GV_SLOT);
    const int k = (int)(e % GV_SLOT), row = k >> 5, i = (k >> 4) & 1, j = k & 15;
    arena[n + e] = arena[base + (j * 2 + i) * 16 + row];
  }
}

// Term ends of the deferred prefix leaves: one wave per (instance, leaf),
// several suffix environments in flight.  TERMS: <P_k> per term; otherwise
// one partial energy per leaf after the chain partials.
template <bool TERMS>
__device__ __forceinline__ void leaf_body(const int2 *__restrict__ items, int nitems,
                                          const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
                                          const double2 *__restrict__ leafs, const double2 *__restrict__ spills,
                                          double2 *__restrict__ terms, double *__restrict__ parts) {
  const int lane = threadIdx.x & 63;
  const int w = xcd_item();
  if (w >= nitems) return;
  const int2 it = uload(items + w);
  const InstDev inst = uload(insts + it.x);
  const ProbDev P = uload(probs + inst.prob);
  const LeafDev lf = uload(P.leaves + it.y);
  const double2 *spill = spills + inst.spill_off;
  Env e;
  env_load_cm(leafs + (inst.leaf_off + it.y) * GV_ENV, lane, e);
  double acc = 0.0;
#pragma unroll 4
  for (int t = 0; t < lf.nterm; ++t) {
    const TermEnd te = uload(P.term_list + lf.term0 + t);
    Env r;
    if (te.rslot >= 0) env_load_cm(spill + (size_t)te.rslot * GV_ENV, lane, r);
    else env_unit(lane, r);
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      sr += e.re[v] * r.re[v] - e.im[v] * r.im[v];
      si += e.re[v] * r.im[v] + e.im[v] * r.re[v];
    }
    if (TERMS) {
      sr = wave_sum(sr);
      si = wave_sum(si);
      if (lane == 0) terms[inst.term_off + te.term] = make_double2(sr, si);
    } else {
      const double2 cf = uload((const double2 *)P.coeffs + te.term);
      acc += cf.x * sr - cf.y * si;
    }
  }
  if (!TERMS) {
    acc = wave_sum(acc);
    if (lane == 0) parts[inst.part_off + P.nchains + it.y] = acc;
  }
}

extern "C" __global__ void __launch_bounds__(256)
k_leaf_terms(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
             const ProbDev *__restrict__ probs, const double2 *__restrict__ leafs, const double2 *__restrict__ spills,
             double2 *__restrict__ terms) {
  leaf_body<true>(items, nitems, insts, probs, leafs, spills, terms, nullptr);
}

extern "C" __global__ void __launch_bounds__(256)
k_leaf_energy(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
              const ProbDev *__restrict__ probs, const double2 *__restrict__ leafs, const double2 *__restrict__ spills,
              double *__restrict__ parts) {
  leaf_body<false>(items, nitems, insts, probs, leafs, spills, nullptr, parts);
}

// E = Re sum_k c_k <P_k> (+ identity words counted as <P> = 1,
// simulator_mps_dist.cu:1031-1037, 1467-1481).  Neumaier summation per thread,
// then a fixed-order pairwise combination: deterministic.
extern "C" __global__ void __launch_bounds__(256)
k_reduce(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs, double2 *__restrict__ terms,
         double *__restrict__ energies) {
  __shared__ double ssum[256], scomp[256];
  const int tid = threadIdx.x;
  const InstDev inst = insts[blockIdx.x];
  const ProbDev P = probs[inst.prob];
  const int *act = P.active;
  double2 *tv = terms + inst.term_off;
  double sum = 0.0, comp = 0.0;
  for (int k = tid; k < P.nterms; k += 256) {
    double2 v = tv[k];
    if (!act[k]) {
      v = make_double2(1.0, 0.0);
      tv[k] = v;
    }
    const double term = P.coeffs[2 * k] * v.x - P.coeffs[2 * k + 1] * v.y;
    const double t = sum + term;
    comp += (fabs(sum) >= fabs(term)) ? (sum - t) + term : (term - t) + sum;
    sum = t;
  }
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
}

// Energy from chain partials: identity words (<P> = 1) plus the partial sums
// of every chain, compensated and in a fixed order.
extern "C" __global__ void __launch_bounds__(256)
k_reduce_energy(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
                const double *__restrict__ parts, double *__restrict__ energies, int with_leaves) {
  __shared__ double ssum[256], scomp[256];
  const InstDev inst = insts[blockIdx.x];
  reduce_energy_block(inst, probs[inst.prob], parts, energies, with_leaves, ssum, scomp);
}

// ------------------------------------------------------------ host launchers
namespace gv {

hipError_t kl_envs(int ninst, hipStream_t s, const InstDev *insts, const double2 *arena, double2 *envs) {
  hipLaunchKernelGGL(k_envs, dim3((2 * ninst + 3) / 4), dim3(256), 0, s, insts, ninst, arena, envs);
  return hipGetLastError();
}

hipError_t kl_transpose_sites(long long n, hipStream_t s, double2 *arena) {
  const long long blocks = std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_transpose_sites, dim3((unsigned)blocks), dim3(256), 0, s, arena, n);
  return hipGetLastError();
}

hipError_t kl_chains(int nitems, int wpb, bool iz, hipStream_t s, const int2 *items, const InstDev *insts,
                     const ProbDev *probs, const double2 *arena, long long nsite, double2 *leafs, double2 *spills,
                     double2 *terms) {
  hipLaunchKernelGGL(iz ? k_chains_iz : k_chains, dim3((nitems + wpb - 1) / wpb), dim3(64 * wpb), 0, s, items,
                     nitems, insts, probs, arena, nsite, leafs, spills, terms);
  return hipGetLastError();
}

hipError_t kl_leaf_terms(int nitems, hipStream_t s, const int2 *items, const InstDev *insts, const ProbDev *probs,
                         const double2 *leafs, const double2 *spills, double2 *terms, double *parts) {
  if (nitems == 0) return hipSuccess;
  if (terms)
    hipLaunchKernelGGL(k_leaf_terms, dim3((nitems + 3) / 4), dim3(256), 0, s, items, nitems, insts, probs, leafs,
                       spills, terms);
  else
    hipLaunchKernelGGL(k_leaf_energy, dim3((nitems + 3) / 4), dim3(256), 0, s, items, nitems, insts, probs, leafs,
                       spills, parts);
  return hipGetLastError();
}

hipError_t kl_chains_energy(int nitems, int wpb, bool iz, hipStream_t s, const int2 *items, const InstDev *insts,
                            const ProbDev *probs, const double2 *arena, long long nsite, double2 *leafs,
                            double2 *spills, double *parts) {
  hipLaunchKernelGGL(iz ? k_chains_energy_iz : k_chains_energy, dim3((nitems + wpb - 1) / wpb), dim3(64 * wpb), 0,
                     s, items, nitems, insts, probs, arena, nsite, leafs, spills, parts);
  return hipGetLastError();
}

hipError_t kl_reduce_energy(int ninst, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                            const double *parts, double *energies, int with_leaves) {
  hipLaunchKernelGGL(k_reduce_energy, dim3(ninst), dim3(256), 0, s, insts, probs, parts, energies, with_leaves);
  return hipGetLastError();
}

hipError_t kl_reduce(int ninst, hipStream_t s, const InstDev *insts, const ProbDev *probs, double2 *terms,
                     double *energies) {
  hipLaunchKernelGGL(k_reduce, dim3(ninst), dim3(256), 0, s, insts, probs, terms, energies);
  return hipGetLastError();
}

}  // namespace gv
