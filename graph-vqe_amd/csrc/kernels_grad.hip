// Reverse-gradient kernels (gfx950): value and dE/dx of every instance.
//
// The reverse engine of the reference (test/tools/reverse_mps.py:28-99,
// 141-184) differentiates an MPS whose CNOTs are applied as exact bond
// doublings, so every site tensor is a function of its own gates only.  The
// pass here:
//   k_build_local   one wave per (instance, site): the site's local program
//                   (rotations, fixed gates, CNOT halves) from [1, 0]; keeps
//                   each rotation's input tensor for the backward pass
//   k_chains_gfwd   the energy forests of kernels_energy.hip, storing every
//                   step's input environment and the prefix leaves whose
//                   terms end in a suffix-forest slot
//   k_chains_bwd    one wave per (instance, chain), generations in reverse:
//                   adjoint environments ride in registers down the chain,
//                   spilled children hand theirs over through HBM; every
//                   step writes its site-tensor adjoint to a slot of its own
//   k_site_reduce   per (instance, site): sum of its steps' adjoints, in a
//                   fixed order (deterministic)
//   k_site_bwd      one wave per (instance, site): the local program in
//                   reverse -> d E / d angle of every rotation
//   k_param_grad    per (instance, parameter): sum over its rotations
//
// Adjoint convention: for a real loss L and complex z, zbar = dL/dRe z +
// i dL/dIm z, so dL = Re sum conj(zbar) dz.  The loss is
// sum_k Re(c_k) Re<P_k> (real coefficients, reverse_mps.py:78-81).
#include "device_common.h"

using namespace gv;

namespace {

// U(theta) = cos(theta/2) I - i sin(theta/2) sigma_axis and dU/dtheta, in the
// branch-free form of the state-preparation kernel.
__device__ __forceinline__ void rot_pair(int axis, double theta, double2 U[4], double2 dU[4]) {
  double sn, cs;
  sincos(0.5 * theta, &sn, &cs);
  const double fx = axis == 0 ? 1.0 : 0.0, fy = axis == 1 ? 1.0 : 0.0, fz = axis == 2 ? 1.0 : 0.0;
  U[0] = make_double2(cs, -sn * fz);
  U[1] = make_double2(-sn * fy, -sn * fx);
  U[2] = make_double2(sn * fy, -sn * fx);
  U[3] = make_double2(cs, sn * fz);
  const double h = 0.5;
  dU[0] = make_double2(-h * sn, -h * cs * fz);
  dU[1] = make_double2(-h * cs * fy, -h * cs * fx);
  dU[2] = make_double2(h * cs * fy, -h * cs * fx);
  dU[3] = make_double2(-h * sn, h * cs * fz);
}

__device__ __forceinline__ void local_matrix(const LocalOp &op, const double *x, const double *fixes,
                                             double2 U[4], double2 dU[4]) {
  if (op.kind == LO_ROT) {
    const double theta = (op.param >= 0 ? op.scale * x[op.param] : 0.0) + op.offset;
    rot_pair(op.arg, theta, U, dU);
  } else {
    const double *F = fixes + 8 * (op.arg < 0 ? 0 : op.arg);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      U[e] = make_double2(F[2 * e], F[2 * e + 1]);
      dU[e] = make_double2(0.0, 0.0);
    }
  }
}

// Slot element e = (l*2 + i)*16 + r.
__device__ __forceinline__ int sidx(int l, int i, int r) { return (l * 2 + i) * 16 + r; }

}  // namespace

// ------------------------------------------------------------------ build
// grid (max sites, instances), one wave per block.
extern "C" __global__ void __launch_bounds__(64)
k_build_local(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
              const double *__restrict__ params, double2 *__restrict__ garena, double2 *__restrict__ rotin,
              int *__restrict__ status) {
  __shared__ double2 buf[GV_SLOT];  // one wave: read into registers, barrier, write back
  const int lane = threadIdx.x;
  const InstDev inst = insts[blockIdx.y];
  const int s = blockIdx.x;
  if (status && s == 0 && lane == 0) status[blockIdx.y] = ST_OK;  // no SVD, no guard
  if (s >= inst.n) return;
  const ProbDev &P = probs[inst.prob];
  const double *x = params + inst.param_off;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int e = lane + 64 * k;
    buf[e] = make_double2(e == 0 ? 1.0 : 0.0, 0.0);  // |0> of a (1,2,1) site
  }
  __syncthreads();
  const int o0 = P.lop_begin[s], o1 = P.lop_begin[s + 1];
  for (int o = o0; o < o1; ++o) {
    const LocalOp op = P.lops[o];
    double2 U[4], dU[4];
    local_matrix(op, x, P.fixes, U, dU);
    // rotation inputs are kept only for the reverse pass (rotin == nullptr: energy-path preparation)
    double2 *keep = (op.kind == LO_ROT && rotin != nullptr) ? rotin + (inst.rot_off + op.rot) * GV_SLOT : nullptr;
    const double2 *in = buf;
    double2 w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int e = lane + 64 * k;
      const int l = e >> 5, i = (e >> 4) & 1, r = e & 15;
      if (keep) keep[e] = in[e];
      double2 v;
      switch (op.kind) {
        case LO_CTRL_R:  // A'[l,i,2r+s] = A[l,i,r] d(i,s)
          v = ((r & 1) == i) ? in[sidx(l, i, r >> 1)] : make_double2(0.0, 0.0);
          break;
        case LO_CTRL_L:  // A'[2l+s,i,r] = A[l,i,r] d(i,s)
          v = ((l & 1) == i) ? in[sidx(l >> 1, i, r)] : make_double2(0.0, 0.0);
          break;
        case LO_TARG_R:  // A'[l,i,2r+s] = A[l,i^s,r]
          v = in[sidx(l, i ^ (r & 1), r >> 1)];
          break;
        case LO_TARG_L:  // A'[2l+s,i,r] = A[l,i^s,r]
          v = in[sidx(l >> 1, i ^ (l & 1), r)];
          break;
        default:  // 2x2 gate on the physical index
          v = cadd(cmul(i ? U[2] : U[0], in[sidx(l, 0, r)]), cmul(i ? U[3] : U[1], in[sidx(l, 1, r)]));
      }
      w[k] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) buf[lane + 64 * k] = w[k];
    __syncthreads();
  }
  double2 *dst = garena + inst.site_off + (size_t)s * GV_SLOT;
#pragma unroll
  for (int k = 0; k < 8; ++k) dst[lane + 64 * k] = buf[lane + 64 * k];
}

// --------------------------------------------------------------- forward
// Energy chains on the bond-doubling state, keeping what the backward pass
// reads: the input environment of every step and the prefix leaves that end
// terms against a suffix-forest slot.
extern "C" __global__ void __launch_bounds__(256) GV_TWO_WAVES
k_chains_gfwd(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
              const ProbDev *__restrict__ probs, const double2 *__restrict__ garena, long long nsite,
              double2 *__restrict__ spills, double *__restrict__ parts, double2 *__restrict__ estore,
              double2 *__restrict__ gleafs) {
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
  const double2 *sites = garena + (ch.dir != 0 ? nsite : 0) + inst.site_off;  // transposed copy for suffix chains
  double2 *spill = spills + inst.spill_off;
  double2 *es = estore + inst.ck_off * GV_ENV;
  double2 *gl = gleafs + inst.gleaf_off * GV_ENV;
  Env e;
  if (ch.src >= 0) env_load_herm(spill + (size_t)ch.src * GV_ENV, lane, e);
  else env_unit(lane, e);
  double acc = 0.0;
  for (int st = 0; st < ch.nsteps; ++st) {
    const int step = ch.step0 + st;
    const StepDev sd = uload(P.steps + step);
    Frag f;
    frag_load(sites + (size_t)sd.site * GV_SLOT, lane, f, false);
    const int cs = uload(P.ck_slot + step);
    if (cs >= 0) env_store_cm_nt(es + (size_t)cs * GV_ENV, lane, e);
    d4 x0r, x0i, x1r, x1i;
    prod_x(e, f, 0, x0r, x0i);
    prod_x(e, f, 1, x1r, x1i);
    d4 y00r, y00i, y11r, y11i, y01r, y01i, y10r, y10i;
    if (sd.need & 1) {
      prod_y(x0r, x0i, f, 0, y00r, y00i);
      prod_y(x1r, x1i, f, 1, y11r, y11i);
    }
    if (sd.need & 2) {
      prod_y(x1r, x1i, f, 0, y01r, y01i);
      prod_y(x0r, x0i, f, 1, y10r, y10i);
    }
    Env cont;
    for (int c = 0; c < sd.nchild; ++c) {
      const ChildDev cd = uload(P.children + sd.child0 + c);
      Env ce;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if (cd.kind == CK_I) { ce.re[v] = y00r[v] + y11r[v]; ce.im[v] = y00i[v] + y11i[v]; }
        else if (cd.kind == CK_Z) { ce.re[v] = y00r[v] - y11r[v]; ce.im[v] = y00i[v] - y11i[v]; }
        else if (cd.kind == CK_X) { ce.re[v] = y01r[v] + y10r[v]; ce.im[v] = y01i[v] + y10i[v]; }
        else { ce.re[v] = y01i[v] - y10i[v]; ce.im[v] = -y01r[v] + y10r[v]; }
      }
      if (cd.store >= 0) env_store_herm(spill + (size_t)cd.store * GV_ENV, lane, ce);
      if (cd.gleaf >= 0) env_store_herm(gl + (size_t)cd.gleaf * GV_ENV, lane, ce);
      for (int t = 0; t < cd.nterm; ++t) {
        const TermEnd te = uload(P.term_list + cd.term0 + t);
        Env r;
        if (te.rslot >= 0) env_load_herm(spill + (size_t)te.rslot * GV_ENV, lane, r);
        else env_unit(lane, r);
        double sr = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) sr += ce.re[v] * r.re[v] - ce.im[v] * r.im[v];
        acc += te.cre * sr;  // Re(c_k) Re<P_k>
      }
      if (cd.action == -2) cont = ce;
      else if (cd.action >= 0) env_store_herm(spill + (size_t)cd.action * GV_ENV, lane, ce);
    }
    e = cont;
  }
  acc = wave_sum(acc);
  if (lane == 0) parts[inst.part_off + it.y] = acc;
}

// -------------------------------------------------------------- backward
namespace {

// Sum of complex 16x16 products in Gauss form: t1 = sum Ar Br, t2 = sum Ai Bi,
// t3 = sum (Ar + Ai)(Br + Bi); the sum is (t1 - t2) + i (t3 - t1 - t2).
// Three MFMA products per complex product instead of four.
struct CAcc {
  d4 t1, t2, t3;
};

__device__ __forceinline__ void cacc_zero(CAcc &c) {
  c.t1 = (d4){0.0, 0.0, 0.0, 0.0};
  c.t2 = c.t1;
  c.t3 = c.t1;
}

// C += A B for complex 16x16 operands given as MFMA registers: `a` in E layout
// (A operand), `b` in D layout (B operand).  ca / cb conjugate the operand.
__device__ __forceinline__ void cmm(const double *ar, const double *ai, const double *br, const double *bi,
                                    bool ca, bool cb, CAcc &c) {
  const double sa = ca ? -1.0 : 1.0, sb = cb ? -1.0 : 1.0;
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const double ait = sa * ai[t], bit = sb * bi[t];
    c.t1 = mfma(ar[t], br[t], c.t1);
    c.t2 = mfma(ait, bit, c.t2);
    c.t3 = mfma(ar[t] + ait, br[t] + bit, c.t3);
  }
}

__device__ __forceinline__ void cacc_get(const CAcc &c, d4 &dr, d4 &di) {
  dr = c.t1 - c.t2;
  di = c.t3 - c.t1 - c.t2;
}

// Input environment of the step after `step` in its chain: the forward
// transfer of `step`'s continuing child (action -2), computed with the same
// products in the same order as k_chains_gfwd (so bit for bit the same).
__device__ __forceinline__ void advance(const ProbDev &P, const double2 *sites, int step,
                                        int lane, Env &e) {
  const StepDev sd = uload(P.steps + step);
  Frag f;
  frag_load(sites + (size_t)sd.site * GV_SLOT, lane, f, false);
  int kind = CK_I;
  for (int c = 0; c < sd.nchild; ++c) {
    const ChildDev cd = uload(P.children + sd.child0 + c);
    if (cd.action == -2) kind = cd.kind;
  }
  d4 x0r, x0i, x1r, x1i, ar, ai, br, bi;
  prod_x(e, f, 0, x0r, x0i);
  prod_x(e, f, 1, x1r, x1i);
  if (kind == CK_I || kind == CK_Z) {
    prod_y(x0r, x0i, f, 0, ar, ai);  // Y00
    prod_y(x1r, x1i, f, 1, br, bi);  // Y11
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      e.re[v] = kind == CK_I ? ar[v] + br[v] : ar[v] - br[v];
      e.im[v] = kind == CK_I ? ai[v] + bi[v] : ai[v] - bi[v];
    }
  } else {
    prod_y(x1r, x1i, f, 0, ar, ai);  // Y01 = A0^H E A1
    prod_y(x0r, x0i, f, 1, br, bi);  // Y10 = A1^H E A0
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      if (kind == CK_X) {
        e.re[v] = ar[v] + br[v];
        e.im[v] = ai[v] + bi[v];
      } else {
        e.re[v] = ai[v] - bi[v];
        e.im[v] = -ar[v] + br[v];
      }
    }
  }
}

__device__ __forceinline__ void env_zero(Env &e) {
#pragma unroll
  for (int t = 0; t < 4; ++t) e.re[t] = e.im[t] = 0.0;
}

// e += w * conj(q) elementwise (same layout)
__device__ __forceinline__ void add_conj(Env &e, const Env &q, double w) {
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    e.re[t] += w * q.re[t];
    e.im[t] -= w * q.im[t];
  }
}

}  // namespace

// One generation of chains, backward.  For a step with input environment E
// and site matrices M_i (A_i, or A_i^T in the suffix forest) the forward
// computed Y_ji = M_j^H E M_i and children I = Y00+Y11, Z = Y00-Y11,
// X = Y01+Y10, Y = -i Y01 + i Y10 (analytic_batch.cuh:11-14).  With child
// adjoints Cbar: Ybar00 = I+Z, Ybar11 = I-Z, Ybar01 = X + iY, Ybar10 = X - iY
// (Hermitian pairs, Ybar10 = Ybar01^H), and with G_i = sum_j M_j Ybar_ji
//   Ebar   = sum_i G_i M_i^H
//   Mbar_i = 2 E G_i            (E Hermitian)
extern "C" __global__ void __launch_bounds__(256) GV_TWO_WAVES
k_chains_bwd(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
             const ProbDev *__restrict__ probs, const double2 *__restrict__ garena, long long nsite,
             const double2 *__restrict__ spills, double2 *__restrict__ adjspill,
             const double2 *__restrict__ estore, const double2 *__restrict__ gleafs, double2 *__restrict__ aslot) {
  __shared__ double2 tr[4][16][17];
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
  // M_i as the forward pass read it (advance), and its transpose (A-operand form)
  const double2 *sites = garena + (ch.dir != 0 ? nsite : 0) + inst.site_off;
  const double2 *sites_a = garena + (ch.dir != 0 ? 0 : nsite) + inst.site_off;
  const double2 *spill = spills + inst.spill_off;
  double2 *aspill = adjspill + inst.spill_off;
  const double2 *es = estore + inst.ck_off * GV_ENV;
  const double2 *gl = gleafs + inst.gleaf_off * GV_ENV;
  double2 *as = aslot + inst.step_off * GV_SLOT;
  const bool transposed = ch.dir != 0;
  const int a16 = lane & 15, g4 = lane >> 4;

  Env eb;  // adjoint of the environment continuing the chain
  env_zero(eb);
  for (int st = ch.nsteps - 1; st >= 0; --st) {
    const int step = ch.step0 + st;
    const StepDev sd = uload(P.steps + step);
    Env E;
    const int cs = uload(P.ck_slot + step);
    if (cs >= 0) {
      env_load_cm_nt(es + (size_t)cs * GV_ENV, lane, E);
    } else {  // recompute from the nearest checkpoint (step 0 of a chain is always kept)
      int s0 = st;
      while (uload(P.ck_slot + ch.step0 + s0) < 0) --s0;
      env_load_cm_nt(es + (size_t)uload(P.ck_slot + ch.step0 + s0) * GV_ENV, lane, E);
      for (int q = s0; q < st; ++q) advance(P, sites, ch.step0 + q, lane, E);
    }
    Frag fa;  // E layout of M_i (A operand form)
    frag_load(sites_a + (size_t)sd.site * GV_SLOT, lane, fa, false);
    Env y00, y11, y01, y10;
    env_zero(y00);
    env_zero(y11);
    env_zero(y01);
    env_zero(y10);
    for (int c = 0; c < sd.nchild; ++c) {
      const ChildDev cd = uload(P.children + sd.child0 + c);
      Env cb;
      if (cd.action == -2) cb = eb;
      else if (cd.action >= 0) env_load_herm(aspill + (size_t)cd.action * GV_ENV, lane, cb);
      else env_zero(cb);
      // term ends (prefix forest): Cbar += Re(c_k) conj(Rt_k)
      for (int t = 0; t < cd.nterm; ++t) {
        const TermEnd te = uload(P.term_list + cd.term0 + t);
        const double ck = te.cre;
        if (te.rslot >= 0) {
          Env r;
          env_load_herm(spill + (size_t)te.rslot * GV_ENV, lane, r);
          add_conj(cb, r, ck);
        } else if (lane == 0) {
          cb.re[0] += ck;  // unit right environment
        }
      }
      // suffix-forest slot read by prefix term ends: Cbar += Re(c_k) conj(E_left(k))
      if (cd.store >= 0) {
        const int q1 = uload(P.seed_begin + cd.store + 1);
        for (int q = uload(P.seed_begin + cd.store); q < q1; ++q) {
          const SeedTerm sdt = uload(P.seed_terms + q);
          Env le;
          env_load_herm(gl + (size_t)sdt.gleaf * GV_ENV, lane, le);
          add_conj(cb, le, uload(P.coeffs + 2 * sdt.term));
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const double re = cb.re[t], im = cb.im[t];
        if (cd.kind == CK_I) { y00.re[t] += re; y00.im[t] += im; y11.re[t] += re; y11.im[t] += im; }
        else if (cd.kind == CK_Z) { y00.re[t] += re; y00.im[t] += im; y11.re[t] -= re; y11.im[t] -= im; }
        else if (cd.kind == CK_X) { y01.re[t] += re; y01.im[t] += im; y10.re[t] += re; y10.im[t] += im; }
        else {  // Ybar01 += i Cbar, Ybar10 -= i Cbar
          y01.re[t] -= im; y01.im[t] += re; y10.re[t] += im; y10.im[t] -= re;
        }
      }
    }
    // G_i in E layout, computed as G_i^T = sum_j Ybar_ji^T M_j^T with
    // Ybar00^T = conj(Ybar00), Ybar10^T = conj(Ybar01), ...
    d4 g0r, g0i, g1r, g1i;
    {
      CAcc g0, g1;
      cacc_zero(g0);
      cacc_zero(g1);
      if (sd.need & 1) {
        cmm(y00.re, y00.im, fa.re[0], fa.im[0], true, false, g0);
        cmm(y11.re, y11.im, fa.re[1], fa.im[1], true, false, g1);
      }
      if (sd.need & 2) {
        cmm(y01.re, y01.im, fa.re[1], fa.im[1], true, false, g0);
        cmm(y10.re, y10.im, fa.re[0], fa.im[0], true, false, g1);
      }
      cacc_get(g0, g0r, g0i);
      cacc_get(g1, g1r, g1i);
    }
    // Ebar^T = sum_i (G_i M_i^H)^T -> conj gives Ebar in E layout (Hermitian)
    d4 er, ei;
    {
      double t0r[4], t0i[4], t1r[4], t1i[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        t0r[t] = g0r[t]; t0i[t] = g0i[t]; t1r[t] = g1r[t]; t1i[t] = g1i[t];
      }
      CAcc ea;
      cacc_zero(ea);
      cmm(t0r, t0i, fa.re[0], fa.im[0], false, true, ea);
      cmm(t1r, t1i, fa.re[1], fa.im[1], false, true, ea);
      cacc_get(ea, er, ei);
    }
    // Mbar_i = 2 E G_i: G_i to D layout through LDS
    d4 mbr[2], mbi[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const d4 &gr = i ? g1r : g0r;
      const d4 &gi = i ? g1i : g0i;
#pragma unroll
      for (int t = 0; t < 4; ++t) tr[wv][a16][kdx(g4, t)] = make_double2(gr[t], gi[t]);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      double dr[4], di[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double2 q = tr[wv][kdx(g4, v)][a16];
        dr[v] = 2.0 * q.x;
        di[v] = 2.0 * q.y;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      // d is 2 G_i as the B operand and 2 G_i^T as the A operand, and E in
      // its layout is E^T as the B operand: Mbar_i = E (2 G_i), and for suffix
      // chains (whose slot holds the adjoint of A_i^T) Mbar_i^T = (2 G_i^T) E^T.
      // D layout either way, so lanes with consecutive columns store
      // consecutive slot elements.
      d4 mr, mi;
      {
        CAcc ma;
        cacc_zero(ma);
        if (transposed) cmm(dr, di, E.re, E.im, false, false, ma);
        else cmm(E.re, E.im, dr, di, false, false, ma);
        cacc_get(ma, mr, mi);
      }
      mbr[i] = mr;
      mbi[i] = mi;
    }
    // pulled back through the site's trailing CNOT halves (tail_pull), compact;
    // non-temporal, like the checkpoints: read once, by the site reduction
    {
      double2 *slot = as + (size_t)step * GV_SLOT;
      const int tail = uload(P.site_tail + sd.site);
      if ((tail & 3) == 0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          nt_store2(slot + sidx(kdx(g4, v), 0, a16), mbr[0][v], mbi[0][v]);
          nt_store2(slot + sidx(kdx(g4, v), 1, a16), mbr[1][v], mbi[1][v]);
        }
      } else {
        double yr[4], yi[4];
        const TailMap tm = tail_pull(tail, a16, g4, mbr[0], mbr[1], yr);
        tail_pull(tail, a16, g4, mbi[0], mbi[1], yi);
        if (tm.keep) {
#pragma unroll
          for (int v = 0; v < 4; ++v) nt_store2(slot + tail_index(tm, g4, v), yr[v], yi[v]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      eb.re[t] = er[t];
      eb.im[t] = -ei[t];
    }
  }
  if (ch.src >= 0) env_store_herm(aspill + (size_t)ch.src * GV_ENV, lane, eb);
}

// ------------------------------------------------------- site adjoints
// grid (max sites, instances), 256 threads: thread t sums complex elements
// 2t, 2t+1 of the compact pulled-back adjoints (512 >> tail count elements,
// tail_pull) over the site's steps in step order, then the block
// writes the 16x2x16 slot, zero beyond the pulled-back dimensions.
extern "C" __global__ void __launch_bounds__(256)
k_site_reduce(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
              const double2 *__restrict__ aslot, double2 *__restrict__ abar) {
  __shared__ double2 red[512];
  const InstDev inst = uload(insts + blockIdx.y);
  const int s = blockIdx.x;
  if (s >= inst.n) return;
  const ProbDev P = uload(probs + inst.prob);
  const int t = threadIdx.x;
  const int tail = uload(P.site_tail + s);
  const double2 *as = aslot + inst.step_off * GV_SLOT;
  double2 s0 = make_double2(0.0, 0.0), s1 = make_double2(0.0, 0.0);
  if (2 * t < (512 >> (tail & 3))) {
    const int q1 = uload(P.site_steps_begin + s + 1);
    int q = uload(P.site_steps_begin + s);
    // eight slots in flight per thread; the additions keep step order
    for (; q + 8 <= q1; q += 8) {
      double2 a[8], c[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const double2 *src = as + (size_t)uload(P.site_steps + q + u) * GV_SLOT;
        a[u] = src[2 * t];
        c[u] = src[2 * t + 1];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        s0 = cadd(s0, a[u]);
        s1 = cadd(s1, c[u]);
      }
    }
    for (; q < q1; ++q) {
      const double2 *src = as + (size_t)uload(P.site_steps + q) * GV_SLOT;
      s0 = cadd(s0, src[2 * t]);
      s1 = cadd(s1, src[2 * t + 1]);
    }
  }
  red[2 * t] = s0;
  red[2 * t + 1] = s1;
  __syncthreads();
  double2 *dst = abar + inst.site_off + (size_t)s * GV_SLOT;
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int k = tail_compact_index(tail, 2 * t + u);
    dst[2 * t + u] = k >= 0 ? red[k] : make_double2(0.0, 0.0);
  }
}

// Local program in reverse.  Rotation k: T_{k+1} = U T_k, so
// dE/dtheta = Re sum conj(Tbar_{k+1}) (dU/dtheta T_k) and Tbar_k = U^H Tbar_{k+1};
// CNOT halves pull the adjoint back through their index maps.  The adjoint
// in abar has already been pulled back through the site's trailing CNOT
// halves (ProbDev::site_tail) by the backward chains, so the program resumes
// before them.
extern "C" __global__ void __launch_bounds__(64)
k_site_bwd(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs, const double *__restrict__ params,
           const double2 *__restrict__ abar, const double2 *__restrict__ rotin, double *__restrict__ rgrad) {
  __shared__ double2 buf[GV_SLOT];  // one wave: read into registers, barrier, write back
  const int lane = threadIdx.x;
  const InstDev inst = insts[blockIdx.y];
  const int s = blockIdx.x;
  if (s >= inst.n) return;
  const ProbDev &P = probs[inst.prob];
  const double *x = params + inst.param_off;
  const double2 *src = abar + inst.site_off + (size_t)s * GV_SLOT;
#pragma unroll
  for (int k = 0; k < 8; ++k) buf[lane + 64 * k] = src[lane + 64 * k];
  __syncthreads();
  const int o0 = P.lop_begin[s], o1 = P.lop_begin[s + 1] - (P.site_tail[s] & 3);
  for (int o = o1 - 1; o >= o0; --o) {
    const LocalOp op = P.lops[o];
    double2 U[4], dU[4];
    local_matrix(op, x, P.fixes, U, dU);
    const double2 *in = buf;
    double2 w[8];
    if (op.kind == LO_ROT) {
      const double2 *T = rotin + (inst.rot_off + op.rot) * GV_SLOT;
      double part = 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int e = lane + 64 * k;
        const int l = e >> 5, i = (e >> 4) & 1, r = e & 15;
        const double2 d = cadd(cmul(i ? dU[2] : dU[0], T[sidx(l, 0, r)]), cmul(i ? dU[3] : dU[1], T[sidx(l, 1, r)]));
        const double2 tb = in[e];
        part += tb.x * d.x + tb.y * d.y;  // Re(conj(tb) d)
      }
      part = wave_sum(part);
      if (lane == 0) rgrad[inst.rot_off + op.rot] = part;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int e = lane + 64 * k;
      const int l = e >> 5, i = (e >> 4) & 1, r = e & 15;
      double2 v = make_double2(0.0, 0.0);
      switch (op.kind) {
        case LO_CTRL_R:
          if (r < 8) v = in[sidx(l, i, 2 * r + i)];
          break;
        case LO_CTRL_L:
          if (l < 8) v = in[sidx(2 * l + i, i, r)];
          break;
        case LO_TARG_R:
          if (r < 8) v = cadd(in[sidx(l, i, 2 * r)], in[sidx(l, i ^ 1, 2 * r + 1)]);
          break;
        case LO_TARG_L:
          if (l < 8) v = cadd(in[sidx(2 * l, i, r)], in[sidx(2 * l + 1, i ^ 1, r)]);
          break;
        default:  // Tbar_k[l,j,r] = sum_i conj(U[i][j]) Tbar[l,i,r]
          v = cadd(cmul(cconj(i ? U[1] : U[0]), in[sidx(l, 0, r)]), cmul(cconj(i ? U[3] : U[2]), in[sidx(l, 1, r)]));
      }
      w[k] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) buf[lane + 64 * k] = w[k];
    __syncthreads();
  }
}

// dE/dx_p = sum over p's rotations of scale * dE/dangle, in rotation order.
extern "C" __global__ void __launch_bounds__(256)
k_param_grad(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs, const double *__restrict__ rgrad,
             double *__restrict__ grads) {
  const InstDev inst = insts[blockIdx.y];
  const ProbDev &P = probs[inst.prob];
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= P.nparams) return;
  double g = 0.0;
  for (int q = P.param_rot_begin[p]; q < P.param_rot_begin[p + 1]; ++q) {
    const int r = P.param_rots[q];
    g += P.rot_scale[r] * rgrad[inst.rot_off + r];
  }
  grads[inst.param_off + p] = g;
}

// ------------------------------------------------------------ host launchers
namespace gv {

hipError_t kl_build_local(int ninst, int max_n, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                          const double *params, double2 *garena, double2 *rotin, int *status) {
  hipLaunchKernelGGL(k_build_local, dim3(max_n, ninst), dim3(64), 0, s, insts, probs, params, garena, rotin,
                     status);
  return hipGetLastError();
}

hipError_t kl_chains_gfwd(int nitems, int wpb, hipStream_t s, const int2 *items, const InstDev *insts, const ProbDev *probs,
                          const double2 *garena, long long nsite, double2 *spills, double *parts, double2 *estore,
                          double2 *gleafs) {
  hipLaunchKernelGGL(k_chains_gfwd, dim3((nitems + wpb - 1) / wpb), dim3(64 * wpb), 0, s, items, nitems, insts, probs, garena,
                     nsite, spills, parts, estore, gleafs);
  return hipGetLastError();
}

hipError_t kl_chains_bwd(int nitems, int wpb, hipStream_t s, const int2 *items, const InstDev *insts, const ProbDev *probs,
                         const double2 *garena, long long nsite, const double2 *spills, double2 *adjspill,
                         const double2 *estore, const double2 *gleafs, double2 *aslot) {
  hipLaunchKernelGGL(k_chains_bwd, dim3((nitems + wpb - 1) / wpb), dim3(64 * wpb), 0, s, items, nitems, insts, probs, garena,
                     nsite, spills, adjspill, estore, gleafs, aslot);
  return hipGetLastError();
}

hipError_t kl_site_reduce(int ninst, int max_n, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                          const double2 *aslot, double2 *abar) {
  hipLaunchKernelGGL(k_site_reduce, dim3(max_n, ninst), dim3(256), 0, s, insts, probs, aslot, abar);
  return hipGetLastError();
}

hipError_t kl_site_bwd(int ninst, int max_n, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                       const double *params, const double2 *abar, const double2 *rotin, double *rgrad) {
  hipLaunchKernelGGL(k_site_bwd, dim3(max_n, ninst), dim3(64), 0, s, insts, probs, params, abar, rotin, rgrad);
  return hipGetLastError();
}

hipError_t kl_param_grad(int ninst, int max_params, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                         const double *rgrad, double *grads) {
  if (max_params <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_param_grad, dim3((max_params + 255) / 256, ninst), dim3(256), 0, s, insts, probs, rgrad,
                     grads);
  return hipGetLastError();
}

}  // namespace gv
