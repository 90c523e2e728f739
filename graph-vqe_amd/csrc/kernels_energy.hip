// State preparation and analytic-energy kernels (gfx950).
//
//   k_prep    one workgroup per MPS instance runs the compiled site program:
//             single-site gates are fused into per-site pending 2x2 matrices,
//             a two-site gate builds Theta (2Dl x 2Dr) in LDS and splits it by
//             a one-sided row Jacobi SVD held in LDS (complex FP64), with the
//             reference rank rule and zero-cutoff guard.
//   k_envs    one wavefront per (instance, direction): identity left
//             environments L_s and transposed right environments Rt_s.
//   k_chains  one wavefront per (instance, chain) of the prefix forest: the
//             environment stays in MFMA registers along the chain; children
//             that start other chains are spilled to HBM.
//   k_reduce  one workgroup per instance: E = Re sum_k c_k <P_k>, compensated
//             and in a fixed order.
#include "device_common.h"

using namespace gv;

#define PREP_THREADS 256
#define WSTR 64          // row stride of the Theta|P panel (nc + m <= 64)
#define MAX_SWEEPS 40

struct PrepShared {
  double2 pend[GV_NMAX][4];
  int pflag[GV_NMAX];
  int bond[GV_NMAX + 1];
  double2 A[GV_SLOT];
  double2 B[GV_SLOT];
  double2 W[32][WSTR];
  double2 G[16];
  double nrm[32];
  int perm[32];
  int k;
  double frob;
};

__device__ __forceinline__ void rr_pair(int r, int i, int m, int &p, int &q) {
  // round-robin tournament over m (even) rows: m-1 rounds of m/2 disjoint pairs
  const int M1 = m - 1;
  p = (r + i) % M1;
  q = (i == 0) ? M1 : (r - i + M1) % M1;
}

__device__ __forceinline__ double2 mat2_apply_row(const double2 *U, int i, double2 a0, double2 a1) {
  return cadd(cmul(U[i * 2 + 0], a0), cmul(U[i * 2 + 1], a1));
}

// One-sided Jacobi on the rows of W[0..m)[0..nc): orthogonalise rows with
// unitary 2x2 row rotations applied to the whole panel (the P block starts as
// the identity and accumulates Q^H).  Theta = Q * W_final on exit.
__device__ void row_jacobi(PrepShared &sh, int m, int nc, int tid) {
  const int width = nc + m;
  const int grp = tid >> 4, sub = tid & 15;
  const int npairs = m >> 1;
  const double tol = 1e-14;
  const double abs_thr = 1e-28 * sh.frob;
  for (int sweep = 0; sweep < MAX_SWEEPS; ++sweep) {
    int rotated = 0;
    for (int round = 0; round < m - 1; ++round) {
      if (grp < npairs) {
        int p, q;
        rr_pair(round, grp, m, p, q);
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        for (int x = sub; x < nc; x += 16) {
          const double2 a = sh.W[p][x], b = sh.W[q][x];
          al += a.x * a.x + a.y * a.y;
          be += b.x * b.x + b.y * b.y;
          gr += a.x * b.x + a.y * b.y;   // gamma = sum a * conj(b)
          gi += a.y * b.x - a.x * b.y;
        }
#pragma unroll
        for (int off = 8; off >= 1; off >>= 1) {
          al += __shfl_xor(al, off, 16);
          be += __shfl_xor(be, off, 16);
          gr += __shfl_xor(gr, off, 16);
          gi += __shfl_xor(gi, off, 16);
        }
        const double g = sqrt(gr * gr + gi * gi);
        if (g > tol * sqrt(al * be) && g > abs_thr) {
          rotated = 1;
          const double zeta = (be - al) / (2.0 * g);
          const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
          const double er = gr / g, ei = gi / g;
          for (int x = sub; x < width; x += 16) {
            const double2 a = sh.W[p][x], b = sh.W[q][x];
            const double ebr = er * b.x - ei * b.y, ebi = er * b.y + ei * b.x;
            sh.W[p][x] = make_double2(c * a.x - s * ebr, c * a.y - s * ebi);
            sh.W[q][x] = make_double2(s * a.x + c * ebr, s * a.y + c * ebi);
          }
        }
      }
      __syncthreads();
    }
    if (!__syncthreads_or(rotated)) break;
  }
}

__device__ void apply_pending_to_global(PrepShared &sh, double2 *slot, int s, int tid) {
  // slot (l,i,r): new[l,i,r] = sum_i' U[i][i'] old[l,i',r]; thread handles (l, r) pairs
  for (int idx = tid; idx < GV_CHI * GV_CHI; idx += PREP_THREADS) {
    const int l = idx >> 4, r = idx & 15;
    const double2 a0 = slot[(l * 2 + 0) * 16 + r], a1 = slot[(l * 2 + 1) * 16 + r];
    slot[(l * 2 + 0) * 16 + r] = mat2_apply_row(sh.pend[s], 0, a0, a1);
    slot[(l * 2 + 1) * 16 + r] = mat2_apply_row(sh.pend[s], 1, a0, a1);
  }
}

extern "C" __global__ void __launch_bounds__(PREP_THREADS)
k_prep(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs, const double *__restrict__ params,
       double2 *__restrict__ arena, int *__restrict__ bonds_out, int *__restrict__ status) {
  __shared__ PrepShared sh;
  const int tid = threadIdx.x;
  const InstDev inst = insts[blockIdx.x];
  const ProbDev P = probs[inst.prob];
  const int n = P.n;
  double2 *sites = arena + inst.site_off;
  const double *x = params + inst.param_off;

  // |0...0>: every slot = [1,0] at (0,0,0), zero padded.
  for (int e = tid; e < n * GV_SLOT; e += PREP_THREADS)
    sites[e] = make_double2((e % GV_SLOT) == 0 ? 1.0 : 0.0, 0.0);
  for (int s = tid; s < n; s += PREP_THREADS) {
    sh.pflag[s] = 0;
    sh.pend[s][0] = make_double2(1, 0);
    sh.pend[s][1] = make_double2(0, 0);
    sh.pend[s][2] = make_double2(0, 0);
    sh.pend[s][3] = make_double2(1, 0);
  }
  for (int s = tid; s <= n; s += PREP_THREADS) sh.bond[s] = 1;
  if (tid == 0 && status) status[blockIdx.x] = ST_OK;
  __syncthreads();

  int op_i = 0;
  while (op_i < P.nprog) {
    // Single-site gates up to the next two-site gate: fused by thread 0.
    if (tid == 0) {
      while (op_i < P.nprog && P.prog[op_i].kind != OP_G2) {
        const ProgOp o = P.prog[op_i];
        double2 U[4];
        if (o.kind == OP_ROT) {
          const double theta = (o.param >= 0 ? o.scale * x[o.param] : 0.0) + o.offset;
          double sn, cs;
          sincos(0.5 * theta, &sn, &cs);
          // R_sigma(theta) = c I - i s sigma: RX = [[c, -is], [-is, c]],
          // RY = [[c, -s], [s, c]], RZ = diag(c - is, c + is).  Written
          // branch-free: a three-way branch here lost the imaginary part of
          // RZ's first entry in this kernel's gfx950 code.
          const double fx = o.arg == 0 ? 1.0 : 0.0, fy = o.arg == 1 ? 1.0 : 0.0, fz = o.arg == 2 ? 1.0 : 0.0;
          U[0] = make_double2(cs, -sn * fz);
          U[1] = make_double2(-sn * fy, -sn * fx);
          U[2] = make_double2(sn * fy, -sn * fx);
          U[3] = make_double2(cs, sn * fz);
        } else {
          const double *F = P.fixes + 8 * o.arg;
          for (int e = 0; e < 4; ++e) U[e] = make_double2(F[2 * e], F[2 * e + 1]);
        }
        double2 *M = sh.pend[o.site];
        const double2 m0 = M[0], m1 = M[1], m2 = M[2], m3 = M[3];
        M[0] = cadd(cmul(U[0], m0), cmul(U[1], m2));
        M[1] = cadd(cmul(U[0], m1), cmul(U[1], m3));
        M[2] = cadd(cmul(U[2], m0), cmul(U[3], m2));
        M[3] = cadd(cmul(U[2], m1), cmul(U[3], m3));
        sh.pflag[o.site] = 1;
        ++op_i;
      }
    }
    // all threads learn op_i from the program (same scan, no divergence on data)
    while (op_i < P.nprog && P.prog[op_i].kind != OP_G2) ++op_i;
    __syncthreads();
    if (op_i >= P.nprog) break;

    // ---- two-site gate on (s, s+1)
    const ProgOp o = P.prog[op_i];
    ++op_i;
    const int s = o.site;
    const int Dl = sh.bond[s], Dc = sh.bond[s + 1], Dr = sh.bond[s + 2];
    const int m = 2 * Dl, nc = 2 * Dr;
    const double2 *ga = sites + (size_t)s * GV_SLOT;
    const double2 *gb = sites + (size_t)(s + 1) * GV_SLOT;
    for (int idx = tid; idx < GV_CHI * GV_CHI; idx += PREP_THREADS) {
      const int l = idx >> 4, r = idx & 15;
      const double2 a0 = ga[(l * 2) * 16 + r], a1 = ga[(l * 2 + 1) * 16 + r];
      const double2 b0 = gb[(l * 2) * 16 + r], b1 = gb[(l * 2 + 1) * 16 + r];
      sh.A[(l * 2) * 16 + r] = mat2_apply_row(sh.pend[s], 0, a0, a1);
      sh.A[(l * 2 + 1) * 16 + r] = mat2_apply_row(sh.pend[s], 1, a0, a1);
      sh.B[(l * 2) * 16 + r] = mat2_apply_row(sh.pend[s + 1], 0, b0, b1);
      sh.B[(l * 2 + 1) * 16 + r] = mat2_apply_row(sh.pend[s + 1], 1, b0, b1);
    }
    if (tid < 16) sh.G[tid] = make_double2(P.gates[32 * o.arg + 2 * tid], P.gates[32 * o.arg + 2 * tid + 1]);
    __syncthreads();
    if (tid == 0) {
      for (int e = 0; e < 4; ++e) {
        sh.pend[s][e] = make_double2((e == 0 || e == 3) ? 1.0 : 0.0, 0.0);
        sh.pend[s + 1][e] = make_double2((e == 0 || e == 3) ? 1.0 : 0.0, 0.0);
      }
      sh.pflag[s] = 0;
      sh.pflag[s + 1] = 0;
    }
    // Theta[(l,I),(J,r)] = sum_{ij} G[(I,J),(i,j)] sum_c A[l,i,c] B[c,j,r]
    // (k_buildM, sharded_mps.cuh:30-45), and the P block = identity.
    for (int idx = tid; idx < m * WSTR; idx += PREP_THREADS) {
      const int row = idx / WSTR, col = idx % WSTR;
      double2 v = make_double2(0.0, 0.0);
      if (col < nc) {
        const int l = row >> 1, I = row & 1;
        const int J = col / Dr, r = col % Dr;
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) {
            const double2 gv = sh.G[(I * 2 + J) * 4 + (i * 2 + j)];
            if (gv.x == 0.0 && gv.y == 0.0) continue;
            double2 acc = make_double2(0.0, 0.0);
            for (int c = 0; c < Dc; ++c)
              acc = cadd(acc, cmul(sh.A[(l * 2 + i) * 16 + c], sh.B[(c * 2 + j) * 16 + r]));
            v = cadd(v, cmul(gv, acc));
          }
      } else if (col < nc + m) {
        v = make_double2((col - nc) == row ? 1.0 : 0.0, 0.0);
      }
      sh.W[row][col] = v;
    }
    __syncthreads();
    if (tid < 64) {
      double f = 0.0;
      for (int idx = tid; idx < m * nc; idx += 64) {
        const double2 v = sh.W[idx / nc][idx % nc];
        f += v.x * v.x + v.y * v.y;
      }
      f = wave_sum(f);
      if (tid == 0) sh.frob = f;
    }
    __syncthreads();

    row_jacobi(sh, m, nc, tid);

    // singular values = row norms; sort descending (ties by row index)
    for (int p = tid; p < m; p += PREP_THREADS) {
      double f = 0.0;
      for (int c = 0; c < nc; ++c) {
        const double2 v = sh.W[p][c];
        f += v.x * v.x + v.y * v.y;
      }
      sh.nrm[p] = f;
    }
    __syncthreads();
    if (tid < m) {
      const double mine = sh.nrm[tid];
      int rank = 0;
      for (int q = 0; q < m; ++q) {
        const double o2 = sh.nrm[q];
        rank += (o2 > mine) || (o2 == mine && q < tid);
      }
      sh.perm[rank] = tid;
    }
    __syncthreads();
    if (tid == 0) {
      // selectDirectSvdRank, dist_svd.cuh:194-212
      const int count = min(m, nc);
      const double s0 = sqrt(sh.nrm[sh.perm[0]]);
      const double thr = fmax(P.abs_cut, P.rel_cut * s0);
      const bool zero_cut = (P.abs_cut == 0.0 && P.rel_cut == 0.0);
      int selected = 0, k = 0;
      bool bad = false;
      for (int i = 0; i < count; ++i) {
        const double si = sqrt(sh.nrm[sh.perm[i]]);
        if (!isfinite(si)) bad = true;
        if (zero_cut || si > thr) ++selected;
      }
      if (bad) {
        if (status) atomicMax(status + blockIdx.x, ST_NONFINITE);
        k = 1;
      } else if (zero_cut && P.max_bond < count &&
                 sqrt(sh.nrm[sh.perm[P.max_bond]]) > 64.0 * 2.2204460492503131e-16 * count * s0) {
        if (status) atomicMax(status + blockIdx.x, ST_GUARD);
        k = min(P.max_bond, count);
      } else {
        k = max(1, min(P.max_bond, selected));
      }
      sh.k = min(k, GV_CHI);
    }
    __syncthreads();
    const int k = sh.k;
    // left = Q[:, perm[:k]] = conj(P rows);  right = W[perm[:k], :nc]
    double2 *wa = sites + (size_t)s * GV_SLOT;
    double2 *wb = sites + (size_t)(s + 1) * GV_SLOT;
    for (int idx = tid; idx < GV_SLOT; idx += PREP_THREADS) {
      const int row = idx >> 5, I = (idx >> 4) & 1, c = idx & 15;
      double2 va = make_double2(0.0, 0.0), vb = make_double2(0.0, 0.0);
      if (row < Dl && c < k) va = cconj(sh.W[sh.perm[c]][nc + row * 2 + I]);
      // right factor: idx = (c'*2 + J)*16 + r with c' = row
      if (row < k && c < Dr) vb = sh.W[sh.perm[row]][I * Dr + c];
      wa[idx] = va;
      wb[idx] = vb;
    }
    if (tid == 0) sh.bond[s + 1] = k;
    __syncthreads();
  }
  // flush pending single-site gates that no two-site gate absorbed
  for (int s = 0; s < n; ++s) {
    if (sh.pflag[s]) apply_pending_to_global(sh, sites + (size_t)s * GV_SLOT, s, tid);
  }
  __syncthreads();
  if (bonds_out)
    for (int s = tid; s <= n; s += PREP_THREADS) bonds_out[(size_t)blockIdx.x * (GV_NMAX + 1) + s] = sh.bond[s];
}

// Identity environments.  Wave w: instance w/2, direction w&1 (0 = L, 1 = Rt).
//   L_0 = [1],  L_{s+1} = sum_i A_i^H L_s A_i
//   Rt_n = [1], Rt_s = sum_i (A_i^T)^H Rt_{s+1} A_i^T  (Rt = R^T, R as in
//   prepareSamplingRightEnvironments, sharded_mps.cuh:61-90, 814-836)
extern "C" __global__ void __launch_bounds__(256)
k_envs(const InstDev *__restrict__ insts, int ninst, const double2 *__restrict__ arena, double2 *__restrict__ envs) {
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= 2 * ninst) return;
  const InstDev inst = insts[w >> 1];
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

// One generation of prefix chains.  Work item (instance, chain) per wave.
extern "C" __global__ void __launch_bounds__(256)
k_chains(const int2 *__restrict__ items, int nitems, const InstDev *__restrict__ insts,
         const ProbDev *__restrict__ probs, const double2 *__restrict__ arena,
         const double2 *__restrict__ envs, double2 *__restrict__ spills, double2 *__restrict__ terms) {
  const int lane = threadIdx.x & 63;
  // XCD-aware block order: hardware block b runs on XCD b % 8; give each XCD
  // a contiguous range of work items so one instance's chains share an L2.
  const int nb = gridDim.x, b = blockIdx.x;
  const int q8 = nb >> 3, r8 = nb & 7, xcd = b & 7;
  const int logical = xcd * q8 + min(xcd, r8) + (b >> 3);
  const int w = logical * 4 + (threadIdx.x >> 6);
  if (w >= nitems) return;
  const int2 it = items[w];
  const InstDev inst = insts[it.x];
  const ProbDev P = probs[inst.prob];
  const ChainDev ch = P.chains[it.y];
  const int n = inst.n;
  const double2 *sites = arena + inst.site_off;
  const double2 *L = envs + inst.env_off;
  const double2 *Rt = L + (size_t)(n + 1) * GV_ENV;
  double2 *spill = spills + inst.spill_off;
  double2 *tv = terms + inst.term_off;

  Env e;
  if (ch.src < 0) env_load(L + (size_t)ch.site0 * GV_ENV, lane, e);
  else env_load(spill + (size_t)ch.src * GV_ENV, lane, e);

  Frag f;
  for (int st = 0; st < ch.nsteps; ++st) {
    const StepDev sd = P.steps[ch.step0 + st];
    frag_load(sites + (size_t)sd.site * GV_SLOT, lane, f, false);
    d4 x0r, x0i, x1r, x1i;
    prod_x(e, f, 0, x0r, x0i);
    prod_x(e, f, 1, x1r, x1i);
    d4 y00r, y00i, y11r, y11i, y01r, y01i, y10r, y10i;
    if (sd.need & 1) {
      prod_y(x0r, x0i, f, 0, y00r, y00i);
      prod_y(x1r, x1i, f, 1, y11r, y11i);
    }
    if (sd.need & 2) {
      prod_y(x1r, x1i, f, 0, y01r, y01i);  // A_0^H E A_1
      prod_y(x0r, x0i, f, 1, y10r, y10i);  // A_1^H E A_0
    }
    Env cont;
    for (int c = 0; c < sd.nchild; ++c) {
      const ChildDev cd = P.children[sd.child0 + c];
      Env ce;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // analyticLeft phases, analytic_batch.cuh:11-14
        if (cd.kind == CK_I) { ce.re[v] = y00r[v] + y11r[v]; ce.im[v] = y00i[v] + y11i[v]; }
        else if (cd.kind == CK_Z) { ce.re[v] = y00r[v] - y11r[v]; ce.im[v] = y00i[v] - y11i[v]; }
        else if (cd.kind == CK_X) { ce.re[v] = y01r[v] + y10r[v]; ce.im[v] = y01i[v] + y10i[v]; }
        else { ce.re[v] = y01i[v] - y10i[v]; ce.im[v] = -y01r[v] + y10r[v]; }  // -i Y01 + i Y10
      }
      if (cd.nterm > 0) {
        Env r;
        env_load(Rt + (size_t)(sd.site + 1) * GV_ENV, lane, r);
        double sr = 0.0, si = 0.0;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          sr += ce.re[v] * r.re[v] - ce.im[v] * r.im[v];
          si += ce.re[v] * r.im[v] + ce.im[v] * r.re[v];
        }
        sr = wave_sum(sr);
        si = wave_sum(si);
        for (int t = lane; t < cd.nterm; t += 64) tv[P.term_list[cd.term0 + t]] = make_double2(sr, si);
      }
      if (cd.action == -2) cont = ce;
      else if (cd.action >= 0) env_store(spill + (size_t)cd.action * GV_ENV, lane, ce);
    }
    e = cont;
  }
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

// ------------------------------------------------------------ host launchers
namespace gv {

hipError_t kl_prep(int ninst, hipStream_t s, const InstDev *insts, const ProbDev *probs, const double *params,
                   double2 *arena, int *bonds, int *status) {
  hipLaunchKernelGGL(k_prep, dim3(ninst), dim3(PREP_THREADS), 0, s, insts, probs, params, arena, bonds, status);
  return hipGetLastError();
}

hipError_t kl_envs(int ninst, hipStream_t s, const InstDev *insts, const double2 *arena, double2 *envs) {
  hipLaunchKernelGGL(k_envs, dim3((2 * ninst + 3) / 4), dim3(256), 0, s, insts, ninst, arena, envs);
  return hipGetLastError();
}

hipError_t kl_chains(int nitems, hipStream_t s, const int2 *items, const InstDev *insts, const ProbDev *probs,
                     const double2 *arena, const double2 *envs, double2 *spills, double2 *terms) {
  hipLaunchKernelGGL(k_chains, dim3((nitems + 3) / 4), dim3(256), 0, s, items, nitems, insts, probs, arena, envs,
                     spills, terms);
  return hipGetLastError();
}

hipError_t kl_reduce(int ninst, hipStream_t s, const InstDev *insts, const ProbDev *probs, double2 *terms,
                     double *energies) {
  hipLaunchKernelGGL(k_reduce, dim3(ninst), dim3(256), 0, s, insts, probs, terms, energies);
  return hipGetLastError();
}

}  // namespace gv
