// State preparation (gfx950): the compiled gate stream of one MPS instance per
// workgroup, with up to GV_WAVES independent two-site gates in flight.
//
// Schedule: the host groups the two-site gates into time steps of gates on
// disjoint sites whose predecessors are all in earlier steps; wavefront w of
// the workgroup takes slot w of each step, and a workgroup barrier separates
// steps.  For a 4-layer CX chain this runs the layers as a diagonal wavefront
// (n-1 + 2*3 steps instead of 4(n-1) sequential gates).
//
// One gate = one wavefront:
//   1. fold the pending single-site ops of both sites into 2x2 matrices and
//      apply them to the loaded site tensors;
//   2. C_ij = A_i B_j on FP64 MFMA (16x16x16 complex, 4 products), then
//      Theta[(l,I),(J,r)] = sum_ij G[(I,J),(i,j)] C_ij[l][r]  (k_buildM,
//      sharded_mps.cuh:30-45) into an LDS panel [Theta | I_m];
//   3. one-sided Jacobi on the rows of Theta with unitary 2x2 row rotations
//      applied to the whole panel (4 lanes per row pair, quad-DPP sums), so
//      Theta = Q W with Q = (right block)^H unitary and W's rows orthogonal;
//   4. singular values = row norms, sorted; the reference rank rule and
//      zero-cutoff guard (selectDirectSvdRank, dist_svd.cuh:194-212);
//      left site <- Q[:, :k] (isometry), right site <- W[:k, :] = S_k Vh_k
//      (packDirectSvdFactors, dist_svd.cuh:214-231).
// When the problem's cutoffs are zero and no gate of the tape is cut below
// min(2 Dl, 2 Dr) (exact_split), steps 3-4 are replaced by the exact split
// Theta = I Theta or Theta I (run_gate), unless the batch asks for the
// factorisation at every gate.
#include "device_common.h"

using namespace gv;

#define PW 65          // panel row stride in complex elements: nc + m <= 64, padded by one
                       // so that the 16 row pairs of a round start in different LDS banks
#define MAX_SWEEPS 40

// Pending single-site matrices of the first GV_PEND_LDS two-site gates,
// computed by all threads of the workgroup before the gate schedule starts
// (their rotations' sines and cosines are then off the schedule's dependent
// path); later gates of longer tapes fold theirs when they run.
#define GV_PEND_LDS 192

struct PrepLds {
  double2 pend[GV_PEND_LDS][8];  // Ul (4), Ur (4) of gate gi
  double2 W[GV_WAVES][32][PW];
  double nrm[GV_WAVES][32];
  int perm[GV_WAVES][32];
  double2 vb[GV_WAVES][32];   // Householder vector broadcast
  int bond[GV_NMAX + 1];
  int sweeps;   // Jacobi sweeps summed over the instance's gates (statistics)
  unsigned int cyc[4];  // core clocks in Theta build, QR, Jacobi, rank rule + write-back
};

__device__ __forceinline__ unsigned long long clk() { return __builtin_amdgcn_s_memtime(); }

__device__ __forceinline__ double dpp_quad(double v, int ctrl_sel) {
  const long long b = __double_as_longlong(v);
  int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
  if (ctrl_sel == 1) {
    lo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    hi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xF, 0xF, false);
  }
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double quad_sum(double v) {
  v += dpp_quad(v, 1);
  v += dpp_quad(v, 2);
  return v;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void rr_pair(int r, int i, int m, int &p, int &q) {
  // round-robin tournament over m (even) rows: m-1 rounds of m/2 disjoint pairs
  const int M1 = m - 1;
  p = (r + i) % M1;
  q = (i == 0) ? M1 : (r - i + M1) % M1;
}

// 2x2 matrix of a single-site op.  R_sigma(theta) = c I - i s sigma
// (RX = [[c, -is], [-is, c]], RY = [[c, -s], [s, c]], RZ = diag(c-is, c+is);
// test/tools/reverse_mps.py:52-55), written branch-free: a three-way branch
// here once lost the imaginary part of RZ's first entry in gfx950 code.
__device__ __forceinline__ void op_matrix(const ProgOp &o, const double *x, const double *fixes, double2 U[4]) {
  if (o.kind == OP_ROT) {
    const double theta = (o.param >= 0 ? o.scale * uload(x + o.param) : 0.0) + o.offset;
    double sn, cs;
    sincos(0.5 * theta, &sn, &cs);
    const double fx = o.arg == 0 ? 1.0 : 0.0, fy = o.arg == 1 ? 1.0 : 0.0, fz = o.arg == 2 ? 1.0 : 0.0;
    U[0] = make_double2(cs, -sn * fz);
    U[1] = make_double2(-sn * fy, -sn * fx);
    U[2] = make_double2(sn * fy, -sn * fx);
    U[3] = make_double2(cs, sn * fz);
  } else {
    const double *F = fixes + 8 * o.arg;
#pragma unroll
    for (int e = 0; e < 4; ++e) U[e] = uload((const double2 *)F + e);
  }
}

__device__ __forceinline__ void fold(double2 M[4], const double2 U[4]) {  // M <- U M
  const double2 m0 = M[0], m1 = M[1], m2 = M[2], m3 = M[3];
  M[0] = cadd(cmul(U[0], m0), cmul(U[1], m2));
  M[1] = cadd(cmul(U[0], m1), cmul(U[1], m3));
  M[2] = cadd(cmul(U[2], m0), cmul(U[3], m2));
  M[3] = cadd(cmul(U[2], m1), cmul(U[3], m3));
}

__device__ __forceinline__ void pending_matrix(const ProgOp *ops, int count, const double *x, const double *fixes,
                                               double2 M[4]) {
  M[0] = make_double2(1.0, 0.0);
  M[1] = make_double2(0.0, 0.0);
  M[2] = make_double2(0.0, 0.0);
  M[3] = make_double2(1.0, 0.0);
  for (int k = 0; k < count; ++k) {
    double2 U[4];
    op_matrix(uload(ops + k), x, fixes, U);
    fold(M, U);
  }
}

__device__ __forceinline__ double group_sum(double v, int lpp) {
  v = quad_sum(v);
  if (lpp == 8) v += __shfl_xor(v, 4, 64);
  return v;
}

template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_max_u64(unsigned long long v) {
  const int lo = (int)(unsigned)v, hi = (int)(v >> 32);
  const unsigned olo = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  const unsigned ohi = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  const unsigned long long o = ((unsigned long long)ohi << 32) | olo;
  return o > v ? o : v;
}

__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int l) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(v >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
  return __longlong_as_double((long long)readlane_u64((unsigned long long)__double_as_longlong(v), l));
}

template <int CTRL>
__device__ __forceinline__ double dpp_add_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = (int)(b & 0xffffffffLL), hi = (int)(b >> 32);
  const unsigned olo = (unsigned)__builtin_amdgcn_update_dpp(lo, lo, CTRL, 0xF, 0xF, false);
  const unsigned ohi = (unsigned)__builtin_amdgcn_update_dpp(hi, hi, CTRL, 0xF, 0xF, false);
  return v + __longlong_as_double((long long)(((unsigned long long)ohi << 32) | olo));
}

// Wave-uniform sum (fixed order): DPP within each row of 16 lanes, then the
// four row sums by readlane.
__device__ __forceinline__ double wave_sum_uniform(double v) {
  v = dpp_add_f64<0xB1>(v);   // quad_perm [1,0,3,2]
  v = dpp_add_f64<0x4E>(v);   // quad_perm [2,3,0,1]
  v = dpp_add_f64<0x141>(v);  // row_half_mirror
  v = dpp_add_f64<0x140>(v);  // row_mirror
  return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Wave-uniform maximum of a 64-bit key: DPP within each row of 16 lanes
// (quad swaps, half-row and row mirrors), then the four row maxima by readlane.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long k) {
  k = dpp_max_u64<0xB1>(k);   // quad_perm [1,0,3,2]
  k = dpp_max_u64<0x4E>(k);   // quad_perm [2,3,0,1]
  k = dpp_max_u64<0x141>(k);  // row_half_mirror
  k = dpp_max_u64<0x140>(k);  // row_mirror
  const unsigned long long a = readlane_u64(k, 0), b = readlane_u64(k, 16);
  const unsigned long long c = readlane_u64(k, 32), d = readlane_u64(k, 48);
  const unsigned long long ab = a > b ? a : b, cd = c > d ? c : d;
  return ab > cd ? ab : cd;
}

// One Householder step j of the pivoted QR, j in [R0, R0 + 4).  Rows < R0 are
// final and not touched; rows R0..j-1 of the block are masked.  Returns false
// when the Theta block is exhausted (every remaining column norm is zero).
// REAL: the panel is real (imaginary parts zero and not carried).
template <bool REAL, int R0>
__device__ __forceinline__ bool qr_step(double2 (&col)[32], double2 *vb, int j, int nc, int lane, double &piv0) {
  // squared norm of rows >= j of this lane's column (four independent sums)
  double n4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = R0; r < 32; ++r) {
    double2 c = col[r];
    if (r < R0 + 4 && r < j) c = make_double2(0.0, 0.0);
    if constexpr (REAL) n4[r & 3] = fma(c.x, c.x, n4[r & 3]);
    else n4[r & 3] = fma(c.x, c.x, fma(c.y, c.y, n4[r & 3]));
  }
  const double nrm = (n4[0] + n4[1]) + (n4[2] + n4[3]);
  // pivot: the largest norm, ties to the lowest column.  Norms are >= 0, so
  // their bit patterns order as unsigned integers; the low 6 bits carry the
  // column (63 - lane) instead.
  const unsigned long long key =
      lane < nc ? ((unsigned long long)__double_as_longlong(nrm) & ~63ULL) | (unsigned long long)(63 - lane) : 0ULL;
  const unsigned long long best = wave_max_u64(key);
  if ((best >> 6) == 0) return false;  // uniform
  const int bi = 63 - (int)(best & 63);
  const double a2 = readlane_f64(nrm, bi);  // exact squared norm of the pivot column
  if (j == 0) piv0 = a2;
  double2 xl = col[R0];
#pragma unroll
  for (int r = R0 + 1; r < R0 + 4; ++r)
    if (r == j) xl = col[r];
  const double2 xj = make_double2(readlane_f64(xl.x, bi), REAL ? 0.0 : readlane_f64(xl.y, bi));
  // v = x + e |x| e_j, e = x_j / |x_j|;  tau = 2 / (v^H v) = 1 / (|x| (|x| + |x_j|))
  const double alpha = sqrt(a2);
  // REAL: |x_j| = fabs(x_j) and e = sign(x_j) (e = 1 at zero), no divisions
  const double axj = REAL ? fabs(xj.x) : sqrt(xj.x * xj.x + xj.y * xj.y);
  const double2 e = REAL ? make_double2(xj.x < 0.0 ? -1.0 : 1.0, 0.0)
                         : (axj > 0.0 ? make_double2(xj.x / axj, xj.y / axj) : make_double2(1.0, 0.0));
  const double tau = 1.0 / (alpha * (alpha + axj));
  if constexpr (REAL) {
    // the reflector is the pivot column: broadcast it by readlane (SGPRs),
    // then dot and update take it as a scalar operand; no LDS round trip
    double v[32 - R0];
#pragma unroll
    for (int r = R0; r < 32; ++r) {
      v[r - R0] = readlane_f64(col[r].x, bi);
      if (r < R0 + 4) {
        if (r < j) v[r - R0] = 0.0;
        if (r == j) v[r - R0] = xj.x + e.x * alpha;
      }
    }
    double dr[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int r = R0; r < 32; ++r) dr[r & 3] = fma(v[r - R0], col[r].x, dr[r & 3]);
    // 1 / (alpha (alpha + |x_j|)) by the hardware reciprocal and two Newton
    // steps (within an ulp of the quotient; the division's scaling sequence is
    // only needed near the ends of the exponent range)
    const double den = alpha * (alpha + axj);
    double tr;
    if (__builtin_expect(den > 1e-280 && den < 1e280, 1)) {
      tr = __builtin_amdgcn_rcp(den);
      tr = fma(tr, fma(-den, tr, 1.0), tr);
      tr = fma(tr, fma(-den, tr, 1.0), tr);
    } else {
      tr = 1.0 / den;
    }
    const double dx = tr * ((dr[0] + dr[1]) + (dr[2] + dr[3]));
#pragma unroll
    for (int r = R0; r < 32; ++r) col[r].x = __builtin_fma(-v[r - R0], dx, col[r].x);
    return true;
  }
  if (lane == bi) {
#pragma unroll
    for (int r = R0; r < 32; ++r) {
      double2 v = col[r];
      if (r < R0 + 4) {
        if (r < j) v = make_double2(0.0, 0.0);
        if (r == j) v = make_double2(xj.x + e.x * alpha, xj.y + e.y * alpha);
      }
      vb[r] = v;
    }
  }
  wave_sync();
  double dr[4] = {0.0, 0.0, 0.0, 0.0}, di[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = R0; r < 32; ++r) {
    const double2 v = vb[r], c = col[r];
    if constexpr (REAL) {
      dr[r & 3] = fma(v.x, c.x, dr[r & 3]);
    } else {
      dr[r & 3] = fma(v.x, c.x, fma(v.y, c.y, dr[r & 3]));  // conj(v) c
      di[r & 3] = fma(v.x, c.y, fma(-v.y, c.x, di[r & 3]));
    }
  }
  const double dx = tau * ((dr[0] + dr[1]) + (dr[2] + dr[3]));
  const double dy = tau * ((di[0] + di[1]) + (di[2] + di[3]));
#pragma unroll
  for (int r = R0; r < 32; ++r) {
    const double2 w = vb[r];
    if constexpr (REAL) {
      col[r].x -= w.x * dx;
    } else {
      col[r].x -= w.x * dx - w.y * dy;
      col[r].y -= w.x * dy + w.y * dx;
    }
  }
  wave_sync();
  return true;
}

// Householder QR with column pivoting of the Theta block (columns < nc) of
// the panel, applied to every panel column (so the identity block becomes
// Q^H): panel <- Q^H panel, Q^H Theta = R P^T with |R_00| >= |R_11| >= ...
// Lane c keeps panel column c in registers (rows >= m are zero and stay
// zero: the reflectors vanish there); the reflector of each step is broadcast
// through `vb`.  Returns the squared Frobenius norm of the rows >= `steps` of
// Q^H Theta (the trailing block R22).  steps <= 16.  REAL: a real panel
// (imaginary parts zero), real arithmetic.
template <bool REAL>
__device__ double wave_pivoted_qr(double2 (*W)[PW], double2 *vb, int m, int nc, int width, int steps, int lane,
                                  double &piv0) {
  piv0 = 0.0;
  double2 col[32];
  const bool own = lane < width;
#pragma unroll
  for (int r = 0; r < 32; ++r) {
    col[r] = (own && r < m) ? W[r][lane] : make_double2(0.0, 0.0);
    if constexpr (REAL) col[r].y = 0.0;
  }
  for (int j = 0; j < steps; ++j) {
    bool more;
    switch (j >> 2) {
      case 0: more = qr_step<REAL, 0>(col, vb, j, nc, lane, piv0); break;
      case 1: more = qr_step<REAL, 4>(col, vb, j, nc, lane, piv0); break;
      case 2: more = qr_step<REAL, 8>(col, vb, j, nc, lane, piv0); break;
      default: more = qr_step<REAL, 12>(col, vb, j, nc, lane, piv0); break;
    }
    if (!more) break;
  }
  double tr = 0.0;
#pragma unroll
  for (int r = 0; r < 32; ++r)
    if (lane < nc && r >= steps) tr += col[r].x * col[r].x + col[r].y * col[r].y;
#pragma unroll
  for (int r = 0; r < 32; ++r)
    if (own && r < m) W[r][lane] = col[r];
  wave_sync();
  return wave_sum_uniform(tr);
}

// One-sided Jacobi on `rows` rows of panel W (Theta block nc wide, the
// accumulated Q^H block after it, `width` columns in all).  One wavefront,
// 4 lanes per row pair (8 when rows <= 16).  Returns the sweeps used.
__device__ int wave_jacobi(double2 (*W)[PW], int rows, int nc, int width, int lane, double frob2) {
  const int lpp = rows > 16 ? 4 : 8;
  const int pair = lane / lpp, sub = lane % lpp, npairs = rows >> 1;
  const double tol2 = 1e-28;              // |gamma| > 1e-14 sqrt(alpha beta)
  const double abs2 = 1e-56 * frob2 * frob2;  // |gamma| > 1e-28 ||Theta||_F^2
  int sweep = 0;
  for (; sweep < MAX_SWEEPS; ++sweep) {
    int rotated = 0;
    for (int round = 0; round < rows - 1; ++round) {
      if (pair < npairs) {
        int p, q;
        rr_pair(round, pair, rows, p, q);
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        for (int x = sub; x < nc; x += lpp) {
          const double2 a = W[p][x], b = W[q][x];
          al += a.x * a.x + a.y * a.y;
          be += b.x * b.x + b.y * b.y;
          gr += a.x * b.x + a.y * b.y;  // gamma = sum a conj(b)
          gi += a.y * b.x - a.x * b.y;
        }
        al = group_sum(al, lpp);
        be = group_sum(be, lpp);
        gr = group_sum(gr, lpp);
        gi = group_sum(gi, lpp);
        const double g2 = gr * gr + gi * gi;
        if (g2 > tol2 * al * be && g2 > abs2) {
          rotated = 1;
          const double rg = rsqrt(g2);
          const double g = g2 * rg, er = gr * rg, ei = gi * rg;
          const double d = be - al;
          const double h = sqrt(d * d + 4.0 * g2);
          const double t = (d >= 0.0 ? 2.0 * g : -2.0 * g) / (fabs(d) + h);
          const double c = rsqrt(1.0 + t * t), s = c * t;
          for (int x = sub; x < width; x += lpp) {
            const double2 a = W[p][x], b = W[q][x];
            const double ebr = er * b.x - ei * b.y, ebi = er * b.y + ei * b.x;
            W[p][x] = make_double2(c * a.x - s * ebr, c * a.y - s * ebi);
            W[q][x] = make_double2(s * a.x + c * ebr, s * a.y + c * ebi);
          }
        }
      }
      wave_sync();
    }
    if (!__any(rotated)) break;
  }
  return sweep + 1;
}

// REAL: the circuit's gate matrices are real, so the sites, the pending
// single-site matrices and Theta are real; the Theta products take one MFMA
// product instead of four and the QR runs in real arithmetic.
template <bool REAL>
__device__ void run_gate(PrepLds &L, int w, int lane, const ProbDev &P, const int gi, const GateDev gd,
                         const double *x, double2 *sites, int *status, bool exact) {
  const unsigned long long t0 = clk();
  const int s = gd.site;
  const int Dl = L.bond[s], Dc = L.bond[s + 1], Dr = L.bond[s + 2];
  const int m = 2 * Dl, nc = 2 * Dr;
  double2 Ul[4], Ur[4];
  if (gi < GV_PEND_LDS) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      Ul[q] = L.pend[gi][q];
      Ur[q] = L.pend[gi][4 + q];
    }
  } else {
    pending_matrix(P.preops + gd.pre0, gd.npre_l, x, P.fixes, Ul);
    pending_matrix(P.preops + gd.pre0 + gd.npre_l, gd.npre_r, x, P.fixes, Ur);
  }
  const double2 *sa = sites + (size_t)s * GV_SLOT;
  const double2 *sb = sites + (size_t)(s + 1) * GV_SLOT;
  const int j16 = lane & 15, g = lane >> 4;
  // A operand: lane (l + 16g), k-step t holds A'_i[l][kdx(g,t)];
  // B operand: lane (r + 16g), k-step t holds B'_j[kdx(g,t)][r].
  double ar[2][4], ai[2][4], br[2][4], bi[2][4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int c = kdx(g, t);
    const double2 a0 = sa[(j16 * 2 + 0) * 16 + c], a1 = sa[(j16 * 2 + 1) * 16 + c];
    const double2 b0 = sb[(c * 2 + 0) * 16 + j16], b1 = sb[(c * 2 + 1) * 16 + j16];
    const double2 A0 = cadd(cmul(Ul[0], a0), cmul(Ul[1], a1)), A1 = cadd(cmul(Ul[2], a0), cmul(Ul[3], a1));
    const double2 B0 = cadd(cmul(Ur[0], b0), cmul(Ur[1], b1)), B1 = cadd(cmul(Ur[2], b0), cmul(Ur[3], b1));
    ar[0][t] = A0.x; ai[0][t] = A0.y; ar[1][t] = A1.x; ai[1][t] = A1.y;
    br[0][t] = B0.x; bi[0][t] = B0.y; br[1][t] = B1.x; bi[1][t] = B1.y;
  }
  d4 cr[2][2], ci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      d4 zr = {0.0, 0.0, 0.0, 0.0}, zi = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        zr = mfma(ar[i][t], br[j][t], zr);
        if constexpr (!REAL) {
          zr = mfma(-ai[i][t], bi[j][t], zr);
          zi = mfma(ar[i][t], bi[j][t], zi);
          zi = mfma(ai[i][t], br[j][t], zi);
        }
      }
      cr[i][j] = zr;
      ci[i][j] = zi;
    }
  double2 G[16];
  const double *gp = P.gates + 32 * gd.gate;
#pragma unroll
  for (int e = 0; e < 16; ++e) G[e] = uload((const double2 *)gp + e);
  double2(*W)[PW] = L.W[w];
  double frob = 0.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int l = kdx(g, v), r = j16;
    if (l < Dl && r < Dr) {
#pragma unroll
      for (int I = 0; I < 2; ++I)
#pragma unroll
        for (int J = 0; J < 2; ++J) {
          double2 val = make_double2(0.0, 0.0);
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              const double2 gv = G[(I * 2 + J) * 4 + (i * 2 + j)];
              if constexpr (REAL) val.x += gv.x * cr[i][j][v];
              else val = cadd(val, cmul(gv, make_double2(cr[i][j][v], ci[i][j][v])));
            }
          W[l * 2 + I][J * Dr + r] = val;
          frob += val.x * val.x + val.y * val.y;
        }
    }
  }
  if (!exact)
    for (int idx = lane; idx < m * m; idx += 64) {
      const int row = idx / m, col = idx % m;
      W[row][nc + col] = make_double2(row == col ? 1.0 : 0.0, 0.0);
    }
  frob = wave_sum_uniform(frob);
  wave_sync();
  const unsigned long long t1 = clk();

  if (exact) {
    // Exact split: zero cutoffs and no gate of the tape cut below
    // min(2 Dl, 2 Dr), so the rule keeps k = min(m, nc) columns, all of
    // Theta.  Theta = I_m Theta (m <= nc: left site the identity isometry,
    // right site Theta's rows) or Theta = Theta I_nc (nc < m: left site
    // Theta, right site the identity co-isometry) is then the same state as
    // U_k (S_k Vh_k) with no rounding at all, and no singular value is
    // compared with anything, so no decomposition is needed.
    const bool left_id = m <= nc;
    const int k = left_id ? m : nc;
    double2 *wa = sites + (size_t)s * GV_SLOT;
    double2 *wb = sites + (size_t)(s + 1) * GV_SLOT;
    for (int idx = lane; idx < GV_SLOT; idx += 64) {
      const int row = idx >> 5, I = (idx >> 4) & 1, c = idx & 15;
      double2 va = make_double2(0.0, 0.0), vb = make_double2(0.0, 0.0);
      if (row < Dl && c < k) va = left_id ? make_double2(c == row * 2 + I ? 1.0 : 0.0, 0.0) : W[row * 2 + I][c];
      if (row < k && c < Dr) vb = left_id ? W[row][I * Dr + c] : make_double2(row == I * Dr + c ? 1.0 : 0.0, 0.0);
      wa[idx] = va;
      wb[idx] = vb;
    }
    if (lane == 0) {
      if (!isfinite(frob) && status) atomicMax(status, ST_NONFINITE);
      L.bond[s + 1] = k;
      atomicAdd(&L.cyc[0], (unsigned int)(t1 - t0));
      atomicAdd(&L.cyc[3], (unsigned int)(clk() - t1));
    }
    return;
  }

  // Preconditioning: pivoted QR leaves the singular values in the first
  // `rp` rows when the trailing block R22 is below the discard scale of the
  // rank rule (sigma_{rp+1} <= ||R22||_F); the Jacobi then runs on those rows
  // only.  Otherwise it runs on all m (pre-rotated) rows.
  const int count = min(m, nc);
  const bool zero_cut = (P.abs_cut == 0.0 && P.rel_cut == 0.0);
  const int rp = min(min(count, GV_CHI), P.max_bond + (P.max_bond & 1));
  double piv0;
  const double trail = wave_pivoted_qr<REAL>(W, L.vb[w], m, nc, nc + m, rp, lane, piv0);
  const double s0_lower = sqrt(piv0);  // largest column norm <= sigma_0
  const double lim = zero_cut ? 64.0 * 2.2204460492503131e-16 * count * s0_lower
                              : fmax(P.abs_cut, P.rel_cut * s0_lower);
  const int rows = (sqrt(trail) <= lim) ? rp : m;
  // Direct factorisation.  With zero cutoffs the rank rule keeps
  // keep = min(max_bond, 2 Dl, 2 Dr) columns.  When the QR has revealed that
  // rank (rows == rp == keep, ||R22||_F <= lim = 64 eps count |R_00|, and
  // |R_00| <= sigma_0), the factorisation Q[:, :keep] (Q^H Theta)[:keep, :]
  // and the truncated SVD U_k S_k Vh_k are both within ||R22||_F of Theta, and
  // sigma_keep <= ||R22||_F lies below the guard scale, so the guard cannot
  // fire.  The Jacobi sweeps would change the new bond's gauge and move the
  // two-site block by at most 2 lim in Frobenius norm; they are skipped:
  // left site <- Q[:, :keep] (isometry), right site <- (Q^H Theta)[:keep, :].
  const int keep = min(min(P.max_bond, count), GV_CHI);
  const bool direct = zero_cut && rows == rp && rp == keep;
  const unsigned long long t2 = clk();
  const int sweeps = direct ? 0 : wave_jacobi(W, rows, nc, nc + m, lane, frob);
  const unsigned long long t3 = clk();
  if (lane == 0) atomicAdd(&L.sweeps, sweeps);

  int k;
  bool bad = false;
  if (direct) {
    if (lane < rows) L.perm[w][lane] = lane;
    wave_sync();
    bad = !isfinite(frob);
    k = keep;
  } else {
    // singular values (squared) = row norms of the Theta block
    const int pair = lane >> 2, sub = lane & 3;
    for (int p = pair; p < rows; p += 16) {
      double f = 0.0;
      for (int c = sub; c < nc; c += 4) {
        const double2 v = W[p][c];
        f += v.x * v.x + v.y * v.y;
      }
      f = quad_sum(f);
      if (sub == 0) L.nrm[w][p] = f;
    }
    wave_sync();
    if (lane < rows) {
      const double mine = L.nrm[w][lane];
      int rank = 0;
      for (int q = 0; q < rows; ++q) {
        const double o = L.nrm[w][q];
        rank += (o > mine) || (o == mine && q < lane);
      }
      L.perm[w][rank] = lane;
    }
    wave_sync();
    // Rank rule, evaluated identically by every lane (selectDirectSvdRank).
    // Singular values past `rows` are bounded by sqrt(trail) <= lim: they are
    // below the guard scale (zero cutoffs) or below the cutoff threshold.
    const double s0 = sqrt(L.nrm[w][L.perm[w][0]]);
    const double thr = fmax(P.abs_cut, P.rel_cut * s0);
    const int known = min(rows, count);
    int selected = zero_cut ? count : 0;
    for (int i = 0; i < known; ++i) {
      const double si = sqrt(L.nrm[w][L.perm[w][i]]);
      if (!isfinite(si)) bad = true;
      if (!zero_cut && si > thr) ++selected;
    }
    if (bad) {
      k = 1;
    } else if (zero_cut && P.max_bond < count && P.max_bond < rows &&
               sqrt(L.nrm[w][L.perm[w][P.max_bond]]) > 64.0 * 2.2204460492503131e-16 * count * s0) {
      if (lane == 0 && status) atomicMax(status, ST_GUARD);
      k = min(P.max_bond, count);
    } else {
      k = max(1, min(P.max_bond, selected));
    }
  }
  if (bad) {
    if (lane == 0 && status) atomicMax(status, ST_NONFINITE);
    k = 1;
  }
  k = min(min(k, rows), GV_CHI);
  double2 *wa = sites + (size_t)s * GV_SLOT;
  double2 *wb = sites + (size_t)(s + 1) * GV_SLOT;
  for (int idx = lane; idx < GV_SLOT; idx += 64) {
    const int row = idx >> 5, I = (idx >> 4) & 1, c = idx & 15;
    double2 va = make_double2(0.0, 0.0), vb = make_double2(0.0, 0.0);
    // direct factorisation: the row order is the QR's (identity permutation)
    if (row < Dl && c < k) va = cconj(W[direct ? c : L.perm[w][c]][nc + row * 2 + I]);
    if (row < k && c < Dr) vb = W[direct ? row : L.perm[w][row]][I * Dr + c];
    wa[idx] = va;
    wb[idx] = vb;
  }
  if (lane == 0) {
    L.bond[s + 1] = k;
    const unsigned long long t4 = clk();
    atomicAdd(&L.cyc[0], (unsigned int)(t1 - t0));
    atomicAdd(&L.cyc[1], (unsigned int)(t2 - t1));
    atomicAdd(&L.cyc[2], (unsigned int)(t3 - t2));
    atomicAdd(&L.cyc[3], (unsigned int)(t4 - t3));
  }
}

template <bool REAL>
__device__ __forceinline__ void prep_body(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs,
                                          const double *__restrict__ params, double2 *__restrict__ arena,
                                          int *__restrict__ bonds_out, int *__restrict__ status,
                                          int *__restrict__ stats, int factor_always) {
  __shared__ PrepLds L;
  const int tid = threadIdx.x, lane = tid & 63, w = wave_id();
  const InstDev inst = uload(insts + blockIdx.x);
  const ProbDev P = uload(probs + inst.prob);
  const int n = P.n;
  double2 *sites = arena + inst.site_off;
  const double *x = params + inst.param_off;
  int *st = status ? status + blockIdx.x : nullptr;

  for (int e = tid; e < n * GV_SLOT; e += 64 * GV_WAVES)
    sites[e] = make_double2((e % GV_SLOT) == 0 ? 1.0 : 0.0, 0.0);
  for (int s = tid; s <= n; s += 64 * GV_WAVES) L.bond[s] = 1;
  if (tid == 0 && st) *st = ST_OK;
  if (tid == 0) {
    L.sweeps = 0;
    L.cyc[0] = L.cyc[1] = L.cyc[2] = L.cyc[3] = 0;
  }
  for (int gi = tid; gi < min(P.ngates, GV_PEND_LDS); gi += 64 * GV_WAVES) {
    const GateDev gd = P.g2[gi];
    double2 U[4];
    pending_matrix(P.preops + gd.pre0, gd.npre_l, x, P.fixes, U);
#pragma unroll
    for (int q = 0; q < 4; ++q) L.pend[gi][q] = U[q];
    pending_matrix(P.preops + gd.pre0 + gd.npre_l, gd.npre_r, x, P.fixes, U);
#pragma unroll
    for (int q = 0; q < 4; ++q) L.pend[gi][4 + q] = U[q];
  }
  __syncthreads();

  const bool exact = P.exact_split && !factor_always;
  for (int row = 0; row < P.nsched; ++row) {
    const int gi = uload(P.sched + row * GV_WAVES + w);
    if (gi >= 0) run_gate<REAL>(L, w, lane, P, gi, uload(P.g2 + gi), x, sites, st, exact);
    __syncthreads();
  }
  // single-site ops left pending after the last two-site gate on each site
  for (int s = 0; s < n; ++s) {
    const int b = uload(P.flush + s), e = uload(P.flush + s + 1);
    if (e == b) continue;
    double2 U[4];
    pending_matrix(P.preops + b, e - b, x, P.fixes, U);
    double2 *slot = sites + (size_t)s * GV_SLOT;
    for (int idx = tid; idx < GV_CHI * GV_CHI; idx += 64 * GV_WAVES) {
      const int l = idx >> 4, r = idx & 15;
      const double2 a0 = slot[(l * 2 + 0) * 16 + r], a1 = slot[(l * 2 + 1) * 16 + r];
      slot[(l * 2 + 0) * 16 + r] = cadd(cmul(U[0], a0), cmul(U[1], a1));
      slot[(l * 2 + 1) * 16 + r] = cadd(cmul(U[2], a0), cmul(U[3], a1));
    }
  }
  if (bonds_out)
    for (int s = tid; s <= n; s += 64 * GV_WAVES) bonds_out[(size_t)blockIdx.x * (GV_NMAX + 1) + s] = L.bond[s];
  if (stats && tid == 0) {
    int *o = stats + (size_t)blockIdx.x * GV_STATS;
    o[0] = L.sweeps;
    for (int i = 0; i < 4; ++i) o[1 + i] = (int)(L.cyc[i] >> 4);  // units of 16 clocks
  }
}

extern "C" __global__ void __launch_bounds__(64 * GV_WAVES)
k_prep(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs, const double *__restrict__ params,
       double2 *__restrict__ arena, int *__restrict__ bonds_out, int *__restrict__ status, int *__restrict__ stats,
       int factor_always) {
  prep_body<false>(insts, probs, params, arena, bonds_out, status, stats, factor_always);
}

// Real circuits (every gate matrix real): the same preparation in real arithmetic.
extern "C" __global__ void __launch_bounds__(64 * GV_WAVES)
k_prep_real(const InstDev *__restrict__ insts, const ProbDev *__restrict__ probs, const double *__restrict__ params,
            double2 *__restrict__ arena, int *__restrict__ bonds_out, int *__restrict__ status,
            int *__restrict__ stats, int factor_always) {
  prep_body<true>(insts, probs, params, arena, bonds_out, status, stats, factor_always);
}

namespace gv {

hipError_t kl_prep(int ninst, bool real, bool factor_always, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                   const double *params, double2 *arena, int *bonds, int *status, int *stats) {
  if (real)
    hipLaunchKernelGGL(k_prep_real, dim3(ninst), dim3(64 * GV_WAVES), 0, s, insts, probs, params, arena, bonds,
                       status, stats, (int)factor_always);
  else
    hipLaunchKernelGGL(k_prep, dim3(ninst), dim3(64 * GV_WAVES), 0, s, insts, probs, params, arena, bonds, status,
                       stats, (int)factor_always);
  return hipGetLastError();
}

}  // namespace gv
