// Internal data structures shared by the host planner and the gfx950 kernels.
//
// Memory layout in HBM (one batch = many MPS instances, possibly of different
// problems):
//   site arena   : per instance, n fixed slots of 16 x 2 x 16 complex128
//                  (element (l,i,r) at (l*2+i)*16+r, zero padded) = 8 KiB/site
//   env arena    : per instance, L_0..L_n then Rt_0..Rt_n, 16 x 16 complex128
//                  row-major (L[bra][ket], Rt = R^T = [bra][ket])
//   spill arena  : per instance, one 16 x 16 slot per spilled prefix node
//   term values  : per instance, one complex128 <P_k> per Hamiltonian term
#pragma once
#include <cstdint>

#define GV_CHI 16            // fixed bond slot (max bond supported on the GPU path)
#define GV_SLOT (GV_CHI * 2 * GV_CHI)
#define GV_ENV (GV_CHI * GV_CHI)
#define GV_NMAX 128          // max sites per instance

namespace gv {

enum OpKind : int { OP_ROT = 1, OP_FIX1 = 2, OP_G2 = 3 };

// One instruction of a compiled site program (qubits already mapped to sites).
struct ProgOp {
  int kind;     // OpKind
  int site;     // ROT/FIX1: site; G2: left site s of the pair (s, s+1)
  int arg;      // ROT: axis 0=X 1=Y 2=Z; FIX1: index in fix table; G2: index in gate table
  int param;    // ROT: parameter index, -1 for a constant angle
  double scale; // ROT: angle = scale * x[param] + offset
  double offset;
};

// Children kinds of a prefix node expansion.
enum ChildKind : int { CK_I = 0, CK_Z = 1, CK_X = 2, CK_Y = 3 };

struct ChainDev {      // one chain of expansions
  int site0;           // site of the first expansion
  int nsteps;
  int step0;           // index of the first step
  int src;             // -1: start from L[site0]; >= 0: spill slot holding the start env
};

struct StepDev {       // one expansion: E -> children at site `site`
  int site;
  int need;            // bit0: Y00/Y11 needed (I,Z children); bit1: Y01/Y10 needed (X,Y)
  int child0;          // first child record
  int nchild;
};

struct ChildDev {
  int kind;            // ChildKind
  int action;          // -1: none, -2: continue the chain, >= 0: spill slot
  int term0;           // range into the term-end list
  int nterm;
};

// Device view of one compiled problem (circuit + Hamiltonian).
struct ProbDev {
  int n;               // sites (= qubits)
  int nparams;
  int nprog;
  int max_bond;
  double abs_cut, rel_cut;
  int nterms;
  int nspill;
  const ProgOp* prog;
  const double* gates;     // 2q gate table, 16 complex each (interleaved re,im)
  const double* fixes;     // 1q fixed table, 4 complex each
  const ChainDev* chains;
  const StepDev* steps;
  const ChildDev* children;
  const int* term_list;
  const double* coeffs;    // nterms complex (re,im)
  const int* active;       // nterms flags: 1 = contracted, 0 = identity word (<P> = 1)
};

// One MPS instance of a batch.
struct InstDev {
  int prob;            // index into the problem table
  int n;
  long long site_off;  // in complex elements
  long long env_off;
  long long spill_off;
  long long term_off;  // in complex elements
  long long param_off; // in doubles
  int out;             // output index (energy slot)
  int pad;
};

// Status codes written by kernels (per instance).
enum Status : int { ST_OK = 0, ST_GUARD = 2, ST_NONFINITE = 1 };

}  // namespace gv
