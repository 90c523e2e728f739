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

// Two-site gates of the pipelined state-preparation schedule.  The single-
// site ops that precede a gate on its two sites (in program order) are folded
// into it: pre-ops [pre0, pre0+npre_l) act on the left site, the next npre_r
// on the right site.
#define GV_WAVES 4           // gates in flight per instance (one wavefront each)
#define GV_STATS 8           // state-preparation statistics per instance (ints)
struct GateDev {
  int site;    // left site s of (s, s+1)
  int gate;    // 4x4 gate table index
  int pre0, npre_l, npre_r;
};

// Bond-doubling construction of one site for the reverse gradient (no SVD):
// a CNOT on neighbouring sites doubles their shared bond exactly
// (test/tools/reverse_mps.py:45-50, 161-171).
enum LocalKind : int {
  LO_ROT = 1,     // rotation, arg = axis, param/scale/offset as ProgOp, rot = rotation id
  LO_FIX = 2,     // fixed 2x2, arg = fix table index
  LO_CTRL_R = 3,  // control, partner on the right: A'[l,i,(r,s)] = A[l,i,r] d(i,s)
  LO_CTRL_L = 4,  // control, partner on the left:  A'[(l,s),i,r] = A[l,i,r] d(i,s)
  LO_TARG_R = 5,  // target, partner on the right:  A'[l,i,(r,s)] = A[l,i^s,r]
  LO_TARG_L = 6   // target, partner on the left:   A'[(l,s),i,r] = A[l,i^s,r]
};
struct LocalOp {
  int kind, arg, param, rot;
  double scale, offset;
  int dl, dr;  // bond dimensions after the op
};

// Children kinds of a prefix node expansion.
enum ChildKind : int { CK_I = 0, CK_Z = 1, CK_X = 2, CK_Y = 3 };

// The prefix forest grows terms from their first non-identity site to the
// right (dir 0, left environments, A_i); the suffix forest grows them from
// their last non-identity site to the left (dir 1, transposed environments,
// A_i^T).  Each term is split once: its left part ends at site c, its right
// part starts at c+1 and <P> = Tr(E_left(c) R_right(c+1)).
struct ChainDev {      // one chain of expansions
  int site0;           // site of the first expansion
  int nsteps;
  int step0;           // index of the first step
  int src;             // -1: start from the identity environment; >= 0: spill slot
  int dir;             // 0: prefix forest (left to right), 1: suffix forest (right to left)
  int pad;
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
  int term0;           // range into the term-end list (prefix forest)
  int nterm;
  int store;           // suffix forest: slot receiving this node's environment, -1 none
  int gleaf;           // gradient: slot keeping this prefix leaf's environment, -1 none
  int defer;           // energy: leaf slot whose term ends a separate pass contracts, -1 inline
};

// A prefix node with many term ends: its environment is stored in a leaf
// slot and k_leaf_terms contracts the terms [term0, term0 + nterm) of the
// term-end list in parallel, off the chain's critical path.
#define GV_DEFER_TERMS 3
// Shallow prefix-forest depths with at most this many chains per instance run
// in the suffix forest's energy launches (their term ends deferred).
#define GV_MERGE_CHAINS 64
// Steps of a root chain before its heavy path continues in the next generation.
#ifndef GV_ROOT_STEPS
#define GV_ROOT_STEPS 5
#endif
struct LeafDev {
  int term0, nterm;
};

// A term ending at a prefix-forest node: its index, the slot of its
// suffix-forest environment (-1: identity right environment Rt[site+1]) and
// Re c_k, so a term end's one scalar load brings its coefficient too.
struct TermEnd {
  int term;
  int rslot;
  double cre;
};

// One deferred leaf of a batch, with everything its wavefront needs in one
// scalar load (the real leaf kernels): its term ends, the offsets (in 16-byte
// units) of its leaf environment in the leaf arena and of its instance's
// suffix-forest slots in the spill arena, and where its partial energy and
// (per-term runs) its term values go.
struct LeafItem {
  const TermEnd* tl;
  long long env;
  long long spill;
  long long part;
  long long tv;
  int nterm;
  int pad;
};

// A term whose suffix-forest environment sits in a store slot; its prefix
// leaf environment was kept in gradient slot `gleaf` (reverse gradient seeds).
struct SeedTerm {
  int term;
  int gleaf;
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
  int nchains;
  const ProgOp* prog;
  int ngates, nsched;      // two-site gates; schedule rows of GV_WAVES gate slots
  const GateDev* g2;
  const int* sched;        // nsched * GV_WAVES gate indices, -1 = idle slot
  const ProgOp* preops;
  const int* flush;        // n+1 offsets into preops: ops left pending on site s at the end
  const double* gates;     // 2q gate table, 16 complex each (interleaved re,im)
  const double* fixes;     // 1q fixed table, 4 complex each
  const ChainDev* chains;
  const StepDev* steps;
  const ChildDev* children;
  const TermEnd* term_list;
  const double* coeffs;    // nterms complex (re,im)
  const int* active;       // nterms flags: 1 = contracted, 0 = identity word (<P> = 1)
  // reverse gradient (bond-doubling state)
  int nsteps, ngleaf, nrot;
  const LocalOp* lops;     // per-site local programs
  const int* lop_begin;    // n+1 offsets
  const int* seed_begin;   // nspill+1 offsets: terms seeding each suffix-forest store slot
  const SeedTerm* seed_terms;
  const int* site_steps_begin;  // n+1 offsets
  const int* site_steps;   // steps at each site (fixed summation order)
  const int* ck_slot;           // per step: checkpoint slot of its input environment, -1 = recomputed
  int nleaf;                    // deferred term-end leaves
  const LeafDev* leaves;
  const int* param_rot_begin;   // nparams+1 offsets: rotations driven by each parameter
  const int* param_rots;
  const double* rot_scale;      // per rotation: d angle / d parameter
  int exact_split;              // 1: every two-site split keeps all of Theta (zero cutoffs, no rank reduction)
  // Per site: the CNOT halves that end its local program, at most one per
  // bond side, as count | (last op kind << 4) | (op before it << 8).  The real
  // backward chains pull each step's site adjoint back through them before
  // storing it, so the stored adjoint has 128, 256 or 512 entries.
  const int* site_tail;
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
  long long part_off;  // chain partial energies (one double per chain)
  int out;             // output index (energy slot)
  int pad;
  // reverse gradient arenas (in slots of their own kind)
  long long rot_off;   // rotation inputs (8 KiB) and rotation gradients (one double)
  long long step_off;  // step input environments (4 KiB) and step site adjoints (8 KiB)
  long long gleaf_off; // kept prefix leaf environments (4 KiB)
  long long ck_off;    // checkpointed step input environments (4 KiB)
  long long leaf_off;  // deferred term-end leaf environments (4 KiB)
};


// Status codes written by kernels (per instance).
enum Status : int { ST_OK = 0, ST_GUARD = 2, ST_NONFINITE = 1 };


}  // namespace gv
