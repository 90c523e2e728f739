// Host planner interface (see plan.cpp).
#pragma once
#include <array>
#include <complex>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gvmps.h"
#include "gvmps_internal.h"

namespace gv {

using cd = std::complex<double>;

struct Problem {
  int n = 0, nparams = 0, max_bond = 0;
  double abs_cut = 0, rel_cut = 0;
  // circuit
  std::vector<ProgOp> prog;
  std::vector<double> gates;   // 32 doubles per two-site gate
  std::vector<double> fixes;   // 8 doubles per fixed single-site gate
  std::vector<int> bonds_pred; // n+1, zero-cutoff allocation
  int max_alloc_bond = 1, n_g2 = 0;
  bool rank_reduced = false;   // some two-site gate has min(2 Dl, 2 Dr) > max_bond
  bool exact_split = false;    // zero cutoffs and no rank reduction: every split is exact
  // pipelined schedule of the two-site gates
  std::vector<GateDev> g2;
  std::vector<int> sched;      // rows of GV_WAVES gate indices (-1 idle)
  std::vector<ProgOp> preops;  // gate pre-ops, then the final per-site flush ops
  std::vector<int> flush;      // n+1 offsets into preops
  double flops_svd = 0;
  // Hamiltonian
  int nterms = 0;
  std::vector<double> coeffs;            // 2*nterms
  std::vector<std::string> site_words;   // site order
  std::vector<int> active;               // non-identity terms
  bool real_ok = false;                  // real gate matrices and no Y in any word: real-arithmetic energies
  bool iz_only = false;                  // every forest step has I/Z children only (no X or Y words)
  // prefix forest, chain schedule
  std::vector<ChainDev> chains;
  std::vector<StepDev> steps;
  std::vector<ChildDev> children;
  std::vector<TermEnd> term_list;
  int n_right_nodes = 0;                 // suffix-forest nodes
  std::vector<int> gen_begin;            // energy launch g: chains [gen_begin[g], gen_begin[g+1])
  std::vector<int> ggen_begin, ggen_chains;  // gradient generation g: chains ggen_chains[ggen_begin[g] ..)
  int nggen = 0;
  std::vector<int> bgen_begin, bgen_chains;  // backward launch g (run from the last to 0): chains bgen_chains[bgen_begin[g] ..)
  int nbgen = 0;
  int nspill = 0, ngen = 0, n_nodes = 0;
  std::vector<LeafDev> leaves;           // deferred term-end leaves (energy passes)
  // reverse gradient
  std::string grad_error;                 // non-empty: the circuit has no bond-doubling form here
  std::vector<LocalOp> lops;              // per-site local programs
  std::vector<int> lop_begin;             // n+1
  std::vector<int> rot_param;             // per rotation id
  std::vector<double> rot_scale;
  std::vector<int> param_rot_begin, param_rots;  // CSR: rotations of each parameter
  int ngleaf = 0;
  std::vector<int> seed_begin;            // nspill+1
  std::vector<SeedTerm> seed_terms;
  std::vector<int> site_steps_begin, site_steps;
  std::vector<int> site_tail;  // per site: trailing CNOT halves pulled back in the backward chains
  double flops_grad = 0;
  double flops_exec = 0, flops_nodes = 0;
};

std::unique_ptr<Problem> compile_problem(int n_qubits, const gvmps_op *ops, int n_ops, int n_params,
                                         int max_bond, double abs_cut, double rel_cut,
                                         const char *words, const double *coeffs, int n_terms);

}  // namespace gv
