// Host planner: compiles a circuit tape into a site program and a Pauli-word
// list into the prefix forest + chain schedule consumed by the energy kernels.
// Pure C++ (no HIP calls), so it runs and is tested without a GPU.
#include "plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace gv {

namespace {

using M2 = std::array<cd, 4>;
using M4 = std::array<cd, 16>;

M2 fixed_matrix(int kind) {
  const double r = 1.0 / std::sqrt(2.0);
  switch (kind) {
    case GVMPS_H: return {cd(r, 0), cd(r, 0), cd(r, 0), cd(-r, 0)};
    case GVMPS_X: return {cd(0, 0), cd(1, 0), cd(1, 0), cd(0, 0)};
    case GVMPS_Y: return {cd(0, 0), cd(0, -1), cd(0, 1), cd(0, 0)};
    case GVMPS_Z: return {cd(1, 0), cd(0, 0), cd(0, 0), cd(-1, 0)};
    case GVMPS_S: return {cd(1, 0), cd(0, 0), cd(0, 0), cd(0, 1)};
    case GVMPS_SDG: return {cd(1, 0), cd(0, 0), cd(0, 0), cd(0, -1)};
  }
  throw std::runtime_error("not a fixed single-qubit gate");
}

// Controlled-U with the control first, row-major [(C*2+T)*4+(c*2+t)]
// (simulator_mps_dist.cu:374-384).
M4 controlled(const M2 &U) {
  M4 G{};
  for (int C = 0; C < 2; ++C)
    for (int T = 0; T < 2; ++T)
      for (int c = 0; c < 2; ++c)
        for (int t = 0; t < 2; ++t) {
          if (C != c) continue;
          G[(C * 2 + T) * 4 + (c * 2 + t)] = (c == 0) ? cd(T == t ? 1.0 : 0.0, 0) : U[T * 2 + t];
        }
  return G;
}

M4 swap_matrix() {  // sharded_mps.cuh:551-556
  M4 S{};
  for (int I = 0; I < 2; ++I)
    for (int J = 0; J < 2; ++J)
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) S[(I * 2 + J) * 4 + (i * 2 + j)] = (I == j && J == i) ? 1.0 : 0.0;
  return S;
}

M4 swap_operands(const M4 &G) {  // sharded_mps.cuh:557-562
  M4 H{};
  for (int I = 0; I < 2; ++I)
    for (int J = 0; J < 2; ++J)
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) H[(I * 2 + J) * 4 + (i * 2 + j)] = G[(J * 2 + I) * 4 + (j * 2 + i)];
  return H;
}

int kind_of(char p) {
  switch (p) {
    case 'I': return CK_I;
    case 'Z': return CK_Z;
    case 'X': return CK_X;
    case 'Y': return CK_Y;
  }
  return -1;
}

struct CircuitCompiler {
  Problem &P;
  std::vector<int> bond;
  explicit CircuitCompiler(Problem &p) : P(p), bond(p.n + 1, 1) {}

  int gate_index(const M4 &G) {
    const int count = (int)P.gates.size() / 32;
    for (int g = 0; g < count; ++g) {
      bool same = true;
      for (int e = 0; e < 16 && same; ++e)
        same = P.gates[g * 32 + 2 * e] == G[e].real() && P.gates[g * 32 + 2 * e + 1] == G[e].imag();
      if (same) return g;
    }
    for (int e = 0; e < 16; ++e) {
      P.gates.push_back(G[e].real());
      P.gates.push_back(G[e].imag());
    }
    return count;
  }

  int fix_index(const M2 &U) {
    const int count = (int)P.fixes.size() / 8;
    for (int g = 0; g < count; ++g) {
      bool same = true;
      for (int e = 0; e < 4 && same; ++e)
        same = P.fixes[g * 8 + 2 * e] == U[e].real() && P.fixes[g * 8 + 2 * e + 1] == U[e].imag();
      if (same) return g;
    }
    for (int e = 0; e < 4; ++e) {
      P.fixes.push_back(U[e].real());
      P.fixes.push_back(U[e].imag());
    }
    return count;
  }

  // Adjacent two-site gate on sites (s, s+1).  The allocated bond follows the
  // zero-cutoff rank rule k = min(2 Dl, 2 Dr, max_bond) (dist_svd.cuh:194-212,
  // restated by test/tools/graph_resources.py:19-31).
  void emit2(const M4 &G, int s) {
    ProgOp op{};
    op.kind = OP_G2;
    op.site = s;
    op.arg = gate_index(G);
    op.param = -1;
    P.prog.push_back(op);
    const int m = 2 * bond[s], nc = 2 * bond[s + 2];
    const int k = std::min(std::min(m, nc), P.max_bond);
    bond[s + 1] = k;
    P.max_alloc_bond = std::max(P.max_alloc_bond, k);
    P.n_g2++;
    const double M = std::max(m, nc), N = std::min(m, nc);
    P.flops_svd += 4.0 * (4.0 * M * M * N + 8.0 * M * N * N + 9.0 * N * N * N);
  }

  // SWAP-routed arbitrary pair in site indices (sharded_mps.cuh:565-572).
  void two(const M4 &G, int a, int b) {
    if (a == b) throw std::runtime_error("2q gate on same qubit");
    if (a > b) {
      two(swap_operands(G), b, a);
      return;
    }
    const M4 SW = swap_matrix();
    for (int x = b; x > a + 1; --x) emit2(SW, x - 1);
    emit2(G, a);
    for (int x = a + 1; x < b; ++x) emit2(SW, x);
  }

  void compile(const gvmps_op *ops, int n_ops) {
    const int n = P.n;
    auto site = [n](int q) {
      if (q < 0 || q >= n) throw std::runtime_error("qubit index out of range");
      return n - 1 - q;  // simulator_mps_dist.cu:386-387
    };
    for (int k = 0; k < n_ops; ++k) {
      const gvmps_op &o = ops[k];
      switch (o.kind) {
        case GVMPS_RX:
        case GVMPS_RY:
        case GVMPS_RZ: {
          if (o.param >= P.nparams) throw std::runtime_error("rotation parameter index out of range");
          ProgOp op{};
          op.kind = OP_ROT;
          op.site = site(o.q0);
          op.arg = o.kind - GVMPS_RX;
          op.param = o.param;
          op.scale = o.param >= 0 ? o.scale : 0.0;
          op.offset = o.offset;
          P.prog.push_back(op);
          break;
        }
        case GVMPS_H:
        case GVMPS_X:
        case GVMPS_Y:
        case GVMPS_Z:
        case GVMPS_S:
        case GVMPS_SDG: {
          ProgOp op{};
          op.kind = OP_FIX1;
          op.site = site(o.q0);
          op.arg = fix_index(fixed_matrix(o.kind));
          op.param = -1;
          P.prog.push_back(op);
          break;
        }
        case GVMPS_CX:
          two(controlled(fixed_matrix(GVMPS_X)), site(o.q0), site(o.q1));
          break;
        case GVMPS_SWAP:
          two(swap_matrix(), site(o.q0), site(o.q1));
          break;
        default:
          throw std::runtime_error(
              "tensornet-mps: >2-qubit or multi-control gates not yet supported "
              "(decompose, or extend applyGate).");
      }
    }
    P.bonds_pred = bond;
  }
};

struct TNode {
  int site, kind, parent;
  int child[4];
  std::vector<int> terms;
  long long work = 0;
};

void build_chains(Problem &P) {
  const int n = P.n;
  std::vector<TNode> nodes;
  std::vector<std::array<int, 4>> roots(n);
  for (auto &r : roots) r = {-1, -1, -1, -1};
  auto make = [&](int site, int kind, int parent) {
    TNode t;
    t.site = site;
    t.kind = kind;
    t.parent = parent;
    for (int &c : t.child) c = -1;
    nodes.push_back(std::move(t));
    return (int)nodes.size() - 1;
  };
  for (int k : P.active) {
    const std::string &w = P.site_words[k];
    int a = 0, b = n - 1;
    while (w[a] == 'I') ++a;
    while (w[b] == 'I') --b;
    int kd = kind_of(w[a]);
    int node = roots[a][kd];
    if (node < 0) node = roots[a][kd] = make(a, kd, -1);
    for (int s = a + 1; s <= b; ++s) {
      kd = kind_of(w[s]);
      int nx = nodes[node].child[kd];
      if (nx < 0) {
        nx = make(s, kd, node);
        nodes[node].child[kd] = nx;
      }
      node = nx;
    }
    nodes[node].terms.push_back(k);
  }
  P.n_nodes = (int)nodes.size();
  // Subtree expansion counts (children are created after their parents).
  for (int v = (int)nodes.size() - 1; v >= 0; --v) {
    long long w = 0;
    bool any = false;
    for (int c : nodes[v].child)
      if (c >= 0) {
        any = true;
        w += nodes[c].work;
      }
    nodes[v].work = w + (any ? 1 : 0);
  }
  // Algorithmic node count: 16 (Dl^2 Dr + Dl Dr^2) flops per prefix node.
  const std::vector<int> &bd = P.bonds_pred;
  auto fsite = [&](int s) {
    const double dl = bd[s], dr = bd[s + 1];
    return 16.0 * (dl * dl * dr + dl * dr * dr);
  };
  P.flops_nodes = 0;
  for (const TNode &t : nodes) P.flops_nodes += fsite(t.site);
  for (int s = 0; s < n; ++s) P.flops_nodes += 2.0 * fsite(s);

  struct Start {
    int slot, site, node, gen;
  };
  std::vector<Start> queue;
  for (int a = 0; a < n; ++a)
    if (roots[a][0] >= 0 || roots[a][1] >= 0 || roots[a][2] >= 0 || roots[a][3] >= 0)
      queue.push_back({-1, a, -1, 0});
  std::vector<ChainDev> chains;
  std::vector<int> gens;
  std::vector<std::vector<StepDev>> chain_steps;
  std::vector<std::vector<ChildDev>> chain_children;
  P.nspill = 0;
  const double chi = GV_CHI;
  P.flops_exec = 2.0 * n * 32.0 * chi * chi * chi;  // L and R identity sweeps, 4 chi^3 complex MACs each
  for (size_t qi = 0; qi < queue.size(); ++qi) {
    const Start st = queue[qi];
    ChainDev ch{};
    ch.site0 = st.site;
    ch.src = st.slot;
    std::vector<StepDev> steps;
    std::vector<ChildDev> kids_out;
    std::array<int, 4> kids = st.node < 0 ? roots[st.site] : std::array<int, 4>{nodes[st.node].child[0],
                                                                               nodes[st.node].child[1],
                                                                               nodes[st.node].child[2],
                                                                               nodes[st.node].child[3]};
    int site = st.site;
    for (;;) {
      StepDev sd{};
      sd.site = site;
      sd.child0 = (int)kids_out.size();
      int cont = -1;
      long long best = 0;
      for (int c : kids)
        if (c >= 0 && nodes[c].work > best) {
          best = nodes[c].work;
          cont = c;
        }
      for (int kd = 0; kd < 4; ++kd) {
        const int c = kids[kd];
        if (c < 0) continue;
        ChildDev cdv{};
        cdv.kind = kd;
        if (c == cont) cdv.action = -2;
        else if (nodes[c].work > 0) {
          cdv.action = P.nspill++;
          queue.push_back({cdv.action, site + 1, c, st.gen + 1});
        } else
          cdv.action = -1;
        cdv.term0 = -1;  // filled when flattening
        cdv.nterm = (int)nodes[c].terms.size();
        cdv.term0 = c;   // temporarily: node id
        kids_out.push_back(cdv);
        if (kd <= CK_Z) sd.need |= 1;
        else sd.need |= 2;
      }
      sd.nchild = (int)kids_out.size() - sd.child0;
      P.flops_exec += (sd.need == 3 ? 48.0 : 32.0) * chi * chi * chi;  // (4 or 6) chi^3 complex MACs
      steps.push_back(sd);
      if (cont < 0) break;
      kids = {nodes[cont].child[0], nodes[cont].child[1], nodes[cont].child[2], nodes[cont].child[3]};
      ++site;
    }
    ch.nsteps = (int)steps.size();
    chains.push_back(ch);
    gens.push_back(st.gen);
    chain_steps.push_back(std::move(steps));
    chain_children.push_back(std::move(kids_out));
  }
  // Order: generation ascending, then longest chains first.
  std::vector<int> order(chains.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (gens[x] != gens[y]) return gens[x] < gens[y];
    return chains[x].nsteps > chains[y].nsteps;
  });
  P.chains.clear();
  P.steps.clear();
  P.children.clear();
  P.term_list.clear();
  P.gen_begin.clear();
  int last_gen = -1;
  for (int idx : order) {
    while (last_gen < gens[idx]) {
      P.gen_begin.push_back((int)P.chains.size());
      ++last_gen;
    }
    ChainDev ch = chains[idx];
    ch.step0 = (int)P.steps.size();
    for (StepDev sd : chain_steps[idx]) {
      const int base = (int)P.children.size();
      for (int c = 0; c < sd.nchild; ++c) {
        ChildDev cdv = chain_children[idx][sd.child0 + c];
        const int node = cdv.term0;
        cdv.term0 = (int)P.term_list.size();
        for (int t : nodes[node].terms) P.term_list.push_back(t);
        P.children.push_back(cdv);
      }
      sd.child0 = base;
      P.steps.push_back(sd);
    }
    P.chains.push_back(ch);
  }
  P.gen_begin.push_back((int)P.chains.size());
  P.ngen = (int)P.gen_begin.size() - 1;
}

}  // namespace

std::unique_ptr<Problem> compile_problem(int n_qubits, const gvmps_op *ops, int n_ops, int n_params,
                                         int max_bond, double abs_cut, double rel_cut,
                                         const char *words, const double *coeffs, int n_terms) {
  if (n_qubits < 2) throw std::runtime_error("tensornet-mps needs >= 2 qubits");  // simulator_mps_dist.cu:235-237
  if (n_qubits > GV_NMAX) throw std::runtime_error("more than 128 qubits is not supported by the GPU layout");
  if (n_params < 0 || n_terms < 0 || n_ops < 0) throw std::runtime_error("negative size");
  if (max_bond < 1) max_bond = 1;  // std::max(1, atoi(b)), simulator_mps_dist.cu:363
  auto P = std::make_unique<Problem>();
  P->n = n_qubits;
  P->nparams = n_params;
  P->max_bond = max_bond;
  P->abs_cut = abs_cut;
  P->rel_cut = rel_cut;
  CircuitCompiler cc(*P);
  cc.compile(ops, n_ops);
  if (P->max_alloc_bond > GV_CHI)
    throw std::runtime_error("allocated bond " + std::to_string(P->max_alloc_bond) +
                             " exceeds the 16-wide bond slots of the GPU layout");
  P->nterms = n_terms;
  P->coeffs.assign(coeffs, coeffs + 2 * (size_t)n_terms);
  P->site_words.resize(n_terms);
  for (int k = 0; k < n_terms; ++k) {
    std::string w(words + (size_t)k * n_qubits, words + (size_t)(k + 1) * n_qubits);
    for (char c : w)
      if (c != 'I' && c != 'X' && c != 'Y' && c != 'Z') throw std::runtime_error("Invalid analytic Pauli");
    std::reverse(w.begin(), w.end());  // site order, simulator_mps_dist.cu:575-577
    bool identity = std::all_of(w.begin(), w.end(), [](char c) { return c == 'I'; });
    if (!identity) P->active.push_back(k);
    P->site_words[k] = std::move(w);
  }
  build_chains(*P);
  return P;
}

}  // namespace gv
