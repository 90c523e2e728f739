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
    if (k < std::min(m, nc)) P.rank_reduced = true;
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

// Pipelined gate schedule.  A two-site gate depends on the previous gates on
// either of its sites; gates at the same ASAP level touch disjoint sites and
// run concurrently, GV_WAVES per time step.  The order of operations on every
// single site is the program order, so the state is the same as the
// sequential gate stream of the reference (simulator_mps_dist.cu:1392-1406).
void build_schedule(Problem &P) {
  const int n = P.n;
  std::vector<std::vector<ProgOp>> pending(n);
  std::vector<int> last(n, -1), level;
  P.g2.clear();
  P.preops.clear();
  for (const ProgOp &op : P.prog) {
    if (op.kind != OP_G2) {
      pending[op.site].push_back(op);
      continue;
    }
    const int s = op.site;
    GateDev gd{};
    gd.site = s;
    gd.gate = op.arg;
    gd.pre0 = (int)P.preops.size();
    gd.npre_l = (int)pending[s].size();
    gd.npre_r = (int)pending[s + 1].size();
    P.preops.insert(P.preops.end(), pending[s].begin(), pending[s].end());
    P.preops.insert(P.preops.end(), pending[s + 1].begin(), pending[s + 1].end());
    pending[s].clear();
    pending[s + 1].clear();
    int lv = 0;
    if (last[s] >= 0) lv = std::max(lv, level[last[s]] + 1);
    if (last[s + 1] >= 0) lv = std::max(lv, level[last[s + 1]] + 1);
    level.push_back(lv);
    last[s] = last[s + 1] = (int)P.g2.size();
    P.g2.push_back(gd);
  }
  P.flush.assign(n + 1, 0);
  for (int s = 0; s < n; ++s) {
    P.flush[s] = (int)P.preops.size();
    P.preops.insert(P.preops.end(), pending[s].begin(), pending[s].end());
  }
  P.flush[n] = (int)P.preops.size();
  int maxlv = -1;
  for (int lv : level) maxlv = std::max(maxlv, lv);
  std::vector<std::vector<int>> by_level(maxlv + 1);
  for (int g = 0; g < (int)level.size(); ++g) by_level[level[g]].push_back(g);
  P.sched.clear();
  for (const auto &gs : by_level)
    for (size_t i = 0; i < gs.size(); i += GV_WAVES)
      for (int w = 0; w < GV_WAVES; ++w)
        P.sched.push_back(i + w < gs.size() ? gs[i + w] : -1);
}

struct TNode {
  int site, kind, parent;
  int child[4];
  std::vector<TermEnd> ends;  // prefix forest: terms whose left part ends here
  int store = -1;             // suffix forest: slot receiving this node's environment
  int refs = 0;               // split planning: terms whose path uses this node
  long long work = 0;         // expansions in the subtree (heavy-path choice)
};

// A trie of Pauli paths grown from each start site in one direction.
struct Forest {
  int dir;  // 0: sites increase (prefix forest), 1: sites decrease (suffix forest)
  std::vector<TNode> nodes;
  std::vector<std::array<int, 4>> roots;
  Forest(int n, int d) : dir(d), roots(n) {
    for (auto &r : roots) r = {-1, -1, -1, -1};
  }
  int next(int site) const { return dir == 0 ? site + 1 : site - 1; }
  int find(int parent, int site, int kind) const {
    return parent < 0 ? roots[site][kind] : nodes[parent].child[kind];
  }
  int get(int parent, int site, int kind) {
    int c = find(parent, site, kind);
    if (c >= 0) return c;
    TNode t;
    t.site = site;
    t.kind = kind;
    t.parent = parent;
    for (int &x : t.child) x = -1;
    nodes.push_back(std::move(t));
    c = (int)nodes.size() - 1;
    if (parent < 0) roots[site][kind] = c;
    else nodes[parent].child[kind] = c;
    return c;
  }
  // Path of word w from `from` to `to` (inclusive, in this forest's direction);
  // adds `delta` to every node's refcount; returns the last node.
  int walk(const std::string &w, int from, int to, int delta) {
    int node = -1;
    for (int s = from;; s = next(s)) {
      node = get(node, s, kind_of(w[s]));
      nodes[node].refs += delta;
      if (s == to) break;
    }
    return node;
  }
  // First position (walking from `from` towards `limit`) whose node is
  // absent or unused; returns next(limit) when the whole path exists.
  int first_new(const std::string &w, int from, int limit) const {
    int node = -1;
    for (int s = from;; s = next(s)) {
      node = find(node, s, kind_of(w[s]));
      if (node < 0 || nodes[node].refs <= 0) return s;
      if (s == limit) return next(s);
    }
  }
};

struct TermPath {
  int k, first, last, c;
};

// Split choice: term k's left path is its word on sites [0, c] in the prefix
// forest (leading identities included, so every path starts at the unit
// environment of site 0 and the identity chain is shared), its right path the
// word on sites [c+1, n-1] in the suffix forest (empty for c = n-1), with
// first <= c <= last (first/last non-identity sites).  The number of new
// nodes for a split c is max(0, c - p0 + 1) + max(0, q - c) where the prefix
// forest already holds the path up to p0 - 1 and the suffix forest down to
// q + 1.  Greedy: start every term at its middle non-identity site, then
// re-place each term once at its cheapest split (ties: nearest the middle).
std::vector<TermPath> choose_splits(const Problem &P) {
  const int n = P.n;
  std::vector<TermPath> paths;
  for (int k : P.active) {
    const std::string &w = P.site_words[k];
    std::vector<int> z;
    for (int s = 0; s < n; ++s)
      if (w[s] != 'I') z.push_back(s);
    paths.push_back({k, z.front(), z.back(), z[(z.size() - 1) / 2]});
  }
  Forest L(n, 0), R(n, 1);
  auto insert = [&](const TermPath &t, int delta) {
    const std::string &w = P.site_words[t.k];
    L.walk(w, 0, t.c, delta);
    if (t.c < n - 1) R.walk(w, n - 1, t.c + 1, delta);
  };
  for (const TermPath &t : paths) insert(t, +1);
  for (TermPath &t : paths) {
    insert(t, -1);
    const std::string &w = P.site_words[t.k];
    const int p0 = L.first_new(w, 0, t.last);
    const int q = R.first_new(w, n - 1, t.first);
    const int mid = t.c;
    int best = -1, bc = mid;
    for (int c = t.first; c <= t.last; ++c) {
      const int cost = std::max(0, c - p0 + 1) + std::max(0, q - c);
      if (best < 0 || cost < best || (cost == best && std::abs(c - mid) < std::abs(bc - mid))) {
        best = cost;
        bc = c;
      }
    }
    t.c = bc;
    insert(t, +1);
  }
  return paths;
}

struct ChainOut {
  ChainDev ch;
  int gen;
  std::vector<StepDev> steps;
  std::vector<ChildDev> kids;  // term0 temporarily holds the node id
};

// Heavy-path chain decomposition of one forest.  Each chain keeps its
// environment in registers; a child that starts another chain is spilled and
// that chain belongs to the next generation.
void forest_chains(Forest &F, int n, int gen0, int &nspill, double &flops, std::vector<ChainOut> &out) {
  for (int v = (int)F.nodes.size() - 1; v >= 0; --v) {
    long long w = 0;
    bool any = false;
    for (int c : F.nodes[v].child)
      if (c >= 0) {
        any = true;
        w += F.nodes[c].work;
      }
    F.nodes[v].work = w + (any ? 1 : 0);
  }
  struct Start {
    int slot, site, node, gen;
  };
  std::vector<Start> queue;
  for (int a = 0; a < n; ++a) {
    const auto &r = F.roots[a];
    if (r[0] >= 0 || r[1] >= 0 || r[2] >= 0 || r[3] >= 0) queue.push_back({-1, a, -1, gen0});
  }
  const double chi = GV_CHI;
  for (size_t qi = 0; qi < queue.size(); ++qi) {
    const Start st = queue[qi];
    ChainOut co;
    co.ch = ChainDev{};
    co.ch.site0 = st.site;
    co.ch.src = st.slot;
    co.ch.dir = F.dir;
    co.gen = st.gen;
    std::array<int, 4> kids = st.node < 0 ? F.roots[st.site]
                                          : std::array<int, 4>{F.nodes[st.node].child[0], F.nodes[st.node].child[1],
                                                               F.nodes[st.node].child[2], F.nodes[st.node].child[3]};
    int site = st.site;
    for (;;) {
      StepDev sd{};
      sd.site = site;
      sd.child0 = (int)co.kids.size();
      int cont = -1;
      long long best = 0;
      for (int c : kids)
        if (c >= 0 && F.nodes[c].work > best) {
          best = F.nodes[c].work;
          cont = c;
        }
      // a root chain (the only work of the first launch, a few waves per
      // instance) hands its heavy path on to the next generation after
      // GV_ROOT_STEPS steps, so the first launch is short and its
      // continuation runs beside the other chains of the second generation
      if (st.node < 0 && (int)co.steps.size() + 1 >= GV_ROOT_STEPS) cont = -1;
      for (int kd = 0; kd < 4; ++kd) {
        const int c = kids[kd];
        if (c < 0) continue;
        ChildDev cdv{};
        cdv.kind = kd;
        if (c == cont) cdv.action = -2;
        else if (F.nodes[c].work > 0) {
          cdv.action = nspill++;
          queue.push_back({cdv.action, F.next(site), c, st.gen + 1});
        } else
          cdv.action = -1;
        cdv.nterm = (int)F.nodes[c].ends.size();
        cdv.store = F.nodes[c].store;
        cdv.term0 = c;  // node id until flattened
        co.kids.push_back(cdv);
        sd.need |= (kd <= CK_Z) ? 1 : 2;
      }
      sd.nchild = (int)co.kids.size() - sd.child0;
      flops += (sd.need == 3 ? 48.0 : 32.0) * chi * chi * chi;  // (4 or 6) chi^3 complex MACs
      co.steps.push_back(sd);
      if (cont < 0) break;
      kids = {F.nodes[cont].child[0], F.nodes[cont].child[1], F.nodes[cont].child[2], F.nodes[cont].child[3]};
      site = F.next(site);
    }
    co.ch.nsteps = (int)co.steps.size();
    out.push_back(std::move(co));
  }
}

void build_chains(Problem &P) {
  const int n = P.n;
  const std::vector<TermPath> paths = choose_splits(P);
  Forest L(n, 0), R(n, 1);
  P.nspill = 0;
  std::vector<int> right_leaf(paths.size(), -1);
  for (size_t i = 0; i < paths.size(); ++i) {
    const TermPath &t = paths[i];
    const std::string &w = P.site_words[t.k];
    if (t.c < n - 1) {  // c = n-1: the right environment is the unit [1]
      const int rn = R.walk(w, n - 1, t.c + 1, +1);
      if (R.nodes[rn].store < 0) R.nodes[rn].store = P.nspill++;
      right_leaf[i] = R.nodes[rn].store;
    }
  }
  for (size_t i = 0; i < paths.size(); ++i) {
    const TermPath &t = paths[i];
    const int ln = L.walk(P.site_words[t.k], 0, t.c, +1);
    L.nodes[ln].ends.push_back(TermEnd{t.k, right_leaf[i], P.coeffs[2 * (size_t)t.k]});
  }
  P.n_nodes = (int)L.nodes.size();
  P.n_right_nodes = (int)R.nodes.size();
  // Algorithmic count: 16 (Dl^2 Dr + Dl Dr^2) flops per forest node (the
  // identity environments are the forests' identity chains).
  const std::vector<int> &bd = P.bonds_pred;
  auto fsite = [&](int s) {
    const double dl = bd[s], dr = bd[s + 1];
    return 16.0 * (dl * dl * dr + dl * dr * dr);
  };
  P.flops_nodes = 0;
  for (const TNode &t : L.nodes) P.flops_nodes += fsite(t.site);
  for (const TNode &t : R.nodes) P.flops_nodes += fsite(t.site);

  P.flops_exec = 0.0;
  std::vector<ChainOut> chains;
  forest_chains(R, n, 0, P.nspill, P.flops_exec, chains);
  int right_gens = 0;
  for (const ChainOut &c : chains) right_gens = std::max(right_gens, c.gen + 1);
  const size_t nright = chains.size();
  forest_chains(L, n, right_gens, P.nspill, P.flops_exec, chains);
  // Generations of the complex gradient's forward pass: the suffix forest,
  // then the prefix forest (its inline term ends read suffix-forest slots).
  // Energy launches (and the real gradient's forward pass): the
  // shallow prefix depths with few chains (latency-bound launches) run in the
  // suffix forest's launches of the same depth, with every term end deferred
  // to the leaf pass, which runs after both forests.
  std::vector<int> depth_count;
  for (size_t i = nright; i < chains.size(); ++i) {
    const int d = chains[i].gen - right_gens;
    if ((int)depth_count.size() <= d) depth_count.resize(d + 1, 0);
    ++depth_count[d];
  }
  int merged = 0;
  while (merged < (int)depth_count.size() && merged < right_gens && depth_count[merged] <= GV_MERGE_CHAINS) ++merged;
  std::vector<int> ggen(chains.size()), early(chains.size(), 0);
  for (size_t i = 0; i < chains.size(); ++i) {
    ggen[i] = chains[i].gen;
    if (i >= nright) {
      const int d = chains[i].gen - right_gens;
      if (d < merged) {
        chains[i].gen = d;
        early[i] = 1;
      } else {
        chains[i].gen = right_gens + d - merged;
      }
    }
  }
  // Order: energy launch ascending (suffix forest first within a launch),
  // longest chains first.
  std::vector<int> order(chains.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (chains[x].gen != chains[y].gen) return chains[x].gen < chains[y].gen;
    return chains[x].ch.nsteps > chains[y].ch.nsteps;
  });
  P.chains.clear();
  P.steps.clear();
  P.children.clear();
  P.term_list.clear();
  P.leaves.clear();
  P.gen_begin.clear();
  int last_gen = -1;
  for (int idx : order) {
    const ChainOut &co = chains[idx];
    while (last_gen < co.gen) {
      P.gen_begin.push_back((int)P.chains.size());
      ++last_gen;
    }
    const Forest &F = (size_t)idx < nright ? R : L;
    ChainDev ch = co.ch;
    ch.step0 = (int)P.steps.size();
    for (StepDev sd : co.steps) {
      const int base = (int)P.children.size();
      for (int c = 0; c < sd.nchild; ++c) {
        ChildDev cdv = co.kids[sd.child0 + c];
        const int node = cdv.term0;
        cdv.term0 = (int)P.term_list.size();
        for (const TermEnd &te : F.nodes[node].ends) P.term_list.push_back(te);
        cdv.defer = -1;
        if (cdv.nterm >= GV_DEFER_TERMS || (early[idx] && cdv.nterm > 0)) {
          cdv.defer = (int)P.leaves.size();
          P.leaves.push_back(LeafDev{cdv.term0, cdv.nterm});
        }
        P.children.push_back(cdv);
      }
      sd.child0 = base;
      P.steps.push_back(sd);
    }
    P.chains.push_back(ch);
  }
  P.gen_begin.push_back((int)P.chains.size());
  P.ngen = (int)P.gen_begin.size() - 1;
  // gradient generations over the same chain numbering, longest chains first
  std::vector<int> pos(chains.size());
  for (size_t k = 0; k < order.size(); ++k) pos[order[k]] = (int)k;
  int nggen = 0;
  for (int g : ggen) nggen = std::max(nggen, g + 1);
  std::vector<std::vector<int>> byg(nggen);
  for (int idx : order) byg[ggen[idx]].push_back(pos[idx]);
  P.ggen_begin.assign(1, 0);
  P.ggen_chains.clear();
  for (const auto &v : byg) {
    P.ggen_chains.insert(P.ggen_chains.end(), v.begin(), v.end());
    P.ggen_begin.push_back((int)P.ggen_chains.size());
  }
  P.nggen = nggen;
  // backward launches: a chain's backward needs the adjoints its spilled
  // children hand back (later generations of its own forest) and forward
  // values only (checkpoints, suffix slots, the prefix leaves of the term
  // ends that read its slots), so the two forests' backward passes are
  // independent.  They start together at their deepest generations: the
  // forest with fewer generations then runs its latency-bound root depths
  // (few chains, long heavy paths) beside the other forest's wide ones.
  const int dp = nggen - right_gens, kb = std::max(right_gens, dp);
  std::vector<std::vector<int>> byb(kb);
  for (int idx : order) {
    const int g = ggen[idx];
    byb[(size_t)idx < nright ? g + kb - right_gens : g - right_gens + kb - dp].push_back(pos[idx]);
  }
  P.bgen_begin.assign(1, 0);
  P.bgen_chains.clear();
  for (const auto &v : byb) {
    P.bgen_chains.insert(P.bgen_chains.end(), v.begin(), v.end());
    P.bgen_begin.push_back((int)P.bgen_chains.size());
  }
  P.nbgen = kb;
}

// Per-site programs of the bond-doubling state used by the reverse gradient
// (test/tools/reverse_mps.py:39-57, 152-184): every site tensor is built from
// [1, 0] by its own single-qubit gates and the CNOT halves that touch it, so
// sites are independent.  A neighbouring CNOT doubles the shared bond exactly;
// anything else (SWAP, distant CNOT, a bond beyond the 16-wide slots) leaves
// the problem without a reverse gradient (grad_error says why).
void build_local_programs(Problem &P, const gvmps_op *ops, int n_ops) {
  const int n = P.n;
  std::vector<std::vector<LocalOp>> prog(n);
  std::vector<int> bond(n + 1, 1);
  P.rot_param.clear();
  P.rot_scale.clear();
  P.grad_error.clear();
  auto fail = [&](const std::string &why) {
    if (P.grad_error.empty()) P.grad_error = why;
  };
  for (int k = 0; k < n_ops && P.grad_error.empty(); ++k) {
    const gvmps_op &o = ops[k];
    if (o.kind == GVMPS_RX || o.kind == GVMPS_RY || o.kind == GVMPS_RZ) {
      const int s = n - 1 - o.q0;
      LocalOp lo{};
      lo.kind = LO_ROT;
      lo.arg = o.kind - GVMPS_RX;
      lo.param = o.param;
      lo.scale = o.param >= 0 ? o.scale : 0.0;
      lo.offset = o.offset;
      lo.rot = (int)P.rot_param.size();
      P.rot_param.push_back(o.param);
      P.rot_scale.push_back(lo.scale);
      prog[s].push_back(lo);
    } else if (o.kind == GVMPS_CX) {
      const int sc = n - 1 - o.q0, st = n - 1 - o.q1;
      if (std::abs(sc - st) != 1) {
        fail("reverse gradient: CX(" + std::to_string(o.q0) + "," + std::to_string(o.q1) +
             ") is not between neighbouring qubits (bond doubling needs adjacent CNOTs)");
        break;
      }
      const int lo_s = std::min(sc, st);
      if (2 * bond[lo_s + 1] > GV_CHI) {
        fail("reverse gradient: CNOT bond doubling exceeds the 16-wide bond slots");
        break;
      }
      bond[lo_s + 1] *= 2;
      LocalOp a{}, b{};
      a.param = b.param = -1;
      a.rot = b.rot = -1;
      if (sc == lo_s) {  // control on the left site
        a.kind = LO_CTRL_R;
        b.kind = LO_TARG_L;
      } else {
        a.kind = LO_TARG_R;
        b.kind = LO_CTRL_L;
      }
      prog[lo_s].push_back(a);
      prog[lo_s + 1].push_back(b);
    } else if (o.kind == GVMPS_SWAP) {
      fail("reverse gradient: SWAP has no bond-doubling form");
    } else {
      LocalOp lo{};
      lo.kind = LO_FIX;
      lo.arg = -1;
      const M2 U = fixed_matrix(o.kind);
      // fix table shared with the state-preparation program
      const int count = (int)P.fixes.size() / 8;
      for (int g = 0; g < count && lo.arg < 0; ++g) {
        bool same = true;
        for (int e = 0; e < 4 && same; ++e)
          same = P.fixes[g * 8 + 2 * e] == U[e].real() && P.fixes[g * 8 + 2 * e + 1] == U[e].imag();
        if (same) lo.arg = g;
      }
      if (lo.arg < 0) throw std::runtime_error("internal: fixed gate missing from the table");
      lo.param = lo.rot = -1;
      prog[n - 1 - o.q0].push_back(lo);
    }
  }
  P.lops.clear();
  P.lop_begin.assign(1, 0);
  for (int s = 0; s < n; ++s) {
    int dl = 1, dr = 1;
    // bond dimensions after each op, replayed in order
    for (LocalOp &lo : prog[s]) {
      if (lo.kind == LO_CTRL_R || lo.kind == LO_TARG_R) dr *= 2;
      if (lo.kind == LO_CTRL_L || lo.kind == LO_TARG_L) dl *= 2;
      lo.dl = dl;
      lo.dr = dr;
      P.lops.push_back(lo);
    }
    P.lop_begin.push_back((int)P.lops.size());
  }
  if (!P.grad_error.empty()) {
    P.lops.clear();
    P.lop_begin.assign(n + 1, 0);
  }
  // Trailing CNOT halves of each program, at most one per bond side (each
  // halves one bond of the site adjoint when pulled back).
  P.site_tail.assign(n, 0);
  for (int s = 0; s < n; ++s) {
    int cnt = 0, code = 0;
    bool left = false, right = false;
    for (int o = P.lop_begin[s + 1] - 1; o >= P.lop_begin[s] && cnt < 2; --o) {
      const int k = P.lops[o].kind;
      const bool isr = k == LO_CTRL_R || k == LO_TARG_R, isl = k == LO_CTRL_L || k == LO_TARG_L;
      if (!(isr || isl) || (isr && right) || (isl && left)) break;
      right = right || isr;
      left = left || isl;
      code |= k << (4 + 4 * cnt);
      ++cnt;
    }
    P.site_tail[s] = cnt | code;
  }
  std::vector<std::vector<int>> by_param(P.nparams);
  for (int r = 0; r < (int)P.rot_param.size(); ++r)
    if (P.rot_param[r] >= 0) by_param[P.rot_param[r]].push_back(r);
  P.param_rot_begin.assign(1, 0);
  P.param_rots.clear();
  for (const auto &v : by_param) {
    P.param_rots.insert(P.param_rots.end(), v.begin(), v.end());
    P.param_rot_begin.push_back((int)P.param_rots.size());
  }
}

// Reverse-gradient metadata of the forests: prefix leaves whose term ends use
// a suffix-forest slot keep their environment (gleaf) so the backward pass can
// seed that slot's adjoint with sum_k Re(c_k) conj(E_left(k)); steps are listed
// per site for the fixed-order site-adjoint reduction.
void build_grad_meta(Problem &P) {
  P.ngleaf = 0;
  std::vector<std::vector<SeedTerm>> seeds(P.nspill);
  for (const ChainDev &ch : P.chains) {
    if (ch.dir != 0) continue;
    for (int st = 0; st < ch.nsteps; ++st) {
      const StepDev &sd = P.steps[ch.step0 + st];
      for (int c = 0; c < sd.nchild; ++c) {
        ChildDev &cd = P.children[sd.child0 + c];
        cd.gleaf = -1;
        for (int t = 0; t < cd.nterm; ++t) {
          const TermEnd &te = P.term_list[cd.term0 + t];
          if (te.rslot < 0) continue;
          if (cd.gleaf < 0) cd.gleaf = P.ngleaf++;
          seeds[te.rslot].push_back(SeedTerm{te.term, cd.gleaf});
        }
      }
    }
  }
  for (const ChainDev &ch : P.chains)
    if (ch.dir != 0)
      for (int st = 0; st < ch.nsteps; ++st) {
        const StepDev &sd = P.steps[ch.step0 + st];
        for (int c = 0; c < sd.nchild; ++c) P.children[sd.child0 + c].gleaf = -1;
      }
  P.seed_begin.assign(1, 0);
  P.seed_terms.clear();
  for (const auto &v : seeds) {
    P.seed_terms.insert(P.seed_terms.end(), v.begin(), v.end());
    P.seed_begin.push_back((int)P.seed_terms.size());
  }
  std::vector<std::vector<int>> per_site(P.n);
  for (int s = 0; s < (int)P.steps.size(); ++s) per_site[P.steps[s].site].push_back(s);
  P.site_steps_begin.assign(1, 0);
  P.site_steps.clear();
  for (const auto &v : per_site) {
    P.site_steps.insert(P.site_steps.end(), v.begin(), v.end());
    P.site_steps_begin.push_back((int)P.site_steps.size());
  }
  // forward with stored step inputs (6 products per full step) plus a
  // backward of 8 (I/Z or X/Y only: 4 and 6)
  P.flops_grad = 0;
  const double prod = 8.0 * 16 * 16 * 16;
  for (const StepDev &sd : P.steps) {
    const bool iz = sd.need & 1, xy = sd.need & 2;
    const int nf = 2 + (iz ? 2 : 0) + (xy ? 2 : 0);
    const int nb = (iz && xy) ? 8 : 6;
    P.flops_grad += (nf + nb) * prod;
  }
}

// Every gate matrix of the tape is real: RY rotations (RX and RZ only as
// constant identities), fixed single-site gates and two-site gates with zero
// imaginary parts.  Applied to |0...0>, such a tape keeps the MPS real.
bool real_circuit(const Problem &P) {
  for (const ProgOp &op : P.prog) {
    if (op.kind == OP_ROT) {
      if (op.arg != 1 && !(op.param < 0 && op.offset == 0.0)) return false;
    } else if (op.kind == OP_FIX1) {
      for (int e = 0; e < 4; ++e)
        if (P.fixes[(size_t)op.arg * 8 + 2 * e + 1] != 0.0) return false;
    } else {
      for (int e = 0; e < 16; ++e)
        if (P.gates[(size_t)op.arg * 32 + 2 * e + 1] != 0.0) return false;
    }
  }
  return true;
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
  // With zero cutoffs the rank rule keeps min(2 Dl, 2 Dr, max_bond) columns
  // (dist_svd.cuh:194-212).  When no gate of the tape is cut below
  // min(2 Dl, 2 Dr), every split keeps the whole of Theta, no singular value
  // is compared with anything (the guard needs max_bond < count), and the
  // state does not depend on which exact factorisation is used.
  P->exact_split = abs_cut == 0.0 && rel_cut == 0.0 && !P->rank_reduced;
  build_schedule(*P);
  build_local_programs(*P, ops, n_ops);
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
  P->real_ok = real_circuit(*P) &&
               std::none_of(P->site_words.begin(), P->site_words.end(),
                            [](const std::string &w) { return w.find('Y') != std::string::npos; });
  build_chains(*P);
  P->iz_only = std::none_of(P->steps.begin(), P->steps.end(), [](const StepDev &sd) { return (sd.need & 2) != 0; });
  build_grad_meta(*P);
  return P;
}

}  // namespace gv
