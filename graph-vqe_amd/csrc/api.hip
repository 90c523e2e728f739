// C ABI of the gvmps library (declared in include/gvmps.h).
//
// Problems are host objects (plan.cpp).  A batch uploads the problems it
// uses, lays out the per-instance arenas in HBM and launches the kernels on
// the caller's stream.  Errors become a nonzero return plus a thread-local
// message; nothing here exits the process.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "api_internal.h"

using namespace gv;


namespace {
thread_local std::string g_error;
}

// Scratch of gvmps_batch_shot_energy_noalloc: the packed shot words (two
// 64-bit words per shot per instance) between the sampling and the Z-sum
// kernels.  A file-scope device array, so that entry point needs no
// allocation: every GPU the library runs on holds its own 32 MiB copy until
// the process ends.  Host code must not pass the array's name to a kernel
// (on the host it is only a stand-in for the symbol); it uses the device
// address that hipGetSymbolAddress returns, looked up once per device.
#define GV_SCRATCH_WORDS (4u << 20)
#define GV_SCRATCH_DEVICES 64
static __device__ __align__(256) unsigned long long scratch[GV_SCRATCH_WORDS];

namespace {
std::mutex g_scratch_mu;
unsigned long long *g_scratch_ptr[GV_SCRATCH_DEVICES] = {};

// Device address of this device's copy of `scratch`; `device` must be current.
unsigned long long *scratch_address(int device) {
  if (device < 0 || device >= GV_SCRATCH_DEVICES) throw std::runtime_error("device index out of range for scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    gv::hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[device] = static_cast<unsigned long long *>(ptr);
  }
  return g_scratch_ptr[device];
}
}  // namespace

namespace gv {

void set_error(const std::string &msg) { g_error = msg; }

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
static T *upload(std::vector<void *> &owned, const T *host, size_t count) {
  const size_t n = count;
  if (count == 0) count = 1;  // never a null table
  void *d = nullptr;
  hip_check(hipMalloc(&d, sizeof(T) * count), "hipMalloc");
  owned.push_back(d);
  if (host && n > 0) hip_check(hipMemcpy(d, host, sizeof(T) * n, hipMemcpyHostToDevice), "hipMemcpy H2D");
  else hip_check(hipMemset(d, 0, sizeof(T) * count), "hipMemset");
  return static_cast<T *>(d);
}

Batch::~Batch() {
  if (device >= 0) hipSetDevice(device);
  if (estore) hipFree(estore);
  if (graph_exec) hipGraphExecDestroy(graph_exec);
  if (graph) hipGraphDestroy(graph);
  if (cap_stream) hipStreamDestroy(cap_stream);
  for (hipEvent_t &e : ev)
    if (e) hipEventDestroy(e);
  for (hipEvent_t &e : kev)
    if (e) hipEventDestroy(e);
  for (void *p : owned) hipFree(p);
  if (sbits) hipFree(sbits);
  if (spacked) hipFree(spacked);
}

// Phase boundaries: event i is recorded before phase i; the phase lengths
// are read after the stream synchronises (collect).
void Batch::mark(int i, hipStream_t s) {
  if (!timing) return;
  if (!ev[i]) hip_check(hipEventCreate(&ev[i]), "hipEventCreate");
  hip_check(hipEventRecord(ev[i], s), "hipEventRecord");
  ev_used = std::max(ev_used, i + 1);
}

void Batch::kbegin(hipStream_t s) {
  if (!timing) return;
  if ((int)kev.size() < kev_used + 2) kev.resize(kev_used + 2, nullptr);
  if (!kev[kev_used]) hip_check(hipEventCreate(&kev[kev_used]), "hipEventCreate");
  hip_check(hipEventRecord(kev[kev_used], s), "hipEventRecord");
}

void Batch::kend(hipStream_t s) {
  if (!timing) return;
  if (!kev[kev_used + 1]) hip_check(hipEventCreate(&kev[kev_used + 1]), "hipEventCreate");
  hip_check(hipEventRecord(kev[kev_used + 1], s), "hipEventRecord");
  kev_used += 2;
}

void Batch::collect() {
  if (!timing) return;
  for (int i = 0; i + 1 < kev_used; i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, kev[i], kev[i + 1]) == hipSuccess) {
      kern_ms += ms;
      ++kern_launches;
    }
  }
  kev_used = 0;
  if (ev_used < 2) return;
  for (int i = 0; i + 1 < ev_used; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) phase_ms[i] += ms;
  }
  ++phase_calls;
  ev_used = 0;
}

// Longest chains first within each XCD's share of the blocks.  The chain
// kernels give XCD x (block b runs on XCD b % 8) a contiguous range of
// logical blocks of `wpb` items, so sorting inside each range keeps every
// item on the XCD, and in the L2, it has in instance-major order while the
// long chains start first and the short ones fill the launch's tail.
template <typename Len>
static void xcd_longest_first(std::vector<int2> &items, int wpb, Len len) {
  const int n = (int)items.size(), nb = (n + wpb - 1) / wpb, q8 = nb >> 3, r8 = nb & 7;
  for (int x = 0; x < 8; ++x) {
    const int b0 = x * q8 + std::min(x, r8), b1 = b0 + q8 + (x < r8 ? 1 : 0);
    const int i0 = std::min(n, b0 * wpb), i1 = std::min(n, b1 * wpb);
    std::stable_sort(items.begin() + i0, items.begin() + i1,
                     [&](const int2 &a, const int2 &b) { return len(a) > len(b); });
  }
}

static auto chain_len(const Batch &B) {
  return [&B](const int2 &a) { return B.probs[B.inst[a.x].prob]->chains[a.y].nsteps; };
}

Batch *make_batch(gvmps_problem *const *problems, const int *counts, int n_problems, int device) {
  if (n_problems <= 0) throw std::runtime_error("empty batch");
  int ndev = 0;
  hip_check(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
  if (device < 0 || device >= ndev) throw std::runtime_error("invalid HIP device");
  hip_check(hipSetDevice(device), "hipSetDevice");
  auto B = std::make_unique<Batch>();
  B->device = device;
  // problem tables
  std::vector<ProbDev> pd(n_problems);
  for (int j = 0; j < n_problems; ++j) {
    const Problem &P = *problems[j]->p;
    B->probs.push_back(problems[j]->p);
    ProbDev &d = pd[j];
    d.n = P.n;
    d.nparams = P.nparams;
    d.nprog = (int)P.prog.size();
    d.max_bond = P.max_bond;
    d.abs_cut = P.abs_cut;
    d.rel_cut = P.rel_cut;
    d.nterms = P.nterms;
    d.nspill = P.nspill;
    d.nchains = (int)P.chains.size();
    d.prog = upload(B->owned, P.prog.data(), P.prog.size());
    d.ngates = (int)P.g2.size();
    d.nsched = (int)P.sched.size() / GV_WAVES;
    d.g2 = upload(B->owned, P.g2.data(), P.g2.size());
    d.sched = upload(B->owned, P.sched.data(), P.sched.size());
    d.preops = upload(B->owned, P.preops.data(), P.preops.size());
    d.flush = upload(B->owned, P.flush.data(), P.flush.size());
    d.gates = upload(B->owned, P.gates.data(), P.gates.size());
    d.fixes = upload(B->owned, P.fixes.data(), P.fixes.size());
    d.chains = upload(B->owned, P.chains.data(), P.chains.size());
    d.steps = upload(B->owned, P.steps.data(), P.steps.size());
    {
      // one zero record past the end: chain loops fetch a step's first two
      // child records unconditionally, one step ahead
      std::vector<ChildDev> ch(P.children);
      ch.push_back(ChildDev{});
      d.children = upload(B->owned, ch.data(), ch.size());
    }
    d.term_list = upload(B->owned, P.term_list.data(), P.term_list.size());
    d.coeffs = upload(B->owned, P.coeffs.data(), P.coeffs.size());
    std::vector<int> act(P.nterms, 0);
    for (int k : P.active) act[k] = 1;
    d.active = upload(B->owned, act.data(), act.size());
    d.nsteps = (int)P.steps.size();
    d.ngleaf = P.ngleaf;
    d.nrot = (int)P.rot_param.size();
    d.lops = upload(B->owned, P.lops.data(), P.lops.size());
    d.lop_begin = upload(B->owned, P.lop_begin.data(), P.lop_begin.size());
    d.seed_begin = upload(B->owned, P.seed_begin.data(), P.seed_begin.size());
    d.seed_terms = upload(B->owned, P.seed_terms.data(), P.seed_terms.size());
    d.site_steps_begin = upload(B->owned, P.site_steps_begin.data(), P.site_steps_begin.size());
    d.site_steps = upload(B->owned, P.site_steps.data(), P.site_steps.size());
    d.site_tail = upload(B->owned, P.site_tail.data(), P.site_tail.size());
    d.param_rot_begin = upload(B->owned, P.param_rot_begin.data(), P.param_rot_begin.size());
    d.param_rots = upload(B->owned, P.param_rots.data(), P.param_rots.size());
    d.rot_scale = upload(B->owned, P.rot_scale.data(), P.rot_scale.size());
    d.nleaf = (int)P.leaves.size();
    d.leaves = upload(B->owned, P.leaves.data(), P.leaves.size());
    d.exact_split = P.exact_split ? 1 : 0;
  }
  B->h_probs = pd;
  B->d_probs = upload(B->owned, pd.data(), pd.size());
  // Z masks (site order) for the shot estimator
  std::vector<const unsigned long long *> mask_ptrs(n_problems);
  for (int j = 0; j < n_problems; ++j) {
    const Problem &P = *problems[j]->p;
    std::vector<unsigned long long> m(2 * (size_t)P.nterms, 0ULL);
    bool zonly = true;
    for (int k = 0; k < P.nterms; ++k)
      for (int s = 0; s < P.n; ++s) {
        const char c = P.site_words[k][s];
        if (c == 'Z') m[2 * k + (s >> 6)] |= 1ULL << (s & 63);
        else if (c != 'I') zonly = false;
      }
    B->zonly.push_back(zonly ? 1 : 0);
    mask_ptrs[j] = upload(B->owned, m.data(), m.size());
    B->max_terms = std::max(B->max_terms, P.nterms);
  }
  B->d_masks = upload(B->owned, mask_ptrs.data(), mask_ptrs.size());
  // instances
  long long site = 0, env = 0, spill = 0, term = 0, param = 0, part = 0, rot = 0, step = 0, gleaf = 0, leaf = 0;
  std::vector<int2> leaf_items;
  for (int j = 0; j < n_problems; ++j) {
    const Problem &P = *problems[j]->p;
    if (counts[j] < 0) throw std::runtime_error("negative instance count");
    for (int c = 0; c < counts[j]; ++c) {
      InstDev in{};
      in.prob = j;
      in.n = P.n;
      in.site_off = site;
      in.env_off = env;
      in.spill_off = spill;
      in.term_off = term;
      in.param_off = param;
      in.part_off = part;
      in.out = (int)B->inst.size();
      in.rot_off = rot;
      in.step_off = step;
      in.gleaf_off = gleaf;
      in.leaf_off = leaf;
      for (int l = 0; l < (int)P.leaves.size(); ++l) leaf_items.push_back(make_int2((int)B->inst.size(), l));
      B->inst.push_back(in);
      leaf += (long long)P.leaves.size();
      rot += (long long)P.rot_param.size();
      step += (long long)P.steps.size();
      gleaf += P.ngleaf;
      B->max_n = std::max(B->max_n, P.n);
      B->max_params = std::max(B->max_params, P.nparams);
      B->max_chains = std::max(B->max_chains, (int)P.chains.size());
      site += (long long)P.n * GV_SLOT;
      env += 2LL * (P.n + 1) * GV_ENV;
      spill += (long long)P.nspill * GV_ENV;
      term += P.nterms;
      param += P.nparams;
      part += (long long)P.chains.size() + (long long)P.leaves.size();
      B->max_gen = std::max(B->max_gen, P.ngen);
      B->max_ggen = std::max(B->max_ggen, P.nggen);
      B->max_bgen = std::max(B->max_bgen, P.nbgen);
    }
  }
  if (B->inst.empty()) throw std::runtime_error("empty batch");
  B->n_terms_total = term;
  B->n_params_total = param;
  B->n_rot_total = rot;
  B->n_steps_total = step;
  B->n_gleaf_total = gleaf;
  B->n_site_total = site;
  B->n_spill_total = spill;
  B->d_inst = upload(B->owned, B->inst.data(), B->inst.size());
  B->arena = upload<double2>(B->owned, nullptr, 2 * site);
  B->real = std::all_of(B->probs.begin(), B->probs.end(), [](const Problem *P) { return P->real_ok; });
  if (B->real) {
    B->rarena = upload<double>(B->owned, nullptr, 2 * site);
    B->smask = upload<unsigned>(B->owned, nullptr, 2 * (size_t)(site / GV_SLOT));
    B->rfarena = upload<double>(B->owned, nullptr, 2 * (size_t)(site / GV_SLOT) * GV_CHI * GV_CHI);
    B->cflags = upload<int>(B->owned, nullptr, std::max(1LL, part));
  }
  B->iz = std::all_of(B->probs.begin(), B->probs.end(), [](const Problem *P) { return P->iz_only; });
  B->envs = upload<double2>(B->owned, nullptr, env);
  B->spills = upload<double2>(B->owned, nullptr, spill);
  B->terms = upload<double2>(B->owned, nullptr, term);
  B->status = upload<int>(B->owned, nullptr, B->inst.size());
  B->status_host.assign(B->inst.size(), 0);
  B->bonds = upload<int>(B->owned, nullptr, B->inst.size() * (GV_NMAX + 1));
  B->stats = upload<int>(B->owned, nullptr, B->inst.size() * GV_STATS);
  B->energies = upload<double>(B->owned, nullptr, B->inst.size());
  B->seeds = upload<unsigned>(B->owned, nullptr, B->inst.size());
  B->parts = upload<double>(B->owned, nullptr, part);
  B->leafs = upload<double2>(B->owned, nullptr, (size_t)leaf * GV_ENV);
  // the leaf kernels use the XCD-aware block order of the chain kernels (4
  // waves per block): each instance's leaves run on the XCD its chains used,
  // the leaves with most term ends first
  xcd_longest_first(leaf_items, 4,
                    [&](const int2 &a) { return B->probs[B->inst[a.x].prob]->leaves[a.y].nterm; });
  B->leaf_items = upload(B->owned, leaf_items.data(), leaf_items.size());
  B->n_leaf_items = (int)leaf_items.size();
  {
    std::vector<LeafItem> desc(leaf_items.size());
    for (size_t k = 0; k < leaf_items.size(); ++k) {
      const InstDev &in = B->inst[leaf_items[k].x];
      const Problem &P = *B->probs[in.prob];
      const LeafDev &lf = P.leaves[leaf_items[k].y];
      LeafItem &d = desc[k];
      d.tl = pd[in.prob].term_list + lf.term0;
      d.env = (in.leaf_off + leaf_items[k].y) * GV_ENV;
      d.spill = in.spill_off;
      d.part = in.part_off + (long long)P.chains.size() + leaf_items[k].y;
      d.tv = in.term_off;
      d.nterm = lf.nterm;
    }
    B->leaf_desc = upload(B->owned, desc.data(), desc.size());
  }
  // chain work items per generation, instance-major, and the same items
  // longest chains first within each XCD's blocks
  for (int g = 0; g < B->max_gen; ++g) {
    std::vector<int2> items;
    for (size_t i = 0; i < B->inst.size(); ++i) {
      const Problem &P = *B->probs[B->inst[i].prob];
      if (g >= P.ngen) continue;
      for (int c = P.gen_begin[g]; c < P.gen_begin[g + 1]; ++c) items.push_back(make_int2((int)i, c));
    }
    B->items.push_back(upload(B->owned, items.data(), items.size()));
    B->n_items.push_back((int)items.size());
    B->items_host.push_back(items);
    xcd_longest_first(items, B->wpb, chain_len(*B));
    B->items_len.push_back(upload(B->owned, items.data(), items.size()));
  }
  for (int g = 0; g < B->max_ggen; ++g) {
    std::vector<int2> items;
    for (size_t i = 0; i < B->inst.size(); ++i) {
      const Problem &P = *B->probs[B->inst[i].prob];
      if (g >= P.nggen) continue;
      for (int q = P.ggen_begin[g]; q < P.ggen_begin[g + 1]; ++q)
        items.push_back(make_int2((int)i, P.ggen_chains[q]));
    }
    B->gitems.push_back(upload(B->owned, items.data(), items.size()));
    B->gn_items.push_back((int)items.size());
    B->gitems_host.push_back(items);
    xcd_longest_first(items, B->wpb, chain_len(*B));
    B->gitems_len.push_back(upload(B->owned, items.data(), items.size()));
  }
  for (int g = 0; g < B->max_bgen; ++g) {
    std::vector<int2> items;
    for (size_t i = 0; i < B->inst.size(); ++i) {
      const Problem &P = *B->probs[B->inst[i].prob];
      if (g >= P.nbgen) continue;
      for (int q = P.bgen_begin[g]; q < P.bgen_begin[g + 1]; ++q)
        items.push_back(make_int2((int)i, P.bgen_chains[q]));
    }
    B->bitems.push_back(upload(B->owned, items.data(), items.size()));
    B->bn_items.push_back((int)items.size());
    B->bitems_host.push_back(items);
    xcd_longest_first(items, B->wpb, chain_len(*B));
    B->bitems_len.push_back(upload(B->owned, items.data(), items.size()));
  }
  B->set_checkpoint(1);
  return B.release();
}

// Checkpoint layout of the step input environments: in every chain, steps
// 0, k, 2k, ... keep their input environment (compact slots per problem);
// the backward pass recomputes the others from the nearest kept one.
void Batch::set_checkpoint(int stride) {
  if (stride < 1 || stride > 64) throw std::runtime_error("checkpoint stride must be in 1..64");
  std::vector<long long> n_ck(probs.size(), 0);
  ck_slots.assign(probs.size(), nullptr);
  for (size_t j = 0; j < probs.size(); ++j) {
    const Problem &P = *probs[j];
    std::vector<int> cs(P.steps.size(), -1);
    int count = 0;
    for (const ChainDev &ch : P.chains)
      for (int st = 0; st < ch.nsteps; ++st)
        if (st % stride == 0) cs[ch.step0 + st] = count++;
    n_ck[j] = count;
    ck_slots[j] = upload(owned, cs.data(), cs.size());
    h_probs[j].ck_slot = ck_slots[j];
  }
  hip_check(hipMemcpy(d_probs, h_probs.data(), sizeof(ProbDev) * h_probs.size(), hipMemcpyHostToDevice),
            "upload problem table");
  long long off = 0;
  for (InstDev &in : inst) {
    in.ck_off = off;
    off += n_ck[in.prob];
  }
  hip_check(hipMemcpy(d_inst, inst.data(), sizeof(InstDev) * inst.size(), hipMemcpyHostToDevice),
            "upload instance table");
  n_ck_total = off;
  ck_stride = stride;
  if (estore) {
    hip_check(hipFree(estore), "hipFree");
    estore = nullptr;
  }
  if (grad_ready) {
    hip_check(hipMalloc(&estore, sizeof(double2) * (size_t)std::max(1LL, n_ck_total) * GV_ENV), "hipMalloc");
  }
}

void check_status(Batch &B, hipStream_t stream) {
  std::vector<int> &st = B.status_host;
  hip_check(hipMemcpyAsync(st.data(), B.status, sizeof(int) * st.size(), hipMemcpyDeviceToHost, stream),
            "hipMemcpyAsync status");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  B.collect();
  for (size_t i = 0; i < st.size(); ++i) {
    if (st[i] == ST_GUARD) {
      const Problem &P = *B.probs[B.inst[i].prob];
      throw std::runtime_error("CUDAQ_MPS_MAX_BOND=" + std::to_string(P.max_bond) +
                               " is insufficient for zero-cutoff simulation: resolved singular values would be "
                               "discarded");  // dist_svd.cuh:277-279
    }
    if (st[i] != ST_OK)
      throw std::runtime_error("Direct MPS SVD failed or produced nonfinite singular values; devInfo=0");
  }
}

void launch_prepare(Batch &B, const double *params, hipStream_t stream) {
  const int ninst = (int)B.inst.size();
  B.mark(0, stream);
  if (B.prep_mode == 1)
    hip_check(kl_build_local(ninst, B.max_n, stream, B.d_inst, B.d_probs, params, B.arena, nullptr, B.status),
              "k_build_local launch");
  else
    hip_check(kl_prep(ninst, B.real && B.arith == 0, B.prep_mode == 2, stream, B.d_inst, B.d_probs, params, B.arena,
                      B.bonds, B.status, B.stats),
              "k_prep launch");
  B.mark(1, stream);
}

// terms == true: per-term <P_k> into B.terms (observe's per-term data);
// otherwise chain partial energies only.
void launch_energy_tail(Batch &B, hipStream_t stream, bool terms) {
  const int ninst = (int)B.inst.size();
  B.mark(2, stream);
  if (B.real && B.arith == 0) {
    // real circuits, no Y words: the same forests in real arithmetic
    hip_check(kl_pack_real(B.n_site_total, stream, B.arena, B.rarena), "k_pack_real launch");
    if (B.use_smask) {
      hip_check(kl_site_masks(B.n_site_total / GV_SLOT, stream, B.rarena, B.smask, B.rfarena), "k_site_masks launch");
      hip_check(kl_chain_flags(ninst, B.max_chains, stream, B.d_inst, B.d_probs, B.smask, B.cflags, 0),
                "k_chain_flags launch");
      B.mask_kind = 0;
    }
    for (int g = 0; g < B.max_gen; ++g) {
      const int ni = B.n_items[g];
      if (ni == 0) continue;
      B.kbegin(stream);
      hip_check(kl_rchains(ni, B.wpb, B.iz, stream, B.gen_items(g), B.d_inst, B.d_probs, B.rarena, B.n_site_total,
                           B.use_smask ? B.smask : nullptr, B.rfarena, B.cflags, B.leafs,
                           B.spills,
                           terms ? B.terms : nullptr, B.parts),
                "k_rchains launch");
      B.kend(stream);
    }
    hip_check(kl_rleaf_terms(B.n_leaf_items, stream, B.leaf_desc, B.leafs, B.spills,
                             terms ? B.terms : nullptr, B.parts),
              "k_rleaf_terms launch");
    B.mark(3, stream);
    if (terms) hip_check(kl_reduce(ninst, stream, B.d_inst, B.d_probs, B.terms, B.energies), "k_reduce launch");
    else hip_check(kl_reduce_energy(ninst, stream, B.d_inst, B.d_probs, B.parts, B.energies, 1), "k_reduce launch");
    B.mark(4, stream);
    return;
  }
  hip_check(kl_transpose_sites(B.n_site_total, stream, B.arena), "k_transpose_sites launch");
  for (int g = 0; g < B.max_gen; ++g) {
    const int ni = B.n_items[g];
    if (ni == 0) continue;
    B.kbegin(stream);
    if (terms)
      hip_check(kl_chains(ni, B.wpb, B.iz, stream, B.gen_items(g), B.d_inst, B.d_probs, B.arena, B.n_site_total, B.leafs,
                          B.spills, B.terms),
                "k_chains launch");
    else
      hip_check(kl_chains_energy(ni, B.wpb, B.iz, stream, B.gen_items(g), B.d_inst, B.d_probs, B.arena, B.n_site_total,
                                 B.leafs, B.spills,
                                 B.parts),
                "k_chains_energy launch");
    B.kend(stream);
  }
  hip_check(kl_leaf_terms(B.n_leaf_items, stream, B.leaf_items, B.d_inst, B.d_probs, B.leafs, B.spills,
                          terms ? B.terms : nullptr, B.parts),
            "k_leaf_terms launch");
  B.mark(3, stream);
  if (terms) hip_check(kl_reduce(ninst, stream, B.d_inst, B.d_probs, B.terms, B.energies), "k_reduce launch");
  else hip_check(kl_reduce_energy(ninst, stream, B.d_inst, B.d_probs, B.parts, B.energies, 1), "k_reduce launch");
  B.mark(4, stream);
}

// Prepare + identity environments + perfect sampling into the batch's own
// (or the caller's) buffers.  Returns the packed-word buffer used.
unsigned long long *launch_sample(Batch &B, const double *params, int shots, const uint32_t *seeds,
                                  unsigned char *bits, unsigned long long *packed, hipStream_t stream) {
  if (shots <= 0) throw std::runtime_error("shared bitstring sampling requires a positive shot count");
  if (!seeds) throw std::runtime_error("null seed array");
  const int ninst = (int)B.inst.size();
  int max_n = 0;
  for (const InstDev &in : B.inst) max_n = std::max(max_n, in.n);
  const size_t nbits = (size_t)ninst * shots * max_n, nwords = (size_t)ninst * shots * 2;
  if (!bits) {
    if (B.sbits_cap < nbits) {
      if (B.sbits) hip_check(hipFree(B.sbits), "hipFree");
      hip_check(hipMalloc(&B.sbits, nbits), "hipMalloc bits");
      B.sbits_cap = nbits;
    }
    bits = B.sbits;
  }
  if (!packed) {
    if (B.spacked_cap < nwords) {
      if (B.spacked) hip_check(hipFree(B.spacked), "hipFree");
      hip_check(hipMalloc(&B.spacked, nwords * sizeof(unsigned long long)), "hipMalloc packed");
      B.spacked_cap = nwords;
    }
    packed = B.spacked;
  }
  hip_check(hipMemcpyAsync(B.seeds, seeds, sizeof(unsigned) * ninst, hipMemcpyHostToDevice, stream), "seeds H2D");
  launch_prepare(B, params, stream);
  launch_sample_tail(B, shots, B.seeds, bits, packed, stream, true);
  return packed;
}

// Right environments and the shot sweep of a prepared batch.  Real batches
// (every problem real, real arithmetic selected) run the real kernels on the
// packed real arena, packed here unless the energy tail already did.
void launch_sample_tail(Batch &B, int shots, const unsigned *seeds_dev, unsigned char *bits,
                        unsigned long long *packed, hipStream_t stream, bool pack) {
  const int ninst = (int)B.inst.size();
  if (B.real && B.arith == 0) {
    if (pack) hip_check(kl_pack_real(B.n_site_total, stream, B.arena, B.rarena), "k_pack_real launch");
    hip_check(kl_renvs_right(ninst, stream, B.d_inst, B.rarena, B.n_site_total, B.envs), "k_renvs_right launch");
    B.mark(2, stream);
    hip_check(kl_sample_real(ninst, shots, B.max_n, stream, B.d_inst, B.rarena, B.envs, seeds_dev, bits, packed),
              "k_sample_real launch");
  } else {
    hip_check(kl_envs(ninst, stream, B.d_inst, B.arena, B.envs), "k_envs launch");
    B.mark(2, stream);
    hip_check(kl_sample(ninst, shots, B.max_n, stream, B.d_inst, B.arena, B.envs, seeds_dev, bits, packed),
              "k_sample launch");
  }
  B.mark(3, stream);
}

// Energies and, for shots > 0, samples of every instance from one state
// preparation.  This is the sequence gvmps_batch_capture records: kernels and
// one device copy only (no memsets, no host copies), so it replays safely.
void launch_evaluate(Batch &B, const double *params, int shots, const unsigned *seeds_dev, double *energies_dev,
                     unsigned char *bits, unsigned long long *packed, hipStream_t stream) {
  const int ninst = (int)B.inst.size();
  if (shots > 0 && (!seeds_dev || !bits || !packed))
    throw std::runtime_error("sampling needs device seeds and both bit buffers");
  launch_prepare(B, params, stream);
  launch_energy_tail(B, stream, false);
  if (energies_dev)
    hip_check(hipMemcpyAsync(energies_dev, B.energies, sizeof(double) * ninst, hipMemcpyDeviceToDevice, stream),
              "copy energies");
  // the energy tail packed the real arena when the batch is real
  if (shots > 0) launch_sample_tail(B, shots, seeds_dev, bits, packed, stream, false);
  B.mark(5, stream);
}

void capture_evaluate(Batch &B, const double *params, int shots, const unsigned *seeds_dev, double *energies_dev,
                      unsigned char *bits, unsigned long long *packed) {
  if (B.graph_exec) hip_check(hipGraphExecDestroy(B.graph_exec), "hipGraphExecDestroy");
  if (B.graph) hip_check(hipGraphDestroy(B.graph), "hipGraphDestroy");
  B.graph_exec = nullptr;
  B.graph = nullptr;
  if (!B.cap_stream) hip_check(hipStreamCreateWithFlags(&B.cap_stream, hipStreamNonBlocking), "hipStreamCreate");
  const bool timing = B.timing;
  B.timing = false;  // no event records inside the graph
  hip_check(hipStreamBeginCapture(B.cap_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  try {
    launch_evaluate(B, params, shots, seeds_dev, energies_dev, bits, packed, B.cap_stream);
  } catch (...) {
    hipGraph_t g = nullptr;
    hipStreamEndCapture(B.cap_stream, &g);
    if (g) hipGraphDestroy(g);
    B.timing = timing;
    throw;
  }
  B.timing = timing;
  hip_check(hipStreamEndCapture(B.cap_stream, &B.graph), "hipStreamEndCapture");
  hip_check(hipGraphInstantiate(&B.graph_exec, B.graph, nullptr, nullptr, 0), "hipGraphInstantiate");
}

// Reverse gradient over the bond-doubling state.  Phases (timing slots):
// 0 local build, 1 forward chains, 2 energy reduce, 3 backward chains,
// 4 site-adjoint reduce, 5 local backward + parameter sums.
void launch_value_grad(Batch &B, const double *params, hipStream_t stream) {
  for (const Problem *P : B.probs)
    if (!P->grad_error.empty()) throw std::runtime_error(P->grad_error);
  if (!params && B.n_params_total > 0) throw std::runtime_error("null parameter array");
  const int ninst = (int)B.inst.size();
  if (!B.grad_ready) {
    B.garena = upload<double2>(B.owned, nullptr, 2 * B.n_site_total);
    B.abar = upload<double2>(B.owned, nullptr, B.n_site_total);
    B.rotin = upload<double2>(B.owned, nullptr, (size_t)B.n_rot_total * GV_SLOT);
    B.rgrad = upload<double>(B.owned, nullptr, B.n_rot_total);
    hip_check(hipMalloc(&B.estore, sizeof(double2) * (size_t)std::max(1LL, B.n_ck_total) * GV_ENV), "hipMalloc");
    B.aslot = upload<double2>(B.owned, nullptr, (size_t)B.n_steps_total * GV_SLOT);
    B.gleafs = upload<double2>(B.owned, nullptr, (size_t)B.n_gleaf_total * GV_ENV);
    B.adjspill = upload<double2>(B.owned, nullptr, B.n_spill_total);
    B.grads = upload<double>(B.owned, nullptr, B.n_params_total);
    B.grad_ready = true;
  }
  B.mark(0, stream);
  hip_check(kl_build_local(ninst, B.max_n, stream, B.d_inst, B.d_probs, params, B.garena, B.rotin, nullptr),
            "k_build_local launch");
  if (B.real && B.arith == 0) {
    // real tape, no Y words: the forests and their adjoints in real arithmetic
    hip_check(kl_pack_real(B.n_site_total, stream, B.garena, B.rarena), "k_pack_real launch");
    // the zero structure of the bond-doubling state for the I/Z forward chains
    const bool gs = B.iz && B.use_smask;
    if (gs) {
      hip_check(kl_site_masks(B.n_site_total / GV_SLOT, stream, B.rarena, B.smask, B.rfarena), "k_site_masks launch");
      hip_check(kl_chain_flags(ninst, B.max_chains, stream, B.d_inst, B.d_probs, B.smask, B.cflags, 1),
                "k_chain_flags launch");
      B.mask_kind = 1;
    }
    B.mark(1, stream);
    // the energy launches, deferred term ends by the leaf kernel
    for (int g = 0; g < B.max_gen; ++g) {
      const int ni = B.n_items[g];
      if (ni == 0) continue;
      hip_check(kl_rchains_gfwd(ni, B.wpb, B.iz, stream, B.gen_items(g), B.d_inst, B.d_probs, B.rarena, B.n_site_total,
                                B.spills, B.parts, B.estore, B.gleafs, B.leafs, B.rfarena, gs ? B.smask : nullptr,
                                B.cflags),
                "k_rchains_gfwd launch");
    }
    hip_check(kl_rleaf_terms(B.n_leaf_items, stream, B.leaf_desc, B.leafs, B.spills, nullptr,
                             B.parts),
              "k_rleaf_terms launch");
    B.mark(2, stream);
    hip_check(kl_reduce_energy(ninst, stream, B.d_inst, B.d_probs, B.parts, B.energies, 1), "k_reduce launch");
    B.mark(3, stream);
    for (int g = B.max_bgen - 1; g >= 0; --g) {
      const int ni = B.bn_items[g];
      if (ni == 0) continue;
      B.kbegin(stream);
      hip_check(kl_rchains_bwd(ni, B.wpb, B.iz, stream, B.bgen_items(g), B.d_inst, B.d_probs, B.rarena, B.n_site_total,
                               B.spills, B.adjspill, B.estore, B.gleafs, B.aslot, gs ? B.smask : nullptr, B.cflags),
                "k_rchains_bwd launch");
      B.kend(stream);
    }
    B.mark(4, stream);
    hip_check(kl_rsite_reduce(ninst, B.max_n, stream, B.d_inst, B.d_probs, B.aslot, B.abar), "k_rsite_reduce launch");
    B.mark(5, stream);
    hip_check(kl_site_bwd(ninst, B.max_n, stream, B.d_inst, B.d_probs, params, B.abar, B.rotin, B.rgrad),
              "k_site_bwd launch");
    hip_check(kl_param_grad(ninst, B.max_params, stream, B.d_inst, B.d_probs, B.rgrad, B.grads),
              "k_param_grad launch");
    B.mark(6, stream);
    return;
  }
  hip_check(kl_transpose_sites(B.n_site_total, stream, B.garena), "k_transpose_sites launch");
  B.mark(1, stream);
  for (int g = 0; g < B.max_ggen; ++g) {
    const int ni = B.gn_items[g];
    if (ni == 0) continue;
    hip_check(kl_chains_gfwd(ni, B.wpb, stream, B.ggen_items(g), B.d_inst, B.d_probs, B.garena, B.n_site_total, B.spills,
                             B.parts, B.estore,
                             B.gleafs),
              "k_chains_gfwd launch");
  }
  B.mark(2, stream);
  hip_check(kl_reduce_energy(ninst, stream, B.d_inst, B.d_probs, B.parts, B.energies, 0), "k_reduce launch");
  B.mark(3, stream);
  for (int g = B.max_bgen - 1; g >= 0; --g) {
    const int ni = B.bn_items[g];
    if (ni == 0) continue;
    B.kbegin(stream);
    hip_check(kl_chains_bwd(ni, B.wpb, stream, B.bgen_items(g), B.d_inst, B.d_probs, B.garena, B.n_site_total, B.spills,
                            B.adjspill, B.estore,
                            B.gleafs, B.aslot),
              "k_chains_bwd launch");
    B.kend(stream);
  }
  B.mark(4, stream);
  hip_check(kl_site_reduce(ninst, B.max_n, stream, B.d_inst, B.d_probs, B.aslot, B.abar), "k_site_reduce launch");
  B.mark(5, stream);
  hip_check(kl_site_bwd(ninst, B.max_n, stream, B.d_inst, B.d_probs, params, B.abar, B.rotin, B.rgrad),
            "k_site_bwd launch");
  hip_check(kl_param_grad(ninst, B.max_params, stream, B.d_inst, B.d_probs, B.rgrad, B.grads),
            "k_param_grad launch");
  B.mark(6, stream);
}

}  // namespace gv

// ------------------------------------------------------------------ C ABI
#define GV_TRY(...)                         \
  try {                                     \
    g_error.clear();                        \
    __VA_ARGS__;                            \
    return 0;                               \
  } catch (const std::exception &ex) {      \
    g_error = ex.what();                    \
    return 1;                               \
  } catch (...) {                           \
    g_error = "unknown error";              \
    return 1;                               \
  }

extern "C" {

const char *gvmps_last_error(void) { return g_error.c_str(); }

int gvmps_problem_create(gvmps_problem **out, int n_qubits, const gvmps_op *ops, int n_ops, int n_params,
                         int max_bond, double abs_cut, double rel_cut, const char *words, const double *coeffs,
                         int n_terms) {
  GV_TRY({
    if (!out) throw std::runtime_error("null output pointer");
    auto p = std::make_unique<gvmps_problem>();
    p->p = compile_problem(n_qubits, ops, n_ops, n_params, max_bond, abs_cut, rel_cut, words, coeffs, n_terms)
               .release();
    *out = p.release();
  })
}

int gvmps_problem_destroy(gvmps_problem *p) {
  GV_TRY({
    if (p) {
      delete p->p;
      delete p;
    }
  })
}

int gvmps_problem_get_generations(const gvmps_problem *p, int *out, int max_gens) {
  GV_TRY({
    if (!p || !out) throw std::runtime_error("null pointer");
    const Problem &P = *p->p;
    for (int g = 0; g < P.ngen && g < max_gens; ++g) {
      int steps = 0, ends = 0, spills = 0;
      for (int c = P.gen_begin[g]; c < P.gen_begin[g + 1]; ++c) {
        const ChainDev &ch = P.chains[c];
        steps += ch.nsteps;
        for (int st = 0; st < ch.nsteps; ++st) {
          const StepDev &sd = P.steps[ch.step0 + st];
          for (int k = 0; k < sd.nchild; ++k) {
            const ChildDev &cd = P.children[sd.child0 + k];
            ends += cd.nterm;
            spills += (cd.action >= 0) + (cd.store >= 0);
          }
        }
      }
      out[4 * g] = P.gen_begin[g + 1] - P.gen_begin[g];
      out[4 * g + 1] = steps;
      out[4 * g + 2] = ends;
      out[4 * g + 3] = spills;
    }
  })
}

int gvmps_problem_get_info(const gvmps_problem *p, gvmps_problem_info *info) {
  GV_TRY({
    if (!p || !info) throw std::runtime_error("null pointer");
    const Problem &P = *p->p;
    std::memset(info, 0, sizeof(*info));
    info->n_qubits = P.n;
    info->n_params = P.nparams;
    info->n_terms = P.nterms;
    info->n_active_terms = (int)P.active.size();
    info->n_prog_ops = (int)P.prog.size();
    info->n_two_site_gates = P.n_g2;
    info->max_alloc_bond = P.max_alloc_bond;
    info->n_chains = (int)P.chains.size();
    info->n_steps = (int)P.steps.size();
    info->n_spill = P.nspill;
    info->n_generations = P.ngen;
    info->n_child_nodes = P.n_nodes;
    info->n_right_nodes = P.n_right_nodes;
    info->n_sched_steps = (int)P.sched.size() / GV_WAVES;
    info->flops_energy_executed = P.flops_exec;
    info->flops_energy_nodes = P.flops_nodes;
    info->flops_svd_nominal = P.flops_svd;
    info->flops_grad_executed = P.flops_grad;
    info->n_rotations = (int)P.rot_param.size();
    info->grad_ok = P.grad_error.empty() ? 1 : 0;
    info->real_arith = P.real_ok ? 1 : 0;
    info->iz_forest = P.iz_only ? 1 : 0;
    info->exact_split = P.exact_split ? 1 : 0;
  })
}

int gvmps_batch_create(gvmps_batch **out, gvmps_problem *const *problems, const int *counts, int n_problems,
                       int device) {
  GV_TRY({
    if (!out || !problems || !counts) throw std::runtime_error("null pointer");
    *out = reinterpret_cast<gvmps_batch *>(make_batch(problems, counts, n_problems, device));
  })
}

int gvmps_batch_destroy(gvmps_batch *b) { GV_TRY({ delete reinterpret_cast<Batch *>(b); }) }

int gvmps_batch_size(const gvmps_batch *b) {
  return b ? (int)reinterpret_cast<const Batch *>(b)->inst.size() : 0;
}

int gvmps_batch_prepare(gvmps_batch *b, const double *params_dev, void *stream) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    launch_prepare(B, params_dev, (hipStream_t)stream);
    check_status(B, (hipStream_t)stream);
  })
}

int gvmps_batch_energy(gvmps_batch *b, const double *params_dev, double *energies_dev, double *terms_dev,
                       void *stream) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    hipStream_t s = (hipStream_t)stream;
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    launch_prepare(B, params_dev, s);
    launch_energy_tail(B, s, terms_dev != nullptr);
    if (energies_dev)
      hip_check(hipMemcpyAsync(energies_dev, B.energies, sizeof(double) * B.inst.size(),
                               hipMemcpyDeviceToDevice, s),
                "copy energies");
    if (terms_dev)
      hip_check(hipMemcpyAsync(terms_dev, B.terms, sizeof(double2) * B.n_terms_total, hipMemcpyDeviceToDevice, s),
                "copy terms");
    check_status(B, s);
  })
}

int gvmps_batch_sample(gvmps_batch *b, const double *params_dev, int shots, const uint32_t *seeds,
                       uint8_t *bits_dev, uint64_t *packed_dev, void *stream) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    hipStream_t s = (hipStream_t)stream;
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    launch_sample(B, params_dev, shots, seeds, bits_dev, reinterpret_cast<unsigned long long *>(packed_dev), s);
    check_status(B, s);
  })
}

int gvmps_batch_shot_energy(gvmps_batch *b, const double *params_dev, int shots, const uint32_t *seeds,
                            double *energies_dev, double *terms_dev, void *stream) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    hipStream_t s = (hipStream_t)stream;
    for (int z : B.zonly)
      if (!z) throw std::runtime_error("finite-shot observe supports I/Z Pauli words only");
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    unsigned long long *packed = launch_sample(B, params_dev, shots, seeds, nullptr, nullptr, s);
    const int ninst = (int)B.inst.size();
    hip_check(kl_zsums(ninst, B.max_terms, shots, s, B.d_inst, B.d_probs, B.d_masks, packed, B.terms), "k_zsums");
    hip_check(kl_reduce(ninst, s, B.d_inst, B.d_probs, B.terms, B.energies), "k_reduce launch");
    B.mark(4, s);
    if (energies_dev)
      hip_check(hipMemcpyAsync(energies_dev, B.energies, sizeof(double) * B.inst.size(),
                               hipMemcpyDeviceToDevice, s),
                "copy energies");
    if (terms_dev)
      hip_check(hipMemcpyAsync(terms_dev, B.terms, sizeof(double2) * B.n_terms_total, hipMemcpyDeviceToDevice, s),
                "copy terms");
    check_status(B, s);
  })
}

int gvmps_batch_shot_energy_noalloc(gvmps_batch *b, const double *params_dev, int shots, const uint32_t *seeds,
                                    double *energies_dev, double *terms_dev, void *stream) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    hipStream_t s = (hipStream_t)stream;
    for (int z : B.zonly)
      if (!z) throw std::runtime_error("finite-shot observe supports I/Z Pauli words only");
    if (shots <= 0) throw std::runtime_error("shared bitstring sampling requires a positive shot count");
    if (!seeds) throw std::runtime_error("null seed array");
    const int ninst = (int)B.inst.size();
    if ((unsigned long long)ninst * (unsigned long long)shots * 2ULL > GV_SCRATCH_WORDS)
      throw std::runtime_error("gvmps_batch_shot_energy_noalloc: instances x shots exceeds the scratch maximum of " +
                               std::to_string(GV_SCRATCH_WORDS / 2) + " shot words (use gvmps_batch_shot_energy)");
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    unsigned long long *packed = scratch_address(B.device);
    hip_check(hipMemcpyAsync(B.seeds, seeds, sizeof(unsigned) * ninst, hipMemcpyHostToDevice, s), "seeds H2D");
    launch_prepare(B, params_dev, s);
    // every live shot's two words are written before k_zsums reads them, so
    // nothing in the scratch needs clearing
    launch_sample_tail(B, shots, B.seeds, nullptr, packed, s, true);
    hip_check(kl_zsums(ninst, B.max_terms, shots, s, B.d_inst, B.d_probs, B.d_masks, packed, B.terms), "k_zsums");
    hip_check(kl_reduce(ninst, s, B.d_inst, B.d_probs, B.terms, B.energies), "k_reduce launch");
    B.mark(4, s);
    if (energies_dev)
      hip_check(hipMemcpyAsync(energies_dev, B.energies, sizeof(double) * B.inst.size(),
                               hipMemcpyDeviceToDevice, s),
                "copy energies");
    if (terms_dev)
      hip_check(hipMemcpyAsync(terms_dev, B.terms, sizeof(double2) * B.n_terms_total, hipMemcpyDeviceToDevice, s),
                "copy terms");
    check_status(B, s);
  })
}

int gvmps_batch_set_timing(gvmps_batch *b, int enable) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    B.timing = enable != 0;
    B.ev_used = 0;
    B.kev_used = 0;
    B.phase_calls = 0;
    B.kern_ms = 0.0;
    B.kern_launches = 0;
    for (double &v : B.phase_ms) v = 0.0;
  })
}

int gvmps_batch_get_timing(gvmps_batch *b, double *phase_ms, long long *calls) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    for (int i = 0; i < 6; ++i) phase_ms[i] = B.phase_ms[i];
    phase_ms[6] = B.kern_ms;
    phase_ms[7] = (double)B.kern_launches;
    if (calls) *calls = B.phase_calls;
  })
}

int gvmps_batch_prep_stats(gvmps_batch *b, int *sweeps_host) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hip_check(hipMemcpy(sweeps_host, B.stats, sizeof(int) * B.inst.size() * GV_STATS, hipMemcpyDeviceToHost),
              "read stats");
  })
}

int gvmps_batch_value_grad(gvmps_batch *b, const double *params_dev, double *energies_dev, double *grads_dev,
                           void *stream) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    hipStream_t s = (hipStream_t)stream;
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    launch_value_grad(B, params_dev, s);
    if (energies_dev)
      hip_check(hipMemcpyAsync(energies_dev, B.energies, sizeof(double) * B.inst.size(),
                               hipMemcpyDeviceToDevice, s),
                "copy energies");
    if (grads_dev && B.n_params_total > 0)
      hip_check(hipMemcpyAsync(grads_dev, B.grads, sizeof(double) * B.n_params_total, hipMemcpyDeviceToDevice, s),
                "copy gradients");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    B.collect();
  })
}

int gvmps_batch_set_launch(gvmps_batch *b, int waves_per_group, int order) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    if (waves_per_group != 1 && waves_per_group != 2 && waves_per_group != 4)
      throw std::runtime_error("waves per workgroup must be 1, 2 or 4");
    if (order != 0 && order != 1) throw std::runtime_error("order must be 0 (instance-major) or 1 (longest first)");
    if (waves_per_group != B.wpb) {  // the XCD ranges depend on the workgroup size
      hip_check(hipSetDevice(B.device), "hipSetDevice");
      hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
      for (size_t g = 0; g < B.items_host.size(); ++g) {
        std::vector<int2> items = B.items_host[g];
        xcd_longest_first(items, waves_per_group, chain_len(B));
        hip_check(hipMemcpy(B.items_len[g], items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice),
                  "hipMemcpy");
      }
      for (size_t g = 0; g < B.gitems_host.size(); ++g) {
        std::vector<int2> items = B.gitems_host[g];
        xcd_longest_first(items, waves_per_group, chain_len(B));
        hip_check(hipMemcpy(B.gitems_len[g], items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice),
                  "hipMemcpy");
      }
      for (size_t g = 0; g < B.bitems_host.size(); ++g) {
        std::vector<int2> items = B.bitems_host[g];
        xcd_longest_first(items, waves_per_group, chain_len(B));
        hip_check(hipMemcpy(B.bitems_len[g], items.data(), items.size() * sizeof(int2), hipMemcpyHostToDevice),
                  "hipMemcpy");
      }
    }
    B.wpb = waves_per_group;
    B.order = order;
  })
}

int gvmps_batch_chain_flops(gvmps_batch *b, double *flops) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    if (!flops) throw std::runtime_error("null output");
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    std::vector<unsigned> mask;
    std::vector<int> flag;  // I/Z forests: chains on the merged fragments (the rest run full products)
    const bool real = B.real && B.arith == 0;
    if (real && B.use_smask) {
      if (B.mask_kind != 0) throw std::runtime_error("chain flops: the batch's last call was not an energy call");
      mask.resize(2 * (size_t)(B.n_site_total / GV_SLOT));
      hip_check(hipMemcpy(mask.data(), B.smask, mask.size() * sizeof(unsigned), hipMemcpyDeviceToHost), "hipMemcpy");
      if (B.iz) {
        long long nparts = 0;
        for (const InstDev &in : B.inst) {
          const Problem &P = *B.probs[in.prob];
          nparts = std::max(nparts, in.part_off + (long long)P.chains.size());
        }
        flag.resize(nparts);
        hip_check(hipMemcpy(flag.data(), B.cflags, flag.size() * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
      }
    }
    // FP64 flops of one real 16x16x16 product; the complex kernels execute three
    // (Gauss form) per complex product
    const double prod = 2.0 * GV_CHI * GV_CHI * GV_CHI;
    double total = 0.0;
    for (const InstDev &in : B.inst) {
      const Problem &P = *B.probs[in.prob];
      const long long slot0 = in.site_off / GV_SLOT;
      for (size_t c = 0; c < P.chains.size(); ++c) {
        const ChainDev &ch = P.chains[c];
        const bool full_chain = !flag.empty() && flag[in.part_off + c] == 0;
        for (int st = 0; st < ch.nsteps; ++st) {
          const StepDev &sd = P.steps[ch.step0 + st];
          const bool both = sd.need == 3;
          double products;  // real 16^3 products of the step
          if (!real) {
            products = 3.0 * (both ? 6.0 : 4.0);
          } else if (mask.empty() || full_chain) {
            products = both ? 6.0 : 4.0;
          } else {
            const unsigned mw = mask[2 * (slot0 + sd.site) + (ch.dir != 0 ? 1 : 0)];
            if (mw & 0x100u) products = 2.0;                        // merged pair serves both passes
            else if ((mw & 0xffu) == 0xc3u) products = both ? 3.0 : 2.0;  // row split: half k-steps
            else products = both ? 6.0 : 4.0;
          }
          total += products * prod;
        }
      }
    }
    *flops = total;
  })
}

int gvmps_batch_grad_flops(gvmps_batch *b, double *flops) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    if (!flops) throw std::runtime_error("null output");
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    const bool real = B.real && B.arith == 0;
    double total = 0.0;
    if (!real) {  // complex kernels: the planner's count (4 real products per complex one)
      for (const InstDev &in : B.inst) total += B.probs[in.prob]->flops_grad;
      *flops = total;
      return 0;
    }
    std::vector<unsigned> mask;
    std::vector<int> flag;  // chains the forward ran over the zero structure
    if (B.iz && B.use_smask && B.mask_kind == 1) {
      mask.resize(2 * (size_t)(B.n_site_total / GV_SLOT));
      hip_check(hipMemcpy(mask.data(), B.smask, mask.size() * sizeof(unsigned), hipMemcpyDeviceToHost), "hipMemcpy");
      long long nparts = 0;
      for (const InstDev &in : B.inst)
        nparts = std::max(nparts, in.part_off + (long long)B.probs[in.prob]->chains.size());
      flag.resize(nparts);
      hip_check(hipMemcpy(flag.data(), B.cflags, flag.size() * sizeof(int), hipMemcpyDeviceToHost), "hipMemcpy");
    }
    const double prod = 2.0 * GV_CHI * GV_CHI * GV_CHI;  // one real 16^3 product
    for (const InstDev &in : B.inst) {
      const Problem &P = *B.probs[in.prob];
      const long long slot0 = in.site_off / GV_SLOT;
      for (size_t c = 0; c < P.chains.size(); ++c) {
        const ChainDev &ch = P.chains[c];
        const bool structured = !flag.empty() && flag[in.part_off + c] != 0;
        for (int st = 0; st < ch.nsteps; ++st) {
          const StepDev &sd = P.steps[ch.step0 + st];
          const bool iz = sd.need & 1, xy = sd.need & 2;
          // forward: X_0, X_1 and a Y pair per pass; backward: G, Ebar, Mbar pairs
          double fwd = 2.0 + (iz ? 2.0 : 0.0) + (xy ? 2.0 : 0.0);
          const double bwd = (iz && xy) ? 8.0 : 6.0;
          if (structured) {
            const unsigned mw = mask[2 * (slot0 + sd.site) + (ch.dir != 0 ? 1 : 0)];
            if (mw & (ch.dir != 0 ? 0x200u : 0x100u)) fwd = 2.0;  // row-split over permuted rows / merged pair
          }
          total += (fwd + bwd) * prod;
        }
      }
    }
    *flops = total;
  })
}

int gvmps_batch_set_sparsity(gvmps_batch *b, int on) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    B.use_smask = on != 0;
  })
}

int gvmps_batch_set_checkpoint(gvmps_batch *b, int stride) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    B.set_checkpoint(stride);
  })
}

int gvmps_batch_set_prep(gvmps_batch *b, int mode) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    if (mode < 0 || mode > 2)
      throw std::runtime_error("preparation mode must be 0 (gates + splits), 1 (bond doubling) or 2 (gates + a "
                               "decomposition at every split)");
    if (mode == 1)
      for (const Problem *P : B.probs)
        if (!P->grad_error.empty()) throw std::runtime_error(P->grad_error);
    B.prep_mode = mode;
  })
}

int gvmps_batch_set_arith(gvmps_batch *b, int mode) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    if (mode != 0 && mode != 1) throw std::runtime_error("arithmetic mode must be 0 (automatic) or 1 (complex)");
    B.arith = mode;
  })
}

int gvmps_batch_evaluate(gvmps_batch *b, const double *params_dev, int shots, const uint32_t *seeds_dev,
                         double *energies_dev, uint8_t *bits_dev, uint64_t *packed_dev, void *stream) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    hipStream_t s = (hipStream_t)stream;
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    launch_evaluate(B, params_dev, shots, seeds_dev, energies_dev, bits_dev,
                    reinterpret_cast<unsigned long long *>(packed_dev), s);
    check_status(B, s);
  })
}

int gvmps_batch_capture(gvmps_batch *b, const double *params_dev, int shots, const uint32_t *seeds_dev,
                        double *energies_dev, uint8_t *bits_dev, uint64_t *packed_dev) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    capture_evaluate(B, params_dev, shots, seeds_dev, energies_dev, bits_dev,
                     reinterpret_cast<unsigned long long *>(packed_dev));
  })
}

int gvmps_batch_replay(gvmps_batch *b, void *stream) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    if (!B.graph_exec) throw std::runtime_error("no captured sequence (call gvmps_batch_capture first)");
    hipStream_t s = (hipStream_t)stream;
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    hip_check(hipGraphLaunch(B.graph_exec, s), "hipGraphLaunch");
    check_status(B, s);
  })
}

int gvmps_batch_read_sites(gvmps_batch *b, int inst, double *sites_host, int *bonds_host) {
  GV_TRY({
    Batch &B = *reinterpret_cast<Batch *>(b);
    if (inst < 0 || inst >= (int)B.inst.size()) throw std::runtime_error("instance index out of range");
    hip_check(hipSetDevice(B.device), "hipSetDevice");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const InstDev &in = B.inst[inst];
    if (sites_host)
      hip_check(hipMemcpy(sites_host, B.arena + in.site_off, sizeof(double2) * (size_t)in.n * GV_SLOT,
                          hipMemcpyDeviceToHost),
                "read sites");
    if (bonds_host)
      hip_check(hipMemcpy(bonds_host, B.bonds + (size_t)inst * (GV_NMAX + 1), sizeof(int) * (in.n + 1),
                          hipMemcpyDeviceToHost),
                "read bonds");
  })
}

}  // extern "C"
