// Host-side batch object shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "plan.h"

struct gvmps_problem {
  gv::Problem *p = nullptr;
};

namespace gv {

struct Batch {
  int device = -1;
  std::vector<const Problem *> probs;
  std::vector<InstDev> inst;
  std::vector<void *> owned;  // every device allocation of this batch
  ProbDev *d_probs = nullptr;
  InstDev *d_inst = nullptr;
  // site arena: n_site_total elements of the prepared MPS, then (complex
  // forests) the same sites transposed
  double2 *arena = nullptr, *envs = nullptr, *spills = nullptr, *terms = nullptr;
  double *parts = nullptr;  // chain partial energies
  int *status = nullptr, *bonds = nullptr, *stats = nullptr;
  std::vector<int> status_host;  // read-back of `status`, sized once at creation
  double *energies = nullptr;
  // sampling
  const unsigned long long **d_masks = nullptr;  // per problem: 2 words per term (site order)
  std::vector<int> zonly;                        // per problem: all words are I/Z
  unsigned *seeds = nullptr;
  unsigned char *sbits = nullptr;
  unsigned long long *spacked = nullptr;
  size_t sbits_cap = 0, spacked_cap = 0;
  int max_terms = 0;
  double2 *leafs = nullptr;       // deferred term-end leaf environments
  int2 *leaf_items = nullptr;     // (instance, leaf) work items
  LeafItem *leaf_desc = nullptr;  // the same items as descriptors (real leaf kernels)
  int n_leaf_items = 0;
  std::vector<int2 *> items;      // per generation, instance-major
  std::vector<int2 *> items_len;  // per generation, longest chains first within each XCD's blocks (for wpb)
  std::vector<std::vector<int2>> items_host;  // instance-major, host copy
  std::vector<int> n_items;
  int max_gen = 0;
  // complex gradient forward generations (suffix forest, then prefix forest)
  std::vector<int2 *> gitems;
  std::vector<int2 *> gitems_len;
  std::vector<std::vector<int2>> gitems_host;
  std::vector<int> gn_items;
  int max_ggen = 0;
  // backward launches (the two forests' backward passes side by side)
  std::vector<int2 *> bitems;
  std::vector<int2 *> bitems_len;
  std::vector<std::vector<int2>> bitems_host;
  std::vector<int> bn_items;
  int max_bgen = 0;
  // launch configuration of the chain kernels (set by the autotuner)
  int wpb = 4;    // wavefronts per workgroup: 1, 2 or 4
  int order = 1;  // 0: instance-major items, 1: longest chains first within each XCD's share
  int prep_mode = 0;  // 0: gates + splits (k_prep), 1: exact bond doubling (k_build_local),
                      // 2: gates + a decomposition at every split, exact ones included
  // real-arithmetic energies: every problem has real gates and no Y words
  bool real = false;
  bool iz = false;    // every problem's forest has I/Z children only: k_rchains*_iz
  int arith = 0;              // 0: real path when `real`, 1: complex path always
  double *rarena = nullptr;   // real site arena: n_site_total doubles as A_i, then n_site_total as A_i^T
  unsigned *smask = nullptr;  // zero structure of each real site slot, direct and transposed (k_site_masks)
  double *rfarena = nullptr;  // merged fragments A_0 + A_1 of each slot, direct then transposed (k_site_masks)
  int *cflags = nullptr;      // per (instance, chain): its steps run on the merged fragments (k_chain_flags)
  bool use_smask = true;      // chain steps skip exact-zero products (gvmps_batch_set_sparsity)
  int mask_kind = -1;         // smask / cflags last written by: 0 an energy call, 1 a value_grad call
  const int2 *gen_items(int g) const { return order ? items_len[g] : items[g]; }
  const int2 *ggen_items(int g) const { return order ? gitems_len[g] : gitems[g]; }
  const int2 *bgen_items(int g) const { return order ? bitems_len[g] : bitems[g]; }
  // captured evaluate + sample sequence
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  long long n_terms_total = 0, n_params_total = 0;
  int max_n = 0, max_params = 0, max_chains = 0;
  // reverse gradient (allocated on first use)
  long long n_rot_total = 0, n_steps_total = 0, n_gleaf_total = 0, n_site_total = 0, n_spill_total = 0;
  bool grad_ready = false;
  // step-environment checkpoints: every ck_stride-th step of a chain keeps
  // its input environment, the others are recomputed in the backward pass
  int ck_stride = 1;
  long long n_ck_total = 0;
  std::vector<ProbDev> h_probs;           // host copy of the problem table
  std::vector<int *> ck_slots;            // per problem, device
  void set_checkpoint(int stride);
  double2 *garena = nullptr, *rotin = nullptr, *estore = nullptr, *gleafs = nullptr, *adjspill = nullptr,
          *aslot = nullptr, *abar = nullptr;
  double *rgrad = nullptr, *grads = nullptr;
  // optional per-phase HIP-event timing (prep, envs, chains, reduce, sample, grad)
  bool timing = false;
  hipEvent_t ev[8] = {};
  int ev_used = 0;
  double phase_ms[8] = {};
  long long phase_calls = 0;
  void mark(int i, hipStream_t s);
  void collect();
  // per-launch timing of the dominant kernel (the forest-chain kernel of
  // energy calls, the backward chain kernel of value_grad): events recorded
  // before and after each of its launches on the launch stream
  std::vector<hipEvent_t> kev;
  int kev_used = 0;
  double kern_ms = 0.0;
  long long kern_launches = 0;
  void kbegin(hipStream_t s);
  void kend(hipStream_t s);
  ~Batch();
};

void set_error(const std::string &msg);
void hip_check(hipError_t e, const char *what);
Batch *make_batch(gvmps_problem *const *problems, const int *counts, int n_problems, int device);
void check_status(Batch &B, hipStream_t stream);
void launch_prepare(Batch &B, const double *params, hipStream_t stream);
void launch_energy_tail(Batch &B, hipStream_t stream, bool terms);
void launch_sample_tail(Batch &B, int shots, const unsigned *seeds_dev, unsigned char *bits,
                        unsigned long long *packed, hipStream_t stream, bool pack);

// kernel launchers (kernels_energy.hip)
hipError_t kl_prep(int ninst, bool real, bool factor_always, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                   const double *params, double2 *arena, int *bonds, int *status, int *stats);
hipError_t kl_envs(int ninst, hipStream_t s, const InstDev *insts, const double2 *arena, double2 *envs);
hipError_t kl_transpose_sites(long long n, hipStream_t s, double2 *arena);
hipError_t kl_chains(int nitems, int wpb, bool iz, hipStream_t s, const int2 *items, const InstDev *insts, const ProbDev *probs,
                     const double2 *arena, long long nsite, double2 *leafs, double2 *spills, double2 *terms);
hipError_t kl_leaf_terms(int nitems, hipStream_t s, const int2 *items, const InstDev *insts, const ProbDev *probs,
                         const double2 *leafs, const double2 *spills, double2 *terms, double *parts);
hipError_t kl_reduce(int ninst, hipStream_t s, const InstDev *insts, const ProbDev *probs, double2 *terms,
                     double *energies);
hipError_t kl_chains_energy(int nitems, int wpb, bool iz, hipStream_t s, const int2 *items, const InstDev *insts,
                            const ProbDev *probs, const double2 *arena, long long nsite, double2 *leafs,
                            double2 *spills, double *parts);
hipError_t kl_reduce_energy(int ninst, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                            const double *parts, double *energies, int with_leaves);
void launch_value_grad(Batch &B, const double *params, hipStream_t stream);
// kernel launchers (kernels_real.hip)
hipError_t kl_pack_real(long long n, hipStream_t s, const double2 *arena, double *rarena);
hipError_t kl_rchains(int nitems, int wpb, bool iz, hipStream_t s, const int2 *items, const InstDev *insts,
                      const ProbDev *probs, const double *rarena, long long nsite, const unsigned *smask,
                      const double *farena, const int *cflags, double2 *leafs, double2 *spills, double2 *terms,
                      double *parts);
hipError_t kl_site_masks(long long nslots, hipStream_t s, const double *rarena, unsigned *smask, double *farena);
hipError_t kl_chain_flags(int ninst, int max_chains, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                          const unsigned *smask, int *cflags, int grad);
hipError_t kl_rleaf_terms(int nitems, hipStream_t s, const LeafItem *items, const double2 *leafs,
                          const double2 *spills, double2 *terms, double *parts);
hipError_t kl_rchains_gfwd(int nitems, int wpb, bool iz, hipStream_t s, const int2 *items, const InstDev *insts,
                           const ProbDev *probs, const double *rgarena, long long nsite, double2 *spills, double *parts,
                           double2 *estore, double2 *gleafs, double2 *leafs, const double *farena,
                           const unsigned *smask, const int *cflags);
hipError_t kl_rchains_bwd(int nitems, int wpb, bool iz, hipStream_t s, const int2 *items, const InstDev *insts,
                          const ProbDev *probs, const double *rgarena, long long nsite, const double2 *spills,
                          double2 *adjspill, const double2 *estore, const double2 *gleafs, double2 *aslot,
                          const unsigned *smask, const int *cflags);
hipError_t kl_rsite_reduce(int ninst, int max_n, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                           const double2 *aslot, double2 *abar);
// kernel launchers (kernels_grad.hip)
hipError_t kl_build_local(int ninst, int max_n, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                          const double *params, double2 *garena, double2 *rotin, int *status);
hipError_t kl_chains_gfwd(int nitems, int wpb, hipStream_t s, const int2 *items, const InstDev *insts, const ProbDev *probs,
                          const double2 *garena, long long nsite, double2 *spills, double *parts, double2 *estore,
                          double2 *gleafs);
hipError_t kl_chains_bwd(int nitems, int wpb, hipStream_t s, const int2 *items, const InstDev *insts, const ProbDev *probs,
                         const double2 *garena, long long nsite, const double2 *spills, double2 *adjspill,
                         const double2 *estore, const double2 *gleafs, double2 *aslot);
hipError_t kl_site_reduce(int ninst, int max_n, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                          const double2 *aslot, double2 *abar);
hipError_t kl_site_bwd(int ninst, int max_n, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                       const double *params, const double2 *abar, const double2 *rotin, double *rgrad);
hipError_t kl_param_grad(int ninst, int max_params, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                         const double *rgrad, double *grads);
// kernel launchers (kernels_sample.hip)
hipError_t kl_sample(int ninst, int shots, int stride, hipStream_t s, const InstDev *insts, const double2 *arena,
                     const double2 *envs, const unsigned *seeds, unsigned char *bits, unsigned long long *packed);
hipError_t kl_renvs_right(int ninst, hipStream_t s, const InstDev *insts, const double *rarena, long long nsite,
                          double2 *envs);
hipError_t kl_sample_real(int ninst, int shots, int stride, hipStream_t s, const InstDev *insts, const double *rarena,
                          const double2 *envs, const unsigned *seeds, unsigned char *bits, unsigned long long *packed);
hipError_t kl_zsums(int ninst, int max_terms, int shots, hipStream_t s, const InstDev *insts, const ProbDev *probs,
                    const unsigned long long *const *masks, const unsigned long long *packed, double2 *terms);

}  // namespace gv
