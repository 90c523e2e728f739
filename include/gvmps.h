/*
 * gvmps -- MI355X (gfx950) MPS simulator backend: C ABI.
 *
 * Drop-in boundary for the hot path of the reference's custom MPS NVQIR
 * plugin (test/backend/source/simulator_mps_dist.cu) and its reverse-gradient
 * engine (test/tools/reverse_mps.py).  Plain pointers and sizes only; device
 * pointers are HIP device memory, `stream` is a hipStream_t (NULL = default).
 *
 * Every entry point returns 0 on success and a nonzero status on failure; the
 * message is then available from gvmps_last_error() (thread-local).  No entry
 * point calls exit() (the reference's CUDA_CHECK does, dist_svd.cuh:22-27).
 */
#ifndef GVMPS_API_H_
#define GVMPS_API_H_
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gate kinds of a circuit tape (CUDA-Q builder ops recorded by
 * test/tools/circuit_tape.py:56-66). */
enum gvmps_gate {
  GVMPS_RX = 0, GVMPS_RY = 1, GVMPS_RZ = 2,
  GVMPS_H = 3, GVMPS_X = 4, GVMPS_Y = 5, GVMPS_Z = 6, GVMPS_S = 7, GVMPS_SDG = 8,
  GVMPS_CX = 9, GVMPS_SWAP = 10
};

/* One tape entry, in CUDA-Q qubit indices.  Rotations take the angle
 * scale * x[param] + offset (param = -1: the constant `offset`); this covers
 * the affine parameter expressions of circuit_tape.Parameter. */
typedef struct gvmps_op {
  int32_t kind;
  int32_t q0, q1;
  int32_t param;
  double scale, offset;
} gvmps_op;

typedef struct gvmps_problem gvmps_problem;
typedef struct gvmps_batch gvmps_batch;

typedef struct gvmps_problem_info {
  int32_t n_qubits, n_params, n_terms, n_active_terms;
  int32_t n_prog_ops, n_two_site_gates;
  int32_t max_alloc_bond;          /* predicted allocated bond (zero-cutoff rule) */
  int32_t n_chains, n_steps, n_spill, n_generations;
  int32_t n_child_nodes;           /* prefix-forest nodes */
  int32_t n_sched_steps;           /* time steps of the pipelined gate schedule */
  int32_t n_right_nodes;           /* suffix-forest nodes */
  double flops_energy_executed;    /* FP64 flops of one energy eval, as executed */
  double flops_energy_nodes;       /* prefix-node algorithmic count, 16(Dl^2Dr+DlDr^2) per node */
  double flops_svd_nominal;        /* Golub-Reinsch nominal count of all SVDs */
  double flops_grad_executed;      /* FP64 flops of one reverse-gradient pass (forest part) */
  int32_t n_rotations;             /* parameterised + constant rotations of the tape */
  int32_t grad_ok;                 /* 1: the tape has a bond-doubling reverse gradient */
  int32_t real_arith;              /* 1: real gate matrices and no Y words (real-arithmetic energies) */
  int32_t iz_forest;               /* 1: no X or Y words: every forest step has I/Z children only */
  int32_t exact_split;             /* 1: zero cutoffs and no two-site gate cut below min(2 Dl, 2 Dr):
                                      every split keeps all of Theta and needs no decomposition */
} gvmps_problem_info;

/* Last error message of the calling thread ("" if none). */
const char *gvmps_last_error(void);

/* Compile a circuit + Hamiltonian.  Replaces the plugin's per-gate
 * applyGate (simulator_mps_dist.cu:1392-1406, routing :392-416, controlled
 * gates :374-384, site = n-1-q :386-387), its engine configuration
 * (CUDAQ_MPS_MAX_BOND / ABS_CUTOFF / RELATIVE_CUTOFF, :363-366) and the Pauli
 * word construction of observe (buildPauliTermData :571-583).
 *   words : n_terms * n_qubits chars 'I','X','Y','Z', word[q] acts on qubit q
 *   coeffs: 2 * n_terms doubles (re, im)
 * Host-only: needs no GPU. */
int gvmps_problem_create(gvmps_problem **out, int n_qubits, const gvmps_op *ops, int n_ops,
                         int n_params, int max_bond, double abs_cut, double rel_cut,
                         const char *words, const double *coeffs, int n_terms);
int gvmps_problem_destroy(gvmps_problem *p);
/* Chain schedule per generation, 4 ints each: chains, expansion steps, term
 * ends and environment stores of generation g, for g < min(n_generations,
 * max_gens).  Diagnostic. */
int gvmps_problem_get_generations(const gvmps_problem *p, int *out, int max_gens);
int gvmps_problem_get_info(const gvmps_problem *p, gvmps_problem_info *info);

/* A batch of MPS instances: counts[j] instances of problems[j], in order.
 * Allocates the device workspace on `device`.  Replaces the per-thread
 * ShardedMPS / workspace ownership (sharded_mps.cuh:309-353, 614-665): one
 * batch holds every instance of a persistent launch. */
int gvmps_batch_create(gvmps_batch **out, gvmps_problem *const *problems, const int *counts,
                       int n_problems, int device);
int gvmps_batch_destroy(gvmps_batch *b);
int gvmps_batch_size(const gvmps_batch *b);

/* Analytic energies (shots = -1 observe).  Replaces observe()'s analytic path
 * (simulator_mps_dist.cu:1434-1485) over expectPauliShared
 * (sharded_mps.cuh:614-687, analytic_batch.cuh:4-36), for every instance:
 *   params_dev  : concatenated parameter vectors (instance order), doubles
 *   energies_dev: one double per instance, Re sum_k c_k <P_k>
 *   terms_dev   : optional, per instance n_terms complex <P_k> (re, im)  */
int gvmps_batch_energy(gvmps_batch *b, const double *params_dev, double *energies_dev,
                       double *terms_dev, void *stream);

/* State preparation only (gates + SVDs).  Replaces the applyGate stream
 * (sharded_mps.cuh:465-572 with dist_svd.cuh:194-283).  Returns nonzero with
 * the reference's guard message when a zero-cutoff SVD would discard a
 * resolved singular value (dist_svd.cuh:277-279). */
int gvmps_batch_prepare(gvmps_batch *b, const double *params_dev, void *stream);

/* Perfect sampling of every instance.  Replaces sample()'s noiseless path
 * (simulator_mps_dist.cu:1587-1612) over sampleBitsGpu / sampleBitsDevice
 * (sharded_mps.cuh:814-902):
 *   seeds      : host array, one kernel seed per instance (the reference
 *                passes hbatchSeedForShotShard(sampleSeed++, 0), :596-608)
 *   bits_dev   : optional, instances*shots*n bytes 0/1, SITE order (site s =
 *                qubit n-1-s), as k_chooseSample writes them
 *   packed_dev : optional, instances*shots*2 uint64, bit s of word s/64 = site s */
int gvmps_batch_sample(gvmps_batch *b, const double *params_dev, int shots, const uint32_t *seeds,
                       uint8_t *bits_dev, uint64_t *packed_dev, void *stream);

/* Finite-shot energies from one shared sample set per instance for all-Z/I
 * Hamiltonians.  Replaces observe()'s diagonal shot path
 * (simulator_mps_dist.cu:1493-1514, runSharedDiagonalShotBatch :1157-1171)
 * over sampleZExpectationsGpu (sharded_mps.cuh:906-935).  Per-term values are
 * (sum of parities)/shots; identity words give 1.  Fails for X/Y words. */
int gvmps_batch_shot_energy(gvmps_batch *b, const double *params_dev, int shots, const uint32_t *seeds,
                            double *energies_dev, double *terms_dev, void *stream);

/* gvmps_batch_shot_energy without allocation: the same kernels and the same
 * results (bitwise), with the shots' packed words staged in a file-scope
 * device array (32 MiB per GPU, held until the process ends) instead of a
 * buffer the batch allocates on first use; the call makes no hipMalloc,
 * hipFree or other allocating call.  instances x shots may not exceed
 * 2,097,152 (error otherwise).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU race.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation). */
int gvmps_batch_shot_energy_noalloc(gvmps_batch *b, const double *params_dev, int shots, const uint32_t *seeds,
                                    double *energies_dev, double *terms_dev, void *stream);

/* Energy and gradient dE/dx of every instance by one reverse (adjoint) pass.
 * Replaces make_problem(kernel, H).value_gradient(x) of the reference's
 * reverse engine (test/tools/reverse_mps.py:84-121; tape variant :141-184):
 * the state is rebuilt with every CNOT as an exact bond doubling (no SVD,
 * :45-50), so only neighbouring CX and single-qubit gates are accepted and
 * the doubled bonds must fit 16; otherwise nonzero with the reason.  The
 * loss is sum_k Re(c_k) Re<P_k> (:78-81).
 *   energies_dev: one double per instance (optional)
 *   grads_dev   : n_params doubles per instance, instance order (optional)
 * Synchronises the stream. */
int gvmps_batch_value_grad(gvmps_batch *b, const double *params_dev, double *energies_dev,
                           double *grads_dev, void *stream);

/* Memory/recompute trade-off of gvmps_batch_value_grad: the forward pass
 * keeps the input environment (4 KiB) of every `stride`-th step of each chain
 * and the backward pass recomputes the others from the nearest kept one
 * (stride - 1 forward transfers at most).  stride 1 (default) keeps all;
 * results are bitwise the same for every stride.  The reference's reverse
 * engine keeps autograd's whole tape instead (test/tools/reverse_mps.py:86-99). */
int gvmps_batch_set_checkpoint(gvmps_batch *b, int stride);

/* Zero-structure use of the real energy forests (default on): after the
 * state is prepared, one pass records for every site which 4-row groups of
 * A_i (and of A_i^T) hold a nonzero and whether the two physical components
 * have disjoint column supports; each chain step then skips the k-steps of
 * its products over all-zero row groups, and where the columns are disjoint
 * it forms X = E (M_0 + M_1) and Y = (M_0 + M_1)^T X once instead of the
 * four products X_i = E M_i, Y_ab = M_a^T X_b (Y_ab is a block of Y).  Only
 * exact zeros are dropped, so energies are the same with it off (tested).
 * The same pass stores F = A_0 + A_1 of every site (both copies), which
 * holds both components exactly where their supports are disjoint, and
 * marks each chain whose steps mostly have such structure; the I/Z energy
 * chains read F (half the fragment bytes) and the rest keep the full sites.
 * The states of CNOT-entangled ansatze have this structure: in the 40-qubit
 * workload every bulk site's A_0 and A_1 occupy disjoint halves of the left
 * bond.  The real reverse gradient's forward chains (I/Z forests) use the
 * structure of its bond-doubling state too, the suffix chains contracting
 * over their rows in even-then-odd order: with it off, gradients agree to
 * rounding rather than bit for bit.  No reference counterpart
 * (simulator_mps_dist.cu contracts dense sites). */
int gvmps_batch_set_sparsity(gvmps_batch *b, int on);

/* FP64 flops the forest-chain launches of the batch's last energy call
 * executed, summed over its instances: per chain step the products the
 * kernel issues (full, row-split or merged, from the sites' zero structure
 * and the chain marks of that call; three real products per complex product
 * on the complex path).  For the roofline of the chain kernel; reads the
 * marks back with a blocking copy.  No reference counterpart. */
int gvmps_batch_chain_flops(gvmps_batch *b, double *flops);

/* FP64 flops the forward and backward chain launches of the batch's last
 * value_grad call executed, summed over its instances: per chain step the
 * forward's products (merged or row-split on the steps that ran over the
 * zero structure) and the backward's.  The complex path reports the
 * planner's count.  For the adjoint's MFMA fraction; reads the marks back
 * with a blocking copy.  No reference counterpart. */
int gvmps_batch_grad_flops(gvmps_batch *b, double *flops);

/* Launch configuration of the forest-chain kernels: wavefronts per workgroup
 * (1, 2 or 4) and work-item order (0 instance-major, 1 (default) longest
 * chains first within each XCD's share of the instance-major items, so each
 * XCD keeps the same instances in its L2); other orders are rejected.
 * Results do not depend on the configuration.
 * No reference counterpart
 * (the reference launches fixed shapes, sharded_mps.cuh:670-684); set by the
 * autotuner (gvmps/autotune.py). */
int gvmps_batch_set_launch(gvmps_batch *b, int waves_per_group, int order);

/* State-preparation method of the batch's energy, sample and evaluate calls:
 * 0 (default) applies the gates with the exact-bond split of the plugin
 * (sharded_mps.cuh:465-572, dist_svd.cuh:194-283): the pivoted QR / Jacobi
 * SVD and the rank rule at every gate, except for problems whose splits all
 * keep the whole of Theta (exact_split in the problem info: zero cutoffs and
 * no gate cut below min(2 Dl, 2 Dr)), which take the exact split
 * Theta = I Theta or Theta I with no decomposition; 2 runs the decomposition
 * at every gate even then; 1 builds the same state by exact CNOT bond
 * doubling, the representation of the reference's reverse engine
 * (test/tools/reverse_mps.py:45-50), with no SVD.  Mode 1 needs every
 * problem's tape to have a doubling form within 16-wide bonds (grad_ok).
 * Energies and probabilities are gauge-invariant: all modes give them for
 * the same state, equal within rounding.  read_sites returns the mode's own
 * tensors (no bond profile in mode 1). */
int gvmps_batch_set_prep(gvmps_batch *b, int mode);

/* Arithmetic of state preparation, energy forests and reverse gradients.
 * When every problem of the batch has real gate matrices and no Y in its
 * Pauli words (real_arith in the problem info), mode 0 (default) runs them in
 * real arithmetic: the state is exactly real, so each complex product reduces
 * to one real product.  A batch that mixes such problems with others runs
 * all of its instances in complex arithmetic.  Mode 1 forces the complex
 * kernels.  Results of the two agree to rounding; bitwise equality between a
 * batch and one-at-a-time calls holds for the same arithmetic.  No reference
 * counterpart (the plugin is complex throughout, sharded_mps.cuh:614-687). */
int gvmps_batch_set_arith(gvmps_batch *b, int mode);

/* Energies plus, for shots > 0, `shots` samples per instance from ONE state
 * preparation.  Replaces an observe() followed by a sample() on the same
 * circuit (simulator_mps_dist.cu:1434-1485, 1587-1612), which prepares the
 * state twice.  seeds_dev: one uint32 per instance on the device; bits_dev
 * (instances*shots*max_n bytes, site order) and packed_dev (instances*shots*2
 * uint64) are both required when shots > 0. */
int gvmps_batch_evaluate(gvmps_batch *b, const double *params_dev, int shots, const uint32_t *seeds_dev,
                         double *energies_dev, uint8_t *bits_dev, uint64_t *packed_dev, void *stream);

/* Record gvmps_batch_evaluate's launch sequence (kernels and one device copy,
 * no memsets) into a HIP graph bound to these device buffers; replay runs it
 * on `stream` and checks the guard status.  Update the parameter / seed
 * buffers in place between replays.  No reference counterpart: the
 * reference issues every gate, SVD and chunk as separate synchronous calls. */
int gvmps_batch_capture(gvmps_batch *b, const double *params_dev, int shots, const uint32_t *seeds_dev,
                        double *energies_dev, uint8_t *bits_dev, uint64_t *packed_dev);
int gvmps_batch_replay(gvmps_batch *b, void *stream);

/* Per-phase device timing with HIP events recorded on the launch stream.
 * Energy calls: 0 state preparation, 2 forest chains, 3 energy reduction.
 * value_grad: 0 bond-doubling build, 1 forward chains, 2 energy reduction,
 * 3 backward chains, 4 site-adjoint reduction, 5 local backward.
 * Slots 6 and 7: the summed device time (ms) and the number of launches of
 * the dominant kernel alone, from events recorded right before and after
 * each of its launches (energy calls: the forest-chain kernel; value_grad:
 * the backward chain kernel).  enable != 0 resets the accumulators.  No
 * reference counterpart (the reference times whole calls from Python,
 * test/tools/analytic_execution.py:53-62); used by the benchmark. */
int gvmps_batch_set_timing(gvmps_batch *b, int enable);
int gvmps_batch_get_timing(gvmps_batch *b, double *phase_ms /* 8 */, long long *calls);

/* Statistics of each instance's last state preparation, 8 ints per instance:
 * Jacobi sweeps summed over its two-site gates, then the wavefront clock
 * counts (in units of 16 clocks, summed over gates) of the Theta build, the
 * pivoted QR, the Jacobi sweeps and the rank rule + write-back.  Diagnostic;
 * no reference counterpart. */
int gvmps_batch_prep_stats(gvmps_batch *b, int *stats_host);

/* Read back the prepared site tensors of instance `inst` (host buffer of
 * n*16*2*16 complex) and its bond dimensions (n+1 ints).  Test/debug aid. */
int gvmps_batch_read_sites(gvmps_batch *b, int inst, double *sites_host, int *bonds_host);

/* ------------------------------------------------------------------
 * Learned warm start: a message-passing network over Hamiltonian
 * interaction graphs, FP32, trained and run by the library's own kernels.
 * No reference counterpart: the reference's warm start is the partition
 * heuristic (Qiskit_VQE/independent_vqe.py:140-305), which stays the
 * fallback.  Network: H0 = relu(X W_in^T + b), then `layers` times
 * H_l = relu(H_{l-1} Ws^T + A_l Wn^T + b) with A_l[v] = sum_u w_uv H_{l-1}[u],
 * then Y = H_L W_out^T + b.  Loss: mean(1 - cos(Y - target)).
 * All graph arrays are device pointers; a batch of graphs is one graph with
 * concatenated nodes. */
typedef struct gvmps_gnn gvmps_gnn;
typedef struct gvmps_graph {
  int32_t n_nodes, n_edges;
  const float *features;                                  /* n_nodes x in_features */
  const int32_t *rows, *cols; const float *weights;       /* CSR by destination node */
  const int32_t *rows_t, *cols_t; const float *weights_t; /* the same edges by source node */
} gvmps_graph;

int gvmps_gnn_create(gvmps_gnn **out, int in_features, int hidden, int layers, int outputs,
                     uint64_t seed, int device);
int gvmps_gnn_destroy(gvmps_gnn *m);
int gvmps_gnn_param_count(const gvmps_gnn *m);
/* Flat FP32 parameters: W_in, b_in, per layer (Ws, Wn, b), W_out, b_out
 * (row-major, out x in).  Either pointer may be NULL; writing resets Adam. */
int gvmps_gnn_params(gvmps_gnn *m, float *host_out, const float *host_in);
int gvmps_gnn_forward(gvmps_gnn *m, const gvmps_graph *g, float *out_dev, void *stream);
/* Forward + loss + backward into the handle's gradient; optional host copies
 * of the loss and of the gradient (same layout as the parameters). */
int gvmps_gnn_loss_grad(gvmps_gnn *m, const gvmps_graph *g, const float *targets_dev,
                        float *loss_host, float *grad_host, void *stream);
/* Forward + backward from a caller-supplied dL/dY (n_nodes x outputs, device),
 * e.g. the MPS energy gradient with respect to the predicted angles. */
int gvmps_gnn_backward(gvmps_gnn *m, const gvmps_graph *g, const float *dy_dev, float *grad_host,
                       void *stream);
int gvmps_gnn_adam(gvmps_gnn *m, float lr, float beta1, float beta2, float eps, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GVMPS_API_H_ */
