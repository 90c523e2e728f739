// C ABI of the warm-start graph network (declared in include/gvmps.h).  The
// handle owns the FP32 parameters, their gradient and Adam moments, and the
// activation workspace; the caller owns the graph arrays and targets.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gvmps.h"
#include "api_internal.h"

namespace gv {

hipError_t gl_dense(hipStream_t s, int N, int I, int O, const float *X1, const float *W1, const float *X2,
                    const float *W2, const float *b, float *Z, float *Y, int relu);
hipError_t gl_aggregate(hipStream_t s, int N, int D, const int *rows, const int *cols, const float *w,
                        const float *H, float *A, int acc);
hipError_t gl_loss(hipStream_t s, int count, const float *Y, const float *T, float *dY, float *partial,
                   float *loss);
hipError_t gl_relu_bwd(hipStream_t s, long long count, const float *Z, const float *dY, float *dZ);
hipError_t gl_dense_bwd_x(hipStream_t s, int N, int I, int O, const float *dZ, const float *W, float *dX, int acc);
hipError_t gl_dense_bwd_w(hipStream_t s, int N, int I, int O, const float *dZ, const float *X, float *dW,
                          float *db);
hipError_t gl_adam(hipStream_t s, int P, float *theta, const float *g, float *m, float *v, float lr, float b1,
                   float b2, float eps, float c1, float c2);

}  // namespace gv

using namespace gv;

struct gvmps_gnn {
  int device = 0, F = 0, H = 0, L = 0, O = 0, P = 0;
  float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  long long step = 0;
  // workspace for up to cap nodes
  int cap = 0;
  std::vector<float *> Z, Hs, A;  // Z[0..L], Hs[0..L], A[1..L] (A[0] unused)
  float *Y = nullptr, *dY = nullptr, *dH = nullptr, *dH2 = nullptr, *dZ = nullptr, *dA = nullptr;
  float *partial = nullptr, *loss = nullptr;
  std::vector<void *> ws;

  // parameter offsets
  size_t w_in() const { return 0; }
  size_t b_in() const { return (size_t)H * F; }
  size_t w_s(int l) const { return b_in() + H + (size_t)l * (2 * H * H + H); }
  size_t w_n(int l) const { return w_s(l) + (size_t)H * H; }
  size_t b_l(int l) const { return w_n(l) + (size_t)H * H; }
  size_t w_out() const { return w_s(L); }
  size_t b_out() const { return w_out() + (size_t)O * H; }

  void free_ws() {
    for (void *p : ws) (void)hipFree(p);
    ws.clear();
    Z.clear();
    Hs.clear();
    A.clear();
    cap = 0;
  }
  float *alloc(size_t n) {
    void *p = nullptr;
    hip_check(hipMalloc(&p, sizeof(float) * (n ? n : 1)), "hipMalloc gnn workspace");
    ws.push_back(p);
    return static_cast<float *>(p);
  }
  void ensure(int N) {
    if (N <= cap) return;
    free_ws();
    const size_t nh = (size_t)N * H;
    for (int l = 0; l <= L; ++l) {
      Z.push_back(alloc(nh));
      Hs.push_back(alloc(nh));
      A.push_back(alloc(nh));
    }
    Y = alloc((size_t)N * O);
    dY = alloc((size_t)N * O);
    dH = alloc(nh);
    dH2 = alloc(nh);
    dZ = alloc(nh);
    dA = alloc(nh);
    partial = alloc((size_t)(N * O + 255) / 256 + 1);
    loss = alloc(1);
    cap = N;
  }
  ~gvmps_gnn() {
    (void)hipSetDevice(device);
    free_ws();
    for (float *p : {theta, grad, m, v})
      if (p) (void)hipFree(p);
  }
};

namespace {
thread_local std::string g_gnn_error;

uint64_t mix64(uint64_t x) {  // SplitMix64 finaliser, for reproducible initial weights
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

void check_graph(const gvmps_gnn &M, const gvmps_graph *g) {
  if (!g || g->n_nodes < 0 || g->n_edges < 0) throw std::runtime_error("invalid graph");
  if (g->n_nodes > 0 && (!g->features || !g->rows || !g->rows_t))
    throw std::runtime_error("graph arrays missing");
  if (g->n_edges > 0 && (!g->cols || !g->weights || !g->cols_t || !g->weights_t))
    throw std::runtime_error("graph edge arrays missing");
  (void)M;
}

void forward(gvmps_gnn &M, const gvmps_graph *g, hipStream_t s) {
  const int N = g->n_nodes, F = M.F, H = M.H;
  M.ensure(N);
  const float *th = M.theta;
  hip_check(gl_dense(s, N, F, H, g->features, th + M.w_in(), nullptr, nullptr, th + M.b_in(), M.Z[0], M.Hs[0], 1),
            "gnn input layer");
  for (int l = 0; l < M.L; ++l) {
    hip_check(gl_aggregate(s, N, H, g->rows, g->cols, g->weights, M.Hs[l], M.A[l + 1], 0), "gnn aggregate");
    hip_check(gl_dense(s, N, H, H, M.Hs[l], th + M.w_s(l), M.A[l + 1], th + M.w_n(l), th + M.b_l(l), M.Z[l + 1],
                       M.Hs[l + 1], 1),
              "gnn message layer");
  }
  hip_check(gl_dense(s, N, H, M.O, M.Hs[M.L], th + M.w_out(), nullptr, nullptr, th + M.b_out(), nullptr, M.Y, 0),
            "gnn output layer");
}

// targets != NULL: loss and dL/dY from the angle loss; otherwise dY_in is dL/dY.
void backward(gvmps_gnn &M, const gvmps_graph *g, const float *targets, const float *dY_in, hipStream_t s) {
  const int N = g->n_nodes, F = M.F, H = M.H, O = M.O;
  const float *th = M.theta;
  float *gr = M.grad;
  hip_check(hipMemsetAsync(gr, 0, sizeof(float) * M.P, s), "zero gradient");
  if (targets)
    hip_check(gl_loss(s, N * O, M.Y, targets, M.dY, M.partial, M.loss), "gnn loss");
  else
    hip_check(hipMemcpyAsync(M.dY, dY_in, sizeof(float) * N * O, hipMemcpyDeviceToDevice, s), "copy dY");
  // output layer
  hip_check(gl_dense_bwd_w(s, N, H, O, M.dY, M.Hs[M.L], gr + M.w_out(), gr + M.b_out()), "gnn dW_out");
  hip_check(gl_dense_bwd_x(s, N, H, O, M.dY, th + M.w_out(), M.dH, 0), "gnn dH_L");
  for (int l = M.L - 1; l >= 0; --l) {
    const long long nh = (long long)N * H;
    hip_check(gl_relu_bwd(s, nh, M.Z[l + 1], M.dH, M.dZ), "gnn relu bwd");
    hip_check(gl_dense_bwd_w(s, N, H, H, M.dZ, M.Hs[l], gr + M.w_s(l), gr + M.b_l(l)), "gnn dWs");
    hip_check(gl_dense_bwd_w(s, N, H, H, M.dZ, M.A[l + 1], gr + M.w_n(l), nullptr), "gnn dWn");
    hip_check(gl_dense_bwd_x(s, N, H, H, M.dZ, th + M.w_s(l), M.dH2, 0), "gnn dH self");
    hip_check(gl_dense_bwd_x(s, N, H, H, M.dZ, th + M.w_n(l), M.dA, 0), "gnn dA");
    hip_check(gl_aggregate(s, N, H, g->rows_t, g->cols_t, g->weights_t, M.dA, M.dH2, 1), "gnn aggregate bwd");
    std::swap(M.dH, M.dH2);
  }
  hip_check(gl_relu_bwd(s, (long long)N * H, M.Z[0], M.dH, M.dZ), "gnn relu bwd");
  hip_check(gl_dense_bwd_w(s, N, F, H, M.dZ, g->features, gr + M.w_in(), gr + M.b_in()), "gnn dW_in");
}
}  // namespace

#define GNN_TRY(...)                        \
  try {                                     \
    g_gnn_error.clear();                    \
    __VA_ARGS__;                            \
    return 0;                               \
  } catch (const std::exception &ex) {      \
    g_gnn_error = ex.what();                \
    gv::set_error(ex.what());               \
    return 1;                               \
  }

extern "C" {

int gvmps_gnn_create(gvmps_gnn **out, int in_features, int hidden, int layers, int outputs, uint64_t seed,
                     int device) {
  GNN_TRY({
    if (!out) throw std::runtime_error("null output pointer");
    if (in_features < 1 || in_features > 256 || hidden < 1 || hidden > 256 || layers < 0 || outputs < 1 ||
        outputs > 256)
      throw std::runtime_error("network sizes out of range (1..256)");
    hip_check(hipSetDevice(device), "hipSetDevice");
    auto M = std::make_unique<gvmps_gnn>();
    M->device = device;
    M->F = in_features;
    M->H = hidden;
    M->L = layers;
    M->O = outputs;
    M->P = (int)M->b_out() + outputs;
    std::vector<float> h(M->P, 0.f);
    auto fill = [&](size_t off, int fan_out, int fan_in, float gain) {
      const float a = gain * std::sqrt(6.f / (float)fan_in);
      for (size_t k = 0; k < (size_t)fan_out * fan_in; ++k) {
        const float u = (float)(mix64(seed * 0x100000001B3ULL + off + k) >> 40) * (1.f / 16777216.f);
        h[off + k] = a * (2.f * u - 1.f);
      }
    };
    fill(M->w_in(), hidden, in_features, 1.f);
    for (int l = 0; l < layers; ++l) {
      fill(M->w_s(l), hidden, hidden, 0.7f);
      fill(M->w_n(l), hidden, hidden, 0.7f);
    }
    fill(M->w_out(), outputs, hidden, 0.3f);
    for (float **p : {&M->theta, &M->grad, &M->m, &M->v}) {
      hip_check(hipMalloc((void **)p, sizeof(float) * M->P), "hipMalloc gnn");
      hip_check(hipMemset(*p, 0, sizeof(float) * M->P), "hipMemset gnn");
    }
    hip_check(hipMemcpy(M->theta, h.data(), sizeof(float) * M->P, hipMemcpyHostToDevice), "upload gnn");
    *out = M.release();
  })
}

int gvmps_gnn_destroy(gvmps_gnn *m) { GNN_TRY({ delete m; }) }

int gvmps_gnn_param_count(const gvmps_gnn *m) { return m ? m->P : 0; }

int gvmps_gnn_params(gvmps_gnn *m, float *host_out, const float *host_in) {
  GNN_TRY({
    if (!m) throw std::runtime_error("null network");
    hip_check(hipSetDevice(m->device), "hipSetDevice");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (host_out)
      hip_check(hipMemcpy(host_out, m->theta, sizeof(float) * m->P, hipMemcpyDeviceToHost), "read params");
    if (host_in) {
      hip_check(hipMemcpy(m->theta, host_in, sizeof(float) * m->P, hipMemcpyHostToDevice), "write params");
      hip_check(hipMemset(m->m, 0, sizeof(float) * m->P), "reset adam");
      hip_check(hipMemset(m->v, 0, sizeof(float) * m->P), "reset adam");
      m->step = 0;
    }
  })
}

int gvmps_gnn_forward(gvmps_gnn *m, const gvmps_graph *g, float *out_dev, void *stream) {
  GNN_TRY({
    if (!m) throw std::runtime_error("null network");
    check_graph(*m, g);
    hipStream_t s = (hipStream_t)stream;
    hip_check(hipSetDevice(m->device), "hipSetDevice");
    forward(*m, g, s);
    if (out_dev && g->n_nodes)
      hip_check(hipMemcpyAsync(out_dev, m->Y, sizeof(float) * g->n_nodes * m->O, hipMemcpyDeviceToDevice, s),
                "copy outputs");
  })
}

int gvmps_gnn_loss_grad(gvmps_gnn *m, const gvmps_graph *g, const float *targets_dev, float *loss_host,
                        float *grad_host, void *stream) {
  GNN_TRY({
    if (!m) throw std::runtime_error("null network");
    check_graph(*m, g);
    if (!targets_dev && g->n_nodes) throw std::runtime_error("null targets");
    hipStream_t s = (hipStream_t)stream;
    hip_check(hipSetDevice(m->device), "hipSetDevice");
    if (g->n_nodes == 0) throw std::runtime_error("empty graph batch");
    forward(*m, g, s);
    backward(*m, g, targets_dev, nullptr, s);
    if (loss_host)
      hip_check(hipMemcpyAsync(loss_host, m->loss, sizeof(float), hipMemcpyDeviceToHost, s), "read loss");
    if (grad_host)
      hip_check(hipMemcpyAsync(grad_host, m->grad, sizeof(float) * m->P, hipMemcpyDeviceToHost, s), "read grad");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  })
}

int gvmps_gnn_backward(gvmps_gnn *m, const gvmps_graph *g, const float *dy_dev, float *grad_host, void *stream) {
  GNN_TRY({
    if (!m) throw std::runtime_error("null network");
    check_graph(*m, g);
    if (g->n_nodes == 0 || !dy_dev) throw std::runtime_error("empty graph batch or null dL/dY");
    hipStream_t s = (hipStream_t)stream;
    hip_check(hipSetDevice(m->device), "hipSetDevice");
    forward(*m, g, s);
    backward(*m, g, nullptr, dy_dev, s);
    if (grad_host)
      hip_check(hipMemcpyAsync(grad_host, m->grad, sizeof(float) * m->P, hipMemcpyDeviceToHost, s), "read grad");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  })
}

int gvmps_gnn_adam(gvmps_gnn *m, float lr, float beta1, float beta2, float eps, void *stream) {
  GNN_TRY({
    if (!m) throw std::runtime_error("null network");
    hip_check(hipSetDevice(m->device), "hipSetDevice");
    m->step += 1;
    const float c1 = 1.f - std::pow(beta1, (float)m->step);
    const float c2 = 1.f - std::pow(beta2, (float)m->step);
    hip_check(gl_adam((hipStream_t)stream, m->P, m->theta, m->grad, m->m, m->v, lr, beta1, beta2, eps, c1, c2),
              "gnn adam");
  })
}

}  // extern "C"
