// Graph network of the learned warm start (gfx950, FP32): forward, backward
// and Adam kernels.  No reference counterpart (the reference's warm start is
// the partition heuristic, Qiskit_VQE/independent_vqe.py:140-305).
//
// Nodes are the qubits of a batch of interaction graphs (concatenated); edges
// come as CSR by destination (aggregation) and CSR by source (its adjoint),
// each with the row-normalised coupling weight.  The network:
//   H0 = relu(X W_in^T + b_in)
//   A_l = sum_u w_uv H_{l-1}[u];  H_l = relu(H_{l-1} Ws_l^T + A_l Wn_l^T + b_l)
//   Y  = H_L W_out^T + b_out                       (angles, 4 per node)
// Loss = mean over nodes and outputs of (1 - cos(Y - target)), the 2*pi-
// periodic distance of rotation angles.  Every reduction runs in a fixed
// order, so results are reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>

// Z = X1 W1^T (+ X2 W2^T) + b;  Y = relu(Z) or Z.  One thread per (n, o).
extern "C" __global__ void __launch_bounds__(256)
k_gnn_dense(int N, int I, int O, const float *__restrict__ X1, const float *__restrict__ W1,
            const float *__restrict__ X2, const float *__restrict__ W2, const float *__restrict__ b,
            float *__restrict__ Z, float *__restrict__ Y, int relu) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * O) return;
  const int n = (int)(t / O), o = (int)(t % O);
  float acc = b ? b[o] : 0.f;
  const float *x = X1 + (size_t)n * I, *w = W1 + (size_t)o * I;
  for (int i = 0; i < I; ++i) acc = fmaf(x[i], w[i], acc);
  if (X2) {
    const float *x2 = X2 + (size_t)n * I, *w2 = W2 + (size_t)o * I;
    for (int i = 0; i < I; ++i) acc = fmaf(x2[i], w2[i], acc);
  }
  if (Z) Z[t] = acc;
  Y[t] = relu ? fmaxf(acc, 0.f) : acc;
}

// A[v, d] = sum_{e in row v} w[e] H[col[e], d]  (or += into A when acc).
extern "C" __global__ void __launch_bounds__(256)
k_gnn_aggregate(int N, int D, const int *__restrict__ rows, const int *__restrict__ cols,
                const float *__restrict__ w, const float *__restrict__ H, float *__restrict__ A, int acc) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * D) return;
  const int v = (int)(t / D), d = (int)(t % D);
  float s = 0.f;
  for (int e = rows[v]; e < rows[v + 1]; ++e) s = fmaf(w[e], H[(size_t)cols[e] * D + d], s);
  A[t] = acc ? A[t] + s : s;
}

// Loss terms and dL/dY: per element 1 - cos(y - t) and sin(y - t) / count.
// Per-block partial sums of the loss in a fixed tree order.
extern "C" __global__ void __launch_bounds__(256)
k_gnn_loss(int count, const float *__restrict__ Y, const float *__restrict__ T, float *__restrict__ dY,
           float *__restrict__ partial) {
  __shared__ float s[256];
  const int t = blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (t < count) {
    const float d = Y[t] - T[t];
    l = 1.f - cosf(d);
    dY[t] = sinf(d) / (float)count;
  }
  s[threadIdx.x] = l;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

extern "C" __global__ void __launch_bounds__(256)
k_gnn_sum(int n, const float *__restrict__ partial, float scale, float *__restrict__ out) {
  __shared__ float s[256];
  float a = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) a += partial[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0] * scale;
}

// dZ = dY * (Z > 0)
extern "C" __global__ void __launch_bounds__(256)
k_gnn_relu_bwd(long long count, const float *__restrict__ Z, const float *__restrict__ dY, float *__restrict__ dZ) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) dZ[t] = Z[t] > 0.f ? dY[t] : 0.f;
}

// dX[n, i] (+)= sum_o dZ[n, o] W[o, i].  One thread per (n, i).
extern "C" __global__ void __launch_bounds__(256)
k_gnn_dense_bwd_x(int N, int I, int O, const float *__restrict__ dZ, const float *__restrict__ W,
                  float *__restrict__ dX, int acc) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)N * I) return;
  const int n = (int)(t / I), i = (int)(t % I);
  float s = 0.f;
  const float *dz = dZ + (size_t)n * O;
  for (int o = 0; o < O; ++o) s = fmaf(dz[o], W[(size_t)o * I + i], s);
  dX[t] = acc ? dX[t] + s : s;
}

// dW[o, i] += sum_n dZ[n, o] X[n, i]; db[o] += sum_n dZ[n, o].  One block
// per o; thread (i + I*c) sums nodes n = c, c+C, ... (C = 256 / I chunks),
// then the C partials are added in chunk order.
extern "C" __global__ void __launch_bounds__(256)
k_gnn_dense_bwd_w(int N, int I, int O, const float *__restrict__ dZ, const float *__restrict__ X,
                  float *__restrict__ dW, float *__restrict__ db) {
  __shared__ float part[256];
  __shared__ float bpart[256];
  const int o = blockIdx.x;
  const int C = 256 / I;
  const int i = threadIdx.x % I, c = threadIdx.x / I;
  float s = 0.f, sb = 0.f;
  if (c < C) {
    for (int n = c; n < N; n += C) {
      const float g = dZ[(size_t)n * O + o];
      s = fmaf(g, X[(size_t)n * I + i], s);
      if (i == 0) sb += g;
    }
  }
  part[threadIdx.x] = s;
  bpart[threadIdx.x] = sb;
  __syncthreads();
  if (threadIdx.x < I) {
    float tot = 0.f;
    for (int k = 0; k < C; ++k) tot += part[k * I + threadIdx.x];
    dW[(size_t)o * I + threadIdx.x] += tot;
  }
  if (threadIdx.x == 0 && db) {
    float tb = 0.f;
    for (int k = 0; k < C; ++k) tb += bpart[k * I];
    db[o] += tb;
  }
}

// Adam (bias-corrected), one thread per parameter.
extern "C" __global__ void __launch_bounds__(256)
k_gnn_adam(int P, float *__restrict__ theta, const float *__restrict__ g, float *__restrict__ m,
           float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= P) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.f - b1) * gt;
  const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  theta[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
}

namespace gv {

static inline dim3 grid_for(long long count) { return dim3((unsigned)((count + 255) / 256)); }

hipError_t gl_dense(hipStream_t s, int N, int I, int O, const float *X1, const float *W1, const float *X2,
                    const float *W2, const float *b, float *Z, float *Y, int relu) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gnn_dense, grid_for((long long)N * O), dim3(256), 0, s, N, I, O, X1, W1, X2, W2, b, Z, Y,
                     relu);
  return hipGetLastError();
}

hipError_t gl_aggregate(hipStream_t s, int N, int D, const int *rows, const int *cols, const float *w,
                        const float *H, float *A, int acc) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gnn_aggregate, grid_for((long long)N * D), dim3(256), 0, s, N, D, rows, cols, w, H, A, acc);
  return hipGetLastError();
}

hipError_t gl_loss(hipStream_t s, int count, const float *Y, const float *T, float *dY, float *partial,
                   float *loss) {
  const int nb = (count + 255) / 256;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gnn_loss, dim3(nb), dim3(256), 0, s, count, Y, T, dY, partial);
  hipLaunchKernelGGL(k_gnn_sum, dim3(1), dim3(256), 0, s, nb, partial, 1.f / (float)count, loss);
  return hipGetLastError();
}

hipError_t gl_relu_bwd(hipStream_t s, long long count, const float *Z, const float *dY, float *dZ) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gnn_relu_bwd, grid_for(count), dim3(256), 0, s, count, Z, dY, dZ);
  return hipGetLastError();
}

hipError_t gl_dense_bwd_x(hipStream_t s, int N, int I, int O, const float *dZ, const float *W, float *dX, int acc) {
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gnn_dense_bwd_x, grid_for((long long)N * I), dim3(256), 0, s, N, I, O, dZ, W, dX, acc);
  return hipGetLastError();
}

hipError_t gl_dense_bwd_w(hipStream_t s, int N, int I, int O, const float *dZ, const float *X, float *dW,
                          float *db) {
  hipLaunchKernelGGL(k_gnn_dense_bwd_w, dim3(O), dim3(256), 0, s, N, I, O, dZ, X, dW, db);
  return hipGetLastError();
}

hipError_t gl_adam(hipStream_t s, int P, float *theta, const float *g, float *m, float *v, float lr, float b1,
                   float b2, float eps, float c1, float c2) {
  hipLaunchKernelGGL(k_gnn_adam, grid_for(P), dim3(256), 0, s, P, theta, g, m, v, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

}  // namespace gv
