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
// order (MFMA accumulation included), so results are reproducible bit for
// bit.
#include <hip/hip_runtime.h>

#include <cmath>

typedef float f4 __attribute__((ext_vector_type(4)));

// The dense products run on FP32 MFMA (v_mfma_f32_16x16x4f32), one wavefront
// per 16 x 16 output tile: the A operand (16 x 4) has lane l on row l % 16,
// column l / 16; the B operand (4 x 16) lane l on row l / 16, column l % 16;
// the result D lane j + 16 g, register v holds D[4 g + v][j].  Operands
// outside the matrices are zero, so ragged sizes need no special tiles.
__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Z = X1 W1^T (+ X2 W2^T) + b;  Y = relu(Z) or Z.  X: N x I, W: O x I.
extern "C" __global__ void __launch_bounds__(256)
k_gnn_dense(int N, int I, int O, const float *__restrict__ X1, const float *__restrict__ W1,
            const float *__restrict__ X2, const float *__restrict__ W2, const float *__restrict__ b,
            float *__restrict__ Z, float *__restrict__ Y, int relu) {
  const int lane = threadIdx.x & 63;
  const int tiles_o = (O + 15) >> 4;
  const int tile = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int n0 = (tile / tiles_o) * 16, o0 = (tile % tiles_o) * 16;
  if (n0 >= N) return;  // whole wave
  const int r = lane & 15, k = lane >> 4;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int pass = 0; pass < (X2 ? 2 : 1); ++pass) {
    const float *X = pass ? X2 : X1, *W = pass ? W2 : W1;
    for (int i0 = 0; i0 < I; i0 += 4) {
      const int i = i0 + k;
      const float a = (n0 + r < N && i < I) ? X[(size_t)(n0 + r) * I + i] : 0.f;
      const float w = (o0 + r < O && i < I) ? W[(size_t)(o0 + r) * I + i] : 0.f;
      acc = mfma4(a, w, acc);
    }
  }
  const int o = o0 + r, g = k;
  if (o >= O) return;
  const float bias = b ? b[o] : 0.f;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int n = n0 + 4 * g + v;
    if (n < N) {
      const float z = acc[v] + bias;
      if (Z) Z[(size_t)n * O + o] = z;
      Y[(size_t)n * O + o] = relu ? fmaxf(z, 0.f) : z;
    }
  }
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

// dX[n, i] (+)= sum_o dZ[n, o] W[o, i]: a wavefront per 16 x 16 tile of dX,
// MFMA over o.
extern "C" __global__ void __launch_bounds__(256)
k_gnn_dense_bwd_x(int N, int I, int O, const float *__restrict__ dZ, const float *__restrict__ W,
                  float *__restrict__ dX, int acc_out) {
  const int lane = threadIdx.x & 63;
  const int tiles_i = (I + 15) >> 4;
  const int tile = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  const int n0 = (tile / tiles_i) * 16, i0 = (tile % tiles_i) * 16;
  if (n0 >= N) return;
  const int r = lane & 15, k = lane >> 4;
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int q0 = 0; q0 < O; q0 += 4) {
    const int q = q0 + k;
    const float a = (n0 + r < N && q < O) ? dZ[(size_t)(n0 + r) * O + q] : 0.f;
    const float w = (i0 + r < I && q < O) ? W[(size_t)q * I + i0 + r] : 0.f;
    acc = mfma4(a, w, acc);
  }
  const int i = i0 + r, g = k;
  if (i >= I) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int n = n0 + 4 * g + v;
    if (n < N) {
      float *d = dX + (size_t)n * I + i;
      *d = acc_out ? *d + acc[v] : acc[v];
    }
  }
}

// dW[o, i] += sum_n dZ[n, o] X[n, i]; db[o] += sum_n dZ[n, o].  A workgroup
// per 16 x 16 tile of dW: its 4 wavefronts take consecutive quarters of the
// nodes (MFMA over n), and the quarters are added in order; db by 16 threads
// of the tiles at i = 0, each over the nodes in order.
extern "C" __global__ void __launch_bounds__(256)
k_gnn_dense_bwd_w(int N, int I, int O, const float *__restrict__ dZ, const float *__restrict__ X,
                  float *__restrict__ dW, float *__restrict__ db) {
  __shared__ float part[4][4][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tiles_i = (I + 15) >> 4;
  const int o0 = (blockIdx.x / tiles_i) * 16, i0 = (blockIdx.x % tiles_i) * 16;
  const int r = lane & 15, k = lane >> 4;
  const int quarter = ((N + 15) >> 4) * 4;  // a multiple of 4 nodes
  const int c0 = min(N, wv * quarter), c1 = min(N, c0 + quarter);
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int m0 = c0; m0 < c1; m0 += 4) {
    const int n = m0 + k;
    const float a = (n < c1 && o0 + r < O) ? dZ[(size_t)n * O + o0 + r] : 0.f;
    const float x = (n < c1 && i0 + r < I) ? X[(size_t)n * I + i0 + r] : 0.f;
    acc = mfma4(a, x, acc);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) part[wv][v][lane] = acc[v];
  __syncthreads();
  if (wv == 0) {
    const int i = i0 + r, g = k;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int o = o0 + 4 * g + v;
      const float tot = ((part[0][v][lane] + part[1][v][lane]) + part[2][v][lane]) + part[3][v][lane];
      if (o < O && i < I) dW[(size_t)o * I + i] += tot;
    }
  }
  if (db && i0 == 0 && threadIdx.x < 16 && o0 + (int)threadIdx.x < O) {
    const int o = o0 + threadIdx.x;
    float sb = 0.f;
    for (int n = 0; n < N; ++n) sb += dZ[(size_t)n * O + o];
    db[o] += sb;
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
  const long long tiles = (long long)((N + 15) / 16) * ((O + 15) / 16);
  hipLaunchKernelGGL(k_gnn_dense, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, s, N, I, O, X1, W1, X2, W2, b, Z,
                     Y, relu);
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
  const long long tiles = (long long)((N + 15) / 16) * ((I + 15) / 16);
  hipLaunchKernelGGL(k_gnn_dense_bwd_x, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, s, N, I, O, dZ, W, dX, acc);
  return hipGetLastError();
}

hipError_t gl_dense_bwd_w(hipStream_t s, int N, int I, int O, const float *dZ, const float *X, float *dW,
                          float *db) {
  hipLaunchKernelGGL(k_gnn_dense_bwd_w, dim3((unsigned)(((O + 15) / 16) * ((I + 15) / 16))), dim3(256), 0, s, N, I, O,
                     dZ, X, dW, db);
  return hipGetLastError();
}

hipError_t gl_adam(hipStream_t s, int P, float *theta, const float *g, float *m, float *v, float lr, float b1,
                   float b2, float eps, float c1, float c2) {
  hipLaunchKernelGGL(k_gnn_adam, grid_for(P), dim3(256), 0, s, P, theta, g, m, v, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

}  // namespace gv
