// Learned T2 initial guess: a per-amplitude MLP (6 features -> 3 tanh layers of width h -> 1),
// trained and evaluated on the GPU.
//
// Parameter vector (FP32, flat): W1[h,6] b1[h] W2[h,h] b2[h] W3[h,h] b3[h] w4[h] b4, every W
// column-major [out, in].  Activations of a batch of B samples are column-major [h, B] (the h
// values of one sample contiguous), so every layer is one GEMM of the FP32-MFMA kernel
//     Z_l = W_l A_{l-1}  (M = h, N = B, K = in)
// followed by a fused bias + tanh pass; the backward pass is the transposed GEMMs
//     dW_l = dZ_l A_{l-1}^T,  dA_{l-1} = W_l^T dZ_l
// plus column reductions for the bias gradients (deterministic two-stage sums) and a fused
// tanh' pass.  Loss: L = 1/(2B) sum_s (y_s - t_s)^2.  Optimiser: Adam.
//
// Features of amplitude (i,j,a,b): T2_MP2[i,j,a,b], eps_i, eps_j, eps_a, eps_b, d_ij (occupied pair
// distance), each standardised with the model's (mean, 1/scale); the network predicts
// (T2_CCSD - T2_MP2) / y_scale, and the guess is symmetrised in (ij)(ab).
#include "fh_internal.h"

#include <cmath>

namespace fh {

size_t mlp_nparams(int h) { return (size_t)h * 6 + h + 2 * ((size_t)h * h + h) + h + 1; }

MlpOffsets mlp_offsets(int h) {
  MlpOffsets o;
  o.W1 = 0;
  o.b1 = o.W1 + (size_t)h * 6;
  o.W2 = o.b1 + h;
  o.b2 = o.W2 + (size_t)h * h;
  o.W3 = o.b2 + h;
  o.b3 = o.W3 + (size_t)h * h;
  o.w4 = o.b3 + h;
  o.b4 = o.w4 + h;
  return o;
}

static unsigned grid1(int64_t n, int bs = 256) {
  int64_t b = (n + bs - 1) / bs;
  return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

// A[h,B] = tanh(A + b[row])
__global__ void k_bias_tanh(int h, int64_t n, const float* __restrict__ b, float* __restrict__ A) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
    A[e] = tanhf(A[e] + b[e % h]);
}

// y[s] += b4  (after y = w4^T A3 as a GEMM with M = 1)
__global__ void k_add_scalar(int64_t n, const float* __restrict__ b, float* __restrict__ y) {
  for (int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x; s < n; s += (int64_t)gridDim.x * 256) y[s] += b[0];
}

// dZ3[r,s] = w4[r] dy[s] (1 - A3[r,s]^2), element order of [h,B] (coalesced)
__global__ void k_out_dz(int h, int64_t n, const float* __restrict__ w4, const float* __restrict__ A3,
                         const float* __restrict__ dy, float* __restrict__ dZ3) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const float a = A3[e];
    dZ3[e] = w4[e % h] * dy[e / h] * (1.0f - a * a);
  }
}

// dy[s] = (y[s] - t[s]) / B;  per-block partial sums of (y - t)^2 / 2
__global__ void k_out_back(int64_t B, const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ dy,
                           double* __restrict__ lpart) {
  __shared__ double red[256];
  double l = 0.0;
  for (int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x; s < B; s += (int64_t)gridDim.x * 256) {
    const float r = y[s] - t[s];
    l += 0.5 * (double)r * (double)r;
    dy[s] = r / (float)B;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

// dZ = dA * (1 - A^2)  (in place on dA)
__global__ void k_tanh_back(int64_t n, const float* __restrict__ A, float* __restrict__ dA) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
    dA[e] *= 1.0f - A[e] * A[e];
}

// partial row sums of M[h,B] (h divides 256): block b sums columns [b*cols, (b+1)*cols)
__global__ void k_rowsum_part(int h, int64_t B, int64_t cols, const float* __restrict__ Mx, float* __restrict__ part) {
  __shared__ float red[256];
  const int lanes = 256 / h, r = threadIdx.x % h, c0 = threadIdx.x / h;
  const int64_t beg = (int64_t)blockIdx.x * cols, end = beg + cols < B ? beg + cols : B;
  float t = 0.0f;
  for (int64_t c = beg + c0; c < end; c += lanes) t += Mx[r + (int64_t)h * c];
  red[threadIdx.x] = t;
  __syncthreads();
  if (threadIdx.x < h) {
    float u = 0.0f;
    for (int q = 0; q < lanes; ++q) u += red[threadIdx.x + q * h];
    part[(int64_t)blockIdx.x * h + threadIdx.x] = u;
  }
}
// out[r] = sum_b part[b, r]
__global__ void k_rowsum_final(int h, int nb, const float* __restrict__ part, float* __restrict__ out) {
  for (int r = blockIdx.x * 256 + threadIdx.x; r < h; r += gridDim.x * 256) {
    float u = 0.0f;
    for (int b = 0; b < nb; ++b) u += part[(int64_t)b * h + r];
    out[r] = u;
  }
}

static constexpr int64_t kRowsumCols = 4096;

static void rowsum(hipStream_t s, int h, int64_t B, const float* Mx, float* part, float* out) {
  const int nb = (int)((B + kRowsumCols - 1) / kRowsumCols);
  hipLaunchKernelGGL(k_rowsum_part, dim3(nb), dim3(256), 0, s, h, B, kRowsumCols, Mx, part);
  hipLaunchKernelGGL(k_rowsum_final, dim3(grid1(h)), dim3(256), 0, s, h, nb, part, out);
  FH_HIP(hipGetLastError());
}

int64_t mlp_partial_elems(int h, int64_t B) { return ((B + kRowsumCols - 1) / kRowsumCols) * std::max(h, 1) + 1024; }

void mlp_forward(hipStream_t s, int h, const float* P, int64_t B, const float* X, float* A1, float* A2, float* A3,
                 float* y) {
  FH_CHECK(h > 0 && 256 % h == 0, -1, "mlp: width must divide 256");
  const MlpOffsets o = mlp_offsets(h);
  gemm_f32(s, false, false, h, B, 6, 1.0, P + o.W1, h, X, 6, 0.0, A1, h);
  hipLaunchKernelGGL(k_bias_tanh, dim3(grid1(h * B)), dim3(256), 0, s, h, (int64_t)h * B, P + o.b1, A1);
  gemm_f32(s, false, false, h, B, h, 1.0, P + o.W2, h, A1, h, 0.0, A2, h);
  hipLaunchKernelGGL(k_bias_tanh, dim3(grid1(h * B)), dim3(256), 0, s, h, (int64_t)h * B, P + o.b2, A2);
  gemm_f32(s, false, false, h, B, h, 1.0, P + o.W3, h, A2, h, 0.0, A3, h);
  hipLaunchKernelGGL(k_bias_tanh, dim3(grid1(h * B)), dim3(256), 0, s, h, (int64_t)h * B, P + o.b3, A3);
  gemm_f32(s, false, false, 1, B, h, 1.0, P + o.w4, 1, A3, h, 0.0, y, 1);
  hipLaunchKernelGGL(k_add_scalar, dim3(grid1(B)), dim3(256), 0, s, B, P + o.b4, y);
  FH_HIP(hipGetLastError());
}

void mlp_backward(hipStream_t s, int h, const float* P, int64_t B, const float* X, const float* T, const float* A1,
                  const float* A2, const float* A3, const float* y, float* dy, float* dA, float* dZ, float* G,
                  float* part, double* lpart, int64_t* nlpart) {
  const MlpOffsets o = mlp_offsets(h);
  const unsigned gb = grid1(B);
  hipLaunchKernelGGL(k_out_back, dim3(gb), dim3(256), 0, s, B, y, T, dy, lpart);
  *nlpart = gb;
  hipLaunchKernelGGL(k_out_dz, dim3(grid1(h * B)), dim3(256), 0, s, h, (int64_t)h * B, P + o.w4, A3, dy, dZ);
  // output layer: g w4 = A3 dy, g b4 = sum dy
  gemm_f32(s, false, false, h, 1, B, 1.0, A3, h, dy, B, 0.0, G + o.w4, h);
  rowsum(s, 1, B, dy, part, G + o.b4);
  // layer 3
  gemm_f32(s, false, true, h, h, B, 1.0, dZ, h, A2, h, 0.0, G + o.W3, h);
  rowsum(s, h, B, dZ, part, G + o.b3);
  gemm_f32(s, true, false, h, B, h, 1.0, P + o.W3, h, dZ, h, 0.0, dA, h);
  hipLaunchKernelGGL(k_tanh_back, dim3(grid1(h * B)), dim3(256), 0, s, (int64_t)h * B, A2, dA);
  // layer 2 (dA now holds dZ2)
  gemm_f32(s, false, true, h, h, B, 1.0, dA, h, A1, h, 0.0, G + o.W2, h);
  rowsum(s, h, B, dA, part, G + o.b2);
  gemm_f32(s, true, false, h, B, h, 1.0, P + o.W2, h, dA, h, 0.0, dZ, h);
  hipLaunchKernelGGL(k_tanh_back, dim3(grid1(h * B)), dim3(256), 0, s, (int64_t)h * B, A1, dZ);
  // layer 1 (dZ holds dZ1)
  gemm_f32(s, false, true, h, 6, B, 1.0, dZ, h, X, 6, 0.0, G + o.W1, h);
  rowsum(s, h, B, dZ, part, G + o.b1);
  FH_HIP(hipGetLastError());
}

__global__ void k_adam(int64_t n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.0f - b1) * gg;
    const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
  }
}

void mlp_adam(hipStream_t s, int64_t n, float* P, const float* G, float* m, float* v, float lr, int64_t t) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)t), c2 = 1.0f - std::pow(b2, (float)t);
  hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(256), 0, s, n, P, G, m, v, lr, b1, b2, 1e-8f, c1, c2);
  FH_HIP(hipGetLastError());
}

// Xb[:, q] = X[:, perm[off + q]], Tb[q] = T[perm[off + q]]  for q < B
__global__ void k_gather(int64_t B, int64_t off, const int64_t* __restrict__ perm, const float* __restrict__ X,
                         const float* __restrict__ T, float* __restrict__ Xb, float* __restrict__ Tb) {
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < B; q += (int64_t)gridDim.x * 256) {
    const int64_t src = perm[off + q];
#pragma unroll
    for (int f = 0; f < 6; ++f) Xb[6 * q + f] = X[6 * src + f];
    Tb[q] = T[src];
  }
}

void mlp_gather(hipStream_t s, int64_t B, int64_t off, const int64_t* perm, const float* X, const float* T, float* Xb,
                float* Tb) {
  hipLaunchKernelGGL(k_gather, dim3(grid1(B)), dim3(256), 0, s, B, off, perm, X, T, Xb, Tb);
  FH_HIP(hipGetLastError());
}

// features of the amplitudes e0 .. e0+n-1 of T2[o,o,v,v] (column-major linear index)
__global__ void k_t2_features(int o, int v, int64_t e0, int64_t n, const double* __restrict__ t2,
                              const double* __restrict__ eo, const double* __restrict__ ev,
                              const double* __restrict__ dij, const float* __restrict__ xm,
                              const float* __restrict__ xs, float* __restrict__ X) {
  for (int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (int64_t)gridDim.x * 256) {
    const int64_t e = e0 + q;
    int64_t r = e;
    const int i = (int)(r % o); r /= o;
    const int j = (int)(r % o); r /= o;
    const int a = (int)(r % v);
    const int b = (int)(r / v);
    const float f[6] = {(float)t2[e], (float)eo[i], (float)eo[j], (float)ev[a], (float)ev[b],
                        dij ? (float)dij[i + (int64_t)o * j] : 0.0f};
#pragma unroll
    for (int k = 0; k < 6; ++k) X[6 * q + k] = (f[k] - xm[k]) * xs[k];
  }
}

void t2_features(hipStream_t s, int o, int v, int64_t e0, int64_t n, const double* t2, const double* eo,
                 const double* ev, const double* dij, const float* xm, const float* xs, float* X) {
  hipLaunchKernelGGL(k_t2_features, dim3(grid1(n)), dim3(256), 0, s, o, v, e0, n, t2, eo, ev, dij, xm, xs, X);
  FH_HIP(hipGetLastError());
}

// T2[i,j,a,b] += y_scale * (Y[i,j,a,b] + Y[j,i,b,a]) / 2
__global__ void k_t2_learned(int o, int v, double ys, const float* __restrict__ Y, double* __restrict__ t2) {
  const int64_t n = (int64_t)o * o * v * v;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    int64_t r = e;
    const int64_t i = r % o; r /= o;
    const int64_t j = r % o; r /= o;
    const int64_t a = r % v;
    const int64_t b = r / v;
    const int64_t et = j + o * (i + o * (b + v * a));
    t2[e] += ys * 0.5 * ((double)Y[e] + (double)Y[et]);
  }
}

void t2_learned_apply(hipStream_t s, int o, int v, double ys, const float* Y, double* t2) {
  hipLaunchKernelGGL(k_t2_learned, dim3(grid1((int64_t)o * o * v * v)), dim3(256), 0, s, o, v, ys, Y, t2);
  FH_HIP(hipGetLastError());
}

}  // namespace fh
