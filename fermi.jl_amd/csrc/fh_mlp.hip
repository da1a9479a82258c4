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

// ------------------------------------------------------------------ fused guess inference, h = 128
// One workgroup (256 threads = 4 wavefronts, 2 x 2 over hidden units x amplitudes) evaluates the
// network on 128 amplitudes: features and layer 1 (K = 6) on the VALU, layers 2 and 3 as
// [128 x 128] x [128 x 128] products on split-BF16x3 MFMA (hi/lo bf16 parts of weights and
// activations, three products, FP32 accumulation: about 2^-16 relative per product), the output
// layer as an in-register dot product.  Activations stay in LDS (amplitude-major rows, hi/lo
// planes); the weights of layers 2 and 3 are split once per call into bf16 planes, row-major
// [out][in], and read by every workgroup from L2.  tanh(x) = 1 - 2 / (e^{2x} + 1) with the
// hardware exp2 and reciprocal (absolute error ~1e-7).
namespace {
constexpr int kFH = 128;        // hidden width of the fused kernel
constexpr int kFN = 128;        // amplitudes per workgroup
constexpr int kFRow = kFH + 8;  // LDS row of one amplitude's activations (bf16; 16-byte aligned rows)

typedef __bf16 fbf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 fbf16x2 __attribute__((ext_vector_type(2)));
typedef float ff32x16 __attribute__((ext_vector_type(16)));
typedef float ff32x2 __attribute__((ext_vector_type(2)));

// (a, b) -> packed bf16 hi parts and packed bf16 lo parts (x - hi), round to nearest even
__device__ inline void split_pair(float a, float b, uint32_t& hi, uint32_t& lo) {
  const ff32x2 x = {a, b};
  const fbf16x2 h = __builtin_convertvector(x, fbf16x2);
  uint32_t hb;
  __builtin_memcpy(&hb, &h, 4);
  const ff32x2 r = {a - __uint_as_float(hb << 16), b - __uint_as_float(hb & 0xffff0000u)};
  const fbf16x2 l = __builtin_convertvector(r, fbf16x2);
  __builtin_memcpy(&lo, &l, 4);
  hi = hb;
}

__device__ inline float fast_tanh(float x) {
  const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);   // e^{2x} = 2^{2x log2 e}
  return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}
}  // namespace

// Wp[layer][hi/lo][out][in] from the column-major W2 / W3 of the parameter vector
__global__ void k_mlp_split_w(const float* __restrict__ P, size_t offW2, size_t offW3, uint32_t* __restrict__ Wp) {
  constexpr int pairs = kFH * kFH / 2;
  for (int e = blockIdx.x * 256 + threadIdx.x; e < 2 * pairs; e += gridDim.x * 256) {
    const int L = e / pairs, r = e - L * pairs, out = r / (kFH / 2), in = 2 * (r % (kFH / 2));
    const float* W = P + (L ? offW3 : offW2);
    uint32_t hi, lo;
    split_pair(W[out + (size_t)kFH * in], W[out + (size_t)kFH * (in + 1)], hi, lo);
    Wp[(2 * L) * pairs + r] = hi;
    Wp[(2 * L + 1) * pairs + r] = lo;
  }
}

__global__ __launch_bounds__(256, 2) void k_t2_mlp_fused(int o, int v, int64_t n, const double* __restrict__ t2,
                                                         const double* __restrict__ eo, const double* __restrict__ ev,
                                                         const double* __restrict__ dij, const float* __restrict__ xm,
                                                         const float* __restrict__ xs, const float* __restrict__ P,
                                                         MlpOffsets off, const uint16_t* __restrict__ Wp,
                                                         float* __restrict__ Y) {
  __shared__ __attribute__((aligned(16))) uint16_t Hs[2][kFN * kFRow];  // [hi/lo][amplitude][hidden]
  __shared__ float Xs[kFN][6];
  __shared__ float red[2][kFN];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6, wm = wid & 1, wn = wid >> 1;
  const int64_t e0 = (int64_t)blockIdx.x * kFN;
  if (tid < kFN) {
    const int64_t e = e0 + tid;
    float f[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (e < n) {
      int64_t r = e;
      const int i = (int)(r % o); r /= o;
      const int j = (int)(r % o); r /= o;
      const int a = (int)(r % v);
      const int b = (int)(r / v);
      f[0] = (float)t2[e]; f[1] = (float)eo[i]; f[2] = (float)eo[j]; f[3] = (float)ev[a]; f[4] = (float)ev[b];
      f[5] = dij ? (float)dij[i + (int64_t)o * j] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) Xs[tid][k] = (f[k] - xm[k]) * xs[k];
  }
  __syncthreads();
  {
    // layer 1: amplitude nn, hidden units h0 .. h0+63
    const int nn = tid & (kFN - 1), h0 = (tid >> 7) * 64;
    float x[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) x[k] = Xs[nn][k];
    const float* W1 = P + off.W1;
    const float* b1 = P + off.b1;
#pragma unroll 4
    for (int q = 0; q < 64; q += 2) {
      float z[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int hh = h0 + q + t;
        float acc = b1[hh];
#pragma unroll
        for (int k = 0; k < 6; ++k) acc = fmaf(W1[hh + kFH * k], x[k], acc);
        z[t] = fast_tanh(acc);
      }
      uint32_t hi, lo;
      split_pair(z[0], z[1], hi, lo);
      *reinterpret_cast<uint32_t*>(&Hs[0][nn * kFRow + h0 + q]) = hi;
      *reinterpret_cast<uint32_t*>(&Hs[1][nn * kFRow + h0 + q]) = lo;
    }
  }
  __syncthreads();
  const int li = lane & 31, kq = 8 * (lane >> 5);
  for (int L = 0; L < 2; ++L) {
    // Z[hid, amp] = W[hid, :] . H[:, amp]; wavefront (wm, wn): hid 64 wm .. +63, amp 64 wn .. +63
    ff32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int jj = 0; jj < 2; ++jj)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][jj][r] = 0.0f;
    const uint16_t* Wh = Wp + (size_t)(2 * L) * kFH * kFH;
    const uint16_t* Wl = Wh + (size_t)kFH * kFH;
#pragma unroll 2
    for (int k0 = 0; k0 < kFH; k0 += 16) {
      fbf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const int row = wm * 64 + i * 32 + li;
        ah[i] = *reinterpret_cast<const fbf16x8*>(Wh + row * kFH + k0 + kq);
        al[i] = *reinterpret_cast<const fbf16x8*>(Wl + row * kFH + k0 + kq);
      }
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        const int col = wn * 64 + jj * 32 + li;
        bh[jj] = *reinterpret_cast<const fbf16x8*>(&Hs[0][col * kFRow + k0 + kq]);
        bl[jj] = *reinterpret_cast<const fbf16x8*>(&Hs[1][col * kFRow + k0 + kq]);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          acc[i][jj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[jj], acc[i][jj], 0, 0, 0);
          acc[i][jj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[jj], acc[i][jj], 0, 0, 0);
          acc[i][jj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[jj], acc[i][jj], 0, 0, 0);
        }
    }
    __syncthreads();  // every wavefront has read H
    // accumulator element r of lane l: hidden row 8 (r / 4) + 4 (l / 32) + r % 4 of the block,
    // amplitude column l % 32
    const float* bL = P + (L ? off.b3 : off.b2);
    if (L == 0) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int col = wn * 64 + jj * 32 + li;
#pragma unroll
          for (int r = 0; r < 16; r += 2) {
            const int hid = wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            uint32_t hi, lo;
            split_pair(fast_tanh(acc[i][jj][r] + bL[hid]), fast_tanh(acc[i][jj][r + 1] + bL[hid + 1]), hi, lo);
            *reinterpret_cast<uint32_t*>(&Hs[0][col * kFRow + hid]) = hi;
            *reinterpret_cast<uint32_t*>(&Hs[1][col * kFRow + hid]) = lo;
          }
        }
      __syncthreads();
    } else {
      const float* w4 = P + off.w4;
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        float part = 0.0f;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const int hid = wm * 64 + i * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            part = fmaf(w4[hid], fast_tanh(acc[i][jj][r] + bL[hid]), part);
          }
        part += __shfl_xor(part, 32, 64);
        if (lane < 32) red[wm][wn * 64 + jj * 32 + li] = part;
      }
      __syncthreads();
      if (tid < kFN && e0 + tid < n) Y[e0 + tid] = red[0][tid] + red[1][tid] + P[off.b4];
    }
  }
}

bool t2_mlp_fused_available(int h) { return h == kFH; }
size_t t2_mlp_fused_scratch_bytes() { return (size_t)4 * kFH * kFH * sizeof(uint16_t); }

void t2_mlp_fused(hipStream_t s, int o, int v, const double* t2, const double* eo, const double* ev, const double* dij,
                  const float* xm, const float* xs, const float* P, void* wsplit, float* Y) {
  const MlpOffsets off = mlp_offsets(kFH);
  hipLaunchKernelGGL(k_mlp_split_w, dim3(64), dim3(256), 0, s, P, off.W2, off.W3, static_cast<uint32_t*>(wsplit));
  const int64_t n = (int64_t)o * o * v * v;
  const int64_t nb = (n + kFN - 1) / kFN;
  FH_CHECK(nb < ((int64_t)1 << 31), -3, "t2_mlp_fused: too many amplitudes");
  hipLaunchKernelGGL(k_t2_mlp_fused, dim3((unsigned)nb), dim3(256), 0, s, o, v, n, t2, eo, ev, dij, xm, xs, P, off,
                     static_cast<const uint16_t*>(wsplit), Y);
  FH_HIP(hipGetLastError());
}

}  // namespace fh
