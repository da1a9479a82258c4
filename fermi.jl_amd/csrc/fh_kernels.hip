// Bandwidth-bound kernels of the DF-RCCSD(T) engine: tensor permutes, amplitude
// updates with on-the-fly denominators, deterministic FP64 reductions, DIIS linear
// combinations, the CC energy and the (T) energy epilogue.
//
// Reductions are deterministic: every kernel writes one FP64 partial per workgroup
// (fixed grid for a given length) and a single workgroup adds the partials in a fixed
// order, so repeated runs give bitwise-identical results.
#include "fh_internal.h"

namespace fh {

static constexpr int RB = 256;        // threads per reduction block
static constexpr int64_t RMAX = 2048; // max partials per reduction

int64_t reduce_blocks_needed(int64_t n) {
  int64_t b = (n + RB * 4 - 1) / (RB * 4);
  if (b < 1) b = 1;
  if (b > RMAX) b = RMAX;
  return b;
}

static __device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// block-wide sum, result valid in thread 0
static __device__ __forceinline__ double block_sum(double v) {
  __shared__ double red[RB / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < RB / 64; ++i) t += red[i];
  }
  return t;
}

// ------------------------------------------------------------------ permute
struct PermArgs {
  int nd;
  int64_t odim[6], istr[6], ostr[6];
};

// out element per thread; used when output dim 0 reads the input with unit stride
__global__ __launch_bounds__(256) void k_permute_simple(PermArgs a, int64_t n, double alpha, const double* __restrict__ in,
                                                        double beta, double* __restrict__ out) {
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * 256) {
    int64_t rem = idx, off = 0;
#pragma unroll
    for (int d = 0; d < 6; ++d) {
      if (d < a.nd) {
        const int64_t c = rem % a.odim[d];
        rem /= a.odim[d];
        off += c * a.istr[d];
      }
    }
    double v = alpha * in[off];
    if (beta != 0.0) v += beta * out[idx];
    out[idx] = v;
  }
}

// 32x32 tiles over (output dim 0, output dim p) where dim p is the input's unit-stride dim:
// reads are coalesced along p, writes along output dim 0.
__global__ __launch_bounds__(256) void k_permute_tiled(PermArgs a, int p, int64_t nrest, int64_t t0, int64_t tp,
                                                       double alpha, const double* __restrict__ in, double beta,
                                                       double* __restrict__ out) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int64_t blk = blockIdx.x; blk < nrest * t0 * tp; blk += gridDim.x) {
    const int64_t b0 = blk % t0, bp = (blk / t0) % tp, rest = blk / (t0 * tp);
    // decompose the remaining dims
    int64_t rem = rest, ioff = 0, ooff = 0;
    for (int d = 1; d < a.nd; ++d) {
      if (d == p) continue;
      const int64_t c = rem % a.odim[d];
      rem /= a.odim[d];
      ioff += c * a.istr[d];
      ooff += c * a.ostr[d];
    }
    // load: threads consecutive along p
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t i0 = b0 * 32 + ty + 8 * r, ip = bp * 32 + tx;
      if (i0 < a.odim[0] && ip < a.odim[p]) tile[ty + 8 * r][tx] = in[ioff + i0 * a.istr[0] + ip];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t i0 = b0 * 32 + tx, ip = bp * 32 + ty + 8 * r;
      if (i0 < a.odim[0] && ip < a.odim[p]) {
        const int64_t o = ooff + i0 + ip * a.ostr[p];
        double v = alpha * tile[tx][ty + 8 * r];
        if (beta != 0.0) v += beta * out[o];
        out[o] = v;
      }
    }
    __syncthreads();
  }
}

void permute_f64(hipStream_t s, int nd, const int64_t* odim, const int64_t* istr, double alpha, const double* in,
                 double beta, double* out) {
  FH_CHECK(nd >= 1 && nd <= 6, -3, "permute: rank out of range");
  PermArgs a;
  a.nd = nd;
  int64_t n = 1;
  for (int d = 0; d < 6; ++d) { a.odim[d] = 1; a.istr[d] = 0; a.ostr[d] = 0; }
  for (int d = 0; d < nd; ++d) {
    a.odim[d] = odim[d];
    a.istr[d] = istr[d];
    a.ostr[d] = n;
    n *= odim[d];
  }
  if (n == 0) return;
  int p = -1;
  if (a.istr[0] != 1 || a.odim[0] == 1) {
    for (int d = 1; d < nd; ++d)
      if (a.istr[d] == 1 && a.odim[d] > 1) p = d;
  }
  if (p < 0 || a.odim[0] < 8) {
    int64_t g = (n + 255) / 256;
    if (g > 65536 * 4) g = 65536 * 4;
    hipLaunchKernelGGL(k_permute_simple, dim3((unsigned)g), dim3(256), 0, s, a, n, alpha, in, beta, out);
  } else {
    const int64_t t0 = (a.odim[0] + 31) / 32, tp = (a.odim[p] + 31) / 32;
    const int64_t nrest = n / (a.odim[0] * a.odim[p]);
    int64_t g = nrest * t0 * tp;
    if (g > 65536 * 4) g = 65536 * 4;
    hipLaunchKernelGGL(k_permute_tiled, dim3((unsigned)g), dim3(256), 0, s, a, p, nrest, t0, tp, alpha, in, beta, out);
  }
  FH_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ elementwise
__global__ void k_axpby(int64_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    y[i] = (b == 0.0) ? a * x[i] : a * x[i] + b * y[i];
}
__global__ void k_fill(int64_t n, double v, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) y[i] = v;
}
static unsigned grid_for(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}
void axpby_f64(hipStream_t s, int64_t n, double a, const double* x, double b, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(256), 0, s, n, a, x, b, y);
  FH_HIP(hipGetLastError());
}
void fill_f64(hipStream_t s, int64_t n, double v, double* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(256), 0, s, n, v, y);
  FH_HIP(hipGetLastError());
}

// tau[i,j,a,b] = t2[i,j,a,b] + c * t1[i,a] t1[j,b]
__global__ void k_tau(int o, int v, const double* __restrict__ t1, const double* __restrict__ t2, double c,
                      double* __restrict__ out) {
  const int64_t n = (int64_t)o * o * v * v;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * 256) {
    const int64_t i = idx % o, j = (idx / o) % o, a = (idx / ((int64_t)o * o)) % v, b = idx / ((int64_t)o * o * v);
    out[idx] = t2[idx] + c * t1[i + o * a] * t1[j + o * b];
  }
}
void tau_f64(hipStream_t s, int o, int v, const double* t1, const double* t2, double c, double* out) {
  const int64_t n = (int64_t)o * o * v * v;
  hipLaunchKernelGGL(k_tau, dim3(grid_for(n)), dim3(256), 0, s, o, v, t1, t2, c, out);
  FH_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ Jacobi update (RCCSDHelper.jl:347-348)
template <bool T2, bool F32>
__global__ __launch_bounds__(RB) void k_jacobi(int o, int v, const double* __restrict__ eo, const double* __restrict__ ev,
                                               double* __restrict__ R, const double* __restrict__ old, void* err,
                                               double* __restrict__ partials) {
  const int64_t n = T2 ? (int64_t)o * o * v * v : (int64_t)o * v;
  double acc = 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * RB + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * RB) {
    double d;
    if (T2) {
      const int64_t i = idx % o, j = (idx / o) % o, a = (idx / ((int64_t)o * o)) % v, b = idx / ((int64_t)o * o * v);
      d = eo[i] + eo[j] - ev[a] - ev[b];
    } else {
      const int64_t i = idx % o, a = idx / o;
      d = eo[i] - ev[a];
    }
    const double nv = R[idx] / d;
    R[idx] = nv;
    const double e = nv - old[idx];
    if (err) {
      if (F32) static_cast<float*>(err)[idx] = (float)e;
      else static_cast<double*>(err)[idx] = e;
    }
    acc += e * e;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}
void jacobi_update(hipStream_t s, int o, int v, bool is_t2, const double* eo, const double* ev, double* R,
                   const double* old, void* err, bool err_f32, double* partials, int64_t* nblocks) {
  const int64_t n = is_t2 ? (int64_t)o * o * v * v : (int64_t)o * v;
  const int64_t g = reduce_blocks_needed(n);
  *nblocks = g;
  if (is_t2) {
    if (err_f32) hipLaunchKernelGGL((k_jacobi<true, true>), dim3(g), dim3(RB), 0, s, o, v, eo, ev, R, old, err, partials);
    else hipLaunchKernelGGL((k_jacobi<true, false>), dim3(g), dim3(RB), 0, s, o, v, eo, ev, R, old, err, partials);
  } else {
    if (err_f32) hipLaunchKernelGGL((k_jacobi<false, true>), dim3(g), dim3(RB), 0, s, o, v, eo, ev, R, old, err, partials);
    else hipLaunchKernelGGL((k_jacobi<false, false>), dim3(g), dim3(RB), 0, s, o, v, eo, ev, R, old, err, partials);
  }
  FH_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ reductions
template <typename T>
__global__ __launch_bounds__(RB) void k_dot(int64_t n, const T* __restrict__ x, const T* __restrict__ y,
                                            double* __restrict__ partials) {
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (int64_t)gridDim.x * RB)
    acc += (double)x[i] * (double)y[i];
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}
__global__ __launch_bounds__(RB) void k_sum(int64_t n, const double* __restrict__ partials, double* __restrict__ result) {
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += RB) acc += partials[i];
  const double t = block_sum(acc);
  if (threadIdx.x == 0) *result = t;
}
void sum_partials(hipStream_t s, int64_t n, const double* partials, double* result) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(RB), 0, s, n, partials, result);
  FH_HIP(hipGetLastError());
}
void dot_f64(hipStream_t s, int64_t n, const double* x, const double* y, double* partials, double* result) {
  const int64_t g = reduce_blocks_needed(n);
  hipLaunchKernelGGL((k_dot<double>), dim3(g), dim3(RB), 0, s, n, x, y, partials);
  FH_HIP(hipGetLastError());
  sum_partials(s, g, partials, result);
}
void dot_f32(hipStream_t s, int64_t n, const float* x, const float* y, double* partials, double* result) {
  const int64_t g = reduce_blocks_needed(n);
  hipLaunchKernelGGL((k_dot<float>), dim3(g), dim3(RB), 0, s, n, x, y, partials);
  FH_HIP(hipGetLastError());
  sum_partials(s, g, partials, result);
}

// ------------------------------------------------------------------ DIIS extrapolation (DIIS.jl:86-93)
struct LinArgs {
  int nv;
  const double* x[8];
  const void* e[8];
  double c[8];
};
template <bool F32>
__global__ void k_lincomb(int64_t n, LinArgs a, bool add_err, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    double acc = 0.0;
    for (int q = 0; q < a.nv; ++q) acc += a.c[q] * a.x[q][i];
    if (add_err) {
      double ae = 0.0;
      for (int q = 0; q < a.nv; ++q)
        ae += a.c[q] * (F32 ? (double)static_cast<const float*>(a.e[q])[i] : static_cast<const double*>(a.e[q])[i]);
      acc += ae;
    }
    y[i] = acc;
  }
}
void lincomb(hipStream_t s, int64_t n, int nv, const double* const* xs, const double* cs, const void* const* es,
             bool err_f32, bool add_err, double* y) {
  FH_CHECK(nv >= 1 && nv <= 8, -3, "lincomb: 1..8 vectors");
  LinArgs a;
  a.nv = nv;
  for (int q = 0; q < 8; ++q) { a.x[q] = q < nv ? xs[q] : nullptr; a.e[q] = q < nv ? es[q] : nullptr; a.c[q] = q < nv ? cs[q] : 0.0; }
  if (err_f32) hipLaunchKernelGGL((k_lincomb<true>), dim3(grid_for(n)), dim3(256), 0, s, n, a, add_err, y);
  else hipLaunchKernelGGL((k_lincomb<false>), dim3(grid_for(n)), dim3(256), 0, s, n, a, add_err, y);
  FH_HIP(hipGetLastError());
}

__global__ void k_cast(int64_t n, const double* __restrict__ x, float* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) y[i] = (float)x[i];
}
void cast_f64_to_f32(hipStream_t s, int64_t n, const double* x, float* y) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cast, dim3(grid_for(n)), dim3(256), 0, s, n, x, y);
  FH_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ CC energy (RCCSDHelper.jl:8-23)
// E = sum_{k,c,l,d} (kc|ld) * (2 T2[k,l,c,d] - T2[l,k,c,d] - T1[l,c] T1[k,d] + 2 T1[k,c] T1[l,d]),
// iterated in the memory order of OVOV[k,c,l,d] so that the integral stream is coalesced.
__global__ __launch_bounds__(RB) void k_cc_energy(int o, int v, const double* __restrict__ t1, const double* __restrict__ t2,
                                                  const double* __restrict__ ovov, double* __restrict__ partials) {
  const int64_t n = (int64_t)o * v * o * v;
  const int64_t oo = (int64_t)o * o;
  double acc = 0.0;
  for (int64_t idx = (int64_t)blockIdx.x * RB + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * RB) {
    const int64_t k = idx % o, c = (idx / o) % v, l = (idx / ((int64_t)o * v)) % o, d = idx / ((int64_t)o * v * o);
    const double t2klcd = t2[k + o * l + oo * c + oo * v * d];
    const double t2lkcd = t2[l + o * k + oo * c + oo * v * d];
    const double b = 2.0 * t2klcd - t2lkcd - t1[l + o * c] * t1[k + o * d] + 2.0 * t1[k + o * c] * t1[l + o * d];
    acc += b * ovov[idx];
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}
void cc_energy(hipStream_t s, int o, int v, const double* t1, const double* t2, const double* ovov, double* partials,
               double* result) {
  const int64_t n = (int64_t)o * v * o * v;
  const int64_t g = reduce_blocks_needed(n);
  hipLaunchKernelGGL(k_cc_energy, dim3(g), dim3(RB), 0, s, o, v, t1, t2, ovov, partials);
  FH_HIP(hipGetLastError());
  sum_partials(s, g, partials, result);
}

// ------------------------------------------------------------------ (T) energy epilogue (ijk.jl:116-136)
// For one occupied triple the reference's restricted sum over a>=b>=c with weights
// 1/(1+d_ab+d_bc) equals the full sum over ordered (a,b,c) of
//     W_abc * (4V_abc + V_bca + V_cab - 2V_acb - 2V_bac - 2V_cba) / D_abc
// (the summand Ef of ijk.jl:132 is symmetric in (a,b,c); a=b=c terms vanish), times
// (2 - d_ij - d_jk).  W[a,b,c] for the k-chunk lies in one buffer with a fastest
// (W + a + v*b + v^2*c + v^3*kk).  The virtual space is cut into 8x8x8 cubes; one workgroup
// takes an orbit {A,B,C} (A>=B>=C) of cube blocks, loads W on the (up to) six permuted cubes
// into LDS (8-element runs along a), forms V = W + T1 x (ia|jb) terms and sums the orbit, so
// every element of W is read once per triple.
static constexpr int CT = 8, CT3 = CT * CT * CT;

template <typename TW>
__global__ __launch_bounds__(256) void k_triples_energy(int o, int v, int i, int j, int k0, const TW* __restrict__ W,
                                                        const int* __restrict__ orbits,
                                                        const double* __restrict__ t1p, const double* __restrict__ vvoo,
                                                        const double* __restrict__ eo, const double* __restrict__ ev,
                                                        double* __restrict__ partials) {
  __shared__ double Wc[6][CT3];
  __shared__ double Vc[6][CT3];
  __shared__ int slot_of[6];
  const int tid = threadIdx.x;
  const int kk = blockIdx.y, k = k0 + kk;
  const int ob = blockIdx.x;
  const double wocc = 2.0 - (i == j ? 1.0 : 0.0) - (j == k ? 1.0 : 0.0);
  const int A = orbits[3 * ob], B = orbits[3 * ob + 1], C = orbits[3 * ob + 2];
  // the six orderings of the orbit's blocks: (abc) (acb) (bac) (bca) (cab) (cba)
  const int pb[6][3] = {{A, B, C}, {A, C, B}, {B, A, C}, {B, C, A}, {C, A, B}, {C, B, A}};
  if (tid < 6) {
    int s = tid;
    for (int q = 0; q < tid; ++q)
      if (pb[q][0] == pb[tid][0] && pb[q][1] == pb[tid][1] && pb[q][2] == pb[tid][2]) { s = q; break; }
    slot_of[tid] = s;
  }
  __syncthreads();
  double acc = 0.0;
  if (wocc != 0.0) {
    const int64_t vv = (int64_t)v * v;
    const TW* Wk = W + vv * v * kk;
    const double* vjk = vvoo + vv * (j + (int64_t)o * k);
    const double* vik = vvoo + vv * (i + (int64_t)o * k);
    const double* vij = vvoo + vv * (i + (int64_t)o * j);
    // load W and form V = W + T1p[x,i] vvoo[y,z,j,k] + vvoo[x,z,i,k] T1p[y,j] + vvoo[x,y,i,j] T1p[z,k]  (ijk.jl:116)
    for (int p = 0; p < 6; ++p) {
      if (slot_of[p] != p) continue;
      const int P0 = pb[p][0] * CT, Q0 = pb[p][1] * CT, R0 = pb[p][2] * CT;
      for (int e = tid; e < CT3; e += 256) {
        const int a = P0 + (e & 7), b = Q0 + ((e >> 3) & 7), c = R0 + (e >> 6);
        double w = 0.0, val = 0.0;
        if (a < v && b < v && c < v) {
          w = (double)Wk[a + (int64_t)v * b + vv * c];
          val = w + t1p[a + (int64_t)v * i] * vjk[b + (int64_t)v * c] + vik[a + (int64_t)v * c] * t1p[b + (int64_t)v * j] +
                vij[a + (int64_t)v * b] * t1p[c + (int64_t)v * k];
        }
        Wc[p][e] = w;
        Vc[p][e] = val;
      }
    }
    __syncthreads();
    const double Dijk = eo[i] + eo[j] + eo[k];
    for (int p = 0; p < 6; ++p) {
      if (slot_of[p] != p) continue;
      const int X = pb[p][0], Y = pb[p][1], Z = pb[p][2];
      // slots holding the cubes (Y,Z,X) (Z,X,Y) (X,Z,Y) (Y,X,Z) (Z,Y,X)
      auto find = [&](int x0, int y0, int z0) {
        int r = 0;
        for (int q = 0; q < 6; ++q)
          if (pb[q][0] == x0 && pb[q][1] == y0 && pb[q][2] == z0) { r = slot_of[q]; break; }
        return r;
      };
      const int s_bca = find(Y, Z, X), s_cab = find(Z, X, Y), s_acb = find(X, Z, Y), s_bac = find(Y, X, Z),
                s_cba = find(Z, Y, X);
      for (int e = tid; e < CT3; e += 256) {
        const int la = e & 7, lb = (e >> 3) & 7, lc = e >> 6;
        const int a = X * CT + la, b = Y * CT + lb, c = Z * CT + lc;
        if (a < v && b < v && c < v) {
          const double Vabc = Vc[p][e];
          const double Vbca = Vc[s_bca][lb + CT * lc + CT * CT * la];
          const double Vcab = Vc[s_cab][lc + CT * la + CT * CT * lb];
          const double Vacb = Vc[s_acb][la + CT * lc + CT * CT * lb];
          const double Vbac = Vc[s_bac][lb + CT * la + CT * CT * lc];
          const double Vcba = Vc[s_cba][lc + CT * lb + CT * CT * la];
          const double Zs = 4.0 * Vabc + Vbca + Vcab - 2.0 * (Vacb + Vbac + Vcba);
          acc += Wc[p][e] * Zs / (Dijk - ev[a] - ev[b] - ev[c]);
        }
      }
    }
    acc *= wocc;
  }
  const double t = block_sum(acc);
  if (tid == 0) partials[blockIdx.x + (int64_t)gridDim.x * kk] = t;
}

int triples_orbit_count(int v) {
  const int nb = (v + CT - 1) / CT;
  return nb * (nb + 1) * (nb + 2) / 6;
}

void triples_orbits(int v, std::vector<int>& out) {
  const int nb = (v + CT - 1) / CT;
  out.clear();
  for (int a = 0; a < nb; ++a)
    for (int b = 0; b <= a; ++b)
      for (int c = 0; c <= b; ++c) { out.push_back(a); out.push_back(b); out.push_back(c); }
}

template <typename TW>
void triples_energy(hipStream_t s, int o, int v, int i, int j, int k0, int nk, const TW* W, const int* orbits,
                    const double* t1p, const double* vvoo, const double* eo, const double* ev, double* partials,
                    int64_t* nblocks) {
  const int norb = triples_orbit_count(v);
  *nblocks = (int64_t)norb * nk;
  hipLaunchKernelGGL(k_triples_energy<TW>, dim3((unsigned)norb, (unsigned)nk), dim3(256), 0, s, o, v, i, j, k0, W,
                     orbits, t1p, vvoo, eo, ev, partials);
  FH_HIP(hipGetLastError());
}
template void triples_energy<double>(hipStream_t, int, int, int, int, int, int, const double*, const int*, const double*,
                                     const double*, const double*, const double*, double*, int64_t*);
template void triples_energy<float>(hipStream_t, int, int, int, int, int, int, const float*, const int*, const double*,
                                    const double*, const double*, const double*, double*, int64_t*);

}  // namespace fh
