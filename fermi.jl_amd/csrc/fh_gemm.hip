// Dense GEMM on CDNA4 matrix cores: D = alpha * op(A) * op(B) + beta * C, column-major.
//
// Workgroup: 256 threads = 4 wavefronts in a 2x2 arrangement; each wavefront owns a
// (BM/2)x(BN/2) block of the output as 16x16 MFMA accumulators.  Operand tiles are staged
// through LDS with double buffering (global loads for stage s+1 are in flight while stage s
// is multiplied).  LDS row paddings put the fragment reads of a half (FP64) or full (FP32)
// wavefront on distinct bank pairs.
//
// Precision variants (template F32 and the global element types TI / TO):
//   F32 = false: FP64 operands, v_mfma_f64_16x16x4f64, FP64 accumulation (TI = TO = double);
//   F32 = true : operands rounded to FP32 when staged into LDS (TI may be double or float),
//                v_mfma_f32_16x16x4f32 with FP32 accumulation, alpha/beta epilogue in FP64,
//                output stored as TO (double or float).
//
// Tile -> workgroup mapping: consecutive workgroup ids land on different XCDs (round robin
// over 8), so the linear id is first remapped such that each XCD receives a contiguous range
// of tiles.  Within that order tiles are rastered M-fastest (tiles sharing a column block of
// B are neighbours), or -- when A is the larger operand -- in groups of gn N-tiles whose B
// blocks fit in L2: each group sweeps all M-tiles, so B stays in L2 and A is streamed once
// per group instead of once per N-tile.
#include "fh_internal.h"

#include <algorithm>
#include <type_traits>

namespace fh {

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int BM, int BN, int BK, bool TA, bool TB, int OCC, int TAG, bool F32, typename TI, typename TO, bool LM,
          bool SEG>
__global__ __launch_bounds__(256, OCC) void k_gemm(
    int64_t M, int64_t N, int64_t K, double alpha, const TI* __restrict__ A, int64_t lda, int64_t sA,
    const TI* __restrict__ B, int64_t ldb, int64_t sB, double beta, const TO* C, int64_t ldc, int64_t sC, TO* D,
    int64_t ldd, int64_t sD, int tilesM, int tilesN, int64_t kchunk, int gn, OutMap om, KSeg sg) {
  static_assert(F32 || std::is_same<TI, double>::value, "FP64 MFMA needs FP64 operands");
  using T = typename std::conditional<F32, float, double>::type;
  using Acc = typename std::conditional<F32, f32x4, f64x4>::type;
  constexpr int PADA = 16, PADB = F32 ? 4 : 2;
  constexpr int WM = BM / 2, WN = BN / 2, MB = WM / 16, NB = WN / 16;
  constexpr int LA = BM * BK / 256, LB = BN * BK / 256;
  __shared__ T As[2][BK][BM + PADA];
  __shared__ T Bs[2][BN][BK + PADB];

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid & 1, wn = wid >> 1;

  // XCD-aware remap of the linear tile id, then the raster
  const int64_t nwg = (int64_t)tilesM * tilesN;
  const int64_t w = blockIdx.x;
  const int64_t q = nwg >> 3, r = nwg & 7, x = w & 7;
  const int64_t L = x * q + (x < r ? x : r) + (w >> 3);
  int64_t mt, nt;
  if (gn <= 0) {
    mt = L % tilesM;
    nt = L / tilesM;
  } else {
    const int64_t g = L / ((int64_t)tilesM * gn);
    const int64_t rem = L - g * (int64_t)tilesM * gn;
    const int64_t wdt = (tilesN - g * gn < gn) ? tilesN - g * gn : gn;
    mt = rem / wdt;
    nt = g * gn + rem % wdt;
  }
  const int64_t m0 = mt * BM, n0 = nt * BN;

  const int64_t bz = blockIdx.y;
  if (kchunk > 0) {
    // split-K: slice bz covers k in [bz*kchunk, min(K, (bz+1)*kchunk)) and writes its own D slice
    const int64_t kb = bz * kchunk;
    K = (K - kb < kchunk) ? K - kb : kchunk;
    A += TA ? kb : lda * kb;
    B += TB ? ldb * kb : kb;
  } else {
    A += bz * sA;
    B += bz * sB;
  }
  if (C) C += bz * sC;
  D += bz * sD;
  // second K segment (SEG; split-K is never combined with it)
  const TI* A2 = SEG ? static_cast<const TI*>(sg.A) + bz * sg.sA : nullptr;
  const TI* B2 = SEG ? static_cast<const TI*>(sg.B) + bz * sg.sB : nullptr;
  const int64_t K1 = SEG ? sg.K1 : K;

  TI ra[LA], rb[LB];
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int t = 0; t < LA; ++t) {
      const int e = tid + 256 * t;
      int mm, kk;
      if (!TA) { mm = e % BM; kk = e / BM; } else { kk = e % BK; mm = e / BK; }
      const int64_t gm = m0 + mm, gk = k0 + kk;
      TI val = 0;
      if (gm < M && gk < K) {
        if (!SEG || gk < K1) val = TA ? A[gk + lda * gm] : A[gm + lda * gk];
        else val = TA ? A2[(gk - K1) + sg.lda * gm] : A2[gm + sg.lda * (gk - K1)];
      }
      ra[t] = val;
    }
#pragma unroll
    for (int t = 0; t < LB; ++t) {
      const int e = tid + 256 * t;
      int nn, kk;
      if (!TB) { kk = e % BK; nn = e / BK; } else { nn = e % BN; kk = e / BN; }
      const int64_t gcol = n0 + nn, gk = k0 + kk;
      TI val = 0;
      if (gcol < N && gk < K) {
        if (!SEG || gk < K1) val = TB ? B[gcol + ldb * gk] : B[gk + ldb * gcol];
        else val = TB ? B2[gcol + sg.ldb * (gk - K1)] : B2[(gk - K1) + sg.ldb * gcol];
      }
      rb[t] = val;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int t = 0; t < LA; ++t) {
      const int e = tid + 256 * t;
      int mm, kk;
      if (!TA) { mm = e % BM; kk = e / BM; } else { kk = e % BK; mm = e / BK; }
      As[buf][kk][mm] = (T)ra[t];
    }
#pragma unroll
    for (int t = 0; t < LB; ++t) {
      const int e = tid + 256 * t;
      int nn, kk;
      if (!TB) { kk = e % BK; nn = e / BK; } else { nn = e % BN; kk = e / BN; }
      Bs[buf][nn][kk] = (T)rb[t];
    }
  };

  Acc acc[MB][NB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = Acc{0, 0, 0, 0};

  const int64_t nst = (K + BK - 1) / BK;
  gload(0);
  sstore(0);
  __syncthreads();
  for (int64_t s = 0; s < nst; ++s) {
    const int buf = s & 1;
    if (s + 1 < nst) gload((s + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      T af[MB], bf[NB];
#pragma unroll
      for (int i = 0; i < MB; ++i) af[i] = As[buf][kk * 4 + (lane >> 4)][wm * WM + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < NB; ++j) bf[j] = Bs[buf][wn * WN + j * 16 + (lane & 15)][kk * 4 + (lane >> 4)];
#pragma unroll
      for (int i = 0; i < MB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          // LM: B fragment as the first MFMA operand, so the accumulator lanes run along M
          if constexpr (F32) {
            if constexpr (LM) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j], af[i], acc[i][j], 0, 0, 0);
            else acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
          } else {
            if constexpr (LM) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
            else acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
          }
        }
    }
    if (s + 1 < nst) sstore(buf ^ 1);
    __syncthreads();
  }

  // epilogue: the accumulator of block (i, j) holds, in lane l and register r, the element at
  // (lane index = l & 15, register index = (l >> 4) + 4 r for FP64 MFMA or 4 (l >> 4) + r for FP32
  // MFMA) of the 16x16 block; the lane index runs along M when LM (rows of D), else along N.
  // Output address: column-major with ldd, or (om.rmod > 0) the two-mode map
  // row r -> (r % rmod)*rs0 + (r / rmod)*rs1, column c -> (c % cmod)*cs0 + (c / cmod)*cs1
  // (cmod = 0: c*cs0), which lets a GEMM accumulate into a permuted tensor.  LM is chosen so
  // that the lanes of one store run along the unit-stride axis of the output.
  // The map's div/mod is done once per row and per column the thread owns (rows depend on (i,
  // lane) and columns on (j, register) when LM, the other way round otherwise); with beta != 0
  // every C value of the thread is loaded (unconditionally, an element outside D reading C[0])
  // before the first store.
  const int li = lane & 15;
  auto rix = [&](int rr) { return F32 ? 4 * (lane >> 4) + rr : (lane >> 4) + 4 * rr; };
  auto map_row = [&](int64_t row) { return om.rmod > 0 ? (row % om.rmod) * om.rs0 + (row / om.rmod) * om.rs1 : row; };
  auto map_col = [&](int64_t col, int64_t ld) {
    return om.rmod > 0 ? (om.cmod > 0 ? (col % om.cmod) * om.cs0 + (col / om.cmod) * om.cs1 : col * om.cs0) : ld * col;
  };
  constexpr int RR = LM ? 1 : 4, CR = LM ? 4 : 1;
  int64_t roff[MB][RR], coff[NB][CR], ccoff[NB][CR];
  bool rok[MB][RR], cok[NB][CR];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int q = 0; q < RR; ++q) {
      const int64_t row = m0 + wm * WM + i * 16 + (LM ? li : rix(q));
      rok[i][q] = row < M;
      roff[i][q] = map_row(rok[i][q] ? row : 0);
    }
#pragma unroll
  for (int j = 0; j < NB; ++j)
#pragma unroll
    for (int q = 0; q < CR; ++q) {
      const int64_t col = n0 + wn * WN + j * 16 + (LM ? rix(q) : li);
      cok[j][q] = col < N;
      coff[j][q] = map_col(cok[j][q] ? col : 0, ldd);
      ccoff[j][q] = map_col(cok[j][q] ? col : 0, ldc);
    }
  auto at = [&](int i, int j, int rr, int64_t& off, int64_t& offc) {
    const int qr = LM ? 0 : rr, qc = LM ? rr : 0;
    off = roff[i][qr] + coff[j][qc];
    offc = om.rmod > 0 ? off : roff[i][qr] + ccoff[j][qc];
    return rok[i][qr] && cok[j][qc];
  };
  if (beta != 0.0) {
    TO cv[MB][NB][4];
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          int64_t off, offc;
          const bool in = at(i, j, rr, off, offc);
          cv[i][j][rr] = C[in ? offc : 0];
        }
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          int64_t off, offc;
          if (at(i, j, rr, off, offc)) D[off] = (TO)(alpha * (double)acc[i][j][rr] + beta * (double)cv[i][j][rr]);
        }
  } else {
#pragma unroll
    for (int i = 0; i < MB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          int64_t off, offc;
          if (at(i, j, rr, off, offc)) D[off] = (TO)(alpha * (double)acc[i][j][rr]);
        }
  }
}

// lanes along M unless only the output's column axis has unit stride
static bool lanes_along_m(const OutMap& om) {
  if (om.rmod <= 0) return true;
  if (om.rs0 == 1) return true;
  return om.cs0 != 1;
}

// L2 budget of the B blocks of one raster group (bytes)
static constexpr int64_t kGroupBytes = 1 << 20;

template <int BM, int BN, int BK, bool TA, bool TB, int TAG, bool F32, typename TI, typename TO>
static void launch(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const TI* A, int64_t lda, int64_t sA,
                   const TI* B, int64_t ldb, int64_t sB, double beta, const TO* C, int64_t ldc, int64_t sC, TO* D,
                   int64_t ldd, int64_t sD, int64_t batch, int64_t kchunk, const OutMap& om, const KSeg& sg) {
  const int64_t tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  FH_CHECK(tm * tn < (int64_t)1 << 31 && batch < 65536, -3, "gemm grid too large");
  dim3 grid((unsigned)(tm * tn), (unsigned)batch);
  // A larger than B: group N-tiles so that a group's B blocks stay in L2
  int gn = 0;
  if (M > N) {
    const int64_t kl = kchunk > 0 ? kchunk : K;
    const int64_t per = std::max<int64_t>(1, kl * BN * (int64_t)sizeof(TI));
    gn = (int)std::max<int64_t>(1, std::min<int64_t>(tn, kGroupBytes / per));
  }
  constexpr int OCC = BM * BN > 8192 ? 1 : 4;
#define FH_GO(LM_, SEG_)                                                                                           \
  hipLaunchKernelGGL((k_gemm<BM, BN, BK, TA, TB, OCC, TAG, F32, TI, TO, LM_, SEG_>), grid, dim3(256), 0, s, M, N, K, \
                     alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, D, ldd, sD, (int)tm, (int)tn, kchunk, gn, om, sg)
  if constexpr (TAG == 2) {
    // the (T) W build: output maps (lane orientation along N when only the columns are unit-stride)
    // and two-segment K
    const bool lm = lanes_along_m(om);
    if (sg.A) {
      if (lm) FH_GO(true, true); else FH_GO(false, true);
    } else {
      if (lm) FH_GO(true, false); else FH_GO(false, false);
    }
  } else {
    FH_CHECK(!sg.A, -3, "gemm: two-segment K is only instantiated for the (T) W build");
    FH_GO(true, false);
  }
#undef FH_GO
  FH_HIP(hipGetLastError());
}

// 64x64 tiles at up to four workgroups per CU measured fastest on every shape of this engine
// (the latency of the next operand tile is hidden by the other workgroups); BK = 8 when the
// per-tile K loop is short.
template <bool TA, bool TB, int TAG, bool F32, typename TI, typename TO>
static void pick(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const TI* A, int64_t lda, int64_t sA,
                 const TI* B, int64_t ldb, int64_t sB, double beta, const TO* C, int64_t ldc, int64_t sC, TO* D,
                 int64_t ldd, int64_t sD, int64_t batch, int64_t kchunk, const OutMap& om, const KSeg& sg) {
  const int64_t kl = kchunk > 0 ? kchunk : K;
  if (kl <= 512)
    launch<64, 64, 8, TA, TB, TAG, F32, TI, TO>(s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, D, ldd, sD,
                                                batch, kchunk, om, sg);
  else
    launch<64, 64, 16, TA, TB, TAG, F32, TI, TO>(s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, D, ldd,
                                                 sD, batch, kchunk, om, sg);
}

// split-K sizing: few output tiles and a long K leave most CUs idle; cut K into slices
static int64_t splitk_slices(int64_t M, int64_t N, int64_t K, int64_t batch) {
  if (batch != 1 || K < 4096) return 1;
  const int64_t tiles = ((M + 63) / 64) * ((N + 63) / 64);
  if (tiles >= 256) return 1;
  int64_t sl = (1024 + tiles - 1) / tiles;
  sl = std::min<int64_t>(sl, K / 512);
  return std::max<int64_t>(sl, 1);
}

int64_t gemm_splitk_elems(int64_t M, int64_t N, int64_t K, int64_t batch) {
  const int64_t sl = splitk_slices(M, N, K, batch);
  return sl > 1 ? sl * M * N : 0;
}

__global__ void k_splitk_reduce(int64_t M, int64_t N, int64_t S, const double* __restrict__ W, double alpha, double beta,
                                const double* C, int64_t ldc, double* D, int64_t ldd) {
  const int64_t n = M * N;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < n; idx += (int64_t)gridDim.x * 256) {
    const int64_t m = idx % M, c = idx / M;
    double acc = 0.0;
    for (int64_t q = 0; q < S; ++q) acc += W[idx + q * n];
    double val = alpha * acc;
    if (beta != 0.0) val += beta * C[m + ldc * c];
    D[m + ldd * c] = val;
  }
}

static void check_args(int64_t M, int64_t K, int64_t N, bool ta, bool tb, int64_t lda, int64_t ldb, int64_t ldd,
                       double beta, const void* C, const void* D, int64_t ldc, const OutMap* map) {
  FH_CHECK(K >= 0, -3, "gemm: negative K");
  FH_CHECK(lda >= (ta ? (K > 0 ? K : 1) : M) && ldb >= (tb ? N : (K > 0 ? K : 1)) && (map || ldd >= M), -3,
           "gemm: leading dimension too small");
  FH_CHECK(beta == 0.0 || (C != nullptr && (map || ldc >= M)), -3, "gemm: C missing");
  FH_CHECK(!map || (map->rmod > 0 && (beta == 0.0 || C == D)), -3, "gemm: mapped output needs rmod > 0 and C == D");
}

// the second K segment covers k in [K1, K): its operands must hold K - K1 values along k
void check_seg(int64_t M, int64_t N, int64_t K, bool ta, bool tb, const KSeg& sg) {
  if (!sg.A) return;
  FH_CHECK(sg.B != nullptr && sg.K1 > 0 && sg.K1 < K && sg.K1 % 4 == 0, -3,
           "gemm: second K segment must start at a positive multiple of 4 below K");
  const int64_t K2 = K - sg.K1;
  FH_CHECK(sg.lda >= (ta ? K2 : M) && sg.ldb >= (tb ? N : K2), -3, "gemm: second K segment leading dimension too small");
}

void gemm_f64(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K, double alpha, const double* A,
              int64_t lda, int64_t sA, const double* B, int64_t ldb, int64_t sB, double beta, const double* C,
              int64_t ldc, int64_t sC, double* D, int64_t ldd, int64_t sD, int64_t batch, int tag, double* work,
              int64_t work_elems, const OutMap* map, bool f32, const KSeg* seg) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  check_args(M, seg ? seg->K1 : K, N, ta, tb, lda, ldb, ldd, beta, C, D, ldc, map);
  const OutMap om = map ? *map : OutMap{};
  const KSeg sg = seg ? *seg : KSeg{};
  check_seg(M, N, K, ta, tb, sg);
  const int64_t sl = (map || sg.A) ? 1 : splitk_slices(M, N, K, batch);
  const bool split = sl > 1 && work != nullptr && work_elems >= sl * M * N;
  const int64_t kchunk = split ? (K + sl - 1) / sl : 0;
  const int64_t nb = split ? (K + kchunk - 1) / kchunk : batch;
  const double al = split ? 1.0 : alpha, be = split ? 0.0 : beta;
  const double* Cc = split ? nullptr : C;
  double* Dd = split ? work : D;
  const int64_t lddd = split ? M : ldd, sDD = split ? M * N : sD;
#define FH_PICK2(TA_, TB_, F_)                                                                                      \
  do {                                                                                                              \
    if (tag == 1)                                                                                                   \
      pick<TA_, TB_, 1, F_, double, double>(s, M, N, K, al, A, lda, sA, B, ldb, sB, be, Cc, ldc, sC, Dd, lddd, sDD, nb, \
                                            kchunk, om, sg);                                                        \
    else if (tag == 2)                                                                                              \
      pick<TA_, TB_, 2, F_, double, double>(s, M, N, K, al, A, lda, sA, B, ldb, sB, be, Cc, ldc, sC, Dd, lddd, sDD, nb, \
                                            kchunk, om, sg);                                                        \
    else                                                                                                            \
      pick<TA_, TB_, 0, F_, double, double>(s, M, N, K, al, A, lda, sA, B, ldb, sB, be, Cc, ldc, sC, Dd, lddd, sDD, nb, \
                                            kchunk, om, sg);                                                        \
  } while (0)
#define FH_PICK(TA_, TB_)              \
  do {                                 \
    if (f32) FH_PICK2(TA_, TB_, true); \
    else FH_PICK2(TA_, TB_, false);    \
  } while (0)
  if (ta) {
    if (tb) FH_PICK(true, true); else FH_PICK(true, false);
  } else {
    if (tb) FH_PICK(false, true); else FH_PICK(false, false);
  }
#undef FH_PICK
#undef FH_PICK2
  if (split) {
    int64_t g = (M * N + 255) / 256;
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_splitk_reduce, dim3((unsigned)g), dim3(256), 0, s, M, N, nb, work, alpha, beta, C, ldc, D, ldd);
    FH_HIP(hipGetLastError());
  }
}

// FP32-MFMA GEMMs with FP32 operands and/or outputs in HBM (NN, no split-K)
template <int TAG, typename TI, typename TO>
static void gemm_nn_f32(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const TI* A, int64_t lda,
                        int64_t sA, const TI* B, int64_t ldb, int64_t sB, double beta, const TO* C, int64_t ldc,
                        int64_t sC, TO* D, int64_t ldd, int64_t sD, int64_t batch, const OutMap* map,
                        const KSeg* seg = nullptr) {
  if (M <= 0 || N <= 0 || batch <= 0) return;
  check_args(M, seg ? seg->K1 : K, N, false, false, lda, ldb, ldd, beta, C, D, ldc, map);
  const OutMap om = map ? *map : OutMap{};
  const KSeg sg = seg ? *seg : KSeg{};
  check_seg(M, N, K, false, false, sg);
  pick<false, false, TAG, true, TI, TO>(s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, D, ldd, sD, batch,
                                        0, om, sg);
}

void gemm_f32_fd(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const float* A, int64_t lda,
                 const float* B, int64_t ldb, double beta, const double* C, int64_t ldc, double* D, int64_t ldd, int tag) {
  if (tag == 1) gemm_nn_f32<1>(s, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, D, ldd, 0, 1, nullptr);
  else gemm_nn_f32<0>(s, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, D, ldd, 0, 1, nullptr);
}

void gemm_f32_df(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const double* A, int64_t lda,
                 const double* B, int64_t ldb, double beta, const float* C, int64_t ldc, float* D, int64_t ldd) {
  gemm_nn_f32<0>(s, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, D, ldd, 0, 1, nullptr);
}

void gemm_f32_ff(hipStream_t s, int64_t M, int64_t N, int64_t K, double alpha, const float* A, int64_t lda, int64_t sA,
                 const float* B, int64_t ldb, int64_t sB, double beta, const float* C, int64_t ldc, int64_t sC, float* D,
                 int64_t ldd, int64_t sD, int64_t batch, int tag, const OutMap* map, const KSeg* seg) {
  if (tag == 2) gemm_nn_f32<2>(s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, D, ldd, sD, batch, map, seg);
  else gemm_nn_f32<0>(s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, D, ldd, sD, batch, map, seg);
}

// FP32 operands and output, any transposition (the T2-guess network's layers; tag 3)
void gemm_f32(hipStream_t s, bool ta, bool tb, int64_t M, int64_t N, int64_t K, double alpha, const float* A,
              int64_t lda, const float* B, int64_t ldb, double beta, float* C, int64_t ldc) {
  if (M <= 0 || N <= 0) return;
  check_args(M, K, N, ta, tb, lda, ldb, ldc, beta, C, C, ldc, nullptr);
  const OutMap om{};
#define FH_F32(TA_, TB_) \
  pick<TA_, TB_, 3, true, float, float>(s, M, N, K, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, C, ldc, 0, 1, 0, om, \
                                        KSeg{})
  if (ta) {
    if (tb) FH_F32(true, true); else FH_F32(true, false);
  } else {
    if (tb) FH_F32(false, true); else FH_F32(false, false);
  }
#undef FH_F32
}

}  // namespace fh
