// FP64 MFMA GEMM variant sweep (tile, BK, occupancy) on the shapes of the DF-RCCSD(T) engine.
// Build: hipcc --offload-arch=gfx950 -O3 -std=c++17 -Xclang -target-feature -Xclang -packed-fp32-ops \
//        fermi.jl_amd/csrc/tools/gemm_sweep.hip -o gemm_sweep
// All buffers are sized M*K, K*N, M*N for the largest shape, so every launch stays in bounds.
#include "../fh_gemm.hip"
#include <cstdio>
#include <vector>
using namespace fh;
template <int BM, int BN, int BK, int OCC, bool F32, typename TI>
double run(int64_t M, int64_t N, int64_t K, const TI* A, const TI* B, double* D, int reps, int gn) {
  const int64_t tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  hipEvent_t a, b; hipEventCreate(&a); hipEventCreate(&b);
  auto go = [&]() { hipLaunchKernelGGL((k_gemm<BM, BN, BK, false, false, OCC, 9, F32, TI, double, true, false>), dim3(tm * tn, 1), dim3(256), 0, 0,
                     M, N, K, 1.0, A, M, 0, B, K, 0, 0.0, nullptr, M, 0, D, M, 0, (int)tm, (int)tn, (int64_t)0, gn, OutMap{}, KSeg{}); };
  go(); hipDeviceSynchronize();
  hipEventRecord(a);
  for (int r = 0; r < reps; ++r) go();
  hipEventRecord(b); hipEventSynchronize(b);
  float ms; hipEventElapsedTime(&ms, a, b);
  return 2.0 * M * N * K * reps / (ms * 1e-3) / 1e12;
}
int main() {
  struct S { const char* n; int64_t M, N, K; int reps; };
  std::vector<S> shapes = {{"ladder", 1600, 36100, 36100, 2}, {"ring", 7600, 7600, 7600, 3},
                           {"T_next", 36100, 3040, 230, 10}, {"T_mext", 577600, 190, 230, 10}};
  size_t maxA = 0, maxB = 0, maxD = 0;
  for (auto& s : shapes) { maxA = std::max(maxA, (size_t)(s.M * s.K)); maxB = std::max(maxB, (size_t)(s.K * s.N)); maxD = std::max(maxD, (size_t)(s.M * s.N)); }
  double *A, *B, *D; float *Af, *Bf;
  if (hipMalloc(&A, maxA * 8) || hipMalloc(&B, maxB * 8) || hipMalloc(&D, maxD * 8) || hipMalloc(&Af, maxA * 4) || hipMalloc(&Bf, maxB * 4)) { printf("alloc failed\n"); return 1; }
  hipMemset(A, 0, maxA * 8); hipMemset(B, 0, maxB * 8); hipMemset(Af, 0, maxA * 4); hipMemset(Bf, 0, maxB * 4);
  for (auto& s : shapes) {
    printf("%-8s M=%ld N=%ld K=%ld\n", s.n, (long)s.M, (long)s.N, (long)s.K); fflush(stdout);
    const int g8 = s.M > s.N ? 8 : 0;
    printf("  fp64 64x64x16 occ4           %6.2f TF/s\n", run<64, 64, 16, 4, false, double>(s.M, s.N, s.K, A, B, D, s.reps, 0)); fflush(stdout);
    printf("  fp64 64x64x8  occ4 group    %6.2f TF/s\n", run<64, 64, 8, 4, false, double>(s.M, s.N, s.K, A, B, D, s.reps, g8)); fflush(stdout);
    printf("  fp32(d in) 64x64x16 occ4    %6.2f TF/s\n", run<64, 64, 16, 4, true, double>(s.M, s.N, s.K, A, B, D, s.reps, 0)); fflush(stdout);
    printf("  fp32(f in) 64x64x16 occ4    %6.2f TF/s\n", run<64, 64, 16, 4, true, float>(s.M, s.N, s.K, Af, Bf, D, s.reps, 0)); fflush(stdout);
    printf("  fp32(f in) 64x64x8 group    %6.2f TF/s\n", run<64, 64, 8, 4, true, float>(s.M, s.N, s.K, Af, Bf, D, s.reps, g8)); fflush(stdout);
    printf("  fp32(f in) 128x64x16 occ2   %6.2f TF/s\n", run<128, 64, 16, 2, true, float>(s.M, s.N, s.K, Af, Bf, D, s.reps, 0)); fflush(stdout);
    printf("  fp32(f in) 128x128x16 occ1  %6.2f TF/s\n", run<128, 128, 16, 1, true, float>(s.M, s.N, s.K, Af, Bf, D, s.reps, 0)); fflush(stdout);
    printf("  fp32(f in) 128x128x32 occ1  %6.2f TF/s\n", run<128, 128, 32, 1, true, float>(s.M, s.N, s.K, Af, Bf, D, s.reps, 0)); fflush(stdout);
  }
  printf("status %s\n", hipGetErrorString(hipDeviceSynchronize()));
  return 0;
}
