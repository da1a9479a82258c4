// FP64 MFMA GEMM variant sweep (tile, BK, occupancy) on the shapes of the DF-RCCSD(T) engine.
// Build: hipcc --offload-arch=gfx950 -O3 -std=c++17 -Xclang -target-feature -Xclang -packed-fp32-ops \
//        fermi.jl_amd/csrc/tools/gemm_sweep.hip -o gemm_sweep
// All buffers are sized M*K, K*N, M*N for the largest shape, so every launch stays in bounds.
#include "../fh_gemm.hip"
#include <cstdio>
#include <vector>
using namespace fh;
template <int BM, int BN, int BK, int OCC>
double run(int64_t M, int64_t N, int64_t K, const double* A, const double* B, double* D, int reps) {
  const int64_t tm = (M + BM - 1) / BM, tn = (N + BN - 1) / BN;
  hipEvent_t a, b; hipEventCreate(&a); hipEventCreate(&b);
  auto go = [&]() { hipLaunchKernelGGL((k_gemm_f64<BM, BN, BK, false, false, OCC, 9>), dim3(tm * tn, 1), dim3(256), 0, 0,
                     M, N, K, 1.0, A, M, 0, B, K, 0, 0.0, nullptr, M, 0, D, M, 0, (int)tm, (int)tn, (int64_t)0, M > N ? 1 : 0); };
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
  double *A, *B, *D;
  if (hipMalloc(&A, maxA * 8) || hipMalloc(&B, maxB * 8) || hipMalloc(&D, maxD * 8)) { printf("alloc failed\n"); return 1; }
  hipMemset(A, 0, maxA * 8); hipMemset(B, 0, maxB * 8);
  for (auto& s : shapes) {
    printf("%-8s M=%ld N=%ld K=%ld\n", s.n, (long)s.M, (long)s.N, (long)s.K); fflush(stdout);
    printf("  64x64x16 occ2   %6.2f TF/s\n", run<64, 64, 16, 2>(s.M, s.N, s.K, A, B, D, s.reps)); fflush(stdout);
    printf("  64x64x16 occ3   %6.2f TF/s\n", run<64, 64, 16, 3>(s.M, s.N, s.K, A, B, D, s.reps)); fflush(stdout);
    printf("  64x64x16 occ4   %6.2f TF/s\n", run<64, 64, 16, 4>(s.M, s.N, s.K, A, B, D, s.reps)); fflush(stdout);
    printf("  64x64x32 occ2   %6.2f TF/s\n", run<64, 64, 32, 2>(s.M, s.N, s.K, A, B, D, s.reps)); fflush(stdout);
    printf("  64x64x32 occ3   %6.2f TF/s\n", run<64, 64, 32, 3>(s.M, s.N, s.K, A, B, D, s.reps)); fflush(stdout);
    printf("  64x64x8  occ4   %6.2f TF/s\n", run<64, 64, 8, 4>(s.M, s.N, s.K, A, B, D, s.reps)); fflush(stdout);
    printf("  128x64x16 occ3  %6.2f TF/s\n", run<128, 64, 16, 3>(s.M, s.N, s.K, A, B, D, s.reps)); fflush(stdout);
    printf("  64x128x16 occ2  %6.2f TF/s\n", run<64, 128, 16, 2>(s.M, s.N, s.K, A, B, D, s.reps)); fflush(stdout);
  }
  printf("status %s\n", hipGetErrorString(hipDeviceSynchronize()));
  return 0;
}
