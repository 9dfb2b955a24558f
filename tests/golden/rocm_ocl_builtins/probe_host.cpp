#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <random>
int main(int argc, char** argv) {
  const int n = 1 << 18;
  std::mt19937 g(7); std::uniform_real_distribution<float> u(-1.f, 1.f), e(-6.f, 6.f);
  std::vector<float> a(3*n), b(3*n), out(9*n);
  for (int i = 0; i < 3*n; ++i) { a[i] = u(g) * std::exp2(e(g)); b[i] = u(g) * std::exp2(e(g)); }
  for (int i = 0; i < n/4; ++i) for (int k = 0; k < 3; ++k) b[3*i+k] = a[3*i+k] * (1.f + 1e-3f * u(g));  // near-parallel
  hipModule_t m; hipFunction_t f;
  if (hipModuleLoad(&m, argv[1]) != hipSuccess) { printf("load failed\n"); return 1; }
  if (hipModuleGetFunction(&f, m, "probe") != hipSuccess) { printf("fn failed\n"); return 1; }
  float *da, *db, *dout;
  hipMalloc(&da, 12*n); hipMalloc(&db, 12*n); hipMalloc(&dout, 36*n);
  hipMemcpy(da, a.data(), 12*n, hipMemcpyHostToDevice); hipMemcpy(db, b.data(), 12*n, hipMemcpyHostToDevice);
  int nn = n; void* args[] = {&da, &db, &dout, &nn};
  if (hipModuleLaunchKernel(f, n/256, 1, 1, 256, 1, 1, 0, nullptr, args, nullptr) != hipSuccess) { printf("launch failed\n"); return 1; }
  hipDeviceSynchronize();
  hipMemcpy(out.data(), dout, 36*n, hipMemcpyDeviceToHost);
  FILE* fo = fopen(argv[2], "wb"); fwrite(a.data(), 4, 3*n, fo); fwrite(b.data(), 4, 3*n, fo); fwrite(out.data(), 4, 9*n, fo); fclose(fo);
  printf("probe ok n=%d\n", n);
  return 0;
}
