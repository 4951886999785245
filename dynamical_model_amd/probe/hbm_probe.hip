// What streaming rate one MI355X reaches for the access shapes of the MOLOCH
// kernels, on arrays of the 2048 x 2048 x 60 FP32 field layout (row pitch
// 2112, two-point halos, j = 1 at element 32): the practical ceilings next to
// the 8 TB/s HBM peak.
//   read        a float4 grid-stride sum over one field (1.07 GB)
//   copy        float4 and float copies of one field, hipMemcpyAsync
//   7r4w column 7 fields read, 4 written per point, one column per lane
//               walking k in (64, 4) blocks (the column kernels' shape)
//   7r4w point  the same with k in the grid (one point per thread)
//   stagger S   7r4w column and point on the eleven arrays carved out of one
//               allocation, array q starting q (F + S) bytes in (F the
//               array's size rounded up to 2 MiB): the rate's dependence on
//               the arrays' relative placement
// Every access is within the allocations by construction.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                              \
  do {                                                                     \
    hipError_t e = (x);                                                    \
    if (e != hipSuccess) {                                                 \
      std::printf("HIP error %s at line %d\n", hipGetErrorString(e), __LINE__); \
      std::exit(1);                                                        \
    }                                                                      \
  } while (0)

constexpr int NJ = 2048, NI = 2048, NK = 60;
constexpr long PJ = 2112;
constexpr long PLANE = PJ * (NI + 4);
constexpr long NEL = PLANE * NK;

__global__ void k_read4(const float4 *__restrict__ a, long n4, float *out) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
  float s = 0;
  for (long x = t; x < n4; x += stride) {
    float4 v = a[x];
    s += v.x + v.y + v.z + v.w;
  }
  if (s == 1.2345f) out[0] = s;  // keeps the loads
}
__global__ void k_copy4(const float4 *__restrict__ a, float4 *__restrict__ b, long n4) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
  for (long x = t; x < n4; x += stride) b[x] = a[x];
}
__global__ void k_copy1(const float *__restrict__ a, float *__restrict__ b, long n) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x, stride = (long)gridDim.x * blockDim.x;
  for (long x = t; x < n; x += stride) b[x] = a[x];
}
struct F7 {
  const float *r[7];
  float *w[4];
};
__device__ __forceinline__ long at(int j, int i, int k) { return (long)(k - 1) * PLANE + (long)(i + 1) * PJ + (j + 31); }
__global__ void __launch_bounds__(256) k_col(F7 f) {
  int j = blockIdx.x * 64 + threadIdx.x + 1, i = blockIdx.y * 4 + threadIdx.y + 1;
  for (int k = 1; k <= NK; k++) {
    long x = at(j, i, k);
    float a0 = f.r[0][x], a1 = f.r[1][x], a2 = f.r[2][x], a3 = f.r[3][x], a4 = f.r[4][x], a5 = f.r[5][x],
          a6 = f.r[6][x];
    f.w[0][x] = a0 + a1;
    f.w[1][x] = a2 * a3;
    f.w[2][x] = a4 - a5;
    f.w[3][x] = a6 + a0;
  }
}
__global__ void __launch_bounds__(256) k_pt(F7 f) {
  int j = blockIdx.x * 64 + threadIdx.x + 1, i = blockIdx.y * 4 + threadIdx.y + 1, k = blockIdx.z + 1;
  long x = at(j, i, k);
  float a0 = f.r[0][x], a1 = f.r[1][x], a2 = f.r[2][x], a3 = f.r[3][x], a4 = f.r[4][x], a5 = f.r[5][x],
        a6 = f.r[6][x];
  f.w[0][x] = a0 + a1;
  f.w[1][x] = a2 * a3;
  f.w[2][x] = a4 - a5;
  f.w[3][x] = a6 + a0;
}

int main() {
  float *buf[11];
  for (int q = 0; q < 11; q++) {
    CK(hipMalloc(&buf[q], NEL * sizeof(float)));
    CK(hipMemset(buf[q], 0, NEL * sizeof(float)));
  }
  float *out;
  CK(hipMalloc(&out, 64));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  auto timeit = [&](const char *name, double bytes, auto launch) {
    launch();
    CK(hipDeviceSynchronize());
    const int R = 10;
    CK(hipEventRecord(e0));
    for (int r = 0; r < R; r++) launch();
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    float ms;
    CK(hipEventElapsedTime(&ms, e0, e1));
    ms /= R;
    std::printf("%-22s %8.3f ms %8.1f GB/s %6.3f of 8 TB/s\n", name, ms, bytes / ms * 1e-6, bytes / ms * 1e-6 / 8000.0);
    std::fflush(stdout);
  };
  const long n = NEL, n4 = NEL / 4;
  const double B = n * 4.0;
  timeit("read float4", B, [&] { k_read4<<<2048, 256>>>((const float4 *)buf[0], n4, out); });
  timeit("copy float4", 2 * B, [&] { k_copy4<<<32768, 256>>>((const float4 *)buf[0], (float4 *)buf[1], n4); });
  timeit("copy float", 2 * B, [&] { k_copy1<<<2048, 256>>>(buf[0], buf[1], n); });
  timeit("hipMemcpyAsync", 2 * B, [&] { CK(hipMemcpyAsync(buf[1], buf[0], n * 4, hipMemcpyDeviceToDevice)); });
  F7 f;
  for (int q = 0; q < 7; q++) f.r[q] = buf[q];
  for (int q = 0; q < 4; q++) f.w[q] = buf[7 + q];
  const double Bc = 11.0 * NJ * NI * NK * 4.0;
  timeit("7r4w column (64,4)", Bc, [&] { k_col<<<dim3(NJ / 64, NI / 4), dim3(64, 4)>>>(f); });
  timeit("7r4w point", Bc, [&] { k_pt<<<dim3(NJ / 64, NI / 4, NK), dim3(64, 4)>>>(f); });
  for (int q = 0; q < 11; q++) CK(hipFree(buf[q]));
  // one slab, array q at q (F + S) bytes
  const size_t F = (NEL * sizeof(float) + (2u << 20) - 1) / (2u << 20) * (2u << 20);
  const size_t smax = 262144;
  char *slab;
  CK(hipMalloc(&slab, 11 * (F + smax)));
  CK(hipMemset(slab, 0, 11 * (F + smax)));
  for (size_t S : {(size_t)0, (size_t)256, (size_t)512, (size_t)1024, (size_t)2048, (size_t)4096, (size_t)8192,
                   (size_t)16384, (size_t)65536, smax}) {
    for (int q = 0; q < 7; q++) f.r[q] = (const float *)(slab + q * (F + S));
    for (int q = 0; q < 4; q++) f.w[q] = (float *)(slab + (7 + q) * (F + S));
    char nm[64];
    std::snprintf(nm, sizeof nm, "stagger %zu column", S);
    timeit(nm, Bc, [&] { k_col<<<dim3(NJ / 64, NI / 4), dim3(64, 4)>>>(f); });
    std::snprintf(nm, sizeof nm, "stagger %zu point", S);
    timeit(nm, Bc, [&] { k_pt<<<dim3(NJ / 64, NI / 4, NK), dim3(64, 4)>>>(f); });
  }
  CK(hipFree(slab));
  CK(hipFree(out));
  std::printf("done\n");
  return 0;
}
