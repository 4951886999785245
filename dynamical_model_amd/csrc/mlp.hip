// mlp.hip -- column-stencil MLP of the learned acoustic corrector on MFMA.
//
// The reference has no learned component; this network is new.  It maps a
// 3 x 3 stencil of columns (6 fields x kz levels each) to corrections of u,
// v, w and the Exner function (4 x kz) that move a reduced acoustic loop
// toward the full one.  Layers: nin -> nhid -> nhid -> nout, ReLU between,
// BF16 operands on the matrix cores with FP32 accumulation, FP32 master
// weights and Adam.
//
// GEMM: C[M x N] = A[M x K] . B[K x N] with arbitrary element strides for A
// and B (so transposed operands need no copy), on
// v_mfma_f32_32x32x16_bf16.  k_gemm_v takes the operands that run
// contiguously in aligned runs of 8 (every product of training and of the
// generic forward at the corrector's shapes): 64 x 64 tiles, 16-byte loads
// eight k tiles ahead, and, for the gradient products, a deterministic
// split of k when the tiles alone give few blocks.  k_gemm_bf16 (128 x 128
// tiles, element loads) takes any other strides.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "mlc_internal.h"
#include "moloch_lac.h"

namespace mlc {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }
__device__ __forceinline__ float from_bf16(__bf16 x) { return (float)x; }

constexpr int GBM = 128, GBN = 128, GBK = 32;
constexpr int GLDK = GBK + 8;  // padded LDS row (bf16 elements)

struct GemmArgs {
  const __bf16 *A;
  long sam, sak;  // element (m,k) at A[m*sam + k*sak]
  const __bf16 *B;
  long sbk, sbn;  // element (k,n) at B[k*sbk + n*sbn]
  int M, N, K;
  const float *bias;  // per n, or null
  int relu;           // apply max(0, .) after the bias
  float *Cf;          // fp32 output (ldc), or null
  __bf16 *Cb;         // bf16 output (ldc), or null
  long ldc;
  int accumulate;     // Cf += result instead of Cf = result
  float *ws;          // workspace for split-k partial sums (ws_cap floats), or null
  long ws_cap;
};

// Threads stage a 128 x 32 tile of A as LDS [m][k] and a 32 x 128 tile of B
// as LDS [n][k], so every MFMA lane reads 8 consecutive k with one 16-byte
// load.  Global reads follow whichever index of the operand is contiguous.
__global__ void __launch_bounds__(256) k_gemm_bf16(GemmArgs a) {
  __shared__ __bf16 sA[2][GBM][GLDK];
  __shared__ __bf16 sB[2][GBN][GLDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
  const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
  const bool a_kfast = a.sak == 1, b_kfast = a.sbk == 1;

  auto load_tiles = [&](int kt, int buf) {
    const int k0 = kt * GBK;
    // 4096 elements of each operand, 16 per thread
    for (int e = 0; e < 16; e++) {
      int idx = tid + 256 * e;
      int mm, kk;
      if (a_kfast) {
        kk = idx % GBK;
        mm = idx / GBK;
      } else {
        mm = idx % GBM;
        kk = idx / GBM;
      }
      int gm = m0 + mm, gk = k0 + kk;
      __bf16 v = (gm < a.M && gk < a.K) ? a.A[(long)gm * a.sam + (long)gk * a.sak] : to_bf16(0.f);
      sA[buf][mm][kk] = v;
      int nn;
      if (b_kfast) {
        kk = idx % GBK;
        nn = idx / GBK;
      } else {
        nn = idx % GBN;
        kk = idx / GBN;
      }
      int gn = n0 + nn;
      gk = k0 + kk;
      __bf16 w = (gn < a.N && gk < a.K) ? a.B[(long)gk * a.sbk + (long)gn * a.sbn] : to_bf16(0.f);
      sB[buf][nn][kk] = w;
    }
  };

  f32x16 acc[2][2];
  for (int x = 0; x < 2; x++)
    for (int y = 0; y < 2; y++)
      for (int r = 0; r < 16; r++) acc[x][y][r] = 0.f;

  const int nkt = (a.K + GBK - 1) / GBK;
  load_tiles(0, 0);
  __syncthreads();
  for (int kt = 0; kt < nkt; kt++) {
    const int buf = kt & 1;
    if (kt + 1 < nkt) load_tiles(kt + 1, buf ^ 1);
    // two MFMA K-steps of 16 per LDS tile
    for (int ks = 0; ks < GBK; ks += 16) {
      const int kq = ks + 8 * (lane >> 5);
      bf16x8 av[2], bv[2];
      for (int x = 0; x < 2; x++) av[x] = *reinterpret_cast<const bf16x8 *>(&sA[buf][wm + 32 * x + (lane & 31)][kq]);
      for (int y = 0; y < 2; y++) bv[y] = *reinterpret_cast<const bf16x8 *>(&sB[buf][wn + 32 * y + (lane & 31)][kq]);
      for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++) acc[x][y] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[x], bv[y], acc[x][y], 0, 0, 0);
    }
    __syncthreads();
  }

  // D[m][n]: n = lane % 32, m = 8*(r/4) + 4*(lane/32) + r%4
  for (int x = 0; x < 2; x++)
    for (int y = 0; y < 2; y++)
      for (int r = 0; r < 16; r++) {
        int m = m0 + wm + 32 * x + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        int n = n0 + wn + 32 * y + (lane & 31);
        if (m >= a.M || n >= a.N) continue;
        float v = acc[x][y][r];
        if (a.bias) v += a.bias[n];
        if (a.relu) v = v > 0.f ? v : 0.f;
        long o = (long)m * a.ldc + n;
        if (a.Cf) a.Cf[o] = a.accumulate ? a.Cf[o] + v : v;
        if (a.Cb) a.Cb[o] = to_bf16(v);
      }
}

// The training-size GEMMs (1024-4096 rows) on 64 x 64 tiles with 16-byte
// operand loads.  Each operand runs contiguously either along k or along its
// other index (m for A, n for B) in aligned runs of 8 elements, which the
// caller checks (gemm below): a thread loads one run of 8 of each operand
// per 32-deep k tile -- along k it lands as one 16-byte LDS store in the
// [row][k] layout of k_gemm_bf16, along m or n as eight 2-byte stores into
// eight rows -- and runs are either wholly inside the matrix or wholly
// outside (zeros).  The loads run VD k tiles ahead of the MFMAs in a ring of
// registers (the global latency of a few blocks per CU is long next to the
// two k-steps of a tile), the LDS tiles are double-buffered, one barrier
// per tile.  Four waves of 32 x 32 outputs; every output sees the MFMAs of
// its k-steps of 16 in increasing order, the fragments of k_gemm_bf16, so
// without a k split the results are those of k_gemm_bf16 bit for bit.
// With part != null the block's k tiles are slice blockIdx.z of a split of
// the k range: it stores its raw partial sums in part[slice][m][n], and
// k_gemm_reduce adds the slices in slice order (deterministic).
constexpr int VBM = 64, VBN = 64, VBK = 32, VLDK = VBK + 8, VD = 8;

template <bool AK, bool BK>
__global__ void __launch_bounds__(256) k_gemm_v(GemmArgs a, int kps, float *part) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[2][VBM][VLDK];
  __shared__ __attribute__((aligned(16))) __bf16 sB[2][VBN][VLDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const int m0 = blockIdx.y * VBM, n0 = blockIdx.x * VBN;
  const int nkt = (a.K + VBK - 1) / VBK;
  const int kt0 = blockIdx.z * kps;
  const int kt1 = kt0 + kps < nkt ? kt0 + kps : nkt;
  // this thread's run of each tile: along k, row tid/4 and k offset
  // 8 (tid%4) (64 rows x 4 runs); along m or n, k row tid/8 and offset
  // 8 (tid%8) (32 k x 8 runs)
  const int r_row = AK ? tid >> 2 : tid >> 3, r_off = AK ? (tid & 3) * 8 : (tid & 7) * 8;
  const int q_row = BK ? tid >> 2 : tid >> 3, q_off = BK ? (tid & 3) * 8 : (tid & 7) * 8;
  const bf16x8 zero8 = {};
  // Loads are unconditional, from addresses clamped into the matrix; a run
  // outside it is replaced by zeros only when the tile is stored to LDS,
  // VD tiles later (a select right after the load would wait for it).  A
  // tile past the slice reloads the slice's last tile (never read).
  // Branch-free loads let the compiler wait for each ring slot alone.
  auto clampt = [&](int kt) { return kt < kt1 ? kt : kt1 - 1; };
  auto okA = [&](int kt) {
    const int k0 = clampt(kt) * VBK;
    return AK ? (m0 + r_row < a.M && k0 + r_off < a.K) : (m0 + r_off < a.M && k0 + r_row < a.K);
  };
  auto okB = [&](int kt) {
    const int k0 = clampt(kt) * VBK;
    return BK ? (n0 + q_row < a.N && k0 + q_off < a.K) : (n0 + q_off < a.N && k0 + q_row < a.K);
  };
  auto loadA = [&](int kt) -> bf16x8 {
    const int k0 = clampt(kt) * VBK;
    if (AK) {
      const int m = m0 + r_row, k = k0 + r_off;
      return *reinterpret_cast<const bf16x8 *>(a.A + (long)(m < a.M ? m : a.M - 1) * a.sam + (k < a.K ? k : a.K - 8));
    }
    const int k = k0 + r_row, m = m0 + r_off;
    return *reinterpret_cast<const bf16x8 *>(a.A + (m < a.M ? m : a.M - 8) + (long)(k < a.K ? k : a.K - 1) * a.sak);
  };
  auto loadB = [&](int kt) -> bf16x8 {
    const int k0 = clampt(kt) * VBK;
    if (BK) {
      const int n = n0 + q_row, k = k0 + q_off;
      return *reinterpret_cast<const bf16x8 *>(a.B + (long)(n < a.N ? n : a.N - 1) * a.sbn + (k < a.K ? k : a.K - 8));
    }
    const int k = k0 + q_row, n = n0 + q_off;
    return *reinterpret_cast<const bf16x8 *>(a.B + (n < a.N ? n : a.N - 8) + (long)(k < a.K ? k : a.K - 1) * a.sbk);
  };
  // tile kt's runs (raw loads va, vb) into LDS buffer buf
  auto store = [&](bf16x8 va, bf16x8 vb, int buf, int kt) {
    va = okA(kt) ? va : zero8;
    vb = okB(kt) ? vb : zero8;
    if (AK) {
      *reinterpret_cast<bf16x8 *>(&sA[buf][r_row][r_off]) = va;
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) sA[buf][r_off + e][r_row] = va[e];
    }
    if (BK) {
      *reinterpret_cast<bf16x8 *>(&sB[buf][q_row][q_off]) = vb;
    } else {
#pragma unroll
      for (int e = 0; e < 8; e++) sB[buf][q_off + e][q_row] = vb[e];
    }
  };

  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; r++) acc[r] = 0.f;
  // ring: tile kt in slot (kt - kt0) % VD
  bf16x8 ra[VD], rb[VD];
#pragma unroll
  for (int u = 0; u < VD; u++) {
    ra[u] = loadA(kt0 + u);
    rb[u] = loadB(kt0 + u);
  }
  store(ra[0], rb[0], 0, kt0);
  ra[0] = loadA(kt0 + VD);
  rb[0] = loadB(kt0 + VD);
  __syncthreads();
  // tile kt (ring slot u): its two k-steps from LDS buffer (kt - kt0) & 1;
  // then tile kt+1 into the other buffer (its last reads, tile kt-1, ended
  // before the previous barrier; after the last tile a copy of it that
  // nobody reads) and the slot refilled with tile kt+1+VD
  auto step = [&](int kt, int u) {
    const int buf = (kt - kt0) & 1;
#pragma unroll
    for (int ks = 0; ks < VBK; ks += 16) {
      const int kq = ks + 8 * (lane >> 5);
      const bf16x8 av = *reinterpret_cast<const bf16x8 *>(&sA[buf][wm + (lane & 31)][kq]);
      const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(&sB[buf][wn + (lane & 31)][kq]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    store(ra[(u + 1) % VD], rb[(u + 1) % VD], buf ^ 1, kt + 1);
    ra[(u + 1) % VD] = loadA(kt + 1 + VD);
    rb[(u + 1) % VD] = loadB(kt + 1 + VD);
    __syncthreads();
  };
  // whole groups of VD tiles in a straight-line body (the ring slots stay
  // in their registers across the loop), then the fewer than VD left
  const int nfull = (kt1 - kt0) / VD * VD;
  int kt = kt0;
  for (; kt < kt0 + nfull; kt += VD) {
#pragma unroll
    for (int u = 0; u < VD; u++) step(kt + u, u);
  }
#pragma unroll
  for (int u = 0; u < VD - 1; u++)
    if (kt + u < kt1) step(kt + u, u);  // uniform over the block

  // D[m][n]: n = lane % 32, m = 8*(r/4) + 4*(lane/32) + r%4
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int m = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int n = n0 + wn + (lane & 31);
    if (m >= a.M || n >= a.N) continue;
    float v = acc[r];
    if (part) {
      part[((long)blockIdx.z * a.M + m) * a.N + n] = v;
      continue;
    }
    if (a.bias) v += a.bias[n];
    if (a.relu) v = v > 0.f ? v : 0.f;
    const long o = (long)m * a.ldc + n;
    if (a.Cf) a.Cf[o] = a.accumulate ? a.Cf[o] + v : v;
    if (a.Cb) a.Cb[o] = to_bf16(v);
  }
}

// Cf = sum of the S slices of part[S][M][N], added in slice order.
__global__ void __launch_bounds__(256) k_gemm_reduce(const float *part, int S, int M, int N, float *Cf, long ldc,
                                                     int accumulate) {
  const long t = blockIdx.x * 256L + threadIdx.x, mn = (long)M * N;
  if (t >= mn) return;
  float s = part[t];
  for (int q = 1; q < S; q++) s += part[q * mn + t];
  const long o = (t / N) * ldc + t % N;
  Cf[o] = accumulate ? Cf[o] + s : s;
}

static bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }

template <bool AK, bool BK>
static void launch_gemm_v(hipStream_t st, const GemmArgs &a, dim3 grid, int kps, float *part) {
  hipLaunchKernelGGL((k_gemm_v<AK, BK>), grid, dim3(256), 0, st, a, kps, part);
}

// C = A . B.  Operands that run contiguously along k, or along m (A) / n (B)
// in aligned runs of 8, go to k_gemm_v; others to k_gemm_bf16.  Only a
// plain fp32 product (no bias, ReLU or bf16 copy: the gradients) with a
// workspace is split along k, into enough slices to give about 256 blocks.
static int gemm(hipStream_t st, const GemmArgs &a) {
  if (a.M <= 0 || a.N <= 0 || a.K <= 0) return MLC_OK;
  const bool ak = a.sak == 1 && a.K % 8 == 0 && a.sam % 8 == 0;
  const bool am = !ak && a.sam == 1 && a.M % 8 == 0 && a.sak % 8 == 0;
  const bool bk = a.sbk == 1 && a.K % 8 == 0 && a.sbn % 8 == 0;
  const bool bn = !bk && a.sbn == 1 && a.N % 8 == 0 && a.sbk % 8 == 0;
  if ((ak || am) && (bk || bn) && aligned16(a.A) && aligned16(a.B)) {
    const int gm = (a.M + VBM - 1) / VBM, gn = (a.N + VBN - 1) / VBN, nkt = (a.K + VBK - 1) / VBK;
    int S = 1;
    if (a.ws && a.Cf && !a.Cb && !a.bias && !a.relu && (long)gm * gn < 256) {
      S = (int)((256 + (long)gm * gn - 1) / ((long)gm * gn));
      S = S < 16 ? S : 16;
      S = S < nkt / 4 ? S : (nkt / 4 > 1 ? nkt / 4 : 1);
      while (S > 1 && (long)S * a.M * a.N > a.ws_cap) S--;
    }
    const int kps = (nkt + S - 1) / S;
    S = (nkt + kps - 1) / kps;  // every slice non-empty
    float *part = S > 1 ? a.ws : nullptr;
    const dim3 grid((unsigned)gn, (unsigned)gm, (unsigned)S);
    if (ak && bk) launch_gemm_v<true, true>(st, a, grid, kps, part);
    else if (ak) launch_gemm_v<true, false>(st, a, grid, kps, part);
    else if (bk) launch_gemm_v<false, true>(st, a, grid, kps, part);
    else launch_gemm_v<false, false>(st, a, grid, kps, part);
    if (part) {
      const long mn = (long)a.M * a.N;
      hipLaunchKernelGGL(k_gemm_reduce, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, (const float *)part, S,
                         a.M, a.N, a.Cf, a.ldc, a.accumulate);
    }
  } else {
    dim3 grid((a.N + GBN - 1) / GBN, (a.M + GBM - 1) / GBM);
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, a);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mlc_fail(MLC_EHIP, "gemm launch: %s", hipGetErrorString(e));
  return MLC_OK;
}

// ------------------------------------------------------ elementwise kernels
__global__ void k_f32_to_bf16(const float *x, __bf16 *y, long n) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < n) y[t] = to_bf16(x[t]);
}

// loss = sum (y - t)^2 / (rows * cols); dY = 2 (y - t) / (rows * cols) as bf16;
// per-block partial sums of the squared error.
__global__ void __launch_bounds__(256) k_mse(const float *y, const float *t, __bf16 *dy, float *dyf, long n,
                                             float scale, double *partial) {
  __shared__ double sh[256];
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (idx < n) {
    float d = y[idx] - t[idx];
    e2 = (double)d * d;
    float g = 2.f * d * scale;
    dy[idx] = to_bf16(g);
    if (dyf) dyf[idx] = g;
  }
  sh[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// dH = dH_in * (H > 0), in place on a bf16 gradient, H the post-ReLU bf16
__global__ void k_relu_bwd(__bf16 *dh, const __bf16 *h, float *dhf, long n) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  float g = from_bf16(h[t]) > 0.f ? (dhf ? dhf[t] : from_bf16(dh[t])) : 0.f;
  dh[t] = to_bf16(g);
  if (dhf) dhf[t] = g;
}

// db[n] = sum over rows of dY[row][n] (fp32 dY)
__global__ void __launch_bounds__(256) k_colsum(const float *dy, float *db, int rows, int cols) {
  int n = blockIdx.x;
  double s = 0.0;
  for (int r = threadIdx.x; r < rows; r += 256) s += dy[(long)r * cols + n];
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[n] = (float)sh[0];
}

// Adam on fp32 master weights, refreshing the bf16 copy
__global__ void k_adam(float *w, __bf16 *wb, const float *g, float *m, float *v, long n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n) return;
  float gt = g[t];
  float mt = b1 * m[t] + (1.f - b1) * gt;
  float vt = b2 * v[t] + (1.f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  float wt = w[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
  w[t] = wt;
  if (wb) wb[t] = to_bf16(wt);
}

// Pack bf16 weights W [in][out] into MFMA B fragments (layout in
// mlc_internal.h): one thread per fragment lane.
__global__ void k_pack_frag(const __bf16 *W, int in, int out, int nt_total, long nfrag_lanes, __bf16 *P) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= nfrag_lanes) return;
  const int lane = (int)(t & 63);
  const long frag = t >> 6;
  const int nt = (int)(frag % nt_total), ks = (int)(frag / nt_total);
  const int n = nt * 32 + (lane & 31), k0 = ks * 16 + 8 * (lane >> 5);
  for (int e = 0; e < 8; e++) {
    const int k = k0 + e;
    P[t * 8 + e] = (k < in && n < out) ? W[(long)k * out + n] : to_bf16(0.f);
  }
}

}  // namespace mlc

using namespace mlc;

// The MLP object (struct mlc_mlp) is declared in mlc_internal.h.

static int mlp_alloc(mlc_mlp *p, void **ptr, size_t bytes) {
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) return mlc_fail(MLC_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  p->allocs.push_back(*ptr);
  e = hipMemsetAsync(*ptr, 0, bytes, p->stream);
  if (e != hipSuccess) return mlc_fail(MLC_EHIP, "hipMemsetAsync: %s", hipGetErrorString(e));
  return MLC_OK;
}

int mlp_reserve_rows(mlc_mlp *p, int rows);
static int mlp_reserve(mlc_mlp *p, int rows) { return mlp_reserve_rows(p, rows); }

int mlp_reserve_rows(mlc_mlp *p, int rows) {
  if (rows <= p->cap) return MLC_OK;
  for (void *q : {(void *)p->H1, (void *)p->H2, (void *)p->dH1, (void *)p->dH2, (void *)p->dYb, (void *)p->Y,
                  (void *)p->dYf, (void *)p->dH1f, (void *)p->dH2f, (void *)p->partial}) {
    if (!q) continue;
    for (size_t i = 0; i < p->allocs.size(); i++)
      if (p->allocs[i] == q) {
        (void)hipFree(q);
        p->allocs.erase(p->allocs.begin() + i);
        break;
      }
  }
  long r = rows;
  int rc;
  if ((rc = mlp_alloc(p, (void **)&p->H1, r * p->nhid * sizeof(__bf16)))) return rc;
  if ((rc = mlp_alloc(p, (void **)&p->H2, r * p->nhid * sizeof(__bf16)))) return rc;
  if ((rc = mlp_alloc(p, (void **)&p->dH1, r * p->nhid * sizeof(__bf16)))) return rc;
  if ((rc = mlp_alloc(p, (void **)&p->dH2, r * p->nhid * sizeof(__bf16)))) return rc;
  if ((rc = mlp_alloc(p, (void **)&p->dYb, r * p->nout * sizeof(__bf16)))) return rc;
  if ((rc = mlp_alloc(p, (void **)&p->Y, r * p->nout * sizeof(float)))) return rc;
  if ((rc = mlp_alloc(p, (void **)&p->dYf, r * p->nout * sizeof(float)))) return rc;
  if ((rc = mlp_alloc(p, (void **)&p->dH1f, r * p->nhid * sizeof(float)))) return rc;
  if ((rc = mlp_alloc(p, (void **)&p->dH2f, r * p->nhid * sizeof(float)))) return rc;
  if ((rc = mlp_alloc(p, (void **)&p->partial, ((r * p->nout + 255) / 256 + 1) * sizeof(double)))) return rc;
  p->cap = rows;
  return MLC_OK;
}

#define MLP_TRY(x)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) return mlc_fail(MLC_EHIP, "%s: %s", #x, hipGetErrorString(e_));    \
  } while (0)

// The three layers of rows X[rows x nin] (bf16, row stride ldx) through the
// hidden activations H1, H2 (bf16, rows x nhid) into Y (fp32, row stride ldy).
static int mlp_layers(mlc_mlp *p, const __bf16 *X, long ldx, int rows, __bf16 *H1, __bf16 *H2, float *Y, long ldy) {
  int rc;
  GemmArgs a{};
  a.A = X, a.sam = ldx, a.sak = 1, a.B = p->Wb[0], a.sbk = p->nhid, a.sbn = 1;
  a.M = rows, a.N = p->nhid, a.K = p->nin, a.bias = p->bvec[0], a.relu = 1, a.Cb = H1, a.ldc = p->nhid;
  if ((rc = gemm(p->stream, a))) return rc;
  a = GemmArgs{};
  a.A = H1, a.sam = p->nhid, a.sak = 1, a.B = p->Wb[1], a.sbk = p->nhid, a.sbn = 1;
  a.M = rows, a.N = p->nhid, a.K = p->nhid, a.bias = p->bvec[1], a.relu = 1, a.Cb = H2, a.ldc = p->nhid;
  if ((rc = gemm(p->stream, a))) return rc;
  a = GemmArgs{};
  a.A = H2, a.sam = p->nhid, a.sak = 1, a.B = p->Wb[2], a.sbk = p->nout, a.sbn = 1;
  a.M = rows, a.N = p->nout, a.K = p->nhid, a.bias = p->bvec[2], a.relu = 0, a.Cf = Y, a.ldc = ldy;
  return gemm(p->stream, a);
}

// Forward of rows X[rows x nin] (bf16, row stride ldx): H1, H2, Y (fp32)
int mlp_forward_dev(mlc_mlp *p, const __bf16 *X, long ldx, int rows, float *Yout, long ldy) {
  int rc = mlp_reserve(p, rows);
  if (rc) return rc;
  return mlp_layers(p, X, ldx, rows, p->H1, p->H2, Yout ? Yout : p->Y, Yout ? ldy : p->nout);
}

// Loss and gradients of one batch: X (bf16) and targets T (fp32, rows x nout).
int mlp_backward_dev(mlc_mlp *p, const __bf16 *X, long ldx, const float *T, int rows, double *loss) {
  int rc = mlp_forward_dev(p, X, ldx, rows, nullptr, 0);
  if (rc) return rc;
  // the gradient products may split k: partial sums in a workspace
  constexpr long kWsFloats = 2L << 20;
  if (!p->ws) {
    if ((rc = mlp_alloc(p, (void **)&p->ws, kWsFloats * sizeof(float)))) return rc;
    p->ws_cap = kWsFloats;
  }
  auto G = [&]() {
    GemmArgs g{};
    g.ws = p->ws;
    g.ws_cap = p->ws_cap;
    return g;
  };
  const long ny = (long)rows * p->nout, nh = (long)rows * p->nhid;
  const int nblk = (int)((ny + 255) / 256);
  hipLaunchKernelGGL(k_mse, dim3(nblk), dim3(256), 0, p->stream, p->Y, T, p->dYb, p->dYf, ny,
                     1.0f / (float)ny, p->partial);
  // layer 3: gW3 = H2^T dY, gb3 = colsum dY, dH2 = dY W3^T
  GemmArgs a = G();
  a.A = p->H2, a.sam = 1, a.sak = p->nhid, a.B = p->dYb, a.sbk = p->nout, a.sbn = 1;
  a.M = p->nhid, a.N = p->nout, a.K = rows, a.Cf = p->gW[2], a.ldc = p->nout;
  if ((rc = gemm(p->stream, a))) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(p->nout), dim3(256), 0, p->stream, p->dYf, p->gb[2], rows, p->nout);
  a = G();
  a.A = p->dYb, a.sam = p->nout, a.sak = 1, a.B = p->Wb[2], a.sbk = 1, a.sbn = p->nout;
  a.M = rows, a.N = p->nhid, a.K = p->nout, a.Cf = p->dH2f, a.ldc = p->nhid;
  if ((rc = gemm(p->stream, a))) return rc;
  hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, p->stream, p->dH2, p->H2,
                     p->dH2f, nh);
  // layer 2
  a = G();
  a.A = p->H1, a.sam = 1, a.sak = p->nhid, a.B = p->dH2, a.sbk = p->nhid, a.sbn = 1;
  a.M = p->nhid, a.N = p->nhid, a.K = rows, a.Cf = p->gW[1], a.ldc = p->nhid;
  if ((rc = gemm(p->stream, a))) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(p->nhid), dim3(256), 0, p->stream, p->dH2f, p->gb[1], rows, p->nhid);
  a = G();
  a.A = p->dH2, a.sam = p->nhid, a.sak = 1, a.B = p->Wb[1], a.sbk = 1, a.sbn = p->nhid;
  a.M = rows, a.N = p->nhid, a.K = p->nhid, a.Cf = p->dH1f, a.ldc = p->nhid;
  if ((rc = gemm(p->stream, a))) return rc;
  hipLaunchKernelGGL(k_relu_bwd, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, p->stream, p->dH1, p->H1,
                     p->dH1f, nh);
  // layer 1
  a = G();
  a.A = X, a.sam = 1, a.sak = ldx, a.B = p->dH1, a.sbk = p->nhid, a.sbn = 1;
  a.M = p->nin, a.N = p->nhid, a.K = rows, a.Cf = p->gW[0], a.ldc = p->nhid;
  if ((rc = gemm(p->stream, a))) return rc;
  hipLaunchKernelGGL(k_colsum, dim3(p->nhid), dim3(256), 0, p->stream, p->dH1f, p->gb[0], rows, p->nhid);
  if (loss) {
    std::vector<double> part(nblk);
    MLP_TRY(hipMemcpyAsync(part.data(), p->partial, nblk * sizeof(double), hipMemcpyDeviceToHost, p->stream));
    MLP_TRY(hipStreamSynchronize(p->stream));
    double s = 0.0;
    for (double x : part) s += x;
    *loss = s / (double)ny;
  }
  return MLC_OK;
}

int mlp_pack_dev(mlc_mlp *p) {
  if (!p->pack_dirty) return MLC_OK;
  for (int l = 0; l < 3; l++) {
    const int in = p->in_dim(l), out = p->out_dim(l);
    const int nt_total = ((out + 255) / 256) * 8, ks_total = (in + 15) / 16;
    const long lanes = (long)nt_total * ks_total * 64;
    if (!p->Wp[l]) {
      int rc = mlp_alloc(p, (void **)&p->Wp[l], lanes * 8 * sizeof(__bf16));
      if (rc) return rc;
      p->wp_nt[l] = nt_total;
      p->wp_ks[l] = ks_total;
    }
    hipLaunchKernelGGL(k_pack_frag, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, p->stream, p->Wb[l], in, out,
                       nt_total, lanes, p->Wp[l]);
  }
  MLP_TRY(hipGetLastError());
  p->pack_dirty = false;
  return MLC_OK;
}

int mlp_adam_dev(mlc_mlp *p, double lr) {
  p->step++;
  p->pack_dirty = true;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.f - powf(b1, (float)p->step), c2 = 1.f - powf(b2, (float)p->step);
  for (int l = 0; l < 3; l++) {
    long nw = (long)p->in_dim(l) * p->out_dim(l), nb = p->out_dim(l);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, p->stream, p->W[l], p->Wb[l],
                       p->gW[l], p->mW[l], p->vW[l], nw, (float)lr, b1, b2, eps, c1, c2);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, p->stream, p->bvec[l],
                       (__bf16 *)nullptr, p->gb[l], p->mb[l], p->vb[l], nb, (float)lr, b1, b2, eps, c1, c2);
  }
  MLP_TRY(hipGetLastError());
  return MLC_OK;
}

static int mlp_sync_bf16(mlc_mlp *p) {
  p->pack_dirty = true;
  for (int l = 0; l < 3; l++) {
    long nw = (long)p->in_dim(l) * p->out_dim(l);
    hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, p->stream, p->W[l],
                       p->Wb[l], nw);
  }
  MLP_TRY(hipGetLastError());
  return MLC_OK;
}

// ---------------------------------------- scratch of the allocation-free forward
// mlc_mlp_forward_noalloc stages the bf16 inputs and the two hidden
// activations of one call here instead of in memory it allocates.  Each
// device the library runs on holds its own copy for the life of the process;
// host code reaches it only through the address hipGetSymbolAddress gives
// on that device (the array's name is a host-side stand-in, not a device
// address).
constexpr size_t kFwdScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kFwdScratchBytes];
static void *g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// Look up this device's copy of the scratch once (device current).
static int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return mlc_fail(MLC_EINVAL, "device %d out of range", device);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return MLC_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return mlc_fail(MLC_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
  g_scratch[device] = ptr;
  return MLC_OK;
}

int mlp_new(int nin, int nhid, int nout, unsigned long long seed, int device, hipStream_t stream, mlc_mlp **out) {
  if (!out || nin <= 0 || nhid <= 0 || nout <= 0) return mlc_fail(MLC_EINVAL, "mlc_mlp_create: bad arguments");
  *out = nullptr;
  MLP_TRY(hipSetDevice(device));
  int src = scratch_setup(device);
  if (src) return src;
  mlc_mlp *p = new mlc_mlp();
  p->nin = nin, p->nhid = nhid, p->nout = nout;
  p->device = device;
  if (stream) {
    p->stream = stream;
    p->own_stream = false;
  } else if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    delete p;
    return mlc_fail(MLC_EHIP, "mlc_mlp_create: stream");
  }
  int rc = MLC_OK;
  // He-uniform initial weights from splitmix64, zero biases
  uint64_t s = seed;
  auto rnd = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < 3 && !rc; l++) {
    long nw = (long)p->in_dim(l) * p->out_dim(l), nb = p->out_dim(l);
    if ((rc = mlp_alloc(p, (void **)&p->W[l], nw * 4)) || (rc = mlp_alloc(p, (void **)&p->Wb[l], nw * 2)) ||
        (rc = mlp_alloc(p, (void **)&p->gW[l], nw * 4)) || (rc = mlp_alloc(p, (void **)&p->mW[l], nw * 4)) ||
        (rc = mlp_alloc(p, (void **)&p->vW[l], nw * 4)) || (rc = mlp_alloc(p, (void **)&p->bvec[l], nb * 4)) ||
        (rc = mlp_alloc(p, (void **)&p->gb[l], nb * 4)) || (rc = mlp_alloc(p, (void **)&p->mb[l], nb * 4)) ||
        (rc = mlp_alloc(p, (void **)&p->vb[l], nb * 4)))
      break;
    std::vector<float> w(nw);
    double lim = std::sqrt(6.0 / (double)p->in_dim(l)) * (l == 2 ? 0.1 : 1.0);
    for (long i = 0; i < nw; i++) w[i] = (float)((2.0 * rnd() - 1.0) * lim);
    if (hipMemcpyAsync(p->W[l], w.data(), nw * 4, hipMemcpyHostToDevice, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
      rc = mlc_fail(MLC_EHIP, "mlc_mlp_create: upload");
  }
  if (!rc) rc = mlp_sync_bf16(p);
  if (!rc && hipStreamSynchronize(p->stream) != hipSuccess) rc = mlc_fail(MLC_EHIP, "mlc_mlp_create: sync");
  if (rc) {
    mlp_free(p);
    return rc;
  }
  *out = p;
  return MLC_OK;
}

void mlp_free(mlc_mlp *p) {
  if (!p) return;
  (void)hipStreamSynchronize(p->stream);
  for (void *q : p->allocs) (void)hipFree(q);
  if (p->own_stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

// ================================================================== C ABI
extern "C" int mlc_mlp_create(int nin, int nhid, int nout, unsigned long long seed, int device, mlc_mlp **out) {
  return mlp_new(nin, nhid, nout, seed, device, nullptr, out);
}

extern "C" int mlc_mlp_destroy(mlc_mlp *p) {
  mlp_free(p);
  return MLC_OK;
}

// Copy fp32 weights and biases of layer l (0..2) in or out: w is [in][out].
extern "C" int mlc_mlp_weights(mlc_mlp *p, int layer, float *w, float *b, int set) {
  if (!p || layer < 0 || layer > 2) return mlc_fail(MLC_EINVAL, "mlc_mlp_weights: bad arguments");
  long nw = (long)p->in_dim(layer) * p->out_dim(layer), nb = p->out_dim(layer);
  hipMemcpyKind kd = set ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  if (w) MLP_TRY(hipMemcpyAsync(set ? (void *)p->W[layer] : (void *)w, set ? (void *)w : (void *)p->W[layer], nw * 4,
                                kd, p->stream));
  if (b) MLP_TRY(hipMemcpyAsync(set ? (void *)p->bvec[layer] : (void *)b, set ? (void *)b : (void *)p->bvec[layer],
                                nb * 4, kd, p->stream));
  if (set) {
    int rc = mlp_sync_bf16(p);
    if (rc) return rc;
  }
  MLP_TRY(hipStreamSynchronize(p->stream));
  return MLC_OK;
}

// Gradients of layer l from the last mlc_mlp_grad call.
extern "C" int mlc_mlp_gradients(mlc_mlp *p, int layer, float *gw, float *gb) {
  if (!p || layer < 0 || layer > 2) return mlc_fail(MLC_EINVAL, "mlc_mlp_gradients: bad arguments");
  long nw = (long)p->in_dim(layer) * p->out_dim(layer), nb = p->out_dim(layer);
  if (gw) MLP_TRY(hipMemcpyAsync(gw, p->gW[layer], nw * 4, hipMemcpyDeviceToHost, p->stream));
  if (gb) MLP_TRY(hipMemcpyAsync(gb, p->gb[layer], nb * 4, hipMemcpyDeviceToHost, p->stream));
  MLP_TRY(hipStreamSynchronize(p->stream));
  return MLC_OK;
}

// Host-array convenience entry points (tests): X [rows x nin] fp32 (rounded
// to bf16 on upload), Y [rows x nout] fp32, T [rows x nout] fp32.
static int upload_bf16(mlc_mlp *p, const float *x, long n, __bf16 **dev) {
  float *tmp;
  int rc = mlp_alloc(p, (void **)&tmp, n * 4);
  if (rc) return rc;
  rc = mlp_alloc(p, (void **)dev, n * 2);
  if (rc) return rc;
  MLP_TRY(hipMemcpyAsync(tmp, x, n * 4, hipMemcpyHostToDevice, p->stream));
  hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, tmp, *dev, n);
  MLP_TRY(hipStreamSynchronize(p->stream));
  for (size_t i = 0; i < p->allocs.size(); i++)
    if (p->allocs[i] == tmp) {
      (void)hipFree(tmp);
      p->allocs.erase(p->allocs.begin() + i);
      break;
    }
  return MLC_OK;
}
static void release(mlc_mlp *p, void *q) {
  for (size_t i = 0; i < p->allocs.size(); i++)
    if (p->allocs[i] == q) {
      (void)hipFree(q);
      p->allocs.erase(p->allocs.begin() + i);
      return;
    }
}

extern "C" int mlc_mlp_forward(mlc_mlp *p, const float *X, int rows, float *Y) {
  if (!p || !X || !Y || rows <= 0) return mlc_fail(MLC_EINVAL, "mlc_mlp_forward: bad arguments");
  __bf16 *xd = nullptr;
  int rc = upload_bf16(p, X, (long)rows * p->nin, &xd);
  if (rc) return rc;
  rc = mlp_forward_dev(p, xd, p->nin, rows, nullptr, 0);
  if (!rc) {
    if (hipMemcpyAsync(Y, p->Y, (long)rows * p->nout * 4, hipMemcpyDeviceToHost, p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess)
      rc = mlc_fail(MLC_EHIP, "mlc_mlp_forward: download");
  }
  release(p, xd);
  return rc;
}

extern "C" int mlc_mlp_grad(mlc_mlp *p, const float *X, const float *T, int rows, double *loss) {
  if (!p || !X || !T || rows <= 0) return mlc_fail(MLC_EINVAL, "mlc_mlp_grad: bad arguments");
  __bf16 *xd = nullptr;
  float *td = nullptr;
  int rc = upload_bf16(p, X, (long)rows * p->nin, &xd);
  if (rc) return rc;
  if ((rc = mlp_alloc(p, (void **)&td, (long)rows * p->nout * 4))) return rc;
  MLP_TRY(hipMemcpyAsync(td, T, (long)rows * p->nout * 4, hipMemcpyHostToDevice, p->stream));
  rc = mlp_backward_dev(p, xd, p->nin, td, rows, loss);
  if (!rc && hipStreamSynchronize(p->stream) != hipSuccess) rc = mlc_fail(MLC_EHIP, "mlc_mlp_grad: sync");
  release(p, xd);
  release(p, td);
  return rc;
}

extern "C" int mlc_mlp_adam(mlc_mlp *p, double lr) {
  if (!p) return mlc_fail(MLC_EINVAL, "mlc_mlp_adam: NULL");
  int rc = mlp_adam_dev(p, lr);
  if (rc) return rc;
  MLP_TRY(hipStreamSynchronize(p->stream));
  return MLC_OK;
}

// C = A . B for host fp32 matrices rounded to bf16 (tests of the GEMM):
// A [M x K] with strides (sam, sak), B [K x N] with (sbk, sbn), C [M x N].
extern "C" int mlc_gemm_bf16_host(const float *A, long sam, long sak, long asize, const float *B, long sbk, long sbn,
                                  long bsize, int M, int N, int K, float *C) {
  mlc_mlp tmp;
  tmp.nin = tmp.nhid = tmp.nout = 1;
  MLP_TRY(hipStreamCreateWithFlags(&tmp.stream, hipStreamNonBlocking));
  __bf16 *ad = nullptr, *bd = nullptr;
  float *cd = nullptr;
  int rc = upload_bf16(&tmp, A, asize, &ad);
  if (!rc) rc = upload_bf16(&tmp, B, bsize, &bd);
  if (!rc) rc = mlp_alloc(&tmp, (void **)&cd, (long)M * N * 4);
  if (!rc) {
    GemmArgs a{};
    a.A = ad, a.sam = sam, a.sak = sak, a.B = bd, a.sbk = sbk, a.sbn = sbn, a.M = M, a.N = N, a.K = K;
    a.Cf = cd, a.ldc = N;
    rc = gemm(tmp.stream, a);
  }
  if (!rc && (hipMemcpyAsync(C, cd, (long)M * N * 4, hipMemcpyDeviceToHost, tmp.stream) != hipSuccess ||
              hipStreamSynchronize(tmp.stream) != hipSuccess))
    rc = mlc_fail(MLC_EHIP, "mlc_gemm_bf16_host: download");
  for (void *q : tmp.allocs) (void)hipFree(q);
  (void)hipStreamDestroy(tmp.stream);
  return rc;
}

// Forward of device arrays without allocating: X [rows x nin] fp32 and
// Y [rows x nout] fp32 in device memory; the bf16 inputs and the hidden
// activations go to the device's scratch.  The same kernels, in the same
// order, as mlc_mlp_forward, so the results are the same bits.
extern "C" int mlc_mlp_forward_noalloc(mlc_mlp *p, const float *X, int rows, float *Y) {
  if (!p || !X || !Y || rows <= 0) return mlc_fail(MLC_EINVAL, "mlc_mlp_forward_noalloc: bad arguments");
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t xb = up((size_t)rows * p->nin * sizeof(__bf16));
  const size_t hb = up((size_t)rows * p->nhid * sizeof(__bf16));
  const size_t need = xb + 2 * hb;
  if (need > kFwdScratchBytes) {
    const size_t per_row = (size_t)p->nin * 2 + 2 * (size_t)p->nhid * 2;
    return mlc_fail(MLC_EINVAL,
                    "mlc_mlp_forward_noalloc: %d rows need %zu bytes of scratch, more than its %zu (at most about "
                    "%zu rows of this network)",
                    rows, need, kFwdScratchBytes, kFwdScratchBytes / per_row);
  }
  void *base = (p->device >= 0 && p->device < kMaxDevices) ? g_scratch[p->device] : nullptr;
  if (!base) return mlc_fail(MLC_ESTATE, "mlc_mlp_forward_noalloc: no scratch on device %d", p->device);
  unsigned char *sc = static_cast<unsigned char *>(base);
  __bf16 *Xb = reinterpret_cast<__bf16 *>(sc);
  __bf16 *H1 = reinterpret_cast<__bf16 *>(sc + xb);
  __bf16 *H2 = reinterpret_cast<__bf16 *>(sc + xb + hb);
  MLP_TRY(hipSetDevice(p->device));
  const long n = (long)rows * p->nin;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, X, Xb, n);
  int rc = mlp_layers(p, Xb, p->nin, rows, H1, H2, Y, p->nout);
  if (rc) return rc;
  MLP_TRY(hipStreamSynchronize(p->stream));
  return MLC_OK;
}
