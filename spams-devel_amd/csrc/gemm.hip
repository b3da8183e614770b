// MFMA GEMMs of the engine: C = alpha * A' * B + beta * E.
//
// A is K x M and B is K x N, both column-major with K contiguous, so every
// product the decomposition path needs is of this one form:
//   D'X        (A = D,  m x p;  B = X,  m x n)      ProdMatrix DtX
//   G = D'D    (A = B = D)                           Matrix::XtX (ssyrk)
//   D*Y - X    (A = D', p x m; B = Y, p x n; E = X)  SqLoss::grad residual
//   D'R        (A = D;  B = R)                       SqLoss::grad
// (linalg/linalg.h:1751-1756, :1967-1972, prox/fista.h:390-402).
//
// Kernels:
//   gemm_tn<float>   v_mfma_f32_16x16x4_f32   (FP32 in, FP32 accumulate)
//   gemm_tn<double>  v_mfma_f64_16x16x4_f64   (FP64 in, FP64 accumulate)
//   gemm_tn_bf16x3   v_mfma_f32_16x16x32_bf16 x3 on a hi/lo BF16 split of
//                    both operands (a*b ~ ah*bh + ah*bl + al*bh), FP32
//                    accumulate.
// Block tile 128 x 128, four wavefronts in 2 x 2, each owning a 64 x 64
// sub-tile (4 x 4 MFMA 16x16 blocks); K staged through LDS.
#include "spx_common.h"
#include "spx_internal.h"

namespace spx {

namespace {

constexpr int BM = 128, BN = 128;

// Output tile of this block.  Blocks are dispatched round-robin over the 8
// XCDs (block b on XCD b % 8), each with its own L2; the remap gives each XCD
// a contiguous run of tiles, row tiles fastest, so the row tiles that share
// one column tile of B (a panel of signals) read it into the same L2.
__device__ __forceinline__ void tile_origin(int& bm, int& bn) {
  constexpr int kXcd = 8;
  const int mt = gridDim.x;
  const int total = mt * gridDim.y;
  int b = blockIdx.x + blockIdx.y * mt;
  const int full = total - total % kXcd;
  if (b < full) b = (b % kXcd) * (full / kXcd) + b / kXcd;
  bm = (b % mt) * BM;
  bn = (b / mt) * BN;
}

// Split over the inner dimension (grid z > 1): slice z takes the inner
// indices [z kz, z kz + kz) and writes its own partial product, zstride
// elements after the previous slice's.
// Returns the slice's first inner index and shortens K to the slice.
__device__ __forceinline__ int split_k_begin(int& K, int kz) {
  if (gridDim.z <= 1) return 0;
  const int kb = (int)blockIdx.z * kz;
  const int rest = K - kb;
  K = rest < kz ? (rest > 0 ? rest : 0) : kz;
  return kb;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <typename T>
struct Acc;
template <>
struct Acc<float> {
  typedef f32x4 type;
  static __device__ __forceinline__ f32x4 mfma(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct Acc<double> {
  typedef f64x4 type;
  static __device__ __forceinline__ f64x4 mfma(double a, double b, f64x4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

// Writes one 16 x 16 accumulator block.  Output layouts (column l%16 for
// lane l in both): v_mfma_f32_16x16x4_f32 / _16x16x32_bf16 keep 4 consecutive
// rows per lane group, row 4*(l/16) + r; v_mfma_f64_16x16x4_f64 interleaves
// the groups, row (l/16) + 4*r.
template <typename T, typename V>
__device__ __forceinline__ void store_block(const V& acc, int i0, int j0,
                                            int M, int N, T* C, int ldc,
                                            T alpha, T beta, const T* E,
                                            int lde) {
  const int lane = threadIdx.x & 63;
  const int j = j0 + (lane & 15);
  if (j >= N) return;
  if constexpr (sizeof(T) == 4) {
    // FP32: the lane's four rows are consecutive; one 16-byte store when
    // they are in range and aligned
    typedef float f4 __attribute__((ext_vector_type(4)));
    const int ib = i0 + 4 * (lane >> 4);
    T* cp = C + (size_t)ib + (size_t)j * ldc;
    const T* ep = E ? E + (size_t)ib + (size_t)j * lde : nullptr;
    if (ib + 3 < M && (reinterpret_cast<uintptr_t>(cp) & 15) == 0 &&
        (!ep || (reinterpret_cast<uintptr_t>(ep) & 15) == 0)) {
      f4 v;
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = alpha * static_cast<T>(acc[r]);
      if (ep) {
        const f4 e = *reinterpret_cast<const f4*>(ep);
#pragma unroll
        for (int r = 0; r < 4; ++r) v[r] += beta * e[r];
      }
      *reinterpret_cast<f4*>(cp) = v;
      return;
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = sizeof(T) == 8 ? i0 + (lane >> 4) + 4 * r : i0 + 4 * (lane >> 4) + r;
    if (i < M) {
      T v = alpha * static_cast<T>(acc[r]);
      if (E) v += beta * E[(size_t)i + (size_t)j * lde];
      C[(size_t)i + (size_t)j * ldc] = v;
    }
  }
}

// ----------------------------------------------------------- FP32 / FP64
//
// K is staged in chunks of BK (32 in FP32, 16 in FP64), each taken as groups
// of 16.  Within a group, MFMA step t (0..3) of lane l uses k = 4*(l/16) + t,
// so a lane's four k values are contiguous and are read from LDS as one
// 16-byte (FP32) / two 16-byte (FP64) loads; the four steps together cover the
// 16 k of the group.  LDS tiles are [row][k]
// with a row stride of BK + 4 elements.  The next chunk is loaded into
// registers while the current one is multiplied.  Two LDS buffers: the
// registers go to the idle buffer after the multiply, one barrier per chunk,
// and the MFMA block runs at raised wave priority so that the other
// workgroup's loads and LDS stores issue between its MFMAs (1.2-1.5x the
// single-buffer, two-barrier form on m=512..2048, n=8192..200000 products).

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec4<double> {
  typedef double type __attribute__((ext_vector_type(4)));
};

// Two workgroups per CU (two LDS buffer pairs of 36 KiB in FP32, 40 KiB in
// FP64, each).
template <typename T, bool VECLOAD>
__global__ __launch_bounds__(256, 2) void gemm_tn_kernel(
    int M, int N, int K, const T* __restrict__ A, int lda,
    const T* __restrict__ B, int ldb, T* __restrict__ C, int ldc, T alpha,
    T beta, const T* __restrict__ E, int lde, int kz, size_t zstride) {
  {
    const int kb = split_k_begin(K, kz);
    A += kb;
    B += kb;
    C += (size_t)blockIdx.z * zstride;
  }
  // K chunk per barrier pair: 32 in FP32 (16 KiB of operands per chunk),
  // 16 in FP64 (same bytes)
  constexpr int BK = sizeof(T) == 4 ? 32 : 16;
  constexpr int LD = BK + 4;  // LDS row stride (elements)
  auto pos = [](int r, int e) { return r * LD + e; };  // element e of row r
  constexpr int NV = BK / 16;  // float4 (double4) loads per row half and operand
  typedef typename Vec4<T>::type V4;
  __shared__ __attribute__((aligned(16))) T As[2][BM * LD];
  __shared__ __attribute__((aligned(16))) T Bs[2][BN * LD];
  typedef typename Acc<T>::type V;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int wm = (wid & 1) * 64, wn = (wid >> 1) * 64;
  int bm, bn;
  tile_origin(bm, bn);

  V acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[a][b][r] = 0;

  // loader: thread t loads rows (t >> 2) and (t >> 2) + 64, k quads
  // (t & 3) + 4 v for v < NV
  const int lr = tid >> 2, lq = (tid & 3) * 4;
  V4 ra[2][NV], rb[2][NV];
  // whole chunks take the vector loads (one uniform branch per chunk, so
  // the loads of a chunk issue together); the tail chunk goes element-wise
  auto load = [&](int k0) {
    if (VECLOAD && k0 + BK <= K) {
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const int i = bm + lr + 64 * h, j = bn + lr + 64 * h, k = k0 + lq + 16 * v;
          ra[h][v] = i < M ? *reinterpret_cast<const V4*>(A + (size_t)i * lda + k) : V4(0);
          rb[h][v] = j < N ? *reinterpret_cast<const V4*>(B + (size_t)j * ldb + k) : V4(0);
        }
    } else {
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const int i = bm + lr + 64 * h, j = bn + lr + 64 * h, k = k0 + lq + 16 * v;
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            ra[h][v][e] = (i < M && k + e < K) ? A[(size_t)i * lda + k + e] : T(0);
            rb[h][v][e] = (j < N && k + e < K) ? B[(size_t)j * ldb + k + e] : T(0);
          }
        }
    }
  };
  auto store = [&](int sb) {
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        *reinterpret_cast<V4*>(&As[sb][pos(lr + 64 * h, lq + 16 * v)]) = ra[h][v];
        *reinterpret_cast<V4*>(&Bs[sb][pos(lr + 64 * h, lq + 16 * v)]) = rb[h][v];
      }
  };
  const int kq = 4 * (lane >> 4);
  auto multiply = [&](int sb) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      V4 a4[4], b4[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        a4[x] = *reinterpret_cast<const V4*>(&As[sb][pos(wm + 16 * x + (lane & 15), kq + 16 * v)]);
        b4[x] = *reinterpret_cast<const V4*>(&Bs[sb][pos(wn + 16 * x + (lane & 15), kq + 16 * v)]);
      }
      __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y) acc[x][y] = Acc<T>::mfma(a4[x][t], b4[y][t], acc[x][y]);
      __builtin_amdgcn_s_setprio(0);
    }
  };

  load(0);
  // buffer cur holds chunk k0; the registers fill buffer cur ^ 1, which
  // every wave finished reading before the previous barrier
  store(0);
  __syncthreads();
  int cur = 0;
  for (int k0 = 0; k0 < K; k0 += BK) {
    const bool more = k0 + BK < K;
    if (more) load(k0 + BK);
    multiply(cur);
    if (more) store(cur ^ 1);
    __syncthreads();
    cur ^= 1;
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
      store_block<T>(acc[x][y], bm + wm + 16 * x, bn + wn + 16 * y, M, N, C,
                     ldc, alpha, beta, E, lde);
}

// ------------------------------------------------------------- BF16 x 3

// Each FP32 operand is split into hi = RNE(x) and lo = RNE(x - hi) BF16
// values, two at a time (v_cvt_pk_bf16_f32), as it goes from registers to
// LDS; 3 x v_mfma_f32_16x16x32_bf16 (lo.hi + hi.lo + hi.hi) per 16 x 16
// block.  The global loads run two 32-deep chunks ahead (two register sets,
// the chunk loop unrolled by two), through one row pointer per thread (rows
// past M / N read the last row: their products land in C elements that are
// never stored).  With the same LDS reads and MFMAs but no global loads,
// LDS stores or splits it runs 1.8x faster; operands split beforehand (no
// split in the kernel) run no faster, so the loads' latency and the loop's
// issue, not the split's VALU work, hold it back.
template <bool VECLOAD>
__global__ __launch_bounds__(256, 2) void gemm_tn_bf16x3_kernel(
    int M, int N, int K, const float* __restrict__ A, int lda,
    const float* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
    float alpha, float beta, const float* __restrict__ E, int lde, int kz, size_t zstride) {
  {
    const int kb = split_k_begin(K, kz);
    A += kb;
    B += kb;
    C += (size_t)blockIdx.z * zstride;
  }
  constexpr int BK = 32;
  // LDS rows of BK = 32 BF16 (64 bytes, no padding): row r's 16-byte slot s
  // (k 8 s .. 8 s + 7) sits at slot s ^ ((r >> 1) & 3).  The 8 lanes an LDS
  // access serves together then hit distinct banks both when the MFMA
  // operands are read (8 consecutive rows, one slot) and when the loaded
  // chunk is stored (4 rows, two slots 2 apart).
  constexpr int LD = BK;
  typedef float f4 __attribute__((ext_vector_type(4)));
  __shared__ __attribute__((aligned(16))) __bf16 Ah[2][BM * LD], Al[2][BM * LD];
  __shared__ __attribute__((aligned(16))) __bf16 Bh[2][BN * LD], Bl[2][BN * LD];
  auto swz = [](int r, int slot) { return r * LD + ((slot ^ ((r >> 1) & 3)) << 3); };
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wid = tid >> 6;
  const int wm = (wid & 1) * 64, wn = (wid >> 1) * 64;
  int bm, bn;
  tile_origin(bm, bn);
  f32x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[a][b][r] = 0;
  const int lr = tid >> 1, lk = (tid & 1) * 16;
  const int o0 = swz(lr, lk >> 3), o1 = swz(lr, (lk >> 3) + 1);  // LDS offsets of its 2 slots
  struct Regs { f4 a[4], b[4]; };
  // this thread's rows of A and B (rows past M / N read row M-1 / N-1,
  // whose products land in C rows / columns that are never stored)
  const float* pa = A + (size_t)(bm + lr < M ? bm + lr : M - 1) * lda + lk;
  const float* pb = B + (size_t)(bn + lr < N ? bn + lr : N - 1) * ldb + lk;
  auto load = [&](int k0, Regs& R) {
    if (VECLOAD && k0 + BK <= K) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        R.a[v] = *reinterpret_cast<const f4*>(pa + k0 + 4 * v);
        R.b[v] = *reinterpret_cast<const f4*>(pb + k0 + 4 * v);
      }
    } else {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int k = k0 + lk + 4 * v + e;
          R.a[v][e] = k < K ? pa[k0 + 4 * v + e] : 0.f;
          R.b[v][e] = k < K ? pb[k0 + 4 * v + e] : 0.f;
        }
      }
    }
  };
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  typedef float f2 __attribute__((ext_vector_type(2)));
  // hi = RNE(x) per pair (v_cvt_pk_bf16_f32), lo = RNE(x - hi)
  auto split8 = [&](const f4& p, const f4& q, bf16x8& h, bf16x8& l) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const f2 x = c < 2 ? f2{p[2 * c], p[2 * c + 1]} : f2{q[2 * c - 4], q[2 * c - 3]};
      const bf16x2 hh = __builtin_convertvector(x, bf16x2);
      const unsigned hb = __builtin_bit_cast(unsigned, hh);
      const f2 hf = {__uint_as_float(hb << 16), __uint_as_float(hb & 0xffff0000u)};
      const bf16x2 ll = __builtin_convertvector(x - hf, bf16x2);
      h[2 * c] = hh[0];
      h[2 * c + 1] = hh[1];
      l[2 * c] = ll[0];
      l[2 * c + 1] = ll[1];
    }
  };
  auto store = [&](int sb, const Regs& R) {
    bf16x8 h, l;
    split8(R.a[0], R.a[1], h, l);
    *reinterpret_cast<bf16x8*>(&Ah[sb][o0]) = h;
    *reinterpret_cast<bf16x8*>(&Al[sb][o0]) = l;
    split8(R.a[2], R.a[3], h, l);
    *reinterpret_cast<bf16x8*>(&Ah[sb][o1]) = h;
    *reinterpret_cast<bf16x8*>(&Al[sb][o1]) = l;
    split8(R.b[0], R.b[1], h, l);
    *reinterpret_cast<bf16x8*>(&Bh[sb][o0]) = h;
    *reinterpret_cast<bf16x8*>(&Bl[sb][o0]) = l;
    split8(R.b[2], R.b[3], h, l);
    *reinterpret_cast<bf16x8*>(&Bh[sb][o1]) = h;
    *reinterpret_cast<bf16x8*>(&Bl[sb][o1]) = l;
  };
  // this lane's operand offsets: row (lane & 15) of each 16-row block, slot
  // lane >> 4 (rows 16 x apart share the swizzle)
  const int ro = swz(lane & 15, lane >> 4);
  auto mma = [&](int cur) {
    bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int ra_ = (wm + 16 * x) * LD + ro;
      const int rb_ = (wn + 16 * x) * LD + ro;
      ah[x] = *reinterpret_cast<const bf16x8*>(&Ah[cur][ra_]);
      al[x] = *reinterpret_cast<const bf16x8*>(&Al[cur][ra_]);
      bh[x] = *reinterpret_cast<const bf16x8*>(&Bh[cur][rb_]);
      bl[x] = *reinterpret_cast<const bf16x8*>(&Bl[cur][rb_]);
    }
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[x], bh[y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bl[y], acc[x][y], 0, 0, 0);
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[x], bh[y], acc[x][y], 0, 0, 0);
      }
    __builtin_amdgcn_s_setprio(0);
  };
  Regs R0, R1;
  // chunk c lives in registers R(c % 2) until stored into LDS buffer c % 2
  load(0, R0);
  if (BK < K) load(BK, R1);
  store(0, R0);
  if (2 * BK < K) load(2 * BK, R0);
  __syncthreads();
  for (int k0 = 0; k0 < K; k0 += 2 * BK) {
    // even chunk k0 in buffer 0; R1 holds chunk k0 + BK, R0 chunk k0 + 2 BK
    mma(0);
    if (k0 + BK < K) {
      store(1, R1);
      if (k0 + 3 * BK < K) load(k0 + 3 * BK, R1);
    }
    __syncthreads();
    if (k0 + BK >= K) break;
    mma(1);
    if (k0 + 2 * BK < K) {
      store(0, R0);
      if (k0 + 4 * BK < K) load(k0 + 4 * BK, R0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
      store_block<float>(acc[x][y], bm + wm + 16 * x, bn + wn + 16 * y, M, N, C, ldc, alpha, beta,
                         E, lde);
}

// ----------------------------------------------------- column sq. norms

template <typename T>
__global__ __launch_bounds__(256) void col_sqnorm_kernel(
    int m, int n, const T* __restrict__ X, int ldx, T* __restrict__ out) {
  const int col = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (col >= n) return;
  const T* x = X + (size_t)col * ldx;
  T s = 0;
  for (int i = lane; i < m; i += 64) s += x[i] * x[i];
  s = wave_sum(s);
  if (lane == 0) out[col] = s;
}

}  // namespace

template <typename T>
int gemm_tn(hipStream_t st, int M, int N, int K, const T* A, int lda,
            const T* B, int ldb, T* C, int ldc, T alpha, T beta, const T* E,
            int lde, int precision) {
  if (M <= 0 || N <= 0) return SPX_OK;
  dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN);
  if (grid.y > 65535) {
    // split very wide products into column panels
    const int panel = 65535 * BN;
    for (int j0 = 0; j0 < N; j0 += panel) {
      const int nb = (N - j0) < panel ? (N - j0) : panel;
      int rc = gemm_tn<T>(st, M, nb, K, A, lda, B + (size_t)j0 * ldb, ldb,
                          C + (size_t)j0 * ldc, ldc, alpha, beta,
                          E ? E + (size_t)j0 * lde : nullptr, lde, precision);
      if (rc) return rc;
    }
    return SPX_OK;
  }
  if constexpr (sizeof(T) == 4) {
    if (precision & SPX_PREC_BF16X3) {
      const bool vec = (lda % 4 == 0) && (ldb % 4 == 0) &&
                       (reinterpret_cast<uintptr_t>(A) % 16 == 0) &&
                       (reinterpret_cast<uintptr_t>(B) % 16 == 0);
      if (vec)
        hipLaunchKernelGGL(gemm_tn_bf16x3_kernel<true>, grid, dim3(256), 0, st, M, N,
                           K, A, lda, B, ldb, C, ldc, alpha, beta, E, lde, 0, (size_t)0);
      else
        hipLaunchKernelGGL(gemm_tn_bf16x3_kernel<false>, grid, dim3(256), 0, st, M, N,
                           K, A, lda, B, ldb, C, ldc, alpha, beta, E, lde, 0, (size_t)0);
      SPX_HIP(hipGetLastError());
      return SPX_OK;
    }
  }
  // 16-byte-aligned rows of whole quads allow the vector loads
  const bool vec = (lda % 4 == 0) && (ldb % 4 == 0) &&
                   (reinterpret_cast<uintptr_t>(A) % (4 * sizeof(T)) == 0) &&
                   (reinterpret_cast<uintptr_t>(B) % (4 * sizeof(T)) == 0);
  if (vec)
    hipLaunchKernelGGL((gemm_tn_kernel<T, true>), grid, dim3(256), 0, st, M, N, K, A,
                       lda, B, ldb, C, ldc, alpha, beta, E, lde, 0, (size_t)0);
  else
    hipLaunchKernelGGL((gemm_tn_kernel<T, false>), grid, dim3(256), 0, st, M, N, K, A,
                       lda, B, ldb, C, ldc, alpha, beta, E, lde, 0, (size_t)0);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

namespace {

// C = (accumulate ? C : 0) + sum over z of part[z] (n elements each, zstride
// apart), summed in slice order.
template <typename T>
__global__ void split_sum_kernel(const T* __restrict__ part, int S, size_t zstride, long long n,
                                 T* C, int accumulate) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  T s = accumulate ? C[e] : T(0);
  for (int z = 0; z < S; ++z) s += part[(size_t)z * zstride + e];
  C[e] = s;
}

// C(i, j) = alpha * sum over z of part[z](i, j) + beta * E(i, j), the slices
// summed in order (part: M x N with leading dimension M per slice).
template <typename T>
__global__ void split_sum_epi_kernel(const T* __restrict__ part, int S, size_t zstride, int M,
                                     long long n, T* __restrict__ C, int ldc, T alpha, T beta,
                                     const T* __restrict__ E, int lde) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const long long j = e / M;
  const int i = (int)(e - j * M);
  T s = T(0);
  for (int z = 0; z < S; ++z) s += part[(size_t)z * zstride + e];
  T v = alpha * s;
  if (E) v += beta * E[(size_t)i + (size_t)j * lde];
  C[(size_t)i + (size_t)j * ldc] = v;
}

}  // namespace

// Launches the slices of a split product into `part`; returns the slice
// count used (whole 32-deep chunks per slice), 0 on a launch error.
template <typename T>
static int split_partials(hipStream_t st, int M, int N, int K, const T* A, int lda, const T* B,
                          int ldb, T* part, int S, int precision);

template <typename T>
int gemm_tn_split(hipStream_t st, int M, int N, int K, const T* A, int lda, const T* B, int ldb,
                  T* C, int accumulate, T* part, int S, int precision) {
  if (M <= 0 || N <= 0) return SPX_OK;
  if ((N + BN - 1) / BN > 65535) {
    set_error("gemm_tn_split: N=%d too wide", N);
    return SPX_ERR_UNSUP;
  }
  S = split_partials<T>(st, M, N, K, A, lda, B, ldb, part, S, precision);
  if (S == 0) return SPX_ERR_HIP;
  const size_t zstride = (size_t)M * N;
  const long long n = (long long)M * N;
  hipLaunchKernelGGL(split_sum_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     static_cast<const T*>(part), S, zstride, n, C, accumulate);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

template <typename T>
int gemm_tn_product(hipStream_t st, int M, int N, int K, const T* A, int lda, const T* B,
                    int ldb, T* C, int ldc, T alpha, T beta, const T* E, int lde, T* part,
                    int S, int precision) {
  if (S <= 1 || !part || (N + BN - 1) / BN > 65535)
    return gemm_tn<T>(st, M, N, K, A, lda, B, ldb, C, ldc, alpha, beta, E, lde, precision);
  if (M <= 0 || N <= 0) return SPX_OK;
  S = split_partials<T>(st, M, N, K, A, lda, B, ldb, part, S, precision);
  if (S == 0) return SPX_ERR_HIP;
  const long long n = (long long)M * N;
  hipLaunchKernelGGL(split_sum_epi_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     static_cast<const T*>(part), S, (size_t)M * N, M, n, C, ldc, alpha, beta, E,
                     lde);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

template <typename T>
static int split_partials(hipStream_t st, int M, int N, int K, const T* A, int lda, const T* B,
                          int ldb, T* part, int S, int precision) {
  const int mt = (M + BM - 1) / BM, nt = (N + BN - 1) / BN;
  if (S < 1) S = 1;
  int kz = (K + S - 1) / S;
  kz = (kz + 31) / 32 * 32;  // whole 32-deep chunks; 16-byte aligned slice starts
  S = (K + kz - 1) / kz;
  if (S < 1) S = 1;
  const size_t zstride = (size_t)M * N;
  dim3 grid(mt, nt, S);
  const bool vec = (lda % 4 == 0) && (ldb % 4 == 0) &&
                   (reinterpret_cast<uintptr_t>(A) % (4 * sizeof(T)) == 0) &&
                   (reinterpret_cast<uintptr_t>(B) % (4 * sizeof(T)) == 0);
  bool launched = false;
  if constexpr (sizeof(T) == 4) {
    if (precision & SPX_PREC_BF16X3) {
      if (vec)
        hipLaunchKernelGGL(gemm_tn_bf16x3_kernel<true>, grid, dim3(256), 0, st, M, N, K, A, lda,
                           B, ldb, part, M, 1.f, 0.f, static_cast<const float*>(nullptr), 0, kz,
                           zstride);
      else
        hipLaunchKernelGGL(gemm_tn_bf16x3_kernel<false>, grid, dim3(256), 0, st, M, N, K, A,
                           lda, B, ldb, part, M, 1.f, 0.f, static_cast<const float*>(nullptr),
                           0, kz, zstride);
      launched = true;
    }
  }
  if (!launched) {
    if (vec)
      hipLaunchKernelGGL((gemm_tn_kernel<T, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B,
                         ldb, part, M, T(1), T(0), static_cast<const T*>(nullptr), 0, kz,
                         zstride);
    else
      hipLaunchKernelGGL((gemm_tn_kernel<T, false>), grid, dim3(256), 0, st, M, N, K, A, lda,
                         B, ldb, part, M, T(1), T(0), static_cast<const T*>(nullptr), 0, kz,
                         zstride);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("gemm split launch failed: %s", hipGetErrorString(e));
    return 0;
  }
  return S;
}

// Slices for gemm_tn_split: enough workgroups to fill the chip (about 512
// output-tile slices), each slice at least 256 deep.
int gemm_split_count(int M, int N, int K) {
  const int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  int S = (512 + tiles - 1) / tiles;
  const int maxS = K / 256 > 1 ? K / 256 : 1;
  if (S > maxS) S = maxS;
  return S < 1 ? 1 : S;
}

// Slices for a product that leaves most of the chip idle (fewer than 128
// output tiles, e.g. a small batch of signals); 1 otherwise.
int gemm_small_split(int M, int N, int K) {
  const int tiles = ((M + BM - 1) / BM) * ((N + BN - 1) / BN);
  return tiles < 128 ? gemm_split_count(M, N, K) : 1;
}

template <typename T>
int col_sqnorms(hipStream_t st, int m, int n, const T* X, int ldx, T* out) {
  if (n <= 0) return SPX_OK;
  hipLaunchKernelGGL(col_sqnorm_kernel<T>, dim3((n + 3) / 4), dim3(256), 0, st,
                     m, n, X, ldx, out);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

template int gemm_tn<float>(hipStream_t, int, int, int, const float*, int,
                            const float*, int, float*, int, float, float,
                            const float*, int, int);
template int gemm_tn<double>(hipStream_t, int, int, int, const double*, int,
                             const double*, int, double*, int, double, double,
                             const double*, int, int);
template int col_sqnorms<float>(hipStream_t, int, int, const float*, int,
                                float*);
template int gemm_tn_split<float>(hipStream_t, int, int, int, const float*, int, const float*,
                                  int, float*, int, float*, int, int);
template int gemm_tn_product<float>(hipStream_t, int, int, int, const float*, int, const float*,
                                    int, float*, int, float, float, const float*, int, float*,
                                    int, int);
template int gemm_tn_product<double>(hipStream_t, int, int, int, const double*, int,
                                     const double*, int, double*, int, double, double,
                                     const double*, int, double*, int, int);
template int gemm_tn_split<double>(hipStream_t, int, int, int, const double*, int,
                                   const double*, int, double*, int, double*, int, int);
template int col_sqnorms<double>(hipStream_t, int, int, const double*, int,
                                 double*);

}  // namespace spx
