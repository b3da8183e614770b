// Online dictionary learning (trainDL): Trainer::train, dictLearn/dicts.h:346-759,
// behind the facade _alltrainDL (cpp_library/spams.h:1153-1340).
//
// One iteration on the device:
//   G = D'D                      MFMA GEMM (Matrix::XtX, dicts.h:553)
//   cleaning                     scan for the first offending atom pair, replace,
//                                continue the scan (Trainer::cleanDict, :296-334)
//   G += max(lambda2, 1e-10) I   (dicts.h:557)
//   code a minibatch             gather columns perm[(j + i bs) mod M], D'X on MFMA,
//                                the engine's LARS (coreLARS2) or ORMP (coreORMPB)
//   A_t = sum a a', B_t = sum x a'   dense codes, two MFMA GEMMs (rank1Update, :601-602)
//   A, B statistics              online scaling with the even/odd epoch buffers
//                                (:668-715) or batch mode (:604-609)
//   dictionary update            block-coordinate descent over the atoms (:717-741)
//
// The dictionary update is sequential over atoms: d_k depends on every d_c
// updated before it.  It runs in blocks of `b` atoms: a panel kernel forms
// P = D A(:, block) with the dictionary as it stands at the block's start
// (many workgroups), then one workgroup walks the block's atoms in order,
// adding the changes of the atoms it already updated in this block
// (sum_c (d_c^new - d_c^old) A(c, k)), so that every d_k sees exactly
// D A(:, k) of the reference's gemv.
#include <sys/time.h>
#include <unistd.h>

#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "spx.h"
#include "spx_common.h"
#include "spx_internal.h"

namespace spx {

namespace {

constexpr int kDlMaxRowsPerThread = 8;  // block kernel: m <= 256 * 8
constexpr int kDlBlockThreads = 256;
constexpr int kDlMaxSplit = 16;         // panel kernel: at most this many inner chunks
// panel kernel: inner atoms per LDS slice (the D and A slices fill 33 KiB)
template <typename T>
constexpr int dl_panel_inner() { return sizeof(T) == 4 ? 64 : 32; }
constexpr int kDlPanelInner = 64;       // chunk granularity (a multiple of both slices)
constexpr int kDlChunk = 16384;         // signals coded per solver launch

// The C library's additive-feedback generator (random_r) with a private
// 128-byte state seeded like srandom(0): the stream the reference draws its
// initial atoms, permutation and replacement signals from (dicts.h:400, :459).
struct Rng {
  random_data d;
  char state[128];
  Rng() { seed(); }
  Rng(const Rng&) = delete;
  Rng& operator=(const Rng&) = delete;
  void seed() {
    std::memset(&d, 0, sizeof d);
    std::memset(state, 0, sizeof state);
    initstate_r(0, state, sizeof state, &d);
  }
  long next() {
    int32_t r;
    random_r(&d, &r);
    return (long)r;
  }
};

// Makes ctx's device current for the lifetime of the guard.
struct DlDevice {
  int prev = -1;
  explicit DlDevice(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DlDevice() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

double wall_seconds() {
  timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

// ---------------------------------------------------------------- kernels

// d0[e] = s0[e], d1[e] = s1[e], d2[e] = s2[e] for e < n: the three copies'
// loads issued together, U per array and thread, before their LDS stores.
template <typename T, int U, int NT>
__device__ __forceinline__ void dl_fill3(T* d0, const T* __restrict__ s0, T* d1,
                                         const T* __restrict__ s1, T* d2,
                                         const T* __restrict__ s2, int n) {
  for (int base = threadIdx.x; base < n; base += U * NT) {
    T v0[U], v1[U], v2[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = base + u * NT;
      const bool in = e < n;
      v0[u] = in ? s0[e] : T(0);
      v1[u] = in ? s1[e] : T(0);
      v2[u] = in ? s2[e] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = base + u * NT;
      if (e < n) {
        d0[e] = v0[u];
        d1[e] = v1[u];
        d2[e] = v2[u];
      }
    }
  }
}

// dst[e] = src[idx(e)] for e < n, threads striding by NT, with U loads in
// flight per thread before their LDS stores (a load-store loop with a
// runtime trip count would wait for every load in turn).
template <typename T, int U, int NT, typename Idx>
__device__ __forceinline__ void dl_fill(T* dst, const T* __restrict__ src, int n, Idx idx) {
  for (int base = threadIdx.x; base < n; base += U * NT) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = base + u * NT;
      v[u] = e < n ? src[idx(e)] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = base + u * NT;
      if (e < n) dst[e] = v[u];
    }
  }
}

// Xb(:, j) = X(:, perm[(base + j) mod M]) for j < nb, and its transpose XbT
// (ldt x m, rows nb..ldt-1 zero); 64 x 64 tiles through LDS.
template <typename T>
__global__ __launch_bounds__(256) void dl_gather_kernel(const T* __restrict__ X, int ldx,
                                                        int m, int M,
                                                        const int* __restrict__ perm,
                                                        long long base, int nb, T* Xb,
                                                        T* XbT, int ldt) {
  __shared__ T tile[64][65];
  const int j0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int jj = ty; jj < 64; jj += 4) {
    const int j = j0 + jj;
    const int r = r0 + tx;
    T v = 0;
    if (j < nb && r < m) {
      const int col = perm[(int)((base + j) % M)];
      v = X[(size_t)col * ldx + r];
      Xb[(size_t)j * m + r] = v;
    }
    tile[jj][tx] = v;
  }
  __syncthreads();
  for (int rr = ty; rr < 64; rr += 4) {
    const int r = r0 + rr;
    const int j = j0 + tx;
    if (r < m && j < ldt) XbT[(size_t)r * ldt + j] = tile[tx][rr];
  }
}

// aT(j, r[t, j]) = v[t, j] for t < count[j] (aT zeroed beforehand).
template <typename T>
__global__ void dl_scatter_kernel(const T* __restrict__ v, const int* __restrict__ r,
                                  const int* __restrict__ count, int L, int nb, T* aT,
                                  int ldt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  const int c = count[j];
  for (int t = 0; t < c; ++t) {
    const int k = r[(size_t)j * L + t];
    aT[(size_t)k * ldt + j] = v[(size_t)j * L + t];
  }
}

// lower triangle <- upper triangle (K x K, leading dimension K): the batch
// statistics A_t = sum a a' exactly symmetric, as the reference's sparse
// rank-one updates make it.
template <typename T>
__global__ void dl_mirror_kernel(T* A, int K) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)K * K) return;
  const int row = (int)(e % K), col = (int)(e / K);
  if (row > col) A[e] = A[(size_t)row * K + col];
}

template <typename T>
__global__ void dl_ridge_kernel(T* G, int ldg, int K, T diag) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < K) G[(size_t)k * ldg + k] += diag;
}

// Online statistics (dicts.h:670-715), element-wise over n entries:
//   swap:     odd <- even, even <- 0 (a new epoch parity)
//   even, odd *= scal; orig *= scal when use_orig
//   even += scal2 * bat;  S = (use_orig ? orig : 0) + odd + even
template <typename T>
__global__ void dl_combine_kernel(long long n, T* S, T* orig, T* odd, T* even,
                                  const T* __restrict__ bat, int swap, T scal, int use_orig,
                                  T scal2) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  T ev = even[e], od = odd[e];
  if (swap) {
    od = ev;
    ev = 0;
  }
  ev = ev * scal;
  od = od * scal;
  T s = 0;
  if (use_orig) {
    const T o = orig[e] * scal;
    orig[e] = o;
    s = o;
  }
  ev = ev + scal2 * bat[e];
  s = s + od;
  s = s + ev;
  even[e] = ev;
  odd[e] = od;
  S[e] = s;
}

// A = t0 I, B = t0 D (the fresh online model, dicts.h:434-441); zero both in
// batch mode.
template <typename T>
__global__ void dl_init_model_kernel(T* A, T* B, const T* __restrict__ D, int m, int K,
                                     T t0, int online) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nA = (long long)K * K, nB = (long long)m * K;
  if (e < nA) A[e] = (online && e % (K + 1) == 0) ? t0 : T(0);
  if (e < nB) B[e] = online ? t0 * D[e] : T(0);
}

// D(:, k) = X(:, idx[k]) (initial atoms, dicts.h:403-410).
template <typename T>
__global__ void dl_copy_cols_kernel(const T* __restrict__ X, int ldx, int m,
                                    const int* __restrict__ idx, int K, T* D) {
  const int k = blockIdx.y;
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < K && r < m) D[(size_t)k * m + r] = X[(size_t)idx[k] * ldx + r];
}

// Column-wise thrsPos (when pos) and normalisation, one wavefront per column:
// scale to unit norm when the norm exceeds `thr` (Matrix::normalize with
// thr = 1e-10, linalg/linalg.h:1297-1312; Vector::normalize with EPSILON =
// 10e-10, :3155-3158).  cols: column indices (null = 0..ncols-1).
template <typename T>
__global__ __launch_bounds__(256) void dl_normalize_kernel(T* D, int m, int ncols,
                                                           const int* __restrict__ cols,
                                                           int pos, T thr) {
  const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= ncols) return;
  T* d = D + (size_t)(cols ? cols[w] : w) * m;
  T s = 0;
  for (int r = lane; r < m; r += 64) {
    T v = d[r];
    if (pos && v < T(0)) {
      v = 0;
      d[r] = v;
    }
    s += v * v;
  }
  s = wave_sum(s);
  const T nrm = sqrt(s);
  if (nrm > thr) {
    const T inv = T(1) / nrm;
    for (int r = lane; r < m; r += 64) d[r] = d[r] * inv;
  }
}

// Cleaning scan: the first pair (i, j >= i), in the reference's row-major
// order and at linear position i K + j >= start, whose atoms are correlated
// beyond 0.999999 (j > i) or whose atom has a squared norm below 1e-4
// (j == i); atomicMin of its position into *out (INT_MAX when none).
template <typename T>
__global__ void dl_clean_find_kernel(const T* __restrict__ G, int ldg, int K, int start,
                                     int* out) {
  const long long total = (long long)K * K;
  for (long long e = start + (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / K), j = (int)(e % K);
    if (j < i) continue;
    const T gij = G[(size_t)i * ldg + j];
    bool hit;
    if (j == i) {
      hit = fabs(gij) < T(1e-4);
    } else {
      hit = fabs(gij) / sqrt(G[(size_t)i * ldg + i] * G[(size_t)j * ldg + j]) > T(0.999999);
    }
    if (hit) atomicMin(out, (int)e);
  }
}

// Row and column j of G from the replaced atom d_j: G(k, j) = G(j, k) = d_k'd_j
// (the gemv and the copy loop of dicts.h:326-329); one wavefront per k.
template <typename T>
__global__ __launch_bounds__(256) void dl_gram_col_kernel(const T* __restrict__ D, int m,
                                                          int K, int j, T* G, int ldg) {
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= K) return;
  const T* dk = D + (size_t)k * m;
  const T* dj = D + (size_t)j * m;
  T s = 0;
  for (int r = lane; r < m; r += 64) s += dk[r] * dj[r];
  s = wave_sum(s);
  if (lane == 0) {
    G[(size_t)j * ldg + k] = s;
    G[(size_t)k * ldg + j] = s;
  }
}

// Partial panels of the dictionary update: Pp(s, kk, r) = sum over the
// inner chunk s (atoms s kc .. s kc + kc - 1) of D(r, c) A(c, k0 + kk), for
// kk < nk (<= 64).  Workgroup = 64 rows x one chunk; its 4 waves take the
// atoms kk = w, w + 4, ...; the slices of D (64 rows) and of A (columns
// k0 .. k0 + nk - 1) go through LDS, IN inner atoms at a time.
template <typename T>
__global__ __launch_bounds__(256) void dl_panel_kernel(const T* __restrict__ D, int m, int K,
                                                       const T* __restrict__ A, int k0, int nk,
                                                       int kc, T* __restrict__ Pp) {
  constexpr int IN = dl_panel_inner<T>();
  __shared__ T As[64][IN + 1];
  __shared__ T Ds[IN][64 + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r0 = blockIdx.x * 64;
  const int r = r0 + lane;
  const int s = blockIdx.y;
  const int cb = s * kc;
  const int ce = cb + kc < K ? cb + kc : K;
  T acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0;
  for (int c0 = cb; c0 < ce; c0 += IN) {
    const int nc = ce - c0 < IN ? ce - c0 : IN;
    __syncthreads();  // the previous slices are consumed
    {
      T va[64 * IN / 256], vd[IN * 64 / 256];
#pragma unroll
      for (int u = 0; u < 64 * IN / 256; ++u) {
        const int e = tid + u * 256;
        const int q = e / IN, c = e % IN;
        va[u] = (q < nk && c < nc) ? A[(size_t)(k0 + q) * K + c0 + c] : T(0);
      }
#pragma unroll
      for (int u = 0; u < IN * 64 / 256; ++u) {
        const int e = tid + u * 256;
        const int c = e / 64, rr = e % 64;
        vd[u] = (c < nc && r0 + rr < m) ? D[(size_t)(c0 + c) * m + r0 + rr] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 64 * IN / 256; ++u) {
        const int e = tid + u * 256;
        As[e / IN][e % IN] = va[u];
      }
#pragma unroll
      for (int u = 0; u < IN * 64 / 256; ++u) {
        const int e = tid + u * 256;
        Ds[e / 64][e % 64] = vd[u];
      }
    }
    __syncthreads();
    // rows q >= nk of As are zero: no guard, so the LDS reads batch
#pragma unroll 4
    for (int c = 0; c < nc; ++c) {
      const T d = Ds[c][lane];
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[j] += d * As[w + 4 * j][c];
    }
  }
  if (r < m) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int q = w + 4 * j;
      if (q < nk) Pp[((size_t)s * nk + q) * m + r] = acc[j];
    }
  }
}

// P(r, kk) = sum of the S partial panels, in chunk order.
template <typename T>
__global__ __launch_bounds__(256) void dl_panel_sum_kernel(const T* __restrict__ Pp, int S,
                                                           int n, T* __restrict__ P) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  T t = 0;
  for (int s = 0; s < S; ++s) t += Pp[(size_t)s * n + e];
  P[e] = t;
}

// The atoms k0 .. k0 + nk - 1 of the dictionary update in order (one
// workgroup; thread t owns rows t, t + 256, ...):
//   t = P(:, kk) + sum_{c < kk} dD(:, c) A(k0 + c, k)
//   A_kk > 1e-6:  d = (b_k - t) / A_kk + d_k, thrsPos when posD, then scaled to
//                 norm 1 when its norm exceeds 1 (Vector::normalize2)
//   otherwise:    d = 0 when zero_unused, else unchanged
// The block's P, B and D columns are staged in LDS first; dD(:, kk) holds
// P(:, kk), then d - d_k^old for the atoms after it; the new columns go back
// to D at the end, so the steps touch LDS only.
template <typename T>
__global__ __launch_bounds__(kDlBlockThreads) void dl_block_kernel(
    T* D, int m, int K, const T* __restrict__ A, const T* __restrict__ B, int k0, int nk,
    const T* __restrict__ P, int posD, int zero_unused) {
  extern __shared__ unsigned char dl_lds[];
  const int nm = nk * m;
  T* dD = reinterpret_cast<T*>(dl_lds);  // nk x m (atom-major)
  T* Bs = dD + nm;                       // B(:, k0 ..) nk x m
  T* Db = Bs + nm;                       // D(:, k0 ..) nk x m, updated in place
  T* Ab = Db + nm;                       // nk x nk block of A
  T* red = Ab + (size_t)nk * nk;         // 2 x 4 wave partials
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const T* Bb = B + (size_t)k0 * m;
  T* Dg = D + (size_t)k0 * m;
  dl_fill3<T, 16, kDlBlockThreads>(dD, P, Bs, Bb, Db, Dg, nm);
  // Ab[q * nk + c] = A(k0 + c, k0 + q)
  dl_fill<T, 16, kDlBlockThreads>(Ab, A, nk * nk, [=](int e) {
    return (size_t)(k0 + e / nk) * K + k0 + e % nk;
  });
  __syncthreads();
  const int R = (m + kDlBlockThreads - 1) / kDlBlockThreads;
  int phase = 0;  // reduction slot parity, flipped on every step with a barrier
  for (int kk = 0; kk < nk; ++kk) {
    const T akk = Ab[kk * nk + kk];
    T* dcol = dD + (size_t)kk * m;
    T* Dk = Db + (size_t)kk * m;
    const T* bk = Bs + (size_t)kk * m;
    if (akk > T(1e-6)) {
      const T inv = T(1) / akk;
      T nd[kDlMaxRowsPerThread];
      T s = 0;
#pragma unroll
      for (int q = 0; q < kDlMaxRowsPerThread; ++q) {
        nd[q] = 0;
        const int r = tid + q * kDlBlockThreads;
        if (q < R && r < m) {
          // four partial sums, so that the LDS reads of the chain overlap
          const T* ab = Ab + kk * nk;
          T t0 = dcol[r], t1 = 0, t2 = 0, t3 = 0;
          int c = 0;
          for (; c + 4 <= kk; c += 4) {
            t0 += dD[(size_t)c * m + r] * ab[c];
            t1 += dD[(size_t)(c + 1) * m + r] * ab[c + 1];
            t2 += dD[(size_t)(c + 2) * m + r] * ab[c + 2];
            t3 += dD[(size_t)(c + 3) * m + r] * ab[c + 3];
          }
          for (; c < kk; ++c) t0 += dD[(size_t)c * m + r] * ab[c];
          const T t = (t0 + t1) + (t2 + t3);
          T v = (bk[r] - t) * inv + Dk[r];
          if (posD && v < T(0)) v = 0;
          nd[q] = v;
          s += v * v;
        }
      }
      s = wave_sum(s);
      T* slot = red + phase * 4;
      phase ^= 1;
      if (lane == 0) slot[w] = s;
      __syncthreads();
      const T nrm = sqrt((slot[0] + slot[1]) + (slot[2] + slot[3]));
      const T sc = nrm > T(1) ? T(1) / nrm : T(1);
#pragma unroll
      for (int q = 0; q < kDlMaxRowsPerThread; ++q) {
        const int r = tid + q * kDlBlockThreads;
        if (q < R && r < m) {
          const T v = nrm > T(1) ? nd[q] * sc : nd[q];
          dcol[r] = v - Dk[r];
          Dk[r] = v;
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < kDlMaxRowsPerThread; ++q) {
        const int r = tid + q * kDlBlockThreads;
        if (q < R && r < m) {
          if (zero_unused) {
            dcol[r] = -Dk[r];
            Dk[r] = 0;
          } else {
            dcol[r] = 0;
          }
        }
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < nm; e += kDlBlockThreads) Dg[e] = Db[e];
}

// Atoms per block of the register form below: its per-thread row of the
// block's P, B and D columns (3 x NK values) stays in registers.
template <typename T>
constexpr int dl_push_atoms() {
  return sizeof(T) == 4 ? 64 : 32;
}

// dl_block_kernel for m <= kDlBlockThreads (one row per thread), with the
// block's columns in registers and the corrections pushed forward: once
// atom c is updated by dd = d_c^new - d_c^old, every later atom j of the
// block takes t_j += dd(r) A(k0 + c, k0 + j) at once, so that atom kk finds
// t_kk = P(:, kk) + sum_{c < kk} dd_c A(k0 + c, k0 + kk) ready (the sum in
// atom order) and its step is one element per thread, a group-wide norm and
// the pushes, whose A row comes from LDS.  The atom loop is unrolled, so
// the register arrays are indexed by constants.
template <typename T>
__global__ __launch_bounds__(kDlBlockThreads) void dl_block_push_kernel(
    T* D, int m, int K, const T* __restrict__ A, const T* __restrict__ B, int k0, int nk,
    const T* __restrict__ P, int posD, int zero_unused) {
  constexpr int NK = dl_push_atoms<T>();
  __shared__ __attribute__((aligned(16))) T Ar[NK * NK];  // Ar[c NK + j] = A(k0 + c, k0 + j)
  __shared__ T red[2 * 4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r = tid;
  const bool live = r < m;
  T t[NK], bb[NK], d[NK];
  // unconditional loads (row and atom clamped into range), then zeros
  // outside: no branch between the loads, so they are all in flight
  const int rr = live ? r : m - 1;
#pragma unroll
  for (int j = 0; j < NK; ++j) {
    const bool in = live && j < nk;
    const int jj = j < nk ? j : nk - 1;
    const T pv = P[(size_t)jj * m + rr];
    const T bv = B[(size_t)(k0 + jj) * m + rr];
    const T dv = D[(size_t)(k0 + jj) * m + rr];
    t[j] = in ? pv : T(0);
    bb[j] = in ? bv : T(0);
    d[j] = in ? dv : T(0);
  }
  // entries past nk are zero, so the pushes need no guard
  for (int e = tid; e < NK * NK; e += kDlBlockThreads) {
    const int c = e / NK, j = e - c * NK;
    Ar[e] = (c < nk && j < nk) ? A[(size_t)(k0 + j) * K + k0 + c] : T(0);
  }
  __syncthreads();
  // 1 / A_kk of the next atom is formed during the step before it, off the
  // step's dependency chain
  T akk = Ar[0];
  T inv = T(1) / akk;
  // reduction slot parity, flipped on every step that has a barrier (an
  // atom with A_kk <= 1e-6 has none: a parity by atom would let a wave
  // overwrite a slot another wave has still to read)
  int phase = 0;
#pragma unroll
  for (int kk = 0; kk < NK; ++kk) {
    if (kk < nk) {
      // the pushes' row of A, read before the step's barrier
      const T* ar = Ar + kk * NK;
      T arow[NK];
#pragma unroll
      for (int j = kk + 1; j < NK; ++j) arow[j] = ar[j];
      const T akn = kk + 1 < NK ? Ar[(kk + 1) * NK + kk + 1] : T(1);
      const T invn = T(1) / akn;
      T dd;
      if (akk > T(1e-6)) {
        T v = live ? (bb[kk] - t[kk]) * inv + d[kk] : T(0);
        if (posD && v < T(0)) v = 0;
        const T sw = wave_sum(v * v);
        T* slot = red + phase * 4;
        phase ^= 1;
        if (lane == 0) slot[w] = sw;
        __syncthreads();
        const T nrm = sqrt((slot[0] + slot[1]) + (slot[2] + slot[3]));
        if (nrm > T(1)) v = v * (T(1) / nrm);
        dd = v - d[kk];
        d[kk] = v;
      } else {
        dd = zero_unused ? -d[kk] : T(0);
        if (zero_unused) d[kk] = 0;
      }
#pragma unroll
      for (int j = kk + 1; j < NK; ++j) t[j] += dd * arow[j];
      akk = akn;
      inv = invn;
    }
  }
  if (live) {
#pragma unroll
    for (int j = 0; j < NK; ++j)
      if (j < nk) D[(size_t)(k0 + j) * m + r] = d[j];
  }
}

// ------------------------------------------------------------------ host

inline unsigned blocks_for(long long n, int threads) {
  return (unsigned)((n + threads - 1) / threads);
}

constexpr size_t kDlBlockLds = 150 * 1024;  // of the 160 KiB of a gfx950 CU

size_t dl_block_lds(int nk, int m, size_t elem) {
  return (3 * (size_t)nk * m + (size_t)nk * nk + 8) * elem;
}

// Atoms per dictionary-update block: the block's P, B and D columns (3 of
// m x b) and its A block (b x b) in kDlBlockLds of LDS, at most 64 atoms.
template <typename T>
int dl_block_atoms(int m) {
  if (m <= kDlBlockThreads) return dl_push_atoms<T>();
  int b = 64;
  while (b > 1 && dl_block_lds(b, m, sizeof(T)) > kDlBlockLds) --b;
  return b;
}

// Inner chunk of the partial panels: a multiple of kDlPanelInner giving at
// most kDlMaxSplit chunks.
inline int dl_panel_chunk(int K) {
  int kc = (K + kDlMaxSplit - 1) / kDlMaxSplit;
  kc = (kc + kDlPanelInner - 1) / kDlPanelInner * kDlPanelInner;
  return kc;
}

// One sweep of the dictionary update.  Pp: S x b x m partial panels, then
// (from offset S b m) the summed panel b x m.
template <typename T>
int dl_update_dict(hipStream_t st, T* D, int m, int K, const T* A, const T* B, T* Pp, int b,
                   int posD, int zero_unused) {
  const int kc = dl_panel_chunk(K);
  const int S = (K + kc - 1) / kc;
  T* P = Pp + (size_t)S * b * m;
  for (int k0 = 0; k0 < K; k0 += b) {
    const int nk = K - k0 < b ? K - k0 : b;
    hipLaunchKernelGGL(dl_panel_kernel<T>, dim3((m + 63) / 64, S), dim3(256), 0, st,
                       static_cast<const T*>(D), m, K, A, k0, nk, kc, Pp);
    SPX_HIP(hipGetLastError());
    const int np = nk * m;
    hipLaunchKernelGGL(dl_panel_sum_kernel<T>, dim3((np + 255) / 256), dim3(256), 0, st,
                       static_cast<const T*>(Pp), S, np, P);
    SPX_HIP(hipGetLastError());
    if (m <= kDlBlockThreads) {
      hipLaunchKernelGGL(dl_block_push_kernel<T>, dim3(1), dim3(kDlBlockThreads), 0, st, D, m,
                         K, A, B, k0, nk, static_cast<const T*>(P), posD, zero_unused);
    } else {
      const size_t lds = dl_block_lds(nk, m, sizeof(T));
      hipLaunchKernelGGL(dl_block_kernel<T>, dim3(1), dim3(kDlBlockThreads), lds, st, D, m, K,
                         A, B, k0, nk, static_cast<const T*>(P), posD, zero_unused);
    }
    SPX_HIP(hipGetLastError());
  }
  return SPX_OK;
}

// The instantiated update sweeps of one trainer call (per zero_unused).
struct DlGraphs {
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  ~DlGraphs() {
    for (hipGraphExec_t e : exec)
      if (e) (void)hipGraphExecDestroy(e);
  }
};

struct DlMem {
  std::vector<void*> ptrs;
  ~DlMem() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename U>
  int get(U** out, size_t count) {
    void* p = nullptr;
    if (hipMalloc(&p, (count ? count : 1) * sizeof(U)) != hipSuccess) {
      set_error("trainDL: device allocation of %zu bytes failed", count * sizeof(U));
      return SPX_ERR_NOMEM;
    }
    ptrs.push_back(p);
    *out = static_cast<U*>(p);
    return SPX_OK;
  }
};

#define DL_TRY(expr)          \
  do {                        \
    int rc_ = (expr);         \
    if (rc_) return rc_;      \
  } while (0)

// The trainer on device buffers: X (m x M, leading dimension ldx), D (m x K),
// A (K x K), B (m x K), all device memory; D, A, B and *itercount in/out.
template <typename T>
int train_dl_dev(spx_ctx* ctx, const T* X, int ldx, int m, int M, T* D, int K, int D_given,
                 T* A, T* B, int* itercount, const spx_train_dl_opts* o) {
  if (!ctx || !X || !D || !A || !B || !itercount || !o || m <= 0 || M <= 0 || K <= 0 ||
      ldx < m) {
    set_error("trainDL: invalid arguments (m=%d, n=%d, K=%d)", m, M, K);
    return SPX_ERR_ARG;
  }
  const int mode = o->mode;
  const T lambda1 = static_cast<T>(o->lambda1);
  const T lambda2 = static_cast<T>(o->lambda2);
  if (mode < 0 || mode > 5) {
    set_error("trainDL: unsupported mode %d", mode);
    return SPX_ERR_UNSUP;
  }
  if (mode == 2 && lambda1 == T(0) && lambda2 > T(0) && !o->posAlpha) {
    set_error("trainDL: the ridge-only problem (lambda1 = 0, lambda2 > 0) is not supported");
    return SPX_ERR_UNSUP;
  }
  if (o->modeParam < 0 || o->modeParam > 2 || o->iter_updateD < 0 || *itercount < 0) {
    set_error("trainDL: invalid modeParam / iter_updateD / iteration count");
    return SPX_ERR_ARG;
  }
  if (*itercount > 0 && !D_given) {
    // the facade continues a model only with its dictionary (cpp_library/spams.h:1201-1210)
    set_error("trainDL: a model (itercount > 0) needs its dictionary (D_given)");
    return SPX_ERR_ARG;
  }
  if (m > kDlBlockThreads * kDlMaxRowsPerThread) {
    set_error("trainDL: signal length m=%d exceeds %d", m, kDlBlockThreads * kDlMaxRowsPerThread);
    return SPX_ERR_UNSUP;
  }
  if ((long long)K * K > INT_MAX) {
    set_error("trainDL: K=%d too large", K);
    return SPX_ERR_UNSUP;
  }
  // L of Trainer::train (dicts.h:386-387); SPARSITY takes lambda1 atoms,
  // at most min(m, K) (a further atom's residual norm is zero)
  const int Lmk = m < K ? m : K;
  int L = mode == 3 ? (int)o->lambda1 : Lmk;
  if (L > Lmk) L = Lmk;
  if (L <= 0) {
    set_error("trainDL: SPARSITY mode needs lambda1 >= 1 atoms");
    return SPX_ERR_ARG;
  }
  int threads = o->num_threads > 0 ? o->num_threads : (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (threads < 1) threads = 1;
  const int batchsize = o->batchsize > 0 ? o->batchsize : 256 * (threads + 1);
  const int bs = o->batch ? M : (batchsize < M ? batchsize : M);
  const int itc = *itercount;
  const bool online = !o->batch;
  const int prec = o->precision & SPX_PREC_BF16X3;
  DlDevice g(ctx->device);
  hipStream_t st = ctx->stream;

  Rng rng;
  DlMem mem;
  const size_t nA = (size_t)K * K, nB = (size_t)m * K;
  const int ldg = pad_ld(K);
  const int CH = bs < kDlChunk ? bs : kDlChunk;
  const int ldt = (CH + 3) & ~3;
  T *G, *Aorig, *Aodd, *Aeven, *Borig, *Bodd, *Beven, *Xb, *XbT, *aT, *DtX, *norms, *P, *lv,
      *scal_dev;
  T *Abat, *Bbat, *Apart, *Bpart;  // batch statistics and their split-K partial products
  int *perm, *lr, *lc, *idx, *hit, *Ldev;
  DL_TRY(mem.get(&G, (size_t)ldg * K));
  DL_TRY(mem.get(&Aorig, nA));
  DL_TRY(mem.get(&Aodd, nA));
  DL_TRY(mem.get(&Aeven, nA));
  DL_TRY(mem.get(&Abat, nA));
  DL_TRY(mem.get(&Borig, nB));
  DL_TRY(mem.get(&Bodd, nB));
  DL_TRY(mem.get(&Beven, nB));
  DL_TRY(mem.get(&Bbat, nB));
  // A_t (K x K) and B_t (m x K) have inner dimension = the chunk's signals;
  // their few output tiles are split over it to fill the chip
  const int SA = gemm_split_count(K, K, ldt), SB = gemm_split_count(m, K, ldt);
  DL_TRY(mem.get(&Apart, nA * SA));
  DL_TRY(mem.get(&Bpart, nB * SB));
  DL_TRY(mem.get(&Xb, (size_t)m * CH));
  DL_TRY(mem.get(&XbT, (size_t)ldt * m));
  DL_TRY(mem.get(&aT, (size_t)ldt * K));
  DL_TRY(mem.get(&DtX, (size_t)ldg * CH));
  DL_TRY(mem.get(&norms, (size_t)CH));
  DL_TRY(mem.get(&lv, (size_t)L * CH));
  DL_TRY(mem.get(&lr, (size_t)L * CH));
  DL_TRY(mem.get(&lc, (size_t)CH));
  DL_TRY(mem.get(&perm, (size_t)M));
  DL_TRY(mem.get(&idx, (size_t)K));
  DL_TRY(mem.get(&hit, 1));
  DL_TRY(mem.get(&scal_dev, 3));
  DL_TRY(mem.get(&Ldev, 1));
  const int bA = dl_block_atoms<T>(m);
  SPX_HIP(hipFuncSetAttribute((const void*)dl_block_kernel<T>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kDlBlockLds));
  DL_TRY(mem.get(&P, (size_t)m * bA * ((K + dl_panel_chunk(K) - 1) / dl_panel_chunk(K) + 1)));
  SPX_HIP(hipMemsetAsync(G, 0, sizeof(T) * (size_t)ldg * K, st));
  SPX_HIP(hipMemsetAsync(DtX, 0, sizeof(T) * (size_t)ldg * CH, st));
  // solver scalars: lambda (LARS constraint), eps and lambda (ORMP), L; the
  // host copies stay alive until the stream synchronises below
  const T sc[3] = {lambda1, mode == 4 ? lambda1 : T(0), mode == 5 ? lambda1 : T(0)};
  SPX_HIP(hipMemcpyAsync(scal_dev, sc, sizeof sc, hipMemcpyHostToDevice, st));
  SPX_HIP(hipMemcpyAsync(Ldev, &L, sizeof(int), hipMemcpyHostToDevice, st));

  // initial dictionary: K random signals (dicts.h:400-411)
  if (!D_given) {
    std::vector<int> h(K);
    for (int k = 0; k < K; ++k) h[k] = (int)(rng.next() % M);
    SPX_HIP(hipMemcpyAsync(idx, h.data(), sizeof(int) * K, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(dl_copy_cols_kernel<T>, dim3((m + 255) / 256, K), dim3(256), 0, st, X,
                       ldx, m, static_cast<const int*>(idx), K, D);
    SPX_HIP(hipGetLastError());
    SPX_HIP(hipStreamSynchronize(st));  // h leaves scope
  }
  // thrsPos when posD, then Matrix::normalize (dicts.h:424-427)
  hipLaunchKernelGGL(dl_normalize_kernel<T>, dim3((K + 3) / 4), dim3(256), 0, st, D, m, K,
                     static_cast<const int*>(nullptr), o->posD, T(1e-10));
  SPX_HIP(hipGetLastError());
  if (itc == 0) {
    const long long ne = (long long)(nA > nB ? nA : nB);
    hipLaunchKernelGGL(dl_init_model_kernel<T>, dim3(blocks_for(ne, 256)), dim3(256), 0, st, A,
                       B, static_cast<const T*>(D), m, K, static_cast<T>(std::fabs(o->t0)),
                       online ? 1 : 0);
    SPX_HIP(hipGetLastError());
  }
  // srandom(0); perm.randperm(M) (dicts.h:459-461, linalg/linalg.h:2906-2918)
  {
    rng.seed();
    std::vector<int> table(M), h(M);
    for (int i = 0; i < M; ++i) table[i] = i;
    int size = M;
    for (int i = 0; i < M; ++i) {
      const int pick = (int)(rng.next() % size);
      h[i] = table[pick];
      table[pick] = table[size - 1];
      --size;
    }
    SPX_HIP(hipMemcpyAsync(perm, h.data(), sizeof(int) * M, hipMemcpyHostToDevice, st));
    SPX_HIP(hipStreamSynchronize(st));
  }
  SPX_HIP(hipMemsetAsync(Aodd, 0, sizeof(T) * nA, st));
  SPX_HIP(hipMemsetAsync(Aeven, 0, sizeof(T) * nA, st));
  SPX_HIP(hipMemsetAsync(Bodd, 0, sizeof(T) * nB, st));
  SPX_HIP(hipMemsetAsync(Beven, 0, sizeof(T) * nB, st));
  SPX_HIP(hipMemcpyAsync(Aorig, A, sizeof(T) * nA, hipMemcpyDeviceToDevice, st));
  SPX_HIP(hipMemcpyAsync(Borig, B, sizeof(T) * nB, hipMemcpyDeviceToDevice, st));

  // solver set-up over the trainer's Gram (ridge stored in G, so no shift)
  spx_dict dict;  // view for the width rule; owns nothing
  dict.m = m;
  dict.p = K;
  dict.ldg = ldg;
  dict.elem = sizeof(T);
  dict.G = G;
  const int waves = o->waves_per_signal > 0 ? o->waves_per_signal : tuned_waves(&dict, mode < 3 ? 0 : 1);

  // iter_updateD sweeps of the dictionary update, captured once per
  // zero_unused flag as a hipGraph and replayed (or launched one by one
  // when eager_update)
  DlGraphs graphs;
  auto update = [&](int zu) -> int {
    if (o->iter_updateD <= 0) return SPX_OK;
    if (o->eager_update) {
      for (int s = 0; s < o->iter_updateD; ++s)
        DL_TRY(dl_update_dict<T>(st, D, m, K, A, B, P, bA, o->posD ? 1 : 0, zu));
      return SPX_OK;
    }
    if (!graphs.exec[zu]) {
      if (!ctx->capture_stream)
        SPX_HIP(hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
      hipStream_t cs = ctx->capture_stream;
      SPX_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      int rc = SPX_OK;
      for (int s = 0; s < o->iter_updateD && !rc; ++s)
        rc = dl_update_dict<T>(cs, D, m, K, A, B, P, bA, o->posD ? 1 : 0, zu);
      hipGraph_t graph = nullptr;
      const hipError_t ec = hipStreamEndCapture(cs, &graph);
      if (rc || ec != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        if (!rc) set_error("trainDL: graph capture failed: %s", hipGetErrorString(ec));
        return rc ? rc : SPX_ERR_HIP;
      }
      const hipError_t ei = hipGraphInstantiate(&graphs.exec[zu], graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        graphs.exec[zu] = nullptr;
        set_error("trainDL: graph instantiation failed: %s", hipGetErrorString(ei));
        return SPX_ERR_HIP;
      }
    }
    SPX_HIP(hipGraphLaunch(graphs.exec[zu], st));
    return SPX_OK;
  };

  const long long JJ = o->iter < 0 ? 100000000LL : o->iter;
  const double t_start = wall_seconds();
  const T ridge = static_cast<T>(o->lambda2 > 1e-10 ? o->lambda2 : 1e-10);
  bool even = true;
  int count = 0, countPrev = 0;
  long long i;
  for (i = 0; i < JJ; ++i) {
    if (o->iter < 0) {
      SPX_HIP(hipStreamSynchronize(st));
      if (wall_seconds() - t_start > -(double)o->iter) break;
    }
    // G = D'D
    DL_TRY(gemm_tn<T>(st, K, K, m, D, m, D, m, G, ldg, T(1), T(0), nullptr, 0, prec));
    if (o->clean) {
      int start = 0;
      for (;;) {
        const int none = INT_MAX;
        SPX_HIP(hipMemcpyAsync(hit, &none, sizeof(int), hipMemcpyHostToDevice, st));
        const long long rem = (long long)K * K - start;
        unsigned nbk = blocks_for(rem, 256);
        if (nbk > 4096) nbk = 4096;
        hipLaunchKernelGGL(dl_clean_find_kernel<T>, dim3(nbk), dim3(256), 0, st,
                           static_cast<const T*>(G), ldg, K, start, hit);
        SPX_HIP(hipGetLastError());
        int h = INT_MAX;
        SPX_HIP(hipMemcpyAsync(&h, hit, sizeof(int), hipMemcpyDeviceToHost, st));
        SPX_HIP(hipStreamSynchronize(st));
        if (h == INT_MAX) break;
        const int j = h % K;
        const int ind = (int)(rng.next() % M);
        SPX_HIP(hipMemcpyAsync(idx, &ind, sizeof(int), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(dl_copy_cols_kernel<T>, dim3((m + 255) / 256, 1), dim3(256), 0, st,
                           X, ldx, m, static_cast<const int*>(idx), 1, D + (size_t)j * m);
        SPX_HIP(hipGetLastError());
        hipLaunchKernelGGL(dl_normalize_kernel<T>, dim3(1), dim3(256), 0, st, D + (size_t)j * m,
                           m, 1, static_cast<const int*>(nullptr), o->posD, T(10e-10));
        SPX_HIP(hipGetLastError());
        hipLaunchKernelGGL(dl_gram_col_kernel<T>, dim3((K + 3) / 4), dim3(256), 0, st,
                           static_cast<const T*>(D), m, K, j, G, ldg);
        SPX_HIP(hipGetLastError());
        SPX_HIP(hipStreamSynchronize(st));  // ind leaves scope
        start = h + 1;
        if (start >= K * K) break;
      }
    }
    hipLaunchKernelGGL(dl_ridge_kernel<T>, dim3((K + 255) / 256), dim3(256), 0, st, G, ldg, K,
                       ridge);
    SPX_HIP(hipGetLastError());
    // code the minibatch in chunks; A_t, B_t accumulate over the chunks
    for (int c0 = 0; c0 < bs; c0 += CH) {
      const int nb = bs - c0 < CH ? bs - c0 : CH;
      const int ld = (nb + 3) & ~3;
      const long long base = i * (long long)bs + c0;
      hipLaunchKernelGGL(dl_gather_kernel<T>, dim3((ld + 63) / 64, (m + 63) / 64), dim3(256), 0,
                         st, X, ldx, m, M, static_cast<const int*>(perm), base, nb, Xb, XbT,
                         ld);
      SPX_HIP(hipGetLastError());
      DL_TRY(gemm_tn<T>(st, K, nb, m, D, m, Xb, m, DtX, ldg, T(1), T(0), nullptr, 0, prec));
      DL_TRY(col_sqnorms<T>(st, m, nb, Xb, m, norms));
      if (mode < 3) {
        LarsArgs<T> a{};
        a.tol = T(1e-15);
        a.G = G;
        a.ldg = ldg;
        a.diag = T(0);
        a.DtX = DtX;
        a.ldd = ldg;
        a.norms = norms;
        a.lambda = scal_dev;
        a.nlam = 1;
        a.p = K;
        a.n = nb;
        a.L = L;
        a.mode = mode;
        a.pos = o->posAlpha ? 1 : 0;
        a.length_path = 4 * L;  // coreLARS2's default cap (decomp/decomp.h:1724)
        a.outV = lv;
        a.outR = lr;
        a.outCount = lc;
        DL_TRY(lars_solve<T>(ctx, a, SPX_FORM_BATCHED, waves));
      } else {
        OmpArgs<T> a{};
        a.G = G;
        a.ldg = ldg;
        a.DtX = DtX;
        a.ldd = ldg;
        a.norms = norms;
        a.Ls = Ldev;
        a.nL = 1;
        a.eps = scal_dev + 1;
        a.nEps = 1;
        a.lambda = scal_dev + 2;
        a.nLambda = 1;
        a.scalarL = Lmk;
        a.p = K;
        a.n = nb;
        a.L = L;
        a.outV = lv;
        a.outR = lr;
        a.outCount = lc;
        DL_TRY(omp_solve<T>(ctx, a, SPX_FORM_BATCHED, waves));
      }
      SPX_HIP(hipMemsetAsync(aT, 0, sizeof(T) * (size_t)ld * K, st));
      hipLaunchKernelGGL(dl_scatter_kernel<T>, dim3((nb + 255) / 256), dim3(256), 0, st,
                         static_cast<const T*>(lv), static_cast<const int*>(lr),
                         static_cast<const int*>(lc), L, nb, aT, ld);
      SPX_HIP(hipGetLastError());
      const int acc = c0 ? 1 : 0;
      DL_TRY(gemm_tn_split<T>(st, K, K, ld, aT, ld, aT, ld, Abat, acc, Apart,
                              gemm_split_count(K, K, ld), prec));
      DL_TRY(gemm_tn_split<T>(st, m, K, ld, XbT, ld, aT, ld, Bbat, acc, Bpart,
                              gemm_split_count(m, K, ld), prec));
    }
    hipLaunchKernelGGL(dl_mirror_kernel<T>, dim3(blocks_for((long long)nA, 256)), dim3(256), 0,
                       st, Abat, K);
    SPX_HIP(hipGetLastError());
    int zero_unused;
    if (!online) {
      SPX_HIP(hipMemcpyAsync(A, Abat, sizeof(T) * nA, hipMemcpyDeviceToDevice, st));
      SPX_HIP(hipMemcpyAsync(B, Bbat, sizeof(T) * nB, hipMemcpyDeviceToDevice, st));
      zero_unused = o->clean ? 1 : 0;
    } else {
      const int epoch = (int)((((i + 1) % M) * (long long)bs) / M);
      int swap = 0;
      if ((even && (epoch % 2) == 1) || (!even && (epoch % 2) == 0)) {
        swap = 1;
        count = countPrev;
        countPrev = 0;
        even = !even;
      }
      const long long ii = itc + i;
      long long num_elem = ii < bs ? ii * bs : (long long)bs * bs + ii - bs;
      if (num_elem > 2LL * M) num_elem = 2LL * M;
      const T scal2 = T(T(1.0) / bs);
      const int totaliter = itc + count;
      T scal;
      if (o->modeParam == 2) {
        scal = static_cast<T>(o->rho);
      } else if (o->modeParam == 1) {
        const T pw = static_cast<T>(
            std::pow((double)(T(totaliter) / T(totaliter + 1)), -(double)static_cast<T>(o->rho)));
        scal = static_cast<T>(0.95 > (double)pw ? 0.95 : (double)pw);
      } else {
        scal = T(itc + num_elem + 1 - bs) / T(itc + num_elem + 1);
      }
      const int use_orig = (itc > 0 && i * bs < M) || (itc == 0 && o->t0 != 0 && i * bs < 10000);
      hipLaunchKernelGGL(dl_combine_kernel<T>, dim3(blocks_for((long long)nA, 256)), dim3(256), 0,
                         st, (long long)nA, A, Aorig, Aodd, Aeven, static_cast<const T*>(Abat),
                         swap, scal, use_orig, scal2);
      SPX_HIP(hipGetLastError());
      hipLaunchKernelGGL(dl_combine_kernel<T>, dim3(blocks_for((long long)nB, 256)), dim3(256), 0,
                         st, (long long)nB, B, Borig, Bodd, Beven, static_cast<const T*>(Bbat),
                         swap, scal, use_orig, scal2);
      SPX_HIP(hipGetLastError());
      ++count;
      ++countPrev;
      zero_unused = (o->clean && (itc + i) * (long long)bs > 10000) ? 1 : 0;
    }
    DL_TRY(update(zero_unused));
  }
  SPX_HIP(hipStreamSynchronize(st));
  *itercount = (int)(itc + i);
  return SPX_OK;
}

// Host buffers: copies in, trains on the device, copies D (and A, B) out.
template <typename T>
int train_dl_host(spx_ctx* ctx, const T* X, int m, int M, T* D, int K, int D_given, T* A,
                  T* B, int* itercount, const spx_train_dl_opts* o) {
  if (!ctx || !X || !D || !itercount || !o || m <= 0 || M <= 0 || K <= 0) {
    set_error("trainDL: invalid arguments (m=%d, n=%d, K=%d)", m, M, K);
    return SPX_ERR_ARG;
  }
  if (*itercount > 0 && (!A || !B)) {
    set_error("trainDL: a model (itercount > 0) needs A and B");
    return SPX_ERR_ARG;
  }
  DlDevice g(ctx->device);
  hipStream_t st = ctx->stream;
  DlMem mem;
  T *Xd, *Dd, *Ad, *Bd;
  const size_t nA = (size_t)K * K, nB = (size_t)m * K;
  DL_TRY(mem.get(&Xd, (size_t)m * M));
  DL_TRY(mem.get(&Dd, nB));
  DL_TRY(mem.get(&Ad, nA));
  DL_TRY(mem.get(&Bd, nB));
  SPX_HIP(hipMemcpyAsync(Xd, X, sizeof(T) * (size_t)m * M, hipMemcpyHostToDevice, st));
  if (D_given) SPX_HIP(hipMemcpyAsync(Dd, D, sizeof(T) * nB, hipMemcpyHostToDevice, st));
  if (*itercount > 0) {
    SPX_HIP(hipMemcpyAsync(Ad, A, sizeof(T) * nA, hipMemcpyHostToDevice, st));
    SPX_HIP(hipMemcpyAsync(Bd, B, sizeof(T) * nB, hipMemcpyHostToDevice, st));
  }
  DL_TRY(train_dl_dev<T>(ctx, Xd, m, m, M, Dd, K, D_given, Ad, Bd, itercount, o));
  SPX_HIP(hipMemcpyAsync(D, Dd, sizeof(T) * nB, hipMemcpyDeviceToHost, st));
  if (A) SPX_HIP(hipMemcpyAsync(A, Ad, sizeof(T) * nA, hipMemcpyDeviceToHost, st));
  if (B) SPX_HIP(hipMemcpyAsync(B, Bd, sizeof(T) * nB, hipMemcpyDeviceToHost, st));
  SPX_HIP(hipStreamSynchronize(st));
  return SPX_OK;
}

}  // namespace

}  // namespace spx

extern "C" {

int spx_train_dl_f32(spx_ctx* ctx, const float* X, int m, int n, float* D, int K, int D_given,
                     float* A, float* B, int* itercount, const spx_train_dl_opts* opts) {
  return spx::train_dl_host<float>(ctx, X, m, n, D, K, D_given, A, B, itercount, opts);
}

int spx_train_dl_f64(spx_ctx* ctx, const double* X, int m, int n, double* D, int K,
                     int D_given, double* A, double* B, int* itercount,
                     const spx_train_dl_opts* opts) {
  return spx::train_dl_host<double>(ctx, X, m, n, D, K, D_given, A, B, itercount, opts);
}

int spx_train_dl_dev_f32(spx_ctx* ctx, const float* X, int ldx, int m, int n, float* D, int K,
                         int D_given, float* A, float* B, int* itercount,
                         const spx_train_dl_opts* opts) {
  return spx::train_dl_dev<float>(ctx, X, ldx, m, n, D, K, D_given, A, B, itercount, opts);
}

int spx_train_dl_dev_f64(spx_ctx* ctx, const double* X, int ldx, int m, int n, double* D,
                         int K, int D_given, double* A, double* B, int* itercount,
                         const spx_train_dl_opts* opts) {
  return spx::train_dl_dev<double>(ctx, X, ldx, m, n, D, K, D_given, A, B, itercount, opts);
}

}  // extern "C"
