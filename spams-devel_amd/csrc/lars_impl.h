#pragma once
// LARS-Lasso solver kernel: coreLARS2 semantics (decomp/decomp.h:1706-1915).
//
// One group of W wavefronts (one workgroup) solves one signal at a time:
//   - the correlation vector DtR (p) lives in LDS, the active-atom flags in a
//     bit mask of the thread that owns the rows;
//   - the correlation change A = G(:, active) * u is accumulated
//     in registers (each thread owns NQ chunks of VEC = 16 bytes of rows),
//     streaming the active Gram columns from the resident G (L2 / MALL),
//     several columns in flight per round;
//   - the explicit inverse of the active Gram block, invGs (s x s, stored full
//     and exactly symmetric), is double buffered for the downdate;
//   - arg-min / arg-max searches are deterministic (value, index) wavefront
//     reductions with the reference's tie rules (first index for i?amax and
//     the custom iamin; last index for the zero-crossing search).
//
// The kernel is instantiated with an active-set capacity CAP:
//   CAP > 0  the active-set arrays and invGs live in LDS (capacity CAP); a
//            signal whose active set would outgrow CAP is appended to an
//            overflow list, with its state (DtR, coefficients, invGs,
//            counters) saved in the signal's hand-over slot;
//   CAP == 0 capacity L: active-set arrays in LDS, invGs in a per-workgroup
//            HBM slot; listed signals resume from their hand-over state.
// A signal's arithmetic depends only on the two group widths, never on the
// batch around it.
//
// Batched form: persistent grids pulling signal indices from atomic counters
// (fast pass over 0..n-1, then the full pass over the overflow list, on the
// same stream).  Plain form: a grid of resident workgroups, each taking
// signals blockIdx.x + k * gridDim.x in turn, full capacity.
#include <climits>
#include <type_traits>

#include "mask_gram.h"
#include "spx_common.h"
#include "spx_internal.h"

namespace spx {

namespace {

template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int N = 4;
};
template <>
struct Vec16<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int N = 2;
};

// LDS capacity of the fast pass (active-set size it solves without HBM).
template <typename T>
struct FastCap {
  static constexpr int C = sizeof(T) == 4 ? 24 : 16;
};

// State of a signal handed from the LDS pass to the full-capacity pass,
// taken just before the admission that would outgrow the LDS capacity C:
// DtR (ldg), coefficients (C), the inverse Gram block (C x C, leading
// dimension C), normX and thrs, the active indices (C) and the counters
// (iterations, admissions, active-set size, atom to admit).  The full pass
// resumes from it instead of restarting the path.
template <typename T>
struct LarsState {
  static constexpr int C = FastCap<T>::C;
  static size_t bytes(int ldg) {
    const size_t b = sizeof(T) * ((size_t)ldg + C + (size_t)C * C + 2) + sizeof(int) * (C + 4);
    return (b + 15) & ~size_t(15);
  }
  T* dtr;
  T* coef;
  T* M;
  T* sc;
  int* ind;
  int* isc;
  __device__ LarsState(void* base, int ldg) {
    dtr = static_cast<T*>(base);
    coef = dtr + ldg;
    M = coef + C;
    sc = M + C * C;
    ind = reinterpret_cast<int*>(sc + 2);
    isc = ind + C;
  }
};

template <typename T, int W, int NQ, int CAP>
struct LarsLds {
  static constexpr int GT = 64 * W;
  static constexpr int VEC = Vec16<T>::N;
  static constexpr int PT = VEC * GT * NQ;
  // byte size for active-set capacity LC; variant 1 (weighted) adds the
  // per-row weights, variant 2 (masked) the Gram-column slot map
  static size_t bytes(int LC, int CL = 0, int var = 0) {
    const size_t LCA = ((size_t)LC + 3) & ~size_t(3);  // array stride (16-byte multiple)
    size_t b = sizeof(T) * (PT + GT + 5 * LCA) + sizeof(int) * 2 * LCA;
    b = (b + 15) & ~size_t(15);
    b += sizeof(KeyIdx<T>) * W + sizeof(T) * 3 * W + 16;
    b += sizeof(unsigned long long) * 2 * W + sizeof(T) * 4 * W + 16;  // parity buffers
    b = (b + 15) & ~size_t(15);
    if (CAP > 0) b += 2 * sizeof(T) * CAP * CAP + 16;
    b = (b + 15) & ~size_t(15);
    // FP64 refit vectors (the LDS pass keeps them in its idle inverse buffer)
    if (CAP == 0) b += sizeof(double) * (size_t)(LC + (LC > 64 ? LC : 64)) + 16;
    b = (b + 15) & ~size_t(15);
    b += sizeof(T) * (size_t)CL * CL + 16;  // LDS copy of the inverse (full pass)
    b = (b + 15) & ~size_t(15);
    if (var == 1) b += sizeof(T) * PT + 16;  // weights of the group's rows
    if (var == 2) b += sizeof(int) * 2 * (size_t)LC + 16;  // slot map + shift buffer
    return b;
  }
};

// y(0:n) = M(0:n, 0:n) x for a symmetric M (leading dimension ldm), with
// all GT threads; DOT: also returns this thread's part of y'x (its rows, in
// order), and a group of several waves leaves the closing barrier to the
// caller's reduction of that sum.  One thread per row, the sum over q in order, eight loads
// in flight; when M is in HBM and has fewer rows than threads (SPLIT), row r
// is split into P = GT / n parts whose partial sums are combined in part
// order (deterministic).  x and y in LDS, part a GT-entry LDS buffer.  Ends
// with a barrier.
template <typename T, int GT, bool SPLIT, bool DOT = false>
__device__ __forceinline__ T symv_dist(const T* M, int ldm, const T* x, T* y, T* part, int n) {
  const int tid = threadIdx.x;
  const int P = (!SPLIT || n >= GT) ? 1 : GT / n;
  T dot = 0;
  if (P == 1) {
    for (int r = tid; r < n; r += GT) {
      T acc = 0;
      int q = 0;
      for (; q + 8 <= n; q += 8) {
        T mv[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          mv[u] = M[r + (q + u) * ldm];
          xv[u] = x[q + u];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += mv[u] * xv[u];
      }
      for (; q < n; ++q) acc += M[r + q * ldm] * x[q];
      y[r] = acc;
      if (DOT) dot += acc * x[r];
    }
    if (!DOT || GT == 64) __syncthreads();
    return dot;
  }
  // n * P <= GT: one (row, part) per thread
  const int seg = (n + P - 1) / P;
  if (tid < n * P) {
    const int k = tid / n, r = tid - k * n;
    const int q0 = k * seg, q1 = q0 + seg < n ? q0 + seg : n;
    T acc = 0;
    int q = q0;
    for (; q + 8 <= q1; q += 8) {
      T mv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) mv[u] = M[r + (q + u) * ldm];
#pragma unroll
      for (int u = 0; u < 8; ++u) acc += mv[u] * x[q + u];
    }
    for (; q < q1; ++q) acc += M[r + q * ldm] * x[q];
    part[k * n + r] = acc;
  }
  __syncthreads();
  for (int r = tid; r < n; r += GT) {
    T acc = part[r];
    for (int k = 1; k < P; ++k) acc += part[k * n + r];
    y[r] = acc;
    if (DOT) dot += acc * x[r];
  }
  if (!DOT || GT == 64) __syncthreads();
  return dot;
}

// Admits the i-th active atom (i >= 1) into the explicit inverse M of the
// active Gram block, i x i -> (i+1) x (i+1), by its Schur complement
// (decomp/decomp.h:1756-1770); sGs(0:i+1) = G(active, new atom).
// v = M gs lands in sV.  An FP32 pivot below 1e-3 of the diagonal (or not
// positive) is recomputed with FP64 accumulation.  Elements (r,q) and (q,r)
// get the identical update v[min] * (schur * v[max]), so M stays exactly
// symmetric.  The same pass forms the direction u(0:i+1) = M_new w from the
// rows it updates (SPLIT: the row segments of symv_dist), so M is read once
// more instead of twice.  Returns true (u formed).  Ends with a barrier.
template <typename T, int W, bool SPLIT>
__device__ __forceinline__ bool inverse_admit(T* M, int ldm, int i, const T* sGs, T* sV,
                                              T* sPart, T* sRedS, int& par, double* sRes,
                                              const T* w, T* u) {
  constexpr int GT = 64 * W;
  const int tid = threadIdx.x;
  // v = invGs(i x i) * gs(0:i) and gs'v; the reduction's barrier publishes v
  T dv = symv_dist<T, GT, SPLIT, true>(M, ldm, sGs, sV, sPart, i);
  dv = group_sum1<T, W>(dv, sRedS, par);
  const T gii = uniform_v(sGs[i]);
  T schur = T(1) / (gii - dv);
  if (sizeof(T) == 4 && !(gii - dv > T(1e-3) * gii)) {
    __syncthreads();
    double pd = 0;
    for (int r = tid; r < i; r += GT) {
      double acc = 0;
      for (int q = 0; q < i; ++q)
        acc += static_cast<double>(M[r + q * ldm]) * static_cast<double>(sGs[q]);
      sV[r] = static_cast<T>(acc);
      pd += acc * static_cast<double>(sGs[r]);
    }
    double z1d = 0, z2d = 0;
    group_sum3<double, W>(pd, z1d, z2d, sRes);
    schur = static_cast<T>(1.0 / (static_cast<double>(gii) - pd));
    __syncthreads();
  }
  // Rank-one update, border and direction in one pass: thread (r, k) takes
  // row r of the new (i+1) x (i+1) block over the columns of segment k, the
  // segments symv_dist uses for that size, and sums M_new(r, q) w(q) over
  // them in order; the part sums are combined in part order, so u comes out
  // exactly as symv_dist would form it from M_new.  Each element is written
  // once, by the thread of its row and segment.
  const int n = i + 1;
  const int P = (!SPLIT || n >= GT) ? 1 : GT / n;
  const int seg = (n + P - 1) / P;
  for (int t = tid; t < n * P; t += GT) {
    const int k = P == 1 ? 0 : t / n, r = t - k * n;
    const int q0 = k * seg, q1 = q0 + seg < n ? q0 + seg : n;
    const int qe = q1 < i ? q1 : i;      // columns of the old block
    const bool hasi = q0 <= i && i < q1;  // the segment holds column i
    T acc = 0;
    if (r < i) {
      const T vr = sV[r];
      int q = q0;
      for (; q + 8 <= qe; q += 8) {
        T mv[8], vq[8], wq[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          mv[e] = M[r + (q + e) * ldm];
          vq[e] = sV[q + e];
          wq[e] = w[q + e];
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const T upd = r <= q + e ? vr * (schur * vq[e]) : vq[e] * (schur * vr);
          const T mn = mv[e] + upd;
          M[r + (q + e) * ldm] = mn;
          acc += mn * wq[e];
        }
      }
      for (; q < qe; ++q) {
        const T vq = sV[q];
        const T upd = r <= q ? vr * (schur * vq) : vq * (schur * vr);
        const T mn = M[r + q * ldm] + upd;
        M[r + q * ldm] = mn;
        acc += mn * w[q];
      }
      if (hasi) {
        // the border column
        const T off = -schur * vr;
        M[r + i * ldm] = off;
        acc += off * w[i];
      }
    } else {
      // the border row
      for (int q = q0; q < qe; ++q) {
        const T off = -schur * sV[q];
        M[i + q * ldm] = off;
        acc += off * w[q];
      }
      if (hasi) {
        M[i + i * ldm] = schur;
        acc += schur * w[i];
      }
    }
    if (P == 1) u[r] = acc;
    else sPart[k * n + r] = acc;
  }
  __syncthreads();
  if (P > 1) {
    for (int r = tid; r < n; r += GT) {
      T acc = sPart[r];
      for (int k = 1; k < P; ++k) acc += sPart[k * n + r];
      u[r] = acc;
    }
    __syncthreads();
  }
  return true;
}

// Md (i x i) = Ms ((i+1) x (i+1)) without row and column fz, downdated by
// -1/Ms(fz,fz) times the outer product of that column
// (decomp/decomp.h:1875-1903).  Ms and Md are distinct buffers.  No barrier
// after the writes.
template <typename T, int W, bool ELEM>
__device__ __forceinline__ void inverse_drop(const T* Ms, int lds, T* Md, int ldd, int i, int fz,
                                             T* sV) {
  constexpr int GT = 64 * W;
  const int tid = threadIdx.x;
  const T schur = Ms[fz + fz * lds];
  for (int r = tid; r < i; r += GT) {
    const int r2 = r + (r >= fz);
    sV[r] = Ms[r2 + fz * lds];
  }
  __syncthreads();
  const T alpha = T(-1) / schur;
  if (ELEM && i > 0) {
    // (r, q) advanced incrementally as in inverse_admit
    const int nn = i * i;
    const int dqG = GT / i, drG = GT - dqG * i;
    int r0 = tid % i, q0 = tid / i;
    for (int e0 = tid; e0 < nn; e0 += 8 * GT) {
      T mv[8];
      int rr[8], qq[8];
      int r = r0, q = q0;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        rr[u] = r;
        qq[u] = q;
        const int e = e0 + u * GT;
        mv[u] = e < nn ? Ms[r + (r >= fz) + (q + (q >= fz)) * lds] : T(0);
        r += drG;
        q += dqG;
        if (r >= i) {
          r -= i;
          ++q;
        }
      }
      r0 = r;
      q0 = q;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * GT;
        if (e < nn) {
          const int ru = rr[u], qu = qq[u];
          const T ur = sV[ru], uq = sV[qu];
          const T upd = ru <= qu ? ur * (alpha * uq) : uq * (alpha * ur);
          Md[ru + qu * ldd] = mv[u] + upd;
        }
      }
    }
  } else {
    for (int r = tid; r < i; r += GT) {
      const int r2 = r + (r >= fz);
      const T ur = sV[r];
      for (int q = 0; q < i; ++q) {
        const int q2 = q + (q >= fz);
        const T uq = sV[q];
        const T upd = r <= q ? ur * (alpha * uq) : uq * (alpha * ur);
        Md[r + q * ldd] = Ms[r2 + q2 * lds] + upd;
      }
    }
  }
}

// Squared residual of the least-squares fit on the active set,
// normX0 - b' M b with b = D'x over the s active atoms and M the inverse of
// their Gram block.  coreLARS with ols tracks this quantity in
// place of the homotopy residual (decomp/decomp.h:1044-1050 on admission,
// :1191-1201 after a drop); FP64 accumulation.  Every thread returns it.
template <typename T, int W>
__device__ __forceinline__ T ls_residual(const T* Mf, int ldf, int s, const T* dtx,
                                         const int* sInd, T normX0, double* sRes) {
  constexpr int GT = 64 * W;
  double acc = 0;
  for (int r = threadIdx.x; r < s; r += GT) {
    double row = 0;
    for (int q = 0; q < s; ++q)
      row += static_cast<double>(Mf[r + q * ldf]) * static_cast<double>(dtx[sInd[q]]);
    acc += row * static_cast<double>(dtx[sInd[r]]);
  }
  double z1 = 0, z2 = 0;
  group_sum3<double, W>(acc, z1, z2, sRes);
  return static_cast<T>(static_cast<double>(normX0) - acc);
}

// The full-capacity pass at 4 wavefronts is held to 4 workgroups per CU
// (128 VGPRs): its signals are latency-bound, so occupancy pays more than
// registers.  The LDS pass at 1 or 2 wavefronts with up to 4 chunks per
// lane is held to 4 waves per SIMD (128 VGPRs, a few 64-bit values spilled
// outside the step loop) for the same reason.
// VAR: 0 plain penalty, 1 weighted penalty (a.weights), 2 masked signals
// (a.mobs).
template <typename T, int W, int NQ, int CAP, int VAR>
__global__ __launch_bounds__(64 * W, (CAP == 0 && W == 4) || (CAP > 0 && W <= 2 && NQ <= 4) ? 4 : 1) void lars_kernel(LarsArgs<T> a) {
  constexpr int GT = 64 * W;
  constexpr int VEC = Vec16<T>::N;
  constexpr int PT = VEC * GT * NQ;  // rows covered by the group
  constexpr int UNR = NQ >= 4 ? 2 : (NQ == 2 ? 4 : 8);  // Gram columns in flight
  typedef typename Vec16<T>::type V;
  typedef typename std::conditional<(NQ * VEC <= 32), unsigned, unsigned long long>::type ActMask;
  const T INF = dev_inf<T>();

  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int L = a.L;
  const int LC = CAP > 0 ? (CAP < L ? CAP : L) : L;  // active-set capacity
  // LDS layout capacity: CAP in the LDS pass (a compile-time layout, so the
  // arrays below sit at constant offsets), L in the full-capacity pass
  const int LCL = CAP > 0 ? CAP : L;
  const int p = a.p;
  const int ldg = a.ldg;
  T* sDtR = reinterpret_cast<T*>(smem);  // PT
  T* sPart = sDtR + PT;                   // GT partial sums
  T* sCoef = sPart + GT;                  // LC
  // the per-position arrays are LCA apart (a multiple of 4 entries), so
  // runs of 2 or 4 entries starting at an even / multiple-of-4 position are
  // read with one 8- / 16-byte LDS load
  const int LCA = (LCL + 3) & ~3;
  T* sU = sCoef + LCA;                    // LC
  T* sW = sU + LCA;                       // LC
  T* sGs = sW + LCA;                      // LC
  T* sV = sGs + LCA;                      // LC
  int* sInd = reinterpret_cast<int*>(sV + LCA);  // LC
  int* sTmpI = sInd + LCA;                       // LC
  KeyIdx<T>* sRed = lds_next16<KeyIdx<T>>(smem, sTmpI + LCA);
  T* sRed3 = reinterpret_cast<T*>(sRed + W);  // 3W
  // single-barrier reductions: 2 buffers of W packed keys, 2 of 2W sums
  unsigned long long* sRedU = lds_next16<unsigned long long>(smem, sRed3 + 3 * W);
  T* sRedS = reinterpret_cast<T*>(sRedU + 2 * W);  // 4W
  int* sSig = reinterpret_cast<int*>(sRedS + 4 * W);
  T* sM0 = lds_next16<T>(smem, sSig + 1);  // CAP x CAP
  T* sM1 = sM0 + (CAP > 0 ? CAP * CAP : 0);
  // FP64 refit vectors (alpha LC, residual / reduction scratch max(LC, 64)):
  // the full-capacity pass has them here; the LDS pass keeps them in its
  // idle inverse buffer, which no live data occupies outside a drop
  double* const sAlphaF =
      lds_next16<double>(smem, sM1 + (CAP > 0 ? CAP * CAP : 0));  // used when CAP == 0
  double* const sResF = sAlphaF + LCL;
  // full-capacity pass: LDS copy of the inverse while the active set has at
  // most CL atoms (CL x CL, leading dimension CL)
  const int CL = CAP == 0 ? a.lds_cap : 0;
  T* sMl = CAP > 0 ? reinterpret_cast<T*>(sAlphaF) : lds_next16<T>(smem, sResF + (LCL > 64 ? LCL : 64));
  // weighted penalty: the signal's weights of the group's rows (1 past p)
  T* sWt = lds_next16<T>(smem, sMl + CL * CL);
  constexpr bool wt = VAR == 1;  // weighted penalty
  constexpr bool mk = VAR == 2;  // masked signals
  // masked: active position -> Gram-column slot, and its shift buffer
  int* sSlot = lds_next16<int>(smem, sMl + CL * CL);
  int* sSlotT = sSlot + LCA;
  T* const gaw = mk ? a.ga + (size_t)blockIdx.x * a.ga_stride : nullptr;

  const int tid = threadIdx.x;
  int par = 0;  // buffer parity of the single-barrier reductions
  T* const M0 = CAP > 0 ? sM0 : a.scratch + (size_t)blockIdx.x * 2 * L * L;
  T* const M1 = CAP > 0 ? sM1 : M0 + (size_t)L * L;
  const int ldm = CAP > 0 ? CAP : L;
  const int Lm = L < p ? L : p;
  const int nwork = a.list ? *a.list_count : a.n;

  // work distribution: chunks of `chunk` consecutive work items per atomic
  // grab (one device-scope atomic per signal serialises on its address)
  const int chunk = a.chunk > 0 ? a.chunk : 1;
  // The LDS pass (many short signals) requests its next grab as soon as it
  // starts on the current one, so the atomic's latency overlaps the set-up
  // loads; the full-capacity pass (few long signals) grabs when free, so no
  // long signal waits behind another.
  constexpr bool PREFETCH = CAP > 0;
  int w = 0, wend = 0;
  int pre = 0;  // thread 0: the prefetched grab
  if (PREFETCH && a.queue && tid == 0) pre = atomicAdd(a.queue, chunk);
  for (int round = 0;; ++round) {
    if (a.queue) {
      if (w >= wend) {
        __syncthreads();
        if (tid == 0) *sSig = PREFETCH ? pre : atomicAdd(a.queue, chunk);
        __syncthreads();
        w = uniform_v(*sSig);
        wend = w + chunk;
        if (PREFETCH && tid == 0 && w < nwork) pre = atomicAdd(a.queue, chunk);
      }
    } else {
      // plain form: a static grid-stride assignment (no work queue, no order)
      const long wl = (long)blockIdx.x + (long)round * gridDim.x;
      w = wl < nwork ? (int)wl : nwork;
      wend = w + 1;
    }
    if (w >= nwork) break;
    const int wcur = w++;
    // Values the whole group shares (signal, lambda, norms, counters, the
    // step quantities) are moved to scalar registers where they come from
    // memory: the compiler then keeps the step loop's control flow uniform
    // (scalar branches and loop counters instead of exec-mask bookkeeping).
    const int sig = uniform_v(a.list ? a.list[wcur] : (a.order ? a.order[wcur] : wcur));

    // ---------------------------------------------------------- set-up
    __syncthreads();  // the previous signal's LDS state is no longer read
    const T* dtx = a.DtX + (size_t)sig * a.ldd;
    // the full-capacity pass resumes the signals handed over by the LDS pass
    const bool resume = CAP == 0 && a.state && a.list;
    const LarsState<T> st(resume ? static_cast<void*>(static_cast<char*>(a.state) +
                                                      (size_t)sig * a.state_bytes)
                                 : nullptr, ldg);
    const bool vec_ok = resume || ((a.ldd % VEC == 0) &&
                                   (reinterpret_cast<uintptr_t>(a.DtX) % (sizeof(T) * VEC) == 0));
    const T* src = resume ? st.dtr : dtx;
    const int nsrc = resume ? ldg : p;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int j0 = VEC * (tid + GT * q);
      V v;
      if (vec_ok && j0 + VEC <= nsrc) {
        v = *reinterpret_cast<const V*>(src + j0);
      } else {
#pragma unroll
        for (int e = 0; e < VEC; ++e) v[e] = j0 + e < p ? src[j0 + e] : T(0);
      }
      if (wt) {
        // DtR ./= weights (decomp/decomp.h:1949); a resumed DtR is divided
        const T* wcol = a.weights + (size_t)sig * a.ldw;
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          const T we = j0 + e < p ? wcol[j0 + e] : T(1);
          sWt[j0 + e] = we;
          if (!resume) v[e] = v[e] / we;
        }
      }
      // rows past p hold +INF: both step candidates of such a row are +INF
      // (or NaN, which never wins), so the search needs no row bound
#pragma unroll
      for (int e = 0; e < VEC; ++e)
        if (j0 + e >= p) v[e] = INF;
      *reinterpret_cast<V*>(&sDtR[j0]) = v;
    }
    // active flags of this thread's rows (bit q * VEC + e)
    ActMask act = 0;
    const T lam = uniform_v(a.lambda[a.nlam > 1 ? sig : 0]);
    T normX;
    T thrs0 = 0;
    int currentInd;
    int i0 = 0;
    int s_active = 0;
    int iter = 0, admissions = 0;
    bool empty = false, overflow = false, reached = false;
    // masked: partly observed signals use Gram columns over their observed
    // rows; a signal with no observed row has an empty column
    const int nobs = mk ? uniform_v(a.mobs[sig]) : 0;
    const bool own = mk && nobs > 0 && nobs < a.m;
    const int* mrl = own ? a.mrows + (size_t)sig * a.m : nullptr;
    const bool mcomp = own && masked_gram_comp(nobs, a.m);
    // the ridge of a masked Gram column summed over the observed rows (the
    // complement form starts from G, which carries it)
    const T mdiag = mcomp ? T(0) : a.diag;
    const int mnsel = own ? masked_gram_nsel(nobs, a.m) : 0;
    if (mk)
      for (int j = tid; j < LC; j += GT) sSlot[j] = j;
    if (resume) {
      const int s0 = uniform_v(st.isc[2]);
      if (own) {
        // the Gram columns of the handed-over active set, slots 0..s0-1
        for (int k = 0; k < s0; ++k)
          masked_gram_col<T, GT, NQ>(a.Dt, a.Dd, a.G, a.m, p, ldg, mrl, mnsel, mcomp, st.ind[k],
                                     gaw + (size_t)k * ldg, mdiag);
      }
      for (int j = tid; j < LC; j += GT) {
        sCoef[j] = j < s0 ? st.coef[j] : T(0);
        sInd[j] = j < s0 ? st.ind[j] : -1;
      }
      for (int e = tid; e < s0 * s0; e += GT) {
        const int r = e % s0, q = e / s0;
        if (CL >= s0) sMl[r + q * CL] = st.M[r + q * LarsState<T>::C];
        else M0[r + q * ldm] = st.M[r + q * LarsState<T>::C];
      }
      for (int k = 0; k < s0; ++k) {
        const int c = st.ind[k];
        if (tid == (c / VEC) % GT) act |= ActMask(1) << ((c / VEC / GT) * VEC + c % VEC);
      }
      normX = uniform_v(st.sc[0]);
      thrs0 = uniform_v(st.sc[1]);
      iter = uniform_v(st.isc[0]);
      admissions = uniform_v(st.isc[1]);
      i0 = s0;
      s_active = s0;
      currentInd = uniform_v(st.isc[3]);
      __syncthreads();
    } else {
      for (int j = tid; j < LC; j += GT) {
        sCoef[j] = 0;
        sInd[j] = -1;
      }
      normX = uniform_v(a.norms[sig]);
      if (mk && nobs == 0) empty = true;
      __syncthreads();

      // first atom: i?amax of DtR, or the plain max when pos
      {
        T key = INF, pay = 0;
        int idx = INT_MAX;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
          for (int e = 0; e < VEC; ++e) {
            const int j = VEC * (tid + GT * q) + e;
            const T d = sDtR[j];
            const T k = j < p ? (a.pos ? -d : -fabs(d)) : INF;
            take_if_better(key, idx, pay, k, j < p ? j : INT_MAX, T(0));
          }
        group_argmin<T, W>(key, idx, pay, sRed);
        currentInd = uniform_v(idx == INT_MAX ? 0 : idx);
      }
      const T dc = uniform_v(sDtR[currentInd]);
      if (a.mode == SPX_PENALTY && fabs(dc) < lam) empty = true;
      if (a.mode == SPX_L2ERROR && normX < lam) empty = true;
    }
    int ovf_i = 0;
    T ovf_thrs = 0;

    T* Mc = M0;
    T* Mn = M1;
    auto sAlpha_ = [&]() { return CAP > 0 ? reinterpret_cast<double*>(Mn) : sAlphaF; };
    auto sRes_ = [&]() { return CAP > 0 ? reinterpret_cast<double*>(Mn) + LCL : sResF; };
    // full-capacity pass: the inverse starts in LDS (capacity CL) and moves
    // to the HBM slot when the active set outgrows it
    bool inl = CAP == 0 && CL > 0 && (!resume || CL >= i0);
    if (!empty) {
      bool newAtom = true;
      T thrs = thrs0;
      T gpre = 0;          // prefetched entry of the next admission's column
      bool pre_ok = false;
      for (int i = i0; i < Lm; ++i) {
        ++iter;
        bool have_u = false;  // the admission formed the direction
        if (newAtom) {
          if (CAP > 0 && i >= LC) {
            // the active set outgrows the fast pass: hand the signal over
            overflow = true;
            ovf_i = i;
            ovf_thrs = thrs;
            break;
          }
          // ---------------------------------------------- admit an atom
          const int c = currentInd;
          ++admissions;
          ++s_active;
          if (tid == 0) sInd[i] = c;
          if (tid == (c / VEC) % GT) act |= ActMask(1) << ((c / VEC / GT) * VEC + c % VEC);
          if (own) __syncthreads();
          const T* gcol = a.G + (size_t)c * ldg;
          if (own) {
            T* slot = gaw + (size_t)sSlot[i] * ldg;
            masked_gram_col<T, GT, NQ>(a.Dt, a.Dd, a.G, a.m, p, ldg, mrl, mnsel, mcomp, c, slot,
                                       mdiag);
            gcol = slot;
          }
          // the new Gram column's active entries, and the signed weights of
          // the direction (coreLARS2W: the weights, decomp/decomp.h:1986-1987)
          for (int j = tid; j <= i; j += GT) {
            const int ij = j < i ? sInd[j] : c;  // sInd[i] is published below
            // G(ij, c): from column ij (just streamed, cached) when G is
            // exactly symmetric
            sGs[j] = pre_ok ? gpre : (a.gsym && !own ? a.G[(size_t)ij * ldg + c] : gcol[ij]);
            const T one = wt ? sWt[ij] : T(1);
            sW[j] = sDtR[ij] > 0 ? one : -one;
          }
          __syncthreads();
          if (i == 0) {
            if (tid == 0) {
              if (inl) sMl[0] = T(1) / sGs[0];
              else Mc[0] = T(1) / sGs[0];
            }
          } else {
            if (inl && i >= CL) {
              // the block outgrows the LDS copy: continue in the HBM slot
              for (int e = tid; e < i * i; e += GT) {
                const int r = e % i, q = e / i;
                Mc[r + q * ldm] = sMl[r + q * CL];
              }
              __syncthreads();
              inl = false;
            }
            if (inl)
              have_u = inverse_admit<T, W, true>(sMl, CL, i, sGs, sV, sPart, sRedS, par,
                                                       sRes_(), sW, sU);
            else
              have_u = inverse_admit<T, W, CAP == 0>(Mc, ldm, i, sGs, sV, sPart,
                                                               sRedS, par, sRes_(), sW, sU);
          }
          // ols: normX is the least-squares residual of the active set
          // (decomp/decomp.h:1044-1050)
          if (a.ols)
            normX = ls_residual<T, W>(inl ? sMl : Mc, inl ? CL : ldm, i + 1, dtx, sInd,
                                      a.norms[sig], sRes_());
        }

        const int s = i + 1;
        // ------------------------------------------ direction u = invGs*w
        // (coreLARS2W: the signed weights, decomp/decomp.h:1986-1987); an
        // admission has set the weights, and formed u with the row-wise update
        if (!newAtom) {
          for (int j = tid; j < s; j += GT) {
            const T one = wt ? sWt[sInd[j]] : T(1);
            sW[j] = sDtR[sInd[j]] > 0 ? one : -one;
          }
          __syncthreads();
        }
        if (!have_u) {
          if (inl) symv_dist<T, GT, true>(sMl, CL, sW, sU, sPart, s);
          else symv_dist<T, GT, CAP == 0>(Mc, ldm, sW, sU, sPart, s);
        }

        // ------------------ zero-crossing step, polynomial coefficients
        T step_max;
        int first_zero;
        T coeff1 = 0, coeff2 = 0;
        {
          T key = INF, pay = 0;
          int idx = INT_MAX;
          unsigned long long kz = pack_key(__builtin_huge_valf(), INT_MAX);
          for (int j = tid; j < s; j += GT) {
            const T uj = sU[j];
            const T ratio = -sCoef[j] / uj;
            // the sequential scan keeps the LAST j on ties: index L - j
            if constexpr (sizeof(T) == 4) {
              const unsigned long long k =
                  ratio > 0 ? pack_key(static_cast<float>(ratio), L - j)
                            : pack_key(__builtin_huge_valf(), INT_MAX);
              kz = k < kz ? k : kz;
            } else {
              take_if_better(key, idx, pay, ratio > 0 ? ratio : INF,
                             ratio > 0 ? L - j : INT_MAX, T(0));
            }
            const T dj = sDtR[sInd[j]];
            if (wt) {
              // decomp/decomp.h:2032-2037
              const T wj = sWt[sInd[j]];
              coeff1 += dj > 0 ? wj * uj : -(wj * uj);
              coeff2 += dj * uj * wj;
            } else {
              coeff1 += dj > 0 ? uj : -uj;
              coeff2 += dj * uj;
            }
          }
          if constexpr (sizeof(T) == 4) {
            kz = group_min_u64<W>(kz, sRedU, par);
            key = packed_key(kz);
            idx = static_cast<int>(static_cast<unsigned>(kz));
          } else {
            group_argmin<T, W>(key, idx, pay, sRed);
          }
          group_sum2<T, W>(coeff1, coeff2, sRedS, par);
          step_max = key;
          first_zero = idx == INT_MAX ? -1 : L - idx;
        }
        const T cc = uniform_v(T(fabs(sDtR[sInd[0]])));

        // ------------------------------------- A = G(:, active) * u
        V Areg[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) Areg[q] = V(0);
        // UNR Gram columns per round so their loads are in flight together;
        // the sum over k keeps its sequential order.  The loads are
        // unconditional, so no branch or register reuse serialises them:
        // when the group's rows fit the column (PT <= ldg) they address it
        // directly (immediate offsets); else rows past ldg read the column's
        // first chunk, which only rows past p, never candidates, accumulate.
        auto stream = [&](auto offset) {
          for (int k0 = 0; k0 < s; k0 += UNR) {
            T uk[UNR];
            const T* col[UNR];
            // the round's multipliers and column numbers: one vector LDS read
            // of consecutive entries each (k0 is a multiple of UNR), then
            // scalar registers, so the column addresses and multipliers cost
            // no vector instructions; entries past s repeat column k0 with
            // multiplier 0
            constexpr int AL = UNR * 4 < 16 ? UNR * 4 : 16;
            const int* cp = static_cast<const int*>(
                __builtin_assume_aligned((own ? sSlot : sInd) + k0, AL));
            const T* up = static_cast<const T*>(__builtin_assume_aligned(sU + k0, AL));
            int cv[UNR];
            T uv[UNR];
#pragma unroll
            for (int e = 0; e < UNR; ++e) {
              cv[e] = cp[e];
              uv[e] = up[e];
            }
            int cs[UNR];
            T us[UNR];
#pragma unroll
            for (int e = 0; e < UNR; ++e) {
              cs[e] = uniform_v(cv[e]);
              us[e] = uniform_v(uv[e]);
            }
#pragma unroll
            for (int e = 0; e < UNR; ++e) {
              const bool live = k0 + e < s;
              uk[e] = live ? us[e] : T(0);
              col[e] = (own ? gaw : a.G) + (size_t)(live ? cs[e] : cs[0]) * ldg;
            }
            V g[UNR][NQ];
#pragma unroll
            for (int e = 0; e < UNR; ++e)
#pragma unroll
              for (int q = 0; q < NQ; ++q)
                g[e][q] = *reinterpret_cast<const V*>(col[e] + offset(q));
#pragma unroll
            for (int e = 0; e < UNR; ++e)
#pragma unroll
              for (int q = 0; q < NQ; ++q) Areg[q] += uk[e] * g[e][q];
          }
        };
        if (PT <= ldg) {
          stream([&](int q) { return VEC * (tid + GT * q); });
        } else {
          stream([&](int q) {
            const int j0 = VEC * (tid + GT * q);
            return j0 < ldg ? j0 : 0;
          });
        }
        if (wt) {
          // A ./= weights (decomp/decomp.h:2005)
#pragma unroll
          for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int e = 0; e < VEC; ++e) Areg[q][e] = Areg[q][e] / sWt[VEC * (tid + GT * q) + e];
        }

        // ------------------------------- next atom: custom iamin over 2p
        // candidates (decomp/decomp.h:1803-1820): c1 = (d + cc) / (1 + A) and
        // c2 = (cc - d) / (1 - A) over the inactive atoms, key |c|, first
        // index of the minimum, IEEE quotients.
        T bestk = INF, bestv = INF;
        int besti = INT_MAX;
        // exact keys of every row of this thread (IEEE quotients and the
        // reference's validity rules), folded into (bestk, besti, bestv)
        auto exact_rows = [&]() {
#pragma unroll
          for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
              const int j = VEC * (tid + GT * q) + e;
              const bool cand = (j < p) & !((act >> (q * VEC + e)) & 1);
              const T Aj = Areg[q][e];
              const T d = sDtR[j];
              T c1 = (Aj < INF && Aj > T(-1)) ? (d + cc) / (T(1) + Aj) : INF;
              c1 = a.pos ? INF : c1;
              const T c2 = (Aj < INF && Aj < T(1)) ? (cc - d) / (T(1) - Aj) : INF;
              take_if_better(bestk, besti, bestv, cand ? T(fabs(c1)) : INF,
                             cand ? j : INT_MAX, c1);
              take_if_better(bestk, besti, bestv, cand ? T(fabs(c2)) : INF,
                             cand ? p + j : INT_MAX, c2);
            }
        };
        if constexpr (sizeof(T) == 4) {
          // pass 1: one key per row, min(|c1|, |c2|), from hardware
          // reciprocals (within 4 ulp of the IEEE quotients) of denominators
          // clamped at zero: an invalid candidate (denominator <= 0 or NaN)
          // gets +INF or NaN, which never wins; under pos c1 gets a NaN
          // numerator.  The key's low SB bits are replaced by the row's slot
          // (a truncation below 2^-17 relative) and an active row's key by
          // all ones.  Each thread keeps its two smallest packed keys.
          constexpr int NR = NQ * VEC;  // rows per thread
          constexpr int SB = NR <= 16 ? 4 : (NR <= 32 ? 5 : 6);
          constexpr unsigned SMASK = (1u << SB) - 1u;
          const T cc1 = a.pos ? __builtin_nanf("") : cc;
          unsigned m1 = 0xFFFFFFFFu, m2 = 0xFFFFFFFFu;
#pragma unroll
          for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
              const int sl = q * VEC + e;
              const T Aj = Areg[q][e];
              const T d = sDtR[VEC * (tid + GT * q) + e];
              const T r1 = __builtin_amdgcn_rcpf(fmaxf(T(1) + Aj, T(0)));
              const T r2 = __builtin_amdgcn_rcpf(fmaxf(T(1) - Aj, T(0)));
              const T k = fminf(fabsf((d + cc1) * r1), fabsf((cc - d) * r2));
              const unsigned ex = 0u - static_cast<unsigned>((act >> sl) & 1);
              const unsigned kt = ((__float_as_uint(k) & ~SMASK) | static_cast<unsigned>(sl)) | ex;
              const unsigned hi = m1 > kt ? m1 : kt;
              m2 = m2 < hi ? m2 : hi;
              m1 = m1 < kt ? m1 : kt;
            }
          // pass 2 recomputes what it needs: no pass-1 temporaries stay live
          // (register pressure sets the signals in flight per CU)
#pragma unroll
          for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
              T x = Areg[q][e];
              asm volatile("" : "+v"(x));
              Areg[q][e] = x;
            }
          asm volatile("" ::: "memory");
          // every row whose exact key may be the minimum has its truncated
          // key within 2^-16 of the group minimum (4 ulp of reciprocal error
          // and the 2^-17 truncation, both ways): those rows get the exact
          // quotients.  No finite key anywhere (NaN minimum): every row.
          const T amin = group_min_f32<W>(__uint_as_float(m1 & ~SMASK), sRedS, par);
          const bool anyk = amin == amin;
          const T thr = anyk ? amin * T(1.0f + 0x1.0p-16f) + T(1e-37f) : INF;
          const unsigned thb = anyk ? __float_as_uint(thr) : 0xFFFFFFFFu;
          if ((m2 & ~SMASK) <= thb) {
            // two rows of this thread near the minimum: all its rows
            exact_rows();
          } else if ((m1 & ~SMASK) <= thb) {
            // one row of this thread near the minimum
            const int sl = static_cast<int>(m1 & SMASK);
            T Aj = 0;
#pragma unroll
            for (int q = 0; q < NQ; ++q)
#pragma unroll
              for (int e = 0; e < VEC; ++e) Aj = sl == q * VEC + e ? Areg[q][e] : Aj;
            const int j = VEC * (tid + GT * (sl / VEC)) + sl % VEC;
            const bool cand = j < p;
            const T d = sDtR[j];
            T c1 = (Aj < INF && Aj > T(-1)) ? (d + cc) / (T(1) + Aj) : INF;
            c1 = a.pos ? INF : c1;
            const T c2 = (Aj < INF && Aj < T(1)) ? (cc - d) / (T(1) - Aj) : INF;
            take_if_better(bestk, besti, bestv, cand ? T(fabs(c1)) : INF, cand ? j : INT_MAX,
                           c1);
            take_if_better(bestk, besti, bestv, cand ? T(fabs(c2)) : INF,
                           cand ? p + j : INT_MAX, c2);
          }
          // (|c|, index, sign of c) packed: the sign rides below the index
          auto reduce = [&]() {
            unsigned long long kb = pack_key(
                bestk, besti == INT_MAX
                           ? 0xFFFFFFFFu
                           : 2u * static_cast<unsigned>(besti) +
                                 (__float_as_uint(static_cast<float>(bestv)) >> 31));
            kb = group_min_u64<W>(kb, sRedU, par);
            const unsigned lo = static_cast<unsigned>(kb);
            bestk = packed_key(kb);
            besti = lo == 0xFFFFFFFFu ? INT_MAX : static_cast<int>(lo >> 1);
            bestv = (lo & 1u) ? -bestk : bestk;
          };
          reduce();
          if (!(bestk <= thr)) {
            // a pass-1 key below its exact value (a correlation change of
            // +INF, whose clamped reciprocal is 0): every row exactly
            bestk = INF;
            bestv = INF;
            besti = INT_MAX;
            exact_rows();
            reduce();
          }
        } else {
          exact_rows();
          group_argmin<T, W>(bestk, besti, bestv, sRed);
        }
        T step = besti == INT_MAX ? INF : bestv;
        const int nextInd = besti == INT_MAX ? 0 : besti % p;
        // the next admission's Gram entries G(active, nextInd), loaded now so
        // that their latency overlaps the updates (unused if the step drops)
        pre_ok = !own && s < GT;
        if (pre_ok && tid <= s) {
          const int ij = tid < s ? sInd[tid] : nextInd;
          gpre = a.gsym ? a.G[(size_t)ij * ldg + nextInd] : a.G[(size_t)nextInd * ldg + ij];
        }

        const T coeff3 = normX - lam;
        T step_max2;
        if (a.mode == SPX_PENALTY) {
          step_max2 = cc - lam;
        } else if (a.mode == SPX_L2ERROR) {
          const T delta = coeff2 * coeff2 - coeff1 * coeff3;
          step_max2 = delta < 0 ? INF : (coeff2 - sqrt(delta)) / coeff1;
          if (!wt) step_max2 = step_max2 < cc ? step_max2 : cc;
          // the Cholesky variant (coreLARS) takes the whole step past the
          // residual target (decomp/decomp.h:997-1033)
          if (a.chol) step_max2 = INF;
        } else {
          step_max2 = coeff1 < 0 ? INF : (lam - thrs) / coeff1;
          if (!wt) step_max2 = step_max2 < cc ? step_max2 : cc;
        }
        step = step < step_max2 ? step : step_max2;
        step = step < step_max ? step : step_max;
        if (step == INF) break;
        if (a.chol && !a.ols) {
          // Cholesky variant: a step that would increase the residual ends
          // the path and drops the entry at the current position
          // (decomp/decomp.h:1036-1043; not with ols, whose residual is the
          // least-squares one)
          const T t = coeff1 * step * step - 2 * coeff2 * step;
          if (t > T(0) || !(fabs(t) < INF)) {
            s_active = i;
            break;
          }
        }
        reached = (a.mode == SPX_PENALTY) && (step == step_max2);

        // -------------------------------------------------- updates
        for (int j = tid; j < s; j += GT) {
          T cj = sCoef[j] + step * sU[j];
          if (a.pos && !wt && cj < 0) cj = 0;
          sCoef[j] = cj;
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int j0 = VEC * (tid + GT * q);
#pragma unroll
          for (int e = 0; e < VEC; ++e) sDtR[j0 + e] -= step * Areg[q][e];
        }
        if (!a.ols) normX += coeff1 * step * step - 2 * coeff2 * step;
        thrs += step * coeff1;
        __syncthreads();
        if (a.path && sig == 0 && iter < a.path_cols) {
          // regularisation path of the first signal (decomp/decomp.h:1868-1871)
          for (int k = tid; k <= i; k += GT) a.path[(size_t)iter * p + sInd[k]] = sCoef[k];
        }

        if (step == step_max) {
          // ------------------- drop atom first_zero (decomp/decomp.h:1875-1903)
          const int fz = first_zero;
          const int dropped = uniform_v(sInd[fz]);
          if (inl) {
            // LDS copy: stage the block in the HBM slot, downdate back
            for (int e = tid; e < (i + 1) * (i + 1); e += GT) {
              const int r = e % (i + 1), q = e / (i + 1);
              Mn[r + q * ldm] = sMl[r + q * CL];
            }
            __syncthreads();
            inverse_drop<T, W, true>(Mn, ldm, sMl, CL, i, fz, sV);
          } else {
            inverse_drop<T, W, CAP == 0>(Mc, ldm, Mn, ldm, i, fz, sV);
          }
          for (int j = tid; j < i; j += GT) {
            const int j2 = j + (j >= fz);
            sTmpI[j] = sInd[j2];
            sGs[j] = sCoef[j2];
          }
          int freed = 0;
          if (mk) {
            // Gram-column slots follow their atoms; the freed one goes last
            freed = sSlot[fz];
            for (int j = tid; j < i; j += GT) sSlotT[j] = sSlot[j + (j >= fz)];
          }
          __syncthreads();
          if (mk) {
            for (int j = tid; j < i; j += GT) sSlot[j] = sSlotT[j];
            if (tid == 0) sSlot[i] = freed;
          }
          if (tid == (dropped / VEC) % GT)
            act &= ~(ActMask(1) << ((dropped / VEC / GT) * VEC + dropped % VEC));
          for (int j = tid; j < i; j += GT) {
            sInd[j] = sTmpI[j];
            sCoef[j] = sGs[j];
          }
          if (tid == 0) {
            sInd[i] = -1;
            sCoef[i] = 0;
          }
          __syncthreads();
          if (!inl) {
            T* t = Mc;
            Mc = Mn;
            Mn = t;
          }
          --s_active;
          newAtom = false;
          // ols: the least-squares residual of the reduced active set
          // (downDateLasso, decomp/decomp.h:1191-1201)
          if (a.ols && s_active > 0)
            normX = ls_residual<T, W>(inl ? sMl : Mc, inl ? CL : ldm, s_active, dtx, sInd,
                                      a.norms[sig], sRes_());
          i = i - 2;
          // an emptied active set has no direction: end the path (the
          // reference would read ind[-1] here)
          if (s_active == 0) break;
        } else {
          newAtom = true;
          currentInd = nextInd;
        }
        // Stop rules of decomp/decomp.h:1907-1911 (tolerance 1e-15), or of
        // coreLARS2W, decomp/decomp.h:2099-2103 (1e-10, iter > 4L through
        // length_path = 4L + 2).  In FP32 a vanishing step
        // that admits an atom is a tie between two atoms at FP32 resolution
        // (both were due at the same point of the path): the atom is admitted
        // and the path goes on, where the rule would end it short of lambda.
        const bool tie = sizeof(T) == 4 && newAtom;
        const T tol = a.tol;
        if ((iter >= a.length_path - 1) || (fabs(step) < tol && !tie) ||
            step == step_max2 || (normX < tol) || (i == (Lm - 1)) ||
            (a.mode == SPX_L2ERROR && (newAtom || !a.chol) && normX - lam < tol) ||
            (a.mode == SPX_L1COEFFS && (lam - thrs < tol)))
          break;
      }
    }

    // ---------- optional FP64 refit (PENALTY, at lambda) or least-squares fit
    // ols: alpha_A = G_AA^-1 (D'x)_A, the coefficients the
    // reference's ols option returns (decomp/decomp.h:1087-1091)
    if (!mk && ((a.refit && reached) || a.ols) && !overflow && s_active > 0) {
      __syncthreads();
      const int s = s_active;
      const T* Mf = inl ? sMl : Mc;  // the inverse of the final active block
      const int ldf = inl ? CL : ldm;
      double* const sAlpha = sAlpha_();
      double* const sRes = sRes_();
      for (int r = tid; r < s; r += GT) sAlpha[r] = a.ols ? 0.0 : static_cast<double>(sCoef[r]);
      __syncthreads();
      for (int it = 0; it < (a.ols ? 3 : 2); ++it) {
        // res = (D'x)_A - lam sign(alpha) - (G_AA + d I) alpha   (FP64)
        for (int r = tid; r < s; r += GT) {
          const int ir = sInd[r];
          const T* gcol = a.G + (size_t)ir * ldg;
          double acc = static_cast<double>(dtx[ir]);
          if (!a.ols) acc -= static_cast<double>(lam) * (sCoef[r] > 0 ? 1.0 : -1.0);
          for (int q = 0; q < s; ++q) acc -= static_cast<double>(gcol[sInd[q]]) * sAlpha[q];
          sRes[r] = acc;
        }
        __syncthreads();
        // alpha += invGs res
        for (int r = tid; r < s; r += GT) {
          double acc = 0;
          for (int q = 0; q < s; ++q) acc += static_cast<double>(Mf[r + q * ldf]) * sRes[q];
          sAlpha[r] += acc;
        }
        __syncthreads();
      }
      for (int r = tid; r < s; r += GT) sCoef[r] = static_cast<T>(sAlpha[r]);
    }

    // ------------------------------------------------------- outputs
    __syncthreads();
    if (overflow) {
      if (tid == 0) {
        const int pos = atomicAdd(a.ovf_count, 1);
        a.ovf_list[pos] = sig;
        // remaining path: from the current correlation down to lambda (the
        // smaller the ratio, the longer; the full pass takes these first)
        if (a.ovf_key) {
          const T kk = lam / fabs(sDtR[sInd[0]]);
          a.ovf_key[pos] = kk == kk ? kk : T(0);  // NaN sorts first, not past the list
        }
      }
      if (CAP > 0 && a.state) {
        // hand-over state for the full-capacity pass
        LarsState<T> sv(static_cast<char*>(a.state) + (size_t)sig * a.state_bytes, ldg);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int j0 = VEC * (tid + GT * q);
          if (j0 < ldg) *reinterpret_cast<V*>(sv.dtr + j0) = *reinterpret_cast<const V*>(&sDtR[j0]);
        }
        for (int j = tid; j < ovf_i; j += GT) {
          sv.coef[j] = sCoef[j];
          sv.ind[j] = sInd[j];
        }
        for (int e = tid; e < ovf_i * ovf_i; e += GT) {
          const int r = e % ovf_i, q = e / ovf_i;
          sv.M[r + q * LarsState<T>::C] = Mc[r + q * ldm];
        }
        if (tid == 0) {
          sv.sc[0] = normX;
          sv.sc[1] = ovf_thrs;
          sv.isc[0] = iter - 1;  // the interrupted iteration is redone
          sv.isc[1] = admissions;
          sv.isc[2] = ovf_i;
          sv.isc[3] = currentInd;
        }
      }
      continue;
    }
    T* ov = a.outV + (size_t)sig * L;
    int* orr = a.outR + (size_t)sig * L;
    // the live part of the list and its -1 terminator (entries past the
    // terminator are left unspecified; consumers read `count` entries)
    const int nout = s_active < L ? s_active + 1 : L;
    for (int j = tid; j < nout; j += GT) {
      ov[j] = j < s_active ? sCoef[j] : T(0);
      orr[j] = j < s_active ? sInd[j] : -1;
    }
    if (tid == 0) {
      a.outCount[sig] = s_active;
      if (a.stats) {
        a.stats[2 * (size_t)sig] = iter;
        a.stats[2 * (size_t)sig + 1] = admissions;
      }
    }
  }
}

// Launches one instantiation: persistent grid over a counter (batched) or
// a grid-stride grid with one signal per workgroup at a time (plain; CAP
// must be 0).
template <typename T, int W, int NQ, int CAP, int VAR>
int lars_launch(Ctx* ctx, const LarsArgs<T>& a, bool persistent, int* counter,
                long max_blocks) {
  const int LC = CAP > 0 ? (CAP < a.L ? CAP : a.L) : a.L;
  const int LCL = CAP > 0 ? CAP : a.L;  // the kernel's LDS layout capacity
  size_t lds = LarsLds<T, W, NQ, CAP>::bytes(LCL, 0, VAR);
  if (lds > 160 * 1024) {
    set_error("lasso: L=%d, p=%d need %zu bytes of LDS per signal", a.L, a.p, lds);
    return SPX_ERR_UNSUP;
  }
  auto kern = lars_kernel<T, W, NQ, CAP, VAR>;
  SPX_HIP(hipFuncSetAttribute((const void*)kern,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  hipStream_t st = ctx->stream;
  LarsArgs<T> b = a;
  b.lds_cap = 0;
  if (CAP == 0) {
    // LDS copy of the inverse: the LDS the register-limited occupancy leaves
    // free, if it holds at least the LDS pass's capacity
    int per_cu = 0;
    SPX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * W, lds));
    if (per_cu < 1) per_cu = 1;
    const size_t budget = (size_t)156 * 1024 / per_cu;
    int cl = 0;
    while (cl + 4 <= a.L && LarsLds<T, W, NQ, CAP>::bytes(LCL, cl + 4, VAR) <= budget) cl += 4;
    if (cl >= a.L) cl = a.L;
    if (cl >= FastCap<T>::C) {
      b.lds_cap = cl;
      lds = LarsLds<T, W, NQ, CAP>::bytes(LCL, cl, VAR);
    }
  }
  if (persistent) {
    int per_cu = 0;
    SPX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * W, lds));
    if (per_cu < 1) per_cu = 1;
    long blocks = (long)per_cu * ctx->num_cus;
    if (blocks > max_blocks) blocks = max_blocks;
    if (blocks < 1) blocks = 1;
    if (CAP == 0) {
      const size_t slot = 2 * sizeof(T) * (size_t)a.L * a.L;
      // an allocation-free call runs as many workgroups as inverse slots
      // fit its static scratch (the results do not depend on the grid)
      const size_t room = ctx->scratch_room();
      if (slot > 0 && room / slot < (size_t)blocks) blocks = (long)(room / slot);
      if (blocks < 1) blocks = 1;
      T* scratch = nullptr;
      int rc = ctx->scratch_get(0, slot * blocks, reinterpret_cast<void**>(&scratch));
      if (rc) return rc;
      b.scratch = scratch;
    }
    if (VAR == 2) {
      // masked signals: LC Gram columns per workgroup
      b.ga_stride = (size_t)a.ldg * LC;
      int rc = ctx->scratch_get(CAP > 0 ? 16 : 17, sizeof(T) * b.ga_stride * blocks,
                                reinterpret_cast<void**>(&b.ga));
      if (rc) return rc;
    }
    b.queue = counter;
    // the LDS pass takes 4 signals per grab of the shared counter when the
    // batch gives every workgroup 32 or more: one device-scope atomic per
    // signal on one address queues up at the L2 and its latency reaches the
    // set-up of every signal; the batch's last signals (its shortest, in
    // the longest-first order) bound the extra tail
    if (CAP > 0 && b.chunk <= 1 && a.n >= 32 * blocks) b.chunk = 4;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(64 * W), lds, st, b);
    SPX_HIP(hipGetLastError());
    return SPX_OK;
  }
  // plain form: one grid of as many workgroups as fit the CUs at once (at
  // most one per signal), each walking the signals blockIdx.x, blockIdx.x +
  // gridDim.x, ... with its own inverse slot: no work queue, no ordering
  int per_cu = 0;
  SPX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * W, lds));
  if (per_cu < 1) per_cu = 1;
  long blocks = (long)per_cu * ctx->num_cus;
  if (blocks > a.n) blocks = a.n;
  const size_t slot = 2 * sizeof(T) * (size_t)a.L * a.L;
  const size_t per_block = slot + (VAR == 2 ? sizeof(T) * (size_t)a.ldg * a.L : 0);
  const size_t room = ctx->scratch_room();
  if (per_block > 0 && room / per_block < (size_t)blocks) blocks = (long)(room / per_block);
  if (blocks < 1) blocks = 1;
  T* scratch = nullptr;
  int rc = ctx->scratch_get(0, slot * blocks, reinterpret_cast<void**>(&scratch));
  if (rc) return rc;
  b.queue = nullptr;
  b.list = nullptr;
  b.scratch = scratch;
  if (VAR == 2) {
    // masked signals: L Gram columns per workgroup
    b.ga_stride = (size_t)a.ldg * a.L;
    rc = ctx->scratch_get(17, sizeof(T) * b.ga_stride * blocks, reinterpret_cast<void**>(&b.ga));
    if (rc) return rc;
  }
  hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(64 * W), lds, st, b);
  SPX_HIP(hipGetLastError());
  return SPX_OK;
}

template <typename T, int W>
int lars_nq(int p) {
  const int per = Vec16<T>::N * 64 * W;
  const int nq = (p + per - 1) / per;
  return nq <= 1 ? 1 : nq <= 2 ? 2 : nq <= 4 ? 4 : nq <= 8 ? 8 : nq <= 16 ? 16 : -1;
}

// One pass of group width W: CAP = FastCap (fast) or 0 (full capacity);
// the weighted penalty has kernels of its own (the plain ones keep their
// registers).
template <typename T, int W, bool FAST, int VAR>
int lars_pass_t(Ctx* ctx, const LarsArgs<T>& a, bool persistent, int* counter, long max_blocks) {
  constexpr int C = FAST ? FastCap<T>::C : 0;
  switch (lars_nq<T, W>(a.p)) {
    case 1: return lars_launch<T, W, 1, C, VAR>(ctx, a, persistent, counter, max_blocks);
    case 2: return lars_launch<T, W, 2, C, VAR>(ctx, a, persistent, counter, max_blocks);
    case 4: return lars_launch<T, W, 4, C, VAR>(ctx, a, persistent, counter, max_blocks);
    case 8: return lars_launch<T, W, 8, C, VAR>(ctx, a, persistent, counter, max_blocks);
    case 16: return lars_launch<T, W, 16, C, VAR>(ctx, a, persistent, counter, max_blocks);
    default: return -1;  // does not fit this group width
  }
}

// Pass of group width w (1, 2, 4 or 16) of the kernels of variant VAR (0
// plain, 1 weighted, 2 masked); -1 when the width does not cover p.
template <typename T, int VAR>
int lars_pass_wt(Ctx* ctx, int w, bool fast, const LarsArgs<T>& a, bool persistent,
                 int* counter, long max_blocks) {
  switch (w) {
    case 1: return fast ? lars_pass_t<T, 1, true, VAR>(ctx, a, persistent, counter, max_blocks)
                        : lars_pass_t<T, 1, false, VAR>(ctx, a, persistent, counter, max_blocks);
    case 2: return fast ? lars_pass_t<T, 2, true, VAR>(ctx, a, persistent, counter, max_blocks)
                        : lars_pass_t<T, 2, false, VAR>(ctx, a, persistent, counter, max_blocks);
    case 4: return fast ? lars_pass_t<T, 4, true, VAR>(ctx, a, persistent, counter, max_blocks)
                        : lars_pass_t<T, 4, false, VAR>(ctx, a, persistent, counter, max_blocks);
    default: return fast ? lars_pass_t<T, 16, true, VAR>(ctx, a, persistent, counter, max_blocks)
                         : lars_pass_t<T, 16, false, VAR>(ctx, a, persistent, counter, max_blocks);
  }
}

}  // namespace
}  // namespace spx
