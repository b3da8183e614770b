// ORMP ("omp") solver kernel: coreORMP semantics (decomp/decomp.h:538-640).
//
// One group of W wavefronts (one workgroup) per signal at a time.  Per-row
// state (the residual correlations Rdn and the projected atom norms `norm`)
// lives in registers of the owning thread, the selected-atom flags in a bit
// mask of that thread; RUn, the selection order and (for L up to the LDS
// capacity) the inverse Cholesky factor Un (upper triangular, L x L) live in
// LDS, a larger Un in a per-workgroup HBM slot.
//
// The reference keeps Undn = Gs * Un (p x L) only to read row c of it when
// atom c is selected (decomp/decomp.h:596).  That row equals
// G(c, active) * Un, an O(j^2) product, so it is recomputed from the resident
// Gram column instead of storing p x L per signal.  Each step streams the
// j+1 selected Gram columns once (several in flight) to form
// Undn_j = G(:, active) * Un_j.
//
// The arg-max of the scores Rdn^2 / norm (i?amax: first index of the largest
// magnitude) takes two passes in FP32: keys from the hardware reciprocal
// (within 3 ulp of the IEEE quotient), then the IEEE quotient for the keys
// within a margin of the largest -- the same winner as IEEE quotients
// everywhere, at one division per thread instead of one per row.
#include <climits>
#include <type_traits>

#include "mask_gram.h"
#include "spx_common.h"
#include "spx_internal.h"

namespace spx {

namespace {

template <typename T>
struct Vec16o;
template <>
struct Vec16o<float> {
  typedef float type __attribute__((ext_vector_type(4)));
  static constexpr int N = 4;
};
template <>
struct Vec16o<double> {
  typedef double type __attribute__((ext_vector_type(2)));
  static constexpr int N = 2;
};

// LDS capacity for Un: L x L elements of at most this many bytes.
constexpr size_t kUnLdsBytes = 32 * 1024;

// Held to 4 wavefronts per SIMD up to 16 rows per lane: the signals are
// latency-bound, so occupancy pays more than registers.  MK: masked signals
// (omp_mask, decomp/decomp.h:406-519; a.mobs set).
template <typename T, int W, int NQ, bool ULDS, bool MK>
__global__ __launch_bounds__(64 * W, NQ <= 4 ? 4 : 1) void ormp_kernel(OmpArgs<T> a) {
  constexpr int GT = 64 * W;
  constexpr int VEC = Vec16o<T>::N;
  constexpr int UNR = NQ >= 4 ? 2 : (NQ == 2 ? 4 : 8);  // Gram columns in flight
  typedef typename Vec16o<T>::type V;
  typedef typename std::conditional<(NQ * VEC <= 32), unsigned, unsigned long long>::type ActMask;
  const T INF = dev_inf<T>();

  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int L = a.L;
  const int p = a.p;
  const int ldg = a.ldg;
  T* sRUn = reinterpret_cast<T*>(smem);  // L
  T* sX = sRUn + L;                       // L  (new Un column / scratch)
  T* sGc = sX + L;                        // L  (G(c, active))
  T* sScal = sGc + L;                     // 4  (Rdn[c], norm[c])
  int* sInd = reinterpret_cast<int*>(sScal + 4);  // L
  KeyIdx<T>* sRed = lds_next16<KeyIdx<T>>(smem, sInd + L);
  int* sSig = reinterpret_cast<int*>(sRed + W);
  T* sUn = lds_next16<T>(smem, sSig + 1);

  const int tid = threadIdx.x;
  T* const Un = ULDS ? sUn : a.scratch + (size_t)blockIdx.x * L * L;  // column-major L x L
  T* const gaw = MK ? a.ga + (size_t)blockIdx.x * a.ga_stride : nullptr;

  // the next grab is requested when the current signal starts, so its
  // latency overlaps the set-up loads
  int pre = 0;
  if (a.queue && tid == 0) pre = atomicAdd(a.queue, 1);
  for (int round = 0;; ++round) {
    int sig;
    if (a.queue) {
      __syncthreads();
      if (tid == 0) *sSig = pre;
      __syncthreads();
      sig = *sSig;
      if (tid == 0 && sig < a.n) pre = atomicAdd(a.queue, 1);
    } else {
      sig = round == 0 ? (int)blockIdx.x : a.n;
    }
    if (sig >= a.n) break;

    const T* dtx = a.DtX + (size_t)sig * a.ldd;
    const bool vec_ok = (a.ldd % VEC == 0) &&
                        (reinterpret_cast<uintptr_t>(a.DtX) % (sizeof(T) * VEC) == 0);
    V Rdn[NQ], nrm[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int j0 = VEC * (tid + GT * q);
      if (vec_ok && j0 + VEC <= p) {
        Rdn[q] = *reinterpret_cast<const V*>(dtx + j0);
      } else {
#pragma unroll
        for (int e = 0; e < VEC; ++e) Rdn[q][e] = j0 + e < p ? dtx[j0 + e] : T(0);
      }
      nrm[q] = V(1);
    }
    ActMask act = 0;  // selected atoms among this thread's rows (bit q * VEC + e)
    for (int j = tid; j < L; j += GT) sInd[j] = -1;
    T normX = a.norms[sig];
    const T eps = fabs(a.nEps == 0 ? T(0) : a.eps[a.nEps > 1 ? sig : 0]);
    const T lambda = a.nLambda == 0 ? T(0) : a.lambda[a.nLambda > 1 ? sig : 0];
    int Li = a.nL == 0 ? a.scalarL : a.Ls[a.nL > 1 ? sig : 0];
    if (Li > L) Li = L;
    // masked: no observed row -> empty; partly observed -> the Gram columns
    // of its observed rows plus 1e-10 on the diagonal
    const int nobs = MK ? a.mobs[sig] : 0;
    const bool own = MK && nobs > 0 && nobs < a.m;
    const int* mrl = own ? a.mrows + (size_t)sig * a.m : nullptr;
    const bool mcomp = own && masked_gram_comp(nobs, a.m);
    const int mnsel = own ? masked_gram_nsel(nobs, a.m) : 0;
    if (MK && nobs == 0) Li = 0;
    __syncthreads();

    int j = 0;
    if (!((normX <= eps) || Li <= 0)) {
      for (j = 0; j < Li; ++j) {
        // ------------------------------------------ c = i?amax(scores)
        // scores = Rdn before the first update, then Rdn^2 / norm with the
        // selected atoms zeroed (decomp/decomp.h:550, :626-629); key -|score|
        T key = INF, pay = 0;
        int idx = INT_MAX;
        if (j == 0 || sizeof(T) != 4) {
#pragma unroll
          for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
              const int r = VEC * (tid + GT * q) + e;
              const bool sel = (act >> (q * VEC + e)) & 1;
              const T sc = j == 0 ? Rdn[q][e] : (sel ? T(0) : (Rdn[q][e] * Rdn[q][e]) / nrm[q][e]);
              take_if_better(key, idx, pay, r < p ? T(-fabs(sc)) : INF, r < p ? r : INT_MAX, T(0));
            }
        } else {
          // pass 1: approximate magnitudes; per thread the largest (slot i1)
          // and the second largest
          T m1 = -INF, m2 = -INF;
          int i1 = 0;
#pragma unroll
          for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
              const int r = VEC * (tid + GT * q) + e;
              const bool sel = (act >> (q * VEC + e)) & 1;
              const T ka = (r < p) ? (sel ? T(0)
                                          : T(fabs(Rdn[q][e] * Rdn[q][e] *
                                                   __builtin_amdgcn_rcpf(nrm[q][e]))))
                                   : -INF;
              const bool gt = ka > m1;
              m2 = gt ? m1 : fmaxf(m2, ka);
              i1 = gt ? q * VEC + e : i1;
              m1 = gt ? ka : m1;
            }
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            V x = Rdn[q], y = nrm[q];
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
              T u = x[e], w = y[e];
              asm volatile("" : "+v"(u), "+v"(w));
              x[e] = u;
              y[e] = w;
            }
            Rdn[q] = x;
            nrm[q] = y;
          }
          T amax = -m1, zp = 0;
          int zi = 0;
          group_argmin<T, W>(amax, zi, zp, sRed);  // minimum of -m1
          amax = -amax;
          const T thr = amax * T(1.0f - 0x1.0p-18f);
          if (m2 >= thr) {
            // several of this thread's rows near the maximum (or ties at
            // zero): every row exactly, one at a time
#pragma unroll 1
            for (int sl = 0; sl < NQ * VEC; ++sl) {
              T rd = 0, nr = 1;
#pragma unroll
              for (int q = 0; q < NQ; ++q)
#pragma unroll
                for (int e = 0; e < VEC; ++e) {
                  rd = sl == q * VEC + e ? Rdn[q][e] : rd;
                  nr = sl == q * VEC + e ? nrm[q][e] : nr;
                }
              const int r = VEC * (tid + GT * (sl / VEC)) + sl % VEC;
              const bool sel = (act >> sl) & 1;
              const T sc = sel ? T(0) : (rd * rd) / nr;
              take_if_better(key, idx, pay, r < p ? T(-fabs(sc)) : INF, r < p ? r : INT_MAX, T(0));
            }
          } else if (m1 >= thr) {
            T rd = 0, nr = 1;
#pragma unroll
            for (int q = 0; q < NQ; ++q)
#pragma unroll
              for (int e = 0; e < VEC; ++e) {
                rd = i1 == q * VEC + e ? Rdn[q][e] : rd;
                nr = i1 == q * VEC + e ? nrm[q][e] : nr;
              }
            const int r = VEC * (tid + GT * (i1 / VEC)) + i1 % VEC;
            const bool sel = (act >> i1) & 1;
            const T sc = sel ? T(0) : (rd * rd) / nr;
            take_if_better(key, idx, pay, T(-fabs(sc)), r, T(0));
          }
        }
        group_argmin<T, W>(key, idx, pay, sRed);
        const int c = idx == INT_MAX ? 0 : idx;
        // owner publishes Rdn[c], norm[c]
        {
          const int chunk = c / VEC;
          if (chunk % GT == tid) {
            const int qc = chunk / GT;
#pragma unroll
            for (int q = 0; q < NQ; ++q)
              if (q == qc) {
                sScal[0] = Rdn[q][c % VEC];
                sScal[1] = nrm[q][c % VEC];
              }
          }
        }
        __syncthreads();
        const T normc = sScal[1];
        if (normc < T(1e-8)) {
          if (tid == 0) sInd[j] = -1;
          break;
        }
        const T invNorm = T(1) / sqrt(normc);
        const T RU = sScal[0] * invNorm;
        const T delta = RU * RU;
        if (delta < 2 * lambda) break;
        normX -= delta;
        // G(c, active) for the Un column
        const T* gcol = a.G + (size_t)c * ldg;
        if (own) {
          // the new atom's Gram column, slot j (selection order)
          T* slot = gaw + (size_t)j * ldg;
          masked_gram_col<T, GT, NQ>(a.Dt, a.Dd, a.G, a.m, p, ldg, mrl, mnsel, mcomp, c, slot, T(1e-10));
          gcol = slot;
        }
        // G(c, active): from the active columns (streamed in the last step,
        // cached) when G is exactly symmetric
        for (int k = tid; k < j; k += GT)
          sGc[k] = a.gsym && !own ? a.G[(size_t)sInd[k] * ldg + c] : gcol[sInd[k]];
        __syncthreads();
        // Undn(c, k) = sum_{k' <= k} G(c, ind k') Un(k', k), k < j
        for (int k = tid; k < j; k += GT) {
          T acc = 0;
          for (int kp = 0; kp <= k; ++kp) acc += sGc[kp] * Un[kp + (size_t)k * L];
          sX[k] = acc;
        }
        __syncthreads();
        // Un(:, j) = -invNorm * [Un(0:j,0:j) * Undn(c, 0:j); -1]
        for (int r = tid; r < j; r += GT) {
          T acc = 0;
          for (int q = r; q < j; ++q) acc += Un[r + (size_t)q * L] * sX[q];
          Un[r + (size_t)j * L] = -invNorm * acc;
        }
        if (tid == 0) {
          Un[j + (size_t)j * L] = invNorm;
          sRUn[j] = RU;
          sInd[j] = c;
        }
        if (tid == (c / VEC) % GT) act |= ActMask(1) << ((c / VEC / GT) * VEC + c % VEC);
        __syncthreads();
        if (j == Li - 1 || normX <= eps) {
          ++j;
          break;
        }
        if (a.path && sig == 0 && j < a.path_cols) {
          // path of the first signal: column j = Un(0:j+1,0:j+1) RUn(0:j+1)
          // scattered to the selected atoms (decomp/decomp.h:605-613)
          for (int r = tid; r <= j; r += GT) {
            T acc = 0;
            for (int q = r; q <= j; ++q) acc += Un[r + (size_t)q * L] * sRUn[q];
            sX[r] = acc;
          }
          __syncthreads();
          for (int k = tid; k <= j; k += GT) a.path[(size_t)j * p + sInd[k]] = sX[k];
          __syncthreads();
        }
        // ----------------------- Undn_j = G(:, active) * Un(0:j+1, j)
        V und[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) und[q] = V(0);
        // UNR Gram columns per round so their loads are in flight together;
        // the sum over k keeps its sequential order
        for (int k0 = 0; k0 <= j; k0 += UNR) {
          T uk[UNR];
          const T* col[UNR];
#pragma unroll
          for (int e = 0; e < UNR; ++e) {
            const int k = k0 + e <= j ? k0 + e : k0;
            uk[e] = k0 + e <= j ? Un[k + (size_t)j * L] : T(0);
            col[e] = own ? gaw + (size_t)k * ldg : a.G + (size_t)sInd[k] * ldg;
          }
          V g[UNR][NQ];
#pragma unroll
          for (int e = 0; e < UNR; ++e)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
              const int j0 = VEC * (tid + GT * q);
              g[e][q] = j0 < ldg ? *reinterpret_cast<const V*>(col[e] + j0) : V(0);
            }
#pragma unroll
          for (int e = 0; e < UNR; ++e)
            if (k0 + e <= j) {
#pragma unroll
              for (int q = 0; q < NQ; ++q) und[q] += uk[e] * g[e][q];
            }
        }
        // Rdn -= RU * Undn_j ; norm -= Undn_j^2
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          Rdn[q] -= RU * und[q];
          nrm[q] -= und[q] * und[q];
        }
      }
    }

    // ------------------- coefficients: RUn := Un(0:j,0:j) * RUn (trmv)
    __syncthreads();
    for (int r = tid; r < j; r += GT) {
      T acc = 0;
      for (int q = r; q < j; ++q) acc += Un[r + (size_t)q * L] * sRUn[q];
      sX[r] = acc;
    }
    __syncthreads();
    T* ov = a.outV + (size_t)sig * L;
    int* orr = a.outR + (size_t)sig * L;
    for (int k = tid; k < L; k += GT) {
      const int id = k < j ? sInd[k] : -1;
      orr[k] = id;
      ov[k] = k < j ? sX[k] : T(0);
    }
    if (a.path && sig == 0 && j > 0 && j - 1 < a.path_cols) {
      // final coefficients in column j-1 (decomp/decomp.h:634-639)
      for (int k = tid; k < j; k += GT)
        if (sInd[k] >= 0) a.path[(size_t)(j - 1) * p + sInd[k]] = sX[k];
    }
    if (tid == 0) {
      int cnt = 0;
      while (cnt < j && sInd[cnt] >= 0) ++cnt;
      a.outCount[sig] = cnt;
    }
  }
}

template <typename T, int W, int NQ, bool ULDS>
size_t ormp_lds_bytes(int L) {
  size_t b = sizeof(T) * (3 * (size_t)L + 4) + sizeof(int) * (size_t)L;
  b = (b + 15) & ~size_t(15);
  b += sizeof(KeyIdx<T>) * W + 16;
  b = (b + 15) & ~size_t(15);
  if (ULDS) b += sizeof(T) * (size_t)L * L + 16;
  return b;
}

template <typename T, int W, int NQ, bool ULDS, bool MK>
int ormp_run(Ctx* ctx, const OmpArgs<T>& a0, int form) {
  OmpArgs<T> a = a0;
  const size_t lds = ormp_lds_bytes<T, W, NQ, ULDS>(a.L);
  auto kern = ormp_kernel<T, W, NQ, ULDS, MK>;
  const size_t ga_stride = MK ? (size_t)a.ldg * a.L : 0;  // masked Gram columns
  SPX_HIP(hipFuncSetAttribute((const void*)kern,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  const size_t slot = ULDS ? 0 : sizeof(T) * (size_t)a.L * a.L;
  hipStream_t st = ctx->stream;
  if (form == SPX_FORM_BATCHED) {
    int per_cu = 0;
    SPX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * W, lds));
    if (per_cu < 1) per_cu = 1;
    long blocks = (long)per_cu * ctx->num_cus;
    if (blocks > a.n) blocks = a.n;
    if (blocks < 1) blocks = 1;
    T* scratch = nullptr;
    int* queue = nullptr;
    int rc = SPX_OK;
    if (slot) {
      rc = ctx->scratch_get(0, slot * blocks, reinterpret_cast<void**>(&scratch));
      if (rc) return rc;
    }
    rc = ctx->scratch_get(1, sizeof(int), reinterpret_cast<void**>(&queue));
    if (rc) return rc;
    rc = zero_ints(st, queue, 1);
    if (rc) return rc;
    a.scratch = scratch;
    a.queue = queue;
    if (MK) {
      a.ga_stride = ga_stride;
      rc = ctx->scratch_get(21, sizeof(T) * ga_stride * blocks, reinterpret_cast<void**>(&a.ga));
      if (rc) return rc;
    }
    ScopedTimer tm(ctx, SPX_TIMER_ORMP);
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(64 * W), lds, st, a);
    SPX_HIP(hipGetLastError());
    return SPX_OK;
  }
  const size_t per_block = slot + sizeof(T) * ga_stride;
  long chunk = (long)(512ull * 1024 * 1024 / (per_block ? per_block : 1));
  if (chunk > 65535) chunk = 65535;
  if (chunk < 1) chunk = 1;
  const long alloc_chunk = chunk < a0.n ? chunk : a0.n;
  T* scratch = nullptr;
  if (slot) {
    int rc = ctx->scratch_get(0, slot * (alloc_chunk > 0 ? alloc_chunk : 1),
                              reinterpret_cast<void**>(&scratch));
    if (rc) return rc;
  }
  T* ga = nullptr;
  if (MK) {
    int rc = ctx->scratch_get(21, sizeof(T) * ga_stride * (alloc_chunk > 0 ? alloc_chunk : 1),
                              reinterpret_cast<void**>(&ga));
    if (rc) return rc;
  }
  for (long s0 = 0; s0 < a0.n; s0 += chunk) {
    const long cnt = (a0.n - s0) < chunk ? (a0.n - s0) : chunk;
    a = a0;
    a.queue = nullptr;
    a.scratch = scratch;
    a.n = (int)cnt;
    a.DtX = a0.DtX + (size_t)s0 * a0.ldd;
    a.norms = a0.norms + s0;
    if (a0.nL > 1) a.Ls = a0.Ls + s0;
    if (a0.nEps > 1) a.eps = a0.eps + s0;
    if (a0.nLambda > 1) a.lambda = a0.lambda + s0;
    a.outV = a0.outV + (size_t)s0 * a0.L;
    a.outR = a0.outR + (size_t)s0 * a0.L;
    a.outCount = a0.outCount + s0;
    if (MK) {
      a.mobs = a0.mobs + s0;
      a.mrows = a0.mrows + (size_t)s0 * a0.m;
      a.ga = ga;
      a.ga_stride = ga_stride;
    }
    hipLaunchKernelGGL(kern, dim3((unsigned)cnt), dim3(64 * W), lds, st, a);
    SPX_HIP(hipGetLastError());
  }
  return SPX_OK;
}

template <typename T, int W, bool MK>
int ormp_dispatch_nq_t(Ctx* ctx, const OmpArgs<T>& a, int form) {
  constexpr int VEC = Vec16o<T>::N;
  const int per = VEC * 64 * W;
  const int nq = (a.p + per - 1) / per;
  if (sizeof(T) * (size_t)a.L * a.L <= kUnLdsBytes) {
    if (nq <= 1) return ormp_run<T, W, 1, true, MK>(ctx, a, form);
    if (nq <= 2) return ormp_run<T, W, 2, true, MK>(ctx, a, form);
    if (nq <= 4) return ormp_run<T, W, 4, true, MK>(ctx, a, form);
    if (nq <= 8) return ormp_run<T, W, 8, true, MK>(ctx, a, form);
    if (nq <= 16) return ormp_run<T, W, 16, true, MK>(ctx, a, form);
    return -1;
  }
  if (nq <= 1) return ormp_run<T, W, 1, false, MK>(ctx, a, form);
  if (nq <= 2) return ormp_run<T, W, 2, false, MK>(ctx, a, form);
  if (nq <= 4) return ormp_run<T, W, 4, false, MK>(ctx, a, form);
  if (nq <= 8) return ormp_run<T, W, 8, false, MK>(ctx, a, form);
  if (nq <= 16) return ormp_run<T, W, 16, false, MK>(ctx, a, form);
  return -1;
}

template <typename T, int W>
int ormp_dispatch_nq(Ctx* ctx, const OmpArgs<T>& a, int form) {
  return a.mobs ? ormp_dispatch_nq_t<T, W, true>(ctx, a, form)
                  : ormp_dispatch_nq_t<T, W, false>(ctx, a, form);
}

}  // namespace

template <typename T>
int omp_solve(Ctx* ctx, const OmpArgs<T>& a, int form, int waves) {
  if (waves <= 0) waves = 4;
  static const int widths[] = {1, 2, 4, 16};
  for (int w : widths) {
    if (w < waves) continue;
    int rc;
    switch (w) {
      case 1: rc = ormp_dispatch_nq<T, 1>(ctx, a, form); break;
      case 2: rc = ormp_dispatch_nq<T, 2>(ctx, a, form); break;
      case 4: rc = ormp_dispatch_nq<T, 4>(ctx, a, form); break;
      default: rc = ormp_dispatch_nq<T, 16>(ctx, a, form); break;
    }
    if (rc >= 0) return rc;
  }
  set_error("omp: p=%d exceeds the largest supported dictionary", a.p);
  return SPX_ERR_UNSUP;
}

template <typename T>
bool ormp_fits(int w, int p) {
  const int per = Vec16o<T>::N * 64 * w;
  return (p + per - 1) / per <= 16;
}

template bool ormp_fits<float>(int, int);
template bool ormp_fits<double>(int, int);
template int omp_solve<float>(Ctx*, const OmpArgs<float>&, int, int);
template int omp_solve<double>(Ctx*, const OmpArgs<double>&, int, int);

}  // namespace spx
