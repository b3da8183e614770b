// LARS-Lasso solver kernel: coreLARS2 semantics (decomp/decomp.h:1706-1915).
//
// One group of W wavefronts (one workgroup) solves one signal at a time:
//   - the correlation vector DtR (p) and the active-atom flags live in LDS;
//   - the correlation change A = G(:,active) * u is accumulated in registers
//     (each thread owns NQ chunks of VEC = 16 bytes of rows), streaming the
//     active Gram columns from the resident G (L2 / MALL);
//   - the explicit inverse of the active Gram block, invGs (s x s, stored full
//     and exactly symmetric), lives in a per-workgroup HBM slot, double
//     buffered for the downdate;
//   - arg-min / arg-max searches are deterministic (value, index) wavefront
//     reductions with the reference's tie rules (first index for i?amax and
//     the custom iamin; last index for the zero-crossing search).
// Batched form: persistent grid that pulls signal indices from an atomic
// counter.  Plain form: one workgroup per signal over a chunk of signals.
#include <climits>

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

template <typename T, int W, int NQ>
__global__ __launch_bounds__(64 * W) void lars_kernel(LarsArgs<T> a) {
  constexpr int GT = 64 * W;
  constexpr int VEC = Vec16<T>::N;
  constexpr int PT = VEC * GT * NQ;  // rows covered by the group
  typedef typename Vec16<T>::type V;
  const T INF = dev_inf<T>();

  extern __shared__ __attribute__((aligned(16))) char smem[];
  const int L = a.L;
  const int p = a.p;
  const int ldg = a.ldg;
  T* sDtR = reinterpret_cast<T*>(smem);  // PT
  T* sCoef = sDtR + PT;                   // L
  T* sU = sCoef + L;                      // L
  T* sW = sU + L;                         // L
  T* sGs = sW + L;                        // L
  T* sV = sGs + L;                        // L
  int* sInd = reinterpret_cast<int*>(sV + L);  // L
  int* sTmpI = sInd + L;                       // L
  unsigned char* sAct = reinterpret_cast<unsigned char*>(sTmpI + L);  // PT
  KeyIdx<T>* sRed = reinterpret_cast<KeyIdx<T>*>(
      (reinterpret_cast<uintptr_t>(sAct + PT) + 15) & ~uintptr_t(15));
  T* sRed3 = reinterpret_cast<T*>(sRed + W);  // 3W
  int* sSig = reinterpret_cast<int*>(sRed3 + 3 * W);

  const int tid = threadIdx.x;
  T* Mbuf0 = a.scratch + (size_t)blockIdx.x * 2 * L * L;
  T* Mbuf1 = Mbuf0 + (size_t)L * L;
  const int Lm = L < p ? L : p;
  const T diag = a.diag;

  for (int round = 0;; ++round) {
    int sig;
    if (a.queue) {
      __syncthreads();
      if (tid == 0) *sSig = atomicAdd(a.queue, 1);
      __syncthreads();
      sig = *sSig;
    } else {
      sig = round == 0 ? (int)blockIdx.x : a.n;
    }
    if (sig >= a.n) break;

    // ---------------------------------------------------------- set-up
    const T* dtx = a.DtX + (size_t)sig * a.ldd;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int j0 = VEC * (tid + GT * q);
#pragma unroll
      for (int e = 0; e < VEC; ++e) {
        const int j = j0 + e;
        sDtR[j] = j < p ? dtx[j] : T(0);
        sAct[j] = j < p ? 0 : 1;
      }
    }
    for (int j = tid; j < L; j += GT) {
      sCoef[j] = 0;
      sInd[j] = -1;
    }
    T normX = a.norms[sig];
    const T lam = a.lambda[a.nlam > 1 ? sig : 0];
    __syncthreads();

    // first atom: i?amax of DtR, or the plain max when pos
    int currentInd;
    {
      T key = INF;
      int idx = INT_MAX;
      T pay = 0;
#pragma unroll
      for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
          const int j = VEC * (tid + GT * q) + e;
          if (j < p) {
            const T d = sDtR[j];
            const T k = a.pos ? -d : -fabs(d);
            if (key_better(k, j, key, idx)) {
              key = k;
              idx = j;
            }
          }
        }
      group_argmin<T, W>(key, idx, pay, sRed);
      currentInd = idx == INT_MAX ? 0 : idx;
    }
    int s_active = 0;
    int iter = 0, admissions = 0;
    bool empty = false;
    {
      const T dc = sDtR[currentInd];
      if (a.mode == SPX_PENALTY && fabs(dc) < lam) empty = true;
      if (a.mode == SPX_L2ERROR && normX < lam) empty = true;
    }

    if (!empty) {
      bool newAtom = true;
      T thrs = 0;
      T* Mc = Mbuf0;
      T* Mn = Mbuf1;
      for (int i = 0; i < Lm; ++i) {
        ++iter;
        if (newAtom) {
          // ---------------------------------------------- admit an atom
          const int c = currentInd;
          ++admissions;
          ++s_active;
          if (tid == 0) {
            sInd[i] = c;
            sAct[c] = 1;
          }
          __syncthreads();
          const T* gcol = a.G + (size_t)c * ldg;
          for (int j = tid; j <= i; j += GT) {
            const int r = sInd[j];
            sGs[j] = gcol[r] + (r == c ? diag : T(0));
          }
          __syncthreads();
          if (i == 0) {
            if (tid == 0) Mc[0] = T(1) / sGs[0];
          } else {
            // v = invGs(i x i) * gs(0:i)
            for (int r = tid; r < i; r += GT) {
              T acc = 0;
              for (int q = 0; q < i; ++q) acc += Mc[r + (size_t)q * L] * sGs[q];
              sV[r] = acc;
            }
            __syncthreads();
            T dv = 0, z1 = 0, z2 = 0;
            for (int r = tid; r < i; r += GT) dv += sV[r] * sGs[r];
            group_sum3<T, W>(dv, z1, z2, sRed3);
            const T schur = T(1) / (sGs[i] - dv);
            // Schur-complement update of the inverse (decomp/decomp.h:1756-1770)
            for (int r = tid; r < i; r += GT) {
              const T vr = sV[r];
              for (int q = 0; q < i; ++q) {
                const T vq = sV[q];
                // symmetric rounding: element (r,q) and (q,r) get the
                // identical update v[min] * (schur * v[max])
                const T upd = r <= q ? vr * (schur * vq) : vq * (schur * vr);
                Mc[r + (size_t)q * L] += upd;
              }
              const T off = -schur * vr;
              Mc[r + (size_t)i * L] = off;
              Mc[i + (size_t)r * L] = off;
            }
            if (tid == 0) Mc[i + (size_t)i * L] = schur;
            __syncthreads();
          }
        }

        const int s = i + 1;
        // ------------------------------------------ direction u = invGs*w
        for (int j = tid; j < s; j += GT) sW[j] = sDtR[sInd[j]] > 0 ? T(1) : T(-1);
        __syncthreads();
        for (int r = tid; r < s; r += GT) {
          T acc = 0;
          for (int q = 0; q < s; ++q) acc += Mc[r + (size_t)q * L] * sW[q];
          sU[r] = acc;
        }
        __syncthreads();

        // ------------------ zero-crossing step, polynomial coefficients
        T step_max;
        int first_zero;
        T coeff1 = 0, coeff2 = 0;
        {
          T key = INF;
          int idx = INT_MAX;
          T pay = 0, z = 0;
          for (int j = tid; j < s; j += GT) {
            const T uj = sU[j];
            const T ratio = -sCoef[j] / uj;
            // sequential scan keeps the LAST j on ties: encode index L - j
            if (ratio > 0 && key_better(ratio, L - j, key, idx)) {
              key = ratio;
              idx = L - j;
            }
            const T dj = sDtR[sInd[j]];
            coeff1 += dj > 0 ? uj : -uj;
            coeff2 += dj * uj;
          }
          group_argmin<T, W>(key, idx, pay, sRed);
          group_sum3<T, W>(coeff1, coeff2, z, sRed3);
          step_max = key;
          first_zero = idx == INT_MAX ? -1 : L - idx;
        }
        const T cc = fabs(sDtR[sInd[0]]);

        // ------------------------------- A = (G + diag I)(:,active) * u
        V Areg[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) Areg[q] = V(0);
        for (int k = 0; k < s; ++k) {
          const T uk = sU[k];
          const T* col = a.G + (size_t)sInd[k] * ldg;
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const int j0 = VEC * (tid + GT * q);
            if (j0 < ldg) {
              const V g = *reinterpret_cast<const V*>(col + j0);
              Areg[q] += uk * g;
            }
          }
        }
        for (int k = 0; k < s; ++k) {
          const int r = sInd[k];
          const int chunk = r / VEC;
          if (chunk % GT == tid) {
            const int q = chunk / GT;
            const T add = diag * sU[k];
#pragma unroll
            for (int qq = 0; qq < NQ; ++qq)
              if (qq == q) Areg[qq][r % VEC] += add;
          }
        }

        // ------------------------------- next atom: custom iamin over 2p
        T bestk = INF, bestv = INF;
        int besti = INT_MAX;
#pragma unroll
        for (int q = 0; q < NQ; ++q)
#pragma unroll
          for (int e = 0; e < VEC; ++e) {
            const int j = VEC * (tid + GT * q) + e;
            if (j < p && !sAct[j]) {
              const T Aj = Areg[q][e];
              const T d = sDtR[j];
              T c1 = (Aj < INF && Aj > T(-1)) ? (d + cc) / (T(1) + Aj) : INF;
              if (a.pos) c1 = INF;
              const T c2 = (Aj < INF && Aj < T(1)) ? (cc - d) / (T(1) - Aj) : INF;
              if (key_better(T(fabs(c1)), j, bestk, besti)) {
                bestk = fabs(c1);
                besti = j;
                bestv = c1;
              }
              if (key_better(T(fabs(c2)), p + j, bestk, besti)) {
                bestk = fabs(c2);
                besti = p + j;
                bestv = c2;
              }
            }
          }
        group_argmin<T, W>(bestk, besti, bestv, sRed);
        T step = besti == INT_MAX ? INF : bestv;
        const int nextInd = besti == INT_MAX ? 0 : besti % p;

        const T coeff3 = normX - lam;
        T step_max2;
        if (a.mode == SPX_PENALTY) {
          step_max2 = cc - lam;
        } else if (a.mode == SPX_L2ERROR) {
          const T delta = coeff2 * coeff2 - coeff1 * coeff3;
          step_max2 = delta < 0 ? INF : (coeff2 - sqrt(delta)) / coeff1;
          step_max2 = step_max2 < cc ? step_max2 : cc;
        } else {
          step_max2 = coeff1 < 0 ? INF : (lam - thrs) / coeff1;
          step_max2 = step_max2 < cc ? step_max2 : cc;
        }
        step = step < step_max2 ? step : step_max2;
        step = step < step_max ? step : step_max;
        if (step == INF) break;

        // -------------------------------------------------- updates
        for (int j = tid; j < s; j += GT) {
          T cj = sCoef[j] + step * sU[j];
          if (a.pos && cj < 0) cj = 0;
          sCoef[j] = cj;
        }
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const int j0 = VEC * (tid + GT * q);
#pragma unroll
          for (int e = 0; e < VEC; ++e) sDtR[j0 + e] -= step * Areg[q][e];
        }
        normX += coeff1 * step * step - 2 * coeff2 * step;
        thrs += step * coeff1;
        __syncthreads();

        if (step == step_max) {
          // ------------------- drop atom first_zero (decomp/decomp.h:1875-1903)
          const int fz = first_zero;
          const int dropped = sInd[fz];
          const T schur = Mc[fz + (size_t)fz * L];
          for (int r = tid; r < i; r += GT) {
            const int r2 = r + (r >= fz);
            sV[r] = Mc[r2 + (size_t)fz * L];
          }
          __syncthreads();
          const T alpha = T(-1) / schur;
          for (int r = tid; r < i; r += GT) {
            const int r2 = r + (r >= fz);
            const T ur = sV[r];
            for (int q = 0; q < i; ++q) {
              const int q2 = q + (q >= fz);
              const T uq = sV[q];
              const T upd = r <= q ? ur * (alpha * uq) : uq * (alpha * ur);
              Mn[r + (size_t)q * L] = Mc[r2 + (size_t)q2 * L] + upd;
            }
          }
          for (int j = tid; j < i; j += GT) {
            const int j2 = j + (j >= fz);
            sTmpI[j] = sInd[j2];
            sGs[j] = sCoef[j2];
          }
          __syncthreads();
          if (tid == 0) sAct[dropped] = 0;
          for (int j = tid; j < i; j += GT) {
            sInd[j] = sTmpI[j];
            sCoef[j] = sGs[j];
          }
          if (tid == 0) {
            sInd[i] = -1;
            sCoef[i] = 0;
          }
          __syncthreads();
          T* t = Mc;
          Mc = Mn;
          Mn = t;
          --s_active;
          newAtom = false;
          i = i - 2;
          // an emptied active set has no direction: end the path (the
          // reference would read ind[-1] here)
          if (s_active == 0) break;
        } else {
          newAtom = true;
          currentInd = nextInd;
        }
        if ((iter >= a.length_path - 1) || fabs(step) < 1e-15 ||
            step == step_max2 || (normX < 1e-15) || (i == (Lm - 1)) ||
            (a.mode == SPX_L2ERROR && normX - lam < 1e-15) ||
            (a.mode == SPX_L1COEFFS && (lam - thrs < 1e-15)))
          break;
      }
    }

    // ------------------------------------------------------- outputs
    __syncthreads();
    T* ov = a.outV + (size_t)sig * L;
    int* orr = a.outR + (size_t)sig * L;
    for (int j = tid; j < L; j += GT) {
      ov[j] = sCoef[j];
      orr[j] = sInd[j];
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

template <typename T, int W, int NQ>
size_t lars_lds_bytes(int L) {
  constexpr int GT = 64 * W;
  constexpr int VEC = Vec16<T>::N;
  constexpr int PT = VEC * GT * NQ;
  size_t b = sizeof(T) * (PT + 5 * (size_t)L) + sizeof(int) * 2 * (size_t)L + PT;
  b = (b + 15) & ~size_t(15);
  b += sizeof(KeyIdx<T>) * W + sizeof(T) * 3 * W + 16;
  return b;
}

template <typename T, int W, int NQ>
int lars_run(Ctx* ctx, const LarsArgs<T>& a0, int form) {
  LarsArgs<T> a = a0;
  const size_t lds = lars_lds_bytes<T, W, NQ>(a.L);
  if (lds > 160 * 1024) {
    set_error("lasso: L=%d, p=%d need %zu bytes of LDS per signal", a.L, a.p, lds);
    return SPX_ERR_UNSUP;
  }
  auto kern = lars_kernel<T, W, NQ>;
  SPX_HIP(hipFuncSetAttribute((const void*)kern,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds));
  const size_t slot = 2 * sizeof(T) * (size_t)a.L * a.L;
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
    int rc = ctx->scratch_get(0, slot * blocks, reinterpret_cast<void**>(&scratch));
    if (rc) return rc;
    rc = ctx->scratch_get(1, sizeof(int), reinterpret_cast<void**>(&queue));
    if (rc) return rc;
    rc = zero_ints(st, queue, 1);
    if (rc) return rc;
    a.scratch = scratch;
    a.queue = queue;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(64 * W), lds, st, a);
    SPX_HIP(hipGetLastError());
    return SPX_OK;
  }
  // plain form: one workgroup per signal, chunks bounded by scratch size
  long chunk = (long)(512ull * 1024 * 1024 / (slot ? slot : 1));
  if (chunk > 65535) chunk = 65535;
  if (chunk < 1) chunk = 1;
  T* scratch = nullptr;
  const long alloc_chunk = chunk < a0.n ? chunk : a0.n;
  int rc = ctx->scratch_get(0, slot * (alloc_chunk > 0 ? alloc_chunk : 1),
                            reinterpret_cast<void**>(&scratch));
  if (rc) return rc;
  for (long s0 = 0; s0 < a0.n; s0 += chunk) {
    const long cnt = (a0.n - s0) < chunk ? (a0.n - s0) : chunk;
    a = a0;
    a.queue = nullptr;
    a.scratch = scratch;
    a.n = (int)cnt;
    a.DtX = a0.DtX + (size_t)s0 * a0.ldd;
    a.norms = a0.norms + s0;
    if (a0.nlam > 1) a.lambda = a0.lambda + s0;
    a.outV = a0.outV + (size_t)s0 * a0.L;
    a.outR = a0.outR + (size_t)s0 * a0.L;
    a.outCount = a0.outCount + s0;
    if (a0.stats) a.stats = a0.stats + 2 * s0;
    hipLaunchKernelGGL(kern, dim3((unsigned)cnt), dim3(64 * W), lds, st, a);
    SPX_HIP(hipGetLastError());
  }
  return SPX_OK;
}

template <typename T, int W>
int lars_dispatch_nq(Ctx* ctx, const LarsArgs<T>& a, int form) {
  constexpr int VEC = Vec16<T>::N;
  const int per = VEC * 64 * W;
  const int nq = (a.p + per - 1) / per;
  if (nq <= 1) return lars_run<T, W, 1>(ctx, a, form);
  if (nq <= 2) return lars_run<T, W, 2>(ctx, a, form);
  if (nq <= 4) return lars_run<T, W, 4>(ctx, a, form);
  if (nq <= 8) return lars_run<T, W, 8>(ctx, a, form);
  if (nq <= 16) return lars_run<T, W, 16>(ctx, a, form);
  return -1;  // does not fit this group width
}

}  // namespace

// Group width: explicit (waves_per_signal), else the autotuner's choice for
// the batched form, else 4 wavefronts for the plain form; widened until the
// group covers p.
template <typename T>
int lars_solve(Ctx* ctx, const LarsArgs<T>& a, int form, int waves) {
  if (waves <= 0) waves = form == SPX_FORM_PLAIN ? 4 : tuned_waves(ctx, 0, a.p, sizeof(T));
  static const int widths[] = {1, 2, 4, 16};
  for (int w : widths) {
    if (w < waves) continue;
    int rc;
    switch (w) {
      case 1: rc = lars_dispatch_nq<T, 1>(ctx, a, form); break;
      case 2: rc = lars_dispatch_nq<T, 2>(ctx, a, form); break;
      case 4: rc = lars_dispatch_nq<T, 4>(ctx, a, form); break;
      default: rc = lars_dispatch_nq<T, 16>(ctx, a, form); break;
    }
    if (rc >= 0) return rc;
  }
  set_error("lasso: p=%d exceeds the largest supported dictionary", a.p);
  return SPX_ERR_UNSUP;
}

template int lars_solve<float>(Ctx*, const LarsArgs<float>&, int, int);
template int lars_solve<double>(Ctx*, const LarsArgs<double>&, int, int);

}  // namespace spx
