// Kernels of the weighted-penalty LARS (coreLARS2W, decomp/decomp.h:1917-2107),
// compiled apart from the plain ones so that the two sets build in parallel.
#include "lars_impl.h"

namespace spx {

template <typename T>
int lars_weighted_pass(Ctx* ctx, int w, bool fast, const LarsArgs<T>& a, bool persistent,
                       int* counter, long max_blocks) {
  return lars_pass_wt<T, 1>(ctx, w, fast, a, persistent, counter, max_blocks);
}

template int lars_weighted_pass<float>(Ctx*, int, bool, const LarsArgs<float>&, bool, int*, long);
template int lars_weighted_pass<double>(Ctx*, int, bool, const LarsArgs<double>&, bool, int*,
                                        long);

}  // namespace spx
