// Kernels of the masked-signal LARS (lasso_mask, decomp/decomp.h:1466-1561):
// a signal with some rows unobserved runs coreLARS2 over the Gram columns of
// its observed rows, formed on the fly; compiled apart from the plain kernels
// so that the sets build in parallel.
#include "lars_impl.h"

namespace spx {

template <typename T>
int lars_masked_pass(Ctx* ctx, int w, bool fast, const LarsArgs<T>& a, bool persistent,
                     int* counter, long max_blocks) {
  return lars_pass_wt<T, 2>(ctx, w, fast, a, persistent, counter, max_blocks);
}

template int lars_masked_pass<float>(Ctx*, int, bool, const LarsArgs<float>&, bool, int*, long);
template int lars_masked_pass<double>(Ctx*, int, bool, const LarsArgs<double>&, bool, int*,
                                      long);

}  // namespace spx
