// Short-row reductions of binary64 elements (the kernel is in reduce_rows.hpp).
#include "reduce_rows.hpp"

namespace nrk {

template nrk_status run_rows<double>(const Chain &, const double *, size_t, size_t, double, const OutPtrs<double> &,
                                 const OutPtrs<double> &, bool, hipStream_t, const ExtScratch *);

}  // namespace nrk
