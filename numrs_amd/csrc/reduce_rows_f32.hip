// Short-row reductions of binary32 elements (the kernel is in reduce_rows.hpp).
#include "reduce_rows.hpp"

namespace nrk {

template nrk_status run_rows<float>(const Chain &, const float *, size_t, size_t, double, const OutPtrs<float> &,
                                 const OutPtrs<float> &, bool, hipStream_t, const ExtScratch *);

}  // namespace nrk
