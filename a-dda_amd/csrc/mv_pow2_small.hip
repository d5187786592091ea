// mv_pow2_small.hip -- matvec pass kernels with compile-time FFT plans of lengths 32, 64 and 128.
#include "mv_kernels.h"

namespace adda {

using PlansPow2Small = PlanList<Fix<16, 8>, Fix<8, 16>, Fix<8, 4, 4>, Fix<4, 4, 8>, Fix<16, 4>, Fix<8, 8>, Fix<4, 16>,
	Fix<8, 4>, Fix<4, 8>>;

bool mv_find_pow2_small(const FftPlan2 &p, MvKernels &m) { return mv_find(p, m, PlansPow2Small{}); }
hipError_t mv_lds_pow2_small() { return mv_lds_all(PlansPow2Small{}); }

}  // namespace adda
