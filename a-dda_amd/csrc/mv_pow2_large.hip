// mv_pow2_large.hip -- matvec pass kernels with compile-time FFT plans of lengths 256 and 512.
#include "mv_kernels.h"

namespace adda {

using PlansPow2Large = PlanList<Fix<16, 16>, Fix<8, 8, 4>, Fix<4, 8, 8>, Fix<16, 4, 4>, Fix<16, 8, 4>, Fix<8, 8, 8>,
	Fix<4, 8, 16>>;

bool mv_find_pow2_large(const FftPlan2 &p, MvKernels &m) { return mv_find(p, m, PlansPow2Large{}); }
hipError_t mv_lds_pow2_large() { return mv_lds_all(PlansPow2Large{}); }

}  // namespace adda
