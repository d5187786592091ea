// mv_mixed.hip -- matvec pass kernels with compile-time FFT plans of the mixed-radix lengths of the workloads:
// 96 = 2^5*3, 108 = 2^2*3^3, 640 = 2^7*5.
#include "mv_kernels.h"

namespace adda {

using PlansMixed = PlanList<Fix<8, 4, 3>, Fix<16, 2, 3>, Fix<3, 8, 4>, Fix<4, 3, 3, 3>, Fix<3, 3, 3, 4>,
	Fix<16, 8, 5>, Fix<8, 16, 5>, Fix<5, 16, 8>, Fix<16, 5, 8>, Fix<8, 5, 16>>;

bool mv_find_mixed(const FftPlan2 &p, MvKernels &m) { return mv_find(p, m, PlansMixed{}); }
hipError_t mv_lds_mixed() { return mv_lds_all(PlansMixed{}); }

}  // namespace adda
