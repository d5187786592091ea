// mv_dyn.hip -- matvec pass kernels with run-time FFT plans (any length of ADDA's fftFit rules; RM = bound on the
// radix).
#include "mv_kernels.h"

namespace adda {

bool mv_find_dyn(const FftPlan2 &p, MvKernels &m)
{
	bool big = p.rmax > 8;
	for (int s = 0; s < p.nst; s++) big |= p.rad[s] == 7;  // radix 7 is compiled into the RM = 16 instances only
	if (big) mv_fill<Dyn<16>>(m);
	else mv_fill<Dyn<8>>(m);
	return true;
}

hipError_t mv_lds_dyn()
{
	hipError_t e;
	if ((e = mv_lds_t<double, Dyn<8>>()) || (e = mv_lds_t<double, Dyn<16>>()) || (e = mv_lds_t<float, Dyn<8>>()))
		return e;
	return mv_lds_t<float, Dyn<16>>();
}

}  // namespace adda
