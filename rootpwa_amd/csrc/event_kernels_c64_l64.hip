// event_kernels_c64_l64.hip -- instantiations of the fused event kernel
// (complex64 amplitudes, 64 lanes per event).
#include "event_kernel.h"

namespace rpwa_amd {

template <int K>
static EventKernel pick_r(int ranks)
{
	if (ranks <= 1)
		return ek::event_kernel<true, 64, K, 1>;
	if (ranks <= 2)
		return ek::event_kernel<true, 64, K, 2>;
	return ek::event_kernel<true, 64, K, 4>;
}

EventKernel event_family_c64_l64(int slots, int ranks)
{
	switch (slots) {
	case 1: return pick_r<1>(ranks);
	case 2: return pick_r<2>(ranks);
	case 3: return pick_r<3>(ranks);
	case 4: return pick_r<4>(ranks);
	case 5: return pick_r<5>(ranks);
	case 6: return pick_r<6>(ranks);
	case 7: return pick_r<7>(ranks);
	default: return pick_r<8>(ranks);
	}
}

}  // namespace rpwa_amd
