// Instantiates the tiled MFMA GEMM (qs_gemm_impl.h) for one compute type.
#include "qs_gemm_impl.h"

namespace qs {
hipError_t launch_gemm_c32(const GemmArgs &g, hipStream_t s) { return launch_gemm_typed<float, true>(g, s); }
}  // namespace qs
