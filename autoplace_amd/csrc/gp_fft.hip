// FFT form of the Poisson solve (power-of-two grids). Not built yet: AUTO falls back to the GEMM form.
#include "gp_common.h"
#include "gp_poisson.h"

namespace ap {

bool Poisson::fft_supported(int) { return false; }
bool Poisson::fft_max_ok() { return false; }
int Poisson::fft_create() {
    set_error("poisson: FFT method not available in this build");
    return AP_ERR_UNSUPPORTED;
}
void Poisson::fft_destroy() {}
int Poisson::fft_solve(const float*, int, float*, float*, int, hipStream_t) {
    set_error("poisson: FFT method not available in this build");
    return AP_ERR_UNSUPPORTED;
}

}  // namespace ap
