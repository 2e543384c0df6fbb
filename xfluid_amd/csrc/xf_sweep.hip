// xf_sweep.hip -- pipelined MIC(0) triangular sweeps (persistent line tiles).
// Placeholder schedule selector: returns false so the caller falls back to the
// per-hyperplane launches until the pipelined kernel is enabled.
#include "xf_internal.h"

template <class T>
bool sweep_pipelined(const Geo& g, bool forward, const uint16_t* code, const T* inv, const T* in, T* out,
                     const PcgScalars* S, int* flags, hipStream_t st) {
    (void)g; (void)forward; (void)code; (void)inv; (void)in; (void)out; (void)S; (void)flags; (void)st;
    return false;
}
template bool sweep_pipelined<double>(const Geo&, bool, const uint16_t*, const double*, const double*, double*,
                                      const PcgScalars*, int*, hipStream_t);
template bool sweep_pipelined<float>(const Geo&, bool, const uint16_t*, const float*, const float*, float*,
                                     const PcgScalars*, int*, hipStream_t);
