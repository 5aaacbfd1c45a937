// Internal helpers shared by the engine and the drop-in ABI shims.
#pragma once
#include "../../include/seismic_xcorr.h"

namespace sxc {
void set_error(const char* fmt, ...);
void clear_error();
const char* last_error();
int validate_desc(const sxc_plan_desc& d);
// Maps the reference ABI arguments to a plan description (lags -shift+1..shift).
int reference_desc(int n_events, int event_length, int shift, int padded, float threshold,
                   sxc_plan_desc* d);
// Full drop-in call on host buffers; errors are reported via last_error().
void reference_call(float* events, int n_events, int event_length, int shift, int padded,
                    float threshold, float* vp, int* lp, float* vn, int* ln);
int warm_up();
}  // namespace sxc

extern "C" {
// Exported for the drop-in shims (same library namespace).
void sxc_reference_call(float* events, int n_events, int event_length, int shift, int padded,
                        float threshold, float* vp, int* lp, float* vn, int* ln);
int sxc_warm_up(void);
}
