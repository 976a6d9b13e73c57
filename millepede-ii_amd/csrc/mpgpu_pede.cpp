// placeholder, replaced below
#include "mpgpu_internal.h"
extern "C" int mpgpu_run_steering(const char *, const char *, int32_t, mpgpu_run_stats *) { return 36; }
