// hr_ctx_internal.h -- what other parts of the library (the request server)
// use of a context beyond the public C ABI.
#pragma once

#include "../../include/happyray.h"
#include "hr_trace.h"

namespace hr {

// the context's current grid as the traversal kernels read it
TraceScene ctx_trace_scene(hr_ctx* c);
// device bytes the context holds
size_t ctx_device_bytes(hr_ctx* c);
// path tracing (depth vertices) from a 7-plane hit buffer of n rays with their
// primitive ids into rgb (3 floats per ray), with the context's materials,
// lights and normals, enqueued on stream s
hr_status ctx_path_trace(hr_ctx* c, const float* hits, const uint32_t* prims, uint32_t n, int depth, float* rgb,
                         hipStream_t s);

}  // namespace hr
