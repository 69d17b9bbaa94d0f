// Library-internal entry points shared by the engine and the serving layer.
#pragma once

#include <string>

#include <hip/hip_runtime_api.h>

#include "titan_mi355x.h"

namespace titan_mi355x {

// A body laid out on the host, not yet on the device.
struct Prepared;

// Host half of titan_body_create under a given schedule: validates the
// descriptor and builds the layout on the host.  It reads no context state
// at all, so calls for different bodies may run on several threads at once,
// beside any other call.  Returns null with *rc and *err set on failure.
__attribute__((visibility("hidden"))) Prepared *prepare_body_sched(const titan_schedule &sched,
                                                                   const titan_body_desc *d, int *rc,
                                                                   std::string *err);
// Device half: places the prepared body (consumed) and returns its id in *id;
// on failure returns the error code with the context's last error set.
__attribute__((visibility("hidden"))) int commit_body_device(titan_ctx *ctx, Prepared *p, int32_t *id);
__attribute__((visibility("hidden"))) void free_prepared(Prepared *p);

// Records `ev` on the context's step stream, after the work enqueued so far.
__attribute__((visibility("hidden"))) int record_step_event(titan_ctx *ctx, hipEvent_t ev);

}  // namespace titan_mi355x
