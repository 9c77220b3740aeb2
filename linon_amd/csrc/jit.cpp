// Run-time field bodies: WGSL text -> HIP source -> hiprtc -> module.
#include <hip/hip_runtime.h>

#include "../../include/linon.h"
#include "context.hpp"

namespace linon {

void Context::release_jit() {
    if (jit_module) hipModuleUnload(jit_module);
    jit_module = nullptr;
    jit_fn = nullptr;
    jit_body.clear();
}

}  // namespace linon

extern "C" linon_status linon_set_field_wgsl(linon_context* ctx, const char* body) {
    linon::Context* c = reinterpret_cast<linon::Context*>(ctx);
    if (!c || !body) return LINON_ERR_ARGUMENT;
    return c->fail(LINON_ERR_UNSUPPORTED, "WGSL field bodies are not compiled by this build yet");
}
