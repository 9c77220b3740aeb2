// Run-time field bodies: WGSL text -> HIP source (wgsl.cpp) -> hiprtc -> code object -> module.
//
// This stands in for MainView::reload_shader (src/main_view.rs:649-690): the reference splices
// the body into the WGSL shader and lets naga/the driver compile it, reporting failures as a
// wgpu::Error string.  Here a failed translation or compilation returns LINON_ERR_COMPILE with
// the translator message or the hiprtc log, and the previously selected field stays active.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/linon.h"
#include "context.hpp"
#include "wgsl.hpp"

extern "C" const char linon_device_source[];

namespace linon {

namespace {

std::mutex g_cache_mu;
std::map<std::string, std::vector<char>> g_code_cache;  // translated source -> gfx950 code object

// Options mirror the static build: gfx950, -O3, packed-FP32 instructions disabled; a body
// that can fall off its end without returning is an error, as it is for naga.
const char* kOptions[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Werror=return-type",
                          "-Wno-unused-parameter", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
// added for the WGSL-accuracy build (linon_set_math(LINON_MATH_WGSL))
const char* kWgslMathOptions[] = {"-DLINON_FAST_MATH", "-fno-hip-fp32-correctly-rounded-divide-sqrt",
                                  "-fgpu-approx-transcendentals"};

bool compile_source(const std::string& src, bool wgsl_math, std::vector<char>& code, std::string& log) {
    const std::string key = (wgsl_math ? "W|" : "I|") + src;
    {
        std::lock_guard<std::mutex> lk(g_cache_mu);
        auto it = g_code_cache.find(key);
        if (it != g_code_cache.end()) {
            code = it->second;
            return true;
        }
    }
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "linon_user_field.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        log = "hiprtcCreateProgram failed";
        return false;
    }
    std::vector<const char*> opts(kOptions, kOptions + sizeof(kOptions) / sizeof(kOptions[0]));
    if (wgsl_math) opts.insert(opts.end(), kWgslMathOptions, kWgslMathOptions + 3);
    hiprtcResult rc = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    if (log_size > 1) {
        std::string l(log_size, '\0');
        hiprtcGetProgramLog(prog, &l[0]);
        l.resize(std::strlen(l.c_str()));
        log = l;
    }
    if (rc != HIPRTC_SUCCESS) {
        if (log.empty()) log = hiprtcGetErrorString(rc);
        hiprtcDestroyProgram(&prog);
        return false;
    }
    size_t n = 0;
    hiprtcGetCodeSize(prog, &n);
    code.assign(n, 0);
    hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    std::lock_guard<std::mutex> lk(g_cache_mu);
    g_code_cache[key] = code;
    return true;
}

bool build_user_kernel(const char* body, bool wgsl_math, std::vector<char>& code, std::string& msg,
                       std::string* translated_out) {
    std::string translated, err;
    if (!translate_wgsl_body(body, translated, err)) {
        msg = err;
        return false;
    }
    if (translated_out) *translated_out = translated;
    std::string src = std::string(linon_device_source) + "\n" + user_field_source(translated);
    std::string log;
    if (!compile_source(src, wgsl_math, code, log)) {
        msg = "hiprtc: " + log;
        return false;
    }
    return true;
}

void copy_out(const std::string& s, char* out, size_t cap) {
    if (!out || cap == 0) return;
    size_t n = std::min(cap - 1, s.size());
    std::memcpy(out, s.data(), n);
    out[n] = '\0';
}

}  // namespace

void Context::release_jit() {
    if (jit_module) {
        // a render on any stream may still be running the module's kernel
        (void)hipSetDevice(device);
        (void)hipDeviceSynchronize();
        (void)hipModuleUnload(jit_module);
    }
    jit_module = nullptr;
    jit_fn = nullptr;
    jit_body.clear();
}

}  // namespace linon

extern "C" {

linon_status linon_set_field_wgsl(linon_context* ctx, const char* body) {
    linon::Context* c = reinterpret_cast<linon::Context*>(ctx);
    if (!c) return LINON_ERR_ARGUMENT;
    c->err.clear();
    if (!body) return c->fail(LINON_ERR_ARGUMENT, "set_field_wgsl: null body");
    std::vector<char> code;
    std::string msg;
    const bool wgsl_math = c->math == LINON_MATH_WGSL;  // AUTO compiles WGSL bodies with IEEE arithmetic
    if (!linon::build_user_kernel(body, wgsl_math, code, msg, nullptr)) return c->fail(LINON_ERR_COMPILE, msg);
    if (hipSetDevice(c->device) != hipSuccess) return c->fail(LINON_ERR_HIP, "set_field_wgsl: hipSetDevice");
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    hipError_t e = hipModuleLoadData(&mod, code.data());
    if (e != hipSuccess) return c->fail(LINON_ERR_HIP, std::string("hipModuleLoadData: ") + hipGetErrorString(e));
    e = hipModuleGetFunction(&fn, mod, "linon_raycast_user");
    if (e != hipSuccess) {
        (void)hipModuleUnload(mod);
        return c->fail(LINON_ERR_HIP, std::string("hipModuleGetFunction: ") + hipGetErrorString(e));
    }
    // the previous kernel may still be running on the stream
    (void)hipDeviceSynchronize();
    c->release_jit();
    c->jit_module = mod;
    c->jit_fn = fn;
    c->jit_body = body;
    c->jit_wgsl_math = wgsl_math;
    c->field_id = -1;
    return LINON_OK;
}

linon_status linon_wgsl_translate(const char* body, char* out, size_t capacity, size_t* needed) {
    if (!body) return LINON_ERR_ARGUMENT;
    std::string translated, err;
    if (!linon::translate_wgsl_body(body, translated, err)) {
        linon::copy_out(err, out, capacity);
        if (needed) *needed = err.size() + 1;
        return LINON_ERR_COMPILE;
    }
    linon::copy_out(translated, out, capacity);
    if (needed) *needed = translated.size() + 1;
    return LINON_OK;
}

linon_status linon_wgsl_compile_check(const char* body, char* log, size_t capacity) {
    if (!body) return LINON_ERR_ARGUMENT;
    std::vector<char> code;
    std::string msg;
    if (!linon::build_user_kernel(body, false, code, msg, nullptr)) {
        linon::copy_out(msg, log, capacity);
        return LINON_ERR_COMPILE;
    }
    linon::copy_out("ok: " + std::to_string(code.size()) + " byte gfx950 code object", log, capacity);
    return LINON_OK;
}

}  // extern "C"
