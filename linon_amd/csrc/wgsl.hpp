// WGSL field-body translation (wgsl.cpp).
#pragma once
#include <string>

namespace linon {

// Translates a WGSL function body (the text linon splices into field_function) into the body
// of a C++ device functor over lw:: types.  On failure returns false with "field_function:
// line:col: message" in err.
bool translate_wgsl_body(const std::string& body, std::string& out, std::string& err);

// Full hiprtc translation unit tail for a translated body: the UserField functor and the
// extern "C" kernel linon_raycast_user.
std::string user_field_source(const std::string& translated);

}  // namespace linon
