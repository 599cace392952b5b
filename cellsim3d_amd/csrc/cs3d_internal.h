// cs3d_internal.h - declarations shared by the translation units of libcs3d.
#pragma once
#include <string>

#include "../../include/cs3d.h"

namespace cs3d {
extern thread_local std::string g_last_error;
cs3d_status fail(cs3d_status s, const std::string& msg);
}  // namespace cs3d
