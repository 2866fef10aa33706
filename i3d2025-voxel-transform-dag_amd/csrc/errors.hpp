// Thread-local error message shared by every C-ABI entry point (tdag_last_error).
#pragma once
#include <string>

#include "tdag.h"

namespace tdag {
tdag_status setError(tdag_status status, const std::string& message);
}
