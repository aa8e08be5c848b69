// Test double of std_compat/memory.h (LibPressio's C++14 shims): make_unique.
#pragma once
#include <memory>

namespace compat {
using std::make_unique;
} // namespace compat
