// dispatch.h -- map a runtime line length onto the compiled FFT sizes.
#pragma once

#include <type_traits>

namespace ihg {

int set_error(const char *fmt, ...);

template <class F>
int dispatch_len(int L, F &&f)
{
    switch (L) {
    case 4: return f(std::integral_constant<int, 4>{});
    case 8: return f(std::integral_constant<int, 8>{});
    case 16: return f(std::integral_constant<int, 16>{});
    case 32: return f(std::integral_constant<int, 32>{});
    case 64: return f(std::integral_constant<int, 64>{});
    case 128: return f(std::integral_constant<int, 128>{});
    case 256: return f(std::integral_constant<int, 256>{});
    case 512: return f(std::integral_constant<int, 512>{});
    default: return set_error("unsupported FFT length %d", L);
    }
}

} // namespace ihg
