// static_for.h -- compile-time loop for device code: fn is called with
// std::integral_constant<int, L> for L in [B, E), so that table lookups
// indexed by L fold into constants and terms with a zero coefficient drop
// out at compile time.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace sol {

template <int B, int E, typename Fn>
__device__ __forceinline__ void static_for(Fn &&fn)
{
    if constexpr (B < E) {
        fn(std::integral_constant<int, B>{});
        static_for<B + 1, E>(fn);
    }
}

}  // namespace sol
