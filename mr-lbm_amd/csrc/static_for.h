// Compile-time loop: f(std::integral_constant<int, I>) for I in [0, N).
// Used wherever a loop index selects a population register, so the index is
// a constant expression and the population array never leaves registers.
#pragma once

#include <type_traits>
#include <utility>

namespace mrlbm {

template <typename F, int... Is>
__host__ __device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__host__ __device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// f(std::integral_constant<int, W>) for the one W in [0, N) equal to w (an
// else-if chain: exactly one branch runs on any path, so the compiler's
// memory-wait analysis never sees two branches' loads in flight into the
// same registers).
template <typename F, int... Ws>
__host__ __device__ __forceinline__ void dispatch_impl(int w, F&& f, std::integer_sequence<int, Ws...>) {
    (void)((w == Ws && (f(std::integral_constant<int, Ws>{}), true)) || ...);
}
template <int N, typename F>
__host__ __device__ __forceinline__ void dispatch(int w, F&& f) {
    dispatch_impl(w, f, std::make_integer_sequence<int, N>{});
}

}  // namespace mrlbm
