// Register-resident decode kernels for the 12 IEEE 802.11 matrices (f32 and
// int8), each instantiated at its own lifting size so every shift is an
// immediate.
#include <utility>

#include "ldpc_fast.h"

namespace syn {

namespace {

template <int... M>
const void *pick(int m, bool i8, std::integer_sequence<int, M...>)
{
    static const void *const f32[] = {
        (const void *)fast_f32<SYN_KIND_WIFI0 + M, Graph<SYN_KIND_WIFI0 + M>::Z>...};
    static const void *const i8k[] = {
        (const void *)fast_i8<SYN_KIND_WIFI0 + M, Graph<SYN_KIND_WIFI0 + M>::Z>...};
    return i8 ? i8k[m] : f32[m];
}

}  // namespace

const void *fast_kernel_wifi(int kind, bool i8)
{
    const int m = kind - SYN_KIND_WIFI0;
    if (m < 0 || m >= SYN_WIFI_COUNT) return nullptr;
    return pick(m, i8, std::make_integer_sequence<int, SYN_WIFI_COUNT>{});
}

}  // namespace syn
