// Element operations of the built-in subtask reductions, shared by the host fold (the runtime's
// reduce.cpp) and the HBM fold (reduce.hip) so that both give the same bits. The operations are
// those of the reference's pmReducer::ReduceMemories (pmReducer.cpp:593-703):
//   ADD a + b, MIN std::min(a, b), MAX std::max(a, b), PRODUCT a * b,
//   LOGICAL_AND (a && b), LOGICAL_OR (a || b), LOGICAL_XOR (a != b) (a comparison of the values,
//   as the reference writes it, not of their truth values),
//   BITWISE_AND / OR / XOR on the values converted to an unsigned integer of the type's width
//   (uint for float, ulong for double, pmReducer.cpp:408-425) and back.
// The reference converts floating values to unsigned integers with a C++ cast, which is only
// defined for values in [0, 2^w); here the conversion saturates (below 0 gives 0, at or above 2^w
// gives 2^w - 1) on host and device alike, which agrees with the reference wherever its cast is
// defined.
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#define RO_FN __host__ __device__ inline
#else
#define RO_FN inline
#endif

namespace redop {

enum Op { ADD, MIN, MAX, PRODUCT, LOGICAL_AND, BITWISE_AND, LOGICAL_OR, BITWISE_OR, LOGICAL_XOR, BITWISE_XOR, OP_COUNT };

template <typename T> struct Bits { typedef T type; };
template <> struct Bits<float> { typedef uint32_t type; };
template <> struct Bits<double> { typedef uint64_t type; };

template <typename T>
RO_FN typename Bits<T>::type to_bits(T v) { return (typename Bits<T>::type)v; }
template <>
RO_FN uint32_t to_bits<float>(float v) {
    return !(v > 0.f) ? 0u : v >= 4294967296.f ? 0xffffffffu : (uint32_t)v;
}
template <>
RO_FN uint64_t to_bits<double>(double v) {
    return !(v > 0.0) ? 0ull : v >= 18446744073709551616.0 ? ~0ull : (uint64_t)v;
}

template <int OP, typename T>
RO_FN T apply(T a, T b) {
    if (OP == ADD) return (T)(a + b);
    if (OP == MIN) return b < a ? b : a;
    if (OP == MAX) return a < b ? b : a;
    if (OP == PRODUCT) return (T)(a * b);
    if (OP == LOGICAL_AND) return (T)(a && b);
    if (OP == LOGICAL_OR) return (T)(a || b);
    if (OP == LOGICAL_XOR) return (T)(a != b);
    if (OP == BITWISE_AND) return (T)(to_bits(a) & to_bits(b));
    if (OP == BITWISE_OR) return (T)(to_bits(a) | to_bits(b));
    return (T)(to_bits(a) ^ to_bits(b));
}

}  // namespace redop
