// Dispatch of a DataType code to a C++ value type.
#pragma once

#include <cstdint>

#include "ddj_internal.h"

namespace ddj {

// Native element type of a DataType (time_t = int64, char = int8).
template <typename F> auto dispatch_native(int dtype, F&& f) {
    switch (dtype) {
    case DDJ_D_INT: return f(int32_t{});
    case DDJ_D_TIME: return f(int64_t{});
    case DDJ_D_FLOAT: return f(float{});
    case DDJ_D_DOUBLE: return f(double{});
    case DDJ_D_SHORT: return f(int16_t{});
    case DDJ_D_CHAR: return f(int8_t{});
    default: fail(DDJ_EINVAL, "unknown data type " + std::to_string(dtype));
    }
}

// Signed integer type of a given width in bytes (the integer view of a column).
template <typename F> auto dispatch_int(int elem_size, F&& f) {
    switch (elem_size) {
    case 1: return f(int8_t{});
    case 2: return f(int16_t{});
    case 4: return f(int32_t{});
    case 8: return f(int64_t{});
    default: fail(DDJ_EINVAL, "bad element size " + std::to_string(elem_size));
    }
}

}  // namespace ddj
