// ft_internal.hpp -- shared host-side helpers of libftreinit (error state).
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

namespace ft {

// Per-thread message of the last failure (ft_last_error).
inline std::string& error_slot()
{
    static thread_local std::string msg;
    return msg;
}

inline int set_error(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    error_slot() = buf;
    return code;
}

}  // namespace ft
