// Thread-local error message behind mph_last_error().
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

namespace mph {

std::string& last_error();

inline int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    last_error() = buf;
    return code;
}

}  // namespace mph
