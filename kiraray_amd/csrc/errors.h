// errors.h - thread-local error reporting for the C ABI.
//
// The reference logs and exits on every failure (src/util/check.h:8-42);
// the ABI instead returns a negative krr_status and keeps a message that
// krr_last_error() returns.
#pragma once

#include <cstdarg>
#include <cstdio>
#include <string>

#include "krr_wavefront.h"

namespace krr {

std::string &lastErrorStorage();

inline int fail(int code, const char *fmt, ...) {
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	std::vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	lastErrorStorage() = buf;
	return code;
}

inline void clearError() { lastErrorStorage().clear(); }

} // namespace krr
