// Internal (not part of the C ABI): records the calling thread's error message that
// fz_last_error() returns, and passes the code through.
#pragma once

int fzi_fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
