// Error reporting and version of the pxff C ABI.
#include <cstring>
#include <string>

#include "pxff_common.h"

namespace pxff {
static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
}  // namespace pxff

extern "C" int pxff_last_error(char *buf, size_t len) {
    if (!buf || len == 0) return PXFF_EINVAL;
    const std::string &m = pxff::g_last_error;
    size_t n = m.size() < len - 1 ? m.size() : len - 1;
    std::memcpy(buf, m.data(), n);
    buf[n] = '\0';
    return 0;
}

extern "C" const char *pxff_version(void) { return "pxff 0.1.0 gfx950"; }
