// Error reporting, version and workspace sizes of the pxff C ABI.
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

// Sizes of the caller-owned buffers that are neither inputs nor outputs of
// their entry point, from the same size calls those entry points document.
extern "C" int pxff_workspace_size(int kind, const void *params, int64_t n0, int64_t n1, int64_t *bytes) {
    PXFF_REQUIRE(bytes != nullptr, "null bytes");
    PXFF_REQUIRE(n0 >= 0 && n1 >= 0, "negative size");
    int rc = 0;
    switch (kind) {
    case PXFF_WS_NLIST:
        // bin_count [nbins], bin_start [nbins + 1], atom_bin [natoms],
        // wrap [3 natoms], bin_atoms [natoms], all int32
        *bytes = 4 * (2 * n1 + 1) + 4 * 5 * n0;
        return 0;
    case PXFF_WS_SO4_Z: {
        PXFF_REQUIRE(params != nullptr, "null params");
        int64_t len = 0;
        if ((rc = pxff_so4_zlen(static_cast<const pxff_so4_params *>(params)->twol, n0, &len))) return rc;
        *bytes = 8 * len;
        return 0;
    }
    case PXFF_WS_SO4_TANGENT: {
        PXFF_REQUIRE(params != nullptr, "null params");
        int nu = 0, nhalf = 0, nz = 0, nb = 0;
        if ((rc = pxff_so4_sizes(static_cast<const pxff_so4_params *>(params)->twol, &nu, &nhalf, &nz, &nb)))
            return rc;
        *bytes = ((n0 + 63) / 64) * 64 * (int64_t)nhalf * 16;
        return 0;
    }
    case PXFF_WS_SO3_PAIRS:
        *bytes = 8 * (256 * 64 + 1 + n1);
        return 0;
    case PXFF_WS_NLIST_SLOTS:
        // slot_j int64, slot_T 3 x int32, slot_r 3 x double per entry
        *bytes = (8 + 12 + 24) * n0 * n1;
        return 0;
    case PXFF_WS_SO4_UI:
        *bytes = 8 * 6 * n1;
        return 0;
    case PXFF_WS_SO3_TABLE: {
        PXFF_REQUIRE(params != nullptr, "null params");
        int nk = 0;
        int64_t len = 0;
        if ((rc = pxff_so3_radial_table_size(static_cast<const pxff_so3_params *>(params), &nk, &len))) return rc;
        *bytes = 8 * len;
        return 0;
    }
    default:
        pxff::set_error("unknown workspace kind " + std::to_string(kind));
        return PXFF_EINVAL;
    }
}
