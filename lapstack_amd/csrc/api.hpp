// api.hpp -- the C-ABI boundary's exception fence: every extern "C" entry
// point runs its body inside LSH_API_BEGIN / LSH_API_END, so a C++ exception
// (a HIP error, a size mismatch) comes back as a library-owned error string
// and never unwinds into the caller (the convention of the reference's
// lapstack-raw ffi.rs:8-9, 82-90).
#pragma once
#include <string>

namespace lsh {
char *make_err(const std::string &msg);
}

#define LSH_API_BEGIN try {
#define LSH_API_END                                    \
    }                                                  \
    catch (const std::exception &e) {                  \
        return ::lsh::make_err(e.what());              \
    }                                                  \
    catch (...) {                                      \
        return ::lsh::make_err("unknown C++ exception"); \
    }                                                  \
    return nullptr;
