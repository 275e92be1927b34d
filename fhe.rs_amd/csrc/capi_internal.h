// Handle definitions and error plumbing shared by the C ABI translation units.
#pragma once

#include <string>

#include "fhehip_internal.h"

struct fhehip_context {
    std::shared_ptr<fhehip::DeviceContext> c;
};

namespace fhehip {
int fail(int code, const std::string& msg);
}

#define FHEHIP_API_BEGIN try {
#define FHEHIP_API_END                                                  \
    return 0;                                                           \
    }                                                                   \
    catch (const ::fhehip::Error& e) {                                  \
        return ::fhehip::fail(e.code, e.what());                        \
    }                                                                   \
    catch (const std::bad_alloc& e) {                                   \
        return ::fhehip::fail(-5, "host allocation failed");            \
    }                                                                   \
    catch (const std::exception& e) {                                   \
        return ::fhehip::fail(-1, e.what());                            \
    }                                                                   \
    catch (...) {                                                       \
        return ::fhehip::fail(-6, "unknown error");                     \
    }
