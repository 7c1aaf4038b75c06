#include "device_context.h"

#include <cstdlib>
#include <mutex>

namespace uepm::pseudopotential {

void throw_on_error(int status, const char* where) {
    if (status == UEPM_OK) return;
    const std::string msg = std::string(where) + ": " + uepm_last_error();
    switch (status) {
        case UEPM_ERR_INVALID_ARG: throw std::invalid_argument(msg);
        case UEPM_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
        default: throw std::runtime_error(msg);
    }
}

uepm_ctx* device_context() {
    static std::once_flag once;
    static uepm_ctx* ctx = nullptr;
    static int status = UEPM_OK;
    static std::string err;
    std::call_once(once, [] {
        const char* env = std::getenv("UEPM_DEVICE");
        const int dev = env ? std::atoi(env) : 0;
        status = uepm_ctx_create(dev, &ctx);
        if (status != UEPM_OK) err = uepm_last_error();
    });
    if (status != UEPM_OK) throw std::runtime_error("uepm device context: " + err);
    return ctx;
}

}  // namespace uepm::pseudopotential
