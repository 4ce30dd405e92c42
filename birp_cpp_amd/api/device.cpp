// device.cpp -- see device.h.
#include "device.h"

#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>

namespace birp_api {

birp_device* device() {
    static std::mutex mu;
    static birp_device* dev = nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!dev) {
        const char* e = std::getenv("BIRP_DEVICE");
        int ordinal = e ? std::atoi(e) : 0;
        int st = birp_device_open(ordinal, &dev);
        if (st != BIRP_OK) {
            dev = nullptr;
            throw std::runtime_error(std::string("birp: cannot open GPU ") + std::to_string(ordinal) + ": " + birp_strerror(st) +
                                     " (" + birp_last_error() + ")");
        }
    }
    return dev;
}

void check(int status, const char* where) {
    if (status == BIRP_OK) return;
    std::string msg = std::string(where) + ": " + birp_strerror(status) + " (" + birp_last_error() + ")";
    if (status == BIRP_ERR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
    if (status == BIRP_ERR_SINGULAR) throw std::runtime_error("Determinant is 0");
    throw std::runtime_error(msg);
}

}  // namespace birp_api
