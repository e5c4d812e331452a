// Exception -> status-code translation at the C-ABI boundary.
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "../../../include/aphi_gpu.h"

namespace aphi_capi {

// A HIP runtime failure (bad launch, fault, out of device memory).
struct DeviceError : std::runtime_error {
    int code;
    DeviceError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};
// A pivot at or below the floor: the call completed and filled its stats.
struct PivotFailure : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void set_error(const std::string& message);
void clear_error();

template <typename F>
int guarded(F&& f) {
    clear_error();
    try {
        f();
        return APHI_OK;
    } catch (const std::invalid_argument& e) {
        set_error(e.what());
        return APHI_INVALID_ARGUMENT;
    } catch (const std::logic_error& e) {
        set_error(e.what());
        return APHI_LOGIC_ERROR;
    } catch (const PivotFailure& e) {
        set_error(e.what());
        return APHI_PIVOT_FAILED;
    } catch (const DeviceError& e) {
        set_error(e.what());
        return e.code == 2 ? APHI_OUT_OF_MEMORY : APHI_DEVICE_ERROR;
    } catch (const std::bad_alloc&) {
        set_error("out of host memory");
        return APHI_OUT_OF_MEMORY;
    } catch (const std::runtime_error& e) {
        set_error(e.what());
        return APHI_RUNTIME_ERROR;
    } catch (...) {
        set_error("unknown exception");
        return APHI_RUNTIME_ERROR;
    }
}

}  // namespace aphi_capi
