// Status strings and version of the gsx C ABI.
#include "gsx_internal.h"

extern "C" {

const char* gsx_status_string(int status) {
    switch (status) {
        case GSX_OK: return "ok";
        case GSX_E_INVALID: return "invalid argument (dimension, size or null pointer)";
        case GSX_E_LAUNCH: return "kernel launch or HIP runtime failure";
        case GSX_E_NO_DEVICE: return "no usable HIP device";
        case GSX_E_WORKSPACE: return "workspace missing or too small";
        case GSX_E_UNSUPPORTED: return "request not supported by this build";
        case GSX_E_SCRATCH: return "request exceeds the library's static scratch array";
        default: return "unknown status";
    }
}

int gsx_version(void) { return 100; }

}  // extern "C"
