// abi.hpp -- how every extern "C" entry point of libsps turns C++ exceptions
// into an sps_status and keeps the message for sps_last_error().
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "common.hpp"

namespace sps {

// message of the calling thread's last failure (defined in api.cpp)
extern thread_local std::string g_last_error;

template <class F>
sps_status guarded(F&& f) {
    try {
        f();
        return SPS_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.status;
    } catch (const std::invalid_argument& e) {
        g_last_error = e.what();
        return SPS_ERR_INVALID_ARGUMENT;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return SPS_ERR_NO_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return SPS_ERR_RUNTIME;
    } catch (...) {
        g_last_error = "unknown error";
        return SPS_ERR_RUNTIME;
    }
}

}  // namespace sps
