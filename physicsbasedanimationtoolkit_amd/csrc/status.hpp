// Error reporting shared by the C ABI entry points of libpbat_vbd.so.
#pragma once

#include "pbat_vbd.h"

#include <stdexcept>
#include <string>

namespace pbat_vbd {

// Thread-local message of the last failing C ABI call.
void SetLastError(std::string const& what);
void ClearLastError();

// Exception carrying a C ABI status code.
struct Error : std::runtime_error
{
    int code;
    Error(int c, std::string const& what) : std::runtime_error(what), code(c) {}
};

[[noreturn]] inline void Fail(int code, std::string const& what) { throw Error(code, what); }

// Run f() and translate exceptions into status codes.
template <class F>
int Guard(F&& f) noexcept
{
    try
    {
        ClearLastError();
        f();
        return PBAT_VBD_OK;
    }
    catch (Error const& e)
    {
        SetLastError(e.what());
        return e.code;
    }
    catch (std::bad_alloc const&)
    {
        SetLastError("out of host memory");
        return PBAT_VBD_ERROR_OUT_OF_MEMORY;
    }
    catch (std::invalid_argument const& e)
    {
        SetLastError(e.what());
        return PBAT_VBD_ERROR_INVALID_ARGUMENT;
    }
    catch (std::exception const& e)
    {
        SetLastError(e.what());
        return PBAT_VBD_ERROR_HIP;
    }
    catch (...)
    {
        SetLastError("unknown error");
        return PBAT_VBD_ERROR_HIP;
    }
}

} // namespace pbat_vbd
