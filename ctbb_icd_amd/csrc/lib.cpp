// lib.cpp -- library-wide pieces of libctbb_icd.so: last-error slot, device
// selection, allocation helper.
#include "common.hpp"

#include <iostream>
#include <streambuf>

namespace ctbb {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int dev_alloc(void **p, size_t bytes, const char *what)
{
    *p = nullptr;
    if (bytes == 0)
        bytes = 1;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(CTBB_E_NOMEM, std::string("hipMalloc(") + what + ", " +
                                      std::to_string(bytes) + " bytes): " +
                                      hipGetErrorString(e));
    }
    return CTBB_OK;
}

static bool g_verbose = true;

namespace {
struct NullBuf : std::streambuf {
    int overflow(int c) override { return c; }
};
NullBuf g_null_buf;
std::ostream g_null_stream(&g_null_buf);
}  // namespace

std::ostream &out() { return g_verbose ? std::cout : g_null_stream; }

}  // namespace ctbb

extern "C" void ctbb_set_verbose(int on) { ctbb::g_verbose = on != 0; }

extern "C" const char *ctbb_last_error(void)
{
    return ctbb::g_last_error.c_str();
}

extern "C" const char *ctbb_version(void) { return "ctbb_icd_amd 0.1 (gfx950)"; }

extern "C" int ctbb_init(int device)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return ctbb::fail(CTBB_E_HIP, "no HIP device available");
    if (device < 0 || device >= n)
        return ctbb::fail(CTBB_E_ARG, "device index out of range");
    CTBB_HIP(hipSetDevice(device));
    return CTBB_OK;
}

