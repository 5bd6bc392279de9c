// Exception fence of the C ABI: no C++ exception crosses an extern "C" entry point. A function
// body between MRT_GUARD_BEGIN and MRT_GUARD_END turns std::bad_alloc into MRT_ERR_LIMIT and any
// other std::exception into MRT_ERR_INTERNAL, with the reason in mrt_last_error().
#pragma once
#include <new>
#include <stdexcept>
#include <string>

namespace mrt {
int fail(int code, const std::string& msg);
}

#define MRT_GUARD_BEGIN try {
#define MRT_GUARD_END                                        \
  }                                                          \
  catch (const std::bad_alloc&) {                            \
    return ::mrt::fail(MRT_ERR_LIMIT, "out of host memory"); \
  }                                                          \
  catch (const std::exception& ex) {                         \
    return ::mrt::fail(MRT_ERR_INTERNAL, ex.what());         \
  }
