// Exception barrier of the C ABI: no C++ exception leaves an extern "C" entry point.
//
// Every entry point's body sits between FEMM_API_BEGIN and FEMM_API_END.  A failed host
// allocation (std::bad_alloc, or std::length_error from a container sized beyond its
// limit) becomes FEMM_GPU_ALLOCATION_FAILED, the reference's allocation-failure status;
// any other exception becomes FEMM_INTERNAL_ERROR.  Entry points returning void swallow
// the exception (they only release resources).
#pragma once

#include <new>
#include <stdexcept>

#include "femm_model.h"

namespace femm
{

inline int StatusOfCurrentException() noexcept
{
  try {
    throw;
  } catch (const std::bad_alloc &) {
    return FEMM_GPU_ALLOCATION_FAILED;
  } catch (const std::length_error &) {
    return FEMM_GPU_ALLOCATION_FAILED;
  } catch (...) {
    return FEMM_INTERNAL_ERROR;
  }
}

}  // namespace femm

#define FEMM_API_BEGIN try {
#define FEMM_API_END                              \
  }                                               \
  catch (...)                                     \
  {                                               \
    return ::femm::StatusOfCurrentException();    \
  }
#define FEMM_API_END_VOID \
  }                       \
  catch (...) {}
