// Entry points of the C ABI return a status and never let a C++ exception out:
// a failed allocation (std::bad_alloc, std::length_error) is DFN_ERR_NOMEM, any
// other exception DFN_ERR_ARG.  An entry point's body sits between
// DFN_API_BEGIN and DFN_API_END.
#pragma once
#include <new>
#include <stdexcept>

#include "../../include/dfn.h"

#define DFN_API_BEGIN try {
#define DFN_API_END                          \
  }                                          \
  catch (const std::bad_alloc &) {           \
    return DFN_ERR_NOMEM;                    \
  }                                          \
  catch (const std::length_error &) {        \
    return DFN_ERR_NOMEM;                    \
  }                                          \
  catch (...) {                              \
    return DFN_ERR_ARG;                      \
  }
