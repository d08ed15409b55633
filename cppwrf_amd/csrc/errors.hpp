// Error plumbing of the C ABI: every entry point runs its body through guard(), which turns a C++
// exception into a status code and keeps the message for mi_last_error() (per host thread).
#pragma once

#include <new>
#include <stdexcept>
#include <string>

#include "../../include/cppwrf_amd.h"

namespace mi_detail {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& m);

template <class F>
int guard(F&& f) noexcept {
  try {
    f();
    return MI_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("out of host memory");
    return MI_ENOMEM;
  } catch (const std::invalid_argument& e) {
    set_last_error(e.what());
    return MI_EINVAL;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return MI_EINVAL;
  } catch (...) {
    set_last_error("unknown error");
    return MI_EINVAL;
  }
}

}  // namespace mi_detail
