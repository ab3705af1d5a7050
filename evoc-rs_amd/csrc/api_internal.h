// Host-side internals shared by the C ABI translation units.
#pragma once
#include <functional>
#include <string>

#include "../../include/evoc_hip.h"
#include "kernels.h"

namespace evoc {

struct NoDevice : std::runtime_error {
  explicit NoDevice(const std::string& s) : std::runtime_error(s) {}
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t aux = nullptr;  // copies of finished results back to the host, beside stream
  DevPool pool;
};

Ctx& context(int device);
void release_all();
// runs f under the library-wide call lock, mapping exceptions to status codes
int guarded(const std::function<void()>& f);
const char* last_error();

}  // namespace evoc
