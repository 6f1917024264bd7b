// Shared helpers of the HIP library: error reporting for the C-ABI and small
// complex-arithmetic device functions.
#ifndef DP3_AMD_COMMON_H
#define DP3_AMD_COMMON_H

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdio>
#include <stdexcept>
#include <string>

namespace ddh {

// Thread-local last error message, returned by ddh_last_error().
void SetLastError(const std::string& message);
const char* LastError();

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

#define DDH_HIP_CHECK(expr)                                                    \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess)                                                      \
      throw ::ddh::Error(-3, std::string(#expr " failed: ") +                  \
                                 hipGetErrorString(_e) + " at " __FILE__ ":" + \
                                 std::to_string(__LINE__));                    \
  } while (0)

#define DDH_REQUIRE(cond, code, msg)              \
  do {                                            \
    if (!(cond)) throw ::ddh::Error((code), (msg)); \
  } while (0)

// Runs f, translating exceptions into a status code and the last error.
template <typename F>
int Guard(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    SetLastError(e.what());
    return e.code();
  } catch (const std::exception& e) {
    SetLastError(e.what());
    return -1;
  }
}

// ---- device complex helpers (float2 = complex64, double2 = complex128) ----
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ double cnorm(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ bool cfinite(double2 a) {
  return isfinite(a.x) && isfinite(a.y);
}

}  // namespace ddh

#endif
