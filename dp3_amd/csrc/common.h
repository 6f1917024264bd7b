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

// ---- address-space views of device pointers ----
// Pointers held in argument structs in device memory are generic to the
// compiler, which then emits flat loads. Flat loads also count as LDS
// operations, so every wait for an LDS read drains the memory loads in
// flight as well. Global views give global loads (a scalar base plus a
// 32-bit offset register where the base is uniform); constant views give
// scalar loads for uniform, read-only values.
template <typename T>
using GlobalT = __attribute__((address_space(1))) T;
template <typename T>
using ConstT = __attribute__((address_space(4))) T;
template <typename T>
__device__ __forceinline__ const GlobalT<T>* AsGlobal(const T* p) {
  return (const GlobalT<T>*)p;
}
template <typename T>
__device__ __forceinline__ GlobalT<T>* AsGlobal(T* p) {
  return (GlobalT<T>*)p;
}
template <typename T>
__device__ __forceinline__ const ConstT<T>* AsConst(const T* p) {
  return (const ConstT<T>*)p;
}

// Copies of trivially copyable structs out of those views, word by word
// (implicit copy constructors only bind generic references).
template <typename T>
__device__ __forceinline__ T LoadConst(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
  T out;
  const ConstT<uint32_t>* src = (const ConstT<uint32_t>*)p;
  uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
  for (int i = 0; i < int(sizeof(T) / 4); ++i) dst[i] = src[i];
  return out;
}
__device__ __forceinline__ double2 LoadGlobal(const double2* p) {
  const GlobalT<double>* q = (const GlobalT<double>*)p;
  return make_double2(q[0], q[1]);
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
