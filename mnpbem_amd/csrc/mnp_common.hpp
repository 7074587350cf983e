// Shared definitions for the MI355X (gfx950) retarded-BEM library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "../../include/mnpbem_hip.h"

namespace mnp {

// ----------------------------------------------------------------- errors
void set_error(const std::string& msg);

// Runs independent host tasks that enqueue GPU work: tasks[0] on the calling
// thread and stream, each other task on a worker thread with a stream of its
// own that starts after the work the caller's stream holds so far; the
// caller's stream then waits for all of them.  At most 4 tasks; called from a
// worker (nested) the tasks run one after another.  The first exception is
// rethrown after every task has ended and its stream has drained.
void run_concurrent(std::vector<std::function<void()>>& tasks);

struct Error {
  int code;
  std::string msg;
};

#define MNP_HIP(call)                                                             \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      throw ::mnp::Error{MNP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                          " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"}; \
    }                                                                             \
  } while (0)

#define MNP_CHECK(cond, code, msg)                 \
  do {                                             \
    if (!(cond)) throw ::mnp::Error{(code), (msg)}; \
  } while (0)

// Runs `body`, converts exceptions into a status code + last-error string.
template <class F>
int guarded(F&& body) {
  try {
    body();
    return MNP_OK;
  } catch (const Error& e) {
    set_error(e.msg);
    return e.code;
  } catch (const std::exception& e) {
    set_error(e.what());
    return MNP_ERR_INTERNAL;
  }
}

// ------------------------------------------------------------ complex types
template <class T> struct cplx_t;
template <> struct cplx_t<double> { using type = double2; };
template <> struct cplx_t<float> { using type = float2; };
template <class T> using cplx = typename cplx_t<T>::type;

template <class T> __host__ __device__ inline cplx<T> mk(T re, T im) {
  cplx<T> z;
  z.x = re;
  z.y = im;
  return z;
}
template <class T> __host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <class T> __host__ __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
template <class T> __host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <class T> __host__ __device__ inline cplx<T> cscale(cplx<T> a, T s) { return mk<T>(a.x * s, a.y * s); }
// a / b (Smith's algorithm, no spurious overflow)
template <class T> __host__ __device__ inline cplx<T> cdiv(cplx<T> a, cplx<T> b) {
  T ar = b.x >= 0 ? b.x : -b.x, ai = b.y >= 0 ? b.y : -b.y;
  if (ar >= ai) {
    T r = b.y / b.x, d = b.x + b.y * r;
    return mk<T>((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    T r = b.x / b.y, d = b.x * r + b.y;
    return mk<T>((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
// |Re| + |Im|, the pivot measure of LAPACK's izamax
template <class T> __host__ __device__ inline T cabs1(cplx<T> a) {
  return (a.x >= 0 ? a.x : -a.x) + (a.y >= 0 ? a.y : -a.y);
}

// exp(i k d) for complex k and real d >= 0
__device__ inline double2 cexp_ikd(double2 k, double d) {
  double s, c;
  sincos(k.x * d, &s, &c);
  double m = exp(-k.y * d);
  return mk<double>(m * c, m * s);
}

inline int64_t cdiv_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Brackets one kernel launch with HIP events when the timers are enabled
// (mnp_timers_enable); `work` is its algorithmic flops or bytes and `bytes`
// its algorithmic global-memory bytes (operands read once, results written
// once).  A negative kind records nothing.
struct TimerScope {
  TimerScope(int kind, double work, double bytes = 0.0);
  ~TimerScope();
  int kind;
  double work, bytes;
  void* ev0 = nullptr;
};

// Converts a user-facing complex scalar into the device type.
// "1.234e-05": residuals and tolerances in error messages
inline std::string sci(double v) {
  char b[32];
  std::snprintf(b, sizeof b, "%.3e", v);
  return b;
}
inline double2 to_d2(mnp_complex z) { return mk<double>(z.re, z.im); }

}  // namespace mnp
