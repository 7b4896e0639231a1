// cuda_abstract_vector.cuh -- drop-in SoA device 3-vector for gfx950.
//
// Same name, members and operations as the reference's
// magnetic_solver/cuda_abstract_vector.cuh:166-243 (np, dev_x, dev_y, dev_z,
// cudaStatus; value-returning + - unary- * and T*vec; += -= *=; import/export;
// reserve, clear, norm2, norm, dot, add_to_element).  Differences:
//  * the three components live in one allocation (dev_y = dev_x + np,
//    dev_z = dev_x + 2 np); free it only through this class;
//  * every operation is one fused launch through the C ABI (mf_capi.h) on the
//    default context, which runs on the HIP null stream, so it is ordered
//    with synchronous hipMemcpy calls exactly like the reference's kernels;
//  * dot/norm2 reduce in FP64 with a fixed order (the reference reduces 64
//    FP32 partials on the host);
//  * cudaStatus is a hipError_t; failures of the C ABI are thrown as
//    std::runtime_error (the reference stored and ignored them);
//  * T is float, double or int.  float and double have every operation; int
//    holds data (copies, import/export, reserve, add_to_element) and converts
//    to the other types with the templated conversion operator
//    (cuda_abstract_vector.cuh:202, 367-379), provided for all three.
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "mf_capi.h"

namespace Libo {

namespace detail {
inline mf_ctx *default_ctx() {
  mf_ctx *ctx = nullptr;
  const int rc = mf_ctx_get_default(&ctx);
  if (rc != MF_OK) throw std::runtime_error(std::string("magforce: ") + mf_last_error());
  return ctx;
}
inline void check(int rc, mf_ctx *ctx) {
  if (rc != MF_OK)
    throw std::runtime_error(std::string("magforce: ") +
                             (ctx ? mf_ctx_last_error(ctx) : mf_last_error()));
}
inline void check_hip(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string("magforce: ") + what + ": " + hipGetErrorString(e));
}

// Element type code of mf_vec_convert.
template <typename T> struct dtype_of;
template <> struct dtype_of<float> { static constexpr int value = MF_DTYPE_F32; };
template <> struct dtype_of<double> { static constexpr int value = MF_DTYPE_F64; };
template <> struct dtype_of<int> { static constexpr int value = MF_DTYPE_I32; };

// Arithmetic of the floating element types through the C ABI.
template <typename T> struct vec_ops;
template <> struct vec_ops<float> {
  static int axpby(mf_ctx *c, int64_t n, float a, const float *ax, const float *ay,
                   const float *az, float b, const float *bx, const float *by, const float *bz,
                   float *ox, float *oy, float *oz) {
    return mf_vec_axpby(c, n, a, ax, ay, az, b, bx, by, bz, ox, oy, oz);
  }
  static int dot(mf_ctx *c, int64_t n, const float *ax, const float *ay, const float *az,
                 const float *bx, const float *by, const float *bz, double *r) {
    return mf_vec_dot(c, n, ax, ay, az, bx, by, bz, r);
  }
};
template <> struct vec_ops<double> {
  static int axpby(mf_ctx *c, int64_t n, double a, const double *ax, const double *ay,
                   const double *az, double b, const double *bx, const double *by,
                   const double *bz, double *ox, double *oy, double *oz) {
    return mf_vec_axpby_f64(c, n, a, ax, ay, az, b, bx, by, bz, ox, oy, oz);
  }
  static int dot(mf_ctx *c, int64_t n, const double *ax, const double *ay, const double *az,
                 const double *bx, const double *by, const double *bz, double *r) {
    return mf_vec_dot_f64(c, n, ax, ay, az, bx, by, bz, r);
  }
};
}  // namespace detail

template <typename T>
class cuda_abstract_vector {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value ||
                    std::is_same<T, int>::value,
                "cuda_abstract_vector is provided for float, double and int");
  template <typename U>
  friend class cuda_abstract_vector;

 public:
  typedef T value_type;

  cuda_abstract_vector() : cudaStatus(hipSuccess), np(0), dev_x(nullptr), dev_y(nullptr), dev_z(nullptr) {}
  explicit cuda_abstract_vector(size_t size) : cuda_abstract_vector() { allocate(size); }
  cuda_abstract_vector(const cuda_abstract_vector &old) : cuda_abstract_vector() {
    allocate(old.np);
    copy_from(old);
  }
  cuda_abstract_vector(cuda_abstract_vector &&old) noexcept
      : cudaStatus(old.cudaStatus), np(old.np), dev_x(old.dev_x), dev_y(old.dev_y), dev_z(old.dev_z) {
    old.np = 0;
    old.dev_x = old.dev_y = old.dev_z = nullptr;
  }
  ~cuda_abstract_vector() { clear(); }

  cuda_abstract_vector &operator=(const cuda_abstract_vector &old) {
    if (this == &old) return *this;
    if (np != old.np) {
      clear();
      allocate(old.np);
    }
    copy_from(old);
    return *this;
  }
  cuda_abstract_vector &operator=(cuda_abstract_vector &&old) noexcept {
    if (this != &old) {
      clear();
      cudaStatus = old.cudaStatus;
      np = old.np;
      dev_x = old.dev_x;
      dev_y = old.dev_y;
      dev_z = old.dev_z;
      old.np = 0;
      old.dev_x = old.dev_y = old.dev_z = nullptr;
    }
    return *this;
  }

  cuda_abstract_vector operator+(const cuda_abstract_vector &o) const { return combine(T(1), o, T(1)); }
  cuda_abstract_vector operator-(const cuda_abstract_vector &o) const { return combine(T(1), o, T(-1)); }
  cuda_abstract_vector operator-() const { return scaled(T(-1)); }
  cuda_abstract_vector &operator+=(const cuda_abstract_vector &o) { return accumulate(o, T(1)); }
  cuda_abstract_vector &operator-=(const cuda_abstract_vector &o) { return accumulate(o, T(-1)); }
  cuda_abstract_vector &operator*=(const T C) {
    mf_ctx *ctx = detail::default_ctx();
    detail::check(detail::vec_ops<T>::axpby(ctx, static_cast<int64_t>(np), C, dev_x, dev_y, dev_z,
                                            T(0), nullptr, nullptr, nullptr, dev_x, dev_y, dev_z),
                  ctx);
    return *this;
  }
  const cuda_abstract_vector operator*(const T C) const { return scaled(C); }
  // Element-wise conversion to another element type (cuda_abstract_vector.cuh:367-379):
  // one mf_vec_convert launch per component; float or double to int truncates.
  template <typename other_T>
  operator cuda_abstract_vector<other_T>() const {
    cuda_abstract_vector<other_T> r(np);
    if (np == 0) return r;
    mf_ctx *ctx = detail::default_ctx();
    const T *src[3] = {dev_x, dev_y, dev_z};
    other_T *dst[3] = {r.dev_x, r.dev_y, r.dev_z};
    for (int c = 0; c < 3; ++c)
      detail::check(mf_vec_convert(ctx, static_cast<int64_t>(np), detail::dtype_of<T>::value, src[c],
                                   detail::dtype_of<other_T>::value, dst[c]),
                    ctx);
    return r;
  }

  void import_data(const T *hx, const T *hy, const T *hz) {
    if (np == 0) return;
    detail::check_hip(hipMemcpy(dev_x, hx, np * sizeof(T), hipMemcpyHostToDevice), "import_data");
    detail::check_hip(hipMemcpy(dev_y, hy, np * sizeof(T), hipMemcpyHostToDevice), "import_data");
    detail::check_hip(hipMemcpy(dev_z, hz, np * sizeof(T), hipMemcpyHostToDevice), "import_data");
  }
  void import_data(const std::vector<T> &hx, const std::vector<T> &hy, const std::vector<T> &hz) {
    if (hy.size() != hx.size() || hz.size() != hx.size())
      throw std::invalid_argument("import_data: component sizes differ");
    reserve(hx.size());
    import_data(hx.data(), hy.data(), hz.data());
  }
  void import_data(const std::vector<float3> &xyz) {
    reserve(xyz.size());
    std::vector<T> hx(np), hy(np), hz(np);
    for (size_t i = 0; i < np; ++i) {
      hx[i] = xyz[i].x;
      hy[i] = xyz[i].y;
      hz[i] = xyz[i].z;
    }
    import_data(hx.data(), hy.data(), hz.data());
  }
  void export_data(T *hx, T *hy, T *hz) const {
    if (np == 0) return;
    detail::check_hip(hipMemcpy(hx, dev_x, np * sizeof(T), hipMemcpyDeviceToHost), "export_data");
    detail::check_hip(hipMemcpy(hy, dev_y, np * sizeof(T), hipMemcpyDeviceToHost), "export_data");
    detail::check_hip(hipMemcpy(hz, dev_z, np * sizeof(T), hipMemcpyDeviceToHost), "export_data");
  }
  void export_data(std::vector<T> &hx, std::vector<T> &hy, std::vector<T> &hz) const {
    hx.resize(np);
    hy.resize(np);
    hz.resize(np);
    export_data(hx.data(), hy.data(), hz.data());
  }
  void add_to_element(size_t id, T x, T y, T z) {
    if (id >= np) throw std::out_of_range("add_to_element: index out of range");
    T v[3];
    for (int c = 0; c < 3; ++c)
      detail::check_hip(hipMemcpy(&v[c], dev_x + c * np + id, sizeof(T), hipMemcpyDeviceToHost),
                        "add_to_element");
    v[0] += x;
    v[1] += y;
    v[2] += z;
    for (int c = 0; c < 3; ++c)
      detail::check_hip(hipMemcpy(dev_x + c * np + id, &v[c], sizeof(T), hipMemcpyHostToDevice),
                        "add_to_element");
  }
  void clear() {
    if (dev_x) (void)hipFree(dev_x);
    dev_x = dev_y = dev_z = nullptr;
    np = 0;
  }
  // Reallocates when the size changes (contents are not kept), as the reference.
  void reserve(size_t n) {
    if (n == np && (n == 0 || dev_x)) return;
    clear();
    allocate(n);
  }
  T norm2() const { return dot_impl(*this); }
  T norm() const { return std::sqrt(norm2()); }
  T dot(const cuda_abstract_vector &o) const { return dot_impl(o); }
  // FP64 value of the reductions (not in the reference).
  double dot_fp64(const cuda_abstract_vector &o) const {
    double r = 0.0;
    mf_ctx *ctx = detail::default_ctx();
    detail::check(detail::vec_ops<T>::dot(ctx, static_cast<int64_t>(np), dev_x, dev_y, dev_z,
                                          o.dev_x, o.dev_y, o.dev_z, &r),
                  ctx);
    return r;
  }

  hipError_t cudaStatus;
  size_t np;
  T *dev_x;
  T *dev_y;
  T *dev_z;

 private:
  void allocate(size_t n) {
    np = n;
    if (n == 0) return;
    void *p = nullptr;
    cudaStatus = hipMalloc(&p, 3 * n * sizeof(T));
    detail::check_hip(cudaStatus, "hipMalloc");
    dev_x = static_cast<T *>(p);
    dev_y = dev_x + n;
    dev_z = dev_x + 2 * n;
  }
  void copy_from(const cuda_abstract_vector &old) {
    if (np == 0) return;
    if (old.dev_y == old.dev_x + old.np && old.dev_z == old.dev_x + 2 * old.np) {
      cudaStatus = hipMemcpy(dev_x, old.dev_x, 3 * np * sizeof(T), hipMemcpyDeviceToDevice);
    } else {
      cudaStatus = hipMemcpy(dev_x, old.dev_x, np * sizeof(T), hipMemcpyDeviceToDevice);
      if (cudaStatus == hipSuccess)
        cudaStatus = hipMemcpy(dev_y, old.dev_y, np * sizeof(T), hipMemcpyDeviceToDevice);
      if (cudaStatus == hipSuccess)
        cudaStatus = hipMemcpy(dev_z, old.dev_z, np * sizeof(T), hipMemcpyDeviceToDevice);
    }
    detail::check_hip(cudaStatus, "hipMemcpy");
  }
  cuda_abstract_vector combine(T a, const cuda_abstract_vector &o, T b) const {
    if (o.np != np) throw std::invalid_argument("cuda_abstract_vector: size mismatch");
    cuda_abstract_vector r(np);
    mf_ctx *ctx = detail::default_ctx();
    detail::check(detail::vec_ops<T>::axpby(ctx, static_cast<int64_t>(np), a, dev_x, dev_y, dev_z,
                                            b, o.dev_x, o.dev_y, o.dev_z, r.dev_x, r.dev_y,
                                            r.dev_z),
                  ctx);
    return r;
  }
  cuda_abstract_vector scaled(T a) const {
    cuda_abstract_vector r(np);
    mf_ctx *ctx = detail::default_ctx();
    detail::check(detail::vec_ops<T>::axpby(ctx, static_cast<int64_t>(np), a, dev_x, dev_y, dev_z,
                                            T(0), nullptr, nullptr, nullptr, r.dev_x, r.dev_y,
                                            r.dev_z),
                  ctx);
    return r;
  }
  cuda_abstract_vector &accumulate(const cuda_abstract_vector &o, T b) {
    if (o.np != np) throw std::invalid_argument("cuda_abstract_vector: size mismatch");
    mf_ctx *ctx = detail::default_ctx();
    detail::check(detail::vec_ops<T>::axpby(ctx, static_cast<int64_t>(np), T(1), dev_x, dev_y,
                                            dev_z, b, o.dev_x, o.dev_y, o.dev_z, dev_x, dev_y,
                                            dev_z),
                  ctx);
    return *this;
  }
  T dot_impl(const cuda_abstract_vector &o) const {
    if (o.np != np) throw std::invalid_argument("cuda_abstract_vector: size mismatch");
    return static_cast<T>(dot_fp64(o));
  }
};

template <typename T>
const cuda_abstract_vector<T> operator*(const T &C, const cuda_abstract_vector<T> &in) {
  return in * C;
}

}  // namespace Libo
