// vectorspace_hip.h -- device-memory ownership for the gfx950 backend.
//
// Drop-in counterpart of qsim's VectorSpaceCUDA (lib/vectorspace_cuda.h:
// 42-167): same Vector type shape, Create/Null/IsNull/Free/Copy/DeviceSync.
// Host-only C++; every device operation goes through the C ABI of qsim_hip.h.

#ifndef QSIM_AMD_VECTORSPACE_HIP_H_
#define QSIM_AMD_VECTORSPACE_HIP_H_

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <type_traits>
#include <utility>

#include "qsim_hip.h"

namespace qsim {

namespace hip_detail {

// Device errors end the process, as lib/util_cuda.h:31-39 (ErrorAssert) does.
inline void Check(int rc, const char* what) {
  if (rc != QSIM_HIP_OK) {
    std::fprintf(stderr, "qsim_hip error in %s: %s\n", what,
                 qsim_hip_last_error());
    std::exit(EXIT_FAILURE);
  }
}

inline void do_not_free(void*) {}

inline void free_device(void* ptr) { Check(qsim_hip_free(ptr), "Free"); }

template <typename FP>
constexpr int Prec() {
  static_assert(std::is_same<FP, float>::value ||
                    std::is_same<FP, double>::value,
                "complex64 or complex128 states only");
  return std::is_same<FP, float>::value ? QSIM_HIP_F32 : QSIM_HIP_F64;
}

}  // namespace hip_detail

template <typename Impl, typename FP>
class VectorSpaceHIP {
 public:
  using fp_type = FP;

 private:
  using Pointer = std::unique_ptr<fp_type, decltype(&hip_detail::free_device)>;

 public:
  class Vector {
   public:
    Vector() = delete;

    Vector(Pointer&& ptr, unsigned num_qubits)
        : ptr_(std::move(ptr)), num_qubits_(num_qubits) {}

    fp_type* get() { return ptr_.get(); }

    const fp_type* get() const { return ptr_.get(); }

    fp_type* release() {
      num_qubits_ = 0;
      return ptr_.release();
    }

    unsigned num_qubits() const { return num_qubits_; }

    static constexpr bool requires_copy_to_host() { return true; }

   private:
    Pointer ptr_;
    unsigned num_qubits_;
  };

  template <typename... Args>
  VectorSpaceHIP(Args&&...) {}

  // Device allocation; Null() when the device is out of memory.
  static Vector Create(unsigned num_qubits) {
    void* p = nullptr;
    int rc = qsim_hip_alloc(hip_detail::Prec<FP>(), num_qubits, &p);
    if (rc == QSIM_HIP_OK) {
      return Vector{Pointer{static_cast<fp_type*>(p), &hip_detail::free_device},
                    num_qubits};
    }
    return Null();
  }

  // Wraps caller-owned device memory of at least Impl::MinSize(num_qubits)
  // elements; it is not freed.
  static Vector Create(fp_type* p, unsigned num_qubits) {
    return Vector{Pointer{p, &hip_detail::do_not_free}, num_qubits};
  }

  static Vector Null() {
    return Vector{Pointer{nullptr, &hip_detail::free_device}, 0};
  }

  static bool IsNull(const Vector& vector) { return vector.get() == nullptr; }

  static void Free(fp_type* ptr) { hip_detail::free_device(ptr); }

  bool Copy(const Vector& src, Vector& dest) const {
    if (src.num_qubits() != dest.num_qubits()) return false;
    hip_detail::Check(
        qsim_hip_copy_d2d(dest.get(), src.get(),
                          sizeof(fp_type) * Impl::MinSize(src.num_qubits())),
        "Copy");
    return true;
  }

  // dest (host) holds at least Impl::MinSize(src.num_qubits()) elements.
  bool Copy(const Vector& src, fp_type* dest) const {
    hip_detail::Check(
        qsim_hip_copy_d2h(dest, src.get(),
                          sizeof(fp_type) * Impl::MinSize(src.num_qubits())),
        "Copy");
    return true;
  }

  // src (host) holds at least Impl::MinSize(dest.num_qubits()) elements.
  bool Copy(const fp_type* src, Vector& dest) const {
    hip_detail::Check(
        qsim_hip_copy_h2d(dest.get(), src,
                          sizeof(fp_type) * Impl::MinSize(dest.num_qubits())),
        "Copy");
    return true;
  }

  // src (host) holds at least min(size, Impl::MinSize(dest.num_qubits())).
  bool Copy(const fp_type* src, uint64_t size, Vector& dest) const {
    size = std::min(size, Impl::MinSize(dest.num_qubits()));
    hip_detail::Check(qsim_hip_copy_h2d(dest.get(), src, sizeof(fp_type) * size),
                      "Copy");
    return true;
  }

  static void DeviceSync() {
    hip_detail::Check(qsim_hip_synchronize(), "DeviceSync");
  }
};

}  // namespace qsim

#endif  // QSIM_AMD_VECTORSPACE_HIP_H_
