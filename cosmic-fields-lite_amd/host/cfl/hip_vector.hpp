// hip_vector.hpp -- RAII device array over the C ABI (libcfhip.so).
//
// Plays the role thrust::device_vector<double> has in the reference's CUDA path
// (src/cuda_wrapper.cuh:39-45): the State type of HIP equations, the odeint
// state, and the argument of the spectrum functions. Host code only; every
// device operation goes through include/cfhip.h.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>

#include "cfhip.h"

namespace cfl {

// Throws std::runtime_error with the library's message on failure.
inline void check(cf_status s, const char *what) {
  if (s != CF_OK) throw std::runtime_error(std::string(what) + ": " + cf_last_error());
}

// The stream all host-layer launches use. The reference runs everything on the
// legacy default stream; so does this layer unless the caller sets another.
inline cf_stream &default_stream() {
  static cf_stream s = nullptr;
  return s;
}

template <class T>
class hip_vector {
 public:
  typedef T value_type;
  typedef std::size_t size_type;

  hip_vector() = default;
  explicit hip_vector(size_type n) { resize(n); }
  hip_vector(const hip_vector &o) {
    resize(o.n_);
    copy_device(o);
  }
  hip_vector(hip_vector &&o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  hip_vector &operator=(const hip_vector &o) {
    if (this != &o) {
      if (n_ != o.n_) {
        release();
        resize(o.n_);
      }
      copy_device(o);
    }
    return *this;
  }
  hip_vector &operator=(hip_vector &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  ~hip_vector() { release(); }

  size_type size() const { return n_; }
  bool empty() const { return n_ == 0; }
  T *data() { return p_; }
  const T *data() const { return p_; }

  // Contents up to min(old, new) size are kept (thrust::device_vector::resize).
  void resize(size_type n) {
    if (n == n_) return;
    T *q = nullptr;
    if (n) check(cf_alloc(n * sizeof(T), reinterpret_cast<void **>(&q)), "hip_vector::resize");
    if (p_ && q) {
      const size_type k = n < n_ ? n : n_;
      cf_status s = cf_d2d(q, p_, k * sizeof(T), default_stream());
      if (s != CF_OK) {
        cf_free(q);
        check(s, "hip_vector::resize copy");
      }
      check(cf_sync(default_stream()), "hip_vector::resize sync");
    }
    release();
    p_ = q;
    n_ = n;
  }
  void clear() { release(); }
  void swap(hip_vector &o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
  }
  void fill_zero() { check(cf_fill_zero(p_, n_ * sizeof(T), default_stream()), "hip_vector::fill_zero"); }

  // Host <-> device copies for any contiguous host container with data()/size()
  // (Eigen::VectorXd, std::vector<T>, ...).
  template <class Host>
  void copy_from_host(const Host &h) {
    if ((size_type)h.size() != n_) resize((size_type)h.size());
    check(cf_h2d(p_, h.data(), n_ * sizeof(T), default_stream()), "hip_vector::copy_from_host");
  }
  template <class Host>
  void copy_to_host(Host &h) const {
    if ((size_type)h.size() < n_) throw std::invalid_argument("copy_to_host: host container too small");
    check(cf_d2h(h.data(), p_, n_ * sizeof(T), default_stream()), "hip_vector::copy_to_host");
  }

 private:
  void copy_device(const hip_vector &o) {
    check(cf_d2d(p_, o.p_, n_ * sizeof(T), default_stream()), "hip_vector copy");
  }
  void release() noexcept {
    if (p_) cf_free(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T *p_ = nullptr;
  size_type n_ = 0;
};

template <class T>
inline void swap(hip_vector<T> &a, hip_vector<T> &b) noexcept {
  a.swap(b);
}

// copy_vector(out, in) -- src/cuda_wrapper.cu:45-54 (device to host; out is
// allocated by the caller and must be at least as long as in).
template <class Host>
inline void copy_vector(Host &out, const hip_vector<double> &in) {
  in.copy_to_host(out);
}

// ALGORITHM_NAMESPACE::copy(host_begin, host_end, state.begin()) of the
// initializers (src/initializer.hpp:24-30): host range into a device vector at
// an element offset.
template <class T>
inline void copy_host_range(const T *first, const T *last, hip_vector<T> &dst, std::size_t offset = 0) {
  const std::size_t n = (std::size_t)(last - first);
  if (offset + n > dst.size()) throw std::invalid_argument("copy_host_range: destination too small");
  check(cf_h2d(dst.data() + offset, first, n * sizeof(T), default_stream()), "copy_host_range");
}

}  // namespace cfl
