// A small value-or-error type with the parts of tl::expected's interface the
// reference's operators use (has_value, value, error, operator*, make_unexpected).
#pragma once

#include <string>
#include <utility>
#include <variant>

namespace fpdb::amd {

struct Unexpected {
  std::string error;
};

inline Unexpected make_unexpected(std::string e) { return Unexpected{std::move(e)}; }

template <typename T>
class Expected {
 public:
  Expected(T v) : v_(std::in_place_index<0>, std::move(v)) {}
  Expected(Unexpected u) : v_(std::in_place_index<1>, std::move(u.error)) {}
  bool has_value() const { return v_.index() == 0; }
  explicit operator bool() const { return has_value(); }
  T& value() { return std::get<0>(v_); }
  const T& value() const { return std::get<0>(v_); }
  T& operator*() { return value(); }
  const T& operator*() const { return value(); }
  T* operator->() { return &value(); }
  const std::string& error() const { return std::get<1>(v_); }

 private:
  std::variant<T, std::string> v_;
};

template <>
class Expected<void> {
 public:
  Expected() = default;
  Expected(Unexpected u) : ok_(false), err_(std::move(u.error)) {}
  bool has_value() const { return ok_; }
  explicit operator bool() const { return ok_; }
  const std::string& error() const { return err_; }

 private:
  bool ok_ = true;
  std::string err_;
};

}  // namespace fpdb::amd
