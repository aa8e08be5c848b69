// Test double of libpressio_ext/cpp/options.h: a map from option names to
// values of a few types, read back with arithmetic conversion.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <type_traits>
#include <variant>

enum pressio_thread_safety
{
  pressio_thread_safety_single = 0,
  pressio_thread_safety_serialized = 1,
  pressio_thread_safety_multiple = 2
};

using pressio_option = std::variant<std::monostate, std::int32_t, std::uint32_t, std::int64_t, std::uint64_t, double,
                                    std::string>;

struct pressio_options
{
  std::map<std::string, pressio_option> values;

  bool contains(std::string const& key) const { return values.count(key) != 0; }
};

// store a value: integers keep their width and signedness, enums go in as
// int32, text as std::string
template<class T>
void
pressio_options_set(pressio_options& o, std::string const& key, T const& v)
{
  if constexpr (std::is_enum<T>::value) {
    o.values[key] = static_cast<std::int32_t>(v);
  } else if constexpr (std::is_same<T, std::int32_t>::value || std::is_same<T, std::uint32_t>::value ||
                       std::is_same<T, std::int64_t>::value || std::is_same<T, std::uint64_t>::value ||
                       std::is_same<T, double>::value) {
    o.values[key] = v;
  } else if constexpr (std::is_integral<T>::value) {
    o.values[key] = static_cast<std::int64_t>(v);
  } else {
    o.values[key] = std::string(v);
  }
}

// 0: found and converted; 1: absent (value untouched); 2: wrong type
template<class T>
int
pressio_options_get(pressio_options const& o, std::string const& key, T* v)
{
  auto it = o.values.find(key);
  if (it == o.values.end() || std::holds_alternative<std::monostate>(it->second)) {
    return 1;
  }
  return std::visit(
    [v](auto const& x) -> int {
      using X = std::decay_t<decltype(x)>;
      if constexpr (std::is_same<X, std::monostate>::value) {
        return 1;
      } else if constexpr (std::is_same<X, std::string>::value) {
        if constexpr (std::is_same<T, std::string>::value) {
          *v = x;
          return 0;
        } else {
          return 2;
        }
      } else if constexpr (std::is_arithmetic<T>::value) {
        *v = static_cast<T>(x);
        return 0;
      } else {
        return 2;
      }
    },
    it->second);
}
