// check.h -- a minimal stand-in for the parts of Catch2 v2 the reference's
// tests use (REQUIRE, REQUIRE_FALSE, Approx with its default tolerance:
// relative epsilon 100*FLT_EPSILON of the expected value, margin 0).
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <memory>
#include <vector>

namespace check {

struct Approx {
  explicit Approx(double v) : value(v) {}
  double value;
  double epsilon = FLT_EPSILON * 100.0;
  double margin = 0.0;
  bool matches(double other) const {
    auto within = [](double a, double b, double m) { return a + m >= b && b + m >= a; };
    return within(value, other, margin) ||
           within(value, other, epsilon * std::fabs(std::isinf(value) ? 0.0 : value));
  }
};
inline bool operator==(double lhs, const Approx& rhs) { return rhs.matches(lhs); }

inline int& failures() {
  static int n = 0;
  return n;
}

struct Case {
  const char* name;
  std::function<void()> fn;
};
inline std::vector<Case>& cases() {
  static std::vector<Case> c;
  return c;
}
struct Register {
  Register(const char* name, std::function<void()> fn) { cases().push_back({name, std::move(fn)}); }
};

inline int run_all() {
  int failed_cases = 0;
  for (auto& c : cases()) {
    const int before = failures();
    std::printf("[case] %s\n", c.name);
    std::fflush(stdout);
    try {
      c.fn();
    } catch (const std::exception& e) {
      std::printf("  exception: %s\n", e.what());
      ++failures();
    }
    if (failures() != before) ++failed_cases;
    std::printf("[%s] %s\n", failures() == before ? "pass" : "FAIL", c.name);
  }
  std::printf("%zu cases, %d failed\n", cases().size(), failed_cases);
  return failed_cases ? 1 : 0;
}

}  // namespace check

#define CHECK_CAT2(a, b) a##b
#define CHECK_CAT(a, b) CHECK_CAT2(a, b)
#define TEST_CASE(name)                                                       \
  static void CHECK_CAT(test_fn_, __LINE__)();                                \
  static check::Register CHECK_CAT(test_reg_, __LINE__)(name, CHECK_CAT(test_fn_, __LINE__)); \
  static void CHECK_CAT(test_fn_, __LINE__)()
#define REQUIRE(expr)                                                          \
  do {                                                                         \
    if (!(expr)) {                                                             \
      std::printf("  %s:%d REQUIRE(%s) failed\n", __FILE__, __LINE__, #expr);  \
      ++check::failures();                                                     \
      return;                                                                  \
    }                                                                          \
  } while (0)
#define REQUIRE_FALSE(expr) REQUIRE(!(expr))
#define REQUIRE_THROWS(expr)                                                   \
  do {                                                                         \
    bool threw_ = false;                                                       \
    try {                                                                      \
      expr;                                                                    \
    } catch (...) {                                                            \
      threw_ = true;                                                           \
    }                                                                          \
    REQUIRE(threw_);                                                           \
  } while (0)
using check::Approx;
