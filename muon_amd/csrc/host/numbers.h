// Exact fast reading of the decimal numbers in scene and mesh files: the
// same values as std::istream's operator>> (the reference parser's reader)
// for plain tokens, without building a stream per line.
#ifndef MUONX_HOST_NUMBERS_H_
#define MUONX_HOST_NUMBERS_H_

#include <cerrno>
#include <climits>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <type_traits>

namespace muonx {
namespace host {

inline bool IsSpace(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }

// Decimal token -> float with the correctly rounded result of strtof, for
// tokens [sign] digits [. digits] [(e|E) [sign] digits] whose significant
// digits fit 2^53 and whose decimal exponent is within +-22: the digits are
// an exact double, 10^|k| is an exact double, so one multiplication or
// division gives the correctly rounded double d of the value; rounding d to
// float gives the correctly rounded float unless d lies exactly halfway
// between two floats (a float midpoint is a double, so the value and d
// cannot straddle one otherwise). Returns false for every other token and
// for that halfway case.
inline bool FastDecimal(const char *b, size_t len, float &out) {
  static const double kPow10[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                    1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
  size_t i = 0;
  bool neg = false;
  if (i < len && (b[i] == '+' || b[i] == '-')) neg = b[i++] == '-';
  uint64_t m = 0;
  int digits = 0, frac = 0;
  bool any = false;
  while (i < len && b[i] >= '0' && b[i] <= '9') {
    any = true;
    if (m || b[i] != '0') {
      if (++digits > 17) return false;
      m = m * 10 + uint64_t(b[i] - '0');
    }
    ++i;
  }
  if (i < len && b[i] == '.') {
    ++i;
    while (i < len && b[i] >= '0' && b[i] <= '9') {
      any = true;
      if (m || b[i] != '0') {
        if (++digits > 17) return false;
        m = m * 10 + uint64_t(b[i] - '0');
      }
      ++frac;
      ++i;
    }
  }
  if (!any) return false;
  int ex = 0;
  if (i < len && (b[i] == 'e' || b[i] == 'E')) {
    ++i;
    bool eneg = false;
    if (i < len && (b[i] == '+' || b[i] == '-')) eneg = b[i++] == '-';
    if (i == len) return false;
    while (i < len && b[i] >= '0' && b[i] <= '9') {
      if (ex > 1000) return false;
      ex = ex * 10 + (b[i] - '0');
      ++i;
    }
    if (eneg) ex = -ex;
  }
  if (i != len || m >= (uint64_t(1) << 53)) return false;
  const int k = ex - frac;  // value = m * 10^k
  if (k < -22 || k > 22) return false;
  double d = double(m);
  d = k >= 0 ? d * kPow10[k] : d / kPow10[-k];
  uint64_t bits;
  std::memcpy(&bits, &d, sizeof(bits));
  if (m != 0 && (bits & ((uint64_t(1) << 29) - 1)) == (uint64_t(1) << 28)) return false;  // float midpoint
  const float f = float(d);
  out = neg ? -f : f;
  return true;
}

// Reads `n` whitespace-separated numbers from [s, e) for the hot commands
// (vertex, tri) without building a stream. A token is accepted only when it
// is made of digits, signs, '.', 'e' / 'E', and strtof / strtol reads all of
// it without a range error; it then has the value std::istream reads. On
// anything else the caller takes the istream path, which decides as before.
// Content after the n-th number is ignored, as the istream path ignores it.
template <class T>
inline bool FastNumbers(const char *s, const char *e, T *out, int n) {
  for (int k = 0; k < n; ++k) {
    while (s < e && IsSpace(*s)) ++s;
    const char *b = s;
    while (s < e && !IsSpace(*s)) {
      const char c = *s;
      const bool ok = (c >= '0' && c <= '9') || c == '+' || c == '-' ||
                      (std::is_floating_point<T>::value && (c == '.' || c == 'e' || c == 'E'));
      if (!ok) return false;
      ++s;
    }
    const size_t len = size_t(s - b);
    if (len == 0 || len > 63) return false;
    if constexpr (std::is_floating_point<T>::value) {
      if (FastDecimal(b, len, out[k])) continue;
    } else if (len <= 9 && b[0] != '+' && b[0] != '-') {
      // up to 9 plain digits: below 2^31, read exactly
      int64_t v = 0;
      size_t j = 0;
      while (j < len && b[j] >= '0' && b[j] <= '9') v = v * 10 + (b[j++] - '0');
      if (j == len) {
        out[k] = T(v);
        continue;
      }
    }
    char buf[64];
    std::memcpy(buf, b, len);
    buf[len] = '\0';
    char *end = nullptr;
    errno = 0;
    if constexpr (std::is_floating_point<T>::value) {
      const float v = std::strtof(buf, &end);
      if (end != buf + len || errno == ERANGE) return false;
      out[k] = v;
    } else {
      const long v = std::strtol(buf, &end, 10);
      if (end != buf + len || errno == ERANGE || v < INT_MIN || v > INT_MAX) return false;
      out[k] = T(v);
    }
  }
  return true;
}

}  // namespace host
}  // namespace muonx

#endif
