// Checks librbg's decimal conversion (rabbitbin_amd/csrc/rbg_decimal.h, the
// depth-file parser's strtof / strtod) against the C library on the host:
// random numbers in the forms depth files hold (fixed, %g, %e, integers,
// signs, leading white space, trailing text) plus inf / nan spellings.  Every
// conversion the header accepts must give glibc's bits and end position; it
// may refuse only the documented cases (more than 19 significant digits,
// magnitudes outside its exact range).  Prints "ok <checked> <refused>" or the
// first mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>

#include "../../rabbitbin_amd/csrc/rbg_decimal.h"

static bool check(const std::string& s, long& refused) {
  const uint8_t* b = (const uint8_t*)s.data();
  const uint8_t* e = b + s.size();
  for (int p : {24, 53}) {
    uint64_t bits = 0, used = 0;
    const int st = rbg::dec::parse_decimal(b, e, p, bits, &used);
    char* ep = nullptr;
    uint64_t want = 0;
    if (p == 24) {
      const float f = std::strtof(s.c_str(), &ep);
      uint32_t w;
      std::memcpy(&w, &f, 4);
      want = w;
    } else {
      const double f = std::strtod(s.c_str(), &ep);
      std::memcpy(&want, &f, 8);
    }
    const uint64_t want_used = (uint64_t)(ep - s.c_str());
    if (st != rbg::dec::kOk) {
      ++refused;
      continue;
    }
    if (bits != want || used != want_used) {
      std::printf("MISMATCH p=%d '%s': got %016llx used %llu, glibc %016llx used %llu\n", p, s.c_str(),
                  (unsigned long long)bits, (unsigned long long)used, (unsigned long long)want,
                  (unsigned long long)want_used);
      return false;
    }
  }
  return true;
}

int main(int argc, char** argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 200000;
  std::mt19937_64 rng(12345);
  long refused = 0, checked = 0;
  const char* fixed[] = {"0", "-0", "0.0", "+5", "1e", "1e+", "1e-", ".5", "5.", ".", "-", "+", "abc", "inf", "-INF",
                         "Infinity", "infinit", "nan", "NaN", "-nan", "in", "na", " \t\n 3.25", "14.2158", "16.817",
                         "3.4028235e38", "3.4028236e38", "1e39", "0.1", "2.52094e+06", "999999999999999999",
                         "1.17549435e-38", "4.9e-324", "1e308", "1.7976931348623157e308", "0.000001234", "7e-5x",
                         "1.5\r", "12\t34", "0012.50", "1.2.3", "5e2.5"};
  for (const char* f : fixed) {
    if (!check(f, refused)) return 1;
    ++checked;
  }
  char buf[96];
  for (long t = 0; t < n; ++t) {
    const int kind = (int)(rng() % 6);
    const double mant = std::ldexp((double)(rng() >> 11), -53);  // [0, 1)
    const int ex = (int)(rng() % 61) - 30;
    const double v = (mant + (double)(rng() % 1000)) * std::pow(10.0, ex / 3);
    const int prec = (int)(rng() % 12);
    switch (kind) {
      case 0: std::snprintf(buf, sizeof buf, "%.*f", prec, v); break;
      case 1: std::snprintf(buf, sizeof buf, "%.*g", prec + 1, v); break;
      case 2: std::snprintf(buf, sizeof buf, "%.*e", prec, v); break;
      case 3: std::snprintf(buf, sizeof buf, "%llu", (unsigned long long)(rng() % 100000000000ull)); break;
      case 4: {  // random digit strings with a point and an exponent
        std::string s;
        const int nd = 1 + (int)(rng() % 22);
        for (int k = 0; k < nd; ++k) s += (char)('0' + rng() % 10);
        if (rng() % 2) s.insert(rng() % (s.size() + 1), ".");
        if (rng() % 2) s += "e" + std::to_string((int)(rng() % 70) - 35);
        std::snprintf(buf, sizeof buf, "%s", s.c_str());
        break;
      }
      default: std::snprintf(buf, sizeof buf, "%.17g", v); break;
    }
    std::string s = buf;
    if (rng() % 4 == 0) s = "-" + s;
    if (rng() % 8 == 0) s = "  " + s;
    if (rng() % 8 == 0) s += (rng() % 2) ? "\t9" : "x";
    if (!check(s, refused)) return 1;
    ++checked;
  }
  std::printf("ok %ld %ld\n", checked, refused);
  return 0;
}
