// fortran_format.hpp -- the Fortran edit descriptors the reference's ASCII
// output uses, as plain C++ (no device code), so that they can be checked
// against a Fortran compiler's own output.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace pt {

// Fortran Gw.dEe output editing with w=14, d=5, e=3 (Fortran 2008
// 10.7.5.2.2): a value whose 5-significant-digit rounding lies in [0.1, 1e5)
// is written F9.(5-k), k its integer digits, followed by 5 blanks; others
// E14.5E3 (0.ddddd and a 3-digit exponent); zero F9.4 and 5 blanks.
inline void g14_5e3(double x, std::string& out) {
  char buf[48];
  if (x == 0.0) {
    std::snprintf(buf, sizeof buf, "%9.4f     ", x);
    out += buf;
    return;
  }
  char e[32];
  std::snprintf(e, sizeof e, "%.4e", std::fabs(x));   // d.dddde+XX, correctly rounded
  const int e10 = std::atoi(std::strchr(e, 'e') + 1);
  const int k = e10 + 1;
  if (k >= 0 && k <= 4) {
    std::snprintf(buf, sizeof buf, "%9.*f     ", 5 - k, x);
  } else if (k == 5) {
    std::snprintf(buf, sizeof buf, "%#9.0f     ", x);
  } else {
    const char digits[6] = {e[0], e[2], e[3], e[4], e[5], 0};
    const int ex = e10 + 1;
    char m[24];
    std::snprintf(m, sizeof m, "%s0.%sE%c%03d", x < 0 ? "-" : "", digits, ex < 0 ? '-' : '+', std::abs(ex));
    std::snprintf(buf, sizeof buf, "%14s", m);
  }
  out += buf;
}

}  // namespace pt
