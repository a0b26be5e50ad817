// fugu2d_io.hpp -- host-side file formats of the reference's I/O module (src/mod_io.f90) at one rank:
// the stream checkpoint and the legacy-ASCII VTK file, written byte for byte as the reference's
// Fortran edit descriptors produce them.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace fugu {

// Fortran ES editing (Fortran 2008, 10.7.2.3.2) of the reference's '(ESw.d)': the correctly rounded
// significand d.ddd (d digits after the point), an exponent E+dd for |e| <= 99 and +ddd beyond, right
// justified in w columns (w asterisks when it does not fit).  NaN and infinities as the Fortran
// runtime of this image spells them ("NaN", "Inf", "-Inf").
inline void fortran_es(std::string& out, double x, int w = 15, int d = 7) {
  char s[64];
  if (std::isnan(x)) {
    std::snprintf(s, sizeof s, "NaN");
  } else if (std::isinf(x)) {
    std::snprintf(s, sizeof s, "%s", x > 0 ? "Inf" : "-Inf");
  } else {
    char m[64];
    std::snprintf(m, sizeof m, "%.*E", d, x);   // correctly rounded (ties to even), e.g. -1.2345678E-100
    char* e = std::strchr(m, 'E');
    const int ex = std::atoi(e + 1);
    *e = '\0';
    if (ex >= -99 && ex <= 99)
      std::snprintf(s, sizeof s, "%sE%c%02d", m, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
    else
      std::snprintf(s, sizeof s, "%s%c%03d", m, ex < 0 ? '-' : '+', ex < 0 ? -ex : ex);
  }
  const int n = (int)std::strlen(s);
  if (n > w) {
    out.append((size_t)w, '*');
  } else {
    out.append((size_t)(w - n), ' ');
    out.append(s, (size_t)n);
  }
}
inline void fortran_i(std::string& out, long long v, int w) {
  char s[32];
  std::snprintf(s, sizeof s, "%lld", v);
  const int n = (int)std::strlen(s);
  if (n > w) out.append((size_t)w, '*');
  else { out.append((size_t)(w - n), ' '); out.append(s, (size_t)n); }
}

// "<prefix>_r0000.bin": the rank-0 name of src/mod_io.f90:181 ('(A,A,I4.4,A)').
inline std::string checkpoint_name(const char* prefix) { return std::string(prefix) + "_r0000.bin"; }

// The 24-byte stream header of src/mod_io.f90:184: int32 step, real64 t, int32 nx, ny, nh.
struct CkptHeader {
  int32_t step;
  double t;
  int32_t nx, ny, nh;
};
inline void header_bytes(const CkptHeader& h, unsigned char (&b)[24]) {
  std::memcpy(b, &h.step, 4);
  std::memcpy(b + 4, &h.t, 8);
  std::memcpy(b + 12, &h.nx, 4);
  std::memcpy(b + 16, &h.ny, 4);
  std::memcpy(b + 20, &h.nh, 4);
}
inline CkptHeader header_from(const unsigned char (&b)[24]) {
  CkptHeader h;
  std::memcpy(&h.step, b, 4);
  std::memcpy(&h.t, b + 4, 8);
  std::memcpy(&h.nx, b + 12, 4);
  std::memcpy(&h.ny, b + 16, 4);
  std::memcpy(&h.nh, b + 20, 4);
  return h;
}

}  // namespace fugu
