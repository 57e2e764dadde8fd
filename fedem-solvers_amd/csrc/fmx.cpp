// .fmx / .fsm binary matrix files read by fedem_solver / fedem_stress
// (writeDoubleDBfile, binaryDBInterface.f90:452-500; writetagdb,
// binaryDB.c:643-674): a file tag "#FEDEM <type>" with the part checksum,
// the ";1.0;\n" version field, then the column-major FP64 data.
//
// The tag itself is written by FFaTag, which lives in the absent
// fedem-foundation; its byte layout is therefore not reproduced.  Here the
// checksum follows the tag text as ";<checksum>" (decimal), which this
// library's reader parses back.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/fedem_reducer_gpu.h"

namespace fgpu {

int write_fmx(const std::string& file, const std::string& dtype, unsigned checksum, const double* data, size_t n) {
  FILE* f = std::fopen(file.c_str(), "wb");
  if (!f) return -1;
  std::string head = "#FEDEM " + dtype + ";" + std::to_string(checksum) + ";1.0;\n";
  bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size();
  ok = ok && (n == 0 || std::fwrite(data, sizeof(double), n, f) == n);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : -2;
}

int read_fmx(const std::string& file, const std::string& dtype, double* data, size_t n, unsigned* checksum) {
  FILE* f = std::fopen(file.c_str(), "rb");
  if (!f) return -1;
  std::string head;
  int c;
  while ((c = std::fgetc(f)) != EOF && c != '\n') head.push_back((char)c);
  std::string want = "#FEDEM " + dtype + ";";
  if (head.compare(0, want.size(), want) != 0) {
    std::fclose(f);
    return -3;
  }
  if (checksum) *checksum = (unsigned)std::strtoul(head.c_str() + want.size(), nullptr, 10);
  size_t got = n ? std::fread(data, sizeof(double), n, f) : 0;
  std::fclose(f);
  return got == n ? 0 : -4;
}

}  // namespace fgpu

// Fortran-callable FMX I/O of the reference's fmxWriter library
// (src/vpmReducer/fmxWriter.f90, loaded by fedempy.write_fmx as
// writefmx_/readfmx_): itype 1 stiffness (_K.fmx), 2 mass (_M.fmx),
// 3 gravity force vectors (_G.fmx); the trailing hidden argument is the
// Fortran length of the prefix string.
static int fmx_io(const char* prefix, const int* itype, double* data, const int* nval, size_t len, bool wr) {
  if (!prefix || !itype || !data || !nval) return -1;
  std::string p(prefix, strnlen(prefix, len));
  while (!p.empty() && p.back() == ' ') p.pop_back();
  const char* suffix = *itype == 1 ? "_K.fmx" : *itype == 2 ? "_M.fmx" : *itype == 3 ? "_G.fmx" : nullptr;
  const char* tag = *itype == 1 ? "stiffness matrix" : *itype == 2 ? "mass matrix" : "gravity force vectors";
  if (!suffix) return -std::max(1, *itype);
  if (wr) return fgpu::write_fmx(p + suffix, tag, 0u, data, (size_t)*nval);
  return fgpu::read_fmx(p + suffix, tag, data, (size_t)*nval, nullptr);
}

extern "C" int writefmx_(const char* prefix, const int* itype, double* data, const int* nval, size_t len) {
  return fmx_io(prefix, itype, data, nval, len, true);
}

extern "C" int readfmx_(const char* prefix, const int* itype, double* data, const int* nval, size_t len) {
  return fmx_io(prefix, itype, data, nval, len, false);
}
