// Binary files read by fedem_solver / fedem_stress: the .fmx matrices
// (writeDoubleDBfile, binaryDBInterface.f90:452-500) and the _SAM.fsm file
// (saveSAM, samReducerModule.f90:586-660).  Each starts with the file tag
// that writetagdb (binaryDB.c:643-674) writes through FFa_writeTag, followed
// by the ";1.0;\n" version field and the data.
//
// Tag layout (40 bytes; the reference's data file
// src/vpmSolver/vpmSolverTests/models/motions.fnd, written by
// createTestFnd.C:38-40 with the tag "#FEDEM nodal displacements" and the
// checksum 12345678, holds exactly these bytes):
//   [0, 30)   the tag text, blank-padded
//   [30, 36)  endian marker: int32 0x1234 in the writer's byte order, then
//             two zero bytes (34 12 00 00 00 00 from a little-endian host)
//   [36, 40)  the checksum, uint32 in the writer's byte order
// readtagdb (binaryDB.c:676-733) returns the marker and byte-swaps the
// checksum when the marker is the other order; the version field is read
// with fgets up to the newline and parsed as ";%f;%d;".
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/fedem_reducer_gpu.h"

namespace fgpu {

constexpr int TAG_CHARS = 30;   // blank-padded tag text
constexpr int TAG_BYTES = 40;   // text + endian marker + checksum
constexpr uint32_t ENDIAN_MARK = 0x1234;

static uint32_t bswap32(uint32_t v) {
  return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

// The 40 tag bytes of `tag` (at most 30 characters) and checksum cs.
int write_tag(FILE* f, const std::string& tag, unsigned cs) {
  if (tag.size() > (size_t)TAG_CHARS) return -1;
  unsigned char h[TAG_BYTES];
  std::memset(h, ' ', TAG_CHARS);
  std::memcpy(h, tag.data(), tag.size());
  const uint32_t mark = ENDIAN_MARK, c = cs;
  std::memcpy(h + TAG_CHARS, &mark, 4);
  h[TAG_CHARS + 4] = h[TAG_CHARS + 5] = 0;
  std::memcpy(h + TAG_CHARS + 6, &c, 4);
  return std::fwrite(h, 1, TAG_BYTES, f) == (size_t)TAG_BYTES ? 0 : -2;
}

// Reads a tag: its text with trailing blanks removed, the checksum in host
// order and the version line (up to and including the newline, or up to
// 63 characters).  Returns 0, -3 when the file is not a tagged binary file
// (no endian marker in either byte order), -4 on a short file.
int read_tag(FILE* f, std::string& tag, unsigned& cs, std::string& version) {
  unsigned char h[TAG_BYTES];
  if (std::fread(h, 1, TAG_BYTES, f) != (size_t)TAG_BYTES) return -4;
  uint32_t mark, c;
  std::memcpy(&mark, h + TAG_CHARS, 4);
  std::memcpy(&c, h + TAG_CHARS + 6, 4);
  if (mark == bswap32(ENDIAN_MARK)) c = bswap32(c);
  else if (mark != ENDIAN_MARK) return -3;
  if (h[TAG_CHARS + 4] || h[TAG_CHARS + 5]) return -3;
  int n = TAG_CHARS;
  while (n > 0 && (h[n - 1] == ' ' || h[n - 1] == 0)) n--;
  tag.assign((const char*)h, n);
  cs = c;
  version.clear();
  int ch;
  while (version.size() < 63 && (ch = std::fgetc(f)) != EOF) {
    version.push_back((char)ch);
    if (ch == '\n') break;
  }
  return 0;
}

int write_fmx(const std::string& file, const std::string& dtype, unsigned checksum, const double* data, size_t n) {
  FILE* f = std::fopen(file.c_str(), "wb");
  if (!f) return -1;
  static const char ver[] = ";1.0;\n";
  bool ok = write_tag(f, "#FEDEM " + dtype, checksum) == 0;
  ok = ok && std::fwrite(ver, 1, sizeof ver - 1, f) == sizeof ver - 1;
  ok = ok && (n == 0 || std::fwrite(data, sizeof(double), n, f) == n);
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : -2;
}

int read_fmx(const std::string& file, const std::string& dtype, double* data, size_t n, unsigned* checksum) {
  FILE* f = std::fopen(file.c_str(), "rb");
  if (!f) return -1;
  std::string tag, ver;
  unsigned cs = 0;
  int rc = read_tag(f, tag, cs, ver);
  if (rc == 0 && (tag != "#FEDEM " + dtype || ver != ";1.0;\n")) rc = -3;
  if (rc) {
    std::fclose(f);
    return rc;
  }
  if (checksum) *checksum = cs;
  size_t got = n ? std::fread(data, sizeof(double), n, f) : 0;
  std::fclose(f);
  return got == n ? 0 : -4;
}

}  // namespace fgpu

extern "C" {

// writetagdb (binaryDB.c:643-674) as a file service: a new file holding the
// tag and the version field `version` (the reducer's files use ";1.0;\n").
int fgpu_write_tag(const char* file, const char* tag, unsigned checksum, const char* version) {
  if (!file || !tag || !version) return -1;
  FILE* f = std::fopen(file, "wb");
  if (!f) return -1;
  const size_t nv = std::strlen(version);
  bool ok = fgpu::write_tag(f, tag, checksum) == 0;
  ok = ok && std::fwrite(version, 1, nv, f) == nv;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? 0 : -2;
}

// readtagdb (binaryDB.c:676-733): the tag text (NUL-terminated, trailing
// blanks removed), the checksum and the version line of a tagged binary
// file.  Returns the byte offset of the data after the version line, or a
// negative error (-1 open/arguments, -3 not a tagged file, -4 short file).
int64_t fgpu_read_tag(const char* file, char* tag, int tag_len, unsigned* checksum, char* version, int ver_len) {
  if (!file || !tag || tag_len < 1 || !checksum || !version || ver_len < 1) return -1;
  FILE* f = std::fopen(file, "rb");
  if (!f) return -1;
  std::string t, v;
  unsigned cs = 0;
  int rc = fgpu::read_tag(f, t, cs, v);
  const long pos = std::ftell(f);
  std::fclose(f);
  if (rc) return rc;
  std::snprintf(tag, (size_t)tag_len, "%s", t.c_str());
  std::snprintf(version, (size_t)ver_len, "%s", v.c_str());
  *checksum = cs;
  return pos;
}

}  // extern "C"

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
