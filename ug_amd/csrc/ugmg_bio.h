// The three encodings of UG's basic file i/o (low/bio.cc): XDR -- 4-byte
// big-endian ints, 8-byte big-endian IEEE doubles, strings "%d " + chars +
// "\n"; ASCII -- ints "%d\n", doubles "%g\n", strings "%d\n" + chars + " ";
// BIN -- native ints and doubles, strings "%d " + chars + " ".  Shared by
// the data-file (ugmg_dio.cc) and multigrid-file (ugmg_mgio.cc) readers.
// Host only.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>

namespace ugmg_bio {

enum { XDR = 0, ASCII = 1, BIN = 2 };   // BIO_XDR, BIO_ASCII, BIO_BIN (low/bio.h:68-70)

struct FileCloser {
  void operator()(FILE *f) const {
    if (f) std::fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

inline uint32_t bswap32(uint32_t v) {
  return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}
inline uint64_t bswap64(uint64_t v) {
  return ((uint64_t)bswap32((uint32_t)v) << 32) | bswap32((uint32_t)(v >> 32));
}
inline bool host_little_endian() {
  const uint16_t one = 1;
  unsigned char c;
  std::memcpy(&c, &one, 1);
  return c == 1;
}
// XDR: big-endian on the wire (RFC 4506 unsigned int and double)
inline uint32_t to_be32(uint32_t v) { return host_little_endian() ? bswap32(v) : v; }
inline uint64_t to_be64(uint64_t v) { return host_little_endian() ? bswap64(v) : v; }

// one encoding of low/bio.cc over a FILE*
struct Bio {
  FILE *f;
  int mode;

  bool read_ints(int n, int *v) {
    for (int i = 0; i < n; ++i) {
      if (mode == ASCII) {
        if (std::fscanf(f, "%d\n", v + i) != 1) return false;
      } else if (mode == BIN) {
        if (std::fread(v + i, sizeof(int), 1, f) != 1) return false;
      } else {
        uint32_t w;
        if (std::fread(&w, 4, 1, f) != 1) return false;
        const uint32_t u = to_be32(w);
        std::memcpy(v + i, &u, 4);
      }
    }
    return true;
  }
  bool write_ints(int n, const int *v) {
    for (int i = 0; i < n; ++i) {
      if (mode == ASCII) {
        if (std::fprintf(f, "%d\n", v[i]) < 0) return false;
      } else if (mode == BIN) {
        if (std::fwrite(v + i, sizeof(int), 1, f) != 1) return false;
      } else {
        uint32_t u;
        std::memcpy(&u, v + i, 4);
        const uint32_t w = to_be32(u);
        if (std::fwrite(&w, 4, 1, f) != 1) return false;
      }
    }
    return true;
  }
  bool read_doubles(long long n, double *v) {
    if (mode == BIN) return n == 0 || std::fread(v, sizeof(double), n, f) == (size_t)n;
    for (long long i = 0; i < n; ++i) {
      if (mode == ASCII) {
        if (std::fscanf(f, "%lg\n", v + i) != 1) return false;
      } else {
        uint64_t w;
        if (std::fread(&w, 8, 1, f) != 1) return false;
        const uint64_t u = to_be64(w);
        std::memcpy(v + i, &u, 8);
      }
    }
    return true;
  }
  bool write_doubles(long long n, const double *v) {
    if (mode == BIN) return n == 0 || std::fwrite(v, sizeof(double), n, f) == (size_t)n;
    for (long long i = 0; i < n; ++i) {
      if (mode == ASCII) {
        if (std::fprintf(f, "%g\n", v[i]) < 0) return false;
      } else {
        uint64_t u;
        std::memcpy(&u, v + i, 8);
        const uint64_t w = to_be64(u);
        if (std::fwrite(&w, 8, 1, f) != 1) return false;
      }
    }
    return true;
  }
  // length, then the characters, then one terminator: ASCII "%d\n" ... ' ',
  // BIN "%d " ... ' ', XDR "%d " ... '\n'
  bool read_string(char *s, int cap) {
    int len;
    if (std::fscanf(f, mode == ASCII ? "%d\n" : "%d ", &len) != 1) return false;
    if (len < 0 || len >= cap) return false;
    for (int i = 0; i < len; ++i) {
      const int c = std::fgetc(f);
      if (c == EOF) return false;
      s[i] = (char)c;
    }
    const int t = std::fgetc(f);
    if (t != (mode == XDR ? '\n' : ' ')) return false;
    s[len] = '\0';
    return true;
  }
  bool write_string(const char *s) {
    const int len = (int)std::strlen(s);
    if (std::fprintf(f, mode == ASCII ? "%d\n" : "%d ", len) < 0) return false;
    if (len && std::fwrite(s, 1, len, f) != (size_t)len) return false;
    return std::fputc(mode == XDR ? '\n' : ' ', f) != EOF;
  }
  // Bio_Jump(1) (bio.cc:456-470): the byte count Bio_Jump_From/_To left as
  // " %20d " in front of a section, then that many bytes skipped
  bool skip_jump() {
    int n;
    if (std::fscanf(f, " %d", &n) != 1 || n < 0) return false;
    if (std::fgetc(f) != ' ') return false;
    return std::fseek(f, n, SEEK_CUR) == 0;
  }
};

inline bool valid_mode(int m) { return m == XDR || m == ASCII || m == BIN; }

}  // namespace ugmg_bio
