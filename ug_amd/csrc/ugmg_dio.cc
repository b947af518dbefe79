// UG's data files (SaveData / LoadData, np/udm/data_io.cc:408-1019): the
// DATA_IO header (Read_DT_General / Write_DT_General, np/udm/dio.cc:269-370)
// and the node data, in the three encodings of low/bio.cc.  Host only.
#include "../../include/ugmg_dio.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>

namespace {

constexpr const char *kTitle = "####.sparse.data.storage.format.####";   // dio.cc:61

struct FileCloser {
  void operator()(FILE *f) const {
    if (f) std::fclose(f);
  }
};
using File = std::unique_ptr<FILE, FileCloser>;

uint32_t bswap32(uint32_t v) {
  return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}
uint64_t bswap64(uint64_t v) {
  return ((uint64_t)bswap32((uint32_t)v) << 32) | bswap32((uint32_t)(v >> 32));
}
bool host_little_endian() {
  const uint16_t one = 1;
  unsigned char c;
  std::memcpy(&c, &one, 1);
  return c == 1;
}
// XDR: big-endian on the wire (RFC 4506 unsigned int and double)
uint32_t to_be32(uint32_t v) { return host_little_endian() ? bswap32(v) : v; }
uint64_t to_be64(uint64_t v) { return host_little_endian() ? bswap64(v) : v; }

// one encoding of low/bio.cc over a FILE*
struct Bio {
  FILE *f;
  int mode;

  bool read_ints(int n, int *v) {
    for (int i = 0; i < n; ++i) {
      if (mode == UGMG_DIO_ASCII) {
        if (std::fscanf(f, "%d\n", v + i) != 1) return false;
      } else if (mode == UGMG_DIO_BIN) {
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
      if (mode == UGMG_DIO_ASCII) {
        if (std::fprintf(f, "%d\n", v[i]) < 0) return false;
      } else if (mode == UGMG_DIO_BIN) {
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
    if (mode == UGMG_DIO_BIN) return n == 0 || std::fread(v, sizeof(double), n, f) == (size_t)n;
    for (long long i = 0; i < n; ++i) {
      if (mode == UGMG_DIO_ASCII) {
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
    if (mode == UGMG_DIO_BIN) return n == 0 || std::fwrite(v, sizeof(double), n, f) == (size_t)n;
    for (long long i = 0; i < n; ++i) {
      if (mode == UGMG_DIO_ASCII) {
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
    if (std::fscanf(f, mode == UGMG_DIO_ASCII ? "%d\n" : "%d ", &len) != 1) return false;
    if (len < 0 || len >= cap) return false;
    for (int i = 0; i < len; ++i) {
      const int c = std::fgetc(f);
      if (c == EOF) return false;
      s[i] = (char)c;
    }
    const int t = std::fgetc(f);
    if (t != (mode == UGMG_DIO_XDR ? '\n' : ' ')) return false;
    s[len] = '\0';
    return true;
  }
  bool write_string(const char *s) {
    const int len = (int)std::strlen(s);
    if (std::fprintf(f, mode == UGMG_DIO_ASCII ? "%d\n" : "%d ", len) < 0) return false;
    if (len && std::fwrite(s, 1, len, f) != (size_t)len) return false;
    return std::fputc(mode == UGMG_DIO_XDR ? '\n' : ' ', f) != EOF;
  }
};

bool valid_mode(int m) { return m == UGMG_DIO_XDR || m == UGMG_DIO_ASCII || m == UGMG_DIO_BIN; }

}  // namespace

extern "C" int ugmg_dio_read(const char *path, ugmg_dio_header *h, double *data,
                             long long capacity) {
  if (!path || !h) return -1;
  File f(std::fopen(path, "rb"));
  if (!f) return -2;
  std::memset(h, 0, sizeof(*h));
  Bio io{f.get(), UGMG_DIO_ASCII};   // the head is always ASCII
  char title[UGMG_DIO_NAMELEN];
  if (!io.read_string(title, sizeof title) || std::strcmp(title, kTitle) != 0) return -1;
  int mode;
  if (!io.read_ints(1, &mode) || !valid_mode(mode)) return -1;
  h->mode = mode;
  io.mode = mode;
  if (!io.read_string(h->version, sizeof h->version)) return -1;
  if (std::strcmp(h->version, "DATA_IO_1.6") == 0) {   // no ident in 1.6 files
    std::strcpy(h->version, "DATA_IO_1.7");
  } else if (!io.read_string(h->ident, sizeof h->ident)) {
    return -1;
  }
  if (!io.read_string(h->mgfile, sizeof h->mgfile)) return -1;
  if (!io.read_doubles(1, &h->time) || !io.read_doubles(1, &h->dt) || !io.read_doubles(1, &h->ndt))
    return -1;
  int iv[4];
  if (!io.read_ints(4, iv)) return -1;
  h->nparfiles = iv[0];
  h->me = iv[1];
  h->magic_cookie = iv[2];
  h->nvd = iv[3];
  if (h->nvd < 0 || h->nvd > UGMG_DIO_VDMAX) return -1;
  for (int i = 0; i < h->nvd; ++i) {
    if (!io.read_string(h->vd_name[i], UGMG_DIO_NAMELEN) || !io.read_ints(1, h->vd_ncomp + i) ||
        !io.read_ints(1, h->vd_type + i) || !io.read_string(h->vd_comp_names[i], UGMG_DIO_NAMELEN))
      return -1;
  }
  if (!io.read_ints(1, &h->ndata) || h->ndata < 0) return -1;
  if (!data) return 0;
  if (capacity < h->ndata) return -1;
  return io.read_doubles(h->ndata, data) ? 0 : -1;
}

extern "C" int ugmg_dio_write(const char *path, const ugmg_dio_header *h, const double *data) {
  if (!path || !h || !valid_mode(h->mode) || h->nvd < 0 || h->nvd > UGMG_DIO_VDMAX ||
      h->ndata < 0 || (h->ndata > 0 && !data))
    return -1;
  File f(std::fopen(path, "wb"));
  if (!f) return -2;
  Bio io{f.get(), UGMG_DIO_ASCII};
  if (!io.write_string(kTitle) || !io.write_ints(1, &h->mode)) return -1;
  io.mode = h->mode;
  if (!io.write_string(h->version) || !io.write_string(h->ident) || !io.write_string(h->mgfile))
    return -1;
  if (!io.write_doubles(1, &h->time) || !io.write_doubles(1, &h->dt) || !io.write_doubles(1, &h->ndt))
    return -1;
  const int iv[4] = {h->nparfiles, h->me, h->magic_cookie, h->nvd};
  if (!io.write_ints(4, iv)) return -1;
  for (int i = 0; i < h->nvd; ++i) {
    if (!io.write_string(h->vd_name[i]) || !io.write_ints(1, h->vd_ncomp + i) ||
        !io.write_ints(1, h->vd_type + i) || !io.write_string(h->vd_comp_names[i]))
      return -1;
  }
  if (!io.write_ints(1, &h->ndata) || !io.write_doubles(h->ndata, data)) return -1;
  return std::fflush(f.get()) == 0 ? 0 : -1;
}
