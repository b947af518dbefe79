// UG's data files (SaveData / LoadData, np/udm/data_io.cc:408-1019): the
// DATA_IO header (Read_DT_General / Write_DT_General, np/udm/dio.cc:269-370)
// and the node data, in the three encodings of low/bio.cc.  Host only.
#include "../../include/ugmg_dio.h"
#include "ugmg_bio.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>

namespace {

constexpr const char *kTitle = "####.sparse.data.storage.format.####";   // dio.cc:61
using ugmg_bio::Bio;
using ugmg_bio::File;

bool valid_mode(int m) { return ugmg_bio::valid_mode(m); }

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
