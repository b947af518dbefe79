/*
 * ugmg_dio.h -- UG's data files ("<name>.ug.data.<xdr|asc|bin>", written by
 * SaveData and read by LoadData, np/udm/data_io.cc:408-1019), host only.
 *
 * Layout (DATA_IO_1.7, np/udm/dio.cc:269-370, encodings of low/bio.cc):
 *   title string "####.sparse.data.storage.format.####" and the mode, both
 *   in the ASCII encoding; then, in the file's mode: version, ident (absent
 *   in DATA_IO_1.6 files), mgfile, time, dt, ndt, {nparfiles, me,
 *   magic_cookie, nVD}, per vector descriptor {name, ncomp, type,
 *   component names}, ndata; then ndata doubles: for every node in ID order
 *   the components of every descriptor in order (SaveData's node blocks).
 * Encodings (bio.cc): XDR -- 4-byte big-endian ints, 8-byte big-endian IEEE
 *   doubles, strings "%d " + chars + "\n"; ASCII -- ints "%d\n", doubles
 *   "%g\n" (6 significant digits, as UG writes them), strings "%d\n" + chars
 *   + " "; BIN -- native ints and doubles, strings "%d " + chars + " ".
 * Element and edge data that may follow the node data are not read.
 *
 * Which vector a node ID denotes is fixed by UG's RenumberMultiGrid
 * (gm/ugio.cc:660-760) on the grid that wrote the file; the caller maps IDs
 * to its level vectors.
 */
#ifndef UGMG_DIO_H
#define UGMG_DIO_H

#ifdef __cplusplus
extern "C" {
#endif

#define UGMG_DIO_XDR 0     /* BIO_XDR, low/bio.h:68 */
#define UGMG_DIO_ASCII 1   /* BIO_ASCII */
#define UGMG_DIO_BIN 2     /* BIO_BIN */

#define UGMG_DIO_VDMAX 100     /* DIO_VDMAX, np/udm/dio.h */
#define UGMG_DIO_NAMELEN 128   /* DIO_NAMELEN */
#define UGMG_DIO_IDENTLEN 4096 /* DIO_IDENTLEN */

/* DIO_GENERAL (np/udm/dio.h) */
typedef struct {
  int mode;
  char version[UGMG_DIO_NAMELEN];
  char mgfile[UGMG_DIO_NAMELEN];
  int nparfiles, me;
  double time, dt, ndt;
  int magic_cookie;
  char ident[UGMG_DIO_IDENTLEN];
  int nvd;
  char vd_name[UGMG_DIO_VDMAX][UGMG_DIO_NAMELEN];
  int vd_ncomp[UGMG_DIO_VDMAX];
  int vd_type[UGMG_DIO_VDMAX];   /* 0 scalar, 1 vector, 2 multiple scalar */
  char vd_comp_names[UGMG_DIO_VDMAX][UGMG_DIO_NAMELEN];
  int ndata;
} ugmg_dio_header;

/* Read_DT_General + the node data: the header into *h and, when data is
 * not NULL, the ndata doubles (capacity >= ndata).  0, or -1 for a file that
 * does not follow the layout (or capacity too small), -2 when it cannot be
 * opened. */
int ugmg_dio_read(const char *path, ugmg_dio_header *h, double *data, long long capacity);
/* Write_DT_General + ndata doubles, as SaveData writes them (h->version is
 * written as given; DATA_IO_1.7 for files UG's LoadData accepts). */
int ugmg_dio_write(const char *path, const ugmg_dio_header *h, const double *data);

#ifdef __cplusplus
}
#endif
#endif
