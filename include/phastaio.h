/*
 * phastaio.h -- C-ABI of the host library for PHASTA's POSIX "phastaIO"
 * file format (the geombc.dat.<part> and restart.<step>.<part> files).
 *
 * A file is a sequence of records.  Each record is one text header line
 *     <keyphrase> : < <bytes> > <int> <int> ... \n
 * followed, when <bytes> > 0, by a data block: in binary files <bytes> =
 * item size x items + 1 bytes of raw native-order data ending in '\n'; in
 * ASCII files <bytes> lines of one item each.  Lines starting with '#' are
 * comments.  A record "byteorder magic number" holding the integer 362436
 * tells the reader whether the data blocks need their bytes swapped.
 *
 * The entry points mirror the reference's POSIX branch of phastaIO.cc
 * (same argument meaning; the reference binds them to Fortran as openfile,
 * closefile, readheader, readdatablock, writeheader, writedatablock,
 * writestring, togglestrictmode, isLittleEndian).  They return 0 on success
 * and a nonzero code otherwise (the reference prints and carries on;
 * phastaio_last_error() holds the message here).  Not thread-safe: one
 * table of open files per process, as in the reference.
 */
#ifndef PHASTAIO_H
#define PHASTAIO_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PHASTAIO_OK = 0,
  PHASTAIO_NOTFOUND = 1, /* readheader: keyphrase not in the file            */
  PHASTAIO_EOPEN = 2,    /* the file could not be opened                     */
  PHASTAIO_EBADFD = 3,   /* no open file behind the descriptor               */
  PHASTAIO_EINVAL = 4,   /* unknown datatype / iotype / mode                 */
  PHASTAIO_EKEY = 5,     /* strict mode: data block key differs from header  */
  PHASTAIO_EIO = 6       /* short read or write                              */
};

/* openfile (phastaIO.cc:916-952): mode "read", "write" or "append";
 * *fd > 0 on success, 0 on failure. */
int phastaio_openfile(const char* filename, const char* mode, int* fd);
/* closefile (phastaIO.cc:1120-1135) */
int phastaio_closefile(int* fd, const char* mode);
/* readheader (phastaIO.cc:1271-1300, readHeader :222-301): searches
 * forward for the record whose keyphrase matches (case-insensitive, blanks
 * ignored, the search key matching a prefix of the file's; a '?' ends the
 * comparison), wrapping to the start of the file once; copies up to
 * *nItems header integers into valueArray.  datatype "integer" / "double"
 * and iotype "binary" / "ascii" describe the data block that follows. */
int phastaio_readheader(int* fd, const char* keyphrase, void* valueArray, int* nItems,
                        const char* datatype, const char* iotype);
/* readdatablock (phastaIO.cc:1483-1550, readDataBlock :1450-1481): reads
 * *nItems items of the block after the last header read; swaps bytes when
 * the file's magic number came out reversed. */
int phastaio_readdatablock(int* fd, const char* keyphrase, void* valueArray, int* nItems,
                           const char* datatype, const char* iotype);
/* writeheader (phastaIO.cc:1646-1680, writeHeader :1624-1644): *nItems
 * header integers; the block that follows holds *ndataItems items. */
int phastaio_writeheader(const int* fd, const char* keyphrase, const void* valueArray,
                         const int* nItems, const int* ndataItems, const char* datatype,
                         const char* iotype);
/* writedatablock (phastaIO.cc:1852-1890, writeDataBlock :1819-1850);
 * ASCII doubles are printed with "%f", as in the reference. */
int phastaio_writedatablock(const int* fd, const char* keyphrase, const void* valueArray,
                            const int* nItems, const char* datatype, const char* iotype);
/* writestring (phastaIO.cc:2011-2017): raw text, e.g. a '#' comment line */
int phastaio_writestring(int* fd, const char* text);
/* togglestrictmode (phastaIO.cc:2042): when on, readdatablock refuses a
 * keyphrase that differs from the last header's (off by default) */
void phastaio_togglestrictmode(void);
/* SwapArrayByteOrder (phastaIO.cc:1998-2009) */
void phastaio_swap_byte_order(void* array, int nbytes, int nItems);
/* isLittleEndian (phastaIO.cc:2044-2056) */
int phastaio_is_little_endian(void);
const char* phastaio_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
