// Host implementation of include/phastaio.h: PHASTA's POSIX phastaIO format
// (the reference's POSIX branch of phastaIO/phastaIO.cc; the MPI-IO
// "SyncIO" variant is not part of this library).
#include "phastaio.h"

#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace {

constexpr int kMagic = 362436;  // "byteorder magic number" (phastaIO.cc:272)

struct OpenFile {
  FILE* f = nullptr;
  bool swap = false;          // the file's magic number came out reversed
  std::string last_key;       // keyphrase of the last header read or written
  bool have_key = false;
  bool last_not_found = false;
  size_t header_type = sizeof(int);
};

std::vector<OpenFile> g_files;
bool g_strict = false;
std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// first blank-delimited word (StringStripper, phastaIO.cc:167-176)
std::string first_word(const char* s) {
  std::string t(s);
  const size_t p = t.find(' ');
  return p == std::string::npos ? t : t.substr(0, p);
}

// keyphrase comparison (cscompare, phastaIO.cc:178-198): blanks skipped,
// case-insensitive, stops at a '?' in the target; true when the test
// string is used up (a prefix of the target) or stops at a '?'
bool key_match(const char* test, const char* target) {
  const char* s1 = test;
  const char* s2 = target;
  while (*s1 == ' ') s1++;
  while (*s2 == ' ') s2++;
  while (*s1 && *s2 && *s2 != '?' &&
         std::tolower((unsigned char)*s1) == std::tolower((unsigned char)*s2)) {
    s1++;
    s2++;
    while (*s1 == ' ') s1++;
    while (*s2 == ' ') s2++;
  }
  return !*s1 || *s1 == '?';
}

int type_size(const char* datatype, size_t* out) {
  const std::string t = first_word(datatype);
  if (key_match("integer", t.c_str())) {
    *out = sizeof(int);
    return 0;
  }
  if (key_match("double", t.c_str())) {
    *out = sizeof(double);
    return 0;
  }
  return fail(PHASTAIO_EINVAL, "unknown datatype '%s'", datatype);
}

int is_binary(const char* iotype, bool* out) {
  const std::string t = first_word(iotype);
  *out = key_match(t.c_str(), "binary");
  return 0;
}

OpenFile* lookup(int fd) {
  if (fd < 1 || fd > (int)g_files.size() || !g_files[fd - 1].f) return nullptr;
  return &g_files[fd - 1];
}

// next line; on end of file, rewind once more (counted)
bool next_line(FILE* f, char* line, int* rewinds) {
  if (fgets(line, 1024, f)) return true;
  if (!feof(f)) return false;
  rewind(f);
  clearerr(f);
  ++*rewinds;
  line[0] = '\0';
  return fgets(line, 1024, f) != nullptr;
}

// readHeader (phastaIO.cc:222-301)
int read_header(OpenFile& of, const char* phrase, int* params, int expect, bool binary) {
  FILE* f = of.f;
  char line[1024] = "";
  int rewinds = 0;
  next_line(f, line, &rewinds);
  bool found = false;
  while (!found && rewinds < 2) {
    const size_t len = strcspn(line, "#");
    if (line[0] != '\n' && line[0] != '\0' && len > 0) {
      std::string text(line, len);
      std::vector<char> buf(text.begin(), text.end());
      buf.push_back('\0');
      char* save = nullptr;
      char* token = strtok_r(buf.data(), ":", &save);
      if (token && key_match(phrase, token)) {
        found = true;
        token = strtok_r(nullptr, " ,;<>", &save);
        int i = 0;
        for (; i < expect && (token = strtok_r(nullptr, " ,;<>", &save)); ++i) params[i] = atoi(token);
        (void)i;  // fewer integers than asked: the rest are left as they were
      } else if (token && key_match(token, "byteorder magic number")) {
        if (binary) {
          int v = 0;
          char junk;
          if (fread(&v, sizeof(int), 1, f) != 1 || fread(&junk, 1, 1, f) != 1)
            return fail(PHASTAIO_EIO, "short read of the magic number");
          if (v != kMagic) of.swap = true;
        } else {
          int v = 0;
          if (fscanf(f, "%d\n", &v) != 1) return fail(PHASTAIO_EIO, "bad ASCII magic number");
        }
      } else if (token) {
        token = strtok_r(nullptr, " ,;<>", &save);
        const long skip = token ? atol(token) : 0;
        if (binary) {
          fseek(f, skip, SEEK_CUR);
        } else {
          char tmp[1024];
          for (long g = 0; g < skip; ++g)
            if (!fgets(tmp, 1024, f)) break;
        }
      }
    }
    if (!found && !next_line(f, line, &rewinds) && rewinds >= 2) break;
  }
  return found ? 0 : PHASTAIO_NOTFOUND;
}

}  // namespace

extern "C" {

const char* phastaio_last_error(void) { return g_err.c_str(); }

int phastaio_openfile(const char* filename, const char* mode, int* fd) {
  *fd = 0;
  const std::string m = first_word(mode);
  const char* fm = nullptr;
  if (key_match("read", m.c_str()))
    fm = "rb";
  else if (key_match("write", m.c_str()))
    fm = "wb";
  else if (key_match("append", m.c_str()))
    fm = "ab";
  else
    return fail(PHASTAIO_EINVAL, "openfile: unknown mode '%s'", mode);
  const std::string name = first_word(filename);
  FILE* f = fopen(name.c_str(), fm);
  if (!f) return fail(PHASTAIO_EOPEN, "openfile: unable to open file %s", name.c_str());
  OpenFile of;
  of.f = f;
  g_files.push_back(of);
  *fd = (int)g_files.size();
  return 0;
}

int phastaio_closefile(int* fd, const char* mode) {
  OpenFile* of = lookup(*fd);
  if (!of) return fail(PHASTAIO_EBADFD, "closefile: no file behind descriptor %d", *fd);
  (void)mode;
  fflush(of->f);
  fclose(of->f);
  *of = OpenFile();
  return 0;
}

int phastaio_readheader(int* fd, const char* keyphrase, void* valueArray, int* nItems,
                        const char* datatype, const char* iotype) {
  OpenFile* of = lookup(*fd);
  if (!of) return fail(PHASTAIO_EBADFD, "readheader: no file behind descriptor %d", *fd);
  bool binary;
  size_t ts;
  is_binary(iotype, &binary);
  if (int rc = type_size(datatype, &ts)) return rc;
  of->last_key = keyphrase;
  of->have_key = true;
  const int rc = read_header(*of, keyphrase, static_cast<int*>(valueArray), *nItems, binary);
  of->last_not_found = rc == PHASTAIO_NOTFOUND;
  if (rc == PHASTAIO_NOTFOUND) return fail(rc, "readheader: could not find '%s'", keyphrase);
  return rc;
}

int phastaio_readdatablock(int* fd, const char* keyphrase, void* valueArray, int* nItems,
                           const char* datatype, const char* iotype) {
  OpenFile* of = lookup(*fd);
  if (!of) return fail(PHASTAIO_EBADFD, "readdatablock: no file behind descriptor %d", *fd);
  if (!of->have_key || !key_match(of->last_key.c_str(), keyphrase)) {
    if (g_strict)
      return fail(PHASTAIO_EKEY, "readdatablock: header '%s' not consistent with block '%s'",
                  of->last_key.c_str(), keyphrase);
  }
  if (of->last_not_found)
    return fail(PHASTAIO_NOTFOUND, "readdatablock: header '%s' was not found", keyphrase);
  of->have_key = false;
  bool binary;
  size_t ts;
  is_binary(iotype, &binary);
  if (int rc = type_size(datatype, &ts)) return rc;
  const int n = *nItems;
  if (binary) {
    char junk;
    if (fread(valueArray, ts, (size_t)n, of->f) != (size_t)n || fread(&junk, 1, 1, of->f) != 1)
      return fail(PHASTAIO_EIO, "readdatablock: short read of '%s'", keyphrase);
    if (of->swap) phastaio_swap_byte_order(valueArray, (int)ts, n);
  } else if (ts == sizeof(int)) {
    int* v = static_cast<int*>(valueArray);
    for (int i = 0; i < n; ++i)
      if (fscanf(of->f, "%d\n", v + i) != 1) return fail(PHASTAIO_EIO, "readdatablock: bad integer");
  } else {
    double* v = static_cast<double*>(valueArray);
    for (int i = 0; i < n; ++i)
      if (fscanf(of->f, "%lf\n", v + i) != 1) return fail(PHASTAIO_EIO, "readdatablock: bad double");
  }
  return 0;
}

int phastaio_writeheader(const int* fd, const char* keyphrase, const void* valueArray,
                         const int* nItems, const int* ndataItems, const char* datatype,
                         const char* iotype) {
  OpenFile* of = lookup(*fd);
  if (!of) return fail(PHASTAIO_EBADFD, "writeheader: no file behind descriptor %d", *fd);
  bool binary;
  size_t ts;
  is_binary(iotype, &binary);
  if (int rc = type_size(datatype, &ts)) return rc;
  of->last_key = keyphrase;
  of->have_key = true;
  of->header_type = ts;
  const int nd = *ndataItems;
  const long size = binary ? (long)ts * nd + (nd > 0 ? 1 : 0) : nd;
  fprintf(of->f, "%s : < %ld > ", keyphrase, size);
  const int* v = static_cast<const int*>(valueArray);
  for (int i = 0; i < *nItems; ++i) fprintf(of->f, "%d ", v[i]);
  fprintf(of->f, "\n");
  return ferror(of->f) ? fail(PHASTAIO_EIO, "writeheader: write error") : 0;
}

int phastaio_writedatablock(const int* fd, const char* keyphrase, const void* valueArray,
                            const int* nItems, const char* datatype, const char* iotype) {
  OpenFile* of = lookup(*fd);
  if (!of) return fail(PHASTAIO_EBADFD, "writedatablock: no file behind descriptor %d", *fd);
  if (!of->have_key || !key_match(of->last_key.c_str(), keyphrase)) {
    if (g_strict)
      return fail(PHASTAIO_EKEY, "writedatablock: header '%s' not consistent with block '%s'",
                  of->last_key.c_str(), keyphrase);
  }
  of->have_key = false;
  bool binary;
  size_t ts;
  is_binary(iotype, &binary);
  if (int rc = type_size(datatype, &ts)) return rc;
  if (of->header_type != ts)
    return fail(PHASTAIO_EINVAL, "writedatablock: header and block differ in type for '%s'",
                keyphrase);
  const int n = *nItems;
  if (binary) {
    if (fwrite(valueArray, ts, (size_t)n, of->f) != (size_t)n)
      return fail(PHASTAIO_EIO, "writedatablock: short write of '%s'", keyphrase);
    fprintf(of->f, "\n");
  } else if (ts == sizeof(int)) {
    const int* v = static_cast<const int*>(valueArray);
    for (int i = 0; i < n; ++i) fprintf(of->f, "%d\n", v[i]);
  } else {
    const double* v = static_cast<const double*>(valueArray);
    for (int i = 0; i < n; ++i) fprintf(of->f, "%f\n", v[i]);
  }
  return ferror(of->f) ? fail(PHASTAIO_EIO, "writedatablock: write error") : 0;
}

int phastaio_writestring(int* fd, const char* text) {
  OpenFile* of = lookup(*fd);
  if (!of) return fail(PHASTAIO_EBADFD, "writestring: no file behind descriptor %d", *fd);
  fprintf(of->f, "%s", text);
  return 0;
}

void phastaio_togglestrictmode(void) { g_strict = !g_strict; }

void phastaio_swap_byte_order(void* array, int nbytes, int nItems) {
  unsigned char* p = static_cast<unsigned char*>(array);
  for (int i = 0; i < nItems; ++i, p += nbytes)
    for (int j = 0; j < nbytes / 2; ++j) std::swap(p[j], p[nbytes - 1 - j]);
}

int phastaio_is_little_endian(void) {
  const unsigned int one = 1;
  return *reinterpret_cast<const unsigned char*>(&one) == 1;
}

}  // extern "C"
