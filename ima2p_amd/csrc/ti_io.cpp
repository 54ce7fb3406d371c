// ti_io.cpp -- the genealogy-table file (.ti) that IMa2p writes in M mode and
// loads in L mode, host side, no device needed.
//
// Writer (ima_main_mpi.cpp:2121-2142 header, output.cpp:662-685 rows): a
// header block of dashed lines around the caller's text, a line
// "VALUESSTART", then one line per genealogy of K values, each printed as
// "%.6f\t" from a float, ending with "\n".
//
// Loader (loadgenealogyvalues, ima_main_mpi.cpp:3216-3440): lines up to
// the one containing VALUESSTART are header; every later line is one
// genealogy whose K whitespace-separated values are read with "%f" into
// floats.  A line with fewer or more than K values is an error.  When more
// genealogies are present than the caller wants, a genealogy is kept while
// loaded / (loaded + skipped) <= wanted / total (the first always), in
// file order across the files.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ima_gpu.h"

namespace {

// One line of any length (no newline), false at end of file.
bool read_line(FILE *f, std::string &s) {
  s.clear();
  int ch;
  bool any = false;
  while ((ch = fgetc(f)) != EOF) {
    any = true;
    if (ch == '\n') return true;
    s.push_back((char)ch);
  }
  return any;
}

// Skips the header; returns false if the file has no VALUESSTART line.
bool skip_header(FILE *f) {
  std::string s;
  while (read_line(f, s))
    if (s.find("VALUESSTART") != std::string::npos) return true;
  return false;
}

// Parses exactly K floats from a line: 0 ok, 1 too few, 2 too many.
int parse_row(const std::string &s, int K, float *out) {
  const char *c = s.c_str();
  for (int i = 0; i < K; ++i) {
    while (*c == ' ' || *c == '\t' || *c == '\r') ++c;
    if (!*c) return 1;
    char *end = nullptr;
    errno = 0;
    const float v = strtof(c, &end);
    if (end == c) return 1;
    out[i] = v;
    c = end;
  }
  while (*c == ' ' || *c == '\t' || *c == '\r') ++c;
  return *c ? 2 : 0;
}

}  // namespace

extern "C" {

int ima_ti_write(const char *path, const char *header, const float *rows, int G, int K,
                 int append) {
  if (!path || (!rows && G > 0) || G < 0 || K < 1) return IMA_ERR_ARG;
  FILE *f = fopen(path, append ? "a" : "w");
  if (!f) return IMA_ERR_ARG;
  if (!append) {
    const char *dash = "-------------------------------------------\n\n";
    fprintf(f, "%s", dash);
    fprintf(f, "Header for genealogy file:  %s\n\n", path);
    fprintf(f, "%s", dash);
    fprintf(f, "%s\n", header ? header : "");
    fprintf(f, "%s", dash);
    fprintf(f, "End of header for genealogy file:  %s\n\n", path);
    fprintf(f, "%s", dash);
    fprintf(f, "VALUESSTART\n");
  }
  for (int g = 0; g < G; ++g) {
    for (int i = 0; i < K; ++i) fprintf(f, "%.6f\t", (double)rows[(size_t)g * K + i]);
    fprintf(f, "\n");
  }
  const int bad = ferror(f);
  if (fclose(f) != 0 || bad) return IMA_ERR_ARG;
  return IMA_OK;
}

int ima_ti_count(const char *path, int *G) {
  if (!path || !G) return IMA_ERR_ARG;
  FILE *f = fopen(path, "r");
  if (!f) return IMA_ERR_ARG;
  if (!skip_header(f)) {
    fclose(f);
    return IMA_ERR_ARG;
  }
  int n = 0, ch;
  while ((ch = fgetc(f)) != EOF) n += ch == '\n';
  fclose(f);
  *G = n;
  return IMA_OK;
}

int ima_ti_load(const char *const *paths, int npaths, int K, int max_rows, float *rows,
                int *loaded_out, int *total_out) {
  if (!paths || npaths < 1 || K < 1 || max_rows < 1 || !rows || !loaded_out) return IMA_ERR_ARG;
  long total = 0;
  for (int p = 0; p < npaths; ++p) {
    int n = 0;
    const int rc = ima_ti_count(paths[p], &n);
    if (rc != IMA_OK) return rc;
    total += n;
  }
  if (total_out) *total_out = (int)total;
  if (total < 1) return IMA_ERR_ARG;
  const long want = total < max_rows ? total : max_rows;
  const bool all = want >= total;
  const float frac = (float)want / (float)total;
  long loaded = 0, skipped = 0;
  std::string s;
  std::vector<float> tmp(K);
  for (int p = 0; p < npaths; ++p) {
    FILE *f = fopen(paths[p], "r");
    if (!f) return IMA_ERR_ARG;
    if (!skip_header(f)) {
      fclose(f);
      return IMA_ERR_ARG;
    }
    while (read_line(f, s)) {
      if (all || loaded == 0 || (float)loaded / (float)(loaded + skipped) <= frac) {
        if (loaded >= want) break;
        if (parse_row(s, K, tmp.data()) != 0) {
          fclose(f);
          return IMA_ERR_ARG;
        }
        std::memcpy(rows + (size_t)loaded * K, tmp.data(), (size_t)K * sizeof(float));
        ++loaded;
      } else {
        ++skipped;
      }
    }
    fclose(f);
  }
  *loaded_out = (int)loaded;
  return IMA_OK;
}

}  // extern "C"
