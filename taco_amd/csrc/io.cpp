// Tensor files: MatrixMarket (.mtx, .ttx; reference src/storage/
// file_io_mtx.cpp) and FROSTT (.tns; src/storage/file_io_tns.cpp), read into
// coordinate lists that the caller packs (taco's read() inserts every
// coordinate and packs, src/storage/file_io_mtx.cpp:84-160), and written
// back from coordinate lists.
//
// Reading follows taco's rules:
//   MatrixMarket: header "%%MatrixMarket matrix|tensor coordinate|array real
//     general|symmetric" (any other field or symmetry is an error, as in
//     taco); comment lines (first non-blank character '%') and blank lines
//     before the size line are skipped; coordinate files give dimensions then
//     the entry count, entries are 1-based coordinates and a value (lines
//     beyond the announced count are ignored, as taco does); array
//     files give dimensions, then the values with the first mode varying
//     fastest (values beyond the tensor's size are ignored, as taco does);
//     symmetric matrices also get the mirrored entry (i != j).
//   FROSTT: every line is 1-based coordinates then a value; the order is the
//     first line's token count minus one; each dimension is the largest
//     coordinate seen in that mode.
// Writing: coordinate MatrixMarket "real general" (or "array" for dense
// tensors, first mode fastest, as the reader expects) and FROSTT lines.
// Values are written with 17 significant digits so that a file read back
// gives the same doubles (taco writes the stream default of 6 digits).

#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"

struct taco_hip_coo {
  int32_t order = 0;
  int32_t dense = 0;  // 1: read from an array (dense) MatrixMarket file
  std::vector<int32_t> dims;
  std::vector<int32_t> coords;  // nnz * order, 0-based
  std::vector<double> vals;
};

namespace {

using taco_hip::fail;

bool read_line(FILE* f, std::string& line) {
  line.clear();
  int c;
  bool any = false;
  while ((c = std::fgetc(f)) != EOF) {
    any = true;
    if (c == '\n') break;
    line.push_back((char)c);
  }
  if (!line.empty() && line.back() == '\r') line.pop_back();
  return any;
}

bool blank_or_comment(const std::string& line) {
  for (char ch : line) {
    if (std::isspace((unsigned char)ch)) continue;
    return ch == '%';
  }
  return true;
}

std::vector<std::string> tokens(const std::string& line) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && std::isspace((unsigned char)line[i])) ++i;
    size_t j = i;
    while (j < line.size() && !std::isspace((unsigned char)line[j])) ++j;
    if (j > i) out.push_back(line.substr(i, j - i));
    i = j;
  }
  return out;
}

// Parses a 1-based coordinate into [0, INT_MAX).
bool parse_index(const char*& p, long* out) {
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(p, &end, 10);
  if (end == p || errno) return false;
  p = end;
  *out = v;
  return true;
}

int read_mtx(FILE* f, const char* path, taco_hip_coo* c) {
  std::string line;
  if (!read_line(f, line)) return fail("%s: empty file", path);
  const std::vector<std::string> head = tokens(line);
  if (head.size() < 5 || head[0] != "%%MatrixMarket")
    return fail("%s: unknown header of MatrixMarket", path);
  if (head[1] != "matrix" && head[1] != "tensor")
    return fail("%s: unknown type of MatrixMarket: %s", path, head[1].c_str());
  if (head[3] != "real") return fail("%s: MatrixMarket field not available: %s", path, head[3].c_str());
  if (head[4] != "general" && head[4] != "symmetric")
    return fail("%s: MatrixMarket symmetry not available: %s", path, head[4].c_str());
  const bool symm = head[4] == "symmetric";
  const bool coordinate = head[2] == "coordinate";
  if (!coordinate && head[2] != "array")
    return fail("%s: MatrixMarket format not available: %s", path, head[2].c_str());
  bool have = false;
  while (read_line(f, line))
    if (!blank_or_comment(line)) {
      have = true;
      break;
    }
  if (!have) return fail("%s: no size line", path);
  std::vector<long> sizes;
  for (const std::string& t : tokens(line)) {
    char* end = nullptr;
    errno = 0;
    const unsigned long long v = std::strtoull(t.c_str(), &end, 10);
    if (*end || errno) return fail("%s: bad size line", path);
    if (v == 0) break;  // taco stops at the first zero
    sizes.push_back((long)(v > (unsigned long long)LONG_MAX ? LONG_MAX : v));
  }
  int64_t nnz = 0;
  if (coordinate) {
    if (sizes.size() < 2) return fail("%s: bad size line", path);
    nnz = sizes.back();
    sizes.pop_back();
  }
  if (sizes.empty()) return fail("%s: no dimensions", path);
  for (long d : sizes)
    if (d > INT_MAX) return fail("%s: dimension exceeds INT_MAX", path);
  if (symm && sizes.size() != 2) return fail("%s: symmetry only available for matrix", path);
  c->order = (int32_t)sizes.size();
  c->dims.assign(sizes.begin(), sizes.end());
  const int32_t order = c->order;
  if (coordinate) {
    c->coords.reserve((size_t)nnz * order * (symm ? 2 : 1));
    c->vals.reserve((size_t)nnz * (symm ? 2 : 1));
    int64_t seen = 0;
    std::vector<int32_t> co((size_t)order);
    // taco inserts the first nnz entries and ignores lines beyond them
    while (seen < nnz && read_line(f, line)) {
      if (blank_or_comment(line)) continue;
      const char* p = line.c_str();
      for (int32_t k = 0; k < order; ++k) {
        long v = 0;
        if (!parse_index(p, &v) || v < 1 || v > c->dims[(size_t)k])
          return fail("%s: entry %lld has a coordinate out of range", path, (long long)seen + 1);
        co[(size_t)k] = (int32_t)(v - 1);
      }
      char* end = nullptr;
      const double val = std::strtod(p, &end);
      if (end == p) return fail("%s: entry %lld has no value", path, (long long)seen + 1);
      c->coords.insert(c->coords.end(), co.begin(), co.end());
      c->vals.push_back(val);
      if (symm && co.front() != co.back()) {
        c->coords.push_back(co[1]);
        c->coords.push_back(co[0]);
        c->vals.push_back(val);
      }
      ++seen;
    }
    if (seen < nnz)
      return fail("%s: the size line announces %lld entries, the file holds %lld", path,
                  (long long)nnz, (long long)seen);
    return 0;
  }
  // array: every value, first mode fastest
  int64_t size = 1;
  for (int32_t d : c->dims) {
    if (size > INT64_MAX / d) return fail("%s: too many values", path);
    size *= d;
  }
  c->dense = 1;
  c->vals.reserve((size_t)size);
  while (read_line(f, line)) {
    if (blank_or_comment(line)) continue;
    const char* p = line.c_str();
    char* end = nullptr;
    const double val = std::strtod(p, &end);
    if (end == p) return fail("%s: bad value line", path);
    c->vals.push_back(val);
  }
  // taco inserts the first `size` values and ignores any beyond them
  if ((int64_t)c->vals.size() < size)
    return fail("%s: %lld values for %lld positions", path, (long long)c->vals.size(),
                (long long)size);
  c->vals.resize((size_t)size);
  c->coords.resize((size_t)size * order);
  std::vector<double> vals = c->vals;
  for (int64_t n = 0; n < size; ++n) {
    int64_t idx = n;
    for (int32_t k = 0; k < order; ++k) {
      c->coords[(size_t)n * order + k] = (int32_t)(idx % c->dims[(size_t)k]);
      idx /= c->dims[(size_t)k];
    }
  }
  if (symm) {
    for (int64_t n = 0; n < size; ++n) {
      const int32_t i = c->coords[(size_t)n * 2], j = c->coords[(size_t)n * 2 + 1];
      if (i != j) {
        c->coords.push_back(j);
        c->coords.push_back(i);
        c->vals.push_back(vals[(size_t)n]);
      }
    }
  }
  return 0;
}

int read_tns(FILE* f, const char* path, taco_hip_coo* c) {
  std::string line;
  int32_t order = -1;
  int64_t seen = 0;
  std::vector<int32_t> co;
  while (read_line(f, line)) {
    if (blank_or_comment(line)) continue;
    if (order < 0) {
      const size_t n = tokens(line).size();
      if (n < 2) return fail("%s: first line has no coordinates", path);
      order = (int32_t)(n - 1);
      c->order = order;
      c->dims.assign((size_t)order, 0);
      co.resize((size_t)order);
    }
    const char* p = line.c_str();
    for (int32_t k = 0; k < order; ++k) {
      long v = 0;
      if (!parse_index(p, &v) || v < 1 || v > INT_MAX)
        return fail("%s: line %lld has a coordinate out of range", path, (long long)seen + 1);
      co[(size_t)k] = (int32_t)(v - 1);
      if (v > c->dims[(size_t)k]) c->dims[(size_t)k] = (int32_t)v;
    }
    char* end = nullptr;
    const double val = std::strtod(p, &end);
    if (end == p) return fail("%s: line %lld has no value", path, (long long)seen + 1);
    c->coords.insert(c->coords.end(), co.begin(), co.end());
    c->vals.push_back(val);
    ++seen;
  }
  if (order < 0) return fail("%s: empty file", path);
  return 0;
}

bool ends_with(const std::string& s, const char* suffix) {
  const size_t n = std::strlen(suffix);
  return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

}  // namespace

extern "C" {

int taco_hip_read_tensor_file(const char* path, taco_hip_coo** out) {
  TH_REQUIRE(path && out, "read: null argument");
  *out = nullptr;
  const std::string p(path);
  const bool mtx = ends_with(p, ".mtx") || ends_with(p, ".ttx");
  const bool tns = ends_with(p, ".tns");
  TH_REQUIRE(mtx || tns, "%s: file format not supported (.mtx, .ttx, .tns)", path);
  FILE* f = std::fopen(path, "rb");
  TH_REQUIRE(f, "%s: cannot open: %s", path, std::strerror(errno));
  auto* c = new taco_hip_coo();
  const int rc = mtx ? read_mtx(f, path, c) : read_tns(f, path, c);
  std::fclose(f);
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

int taco_hip_coo_info(const taco_hip_coo* c, int32_t* order, int32_t* dims, int32_t max_dims,
                      int64_t* nnz, int32_t* dense) {
  TH_REQUIRE(c && order && nnz && dense, "coo: null argument");
  TH_REQUIRE(dims || max_dims == 0, "coo: null dims");
  *order = c->order;
  *nnz = (int64_t)c->vals.size();
  *dense = c->dense;
  for (int32_t k = 0; k < c->order && k < max_dims; ++k) dims[k] = c->dims[(size_t)k];
  return 0;
}

int taco_hip_coo_data(const taco_hip_coo* c, int32_t* coords, double* vals) {
  TH_REQUIRE(c && (coords || c->coords.empty()) && (vals || c->vals.empty()),
             "coo: null argument");
  if (!c->coords.empty()) std::memcpy(coords, c->coords.data(), sizeof(int32_t) * c->coords.size());
  if (!c->vals.empty()) std::memcpy(vals, c->vals.data(), sizeof(double) * c->vals.size());
  return 0;
}

void taco_hip_coo_destroy(taco_hip_coo* c) { delete c; }

int taco_hip_write_tensor_file(const char* path, int32_t order, const int32_t* dims,
                               int64_t nnz, const int32_t* coords, const double* vals,
                               int32_t dense) {
  TH_REQUIRE(path && order > 0 && dims && nnz >= 0 && (vals || nnz == 0),
             "write: bad arguments");
  const std::string p(path);
  const bool mtx = ends_with(p, ".mtx") || ends_with(p, ".ttx");
  const bool tns = ends_with(p, ".tns");
  TH_REQUIRE(mtx || tns, "%s: file format not supported (.mtx, .ttx, .tns)", path);
  TH_REQUIRE(!(dense && tns), "%s: FROSTT files are coordinate lists", path);
  TH_REQUIRE(dense || coords || nnz == 0, "write: coordinates missing");
  FILE* f = std::fopen(path, "wb");
  TH_REQUIRE(f, "%s: cannot open for writing: %s", path, std::strerror(errno));
  const char* kind = order == 2 ? "matrix" : "tensor";
  if (mtx) {
    std::fprintf(f, "%%%%MatrixMarket %s %s real general\n%%\n", kind,
                 dense ? "array" : "coordinate");
    for (int32_t k = 0; k < order; ++k) std::fprintf(f, "%s%d", k ? " " : "", dims[k]);
    if (!dense) std::fprintf(f, " %lld", (long long)nnz);
    std::fprintf(f, "\n");
  }
  for (int64_t n = 0; n < nnz; ++n) {
    if (!dense)
      for (int32_t k = 0; k < order; ++k)
        std::fprintf(f, "%d ", coords[(size_t)n * order + k] + 1);
    std::fprintf(f, "%.17g\n", vals[n]);
  }
  const bool ok = std::fflush(f) == 0 && !std::ferror(f);
  std::fclose(f);
  TH_REQUIRE(ok, "%s: write failed", path);
  return 0;
}

}  // extern "C"
