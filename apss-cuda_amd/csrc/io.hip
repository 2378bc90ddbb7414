// Input and output formats (host code; no device calls).
//
//  - the reference's text format (main.cu:286-323): one row per line, a
//    leading separator character, then "d v" pairs up to the end of the line.
//    Like the reference (main.cu:301 reads one character before the first
//    pair), the first character of a line is always taken as the separator,
//    whatever it is.  In addition pairs may be separated by any run of
//    blanks, and "\r\n" line ends are accepted.  Ids need not be sorted in a
//    line (they are sorted here); a repeated id in one line is an error;
//  - a binary CSR file: "APSSCSR1", int64 n_rows, n_cols, nnz, then row_ptr
//    (int64), col (int32) and val (float32), little-endian;
//  - pair files: one "i j score" line per pair (the lines compare() reads,
//    main.cu:143-171), and the set difference of two such files.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>
#include <utility>
#include <vector>

#include "engine.h"

namespace apss {
extern thread_local std::string g_last_error;

namespace {

int fail(int code, const std::string& what) {
  g_last_error = what;
  return code;
}

struct File {
  FILE* f = nullptr;
  File(const char* path, const char* mode) : f(path ? std::fopen(path, mode) : nullptr) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

// Owned CSR arrays handed to the caller through apss_csr (freed by apss_csr_free).
int hand_over(std::vector<int64_t>& rp, std::vector<int32_t>& col, std::vector<float>& val,
              int64_t n_cols, apss_csr* out) {
  const int64_t n = (int64_t)rp.size() - 1, nnz = (int64_t)col.size();
  int64_t* r = (int64_t*)std::malloc(sizeof(int64_t) * rp.size());
  int32_t* c = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(col.size(), 1));
  float* v = (float*)std::malloc(sizeof(float) * std::max<size_t>(val.size(), 1));
  if (!r || !c || !v) {
    std::free(r);
    std::free(c);
    std::free(v);
    return fail(APSS_E_NOMEM, "host allocation of the matrix failed");
  }
  std::memcpy(r, rp.data(), sizeof(int64_t) * rp.size());
  if (nnz) {
    std::memcpy(c, col.data(), sizeof(int32_t) * col.size());
    std::memcpy(v, val.data(), sizeof(float) * val.size());
  }
  *out = apss_csr{n, n_cols, nnz, r, c, v, 0, 0};
  return APSS_OK;
}

// Full structural check of a host CSR: row_ptr from 0 to nnz and
// non-decreasing, column ids in [0, n_cols) and strictly increasing in a row.
int check_structure(const apss_csr& X) {
  try {
    check_csr_host(X, 0);
  } catch (const Error& e) {
    return fail(e.code, e.what);
  }
  for (int64_t r = 0; r < X.n_rows; ++r) {
    const int64_t a = X.row_ptr[r], b = X.row_ptr[r + 1];
    if (b < a || b > X.nnz) return fail(APSS_E_INVALID, "row_ptr not non-decreasing at row " + std::to_string(r));
    for (int64_t k = a; k < b; ++k)
      if (X.col[k] < 0 || X.col[k] >= X.n_cols || (k > a && X.col[k] <= X.col[k - 1]))
        return fail(APSS_E_INVALID, "column ids out of range or not increasing in row " + std::to_string(r));
  }
  return APSS_OK;
}

}  // namespace
}  // namespace apss

using namespace apss;

extern "C" {

int apss_load_text(const char* path, int64_t n_cols, apss_csr* out) {
  if (!path || !out) return fail(APSS_E_INVALID, "null argument");
  std::memset(out, 0, sizeof(*out));
  File in(path, "rb");
  if (!in.f) return fail(APSS_E_INVALID, std::string("cannot open ") + path + ": " + std::strerror(errno));
  std::vector<int64_t> rp(1, 0);
  std::vector<int32_t> col;
  std::vector<float> val;
  std::vector<std::pair<int32_t, float>> row;
  int64_t max_id = -1, line = 0;
  char* buf = nullptr;
  size_t cap = 0;
  ssize_t len;
  auto bad = [&](const std::string& why) {
    std::free(buf);
    return fail(APSS_E_INVALID, "line " + std::to_string(line) + ": " + why);
  };
  auto blank = [](char c) { return c == ' ' || c == '\t' || c == '\r'; };
  while ((len = getline(&buf, &cap, in.f)) > 0) {
    ++line;
    if (buf[len - 1] == '\n') buf[--len] = 0;
    const char* s = buf;
    const char* end = buf + len;
    if (s < end) ++s;  // the leading separator, consumed unread (main.cu:301)
    row.clear();
    for (;;) {
      while (s < end && blank(*s)) ++s;
      if (s == end) break;
      char* e;
      errno = 0;
      const long long d = std::strtoll(s, &e, 10);
      if (e == s || errno || e == end || !blank(*e)) return bad("expected \"id value\"");
      if (d < 0 || d > INT32_MAX - 1) return bad("feature id out of range");
      s = e;
      while (s < end && blank(*s)) ++s;
      const float v = std::strtof(s, &e);
      if (e == s || (e < end && !blank(*e))) return bad("expected a value after id " + std::to_string(d));
      s = e;
      row.push_back({(int32_t)d, v});
    }
    std::sort(row.begin(), row.end(),
              [](const std::pair<int32_t, float>& a, const std::pair<int32_t, float>& b) {
                return a.first < b.first;
              });
    for (size_t k = 0; k < row.size(); ++k) {
      if (k && row[k].first == row[k - 1].first)
        return bad("feature id " + std::to_string(row[k].first) + " repeated");
      col.push_back(row[k].first);
      val.push_back(row[k].second);
      max_id = std::max<int64_t>(max_id, row[k].first);
    }
    rp.push_back((int64_t)col.size());
  }
  std::free(buf);
  if ((int64_t)rp.size() - 1 >= (int64_t)INT32_MAX)
    return fail(APSS_E_UNSUPPORTED, "more than 2^31-1 rows");
  const int64_t d = n_cols > 0 ? n_cols : max_id + 1;
  if (max_id >= d) return fail(APSS_E_INVALID, "feature id beyond n_cols");
  return hand_over(rp, col, val, d, out);
}

int apss_save_text(const char* path, const apss_csr* X) {
  if (!path || !X) return fail(APSS_E_INVALID, "null argument");
  if (X->on_device) return fail(APSS_E_UNSUPPORTED, "save a host matrix");
  const int rc = check_structure(*X);
  if (rc != APSS_OK) return rc;
  File f(path, "w");
  if (!f.f) return fail(APSS_E_INVALID, std::string("cannot create ") + path);
  for (int64_t r = 0; r < X->n_rows; ++r) {
    std::fputc(' ', f.f);
    for (int64_t k = X->row_ptr[r]; k < X->row_ptr[r + 1]; ++k)
      std::fprintf(f.f, k + 1 < X->row_ptr[r + 1] ? "%d %.9g " : "%d %.9g", X->col[k], (double)X->val[k]);
    if (std::fputc('\n', f.f) == EOF) return fail(APSS_E_INVALID, "write failed");
  }
  return APSS_OK;
}

int apss_save_binary(const char* path, const apss_csr* X) {
  if (!path || !X) return fail(APSS_E_INVALID, "null argument");
  if (X->on_device) return fail(APSS_E_UNSUPPORTED, "save a host matrix");
  const int rc = check_structure(*X);
  if (rc != APSS_OK) return rc;
  File f(path, "wb");
  if (!f.f) return fail(APSS_E_INVALID, std::string("cannot create ") + path);
  const int64_t hdr[3] = {X->n_rows, X->n_cols, X->nnz};
  bool ok = std::fwrite("APSSCSR1", 1, 8, f.f) == 8 && std::fwrite(hdr, sizeof(hdr), 1, f.f) == 1 &&
            std::fwrite(X->row_ptr, sizeof(int64_t), (size_t)X->n_rows + 1, f.f) == (size_t)X->n_rows + 1;
  if (ok && X->nnz)
    ok = std::fwrite(X->col, sizeof(int32_t), (size_t)X->nnz, f.f) == (size_t)X->nnz &&
         std::fwrite(X->val, sizeof(float), (size_t)X->nnz, f.f) == (size_t)X->nnz;
  return ok ? APSS_OK : fail(APSS_E_INVALID, "write failed");
}

int apss_load_binary(const char* path, apss_csr* out) {
  if (!path || !out) return fail(APSS_E_INVALID, "null argument");
  std::memset(out, 0, sizeof(*out));
  File f(path, "rb");
  if (!f.f) return fail(APSS_E_INVALID, std::string("cannot open ") + path);
  char magic[8];
  int64_t hdr[3];
  if (std::fread(magic, 1, 8, f.f) != 8 || std::memcmp(magic, "APSSCSR1", 8) != 0 ||
      std::fread(hdr, sizeof(hdr), 1, f.f) != 1)
    return fail(APSS_E_INVALID, "not an APSS binary CSR file");
  const int64_t n = hdr[0], d = hdr[1], nnz = hdr[2];
  if (n < 0 || d < 0 || nnz < 0 || n >= INT32_MAX || d >= INT32_MAX || nnz > (INT64_MAX >> 4))
    return fail(APSS_E_INVALID, "bad header");
  // The file must hold exactly the arrays the header announces.
  const long here = std::ftell(f.f);
  if (here < 0 || std::fseek(f.f, 0, SEEK_END) != 0) return fail(APSS_E_INVALID, "file not seekable");
  const int64_t size = std::ftell(f.f);
  if (size != here + 8 * (n + 1) + 8 * nnz) return fail(APSS_E_INVALID, "file size does not match the header");
  std::fseek(f.f, here, SEEK_SET);
  int64_t* rp = (int64_t*)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
  int32_t* col = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1));
  float* val = (float*)std::malloc(sizeof(float) * (size_t)std::max<int64_t>(nnz, 1));
  *out = apss_csr{n, d, nnz, rp, col, val, 0, 0};
  if (!rp || !col || !val) {
    apss_csr_free(out);
    return fail(APSS_E_NOMEM, "host allocation of the matrix failed");
  }
  if (std::fread(rp, sizeof(int64_t), (size_t)n + 1, f.f) != (size_t)n + 1 ||
      (nnz && (std::fread(col, sizeof(int32_t), (size_t)nnz, f.f) != (size_t)nnz ||
               std::fread(val, sizeof(float), (size_t)nnz, f.f) != (size_t)nnz))) {
    apss_csr_free(out);
    return fail(APSS_E_INVALID, "read failed");
  }
  const int rc = check_structure(*out);
  if (rc != APSS_OK) apss_csr_free(out);
  return rc;
}

void apss_csr_free(apss_csr* X) {
  if (!X || X->on_device) return;
  std::free((void*)X->row_ptr);
  std::free((void*)X->col);
  std::free((void*)X->val);
  X->row_ptr = nullptr;
  X->col = nullptr;
  X->val = nullptr;
  X->n_rows = X->nnz = 0;
}

int apss_write_pairs(const char* path, const apss_pair* pairs, int64_t count) {
  if (!path || (count > 0 && !pairs) || count < 0) return fail(APSS_E_INVALID, "bad argument");
  File f(path, "w");
  if (!f.f) return fail(APSS_E_INVALID, std::string("cannot create ") + path);
  for (int64_t k = 0; k < count; ++k)
    if (std::fprintf(f.f, "%u %u %.9g\n", pairs[k].i, pairs[k].j, (double)pairs[k].score) < 0)
      return fail(APSS_E_INVALID, "write failed");
  return APSS_OK;
}

int64_t apss_compare_pair_files(const char* a, const char* b, int64_t* only_a, int64_t* only_b) {
  if (!a || !b) return fail(APSS_E_INVALID, "null argument");
  auto read = [](const char* path, std::unordered_set<uint64_t>& s) -> int {
    File f(path, "r");
    if (!f.f) return fail(APSS_E_INVALID, std::string("cannot open ") + path);
    char* buf = nullptr;
    size_t cap = 0;
    while (getline(&buf, &cap, f.f) > 0) {
      unsigned long i, j;
      if (std::sscanf(buf, "%lu %lu", &i, &j) != 2) continue;  // blank or malformed line
      const uint64_t x = std::min(i, j), y = std::max(i, j);
      s.insert(x << 32 | y);
    }
    std::free(buf);
    return APSS_OK;
  };
  std::unordered_set<uint64_t> sa, sb;
  int rc = read(a, sa);
  if (rc != APSS_OK) return rc;
  rc = read(b, sb);
  if (rc != APSS_OK) return rc;
  int64_t oa = 0, ob = 0;
  for (uint64_t x : sa) oa += !sb.count(x);
  for (uint64_t x : sb) ob += !sa.count(x);
  if (only_a) *only_a = oa;
  if (only_b) *only_b = ob;
  return oa + ob;
}

}  // extern "C"
