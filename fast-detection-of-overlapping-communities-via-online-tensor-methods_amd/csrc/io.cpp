// On-disk formats of the reference: COO partition files in, one-value-per-line and triplet files out.
//
// read: the GPU reader read_G (pre_proc.c:49-80) parses numbers with %lf, 1-based, and stores the
// entry; BINARY sets it to 1, EXPECTED reads a third value (last value wins on repeats). The CPU
// reader read_G_sparse (IOfn.cpp:21-54) takes the first number as the row and sums repeated
// triplets (setFromTriplets) then prunes |v| <= 1e-18. Both `while(!feof)` loops apply the last
// entry once more when the file ends with whitespace (pre_proc.c:60-78, IOfn.cpp:36-49).
// The result here is a CSR block (columns ascending, duplicates merged by the selected policy).
// write: "%5.25e\n" per value (write_mat, post_proc.c:45-59; stpm.c:926-951), "row\tcol\tvalue"
// triplets with 6 significant digits (write_pi, IOfn.cpp:57-67), sim_info lines (stpm.c:957-975).
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "host_util.h"
#include "mmsb_stpm.h"

namespace mmsb {

static bool read_file(const char* path, std::string& buf) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize(sz > 0 ? (size_t)sz : 0);
  const size_t got = sz > 0 ? std::fread(&buf[0], 1, (size_t)sz, f) : 0;
  std::fclose(f);
  return got == buf.size();
}

}  // namespace mmsb

using namespace mmsb;

extern "C" int mmsb_read_coo(const char* path, int32_t n_rows, int32_t n_cols, int32_t first_is_row, int32_t weighted,
                             int32_t dup_policy, int32_t feof_repeat, mmsb_csr* out) {
  clear_error();
  MMSB_CHECK_ARG(path && out && n_rows >= 0 && n_cols >= 0, "read_coo: bad arguments");
  std::memset(out, 0, sizeof(*out));
  std::string buf;
  if (!read_file(path, buf)) return set_error(MMSB_ERR_IO, "reading %s failed", path);
  const int w = weighted ? 3 : 2;
  std::vector<double> nums;
  nums.reserve(buf.size() / 4);
  const char* p = buf.c_str();
  char* end = nullptr;
  for (;;) {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f') ++p;
    if (!*p) break;
    const double v = std::strtod(p, &end);
    if (end == p) return set_error(MMSB_ERR_IO, "%s: unparsable text at byte %ld", path, (long)(p - buf.c_str()));
    nums.push_back(v);
    p = end;
  }
  int64_t ne = (int64_t)nums.size() / w;
  if (feof_repeat && ne > 0 && !buf.empty()) {
    const char last = buf.back();
    if (last == '\n' || last == ' ' || last == '\t' || last == '\r') {
      for (int c = 0; c < w; ++c) nums.push_back(nums[(ne - 1) * w + c]);
      ++ne;
    }
  }
  struct Ent {
    int64_t key;
    double v;
    int64_t order;
  };
  std::vector<Ent> ents;
  ents.reserve(ne);
  for (int64_t e = 0; e < ne; ++e) {
    const int64_t a = (int64_t)nums[e * w] - 1, b = (int64_t)nums[e * w + 1] - 1;
    const int64_t row = first_is_row ? a : b, col = first_is_row ? b : a;
    if (row < 0 || row >= n_rows || col < 0 || col >= n_cols)
      return set_error(MMSB_ERR_IO, "%s: entry %lld (%lld, %lld) outside %d x %d", path, (long long)e + 1,
                       (long long)row + 1, (long long)col + 1, n_rows, n_cols);
    ents.push_back({row * (int64_t)n_cols + col, weighted ? nums[e * w + 2] : 1.0, e});
  }
  std::sort(ents.begin(), ents.end(), [](const Ent& x, const Ent& y) { return x.key < y.key || (x.key == y.key && x.order < y.order); });
  // Values are kept for weighted files (unless every entry is forced to 1) and for summed two-column
  // files, where read_G_sparse adds 1.0 per line (IOfn.cpp:45-46) so that a repeated line -- the
  // while(!feof) repeat of the last one included -- counts twice.
  const bool with_vals = dup_policy != MMSB_DUP_ONE && (weighted || dup_policy == MMSB_DUP_SUM);
  std::vector<int64_t> keys;
  std::vector<double> vals;
  for (size_t i = 0; i < ents.size();) {
    size_t j = i;
    double v = 0.0;
    while (j < ents.size() && ents[j].key == ents[i].key) {
      if (dup_policy == MMSB_DUP_ONE) v = 1.0;
      else if (dup_policy == MMSB_DUP_LAST) v = ents[j].v;
      else v += ents[j].v;
      ++j;
    }
    if (!with_vals || std::fabs(v) > 1e-18) {
      keys.push_back(ents[i].key);
      vals.push_back(v);
    }
    i = j;
  }
  const int64_t nnz = (int64_t)keys.size();
  if (nnz >= INT32_MAX) return set_error(MMSB_ERR_ARG, "%s: too many nonzeros for int32 CSR", path);
  out->n_rows = n_rows;
  out->n_cols = n_cols;
  out->nnz = nnz;
  out->rowptr = (int32_t*)std::calloc((size_t)n_rows + 1, sizeof(int32_t));
  out->colidx = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1));
  out->vals = with_vals ? (float*)std::malloc(sizeof(float) * (size_t)std::max<int64_t>(nnz, 1)) : nullptr;
  if (!out->rowptr || !out->colidx || (with_vals && !out->vals))
    return set_error(MMSB_ERR_NOMEM, "read_coo: out of host memory");
  for (int64_t e = 0; e < nnz; ++e) {
    out->rowptr[keys[e] / n_cols + 1]++;
    out->colidx[e] = (int32_t)(keys[e] % n_cols);
    if (out->vals) out->vals[e] = (float)vals[e];
  }
  for (int32_t i = 0; i < n_rows; ++i) out->rowptr[i + 1] += out->rowptr[i];
  return MMSB_OK;
}

extern "C" void mmsb_csr_free(mmsb_csr* m) {
  if (!m) return;
  std::free(m->rowptr);
  std::free(m->colidx);
  std::free(m->vals);
  std::memset(m, 0, sizeof(*m));
}

extern "C" int mmsb_write_values(const char* path, const double* v, int64_t n) {
  clear_error();
  FILE* f = std::fopen(path, "w");
  if (!f) return set_error(MMSB_ERR_IO, "writing %s failed", path);
  for (int64_t i = 0; i < n; ++i) std::fprintf(f, "%5.25e\n", v[i]);
  return std::fclose(f) == 0 ? MMSB_OK : set_error(MMSB_ERR_IO, "closing %s failed", path);
}

extern "C" int mmsb_write_coo(const char* path, const mmsb_csr* m) {
  clear_error();
  FILE* f = std::fopen(path, "w");
  if (!f) return set_error(MMSB_ERR_IO, "writing %s failed", path);
  for (int32_t i = 0; i < m->n_rows; ++i)
    for (int32_t e = m->rowptr[i]; e < m->rowptr[i + 1]; ++e) {
      if (m->vals)
        std::fprintf(f, "%d\t%d\t%.9g\n", i + 1, m->colidx[e] + 1, (double)m->vals[e]);
      else
        std::fprintf(f, "%d\t%d\n", i + 1, m->colidx[e] + 1);
    }
  return std::fclose(f) == 0 ? MMSB_OK : set_error(MMSB_ERR_IO, "closing %s failed", path);
}

// k x n matrix given node-major ([n][k], i.e. the k x n column-major buffer of the reference):
// nonzeros in column-major order as "row+1 \t col+1 \t value" with 6 significant digits.
extern "C" int mmsb_write_pi_triplets(const char* path, const double* pi_nk, int64_t n, int32_t k) {
  clear_error();
  FILE* f = std::fopen(path, "w");
  if (!f) return set_error(MMSB_ERR_IO, "writing %s failed", path);
  for (int64_t col = 0; col < n; ++col)
    for (int32_t row = 0; row < k; ++row) {
      const double v = pi_nk[col * k + row];
      if (v != 0.0) std::fprintf(f, "%d\t%lld\t%g\n", row + 1, (long long)col + 1, v);
    }
  return std::fclose(f) == 0 ? MMSB_OK : set_error(MMSB_ERR_IO, "closing %s failed", path);
}
