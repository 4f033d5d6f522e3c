// ffhip_io.hip — matrix exchange with FreeFEM scripts and files:
//   [I, J, C] = A            (lgmat.cpp:2670-2710)  -> ffhip_mat_coo
//   ofstream f(..); f << A;  (HashMatrix.hpp:485-521, the Morse text of a
//                             CSR-state HashMatrix)  -> ffhip_csr_write_morse /
//                                                       ffhip_mat_write_morse
//   ifstream f(..); f >> A;  (HashMatrix(istream), HashMatrix.cpp:7-30,
//                             136-214: Morse or COO text) -> ffhip_csr_read_text /
//                                                           ffhip_mat_read_text
// The triplet export runs on the device (one lane per entry, coalesced
// writes); the text formats are host byte work (formatted by several threads,
// written in order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "ffhip_internal.hpp"

namespace ffhip {
namespace {

unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// I[k] = row of entry k (binary search of rowptr), J[k] = colidx[k]
__global__ void k_coo_rows(int64_t n, int64_t nnz, const int64_t *__restrict__ rowptr,
                           const int32_t *__restrict__ colidx, int64_t *__restrict__ I,
                           int64_t *__restrict__ J) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  int64_t lo = 0, hi = n;  // last row with rowptr[row] <= k
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rowptr[mid] <= k) lo = mid; else hi = mid;
  }
  I[k] = lo;
  J[k] = colidx[k];
}

// one data line of the Morse text: setw(9) i+1, ' ', setw(9) j+1, ' ',
// setprecision(20) removeeps(a) (|a| < 1e-305 printed as 0), '\n'
inline int morse_line(char *buf, size_t cap, int64_t i, int64_t j, double a) {
  if (std::fabs(a) < 1e-305) a = 0.;
  return snprintf(buf, cap, "%9" PRId64 " %9" PRId64 " %.20g\n", i + 1, j + 1, a);
}

int threads_for(int64_t work) {
  unsigned hc = std::thread::hardware_concurrency();
  int t = (int)std::min<unsigned>(hc ? hc : 1, 16);
  if (work < 100000) t = 1;
  return t;
}

void write_morse(FILE *f, int64_t n, int64_t m, int half, const int64_t *rowptr,
                 const int32_t *colidx, const double *vals, const void *tag) {
  const int64_t nnz = rowptr[n] - rowptr[0];
  fprintf(f, "# Sparse Matrix (Morse)  %p\n", tag);
  fprintf(f, "# first line: n m (is symmetic) nnz \n");
  fprintf(f, "# after for each nonzero coefficient:   i j a_ij where (i,j) \\in  {1,...,n}x{1,...,m} \n");
  fprintf(f, "%" PRId64 " %" PRId64 " %d  %" PRId64 "\n", n, m, half, nnz);
  // rows in blocks; each block formatted by T threads into their own strings
  const int T = threads_for(nnz);
  const int64_t block = 1 << 20;
  std::vector<std::string> out(T);
  for (int64_t r0 = 0; r0 < n; r0 += block) {
    const int64_t r1 = std::min(n, r0 + block);
    auto work = [&](int t) {
      const int64_t a = r0 + (r1 - r0) * t / T, b = r0 + (r1 - r0) * (t + 1) / T;
      std::string &s = out[t];
      s.clear();
      s.reserve((size_t)(rowptr[b] - rowptr[a]) * 48);
      char line[96];
      for (int64_t i = a; i < b; ++i)
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
          const int len = morse_line(line, sizeof line, i, colidx[k], vals[k]);
          s.append(line, (size_t)len);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    for (int t = 0; t < T; ++t) fwrite(out[t].data(), 1, out[t].size(), f);
  }
}

// ---- text reader
struct Text {
  std::vector<char> buf;
  size_t p = 0;
  bool eof() const { return p >= buf.size(); }
  void skip_space() {
    while (!eof() && isspace((unsigned char)buf[p])) ++p;
  }
  std::string line() {  // up to '\n' or '\r', consumed
    std::string s;
    while (!eof() && buf[p] != '\n' && buf[p] != '\r') s += buf[p++];
    if (!eof()) ++p;
    if (!eof() && (buf[p] == '\n' || buf[p] == '\r')) ++p;
    return s;
  }
  bool i64(int64_t &v) {
    skip_space();
    if (eof()) return false;
    char *e;
    v = strtoll(buf.data() + p, &e, 10);
    if (e == buf.data() + p) return false;
    p = (size_t)(e - buf.data());
    return true;
  }
  bool f64(double &v) {
    skip_space();
    if (eof()) return false;
    char *e;
    v = strtod(buf.data() + p, &e);
    if (e == buf.data() + p) return false;
    p = (size_t)(e - buf.data());
    return true;
  }
};

struct Parsed {
  int64_t n = 0, m = 0;
  int half = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> colidx;
  std::vector<double> vals;
};

// HashMatrix(istream, -1): the first comment line names the format, every
// '#' line is skipped, entries are assigned (a repeated (i,j) keeps its last
// value), then the rows are sorted as HashMatrix::CSR sorts them
void parse_text(const char *path, Parsed &P) {
  FILE *f = fopen(path, "rb");
  FFHIP_REQUIRE(f, FFHIP_ERR_ARG, "matrix file %s: cannot open", path);
  Text t;
  fseek(f, 0, SEEK_END);
  const long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  t.buf.resize(sz > 0 ? (size_t)sz : 0);
  const size_t got = sz > 0 ? fread(t.buf.data(), 1, (size_t)sz, f) : 0;
  fclose(f);
  FFHIP_REQUIRE(got == t.buf.size(), FFHIP_ERR_ARG, "matrix file %s: read error", path);
  t.buf.push_back('\0');  // strtod / strtoll stop here
  t.skip_space();
  int cas = 0;
  if (!t.eof() && t.buf[t.p] == '#') {
    const std::string first = t.line();
    if (first.find("(Morse)") != std::string::npos) cas = 2;
    else if (first.find("(COO)") != std::string::npos) cas = 3;
  }
  FFHIP_REQUIRE(cas == 2 || cas == 3, FFHIP_ERR_ARG, "matrix file %s: unknown matrix type", path);
  t.skip_space();
  while (!t.eof() && t.buf[t.p] == '#') {
    t.line();
    t.skip_space();
  }
  int64_t n = 0, m = 0, half = 0, cnt = 0, f77 = 1;
  bool ok;
  if (cas == 2) {
    ok = t.i64(n) && t.i64(m) && t.i64(half) && t.i64(cnt);
    FFHIP_REQUIRE(ok && n >= 0 && m >= 0 && cnt >= 0, FFHIP_ERR_ARG, "matrix file %s: bad header", path);
  } else {
    int64_t st, ts;
    ok = t.i64(n) && t.i64(m) && t.i64(cnt) && t.i64(half) && t.i64(f77) && t.i64(st) && t.i64(ts);
    FFHIP_REQUIRE(ok && n > 0 && m > 0 && cnt > 0, FFHIP_ERR_ARG, "matrix file %s: bad header", path);
  }
  FFHIP_REQUIRE(n < INT32_MAX && m < INT32_MAX, FFHIP_ERR_UNSUPPORTED, "matrix file %s: too large", path);
  std::vector<int64_t> key((size_t)cnt);
  std::vector<double> a((size_t)cnt);
  for (int64_t k = 0; k < cnt; ++k) {
    int64_t i, j;
    double v;
    ok = t.i64(i) && t.i64(j) && t.f64(v);
    FFHIP_REQUIRE(ok, FFHIP_ERR_ARG, "matrix file %s: entry %" PRId64 " unreadable", path, k);
    i -= f77;
    j -= f77;
    FFHIP_REQUIRE(i >= 0 && i < n && j >= 0 && j < m, FFHIP_ERR_ARG,
                  "matrix file %s: entry %" PRId64 " out of range", path, k);
    key[(size_t)k] = i * m + j;
    a[(size_t)k] = v;
  }
  std::vector<int64_t> ord((size_t)cnt);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return key[x] < key[y]; });
  P.n = n;
  P.m = m;
  P.half = (int)half;
  P.rowptr.assign((size_t)n + 1, 0);
  P.colidx.clear();
  P.vals.clear();
  for (size_t q = 0; q < ord.size(); ++q) {
    const int64_t kq = key[ord[q]];
    if (q + 1 < ord.size() && key[ord[q + 1]] == kq) continue;  // a later assignment wins
    P.rowptr[(size_t)(kq / m) + 1]++;
    P.colidx.push_back((int32_t)(kq % m));
    P.vals.push_back(a[ord[q]]);
  }
  for (int64_t i = 0; i < n; ++i) P.rowptr[i + 1] += P.rowptr[i];
}

}  // namespace
}  // namespace ffhip

using namespace ffhip;

extern "C" {

int ffhip_mat_coo(ffhip_mat A, int64_t *count, int64_t *I, int64_t *J, double *C, int on_device) {
  return guard([&] {
    FFHIP_REQUIRE(A && count, FFHIP_ERR_ARG, "ffhip_mat_coo: bad arguments");
    ffhip_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = A->n, nnz = A->nnz;
    FFHIP_REQUIRE(n > 0, FFHIP_ERR_ARG, "ffhip_mat_coo: matrix without rows");
    // HashMatrix has no (n-1, m-1) entry: [I,J,C] = A appends (n, m, 0)
    int64_t last[2] = {0, 0};
    FFHIP_CHECK(hipMemcpyAsync(last, A->rowptr + n - 1, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    FFHIP_CHECK(hipStreamSynchronize(s));
    int32_t lastcol = -1;
    if (last[1] > last[0])
      FFHIP_CHECK(hipMemcpy(&lastcol, A->colidx + last[1] - 1, sizeof(int32_t), hipMemcpyDeviceToHost));
    const bool extra = lastcol != (int32_t)(n - 1);
    *count = nnz + (extra ? 1 : 0);
    if (!I && !J && !C) return;
    FFHIP_REQUIRE(I && J && C, FFHIP_ERR_ARG, "ffhip_mat_coo: I, J and C are all needed");
    int64_t *Id = on_device ? I : dalloc<int64_t>(nnz);
    int64_t *Jd = on_device ? J : dalloc<int64_t>(nnz);
    ctx->tic();
    if (nnz)
      k_coo_rows<<<nblocks(nnz, 256), 256, 0, s>>>(n, nnz, A->rowptr, A->colidx, Id, Jd);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    FFHIP_CHECK(hipMemcpyAsync(C, A->vals, 8 * (size_t)nnz, kind, s));
    if (!on_device) {
      FFHIP_CHECK(hipMemcpyAsync(I, Id, 8 * (size_t)nnz, hipMemcpyDeviceToHost, s));
      FFHIP_CHECK(hipMemcpyAsync(J, Jd, 8 * (size_t)nnz, hipMemcpyDeviceToHost, s));
    }
    const int64_t tail[2] = {n, n};  // (n, m) with m = n
    const double zero = 0.;
    if (extra && on_device) {
      FFHIP_CHECK(hipMemcpyAsync(I + nnz, &tail[0], 8, hipMemcpyHostToDevice, s));
      FFHIP_CHECK(hipMemcpyAsync(J + nnz, &tail[1], 8, hipMemcpyHostToDevice, s));
      FFHIP_CHECK(hipMemcpyAsync(C + nnz, &zero, 8, hipMemcpyHostToDevice, s));
    }
    FFHIP_CHECK(hipStreamSynchronize(s));  // the host sources above stay alive until here
    if (!on_device) {
      if (extra) { I[nnz] = tail[0]; J[nnz] = tail[1]; C[nnz] = zero; }
      dfree(Id);
      dfree(Jd);
    }
  });
}

int ffhip_csr_write_morse(int64_t n, int64_t m, int half, const int64_t *rowptr,
                          const int32_t *colidx, const double *vals, const char *path) {
  return guard([&] {
    FFHIP_REQUIRE(n >= 0 && m >= 0 && rowptr && path && (rowptr[n] == rowptr[0] || (colidx && vals)),
                  FFHIP_ERR_ARG, "ffhip_csr_write_morse: bad arguments");
    FILE *f = fopen(path, "wb");
    FFHIP_REQUIRE(f, FFHIP_ERR_ARG, "ffhip_csr_write_morse: cannot open %s", path);
    write_morse(f, n, m, half, rowptr, colidx, vals, rowptr);
    const bool ok = !ferror(f);
    fclose(f);
    FFHIP_REQUIRE(ok, FFHIP_ERR_ARG, "ffhip_csr_write_morse: write error on %s", path);
  });
}

int ffhip_mat_write_morse(ffhip_mat A, const char *path) {
  return guard([&] {
    FFHIP_REQUIRE(A && path, FFHIP_ERR_ARG, "ffhip_mat_write_morse: bad arguments");
    std::vector<int64_t> rp((size_t)A->n + 1);
    std::vector<int32_t> ci((size_t)A->nnz);
    std::vector<double> va((size_t)A->nnz);
    const int rc = ffhip_mat_download(A, rp.data(), ci.data(), va.data());
    if (rc != FFHIP_OK) throw Error{rc};
    const int rc2 = ffhip_csr_write_morse(A->n, A->n, A->half, rp.data(), ci.data(), va.data(), path);
    if (rc2 != FFHIP_OK) throw Error{rc2};
  });
}

int ffhip_csr_read_text(const char *path, int64_t cap, int64_t *n, int64_t *m, int *half,
                        int64_t *nnz, int64_t *rowptr, int32_t *colidx, double *vals) {
  return guard([&] {
    FFHIP_REQUIRE(path && n && m && half && nnz, FFHIP_ERR_ARG, "ffhip_csr_read_text: bad arguments");
    Parsed P;
    parse_text(path, P);
    *n = P.n;
    *m = P.m;
    *half = P.half;
    *nnz = (int64_t)P.vals.size();
    if (!rowptr && !colidx && !vals) return;
    FFHIP_REQUIRE(rowptr && colidx && vals, FFHIP_ERR_ARG, "ffhip_csr_read_text: output arrays");
    FFHIP_REQUIRE(*nnz <= cap, FFHIP_ERR_ARG, "ffhip_csr_read_text: %" PRId64 " entries, capacity %" PRId64,
                  *nnz, cap);
    std::copy(P.rowptr.begin(), P.rowptr.end(), rowptr);
    std::copy(P.colidx.begin(), P.colidx.end(), colidx);
    std::copy(P.vals.begin(), P.vals.end(), vals);
  });
}

int ffhip_mat_read_text(ffhip_ctx ctx, const char *path, ffhip_mat *mat) {
  return guard([&] {
    FFHIP_REQUIRE(ctx && path && mat, FFHIP_ERR_ARG, "ffhip_mat_read_text: bad arguments");
    Parsed P;
    parse_text(path, P);
    FFHIP_REQUIRE(P.n == P.m, FFHIP_ERR_UNSUPPORTED, "ffhip_mat_read_text: %s is not square", path);
    if (P.half) {  // lower triangle stored: the full symmetric matrix
      std::vector<int64_t> cnt((size_t)P.n + 1, 0);
      for (int64_t i = 0; i < P.n; ++i)
        for (int64_t k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
          cnt[(size_t)i + 1]++;
          if (P.colidx[k] != i) cnt[(size_t)P.colidx[k] + 1]++;
        }
      for (int64_t i = 0; i < P.n; ++i) cnt[i + 1] += cnt[i];
      std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
      std::vector<int32_t> ci((size_t)cnt[P.n]);
      std::vector<double> va((size_t)cnt[P.n]);
      for (int64_t i = 0; i < P.n; ++i)  // each stored entry and its mirror ...
        for (int64_t k = P.rowptr[i]; k < P.rowptr[i + 1]; ++k) {
          const int64_t j = P.colidx[k];
          ci[pos[i]] = (int32_t)j; va[pos[i]++] = P.vals[k];
          if (j != i) { ci[pos[j]] = (int32_t)i; va[pos[j]++] = P.vals[k]; }
        }
      for (int64_t i = 0; i < P.n; ++i) {  // ... then each row sorted by column
        std::vector<std::pair<int32_t, double>> row;
        for (int64_t k = cnt[i]; k < cnt[i + 1]; ++k) row.emplace_back(ci[k], va[k]);
        std::sort(row.begin(), row.end(),
                  [](const std::pair<int32_t, double> &x, const std::pair<int32_t, double> &y) {
                    return x.first < y.first;
                  });
        for (size_t q = 0; q < row.size(); ++q) { ci[cnt[i] + q] = row[q].first; va[cnt[i] + q] = row[q].second; }
      }
      P.rowptr.swap(cnt);
      P.colidx.swap(ci);
      P.vals.swap(va);
    }
    const int rc = ffhip_mat_upload(ctx, P.n, (int64_t)P.vals.size(), P.rowptr.data(), P.colidx.data(),
                                    P.vals.data(), mat);
    if (rc != FFHIP_OK) throw Error{rc};
  });
}

}  // extern "C"
