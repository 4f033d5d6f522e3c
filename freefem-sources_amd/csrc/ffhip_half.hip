// ffhip_half.hip — symmetric half storage, FreeFEM's `sym=1` matrices
// (HashMatrix::half): only the lower triangle (row >= column) is stored.
//
//  * ffhip_mat_lower: the half matrix of a symmetric device matrix — what
//    `matrix A = va(Vh, Vh, sym=1)` holds after the symmetric element
//    matrices are added (HashMatrix::operator+= with
//    MatriceElementaire::Symmetric stores (max(i,j), min(i,j)),
//    HashMatrix.cpp:1312-1318);
//  * ffhip_spmv on a half matrix: y = (L + L^T - diag(L)) x, the half branch
//    of HashMatrix::addMatMul (HashMatrix.cpp:1128-1136);
//  * ffhip_mat_full: the full symmetric matrix of a half one.
//
// A half matrix keeps a "full view" built once: for every full row i, its
// lower entries (columns <= i, from row i of L) followed by the transposed
// entries L(r, i), r > i, in ascending r — the sorted columns of the full row
// — each as a slot of L's values.  A row sum over the view adds in the order
// FreeFEM's half product accumulates y_i (row i's own entries, then the later
// rows), and L's values stay the only copy.  The transposed lists come from a
// radix sort of (column, row) keys, so the view is deterministic.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "ffhip_internal.hpp"

namespace ffhip {

int read_flag(ffhip_ctx ctx, int *d);

namespace {

unsigned nblocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

// entries with column <= row, per row
__global__ void k_lower_count(int64_t n, const int64_t *__restrict__ rowptr,
                              const int32_t *__restrict__ colidx, int64_t *__restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t c = 0;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) c += colidx[k] <= i;
  cnt[i] = c;
}

__global__ void k_lower_fill(int64_t n, const int64_t *__restrict__ rowptr,
                             const int32_t *__restrict__ colidx, const double *__restrict__ vals,
                             const int64_t *__restrict__ lptr, int32_t *__restrict__ lcol,
                             double *__restrict__ lval, int64_t *__restrict__ ldiag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t o = lptr[i];
  ldiag[i] = -1;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1] && colidx[k] <= i; ++k, ++o) {
    lcol[o] = colidx[k];
    lval[o] = vals[k];
    if (colidx[k] == i) ldiag[i] = o;
  }
}

// a half matrix must hold no entry above the diagonal
__global__ void k_check_lower(int64_t n, const int64_t *__restrict__ rowptr,
                              const int32_t *__restrict__ colidx, int *__restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t e = rowptr[i + 1];
  if (e > rowptr[i] && colidx[e - 1] > i) atomicOr(bad, 1);
}

// strictly lower entries: per-column counts and (column, row) keys with their slots
__global__ void k_half_keys(int64_t n, const int64_t *__restrict__ rowptr,
                            const int32_t *__restrict__ colidx, const int64_t *__restrict__ spos,
                            unsigned long long *__restrict__ ccnt, uint64_t *__restrict__ keys,
                            int64_t *__restrict__ slots) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t o = spos[i];
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
    const int32_t c = colidx[k];
    if (c >= i) continue;
    atomicAdd(&ccnt[c], 1ull);  // a count: the order of the additions does not matter
    keys[o] = ((uint64_t)(uint32_t)c << 32) | (uint32_t)i;
    slots[o++] = k;
  }
}

__global__ void k_strict_count(int64_t n, const int64_t *__restrict__ rowptr,
                               const int32_t *__restrict__ colidx, int64_t *__restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t c = 0;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) c += colidx[k] < i;
  cnt[i] = c;
}

__global__ void k_view_count(int64_t n, const int64_t *__restrict__ rowptr,
                             const unsigned long long *__restrict__ ccnt, int64_t *__restrict__ fcnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  fcnt[i] = (rowptr[i + 1] - rowptr[i]) + (int64_t)ccnt[i];
}

// row i of the view: its own lower entries, then (after the sort) the transposed ones
__global__ void k_view_own(int64_t n, const int64_t *__restrict__ rowptr,
                           const int32_t *__restrict__ colidx, const int64_t *__restrict__ fptr,
                           int32_t *__restrict__ fcol, int64_t *__restrict__ fslot) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int64_t o = fptr[i];
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k, ++o) {
    fcol[o] = colidx[k];
    fslot[o] = k;
  }
}

// sorted strictly-lower entry t = (column c, row r): the (t - cstart[c])-th
// transposed entry of view row c
__global__ void k_view_trans(int64_t m, const uint64_t *__restrict__ skeys,
                             const int64_t *__restrict__ sslots, const int64_t *__restrict__ cstart,
                             const int64_t *__restrict__ rowptr, const int64_t *__restrict__ fptr,
                             int32_t *__restrict__ fcol, int64_t *__restrict__ fslot) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int64_t c = (int64_t)(skeys[t] >> 32);
  const int32_t r = (int32_t)(skeys[t] & 0xffffffffu);
  const int64_t o = fptr[c] + (rowptr[c + 1] - rowptr[c]) + (t - cstart[c]);
  fcol[o] = r;
  fslot[o] = sslots[t];
}

__global__ void k_gather_vals(int64_t m, const int64_t *__restrict__ slot,
                              const double *__restrict__ v, double *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < m) out[k] = v[slot[k]];
}

__global__ void k_diag_slots(int64_t n, const int64_t *__restrict__ rowptr,
                             const int32_t *__restrict__ colidx, int64_t *__restrict__ diag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  diag[i] = -1;
  for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k)
    if (colidx[k] == i) diag[i] = k;
}

__global__ void k_u64_to_i64(int64_t n, const unsigned long long *__restrict__ a,
                             int64_t *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (int64_t)a[i];
}

}  // namespace

// the full view of a half matrix (built once; the pattern of L never changes)
void half_view(ffhip_mat L) {
  if (L->fptr) return;
  ffhip_ctx ctx = L->ctx;
  hipStream_t s = ctx->stream;
  const int64_t n = L->n;
  int64_t *scnt = dalloc<int64_t>(n + 1), *spos = dalloc<int64_t>(n + 1);
  FFHIP_CHECK(hipMemsetAsync(scnt + n, 0, 8, s));
  k_strict_count<<<nblocks(n, 256), 256, 0, s>>>(n, L->rowptr, L->colidx, scnt);
  exclusive_scan_i64(ctx, scnt, spos, n + 1);
  const int64_t m = read_i64(ctx, spos + n);  // strictly lower entries
  if (m >= INT32_MAX) {
    dfree(scnt); dfree(spos);
    FFHIP_REQUIRE(false, FFHIP_ERR_UNSUPPORTED, "half matrix: too many entries");
  }
  unsigned long long *ccnt = dalloc<unsigned long long>(n);
  uint64_t *keys = dalloc<uint64_t>(m), *skeys = dalloc<uint64_t>(m);
  int64_t *slots = dalloc<int64_t>(m), *sslots = dalloc<int64_t>(m);
  FFHIP_CHECK(hipMemsetAsync(ccnt, 0, 8 * n, s));
  k_half_keys<<<nblocks(n, 256), 256, 0, s>>>(n, L->rowptr, L->colidx, spos, ccnt, keys, slots);
  size_t tb = 0;
  char *tmp = nullptr;
  if (m) {
    FFHIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys, skeys, slots, sslots, (int)m, 0,
                                                   64, s));
    tmp = dalloc<char>(tb);
    FFHIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, tb, keys, skeys, slots, sslots, (int)m, 0,
                                                   64, s));
  }
  // column starts of the sorted list, and the view's rows
  int64_t *cstart = dalloc<int64_t>(n + 1), *fcnt = dalloc<int64_t>(n + 1);
  k_view_count<<<nblocks(n, 256), 256, 0, s>>>(n, L->rowptr, ccnt, fcnt);
  FFHIP_CHECK(hipMemsetAsync(fcnt + n, 0, 8, s));
  L->fptr = dalloc<int64_t>(n + 1);
  exclusive_scan_i64(ctx, fcnt, L->fptr, n + 1);
  k_u64_to_i64<<<nblocks(n, 256), 256, 0, s>>>(n, ccnt, fcnt);  // reuse: per-column counts
  FFHIP_CHECK(hipMemsetAsync(fcnt + n, 0, 8, s));
  exclusive_scan_i64(ctx, fcnt, cstart, n + 1);
  L->fnnz = read_i64(ctx, L->fptr + n);
  L->fcol = dalloc<int32_t>(L->fnnz);
  L->fslot = dalloc<int64_t>(L->fnnz);
  k_view_own<<<nblocks(n, 256), 256, 0, s>>>(n, L->rowptr, L->colidx, L->fptr, L->fcol, L->fslot);
  if (m)
    k_view_trans<<<nblocks(m, 256), 256, 0, s>>>(m, skeys, sslots, cstart, L->rowptr, L->fptr,
                                                  L->fcol, L->fslot);
  FFHIP_CHECK(hipGetLastError());
  FFHIP_CHECK(hipStreamSynchronize(s));
  dfree(scnt); dfree(spos); dfree(ccnt); dfree(keys); dfree(skeys); dfree(slots); dfree(sslots);
  dfree(tmp); dfree(cstart); dfree(fcnt);
}

}  // namespace ffhip

using namespace ffhip;

extern "C" {

int ffhip_mat_lower(ffhip_mat A, ffhip_mat *out) {
  return guard([&] {
    FFHIP_REQUIRE(A && out && !A->half, FFHIP_ERR_ARG, "ffhip_mat_lower: needs a full matrix");
    ffhip_ctx ctx = A->ctx;
    hipStream_t s = ctx->stream;
    const int64_t n = A->n;
    int64_t *cnt = dalloc<int64_t>(n + 1);
    FFHIP_CHECK(hipMemsetAsync(cnt + n, 0, 8, s));
    ffhip_mat L = new ffhip_mat_s();
    L->ctx = ctx;
    L->n = n;
    L->half = 1;
    L->rowptr = dalloc<int64_t>(n + 1);
    L->diag = dalloc<int64_t>(n);
    ctx->tic();
    k_lower_count<<<nblocks(n, 256), 256, 0, s>>>(n, A->rowptr, A->colidx, cnt);
    exclusive_scan_i64(ctx, cnt, L->rowptr, n + 1);
    L->nnz = read_i64(ctx, L->rowptr + n);
    L->colidx = dalloc<int32_t>(L->nnz);
    L->vals = dalloc<double>(L->nnz);
    k_lower_fill<<<nblocks(n, 256), 256, 0, s>>>(n, A->rowptr, A->colidx, A->vals, L->rowptr,
                                                 L->colidx, L->vals, L->diag);
    FFHIP_CHECK(hipGetLastError());
    ctx->toc();
    dfree(cnt);
    *out = L;
  });
}

int ffhip_mat_set_half(ffhip_mat A, int half) {
  return guard([&] {
    FFHIP_REQUIRE(A, FFHIP_ERR_ARG, "ffhip_mat_set_half: null matrix");
    if (!half) {
      FFHIP_REQUIRE(!A->half, FFHIP_ERR_UNSUPPORTED,
                    "ffhip_mat_set_half: a half matrix stays half (ffhip_mat_full expands it)");
      return;
    }
    int *bad = dalloc<int>(1);
    FFHIP_CHECK(hipMemsetAsync(bad, 0, sizeof(int), A->ctx->stream));
    k_check_lower<<<nblocks(A->n, 256), 256, 0, A->ctx->stream>>>(A->n, A->rowptr, A->colidx, bad);
    FFHIP_CHECK(hipGetLastError());
    const int b = read_flag(A->ctx, bad);
    dfree(bad);
    FFHIP_REQUIRE(!b, FFHIP_ERR_ARG, "ffhip_mat_set_half: the matrix has entries above the diagonal");
    A->half = 1;
  });
}

int ffhip_mat_full(ffhip_mat L, ffhip_mat *out) {
  return guard([&] {
    FFHIP_REQUIRE(L && out && L->half, FFHIP_ERR_ARG, "ffhip_mat_full: needs a half matrix");
    ffhip_ctx ctx = L->ctx;
    hipStream_t s = ctx->stream;
    half_view(L);
    ffhip_mat A = new ffhip_mat_s();
    A->ctx = ctx;
    A->n = L->n;
    A->nnz = L->fnnz;
    A->rowptr = dalloc<int64_t>(L->n + 1);
    A->colidx = dalloc<int32_t>(L->fnnz);
    A->vals = dalloc<double>(L->fnnz);
    A->diag = dalloc<int64_t>(L->n);
    FFHIP_CHECK(hipMemcpyAsync(A->rowptr, L->fptr, 8 * (L->n + 1), hipMemcpyDeviceToDevice, s));
    FFHIP_CHECK(hipMemcpyAsync(A->colidx, L->fcol, 4 * L->fnnz, hipMemcpyDeviceToDevice, s));
    if (L->fnnz)
      k_gather_vals<<<nblocks(L->fnnz, 256), 256, 0, s>>>(L->fnnz, L->fslot, L->vals, A->vals);
    k_diag_slots<<<nblocks(L->n, 256), 256, 0, s>>>(L->n, A->rowptr, A->colidx, A->diag);
    FFHIP_CHECK(hipGetLastError());
    FFHIP_CHECK(hipStreamSynchronize(s));
    *out = A;
  });
}

int ffhip_mat_half(ffhip_mat A, int *half) {
  return guard([&] {
    FFHIP_REQUIRE(A && half, FFHIP_ERR_ARG, "ffhip_mat_half: bad arguments");
    *half = A->half;
  });
}

}  // extern "C"
