// Device construction of the vev sparsity index and the ve incidence lists.
//
// The reference evaluates the path expression vev = (ve and ev, exists e) on
// the host with map<unsigned, set<unsigned>> per row (src/path_indices.cpp:
// 359-429) and packs sorted rows (:172-203).  Here:
//   1. count incidences per vertex (atomics only on integer counters),
//   2. exclusive scan -> ve_ptr, scatter element ids -> ve_elem,
//   3. sort every ve list ascending (element order is what the deterministic
//      assembly sums in),
//   4. per row, merge the endpoints of its elements into a sorted unique list
//      (thread per row in an LDS slice; rows that overflow the slice go to a
//      workgroup-per-row bitonic path), count, scan -> rowptr, fill colidx.
// Output is independent of scheduling, so it is bit-identical to the host
// builder's.
#include "smh_internal.h"

namespace {

constexpr int kRowCap = 64;     // per-thread LDS slice (entries)
constexpr int kRowThreads = 64; // threads per block in the thread-per-row path
constexpr int kBigCap = 8192;   // candidates per row in the workgroup path

struct SetsDev {
  const int32_t *ep[4];
  const uint32_t *ve_ptr[4];
  const uint32_t *ve_elem[4];
  uint32_t card[4];
  uint32_t nsets;
};

__global__ void k_count_inc(const int32_t *__restrict__ ep, uint64_t n, uint32_t nv,
                            uint32_t *__restrict__ deg, int *__restrict__ err) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    int32_t v = ep[i];
    if (v < 0 || (uint32_t)v >= nv) {
      atomicOr(err, 1);
      continue;
    }
    atomicAdd(&deg[v], 1u);
  }
}

__global__ void k_fill_inc(const int32_t *__restrict__ ep, uint64_t n, uint32_t card,
                           uint32_t *__restrict__ cursor, uint32_t *__restrict__ ve_elem) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    int32_t v = ep[i];
    uint32_t pos = atomicAdd(&cursor[v], 1u);
    ve_elem[pos] = (uint32_t)(i / card);
  }
}

// Sort each vertex's incidence list ascending, keeping duplicates (an element
// that holds the vertex twice is listed twice, as Link::ve does).  Lists
// longer than kRowCap are queued for the workgroup kernel.
__global__ void __launch_bounds__(kRowThreads)
k_sort_ve_small(const uint32_t *__restrict__ ptr, uint32_t *__restrict__ elem, uint32_t nv,
                uint32_t *__restrict__ big, uint32_t *__restrict__ nbig) {
  __shared__ uint32_t s[kRowCap * kRowThreads];
  const uint32_t t = threadIdx.x;
  const uint32_t v = blockIdx.x * kRowThreads + t;
  if (v >= nv) return;
  const uint32_t b = ptr[v], e = ptr[v + 1], n = e - b;
  if (n > (uint32_t)kRowCap) {
    big[atomicAdd(nbig, 1u)] = v;
    return;
  }
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t u = elem[b + k];
    int pos = (int)k;
    while (pos > 0 && s[(pos - 1) * kRowThreads + t] > u) {
      s[pos * kRowThreads + t] = s[(pos - 1) * kRowThreads + t];
      --pos;
    }
    s[pos * kRowThreads + t] = u;
  }
  for (uint32_t k = 0; k < n; ++k) elem[b + k] = s[k * kRowThreads + t];
}

__device__ void bitonic_sort_lds(uint32_t *a, int P) {
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        int l = i ^ j;
        if (l > i) {
          bool up = (i & k) == 0;
          uint32_t x = a[i], y = a[l];
          if ((x > y) == up) {
            a[i] = y;
            a[l] = x;
          }
        }
      }
      __syncthreads();
    }
}

__global__ void __launch_bounds__(256)
k_sort_ve_big(const uint32_t *__restrict__ ptr, uint32_t *__restrict__ elem,
              const uint32_t *__restrict__ big, int *__restrict__ err) {
  __shared__ uint32_t a[kBigCap];
  const uint32_t v = big[blockIdx.x];
  const uint32_t b = ptr[v], n = ptr[v + 1] - b;
  if (n > (uint32_t)kBigCap) {
    if (threadIdx.x == 0) atomicOr(err, 2);
    return;
  }
  int P = 1;
  while (P < (int)n) P <<= 1;
  for (int i = threadIdx.x; i < P; i += 256) a[i] = i < (int)n ? elem[b + i] : 0xffffffffu;
  __syncthreads();
  bitonic_sort_lds(a, P);
  for (int i = threadIdx.x; i < (int)n; i += 256) elem[b + i] = a[i];
}

// ---- vev rows ---------------------------------------------------------------
__device__ __forceinline__ bool insert_unique(uint32_t *s, int &n, uint32_t u, uint32_t t) {
  int pos = n;
  while (pos > 0 && s[(pos - 1) * kRowThreads + t] > u) --pos;
  if (pos > 0 && s[(pos - 1) * kRowThreads + t] == u) return true;
  if (n == kRowCap) return false;
  for (int k = n; k > pos; --k) s[k * kRowThreads + t] = s[(k - 1) * kRowThreads + t];
  s[pos * kRowThreads + t] = u;
  ++n;
  return true;
}

template <bool FILL>
__global__ void __launch_bounds__(kRowThreads)
k_vev_small(SetsDev S, uint32_t nv, int force_diag, uint32_t *__restrict__ cnt,
            const uint32_t *__restrict__ rowptr, uint32_t *__restrict__ colidx,
            uint32_t *__restrict__ big, uint32_t *__restrict__ nbig) {
  __shared__ uint32_t s[kRowCap * kRowThreads];
  const uint32_t t = threadIdx.x;
  const uint32_t v = blockIdx.x * kRowThreads + t;
  if (v >= nv) return;
  int n = 0;
  bool ok = true;
  for (uint32_t q = 0; q < S.nsets && ok; ++q) {
    const uint32_t card = S.card[q];
    const int32_t *ep = S.ep[q];
    for (uint32_t i = S.ve_ptr[q][v]; i < S.ve_ptr[q][v + 1] && ok; ++i) {
      const uint64_t e = S.ve_elem[q][i];
      for (uint32_t a = 0; a < card && ok; ++a) ok = insert_unique(s, n, (uint32_t)ep[e * card + a], t);
    }
  }
  if (ok && force_diag) ok = insert_unique(s, n, v, t);
  if (!ok) {
    if (!FILL) {
      big[atomicAdd(nbig, 1u)] = v;
      cnt[v] = 0;
    }
    return;
  }
  if (!FILL) {
    cnt[v] = (uint32_t)n;
  } else {
    const uint32_t b = rowptr[v];
    for (int k = 0; k < n; ++k) colidx[b + k] = s[k * kRowThreads + t];
  }
}

template <bool FILL>
__global__ void __launch_bounds__(256)
k_vev_big(SetsDev S, int force_diag, const uint32_t *__restrict__ big,
          uint32_t *__restrict__ cnt, const uint32_t *__restrict__ rowptr,
          uint32_t *__restrict__ colidx, int *__restrict__ err) {
  __shared__ uint32_t a[kBigCap];
  __shared__ uint32_t part[256];
  const uint32_t v = big[blockIdx.x];
  uint32_t m = force_diag ? 1u : 0u;
  for (uint32_t q = 0; q < S.nsets; ++q) m += (S.ve_ptr[q][v + 1] - S.ve_ptr[q][v]) * S.card[q];
  if (m > (uint32_t)kBigCap) {
    if (threadIdx.x == 0) atomicOr(err, 4);
    return;
  }
  int P = 1;
  while (P < (int)m) P <<= 1;
  for (int i = threadIdx.x; i < P; i += 256) a[i] = 0xffffffffu;
  __syncthreads();
  uint32_t off = 0;
  for (uint32_t q = 0; q < S.nsets; ++q) {
    const uint32_t card = S.card[q];
    const uint32_t b = S.ve_ptr[q][v], n = (S.ve_ptr[q][v + 1] - b) * card;
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
      const uint64_t e = S.ve_elem[q][b + i / card];
      a[off + i] = (uint32_t)S.ep[q][e * card + i % card];
    }
    off += n;
  }
  if (force_diag && threadIdx.x == 0) a[off] = v;
  __syncthreads();
  bitonic_sort_lds(a, P);
  // unique count per contiguous chunk, then a block scan of the chunk counts
  const int chunk = (P + 255) / 256;
  const int lo = threadIdx.x * chunk, hi = min(P, lo + chunk);
  uint32_t c = 0;
  for (int i = lo; i < hi; ++i)
    if (a[i] != 0xffffffffu && (i == 0 || a[i] != a[i - 1])) ++c;
  part[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int i = 0; i < 256; ++i) {
      uint32_t x = part[i];
      part[i] = run;
      run += x;
    }
    if (!FILL) cnt[v] = run;
  }
  __syncthreads();
  if (FILL) {
    uint32_t w = rowptr[v] + part[threadIdx.x];
    for (int i = lo; i < hi; ++i)
      if (a[i] != 0xffffffffu && (i == 0 || a[i] != a[i - 1])) colidx[w++] = a[i];
  }
}

__global__ void k_locs(const int32_t *__restrict__ ep, uint64_t ne, uint32_t card,
                       const uint32_t *__restrict__ rowptr, const uint32_t *__restrict__ colidx,
                       int32_t *__restrict__ locs) {
  const uint64_t n = ne * card * card;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = i / (card * card);
    const uint32_t r = (uint32_t)(i % (card * card));
    const uint32_t a = r / card, b = r % card;
    const uint32_t row = (uint32_t)ep[e * card + a], col = (uint32_t)ep[e * card + b];
    uint32_t l = rowptr[row];
    const uint32_t end = rowptr[row + 1];
    while (l < end && colidx[l] != col) ++l;
    locs[i] = l < end ? (int32_t)l : -1;
  }
}

// row of the first block of every 64-block assembly wave
__global__ void k_wave_rows(const uint32_t *__restrict__ rowptr, uint32_t nrows, uint64_t nwaves,
                            uint32_t *__restrict__ wave_row) {
  const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwaves) return;
  const uint64_t b0 = w * 64;
  uint32_t lo = 0, hi = nrows;  // rowptr[lo] <= b0 < rowptr[hi]
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if ((uint64_t)rowptr[mid] <= b0) lo = mid; else hi = mid;
  }
  wave_row[w] = lo;
}

void free_index(smh_index *ix) {
  if (!ix) return;
  smh::dfree(ix->rowptr);
  smh::dfree(ix->colidx);
  smh::dfree(ix->wave_row);
  for (auto p : ix->ve_ptr) smh::dfree(p);
  for (auto p : ix->ve_elem) smh::dfree(p);
  for (auto &a : ix->asm_plan) {
    smh::dfree(a.slot);
    smh::dfree(a.mask);
    smh::dfree(a.rank);
    smh::dfree(a.wnd);
    smh::dfree(a.wnc);
  }
  delete ix;
}

}  // namespace

using namespace smh;

extern "C" int smh_index_build(smh_ctx *ctx, uint32_t nv, uint32_t nsets, const uint64_t *ne,
                               const uint32_t *card, const int32_t *const *d_ep,
                               uint32_t flags, smh_index **out) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, out != nullptr, SMH_ERR_INVALID, "smh_index_build: out is null");
  *out = nullptr;
  SMH_CHECK(ctx, nsets >= 1 && nsets <= 4, SMH_ERR_UNSUPPORTED,
            "smh_index_build: 1 to 4 edge sets supported, got %u", nsets);
  SMH_CHECK(ctx, ne && card && d_ep, SMH_ERR_INVALID, "smh_index_build: null set arrays");
  for (uint32_t s = 0; s < nsets; ++s) {
    SMH_CHECK(ctx, card[s] >= 1 && card[s] <= 64, SMH_ERR_UNSUPPORTED,
              "smh_index_build: set %u has cardinality %u (1..64 supported)", s, card[s]);
    SMH_CHECK(ctx, ne[s] == 0 || d_ep[s] != nullptr, SMH_ERR_INVALID,
              "smh_index_build: set %u has null endpoints", s);
    SMH_CHECK(ctx, ne[s] * card[s] < 0xffffffffull, SMH_ERR_UNSUPPORTED,
              "smh_index_build: set %u has too many incidences", s);
  }
  SMH_HIP(ctx, hipSetDevice(ctx->device));
  smh_index *ix = new smh_index();
  ix->nrows = nv;
  ix->nsets = nsets;
  int rc = SMH_OK;
  uint32_t *deg = nullptr, *cursor = nullptr, *big = nullptr, *nbig = nullptr, *cnt = nullptr;
  int *err = nullptr;
  SetsDev S{};
  S.nsets = nsets;
#define TRY(x)            \
  do {                    \
    rc = (x);             \
    if (rc) goto cleanup; \
  } while (0)
  TRY(dmalloc(ctx, (void **)&deg, sizeof(uint32_t) * ((size_t)nv + 1)));
  TRY(dmalloc(ctx, (void **)&cursor, sizeof(uint32_t) * ((size_t)nv + 1)));
  TRY(dmalloc(ctx, (void **)&big, sizeof(uint32_t) * ((size_t)nv + 1)));
  TRY(dmalloc(ctx, (void **)&nbig, sizeof(uint32_t) * 4));
  TRY(dmalloc(ctx, (void **)&err, sizeof(int) * 4));
  TRY(dmalloc(ctx, (void **)&cnt, sizeof(uint32_t) * ((size_t)nv + 1)));
  if (hipMemsetAsync(err, 0, sizeof(int) * 4, ctx->stream) != hipSuccess) {
    rc = fail(ctx, SMH_ERR_HIP, "memset failed");
    goto cleanup;
  }
  for (uint32_t s = 0; s < nsets; ++s) {
    const uint64_t n = ne[s] * card[s];
    uint32_t *vp = nullptr, *vel = nullptr;
    TRY(dmalloc(ctx, (void **)&vp, sizeof(uint32_t) * ((size_t)nv + 1)));
    ix->ve_ptr.push_back(vp);
    TRY(dmalloc(ctx, (void **)&vel, sizeof(uint32_t) * (n ? n : 1)));
    ix->ve_elem.push_back(vel);
    ix->ne.push_back(ne[s]);
    ix->card.push_back(card[s]);
    ix->ep.push_back(d_ep[s]);
    hipMemsetAsync(deg, 0, sizeof(uint32_t) * ((size_t)nv + 1), ctx->stream);
    if (n)
      hipLaunchKernelGGL(k_count_inc, dim3(grid_for(n, 256, 65536)), dim3(256), 0, ctx->stream,
                         d_ep[s], n, nv, deg, err);
    uint64_t tot = 0;
    TRY(scan_u32(ctx, deg, vp, nv, &tot));
    ix->n_inc.push_back(tot);
    int herr = 0;
    hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    hipStreamSynchronize(ctx->stream);
    if (herr) {
      rc = fail(ctx, SMH_ERR_INVALID, "smh_index_build: set %u has an endpoint outside [0, %u)",
                s, nv);
      goto cleanup;
    }
    hipMemcpyAsync(cursor, vp, sizeof(uint32_t) * ((size_t)nv + 1), hipMemcpyDeviceToDevice,
                   ctx->stream);
    if (n)
      hipLaunchKernelGGL(k_fill_inc, dim3(grid_for(n, 256, 65536)), dim3(256), 0, ctx->stream,
                         d_ep[s], n, card[s], cursor, vel);
    hipMemsetAsync(nbig, 0, sizeof(uint32_t), ctx->stream);
    if (nv)
      hipLaunchKernelGGL(k_sort_ve_small, dim3(grid_for(nv, kRowThreads)), dim3(kRowThreads), 0,
                         ctx->stream, vp, vel, nv, big, nbig);
    uint32_t hbig = 0;
    hipMemcpyAsync(&hbig, nbig, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    hipStreamSynchronize(ctx->stream);
    if (hbig)
      hipLaunchKernelGGL(k_sort_ve_big, dim3(hbig), dim3(256), 0, ctx->stream, vp, vel, big, err);
    S.ep[s] = d_ep[s];
    S.ve_ptr[s] = vp;
    S.ve_elem[s] = vel;
    S.card[s] = card[s];
  }
  {
    const int fd = (flags & SMH_INDEX_FORCE_DIAG) ? 1 : 0;
    hipMemsetAsync(nbig, 0, sizeof(uint32_t), ctx->stream);
    if (nv)
      hipLaunchKernelGGL(k_vev_small<false>, dim3(grid_for(nv, kRowThreads)), dim3(kRowThreads),
                         0, ctx->stream, S, nv, fd, cnt, (const uint32_t *)nullptr,
                         (uint32_t *)nullptr, big, nbig);
    uint32_t hbig = 0;
    hipMemcpyAsync(&hbig, nbig, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream);
    hipStreamSynchronize(ctx->stream);
    if (hbig)
      hipLaunchKernelGGL(k_vev_big<false>, dim3(hbig), dim3(256), 0, ctx->stream, S, fd,
                         (const uint32_t *)big, cnt, (const uint32_t *)nullptr,
                         (uint32_t *)nullptr, err);
    TRY(dmalloc(ctx, (void **)&ix->rowptr, sizeof(uint32_t) * ((size_t)nv + 1)));
    uint64_t nnz = 0;
    TRY(scan_u32(ctx, cnt, ix->rowptr, nv, &nnz));
    ix->nnz = nnz;
    TRY(dmalloc(ctx, (void **)&ix->colidx, sizeof(uint32_t) * (nnz ? nnz : 1)));
    if (nv)
      hipLaunchKernelGGL(k_vev_small<true>, dim3(grid_for(nv, kRowThreads)), dim3(kRowThreads),
                         0, ctx->stream, S, nv, fd, cnt, (const uint32_t *)ix->rowptr,
                         ix->colidx, big, nbig);
    if (hbig)
      hipLaunchKernelGGL(k_vev_big<true>, dim3(hbig), dim3(256), 0, ctx->stream, S, fd,
                         (const uint32_t *)big, cnt, (const uint32_t *)ix->rowptr, ix->colidx,
                         err);
    const uint64_t nwaves = (nnz + 63) / 64;
    TRY(dmalloc(ctx, (void **)&ix->wave_row, sizeof(uint32_t) * (nwaves ? nwaves : 1)));
    if (nwaves && nv)
      hipLaunchKernelGGL(k_wave_rows, dim3(grid_for(nwaves, 256)), dim3(256), 0, ctx->stream,
                         (const uint32_t *)ix->rowptr, nv, nwaves, ix->wave_row);
    int herr = 0;
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
      rc = fail(ctx, SMH_ERR_HIP, "index kernels failed: %s", hipGetErrorString(le));
      goto cleanup;
    }
    hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
      rc = fail(ctx, SMH_ERR_HIP, "index build failed on the device");
      goto cleanup;
    }
    if (herr) {
      rc = fail(ctx, SMH_ERR_UNSUPPORTED,
                "smh_index_build: a row has more than %d candidate neighbours", kBigCap);
      goto cleanup;
    }
  }
cleanup:
#undef TRY
  dfree(deg);
  dfree(cursor);
  dfree(big);
  dfree(nbig);
  dfree(err);
  dfree(cnt);
  if (rc) {
    free_index(ix);
    return rc;
  }
  *out = ix;
  return SMH_OK;
}

extern "C" int smh_index_get(const smh_index *idx, smh_index_view *view) {
  if (!idx || !view) return SMH_ERR_INVALID;
  view->nrows = idx->nrows;
  view->nnz = idx->nnz;
  view->rowptr = idx->rowptr;
  view->colidx = idx->colidx;
  return SMH_OK;
}

extern "C" int smh_index_get_ve(const smh_index *idx, uint32_t set, const uint32_t **ve_ptr,
                                const uint32_t **ve_elem, uint64_t *n_inc) {
  if (!idx || set >= idx->nsets) return SMH_ERR_INVALID;
  if (ve_ptr) *ve_ptr = idx->ve_ptr[set];
  if (ve_elem) *ve_elem = idx->ve_elem[set];
  if (n_inc) *n_inc = idx->n_inc[set];
  return SMH_OK;
}

extern "C" int smh_index_destroy(smh_ctx *ctx, smh_index *idx) {
  if (!ctx) return SMH_ERR_INVALID;
  if (!idx) return SMH_OK;
  hipStreamSynchronize(ctx->stream);
  smh::forget_plans(ctx, idx->rowptr);
  free_index(idx);
  return SMH_OK;
}

extern "C" int smh_index_locs(smh_ctx *ctx, const smh_index *idx, uint32_t set, int32_t *d_locs) {
  if (!ctx) return SMH_ERR_INVALID;
  SMH_CHECK(ctx, idx && set < idx->nsets && d_locs, SMH_ERR_INVALID, "smh_index_locs: bad argument");
  const uint64_t n = idx->ne[set] * idx->card[set] * idx->card[set];
  if (!n) return SMH_OK;
  hipLaunchKernelGGL(k_locs, dim3(grid_for(n, 256, 65536)), dim3(256), 0, ctx->stream,
                     idx->ep[set], idx->ne[set], idx->card[set], idx->rowptr, idx->colidx, d_locs);
  SMH_LAUNCH_CHECK(ctx);
  return SMH_OK;
}
