// Hub-first vertex order for the batched unweighted engine (no counterpart
// in the reference, whose vertices keep their file order).
//
// The dense pull gathers the frontier record of every in-neighbour, and on
// skewed graphs most gathers hit the few vertices of largest out-degree.
// Scattered over the vertex range, each such record shares its cache line
// with cold vertices; placed first, the hot records fill whole lines and a
// few MB of L2 / Infinity Cache hold them.  The view renumbers the k
// vertices of largest out-degree to [0, k) (descending degree, ties by id)
// and keeps every other vertex in its original order after them, so the
// distance pass, which writes rows in the original order, still reads
// almost contiguous records.  Results are identical (the order only moves
// work in memory); the engine maps sources in and rows out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "common.hpp"

namespace dawn {

int upload_i32(int32_t** dst, const std::vector<int32_t>& src);
int build_chunk_list(const std::vector<int32_t>& rp, int heavy, int chunk, int32_t** dv,
                     int32_t** db, int32_t** de, int64_t* count);

namespace {

// Row r of the new adjacency = row inv[r] of the old one, columns mapped
// through perm: one wave per row, lanes stride the edges.
__global__ void view_copy_kernel(int32_t n, const int32_t* old_rp, const int32_t* old_col,
                                 const int32_t* new_rp, int32_t* new_col, const int32_t* inv,
                                 const int32_t* perm) {
  const int lane = lane_id();
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nwaves = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t r = wave; r < n; r += nwaves) {
    const int32_t o = inv[r];
    const int32_t b = old_rp[o], e = old_rp[o + 1], nb = new_rp[r];
    for (int32_t k = b + lane; k < e; k += kWave) new_col[nb + (k - b)] = perm[old_col[k]];
  }
}

void free_ptr(int32_t*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

void ms_view_free(dawn_graph* g) {
  MsView* v = g->view;
  if (!v) return;
  (void)hipSetDevice(g->device);
  if (v->in_row_ptr == v->row_ptr) v->in_row_ptr = nullptr;
  if (v->in_col == v->col) v->in_col = nullptr;
  if (v->oc_v == v->hc_v) v->oc_v = v->oc_beg = v->oc_end = nullptr;
  for (int32_t** p : {&v->row_ptr, &v->col, &v->in_row_ptr, &v->in_col, &v->perm, &v->hc_v,
                      &v->hc_beg, &v->hc_end, &v->oc_v, &v->oc_beg, &v->oc_end})
    free_ptr(*p);
  delete v;
  g->view = nullptr;
}

// Builds the view if the knobs ask for one and the graph is large enough
// for cache placement to matter; returns DAWN_OK with g->view == null
// otherwise.
int ms_view_ensure(dawn_graph* g) {
  if (!g->knobs.hub_order || g->n < kViewMinVertices) {
    ms_view_free(g);
    return DAWN_OK;
  }
  if (g->view) return DAWN_OK;
  const int32_t n = g->n;
  const int32_t k = std::max<int32_t>(1, n / 16);
  std::vector<int32_t> deg(n), order(n);
  for (int32_t v = 0; v < n; v++) deg[v] = g->h_row_ptr[v + 1] - g->h_row_ptr[v];
  std::iota(order.begin(), order.end(), 0);
  auto hotter = [&](int32_t x, int32_t y) { return deg[x] != deg[y] ? deg[x] > deg[y] : x < y; };
  std::nth_element(order.begin(), order.begin() + (k - 1), order.end(), hotter);
  std::sort(order.begin(), order.begin() + k, hotter);
  std::vector<char> hub(n, 0);
  for (int32_t i = 0; i < k; i++) hub[order[i]] = 1;
  MsView* view = new MsView();
  view->k = k;
  view->h_perm.resize(n);
  std::vector<int32_t> inv(n);
  for (int32_t i = 0; i < k; i++) {
    view->h_perm[order[i]] = i;
    inv[i] = order[i];
  }
  int32_t next = k;
  for (int32_t v = 0; v < n; v++)
    if (!hub[v]) {
      view->h_perm[v] = next;
      inv[next++] = v;
    }
  auto offsets = [&](const std::vector<int32_t>& rp) {
    std::vector<int32_t> out((size_t)n + 1);
    out[0] = 0;
    for (int32_t r = 0; r < n; r++) out[r + 1] = out[r] + (rp[inv[r] + 1] - rp[inv[r]]);
    return out;
  };
  view->h_row_ptr = offsets(g->h_row_ptr);
  if (g->directed) view->h_in_row_ptr = offsets(g->h_in_row_ptr);
  g->view = view;
  auto fail = [&](int rc) {
    ms_view_free(g);
    return rc;
  };
  int rc;
  int32_t* d_inv = nullptr;
  if ((rc = upload_i32(&view->perm, view->h_perm)) || (rc = upload_i32(&d_inv, inv))) {
    free_ptr(d_inv);
    return fail(rc);
  }
  auto relabel = [&](const std::vector<int32_t>& new_rp, const int32_t* old_rp,
                     const int32_t* old_col, int32_t** rp_out, int32_t** col_out) -> int {
    DAWN_TRY(upload_i32(rp_out, new_rp));
    const hipError_t e = hipMalloc((void**)col_out, sizeof(int32_t) * std::max<int64_t>(g->m, 1));
    if (e != hipSuccess) {
      *col_out = nullptr;
      set_hip_error(e, "hipMalloc", __FILE__, __LINE__);
      return e == hipErrorOutOfMemory ? DAWN_ERR_OUT_OF_MEMORY : DAWN_ERR_HIP;
    }
    view_copy_kernel<<<2048, kBlock, 0, g->stream>>>(n, old_rp, old_col, *rp_out, *col_out, d_inv,
                                                     view->perm);
    DAWN_HIP(hipGetLastError());
    return DAWN_OK;
  };
  rc = relabel(view->h_row_ptr, g->row_ptr, g->col, &view->row_ptr, &view->col);
  if (!rc && g->directed)
    rc = relabel(view->h_in_row_ptr, g->in_row_ptr, g->in_col, &view->in_row_ptr, &view->in_col);
  if (!rc && hipStreamSynchronize(g->stream) != hipSuccess) rc = DAWN_ERR_HIP;
  free_ptr(d_inv);
  if (rc) return fail(rc);
  if (!g->directed) {
    view->in_row_ptr = view->row_ptr;
    view->in_col = view->col;
  }
  const int heavy = g->knobs.heavy_degree, chunk = g->knobs.pull_chunk;
  if ((rc = build_chunk_list(g->directed ? view->h_in_row_ptr : view->h_row_ptr, heavy, chunk,
                             &view->hc_v, &view->hc_beg, &view->hc_end, &view->num_hc)))
    return fail(rc);
  if (g->directed) {
    if ((rc = build_chunk_list(view->h_row_ptr, heavy, chunk, &view->oc_v, &view->oc_beg,
                               &view->oc_end, &view->num_oc)))
      return fail(rc);
  } else {
    view->oc_v = view->hc_v;
    view->oc_beg = view->hc_beg;
    view->oc_end = view->hc_end;
    view->num_oc = view->num_hc;
  }
  return DAWN_OK;
}

}  // namespace dawn
