// Device graph: upload of a host CSR, in-edge CSC for directed graphs,
// degree statistics; runtime helpers (errors, pointer kinds).
//
// Reference: the device copies of the CSR made per call by
// DAWN::MSSP_GPU::run (src/algorithm/gpu/mssp.cu:14-17) and passed by value
// to every per-source kernel (bfs.cuh:17-18).  Here the graph is uploaded
// once and stays resident until dawn_graph_destroy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace dawn {

int validate_csr(const dawn_csr* csr);
int build_csc(const dawn_csr* csr, std::vector<int32_t>& ip, std::vector<int32_t>& ic,
              std::vector<int32_t>& iw, int32_t* max_in_degree);

static thread_local std::string g_last_error;

void set_hip_error(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  g_last_error = buf;
  if (std::getenv("DAWN_VERBOSE")) std::fprintf(stderr, "dawn: %s\n", buf);
}

const char* last_error() { return g_last_error.c_str(); }

void set_error(const char* text) {
  g_last_error = text;
  if (std::getenv("DAWN_VERBOSE")) std::fprintf(stderr, "dawn: %s\n", text);
}

bool is_device_pointer(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t attr;
  hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // clear the sticky error from an unknown host pointer
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

int upload_i32(int32_t** dst, const std::vector<int32_t>& src) {
  hipError_t e = hipMalloc((void**)dst, sizeof(int32_t) * std::max<size_t>(src.size(), 1));
  if (e != hipSuccess) {
    *dst = nullptr;
    set_hip_error(e, "hipMalloc", __FILE__, __LINE__);
    return e == hipErrorOutOfMemory ? DAWN_ERR_OUT_OF_MEMORY : DAWN_ERR_HIP;
  }
  if (!src.empty()) DAWN_HIP(hipMemcpy(*dst, src.data(), sizeof(int32_t) * src.size(), hipMemcpyHostToDevice));
  return DAWN_OK;
}

namespace {
void free_chunks(dawn_graph* g) {
  if (g->oc_v && g->oc_v != g->hc_v) {
    (void)hipFree(g->oc_v);
    (void)hipFree(g->oc_beg);
    (void)hipFree(g->oc_end);
  }
  if (g->hc_v) (void)hipFree(g->hc_v);
  if (g->hc_beg) (void)hipFree(g->hc_beg);
  if (g->hc_end) (void)hipFree(g->hc_end);
  g->hc_v = g->hc_beg = g->hc_end = g->oc_v = g->oc_beg = g->oc_end = nullptr;
  g->num_hc = g->num_oc = 0;
}
}  // namespace

// Chunk list of the rows of `rp` with at least `heavy` entries, cut into
// pieces of at most `chunk`: (row, first, end) triples on the device.
int build_chunk_list(const std::vector<int32_t>& rp, int heavy, int chunk, int32_t** dv,
                     int32_t** db, int32_t** de, int64_t* count) {
  std::vector<int32_t> hv, hb, he;
  const int32_t n = (int32_t)rp.size() - 1;
  for (int32_t v = 0; v < n; v++) {
    const int32_t b = rp[v], e = rp[v + 1];
    if (e - b < heavy) continue;
    for (int32_t k = b; k < e; k += chunk) {
      hv.push_back(v);
      hb.push_back(k);
      he.push_back(std::min(e, k + chunk));
    }
  }
  *count = (int64_t)hv.size();
  DAWN_TRY(upload_i32(dv, hv));
  DAWN_TRY(upload_i32(db, hb));
  return upload_i32(de, he);
}

// Heavy in-edge chunks for the pull step (dawn_graph::hc_v) and heavy
// out-edge chunks for the dense top-down step, from g->knobs.
int build_chunks(dawn_graph* g) {
  free_chunks(g);
  const int heavy = g->knobs.heavy_degree, chunk = g->knobs.pull_chunk;
  DAWN_TRY(build_chunk_list(g->directed ? g->h_in_row_ptr : g->h_row_ptr, heavy, chunk, &g->hc_v,
                            &g->hc_beg, &g->hc_end, &g->num_hc));
  if (g->directed) {
    DAWN_TRY(build_chunk_list(g->h_row_ptr, heavy, chunk, &g->oc_v, &g->oc_beg, &g->oc_end,
                              &g->num_oc));
  } else {
    g->oc_v = g->hc_v;
    g->oc_beg = g->hc_beg;
    g->oc_end = g->hc_end;
    g->num_oc = g->num_hc;
  }
  return DAWN_OK;
}

}  // namespace dawn

using namespace dawn;

extern "C" {

const char* dawn_last_error(void) { return dawn::last_error(); }

int dawn_device_count(int32_t* count) {
  if (!count) return DAWN_ERR_INVALID_ARG;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    n = 0;
  }
  *count = n;
  return DAWN_OK;
}

int dawn_graph_create(const dawn_csr* csr, dawn_graph** out) {
  if (!csr || !out || csr->n < 0 || csr->m < 0 || !csr->row_ptr ||
      (csr->m > 0 && !csr->col))
    return DAWN_ERR_INVALID_ARG;
  if (csr->m > (int64_t)INT32_MAX) return DAWN_ERR_OVERFLOW;
  const int32_t n = csr->n;
  const int64_t m = csr->m;
  if (csr->row_ptr[0] != 0 || (int64_t)csr->row_ptr[n] != m) return DAWN_ERR_INVALID_ARG;
  DAWN_TRY(validate_csr(csr));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return DAWN_ERR_NO_DEVICE;
  }

  dawn_graph* g = new dawn_graph();
  g->n = n;
  g->m = m;
  g->directed = csr->directed ? 1 : 0;
  g->weighted = csr->weight ? 1 : 0;
  DAWN_HIP(hipGetDevice(&g->device));
  g->h_row_ptr.assign(csr->row_ptr, csr->row_ptr + n + 1);
  int32_t maxd = 0;
  for (int32_t i = 0; i < n; i++) maxd = std::max(maxd, csr->row_ptr[i + 1] - csr->row_ptr[i]);
  g->max_out_degree = maxd;

  auto fail = [&](int rc) {
    dawn_graph_destroy(g);
    return rc;
  };
  auto up = [&](int32_t** dst, const int32_t* src, size_t count) -> int {
    hipError_t e = hipMalloc((void**)dst, sizeof(int32_t) * std::max<size_t>(count, 1));
    if (e != hipSuccess) {
      *dst = nullptr;
      set_hip_error(e, "hipMalloc", __FILE__, __LINE__);
      return e == hipErrorOutOfMemory ? DAWN_ERR_OUT_OF_MEMORY : DAWN_ERR_HIP;
    }
    if (count) {
      e = hipMemcpy(*dst, src, sizeof(int32_t) * count, hipMemcpyHostToDevice);
      if (e != hipSuccess) {
        set_hip_error(e, "hipMemcpy", __FILE__, __LINE__);
        return DAWN_ERR_HIP;
      }
    }
    return DAWN_OK;
  };
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(DAWN_ERR_HIP);
  int rc;
  std::vector<int32_t> in_rp_host;  // CSC offsets (directed graphs)
  if ((rc = up(&g->row_ptr, csr->row_ptr, (size_t)n + 1))) return fail(rc);
  if ((rc = up(&g->col, csr->col, (size_t)m))) return fail(rc);
  if (csr->weight && (rc = up(&g->weight, csr->weight, (size_t)m))) return fail(rc);

  if (!g->directed) {
    g->in_row_ptr = g->row_ptr;
    g->in_col = g->col;
    g->in_weight = g->weight;
    g->max_in_degree = maxd;
  } else {
    // In-edge CSC on the host (rows ascending within each column).
    std::vector<int32_t> ip, ic, iw;
    int32_t maxi = 0;
    if ((rc = build_csc(csr, ip, ic, iw, &maxi))) return fail(rc);
    g->max_in_degree = maxi;
    in_rp_host = ip;
    if ((rc = up(&g->in_row_ptr, ip.data(), (size_t)n + 1))) return fail(rc);
    if ((rc = up(&g->in_col, ic.data(), (size_t)m))) return fail(rc);
    if (csr->weight && (rc = up(&g->in_weight, iw.data(), (size_t)m))) return fail(rc);
  }
  if (g->directed) g->h_in_row_ptr = in_rp_host;
  {
    const std::vector<int32_t>& ip = g->directed ? g->h_in_row_ptr : g->h_row_ptr;
    int64_t z = 0;
    for (int32_t v = 0; v < n; v++) z += ip[v + 1] == ip[v];
    g->zero_in = z;
  }
  (void)tuned_knobs_for(g, &g->knobs);
  if ((rc = build_chunks(g))) return fail(rc);
  *out = g;
  return DAWN_OK;
}

int dawn_graph_destroy(dawn_graph* g) {
  if (!g) return DAWN_OK;
  (void)hipSetDevice(g->device);
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  if (g->stream2) (void)hipStreamSynchronize(g->stream2);
  ms_workspace_free(g->ms);
  ms_workspace_free(g->ms2);
  ms_workspace_free(g->ms_tail);
  ss_workspace_free(g->ss);
  wms_workspace_free(g->wms);
  ms_view_free(g);
  if (g->directed) {
    if (g->in_row_ptr) (void)hipFree(g->in_row_ptr);
    if (g->in_col) (void)hipFree(g->in_col);
    if (g->in_weight) (void)hipFree(g->in_weight);
  }
  free_chunks(g);
  if (g->row_ptr) (void)hipFree(g->row_ptr);
  if (g->col) (void)hipFree(g->col);
  if (g->weight) (void)hipFree(g->weight);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  if (g->stream2) (void)hipStreamDestroy(g->stream2);
  delete g;
  return DAWN_OK;
}

int dawn_graph_synchronize(const dawn_graph* g) {
  if (!g) return DAWN_ERR_INVALID_ARG;
  DAWN_HIP(hipSetDevice(g->device));
  if (g->stream) DAWN_HIP(hipStreamSynchronize(g->stream));
  return DAWN_OK;
}

int dawn_graph_info(const dawn_graph* g, int32_t* n, int64_t* m, int32_t* directed,
                    int32_t* weighted) {
  if (!g) return DAWN_ERR_INVALID_ARG;
  if (n) *n = g->n;
  if (m) *m = g->m;
  if (directed) *directed = g->directed;
  if (weighted) *weighted = g->weighted;
  return DAWN_OK;
}

}  // extern "C"
