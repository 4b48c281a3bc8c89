// Internal declarations shared by the host layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/qj_c_api.h"
#include "qj_types.h"

namespace qj {

inline size_t roots32_bytes(int n) { return (size_t)n * (sizeof(F4) + 2 * sizeof(int32_t)); }
// Packs nt top records followed by nb bottom records into the block layout.
void pack_roots32(const RootRec32* rec, int nt, int nb, std::vector<uint8_t>& blk);

// Division of a 32-bit unsigned integer by an invariant divisor d >= 1 with a
// multiply-high and two shifts (Granlund and Montgomery, "Division by
// invariant integers using multiplication", PLDI 1994, section 4):
// q = (t + ((n - t) >> s1)) >> s2, t = mulhi(m, n), with l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, s1 = min(l, 1), s2 = max(l - 1, 0).
struct FastDiv {
  uint32_t m, s1, s2;
};
inline FastDiv make_fastdiv(uint32_t d) {
  uint32_t l = 0;
  while (l < 32 && (1ull << l) < d) l++;
  const unsigned __int128 num = ((unsigned __int128)1 << 32) * (((unsigned __int128)1 << l) - d);
  return FastDiv{(uint32_t)(num / d + 1), l < 1 ? l : 1u, l > 1 ? l - 1 : 0u};
}
__host__ __device__ inline uint32_t fast_div(uint32_t n, const FastDiv& f) {
  const uint32_t t = (uint32_t)(((uint64_t)f.m * n) >> 32);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

// FIELD_3D geometry alone (marching cubes, weight-field lookups). div_x and
// div_xy divide point indices by X and X*Y (valid while X*Y < 2^32).
struct GridGeom {
  int32_t res[3];
  double lo[3];
  double dx[3];
  double center[3];
  double rot[9];
  double lengths[3];
  FastDiv div_x, div_xy;
};

}  // namespace qj

struct qj_rational {
  int32_t nt = 0, nb = 0;
  double sc_top = 1, sc_bottom = 1;
  qj::RootRec64* d64 = nullptr;
  void* d32 = nullptr;  // FP32 root block (roots32_view)
  std::vector<qj::RootRec64> h64;
  std::vector<qj::RootRec32> h32;
  std::vector<uint8_t> blk32;
  int device = 0;
};

struct qj_workspace {
  int device = 0;
  unsigned long long* d_counters = nullptr;  // CNT_N x kCounterSpread
  uint32_t* d_list = nullptr;                // band list (list_cap entries), then the adaptive list
  uint32_t* d_alist = nullptr;               // adaptive list (alist_cap entries)
  int64_t list_cap = 0, alist_cap = 0;
  uint32_t* d_list_count = nullptr;          // [0] band count, [1] adaptive count
  // marching cubes
  void* mc = nullptr;
  // gradient partials
  double* d_partials = nullptr;
  int64_t partials_cap = 0;
  // sampled FP64 cross-check of every field (qj_workspace_check); 0 = off
  int64_t check_samples = 0;
  // kernel timing (qj_workspace_timing)
  bool timing = false;
  std::vector<hipEvent_t> ev;  // pairs: bulk start/stop, band start/stop per launch
  std::mutex mu;
};

namespace qj {

void set_error(const char* fmt, ...);
qj_status hip_fail(hipError_t e, const char* what);
#define QJ_HIP(call)                                 \
  do {                                               \
    hipError_t _e = (call);                          \
    if (_e != hipSuccess) return qj::hip_fail(_e, #call); \
  } while (0)

void rotation_matrix(const double* q, double* m);
void make_geom(const qj_grid* g, GridGeom* out);
// Fill MapArgs from the C structs (host arithmetic mirrors FIELD_3D).
qj_status make_map_args(const qj_grid* g, const qj_rational* f, const qj_map_params* p, MapArgs* a);
double term_bits(const std::vector<RootRec64>& roots, double escape);
double near_bits(const std::vector<RootRec64>& roots);
qj_status ensure_list(qj_workspace* ws, int64_t n, bool adaptive, hipStream_t s);
qj_status ensure_partials(qj_workspace* ws, int64_t n, hipStream_t s);
// Records an event on s when timing is enabled (slot order: bulk0, bulk1, band0, band1).
void timing_mark(qj_workspace* ws, hipStream_t s);

// kernels (qj_field.hip)
qj_status launch_field(qj_workspace* ws, const MapArgs& a, const qj_rational* f, qj_precision prec, double band,
                       double* d_out, hipStream_t s);
qj_status launch_points(qj_workspace* ws, const MapArgs& a, const qj_rational* f, qj_precision prec,
                        double band, const double* d_pts, int64_t n, double* d_out, hipStream_t s);

Consts32 consts32(const MapArgs& a, double band);

// batched chunks for the request loop (qj_field.hip)
struct ChunkDesc {
  MapArgs a;
  const RootRec64* r64;
  const F4* r32;  // FP32 root block of the chunk's request (roots32_view)
  double* out;
  int64_t start;  // first voxel (linear, x fastest) of this chunk in its request
  int32_t count;  // voxels in this chunk (0 = unused slot)
  Consts32 k;     // FP32 constants of the chunk's request (consts32)
  int32_t total;  // chunks in use in this launch (read from the first descriptor)
  int32_t g1, g2, g3;  // variant groups (first descriptor): [0,g1) general, [g1,g2) RN4 integer, [g2,g3) RN2
                       // integer, [g3,total) RN8 integer
};
// The request loop's descriptors copied from pinned host memory to the
// device by a kernel (the first node of its graphs).
qj_status launch_desc_fetch(const ChunkDesc* h_desc, ChunkDesc* d_desc, int max_chunks, hipStream_t s);
// One field's sampled FP64 cross-check (the kernel of qj_workspace_check)
// into d_out[CNT_N] (CNT_CHECKED ... CNT_CHECK_MAXREL), cleared first.
qj_status launch_check_one(const MapArgs& a, const RootRec64* roots, const double* field, int64_t samples,
                           unsigned long long* d_out, hipStream_t s);
// FP32 variant of a chunk's root set: 0 general exponents; integer exponents
// renormalised every 4 terms (1), every 2 (2) or every 8 (3) (rn_of).
int batch_variant(const MapArgs& a);
// Enqueue one batch: FP32 bulk over every chunk then FP64 for the band (or
// FP64 for all when prec == QJ_FP64). d_band/d_band_count: scratch of
// nchunks * chunk_voxels entries; the count is cleared by a kernel, not a
// memset, so the sequence can be captured in a graph and replayed.
// d_counters: 8 uint32 (band counts and work queues), cleared at the start of
// every batch; d_band: nchunks * chunk_voxels entries. variants: bit v set when
// the batch holds chunks of FP32 variant v (kernels of absent variants are not
// enqueued; ignored for QJ_FP64).
qj_status launch_batch(const ChunkDesc* d_desc, int nchunks, int chunk_voxels, qj_precision prec, uint32_t* d_band,
                       uint32_t* d_counters, unsigned variants, hipStream_t s);

// marching cubes (qj_mc.hip)
qj_status mc_run(qj_workspace* ws, const double* d_field, const qj_grid* g, double iso, int64_t* nv,
                 int64_t* nt, hipStream_t s);
qj_status mc_run_tiles(qj_workspace* ws, int n, const double* const* fields, const qj_grid* grids, double iso,
                       int64_t* nv_out, int64_t* nt_out, hipStream_t s, double* ext_v = nullptr,
                       int64_t ext_vcap = 0, int32_t* ext_t = nullptr, int64_t ext_tcap = 0);
qj_status mc_copy(qj_workspace* ws, double* v, int32_t* t, hipStream_t s);
qj_status nlmc_cache(qj_workspace* ws, int64_t* n_edges, int32_t* edges, int64_t* n_cells, int32_t* cells,
                     hipStream_t s);
qj_status nlmc_resume(qj_workspace* ws, const MapArgs& a, const qj_rational* f, double iso, int64_t ne,
                      const int32_t* h_edges, int64_t nc, const int32_t* h_cells, int64_t* nv_out, int64_t* nt_out,
                      hipStream_t s);
qj_status nlmc_run(qj_workspace* ws, const MapArgs& a, const qj_rational* f, double iso, qj_precision prec,
                   double band, int chunk, int64_t* nv, int64_t* nt, int64_t* n_bad, hipStream_t s, const double* src_field = nullptr);
qj_status mc_ptrs(qj_workspace* ws, const double** v, const int32_t** t);
void mc_free(qj_workspace* ws);

// gradients (qj_grad.hip)
qj_status ap_grads(qj_workspace* ws, const MapArgs& a, const qj_rational* f, const qj_grid* wg,
                   const double* d_weight, double alpha, double thr, double* d_field, double* gt, double* gb,
                   qj_ap_grads* out, hipStream_t s);
qj_status points_adjoint(qj_workspace* ws, const MapArgs& a, const qj_rational* f, qj_precision prec,
                         const double* d_pts, int64_t n, const double* d_dJdF, double* d_gr, double* d_gp,
                         hipStream_t s);

// address of this GPU's copy of the allocation-free entry point's scratch
// array (qj_field.hip); `device` must be current
qj_status scratch_setup(int device, unsigned char** out);

// FIELD_3D clamp / border passes (qj_fieldops.hip)
qj_status field_clamp_to_mean(qj_workspace* ws, double* d_field, int64_t n, bool nans, double* mean_out,
                              hipStream_t s);
qj_status field_set_borders(const int32_t* res, double value, double* d_field, hipStream_t s);
// invertedDistance / invertedDistanceNormalized of a signed distance field
// (qj_fieldops.hip); abs_sum (host, optional): sum of |result|
qj_status field_inverted_distance(qj_workspace* ws, const double* d_dist, const qj_grid* grid, bool normalized,
                                  double* d_out, double* abs_sum, hipStream_t s);
// fieldMin / fieldMax with the serial scan's NaN and tie behaviour (host results)
qj_status field_min_max(qj_workspace* ws, const double* d_field, int64_t n, double* mn, double* mx, hipStream_t s);
// weightScore for each threshold (and the maximum score) of a field
qj_status weight_scores(qj_workspace* ws, const double* d_field, const qj_grid* grid, const qj_grid* wgrid,
                        const double* d_weight, double alpha, const double* thresholds, int nthr, double* scores,
                        double* max_score, hipStream_t s);

// exclusive scan of uint32 (qj_scan.hip); n elements, temp provided by ws
qj_status scan_u32(const uint32_t* d_in, uint32_t* d_out, int64_t n, uint32_t* d_total, void* d_tmp,
                   size_t tmp_bytes, hipStream_t s);
size_t scan_tmp_bytes(int64_t n);

}  // namespace qj
