// Marching cubes on a resident field (TRIANGLE_MESH(const FIELD_3D&) ->
// computeStagedMarchingCubes, TRIANGLE_MESH.cpp:989-1114), bit-exact with the
// CPU mesher: same flags, same vertex set and numbering, same vertex
// arithmetic, same triangle order, same degenerate filter.
//
// Pipeline per field (one lane per grid point / cell, point blocks dealt to
// the XCDs in contiguous runs; no host synchronisation between fields, see
// mc_run_tiles):
//   k_mc_cells      cell case from the float sign test (the 8 corners read
//                   once), and which of the case row's triangles survive the
//                   degenerate filter (edge < 1e-7 dropped): an FP32 estimate
//                   of where each vertex sits on its edge clears every
//                   triangle with fewer than two vertices near an edge end,
//                   the other cells go to a list; per block of cells:
//                   kept-triangle count
//   k_mc_measure    the listed cells' triangles measured with the FP64
//                   vertices, one cell per lane
//   k_mc_edges      per grid point: which of its +x,+y,+z edges some adjacent
//                   cell's case row references (the edge-use masks of the four
//                   cell rows around a block staged in LDS), packed with the
//                   point's first vertex id within its block; per block of
//                   points: vertex count
//   scan            block vertex counts -> first vertex id per block (the
//                   reference numbers vertices by std::map<pair<int,int>> order
//                   = lower endpoint, then x<y<z)
//   k_mc_vertices   ids within the block, p0 + d0/(d0-d1) (p1-p0) at cell
//                   centres, no contraction
//   scan            block triangle counts -> first triangle per block
//   k_mc_tris       each kept triangle written once, in cell then table order
#include <float.h>
#include <math.h>

#include <algorithm>
#include <vector>

#include "qj_eval.h"
#include "qj_geom.h"
#include "qj_internal.h"
#include "qj_mc_table.h"

namespace qj {

constexpr int kMcBlock = 256;

// edge id -> lower endpoint offset (dx,dy,dz) and axis
constexpr void edge_parts(int e, int& dx, int& dy, int& dz, int& axis) {
  if (e < 4) {
    axis = 0; dx = 0; dy = e & 1; dz = e >> 1;
  } else if (e < 8) {
    axis = 1; dx = e & 1; dy = 0; dz = (e >> 1) & 1;
  } else {
    axis = 2; dx = e & 1; dy = (e >> 1) & 1; dz = 0;
  }
}

struct McTables {
  signed char tri[256][16];
  unsigned char ntri[256];
  unsigned short emask[256];
  // edge mask | triangle count << 12
  unsigned short emn[256];
  // per case row and triangle: the bit mask of its 3 edges, and its 3 corners
  // in table order, 5 bits each: dx | (dy + 2 dz) << 1 | axis << 3 (the row
  // of owner points and the axis the triangle kernel looks up)
  unsigned short tmask[256][5];
  unsigned short tcorner[256][5];
};

constexpr McTables make_tables() {
  McTables t{QJ_MC_TRI_INIT, {}, {}, {}, {}, {}};
  for (int c = 0; c < 256; c++) {
    int n = 0;
    unsigned m = 0;
    for (int k = 0; k < 16 && t.tri[c][k] >= 0; k++) {
      n++;
      m |= 1u << t.tri[c][k];
    }
    t.ntri[c] = (unsigned char)(n / 3);
    t.emask[c] = (unsigned short)m;
    t.emn[c] = (unsigned short)(m | (unsigned)(n / 3) << 12);
    for (int q = 0; q < n / 3; q++) {
      unsigned tm = 0, tc = 0;
      for (int j = 0; j < 3; j++) {
        const int e = t.tri[c][3 * q + j];
        int dx = 0, dy = 0, dz = 0, axis = 0;
        edge_parts(e, dx, dy, dz, axis);
        tm |= 1u << e;
        tc |= (unsigned)(dx | (dy + 2 * dz) << 1 | axis << 3) << (5 * j);
      }
      t.tmask[c][q] = (unsigned short)tm;
      t.tcorner[c][q] = (unsigned short)tc;
    }
  }
  return t;
}

__constant__ McTables kMc = make_tables();

// Case tables staged in LDS for the kernels that look them up per lane: with
// divergent indices, reads of __constant__ memory are scattered vector loads
// (12 per grid point in the edge kernels); from LDS they are ds_reads.
static_assert(kMcBlock == 256, "one table entry per thread");

__device__ __forceinline__ void stage_emask(unsigned short* s_em) {
  s_em[threadIdx.x] = kMc.emask[threadIdx.x];
  __syncthreads();
}

// s_tri: the 256 x 16 triangle table as 1024 words; s_ntri: triangles per case
__device__ __forceinline__ void stage_tri(uint32_t* s_tri, uint8_t* s_ntri) {
  const uint32_t* src = (const uint32_t*)&kMc.tri[0][0];
  for (int i = threadIdx.x; i < 1024; i += kMcBlock) s_tri[i] = src[i];
  s_ntri[threadIdx.x] = kMc.ntri[threadIdx.x];
  __syncthreads();
}

__device__ __forceinline__ int tri_edge(const uint32_t* s_tri, int flag, int k) {
  return ((const signed char*)s_tri)[flag * 16 + k];
}

__device__ __forceinline__ void edge_geom(int e, int& dx, int& dy, int& dz, int& axis) {
  edge_parts(e, dx, dy, dz, axis);
}

// per-triangle tables (5 per case row) staged in LDS as 640 words
__device__ __forceinline__ void stage_tri5(unsigned short* dst, const unsigned short (*src)[5]) {
  const uint32_t* s32 = (const uint32_t*)&src[0][0];
  uint32_t* d32 = (uint32_t*)dst;
  for (int i = threadIdx.x; i < 640; i += kMcBlock) d32[i] = s32[i];
}

__device__ __forceinline__ double sq3(double e0, double e1, double e2) {
#pragma clang fp contract(off)
  return e0 * e0 + e1 * e1 + e2 * e2;
}

struct McState {
  int64_t cap_flags = 0, cap_emask = 0, cap_cnt = 0, cap_active = 0, cap_toff = 0, cap_blk = 0;
  int64_t cap_v = 0, cap_tp = 0, cap_keep = 0, cap_t = 0;
  uint8_t* flags = nullptr;
  uint8_t* emask = nullptr;
  uint32_t* cnt = nullptr;     // per point edge count -> vertex base (scanned in place; non-linear mesher)
  uint16_t* pk16 = nullptr;    // linear mesher, per point: first vertex id within its block << 3 | edge mask
  uint8_t* kmask = nullptr;    // per cell: bit t set when triangle t of its case row is kept
  int64_t cap_kmask = 0;
  uint32_t* mlist = nullptr;   // cells that await the degenerate measurement
  int64_t cap_mlist = 0;
  uint32_t* d_mcount = nullptr;
  uint32_t* blkt = nullptr;    // per block of cells triangle count -> first triangle of the block
  int64_t cap_blkt = 0;
  uint32_t* blkv = nullptr;    // per block vertex count -> first vertex id of the block
  int64_t cap_pk16 = 0, cap_blkv = 0;
  uint32_t* blk = nullptr;     // per block active count -> offset
  uint32_t* active = nullptr;  // active cells
  uint32_t* tri_off = nullptr; // per active cell triangle count -> offset
  double* verts = nullptr;
  int32_t* tris_pre = nullptr;
  uint32_t* keep = nullptr;    // per pre-filter triangle keep -> offset
  int32_t* tris = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  uint32_t* totals = nullptr;  // [0] active [1] vertices [2] triangles pre-filter [3] triangles
  int64_t nv = 0, nt = 0;
  // non-linear mesher: two slabs of field values (chunk i's in slab[i % 2]; the
  // next chunk's is evaluated on stream s2 while this one is meshed), slice
  // offsets, error counter
  double* slab[2] = {nullptr, nullptr};
  int64_t cap_slab[2] = {0, 0};
  hipStream_t s2 = nullptr;
  hipEvent_t ev_start = nullptr, ev_field[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
  uint32_t* slice_off = nullptr;
  int64_t cap_slice = 0;
  uint32_t* vpt = nullptr;  // vertices of a range of points: owning point, edge axis
  uint8_t* vax = nullptr;
  int64_t cap_vpt = 0, cap_vax = 0;
  unsigned int* d_bad = nullptr;
  // edge cache of the last non-linear mesh (TRIANGLE_MESH.cpp:431-586): per
  // vertex its edge (lower endpoint x, y, z, axis), per active cell (x, y, z, flag)
  int32_t* c_edges = nullptr;
  int32_t* c_cells = nullptr;
  int64_t cap_ce = 0, cap_cc = 0, ncells = 0;
  bool cache_valid = false;
  // linear mesher over several fields: running vertex / triangle totals and
  // per-field counts on the device
  uint64_t* d_base = nullptr;
  int64_t* d_counts = nullptr;
  int64_t cap_counts = 0;
};

__device__ __forceinline__ int cell_flag(const double* __restrict__ f, int64_t i, int64_t X, int64_t XY, double iso) {
  // CUBE corners are floats (TRIANGLE_MESH.h:283-288); test at TRIANGLE_MESH.cpp:1037-1039
  int flag = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int dx = k >> 2, dy = (k >> 1) & 1, dz = k & 1;
    const float v = (float)(f[i + dx + dy * X + dz * XY] - iso);
    flag |= (v > 0.0f) << k;
  }
  return flag;
}

// grid coordinates of a point index (the mesher runs on fields of < 2^32 points)
__device__ __forceinline__ void point_xyz(const GridGeom& g, int64_t p, int64_t& x, int64_t& y, int64_t& z) {
  const uint32_t pu = (uint32_t)p;
  const uint32_t zu = fast_div(pu, g.div_xy);
  const uint32_t r = pu - zu * (uint32_t)g.res[0] * (uint32_t)g.res[1];
  const uint32_t yu = fast_div(r, g.div_x);
  z = zu;
  y = yu;
  x = r - yu * (uint32_t)g.res[0];
}

// Workgroups go to the 8 XCDs round-robin (block b to XCD b % 8); the point
// kernels deal each XCD a contiguous run of point blocks instead, so the
// field rows a block shares with its neighbours are fetched into one L2.
__device__ __forceinline__ int64_t xcd_block(int64_t nb) {
  const int64_t b = blockIdx.x, x = b & 7, k = b >> 3, per = nb >> 3, rem = nb & 7;
  return x * per + (x < rem ? x : rem) + k;
}

// Block-wide exclusive prefix and total of a small per-thread count given as
// three bit flags of weights W0, W1, W2 (edge bits 1,1,1; a triangle count
// 0..5 in binary 1,2,4): three ballots and masked popcounts, no shuffles.
struct BlockPrefix {
  uint32_t excl;
  uint32_t total;
};
template <int W0, int W1, int W2>
__device__ __forceinline__ BlockPrefix mc_block_prefix3(unsigned v) {
  __shared__ uint32_t s_w[kMcBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const unsigned long long b0 = __ballot(v & 1u), b1 = __ballot(v & 2u), b2 = __ballot(v & 4u);
  const uint32_t within = W0 * __popcll(b0 & lt) + W1 * __popcll(b1 & lt) + W2 * __popcll(b2 & lt);
  if (lane == 0) s_w[wave] = W0 * __popcll(b0) + W1 * __popcll(b1) + W2 * __popcll(b2);
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (int i = 0; i < kMcBlock / 64; i++) {
    const uint32_t c = s_w[i];
    if (i < wave) before += c;
    total += c;
  }
  return BlockPrefix{before + within, total};
}

// ---- linear mesher: one wave per block of 256 points (the unit of the
// vertex and triangle numbering), four blocks per workgroup. Lane l of a wave
// takes points l, l + 64, l + 128 and l + 192 of its block: every load
// instruction is contiguous over the wave, each lane has four independent
// points in flight, and prefixes in point order are wave ballots with no
// barrier.
constexpr int kMcWaves = 4;              // blocks per workgroup
constexpr int kMcGroup = 64 * kMcWaves;  // threads per workgroup
static_assert(kMcGroup == kMcBlock, "the table stagings use one thread per case");

struct WavePrefix {
  uint32_t excl;
  uint32_t total;
};
// exclusive prefix over the wave's lanes and wave total of a count 0..7 given
// as three bits of weights W0, W1, W2
template <int W0, int W1, int W2>
__device__ __forceinline__ WavePrefix wave_prefix3(unsigned v) {
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (1ull << lane) - 1ull;
  const unsigned long long b0 = __ballot(v & 1u), b1 = __ballot(v & 2u), b2 = __ballot(v & 4u);
  return WavePrefix{(uint32_t)(W0 * __popcll(b0 & lt) + W1 * __popcll(b1 & lt) + W2 * __popcll(b2 & lt)),
                    (uint32_t)(W0 * __popcll(b0) + W1 * __popcll(b1) + W2 * __popcll(b2))};
}

// this wave's block of points (blocks of a workgroup are consecutive), or -1
// past the last block
__device__ __forceinline__ int64_t wave_block(int64_t nb) {
  const int64_t B = xcd_block((nb + kMcWaves - 1) / kMcWaves) * kMcWaves + (threadIdx.x >> 6);
  return B < nb ? B : -1;
}

// Edge masks (bit axis set when the +axis edge from the point carries a
// vertex) and each point's first vertex id within its block, packed as
// pk16[p] = id << 3 | mask (ids within a block stay below 3 x 256), and each
// block's vertex count. The edges of point p are referenced by the cells
// p - a X - b XY (x edge), p - a - b XY (y edge) and p - a - b X (z edge),
// a, b in {0, 1}: entries -1..255 of the four cell rows that start at
// p0 - a X - b XY. Each wave stages those rows' edge-use masks (the case
// table's emask of each cell's flag) in LDS, so a point's 12 tests are LDS
// reads and each flag is looked up once.
__global__ __launch_bounds__(kMcGroup) void k_mc_edges(const uint8_t* __restrict__ flags, GridGeom g,
                                                       uint16_t* __restrict__ pk16, uint32_t* __restrict__ blkv) {
  __shared__ unsigned short s_em[256];
  __shared__ unsigned short s_ce[kMcWaves][4][kMcBlock + 1];
  s_em[threadIdx.x] = kMc.emask[threadIdx.x];
  __syncthreads();
  const int64_t X = g.res[0], Y = g.res[1], Z = g.res[2], XY = X * Y, N = XY * Z;
  const int64_t B = wave_block((N + kMcBlock - 1) / kMcBlock);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (B >= 0) {
    // all 20 loads issued before the first table lookup (case 0 uses no edge)
    const int64_t p0 = B * kMcBlock;
    unsigned f[4][5];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int64_t c0 = p0 - (r & 1) * X - (r >> 1) * XY - 1;
#pragma unroll
      for (int k = 0; k < 5; k++) {
        const int64_t c = c0 + lane + 64 * k;
        f[r][k] = ((k < 4 || lane == 0) && c >= 0 && c < N) ? (unsigned)flags[c] : 0u;
      }
    }
#pragma unroll
    for (int r = 0; r < 4; r++)
#pragma unroll
      for (int k = 0; k < 5; k++)
        if (k < 4 || lane == 0) s_ce[w][r][lane + 64 * k] = s_em[f[r][k]];
  }
  __syncthreads();
  if (B < 0) return;
  uint32_t base = 0;
  for (int sub = 0; sub < 4; sub++) {
    const int j = sub * 64 + lane;
    const int64_t p = B * kMcBlock + j;
    unsigned bits = 0;
    if (p < N) {
      int64_t x, y, z;
      point_xyz(g, p, x, y, z);
      // x-edge (p, p+1): cells p - (0,a,b), row a + 2b, local edge a + 2b
      if (x < X - 1) {
        unsigned ref = 0;
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++)
            if (y >= a && z >= b) ref |= (s_ce[w][a + 2 * b][j + 1] >> (a + 2 * b)) & 1u;
        bits |= ref;
      }
      // y-edge (p, p+X): cells p - (a,0,b), row 2b, local edge 4 + a + 2b
      if (y < Y - 1) {
        unsigned ref = 0;
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++)
            if (x >= a && z >= b) ref |= (s_ce[w][2 * b][j + 1 - a] >> (4 + a + 2 * b)) & 1u;
        bits |= ref << 1;
      }
      // z-edge (p, p+XY): cells p - (a,b,0), row b, local edge 8 + a + 2b
      if (z < Z - 1) {
        unsigned ref = 0;
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++)
            if (x >= a && y >= b) ref |= (s_ce[w][b][j + 1 - a] >> (8 + a + 2 * b)) & 1u;
        bits |= ref << 2;
      }
    }
    const WavePrefix vp = wave_prefix3<1, 1, 1>(bits);
    if (p < N) pk16[p] = (uint16_t)((base + vp.excl) << 3 | bits);
    base += vp.total;
  }
  if (lane == 0) blkv[B] = base;
}

// One vertex of computeEdgeInterpolations (TRIANGLE_MESH.cpp:2373-2425): the
// edge from point (x, y, z) along `axis`, p0 + d0/(d0-d1) (p1-p0) between the
// two cell centres, in the reference's operation order with no contraction.
// k_mc_vertices writes it; k_mc_cells measures triangles with it, so both
// see the same bits.
__device__ __forceinline__ void mc_vertex(const GridGeom& g, int x, int y, int z, int axis, double f0, double f1,
                                          double iso, double* v) {
#pragma clang fp contract(off)
  double a[3], b[3];
  cell_center(g, x, y, z, a);
  cell_center(g, x + (axis == 0), y + (axis == 1), z + (axis == 2), b);
  const double d0 = f0 - iso;
  const double off0 = b[0] - a[0], off1 = b[1] - a[1], off2 = b[2] - a[2];
  const double d1 = f1 - iso;
  const double interpolation = d0 / (d0 - d1);
  v[0] = a[0] + interpolation * off0;
  v[1] = a[1] + interpolation * off1;
  v[2] = a[2] + interpolation * off2;
}

__device__ __forceinline__ double edge_len(const double* a, const double* b) {
#pragma clang fp contract(off)
  const double e0 = a[0] - b[0], e1 = a[1] - b[1], e2 = a[2] - b[2];
  return sqrt(e0 * e0 + e1 * e1 + e2 * e2);  // norm(v0 - v1), VEC3.h:215-216
}

// The degenerate filter (TRIANGLE_MESH.cpp:1088-1102) drops triangles with a
// side shorter than 1e-7. Two vertices of one triangle lie on edges of one
// cell: parallel or skew edges are at least one grid spacing h apart, and on
// two edges meeting at a corner c the vertices satisfy |v1 - v2| >=
// max |vi - c| (the edges are perpendicular). So with h >= kPrefilterSpacing
// a triangle with at most one vertex within 2e-7 of an end of its edge has
// every side longer than 1e-7. k_mc_cells marks a vertex "near" from an FP32
// estimate t' of its place t on the edge: near unless t' h >= kNearLimit and
// (1 - t') h >= kNearLimit. |t' - t| stays below 2.5e-7 (corner values
// rounded to float, one division and one subtraction), so with
// h <= kPrefilterMaxSpacing every vertex within 2e-7 of an end is marked
// (2e-7 + 2.5e-7 h < 1e-6); NaN or infinite corners mark it too. Triangles
// with two or three marked vertices are measured with the FP64 vertices.
constexpr float kNearLimit = 1e-6f;
constexpr double kPrefilterSpacing = 1e-6;
constexpr double kPrefilterMaxSpacing = 3.0;

// The 12-bit mask of a cell's edges whose vertex may lie within 2e-7 of an
// end of its edge, from the float corner values (corner k = 4 dx + 2 dy + dz):
// near unless |v0| h >= 1e-6 |v0 - v1| and |v1| h >= 1e-6 |v0 - v1| (a vertex
// at place t has t = |v0| / |v0 - v1| and 1 - t = |v1| / |v0 - v1|, so no
// division); NaN and infinite corners make it near. prefilter = 0: all 12.
__device__ __forceinline__ unsigned mc_near_edges(const float* v, const GridGeom& g, int prefilter) {
  if (!prefilter) return 0xfffu;
  const float h[3] = {fabsf((float)g.dx[0]), fabsf((float)g.dx[1]), fabsf((float)g.dx[2])};
  unsigned near = 0;
#pragma unroll
  for (int e = 0; e < 12; e++) {
    int dx = 0, dy = 0, dz = 0, axis = 0;
    edge_parts(e, dx, dy, dz, axis);
    const int k0 = 4 * dx + 2 * dy + dz, k1 = k0 + (axis == 0 ? 4 : (axis == 1 ? 2 : 1));
    const float a0 = fabsf(v[k0]), a1 = fabsf(v[k1]), lim = kNearLimit * fabsf(v[k0] - v[k1]);
    near |= (a0 * h[axis] >= lim && a1 * h[axis] >= lim) ? 0u : (1u << e);
  }
  return near;
}

// The 8 corner values of the cell at point p as floats (CUBE corners are
// floats, TRIANGLE_MESH.h:283-288; sign test at TRIANGLE_MESH.cpp:1037-1039),
// and its case.
__device__ __forceinline__ int mc_corners(const double* __restrict__ field, int64_t p, int64_t X, int64_t XY,
                                          double iso, float* v) {
  int flag = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int dx = k >> 2, dy = (k >> 1) & 1, dz = k & 1;
    v[k] = (float)(field[p + dx + dy * X + dz * XY] - iso);
    flag |= (v[k] > 0.0f) << k;
  }
  return flag;
}

// Bit 7 of kmask: the cell awaits measurement and did not fit the list.
constexpr unsigned kMeasureMark = 0x80u;
// The measure list is split in kMeasureLists sub-lists of equal capacity, one
// counter each (block lb appends to sub-list lb % kMeasureLists): one counter
// for all blocks serialised the appends (65536 atomics on one address per
// 256^3 field). Counter kMeasureLists is set when a cell did not fit.
constexpr int kMeasureLists = 251;

// Cell case and provisional kept triangles of every cell. flags[p] is the
// case of the cell with lowest corner p (0 on the last layer of each axis),
// kmask[p] bit t is set when triangle t of that case row is kept, blkt[block]
// counts the block's kept triangles. A cell with two or more near vertices
// may hold a degenerate triangle: it keeps all its triangles here and goes to
// the measure list (mlist: kMeasureLists sub-lists of mcap entries; beyond
// that its kmask gets kMeasureMark), which k_mc_measure works through with
// full waves. The FP64 measurement is rare but long, and one lane of a wave
// taking it held up the whole wave here (the borders of a field set by
// setBorders put such cells in every row). One cell per thread, a block of
// 256 cells per block of the numbering.
__global__ __launch_bounds__(kMcBlock) void k_mc_cells(const double* __restrict__ field, GridGeom g, double iso,
                                                       int prefilter, uint8_t* __restrict__ flags,
                                                       uint8_t* __restrict__ kmask, uint32_t* __restrict__ blkt,
                                                       uint32_t* __restrict__ mlist, uint32_t* __restrict__ mcount,
                                                       uint32_t mcap) {
  __shared__ unsigned short s_emn[256];
  __shared__ uint32_t s_w[kMcBlock / 64], s_m[kMcBlock / 64];
  __shared__ uint32_t s_mbase;
  s_emn[threadIdx.x] = kMc.emn[threadIdx.x];
  __syncthreads();
  const int64_t X = g.res[0], Y = g.res[1], Z = g.res[2], XY = X * Y, N = XY * Z;
  const int64_t lb = xcd_block((N + kMcBlock - 1) / kMcBlock);
  const int64_t p = lb * kMcBlock + threadIdx.x;
  int flag = 0;
  unsigned keep = 0;
  bool measure = false;
  if (p < N) {
    int64_t x, y, z;
    point_xyz(g, p, x, y, z);
    if (x < X - 1 && y < Y - 1 && z < Z - 1) {
      float v[8];
      flag = mc_corners(field, p, X, XY, iso, v);
      const unsigned emn = s_emn[flag];
      const int n = (int)(emn >> 12);
      if (n) {
        keep = (1u << n) - 1u;
        measure = __popc(mc_near_edges(v, g, prefilter) & emn & 0xfffu) >= 2;
      }
    }
  }
  // kept-triangle count of the block and this block's share of the measure list
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned cnt = (unsigned)__popc(keep);
  const unsigned long long b0 = __ballot(cnt & 1u), b1 = __ballot(cnt & 2u), b2 = __ballot(cnt & 4u);
  const unsigned long long bm = __ballot(measure);
  if (lane == 0) {
    s_w[wave] = (uint32_t)(__popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2));
    s_m[wave] = (uint32_t)__popcll(bm);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0, m = 0;
    for (int i = 0; i < kMcBlock / 64; i++) {
      t += s_w[i];
      const uint32_t c = s_m[i];
      s_m[i] = m;
      m += c;
    }
    blkt[lb] = t;
    const int sub = (int)(lb % kMeasureLists);
    const uint32_t base = m ? atomicAdd(&mcount[sub], m) : 0u;
    s_mbase = base;
    if (m && base + m > mcap) mcount[kMeasureLists] = 1u;  // some cells did not fit
  }
  __syncthreads();
  if (p < N) {
    const uint32_t slot = s_mbase + s_m[wave] + (uint32_t)__popcll(bm & ((1ull << lane) - 1ull));
    const bool listed = measure && slot < mcap;
    if (listed) mlist[(int64_t)(lb % kMeasureLists) * mcap + slot] = (uint32_t)p;
    flags[p] = (uint8_t)flag;
    kmask[p] = (uint8_t)(keep | (measure && !listed ? kMeasureMark : 0u));
  }
}

// The degenerate filter (TRIANGLE_MESH.cpp:1088-1102) for one listed cell:
// triangles with two or three near vertices are measured with the FP64
// vertices (the function the vertex kernel uses, so both see the same bits).
// Only a side between two near vertices can be short: a vertex that is not
// near is at least 2e-7 from both ends of its edge, so at least that far from
// a vertex on an edge through either end, and a grid spacing from vertices on
// the other edges. Dropped triangles leave kmask and the block's count.
__device__ __forceinline__ void mc_measure_cell(const double* __restrict__ field, const GridGeom& g, double iso,
                                                int prefilter, int64_t p, uint8_t* __restrict__ kmask,
                                                uint32_t* __restrict__ blkt) {
  const int64_t X = g.res[0], XY = X * g.res[1];
  int64_t x, y, z;
  point_xyz(g, p, x, y, z);
  float v[8];
  const int flag = mc_corners(field, p, X, XY, iso, v);
  const unsigned near = mc_near_edges(v, g, prefilter) & kMc.emask[flag];
  const int n = kMc.ntri[flag];
  unsigned keep = kmask[p] & 0x1fu, dropped = 0;
  for (int t = 0; t < n; t++) {
    const unsigned nm = near & kMc.tmask[flag][t];
    const int m = __popc(nm);
    if (m < 2) continue;
    double w[3][3] = {};
    unsigned r = nm;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (j < m) {  // constant indices into w keep it in registers
        const int e = __ffs(r) - 1;
        r &= r - 1;
        int dx, dy, dz, axis;
        edge_geom(e, dx, dy, dz, axis);
        const int64_t q0 = p + dx + dy * X + dz * XY;
        const int64_t q1 = q0 + (axis == 0 ? 1 : (axis == 1 ? X : XY));
        mc_vertex(g, (int)x + dx, (int)y + dy, (int)z + dz, axis, field[q0], field[q1], iso, w[j]);
      }
    }
    const double eps = 1e-7;
    bool degenerate = edge_len(w[0], w[1]) < eps;
    if (m == 3) degenerate = degenerate || edge_len(w[0], w[2]) < eps || edge_len(w[1], w[2]) < eps;
    if (degenerate) {
      keep &= ~(1u << t);
      dropped++;
    }
  }
  kmask[p] = (uint8_t)keep;
  if (dropped) atomicSub(&blkt[p / kMcBlock], dropped);
}

// The listed cells, one per thread; then, if the list overflowed, every cell
// that carries kMeasureMark.
__global__ __launch_bounds__(kMcBlock) void k_mc_measure(const double* __restrict__ field, GridGeom g, double iso,
                                                         int prefilter, const uint32_t* __restrict__ mlist,
                                                         const uint32_t* __restrict__ mcount, uint32_t mcap,
                                                         uint8_t* __restrict__ kmask, uint32_t* __restrict__ blkt) {
  const int64_t stride = (int64_t)gridDim.x * kMcBlock;
  const int64_t total = (int64_t)kMeasureLists * mcap;
  for (int64_t i = (int64_t)blockIdx.x * kMcBlock + threadIdx.x; i < total; i += stride) {
    const int64_t sub = i / mcap, j = i - sub * mcap;
    if (j < (int64_t)mcount[sub]) mc_measure_cell(field, g, iso, prefilter, (int64_t)mlist[i], kmask, blkt);
  }
  if (!mcount[kMeasureLists]) return;  // the same for every thread
  const int64_t N = (int64_t)g.res[0] * g.res[1] * g.res[2];
  for (int64_t p = (int64_t)blockIdx.x * kMcBlock + threadIdx.x; p < N; p += stride)
    if (kmask[p] & kMeasureMark) mc_measure_cell(field, g, iso, prefilter, p, kmask, blkt);
}

// computeEdgeInterpolations (TRIANGLE_MESH.cpp:2373-2425). Vertex ids follow
// the point order: the block's first id from blk_off, then pk16 >> 3. The tile's
// vertices go to verts from the running vertex total base[0] (several tiles
// share the buffer); ids at or beyond cap are not written (the host sees the
// overflow in the totals and meshes again).
__global__ __launch_bounds__(kMcGroup) void k_mc_vertices(const double* __restrict__ field, GridGeom g, double iso,
                                                          const uint16_t* __restrict__ pk16,
                                                          const uint32_t* __restrict__ blk_off,
                                                          const uint64_t* __restrict__ base, int64_t cap,
                                                          double* __restrict__ verts) {
  // each wave lists its block's vertices (at most 3 per point) in LDS in id
  // order and computes them one per lane, so the FP64 work is not spread
  // over mostly idle lanes
  __shared__ uint8_t s_pt[kMcWaves][3 * kMcBlock];
  __shared__ uint8_t s_ax[kMcWaves][3 * kMcBlock];
  const int64_t X = g.res[0], XY = X * g.res[1], N = XY * g.res[2];
  const int64_t B = wave_block((N + kMcBlock - 1) / kMcBlock);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t total = 0;
  if (B >= 0) {
    unsigned pks[4];
#pragma unroll
    for (int sub = 0; sub < 4; sub++) {
      const int64_t p = B * kMcBlock + sub * 64 + lane;
      pks[sub] = p < N ? (unsigned)pk16[p] : 0u;
    }
#pragma unroll
    for (int sub = 0; sub < 4; sub++) {
      const int64_t p = B * kMcBlock + sub * 64 + lane;
      unsigned bits = 0;
      if (p < N) {
        const unsigned pk = pks[sub];
        bits = pk & 7u;
        uint32_t k = pk >> 3;
        for (int axis = 0; axis < 3; axis++)
          if (bits & (1u << axis)) {
            s_pt[w][k] = (uint8_t)(sub * 64 + lane);
            s_ax[w][k] = (uint8_t)axis;
            k++;
          }
      }
      total += wave_prefix3<1, 1, 1>(bits).total;
    }
  }
  __syncthreads();
  if (B < 0) return;
  const int64_t first = (int64_t)base[0] + blk_off[B];
  for (uint32_t j = lane; j < total; j += 64) {
    const int64_t id = first + j;
    if (id >= cap) break;
    const int64_t q = B * kMcBlock + s_pt[w][j];
    const int axis = s_ax[w][j];
    int64_t x64, y64, z64;
    point_xyz(g, q, x64, y64, z64);
    double v[3];
    mc_vertex(g, (int)x64, (int)y64, (int)z64, axis, field[q], field[q + (axis == 0 ? 1 : (axis == 1 ? X : XY))], iso,
              v);
    verts[3 * id + 0] = v[0];
    verts[3 * id + 1] = v[1];
    verts[3 * id + 2] = v[2];
  }
}

// The vertex ids of the points that own the edges of a block's cells: the
// owner of an edge of cell p is p + dx + dy*X + dz*XY, so they lie in 4 rows
// of 257 points from the block's first point (this row, +y, +z, +y+z). Per
// owner its pk16 word (first vertex id within its block << 3 | edge mask);
// per row the first ids of the (at most two) blocks it spans and the entry
// where the second begins.
struct OwnerRows {
  uint16_t pk[4][kMcBlock + 1];
  uint32_t bo[4][2];
  int32_t jb[4];
};

// The block's kept triangles in id order, (cell << 3) | triangle of the case
// row, and the cells' cases.
struct TriList {
  uint16_t e[5 * kMcBlock];
  uint8_t flag[kMcBlock];
};

// addVertexPairTriangle (TRIANGLE_MESH.cpp:2490-2503) in cell order then
// table order, degenerate triangles dropped (kmask from k_mc_cells): the
// block's kept triangles go to the running triangle total base[1] + the
// block's first triangle. Each wave lists its block's kept triangles in LDS
// and then writes them one per lane, so consecutive lanes write consecutive
// triangles. Ids at or beyond tcap are not written.
__global__ __launch_bounds__(kMcGroup) void k_mc_tris(const uint8_t* __restrict__ flags,
                                                      const uint8_t* __restrict__ kmask, GridGeom g,
                                                      const uint16_t* __restrict__ pk16,
                                                      const uint32_t* __restrict__ blk_off,
                                                      const uint64_t* __restrict__ base,
                                                      const uint32_t* __restrict__ blk_toff, int64_t tcap,
                                                      int32_t* __restrict__ tris) {
  __shared__ uint32_t s_tc32[256 * 5 / 2];
  __shared__ OwnerRows s_own[kMcWaves];
  __shared__ TriList s_tl[kMcWaves];
  unsigned short* s_tc = (unsigned short*)s_tc32;
  stage_tri5(s_tc, kMc.tcorner);
  const int64_t X = g.res[0], XY = X * g.res[1], N = XY * g.res[2];
  const int64_t nb = (N + kMcBlock - 1) / kMcBlock;
  const int64_t B = wave_block(nb);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  OwnerRows& own = s_own[w];
  TriList& tl = s_tl[w];
  uint32_t total = 0;
  if (B >= 0) {
    // every load of the staging issued before the first LDS write
    const int64_t p0 = B * kMcBlock;
    unsigned v[4][5], fl[4], km[4], bo[4][2] = {{0u, 0u}, {0u, 0u}, {0u, 0u}, {0u, 0u}};
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int64_t row = p0 + (r & 1) * X + (r >> 1) * XY;
#pragma unroll
      for (int k = 0; k < 5; k++) {
        const int64_t q = row + lane + 64 * k;
        v[r][k] = ((k < 4 || lane == 0) && q < N) ? (unsigned)pk16[q] : 0u;
      }
      const int64_t b0 = row / kMcBlock;
      if (lane == 0) {
        bo[r][0] = b0 < nb ? blk_off[b0] : 0u;
        bo[r][1] = b0 + 1 < nb ? blk_off[b0 + 1] : 0u;
      }
    }
#pragma unroll
    for (int sub = 0; sub < 4; sub++) {
      const int64_t p = p0 + sub * 64 + lane;
      fl[sub] = p < N ? (unsigned)flags[p] : 0u;
      km[sub] = p < N ? (unsigned)kmask[p] : 0u;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int64_t row = p0 + (r & 1) * X + (r >> 1) * XY;
#pragma unroll
      for (int k = 0; k < 5; k++)
        if (k < 4 || lane == 0) own.pk[r][lane + 64 * k] = (uint16_t)v[r][k];
      if (lane == 0) {
        own.bo[r][0] = bo[r][0];
        own.bo[r][1] = bo[r][1];
        own.jb[r] = (int32_t)(kMcBlock - (row - (row / kMcBlock) * kMcBlock));
      }
    }
#pragma unroll
    for (int sub = 0; sub < 4; sub++) {
      const int c = sub * 64 + lane;
      const unsigned flag = fl[sub];
      const unsigned keep = km[sub];
      tl.flag[c] = (uint8_t)flag;
      const WavePrefix tp = wave_prefix3<1, 2, 4>((unsigned)__popc(keep));
      uint32_t k = total + tp.excl;
      for (unsigned m = keep; m; m &= m - 1) tl.e[k++] = (uint16_t)(c << 3 | (__ffs(m) - 1));
      total += tp.total;
    }
  }
  __syncthreads();
  if (B < 0) return;
  const int64_t o = (int64_t)base[1] + blk_toff[B];
  for (uint32_t j = lane; j < total; j += 64) {
    const int64_t t_id = o + j;
    if (t_id >= tcap) break;
    const unsigned e = tl.e[j];
    const int c = (int)(e >> 3);
    const unsigned tc = s_tc[tl.flag[c] * 5 + (e & 7u)];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const unsigned d = (tc >> (5 * k)) & 31u;
      const int dx = d & 1, r = (d >> 1) & 3, axis = d >> 3;
      const int jj = c + dx;
      const unsigned pk = own.pk[r][jj];
      tris[3 * t_id + k] = (int32_t)(own.bo[r][jj >= own.jb[r]] + (pk >> 3) + __popc(pk & ((1u << axis) - 1u)));
    }
  }
}

// After a tile: its vertex and triangle counts (the scans' totals) recorded
// and added to the running totals.
// (Both also clear the measure-list counters for the next field.)
__global__ void k_mc_advance(const uint32_t* __restrict__ totals, uint64_t* __restrict__ base, int tile,
                             int64_t* __restrict__ counts, uint32_t* __restrict__ mcount) {
  for (int i = threadIdx.x; i <= kMeasureLists; i += blockDim.x) mcount[i] = 0u;
  if (threadIdx.x != 0) return;
  const uint64_t nv = totals[1], nt = totals[2];
  counts[2 * tile + 0] = (int64_t)nv;
  counts[2 * tile + 1] = (int64_t)nt;
  base[0] += nv;
  base[1] += nt;
}

__global__ void k_mc_reset(uint64_t* __restrict__ base, uint32_t* __restrict__ mcount) {
  for (int i = threadIdx.x; i <= kMeasureLists; i += blockDim.x) mcount[i] = 0u;
  if (threadIdx.x < 2) base[threadIdx.x] = 0;
}

// koff is the exclusive scan of the keep flags: triangle i is kept when its
// offset differs from the next one (or from the total for the last).
__global__ __launch_bounds__(kMcBlock) void k_mc_squeeze(const int32_t* __restrict__ tris_pre,
                                                         const uint32_t* __restrict__ koff, int64_t n,
                                                         const uint32_t* __restrict__ total,
                                                         int32_t* __restrict__ tris) {
  const int64_t i = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t o = koff[i];
  const uint32_t next = (i + 1 < n) ? koff[i + 1] : *total;
  if (next == o) return;
  tris[3 * (int64_t)o + 0] = tris_pre[3 * i + 0];
  tris[3 * (int64_t)o + 1] = tris_pre[3 * i + 1];
  tris[3 * (int64_t)o + 2] = tris_pre[3 * i + 2];
}

// ---------------------------------------------------------------- host side

template <class T>
static qj_status grow(T*& p, int64_t& cap, int64_t n, hipStream_t s) {
  if (n <= cap) return QJ_OK;
  if (p) {
    QJ_HIP(hipStreamSynchronize(s));
    QJ_HIP(hipFree(p));
    p = nullptr;
    cap = 0;
  }
  const int64_t want = n + n / 8 + 1024;
  QJ_HIP(hipMalloc(&p, sizeof(T) * (size_t)want));
  cap = want;
  return QJ_OK;
}

static unsigned nblocks(int64_t n) { return (unsigned)((n + kMcBlock - 1) / kMcBlock); }

__global__ void k_mc_zero_totals(uint32_t* __restrict__ totals) {
  if (threadIdx.x < 4) totals[threadIdx.x] = 0u;
}


// Meshes n fields one after another on the stream without waiting on the
// host: every count stays on the device (running vertex and triangle totals,
// per-tile counts) and is read back once at the end. The vertex and triangle
// buffers keep their size from earlier calls; a pass that outgrows them
// writes nothing beyond them, and the host then sizes them from the totals it
// read and runs the pass again.
// ext_v / ext_t (with their capacities in vertices and triangles): the
// caller's buffers, written directly in one pass; null = the workspace's.
qj_status mc_run_tiles(qj_workspace* ws, int n, const double* const* fields, const qj_grid* grids, double iso,
                       int64_t* nv_out, int64_t* nt_out, hipStream_t s, double* ext_v, int64_t ext_vcap,
                       int32_t* ext_t, int64_t ext_tcap) {
  const bool ext = ext_v != nullptr;
  if (!ws->mc) ws->mc = new McState();
  McState& m = *(McState*)ws->mc;
  m.nv = m.nt = 0;
  m.cache_valid = false;
  if (n <= 0) return QJ_OK;
  std::vector<GridGeom> geo((size_t)n);
  std::vector<bool> prefilter((size_t)n);
  int64_t nmax = 0, total_pts = 0;
  for (int k = 0; k < n; k++) {
    GridGeom& g = geo[(size_t)k];
    make_geom(&grids[k], &g);
    const int64_t N = (int64_t)g.res[0] * g.res[1] * g.res[2];
    const int64_t cells = (int64_t)(g.res[0] - 1) * (g.res[1] - 1) * (g.res[2] - 1);
    if (N > (int64_t)UINT32_MAX || 5 * cells > (int64_t)INT32_MAX) {
      set_error("marching cubes on %lld points exceeds 32-bit vertex/triangle ids; split the field in tiles",
                (long long)N);
      return QJ_ERR_CAPACITY;
    }
    nmax = N > nmax ? N : nmax;
    total_pts += N;
    // the degenerate-test prefilter holds when the cell-centre map is a
    // rotation (edges meeting at a corner stay perpendicular) and every grid
    // spacing lies in [kPrefilterSpacing, kPrefilterMaxSpacing]
    bool pf = true;
    for (int i = 0; i < 3; i++) {
      pf &= fabs(g.dx[i]) >= kPrefilterSpacing && fabs(g.dx[i]) <= kPrefilterMaxSpacing;
      for (int j = 0; j < 3; j++) {
        const double d = g.rot[i] * g.rot[j] + g.rot[3 + i] * g.rot[3 + j] + g.rot[6 + i] * g.rot[6 + j];
        pf &= fabs(d - (i == j ? 1.0 : 0.0)) < 1e-9;
      }
    }
    prefilter[(size_t)k] = pf;
  }
  qj_status st;
  const int64_t nbmax = nblocks(nmax);
  if ((st = grow(m.flags, m.cap_flags, nmax, s)) || (st = grow(m.kmask, m.cap_kmask, nmax, s)) ||
      (st = grow(m.pk16, m.cap_pk16, nmax, s)) || (st = grow(m.blkv, m.cap_blkv, nbmax, s)) ||
      (st = grow(m.mlist, m.cap_mlist, std::max(nmax / 16, (int64_t)kMeasureLists * 256), s)) ||
      (st = grow(m.blkt, m.cap_blkt, nbmax, s)) || (st = grow(m.d_counts, m.cap_counts, 2 * (int64_t)n, s)))
    return st;
  const size_t need_tmp = scan_tmp_bytes(nbmax > 4096 ? nbmax : 4096);
  if (need_tmp > m.tmp_bytes) {
    if (m.tmp) {
      QJ_HIP(hipStreamSynchronize(s));
      QJ_HIP(hipFree(m.tmp));
    }
    m.tmp = nullptr;
    m.tmp_bytes = 0;
    QJ_HIP(hipMalloc(&m.tmp, need_tmp));
    m.tmp_bytes = need_tmp;
  }
  if (!m.totals) QJ_HIP(hipMalloc(&m.totals, 4 * sizeof(uint32_t)));
  if (!m.d_base) QJ_HIP(hipMalloc(&m.d_base, 2 * sizeof(uint64_t)));
  if (!m.d_mcount) QJ_HIP(hipMalloc(&m.d_mcount, sizeof(uint32_t) * (kMeasureLists + 1)));
  // first sizes: a quarter of the points as vertices, half as triangles
  // (a 256^3 torus tile holds 0.42 and 0.73 of its points)
  if (!ext && m.cap_v == 0 && (st = grow(m.verts, m.cap_v, 3 * (total_pts / 4 + 1024), s))) return st;
  if (!ext && m.cap_t == 0 && (st = grow(m.tris, m.cap_t, 3 * (total_pts / 2 + 1024), s))) return st;
  std::vector<int64_t> counts(2 * (size_t)n);
  uint64_t totals[2] = {0, 0};
  for (int pass = 0; pass < 2; pass++) {
    const int64_t vcap = ext ? ext_vcap : m.cap_v / 3, tcap = ext ? ext_tcap : m.cap_t / 3;
    double* const verts = ext ? ext_v : m.verts;
    int32_t* const tris = ext ? ext_t : m.tris;
    hipLaunchKernelGGL(k_mc_reset, dim3(1), dim3(64), 0, s, m.d_base, m.d_mcount);
    QJ_HIP(hipGetLastError());
    for (int k = 0; k < n; k++) {
      const GridGeom& g = geo[(size_t)k];
      const int64_t N = (int64_t)g.res[0] * g.res[1] * g.res[2];
      const int64_t cells = (int64_t)(g.res[0] - 1) * (g.res[1] - 1) * (g.res[2] - 1);
      const dim3 nb(nblocks(N));
      if (cells <= 0) {
        hipLaunchKernelGGL(k_mc_zero_totals, dim3(1), dim3(64), 0, s, m.totals);
        QJ_HIP(hipGetLastError());
      } else {
        const int pf = prefilter[(size_t)k] ? 1 : 0;
        const dim3 nwg((nb.x + kMcWaves - 1) / kMcWaves);
        const uint32_t mcap = (uint32_t)(m.cap_mlist / kMeasureLists);  // entries per sub-list
        hipLaunchKernelGGL(k_mc_cells, nb, dim3(kMcBlock), 0, s, fields[k], g, iso, pf, m.flags, m.kmask, m.blkt,
                           m.mlist, m.d_mcount, mcap);
        QJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mc_measure, dim3(std::min<unsigned>(nb.x, 1024u)), dim3(kMcBlock), 0, s, fields[k], g,
                           iso, pf, (const uint32_t*)m.mlist, (const uint32_t*)m.d_mcount, mcap, m.kmask, m.blkt);
        QJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mc_edges, nwg, dim3(kMcGroup), 0, s, m.flags, g, m.pk16, m.blkv);
        QJ_HIP(hipGetLastError());
        if ((st = scan_u32(m.blkv, m.blkv, nb.x, m.totals + 1, m.tmp, m.tmp_bytes, s))) return st;
        hipLaunchKernelGGL(k_mc_vertices, nwg, dim3(kMcGroup), 0, s, fields[k], g, iso, m.pk16, m.blkv,
                           (const uint64_t*)m.d_base, vcap, verts);
        QJ_HIP(hipGetLastError());
        if ((st = scan_u32(m.blkt, m.blkt, nb.x, m.totals + 2, m.tmp, m.tmp_bytes, s))) return st;
        hipLaunchKernelGGL(k_mc_tris, nwg, dim3(kMcGroup), 0, s, m.flags, m.kmask, g, m.pk16, m.blkv,
                           (const uint64_t*)m.d_base, m.blkt, tcap, tris);
        QJ_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(k_mc_advance, dim3(1), dim3(64), 0, s, m.totals, m.d_base, k, m.d_counts, m.d_mcount);
      QJ_HIP(hipGetLastError());
    }
    QJ_HIP(hipMemcpyAsync(counts.data(), m.d_counts, sizeof(int64_t) * 2 * (size_t)n, hipMemcpyDeviceToHost, s));
    QJ_HIP(hipMemcpyAsync(totals, m.d_base, sizeof(totals), hipMemcpyDeviceToHost, s));
    QJ_HIP(hipStreamSynchronize(s));
    if ((int64_t)totals[0] <= vcap && (int64_t)totals[1] <= tcap) break;
    if (ext) {
      // the caller's buffers are too small: report the sizes needed
      for (int k = 0; k < n; k++) {
        if (nv_out) nv_out[k] = counts[2 * (size_t)k];
        if (nt_out) nt_out[k] = counts[2 * (size_t)k + 1];
      }
      set_error("marching cubes: the meshes need %llu vertices and %llu triangles (buffers for %lld and %lld)",
                (unsigned long long)totals[0], (unsigned long long)totals[1], (long long)vcap, (long long)tcap);
      return QJ_ERR_CAPACITY;
    }
    if (pass == 1) {
      set_error("marching cubes: counts changed between two passes over the same fields");
      return QJ_ERR_NUMERIC;
    }
    // outgrown: size the buffers from this pass's totals and mesh again
    if ((st = grow(m.verts, m.cap_v, 3 * (int64_t)totals[0], s)) ||
        (st = grow(m.tris, m.cap_t, 3 * (int64_t)totals[1], s)))
      return st;
  }
  m.nv = ext ? 0 : (int64_t)totals[0];
  m.nt = ext ? 0 : (int64_t)totals[1];
  for (int k = 0; k < n; k++) {
    if (nv_out) nv_out[k] = counts[2 * (size_t)k];
    if (nt_out) nt_out[k] = counts[2 * (size_t)k + 1];
  }
  return QJ_OK;
}

qj_status mc_run(qj_workspace* ws, const double* d_field, const qj_grid* grid, double iso, int64_t* nv_out,
                 int64_t* nt_out, hipStream_t s) {
  *nv_out = *nt_out = 0;
  return mc_run_tiles(ws, 1, &d_field, grid, iso, nv_out, nt_out, s);
}

qj_status mc_copy(qj_workspace* ws, double* v, int32_t* t, hipStream_t s) {
  McState* m = (McState*)ws->mc;
  if (!m) return QJ_ERR_INVALID;
  if (v && m->nv) QJ_HIP(hipMemcpyAsync(v, m->verts, sizeof(double) * 3 * m->nv, hipMemcpyDefault, s));
  if (t && m->nt) QJ_HIP(hipMemcpyAsync(t, m->tris, sizeof(int32_t) * 3 * m->nt, hipMemcpyDefault, s));
  QJ_HIP(hipStreamSynchronize(s));
  return QJ_OK;
}

qj_status mc_ptrs(qj_workspace* ws, const double** v, const int32_t** t) {
  McState* m = (McState*)ws->mc;
  if (!m) return QJ_ERR_INVALID;
  if (v) *v = m->verts;
  if (t) *t = m->tris;
  return QJ_OK;
}

void mc_free(qj_workspace* ws) {
  McState* m = (McState*)ws->mc;
  if (!m) return;
  hipFree(m->flags);
  hipFree(m->emask);
  hipFree(m->cnt);
  hipFree(m->pk16);
  hipFree(m->mlist);
  hipFree(m->d_mcount);
  hipFree(m->kmask);
  hipFree(m->blkt);
  hipFree(m->blkv);
  hipFree(m->blk);
  hipFree(m->active);
  hipFree(m->tri_off);
  hipFree(m->verts);
  hipFree(m->tris_pre);
  hipFree(m->keep);
  hipFree(m->tris);
  hipFree(m->tmp);
  hipFree(m->totals);
  hipFree(m->slab[0]);
  hipFree(m->slab[1]);
  if (m->s2) {
    hipStreamSynchronize(m->s2);
    hipStreamDestroy(m->s2);
  }
  for (hipEvent_t e : {m->ev_start, m->ev_field[0], m->ev_field[1], m->ev_free[0], m->ev_free[1]})
    if (e) hipEventDestroy(e);
  hipFree(m->slice_off);
  hipFree(m->vpt);
  hipFree(m->vax);
  hipFree(m->d_bad);
  hipFree(m->c_edges);
  hipFree(m->c_cells);
  hipFree(m->d_base);
  hipFree(m->d_counts);
  delete m;
  ws->mc = nullptr;
}


// ---------------------------------------------------------------- non-linear low-memory mesher

// TRIANGLE_MESH::computeNonlinearMarchingCubesLowMemoryHuge (TRIANGLE_MESH.cpp:681-768)
// streamed over z-chunks: for chunk [z0, z1) the map is evaluated on slices
// [z0-1, z1+1] (one halo slice below for the edges at z0, two above for the
// edges of the points at z1, whose ids belong to the next chunk), cases come
// from the raw slab values log|q| (TRIANGLE_MESH.cpp:379), and every vertex is
// placed by bisection on re-evaluated values log|q| - iso (nonlinearValue
// :1755-1795, midpointSearch :2232-2271, at most 6 halvings, early exit at
// |value| < 1e-8). Vertex numbering, triangle order and the degenerate filter
// are those of the linear mesher. The re-evaluations take exponents 1 and 2
// as exact products (b and b b) and others in the polar form, as the FP64
// band does: the values differ from the polar form's in the last few ulps,
// the same order as the device math library against the host's.

// kPow: kExactOnly when every exponent is 1 or 2 (no polar code, far fewer
// registers), kExactInt otherwise; both give the same values for exponents 1
// and 2. kLag: the map in Lagrange order (QJ_MAP_LAGRANGE).
template <int kPow, bool kLag>
__device__ __forceinline__ double nonlinear_value(const MapArgs& a, const RootRec64* __restrict__ roots,
                                                  const double* p, double iso) {
  if (kLag) {
    // lagrangeNonlinearValue (TRIANGLE_MESH.cpp:1684-1730): an infinite |q|
    // restores the previous iterate and stops. Before the first iteration the
    // reference's previous magnitude is uninitialised; here it is |q0|.
    q64 q{p[0], p[1], p[2], a.slice};
    double magnitude = qmag(q), prev = magnitude;
    int it = 0;
    while (magnitude < a.escape && it < a.max_iter) {
      q = qscale(lagrange64<kPow>(roots, a.nt, a.nb, q), a.exp_scaling);
      magnitude = qmag(q);
      if (isinf(magnitude)) {
        magnitude = prev;
        break;
      }
      prev = magnitude;
      it++;
      if (magnitude < kBlackHole) it = a.max_iter;
    }
    // an infinite value (|q| = 0) is reported as FLT_MAX (TRIANGLE_MESH.cpp:1733-1742)
    const double v = log(magnitude) - iso;
    return isinf(v) ? (double)FLT_MAX : v;
  }
  const Res64 r = eval64_any<kPow>(a, roots, q64{p[0], p[1], p[2], a.slice});
  return r.F - iso;
}

// cases of cells in slab slices [0, ns-1) (global z = s0 + local z)
__global__ __launch_bounds__(kMcBlock) void k_nl_flags(const double* __restrict__ slab, int64_t X, int64_t Y,
                                                       int64_t Z, int64_t s0, int64_t ns,
                                                       uint8_t* __restrict__ flags) {
  const int64_t XY = X * Y, N = XY * ns;
  const int64_t p = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  if (p >= N) return;
  const int64_t lz = p / XY, r = p - lz * XY, y = r / X, x = r - y * X;
  int flag = 0;
  if (x < X - 1 && y < Y - 1 && s0 + lz < Z - 1 && lz + 1 < ns) flag = cell_flag(slab, p, X, XY, 0.0);
  flags[p] = (uint8_t)flag;
}

// edge masks of points in global slices [p0, p0 + np); flags are slab-local with slab start s0
__global__ __launch_bounds__(kMcBlock) void k_nl_edges(const uint8_t* __restrict__ flags, int64_t X, int64_t Y,
                                                       int64_t Z, int64_t s0, int64_t p0, int64_t np,
                                                       uint8_t* __restrict__ emask, uint32_t* __restrict__ cnt) {
  __shared__ unsigned short s_em[256];
  stage_emask(s_em);
  const int64_t XY = X * Y, N = XY * np;
  const int64_t j = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  if (j >= N) return;
  const int64_t lz = j / XY, r = j - lz * XY, y = r / X, x = r - y * X;
  const int64_t z = p0 + lz;
  const int64_t f = (z - s0) * XY + y * X + x;  // slab-local index of the point
  unsigned bits = 0;
  if (x < X - 1) {
    unsigned ref = 0;
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++)
        if (y >= a && z - b >= s0) ref |= (s_em[flags[f - a * X - b * XY]] >> (a + 2 * b)) & 1u;
    bits |= ref;
  }
  if (y < Y - 1) {
    unsigned ref = 0;
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++)
        if (x >= a && z - b >= s0) ref |= (s_em[flags[f - a - b * XY]] >> (4 + a + 2 * b)) & 1u;
    bits |= ref << 1;
  }
  if (z < Z - 1) {
    unsigned ref = 0;
    for (int b = 0; b < 2; b++)
      for (int a = 0; a < 2; a++)
        if (x >= a && y >= b) ref |= (s_em[flags[f - a - b * X]] >> (8 + a + 2 * b)) & 1u;
    bits |= ref << 2;
  }
  emask[j] = (uint8_t)bits;
  cnt[j] = (bits & 1u) + ((bits >> 1) & 1u) + ((bits >> 2) & 1u);
}

// One vertex by bisection between grid points (x, y, z) and its +axis
// neighbour (computeNonlinearEdgeInterpolationsHuge :2034-2123 with
// midpointSearch :2232-2271). va: value at (x, y, z). Counts edges whose end
// values do not straddle the surface in *bad (the reference exits there).
// fb: the field value at the +axis neighbour when the caller has it (the value
// minus iso is the neighbour's end value), else null and it is evaluated.
template <int kPow, bool kLag>
__device__ __forceinline__ void place_vertex(const MapArgs& a, const RootRec64* __restrict__ roots, double iso,
                                             int x, int y, int z, int axis, const double* pa, double va,
                                             const double* fb, double* __restrict__ out,
                                             unsigned int* __restrict__ bad) {
#pragma clang fp contract(off)
  double pb[3];
  cell_center(a, x + (axis == 0), y + (axis == 1), z + (axis == 2), pb);
  const double vb = fb ? *fb - iso : nonlinear_value<kPow, kLag>(a, roots, pb, iso);
  // first = lower endpoint; the positive end is the first unless firstValue < 0
  double pp[3], pn[3], vp, vn;
  if (va < 0) {
    pp[0] = pb[0]; pp[1] = pb[1]; pp[2] = pb[2]; vp = vb;
    pn[0] = pa[0]; pn[1] = pa[1]; pn[2] = pa[2]; vn = va;
  } else {
    pp[0] = pa[0]; pp[1] = pa[1]; pp[2] = pa[2]; vp = va;
    pn[0] = pb[0]; pn[1] = pb[1]; pn[2] = pb[2]; vn = vb;
  }
  if (va * vb >= 0.0) atomicAdd(bad, 1u);  // the reference exits here (:2082-2097)
  double m[3];
  for (int rec = 0;; rec++) {
    m[0] = pp[0] + pn[0];
    m[1] = pp[1] + pn[1];
    m[2] = pp[2] + pn[2];
    m[0] *= 0.5;
    m[1] *= 0.5;
    m[2] *= 0.5;
    if (vp * vn >= 0.0) {
      if (rec > 0) atomicAdd(bad, 1u);
      break;
    }
    if (rec >= 6) break;
    const double vm = nonlinear_value<kPow, kLag>(a, roots, m, iso);
    if (fabs(vm) < 1e-8) break;
    if (vm < 0) {
      pn[0] = m[0]; pn[1] = m[1]; pn[2] = m[2]; vn = vm;
    } else {
      pp[0] = m[0]; pp[1] = m[1]; pp[2] = m[2]; vp = vm;
    }
  }
  out[0] = m[0];
  out[1] = m[1];
  out[2] = m[2];
}

// The vertices of a range of points (off: the exclusive scan of the edge
// counts, from the range's first point) in id order: the owning point
// (relative to the range) and the edge's axis.
__global__ __launch_bounds__(kMcBlock) void k_nl_vlist(const uint8_t* __restrict__ emask,
                                                       const uint32_t* __restrict__ off, int64_t n,
                                                       uint32_t* __restrict__ vpt, uint8_t* __restrict__ vax) {
  const int64_t j = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  if (j >= n) return;
  const unsigned bits = emask[j];
  if (!bits) return;
  uint32_t k = off[j] - off[0];
  for (int axis = 0; axis < 3; axis++)
    if (bits & (1u << axis)) {
      vpt[k] = (uint32_t)j;
      vax[k] = (uint8_t)axis;
      k++;
    }
}

// Vertices by bisection, one per thread over the list: both end values
// (evaluated, or read from the slab), then up to 6 midpoints (FP64 map
// evaluations on full waves; about one point in ten owns an edge, so a thread
// per point would leave most lanes idle). One launch takes two ranges of
// points: the chunk's own points from slice p0a (list entries [0, nA), nA the
// scan total minus the scan at the range's first point, off_a null when the
// range is empty; vertex t gets id base + off_a[0] + t) and the next chunk's
// first slice p0b (entries [b_start, b_start + nb), ids base + t), so the
// slice's few vertices do not take a launch of their own. List entries hold
// the point relative to their range's first point. ends_a / ends_b: null, or
// the slab's values from the range's first point on, when they are the end
// values' own (see slab_holds_end_values); the midpoints are always evaluated.
template <int kPow, bool kLag>
__global__ __launch_bounds__(kMcBlock) void k_nl_vertices(MapArgs a, const RootRec64* __restrict__ roots,
                                                          double iso, int64_t p0a, const double* __restrict__ ends_a,
                                                          const uint32_t* __restrict__ off_a,
                                                          const uint32_t* __restrict__ total_a, int64_t p0b,
                                                          const double* __restrict__ ends_b, uint32_t b_start,
                                                          uint32_t nb, const uint32_t* __restrict__ vpt,
                                                          const uint8_t* __restrict__ vax, uint32_t base,
                                                          double* __restrict__ verts, int32_t* __restrict__ edges,
                                                          unsigned int* __restrict__ bad) {
  const uint32_t t = blockIdx.x * kMcBlock + threadIdx.x;
  int64_t p0, k;
  const double* ends;
  if (t >= b_start) {
    if (t - b_start >= nb) return;
    p0 = p0b;
    ends = ends_b;
    k = (int64_t)base + t;
  } else {
    if (!off_a) return;
    const uint32_t first = off_a[0];
    if (t >= *total_a - first) return;
    p0 = p0a;
    ends = ends_a;
    k = (int64_t)base + first + t;
  }
  const int64_t X = a.res[0], XY = X * a.res[1];
  const int64_t q = vpt[t];
  const int axis = vax[t];
  const int64_t lz = q / XY, r = q - lz * XY;
  const int y = (int)(r / X), x = (int)(r - (int64_t)y * X), z = (int)(p0 + lz);
  double pa[3];
  cell_center(a, x, y, z, pa);
  const double va = ends ? ends[q] - iso : nonlinear_value<kPow, kLag>(a, roots, pa, iso);
  const double* fb = ends ? ends + q + (axis == 0 ? 1 : axis == 1 ? X : XY) : nullptr;
  place_vertex<kPow, kLag>(a, roots, iso, x, y, z, axis, pa, va, fb, verts + 3 * k, bad);
  edges[4 * k + 0] = x;
  edges[4 * k + 1] = y;
  edges[4 * k + 2] = z;
  edges[4 * k + 3] = axis;
}

// Launch a mesher vertex kernel instantiated for the map's power mode and
// order (see nonlinear_value).
template <class... Args>
static void launch_nl_vertices(const MapArgs& a, dim3 grid, hipStream_t s, Args... args) {
  const bool lag = (a.flags & QJ_MAP_LAGRANGE) != 0;
  if (a.gen && lag) hipLaunchKernelGGL((k_nl_vertices<kExactInt, true>), grid, dim3(kMcBlock), 0, s, a, args...);
  else if (a.gen) hipLaunchKernelGGL((k_nl_vertices<kExactInt, false>), grid, dim3(kMcBlock), 0, s, a, args...);
  else if (lag) hipLaunchKernelGGL((k_nl_vertices<kExactOnly, true>), grid, dim3(kMcBlock), 0, s, a, args...);
  else hipLaunchKernelGGL((k_nl_vertices<kExactOnly, false>), grid, dim3(kMcBlock), 0, s, a, args...);
}
// (x, y, z, flag) of the chunk's active cells, appended in cell order
__global__ __launch_bounds__(kMcBlock) void k_nl_cells(const uint32_t* __restrict__ active, int64_t na,
                                                       const uint8_t* __restrict__ flags, int64_t fo, int64_t X,
                                                       int64_t XY, int64_t z0, int32_t* __restrict__ cells) {
  const int64_t i = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  if (i >= na) return;
  const int64_t c = active[i];
  const int64_t lz = c / XY, r = c - lz * XY, y = r / X, x = r - y * X;
  cells[4 * i + 0] = (int32_t)x;
  cells[4 * i + 1] = (int32_t)y;
  cells[4 * i + 2] = (int32_t)(z0 + lz);
  cells[4 * i + 3] = flags[fo + c];
}

// per-block active counts of cells in global slices [c0, c0 + nc) (slab-local flags from s0)
__global__ __launch_bounds__(kMcBlock) void k_nl_active_count(const uint8_t* __restrict__ flags, int64_t XY,
                                                              int64_t c0, int64_t s0, int64_t nc,
                                                              uint32_t* __restrict__ blk) {
  __shared__ uint32_t s_n[kMcBlock / 64];
  const int64_t j = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  const bool act = j < XY * nc && kMc.ntri[flags[(c0 - s0) * XY + j]] > 0;
  const unsigned long long m = __ballot(act);
  if ((threadIdx.x & 63) == 0) s_n[threadIdx.x >> 6] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < kMcBlock / 64; i++) t += s_n[i];
    blk[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kMcBlock) void k_nl_active(const uint8_t* __restrict__ flags, int64_t XY, int64_t c0,
                                                        int64_t s0, int64_t nc, const uint32_t* __restrict__ blk_off,
                                                        uint32_t* __restrict__ active) {
  __shared__ uint32_t s_n[kMcBlock / 64];
  const int64_t j = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  const bool act = j < XY * nc && kMc.ntri[flags[(c0 - s0) * XY + j]] > 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(act);
  if (lane == 0) s_n[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  uint32_t o = blk_off[blockIdx.x];
  for (int i = 0; i < wave; i++) o += s_n[i];
  if (act) active[o + __popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)j;  // chunk-local cell index
}

__global__ __launch_bounds__(kMcBlock) void k_nl_tricount(const uint32_t* __restrict__ active, int64_t na,
                                                          const uint8_t* __restrict__ flags, int64_t fo,
                                                          uint32_t* __restrict__ tcount) {
  const int64_t i = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  if (i < na) tcount[i] = kMc.ntri[flags[fo + active[i]]];
}

// triangles of the chunk's active cells; owners at global z < z1 use base + off,
// owners at z == z1 use base_next + slice_off
__global__ __launch_bounds__(kMcBlock) void k_nl_triangles(
    const uint32_t* __restrict__ active, int64_t na, const uint8_t* __restrict__ flags, int64_t fo, int64_t X,
    int64_t XY, int64_t z0, int64_t z1, const uint8_t* __restrict__ emask, const uint32_t* __restrict__ off,
    uint32_t base, const uint8_t* __restrict__ emask_next, const uint32_t* __restrict__ slice_off,
    uint32_t base_next, const double* __restrict__ verts, const uint32_t* __restrict__ toff,
    int32_t* __restrict__ tris_pre, uint32_t* __restrict__ keep) {
  __shared__ uint32_t s_tri[1024];
  __shared__ uint8_t s_ntri[256];
  stage_tri(s_tri, s_ntri);
  const int64_t i = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  if (i >= na) return;
  const int64_t c = active[i];  // chunk-local cell index: (z - z0) * XY + y * X + x
  const int flag = flags[fo + c];
  const int n = s_ntri[flag];
  uint32_t o = toff[i];
  for (int t = 0; t < n; t++) {
    int32_t v[3];
    for (int q = 0; q < 3; q++) {
      int dx, dy, dz, axis;
      edge_geom(tri_edge(s_tri, flag, 3 * t + q), dx, dy, dz, axis);
      const int64_t owner = c + dx + dy * X + dz * XY;  // chunk-local point index
      const unsigned lowbits = (1u << axis) - 1u;
      if (z0 + owner / XY < z1) {
        v[q] = (int32_t)(base + off[owner] + __popc(emask[owner] & lowbits));
      } else {
        const int64_t so = owner - (z1 - z0) * XY;
        v[q] = (int32_t)(base_next + slice_off[so] + __popc(emask_next[so] & lowbits));
      }
    }
    const double eps = 1e-7;
    const double* p0 = verts + 3 * (int64_t)v[0];
    const double* p1 = verts + 3 * (int64_t)v[1];
    const double* p2 = verts + 3 * (int64_t)v[2];
    const bool degenerate = edge_len(p0, p1) < eps || edge_len(p0, p2) < eps || edge_len(p1, p2) < eps;
    tris_pre[3 * (int64_t)o + 0] = v[0];
    tris_pre[3 * (int64_t)o + 1] = v[1];
    tris_pre[3 * (int64_t)o + 2] = v[2];
    keep[o] = degenerate ? 0u : 1u;
    o++;
  }
}

template <class T>
static qj_status grow_keep(T*& p, int64_t& cap, int64_t n, int64_t used, hipStream_t s) {
  if (n <= cap) return QJ_OK;
  const int64_t want = n + n / 2 + 4096;
  T* q = nullptr;
  QJ_HIP(hipMalloc(&q, sizeof(T) * (size_t)want));
  if (p && used > 0) QJ_HIP(hipMemcpyAsync(q, p, sizeof(T) * (size_t)used, hipMemcpyDeviceToDevice, s));
  QJ_HIP(hipStreamSynchronize(s));
  if (p) QJ_HIP(hipFree(p));
  p = q;
  cap = want;
  return QJ_OK;
}

// Whether the slab's values can stand in for the bisection's end values
// (TRIANGLE_MESH.cpp:2063-2067 evaluates them again with nonlinearValue,
// whose values the reference's slices hold too: computeNonlinearSlice
// :330-378). In the FP32 + FP64 band field every cell with |F| < band holds
// the FP64 value of the evaluator nonlinear_value calls (k_list64 and
// k_band_sweep with the same power mode and range), and every other cell an
// FP32 value of the FP64 value's sign: with iso 0 the end values enter the
// bisection only through their signs (the orientation and the straddle
// tests; products of values of magnitude band or more do not underflow), so
// the vertices are those of the evaluated end values. Not for the caller's
// field (its values are not the map's), the Lagrange order (its value has
// its own infinite case), the plain FP64 slabs (polar powers, where the
// bisection evaluates exact ones), the AP flags or a band of 0.
static bool slab_holds_end_values(const MapArgs& a, const double* src_field, double iso, qj_precision prec,
                                  double band) {
  return !src_field && prec == QJ_FP32_BAND && iso == 0.0 && band > 0.0 &&
         !(a.flags & (QJ_MAP_LAGRANGE | QJ_MAP_SKIP_ORIGIN | QJ_MAP_ADAPTIVE));
}

qj_status nlmc_run(qj_workspace* ws, const MapArgs& a0, const qj_rational* f, double iso, qj_precision prec,
                   double band, int chunk, int64_t* nv_out, int64_t* nt_out, int64_t* n_bad, hipStream_t s,
                   const double* src_field) {
  if (!ws->mc) ws->mc = new McState();
  McState& m = *(McState*)ws->mc;
  m.nv = m.nt = 0;
  const int64_t X = a0.res[0], Y = a0.res[1], Z = a0.res[2], XY = X * Y;
  *nv_out = *nt_out = 0;
  *n_bad = 0;
  m.ncells = 0;
  m.cache_valid = false;
  if (X < 2 || Y < 2 || Z < 2) return QJ_OK;
  if (chunk < 1) chunk = 64;
  if (XY * (int64_t)(chunk + 3) > (int64_t)UINT32_MAX) {
    set_error("non-linear mesher: slab of %lld points exceeds 32-bit indices", (long long)(XY * (chunk + 3)));
    return QJ_ERR_CAPACITY;
  }
  const int64_t slab_pts = XY * (chunk + 3), own_pts = XY * (chunk + 1);
  qj_status st;
  // the slab stream is idle whenever the call returns, early error returns
  // included (its slab evaluations read the caller's roots)
  struct JoinOnExit {
    hipStream_t s = nullptr;
    ~JoinOnExit() {
      if (s) hipStreamSynchronize(s);
    }
  } join;
  if (!src_field) {
    if (m.s2) QJ_HIP(hipStreamSynchronize(m.s2));
    if ((st = grow(m.slab[0], m.cap_slab[0], slab_pts, s)) || (st = grow(m.slab[1], m.cap_slab[1], slab_pts, s)))
      return st;
    // the band lists at the largest slab's size, so no slab evaluation regrows them
    if ((st = ensure_list(ws, slab_pts, (a0.flags & QJ_MAP_ADAPTIVE) != 0, s))) return st;
    // created on the first mesh (what a failed creation left out, on the next)
    if (!m.s2) QJ_HIP(hipStreamCreateWithFlags(&m.s2, hipStreamNonBlocking));
    join.s = m.s2;
    for (hipEvent_t* e : {&m.ev_start, &m.ev_field[0], &m.ev_field[1], &m.ev_free[0], &m.ev_free[1]})
      if (!*e) QJ_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
  }
  if ((st = grow(m.flags, m.cap_flags, slab_pts, s)) ||
      (st = grow(m.emask, m.cap_emask, own_pts, s)) || (st = grow(m.cnt, m.cap_cnt, own_pts, s)) ||
      (st = grow(m.active, m.cap_active, own_pts, s)) || (st = grow(m.tri_off, m.cap_toff, own_pts, s)) ||
      (st = grow(m.blk, m.cap_blk, (int64_t)nblocks(own_pts), s)) ||
      (st = grow(m.slice_off, m.cap_slice, XY, s)))
    return st;
  const size_t need_tmp = scan_tmp_bytes(5 * own_pts);
  if (need_tmp > m.tmp_bytes) {
    if (m.tmp) {
      QJ_HIP(hipStreamSynchronize(s));
      QJ_HIP(hipFree(m.tmp));
    }
    m.tmp = nullptr;
    m.tmp_bytes = 0;
    QJ_HIP(hipMalloc(&m.tmp, need_tmp));
    m.tmp_bytes = need_tmp;
  }
  if (!m.totals) QJ_HIP(hipMalloc(&m.totals, 4 * sizeof(uint32_t)));
  if (!m.d_bad) QJ_HIP(hipMalloc(&m.d_bad, sizeof(unsigned int)));
  QJ_HIP(hipMemsetAsync(m.d_bad, 0, sizeof(unsigned int), s));
  // the mesher's map: escape 2000 and the black-hole stop (TRIANGLE_MESH.cpp:97, 1789-1791)
  MapArgs a = a0;
  const bool reuse_ends = slab_holds_end_values(a, src_field, iso, prec, band);
  // slices [s0, s0 + ns) of the slab of the chunk that starts at slice zc: the
  // chunk's points and one halo slice on each side
  auto slab_range = [&](int64_t zc, int64_t& s0, int64_t& ns) {
    const int64_t zc1 = zc + chunk < Z ? zc + chunk : Z;
    s0 = zc > 0 ? zc - 1 : 0;
    const int64_t s1 = zc1 + 1 < Z - 1 ? zc1 + 1 : Z - 1;
    ns = s1 - s0 + 1;
  };
  // 1. slab of log|q| (or the caller's field, TRIANGLE_MESH.cpp:152-215): the
  // chunk at slice zc into slab[b] on stream s2 (the Lagrange order has no FP32
  // path: its slabs are FP64)
  auto eval_slab = [&](int64_t zc, int b) -> qj_status {
    int64_t fs0, fns;
    slab_range(zc, fs0, fns);
    MapArgs sa = a;
    sa.res[2] = (int32_t)fns;
    sa.zoff = (int32_t)fs0;
    qj_status e = launch_field(ws, sa, f, (a.flags & QJ_MAP_LAGRANGE) ? QJ_FP64 : prec, band, m.slab[b], m.s2);
    if (e) return e;
    QJ_HIP(hipEventRecord(m.ev_field[b], m.s2));
    return QJ_OK;
  };
  if (!src_field) {
    // the first slab after the work the caller queued on s
    QJ_HIP(hipEventRecord(m.ev_start, s));
    QJ_HIP(hipStreamWaitEvent(m.s2, m.ev_start, 0));
    if ((st = eval_slab(0, 0))) return st;
  }
  uint64_t nv = 0, nt = 0, placed_next = 0;
  for (int64_t z0 = 0, ci = 0; z0 < Z; z0 += chunk, ci++) {
    const int64_t z1 = z0 + chunk < Z ? z0 + chunk : Z;
    int64_t s0, ns;
    slab_range(z0, s0, ns);
    const int b = (int)(ci & 1);
    const double* slab = src_field ? src_field + s0 * XY : m.slab[b];
    if (!src_field) {
      // the next chunk's slab goes into the other buffer once the previous
      // chunk's vertex kernel, its last reader, is done; this chunk's cases
      // wait for this chunk's slab
      if (z1 < Z) {
        if (ci > 0) QJ_HIP(hipStreamWaitEvent(m.s2, m.ev_free[1 - b], 0));
        if ((st = eval_slab(z1, 1 - b))) return st;
      }
      QJ_HIP(hipStreamWaitEvent(s, m.ev_field[b], 0));
    }
    const double* ends = reuse_ends ? slab : nullptr;
    // 2. cases
    hipLaunchKernelGGL(k_nl_flags, dim3(nblocks(XY * ns)), dim3(kMcBlock), 0, s, slab, X, Y, Z, s0, ns, m.flags);
    QJ_HIP(hipGetLastError());
    // 3. edge masks of points [z0, z1] (z1 only when it exists)
    const int64_t np = (z1 < Z ? z1 + 1 : Z) - z0;
    hipLaunchKernelGGL(k_nl_edges, dim3(nblocks(XY * np)), dim3(kMcBlock), 0, s, m.flags, X, Y, Z, s0, z0, np,
                       m.emask, m.cnt);
    QJ_HIP(hipGetLastError());
    // 4. vertex ids: owned points [z0, z1), and the next chunk's first slice separately
    const int64_t nown = XY * (z1 - z0);
    uint32_t h[4] = {0, 0, 0, 0};
    if (z1 < Z) {
      if ((st = scan_u32(m.cnt + nown, m.slice_off, XY, m.totals + 3, m.tmp, m.tmp_bytes, s))) return st;
      QJ_HIP(hipMemcpyAsync(h + 3, m.totals + 3, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    }
    if ((st = scan_u32(m.cnt, m.cnt, nown, m.totals + 1, m.tmp, m.tmp_bytes, s))) return st;
    QJ_HIP(hipMemcpyAsync(h + 1, m.totals + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    QJ_HIP(hipStreamSynchronize(s));
    const uint64_t cnt_own = h[1], cnt_next = h[3];
    if (nv + cnt_own + cnt_next > (uint64_t)INT32_MAX) return QJ_ERR_CAPACITY;
    // The chunk's triangles reach the first slice of the next chunk, whose
    // vertices the degenerate filter reads: place them now (ids base_next +
    // slice_off, the ids the next chunk's numbering gives them) and skip the
    // chunk's own first slice, placed by the previous chunk.
    // keep what is already placed: ids < nv plus this chunk's first slice (placed by the previous chunk)
    if ((st = grow_keep(m.verts, m.cap_v, 3 * (int64_t)(nv + cnt_own + cnt_next), 3 * (int64_t)(nv + placed_next),
                        s)))
      return st;
    if ((st = grow_keep(m.c_edges, m.cap_ce, 4 * (int64_t)(nv + cnt_own + cnt_next),
                        4 * (int64_t)(nv + placed_next), s)))
      return st;
    const int64_t skip = z0 > 0 ? 1 : 0;
    if ((st = grow(m.vpt, m.cap_vpt, (int64_t)(cnt_own + cnt_next), s))) return st;
    if ((st = grow(m.vax, m.cap_vax, (int64_t)(cnt_own + cnt_next), s))) return st;
    const bool own = cnt_own > 0 && z1 - z0 > skip;
    if (own) {
      const int64_t n = XY * (z1 - z0 - skip);
      hipLaunchKernelGGL(k_nl_vlist, dim3(nblocks(n)), dim3(kMcBlock), 0, s, m.emask + skip * XY,
                         m.cnt + skip * XY, n, m.vpt, m.vax);
      QJ_HIP(hipGetLastError());
    }
    if (cnt_next > 0) {
      hipLaunchKernelGGL(k_nl_vlist, dim3(nblocks(XY)), dim3(kMcBlock), 0, s, m.emask + nown, m.slice_off, XY,
                         m.vpt + cnt_own, m.vax + cnt_own);
      QJ_HIP(hipGetLastError());
    }
    if (own || cnt_next > 0) {
      launch_nl_vertices(a, dim3(nblocks((int64_t)(cnt_own + cnt_next))), s, f->d64, iso, z0 + skip,
                         ends ? ends + (z0 + skip - s0) * XY : (const double*)nullptr,
                         own ? (const uint32_t*)(m.cnt + skip * XY) : (const uint32_t*)nullptr,
                         (const uint32_t*)(m.totals + 1), z1, ends ? ends + (z1 - s0) * XY : (const double*)nullptr,
                         (uint32_t)cnt_own, (uint32_t)cnt_next, (const uint32_t*)m.vpt, (const uint8_t*)m.vax,
                         (uint32_t)nv, m.verts, m.c_edges, m.d_bad);
      QJ_HIP(hipGetLastError());
    }
    if (!src_field) QJ_HIP(hipEventRecord(m.ev_free[b], s));  // the last reader of slab[b] is queued
    // 5. active cells of the chunk, in cell order
    const int64_t cz1 = z1 < Z - 1 ? z1 : Z - 1;  // cells exist for z <= Z-2
    const int64_t nc = cz1 > z0 ? cz1 - z0 : 0;
    int64_t na = 0;
    if (nc > 0) {
      const int64_t nb = nblocks(XY * nc);
      hipLaunchKernelGGL(k_nl_active_count, dim3((unsigned)nb), dim3(kMcBlock), 0, s, m.flags, XY, z0, s0, nc,
                         m.blk);
      QJ_HIP(hipGetLastError());
      if ((st = scan_u32(m.blk, m.blk, nb, m.totals + 0, m.tmp, m.tmp_bytes, s))) return st;
      hipLaunchKernelGGL(k_nl_active, dim3((unsigned)nb), dim3(kMcBlock), 0, s, m.flags, XY, z0, s0, nc, m.blk,
                         m.active);
      QJ_HIP(hipGetLastError());
      QJ_HIP(hipMemcpyAsync(h, m.totals, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      QJ_HIP(hipStreamSynchronize(s));
      na = h[0];
    }
    if (na > 0) {
      const int64_t fo = (z0 - s0) * XY;
      if ((st = grow_keep(m.c_cells, m.cap_cc, 4 * (m.ncells + na), 4 * m.ncells, s))) return st;
      hipLaunchKernelGGL(k_nl_cells, dim3(nblocks(na)), dim3(kMcBlock), 0, s, m.active, na, m.flags, fo, X, XY, z0,
                         m.c_cells + 4 * m.ncells);
      QJ_HIP(hipGetLastError());
      m.ncells += na;
      hipLaunchKernelGGL(k_nl_tricount, dim3(nblocks(na)), dim3(kMcBlock), 0, s, m.active, na, m.flags, fo,
                         m.tri_off);
      QJ_HIP(hipGetLastError());
      if ((st = scan_u32(m.tri_off, m.tri_off, na, m.totals + 2, m.tmp, m.tmp_bytes, s))) return st;
      QJ_HIP(hipMemcpyAsync(h + 2, m.totals + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      QJ_HIP(hipStreamSynchronize(s));
      const int64_t ntp = h[2];
      if ((st = grow(m.tris_pre, m.cap_tp, 3 * ntp, s)) || (st = grow(m.keep, m.cap_keep, ntp, s))) return st;
      hipLaunchKernelGGL(k_nl_triangles, dim3(nblocks(na)), dim3(kMcBlock), 0, s, m.active, na, m.flags, fo, X, XY,
                         z0, z1, m.emask, m.cnt, (uint32_t)nv, m.emask + nown, m.slice_off,
                         (uint32_t)(nv + cnt_own), m.verts, m.tri_off, m.tris_pre, m.keep);
      QJ_HIP(hipGetLastError());
      if ((st = scan_u32(m.keep, m.keep, ntp, m.totals + 3, m.tmp, m.tmp_bytes, s))) return st;
      QJ_HIP(hipMemcpyAsync(h + 3, m.totals + 3, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      QJ_HIP(hipStreamSynchronize(s));
      const int64_t ntk = h[3];
      if ((st = grow_keep(m.tris, m.cap_t, 3 * (int64_t)(nt + ntk), 3 * (int64_t)nt, s))) return st;
      if (ntk > 0) {
        hipLaunchKernelGGL(k_mc_squeeze, dim3(nblocks(ntp)), dim3(kMcBlock), 0, s, m.tris_pre, m.keep, ntp,
                           m.totals + 3, m.tris + 3 * nt);
        QJ_HIP(hipGetLastError());
      }
      nt += ntk;
    }
    nv += cnt_own;
    placed_next = cnt_next;
  }
  unsigned int bad = 0;
  QJ_HIP(hipMemcpyAsync(&bad, m.d_bad, sizeof(bad), hipMemcpyDeviceToHost, s));
  QJ_HIP(hipStreamSynchronize(s));
  m.nv = (int64_t)nv;
  m.nt = (int64_t)nt;
  m.cache_valid = true;
  *nv_out = m.nv;
  *nt_out = m.nt;
  *n_bad = bad;
  if (bad) {
    set_error("non-linear mesher: %u edges whose end values do not straddle the isosurface "
              "(the reference exits at TRIANGLE_MESH.cpp:2082-2097)", bad);
    return QJ_ERR_NUMERIC;
  }
  return QJ_OK;
}

// ---------------------------------------------------------------- edge cache

qj_status nlmc_cache(qj_workspace* ws, int64_t* n_edges, int32_t* edges, int64_t* n_cells, int32_t* cells,
                     hipStream_t s) {
  McState* m = (McState*)ws->mc;
  if (!m || !m->cache_valid) {
    set_error("qj_nlmc_cache: no non-linear mesh in this workspace");
    return QJ_ERR_INVALID;
  }
  if (n_edges) *n_edges = m->nv;
  if (n_cells) *n_cells = m->ncells;
  if (edges && m->nv) QJ_HIP(hipMemcpyAsync(edges, m->c_edges, sizeof(int32_t) * 4 * m->nv, hipMemcpyDefault, s));
  if (cells && m->ncells)
    QJ_HIP(hipMemcpyAsync(cells, m->c_cells, sizeof(int32_t) * 4 * m->ncells, hipMemcpyDefault, s));
  QJ_HIP(hipStreamSynchronize(s));
  return QJ_OK;
}

template <int kPow, bool kLag>
__global__ __launch_bounds__(kMcBlock) void k_cache_vertices(MapArgs a, const RootRec64* __restrict__ roots,
                                                             double iso, const int32_t* __restrict__ edges, int64_t ne,
                                                             double* __restrict__ verts,
                                                             unsigned int* __restrict__ bad) {
  const int64_t i = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  if (i >= ne) return;
  const int x = edges[4 * i], y = edges[4 * i + 1], z = edges[4 * i + 2], axis = edges[4 * i + 3];
  double pa[3];
  cell_center(a, x, y, z, pa);
  const double va = nonlinear_value<kPow, kLag>(a, roots, pa, iso);
  place_vertex<kPow, kLag>(a, roots, iso, x, y, z, axis, pa, va, nullptr, verts + 3 * i, bad);
}

template <class... Args>
static void launch_cache_vertices(const MapArgs& a, dim3 grid, hipStream_t s, Args... args) {
  const bool lag = (a.flags & QJ_MAP_LAGRANGE) != 0;
  if (a.gen && lag) hipLaunchKernelGGL((k_cache_vertices<kExactInt, true>), grid, dim3(kMcBlock), 0, s, a, args...);
  else if (a.gen) hipLaunchKernelGGL((k_cache_vertices<kExactInt, false>), grid, dim3(kMcBlock), 0, s, a, args...);
  else if (lag) hipLaunchKernelGGL((k_cache_vertices<kExactOnly, true>), grid, dim3(kMcBlock), 0, s, a, args...);
  else hipLaunchKernelGGL((k_cache_vertices<kExactOnly, false>), grid, dim3(kMcBlock), 0, s, a, args...);
}

__global__ __launch_bounds__(kMcBlock) void k_cache_tricount(const int32_t* __restrict__ cells, int64_t nc,
                                                             uint32_t* __restrict__ tcount) {
  const int64_t i = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  if (i < nc) tcount[i] = kMc.ntri[cells[4 * i + 3]];
}

// position of key in the ascending keys, or -1
__device__ __forceinline__ int64_t find_key(const int64_t* __restrict__ keys, int64_t n, int64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return (lo < n && keys[lo] == key) ? lo : -1;
}

__global__ __launch_bounds__(kMcBlock) void k_cache_triangles(
    const int32_t* __restrict__ cells, int64_t nc, int64_t X, int64_t XY, const int64_t* __restrict__ keys,
    int64_t ne, const double* __restrict__ verts, const uint32_t* __restrict__ toff, int32_t* __restrict__ tris_pre,
    uint32_t* __restrict__ keep, unsigned int* __restrict__ missing) {
  const int64_t i = (int64_t)blockIdx.x * kMcBlock + threadIdx.x;
  if (i >= nc) return;
  const int64_t x = cells[4 * i], y = cells[4 * i + 1], z = cells[4 * i + 2];
  const int flag = cells[4 * i + 3];
  const int n = kMc.ntri[flag];
  uint32_t o = toff[i];
  for (int t = 0; t < n; t++) {
    int32_t v[3];
    bool ok = true;
    for (int q = 0; q < 3; q++) {
      int dx, dy, dz, axis;
      edge_geom(kMc.tri[flag][3 * t + q], dx, dy, dz, axis);
      const int64_t key = ((z + dz) * XY + (y + dy) * X + (x + dx)) * 3 + axis;
      const int64_t id = find_key(keys, ne, key);
      if (id < 0) ok = false;
      v[q] = id < 0 ? 0 : (int32_t)id;
    }
    bool degenerate = false;
    if (!ok) {
      atomicAdd(missing, 1u);
    } else {
      // vertex coordinates are read only for triangles whose three edges are listed
      const double eps = 1e-7;
      const double* p0 = verts + 3 * (int64_t)v[0];
      const double* p1 = verts + 3 * (int64_t)v[1];
      const double* p2 = verts + 3 * (int64_t)v[2];
      degenerate = edge_len(p0, p1) < eps || edge_len(p0, p2) < eps || edge_len(p1, p2) < eps;
    }
    tris_pre[3 * (int64_t)o + 0] = v[0];
    tris_pre[3 * (int64_t)o + 1] = v[1];
    tris_pre[3 * (int64_t)o + 2] = v[2];
    keep[o] = (degenerate || !ok) ? 0u : 1u;
    o++;
  }
}

qj_status nlmc_resume(qj_workspace* ws, const MapArgs& a, const qj_rational* f, double iso, int64_t ne,
                      const int32_t* h_edges, int64_t nc, const int32_t* h_cells, int64_t* nv_out, int64_t* nt_out,
                      hipStream_t s) {
  if (!ws->mc) ws->mc = new McState();
  McState& m = *(McState*)ws->mc;
  m.nv = m.nt = 0;
  m.cache_valid = false;
  *nv_out = *nt_out = 0;
  const int64_t X = a.res[0], Y = a.res[1], Z = a.res[2], XY = X * Y;
  if (ne < 0 || nc < 0 || (ne && !h_edges) || (nc && !h_cells) || ne > (int64_t)INT32_MAX ||
      nc > (int64_t)UINT32_MAX / 5) {
    set_error("qj_nlmc_resume: bad edge or cell arrays");
    return QJ_ERR_INVALID;
  }
  // every flag in 1..254 has at least one triangle, so active cells need edges
  if (nc > 0 && ne == 0) {
    set_error("qj_nlmc_resume: %lld active cells but no edges", (long long)nc);
    return QJ_ERR_INVALID;
  }
  // host checks: every edge inside the grid with its far end, keys strictly
  // ascending (the cache's vertex order); every cell inside the grid
  std::vector<int64_t> keys((size_t)ne);
  for (int64_t i = 0; i < ne; i++) {
    const int32_t x = h_edges[4 * i], y = h_edges[4 * i + 1], z = h_edges[4 * i + 2], ax = h_edges[4 * i + 3];
    if (x < 0 || y < 0 || z < 0 || ax < 0 || ax > 2 || x + (ax == 0) >= X || y + (ax == 1) >= Y ||
        z + (ax == 2) >= Z) {
      set_error("qj_nlmc_resume: edge %lld lies outside the grid", (long long)i);
      return QJ_ERR_INVALID;
    }
    keys[i] = ((int64_t)z * XY + (int64_t)y * X + x) * 3 + ax;
    if (i && keys[i] <= keys[i - 1]) {
      set_error("qj_nlmc_resume: edges must be in vertex order (lower endpoint, then axis)");
      return QJ_ERR_INVALID;
    }
  }
  for (int64_t i = 0; i < nc; i++) {
    const int32_t x = h_cells[4 * i], y = h_cells[4 * i + 1], z = h_cells[4 * i + 2], fl = h_cells[4 * i + 3];
    if (x < 0 || y < 0 || z < 0 || x >= X - 1 || y >= Y - 1 || z >= Z - 1 || fl < 1 || fl > 254) {
      set_error("qj_nlmc_resume: cell %lld lies outside the grid or has flag %d", (long long)i, (int)fl);
      return QJ_ERR_INVALID;
    }
  }
  qj_status st;
  if ((st = grow_keep(m.verts, m.cap_v, 3 * ne, 0, s)) || (st = grow_keep(m.c_edges, m.cap_ce, 4 * ne, 0, s)) ||
      (st = grow_keep(m.c_cells, m.cap_cc, 4 * nc, 0, s)) || (st = grow(m.tri_off, m.cap_toff, nc, s)))
    return st;
  if (!m.totals) QJ_HIP(hipMalloc(&m.totals, 4 * sizeof(uint32_t)));
  if (!m.d_bad) QJ_HIP(hipMalloc(&m.d_bad, sizeof(unsigned int)));
  int64_t* d_keys = nullptr;
  unsigned int* d_missing = nullptr;
  QJ_HIP(hipMalloc(&d_keys, sizeof(int64_t) * (size_t)(ne > 0 ? ne : 1)));
  QJ_HIP(hipMalloc(&d_missing, sizeof(unsigned int)));
  struct Free {
    void* a;
    void* b;
    ~Free() {
      (void)hipFree(a);
      (void)hipFree(b);
    }
  } guard{d_keys, d_missing};
  QJ_HIP(hipMemsetAsync(m.d_bad, 0, sizeof(unsigned int), s));
  QJ_HIP(hipMemsetAsync(d_missing, 0, sizeof(unsigned int), s));
  if (ne) {
    QJ_HIP(hipMemcpyAsync(d_keys, keys.data(), sizeof(int64_t) * ne, hipMemcpyHostToDevice, s));
    QJ_HIP(hipMemcpyAsync(m.c_edges, h_edges, sizeof(int32_t) * 4 * ne, hipMemcpyHostToDevice, s));
    launch_cache_vertices(a, dim3(nblocks(ne)), s, f->d64, iso, m.c_edges, ne, m.verts, m.d_bad);
    QJ_HIP(hipGetLastError());
  }
  int64_t nt = 0;
  if (nc) {
    QJ_HIP(hipMemcpyAsync(m.c_cells, h_cells, sizeof(int32_t) * 4 * nc, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_cache_tricount, dim3(nblocks(nc)), dim3(kMcBlock), 0, s, m.c_cells, nc, m.tri_off);
    QJ_HIP(hipGetLastError());
    const size_t need_tmp = scan_tmp_bytes(nc > 4096 ? nc : 4096);
    if (need_tmp > m.tmp_bytes) {
      if (m.tmp) {
        QJ_HIP(hipStreamSynchronize(s));
        QJ_HIP(hipFree(m.tmp));
      }
      m.tmp = nullptr;
      m.tmp_bytes = 0;
      QJ_HIP(hipMalloc(&m.tmp, need_tmp));
      m.tmp_bytes = need_tmp;
    }
    if ((st = scan_u32(m.tri_off, m.tri_off, nc, m.totals + 2, m.tmp, m.tmp_bytes, s))) return st;
    uint32_t h = 0;
    QJ_HIP(hipMemcpyAsync(&h, m.totals + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    QJ_HIP(hipStreamSynchronize(s));
    const int64_t ntp = h;
    if ((st = grow(m.tris_pre, m.cap_tp, 3 * ntp, s)) || (st = grow(m.keep, m.cap_keep, ntp, s)) ||
        (st = grow(m.tris, m.cap_t, 3 * ntp, s)))
      return st;
    if (scan_tmp_bytes(ntp) > m.tmp_bytes) {
      QJ_HIP(hipStreamSynchronize(s));
      QJ_HIP(hipFree(m.tmp));
      m.tmp = nullptr;
      m.tmp_bytes = 0;
      QJ_HIP(hipMalloc(&m.tmp, scan_tmp_bytes(ntp)));
      m.tmp_bytes = scan_tmp_bytes(ntp);
    }
    if (ntp > 0) {
      hipLaunchKernelGGL(k_cache_triangles, dim3(nblocks(nc)), dim3(kMcBlock), 0, s, m.c_cells, nc, X, XY, d_keys,
                         ne, m.verts, m.tri_off, m.tris_pre, m.keep, d_missing);
      QJ_HIP(hipGetLastError());
      if ((st = scan_u32(m.keep, m.keep, ntp, m.totals + 3, m.tmp, m.tmp_bytes, s))) return st;
      QJ_HIP(hipMemcpyAsync(&h, m.totals + 3, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      QJ_HIP(hipStreamSynchronize(s));
      nt = h;
      if (nt > 0) {
        hipLaunchKernelGGL(k_mc_squeeze, dim3(nblocks(ntp)), dim3(kMcBlock), 0, s, m.tris_pre, m.keep, ntp,
                           m.totals + 3, m.tris);
        QJ_HIP(hipGetLastError());
      }
    }
  }
  unsigned int bad = 0, missing = 0;
  QJ_HIP(hipMemcpyAsync(&bad, m.d_bad, sizeof(bad), hipMemcpyDeviceToHost, s));
  QJ_HIP(hipMemcpyAsync(&missing, d_missing, sizeof(missing), hipMemcpyDeviceToHost, s));
  QJ_HIP(hipStreamSynchronize(s));
  m.nv = ne;
  m.nt = nt;
  m.ncells = nc;
  m.cache_valid = true;
  *nv_out = m.nv;
  *nt_out = m.nt;
  if (missing) {
    set_error("qj_nlmc_resume: %u triangles reference edges the cache does not list", missing);
    return QJ_ERR_INVALID;
  }
  if (bad) {
    set_error("non-linear mesher: %u edges whose end values do not straddle the isosurface "
              "(the reference exits at TRIANGLE_MESH.cpp:2082-2097)", bad);
    return QJ_ERR_NUMERIC;
  }
  return QJ_OK;
}

}  // namespace qj
