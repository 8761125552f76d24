// MLS-MPM growth step and its hand-written adjoint for gfx950 (MI355X).
//
// Per forward step (transfer.step, python/morphompm/transfer.py:33-79):
//   prep     : stress, grown volume and the APIC payload of each particle
//              (slot order), its bin key and the bin counts -- done by the
//              previous step's G2P when steps are chained;
//   sort     : counting sort over this step's occupied bins (integer atomics
//              only, for counts and places in a bin) -> per-bin sort by
//              slot, so the order inside every bin is fixed;
//   p2g tile : one thread per bin sums its particles' 27 stencil contributions
//              in registers, then adds them into its wave's LDS node tile in
//              27 rounds (one writer per node per round); the waves' tiles
//              are summed in wave order: a cell-binned segmented reduction
//              with no float atomics;
//   grid     : each node sums the <= 8 tile partials covering it in a fixed
//              order, then applies v = (1 - damping) mom / m;
//   g2p      : 27-node gather, new v, C, x, F = (I + dt C) F.
// Per reverse step (transfer.step_vjp, python/morphompm/transfer.py:82-158):
//   prep_adj : G2P-adjoint payload (nv_bar, C_bar') in sorted order;
//   p2g tile : the G2P-adjoint scatter as the same segmented reduction;
//   grid_adj : gmom_bar = s gv_bar / m, gm_bar = -(gv_bar . gv) / m;
//   adj      : per particle, in sorted order: 27-node gather of (gmom_bar,
//              gm_bar, gv) and the advection terms, then the constitutive VJP,
//              with the gathered values kept in registers.
// Every sum has a fixed order, so results are bitwise reproducible.
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "../../include/morphompm_hip.h"
#include "mpm_internal.h"
#include "mpm_math.h"

using namespace mpm;

namespace mpm_internal {

MPM_HIDDEN thread_local char g_err[512] = "";

int set_err(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(MPM_ELAUNCH, "%s: %s", what, hipGetErrorString(e));
  return MPM_OK;
}

}  // namespace mpm_internal

namespace {

using mpm_internal::check_launch;
using mpm_internal::g_err;
using mpm_internal::set_err;

constexpr int kChunk = MPM_SLOT_ALIGN;  // slots per instance-table entry
constexpr int kPayload = 15;            // fx[3], Q[3], A'[9]
constexpr int kRecord = 16;             // payload record, padded to 4 quads
constexpr int kScanBlock = 256;
constexpr int kScanPer = 4;             // elements per thread in the bin scan

// Per-instance constants in working precision.
template <typename T>
struct Inst {
  int64_t p_off, g_off, t_off;
  int n, nx, ny, nz, tx, ty, tz, sx, sy, sz, ox, oy, oz;
  T dx, inv_dx, dt, s, pmass, vol0, k4;
  ModelParams<T> mp;
};

// MODEL >= 0 fixes the constitutive model at compile time (a batch with one
// model); -1 reads it from the instance table.
template <typename T, int MODEL = -1>
__device__ __forceinline__ Inst<T> load_inst(const mpm_instance* inst, int i) {
  const mpm_instance& I = inst[i];
  Inst<T> r;
  r.p_off = I.p_off; r.g_off = I.g_off; r.t_off = I.t_off;
  r.n = I.n; r.nx = I.nx; r.ny = I.ny; r.nz = I.nz; r.tx = I.tx; r.ty = I.ty; r.tz = I.tz;
  r.sx = I.shift[0]; r.sy = I.shift[1]; r.sz = I.shift[2];
  r.ox = I.origin[0]; r.oy = I.origin[1]; r.oz = I.origin[2];
  r.dx = (T)I.dx;
  r.inv_dx = (T)(1.0 / I.dx);
  r.dt = (T)I.dt;
  r.s = (T)(1.0 - I.damping);
  r.pmass = (T)I.pmass;
  r.vol0 = (T)I.vol0;
  r.k4 = (T)(4.0 / (I.dx * I.dx));
  r.mp.kind = MODEL >= 0 ? MODEL : I.model;
  for (int k = 0; k < 6; ++k) r.mp.p[k] = (T)I.prm[k];
  return r;
}

template <typename T>
__device__ __forceinline__ void ld3(const T* base, int64_t stride, int64_t p, int row, T v[3]) {
  for (int k = 0; k < 3; ++k) v[k] = base[(int64_t)(row + k) * stride + p];
}
template <typename T>
__device__ __forceinline__ M3<T> ld9(const T* base, int64_t stride, int64_t p, int row) {
  M3<T> m;
  for (int k = 0; k < 9; ++k) m.a[k] = base[(int64_t)(row + k) * stride + p];
  return m;
}
template <typename T>
__device__ __forceinline__ void st3(T* base, int64_t stride, int64_t p, int row, const T v[3]) {
  for (int k = 0; k < 3; ++k) base[(int64_t)(row + k) * stride + p] = v[k];
}
template <typename T>
__device__ __forceinline__ void st9(T* base, int64_t stride, int64_t p, int row, const M3<T>& m) {
  for (int k = 0; k < 9; ++k) base[(int64_t)(row + k) * stride + p] = m.a[k];
}

// Non-temporal variants for rows a kernel reads or writes once in a stream:
// they do not displace the grid nodes the same kernel gathers from the caches.
template <typename T>
__device__ __forceinline__ void ld3_nt(const T* base, int64_t stride, int64_t p, int row,
                                       T v[3]) {
  for (int k = 0; k < 3; ++k)
    v[k] = __builtin_nontemporal_load(&base[(int64_t)(row + k) * stride + p]);
}
template <typename T>
__device__ __forceinline__ M3<T> ld9_nt(const T* base, int64_t stride, int64_t p, int row) {
  M3<T> m;
  for (int k = 0; k < 9; ++k)
    m.a[k] = __builtin_nontemporal_load(&base[(int64_t)(row + k) * stride + p]);
  return m;
}
template <typename T>
__device__ __forceinline__ void st3_nt(T* base, int64_t stride, int64_t p, int row,
                                       const T v[3]) {
  for (int k = 0; k < 3; ++k)
    __builtin_nontemporal_store(v[k], &base[(int64_t)(row + k) * stride + p]);
}
template <typename T>
__device__ __forceinline__ void st9_nt(T* base, int64_t stride, int64_t p, int row,
                                       const M3<T>& m) {
  for (int k = 0; k < 9; ++k)
    __builtin_nontemporal_store(m.a[k], &base[(int64_t)(row + k) * stride + p]);
}

// 4-wide records (grid nodes, payload quads): one vector load per record.
template <typename T> struct V4 { T x, y, z, w; };
__device__ __forceinline__ V4<float> ld4(const float* p) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  return {v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ V4<double> ld4(const double* p) {
  const double2 a = *reinterpret_cast<const double2*>(p);
  const double2 b = *reinterpret_cast<const double2*>(p + 2);
  return {a.x, a.y, b.x, b.y};
}
__device__ __forceinline__ void st4(float* p, float x, float y, float z, float w) {
  *reinterpret_cast<float4*>(p) = make_float4(x, y, z, w);
}
__device__ __forceinline__ void st4(double* p, double x, double y, double z, double w) {
  *reinterpret_cast<double2*>(p) = make_double2(x, y);
  *reinterpret_cast<double2*>(p + 2) = make_double2(z, w);
}

// Stencil nodes of a particle along each axis, clamped into the grid, with a
// validity flag: out-of-grid nodes contribute nothing (transfer.py:51).
struct AxisNodes {
  int idx[3];
  bool ok[3];
};
__device__ __forceinline__ AxisNodes axis_nodes(int base, int n) {
  AxisNodes a;
  for (int i = 0; i < 3; ++i) {
    const int v = base + i;
    a.ok[i] = v >= 0 && v < n;
    a.idx[i] = v < 0 ? 0 : (v >= n ? n - 1 : v);
  }
  return a;
}

// Growth tensor of this step: Fg = I + a (G - I); a == 1 keeps G exactly.
template <typename T>
__device__ __forceinline__ M3<T> grown(const M3<T>& G, T a) {
  if (a == T(1)) return G;
  M3<T> Fg;
  for (int k = 0; k < 9; ++k) Fg.a[k] = a * (G.a[k] - ((k % 4 == 0) ? T(1) : T(0)));
  Fg.a[0] += T(1); Fg.a[4] += T(1); Fg.a[8] += T(1);
  return Fg;
}

// Stencil base node (in grid nodes) of a position stored relative to the
// instance origin: base = floor(x/dx - 1/2) + origin.
template <typename T>
__device__ __forceinline__ void stencil_base(const T x[3], T inv_dx, const int origin[3],
                                             int base[3], T fx[3]) {
  for (int a = 0; a < 3; ++a) base[a] = cell_base(x[a], inv_dx, fx[a]) + origin[a];
}
template <typename T>
__device__ __forceinline__ void stencil_base(const T x[3], const Inst<T>& I, int base[3], T fx[3]) {
  const int o[3] = {I.ox, I.oy, I.oz};
  stencil_base(x, I.inv_dx, o, base, fx);
}

// Instance of this block: blocks never straddle a 256-slot chunk (block sizes
// divide 256), so the index is block-uniform and lives in a scalar register.
__device__ __forceinline__ int block_inst(const int32_t* chunk_inst) {
  const int64_t first = (int64_t)blockIdx.x * blockDim.x;
  return __builtin_amdgcn_readfirstlane(chunk_inst[first / kChunk]);
}

// First slot of this block.  chunks == NULL: blocks cover the slots in
// order; else block b works in the 256-slot chunk chunks[b / (256 / blockDim)]
// (a launch over one model's chunks).
__device__ __forceinline__ int64_t block_first_slot(const int32_t* chunks) {
  const int64_t first = (int64_t)blockIdx.x * blockDim.x;
  if (chunks == nullptr) return first;
  return (int64_t)__builtin_amdgcn_readfirstlane(chunks[first / kChunk]) * kChunk + first % kChunk;
}

// dyn row offsets
constexpr int RX = 0, RV = 3, RC = 6, RF = 15;

// ---------------------------------------------------------------------------
// sort
//
// Particles are binned by cell (base+2 per axis, (nx+2)(ny+2)(nz+2) cells per
// instance).  Cells are grouped into P2G tiles of TB = TX*TY*TZ cells, and
// every instance owns tx*ty*tz tile units followed by one dead unit (particles
// with no in-grid node, then the padding slots).  Per step only the occupied
// units get bins: unit_rank[u] is the compact index of an occupied unit (-1
// otherwise), active_unit[a] its inverse and n_active their number; a
// particle's key is unit_rank[unit] * TB + its cell's place in the tile.  Units
// are ranked in order, so each instance's sorted range is its own slot range.
// ---------------------------------------------------------------------------

// Unit (tile) of in-grid bin (bx, by, bz) and the bin's place in the tile.
// Every built tile shape has power-of-two sides and 0 <= shift < tile, so the
// shifted bin is non-negative and / and % by a side are a shift and a mask.
__device__ __forceinline__ void bin_unit(const mpm_instance& I, int bx, int by, int bz, int TX,
                                         int TY, int TZ, int64_t& unit, int& local) {
  bx += I.shift[0];
  by += I.shift[1];
  bz += I.shift[2];
  const int itx = bx >> __builtin_ctz(TX), ity = by >> __builtin_ctz(TY),
            itz = bz >> __builtin_ctz(TZ);
  unit = I.t_off + ((int64_t)itx * I.ty + ity) * I.tz + itz;
  local = ((bx & (TX - 1)) * TY + (by & (TY - 1))) * TZ + (bz & (TZ - 1));
}

__device__ __forceinline__ int64_t dead_unit(const mpm_instance& I) {
  return I.t_off + (int64_t)I.tx * I.ty * I.tz;
}

// Units are ranked in two passes over 256-unit blocks: k_unit_count counts
// each block's live units (flagged, or the dead unit -- the last unit of its
// instance), k_unit_rank adds the counts of the blocks before it, ranks the
// live units with a ballot, writes unit_rank, active_unit and n_active, and
// clears the flags for the next step.
__device__ __forceinline__ int32_t unit_live(const int32_t* unit_inst, int64_t n_units,
                                             const int32_t* unit_flag, int64_t u) {
  if (u >= n_units) return 0;
  const bool dead = (u + 1 == n_units) || (unit_inst[u + 1] != unit_inst[u]);
  return (unit_flag[u] != 0 || dead) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_unit_count(const int32_t* unit_inst, int64_t n_units,
                                                    const int32_t* unit_flag, int32_t* ucount) {
  __shared__ int32_t wsum[4];
  const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const unsigned long long m = __ballot(unit_live(unit_inst, n_units, unit_flag, u));
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) ucount[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Sum of v[0, end) by the 256 threads of a block (result valid in every thread).
__device__ __forceinline__ int32_t block_prefix(const int32_t* v, int64_t end, int32_t* red) {
  int32_t s = 0;
  for (int64_t i = threadIdx.x; i < end; i += 256) s += v[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  const int32_t r = red[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void k_unit_rank(const int32_t* unit_inst, int64_t n_units,
                                                   int32_t* unit_flag, const int32_t* ucount,
                                                   int32_t* unit_rank, int32_t* active_unit,
                                                   int32_t* n_active) {
  __shared__ int32_t red[256];
  __shared__ int32_t wsum[4];
  const int32_t base = block_prefix(ucount, blockIdx.x, red);
  const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int32_t f = unit_live(unit_inst, n_units, unit_flag, u);
  const unsigned long long m = __ballot(f);
  if (lane == 0) wsum[wave] = __popcll(m);
  __syncthreads();
  int32_t r = base + __popcll(m & ((1ull << lane) - 1));
  for (int w = 0; w < wave; ++w) r += wsum[w];
  if (u < n_units) {
    unit_rank[u] = f ? r : -1;
    if (f) active_unit[r] = (int32_t)u;
    unit_flag[u] = 0;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
    *n_active = base + wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Bin counts are kept per unit-level key (unit * TB + place) by the kernels
// that compute the keys (no-return atomics: the producer never waits on
// them); the scan reads the counts of this step's occupied units only.  A
// wave adds one atomic per run of equal keys in adjacent lanes (particles
// arrive nearly sorted).  Every lane of the wave must call these; k < 0
// counts nothing.
__device__ __forceinline__ void run_heads(int32_t k, bool& head, int& my_head, int& run) {
  const int lane = threadIdx.x & 63;
  const int32_t prev = __shfl_up(k, 1);
  head = (lane == 0) || (prev != k);
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = heads & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
  my_head = 63 - __clzll(upto);
  const unsigned long long after = (lane == 63) ? 0ull : (heads >> (lane + 1));
  run = (after ? lane + __ffsll(after) : 64) - lane;
}

__device__ __forceinline__ void count_keys(int32_t k, int32_t* count) {
  bool head;
  int my_head, run;
  run_heads(k, head, my_head, run);
  if (head && k >= 0) atomicAdd(&count[k], run);
}

// Clears the bin counts and unit flags of the whole batch (before keys are
// computed from scratch: a chained G2P whose counts no sort consumed may have
// left some behind).
__global__ __launch_bounds__(256) void k_clear_bins(int32_t* count, int64_t n_bins,
                                                    int32_t* unit_flag, int64_t n_units) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n_bins) count[i] = 0;
  if (i < n_units) unit_flag[i] = 0;
}

// The dead unit's bin 0 holds live particles with no in-grid node and its bin
// 1 the padding slots, so each instance's sorted range is exactly its own
// slot range [p_off, p_off + padded count): padding slots keep key -1, take
// the tail of that range in slot order (k_place) and no place from the
// cursor; the first padding slot (local == n) sets bin 1's count once.
__device__ __forceinline__ void count_padding(const mpm_instance& I, int64_t local, int TB,
                                              int32_t* count) {
  if (local != I.n) return;
  const int64_t padded = ((int64_t)I.n + kChunk - 1) / kChunk * kChunk;
  atomicAdd(&count[dead_unit(I) * TB + 1], (int32_t)((padded > 0 ? padded : kChunk) - I.n));
}

// Stores the unit-level key of slot p and counts it.
__device__ __forceinline__ void put_key(int64_t p, int32_t k, int32_t* key, int32_t* count) {
  count_keys(k, count);
  key[p] = k;
}

// Exclusive scan of the counts of this step's n = n_active * TB bins (bin
// a * TB + l is place l of unit active_unit[a]) -> start[n + 1], in two
// passes; the counts are cleared for the next step.
__global__ void k_scan_blocks(int32_t* count, const int32_t* active_unit, int32_t* start,
                              int32_t* cursor, int32_t* bsum, const int32_t* n_active, int TB) {
  __shared__ int32_t sh[kScanBlock];
  const int64_t n = (int64_t)__builtin_amdgcn_readfirstlane(*n_active) * TB;
  const int64_t i0 = ((int64_t)blockIdx.x * kScanBlock + threadIdx.x) * kScanPer;
  if ((int64_t)blockIdx.x * kScanBlock * kScanPer >= n) return;
  static_assert(kScanPer == 4, "one 16-byte quad of bins per thread");
  int32_t v[kScanPer] = {0, 0, 0, 0};
  if (i0 < n) {
    // TB (a product of power-of-two tile sides) is a multiple of 4 and so is
    // n, so the thread's 4 bins are places l0 .. l0 + 3 of one unit
    const int32_t a = (int32_t)i0 / TB, l0 = (int32_t)i0 - a * TB;
    int32_t* cp = count + (int64_t)active_unit[a] * TB + l0;
    const int4 c4 = *reinterpret_cast<const int4*>(cp);
    v[0] = c4.x;
    v[1] = c4.y;
    v[2] = c4.z;
    v[3] = c4.w;
    *reinterpret_cast<int4*>(cp) = make_int4(0, 0, 0, 0);
    *reinterpret_cast<int4*>(cursor + i0) = make_int4(0, 0, 0, 0);
  }
  const int32_t tot = v[0] + v[1] + v[2] + v[3];
  sh[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int32_t add = (threadIdx.x >= (unsigned)off) ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += add;
    __syncthreads();
  }
  int32_t run = sh[threadIdx.x] - tot;  // exclusive prefix of this thread
  for (int k = 0; k < kScanPer; ++k) {
    if (i0 + k < n) start[i0 + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kScanBlock - 1) bsum[blockIdx.x] = sh[threadIdx.x];
}

// Adds the bins of the scan blocks before this one (bsum) to the block's
// local prefixes; the last block also writes the total start[n].
__global__ __launch_bounds__(kScanBlock) void k_scan_add(int32_t* start, const int32_t* bsum,
                                                         const int32_t* n_active, int TB) {
  __shared__ int32_t red[kScanBlock];
  const int64_t n = (int64_t)__builtin_amdgcn_readfirstlane(*n_active) * TB;
  const int64_t blk = blockIdx.x;
  if (blk * kScanBlock * kScanPer >= n) return;
  const int32_t base = block_prefix(bsum, blk, red);
  const int64_t i0 = (blk * kScanBlock + threadIdx.x) * kScanPer;
  for (int k = 0; k < kScanPer; ++k)
    if (i0 + k < n) start[i0 + k] += base;
  if ((blk + 1) * kScanBlock * kScanPer >= n && threadIdx.x == 0) start[n] = base + bsum[blk];
}

// Bins in arbitrary order within each bin: placed[start[k] + place] = p,
// with k the compact key unit_rank[unit] * TB + place in tile, written back
// to key, and the place in the bin from an atomic cursor (one atomic per run
// of equal keys in adjacent lanes).  Each wave takes kPlaceGroups runs of 64
// slots, so their loads and atomics are in flight together.  A key outside
// the occupied units (only possible when a chained step is told its keys are
// ready and they are not) is dropped: wrong, but every index stays in bounds.
constexpr int kPlaceGroups = 4;
__global__ __launch_bounds__(256) void k_place(const int32_t* unit_rank, int TB, int64_t n_units,
                                               int32_t* key, const int32_t* start,
                                               int32_t* cursor, int32_t* placed,
                                               int64_t n_slots) {
  constexpr int G = kPlaceGroups;
  const int lane = threadIdx.x & 63;
  const int64_t w0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * (64 * G);
  if (w0 >= n_slots) return;  // whole waves (n_slots is a multiple of 256)
  int32_t raw[G], k[G], s0[G], my_head[G], base[G];
  const int tb_shift = __builtin_ctz(TB);  // TB is a power of two
#pragma unroll
  for (int i = 0; i < G; ++i) raw[i] = key[w0 + i * 64 + lane];
#pragma unroll
  for (int i = 0; i < G; ++i) {
    k[i] = -1;
    if (raw[i] >= 0) {
      const int32_t u = raw[i] >> tb_shift;
      const int32_t ur = u < n_units ? unit_rank[u] : -1;
      if (ur >= 0) k[i] = ur * TB + (raw[i] & (TB - 1));
    }
    s0[i] = k[i] >= 0 ? start[k[i]] : 0;
  }
#pragma unroll
  for (int i = 0; i < G; ++i) {
    bool head;
    int run;
    run_heads(k[i], head, my_head[i], run);
    base[i] = 0;
    if (head && k[i] >= 0) base[i] = atomicAdd(&cursor[k[i]], run);
  }
#pragma unroll
  for (int i = 0; i < G; ++i) {
    const int64_t p = w0 + i * 64 + lane;
    const int32_t place = __shfl(base[i], my_head[i]) + (lane - my_head[i]);
    if (raw[i] < 0) {
      placed[p] = (int32_t)p;  // padding slot: the tail of its instance's range, in order
      continue;
    }
    key[p] = k[i];
    if (k[i] < 0) continue;
    const int64_t pos = (int64_t)s0[i] + place;
    if (pos >= 0 && pos < n_slots) placed[pos] = (int32_t)p;
  }
}

// Stable order within each bin: the particle at placed position j goes to
// start[k] + (number of its bin's particles with a smaller slot), so every
// bin ends up ascending by slot; rank is the inverse permutation.  Work per
// bin is quadratic in its size, with independent loads, so a bin of more
// than kStableBin particles (no physical cell holds that many; a state that
// has blown up can put everything in one bin) keeps its placement order.
constexpr int32_t kStableBin = 2048;
constexpr int kRankGroups = 4;  // positions per thread, their first loads in flight together
__global__ __launch_bounds__(256) void k_bin_rank(const int32_t* key, const int32_t* start,
                                                  const int32_t* placed, int32_t* sorted,
                                                  int32_t* rank, int64_t n_slots) {
  constexpr int G = kRankGroups;
  const int64_t j0 = (int64_t)blockIdx.x * (256 * G) + threadIdx.x;
  int32_t v[G], k[G], s0[G], s1[G];
#pragma unroll
  for (int g = 0; g < G; ++g) v[g] = j0 + g * 256 < n_slots ? placed[j0 + g * 256] : 0;
#pragma unroll
  for (int g = 0; g < G; ++g) k[g] = j0 + g * 256 < n_slots ? key[v[g]] : -1;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    s0[g] = k[g] >= 0 ? start[k[g]] : 0;
    s1[g] = k[g] >= 0 ? start[k[g] + 1] : 0;
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int64_t j = j0 + g * 256;
    if (j >= n_slots) continue;
    int32_t r = (int32_t)j;
    if (k[g] >= 0 && s1[g] - s0[g] <= kStableBin) {
      r = s0[g];
#pragma unroll 4
      for (int32_t i = s0[g]; i < s1[g]; ++i) r += placed[i] < v[g];
    }
    sorted[r] = v[g];
    rank[v[g]] = r;
  }
}

// Record slot 15 carries the particle's cell as bits: (base+2) per axis in 10
// bits each, or all ones when no stencil node is in the grid.
constexpr uint32_t kNoCell = 0xffffffffu;
__device__ __forceinline__ uint32_t pack_cell(const int base[3], int nx, int ny, int nz) {
  const int bx = base[0] + 2, by = base[1] + 2, bz = base[2] + 2;
  if (bx < 0 || bx >= nx + 2 || by < 0 || by >= ny + 2 || bz < 0 || bz >= nz + 2) return kNoCell;
  return (uint32_t)bx | ((uint32_t)by << 10) | ((uint32_t)bz << 20);
}
__device__ __forceinline__ float cell_as(float, uint32_t c) { return __uint_as_float(c); }
__device__ __forceinline__ double cell_as(double, uint32_t c) {
  return __longlong_as_double((long long)c);
}
__device__ __forceinline__ uint32_t cell_of(float v) { return __float_as_uint(v); }
__device__ __forceinline__ uint32_t cell_of(double v) { return (uint32_t)__double_as_longlong(v); }

template <typename T>
__device__ __forceinline__ void store_payload(T* r, const T fx[3], const T q[3], const M3<T>& A,
                                              uint32_t cell) {
  st4(r, fx[0], fx[1], fx[2], q[0]);
  st4(r + 4, q[1], q[2], A.a[0], A.a[1]);
  st4(r + 8, A.a[2], A.a[3], A.a[4], A.a[5]);
  st4(r + 12, A.a[6], A.a[7], A.a[8], cell_as(T(0), cell));
}

// Sort key of a particle with stencil base `base`, the input of the next
// sort: unit * TB + place in the unit, or dead unit * TB when no stencil node
// is in the grid.  Flags the unit as occupied.
__device__ __forceinline__ int32_t cell_key(const mpm_instance& I, const int base[3], int TX,
                                            int TY, int TZ, int32_t* unit_flag) {
  const int TB = TX * TY * TZ;
  const int bx = base[0] + 2, by = base[1] + 2, bz = base[2] + 2;
  if (bx < 0 || bx >= I.nx + 2 || by < 0 || by >= I.ny + 2 || bz < 0 || bz >= I.nz + 2)
    return (int32_t)(dead_unit(I) * TB);
  int64_t unit;
  int local;
  bin_unit(I, bx, by, bz, TX, TY, TZ, unit, local);
  unit_flag[unit] = 1;
  return (int32_t)(unit * TB + local);
}

// ---------------------------------------------------------------------------
// forward prep: stress, grown volume, APIC payload (transfer.py:43-45,53)
//   m_i += w pmass;  mom_i += w (P + A dpos),  P = pmass v,  A = pmass C + dt c tau,
//   c = -4 inv_dx^2 vol0 det Fg,  dpos = (o - fx) dx  ->  A dpos = A' (o - fx),
//   A' = A dx;  stored: fx, Q = P - A' fx, A'.  Payload records are in slot
//   order; P2G reads them through the sort.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void make_payload(const Inst<T>& I, const T x[3], const T v[3],
                                             const M3<T>& C, const M3<T>& F, const M3<T>& Fg,
                                             T* rec, int base[3]) {
  T fx[3];
  stencil_base(x, I, base, fx);
  const M3<T> tau = kirchhoff(I.mp, F, Fg, C);
  const T c = -I.k4 * I.vol0 * det(Fg);
  M3<T> Ap;
  for (int k = 0; k < 9; ++k) Ap.a[k] = (I.pmass * C.a[k] + I.dt * (c * tau.a[k])) * I.dx;
  T q[3];
  for (int a = 0; a < 3; ++a)
    q[a] = I.pmass * v[a] - (Ap(a, 0) * fx[0] + Ap(a, 1) * fx[1] + Ap(a, 2) * fx[2]);
  store_payload(rec, fx, q, Ap, pack_cell(base, I.nx, I.ny, I.nz));
}

// Payload, sort key, bin count and unit flag of a step from its input state.
// (Whole waves only leave early: count_in_bin needs every lane.)
template <typename T, int MODEL>
__global__ __launch_bounds__(256) void k_prep_fwd(const mpm_instance* inst,
                                                  const int32_t* slot_inst, int64_t n_slots,
                                                  const T* dyn, const T* Fgp, const T* sched,
                                                  int TX, int TY, int TZ, T* payload, int32_t* key,
                                                  int32_t* count, int32_t* unit_flag,
                                                  const int32_t* chunks) {
  const int64_t p0 = block_first_slot(chunks), p = p0 + threadIdx.x;
  if (p0 >= n_slots) return;
  const int ii = __builtin_amdgcn_readfirstlane(slot_inst[p0 / kChunk]);
  const Inst<T> I = load_inst<T, MODEL>(inst, ii);
  const int64_t local = p - I.p_off;
  int32_t k = -1;
  if (local < I.n) {
    T x[3], v[3];
    ld3(dyn, n_slots, p, RX, x);
    ld3(dyn, n_slots, p, RV, v);
    const M3<T> C = ld9(dyn, n_slots, p, RC);
    const M3<T> F = ld9(dyn, n_slots, p, RF);
    const M3<T> Fg = grown(ld9(Fgp, n_slots, p, 0), sched ? sched[ii] : T(1));
    int base[3];
    make_payload(I, x, v, C, F, Fg, payload + p * kRecord, base);
    k = cell_key(inst[ii], base, TX, TY, TZ, unit_flag);
  } else {
    count_padding(inst[ii], local, TX * TY * TZ, count);
  }
  put_key(p, k, key, count);
}

// Sort keys, bin counts and unit flags from positions only (mpm_sort_by_cell).
template <typename T>
__global__ __launch_bounds__(256) void k_cell_keys(const mpm_instance* inst,
                                                   const int32_t* slot_inst, int64_t n_slots,
                                                   const T* dyn, int TX, int TY, int TZ,
                                                   int32_t* key, int32_t* count,
                                                   int32_t* unit_flag) {
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x, p = p0 + threadIdx.x;
  if (p0 >= n_slots) return;
  const int ii = block_inst(slot_inst);
  const Inst<T> I = load_inst<T>(inst, ii);
  const int64_t local = p - I.p_off;
  int32_t k = -1;
  if (local < I.n) {
    T x[3], fx[3];
    int base[3];
    ld3(dyn, n_slots, p, RX, x);
    stencil_base(x, I, base, fx);
    k = cell_key(inst[ii], base, TX, TY, TZ, unit_flag);
  } else {
    count_padding(inst[ii], local, TX * TY * TZ, count);
  }
  put_key(p, k, key, count);
}

// ---------------------------------------------------------------------------
// P2G as a cell-binned segmented reduction into LDS node tiles.
// Tile = TX*TY*TZ bins (contiguous keys), one thread each; node tile
// (TX+2)(TY+2)(TZ+2), one per wave.  One block per occupied unit of this step
// (blocks past n_active and dead units return at once).  Each thread sums its
// bin's particles in sorted order, its wave adds the bins into the wave's
// node tile in 27 rounds (one stencil node per round), and the last wave to
// finish sums the waves' tiles in wave order: a fixed order throughout, no
// float atomics, no block barrier between rounds.  partial: [active unit][TN]
// node records (sum of weights (MASS), sum of w (Q + A' o)), one quad per
// node.
// ---------------------------------------------------------------------------
// INDIRECT: payload records in slot order, read through sorted[j]; else in
// sorted order.
constexpr int kTileIdxCap = 2048;   // staged slot indices per tile (8 KB of LDS)
template <typename T, int TX, int TY, int TZ, bool MASS, bool INDIRECT>
__global__ __launch_bounds__(TX* TY* TZ) void k_p2g_tile(const mpm_instance* inst,
                                                        const int32_t* unit_inst,
                                                        const int32_t* active_unit,
                                                        const int32_t* n_active,
                                                        const int32_t* bin_start,
                                                        const int32_t* sorted,
                                                        const T* payload, T* partial) {
  constexpr int NY = TY + 2, NZ = TZ + 2, TN = (TX + 2) * NY * NZ;
  constexpr int NT = TX * TY * TZ;
  constexpr int NW = NT / 64;  // waves; each has its own node tile
  static_assert(NT % 64 == 0, "whole waves");
  __shared__ V4<T> ldsw[NW][TN];
  __shared__ int waves_done;
  const int64_t a = blockIdx.x;
  if (a >= *n_active) return;
  const int32_t u = active_unit[a];
  const mpm_instance& I = inst[__builtin_amdgcn_readfirstlane(unit_inst[u])];
  if (u == dead_unit(I)) return;
  const int64_t key0 = a * NT;
  const int tid = threadIdx.x;
  const int lz = tid % TZ, ly = (tid / TZ) % TY, lx = tid / (TY * TZ);

  const int lane = tid & 63, wave = tid >> 6;
  V4<T>* lds = ldsw[wave];
  for (int i = lane; i < TN; i += 64) lds[i] = V4<T>{T(0), T(0), T(0), T(0)};
  if (tid == 0) waves_done = 0;
  // INDIRECT: the tile's particles are the contiguous sorted range [r0, r1);
  // their slot indices are staged in LDS with coalesced loads, so the record
  // loads do not wait on a dependent index load each.
  __shared__ int32_t sidx[INDIRECT ? kTileIdxCap : 1];
  const int32_t r0 = bin_start[key0];
  if (INDIRECT) {
    const int32_t n_idx = min(bin_start[key0 + NT] - r0, kTileIdxCap);
    for (int i = tid; i < n_idx; i += NT) sidx[i] = sorted[r0 + i];
  }
  __syncthreads();

  T aw[27], am[27][3];
#pragma unroll
  for (int o = 0; o < 27; ++o) {
    aw[o] = T(0);
    am[o][0] = am[o][1] = am[o][2] = T(0);
  }
  const int32_t s0 = bin_start[key0 + tid], s1 = bin_start[key0 + tid + 1];
  for (int32_t j = s0; j < s1; ++j) {
    const int32_t src = INDIRECT ? (j - r0 < kTileIdxCap ? sidx[j - r0] : sorted[j]) : j;
    const T* r = payload + (int64_t)src * kRecord;
    const V4<T> r0 = ld4(r), r1 = ld4(r + 4), r2 = ld4(r + 8), r3 = ld4(r + 12);
    // fx = r0.xyz, Q = (r0.w, r1.x, r1.y), A' row-major = r1.zw r2.xyzw r3.xyz
    const T q0 = r0.w, q1 = r1.x, q2 = r1.y;
    const T A00 = r1.z, A01 = r1.w, A02 = r2.x, A10 = r2.y, A11 = r2.z, A12 = r2.w,
            A20 = r3.x, A21 = r3.y, A22 = r3.z;
    T w0[3], w1[3], w2[3];
    bspline_w(r0.x, w0);
    bspline_w(r0.y, w1);
    bspline_w(r0.z, w2);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) {
        const T wij = w0[i] * w1[jj];
        const T u0 = q0 + T(i) * A00 + T(jj) * A01;
        const T u1 = q1 + T(i) * A10 + T(jj) * A11;
        const T u2 = q2 + T(i) * A20 + T(jj) * A21;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int o = (i * 3 + jj) * 3 + k;
          const T wt = wij * w2[k];
          if (MASS) aw[o] += wt;
          am[o][0] += wt * (u0 + T(k) * A02);
          am[o][1] += wt * (u1 + T(k) * A12);
          am[o][2] += wt * (u2 + T(k) * A22);
        }
      }
  }
  // Each wave adds its bins into its own node tile, 27 rounds with no block
  // barrier: a wave's LDS accesses complete in order, and the fences keep the
  // compiler from moving one round's accesses across another's.
#pragma unroll
  for (int o = 0; o < 27; ++o) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int nl = ((lx + o / 9) * NY + (ly + (o / 3) % 3)) * NZ + (lz + o % 3);
    V4<T> t = lds[nl];
    if (MASS) t.x += aw[o];
    t.y += am[o][0];
    t.z += am[o][1];
    t.w += am[o][2];
    lds[nl] = t;
  }
  // The last wave to finish sums the waves' tiles in wave order and writes
  // the partial; the others end here.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  int prev = 0;
  if (lane == 0) prev = __hip_atomic_fetch_add(&waves_done, 1, __ATOMIC_ACQ_REL,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
  prev = __shfl(prev, 0);
  if (prev != NW - 1) return;
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  T* out = partial + a * (int64_t)(4 * TN);
  for (int i = lane; i < TN; i += 64) {
    V4<T> t = ldsw[0][i];
#pragma unroll
    for (int w = 1; w < NW; ++w) {
      const V4<T> s2 = ldsw[w][i];
      t.x += s2.x;
      t.y += s2.y;
      t.z += s2.z;
      t.w += s2.w;
    }
    st4(out + 4 * i, t.x, t.y, t.z, t.w);
  }
}

// The tiles of side TS along one axis that cover node n (s = n + 2 + shift):
// tile ix covers nodes [ix TS - shift - 2, ix TS - shift + TS - 1], so node n
// lies in tile hi = s / TS at place lo = s - hi TS (0 <= lo < TS), and also in
// tile hi - 1 at place lo + TS when lo <= 1: at most two, in ascending order.
template <int TS>
__device__ __forceinline__ void tile_cover(int s, int t[2], int l[2], bool v[2]) {
  const int hi = s / TS, lo = s - hi * TS;
  t[0] = hi - 1;
  l[0] = lo + TS;
  v[0] = lo <= 1 && hi >= 1;
  t[1] = hi;
  l[1] = lo;
  v[1] = true;
}

// Sum of the occupied tile partials covering node (i, j, k), tiles in
// ascending order.  The (at most 8) unit ranks are loaded together, then the
// partial records, so the dependent loads of different tiles overlap.
template <typename T, int TX, int TY, int TZ>
__device__ __forceinline__ void gather_tiles(const Inst<T>& I, const int32_t* unit_rank,
                                             const T* partial, int i, int j, int k, T acc[4]) {
  constexpr int NY = TY + 2, NZ = TZ + 2, TN = (TX + 2) * NY * NZ;
  int ct[3][2], cl[3][2];
  bool cv[3][2];
  tile_cover<TX>(i + 2 + I.sx, ct[0], cl[0], cv[0]);
  tile_cover<TY>(j + 2 + I.sy, ct[1], cl[1], cv[1]);
  tile_cover<TZ>(k + 2 + I.sz, ct[2], cl[2], cv[2]);
  int32_t ar[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int x = c >> 2, y = (c >> 1) & 1, z = c & 1;
    ar[c] = -1;
    if (cv[0][x] && cv[1][y] && cv[2][z])
      ar[c] = unit_rank[I.t_off + ((int64_t)ct[0][x] * I.ty + ct[1][y]) * I.tz + ct[2][z]];
  }
  V4<T> pv[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int x = c >> 2, y = (c >> 1) & 1, z = c & 1;
    pv[c] = V4<T>{T(0), T(0), T(0), T(0)};
    if (ar[c] >= 0)
      pv[c] = ld4(partial + ar[c] * (int64_t)(4 * TN) + 4 * ((cl[0][x] * NY + cl[1][y]) * NZ + cl[2][z]));
  }
  acc[0] = acc[1] = acc[2] = acc[3] = T(0);
#pragma unroll
  for (int c = 0; c < 8; ++c)
    if (ar[c] >= 0) {
      acc[0] += pv[c].x;
      acc[1] += pv[c].y;
      acc[2] += pv[c].z;
      acc[3] += pv[c].w;
    }
}

// Node (i, j, k) of global node g in its instance's grid; false for the
// padding nodes past the grid.  An instance grid holds at most 1022^3 < 2^31
// nodes, so 32-bit division is exact (the 64-bit one costs several times more).
template <typename T>
__device__ __forceinline__ bool node_of(const Inst<T>& I, int64_t g, int& i, int& j, int& k) {
  const int64_t local = g - I.g_off;
  const int plane = I.ny * I.nz;
  if (local >= (int64_t)I.nx * plane) return false;
  const int l = (int)local;
  i = l / plane;
  const int r = l - i * plane;
  j = r / I.nz;
  k = r - j * I.nz;
  return true;
}

// Grid update (transfer.py:56-59): v = (1 - damping) mom / m where m > 1e-12.
// grid record per node: (m, vx, vy, vz).
template <typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_grid_fwd(const mpm_instance* inst, const int32_t* node_inst,
                                                  int64_t n_nodes, const int32_t* unit_rank,
                                                  const T* partial, T* grid) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_nodes) return;
  const Inst<T> I = load_inst<T>(inst, block_inst(node_inst));
  int i, j, k;
  if (!node_of(I, g, i, j, k)) return;
  T acc[4];
  gather_tiles<T, TX, TY, TZ>(I, unit_rank, partial, i, j, k, acc);
  const T m = I.pmass * acc[0];
  T v[3] = {T(0), T(0), T(0)};
  if (m > T(kMassEps))
    for (int a = 0; a < 3; ++a) v[a] = I.s * acc[1 + a] / m;
  st4(grid + 4 * g, m, v[0], v[1], v[2]);
}

// ---------------------------------------------------------------------------
// G2P (transfer.py:61-77)
// ---------------------------------------------------------------------------
// Stencil nodes of a sorted 256-slot chunk, staged in LDS.  Particles in
// sorted order are spatially compact, so the box spanned by the chunk's
// stencils (base .. base+2 per axis over its particles with an in-grid node)
// is small (a few hundred nodes).  When it exceeds the LDS capacity the
// chunk reads its nodes from the grid in global memory instead.  Nodes
// outside the grid read as zero either way.
template <typename T>
struct BoxCap {
  static constexpr int value = sizeof(T) == 4 ? 1024 : 640;
};
// FP32 capacities sized to each kernel's register occupancy, so the LDS of
// its resident blocks still fits in 160 KB: G2P (one box, six blocks of 256
// per CU) and the Hencky reverse kernel (two boxes, three blocks) take 1536
// nodes, the other reverse kernels (four blocks) 1024.  Fewer blocks fall
// back to reading the grid: 1.5 ms less per batched iteration.
template <typename T>
constexpr int kG2pBoxCap = sizeof(T) == 4 ? 1536 : BoxCap<T>::value;
template <typename T, int MODEL>
constexpr int kAdjBoxCap = (sizeof(T) == 4 && MODEL == MPM_HENCKY) ? 1536 : BoxCap<T>::value;

struct NodeBox {
  int lo[3], d[3];
  int vol;
  bool staged;
};

// Block-wide bounding box of the valid particles' bases.  NT: the block size
// when fixed at compile time, else 0 (blockDim 64, 128 or 256); red holds 6
// ints per wave.
template <int CAP, int NT>
__device__ __forceinline__ NodeBox chunk_box(bool valid, const int base[3], int* red) {
  int v[6];
  for (int c = 0; c < 3; ++c) {
    v[c] = valid ? base[c] : INT_MAX;
    v[3 + c] = valid ? base[c] : INT_MIN;
  }
  for (int off = 32; off > 0; off >>= 1)
    for (int c = 0; c < 3; ++c) {
      v[c] = min(v[c], __shfl_xor(v[c], off));
      v[3 + c] = max(v[3 + c], __shfl_xor(v[3 + c], off));
    }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int c = 0; c < 6; ++c) red[wave * 6 + c] = v[c];
  __syncthreads();
  NodeBox b;
  int hi[3];
  const int nw = (NT ? NT : (int)blockDim.x) >> 6;
  for (int c = 0; c < 3; ++c) {
    b.lo[c] = red[c];
    hi[c] = red[3 + c];
    for (int w = 1; w < nw; ++w) {
      b.lo[c] = min(b.lo[c], red[6 * w + c]);
      hi[c] = max(hi[c], red[6 * w + 3 + c]);
    }
  }
  const bool any = hi[0] >= b.lo[0];
  for (int c = 0; c < 3; ++c) b.d[c] = any ? hi[c] - b.lo[c] + 3 : 0;
  b.vol = b.d[0] * b.d[1] * b.d[2];
  b.staged = any && b.vol <= CAP;
  return b;
}

// n / d for 0 <= n < 2^20, given inv_d = 1 / d: (n + 1/2) / d lies at
// least 1 / (2d) from an integer, and the float product is within
// (n + 1/2) / d * 2^-23 of it.
__device__ __forceinline__ int small_div(int n, float inv_d) {
  return (int)(((float)n + 0.5f) * inv_d);
}

// lds[n] = record of box node n (x-major), zero outside the grid; NT as in
// chunk_box.
template <int NT, typename T>
__device__ __forceinline__ void stage_box(const Inst<T>& I, const NodeBox& b, const T* rec,
                                          V4<T>* lds) {
  const int plane = b.d[1] * b.d[2];  // <= b.vol <= the box capacity
  const float inv_plane = 1.0f / (float)plane, inv_d2 = 1.0f / (float)b.d[2];
  for (int n = threadIdx.x; n < b.vol; n += (NT ? NT : (int)blockDim.x)) {
    const int i = small_div(n, inv_plane), r = n - i * plane;
    const int j = small_div(r, inv_d2), k = r - j * b.d[2];
    const int gx = b.lo[0] + i, gy = b.lo[1] + j, gz = b.lo[2] + k;
    V4<T> g = {T(0), T(0), T(0), T(0)};
    if (gx >= 0 && gx < I.nx && gy >= 0 && gy < I.ny && gz >= 0 && gz < I.nz)
      g = ld4(rec + 4 * (I.g_off + ((int64_t)gx * I.ny + gy) * I.nz + gz));
    lds[n] = g;
  }
}

__device__ __forceinline__ bool base_in_grid(const int base[3], int nx, int ny, int nz) {
  return base[0] >= -2 && base[0] < nx && base[1] >= -2 && base[1] < ny && base[2] >= -2 &&
         base[2] < nz;
}

// Node records of one particle's 3x3x3 stencil: from the staged box, or from
// global memory (zero outside the grid).
template <typename T>
struct BoxNodes {
  const V4<T>* lds;
  int l0, sy, sz;
  __device__ __forceinline__ V4<T> operator()(int i, int j, int k) const {
    return lds[l0 + i * sy + j * sz + k];
  }
};
template <typename T>
struct GridNodes {
  const T* rec;
  int64_t g_off;
  int ny, nz;
  AxisNodes ax, ay, az;
  // select, not array indexing: i, j may be loop variables of a rolled loop
  template <typename E>
  static __device__ __forceinline__ E pick(const E a[3], int i) {
    return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]);
  }
  __device__ __forceinline__ V4<T> operator()(int i, int j, int k) const {
    const int64_t gi =
        g_off + ((int64_t)pick(ax.idx, i) * ny + pick(ay.idx, j)) * nz + pick(az.idx, k);
    const V4<T> g = ld4(rec + 4 * gi);
    const bool ok = pick(ax.ok, i) && pick(ay.ok, j) && pick(az.ok, k);
    const V4<T> z = {T(0), T(0), T(0), T(0)};
    return ok ? g : z;
  }
};

template <typename T>
__device__ __forceinline__ BoxNodes<T> box_nodes(const V4<T>* lds, const NodeBox& b,
                                                 const int base[3]) {
  BoxNodes<T> s;
  s.lds = lds;
  s.sz = b.d[2];
  s.sy = b.d[1] * b.d[2];
  s.l0 = (base[0] - b.lo[0]) * s.sy + (base[1] - b.lo[1]) * s.sz + (base[2] - b.lo[2]);
  return s;
}
template <typename T>
__device__ __forceinline__ GridNodes<T> grid_nodes(const T* rec, const Inst<T>& I,
                                                   const int base[3]) {
  GridNodes<T> s;
  s.rec = rec;
  s.g_off = I.g_off;
  s.ny = I.ny;
  s.nz = I.nz;
  s.ax = axis_nodes(base[0], I.nx);
  s.ay = axis_nodes(base[1], I.ny);
  s.az = axis_nodes(base[2], I.nz);
  return s;
}

// ---------------------------------------------------------------------------
// G2P (transfer.py:61-77), one thread per particle slot (coalesced state
// reads and writes); stencil nodes come from the block's box staged in LDS.
// NEXT: also writes the next step's payload, sort key, bin count and unit
// flag from the new state (what k_prep_fwd would compute from dyn_out, bit for
// bit), with the next step's growth factor sched_next.
// ---------------------------------------------------------------------------
template <typename T, int MODEL, bool NEXT>
__global__ __launch_bounds__(256) void k_g2p(const mpm_instance* inst, const int32_t* slot_inst,
                                             int64_t n_slots, const T* dyn_in, const T* grid,
                                             T* dyn_out, int advect, const T* Fgp,
                                             const T* sched_next, int TX, int TY, int TZ,
                                             T* payload_next, int32_t* key, int32_t* count,
                                             int32_t* unit_flag, int out,
                                             const int32_t* chunks) {
  constexpr int CAP = kG2pBoxCap<T>;
  __shared__ V4<T> lg[CAP];  // the block's stencil box of the grid
  __shared__ int red[24];
  const int64_t p0 = block_first_slot(chunks), p = p0 + threadIdx.x;
  if (p0 >= n_slots) return;  // whole blocks
  const int ii = __builtin_amdgcn_readfirstlane(slot_inst[p0 / kChunk]);
  const Inst<T> I = load_inst<T, MODEL>(inst, ii);
  const int64_t local = p - I.p_off;
  const bool live = local < I.n;
  T x[3] = {T(0), T(0), T(0)}, fx[3] = {T(0), T(0), T(0)};
  int base[3] = {0, 0, 0};
  bool valid = false;
  if (live) {
    ld3_nt(dyn_in, n_slots, p, RX, x);
    stencil_base(x, I, base, fx);
    valid = base_in_grid(base, I.nx, I.ny, I.nz);
  }
  // Particles of consecutive slots are spatially close, so the box spanned by
  // the block's stencils is small; it is staged once in LDS and the 27-node
  // gathers read it there (a box beyond the LDS capacity reads the grid).
  const NodeBox bx = chunk_box<CAP, 0>(valid, base, red);
  if (bx.staged) stage_box<0>(I, bx, grid, lg);
  __syncthreads();
  if (!NEXT && !live) return;
  // (with NEXT, every lane stays: count_in_bin needs every lane)
  int32_t next_key = -1;
  if (live) {
    // Separable weights: with d = (o - fx) dx,
    //   nv = sum w v,   nC = k4 dx (sum w v o^T - nv fx^T),
    // and each z-row (i, j) of 3 nodes is summed first.
    T nv[3] = {T(0), T(0), T(0)};
    M3<T> nC = zero3<T>();
    auto gather = [&](const auto& nodes) {
      T w[3][3];
      for (int c = 0; c < 3; ++c) bspline_w(fx[c], w[c]);
      const T wz0 = w[2][0], wz1 = w[2][1], wz2 = w[2][2], wz2k = T(2) * w[2][2];
      T Si[3] = {T(0), T(0), T(0)}, Sj[3] = {T(0), T(0), T(0)}, Sk[3] = {T(0), T(0), T(0)};
      // x planes stay rolled to bound the gathered values in flight
#pragma unroll 1
      for (int i = 0; i < 3; ++i) {
        const T wxi = i == 0 ? w[0][0] : (i == 1 ? w[0][1] : w[0][2]);
        const T Ti = T(i);
        // z-rows stay rolled too (77-79 VGPRs in FP32, occupancy 6); the
        // row's y weight is picked by select
#pragma unroll 1
        for (int jj = 0; jj < 3; ++jj) {
          const V4<T> g0 = nodes(i, jj, 0), g1 = nodes(i, jj, 1), g2 = nodes(i, jj, 2);
          const T v0[3] = {g0.y, g0.z, g0.w}, v1[3] = {g1.y, g1.z, g1.w},
                  v2[3] = {g2.y, g2.z, g2.w};
          const T wyj = jj == 0 ? w[1][0] : (jj == 1 ? w[1][1] : w[1][2]);
          const T wxy = wxi * wyj;
          for (int c = 0; c < 3; ++c) {
            const T r0 = wz0 * v0[c] + wz1 * v1[c] + wz2 * v2[c];
            const T r1 = wz1 * v1[c] + wz2k * v2[c];
            nv[c] += wxy * r0;
            Sk[c] += wxy * r1;
            Si[c] += (Ti * wxy) * r0;
            if (jj) Sj[c] += (T(jj) * wxy) * r0;
          }
        }
      }
      const T s = I.k4 * I.dx;
      for (int c = 0; c < 3; ++c) {
        nC(c, 0) = s * (Si[c] - nv[c] * fx[0]);
        nC(c, 1) = s * (Sj[c] - nv[c] * fx[1]);
        nC(c, 2) = s * (Sk[c] - nv[c] * fx[2]);
      }
    };
    if (valid) {
      if (bx.staged)
        gather(box_nodes(lg, bx, base));
      else
        gather(grid_nodes(grid, I, base));
    }
    const M3<T> F = ld9_nt(dyn_in, n_slots, p, RF);
    // F_new = (I + dt C) F, formed as F + dt (C F): I + dt C would round away
    // the low bits of the small increment the same way for every particle.
    const M3<T> CF = mul(nC, F);
    M3<T> Fn;
    for (int k = 0; k < 9; ++k) Fn.a[k] = F.a[k] + I.dt * CF.a[k];
    T xn[3];
    for (int c = 0; c < 3; ++c) xn[c] = advect ? x[c] + I.dt * nv[c] : x[c];
    st3_nt(dyn_out, n_slots, p, RX, xn);
    if (out == MPM_OUT_FULL) st3_nt(dyn_out, n_slots, p, RV, nv);
    if (out != MPM_OUT_NEXT) st9_nt(dyn_out, n_slots, p, RC, nC);
    st9_nt(dyn_out, n_slots, p, RF, Fn);
    if (NEXT) {
      const M3<T> Fg = grown(ld9_nt(Fgp, n_slots, p, 0), sched_next ? sched_next[ii] : T(1));
      int nbase[3];
      make_payload(I, xn, nv, nC, Fn, Fg, payload_next + p * kRecord, nbase);
      next_key = cell_key(inst[ii], nbase, TX, TY, TZ, unit_flag);
    }
  } else if (NEXT) {
    count_padding(inst[ii], local, TX * TY * TZ, count);
  }
  if (NEXT) put_key(p, next_key, key, count);
}

// ---------------------------------------------------------------------------
// reverse step, part 1 (transfer.py:100-114): G2P-adjoint payload.
//   gv_bar_i += w nv_b + 4 inv_dx^2 w C_b' dpos,  nv_b = vn_b + dt xn_b,
//   C_b' = Cn_b + dt Fn_b F^T;  stored like the forward payload with
//   P = nv_b, A = 4 inv_dx^2 C_b'.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void make_adj_payload(const Inst<T>& I, const T x[3], const M3<T>& F,
                                                 const T xb[3], const T vb[3], const M3<T>& Cnb,
                                                 const M3<T>& Fnb, int advect, T* rec) {
  T fx[3], nvb[3];
  int base[3];
  stencil_base(x, I, base, fx);
  for (int a = 0; a < 3; ++a) nvb[a] = advect ? vb[a] + I.dt * xb[a] : vb[a];
  const M3<T> FF = mul_nt(Fnb, F);
  M3<T> Ap;
  for (int k = 0; k < 9; ++k) Ap.a[k] = I.k4 * (Cnb.a[k] + I.dt * FF.a[k]) * I.dx;
  T q[3];
  for (int a = 0; a < 3; ++a)
    q[a] = nvb[a] - (Ap(a, 0) * fx[0] + Ap(a, 1) * fx[1] + Ap(a, 2) * fx[2]);
  store_payload(rec, fx, q, Ap, pack_cell(base, I.nx, I.ny, I.nz));
}

template <typename T>
__global__ __launch_bounds__(256) void k_prep_adj(const mpm_instance* inst, const int32_t* slot_inst, int64_t n_slots,
                           const T* dyn_in, const T* bar_out, const int32_t* rank, T* payload,
                           int advect) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_slots) return;
  const Inst<T> I = load_inst<T>(inst, block_inst(slot_inst));
  if (p - I.p_off >= I.n) return;
  T x[3], xb[3], vb[3];
  ld3(dyn_in, n_slots, p, RX, x);
  ld3(bar_out, n_slots, p, RX, xb);
  ld3(bar_out, n_slots, p, RV, vb);
  make_adj_payload(I, x, ld9(dyn_in, n_slots, p, RF), xb, vb, ld9(bar_out, n_slots, p, RC),
                   ld9(bar_out, n_slots, p, RF), advect, payload + (int64_t)rank[p] * kRecord);
}

// reverse step, part 2 (transfer.py:121-127): grid adjoint.
//   gadj record per node: (gm_bar, gmom_bar x, y, z).
template <typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__(256) void k_grid_adj(const mpm_instance* inst, const int32_t* node_inst,
                                                  int64_t n_nodes, const int32_t* unit_rank,
                                                  const T* partial, const T* grid, T* gadj,
                                                  int advect) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_nodes) return;
  const Inst<T> I = load_inst<T>(inst, block_inst(node_inst));
  int i, j, k;
  if (!node_of(I, g, i, j, k)) return;
  T acc[4];
  gather_tiles<T, TX, TY, TZ>(I, unit_rank, partial, i, j, k, acc);
  const V4<T> gr = ld4(grid + 4 * g);
  const T m = gr.x;
  T mb[3] = {T(0), T(0), T(0)};
  T gmb = T(0);
  if (m > T(kMassEps)) {
    for (int a = 0; a < 3; ++a) mb[a] = I.s * acc[1 + a] / m;
    if (advect) gmb = -(acc[1] * gr.y + acc[2] * gr.z + acc[3] * gr.w) / m;
  }
  st4(gadj + 4 * g, gmb, mb[0], mb[1], mb[2]);
}

// reverse step, part 3 (transfer.py:100-119 x_bar terms, :129-156), one
// thread per particle slot p (coalesced state rows, read and written once,
// so non-temporal: they leave the caches to the staged boxes and the
// gathered payloads), at sorted position j = rank[p]; each 256-slot chunk
// stages its stencil box of the grid adjoint (and, with advection, of the
// grid) in LDS.
//
// The P2G-adjoint gather reads the forward payload pf (fx, q2, A', cell) of p
// and the adjoint payload pa (fx, q1, P1) at j.  With dpos = (o - fx) dx for
// stencil offset o, the per-node factors are affine in o:
//   nv_b + 4 inv_dx^2 C_b' dpos = q1 + P1 o   (P1 = 4 inv_dx^2 dx C_b')
//   pmass v + A dpos            = q2 + A' o   (A' = A dx, A = pmass C + dt c tau)
// so x_bar's grad-w coefficient is v_i.(q1 + P1 o) + mb_i.(q2 + A' o) + gm_b,i pmass.
// It gives x_bar, v_bar and MD = sum w mb dpos^T.
//
// The constitutive VJP then follows in the same thread, with MD in registers:
// tau_b = c sk_b with sk_b = dt MD, c = -4 inv_dx^2 vol0 det Fg;
//   C_bar = pmass MD + C_bar(stress),  F_bar = (I + dt Cn)^T Fn_b + F_bar(stress),
//   G_bar += a (Fg_bar(stress) + <sk_b, tau> (-4 inv_dx^2 vol0) det Fg Fg^-T).
// PREV: with the state's full cotangent at hand, also writes the previous
// step's G2P-adjoint payload, what k_prep_adj computes for that step, at its
// sorted position rank_prev[p]; C_bar then goes only there.
template <typename T, int MODEL, bool PREV>
__global__ __launch_bounds__(256) void k_adj_particle(
    const mpm_instance* inst, const int32_t* slot_inst, int64_t n_slots, const int32_t* rank,
    const T* pf, const T* pa, const T* grid, const T* gadj, const T* dyn_in, const T* Fgp,
    const T* sched, const T* dyn_out, const T* bar_out, T* bar_in, T* Fg_bar, double* theta_acc,
    const T* dyn_prev, const int32_t* rank_prev, T* payload_prev, int advect,
    const int32_t* chunks) {
  constexpr int CAP = kAdjBoxCap<T, MODEL>;
  __shared__ V4<T> lg[CAP];  // grid adjoint (gm_b, mb)
  __shared__ V4<T> lv[CAP];  // grid (m, v), advection only
  __shared__ int red[24];
  const int64_t p0 = block_first_slot(chunks), p = p0 + threadIdx.x;  // blockDim = 256
  const int ii = __builtin_amdgcn_readfirstlane(slot_inst[p0 / kChunk]);
  const Inst<T> I = load_inst<T, MODEL>(inst, ii);
  const bool live = p - I.p_off < I.n;
  const int64_t j = live ? rank[p] : 0;
  T fx[3] = {T(0), T(0), T(0)}, q1[3], q2[3];
  M3<T> P1, Ap;
  uint32_t cell = kNoCell;
  if (live) {
    const T* rf = pf + p * kRecord;  // forward payload: slot order
    const T* ra = pa + j * kRecord;  // adjoint payload: sorted order
    const V4<T> f0 = ld4(rf), f1 = ld4(rf + 4), f2 = ld4(rf + 8), f3 = ld4(rf + 12);
    fx[0] = f0.x; fx[1] = f0.y; fx[2] = f0.z;
    q2[0] = f0.w; q2[1] = f1.x; q2[2] = f1.y;
    Ap.a[0] = f1.z; Ap.a[1] = f1.w; Ap.a[2] = f2.x; Ap.a[3] = f2.y; Ap.a[4] = f2.z;
    Ap.a[5] = f2.w; Ap.a[6] = f3.x; Ap.a[7] = f3.y; Ap.a[8] = f3.z;
    cell = cell_of(f3.w);
    const V4<T> a0 = ld4(ra), a1 = ld4(ra + 4), a2 = ld4(ra + 8), a3 = ld4(ra + 12);
    q1[0] = a0.w; q1[1] = a1.x; q1[2] = a1.y;
    P1.a[0] = a1.z; P1.a[1] = a1.w; P1.a[2] = a2.x; P1.a[3] = a2.y; P1.a[4] = a2.z;
    P1.a[5] = a2.w; P1.a[6] = a3.x; P1.a[7] = a3.y; P1.a[8] = a3.z;
  }
  const bool valid = live && cell != kNoCell;
  const int base[3] = {(int)(cell & 1023u) - 2, (int)((cell >> 10) & 1023u) - 2,
                       (int)(cell >> 20) - 2};
  const NodeBox bx = chunk_box<CAP, 256>(valid, base, red);
  if (bx.staged) {
    stage_box<256>(I, bx, gadj, lg);
    if (advect) stage_box<256>(I, bx, grid, lv);
  }
  __syncthreads();
  T th[2] = {T(0), T(0)};
  if (live) {
    T xb[3] = {T(0), T(0), T(0)};
    T vb[3] = {T(0), T(0), T(0)};
    M3<T> MD = zero3<T>();
    {
      T Sv[3] = {T(0), T(0), T(0)};  // sum w v_i
      // Separable weights: with d = (o - fx) dx, vb = sum w mb and
      //   MD = dx (sum w mb o^T - vb fx^T);
      // x_bar's terms sum alpha grad w, grad w = (dwx wy wz, wx dwy wz, wx wy dwz) / dx,
      // over each z-row (i, j) of 3 nodes first.
      M3<T> Mo = zero3<T>();  // sum w mb o^T
      auto gather = [&](const auto& adj_nodes, const auto& vel_nodes) {
        T w[3][3], dw[3][3];
        for (int c = 0; c < 3; ++c) {
          bspline_w(fx[c], w[c]);
          bspline_dw(fx[c], dw[c]);
        }
        const T wz0 = w[2][0], wz1 = w[2][1], wz2 = w[2][2], wz2k = T(2) * w[2][2];
        // The (i, j) loop stays rolled (9 rows of 3 nodes) to bound the live
        // values; per-axis values of row (i, j) are picked by select.
#pragma unroll 1
        for (int ij = 0; ij < 9; ++ij) {
          const int i = ij / 3, jj = ij - 3 * (ij / 3);
          const T wxi = i == 0 ? w[0][0] : (i == 1 ? w[0][1] : w[0][2]);
          const T dxi = i == 0 ? dw[0][0] : (i == 1 ? dw[0][1] : dw[0][2]);
          const T wyj = jj == 0 ? w[1][0] : (jj == 1 ? w[1][1] : w[1][2]);
          const T dyj = jj == 0 ? dw[1][0] : (jj == 1 ? dw[1][1] : dw[1][2]);
          const T Ti = T(i), Tj = T(jj);
          const T wxy = wxi * wyj;
          V4<T> ga[3];
          for (int k = 0; k < 3; ++k) ga[k] = adj_nodes(i, jj, k);
          for (int c = 0; c < 3; ++c) {
            const T m0 = c == 0 ? ga[0].y : (c == 1 ? ga[0].z : ga[0].w);
            const T m1 = c == 0 ? ga[1].y : (c == 1 ? ga[1].z : ga[1].w);
            const T m2 = c == 0 ? ga[2].y : (c == 1 ? ga[2].z : ga[2].w);
            const T r0 = wz0 * m0 + wz1 * m1 + wz2 * m2;
            const T r1 = wz1 * m1 + wz2k * m2;
            vb[c] += wxy * r0;
            Mo(c, 0) += (Ti * wxy) * r0;
            Mo(c, 1) += (Tj * wxy) * r0;
            Mo(c, 2) += wxy * r1;
          }
          if (advect) {
            T u1[3], u2[3];
            for (int c = 0; c < 3; ++c) {
              u1[c] = q1[c] + Ti * P1(c, 0) + Tj * P1(c, 1);
              u2[c] = q2[c] + Ti * Ap(c, 0) + Tj * Ap(c, 1);
            }
            T A0 = T(0), A1 = T(0);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
              const V4<T> gv = vel_nodes(i, jj, k);
              const T vi[3] = {gv.y, gv.z, gv.w};
              const T mb[3] = {ga[k].y, ga[k].z, ga[k].w};
              T alpha = ga[k].x * I.pmass;
              for (int c = 0; c < 3; ++c)
                alpha += vi[c] * (u1[c] + T(k) * P1(c, 2)) + mb[c] * (u2[c] + T(k) * Ap(c, 2));
              A0 += w[2][k] * alpha;
              A1 += dw[2][k] * alpha;
              const T wt = wxy * w[2][k];
              Sv[0] += wt * gv.y;
              Sv[1] += wt * gv.z;
              Sv[2] += wt * gv.w;
            }
            xb[0] += (dxi * wyj * I.inv_dx) * A0;
            xb[1] += (wxi * dyj * I.inv_dx) * A0;
            xb[2] += (wxy * I.inv_dx) * A1;
          }
        }
      };
      if (valid) {
        if (bx.staged)
          gather(box_nodes(lg, bx, base), box_nodes(lv, bx, base));
        else
          gather(grid_nodes(gadj, I, base), grid_nodes(grid, I, base));
      }
      if (advect) {
        // x_bar += xn_b - (4 inv_dx^2 C_b')^T (sum w v_i) - A^T (sum w mb_i), A = A'/dx
        T xnb[3];
        ld3_nt(bar_out, n_slots, p, RX, xnb);
        for (int c = 0; c < 3; ++c)
          xb[c] += xnb[c] - ((P1(0, c) * Sv[0] + P1(1, c) * Sv[1] + P1(2, c) * Sv[2]) +
                             (Ap(0, c) * vb[0] + Ap(1, c) * vb[1] + Ap(2, c) * vb[2])) * I.inv_dx;
      }
      for (int c = 0; c < 3; ++c)
        for (int e = 0; e < 3; ++e) MD(c, e) = I.dx * (Mo(c, e) - vb[c] * fx[e]);
      for (int c = 0; c < 3; ++c) vb[c] *= I.pmass;
    }
    st3_nt(bar_in, n_slots, p, RX, xb);
    st3_nt(bar_in, n_slots, p, RV, vb);

    const T a_t = sched ? sched[ii] : T(1);
    const M3<T> Fg = grown(ld9_nt(Fgp, n_slots, p, 0), a_t);
    const T detG = det(Fg);
    const T c = -I.k4 * I.vol0 * detG;
    const M3<T> skb = scaled(MD, I.dt);
    const M3<T> C = ld9_nt(dyn_in, n_slots, p, RC);
    M3<T> Fpb, Fgb, Cbar, tau;
    {
      const M3<T> F = ld9_nt(dyn_in, n_slots, p, RF);
      kirchhoff_vjp(I.mp, F, Fg, C, scaled(skb, c), Fpb, Fgb, Cbar, th, &tau);
    }
    M3<T> Cbo = scaled(MD, I.pmass);
    for (int k = 0; k < 9; ++k) Cbo.a[k] += Cbar.a[k];
    // with PREV, C_bar goes into the previous step's payload below and no
    // later kernel of the chain reads it from bar_in
    if (!PREV) st9_nt(bar_in, n_slots, p, RC, Cbo);
    {
      const M3<T> Cn = ld9_nt(dyn_out, n_slots, p, RC);
      const M3<T> Fnb = ld9_nt(bar_out, n_slots, p, RF);
      // (I + dt C_new)^T Fn_b, formed as Fn_b + dt C_new^T Fn_b
      const M3<T> CtF = mul_tn(Cn, Fnb);
      for (int k = 0; k < 9; ++k) Fpb.a[k] += Fnb.a[k] + I.dt * CtF.a[k];
      st9_nt(bar_in, n_slots, p, RF, Fpb);
    }
    if (PREV) {
      T xp[3];
      ld3_nt(dyn_prev, n_slots, p, RX, xp);
      make_adj_payload(I, xp, ld9_nt(dyn_prev, n_slots, p, RF), xb, vb, Cbo, Fpb, advect,
                       payload_prev + (int64_t)rank_prev[p] * kRecord);
    }
    const M3<T> GiT = inv_t(Fg);
    const T vol = frob(skb, tau) * (-I.k4 * I.vol0) * detG;
    M3<T> Gb = ld9_nt(Fg_bar, n_slots, p, 0);
    for (int k = 0; k < 9; ++k) Gb.a[k] += a_t * (Fgb.a[k] + vol * GiT.a[k]);
    st9_nt(Fg_bar, n_slots, p, 0, Gb);
  }
  if (theta_acc != nullptr) {  // fixed-order block reduction, one slot per 256 particles
    __shared__ double tred[2][256];
    tred[0][threadIdx.x] = (double)th[0];
    tred[1][threadIdx.x] = (double)th[1];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
      if (threadIdx.x < (unsigned)off) {
        tred[0][threadIdx.x] += tred[0][threadIdx.x + off];
        tred[1][threadIdx.x] += tred[1][threadIdx.x + off];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      theta_acc[2 * (p0 / kChunk)] += tred[0][0];
      theta_acc[2 * (p0 / kChunk) + 1] += tred[1][0];
    }
  }
}

// ---------------------------------------------------------------------------
// constitutive seam kernels
// ---------------------------------------------------------------------------
template <typename T, int MODEL>
__global__ __launch_bounds__(256) void k_stress(ModelParams<T> mp, int64_t n, int64_t stride, const T* F, const T* Fg,
                         const T* C, T* tau) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const M3<T> Cm = C ? ld9(C, stride, p, 0) : zero3<T>();
  mp.kind = MODEL;
  st9(tau, stride, p, 0, kirchhoff(mp, ld9(F, stride, p, 0), ld9(Fg, stride, p, 0), Cm));
}

template <typename T, int MODEL>
__global__ __launch_bounds__(256) void k_stress_vjp(ModelParams<T> mp, int64_t n, int64_t stride, const T* F,
                             const T* Fg, const T* C, const T* tb, T* Fb, T* Fgb, T* Cb, T* thb) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const M3<T> Cm = C ? ld9(C, stride, p, 0) : zero3<T>();
  M3<T> fb, gb, cb;
  T th[2];
  mp.kind = MODEL;
  kirchhoff_vjp(mp, ld9(F, stride, p, 0), ld9(Fg, stride, p, 0), Cm, ld9(tb, stride, p, 0), fb, gb,
                cb, th);
  st9(Fb, stride, p, 0, fb);
  st9(Fgb, stride, p, 0, gb);
  if (Cb) st9(Cb, stride, p, 0, cb);
  if (thb) {
    thb[p] = th[0];
    thb[stride + p] = th[1];
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct TileShape {
  int x, y, z;
};
constexpr TileShape kTiles[] = {{4, 8, 8}, {4, 4, 16}, {8, 8, 4}, {4, 4, 8}, {4, 4, 4}, {2, 4, 8}};
constexpr int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

int tile_index(const mpm_batch* b) {
  for (int i = 0; i < kNumTiles; ++i)
    if (kTiles[i].x == b->tile[0] && kTiles[i].y == b->tile[1] && kTiles[i].z == b->tile[2]) return i;
  return -1;
}

int64_t tile_nodes(const mpm_batch* b) {
  return (int64_t)(b->tile[0] + 2) * (b->tile[1] + 2) * (b->tile[2] + 2);
}

inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// payload and payload2: the two areas of the G2P-adjoint payload in a chain
// of reverse steps, each step reading one and writing the other.
struct Workspace {
  int32_t *key, *cursor, *placed, *count, *bsum, *unit_flag, *ucount;
  void *payload, *payload2, *partial, *gadj;
};

// Views into a step's sort record:
// rank[n_slots] | sorted[n_slots] | start[n_bins + 1] | unit_rank[n_tiles] |
// active_unit[n_tiles] | n_active[1].
template <typename I32>
struct SortView {
  I32 *rank, *sorted, *start, *unit_rank, *active_unit, *n_active;
  SortView(const mpm_batch* b, I32* s) {
    rank = s;
    sorted = s + b->n_slots;
    start = s + 2 * b->n_slots;
    unit_rank = start + b->n_bins + 1;
    active_unit = unit_rank + b->n_tiles;
    n_active = active_unit + b->n_tiles;
  }
};

int64_t scan_blocks(int64_t n_bins) {
  return (n_bins + kScanBlock * kScanPer - 1) / (kScanBlock * kScanPer);
}

// The regions a forward step uses (sort scratch and tile partials) come
// first; *fwd_bytes (optional) receives their extent.  The reverse step's
// regions (two adjoint payload areas, the adjoint grid) follow.
int64_t carve(const mpm_batch* b, size_t tsz, char* base, Workspace* w,
              int64_t* fwd_bytes = nullptr) {
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* p = base ? base + off : nullptr;
    off = align_up(off + bytes, 256);
    return p;
  };
  Workspace ws;
  ws.key = (int32_t*)take(4 * b->n_slots);
  ws.cursor = (int32_t*)take(4 * b->n_bins);
  ws.count = (int32_t*)take(4 * b->n_bins);
  ws.bsum = (int32_t*)take(4 * (scan_blocks(b->n_bins) + 1));
  ws.unit_flag = (int32_t*)take(4 * b->n_tiles);
  ws.ucount = (int32_t*)take(4 * ((b->n_tiles + 255) / 256));
  ws.placed = (int32_t*)take(4 * b->n_slots);
  ws.partial = take((int64_t)tsz * 4 * tile_nodes(b) * b->n_tiles);
  if (fwd_bytes) *fwd_bytes = off;
  ws.payload = take((int64_t)tsz * kRecord * b->n_slots);
  ws.gadj = take((int64_t)tsz * 4 * b->n_nodes);
  ws.payload2 = take((int64_t)tsz * kRecord * b->n_slots);
  if (w) *w = ws;
  return off;
}

int validate(const mpm_batch* b) {
  if (!b) return set_err(MPM_EINVAL, "batch is NULL");
  if (b->n_inst <= 0 || !b->inst || !b->slot_inst || !b->node_inst || !b->tile_inst)
    return set_err(MPM_EINVAL, "batch tables missing");
  if (b->n_slots <= 0 || b->n_slots % kChunk || b->n_nodes <= 0 || b->n_nodes % kChunk)
    return set_err(MPM_EINVAL, "n_slots / n_nodes must be positive multiples of %d", kChunk);
  if (b->n_slots >= (int64_t)1 << 31 || b->n_bins >= (int64_t)1 << 31)
    return set_err(MPM_EINVAL, "n_slots and n_bins must be below 2^31");
  if (b->n_tiles <= 0 || b->n_tiles < b->n_inst)
    return set_err(MPM_EINVAL, "unit table must hold every instance's tiles and dead unit");
  if (tile_index(b) >= 0 &&
      b->n_bins != b->n_tiles * (int64_t)b->tile[0] * b->tile[1] * b->tile[2])
    return set_err(MPM_EINVAL, "n_bins must be n_tiles * tile cells");
  if (b->block != 64 && b->block != 128 && b->block != 256)
    return set_err(MPM_EINVAL, "block must be 64, 128 or 256 (got %d)", b->block);
  if (tile_index(b) < 0)
    return set_err(MPM_ENOTSUP, "tile shape %dx%dx%d not built", b->tile[0], b->tile[1], b->tile[2]);
  if (b->model_chunk != nullptr) {
    const int64_t* o = b->model_chunk_off;
    if (o[0] != 0 || o[1] < o[0] || o[2] < o[1] || o[3] < o[2] || o[3] != b->n_slots / kChunk)
      return set_err(MPM_EINVAL, "model_chunk_off must rise from 0 to n_slots / %d", kChunk);
  }
  return MPM_OK;
}

inline dim3 grid1(int64_t n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }

// Calls f with std::integral_constant<int, M>: the single model of the batch,
// or -1 (read per instance) when the batch mixes models.
template <typename F>
void with_model(int32_t mask, F&& f) {
  switch (mask) {
    case 1 << MPM_NEO_HOOKEAN: f(std::integral_constant<int, MPM_NEO_HOOKEAN>()); break;
    case 1 << MPM_HENCKY: f(std::integral_constant<int, MPM_HENCKY>()); break;
    case 1 << MPM_HERSCHEL_BULKLEY: f(std::integral_constant<int, MPM_HERSCHEL_BULKLEY>()); break;
    default: f(std::integral_constant<int, -1>()); break;
  }
}

// Launches f(std::integral_constant<int, M>, chunks, grid) for the batch's
// particle slots: one launch per model over that model's chunks when the
// batch lists them (each launch runs kernels specialised for its model), else
// one launch over all slots (chunks = NULL).
template <typename F>
void per_model(const mpm_batch* b, int bs, F&& f) {
  if (b->model_chunk == nullptr || __builtin_popcount(b->model_mask) < 2) {
    with_model(b->model_mask, [&](auto M) { f(M, (const int32_t*)nullptr, grid1(b->n_slots, bs)); });
    return;
  }
  for (int m = 0; m < 3; ++m) {
    const int64_t n = b->model_chunk_off[m + 1] - b->model_chunk_off[m];
    if (n <= 0) continue;
    const int32_t* ch = b->model_chunk + b->model_chunk_off[m];
    const dim3 g((unsigned)(n * (kChunk / bs)));
    if (m == MPM_NEO_HOOKEAN) f(std::integral_constant<int, MPM_NEO_HOOKEAN>(), ch, g);
    if (m == MPM_HENCKY) f(std::integral_constant<int, MPM_HENCKY>(), ch, g);
    if (m == MPM_HERSCHEL_BULKLEY) f(std::integral_constant<int, MPM_HERSCHEL_BULKLEY>(), ch, g);
  }
}

// Kernel timer (mpm_kernel_timer): events around the armed kernel's launches.
struct KernelTimer {
  int id = -1;
  hipEvent_t* ev = nullptr;
  int n_pairs = 0;
  int count = 0;
};
// Per host thread: arming the timer times the launches the same thread
// issues, and other threads' calls neither see nor disturb it.
thread_local KernelTimer g_timer;

// Launches under stream capture are not timed: this stack refuses event
// records (external or not) inside a capture.
bool capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

struct TimedLaunch {
  bool on;
  hipStream_t st;
  hipEvent_t stop;
  TimedLaunch(int id, hipStream_t s) : on(false), st(s), stop(nullptr) {
    if (g_timer.id == id && g_timer.count < g_timer.n_pairs && !capturing(st)) {
      on = true;
      (void)hipEventRecord(g_timer.ev[2 * g_timer.count], st);
      stop = g_timer.ev[2 * g_timer.count + 1];
      ++g_timer.count;
    }
  }
  ~TimedLaunch() {
    if (on) (void)hipEventRecord(stop, st);
  }
};

void clear_bins(const mpm_batch* b, const Workspace& w, hipStream_t st) {
  const int64_t n = b->n_bins > b->n_tiles ? b->n_bins : b->n_tiles;
  k_clear_bins<<<dim3((unsigned)((n + 255) / 256)), 256, 0, st>>>(w.count, b->n_bins,
                                                                  w.unit_flag, b->n_tiles);
}

// The sort from the unit-level keys, bin counts and unit flags a prep, G2P or
// key kernel left in the workspace.
int sort_from_keys(const mpm_batch* b, int32_t* sortbuf, const Workspace& w, hipStream_t st) {
  const SortView<int32_t> v(b, sortbuf);
  const int bs = b->block;
  const int TB = b->tile[0] * b->tile[1] * b->tile[2];
  const int64_t nb = scan_blocks(b->n_bins);
  const int64_t nu = (b->n_tiles + 255) / 256;
  TimedLaunch tl(MPM_K_SORT, st);
  k_unit_count<<<dim3((unsigned)nu), 256, 0, st>>>(b->tile_inst, b->n_tiles, w.unit_flag,
                                                   w.ucount);
  k_unit_rank<<<dim3((unsigned)nu), 256, 0, st>>>(b->tile_inst, b->n_tiles, w.unit_flag, w.ucount,
                                                  v.unit_rank, v.active_unit, v.n_active);
  k_scan_blocks<<<dim3((unsigned)nb), kScanBlock, 0, st>>>(w.count, v.active_unit, v.start,
                                                            w.cursor, w.bsum, v.n_active, TB);
  k_scan_add<<<dim3((unsigned)nb), kScanBlock, 0, st>>>(v.start, w.bsum, v.n_active, TB);
  k_place<<<grid1(b->n_slots, 256 * kPlaceGroups), 256, 0, st>>>(
      v.unit_rank, TB, b->n_tiles, w.key, v.start, w.cursor, w.placed, b->n_slots);
  k_bin_rank<<<grid1(b->n_slots, 256 * kRankGroups), 256, 0, st>>>(w.key, v.start, w.placed,
                                                                   v.sorted, v.rank, b->n_slots);
  return check_launch("sort_by_cell");
}

template <typename T>
int sort_impl(const mpm_batch* b, const T* dyn, int32_t* sortbuf, const Workspace& w,
              hipStream_t st) {
  const int bs = b->block;
  {
  TimedLaunch tl(MPM_K_SORT, st);
  clear_bins(b, w, st);
  k_cell_keys<T><<<grid1(b->n_slots, bs), bs, 0, st>>>(b->inst, b->slot_inst, b->n_slots, dyn,
                                                        b->tile[0], b->tile[1], b->tile[2], w.key,
                                                        w.count, w.unit_flag);
  }
  return sort_from_keys(b, sortbuf, w, st);
}

template <typename T, int TX, int TY, int TZ>
void launch_tile_fwd(const mpm_batch* b, const SortView<const int32_t>& v, const Workspace& w,
                     const T* payload, T* grid, hipStream_t st) {
  {
  TimedLaunch tl(MPM_K_P2G_FWD, st);
  k_p2g_tile<T, TX, TY, TZ, true, true><<<dim3((unsigned)b->n_tiles), TX * TY * TZ, 0, st>>>(
      b->inst, b->tile_inst, v.active_unit, v.n_active, v.start, v.sorted, payload,
      (T*)w.partial);
  }
  {
  TimedLaunch tl(MPM_K_GRID_FWD, st);
  k_grid_fwd<T, TX, TY, TZ><<<grid1(b->n_nodes, 256), 256, 0, st>>>(
      b->inst, b->node_inst, b->n_nodes, v.unit_rank, (const T*)w.partial, grid);
  }
}

template <typename T, int TX, int TY, int TZ>
void launch_tile_adj(const mpm_batch* b, const SortView<const int32_t>& v, const Workspace& w,
                     const T* pa, const T* grid, int advect, hipStream_t st) {
  {
  TimedLaunch tl(MPM_K_P2G_ADJ_TILE, st);
  k_p2g_tile<T, TX, TY, TZ, false, false><<<dim3((unsigned)b->n_tiles), TX * TY * TZ, 0, st>>>(
      b->inst, b->tile_inst, v.active_unit, v.n_active, v.start, v.sorted, pa, (T*)w.partial);
  }
  {
  TimedLaunch tl(MPM_K_GRID_ADJ, st);
  k_grid_adj<T, TX, TY, TZ><<<grid1(b->n_nodes, 256), 256, 0, st>>>(
      b->inst, b->node_inst, b->n_nodes, v.unit_rank, (const T*)w.partial, grid, (T*)w.gadj,
      advect);
  }
}

template <typename T>
void dispatch_fwd(int ti, const mpm_batch* b, const SortView<const int32_t>& sv,
                  const Workspace& w, const T* payload, T* grid, hipStream_t st) {
  switch (ti) {
    case 0: launch_tile_fwd<T, 4, 8, 8>(b, sv, w, payload, grid, st); break;
    case 1: launch_tile_fwd<T, 4, 4, 16>(b, sv, w, payload, grid, st); break;
    case 2: launch_tile_fwd<T, 8, 8, 4>(b, sv, w, payload, grid, st); break;
    case 4: launch_tile_fwd<T, 4, 4, 4>(b, sv, w, payload, grid, st); break;
    case 5: launch_tile_fwd<T, 2, 4, 8>(b, sv, w, payload, grid, st); break;
    default: launch_tile_fwd<T, 4, 4, 8>(b, sv, w, payload, grid, st); break;
  }
}

template <typename T>
void dispatch_adj(int ti, const mpm_batch* b, const SortView<const int32_t>& sv,
                  const Workspace& w, const T* pa, const T* grid, int advect, hipStream_t st) {
  switch (ti) {
    case 0: launch_tile_adj<T, 4, 8, 8>(b, sv, w, pa, grid, advect, st); break;
    case 1: launch_tile_adj<T, 4, 4, 16>(b, sv, w, pa, grid, advect, st); break;
    case 2: launch_tile_adj<T, 8, 8, 4>(b, sv, w, pa, grid, advect, st); break;
    case 4: launch_tile_adj<T, 4, 4, 4>(b, sv, w, pa, grid, advect, st); break;
    case 5: launch_tile_adj<T, 2, 4, 8>(b, sv, w, pa, grid, advect, st); break;
    default: launch_tile_adj<T, 4, 4, 8>(b, sv, w, pa, grid, advect, st); break;
  }
}

// One forward step.  ready: this step's payload, sort keys and unit flags
// were written by the previous chained step (G2P with NEXT) into `payload`
// and the workspace; else k_prep_fwd computes them from dyn_in.
// payload_next (optional): G2P also writes the next step's payload there,
// and its keys and flags into the workspace, with growth factor sched_next.
template <typename T>
int step_fwd_impl(const mpm_batch* b, const T* dyn_in, const T* Fg, const T* sched,
                  const T* sched_next, T* dyn_out, T* grid, int32_t* sortbuf, T* payload,
                  T* payload_next, int advect, int ready, int out, void* ws, void* stream) {
  int rc = validate(b);
  if (rc) return rc;
  if (out != MPM_OUT_FULL && out != MPM_OUT_RECORD && out != MPM_OUT_NEXT &&
      out != MPM_OUT_NONE)
    return set_err(MPM_EINVAL,
                   "step_fwd: out must be MPM_OUT_FULL, _RECORD, _NEXT or _NONE (got %d)", out);
  if (out == MPM_OUT_NONE && payload_next)
    return set_err(MPM_EINVAL, "step_fwd: MPM_OUT_NONE ends a chain (payload_next must be NULL)");
  if (!dyn_in || !Fg || (!dyn_out && out != MPM_OUT_NONE) || !grid || !sortbuf || !payload ||
      !ws)
    return set_err(MPM_EINVAL, "step_fwd: NULL buffer");
  if (dyn_in == dyn_out) return set_err(MPM_EINVAL, "step_fwd: dyn_out must not alias dyn_in");
  if (payload_next == payload) return set_err(MPM_EINVAL, "step_fwd: payload_next aliases payload");
  hipStream_t st = (hipStream_t)stream;
  Workspace w;
  carve(b, sizeof(T), (char*)ws, &w);
  const int bs = b->block;
  const int TX = b->tile[0], TY = b->tile[1], TZ = b->tile[2];
  if (!ready) {
    TimedLaunch tl(MPM_K_PREP_FWD, st);
    clear_bins(b, w, st);
    per_model(b, bs, [&](auto M, const int32_t* ch, dim3 g) {
      k_prep_fwd<T, decltype(M)::value><<<g, bs, 0, st>>>(
          b->inst, b->slot_inst, b->n_slots, dyn_in, Fg, sched, TX, TY, TZ, payload, w.key,
          w.count, w.unit_flag, ch);
    });
  }
  rc = sort_from_keys(b, sortbuf, w, st);
  if (rc) return rc;
  const SortView<const int32_t> sv(b, sortbuf);
  dispatch_fwd<T>(tile_index(b), b, sv, w, payload, grid, st);
  if (out == MPM_OUT_NONE) return check_launch("step_fwd");   // the step record only
  TimedLaunch tl(MPM_K_G2P, st);
  if (payload_next)
    per_model(b, bs, [&](auto M, const int32_t* ch, dim3 g) {
      k_g2p<T, decltype(M)::value, true><<<g, bs, 0, st>>>(
          b->inst, b->slot_inst, b->n_slots, dyn_in, grid, dyn_out, advect, Fg, sched_next, TX,
          TY, TZ, payload_next, w.key, w.count, w.unit_flag, out, ch);
    });
  else   // the G2P part does not depend on the model
    k_g2p<T, MPM_NEO_HOOKEAN, false><<<grid1(b->n_slots, bs), bs, 0, st>>>(
        b->inst, b->slot_inst, b->n_slots, dyn_in, grid, dyn_out, advect, nullptr, nullptr, TX,
        TY, TZ, nullptr, nullptr, nullptr, nullptr, out, nullptr);
  return check_launch("step_fwd");
}

// One reverse step.  ready = 1 or 2: this step's G2P-adjoint payload was
// written into payload area ready - 1 of the workspace by the previous chained
// reverse call (the later step's); ready = 0: k_prep_adj computes it into area
// 0.  dyn_prev / sortbuf_prev (both or neither): the previous step's input
// state and sort record; the per-particle kernel then also writes that step's
// G2P-adjoint payload, into the other area (its blocks still read this step's
// payload while others write the next one).
template <typename T>
int step_bwd_impl(const mpm_batch* b, const T* dyn_in, const T* Fg, const T* sched,
                  const T* dyn_out, const T* grid, const int32_t* sortbuf, const T* payload,
                  const T* bar_out, T* bar_in, T* Fg_bar, double* theta_acc, int advect, int ready,
                  const T* dyn_prev, const int32_t* sortbuf_prev, void* ws, void* stream) {
  int rc = validate(b);
  if (rc) return rc;
  if (!dyn_in || !Fg || !dyn_out || !grid || !sortbuf || !payload || !bar_out || !bar_in ||
      !Fg_bar || !ws)
    return set_err(MPM_EINVAL, "step_bwd: NULL buffer");
  if (bar_in == bar_out) return set_err(MPM_EINVAL, "step_bwd: bar_in must not alias bar_out");
  if ((dyn_prev == nullptr) != (sortbuf_prev == nullptr))
    return set_err(MPM_EINVAL, "step_bwd: dyn_prev and sortbuf_prev go together");
  if (ready < 0 || ready > 2) return set_err(MPM_EINVAL, "step_bwd: ready must be 0, 1 or 2");
  hipStream_t st = (hipStream_t)stream;
  Workspace w;
  carve(b, sizeof(T), (char*)ws, &w);
  const SortView<const int32_t> sv(b, sortbuf);
  const int bs = b->block;
  T* const area[2] = {(T*)w.payload, (T*)w.payload2};
  T* const pa = area[ready == 2 ? 1 : 0];
  T* const pnext = area[ready == 2 ? 0 : 1];
  if (!ready) {
    TimedLaunch tl(MPM_K_PREP_ADJ, st);
    k_prep_adj<T><<<grid1(b->n_slots, bs), bs, 0, st>>>(b->inst, b->slot_inst, b->n_slots, dyn_in,
                                                        bar_out, sv.rank, pa, advect);
  }
  dispatch_adj<T>(tile_index(b), b, sv, w, pa, grid, advect, st);
  TimedLaunch tl(MPM_K_ADJ_PARTICLE, st);
  per_model(b, 256, [&](auto M, const int32_t* ch, dim3 g) {
    constexpr int MODEL = decltype(M)::value;
    if (dyn_prev)
      k_adj_particle<T, MODEL, true><<<g, 256, 0, st>>>(
          b->inst, b->slot_inst, b->n_slots, sv.rank, payload, pa, grid, (const T*)w.gadj,
          dyn_in, Fg, sched, dyn_out, bar_out, bar_in, Fg_bar, theta_acc, dyn_prev,
          SortView<const int32_t>(b, sortbuf_prev).rank, pnext, advect, ch);
    else
      k_adj_particle<T, MODEL, false><<<g, 256, 0, st>>>(
          b->inst, b->slot_inst, b->n_slots, sv.rank, payload, pa, grid, (const T*)w.gadj,
          dyn_in, Fg, sched, dyn_out, bar_out, bar_in, Fg_bar, theta_acc, nullptr, nullptr,
          nullptr, advect, ch);
  });
  return check_launch("step_bwd");
}

template <typename T>
int stress_impl(int32_t model, const double* prm6, int64_t n, int64_t stride, const T* F,
                const T* Fg, const T* C, T* tau, void* stream) {
  if (model < 0 || model > 2) return set_err(MPM_EINVAL, "unknown model %d", model);
  if (!prm6 || !F || !Fg || !tau || n < 0 || stride < n)
    return set_err(MPM_EINVAL, "stress: bad arguments");
  if (model == MPM_HERSCHEL_BULKLEY && !C) return set_err(MPM_EINVAL, "stress: HB needs C");
  if (n == 0) return MPM_OK;
  ModelParams<T> mp;
  mp.kind = model;
  for (int k = 0; k < 6; ++k) mp.p[k] = (T)prm6[k];
  with_model(1 << model, [&](auto M) {
    k_stress<T, decltype(M)::value><<<grid1(n, 256), 256, 0, (hipStream_t)stream>>>(
        mp, n, stride, F, Fg, C, tau);
  });
  return check_launch("stress");
}

template <typename T>
int stress_vjp_impl(int32_t model, const double* prm6, int64_t n, int64_t stride, const T* F,
                    const T* Fg, const T* C, const T* tb, T* Fb, T* Fgb, T* Cb, T* thb,
                    void* stream) {
  if (model < 0 || model > 2) return set_err(MPM_EINVAL, "unknown model %d", model);
  if (!prm6 || !F || !Fg || !tb || !Fb || !Fgb || n < 0 || stride < n)
    return set_err(MPM_EINVAL, "stress_vjp: bad arguments");
  if (model == MPM_HERSCHEL_BULKLEY && !C) return set_err(MPM_EINVAL, "stress_vjp: HB needs C");
  if (n == 0) return MPM_OK;
  ModelParams<T> mp;
  mp.kind = model;
  for (int k = 0; k < 6; ++k) mp.p[k] = (T)prm6[k];
  with_model(1 << model, [&](auto M) {
    k_stress_vjp<T, decltype(M)::value><<<grid1(n, 256), 256, 0, (hipStream_t)stream>>>(
        mp, n, stride, F, Fg, C, tb, Fb, Fgb, Cb, thb);
  });
  return check_launch("stress_vjp");
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free forward step (mpm_step_fwd_noalloc_f32): a
// fixed device array, one copy per GPU, holding the forward step's workspace
// regions (sort scratch and tile partials).  Its device address on each GPU
// comes from hipGetSymbolAddress, looked up once per device on the first
// call there; host code never passes the array's own name to a kernel.
// Every call on a GPU shares that GPU's array: calls that overlap on one GPU
// (other streams, other threads) race.  Each call clears the counters and
// flags it needs before use (a fill kernel, not a memset node).
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocScratchBytes = int64_t(64) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

int scratch_on_current_device(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return set_err(MPM_ELAUNCH, "step_fwd_noalloc: hipGetDevice: %s",
                                      hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices)
    return set_err(MPM_ENOTSUP, "step_fwd_noalloc: device %d beyond the %d supported", dev,
                   kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      return set_err(MPM_ELAUNCH, "step_fwd_noalloc: hipGetSymbolAddress(scratch): %s",
                     hipGetErrorString(e));
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return MPM_OK;
}

}  // namespace

int mpm_internal::validate_batch(const mpm_batch* b) { return validate(b); }

extern "C" {

int32_t mpm_abi_version(void) { return MPM_ABI_VERSION; }

int32_t mpm_kernel_timer(int32_t kernel_id, void* const* events, int32_t n_pairs) {
  if (kernel_id >= 0 && (events == nullptr || n_pairs <= 0))
    return set_err(MPM_EINVAL, "kernel_timer: events required");
  g_timer.id = kernel_id;
  g_timer.ev = (hipEvent_t*)events;
  g_timer.n_pairs = kernel_id >= 0 ? n_pairs : 0;
  g_timer.count = 0;
  return MPM_OK;
}
int32_t mpm_kernel_timer_count(void) { return g_timer.count; }
const char* mpm_last_error(void) { return g_err; }

int64_t mpm_workspace_bytes(const mpm_batch* b, int32_t dtype) {
  const int rc = validate(b);
  if (rc) return rc;
  if (dtype != 0 && dtype != 1) return set_err(MPM_EINVAL, "dtype must be 0 (f32) or 1 (f64)");
  return carve(b, dtype ? sizeof(double) : sizeof(float), nullptr, nullptr);
}

int32_t mpm_tile_shapes(int32_t* out, int32_t max_shapes) {
  const int n = kNumTiles < max_shapes ? kNumTiles : max_shapes;
  for (int i = 0; i < n && out; ++i) {
    out[3 * i] = kTiles[i].x;
    out[3 * i + 1] = kTiles[i].y;
    out[3 * i + 2] = kTiles[i].z;
  }
  return kNumTiles;
}

int32_t mpm_stress_f32(int32_t model, const double* prm6, int64_t n, int64_t stride, const float* F,
                       const float* Fg, const float* C, float* tau, void* stream) {
  return stress_impl<float>(model, prm6, n, stride, F, Fg, C, tau, stream);
}
int32_t mpm_stress_f64(int32_t model, const double* prm6, int64_t n, int64_t stride,
                       const double* F, const double* Fg, const double* C, double* tau,
                       void* stream) {
  return stress_impl<double>(model, prm6, n, stride, F, Fg, C, tau, stream);
}
int32_t mpm_stress_vjp_f32(int32_t model, const double* prm6, int64_t n, int64_t stride,
                           const float* F, const float* Fg, const float* C, const float* tau_bar,
                           float* F_bar, float* Fg_bar, float* C_bar, float* theta_bar,
                           void* stream) {
  return stress_vjp_impl<float>(model, prm6, n, stride, F, Fg, C, tau_bar, F_bar, Fg_bar, C_bar,
                                theta_bar, stream);
}
int32_t mpm_stress_vjp_f64(int32_t model, const double* prm6, int64_t n, int64_t stride,
                           const double* F, const double* Fg, const double* C,
                           const double* tau_bar, double* F_bar, double* Fg_bar, double* C_bar,
                           double* theta_bar, void* stream) {
  return stress_vjp_impl<double>(model, prm6, n, stride, F, Fg, C, tau_bar, F_bar, Fg_bar, C_bar,
                                 theta_bar, stream);
}

int32_t mpm_sort_by_cell_f32(const mpm_batch* b, const float* dyn, int32_t* sortbuf, void* ws,
                             void* stream) {
  int rc = validate(b);
  if (rc) return rc;
  if (!dyn || !sortbuf || !ws) return set_err(MPM_EINVAL, "sort_by_cell: NULL buffer");
  Workspace w;
  carve(b, sizeof(float), (char*)ws, &w);
  return sort_impl<float>(b, dyn, sortbuf, w, (hipStream_t)stream);
}
int32_t mpm_sort_by_cell_f64(const mpm_batch* b, const double* dyn, int32_t* sortbuf, void* ws,
                             void* stream) {
  int rc = validate(b);
  if (rc) return rc;
  if (!dyn || !sortbuf || !ws) return set_err(MPM_EINVAL, "sort_by_cell: NULL buffer");
  Workspace w;
  carve(b, sizeof(double), (char*)ws, &w);
  return sort_impl<double>(b, dyn, sortbuf, w, (hipStream_t)stream);
}

int32_t mpm_step_fwd_f32(const mpm_batch* b, const float* dyn_in, const float* Fg,
                         const float* sched, float* dyn_out, float* grid, int32_t* sortbuf,
                         float* payload, int32_t advect, void* ws, void* stream) {
  return step_fwd_impl<float>(b, dyn_in, Fg, sched, nullptr, dyn_out, grid, sortbuf, payload,
                              nullptr, advect, 0, MPM_OUT_FULL, ws, stream);
}
int32_t mpm_step_fwd_f64(const mpm_batch* b, const double* dyn_in, const double* Fg,
                         const double* sched, double* dyn_out, double* grid, int32_t* sortbuf,
                         double* payload, int32_t advect, void* ws, void* stream) {
  return step_fwd_impl<double>(b, dyn_in, Fg, sched, nullptr, dyn_out, grid, sortbuf, payload,
                               nullptr, advect, 0, MPM_OUT_FULL, ws, stream);
}
int32_t mpm_step_fwd_noalloc_f32(const mpm_batch* b, const float* dyn_in, const float* Fg,
                                 const float* sched, float* dyn_out, float* grid,
                                 int32_t* sortbuf, float* payload, int32_t advect,
                                 void* stream) {
  int rc = validate(b);
  if (rc) return rc;
  int64_t need = 0;
  carve(b, sizeof(float), nullptr, nullptr, &need);
  if (need > kNoAllocScratchBytes)
    return set_err(MPM_EINVAL,
                   "step_fwd_noalloc: this batch needs %lld bytes of scratch, more than the "
                   "%lld the library holds; use mpm_step_fwd_f32 with a workspace",
                   (long long)need, (long long)kNoAllocScratchBytes);
  void* ws = nullptr;
  rc = scratch_on_current_device(&ws);
  if (rc) return rc;
  // ready = 0: the step clears the bin counters and unit flags it uses first
  return step_fwd_impl<float>(b, dyn_in, Fg, sched, nullptr, dyn_out, grid, sortbuf, payload,
                              nullptr, advect, 0, MPM_OUT_FULL, ws, stream);
}

int32_t mpm_step_fwd_chain_f32(const mpm_batch* b, const float* dyn_in, const float* Fg,
                               const float* sched, const float* sched_next, float* dyn_out,
                               float* grid, int32_t* sortbuf, float* payload, float* payload_next,
                               int32_t advect, int32_t ready, int32_t out, void* ws,
                               void* stream) {
  return step_fwd_impl<float>(b, dyn_in, Fg, sched, sched_next, dyn_out, grid, sortbuf, payload,
                              payload_next, advect, ready, out, ws, stream);
}
int32_t mpm_step_fwd_chain_f64(const mpm_batch* b, const double* dyn_in, const double* Fg,
                               const double* sched, const double* sched_next, double* dyn_out,
                               double* grid, int32_t* sortbuf, double* payload,
                               double* payload_next, int32_t advect, int32_t ready, int32_t out,
                               void* ws, void* stream) {
  return step_fwd_impl<double>(b, dyn_in, Fg, sched, sched_next, dyn_out, grid, sortbuf, payload,
                               payload_next, advect, ready, out, ws, stream);
}
int32_t mpm_step_bwd_f32(const mpm_batch* b, const float* dyn_in, const float* Fg,
                         const float* sched, const float* dyn_out, const float* grid,
                         const int32_t* sortbuf, const float* payload, const float* bar_out,
                         float* bar_in, float* Fg_bar, double* theta_acc, int32_t advect, void* ws,
                         void* stream) {
  return step_bwd_impl<float>(b, dyn_in, Fg, sched, dyn_out, grid, sortbuf, payload, bar_out,
                              bar_in, Fg_bar, theta_acc, advect, 0, nullptr, nullptr, ws, stream);
}
int32_t mpm_step_bwd_f64(const mpm_batch* b, const double* dyn_in, const double* Fg,
                         const double* sched, const double* dyn_out, const double* grid,
                         const int32_t* sortbuf, const double* payload, const double* bar_out,
                         double* bar_in, double* Fg_bar, double* theta_acc, int32_t advect,
                         void* ws, void* stream) {
  return step_bwd_impl<double>(b, dyn_in, Fg, sched, dyn_out, grid, sortbuf, payload, bar_out,
                               bar_in, Fg_bar, theta_acc, advect, 0, nullptr, nullptr, ws, stream);
}
int32_t mpm_step_bwd_chain_f32(const mpm_batch* b, const float* dyn_in, const float* Fg,
                               const float* sched, const float* dyn_out, const float* grid,
                               const int32_t* sortbuf, const float* payload, const float* bar_out,
                               float* bar_in, float* Fg_bar, double* theta_acc, int32_t advect,
                               int32_t ready, const float* dyn_prev, const int32_t* sortbuf_prev,
                               void* ws, void* stream) {
  return step_bwd_impl<float>(b, dyn_in, Fg, sched, dyn_out, grid, sortbuf, payload, bar_out,
                              bar_in, Fg_bar, theta_acc, advect, ready, dyn_prev, sortbuf_prev, ws,
                              stream);
}
int32_t mpm_step_bwd_chain_f64(const mpm_batch* b, const double* dyn_in, const double* Fg,
                               const double* sched, const double* dyn_out, const double* grid,
                               const int32_t* sortbuf, const double* payload,
                               const double* bar_out, double* bar_in, double* Fg_bar,
                               double* theta_acc, int32_t advect, int32_t ready,
                               const double* dyn_prev, const int32_t* sortbuf_prev, void* ws,
                               void* stream) {
  return step_bwd_impl<double>(b, dyn_in, Fg, sched, dyn_out, grid, sortbuf, payload, bar_out,
                               bar_in, Fg_bar, theta_acc, advect, ready, dyn_prev, sortbuf_prev, ws,
                               stream);
}

}  // extern "C"
