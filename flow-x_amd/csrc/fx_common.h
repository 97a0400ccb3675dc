// Shared device helpers and host plumbing for libflowx_hip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "flowx.h"

namespace fx {

// ---- error handling (thread-local last error, returned as codes) ---------
void set_error(const std::string& msg);
int fail(int code, const char* fmt, ...);
int check_launch(const char* what);

#define FX_HIP_CHECK(expr)                                                     \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::fx::fail(FX_EHIP, "%s: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

#define FX_TRY(expr)              \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != FX_OK) return rc_; \
  } while (0)

constexpr float kPi = 3.14159265358979323846f;  // GLSL float literal PI

// Derived per-launch constants, computed on the host in float exactly as the
// GLSL computes them per invocation (shaders/sph_common.glsl:68-81).
struct Consts {
  float h, h2, mass, rho0, eps;
  float poly6, spiky, visc_lap;
  float w_q;          // W_poly6((0.2h)^2) of shaders/sph_delta.glsl:44
  float scorr_k;
  float visc;
  float dt;
  int max_scan;
  float delta_scale;  // 1, or the mass with FX_FLAG_MASS_WEIGHTED_DELTA
};

Consts make_consts(const FxParams& p);
float poly6_coef(float h);

// ---- grid geometry --------------------------------------------------------
struct Grid {
  float lox, loy, loz, cell;
  int cx, cy, cz, nc;
};

inline Grid make_grid(const FxParams& p) {
  return Grid{p.lo[0], p.lo[1], p.lo[2], p.cell_size,
              p.cells[0], p.cells[1], p.cells[2], p.n_cells};
}

// shaders/sph_common.glsl:50-54 cell_coord: clamp(ivec3(floor((p-lo)/cell)))
// with IEEE division (no contraction possible: sub, div, floor).
__device__ __forceinline__ int floor_to_int(float v) {
  float f = floorf(v);
  // keep the conversion defined for huge / NaN values; clamp() follows anyway
  f = fminf(fmaxf(f, -1.0e9f), 1.0e9f);
  return (int)f;
}

__device__ __forceinline__ void cell_coord(const Grid& g, float px, float py, float pz,
                                           int& x, int& y, int& z) {
  x = min(max(floor_to_int((px - g.lox) / g.cell), 0), g.cx - 1);
  y = min(max(floor_to_int((py - g.loy) / g.cell), 0), g.cy - 1);
  z = min(max(floor_to_int((pz - g.loz) / g.cell), 0), g.cz - 1);
}

__device__ __forceinline__ int cell_index(const Grid& g, int x, int y, int z) {
  return (z * g.cy + y) * g.cx + x;
}

// ---- collider (shaders/sph_common.glsl:132-160) ---------------------------
struct Collider {
  const uint8_t* occ;  // NULL = none
  float lox, loy, loz, voxel;
  int kx, ky, kz;
};

inline Collider make_collider(const FxParams& p, const uint8_t* occ) {
  Collider c{occ, p.lo[0], p.lo[1], p.lo[2], p.collider_voxel,
             p.collider_dims[0], p.collider_dims[1], p.collider_dims[2]};
  if (!(p.collider_voxel > 0.0f)) c.occ = nullptr;
  return c;
}

__device__ __forceinline__ bool collider_occupied(const Collider& c, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= c.kx || y >= c.ky || z >= c.kz) return false;
  return c.occ[((size_t)z * c.ky + y) * c.kx + x] != 0;
}

__device__ __forceinline__ void collider_coord(const Collider& c, float px, float py, float pz,
                                               int& x, int& y, int& z) {
  x = floor_to_int((px - c.lox) / c.voxel);
  y = floor_to_int((py - c.loy) / c.voxel);
  z = floor_to_int((pz - c.loz) / c.voxel);
}

// ---- exact pair arithmetic ---------------------------------------------------
// r^2 = dot(d, d) evaluated as (x*x + y*y) + z*z with every operation rounded
// (no FMA), so the r < h decisions and the density sums below match a
// float32 CPU evaluation of the GLSL bit for bit.
__device__ __forceinline__ float dist2_exact(float dx, float dy, float dz) {
#pragma clang fp contract(off)
  return dx * dx + dy * dy + dz * dz;
}

// mass * W_poly6 with W = ((coef * d) * d) * d, d = h^2 - r^2
// (shaders/sph_common.glsl:84-91, shaders/sph_lambda.glsl:57), and its
// addition to a running sum, each rounded as written.
__device__ __forceinline__ float poly6_term_exact(float mass, float coef, float d) {
#pragma clang fp contract(off)
  return mass * (((coef * d) * d) * d);
}
__device__ __forceinline__ float add_exact(float acc, float t) {
#pragma clang fp contract(off)
  return acc + t;
}
__device__ __forceinline__ float poly6_accum_exact(float acc, float mass, float coef, float d) {
  return add_exact(acc, poly6_term_exact(mass, coef, d));
}

// ---- neighbour iteration ----------------------------------------------------
// Visits the reference's 27-cell stencil around (bx,by,bz) in its order
// (dz, dy, dx from -1 to 1, z outermost) as 9 contiguous x-runs of sorted
// slots.  A run of 3 cells whose total is <= max_scan cannot contain a cell
// above MAX_CELL_SCAN, so it is visited whole; otherwise each cell is visited
// separately with end = min(end, start + max_scan)
// (shaders/sph_lambda.glsl:49-56).  off[c] is the first slot of cell c.
template <int R, class F>
__device__ __forceinline__ void for_each_run(const Grid& g, const int32_t* __restrict__ off,
                                             int max_scan, int bx, int by, int bz, F&& f) {
  const int x0 = max(bx - R, 0), x1 = min(bx + R, g.cx - 1);
  for (int dz = -R; dz <= R; ++dz) {
    const int z = bz + dz;
    if (z < 0 || z >= g.cz) continue;
    for (int dy = -R; dy <= R; ++dy) {
      const int y = by + dy;
      if (y < 0 || y >= g.cy) continue;
      const int c0 = (z * g.cy + y) * g.cx + x0, c1 = c0 + (x1 - x0);
      const int s = off[c0], e = off[c1 + 1];
      if (e - s <= max_scan) {
        f(s, e);
      } else {
        for (int c = c0; c <= c1; ++c) {
          const int cs = off[c], ce = min(off[c + 1], cs + max_scan);
          f(cs, ce);
        }
      }
    }
  }
}

// Scan slots [s, e) of a sorted-space array.  The loads of FX_BATCH
// candidates are issued before any of them is used, so a wave keeps several
// gathers in flight instead of waiting out each one (a load behind the
// previous candidate's divergent branches is otherwise not hoisted).
// refine(q, d) runs on every loaded candidate before any f(q, d) call and may
// issue further (predicated) loads into d.
#ifndef FX_BATCH
#define FX_BATCH 4
#endif
struct NoRefine {
  template <class D>
  __device__ __forceinline__ void operator()(int, D&) const {}
};

template <int B = FX_BATCH, class LD, class F, class RF = NoRefine>
__device__ __forceinline__ void scan_run(int s, int e, LD&& ld, F&& f, RF&& refine = RF()) {
  int q = s;
  for (; q + B <= e; q += B) {
    decltype(ld(q)) d[B];
#pragma unroll
    for (int u = 0; u < B; ++u) d[u] = ld(q + u);
#pragma unroll
    for (int u = 0; u < B; ++u) refine(q + u, d[u]);
#pragma unroll
    for (int u = 0; u < B; ++u) f(q + u, d[u]);
  }
  for (; q < e; ++q) {
    auto d = ld(q);
    refine(q, d);
    f(q, d);
  }
}

// ---- four lanes per item ----------------------------------------------------
// Lane J's value of v, read by every lane of its quad (DPP quad_perm).  All
// four lanes of the quad must be active.
template <int J>
__device__ __forceinline__ float quad_lane(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), J * 0x55, 0xF, 0xF,
                                                              false));
}

// Slots [s, e) shared by the four lanes of a quad (lane j of 0..3): in each
// round lane j takes slot k + j, ld(q) loads it (q clamped into the run),
// tm(q, data, valid) makes its term, and fold(term) is called with the four
// lanes' terms in slot order on every lane (term.lane<J>() reads lane J's).
// B rounds of loads are in flight per lane.  The quad must run this with the
// same s, e.
template <int B, class LD, class TM, class FD>
__device__ __forceinline__ void quad_run(int s, int e, int j, LD&& ld, TM&& tm, FD&& fold) {
#pragma unroll 1
  for (int k = s; k < e; k += 4 * B) {
    decltype(ld(0)) d[B];
#pragma unroll
    for (int u = 0; u < B; ++u) d[u] = ld(min(k + 4 * u + j, e - 1));
    decltype(tm(0, d[0], false)) tt[B];
#pragma unroll
    for (int u = 0; u < B; ++u) tt[u] = tm(k + 4 * u + j, d[u], k + 4 * u + j < e);
#pragma unroll
    for (int u = 0; u < B; ++u) {
      if (k + 4 * u < e) {  // the same for the four lanes
        fold(tt[u].template lane<0>());
        fold(tt[u].template lane<1>());
        fold(tt[u].template lane<2>());
        fold(tt[u].template lane<3>());
      }
    }
  }
}

// ---- block-aggregated counting-sort ranks ---------------------------------------
// The counting sort's rank of a particle in its cell is its arrival order at
// the cell's counter.  One returning global atomic per particle made that the
// larger part of the fused finalize + predict pass (84 us at 1.64M particles,
// 29 us with the atomic taken out).  Here the threads of a block first gather
// their keys in an LDS hash table (one entry per distinct key, its count
// taken with LDS atomics); one thread per distinct key then adds the block's
// count to the global counter and gets the block's base back, and a thread's
// rank is that base plus its arrival order in the block.  Particles in slot
// order (the resident substeps) put a whole cell's run - about eight
// particles - into consecutive lanes, whose same-address atomics serialize;
// a block of 256 holds some 30-40 distinct cells, so the global atomics drop
// several-fold.  (In lattice order a wave holds about two particles per cell
// and the plain atomics are cheaper.)  Ranks within a cell stay in arrival
// order: the canonicalisation after the scatter orders each cell's run.
// Every thread of the block must call it (barriers); `active` false for
// threads past the end.  Keys are >= 0.
template <int BLOCK>
__device__ __forceinline__ int block_rank(int key, bool active, int32_t* __restrict__ cnt) {
  constexpr int H = 2 * BLOCK;  // hash slots: at most BLOCK distinct keys
  static_assert((H & (H - 1)) == 0, "power of two");
  __shared__ int hkey[H];
  __shared__ int hval[H];  // the block's count of the key, then its base
  for (int i = threadIdx.x; i < H; i += BLOCK) {
    hkey[i] = -1;
    hval[i] = 0;
  }
  __syncthreads();
  int slot = 0, local = 0;
  if (active) {
    slot = (int)(((unsigned)key * 2654435761u) >> 16) & (H - 1);
    for (int probe = 0; probe < H; ++probe) {  // ends: fewer keys than slots
      const int prev = atomicCAS(&hkey[slot], -1, key);
      if (prev == -1 || prev == key) break;
      slot = (slot + 1) & (H - 1);
    }
    local = atomicAdd(&hval[slot], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < H; i += BLOCK) {
    const int c = hval[i];
    if (c > 0) hval[i] = atomicAdd(&cnt[hkey[i]], c);
  }
  __syncthreads();
  return hval[slot] + local;
}

// ---- launch helpers ---------------------------------------------------------
constexpr int kBlock = 256;

// Optional per-kernel timing (fx_profile): when enabled, every launch made
// through FX_LAUNCH is bracketed by a pair of HIP events on its own stream.
bool profiling_enabled();
struct ProfScope {
  int slot;
  hipStream_t st;
  ProfScope(const char* name, hipStream_t stream);
  ~ProfScope();
};

#define FX_LAUNCH(kernel, grid, block, st, ...)                  \
  do {                                                          \
    ::fx::ProfScope prof_(#kernel, st);                         \
    hipLaunchKernelGGL(kernel, grid, block, 0, st, __VA_ARGS__); \
  } while (0)
// the same with `shm` bytes of dynamic LDS per workgroup
#define FX_LAUNCH_SHM(kernel, grid, block, shm, st, ...)           \
  do {                                                            \
    ::fx::ProfScope prof_(#kernel, st);                           \
    hipLaunchKernelGGL(kernel, grid, block, shm, st, __VA_ARGS__); \
  } while (0)

// Workgroups are dispatched round-robin over the chip's 8 XCDs, each with
// its own L2.  xcd_block() renumbers the blocks of a 1-D launch so that each
// XCD runs one contiguous range of them: the neighbour gathers of a sorted
// particle range then hit lines its own L2 already holds, instead of every
// L2 pulling the whole position array.  A bijection on [0, gridDim.x).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, G = gridDim.x;
  const int x = b % kXcds, local = b / kXcds;
  const int per = (G + kXcds - 1) / kXcds;
  const int tall = G % kXcds == 0 ? kXcds : G % kXcds;  // XCDs that receive `per` blocks
  return x < tall ? x * per + local : tall * per + (x - tall) * (per - 1) + local;
}

inline unsigned grid_for(long long n, int block = kBlock) {
  long long g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

// Workspace layout for the solver (offsets in bytes, 256-B aligned).
struct Workspace {
  size_t cnt, off, rank, key, sidx, skey, pred_s0, pred_s1, pos_s, vel_s, dl_s, mask, lcell, lfull, disp, dcell,
      partial, flags, normal_next, sdcell;
  // views into the neighbour-list masks (unused while the lists are off):
  // the state carried in sorted order between the substeps of one call
  size_t pair64, pos_r, vel_r;
  // candidate positions as three float arrays (x, y, z), read by lambda's
  // recording walk four slots to a load
  size_t sx, sy, sz;
  size_t total;
};

Workspace workspace_layout(const FxParams& p);

// Library-wide tuning (fx_set_tuning): neighbour-list skin in units of h
// (0 disables the lists) and the block size of the neighbour passes.
struct Tuning {
  float skin;
  int block;
  int persistent;  // batched passes: workgroups per CU of a persistent launch, 0 = flat
  int walk;        // large-scene neighbour passes: 0 default walk, 1 row by row, 2 the other stream walk,
                   // 3 lambda and delta through LDS tiles of the workgroup's stencil rows, 4 the default
                   // walk without the pair masks lambda records for delta
};
Tuning tuning();

// Workspace header: words kHdr.. of the workspace's flags block record what
// the workspace holds.  The kernels that build a grid or compute normals write
// it, so the record travels with the memory, not with its address: a fresh or
// reused allocation carries no claim it cannot back.
//   magic, n, n_cells   a grid (offsets, sorted slots, sorted-space state)
//                       for n particles over n_cells cells is built
//   normals             0: none; 1: the next substep's normals are in the
//                       workspace's normal_next (by particle); 2: they are
//                       already in the caller's normal buffer
//   h, mass, nlo, nhi   the smoothing radius and mass (float bits) and the
//                       normal buffer's address those normals were made for
//   error               set by a kernel that found no grid (a captured launch
//                       on a workspace that never had one); cleared by the
//                       next grid build
// Passes that read the previous grid (the normal pass, the splat) check it on
// the device, and the host entry points read it back when they are not under
// stream capture.
constexpr int32_t kGridMagic = 0x31475846;  // "FXG1"
enum HeaderWord : int {
  kHdr = 16,  // first word, in the flags block
  kHdrMagic = kHdr,
  kHdrN,
  kHdrCells,
  kHdrNormals,
  kHdrH,
  kHdrMass,
  kHdrNormLo,
  kHdrNormHi,
  kHdrError,
  kHdrEnd
};

__device__ __forceinline__ bool header_grid_ok(const int32_t* flags, int n, int nc) {
  return flags[kHdrMagic] == kGridMagic && flags[kHdrN] == n && flags[kHdrCells] == nc;
}

// What a usable normals record must carry (normals computed for this
// normal buffer with these constants).
struct NormalKey {
  int32_t hbits, mbits;
  uint32_t lo, hi;
};
inline NormalKey normal_key(const FxParams& p, const void* normal) {
  NormalKey k;
  memcpy(&k.hbits, &p.h, 4);
  memcpy(&k.mbits, &p.mass, 4);
  const uint64_t a = (uint64_t)(uintptr_t)normal;
  k.lo = (uint32_t)a;
  k.hi = (uint32_t)(a >> 32);
  return k;
}

// Host check that the workspace at `work` (solver layout wl) holds a grid
// for (n, n_cells): a synchronous read of the header, skipped while `st` is
// capturing (the kernels check it themselves then).  `what` names the caller
// in the error.
int require_grid(const Workspace& wl, const void* work, int n, int n_cells, hipStream_t st, const char* what);

// Exclusive scan of int32 counts[0..n) into out[0..n] (out[n] = total).
// partial must hold scan_partials(n) ints.  Launches 3 kernels.
size_t scan_partials(long long n);
int exclusive_scan_i32(const int32_t* counts, int32_t* out, long long n, int32_t* partial,
                       hipStream_t st);

// Reorder each cell's run sidx[off[c], off[c+1]) to ascending particle index
// (the counting sort's atomic ranks leave it in arrival order).  Scratch:
// big_list receives the cells of more than 128 particles (at most n / 129),
// big_count one int.
int canon_cells(const int32_t* off, int32_t* sidx, int n_cells, int32_t* big_list, int32_t* big_count,
                hipStream_t st);
// Sort the queued runs data[off[q], off[q+1]), q in queue[0, *queue_count),
// ascending in place, one workgroup per run.
int sort_queued_runs(const int32_t* off, int32_t* data, const int32_t* queue, const int32_t* queue_count,
                     hipStream_t st);

}  // namespace fx
