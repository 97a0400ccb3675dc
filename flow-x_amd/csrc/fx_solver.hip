// PBF solver kernels for gfx950 and the C ABI of the step / reset / grid path.
//
// Layout.  The caller's state is canonical (particle i at row i, AoS-4, the
// reference texture layout).  Each grid build sorts particles by cell with a
// deterministic counting sort (ascending cell, then ascending particle index)
// and gathers them into "sorted space" inside the workspace, where slot t holds
// particle sidx[t].  All neighbour passes run in sorted space: the particles of
// one cell row are contiguous, so a 3-cell x-run of the reference's 27-cell
// stencil (shaders/sph_lambda.glsl:39-70) is one contiguous slot range, loads
// of neighbours are 16-byte float4 gathers from a few cache lines per wave,
// and the neighbour-sum order is the reference's (dz, dy, dx, slot) order.
// Results go back to canonical rows at the end of the substep.
//
// Exactness.  Streaming passes (predict, grid key, apply, velocity,
// finalize) are compiled without FMA contraction and use IEEE division and
// sqrt, so they match a float32 CPU evaluation of the GLSL bit for bit.  The
// neighbour sums allow contraction and use v_rsq for 1/r; they match to a few
// float32 ulps of each term.
#include <cstdarg>
#include <cstring>
#include <limits>
#include <mutex>
#include <map>
#include <unordered_map>
#include <utility>
#include <vector>

#include "fx_common.h"

#pragma clang fp contract(off)

namespace fx {

// ---- errors ----------------------------------------------------------------
static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(FX_EHIP, "%s: %s", what, hipGetErrorString(e));
  return FX_OK;
}

// ---- per-kernel timing ------------------------------------------------------
struct ProfRec {
  const char* name;
  hipEvent_t a, b;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof_recs;
static size_t g_prof_used = 0;

bool profiling_enabled() { return g_prof_on; }

ProfScope::ProfScope(const char* name, hipStream_t stream) : slot(-1), st(stream) {
  if (!g_prof_on) return;
  std::lock_guard<std::mutex> lock(g_prof_mu);
  if (g_prof_used == g_prof_recs.size()) {
    ProfRec r{name, nullptr, nullptr};
    if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) return;
    g_prof_recs.push_back(r);
  }
  slot = (int)g_prof_used++;
  g_prof_recs[slot].name = name;
  (void)hipEventRecord(g_prof_recs[slot].a, st);
}

ProfScope::~ProfScope() {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lock(g_prof_mu);
  (void)hipEventRecord(g_prof_recs[slot].b, st);
}

static Tuning g_tuning = {0.0f, 64, 0, 0};

Tuning tuning() { return g_tuning; }

int require_grid(const Workspace& wl, const void* work, int n, int n_cells, hipStream_t st, const char* what) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  FX_HIP_CHECK(hipStreamIsCapturing(st, &cs));
  if (cs != hipStreamCaptureStatusNone) return FX_OK;
  int32_t h[kHdrEnd - kHdr];
  const char* src = (const char*)work + wl.flags + kHdr * sizeof(int32_t);
  FX_HIP_CHECK(hipMemcpyAsync(h, src, sizeof(h), hipMemcpyDeviceToHost, st));
  FX_HIP_CHECK(hipStreamSynchronize(st));
  // (the error word is set only by a kernel that found no grid, and every
  // grid build clears it, so it matters only when no grid is recorded)
  if (h[kHdrMagic - kHdr] != kGridMagic || h[kHdrN - kHdr] != n || h[kHdrCells - kHdr] != n_cells)
    return fail(FX_EINVAL,
                "%s reads the previous grid, and this workspace holds none for %d particles / %d cells: "
                "call fx_load_state (or fx_build_grid) first%s",
                what, n, n_cells,
                h[kHdrError - kHdr] == 1 ? " (a captured launch already ran on it: its normals or field are zero)"
                                         : "");
  return FX_OK;
}

// ---- constants (shaders/sph_common.glsl:68-91, sph_delta.glsl:44) --------
float poly6_coef(float h) { return 315.0f / (64.0f * kPi * powf(h, 9.0f)); }

Consts make_consts(const FxParams& p) {
  Consts k;
  k.h = p.h;
  k.h2 = p.h * p.h;
  k.mass = p.mass;
  k.rho0 = p.rest_density;
  k.eps = p.relaxation;
  k.poly6 = poly6_coef(p.h);
  k.spiky = -45.0f / (kPi * powf(p.h, 6.0f));
  k.visc_lap = 45.0f / (kPi * powf(p.h, 6.0f));
  float rq = powf(0.2f * p.h, 2.0f);
  float wq = 0.0f;
  if (rq < k.h2) {
    float d = k.h2 - rq;
    wq = k.poly6 * d * d * d;
  }
  k.w_q = wq;
  k.scorr_k = p.scorr_k;
  k.visc = p.viscosity;
  k.dt = p.dt;
  k.max_scan = p.max_cell_scan > 0 ? p.max_cell_scan : 512;
  k.delta_scale = (p.flags & FX_FLAG_MASS_WEIGHTED_DELTA) ? p.mass : 1.0f;
  return k;
}

// ---- workspace -------------------------------------------------------------
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Slots of padding after each candidate array (sorted positions and
// velocities): the row-batch walk loads whole batches of consecutive slots.
constexpr int kCandPad = 8;

Workspace workspace_layout(const FxParams& p) {
  Workspace w;
  size_t n = (size_t)(p.n > 0 ? p.n : 0), c = (size_t)(p.n_cells > 0 ? p.n_cells : 0);
  size_t o = 0;
  w.cnt = o;      o += align256((c + 1) * 4);
  w.off = o;      o += align256((c + 1) * 4);
  w.rank = o;     o += align256(n * 4 + 4);
  w.key = o;      o += align256(n * 4 + 4);
  w.sidx = o;     o += align256(n * 4 + 4);
  w.skey = o;     o += align256(n * 4 + 4);
  // candidate arrays of the neighbour passes: kCandPad slots of padding
  w.pred_s0 = o;  o += align256(n * 16 + 16 * kCandPad);
  w.pred_s1 = o;  o += align256(n * 16 + 16 * kCandPad);
  w.pos_s = o;    o += align256(n * 16 + 16);
  w.vel_s = o;    o += align256(n * 16 + 16 * kCandPad);
  w.dl_s = o;     o += align256(n * 8 + 8);
  w.mask = o;     o += std::max(align256(n * 8 * 9 + 8), align256(n * 8) + align256(n * 16) + align256(n * 16 + 16));
  w.lcell = o;    o += align256(n * 4 + 4);
  w.lfull = o;    o += align256(n * 4 + 4);
  w.disp = o;     o += align256(n * 4 + 4);
  w.dcell = o;    o += align256(c * 4 + 4);
  w.partial = o;  o += align256(scan_partials((long long)c + 1) * 4 + 4);
  w.flags = o;    o += align256(64 * 4);
  w.normal_next = o; o += align256(n * 16 + 16);
  w.sdcell = o;   o += align256(c * 4 + 4);
  // x, y, z of the candidate positions, each with 16 floats of padding (a
  // recording batch reads up to 7 slots past a run)
  w.sx = o;       o += align256(n * 4 + 64);
  w.sy = o;       o += align256(n * 4 + 64);
  w.sz = o;       o += align256(n * 4 + 64);
  // 8 + 16 + 16 bytes per particle inside the 72 of the masks
  w.pair64 = w.mask;
  w.pos_r = w.pair64 + align256(n * 8);
  w.vel_r = w.pos_r + align256(n * 16);
  w.total = o;
  return w;
}

struct Ws {
  int32_t *cnt, *off, *rank, *key, *sidx, *skey, *partial, *flags;
  float4 *pred_s[2], *pos_s, *vel_s;
  float4* normal_next;  // the next substep's normals, from the XSPH pass
  uint64_t* pair64;     // (particle << 32 | previous slot), sorted order
  float4 *pos_r, *vel_r;  // final state by slot, between substeps of a call
  float2* dl_s;
  uint64_t* mask;
  int32_t *lcell, *lfull;
  float* disp;
  int32_t* dcell;
  float *sx, *sy, *sz;
};

static Ws bind_ws(const Workspace& w, void* base) {
  char* b = (char*)base;
  Ws s;
  s.cnt = (int32_t*)(b + w.cnt);
  s.off = (int32_t*)(b + w.off);
  s.rank = (int32_t*)(b + w.rank);
  s.key = (int32_t*)(b + w.key);
  s.sidx = (int32_t*)(b + w.sidx);
  s.skey = (int32_t*)(b + w.skey);
  s.partial = (int32_t*)(b + w.partial);
  s.flags = (int32_t*)(b + w.flags);
  s.pred_s[0] = (float4*)(b + w.pred_s0);
  s.pred_s[1] = (float4*)(b + w.pred_s1);
  s.pos_s = (float4*)(b + w.pos_s);
  s.vel_s = (float4*)(b + w.vel_s);
  s.normal_next = (float4*)(b + w.normal_next);
  s.pair64 = (uint64_t*)(b + w.pair64);
  s.pos_r = (float4*)(b + w.pos_r);
  s.vel_r = (float4*)(b + w.vel_r);
  s.dl_s = (float2*)(b + w.dl_s);
  s.mask = (uint64_t*)(b + w.mask);
  s.lcell = (int32_t*)(b + w.lcell);
  s.lfull = (int32_t*)(b + w.lfull);
  s.disp = (float*)(b + w.disp);
  s.dcell = (int32_t*)(b + w.dcell);
  s.sx = (float*)(b + w.sx);
  s.sy = (float*)(b + w.sy);
  s.sz = (float*)(b + w.sz);
  return s;
}

// ---- streaming device functions (exact; contraction off at file scope) ----

// shaders/sph_predict.glsl:23-57
__device__ __forceinline__ void predict_one(const FxParams& p, float4 x, float4 v, float4 nrm,
                                            float4& vo, float4& po) {
  float dt = p.dt, sigma = p.surface_tension;
  float vx = v.x, vy = v.y, vz = v.z;
  vz += p.gravity * dt;
  if (sigma > 0.0f) {
    float mag = sqrtf(nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z);
    if (mag > 1e-4f) {
      float s = -sigma * nrm.w;
      vx += s * nrm.x / mag / p.mass * dt;
      vy += s * nrm.y / mag / p.mass * dt;
      vz += s * nrm.z / mag / p.mass * dt;
    }
  }
  float v_max = 0.4f * p.h / fmaxf(dt, 1e-6f);
  float speed = sqrtf(vx * vx + vy * vy + vz * vz);
  if (speed > v_max) {
    float f = v_max / speed;
    vx *= f;
    vy *= f;
    vz *= f;
  }
  po = make_float4(x.x + vx * dt, x.y + vy * dt, x.z + vz * dt, 1.0f);
  vo = make_float4(vx, vy, vz, 0.0f);
}

// shaders/sph_finalize.glsl:15-96 (v already includes the XSPH delta)
__device__ __forceinline__ void finalize_one(const FxParams& p, const Collider& col, float& px,
                                             float& py, float& pz, float& vx, float& vy,
                                             float& vz) {
  const float radius = p.particle_radius, damping = p.boundary_damping;
  if (col.occ) {
    int cx, cy, cz;
    collider_coord(col, px, py, pz, cx, cy, cz);
    if (collider_occupied(col, cx, cy, cz)) {
      const float voxel = col.voxel;
      float nfx = 0.0f, nfy = 0.0f, nfz = 0.0f, best = 1e30f;
      bool found = false;
      // The 5^3 neighbourhood in the reference's order (dz, dy, dx from -2 to
      // 2; the first strict minimum wins), one x-row at a time: the row's 5
      // occupancy bytes are loaded (at clamped coordinates) before any is
      // tested, so a particle inside a collider waits for 25 loads, not 125.
#pragma unroll 1
      for (int rz = 0; rz < 25; ++rz) {
        const int z = cz + rz / 5 - 2, y = cy + rz % 5 - 2;
        const size_t row = ((size_t)min(max(z, 0), col.kz - 1) * col.ky + min(max(y, 0), col.ky - 1)) * col.kx;
        uint8_t o[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) o[i] = col.occ[row + min(max(cx + i - 2, 0), col.kx - 1)];
        if (z < 0 || y < 0 || z >= col.kz || y >= col.ky) continue;
#pragma unroll
        for (int i = 0; i < 5; ++i) {
          const int x = cx + i - 2;
          if (x < 0 || x >= col.kx || o[i] != 0) continue;
          float ctx = col.lox + ((float)x + 0.5f) * voxel;
          float cty = col.loy + ((float)y + 0.5f) * voxel;
          float ctz = col.loz + ((float)z + 0.5f) * voxel;
          float ex = ctx - px, ey = cty - py, ez = ctz - pz;
          float d2 = ex * ex + ey * ey + ez * ez;
          if (d2 < best) {
            best = d2;
            nfx = ctx;
            nfy = cty;
            nfz = ctz;
            found = true;
          }
        }
      }
      if (found) {
        float dist = sqrtf(best);
        float ux, uy, uz;
        if (dist > 1e-6f) {
          ux = (nfx - px) / dist;
          uy = (nfy - py) / dist;
          uz = (nfz - pz) / dist;
        } else {
          ux = 0.0f;
          uy = 0.0f;
          uz = 1.0f;
        }
        float push = dist + radius;
        px += ux * push;
        py += uy * push;
        pz += uz * push;
        float vn = vx * ux + vy * uy + vz * uz;
        if (vn < 0.0f) {
          float f = vn * (1.0f + damping);
          vx -= f * ux;
          vy -= f * uy;
          vz -= f * uz;
        }
      }
    }
  }
  float lox = p.lo[0] + radius, loy = p.lo[1] + radius, loz = p.lo[2] + radius;
  float hix = p.hi[0] - radius, hiy = p.hi[1] - radius, hiz = p.hi[2] - radius;
  if (px < lox) {
    px = lox;
    vx = -vx * damping;
  } else if (px > hix) {
    px = hix;
    vx = -vx * damping;
  }
  if (py < loy) {
    py = loy;
    vy = -vy * damping;
  } else if (py > hiy) {
    py = hiy;
    vy = -vy * damping;
  }
  if (pz < loz) {
    pz = loz;
    vz = -vz * damping;
  } else if (pz > hiz) {
    pz = hiz;
    vz = -vz * damping;
  }
}

// ---- grid build kernels ---------------------------------------------------

__global__ void k_fill_i32(int32_t* __restrict__ a, long long n, int32_t v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Fused sph_predict + sph_grid_key + counting-sort histogram.  mode bit 0:
// run predict (else pred is read as is); bit 1: count into cnt (rank[i] =
// arrival order within the cell, canonicalised later).
__global__ __launch_bounds__(256) void k_predict_key(FxParams p, Grid g, const float4* __restrict__ pos,
                                                     float4* __restrict__ vel,
                                                     const float4* __restrict__ nrm,
                                                     float4* __restrict__ pred,
                                                     int32_t* __restrict__ key,
                                                     int32_t* __restrict__ cnt,
                                                     int32_t* __restrict__ rank, int mode) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < p.n;
  int k = 0;
  if (active) {
    float4 pp;
    if (mode & 1) {
      float4 vo;
      const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
      predict_one(p, pos[i], vel[i], (p.surface_tension > 0.0f && nrm) ? nrm[i] : zero, vo, pp);
      vel[i] = vo;
      pred[i] = pp;
    } else {
      pp = pred[i];
    }
    if (mode & 6) {
      int x, y, z;
      cell_coord(g, pp.x, pp.y, pp.z, x, y, z);
      k = cell_index(g, x, y, z);
      key[i] = k;
    }
  }
  // canonical (lattice) order puts about two particles of a cell in a wave,
  // so one global atomic per particle contends little here (with the block
  // aggregation of block_rank this launch took 100 us instead of 67)
  if ((mode & 2) && active) rank[i] = atomicAdd(&cnt[k], 1);
}

// Counting-sort histogram of given keys (the FX_PASS_SORT entry).
// Keys outside [0, nc) are clamped into range (and flagged) so that a bad
// input cannot index out of bounds.
__global__ void k_count_keys(int n, int nc, int32_t* __restrict__ key, int32_t* __restrict__ cnt,
                             int32_t* __restrict__ rank, int32_t* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k = key[i];
  if (k < 0 || k >= nc) {
    atomicOr(bad, 1);
    k = min(max(k, 0), nc - 1);
    key[i] = k;
  }
  rank[i] = atomicAdd(&cnt[k], 1);
}

// Cell ranges in the reference format from the offsets (-1 for empty cells).
__global__ void k_cell_ranges(int nc, const int32_t* __restrict__ off, int32_t* __restrict__ start,
                              int32_t* __restrict__ end) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int s = off[c], e = off[c + 1];
  start[c] = e > s ? s : -1;
  end[c] = e > s ? e : -1;
}

// Counting-sort scatter.  The counts are dead once scanned: the same launch
// (over max(n, n_cnt) threads) zeroes them for the next counting pass.
__global__ void k_scatter(int n, const int32_t* __restrict__ key, const int32_t* __restrict__ rank,
                          const int32_t* __restrict__ off, int32_t* __restrict__ sidx, int32_t* __restrict__ cnt,
                          int n_cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_cnt) cnt[i] = 0;
  if (i >= n) return;
  const int k = key[i];
  sidx[off[k] + rank[i]] = i;
}

// ---- resident grid build (substeps after the first of one fx_step call) ----
// Between two substeps of one call the state stays in sorted order.  The
// finalize of substep k (k_finalize_predict) also runs substep k+1's
// predict: it has the final position and velocity in registers and, with
// cohesion, the next normals from its own XSPH pass (nrm_r, by slot).  It
// writes the final position (pos_r: the next substep's start position),
// the prediction, the cell key and the rank by slot; the predicted velocity
// is not kept (the velocity pass recomputes it from the positions,
// shaders/sph_velocity.glsl:12-24).  nrm_canon (when substep k+1 is the
// call's last) receives the canonical normals the caller's normal buffer
// holds after that substep.  Bit for bit what a separate finalize and
// predict would compute: the same float values pass through registers
// instead of memory.
__global__ __launch_bounds__(256) void k_finalize_predict(int n, FxParams p, Collider col, Grid g,
                                                          const float4* __restrict__ pred_s,
                                                          const float4* __restrict__ vel_s,
                                                          const float4* __restrict__ dv_s,
                                                          const float4* __restrict__ nrm_r,
                                                          const int32_t* __restrict__ sidx,
                                                          float4* __restrict__ pos_r, float4* __restrict__ pred_r,
                                                          int32_t* __restrict__ key, int32_t* __restrict__ cnt,
                                                          int32_t* __restrict__ rank,
                                                          float4* __restrict__ nrm_canon,
                                                          int32_t* __restrict__ big_count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = t < n;
  if (t == 0) *big_count = 0;  // the next build's big-cell queue (k_scatter_pairs)
  int k = 0;
  if (active) {
    const float4 pi = pred_s[t], vi = vel_s[t], dv = dv_s[t];
    float vx = vi.x + dv.x, vy = vi.y + dv.y, vz = vi.z + dv.z;
    float px = pi.x, py = pi.y, pz = pi.z;
    finalize_one(p, col, px, py, pz, vx, vy, vz);
    const float4 pf = make_float4(px, py, pz, 1.0f), vf = make_float4(vx, vy, vz, 0.0f);
    pos_r[t] = pf;
    const float4 nrm = (p.surface_tension > 0.0f && nrm_r) ? nrm_r[t] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 vo, pp;
    predict_one(p, pf, vf, nrm, vo, pp);
    pred_r[t] = pp;
    int x, y, z;
    cell_coord(g, pp.x, pp.y, pp.z, x, y, z);
    k = cell_index(g, x, y, z);
    key[t] = k;
    if (nrm_canon) nrm_canon[min(max(sidx[t], 0), n - 1)] = nrm;
  }
  const int r = block_rank<kBlock>(k, active, cnt);
  if (active) rank[t] = r;
}

// Counting-sort scatter of (particle << 32 | previous slot) pairs.  The
// counts are dead once scanned, so the same launch (over max(n, n_cnt)
// threads) zeroes them for the next substep's counting, clears the big-cell
// queue's count (flags[0]) for the canonicalisation that follows and marks
// the neighbour lists invalid (flags[2] = kInfBits): three fills fewer.
constexpr int32_t kInfBits = 0x7f800000;  // +inf: neighbour lists invalid
// cell runs longer than this are sorted by a workgroup (k_sort_big_cells);
// shorter ones in registers or, in a resident build, ranked by the gather
constexpr int kSmallCell = 32;

__global__ void k_scatter_pairs(int n, const int32_t* __restrict__ key, const int32_t* __restrict__ rank,
                                const int32_t* __restrict__ off, const int32_t* __restrict__ pid,
                                uint64_t* __restrict__ pairs, int32_t* __restrict__ cnt, int n_cnt,
                                int32_t* __restrict__ flags, int32_t* __restrict__ big_list) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_cnt) cnt[t] = 0;
  if (t == 0) {
    flags[2] = kInfBits;
    // the workspace header: this grid, and no normals by particle (the
    // resident substeps keep theirs by slot)
    flags[kHdrMagic] = kGridMagic;
    flags[kHdrN] = n;
    flags[kHdrCells] = n_cnt - 1;
    flags[kHdrNormals] = 0;
    flags[kHdrError] = 0;
  }
  if (t >= n) return;
  const int k = key[t], r = rank[t];
  pairs[off[k] + r] = ((uint64_t)(uint32_t)pid[t] << 32) | (uint32_t)t;
  // a cell's (kSmallCell + 1)-th member queues it for the block sort (the
  // count was zeroed by the k_finalize_predict before this build)
  if (r == kSmallCell) big_list[atomicAdd(&flags[0], 1)] = k;
}

// The closing mark of a grid build: the neighbour lists are invalid
// (flags[2]) and the workspace header names the new grid, forgets any
// normals computed over the previous one and clears the error word.
__global__ void k_mark_grid(int32_t* __restrict__ flags, int n, int nc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  flags[2] = kInfBits;
  flags[kHdrMagic] = kGridMagic;
  flags[kHdrN] = n;
  flags[kHdrCells] = nc;
  flags[kHdrNormals] = 0;
  flags[kHdrError] = 0;
}

// a[q] addressed as the array plus a 32-bit byte offset (one address add
// instead of a sign extension and a 64-bit add): 16-byte slots stay below
// 2^32 bytes with at most kMaxParticles particles (validated at the entry
// points).  Delta's walk over the pair masks gains from it; lambda's
// recording loads (301 -> 310 us) and XSPH's (518 -> 525) measured slower
// with it and keep the array indexing.
constexpr int kMaxParticles = (1 << 28) - 1;
template <class T>
__device__ __forceinline__ T slot_load(const T* __restrict__ a, int q) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(a) + ((uint32_t)q * (uint32_t)sizeof(T)));
}

// The candidate positions as three float arrays (x, y, z by slot), written
// beside pred_s by the passes that write it when the next lambda pass reads
// its candidates four slots to a load (PairMasks); null x: not written.
struct SoaPos {
  float *x, *y, *z;
};
__device__ __forceinline__ void soa_store(const SoaPos& a, int t, float x, float y, float z) {
  if (a.x) {
    a.x[t] = x;
    a.y[t] = y;
    a.z[t] = z;
  }
}

// The gather of a resident build: slot u of the new order reads its previous
// slot (mostly a nearby one: particles rarely change cells), and records the
// particle in sidx (plus the caller's reference-format copies when given).
// dl_s (zeroed: read before any lambda pass only with no iterations) and the
// sorted keys skey_int (read by the lists, the LDS-tiled walk and, after the
// call, the splat) are skipped when null.
__global__ void k_gather_pairs(int n, const uint64_t* __restrict__ pairs, const int32_t* __restrict__ off,
                               const int32_t* __restrict__ key_r,
                               const float4* __restrict__ pred_r, const float4* __restrict__ pos_r,
                               float4* __restrict__ pred_s, float4* __restrict__ pos_s,
                               float2* __restrict__ dl_s, int32_t* __restrict__ skey_out,
                               int32_t* __restrict__ skey_int, int32_t* __restrict__ sidx,
                               int32_t* __restrict__ sidx_out, int32_t* __restrict__ key_canon,
                               SoaPos soa = SoaPos{nullptr, nullptr, nullptr}) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint64_t e = pairs[v];
  const int pid = min(max((int)(e >> 32), 0), n - 1);
  const int src = min(max((int)(uint32_t)e, 0), n - 1);
  // the pair's canonical slot: a run of up to kSmallCell pairs sits in
  // scatter order and each pair ranks itself by particle index; a longer run
  // was sorted in place by k_sort_big_cells
  const int kc = key_r[src];
  const int c0 = off[kc], m = off[kc + 1] - c0;
  int u = v;
  if (m <= kSmallCell) {
    int r = 0;
#pragma unroll 4
    for (int j = 0; j < m; ++j) r += (uint32_t)(pairs[c0 + j] >> 32) < (uint32_t)pid ? 1 : 0;
    u = c0 + r;
  }
  float4 q = pred_r[src];
  q.w = 0.0f;
  pred_s[u] = q;
  soa_store(soa, u, q.x, q.y, q.z);
  pos_s[u] = pos_r[src];
  if (dl_s) dl_s[u] = make_float2(0.0f, 0.0f);
  sidx[u] = pid;
  if (sidx_out) sidx_out[u] = pid;
  if (skey_int) skey_int[u] = kc;
  if (skey_out) skey_out[u] = kc;
  if (key_canon) key_canon[pid] = kc;
}

// Canonicalise each cell's run to ascending particle index.  Runs of up to
// kSmallCell go through a bitonic network in registers, every load issued at
// once; longer runs are queued for k_sort_big_cells.  A run is sorted by its
// own lane, so the kernel lasts as long as its slowest lane: with a
// per-thread insertion sort (dependent loads, quadratic in the run) a few
// dense cells set the time - 73 us at 1.64M particles for 8-byte pairs, 43 us
// at 25,600 particles.
constexpr int kBigCellBlocks = 256;  // workgroups striding over the queued runs

template <int N, class T>
__device__ __forceinline__ void sort_run_in_registers(T* __restrict__ run, int m) {
  T a[N];
#pragma unroll
  for (int u = 0; u < N; ++u) a[u] = run[min(u, m - 1)];
#pragma unroll
  for (int u = 0; u < N; ++u)
    if (u >= m) a[u] = std::numeric_limits<T>::max();
#pragma unroll
  for (int kk = 2; kk <= N; kk <<= 1)
#pragma unroll
    for (int j = kk >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const bool asc = (i & kk) == 0;
          const T x = a[i], y = a[l];
          const bool sw = asc ? (x > y) : (x < y);
          a[i] = sw ? y : x;
          a[l] = sw ? x : y;
        }
      }
#pragma unroll
  for (int u = 0; u < N; ++u)
    if (u < m) run[u] = a[u];
}

template <class T>
__global__ void k_canon_cells(int nc, const int32_t* __restrict__ off, T* __restrict__ sidx,
                              int32_t* __restrict__ big_list, int32_t* __restrict__ big_count) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const int s = off[c], e = off[c + 1], m = e - s;
  if (m <= 1) return;
  if (m > kSmallCell) {
    big_list[atomicAdd(big_count, 1)] = c;
    return;
  }
  if (m <= 16)
    sort_run_in_registers<16>(sidx + s, m);
  else
    sort_run_in_registers<kSmallCell>(sidx + s, m);
}

// Block-cooperative sort of long runs: a bitonic network in which every
// compare-exchange puts the minimum at the lower slot (first step of each
// merge compares i with i ^ (k-1)), so slots past the run end act as +inf
// and are never touched.  Grid-stride over the queued cells.
template <class T>
__global__ __launch_bounds__(256) void k_sort_big_cells(const int32_t* __restrict__ off,
                                                        T* __restrict__ sidx,
                                                        const int32_t* __restrict__ big_list,
                                                        const int32_t* __restrict__ big_count) {
  const int count = *big_count;
  for (int b = blockIdx.x; b < count; b += gridDim.x) {
    const int c = big_list[b];
    const int s = off[c], m = off[c + 1] - s;
    int P = 1;
    while (P < m) P <<= 1;
    T* a = sidx + s;
    for (int k = 2; k <= P; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < P; i += blockDim.x) {
          const int partner = (j == (k >> 1)) ? (i ^ (k - 1)) : (i ^ j);
          if (partner > i && partner < m) {
            const T x = a[i], y = a[partner];
            if (x > y) {
              a[i] = y;
              a[partner] = x;
            }
          }
        }
        __syncthreads();
      }
    }
    __syncthreads();
  }
}

// Gather canonical state into sorted space.  pred_s.w carries the
// neighbour's density for the normal pass (lambda4.x, or 0 after a reset:
// the reference's zeroed lambda_img, solver/sph.py:716-724).
__global__ void k_gather(int n, const int32_t* __restrict__ sidx, const int32_t* __restrict__ key,
                         const float4* __restrict__ pred, const float4* __restrict__ pos,
                         const float4* __restrict__ lambda4, float4* __restrict__ pred_s,
                         float4* __restrict__ pos_s, float2* __restrict__ dl_s,
                         int32_t* __restrict__ skey_out, int32_t* __restrict__ skey_int,
                         int32_t* __restrict__ sidx_out, SoaPos soa = SoaPos{nullptr, nullptr, nullptr}) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = min(max(sidx[t], 0), n - 1);  // a caller-supplied permutation stays in bounds
  float4 q = pred[i];
  float2 dl = make_float2(0.0f, 0.0f);
  if (lambda4) {
    const float4 l = lambda4[i];
    dl = make_float2(l.x, l.y);
  }
  q.w = dl.x;
  pred_s[t] = q;
  soa_store(soa, t, q.x, q.y, q.z);
  if (pos_s) pos_s[t] = pos[i];
  if (dl_s) dl_s[t] = dl;
  if (key) {
    const int k = key[i];
    if (skey_int) skey_int[t] = k;
    if (skey_out) skey_out[t] = k;
  }
  if (sidx_out) sidx_out[t] = i;
}

// Offsets from sorted keys: off[c] = first slot with key >= c.
__global__ void k_off_from_sorted(int n, int nc, const int32_t* __restrict__ skey,
                                  int32_t* __restrict__ off) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > n) return;
  // clamped so that an unsorted or out-of-range key array cannot index
  // outside off[0..nc]
  const int prev = t == 0 ? -1 : min(max(skey[t - 1], 0), nc);
  const int cur = t == n ? nc : min(max(skey[t], 0), nc);
  for (int c = prev + 1; c <= cur; ++c) off[c] = t;
}

// The reference's own cell-range pass (shaders/sph_cell_range.glsl:10-32)
// over the n sorted slots: start where the key changes from the previous
// slot, end where it changes to the next.
__global__ void k_cell_range_ref(int n, int nc, const int32_t* __restrict__ skey,
                                 int32_t* __restrict__ start, int32_t* __restrict__ end) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int key = skey[i];
  if (key < 0 || key >= nc) return;
  const int prev = (i == 0) ? -1 : skey[i - 1];
  const int next = (i == n - 1) ? -1 : skey[i + 1];
  if (key != prev) start[key] = i;
  if (key != next) end[key] = i + 1;
}

// ---- batched scenes -----------------------------------------------------------
// Many independent scenes share one launch per pass: their particles are
// concatenated (scene s owns particles [poff, poff + n)), their cells are
// offset into one global key space (scene s owns keys [coff, coff + nc)), so
// one counting sort orders every scene at once and scene s keeps the sorted
// slots [poff, poff + n).  A thread finds its scene by binary search over the
// device table and runs the plain pass with that scene's grid, constants and
// cell offsets.
struct SceneDev {
  FxParams p;
  Consts k;
  Grid g;
  Collider col;
  int poff, n, coff, nc;
};

struct Scenes {
  const SceneDev* s;
  int count;
  int32_t* queue = nullptr;  // persistent launches: one work counter per XCD, zeroed before the launch
};

__device__ __forceinline__ int scene_of(const Scenes& S, int t) {
  int lo = 0, hi = S.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (S.s[mid].poff <= t)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// ---- neighbour passes (sorted space) -------------------------------------
//
// Neighbour lists.  The first lambda pass of a substep visits the full
// 27-cell stencil and records, per particle and per stencil row (dz, dy), a
// 64-bit mask of the row's slots within h + skin of the particle
// (build positions).  Later passes of the substep - and the next substep's
// normal pass, which reads the same grid - visit only the masked slots, in
// the same row order and ascending slot order, so every sum has the same
// terms in the same order as the full stencil: results are identical.  This
// holds while (a) the particle is still in the cell it was in at the build
// (the stencil is centred on its current cell, shaders/sph_lambda.glsl:37)
// and (b) for a row, the particle's own accumulated displacement plus the
// largest accumulated displacement of any particle in the row's cells stays
// below the skin (then a pair within h now was within h + skin at the
// build).  Each particle accumulates |dp| over the Jacobi iterations and
// raises its cell's bound with an atomic max; a particle that fails (a), or
// a row that fails (b), was longer than 64 slots or needed the MAX_CELL_SCAN
// cut, is visited in full instead.
struct Lists {
  uint64_t* mask;        // [9][n]
  int32_t* cell;         // [n] cell key at build
  int32_t* full;         // [n] bit r: row r is scanned in full
  float* disp;           // [n] accumulated |dp| since the build
  int32_t* dcell;        // [n_cells] float bits: max accumulated |dp| in the cell
  const int32_t* built;  // 0 while the masks belong to the current grid
  float r2_list;         // (h + skin)^2
  float dlimit;          // a row's masks hold while disp_i + dcell <= dlimit
};

enum { VISIT_FULL = 0, VISIT_BUILD = 1, VISIT_LIST = 2 };
// Mode bits for VISIT_FULL: how a lane walks its nine stencil rows (see visit).
constexpr int VISIT_STREAM = 4;        // one stream per lane; batches run on across row ends
constexpr int VISIT_MASKED_LOADS = 8;  // with VISIT_STREAM: empty batch slots load nothing
constexpr int VISIT_ROW_BATCH = 16;    // with VISIT_STREAM: batches stay in a row
constexpr int VISIT_PACKED = 32;       // with VISIT_ROW_BATCH: row lengths two to a register
// A small scene leaves most SIMDs with one wave or none, so a pass lasts as
// long as one wave's instruction stream (~330 cycles per candidate at 25,600
// particles): the stream walks take about the same time from 26k to 127k
// particles (lambda 54-58 us).  Small scenes run the quad walk (visit_quad:
// four lanes per particle), which costs more instructions in all but
// finishes sooner while SIMDs are idle.  Where it stops paying, per pass
// (h/2 lattice, ~216 candidates per particle; lambda / delta / XSPH with
// normals, us, stream vs quad): 26k 54 / 58 / 76 vs 28 / 32 / 65, 65k
// 56 / 58 / 79 vs 40 / 47 / 101, 90k 56 / 59 / 82 vs 51 / 60 / 141, 127k
// 58 / 60 / 84 vs 65 / 77 / 175.
constexpr int kQuadBelowLambda = 104000;
constexpr int kQuadBelowDelta = 88000;
constexpr int kQuadBelowXsph = 36000;
// Delta over lambda's recorded pairs (pair masks): in quads below this, one
// lane per particle above (h/2 lattices, cohesion on; delta us, quads / one
// lane: 25,600 particles 18.1 / 21.1, 40,000 21.6 / 21.3, 50,000 23.7 /
// 21.5, 64,980 25.6 / 21.4, 89,780 31.7 / 23.2).
constexpr int kQuadMaskBelowDelta = 40000;
// Kernel template bit: between its kQuadBelow* and kRowBelow, delta and
// XSPH walk row by row, whose bookkeeping is lighter than the stream walks'
// (25,600 particles: delta / XSPH 48 / 58 us row by row against 59 / 74 us
// streamed; 127k: 59 / 81 against 60 / 84).  Lambda's stream walk already
// batches within rows and stays ahead (127k: 58 us streamed, 64 by row).
constexpr int kRowWalk = 64;
// Kernel template bit: the large-scene passes' other stream walk (the
// autotuner's walk knob, FxTuning.walk = 2): lambda and delta stream
// without row batches, XSPH streams with unconditional loads.  Every walk
// visits the same candidates in the same order.
constexpr int kWalkAlt = 512;
constexpr int kRowBelow = 1 << 17;

struct PV {
  float4 p, v;
  float r2;
};

// ld(q) loads candidate slot q; f(q, data) is called for each candidate in
// the reference's order and returns whether the candidate is within the list
// radius (used by BUILD).
template <int MODE, class LD, class F, class RF = NoRefine>
__device__ __forceinline__ void visit(const Grid& g, const int32_t* __restrict__ off, int max_scan,
                                      int bx, int by, int bz, int t, int n, const Lists& L, LD&& ld,
                                      F&& f, RF&& rf = RF()) {
  const int x0 = max(bx - 1, 0), x1 = min(bx + 1, g.cx - 1);
  constexpr int M = MODE & 3;
  if ((MODE & VISIT_STREAM) && M == VISIT_FULL) {
    // The nine row runs are loaded up front into a queue S/E that shifts on
    // every row change (which keeps it in registers), and each lane walks
    // them without waiting for the rest of its wave at row ends.  Walked row
    // by row (below), a wave waits at every row end for the lane with the
    // longest run and its one-at-a-time tail: in a developed flow, where cell
    // occupancy varies, that is 368 loads per wave for the ~235 a lane needs.
    // Candidates reach f in the same order either way.
    //   default: batches of B run on across row ends; the empty slots
    //     of a lane's last batch re-load its first candidate, which keeps the
    //     batch's loads unconditional and back to back (with each load behind
    //     its own branch lambda takes 517 us instead of 343).
    //   VISIT_MASKED_LOADS: empty slots load nothing (XSPH, whose second
    //     gather already sits behind the range test: 525 vs 575 us).
    //   VISIT_ROW_BATCH: a batch stays inside one row, the row's last batch
    //     masked, one row check per batch instead of one per slot (lambda
    //     and delta).
    // Per pass, substeps 8-248 of the 1.64M scene (us; row by row / stream /
    // row batches): lambda 351 / 384 / 350, delta 354 / 347 / 360, XSPH
    // 602 / 525 (masked) / 571.  Once the row batches load from one address
    // with immediate offsets and take each batch's first slot unconditionally
    // (below), delta's row batches overtake its stream walk: 306 -> 299 us.
    // A lane with a row past max_scan takes the per-cell path below.
    // loads in flight per lane (substep at 1.64M particles, batch 3 / 4 / 5 /
    // 6 / 7 / 8: lambda 356 / 346 / 337 / 340 / 338 / 343 us, delta 352 /
    // 344 / 341 / 340 / 342 / 349, XSPH 548 / 530 / 543 / 533 / 575 / 563)
    constexpr int B = (MODE & VISIT_MASKED_LOADS) ? 4 : 6;
    int S[9], E[9];
    bool big = false;
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
      const bool in = z >= 0 && z < g.cz && y >= 0 && y < g.cy;
      const int c0 = in ? (z * g.cy + y) * g.cx + x0 : 0;
      S[r] = in ? off[c0] : 0;
      E[r] = in ? off[c0 + (x1 - x0) + 1] : 0;
      big |= E[r] - S[r] > max_scan || E[r] - S[r] > 0xFFFF;
    }
    if (!big && !(MODE & VISIT_ROW_BATCH)) {
      // the queue's row lengths held two to a register (each is at most
      // 0xFFFF): registers fewer, which keeps the XSPH pass at 7 waves per
      // SIMD instead of 6
      uint32_t Lp[5];
#pragma unroll
      for (int i = 0; i < 5; ++i)
        Lp[i] = (uint32_t)(E[2 * i] - S[2 * i]) | (2 * i + 1 < 9 ? (uint32_t)(E[2 * i + 1] - S[2 * i + 1]) << 16 : 0u);
      int k = S[0], e = S[0] + (int)(Lp[0] & 0xFFFFu), left = 8;
#pragma unroll 1
      for (;;) {
        int q[B];
        bool v[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
          while (k >= e && left > 0) {
#pragma unroll
            for (int r = 0; r < 8; ++r) S[r] = S[r + 1];
#pragma unroll
            for (int i = 0; i < 4; ++i) Lp[i] = (Lp[i] >> 16) | (Lp[i + 1] << 16);
            Lp[4] >>= 16;
            k = S[0];
            e = S[0] + (int)(Lp[0] & 0xFFFFu);
            --left;
          }
          v[u] = k < e;
          q[u] = k;
          k += v[u] ? 1 : 0;
        }
        if (!v[0]) break;
        decltype(ld(0)) d[B];
        if (MODE & VISIT_MASKED_LOADS) {
#pragma unroll
          for (int u = 0; u < B; ++u)
            if (v[u]) d[u] = ld(q[u]);
        } else {
#pragma unroll
          for (int u = 0; u < B; ++u) d[u] = ld(v[u] ? q[u] : q[0]);
        }
#pragma unroll
        for (int u = 0; u < B; ++u)
          if (v[u]) rf(q[u], d[u]);
#pragma unroll
        for (int u = 0; u < B; ++u)
          if (v[u]) f(q[u], d[u]);
      }
      return;
    }
    if (!big && (MODE & VISIT_ROW_BATCH)) {
      // VISIT_PACKED: the row lengths two to a register, as below (the
      // batched lambda pass: 65 -> 62 VGPRs, 8 waves per SIMD; the plain one
      // is at 8 already and only pays the shifts)
      constexpr bool P = (MODE & VISIT_PACKED) != 0;
      uint32_t Lp[5];
#pragma unroll
      for (int i = 0; i < 5; ++i)
        Lp[i] = P ? (uint32_t)(E[2 * i] - S[2 * i]) | (2 * i + 1 < 9 ? (uint32_t)(E[2 * i + 1] - S[2 * i + 1]) << 16 : 0u)
                  : 0u;
      int k = S[0], e = P ? S[0] + (int)(Lp[0] & 0xFFFFu) : E[0], left = 8;
      auto next_row = [&]() {
        while (k >= e && left > 0) {
#pragma unroll
          for (int r = 0; r < 8; ++r) {
            S[r] = S[r + 1];
            if (!P) E[r] = E[r + 1];
          }
          if (P) {
#pragma unroll
            for (int i = 0; i < 4; ++i) Lp[i] = (Lp[i] >> 16) | (Lp[i + 1] << 16);
            Lp[4] >>= 16;
          }
          k = S[0];
          e = P ? S[0] + (int)(Lp[0] & 0xFFFFu) : E[0];
          --left;
        }
      };
      next_row();
#pragma unroll 1
      while (k < e) {
        // a batch's slots are consecutive: one address, immediate offsets.
        // Its slots past the row's end load whatever follows (the candidate
        // arrays carry kCandPad slots of padding) and are not used (lambda
        // 333 -> 318 us at 1.64M against loads clamped to the row)
        static_assert(B <= kCandPad, "row batches read up to B - 1 slots past a run");
        decltype(ld(0)) d[B];
#pragma unroll
        for (int u = 0; u < B; ++u) d[u] = ld(k + u);
        // slot k is in the run (k < e inside the loop): its term is taken
        // unconditionally, which keeps its load first in issue order
        // instead of sunk behind the others (lambda 317 -> 313 us)
        const int rem = e - k;  // one subtraction per batch, a constant compare per slot
        rf(k, d[0]);
#pragma unroll
        for (int u = 1; u < B; ++u)
          if (u < rem) rf(k + u, d[u]);
        f(k, d[0]);
#pragma unroll
        for (int u = 1; u < B; ++u)
          if (u < rem) f(k + u, d[u]);
        k += B;
        next_row();
      }
      return;
    }
  }
  if (M == VISIT_FULL) {
    // row by row: mid-size scenes (kRowWalk), passes without a stream walk
    // (the normal pass), and the stream walks' lanes with a row past
    // max_scan, which keep FX_BATCH - this path's registers count against
    // the whole kernel (8 would take XSPH from 73 to 107 VGPRs)
    constexpr int kPlainBatch = (MODE & VISIT_STREAM) ? FX_BATCH : 8;
    // the next row's run bounds are loaded while the current row is scanned
    auto row = [&](int r, int& c0, int& s, int& e) {
      const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
      if (z < 0 || z >= g.cz || y < 0 || y >= g.cy) {
        c0 = -1;
        s = e = 0;
        return;
      }
      c0 = (z * g.cy + y) * g.cx + x0;
      s = off[c0];
      e = off[c0 + (x1 - x0) + 1];
    };
    int c0n, sn, en;
    row(0, c0n, sn, en);
#pragma unroll 1
    for (int r = 0; r < 9; ++r) {
      const int c0 = c0n, s = sn, e = en;
      if (r < 8) row(r + 1, c0n, sn, en);
      if (c0 < 0) continue;
      if (e - s <= max_scan) {
        scan_run<kPlainBatch>(s, e, ld, f, rf);
      } else {
        for (int c = c0; c <= c0 + (x1 - x0); ++c) {
          const int cs = off[c], ce = min(off[c + 1], cs + max_scan);
          scan_run<kPlainBatch>(cs, ce, ld, f, rf);
        }
      }
    }
    return;
  }
  int full = 0;
  float di = 0.0f;
  if (M == VISIT_LIST) {
    full = L.full[t];
    di = L.disp[t];
  }
#pragma unroll 1
  for (int r = 0; r < 9; ++r) {
    const int dz = r / 3 - 1, dy = r % 3 - 1;
    const int z = bz + dz, y = by + dy;
    if (z < 0 || z >= g.cz || y < 0 || y >= g.cy) {
      if (M == VISIT_BUILD) L.mask[(size_t)r * n + t] = 0;
      continue;
    }
    const int c0 = (z * g.cy + y) * g.cx + x0, c1 = c0 + (x1 - x0);
    const int s = off[c0], e = off[c1 + 1];
    bool use_mask = false;
    if (M == VISIT_LIST && !((full >> r) & 1)) {
      float dm = 0.0f;
      for (int c = c0; c <= c1; ++c) dm = fmaxf(dm, __int_as_float(L.dcell[c]));
      use_mask = di + dm <= L.dlimit;
    }
    if (e - s <= max_scan) {
      if (M == VISIT_LIST && use_mask) {
        uint64_t m = L.mask[(size_t)r * n + t];
        while (m) {
          const int b = __ffsll((unsigned long long)m) - 1;
          m &= m - 1;
          auto d = ld(s + b);
          rf(s + b, d);
          f(s + b, d);
        }
      } else if (M == VISIT_BUILD && e - s <= 64) {
        uint64_t m = 0;
        scan_run(
            s, e, ld, [&](int q, const auto& d) {
              if (f(q, d)) m |= 1ull << (q - s);
            },
            rf);
        L.mask[(size_t)r * n + t] = m;
      } else {
        scan_run(s, e, ld, f, rf);
        if (M == VISIT_BUILD) {
          full |= 1 << r;
          L.mask[(size_t)r * n + t] = 0;
        }
      }
    } else {
      for (int c = c0; c <= c1; ++c) {
        const int cs = off[c], ce = min(off[c + 1], cs + max_scan);
        scan_run(cs, ce, ld, f, rf);
      }
      if (M == VISIT_BUILD) {
        full |= 1 << r;
        L.mask[(size_t)r * n + t] = 0;
      }
    }
  }
  if (M == VISIT_BUILD) {
    L.full[t] = full;
    L.cell[t] = cell_index(g, bx, by, bz);
    L.disp[t] = 0.0f;
  }
}

// ---- pair masks: lambda records, delta walks --------------------------------
// The delta pass of a Jacobi iteration visits the candidates lambda visited,
// at the same positions, and keeps the same ones (r^2 < h^2; delta's r^2 >
// 1e-12 test is made again on the candidates themselves).  Lambda therefore
// records, per stencil row r, a 64-bit word whose bit j says that slot
// S[r] + j (S[r]: the row's first slot) lay within h; delta then loads only
// those ~34 of the ~216 candidates, in slot order, so its sums see the same
// terms in the same order and are bit-identical to a full walk.  A particle
// with a row of more than 64 slots (or past max_scan) is not recorded
// (whole[t] = 1) and delta walks its full stencil.  The words live in the
// neighbour-list mask area of the workspace ([9][n] words), which is unused
// while the lists are off; the resident hand-off state that shares it is dead
// between the gather and the substep's final finalize.  Lambda keeps a row's
// word in registers while it walks the row and parks it in LDS at the row's
// end (LDS stores are not counted with the loads in flight), then writes the
// nine words out once.  On the 1.64M scene: delta 302 -> 151 us, lambda 314
// -> 330 us (the recording's registers), substep 3.09 -> 2.56 ms.
struct PairMasks {
  uint64_t* rows;  // [9][n]
  int32_t* whole;  // [n]: 1 = not recorded, walk the full stencil
  SoaPos pos;      // lambda (one lane per particle) reads its candidates here;
                   // delta writes its new predictions here when not null
};
constexpr int kMaskMaxRow = 64;

// The row-batch walk of visit (VISIT_ROW_BATCH, all nine rows' bounds loaded
// up front, row lengths packed two to a register), recording each row's
// in-range bits: f returns whether the candidate lies within h.  rm: this
// lane's first row word in LDS, `stride` words between rows.  Returns false,
// having visited nothing, when a row is longer than kMaskMaxRow or max_scan;
// the caller then walks visit_cells_capped.  (Walking such rows here, or
// falling back to the stream walk, holds 9-13 more registers in the batched
// pass: 6 waves per SIMD instead of 7.)  Loading each row's bounds while the
// row before it is walked instead holds 8 registers fewer but measured slower
// (lambda 335 -> 347-357 us at 1.64M, 4-6 loads per batch; the batch's
// lambda 1215 -> 1192-1262 us).
template <int B, class LD, class F>
__device__ __forceinline__ bool visit_record(const Grid& g, const int32_t* __restrict__ off, int max_scan,
                                            int bx, int by, int bz, LD&& ld, F&& f, uint64_t* rm, int stride) {
  static_assert(B <= kCandPad, "row batches read up to B - 1 slots past a run");
  const int x0 = max(bx - 1, 0), x1 = min(bx + 1, g.cx - 1);
  int S[9];
  uint32_t Lp[5];
  bool big = false;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
    const bool in = z >= 0 && z < g.cz && y >= 0 && y < g.cy;
    const int c0 = in ? (z * g.cy + y) * g.cx + x0 : 0;
    S[r] = in ? off[c0] : 0;
    const int len = in ? off[c0 + (x1 - x0) + 1] - S[r] : 0;
    big |= len > max_scan || len > kMaskMaxRow;
    if (r % 2 == 0)
      Lp[r / 2] = (uint32_t)len;
    else
      Lp[r / 2] |= (uint32_t)len << 16;
  }
  if (big) return false;
  int k = S[0], e = S[0] + (int)(Lp[0] & 0xFFFFu), left = 8;  // the current row is 8 - left
  uint64_t acc = 0;
  auto next_row = [&]() {
    while (k >= e && left > 0) {
      rm[(8 - left) * stride] = acc;
      acc = 0;
#pragma unroll
      for (int r = 0; r < 8; ++r) S[r] = S[r + 1];
#pragma unroll
      for (int i = 0; i < 4; ++i) Lp[i] = (Lp[i] >> 16) | (Lp[i + 1] << 16);
      Lp[4] >>= 16;
      k = S[0];
      e = S[0] + (int)(Lp[0] & 0xFFFFu);
      --left;
    }
  };
  next_row();
#pragma unroll 1
  while (k < e) {
    decltype(ld(0)) d[B];
#pragma unroll
    for (int u = 0; u < B; ++u) d[u] = ld(k + u);
    const int rem = e - k;
    uint32_t bits = f(k, d[0]) ? 1u : 0u;
#pragma unroll
    for (int u = 1; u < B; ++u)
      if (u < rem && f(k + u, d[u])) bits |= 1u << u;
    acc |= (uint64_t)bits << (k - S[0]);
    k += B;
    next_row();
  }
  rm[(8 - left) * stride] = acc;
  return true;
}

// visit_record reading its candidates from the x / y / z arrays, four slots
// to a 16-byte load (the position's w is not needed by lambda): B candidates
// in 3B/4 load instructions instead of B.  On the 1.64M scene, lambda 331 ->
// 302 us with 8 candidates per batch (72 registers, 7 waves per SIMD); 4: 334
// us, 12: 332, 16: 341.  The candidates reach f in the same order as in
// visit_record.  In the batched pass (per-scene constants in registers) it
// holds 84 registers, 6 waves per SIMD, and measured slower (1186 -> 1236 us
// on the 64-scene batch), so the batch keeps visit_record.  Loading each
// row's bounds only while the row before is walked (after an up-front check
// of the nine lengths) holds 62 registers instead of 72 but took 301 -> 320
// us: the row changes then wait on those loads.  Recording alone (the
// distance test, no terms) and then walking the recorded words for the sums
// as delta does is bit-identical but holds 82 registers (5 waves per SIMD)
// and took 560 us: the scan, not the in-range terms, sets the pass's time.
// Held to 64 registers for 8 waves per SIMD (amdgpu_waves_per_eu(8)), four
// values spill to scratch: 320 us.  Branch-free forms of the per-slot tests
// (slots past the run given x = +inf, the own-particle gradient test as a
// select) took 314 / 313 us, both 320: the branches let whole waves skip.
// The row runs in LDS (a row change one LDS read, as delta's words) with
// each word stored to global memory at its row's end: 70 registers, still 7
// waves, 307 us.
template <int B, class F>
__device__ __forceinline__ bool visit_record_xyz(const Grid& g, const int32_t* __restrict__ off, int max_scan,
                                                 int bx, int by, int bz, const SoaPos& a, F&& f, uint64_t* rm,
                                                 int stride) {
  static_assert(B % 4 == 0 && B <= 16, "whole 16-byte loads within the arrays' padding");
  struct __attribute__((aligned(4))) F4 {
    float v[4];
  };
  const int x0 = max(bx - 1, 0), x1 = min(bx + 1, g.cx - 1);
  int S[9];
  uint32_t Lp[5];
  bool big = false;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
    const bool in = z >= 0 && z < g.cz && y >= 0 && y < g.cy;
    const int c0 = in ? (z * g.cy + y) * g.cx + x0 : 0;
    S[r] = in ? off[c0] : 0;
    const int len = in ? off[c0 + (x1 - x0) + 1] - S[r] : 0;
    big |= len > max_scan || len > kMaskMaxRow;
    if (r % 2 == 0)
      Lp[r / 2] = (uint32_t)len;
    else
      Lp[r / 2] |= (uint32_t)len << 16;
  }
  if (big) return false;
  int k = S[0], e = S[0] + (int)(Lp[0] & 0xFFFFu), left = 8;
  uint64_t acc = 0;
  auto next_row = [&]() {
    while (k >= e && left > 0) {
      rm[(8 - left) * stride] = acc;
      acc = 0;
#pragma unroll
      for (int r = 0; r < 8; ++r) S[r] = S[r + 1];
#pragma unroll
      for (int i = 0; i < 4; ++i) Lp[i] = (Lp[i] >> 16) | (Lp[i + 1] << 16);
      Lp[4] >>= 16;
      k = S[0];
      e = S[0] + (int)(Lp[0] & 0xFFFFu);
      --left;
    }
  };
  next_row();
#pragma unroll 1
  while (k < e) {
    F4 X[B / 4], Y[B / 4], Z[B / 4];
#pragma unroll
    for (int i = 0; i < B / 4; ++i) {
      X[i] = *reinterpret_cast<const F4*>(a.x + k + 4 * i);
      Y[i] = *reinterpret_cast<const F4*>(a.y + k + 4 * i);
      Z[i] = *reinterpret_cast<const F4*>(a.z + k + 4 * i);
    }
    const int rem = e - k;
    uint32_t bits = f(k, make_float4(X[0].v[0], Y[0].v[0], Z[0].v[0], 0.0f)) ? 1u : 0u;
#pragma unroll
    for (int u = 1; u < B; ++u)
      if (u < rem && f(k + u, make_float4(X[u / 4].v[u % 4], Y[u / 4].v[u % 4], Z[u / 4].v[u % 4], 0.0f)))
        bits |= 1u << u;
    acc |= (uint64_t)bits << (k - S[0]);
    k += B;
    next_row();
  }
  rm[(8 - left) * stride] = acc;
  return true;
}

// The 27 stencil cells one at a time, each capped at max_scan: the
// reference's order with few registers, for the rare particle with a row the
// pair masks do not record.
template <class LD, class F>
__device__ __forceinline__ void visit_cells_capped(const Grid& g, const int32_t* __restrict__ off, int max_scan,
                                                   int bx, int by, int bz, LD&& ld, F&& f) {
  const int x0 = max(bx - 1, 0), x1 = min(bx + 1, g.cx - 1);
#pragma unroll 1
  for (int r = 0; r < 9; ++r) {
    const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
    if (z < 0 || z >= g.cz || y < 0 || y >= g.cy) continue;
    const int c0 = (z * g.cy + y) * g.cx + x0;
#pragma unroll 1
    for (int c = c0; c <= c0 + (x1 - x0); ++c) {
      const int cs = off[c], ce = min(off[c + 1], cs + max_scan);
      scan_run<4>(cs, ce, ld, f);
    }
  }
}


// Delta's walk over the recorded words: B set bits of the current row per
// batch (the batch's loads issued together, empty slots re-loading its
// first), the candidates reaching f in slot order, each lane running on to
// its next row without waiting for its wave at row ends.  The nine words sit
// in LDS (rt: the lane's first, row r at rt[r * lstride]) and a row change
// reads the next one; the row starts shift down a register queue.  On the
// 1.64M scene (delta us), with the words in a register queue too: 134 with
// 4 bits per batch, 141 / 137 with 3 / 6; walked row by row, the wave
// waiting at each row's end for its lane with the most bits, 152 (3 / 6 / 8
// bits: 156 / 159 / 165), and with only the next row's word loaded ahead
// 155; a z-plane of three rows taken as one run, a batch's bits picked
// across the plane's rows by selects, 155-166; the 64-scene batch 565
// against 592 row by row.  The words in LDS instead of the queue (a row
// change one LDS read instead of shifting eighteen registers): 131 -> 120 us
// (VALU instructions per wave 3907 -> 3267, share of wave cycles waiting on
// a dependency 41% -> 26%), the batch 558 -> 536.
template <int B, class LD, class F>
__device__ __forceinline__ void visit_masked_lds(const Grid& g, const int32_t* __restrict__ off, int bx, int by,
                                                 int bz, int t, int n, const uint64_t* __restrict__ rows,
                                                 uint64_t* rt, int lstride, LD&& ld, F&& f) {
  const int x0 = max(bx - 1, 0);
  int S[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
    const bool in = z >= 0 && z < g.cz && y >= 0 && y < g.cy;
    S[r] = in ? off[(z * g.cy + y) * g.cx + x0] : 0;
    rt[r * lstride] = rows[(size_t)r * n + t];
  }
  int left = 8;
  uint64_t M = rt[0];
  auto next_row = [&]() {
    while (M == 0 && left > 0) {
#pragma unroll
      for (int r = 0; r < 8; ++r) S[r] = S[r + 1];
      --left;
      M = rt[(8 - left) * lstride];
    }
  };
  next_row();
#pragma unroll 1
  while (M) {
    int q[B];
    bool v[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      v[u] = M != 0;
      q[u] = S[0] + __ffsll((unsigned long long)M) - 1;
      M &= M - 1;
    }
    decltype(ld(0)) d[B];
#pragma unroll
    for (int u = 0; u < B; ++u) d[u] = ld(v[u] ? q[u] : q[0]);
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (v[u]) f(q[u], d[u]);
    next_row();
  }
}

// Run `body` over the candidates with the lists when they are valid for this
// particle, else over the full stencil.
template <int MODE, class LD, class F, class RF = NoRefine>
__device__ __forceinline__ void visit_auto(const Grid& g, const int32_t* __restrict__ off, int max_scan,
                                           int bx, int by, int bz, int t, int n, const Lists& L, LD&& ld,
                                           F&& f, RF&& rf = RF()) {
  if ((MODE & 3) == VISIT_LIST) {
    const bool ok = L.mask && *L.built == 0 && L.cell[t] == cell_index(g, bx, by, bz);
    if (ok)
      visit<VISIT_LIST>(g, off, max_scan, bx, by, bz, t, n, L, ld, f, rf);
    else
      visit<VISIT_FULL | (MODE & (VISIT_STREAM | VISIT_MASKED_LOADS | VISIT_ROW_BATCH | VISIT_PACKED))>(g, off, max_scan, bx, by, bz, t, n, L, ld, f, rf);
  } else {
    visit<MODE>(g, off, max_scan, bx, by, bz, t, n, L, ld, f, rf);
  }
}

// LDS-staged stencil rows (kernel template bit kLdsWalk; lambda and delta on
// large scenes).  The workgroup's kTileBlock consecutive sorted slots were
// binned into build cells kf..kl (skey of its first and last slot).  For each
// stencil row (dz, dy), with d its key offset dz*cy*cx + dy*cx, the cells
// kf + d - 1 .. kl + d + 1 hold one contiguous run of slots: the workgroup
// copies those nine runs (the window of row r) into LDS once, with coalesced
// loads, and each lane then reads its candidates from LDS.  A lane whose
// stencil cell is still a build cell of the block finds each of its row runs
// inside the window of the same row; one that moved to a neighbouring
// row of cells since the build (its cell minus one row offset lies in kf..kl)
// finds them in the windows shifted by that offset, where they exist.  Any
// run that is not inside a staged window - the lane moved further, the row
// falls outside the shifted windows, or the windows overflowed the tile - is
// read from global memory instead.  Candidates reach f in the same order
// whichever memory serves them.
// Measured on the 1.64M scene (FxTuning.walk = 3; lambda us, the stream walk
// 334): 425 with 256-particle workgroups and 2,496 tile slots (40 KB, four
// workgroups per CU), 505 / 592 / 703 with 3,200 / 1,920 / 1,152 slots for
// 256 / 128 / 64 particles (three, five, eight workgroups per CU).  The time
// follows the waves per CU (16, 12, 10, 8 against the stream walk's 32): the
// walk is bound by each wave's chain of dependent instructions, not by the
// memory pipeline it unloads (TA busy 84 -> 46%, TD 89 -> 68%).  Two more
// limits: a 256-slot workgroup's nine windows hold 2,483 slots at the median
// and 3,184 at p99, and by the last Jacobi iteration 30% of the particles sit
// in another (z, y) row of cells than their build cell (42% in another
// cell), so 229 of the stream walk's 291 loads per wave remain.  On h/2
// lattices of 204,800 / 819,200 / 1,638,400 particles lambda takes 79 / 223
// / 425 us against the stream walk's 68 / 172 / 334 and delta 70 / 203 /
// 389 against 69 / 168 / 324.  Kept as a walk the autotuner measures
// (bit-identical to the others); it has not won on a scene measured.
constexpr int kLdsWalk = 1024;
// Kernel template bit (pair masks, see PairMasks): lambda records its
// in-range candidates per stencil row, and the delta pass of the same
// iteration walks only those.
constexpr int kMaskRecord = 2048;
constexpr int kMaskWalk = 2048;
// Kernel template bit (with kMaskRecord): lambda reads its candidates from
// the x / y / z arrays (PairMasks::pos, visit_record_xyz).
constexpr int kXyzCand = 4096;
// Kernel template bit (delta over the pair masks): the nine row words sit in
// LDS (visit_masked_lds); launched with row_words_lds bytes of dynamic LDS.
constexpr int kLdsWords = 8192;
constexpr int kTileBlock = 256;
constexpr int kTileCap = 2496;  // float4 slots: four workgroups share a CU's 160 KB
constexpr int kTileGlobal = std::numeric_limits<int>::min();

struct Tile {
  float4 c[kTileCap + 8];  // a batch may read up to 7 slots past a window
  int a[9], len[9], base[9];
  int kf, kl;
};

// Fill T for the workgroup whose first slot is t0.  Every thread of the
// workgroup calls it (it synchronises).
__device__ __forceinline__ void stage_tile(Tile& T, const Grid& g, const int32_t* __restrict__ off,
                                           const int32_t* __restrict__ skey, const float4* __restrict__ src,
                                           int t0, int n) {
  const int tid = threadIdx.x;
  if (tid < 9) {
    const int t1 = min(t0 + kTileBlock, n) - 1;
    const int kf = skey[t0], kl = skey[t1];
    const long long nc = (long long)g.cx * g.cy * g.cz;
    const long long d = (long long)(tid / 3 - 1) * g.cy * g.cx + (long long)(tid % 3 - 1) * g.cx;
    const long long lo = max((long long)kf + d - 1, 0LL), hi = min((long long)kl + d + 1, nc - 1);
    int a = 0, len = 0;
    if (lo <= hi) {
      a = off[lo];
      len = max(off[hi + 1] - a, 0);
    }
    T.a[tid] = a;
    T.len[tid] = len;
    if (tid == 0) {
      T.kf = kf;
      T.kl = kl;
    }
  }
  __syncthreads();
  if (tid == 0) {
    int b = 0;
    for (int r = 0; r < 9; ++r) {
      const int l = min(T.len[r], kTileCap - b);
      T.base[r] = b;
      T.len[r] = l;
      b += l;
    }
  }
  __syncthreads();
  const int total = T.base[8] + T.len[8];
  int wa[9], wb[9];
#pragma unroll
  for (int w = 0; w < 9; ++w) {
    wa[w] = T.a[w];
    wb[w] = T.base[w];
  }
  constexpr int J = (kTileCap + kTileBlock - 1) / kTileBlock;
  float4 v[J];
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int e = tid + j * kTileBlock;
    int r = 0;
#pragma unroll
    for (int w = 1; w < 9; ++w)
      if (e >= wb[w]) r = w;  // the last window starting at or before e (empty ones are skipped)
    // past the staged total the load is clamped to a valid slot and not stored
    v[j] = src[min(max(wa[r] + (e - wb[r]), 0), n - 1)];
  }
#pragma unroll
  for (int j = 0; j < J; ++j) {
    const int e = tid + j * kTileBlock;
    if (e < total) T.c[e] = v[j];
  }
  __syncthreads();
}

// The full stencil of one particle through the staged tile, batches of B
// inside each row.  ld(q) loads slot q from global memory (same type as the
// tile's entries); f as for visit.
template <int B, class LD, class F>
__device__ __forceinline__ void visit_tile(const Grid& g, const int32_t* __restrict__ off, int max_scan, int bx,
                                           int by, int bz, const Tile& T, LD&& ld, F&& f) {
  const int x0 = max(bx - 1, 0), x1 = min(bx + 1, g.cx - 1);
  const int c = cell_index(g, bx, by, bz);
  const int kf = T.kf, kl = T.kl;
  // the row shift (sz, sy) with c - shift inside kf..kl; sz = 2: none
  int sz = 2, sy = 0;
  if (c >= kf && c <= kl) {
    sz = 0;
  } else {
    const int cxy = g.cx * g.cy;
#pragma unroll
    for (int s = 0; s < 9; ++s) {
      const int dz = s / 3 - 1, dy = s % 3 - 1;
      const int ck = c - dz * cxy - dy * g.cx;
      if (sz == 2 && ck >= kf && ck <= kl) {
        sz = dz;
        sy = dy;
      }
    }
  }
  int S[9], E[9], D[9];
  bool big = false;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
    const bool in = z >= 0 && z < g.cz && y >= 0 && y < g.cy;
    const int c0 = in ? (z * g.cy + y) * g.cx + x0 : 0;
    S[r] = in ? off[c0] : 0;
    E[r] = in ? off[c0 + (x1 - x0) + 1] : 0;
    big |= E[r] - S[r] > max_scan;
    D[r] = kTileGlobal;
    const int wz = r / 3 - 1 + sz, wy = r % 3 - 1 + sy;
    if (wz >= -1 && wz <= 1 && wy >= -1 && wy <= 1) {
      const int w = (wz + 1) * 3 + (wy + 1);
      const int a = T.a[w];
      if (S[r] >= a && E[r] <= a + T.len[w]) D[r] = T.base[w] - a;
    }
  }
  if (big) {
    // a row past max_scan: the per-cell cut of the row-by-row walk
    visit<VISIT_FULL>(g, off, max_scan, bx, by, bz, 0, 0, Lists{}, ld, f);
    return;
  }
  int k = S[0], e = E[0], dd = D[0], left = 8;
  auto next_row = [&]() {
    while (k >= e && left > 0) {
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        S[r] = S[r + 1];
        E[r] = E[r + 1];
        D[r] = D[r + 1];
      }
      k = S[0];
      e = E[0];
      dd = D[0];
      --left;
    }
  };
  next_row();
#pragma unroll 1
  while (k < e) {
    decltype(ld(0)) d[B];
    if (dd != kTileGlobal) {
      const float4* tp = T.c + (k + dd);
#pragma unroll
      for (int u = 0; u < B; ++u) d[u] = tp[u];
    } else {
#pragma unroll
      for (int u = 0; u < B; ++u) d[u] = ld(min(k + u, e - 1));
    }
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (k + u < e) f(k + u, d[u]);
    k += B;
    next_row();
  }
}

// Small scenes, full stencil, four lanes per particle (kernel template bit
// kQuadWalk).  In a small scene most SIMDs hold one wave or none, and a pass
// lasts as long as one wave's instruction stream; four lanes per particle
// quarter each stream and fill four times the SIMDs.  The quad walks
// its particle's rows exactly as visit's row-by-row path does (same runs, same
// per-cell cut past max_scan); in each round lane j takes slot k + j of the
// run, computes its term (a zero term past the run's end or out of range),
// and the quad folds the four terms in slot order through DPP broadcasts, each
// lane adding lane 0's, then lane 1's, ... term to its own copy of the sums.
// The terms and the fold are the same explicitly rounded operations the
// one-lane walks use, and folding a zero term (a zero summand, or a zero
// factor of a fused multiply-add) equals skipping the candidate: the sums
// start at +0 and round to nearest, so they are never -0, and x + (+-0) = x.
// So the four copies and the one-lane walks agree bit for bit.  tm(q, data,
// valid) returns the term of slot q; fold(term) adds one term.  All four
// lanes of a quad run the same control flow (same cell, same runs), which
// the DPP reads need.
constexpr int kQuadWalk = 128;

// Persistent launches (kernel template bit kPersistent; the batched passes,
// FxTuning.persistent): a grid of a few workgroups per CU, a multiple of the
// chip's 8 XCDs, works through the launch's virtual blocks.  The workgroups
// an XCD runs (blockIdx % 8) share that XCD's contiguous eighth of the
// virtual blocks, so each XCD's L2 serves one stretch of sorted slots, and
// take them in turn from the XCD's counter in S.queue (zeroed before the
// launch): a workgroup that drew cheap blocks draws again.  A workgroup
// leaves when its XCD's range is used up.
constexpr int kPersistent = 256;
__device__ __forceinline__ int persistent_take(int32_t* queue, long long nthreads) {
  __shared__ int s_vb;
  const int nb = (int)((nthreads + blockDim.x - 1) / blockDim.x);
  const int x = (int)blockIdx.x % kXcds, per = (nb + kXcds - 1) / kXcds;
  __syncthreads();  // the previous block's readers of s_vb are done
  if (threadIdx.x == 0) s_vb = atomicAdd(&queue[x], 1);
  __syncthreads();
  const int j = s_vb, vb = x * per + j;
  return (j < per && vb < nb) ? vb : -1;
}

// B: rounds of loads in flight per lane (dam break, 25,600 particles, B = 2 /
// 4 / 8: lambda 29.9 / 25.3 / 23.9 us, delta 30.5 / 28.7 / 28.2, XSPH 43.0 /
// 44.0 / 47.8; at 65k-90k particles B = 8 costs delta 2% and lambda up to 2%,
// its registers taking the quad kernels from 8 to 6 waves per SIMD): lambda
// takes 8, delta and XSPH 4
template <int B, class LD, class TM, class FD>
__device__ __forceinline__ void visit_quad(const Grid& g, const int32_t* __restrict__ off, int max_scan, int bx,
                                           int by, int bz, int j, LD&& ld, TM&& tm, FD&& fold) {
  const int x0 = max(bx - 1, 0), x1 = min(bx + 1, g.cx - 1);
  auto run = [&](int s, int e) { quad_run<B>(s, e, j, ld, tm, fold); };
  auto row = [&](int r, int& c0, int& s, int& e) {
    const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
    if (z < 0 || z >= g.cz || y < 0 || y >= g.cy) {
      c0 = -1;
      s = e = 0;
      return;
    }
    c0 = (z * g.cy + y) * g.cx + x0;
    s = off[c0];
    e = off[c0 + (x1 - x0) + 1];
  };
  int c0n, sn, en;
  row(0, c0n, sn, en);
#pragma unroll 1
  for (int r = 0; r < 9; ++r) {
    const int c0 = c0n, s = sn, e = en;
    if (r < 8) row(r + 1, c0n, sn, en);
    if (c0 < 0) continue;
    if (e - s <= max_scan) {
      run(s, e);
    } else {
      for (int c = c0; c <= c0 + (x1 - x0); ++c) {
        const int cs = off[c], ce = min(off[c + 1], cs + max_scan);
        run(cs, ce);
      }
    }
  }
}

// Pair masks for the quad walk (small scenes).  Recording: the quad's lanes
// walk each row as visit_quad does; lane j's in-range bits (its slots
// k + 4u + j) gather in its own word, the quad ORs the four words at the row's
// end (DPP broadcasts, every lane active), and lane 0 parks the row's word in
// LDS at rm[r * stride].  tm(q, data, valid, in) also sets `in` (within h).
// Returns whether every row was recorded (none past kMaskMaxRow or max_scan;
// such a row is walked cell by cell or whole without recording).
template <int B, class LD, class TM, class FD>
__device__ __forceinline__ bool visit_quad_record(const Grid& g, const int32_t* __restrict__ off, int max_scan,
                                                  int bx, int by, int bz, int j, LD&& ld, TM&& tm, FD&& fold,
                                                  uint64_t* rm, int stride) {
  const int x0 = max(bx - 1, 0), x1 = min(bx + 1, g.cx - 1);
  auto row = [&](int r, int& c0, int& s, int& e) {
    const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
    if (z < 0 || z >= g.cz || y < 0 || y >= g.cy) {
      c0 = -1;
      s = e = 0;
      return;
    }
    c0 = (z * g.cy + y) * g.cx + x0;
    s = off[c0];
    e = off[c0 + (x1 - x0) + 1];
  };
  auto plain_tm = [&](int q, const auto& d, bool valid) {
    bool in;
    return tm(q, d, valid, in);
  };
  bool whole = true;
  int c0n, sn, en;
  row(0, c0n, sn, en);
#pragma unroll 1
  for (int r = 0; r < 9; ++r) {
    const int c0 = c0n, s = sn, e = en;
    if (r < 8) row(r + 1, c0n, sn, en);
    uint64_t acc = 0;
    if (c0 >= 0) {
      if (e - s > max_scan) {
        whole = false;
        for (int c = c0; c <= c0 + (x1 - x0); ++c) {
          const int cs = off[c], ce = min(off[c + 1], cs + max_scan);
          quad_run<B>(cs, ce, j, ld, plain_tm, fold);
        }
      } else if (e - s > kMaskMaxRow) {
        whole = false;
        quad_run<B>(s, e, j, ld, plain_tm, fold);
      } else {
#pragma unroll 1
        for (int k = s; k < e; k += 4 * B) {
          decltype(ld(0)) d[B];
#pragma unroll
          for (int u = 0; u < B; ++u) d[u] = ld(min(k + 4 * u + j, e - 1));
          decltype(plain_tm(0, d[0], false)) tt[B];
#pragma unroll
          for (int u = 0; u < B; ++u) {
            const int q = k + 4 * u + j;
            bool in;
            tt[u] = tm(q, d[u], q < e, in);
            if (in) acc |= 1ull << (q - s);
          }
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
    }
    uint32_t lo = (uint32_t)acc, hi = (uint32_t)(acc >> 32);
    lo = __builtin_bit_cast(uint32_t, quad_lane<0>(__builtin_bit_cast(float, lo))) |
         __builtin_bit_cast(uint32_t, quad_lane<1>(__builtin_bit_cast(float, lo))) |
         __builtin_bit_cast(uint32_t, quad_lane<2>(__builtin_bit_cast(float, lo))) |
         __builtin_bit_cast(uint32_t, quad_lane<3>(__builtin_bit_cast(float, lo)));
    hi = __builtin_bit_cast(uint32_t, quad_lane<0>(__builtin_bit_cast(float, hi))) |
         __builtin_bit_cast(uint32_t, quad_lane<1>(__builtin_bit_cast(float, hi))) |
         __builtin_bit_cast(uint32_t, quad_lane<2>(__builtin_bit_cast(float, hi))) |
         __builtin_bit_cast(uint32_t, quad_lane<3>(__builtin_bit_cast(float, hi)));
    if (j == 0) rm[r * stride] = ((uint64_t)hi << 32) | lo;
  }
  return whole;
}

// Delta's quad walk over the recorded words: in each round the quad's lanes
// take the next four set bits of the current row (lane j the j-th), so the
// quad folds the terms in slot order as visit_quad does; B rounds of loads in
// flight, all from one row.  Each quad runs on to its next row without
// waiting for the rest of its wave (the nine words and first slots shift down
// a register queue).  A lane past the row's last bit loads the
// row's first slot and its term is the zero term (tm sees valid = false).
// Dam break (25,600 particles), delta us: 17.5 / 17.7 / 20.8 with 2 / 1 / 4
// rounds per batch; row by row with the wave in step (4 rounds) 20.3.
template <int B, class LD, class TM, class FD>
__device__ __forceinline__ void visit_quad_masked(const Grid& g, const int32_t* __restrict__ off, int bx, int by,
                                                  int bz, int j, int t, int n, const uint64_t* __restrict__ rows,
                                                  LD&& ld, TM&& tm, FD&& fold) {
  const int x0 = max(bx - 1, 0);
  int S[9];
  uint64_t M[9];
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
    const bool in = z >= 0 && z < g.cz && y >= 0 && y < g.cy;
    S[r] = in ? off[(z * g.cy + y) * g.cx + x0] : 0;
    M[r] = rows[(size_t)r * n + t];
  }
  int left = 8;
  auto next_row = [&]() {
    while (M[0] == 0 && left > 0) {  // the same for the four lanes
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        M[r] = M[r + 1];
        S[r] = S[r + 1];
      }
      M[8] = 0;
      --left;
    }
  };
  next_row();
#pragma unroll 1
  while (M[0]) {  // the same for the four lanes
    int q[B];
    bool v[B], any[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      uint64_t mm = M[0];
      if (j >= 1) mm &= mm - 1;
      if (j >= 2) mm &= mm - 1;
      if (j >= 3) mm &= mm - 1;
      any[u] = M[0] != 0;
      v[u] = mm != 0;
      q[u] = v[u] ? S[0] + __ffsll((unsigned long long)mm) - 1 : S[0];
#pragma unroll
      for (int i = 0; i < 4; ++i) M[0] &= M[0] - 1;
    }
    decltype(ld(0)) d[B];
#pragma unroll
    for (int u = 0; u < B; ++u) d[u] = ld(q[u]);
    decltype(tm(0, d[0], false)) tt[B];
#pragma unroll
    for (int u = 0; u < B; ++u) tt[u] = tm(q[u], d[u], v[u]);
#pragma unroll
    for (int u = 0; u < B; ++u) {
      if (any[u]) {
        fold(tt[u].template lane<0>());
        fold(tt[u].template lane<1>());
        fold(tt[u].template lane<2>());
        fold(tt[u].template lane<3>());
      }
    }
    next_row();
  }
}

// Lambda's per-candidate terms and their fold, shared by every walk: the
// density term m ((coef d) d) d (the reference's operations, so the density
// sum stays bit-exact), and the gradient c d, c = spiky (h - r)^2 / r without
// the mass (hoisted), with its squared norm.  A term that does not contribute
// has dens = 0 and c = 0 (its gradient and norm are then 0 too).
__device__ __forceinline__ float lam_dens(const Consts& k, float r2) {
#pragma clang fp contract(off)
  const float d = k.h2 - r2;
  return k.mass * (((k.poly6 * d) * d) * d);
}
__device__ __forceinline__ float lam_coef(const Consts& k, float r2) {
#pragma clang fp contract(off)
  const float inv_r = __builtin_amdgcn_rsqf(r2);
  const float w = __builtin_fmaf(-r2, inv_r, k.h);  // h - r
  return ((k.spiky * w) * w) * inv_r;
}
struct LamTerm {
  float dens, ax, ay, az, norm;
  template <int J>
  __device__ __forceinline__ LamTerm lane() const {
    return {quad_lane<J>(dens), quad_lane<J>(ax), quad_lane<J>(ay), quad_lane<J>(az), quad_lane<J>(norm)};
  }
};
// the gradient c d and its squared norm, given c (0 for no term)
__device__ __forceinline__ LamTerm lam_grad(float dens, float c, float dx, float dy, float dz) {
#pragma clang fp contract(off)
  const float ax = c * dx, ay = c * dy, az = c * dz;
  return {dens, ax, ay, az, __builtin_fmaf(az, az, __builtin_fmaf(ay, ay, ax * ax))};
}
struct LamSums {
  float dens = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f, gn2 = 0.0f;
  __device__ __forceinline__ void add_dens(float t) {
#pragma clang fp contract(off)
    dens = dens + t;
  }
  __device__ __forceinline__ void add_grad(const LamTerm& t) {
#pragma clang fp contract(off)
    gx = gx + t.ax;
    gy = gy + t.ay;
    gz = gz + t.az;
    gn2 = gn2 + t.norm;
  }
  __device__ __forceinline__ void add(const LamTerm& t) {
    add_dens(t.dens);
    add_grad(t);
  }
};

// Delta's per-candidate correction c d, c = (lambda_i + lambda_j + s_corr)
// spiky (h - r)^2 / r, and its fold (one fused multiply-add per component),
// shared by every walk.  A term that does not contribute has c = 0.
struct DelTerm {
  float c, dx, dy, dz;
  template <int J>
  __device__ __forceinline__ DelTerm lane() const {
    return {quad_lane<J>(c), quad_lane<J>(dx), quad_lane<J>(dy), quad_lane<J>(dz)};
  }
};
__device__ __forceinline__ float del_coef(const Consts& k, float li, float lj, float inv_wq, bool scorr, float r2) {
#pragma clang fp contract(off)
  const float inv_r = __builtin_amdgcn_rsqf(r2);
  float sc = 0.0f;
  if (scorr) {
    const float d = k.h2 - r2;
    const float ratio = (((k.poly6 * d) * d) * d) * inv_wq;
    const float r4 = ratio * ratio;
    sc = -k.scorr_k * (r4 * r4);
  }
  const float l = (li + lj) + sc;
  const float w = __builtin_fmaf(-r2, inv_r, k.h);  // h - r
  return (l * ((k.spiky * w) * w)) * inv_r;
}
struct DelSums {
  float ox = 0.0f, oy = 0.0f, oz = 0.0f;
  __device__ __forceinline__ void add(const DelTerm& t) {
    ox = __builtin_fmaf(t.dx, t.c, ox);
    oy = __builtin_fmaf(t.dy, t.c, oy);
    oz = __builtin_fmaf(t.dz, t.c, oz);
  }
};

// XSPH's per-candidate terms (the velocity difference weighted by
// ((d d) d) / rho_j, and with the normals the gradient weight (d d) / rho_j
// and the Laplacian weight (h - r) / rho_j) and their fold, one fused
// multiply-add per sum, shared by every walk.  A term that does not
// contribute is all zeros.
struct XsTerm {
  float w, dvx, dvy, dvz, gsc, dx, dy, dz, hr, irho;
  template <int J>
  __device__ __forceinline__ XsTerm lane() const {
    return {quad_lane<J>(w),   quad_lane<J>(dvx), quad_lane<J>(dvy), quad_lane<J>(dvz), quad_lane<J>(gsc),
            quad_lane<J>(dx),  quad_lane<J>(dy),  quad_lane<J>(dz),  quad_lane<J>(hr),  quad_lane<J>(irho)};
  }
};
template <bool NORMAL>
__device__ __forceinline__ XsTerm xs_term(const Consts& k, const float4& vi, const float4& vj, float dx, float dy,
                                          float dz, float r2) {
#pragma clang fp contract(off)
  XsTerm o;
  const float d = k.h2 - r2;
  o.irho = __builtin_amdgcn_rcpf(fmaxf(vj.w, 1e-6f));
  o.w = ((d * d) * d) * o.irho;
  o.dvx = vj.x - vi.x;
  o.dvy = vj.y - vi.y;
  o.dvz = vj.z - vi.z;
  o.gsc = o.dx = o.dy = o.dz = o.hr = 0.0f;
  if (NORMAL) {  // k_normal's pair term (pj.w is the same density)
    o.gsc = (d * d) * o.irho;
    o.dx = dx;
    o.dy = dy;
    o.dz = dz;
    // h - r; a distinct particle at the same position (r2 = +0) gives h, as
    // the reference's w_visc_lap(sqrt(r2)) does (r2 * rsq(r2) would be NaN)
    o.hr = r2 > 0.0f ? __builtin_fmaf(-r2, __builtin_amdgcn_rsqf(r2), k.h) : k.h;
  }
  return o;
}
template <bool NORMAL>
struct XsSums {
  float ax = 0.0f, ay = 0.0f, az = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f, lap = 0.0f;
  __device__ __forceinline__ void add(const XsTerm& t) {
    ax = __builtin_fmaf(t.dvx, t.w, ax);
    ay = __builtin_fmaf(t.dvy, t.w, ay);
    az = __builtin_fmaf(t.dvz, t.w, az);
    if (NORMAL) {
      nx = __builtin_fmaf(t.dx, t.gsc, nx);
      ny = __builtin_fmaf(t.dy, t.gsc, ny);
      nz = __builtin_fmaf(t.dz, t.gsc, nz);
      lap = __builtin_fmaf(t.hr, t.irho, lap);
    }
  }
};

// shaders/sph_lambda.glsl:17-77.  Writes dl_s[t] = (max(rho,1e-6), lambda),
// pred_s[t].w = lambda (read by the delta pass with the position), and the
// canonical lambda4 row if requested.  The density sum is bit exact
// (dist2_exact, then lam_dens terms added in the reference's order).
// V = MODE + 8 (batched scenes) + kRowWalk or kQuadWalk
template <int V>
__device__ __forceinline__ void lambda_body(int gt_in, int n, Grid g_, Consts k_, float4* __restrict__ pred_s,
                                            const int32_t* __restrict__ off_, float2* __restrict__ dl_s,
                                            const int32_t* __restrict__ sidx, float4* __restrict__ lambda4, Lists L,
                                            Scenes S, int iter, const Tile* tile = nullptr,
                                            const PairMasks& pm = PairMasks{nullptr, nullptr}, uint64_t* rm = nullptr) {
  constexpr int MODE = V & 3;
  constexpr bool QUAD = (V & kQuadWalk) != 0;
  constexpr bool REC = (V & kMaskRecord) != 0;
  const int gt = gt_in;
  const int t = QUAD ? gt >> 2 : gt;
  if (t >= n) return;
  Grid g = g_;
  Consts k = k_;
  const int32_t* off = off_;
  if (V & 8) {
    const SceneDev& sc = S.s[scene_of(S, t)];
    // batched: a scene past its own iteration count keeps its last lambda
    // pass (its densities feed XSPH; its delta pass reads no words)
    if (iter >= sc.p.iterations) return;
    g = sc.g;
    k = sc.k;
    off = off_ + sc.coff;
  }
  const float4 pi = pred_s[t];
  int bx, by, bz;
  cell_coord(g, pi.x, pi.y, pi.z, bx, by, bz);
  LamSums sum;
  {
    const float h2 = k.h2, r2l = L.r2_list;
    auto ld = [&](int q) { return pred_s[q]; };
    if constexpr (QUAD) {
      auto tm = [&](int q, const float4& pj, bool valid) {
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dist2_exact(dx, dy, dz);
        const bool in = valid && r2 < h2, grad = in && r2 > 1e-12f;
        return lam_grad(in ? lam_dens(k, r2) : 0.0f, grad ? lam_coef(k, r2) : 0.0f, dx, dy, dz);
      };
      if constexpr (REC) {
        auto tm_rec = [&](int q, const float4& pj, bool valid, bool& in_h) {
          const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
          const float r2 = dist2_exact(dx, dy, dz);
          const bool in = valid && r2 < h2, grad = in && r2 > 1e-12f;
          in_h = in;
          return lam_grad(in ? lam_dens(k, r2) : 0.0f, grad ? lam_coef(k, r2) : 0.0f, dx, dy, dz);
        };
        const int stride = (int)blockDim.x >> 2;
        const bool recorded = visit_quad_record<8>(g, off, k.max_scan, bx, by, bz, gt & 3, ld, tm_rec,
                                                   [&](const LamTerm& tt) { sum.add(tt); }, rm, stride);
        if ((gt & 3) == 0) {
          if (recorded) {
#pragma unroll
            for (int r = 0; r < 9; ++r) pm.rows[(size_t)r * n + t] = rm[r * stride];
          }
          pm.whole[t] = recorded ? 0 : 1;
        }
      } else {
        visit_quad<8>(g, off, k.max_scan, bx, by, bz, gt & 3, ld, tm, [&](const LamTerm& tt) { sum.add(tt); });
      }
      if (gt & 3) return;
    } else {
      auto body = [&](int q, const float4& pj) -> bool {
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dist2_exact(dx, dy, dz);
        if (r2 < h2) {
          sum.add_dens(lam_dens(k, r2));
          // the reference's own-particle test (j != i) is implied: slot t
          // against itself gives r2 = +0 exactly (delta 321 -> 307 us at
          // 1.64M without the extra compare)
          if (r2 > 1e-12f) sum.add_grad(lam_grad(0.0f, lam_coef(k, r2), dx, dy, dz));
        }
        return r2 < r2l;
      };
      if constexpr (REC) {
        // the same body; what it returns is the pair-mask bit (within h)
        auto body_rec = [&](int q, const float4& pj) -> bool {
          const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
          const float r2 = dist2_exact(dx, dy, dz);
          if (r2 < h2) {
            sum.add_dens(lam_dens(k, r2));
            if (r2 > 1e-12f) sum.add_grad(lam_grad(0.0f, lam_coef(k, r2), dx, dy, dz));
          }
          return r2 < h2;
        };
        // (4 loads per batch in the batched pass: 73 -> 69 registers, 7 waves per SIMD;
        // 6 in the plain pass: 331 -> 334 us)
        bool recorded;
        if constexpr ((V & kXyzCand) != 0)
          recorded = visit_record_xyz<8>(g, off, k.max_scan, bx, by, bz, pm.pos, body_rec, rm, (int)blockDim.x);
        else
          recorded = visit_record<(V & 8) ? 4 : 5>(g, off, k.max_scan, bx, by, bz, ld, body_rec, rm, (int)blockDim.x);
        if (recorded) {
#pragma unroll
          for (int r = 0; r < 9; ++r) pm.rows[(size_t)r * n + t] = rm[r * (int)blockDim.x];
        } else {
          visit_cells_capped(g, off, k.max_scan, bx, by, bz, ld, body_rec);
        }
        pm.whole[t] = recorded ? 0 : 1;
      } else if constexpr ((V & kLdsWalk) != 0)
        visit_tile<6>(g, off, k.max_scan, bx, by, bz, *tile, ld, body);
      else
        visit_auto<MODE | ((V & kRowWalk) ? 0 : (V & kWalkAlt) ? VISIT_STREAM : VISIT_STREAM | VISIT_ROW_BATCH | ((V & 8) ? VISIT_PACKED : 0))>(g, off, k.max_scan, bx, by, bz, t, n, L, ld, body);
    }
  }
  float dens = sum.dens, gx = sum.gx, gy = sum.gy, gz = sum.gz, gn2 = sum.gn2;
  // the mass factor of the gradient terms is hoisted out of the sums
  gx *= k.mass;
  gy *= k.mass;
  gz *= k.mass;
  gn2 *= k.mass * k.mass;
  const float constraint = fmaxf(dens, 1e-6f) / k.rho0 - 1.0f;
  const float denom = (gx * gx + gy * gy + gz * gz + gn2) / (k.rho0 * k.rho0) + k.eps;
  const float lambda = -constraint / fmaxf(denom, 1e-9f);
  const float rho = fmaxf(dens, 1e-6f);
  dl_s[t] = make_float2(rho, lambda);
  pred_s[t].w = lambda;
  if (lambda4) lambda4[min(max(sidx[t], 0), n - 1)] = make_float4(rho, lambda, 0.0f, 0.0f);
}

template <int V>
__global__ __launch_bounds__(256) void k_lambda(int n, Grid g_, Consts k_,
                                                float4* __restrict__ pred_s,
                                                const int32_t* __restrict__ off_,
                                                float2* __restrict__ dl_s,
                                                const int32_t* __restrict__ sidx,
                                                float4* __restrict__ lambda4, Lists L, Scenes S,
                                                int iter = -1, const int32_t* __restrict__ tkey = nullptr,
                                                PairMasks pm = PairMasks{nullptr, nullptr}) {
  if constexpr ((V & kMaskRecord) != 0) {
    // each lane's (the quad walk: each quad's) nine row words in LDS, row r
    // at rm[r * blockDim.x] (rm[r * blockDim.x / 4]); launched with
    // row_words_lds bytes of dynamic LDS
    extern __shared__ uint64_t fx_row_words[];
    uint64_t* rm = fx_row_words + (((V & kQuadWalk) != 0) ? threadIdx.x >> 2 : threadIdx.x);
    if constexpr ((V & kPersistent) != 0) {
      constexpr bool QUADW = (V & kQuadWalk) != 0;
      const long long nth = (QUADW ? 4LL * n : (long long)n);
      for (int vb = persistent_take(S.queue, nth); vb >= 0; vb = persistent_take(S.queue, nth))
        lambda_body<V>(vb * (int)blockDim.x + (int)threadIdx.x, n, g_, k_, pred_s, off_, dl_s, sidx, lambda4, L, S,
                       iter, nullptr, pm, rm);
    } else {
      lambda_body<V>(blockIdx.x * blockDim.x + threadIdx.x, n, g_, k_, pred_s, off_, dl_s, sidx, lambda4, L, S, iter,
                     nullptr, pm, rm);
    }
  } else if constexpr ((V & kPersistent) != 0) {
    constexpr bool QUADW = (V & kQuadWalk) != 0;
    const long long nth = (QUADW ? 4LL * n : (long long)n);
    for (int vb = persistent_take(S.queue, nth); vb >= 0; vb = persistent_take(S.queue, nth))
      lambda_body<V>(vb * (int)blockDim.x + (int)threadIdx.x, n, g_, k_, pred_s, off_, dl_s, sidx, lambda4, L, S, iter);
  } else if constexpr ((V & kLdsWalk) != 0) {
    // launched with kTileBlock threads; tkey: the sorted build keys
    __shared__ Tile tile;
    const int t0 = xcd_block() * kTileBlock;
    stage_tile(tile, g_, off_, tkey, pred_s, t0, n);
    lambda_body<V>(t0 + (int)threadIdx.x, n, g_, k_, pred_s, off_, dl_s, sidx, lambda4, L, S, iter, &tile);
  } else {
    lambda_body<V>(blockIdx.x * blockDim.x + threadIdx.x, n, g_, k_, pred_s, off_, dl_s, sidx, lambda4, L, S, iter);
  }
}

// shaders/sph_delta.glsl:22-85 fused with shaders/sph_apply_delta.glsl:10-20
// (Jacobi ping-pong: reads pred_in, writes pred_out) and, on the last
// iteration, shaders/sph_velocity.glsl:12-24 into vel_s (w = own density).
// delta4 (canonical, optional) receives the correction itself.  With lists,
// each particle's accumulated |dp| feeds the global displacement bound.
template <int V>
__device__ __forceinline__ void delta_body(int gt_in, int n, Grid g_, Consts k_, const float4* __restrict__ pred_in,
                                           const float2* __restrict__ dl_s, const int32_t* __restrict__ off_,
                                           float4* __restrict__ pred_out, const float4* __restrict__ pos_s,
                                           float4* __restrict__ vel_s, const int32_t* __restrict__ sidx,
                                           float4* __restrict__ delta4, Lists L, const int32_t* __restrict__ skey,
                                           Scenes S, int iter, const Tile* tile = nullptr,
                                           const PairMasks& pm = PairMasks{nullptr, nullptr}, uint64_t* rt = nullptr) {
  constexpr int MODE = V & 3;
  constexpr bool QUAD = (V & kQuadWalk) != 0;
  const int gt = gt_in;
  const int t = QUAD ? gt >> 2 : gt;
  if (t < n) {
    Grid g = g_;
    Consts k = k_;
    const int32_t* off = off_;
    if (V & 8) {
      const SceneDev& sc = S.s[scene_of(S, t)];
      const int its = sc.p.iterations;
      if (iter >= its) {
        // batched, the scene past its own iteration count: carry its
        // prediction through the ping-pong unchanged (w = its last density);
        // with no iterations at all, the velocity pass as k_velocity_s does
        const float4 pi = pred_in[t];
        const float2 dl = dl_s[t];
        pred_out[t] = make_float4(pi.x, pi.y, pi.z, dl.x);
        if (its == 0 && iter == 0) {
          const float dtc = fmaxf(sc.k.dt, 1e-6f);
          const float4 po = pos_s[t];
          vel_s[t] = make_float4((pi.x - po.x) / dtc, (pi.y - po.y) / dtc, (pi.z - po.z) / dtc, dl.x);
        }
        return;
      }
      if (iter != its - 1) vel_s = nullptr;  // its last iteration writes the velocity
      g = sc.g;
      k = sc.k;
      off = off_ + sc.coff;
    }
    const float4 pi = pred_in[t];
    const float li = pi.w;
    int bx, by, bz;
    cell_coord(g, pi.x, pi.y, pi.z, bx, by, bz);
    DelSums sum;
    {
      const float h2 = k.h2;
      const bool scorr = k.scorr_k > 0.0f && k.w_q > 1e-9f;
      const float inv_wq = scorr ? 1.0f / k.w_q : 0.0f;
      auto ld = [&](int q) { return pred_in[q]; };
      if constexpr (QUAD) {
        auto tm = [&](int q, const float4& pj, bool valid) {
          const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
          const float r2 = dist2_exact(dx, dy, dz);
          const bool in = valid && r2 < h2 && r2 > 1e-12f;  // q == t gives r2 = +0
          return DelTerm{in ? del_coef(k, li, pj.w, inv_wq, scorr, r2) : 0.0f, dx, dy, dz};
        };
        if constexpr ((V & kMaskWalk) != 0) {
          if (pm.whole[t] == 0)
            visit_quad_masked<2>(g, off, bx, by, bz, gt & 3, t, n, pm.rows, ld, tm,
                                 [&](const DelTerm& tt) { sum.add(tt); });
          else
            visit_quad<4>(g, off, k.max_scan, bx, by, bz, gt & 3, ld, tm, [&](const DelTerm& tt) { sum.add(tt); });
        } else {
          visit_quad<4>(g, off, k.max_scan, bx, by, bz, gt & 3, ld, tm, [&](const DelTerm& tt) { sum.add(tt); });
        }
        if (gt & 3) return;
      } else {
        auto body = [&](int q, const float4& pj) -> bool {
          const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
          const float r2 = dist2_exact(dx, dy, dz);
          // q == t gives r2 = +0: the own-particle test is implied
          if (r2 < h2 && r2 > 1e-12f) sum.add(DelTerm{del_coef(k, li, pj.w, inv_wq, scorr, r2), dx, dy, dz});
          return false;
        };
        if constexpr ((V & kMaskWalk) != 0) {
          // a recorded pair lies within h (its bit is this r2 < h2, from the
          // same positions), and a load's address is the array plus a 32-bit
          // byte offset: two compares and one address add fewer per pair
          // (delta 134.5 -> 132.2 us at 1.64M)
          auto body_m = [&](int q, const float4& pj) -> bool {
            const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const float r2 = dist2_exact(dx, dy, dz);
            if (r2 > 1e-12f) sum.add(DelTerm{del_coef(k, li, pj.w, inv_wq, scorr, r2), dx, dy, dz});
            return false;
          };
          auto ld32 = [&](int q) { return slot_load(pred_in, q); };
          static_assert((V & kLdsWords) != 0, "the one-lane walk over the pair masks keeps its words in LDS");
          if (pm.whole[t] == 0)
            visit_masked_lds<4>(g, off, bx, by, bz, t, n, pm.rows, rt, (int)blockDim.x, ld32, body_m);
          else
            visit_auto<MODE | VISIT_STREAM | VISIT_ROW_BATCH | ((V & 8) ? VISIT_PACKED : 0)>(g, off, k.max_scan, bx, by, bz, t, n, L, ld, body);
        } else if constexpr ((V & kLdsWalk) != 0)
          visit_tile<6>(g, off, k.max_scan, bx, by, bz, *tile, ld, body);
        else
          visit_auto<MODE | ((V & kRowWalk) ? 0 : (V & kWalkAlt) ? VISIT_STREAM : VISIT_STREAM | VISIT_ROW_BATCH | ((V & 8) ? VISIT_PACKED : 0))>(g, off, k.max_scan, bx, by, bz, t, n, L, ld, body);
      }
    }
    const float ox = sum.ox, oy = sum.oy, oz = sum.oz;
    const float ddx = (ox * k.delta_scale) / k.rho0, ddy = (oy * k.delta_scale) / k.rho0,
                ddz = (oz * k.delta_scale) / k.rho0;
    const float2 dl = dl_s[t];
    const float4 pn = make_float4(pi.x + ddx, pi.y + ddy, pi.z + ddz, dl.x);
    if (pred_out) pred_out[t] = pn;
    if constexpr ((V & kMaskWalk) != 0) soa_store(pm.pos, t, pn.x, pn.y, pn.z);
    if (vel_s) {
      const float dtc = fmaxf(k.dt, 1e-6f);
      const float4 po = pos_s[t];
      vel_s[t] = make_float4((pn.x - po.x) / dtc, (pn.y - po.y) / dtc, (pn.z - po.z) / dtc, dl.x);
    }
    if (delta4) delta4[min(max(sidx[t], 0), n - 1)] = make_float4(ddx, ddy, ddz, 0.0f);
    if (skey) {
      // the bound is accumulated in float32 with a 1% margin on the limit
      const float moved = L.disp[t] + sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
      L.disp[t] = moved;
      if (moved > 0.0f) atomicMax(&L.dcell[skey[t]], __float_as_int(moved));
    }
  }
}

template <int V>
__global__ __launch_bounds__(256) void k_delta(int n, Grid g_, Consts k_,
                                               const float4* __restrict__ pred_in,
                                               const float2* __restrict__ dl_s,
                                               const int32_t* __restrict__ off_,
                                               float4* __restrict__ pred_out,
                                               const float4* __restrict__ pos_s,
                                               float4* __restrict__ vel_s,
                                               const int32_t* __restrict__ sidx,
                                               float4* __restrict__ delta4, Lists L,
                                               const int32_t* __restrict__ skey, Scenes S,
                                               int iter = -1, const int32_t* __restrict__ tkey = nullptr,
                                               PairMasks pm = PairMasks{nullptr, nullptr}) {
  // kLdsWords: each lane's nine row words, row r at
  // fx_delta_words[r * blockDim.x + lane]
  extern __shared__ uint64_t fx_delta_words[];
  uint64_t* const rt = ((V & kLdsWords) != 0) ? fx_delta_words + threadIdx.x : nullptr;
  if constexpr ((V & kPersistent) != 0) {
    constexpr bool QUADW = (V & kQuadWalk) != 0;
    const long long nth = (QUADW ? 4LL * n : (long long)n);
    for (int vb = persistent_take(S.queue, nth); vb >= 0; vb = persistent_take(S.queue, nth))
      delta_body<V>(vb * (int)blockDim.x + (int)threadIdx.x, n, g_, k_, pred_in, dl_s, off_, pred_out, pos_s, vel_s, sidx, delta4, L, skey, S, iter, nullptr, pm, rt);
  } else if constexpr ((V & kLdsWords) != 0) {
    delta_body<V>(blockIdx.x * blockDim.x + threadIdx.x, n, g_, k_, pred_in, dl_s, off_, pred_out, pos_s, vel_s, sidx, delta4, L, skey, S, iter, nullptr, pm, rt);
  } else if constexpr ((V & kLdsWalk) != 0) {
    // launched with kTileBlock threads; tkey: the sorted build keys
    __shared__ Tile tile;
    const int t0 = xcd_block() * kTileBlock;
    stage_tile(tile, g_, off_, tkey, pred_in, t0, n);
    delta_body<V>(t0 + (int)threadIdx.x, n, g_, k_, pred_in, dl_s, off_, pred_out, pos_s, vel_s, sidx, delta4, L, skey, S, iter, &tile);
  } else {
    delta_body<V>(blockIdx.x * blockDim.x + threadIdx.x, n, g_, k_, pred_in, dl_s, off_, pred_out, pos_s, vel_s, sidx, delta4, L, skey, S, iter, nullptr, pm);
  }
}

// shaders/sph_velocity.glsl:12-24 in sorted space (used when iterations = 0).
__global__ void k_velocity_s(int n, float dt, const float4* __restrict__ pred_s,
                             const float4* __restrict__ pos_s, const float2* __restrict__ dl_s,
                             float4* __restrict__ vel_s) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float dtc = fmaxf(dt, 1e-6f);
  const float4 pn = pred_s[t], po = pos_s[t];
  vel_s[t] = make_float4((pn.x - po.x) / dtc, (pn.y - po.y) / dtc, (pn.z - po.z) / dtc, dl_s[t].x);
}

// shaders/sph_xsph.glsl:28-82 and, fused, shaders/sph_finalize.glsl:15-96.
// pred_s: final predicted positions; vel_s: (raw velocity, density).
// FUSE: apply the correction, collide, clamp and write canonical pos / vel
// (plus pos_s for the splat); otherwise write the correction to delta4 only.
// V = 4 * FUSE + MODE (one template argument keeps launch macros simple)
template <int V>
__device__ __forceinline__ void xsph_body(int gt_in, int n, FxParams p_, Grid g_, Consts k_, Collider col_,
                                          const float4* __restrict__ pred_s, const float4* __restrict__ vel_s,
                                          const int32_t* __restrict__ off_, const int32_t* __restrict__ sidx,
                                          float4* __restrict__ delta4, float4* __restrict__ pos4,
                                          float4* __restrict__ vel4, float4* __restrict__ pos_s,
                                          float4* __restrict__ pred4_out, Lists L, Scenes S,
                                          float4* __restrict__ normal4, int normal_by_slot) {
  constexpr bool FUSE = (V & 4) != 0;
  constexpr bool SPLIT = (V & 16) != 0;  // dv in sorted order for k_finalize_s
  constexpr bool NORMAL = (V & 32) != 0;  // also the next substep's normal pass (k_normal)
  constexpr int MODE = V & 3;
  constexpr bool QUAD = (V & kQuadWalk) != 0;
  const int gt = gt_in;
  const int t = QUAD ? gt >> 2 : gt;
  if (t >= n) return;
  FxParams p = p_;
  Grid g = g_;
  Consts k = k_;
  Collider col = col_;
  const int32_t* off = off_;
  if (V & 8) {
    const SceneDev& sc = S.s[scene_of(S, t)];
    p = sc.p;
    g = sc.g;
    k = sc.k;
    col = sc.col;
    off = off_ + sc.coff;
  }
  const float4 pi = pred_s[t];
  const float4 vi = vel_s[t];
  int bx, by, bz;
  cell_coord(g, pi.x, pi.y, pi.z, bx, by, bz);
  XsSums<NORMAL> sum;
  {
    const float h2 = k.h2;
    if constexpr (QUAD) {
      // each lane loads its candidate's velocity too (valid slot, finite),
      // so a term that does not contribute is a clean zero
      auto ld = [&](int q) { return q; };
      auto tm = [&](int q, int qc, bool valid) {
        const float4 pj = pred_s[qc], vj = vel_s[qc];
        const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
        const float r2 = dist2_exact(dx, dy, dz);
        const XsTerm z{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        return valid && r2 < h2 && q != t ? xs_term<NORMAL>(k, vi, vj, dx, dy, dz, r2) : z;
      };
      visit_quad<4>(g, off, k.max_scan, bx, by, bz, gt & 3, ld, tm, [&](const XsTerm& tt) { sum.add(tt); });
      if (gt & 3) return;
    } else {
      // stage 1 loads positions; stage 2 loads velocities only for candidates
      // within h (predicated, still before any use); stage 3 accumulates.
      // Positions from the x / y / z arrays instead (written by the last
      // delta; bit-identical), on the 1.64M scene: row batches of 4 slots (3
      // loads) 560 us, of 8 718, three scalar loads per candidate in this
      // stream 1013, against 520 here (row batches save position loads but
      // add partly empty batches at row ends, each still issuing its
      // velocity loads).  Loading every candidate's velocity with its
      // position, without waiting for the range test: 749 us.  Each lane's
      // nine row runs kept in LDS instead of the register queue (a row change
      // one LDS read instead of a shift unrolled at every slot): 69 instead
      // of 72 registers, 517 against 522 us; not kept.
      auto ld = [&](int q) {
        PV c;
        c.p = pred_s[q];
        return c;
      };
      auto rf = [&](int q, PV& c) {
        c.r2 = dist2_exact(pi.x - c.p.x, pi.y - c.p.y, pi.z - c.p.z);
        if (c.r2 < h2 && q != t) c.v = vel_s[q];
      };
      // xs_term and XsSums::add, written out so each sum is updated as soon
      // as its weight is known (fewer live registers; the same operations)
      visit_auto<MODE | ((V & kRowWalk) ? 0 : (V & kWalkAlt) ? VISIT_STREAM : VISIT_STREAM | VISIT_MASKED_LOADS)>(g, off, k.max_scan, bx, by, bz, t, n, L, ld, [&](int q, const PV& c) -> bool {
#pragma clang fp contract(off)
        const float4 vj = c.v;
        const float dx = pi.x - c.p.x, dy = pi.y - c.p.y, dz = pi.z - c.p.z;
        const float r2 = c.r2;
        if (r2 < h2 && q != t) {
          const float d = h2 - r2;
          const float irho = __builtin_amdgcn_rcpf(fmaxf(vj.w, 1e-6f));
          const float w = ((d * d) * d) * irho;
          sum.ax = __builtin_fmaf(vj.x - vi.x, w, sum.ax);
          sum.ay = __builtin_fmaf(vj.y - vi.y, w, sum.ay);
          sum.az = __builtin_fmaf(vj.z - vi.z, w, sum.az);
          if (NORMAL) {
            const float gsc = (d * d) * irho;
            sum.nx = __builtin_fmaf(dx, gsc, sum.nx);
            sum.ny = __builtin_fmaf(dy, gsc, sum.ny);
            sum.nz = __builtin_fmaf(dz, gsc, sum.nz);
            const float hr = r2 > 0.0f ? __builtin_fmaf(-r2, __builtin_amdgcn_rsqf(r2), k.h) : k.h;  // as xs_term
            sum.lap = __builtin_fmaf(hr, irho, sum.lap);
          }
        }
        return false;
      }, rf);
    }
  }
  float ax = sum.ax, ay = sum.ay, az = sum.az, nx = sum.nx, ny = sum.ny, nz = sum.nz, lap = sum.lap;
  if (NORMAL) {
    const float gfac = k.mass * (k.poly6 * -6.0f);
    nx *= gfac;
    ny *= gfac;
    nz *= gfac;
    lap *= k.mass * k.visc_lap;
    const float mag = sqrtf(nx * nx + ny * ny + nz * nz);
    const float curvature = (mag > 1e-4f) ? -lap / mag : 0.0f;
    normal4[normal_by_slot ? t : min(max(sidx[t], 0), n - 1)] = make_float4(nx, ny, nz, curvature);
  }
  const float f = k.visc * (k.mass * k.poly6);
  const float dvx = f * ax, dvy = f * ay, dvz = f * az;
  if (SPLIT) {
    delta4[t] = make_float4(dvx, dvy, dvz, 0.0f);
    return;
  }
  const int i = min(max(sidx[t], 0), n - 1);
  if (!FUSE) {
    delta4[i] = make_float4(dvx, dvy, dvz, 0.0f);
    return;
  }
  if (delta4) delta4[i] = make_float4(dvx, dvy, dvz, 0.0f);
  float vx = vi.x + dvx, vy = vi.y + dvy, vz = vi.z + dvz;
  float px = pi.x, py = pi.y, pz = pi.z;
  if (pred4_out) pred4_out[i] = make_float4(px, py, pz, 1.0f);
  finalize_one(p, col, px, py, pz, vx, vy, vz);
  const float4 pf = make_float4(px, py, pz, 1.0f);
  pos4[i] = pf;
  vel4[i] = make_float4(vx, vy, vz, 0.0f);
  if (pos_s) pos_s[t] = pf;
}

// XSPH's blocks in one contiguous range per XCD (xcd_block) against
// round-robin: 515 / 521 us at 1.64M particles, but the 4.1M animated scene's
// substep takes 10.30 / 9.95 ms (lambda and delta measured slower with it at
// both sizes: 1.64M lambda 302 -> 319 us, delta 131 -> 144; 4.1M substep
// 10.3 -> 18.4 ms)
constexpr int kXsphXcdBelow = 1 << 21;

template <int V>
__global__ __launch_bounds__(256) void k_xsph(int n, FxParams p_, Grid g_, Consts k_,
                                              Collider col_, const float4* __restrict__ pred_s,
                                              const float4* __restrict__ vel_s,
                                              const int32_t* __restrict__ off_,
                                              const int32_t* __restrict__ sidx,
                                              float4* __restrict__ delta4,
                                              float4* __restrict__ pos4, float4* __restrict__ vel4,
                                              float4* __restrict__ pos_s,
                                              float4* __restrict__ pred4_out, Lists L, Scenes S,
                                              float4* __restrict__ normal4 = nullptr, int normal_by_slot = 0) {
  if constexpr ((V & kPersistent) != 0) {
    constexpr bool QUADW = (V & kQuadWalk) != 0;
    const long long nth = (QUADW ? 4LL * n : (long long)n);
    for (int vb = persistent_take(S.queue, nth); vb >= 0; vb = persistent_take(S.queue, nth))
      xsph_body<V>(vb * (int)blockDim.x + (int)threadIdx.x, n, p_, g_, k_, col_, pred_s, vel_s, off_, sidx, delta4, pos4, vel4, pos_s, pred4_out, L, S, normal4, normal_by_slot);
  } else {
    // blocks renumbered per XCD below kXsphXcdBelow particles only
    const int b = n < kXsphXcdBelow ? xcd_block() : (int)blockIdx.x;
    xsph_body<V>(b * blockDim.x + threadIdx.x, n, p_, g_, k_, col_, pred_s, vel_s, off_, sidx, delta4, pos4, vel4, pos_s, pred4_out, L, S, normal4, normal_by_slot);
  }
}

// shaders/sph_finalize.glsl:15-96 over sorted slots, after a SPLIT XSPH pass
// left dv in sorted order: apply dv, collide, clamp, write the canonical
// position / velocity (and correction / prediction when asked) and pos_s.
// Keeping the collider search out of the neighbour pass keeps that pass's
// register count, and so its occupancy, down (63 instead of 81 VGPRs).
// V = 8: batched scenes (each particle's params and collider from its scene).
// The record a substep leaves when its XSPH pass computed the next
// substep's normals (see the workspace header in fx_common.h); flags NULL:
// none.
struct NormalMark {
  int32_t* flags;
  int32_t state;  // 1: in normal_next, 2: in the caller's normal buffer
  NormalKey key;
};

template <int V>
__device__ __forceinline__ void finalize_s_body(int gt_in, int n, FxParams p_, Collider col_,
                                                const float4* __restrict__ pred_s, const float4* __restrict__ vel_s,
                                                const float4* __restrict__ dv_s, const int32_t* __restrict__ sidx,
                                                float4* __restrict__ delta4, float4* __restrict__ pos4,
                                                float4* __restrict__ vel4, float4* __restrict__ pos_s,
                                                float4* __restrict__ pred4_out, Scenes S, NormalMark mk) {
  const int t = gt_in;
  if (mk.flags && t == 0) {
    mk.flags[kHdrNormals] = mk.state;
    mk.flags[kHdrH] = mk.key.hbits;
    mk.flags[kHdrMass] = mk.key.mbits;
    mk.flags[kHdrNormLo] = (int32_t)mk.key.lo;
    mk.flags[kHdrNormHi] = (int32_t)mk.key.hi;
  }
  if (t >= n) return;
  FxParams p = p_;
  Collider col = col_;
  if (V & 8) {
    const SceneDev& sc = S.s[scene_of(S, t)];
    p = sc.p;
    col = sc.col;
  }
  const float4 pi = pred_s[t], vi = vel_s[t], dv = dv_s[t];
  const int i = min(max(sidx[t], 0), n - 1);
  if (delta4) delta4[i] = make_float4(dv.x, dv.y, dv.z, 0.0f);
  float vx = vi.x + dv.x, vy = vi.y + dv.y, vz = vi.z + dv.z;
  float px = pi.x, py = pi.y, pz = pi.z;
  if (pred4_out) pred4_out[i] = make_float4(px, py, pz, 1.0f);
  finalize_one(p, col, px, py, pz, vx, vy, vz);
  const float4 pf = make_float4(px, py, pz, 1.0f), vf = make_float4(vx, vy, vz, 0.0f);
  if (pos4) pos4[i] = pf;
  if (vel4) vel4[i] = vf;
  if (pos_s) pos_s[t] = pf;
}

template <int V>
__global__ __launch_bounds__(256) void k_finalize_s(int n, FxParams p_, Collider col_,
                                                    const float4* __restrict__ pred_s,
                                                    const float4* __restrict__ vel_s,
                                                    const float4* __restrict__ dv_s,
                                                    const int32_t* __restrict__ sidx,
                                                    float4* __restrict__ delta4, float4* __restrict__ pos4,
                                                    float4* __restrict__ vel4, float4* __restrict__ pos_s,
                                                    float4* __restrict__ pred4_out, Scenes S,
                                                    NormalMark mk = NormalMark{}) {
  if constexpr ((V & kPersistent) != 0) {
    constexpr bool QUADW = (V & kQuadWalk) != 0;
    const long long nth = (long long)n;
    for (int vb = persistent_take(S.queue, nth); vb >= 0; vb = persistent_take(S.queue, nth))
      finalize_s_body<V>(vb * (int)blockDim.x + (int)threadIdx.x, n, p_, col_, pred_s, vel_s, dv_s, sidx, delta4, pos4, vel4, pos_s, pred4_out, S, mk);
  } else {
    finalize_s_body<V>(blockIdx.x * blockDim.x + threadIdx.x, n, p_, col_, pred_s, vel_s, dv_s, sidx, delta4, pos4, vel4, pos_s, pred4_out, S, mk);
  }
}

// How a substep's normal pass decides on the device what to do (flags NULL:
// always compute).  With no grid for (n, nc) in the workspace header it
// writes zero normals and sets the header's error word instead of reading
// offsets the workspace does not hold; when the header records normals made
// for `key` it takes them (state 1: copies normal_next, state 2: they are in
// place already) instead of traversing the stencil again.  The choice is the
// kernel's, so a captured substep makes it afresh on every replay.
struct NormalPick {
  int32_t* flags;
  const float4* next;
  NormalKey key;
  int nc;
};

// shaders/sph_normal.glsl:24-81 over the previous grid; pred_s.w = the
// neighbour's density from the last lambda pass.  Writes canonical normal4.
template <int V>
__device__ __forceinline__ void normal_body(int gt_in, int n, Grid g_, Consts k_, const float4* __restrict__ pred_s,
                                            const int32_t* __restrict__ off_, const int32_t* __restrict__ sidx,
                                            float4* __restrict__ normal4, Lists L, Scenes S, NormalPick pk) {
  constexpr int MODE = V & 3;
  const int t = gt_in;
  if (pk.flags) {
    if (!header_grid_ok(pk.flags, n, pk.nc)) {
      if (t < n) normal4[t] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (t == 0) pk.flags[kHdrError] = 1;
      return;
    }
    const int state = pk.flags[kHdrNormals];
    if (state != 0 && pk.flags[kHdrH] == pk.key.hbits && pk.flags[kHdrMass] == pk.key.mbits &&
        (uint32_t)pk.flags[kHdrNormLo] == pk.key.lo && (uint32_t)pk.flags[kHdrNormHi] == pk.key.hi) {
      if (state == 1 && t < n) normal4[t] = pk.next[t];
      return;
    }
  }
  if (t >= n) return;
  Grid g = g_;
  Consts k = k_;
  const int32_t* off = off_;
  if (V & 8) {
    const SceneDev& sc = S.s[scene_of(S, t)];
    g = sc.g;
    k = sc.k;
    off = off_ + sc.coff;
  }
  const float4 pi = pred_s[t];
  int bx, by, bz;
  cell_coord(g, pi.x, pi.y, pi.z, bx, by, bz);
  float nx = 0.0f, ny = 0.0f, nz = 0.0f, lap = 0.0f;
  {
#pragma clang fp contract(fast)
    const float h = k.h, h2 = k.h2;
    auto ld = [&](int q) { return pred_s[q]; };
    visit_auto<MODE>(g, off, k.max_scan, bx, by, bz, t, n, L, ld, [&](int q, const float4& pj) -> bool {
      const float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      const float r2 = dist2_exact(dx, dy, dz);
      if (r2 < h2 && q != t) {
        const float inv_rho = __builtin_amdgcn_rcpf(fmaxf(pj.w, 1e-6f));
        const float tt = h2 - r2;
        const float gsc = tt * tt * inv_rho;
        nx += gsc * dx;
        ny += gsc * dy;
        nz += gsc * dz;
        // h - r with one rounding, as the fused XSPH pass computes it; a
        // distinct particle at the same position (r2 = +0) gives h, as the
        // reference's w_visc_lap(sqrt(r2)) does (r2 * rsq(r2) would be NaN)
        const float hr = r2 > 0.0f ? __builtin_fmaf(-r2, __builtin_amdgcn_rsqf(r2), h) : h;
        lap = __builtin_fmaf(hr, inv_rho, lap);
      }
      return false;
    });
  }
  const float gfac = k.mass * (k.poly6 * -6.0f);
  nx *= gfac;
  ny *= gfac;
  nz *= gfac;
  lap *= k.mass * k.visc_lap;
  const float mag = sqrtf(nx * nx + ny * ny + nz * nz);
  const float curvature = (mag > 1e-4f) ? -lap / mag : 0.0f;
  normal4[min(max(sidx[t], 0), n - 1)] = make_float4(nx, ny, nz, curvature);
}

template <int V>
__global__ __launch_bounds__(256) void k_normal(int n, Grid g_, Consts k_,
                                                const float4* __restrict__ pred_s,
                                                const int32_t* __restrict__ off_,
                                                const int32_t* __restrict__ sidx,
                                                float4* __restrict__ normal4, Lists L, Scenes S,
                                                NormalPick pk = NormalPick{}) {
  if constexpr ((V & kPersistent) != 0) {
    constexpr bool QUADW = (V & kQuadWalk) != 0;
    const long long nth = (long long)n;
    for (int vb = persistent_take(S.queue, nth); vb >= 0; vb = persistent_take(S.queue, nth))
      normal_body<V>(vb * (int)blockDim.x + (int)threadIdx.x, n, g_, k_, pred_s, off_, sidx, normal4, L, S, pk);
  } else {
    normal_body<V>(blockIdx.x * blockDim.x + threadIdx.x, n, g_, k_, pred_s, off_, sidx, normal4, L, S, pk);
  }
}


// Batched fused predict + grid key + histogram: particle i belongs to the
// scene whose particle range holds it; keys are offset by the scene's cells.
__global__ __launch_bounds__(256) void k_b_predict_key(int ntot, Scenes S, const float4* __restrict__ pos,
                                                       float4* __restrict__ vel,
                                                       const float4* __restrict__ nrm,
                                                       float4* __restrict__ pred,
                                                       int32_t* __restrict__ key,
                                                       int32_t* __restrict__ cnt,
                                                       int32_t* __restrict__ rank, int do_predict) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < ntot;
  int k = 0;
  if (active) {
    const SceneDev& sc = S.s[scene_of(S, i)];
    float4 pp;
    if (do_predict) {
      float4 vo;
      const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
      predict_one(sc.p, pos[i], vel[i], (sc.p.surface_tension > 0.0f && nrm) ? nrm[i] : zero, vo, pp);
      vel[i] = vo;
      pred[i] = pp;
    } else {
      pp = pos[i];
      pp.w = 1.0f;
      pred[i] = pp;
    }
    int x, y, z;
    cell_coord(sc.g, pp.x, pp.y, pp.z, x, y, z);
    k = sc.coff + cell_index(sc.g, x, y, z);
    key[i] = k;
    rank[i] = atomicAdd(&cnt[k], 1);  // lattice order: little contention (see k_predict_key)
  }
}

__global__ void k_b_load_state(int ntot, float4* __restrict__ pos, float4* __restrict__ vel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntot) return;
  float4 p = pos[i];
  p.w = 1.0f;
  pos[i] = p;
  float4 v = vel[i];
  v.w = 0.0f;
  vel[i] = v;
}

// ---- canonical streaming passes (FX_PASS_*) --------------------------------

__global__ void k_apply_delta_c(int n, float4* __restrict__ pred, const float4* __restrict__ delta) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 p = pred[i], d = delta[i];
  pred[i] = make_float4(p.x + d.x, p.y + d.y, p.z + d.z, 1.0f);
}

__global__ void k_velocity_c(int n, float dt, const float4* __restrict__ pred,
                             const float4* __restrict__ pos, float4* __restrict__ vel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float dtc = fmaxf(dt, 1e-6f);
  const float4 a = pred[i], b = pos[i];
  vel[i] = make_float4((a.x - b.x) / dtc, (a.y - b.y) / dtc, (a.z - b.z) / dtc, 0.0f);
}

__global__ void k_finalize_c(FxParams p, Collider col, float4* __restrict__ vel,
                             const float4* __restrict__ delta, const float4* __restrict__ pred,
                             float4* __restrict__ pos) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const float4 v = vel[i], d = delta[i], q = pred[i];
  float vx = v.x + d.x, vy = v.y + d.y, vz = v.z + d.z;
  float px = q.x, py = q.y, pz = q.z;
  finalize_one(p, col, px, py, pz, vx, vy, vz);
  pos[i] = make_float4(px, py, pz, 1.0f);
  vel[i] = make_float4(vx, vy, vz, 0.0f);
}

// reset / cache load: pred = pos, w lanes as the reference stores them
__global__ void k_load_state(int n, float4* __restrict__ pos, float4* __restrict__ vel,
                             float4* __restrict__ pred, float4* __restrict__ lambda4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 p = pos[i];
  p.w = 1.0f;
  pos[i] = p;
  float4 v = vel[i];
  v.w = 0.0f;
  vel[i] = v;
  if (pred) pred[i] = p;
  if (lambda4) lambda4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}


__global__ void k_w_from_lambda(int n, float4* pred_s, const float2* dl_s) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  pred_s[t].w = dl_s[t].y;
}

__global__ void k_gather_xsph(int n, const int32_t* sidx, const float4* pred, const float4* vel,
                              const float4* lambda4, float4* pred_s, float4* vel_s) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = min(max(sidx[t], 0), n - 1);
  pred_s[t] = pred[i];
  float4 v = vel[i];
  v.w = lambda4[i].x;
  vel_s[t] = v;
}

__global__ void k_sorted_keys(int n, const int32_t* __restrict__ sidx, const int32_t* __restrict__ key,
                              int32_t* __restrict__ skey) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  skey[t] = key[min(max(sidx[t], 0), n - 1)];
}

// ---- host side --------------------------------------------------------------

static int validate(const FxParams* p, const FxState* s, bool need_work) {
  if (!p || !s) return fail(FX_EINVAL, "null params or state");
  if (p->n < 0) return fail(FX_EINVAL, "particle count %d < 0", p->n);
  if (p->n > kMaxParticles)
    return fail(FX_ERANGE, "particle count %d exceeds %d (slots are addressed by 32-bit byte offsets)", p->n,
                kMaxParticles);
  if (p->cells[0] < 1 || p->cells[1] < 1 || p->cells[2] < 1)
    return fail(FX_EINVAL, "cell dims (%d,%d,%d) must be >= 1", p->cells[0], p->cells[1], p->cells[2]);
  long long nc = (long long)p->cells[0] * p->cells[1] * p->cells[2];
  if (nc != p->n_cells) return fail(FX_EINVAL, "n_cells %d != dims product %lld", p->n_cells, nc);
  if (nc >= (1LL << 30)) return fail(FX_ERANGE, "cell count %lld too large", nc);
  if (!(p->cell_size > 0.0f)) return fail(FX_EINVAL, "cell_size must be > 0");
  if (!(p->h > 0.0f)) return fail(FX_EINVAL, "smoothing radius must be > 0");
  if (p->iterations < 0) return fail(FX_EINVAL, "iterations < 0");
  if (p->collider_voxel > 0.0f &&
      (p->collider_dims[0] < 1 || p->collider_dims[1] < 1 || p->collider_dims[2] < 1))
    return fail(FX_EINVAL, "collider dims must be >= 1 when collider_voxel > 0");
  if (need_work) {
    size_t need = workspace_layout(*p).total;
    if (!s->work) return fail(FX_EINVAL, "workspace pointer is NULL");
    if (s->work_bytes < need)
      return fail(FX_ENOMEM, "workspace too small: %zu < %zu bytes", s->work_bytes, need);
  }
  return FX_OK;
}

#define LAUNCH(kernel, n, st, ...) FX_LAUNCH(kernel, dim3(grid_for(n)), dim3(kBlock), st, __VA_ARGS__)
// neighbour passes use the tuned block size (64, 128 or 256)
#define NLAUNCH(kernel, n, st, ...) \
  FX_LAUNCH(kernel, dim3(grid_for(n, g_tuning.block)), dim3(g_tuning.block), st, __VA_ARGS__)
// a full-stencil neighbour pass with the walk chosen by scene size: four
// lanes per particle under quad_below, row by row under row_below, else the
// stream walk (V unparenthesised: the launch's profile label is the kernel
// text)
#define QLAUNCH(kernel, V, n, quad_below, row_below, st, ...)                                             \
  do {                                                                                                    \
    if ((n) >= (quad_below) && (n) >= (row_below)) {                                                      \
      if (g_tuning.walk == 1)                                                                             \
        NLAUNCH(kernel<V + kRowWalk>, n, st, __VA_ARGS__);                                                \
      else if (g_tuning.walk == 2)                                                                        \
        NLAUNCH(kernel<V + kWalkAlt>, n, st, __VA_ARGS__);                                                \
      else                                                                                                \
        NLAUNCH(kernel<V>, n, st, __VA_ARGS__);                                                           \
    } else if ((n) >= (quad_below))                                                                       \
      NLAUNCH(kernel<V + kRowWalk>, n, st, __VA_ARGS__);                                                  \
    else                                                                                                  \
      FX_LAUNCH(kernel<V + kQuadWalk>, dim3(grid_for(4LL * (n), g_tuning.block)), dim3(g_tuning.block), st, \
                __VA_ARGS__);                                                                             \
  } while (0)

// a batched pass: flat, or persistent with FxTuning.persistent workgroups per CU
// (V unparenthesised, as above)
#define BLAUNCH(kernel, V, n, st, ...)                                                                   \
  do {                                                                                                   \
    if (g_tuning.persistent > 0) {                                                                       \
      FX_TRY(fill(S.queue, kXcds, 0, st));                                                               \
      FX_LAUNCH(kernel<V + kPersistent>, dim3(persistent_grid(n, g_tuning.block)), dim3(g_tuning.block),   \
                st, __VA_ARGS__);                                                                        \
    } else                                                                                               \
      NLAUNCH(kernel<V>, n, st, __VA_ARGS__);                                                            \
  } while (0)
// the same with `shm` bytes of dynamic LDS per workgroup
#define BLAUNCH_SHM(kernel, V, n, shm, st, ...)                                                          \
  do {                                                                                                   \
    if (g_tuning.persistent > 0) {                                                                       \
      FX_TRY(fill(S.queue, kXcds, 0, st));                                                               \
      FX_LAUNCH_SHM(kernel<V + kPersistent>, dim3(persistent_grid(n, g_tuning.block)),                    \
                    dim3(g_tuning.block), shm, st, __VA_ARGS__);                                         \
    } else                                                                                               \
      FX_LAUNCH_SHM(kernel<V>, dim3(grid_for(n, g_tuning.block)), dim3(g_tuning.block), shm, st,          \
                    __VA_ARGS__);                                                                        \
  } while (0)

// Workgroups of a persistent launch over n threads: g_tuning.persistent per
// CU, a multiple of the 8 XCDs, no more than the flat launch would use.
static int g_cus = 0;
static unsigned persistent_grid(long long n, int block) {
  if (g_cus <= 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
      g_cus = cus;
    else
      g_cus = 256;
  }
  long long g = (long long)g_cus * g_tuning.persistent;
  const long long flat = (n + block - 1) / block;
  if (g > flat) g = flat;
  g = (g + kXcds - 1) / kXcds * kXcds;
  return (unsigned)(g < kXcds ? kXcds : g);
}

static Lists make_lists(const Ws& w, const Consts& k, bool enabled) {
  Lists L;
  L.mask = enabled ? w.mask : nullptr;
  L.cell = w.lcell;
  L.full = w.lfull;
  L.disp = w.disp;
  L.dcell = w.dcell;
  L.built = w.flags + 2;
  const float skin = g_tuning.skin * k.h;
  const float R = k.h + skin;
  L.r2_list = R * R;
  L.dlimit = 0.99f * skin;
  return L;
}

// Pair masks (PairMasks) serve every walk without neighbour lists (the quad
// walk of small scenes records and walks them in quads), unless the tuning
// asks for the walks without them (walk 4).  Their words take the lists' mask
// area and whole-stencil flags.
static bool pair_masks_on(bool lists) { return !lists && g_tuning.walk == 0; }
static PairMasks pair_masks(const Ws& w, bool xyz = false) {
  return PairMasks{w.mask, w.lfull, xyz ? SoaPos{w.sx, w.sy, w.sz} : SoaPos{nullptr, nullptr, nullptr}};
}
// lambda's one-lane recording walk reads the x / y / z arrays (visit_record_xyz)
static bool xyz_candidates(long long n, bool lists) { return pair_masks_on(lists) && n >= kQuadBelowLambda; }
static SoaPos xyz_arrays(const Ws& w, long long n, bool lists) {
  return xyz_candidates(n, lists) ? SoaPos{w.sx, w.sy, w.sz} : SoaPos{nullptr, nullptr, nullptr};
}
// nine words per lane, or per quad of lanes in the quad walk
static size_t row_words_lds(int block, bool quad = false) {
  return (size_t)9 * sizeof(uint64_t) * (size_t)(quad ? block / 4 : block);
}

static int fill(int32_t* a, long long n, int32_t v, hipStream_t st) {
  if (n <= 0) return FX_OK;
  LAUNCH(k_fill_i32, n, st, a, n, v);
  return FX_OK;
}

int canon_cells(const int32_t* off, int32_t* sidx, int n_cells, int32_t* big_list, int32_t* big_count,
                hipStream_t st) {
  if (n_cells <= 0) return FX_OK;
  FX_TRY(fill(big_count, 1, 0, st));
  LAUNCH(k_canon_cells<int32_t>, n_cells, st, n_cells, off, sidx, big_list, big_count);
  return sort_queued_runs(off, sidx, big_list, big_count, st);
}

int sort_queued_runs(const int32_t* off, int32_t* data, const int32_t* queue, const int32_t* queue_count,
                     hipStream_t st) {
  FX_LAUNCH(k_sort_big_cells<int32_t>, dim3(kBigCellBlocks), dim3(256), st, off, data, queue, queue_count);
  return FX_OK;
}

// Counting sort of ws.key (counts already in ws.cnt, ranks in ws.rank) into
// ws.sidx: scan, scatter (which zeroes the counts again), canonicalise.  Writes reference-format ranges if
// requested.
static int sort_from_counts(const FxParams& p, const FxState& s, const Ws& w, hipStream_t st) {
  const int n = p.n, nc = p.n_cells;
  FX_TRY(exclusive_scan_i32(w.cnt, w.off, nc, w.partial, st));
  if (s.cell_start && s.cell_end) LAUNCH(k_cell_ranges, nc, st, nc, w.off, s.cell_start, s.cell_end);
  if (n > 0) {
    LAUNCH(k_scatter, std::max(n, nc + 1), st, n, w.key, w.rank, w.off, w.sidx, w.cnt, nc + 1);
    // big-cell queue reuses the rank array (dead after the scatter)
    FX_TRY(canon_cells(w.off, w.sidx, nc, w.rank, w.flags, st));
  }
  return check_launch("sort_from_counts");
}

// Full grid build from s.pred (canonical) with optional fused predict.
// target: which pred_s buffer receives the gathered predicted positions.
static int build_grid(const FxParams& p, FxState& s, const Ws& w, bool do_predict, int target,
                      hipStream_t st) {
  const int n = p.n, nc = p.n_cells;
  Grid g = make_grid(p);
  FX_TRY(fill(w.cnt, nc + 1, 0, st));
  int32_t* key = w.key;
  if (n > 0) {
    LAUNCH(k_predict_key, n, st, p, g, (const float4*)s.pos, (float4*)s.vel,
           (const float4*)s.normal, (float4*)s.pred, key, w.cnt, w.rank, do_predict ? 3 : 2);
  }
  FX_TRY(sort_from_counts(p, s, w, st));
  if (n > 0) {
    LAUNCH(k_gather, n, st, n, w.sidx, key, (const float4*)s.pred, (const float4*)s.pos,
           (const float4*)nullptr, w.pred_s[target], w.pos_s, w.dl_s, s.sorted_key, w.skey,
           s.sorted_idx, xyz_arrays(w, n, g_tuning.skin > 0.0f));
    if (s.key) FX_HIP_CHECK(hipMemcpyAsync(s.key, key, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  }
  // a new grid invalidates the neighbour lists until a lambda pass rebuilds
  // them, and any normals computed for the next substep
  FX_LAUNCH(k_mark_grid, dim3(1), dim3(64), st, w.flags, n, nc);
  return check_launch("build_grid");
}

// Grid build of a substep whose state enters by slot of the previous
// substep of the same fx_step call, predicted already by that substep's
// k_finalize_predict (prediction in vel_r, keys and ranks by slot, counts in
// cnt).  The counting sort carries (particle, previous slot) pairs, so each
// cell's run is still put in ascending particle order - the same order as
// build_grid - and the gather reads the previous slots nearly in sequence
// instead of gathering the canonical arrays.  The gather ranks each pair
// within its cell's run itself (runs of up to kSmallCell pairs; longer ones
// are sorted in place first): 25 us of canonicalisation and a 39 us gather
// became a 49 us gather at 1.64M.  The reference-format outputs
// (keys, sorted keys and indices, cell ranges) go to the caller on the
// call's last substep only.
static int build_grid_resident(const FxParams& p, FxState& s, const Ws& w, int target, bool last_in_call,
                               hipStream_t st) {
  const int n = p.n, nc = p.n_cells;
  const float4* pred_r = w.vel_r;  // the prediction by slot (see k_finalize_predict)
  FX_TRY(exclusive_scan_i32(w.cnt, w.off, nc, w.partial, st));
  if (last_in_call && s.cell_start && s.cell_end) LAUNCH(k_cell_ranges, nc, st, nc, w.off, s.cell_start, s.cell_end);
  // also zeroes the counts (nc + 1) and flags[0], and invalidates the lists
  // big-cell queue in the neighbour lists' cell array (the lists are off
  // in a resident build); its count was zeroed by k_finalize_predict
  LAUNCH(k_scatter_pairs, std::max(n, nc + 1), st, n, (const int32_t*)w.key, (const int32_t*)w.rank,
         (const int32_t*)w.off, (const int32_t*)w.sidx, w.pair64, w.cnt, nc + 1, w.flags, w.lcell);
  FX_LAUNCH(k_sort_big_cells<uint64_t>, dim3(kBigCellBlocks), dim3(256), st, (const int32_t*)w.off, w.pair64,
            (const int32_t*)w.lcell, (const int32_t*)w.flags);
  // the sorted keys: for the lists and the LDS-tiled walk, and after the
  // call's last substep for the splat (its per-cell displacement bounds bin
  // each slot by its key)
  const bool lists = g_tuning.skin > 0.0f;
  const bool skey = lists || g_tuning.walk == 3 || last_in_call;
  LAUNCH(k_gather_pairs, n, st, n, (const uint64_t*)w.pair64, (const int32_t*)w.off, (const int32_t*)w.key, pred_r,
         (const float4*)w.pos_r, w.pred_s[target], w.pos_s, p.iterations > 0 ? (float2*)nullptr : w.dl_s,
         last_in_call ? s.sorted_key : nullptr, skey ? w.skey : (int32_t*)nullptr, w.sidx,
         last_in_call ? s.sorted_idx : nullptr, last_in_call ? s.key : nullptr, xyz_arrays(w, n, lists));
  return check_launch("build_grid_resident");
}

// last: this substep's state is what the caller sees next.  Earlier
// substeps of one fx_step call leave out the outputs the next substep
// overwrites before anything reads them: the canonical prediction, lambda and
// correction, and the sorted final positions (read by the surface).
// Whether the substeps of one fx_step call can hand their state on by slot
// (build_grid_resident): not with neighbour lists (their XSPH pass writes the
// canonical state itself) and, with cohesion, only when the XSPH pass
// computes the next normals (iterations > 0).
static bool resident_ok(const FxParams& p, const FxState& s) {
  return p.n > 0 && g_tuning.skin <= 0.0f && (p.surface_tension <= 0.0f || (s.normal && p.iterations > 0));
}

// resident_in: the state enters by slot from the previous substep of the
// call; resident_out: it leaves that way for the next one.
// next_is_last: with resident_out, the next substep is the call's last.
static int substep(const FxParams& p, FxState& s, const Ws& w, hipStream_t st, bool last_in_call = true,
                   bool resident_in = false, bool resident_out = false, bool next_is_last = false) {
  const int n = p.n;
  Grid g = make_grid(p);
  Consts k = make_consts(p);
  Collider col = make_collider(p, s.collider);
  const bool lists = g_tuning.skin > 0.0f;
  Lists L = make_lists(w, k, lists);
  // sph_normal runs first, over the previous substep's grid and state
  // (shaders/sph_normal.glsl:15-21); its output feeds sph_predict.  A
  // resident substep takes the normals by slot from the previous XSPH pass.
  // Whether the previous substep's XSPH pass already computed these normals
  // (then they are copied or already in place) or the pass must run is read
  // from the workspace header by the kernel itself (NormalPick).
  if (p.surface_tension > 0.0f && !resident_in && n > 0) {
    if (!s.normal) return fail(FX_EINVAL, "surface_tension > 0 needs a normal buffer");
    const NormalPick pk{w.flags, w.normal_next, normal_key(p, s.normal), p.n_cells};
    // without lists the one-mode kernels need fewer registers (higher occupancy)
    if (lists)
      NLAUNCH(k_normal<VISIT_LIST>, n, st, n, g, k, w.pred_s[0], w.off, w.sidx, (float4*)s.normal, L,
              Scenes{nullptr, 0}, pk);
    else
      NLAUNCH(k_normal<VISIT_FULL>, n, st, n, g, k, w.pred_s[0], w.off, w.sidx, (float4*)s.normal, L,
              Scenes{nullptr, 0}, pk);
  }
  const int iters = p.iterations;
  int cur = iters & 1;  // the last iteration writes buffer 0
  if (resident_in) {
    FX_TRY(build_grid_resident(p, s, w, cur, last_in_call, st));
  } else {
    // the reference-format grid (keys, sorted keys and indices, cell ranges)
    // is published for the caller only by the call's last substep
    FxState sg = s;
    if (!last_in_call) sg.key = sg.sorted_key = sg.sorted_idx = sg.cell_start = sg.cell_end = nullptr;
    FX_TRY(build_grid(p, sg, w, true, cur, st));
  }
  if (n == 0) return check_launch("substep");
  if (lists && iters > 0) {
    FX_TRY(fill(w.dcell, p.n_cells, 0, st));
    FX_TRY(fill(w.flags + 2, 1, 0, st));
  }
  const bool masks = pair_masks_on(lists);
  const bool xyz = xyz_candidates(n, lists);
  const PairMasks pm = pair_masks(w, xyz);
  for (int it = 0; it < iters; ++it) {
    const bool last = it == iters - 1;
    float4* lam_out = (last && last_in_call && s.lambda) ? (float4*)s.lambda : (float4*)nullptr;
    if (masks && n < kQuadBelowLambda)
      FX_LAUNCH_SHM(k_lambda<VISIT_FULL + kQuadWalk + kMaskRecord>, dim3(grid_for(4LL * n, g_tuning.block)),
                    dim3(g_tuning.block), row_words_lds(g_tuning.block, true), st, n, g, k, w.pred_s[cur], w.off,
                    w.dl_s, w.sidx, lam_out, L, Scenes{nullptr, 0}, -1, (const int32_t*)nullptr, pm);
    else if (masks)
      FX_LAUNCH_SHM(k_lambda<VISIT_FULL + kMaskRecord + kXyzCand>, dim3(grid_for(n, g_tuning.block)),
                    dim3(g_tuning.block), row_words_lds(g_tuning.block), st, n, g, k, w.pred_s[cur], w.off,
                    w.dl_s, w.sidx, lam_out, L, Scenes{nullptr, 0}, -1, (const int32_t*)nullptr, pm);
    else if (!lists && g_tuning.walk == 3 && n >= kRowBelow)
      FX_LAUNCH(k_lambda<VISIT_FULL + kLdsWalk>, dim3(grid_for(n, kTileBlock)), dim3(kTileBlock), st, n, g, k,
                w.pred_s[cur], w.off, w.dl_s, w.sidx, lam_out, L, Scenes{nullptr, 0}, -1, (const int32_t*)w.skey);
    else if (!lists)
      QLAUNCH(k_lambda, VISIT_FULL, n, kQuadBelowLambda, 0, st, n, g, k, w.pred_s[cur], w.off, w.dl_s, w.sidx, lam_out, L, Scenes{nullptr, 0});
    else if (it == 0)
      NLAUNCH(k_lambda<VISIT_BUILD>, n, st, n, g, k, w.pred_s[cur], w.off, w.dl_s, w.sidx, lam_out, L, Scenes{nullptr, 0});
    else
      NLAUNCH(k_lambda<VISIT_LIST>, n, st, n, g, k, w.pred_s[cur], w.off, w.dl_s, w.sidx, lam_out, L, Scenes{nullptr, 0});
    if (masks && n < kQuadMaskBelowDelta)
      FX_LAUNCH(k_delta<VISIT_FULL + kQuadWalk + kMaskWalk>, dim3(grid_for(4LL * n, g_tuning.block)),
                dim3(g_tuning.block), st, n, g, k, (const float4*)w.pred_s[cur], (const float2*)w.dl_s, w.off,
                w.pred_s[cur ^ 1], (const float4*)w.pos_s, last ? w.vel_s : (float4*)nullptr, w.sidx,
                (float4*)nullptr, L, (const int32_t*)nullptr, Scenes{nullptr, 0}, -1, (const int32_t*)nullptr, pm);
    else if (masks)  // the next lambda's x / y / z candidates, unless this is the last iteration
      FX_LAUNCH_SHM(k_delta<VISIT_FULL + kMaskWalk + kLdsWords>, dim3(grid_for(n, g_tuning.block)),
                    dim3(g_tuning.block), row_words_lds(g_tuning.block), st, n, g, k, (const float4*)w.pred_s[cur],
                    (const float2*)w.dl_s, w.off, w.pred_s[cur ^ 1], (const float4*)w.pos_s,
                    last ? w.vel_s : (float4*)nullptr, w.sidx, (float4*)nullptr, L, (const int32_t*)nullptr,
                    Scenes{nullptr, 0}, -1, (const int32_t*)nullptr, pair_masks(w, xyz && !last));
    else if (lists)
      NLAUNCH(k_delta<VISIT_LIST>, n, st, n, g, k, (const float4*)w.pred_s[cur], (const float2*)w.dl_s, w.off,
              w.pred_s[cur ^ 1], (const float4*)w.pos_s, last ? w.vel_s : (float4*)nullptr, w.sidx,
              (float4*)nullptr, L, (const int32_t*)w.skey, Scenes{nullptr, 0});
    else if (g_tuning.walk == 3 && n >= kRowBelow)
      FX_LAUNCH(k_delta<VISIT_FULL + kLdsWalk>, dim3(grid_for(n, kTileBlock)), dim3(kTileBlock), st, n, g, k,
                (const float4*)w.pred_s[cur], (const float2*)w.dl_s, w.off, w.pred_s[cur ^ 1], (const float4*)w.pos_s,
                last ? w.vel_s : (float4*)nullptr, w.sidx, (float4*)nullptr, L, (const int32_t*)nullptr,
                Scenes{nullptr, 0}, -1, (const int32_t*)w.skey);
    else
      QLAUNCH(k_delta, VISIT_FULL, n, kQuadBelowDelta, kRowBelow, st, n, g, k, (const float4*)w.pred_s[cur], (const float2*)w.dl_s, w.off,
              w.pred_s[cur ^ 1], (const float4*)w.pos_s, last ? w.vel_s : (float4*)nullptr, w.sidx,
              (float4*)nullptr, L, (const int32_t*)nullptr, Scenes{nullptr, 0});
    cur ^= 1;
  }
  if (iters == 0) LAUNCH(k_velocity_s, n, st, n, p.dt, w.pred_s[0], w.pos_s, w.dl_s, w.vel_s);
  float4* delta_out = last_in_call ? (float4*)s.delta : nullptr;
  float4* pred_out = last_in_call ? (float4*)s.pred : nullptr;
  float4* pos_s_out = last_in_call ? w.pos_s : nullptr;
  if (lists)
    NLAUNCH(k_xsph<4 + VISIT_LIST>, n, st, n, p, g, k, col, (const float4*)w.pred_s[0], (const float4*)w.vel_s,
            w.off, w.sidx, delta_out, (float4*)s.pos, (float4*)s.vel, pos_s_out, pred_out, L,
            Scenes{nullptr, 0});
  else {
    // dv into the free ping-pong buffer, then the streaming finalize; with
    // cohesion on, the same traversal yields the next substep's normals
    const bool fuse_normals = p.surface_tension > 0.0f && s.normal && iters > 0;
    if (fuse_normals) {
      // resident: by slot for the next substep; last of the call: for the
      // next call; otherwise straight into the caller's buffer
      float4* nrm = (resident_out || last_in_call) ? w.normal_next : (float4*)s.normal;
      QLAUNCH(k_xsph, 32 + 16 + VISIT_FULL, n, kQuadBelowXsph, kRowBelow, st, n, p, g, k, col, (const float4*)w.pred_s[0],
              (const float4*)w.vel_s, w.off, w.sidx, w.pred_s[1], (float4*)nullptr, (float4*)nullptr,
              (float4*)nullptr, (float4*)nullptr, L, Scenes{nullptr, 0}, nrm, resident_out ? 1 : 0);
    } else {
      QLAUNCH(k_xsph, 16 + VISIT_FULL, n, kQuadBelowXsph, kRowBelow, st, n, p, g, k, col, (const float4*)w.pred_s[0], (const float4*)w.vel_s,
              w.off, w.sidx, w.pred_s[1], (float4*)nullptr, (float4*)nullptr, (float4*)nullptr, (float4*)nullptr,
              L, Scenes{nullptr, 0});
    }
    if (resident_out) {
      // finalize, and the next substep's predict (see k_finalize_predict);
      // either grid build's scatter has already zeroed the counts
      const bool cohesion = p.surface_tension > 0.0f;
      LAUNCH(k_finalize_predict, n, st, n, p, col, g, (const float4*)w.pred_s[0], (const float4*)w.vel_s,
             (const float4*)w.pred_s[1], cohesion ? (const float4*)w.normal_next : (const float4*)nullptr,
             (const int32_t*)w.sidx, w.pos_r, w.vel_r, w.key, w.cnt, w.rank,
             (cohesion && next_is_last) ? (float4*)s.normal : (float4*)nullptr, w.flags);
    } else {
      // with fused normals, the header records where they went
      const NormalMark mk = fuse_normals ? NormalMark{w.flags, last_in_call ? 1 : 2, normal_key(p, s.normal)}
                                         : NormalMark{};
      LAUNCH(k_finalize_s<0>, n, st, n, p, col, (const float4*)w.pred_s[0], (const float4*)w.vel_s,
             (const float4*)w.pred_s[1], w.sidx, delta_out, (float4*)s.pos, (float4*)s.vel, pos_s_out, pred_out,
             Scenes{nullptr, 0}, mk);
    }
  }
  return check_launch("substep");
}

// Prepare sorted-space inputs for a canonical neighbour pass: offsets from
// s.sorted_key, the permutation from s.sorted_idx.
static int canonical_grid(const FxParams& p, const FxState& s, const Ws& w, hipStream_t st) {
  if (!s.sorted_key || !s.sorted_idx) return fail(FX_EINVAL, "pass needs sorted_key and sorted_idx");
  const int n = p.n, nc = p.n_cells;
  if (n > 0) {
    FX_HIP_CHECK(hipMemcpyAsync(w.sidx, s.sorted_idx, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    FX_HIP_CHECK(hipMemcpyAsync(w.skey, s.sorted_key, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
  }
  FX_LAUNCH(k_off_from_sorted, dim3(grid_for((long long)n + 1)), dim3(kBlock), st, n, nc, w.skey, w.off);
  FX_LAUNCH(k_mark_grid, dim3(1), dim3(64), st, w.flags, n, nc);
  return check_launch("canonical_grid");
}

static int run_pass(int id, const FxParams& p, FxState& s, const Ws& w, hipStream_t st) {
  const int n = p.n, nc = p.n_cells;
  Grid g = make_grid(p);
  Consts k = make_consts(p);
  Collider col = make_collider(p, s.collider);
  const Lists L0 = make_lists(w, k, false);
  auto need = [&](const void* ptr, const char* name) -> int {
    if (!ptr) return fail(FX_EINVAL, "pass %d needs the %s buffer", id, name);
    return FX_OK;
  };
  switch (id) {
    case FX_PASS_PREDICT:
      FX_TRY(need(s.pos, "pos"));
      FX_TRY(need(s.vel, "vel"));
      FX_TRY(need(s.pred, "pred"));
      if (p.surface_tension > 0.0f) FX_TRY(need(s.normal, "normal"));
      if (n > 0)
        LAUNCH(k_predict_key, n, st, p, g, (const float4*)s.pos, (float4*)s.vel,
               (const float4*)s.normal, (float4*)s.pred, (int32_t*)nullptr, (int32_t*)nullptr,
               (int32_t*)nullptr, 1);
      break;
    case FX_PASS_GRID_KEY:
      FX_TRY(need(s.pred, "pred"));
      FX_TRY(need(s.key, "key"));
      if (n > 0)
        LAUNCH(k_predict_key, n, st, p, g, (const float4*)nullptr, (float4*)nullptr,
               (const float4*)nullptr, (float4*)s.pred, s.key, (int32_t*)nullptr, (int32_t*)nullptr,
               4);
      break;
    case FX_PASS_SORT: {
      FX_TRY(need(s.key, "key"));
      FX_TRY(need(s.sorted_key, "sorted_key"));
      FX_TRY(need(s.sorted_idx, "sorted_idx"));
      FX_TRY(fill(w.cnt, nc + 1, 0, st));
      FX_TRY(fill(w.flags + 1, 1, 0, st));
      if (n > 0) {
        FX_HIP_CHECK(hipMemcpyAsync(w.key, s.key, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        LAUNCH(k_count_keys, n, st, n, nc, w.key, w.cnt, w.rank, w.flags + 1);
      }
      FxState t = s;
      t.cell_start = t.cell_end = nullptr;
      FX_TRY(sort_from_counts(p, t, w, st));
      FX_LAUNCH(k_mark_grid, dim3(1), dim3(64), st, w.flags, n, nc);
      if (n > 0) {
        FX_HIP_CHECK(hipMemcpyAsync(s.sorted_idx, w.sidx, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        LAUNCH(k_sorted_keys, n, st, n, w.sidx, w.key, w.skey);
        FX_HIP_CHECK(hipMemcpyAsync(s.sorted_key, w.skey, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
      }
      break;
    }
    case FX_PASS_CELL_CLEAR:
      FX_TRY(need(s.cell_start, "cell_start"));
      FX_TRY(need(s.cell_end, "cell_end"));
      FX_TRY(fill(s.cell_start, nc, -1, st));
      FX_TRY(fill(s.cell_end, nc, -1, st));
      break;
    case FX_PASS_CELL_RANGE:
      FX_TRY(need(s.sorted_key, "sorted_key"));
      FX_TRY(need(s.cell_start, "cell_start"));
      FX_TRY(need(s.cell_end, "cell_end"));
      if (n > 0) LAUNCH(k_cell_range_ref, n, st, n, nc, s.sorted_key, s.cell_start, s.cell_end);
      break;
    case FX_PASS_LAMBDA:
      FX_TRY(need(s.pred, "pred"));
      FX_TRY(need(s.lambda, "lambda"));
      FX_TRY(canonical_grid(p, s, w, st));
      if (n > 0) {
        LAUNCH(k_gather, n, st, n, w.sidx, (const int32_t*)nullptr, (const float4*)s.pred, (const float4*)nullptr,
               (const float4*)nullptr, w.pred_s[0], (float4*)nullptr, (float2*)nullptr,
               (int32_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr);
        QLAUNCH(k_lambda, VISIT_FULL, n, kQuadBelowLambda, 0, st, n, g, k, w.pred_s[0], w.off, w.dl_s, w.sidx, (float4*)s.lambda, L0, Scenes{nullptr, 0});
      }
      break;
    case FX_PASS_DELTA:
      FX_TRY(need(s.pred, "pred"));
      FX_TRY(need(s.lambda, "lambda"));
      FX_TRY(need(s.delta, "delta"));
      FX_TRY(canonical_grid(p, s, w, st));
      if (n > 0) {
        // pred_s.w must hold lambda_j: gather (rho, lambda) then move lambda to w
        LAUNCH(k_gather, n, st, n, w.sidx, (const int32_t*)nullptr, (const float4*)s.pred, (const float4*)nullptr,
               (const float4*)s.lambda, w.pred_s[0], (float4*)nullptr, w.dl_s, (int32_t*)nullptr,
               (int32_t*)nullptr, (int32_t*)nullptr);
        LAUNCH(k_w_from_lambda, n, st, n, w.pred_s[0], (const float2*)w.dl_s);
        QLAUNCH(k_delta, VISIT_FULL, n, kQuadBelowDelta, kRowBelow, st, n, g, k, (const float4*)w.pred_s[0], (const float2*)w.dl_s, w.off,
                (float4*)nullptr, (const float4*)nullptr, (float4*)nullptr, w.sidx, (float4*)s.delta, L0,
                (const int32_t*)nullptr, Scenes{nullptr, 0});
      }
      break;
    case FX_PASS_APPLY_DELTA:
      FX_TRY(need(s.pred, "pred"));
      FX_TRY(need(s.delta, "delta"));
      if (n > 0) LAUNCH(k_apply_delta_c, n, st, n, (float4*)s.pred, (const float4*)s.delta);
      break;
    case FX_PASS_VELOCITY:
      FX_TRY(need(s.pred, "pred"));
      FX_TRY(need(s.pos, "pos"));
      FX_TRY(need(s.vel, "vel"));
      if (n > 0) LAUNCH(k_velocity_c, n, st, n, p.dt, (const float4*)s.pred, (const float4*)s.pos, (float4*)s.vel);
      break;
    case FX_PASS_XSPH:
      FX_TRY(need(s.pred, "pred"));
      FX_TRY(need(s.vel, "vel"));
      FX_TRY(need(s.lambda, "lambda"));
      FX_TRY(need(s.delta, "delta"));
      FX_TRY(canonical_grid(p, s, w, st));
      if (n > 0) {
        LAUNCH(k_gather_xsph, n, st, n, w.sidx, (const float4*)s.pred, (const float4*)s.vel,
               (const float4*)s.lambda, w.pred_s[0], w.vel_s);
        QLAUNCH(k_xsph, VISIT_FULL, n, kQuadBelowXsph, kRowBelow, st, n, p, g, k, col, (const float4*)w.pred_s[0],
                (const float4*)w.vel_s, w.off, w.sidx, (float4*)s.delta, (float4*)nullptr,
                (float4*)nullptr, (float4*)nullptr, (float4*)nullptr, L0, Scenes{nullptr, 0});
      }
      break;
    case FX_PASS_FINALIZE:
      FX_TRY(need(s.pos, "pos"));
      FX_TRY(need(s.vel, "vel"));
      FX_TRY(need(s.delta, "delta"));
      FX_TRY(need(s.pred, "pred"));
      if (n > 0)
        LAUNCH(k_finalize_c, n, st, p, col, (float4*)s.vel, (const float4*)s.delta,
               (const float4*)s.pred, (float4*)s.pos);
      break;
    case FX_PASS_NORMAL:
      FX_TRY(need(s.pred, "pred"));
      FX_TRY(need(s.lambda, "lambda"));
      FX_TRY(need(s.normal, "normal"));
      FX_TRY(canonical_grid(p, s, w, st));
      if (n > 0) {
        LAUNCH(k_gather, n, st, n, w.sidx, (const int32_t*)nullptr, (const float4*)s.pred, (const float4*)nullptr,
               (const float4*)s.lambda, w.pred_s[0], (float4*)nullptr, (float2*)nullptr,
               (int32_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr);
        NLAUNCH(k_normal<VISIT_FULL>, n, st, n, g, k, w.pred_s[0], w.off, w.sidx, (float4*)s.normal, L0, Scenes{nullptr, 0});
      }
      break;
    default:
      return fail(FX_EINVAL, "unknown pass id %d", id);
  }
  return check_launch("fx_pass");
}


// ---- batched scenes: host side ----------------------------------------------
struct BatchPlan {
  long long ntot, ctot;
  int iterations;
  bool tension;
  FxParams virt;  // n = ntot, n_cells = ctot: sizes the shared workspace
  size_t table_off, normal_off, pred_off, solver_off, total;
};

static int batch_plan(const FxScene* scenes, int count, BatchPlan& b) {
  if (!scenes || count < 1) return fail(FX_EINVAL, "batch: need at least one scene");
  b.ntot = 0;
  b.ctot = 0;
  b.tension = false;
  b.iterations = 0;  // the most any scene runs; the others carry their state through the rest
  for (int i = 0; i < count; ++i) {
    const FxParams& p = scenes[i].params;
    FxState dummy{};
    FX_TRY(validate(&p, &dummy, false));
    b.iterations = std::max(b.iterations, p.iterations);
    if (p.surface_tension > 0.0f) b.tension = true;
    b.ntot += p.n;
    b.ctot += p.n_cells;
  }
  if (b.ntot > kMaxParticles || b.ctot >= (1LL << 30))
    return fail(FX_ERANGE, "batch: %lld particles / %lld cells exceed the index range", b.ntot, b.ctot);
  b.virt = scenes[0].params;
  b.virt.n = (int)b.ntot;
  b.virt.n_cells = (int)b.ctot;
  size_t o = 0;
  b.table_off = o;
  o += align256((size_t)count * sizeof(SceneDev));
  b.normal_off = o;
  o += align256((size_t)b.ntot * 16 + 16);
  b.pred_off = o;
  o += align256((size_t)b.ntot * 16 + 16);
  b.solver_off = o;
  o += workspace_layout(b.virt).total;
  b.total = o;
  return FX_OK;
}

static std::vector<SceneDev> scene_table(const FxScene* scenes, int count, const uint8_t* colliders) {
  std::vector<SceneDev> t(count);
  int poff = 0, coff = 0;
  for (int i = 0; i < count; ++i) {
    const FxParams& p = scenes[i].params;
    SceneDev& d = t[i];
    d.p = p;
    d.k = make_consts(p);
    d.g = make_grid(p);
    const uint8_t* occ = (colliders && scenes[i].collider_offset >= 0) ? colliders + scenes[i].collider_offset : nullptr;
    d.col = make_collider(p, occ);
    d.poff = poff;
    d.n = p.n;
    d.coff = coff;
    d.nc = p.n_cells;
    poff += p.n;
    coff += p.n_cells;
  }
  return t;
}

static int batch_step(const FxScene* scenes, int count, float* pos, float* vel, const uint8_t* colliders,
                      void* work, size_t work_bytes, hipStream_t st, bool load) {
  BatchPlan b;
  FX_TRY(batch_plan(scenes, count, b));
  if (!work || work_bytes < b.total)
    return fail(FX_ENOMEM, "batch workspace too small: %zu < %zu bytes", work_bytes, b.total);
  if (b.ntot > 0 && (!pos || !vel)) return fail(FX_EINVAL, "batch: pos and vel are required");
  char* base = (char*)work;
  SceneDev* table = (SceneDev*)(base + b.table_off);
  float4* normal = (float4*)(base + b.normal_off);
  float4* pred = (float4*)(base + b.pred_off);
  void* wsp = base + b.solver_off;
  Ws w = bind_ws(workspace_layout(b.virt), wsp);
  // persistent launches draw their blocks from eight counters in the flags block
  Scenes S{table, count, w.flags + 32};
  const int n = (int)b.ntot, nc = (int)b.ctot;
  if (load) {
    // the per-scene table lives in the workspace so captured substeps read it
    // from device memory; this upload is the only host-to-device copy
    std::vector<SceneDev> host = scene_table(scenes, count, colliders);
    FX_HIP_CHECK(hipMemcpyAsync(table, host.data(), host.size() * sizeof(SceneDev), hipMemcpyHostToDevice, st));
    FX_HIP_CHECK(hipStreamSynchronize(st));
    if (n > 0) LAUNCH(k_b_load_state, n, st, n, (float4*)pos, (float4*)vel);
  }
  const Lists L0 = make_lists(w, make_consts(scenes[0].params), false);
  if (!load && b.tension) {
    FX_TRY(require_grid(workspace_layout(b.virt), wsp, n, nc, st, "batch: surface tension"));
    // the pick only guards the grid: the batch never fuses normals into XSPH
    const NormalPick pk{w.flags, nullptr, NormalKey{}, nc};
    BLAUNCH(k_normal, 8 + VISIT_FULL, n, st, n, Grid{}, Consts{}, w.pred_s[0], w.off, w.sidx, normal, L0, S, pk);
  }
  const int iters = b.iterations;
  const int cur0 = load ? 0 : (iters & 1);
  FX_TRY(fill(w.cnt, nc + 1, 0, st));
  if (n > 0)
    LAUNCH(k_b_predict_key, n, st, n, S, (const float4*)pos, (float4*)vel, (const float4*)normal, pred, w.key,
           w.cnt, w.rank, load ? 0 : 1);
  FxState fs{};
  FX_TRY(sort_from_counts(b.virt, fs, w, st));
  if (n > 0)
    LAUNCH(k_gather, n, st, n, w.sidx, w.key, (const float4*)pred, (const float4*)pos, (const float4*)nullptr,
           w.pred_s[cur0], w.pos_s, w.dl_s, (int32_t*)nullptr, w.skey, (int32_t*)nullptr);
  FX_LAUNCH(k_mark_grid, dim3(1), dim3(64), st, w.flags, n, nc);
  if (load || n == 0) return check_launch("batch");
  int cur = cur0;
  // the batch streams one lane per particle at any size, so the pair masks
  // serve it whenever the tuning keeps the default walk
  const bool masks = g_tuning.walk == 0;
  const PairMasks pm = pair_masks(w);
  for (int it = 0; it < iters; ++it) {
    // each scene writes its velocity on its own last iteration
    if (masks) {
      BLAUNCH_SHM(k_lambda, 8 + VISIT_FULL + kMaskRecord, n, row_words_lds(g_tuning.block), st, n, Grid{}, Consts{},
                  w.pred_s[cur], w.off, w.dl_s, w.sidx, (float4*)nullptr, L0, S, it, (const int32_t*)nullptr, pm);
      BLAUNCH_SHM(k_delta, 8 + VISIT_FULL + kMaskWalk + kLdsWords, n, row_words_lds(g_tuning.block), st, n, Grid{},
                  Consts{}, (const float4*)w.pred_s[cur], (const float2*)w.dl_s, w.off, w.pred_s[cur ^ 1],
                  (const float4*)w.pos_s, w.vel_s, w.sidx, (float4*)nullptr, L0, (const int32_t*)nullptr, S, it,
                  (const int32_t*)nullptr, pm);
    } else {
      BLAUNCH(k_lambda, 8 + VISIT_FULL, n, st, n, Grid{}, Consts{}, w.pred_s[cur], w.off, w.dl_s, w.sidx,
              (float4*)nullptr, L0, S, it);
      BLAUNCH(k_delta, 8 + VISIT_FULL, n, st, n, Grid{}, Consts{}, (const float4*)w.pred_s[cur],
              (const float2*)w.dl_s, w.off, w.pred_s[cur ^ 1], (const float4*)w.pos_s, w.vel_s, w.sidx,
              (float4*)nullptr, L0, (const int32_t*)nullptr, S, it);
    }
    cur ^= 1;
  }
  // no scene iterates: the batched delta launch as iteration 0 writes each
  // scene's velocity with its own dt (its copy into pred_s[1] is overwritten
  // by XSPH's correction below)
  if (iters == 0)
    NLAUNCH(k_delta<8 + VISIT_FULL>, n, st, n, Grid{}, Consts{}, (const float4*)w.pred_s[0], (const float2*)w.dl_s,
            w.off, w.pred_s[1], (const float4*)w.pos_s, w.vel_s, w.sidx, (float4*)nullptr, L0,
            (const int32_t*)nullptr, S, 0);
  // XSPH leaves dv by slot (in pred_s[1], free after the iterations) and the
  // finalize runs on its own, as in the plain substep: without the collider
  // search the neighbour pass holds fewer registers (91 -> 73 VGPRs)
  BLAUNCH(k_xsph, 8 + 16 + VISIT_FULL, n, st, n, FxParams{}, Grid{}, Consts{}, Collider{}, (const float4*)w.pred_s[0],
          (const float4*)w.vel_s, w.off, w.sidx, w.pred_s[1], (float4*)nullptr, (float4*)nullptr, (float4*)nullptr,
          (float4*)nullptr, L0, S);
  LAUNCH(k_finalize_s<8>, n, st, n, FxParams{}, Collider{}, (const float4*)w.pred_s[0], (const float4*)w.vel_s,
         (const float4*)w.pred_s[1], w.sidx, (float4*)nullptr, (float4*)pos, (float4*)vel, w.pos_s, (float4*)nullptr, S);
  return check_launch("batch");
}
}  // namespace fx

// ---- C ABI -------------------------------------------------------------------
using namespace fx;

extern "C" {

const char* fx_last_error(void) { return g_last_error.c_str(); }

int fx_abi_version(void) { return FX_ABI_VERSION; }

int fx_set_tuning(const FxTuning* t) {
  if (!t) return fail(FX_EINVAL, "fx_set_tuning: null");
  if (!(t->list_skin >= 0.0f) || t->list_skin > 1.0f)
    return fail(FX_EINVAL, "fx_set_tuning: list_skin %g outside [0, 1]", (double)t->list_skin);
  if (t->block != 64 && t->block != 128 && t->block != 256)
    return fail(FX_EINVAL, "fx_set_tuning: block %d not in {64, 128, 256}", t->block);
  if (t->persistent < 0 || t->persistent > 32)
    return fail(FX_EINVAL, "fx_set_tuning: persistent %d not in [0, 32]", t->persistent);
  if (t->walk < 0 || t->walk > 4) return fail(FX_EINVAL, "fx_set_tuning: walk %d not in {0, 1, 2, 3, 4}", t->walk);
  g_tuning.skin = t->list_skin;
  g_tuning.block = t->block;
  g_tuning.persistent = t->persistent;
  g_tuning.walk = t->walk;
  return FX_OK;
}

int fx_get_tuning(FxTuning* t) {
  if (!t) return fail(FX_EINVAL, "fx_get_tuning: null");
  t->list_skin = g_tuning.skin;
  t->block = g_tuning.block;
  t->persistent = g_tuning.persistent;
  t->walk = g_tuning.walk;
  return FX_OK;
}

int fx_profile(int enable) {
  std::lock_guard<std::mutex> lock(g_prof_mu);
  g_prof_on = enable != 0;
  g_prof_used = 0;
  return FX_OK;
}

int fx_profile_report(char* buf, size_t len) {
  if (!buf || len == 0) return fail(FX_EINVAL, "fx_profile_report: null buffer");
  std::lock_guard<std::mutex> lock(g_prof_mu);
  std::map<std::string, std::pair<long, double>> acc;
  for (size_t i = 0; i < g_prof_used; ++i) {
    ProfRec& r = g_prof_recs[i];
    if (hipEventSynchronize(r.b) != hipSuccess) return fail(FX_EHIP, "fx_profile_report: event sync");
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) continue;
    auto& e = acc[r.name];
    e.first += 1;
    e.second += ms;
  }
  std::string out;
  for (auto& kv : acc) {
    char line[256];
    snprintf(line, sizeof(line), "%s %ld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
    out += line;
  }
  g_prof_used = 0;
  snprintf(buf, len, "%s", out.c_str());
  return out.size() < len ? FX_OK : FX_ENOMEM;
}

size_t fx_workspace_bytes(const FxParams* p) {
  if (!p) return 0;
  return workspace_layout(*p).total;
}

int fx_pass(int pass_id, const FxParams* p, FxState* s, void* stream) {
  FX_TRY(validate(p, s, true));
  Ws w = bind_ws(workspace_layout(*p), s->work);
  return run_pass(pass_id, *p, *s, w, (hipStream_t)stream);
}

int fx_build_grid(const FxParams* p, FxState* s, void* stream) {
  FX_TRY(validate(p, s, true));
  if (p->n > 0 && (!s->pred || !s->pos)) return fail(FX_EINVAL, "build_grid needs pred and pos");
  Ws w = bind_ws(workspace_layout(*p), s->work);
  return build_grid(*p, *s, w, false, 0, (hipStream_t)stream);
}

// The normal pass of a call's first substep reads the previous grid
// (shaders/sph_normal.glsl:15-21): refuse a workspace that holds none.
static int check_previous_grid(const FxParams& p, const FxState& s, hipStream_t st, const char* what) {
  if (!(p.surface_tension > 0.0f) || p.n == 0) return FX_OK;
  return require_grid(workspace_layout(p), s.work, p.n, p.n_cells, st, what);
}

int fx_substep(const FxParams* p, FxState* s, void* stream) {
  FX_TRY(validate(p, s, true));
  if (p->n > 0 && (!s->pos || !s->vel)) return fail(FX_EINVAL, "substep needs pos and vel");
  if (p->n > 0 && !s->pred) return fail(FX_EINVAL, "substep needs a pred buffer");
  FX_TRY(check_previous_grid(*p, *s, (hipStream_t)stream, "fx_substep with surface tension"));
  Ws w = bind_ws(workspace_layout(*p), s->work);
  return substep(*p, *s, w, (hipStream_t)stream);
}

int fx_step(const FxParams* p, FxState* s, int substeps, void* stream) {
  if (substeps < 0) return fail(FX_EINVAL, "substeps < 0");
  if (substeps == 0) return FX_OK;
  FX_TRY(validate(p, s, true));
  if (p->n > 0 && (!s->pos || !s->vel)) return fail(FX_EINVAL, "substep needs pos and vel");
  if (p->n > 0 && !s->pred) return fail(FX_EINVAL, "substep needs a pred buffer");
  FX_TRY(check_previous_grid(*p, *s, (hipStream_t)stream, "fx_step with surface tension"));
  Ws w = bind_ws(workspace_layout(*p), s->work);
  const bool resident = resident_ok(*p, *s);
  for (int i = 0; i < substeps; ++i) {
    const bool last = i == substeps - 1;
    FX_TRY(substep(*p, *s, w, (hipStream_t)stream, last, resident && i > 0, resident && !last,
                   i + 1 == substeps - 1));
  }
  return FX_OK;
}

size_t fx_batch_workspace_bytes(const FxScene* scenes, int count) {
  BatchPlan b;
  if (batch_plan(scenes, count, b) != FX_OK) return 0;
  return b.total;
}

int fx_batch_load_state(const FxScene* scenes, int count, float* pos, float* vel, const uint8_t* colliders,
                        void* work, size_t work_bytes, void* stream) {
  return batch_step(scenes, count, pos, vel, colliders, work, work_bytes, (hipStream_t)stream, true);
}

int fx_batch_substep(const FxScene* scenes, int count, float* pos, float* vel, const uint8_t* colliders,
                     void* work, size_t work_bytes, void* stream) {
  return batch_step(scenes, count, pos, vel, colliders, work, work_bytes, (hipStream_t)stream, false);
}

int fx_load_state(const FxParams* p, FxState* s, void* stream) {
  FX_TRY(validate(p, s, true));
  if (p->n > 0 && (!s->pos || !s->vel || !s->pred))
    return fail(FX_EINVAL, "load_state needs pos, vel and pred");
  hipStream_t st = (hipStream_t)stream;
  if (p->n > 0)
    LAUNCH(k_load_state, p->n, st, p->n, (float4*)s->pos, (float4*)s->vel, (float4*)s->pred,
           (float4*)s->lambda);
  Ws w = bind_ws(workspace_layout(*p), s->work);
  return build_grid(*p, *s, w, false, 0, st);
}

}  // extern "C"
