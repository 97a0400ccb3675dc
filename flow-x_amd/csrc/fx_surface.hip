// Surface path on gfx950: density splat (shaders/surface_splat.glsl), GPU
// marching cubes (solver/marching_cubes.py) and analytic collider
// voxelization (collision/__init__.py geometry).
//
// Marching cubes runs as classify -> scan -> emit.  Every lattice edge with a
// sign change owns one vertex, numbered by its lower sample and axis, so
// neighbouring cubes share vertices exactly as the reference's
// (lower-corner sample, axis) key does (solver/marching_cubes.py:260).  The
// case table is derived at compile time from the cube's geometry with the
// reference's construction (solver/marching_cubes.py:60-173): oriented
// faces, per-face contour segments, loops walked in first-insertion order,
// reversed fans.  Vertex positions are computed in float64 without
// contraction and rounded once to float32.
#include <algorithm>
#include <mutex>

#ifndef FX_SPLAT_M
#define FX_SPLAT_M 2  // lattice samples per splat thread (1, 2, 4, 8 timed: 2 fastest)
#endif

#include "fx_common.h"

#pragma clang fp contract(off)

namespace fx {

// splat threads (groups of FX_SPLAT_M samples) below which each thread's
// samples get four lanes (a small lattice leaves most SIMDs idle); the two
// forms cross near 530k threads (lattice side ~102)
constexpr long long kSplatQuadBelow = 1 << 19;

// ---- marching-cubes tables, derived at compile time -------------------------
struct McTables {
  int8_t tri[256][16];  // edge triples, -1 terminated
  uint8_t ntri[256];
  uint16_t edge_mask[256];
  int8_t walk[12][3];   // lower corner, upper corner, axis
  int8_t corner[8][3];
};

constexpr int kCorners[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
constexpr int kEdges[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                               {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
constexpr int kFaces[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                              {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 2, 6, 5}};

constexpr int mc_edge_between(int a, int b) {
  for (int e = 0; e < 12; ++e)
    if ((kEdges[e][0] == a && kEdges[e][1] == b) || (kEdges[e][0] == b && kEdges[e][1] == a)) return e;
  return -1;
}

constexpr McTables build_mc_tables() {
  McTables t{};
  // faces wound counter-clockwise seen from outside (Newell normal . outward)
  int faces[6][4] = {};
  for (int f = 0; f < 6; ++f) {
    double nx = 0, ny = 0, nz = 0, cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < 4; ++i) {
      const int* a = kCorners[kFaces[f][i]];
      const int* b = kCorners[kFaces[f][(i + 1) % 4]];
      nx += double(a[1] - b[1]) * double(a[2] + b[2]);
      ny += double(a[2] - b[2]) * double(a[0] + b[0]);
      nz += double(a[0] - b[0]) * double(a[1] + b[1]);
      cx += a[0];
      cy += a[1];
      cz += a[2];
    }
    const double out = nx * (cx / 4.0 - 0.5) + ny * (cy / 4.0 - 0.5) + nz * (cz / 4.0 - 0.5);
    for (int c = 0; c < 4; ++c) faces[f][c] = out > 0.0 ? kFaces[f][c] : kFaces[f][3 - c];
  }
  for (int cs = 0; cs < 256; ++cs) {
    bool inside[8] = {};
    for (int c = 0; c < 8; ++c) inside[c] = (cs >> c) & 1;
    // successor map in insertion order (face order, then segment order)
    int skey[12] = {}, sval[12] = {}, ns = 0;
    for (int f = 0; f < 6; ++f) {
      bool opening[4] = {};
      int edge[4] = {}, count = 0;
      for (int i = 0; i < 4; ++i) {
        const int a = faces[f][i], b = faces[f][(i + 1) % 4];
        if (inside[a] && !inside[b]) {
          opening[count] = true;
          edge[count++] = mc_edge_between(a, b);
        } else if (inside[b] && !inside[a]) {
          opening[count] = false;
          edge[count++] = mc_edge_between(a, b);
        }
      }
      for (int i = 0; i < count; ++i) {
        if (!opening[i]) continue;
        for (int step = 1; step < count; ++step) {
          const int o = (i + step) % count;
          if (!opening[o]) {
            int found = -1;
            for (int q = 0; q < ns; ++q)
              if (skey[q] == edge[i]) found = q;
            if (found >= 0) {
              sval[found] = edge[o];
            } else {
              skey[ns] = edge[i];
              sval[ns++] = edge[o];
            }
            break;
          }
        }
      }
    }
    bool visited[12] = {};
    int nt = 0;
    for (int i = 0; i < 16; ++i) t.tri[cs][i] = -1;
    for (int q = 0; q < ns; ++q) {
      if (visited[skey[q]]) continue;
      int loop[12] = {}, nl = 0, e = skey[q];
      while (!visited[e]) {
        visited[e] = true;
        loop[nl++] = e;
        int nxt = -1;
        for (int r = 0; r < ns; ++r)
          if (skey[r] == e) nxt = sval[r];
        e = nxt;
      }
      for (int i = 1; i < nl - 1; ++i) {
        t.tri[cs][3 * nt] = (int8_t)loop[0];
        t.tri[cs][3 * nt + 1] = (int8_t)loop[i + 1];
        t.tri[cs][3 * nt + 2] = (int8_t)loop[i];
        ++nt;
      }
    }
    t.ntri[cs] = (uint8_t)nt;
    uint16_t mask = 0;
    for (int e = 0; e < 12; ++e)
      if (inside[kEdges[e][0]] != inside[kEdges[e][1]]) mask = (uint16_t)(mask | (1u << e));
    t.edge_mask[cs] = mask;
  }
  for (int e = 0; e < 12; ++e) {
    const int a = kEdges[e][0], b = kEdges[e][1];
    int axis = 0;
    while (kCorners[a][axis] == kCorners[b][axis]) ++axis;
    const bool a_lo = kCorners[a][axis] < kCorners[b][axis];
    t.walk[e][0] = (int8_t)(a_lo ? a : b);
    t.walk[e][1] = (int8_t)(a_lo ? b : a);
    t.walk[e][2] = (int8_t)axis;
  }
  for (int c = 0; c < 8; ++c)
    for (int a = 0; a < 3; ++a) t.corner[c][a] = (int8_t)kCorners[c][a];
  return t;
}

static constexpr McTables kMcHost = build_mc_tables();
__constant__ McTables kMc = build_mc_tables();

// ---- splat (shaders/surface_splat.glsl:32-97) ------------------------------
// Each sample (x fastest) gathers the 5x5x5 cells around its clamped cell as
// contiguous x-runs of sorted slots; pos_s holds the finalized positions in
// the grid's sorted order.  Cells that cannot hold a particle within the
// kernel radius are skipped: dcell[c] bounds how far cell c's particles lie
// outside its box (the grid is built from pre-constraint positions,
// shaders/surface_splat.glsl:8-9), so a cell whose box is farther than
// kr + dcell[c] contributes only zero terms.  The visited terms and their
// order are the reference's, so the sum is exact.
__device__ __forceinline__ float axis_gap(float p, float lo, float hi) {
  return fmaxf(fmaxf(lo - p, p - hi), 0.0f);
}

// shaders/surface_splat.glsl:32-97 for M consecutive samples of one lattice
// row per thread: they share the stencil rows (same y, z), so each candidate
// is loaded once and tested against all M.  Each sample still visits its own stencil cells (within 2
// cells of its own, shaders/surface_splat.glsl:62-82) in the reference's
// order - cells by increasing x, slots ascending, at most max_scan per cell -
// so every sum is the one-sample kernel's, bit for bit.
// The M samples' density terms of one candidate (zero where it does not
// contribute), for the four-lane form of the splat.
template <int M>
struct SplatTerm {
  float t[M];
  template <int J>
  __device__ __forceinline__ SplatTerm lane() const {
    SplatTerm o;
#pragma unroll
    for (int u = 0; u < M; ++u) o.t[u] = quad_lane<J>(t[u]);
    return o;
  }
};

// MQ = M samples per thread, + kSplatQuad for four lanes per thread
constexpr int kSplatQuad = 16;
template <int MQ>
__global__ __launch_bounds__(256) void k_splat_m(FxSurface sf, Grid g, float mass, float rho0, float h2,
                                                 float coef, int max_scan, Collider col,
                                                 const float4* __restrict__ pos_s,
                                                 const int32_t* __restrict__ off,
                                                 const int32_t* __restrict__ dcell_bits, float kr,
                                                 float* __restrict__ field, const int32_t* __restrict__ hdr,
                                                 int n) {
  constexpr int M = MQ % kSplatQuad;
  constexpr bool QUAD = MQ >= kSplatQuad;  // four lanes per thread's samples
  const int nx = sf.dims[0], ny = sf.dims[1], nz = sf.dims[2];
  const int groups = (nx + M - 1) / M;
  const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long q0 = QUAD ? gt >> 2 : gt;
  const int jq = (int)(gt & 3);
  if (q0 >= (long long)groups * ny * nz) return;
  const int gx0 = (int)(q0 % groups) * M, gy = (int)((q0 / groups) % ny), gz = (int)(q0 / ((long long)groups * ny));
  const int m = min(M, nx - gx0);
  const float py = sf.lo[1] + (float)gy * sf.spacing, pz = sf.lo[2] + (float)gz * sf.spacing;
  float px[M], acc[M];
  int bxs[M];
  int by = 0, bz = 0;
#pragma unroll
  for (int u = 0; u < M; ++u) {
    px[u] = sf.lo[0] + (float)min(gx0 + u, nx - 1) * sf.spacing;
    acc[u] = 0.0f;
    cell_coord(g, px[u], py, pz, bxs[u], by, bz);
  }
  const float pad = 1e-3f * kr;
  const int wlo = max(bxs[0] - 2, 0), whi = min(bxs[M - 1] + 2, g.cx - 1);
  // a workspace without this grid (the header check of fx_common.h) leaves
  // the field zero rather than read offsets it does not hold
  const int zend = header_grid_ok(hdr, n, g.nc) ? 2 : -3;
  for (int dz = -2; dz <= zend; ++dz) {
    const int z = bz + dz;
    if (z < 0 || z >= g.cz) continue;
    const float zlo = g.loz + (float)z * g.cell;
    const float gzd = axis_gap(pz, zlo, zlo + g.cell);
    for (int dy = -2; dy <= 2; ++dy) {
      const int y = by + dy;
      if (y < 0 || y >= g.cy) continue;
      const float ylo = g.loy + (float)y * g.cell;
      const float gyd = axis_gap(py, ylo, ylo + g.cell);
      const float gyz = gzd * gzd + gyd * gyd;
      const int row = (z * g.cy + y) * g.cx;
      if (off[row + wlo] == off[row + whi + 1]) continue;
      // per sample: the cells of its stencil that can hold a particle within
      // kr (kr + the cell's displacement bound), as a slot range [slo, shi)
      int xa[M], xb[M];
#pragma unroll
      for (int u = 0; u < M; ++u) xa[u] = g.cx, xb[u] = -1;
      for (int x = wlo; x <= whi; ++x) {
        const float xlo = g.lox + (float)x * g.cell;
        const float R = kr + __int_as_float(dcell_bits[row + x]) + pad;
#pragma unroll
        for (int u = 0; u < M; ++u) {
          const float gxd = axis_gap(px[u], xlo, xlo + g.cell);
          if (u < m && x >= bxs[u] - 2 && x <= bxs[u] + 2 && gxd * gxd + gyz < R * R) {
            xa[u] = min(xa[u], x);
            xb[u] = max(xb[u], x);
          }
        }
      }
      int ua = g.cx, ub = -1;
#pragma unroll
      for (int u = 0; u < M; ++u) ua = min(ua, xa[u]), ub = max(ub, xb[u]);
      if (ub < ua) continue;
      const int A = off[row + ua], E = off[row + ub + 1];
      int slo[M], shi[M];
#pragma unroll
      for (int u = 0; u < M; ++u) {
        slo[u] = xb[u] < xa[u] ? 0 : off[row + xa[u]];
        shi[u] = xb[u] < xa[u] ? 0 : off[row + xb[u] + 1];
      }
      auto body = [&](int q, const float4& pj) {
#pragma unroll
        for (int u = 0; u < M; ++u) {
          const float dx = px[u] - pj.x, dy2 = py - pj.y, dz2 = pz - pj.z;
          const float r2 = dist2_exact(dx, dy2, dz2);
          if (q >= slo[u] && q < shi[u] && r2 < h2) acc[u] = poly6_accum_exact(acc[u], mass, coef, h2 - r2);
        }
      };
      auto ld = [&](int q) { return pos_s[q]; };
      if constexpr (QUAD) {
        // the quad splits each run's candidates and folds the terms back in
        // slot order (quad_run): the same sums as `body`, bit for bit
        auto tm = [&](int q, const float4& pj, bool valid) {
          SplatTerm<M> o;
#pragma unroll
          for (int u = 0; u < M; ++u) {
            const float dx = px[u] - pj.x, dy2 = py - pj.y, dz2 = pz - pj.z;
            const float r2 = dist2_exact(dx, dy2, dz2);
            o.t[u] = (valid && q >= slo[u] && q < shi[u] && r2 < h2) ? poly6_term_exact(mass, coef, h2 - r2) : 0.0f;
          }
          return o;
        };
        auto fold = [&](const SplatTerm<M>& tt) {
#pragma unroll
          for (int u = 0; u < M; ++u) acc[u] = add_exact(acc[u], tt.t[u]);
        };
        if (E - A <= max_scan) {
          quad_run<4>(A, E, jq, ld, tm, fold);
        } else {
          for (int x = ua; x <= ub; ++x) {
            const int cs = off[row + x], ce = min(off[row + x + 1], cs + max_scan);
            quad_run<4>(cs, ce, jq, ld, tm, fold);
          }
        }
      } else if (E - A <= max_scan) {
        scan_run(A, E, ld, body);  // no cell of the run reaches the per-cell cap
      } else {
        for (int x = ua; x <= ub; ++x) {
          const int cs = off[row + x], ce = min(off[row + x + 1], cs + max_scan);
          scan_run(cs, ce, ld, body);
        }
      }
    }
  }
  if (QUAD && jq != 0) return;
#pragma unroll
  for (int u = 0; u < M; ++u) {
    if (u >= m) break;
    float value = acc[u] / fmaxf(rho0, 1e-6f);
    if (col.occ) {
      int cx, cy, cz;
      collider_coord(col, px[u], py, pz, cx, cy, cz);
      if (collider_occupied(col, cx, cy, cz)) value = 0.0f;
    }
    field[((long long)gz * ny + gy) * nx + gx0 + u] = value;
  }
}

// Gathers the finalized positions into sorted order and bounds, per cell,
// how far its particles lie outside the cell's box (atomic max into the
// cell's slot: contention only among one cell's particles).
__global__ void k_gather_pos(int n, const int32_t* __restrict__ sidx, const int32_t* __restrict__ skey,
                             Grid g, const float4* __restrict__ pos, float4* __restrict__ pos_s,
                             int32_t* __restrict__ dcell_bits, int32_t* __restrict__ hdr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (!header_grid_ok(hdr, n, g.nc)) {
    if (t == 0) hdr[kHdrError] = 1;
    return;
  }
  float d = 0.0f;
  if (t < n) {
    const float4 p = pos[sidx[t]];
    pos_s[t] = p;
    const int c = skey[t];
    const int x = c % g.cx, y = (c / g.cx) % g.cy, z = c / (g.cx * g.cy);
    const float xlo = g.lox + (float)x * g.cell, ylo = g.loy + (float)y * g.cell, zlo = g.loz + (float)z * g.cell;
    const float ex = axis_gap(p.x, xlo, xlo + g.cell), ey = axis_gap(p.y, ylo, ylo + g.cell),
                ez = axis_gap(p.z, zlo, zlo + g.cell);
    d = sqrtf(ex * ex + ey * ey + ez * ez) * 1.000001f;
    if (d > 0.0f && c >= 0 && c < g.nc) atomicMax(&dcell_bits[c], __float_as_int(d));
  }
}

__global__ void k_fill_cells(int32_t* __restrict__ a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// ---- marching cubes -------------------------------------------------------
constexpr int kMcBlock = 256;

struct McLattice {
  int nx, ny, nz;
  long long S;
  float iso;
};

__device__ __forceinline__ int wave_block_scan(int v, int* lds, int* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  if (lane == 63) lds[wave] = v;
  __syncthreads();
  int add = 0;
  for (int w = 0; w < wave; ++w) add += lds[w];
  *total = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  return v + add;
}

// Lattice coordinates of sample q.  fx_mc_extract caps the lattice at 2^28
// samples, so 32-bit arithmetic holds (two 32-bit divisions instead of three
// 64-bit ones, which dominated these kernels' instruction count).
__device__ __forceinline__ void mc_ijk(const McLattice& L, int q, int& i, int& j, int& k) {
  const unsigned row = (unsigned)q / (unsigned)L.nx;
  const unsigned kk = row / (unsigned)L.ny;
  i = q - (int)(row * (unsigned)L.nx);
  j = (int)(row - kk * (unsigned)L.ny);
  k = (int)kk;
}

// Per sample: the sign-change flags of its +x / +y / +z lattice edges (the
// vertices it owns) and, if it is a cube's lower corner, the cube's case.
// Each of the cube's eight corner values is loaded once; corner c of the
// case index sits at kCorners[c] (solver/marching_cubes.py:25-34).
__global__ __launch_bounds__(kMcBlock) void k_mc_classify(McLattice L, const float* __restrict__ f,
                                                          uint16_t* __restrict__ code,
                                                          int32_t* __restrict__ bv,
                                                          int32_t* __restrict__ bt,
                                                          uint8_t* __restrict__ cases) {
  __shared__ int lds[4];
  const int q = blockIdx.x * kMcBlock + threadIdx.x;
  int nv = 0, nt = 0;
  if (q < (int)L.S) {
    const int nx = L.nx, ny = L.ny, nz = L.nz, nxy = nx * ny;
    int i, j, k;
    mc_ijk(L, q, i, j, k);
    const bool hx = i + 1 < nx, hy = j + 1 < ny, hz = k + 1 < nz;
    // v[x + 2 y + 4 z] = the field at corner (x, y, z) of the cube at q
    bool v[8];
    v[0] = f[q] >= L.iso;
    v[1] = hx && f[q + 1] >= L.iso;
    v[2] = hy && f[q + nx] >= L.iso;
    v[4] = hz && f[q + nxy] >= L.iso;
    int flags = 0;
    if (hx && v[1] != v[0]) flags |= 1;
    if (hy && v[2] != v[0]) flags |= 2;
    if (hz && v[4] != v[0]) flags |= 4;
    int cs = 0;
    if (hx && hy && hz) {
      v[3] = f[q + 1 + nx] >= L.iso;
      v[5] = f[q + 1 + nxy] >= L.iso;
      v[6] = f[q + nx + nxy] >= L.iso;
      v[7] = f[q + 1 + nx + nxy] >= L.iso;
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (v[kCorners[c][0] + 2 * kCorners[c][1] + 4 * kCorners[c][2]]) cs |= 1 << c;
      nt = kMc.ntri[cs];
      if (cases) cases[(k * (ny - 1) + j) * (nx - 1) + i] = (uint8_t)cs;
    }
    nv = __popc(flags);
    code[q] = (uint16_t)(flags | (cs << 3));
  }
  int tv, tt;
  wave_block_scan(nv, lds, &tv);
  wave_block_scan(nt, lds, &tt);
  if (threadIdx.x == 0) {
    bv[blockIdx.x] = tv;
    bt[blockIdx.x] = tt;
  }
}

// Per sample: global vertex / triangle bases, and its vertices (float64
// interpolation of solver/marching_cubes.py:263-275, rounded to float32).
__global__ __launch_bounds__(kMcBlock) void k_mc_vertices(McLattice L, const float* __restrict__ f,
                                                          const uint16_t* __restrict__ code,
                                                          const int32_t* __restrict__ bv_off,
                                                          const int32_t* __restrict__ bt_off,
                                                          uint32_t* __restrict__ vbase,
                                                          uint32_t* __restrict__ tbase,
                                                          double ox, double oy, double oz,
                                                          double spacing,
                                                          float* __restrict__ verts,
                                                          long long vcap) {
  __shared__ int lds[4];
  const int q = blockIdx.x * kMcBlock + threadIdx.x;
  int flags = 0, cs = 0;
  if (q < (int)L.S) {
    const int c = code[q];
    flags = c & 7;
    cs = c >> 3;
  }
  const int nv = __popc(flags), nt = kMc.ntri[cs];
  int tv, tt;
  const int iv = wave_block_scan(nv, lds, &tv) - nv + bv_off[blockIdx.x];
  const int it = wave_block_scan(nt, lds, &tt) - nt + bt_off[blockIdx.x];
  if (q >= (int)L.S) return;
  vbase[q] = (uint32_t)iv;
  tbase[q] = (uint32_t)it;
  if (!flags) return;
  const int nx = L.nx, nxy = L.nx * L.ny;
  int i, j, k;
  mc_ijk(L, q, i, j, k);
  const double lo_v = (double)f[q];
  const double iso = (double)L.iso;
  int r = 0;
  for (int axis = 0; axis < 3; ++axis) {
    if (!(flags & (1 << axis))) continue;
    const int stride = axis == 0 ? 1 : (axis == 1 ? nx : nxy);
    const double hi_v = (double)f[q + stride];
    const double delta = hi_v - lo_v;
    const double t = delta == 0.0 ? 0.5 : (iso - lo_v) / delta;
    double pt[3] = {ox + (double)i * spacing, oy + (double)j * spacing, oz + (double)k * spacing};
    pt[axis] += t * spacing;
    const long long v = (long long)iv + r;
    if (v < vcap) {
      verts[3 * v] = (float)pt[0];
      verts[3 * v + 1] = (float)pt[1];
      verts[3 * v + 2] = (float)pt[2];
    }
    ++r;
  }
}

// Per cube: its triangles from the case table, vertex indices looked up by
// (lower-corner sample, axis).
__global__ __launch_bounds__(kMcBlock) void k_mc_triangles(McLattice L,
                                                           const uint16_t* __restrict__ code,
                                                           const uint32_t* __restrict__ vbase,
                                                           const uint32_t* __restrict__ tbase,
                                                           int32_t* __restrict__ tris, long long tcap) {
  const int q = blockIdx.x * kMcBlock + threadIdx.x;
  if (q >= (int)L.S) return;
  const int cs = code[q] >> 3;
  const int nt = kMc.ntri[cs];
  if (!nt) return;
  const int nx = L.nx, nxy = L.nx * L.ny;
  const long long base = tbase[q];
  for (int m = 0; m < nt; ++m) {
    int idx[3];
    for (int e3 = 0; e3 < 3; ++e3) {
      const int e = kMc.tri[cs][3 * m + e3];
      const int lc = kMc.walk[e][0], axis = kMc.walk[e][2];
      const int q2 = q + kMc.corner[lc][0] + kMc.corner[lc][1] * nx + kMc.corner[lc][2] * nxy;
      const int fl = code[q2] & 7;
      idx[e3] = (int)vbase[q2] + __popc(fl & ((1 << axis) - 1));
    }
    const long long tt = base + m;
    if (tt < tcap) {
      tris[3 * tt] = idx[0];
      tris[3 * tt + 1] = idx[1];
      tris[3 * tt + 2] = idx[2];
    }
  }
}

__global__ void k_mc_counts(const int32_t* bv_off, const int32_t* bt_off, int nb, int64_t* counts) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    counts[0] = bv_off[nb];
    counts[1] = bt_off[nb];
  }
}

struct McWs {
  size_t code, vbase, tbase, bv, bt, bv_off, bt_off, part, total;
};

static size_t al(size_t x) { return (x + 255) & ~(size_t)255; }

static McWs mc_layout(long long S) {
  McWs w;
  const long long nb = (S + kMcBlock - 1) / kMcBlock + 1;
  size_t o = 0;
  w.code = o;   o += al((size_t)S * 2 + 2);
  w.vbase = o;  o += al((size_t)S * 4 + 4);
  w.tbase = o;  o += al((size_t)S * 4 + 4);
  w.bv = o;     o += al((size_t)nb * 4);
  w.bt = o;     o += al((size_t)nb * 4);
  w.bv_off = o; o += al((size_t)(nb + 1) * 4);
  w.bt_off = o; o += al((size_t)(nb + 1) * 4);
  w.part = o;   o += al(scan_partials(nb) * 4 + 4);
  w.total = o;
  return w;
}

// ---- voxelizers --------------------------------------------------------------
// Occupancy of voxel centres origin + (i + 0.5) * voxel (float64 as the
// reference computes them, collision/__init__.py:353-366).
struct VoxShape {
  int kind;  // 0 box, 1 sphere, 2 vertical cylinder
  double a[3], b[3];
  double r, height;
};

__global__ void k_voxelize(int nx, int ny, int nz, double ox, double oy, double oz, double voxel,
                           VoxShape sh, uint8_t* __restrict__ grid, int accumulate) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)nx * ny * nz;
  if (q >= total) return;
  const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / ((long long)nx * ny));
  const double x = ox + (i + 0.5) * voxel, y = oy + (j + 0.5) * voxel, z = oz + (k + 0.5) * voxel;
  bool in = false;
  if (sh.kind == 0) {
    in = x > sh.a[0] && x < sh.b[0] && y > sh.a[1] && y < sh.b[1] && z > sh.a[2] && z < sh.b[2];
  } else if (sh.kind == 1) {
    const double dx = x - sh.a[0], dy = y - sh.a[1], dz = z - sh.a[2];
    in = dx * dx + dy * dy + dz * dz < sh.r * sh.r;
  } else {
    const double dx = x - sh.a[0], dy = y - sh.a[1];
    in = dx * dx + dy * dy < sh.r * sh.r && z > sh.a[2] && z < sh.a[2] + sh.height;
  }
  if (accumulate) {
    if (in) grid[q] = 1;
  } else {
    grid[q] = in ? 1 : 0;
  }
}

static int voxelize(const int32_t dims[3], const float origin[3], float voxel, const VoxShape& sh,
                    uint8_t* grid, int accumulate, hipStream_t st) {
  if (!dims || !origin || !grid) return fail(FX_EINVAL, "voxelize: null argument");
  if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1) return fail(FX_EINVAL, "voxelize: dims must be >= 1");
  if (!(voxel > 0.0f)) return fail(FX_EINVAL, "voxelize: voxel size must be > 0");
  const long long total = (long long)dims[0] * dims[1] * dims[2];
  hipLaunchKernelGGL(k_voxelize, dim3(grid_for(total)), dim3(kBlock), 0, st, dims[0], dims[1], dims[2],
                     (double)origin[0], (double)origin[1], (double)origin[2], (double)voxel, sh, grid,
                     accumulate);
  return check_launch("voxelize");
}

// ---- mesh colliders (collision/__init__.py:263-273, :304-367) --------------
//
// The reference casts a ray from every voxel centre inside the mesh's world
// bounds along the world +z axis (the local-space image of it, for a BVH in
// local space), counts hits - restarting 1e-4 (local units) beyond each hit,
// at most 64 times - and marks the voxel when the count is odd.  Here one
// thread owns a column (i, j): it intersects the vertical line through the
// column's voxel centres with every triangle (world space, in chunks staged in
// LDS), keeps the hit heights sorted, and replays the reference's hit chain
// upward from each voxel centre: the first hit at or above it, then each
// next hit more than ray_eps (world units) beyond the previous one.
constexpr int kMeshHits = 128;

__global__ void k_fill_u8(uint8_t* __restrict__ a, long long n, uint8_t v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

constexpr int kTriChunk = 512;

__global__ __launch_bounds__(64) void k_voxelize_mesh(int nx, int ny, int i0, int i1, int j0, int j1, int k0,
                                                      int k1, double ox, double oy, double oz, double voxel,
                                                      const float* __restrict__ verts,
                                                      const int32_t* __restrict__ tris, int ntris, int nverts,
                                                      double eps, uint8_t* __restrict__ grid, int accumulate,
                                                      int32_t* __restrict__ overflow) {
  __shared__ float tri_lds[kTriChunk * 9];
  const int ci = i1 - i0, cols = ci * (j1 - j0);
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = c < cols;
  const int i = i0 + (active ? c % ci : 0), j = j0 + (active ? c / ci : 0);
  const double x = ox + (i + 0.5) * voxel, y = oy + (j + 0.5) * voxel;
  double hits[kMeshHits];
  int nh = 0;
  bool over = false;
  for (int base = 0; base < ntris; base += kTriChunk) {
    const int m = min(kTriChunk, ntris - base);
    __syncthreads();
    for (int q = threadIdx.x; q < m * 3; q += blockDim.x) {
      int v = tris[(size_t)(base + q / 3) * 3 + q % 3];
      v = min(max(v, 0), nverts - 1);
      tri_lds[q * 3 + 0] = verts[(size_t)v * 3 + 0];
      tri_lds[q * 3 + 1] = verts[(size_t)v * 3 + 1];
      tri_lds[q * 3 + 2] = verts[(size_t)v * 3 + 2];
    }
    __syncthreads();
    if (!active) continue;
    for (int t = 0; t < m; ++t) {
      const float* T = tri_lds + t * 9;
      const double ax = T[0], ay = T[1], az = T[2], bx = T[3], by = T[4], bz = T[5], cx = T[6], cy = T[7],
                   cz = T[8];
      // 2-D edge functions of (x, y) against the projected triangle
      const double area = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
      if (area == 0.0) continue;  // edge-on: the vertical line grazes it
      const double w0 = (bx - x) * (cy - y) - (by - y) * (cx - x);
      const double w1 = (cx - x) * (ay - y) - (cy - y) * (ax - x);
      const double w2 = (ax - x) * (by - y) - (ay - y) * (bx - x);
      const bool inside = area > 0.0 ? (w0 >= 0.0 && w1 >= 0.0 && w2 >= 0.0) : (w0 <= 0.0 && w1 <= 0.0 && w2 <= 0.0);
      if (!inside) continue;
      const double z = (w0 * az + w1 * bz + w2 * cz) / area;
      if (nh == kMeshHits) {
        over = true;
        continue;
      }
      int p = nh++;  // insertion keeps the hits sorted
      while (p > 0 && hits[p - 1] > z) {
        hits[p] = hits[p - 1];
        --p;
      }
      hits[p] = z;
    }
  }
  if (!active) return;
  if (over && overflow) atomicOr(overflow, 1);
  int first = 0;
  for (int k = k0; k < k1; ++k) {
    const double z = oz + (k + 0.5) * voxel;
    while (first < nh && hits[first] < z) ++first;
    int count = 0;
    double last = 0.0;
    for (int h = first; h < nh && count < 64; ++h) {
      if (count > 0 && hits[h] < last + eps) continue;
      last = hits[h];
      ++count;
    }
    const size_t q = ((size_t)k * ny + j) * nx + i;
    if (count & 1)
      grid[q] = 1;
    else if (!accumulate)
      grid[q] = 0;
  }
}

// ---- surface hand-off (solver/surface.py:278-301 _rebuild_mesh) -------------
//
// bmesh refuses a face over a vertex set that already has one (any winding),
// so the reference keeps the first of each duplicate triangle marching cubes
// emits where the surface pinches.  Here each triangle's sorted vertex
// triple goes into an open-addressing hash table whose slot records the
// smallest triangle index with that triple (atomicMin), and an ordered
// compaction keeps exactly those triangles.
constexpr unsigned long long kDedupEmpty = ~0ull;

__device__ __forceinline__ unsigned long long tri_key(const int32_t* t) {
  unsigned long long a = (unsigned)t[0], b = (unsigned)t[1], c = (unsigned)t[2], x;
  if (a > b) x = a, a = b, b = x;
  if (b > c) x = b, b = c, c = x;
  if (a > b) x = a, a = b, b = x;
  return a | (b << 21) | (c << 42);
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

__global__ void k_dedup_clear(unsigned long long* __restrict__ keys, int32_t* __restrict__ first, long long cap) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cap) return;
  keys[i] = kDedupEmpty;
  first[i] = 0x7fffffff;
}

__global__ void k_dedup_insert(const int32_t* __restrict__ tris, int n, unsigned long long* __restrict__ keys,
                               int32_t* __restrict__ first, unsigned long long mask) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const unsigned long long k = tri_key(tris + (size_t)t * 3);
  unsigned long long slot = mix64(k) & mask;
  for (;;) {  // the table holds at least 2n slots, so a free or matching slot exists
    const unsigned long long prev = atomicCAS(&keys[slot], kDedupEmpty, k);
    if (prev == kDedupEmpty || prev == k) {
      atomicMin(&first[slot], t);
      return;
    }
    slot = (slot + 1) & mask;
  }
}

__global__ void k_dedup_flag(const int32_t* __restrict__ tris, int n, const unsigned long long* __restrict__ keys,
                             const int32_t* __restrict__ first, unsigned long long mask, int32_t* __restrict__ keep) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const unsigned long long k = tri_key(tris + (size_t)t * 3);
  unsigned long long slot = mix64(k) & mask;
  while (keys[slot] != k) slot = (slot + 1) & mask;  // inserted above, so it is found
  keep[t] = first[slot] == t ? 1 : 0;
}

__global__ void k_dedup_compact(const int32_t* __restrict__ tris, int n, const int32_t* __restrict__ keep,
                                const int32_t* __restrict__ off, int32_t* __restrict__ out,
                                int32_t* __restrict__ count) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) count[0] = off[n];
  if (t >= n || !keep[t]) return;
  const size_t o = (size_t)off[t] * 3;
  out[o] = tris[(size_t)t * 3];
  out[o + 1] = tris[(size_t)t * 3 + 1];
  out[o + 2] = tris[(size_t)t * 3 + 2];
}

struct DedupWs {
  size_t keys, first, keep, off, part, total;
  unsigned long long cap;
};

static DedupWs dedup_layout(long long n) {
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  DedupWs w;
  unsigned long long cap = 64;
  while (cap < 2ull * (unsigned long long)std::max(n, 1LL)) cap <<= 1;
  w.cap = cap;
  size_t o = 0;
  w.keys = o;  o += al(cap * 8);
  w.first = o; o += al(cap * 4);
  w.keep = o;  o += al((size_t)n * 4 + 4);
  w.off = o;   o += al((size_t)n * 4 + 8);
  w.part = o;  o += al(scan_partials(n + 1) * 4 + 4);
  w.total = o;
  return w;
}

// Marching cubes of `field` with its scratch (mc_layout) at `b`: classify,
// two scans, vertices, triangles, counts; asynchronous on st.
static int mc_extract_run(const char* name, const float* field, const int32_t dims[3], const double origin[3],
                          double spacing, float iso, float* verts, int64_t vcap, int32_t* tris, int64_t tcap,
                          int64_t* counts, uint8_t* cases, char* b, size_t work_bytes, hipStream_t st) {
  if (!field || !dims || !origin || !counts) return fail(FX_EINVAL, "%s: null argument", name);
  if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1) return fail(FX_EINVAL, "%s: dims must be >= 1", name);
  if (vcap < 0 || tcap < 0) return fail(FX_EINVAL, "%s: negative capacity", name);
  if ((vcap > 0 && !verts) || (tcap > 0 && !tris)) return fail(FX_EINVAL, "%s: null output", name);
  const long long S = (long long)dims[0] * dims[1] * dims[2];
  if (S > (1LL << 28)) return fail(FX_ERANGE, "%s: %lld samples exceed 2^28", name, S);
  const McWs wl = mc_layout(S);
  if (work_bytes < wl.total) return fail(FX_ENOMEM, "%s: workspace %zu < %zu", name, work_bytes, wl.total);
  uint16_t* code = (uint16_t*)(b + wl.code);
  uint32_t* vbase = (uint32_t*)(b + wl.vbase);
  uint32_t* tbase = (uint32_t*)(b + wl.tbase);
  int32_t* bv = (int32_t*)(b + wl.bv);
  int32_t* bt = (int32_t*)(b + wl.bt);
  int32_t* bv_off = (int32_t*)(b + wl.bv_off);
  int32_t* bt_off = (int32_t*)(b + wl.bt_off);
  int32_t* part = (int32_t*)(b + wl.part);
  const long long nb = (S + kMcBlock - 1) / kMcBlock;
  McLattice L{dims[0], dims[1], dims[2], S, iso};
  FX_LAUNCH(k_mc_classify, dim3((unsigned)nb), dim3(kMcBlock), st, L, field, code, bv, bt, cases);
  FX_TRY(exclusive_scan_i32(bv, bv_off, nb, part, st));
  FX_TRY(exclusive_scan_i32(bt, bt_off, nb, part, st));
  FX_LAUNCH(k_mc_vertices, dim3((unsigned)nb), dim3(kMcBlock), st, L, field,
                     (const uint16_t*)code, (const int32_t*)bv_off, (const int32_t*)bt_off, vbase, tbase,
                     origin[0], origin[1], origin[2], spacing, verts, (long long)vcap);
  FX_LAUNCH(k_mc_triangles, dim3((unsigned)nb), dim3(kMcBlock), st, L, (const uint16_t*)code,
                     (const uint32_t*)vbase, (const uint32_t*)tbase, tris, (long long)tcap);
  hipLaunchKernelGGL(k_mc_counts, dim3(1), dim3(64), 0, st, (const int32_t*)bv_off, (const int32_t*)bt_off,
                     (int)nb, counts);
  return check_launch(name);
}

// ---- marching cubes without a workspace ----------------------------------------
// fx_mc_extract_static runs fx_mc_extract's kernels with their scratch in a
// fixed device array instead of a caller workspace, for callers that cannot
// allocate (code under graph capture, loops that forbid hipMalloc).  The array
// holds the scratch of kMcStaticSamples samples - twice the reference's
// surface cap of 2^20 samples (solver/surface.py:49).  Each GPU the library
// runs on holds its own copy until the process ends; every call on a GPU
// shares that copy, so calls that overlap on one GPU race.  Nothing in it
// needs to start at zero: every kernel writes what the next one reads.
constexpr long long kMcStaticSamples = 1LL << 21;
constexpr size_t kMcStaticBytes = 22u << 20;  // >= mc_layout(kMcStaticSamples).total (checked)
static __device__ __align__(256) uint8_t scratch[kMcStaticBytes];

// The array's device address, looked up once per device with that device
// current (the array's name in host code is only a stand-in for it).
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[64] = {};

static int mc_static_scratch(char** out) {
  int dev = 0;
  FX_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(FX_EINVAL, "fx_mc_extract_static: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    if (mc_layout(kMcStaticSamples).total > kMcStaticBytes)
      return fail(FX_ENOMEM, "fx_mc_extract_static: fixed scratch smaller than its sample limit needs");
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(FX_EHIP, "fx_mc_extract_static: hipGetSymbolAddress: %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = (char*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return FX_OK;
}

}  // namespace fx

using namespace fx;

extern "C" {

int fx_splat(const FxParams* p, const FxSurface* surf, const FxState* s, float* field, void* stream) {
  if (!p || !surf || !s || !field) return fail(FX_EINVAL, "fx_splat: null argument");
  if (surf->dims[0] < 1 || surf->dims[1] < 1 || surf->dims[2] < 1)
    return fail(FX_EINVAL, "fx_splat: surface dims must be >= 1");
  if (!s->work || s->work_bytes < workspace_layout(*p).total)
    return fail(FX_ENOMEM, "fx_splat: workspace missing or too small");
  if (p->n > 0 && !s->pos) return fail(FX_EINVAL, "fx_splat: pos is NULL");
  hipStream_t st = (hipStream_t)stream;
  const Workspace wl = workspace_layout(*p);
  FX_TRY(require_grid(wl, s->work, p->n, p->n_cells, st, "fx_splat (it gathers through the last grid)"));
  char* b = (char*)s->work;
  int32_t* hdr = (int32_t*)(b + wl.flags);
  const int32_t* sidx = (const int32_t*)(b + wl.sidx);
  const int32_t* off = (const int32_t*)(b + wl.off);
  const int32_t* skey = (const int32_t*)(b + wl.skey);
  int32_t* dmax = (int32_t*)(b + wl.sdcell);
  float4* pos_s = (float4*)(b + wl.pos_s);
  FX_LAUNCH(k_fill_cells, dim3(grid_for(p->n_cells)), dim3(kBlock), st, dmax, p->n_cells);
  if (p->n > 0)
    FX_LAUNCH(k_gather_pos, dim3(grid_for(p->n)), dim3(kBlock), st, p->n, sidx, skey, make_grid(*p),
              (const float4*)s->pos, pos_s, dmax, hdr);
  const long long S = (long long)surf->dims[0] * surf->dims[1] * surf->dims[2];
  const float kr = surf->kernel_radius;
  const float coef = poly6_coef(kr);
  Collider col = make_collider(*p, s->collider);
  constexpr int M = FX_SPLAT_M;
  const long long threads = (long long)((surf->dims[0] + M - 1) / M) * surf->dims[1] * surf->dims[2];
  // a small lattice leaves SIMDs idle: four lanes per thread's samples
  if (threads < kSplatQuadBelow)
    FX_LAUNCH(k_splat_m<M + kSplatQuad>, dim3(grid_for(4 * threads)), dim3(256), st, *surf, make_grid(*p), p->mass,
              p->rest_density, kr * kr, coef, p->max_cell_scan > 0 ? p->max_cell_scan : 512, col,
              (const float4*)pos_s, off, (const int32_t*)dmax, kr, field, (const int32_t*)hdr, p->n);
  else
    FX_LAUNCH(k_splat_m<M>, dim3(grid_for(threads)), dim3(256), st, *surf, make_grid(*p), p->mass,
              p->rest_density, kr * kr, coef, p->max_cell_scan > 0 ? p->max_cell_scan : 512, col,
              (const float4*)pos_s, off, (const int32_t*)dmax, kr, field, (const int32_t*)hdr, p->n);
  (void)S;
  return check_launch("fx_splat");
}

size_t fx_mc_workspace_bytes(const int32_t dims[3]) {
  if (!dims || dims[0] < 1 || dims[1] < 1 || dims[2] < 1) return 0;
  return mc_layout((long long)dims[0] * dims[1] * dims[2]).total;
}

int fx_mc_extract(const float* field, const int32_t dims[3], const double origin[3], double spacing,
                  float iso, float* verts, int64_t vcap, int32_t* tris, int64_t tcap, int64_t* counts,
                  uint8_t* cases, void* work, size_t work_bytes, void* stream) {
  if (!work) return fail(FX_EINVAL, "fx_mc_extract: null workspace");
  return mc_extract_run("fx_mc_extract", field, dims, origin, spacing, iso, verts, vcap, tris, tcap, counts, cases,
                        (char*)work, work_bytes, (hipStream_t)stream);
}

int fx_mc_extract_static(const float* field, const int32_t dims[3], const double origin[3], double spacing,
                         float iso, float* verts, int64_t vcap, int32_t* tris, int64_t tcap, int64_t* counts,
                         uint8_t* cases, void* stream) {
  if (!dims) return fail(FX_EINVAL, "fx_mc_extract_static: null dims");
  if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1)
    return fail(FX_EINVAL, "fx_mc_extract_static: dims must be >= 1");
  const long long S = (long long)dims[0] * dims[1] * dims[2];
  if (S > kMcStaticSamples)
    return fail(FX_ERANGE,
                "fx_mc_extract_static: %lld samples exceed its fixed scratch (%lld samples); use fx_mc_extract "
                "with a workspace",
                S, kMcStaticSamples);
  char* ptr = nullptr;
  FX_TRY(mc_static_scratch(&ptr));
  return mc_extract_run("fx_mc_extract_static", field, dims, origin, spacing, iso, verts, vcap, tris, tcap, counts,
                        cases, ptr, kMcStaticBytes, (hipStream_t)stream);
}

size_t fx_mc_dedup_workspace_bytes(int64_t ntris) { return ntris < 0 ? 0 : dedup_layout(ntris).total; }

int fx_mc_dedup(const int32_t* tris, int64_t ntris, int64_t nverts, int32_t* out, int32_t* count, void* work,
                size_t work_bytes, void* stream) {
  if (ntris < 0 || !count || (ntris > 0 && (!tris || !out))) return fail(FX_EINVAL, "fx_mc_dedup: bad arguments");
  if (ntris > (1LL << 30)) return fail(FX_ERANGE, "fx_mc_dedup: %lld triangles exceed 2^30", (long long)ntris);
  if (nverts > (1LL << 21)) return fail(FX_ERANGE, "fx_mc_dedup: %lld vertices exceed the 21-bit key", (long long)nverts);
  const DedupWs w = dedup_layout(ntris);
  if (!work || work_bytes < w.total) return fail(FX_ENOMEM, "fx_mc_dedup: workspace %zu < %zu", work_bytes, w.total);
  hipStream_t st = (hipStream_t)stream;
  char* b = (char*)work;
  unsigned long long* keys = (unsigned long long*)(b + w.keys);
  int32_t *first = (int32_t*)(b + w.first), *keep = (int32_t*)(b + w.keep), *off = (int32_t*)(b + w.off),
          *part = (int32_t*)(b + w.part);
  const int n = (int)ntris;
  FX_LAUNCH(k_dedup_clear, dim3(grid_for((long long)w.cap)), dim3(kBlock), st, keys, first, (long long)w.cap);
  if (n > 0) {
    FX_LAUNCH(k_dedup_insert, dim3(grid_for(n)), dim3(kBlock), st, tris, n, keys, first, w.cap - 1);
    FX_LAUNCH(k_dedup_flag, dim3(grid_for(n)), dim3(kBlock), st, tris, n, (const unsigned long long*)keys,
              (const int32_t*)first, w.cap - 1, keep);
  }
  FX_TRY(exclusive_scan_i32(keep, off, n, part, st));
  FX_LAUNCH(k_dedup_compact, dim3(grid_for(std::max(n, 1))), dim3(kBlock), st, tris, n, (const int32_t*)keep,
            (const int32_t*)off, out, count);
  return check_launch("fx_mc_dedup");
}

// Host copy of the derived case table (edge triples, -1 terminated) and the
// per-case sign-change edge mask, for checking the derivation.
int fx_mc_tables(int8_t* tri_out, uint16_t* edge_mask_out) {
  if (!tri_out || !edge_mask_out) return fail(FX_EINVAL, "fx_mc_tables: null argument");
  for (int c = 0; c < 256; ++c) {
    for (int i = 0; i < 16; ++i) tri_out[c * 16 + i] = kMcHost.tri[c][i];
    edge_mask_out[c] = kMcHost.edge_mask[c];
  }
  return FX_OK;
}

int fx_voxelize_boxes(const int32_t dims[3], const float origin[3], float voxel, const float* boxes,
                      int k, uint8_t* grid, int accumulate, void* stream) {
  if (k < 0 || (k > 0 && !boxes)) return fail(FX_EINVAL, "fx_voxelize_boxes: bad box list");
  hipStream_t st = (hipStream_t)stream;
  if (k == 0 && !accumulate) {
    VoxShape none{0, {1, 1, 1}, {0, 0, 0}, 0, 0};  // empty box: clears the grid
    return voxelize(dims, origin, voxel, none, grid, 0, st);
  }
  for (int b = 0; b < k; ++b) {
    VoxShape sh{0,
                {boxes[6 * b], boxes[6 * b + 1], boxes[6 * b + 2]},
                {boxes[6 * b + 3], boxes[6 * b + 4], boxes[6 * b + 5]},
                0.0,
                0.0};
    FX_TRY(voxelize(dims, origin, voxel, sh, grid, (b > 0 || accumulate) ? 1 : 0, st));
  }
  return FX_OK;
}

int fx_voxelize_sphere(const int32_t dims[3], const float origin[3], float voxel, const float center[3],
                       float radius, uint8_t* grid, int accumulate, void* stream) {
  if (!center) return fail(FX_EINVAL, "fx_voxelize_sphere: null centre");
  VoxShape sh{1, {center[0], center[1], center[2]}, {0, 0, 0}, radius, 0.0};
  return voxelize(dims, origin, voxel, sh, grid, accumulate, (hipStream_t)stream);
}

int fx_voxelize_cylinder(const int32_t dims[3], const float origin[3], float voxel, const float base[3],
                         float radius, float height, uint8_t* grid, int accumulate, void* stream) {
  if (!base) return fail(FX_EINVAL, "fx_voxelize_cylinder: null base");
  VoxShape sh{2, {base[0], base[1], base[2]}, {0, 0, 0}, radius, height};
  return voxelize(dims, origin, voxel, sh, grid, accumulate, (hipStream_t)stream);
}

int fx_voxelize_mesh(const int32_t dims[3], const float origin[3], float voxel, const float* verts, int nverts,
                     const int32_t* tris, int ntris, const float bounds_lo[3], const float bounds_hi[3],
                     float ray_eps, uint8_t* grid, int accumulate, int32_t* overflow, void* stream) {
  if (!dims || !origin || !bounds_lo || !bounds_hi || !grid) return fail(FX_EINVAL, "fx_voxelize_mesh: null argument");
  if (dims[0] < 1 || dims[1] < 1 || dims[2] < 1) return fail(FX_EINVAL, "fx_voxelize_mesh: dims must be >= 1");
  if (!(voxel > 0.0f)) return fail(FX_EINVAL, "fx_voxelize_mesh: voxel size must be > 0");
  if (ntris < 0 || nverts < 0 || (ntris > 0 && (!verts || !tris || nverts == 0)))
    return fail(FX_EINVAL, "fx_voxelize_mesh: bad mesh arrays");
  if (!(ray_eps >= 0.0f)) return fail(FX_EINVAL, "fx_voxelize_mesh: ray_eps must be >= 0");
  hipStream_t st = (hipStream_t)stream;
  // collision/__init__.py:290-294 _index_range
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = std::max(0, (int)std::floor(((double)bounds_lo[a] - origin[a]) / voxel));
    hi[a] = std::min(dims[a], (int)std::ceil(((double)bounds_hi[a] - origin[a]) / voxel));
  }
  if (!accumulate) {
    const long long total = (long long)dims[0] * dims[1] * dims[2];
    hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(total)), dim3(kBlock), 0, st, grid, total, (uint8_t)0);
  }
  if (ntris == 0 || lo[0] >= hi[0] || lo[1] >= hi[1] || lo[2] >= hi[2]) return check_launch("fx_voxelize_mesh");
  const long long cols = (long long)(hi[0] - lo[0]) * (hi[1] - lo[1]);
  hipLaunchKernelGGL(k_voxelize_mesh, dim3(grid_for(cols, 64)), dim3(64), 0, st, dims[0], dims[1], lo[0], hi[0],
                     lo[1], hi[1], lo[2], hi[2], (double)origin[0], (double)origin[1], (double)origin[2],
                     (double)voxel, verts, tris, ntris, nverts, (double)ray_eps, grid, 1, overflow);
  return check_launch("fx_voxelize_mesh");
}

}  // extern "C"
