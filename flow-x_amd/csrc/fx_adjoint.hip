// Differentiable PBF substep and its hand-written adjoint (gfx950), in float
// or double.  No reference counterpart: the reference has no gradients.
//
// The forward is the reference substep (shaders/sph_predict.glsl,
// sph_lambda.glsl, sph_delta.glsl + sph_apply_delta.glsl, sph_velocity.glsl,
// sph_xsph.glsl, sph_finalize.glsl) with two definitions that make it
// differentiable:
//   * the neighbour structure of a substep is recorded once, after predict:
//     every pair within h + skin of the predicted positions among the 27
//     grid cells around the particle's cell, each particle's list sorted by
//     index.  Every pass of the substep sums over these lists
//     (pairs beyond h contribute zero, and the kernels are C1 at r = h).
//   * the collider's voxel frame moves with its translation T: occupancy is
//     sampled at p - T and the free-voxel centres are shifted by T, so a
//     pushed-out particle's position depends smoothly on T through the
//     chosen centre (the inside test and the chosen voxel are recorded
//     decisions, piecewise constant).
// surface_tension and s_corr are not differentiated (the adjoint runs require
// both to be zero).  Backward recomputes the substep from its checkpoint
// (x, v), keeping every Jacobi iterate, then runs the reverse passes.
// Reductions of parameter gradients use a fixed-order tree (deterministic).
#include "fx_common.h"

#pragma clang fp contract(off)

namespace fx {

template <class T>
struct V3 {
  T x, y, z, w;
};

template <class T>
__device__ __forceinline__ V3<T> v3(T x, T y, T z) {
  return V3<T>{x, y, z, T(0)};
}
template <class T>
__device__ __forceinline__ V3<T> operator+(V3<T> a, V3<T> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T>
__device__ __forceinline__ V3<T> operator-(V3<T> a, V3<T> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T>
__device__ __forceinline__ V3<T> operator*(T s, V3<T> a) { return v3(s * a.x, s * a.y, s * a.z); }
template <class T>
__device__ __forceinline__ T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

constexpr int kMaxNbr = 192;  // list capacity: kMaxNbr pairs per particle on average

template <class T>
struct AdjC {
  T h, h2, mass, rho0, eps, poly6, spiky, visc, dt, g, radius, damping, K;  // K = delta scale / rho0
  T lo[3], hi[3];
  T tr[3];  // collider translation
  T voxel;
  int kx, ky, kz;
  const uint8_t* occ;
  int iters;
  T list_r2;
  // recorded discrete decisions, 2 ints per particle: flags (bit 0 pushed
  // out of the collider, bit 1 velocity reflected, bits 2-7 wall clamp per
  // axis (1 low, 2 high), bit 8 speed clamped) and the chosen free voxel
  int32_t* dec;
  int dec_mode;  // 0 decide, 1 decide and record, 2 replay the record
};

enum { DEC_NONE = 0, DEC_RECORD = 1, DEC_REPLAY = 2 };

template <class T>
static AdjC<T> make_adjc(const FxParams& p, double gravity, const double tr[3], const uint8_t* col, double skin,
                         int32_t* dec = nullptr, int dec_mode = DEC_NONE) {
  AdjC<T> c;
  c.dec = dec;
  c.dec_mode = dec ? dec_mode : DEC_NONE;
  const double h = p.h, pi = 3.14159265358979323846;
  c.h = (T)h;
  c.h2 = (T)(h * h);
  c.mass = (T)p.mass;
  c.rho0 = (T)p.rest_density;
  c.eps = (T)p.relaxation;
  c.poly6 = (T)(315.0 / (64.0 * pi * std::pow(h, 9.0)));
  c.spiky = (T)(-45.0 / (pi * std::pow(h, 6.0)));
  c.visc = (T)p.viscosity;
  c.dt = (T)p.dt;
  c.g = (T)gravity;
  c.radius = (T)p.particle_radius;
  c.damping = (T)p.boundary_damping;
  const double scale = (p.flags & FX_FLAG_MASS_WEIGHTED_DELTA) ? p.mass : 1.0;
  c.K = (T)(scale / p.rest_density);
  for (int a = 0; a < 3; ++a) {
    c.lo[a] = (T)p.lo[a];
    c.hi[a] = (T)p.hi[a];
    c.tr[a] = (T)(tr ? tr[a] : 0.0);
  }
  c.voxel = (T)p.collider_voxel;
  c.kx = p.collider_dims[0];
  c.ky = p.collider_dims[1];
  c.kz = p.collider_dims[2];
  c.occ = (p.collider_voxel > 0.0f) ? col : nullptr;
  c.iters = p.iterations;
  const double R = h * (1.0 + skin);
  c.list_r2 = (T)(R * R);
  return c;
}

// ---- kernel functions and their derivatives -----------------------------
// W(r) = poly6 (h^2 - r^2)^3;   dW/dp_i = -6 poly6 (h^2 - r^2)^2 d
// phi(d) = s(r) d / r with s(r) = spiky (h - r)^2   (1e-12 < r^2 < h^2)
// J_phi(d) = f I + (f'/r) d d^T,  f = s / r,  f' = (s' r - s) / r^2
template <class T>
__device__ __forceinline__ void phi_and_jac(const AdjC<T>& c, V3<T> d, T r2, V3<T>& phi, T& f, T& fpr) {
  if (!(r2 < c.h2) || !(r2 > T(1e-12))) {
    phi = v3(T(0), T(0), T(0));
    f = fpr = T(0);
    return;
  }
  const T r = sqrt(r2);
  const T hr = c.h - r;
  const T s = c.spiky * hr * hr;
  const T sp = T(-2) * c.spiky * hr;
  f = s / r;
  const T fp = (sp * r - s) / r2;
  fpr = fp / r;
  phi = f * d;
}

template <class T>
__device__ __forceinline__ V3<T> jac_mul(T f, T fpr, V3<T> d, V3<T> a) {
  // (f I + fpr d d^T) a
  return f * a + (fpr * dot(d, a)) * d;
}

// ---- grid + recorded neighbour lists --------------------------------------

__global__ void k_adj_key(int n, Grid g, const float4* __restrict__ p, int32_t* __restrict__ key,
                          int32_t* __restrict__ cnt, int32_t* __restrict__ rank) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  const float4 q = p[i];
  cell_coord(g, q.x, q.y, q.z, x, y, z);
  const int k = cell_index(g, x, y, z);
  key[i] = k;
  rank[i] = atomicAdd(&cnt[k], 1);
}

__global__ void k_adj_scatter(int n, const int32_t* key, const int32_t* rank, const int32_t* off, int32_t* sidx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  sidx[off[key[i]] + rank[i]] = i;
}

template <class T>
__global__ void k_adj_to_f4(int n, const V3<T>* __restrict__ p, float4* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3<T> q = p[i];
  out[i] = make_float4((float)q.x, (float)q.y, (float)q.z, 1.0f);
}

// Each particle's neighbours within the list radius among the 27 cells
// around its cell, in CSR form: a count pass, an exclusive scan into noff
// (n + 1 entries), a fill pass, then k_adj_sort_lists.  Lists sorted by
// index are deterministic (the cell runs are in atomic arrival order) and
// keep the lanes of a wave on nearby neighbours at every list position.
// The pair total is capped at n * kMaxNbr; beyond it the offsets are clamped
// and the overflow flag is raised.
template <class T, bool FILL>
__global__ void k_adj_lists(int n, Grid g, AdjC<T> c, const V3<T>* __restrict__ p, const float4* __restrict__ pf,
                            const int32_t* __restrict__ off, const int32_t* __restrict__ sidx,
                            int32_t* __restrict__ count, const int32_t* __restrict__ noff,
                            int32_t* __restrict__ nbr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3<T> pi = p[i];
  const float4 qi = pf[i];
  int bx, by, bz;
  cell_coord(g, qi.x, qi.y, qi.z, bx, by, bz);
  int m = 0;
  int32_t* L = nullptr;
  int room = 0;
  if (FILL) {
    L = nbr + noff[i];
    room = noff[i + 1] - noff[i];
  }
  // the 27 cells as 9 runs of 3 adjacent cells (x-1..x+1 are contiguous in
  // sidx), each walked 4 slots at a time: the slots' particle indices, then
  // their positions, are loaded before any is tested
  const int x0 = max(bx - 1, 0), x1 = min(bx + 1, g.cx - 1);
  for (int r = 0; r < 9; ++r) {
    const int z = bz + r / 3 - 1, y = by + r % 3 - 1;
    if (y < 0 || z < 0 || y >= g.cy || z >= g.cz) continue;
    const int c0 = cell_index(g, x0, y, z);
    const int e = off[c0 + (x1 - x0) + 1];
    for (int q = off[c0]; q < e; q += 4) {
      int j[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) j[u] = sidx[min(q + u, e - 1)];
      V3<T> pj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) pj[u] = p[j[u]];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (q + u >= e || j[u] == i) continue;
        const V3<T> d = pi - pj[u];
        if (dot(d, d) < c.list_r2) {
          if (FILL && m < room) L[m] = j[u];
          ++m;
        }
      }
    }
  }
  if (!FILL) count[i] = m;
}

// Sort every list ascending, one wave per list.  Up to kWaveSortMax entries
// sit in registers (element e = 64 r + lane in v[r], padded with INT_MAX) and
// pass through a bitonic network: steps of stride >= 64 exchange registers
// within a lane, shorter strides exchange with lane ^ stride by shuffle.
// Longer lists are queued for sort_queued_runs.
constexpr int kWaveSortMax = 256;

__device__ __forceinline__ void cmpx(int& a, int& b, bool asc) {
  const int lo = min(a, b), hi = max(a, b);
  a = asc ? lo : hi;
  b = asc ? hi : lo;
}

__global__ __launch_bounds__(256) void k_adj_sort_lists(int n, const int32_t* __restrict__ noff,
                                                        int32_t* __restrict__ nbr, int32_t* __restrict__ queue,
                                                        int32_t* __restrict__ queue_count) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;  // whole waves
  const int s = noff[i], m = noff[i + 1] - s;
  if (m <= 1) return;
  if (m > kWaveSortMax) {
    if (lane == 0) queue[atomicAdd(queue_count, 1)] = i;
    return;
  }
  const int R = m <= 64 ? 1 : (m <= 128 ? 2 : 4);  // registers in use
  const int P = 64 * R;
  int v[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int e = 64 * r + lane;
    v[r] = e < m ? nbr[s + e] : INT_MAX;
  }
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j == 128) {
        cmpx(v[0], v[2], (lane & k) == 0);
        cmpx(v[1], v[3], ((64 + lane) & k) == 0);
      } else if (j == 64) {
        cmpx(v[0], v[1], (lane & k) == 0);
        if (R == 4) cmpx(v[2], v[3], ((128 + lane) & k) == 0);
      } else {
        const bool lower = (lane & j) == 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (r < R) {
            const int o = __shfl_xor(v[r], j);
            const bool asc = ((64 * r + lane) & k) == 0;
            v[r] = (asc == lower) ? min(v[r], o) : max(v[r], o);
          }
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int e = 64 * r + lane;
    if (e < m) nbr[s + e] = v[r];
  }
}

// Clamp the list offsets to the capacity; flag an overflow.
__global__ void k_adj_cap(int n, long long cap, int32_t* __restrict__ noff, int32_t* __restrict__ overflow) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > n) return;
  const int v = noff[i];
  if (v > cap) {
    noff[i] = (int32_t)cap;
    if (i == n) atomicOr(overflow, 1);
  }
}

// ---- list walks ---------------------------------------------------------------
// f(j, p[j]) for the entries j of list L[0, m) in order, kAdjBatch at a time:
// the batch's indices, then its positions, are all loaded before any entry
// is used (one entry at a time, every position load waits on its own index
// load).  Empty slots of the last batch re-load its first entry so the loads
// stay unconditional.  The order of the calls, and so of every sum, is the
// list's.
constexpr int kAdjBatch = 4;

template <int B, class T, class F>
__device__ __forceinline__ void walk_list(const int32_t* __restrict__ L, int m, const V3<T>* __restrict__ P,
                                          F&& f) {
  for (int k = 0; k < m; k += B) {
    int j[B];
#pragma unroll
    for (int u = 0; u < B; ++u) j[u] = L[min(k + u, m - 1)];
    V3<T> q[B];
#pragma unroll
    for (int u = 0; u < B; ++u) q[u] = P[j[u]];
#pragma unroll
    for (int u = 0; u < B; ++u)
      if (k + u < m) f(j[u], q[u]);
  }
}

// ---- forward passes ---------------------------------------------------------

// shaders/sph_predict.glsl:23-57 (no cohesion); also stores the clamped
// velocity for the backward pass.
template <class T>
__global__ void k_adj_predict(int n, AdjC<T> c, const V3<T>* __restrict__ x, const V3<T>* __restrict__ v,
                              V3<T>* __restrict__ p0, V3<T>* __restrict__ v1) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3<T> u = v[i];
  u.z = u.z + c.g * c.dt;
  const T vmax = T(0.4) * c.h / (c.dt > T(1e-6) ? c.dt : T(1e-6));
  const T speed = sqrt(dot(u, u));
  const bool clamp = c.dec_mode == DEC_REPLAY ? ((c.dec[2 * i] >> 8) & 1) != 0 : speed > vmax;
  if (c.dec_mode == DEC_RECORD) {
    c.dec[2 * i] = clamp ? 256 : 0;
    c.dec[2 * i + 1] = -1;
  }
  if (clamp) u = (vmax / speed) * u;
  v1[i] = u;
  p0[i] = x[i] + c.dt * u;
}

// shaders/sph_lambda.glsl:17-77 over the recorded lists; stores rho, C, D,
// lambda and the gradient sum G for the backward pass.
template <class T>
__global__ void k_adj_lambda(int n, AdjC<T> c, const V3<T>* __restrict__ p, const int32_t* __restrict__ nbr,
                             const int32_t* __restrict__ noff, T* __restrict__ rho, T* __restrict__ lam,
                             T* __restrict__ Dv, V3<T>* __restrict__ G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3<T> pi = p[i];
  T dens = c.poly6 * c.h2 * c.h2 * c.h2;  // self term W(0)
  V3<T> gs = v3(T(0), T(0), T(0));
  T gn2 = T(0);
  const int32_t* L = nbr + noff[i];
  walk_list<kAdjBatch>(L, noff[i + 1] - noff[i], p, [&](const int j, const V3<T>& pj) {
    const V3<T> d = pi - pj;
    const T r2 = dot(d, d);
    if (r2 < c.h2) {
      const T w = c.h2 - r2;
      dens = dens + c.poly6 * w * w * w;
      V3<T> ph;
      T f, fpr;
      phi_and_jac(c, d, r2, ph, f, fpr);
      const V3<T> gij = c.mass * ph;
      gs = gs + gij;
      gn2 = gn2 + dot(gij, gij);
    }
  });
  dens = c.mass * dens;
  const T rr = dens > T(1e-6) ? dens : T(1e-6);
  const T C = rr / c.rho0 - T(1);
  T D = (dot(gs, gs) + gn2) / (c.rho0 * c.rho0) + c.eps;
  D = D > T(1e-9) ? D : T(1e-9);
  rho[i] = rr;
  lam[i] = -C / D;
  Dv[i] = D;
  G[i] = gs;
}

// shaders/sph_delta.glsl:22-85 + sph_apply_delta.glsl (p_next = p + dp)
template <class T>
__global__ void k_adj_delta(int n, AdjC<T> c, const V3<T>* __restrict__ p, const T* __restrict__ lam,
                            const int32_t* __restrict__ nbr, const int32_t* __restrict__ noff,
                            V3<T>* __restrict__ pn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3<T> pi = p[i];
  const T li = lam[i];
  V3<T> acc = v3(T(0), T(0), T(0));
  const int32_t* L = nbr + noff[i];
  walk_list<kAdjBatch>(L, noff[i + 1] - noff[i], p, [&](const int j, const V3<T>& pj) {
    const V3<T> d = pi - pj;
    const T r2 = dot(d, d);
    V3<T> ph;
    T f, fpr;
    phi_and_jac(c, d, r2, ph, f, fpr);
    acc = acc + (li + lam[j]) * ph;
  });
  pn[i] = pi + c.K * acc;
}

// Finalize with the collider frame shifted by T (shaders/sph_finalize.glsl).
template <class T>
struct FinalRec {
  int inside;      // pushed out of a collider voxel
  V3<T> centre;    // chosen free-voxel centre (moves with T)
  int damped;      // the normal velocity was reflected
  int clamped[3];  // per axis: 0 none, -1 low wall, +1 high wall
};

template <class T>
__device__ __forceinline__ bool adj_occupied(const AdjC<T>& c, int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= c.kx || y >= c.ky || z >= c.kz) return false;
  return c.occ[((size_t)z * c.ky + y) * c.kx + x] != 0;
}

template <class T>
__device__ void adj_finalize(const AdjC<T>& c, int i, V3<T>& p, V3<T>& v, FinalRec<T>& rec) {
  rec.inside = 0;
  rec.damped = 0;
  rec.clamped[0] = rec.clamped[1] = rec.clamped[2] = 0;
  if (c.dec_mode == DEC_REPLAY) {
    const int flags = c.dec[2 * i];
    if ((flags & 1) && c.occ) {
      const int idx = c.dec[2 * i + 1];
      const int x = idx % c.kx, y = (idx / c.kx) % c.ky, z = idx / (c.kx * c.ky);
      const V3<T> ctr = v3(c.lo[0] + ((T)x + T(0.5)) * c.voxel + c.tr[0], c.lo[1] + ((T)y + T(0.5)) * c.voxel + c.tr[1],
                           c.lo[2] + ((T)z + T(0.5)) * c.voxel + c.tr[2]);
      const V3<T> e = ctr - p;
      const T dist = sqrt(dot(e, e));
      const V3<T> u = dist > T(1e-6) ? (T(1) / dist) * e : v3(T(0), T(0), T(1));
      rec.inside = 1;
      rec.centre = ctr;
      p = p + (dist + c.radius) * u;
      if (flags & 2) {
        const T vn = dot(v, u);
        v = v - (vn * (T(1) + c.damping)) * u;
        rec.damped = 1;
      }
    }
    T* pp[3] = {&p.x, &p.y, &p.z};
    T* vv[3] = {&v.x, &v.y, &v.z};
    for (int a = 0; a < 3; ++a) {
      const int code = (flags >> (2 + 2 * a)) & 3;
      if (code == 1) {
        *pp[a] = c.lo[a] + c.radius;
        *vv[a] = -*vv[a] * c.damping;
        rec.clamped[a] = -1;
      } else if (code == 2) {
        *pp[a] = c.hi[a] - c.radius;
        *vv[a] = -*vv[a] * c.damping;
        rec.clamped[a] = 1;
      }
    }
    return;
  }
  int vox = -1;
  if (c.occ) {
    const int cx = (int)floor((p.x - c.tr[0] - c.lo[0]) / c.voxel);
    const int cy = (int)floor((p.y - c.tr[1] - c.lo[1]) / c.voxel);
    const int cz = (int)floor((p.z - c.tr[2] - c.lo[2]) / c.voxel);
    if (adj_occupied(c, cx, cy, cz)) {
      T best = T(1e30);
      bool found = false;
      V3<T> nf = v3(T(0), T(0), T(0));
      for (int dz = -2; dz <= 2; ++dz)
        for (int dy = -2; dy <= 2; ++dy)
          for (int dx = -2; dx <= 2; ++dx) {
            const int x = cx + dx, y = cy + dy, z = cz + dz;
            if (x < 0 || y < 0 || z < 0 || x >= c.kx || y >= c.ky || z >= c.kz) continue;
            if (adj_occupied(c, x, y, z)) continue;
            const V3<T> ctr = v3(c.lo[0] + ((T)x + T(0.5)) * c.voxel + c.tr[0],
                                 c.lo[1] + ((T)y + T(0.5)) * c.voxel + c.tr[1],
                                 c.lo[2] + ((T)z + T(0.5)) * c.voxel + c.tr[2]);
            const V3<T> e = ctr - p;
            const T d2 = dot(e, e);
            if (d2 < best) {
              best = d2;
              nf = ctr;
              found = true;
              vox = (z * c.ky + y) * c.kx + x;
            }
          }
      if (found) {
        const T dist = sqrt(best);
        V3<T> u = dist > T(1e-6) ? (T(1) / dist) * (nf - p) : v3(T(0), T(0), T(1));
        rec.inside = 1;
        rec.centre = nf;
        p = p + (dist + c.radius) * u;
        const T vn = dot(v, u);
        if (vn < T(0)) {
          v = v - (vn * (T(1) + c.damping)) * u;
          rec.damped = 1;
        }
      }
    }
  }
  T* pp[3] = {&p.x, &p.y, &p.z};
  T* vv[3] = {&v.x, &v.y, &v.z};
  for (int a = 0; a < 3; ++a) {
    const T lo = c.lo[a] + c.radius, hi = c.hi[a] - c.radius;
    if (*pp[a] < lo) {
      *pp[a] = lo;
      *vv[a] = -*vv[a] * c.damping;
      rec.clamped[a] = -1;
    } else if (*pp[a] > hi) {
      *pp[a] = hi;
      *vv[a] = -*vv[a] * c.damping;
      rec.clamped[a] = 1;
    }
  }
  if (c.dec_mode == DEC_RECORD) {
    int flags = (c.dec[2 * i] & 256) | (rec.inside ? 1 : 0) | (rec.damped ? 2 : 0);
    for (int a = 0; a < 3; ++a) flags |= (rec.clamped[a] == -1 ? 1 : rec.clamped[a] == 1 ? 2 : 0) << (2 + 2 * a);
    c.dec[2 * i] = flags;
    c.dec[2 * i + 1] = rec.inside ? vox : -1;
  }
}

// velocity + XSPH (density from the last lambda pass) + finalize
template <class T>
__global__ void k_adj_velocity(int n, AdjC<T> c, const V3<T>* __restrict__ pI, const V3<T>* __restrict__ x,
                               V3<T>* __restrict__ v2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T dtc = c.dt > T(1e-6) ? c.dt : T(1e-6);
  v2[i] = (T(1) / dtc) * (pI[i] - x[i]);
}

template <class T>
__global__ void k_adj_xsph_final(int n, AdjC<T> c, const V3<T>* __restrict__ pI, const V3<T>* __restrict__ v2,
                                 const T* __restrict__ rho, const int32_t* __restrict__ nbr,
                                 const int32_t* __restrict__ noff, V3<T>* __restrict__ x_out,
                                 V3<T>* __restrict__ v_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3<T> pi = pI[i], vi = v2[i];
  V3<T> acc = v3(T(0), T(0), T(0));
  const int32_t* L = nbr + noff[i];
  walk_list<kAdjBatch>(L, noff[i + 1] - noff[i], pI, [&](const int j, const V3<T>& pj) {
    const V3<T> d = pi - pj;
    const T r2 = dot(d, d);
    if (r2 < c.h2) {
      const T w = c.h2 - r2;
      acc = acc + ((c.mass / rho[j]) * (c.poly6 * w * w * w)) * (v2[j] - vi);
    }
  });
  V3<T> p = pi, v = vi + c.visc * acc;
  FinalRec<T> rec;
  adj_finalize(c, i, p, v, rec);
  x_out[i] = p;
  v_out[i] = v;
}

// ---- backward passes --------------------------------------------------------

// finalize backward: (gp_f, gv_f) -> (gp_I partial, gv3), collider-translation
// contributions per particle (gt[i].xyz).
template <class T>
__global__ void k_adj_final_bwd(int n, AdjC<T> c, const V3<T>* __restrict__ pI, const V3<T>* __restrict__ v2,
                                const T* __restrict__ rho, const int32_t* __restrict__ nbr,
                                const int32_t* __restrict__ noff, const V3<T>* __restrict__ gxo,
                                const V3<T>* __restrict__ gvo, V3<T>* __restrict__ gpI, V3<T>* __restrict__ gv3,
                                V3<T>* __restrict__ gt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // recompute the pre-finalize state
  const V3<T> pi = pI[i], vi = v2[i];
  V3<T> acc = v3(T(0), T(0), T(0));
  const int32_t* L = nbr + noff[i];
  walk_list<kAdjBatch>(L, noff[i + 1] - noff[i], pI, [&](const int j, const V3<T>& pj) {
    const V3<T> d = pi - pj;
    const T r2 = dot(d, d);
    if (r2 < c.h2) {
      const T w = c.h2 - r2;
      acc = acc + ((c.mass / rho[j]) * (c.poly6 * w * w * w)) * (v2[j] - vi);
    }
  });
  const V3<T> v3in = vi + c.visc * acc;
  V3<T> p = pi, v = v3in;
  FinalRec<T> rec;
  adj_finalize(c, i, p, v, rec);
  V3<T> gp = gxo[i], gv = gvo[i];
  // domain clamp (applied last): clamped position components have no
  // dependence on the input; clamped velocity components scale by -damping
  T* gpp[3] = {&gp.x, &gp.y, &gp.z};
  T* gvv[3] = {&gv.x, &gv.y, &gv.z};
  for (int a = 0; a < 3; ++a)
    if (rec.clamped[a]) {
      *gpp[a] = T(0);
      *gvv[a] = -*gvv[a] * c.damping;
    }
  V3<T> gT = v3(T(0), T(0), T(0));
  if (rec.inside) {
    // p' = centre + r u, u = (centre - p)/|centre - p|; v' = v - (1+dmp)(v.u)u
    const V3<T> e = rec.centre - pi;
    const T dist = sqrt(dot(e, e));
    V3<T> gp_in = v3(T(0), T(0), T(0)), gc = v3(T(0), T(0), T(0));
    V3<T> gv_in = gv;
    if (dist > T(1e-6)) {
      const V3<T> u = (T(1) / dist) * e;
      // d p'/d centre = I + r P/dist, d p'/d p = -r P/dist, P = I - u u^T
      const V3<T> Pg = gp - dot(u, gp) * u;
      gc = gp + (c.radius / dist) * Pg;
      gp_in = (-c.radius / dist) * Pg;
      if (rec.damped) {
        const T vn = dot(v3in, u);
        const T k1 = T(1) + c.damping;
        // v' = v - k1 vn u: dv'/dv = I - k1 u u^T; dv'/du = -k1 (u v^T + vn I)
        gv_in = gv - (k1 * dot(u, gv)) * u;
        const V3<T> gu = -k1 * (dot(gv, u) * v3in + vn * gv);  // (dv'/du)^T gv
        const V3<T> Pgu = gu - dot(u, gu) * u;
        // du/dcentre = P/dist, du/dp = -P/dist
        gc = gc + (T(1) / dist) * Pgu;
        gp_in = gp_in - (T(1) / dist) * Pgu;
      }
    } else {
      gp_in = gp;  // degenerate push along +z: p' = p + (dist + r) e_z
    }
    gT = gc;
    gp = gp_in;
    gv = gv_in;
  }
  gpI[i] = gp;  // partial: the XSPH and velocity terms are added later
  gv3[i] = gv;
  gt[i] = gT;
}

// XSPH backward: gv3 -> gv2 (= gv3 + XSPH terms), gp_I (+=), grho (last lambda pass)
template <class T>
__global__ void k_adj_xsph_bwd(int n, AdjC<T> c, const V3<T>* __restrict__ pI, const V3<T>* __restrict__ v2,
                               const T* __restrict__ rho, const int32_t* __restrict__ nbr,
                               const int32_t* __restrict__ noff, const V3<T>* __restrict__ gv3,
                               V3<T>* __restrict__ gpI, V3<T>* __restrict__ gv2, T* __restrict__ grho) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3<T> pi = pI[i], vi = v2[i], gi = gv3[i];
  const T ri = rho[i];
  V3<T> gp = v3(T(0), T(0), T(0)), gv = gi;
  T gr = T(0);
  const int32_t* L = nbr + noff[i];
  walk_list<kAdjBatch>(L, noff[i + 1] - noff[i], pI, [&](const int j, const V3<T>& pj) {
    const V3<T> d = pi - pj;
    const T r2 = dot(d, d);
    if (!(r2 < c.h2)) return;
    const V3<T> vj = v2[j], gj = gv3[j];
    const T rj = rho[j];
    const T w = c.h2 - r2;
    const T W = c.poly6 * w * w * w;
    const V3<T> gradW = (T(-6) * c.poly6 * w * w) * d;  // dW/dp_i
    // dv_i = visc sum (m/rho_j)(v_j - v_i) W_ij ; dv_j holds (m/rho_i)(v_i - v_j) W_ij
    gv = gv + (c.visc * W) * ((-c.mass / rj) * gi + (c.mass / ri) * gj);
    const T s = c.visc * ((c.mass / rj) * dot(vj - vi, gi) + (c.mass / ri) * dot(vi - vj, gj));
    gp = gp + s * gradW;
    gr = gr - c.visc * (c.mass / (ri * ri)) * W * dot(vi - vj, gj);
  });
  gpI[i] = gpI[i] + gp;
  gv2[i] = gv;
  grho[i] = gr;
}

// velocity backward: v2 = (pI - x)/dt
template <class T>
__global__ void k_adj_velocity_bwd(int n, AdjC<T> c, const V3<T>* __restrict__ gv2, V3<T>* __restrict__ gpI,
                                   V3<T>* __restrict__ gx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T dtc = c.dt > T(1e-6) ? c.dt : T(1e-6);
  const V3<T> g = (T(1) / dtc) * gv2[i];
  gpI[i] = gpI[i] + g;
  gx[i] = T(-1) * g;
}

// Jacobi iteration backward, step 1: per-particle adjoints of lambda,
// rho, the denominator's G and S sums.  a = adjoint of p_{it+1}.
template <class T>
__global__ void k_adj_iter_bwd1(int n, AdjC<T> c, const V3<T>* __restrict__ p, const T* __restrict__ rho,
                                const T* __restrict__ lam, const T* __restrict__ Dv, const V3<T>* __restrict__ G,
                                const int32_t* __restrict__ nbr, const int32_t* __restrict__ noff,
                                const V3<T>* __restrict__ a, const T* __restrict__ grho_extra,
                                T* __restrict__ grho, V3<T>* __restrict__ gG, T* __restrict__ gS) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3<T> pi = p[i], ai = a[i];
  T gl = T(0);
  const int32_t* L = nbr + noff[i];
  walk_list<kAdjBatch>(L, noff[i + 1] - noff[i], p, [&](const int j, const V3<T>& pj) {
    const V3<T> d = pi - pj;
    V3<T> ph;
    T f, fpr;
    phi_and_jac(c, d, dot(d, d), ph, f, fpr);
    gl = gl + dot(ai - a[j], ph);
  });
  gl = c.K * gl;
  const T D = Dv[i];
  const T C = rho[i] / c.rho0 - T(1);
  const T gD = (D > T(1e-9)) ? gl * C / (D * D) : T(0);
  const T gr = (rho[i] > T(1e-6) ? gl * (T(-1) / D) / c.rho0 : T(0)) + (grho_extra ? grho_extra[i] : T(0));
  grho[i] = gr;
  gG[i] = (gD * T(2) / (c.rho0 * c.rho0)) * G[i];
  gS[i] = gD / (c.rho0 * c.rho0);
}

// step 2: adjoint of p_it.
template <class T>
__global__ void k_adj_iter_bwd2(int n, AdjC<T> c, const V3<T>* __restrict__ p, const T* __restrict__ lam,
                                const int32_t* __restrict__ nbr, const int32_t* __restrict__ noff,
                                const V3<T>* __restrict__ a, const T* __restrict__ grho,
                                const V3<T>* __restrict__ gG, const T* __restrict__ gS,
                                V3<T>* __restrict__ gp_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const V3<T> pi = p[i], ai = a[i], gGi = gG[i];
  const T li = lam[i], gSi = gS[i], gri = grho[i];
  V3<T> acc = ai;
  const int32_t* L = nbr + noff[i];
  walk_list<kAdjBatch>(L, noff[i + 1] - noff[i], p, [&](const int j, const V3<T>& pj) {
    const V3<T> d = pi - pj;
    const T r2 = dot(d, d);
    if (!(r2 < c.h2)) return;
    V3<T> ph;
    T f, fpr;
    phi_and_jac(c, d, r2, ph, f, fpr);
    // delta: K (l_i + l_j) J_phi (a_i - a_j)
    acc = acc + (c.K * (li + lam[j])) * jac_mul(f, fpr, d, ai - a[j]);
    // denominator: m J_phi [gG_i - gG_j + 2 m phi (gS_i + gS_j)]
    const V3<T> b = (gGi - gG[j]) + (T(2) * c.mass * (gSi + gS[j])) * ph;
    acc = acc + c.mass * jac_mul(f, fpr, d, b);
    // density: m dW/dp_i (grho_i + grho_j)
    const T w = c.h2 - r2;
    acc = acc + (c.mass * (gri + grho[j]) * T(-6) * c.poly6 * w * w) * d;
  });
  gp_out[i] = acc;
}

// predict backward: gp0 -> gx (+=), gv_in, gravity contributions.
template <class T>
__global__ void k_adj_predict_bwd(int n, AdjC<T> c, const V3<T>* __restrict__ v, const V3<T>* __restrict__ gp0,
                                  V3<T>* __restrict__ gx, V3<T>* __restrict__ gv, T* __restrict__ gg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  V3<T> u = v[i];
  u.z = u.z + c.g * c.dt;
  const T vmax = T(0.4) * c.h / (c.dt > T(1e-6) ? c.dt : T(1e-6));
  const T speed = sqrt(dot(u, u));
  const V3<T> g0 = gp0[i];
  V3<T> g1 = c.dt * g0;  // adjoint of the clamped velocity
  const bool clamp = c.dec_mode == DEC_REPLAY ? ((c.dec[2 * i] >> 8) & 1) != 0 : speed > vmax;
  if (clamp) {
    // u' = vmax u / |u|: J = (vmax/|u|)(I - uhat uhat^T)
    const V3<T> uh = (T(1) / speed) * u;
    g1 = (vmax / speed) * (g1 - dot(uh, g1) * uh);
  }
  gx[i] = gx[i] + g0;
  gv[i] = g1;
  gg[i] = g1.z * c.dt;
}

// ---- splat loss: field and its adjoint ------------------------------------
template <class T>
__global__ void k_adj_splat(long long S, FxSurface sf, Grid g, AdjC<T> c, T kr, T coef,
                            const V3<T>* __restrict__ x, const int32_t* __restrict__ off,
                            const int32_t* __restrict__ sidx, T* __restrict__ field, const uint8_t* __restrict__ mask) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int nx = sf.dims[0], ny = sf.dims[1];
  const int gx = (int)(s % nx), gy = (int)((s / nx) % ny), gz = (int)(s / ((long long)nx * ny));
  const V3<T> ps = v3((T)sf.lo[0] + (T)gx * (T)sf.spacing, (T)sf.lo[1] + (T)gy * (T)sf.spacing,
                      (T)sf.lo[2] + (T)gz * (T)sf.spacing);
  int bx, by, bz;
  cell_coord(g, (float)ps.x, (float)ps.y, (float)ps.z, bx, by, bz);
  const T kr2 = kr * kr;
  T acc = T(0);
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int X = bx + dx, Y = by + dy, Z = bz + dz;
        if (X < 0 || Y < 0 || Z < 0 || X >= g.cx || Y >= g.cy || Z >= g.cz) continue;
        const int ci = cell_index(g, X, Y, Z);
        for (int q = off[ci]; q < off[ci + 1]; ++q) {
          const V3<T> d = ps - x[sidx[q]];
          const T r2 = dot(d, d);
          if (r2 < kr2) {
            const T w = kr2 - r2;
            acc = acc + w * w * w;
          }
        }
      }
  T val = c.mass * coef * acc / c.rho0;
  if (mask && mask[s]) val = T(0);
  field[s] = val;
}

// gx_j += sum over samples within kr of gF(s) (m coef / rho0) 6 (kr^2 - r^2)^2 (s - x_j)
template <class T>
__global__ void k_adj_splat_bwd(int n, FxSurface sf, AdjC<T> c, T kr, T coef, const V3<T>* __restrict__ x,
                                const T* __restrict__ gF, const uint8_t* __restrict__ mask, V3<T>* __restrict__ gx) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const V3<T> xj = x[j];
  const T sp = (T)sf.spacing, kr2 = kr * kr;
  const T lo[3] = {(T)sf.lo[0], (T)sf.lo[1], (T)sf.lo[2]};
  const T xv[3] = {xj.x, xj.y, xj.z};
  int a0[3], a1[3];
  for (int a = 0; a < 3; ++a) {
    a0[a] = max(0, (int)floor((xv[a] - kr - lo[a]) / sp));
    a1[a] = min(sf.dims[a] - 1, (int)ceil((xv[a] + kr - lo[a]) / sp));
  }
  V3<T> acc = v3(T(0), T(0), T(0));
  const T f = c.mass * coef / c.rho0 * T(6);
  for (int k = a0[2]; k <= a1[2]; ++k)
    for (int jj = a0[1]; jj <= a1[1]; ++jj)
      for (int i = a0[0]; i <= a1[0]; ++i) {
        const long long s = ((long long)k * sf.dims[1] + jj) * sf.dims[0] + i;
        if (mask && mask[s]) continue;
        const V3<T> ps = v3(lo[0] + (T)i * sp, lo[1] + (T)jj * sp, lo[2] + (T)k * sp);
        const V3<T> d = ps - xj;
        const T r2 = dot(d, d);
        if (r2 < kr2) {
          const T w = kr2 - r2;
          acc = acc + (gF[s] * f * w * w) * d;
        }
      }
  gx[j] = gx[j] + acc;
}

// ---- deterministic reduction (fixed tree) -----------------------------------
template <class T>
__global__ void k_adj_reduce(long long n, int stride, int comp, const T* __restrict__ in, T* __restrict__ out) {
  __shared__ T buf[256];
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  buf[threadIdx.x] = i < n ? in[i * stride + comp] : T(0);
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) buf[threadIdx.x] = buf[threadIdx.x] + buf[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = buf[0];
}

template <class T>
__global__ void k_adj_reduce_final(int nb, const T* __restrict__ in, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += (double)in[b];
  *out += s;
}

__global__ void k_adj_zero(int32_t* a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// ---- host side --------------------------------------------------------------
struct AdjWs {
  size_t cnt, off, rank, key, sidx, part, pf, nbr, noff, lcnt, flag;
  size_t v1, P, rho, lam, D, G, v2, a, b, grho, gG, gS, gv2, gv3, gt, gg, red;
  size_t total;
};

static AdjWs adj_layout(const FxParams& p, int prec) {
  const size_t n = (size_t)(p.n > 0 ? p.n : 0), c = (size_t)p.n_cells, T = (size_t)prec, I = (size_t)p.iterations;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  AdjWs w;
  size_t o = 0;
  w.cnt = o;  o += al((c + 1) * 4);
  w.off = o;  o += al((c + 1) * 4);
  w.rank = o; o += al(n * 4 + 4);
  w.key = o;  o += al(n * 4 + 4);
  w.sidx = o; o += al(n * 4 + 4);
  w.part = o; o += al(scan_partials((long long)std::max(c, n) + 1) * 4 + 4);
  w.pf = o;   o += al(n * 16 + 16);
  w.nbr = o;  o += al(n * kMaxNbr * 4 + 4);
  w.noff = o; o += al(n * 4 + 8);
  w.lcnt = o; o += al(n * 4 + 4);
  w.flag = o; o += al(64);
  w.v1 = o;   o += al(n * 4 * T + 64);
  w.P = o;    o += al((I + 1) * n * 4 * T + 64);
  w.rho = o;  o += al((I + 1) * n * T + 64);
  w.lam = o;  o += al((I + 1) * n * T + 64);
  w.D = o;    o += al((I + 1) * n * T + 64);
  w.G = o;    o += al((I + 1) * n * 4 * T + 64);
  w.v2 = o;   o += al(n * 4 * T + 64);
  w.a = o;    o += al(n * 4 * T + 64);
  w.b = o;    o += al(n * 4 * T + 64);
  w.grho = o; o += al(n * T + 64);
  w.gG = o;   o += al(n * 4 * T + 64);
  w.gS = o;   o += al(n * T + 64);
  w.gv2 = o;  o += al(n * 4 * T + 64);
  w.gv3 = o;  o += al(n * 4 * T + 64);
  w.gt = o;   o += al(n * 4 * T + 64);
  w.gg = o;   o += al(n * T + 64);
  w.red = o;  o += al(((n + 255) / 256 + 1) * 8 + 64);
  w.total = o;
  return w;
}

static int adj_validate(const FxParams* p, int prec, void* work, size_t work_bytes) {
  if (!p) return fail(FX_EINVAL, "adjoint: null params");
  if (prec != 4 && prec != 8) return fail(FX_EINVAL, "adjoint: precision must be 4 or 8 bytes");
  if (p->surface_tension > 0.0f || p->scorr_k > 0.0f)
    return fail(FX_EINVAL, "adjoint: surface_tension and scorr_k must be 0");
  if (p->n < 0 || p->cells[0] < 1 || p->cells[1] < 1 || p->cells[2] < 1 ||
      (long long)p->cells[0] * p->cells[1] * p->cells[2] != p->n_cells)
    return fail(FX_EINVAL, "adjoint: bad particle count or cell dims");
  if (p->iterations < 1) return fail(FX_EINVAL, "adjoint: iterations must be >= 1");
  if ((long long)p->n * kMaxNbr >= (1LL << 31))
    return fail(FX_ERANGE, "adjoint: %d particles exceed the 32-bit neighbour-list capacity", p->n);
  if (!work || work_bytes < adj_layout(*p, prec).total)
    return fail(FX_ENOMEM, "adjoint workspace too small: %zu < %zu", work_bytes, adj_layout(*p, prec).total);
  return FX_OK;
}

#define ALAUNCH(kernel, n, st, ...) FX_LAUNCH(kernel, dim3(grid_for(n)), dim3(kBlock), st, __VA_ARGS__)

// A substep's record: [0, 2n) branch decisions (flags, voxel), [2n, 3n + 1)
// list offsets, [3n + 1, 3n + 1 + n * kMaxNbr) neighbour lists.  Recording
// keeps the lists there; replaying reuses them instead of rebuilding from
// the new state.
static size_t adj_record_ints(int n) { return n > 0 ? (size_t)n * (3 + kMaxNbr) + 1 : 0; }

template <class T>
static void adj_list_ptrs(const AdjC<T>& c, int n, char* b, const AdjWs& w, int32_t*& nbr, int32_t*& noff) {
  if (c.dec_mode != DEC_NONE) {
    noff = c.dec + 2 * (size_t)n;
    nbr = c.dec + 3 * (size_t)n + 1;
  } else {
    noff = (int32_t*)(b + w.noff);
    nbr = (int32_t*)(b + w.nbr);
  }
}

// Forward substep; keeps every iterate in the workspace.
template <class T>
static int adj_forward(const FxParams& p, const AdjC<T>& c, const V3<T>* x, const V3<T>* v, V3<T>* x_out,
                       V3<T>* v_out, char* b, const AdjWs& w, hipStream_t st) {
  const int n = p.n, nc = p.n_cells, I = p.iterations;
  if (n == 0) return FX_OK;
  V3<T>* v1 = (V3<T>*)(b + w.v1);
  V3<T>* P = (V3<T>*)(b + w.P);
  T* rho = (T*)(b + w.rho);
  T* lam = (T*)(b + w.lam);
  T* D = (T*)(b + w.D);
  V3<T>* G = (V3<T>*)(b + w.G);
  V3<T>* v2 = (V3<T>*)(b + w.v2);
  int32_t *cnt = (int32_t*)(b + w.cnt), *off = (int32_t*)(b + w.off), *rank = (int32_t*)(b + w.rank),
          *key = (int32_t*)(b + w.key), *sidx = (int32_t*)(b + w.sidx), *part = (int32_t*)(b + w.part),
          *flag = (int32_t*)(b + w.flag), *lcnt = (int32_t*)(b + w.lcnt), *nbr, *noff;
  adj_list_ptrs(c, n, b, w, nbr, noff);
  float4* pf = (float4*)(b + w.pf);
  const Grid g = make_grid(p);
  ALAUNCH(k_adj_predict<T>, n, st, n, c, x, v, P, v1);
  if (c.dec_mode != DEC_REPLAY) {
    ALAUNCH(k_adj_to_f4<T>, n, st, n, (const V3<T>*)P, pf);
    FX_LAUNCH(k_adj_zero, dim3(grid_for(nc + 1)), dim3(kBlock), st, cnt, nc + 1);
    ALAUNCH(k_adj_key, n, st, n, g, (const float4*)pf, key, cnt, rank);
    FX_TRY(exclusive_scan_i32(cnt, off, nc, part, st));
    ALAUNCH(k_adj_scatter, n, st, n, (const int32_t*)key, (const int32_t*)rank, (const int32_t*)off, sidx);
    ALAUNCH((k_adj_lists<T, false>), n, st, n, g, c, (const V3<T>*)P, (const float4*)pf, (const int32_t*)off,
            (const int32_t*)sidx, lcnt, (const int32_t*)nullptr, (int32_t*)nullptr);
    FX_TRY(exclusive_scan_i32(lcnt, noff, n, part, st));
    FX_LAUNCH(k_adj_cap, dim3(grid_for(n + 1)), dim3(kBlock), st, n, (long long)n * kMaxNbr, noff, flag);
    ALAUNCH((k_adj_lists<T, true>), n, st, n, g, c, (const V3<T>*)P, (const float4*)pf, (const int32_t*)off,
            (const int32_t*)sidx, (int32_t*)nullptr, (const int32_t*)noff, nbr);
    // lists past kWaveSortMax queue into rank (dead after the scatter), their
    // count in part[0] (dead after the scan)
    FX_LAUNCH(k_adj_zero, dim3(1), dim3(kBlock), st, part, 1);
    FX_LAUNCH(k_adj_sort_lists, dim3((unsigned)((n + 3) / 4)), dim3(256), st, n, (const int32_t*)noff, nbr, rank,
              part);
    FX_TRY(sort_queued_runs(noff, nbr, rank, part, st));
  }
  for (int it = 0; it < I; ++it) {
    V3<T>* Pi = P + (size_t)it * n;
    ALAUNCH(k_adj_lambda<T>, n, st, n, c, (const V3<T>*)Pi, (const int32_t*)nbr, (const int32_t*)noff,
            rho + (size_t)it * n, lam + (size_t)it * n, D + (size_t)it * n, G + (size_t)it * n);
    ALAUNCH(k_adj_delta<T>, n, st, n, c, (const V3<T>*)Pi, (const T*)(lam + (size_t)it * n), (const int32_t*)nbr,
            (const int32_t*)noff, P + (size_t)(it + 1) * n);
  }
  const V3<T>* PI = P + (size_t)I * n;
  const T* rho_last = rho + (size_t)(I - 1) * n;
  ALAUNCH(k_adj_velocity<T>, n, st, n, c, PI, x, v2);
  if (x_out && v_out)
    ALAUNCH(k_adj_xsph_final<T>, n, st, n, c, PI, (const V3<T>*)v2, rho_last, (const int32_t*)nbr,
            (const int32_t*)noff, x_out, v_out);
  return check_launch("adjoint forward");
}

template <class T>
static int adj_reduce(const T* vals, long long n, int stride, int comp, char* b, const AdjWs& w, double* out,
                      hipStream_t st) {
  T* red = (T*)(b + w.red);
  const unsigned nb = grid_for(n, 256);
  FX_LAUNCH(k_adj_reduce<T>, dim3(nb), dim3(256), st, n, stride, comp, vals, red);
  FX_LAUNCH(k_adj_reduce_final<T>, dim3(1), dim3(64), st, (int)nb, (const T*)red, out);
  return FX_OK;
}

template <class T>
static int adj_backward(const FxParams& p, const AdjC<T>& c, const V3<T>* x, const V3<T>* v, const V3<T>* gxo,
                        const V3<T>* gvo, V3<T>* gx, V3<T>* gv, double* gparams, char* b, const AdjWs& w,
                        hipStream_t st) {
  const int n = p.n, I = p.iterations;
  if (n == 0) return FX_OK;
  FX_TRY(adj_forward<T>(p, c, x, v, nullptr, nullptr, b, w, st));
  V3<T>* P = (V3<T>*)(b + w.P);
  T* rho = (T*)(b + w.rho);
  T* lam = (T*)(b + w.lam);
  T* D = (T*)(b + w.D);
  V3<T>* G = (V3<T>*)(b + w.G);
  V3<T>* v2 = (V3<T>*)(b + w.v2);
  V3<T>* A = (V3<T>*)(b + w.a);
  V3<T>* B = (V3<T>*)(b + w.b);
  T* grho = (T*)(b + w.grho);
  V3<T>* gG = (V3<T>*)(b + w.gG);
  T* gS = (T*)(b + w.gS);
  V3<T>* gv2 = (V3<T>*)(b + w.gv2);
  V3<T>* gv3 = (V3<T>*)(b + w.gv3);
  V3<T>* gt = (V3<T>*)(b + w.gt);
  T* gg = (T*)(b + w.gg);
  T* grho_x = (T*)(b + w.gS);  // reuse: XSPH's density adjoint, consumed before gS is written
  int32_t *nbr_w, *noff_w;
  adj_list_ptrs(c, n, b, w, nbr_w, noff_w);
  const int32_t* nbr = nbr_w;
  const int32_t* noff = noff_w;
  const V3<T>* PI = P + (size_t)I * n;
  const T* rho_last = rho + (size_t)(I - 1) * n;
  // finalize, XSPH, velocity
  ALAUNCH(k_adj_final_bwd<T>, n, st, n, c, PI, (const V3<T>*)v2, rho_last, nbr, noff, gxo, gvo, A, gv3, gt);
  for (int a = 0; a < 3; ++a) FX_TRY(adj_reduce<T>((const T*)gt, n, 4, a, b, w, gparams + 1 + a, st));
  T* grho_tmp = (T*)(b + w.grho);
  ALAUNCH(k_adj_xsph_bwd<T>, n, st, n, c, PI, (const V3<T>*)v2, rho_last, nbr, noff, (const V3<T>*)gv3, A, gv2,
          grho_tmp);
  // keep XSPH's density adjoint apart: copy into the gS slot
  FX_HIP_CHECK(hipMemcpyAsync(grho_x, grho_tmp, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, st));
  ALAUNCH(k_adj_velocity_bwd<T>, n, st, n, c, (const V3<T>*)gv2, A, gx);
  // Jacobi iterations, last to first: A = adjoint of P[it+1] -> B = adjoint of P[it]
  T* grho_extra = (T*)(b + w.gv2);  // gv2 is consumed; reuse as the XSPH density adjoint
  FX_HIP_CHECK(hipMemcpyAsync(grho_extra, grho_x, (size_t)n * sizeof(T), hipMemcpyDeviceToDevice, st));
  for (int it = I - 1; it >= 0; --it) {
    const V3<T>* Pi = P + (size_t)it * n;
    const T* li = lam + (size_t)it * n;
    ALAUNCH(k_adj_iter_bwd1<T>, n, st, n, c, Pi, (const T*)(rho + (size_t)it * n), li, (const T*)(D + (size_t)it * n),
            (const V3<T>*)(G + (size_t)it * n), nbr, noff, (const V3<T>*)A,
            (const T*)(it == I - 1 ? grho_extra : nullptr), grho, gG, gS);
    ALAUNCH(k_adj_iter_bwd2<T>, n, st, n, c, Pi, li, nbr, noff, (const V3<T>*)A, (const T*)grho, (const V3<T>*)gG,
            (const T*)gS, B);
    V3<T>* tmp = A;
    A = B;
    B = tmp;
  }
  ALAUNCH(k_adj_predict_bwd<T>, n, st, n, c, v, (const V3<T>*)A, gx, gv, gg);
  FX_TRY(adj_reduce<T>((const T*)gg, n, 1, 0, b, w, gparams, st));
  return check_launch("adjoint backward");
}

}  // namespace fx

using namespace fx;

extern "C" {

size_t fx_adjoint_record_bytes(const FxParams* p) { return p ? adj_record_ints(p->n) * 4 : 0; }

size_t fx_adjoint_workspace_bytes(const FxParams* p, int precision) {
  if (!p || (precision != 4 && precision != 8)) return 0;
  return adj_layout(*p, precision).total;
}

int fx_adjoint_forward(const FxParams* p, int precision, double gravity, const double translation[3],
                       double list_skin, const uint8_t* collider, int32_t* decisions, int decision_mode,
                       const void* x, const void* v, void* x_out, void* v_out, void* work, size_t work_bytes,
                       void* stream) {
  FX_TRY(adj_validate(p, precision, work, work_bytes));
  if (p->n > 0 && (!x || !v || !x_out || !v_out)) return fail(FX_EINVAL, "adjoint forward: null state");
  if (decision_mode < DEC_NONE || decision_mode > DEC_REPLAY) return fail(FX_EINVAL, "adjoint: bad decision mode");
  const AdjWs w = adj_layout(*p, precision);
  hipStream_t st = (hipStream_t)stream;
  if (precision == 8)
    return adj_forward<double>(*p, make_adjc<double>(*p, gravity, translation, collider, list_skin, decisions,
                                                     decision_mode),
                               (const V3<double>*)x, (const V3<double>*)v, (V3<double>*)x_out, (V3<double>*)v_out,
                               (char*)work, w, st);
  return adj_forward<float>(*p, make_adjc<float>(*p, gravity, translation, collider, list_skin, decisions,
                                                   decision_mode),
                            (const V3<float>*)x, (const V3<float>*)v, (V3<float>*)x_out, (V3<float>*)v_out,
                            (char*)work, w, st);
}

int fx_adjoint_backward(const FxParams* p, int precision, double gravity, const double translation[3],
                        double list_skin, const uint8_t* collider, int32_t* decisions, const void* x,
                        const void* v,
                        const void* gx_out, const void* gv_out, void* gx, void* gv, double* grad_params,
                        void* work, size_t work_bytes, void* stream) {
  FX_TRY(adj_validate(p, precision, work, work_bytes));
  if (p->n > 0 && (!x || !v || !gx_out || !gv_out || !gx || !gv || !grad_params))
    return fail(FX_EINVAL, "adjoint backward: null argument");
  const AdjWs w = adj_layout(*p, precision);
  hipStream_t st = (hipStream_t)stream;
  if (precision == 8)
    return adj_backward<double>(*p, make_adjc<double>(*p, gravity, translation, collider, list_skin, decisions,
                                                      DEC_REPLAY),
                                (const V3<double>*)x, (const V3<double>*)v, (const V3<double>*)gx_out,
                                (const V3<double>*)gv_out, (V3<double>*)gx, (V3<double>*)gv, grad_params,
                                (char*)work, w, st);
  return adj_backward<float>(*p, make_adjc<float>(*p, gravity, translation, collider, list_skin, decisions,
                                                    DEC_REPLAY),
                             (const V3<float>*)x, (const V3<float>*)v, (const V3<float>*)gx_out,
                             (const V3<float>*)gv_out, (V3<float>*)gx, (V3<float>*)gv, grad_params,
                             (char*)work, w, st);
}

int fx_adjoint_overflow(const FxParams* p, int precision, const void* work, int32_t* host_flag, void* stream) {
  if (!p || !work || !host_flag) return fail(FX_EINVAL, "fx_adjoint_overflow: null argument");
  const AdjWs w = adj_layout(*p, precision);
  int32_t* flag = (int32_t*)((char*)work + w.flag);
  FX_HIP_CHECK(hipMemcpyAsync(host_flag, flag, 4, hipMemcpyDeviceToHost, (hipStream_t)stream));
  FX_LAUNCH(k_adj_zero, dim3(1), dim3(64), (hipStream_t)stream, flag, 1);
  FX_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
  return FX_OK;
}

int fx_adjoint_splat(const FxParams* p, const FxSurface* surf, int precision, const void* x,
                     const uint8_t* sample_mask, void* field, void* work, size_t work_bytes, void* stream) {
  FX_TRY(adj_validate(p, precision, work, work_bytes));
  if (!surf || !field || (p->n > 0 && !x)) return fail(FX_EINVAL, "fx_adjoint_splat: null argument");
  const AdjWs w = adj_layout(*p, precision);
  hipStream_t st = (hipStream_t)stream;
  char* b = (char*)work;
  const int n = p->n, nc = p->n_cells;
  int32_t *cnt = (int32_t*)(b + w.cnt), *off = (int32_t*)(b + w.off), *rank = (int32_t*)(b + w.rank),
          *key = (int32_t*)(b + w.key), *sidx = (int32_t*)(b + w.sidx), *part = (int32_t*)(b + w.part),
          *flag = (int32_t*)(b + w.flag);
  float4* pf = (float4*)(b + w.pf);
  // grid of cell size = kernel radius over the domain (samples clamp onto it)
  FxParams q = *p;
  const float kr = surf->kernel_radius;
  for (int a = 0; a < 3; ++a) q.cells[a] = std::max(1, (int)std::ceil((p->hi[a] - p->lo[a]) / kr));
  q.n_cells = q.cells[0] * q.cells[1] * q.cells[2];
  q.cell_size = kr;
  if (q.n_cells > nc) return fail(FX_EINVAL, "fx_adjoint_splat: kernel radius below the cell size");
  const Grid g = make_grid(q);
  const long long S = (long long)surf->dims[0] * surf->dims[1] * surf->dims[2];
  FX_LAUNCH(k_adj_zero, dim3(grid_for(q.n_cells + 1)), dim3(kBlock), st, cnt, q.n_cells + 1);
  const double pi = 3.14159265358979323846;
  const double coef = 315.0 / (64.0 * pi * std::pow((double)kr, 9.0));
  if (precision == 8) {
    AdjC<double> c = make_adjc<double>(*p, 0.0, nullptr, nullptr, 0.0);
    if (n > 0) {
      ALAUNCH(k_adj_to_f4<double>, n, st, n, (const V3<double>*)x, pf);
      ALAUNCH(k_adj_key, n, st, n, g, (const float4*)pf, key, cnt, rank);
    }
    FX_TRY(exclusive_scan_i32(cnt, off, q.n_cells, part, st));
    if (n > 0) {
      ALAUNCH(k_adj_scatter, n, st, n, (const int32_t*)key, (const int32_t*)rank, (const int32_t*)off, sidx);
      FX_TRY(canon_cells(off, sidx, q.n_cells, rank, part, st));  // deterministic summation order
    }
    FX_LAUNCH(k_adj_splat<double>, dim3(grid_for(S)), dim3(kBlock), st, S, *surf, g, c, (double)kr, coef,
              (const V3<double>*)x, (const int32_t*)off, (const int32_t*)sidx, (double*)field, sample_mask);
  } else {
    AdjC<float> c = make_adjc<float>(*p, 0.0, nullptr, nullptr, 0.0);
    if (n > 0) {
      ALAUNCH(k_adj_to_f4<float>, n, st, n, (const V3<float>*)x, pf);
      ALAUNCH(k_adj_key, n, st, n, g, (const float4*)pf, key, cnt, rank);
    }
    FX_TRY(exclusive_scan_i32(cnt, off, q.n_cells, part, st));
    if (n > 0) {
      ALAUNCH(k_adj_scatter, n, st, n, (const int32_t*)key, (const int32_t*)rank, (const int32_t*)off, sidx);
      FX_TRY(canon_cells(off, sidx, q.n_cells, rank, part, st));  // deterministic summation order
    }
    FX_LAUNCH(k_adj_splat<float>, dim3(grid_for(S)), dim3(kBlock), st, S, *surf, g, c, kr, (float)coef,
              (const V3<float>*)x, (const int32_t*)off, (const int32_t*)sidx, (float*)field, sample_mask);
  }
  return check_launch("fx_adjoint_splat");
}

int fx_adjoint_splat_backward(const FxParams* p, const FxSurface* surf, int precision, const void* x,
                              const uint8_t* sample_mask, const void* gfield, void* gx, void* stream) {
  if (!p || !surf || !gfield || (p->n > 0 && (!x || !gx))) return fail(FX_EINVAL, "splat backward: null argument");
  if (precision != 4 && precision != 8) return fail(FX_EINVAL, "splat backward: precision must be 4 or 8");
  hipStream_t st = (hipStream_t)stream;
  const int n = p->n;
  const float kr = surf->kernel_radius;
  const double pi = 3.14159265358979323846;
  const double coef = 315.0 / (64.0 * pi * std::pow((double)kr, 9.0));
  if (n == 0) return FX_OK;
  if (precision == 8) {
    AdjC<double> c = make_adjc<double>(*p, 0.0, nullptr, nullptr, 0.0);
    ALAUNCH(k_adj_splat_bwd<double>, n, st, n, *surf, c, (double)kr, coef, (const V3<double>*)x,
            (const double*)gfield, sample_mask, (V3<double>*)gx);
  } else {
    AdjC<float> c = make_adjc<float>(*p, 0.0, nullptr, nullptr, 0.0);
    ALAUNCH(k_adj_splat_bwd<float>, n, st, n, *surf, c, kr, (float)coef, (const V3<float>*)x,
            (const float*)gfield, sample_mask, (V3<float>*)gx);
  }
  return check_launch("fx_adjoint_splat_backward");
}

}  // extern "C"
