/* Oracle (TEST INFRASTRUCTURE ONLY - imported by tests/, by the smoke check
 * and by bench.py's CPU-baseline leg; never by the product path).
 *
 * CPU restatement of the reference's PBF step / grid / surface path:
 *   - every GLSL pass in FP32 (IEEE division and sqrt, built with
 *     -ffp-contract=off) and FP64 (pbf_passes.inc, included twice);
 *   - the bitonic compare-exchange network of shaders/sph_sort.glsl:14-34
 *     driven as solver/sph.py:488-495 drives it, and a stable counting sort;
 *   - shaders/sph_cell_range.glsl:10-32 with the clear of
 *     shaders/sph_cell_clear.glsl:8-16;
 *   - the substep composition of solver/sph.py:501-514;
 *   - marching cubes: the case-table derivation of
 *     solver/marching_cubes.py:60-173 (dict insertion order included) and
 *     extract() of solver/marching_cubes.py:191-280, in float64;
 *   - MT19937 with init_by_array and genrand_res53 (CPython's
 *     random.Random(0), used by solver/sph.py:329-346), and the jittered
 *     lattice seeder of solver/sph.py:316-347.
 * Built by oracle/Makefile into oracle/liboracle.so (OpenMP over all host
 * cores: the same library times the CPU baseline).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "flowx.h"

/* ---- FP32 transliteration ------------------------------------------------ */
#define R float
#define SFX(name) name##_f32
#define RSQRT_EXACT(x) sqrtf(x)
#define RFLOOR(x) floorf(x)
#define RPOW(x, y) powf((x), (y))
#define LIT(x) ((float)(x))
#define RPI 3.14159265358979323846f
#include "pbf_passes.inc"
#undef R
#undef SFX
#undef RSQRT_EXACT
#undef RFLOOR
#undef RPOW
#undef LIT
#undef RPI

/* ---- FP64 transliteration ------------------------------------------------ */
#define R double
#define SFX(name) name##_f64
#define RSQRT_EXACT(x) sqrt(x)
#define RFLOOR(x) floor(x)
#define RPOW(x, y) pow((x), (y))
#define LIT(x) ((double)(x))
#define RPI 3.14159265358979323846
#include "pbf_passes.inc"

/* ---- grid build: bitonic network (shaders/sph_sort.glsl:14-34, host loop
 * solver/sph.py:488-495).  Only the lower element of each pair swaps; the
 * comparison is on the key alone, so equal keys are not kept in order. */
void orc_bitonic(int P, int32_t* keys, int32_t* idx) {
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < P; ++i) {
        int partner = i ^ j;
        if (partner <= i) continue;
        int ascending = (i & k) == 0;
        if ((keys[i] > keys[partner]) == ascending) {
          int32_t tk = keys[i], ti = idx[i];
          keys[i] = keys[partner];
          idx[i] = idx[partner];
          keys[partner] = tk;
          idx[partner] = ti;
        }
      }
    }
  }
}

/* Stable counting sort by key (keys in [0, max_key]). */
int orc_stable_sort(int P, int32_t* keys, int32_t* idx, int max_key) {
  int32_t* count = (int32_t*)calloc((size_t)max_key + 2, sizeof(int32_t));
  int32_t* tk = (int32_t*)malloc((size_t)P * sizeof(int32_t));
  int32_t* ti = (int32_t*)malloc((size_t)P * sizeof(int32_t));
  if (!count || !tk || !ti) {
    free(count);
    free(tk);
    free(ti);
    return -1;
  }
  for (int i = 0; i < P; ++i) count[keys[i] + 1]++;
  for (int c = 0; c <= max_key; ++c) count[c + 1] += count[c];
  for (int i = 0; i < P; ++i) {
    int d = count[keys[i]]++;
    tk[d] = keys[i];
    ti[d] = idx[i];
  }
  memcpy(keys, tk, (size_t)P * sizeof(int32_t));
  memcpy(idx, ti, (size_t)P * sizeof(int32_t));
  free(count);
  free(tk);
  free(ti);
  return 0;
}

/* shaders/sph_cell_clear.glsl:8-16 followed by shaders/sph_cell_range.glsl:10-32 */
void orc_cell_range(int n_cells, int P, const int32_t* keys, int32_t* start, int32_t* end) {
  for (int c = 0; c < n_cells; ++c) start[c] = end[c] = -1;
  for (int i = 0; i < P; ++i) {
    int key = keys[i];
    if (key >= n_cells) continue;
    int prev = (i == 0) ? -1 : keys[i - 1];
    int next = (i == P - 1) ? -1 : keys[i + 1];
    if (key != prev) start[key] = i;
    if (key != next) end[key] = i + 1;
  }
}

/* solver/sph.py:482-498 _build_grid.  sort_mode 0 = bitonic (the reference
 * network), 1 = stable counting sort.  keys / idx hold P slots. */
#define BUILD_GRID(SUF, RT)                                                             \
  int orc_build_grid_##SUF(const FxParams* p, const RT* pred4, int P, int32_t* keys,    \
                           int32_t* idx, int32_t* start, int32_t* end, int sort_mode) { \
    orc_grid_key_##SUF(p, pred4, P, keys, idx);                                          \
    if (sort_mode == 0)                                                                  \
      orc_bitonic(P, keys, idx);                                                         \
    else if (orc_stable_sort(P, keys, idx, p->n_cells) != 0)                             \
      return -1;                                                                         \
    orc_cell_range(p->n_cells, P, keys, start, end);                                     \
    return 0;                                                                            \
  }
BUILD_GRID(f32, float)
BUILD_GRID(f64, double)

/* solver/sph.py:501-514 _substep, one substep in the reference pass order. */
#define SUBSTEP(SUF, RT)                                                                     \
  int orc_substep_##SUF(const FxParams* p, RT* pos4, RT* vel4, RT* pred4, RT* lambda4,      \
                        RT* delta4, RT* normal4, int P, int32_t* keys, int32_t* idx,         \
                        int32_t* start, int32_t* end, const uint8_t* collider,               \
                        int sort_mode) {                                                     \
    if (p->surface_tension > 0.0f)                                                           \
      orc_normal_##SUF(p, pred4, lambda4, idx, start, end, normal4);                        \
    orc_predict_##SUF(p, pos4, vel4, normal4, pred4);                                        \
    if (orc_build_grid_##SUF(p, pred4, P, keys, idx, start, end, sort_mode) != 0) return -1; \
    for (int it = 0; it < p->iterations; ++it) {                                             \
      orc_lambda_##SUF(p, pred4, idx, start, end, lambda4);                                  \
      orc_delta_##SUF(p, pred4, lambda4, idx, start, end, delta4);                          \
      orc_apply_delta_##SUF(p, pred4, delta4);                                               \
    }                                                                                        \
    orc_velocity_##SUF(p, pred4, pos4, vel4);                                                \
    orc_xsph_##SUF(p, pred4, vel4, lambda4, idx, start, end, delta4);                        \
    orc_finalize_##SUF(p, vel4, delta4, pred4, collider, pos4);                              \
    return 0;                                                                                \
  }
SUBSTEP(f32, float)
SUBSTEP(f64, double)

/* ---- MT19937 (Matsumoto & Nishimura 1998, init_by_array of the 2002
 * reference code) and genrand_res53, as CPython's random.Random(seed) uses
 * them for a non-negative int seed (key = the seed's 32-bit words). -------- */
typedef struct {
  uint32_t mt[624];
  int mti;
} Mt;

static void mt_init_genrand(Mt* m, uint32_t s) {
  m->mt[0] = s;
  for (int i = 1; i < 624; ++i)
    m->mt[i] = 1812433253u * (m->mt[i - 1] ^ (m->mt[i - 1] >> 30)) + (uint32_t)i;
  m->mti = 624;
}

static void mt_init_by_array(Mt* m, const uint32_t* key, int len) {
  mt_init_genrand(m, 19650218u);
  int i = 1, j = 0;
  for (int k = (624 > len ? 624 : len); k; --k) {
    m->mt[i] = (m->mt[i] ^ ((m->mt[i - 1] ^ (m->mt[i - 1] >> 30)) * 1664525u)) + key[j] +
               (uint32_t)j;
    ++i;
    ++j;
    if (i >= 624) {
      m->mt[0] = m->mt[623];
      i = 1;
    }
    if (j >= len) j = 0;
  }
  for (int k = 623; k; --k) {
    m->mt[i] = (m->mt[i] ^ ((m->mt[i - 1] ^ (m->mt[i - 1] >> 30)) * 1566083941u)) - (uint32_t)i;
    ++i;
    if (i >= 624) {
      m->mt[0] = m->mt[623];
      i = 1;
    }
  }
  m->mt[0] = 0x80000000u;
}

static uint32_t mt_genrand(Mt* m) {
  static const uint32_t mag01[2] = {0u, 0x9908b0dfu};
  uint32_t y;
  if (m->mti >= 624) {
    int kk;
    for (kk = 0; kk < 624 - 397; ++kk) {
      y = (m->mt[kk] & 0x80000000u) | (m->mt[kk + 1] & 0x7fffffffu);
      m->mt[kk] = m->mt[kk + 397] ^ (y >> 1) ^ mag01[y & 1u];
    }
    for (; kk < 623; ++kk) {
      y = (m->mt[kk] & 0x80000000u) | (m->mt[kk + 1] & 0x7fffffffu);
      m->mt[kk] = m->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
    }
    y = (m->mt[623] & 0x80000000u) | (m->mt[0] & 0x7fffffffu);
    m->mt[623] = m->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
    m->mti = 0;
  }
  y = m->mt[m->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

static double mt_res53(Mt* m) {
  uint32_t a = mt_genrand(m) >> 5, b = mt_genrand(m) >> 6;
  return (a * 67108864.0 + b) * (1.0 / 9007199254740992.0);
}

/* n draws of random.Random(seed).random() for a seed < 2**32. */
void orc_mt_random(uint32_t seed, long n, double* out) {
  Mt m;
  uint32_t key[1] = {seed};
  mt_init_by_array(&m, key, 1);
  for (long i = 0; i < n; ++i) out[i] = mt_res53(&m);
}

/* solver/sph.py:316-347 _seed_positions on an explicit nx x ny x nz lattice:
 * loops k -> j -> i, one uniform(-jitter, jitter) draw each for x, y, z,
 * base + i*spacing + draw in float64, stored as float32 with w = 1.
 * base[] is the (float32-valued) lattice origin lo + spacing/2. */
void orc_seed_lattice(const double base[3], double spacing, double jitter, int nx, int ny, int nz,
                      uint32_t seed, float* out4) {
  Mt m;
  uint32_t key[1] = {seed};
  mt_init_by_array(&m, key, 1);
  long c = 0;
  double a = -jitter, b = jitter;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        double x = base[0] + i * spacing + (a + (b - a) * mt_res53(&m));
        double y = base[1] + j * spacing + (a + (b - a) * mt_res53(&m));
        double z = base[2] + k * spacing + (a + (b - a) * mt_res53(&m));
        out4[4 * c] = (float)x;
        out4[4 * c + 1] = (float)y;
        out4[4 * c + 2] = (float)z;
        out4[4 * c + 3] = 1.0f;
        ++c;
      }
}

/* ---- marching cubes (solver/marching_cubes.py) --------------------------- */
static const int MC_CORNERS[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                     {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
static const int MC_EDGES[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                    {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
static const int MC_FACES[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                   {3, 2, 6, 7}, {0, 3, 7, 4}, {1, 2, 6, 5}};

static int8_t g_tri[256][16];   /* edge triples, -1 terminated */
static uint16_t g_edge_mask[256];
static int g_walk[12][3];       /* lo corner, hi corner, axis */
static int g_table_ready = 0;

static int mc_edge_of(int a, int b) {
  for (int e = 0; e < 12; ++e)
    if ((MC_EDGES[e][0] == a && MC_EDGES[e][1] == b) || (MC_EDGES[e][0] == b && MC_EDGES[e][1] == a))
      return e;
  return -1;
}

/* solver/marching_cubes.py:60-87 _newell_normal / _oriented_faces */
static void mc_oriented_faces(int faces[6][4]) {
  for (int f = 0; f < 6; ++f) {
    double pts[4][3];
    for (int c = 0; c < 4; ++c)
      for (int a = 0; a < 3; ++a) pts[c][a] = MC_CORNERS[MC_FACES[f][c]][a];
    double nx = 0, ny = 0, nz = 0;
    for (int i = 0; i < 4; ++i) {
      const double* A = pts[i];
      const double* B = pts[(i + 1) % 4];
      nx += (A[1] - B[1]) * (A[2] + B[2]);
      ny += (A[2] - B[2]) * (A[0] + B[0]);
      nz += (A[0] - B[0]) * (A[1] + B[1]);
    }
    double cx = (pts[0][0] + pts[1][0] + pts[2][0] + pts[3][0]) / 4.0 - 0.5;
    double cy = (pts[0][1] + pts[1][1] + pts[2][1] + pts[3][1]) / 4.0 - 0.5;
    double cz = (pts[0][2] + pts[1][2] + pts[2][2] + pts[3][2]) / 4.0 - 0.5;
    double outward = nx * cx + ny * cy + nz * cz;
    for (int c = 0; c < 4; ++c) faces[f][c] = outward > 0.0 ? MC_FACES[f][c] : MC_FACES[f][3 - c];
  }
}

/* solver/marching_cubes.py:90-119 _face_segments */
static int mc_face_segments(const int face[4], const int inside[8], int seg[][2]) {
  int tr_open[4], tr_edge[4], count = 0;
  for (int idx = 0; idx < 4; ++idx) {
    int a = face[idx], b = face[(idx + 1) % 4];
    if (inside[a] && !inside[b]) {
      tr_open[count] = 1;
      tr_edge[count++] = mc_edge_of(a, b);
    } else if (inside[b] && !inside[a]) {
      tr_open[count] = 0;
      tr_edge[count++] = mc_edge_of(a, b);
    }
  }
  int ns = 0;
  for (int idx = 0; idx < count; ++idx) {
    if (!tr_open[idx]) continue;
    for (int step = 1; step < count; ++step) {
      int o = (idx + step) % count;
      if (!tr_open[o]) {
        seg[ns][0] = tr_edge[idx];
        seg[ns][1] = tr_edge[o];
        ++ns;
        break;
      }
    }
  }
  return ns;
}

/* solver/marching_cubes.py:122-162 _build_case_table, :166-173 EDGE_TABLE,
 * :180-188 _build_edge_walk.  The successor dict is an insertion-ordered
 * list here; a key is assigned at most once (each sign-change edge opens a
 * segment on exactly one face), so insertion order = first assignment. */
static void mc_build_tables(void) {
  if (g_table_ready) return;
  int faces[6][4];
  mc_oriented_faces(faces);
  for (int cs = 0; cs < 256; ++cs) {
    int inside[8];
    for (int c = 0; c < 8; ++c) inside[c] = (cs >> c) & 1;
    int succ_key[12], succ_val[12], nsucc = 0;
    for (int f = 0; f < 6; ++f) {
      int seg[4][2];
      int ns = mc_face_segments(faces[f], inside, seg);
      for (int s = 0; s < ns; ++s) {
        int found = -1;
        for (int q = 0; q < nsucc; ++q)
          if (succ_key[q] == seg[s][0]) found = q;
        if (found >= 0) {
          succ_val[found] = seg[s][1];
        } else {
          succ_key[nsucc] = seg[s][0];
          succ_val[nsucc++] = seg[s][1];
        }
      }
    }
    int visited[12] = {0}, nt = 0;
    memset(g_tri[cs], -1, sizeof(g_tri[cs]));
    for (int q = 0; q < nsucc; ++q) {
      int first = succ_key[q];
      if (visited[first]) continue;
      int loop[12], nl = 0, edge = first;
      while (!visited[edge]) {
        visited[edge] = 1;
        loop[nl++] = edge;
        int nxt = -1;
        for (int r = 0; r < nsucc; ++r)
          if (succ_key[r] == edge) nxt = succ_val[r];
        edge = nxt;
      }
      for (int i = 1; i < nl - 1; ++i) {
        g_tri[cs][3 * nt] = (int8_t)loop[0];
        g_tri[cs][3 * nt + 1] = (int8_t)loop[i + 1];
        g_tri[cs][3 * nt + 2] = (int8_t)loop[i];
        ++nt;
      }
    }
    uint16_t mask = 0;
    for (int e = 0; e < 12; ++e)
      if (inside[MC_EDGES[e][0]] != inside[MC_EDGES[e][1]]) mask |= (uint16_t)(1u << e);
    g_edge_mask[cs] = mask;
  }
  for (int e = 0; e < 12; ++e) {
    int a = MC_EDGES[e][0], b = MC_EDGES[e][1], axis = 0;
    while (MC_CORNERS[a][axis] == MC_CORNERS[b][axis]) ++axis;
    if (MC_CORNERS[a][axis] < MC_CORNERS[b][axis]) {
      g_walk[e][0] = a;
      g_walk[e][1] = b;
    } else {
      g_walk[e][0] = b;
      g_walk[e][1] = a;
    }
    g_walk[e][2] = axis;
  }
  g_table_ready = 1;
}

void orc_mc_tables(int8_t* tri_out /* 256*16 */, uint16_t* edge_mask_out /* 256 */,
                   int32_t* walk_out /* 12*3 */) {
  mc_build_tables();
  memcpy(tri_out, g_tri, sizeof(g_tri));
  memcpy(edge_mask_out, g_edge_mask, sizeof(g_edge_mask));
  memcpy(walk_out, g_walk, sizeof(g_walk));
}

/* solver/marching_cubes.py:191-280 extract().  field is float32 (the
 * reference reads back an R32F texture); values, iso and the vertex
 * arithmetic are float64.  Writes up to vcap vertices / tcap triangles,
 * returns the totals in counts[0..1]; cases (optional) receives every cube's
 * case index.  Returns 0, or -1 if the edge map cannot be allocated. */
int orc_mc_extract(const float* field, const int32_t dims[3], const double origin[3],
                   double spacing, double iso, double* verts, long vcap, int32_t* tris,
                   long tcap, long* counts, uint8_t* cases) {
  mc_build_tables();
  int nx = dims[0], ny = dims[1], nz = dims[2];
  counts[0] = counts[1] = 0;
  if (nx < 2 || ny < 2 || nz < 2) return 0;
  long nxy = (long)nx * ny, S = nxy * nz;
  int32_t* vmap = (int32_t*)malloc((size_t)S * 3 * sizeof(int32_t));
  if (!vmap) return -1;
  for (long q = 0; q < 3 * S; ++q) vmap[q] = -1;
  long coff[8];
  for (int c = 0; c < 8; ++c)
    coff[c] = MC_CORNERS[c][0] + MC_CORNERS[c][1] * (long)nx + MC_CORNERS[c][2] * nxy;
  long nv = 0, nt = 0;
  for (int k = 0; k < nz - 1; ++k)
    for (int j = 0; j < ny - 1; ++j) {
      long row = j * (long)nx + k * nxy;
      for (int i = 0; i < nx - 1; ++i) {
        long base = row + i;
        double v[8];
        int cs = 0;
        for (int c = 0; c < 8; ++c) {
          v[c] = (double)field[base + coff[c]];
          if (v[c] >= iso) cs |= 1 << c;
        }
        if (cases) cases[((long)k * (ny - 1) + j) * (nx - 1) + i] = (uint8_t)cs;
        if (cs == 0 || cs == 255) continue;
        for (int t = 0; g_tri[cs][3 * t] >= 0; ++t) {
          int32_t face[3];
          for (int e3 = 0; e3 < 3; ++e3) {
            int edge = g_tri[cs][3 * t + e3];
            int lo_c = g_walk[edge][0], hi_c = g_walk[edge][1], axis = g_walk[edge][2];
            long key = (base + coff[lo_c]) * 3 + axis;
            int32_t index = vmap[key];
            if (index < 0) {
              double lo_v = v[lo_c], hi_v = v[hi_c];
              double delta = hi_v - lo_v;
              double tt = delta == 0.0 ? 0.5 : (iso - lo_v) / delta;
              double pt[3] = {origin[0] + (i + MC_CORNERS[lo_c][0]) * spacing,
                              origin[1] + (j + MC_CORNERS[lo_c][1]) * spacing,
                              origin[2] + (k + MC_CORNERS[lo_c][2]) * spacing};
              pt[axis] += tt * spacing;
              index = (int32_t)nv;
              if (nv < vcap) {
                verts[3 * nv] = pt[0];
                verts[3 * nv + 1] = pt[1];
                verts[3 * nv + 2] = pt[2];
              }
              ++nv;
              vmap[key] = index;
            }
            face[e3] = index;
          }
          if (nt < tcap) {
            tris[3 * nt] = face[0];
            tris[3 * nt + 1] = face[1];
            tris[3 * nt + 2] = face[2];
          }
          ++nt;
        }
      }
    }
  free(vmap);
  counts[0] = nv;
  counts[1] = nt;
  return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}
