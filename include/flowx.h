/* flowx.h - C ABI of libflowx_hip.so, the MI355X (gfx950) Position Based Fluids
 * step / reset / readback / surface path.
 *
 * The library replaces the GPU boundary of the Flow-X Blender add-on: the
 * texture + push-constant + gpu.compute.dispatch plumbing of solver/gpu_util.py
 * and the GLSL passes under shaders/ that solver/sph.py and solver/surface.py
 * drive through it.  Every entry point below names the reference interface it
 * replaces (file:line, relative to the reference repository root).
 *
 * Conventions
 *   - Every function returns int: 0 = ok, < 0 = FX_E* error code.  No C++
 *     exception or abort crosses the ABI; fx_last_error() returns a
 *     thread-local message describing the last failure on this thread.
 *   - All array arguments are DEVICE pointers owned by the caller (the Python
 *     host allocates torch tensors on the ROCm device and passes data_ptr()).
 *     The library allocates no device memory: scratch space is a caller-owned
 *     workspace (fx_workspace_bytes / fx_mc_workspace_bytes), and the
 *     marching-cubes table is compile-time __constant__ data.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).  All
 *     calls are asynchronous on that stream; the caller synchronises when it
 *     reads results back (the reference's texture.read(), gpu_util.py:62-89).
 *   - Vector particle state is AoS-4 float32, the reference texture layout
 *     (RGBA32F, solver/gpu_util.py:28-59): positions (x,y,z,1), velocities
 *     (x,y,z,0), predicted (x,y,z,1), lambda (density,lambda,0,0), delta
 *     (x,y,z,0), normal (nx,ny,nz,curvature).  Keys, indices and cell ranges
 *     are int32 (the reference stores them as exact-integer floats,
 *     shaders/sph_grid_key.glsl:8-11); an empty cell has start = end = -1
 *     (shaders/sph_cell_clear.glsl:8-16).
 */
#ifndef FLOWX_H
#define FLOWX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FX_ABI_VERSION 2

/* error codes */
#define FX_OK 0
#define FX_EINVAL -1   /* bad argument (null pointer, size, dims)            */
#define FX_ENOMEM -2   /* workspace / output capacity too small             */
#define FX_EHIP -3     /* HIP runtime error (launch or API call failed)     */
#define FX_ERANGE -4   /* a count exceeds the int32 index range              */

/* Solver parameters as the GPU sees them: float32 / int32, one field per
 * lane of the reference's 128-byte push-constant block
 * (shaders/sph_common.glsl:10-17, solver/sph.py:433-461).  The two values the
 * reference bit-packs into i_sort.zw (scorr_k, surface_tension) and
 * i_collider.w (voxel size) are plain floats here.  n is at most 2^28 - 1
 * (a batch's total too): larger counts are refused with FX_ERANGE. */
typedef struct FxParams {
  int32_t n;                 /* i_layout.x  particle count                        */
  int32_t cells[3];          /* i_grid.xyz  cell dims (x fastest)                 */
  int32_t n_cells;           /* i_grid.w    cells[0]*cells[1]*cells[2]            */
  float lo[3];               /* f_lo.xyz    domain min = grid / collider origin   */
  float cell_size;           /* f_lo.w                                            */
  float hi[3];               /* f_hi.xyz    domain max                            */
  float particle_radius;     /* f_hi.w      spacing / 2                           */
  float h;                   /* f_sph.x     smoothing radius                      */
  float mass;                /* f_sph.y                                           */
  float rest_density;        /* f_sph.z                                           */
  float relaxation;          /* f_sph.w     CFM epsilon of the lambda denominator */
  float viscosity;           /* f_sim.x     XSPH blend                            */
  float dt;                  /* f_sim.y                                           */
  float gravity;             /* f_sim.z     -9.81                                 */
  float boundary_damping;    /* f_sim.w     0.35                                  */
  float scorr_k;             /* i_sort.z    s_corr strength (0 = off)             */
  float surface_tension;     /* i_sort.w    cohesion sigma (0 = normal pass off)  */
  int32_t collider_dims[3];  /* i_collider.xyz                                    */
  float collider_voxel;      /* i_collider.w  0 = no collider                     */
  int32_t iterations;        /* Jacobi iterations per substep (solver/sph.py:510) */
  int32_t max_cell_scan;     /* MAX_CELL_SCAN (shaders/sph_common.glsl:27-30)     */
  int32_t flags;             /* FX_FLAG_* (no reference counterpart; 0 = the      */
                             /* reference's exact pass semantics)                 */
} FxParams;

/* Position correction weighted by the particle mass,
 * dp_i = (m / rho0) sum_j (lambda_i + lambda_j + s_corr) gradW_ij.  The
 * reference's shaders/sph_delta.glsl:81-84 omits m while its lambda
 * (shaders/sph_lambda.glsl:58-74) includes it, so its Jacobi step is 1/m times
 * the Newton step and diverges whenever m < 1 (every default scene).  Off by
 * default: set it only for runs that need the constraint to converge. */
#define FX_FLAG_MASS_WEIGHTED_DELTA 1

/* Surface lattice of the density splat (solver/surface.py:97-140 resolve(),
 * repurposed slots of shaders/surface_splat.glsl:13-27). */
typedef struct FxSurface {
  int32_t dims[3];           /* samples per axis (x fastest)                      */
  float lo[3];               /* lattice low corner (domain lo - kernel radius)    */
  float spacing;             /* sample spacing                                    */
  float kernel_radius;       /* splat Poly6 support                               */
} FxSurface;

/* Particle state.  Every pointer is a caller-owned device buffer.  pos and
 * vel are the persistent state; the others mirror the reference's scratch
 * textures (solver/sph.py:123-158) and may be NULL for fx_substep / fx_step,
 * in which case that canonical copy is not written.  `work` holds the
 * library's internal sorted-order state (see fx_workspace_bytes); its
 * contents carry from one call to the next like the reference's scratch
 * textures do (the normal pass and the splat read the previous grid). */
typedef struct FxState {
  float* pos;                /* n x 4  positions_img                              */
  float* vel;                /* n x 4  velocities_img                             */
  float* pred;               /* n x 4  predicted_img                              */
  float* lambda;             /* n x 4  lambda_img  (density, lambda, 0, 0)        */
  float* delta;              /* n x 4  delta_img                                  */
  float* normal;             /* n x 4  normal_img  (normal, curvature)            */
  int32_t* key;              /* n      canonical cell key of particle i           */
  int32_t* sorted_key;       /* n      keys_img.x  (key of sorted slot s)         */
  int32_t* sorted_idx;       /* n      keys_img.y  (particle in sorted slot s)    */
  int32_t* cell_start;       /* n_cells  cell_start_img  (-1 = empty)             */
  int32_t* cell_end;         /* n_cells  cell_end_img    (-1 = empty)             */
  const uint8_t* collider;   /* collider_dims product, 1 = solid, or NULL         */
  void* work;                /* fx_workspace_bytes(params) bytes                  */
  size_t work_bytes;
} FxState;

/* Reference passes, in solver/sph.py:_PASSES order (solver/sph.py:161-174). */
enum FxPass {
  FX_PASS_NORMAL = 0,        /* shaders/sph_normal.glsl        */
  FX_PASS_PREDICT = 1,       /* shaders/sph_predict.glsl       */
  FX_PASS_GRID_KEY = 2,      /* shaders/sph_grid_key.glsl      */
  FX_PASS_SORT = 3,          /* shaders/sph_sort.glsl (whole sort) */
  FX_PASS_CELL_CLEAR = 4,    /* shaders/sph_cell_clear.glsl    */
  FX_PASS_CELL_RANGE = 5,    /* shaders/sph_cell_range.glsl    */
  FX_PASS_LAMBDA = 6,        /* shaders/sph_lambda.glsl        */
  FX_PASS_DELTA = 7,         /* shaders/sph_delta.glsl         */
  FX_PASS_APPLY_DELTA = 8,   /* shaders/sph_apply_delta.glsl   */
  FX_PASS_VELOCITY = 9,      /* shaders/sph_velocity.glsl      */
  FX_PASS_XSPH = 10,         /* shaders/sph_xsph.glsl          */
  FX_PASS_FINALIZE = 11,     /* shaders/sph_finalize.glsl      */
  FX_PASS_COUNT = 12
};

/* ---- library ----------------------------------------------------------- */

/* Replaces the error reporting of solver/gpu_util.py:92-103 / sph.py:926-933. */
const char* fx_last_error(void);
int fx_abi_version(void);

/* Per-kernel timing with HIP events (no reference counterpart; the reference
 * times only whole frames, solver/sph.py:517-545).  fx_profile(1) starts
 * bracketing every kernel launch with an event pair on its stream;
 * fx_profile_report() waits for the recorded events and writes one line
 * "<kernel> <launches> <total_ms>" per kernel into buf, then clears the
 * record.  Leave it off under HIP graph capture. */
int fx_profile(int enable);

/* Launch tuning of the neighbour passes (no reference counterpart), set by
 * the autotuner.  list_skin: neighbour-list margin in units of h (0 = always
 * scan the full 27-cell stencil); block: 64, 128 or 256 threads; walk: how a
 * large scene's neighbour passes stage the stencil's cell rows into batches
 * of loads in flight (0: per pass default, with lambda recording which
 * candidates lie within h and the same iteration's delta visiting only
 * those, 1: row by row, 2: the other stream walk, 3: lambda and delta read
 * their candidates from LDS tiles of the workgroup's stencil rows,
 * workgroups of 256 particles, 4: the per pass default without the recorded
 * candidates); persistent: the batched passes' launch, 0 = one workgroup per block
 * of particles, k > 0 = k workgroups per CU looping over them (XCD-contiguous
 * ranges).  Results do not depend on any of them. */
typedef struct FxTuning {
  float list_skin;
  int32_t block;
  int32_t persistent;
  int32_t walk;
} FxTuning;
int fx_set_tuning(const FxTuning* t);
int fx_get_tuning(FxTuning* t);
int fx_profile_report(char* buf, size_t len);

/* Bytes of `work` the solver needs for these params (internal sorted-order
 * buffers, counters, scan partials).  Replaces the scratch allocations of
 * solver/sph.py:350-369 (_allocate). */
size_t fx_workspace_bytes(const FxParams* p);

/* ---- single passes, canonical layout, reference semantics ---------------
 * One GLSL pass each, reading and writing the FxState buffers exactly as the
 * pass reads and writes its textures; used by parity tests and by callers that
 * drive passes one at a time like solver/sph.py:_substep (sph.py:501-514).
 * FX_PASS_SORT sorts state->key into sorted_key / sorted_idx, ascending by
 * key and, within a key, ascending by particle index (a deterministic
 * counting sort; the reference's bitonic network orders equal keys
 * differently, shaders/sph_sort.glsl:29-33).  The neighbour passes accept any
 * permutation in sorted_idx that is sorted by key. */
int fx_pass(int pass_id, const FxParams* p, FxState* s, void* stream);

/* ---- the grid build (solver/sph.py:482-498, _build_grid) -----------------
 * grid_key + sort + cell_clear + cell_range from state->pred, and the
 * sorted-order copies the neighbour passes read (kept in `work`). */
int fx_build_grid(const FxParams* p, FxState* s, void* stream);

/* ---- the substep and the frame (solver/sph.py:501-545) -------------------
 * fx_substep runs one substep in the reference's pass order (normal if
 * surface_tension > 0, predict, grid build, iterations x (lambda, delta,
 * apply), velocity, xsph, finalize) with passes fused where the data flow
 * allows; fx_step runs `substeps` of them (solver/sph.py:529-530).  After
 * fx_step every caller buffer holds what `substeps` fx_substep calls would
 * leave, bit for bit; in between, the substeps keep their state in the
 * workspace in sorted order and only the last writes the canonical buffers
 * and the optional scratch copies. */
int fx_substep(const FxParams* p, FxState* s, void* stream);
int fx_step(const FxParams* p, FxState* s, int substeps, void* stream);

/* ---- reset / cache load (solver/sph.py:350-369 _allocate, :700-733
 * _apply_cached): pred = pos, lambda = 0, velocity w lanes set, then the grid
 * is built from pred (the reference's _build_grid(0.0)).  pos/vel are read
 * from s->pos / s->vel. */
int fx_load_state(const FxParams* p, FxState* s, void* stream);

/* ---- batched scenes (no reference counterpart; the reference runs one
 * scene, solver/sph.py:501-514).  Many independent scenes advance in one
 * launch per pass: particles are concatenated in scene order (scene i owns
 * the next scenes[i].params.n rows of pos / vel), cells are offset into one
 * key space, one counting sort orders every scene, and each thread reads its
 * scene's parameters from a table the library keeps in `work`.  Scenes may
 * differ in everything, iteration count included: the Jacobi launches run
 * as many times as the scene with the most iterations needs, and a scene
 * past its own count carries its state through the rest unchanged (results
 * equal the scene's own plain substep).  fx_batch_load_state uploads the table
 * (synchronously) and builds the grid from pos (pred = pos); call it again
 * whenever a scene's parameters change.  fx_batch_substep reads only device
 * memory and may be captured in a HIP graph. */
typedef struct FxScene {
  FxParams params;           /* the scene's parameters; params.n = its particles  */
  int64_t collider_offset;   /* first byte of its grid in `colliders`, -1 = none  */
} FxScene;

size_t fx_batch_workspace_bytes(const FxScene* scenes, int count);
int fx_batch_load_state(const FxScene* scenes, int count, float* pos, float* vel,
                        const uint8_t* colliders, void* work, size_t work_bytes, void* stream);
int fx_batch_substep(const FxScene* scenes, int count, float* pos, float* vel,
                     const uint8_t* colliders, void* work, size_t work_bytes, void* stream);

/* ---- differentiable substep (no reference counterpart) -------------------
 * A substep whose neighbour structure is recorded once after predict (pairs
 * within h*(1 + list_skin) of the predicted positions, each list sorted by
 * index) and whose collider voxel frame moves with `translation`, in float
 * (precision 4) or double (precision 8).  State arrays are N x 4 of that
 * precision (x, y, z, unused).  fx_adjoint_backward recomputes the substep
 * from its checkpoint (x, v), keeping every Jacobi iterate, and writes the
 * adjoints of x and v given those of the outputs; grad_params (device
 * double[4]) accumulates d/dgravity and d/dtranslation.  surface_tension and
 * scorr_k must be 0.  fx_adjoint_overflow reports whether a neighbour list
 * exceeded its 192 entries in the last forward.
 * decisions (device memory of fx_adjoint_record_bytes, or NULL): the
 * substep's record - its discrete branches (collider push-out and the chosen
 * voxel, velocity reflection, wall clamps, speed clamp) and its neighbour
 * lists - written by a forward with decision_mode 1 and replayed with mode 2
 * (mode 0: decide and build lists in the workspace, no record).  Replaying
 * freezes the branches and the neighbour structure of a perturbed run, so
 * finite differences see the same smooth piece the adjoint differentiates,
 * and skips the list build; fx_adjoint_backward always replays when
 * decisions is non-NULL. */
size_t fx_adjoint_workspace_bytes(const FxParams* p, int precision);
size_t fx_adjoint_record_bytes(const FxParams* p);
int fx_adjoint_forward(const FxParams* p, int precision, double gravity, const double translation[3],
                       double list_skin, const uint8_t* collider, int32_t* decisions,
                       int decision_mode, const void* x, const void* v, void* x_out, void* v_out,
                       void* work, size_t work_bytes, void* stream);
int fx_adjoint_backward(const FxParams* p, int precision, double gravity, const double translation[3],
                        double list_skin, const uint8_t* collider, int32_t* decisions,
                        const void* x, const void* v, const void* gx_out, const void* gv_out,
                        void* gx, void* gv, double* grad_params, void* work, size_t work_bytes,
                        void* stream);
int fx_adjoint_overflow(const FxParams* p, int precision, const void* work, int32_t* host_flag,
                        void* stream);
/* Splat of the density field (shaders/surface_splat.glsl arithmetic, sample
 * zeroed where sample_mask is set) at the given precision, and its adjoint:
 * gx += d field / d x ^T gfield. */
int fx_adjoint_splat(const FxParams* p, const FxSurface* surf, int precision, const void* x,
                     const uint8_t* sample_mask, void* field, void* work, size_t work_bytes,
                     void* stream);
int fx_adjoint_splat_backward(const FxParams* p, const FxSurface* surf, int precision, const void* x,
                              const uint8_t* sample_mask, const void* gfield, void* gx, void* stream);

/* ---- surface (solver/surface.py:201-251 update, shaders/surface_splat.glsl)
 * Density field at every surface sample, gathered through the last grid from
 * the finalized positions s->pos (x fastest, S = dims product). */
int fx_splat(const FxParams* p, const FxSurface* surf, const FxState* s,
             float* field, void* stream);

/* Marching cubes (solver/marching_cubes.py:191-280 extract) on the GPU.
 * field: dims product float32, x fastest; sample (i,j,k) sits at
 * origin + (i,j,k)*spacing.  Samples >= iso are inside.  Vertices are
 * computed in float64 and rounded once to float32 (verts: 3 floats each);
 * tris: 3 int32 vertex indices each.  counts (device int64[2]) receives the
 * total vertex and triangle counts even when they exceed vcap / tcap (then
 * only the first vcap / tcap are written and the call returns FX_ENOMEM
 * after the launch; call again with larger buffers).  `cases` (device
 * uint8, (nx-1)(ny-1)(nz-1) cubes) receives every cube's case index if
 * non-NULL.  work: fx_mc_workspace_bytes(dims). */
size_t fx_mc_workspace_bytes(const int32_t dims[3]);
int fx_mc_extract(const float* field, const int32_t dims[3], const double origin[3],
                  double spacing, float iso, float* verts, int64_t vcap, int32_t* tris,
                  int64_t tcap, int64_t* counts, uint8_t* cases, void* work,
                  size_t work_bytes, void* stream);

/* Marching cubes with no workspace argument: fx_mc_extract's arguments and
 * results, its scratch in a fixed device array of the library instead of
 * `work`.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc).  It allocates nothing.  Limit: dims product <= 2^21
 * samples (twice the reference's 2^20 surface cap, solver/surface.py:49);
 * larger lattices return FX_ERANGE.  Each GPU holds one copy of the array
 * (about 22 MB) until the process ends, shared by every call on that GPU:
 * calls that overlap on one GPU race. */
int fx_mc_extract_static(const float* field, const int32_t dims[3], const double origin[3],
                         double spacing, float iso, float* verts, int64_t vcap, int32_t* tris,
                         int64_t tcap, int64_t* counts, uint8_t* cases, void* stream);

/* Surface hand-off (solver/surface.py:278-301): the triangles bmesh would
 * accept, in order - a triangle whose vertex set (any winding) already
 * appeared earlier is dropped, as bmesh.faces.new rejects it.  tris: device
 * int32 ntris x 3 with vertex indices < nverts <= 2^21; out: device int32
 * ntris x 3; *count (device int32) receives the kept count.
 * work: fx_mc_dedup_workspace_bytes(ntris). */
size_t fx_mc_dedup_workspace_bytes(int64_t ntris);
int fx_mc_dedup(const int32_t* tris, int64_t ntris, int64_t nverts, int32_t* out, int32_t* count,
                void* work, size_t work_bytes, void* stream);
/* The derived case table behind fx_mc_extract (edge triples per case, -1
 * terminated, 16 per case) and the per-case sign-change edge masks: the
 * TRI_TABLE / EDGE_TABLE of solver/marching_cubes.py:162-173.  Host memory. */
int fx_mc_tables(int8_t* tri_out /* 256*16 */, uint16_t* edge_mask_out /* 256 */);

/* ---- collider occupancy (collision/__init__.py:245-260 geometry,
 * :304-367 voxelization, :225-242 union).  The voxel centre
 * origin + (i+0.5)*voxel decides occupancy; voxel index (k*ny + j)*nx + i.
 * `accumulate` != 0 ORs into grid (the union over colliders), 0 overwrites. */
int fx_voxelize_boxes(const int32_t dims[3], const float origin[3], float voxel,
                      const float* boxes /* host: k x (lo[3], hi[3]) */, int k,
                      uint8_t* grid, int accumulate, void* stream);
int fx_voxelize_sphere(const int32_t dims[3], const float origin[3], float voxel,
                       const float center[3], float radius, uint8_t* grid, int accumulate,
                       void* stream);
int fx_voxelize_cylinder(const int32_t dims[3], const float origin[3], float voxel,
                         const float base[3], float radius, float height, uint8_t* grid,
                         int accumulate, void* stream);
/* Mesh collider (collision/__init__.py:263-273 _point_inside, :304-367
 * _rebuild_grid): voxels whose centre lies inside [bounds_lo, bounds_hi) (the
 * object's world bounds, as _index_range rounds them) get the parity of the
 * upward ray's hit chain - hits at or above the centre, each next hit at
 * least ray_eps beyond the previous, at most 64 - against the triangles.
 * verts: device float32 nverts x 3 in world space; tris: device int32
 * ntris x 3.  ray_eps is the reference's 1e-4 local-space restart step in
 * world units (1e-4 / |inverse(matrix_world)_3x3 * z|).  *overflow (device,
 * optional) is set when a column has more than 128 hits (the extra hits are
 * dropped). */
int fx_voxelize_mesh(const int32_t dims[3], const float origin[3], float voxel, const float* verts,
                     int nverts, const int32_t* tris, int ntris, const float bounds_lo[3],
                     const float bounds_hi[3], float ray_eps, uint8_t* grid, int accumulate,
                     int32_t* overflow, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FLOWX_H */
