/*
 * exa.h -- C ABI of the MI355X-native exaStitch volume path tracer
 * (libexastitch.so).  Plain C types only; every entry point returns an int
 * status (EXA_OK == 0, negative on error) and never throws.  The message of
 * the last failure on the calling thread is exa_get_last_error().
 *
 * The boundary mirrors the reference's three host-visible layers:
 *   - the model      exa::ExaStitchModel   (model/ExaStitchModel.h:24-57,
 *                                           model/Model.h:29-65)
 *   - the sampler    exa::ExaStitchSampler (sampler/Sampler.h:28-69,
 *                                           sampler/ExaStitchSampler.h:37-141)
 *   - the renderer   exa::OWLRenderer      (OWLRenderer.h:29-175)
 * plus the offline dual-mesh tools (tools/amrMakeDualMesh) and synthetic
 * scene generators.  Device pointers are hipMalloc'd memory of the current
 * device; `stream` is a hipStream_t passed as void* (NULL = default stream).
 * Objects are not thread-safe (the reference's OWLRenderer is used from one
 * thread, OWLRenderer.cpp).
 */
#ifndef EXA_EXA_H
#define EXA_EXA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EXA_OK 0
#define EXA_ERR_INVALID (-1) /* bad argument or invalid model data */
#define EXA_ERR_HIP (-2)     /* HIP runtime failure */
#define EXA_ERR_STATE (-3)   /* call order (e.g. locate before build) */
#define EXA_ERR_ALLOC (-4)   /* host or device allocation failed */

/* Gridlet.h:22-37 (32 bytes) */
typedef struct {
  int32_t lower[3];
  int32_t level;
  int32_t dims[3];
  uint32_t begin;
} exa_gridlet;

/* sampler/Sampler.h:72-76 (12 bytes) */
typedef struct {
  int32_t primID;
  int32_t cellID;
  float value;
} exa_sample;

/* tools/amrMakeDualMesh/makeDual.cpp:75-84 LogicalCell (16 bytes) */
typedef struct {
  int32_t pos[3];
  int32_t level;
} exa_amr_cell;

const char *exa_get_last_error(void);
int exa_version(void);

/* ---- synthetic AMR scenes (no reference counterpart: generators of the
        benchmark inputs; the reference reads .cells/.scalars files) ---- */
typedef struct exa_amr exa_amr;
int exa_amr_create_gaussians(int baseRes, int coarseLevel, int fineLo, int fineHi, int numGaussians,
                             uint64_t seed, exa_amr **out);
int exa_amr_create_noise(int baseRes, int baseLevel, float thresh1, float thresh2, int octaves,
                         float contrast, uint64_t seed, exa_amr **out);
/* Expanding spherical shock shell with turbulence in its wake (the large
   configuration: baseRes 256, baseLevel 4, radius 1200, thickness 40 over
   [0,4096]^3): shell at level 1, level-0 patches where a low-frequency
   noise exceeds thetaCore, inside the shell levels 2 and 3 split along the
   iso-surface thetaWake of a cell-scale turbulence (level 3 below it; about
   16% stitching elements per leaf cell at the default size with thetaWake
   0.43); leaf levels follow a graded distance field, so touching leaves
   differ by at most one level.  Scalar = shock profile x turbulence + wake
   turbulence, in [0,1]. */
int exa_amr_create_shock(int baseRes, int baseLevel, float radius, float thickness, float thetaCore,
                         float thetaWake, uint64_t seed, exa_amr **out);
/* leaf cells per level (counts[16]) of that scene, without building it */
int exa_amr_count_shock(int baseRes, int baseLevel, float radius, float thickness, float thetaCore,
                        float thetaWake, uint64_t seed, int64_t counts[16]);
int exa_amr_create_from_cells(const exa_amr_cell *cells, const float *scalars, int64_t n, exa_amr **out);
int64_t exa_amr_num_cells(const exa_amr *a);
int exa_amr_get(const exa_amr *a, exa_amr_cell *cells, float *scalars);
int exa_amr_get_params(const exa_amr *a, float *out, int maxCount);
void exa_amr_destroy(exa_amr *a);

/* ---- dual mesh: replaces amrMakeDualMesh + amrMakeGrids
        (tools/amrMakeDualMesh/makeDual.cpp:596-1000, makeGrids.cpp:46-343) ---- */
typedef struct exa_dual_mesh exa_dual_mesh;
int exa_dual_mesh_build(const exa_amr *a, exa_dual_mesh **out);
/* counts[11]: tets, pyramids, wedges, hexes, elements, vertices, gridlets,
   gridlet scalars, perfect cubes, conflicts, degenerate cells dropped */
int exa_dual_mesh_counts(const exa_dual_mesh *d, int64_t counts[11]);
int exa_dual_mesh_get(const exa_dual_mesh *d, int32_t *indices, float *vertices, exa_gridlet *gridlets,
                      float *gridletScalars);
void exa_dual_mesh_destroy(exa_dual_mesh *d);

/* ---- on-disk formats ----
   .cells: 16-byte {int32 x, y, z, level} records (makeDual.cpp:932-940);
   .scalars: raw float32 (ExaStitchModel.cpp:44-54); _<L>.cubes: 48-byte
   {float lower[3]; int32 level; int32 scalarIDs[8]} (makeDual.cpp:445-449,
   893-905); .grids: {int32 lower[3], level, dims[3], scalarIDs[(dims+1)^3]}
   per gridlet (makeGrids.cpp:240-246); .xf: headless.cpp:258-283.  The
   element mesh (the reference's umesh file) uses this library's EXAELEM1
   format: vertices, per-vertex scalar IDs and/or values, tets, pyramids,
   wedges, hexes. */
/* reads a .cells file and its .scalars (one per cell) */
int exa_amr_load(const char *cellsPath, const char *scalarsPath, exa_amr **out);
int exa_amr_save(const exa_amr *a, const char *cellsPath, const char *scalarsPath);
/* amrMakeDualMesh + amrMakeGrids to files (makeDual.cpp:908-960,
   makeGrids.cpp:262-343): the element mesh (with scalar IDs and values per
   vertex), optionally <cubesPrefix>_<L>.cubes and the .grids file; NULL
   paths are skipped.  counts[11] as exa_dual_mesh_counts (may be NULL). */
int exa_amr_write_dual_mesh(const exa_amr *a, const char *elementMeshPath, const char *cubesPrefix,
                            const char *gridsPath, int64_t counts[11]);
/* amrMakeGrids (makeGrids.cpp:262-343): bricks of each cubes file (level
   from its "_<L>.cubes" suffix), appended to gridsPath in argument order;
   *conflicts (may be NULL) counts inconsistent brick-vertex writes */
int exa_make_grids(const char *const *cubesPaths, int numPaths, const char *gridsPath, int64_t *conflicts);
/* ExaStitchModel::load(umesh, grids, scalars) (ExaStitchModel.cpp:24-257):
   any path may be NULL or empty.  Vertex value = scalars[vertexTag] when
   both exist, else the stored per-vertex value; unreferenced vertices are
   compacted when a grids file is given; a gridlet scalar ID beyond the
   scalars file is a hole (NaN).  Unlike the reference, a path that cannot
   be read, a truncated record or a vertex tag beyond the scalars is an
   error. */
typedef struct exa_model exa_model;
int exa_model_load(const char *elementMeshPath, const char *gridsPath, const char *scalarsPath, exa_model **out);
/* writes a .xf file (read with exa_xf_load): params = opacityScale,
   absDomain lo/hi, relDomain lo/hi */
int exa_xf_write(const char *path, const float params[5], const float *rgba, int numColors);

/* ---- model: exa::ExaStitchModel (ExaStitchModel.cpp:24-257) ---- */
typedef struct {
  const float *vertices; /* float4: xyz + scalar */
  int64_t numVertices;
  const int32_t *indices; /* int[8] per element, -1 padded; tets, pyrs, wedges, hexes */
  int64_t numElements;
  const exa_gridlet *gridlets;
  int64_t numGridlets;
  const float *gridletScalars; /* NaN = empty */
  int64_t numGridletScalars;
} exa_stitch_model_desc;
int exa_model_create(const exa_stitch_model_desc *desc, exa_model **out); /* copies */
int exa_model_create_from_dual_mesh(const exa_dual_mesh *d, exa_model **out);
/* cellBounds (lo xyz, hi xyz) and valueRange (Model.h:42-43) */
int exa_model_get_bounds(const exa_model *m, float cellBounds[6], float valueRange[2]);
int exa_model_get_counts(const exa_model *m, int64_t counts[4]); /* vertices, elements, gridlets, scalars */
/* copies the model buffers (ExaStitchModel.h members vertices, indices,
   gridlets, gridletScalars), sized by exa_model_get_counts; NULL skips one */
int exa_model_get(const exa_model *m, float *vertices, int32_t *indices, exa_gridlet *gridlets,
                  float *gridletScalars);
int exa_model_set_num_grid_cells(exa_model *m, const int32_t numMCs[3]);                     /* Model.cpp:25-33 */
int exa_model_set_voxel_space_transform(exa_model *m, const float from[6], const float to[6]); /* Model.cpp:35-48 */
void exa_model_destroy(exa_model *m);

/* ---- sampler: exa::ExaStitchSampler ---- */
typedef struct exa_sampler exa_sampler;
int exa_sampler_create(exa_sampler **out);
/* Sampler::build (ExaStitchSampler.cpp:23-340): uploads the model, builds the
   point-location BVHs and the majorant grid over the model's cellBounds with
   the model's numMCs (Grid.cu:457-594).  Fails with EXA_ERR_STATE when numMCs
   is unset or zero (ExaStitchSampler.cpp:326-327). */
int exa_sampler_build(exa_sampler *s, const exa_model *m, void *stream);
/* Sampler::computeMaxOpacities (ExaStitchSampler.cu:279-376): rgba is host
   float4[numColors]; xfRange = [lo, hi]. */
int exa_sampler_compute_max_opacities(exa_sampler *s, const float *rgba, int numColors, float lo, float hi,
                                      void *stream);
/* device sample() (ExaStitchSampler.h:128-140) over n device points (float3). */
int exa_sampler_locate(const exa_sampler *s, const float *d_pos, int64_t n, exa_sample *d_out, void *stream);
/* same, accumulating per-result counts into host counters[6]: gridlet, tet,
   pyramid, wedge, hex, miss (synchronises the stream). */
int exa_sampler_locate_counted(const exa_sampler *s, const float *d_pos, int64_t n, exa_sample *d_out,
                               uint64_t counters[6], void *stream);
/* exa_sampler_locate without any allocation, for callers that cannot
   allocate (code under HIP graph capture, tight serving loops, harnesses that
   forbid hipMalloc); no counterpart in the reference.  The queue of the
   element stage is a library-owned device array of EXA_LOCATE_NOALLOC_MAX + 1
   uint32 (64 MiB per GPU, held until the process ends); n beyond the maximum
   fails with EXA_ERR_INVALID.  Calls that overlap on one GPU (other streams,
   other host threads) share the array and race; results equal
   exa_sampler_locate's. */
#define EXA_LOCATE_NOALLOC_MAX (1 << 24)
int exa_sampler_locate_noalloc(const exa_sampler *s, const float *d_pos, int64_t n, exa_sample *d_out,
                               void *stream);
/* device testSample() (ExaStitchSampler.h:109-126) with per-point gridlet ids. */
int exa_sampler_test_sample(const exa_sampler *s, const float *d_pos, const int32_t *d_primIDs, int64_t n,
                            exa_sample *d_out, void *stream);
int exa_sampler_get_grid_info(const exa_sampler *s, int32_t dims[3], float bounds[6]);
int exa_sampler_get_value_ranges(const exa_sampler *s, float *out);          /* host, 2 per macrocell */
int exa_sampler_get_max_opacities(const exa_sampler *s, float *out);         /* host, 1 per macrocell */
int exa_sampler_get_gridlet_value_ranges(const exa_sampler *s, float *out);  /* host, 2 per gridlet */
int exa_sampler_get_prim_max_opacities(const exa_sampler *s, float *gridletOut, float *elementOut);
/* Point-location accelerator (replaces the OptiX sampling BVH,
   ExaStitchSampler.cpp:71-323): EXA_ACCEL_TABLES (default) = per-level
   macrocell lists of gridlets + uniform element bins; EXA_ACCEL_BVH = binary
   BVHs.  Both return the same samples.  Selecting an accelerator that was not
   built fails with EXA_ERR_STATE after build. */
#define EXA_ACCEL_TABLES 0
#define EXA_ACCEL_BVH 1
int exa_sampler_set_accel(exa_sampler *s, int accel);
int exa_sampler_get_accel(const exa_sampler *s);
/* info[8]: gridlet BVH nodes, depth, refs; element BVH nodes, depth, refs;
   device bytes of the model; device bytes of the accelerators */
int exa_sampler_get_info(const exa_sampler *s, int64_t info[8]);
void exa_sampler_destroy(exa_sampler *s);

/* ---- renderer: exa::OWLRenderer ---- */
typedef struct exa_renderer exa_renderer;
/* OWLRenderer(...) + initGPU (OWLRenderer.cpp:123-181, 309-582); the model is
   copied; numMCs = macrocell grid dims; remapFrom/remapTo as lo xyz, hi xyz
   (NULL = unit boxes). */
int exa_renderer_create(const exa_model *m, const int32_t numMCs[3], const float remapFrom[6],
                        const float remapTo[6], exa_renderer **out);
int exa_renderer_set_camera(exa_renderer *r, const float org[3], const float dir00[3], const float du[3],
                            const float dv[3]);
int exa_renderer_resize(exa_renderer *r, int width, int height);
/* render(uint32_t *fbPointer) (OWLRenderer.cpp:610-622): one launch of
   width*height pixels; d_fb may be NULL (accumulation only).  Asynchronous
   on `stream`; accumID advances by one. */
int exa_renderer_render(exa_renderer *r, uint32_t *d_fb, void *stream);
/* render() with the instrumented kernel; counters[11] receive: gridlet, tet,
   pyramid, wedge, hex hits, misses, cached gridlet hits, DDA cells, pixels,
   TF fetches, neural-field lookups (synchronises the stream). */
int exa_renderer_render_counted(exa_renderer *r, uint32_t *d_fb, uint64_t counters[11], void *stream);
int exa_renderer_get_accum(const exa_renderer *r, float *hostOut); /* float4 per pixel */
/* OWLRenderer::Type (OWLRenderer.h:32): 0 path tracer, 1 direct lighting,
   2 ray marcher (deviceCode.cu:1213-1260; runs in the one-thread-per-pixel
   kernel whatever the render scheme) */
int exa_renderer_set_type(exa_renderer *r, int type);
int exa_renderer_set_color_map(exa_renderer *r, const float *rgba, int numColors);
int exa_renderer_set_range(exa_renderer *r, float lo, float hi);
int exa_renderer_set_rel_domain(exa_renderer *r, float lo, float hi);
int exa_renderer_set_opacity_scale(exa_renderer *r, float scale); /* stores 1.1^(scale-100) */
int exa_renderer_set_clip_plane(exa_renderer *r, int id, int enabled, const float N[3], float d);
/* OWLRenderer::setShadeMode / shadeModes (OWLRenderer.cpp:769-774):
   0 default, 1 gridlets (checkerboard over gridlet cells, elements red),
   2 teaser (deviceCode.cu:625-697) */
int exa_renderer_set_shade_mode(exa_renderer *r, int mode);
/* ray-marcher step in voxel units (LaunchParams render.dt; the reference
   fixes it at 1, OWLRenderer.cpp:616) */
int exa_renderer_set_dt(exa_renderer *r, float dt);
/* OWLRenderer heatMapEnabled / heatMapScale (OWLRenderer.h:160-161,
   deviceCode.cu:1323-1327): each pixel blended with the colour of the clock
   cycles it took times scale / spp; frames with it run in the
   one-thread-per-pixel kernel */
int exa_renderer_set_heat_map(exa_renderer *r, int enabled, float scale);
int exa_renderer_set_light_source(exa_renderer *r, int id, const float pos[3], float intensity, int on);
/* build-side addition: a light at infinity in direction dir (towards the light) */
int exa_renderer_set_directional_light(exa_renderer *r, int id, const float dir[3], float intensity, int on);
int exa_renderer_set_light_space_transform(exa_renderer *r, const float xfm[12]);
int exa_renderer_set_sub_image(exa_renderer *r, const float box[4], int active);
int exa_renderer_set_sub_image_selection(exa_renderer *r, const float box[4], int active);
int exa_renderer_reset_accum(exa_renderer *r);
int exa_renderer_set_spp(exa_renderer *r, int spp);
/* deviceCode.cu:1050 hard-codes 1024 bounces; the default here is 1024 too */
int exa_renderer_set_max_bounces(exa_renderer *r, int maxBounces);
/* deviceCode.cu:1334 inverts the frame-buffer pixels of the centre row and
   column (DEBUGGING == 1 in common.h); on by default, as in the reference */
int exa_renderer_set_debug_crosshair(exa_renderer *r, int on);
int exa_renderer_get_bounds(const exa_renderer *r, float modelBounds[6]);
int exa_renderer_get_value_range(const exa_renderer *r, float valueRange[2]);
int exa_renderer_get_accum_id(const exa_renderer *r);
/* OWLRenderer::accumID is a public field (OWLRenderer.h:163); the next
   render() seeds its RNG with (pixelID, accumID) */
int exa_renderer_set_accum_id(exa_renderer *r, int accumID);
/* the exaStitch sampler the renderer built (its OWLRenderer::sampler,
   OWLRenderer.h:171); valid while the renderer lives */
const exa_sampler *exa_renderer_get_sampler(const exa_renderer *r);
/* selects the sampler's accelerator (building it from the renderer's model
   copy when needed) */
int exa_renderer_set_accel(exa_renderer *r, int accel);
/* Launch scheme of render(); both give bit-identical images (no reference
   counterpart: the reference launches one OptiX raygen thread per pixel,
   deviceCode.cu:1268-1360).
   EXA_RENDER_MEGAKERNEL: one thread per pixel, 16x16 tiles.
   EXA_RENDER_PERSISTENT (default): persistent lanes that take pixels from a
   work counter and run their paths as a state machine, in per-wave passes.
   Tuning: ctrlSteps = tracking iterations (one free flight, and the DDA
   step into the next macrocell when it leaves the cell) per lane and pass
   (>= 1); elemBatch and rareBatch in [0, 64] = wave fraction (x/64) of live
   lanes that must wait for an element test, resp. for per-segment work,
   before that pass runs.  Defaults 4, 16, 24. */
#define EXA_RENDER_MEGAKERNEL 0
#define EXA_RENDER_PERSISTENT 1
int exa_renderer_set_scheme(exa_renderer *r, int scheme);
int exa_renderer_get_scheme(const exa_renderer *r);
int exa_renderer_set_persistent_tuning(exa_renderer *r, int ctrlSteps, int elemBatch, int rareBatch);
/* Launch shape of the persistent scheme (no reference counterpart; images
   do not depend on it).  elementStage: where a path's element-stage lookup
   runs when its gridlet stage misses:
   EXA_ELEMENT_IN_WAVE (default): in its own wave, once elemBatch/64 of the
     wave's live lanes wait for it;
   EXA_ELEMENT_WORKGROUP_QUEUE: the lane posts the query to its workgroup's
     queue in LDS and waits; any wave of the workgroup answers up to 64
     posted queries at once when elemBatch (>= 1) of them are posted, or
     when all its own live lanes wait;
   EXA_ELEMENT_QUEUE_ROUNDS: the lane leaves its whole path in a device
     queue and takes other work; between launches the element stage runs over
     the queue with every lane busy and the next launch resumes the paths;
     queueRounds in [1, 32] such rounds per frame, the last launch running
     the element stage in its own lanes; queueDepth in [0, 2^24] entries per
     queue (0 = one per pixel, at most 2^21; paths that find it full run the
     element stage in their wave).
   The queues need the table accelerator without a neural field (otherwise
   the element stage runs in the wave).  blockSize: 64, 128 or 256 threads
   per workgroup.  Defaults: in wave, 1, 0, 256. */
#define EXA_ELEMENT_IN_WAVE 0
#define EXA_ELEMENT_WORKGROUP_QUEUE 1
#define EXA_ELEMENT_QUEUE_ROUNDS 2
int exa_renderer_set_launch_tuning(exa_renderer *r, int elementStage, int queueRounds, int queueDepth,
                                   int blockSize);
/* ctrlSteps, elemBatch, rareBatch, elementStage, queueRounds, queueDepth, blockSize */
int exa_renderer_get_tuning(const exa_renderer *r, int tuning[7]);

/* Snapshot of the launch parameters of the next render() call (for checking
   against the CPU oracle). */
typedef struct {
  int32_t W, H, spp, accumID;
  int32_t integrator, shadeMode, maxBounces, debugCrosshair;
  float camOrg[3], camDir00[3], camDu[3], camDv[3];
  float worldBounds[6];
  float voxelXfm[12], voxelXfmInv[12], lightXfm[12];
  int32_t gridDims[3];
  float gridBounds[6];
  float tfDomain[2];
  float opacityScale;
  int32_t tfN;
  int32_t clipEnabled;
  float clipN[3];
  float clipD;
  float lightPos[4][3];
  float lightIntensity[4];
  int32_t lightOn[4];
  int32_t lightType[4];
  float dt; /* ray-marcher step */
  int32_t heatMapEnabled;
  float heatMapScale;
  int32_t subImageActive; /* setSubImage: lower.xy, upper.xy in [0,1] */
  float subImage[4];
  int32_t subImageSelecting; /* setSubImageSelection */
  float subImageSel[4];
} exa_render_state;
int exa_renderer_get_state(const exa_renderer *r, exa_render_state *out);
void exa_renderer_destroy(exa_renderer *r);

/* ---- batched requests (build-side; no reference counterpart) ----
   Many views, transfer functions and image sizes of uneven size rendered in
   one persistent launch per frame, sharing the renderer's model, sampler and
   tables (the reference renders one OWLRenderer view per owlLaunch2D,
   OWLRenderer.cpp:610-622).  Each request keeps its own accumulation buffer
   and accumID (seeding as in deviceCode.cu:1278) and gives the image the
   renderer would give for the same settings.  Lights, clip plane, bounce cap,
   crosshair and model transforms are taken from the renderer when the
   request is added. */
typedef struct exa_batch exa_batch;
typedef struct {
  int32_t width, height, spp, integrator; /* integrator: 0 path tracer, 1 direct lighting */
  int32_t tf;                              /* index from exa_batch_add_transfer_function */
  float camOrg[3], camDir00[3], camDu[3], camDv[3];
} exa_batch_request;
int exa_batch_create(exa_renderer *r, exa_batch **out);
void exa_batch_destroy(exa_batch *b);
/* rgba: host float4[numColors]; [lo, hi] = absolute TF domain; opacityScale
   multiplies extinction directly (the renderer's setOpacityScale slider value
   s corresponds to 1.1^(s-100)). */
int exa_batch_add_transfer_function(exa_batch *b, const float *rgba, int numColors, float lo, float hi,
                                    float opacityScale, int *index);
int exa_batch_add_request(exa_batch *b, const exa_batch_request *req, int *index);
int exa_batch_num_requests(const exa_batch *b);
int exa_batch_set_tuning(exa_batch *b, int ctrlSteps, int elemBatch, int rareBatch);
int exa_batch_get_tuning(const exa_batch *b, int tuning[3]);
/* element stage, queue rounds, queue depth and workgroup size as
   exa_renderer_set_launch_tuning (depth 0: one entry per pixel of all
   requests, at most 2^21) */
int exa_batch_set_launch_tuning(exa_batch *b, int elementStage, int queueRounds, int queueDepth, int blockSize);
int exa_batch_get_launch_tuning(const exa_batch *b, int tuning[4]);
/* one frame of every request (asynchronous on stream); accumIDs advance */
int exa_batch_render(exa_batch *b, void *stream);
int exa_batch_render_counted(exa_batch *b, uint64_t counters[11], void *stream);
/* Captures one frame of the batch into a HIP graph on a non-default stream;
   exa_batch_render_graph replays it (the accumIDs live on the device and
   advance inside the graph).  Any change to the batch drops the graph. */
int exa_batch_capture_graph(exa_batch *b, void *stream);
int exa_batch_render_graph(exa_batch *b, void *stream);
int exa_batch_reset_accum(exa_batch *b);
int exa_batch_get_accum_id(const exa_batch *b, int req);
int exa_batch_get_accum(const exa_batch *b, int req, float *hostOut);     /* float4 per pixel */
int exa_batch_get_frame(const exa_batch *b, int req, uint32_t *hostOut);  /* RGBA8 per pixel */
/* Autotuner: one timed frame per combination of candidate knob values;
   keeps the fastest, returns it in best[3] (and its time), resets the
   accumulation. */
int exa_batch_autotune(exa_batch *b, const int *ctrlSteps, int nCtrl, const int *elemBatch, int nElem,
                       const int *rareBatch, int nRare, void *stream, int best[3], float *bestMs);
/* the same over seven axes: the three scheduling knobs, where the element
   stage runs, the queue rounds and queue depth, and the workgroup size;
   best[7] */
int exa_batch_autotune_launch(exa_batch *b, const int *ctrlSteps, int nCtrl, const int *elemBatch, int nElem,
                              const int *rareBatch, int nRare, const int *elementStage, int nStage,
                              const int *queueRounds, int nRounds, const int *queueDepth, int nDepth,
                              const int *blockSize, int nBlock, void *stream, int best[7], float *bestMs);

/* ---- neural field sampler (build-side; no reference counterpart) ----
   A multiresolution hash encoding (levels x 2^log2T entries x 2 bf16
   features, resolutions nMin..nMax; Mueller et al. 2022) feeding an MLP
   32-64-64-1 (bf16 weights, float32 master copies, bf16 MFMA with float32
   accumulation), trained with Adam on samples the exact sampler draws.  After
   per-macrocell validation it answers the element stage of point location
   (the stitching elements, ExaStitchSampler.h:128-140) in the macrocells
   where its max error stays within a bound; the majorants of those
   macrocells are widened by the bound so that tracking stays conservative. */
typedef struct exa_neural exa_neural;
int exa_neural_create(exa_renderer *r, int levels, int log2T, float nMin, float nMax, uint32_t seed,
                      exa_neural **out);
void exa_neural_destroy(exa_neural *n);
/* info[4]: levels, log2T, Adam steps taken, validated; res[16]: level resolutions */
int exa_neural_get_config(const exa_neural *n, int32_t info[4], int32_t res[16]);
/* float master parameters: table (levels x 2^log2T x 2), W1 (32x64), W2 (64x64), w3 (64), row-major */
int exa_neural_get_params(const exa_neural *n, float *table, float *W1, float *W2, float *w3);
int exa_neural_set_params(exa_neural *n, const float *table, const float *W1, const float *W2, const float *w3);
/* training pool of count samples drawn with the exact sampler (no culling):
   elemFraction inside random stitching elements, the rest in random gridlets */
int exa_neural_make_pool(exa_neural *n, exa_renderer *r, int64_t count, float elemFraction, uint32_t seed,
                         int64_t *numValid);
/* Adam steps on random batches of the pool, learning rate decaying
   exponentially from lr to lrEnd; lossOut[steps] (may be NULL) */
int exa_neural_train(exa_neural *n, int steps, int batch, float lr, float lrEnd, float *lossOut);
/* loss and gradients of one batch given on the host (x in [0,1]^3 over the
   model's cell bounds, value normalised to the value range; count % 64 == 0) */
int exa_neural_gradients(exa_neural *n, const float *x01, const float *target, int count, float *loss,
                         float *dTable, float *dW1, float *dW2, float *dw3);
int exa_neural_apply_adam(exa_neural *n, float lr);
/* values at count device positions (float3, voxel space) */
int exa_neural_evaluate(const exa_neural *n, const float *d_pos, int64_t count, float *d_out, void *stream);
/* out[3]: macrocells answered by the network, macrocells with element-stage
   points, max error over the accepted macrocells (scalar units) */
int exa_neural_validate(exa_neural *n, exa_renderer *r, int pointsPerMC, float bound, uint32_t seed,
                        float out[3]);
int exa_neural_get_flags(const exa_neural *n, uint8_t *out); /* 1 per macrocell */
/* n validated, or NULL for the exact sampler everywhere (persistent scheme,
   table accelerator) */
int exa_renderer_set_neural(exa_renderer *r, exa_neural *n);

/* ---- ExaBricks comparison sampler (model/ExaBrickModel.h:27-56,
        model/ABRs.h:28-100, sampler/ExaBrickSampler.h:30-241) ----
   The reference's basis-function sampler over overlapping bricks of AMR
   cells, in its default configuration (EXABRICK_ABR_TRAVERSAL with the ABR
   sampling BVH, CMakeLists.txt:189-195): active brick regions (ABRs, boxes
   in which one set of bricks overlaps) carry the majorants; a sample is the
   blend of the tent basis functions of the region's bricks.  The regions are
   found on the kd-tree that their build's recursion describes (in place of
   the OptiX BVH over them). */
/* model/ABRs.h:28-63 ExaBrick (32 bytes): size[0]*size[1]*size[2] cells of
   width 2^level from lower (finest-cell units); cell (x,y,z) has scalar
   scalars[begin + x + size[0]*(y + size[1]*z)] */
typedef struct {
  int32_t lower[3];
  int32_t size[3];
  int32_t level;
  uint32_t begin;
} exa_brick;
/* model/ABRs.h:65-75 ABR (44 bytes) */
typedef struct {
  float domain[6]; /* lower xyz, upper xyz */
  float valueRange[2];
  int32_t leafListBegin;
  int32_t leafListSize;
  float finestLevelCellWidth;
} exa_abr;
typedef struct exa_brick_model exa_brick_model;
/* ExaBrickModel::load(bricks, scalars, numBricks) (ExaBrickModel.cpp:119-145)
   and init (148-213): copies; numCells = the last brick's begin + cells */
int exa_brick_model_create(const exa_brick *bricks, const float *scalars, size_t numBricks, exa_brick_model **out);
/* ExaBrickModel::load(brickFile, scalarFile, "") (ExaBrickModel.cpp:27-117):
   .bricks = per brick size (3 int32), lower (3 int32), level (int32), then
   its cell IDs (int32, x fastest) into the .scalars file */
int exa_brick_model_load(const char *bricksPath, const char *scalarsPath, exa_brick_model **out);
/* Brick maker (no reference counterpart; the reference reads bricks made by
   external tools): each level's leaf cells, in blocks of blockCells^3 cells
   (1..16), covered greedily by dense boxes; cell IDs = input order */
int exa_brick_model_from_amr(const exa_amr *a, int blockCells, exa_brick_model **out);
int exa_amr_write_bricks(const exa_amr *a, int blockCells, const char *bricksPath, const char *scalarsPath);
/* counts[6]: bricks, brick cells, ABRs, leaf-list entries, kd nodes, kd depth */
int exa_brick_model_get_counts(const exa_brick_model *m, int64_t counts[6]);
/* copies, sized by exa_brick_model_get_counts; NULL skips one */
int exa_brick_model_get(const exa_brick_model *m, exa_brick *bricks, float *scalars, exa_abr *abrs,
                        int32_t *leafList);
/* kd nodes, 2 int32 each: split (float bits), info = (lower child << 2) | dim
   for a plane (upper child = next node), ((ABR + 1) << 2) | 3 for a leaf */
int exa_brick_model_get_kd_nodes(const exa_brick_model *m, int32_t *nodes);
/* cellBounds = union of brick bounds, valueRange = union of ABR ranges
   (ExaBrickModel.cpp:177-187) */
int exa_brick_model_get_bounds(const exa_brick_model *m, float cellBounds[6], float valueRange[2]);
void exa_brick_model_destroy(exa_brick_model *m);

typedef struct exa_brick_sampler exa_brick_sampler;
/* ExaBrickSampler::build in ABR mode (ExaBrickSampler.cpp:26-247) */
int exa_brick_sampler_create(const exa_brick_model *m, void *stream, exa_brick_sampler **out);
/* computeMaxOpacitiesForABRs (ExaBrickSampler.cu:65-98, 100-118); a region
   with max opacity 0 is culled (ExaBrickABRGeomBounds,
   sampler/programs/ExaBrickSampler.cu:30-43) */
int exa_brick_sampler_compute_max_opacities(exa_brick_sampler *s, const float *rgba, int numColors, float lo,
                                            float hi, void *stream);
int exa_brick_sampler_get_max_opacities(const exa_brick_sampler *s, float *hostOut); /* 1 per ABR */
/* sample() by point (ExaBrickSampler.h:221-236, ExaBrickABRGeomSamplingIsect
   sampler/programs/ExaBrickSampler.cu:69-94): {0, -1, value} where a region
   that is not culled contains the point, else {0, 0, 0} */
int exa_brick_sampler_locate(const exa_brick_sampler *s, const float *d_pos, int64_t n, exa_sample *d_out,
                             void *stream);
void exa_brick_sampler_destroy(exa_brick_sampler *s);
/* OWLRenderer(const ExaBrick *, const float *, size_t, remap_from, remap_to,
   numMCs) (OWLRenderer.cpp:269-291) from a brick model (copied); the
   exa_renderer_* setters, render, render_counted, get_accum and get_state
   apply.  renderFrame_ExaBrickSampler shades with Default only (shade mode
   ignored); the ray marcher is the ExaBricks integrateDVR (deviceCode.cu:
   401-499).  Render counters: [0] samples blended, [1] brick records read,
   [2] kd leaves visited, [7] regions entered, [8] pixels, [9] TF fetches.
   The scheme selects the launch as for exaStitch: EXA_RENDER_PERSISTENT
   (default) runs the path tracer and direct lighting in persistent lanes
   that take pixels from a work counter, EXA_RENDER_MEGAKERNEL one thread per
   pixel; the ray marcher and heat-map frames always run one thread per
   pixel; the frames are the same.  The persistent tuning is ignored;
   accelerator, batch and neural entry points return EXA_ERR_STATE. */
int exa_renderer_create_bricks(const exa_brick_model *m, const int32_t numMCs[3], const float remapFrom[6],
                               const float remapTo[6], exa_renderer **out);
/* the renderer's brick sampler, or NULL for an exaStitch renderer */
const exa_brick_sampler *exa_renderer_get_brick_sampler(const exa_renderer *r);

/* ---- host helpers ---- */
/* Camera from eye/point-of-interest/up/fovy for a W x H image
   (viewer.cpp:79-107 SimpleCamera, 550-560 cameraChanged). */
int exa_camera_look_at(const float from[3], const float at[3], const float up[3], float fovyDeg, int W, int H,
                       float org[3], float dir00[3], float du[3], float dv[3]);
/* .xf transfer-function file (headless.cpp:258-283).  Pass rgba == NULL to
   query numColors. */
int exa_xf_load(const char *path, float *opacityScale, float absDomain[2], float relDomain[2], int *numColors,
                float *rgba, int maxColors);
/* Piecewise-linear RGBA transfer function: points are (x, r, g, b, a) with x
   ascending in [0,1]; rasterised to N texels at x = i/(N-1). */
int exa_tf_rasterize(const float *points, int numPoints, int N, float *rgba);

#ifdef __cplusplus
}
#endif
#endif
