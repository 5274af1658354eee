/*
 * exa_oracle.h -- CPU oracle for the exaStitch volume path-tracer hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference's algorithms (owl-project/owlExaStitcher), used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 * No product code links, loads or calls it.
 *
 * Every function cites the reference file:line it follows.  Two builds of
 * the numeric core exist:
 *   *_f32  -- single precision, operation order and literals as in the
 *             reference source (compiled with -ffp-contract=off), used for
 *             bit-exact parity with the HIP kernels;
 *   *_f64  -- the same algorithms in double precision (the "FP64 high-spp
 *             oracle" for converged-image RMSE checks).
 *
 * Parity pinning (what checks each function; see tests/test_oracle_*.py):
 *   intersect_gridlet / tet / pyr / wedge / hex : affine-field reproduction,
 *       partition of unity, value bounds, inside/outside on known points,
 *       agreement of the f32 and f64 builds, float-vs-double Newton.
 *   locate_brute / locate_binned                 : binned == brute force,
 *       containing-count census, miss outside the mesh.
 *   grid_build / max_opacity                     : brute-force recomputation
 *       of per-macrocell ranges from the definition, exact min/max.
 *   dda3                                         : t-intervals tile the
 *       entry..exit segment, visited cell contains each segment midpoint.
 *   tf_lookup                                    : texel centres return the
 *       table value, border halving, linearity between texel centres.
 *   lcg                                          : first values of the TEA/LCG
 *       stream against an independent numpy restatement; uniformity.
 *   logf / sincosf                               : error bound vs libm (<= 2 ulp).
 *   render (Woodcock, ratio tracking, HG)        : homogeneous-medium closed
 *       form exp(-sigma*d), HG uniformity, f32 vs f64 image RMSE.
 *   dual_build                                   : cell-count invariants,
 *       perfect cubes == gridlet cells, element volumes tile the
 *       stitching region (sum of volumes), analytic cases (one-level box).
 *   expf / powf                                  : error bound vs libm.
 *   abrs_build (ExaBricks regions)               : the regions tile the union
 *       of brick domains without overlap, each leaf list is exactly the set
 *       of bricks overlapping the region, no brick face cuts a region, value
 *       ranges recomputed from their definition, split-order invariants.
 *   brick blend / locate                         : the reference's 2-brick
 *       test data (tools/makeTestData.cpp) at known points, constant and
 *       affine fields inside single bricks, partition of unity of the tent
 *       weights, region blend == brute-force blend over all bricks.
 *   brick rendering (region walk)                : closest-hit regions by
 *       brute force over all regions; homogeneous closed forms.
 */
#ifndef EXA_ORACLE_H
#define EXA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Gridlet.h:22-37 -- 32-byte gridlet record */
typedef struct {
  int32_t  lower[3];
  int32_t  level;
  int32_t  dims[3];
  uint32_t begin;
} ora_gridlet;

/* sampler/Sampler.h:72-76 */
typedef struct {
  int32_t primID;
  int32_t cellID;
  float   value;
} ora_sample;

/* In-memory ExaStitchModel (model/ExaStitchModel.h:40-50) plus the
   per-primitive max opacities used for culling
   (sampler/programs/ExaStitchSampler.cu:35-37,81-83).  NULL opacity arrays
   mean "nothing culled". */
typedef struct {
  const float       *vertices;      /* float4 xyz+scalar, numVertices */
  int64_t            numVertices;
  const int32_t     *indices;       /* int[8] per element, -1 padded */
  int64_t            numElements;
  const ora_gridlet *gridlets;
  int64_t            numGridlets;
  const float       *gridletScalars;
  int64_t            numGridletScalars;
  const float       *gridletMaxOpacity;
  const float       *elementMaxOpacity;
} ora_model;

/* ---- element math ------------------------------------------------------ */
int ora_intersect_gridlet_f32(float *value, int32_t *cellID, const float pos[3],
                              const ora_gridlet *g, const float *scalars);
int ora_intersect_tet_f32(float *value, const float P[3], const float *v4x4);
int ora_intersect_pyr_f32(float *value, const float P[3], const float *v5x4);
int ora_intersect_wedge_f32(float *value, const float P[3], const float *v6x4);
int ora_intersect_hex_f32(float *value, const float P[3], const float *v8x4);
int ora_intersect_element_f32(float *value, const float P[3], const int32_t *idx8,
                              const float *vertices);
int ora_intersect_gridlet_f64(double *value, int32_t *cellID, const double pos[3],
                              const ora_gridlet *g, const float *scalars);
int ora_intersect_element_f64(double *value, const double P[3], const int32_t *idx8,
                              const float *vertices);

/* ---- point location ----------------------------------------------------- */
/* Deterministic rule (the reference's winner among several accepting
   primitives is OptiX-order dependent): the lowest-ID accepting gridlet, else
   the lowest-ID accepting element, else a miss {-1,-1,0}. */
void ora_locate_brute_f32(const ora_model *m, const float *pos, int64_t n,
                          ora_sample *out, int32_t *numContaining);

typedef struct ora_bins ora_bins;
ora_bins *ora_bins_build(const ora_model *m, int res);
void      ora_bins_free(ora_bins *b);
void      ora_locate_binned_f32(const ora_bins *b, const ora_model *m,
                                const float *pos, int64_t n, ora_sample *out);

/* ---- majorant grid ------------------------------------------------------ */
void ora_project_on_grid(const float V[3], const int32_t dims[3],
                         const float bounds[6], int32_t out[3]);
void ora_grid_build(const ora_model *m, const int32_t dims[3],
                    const float bounds[6], float *ranges /* 2 per MC */);
void ora_gridlet_value_ranges(const ora_model *m, float *ranges /* 2 per gridlet */);
void ora_map_max_opacity(const float *ranges, int64_t n, const float *rgba,
                         int32_t numColors, const float xfRange[2], float *out);
void ora_element_max_opacity(const ora_model *m, const float *rgba,
                             int32_t numColors, const float xfRange[2], float *out);

/* ---- DDA ---------------------------------------------------------------- */
int ora_dda_trace_f32(const float org[3], const float dir[3], float tmin, float tmax,
                      const int32_t dims[3], const float bounds[6],
                      int64_t *cells, float *t0s, float *t1s, int maxCells);

/* ---- transfer function -------------------------------------------------- */
void ora_tf_lookup_f32(const float *rgba, int32_t N, const float domain[2],
                       float v, float out[4]);

/* ---- phase function (deviceCode.cu:144-149, 957-985) -------------------- */
void ora_orthonormal_basis_f32(const float w[3], float u[3], float v[3]);
/* n directions hg_sample(wo) (g = 0) drawn from the LCG stream seeded with
   (v0, v1), 3 floats each */
void ora_hg_sample_f32(const float wo[3], uint32_t v0, uint32_t v1, int64_t n, float *out);

/* ---- RNG and math ------------------------------------------------------- */
uint32_t ora_lcg_init(uint32_t v0, uint32_t v1);
float    ora_lcg_next(uint32_t *state);
float    ora_logf(float x);
void     ora_sincosf(float x, float *s, float *c);
float    ora_expf(float x);
float    ora_powf(float x, float y);

/* ---- renderer ----------------------------------------------------------- */
#define ORA_LIGHTS_MAX 4
typedef struct {
  float   pos[3];      /* point light: position; directional: direction to light */
  float   intensity;
  int32_t on;
  int32_t type;        /* 0 = point (deviceCode.cu:856-926), 1 = directional */
} ora_light;

typedef struct {
  int32_t W, H, spp, accumID;
  int32_t integrator;      /* 0 path tracer, 1 direct lighting, 2 ray marcher */
  int32_t maxBounces;      /* deviceCode.cu:1050 hard-codes 1024 */
  float   camOrg[3], camDir00[3], camDu[3], camDv[3];
  float   worldBounds[6];
  float   voxelXfm[12];    /* affine3f: vx,vy,vz,p (column vectors) */
  float   voxelXfmInv[12];
  float   lightXfm[12];
  int32_t gridDims[3];
  float   gridBounds[6];
  const float *majorants;
  const float *tfRGBA;
  int32_t tfN;
  float   tfDomain[2];
  float   opacityScale;
  int32_t clipEnabled;
  float   clipN[3];
  float   clipD;
  ora_light lights[ORA_LIGHTS_MAX];
  int32_t debugCrosshair;  /* deviceCode.cu:1334-1342 (DEBUGGING == 1) */
  int32_t shadeMode;       /* deviceCode.cu:619-623: 0 default, 1 gridlets, 2 teaser */
  float   dt;              /* ray-marcher step (LaunchParams render.dt) */
  /* ExaBricks scene (renderFrame_ExaBrickSampler, deviceCode.cu:1378-1388);
     NULL: the exaStitch model passed to the render call */
  const struct ora_brick_scene *bricks;
  /* deviceCode.cu:1285-1296 sub-image (lower.xy, upper.xy in [0,1]) and
     1337-1340 selection box */
  int32_t subImageActive;
  float   subImage[4];
  int32_t subImageSelecting;
  float   subImageSel[4];
} ora_render_params;

/* Render rows [y0,y1) of one frame.  accum is float4[W*H] (in/out),
   fb is uint32[W*H] (may be NULL). */
void ora_render_f32(const ora_render_params *p, const ora_model *m, const ora_bins *b,
                    float *accum, uint32_t *fb, int32_t y0, int32_t y1);
/* ora_render_f32 that also records the position of every point-location
   query (tentative collision) into out (3 floats each, at most cap; *count,
   shared by threads, counts all of them) */
void ora_render_queries_f32(const ora_render_params *p, const ora_model *m, const ora_bins *b, float *accum,
                            int32_t y0, int32_t y1, float *out, int64_t cap, int64_t *count);
/* FP64 build: same stream of random numbers, double arithmetic; accum is
   double4[W*H]. */
void ora_render_f64(const ora_render_params *p, const ora_model *m, const ora_bins *b,
                    double *accum, int32_t y0, int32_t y1);

/* ---- dual mesh (tools/amrMakeDualMesh) ---------------------------------- */
typedef struct { int32_t pos[3]; int32_t level; } ora_cell;
typedef struct ora_dual ora_dual;
ora_dual *ora_dual_build(const ora_cell *cells, const float *scalars, int64_t n);
/* what: 0 tets, 1 pyrs, 2 wedges, 3 hexes, 4 elements, 5 vertices,
         6 gridlets, 7 gridlet scalars, 8 perfect cubes, 9 brick write conflicts,
         10 degenerate dual cells dropped */
int64_t ora_dual_count(const ora_dual *d, int what);
void ora_dual_get(const ora_dual *d, int32_t *indices, float *vertices,
                  ora_gridlet *gridlets, float *gridletScalars);
void ora_dual_free(ora_dual *d);

/* model/ExaStitchModel.cpp:78-99,224-245: cellBounds and valueRange */
void ora_model_bounds(const ora_model *m, float cellBounds[6], float valueRange[2]);

/* ---- ExaBricks (model/ABRs.h, model/ABRs.cpp, sampler/ExaBrickSampler.h) -- */
/* model/ABRs.h:28-63 ExaBrick */
typedef struct {
  int32_t  lower[3];
  int32_t  size[3];
  int32_t  level;
  uint32_t begin;
} ora_brick;
/* model/ABRs.h:65-75 ABR */
typedef struct {
  float   domain[6];
  float   valueRange[2];
  int32_t leafListBegin;
  int32_t leafListSize;
  float   finestLevelCellWidth;
} ora_abr;
typedef struct ora_brick_scene {
  const ora_brick *bricks;
  int64_t          numBricks;
  const float     *scalars;
  const ora_abr   *abrs;
  int64_t          numAbrs;
  const int32_t   *leafList;
  const float     *abrMaxOpacity; /* majorants; 0 = culled; NULL = none culled */
} ora_brick_scene;

/* ABRs::buildFrom (ABRs.cpp:272-352): regions in the reference's order */
typedef struct ora_abrs ora_abrs;
ora_abrs *ora_abrs_build(const ora_brick *bricks, int64_t n, const float *scalars);
int64_t   ora_abrs_count(const ora_abrs *a, int what); /* 0 regions, 1 leaf-list entries */
void      ora_abrs_get(const ora_abrs *a, ora_abr *abrs, int32_t *leafList);
void      ora_abrs_free(ora_abrs *a);
/* addBasisFunctions (ExaBrickSampler.h:102-202) summed over every brick
   whose closed domain contains the point, ascending IDs: value and weight */
void ora_brick_blend_brute_f32(const ora_brick *bricks, int64_t n, const float *scalars,
                               const float *pos, int64_t npts, float *value, float *sumWeights);
/* sample() by point (ExaBrickSampler.h:221-236): the lowest-ID region not
   culled that contains the point, blend over its bricks; {0,-1,v}, or
   {0,0,0} without weight */
void ora_brick_locate_f32(const ora_brick_scene *s, const float *pos, int64_t n, ora_sample *out);

#ifdef __cplusplus
}
#endif
#endif
