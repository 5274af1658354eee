/* OWLRenderer.h -- C++ host classes over the C ABI of exa.h, with the names,
 * argument meaning and error behaviour of the reference's host interface:
 *
 *   exa::OWLRenderer       OWLRenderer.h:29-175, OWLRenderer.cpp:123-291, 584-808
 *   exa::Model             model/Model.h:29-65
 *   exa::ExaStitchModel    model/ExaStitchModel.h, ExaStitchModel.cpp:24-257
 *   exa::ExaBrickModel     model/ExaBrickModel.h, ExaBrickModel.cpp:27-213
 *   exa::Sampler           sampler/Sampler.h:27-69 (abstract plugin base)
 *   exa::ExaStitchSampler  sampler/Sampler.h:28-94, sampler/ExaStitchSampler.h:108-141
 *   exa::ExaBrickSampler   sampler/ExaBrickSampler.h:98-241, ExaBrickSampler.cpp:26-247
 *   exa::ExaBrick          model/ABRs.h:28-63
 *
 * Header only; link with libexastitch.so.  Failures throw std::runtime_error
 * with the library's message, as the reference's constructors do
 * (OWLRenderer.cpp:163, 306).  The frame buffer given to render() is device
 * memory owned by the caller (headless.cpp:66).  The math types are plain
 * structs with the member names of owl::common's (x/y/z/w, lower/upper,
 * l.vx/l.vy/l.vz/p), so caller code that fills them reads the same.
 */
#pragma once
#include <stdint.h>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include "exa.h"

namespace exa {

struct vec2i { int x = 0, y = 0; };
struct vec3i { int x = 0, y = 0, z = 0; };
struct vec2f { float x = 0.f, y = 0.f; };
struct vec3f { float x = 0.f, y = 0.f, z = 0.f; };
struct vec4f { float x = 0.f, y = 0.f, z = 0.f, w = 0.f; };
template <typename T>
struct interval {
  T lower, upper;
};
using range1f = interval<float>;
struct box2f {
  vec2f lower, upper;
};
struct box3i {
  vec3i lower, upper;
};
struct box3f {
  vec3f lower, upper;
  vec3f size() const { return {upper.x - lower.x, upper.y - lower.y, upper.z - lower.z}; }
  vec3f center() const { return {.5f * (lower.x + upper.x), .5f * (lower.y + upper.y), .5f * (lower.z + upper.z)}; }
};
struct linear3f {
  vec3f vx{1.f, 0.f, 0.f}, vy{0.f, 1.f, 0.f}, vz{0.f, 0.f, 1.f};
};
struct affine3f {
  linear3f l;
  vec3f p;
};

inline void check(int rc) {
  if (rc != EXA_OK) throw std::runtime_error(exa_get_last_error());
}

/* model/ABRs.h:28-63 ExaBrick; the layout of exa_brick */
struct ExaBrick {
  vec3i lower;
  vec3i size;
  int level = 0;
  uint32_t begin = 0;
  size_t numCells() const { return (size_t)size.x * size.y * size.z; }
  box3f getBounds() const {
    const float w = (float)(1 << level);
    return {{(float)lower.x, (float)lower.y, (float)lower.z},
            {(float)lower.x + size.x * w, (float)lower.y + size.y * w, (float)lower.z + size.z * w}};
  }
};
static_assert(sizeof(ExaBrick) == sizeof(exa_brick), "ExaBrick must have the layout of exa_brick");

inline void boxArray(const box3f &b, float o[6]) {
  o[0] = b.lower.x, o[1] = b.lower.y, o[2] = b.lower.z, o[3] = b.upper.x, o[4] = b.upper.y, o[5] = b.upper.z;
}

/* model/Model.h:29-65: the data a sampler is built from.  The macrocell
   grid dims and the voxel-space remap are stored until a renderer or a
   sampler is built from the model; the remapped bounds (Model::getBounds,
   Model.cpp:50-70) are OWLRenderer::modelBounds, computed by the library. */
struct Model : std::enable_shared_from_this<Model> {
  typedef std::shared_ptr<Model> SP;
  virtual ~Model() = default;
  template <typename T>
  std::shared_ptr<T> as() {
    return std::dynamic_pointer_cast<T>(shared_from_this());
  }
  box3f cellBounds;
  range1f valueRange{1e30f, -1e30f};
  vec3i numGridCells{128, 128, 128};
  box3f remapFrom{{0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}};
  box3f remapTo{{0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}};
  void setNumGridCells(const vec3i &n) { numGridCells = n; }
  void setVoxelSpaceTransform(const box3f from, const box3f to) { remapFrom = from, remapTo = to; }
};

/* ExaStitchModel::load(umesh, grids, scalars) (ExaStitchModel.cpp:24-257);
   the element mesh is this library's EXAELEM1 file */
struct ExaStitchModel : Model {
  typedef std::shared_ptr<ExaStitchModel> SP;
  exa_model *handle = nullptr;
  ~ExaStitchModel() override { exa_model_destroy(handle); }
  static SP load(const std::string umeshFileName, const std::string gridsFileName,
                 const std::string scalarFileName) {
    SP m = std::make_shared<ExaStitchModel>();
    check(exa_model_load(umeshFileName.empty() ? nullptr : umeshFileName.c_str(),
                         gridsFileName.empty() ? nullptr : gridsFileName.c_str(),
                         scalarFileName.empty() ? nullptr : scalarFileName.c_str(), &m->handle));
    m->fetchBounds();
    return m;
  }
  /* an AMR .cells/.scalars pair through the dual mesh (amrMakeDualMesh +
     amrMakeGrids in memory) */
  static SP fromCells(const std::string cellFileName, const std::string scalarFileName) {
    exa_amr *a = nullptr;
    check(exa_amr_load(cellFileName.c_str(), scalarFileName.c_str(), &a));
    exa_dual_mesh *d = nullptr;
    int rc = exa_dual_mesh_build(a, &d);
    exa_amr_destroy(a);
    check(rc);
    SP m = std::make_shared<ExaStitchModel>();
    rc = exa_model_create_from_dual_mesh(d, &m->handle);
    exa_dual_mesh_destroy(d);
    check(rc);
    m->fetchBounds();
    return m;
  }
  void fetchBounds() {
    float cb[6], vr[2];
    check(exa_model_get_bounds(handle, cb, vr));
    cellBounds = {{cb[0], cb[1], cb[2]}, {cb[3], cb[4], cb[5]}};
    valueRange = {vr[0], vr[1]};
  }
};

/* ExaBrickModel::load (ExaBrickModel.cpp:27-145) and init (148-213) */
struct ExaBrickModel : Model {
  typedef std::shared_ptr<ExaBrickModel> SP;
  exa_brick_model *handle = nullptr;
  ~ExaBrickModel() override { exa_brick_model_destroy(handle); }
  static SP load(const std::string brickFileName, const std::string scalarFileName,
                 const std::string kdTreeFileName = "") {
    if (!kdTreeFileName.empty()) throw std::runtime_error("ExaBricks kd-tree majorant files are not supported");
    SP m = std::make_shared<ExaBrickModel>();
    check(exa_brick_model_load(brickFileName.c_str(), scalarFileName.c_str(), &m->handle));
    m->fetchBounds();
    return m;
  }
  static SP load(const ExaBrick *bricks, const float *scalars, size_t numBricks) {
    SP m = std::make_shared<ExaBrickModel>();
    check(exa_brick_model_create((const exa_brick *)bricks, scalars, numBricks, &m->handle));
    m->fetchBounds();
    return m;
  }
  void fetchBounds() {
    float cb[6], vr[2];
    check(exa_brick_model_get_bounds(handle, cb, vr));
    cellBounds = {{cb[0], cb[1], cb[2]}, {cb[3], cb[4], cb[5]}};
    valueRange = {vr[0], vr[1]};
  }
};

struct OWLRenderer;

/* sampler/Sampler.h:27-69: the plugin base a renderer dispatches on.  A
   sampler is built from a model (false when the model is not of its kind),
   maps its value ranges through a colour map into majorants, and answers
   point queries (the device contract sample(), here over device-resident
   points in bulk).  OWLRenderer picks the subclass by the model's type
   (OWLRenderer.cpp:315-325) and lets it make the renderer whose launch reads
   its device data, as "renderFrame_" + className() selects the raygen program
   (OWLRenderer.cpp:334-337) and getLPVariables/setLPs pack the sampler's
   launch parameters; a sampler made by a renderer answers for that
   renderer's data until the renderer is destroyed. */
struct Sampler : std::enable_shared_from_this<Sampler> {
  typedef std::shared_ptr<Sampler> SP;
  virtual ~Sampler() = default;
  template <typename T>
  std::shared_ptr<T> as() {
    return std::dynamic_pointer_cast<T>(shared_from_this());
  }
  virtual bool build(Model::SP model) = 0;
  virtual void computeMaxOpacities(const std::vector<vec4f> &colorMap, range1f xfRange) = 0;
  //! simple reflection: the exact name of the subclass
  virtual std::string className() = 0;
  //! sample() over n device-resident points; exa_sample = Sample (Sampler.h:72-76)
  virtual void sample(const float *d_pos, int64_t n, exa_sample *d_out, void *stream = nullptr) const = 0;
  //! the majorants (max opacities), one per macrocell or brick region, on the host
  virtual std::vector<float> maxOpacities() const = 0;

  /*! the accel used for space skipping; after a successful build exactly
      one of these is set (Sampler.h:55-62) */
  struct Grid {
    vec3i dims;
    box3f bounds;
  };
  struct KDTree {
    int64_t numNodes = 0, depth = 0;
  };
  struct {
    std::shared_ptr<Grid> grid;
    std::shared_ptr<KDTree> kdtree;
  } majorantAccel;

 protected:
  friend struct OWLRenderer;
  virtual exa_renderer *createRenderer(Model::SP model, const int32_t numMCs[3], const float from[6],
                                       const float to[6]) = 0;
  virtual void releaseRenderer() = 0;
};

/* sampler/ExaStitchSampler: build() makes the majorant grid, the
   primitive value ranges and the point-location structures;
   computeMaxOpacities maps them through a colour map; sample() answers
   point queries in bulk (device pointers, exa_sample records). */
struct ExaStitchSampler : Sampler {
  typedef std::shared_ptr<ExaStitchSampler> SP;
  exa_sampler *handle = nullptr;  // built by build(); owned
  ExaStitchSampler() = default;
  ~ExaStitchSampler() override { exa_sampler_destroy(handle); }
  ExaStitchSampler(const ExaStitchSampler &) = delete;
  ExaStitchSampler &operator=(const ExaStitchSampler &) = delete;
  // Sampler::build (Sampler.h:44): false when the model is not an exaStitch model
  bool build(Model::SP model) override {
    auto m = model ? model->as<ExaStitchModel>() : nullptr;
    if (!m) return false;
    if (rendererOwned) throw std::runtime_error("this sampler belongs to a renderer");
    if (!handle) check(exa_sampler_create(&handle));
    const int32_t n[3] = {m->numGridCells.x, m->numGridCells.y, m->numGridCells.z};
    check(exa_model_set_num_grid_cells(m->handle, n));
    float from[6], to[6];
    boxArray(m->remapFrom, from);
    boxArray(m->remapTo, to);
    check(exa_model_set_voxel_space_transform(m->handle, from, to));
    check(exa_sampler_build(handle, m->handle, nullptr));
    view = handle;
    fetchAccel();
    return true;
  }
  /* the majorants of a renderer's sampler follow the renderer's colour map
     and range (OWLRenderer::setColorMap/setRange call this themselves) */
  void computeMaxOpacities(const std::vector<vec4f> &colorMap, range1f xfRange) override {
    if (colorMap.empty()) throw std::runtime_error("empty colour map");
    if (rendererOwned) throw std::runtime_error("set the colour map and range on the renderer that owns this sampler");
    if (!handle) throw std::runtime_error("sampler not built");
    check(exa_sampler_compute_max_opacities(handle, &colorMap[0].x, (int)colorMap.size(), xfRange.lower,
                                            xfRange.upper, nullptr));
  }
  // sample() (ExaStitchSampler.h:128-140) over n device-resident points
  void sample(const float *d_pos, int64_t n, exa_sample *d_out, void *stream = nullptr) const override {
    check(exa_sampler_locate(live(), d_pos, n, d_out, stream));
  }
  std::vector<float> maxOpacities() const override {
    const exa_sampler *s = live();
    int32_t d[3];
    check(exa_sampler_get_grid_info(s, d, nullptr));
    std::vector<float> out((size_t)d[0] * d[1] * d[2]);
    check(exa_sampler_get_max_opacities(s, out.data()));
    return out;
  }
  std::string className() override { return "ExaStitchSampler"; }

 protected:
  exa_renderer *createRenderer(Model::SP model, const int32_t numMCs[3], const float from[6],
                               const float to[6]) override {
    auto m = model->as<ExaStitchModel>();
    if (!m) throw std::runtime_error("not an exaStitch model");
    exa_renderer *r = nullptr;
    check(exa_renderer_create(m->handle, numMCs, from, to, &r));
    view = exa_renderer_get_sampler(r);
    rendererOwned = true;
    fetchAccel();
    return r;
  }
  void releaseRenderer() override {
    view = nullptr;
    majorantAccel.grid.reset();
  }

 private:
  const exa_sampler *view = nullptr;  // the built sampler: own or a renderer's
  bool rendererOwned = false;
  const exa_sampler *live() const {
    if (!view) throw std::runtime_error("sampler not built (or its renderer was destroyed)");
    return view;
  }
  void fetchAccel() {
    auto g = std::make_shared<Grid>();
    float b[6];
    check(exa_sampler_get_grid_info(view, &g->dims.x, b));
    g->bounds = {{b[0], b[1], b[2]}, {b[3], b[4], b[5]}};
    majorantAccel.grid = g;
    majorantAccel.kdtree.reset();
  }
};

/* sampler/ExaBrickSampler in its default configuration (ABR traversal, ABR
   sampling; ExaBrickSampler.cpp:26-247): build() makes the active brick
   regions' device tables, computeMaxOpacities the per-region majorants
   (regions at 0 are culled), sample() the region blend by point. */
struct ExaBrickSampler : Sampler {
  typedef std::shared_ptr<ExaBrickSampler> SP;
  exa_brick_sampler *handle = nullptr;  // built by build(); owned
  ExaBrickSampler() = default;
  ~ExaBrickSampler() override { exa_brick_sampler_destroy(handle); }
  ExaBrickSampler(const ExaBrickSampler &) = delete;
  ExaBrickSampler &operator=(const ExaBrickSampler &) = delete;
  bool build(Model::SP model) override {
    auto m = model ? model->as<ExaBrickModel>() : nullptr;
    if (!m) return false;
    if (rendererOwned) throw std::runtime_error("this sampler belongs to a renderer");
    exa_brick_sampler_destroy(handle);
    handle = nullptr;
    view = nullptr;
    check(exa_brick_sampler_create(m->handle, nullptr, &handle));
    view = handle;
    fetchAccel(*m);
    return true;
  }
  void computeMaxOpacities(const std::vector<vec4f> &colorMap, range1f xfRange) override {
    if (colorMap.empty()) throw std::runtime_error("empty colour map");
    if (rendererOwned) throw std::runtime_error("set the colour map and range on the renderer that owns this sampler");
    if (!handle) throw std::runtime_error("sampler not built");
    check(exa_brick_sampler_compute_max_opacities(handle, &colorMap[0].x, (int)colorMap.size(), xfRange.lower,
                                                  xfRange.upper, nullptr));
  }
  // sample() by point (ExaBrickSampler.h:221-236) over n device-resident points
  void sample(const float *d_pos, int64_t n, exa_sample *d_out, void *stream = nullptr) const override {
    check(exa_brick_sampler_locate(live(), d_pos, n, d_out, stream));
  }
  std::vector<float> maxOpacities() const override {
    const exa_brick_sampler *s = live();
    std::vector<float> out((size_t)numRegions);
    check(exa_brick_sampler_get_max_opacities(s, out.data()));
    return out;
  }
  std::string className() override { return "ExaBrickSampler"; }

 protected:
  exa_renderer *createRenderer(Model::SP model, const int32_t numMCs[3], const float from[6],
                               const float to[6]) override {
    auto m = model->as<ExaBrickModel>();
    if (!m) throw std::runtime_error("not an ExaBricks model");
    exa_renderer *r = nullptr;
    check(exa_renderer_create_bricks(m->handle, numMCs, from, to, &r));
    view = exa_renderer_get_brick_sampler(r);
    rendererOwned = true;
    fetchAccel(*m);
    return r;
  }
  void releaseRenderer() override {
    view = nullptr;
    majorantAccel.kdtree.reset();
  }

 private:
  const exa_brick_sampler *view = nullptr;  // the built sampler: own or a renderer's
  bool rendererOwned = false;
  int64_t numRegions = 0;
  const exa_brick_sampler *live() const {
    if (!view) throw std::runtime_error("sampler not built (or its renderer was destroyed)");
    return view;
  }
  void fetchAccel(const ExaBrickModel &m) {
    int64_t c[6];  // bricks, cells, regions, leaf list, kd nodes, kd depth
    check(exa_brick_model_get_counts(m.handle, c));
    numRegions = c[2];
    auto k = std::make_shared<KDTree>();
    k->numNodes = c[4];
    k->depth = c[5];
    majorantAccel.kdtree = k;
    majorantAccel.grid.reset();
  }
};

struct OWLRenderer {
  //! Different modes the renderer will run in (OWLRenderer.h:32)
  enum class Type { PathTracer, DirectLighting, RayMarcher };

  /* Construct from files (OWLRenderer.cpp:123-181): an element mesh and/or
     grids file makes an exaStitch model; else bricks + scalars an ExaBricks
     model; else cells + scalars an exaStitch model over their dual mesh (the
     reference builds its AMR-cell sampler there).  The BigMesh,
     QuickClusters, triangle-mesh, kd-tree and own-majorant inputs select
     modes that are not built and throw. */
  OWLRenderer(const std::string umeshFileName, const std::string gridsFileName = "",
              const std::string amrCellFileName = "", const std::string exaBrickFileName = "",
              const std::string bigMeshFileName = "", const std::string quickClustersFileName = "",
              const std::string meshFileName = "", const std::string scalarFileName = "",
              const std::string kdtreeFileName = "", const std::string majorantsFileName = "",
              const box3f remap_from = {{0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}},
              const box3f remap_to = {{0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}}, const vec3i numMCs = {128, 128, 128}) {
    if (!bigMeshFileName.empty() || !quickClustersFileName.empty() || !meshFileName.empty() ||
        !majorantsFileName.empty())
      throw std::runtime_error("BigMesh, QuickClusters, triangle meshes and own majorants are not built");
    if (!umeshFileName.empty() || !gridsFileName.empty())
      model = ExaStitchModel::load(umeshFileName, gridsFileName, scalarFileName);
    else if (!exaBrickFileName.empty() && !scalarFileName.empty())
      model = ExaBrickModel::load(exaBrickFileName, scalarFileName, kdtreeFileName);
    else if (!amrCellFileName.empty() && !scalarFileName.empty())
      model = ExaStitchModel::fromCells(amrCellFileName, scalarFileName);
    if (!model) throw std::runtime_error("Could not load module");
    create(remap_from, remap_to, numMCs);
  }

  //! Construct as ExaBricks (OWLRenderer.cpp:269-291)
  OWLRenderer(const ExaBrick *bricks, const float *scalars, size_t numBricks,
              const box3f remap_from = {{0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}},
              const box3f remap_to = {{0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}}, const vec3i numMCs = {128, 128, 128}) {
    model = ExaBrickModel::load(bricks, scalars, numBricks);
    create(remap_from, remap_to, numMCs);
  }

  /*! Block-structured input (OWLRenderer.h:61-67): not implemented in the
      reference either (OWLRenderer.cpp:300-307) */
  struct BlockData {
    vec3i size;
    float *data;
  };
  OWLRenderer(const box3i *, const int *, const BlockData *, size_t, const box3f = {}, const box3f = {},
              const vec3i = {128, 128, 128}) {
    throw std::runtime_error("block array constructor not implemented yet");
  }

  ~OWLRenderer() {
    if (sampler) sampler->releaseRenderer();
    exa_renderer_destroy(r);
  }
  OWLRenderer(const OWLRenderer &) = delete;
  OWLRenderer &operator=(const OWLRenderer &) = delete;

  //! The device state is built by the constructor; kept for callers that call it
  void initGPU() {}

  void setCamera(const vec3f &org, const vec3f &dir_00, const vec3f &dir_du, const vec3f &dir_dv) {
    push();
    call(exa_renderer_set_camera(r, &org.x, &dir_00.x, &dir_du.x, &dir_dv.x));
  }
  void resize(const vec2i &newSize) {
    push();
    call(exa_renderer_resize(r, newSize.x, newSize.y));
    fbSize = newSize;
  }
  /* render(fbPointer) (OWLRenderer.cpp:610-622): the public fields accumID,
     spp and the heat map go into this launch; accumID advances by one.
     Asynchronous on the default stream. */
  void render(uint32_t *fbPointer) {
    check(exa_renderer_set_spp(r, spp > 1 ? spp : 1));
    check(exa_renderer_set_heat_map(r, heatMapEnabled ? 1 : 0, heatMapScale));
    check(exa_renderer_set_accum_id(r, accumID));  // last: the setters above restart it
    check(exa_renderer_render(r, fbPointer, nullptr));
    accumID++;
  }
  void setType(const Type t) {
    type = t;
    push();
    call(exa_renderer_set_type(r, (int)t));
  }
  void setColorMap(const std::vector<vec4f> &newCM) {
    if (newCM.empty()) throw std::runtime_error("empty colour map");
    push();
    call(exa_renderer_set_color_map(r, &newCM[0].x, (int)newCM.size()));
  }
  void setRange(interval<float> xfDomain) {
    push();
    call(exa_renderer_set_range(r, xfDomain.lower, xfDomain.upper));
  }
  void setRelDomain(interval<float> relDomain) {
    push();
    call(exa_renderer_set_rel_domain(r, relDomain.lower, relDomain.upper));
  }
  void setOpacityScale(float scale) {
    push();
    call(exa_renderer_set_opacity_scale(r, scale));
  }
  void setClipPlane(int id, bool enabled, vec3f N, float d) {
    push();
    call(exa_renderer_set_clip_plane(r, id, enabled ? 1 : 0, &N.x, d));
  }
  void setShadeMode(int sm) {
    push();
    call(exa_renderer_set_shade_mode(r, sm));
  }
  // OWLRenderer.cpp:769-774
  std::map<int, std::string> shadeModes() { return {{0, "default"}, {1, "gridlets"}, {2, "teaser"}}; }
  void setSubImage(const box2f si, bool active) {
    push();
    call(exa_renderer_set_sub_image(r, &si.lower.x, active ? 1 : 0));
  }
  void setSubImageSelection(const box2f si, bool active) {
    push();
    call(exa_renderer_set_sub_image_selection(r, &si.lower.x, active ? 1 : 0));
  }
  void setLightSource(int lightID, const vec3f &pos, float intensity, bool on) {
    push();
    call(exa_renderer_set_light_source(r, lightID, &pos.x, intensity, on ? 1 : 0));
  }
  void setLightSpaceTransform(const affine3f xform) {
    push();
    call(exa_renderer_set_light_space_transform(r, &xform.l.vx.x));
    lightSpaceTransform = xform;
  }
  //! Call resetAccum() to reset Monte Carlo convergence frame rendering
  void resetAccum() { accumID = 0; }

  // build-side additions: a light at infinity, the ray-marcher step, the
  // path-length cap
  void setDirectionalLight(int lightID, const vec3f &dir, float intensity, bool on) {
    push();
    call(exa_renderer_set_directional_light(r, lightID, &dir.x, intensity, on ? 1 : 0));
  }
  void setDt(float dt) {
    push();
    call(exa_renderer_set_dt(r, dt));
  }
  void setMaxBounces(int n) {
    push();
    call(exa_renderer_set_max_bounces(r, n));
  }
  //! the float4 accumulation buffer, copied to host memory
  std::vector<vec4f> accum() const {
    std::vector<vec4f> out((size_t)fbSize.x * fbSize.y);
    check(exa_renderer_get_accum(r, &out[0].x));
    return out;
  }

  Type type = Type::PathTracer;
  Model::SP model;
  Sampler::SP sampler;
  exa_renderer *r = nullptr;
  int accumID = 0;
  vec2i fbSize{1, 1};
  int spp = 1;
  bool heatMapEnabled = false;
  float heatMapScale = 1.f;
  box3f modelBounds;
  range1f valueRange;
  affine3f lightSpaceTransform;

 private:
  void create(const box3f &remap_from, const box3f &remap_to, const vec3i &numMCs) {
    float from[6], to[6];
    boxArray(remap_from, from);
    boxArray(remap_to, to);
    model->setVoxelSpaceTransform(remap_from, remap_to);
    model->setNumGridCells(numMCs);
    const int32_t n[3] = {numMCs.x, numMCs.y, numMCs.z};
    // initGPU (OWLRenderer.cpp:315-337): the sampler by the model's type,
    // which then makes the renderer whose launch reads its device data
    if (model->as<ExaBrickModel>())
      sampler = std::make_shared<ExaBrickSampler>();
    else if (model->as<ExaStitchModel>())
      sampler = std::make_shared<ExaStitchSampler>();
    if (!sampler) throw std::runtime_error("Could not load module");
    r = sampler->createRenderer(model, n, from, to);
    float mb[6], vr[2];
    check(exa_renderer_get_bounds(r, mb));
    check(exa_renderer_get_value_range(r, vr));
    modelBounds = {{mb[0], mb[1], mb[2]}, {mb[3], mb[4], mb[5]}};
    valueRange = {vr[0], vr[1]};
  }
  // The public accumID is the frame index of the next render.  Setters that
  // restart convergence in the reference (setType, setColorMap, changed clip
  // planes, setShadeMode, sub-image, lights) reset it in the library, so
  // each call hands the field to the library and reads it back.
  void call(int rc_of_setter) {
    check(rc_of_setter);
    accumID = exa_renderer_get_accum_id(r);
  }
  void push() { check(exa_renderer_set_accum_id(r, accumID)); }
};

}  // namespace exa
