// exa_headless.cpp -- the reference's headless benchmark viewer in C++, on
// the host classes of include/exa/OWLRenderer.h.
//
//   exa_headless model.exum -grids model.grids -scalars model.scalars
//       -xf cloud.xf --camera px py pz ix iy iz ux uy uz -fovy 70
//       --size 1024 1024 -o out -fps fps.log
//
// Follows viewer.cpp:119-158 (defaults), 755-885 (flags), 919-975 (renderer
// set-up), 1476-1513 (headless set-up) and headless.cpp:97-186 (frame loop,
// log and fps files), 250-298 (.xf file, xfRange).  The element mesh is this
// library's EXAELEM1 file; -cells with -scalars builds the dual mesh in
// memory; -bricks with -scalars renders with the ExaBricks sampler when no
// mesh or grids file is given.  Flags the reference's headless binary parses
// without applying (-spp, --heat-map, -dt, --range) are applied, as its
// interactive viewer does through its widgets; --frames N (default 510) and
// --screenshot-frame N (default 10) are additions.  Flags of samplers or
// modes that are not built (-bm, -qc, -kdtree, -majorants, -mesh) are
// rejected.  The same harness exists in Python (owlexastitcher_amd.headless);
// both give the same images for the same arguments.
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <time.h>
#include <zlib.h>
#include <algorithm>
#include <chrono>
#include <fstream>
#include <string>
#include <vector>
#include "exa/OWLRenderer.h"

using namespace exa;

namespace {

constexpr int kClipPlanesMax = 1;  // common.h:34

struct UsageError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

struct ClipPlane {
  bool enabled = false;
  vec3f N{0.f, 0.f, 1.f};
  float d = 1000.f;
};

// viewer.cpp:119-158 defaults
struct Cmdline {
  std::string inFileName, scalarFileName, gridsFileName, amrCellFileName, exaBrickFileName, xfFileName;
  std::string outFileName = "Witcher3.png";
  std::string fpsFileName = "fps.log";
  range1f valueRange{1e30f, -1e30f};
  box2f subImage{{1e30f, 1e30f}, {-1e30f, -1e30f}};
  vec3f vp, vi, vu;
  float fovy = 70.f;
  ClipPlane clipPlanes[kClipPlanesMax];
  box3f remap_from{{0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}};
  box3f remap_to{{0.f, 0.f, 0.f}, {1.f, 1.f, 1.f}};
  int rendererType = 0;
  int shadeMode = 0;
  vec3f lightPos;
  float lightIntensity = 0.f;
  bool lightOn = false;
  vec3i numMCs{128, 128, 128};
  vec2i windowSize{1024, 1024};
  bool haveDt = false;
  float dt = 1.f;
  int spp = 1;
  bool heatMapEnabled = false;
  float heatMapScale = 1.f;
  int frames = 510;
  int screenshotFrame = 10;
};

float toFloat(const std::string &s) {
  size_t n = 0;
  float v = 0.f;
  try {
    v = std::stof(s, &n);
  } catch (const std::exception &) {
    n = 0;
  }
  if (n == 0 || n != s.size()) throw UsageError("not a number: '" + s + "'");
  return v;
}
int toInt(const std::string &s) {
  size_t n = 0;
  int v = 0;
  try {
    v = std::stoi(s, &n);
  } catch (const std::exception &) {
    n = 0;
  }
  if (n == 0 || n != s.size()) throw UsageError("not an integer: '" + s + "'");
  return v;
}

// viewer.cpp:755-885 flag parsing
Cmdline parse(const std::vector<std::string> &argv) {
  Cmdline c;
  int clip = 0;
  for (size_t i = 0; i < argv.size(); ++i) {
    const std::string &a = argv[i];
    auto take = [&](size_t n) {
      if (i + n >= argv.size()) throw UsageError("missing value(s) after '" + a + "'");
      std::vector<std::string> v(argv.begin() + i + 1, argv.begin() + i + 1 + n);
      i += n;
      return v;
    };
    auto takeF = [&](size_t n) {
      std::vector<float> f;
      for (const std::string &s : take(n)) f.push_back(toFloat(s));
      return f;
    };
    if (a.empty() || a[0] != '-') {
      c.inFileName = a;
    } else if (a == "-scalar" || a == "-scalars") {
      c.scalarFileName = take(1)[0];
    } else if (a == "-grids" || a == "-gridlets") {
      c.gridsFileName = take(1)[0];
    } else if (a == "-cells") {
      c.amrCellFileName = take(1)[0];
    } else if (a == "-bricks") {
      c.exaBrickFileName = take(1)[0];
    } else if (a == "-bm" || a == "-qc" || a == "-kdtree" || a == "-majorants" || a == "-mesh") {
      throw UsageError("'" + a + "' selects a sampler or mode that is not built");
    } else if (a == "-xf") {
      c.xfFileName = take(1)[0];
    } else if (a == "--range") {
      const auto v = takeF(2);
      c.valueRange = {v[0], v[1]};
    } else if (a == "--subimg") {
      const auto v = takeF(4);
      c.subImage = {{v[0], v[1]}, {v[2], v[3]}};
    } else if (a == "-fovy" || a == "--fov") {
      c.fovy = takeF(1)[0];
    } else if (a == "--camera") {
      const auto v = takeF(9);
      c.vp = {v[0], v[1], v[2]};
      c.vi = {v[3], v[4], v[5]};
      c.vu = {v[6], v[7], v[8]};
    } else if (a == "-win" || a == "--win" || a == "--size") {
      const auto v = take(2);
      c.windowSize = {toInt(v[0]), toInt(v[1])};
    } else if (a == "-o") {
      c.outFileName = take(1)[0];
    } else if (a == "-fps") {
      c.fpsFileName = take(1)[0];
    } else if (a == "-spp" || a == "--spp") {
      c.spp = toInt(take(1)[0]);
    } else if (a == "--heat-map") {
      c.heatMapEnabled = true;
      c.heatMapScale = takeF(1)[0];
    } else if (a == "-dt") {
      c.haveDt = true;
      c.dt = takeF(1)[0];
    } else if (a == "--remap-from" || a == "--remap-to") {
      const auto v = takeF(6);
      box3f &b = a == "--remap-from" ? c.remap_from : c.remap_to;
      b = {{v[0], v[1], v[2]}, {v[3], v[4], v[5]}};
    } else if (a == "--clip-plane") {
      if (clip >= kClipPlanesMax) throw UsageError("at most 1 clip plane");
      const auto v = takeF(4);
      c.clipPlanes[clip].enabled = true;
      c.clipPlanes[clip].N = {v[0], v[1], v[2]};
      c.clipPlanes[clip].d = v[3];
      ++clip;
    } else if (a == "-rt") {
      c.rendererType = toInt(take(1)[0]);
    } else if (a == "-sm") {
      c.shadeMode = toInt(take(1)[0]);
    } else if (a == "--num-mcs") {
      const auto v = take(3);
      c.numMCs = {toInt(v[0]), toInt(v[1]), toInt(v[2])};
    } else if (a == "--light") {
      const auto v = takeF(4);
      c.lightPos = {v[0], v[1], v[2]};
      c.lightIntensity = v[3];
      c.lightOn = true;
    } else if (a == "--frames") {
      c.frames = toInt(take(1)[0]);
    } else if (a == "--screenshot-frame") {
      c.screenshotFrame = toInt(take(1)[0]);
    } else {
      throw UsageError("unknown cmdline arg '" + a + "'");
    }
  }
  if (c.amrCellFileName.empty() && c.inFileName.empty() && c.gridsFileName.empty() && c.exaBrickFileName.empty())
    throw UsageError("no model: give an element-mesh file and/or -grids, -bricks, or -cells");
  if (c.frames < 1) throw UsageError("--frames must be at least 1");
  if (c.windowSize.x < 1 || c.windowSize.y < 1) throw UsageError("--size must be positive");
  return c;
}

// The library's default colour map: diverging blue-grey-red with a linear
// alpha ramp over 128 entries (the reference starts from map 0 of its TF
// editor's library, absent from the reference tree).  Piecewise-linear in
// double between the stops at 0, 1/2, 1, rounded to float.
std::vector<vec4f> defaultColorMap(int n = 128) {
  const double stops[3][3] = {{59, 76, 192}, {221, 221, 221}, {180, 4, 38}};
  const double xs[3] = {0.0, 0.5, 1.0};
  std::vector<vec4f> cm(n);
  const double step = 1.0 / (n - 1);
  for (int i = 0; i < n; ++i) {
    const double x = i == n - 1 ? 1.0 : i * step;
    const int j = x < xs[1] ? 0 : 1;
    double rgb[3];
    for (int k = 0; k < 3; ++k) {
      const double f0 = stops[j][k] / 255.0, f1 = stops[j + 1][k] / 255.0;
      rgb[k] = x == xs[2] ? stops[2][k] / 255.0 : (f1 - f0) / (xs[j + 1] - xs[j]) * (x - xs[j]) + f0;
    }
    cm[i] = {(float)rgb[0], (float)rgb[1], (float)rgb[2], (float)x};
  }
  return cm;
}

// headless.cpp:285-298 xfRange: the ordered absolute domain lerped at the
// ordered relative domain / 100, in float
range1f xfRange(range1f absDomain, range1f relDomain) {
  const float a0 = std::min(absDomain.lower, absDomain.upper), a1 = std::max(absDomain.lower, absDomain.upper);
  const float r0 = .01f * relDomain.lower, r1 = .01f * relDomain.upper;
  const float t0 = std::min(r0, r1), t1 = std::max(r0, r1);
  auto lerp = [&](float t) { return (1.f - t) * a0 + t * a1; };
  return {lerp(t0), lerp(t1)};
}

// 8-bit RGBA PNG, no filtering, zlib level 6
void writePNG(const std::string &path, const std::vector<uint32_t> &px, int w, int h) {
  std::vector<unsigned char> raw;
  raw.reserve((size_t)h * (4 * (size_t)w + 1));
  for (int y = 0; y < h; ++y) {
    raw.push_back(0);
    const unsigned char *row = (const unsigned char *)&px[(size_t)y * w];
    raw.insert(raw.end(), row, row + 4 * (size_t)w);
  }
  uLongf zlen = compressBound(raw.size());
  std::vector<unsigned char> z(zlen);
  if (compress2(z.data(), &zlen, raw.data(), raw.size(), 6) != Z_OK) throw std::runtime_error("zlib failed");
  std::ofstream f(path, std::ios::binary);
  if (!f.good()) throw std::runtime_error("cannot write " + path);
  auto be32 = [&](uint32_t v) {
    const unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8),
                                (unsigned char)v};
    f.write((const char *)b, 4);
  };
  auto chunk = [&](const char *tag, const unsigned char *data, size_t n) {
    be32((uint32_t)n);
    uLong crc = crc32(0L, (const Bytef *)tag, 4);
    if (n) crc = crc32(crc, data, (uInt)n);
    f.write(tag, 4);
    f.write((const char *)data, n);
    be32((uint32_t)crc);
  };
  f.write("\x89PNG\r\n\x1a\n", 8);
  unsigned char ihdr[13] = {0};
  for (int k = 0; k < 4; ++k) {
    ihdr[k] = (unsigned char)((uint32_t)w >> (24 - 8 * k));
    ihdr[4 + k] = (unsigned char)((uint32_t)h >> (24 - 8 * k));
  }
  ihdr[8] = 8;  // bit depth
  ihdr[9] = 6;  // RGBA
  chunk("IHDR", ihdr, 13);
  chunk("IDAT", z.data(), zlen);
  chunk("IEND", nullptr, 0);
  if (!f.good()) throw std::runtime_error("write error on " + path);
}

// The reference logs nvidia-smi; the ROCm counterpart is rocm-smi
std::string gpuInfo() {
  std::string out;
  FILE *p = popen("rocm-smi --showproductname 2>/dev/null", "r");
  if (!p) return out;
  char buf[512];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), p)) > 0) out.append(buf, n);
  pclose(p);
  return out;
}

#define HIP_OK(x)                                                                              \
  do {                                                                                         \
    hipError_t e_ = (x);                                                                       \
    if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// headless.cpp: a fixed-size device frame buffer, a camera, the .xf state
// and the benchmark loop
struct Headless {
  OWLRenderer &renderer;
  std::string outFileName, outFileNameFPS;
  float opacityScale = 100.f;
  range1f absDomain{0.f, 1.f}, relDomain{0.f, 99.f};
  std::vector<vec4f> colorMap;
  uint32_t *fb = nullptr;
  vec2i fbSize;
  bool haveCamera = false;
  vec3f from, at, up;
  float fovy = 70.f;

  Headless(OWLRenderer &r, vec2i size) : renderer(r) { resize(size); }
  ~Headless() {
    if (fb) (void)hipFree(fb);
  }
  void resize(vec2i size) {
    if (fb) HIP_OK(hipFree(fb));
    fb = nullptr;
    fbSize = size;
    HIP_OK(hipMalloc((void **)&fb, (size_t)size.x * size.y * sizeof(uint32_t)));
    HIP_OK(hipMemset(fb, 0, (size_t)size.x * size.y * sizeof(uint32_t)));
    renderer.resize(size);
    renderer.resetAccum();
    if (haveCamera) cameraChanged();
  }
  void setCameraOrientation(vec3f o, vec3f i, vec3f u, float fov) {
    from = o, at = i, up = u, fovy = fov;
    haveCamera = true;
    cameraChanged();
  }
  // viewer.cpp:550-560: SimpleCamera -> renderer camera, accumulation reset
  void cameraChanged() {
    vec3f org, d00, du, dv;
    check(exa_camera_look_at(&from.x, &at.x, &up.x, fovy, fbSize.x, fbSize.y, &org.x, &d00.x, &du.x, &dv.x));
    renderer.setCamera(org, d00, du, dv);
    renderer.resetAccum();
  }
  // headless.cpp:258-283
  void loadTransferFunction(const std::string &path) {
    float os, ad[2], rd[2];
    int n = 0;
    check(exa_xf_load(path.c_str(), &os, ad, rd, &n, nullptr, 0));
    std::vector<vec4f> cm(n);
    check(exa_xf_load(path.c_str(), &os, ad, rd, &n, &cm[0].x, n));
    opacityScale = os;
    absDomain = {ad[0], ad[1]};
    relDomain = {rd[0], rd[1]};
    colorMap = cm;
  }
  // rows flipped so that the file starts with the top row, alpha 255
  void screenShot(const std::string &path) {
    std::vector<uint32_t> px((size_t)fbSize.x * fbSize.y), flipped(px.size());
    HIP_OK(hipMemcpy(px.data(), fb, px.size() * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int y = 0; y < fbSize.y; ++y)
      for (int x = 0; x < fbSize.x; ++x)
        flipped[(size_t)y * fbSize.x + x] = px[(size_t)(fbSize.y - 1 - y) * fbSize.x + x] | 0xff000000u;
    writePNG(path, flipped, fbSize.x, fbSize.y);
  }
  // headless.cpp:97-186: render `frames` frames, time each, save a PNG at
  // screenshotFrame, average the frames after it
  void run(int frames, int screenshotFrame) {
    std::string log;
    auto out = [&](const std::string &s) {
      log += s;
      fputs(s.c_str(), stdout);
      fflush(stdout);
    };
    auto fmt = [](const char *f, double v) {
      char b[64];
      snprintf(b, sizeof(b), f, v);
      return std::string(b);
    };
    out("Benchmark summary:\n==================\n\n");
    out("\n");
    out(gpuInfo());
    out("\nBechmark:\n");
    out("FRAME_ID;SEC.\n");
    const std::string shot = outFileName.empty() ? "offline.png" : outFileName + ".png";
    double total = 0.0;
    int count = 0;
    for (int frameID = 1; frameID <= frames; ++frameID) {
      HIP_OK(hipDeviceSynchronize());
      const auto t1 = std::chrono::steady_clock::now();
      renderer.render(fb);
      HIP_OK(hipDeviceSynchronize());
      const auto t2 = std::chrono::steady_clock::now();
      const double sec = std::chrono::duration<double>(t2 - t1).count();
      out(std::to_string(frameID) + ";" + fmt("%g", sec) + "\n");
      if (frameID == screenshotFrame) {
        screenShot(shot);
        printf("#owl.viewer: frame buffer written to %s\n", shot.c_str());
      }
      if (frameID > screenshotFrame) {
        total += sec;
        ++count;
      }
    }
    const double avg = count ? total / count : NAN;
    const double fps = count ? 1.0 / avg : NAN;
    out(fmt("%g", avg) + " avg. secs over " + std::to_string(count) + " time\n");
    out(fmt("%g", fps) + " FPS\n");
    {
      std::ofstream f(outFileName.empty() ? "benchmark.log" : outFileName + ".log");
      f << log;
    }
    std::ofstream f(outFileName.empty() ? "fps.log" : outFileNameFPS, std::ios::app);
    char stamp[64];
    const time_t now = time(nullptr);
    strftime(stamp, sizeof(stamp), "%Y-%m-%d %H:%M:%S", localtime(&now));
    f << stamp << "," << outFileName << "," << fmt("%g", fps) << "\n";
  }
};

int run(const std::vector<std::string> &argv) {
  const Cmdline c = parse(argv);
  // OWLRenderer.cpp:141-148 order: mesh or grids, then bricks, then cells
  const bool stitchFiles = !c.inFileName.empty() || !c.gridsFileName.empty();
  if (!stitchFiles && !c.exaBrickFileName.empty() && c.scalarFileName.empty())
    throw UsageError("-bricks needs -scalars");
  if (!stitchFiles && c.exaBrickFileName.empty() && !c.amrCellFileName.empty() && c.scalarFileName.empty())
    throw UsageError("-cells needs -scalars");
  // viewer.cpp:919-975 without the GUI
  OWLRenderer r(c.inFileName, c.gridsFileName, stitchFiles ? "" : c.amrCellFileName,
                stitchFiles ? "" : c.exaBrickFileName, "", "", "", c.scalarFileName, "", "", c.remap_from, c.remap_to,
                c.numMCs);
  r.setColorMap(defaultColorMap());
  for (int i = 0; i < kClipPlanesMax; ++i)
    r.setClipPlane(i, c.clipPlanes[i].enabled, c.clipPlanes[i].N, c.clipPlanes[i].d);
  r.setType((OWLRenderer::Type)c.rendererType);
  r.setShadeMode(c.shadeMode);
  if (c.lightOn) r.setLightSource(0, c.lightPos, c.lightIntensity, true);
  else r.setLightSource(0, r.modelBounds.upper, 60.f, false);
  if (c.subImage.lower.x <= c.subImage.upper.x && c.subImage.lower.y <= c.subImage.upper.y)
    r.setSubImage(c.subImage, true);
  r.spp = c.spp;
  if (c.heatMapEnabled) {
    r.heatMapEnabled = true;
    r.heatMapScale = c.heatMapScale;
  }
  if (c.haveDt) r.setDt(c.dt);

  Headless viewer(r, c.windowSize);
  viewer.outFileName = c.outFileName;
  viewer.outFileNameFPS = c.fpsFileName;
  if (c.vu.x != 0.f || c.vu.y != 0.f || c.vu.z != 0.f) {
    viewer.setCameraOrientation(c.vp, c.vi, c.vu, c.fovy);
  } else {
    // centre and span in float, the offset in double, rounded once
    const box3f &b = r.modelBounds;
    const vec3f center = b.center(), span = b.size();
    const vec3f eye{(float)((double)center.x + -0.3 * (double)span.x), (float)((double)center.y + 0.7 * (double)span.y),
                    (float)((double)center.z + 1.0 * (double)span.z)};
    viewer.setCameraOrientation(eye, center, {0.f, 1.f, 0.f}, 70.f);
  }
  range1f absDomain = r.valueRange;
  if (c.valueRange.lower <= c.valueRange.upper) absDomain = c.valueRange;
  if (!c.xfFileName.empty()) {
    viewer.loadTransferFunction(c.xfFileName);
    r.setColorMap(viewer.colorMap);
    viewer.absDomain = absDomain;
    const range1f xr = xfRange(viewer.absDomain, viewer.relDomain);
    r.setRange({std::min(xr.lower, xr.upper), std::max(xr.lower, xr.upper)});
    r.setRelDomain(viewer.relDomain);
    r.setOpacityScale(viewer.opacityScale);
    r.resetAccum();
  } else {
    r.setRange(absDomain);
  }
  viewer.run(c.frames, c.screenshotFrame);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::vector<std::string> args(argv + 1, argv + argc);
  try {
    return run(args);
  } catch (const UsageError &e) {
    fprintf(stderr,
            "\nFatal error: %s\n\nUsage: exa_headless <elements.exum> [-grids f.grids] [-scalars f.scalars] "
            "[-cells f.cells] [-bricks f.bricks] [-xf f.xf] [--camera 9 floats] [-fovy deg] [--size W H] "
            "[-o out] [-fps file] [-spp n] [-rt type] [-sm mode] [--num-mcs x y z] [--light x y z I] "
            "[--clip-plane nx ny nz d] [--frames n]\n",
            e.what());
    return 1;
  } catch (const std::exception &e) {
    fprintf(stderr, "exa_headless: %s\n", e.what());
    return 2;
  }
}
