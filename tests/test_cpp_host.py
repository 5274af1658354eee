"""The C++ host side over the C ABI, on the CPU: include/exa/OWLRenderer.h
compiles on its own and links against every library symbol it names, and the
C++ headless harness rejects bad command lines and unreadable model files
before it touches a GPU (viewer.cpp:755-885 flags, OWLRenderer.cpp:163 "Could
not load module" as an exception)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADLESS = os.path.join(ROOT, "owlexastitcher_amd", "exa_headless")

PROGRAM = r'''
#include "exa/OWLRenderer.h"
#include <cstdio>
int main(int argc, char **argv) {
  using namespace exa;
  if (argc < 2) return 0;  // only the link is under test
  OWLRenderer r(argv[1], "", "", "", "", "", "", argc > 2 ? argv[2] : "");
  r.setColorMap({{0, 0, 0, 0}, {1, 1, 1, 1}});
  r.setRange(r.valueRange);
  r.setRelDomain({0.f, 100.f});
  r.setOpacityScale(100.f);
  r.setType(OWLRenderer::Type::DirectLighting);
  r.setShadeMode(0);
  r.setClipPlane(0, false, {0, 0, 1}, 0.f);
  r.setLightSource(0, r.modelBounds.upper, 60.f, false);
  r.setLightSpaceTransform(r.lightSpaceTransform);
  r.setSubImage({{0, 0}, {1, 1}}, false);
  r.setSubImageSelection({{0, 0}, {1, 1}}, false);
  r.setDirectionalLight(0, {0, 1, 0}, 1.f, false);
  r.resize({4, 4});
  r.setCamera({0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {0, 1, 0});
  r.resetAccum();
  r.render(nullptr);
  std::printf("%d %zu %zu\n", r.accumID, r.accum().size(), r.shadeModes().size());
  ExaStitchSampler s;
  auto m = ExaStitchModel::load(argv[1], "", argc > 2 ? argv[2] : "");
  m->setNumGridCells({4, 4, 4});
  std::printf("%d\n", (int)s.build(m));
  std::vector<ExaBrick> b(1);
  b[0].size = {1, 1, 1};
  float v = 1.f;
  OWLRenderer rb(b.data(), &v, 1);
  ExaBrickSampler bs;
  auto bm = ExaBrickModel::load(b.data(), &v, 1);
  std::printf("%d %d\n", (int)bs.build(bm), (int)bs.build(m));
  bs.computeMaxOpacities({{0, 0, 0, 1}, {1, 1, 1, 1}}, bm->valueRange);
  // a caller written against the plugin base (sampler/Sampler.h:27-69)
  std::vector<Sampler::SP> plugins = {std::make_shared<ExaBrickSampler>(), std::make_shared<ExaStitchSampler>()};
  for (const Sampler::SP &sp : plugins) {
    const bool built = sp->build(m);  // false: not a model of its kind
    std::printf("%s %d %d", sp->className().c_str(), (int)built, (int)(sp->as<ExaStitchSampler>() != nullptr));
    if (built) {
      sp->computeMaxOpacities({{0, 0, 0, 0}, {1, 1, 1, 1}}, m->valueRange);
      std::printf(" %zu %d", sp->maxOpacities().size(), (int)(sp->majorantAccel.grid != nullptr));
    }
    std::printf("\n");
  }
  // the renderers picked their samplers by the model type (OWLRenderer.cpp:315-325);
  // a renderer's sampler maps its majorants from the renderer's colour map
  rb.setColorMap({{0, 0, 0, 1}, {1, 1, 1, 1}});
  rb.setRange(rb.valueRange);
  std::printf("%s %s %zu\n", r.sampler->className().c_str(), rb.sampler->className().c_str(),
              rb.sampler->maxOpacities().size());
  return 0;
}
'''


def test_header_compiles_and_links(tmp_path):
    src = tmp_path / "use_header.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "use_header"
    lib_dir = os.path.join(ROOT, "owlexastitcher_amd")
    cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
           str(src), "-o", str(exe), "-L", lib_dir, "-lexastitch", f"-Wl,-rpath,{lib_dir}"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    # without arguments the program returns before any library call
    assert subprocess.run([str(exe)], timeout=60).returncode == 0


@pytest.mark.parametrize("args,message", [
    ([], "no model"),
    (["m.exum", "--bogus"], "unknown cmdline arg '--bogus'"),
    (["m.exum", "-bm", "x"], "'-bm' selects a sampler or mode that is not built"),
    (["m.exum", "--size", "64", "x"], "not an integer: 'x'"),
    (["m.exum", "--camera", "1", "2"], "missing value(s) after '--camera'"),
    (["-bricks", "m.bricks"], "-bricks needs -scalars"),
    (["-cells", "m.cells"], "-cells needs -scalars"),
    (["m.exum", "--frames", "0"], "--frames must be at least 1"),
])
def test_headless_usage_errors(args, message):
    r = subprocess.run([HEADLESS] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Fatal error: " + message in r.stderr, r.stderr
    assert "Usage: exa_headless" in r.stderr


def test_headless_unreadable_model(tmp_path):
    missing = str(tmp_path / "none.scalars")
    r = subprocess.run([HEADLESS, str(tmp_path / "none.exum"), "-scalars", missing], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "cannot open" in r.stderr and missing in r.stderr
