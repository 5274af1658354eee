// amr.cpp -- seeded synthetic AMR scenes for the exaStitch benchmarks.
//
// Cells are emitted in octree depth-first order over the base grid (base
// cells z-major, children in z,y,x order); the emission index is the
// cell's scalar ID, as in the reference's .cells/.scalars pair
// (tools/amrMakeDualMesh/makeDual.cpp:90-96,932-940).
#include <math.h>
#include <stdlib.h>
#include <algorithm>
#include <atomic>
#include <stdexcept>
#include <thread>
#include "exa_host.h"

namespace exa {

int hostThreads() {
  static int n = [] {
    const char *e = getenv("EXA_NUM_THREADS");
    int t = e ? atoi(e) : (int)std::thread::hardware_concurrency();
    if (t < 1) t = 1;
    if (t > 64) t = 64;
    return t;
  }();
  return n;
}

template <typename F>
static void parallelFor(int64_t n, int64_t grain, F &&f) {
  const int T = hostThreads();
  if (T <= 1 || n <= grain) {
    f((int64_t)0, n);
    return;
  }
  std::atomic<int64_t> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < T; ++t)
    pool.emplace_back([&] {
      for (;;) {
        const int64_t b = next.fetch_add(grain);
        if (b >= n) break;
        f(b, std::min(n, b + grain));
      }
    });
  for (auto &th : pool) th.join();
}

static inline uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static inline float u01(uint64_t &s) { return (float)(splitmix64(s) >> 40) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------------------
void amrGaussians(AmrScene &out, int baseRes, int coarseLevel, int fineLo, int fineHi,
                  int numGaussians, uint64_t seed) {
  if (baseRes <= 0 || coarseLevel < 1 || coarseLevel > 10 || numGaussians < 0)
    throw std::runtime_error("amrGaussians: invalid parameters");
  const int w = 1 << coarseLevel, fw = 1 << (coarseLevel - 1);
  const double domain = (double)baseRes * w;
  uint64_t s = seed;
  out.params.clear();
  for (int k = 0; k < numGaussians; ++k) {
    const float mx = u01(s) * (float)domain, my = u01(s) * (float)domain, mz = u01(s) * (float)domain;
    const float sigma = 6.f + 14.f * u01(s);
    const float amp = 0.2f + 0.8f * u01(s);
    out.params.insert(out.params.end(), {mx, my, mz, sigma, amp});
  }
  out.cells.clear();
  out.scalars.clear();
  for (int z = 0; z < baseRes; ++z)
    for (int y = 0; y < baseRes; ++y)
      for (int x = 0; x < baseRes; ++x) {
        const int p[3] = {x * w, y * w, z * w};
        const bool fine = p[0] >= fineLo && p[1] >= fineLo && p[2] >= fineLo && p[0] + w <= fineHi &&
                          p[1] + w <= fineHi && p[2] + w <= fineHi;
        if (!fine) {
          out.cells.push_back(AmrCell{{p[0], p[1], p[2]}, coarseLevel});
        } else {
          for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
              for (int dx = 0; dx < 2; ++dx)
                out.cells.push_back(
                    AmrCell{{p[0] + dx * fw, p[1] + dy * fw, p[2] + dz * fw}, coarseLevel - 1});
        }
      }
  out.scalars.resize(out.cells.size());
  const std::vector<float> &P = out.params;
  parallelFor((int64_t)out.cells.size(), 1 << 14, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const AmrCell &c = out.cells[i];
      const double hw = 0.5 * (double)(1 << c.level);
      const double cx = c.pos[0] + hw, cy = c.pos[1] + hw, cz = c.pos[2] + hw;
      double v = 0.0;
      for (int k = 0; k < numGaussians; ++k) {
        const double dx = cx - P[5 * k], dy = cy - P[5 * k + 1], dz = cz - P[5 * k + 2];
        const double sg = P[5 * k + 3];
        v += P[5 * k + 4] * exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * sg * sg));
      }
      out.scalars[i] = (float)v;
    }
  });
}

// ---------------------------------------------------------------------------
// seeded value noise: lattice values U[0,1) from a hash of (i,j,k,octave),
// smoothstep-weighted trilinear interpolation, fractal sum with lacunarity 2
// and gain 0.5 normalised by the amplitude sum, then a contrast stretch about
// 0.5 clamped to [0,1].
// ---------------------------------------------------------------------------
static inline double lattice(int64_t i, int64_t j, int64_t k, uint64_t seed) {
  uint64_t key = ((uint64_t)(i & 0x1fffff)) | ((uint64_t)(j & 0x1fffff) << 21) |
                 ((uint64_t)(k & 0x1fffff) << 42);
  uint64_t s = key ^ (seed * 0x632be59bd9b4e019ull);
  return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}
static inline double smooth(double t) { return t * t * (3.0 - 2.0 * t); }
static double valueNoise(double x, double y, double z, uint64_t seed) {
  const double fx = floor(x), fy = floor(y), fz = floor(z);
  const int64_t i = (int64_t)fx, j = (int64_t)fy, k = (int64_t)fz;
  const double tx = smooth(x - fx), ty = smooth(y - fy), tz = smooth(z - fz);
  double c[2][2][2];
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) c[dz][dy][dx] = lattice(i + dx, j + dy, k + dz, seed);
  const double x00 = c[0][0][0] + tx * (c[0][0][1] - c[0][0][0]);
  const double x01 = c[0][1][0] + tx * (c[0][1][1] - c[0][1][0]);
  const double x10 = c[1][0][0] + tx * (c[1][0][1] - c[1][0][0]);
  const double x11 = c[1][1][0] + tx * (c[1][1][1] - c[1][1][0]);
  const double y0 = x00 + ty * (x01 - x00);
  const double y1 = x10 + ty * (x11 - x10);
  return y0 + tz * (y1 - y0);
}
struct Fbm {
  int octaves;
  double freq0, contrast;
  uint64_t seed;
  double operator()(double x, double y, double z) const {
    double sum = 0.0, amp = 1.0, norm = 0.0, f = freq0;
    for (int o = 0; o < octaves; ++o) {
      sum += amp * valueNoise(x * f, y * f, z * f, seed + 1000003ull * (uint64_t)o);
      norm += amp;
      amp *= 0.5;
      f *= 2.0;
    }
    const double n = 0.5 + (sum / norm - 0.5) * contrast;
    return n < 0.0 ? 0.0 : (n > 1.0 ? 1.0 : n);
  }
};

void amrNoise(AmrScene &out, int baseRes, int baseLevel, float thresh1, float thresh2,
              int octaves, float contrast, uint64_t seed) {
  if (baseRes <= 0 || baseLevel < 2 || baseLevel > 12 || octaves < 1)
    throw std::runtime_error("amrNoise: invalid parameters");
  const int64_t B = baseRes;
  const int64_t nb = B * B * B;
  const int W = 1 << baseLevel, W1 = W / 2, W0 = W / 4;
  const double domain = (double)B * W;
  Fbm fbm{octaves, 8.0 / domain, (double)contrast, seed};
  // refinement flags: bit 8 = base cell split; bits 0..7 = child split
  std::vector<uint16_t> flag(nb, 0);
  parallelFor(nb, 4096, [&](int64_t b, int64_t e) {
    for (int64_t id = b; id < e; ++id) {
      const int64_t x = id % B, y = (id / B) % B, z = id / (B * B);
      const double cx = x * W + 0.5 * W, cy = y * W + 0.5 * W, cz = z * W + 0.5 * W;
      if (fbm(cx, cy, cz) <= thresh1) continue;
      uint16_t f = 0x100;
      for (int c = 0; c < 8; ++c) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        const double ccx = x * W + dx * W1 + 0.5 * W1, ccy = y * W + dy * W1 + 0.5 * W1,
                     ccz = z * W + dz * W1 + 0.5 * W1;
        if (fbm(ccx, ccy, ccz) > thresh2) f |= (uint16_t)(1u << c);
      }
      flag[id] = f;
    }
  });
  // vertex 2:1 balance: an unsplit base cell touching (face, edge or corner)
  // a base cell that holds finest cells is split once.
  std::vector<uint8_t> force(nb, 0);
  parallelFor(nb, 4096, [&](int64_t b, int64_t e) {
    for (int64_t id = b; id < e; ++id) {
      if (flag[id] & 0x100) continue;
      const int64_t x = id % B, y = (id / B) % B, z = id / (B * B);
      bool need = false;
      for (int dz = -1; dz <= 1 && !need; ++dz)
        for (int dy = -1; dy <= 1 && !need; ++dy)
          for (int dx = -1; dx <= 1 && !need; ++dx) {
            const int64_t nx = x + dx, ny = y + dy, nz = z + dz;
            if (nx < 0 || ny < 0 || nz < 0 || nx >= B || ny >= B || nz >= B) continue;
            if (flag[(nz * B + ny) * B + nx] & 0xff) need = true;
          }
      force[id] = need;
    }
  });
  for (int64_t id = 0; id < nb; ++id)
    if (force[id]) flag[id] = 0x100;
  // emit in depth-first order
  out.cells.clear();
  for (int64_t id = 0; id < nb; ++id) {
    const int x = (int)(id % B), y = (int)((id / B) % B), z = (int)(id / (B * B));
    const int p[3] = {x * W, y * W, z * W};
    if (!(flag[id] & 0x100)) {
      out.cells.push_back(AmrCell{{p[0], p[1], p[2]}, baseLevel});
      continue;
    }
    for (int c = 0; c < 8; ++c) {
      const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
      const int q[3] = {p[0] + dx * W1, p[1] + dy * W1, p[2] + dz * W1};
      if (!(flag[id] & (1u << c))) {
        out.cells.push_back(AmrCell{{q[0], q[1], q[2]}, baseLevel - 1});
        continue;
      }
      for (int g = 0; g < 8; ++g) {
        const int gx = g & 1, gy = (g >> 1) & 1, gz = g >> 2;
        out.cells.push_back(
            AmrCell{{q[0] + gx * W0, q[1] + gy * W0, q[2] + gz * W0}, baseLevel - 2});
      }
    }
  }
  out.scalars.resize(out.cells.size());
  const double cc = 0.5 * domain, R = 0.5 * domain;
  parallelFor((int64_t)out.cells.size(), 1 << 14, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const AmrCell &c = out.cells[i];
      const double hw = 0.5 * (double)(1 << c.level);
      const double x = c.pos[0] + hw, y = c.pos[1] + hw, z = c.pos[2] + hw;
      const double r2 = ((x - cc) * (x - cc) + (y - cc) * (y - cc) + (z - cc) * (z - cc)) / (R * R);
      out.scalars[i] = (float)(fbm(x, y, z) * exp(-r2));
    }
  });
  out.params = {(float)baseRes, (float)baseLevel, thresh1, thresh2, (float)octaves, contrast};
}

// ---------------------------------------------------------------------------
// Expanding spherical shock shell with turbulence in its wake.
//
// The leaf level of a point follows a graded distance field
//   T(x) = floor(log2(max(1, D(x) / G))),  clamped to [0, baseLevel],
// D = min(D_shell, D_core, D_wake), with ds the distance to a noise-perturbed
// sphere minus the shell half thickness, n a low-frequency noise and w a
// cell-scale turbulence:
//   D_shell = 2.5G + k ds                          (shell band at level 1),
//   D_core  = G + k ds + S max(0, thetaCore - n)   (level-0 patches in it),
//   D_wake  = max(4.75G, Dw + Sw (thetaWake - w)) + max(0, r - R)
//                                 (levels 2 and 3 inside, split along w);
// the offsets above 2G and 4G leave room for the split test's margin, and Dw
// is the level-3 split threshold, so the wake's level-2/3 boundary follows
// the iso-surface w = thetaWake (features of about 10 units, which gives the
// dual mesh about 16% stitching elements per leaf cell at the default size
// with thetaWake 0.43).
// A cell of level L is split when a lower bound of D over the cell (D at the
// centre minus Lip * half diagonal) is below G 2^L.  For a field with
// Lipschitz constant Lip and G > 6.1 Lip, touching leaves differ by at most
// one level (vertex 2:1 balance, which the dual mesh needs).  With value
// noise of slope at most 2.6 f per octave, k = 1.5, S = 400 (per 4096 of
// domain), the frequencies of n below, w of 2 octaves from 1.5/G per unit and
// Sw = 4.3 (slope at most 4.3 * 2.6 * 0.09375 * 2/1.5 = 1.40, plus 1 from
// r - R) bound Lip by 2.5 (G = 16).  The scalar field uses a finer turbulence
// than the refinement.
// ---------------------------------------------------------------------------
namespace {
struct ShockField {
  double cx, cy, cz, R, h, G, lip, k, slope, wobble, wakeSlope, wakeOffset;
  Fbm shellNoise, refine, turb, wakeTurb;
  double thetaCore, thetaWake;
  double dShell(double x, double y, double z, double &r) const {
    const double dx = x - cx, dy = y - cy, dz = z - cz;
    r = sqrt(dx * dx + dy * dy + dz * dz);
    const double Rn = R + wobble * (shellNoise(x, y, z) - 0.5);
    return std::max(0.0, fabs(r - Rn) - h);
  }
  double D(double x, double y, double z) const {
    double r;
    const double ds = dShell(x, y, z, r);
    const double n = refine(x, y, z);
    const double shell = 2.5 * G + k * ds;
    const double core = G + k * ds + slope * std::max(0.0, thetaCore - n);
    const double wake =
        std::max(4.75 * G, wakeOffset + wakeSlope * (thetaWake - wakeTurb(x, y, z))) + std::max(0.0, r - R);
    return std::min(std::min(shell, core), wake);
  }
  double scalar(double x, double y, double z) const {
    double r;
    const double ds = dShell(x, y, z, r);
    const double t = turb(x, y, z);
    const double shock = exp(-(ds / (2.0 * h)) * (ds / (2.0 * h)));
    const double wake = r < R ? 0.35 * t * (r / R) : 0.05 * t;
    return std::min(1.0, 0.8 * shock * (0.75 + 0.5 * t) + wake);
  }
};

ShockField makeShock(double domain, float radius, float thickness, float thetaCore, float thetaWake,
                     uint64_t seed) {
  ShockField F;
  F.cx = F.cy = F.cz = 0.5 * domain;
  F.R = radius;
  F.h = 0.5 * thickness;
  F.G = 16.0;
  F.lip = 2.5;
  F.k = 1.5;
  // amplitudes scale with the domain so that the slopes stay domain-free
  F.slope = 400.0 * domain / 4096.0;
  F.wobble = 60.0 * domain / 4096.0;
  F.shellNoise = Fbm{2, 3.0 / domain, 1.0, seed};
  F.refine = Fbm{2, 2.0 / domain, 1.0, seed + 5};
  F.turb = Fbm{3, 6.0 / domain, 1.0, seed + 7};
  F.wakeTurb = Fbm{2, 1.5 / F.G, 1.0, seed + 11};
  F.wakeSlope = 4.3;
  F.wakeOffset = 8.0 * F.G + F.lip * 0.8660254037844386 * 8.0;
  F.thetaCore = thetaCore;
  F.thetaWake = thetaWake;
  return F;
}

template <typename Emit>
void shockCell(const ShockField &F, int x, int y, int z, int L, Emit &&emit) {
  const double s = (double)(1 << L);
  const double dLow = F.D(x + 0.5 * s, y + 0.5 * s, z + 0.5 * s) - F.lip * 0.8660254037844386 * s;
  if (L == 0 || dLow >= F.G * s) {
    emit(x, y, z, L);
    return;
  }
  const int hs = 1 << (L - 1);
  for (int dz = 0; dz < 2; ++dz)
    for (int dy = 0; dy < 2; ++dy)
      for (int dx = 0; dx < 2; ++dx) shockCell(F, x + dx * hs, y + dy * hs, z + dz * hs, L - 1, emit);
}
}  // namespace

void amrShockCount(int baseRes, int baseLevel, float radius, float thickness, float thetaCore, float thetaWake,
                   uint64_t seed, int64_t counts[16]) {
  if (baseRes <= 0 || baseLevel < 1 || baseLevel > 15 || radius <= 0.f || thickness <= 0.f)
    throw std::runtime_error("amrShock: invalid parameters");
  const int64_t B = baseRes;
  const int W = 1 << baseLevel;
  const ShockField F = makeShock((double)B * W, radius, thickness, thetaCore, thetaWake, seed);
  std::vector<std::atomic<int64_t>> c(16);
  for (auto &v : c) v.store(0);
  parallelFor(B * B * B, 4096, [&](int64_t b, int64_t e) {
    int64_t local[16] = {};
    for (int64_t id = b; id < e; ++id)
      shockCell(F, (int)(id % B) * W, (int)((id / B) % B) * W, (int)(id / (B * B)) * W, baseLevel,
                [&](int, int, int, int L) { local[L]++; });
    for (int l = 0; l < 16; ++l) c[l] += local[l];
  });
  for (int l = 0; l < 16; ++l) counts[l] = c[l].load();
}

void amrShock(AmrScene &out, int baseRes, int baseLevel, float radius, float thickness, float thetaCore,
              float thetaWake, uint64_t seed) {
  if (baseRes <= 0 || baseLevel < 1 || baseLevel > 12 || radius <= 0.f || thickness <= 0.f)
    throw std::runtime_error("amrShock: invalid parameters");
  const int64_t B = baseRes;
  const int W = 1 << baseLevel;
  const ShockField F = makeShock((double)B * W, radius, thickness, thetaCore, thetaWake, seed);
  // base cells in z-major order, chunks of base cells in parallel
  const int64_t nb = B * B * B;
  const int64_t chunk = 4096;
  const int64_t nchunks = (nb + chunk - 1) / chunk;
  std::vector<std::vector<AmrCell>> parts(nchunks);
  parallelFor(nchunks, 1, [&](int64_t b, int64_t e) {
    for (int64_t ci = b; ci < e; ++ci) {
      std::vector<AmrCell> &v = parts[ci];
      for (int64_t id = ci * chunk; id < std::min(nb, (ci + 1) * chunk); ++id)
        shockCell(F, (int)(id % B) * W, (int)((id / B) % B) * W, (int)(id / (B * B)) * W, baseLevel,
                  [&](int x, int y, int z, int L) { v.push_back(AmrCell{{x, y, z}, L}); });
    }
  });
  size_t total = 0;
  for (auto &v : parts) total += v.size();
  out.cells.clear();
  out.cells.reserve(total);
  for (auto &v : parts) {
    out.cells.insert(out.cells.end(), v.begin(), v.end());
    std::vector<AmrCell>().swap(v);
  }
  out.scalars.resize(out.cells.size());
  parallelFor((int64_t)out.cells.size(), 1 << 14, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) {
      const AmrCell &c = out.cells[i];
      const double hw = 0.5 * (double)(1 << c.level);
      out.scalars[i] = (float)F.scalar(c.pos[0] + hw, c.pos[1] + hw, c.pos[2] + hw);
    }
  });
  out.params = {(float)baseRes, (float)baseLevel, radius, thickness, thetaCore, thetaWake};
}

}  // namespace exa
